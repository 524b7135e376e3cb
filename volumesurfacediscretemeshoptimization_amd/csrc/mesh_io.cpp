// mesh_io.cpp -- mesh files for the DMO factories (host code).
//
// Formats and their conventions follow the reference's readers and writers
// (VSDMO/src/io/ParserOFF.cu:15-83, 86-160, ParserTetGen.cu:13-120,
// ParserHex.cu:115-240, FileWriter.cu:9-147):
//   OFF      "OFF", then "nV nF nE", nV lines "x y z", nF lines "3 a b c"
//            (triangles) or "4 a b c d" (quads); a mesh with any z != 0 is
//            not flat;
//   TetGen   <base>.node: "nV 3 nAttr nMarker", lines "i x y z [attrs]
//            [marker]"; <base>.ele: "nT 4 nAttr", lines "i a b c d"; the
//            numbering is 1-based when the first point index is 1; '#'
//            starts a comment;
//   MEDIT    token stream MeshVersionFormatted / Dimension / Vertices
//            (x y z ref) / Quadrilaterals|Quads (skipped) / Hexahedra
//            (8 one-based indices + ref) / End; the file's bottom face
//            (v0..v3) is the internal top face: internal v0..v3 = file
//            v4..v7 and v4..v7 = file v0..v3; indices out of range or
//            unreferenced vertices are errors.
// The writers print coordinates with 9 significant digits so that a file
// written here reads back bit-identically (the reference prints the stream
// default of 6).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "vsdmo_internal.h"

namespace vsdmo {

namespace {

struct MeshData {
    std::vector<float> xyz;
    std::vector<int32_t> cells;
    int npe = 0;
    bool flat = true;
};

void strip(std::string& line)
{
    size_t c = line.find('#');
    if (c != std::string::npos) line.erase(c);
    size_t b = line.find_first_not_of(" \t\r\n");
    line = (b == std::string::npos) ? std::string() : line.substr(b);
}

bool next_line(std::ifstream& in, std::string& line)
{
    while (std::getline(in, line)) {
        strip(line);
        if (!line.empty()) return true;
    }
    return false;
}

std::string read_off(const std::string& path, MeshData& d, int k)
{
    std::ifstream in(path);
    if (!in.is_open()) return "cannot open " + path;
    std::string line;
    if (!std::getline(in, line)) return "empty file " + path;
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    if (line != "OFF") return "first line of " + path + " is not OFF";
    if (!next_line(in, line)) return "missing counts in " + path;
    long long nV = -1, nF = -1, nE = 0;
    {
        std::istringstream s(line);
        s >> nV >> nF >> nE;
        if (!s || nV < 0 || nF < 0) return "bad counts in " + path;
    }
    d.npe = k;
    d.xyz.resize((size_t)nV * 3);
    for (long long i = 0; i < nV; ++i) {
        if (!next_line(in, line)) return "missing vertices in " + path;
        std::istringstream s(line);
        float x, y, z;
        if (!(s >> x >> y >> z)) return "bad vertex line in " + path;
        d.xyz[3 * i] = x; d.xyz[3 * i + 1] = y; d.xyz[3 * i + 2] = z;
        if (z != 0.0f) d.flat = false;
    }
    d.cells.resize((size_t)nF * k);
    for (long long i = 0; i < nF; ++i) {
        if (!next_line(in, line)) return "missing faces in " + path;
        std::istringstream s(line);
        int kf = 0;
        if (!(s >> kf) || kf != k)
            return "face line of " + path + (k == 3 ? " is not a triangle" : " is not a quad");
        for (int j = 0; j < k; ++j)
            if (!(s >> d.cells[(size_t)k * i + j])) return "bad face line in " + path;
    }
    return "";
}

std::string read_tetgen(const std::string& base, MeshData& d)
{
    std::ifstream nodes(base + ".node"), eles(base + ".ele");
    if (!nodes.is_open()) return "cannot open " + base + ".node";
    if (!eles.is_open()) return "cannot open " + base + ".ele";
    std::string line;
    if (!next_line(nodes, line)) return "empty " + base + ".node";
    long long nV = -1;
    int dim = 0, nAttr = 0, nMark = 0;
    {
        std::istringstream s(line);
        s >> nV >> dim >> nAttr >> nMark;
        if (!s || nV < 0 || nAttr < 0) return "bad header in " + base + ".node";
    }
    d.npe = 4;
    d.xyz.resize((size_t)nV * 3);
    bool oneBased = false;
    for (long long j = 0; j < nV; ++j) {
        if (!next_line(nodes, line)) return "missing points in " + base + ".node";
        std::istringstream s(line);
        long long idx;
        float x, y, z;
        if (!(s >> idx >> x >> y >> z)) return "bad point line in " + base + ".node";
        if (j == 0 && idx == 1) oneBased = true;
        d.xyz[3 * j] = x; d.xyz[3 * j + 1] = y; d.xyz[3 * j + 2] = z;
        // attributes and the boundary marker are read past; the boundary is
        // recomputed from the elements
    }
    if (!next_line(eles, line)) return "empty " + base + ".ele";
    long long nT = -1;
    int npe = 0;
    {
        std::istringstream s(line);
        s >> nT >> npe;
        if (!s || nT < 0) return "bad header in " + base + ".ele";
        if (npe != 4) return "only 4-node tetrahedra are supported";
    }
    d.cells.resize((size_t)nT * 4);
    for (long long j = 0; j < nT; ++j) {
        if (!next_line(eles, line)) return "missing tetrahedra in " + base + ".ele";
        std::istringstream s(line);
        long long idx;
        int v[4];
        if (!(s >> idx >> v[0] >> v[1] >> v[2] >> v[3])) return "bad tetrahedron line in " + base + ".ele";
        for (int k = 0; k < 4; ++k) d.cells[4 * j + k] = oneBased ? v[k] - 1 : v[k];
    }
    return "";
}

std::string read_medit(const std::string& path, MeshData& d)
{
    std::ifstream in(path);
    if (!in.is_open()) return "cannot open " + path;
    d.npe = 8;
    std::string tag;
    while (in >> tag) {
        if (tag == "MeshVersionFormatted" || tag == "Dimension") {
            int v;
            in >> v;
        } else if (tag == "Vertices") {
            long long n;
            if (!(in >> n) || n < 0) return "bad Vertices count";
            d.xyz.resize((size_t)n * 3);
            for (long long i = 0; i < n; ++i) {
                float x, y, z;
                int ref;
                if (!(in >> x >> y >> z >> ref)) return "bad vertex";
                d.xyz[3 * i] = x; d.xyz[3 * i + 1] = y; d.xyz[3 * i + 2] = z;
            }
        } else if (tag == "Quadrilaterals" || tag == "Quads") {
            long long n;
            if (!(in >> n) || n < 0) return "bad Quadrilaterals count";
            for (long long i = 0; i < n; ++i) {
                int q[5];
                if (!(in >> q[0] >> q[1] >> q[2] >> q[3] >> q[4])) return "bad quadrilateral";
            }
        } else if (tag == "Hexahedra") {
            long long n;
            if (!(in >> n) || n < 0) return "bad Hexahedra count";
            d.cells.resize((size_t)n * 8);
            for (long long h = 0; h < n; ++h) {
                int f[8], ref;
                for (int k = 0; k < 8; ++k)
                    if (!(in >> f[k])) return "bad hexahedron";
                if (!(in >> ref)) return "bad hexahedron";
                for (int k = 0; k < 4; ++k) {
                    d.cells[8 * h + k] = f[k + 4] - 1;
                    d.cells[8 * h + k + 4] = f[k] - 1;
                }
            }
        } else if (tag == "End") {
            break;
        } else {
            return "unexpected tag " + tag + " in " + path;
        }
    }
    const long long nV = (long long)d.xyz.size() / 3;
    std::vector<int> used((size_t)nV, 0);
    for (int32_t v : d.cells) {
        if (v < 0 || v >= nV) return "hexahedron index out of range";
        used[v] = 1;
    }
    for (int u : used)
        if (!u) return "the file has unreferenced vertices";
    return "";
}

std::string read_any(vsdmo_elem type, const std::string& path, MeshData& d)
{
    if (type == VSDMO_TRI) return read_off(path, d, 3);
    if (type == VSDMO_QUAD) return read_off(path, d, 4);
    if (type == VSDMO_TET) return read_tetgen(path, d);
    if (type == VSDMO_HEX) return read_medit(path, d);
    return "unsupported element type";
}

std::string write_any(vsdmo_elem type, const std::string& path, const float* xyz, int32_t nV, const int32_t* cells,
                      int32_t nC)
{
    char buf[128];
    if (type == VSDMO_TRI || type == VSDMO_QUAD) {
        std::FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return "cannot write " + path;
        std::fprintf(f, "OFF\n%d %d 0\n", nV, nC);
        for (int i = 0; i < nV; ++i) std::fprintf(f, "%.9g %.9g %.9g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        const int k = type == VSDMO_TRI ? 3 : 4;
        for (int i = 0; i < nC; ++i) {
            const int32_t* c = cells + (size_t)k * i;
            if (k == 3) std::fprintf(f, "3 %d %d %d\n", c[0], c[1], c[2]);
            else std::fprintf(f, "4 %d %d %d %d\n", c[0], c[1], c[2], c[3]);
        }
        std::fclose(f);
        return "";
    }
    if (type == VSDMO_TET) {
        std::FILE* f = std::fopen((path + ".node").c_str(), "w");
        if (!f) return "cannot write " + path + ".node";
        std::fprintf(f, "# vsdmo\n%d 3 0 0\n", nV);
        for (int i = 0; i < nV; ++i)
            std::fprintf(f, "%d %.9g %.9g %.9g\n", i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        std::fclose(f);
        f = std::fopen((path + ".ele").c_str(), "w");
        if (!f) return "cannot write " + path + ".ele";
        std::fprintf(f, "# vsdmo\n%d 4 0\n", nC);
        for (int i = 0; i < nC; ++i)
            std::fprintf(f, "%d %d %d %d %d\n", i, cells[4 * i], cells[4 * i + 1], cells[4 * i + 2], cells[4 * i + 3]);
        std::fclose(f);
        return "";
    }
    if (type == VSDMO_HEX) {
        std::FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return "cannot write " + path;
        std::fprintf(f, "MeshVersionFormatted 1\nDimension 3\nVertices %d\n", nV);
        for (int i = 0; i < nV; ++i) std::fprintf(f, "%.9g %.9g %.9g 0\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        std::fprintf(f, "Hexahedra %d\n", nC);
        for (int i = 0; i < nC; ++i) {
            const int32_t* c = cells + 8 * (size_t)i;
            std::snprintf(buf, sizeof(buf), "%d %d %d %d %d %d %d %d 0\n", c[4] + 1, c[5] + 1, c[6] + 1, c[7] + 1,
                          c[0] + 1, c[1] + 1, c[2] + 1, c[3] + 1);
            std::fputs(buf, f);
        }
        std::fprintf(f, "End\n");
        std::fclose(f);
        return "";
    }
    return "unsupported element type";
}

}  // namespace
}  // namespace vsdmo

using namespace vsdmo;

extern "C" {

vsdmo_status vsdmo_mesh_file_read(vsdmo_elem type, const char* path, int32_t* nV, int32_t* nCells, float* xyz,
                                  int32_t* cells)
{
    try {
        if (!path || !nV || !nCells) {
            set_error("NULL argument");
            return VSDMO_INVALID_ARG;
        }
        MeshData d;
        std::string err = read_any(type, path, d);
        if (!err.empty()) {
            set_error(err);
            return VSDMO_INVALID_ARG;
        }
        const long long nv = (long long)d.xyz.size() / 3, nc = d.npe ? (long long)d.cells.size() / d.npe : 0;
        if (nv >= (1LL << 31) || nc >= (1LL << 31)) {
            set_error("mesh too large");
            return VSDMO_INVALID_ARG;
        }
        *nV = (int32_t)nv;
        *nCells = (int32_t)nc;
        if (xyz) std::memcpy(xyz, d.xyz.data(), sizeof(float) * d.xyz.size());
        if (cells) std::memcpy(cells, d.cells.data(), sizeof(int32_t) * d.cells.size());
        return VSDMO_OK;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return VSDMO_HIP_ERROR;
    }
}

vsdmo_status vsdmo_mesh_file_write(vsdmo_elem type, const char* path, const float* xyz, int32_t nV,
                                   const int32_t* cells, int32_t nCells)
{
    if (!path || (nV > 0 && !xyz) || (nCells > 0 && !cells) || nV < 0 || nCells < 0) {
        set_error("bad argument");
        return VSDMO_INVALID_ARG;
    }
    std::string err = write_any(type, path, xyz, nV, cells, nCells);
    if (!err.empty()) {
        set_error(err);
        return VSDMO_INVALID_ARG;
    }
    return VSDMO_OK;
}

vsdmo_status vsdmo_mesh_create_from_file(vsdmo_elem type, const char* path, const vsdmo_mesh_params* params,
                                         vsdmo_mesh** out)
{
    if (!out || !path) {
        set_error("NULL argument");
        return VSDMO_INVALID_ARG;
    }
    *out = nullptr;
    int32_t nV = 0, nC = 0;
    vsdmo_status st = vsdmo_mesh_file_read(type, path, &nV, &nC, nullptr, nullptr);
    if (st != VSDMO_OK) return st;
    const int npe = type == VSDMO_TRI ? 3 : (type == VSDMO_TET || type == VSDMO_QUAD) ? 4 : 8;
    std::vector<float> xyz((size_t)nV * 3);
    std::vector<int32_t> cells((size_t)nC * npe);
    st = vsdmo_mesh_file_read(type, path, &nV, &nC, xyz.data(), cells.data());
    if (st != VSDMO_OK) return st;
    return vsdmo_mesh_create(type, xyz.data(), nV, cells.data(), nC, nullptr, params, out);
}

vsdmo_status vsdmo_mesh_write(vsdmo_mesh* m, const char* path, int32_t original_order)
{
    if (!m || !path) {
        set_error("NULL argument");
        return VSDMO_INVALID_ARG;
    }
    vsdmo_counts c{};
    vsdmo_status st = vsdmo_mesh_counts(m, &c);
    if (st != VSDMO_OK) return st;
    const int npe = m->npe;
    std::vector<float> xyz((size_t)c.n_vertices * 3);
    std::vector<int32_t> cells((size_t)c.n_elements * npe);
    if ((st = vsdmo_mesh_get_positions(m, xyz.data(), original_order)) != VSDMO_OK) return st;
    if ((st = vsdmo_mesh_get_cells(m, cells.data())) != VSDMO_OK) return st;
    if (original_order) {
        std::vector<int32_t> perm((size_t)c.n_vertices);   // new id of each input vertex
        if ((st = vsdmo_mesh_get_permutation(m, perm.data())) != VSDMO_OK) return st;
        std::vector<int32_t> old((size_t)c.n_vertices);
        for (int32_t i = 0; i < c.n_vertices; ++i) old[perm[i]] = i;
        for (auto& v : cells) v = old[v];
    }
    return vsdmo_mesh_file_write((vsdmo_elem)m->elem, path, xyz.data(), c.n_vertices, cells.data(), c.n_elements);
}

}  // extern "C"
