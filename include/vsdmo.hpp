// vsdmo.hpp -- header-only C++ drop-in over the C ABI (include/vsdmo.h) that
// keeps the reference's mesh / optimiser API shape:
//   DMOMeshTetFactory::create(...) -> std::shared_ptr<DMOMeshTet>
//       (reference VSDMO/src/mesh/MeshFactory.h:17-41)
//   DMO::DMOTetClass(mesh, QualityCriterium, gridScale, n_iter) with
//       optimize(), doIteration(), isDone(), getQuality(),
//       getQualityHistogram(), getElementQualities(), displayQualityGPU()
//       (reference VSDMO/src/dmo/DMOBase.h:23-88, DMOTet.h:24), and the
//       DMOHex / DMOTriFlat / DMOTri / DMOQuad classes
//   laplaceSmoothing(mesh, n_iter) (reference VSDMO/src/dmo/Laplace.h:10)
// plus the learned proposal (Dataset, MLP, DMOBaseClass::record /
// setProposal / setFallback), which has no counterpart in the reference.
// The file-name overloads read OFF / TetGen / MEDIT files (vsdmo_mesh_file_read).
// Errors become std::runtime_error carrying vsdmo_last_error().
#pragma once

#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vsdmo.h"

namespace vsdmo_cpp {

inline void check(vsdmo_status s, const char* where)
{
    if (s != VSDMO_OK) throw std::runtime_error(std::string(where) + ": " + vsdmo_last_error());
}

class DMOMesh {
public:
    DMOMesh(vsdmo_elem type, const std::vector<float>& xyz, const std::vector<int32_t>& cells, int npe,
            const vsdmo_mesh_params* prm = nullptr)
    {
        check(vsdmo_mesh_create(type, xyz.data(), (int32_t)(xyz.size() / 3), cells.data(),
                                (int32_t)(cells.size() / npe), nullptr, prm, &h_),
              "vsdmo_mesh_create");
    }
    ~DMOMesh() { if (h_) vsdmo_mesh_destroy(h_); }
    DMOMesh(const DMOMesh&) = delete;
    DMOMesh& operator=(const DMOMesh&) = delete;
    vsdmo_mesh* handle() const { return h_; }
    vsdmo_counts counts() const
    {
        vsdmo_counts c{};
        check(vsdmo_mesh_counts(h_, &c), "vsdmo_mesh_counts");
        return c;
    }
    int nVertices() const { return counts().n_vertices; }
    std::vector<float> positions(bool original_order = true) const
    {
        std::vector<float> out(3 * (size_t)nVertices());
        check(vsdmo_mesh_get_positions(h_, out.data(), original_order), "vsdmo_mesh_get_positions");
        return out;
    }

private:
    vsdmo_mesh* h_ = nullptr;
};

using DMOMeshTri = DMOMesh;
using DMOMeshTet = DMOMesh;
using DMOMeshHex = DMOMesh;
using DMOMeshQuad = DMOMesh;

// the file overloads return nullptr when the file cannot be read, like the
// reference factories
inline std::shared_ptr<DMOMesh> create_from_file(vsdmo_elem type, int npe, const std::string& path)
{
    int32_t nV = 0, nC = 0;
    if (vsdmo_mesh_file_read(type, path.c_str(), &nV, &nC, nullptr, nullptr) != VSDMO_OK) return nullptr;
    std::vector<float> xyz(3 * (size_t)nV);
    std::vector<int32_t> cells((size_t)npe * nC);
    if (vsdmo_mesh_file_read(type, path.c_str(), &nV, &nC, xyz.data(), cells.data()) != VSDMO_OK) return nullptr;
    return std::make_shared<DMOMesh>(type, xyz, cells, npe);
}

struct DMOMeshTriFactory {
    static std::shared_ptr<DMOMeshTri> create(const std::vector<float>& xyz, const std::vector<int32_t>& tris)
    {
        return std::make_shared<DMOMeshTri>(VSDMO_TRI, xyz, tris, 3);
    }
    static std::shared_ptr<DMOMeshTri> create(const std::string& offFile) { return create_from_file(VSDMO_TRI, 3, offFile); }
};
struct DMOMeshQuadFactory {
    static std::shared_ptr<DMOMeshQuad> create(const std::vector<float>& xyz, const std::vector<int32_t>& quads)
    {
        return std::make_shared<DMOMeshQuad>(VSDMO_QUAD, xyz, quads, 4);
    }
    static std::shared_ptr<DMOMeshQuad> create(const std::string& offFile) { return create_from_file(VSDMO_QUAD, 4, offFile); }
};
struct DMOMeshTetFactory {
    static std::shared_ptr<DMOMeshTet> create(const std::vector<float>& xyz, const std::vector<int32_t>& tets)
    {
        return std::make_shared<DMOMeshTet>(VSDMO_TET, xyz, tets, 4);
    }
    // TetGen base name: <name>.node and <name>.ele
    static std::shared_ptr<DMOMeshTet> create(const std::string& tetgenBase)
    {
        return create_from_file(VSDMO_TET, 4, tetgenBase);
    }
};
struct DMOMeshHexFactory {
    static std::shared_ptr<DMOMeshHex> create(const std::vector<float>& xyz, const std::vector<int32_t>& hexes)
    {
        return std::make_shared<DMOMeshHex>(VSDMO_HEX, xyz, hexes, 8);
    }
    static std::shared_ptr<DMOMeshHex> create(const std::string& meditFile) { return create_from_file(VSDMO_HEX, 8, meditFile); }
};

// The learned proposal (no counterpart in the reference): training pairs
// recorded from plain DMO, and the 96-128-128-3 network trained on them.
class Dataset {
public:
    explicit Dataset(int64_t capacity) { check(vsdmo_dataset_create(capacity, &h_), "vsdmo_dataset_create"); }
    ~Dataset() { if (h_) vsdmo_dataset_destroy(h_); }
    Dataset(const Dataset&) = delete;
    Dataset& operator=(const Dataset&) = delete;
    int64_t size() const
    {
        int64_t n = 0;
        check(vsdmo_dataset_size(h_, &n), "vsdmo_dataset_size");
        return n;
    }
    vsdmo_dataset* handle() const { return h_; }

private:
    vsdmo_dataset* h_ = nullptr;
};

class MLP {
public:
    explicit MLP(uint64_t seed = 0) { check(vsdmo_mlp_create(seed, &h_), "vsdmo_mlp_create"); }
    ~MLP() { if (h_) vsdmo_mlp_destroy(h_); }
    MLP(const MLP&) = delete;
    MLP& operator=(const MLP&) = delete;
    // Adam training; returns the mean batch loss of every epoch
    std::vector<double> train(Dataset& ds, int epochs = 20, int batch = 4096, float lr = 1e-3f, uint64_t seed = 0)
    {
        std::vector<double> loss((size_t)epochs);
        check(vsdmo_mlp_train(h_, ds.handle(), epochs, batch, lr, seed, loss.data()), "vsdmo_mlp_train");
        return loss;
    }
    vsdmo_mlp* handle() const { return h_; }

private:
    vsdmo_mlp* h_ = nullptr;
};

// Batched DMO over many meshes (no counterpart in the reference): each mesh
// ends where one optimiser per mesh with fixed_boundary = 1 ends.  The
// meshes, and the networks set as proposals, must outlive the batch.
class Batch {
public:
    explicit Batch(const std::vector<DMOMesh*>& meshes, const vsdmo_batch_params* prm = nullptr)
    {
        std::vector<vsdmo_mesh*> h;
        for (DMOMesh* m : meshes) h.push_back(m->handle());
        check(vsdmo_batch_create(h.data(), (int32_t)h.size(), prm, &h_), "vsdmo_batch_create");
        n_ = h.size();
    }
    ~Batch() { if (h_) vsdmo_batch_destroy(h_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void iterate(int n) { check(vsdmo_batch_iterate(h_, n), "vsdmo_batch_iterate"); }
    std::vector<float> minQuality() const
    {
        std::vector<float> q(n_);
        check(vsdmo_batch_min_quality(h_, q.data()), "vsdmo_batch_min_quality");
        return q;
    }
    void writeback() { check(vsdmo_batch_writeback(h_), "vsdmo_batch_writeback"); }
    vsdmo_batch_params autotune(int iters = 3, float tol = 1e-3f, const char* log = "")
    {
        vsdmo_batch_params p;
        check(vsdmo_batch_autotune(h_, iters, tol, log, &p), "vsdmo_batch_autotune");
        return p;
    }
    // learned proposal of the batch's VSDMO_TET or VSDMO_HEX meshes
    void setProposal(int32_t elem, MLP* mlp, int mode, int depth, float windowFrac)
    {
        check(vsdmo_batch_set_proposal(h_, elem, mlp ? mlp->handle() : nullptr, mode, depth, windowFrac),
              "vsdmo_batch_set_proposal");
    }
    vsdmo_batch* handle() const { return h_; }

private:
    vsdmo_batch* h_ = nullptr;
    size_t n_ = 0;
};

}  // namespace vsdmo_cpp

namespace DMO {

// the reference's criteria (dmo/DMOQualityUtil.h:10-18), same names and
// order; this library optimises MEAN_RATIO only, and an optimiser created
// with any other criterion throws (VSDMO_NOT_SUPPORTED)
enum class QualityCriterium { MEAN_RATIO = 0, AREA, RIGHT_ANGLE, JACOBIAN, MIN_ANGLE, RADIUS_RATIO, MAX_ANGLE };

class DMOBaseClass {
public:
    DMOBaseClass(vsdmo_cpp::DMOMesh& mesh, QualityCriterium q = QualityCriterium::MEAN_RATIO, float gridScale = 0.5f,
                 int n_iter = 100, const vsdmo_params* prm = nullptr, int32_t surfacePath = -1)
        : mesh_(mesh)
    {
        vsdmo_params p;
        if (prm) p = *prm;
        else vsdmo_params_default(&p);
        if (surfacePath >= 0) p.surface_path = surfacePath;
        vsdmo_cpp::check(vsdmo_opt_create(mesh.handle(), (int32_t)q, gridScale, n_iter, &p, &h_), "vsdmo_opt_create");
    }
    virtual ~DMOBaseClass() { if (h_) vsdmo_opt_destroy(h_); }
    DMOBaseClass(const DMOBaseClass&) = delete;
    DMOBaseClass& operator=(const DMOBaseClass&) = delete;

    void optimize() { vsdmo_cpp::check(vsdmo_opt_optimize(h_), "vsdmo_opt_optimize"); }
    virtual void doIteration() { vsdmo_cpp::check(vsdmo_opt_iterate(h_, 1), "vsdmo_opt_iterate"); }
    bool isDone() const { return vsdmo_opt_is_done(h_) != 0; }
    float getQuality() const
    {
        float q = 0.0f;
        vsdmo_cpp::check(vsdmo_opt_min_quality(h_, &q), "vsdmo_opt_min_quality");
        return q;
    }
    void getQualityHistogram(std::vector<int>& vec, int n_cols)
    {
        vec.assign(n_cols, 0);
        vsdmo_cpp::check(vsdmo_opt_histogram(h_, n_cols, vec.data()), "vsdmo_opt_histogram");
    }
    // host buffer of n_elements floats (the reference fills a device buffer)
    void getElementQualities(void* outFloatBuffer)
    {
        vsdmo_cpp::check(vsdmo_opt_element_qualities(h_, (float*)outFloatBuffer), "vsdmo_opt_element_qualities");
    }
    void displayQualityGPU(int n_cols = 10)
    {
        std::vector<int> h;
        getQualityHistogram(h, n_cols);
        std::printf("interval ");
        for (int i = 0; i < n_cols; ++i) std::printf("%.2f ", (float)i / n_cols);
        std::printf("\namount   ");
        for (int v : h) std::printf("%d ", v);
        std::printf("\nq_min = %g\n", getQuality());
    }
    // host buffers of 3*nu*nv floats (the reference fills device buffers)
    void getEstimateLocalSurfacePoints(int vid, int nu, int nv, void* outSurfacePoints)
    {
        vsdmo_cpp::check(vsdmo_opt_surface_points(h_, vid, nu, nv, 0, 0, (float*)outSurfacePoints),
                         "vsdmo_opt_surface_points");
    }
    void getLocalSurfacePoints(int vid, int nu, int nv, void* outSurfacePoints, int featureSid)
    {
        vsdmo_cpp::check(vsdmo_opt_surface_points(h_, vid, nu, nv, 1, featureSid, (float*)outSurfacePoints),
                         "vsdmo_opt_surface_points");
    }
    // mesh file written when the monotone-quality guard fails (reference res/fail_debug)
    void setFailureDump(const std::string& path)
    {
        vsdmo_cpp::check(vsdmo_opt_set_failure_dump(h_, path.c_str()), "vsdmo_opt_set_failure_dump");
    }
    // learned proposal (no counterpart in the reference): n plain iterations
    // recording training pairs; the proposal mode (1 network, 2 the vertex
    // itself, 0 off) and its window (measured defaults: tets 1 level on 0.025,
    // hexes 1 level on 2.0); the fallback policy (VSDMO_FALLBACK_IMMEDIATE /
    // VSDMO_FALLBACK_DEFERRED)
    void record(vsdmo_cpp::Dataset& ds, int n) { vsdmo_cpp::check(vsdmo_opt_record(h_, ds.handle(), n), "vsdmo_opt_record"); }
    void setProposal(vsdmo_cpp::MLP* mlp, int mode, int depth, float windowFrac)
    {
        vsdmo_cpp::check(vsdmo_opt_set_proposal(h_, mlp ? mlp->handle() : nullptr, mode, depth, windowFrac),
                         "vsdmo_opt_set_proposal");
    }
    void setFallback(int policy) { vsdmo_cpp::check(vsdmo_opt_set_fallback(h_, policy), "vsdmo_opt_set_fallback"); }
    vsdmo_opt* handle() const { return h_; }

protected:
    vsdmo_cpp::DMOMesh& mesh_;
    vsdmo_opt* h_ = nullptr;
};

class DMOTetClass : public DMOBaseClass {
public:
    DMOTetClass(vsdmo_cpp::DMOMeshTet& m, QualityCriterium q = QualityCriterium::MEAN_RATIO, float gridScale = 0.5f,
                int n_iter = 30, const vsdmo_params* prm = nullptr)
        : DMOBaseClass(m, q, gridScale, n_iter, prm) {}
};
class DMOHexClass : public DMOBaseClass {
public:
    DMOHexClass(vsdmo_cpp::DMOMeshHex& m, QualityCriterium q = QualityCriterium::MEAN_RATIO, float gridScale = 0.5f,
                int n_iter = 30, const vsdmo_params* prm = nullptr)
        : DMOBaseClass(m, q, gridScale, n_iter, prm) {}
};
class DMOTriFlatClass : public DMOBaseClass {
public:
    // the reference's default criterion for flat triangle meshes is MIN_ANGLE
    // (dmo/DMOTriFlat.h:25), which this library does not optimise: the
    // criterion has no default here, so reference code that relied on it
    // fails to compile instead of silently optimising another objective
    DMOTriFlatClass(vsdmo_cpp::DMOMeshTri& m, QualityCriterium q, float gridScale = 0.5f, int n_iter = 30,
                    const vsdmo_params* prm = nullptr)
        : DMOBaseClass(m, q, gridScale, n_iter, prm, VSDMO_SURFACE_PATH_FLAT) {}
};
// triangle / quad surface meshes: vertices move on their local surfaces
// (reference dmo/DMOTri.h:24, dmo/DMOQuad.h:18)
class DMOTriClass : public DMOBaseClass {
public:
    DMOTriClass(vsdmo_cpp::DMOMeshTri& m, QualityCriterium q = QualityCriterium::MEAN_RATIO, float gridScale = 0.5f,
                int n_iter = 30, const vsdmo_params* prm = nullptr)
        : DMOBaseClass(m, q, gridScale, n_iter, prm, VSDMO_SURFACE_PATH_SURFACE) {}
};
class DMOQuadClass : public DMOBaseClass {
public:
    DMOQuadClass(vsdmo_cpp::DMOMeshQuad& m, QualityCriterium q = QualityCriterium::MEAN_RATIO, float gridScale = 0.5f,
                 int n_iter = 30, const vsdmo_params* prm = nullptr)
        : DMOBaseClass(m, q, gridScale, n_iter, prm, VSDMO_SURFACE_PATH_SURFACE) {}
};

}  // namespace DMO

// reference dmo/Laplace.h:10 (free vertices of a triangle / quad mesh)
inline void laplaceSmoothing(vsdmo_cpp::DMOMesh& mesh, int n_iter)
{
    vsdmo_cpp::check(vsdmo_mesh_laplace(mesh.handle(), n_iter), "vsdmo_mesh_laplace");
}
