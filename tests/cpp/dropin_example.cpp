// Drop-in usage of the C++ wrapper exactly as the reference's CLI drives DMO
// (reference VSDMO/src/main.cu:51-73: factory create, DMOTetClass, optimize).
// Builds a 4^3 Kuhn cube, runs 3 iterations and prints the minimum quality.
#include <cstdio>
#include <vector>

#include "vsdmo.hpp"

int main()
{
    const int n = 4, m = n + 1;
    std::vector<float> xyz;
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                bool inner = i > 0 && j > 0 && k > 0 && i < n && j < n && k < n;
                float d = inner ? 0.05f * (float)((i * 7 + j * 3 + k) % 5 - 2) / n : 0.0f;
                xyz.push_back((float)i / n + d);
                xyz.push_back((float)j / n - d);
                xyz.push_back((float)k / n + 0.5f * d);
            }
    auto id = [&](int i, int j, int k) { return i + m * (j + m * k); };
    std::vector<int32_t> tets;
    const int perms[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (auto& p : perms) {
                    int c[3] = { i, j, k };
                    int v0 = id(i, j, k);
                    c[p[0]]++;
                    int v1 = id(c[0], c[1], c[2]);
                    c[p[1]]++;
                    int v2 = id(c[0], c[1], c[2]);
                    int v3 = id(i + 1, j + 1, k + 1);
                    bool even = (p[0] == 0 && p[1] == 1) || (p[0] == 1 && p[1] == 2) || (p[0] == 2 && p[1] == 0);
                    if (even) tets.insert(tets.end(), { v0, v1, v2, v3 });
                    else tets.insert(tets.end(), { v0, v1, v3, v2 });
                }
    try {
        auto mesh = vsdmo_cpp::DMOMeshTetFactory::create(xyz, tets);
        DMO::DMOTetClass dmo(*mesh, DMO::QualityCriterium::MEAN_RATIO, 0.5f, 3);
        dmo.optimize();
        std::printf("min quality %g done %d\n", dmo.getQuality(), (int)dmo.isDone());
        dmo.displayQualityGPU(10);
        // the learned proposal (no counterpart in the reference): training
        // pairs from plain iterations, a network, the proposal search
        vsdmo_params prm;
        vsdmo_params_default(&prm);
        prm.fixed_boundary = 1;
        vsdmo_cpp::Dataset ds(100000);
        auto recMesh = vsdmo_cpp::DMOMeshTetFactory::create(xyz, tets);
        DMO::DMOTetClass rec(*recMesh, DMO::QualityCriterium::MEAN_RATIO, 0.5f, 3, &prm);
        rec.record(ds, 3);
        vsdmo_cpp::MLP mlp(1);
        std::vector<double> loss = mlp.train(ds, 3, 64, 1e-3f, 2);
        auto learnMesh = vsdmo_cpp::DMOMeshTetFactory::create(xyz, tets);
        DMO::DMOTetClass learned(*learnMesh, DMO::QualityCriterium::MEAN_RATIO, 0.5f, 3, &prm);
        learned.setProposal(&mlp, 1, 1, 0.025f);
        learned.setFallback(VSDMO_FALLBACK_DEFERRED);
        const float q0 = learned.getQuality();
        learned.optimize();
        std::printf("learned: %lld pairs, loss %g -> %g, min quality %g -> %g\n", (long long)ds.size(), loss.front(),
                    loss.back(), q0, learned.getQuality());
        if (!(learned.getQuality() >= q0)) return 2;
        // the batch (no counterpart in the reference) over two copies of the
        // cube, the network as the proposal of its tet meshes
        auto b0 = vsdmo_cpp::DMOMeshTetFactory::create(xyz, tets);
        auto b1 = vsdmo_cpp::DMOMeshTetFactory::create(xyz, tets);
        vsdmo_cpp::Batch batch({ b0.get(), b1.get() });
        batch.setProposal(VSDMO_TET, &mlp, 1, 1, 0.025f);
        batch.iterate(3);
        const std::vector<float> bq = batch.minQuality();
        std::printf("batch: min quality %g %g\n", bq[0], bq[1]);
        if (!(bq[0] == bq[1] && bq[0] >= q0)) return 3;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
