// Drop-in check of include/tmc_dmc.hpp: the reference's own driver programs
// (cpp_dmc/main.cpp:8-31 and dmc/kernel.cu:14-60) written against this
// engine.  Prints one JSON line per case and, when given a directory, dumps
// the cpp-surface input grid and mesh as raw arrays for the Python test.
#include <cstdio>
#include <fstream>
#include <string>

#include "tmc_dmc.hpp"

template <class T>
static void dump(const std::string& path, const T* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), n * sizeof(T));
}

int main(int argc, char** argv) {
    const std::string out = argc > 1 ? argv[1] : "";
    // cpp_dmc/main.cpp: 64^3 GenusTwo from Volumes, iso 0, sFlag = true
    dmc::UniformGrid ugrid;
    dmc::Volumes vol;
    dmc::DualMarchingCubes dmc;
    vol.scalar<dmc::Volumes::Surface::GenusTwo>(ugrid, 64, 64, 64);
    std::vector<dmc::Vector> v, n;
    std::vector<int> t, q;
    dmc.dualMC(0.0, ugrid, v, n, t, q, true);
    std::vector<int> bnd;
    dmc.getBndVertices(bnd);
    std::printf("{\"case\": \"cpp_genus2_64\", \"nv\": %zu, \"nq\": %zu, \"nt\": %zu, \"nbnd\": %zu}\n", v.size(),
                q.size() / 4, t.size() / 3, bnd.size());
    if (!out.empty()) {
        dump(out + "/scalars.f64", ugrid.scalars_data(), (size_t)ugrid.size());
        dump(out + "/grads.f64", ugrid.gradients_data(), 3 * (size_t)ugrid.size());
        const double geo[6] = {ugrid.minX(), ugrid.minY(), ugrid.minZ(), ugrid.dx(), ugrid.dy(), ugrid.dz()};
        dump(out + "/geo.f64", geo, 6);
        dump(out + "/v.f64", v.empty() ? nullptr : v[0].p, 3 * v.size());
        dump(out + "/n.f64", n.empty() ? nullptr : n[0].p, 3 * n.size());
        dump(out + "/quads.i32", q.data(), q.size());
        dump(out + "/tris.i32", t.data(), t.size());
    }
    // same grid written to the cpp_dmc binary format and read back: the
    // gradients then come from estimateGradient + flip_gradient
    if (!out.empty()) {
        const std::string file = out + "/grid.bin";
        {
            std::ofstream f(file, std::ios::binary);
            const int dims[3] = {ugrid.x_size(), ugrid.y_size(), ugrid.z_size()};
            const double sp[3] = {ugrid.dx(), ugrid.dy(), ugrid.dz()};
            f.write(reinterpret_cast<const char*>(dims), sizeof(dims));
            f.write(reinterpret_cast<const char*>(sp), sizeof(sp));
            for (int a = 0; a < ugrid.size(); a++) {
                const float x = (float)ugrid.scalar(a);
                f.write(reinterpret_cast<const char*>(&x), sizeof(x));
            }
        }
        dmc::UniformGrid g2;
        g2.init(file);
        std::vector<dmc::Vector> v2, n2;
        std::vector<int> t2, q2;
        dmc.dualMC(0.0, g2, v2, n2, t2, q2, true);
        std::printf("{\"case\": \"cpp_file_64\", \"nv\": %zu, \"nq\": %zu, \"on_device_gradients\": %d}\n", v2.size(),
                    q2.size() / 4, (int)g2.gradients_from_field());
        dump(out + "/file_v.f64", v2.empty() ? nullptr : v2[0].p, 3 * v2.size());
        dump(out + "/file_quads.i32", q2.data(), q2.size());
    }
    // dmc/kernel.cu: CUDA surface, 64^3 GenusTwo, all simplifications
    p_mc::DualMarchingCubes cu;
    cu.init({64, 64, 64}, p_mc::DualMarchingCubes::SurfaceCase::GenusTwo);
    std::map<std::string, int> config{{"valence", 0}, {"element-quality", 0}, {"p3X3YColor", 1}, {"p3X3YOld", 1},
                                      {"p3333", 1}, {"halfedge-datastructure", 1}, {"non-manifold", 0}};
    std::vector<p_mc::DualMarchingCubes::Vertex> cv, cn;
    std::vector<p_mc::DualMarchingCubes::Triangle> ct;
    std::vector<p_mc::DualMarchingCubes::Quadrilateral> cq;
    std::vector<p_mc::DualMarchingCubes::Halfedge> he;
    std::vector<int> hef, hev;
    cu.dualMC(0.0f, cv, cn, ct, cq, he, hef, hev, config);
    int twins_ok = 1;
    for (size_t h = 0; h < he.size(); h++)
        if (he[h][3] >= 0 && he[(size_t)he[h][3]][3] != (int)h) twins_ok = 0;
    std::printf("{\"case\": \"cuda_genus2_64\", \"nv\": %zu, \"nq\": %zu, \"nhe\": %zu, \"twins_ok\": %d}\n",
                cv.size(), cq.size(), he.size(), twins_ok);
    // the same with the CUDA path's diagnostics switched on (dmc/kernel.cu:26-32)
    config["valence"] = 1;
    config["element-quality"] = 1;
    config["non-manifold"] = 1;
    cu.dualMC(0.0f, cv, cn, ct, cq, he, hef, hev, config);
    auto total = [](const std::vector<unsigned int>& d) {
        size_t s = 0;
        for (unsigned int x : d) s += x;
        return s;
    };
    std::printf("{\"case\": \"cuda_diag_64\", \"nv\": %zu, \"nq\": %zu, \"dist_dmc\": %zu, \"dist_color\": %zu, "
                "\"dist_old\": %zu, \"dist_3333\": %zu, \"dist_tris\": %zu, \"nquality_q\": %zu, "
                "\"nquality_t\": %zu, \"nonmanifold\": %d}\n",
                cv.size(), cq.size(), total(cu.valenceDistDMC), total(cu.valenceDistP3X3YColor),
                total(cu.valenceDistP3X3YOld), total(cu.valenceDistP3333), total(cu.valenceDistTris),
                cu.elementQualityQuads.size(), cu.elementQualityTris.size(), cu.nrNonManifoldEdges);
    if (!out.empty()) {
        dump(out + "/cu_v.f32", cv.empty() ? nullptr : cv[0].data(), 3 * cv.size());
        dump(out + "/cu_quads.i32", cq.empty() ? nullptr : cq[0].data(), 4 * cq.size());
        dump(out + "/cu_tris.i32", ct.empty() ? nullptr : ct[0].data(), 3 * ct.size());
        dump(out + "/cu_dist_3333.u32", cu.valenceDistP3333.data(), cu.valenceDistP3333.size());
        dump(out + "/cu_dist_tris.u32", cu.valenceDistTris.data(), cu.valenceDistTris.size());
        dump(out + "/cu_quality_q.f32", cu.elementQualityQuads.data(), cu.elementQualityQuads.size());
        dump(out + "/cu_quality_t.f32", cu.elementQualityTris.data(), cu.elementQualityTris.size());
        std::vector<std::array<int, 2>> e;
        cu.edges(e);
        dump(out + "/cu_edges.i32", e.empty() ? nullptr : e[0].data(), 2 * e.size());
    }
    return 0;
}
