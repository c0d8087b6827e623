// tmc_ref_harness.cpp -- C entry point around the reference's topologically
// correct Marching Cubes (mc/MarchingCubes.cpp), compiled from the reference
// sources in place (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: it pins
// the C restatement (tmc_oracle.c).  Built with ref_prelude.h (which opens up
// `private`) so the grid can be set and t_mc called directly; the file
// reader and UGrid::init(nx, ny, nz, bb) are not used.
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <vector>

#include "MarchingCubes.h"

namespace {

struct Mute {   // the reference prints progress on std::cout
    std::ostringstream sink;
    std::streambuf* old;
    Mute() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~Mute() { std::cout.rdbuf(old); }
};

tmc::MarchingCubes* g_mc = nullptr;

}  // namespace

extern "C" {

// t_mc on an nx*ny*nz field of doubles (x fastest), point (i, j, k) at
// (x0 + i dx, ...); gradients from UniformGrid::gradient; with c_flag then
// connectivity() (the triangle orientation pass).  Returns the vertex and
// triangle counts; tmcref_fetch copies the lists.
int tmcref_run(int nx, int ny, int nz, double x0, double y0, double z0, double dx, double dy, double dz,
               const double* field, double iso, int c_flag, int64_t* nv, int64_t* nt) {
    delete g_mc;
    g_mc = new tmc::MarchingCubes();
    auto& g = g_mc->m_ugrid;
    g.m_nx = nx;
    g.m_ny = ny;
    g.m_nz = nz;
    g.m_dx = dx;
    g.m_dy = dy;
    g.m_dz = dz;
    for (auto& p : g.m_bbox) p = {{x0, y0, z0}};
    const size_t n = (size_t)nx * ny * nz;
    g.m_scalars.assign(field, field + n);
    g.m_normals.assign(n, {{0, 0, 0}});
    {
        Mute m;
        g.gradient();
        g_mc->t_mc(iso);
        if (c_flag) {
            std::streambuf* err = std::cerr.rdbuf(m.sink.rdbuf());
            g_mc->connectivity();   // operator()'s orientation pass (c_flag)
            std::cerr.rdbuf(err);
        }
    }
    *nv = (int64_t)g_mc->m_points.size();
    *nt = (int64_t)g_mc->m_triangles.size();
    return 0;
}

void tmcref_fetch(double* v, double* n, int32_t* t) {
    const auto& P = g_mc->m_points;
    const auto& N = g_mc->m_pnorms;
    for (size_t i = 0; i < P.size(); i++)
        for (int x = 0; x < 3; x++) {
            v[3 * i + x] = P[i][x];
            n[3 * i + x] = N[i][x];
        }
    const auto& T = g_mc->m_triangles;
    for (size_t i = 0; i < T.size(); i++)
        for (int x = 0; x < 3; x++) t[3 * i + x] = T[i].v[x];
}

}  // extern "C"
