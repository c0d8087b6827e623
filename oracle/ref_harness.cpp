// ref_harness.cpp -- C entry points around the reference cpp_dmc, compiled
// from the reference sources in place (see oracle/Makefile).  TEST
// INFRASTRUCTURE ONLY: it validates the C restatement (dmc_oracle.c) and can
// serve as the timed CPU baseline.  Built with ref_prelude.h (which opens up
// `private`) so the intermediate stages can be called one by one.
// Nothing here changes the reference; this file only drives it.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <vector>

#include "DualMarchingCubes.h"
#include "UniformGrid.h"

// The tables are only read, for the table check.  LookupTables.h defines a
// non-const global (l_edges) that DualMarchingCubes.cpp also defines, so this
// translation unit takes its own copy under a separate namespace.
namespace reftab {
#include "LookupTables.h"
}

namespace {

struct Mute {   // the reference prints progress on std::cout
    std::ostringstream sink;
    std::streambuf* old;
    Mute() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~Mute() { std::cout.rdbuf(old); }
};

struct Grid { int32_t nx, ny, nz; double x0, y0, z0, dx, dy, dz; };

void make_grid(dmc::UniformGrid& ug, const Grid* g, const double* s, const double* grad) {
    dmc::UniformGrid::BBox bb;
    const double x1 = g->x0 + (g->nx - 1) * g->dx, y1 = g->y0 + (g->ny - 1) * g->dy, z1 = g->z0 + (g->nz - 1) * g->dz;
    for (int c = 0; c < 8; c++) {
        bb[c][0] = (c & 1) ? x1 : g->x0;
        bb[c][1] = (c & 2) ? y1 : g->y0;
        bb[c][2] = (c & 4) ? z1 : g->z0;
    }
    ug.init(g->nx, g->ny, g->nz, bb);
    ug.set_dx(g->dx);
    ug.set_dy(g->dy);
    ug.set_dz(g->dz);
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t q = 0; q < n; q++) ug.m_scalars[q] = s[q];
    if (grad)
        for (size_t q = 0; q < n; q++)
            for (int c = 0; c < 3; c++) ug.m_gradient[q][c] = grad[3 * q + c];
}

template <class T> T* dup(const std::vector<T>& v) {
    T* p = (T*)malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (!v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

double* flatten(const std::vector<dmc::Vector>& v) {
    double* p = (double*)malloc(sizeof(double) * 3 * (v.size() ? v.size() : 1));
    for (size_t a = 0; a < v.size(); a++)
        for (int c = 0; c < 3; c++) p[3 * a + c] = v[a][c];
    return p;
}

std::vector<dmc::Vector> unflatten(const double* p, int64_t n) {
    std::vector<dmc::Vector> v((size_t)n);
    for (int64_t a = 0; a < n; a++)
        for (int c = 0; c < 3; c++) v[(size_t)a][c] = p[3 * a + c];
    return v;
}

std::vector<dmc::DualMarchingCubes::Halfedge> he_in(const int32_t* he, int64_t nhe) {
    std::vector<dmc::DualMarchingCubes::Halfedge> h((size_t)nhe);
    for (int64_t e = 0; e < nhe; e++)
        for (int c = 0; c < 5; c++) h[(size_t)e][c] = he[5 * e + c];
    return h;
}

}  // namespace

extern "C" {

void dmcref_free(void* p) { free(p); }

const char* dmcref_r_pattern() { return reftab::dmc::r_pattern.data(); }
const unsigned char* dmcref_t_ambig() { return reftab::dmc::t_ambig; }

void dmcref_estimate_gradient(const Grid* g, const double* s, double* grad, int flip) {
    dmc::UniformGrid ug;
    make_grid(ug, g, s, nullptr);
    ug.estimateGradient();
    if (flip) ug.flip_gradient();
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t q = 0; q < n; q++)
        for (int c = 0; c < 3; c++) grad[3 * q + c] = ug.m_gradient[q][c];
}

// The cell sweep of DualMarchingCubes::dualMC (DualMarchingCubes.cpp:16-58),
// driven here so that the quad colours can be returned as well.
int dmcref_cells(const Grid* g, const double* s, const double* grad, double iso,
                 int64_t* nv, double** v, double** n, int64_t* nq, int32_t** quads,
                 int32_t** colors, uint32_t* proj1, uint32_t* proj2, int64_t* nv_before) {
    Mute mute;
    dmc::UniformGrid ug;
    make_grid(ug, g, s, grad);
    dmc::DualMarchingCubes d;
    std::vector<dmc::Vector> V, N;
    std::map<int, std::array<int, 5>> m_;
    for (int k = 0; k < g->nz - 1; k++)
        for (int j = 0; j < g->ny - 1; j++)
            for (int i = 0; i < g->nx - 1; i++) {
                double u[8];
                u[0] = ug.scalar(i, j, k);
                u[1] = ug.scalar(i + 1, j, k);
                u[2] = ug.scalar(i, j + 1, k);
                u[3] = ug.scalar(i + 1, j + 1, k);
                u[4] = ug.scalar(i, j, k + 1);
                u[5] = ug.scalar(i + 1, j, k + 1);
                u[6] = ug.scalar(i, j + 1, k + 1);
                u[7] = ug.scalar(i + 1, j + 1, k + 1);
                unsigned ic = 0;
                for (int c = 0; c < 8; c++) ic += ((unsigned)(u[c] >= iso)) << c;
                if (ic == 0 || ic == 255) continue;
                d.slice(iso, ic, i, j, k, u, ug, V, N, m_);
            }
    *nv_before = (int64_t)V.size();
    std::vector<int> Q, C;
    d.collectQuadrilaterals(Q, C, m_);
    d.removeUnusedVertices(V, N, Q);
    *nv = (int64_t)V.size();
    *v = flatten(V);
    *n = flatten(N);
    *nq = (int64_t)Q.size() / 4;
    *quads = dup(Q);
    *colors = dup(C);
    *proj1 = d.nrProj1;
    *proj2 = d.nrProj2;
    return 0;
}

int dmcref_halfedges(int64_t nv, int64_t nq, const int32_t* quads, int32_t** he, int32_t** he_v,
                     int32_t** he_f, int64_t* nbnd, int32_t** bnd) {
    Mute mute;
    dmc::DualMarchingCubes d;
    std::vector<int> Q(quads, quads + 4 * nq);
    std::vector<dmc::DualMarchingCubes::Halfedge> H;
    std::vector<int> hv, hf;
    d.halfedges((int)nv, Q, H, hv, hf);
    std::vector<int32_t> flat(5 * H.size());
    for (size_t e = 0; e < H.size(); e++)
        for (int c = 0; c < 5; c++) flat[5 * e + c] = H[e][c];
    *he = dup(flat);
    *he_v = dup(hv);
    *he_f = dup(hf);
    *nbnd = (int64_t)d.bndVertices.size();
    *bnd = dup(d.bndVertices);
    return 0;
}

void dmcref_color_faces(int64_t nq, const int32_t* he, const int32_t* he_f, int32_t* colors) {
    dmc::DualMarchingCubes d;
    auto H = he_in(he, 4 * nq);
    std::vector<int> hf(he_f, he_f + nq), C(colors, colors + nq);
    d.colorFaces(H, hf, C);
    memcpy(colors, C.data(), sizeof(int32_t) * (size_t)nq);
}

void dmcref_vertex_valence(int64_t nv, int64_t nhe, const int32_t* he, int32_t* vV) {
    Mute mute;
    dmc::DualMarchingCubes d;
    auto H = he_in(he, nhe);
    std::vector<int> out;
    d.vertexValence((int)nv, H, out);
    memcpy(vV, out.data(), sizeof(int32_t) * (size_t)nv);
}

void dmcref_mark3x3y(int64_t nq, const int32_t* quads, const int32_t* vV, int64_t nv, uint8_t* p) {
    dmc::DualMarchingCubes d;
    std::vector<int> Q(quads, quads + 4 * nq), V(vV, vV + nv);
    std::vector<bool> P;
    d.mark3X3Y(Q, V, P);
    for (int64_t f = 0; f < nq; f++) p[f] = P[(size_t)f];
}

void dmcref_merge3x3y(int64_t nv, int64_t nq, double* v, double* n, const uint8_t* p,
                      const int32_t* vV, const int32_t* colors, const int32_t* he,
                      const int32_t* he_f, uint8_t* vm_first, int32_t* vm_second, uint8_t* em) {
    dmc::DualMarchingCubes d;
    auto V = unflatten(v, nv), N = unflatten(n, nv);
    std::vector<bool> P(p, p + nq);
    std::vector<int> val(vV, vV + nv), hf(he_f, he_f + nq);
    auto H = he_in(he, 4 * nq);
    std::vector<std::pair<bool, int>> vm;
    std::vector<bool> E;
    if (colors) {
        std::vector<int> C(colors, colors + nq);
        d.mergeVertices3X3Y(V, N, P, val, C, H, hf, vm, E);
    } else {
        d.mergeVertices3X3Y(V, N, P, val, H, hf, vm, E);
    }
    for (int64_t a = 0; a < nv; a++) {
        for (int c = 0; c < 3; c++) { v[3 * a + c] = V[(size_t)a][c]; n[3 * a + c] = N[(size_t)a][c]; }
        vm_first[a] = vm[(size_t)a].first;
        vm_second[a] = vm[(size_t)a].second;
    }
    for (int64_t f = 0; f < nq; f++) em[f] = E[(size_t)f];
}

void dmcref_mark3333(int64_t nv, int64_t nq, const int32_t* quads, const int32_t* he,
                     const int32_t* he_f, const int32_t* vV, uint8_t* vm_flag, int32_t* vm_twin,
                     uint8_t* p3333, uint8_t* rflag) {
    dmc::DualMarchingCubes d;
    std::vector<int> Q(quads, quads + 4 * nq), hf(he_f, he_f + nq), val(vV, vV + nv);
    auto H = he_in(he, 4 * nq);
    std::vector<std::tuple<bool, int, int>> vm;
    std::vector<bool> P, R;
    d.mark3333((int)nv, Q, H, hf, val, vm, P, R);
    for (int64_t a = 0; a < nv; a++) { vm_flag[a] = std::get<0>(vm[(size_t)a]); vm_twin[a] = std::get<1>(vm[(size_t)a]); }
    for (int64_t f = 0; f < nq; f++) { p3333[f] = P[(size_t)f]; rflag[f] = R[(size_t)f]; }
}

// In-place whole-stage calls on a mesh {v, n, quads, colors}.
int dmcref_simplify(int which, int64_t* nv, double** v, double** n, int64_t* nq, int32_t** quads,
                    int32_t* colors) {
    Mute mute;
    dmc::DualMarchingCubes d;
    auto V = unflatten(*v, *nv), N = unflatten(*n, *nv);
    std::vector<int> Q(*quads, *quads + 4 * *nq);
    std::vector<int> C;
    if (colors) C.assign(colors, colors + *nq);
    if (which == 0) d.simplify3X3Y(V, N, Q, C);
    else d.simplify3333(V, N, Q);
    free(*v); free(*n); free(*quads);
    *nv = (int64_t)V.size();
    *v = flatten(V);
    *n = flatten(N);
    *nq = (int64_t)Q.size() / 4;
    *quads = dup(Q);
    return 0;
}

int dmcref_triangles(int64_t nv, const double* v, int64_t nq, const int32_t* quads, int32_t* tris) {
    dmc::DualMarchingCubes d;
    auto V = unflatten(v, nv);
    std::vector<int> Q(quads, quads + 4 * nq), T;
    d.collectTriangles(T, Q, V);
    memcpy(tris, T.data(), sizeof(int32_t) * T.size());
    return 0;
}

// DualMarchingCubes::dualMC as a whole (DualMarchingCubes.cpp:5-88).
int dmcref_dualmc(const Grid* g, const double* s, const double* grad, double iso, int sflag,
                  int64_t* nv, double** v, double** n, int64_t* nq, int32_t** quads,
                  int64_t* nt, int32_t** tris) {
    Mute mute;
    dmc::UniformGrid ug;
    make_grid(ug, g, s, grad);
    dmc::DualMarchingCubes d;
    std::vector<dmc::Vector> V, N;
    std::vector<int> T, Q;
    d.dualMC(iso, ug, V, N, T, Q, sflag != 0);
    *nv = (int64_t)V.size();
    *v = flatten(V);
    *n = flatten(N);
    *nq = (int64_t)Q.size() / 4;
    *quads = dup(Q);
    *nt = (int64_t)T.size() / 3;
    *tris = dup(T);
    return 0;
}

// Timing entry for the CPU baseline: UniformGrid set-up from a float field,
// estimateGradient + flip_gradient, then dualMC.  Returns the quad count;
// t_grad / t_dualmc (seconds, may be NULL) receive the wall time of the
// gradient (estimateGradient + flip_gradient) and of dualMC alone.
int64_t dmcref_run_float(const Grid* g, const float* f, double iso, int sflag, int64_t* nv_out, double* t_grad,
                         double* t_dualmc) {
    using clk = std::chrono::steady_clock;
    Mute mute;
    dmc::UniformGrid ug;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    std::vector<double> s(n);
    for (size_t q = 0; q < n; q++) s[q] = f[q];
    make_grid(ug, g, s.data(), nullptr);
    const auto t0 = clk::now();
    ug.estimateGradient();
    ug.flip_gradient();
    const auto t1 = clk::now();
    dmc::DualMarchingCubes d;
    std::vector<dmc::Vector> V, N;
    std::vector<int> T, Q;
    d.dualMC(iso, ug, V, N, T, Q, sflag != 0);
    const auto t2 = clk::now();
    if (t_grad) *t_grad = std::chrono::duration<double>(t1 - t0).count();
    if (t_dualmc) *t_dualmc = std::chrono::duration<double>(t2 - t1).count();
    if (nv_out) *nv_out = (int64_t)V.size();
    return (int64_t)Q.size() / 4;
}

// Polygons of one cell, decoded from the reference's packed contour word.
int dmcref_cell_polygons(const double* f, double iso, int32_t* sizes, int32_t* edges) {
    dmc::DualMarchingCubes d;
    unsigned ic = 0;
    for (int c = 0; c < 8; c++) ic += ((unsigned)(f[c] >= iso)) << c;
    for (int t = 0; t < 4; t++) sizes[t] = 0;
    for (int e = 0; e < 12; e++) edges[e] = -1;
    if (ic == 0 || ic == 255) return 0;
    unsigned long long c_ = 0xFFFFFFFFFFFF0000ull;
    double F[8];
    for (int c = 0; c < 8; c++) F[c] = f[c];
    unsigned cnt = (reftab::dmc::t_ambig[ic] == d.MC_AMBIGUOUS) ? d.mc_polygon(iso, F, c_) : d.mc_polygon((int)ic, c_);
    int pos = 0;
    for (unsigned t = 0; t < cnt; t++) {
        sizes[t] = d.get_cnt_size((int)t, c_);
        for (int i = 0; i < sizes[t]; i++) edges[pos++] = d.get_c((int)t, i, c_);
    }
    return (int)cnt;
}

// Local coordinates of the representatives of one cell: slice() on a
// 2x2x2 grid with origin 0 and spacing 1, so positions are local coordinates.
int dmcref_cell_vertices(const double* f, double iso, double* local, uint32_t* fail1, uint32_t* fail2) {
    Mute mute;
    Grid g{2, 2, 2, 0, 0, 0, 1, 1, 1};
    dmc::UniformGrid ug;
    make_grid(ug, &g, f, nullptr);
    unsigned ic = 0;
    for (int c = 0; c < 8; c++) ic += ((unsigned)(f[c] >= iso)) << c;
    if (ic == 0 || ic == 255) return 0;
    dmc::DualMarchingCubes d;
    std::vector<dmc::Vector> V, N;
    std::map<int, std::array<int, 5>> m_;
    double F[8];
    for (int c = 0; c < 8; c++) F[c] = f[c];
    d.slice(iso, ic, 0, 0, 0, F, ug, V, N, m_);
    for (size_t a = 0; a < V.size(); a++)
        for (int c = 0; c < 3; c++) local[3 * a + c] = V[a][c];
    *fail1 = d.nrProj1;
    *fail2 = d.nrProj2;
    return (int)V.size();
}

}  // extern "C"
