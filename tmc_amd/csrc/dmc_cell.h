// dmc_cell.h -- per-cell device math of the DMC engine (gfx950).
//
// Cell corners are numbered c = x | y<<1 | z<<2.  The functions here follow
// the reference cpp_dmc cell path (cpp_dmc/DualMarchingCubes.cpp:97-481,
// DualMarchingCubes.h:301-597, UniformGrid.h:286-305) including where it
// rounds to float; this file must be compiled with -ffp-contract=off so that
// no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmc_tables.h"

namespace tmc {

// ---------------------------------------------------------------------------
// cell geometry
// ---------------------------------------------------------------------------
// Edge end points: v0 is always the lower-coordinate corner
// (cpp_dmc/DualMarchingCubes.cpp:110).
constexpr int kEdgeV0[12] = {0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2};
constexpr int kEdgeV1[12] = {1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6};
// Axis of each edge: 0 = x, 1 = y, 2 = z.
constexpr int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// Slot of a cell's vertex in the quadrilateral around edge e, for the two
// orientations (bit = field at v0 below the iso-value); DualMarchingCubes.h:226-230.
constexpr int kEdgeSlot[12][2] = {{0, 0}, {0, 0}, {3, 1}, {3, 1}, {1, 3}, {1, 3},
                                  {2, 2}, {2, 2}, {0, 0}, {3, 1}, {2, 2}, {1, 3}};
// Oriented faces seen from outside: corners and edges in face order
// (DualMarchingCubes.h:301-303).
constexpr int kFaceV[6][4] = {{0, 2, 1, 3}, {5, 7, 4, 6}, {4, 0, 5, 1},
                              {2, 6, 3, 7}, {4, 6, 0, 2}, {1, 3, 5, 7}};
constexpr int kFaceE[6][4] = {{3, 2, 1, 0}, {5, 6, 7, 4}, {8, 0, 9, 4},
                              {11, 6, 10, 2}, {7, 11, 3, 8}, {1, 10, 5, 9}};

constexpr uint64_t pack4(const int (&a)[12]) {
    uint64_t w = 0;
    for (int i = 0; i < 12; i++) w |= uint64_t(a[i]) << (4 * i);
    return w;
}
constexpr uint64_t kEdgeV0Packed = pack4(kEdgeV0);
constexpr uint64_t kEdgeV1Packed = pack4(kEdgeV1);
constexpr uint64_t kEdgeAxisPacked = pack4(kEdgeAxis);
constexpr uint64_t pack_slots() {
    uint64_t w = 0;
    for (int e = 0; e < 12; e++) w |= uint64_t(kEdgeSlot[e][0] | kEdgeSlot[e][1] << 2) << (4 * e);
    return w;
}
constexpr uint64_t kEdgeSlotPacked = pack_slots();

__device__ __forceinline__ int edge_v0(int e) { return (int)((kEdgeV0Packed >> (4 * e)) & 0xF); }
__device__ __forceinline__ int edge_v1(int e) { return (int)((kEdgeV1Packed >> (4 * e)) & 0xF); }
__device__ __forceinline__ int edge_axis(int e) { return (int)((kEdgeAxisPacked >> (4 * e)) & 0xF); }
__device__ __forceinline__ int edge_slot(int e, int bit) {
    return (int)((kEdgeSlotPacked >> (4 * e + 2 * bit)) & 3);
}
// Does the iso-surface cross edge e of a cell with case `icase`?
__device__ __forceinline__ bool edge_cut(unsigned icase, int e) {
    return ((icase >> edge_v0(e)) ^ (icase >> edge_v1(e))) & 1u;
}

// Is the grid edge e of cell (i,j,k) shared by four existing cells?  Only
// such edges yield complete quadrilaterals (DualMarchingCubes.h:663-676).
__device__ __forceinline__ bool edge_interior(int e, int i, int j, int k, int nx, int ny, int nz) {
    const int v0 = edge_v0(e);
    const int pi = i + (v0 & 1), pj = j + ((v0 >> 1) & 1), pk = k + ((v0 >> 2) & 1);
    const int ax = edge_axis(e);
    const bool okx = pi >= 1 && pi <= nx - 2;
    const bool oky = pj >= 1 && pj <= ny - 2;
    const bool okz = pk >= 1 && pk <= nz - 2;
    return ax == 0 ? (oky && okz) : ax == 1 ? (okx && okz) : (okx && oky);
}

// ---------------------------------------------------------------------------
// MC polygons, packed: bits 0..15 = four 4-bit contour sizes (0 = absent),
// bits 16..63 = edge ids, contour after contour.
// ---------------------------------------------------------------------------
__device__ __constant__ uint64_t c_case_contours[256] = DMC_CASE_CONTOURS_INIT;
__device__ __constant__ uint64_t c_ambiguous[4] = DMC_AMBIGUOUS_MASK_INIT;

__device__ __forceinline__ bool case_ambiguous(unsigned icase) {
    return (c_ambiguous[icase >> 6] >> (icase & 63)) & 1ull;
}
__device__ __forceinline__ int poly_size(uint64_t w, int t) { return (int)((w >> (4 * t)) & 0xF); }
__device__ __forceinline__ int poly_count(uint64_t w) {
    return (poly_size(w, 0) != 0) + (poly_size(w, 1) != 0) + (poly_size(w, 2) != 0) + (poly_size(w, 3) != 0);
}
__device__ __forceinline__ int poly_start(uint64_t w, int t) {
    int s = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) s += (i < t) ? poly_size(w, i) : 0;
    return s;
}
__device__ __forceinline__ int poly_edge_at(uint64_t w, int pos) { return (int)((w >> (16 + 4 * pos)) & 0xF); }

// cpp_dmc/DualMarchingCubes.h:330-332
__device__ __forceinline__ double asymptotic_decider(double f0, double f1, double f2, double f3) {
    return (f0 * f3 - f1 * f2) / (f0 + f3 - f1 - f2);
}

// Polygons of an ambiguous case (cpp_dmc/DualMarchingCubes.cpp:328-461):
// each face contributes oriented segments, faces with two opposite positive
// corners are split by the asymptotic decider, and segments are chained into
// contours starting from the lowest intersected edge.
__device__ inline uint64_t polygons_decider(double iso, const double F[8]) {
    uint64_t outs = 0;   // 4-bit successor per edge
    unsigned set = 0;    // edges touched by a segment
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const double f0 = F[kFaceV[f][0]], f1 = F[kFaceV[f][1]], f2 = F[kFaceV[f][2]], f3 = F[kFaceV[f][3]];
        const int e0 = kFaceE[f][0], e1 = kFaceE[f][1], e2 = kFaceE[f][2], e3 = kFaceE[f][3];
        const unsigned fc = (unsigned)(f0 >= iso) | (unsigned)(f1 >= iso) << 1 |
                            (unsigned)(f2 >= iso) << 2 | (unsigned)(f3 >= iso) << 3;
        int a0 = -1, b0 = -1, a1 = -1, b1 = -1;   // up to two segments a->b
        switch (fc) {
        case 1: a0 = e0; b0 = e3; break;
        case 2: a0 = e1; b0 = e0; break;
        case 3: a0 = e1; b0 = e3; break;
        case 4: a0 = e3; b0 = e2; break;
        case 5: a0 = e0; b0 = e2; break;
        case 6: {
            const double val = asymptotic_decider(f0, f1, f2, f3);
            if (val >= iso) { a0 = e3; b0 = e0; a1 = e1; b1 = e2; }
            else if (val < iso) { a0 = e1; b0 = e0; a1 = e3; b1 = e2; }
        } break;
        case 7: a0 = e1; b0 = e2; break;
        case 8: a0 = e2; b0 = e1; break;
        case 9: {
            const double val = asymptotic_decider(f0, f1, f2, f3);
            if (val >= iso) { a0 = e0; b0 = e1; a1 = e2; b1 = e3; }
            else if (val < iso) { a0 = e0; b0 = e3; a1 = e2; b1 = e1; }
        } break;
        case 10: a0 = e2; b0 = e0; break;
        case 11: a0 = e2; b0 = e3; break;
        case 12: a0 = e3; b0 = e1; break;
        case 13: a0 = e0; b0 = e1; break;
        case 14: a0 = e3; b0 = e0; break;
        default: break;
        }
        if (a0 >= 0) {
            outs = (outs & ~(0xFull << (4 * a0))) | ((uint64_t)b0 << (4 * a0));
            set |= (1u << a0) | (1u << b0);
        }
        if (a1 >= 0) {
            outs = (outs & ~(0xFull << (4 * a1))) | ((uint64_t)b1 << (4 * a1));
            set |= (1u << a1) | (1u << b1);
        }
    }
    uint64_t w = 0;
    int cnt = 0, pos = 0;
    for (int e = 0; e < 12; e++) {
        if (!((set >> e) & 1u)) continue;
        if (cnt >= 4 || pos >= 12) break;
        int size = 1;
        w |= (uint64_t)e << (16 + 4 * pos);
        pos++;
        int to = (int)((outs >> (4 * e)) & 0xF);
        while (to != e && to < 12 && pos < 12) {
            w |= (uint64_t)to << (16 + 4 * pos);
            pos++;
            size++;
            const int cur = to;
            to = (int)((outs >> (4 * cur)) & 0xF);
            set &= ~(1u << cur);
        }
        w |= (uint64_t)size << (4 * cnt);
        cnt++;
    }
    return w;
}

__device__ __forceinline__ uint64_t cell_polygons(unsigned icase, double iso, const double F[8]) {
    return case_ambiguous(icase) ? polygons_decider(iso, F) : c_case_contours[icase];
}

// Pick element t of a 4-array without dynamic register indexing.
template <class T>
__device__ __forceinline__ T pick4(const T (&a)[4], int t) {
    return t == 0 ? a[0] : t == 1 ? a[1] : t == 2 ? a[2] : a[3];
}
template <class T>
__device__ __forceinline__ T pick8(const T* a, int t) {
    const T lo = (t & 2) ? ((t & 1) ? a[3] : a[2]) : ((t & 1) ? a[1] : a[0]);
    const T hi = (t & 2) ? ((t & 1) ? a[7] : a[6]) : ((t & 1) ? a[5] : a[4]);
    return (t & 4) ? hi : lo;
}

// Corner values of one cell kept in LDS (one column per thread) for reads
// with a run-time corner index; register arrays indexed that way end up in
// scratch memory.
constexpr int kCornerStride = 256;
struct CornerRef {
    double* p;   // &corners[0][threadIdx.x] of a [8][kCornerStride] LDS array
    __device__ __forceinline__ double operator()(int c) const { return p[c * kCornerStride]; }
    __device__ __forceinline__ void store(const double F[8]) const {
#pragma unroll
        for (int c = 0; c < 8; c++) p[c * kCornerStride] = F[c];
    }
};

// Edge parameter rounded to float (cpp_dmc/DualMarchingCubes.h:391-407): the
// coordinate along the edge axis is l, the others 0 or 1 by the position of v0.
__device__ __forceinline__ float edge_param(int e, const CornerRef& Fc, double iso) {
    const double f0 = Fc(edge_v0(e)), f1 = Fc(edge_v1(e));
    const float l = (iso - f0) / (f1 - f0);
    return l;
}
__device__ __forceinline__ void local_coords(int e, float l, double u[3]) {
    const int v0 = edge_v0(e), ax = edge_axis(e);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float val = (ax == c) ? 1.f : 0.f;
        const float con = (float)((v0 >> c) & 1);
        u[c] = l * val + con;
    }
}

// Per-contour bounding boxes live in LDS, one column per thread: entry
// (axis, lo/hi, contour) of thread x is at base[((axis*2 + hi)*4 + t) * 256 + x].
// All entries are floats widened to double in the reference, so float
// storage is exact.
constexpr int kBoxStride = 256;
constexpr int kBoxFloats = 24;

struct BoxRef {
    float* p;   // &box[0][threadIdx.x] of a [24][kBoxStride] LDS array
    __device__ __forceinline__ float lo(int ax, int t) const { return p[((ax * 2) * 4 + t) * kBoxStride]; }
    __device__ __forceinline__ float hi(int ax, int t) const { return p[((ax * 2 + 1) * 4 + t) * kBoxStride]; }
    __device__ __forceinline__ void set(int ax, int t, float l, float h) const {
        p[((ax * 2) * 4 + t) * kBoxStride] = l;
        p[((ax * 2 + 1) * 4 + t) * kBoxStride] = h;
    }
};

// cpp_dmc/DualMarchingCubes.h:590-597
__device__ __forceinline__ void trilinear_gradient(const CornerRef& f, double u, double v, double w, double g[3]) {
    g[0] = (1 - w) * ((1 - v) * (f(1) - f(0)) + v * (f(3) - f(2))) + w * ((1 - v) * (f(5) - f(4)) + v * (f(7) - f(6)));
    g[1] = (1 - w) * ((1 - u) * (f(2) - f(0)) + u * (f(3) - f(1))) + w * ((1 - u) * (f(6) - f(4)) + u * (f(7) - f(5)));
    g[2] = (1 - v) * ((1 - u) * (f(4) - f(0)) + u * (f(5) - f(1))) + v * ((1 - u) * (f(6) - f(2)) + u * (f(7) - f(3)));
}

// Projection sampling (cpp_dmc/DualMarchingCubes.cpp:274-326).  The frame
// (a, b, c) maps the projected (u, v, w) axes to cell axes; corner q of the
// frame (q = bit0 along a, bit1 along b, bit2 along c) is cell corner
// (cpk >> 3q) & 7 (corners 0..3 = bilinear face at w = 0, 4..7 at w = 1).
// The blends along u depend on the row i only and are formed once per row,
// exactly as the reference's expression groups them.  Boxes of other size-3
// contours are only consulted for multi-contour cells
// (DualMarchingCubes.h:500-519).
struct ProjFrame {
    double umin, vmin, du, dv, wmin, wmax;
};

__device__ __forceinline__ ProjFrame proj_frame(const BoxRef& B, int axa, int axb, int axc, int t, int ns) {
    ProjFrame f;
    const double umax = B.hi(axa, t), vmax = B.hi(axb, t);
    f.umin = B.lo(axa, t);
    f.vmin = B.lo(axb, t);
    f.du = (umax - f.umin) / (ns - 1);
    f.dv = (vmax - f.vmin) / (ns - 1);
    f.wmin = B.lo(axc, t);
    f.wmax = B.hi(axc, t);
    const double eps = 1e-5;
    if (fabs(f.wmax - f.wmin) < eps) {
        f.wmin -= eps;
        f.wmax += eps;
    }
    return f;
}

// Blends of the four frame edges along u at row i.
__device__ __forceinline__ void proj_row(const CornerRef& Fc, uint32_t cpk, float u, double& u01, double& u23,
                                         double& u45, double& u67) {
    u01 = (1 - u) * Fc((cpk >> 0) & 7) + u * Fc((cpk >> 3) & 7);
    u23 = (1 - u) * Fc((cpk >> 6) & 7) + u * Fc((cpk >> 9) & 7);
    u45 = (1 - u) * Fc((cpk >> 12) & 7) + u * Fc((cpk >> 15) & 7);
    u67 = (1 - u) * Fc((cpk >> 18) & 7) + u * Fc((cpk >> 21) & 7);
}

// 9 x 9 sampling of a cell without a neighbouring size-3 contour box to
// avoid.  The reference keeps the first sample (in (i, j) order) of least
// distance d among those whose w lies in [wmin, wmax]; any evaluation order
// that breaks ties of d by that order finds the same sample.  Rows are
// evaluated nearest to p first, outwards on both sides, and a side stops at
// the first row that cannot hold a winner: with a_i = (p0 - u_i)^2 and
// b_j = (p1 - v_j)^2 as the distance expression rounds them, every d of row
// i is at least float(a_i + min_j b_j) (rounding is monotone and the w term
// is not negative), and a_i grows away from p on each side.  A sample with
// g1 == g2 gives w = +-inf or NaN, which no [wmin, wmax] test accepts,
// exactly as the reference's skip of that sample.
// kUnroll: how many of a row's 7 samples the compiler interleaves (7: the
// most instruction-level parallelism, 1: the fewest registers).
template <int kUnroll>
__device__ inline bool projection9(const CornerRef& Fc, uint32_t cpk, const BoxRef& B, int axa, int axb, int axc,
                                   double iso, int t, double p[3]) {
    constexpr int ns = 9;
    constexpr int kIn = ns - 2;  // interior samples per axis: i, j = 1 .. 7
    const ProjFrame f = proj_frame(B, axa, axb, axc, t, ns);
    float vj[kIn];
    double bj[kIn];
    double bmin = __builtin_huge_val();
#pragma unroll
    for (int j = 1; j <= kIn; j++) {
        const float v = f.vmin + j * f.dv;
        vj[j - 1] = v;
        bj[j - 1] = (p[1] - v) * (p[1] - v);
        bmin = bj[j - 1] < bmin ? bj[j - 1] : bmin;
    }
    auto row_a = [&](int i) {
        const float u = f.umin + i * f.du;
        return (p[0] - u) * (p[0] - u);
    };
    auto row_bound = [&](int i) { return (double)(float)(row_a(i) + bmin); };
    // first row: about the nearest to p0 (any row is correct)
    const float x = fminf(fmaxf((float)((p[0] - f.umin) / f.du) + 0.5f, 1.f), (float)kIn);
    int cur = (int)x, lo = cur - 1, hi = cur + 1;
    // the best sample as one key: the float distance's bits (a taken
    // distance is finite and not negative, so its bits order as the values
    // do) above the sample's index, so that (distance, index) compares in
    // one unsigned compare; starts at distance 100, the reference's bound
    const uint64_t start = (uint64_t)__float_as_uint(100.f) << 32;
    uint64_t bestKey = start;
    float bu = 0.f, bv = 0.f, bw = -1.f;
    while (__any(cur >= 0)) {
        if (cur >= 0) {
            const int i = cur;
            const float u = f.umin + i * f.du;
            const double ai = (p[0] - u) * (p[0] - u);
            double u01, u23, u45, u67;
            proj_row(Fc, cpk, u, u01, u23, u45, u67);
#pragma unroll kUnroll
            for (int j = 1; j <= kIn; j++) {
                const float v = vj[j - 1];
                const float g1 = (1 - v) * u01 + v * u23;
                const float g2 = (1 - v) * u45 + v * u67;
                const float ww = (iso - g1) / (g2 - g1);
                const double s = ai + bj[j - 1];
                const float d = s + (p[2] - ww) * (p[2] - ww);
                const uint64_t key = (uint64_t)__float_as_uint(d) << 32 | (uint32_t)(ns * i + j);
                const bool take = f.wmin <= ww && ww <= f.wmax && key < bestKey;
                bestKey = take ? key : bestKey;
                bu = take ? u : bu;
                bv = take ? v : bv;
                bw = take ? ww : bw;
            }
            // next row: the side whose next row has the smaller bound, among
            // the sides that can still hold a winner
            const double best = (double)__uint_as_float((uint32_t)(bestKey >> 32));
            double bl = 0, bh = 0;
            if (lo >= 1) {
                bl = row_bound(lo);
                if (bl > best) lo = 0;
            }
            if (hi <= kIn) {
                bh = row_bound(hi);
                if (bh > best) hi = kIn + 1;
            }
            const bool okl = lo >= 1, okh = hi <= kIn;
            cur = (okl && (!okh || bl <= bh)) ? lo-- : okh ? hi++ : -1;
        }
    }
    const bool found = bestKey != start;
    if (found) {
        p[0] = bu;
        p[1] = bv;
        p[2] = bw;
    }
    return found;
}

// General sampling with ns x ns samples and the neighbour-box test, in the
// order of projection9: rows nearest to p first, outwards on both sides, a
// side stopping at the first row that cannot hold a winner, and the best
// sample kept as one (distance bits, index) key so that ties resolve as in
// (i, j) order.  The box test only removes samples, so float(a_i + min_j
// b_j) stays a lower bound of every distance row i can contribute.  The
// distance is the reference's expression, ((p0 - u)^2 + (p1 - v)^2) +
// (p2 - w)^2 in double, rounded to float.
__device__ inline bool projection(const CornerRef& Fc, uint32_t cpk, const BoxRef& B, int axa, int axb, int axc,
                                  double iso, int t, unsigned nbmask, double p[3], int ns) {
    const ProjFrame f = proj_frame(B, axa, axb, axc, t, ns);
    const int kIn = ns - 2;   // interior samples per axis: i, j = 1 .. ns - 2
    double bmin = __builtin_huge_val();
#pragma unroll 1
    for (int j = 1; j <= kIn; j++) {
        const float v = f.vmin + j * f.dv;
        const double b = (p[1] - v) * (p[1] - v);
        bmin = b < bmin ? b : bmin;
    }
    auto row_bound = [&](int i) {
        const float u = f.umin + i * f.du;
        return (double)(float)((p[0] - u) * (p[0] - u) + bmin);
    };
    const float x = fminf(fmaxf((float)((p[0] - f.umin) / f.du) + 0.5f, 1.f), (float)kIn);
    int cur = (int)x, lo = cur - 1, hi = cur + 1;
    const uint64_t start = (uint64_t)__float_as_uint(100.f) << 32;
    uint64_t bestKey = start;
    float bu = 0.f, bv = 0.f, bw = -1.f;
    while (__any(cur >= 0)) {
        if (cur >= 0) {
            const int i = cur;
            const float u = f.umin + i * f.du;
            const double ai = (p[0] - u) * (p[0] - u);
            double u01, u23, u45, u67;
            proj_row(Fc, cpk, u, u01, u23, u45, u67);
#pragma unroll 1
            for (int j = 1; j <= kIn; j++) {
                const float v = f.vmin + j * f.dv;
                const float g1 = (1 - v) * u01 + v * u23;
                const float g2 = (1 - v) * u45 + v * u67;
                // g1 == g2: w is +-inf or NaN, which the range test rejects
                const float ww = (iso - g1) / (g2 - g1);
                if (!(f.wmin <= ww && ww <= f.wmax)) continue;
                bool inside = false;
                if (nbmask) {
                    const double du_ = u, dv_ = v, dw_ = ww;
                    for (int q = 0; q < 4; q++) {
                        if (!((nbmask >> q) & 1u) || inside) continue;
                        if (du_ < (double)B.lo(axa, q) || (double)B.hi(axa, q) < du_) continue;
                        if (dv_ < (double)B.lo(axb, q) || (double)B.hi(axb, q) < dv_) continue;
                        if (dw_ < (double)B.lo(axc, q) || (double)B.hi(axc, q) < dw_) continue;
                        inside = true;
                    }
                }
                if (inside) continue;
                const double s = ai + (p[1] - v) * (p[1] - v);
                const float d = s + (p[2] - ww) * (p[2] - ww);
                const uint64_t key = (uint64_t)__float_as_uint(d) << 32 | (uint32_t)(ns * i + j);
                if (key < bestKey) {
                    bestKey = key;
                    bu = u;
                    bv = v;
                    bw = ww;
                }
            }
            const double best = (double)__uint_as_float((uint32_t)(bestKey >> 32));
            double bl = 0, bh = 0;
            if (lo >= 1) {
                bl = row_bound(lo);
                if (bl > best) lo = 0;
            }
            if (hi <= kIn) {
                bh = row_bound(hi);
                if (bh > best) hi = kIn + 1;
            }
            const bool okl = lo >= 1, okh = hi <= kIn;
            cur = (okl && (!okh || bl <= bh)) ? lo-- : okh ? hi++ : -1;
        }
    }
    const bool found = bestKey != start;
    if (found) {
        p[0] = bu;
        p[1] = bv;
        p[2] = bw;
    }
    return found;
}

// Projection frame of contour t (cpp_dmc/DualMarchingCubes.cpp:141-156 and
// :194-272): p is the mean of the contour's edge intersections in local cell
// coordinates and pt the same point in the frame (a, b, c) of the chosen
// projection axis.
struct ContourFrame {
    int axa, axb, axc;
    uint32_t cpk;
};

__device__ inline ContourFrame contour_frame(const CornerRef& Fc, double iso, uint64_t w, int t, const BoxRef& B,
                                             double p[3], double pt[3]) {
    // mean of the edge intersections, summed in contour order
    const int start = poly_start(w, t), size = poly_size(w, t);
    p[0] = 0;
    p[1] = 0;
    p[2] = 0;
    for (int m = 0; m < size; m++) {
        const int e = poly_edge_at(w, start + m);
        double ui[3];
        local_coords(e, edge_param(e, Fc, iso), ui);
        p[0] += ui[0];
        p[1] += ui[1];
        p[2] += ui[2];
    }
    p[0] /= size;
    p[1] /= size;
    p[2] /= size;
    // projection axis: W = 1, V = 2, U = 3
    int prj;
    if (size >= 6) {
        double g[3];
        trilinear_gradient(Fc, p[0], p[1], p[2], g);
        const double dx = fabs(g[0]), dy = fabs(g[1]), dz = fabs(g[2]);
        if (dz >= dx && dz >= dy) prj = 1;
        else if (dy >= dx && dy >= dz) prj = 2;
        else if (dx >= dy && dx >= dz) prj = 3;
        else prj = 1;
    } else {
        const double eu = (double)B.hi(0, t) - (double)B.lo(0, t);
        const double ev = (double)B.hi(1, t) - (double)B.lo(1, t);
        const double ew = (double)B.hi(2, t) - (double)B.lo(2, t);
        if (ew <= eu && ew <= ev) prj = 1;
        else if (ev <= eu && ev <= ew) prj = 2;
        else if (eu <= ev && eu <= ew) prj = 3;
        else prj = 1;
    }
    // frame (a, b, c): W = (x, y, z), V = (x, z, y), U = (y, z, x)
    ContourFrame f;
    f.axa = prj == 3 ? 1 : 0;
    f.axb = prj == 1 ? 1 : 2;
    f.axc = prj == 1 ? 2 : prj == 2 ? 1 : 0;
    f.cpk = 0;
#pragma unroll
    for (int q = 0; q < 8; q++)
        f.cpk |= (uint32_t)(((q & 1) << f.axa) | (((q >> 1) & 1) << f.axb) | (((q >> 2) & 1) << f.axc)) << (3 * q);
    pt[0] = f.axa == 0 ? p[0] : p[1];
    pt[1] = f.axb == 1 ? p[1] : p[2];
    pt[2] = f.axc == 2 ? p[2] : f.axc == 1 ? p[1] : p[0];
    return f;
}

// Back from the frame to cell coordinates.
__device__ __forceinline__ void contour_unframe(const ContourFrame& f, const double pt[3], double p[3]) {
    double q0 = p[0], q1 = p[1], q2 = p[2];
    if (f.axa == 0) q0 = pt[0]; else q1 = pt[0];
    if (f.axb == 1) q1 = pt[1]; else q2 = pt[1];
    if (f.axc == 2) q2 = pt[2]; else if (f.axc == 1) q1 = pt[2]; else q0 = pt[2];
    p[0] = q0;
    p[1] = q1;
    p[2] = q2;
}

// A cell whose contours' samples must avoid the box of another contour: it
// has several contours, at least one of size 3 and at least one of another
// size (DualMarchingCubes.h:500-519 tests the boxes of the other size-3
// contours when the contour itself is not of size 3).
__device__ __forceinline__ bool cell_needs_box_test(uint64_t w, int cnt) {
    if (cnt < 2) return false;
    bool three = false, other = false;
    for (int q = 0; q < cnt; q++) {
        if (poly_size(w, q) == 3) three = true;
        else other = true;
    }
    return three && other;
}

// Representative of contour t of a cell without the box test, from the 9 x 9
// sampling; false when that sampling finds no point (the caller then takes
// the general path, which also samples at 28 x 28).
template <int kUnroll>
__device__ inline bool contour_representative9(const CornerRef& Fc, double iso, uint64_t w, int t, const BoxRef& B,
                                               double p[3]) {
    double pt[3];
    const ContourFrame f = contour_frame(Fc, iso, w, t, B, p, pt);
    if (!projection9<kUnroll>(Fc, f.cpk, B, f.axa, f.axb, f.axc, iso, t, pt)) return false;
    contour_unframe(f, pt, p);
    return true;
}

// Representative of contour t in local cell coordinates
// (cpp_dmc/DualMarchingCubes.cpp:141-156 and :194-272), general form.
// Returns the number of failed projection levels (0, 1 or 2).
__device__ inline int contour_representative(const CornerRef& Fc, double iso, uint64_t w, int cnt, int t,
                                             const BoxRef& B, double p[3]) {
    double pt[3];
    const ContourFrame f = contour_frame(Fc, iso, w, t, B, p, pt);
    // size-3 contours of the cell other than t whose boxes samples must avoid
    unsigned nbmask = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) nbmask |= (q < cnt && q != t && poly_size(w, q) == 3) ? (1u << q) : 0u;
    if (cnt == 1 || poly_size(w, t) == 3) nbmask = 0;
    int fails = 0;
    if (!projection(Fc, f.cpk, B, f.axa, f.axb, f.axc, iso, t, nbmask, pt, 9)) {
        fails = 1;
        if (!projection(Fc, f.cpk, B, f.axa, f.axb, f.axc, iso, t, nbmask, pt, 28)) fails = 2;
    }
    contour_unframe(f, pt, p);
    return fails;
}

// Bounding boxes of all contours (cpp_dmc/DualMarchingCubes.cpp:113-139).
__device__ inline void cell_boxes(const CornerRef& Fc, double iso, uint64_t w, int cnt, const BoxRef& B) {
    for (int t = 0; t < cnt; t++) {
        const int start = poly_start(w, t), size = poly_size(w, t);
        float lo[3] = {1.f, 1.f, 1.f}, hi[3] = {0.f, 0.f, 0.f};
        for (int m = 0; m < size; m++) {
            const int e = poly_edge_at(w, start + m);
            double ui[3];
            local_coords(e, edge_param(e, Fc, iso), ui);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float x = (float)ui[c];
                lo[c] = lo[c] > x ? x : lo[c];
                hi[c] = hi[c] < x ? x : hi[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) B.set(c, t, lo[c], hi[c]);
    }
}

}  // namespace tmc
