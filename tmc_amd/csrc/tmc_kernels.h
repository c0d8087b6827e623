// tmc_kernels.h -- topologically correct Marching Cubes on the GPU
// (mc/MarchingCubes.cpp: t_mc :150-318, p_slice :322-1303; the CUDA
// counterpart is cuda_tmc/MarchingCubes.cu t_slice :1366).
//
// The active cells come from the DMC sweep (same classification, f >= iso,
// same per-(row, segment) lists, whose flat order is the reference's
// (k, j, i) cell order).  Then:
//   k_tmc_cells      one wave per row: each active cell's case, new vertices
//                    (the cut edges it owns) and table triangles; the
//                    ambiguous cells listed;
//   k_tmc_amb_count  the listed cells' inner vertices and triangles from the
//                    decider slicing (kept out of the row walk, where one
//                    ambiguous lane would hold up its wave);
//   scans            over the active cells: each cell's first vertex and
//                    triangle -- the reference's numbering, first sight in
//                    cell order;
//   k_tmc_emit       one thread per non-ambiguous active cell, and
//   k_tmc_emit_amb   one per listed cell: positions and normals of the cell's
//                    new vertices, and its triangles with the vertex ids of
//                    cut edges owned by earlier cells.
// A grid edge is owned by the first cell containing it in (k, j, i) order;
// that cell is active because the edge is cut, so its active index is found
// from the sweep's activity words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmc_extract_kernels.h"
#include "tmc_mc_tables.h"

namespace tmc {

__device__ __constant__ uint64_t c_tmc_tri[256] = TMC_CASE_TRIANGLES_INIT;
__device__ __constant__ uint64_t c_tmc_amb[4] = TMC_AMBIGUOUS_MASK_INIT;

// Edge e: end corners (corner c = i + 2 j + 4 k) packed as v0 | v1 << 4;
// its first grid point relative to the cell (bits 0..2: x, y, z) and axis
// (bits 3..4), the reference's key layout (mc/MarchingCubes.cpp:153-159).
__device__ __constant__ uint8_t c_tmc_ev[12] = {0x10, 0x31, 0x32, 0x20, 0x54, 0x75, 0x76, 0x64, 0x40, 0x51, 0x73, 0x62};
__device__ __constant__ uint8_t c_tmc_ekey[12] = {0, 1 | 8, 2, 8, 4, 1 | 4 | 8, 2 | 4, 4 | 8, 16, 1 | 16, 1 | 2 | 16, 2 | 16};
// faces seen from outside: corners and edges in oriented order
// (mc/MarchingCubes.h:84-107), one nibble each
__device__ __constant__ uint16_t c_tmc_fv[6] = {0x3120, 0x6475, 0x1504, 0x7362, 0x2064, 0x7531};
__device__ __constant__ uint16_t c_tmc_fe[6] = {0x0123, 0x4765, 0x4908, 0x2A6B, 0x83B7, 0x95A1};
// edge of each axis (x, y, z) by the offset of its first point from a cell
// (bits: x, y, z of the offset)
__device__ __constant__ int8_t c_tmc_edge_at[3][8] = {{0, -1, 2, -1, 4, -1, 6, -1},
                                                      {3, 1, -1, -1, 7, 5, -1, -1},
                                                      {8, 9, 11, 10, -1, -1, -1, -1}};
// local coordinates of the edge points per axis: 2 bits per edge, 3 = along
// the edge (mc/MarchingCubes.cpp:896-903)
constexpr unsigned kTmcLCoord[3] = {1324855u, 5299420u, 16733440u};

__device__ __forceinline__ bool tmc_ambiguous(unsigned icase) { return (c_tmc_amb[icase >> 6] >> (icase & 63)) & 1ull; }

__device__ __forceinline__ unsigned tmc_cut_mask(unsigned icase) {
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const unsigned a = c_tmc_ev[e] & 0xF, b = c_tmc_ev[e] >> 4;
        m |= (((icase >> a) ^ (icase >> b)) & 1u) << e;
    }
    return m;
}

__device__ __forceinline__ bool cell_exists(int x, int y, int z, const Dims& d) {
    return x >= 0 && y >= 0 && z >= 0 && x <= d.nx - 2 && y <= d.ny - 2 && z <= d.nz - 2;
}

// The first cell in (k, j, i) order that contains edge e of cell (i, j, k),
// and the edge's id in that cell.  The candidates are the four cells around
// the edge, earliest first.
__device__ __forceinline__ void tmc_edge_owner(int i, int j, int k, int e, const Dims& d, int& oi, int& oj, int& ok,
                                               int& oe) {
    const unsigned key = c_tmc_ekey[e];
    const int gx = i + (key & 1), gy = j + ((key >> 1) & 1), gz = k + ((key >> 2) & 1);
    const int axis = key >> 3;
    // offsets (s, t) of the candidates along the two other axes, earliest first
    const int sa[4] = {1, 0, 1, 0}, ta[4] = {1, 1, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int cx = gx, cy = gy, cz = gz;
        if (axis == 0) { cy -= sa[c]; cz -= ta[c]; }
        else if (axis == 1) { cx -= sa[c]; cz -= ta[c]; }
        else { cx -= sa[c]; cy -= ta[c]; }
        if (cell_exists(cx, cy, cz, d)) {
            oi = cx;
            oj = cy;
            ok = cz;
            oe = c_tmc_edge_at[axis][(gx - cx) | (gy - cy) << 1 | (gz - cz) << 2];
            return;
        }
    }
    oi = i;
    oj = j;
    ok = k;
    oe = e;
}

// Edges of cell (i, j, k) that it owns (it is their first cell).
__device__ __forceinline__ unsigned tmc_owned_mask(int i, int j, int k, const Dims& d) {
    if (i >= 1 && j >= 1 && k >= 1) return (1u << 5) | (1u << 6) | (1u << 10);
    unsigned m = 0;
    for (int e = 0; e < 12; e++) {
        int oi, oj, ok, oe;
        tmc_edge_owner(i, j, k, e, d, oi, oj, ok, oe);
        if (oi == i && oj == j && ok == k) m |= 1u << e;
    }
    return m;
}

__device__ __forceinline__ double tmc_decider(double f0, double f1, double f2, double f3) {
    return (f0 * f3 - f1 * f2) / (f0 + f3 - f1 - f2);
}

// p_slice: contours of an ambiguous cell from the face deciders, then its
// triangles (local refs: 0..11 edge vertices, 12..17 inner vertices) and
// inner vertices (local coordinates), handed to `em` in the reference's
// order.  Returns false where a singular face cannot be triangulated: the
// reference then leaves the cell without triangles.
template <class Emit>
__device__ bool tmc_slice(const double F[8], unsigned icase, double i0, Emit& em) {
    double ecoord[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const unsigned a = c_tmc_ev[e] & 0xF, b = c_tmc_ev[e] >> 4;
        ecoord[e] = (((icase >> a) ^ (icase >> b)) & 1u) ? (i0 - F[a]) / (F[b] - F[a]) : 0.0;
    }
    int8_t to[12];
#pragma unroll
    for (int e = 0; e < 12; e++) to[e] = -1;
    bool f_flag[6];
    for (int f = 0; f < 6; f++) {
        f_flag[f] = false;
        const unsigned fv = c_tmc_fv[f], fe = c_tmc_fe[f];
        const double f0 = F[fv & 0xF], f1 = F[(fv >> 4) & 0xF], f2 = F[(fv >> 8) & 0xF], f3 = F[fv >> 12];
        const int e0 = fe & 0xF, e1 = (fe >> 4) & 0xF, e2 = (fe >> 8) & 0xF, e3 = fe >> 12;
        const int fc = (f0 >= i0) | (f1 >= i0) << 1 | (f2 >= i0) << 2 | (f3 >= i0) << 3;
        switch (fc) {
        case 1: to[e0] = e3; break;
        case 2: to[e1] = e0; break;
        case 3: to[e1] = e3; break;
        case 4: to[e3] = e2; break;
        case 5: to[e0] = e2; break;
        case 6:
        case 9: {
            const double val = tmc_decider(f0, f1, f2, f3);
            bool sep;   // the pairing of val > i0
            if (val > i0) sep = true;
            else if (val < i0) sep = false;
            else {
                f_flag[f] = true;
                const unsigned e_flag = 0x218;
                double c0 = ecoord[e0], c1 = ecoord[e1], c2 = ecoord[e2], c3 = ecoord[e3];
                if ((e_flag >> (f * 2)) & 1u) { c0 = 1 - c0; c2 = 1 - c2; }
                if ((e_flag >> (f * 2)) & 2u) { c1 = 1 - c1; c3 = 1 - c3; }
                if (c1 < c3 && c0 > c2) sep = fc != 6;
                else if (c1 > c3 && c0 < c2) sep = fc == 6;
                else return false;
            }
            if (fc == 6) {
                if (sep) { to[e3] = e0; to[e1] = e2; }
                else { to[e1] = e0; to[e3] = e2; }
            } else {
                if (sep) { to[e0] = e1; to[e2] = e3; }
                else { to[e0] = e3; to[e2] = e1; }
            }
        } break;
        case 7: to[e1] = e2; break;
        case 8: to[e2] = e1; break;
        case 10: to[e2] = e0; break;
        case 11: to[e2] = e3; break;
        case 12: to[e3] = e1; break;
        case 13: to[e0] = e1; break;
        case 14: to[e3] = e0; break;
        default: break;
        }
    }
    // closed contours, each started at its smallest edge: edges packed 4
    // bits each in walk order, contour c from bit 4 * cst[c]
    int cnt = 0;
    int csz[4] = {0, 0, 0, 0}, cst[4] = {0, 0, 0, 0};
    uint64_t cedge = 0;
    int used_slots = 0;
    for (int e = 0; e < 12 && cnt < 4; e++) {
        if (to[e] < 0) continue;
        cst[cnt] = used_slots;
        int pos = 0, t = to[e];
        cedge |= (uint64_t)e << (4 * (used_slots + pos));
        // every cut edge has one outgoing and one incoming segment, so the
        // walk closes; the bounds only keep a malformed case in range
        while (t != e && t >= 0 && used_slots + pos < 11) {
            pos++;
            cedge |= (uint64_t)t << (4 * (used_slots + pos));
            const int nx = to[t];
            to[t] = -1;
            t = nx;
        }
        csz[cnt] = pos + 1;
        used_slots += pos + 1;
        cnt++;
    }
    auto ce = [&](int c, int r) { return (int)((cedge >> (4 * (cst[c] + r))) & 0xF); };

    double ui[2] = {0, 0}, vi[2] = {0, 0}, wi[2] = {0, 0};
    unsigned q_sol = 0;
    const double a = (F[0] - F[1]) * (-F[6] + F[7] + F[4] - F[5]) - (F[4] - F[5]) * (-F[2] + F[3] + F[0] - F[1]);
    const double b = (i0 - F[0]) * (-F[6] + F[7] + F[4] - F[5]) + (F[0] - F[1]) * (F[6] - F[4]) -
                     (i0 - F[4]) * (-F[2] + F[3] + F[0] - F[1]) - (F[4] - F[5]) * (F[2] - F[0]);
    const double c = (i0 - F[0]) * (F[6] - F[4]) - (i0 - F[4]) * (F[2] - F[0]);
    double disc = b * b - 4 * a * c;
    if (disc > 0) {
        disc = sqrt(disc);
        ui[0] = (-b - disc) / (2 * a);
        ui[1] = (-b + disc) / (2 * a);
        for (int s = 0; s < 2; s++) {
            const double g1 = F[0] * (1 - ui[s]) + F[1] * ui[s];
            const double g2 = F[2] * (1 - ui[s]) + F[3] * ui[s];
            vi[s] = (i0 - g1) / (g2 - g1);
            if (isnan(vi[s]) || isinf(vi[s])) vi[s] = -1.0;
        }
        for (int s = 0; s < 2; s++) {
            const double g1 = F[0] * (1 - ui[s]) + F[1] * ui[s];
            const double g2 = F[4] * (1 - ui[s]) + F[5] * ui[s];
            wi[s] = (i0 - g1) / (g2 - g1);
            if (isnan(wi[s]) || isinf(wi[s])) wi[s] = -1.0;
        }
        if (f_flag[0]) { if (wi[0] < wi[1]) wi[0] = 0; else wi[1] = 0; }
        if (f_flag[1]) wi[1] = 1;   // both branches of the reference set wi[1]
        if (f_flag[2]) { if (vi[0] < vi[1]) vi[0] = 0; else vi[1] = 0; }
        if (f_flag[3]) { if (vi[0] > vi[1]) vi[0] = 1; else vi[1] = 1; }
        if (f_flag[4]) { if (ui[0] < ui[1]) ui[0] = 0; else ui[1] = 0; }
        if (f_flag[5]) { if (ui[0] > ui[1]) ui[0] = 1; else ui[1] = 1; }
        q_sol |= (0 < ui[0] && ui[0] < 1) ? 1u : 0u;
        q_sol |= (0 < ui[1] && ui[1] < 1) ? 2u : 0u;
        q_sol |= (0 < vi[0] && vi[0] < 1) ? 4u : 0u;
        q_sol |= (0 < vi[1] && vi[1] < 1) ? 8u : 0u;
        q_sol |= (0 < wi[0] && wi[0] < 1) ? 16u : 0u;
        q_sol |= (0 < wi[1] && wi[1] < 1) ? 32u : 0u;
    }
    auto ev = [&](int e, int x) {
        const unsigned fl = (kTmcLCoord[x] >> (2 * e)) & 3u;
        return fl == 3u ? ecoord[e] : (double)fl;
    };
    if (__popc(q_sol) == 6) {
        // tunnel, or one contour with 12 vertices: the inner hexagon
        const double hu[6] = {ui[0], ui[0], ui[1], ui[1], ui[1], ui[0]};
        const double hv[6] = {vi[0], vi[0], vi[0], vi[1], vi[1], vi[1]};
        const double hw[6] = {wi[0], wi[1], wi[1], wi[1], wi[0], wi[0]};
        int not_tunnel = 0xF;
        if (cnt == 3) {
            const double uc_min = (ui[0] < ui[1]) ? ui[0] : ui[1];
            const double uc_max = (ui[0] < ui[1]) ? ui[1] : ui[0];
            for (int t = 0; t < cnt; t++) {
                if (csz[t] != 3) continue;
                double umin = 2, umax = -2;
                const int x0 = ce(t, 0), x1 = ce(t, 1), x2 = ce(t, 2);
                const double u0 = ev(x0, 0), u1 = ev(x1, 0), u2 = ev(x2, 0);
                umin = (u0 < umin) ? u0 : umin;
                umin = (u1 < umin) ? u1 : umin;
                umin = (u2 < umin) ? u2 : umin;
                umax = (u0 > umax) ? u0 : umax;
                umax = (u1 > umax) ? u1 : umax;
                umax = (u2 > umax) ? u1 : umax;   // the reference takes u1 here
                if (uc_min > umax || uc_max < umin) {
                    not_tunnel = t;
                    em.tri(x0, x1, x2);
                }
            }
        }
        for (int s = 0; s < 6; s++) em.inner(s, hu[s], hv[s], hw[s]);
        uint64_t tcon = 0;   // nearest hexagon vertex of each contour edge, 4 bits per edge
        for (int t = 0; t < cnt; t++) {
            if (not_tunnel == t) continue;
            const int sz = csz[t];
            for (int r = 0; r < sz; r++) {
                unsigned index = 0xFu;
                double dist = 1000.;
                const int ci = ce(t, r);
                const double ue = ev(ci, 0), ve = ev(ci, 1), we = ev(ci, 2);
                for (int s = 0; s < 6; s++) {
                    const double du = ue - hu[s], dv = ve - hv[s], dw = we - hw[s];
                    const double val = du * du + dv * dv + dw * dw;
                    if (dist > val) {
                        index = (unsigned)s;
                        dist = val;
                    }
                }
                tcon = (tcon & ~(0xFull << (4 * ci))) | ((uint64_t)index << (4 * ci));
            }
            for (int r = 0; r < sz; r++) {
                const int t1 = ce(t, r), t2 = ce(t, (r + 1) % sz);
                const int c1 = (int)((tcon >> (4 * t1)) & 0xF), c2 = (int)((tcon >> (4 * t2)) & 0xF);
                if (c1 > 5 || c2 > 5) continue;   // no nearest vertex (not reached with finite coordinates)
                int dst = c1 - c2 < 0 ? c2 - c1 : c1 - c2;
                if (dst > 2) dst = 6 - dst;
                if (dst == 0) {
                    em.tri(t1, t2, 12 + c1);
                } else if (dst == 1) {
                    const double l1 = (ev(t1, 0) - hu[c2]) * (ev(t1, 0) - hu[c2]) + (ev(t1, 1) - hv[c2]) * (ev(t1, 1) - hv[c2]) +
                                      (ev(t1, 2) - hw[c2]) * (ev(t1, 2) - hw[c2]);
                    const double l2 = (ev(t2, 0) - hu[c1]) * (ev(t2, 0) - hu[c1]) + (ev(t2, 1) - hv[c1]) * (ev(t2, 1) - hv[c1]) +
                                      (ev(t2, 2) - hw[c1]) * (ev(t2, 2) - hw[c1]);
                    if (l1 < l2) {
                        em.tri(t1, t2, 12 + c2);
                        em.tri(t1, 12 + c2, 12 + c1);
                    } else {
                        em.tri(t1, t2, 12 + c1);
                        em.tri(t2, 12 + c2, 12 + c1);
                    }
                } else if (dst == 2) {
                    const int dmax = c1 > c2 ? c1 : c2, dmin = c1 < c2 ? c1 : c2;
                    const int cm = ((dmax + 2) % 6 == dmin) ? (dmax + 1) % 6 : (dmax + dmin) / 2;
                    em.tri(t1, t2, 12 + cm);
                    em.tri(t1, 12 + cm, 12 + c1);
                    em.tri(t2, 12 + c2, 12 + cm);
                }
            }
        }
        if (cnt == 1) {
            const bool s_ = tmc_decider(F[0], F[1], F[2], F[3]) <= i0;
            const bool of_ = (wi[1] < wi[0]) ? s_ : !s_;
            if (!of_) {
                em.tri(12, 14, 13);
                em.tri(14, 16, 15);
                em.tri(12, 17, 16);
                em.tri(12, 16, 14);
            } else {
                em.tri(12, 13, 14);
                em.tri(14, 15, 16);
                em.tri(12, 16, 17);
                em.tri(12, 14, 16);
            }
        }
        return true;
    }
    const unsigned nr_u = ((q_sol & 0x5u) == 0x5u) + ((q_sol & 0xAu) == 0xAu);
    const unsigned nr_v = ((q_sol & 0x11u) == 0x11u) + ((q_sol & 0x22u) == 0x22u);
    const unsigned nr_w = ((q_sol & 0x18u) == 0x18u) + ((q_sol & 0x24u) == 0x24u);
    const unsigned nr_t = nr_u + nr_v + nr_w;
    if (nr_t == nr_u || nr_t == nr_v || nr_t == nr_w) {
        // no saddle point inside: fans of the contours of 3 to 6 edges
        for (int t = 0; t < cnt; t++) {
            const int c0 = ce(t, 0), c1 = ce(t, 1), c2 = ce(t, 2);
            switch (csz[t]) {
            case 3: em.tri(c0, c1, c2); break;
            case 4:
                em.tri(c0, c1, c2);
                em.tri(c0, c2, ce(t, 3));
                break;
            case 5:
                em.tri(c0, c1, c2);
                em.tri(c0, c2, ce(t, 3));
                em.tri(c0, ce(t, 3), ce(t, 4));
                break;
            case 6:
                em.tri(c0, c1, ce(t, 3));
                em.tri(c1, c2, ce(t, 3));
                em.tri(c0, ce(t, 3), ce(t, 4));
                em.tri(c0, ce(t, 4), ce(t, 5));
                break;
            default: break;
            }
        }
        return true;
    }
    // one inner vertex, joined to the contours of more than three edges
    const unsigned f00 = q_sol & 1u, f01 = (q_sol >> 1) & 1u, f10 = (q_sol >> 2) & 1u, f11 = (q_sol >> 3) & 1u,
                   f20 = (q_sol >> 4) & 1u, f21 = (q_sol >> 5) & 1u;
    const double d00 = f00, d01 = f01, d10 = f10, d11 = f11, d20 = f20, d21 = f21;
    const unsigned fc1 = f00 * f10 + f01 * f11, fc2 = f00 * f20 + f01 * f21, fc3 = f10 * f21 + f11 * f20;
    double uc = 0, vc = 0, wc = 0;
    switch (fc1 + fc2 + fc3) {
    case 2:
        if (fc1 == 0) {
            uc = d00 * ui[0] + d01 * ui[1];
            vc = d10 * vi[0] + d11 * vi[1];
            wc = d10 * wi[1] + d11 * wi[0];
        } else if (fc2 == 0) {
            uc = d00 * ui[0] + d01 * ui[1];
            vc = d00 * vi[0] + d01 * vi[1];
            wc = d00 * wi[1] + d01 * wi[0];
        } else if (fc3 == 0) {
            uc = d10 * ui[0] + d11 * ui[1];
            vc = d10 * vi[0] + d11 * vi[1];
            wc = d10 * wi[0] + d11 * wi[1];
        }
        break;
    case 3:
        uc = (d00 * ui[0] + d01 * ui[1]) / (double)(f00 + f01);
        vc = (d10 * vi[0] + d11 * vi[1]) / (double)(f10 + f11);
        wc = (d20 * wi[0] + d21 * wi[1]) / (double)(f20 + f21);
        break;
    case 4:
        if (f20 + f21 == 1) {
            uc = d20 * ui[0] + d21 * ui[1];
            vc = d21 * vi[0] + d20 * vi[1];
            wc = d20 * wi[0] + d21 * wi[1];
        } else if (f10 + f11 == 1) {
            uc = d10 * ui[0] + d11 * ui[1];
            vc = d10 * vi[0] + d11 * vi[1];
            wc = d11 * wi[0] + d10 * wi[1];
        } else if (f00 + f01 == 1) {
            uc = d00 * ui[0] + d01 * ui[1];
            vc = d00 * vi[0] + d01 * vi[1];
            wc = d00 * wi[0] + d01 * wi[1];
        }
        break;
    default: break;
    }
    bool used = false;
    for (int t = 0; t < cnt; t++) {
        const int sz = csz[t];
        if (sz == 3) {
            em.tri(ce(t, 0), ce(t, 1), ce(t, 2));
        } else {
            used = true;
            for (int r = 0; r < sz; r++) em.tri(ce(t, r), ce(t, (r + 1) % sz), 12);
        }
    }
    if (used) em.inner(0, uc, vc, wc);
    return true;
}

// Counting emitter: triangles and inner vertices of a cell.
struct TmcCount {
    int nt = 0, ninner = 0;
    __device__ void tri(int, int, int) { nt++; }
    __device__ void inner(int s, double, double, double) { ninner = s + 1 > ninner ? s + 1 : ninner; }
};

// Corner values of cell (i, j, k) as doubles.
template <class T>
__device__ __forceinline__ void tmc_corners(const T* __restrict__ field, const Dims& d, int i, int j, int k,
                                            double F[8]) {
    const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
    const T* p = field + (int64_t)k * sz + (int64_t)j * sy + i;
#pragma unroll
    for (int c = 0; c < 8; c++) F[c] = (double)p[(c & 1) + ((c >> 1) & 1) * sy + (c >> 2) * sz];
}

// K_T1: one wave per row, cells in ascending x (the flat active order):
// per active cell its cell word, case (acode), new vertices (cnt_v: the cut
// edges it owns) and triangles (cnt_t: the case table's); the ambiguous
// cells are listed (amb[0] = count, from amb[1]) for k_tmc_amb_count, which
// adds their inner vertices and sets their triangles.  stats[0] counts
// them.
__global__ void __launch_bounds__(kBlock) k_tmc_cells(VolTable vt, int64_t Rtot, const uint32_t* __restrict__ lists,
                                                       const uint32_t* __restrict__ segCount,
                                                       const uint32_t* __restrict__ segBase,
                                                       uint32_t* __restrict__ acell, uint32_t* __restrict__ acode,
                                                       uint32_t* __restrict__ cnt_v, uint32_t* __restrict__ cnt_t,
                                                       uint32_t* __restrict__ amb,
                                                       unsigned long long* __restrict__ stats) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wpb = kBlock / kWave;
    unsigned long long namb = 0;
    for (int64_t g = (int64_t)blockIdx.x * wpb + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave); g < Rtot;
         g += (int64_t)gridDim.x * wpb) {
        const int v = vol_of(vt.rowBase, vt.n, g);
        const VolDesc& V = vt.desc[v];
        const Dims d = V.d;
        const int S = V.S;
        const int64_t r = g - V.rowBase;
        const int j = (int)(r % (d.ny - 1));
        const int k = (int)(r / (d.ny - 1));
        const int64_t rs0 = V.rsBase + r * S;
        // the row's segment bases, one per lane (rows of at most 64
        // segments; wider rows search memory)
        const bool regs = S <= kWave;
        const uint32_t mybase = (regs && lane < S) ? segBase[rs0 + lane] : 0u;
        const uint32_t first = regs ? (uint32_t)__builtin_amdgcn_readfirstlane((int)mybase) : segBase[rs0];
        const uint32_t last = segBase[rs0 + S - 1] + segCount[rs0 + S - 1];
        for (uint32_t a0 = first; a0 < last; a0 += kWave) {
            const uint32_t a = a0 + lane;
            const bool act = a < last;
            int seg = 0;
            uint32_t sbase = first;
            if (regs) {
                for (int s2 = 1; s2 < S; s2++) {
                    const uint32_t b = (uint32_t)__shfl((int)mybase, s2, kWave);
                    if (b <= a) {
                        seg = s2;
                        sbase = b;
                    }
                }
            }
            bool isamb = false;
            if (act) {
                if (!regs) {
                    int lo = 0, hi = S - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (segBase[rs0 + mid] <= a) lo = mid;
                        else hi = mid - 1;
                    }
                    seg = lo;
                    sbase = segBase[rs0 + lo];
                }
                const uint32_t x = lists[(rs0 + seg) * kSeg + (a - sbase)];
                const int i = (int)(x & 0xFFFFFFu);
                const unsigned icase = x >> 24;
                isamb = tmc_ambiguous(icase);
                namb += isamb;
                acell[a] = (uint32_t)(r * d.C) + (uint32_t)i;
                acode[a] = icase;
                cnt_v[a] = (uint32_t)__popc(tmc_cut_mask(icase) & tmc_owned_mask(i, j, k, d));
                cnt_t[a] = isamb ? 0u : (uint32_t)(c_tmc_tri[icase] >> 60);
            }
            // ambiguous cells go to the list k_tmc_amb_count and k_tmc_emit_amb
            // work through
            const uint64_t m = __ballot(isamb);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(amb, (uint32_t)__popcll(m));
                base = (uint32_t)__shfl((int)base, leader, kWave);
                const uint32_t below =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (isamb) amb[1 + base + below] = a;
            }
        }
    }
    namb = wave_sum(namb);
    if (lane == 0 && namb) atomicAdd(&stats[0], namb);
}

// K_T2: the listed ambiguous cells' inner vertices (added to cnt_v) and
// triangles (cnt_t) from the decider slicing; acode bit 8 marks a cell
// whose singular face could not be triangulated (no triangles), counted in
// stats[1].
template <class T>
__global__ void __launch_bounds__(kBlock) k_tmc_amb_count(VolTable vt, const uint32_t* __restrict__ amb,
                                                           const uint32_t* __restrict__ acell,
                                                           uint32_t* __restrict__ acode, uint32_t* __restrict__ cnt_v,
                                                           uint32_t* __restrict__ cnt_t,
                                                           unsigned long long* __restrict__ stats) {
    const uint32_t n = amb[0];
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        const uint32_t a = amb[1 + s];
        const int vol = vol_of(vt.aBase, vt.n, a);
        const VolDesc& V = vt.desc[vol];
        int i, j, k;
        uint32_t r;
        cell_ijk(acell[a], V.d, i, j, k, r);
        const unsigned icase = acode[a] & 0xFFu;
        double F[8];
        tmc_corners(static_cast<const T*>(V.field), V.d, i, j, k, F);
        TmcCount c;
        if (tmc_slice(F, icase, V.iso, c)) {
            cnt_v[a] += (uint32_t)c.ninner;
            cnt_t[a] = (uint32_t)c.nt;
        } else {
            acode[a] = icase | 0x100u;
            atomicAdd(&stats[1], 1ull);
        }
    }
}

// Gradient at corner c of cell (i, j, k): central differences inside the
// grid, one-sided at its border (mc/MarchingCubes.cpp:1450-1512).  The
// neighbours inside the cell come from its corner values F (the same
// doubles); only the three outside it are loaded.
template <class T>
__device__ __forceinline__ void tmc_corner_gradient(const T* __restrict__ field, const Dims& d, const GridGeom& geo,
                                                    const double F[8], int i, int j, int k, int c, double n[3]) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
    const int x = i + cx, y = j + cy, z = k + cz;
    const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
    const T* p = field + (int64_t)z * sz + (int64_t)y * sy + x;
    // x == 0 only at cx == 0, x == nx - 1 only at cx == 1 (likewise y, z)
    if (x == 0) n[0] = (F[c | 1] - F[c]) / geo.dx;
    else if (x == d.nx - 1) n[0] = (F[c] - F[c & ~1]) / geo.dx;
    else n[0] = (cx ? (double)p[1] - F[c & ~1] : F[c | 1] - (double)p[-1]) / (2 * geo.dx);
    if (y == 0) n[1] = (F[c | 2] - F[c]) / geo.dy;
    else if (y == d.ny - 1) n[1] = (F[c] - F[c & ~2]) / geo.dy;
    else n[1] = (cy ? (double)p[sy] - F[c & ~2] : F[c | 2] - (double)p[-sy]) / (2 * geo.dy);
    if (z == 0) n[2] = (F[c | 4] - F[c]) / geo.dz;
    else if (z == d.nz - 1) n[2] = (F[c] - F[c & ~4]) / geo.dz;
    else n[2] = (cz ? (double)p[sz] - F[c & ~4] : F[c | 4] - (double)p[-sz]) / (2 * geo.dz);
}

// After the scans: per active cell its owned cut edges (x) and first vertex
// id (y, the exclusive scan of cnt_v), one 8-byte word for the owner
// lookups of k_tmc_emit.
__global__ void __launch_bounds__(kBlock) k_tmc_vbase(VolTable vt, const uint32_t* __restrict__ acell,
                                                       const uint32_t* __restrict__ acode,
                                                       const uint32_t* __restrict__ vofs, int64_t A,
                                                       uint2* __restrict__ avb) {
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        const int vol = vol_of(vt.aBase, vt.n, a);
        const VolDesc& V = vt.desc[vol];
        int i, j, k;
        uint32_t r;
        cell_ijk(acell[a], V.d, i, j, k, r);
        const unsigned own = tmc_cut_mask(acode[a] & 0xFFu) & tmc_owned_mask(i, j, k, V.d);
        avb[a] = make_uint2(own, vofs[a]);
    }
}

// Writing emitter: triangles with global vertex ids (edge refs through ev,
// inner refs from ibase) and the inner vertices' positions and normals, the
// trilinear blends of the corners' in the reference's operation order.
struct TmcWrite {
    const double (*P)[3];
    const double (*N)[3];
    const int32_t* ev;
    int32_t ibase;
    double* vpos;
    double* vnrm;
    int32_t* tris;
    int64_t t;
    __device__ int32_t id(int r) const { return r < 12 ? ev[r] : ibase + (r - 12); }
    __device__ void tri(int a, int b, int c) {
        int32_t* o = tris + 3 * t;
        o[0] = id(a);
        o[1] = id(b);
        o[2] = id(c);
        t++;
    }
    __device__ static double blend(const double (*q)[3], int x, double u, double v, double w) {
        return (1 - w) * ((1 - v) * (q[0][x] + u * (q[1][x] - q[0][x])) + v * (q[2][x] + u * (q[3][x] - q[2][x]))) +
               w * ((1 - v) * (q[4][x] + u * (q[5][x] - q[4][x])) + v * (q[6][x] + u * (q[7][x] - q[6][x])));
    }
    __device__ void inner(int s, double u, double v, double w) {
        double* p = vpos + 3 * (int64_t)(ibase + s);
        double* n = vnrm + 3 * (int64_t)(ibase + s);
        double q[3];
#pragma unroll
        for (int x = 0; x < 3; x++) p[x] = blend(P, x, u, v, w);
#pragma unroll
        for (int x = 0; x < 3; x++) q[x] = blend(N, x, u, v, w);
        const double f = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        n[0] = q[0] / f;
        n[1] = q[1] / f;
        n[2] = q[2] / f;
    }
};

// Vertices and triangles of active cell a.  AMB: the cell is ambiguous
// (sliced with the decider, inner vertices); else the case table.  Owned
// edge vertices take the gradients of their two corners only.
template <class T, bool AMB>
__device__ __forceinline__ void tmc_emit_cell(const VolTable& vt, int64_t a, const uint32_t* __restrict__ acell,
                                              const uint32_t* __restrict__ acode,
                                              const uint2* __restrict__ avb,
                                              const uint32_t* __restrict__ tofs,
                                              const unsigned long long* __restrict__ segMask,
                                              const uint32_t* __restrict__ segBase, double* __restrict__ vpos,
                                              double* __restrict__ vnrm, int32_t* __restrict__ tris,
                                              unsigned long long* __restrict__ unresolved) {
    const int vol = vol_of(vt.aBase, vt.n, a);
    const VolDesc& V = vt.desc[vol];
    const Dims d = V.d;
    const GridGeom geo = V.geo;
    const T* __restrict__ field = static_cast<const T*>(V.field);
    const SegLists L{segMask, segBase, V.S, V.rsBase};
    int i, j, k;
    uint32_t r;
    cell_ijk(acell[a], d, i, j, k, r);
    const uint32_t code = acode[a];
    const unsigned icase = code & 0xFFu;
    const unsigned cut = tmc_cut_mask(icase);
    const uint2 self = avb[a];
    const unsigned own = self.x;
    const double i0 = V.iso;
    const int32_t vbase = (int32_t)self.y;
    double F[8];
    tmc_corners(field, d, i, j, k, F);
    auto corner_pos = [&](int c, double p[3]) {
        p[0] = geo.x0 + (i + (c & 1)) * geo.dx;
        p[1] = geo.y0 + (j + ((c >> 1) & 1)) * geo.dy;
        p[2] = geo.z0 + (k + (c >> 2)) * geo.dz;
    };
    int32_t ev[12];
    int nown = 0;
    // active index of each owner cell met, by its offset (i - oi) | (j - oj)
    // << 1 | (k - ok) << 2 (-1: not looked up yet); the cell before in the
    // row is the previous active cell
    int64_t oidx[8];
#pragma unroll
    for (int x = 0; x < 8; x++) oidx[x] = -1;
    oidx[1] = a - 1;
    for (int e = 0; e < 12; e++) {
        ev[e] = -1;
        if (!((cut >> e) & 1u)) continue;
        if ((own >> e) & 1u) {
            ev[e] = vbase + nown++;
            const int v0 = c_tmc_ev[e] & 0xF, v1 = c_tmc_ev[e] >> 4;
            const double l = (i0 - F[v0]) / (F[v1] - F[v0]);
            double P0[3], P1[3], N0[3], N1[3];
            corner_pos(v0, P0);
            corner_pos(v1, P1);
            tmc_corner_gradient(field, d, geo, F, i, j, k, v0, N0);
            tmc_corner_gradient(field, d, geo, F, i, j, k, v1, N1);
            double* p = vpos + 3 * (int64_t)ev[e];
            double* n = vnrm + 3 * (int64_t)ev[e];
            double q[3];
#pragma unroll
            for (int x = 0; x < 3; x++) p[x] = (1 - l) * P0[x] + l * P1[x];
#pragma unroll
            for (int x = 0; x < 3; x++) q[x] = (1 - l) * N0[x] + l * N1[x];
            const double s = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            n[0] = q[0] / s;
            n[1] = q[1] / s;
            n[2] = q[2] / s;
            continue;
        }
        int oi, oj, ok, oe;
        tmc_edge_owner(i, j, k, e, d, oi, oj, ok, oe);
        const int off = (i - oi) | (j - oj) << 1 | (k - ok) << 2;
        int64_t ao = oidx[off];
        if (ao < 0) {
            ao = find_active(L, d, oi, oj, ok);
            oidx[off] = ao;
        }
        if (ao < 0) {
            atomicAdd(unresolved, 1ull);
            continue;
        }
        const uint2 ob = avb[ao];
        ev[e] = (int32_t)(ob.y + (uint32_t)__popc(ob.x & ((1u << oe) - 1u)));
    }
    const int64_t tbase = tofs[a];
    if constexpr (AMB) {
        if (code & 0x100u) return;   // singular face: no triangles
        double P[8][3], N[8][3];
        for (int c = 0; c < 8; c++) {
            corner_pos(c, P[c]);
            tmc_corner_gradient(field, d, geo, F, i, j, k, c, N[c]);
        }
        TmcWrite w{P, N, ev, vbase + nown, vpos, vnrm, tris, tbase};
        tmc_slice(F, icase, i0, w);
    } else {
        const uint64_t wd = c_tmc_tri[icase];
        const int nt = (int)(wd >> 60);
        for (int t = 0; t < nt; t++) {
            const int e0 = (int)((wd >> (12 * t)) & 0xF), e1 = (int)((wd >> (12 * t + 4)) & 0xF),
                      e2 = (int)((wd >> (12 * t + 8)) & 0xF);
            int32_t* o = tris + 3 * (tbase + t);
            o[0] = ev[e2];   // reversed: positive corners outside
            o[1] = ev[e1];
            o[2] = ev[e0];
        }
    }
}

// K_T3: one thread per active cell that is not ambiguous (the case table).
template <class T>
__global__ void __launch_bounds__(kBlock) k_tmc_emit(VolTable vt, const uint32_t* __restrict__ acell,
                                                      const uint32_t* __restrict__ acode,
                                                      const uint2* __restrict__ avb,
                                                      const uint32_t* __restrict__ tofs,
                                                      const unsigned long long* __restrict__ segMask,
                                                      const uint32_t* __restrict__ segBase, int64_t A,
                                                      double* __restrict__ vpos, double* __restrict__ vnrm,
                                                      int32_t* __restrict__ tris,
                                                      unsigned long long* __restrict__ unresolved) {
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        if (tmc_ambiguous(acode[a] & 0xFFu)) continue;
        tmc_emit_cell<T, false>(vt, a, acell, acode, avb, tofs, segMask, segBase, vpos, vnrm, tris, unresolved);
    }
}

// K_T4: the ambiguous cells k_tmc_cells listed (amb[0] = count, cells from
// amb[1]): sliced with the asymptotic decider.
template <class T>
__global__ void __launch_bounds__(kBlock) k_tmc_emit_amb(VolTable vt, const uint32_t* __restrict__ amb,
                                                          const uint32_t* __restrict__ acell,
                                                          const uint32_t* __restrict__ acode,
                                                          const uint2* __restrict__ avb,
                                                          const uint32_t* __restrict__ tofs,
                                                          const unsigned long long* __restrict__ segMask,
                                                          const uint32_t* __restrict__ segBase,
                                                          double* __restrict__ vpos, double* __restrict__ vnrm,
                                                          int32_t* __restrict__ tris,
                                                          unsigned long long* __restrict__ unresolved) {
    const uint32_t n = amb[0];
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock)
        tmc_emit_cell<T, true>(vt, (int64_t)amb[1 + s], acell, acode, avb, tofs, segMask, segBase, vpos, vnrm, tris,
                               unresolved);
}

}  // namespace tmc
