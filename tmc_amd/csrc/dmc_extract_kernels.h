// dmc_extract_kernels.h -- cell-sweep kernels of the DMC engine: classify,
// compact, per-cell polygons, vertex representatives and quadrilaterals.
//
// Data layout in HBM:
//   field      nx*ny*nz scalars, x fastest (T = float or double)
//   cell rows  r = k*(ny-1) + j, C = nx-1 cells per row
//   row lists  rowData[r*C + m]  = i | icase << 24 for the m-th active cell of
//              row r (ascending i); rowCount[r]; rowBase[r] = exclusive scan
//   active     a = rowBase[r] + m enumerates active cells in (k, j, i) order,
//              which is the reference's vertex order
//              (cpp_dmc/DualMarchingCubes.cpp:16-50 with :167-171).
#pragma once
#include "dmc_cell.h"
#include "dmc_kernels.h"

namespace tmc {

struct Dims {
    int nx, ny, nz;
    int64_t C;      // cells per row (nx-1)
    int64_t R;      // rows (ny-1)*(nz-1)
};

template <class T>
__device__ __forceinline__ bool at_or_above(T f, T iso_t) { return f >= iso_t; }

// ---------------------------------------------------------------------------
// K1: classify every cell of a row and write the row's active list.
// One 256-thread block per row, 4 consecutive cells per thread.  For F32
// fields the comparison (double)f >= iso is done as f >= iso_up, iso_up the
// smallest float not below iso, which is the same predicate.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock) k_classify(const T* __restrict__ field, Dims d, T iso_t,
                                                      uint32_t* __restrict__ rowCount,
                                                      uint32_t* __restrict__ rowData) {
    __shared__ uint32_t sm[kBlock / kWave + 1];
    for (int64_t r = blockIdx.x; r < d.R; r += gridDim.x) {
        const int j = (int)(r % (d.ny - 1));
        const int k = (int)(r / (d.ny - 1));
        const int64_t sz = (int64_t)d.nx * d.ny;
        const T* p00 = field + (int64_t)k * sz + (int64_t)j * d.nx;
        const T* p10 = p00 + d.nx;
        const T* p01 = p00 + sz;
        const T* p11 = p01 + d.nx;
        uint32_t* out = rowData + r * d.C;
        uint32_t base = 0;
        for (int i0 = 0; i0 < d.nx - 1; i0 += kBlock * 4) {
            const int i = i0 + threadIdx.x * 4;
            unsigned b00 = 0, b10 = 0, b01 = 0, b11 = 0;
#pragma unroll
            for (int q = 0; q < 5; q++) {
                if (i + q < d.nx) {
                    b00 |= (unsigned)at_or_above(p00[i + q], iso_t) << q;
                    b10 |= (unsigned)at_or_above(p10[i + q], iso_t) << q;
                    b01 |= (unsigned)at_or_above(p01[i + q], iso_t) << q;
                    b11 |= (unsigned)at_or_above(p11[i + q], iso_t) << q;
                }
            }
            unsigned cases[4];
            uint32_t cnt = 0;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const unsigned ic = ((b00 >> s) & 1) | ((b00 >> (s + 1)) & 1) << 1 | ((b10 >> s) & 1) << 2 |
                                    ((b10 >> (s + 1)) & 1) << 3 | ((b01 >> s) & 1) << 4 |
                                    ((b01 >> (s + 1)) & 1) << 5 | ((b11 >> s) & 1) << 6 |
                                    ((b11 >> (s + 1)) & 1) << 7;
                const bool act = (i + s < d.nx - 1) && ic != 0 && ic != 255;
                cases[s] = act ? ic : 256u;
                cnt += act;
            }
            uint32_t tot;
            uint32_t pos = base + block_exclusive_scan<uint32_t, kBlock>(cnt, sm, tot);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if (cases[s] < 256u) out[pos++] = (uint32_t)(i + s) | (cases[s] << 24);
            }
            base += tot;
        }
        if (threadIdx.x == 0) rowCount[r] = base;
    }
}

// ---------------------------------------------------------------------------
// K2: gather the row lists into the flat active-cell arrays.  One wave per row.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_compact_rows(const uint32_t* __restrict__ rowCount,
                                                          const uint32_t* __restrict__ rowBase,
                                                          const uint32_t* __restrict__ rowData, Dims d,
                                                          uint32_t* __restrict__ acell,
                                                          uint8_t* __restrict__ acase) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wpb = kBlock / kWave;
    for (int64_t r = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; r < d.R; r += (int64_t)gridDim.x * wpb) {
        const uint32_t cnt = rowCount[r], base = rowBase[r];
        const uint32_t* src = rowData + r * d.C;
        for (uint32_t m = lane; m < cnt; m += kWave) {
            const uint32_t x = src[m];
            acell[base + m] = (uint32_t)(r * d.C) + (x & 0xFFFFFFu);
            acase[base + m] = (uint8_t)(x >> 24);
        }
    }
}

template <class T>
__device__ __forceinline__ void load_corners(const T* __restrict__ field, const Dims& d, int i, int j, int k,
                                             double F[8]) {
    const int64_t sz = (int64_t)d.nx * d.ny;
    const T* p = field + (int64_t)k * sz + (int64_t)j * d.nx + i;
    F[0] = (double)p[0];
    F[1] = (double)p[1];
    F[2] = (double)p[d.nx];
    F[3] = (double)p[d.nx + 1];
    F[4] = (double)p[sz];
    F[5] = (double)p[sz + 1];
    F[6] = (double)p[sz + d.nx];
    F[7] = (double)p[sz + d.nx + 1];
}

__device__ __forceinline__ void cell_ijk(uint32_t c, const Dims& d, int& i, int& j, int& k) {
    const uint32_t r = c / (uint32_t)d.C;
    i = (int)(c - r * (uint32_t)d.C);
    j = (int)(r % (uint32_t)(d.ny - 1));
    k = (int)(r / (uint32_t)(d.ny - 1));
}

// Used contours: those with at least one interior edge (the others only
// feed incomplete quadrilaterals and are dropped by removeUnusedVertices,
// cpp_dmc/DualMarchingCubes.h:685-729).
__device__ __forceinline__ unsigned used_contours(unsigned icase, uint32_t map, int i, int j, int k, const Dims& d) {
    unsigned used = 0;
#pragma unroll
    for (int e = 0; e < 12; e++)
        if (edge_cut(icase, e) && edge_interior(e, i, j, k, d.nx, d.ny, d.nz)) used |= 1u << ((map >> (2 * e)) & 3u);
    return used;
}

// ---------------------------------------------------------------------------
// K3: polygons of every active cell -> code (icase | edge-to-contour map),
// number of kept vertices and of owned quadrilaterals.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock) k_cell_info(const T* __restrict__ field, Dims d, double iso,
                                                       const uint32_t* __restrict__ acell,
                                                       const uint8_t* __restrict__ acase, int64_t A,
                                                       uint32_t* __restrict__ acode, uint32_t* __restrict__ anv,
                                                       uint32_t* __restrict__ anq,
                                                       unsigned long long* __restrict__ ncontours) {
    unsigned long long total = 0;
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        int i, j, k;
        cell_ijk(acell[a], d, i, j, k);
        const unsigned icase = acase[a];
        uint64_t w;
        if (case_ambiguous(icase)) {
            double F[8];
            load_corners(field, d, i, j, k, F);
            w = polygons_decider(iso, F);
        } else {
            w = c_case_contours[icase];
        }
        uint32_t map = 0;
        int pos = 0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int s = poly_size(w, t);
            for (int m = 0; m < s; m++) {
                const int e = poly_edge_at(w, pos + m);
                map |= (uint32_t)t << (2 * e);
            }
            pos += s;
        }
        const unsigned used = used_contours(icase, map, i, j, k, d);
        total += poly_count(w);
        acode[a] = icase | (map << 8);
        anv[a] = __popc(used);
        uint32_t nq = 0;
        nq += (edge_cut(icase, 0) && j >= 1 && k >= 1);
        nq += (edge_cut(icase, 3) && i >= 1 && k >= 1);
        nq += (edge_cut(icase, 8) && i >= 1 && j >= 1);
        anq[a] = nq;
    }
    total = wave_sum(total);
    if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(ncontours, total);
}

// Gradient of UniformGrid::estimateGradient followed by flip_gradient
// (cpp_dmc/UniformGrid.cpp:156-197) at grid point (i,j,k).  Out-of-range
// neighbours: index -1 clamps to 0, index n maps to 0 (the reference's
// index lambda returns max - i), the divisor stays 2*dx.
template <class T>
__device__ __forceinline__ void point_gradient(const T* __restrict__ field, const Dims& d, int i, int j, int k,
                                               double dx2, double dy2, double dz2, double g[3]) {
    auto idx = [](int a, int m) { return a < 0 ? 0 : (a >= m ? m - a : a); };
    const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
    const int64_t row = (int64_t)k * sz + (int64_t)j * sy;
    const double gx = ((double)field[row + idx(i + 1, d.nx)] - (double)field[row + idx(i - 1, d.nx)]) / dx2;
    const int64_t col = (int64_t)k * sz + i;
    const double gy = ((double)field[col + idx(j + 1, d.ny) * sy] - (double)field[col + idx(j - 1, d.ny) * sy]) / dy2;
    const int64_t pil = (int64_t)j * sy + i;
    const double gz = ((double)field[pil + idx(k + 1, d.nz) * sz] - (double)field[pil + idx(k - 1, d.nz) * sz]) / dz2;
    const double s = -1.;
    g[0] = gx * s;
    g[1] = gy * s;
    g[2] = gz * s;
}

// ---------------------------------------------------------------------------
// K4: vertex representatives, normals and positions
// (cpp_dmc/DualMarchingCubes.cpp:141-171, UniformGrid.h:286-305).
// One thread per active cell; vertices of used contours are written at
// avbase[a] + rank.  Failed projections are counted for every contour.
// ---------------------------------------------------------------------------
struct GridGeom {
    double x0, y0, z0, dx, dy, dz;
};

template <class T>
__global__ void __launch_bounds__(kBlock) k_vertices(const T* __restrict__ field, const double* __restrict__ grads,
                                                      Dims d, GridGeom geo, double iso,
                                                      const uint32_t* __restrict__ acell,
                                                      const uint32_t* __restrict__ acode,
                                                      const uint32_t* __restrict__ avbase, int64_t A,
                                                      double* __restrict__ vpos, double* __restrict__ vnrm,
                                                      uint32_t* __restrict__ fails) {
    uint32_t f1 = 0, f2 = 0;
    const double dx2 = 2 * geo.dx, dy2 = 2 * geo.dy, dz2 = 2 * geo.dz;
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        int i, j, k;
        cell_ijk(acell[a], d, i, j, k);
        const uint32_t code = acode[a];
        const unsigned icase = code & 0xFFu;
        double F[8];
        load_corners(field, d, i, j, k, F);
        const uint64_t w = cell_polygons(icase, iso, F);
        const int cnt = poly_count(w);
        BBoxes B;
        float lpar[12];
        cell_boxes(F, iso, icase, w, cnt, B, lpar);
        const unsigned used = used_contours(icase, code >> 8, i, j, k, d);
        uint32_t out = avbase[a];
        bool have_n = false;
        double G[8][3];
        for (int t = 0; t < cnt; t++) {
            double p[3];
            const int fl = contour_representative(F, iso, w, cnt, t, B, lpar, p);
            f1 += fl >= 1;
            f2 += fl >= 2;
            if (!((used >> t) & 1u)) continue;
            if (!have_n) {
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const int ci = i + (c & 1), cj = j + ((c >> 1) & 1), ck = k + ((c >> 2) & 1);
                    if (grads) {
                        const int64_t q = ((int64_t)ck * d.ny + cj) * d.nx + ci;
                        G[c][0] = grads[3 * q];
                        G[c][1] = grads[3 * q + 1];
                        G[c][2] = grads[3 * q + 2];
                    } else {
                        point_gradient(field, d, ci, cj, ck, dx2, dy2, dz2, G[c]);
                    }
                }
                have_n = true;
            }
            const double u = p[0], v = p[1], ww = p[2];
            double n[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double a0 = G[0][c] * (1 - u) + G[1][c] * u;
                const double a1 = G[2][c] * (1 - u) + G[3][c] * u;
                const double a2 = G[4][c] * (1 - u) + G[5][c] * u;
                const double a3 = G[6][c] * (1 - u) + G[7][c] * u;
                const double lo = a0 * (1 - v) + a1 * v;
                const double hi = a2 * (1 - v) + a3 * v;
                n[c] = lo * (1 - ww) + hi * ww;
            }
            const double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            n[0] /= s;
            n[1] /= s;
            n[2] /= s;
            vpos[3 * (int64_t)out + 0] = geo.x0 + (i + u) * geo.dx;
            vpos[3 * (int64_t)out + 1] = geo.y0 + (j + v) * geo.dy;
            vpos[3 * (int64_t)out + 2] = geo.z0 + (k + ww) * geo.dz;
            vnrm[3 * (int64_t)out + 0] = n[0];
            vnrm[3 * (int64_t)out + 1] = n[1];
            vnrm[3 * (int64_t)out + 2] = n[2];
            out++;
        }
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    if ((threadIdx.x & (kWave - 1)) == 0 && (f1 | f2)) {
        atomicAdd(&fails[0], f1);
        atomicAdd(&fails[1], f2);
    }
}

// Position of cell (i, j, k) in the active arrays: binary search in its
// row's list.  The cell is active whenever one of its edges is cut; -1 is
// returned (and counted by the caller) if it is not found.
__device__ __forceinline__ int64_t find_active(const uint32_t* __restrict__ rowData,
                                               const uint32_t* __restrict__ rowCount,
                                               const uint32_t* __restrict__ rowBase, const Dims& d, int i, int j,
                                               int k) {
    const int64_t r = (int64_t)k * (d.ny - 1) + j;
    const uint32_t* list = rowData + r * d.C;
    const uint32_t n = rowCount[r];
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((list[mid] & 0xFFFFFFu) < (uint32_t)i) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n || (list[lo] & 0xFFFFFFu) != (uint32_t)i) return -1;
    return (int64_t)rowBase[r] + lo;
}

// Final id of the vertex of cell `a` whose contour contains edge e.
__device__ __forceinline__ int32_t vertex_of(const uint32_t* __restrict__ acode, const uint32_t* __restrict__ avbase,
                                             int64_t a, int e, int i, int j, int k, const Dims& d) {
    if (a < 0) return -1;
    const uint32_t code = acode[a];
    const unsigned icase = code & 0xFFu;
    const uint32_t map = code >> 8;
    const unsigned t = (map >> (2 * e)) & 3u;
    const unsigned used = used_contours(icase, map, i, j, k, d);
    return (int32_t)(avbase[a] + __popc(used & ((1u << t) - 1u)));
}

// ---------------------------------------------------------------------------
// K5: quadrilaterals.  Each complete quadrilateral belongs to the grid edge
// e0 (x), e3 (y) or e8 (z) of the cell at its lower corner; ascending
// (cell, direction) is the reference's map order
// (cpp_dmc/DualMarchingCubes.h:215-223, :663-676).  Colour: 3*parity + dir
// (cpp_dmc/DualMarchingCubes.cpp:180-183).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_quads(Dims d, const uint32_t* __restrict__ acell,
                                                   const uint32_t* __restrict__ acode,
                                                   const uint32_t* __restrict__ avbase,
                                                   const uint32_t* __restrict__ aqbase, int64_t A,
                                                   const uint32_t* __restrict__ rowData,
                                                   const uint32_t* __restrict__ rowCount,
                                                   const uint32_t* __restrict__ rowBase,
                                                   int32_t* __restrict__ quads, int32_t* __restrict__ colors,
                                                   uint32_t* __restrict__ unresolved) {
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        int i, j, k;
        cell_ijk(acell[a], d, i, j, k);
        const unsigned icase = acode[a] & 0xFFu;
        const int bit = (icase & 1u) ? 0 : 1;     // field at the edge origin below iso
        const int parity = (i & 1) | (j & 1) << 1 | (k & 1) << 2;
        int64_t q = aqbase[a];
        // x edge: cells (i,j,k) e0, (i,j-1,k) e2, (i,j,k-1) e4, (i,j-1,k-1) e6
        if (edge_cut(icase, 0) && j >= 1 && k >= 1) {
            int32_t out[4];
            const int64_t b = find_active(rowData, rowCount, rowBase, d, i, j - 1, k);
            const int64_t c = find_active(rowData, rowCount, rowBase, d, i, j, k - 1);
            const int64_t e = find_active(rowData, rowCount, rowBase, d, i, j - 1, k - 1);
            out[edge_slot(0, bit)] = vertex_of(acode, avbase, a, 0, i, j, k, d);
            out[edge_slot(2, bit)] = vertex_of(acode, avbase, b, 2, i, j - 1, k, d);
            out[edge_slot(4, bit)] = vertex_of(acode, avbase, c, 4, i, j, k - 1, d);
            out[edge_slot(6, bit)] = vertex_of(acode, avbase, e, 6, i, j - 1, k - 1, d);
            if (out[0] < 0 || out[1] < 0 || out[2] < 0 || out[3] < 0) atomicAdd(unresolved, 1u);
            reinterpret_cast<int4*>(quads)[q] = make_int4(out[0], out[1], out[2], out[3]);
            colors[q] = 3 * parity;
            q++;
        }
        // y edge: cells (i,j,k) e3, (i-1,j,k) e1, (i,j,k-1) e7, (i-1,j,k-1) e5
        if (edge_cut(icase, 3) && i >= 1 && k >= 1) {
            int32_t out[4];
            const int64_t b = a - 1;
            const int64_t c = find_active(rowData, rowCount, rowBase, d, i, j, k - 1);
            const int64_t e = c > 0 ? c - 1 : -1;
            out[edge_slot(3, bit)] = vertex_of(acode, avbase, a, 3, i, j, k, d);
            out[edge_slot(1, bit)] = vertex_of(acode, avbase, b, 1, i - 1, j, k, d);
            out[edge_slot(7, bit)] = vertex_of(acode, avbase, c, 7, i, j, k - 1, d);
            out[edge_slot(5, bit)] = vertex_of(acode, avbase, e, 5, i - 1, j, k - 1, d);
            if (out[0] < 0 || out[1] < 0 || out[2] < 0 || out[3] < 0) atomicAdd(unresolved, 1u);
            reinterpret_cast<int4*>(quads)[q] = make_int4(out[0], out[1], out[2], out[3]);
            colors[q] = 3 * parity + 1;
            q++;
        }
        // z edge: cells (i,j,k) e8, (i-1,j,k) e9, (i,j-1,k) e11, (i-1,j-1,k) e10
        if (edge_cut(icase, 8) && i >= 1 && j >= 1) {
            int32_t out[4];
            const int64_t b = a - 1;
            const int64_t c = find_active(rowData, rowCount, rowBase, d, i, j - 1, k);
            const int64_t e = c > 0 ? c - 1 : -1;
            out[edge_slot(8, bit)] = vertex_of(acode, avbase, a, 8, i, j, k, d);
            out[edge_slot(9, bit)] = vertex_of(acode, avbase, b, 9, i - 1, j, k, d);
            out[edge_slot(11, bit)] = vertex_of(acode, avbase, c, 11, i, j - 1, k, d);
            out[edge_slot(10, bit)] = vertex_of(acode, avbase, e, 10, i - 1, j - 1, k, d);
            if (out[0] < 0 || out[1] < 0 || out[2] < 0 || out[3] < 0) atomicAdd(unresolved, 1u);
            reinterpret_cast<int4*>(quads)[q] = make_int4(out[0], out[1], out[2], out[3]);
            colors[q] = 3 * parity + 2;
            q++;
        }
    }
}

}  // namespace tmc
