// dmc_extract_kernels.h -- cell-sweep kernels of the DMC engine: the field
// sweep, per-row polygon codes, vertex representatives and quadrilaterals.
//
// Data layout in HBM:
//   field     nx*ny*nz scalars, x fastest (T = float or double)
//   rows      r = k*(ny-1) + j holds C = nx-1 cells, cut into S segments of
//             256 cells along x
//   lists     lists[(r*S + s)*256 + m] = i | case << 24: the m-th active
//             cell of segment s of row r (ascending i); segCount[r*S + s];
//             segBase = exclusive scan of segCount
//   active    a = segBase[r*S + s] + m enumerates the active cells in
//             (k, j, i) order, the reference's vertex order
//             (cpp_dmc/DualMarchingCubes.cpp:16-50 with :167-171):
//             acell[a] (linear cell index), acode[a] (case | edge-to-contour
//             map << 8), avloc[a] / aqloc[a] (first vertex / quadrilateral
//             of the cell relative to its row); rowBaseV / rowBaseQ are the
//             scans of the row totals.
#pragma once
#include <utility>

#include "dmc_cell.h"
#include "dmc_kernels.h"

namespace tmc {

struct Dims {
    int nx, ny, nz;
    int64_t C;      // cells per row (nx-1)
    int64_t R;      // rows (ny-1)*(nz-1)
};

struct GridGeom {
    double x0, y0, z0, dx, dy, dz;
};

// ---------------------------------------------------------------------------
// Volume table.  One extraction processes one or many independent volumes
// (a batch) in the same launches: the cell stage looks up the volume of each
// sweep wave, cell row and active cell; the mesh stage works on the union
// mesh, whose vertices and quadrilaterals are the volumes' in volume order.
// ---------------------------------------------------------------------------
struct VolDesc {
    const void* field;     // T* (device)
    const double* grads;   // 3 doubles per point, or null (gradients from the field)
    Dims d;
    GridGeom geo;
    double iso;            // iso-value as the reference uses it (double)
    double iso_t;          // comparison threshold, exact in the field type
    double rdx2[3];        // 1 / (2 dx), 1 / (2 dy), 1 / (2 dz) correctly rounded, or 0 (divide instead)
    int S, ntj, K, ntk;    // sweep tiling: segments per row, row tiles, planes per wave, plane tiles
    int64_t rowBase;       // first global cell row
    int64_t rsBase;        // first global (row, segment)
};

struct VolStat {
    unsigned int fail1, fail2;      // failed projections, levels 1 and 2
    unsigned long long contours;    // every contour (vertices before removeUnusedVertices)
};

struct VolTable {
    const VolDesc* desc;
    const int64_t* waveBase;   // n+1: first sweep wave of each volume
    const int64_t* rowBase;    // n+1: first global cell row
    const int64_t* aBase;      // n+1: first active cell (set after the sweep's scan)
    int n;
};


// ---------------------------------------------------------------------------
// K1: the field sweep.  A wave owns one x-segment of 256 cells for J
// consecutive cell rows and marches over K planes, keeping only the
// comparison bits (field >= iso) of the previous point plane in registers,
// so every field value is read once from HBM (the halo row shared with the
// next row tile and the halo plane shared with the next plane tile are read
// through L2 by both waves).  Lane L reads the 4 points x0+4L..x0+4L+3 with
// one vector load and owns the cells starting at them; the fifth point
// comes from the next lane, and for lane 63 from a wave-uniform scalar load
// of point x0+256.  All loads of a plane are issued before any comparison
// (clamped addresses, no branches), so a wave keeps J+1 loads in flight.  Active cells (case not 0/255) of a
// (row, segment) are written in ascending x using wave ballots.  For F32
// fields (double)f >= iso is evaluated as f >= iso_up, iso_up the smallest
// float not below iso: the same predicate.
// ---------------------------------------------------------------------------
constexpr int kSeg = 256;           // cells per segment (one wave) and list slots per (row, segment)
#ifndef TMC_SWEEP_ROWS
#define TMC_SWEEP_ROWS 16
#endif
constexpr int kSweepRows = TMC_SWEEP_ROWS;  // J
#ifndef TMC_VERTICES_ATTR
#define TMC_VERTICES_ATTR
#endif
// k_vertices' work-queue take: at most this many rounds of a block's threads.
#ifndef TMC_VERTICES_CHUNK
#define TMC_VERTICES_CHUNK 8
#endif
constexpr int kVerticesChunk = TMC_VERTICES_CHUNK;

// Loads of the 4 points i0..i0+3 of a row.  Rows read by one wave only are
// streamed (non-temporal, so they do not evict reused lines); rows that a
// neighbouring wave also reads (KEEP) are loaded through L2.
template <class T, bool VEC, bool KEEP>
__device__ __forceinline__ void row_load(const T* __restrict__ row, int i0, int nx, T v[4]) {
    if constexpr (VEC) {
        const int ic = min(i0, nx - 4);  // nx % 4 == 0: lanes past the end read the last 4 points
        typedef T vec4 __attribute__((ext_vector_type(4)));
        const vec4* src = reinterpret_cast<const vec4*>(row + ic);
        const vec4 t = KEEP ? *src : __builtin_nontemporal_load(src);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = row[min(i0 + q, nx - 1)];
    }
}

// bits 0..3: points i0..i0+3 at or above iso; bit 4: point i0+4 (next
// lane; `edge` is point x0+256 for lane 63)
template <class T>
__device__ __forceinline__ uint32_t row_bits(const T v[4], T edge, T iso_t, int lane) {
    const uint32_t b = (uint32_t)(v[0] >= iso_t) | (uint32_t)(v[1] >= iso_t) << 1 | (uint32_t)(v[2] >= iso_t) << 2 |
                       (uint32_t)(v[3] >= iso_t) << 3;
    // DPP wave_shl:1: lane L receives lane L+1's value
    uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(b & 1u), 0x130, 0xF, 0xF, false);
    nb = lane == kWave - 1 ? (uint32_t)(edge >= iso_t) : nb;
    return b | (nb << 4);
}

// Comparison bits of rows j0..j0+J of one point plane.  All loads are
// issued before any comparison (clamped row addresses, no branches): rows
// past the last one are clamped to it and their bits feed no cell.  The
// first and last rows are shared with the neighbouring row tiles; a whole
// plane is shared when KEEP.
template <class T, bool VEC, bool KEEP>
__device__ __forceinline__ void plane_bits(const T* __restrict__ pl, const Dims& d, int j0, int i0, int ie,
                                           T iso_t, int lane, uint32_t bits[kSweepRows + 1]) {
    constexpr int J = kSweepRows;
    T v[J + 1][4], e[J + 1];
#pragma unroll
    for (int jr = 0; jr <= J; jr++) {
        const T* row = pl + (int64_t)min(j0 + jr, d.ny - 1) * d.nx;
        if (KEEP || jr == 0 || jr == J) row_load<T, VEC, true>(row, i0, d.nx, v[jr]);
        else row_load<T, VEC, false>(row, i0, d.nx, v[jr]);
        e[jr] = row[ie];
    }
#pragma unroll
    for (int jr = 0; jr <= J; jr++) bits[jr] = row_bits<T>(v[jr], e[jr], iso_t, lane);
}

// v_writelane_b32 through its LLVM intrinsic, so that the compiler sees the
// instruction and inserts the wait states it needs after the VALU write of x.
extern "C" __device__ int tmc_writelane_i32(int x, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// Lane L of v takes the wave-uniform value x.
template <int L>
__device__ __forceinline__ void write_lane(uint32_t& v, uint32_t x) {
    static_assert(L >= 0 && L < kWave, "lane index");
    v = (uint32_t)tmc_writelane_i32((int)x, L, (int)v);
}

// State of the cells of one plane of a wave's tile.
struct PlaneCells {
    const uint32_t* lo;   // bits of point plane k, rows j0..j0+J
    const uint32_t* hi;   // bits of point plane k+1
    uint32_t* lists;      // list of the tile's first (row, segment)
    int S, i0, nj;
    uint32_t valid;
    // lane l of stage[h]: dword (l & 7) of the activity words of row
    // 8h + (l >> 3); lane jr of counts: active cells of row jr
    uint32_t stage[2];
    uint32_t counts;
};

// Cell row jr of a plane: the activity ballots, staged in lanes, and the
// list entries of the active cells in ascending x.
template <int JR>
__device__ __forceinline__ void cell_row(PlaneCells& c) {
    const uint32_t A = c.lo[JR], B = c.lo[JR + 1], Cm = c.hi[JR], D = c.hi[JR + 1];
    // bit q: cell i0+q has all 8 corners in / any corner in
    const uint32_t all_in = A & (A >> 1) & B & (B >> 1) & Cm & (Cm >> 1) & D & (D >> 1);
    const uint32_t any_in = A | (A >> 1) | B | (B >> 1) | Cm | (Cm >> 1) | D | (D >> 1);
    uint32_t act = any_in & ~all_in & (JR < c.nj ? c.valid : 0u);
    const uint64_t b0 = __ballot(act & 1u), b1 = __ballot(act & 2u), b2 = __ballot(act & 4u), b3 = __ballot(act & 8u);
    const uint32_t total = (uint32_t)(__popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3));
    uint32_t& st = c.stage[JR >> 3];
    constexpr int l0 = 8 * (JR & 7);
    write_lane<l0 + 0>(st, (uint32_t)b0);
    write_lane<l0 + 1>(st, (uint32_t)(b0 >> 32));
    write_lane<l0 + 2>(st, (uint32_t)b1);
    write_lane<l0 + 3>(st, (uint32_t)(b1 >> 32));
    write_lane<l0 + 4>(st, (uint32_t)b2);
    write_lane<l0 + 5>(st, (uint32_t)(b2 >> 32));
    write_lane<l0 + 6>(st, (uint32_t)b3);
    write_lane<l0 + 7>(st, (uint32_t)(b3 >> 32));
    write_lane<JR>(c.counts, total);
    if (total) {
        // cells of lower lanes (any q) come first, then this lane's lower q
        uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(b0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b0, 0u));
        pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(b1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b1, pos));
        pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(b2 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b2, pos));
        pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(b3 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b3, pos));
        uint32_t* out = c.lists + (int64_t)JR * c.S * kSeg;
        // 5-bit row words side by side: bits q, q+1 of each byte are the two
        // corners of cell q in that row
        const uint32_t X = A | B << 8 | Cm << 16 | D << 24;
        while (act) {
            const int q = __builtin_ctz(act);
            act &= act - 1u;
            const uint32_t m = (X >> q) & 0x03030303u;
            const uint32_t ic = (m | m >> 6 | m >> 12 | m >> 18) & 0xFFu;
            out[pos++] = (uint32_t)(c.i0 + q) | (ic << 24);
        }
    }
}

template <size_t... JR>
__device__ __forceinline__ void cell_rows(PlaneCells& c, std::index_sequence<JR...>) {
    (cell_row<(int)JR>(c), ...);
}

// The cells of plane k of the wave's tile from the bits of point planes k
// (lo) and k+1 (hi); the activity words and active counts of the rows below
// nj are stored once per plane from their lanes.
__device__ __forceinline__ void plane_cells(const uint32_t lo[kSweepRows + 1], const uint32_t hi[kSweepRows + 1],
                                            int64_t k, const Dims& d, int j0, int nj, int S, int s, int i0,
                                            uint32_t valid, int lane, uint32_t* __restrict__ lists,
                                            uint32_t* __restrict__ segCount,
                                            unsigned long long* __restrict__ segMask) {
    static_assert(kSweepRows <= 16, "activity words of the rows are staged in two VGPRs");
    const int64_t rs0 = (k * (d.ny - 1) + j0) * S + s;  // (row, segment) of the tile's first row
    PlaneCells c{lo, hi, lists + rs0 * kSeg, S, i0, nj, valid, {0u, 0u}, 0u};
    cell_rows(c, std::make_index_sequence<kSweepRows>{});
    // activity of the segment: bit L of word q <-> cell x0 + 4L + q
    uint32_t* mask32 = reinterpret_cast<uint32_t*>(segMask);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int row = 8 * h + (lane >> 3);
        if (row < nj && row < kSweepRows) mask32[8 * (rs0 + (int64_t)row * S) + (lane & 7)] = c.stage[h];
    }
    if (lane < nj) segCount[rs0 + (int64_t)lane * S] = c.counts;
}

// March of one wave over its planes k0..k1, upwards (DOWN false) or
// downwards.  Plane tiles alternate direction, so two neighbouring tiles
// read their shared boundary plane both first or both last, at about the
// same time: the second read is served by L2.
template <class T, bool VEC, bool DOWN>
__device__ __forceinline__ void sweep_march(const T* __restrict__ field, const Dims& d, int k0, int k1, int j0,
                                            int nj, int S, int s, int i0, int ie, uint32_t valid,
                                            T iso_t, int lane, uint32_t* __restrict__ lists,
                                            uint32_t* __restrict__ segCount,
                                            unsigned long long* __restrict__ segMask) {
    constexpr int J = kSweepRows;
    const int64_t plane = (int64_t)d.nx * d.ny;
    uint32_t prev[J + 1], cur[J + 1];
    int kp = DOWN ? k1 : k0;
    plane_bits<T, VEC, true>(field + (int64_t)kp * plane, d, j0, i0, ie, iso_t, lane, prev);
    for (int n = k1 - k0; n > 0; n--) {
        const int kn = DOWN ? kp - 1 : kp + 1;
        if (n == 1) plane_bits<T, VEC, true>(field + (int64_t)kn * plane, d, j0, i0, ie, iso_t, lane, cur);
        else plane_bits<T, VEC, false>(field + (int64_t)kn * plane, d, j0, i0, ie, iso_t, lane, cur);
        if (DOWN) plane_cells(cur, prev, kn, d, j0, nj, S, s, i0, valid, lane, lists, segCount, segMask);
        else plane_cells(prev, cur, kp, d, j0, nj, S, s, i0, valid, lane, lists, segCount, segMask);
#pragma unroll
        for (int jr = 0; jr <= J; jr++) prev[jr] = cur[jr];
        kp = kn;
    }
}

template <class T, bool VEC>
__global__ void __launch_bounds__(kBlock) k_sweep(VolTable vt, int64_t waves, uint32_t* __restrict__ lists,
                                                   uint32_t* __restrict__ segCount,
                                                   unsigned long long* __restrict__ segMask) {
    constexpr int J = kSweepRows;
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform from here on (kept in scalar registers)
    const int64_t gwave =
        xcd_block() * (kBlock / kWave) + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    if (gwave >= waves) return;
    const int v = vol_of(vt.waveBase, vt.n, gwave);
    const VolDesc& V = vt.desc[v];
    const T* __restrict__ field = static_cast<const T*>(V.field);
    const Dims d = V.d;
    const T iso_t = (T)V.iso_t;
    const int S = V.S, ntj = V.ntj, K = V.K;
    const int64_t wave = gwave - vt.waveBase[v];
    lists += V.rsBase * kSeg;
    segCount += V.rsBase;
    segMask += 4 * V.rsBase;
    const int s = (int)(wave % S);
    const int64_t rest = wave / S;
    const int tj = (int)(rest % ntj), tk = (int)(rest / ntj);
    const int j0 = tj * J, k0 = tk * K;
    const int nj = min(J, d.ny - 1 - j0);
    const int k1 = min(k0 + K, d.nz - 1);
    const int i0 = s * kSeg + lane * 4;
    const int ie = min(s * kSeg + kSeg, d.nx - 1);  // lane 63's fifth point (uniform)
    // cells i0+q (q < 4) that exist (i < nx-1)
    const int ncell = d.nx - 1 - i0;
    const uint32_t valid = ncell >= 4 ? 0xFu : ncell > 0 ? (1u << ncell) - 1u : 0u;
    if (tk & 1)
        sweep_march<T, VEC, true>(field, d, k0, k1, j0, nj, S, s, i0, ie, valid, iso_t, lane, lists, segCount,
                                  segMask);
    else
        sweep_march<T, VEC, false>(field, d, k0, k1, j0, nj, S, s, i0, ie, valid, iso_t, lane, lists, segCount,
                                   segMask);
}

// A volume's field as a global-memory pointer.  The descriptor keeps a
// generic pointer, through which the compiler emits flat loads; those count
// in the LDS counter too, so an LDS wait would also wait for them.  Global
// loads leave the LDS traffic of the cell kernels free to proceed.
template <class T>
using GPtr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ GPtr<T> gfield(const void* p) {
    return (GPtr<T>)p;
}

template <class T, class P>
__device__ __forceinline__ void load_corners(P field, const Dims& d, int i, int j, int k, double F[8]) {
    const int64_t sz = (int64_t)d.nx * d.ny;
    const P p = field + (int64_t)k * sz + (int64_t)j * d.nx + i;
    F[0] = (double)p[0];
    F[1] = (double)p[1];
    F[2] = (double)p[d.nx];
    F[3] = (double)p[d.nx + 1];
    F[4] = (double)p[sz];
    F[5] = (double)p[sz + 1];
    F[6] = (double)p[sz + d.nx];
    F[7] = (double)p[sz + d.nx + 1];
}

__device__ __forceinline__ void cell_ijk(uint32_t c, const Dims& d, int& i, int& j, int& k, uint32_t& r) {
    r = c / (uint32_t)d.C;
    i = (int)(c - r * (uint32_t)d.C);
    j = (int)(r % (uint32_t)(d.ny - 1));
    k = (int)(r / (uint32_t)(d.ny - 1));
}

// Edge masks in "spread" form: edge e at bit 2e, the layout of the contour
// map (2 bits per edge).  Cut edges of each case from a table; interior
// edges of a cell away from the grid boundary are all twelve.
struct CutSpread {
    uint32_t v[256];
    constexpr CutSpread() : v() {
        for (int c = 0; c < 256; c++) {
            uint32_t m = 0;
            for (int e = 0; e < 12; e++)
                if (((c >> kEdgeV0[e]) ^ (c >> kEdgeV1[e])) & 1) m |= 1u << (2 * e);
            v[c] = m;
        }
    }
};
__device__ __constant__ CutSpread c_cut_spread{};
constexpr uint32_t kAllEdgesSpread = 0x555555u;

__device__ __forceinline__ uint32_t interior_spread(int i, int j, int k, const Dims& d) {
    if (i >= 1 && i <= d.nx - 3 && j >= 1 && j <= d.ny - 3 && k >= 1 && k <= d.nz - 3) return kAllEdgesSpread;
    uint32_t m = 0;
#pragma unroll
    for (int e = 0; e < 12; e++)
        if (edge_interior(e, i, j, k, d.nx, d.ny, d.nz)) m |= 1u << (2 * e);
    return m;
}
// bit 2e set iff edge e belongs to contour t in the map
__device__ __forceinline__ uint32_t contour_spread(uint32_t map, unsigned t) {
    const uint32_t x = map ^ (t * kAllEdgesSpread);
    return ~(x | (x >> 1)) & kAllEdgesSpread;
}
// Per active cell: interior cut edges (spread, bits 0..23) and used
// contours (bits 24..27), formed once in k_rows.  Used contours are those
// with at least one interior edge; the others only feed incomplete
// quadrilaterals and are dropped by removeUnusedVertices
// (cpp_dmc/DualMarchingCubes.h:685-729).  A vertex's quadrilaterals are
// those of its contour's interior edges, one each: its degree and, ordered
// by edge id, its rows in the first halfedge CSR.
__device__ __forceinline__ uint32_t cell_aux(unsigned icase, uint32_t map, int i, int j, int k, const Dims& d) {
    const uint32_t live = c_cut_spread.v[icase] & interior_spread(i, j, k, d);
    uint32_t used = 0;
#pragma unroll
    for (unsigned t = 0; t < 4; t++)
        if (contour_spread(map, t) & live) used |= 1u << t;
    return (live & kAllEdgesSpread) | used << 24;
}
// interior cut edges of contour t below edge e / in all: a vertex's rows
__device__ __forceinline__ int contour_rank(uint32_t map, uint32_t aux, unsigned icase, unsigned t, int e) {
    return __popc(contour_spread(map, t) & aux & c_cut_spread.v[icase] & kAllEdgesSpread & ((1u << (2 * e)) - 1u));
}

// ---------------------------------------------------------------------------
// K2: one wave per row, cells in ascending x: polygons of every active cell
// -> code, kept vertices and owned quadrilaterals (edges e0, e3, e8 of the
// cell with all four neighbour cells present) and their prefix inside the
// row; the flat active arrays; row totals for the row scans.  Counts every
// contour for the reference's pre-removal vertex count.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock) k_rows(VolTable vt, int64_t Rtot, const uint32_t* __restrict__ lists,
                                                  const uint32_t* __restrict__ segCount,
                                                  const uint32_t* __restrict__ segBase,
                                                  uint32_t* __restrict__ acell, uint32_t* __restrict__ acode,
                                                  uint32_t* __restrict__ aaux,
                                                  uint32_t* __restrict__ avloc, uint32_t* __restrict__ aqloc,
                                                  uint32_t* __restrict__ rowV, uint32_t* __restrict__ rowQ,
                                                  VolStat* __restrict__ vstat, unsigned long long* __restrict__ cpart,
                                                  uint32_t acap) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wpb = kBlock / kWave;
    // one volume: its contour count is spread over kContourParts counters
    // (summed by the host), so that the waves' adds do not queue on a few
    // addresses
    auto contours = [&](int vol) {
        return vt.n == 1 ? &cpart[(blockIdx.x * wpb + threadIdx.x / kWave) & (kContourParts - 1)]
                         : &vstat[vol].contours;
    };
    unsigned long long total = 0;
    int cur = -1;  // volume whose contour count `total` holds (wave-uniform)
    for (int64_t g = (int64_t)blockIdx.x * wpb + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave); g < Rtot;
         g += (int64_t)gridDim.x * wpb) {
        const int v = vol_of(vt.rowBase, vt.n, g);
        if (v != cur) {
            total = wave_sum(total);
            if (lane == 0 && total) atomicAdd(contours(cur), total);
            total = 0;
            cur = v;
        }
        const VolDesc& V = vt.desc[v];
        const GPtr<T> field = gfield<T>(V.field);
        const Dims d = V.d;
        const double iso = V.iso;
        const int S = V.S;
        const int64_t r = g - V.rowBase;
        const int j = (int)(r % (d.ny - 1));
        const int k = (int)(r / (d.ny - 1));
        uint32_t cv = 0, cq = 0;
        // the row's active cells are contiguous in the flat order: process
        // them 64 at a time across segment boundaries
        const int64_t rs0 = V.rsBase + r * S;
        // the row's segment bases, one per lane (rows of at most 64 segments;
        // wider rows search memory)
        const bool regs = S <= kWave;
        const uint32_t mybase = (regs && lane < S) ? segBase[rs0 + lane] : 0u;
        const uint32_t first = regs ? (uint32_t)__builtin_amdgcn_readfirstlane((int)mybase) : segBase[rs0];
        const uint32_t last = segBase[rs0 + S - 1] + segCount[rs0 + S - 1];
        for (uint32_t a0 = first; a0 < last; a0 += kWave) {
            const uint32_t a = a0 + lane;
            const bool act = a < last;
            // segment of a: the last s with segBase[rs0 + s] <= a
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
            uint32_t nv = 0, nq = 0, code = 0, aux = 0;
            int i = 0;
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
                i = (int)(x & 0xFFFFFFu);
                const unsigned icase = x >> 24;
                uint64_t w;
                if (case_ambiguous(icase)) {
                    double F[8];
                    load_corners<T>(field, d, i, j, k, F);
                    w = polygons_decider(iso, F);
                } else {
                    w = c_case_contours[icase];
                }
                uint32_t map = 0;
                int pos = 0;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int sz = poly_size(w, t);
                    for (int q = 0; q < sz; q++) map |= (uint32_t)t << (2 * poly_edge_at(w, pos + q));
                    pos += sz;
                }
                total += (unsigned long long)poly_count(w);
                aux = cell_aux(icase, map, i, j, k, d);
                nv = (uint32_t)__popc(aux >> 24);
                nq = (uint32_t)(edge_cut(icase, 0) && j >= 1 && k >= 1) +
                     (uint32_t)(edge_cut(icase, 3) && i >= 1 && k >= 1) +
                     (uint32_t)(edge_cut(icase, 8) && i >= 1 && j >= 1);
                code = icase | (map << 8);
            }
            const uint32_t iv = wave_inclusive_scan(nv), iq = wave_inclusive_scan(nq);
            // the counts are exact whatever the capacity; entries past it are
            // left out (the pass is then run again with room for them)
            if (act && a < acap) {
                acell[a] = (uint32_t)(r * d.C) + (uint32_t)i;
                acode[a] = code;
                aaux[a] = aux;
                avloc[a] = cv + iv - nv;
                aqloc[a] = cq + iq - nq;
            }
            cv += (uint32_t)__shfl((int)iv, kWave - 1, kWave);
            cq += (uint32_t)__shfl((int)iq, kWave - 1, kWave);
        }
        if (lane == 0) {
            rowV[g] = cv;
            rowQ[g] = cq;
        }
    }
    total = wave_sum(total);
    if (lane == 0 && total) atomicAdd(contours(cur), total);
}

// ---------------------------------------------------------------------------
// K3: vertices (cpp_dmc/DualMarchingCubes.cpp:141-171, UniformGrid.h:286-305).
// One thread per active cell; position, normal and count of interior cut
// edges of the vertices of used contours are written from
// rowBaseV[row] + avloc[a] on.  Cells that need the general sampling (a
// contour whose samples must avoid another contour's box, or a 9 x 9
// sampling that finds no point) are appended to `slow` (slow[0] = count,
// entries from slow[1]) and finished by k_vertices_slow, which also counts
// the failed projections: only those cells can have any.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ uint64_t cell_corners(const VolDesc& V, uint32_t cellword, uint32_t code,
                                                 const CornerRef& Fc, int& i, int& j, int& k, uint32_t& r) {
    const GPtr<T> field = gfield<T>(V.field);
    cell_ijk(cellword, V.d, i, j, k, r);
    double F[8];
    load_corners<T>(field, V.d, i, j, k, F);
    Fc.store(F);
    return cell_polygons(code & 0xFFu, V.iso, F);
}

// The 24 field values one step beyond cell (i, j, k) along each axis that
// the corner gradients need besides the cell's corners, with the index rule
// of estimateGradient (cpp_dmc/UniformGrid.cpp:156-197: -1 -> 0, n -> 0, the
// top boundary wraps).  x[b]: row (j + b&1, k + b>>1) at i-1 and i+2;
// y[b]: column (i + b&1, k + b>>1) at j-1 and j+2; z[b]: column
// (i + b&1, j + b>>1) at k-1 and k+2.
template <class T>
struct OuterValues {
    T x[4][2], y[4][2], z[4][2];
};

template <class T>
__device__ __forceinline__ void load_outer(const VolDesc& V, int i, int j, int k, OuterValues<T>& o) {
    const Dims& d = V.d;
    const GPtr<T> field = gfield<T>(V.field);
    auto idx = [](int x, int m) { return x < 0 ? 0 : (x >= m ? m - x : x); };
    const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
    const int xlo = idx(i - 1, d.nx), xhi = idx(i + 2, d.nx);
    const int ylo = idx(j - 1, d.ny), yhi = idx(j + 2, d.ny);
    const int zlo = idx(k - 1, d.nz), zhi = idx(k + 2, d.nz);
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int64_t row = (int64_t)(k + (b >> 1)) * sz + (int64_t)(j + (b & 1)) * sy;
        o.x[b][0] = field[row + xlo];
        o.x[b][1] = field[row + xhi];
        const int64_t colY = (int64_t)(k + (b >> 1)) * sz + i + (b & 1);
        o.y[b][0] = field[colY + (int64_t)ylo * sy];
        o.y[b][1] = field[colY + (int64_t)yhi * sy];
        const int64_t colZ = (int64_t)(j + (b >> 1)) * sy + i + (b & 1);
        o.z[b][0] = field[colZ + (int64_t)zlo * sz];
        o.z[b][1] = field[colZ + (int64_t)zhi * sz];
    }
}

// x / c, correctly rounded, from rc = 1 / c correctly rounded: q = x rc is
// within an ulp of x / c, e = x - q c is exact (one FMA), and q + e rc
// rounds to the correctly rounded quotient (Markstein's theorem) -- three
// FP64 operations for the dozen of a division.  Zeros, infinities, NaN and
// quotients near the exponent range's ends take the division itself, as
// does rc = 0.  tests/cpp/exact_div_check.c checks it against x / c bit for
// bit.
__device__ __forceinline__ double div_by(double x, double c, double rc) {
    const double q = x * rc;
    const double e = fma(-q, c, x);
    const double q2 = fma(e, rc, q);
    const double aq = fabs(q);
    double r = q2;
    // a branch, not a select: the wave runs the full division only when one
    // of its lanes needs it
    if (__builtin_expect(!(rc != 0.0 && aq > 0x1p-960 && aq < 0x1p+960), 0)) r = x / c;
    return r;
}

// Position, normal and interior-edge count of the vertex at local
// coordinates p of cell (i, j, k) (cpp_dmc/DualMarchingCubes.cpp:160-171,
// UniformGrid.h:286-305).  The normal is the trilinear blend of the eight
// corner gradients, normalised.  Corner gradients are the supplied ones, or
// those of estimateGradient + flip_gradient: central differences of the
// cell's corners (in LDS) and the outer values o.
template <class T>
__device__ __forceinline__ void store_vertex(const VolDesc& V, int i, int j, int k, const CornerRef& Fc,
                                             const OuterValues<T>& o, const double p[3], uint32_t code, uint32_t aux,
                                             int t, uint32_t out, double* __restrict__ vpos,
                                             double* __restrict__ vnrm, uint32_t* __restrict__ vdeg) {
    const Dims& d = V.d;
    const double u = p[0], v = p[1], ww = p[2];
    const double dx2[3] = {2 * V.geo.dx, 2 * V.geo.dy, 2 * V.geo.dz};
    double n[3];
    // one component at a time: its eight corner values, then the blend
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double G[8];
        if (V.grads) {
            const GPtr<double> grads = gfield<double>(V.grads);
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int64_t x = ((int64_t)(k + ((c >> 2) & 1)) * d.ny + (j + ((c >> 1) & 1))) * d.nx + i + (c & 1);
                G[c] = grads[3 * x + q];
            }
        } else {
            const double s = -1.;
            const int bit = 1 << q;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
                // the outer value on this axis: the other two corner bits
                // select the row (x), column (y) or pillar (z)
                const int b = q == 0 ? (cy | cz << 1) : q == 1 ? (cx | cz << 1) : (cx | cy << 1);
                const T* ov = q == 0 ? o.x[b] : q == 1 ? o.y[b] : o.z[b];
                const double g = div_by((c & bit) ? (double)ov[1] - Fc(c & ~bit) : Fc(c | bit) - (double)ov[0],
                                        dx2[q], V.rdx2[q]);
                G[c] = g * s;
            }
        }
        // (1-u) n_even + u n_odd, then rows by v, then planes by w
        const double lo0 = (G[0] * (1 - u) + G[1] * u) * (1 - v);
        const double hi0 = lo0 + (G[2] * (1 - u) + G[3] * u) * v;
        const double lo1 = (G[4] * (1 - u) + G[5] * u) * (1 - v);
        const double hi1 = lo1 + (G[6] * (1 - u) + G[7] * u) * v;
        n[q] = hi0 * (1 - ww) + hi1 * ww;
    }
    const double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const GridGeom& geo = V.geo;
    double* vp = vpos + 3 * (int64_t)out;
    vp[0] = geo.x0 + (i + u) * geo.dx;
    vp[1] = geo.y0 + (j + v) * geo.dy;
    vp[2] = geo.z0 + (k + ww) * geo.dz;
    double* vn = vnrm + 3 * (int64_t)out;
    const double rs = 1.0 / s;
    vn[0] = div_by(n[0], s, rs);
    vn[1] = div_by(n[1], s, rs);
    vn[2] = div_by(n[2], s, rs);
    vdeg[out] = (uint32_t)contour_rank(code >> 8, aux, code & 0xFFu, (unsigned)t, 12);
}

template <class T, int kUnroll>
__device__ __forceinline__ void vertices_run(const VolTable& vt, const uint32_t* __restrict__ acell,
                                             const uint32_t* __restrict__ acode, const uint32_t* __restrict__ aaux,
                                             const uint32_t* __restrict__ avloc,
                                             const uint32_t* __restrict__ rowBaseV, DevN dA,
                                             double* __restrict__ vpos, double* __restrict__ vnrm,
                                             uint32_t* __restrict__ vdeg, uint32_t* __restrict__ slow,
                                             unsigned long long* qhead) {
    static_assert(kBlock == kBoxStride && kBlock == kCornerStride, "LDS columns are one per thread");
    const int64_t A = dA.get();
    __shared__ float boxes[kBoxFloats * kBoxStride];
    __shared__ double corners[8 * kCornerStride];
    const BoxRef B{boxes + threadIdx.x};
    const CornerRef Fc{corners + threadIdx.x};
    // Persistent blocks pull chunks of kVerticesChunk x kBlock active cells
    // from a device work queue (*qhead, ascending): cells cost unevenly (one
    // to four contours, searches of different lengths, the volumes of a
    // batch of different sizes), so no block is held to a fixed share.  Each
    // take is one add on a single address for all blocks, so a take covers
    // several rounds of the block's threads: up to kVerticesChunk, fewer
    // when the cells would make under four takes per block (a small volume
    // keeps every block busy instead of leaving most of the grid idle).
    const int rounds = (int)std::min<int64_t>(kVerticesChunk, std::max<int64_t>(1, A / ((int64_t)kBlock * gridDim.x * 4)));
    __shared__ int64_t chunk;
    for (;;) {
        if (threadIdx.x == 0) chunk = (int64_t)atomicAdd(qhead, (unsigned long long)(rounds * kBlock));
        __syncthreads();
        const int64_t c0 = chunk;
        const bool drained = c0 >= A;
        __syncthreads();
        if (drained) break;
        for (int round = 0; round < rounds; round++) {
        const int64_t a = c0 + (int64_t)round * kBlock + threadIdx.x;
        if (a >= A) continue;
        const int vol = vol_of(vt.aBase, vt.n, a);
        const VolDesc& V = vt.desc[vol];
        const double iso = V.iso;
        const uint32_t code = acode[a];
        int i, j, k;
        uint32_t r;
        const uint64_t w = cell_corners<T>(V, acell[a], code, Fc, i, j, k, r);
        const int cnt = poly_count(w);
        bool defer = cell_needs_box_test(w, cnt);
        if (!defer) {
            // issued before the sampling, which hides their latency
            OuterValues<T> o;
            if (!V.grads) load_outer<T>(V, i, j, k, o);
            cell_boxes(Fc, iso, w, cnt, B);
            const uint32_t aux = aaux[a];
            const unsigned used = aux >> 24;
            uint32_t out = rowBaseV[V.rowBase + r] + avloc[a];
            for (int t = 0; t < cnt; t++) {
                double p[3];
                if (!contour_representative9<kUnroll>(Fc, iso, w, t, B, p)) {
                    defer = true;
                    break;
                }
                if (!((used >> t) & 1u)) continue;
                store_vertex<T>(V, i, j, k, Fc, o, p, code, aux, t, out, vpos, vnrm, vdeg);
                out++;
            }
        }
        // one counter update per wave for its deferred cells
        const uint64_t m = __ballot(defer);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(slow, (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, leader, kWave);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (defer) slow[1 + base + below] = (uint32_t)a;
        }
        }
    }
}

// K3 for passes of up to one round of the persistent grid's threads per
// active cell: a row's samples interleaved (the latency of one cell matters
// most), 2 waves per SIMD.
template <class T>
__global__ void __launch_bounds__(kBlock) TMC_VERTICES_ATTR k_vertices(VolTable vt, const uint32_t* __restrict__ acell,
                                                      const uint32_t* __restrict__ acode,
                                                      const uint32_t* __restrict__ aaux,
                                                      const uint32_t* __restrict__ avloc,
                                                      const uint32_t* __restrict__ rowBaseV, DevN dA,
                                                      double* __restrict__ vpos, double* __restrict__ vnrm,
                                                      uint32_t* __restrict__ vdeg, uint32_t* __restrict__ slow,
                                                      unsigned long long* qhead) {
    vertices_run<T, 7>(vt, acell, acode, aaux, avloc, rowBaseV, dA, vpos, vnrm, vdeg, slow, qhead);
}

// K3 for larger passes: samples one at a time, which fits 3 waves per SIMD
// (166 VGPRs for F32 fields, no scratch), so more cells hide each other's
// memory waits (1024^3 Schwarz-P: 20.25 -> 19.85 ms per extraction; one
// 128^3 volume: 0.401 -> 0.407 ms, hence the two forms).  The same
// arithmetic in the same order as k_vertices.
template <class T>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_vertices_wide(VolTable vt, const uint32_t* __restrict__ acell, const uint32_t* __restrict__ acode,
                const uint32_t* __restrict__ aaux, const uint32_t* __restrict__ avloc,
                const uint32_t* __restrict__ rowBaseV, DevN dA, double* __restrict__ vpos,
                double* __restrict__ vnrm, uint32_t* __restrict__ vdeg, uint32_t* __restrict__ slow,
                unsigned long long* qhead) {
    vertices_run<T, 1>(vt, acell, acode, aaux, avloc, rowBaseV, dA, vpos, vnrm, vdeg, slow, qhead);
}

// General path of K3 for the cells k_vertices deferred: every contour is
// sampled with the box test at 9 x 9 and, where that finds no point, at
// 28 x 28; failed projections are counted per volume.
template <class T>
__global__ void __launch_bounds__(kBlock) k_vertices_slow(VolTable vt, const uint32_t* __restrict__ acell,
                                                          const uint32_t* __restrict__ acode,
                                                          const uint32_t* __restrict__ aaux,
                                                          const uint32_t* __restrict__ avloc,
                                                          const uint32_t* __restrict__ rowBaseV,
                                                          const uint32_t* __restrict__ slow,
                                                          double* __restrict__ vpos, double* __restrict__ vnrm,
                                                          uint32_t* __restrict__ vdeg, VolStat* __restrict__ vstat) {
    __shared__ float boxes[kBoxFloats * kBoxStride];
    __shared__ double corners[8 * kCornerStride];
    const BoxRef B{boxes + threadIdx.x};
    const CornerRef Fc{corners + threadIdx.x};
    const uint32_t n = slow[0];
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        const int64_t a = slow[1 + s];
        const int vol = vol_of(vt.aBase, vt.n, a);
        const VolDesc& V = vt.desc[vol];
        const double iso = V.iso;
        const uint32_t code = acode[a];
        int i, j, k;
        uint32_t r;
        const uint64_t w = cell_corners<T>(V, acell[a], code, Fc, i, j, k, r);
        const int cnt = poly_count(w);
        OuterValues<T> o;
        if (!V.grads) load_outer<T>(V, i, j, k, o);
        cell_boxes(Fc, iso, w, cnt, B);
        const uint32_t aux = aaux[a];
        const unsigned used = aux >> 24;
        uint32_t out = rowBaseV[V.rowBase + r] + avloc[a];
        for (int t = 0; t < cnt; t++) {
            double p[3];
            const int fl = contour_representative(Fc, iso, w, cnt, t, B, p);
            // failed projections: one add per wave when its lanes share the
            // volume (always, for one volume)
            const uint64_t m1 = __ballot(fl >= 1), m2 = __ballot(fl >= 2);
            if (m1) {
                const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
                if (__all(vol == __shfl(vol, leader, kWave))) {
                    if ((int)__lane_id() == leader) {
                        atomicAdd(&vstat[vol].fail1, (uint32_t)__popcll(m1));
                        if (m2) atomicAdd(&vstat[vol].fail2, (uint32_t)__popcll(m2));
                    }
                } else {
                    if (fl >= 1) atomicAdd(&vstat[vol].fail1, 1u);
                    if (fl >= 2) atomicAdd(&vstat[vol].fail2, 1u);
                }
            }
            if (!((used >> t) & 1u)) continue;
            store_vertex<T>(V, i, j, k, Fc, o, p, code, aux, t, out, vpos, vnrm, vdeg);
            out++;
        }
    }
}

// Activity of the cells of one volume, for neighbour lookups.
struct SegLists {
    const unsigned long long* mask;   // 4 words per (row, segment): bit L of word q <-> cell 4L + q
    const uint32_t* base;             // first active index of each (row, segment)
    int S;
    int64_t rsBase;                   // first (row, segment) of the volume
};

// Position of cell (i, j, k) in the active arrays from its segment's
// activity words: the active cells before it are those of lower lanes (any
// word) and of its own lane in lower words.  -1 if the cell is not active.
__device__ __forceinline__ int64_t find_active(const SegLists& L, const Dims& d, int i, int j, int k) {
    const int64_t rs = L.rsBase + ((int64_t)k * (d.ny - 1) + j) * L.S + (i / kSeg);
    const int il = i % kSeg, ln = il >> 2, q = il & 3;
    const ulonglong2* m2 = reinterpret_cast<const ulonglong2*>(L.mask + 4 * rs);
    const ulonglong2 w01 = m2[0], w23 = m2[1];
    const unsigned long long w[4] = {w01.x, w01.y, w23.x, w23.y};
    if (!((w[q] >> ln) & 1ull)) return -1;
    const unsigned long long below = (1ull << ln) - 1ull;
    int rank = 0;
#pragma unroll
    for (int t = 0; t < 4; t++) rank += __popcll(w[t] & below) + (t < q ? (int)((w[t] >> ln) & 1ull) : 0);
    return (int64_t)L.base[rs] + rank;
}

// Final id of the vertex of active cell `a` whose contour contains edge e,
// and the row of edge e among that contour's interior edges (its slot in the
// first halfedge CSR).  rowBase points at the volume's first row; r is the
// cell's row in the volume.
__device__ __forceinline__ int32_t vertex_slot_of(const uint32_t* __restrict__ acode,
                                                  const uint32_t* __restrict__ aaux,
                                                  const uint32_t* __restrict__ avloc,
                                                  const uint32_t* __restrict__ rowBaseV, int64_t a, int e,
                                                  int64_t r, int& slot) {
    slot = 0;
    if (a < 0) return -1;
    const uint32_t code = acode[a], aux = aaux[a];
    const unsigned icase = code & 0xFFu;
    const uint32_t map = code >> 8;
    const unsigned t = (map >> (2 * e)) & 3u;
    slot = contour_rank(map, aux, icase, t, e);
    const uint32_t vb = rowBaseV[r] + avloc[a];
    return (int32_t)(vb + __popc((aux >> 24) & ((1u << t) - 1u)));
}

// ---------------------------------------------------------------------------
// K4: quadrilaterals.  Each complete quadrilateral belongs to the grid edge
// e0 (x), e3 (y) or e8 (z) of the cell at its lower corner; ascending
// (cell, direction) is the reference's map order
// (cpp_dmc/DualMarchingCubes.h:215-223, :663-676).  Colour: 3*parity + dir
// (cpp_dmc/DualMarchingCubes.cpp:180-183).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_quads(VolTable vt, const uint32_t* __restrict__ acell,
                                                   const uint32_t* __restrict__ acode,
                                                   const uint32_t* __restrict__ aaux,
                                                   const uint32_t* __restrict__ avloc,
                                                   const uint32_t* __restrict__ aqloc,
                                                   const uint32_t* __restrict__ rowBaseV0,
                                                   const uint32_t* __restrict__ rowBaseQ, DevN dA,
                                                   const unsigned long long* __restrict__ segMask,
                                                   const uint32_t* __restrict__ segBase,
                                                   int32_t* __restrict__ quads, uint8_t* __restrict__ colors,
                                                   const uint32_t* __restrict__ off, int2* __restrict__ adj,
                                                   uint64_t adjcap, uint32_t* unresolved) {
    const int64_t A = dA.get();
    for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < A; a += (int64_t)gridDim.x * kBlock) {
        const int vol = vol_of(vt.aBase, vt.n, a);
        const VolDesc& V = vt.desc[vol];
        const Dims d = V.d;
        const SegLists L{segMask, segBase, V.S, V.rsBase};
        const uint32_t* __restrict__ rowBaseV = rowBaseV0 + V.rowBase;
        int i, j, k;
        uint32_t r;
        cell_ijk(acell[a], d, i, j, k, r);
        const unsigned icase = acode[a] & 0xFFu;
        const int bit = (icase & 1u) ? 0 : 1;     // field at the edge origin below iso
        const int parity = (i & 1) | (j & 1) << 1 | (k & 1) << 2;
        int64_t q = (int64_t)rowBaseQ[V.rowBase + r] + aqloc[a];
        const bool ex = edge_cut(icase, 0) && j >= 1 && k >= 1;
        const bool ey = edge_cut(icase, 3) && i >= 1 && k >= 1;
        const bool ez = edge_cut(icase, 8) && i >= 1 && j >= 1;
        if (!(ex || ey || ez)) continue;
        int s0, s3, s8;
        const int32_t own0 = vertex_slot_of(acode, aaux, avloc, rowBaseV, a, 0, r, s0);
        const int32_t own3 = vertex_slot_of(acode, aaux, avloc, rowBaseV, a, 3, r, s3);
        const int32_t own8 = vertex_slot_of(acode, aaux, avloc, rowBaseV, a, 8, r, s8);
        // rows of the cells (., j-1, k) and (., j, k-1)
        const int64_t rj = (int64_t)r - 1, rk = (int64_t)r - (d.ny - 1);
        // cells of the previous row (j-1) and plane (k-1) found once
        const int64_t cjm = (ex || ez) ? find_active(L, d, i, j - 1, k) : -1;
        const int64_t ckm = (ex || ey) ? find_active(L, d, i, j, k - 1) : -1;
        int32_t out[4];
        int sl[4];
        // quadrilateral q with its vertices in out[]: store it, and each of
        // its halfedges (4q+p: out[p] -> out[p+1]) in its origin's CSR row
        auto emit = [&](int c) {
            if (out[0] < 0 || out[1] < 0 || out[2] < 0 || out[3] < 0) {
                atomicAdd(unresolved, 1u);
            } else {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    // rows total 4 Q (checked by the host); never past the array
                    const uint64_t at = (uint64_t)off[out[p]] + (uint64_t)sl[p];
                    if (at < adjcap) adj[at] = make_int2((int32_t)(4 * q + p), out[(p + 1) & 3]);
                }
            }
            reinterpret_cast<int4*>(quads)[q] = make_int4(out[0], out[1], out[2], out[3]);
            colors[q] = (uint8_t)(3 * parity + c);
            q++;
        };
        // x edge: cells (i,j,k) e0, (i,j-1,k) e2, (i,j,k-1) e4, (i,j-1,k-1) e6
        if (ex) {
            const int64_t e = find_active(L, d, i, j - 1, k - 1);
            out[edge_slot(0, bit)] = own0;
            sl[edge_slot(0, bit)] = s0;
            out[edge_slot(2, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, cjm, 2, rj, sl[edge_slot(2, bit)]);
            out[edge_slot(4, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, ckm, 4, rk, sl[edge_slot(4, bit)]);
            out[edge_slot(6, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, e, 6, rk - 1, sl[edge_slot(6, bit)]);
            emit(0);
        }
        // y edge: cells (i,j,k) e3, (i-1,j,k) e1, (i,j,k-1) e7, (i-1,j,k-1) e5
        if (ey) {
            const int64_t e = ckm > 0 ? ckm - 1 : -1;
            out[edge_slot(3, bit)] = own3;
            sl[edge_slot(3, bit)] = s3;
            out[edge_slot(1, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, a - 1, 1, r, sl[edge_slot(1, bit)]);
            out[edge_slot(7, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, ckm, 7, rk, sl[edge_slot(7, bit)]);
            out[edge_slot(5, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, e, 5, rk, sl[edge_slot(5, bit)]);
            emit(1);
        }
        // z edge: cells (i,j,k) e8, (i-1,j,k) e9, (i,j-1,k) e11, (i-1,j-1,k) e10
        if (ez) {
            const int64_t e = cjm > 0 ? cjm - 1 : -1;
            out[edge_slot(8, bit)] = own8;
            sl[edge_slot(8, bit)] = s8;
            out[edge_slot(9, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, a - 1, 9, r, sl[edge_slot(9, bit)]);
            out[edge_slot(11, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, cjm, 11, rj, sl[edge_slot(11, bit)]);
            out[edge_slot(10, bit)] = vertex_slot_of(acode, aaux, avloc, rowBaseV, e, 10, rj, sl[edge_slot(10, bit)]);
            emit(2);
        }
    }
}

// ---------------------------------------------------------------------------
// Per-volume bookkeeping of a batch: first active cell, vertex and
// quadrilateral of each volume (n+1 entries, the last one the total), kept
// up to date through order-preserving compactions, and the final rebase of
// each volume's ids to start at 0.
// ---------------------------------------------------------------------------
__global__ void k_vol_abase(VolTable vt, const uint32_t* __restrict__ segBase, const uint64_t* A,
                            int64_t* __restrict__ aBase) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v <= vt.n; v += gridDim.x * blockDim.x)
        aBase[v] = v < vt.n ? (int64_t)segBase[vt.desc[v].rsBase] : (int64_t)*A;
}

// After the row scans: which capacities the cell stage exceeded.  A pass
// that exceeded one processes an empty mesh from here on (acnt = nv = nq =
// 0) and is run again by the host with room for the exact sizes.
__global__ void k_cell_guard(Sizes* S, uint64_t acap, uint64_t vcap, uint64_t qcap, uint32_t* slow) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        slow[0] = 0;   // k_vertices' count of deferred cells
        const uint32_t o = (S->A > acap ? 1u : 0u) | (S->V > vcap ? 2u : 0u) | (S->Q > qcap ? 4u : 0u);
        S->ovf |= o;
        S->acnt = o ? 0 : S->A;
        S->nv = o ? 0 : S->V;
        S->nq = o ? 0 : S->Q;
    }
}

// Each volume's first vertex and quadrilateral: [0, n] the live bounds,
// [n+1, 2n+1] the cell stage's (kept for the per-volume counts).
__global__ void k_vol_bounds(VolTable vt, const uint32_t* __restrict__ rowBaseV, const uint32_t* __restrict__ rowBaseQ,
                             const Sizes* S, int64_t* __restrict__ vb, int64_t* __restrict__ qb) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v <= vt.n; v += gridDim.x * blockDim.x) {
        const int64_t x = v < vt.n ? (int64_t)rowBaseV[vt.desc[v].rowBase] : (int64_t)S->V;
        const int64_t y = v < vt.n ? (int64_t)rowBaseQ[vt.desc[v].rowBase] : (int64_t)S->Q;
        vb[v] = x;
        qb[v] = y;
        vb[vt.n + 1 + v] = x;
        qb[vt.n + 1 + v] = y;
    }
}

// b[v] -> number of kept elements before it (newIdx = exclusive scan of the
// keep flags over *nOld elements, *nNew kept).
// newIdx: the exclusive scan of the keep flags, or a compaction map
// (ScanKeepMap: ~rank among the removed at a removed item), from which the
// kept count before x is x - ~map[x].
__global__ void k_vol_remap(int64_t* __restrict__ b, int n1, const uint32_t* __restrict__ newIdx,
                            const uint64_t* nOld, const uint64_t* nNew) {
    const int64_t no = (int64_t)*nOld, nn = (int64_t)*nNew;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n1; v += gridDim.x * blockDim.x) {
        const int64_t x = b[v];
        const int32_t m = x < no ? (int32_t)newIdx[x] : 0;
        b[v] = x < no ? (m >= 0 ? (int64_t)m : x - (int64_t)~m) : nn;
    }
}

__global__ void __launch_bounds__(kBlock) k_rebase(const int64_t* __restrict__ vb, const int64_t* __restrict__ qb,
                                                    int n, int32_t* __restrict__ quads, int32_t* __restrict__ twin,
                                                    int32_t* __restrict__ tris, const Sizes* S,
                                                    int32_t* __restrict__ hev) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    const int64_t m = nq > nv ? nq : nv;
    for (int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x; x < m; x += (int64_t)gridDim.x * kBlock) {
        if (x < nq) {
            const int vol = vol_of(qb, n, x);
            const int32_t v0 = (int32_t)vb[vol], h0 = (int32_t)(4 * qb[vol]);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                quads[4 * x + k] -= v0;
                const int32_t t = twin[4 * x + k];
                if (t >= 0) twin[4 * x + k] = t - h0;
            }
            if (tris)
#pragma unroll
                for (int k = 0; k < 6; k++) tris[6 * x + k] -= v0;
        }
        if (x < nv) {
            const int vol = vol_of(vb, n, x);
            const int32_t e = hev[x];
            if (e >= 0) hev[x] = e - (int32_t)(4 * qb[vol]);
        }
    }
}

}  // namespace tmc
