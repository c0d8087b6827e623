// dmc_mesh_kernels.h -- quad-mesh stages of the DMC engine: halfedges,
// vertex valence, face colouring, 3X3Y and 3333 simplification, triangles.
//
// Halfedge h = 4q + k runs from quads[4q+k] to quads[4q+(k+1)%4]; face and
// next are implicit, twins and flags are stored (4 B + 1 B per halfedge).
// Twins are found through a per-vertex CSR of outgoing (halfedge,
// destination) pairs instead of the reference's map keyed by the vertex pair
// (cpp_dmc/DualMarchingCubes.cpp:488-671): the halfedges sharing an edge
// {a, b} are the out-halfedges of a ending at b plus those of b ending at a,
// taken in ascending halfedge order, which is the map's insertion order.
#pragma once
#include "dmc_kernels.h"

namespace tmc {

__device__ __forceinline__ int32_t he_next(int32_t h) { return (h & ~3) | ((h + 1) & 3); }
__device__ __forceinline__ int32_t he_prev(int32_t h) { return (h & ~3) | ((h + 3) & 3); }

#define TMC_GRID_STRIDE(idx, n) \
    for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < (n); idx += (int64_t)gridDim.x * kBlock)

// Quad vertex ids must be in [0, nv); anything else (possible only after a
// degenerate simplification) is counted and skipped.
__device__ __forceinline__ bool vid_ok(int32_t v, int64_t nv) { return v >= 0 && (int64_t)v < nv; }

__global__ void __launch_bounds__(kBlock) k_he_degree(const int32_t* __restrict__ quads, const Sizes* S,
                                                       uint32_t* __restrict__ deg, uint32_t* __restrict__ bad,
                                                       const uint32_t* __restrict__ gate) {
    if (gated_off(gate)) return;
    const int64_t nhe = 4 * (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(h, nhe) {
        const int32_t v = quads[h];
        if (vid_ok(v, nv)) atomicAdd(&deg[v], 1u);
        else atomicAdd(bad, 1u);
    }
}

// Scatter (halfedge, destination) pairs into the CSR rows; the order
// inside a row is arbitrary (the group search sorts what it needs).
__global__ void __launch_bounds__(kBlock) k_he_scatter(const int32_t* __restrict__ quads, const Sizes* S,
                                                        const uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
                                                        int2* __restrict__ adj, const uint32_t* __restrict__ gate) {
    if (gated_off(gate)) return;
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        const int32_t v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int32_t a = v[k];
            if (!vid_ok(a, nv)) continue;
            const uint32_t p = off[a] + atomicAdd(&cursor[a], 1u);
            adj[p] = make_int2((int32_t)(4 * f + k), v[(k + 1) & 3]);
        }
    }
}

__device__ __forceinline__ int32_t he_dest(const int32_t* __restrict__ quads, int32_t h) { return quads[he_next(h)]; }

// Visit the entries of CSR row [s, e) four at a time: the four loads of a
// chunk are issued together (indices clamped into the row), then used.
template <class F>
__device__ __forceinline__ void for_row(const int2* __restrict__ adj, uint32_t s, uint32_t e, F&& fn) {
    for (uint32_t x = s; x < e; x += 4) {
        int2 r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = adj[min(x + (uint32_t)i, e - 1)];
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (x + (uint32_t)i < e) fn(r[i]);
    }
}

// Keep the 5 smallest members in ascending order (unrolled insertion).
__device__ __forceinline__ void group_insert(int32_t g[5], int32_t h) {
#pragma unroll
    for (int p = 0; p < 5; p++) {
        if (h < g[p]) {
            const int32_t t = g[p];
            g[p] = h;
            h = t;
        }
    }
}

// Members of the edge group of halfedge h = (a -> b): out-halfedges of a
// ending at b and of b ending at a, in ascending order (5 smallest kept);
// the true count is returned.  The CSR rows are off[v] .. off[v+1].
__device__ __forceinline__ int he_group(const int32_t* __restrict__ quads, const uint32_t* __restrict__ off,
                                        const int2* __restrict__ adj, int32_t h, int64_t nv, int32_t g[5]) {
#pragma unroll
    for (int p = 0; p < 5; p++) g[p] = 0x7FFFFFFF;
    const int32_t a = quads[h], b = he_dest(quads, h);
    if (!vid_ok(a, nv) || !vid_ok(b, nv)) return 0;
    int c = 0;
    const uint32_t sa = off[a], ea = off[a + 1], sb = off[b], eb = off[b + 1];
    for_row(adj, sa, ea, [&](int2 r) {
        if (r.y == b) {
            group_insert(g, r.x);
            c++;
        }
    });
    if (b != a)
        for_row(adj, sb, eb, [&](int2 r) {
            if (r.y == a) {
                group_insert(g, r.x);
                c++;
            }
        });
    return c;
}

// Flag bytes are the reference's: 0 manifold, 1 non-manifold (4 faces),
// 2 boundary multiple edge (3 faces).  A group of more than 4 faces (the
// reference's error case) keeps 0 in the output; inside the engine its
// members carry kFlagErr (its first member kFlagErrFirst) so that a build
// that carries the group over can still count it.  Every test of a flag
// compares with 1 or 2, which the markers never equal; the final build
// clears them.
constexpr uint8_t kFlagErr = 0x80, kFlagErrFirst = 0xC0;

// Twin and flag from a sorted group (cpp_dmc/DualMarchingCubes.cpp:574-669).
__device__ __forceinline__ void twin_rule(const int32_t* __restrict__ quads, int c, const int32_t g[5], int32_t h,
                                          int32_t& tw, uint8_t& fl) {
    tw = -1;
    fl = c > 4 ? (h == g[0] ? kFlagErrFirst : kFlagErr) : 0;
    if (c == 2) {
        tw = (g[0] == h) ? g[1] : g[0];
    } else if (c == 3) {
        fl = 2;
        const int32_t v0 = quads[g[0]], v1 = quads[g[1]], v2 = quads[g[2]];
        int32_t pa, pb;
        if (v0 != v1) { pa = g[0]; pb = g[1]; }
        else if (v0 != v2) { pa = g[0]; pb = g[2]; }
        else { pa = g[1]; pb = g[2]; }
        tw = (h == pa) ? pb : (h == pb) ? pa : -1;
    } else if (c == 4) {
        fl = 1;
        if (quads[g[0]] != quads[g[1]]) tw = (h == g[0]) ? g[1] : (h == g[1]) ? g[0] : (h == g[2]) ? g[3] : g[2];
        else tw = (h == g[0]) ? g[2] : (h == g[2]) ? g[0] : (h == g[1]) ? g[3] : g[1];
    }
}

__device__ __forceinline__ uint64_t he_key(int32_t v0, int32_t v1) {
    return v0 < v1 ? ((uint64_t)(uint32_t)v0 << 32) | (uint32_t)v1 : ((uint64_t)(uint32_t)v1 << 32) | (uint32_t)v0;
}

// A 3-face edge (cpp_dmc/DualMarchingCubes.cpp:603-637): its map key, the
// two boundary vertices (getBndVertices), and the unpaired halfedge, which
// overrides he_v of its origin -- the map is walked in key order, so the
// largest key wins.  Written by the group's first member.
struct BndRec {
    unsigned long long key;
    int32_t a, b;
    int32_t lone, pad;
};

__device__ __forceinline__ BndRec three_record(const int32_t* __restrict__ quads, const int32_t g[5], int32_t h) {
    const int32_t v0 = quads[g[0]], v1 = quads[g[1]], v2 = quads[g[2]];
    BndRec r;
    if (v0 != v1) { r.a = v0; r.b = v1; r.lone = g[2]; }
    else if (v0 != v2) { r.a = v0; r.b = v2; r.lone = g[1]; }
    else { r.a = v1; r.b = v2; r.lone = g[0]; }
    r.key = he_key(quads[h], he_dest(quads, h));
    r.pad = 0;
    return r;
}

// A record appended to recs (order free: the host sorts them by key).  The
// lanes of a wave that append together take their slots with one add.
__device__ __forceinline__ void put_record(BndRec* __restrict__ recs, uint32_t* nrec, uint32_t reccap, uint32_t* ovf,
                                           const BndRec& r) {
    const uint64_t m = __ballot(1);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(nrec, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader, kWave);
    const uint32_t p = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (p < reccap) recs[p] = r;
    else atomicOr(ovf, 8u);
}

// Per-thread counts of flagged halfedges, added to the build's statistics
// once per wave (every lane of the wave must call flush).
struct HeCount {
    uint32_t n4 = 0, n3 = 0, err = 0;
    __device__ __forceinline__ void add(uint8_t fl) {
        n4 += fl == 1;
        n3 += fl == 2;
        err += fl == kFlagErrFirst;
    }
    // One add per block and counter, into the block's slot: the waves' sums
    // meet in LDS.  Every thread of the block calls this.
    __device__ __forceinline__ void flush(unsigned long long (*hep)[3]) const {
        __shared__ uint32_t sm[3];
        if (threadIdx.x < 3) sm[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t a = wave_sum(n4), b = wave_sum(n3), e = wave_sum(err);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            if (a) atomicAdd(&sm[0], a);
            if (b) atomicAdd(&sm[1], b);
            if (e) atomicAdd(&sm[2], e);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long* slot = hep[blockIdx.x & (kHeParts - 1)];
            if (sm[0]) atomicAdd(&slot[0], (unsigned long long)sm[0]);
            if (sm[1]) atomicAdd(&slot[1], (unsigned long long)sm[1]);
            if (sm[2]) atomicAdd(&slot[2], (unsigned long long)sm[2]);
        }
    }
};

// Per volume of a batch (rare events): [3v] non-manifold halfedges, [3v+1]
// 3-face halfedges, [3v+2] groups of more than 4.
__device__ __forceinline__ void vol_count(unsigned long long* __restrict__ vst, const int64_t* __restrict__ qb,
                                          int nvol, int32_t h, uint8_t fl) {
    if (!vst || (fl != 1 && fl != 2 && fl != kFlagErrFirst)) return;
    const int vol = vol_of(qb, nvol, h >> 2);
    atomicAdd(&vst[3 * vol + (fl == 1 ? 0 : fl == 2 ? 1 : 2)], 1ull);
}

// Twin and flag of every halfedge (cpp_dmc/DualMarchingCubes.cpp:574-669),
// one thread per halfedge.  The four halfedges of a quadrilateral sit in
// four adjacent lanes (4q+k in lane 4m+k), and halfedge k needs the rows of
// its origin v[k] and of v[k+1] -- the row lane k+1 loads for itself.  So
// every lane loads only its own row (up to kRowRegs entries in registers)
// and takes its neighbour's with a DPP quad permutation; longer rows are
// finished from memory.  Each lane also counts its flag, adds a boundary
// arrival to its destination's valence, and writes he_v of its origin if
// it is the row's largest halfedge; the first member of a 3-face group
// writes the group's record.
constexpr int kRowRegs = 6;

__device__ __forceinline__ int next_in_quad(int x) {
    // quad_perm(1, 2, 3, 0): lane k receives lane (k+1) & 3 of its quad
    return __builtin_amdgcn_update_dpp(0, x, 0x39, 0xF, 0xF, false);
}

__global__ void __launch_bounds__(kBlock) k_he_twins(const int32_t* __restrict__ quads, const Sizes* S,
                                                      const uint32_t* __restrict__ off,
                                                      const int2* __restrict__ adj, int32_t* __restrict__ twin,
                                                      uint8_t* __restrict__ flag, int32_t* __restrict__ hev,
                                                      uint32_t* __restrict__ arr, unsigned long long (*hep)[3],
                                                      BndRec* __restrict__ recs, uint32_t* nrec, uint32_t reccap,
                                                      uint32_t* ovf, const int64_t* __restrict__ qb, int nvol,
                                                      unsigned long long* __restrict__ vst,
                                                      const uint32_t* __restrict__ gate) {
    static_assert(kBlock % 4 == 0, "quads must not straddle blocks");
    if (gated_off(gate)) return;
    const int64_t nhe = 4 * (int64_t)S->nq, nv = (int64_t)S->nv;
    HeCount cnt;
    TMC_GRID_STRIDE(h64, nhe) {
        const int32_t h = (int32_t)h64;
        const int32_t a = quads[h];
        const bool aok = vid_ok(a, nv);
        uint32_t s = 0, n = 0;
        if (aok) {
            s = off[a];
            n = off[a + 1] - s;
        }
        int2 r[kRowRegs];
        if (n > 0) {
#pragma unroll
            for (int i = 0; i < kRowRegs; i++) r[i] = adj[s + min((uint32_t)i, n - 1)];
        } else {
#pragma unroll
            for (int i = 0; i < kRowRegs; i++) r[i] = make_int2(-1, -1);
        }
        // he_v (cpp_dmc/DualMarchingCubes.cpp:561-564): the largest halfedge
        // of the row; its own lane writes it
        int32_t hmax = -1;
#pragma unroll
        for (int i = 0; i < kRowRegs; i++) hmax = max(hmax, r[i].x);
        for (uint32_t x = kRowRegs; x < n; x++) hmax = max(hmax, adj[s + x].x);
        if (n > 0 && h == hmax) hev[a] = h;
        // the next lane of the quad: its origin is this halfedge's destination
        const int32_t b = next_in_quad(a);
        const uint32_t nb = (uint32_t)next_in_quad((int)n);
        int2 q[kRowRegs];
#pragma unroll
        for (int i = 0; i < kRowRegs; i++) q[i] = make_int2(next_in_quad(r[i].x), next_in_quad(r[i].y));
        int32_t g[5];
#pragma unroll
        for (int p = 0; p < 5; p++) g[p] = 0x7FFFFFFF;
        int c = 0;
        // The common edge: both rows fit the registers, one a -> b entry and
        // one b -> a entry.  Its group is those two, in ascending order --
        // what the sorted inserts below would build, at a fraction of the
        // VALU work.
        uint32_t m1 = 0, m2 = 0;
        int32_t x1 = 0, x2 = 0;
#pragma unroll
        for (int i = 0; i < kRowRegs; i++)
            if ((uint32_t)i < n && r[i].y == b) {
                m1 |= 1u << i;
                x1 = r[i].x;
            }
#pragma unroll
        for (int i = 0; i < kRowRegs; i++)
            if ((uint32_t)i < nb && q[i].y == a) {
                m2 |= 1u << i;
                x2 = q[i].x;
            }
        const bool pair = aok && vid_ok(b, nv) && b != a && n <= (uint32_t)kRowRegs && nb <= (uint32_t)kRowRegs &&
                          __popc(m1) == 1 && __popc(m2) == 1;
        if (pair) {
            g[0] = min(x1, x2);
            g[1] = max(x1, x2);
            c = 2;
        } else if (aok && vid_ok(b, nv)) {
#pragma unroll
            for (int i = 0; i < kRowRegs; i++)
                if ((uint32_t)i < n && r[i].y == b) {
                    group_insert(g, r[i].x);
                    c++;
                }
            for (uint32_t x = kRowRegs; x < n; x++) {
                const int2 e = adj[s + x];
                if (e.y == b) {
                    group_insert(g, e.x);
                    c++;
                }
            }
            if (b != a) {
#pragma unroll
                for (int i = 0; i < kRowRegs; i++)
                    if ((uint32_t)i < nb && q[i].y == a) {
                        group_insert(g, q[i].x);
                        c++;
                    }
                if (nb > kRowRegs) {
                    const uint32_t sb = off[b];
                    for (uint32_t x = kRowRegs; x < nb; x++) {
                        const int2 e = adj[sb + x];
                        if (e.y == a) {
                            group_insert(g, e.x);
                            c++;
                        }
                    }
                }
            }
        }
        int32_t tw;
        uint8_t fl;
        twin_rule(quads, c, g, h, tw, fl);
        twin[h] = tw;
        flag[h] = fl;
        // valence: a boundary halfedge adds one to the vertex it ends at
        if (tw == -1 && arr && vid_ok(b, nv)) atomicAdd(&arr[b], 1u);
        cnt.add(fl);
        vol_count(vst, qb, nvol, h, fl);
        if (c == 3 && g[0] == h) put_record(recs, nrec, reccap, ovf, three_record(quads, g, h));
    }
    cnt.flush(hep);
}

// The 3-face overrides of he_v (cpp_dmc/DualMarchingCubes.cpp:603-637) from
// the final build's records: pass 0 finds each lone vertex's largest key,
// pass 1 lets that record write.  ovr (2 words per vertex) starts at 0.
__global__ void __launch_bounds__(kBlock) k_three_apply(const BndRec* __restrict__ recs, const Sizes* S,
                                                         uint32_t reccap, const int32_t* __restrict__ quads,
                                                         unsigned long long* __restrict__ ovr, int pass,
                                                         int32_t* __restrict__ he_v) {
    const int64_t n = S->nrec < reccap ? S->nrec : reccap, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(i, n) {
        const BndRec r = recs[i];
        const int32_t vl = quads[r.lone];
        if (!vid_ok(vl, nv)) continue;
        if (pass == 0) atomicMax(&ovr[vl], r.key + 1ull);
        else if (ovr[vl] == r.key + 1ull) he_v[vl] = r.lone;
    }
}

// The final mesh's flags as the reference leaves them: the markers of
// groups of more than 4 faces cleared (only when there are any).
__global__ void __launch_bounds__(kBlock) k_strip_flags(const Sizes* S, uint8_t* __restrict__ flag) {
    if (he_errors(S) == 0) return;
    const int64_t nhe = 4 * (int64_t)S->nq;
    TMC_GRID_STRIDE(h, nhe) {
        if (flag[h] & kFlagErr) flag[h] = 0;
    }
}

// Arrays the simplification after a halfedge build needs cleared, cleared
// by the build's valence pass (one launch fewer per stage): bytes to 0 and
// words to 0xFFFFFFFF per vertex, bytes to 0 per face (null: none).
struct StageClears {
    uint8_t* v8 = nullptr;
    uint32_t* v32[3] = {nullptr, nullptr, nullptr};
    uint8_t* f8[2] = {nullptr, nullptr};
    int64_t vcap = 0, qcap = 0;
};

// Valence (cpp_dmc/DualMarchingCubes.cpp:847-871): outgoing halfedges plus
// boundary halfedges ending at the vertex -- the out-degree and the
// arrivals the pairing counted.  Counts the vertices of valence 0.
__global__ void __launch_bounds__(kBlock) k_valence(const uint32_t* __restrict__ deg,
                                                     const uint32_t* __restrict__ arr, const Sizes* S,
                                                     uint32_t* __restrict__ val, unsigned int* zero,
                                                     StageClears c) {
    const int64_t nv = (int64_t)S->nv;
    unsigned int z = 0;
    TMC_GRID_STRIDE(v, nv) {
        const uint32_t x = deg[v] + arr[v];
        val[v] = x;
        z += x == 0;
    }
    z = wave_sum(z);
    if ((threadIdx.x & (kWave - 1)) == 0 && z) atomicAdd(zero, z);
    const int64_t nc = nv < c.vcap ? nv : c.vcap;
    if (c.v8) TMC_GRID_STRIDE(v, nc) c.v8[v] = 0;
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (c.v32[k]) TMC_GRID_STRIDE(v, nc) c.v32[k][v] = 0xFFFFFFFFu;
    const int64_t nq = (int64_t)S->nq < c.qcap ? (int64_t)S->nq : c.qcap;
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (c.f8[k]) TMC_GRID_STRIDE(f, nq) c.f8[k][f] = 0;
}

// The four edge neighbours of face f (-1 where an edge has no twin).
__device__ __forceinline__ int4 face_neighbors(const int32_t* __restrict__ twin, int64_t f) {
    const int4 t = reinterpret_cast<const int4*>(twin)[f];
    return make_int4(t.x >= 0 ? t.x >> 2 : -1, t.y >= 0 ? t.y >> 2 : -1, t.z >= 0 ? t.z >> 2 : -1,
                     t.w >= 0 ? t.w >> 2 : -1);
}

// First colour in 0..ncol-1 that none of the neighbours nb has.
__device__ __forceinline__ void take_free_colour(const int4 nb, int ncol, int32_t f, uint8_t* __restrict__ colors) {
    const int c0 = nb.x >= 0 ? colors[nb.x] : -1, c1 = nb.y >= 0 ? colors[nb.y] : -1;
    const int c2 = nb.z >= 0 ? colors[nb.z] : -1, c3 = nb.w >= 0 ? colors[nb.w] : -1;
    for (int col = 0; col < ncol; col++) {
        if (c0 != col && c1 != col && c2 != col && c3 != col) {
            colors[f] = (uint8_t)col;
            break;
        }
    }
}

// Face colouring (cpp_dmc/DualMarchingCubes.cpp:714-844).  Colours are
// bytes (initial colours 0..23, final ones 0..4).  Faces are bucketed by
// initial colour; classes 5..23 are processed in order, each face taking the
// first colour in 0..4 unused by its edge neighbours, then the initially
// colour-4 faces try 0..3.  Faces of one initial colour are never edge
// neighbours (their grid edges are parallel edges of different parity), so
// each class is processed in parallel.
// Bucketing by initial colour without hot global counters: each block keeps
// an LDS histogram of its grid-stride range, one block turns the histograms
// into per-(block, colour) offsets, and the same blocks scatter again,
// writing each face's id and its neighbours' ids (one int4) in class order,
// so that a class reads its faces' neighbours contiguously instead of one
// twin row in every 24.  The order inside a class does not matter (its
// faces are independent).
constexpr int kColorBlocks = 1024;

__global__ void __launch_bounds__(kBlock) k_color_hist(const uint8_t* __restrict__ colors, const Sizes* S,
                                                        uint32_t* __restrict__ hist) {
    __shared__ uint32_t sm[24];
    const int64_t nq = (int64_t)S->nq;
    if (threadIdx.x < 24) sm[threadIdx.x] = 0;
    __syncthreads();
    TMC_GRID_STRIDE(f, nq) {
        const int c = colors[f];
        if (c < 24) atomicAdd(&sm[c], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 24) hist[blockIdx.x * 24 + threadIdx.x] = sm[threadIdx.x];
}

// One block of 1024 threads (16 waves): totals per colour, class starts, and
// per-(block, colour) offsets in place of the histogram.  Wave w takes
// colours w and w + 16; each lane holds the counts of blocks lane + 64 q
// (nblocks <= kColorBlocks = 1024), so every load is issued before the
// wave's scans.
__global__ void __launch_bounds__(1024) k_color_offsets(uint32_t* __restrict__ hist, int nblocks,
                                                         uint32_t* __restrict__ totals) {
    static_assert(kColorBlocks <= 16 * kWave, "16 counts per lane");
    __shared__ uint32_t tot[24], start[24];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t v[2][16];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int c = w + 16 * h;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int b = lane + kWave * q;
            v[h][q] = (c < 24 && b < nblocks) ? hist[b * 24 + c] : 0u;
        }
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) sum += v[h][q];
        sum = wave_sum(sum);
        if (c < 24 && lane == 0) tot[c] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int c = 0; c < 24; c++) {
            start[c] = run;
            totals[c] = tot[c];
            totals[24 + c] = run;
            run += tot[c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int c = w + 16 * h;
        if (c >= 24) continue;   // wave-uniform
        uint32_t run = start[c];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int b = lane + kWave * q;
            const uint32_t inc = wave_inclusive_scan(v[h][q]);
            if (b < nblocks) hist[b * 24 + c] = run + inc - v[h][q];
            run += __shfl(inc, kWave - 1, kWave);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_color_scatter(const uint8_t* __restrict__ colors, const Sizes* S,
                                                           const uint32_t* __restrict__ offs,
                                                           const int32_t* __restrict__ twin,
                                                           int32_t* __restrict__ list, int4* __restrict__ nbr) {
    __shared__ uint32_t cur[24];
    const int64_t nq = (int64_t)S->nq;
    if (threadIdx.x < 24) cur[threadIdx.x] = offs[blockIdx.x * 24 + threadIdx.x];
    __syncthreads();
    TMC_GRID_STRIDE(f, nq) {
        const int c = colors[f];
        if (c >= 24) continue;
        const uint32_t at = atomicAdd(&cur[c], 1u);
        list[at] = (int32_t)f;
        if (c >= 4) nbr[at] = face_neighbors(twin, f);   // classes 0..3 are never recoloured
    }
}

// Classes ca and cb (cb < 0: none) in one launch, two classes with no edge
// neighbours in common: their counts and starts in totals
// (k_color_offsets).  Class 4 takes a colour in 0..3, the others in 0..4.
__global__ void __launch_bounds__(kBlock) k_color_class(const int32_t* __restrict__ list,
                                                         const int4* __restrict__ nbr,
                                                         const uint32_t* __restrict__ totals, int ca, int cb,
                                                         uint8_t* __restrict__ colors) {
    const int64_t na = totals[ca], nb2 = cb >= 0 ? totals[cb] : 0;
    TMC_GRID_STRIDE(x, na + nb2) {
        const bool first = x < na;
        const int ncol = (first ? ca : cb) == 4 ? 4 : 5;
        const int64_t at = first ? (int64_t)totals[24 + ca] + x : (int64_t)totals[24 + cb] + (x - na);
        take_free_colour(nbr[at], ncol, list[at], colors);
    }
}

// All classes in one block, in order, for a small mesh (one launch instead
// of twenty): the same first-free-colour rule, a barrier between classes.
__global__ void __launch_bounds__(1024) k_color_all(const int32_t* __restrict__ list,
                                                     const int4* __restrict__ nbr,
                                                     const uint32_t* __restrict__ totals,
                                                     uint8_t* __restrict__ colors) {
    for (int step = 0; step < 20; step++) {
        const int cls = step < 19 ? 5 + step : 4;
        const int ncol = step < 19 ? 5 : 4;
        const int64_t n = totals[cls], start = totals[24 + cls];
        for (int64_t x = threadIdx.x; x < n; x += blockDim.x)
            take_free_colour(nbr[start + x], ncol, list[start + x], colors);
        // the next class reads this one's colours
        __threadfence_block();
        __syncthreads();
    }
}

// 3X3Y pattern (cpp_dmc/DualMarchingCubes.cpp:889-913)
__global__ void __launch_bounds__(kBlock) k_mark3x3y(const int32_t* __restrict__ quads, const Sizes* S,
                                                      const int32_t* __restrict__ val, uint8_t* __restrict__ p) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        bool ok = vid_ok(q.x, nv) && vid_ok(q.y, nv) && vid_ok(q.z, nv) && vid_ok(q.w, nv);
        uint8_t r = 0;
        if (ok) {
            const int a = val[q.x], b = val[q.y], c = val[q.z], d = val[q.w];
            r = (a == 3 && b >= 5 && c == 3 && d >= 5) || (a >= 5 && b == 3 && c >= 5 && d == 3);
        }
        p[f] = r;
    }
}

__device__ __forceinline__ bool face_nonmanifold(const uint8_t* __restrict__ flag, int32_t f) {
    return flag[4 * f] == 1 || flag[4 * f + 1] == 1 || flag[4 * f + 2] == 1 || flag[4 * f + 3] == 1;
}

// any of face g's four halfedges flagged NON_MANIFOLD (flag bytes 4g..4g+3)
__device__ __forceinline__ bool flags_nonmanifold(uint32_t w) {
    return (w & 0xFFu) == 1u || ((w >> 8) & 0xFFu) == 1u || ((w >> 16) & 0xFFu) == 1u || (w >> 24) == 1u;
}

// Decision of mergeVertices3X3Y for face f (colored: :916-1012, isolated:
// :1014-1094).  Depends only on data fixed before the merge, so all faces
// decide in parallel.  Returns 0 (keep), 1 (move v0, drop v2), 2 (move v1,
// drop v3).  Every lookup of the neighbours is issued before any is tested.
__device__ __forceinline__ int merge_decision(int64_t f, const int32_t* __restrict__ quads,
                                              const int32_t* __restrict__ val, const uint8_t* __restrict__ p,
                                              const uint8_t* __restrict__ colors, const int32_t* __restrict__ twin,
                                              const uint8_t* __restrict__ flag, int64_t nv) {
    if (!p[f]) return 0;
    const int4 tw = reinterpret_cast<const int4*>(twin)[f];
    const int32_t nb[4] = {tw.x >= 0 ? tw.x >> 2 : -1, tw.y >= 0 ? tw.y >> 2 : -1, tw.z >= 0 ? tw.z >> 2 : -1,
                           tw.w >= 0 ? tw.w >> 2 : -1};
    const int4 q = reinterpret_cast<const int4*>(quads)[f];
    uint32_t fw[4];
    uint8_t pn[4];
    int cn[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int32_t g = nb[k] >= 0 ? nb[k] : (int32_t)f;   // any valid face for a missing neighbour
        fw[k] = reinterpret_cast<const uint32_t*>(flag)[g];
        pn[k] = p[g];
        cn[k] = colors ? colors[g] : 0;
    }
    const int a = val[q.x], b = val[q.y], c = val[q.z], d = val[q.w];
    if (colors) {
        const int cf = colors[f];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (nb[k] < 0 || flags_nonmanifold(fw[k])) return 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (pn[k] && cn[k] <= cf) return 0;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (nb[k] >= 0 && flags_nonmanifold(fw[k])) return 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (nb[k] >= 0 && pn[k]) return 0;
    }
    if (a == 3 && b >= 5 && c == 3 && d >= 5) return 1;
    if (a >= 5 && b == 3 && c >= 5 && d == 3) return 2;
    return 0;
}

// Pass A: every merging face claims its moved and dropped vertex; the
// reference loops over faces in order, so the largest face index wins.
__global__ void __launch_bounds__(kBlock) k_merge_claim(const int32_t* __restrict__ quads, const Sizes* S,
                                                         const int32_t* __restrict__ val, const uint8_t* __restrict__ p,
                                                         const uint8_t* __restrict__ colors,
                                                         const int32_t* __restrict__ twin,
                                                         const uint8_t* __restrict__ flag, int32_t* __restrict__ lastT,
                                                         int32_t* __restrict__ lastR, uint8_t* __restrict__ em,
                                                         uint8_t* __restrict__ vfirst) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        const int m = merge_decision(f, quads, val, p, colors, twin, flag, nv);
        em[f] = (uint8_t)m;   // 0 kept; 1 / 2 which pair of vertices merges
        if (!m) continue;
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        const int32_t t = m == 1 ? q.x : q.y;
        const int32_t r = m == 1 ? q.z : q.w;
        atomicMax(&lastT[t], (int32_t)f);
        atomicMax(&lastR[r], (int32_t)f);
        vfirst[r] = 1;
    }
}

// Pass B: the winning face writes the midpoint position and the normalised
// mean normal (reads only valence >= 5 vertices, which are never written)
// and the drop map.
__global__ void __launch_bounds__(kBlock) k_merge_apply(const int32_t* __restrict__ quads,
                                                         const int32_t* __restrict__ mlist, const Sizes* S,
                                                         const uint8_t* __restrict__ em,
                                                         const int32_t* __restrict__ val,
                                                         const int32_t* __restrict__ lastT,
                                                         const int32_t* __restrict__ lastR, double* __restrict__ pos,
                                                         double* __restrict__ nrm, const uint32_t* __restrict__ vslot,
                                                         int32_t* __restrict__ vsecond) {
    const int64_t n = (int64_t)(S->nq - S->qn);   // the merging faces
    TMC_GRID_STRIDE(i, n) {
        const int64_t f = mlist[i];
        const int m = em[f];
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        const bool first = m == 1;     // pattern (3, >=5, 3, >=5)
        const int32_t t = first ? q.x : q.y;
        const int32_t r = first ? q.z : q.w;
        const int32_t s0 = first ? q.y : q.x;
        const int32_t s1 = first ? q.w : q.z;
        // storage slots of the vertices (positions and normals are not moved
        // by the compactions until the end); loaded with the claim
        const int32_t lt = lastT[t];
        const int64_t a0 = vslot ? vslot[s0] : s0, a1 = vslot ? vslot[s1] : s1;
        const int64_t at = vslot ? vslot[t] : t;
        if (lt == (int32_t)f) {
            double np[3], nn[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                np[c] = pos[3 * a0 + c] + 0.5 * (pos[3 * a1 + c] - pos[3 * a0 + c]);
                nn[c] = nrm[3 * a0 + c] + 0.5 * (nrm[3 * a1 + c] - nrm[3 * a0 + c]);
            }
            const double s = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            nn[0] /= s;
            nn[1] /= s;
            nn[2] /= s;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                pos[3 * at + c] = np[c];
                nrm[3 * at + c] = nn[c];
            }
        }
        if (lastR[r] == (int32_t)f) vsecond[r] = t;
    }
}

// The flagged items, ascending, at i - newidx[i] (newidx = exclusive scan of
// the keep flags): a dense list without atomics.
__global__ void __launch_bounds__(kBlock) k_list_flagged(const uint8_t* __restrict__ flag,
                                                          const int32_t* __restrict__ map, DevN dn,
                                                          int32_t* __restrict__ list) {
    const int64_t n = dn.get();
    TMC_GRID_STRIDE(i, n) {
        if (flag[i]) list[~map[i]] = (int32_t)i;
    }
}

// Vertex compaction without moving positions and normals: vertex v's data
// lives at slot vslot[v] (the identity when vslot is null); the compaction
// only carries the slots along (4 B per vertex instead of 48 B), and one
// gather at the end lays the survivors out densely.
__global__ void __launch_bounds__(kBlock) k_compact_slots(const uint8_t* __restrict__ removed,
                                                           const uint32_t* __restrict__ newidx, const Sizes* S,
                                                           const uint32_t* __restrict__ vslot,
                                                           uint32_t* __restrict__ vslot2,
                                                           int32_t* __restrict__ vsecond) {
    const int64_t nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(v, nv) {
        if (removed[v]) continue;
        const uint32_t o = newidx[v];
        vslot2[o] = vslot ? vslot[v] : (uint32_t)v;
        if (vsecond) vsecond[v] = (int32_t)o;
    }
}

__global__ void __launch_bounds__(kBlock) k_gather_vertices(const uint32_t* __restrict__ vslot, const Sizes* S,
                                                             const double* __restrict__ pos,
                                                             const double* __restrict__ nrm,
                                                             double* __restrict__ pos2, double* __restrict__ nrm2) {
    const int64_t nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(v, nv) {
        const int64_t s = vslot[v];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pos2[3 * v + c] = pos[3 * s + c];
            nrm2[3 * v + c] = nrm[3 * s + c];
        }
    }
}


// mark3333 (cpp_dmc/DualMarchingCubes.cpp:1190-1301).  Returns true with
// the face's four vertices, outer vertices and neighbour faces if f is a
// removable all-valence-3 face.
__device__ __forceinline__ bool pattern3333(int64_t f, const int32_t* __restrict__ quads,
                                            const int32_t* __restrict__ twin, const uint8_t* __restrict__ flag,
                                            const int32_t* __restrict__ val, int64_t nv, int32_t vq[4],
                                            int32_t vo[4], int32_t fn[4]) {
    if (face_nonmanifold(flag, (int32_t)f)) return false;
    const int4 q = reinterpret_cast<const int4*>(quads)[f];
    vq[0] = q.x; vq[1] = q.y; vq[2] = q.z; vq[3] = q.w;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (!vid_ok(vq[k], nv) || val[vq[k]] != 3) return false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int32_t tw = twin[4 * f + k];
        if (tw == -1) return false;
        fn[k] = tw >> 2;
        const int32_t n2 = he_next(he_next(tw));
        vo[k] = quads[n2];
        if (!vid_ok(vo[k], nv)) return false;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (face_nonmanifold(flag, fn[k])) return false;
    const int a4 = val[vo[0]], a5 = val[vo[1]], a6 = val[vo[2]], a7 = val[vo[3]];
    if ((a4 == 3 && a5 == 3) || (a5 == 3 && a6 == 3) || (a6 == 3 && a7 == 3) || (a7 == 3 && a4 == 3)) return false;
    if (a4 == 3 || a5 == 3 || a6 == 3 || a7 == 3) return false;
    if (a4 == 4 && a5 == 4 && a6 == 4 && a7 == 4) return false;
    return true;
}

__global__ void __launch_bounds__(kBlock) k_mark3333(const int32_t* __restrict__ quads, const Sizes* S,
                                                      const int32_t* __restrict__ twin, const uint8_t* __restrict__ flag,
                                                      const int32_t* __restrict__ val, int32_t* __restrict__ lastF,
                                                      uint8_t* __restrict__ vflag, uint8_t* __restrict__ p3333,
                                                      uint8_t* __restrict__ rflag) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        int32_t vq[4], vo[4], fn[4];
        const bool m = pattern3333(f, quads, twin, flag, val, nv, vq, vo, fn);
        if (!m) continue;
        p3333[f] = 1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            vflag[vq[k]] = 1;
            atomicMax(&lastF[vq[k]], (int32_t)f);
            rflag[fn[k]] = 1;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_twin3333(const int32_t* __restrict__ quads, const Sizes* S,
                                                      const int32_t* __restrict__ twin, const uint8_t* __restrict__ flag,
                                                      const int32_t* __restrict__ val,
                                                      const uint8_t* __restrict__ p3333,
                                                      const int32_t* __restrict__ lastF, int32_t* __restrict__ vtwin) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        if (!p3333[f]) continue;
        int32_t vq[4], vo[4], fn[4];
        pattern3333(f, quads, twin, flag, val, nv, vq, vo, fn);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (lastF[vq[k]] == (int32_t)f) vtwin[vq[k]] = vo[k];
    }
}

// Opens a compaction's carry (thread 0 of the vertex-map kernel, whose other
// threads do not read these): whether the build it starts from met invalid
// vertex references, and fresh counters for the carried build.
__device__ __forceinline__ void open_carry(Sizes* S) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    S->incr_ok = S->he.bad == 0u ? 1u : 0u;
    he_clear(S);
    S->nrec_prev = S->nrec;
    S->recmax = S->nrec > S->recmax ? S->nrec : S->recmax;
    S->nrec = 0;
    S->hq = 4 * S->qn;
    S->cbad = 0;
}

// A kept vertex x hands its out-degree and he_v (renumbered by the face
// map) to its new id y; the changed vertices and those that lost a corner
// are corrected afterwards (k_changed*).
__device__ __forceinline__ void carry_vertex(int64_t x, int32_t y, int64_t nn, const uint32_t* __restrict__ deg,
                                             const int32_t* __restrict__ hev, const int32_t* __restrict__ qmap,
                                             uint32_t* __restrict__ deg_n, int32_t* __restrict__ hev_n) {
    if (!vid_ok(y, nn)) return;
    deg_n[y] = deg[x];
    const int32_t e = hev[x];
    const int32_t m = e < 0 ? -1 : qmap[e >> 2];
    hev_n[y] = m < 0 ? -1 : 4 * m + (e & 3);
}

// The zeroed state of a carry, cleared by the vertex map kernel that starts
// it (which reads none of it): the changed-vertex marks (whole words, set
// with word atomics), the new vertices' arrivals and per-origin counts, and
// the per-turn counts of the compaction.
struct CarryClears {
    uint32_t* cflag;
    int64_t cwords;
    uint32_t* arr;
    uint32_t* cnt;
    uint32_t* bcount;
    int64_t turns;
};

__device__ __forceinline__ void carry_clear(const CarryClears& z, int64_t nn) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = t0; i < z.cwords; i += stride) z.cflag[i] = 0u;
    for (int64_t i = t0; i < nn; i += stride) {
        z.arr[i] = 0u;
        z.cnt[i] = 0u;
    }
    for (int64_t i = t0; i < z.turns; i += stride) z.bcount[i] = 0u;
}

// 3X3Y: vmap[x] = vsecond[x] for a kept x (its new id after the vertex
// compaction), vsecond[vsecond[x]] for a removed one (its target's); and the
// kept vertices' carry.
__global__ void __launch_bounds__(kBlock) k_vmap3x3y(const uint8_t* __restrict__ vfirst,
                                                      const int32_t* __restrict__ vsecond, Sizes* S,
                                                      int32_t* __restrict__ vmap, const uint32_t* __restrict__ deg,
                                                      const int32_t* __restrict__ hev,
                                                      const int32_t* __restrict__ qmap, uint32_t* __restrict__ deg_n,
                                                      int32_t* __restrict__ hev_n, CarryClears z) {
    const int64_t nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    open_carry(S);
    carry_clear(z, nn);
    TMC_GRID_STRIDE(x, nv) {
        int32_t y = vsecond[x];
        const bool removed = vfirst[x] != 0;
        if (removed) y = vid_ok(y, nv) ? vsecond[y] : -1;
        vmap[x] = y;
        if (!removed) carry_vertex(x, y, nn, deg, hev, qmap, deg_n, hev_n);
    }
}

// 3333: vmap[x] = new id of a kept x, -1 for a removed one; and the kept
// vertices' carry.
__global__ void __launch_bounds__(kBlock) k_vmap3333(const uint8_t* __restrict__ vflag,
                                                      const uint32_t* __restrict__ newv, Sizes* S,
                                                      int32_t* __restrict__ vmap, const uint32_t* __restrict__ deg,
                                                      const int32_t* __restrict__ hev,
                                                      const int32_t* __restrict__ qmap, uint32_t* __restrict__ deg_n,
                                                      int32_t* __restrict__ hev_n, CarryClears z) {
    const int64_t nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    open_carry(S);
    carry_clear(z, nn);
    TMC_GRID_STRIDE(x, nv) {
        const bool removed = vflag[x] != 0;
        const int32_t y = removed ? -1 : (int32_t)newv[x];
        vmap[x] = y;
        if (!removed) carry_vertex(x, y, nn, deg, hev, qmap, deg_n, hev_n);
    }
}

// ---------------------------------------------------------------------------
// The halfedges of a compacted mesh, carried over from the build before it.
//
// A compaction keeps face f as face newq[f] (unless dropped) with corner k
// in place, each vertex renamed to its new id.  Call a new vertex changed
// (cflag bit 1) if a removed vertex maps to it, or if it is the image of a
// corner of a dropped face that the compaction moved or merged (3X3Y: the
// two valence-3 corners of a merging face; 3333: every kept corner of a
// dropped face, which covers the pattern face's new corners too).  For a
// new edge {a, b} with neither end changed, a and b each have exactly one
// preimage x, y (kept vertices), every halfedge on {x, y} sits in a
// surviving face that is not a pattern face, and every halfedge on {a, b}
// comes from one of them: the edge's group is the old group renumbered by
// newq, which keeps its order, so its twins (renumbered) and flags carry
// over, and so does its 3-face record.  Halfedges with a changed end are
// paired again through rows of those halfedges only (keyed by origin): a
// group with a changed end has all its members there.
//
// Out-degree and he_v (the largest halfedge leaving the vertex) carry over
// for a vertex with no corner in a dropped face.  A changed vertex has all
// its out-halfedges in its row.  A vertex that only lost corners (the
// valence >= 5 corners of a merging face) loses one out-halfedge per
// corner, and takes he_v again (with atomics over its surviving
// out-halfedges) only if its largest one was dropped (cflag bit 4).
// ---------------------------------------------------------------------------
struct Carry {
    const int32_t* twin;          // the build before the compaction
    const uint8_t* flag;
    const int32_t* qmap;          // new index of each face; negative if dropped (ScanKeepMap)
    const uint8_t* cflag;         // changed new vertices (bit 1), vertices whose he_v is taken again (bit 4)
    int32_t* twin_n;              // the carried build
    uint8_t* flag_n;
    uint32_t* deg_n;
    int32_t* hev_n;
    uint32_t* arr;                // boundary arrivals per new vertex (valence)
    uint8_t* aflag;               // new halfedges to pair again: 1 + rank in the origin's row, 0 if carried
    uint32_t* cnt;                // ... per origin
    uint32_t* bcount;             // ... per 256 old faces (a block's turn of the compaction loop)
    Sizes* S;
    unsigned long long* vst;      // per-volume counts (last compaction of a batch)
    const int64_t* qb;            // first new face of each volume
    int nvol;
};

// Mark a new vertex (bits of cflag, one byte per vertex, updated as words).
__device__ __forceinline__ void mark_vertex(uint8_t* cflag, int32_t y, uint32_t bits) {
    atomicOr(reinterpret_cast<uint32_t*>(cflag) + (y >> 2), bits << (8 * (y & 3)));
}


// A changed vertex: paired again, its out-degree and he_v from its row.
__device__ __forceinline__ void mark_changed(uint8_t* cflag, uint32_t* deg_n, int32_t* hev_n, int32_t y) {
    mark_vertex(cflag, y, 1u);
    deg_n[y] = 0;
    hev_n[y] = -1;
}

// 3X3Y: the merging faces (mlist, S->nq - S->qn of them).  The moved and
// the dropped corner are changed.  The other two (valence >= 5, never moved)
// each lose the halfedge they start in the face; if that was their largest,
// their he_v is taken again over what survives (cflag bit 4).
__global__ void __launch_bounds__(kBlock) k_changed3x3y(const int32_t* __restrict__ quads,
                                                         const int32_t* __restrict__ mlist, const Sizes* S,
                                                         const uint8_t* __restrict__ em,
                                                         const int32_t* __restrict__ vmap,
                                                         const int32_t* __restrict__ hev, uint8_t* __restrict__ cflag,
                                                         uint32_t* __restrict__ deg_n, int32_t* __restrict__ hev_n) {
    const int64_t n = (int64_t)(S->nq - S->qn), nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    TMC_GRID_STRIDE(i, n) {
        const int32_t f = mlist[i];
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        const bool first = em[f] == 1;
        const int32_t x[4] = {first ? q.x : q.y, first ? q.z : q.w, first ? q.y : q.x, first ? q.w : q.z};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!vid_ok(x[k], nv)) continue;
            const int32_t y = vmap[x[k]];
            if (!vid_ok(y, nn)) continue;
            if (k < 2) {
                mark_changed(cflag, deg_n, hev_n, y);
            } else {
                atomicSub(&deg_n[y], 1u);
                if ((hev[x[k]] >> 2) == f) {
                    mark_vertex(cflag, y, 4u);
                    hev_n[y] = -1;
                }
            }
        }
    }
}

// 3333: the dropped faces (dlist, S->nq - S->qn of them); every kept corner
// is changed.
__global__ void __launch_bounds__(kBlock) k_changed3333(const int32_t* __restrict__ quads,
                                                         const int32_t* __restrict__ dlist, const Sizes* S,
                                                         const int32_t* __restrict__ vmap, uint8_t* __restrict__ cflag,
                                                         uint32_t* __restrict__ deg_n, int32_t* __restrict__ hev_n) {
    const int64_t n = (int64_t)(S->nq - S->qn), nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    TMC_GRID_STRIDE(i, n) {
        const int4 q = reinterpret_cast<const int4*>(quads)[dlist[i]];
        const int32_t x[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!vid_ok(x[k], nv)) continue;
            const int32_t y = vmap[x[k]];
            if (vid_ok(y, nn)) mark_changed(cflag, deg_n, hev_n, y);
        }
    }
}

// The carry of surviving face f (new index nf, new corners r): every
// halfedge either keeps its group (twin renumbered, flag, arrival) or is
// flagged for pairing, with its rank in its origin's row (a rank past 253,
// never met in practice, sends the next build to the generic rows).
// Returns the number flagged.
__device__ __forceinline__ uint32_t carry_face(const Carry& c, bool act, int64_t f, uint32_t nf, const int32_t r[4],
                                               int64_t nn, HeCount& cnt) {
    int32_t tw[4] = {-1, -1, -1, -1};
    uint32_t fl = 0, af = 0, naff = 0;
    if (act) {
        const int4 t4 = reinterpret_cast<const int4*>(c.twin)[f];
        const uint32_t f4 = reinterpret_cast<const uint32_t*>(c.flag)[f];
        const int32_t t[4] = {t4.x, t4.y, t4.z, t4.w};
        uint8_t cf[4];
#pragma unroll
        for (int k = 0; k < 4; k++) cf[k] = vid_ok(r[k], nn) ? c.cflag[r[k]] : 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int32_t a = r[k], b = r[(k + 1) & 3];
            const int32_t h = (int32_t)(4 * nf) + k;
            if (!vid_ok(a, nn) || !vid_ok(b, nn)) continue;   // invalid ids: the next build starts over
            const uint8_t ca = cf[k], cb = cf[(k + 1) & 3];
            if (ca & 4u) atomicMax(&c.hev_n[a], h);   // its largest halfedge was dropped
            if ((ca | cb) & 1u) {
                const uint32_t rk = atomicAdd(&c.cnt[a], 1u);
                if (rk >= 254u) atomicOr(&c.S->cbad, 0x80000000u);
                af |= (rk < 254u ? rk + 1u : 255u) << (8 * k);
                naff++;
                continue;
            }
            const int32_t nq = t[k] >= 0 ? c.qmap[t[k] >> 2] : -1;
            tw[k] = nq < 0 ? -1 : 4 * nq + (t[k] & 3);
            const uint8_t x = (uint8_t)(f4 >> (8 * k));
            fl |= (uint32_t)x << (8 * k);
            if (tw[k] < 0) atomicAdd(&c.arr[b], 1u);
            cnt.add(x);
            vol_count(c.vst, c.qb, c.nvol, h, x);
        }
        reinterpret_cast<int4*>(c.twin_n)[nf] = make_int4(tw[0], tw[1], tw[2], tw[3]);
        reinterpret_cast<uint32_t*>(c.flag_n)[nf] = fl;
        reinterpret_cast<uint32_t*>(c.aflag)[nf] = af;
    }
    return naff;
}

// The block's count of flagged halfedges for its turn j of the compaction
// loop (every thread of the block calls this).
__device__ __forceinline__ void block_count(uint32_t* __restrict__ bcount, int64_t j, uint32_t k) {
    __shared__ uint32_t sm[kBlock / kWave];
    const uint32_t w = wave_sum(k);
    if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kBlock / kWave; i++) t += sm[i];
        bcount[j] = t;
    }
    __syncthreads();
}

// removeQuadrilaterals3X3Y (cpp_dmc/DualMarchingCubes.cpp:1114-1137), with
// the carry of the halfedges.  The vertex ids come from vmap (k_vmap3x3y):
// vm[vm[x]]-style lookups of a removed vertex x, exactly as the reference
// composes them.  Blocks stride over the faces in whole turns of 256, each
// turn writing its count of flagged halfedges.
__global__ void __launch_bounds__(kBlock) k_compact_quads3x3y(const int32_t* __restrict__ quads, const Sizes* S,
                                                               const int32_t* __restrict__ vmap,
                                                               int32_t* __restrict__ quads2, Carry c) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    HeCount cnt;
    for (int64_t f0 = (int64_t)blockIdx.x * kBlock; f0 < nq; f0 += (int64_t)gridDim.x * kBlock) {
        const int64_t f = f0 + threadIdx.x;
        const int32_t m = f < nq ? c.qmap[f] : -1;
        const bool act = m >= 0;
        int32_t r[4] = {-1, -1, -1, -1};
        const uint32_t nf = act ? (uint32_t)m : 0u;
        if (act) {
            const int4 q = reinterpret_cast<const int4*>(quads)[f];
            const int32_t x[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int k = 0; k < 4; k++) r[k] = vid_ok(x[k], nv) ? vmap[x[k]] : -1;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (r[k] < 0) atomicAdd(&c.S->cbad, 1u);
            reinterpret_cast<int4*>(quads2)[nf] = make_int4(r[0], r[1], r[2], r[3]);
        }
        block_count(c.bcount, f0 / kBlock, carry_face(c, act, f, nf, r, nn, cnt));
    }
    cnt.flush(c.S->hep);
}

// removeQuadrilaterals3333 (cpp_dmc/DualMarchingCubes.cpp:1320-1360), with
// the carry of the halfedges (a pattern face's halfedges all end at
// changed vertices, so none of them is carried).
__global__ void __launch_bounds__(kBlock) k_compact_quads3333(const int32_t* __restrict__ quads, const Sizes* S,
                                                               const uint8_t* __restrict__ p3333,
                                                               const int32_t* __restrict__ vtwin,
                                                               const int32_t* __restrict__ vmap,
                                                               int32_t* __restrict__ quads2, Carry c) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    HeCount cnt;
    for (int64_t f0 = (int64_t)blockIdx.x * kBlock; f0 < nq; f0 += (int64_t)gridDim.x * kBlock) {
        const int64_t f = f0 + threadIdx.x;
        const int32_t m = f < nq ? c.qmap[f] : -1;
        const bool act = m >= 0;
        int32_t r[4] = {-1, -1, -1, -1};
        const uint32_t nf = act ? (uint32_t)m : 0u;
        if (act) {
            const int4 q = reinterpret_cast<const int4*>(quads)[f];
            int32_t x[4] = {q.x, q.y, q.z, q.w};
            if (p3333[f]) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (vid_ok(x[k], nv)) x[k] = vtwin[x[k]];
            }
            // vmap[x] = new id of a kept x, -1 for a removed one (k_vmap3333)
#pragma unroll
            for (int k = 0; k < 4; k++) r[k] = vid_ok(x[k], nv) ? vmap[x[k]] : -1;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (r[k] < 0) atomicAdd(&c.S->cbad, 1u);
            reinterpret_cast<int4*>(quads2)[nf] = make_int4(r[0], r[1], r[2], r[3]);
        }
        block_count(c.bcount, f0 / kBlock, carry_face(c, act, f, nf, r, nn, cnt));
    }
    cnt.flush(c.S->hep);
}

// The flagged halfedges as a dense list, and their rows by origin.  Walks
// the old faces in the compaction's turns of 256: boff = exclusive scan of
// the turns' counts, ranks inside a turn from a block scan.  aoff =
// exclusive scan of the rows' counts; a halfedge's place in its row is the
// rank its flag holds.
__global__ void __launch_bounds__(kBlock) k_aff_scatter(const Sizes* S, const int32_t* __restrict__ qmap,
                                                         const uint8_t* __restrict__ aflag,
                                                         const uint32_t* __restrict__ bcount,
                                                         const uint32_t* __restrict__ boff,
                                                         const int32_t* __restrict__ quads,
                                                         const uint32_t* __restrict__ aoff,
                                                         int2* __restrict__ aadj, int32_t* __restrict__ alist) {
    __shared__ uint32_t sm[kBlock / kWave + 1];
    const int64_t nq = (int64_t)S->nq;
    for (int64_t f0 = (int64_t)blockIdx.x * kBlock; f0 < nq; f0 += (int64_t)gridDim.x * kBlock) {
        if (bcount[f0 / kBlock] == 0u) continue;   // a turn with nothing flagged (block-uniform)
        const int64_t f = f0 + threadIdx.x;
        const int32_t m = f < nq ? qmap[f] : -1;
        const uint32_t af = m >= 0 ? reinterpret_cast<const uint32_t*>(aflag)[m] : 0u;
        uint32_t k4 = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) k4 += ((af >> (8 * k)) & 0xFFu) != 0u;
        uint32_t tot;
        uint32_t p = boff[f0 / kBlock] + block_exclusive_scan<uint32_t, kBlock>(k4, sm, tot);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t code = (af >> (8 * k)) & 0xFFu;
            if (!code) continue;
            const int32_t h = 4 * m + k;
            alist[p++] = h;
            if (code == 255u) continue;   // rank unknown: this build is redone from generic rows
            const int32_t a = quads[h], b = he_dest(quads, h);
            aadj[aoff[a] + code - 1u] = make_int2(h, b);
        }
    }
}

// Pair the flagged halfedges (cpp_dmc/DualMarchingCubes.cpp:574-669) through
// their rows; a changed origin also takes its out-degree and he_v from its
// row, which holds all its out-halfedges.
__global__ void __launch_bounds__(kBlock) k_aff_twins(const int32_t* __restrict__ alist, const Sizes* S,
                                                       const int32_t* __restrict__ quads,
                                                       const uint32_t* __restrict__ aoff,
                                                       const int2* __restrict__ aadj, Carry c, BndRec* __restrict__ recs,
                                                       uint32_t reccap) {
    // rows incomplete (a rank past 253): the next build starts over
    if (S->cbad & 0x80000000u) return;
    const int64_t n = (int64_t)S->naff, nn = (int64_t)S->nn;
    HeCount cnt;
    TMC_GRID_STRIDE(i, n) {
        const int32_t h = alist[i];
        const int32_t a = quads[h], b = he_dest(quads, h);
        int32_t g[5];
        const int k = he_group(quads, aoff, aadj, h, nn, g);
        int32_t tw;
        uint8_t fl;
        twin_rule(quads, k, g, h, tw, fl);
        c.twin_n[h] = tw;
        c.flag_n[h] = fl;
        if (tw == -1) atomicAdd(&c.arr[b], 1u);
        cnt.add(fl);
        vol_count(c.vst, c.qb, c.nvol, h, fl);
        if (k == 3 && g[0] == h) put_record(recs, &c.S->nrec, reccap, &c.S->ovf, three_record(quads, g, h));
        if (c.cflag[a] & 1u) {
            const uint32_t s = aoff[a], e = aoff[a + 1];
            int32_t hmax = -1;
            for_row(aadj, s, e, [&](int2 x) { hmax = max(hmax, x.x); });
            c.deg_n[a] = e - s;
            if (h == hmax) c.hev_n[a] = h;
        }
    }
    cnt.flush(c.S->hep);
}

// 3-face records of groups the compaction left alone, renumbered.
__global__ void __launch_bounds__(kBlock) k_carry_records(const BndRec* __restrict__ recs, const Sizes* S,
                                                           const int32_t* __restrict__ vmap,
                                                           const uint8_t* __restrict__ cflag,
                                                           const int32_t* __restrict__ qmap,
                                                           BndRec* __restrict__ recs_n, uint32_t* nrec,
                                                           uint32_t reccap, uint32_t* ovf) {
    const int64_t n = S->nrec_prev < reccap ? S->nrec_prev : reccap;
    const int64_t nv = (int64_t)S->nv, nn = (int64_t)S->nn;
    TMC_GRID_STRIDE(i, n) {
        const BndRec r = recs[i];
        const int32_t v0 = (int32_t)(r.key >> 32), v1 = (int32_t)(r.key & 0xFFFFFFFFull);
        if (!vid_ok(v0, nv) || !vid_ok(v1, nv) || !vid_ok(r.a, nv) || !vid_ok(r.b, nv)) continue;
        const int32_t y0 = vmap[v0], y1 = vmap[v1];
        if (!vid_ok(y0, nn) || !vid_ok(y1, nn) || ((cflag[y0] | cflag[y1]) & 1u)) continue;
        BndRec o;
        o.key = he_key(y0, y1);
        o.a = vmap[r.a];
        o.b = vmap[r.b];
        const int32_t m = qmap[r.lone >> 2];
        if (m < 0) continue;
        o.lone = 4 * m + (r.lone & 3);
        o.pad = 0;
        put_record(recs_n, nrec, reccap, ovf, o);
    }
}

// Before a halfedge build: which way it pairs.  ready: 0 generic rows
// (launched only in a pass run again for them), 1 the cell stage's rows, 2
// the build a compaction carried over.  The cell stage's rows are whole
// unless a quadrilateral has an unresolved neighbour cell; a carried build
// stands only if the build before it and the compaction met no invalid
// vertex reference.  Otherwise the build needs generic rows, which this
// pass did not launch: it flags the pass to be run again with them.
__global__ void k_rows_gate(Sizes* S, int ready, int commit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (commit) {   // the compaction before this build: its counts become the live mesh
            S->nv = S->nn;
            S->nq = S->qn;
        }
        const bool incr = ready == 2 && S->incr_ok != 0u && S->cbad == 0u;
        const bool need = (ready == 1 && S->unresolved != 0u) || (ready == 2 && !incr);
        if (need) S->need_generic = 1u;
        S->incr = incr ? 1u : 0u;
        S->pair_rows = ready == 0 || (ready == 1 && !need) ? 1u : 0u;
        S->generic = ready == 0 ? 1u : 0u;
        if (!incr) {
            S->recmax = S->nrec > S->recmax ? S->nrec : S->recmax;
            he_clear(S);
            S->nrec = 0;
        }
    }
}



// ---------------------------------------------------------------------------
// Diagnostics of a finished mesh
// ---------------------------------------------------------------------------
// Unique undirected edges (dmc/HalfedgeMesh.cu:430-450): one per group of
// halfedges sharing a vertex pair -- the reference's hash key (min, max)
// (dmc/EdgeHashTable.h:101-127) -- represented by the group's smallest
// halfedge.  A halfedge with a twin and no flag is in a group of two; the
// others (boundary, three or four faces, more) take the group from the rows.
__global__ void __launch_bounds__(kBlock) k_edge_reps(const int32_t* __restrict__ quads, int64_t nhe, int64_t nv,
                                                       const uint32_t* __restrict__ off,
                                                       const int2* __restrict__ adj,
                                                       const int32_t* __restrict__ twin,
                                                       const uint8_t* __restrict__ flag, uint32_t* __restrict__ rep) {
    TMC_GRID_STRIDE(h64, nhe) {
        const int32_t h = (int32_t)h64;
        const int32_t a = quads[h], b = he_dest(quads, h);
        bool r = false;
        if (vid_ok(a, nv) && vid_ok(b, nv)) {
            const int32_t tw = twin[h];
            if (flag[h] == 0 && tw >= 0) {
                r = h < tw;
            } else {
                int32_t g[5];
                const int c = he_group(quads, off, adj, h, nv, g);
                r = c > 0 && g[0] == h;
            }
        }
        rep[h] = r;
    }
}

// Edge list in ascending order of the representatives (newidx = exclusive
// scan of rep): (origin, destination) of each.
__global__ void __launch_bounds__(kBlock) k_edge_emit(const int32_t* __restrict__ quads, int64_t nhe,
                                                       const uint32_t* __restrict__ rep,
                                                       const uint32_t* __restrict__ newidx,
                                                       int2* __restrict__ edges) {
    TMC_GRID_STRIDE(h, nhe) {
        if (!rep[h]) continue;
        edges[newidx[h]] = make_int2(quads[h], he_dest(quads, (int32_t)h));
    }
}

// Valence as dmc/VertexValence.cu:1-125 counts it: every distinct edge key
// (min, max) adds one to both end vertices, so a vertex's valence is its
// number of distinct neighbours, a loop counting twice.  One thread per
// vertex walks the corners in its halfedge row: with tris == nullptr the
// neighbours are the quad's vertices before and after the corner, else the
// other vertices of the quad's two triangles (tris[6q..6q+5]) that contain
// it.  Rows with more than kValRow corners take a slow loop.
constexpr int kValRow = 8;
__device__ __forceinline__ void corner_neighbours(const int32_t* __restrict__ quads,
                                                  const int32_t* __restrict__ tris, int32_t h, int32_t v,
                                                  int32_t nb[4]) {
    const int32_t q = h >> 2, k = h & 3;
    if (!tris) {
        nb[0] = quads[4 * q + ((k + 1) & 3)];
        nb[1] = quads[4 * q + ((k + 3) & 3)];
        nb[2] = nb[3] = -1;
        return;
    }
    const int32_t* t = tris + 6 * (int64_t)q;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int32_t a = t[3 * s], b = t[3 * s + 1], c = t[3 * s + 2];
        const bool in = a == v || b == v || c == v;
        // the two other corners of a triangle containing v (v itself where
        // the triangle repeats it)
        nb[2 * s] = !in ? -1 : (a == v ? b : a);
        nb[2 * s + 1] = !in ? -1 : (a == v || b == v ? c : b);
    }
}

__global__ void __launch_bounds__(kBlock) k_vertex_valence(const int32_t* __restrict__ quads,
                                                            const int32_t* __restrict__ tris, int64_t nv,
                                                            const uint32_t* __restrict__ off,
                                                            const int2* __restrict__ adj,
                                                            uint32_t* __restrict__ vval) {
    TMC_GRID_STRIDE(v64, nv) {
        const int32_t v = (int32_t)v64;
        const uint32_t s = off[v], e = off[v + 1], n = e - s;
        uint32_t val = 0;
        if (n <= (uint32_t)kValRow) {
            int32_t nb[4 * kValRow];
#pragma unroll
            for (int i = 0; i < kValRow; i++) {
                int32_t c[4] = {-1, -1, -1, -1};
                if ((uint32_t)i < n) corner_neighbours(quads, tris, adj[s + i].x, v, c);
#pragma unroll
                for (int x = 0; x < 4; x++) nb[4 * i + x] = c[x];
            }
            // count each neighbour at its first occurrence
#pragma unroll
            for (int i = 0; i < 4 * kValRow; i++) {
                const int32_t u = nb[i];
                bool first = u >= 0;
#pragma unroll
                for (int j = 0; j < i; j++) first = first && nb[j] != u;
                if (first) val += u == v ? 2u : 1u;
            }
        } else {
            for (uint32_t i = s; i < e; i++) {
                int32_t c[4];
                corner_neighbours(quads, tris, adj[i].x, v, c);
                for (int x = 0; x < 4; x++) {
                    const int32_t u = c[x];
                    if (u < 0) continue;
                    bool first = true;
                    // earlier corners of the row, then earlier slots of this one
                    for (uint32_t j = s; j < i && first; j++) {
                        int32_t d[4];
                        corner_neighbours(quads, tris, adj[j].x, v, d);
                        first = d[0] != u && d[1] != u && d[2] != u && d[3] != u;
                    }
                    for (int y = 0; y < x && first; y++) first = c[y] != u;
                    if (first) val += u == v ? 2u : 1u;
                }
            }
        }
        vval[v] = val;
    }
}

// Histogram of the valence counts: bins 0 .. kValenceBins-1 and, in the
// last entry, the vertices beyond them.
constexpr int kValenceBins = 100;
__global__ void __launch_bounds__(kBlock) k_valence_hist(const uint32_t* __restrict__ vval, int64_t nv,
                                                          uint32_t* __restrict__ hist) {
    __shared__ uint32_t sm[kValenceBins + 1];
    for (int b = threadIdx.x; b <= kValenceBins; b += kBlock) sm[b] = 0;
    __syncthreads();
    TMC_GRID_STRIDE(v, nv) atomicAdd(&sm[min(vval[v], (uint32_t)kValenceBins)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b <= kValenceBins; b += kBlock)
        if (sm[b]) atomicAdd(&hist[b], sm[b]);
}

// Element quality (dmc/ElementQuality.h:65-91, dmc/EstimateElementQuality.cu)
// in FP32 on the positions rounded to float, with the reference's operand
// order (its cross product's x component has the opposite sign, which the
// norm does not see).
struct QVec {
    float x, y, z;
};
__device__ __forceinline__ QVec q_vertex(const double* __restrict__ pos, int32_t v) {
    const double* p = pos + 3 * (int64_t)v;
    return QVec{(float)p[0], (float)p[1], (float)p[2]};
}
__device__ __forceinline__ QVec q_diff(const QVec& a, const QVec& b) { return QVec{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ QVec q_cross(const QVec& a, const QVec& b) {
    return QVec{a.z * b.y - a.y * b.z, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float q_norm2(const QVec& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ __forceinline__ float q_norm(const QVec& v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }

__global__ void __launch_bounds__(kBlock) k_quality_quads(const int32_t* __restrict__ quads, int64_t nq,
                                                           const double* __restrict__ pos,
                                                           float* __restrict__ quality) {
    TMC_GRID_STRIDE(f, nq) {
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        const QVec v[4] = {q_vertex(pos, q.x), q_vertex(pos, q.y), q_vertex(pos, q.z), q_vertex(pos, q.w)};
        float m = 0.f;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            // corner c: edges to the next and to the previous vertex
            const QVec e1 = q_diff(v[(c + 1) & 3], v[c]), e2 = q_diff(v[(c + 3) & 3], v[c]);
            const float a = q_norm(q_cross(e1, e2));
            const float r = 2 * a / (q_norm2(e1) + q_norm2(e2));
            m = c == 0 ? r : (m < r ? m : r);
        }
        quality[f] = m;
    }
}

__global__ void __launch_bounds__(kBlock) k_quality_tris(const int32_t* __restrict__ tris, int64_t nt,
                                                          const double* __restrict__ pos,
                                                          float* __restrict__ quality) {
    const float factor = 2 * sqrtf(3.f);
    TMC_GRID_STRIDE(f, nt) {
        const int32_t* t = tris + 3 * f;
        const QVec v0 = q_vertex(pos, t[0]), v1 = q_vertex(pos, t[1]), v2 = q_vertex(pos, t[2]);
        const float a = q_norm(q_cross(q_diff(v1, v0), q_diff(v2, v0)));
        const float l0 = q_norm2(q_diff(v1, v0));
        const float l1 = q_norm2(q_diff(v2, v1));
        const float l2 = q_norm2(q_diff(v0, v2));
        float q = 0.f;
        if (l0 != 0 && l1 != 0 && l2 != 0) q = factor * a / (l0 + l1 + l2);
        quality[f] = q;
    }
}

// collectTriangles (cpp_dmc/DualMarchingCubes.cpp:1387-1460) with minAngle
// (DualMarchingCubes.h:746-756): split along the diagonal whose two
// triangles have the larger minimum angle.
// Side lengths as minAngle forms them: the double distance rounded to
// float.  (p - q)^2 == (q - p)^2 exactly, so each of a quad's six vertex
// pairs is computed once and shared by the triangles that use it.
__device__ __forceinline__ float tri_side(const double* __restrict__ pos, int32_t p, int32_t q) {
    const double* a = pos + 3 * (int64_t)p;
    const double* b = pos + 3 * (int64_t)q;
    const double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    return (float)sqrt(d);
}

// minimum angle of the triangle with sides a = |i0 i1|, b = |i1 i2|, c = |i2 i0|
__device__ __forceinline__ double tri_min_angle(float a, float b, float c) {
    const float A = acosf((b * b + c * c - a * a) / (2 * b * c));
    const float B = acosf((a * a + c * c - b * b) / (2 * a * c));
    const float C = acosf((b * b + a * a - c * c) / (2 * b * a));
    const float m = (B < A) ? B : A;
    return (C < m) ? C : m;
}

__global__ void __launch_bounds__(kBlock) k_triangles(const int32_t* __restrict__ quads, const Sizes* S,
                                                       const double* __restrict__ pos, int32_t* __restrict__ tris) {
    const int64_t nq = (int64_t)S->nq, nv = (int64_t)S->nv;
    TMC_GRID_STRIDE(f, nq) {
        const int4 q = reinterpret_cast<const int4*>(quads)[f];
        int32_t* t = tris + 6 * f;
        if (!(vid_ok(q.x, nv) && vid_ok(q.y, nv) && vid_ok(q.z, nv) && vid_ok(q.w, nv))) {
            t[0] = q.x; t[1] = q.y; t[2] = q.z; t[3] = q.x; t[4] = q.z; t[5] = q.w;
            continue;
        }
        const float xy = tri_side(pos, q.x, q.y), yz = tri_side(pos, q.y, q.z), zw = tri_side(pos, q.z, q.w),
                    wx = tri_side(pos, q.w, q.x), xz = tri_side(pos, q.x, q.z), yw = tri_side(pos, q.y, q.w);
        // triangles (x,y,z), (x,z,w), then (y,w,x), (y,z,w)
        double a1 = tri_min_angle(xy, yz, xz);
        double a2 = tri_min_angle(xz, zw, wx);
        const double b1 = (a2 < a1) ? a2 : a1, b2 = (a1 < a2) ? a2 : a1;
        a1 = tri_min_angle(yw, wx, xy);
        a2 = tri_min_angle(yz, zw, yw);
        const double c1 = (a2 < a1) ? a2 : a1, c2 = (a1 < a2) ? a2 : a1;
        if (b1 < c1 || (b1 == c1 && b2 <= c2)) {
            t[0] = q.y; t[1] = q.w; t[2] = q.x; t[3] = q.y; t[4] = q.z; t[5] = q.w;
        } else {
            t[0] = q.x; t[1] = q.y; t[2] = q.z; t[3] = q.x; t[4] = q.z; t[5] = q.w;
        }
    }
}

}  // namespace tmc
