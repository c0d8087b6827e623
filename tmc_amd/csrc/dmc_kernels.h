// dmc_kernels.h -- generic device building blocks: wave/block scans and the
// device-wide exclusive scan used to turn per-item counts into offsets.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmc {

constexpr int kBlock = 256;
constexpr int kWave = 64;
// The 8 XCDs each have their own L2 and receive workgroups round-robin
// (block b runs on XCD b % 8).  xcd_block() renumbers blocks so that each
// XCD works on one contiguous range of logical blocks, keeping the halo
// shared by neighbouring tiles in one L2.  It is a bijection when the grid
// is a multiple of 8 blocks (the launcher rounds up), else the identity.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t nb = gridDim.x, b = blockIdx.x;
    if (nb % kXcds) return b;
    return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// Largest v < n with base[v] <= x (base ascending, base[0] <= x): the
// volume of item x when base holds each volume's first item.
__device__ __forceinline__ int vol_of(const int64_t* __restrict__ base, int n, int64_t x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// An item count that lives in device memory (a scan total, or the size a
// compaction left), clamped to the capacity of the arrays it indexes, so
// that a launch sized on the host by that capacity never reads or writes
// past them.  p == nullptr: the host-known count `cap`.
struct DevN {
    const uint64_t* p;
    int64_t cap;
    __device__ __forceinline__ int64_t get() const {
        if (!p) return cap;
        const uint64_t x = *p;
        return x < (uint64_t)cap ? (int64_t)x : cap;
    }
};

// Statistics of one halfedge build, counted per halfedge so that a build
// that carries most groups over from the last one can count them too:
// non-manifold edges = nm4 / 4 + nm3 / 3, 3-face edges = nm3 / 3.
// Halfedge statistics of a build.  On the device only bad and pad are kept
// here; the counts of flagged halfedges live in Sizes::hep (spread over
// kHeParts slots) and the host writes their sums into nm4, errors, nm3.
struct HeStats {
    unsigned long long nm4;      // halfedges flagged non-manifold (groups of 4)
    unsigned long long errors;   // groups of more than 4 halfedges
    unsigned long long nm3;      // halfedges flagged boundary multiple edge (groups of 3)
    unsigned int bad;            // invalid vertex references
    unsigned int pad;
};

// The flagged-halfedge counts each block of a build adds are spread over
// this many slots (block index modulo the count) and summed where they are
// read, so that a large grid over a noisy mesh does not queue its adds on
// three addresses.
constexpr int kHeParts = 64;

// A single volume's contour count is spread over this many counters (one
// per wave slot, summed by the host), so that the waves' adds at the end of
// their rows do not queue on a few addresses.
constexpr int kContourParts = 256;

// The sizes of one extraction pass, kept on the device: every launch reads
// its item count from here, so a pass needs no count read back between its
// stages.  The host sizes launches and buffers by capacities and reads this
// block once at the end.
struct Sizes {
    uint64_t A, V, Q;            // active cells, vertices, quadrilaterals of the cell stage (exact)
    uint64_t acnt;               // active cells the cell kernels process: A, or 0 once a capacity is exceeded
    uint64_t nv, nq;             // the live mesh
    uint64_t nn, qn;             // vertices / quadrilaterals a compaction keeps
    uint64_t rows;               // first CSR's row total (4 Q)
    uint64_t tot;                // a scan total no one else reads
    uint64_t hq;                 // halfedges of the mesh a compaction leaves (4 qn)
    uint64_t naff;               // halfedges a compaction left to pair again
    unsigned long long qhead;    // work queue of k_vertices: the next chunk of active cells
    HeStats he;                  // the last halfedge build
    uint32_t ovf;                // capacities exceeded: 1 active cells, 2 vertices, 4 quads, 8 3-face records
    uint32_t unresolved;         // quadrilaterals with an unresolved neighbour cell
    uint32_t zero;               // vertices of valence 0 (all valence passes)
    uint32_t generic;            // this halfedge build counts its rows from the quadrilaterals
    uint32_t need_generic;       // a build needed generic rows the pass did not launch: run it again with them
    uint32_t pair_rows;          // this build pairs halfedges through full rows (cell rows or generic)
    uint32_t incr;               // this build is the one the last compaction carried over
    uint32_t incr_ok;            // the build before the compaction saw no invalid vertex reference
    uint32_t cbad;               // invalid references the last quadrilateral compaction met
    uint32_t nrec;               // 3-face edge records of the current build
    uint32_t nrec_prev;          // ... of the build before (carried over by a compaction)
    uint32_t recmax;             // most 3-face records any build of the pass wrote (capacity planning)
    uint32_t colors[48];         // colour classes: counts [0, 24), starts [24, 48)
    unsigned long long cpart[kContourParts];   // a single volume's contour count, in parts
    unsigned long long hep[kHeParts][3];      // the build's flagged halfedges per slot: groups of 4, of 3, error groups
};

// The build's count of error groups (groups of more than 4 halfedges).
__device__ __forceinline__ unsigned long long he_errors(const Sizes* S) {
    unsigned long long e = 0;
    for (int p = 0; p < kHeParts; p++) e += S->hep[p][2];
    return e;
}

// Fresh counts for a new build (one thread).
__device__ __forceinline__ void he_clear(Sizes* S) {
    for (int p = 0; p < kHeParts; p++) S->hep[p][0] = S->hep[p][1] = S->hep[p][2] = 0;
    S->he.bad = 0;
    S->he.pad = 0;
}

// A launch whose work is conditional on device data: it runs only when
// *gate != 0 (gate == nullptr: always).
__device__ __forceinline__ bool gated_off(const uint32_t* gate) { return gate && *gate == 0u; }

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Block-wide exclusive scan; smem needs BS/64 + 1 entries.  Returns the
// exclusive prefix of x and the block total.
template <class T, int BS>
__device__ __forceinline__ T block_exclusive_scan(T x, T* smem, T& total) {
    constexpr int NW = BS / kWave;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const T inc = wave_inclusive_scan(x);
    if (lane == kWave - 1) smem[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const T s = lane < NW ? smem[lane] : T(0);
        const T si = wave_inclusive_scan(s);
        if (lane < NW) smem[lane] = si - s;
        if (lane == NW - 1) smem[NW] = si;
    }
    __syncthreads();
    const T res = inc - x + smem[wid];
    total = smem[NW];
    __syncthreads();
    return res;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of uint32 counts, or of keep flags read from
// byte flags, over tiles of kScanTile items.  Totals are accumulated in 64
// bits so that overflow of the 32-bit offsets can be detected by the host.
// ---------------------------------------------------------------------------
constexpr int kScanIPT = 16;
constexpr int kScanTile = kBlock * kScanIPT;

// Scan inputs: uint32 counts, or keep flags (1 where a byte flag is 0).
// Thread t of a tile holds the 16 consecutive items from base = tile * kScanTile + 16 t.
struct ScanCounts {
    static constexpr bool kMap = false;
    const uint32_t* p;
    __device__ __forceinline__ uint32_t at(int64_t i) const { return p[i]; }
    __device__ __forceinline__ void vec16(int64_t base, uint32_t v[kScanIPT]) const {
#pragma unroll
        for (int q = 0; q < kScanIPT / 4; q++) {
            const uint4 t = reinterpret_cast<const uint4*>(p + base)[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    }
};
struct ScanSet {
    static constexpr bool kMap = false;
    const uint8_t* p;   // set flags
    __device__ __forceinline__ uint32_t at(int64_t i) const { return p[i] ? 1u : 0u; }
    __device__ __forceinline__ void vec16(int64_t base, uint32_t v[kScanIPT]) const {
        const uint4 t = *reinterpret_cast<const uint4*>(p + base);
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) v[q] = ((w[q >> 2] >> (8 * (q & 3))) & 0xFFu) ? 1u : 0u;
    }
};
struct ScanKeep {
    static constexpr bool kMap = false;
    const uint8_t* p;   // removed flags
    __device__ __forceinline__ uint32_t at(int64_t i) const { return p[i] ? 0u : 1u; }
    __device__ __forceinline__ void vec16(int64_t base, uint32_t v[kScanIPT]) const {
        const uint4 t = *reinterpret_cast<const uint4*>(p + base);
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) v[q] = ((w[q >> 2] >> (8 * (q & 3))) & 0xFFu) ? 0u : 1u;
    }
};

// Keep flags whose scan is written as a compaction map: a kept item gets
// its new index, a removed one ~(its rank among the removed), which is
// negative as an int32 (so the map reads -1-or-below for "dropped", and the
// rank is ~map).  The kept count before item i is map[i] or i - ~map[i].
struct ScanKeepMap : ScanKeep {
    static constexpr bool kMap = true;
};

template <class In>
__device__ __forceinline__ void scan_items(const In& in, int64_t n, int64_t base, uint32_t v[kScanIPT]) {
    if (base + kScanIPT <= n) {
        in.vec16(base, v);   // all 16 items in range: vector loads
    } else {
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) v[q] = base + q < n ? in.at(base + q) : 0u;
    }
}

// Device-wide exclusive scan over n = dn.get() items in three launches over
// ceil(dn.cap / kScanTile) tiles (tile sums, their scan, the tiles again);
// skipped entirely when gated off.  For inputs of many tiles.
template <class In>
__device__ __forceinline__ void scan_reduce_tile(const In& in, DevN dn, const uint32_t* __restrict__ gate,
                                                 uint64_t* __restrict__ partial, int64_t tile) {
    __shared__ uint64_t sm[kBlock / kWave + 1];
    if (gated_off(gate)) return;
    const int64_t n = dn.get();
    uint32_t v[kScanIPT];
    scan_items(in, n, tile * kScanTile + (int64_t)threadIdx.x * kScanIPT, v);
    uint64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) s += v[q];
    s = wave_sum(s);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kBlock / kWave; w++) t += sm[w];
        partial[tile] = t;
    }
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(In in, DevN dn, const uint32_t* __restrict__ gate,
                                                       uint64_t* __restrict__ partial) {
    scan_reduce_tile(in, dn, gate, partial, (int64_t)blockIdx.x);
}

// Single block: exclusive scan of the tile sums in place, total to *total
// (and, for a CSR, its low 32 bits to row_end[n], the end of the last row).
__device__ __forceinline__ void scan_partials_run(uint64_t* __restrict__ partial, int64_t ntiles,
                                                  const uint32_t* __restrict__ gate, uint64_t* __restrict__ total,
                                                  uint32_t* __restrict__ row_end, DevN dn) {
    __shared__ uint64_t sm[1024 / kWave + 1];
    if (gated_off(gate)) return;
    uint64_t carry = 0;
    for (int64_t b = 0; b < ntiles; b += 1024) {
        const int64_t idx = b + threadIdx.x;
        const uint64_t x = idx < ntiles ? partial[idx] : 0;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan<uint64_t, 1024>(x, sm, tot);
        if (idx < ntiles) partial[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (row_end) row_end[dn.get()] = (uint32_t)carry;
    }
}

__global__ void __launch_bounds__(1024) k_scan_partials(uint64_t* __restrict__ partial, int64_t ntiles,
                                                         const uint32_t* __restrict__ gate, uint64_t* __restrict__ total,
                                                         uint32_t* __restrict__ row_end, DevN dn) {
    scan_partials_run(partial, ntiles, gate, total, row_end, dn);
}

template <class In>
__device__ __forceinline__ void scan_down_tile(const In& in, uint32_t* __restrict__ out, DevN dn,
                                               const uint32_t* __restrict__ gate,
                                               const uint64_t* __restrict__ partial, int64_t tile) {
    __shared__ uint32_t sm[kBlock / kWave + 1];
    if (gated_off(gate)) return;
    const int64_t n = dn.get();
    const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanIPT;
    const bool full = base + kScanIPT <= n;
    uint32_t v[kScanIPT];
    scan_items(in, n, base, v);
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) s += v[q];
    uint32_t tot;
    uint32_t run = block_exclusive_scan<uint32_t, kBlock>(s, sm, tot) + (uint32_t)partial[tile];
    uint32_t o[kScanIPT];
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) {
        o[q] = (In::kMap && v[q] == 0u) ? ~(uint32_t)(base + q - run) : run;
        run += v[q];
    }
    if (full) {
#pragma unroll
        for (int q = 0; q < kScanIPT / 4; q++)
            reinterpret_cast<uint4*>(out + base)[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < kScanIPT; q++)
            if (base + q < n) out[base + q] = o[q];
    }
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_scan_down(In in, uint32_t* __restrict__ out, DevN dn,
                                                       const uint32_t* __restrict__ gate,
                                                       const uint64_t* __restrict__ partial) {
    scan_down_tile(in, out, dn, gate, partial, (int64_t)blockIdx.x);
}

// The same in one launch of ceil(dn.cap / kScanTile) blocks (decoupled
// look-back), for inputs of up to kScanOnePassTiles tiles, so that a
// look-back never spans more than one wave's window; skipped entirely when
// gated off.  Each block takes the next tile from a ticket counter, so a tile
// only waits on tiles whose blocks already run.  It publishes its sum
// (kScanAgg); wave 0 looks back over the predecessors, adding sums until it
// meets a tile that holds its inclusive prefix (kScanInc); the tile then
// publishes its own.  Waits are bounded: a tile whose predecessors are not
// all published after `spin` polls (at most kScanSpin) sums their items
// straight from the input, which the scan does not write (the host checks
// that it is not the output), so its prefix is exact
// either way, and a late flag can neither hang the device nor corrupt the
// result.  The last block to finish its look-back clears the tickets and
// flags, so the state is zero again for the next scan.
// state st: [0] tickets, [1] look-backs done, [2..4) unused, [4, 4 + ntiles) tile flags
constexpr uint64_t kScanAgg = 1ull << 62, kScanInc = 2ull << 62, kScanVal = (1ull << 62) - 1;
constexpr int kScanSpin = 1 << 22;
constexpr int kScanStateHead = 4;
constexpr int64_t kScanOnePassTiles = 64;

__device__ __forceinline__ uint64_t scan_flag_load(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void scan_flag_store(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class In>
__device__ __forceinline__ void scan_onepass_run(const In& in, uint32_t* __restrict__ out, DevN dn,
                                                 const uint32_t* __restrict__ gate, uint64_t* __restrict__ total,
                                                 uint32_t* __restrict__ row_end, unsigned long long* __restrict__ st,
                                                 int64_t ntiles, int spin) {
    __shared__ uint64_t sm[kBlock / kWave + 1];
    __shared__ int64_t s_tile;
    __shared__ uint64_t s_excl;
    __shared__ int s_stalled, s_last;
    if (gated_off(gate)) return;
    const int64_t n = dn.get();
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(&st[0], 1ull);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanIPT;
    const bool full = base + kScanIPT <= n;
    uint32_t v[kScanIPT];
    scan_items(in, n, base, v);
    uint64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) s += v[q];
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan<uint64_t, kBlock>(s, sm, tot);
    unsigned long long* flags = st + kScanStateHead;
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        if (lane == 0 && tile > 0) scan_flag_store(&flags[tile], kScanAgg | (tot & kScanVal));
        uint64_t excl = 0;
        int64_t p = tile - 1;
        int spins = 0;
        bool stalled = false;
        while (p >= 0) {
            const int64_t q = p - lane;
            // before the first tile: an inclusive prefix of 0
            const uint64_t w = q >= 0 ? scan_flag_load(&flags[q]) : kScanInc;
            const uint64_t inc = __ballot((w >> 62) == 2u);
            const uint64_t notready = __ballot((w >> 62) == 0u);
            const int fi = inc ? __ffsll((unsigned long long)inc) - 1 : kWave;   // nearest inclusive lane
            const uint64_t need = fi >= kWave - 1 ? ~0ull : ((2ull << fi) - 1ull);
            if (notready & need) {
                if (++spins < spin) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                stalled = true;
                break;
            }
            excl += wave_sum(lane <= fi ? (w & kScanVal) : 0ull);
            if (fi < kWave) break;
            p -= kWave;
        }
        if (lane == 0) {
            s_excl = excl;
            s_stalled = stalled ? 1 : 0;
        }
    }
    __syncthreads();
    if (s_stalled) {
        // the predecessors' items straight from the input
        const int64_t end = tile * kScanTile < n ? tile * kScanTile : n;
        uint64_t acc = 0;
        for (int64_t i = threadIdx.x; i < end; i += kBlock) acc += in.at(i);
        uint64_t all;
        (void)block_exclusive_scan<uint64_t, kBlock>(acc, sm, all);
        if (threadIdx.x == 0) s_excl = all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t excl = s_excl;
        scan_flag_store(&flags[tile], kScanInc | ((excl + tot) & kScanVal));
        if (tile == ntiles - 1) {
            *total = excl + tot;
            if (row_end) row_end[n] = (uint32_t)(excl + tot);
        }
        // the flag store completes before the count (the clearing block
        // counts last, so no flag lands after it clears); a plain wait, not
        // a release, which would write back and invalidate the L2
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned long long d = atomicAdd(&st[1], 1ull);
        s_last = d == (unsigned long long)(ntiles - 1) ? 1 : 0;
    }
    __syncthreads();
    uint32_t run = (uint32_t)(s_excl + ex);
    uint32_t o[kScanIPT];
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) {
        o[q] = (In::kMap && v[q] == 0u) ? ~(uint32_t)(base + q - run) : run;
        run += v[q];
    }
    if (full) {
#pragma unroll
        for (int q = 0; q < kScanIPT / 4; q++)
            reinterpret_cast<uint4*>(out + base)[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < kScanIPT; q++)
            if (base + q < n) out[base + q] = o[q];
    }
    // every tile has finished its look-back: clear the state for the next scan
    if (s_last) {
        for (int64_t i = threadIdx.x; i < ntiles; i += kBlock) scan_flag_store(&flags[i], 0ull);
        if (threadIdx.x == 0) {
            scan_flag_store(&st[0], 0ull);
            scan_flag_store(&st[1], 0ull);
        }
    }
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_scan(In in, uint32_t* __restrict__ out, DevN dn,
                                                  const uint32_t* __restrict__ gate, uint64_t* __restrict__ total,
                                                  uint32_t* __restrict__ row_end, unsigned long long* __restrict__ st,
                                                  int64_t ntiles, int spin) {
    scan_onepass_run(in, out, dn, gate, total, row_end, st, ntiles, spin);
}

// Two independent scans in one launch (a pair of the kernels above): the
// first a.ntiles blocks run scan a, the rest scan b, each job with its own
// tickets and flags (one pass) or tile sums (three passes).
template <class In>
struct ScanJob {
    In in;
    uint32_t* out;
    DevN dn;
    const uint32_t* gate;
    uint64_t* total;
    uint32_t* row_end;
    unsigned long long* st;
    uint64_t* partial;
    int64_t ntiles;
};

template <class A, class B>
__global__ void __launch_bounds__(kBlock) k_scan_pair(ScanJob<A> a, ScanJob<B> b, int spin) {
    if ((int64_t)blockIdx.x < a.ntiles)
        scan_onepass_run(a.in, a.out, a.dn, a.gate, a.total, a.row_end, a.st, a.ntiles, spin);
    else
        scan_onepass_run(b.in, b.out, b.dn, b.gate, b.total, b.row_end, b.st, b.ntiles, spin);
}

template <class A, class B>
__global__ void __launch_bounds__(kBlock) k_scan_reduce_pair(ScanJob<A> a, ScanJob<B> b) {
    const int64_t t = blockIdx.x;
    if (t < a.ntiles) scan_reduce_tile(a.in, a.dn, a.gate, a.partial, t);
    else scan_reduce_tile(b.in, b.dn, b.gate, b.partial, t - a.ntiles);
}

template <class A, class B>
__global__ void __launch_bounds__(1024) k_scan_partials_pair(ScanJob<A> a, ScanJob<B> b) {
    if (blockIdx.x == 0) scan_partials_run(a.partial, a.ntiles, a.gate, a.total, a.row_end, a.dn);
    else scan_partials_run(b.partial, b.ntiles, b.gate, b.total, b.row_end, b.dn);
}

template <class A, class B>
__global__ void __launch_bounds__(kBlock) k_scan_down_pair(ScanJob<A> a, ScanJob<B> b) {
    const int64_t t = blockIdx.x;
    if (t < a.ntiles) scan_down_tile(a.in, a.out, a.dn, a.gate, a.partial, t);
    else scan_down_tile(b.in, b.out, b.dn, b.gate, b.partial, t - a.ntiles);
}

// The whole scan in one block, for short inputs (one launch, no state):
// tiles of 1024 x kScanIPT items in turn, carrying the running sum.
template <class In>
__global__ void __launch_bounds__(1024) k_scan_small(In in, uint32_t* __restrict__ out, DevN dn,
                                                      const uint32_t* __restrict__ gate, uint64_t* __restrict__ total,
                                                      uint32_t* __restrict__ row_end) {
    __shared__ uint32_t sm[1024 / kWave + 1];
    if (gated_off(gate)) return;
    const int64_t n = dn.get();
    uint32_t carry = 0;
    uint64_t carry64 = 0;
    for (int64_t t0 = 0; t0 < n; t0 += (int64_t)1024 * kScanIPT) {
        const int64_t base = t0 + (int64_t)threadIdx.x * kScanIPT;
        uint32_t v[kScanIPT];
        scan_items(in, n, base, v);
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) sum += v[q];
        uint32_t tot;
        uint32_t run = block_exclusive_scan<uint32_t, 1024>(sum, sm, tot) + carry;
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) {
            if (base + q < n) out[base + q] = (In::kMap && v[q] == 0u) ? ~(uint32_t)(base + q - run) : run;
            run += v[q];
        }
        carry += tot;
        carry64 += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry64;
        if (row_end) row_end[n] = carry;
    }
}

// Fill n = dn.get() 32-bit words with `value` (a kernel rather than a memset
// node, so that a captured pass replays it as a kernel).
__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* __restrict__ p, DevN dn, uint32_t value,
                                                    const uint32_t* __restrict__ gate) {
    if (gated_off(gate)) return;
    const int64_t n = dn.get();
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = value;
}

__global__ void __launch_bounds__(kBlock) k_fill_u8(uint8_t* __restrict__ p, DevN dn, uint8_t value) {
    const int64_t n = dn.get();
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = value;
}

// Several fills in one launch: fill y of the grid is job y (32-bit words or
// bytes), each gated on its own.
struct FillJob {
    void* p;
    DevN n;
    uint32_t value;
    const uint32_t* gate;
    int bytes;   // 4 or 1
};
constexpr int kMaxFillJobs = 6;
struct FillSet {
    FillJob j[kMaxFillJobs];
};

__global__ void __launch_bounds__(kBlock) k_fill_many(FillSet s) {
    const FillJob& J = s.j[blockIdx.y];
    if (gated_off(J.gate)) return;
    const int64_t n = J.n.get();
    if (J.bytes == 4) {
        uint32_t* p = static_cast<uint32_t*>(J.p);
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
            p[i] = J.value;
    } else {
        uint8_t* p = static_cast<uint8_t*>(J.p);
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
            p[i] = (uint8_t)J.value;
    }
}

// off[n] = the low 32 bits of a scan total (the end of the last CSR row).
__global__ void k_set_row_end(uint32_t* __restrict__ off, DevN dn, const uint64_t* __restrict__ total,
                              const uint32_t* __restrict__ gate) {
    if (gated_off(gate)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) off[dn.get()] = (uint32_t)*total;
}

}  // namespace tmc
