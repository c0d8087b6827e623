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
// byte flags (three passes over tiles of kScanTile items).  Totals are accumulated in 64 bits so that overflow of
// the 32-bit offsets can be detected by the host.
// ---------------------------------------------------------------------------
constexpr int kScanIPT = 16;
constexpr int kScanTile = kBlock * kScanIPT;

// Scan inputs: uint32 counts, or keep flags (1 where a byte flag is 0).
// Thread t of a tile holds the 16 consecutive items from base = tile * kScanTile + 16 t.
struct ScanCounts {
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
struct ScanKeep {
    const uint8_t* p;   // removed flags
    __device__ __forceinline__ uint32_t at(int64_t i) const { return p[i] ? 0u : 1u; }
    __device__ __forceinline__ void vec16(int64_t base, uint32_t v[kScanIPT]) const {
        const uint4 t = *reinterpret_cast<const uint4*>(p + base);
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int q = 0; q < kScanIPT; q++) v[q] = ((w[q >> 2] >> (8 * (q & 3))) & 0xFFu) ? 0u : 1u;
    }
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

template <class In>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(In in, int64_t n, uint64_t* __restrict__ partial) {
    __shared__ uint64_t sm[kBlock / kWave + 1];
    uint32_t v[kScanIPT];
    scan_items(in, n, (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanIPT, v);
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
        partial[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of the tile sums in place, total to *total.
__global__ void __launch_bounds__(1024) k_scan_partials(uint64_t* __restrict__ partial, int64_t ntiles,
                                                         uint64_t* __restrict__ total) {
    __shared__ uint64_t sm[1024 / kWave + 1];
    uint64_t carry = 0;
    for (int64_t b = 0; b < ntiles; b += 1024) {
        const int64_t idx = b + threadIdx.x;
        const uint64_t x = idx < ntiles ? partial[idx] : 0;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan<uint64_t, 1024>(x, sm, tot);
        if (idx < ntiles) partial[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_scan_down(In in, uint32_t* __restrict__ out, int64_t n,
                                                       const uint64_t* __restrict__ partial) {
    __shared__ uint32_t sm[kBlock / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanIPT;
    const bool full = base + kScanIPT <= n;
    uint32_t v[kScanIPT];
    scan_items(in, n, base, v);
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) s += v[q];
    uint32_t tot;
    uint32_t run = block_exclusive_scan<uint32_t, kBlock>(s, sm, tot) + (uint32_t)partial[blockIdx.x];
    uint32_t o[kScanIPT];
#pragma unroll
    for (int q = 0; q < kScanIPT; q++) {
        o[q] = run;
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

}  // namespace tmc
