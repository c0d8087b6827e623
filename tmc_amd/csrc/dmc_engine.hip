// dmc_engine.hip -- host runtime and C ABI of the MI355X DMC engine.
//
// One context = one HIP stream + a pool of device buffers that grow on
// demand and are reused across extractions.  dmc_extract runs
//   classify -> compact -> cell info -> scans -> vertices -> quads
//   [-> 3X3Y colour] [-> 3X3Y isolated] [-> 3333] -> halfedges [-> triangles]
// entirely on the device; the host only reads back the counts that size the
// next launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/tmc_dmc.h"
#include "dmc_extract_kernels.h"
#include "dmc_mesh_kernels.h"
#include "tmc_kernels.h"

using namespace tmc;

namespace {

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

// Default launch shape of a new context (env TMC_GRID_CAP /
// TMC_SWEEP_WAVES override it; dmc_set_launch and dmc_autotune change it per
// context).  Grid-stride kernels run on at most kGridCap blocks (persistent
// style): enough to fill every CU many times over, few enough that waves
// live long and keep many loads in flight (32768 measured fastest at 1024^3,
// 0.7% ahead of 16384, which is 3% ahead on a 512^3 noisy field).  The field
// sweep aims for kSweepWaves waves, which sets its planes per wave.
constexpr int kGridCap = 32768;
constexpr int kSweepWaves = 32768;
constexpr int64_t kGridCapMin = 64, kGridCapMax = 1 << 24;
constexpr int64_t kSweepWavesMin = 256, kSweepWavesMax = 1 << 24;
// captured passes kept per context, shapes remembered, and the passes of a
// shape run kernel by kernel before it is captured
constexpr size_t kMaxGraphs = 64;
constexpr size_t kMaxSeen = 4096;
constexpr int kCaptureAfter = 2;
// capacity floor of a pass: its cells, up to this many
constexpr int64_t kSmallPassCells = (int64_t)1 << 18;
// k_vertices blocks: 256 CUs x 2 resident blocks (40 KB of LDS, 256 VGPRs
// each) x 4 rounds
constexpr int64_t kVerticesBlocks = 2048;

struct StageMark {
    const char* name;
    hipEvent_t ev;
};

// Scratch of dmc_extract_noalloc: one array per device, in the code object,
// so that the entry point needs no device allocation.  Host code never uses
// the name `scratch` as an address: each device's address comes from
// hipGetSymbolAddress, taken once per device at context creation with that
// device current (scratch_init, kept in g_scratch_ptr).  Calls on one
// device share the array; g_scratch_mu serialises them (each call leaves
// the stream drained).
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int64_t kNoAllocMaxPoints = (int64_t)1 << 21;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu[kMaxDevices];
std::mutex g_scratch_init_mu;

// The device's scratch address (device current); nullptr if it cannot be
// had.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

}  // namespace

struct dmc_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // launch shape (results do not depend on it)
    int grid_cap = kGridCap;
    int sweep_waves = kSweepWaves;
    std::string err;
    std::vector<Buf*> bufs;
    // cell sweep
    Buf segCount, segBase, segMask, lists, acell, acode, aaux, avloc, aqloc, rowV, rowQ, rowBaseV, rowBaseQ;
    Buf field, grads, scanPartial, scanState, scanTotal, sizes;
    // volume table of the last extraction (one volume, or a batch) and the
    // per-volume first active cell / vertex / quadrilateral, kept through
    // compactions
    Buf vtab, vstat, vbound, qbound, vhe;
    std::vector<uint8_t> vtab_host;
    int nvol = 0;
    // mesh (double-buffered where compacted)
    Buf pos[2], nrm[2], quads[2], colors, vslot[2];
    int cur = 0;   // which quads buffer holds the live mesh
    int pcur = 0;  // which positions/normals buffer holds the vertex data
    int scur = 0;  // which slot map is live ...
    bool slots = false;  // ... once a compaction has run (else slots are the identity)
    // what the next halfedge build starts from: 1 the cell stage's rows in
    // off/adj, 2 the build a compaction carried over, 0 nothing
    int ready = 0;
    bool rows_live = false;   // off/adj hold the rows of the last extraction's final mesh
    bool force_generic = false;   // this extraction's passes pair every build through generic rows
    bool commit_pending = false;  // a compaction's counts wait for the next build's gate to commit them
    bool always_generic = false;  // every extraction pairs through generic rows (TMC_GENERIC_ROWS=1)
    int scan_spin = kScanSpin;    // polls a scan look-back waits before it sums the input (TMC_SCAN_SPIN)
    // halfedges / simplification scratch.  twin, flag, deg (out-degree), hev
    // and bndrec are the current build's; the ...2 buffers receive the build
    // a compaction carries over, and the pairs swap.
    Buf deg, off, cursor, adj, twin, flag, hev, ovr, bndrec, nbnd, val, p, em, vfirst, vsecond, lastT, lastR,
        newV, newQ, clist, cnbr, ccnt, tris, stats, off2, adj2, vmap, mlist, slow, twin2, flag2, deg2, hev2, bndrec2, arr,
        cflag, aflag, bcount, boff, alist;
    // diagnostics of the last mesh
    Buf drep, dnew, dval, dhist, dedges, dqual;
    // topologically correct MC: per active cell vertex / triangle counts and
    // offsets, the ambiguous-cell list, owner words; stats; the mesh
    Buf tcntv, tcntt, tvofs, ttofs, tamb, tvb, tstats, tpos, tnrm, ttris;
    tmc_mesh tmesh{};
    bool have_tmesh = false;
    // pinned host staging of device-to-host reads (counts, totals, tables)
    void* pin = nullptr;
    size_t pin_cap = 0;
    int64_t syncs = 0;         // stream synchronisations of HostReads so far
    int64_t syncs_at_start = 0;   // ... when the current extraction began
    // capacities of the pass's arrays (active cells, vertices,
    // quadrilaterals, 3-face records), planned from the largest densities of
    // active cells, vertices and quadrilaterals seen in one volume, per cell
    // (rcell) and per cells^(2/3) (rarea), and in the rerun of a pass that
    // overflowed, raised to its exact sizes (the *_min, one extraction long)
    int64_t acap = 0, vcap = 0, qcap = 0, reccap = 0;
    int64_t acap_min = 0, vcap_min = 0, qcap_min = 0;
    double rcell[3] = {0, 0, 0}, rarea[3] = {0, 0, 0};
    Sizes hsz{};               // the last pass's sizes block
    // captured passes (replayed when a pass of the same shape comes again)
    struct PassGraph {
        std::string key;
        hipGraphExec_t exec = nullptr;
        uint64_t use = 0;
        int cur = 0, pcur = 0, scur = 0, ready = 0;
        bool slots = false;
        uint32_t swapped = 0;   // which buffer pairs are in the order opposite to their addresses
    };
    std::vector<PassGraph> graphs;
    std::unordered_map<std::string, int> seen;   // passes run per shape (bounded)
    uint64_t buf_epoch = 0;    // bumped by every buffer (re)allocation
    uint64_t graph_clock = 0;
    bool use_graphs = true;
    bool capturing = false;
    int pass_mode = 0;         // the last pass: 0 launched one by one, 1 captured, 2 replayed
    // results
    dmc_mesh mesh{};
    bool have_mesh = false;
    std::vector<int32_t> bnd;  // sorted boundary vertices of the last final halfedge build
    // per-volume results of the last extraction (ids rebased to each volume)
    std::vector<dmc_mesh> vmesh;
    std::vector<std::vector<int32_t>> vbnd;
    // stage timing
    bool timing = false;
    std::vector<StageMark> marks;
    // kernel timers: events recorded right before and after one kernel (or
    // one scan), summed per name
    struct KMark {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<KMark> kmarks;
    std::vector<hipEvent_t> evpool;   // events reused across extractions
    size_t evused = 0;
    std::vector<std::pair<const char*, float>> stage_ms;
    // arena mode (dmc_extract_noalloc): buffers are carved from a fixed
    // device region instead of allocated
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;

    dmc_context() {
        Buf* all[] = {&segCount, &segBase, &segMask, &lists, &acell, &acode, &aaux, &avloc, &aqloc, &rowV, &rowQ, &rowBaseV, &rowBaseQ,
                      &field, &grads, &scanPartial, &scanState, &scanTotal, &sizes, &pos[0], &pos[1], &nrm[0], &nrm[1], &vslot[0], &vslot[1],
                      &quads[0], &quads[1], &colors, &deg, &off, &cursor, &adj, &twin, &flag, &hev, &ovr,
                      &bndrec, &nbnd, &val, &p, &em, &vfirst, &vsecond, &lastT, &lastR, &newV,
                      &newQ, &clist, &cnbr, &ccnt, &tris, &stats, &vtab, &vstat, &vbound, &qbound, &vhe,
                      &off2, &adj2, &vmap, &mlist, &slow, &twin2, &flag2, &deg2, &hev2, &bndrec2, &arr, &cflag,
                      &aflag, &bcount, &boff, &alist, &drep, &dnew, &dval,
                      &dhist, &dedges, &dqual, &tcntv, &tcntt, &tvofs, &ttofs, &tamb, &tvb, &tstats, &tpos, &tnrm,
                      &ttris};
        for (Buf* b : all) bufs.push_back(b);
    }
};

namespace {

struct Fail {
    dmc_status s;
};

void fail(dmc_context* ctx, dmc_status s, const char* fmt, ...) {
    char msg[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    ctx->err = msg;
    throw Fail{s};
}

#define HIPC(ctx, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) fail(ctx, DMC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
T* ensure(dmc_context* ctx, Buf& b, size_t n) {
    const size_t bytes = std::max<size_t>(n * sizeof(T), 256);
    if (ctx->arena) {
        // carve (never free): a buffer that must grow takes a new piece
        if (b.cap < bytes) {
            const size_t want = (bytes + 255) & ~(size_t)255;
            if (ctx->arena_used + want > ctx->arena_cap)
                fail(ctx, DMC_ERR_CAPACITY, "no-alloc scratch of %zu bytes exhausted (%zu used, %zu more needed)",
                     ctx->arena_cap, ctx->arena_used, want);
            b.p = ctx->arena + ctx->arena_used;
            b.cap = want;
            ctx->arena_used += want;
        }
        return static_cast<T*>(b.p);
    }
    if (b.cap < bytes) {
        if (ctx->capturing) fail(ctx, DMC_ERR_STATE, "a buffer would grow inside a captured pass");
        ctx->buf_epoch++;
        if (b.p) {
            HIPC(ctx, hipStreamSynchronize(ctx->stream));
            HIPC(ctx, hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        const size_t want = bytes + bytes / 4;
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            fail(ctx, DMC_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        b.cap = want;
    }
    return static_cast<T*>(b.p);
}

int grid_for(const dmc_context* ctx, int64_t n, int per_block = kBlock, int64_t cap = 0) {
    if (cap <= 0) cap = ctx->grid_cap;
    int64_t g = (n + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

hipEvent_t take_event(dmc_context* ctx) {
    if (ctx->evused == ctx->evpool.size()) {
        hipEvent_t e = nullptr;
        HIPC(ctx, hipEventCreate(&e));
        ctx->evpool.push_back(e);
    }
    return ctx->evpool[ctx->evused++];
}

// Stage boundary: a stage's time runs to the next mark.
void mark(dmc_context* ctx, const char* name) {
    if (!ctx->timing) return;
    StageMark m{name, take_event(ctx)};
    HIPC(ctx, hipEventRecord(m.ev, ctx->stream));
    ctx->marks.push_back(m);
}

// Kernel timer around the launches between kbegin and kend (same stream).
int kbegin(dmc_context* ctx, const char* name) {
    if (!ctx->timing) return -1;
    dmc_context::KMark k{name, take_event(ctx), take_event(ctx)};
    HIPC(ctx, hipEventRecord(k.a, ctx->stream));
    ctx->kmarks.push_back(k);
    return (int)ctx->kmarks.size() - 1;
}

void kend(dmc_context* ctx, int k) {
    if (k >= 0) HIPC(ctx, hipEventRecord(ctx->kmarks[k].b, ctx->stream));
}

// Every kernel launch of the engine: on the context's stream, timed under
// `name` when the extraction records stage times.
template <class K, class... Args>
void launch(dmc_context* ctx, const char* name, K kernel, dim3 grid, dim3 block, Args&&... args) {
    const int kt = kbegin(ctx, name);
    kernel<<<grid, block, 0, ctx->stream>>>(std::forward<Args>(args)...);
    kend(ctx, kt);
}

void check_launch(dmc_context* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(ctx, DMC_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

// Device-to-host reads, gathered into pinned host memory so that each is a
// direct transfer (a pageable destination would be staged through the
// runtime's bounce buffer), then one stream synchronisation for all.
class HostReads {
   public:
    explicit HostReads(dmc_context* ctx) : ctx_(ctx) {}
    void add(void* dst, const void* src, size_t bytes) {
        if (bytes == 0) return;
        outs_.push_back({dst, src, used_, bytes});
        used_ += (bytes + 255) & ~size_t(255);
    }
    void wait() {
        if (outs_.empty()) return;
        if (ctx_->pin_cap < used_ && ctx_->arena) {
            // no-alloc call: the pinned staging is not grown; copy straight
            // into the (pageable) destinations
            for (const Out& o : outs_)
                HIPC(ctx_, hipMemcpyAsync(o.dst, o.src, o.bytes, hipMemcpyDeviceToHost, ctx_->stream));
            HIPC(ctx_, hipStreamSynchronize(ctx_->stream));
            ctx_->syncs++;
            outs_.clear();
            used_ = 0;
            return;
        }
        if (ctx_->pin_cap < used_) {
            HIPC(ctx_, hipStreamSynchronize(ctx_->stream));
            if (ctx_->pin) {
                (void)hipHostFree(ctx_->pin);
                ctx_->pin = nullptr;
                ctx_->pin_cap = 0;
            }
            const size_t want = std::max<size_t>(used_, 64 << 10);
            if (hipHostMalloc(&ctx_->pin, want) != hipSuccess) {
                (void)hipGetLastError();
                ctx_->pin = nullptr;
                fail(ctx_, DMC_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu) failed", want);
            }
            ctx_->pin_cap = want;
        }
        char* pin = static_cast<char*>(ctx_->pin);
        for (const Out& o : outs_)
            HIPC(ctx_, hipMemcpyAsync(pin + o.at, o.src, o.bytes, hipMemcpyDeviceToHost, ctx_->stream));
        HIPC(ctx_, hipStreamSynchronize(ctx_->stream));
        ctx_->syncs++;
        for (const Out& o : outs_) std::memcpy(o.dst, pin + o.at, o.bytes);
        outs_.clear();
        used_ = 0;
    }

   private:
    struct Out {
        void* dst;
        const void* src;
        size_t at, bytes;
    };
    dmc_context* ctx_;
    std::vector<Out> outs_;
    size_t used_ = 0;
};

uint64_t read_u64(dmc_context* ctx, const void* dptr) {
    uint64_t v = 0;
    HostReads r(ctx);
    r.add(&v, dptr, sizeof(v));
    r.wait();
    return v;
}

// Device sizes block of the running pass (allocated with the pass).
Sizes* dsizes(dmc_context* ctx) { return static_cast<Sizes*>(ctx->sizes.p); }

// The double-buffered halfedge arrays swap at every compaction; a pass
// starts from one order (lower address current) so that passes of one shape
// launch on the same pointers.
Buf dmc_context::* const kPairs[][2] = {{&dmc_context::twin, &dmc_context::twin2},
                                        {&dmc_context::flag, &dmc_context::flag2},
                                        {&dmc_context::deg, &dmc_context::deg2},
                                        {&dmc_context::hev, &dmc_context::hev2},
                                        {&dmc_context::bndrec, &dmc_context::bndrec2}};
constexpr int kNumPairs = sizeof(kPairs) / sizeof(kPairs[0]);

uint32_t pair_order(dmc_context* ctx) {
    uint32_t m = 0;
    for (int i = 0; i < kNumPairs; i++)
        if ((ctx->*kPairs[i][0]).p > (ctx->*kPairs[i][1]).p) m |= 1u << i;
    return m;
}

void set_pair_order(dmc_context* ctx, uint32_t m) {
    for (int i = 0; i < kNumPairs; i++) {
        Buf& a = ctx->*kPairs[i][0];
        Buf& b = ctx->*kPairs[i][1];
        if ((a.p > b.p) != (((m >> i) & 1u) != 0)) std::swap(a, b);
    }
}

void swap_pairs(dmc_context* ctx) {
    for (int i = 0; i < kNumPairs; i++) std::swap(ctx->*kPairs[i][0], ctx->*kPairs[i][1]);
}


DevN dn(const uint64_t* p, int64_t cap) { return DevN{p, cap}; }
DevN hn(int64_t n) { return DevN{nullptr, n}; }

// Fill n (device count, capped) 32-bit words or bytes with a value.
void fill_u32(dmc_context* ctx, void* p, DevN n, uint32_t value, const uint32_t* gate = nullptr) {
    if (n.cap <= 0) return;
    launch(ctx, "k_fill", k_fill_u32, dim3(grid_for(ctx, n.cap)), dim3(kBlock), static_cast<uint32_t*>(p), n, value,
           gate);
}
void fill_u8(dmc_context* ctx, void* p, DevN n, uint8_t value) {
    if (n.cap <= 0) return;
    launch(ctx, "k_fill", k_fill_u8, dim3(grid_for(ctx, n.cap)), dim3(kBlock), static_cast<uint8_t*>(p), n, value);
}

// Fills gathered into one launch (k_fill_many).
struct Fills {
    dmc_context* ctx;
    FillSet s{};
    int n = 0;
    int64_t cap = 0;
    explicit Fills(dmc_context* c) : ctx(c) {}
    Fills& u32(void* p, DevN d, uint32_t value, const uint32_t* gate = nullptr) { return add(p, d, value, gate, 4); }
    Fills& u8(void* p, DevN d, uint8_t value, const uint32_t* gate = nullptr) { return add(p, d, value, gate, 1); }
    Fills& add(void* p, DevN d, uint32_t value, const uint32_t* gate, int bytes) {
        if (d.cap <= 0) return *this;
        if (n == kMaxFillJobs) run();
        s.j[n++] = FillJob{p, d, value, gate, bytes};
        cap = std::max(cap, d.cap);
        return *this;
    }
    void run() {
        if (n == 0) return;
        launch(ctx, "k_fill", k_fill_many, dim3(grid_for(ctx, cap), n), dim3(kBlock), s);
        n = 0;
        cap = 0;
    }
};

// Inputs of at most one tile are scanned by one block in one launch.
constexpr int64_t kScanSmallItems = kScanTile;

// The state of the single-pass scans: zero when a scan starts, and left zero
// by every scan, so it is cleared only when its buffer is (re)allocated or
// carved from the no-alloc scratch.
unsigned long long* scan_state(dmc_context* ctx, int64_t tiles) {
    Buf& b = ctx->scanState;
    const void* p0 = b.p;
    const size_t c0 = b.cap;
    unsigned long long* st = ensure<unsigned long long>(ctx, b, kScanStateHead + (size_t)tiles);
    if (b.p != p0 || b.cap != c0) fill_u32(ctx, reinterpret_cast<uint32_t*>(st), hn((int64_t)(b.cap / 4)), 0u);
    return st;
}

// Exclusive scan of n.get() uint32 items into out (launched for n.cap); the
// 64-bit total goes to *dtotal (device memory); row_end: also the total's low
// 32 bits to row_end[n] (the end of a CSR's last row).  Nothing runs when
// *gate is 0.  One block for a few tiles, one launch of look-back tiles for
// up to kScanOnePassTiles, three launches beyond.
template <class In>
void scan_to(dmc_context* ctx, In in, uint32_t* out, DevN n, uint64_t* dtotal, const char* name = "k_scan",
             const uint32_t* gate = nullptr, uint32_t* row_end = nullptr) {
    if (n.cap <= kScanSmallItems) {
        launch(ctx, name, k_scan_small<In>, dim3(1), dim3(1024), in, out, n, gate, dtotal, row_end);
        check_launch(ctx, "scan");
        return;
    }
    const int64_t tiles = (n.cap + kScanTile - 1) / kScanTile;
    if (tiles <= kScanOnePassTiles) {
        // a stalled look-back re-reads the input
        if (static_cast<const void*>(in.p) == static_cast<const void*>(out))
            fail(ctx, DMC_ERR_STATE, "%s: a scan in place", name);
        launch(ctx, name, k_scan<In>, dim3((unsigned)tiles), dim3(kBlock), in, out, n, gate, dtotal, row_end,
               scan_state(ctx, tiles), tiles, ctx->scan_spin);
    } else {
        uint64_t* part = ensure<uint64_t>(ctx, ctx->scanPartial, tiles);
        launch(ctx, name, k_scan_reduce<In>, dim3((unsigned)tiles), dim3(kBlock), in, n, gate, part);
        launch(ctx, name, k_scan_partials, dim3(1), dim3(1024), part, tiles, gate, dtotal, row_end, n);
        launch(ctx, name, k_scan_down<In>, dim3((unsigned)tiles), dim3(kBlock), in, out, n, gate, part);
    }
    check_launch(ctx, "scan");
}

template <class In>
ScanJob<In> scan_job(In in, uint32_t* out, DevN n, uint64_t* dtotal, const uint32_t* gate = nullptr,
                     uint32_t* row_end = nullptr) {
    ScanJob<In> j{};
    j.in = in;
    j.out = out;
    j.dn = n;
    j.gate = gate;
    j.total = dtotal;
    j.row_end = row_end;
    j.ntiles = (n.cap + kScanTile - 1) / kScanTile;
    return j;
}

// Two independent scans (neither reads what the other writes) in the
// launches of one: one look-back launch when both fit kScanOnePassTiles,
// else the three passes for both.
template <class A, class B>
void scan_pair(dmc_context* ctx, ScanJob<A> a, ScanJob<B> b, const char* name = "k_scan") {
    a.ntiles = std::max<int64_t>(a.ntiles, 1);
    b.ntiles = std::max<int64_t>(b.ntiles, 1);
    const unsigned blocks = (unsigned)(a.ntiles + b.ntiles);
    if (a.ntiles <= kScanOnePassTiles && b.ntiles <= kScanOnePassTiles) {
        if (static_cast<const void*>(a.in.p) == static_cast<const void*>(a.out) ||
            static_cast<const void*>(b.in.p) == static_cast<const void*>(b.out))
            fail(ctx, DMC_ERR_STATE, "%s: a scan in place", name);
        // each job's state: kScanStateHead words, then its tile flags
        a.st = scan_state(ctx, kScanStateHead + a.ntiles + b.ntiles);
        b.st = a.st + kScanStateHead + a.ntiles;
        launch(ctx, name, k_scan_pair<A, B>, dim3(blocks), dim3(kBlock), a, b, ctx->scan_spin);
    } else {
        a.partial = ensure<uint64_t>(ctx, ctx->scanPartial, (size_t)blocks);
        b.partial = a.partial + a.ntiles;
        launch(ctx, name, k_scan_reduce_pair<A, B>, dim3(blocks), dim3(kBlock), a, b);
        launch(ctx, name, k_scan_partials_pair<A, B>, dim3(2), dim3(1024), a, b);
        launch(ctx, name, k_scan_down_pair<A, B>, dim3(blocks), dim3(kBlock), a, b);
    }
    check_launch(ctx, "scan");
}

// Host-counted scan with the total read back (diagnostics and the
// topologically correct MC, which size their launches on the host).
template <class In>
uint64_t scan_read(dmc_context* ctx, In in, uint32_t* out, int64_t n) {
    if (n <= 0) return 0;
    uint64_t* tot = ensure<uint64_t>(ctx, ctx->scanTotal, 1);
    scan_to(ctx, in, out, hn(n), tot);
    return read_u64(ctx, tot);
}

uint64_t scan(dmc_context* ctx, const uint32_t* in, uint32_t* out, int64_t n) {
    return scan_read(ctx, ScanCounts{in}, out, n);
}

// ---------------------------------------------------------------------------
// cell stage
// ---------------------------------------------------------------------------
// One volume of an extraction, with device pointers.
struct HostVol {
    dmc_grid g;
    const void* field;
    const double* grads;
    double iso;
};

float smallest_float_not_below(double iso) {
    float f = (float)iso;
    if (std::isnan(iso)) return f;
    if ((double)f < iso) f = std::nextafter(f, INFINITY);
    return f;
}

int64_t* vol_ptr(dmc_context* ctx, int which) {  // 0 waveBase, 1 rowBase, 2 aBase
    const size_t n = (size_t)ctx->nvol;
    uint8_t* base = static_cast<uint8_t*>(ctx->vtab.p) + n * sizeof(VolDesc);
    return reinterpret_cast<int64_t*>(base) + (size_t)which * (n + 1);
}

VolTable vol_table(dmc_context* ctx) {
    return VolTable{static_cast<const VolDesc*>(ctx->vtab.p), vol_ptr(ctx, 0), vol_ptr(ctx, 1), vol_ptr(ctx, 2),
                    ctx->nvol};
}

// What the field sweep leaves for the cell stages.
struct Swept {
    int64_t R;      // cell rows (all volumes)
    int64_t cells;  // cells (all volumes)
    VolTable vt;
    VolStat* vstat;
    uint32_t *lists, *segCount, *segBase;
    unsigned long long* segMask;
};

// Host side of a pass: the volume table (built in ctx->vtab_host, uploaded
// by upload_vtab) and the sweep's tiling.
struct SweepPlan {
    int64_t W, R, RS, cells;
    bool vec;
    size_t tab_bytes;
};

template <class T>
SweepPlan plan_sweep(dmc_context* ctx, const std::vector<HostVol>& hv) {
    const int n = (int)hv.size();
    ctx->nvol = n;
    // sweep tiling: J rows per wave; one plane count K for the whole
    // extraction, chosen so that about ctx->sweep_waves waves fill the chip
    std::vector<VolDesc> desc(n);
    int64_t units = 0, cells = 0;
    for (int v = 0; v < n; v++) {
        const dmc_grid& g = hv[v].g;
        VolDesc& D = desc[v];
        D.field = hv[v].field;
        D.grads = hv[v].grads;
        D.d = Dims{g.nx, g.ny, g.nz, (int64_t)g.nx - 1, (int64_t)(g.ny - 1) * (g.nz - 1)};
        D.geo = GridGeom{g.x0, g.y0, g.z0, g.dx, g.dy, g.dz};
        // reciprocals for the corner gradients' divisions (k_vertices
        // div_by), where the spacing is in a range that keeps them exact
        const double sp[3] = {g.dx, g.dy, g.dz};
        for (int c = 0; c < 3; c++) {
            const double c2 = 2 * sp[c], a2 = std::fabs(c2);
            D.rdx2[c] = (std::isnormal(c2) && a2 > 0x1p-500 && a2 < 0x1p+500) ? 1.0 / c2 : 0.0;
        }
        D.iso = hv[v].iso;
        D.iso_t = sizeof(T) == 4 ? (double)smallest_float_not_below(hv[v].iso) : hv[v].iso;
        D.S = (int)((D.d.C + kSeg - 1) / kSeg);
        D.ntj = (g.ny - 1 + kSweepRows - 1) / kSweepRows;
        units += (int64_t)(g.nz - 1) * D.S * D.ntj;
        cells += D.d.C * D.d.R;
    }
    // (at least 2 planes per wave, so that the shared boundary plane costs at
    // most half again; small volumes get short marches and many waves)
    const int Kg = (int)std::max<int64_t>(2, (units + ctx->sweep_waves - 1) / ctx->sweep_waves);
    const size_t tab_bytes = (size_t)n * sizeof(VolDesc) + 3 * (size_t)(n + 1) * sizeof(int64_t);
    ctx->vtab_host.assign(tab_bytes, 0);
    int64_t* waveBase = reinterpret_cast<int64_t*>(ctx->vtab_host.data() + (size_t)n * sizeof(VolDesc));
    int64_t* rowBase = waveBase + (n + 1);
    int64_t W = 0, R = 0, RS = 0;
    bool vec = true;
    for (int v = 0; v < n; v++) {
        VolDesc& D = desc[v];
        D.K = std::min(Kg, D.d.nz - 1);
        D.ntk = (D.d.nz - 1 + D.K - 1) / D.K;
        D.rowBase = R;
        D.rsBase = RS;
        waveBase[v] = W;
        rowBase[v] = R;
        W += (int64_t)D.S * D.ntj * D.ntk;
        R += D.d.R;
        RS += D.d.R * D.S;
        vec = vec && (D.d.nx % 4 == 0) && ((uintptr_t)D.field % (4 * sizeof(T)) == 0);
    }
    waveBase[n] = W;
    rowBase[n] = R;
    memcpy(ctx->vtab_host.data(), desc.data(), (size_t)n * sizeof(VolDesc));
    ensure<uint8_t>(ctx, ctx->vtab, tab_bytes);
    return SweepPlan{W, R, RS, cells, vec, tab_bytes};
}

// The volume table to the device (before the pass's launches; a replayed
// pass reads it from there).  Pageable source: the copy has been staged
// when the call returns.
void upload_vtab(dmc_context* ctx, const SweepPlan& P) {
    HIPC(ctx, hipMemcpyAsync(ctx->vtab.p, ctx->vtab_host.data(), P.tab_bytes, hipMemcpyHostToDevice, ctx->stream));
}

// Sizes block, classification sweep and the scan of the active cells
// (total to S->A) -- shared by Dual Marching Cubes and the topologically
// correct MC.
template <class T>
Swept sweep_cells(dmc_context* ctx, const SweepPlan& P) {
    const int n = ctx->nvol;
    const int64_t W = P.W, R = P.R, RS = P.RS;
    const VolTable vt = vol_table(ctx);
    VolStat* vstat = ensure<VolStat>(ctx, ctx->vstat, n);
    Sizes* S = reinterpret_cast<Sizes*>(ensure<uint8_t>(ctx, ctx->sizes, sizeof(Sizes)));
    static_assert(sizeof(Sizes) % 4 == 0 && sizeof(VolStat) % 4 == 0, "filled as words");
    Fills(ctx).u32(S, hn(sizeof(Sizes) / 4), 0u).u32(vstat, hn((int64_t)n * sizeof(VolStat) / 4), 0u).run();

    uint32_t* segCount = ensure<uint32_t>(ctx, ctx->segCount, RS);
    uint32_t* segBase = ensure<uint32_t>(ctx, ctx->segBase, RS);
    uint32_t* lists = ensure<uint32_t>(ctx, ctx->lists, (size_t)RS * kSeg);
    unsigned long long* segMask = ensure<unsigned long long>(ctx, ctx->segMask, 4 * (size_t)RS);
    mark(ctx, "sweep");
    const int64_t sblocks = (W + kBlock / kWave - 1) / (kBlock / kWave);
    const dim3 sgrid((unsigned)((sblocks + kXcds - 1) / kXcds * kXcds));
    if (P.vec)
        launch(ctx, "k_sweep", (k_sweep<T, true>), sgrid, dim3(kBlock), vt, W, lists, segCount, segMask);
    else
        launch(ctx, "k_sweep", (k_sweep<T, false>), sgrid, dim3(kBlock), vt, W, lists, segCount, segMask);
    check_launch(ctx, "k_sweep");
    mark(ctx, "rows");
    scan_to(ctx, ScanCounts{segCount}, segBase, hn(RS), &S->A, "k_scan_cells");
    if (n > 1)
        launch(ctx, "k_vol_abase", k_vol_abase, dim3(grid_for(ctx, n + 1)), dim3(kBlock), vt, segBase, &S->A,
               vol_ptr(ctx, 2));
    return Swept{R, P.cells, vt, vstat, lists, segCount, segBase, segMask};
}

// Capacities of the cell-stage arrays for a pass over the volumes `hv`: per
// volume the smaller of two estimates from the densest volumes this context
// has seen (with a margin) -- count per cell x cells, and count per
// cells^(2/3) x cells^(2/3), since a surface's cells grow with its area, so
// that a small dense volume does not size a large one by its density -- and
// over the pass at least a default share of the cells; a pass that exceeds
// one is run again with the exact size.  Capacities are rounded up to one of
// 8 steps per power of two, so that passes of one shape keep the same
// capacities (and buffers) while the context's densities creep up.
int64_t cap_step(int64_t x) {
    if (x <= 4096) return 4096;
    int b = 63 - __builtin_clzll((unsigned long long)x);
    const int64_t unit = (int64_t)1 << (b - 3);
    return (x + unit - 1) / unit * unit;
}

double area_of(double cells) { return std::cbrt(cells * cells); }

void plan_caps(dmc_context* ctx, const std::vector<HostVol>& hv, int64_t cells) {
    double est[3] = {0, 0, 0};
    for (const HostVol& h : hv) {
        const double c = (double)(h.g.nx - 1) * (double)(h.g.ny - 1) * (double)(h.g.nz - 1);
        const double a = area_of(c);
        for (int k = 0; k < 3; k++) est[k] += std::min(ctx->rcell[k] * c, ctx->rarea[k] * a) * 1.02;
    }
    auto cap = [&](int k) {
        return cap_step((int64_t)std::ceil(std::max(est[k], 0.02 * (double)cells)) + 1024);
    };
    // small passes: room for every cell (a few MB), so that a dense small
    // volume never runs twice -- except in the fixed scratch of a no-alloc
    // call, where a rerun is cheaper than the room
    const int64_t floor = ctx->arena ? 0 : cap_step(std::min<int64_t>(cells, kSmallPassCells));
    ctx->acap = std::max({cap_step(ctx->acap_min), cap(0), floor});
    ctx->vcap = std::max({cap_step(ctx->vcap_min), cap(1), floor});
    ctx->qcap = std::max({cap_step(ctx->qcap_min), cap(2), floor});
    ctx->acap = std::min<int64_t>(std::min<int64_t>(ctx->acap, cells), 0xFFFFFFFF);
    ctx->vcap = std::min<int64_t>(ctx->vcap, 0x7FFFFFFF);
    ctx->qcap = std::min<int64_t>(ctx->qcap, 0x1FFFFFFF);
    ctx->reccap = std::max<int64_t>(ctx->reccap, 1024);
}

template <class T>
void run_cells(dmc_context* ctx, const SweepPlan& P) {
    // canonical buffer roles at the start of every pass, so that a pass of
    // the same shape launches on the same pointers (and can be replayed)
    set_pair_order(ctx, 0);
    const Swept sw = sweep_cells<T>(ctx, P);
    const int n = ctx->nvol;
    const int64_t acap = ctx->acap, vcap = ctx->vcap, qcap = ctx->qcap;
    const int64_t R = sw.R;
    const VolTable vt = sw.vt;
    Sizes* S = dsizes(ctx);
    uint32_t* acell = ensure<uint32_t>(ctx, ctx->acell, acap);
    uint32_t* acode = ensure<uint32_t>(ctx, ctx->acode, acap);
    uint32_t* aaux = ensure<uint32_t>(ctx, ctx->aaux, acap);
    uint32_t* avloc = ensure<uint32_t>(ctx, ctx->avloc, acap);
    uint32_t* aqloc = ensure<uint32_t>(ctx, ctx->aqloc, acap);
    uint32_t* rowV = ensure<uint32_t>(ctx, ctx->rowV, R);
    uint32_t* rowQ = ensure<uint32_t>(ctx, ctx->rowQ, R);
    uint32_t* rowBaseV = ensure<uint32_t>(ctx, ctx->rowBaseV, R);
    uint32_t* rowBaseQ = ensure<uint32_t>(ctx, ctx->rowBaseQ, R);
    launch(ctx, "k_rows", k_rows<T>, dim3(grid_for(ctx, R, kBlock / kWave)), dim3(kBlock), vt, R, sw.lists,
           sw.segCount, sw.segBase, acell, acode, aaux, avloc, aqloc, rowV, rowQ, sw.vstat, S->cpart, (uint32_t)acap);
    check_launch(ctx, "k_rows");
    scan_pair(ctx, scan_job(ScanCounts{rowV}, rowBaseV, hn(R), &S->V), scan_job(ScanCounts{rowQ}, rowBaseQ, hn(R), &S->Q),
              "k_scan_cells");
    // cells deferred to the general sampling: slow[0] = count (zeroed by the
    // guard), then cell ids
    uint32_t* slow = ensure<uint32_t>(ctx, ctx->slow, acap + 1);
    launch(ctx, "k_cell_guard", k_cell_guard, dim3(1), dim3(64), S, (uint64_t)acap, (uint64_t)vcap, (uint64_t)qcap,
           slow);
    // [0, n]: live bounds; [n+1, 2n+1]: the cell stage's (before simplification)
    int64_t* vb = ensure<int64_t>(ctx, ctx->vbound, 2 * (size_t)(n + 1));
    int64_t* qb = ensure<int64_t>(ctx, ctx->qbound, 2 * (size_t)(n + 1));
    if (n > 1)
        launch(ctx, "k_vol_bounds", k_vol_bounds, dim3(grid_for(ctx, n + 1)), dim3(kBlock), vt, rowBaseV, rowBaseQ, S,
               vb, qb);
    ctx->cur = 0;
    ctx->pcur = 0;
    ctx->scur = 0;
    ctx->slots = false;
    double* pos = ensure<double>(ctx, ctx->pos[0], 3 * vcap);
    double* nrm = ensure<double>(ctx, ctx->nrm[0], 3 * vcap);
    int32_t* quads = ensure<int32_t>(ctx, ctx->quads[0], 4 * qcap);
    uint8_t* colors = ensure<uint8_t>(ctx, ctx->colors, qcap);
    // first halfedge CSR straight from the cells: degree of each vertex from
    // k_vertices, rows filled by k_quads
    uint32_t* deg = ensure<uint32_t>(ctx, ctx->deg, vcap);
    uint32_t* off = ensure<uint32_t>(ctx, ctx->off, vcap + 1);
    int2* adj = ensure<int2>(ctx, ctx->adj, 4 * qcap);
    const DevN A = dn(&S->acnt, acap);
    mark(ctx, "vertices");
    // persistent: a few rounds of the blocks the chip holds at once; the
    // 3-wave form once there are more cells than one round of the grid's
    // threads (F32 fields: the F64 form would spill at 3 waves)
    const dim3 gv(grid_for(ctx, acap, kBlock, kVerticesBlocks));
    if (std::is_same<T, float>::value && acap > kVerticesBlocks * kBlock)
        launch(ctx, "k_vertices", k_vertices_wide<T>, gv, dim3(kBlock), vt, acell, acode, aaux, avloc, rowBaseV, A,
               pos, nrm, deg, slow, &S->qhead);
    else
        launch(ctx, "k_vertices", k_vertices<T>, gv, dim3(kBlock), vt, acell, acode, aaux, avloc, rowBaseV, A, pos,
               nrm, deg, slow, &S->qhead);
    check_launch(ctx, "k_vertices");
    launch(ctx, "k_vertices_slow", k_vertices_slow<T>, dim3(grid_for(ctx, acap, kBlock, 1024)), dim3(kBlock), vt,
           acell, acode, aaux, avloc, rowBaseV, slow, pos, nrm, deg, sw.vstat);
    check_launch(ctx, "k_vertices_slow");
    mark(ctx, "quads");
    // the row total must be 4 Q (checked with the pass's sizes at the end)
    scan_to(ctx, ScanCounts{deg}, off, dn(&S->nv, vcap), &S->rows, "k_scan_cells", nullptr, off);
    launch(ctx, "k_quads", k_quads, dim3(grid_for(ctx, acap)), dim3(kBlock), vt, acell, acode, aaux, avloc, aqloc,
           rowBaseV, rowBaseQ, A, sw.segMask, sw.segBase, quads, colors, off, adj, (uint64_t)(4 * qcap),
           &S->unresolved);
    check_launch(ctx, "k_quads");
    ctx->ready = 1;
}

// ---------------------------------------------------------------------------
// halfedges
// ---------------------------------------------------------------------------
// A halfedge build gives every halfedge its twin and flag, every vertex its
// out-degree, valence and he_v, and lists the 3-face edges.  The first build
// pairs through the rows of outgoing (halfedge, destination) pairs the cell
// stage left (off/adj).  A compaction carries the build over to the mesh it
// leaves: most groups keep their members, renumbered, and only halfedges at
// a changed vertex are paired again.  If the device finds the rows or the
// carry unusable (a quadrilateral with an unresolved neighbour cell, an
// invalid vertex reference), the build counts generic rows from the
// quadrilaterals and pairs everything.

// Pairing through full rows (on the device only when S->pair_rows): the
// cell stage's rows, or (generic: launched only in a pass run again for
// them) rows counted from the quadrilaterals.
void pair_rows(dmc_context* ctx, bool final_pass, bool generic) {
    const int64_t vcap = ctx->vcap, qcap = ctx->qcap, hcap = 4 * qcap;
    Sizes* S = dsizes(ctx);
    const int32_t* quads = static_cast<const int32_t*>(ctx->quads[ctx->cur].p);
    uint32_t* deg = ensure<uint32_t>(ctx, ctx->deg, vcap);
    uint32_t* off = ensure<uint32_t>(ctx, ctx->off, vcap + 1);
    uint32_t* cursor = ensure<uint32_t>(ctx, ctx->cursor, vcap);
    int2* adj = ensure<int2>(ctx, ctx->adj, hcap);
    int32_t* twin = ensure<int32_t>(ctx, ctx->twin, hcap);
    uint8_t* flag = ensure<uint8_t>(ctx, ctx->flag, hcap);
    int32_t* hev = ensure<int32_t>(ctx, ctx->hev, vcap);
    uint32_t* arr = final_pass ? nullptr : ensure<uint32_t>(ctx, ctx->arr, vcap);
    BndRec* recs = ensure<BndRec>(ctx, ctx->bndrec, ctx->reccap);
    const uint32_t* gen = &S->generic;
    const uint32_t* rows = &S->pair_rows;
    const DevN nv = dn(&S->nv, vcap);
    if (generic) {
        Fills(ctx).u32(deg, nv, 0u, gen).u32(cursor, nv, 0u, gen).run();
        launch(ctx, "k_he_degree", k_he_degree, dim3(grid_for(ctx, hcap)), dim3(kBlock), quads, (const Sizes*)S,
               deg, &S->he.bad, gen);
        scan_to(ctx, ScanCounts{deg}, off, nv, &S->tot, "k_scan", gen, off);
        launch(ctx, "k_he_scatter", k_he_scatter, dim3(grid_for(ctx, qcap)), dim3(kBlock), quads, (const Sizes*)S,
               off, cursor, adj, gen);
    }
    Fills fl(ctx);
    fl.u32(hev, nv, 0xFFFFFFFFu, rows);
    if (arr) fl.u32(arr, nv, 0u, rows);
    unsigned long long* vst = nullptr;
    if (final_pass && ctx->nvol > 1) {
        vst = ensure<unsigned long long>(ctx, ctx->vhe, 3 * (size_t)ctx->nvol);
        fl.u32(vst, hn(6 * (int64_t)ctx->nvol), 0u, rows);
    }
    fl.run();
    launch(ctx, "k_he_twins", k_he_twins, dim3(grid_for(ctx, hcap)), dim3(kBlock), quads, (const Sizes*)S, off, adj,
           twin, flag, hev, arr, S->hep, recs, &S->nrec, (uint32_t)ctx->reccap, &S->ovf,
           static_cast<const int64_t*>(ctx->qbound.p), ctx->nvol, vst, rows);
    check_launch(ctx, "k_he_twins");
}

// A halfedge build of the live mesh.  Not final: ends with the valence,
// which also clears what the next simplification needs cleared.  Final: the
// 3-face overrides of he_v and the reference's flag values.
void build_halfedges(dmc_context* ctx, bool final_pass, const StageClears& clears = StageClears{}) {
    Sizes* S = dsizes(ctx);
    const int64_t vcap = ctx->vcap;
    const int ready = ctx->force_generic ? 0 : ctx->ready;
    launch(ctx, "k_rows_gate", k_rows_gate, dim3(1), dim3(64), S, ready, ctx->commit_pending ? 1 : 0);
    ctx->commit_pending = false;
    ctx->ready = 0;
    // a carried build pairs nothing here
    if (ready != 2) pair_rows(ctx, final_pass, ready == 0);
    const int gv = grid_for(ctx, vcap);
    if (!final_pass) {
        uint32_t* val = ensure<uint32_t>(ctx, ctx->val, vcap);
        launch(ctx, "k_valence", k_valence, dim3(gv), dim3(kBlock), static_cast<const uint32_t*>(ctx->deg.p),
               static_cast<const uint32_t*>(ctx->arr.p), (const Sizes*)S, val, &S->zero, clears);
        check_launch(ctx, "valence");
        return;
    }
    const int32_t* quads = static_cast<const int32_t*>(ctx->quads[ctx->cur].p);
    int32_t* hev = static_cast<int32_t*>(ctx->hev.p);
    const BndRec* recs = static_cast<const BndRec*>(ctx->bndrec.p);
    unsigned long long* ovr = ensure<unsigned long long>(ctx, ctx->ovr, vcap);
    fill_u32(ctx, ovr, DevN{nullptr, 2 * vcap}, 0u, &S->nrec);
    const int gr = grid_for(ctx, ctx->reccap);
    for (int pass = 0; pass < 2; pass++)
        launch(ctx, "k_three_apply", k_three_apply, dim3(gr), dim3(kBlock), recs, (const Sizes*)S,
               (uint32_t)ctx->reccap, quads, ovr, pass, hev);
    launch(ctx, "k_strip_flags", k_strip_flags, dim3(grid_for(ctx, 4 * ctx->qcap)), dim3(kBlock), (const Sizes*)S,
           static_cast<uint8_t*>(ctx->flag.p));
    check_launch(ctx, "final halfedges");
}

// Meshes of at most this many faces are coloured by one block.
constexpr int64_t kSmallColourFaces = (int64_t)1 << 14;

// Colour classes that run in one launch.  Class c holds the faces of initial
// colour c = 3 * parity + direction of their grid edge, and a face reads the
// colours of its edge neighbours.  Two quadrilaterals share a mesh edge only
// when their grid edges lie on one cell face: the face's other edge of the
// same direction (lower corners differing in the bit of one of the two other
// axes), or an edge of another direction b whose lower corner differs by at
// most the bits of the two axes of the face they span.  A class sees the
// colours it would see in the sequential order (5..23, then 4) when it runs
// after every earlier class it touches and before every later one: it runs
// one level past the highest level of the earlier classes it touches, and
// the classes of a level share a launch.
bool classes_touch(int c1, int c2) {
    const int a = c1 % 3, b = c2 % 3, p = c1 / 3, q = c2 / 3;
    const int ea = 1 << a, eb = 1 << b;
    if (a == b) {
        const int o1 = 1 << ((a + 1) % 3), o2 = 1 << ((a + 2) % 3);
        return q == (p ^ o1) || q == (p ^ o2);
    }
    const int d = p ^ q;
    return (d & ~(ea | eb)) == 0;
}

struct ColourLevels {
    int n = 0;
    int cls[20][2];
};

const ColourLevels& colour_levels() {
    static const ColourLevels L = [] {
        int order[20], level[24] = {0};
        for (int i = 0; i < 19; i++) order[i] = 5 + i;
        order[19] = 4;
        ColourLevels r;
        for (auto& x : r.cls) x[0] = x[1] = -1;
        for (int i = 0; i < 20; i++) {
            int lv = 0;
            for (int j = 0; j < i; j++)
                if (classes_touch(order[i], order[j])) lv = std::max(lv, level[order[j]] + 1);
            level[order[i]] = lv;
            r.n = std::max(r.n, lv + 1);
            int* slot = r.cls[lv];
            if (slot[0] < 0) slot[0] = order[i];
            else if (slot[1] < 0) slot[1] = order[i];
            else throw std::logic_error("three colour classes in one level");
        }
        return r;
    }();
    return L;
}

void color_faces(dmc_context* ctx) {
    const int64_t qcap = ctx->qcap;
    Sizes* S = dsizes(ctx);
    uint8_t* colors = static_cast<uint8_t*>(ctx->colors.p);
    // one histogram block per 4096 faces of capacity, at most kColorBlocks
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kColorBlocks, qcap / 4096));
    uint32_t* hist = ensure<uint32_t>(ctx, ctx->ccnt, (size_t)kColorBlocks * 24);
    int32_t* list = ensure<int32_t>(ctx, ctx->clist, qcap);
    int4* nbr = ensure<int4>(ctx, ctx->cnbr, qcap);
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    launch(ctx, "k_color_hist", k_color_hist, dim3(nb), dim3(kBlock), (const uint8_t*)colors, (const Sizes*)S, hist);
    launch(ctx, "k_color_offsets", k_color_offsets, dim3(1), dim3(1024), hist, nb, S->colors);
    launch(ctx, "k_color_scatter", k_color_scatter, dim3(nb), dim3(kBlock), (const uint8_t*)colors, (const Sizes*)S,
           (const uint32_t*)hist, twin, list, nbr);
    if (qcap <= kSmallColourFaces) {
        launch(ctx, "k_color_class", k_color_all, dim3(1), dim3(1024), (const int32_t*)list, (const int4*)nbr,
               (const uint32_t*)S->colors, colors);
        check_launch(ctx, "colour");
        return;
    }
    // a class holds at most a quarter of the faces (four parities of one
    // edge direction), a launch at most two classes
    const int g = grid_for(ctx, qcap / 2 + 1);
    const ColourLevels& L = colour_levels();
    for (int lv = 0; lv < L.n; lv++)
        launch(ctx, "k_color_class", k_color_class, dim3(g), dim3(kBlock), (const int32_t*)list, (const int4*)nbr,
               (const uint32_t*)S->colors, L.cls[lv][0], L.cls[lv][1], colors);
    check_launch(ctx, "colour");
}

// Per-volume first vertex / quadrilateral after an order-preserving
// compaction (newidx = exclusive scan of the keep flags over *nold items,
// *nnew kept).
void remap_bounds(dmc_context* ctx, Buf& bounds, const uint32_t* newidx, const uint64_t* nold, const uint64_t* nnew) {
    if (ctx->nvol == 1) return;   // one volume: its bounds are [0, count], read from the sizes
    launch(ctx, "k_vol_remap", k_vol_remap, dim3(grid_for(ctx, ctx->nvol + 1)), dim3(kBlock),
           static_cast<int64_t*>(bounds.p), ctx->nvol + 1, newidx, nold, nnew);
    check_launch(ctx, "k_vol_remap");
}

// Compact vertices by `removed` (S->nv of them, S->nn kept); vsecond (if
// non-null) receives new indices of kept vertices.  Positions and normals
// stay where they are; the slot map follows the vertices (gather_vertices
// lays them out at the end).
// The scan of the kept vertices into ctx->newV (S->nn of them), the job that
// compact_vertices(..., scanned = true) expects to have run.
ScanJob<ScanKeep> kept_vertices_job(dmc_context* ctx, const uint8_t* removed) {
    Sizes* S = dsizes(ctx);
    return scan_job(ScanKeep{removed}, ensure<uint32_t>(ctx, ctx->newV, ctx->vcap), dn(&S->nv, ctx->vcap), &S->nn);
}

uint32_t* compact_vertices(dmc_context* ctx, const uint8_t* removed, int32_t* vsecond, bool scanned = false) {
    const int64_t vcap = ctx->vcap;
    Sizes* S = dsizes(ctx);
    uint32_t* newidx = ensure<uint32_t>(ctx, ctx->newV, vcap);
    if (!scanned) scan_to(ctx, ScanKeep{removed}, newidx, dn(&S->nv, vcap), &S->nn);
    const uint32_t* slot = ctx->slots ? static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p) : nullptr;
    uint32_t* slot2 = ensure<uint32_t>(ctx, ctx->vslot[1 - ctx->scur], vcap);
    launch(ctx, "k_compact_slots", k_compact_slots, dim3(grid_for(ctx, vcap)), dim3(kBlock), removed, newidx,
           (const Sizes*)S, slot, slot2, vsecond);
    check_launch(ctx, "k_compact_slots");
    ctx->scur = 1 - ctx->scur;
    ctx->slots = true;
    remap_bounds(ctx, ctx->vbound, newidx, &S->nv, &S->nn);
    return newidx;
}

// Lay the live vertices' positions and normals out densely (after the last
// compaction).
void gather_vertices(dmc_context* ctx) {
    if (!ctx->slots) return;
    const int64_t vcap = ctx->vcap;
    const int src = ctx->pcur, dst = 1 - ctx->pcur;
    double* pos2 = ensure<double>(ctx, ctx->pos[dst], 3 * vcap);
    double* nrm2 = ensure<double>(ctx, ctx->nrm[dst], 3 * vcap);
    launch(ctx, "k_gather_vertices", k_gather_vertices, dim3(grid_for(ctx, vcap)), dim3(kBlock),
           static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p), (const Sizes*)dsizes(ctx),
           static_cast<const double*>(ctx->pos[src].p), static_cast<const double*>(ctx->nrm[src].p), pos2, nrm2);
    check_launch(ctx, "k_gather_vertices");
    ctx->pcur = dst;
    ctx->slots = false;
}

// The carry of a compaction, first half: the face map and the buffers of the
// carried build (the vertex map kernel then opens the carry's counters,
// clears its flags and counts -- carry_clears -- and carries the kept
// vertices' out-degree and he_v).  S->nn and S->qn must be on the device;
// `last`: the compaction before the final build (a batch counts per volume).
Carry carry_begin(dmc_context* ctx, const uint32_t* newq, bool last) {
    const int64_t vcap = ctx->vcap, qcap = ctx->qcap, hcap = 4 * qcap;
    Sizes* S = dsizes(ctx);
    Carry c{};
    c.twin = static_cast<const int32_t*>(ctx->twin.p);
    c.flag = static_cast<const uint8_t*>(ctx->flag.p);
    c.qmap = reinterpret_cast<const int32_t*>(newq);
    uint8_t* cflag = ensure<uint8_t>(ctx, ctx->cflag, vcap + 4);
    c.cflag = cflag;
    c.twin_n = ensure<int32_t>(ctx, ctx->twin2, hcap);
    c.flag_n = ensure<uint8_t>(ctx, ctx->flag2, hcap);
    c.deg_n = ensure<uint32_t>(ctx, ctx->deg2, vcap);
    c.hev_n = ensure<int32_t>(ctx, ctx->hev2, vcap);
    c.arr = ensure<uint32_t>(ctx, ctx->arr, vcap);
    c.aflag = ensure<uint8_t>(ctx, ctx->aflag, hcap);
    c.cnt = ensure<uint32_t>(ctx, ctx->cursor, vcap);
    const int64_t turns = qcap / kBlock + 1;
    c.bcount = ensure<uint32_t>(ctx, ctx->bcount, turns);
    c.S = S;
    c.qb = static_cast<const int64_t*>(ctx->qbound.p);
    c.nvol = ctx->nvol;
    Fills fl(ctx);
    if (last && ctx->nvol > 1) {
        c.vst = ensure<unsigned long long>(ctx, ctx->vhe, 3 * (size_t)ctx->nvol);
        fl.u32(c.vst, hn(6 * (int64_t)ctx->nvol), 0u);
    }
    fl.run();
    return c;
}

// What k_vmap* clears for the carry begun by carry_begin.
CarryClears carry_clears(dmc_context* ctx, const Carry& c) {
    return CarryClears{reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(c.cflag)), (ctx->vcap + 4) / 4, c.arr, c.cnt,
                       c.bcount, ctx->qcap / kBlock + 1};
}

// Second half, after the compaction kernel: the flagged halfedges paired
// through their rows, the 3-face records carried, the counts committed and
// the carried build made current.
void carry_end(dmc_context* ctx, const Carry& c, const int32_t* vmap) {
    const int64_t vcap = ctx->vcap, hcap = 4 * ctx->qcap;
    Sizes* S = dsizes(ctx);
    const DevN nn = dn(&S->nn, vcap);
    const int32_t* quads2 = static_cast<const int32_t*>(ctx->quads[1 - ctx->cur].p);
    uint32_t* aoff = ensure<uint32_t>(ctx, ctx->off2, vcap + 1);
    int2* aadj = ensure<int2>(ctx, ctx->adj2, hcap);
    BndRec* recs_n = ensure<BndRec>(ctx, ctx->bndrec2, ctx->reccap);
    const int gh = grid_for(ctx, hcap);
    const int64_t turns = ctx->qcap / kBlock + 1;
    uint32_t* boff = ensure<uint32_t>(ctx, ctx->boff, turns);
    int32_t* alist = ensure<int32_t>(ctx, ctx->alist, hcap);
    scan_pair(ctx, scan_job(ScanCounts{c.bcount}, boff, hn(turns), &S->naff),
              scan_job(ScanCounts{c.cnt}, aoff, nn, &S->tot, nullptr, aoff));
    launch(ctx, "k_aff_scatter", k_aff_scatter, dim3(grid_for(ctx, ctx->qcap)), dim3(kBlock), (const Sizes*)S,
           c.qmap, (const uint8_t*)c.aflag, (const uint32_t*)c.bcount, (const uint32_t*)boff, quads2,
           (const uint32_t*)aoff, aadj, alist);
    launch(ctx, "k_aff_twins", k_aff_twins, dim3(gh), dim3(kBlock), (const int32_t*)alist, (const Sizes*)S, quads2,
           (const uint32_t*)aoff, (const int2*)aadj, c, recs_n, (uint32_t)ctx->reccap);
    launch(ctx, "k_carry_records", k_carry_records, dim3(grid_for(ctx, ctx->reccap)), dim3(kBlock),
           static_cast<const BndRec*>(ctx->bndrec.p), (const Sizes*)S, vmap, c.cflag, c.qmap, recs_n, &S->nrec,
           (uint32_t)ctx->reccap, &S->ovf);
    check_launch(ctx, "carry");
    // the counts are committed by the next build's gate
    ctx->commit_pending = true;
    swap_pairs(ctx);
    ctx->cur = 1 - ctx->cur;
    ctx->ready = 2;
}

void simplify3x3y(dmc_context* ctx, bool colored, bool last) {
    const int64_t vcap = ctx->vcap, qcap = ctx->qcap;
    uint8_t* p = ensure<uint8_t>(ctx, ctx->p, qcap);
    uint8_t* em = ensure<uint8_t>(ctx, ctx->em, qcap);
    uint8_t* vfirst = ensure<uint8_t>(ctx, ctx->vfirst, vcap);
    int32_t* vsecond = ensure<int32_t>(ctx, ctx->vsecond, vcap);
    int32_t* lastT = ensure<int32_t>(ctx, ctx->lastT, vcap);
    int32_t* lastR = ensure<int32_t>(ctx, ctx->lastR, vcap);
    // vfirst = 0, vsecond = lastT = lastR = -1 over the live vertices
    StageClears cl;
    cl.v8 = vfirst;
    cl.v32[0] = reinterpret_cast<uint32_t*>(vsecond);
    cl.v32[1] = reinterpret_cast<uint32_t*>(lastT);
    cl.v32[2] = reinterpret_cast<uint32_t*>(lastR);
    cl.vcap = vcap;
    cl.qcap = qcap;
    build_halfedges(ctx, false, cl);
    if (colored) color_faces(ctx);
    const int32_t* val = static_cast<const int32_t*>(ctx->val.p);
    Sizes* S = dsizes(ctx);
    const DevN nq = dn(&S->nq, qcap);
    const int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    const uint8_t* flag = static_cast<uint8_t*>(ctx->flag.p);
    const uint8_t* colors = colored ? static_cast<uint8_t*>(ctx->colors.p) : nullptr;
    uint32_t* newQ = ensure<uint32_t>(ctx, ctx->newQ, qcap);
    const int gq = grid_for(ctx, qcap);
    launch(ctx, "k_mark3x3y", k_mark3x3y, dim3(gq), dim3(kBlock), quads, (const Sizes*)S, val, p);
    launch(ctx, "k_merge_claim", k_merge_claim, dim3(gq), dim3(kBlock), quads, (const Sizes*)S, val,
           (const uint8_t*)p, colors, twin, flag, lastT, lastR, em, vfirst);
    check_launch(ctx, "merge3x3y");
    // the quadrilateral compaction's scan (which also lists the merging
    // faces) beside the vertices' (vfirst is final after the claims)
    scan_pair(ctx, scan_job(ScanKeepMap{{em}}, newQ, nq, &S->qn), kept_vertices_job(ctx, vfirst));
    int32_t* mlist = ensure<int32_t>(ctx, ctx->mlist, qcap);
    launch(ctx, "k_list_flagged", k_list_flagged, dim3(gq), dim3(kBlock), (const uint8_t*)em, (const int32_t*)newQ,
           nq, mlist);
    launch(ctx, "k_merge_apply", k_merge_apply, dim3(gq), dim3(kBlock), quads, (const int32_t*)mlist,
           (const Sizes*)S, (const uint8_t*)em, val, (const int32_t*)lastT, (const int32_t*)lastR,
           static_cast<double*>(ctx->pos[ctx->pcur].p), static_cast<double*>(ctx->nrm[ctx->pcur].p),
           ctx->slots ? static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p) : nullptr, vsecond);
    check_launch(ctx, "k_merge_apply");
    compact_vertices(ctx, vfirst, vsecond, true);
    int32_t* q2 = ensure<int32_t>(ctx, ctx->quads[1 - ctx->cur], 4 * qcap);
    int32_t* vmap = ensure<int32_t>(ctx, ctx->vmap, vcap);
    remap_bounds(ctx, ctx->qbound, newQ, &S->nq, &S->qn);
    Carry c = carry_begin(ctx, newQ, last);
    launch(ctx, "k_vmap3x3y", k_vmap3x3y, dim3(grid_for(ctx, vcap)), dim3(kBlock), (const uint8_t*)vfirst,
           (const int32_t*)vsecond, S, vmap, static_cast<const uint32_t*>(ctx->deg.p),
           static_cast<const int32_t*>(ctx->hev.p), c.qmap, c.deg_n, c.hev_n, carry_clears(ctx, c));
    launch(ctx, "k_changed3x3y", k_changed3x3y, dim3(gq), dim3(kBlock), quads, (const int32_t*)mlist,
           (const Sizes*)S, (const uint8_t*)em, (const int32_t*)vmap, static_cast<const int32_t*>(ctx->hev.p),
           const_cast<uint8_t*>(c.cflag), c.deg_n, c.hev_n);
    launch(ctx, "k_compact_quads3x3y", k_compact_quads3x3y, dim3(gq), dim3(kBlock), quads, (const Sizes*)S,
           (const int32_t*)vmap, q2, c);
    check_launch(ctx, "k_compact_quads3x3y");
    carry_end(ctx, c, vmap);
}

void simplify3333(dmc_context* ctx, bool last) {
    const int64_t vcap = ctx->vcap, qcap = ctx->qcap;
    uint8_t* vflag = ensure<uint8_t>(ctx, ctx->vfirst, vcap);
    int32_t* vtwin = ensure<int32_t>(ctx, ctx->vsecond, vcap);
    int32_t* lastF = ensure<int32_t>(ctx, ctx->lastT, vcap);
    uint8_t* p3333 = ensure<uint8_t>(ctx, ctx->p, qcap);
    uint8_t* rflag = ensure<uint8_t>(ctx, ctx->em, qcap);
    // vflag = 0, vtwin = lastF = -1 over the live vertices, the face flags 0
    StageClears cl;
    cl.v8 = vflag;
    cl.v32[0] = reinterpret_cast<uint32_t*>(vtwin);
    cl.v32[1] = reinterpret_cast<uint32_t*>(lastF);
    cl.f8[0] = p3333;
    cl.f8[1] = rflag;
    cl.vcap = vcap;
    cl.qcap = qcap;
    build_halfedges(ctx, false, cl);
    const int32_t* val = static_cast<const int32_t*>(ctx->val.p);
    Sizes* S = dsizes(ctx);
    const DevN nv = dn(&S->nv, vcap), nq = dn(&S->nq, qcap);
    const int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    const uint8_t* flag = static_cast<uint8_t*>(ctx->flag.p);
    const int gq = grid_for(ctx, qcap);
    launch(ctx, "k_mark3333", k_mark3333, dim3(gq), dim3(kBlock), quads, (const Sizes*)S, twin, flag, val, lastF,
           vflag, p3333, rflag);
    launch(ctx, "k_twin3333", k_twin3333, dim3(gq), dim3(kBlock), quads, (const Sizes*)S, twin, flag, val,
           (const uint8_t*)p3333, (const int32_t*)lastF, vtwin);
    check_launch(ctx, "mark3333");
    uint32_t* newQ = ensure<uint32_t>(ctx, ctx->newQ, qcap);
    scan_pair(ctx, scan_job(ScanKeepMap{{rflag}}, newQ, nq, &S->qn), kept_vertices_job(ctx, vflag));
    uint32_t* newv = compact_vertices(ctx, vflag, nullptr, true);
    int32_t* q2 = ensure<int32_t>(ctx, ctx->quads[1 - ctx->cur], 4 * qcap);
    int32_t* vmap = ensure<int32_t>(ctx, ctx->vmap, vcap);
    remap_bounds(ctx, ctx->qbound, newQ, &S->nq, &S->qn);
    // the dropped quadrilaterals, whose corners are the changed vertices
    int32_t* dlist = ensure<int32_t>(ctx, ctx->mlist, qcap);
    launch(ctx, "k_list_flagged", k_list_flagged, dim3(gq), dim3(kBlock), (const uint8_t*)rflag,
           (const int32_t*)newQ, nq, dlist);
    Carry c = carry_begin(ctx, newQ, last);
    launch(ctx, "k_vmap3333", k_vmap3333, dim3(grid_for(ctx, vcap)), dim3(kBlock), (const uint8_t*)vflag,
           (const uint32_t*)newv, S, vmap, static_cast<const uint32_t*>(ctx->deg.p),
           static_cast<const int32_t*>(ctx->hev.p), c.qmap, c.deg_n, c.hev_n, carry_clears(ctx, c));
    launch(ctx, "k_changed3333", k_changed3333, dim3(gq), dim3(kBlock), quads, (const int32_t*)dlist,
           (const Sizes*)S, (const int32_t*)vmap, const_cast<uint8_t*>(c.cflag), c.deg_n, c.hev_n);
    launch(ctx, "k_compact_quads3333", k_compact_quads3333, dim3(gq), dim3(kBlock), quads, (const Sizes*)S,
           (const uint8_t*)p3333, (const int32_t*)vtwin, (const int32_t*)vmap, q2, c);
    check_launch(ctx, "k_compact_quads3333");
    carry_end(ctx, c, vmap);
}

void clear_timing(dmc_context* ctx) {
    ctx->marks.clear();
    ctx->kmarks.clear();
    ctx->evused = 0;
}

// Stage times (consecutive marks), then kernel times summed per name in
// order of first launch.
void finish_timing(dmc_context* ctx) {
    ctx->stage_ms.clear();
    if (!ctx->timing || (ctx->marks.empty() && ctx->kmarks.empty())) {
        clear_timing(ctx);
        return;
    }
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i + 1 < ctx->marks.size(); i++) {
        float ms = 0;
        HIPC(ctx, hipEventElapsedTime(&ms, ctx->marks[i].ev, ctx->marks[i + 1].ev));
        ctx->stage_ms.emplace_back(ctx->marks[i].name, ms);
    }
    const size_t first_kernel = ctx->stage_ms.size();
    for (const auto& k : ctx->kmarks) {
        float ms = 0;
        HIPC(ctx, hipEventElapsedTime(&ms, k.a, k.b));
        size_t i = first_kernel;
        while (i < ctx->stage_ms.size() && strcmp(ctx->stage_ms[i].first, k.name) != 0) i++;
        if (i == ctx->stage_ms.size()) ctx->stage_ms.emplace_back(k.name, 0.0f);
        ctx->stage_ms[i].second += ms;
    }
    clear_timing(ctx);
}

dmc_status guard(dmc_context* ctx, void (*fn)(dmc_context*, void*), void* arg) {
    if (!ctx) return DMC_ERR_INVALID_ARG;
    ctx->err.clear();
    try {
        HIPC(ctx, hipSetDevice(ctx->device));
        fn(ctx, arg);
        return DMC_OK;
    } catch (const Fail& f) {
        clear_timing(ctx);
        return f.s;
    }
}

void check_grid(dmc_context* ctx, const dmc_grid& g) {
    if (g.nx < 2 || g.ny < 2 || g.nz < 2) fail(ctx, DMC_ERR_INVALID_ARG, "grid needs at least 2 points per axis");
    if (g.nx > (1 << 24) || g.ny > (1 << 24) || g.nz > (1 << 24))
        fail(ctx, DMC_ERR_INVALID_ARG, "an axis above 16777216 points");
    const uint64_t ncells = (uint64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    if (ncells > 0xFFFFFFFFull) fail(ctx, DMC_ERR_CAPACITY, "more than 2^32 cells");
}

// Topologically correct MC of one volume (mc/MarchingCubes.cpp t_mc): the
// sweep's active cells, their vertex and triangle counts, scans over the
// active cells (whose order is the reference's cell order), then vertices
// and triangles.
template <class T>
void run_tmc(dmc_context* ctx, const std::vector<HostVol>& hv) {
    const SweepPlan P = plan_sweep<T>(ctx, hv);
    upload_vtab(ctx, P);
    const Swept sw = sweep_cells<T>(ctx, P);
    const uint64_t A = read_u64(ctx, &dsizes(ctx)->A);
    if (A > 0xFFFFFFFFull) fail(ctx, DMC_ERR_CAPACITY, "too many active cells (%llu)", (unsigned long long)A);
    const int64_t R = sw.R;
    uint32_t* acell = ensure<uint32_t>(ctx, ctx->acell, A);
    uint32_t* acode = ensure<uint32_t>(ctx, ctx->acode, A);
    uint32_t* cnt_v = ensure<uint32_t>(ctx, ctx->tcntv, A);
    uint32_t* cnt_t = ensure<uint32_t>(ctx, ctx->tcntt, A);
    uint32_t* vofs = ensure<uint32_t>(ctx, ctx->tvofs, A);
    uint32_t* tofs = ensure<uint32_t>(ctx, ctx->ttofs, A);
    unsigned long long* st = ensure<unsigned long long>(ctx, ctx->tstats, 3);
    uint32_t* amb = ensure<uint32_t>(ctx, ctx->tamb, A + 1);
    fill_u32(ctx, st, hn(6), 0u);
    fill_u32(ctx, amb, hn(1), 0u);
    launch(ctx, "k_tmc_cells", k_tmc_cells, dim3(grid_for(ctx, R, kBlock / kWave)), dim3(kBlock), sw.vt, R, sw.lists,
           sw.segCount, sw.segBase, acell, acode, cnt_v, cnt_t, amb, st);
    check_launch(ctx, "k_tmc_cells");
    if (A > 0) {
        launch(ctx, "k_tmc_amb_count", k_tmc_amb_count<T>, dim3(grid_for(ctx, A, kBlock, 1024)), dim3(kBlock), sw.vt,
               amb, acell, acode, cnt_v, cnt_t, st);
        check_launch(ctx, "k_tmc_amb_count");
    }
    uint64_t* tot = ensure<uint64_t>(ctx, ctx->scanTotal, 2);
    scan_pair(ctx, scan_job(ScanCounts{cnt_v}, vofs, hn((int64_t)A), tot),
              scan_job(ScanCounts{cnt_t}, tofs, hn((int64_t)A), tot + 1));
    uint64_t VT[2] = {0, 0};
    {
        HostReads rd(ctx);
        rd.add(VT, tot, sizeof(VT));
        rd.wait();
    }
    const uint64_t V = VT[0], Tn = VT[1];
    if (V > 0x7FFFFFFFull || Tn > 0x7FFFFFFFull)
        fail(ctx, DMC_ERR_CAPACITY, "mesh too large (V=%llu, T=%llu)", (unsigned long long)V,
             (unsigned long long)Tn);
    double* pos = ensure<double>(ctx, ctx->tpos, 3 * V);
    double* nrm = ensure<double>(ctx, ctx->tnrm, 3 * V);
    int32_t* tris = ensure<int32_t>(ctx, ctx->ttris, 3 * Tn);
    uint2* avb = ensure<uint2>(ctx, ctx->tvb, A);
    mark(ctx, "emit");
    if (A > 0) {
        launch(ctx, "k_tmc_vbase", k_tmc_vbase, dim3(grid_for(ctx, A)), dim3(kBlock), sw.vt, acell, acode, vofs,
               (int64_t)A, avb);
        check_launch(ctx, "k_tmc_vbase");
        launch(ctx, "k_tmc_emit", k_tmc_emit<T>, dim3(grid_for(ctx, A)), dim3(kBlock), sw.vt, acell, acode, avb, tofs,
               sw.segMask, sw.segBase, (int64_t)A, pos, nrm, tris, st + 2);
        check_launch(ctx, "k_tmc_emit");
        launch(ctx, "k_tmc_emit_amb", k_tmc_emit_amb<T>, dim3(grid_for(ctx, A, kBlock, 1024)), dim3(kBlock), sw.vt,
               amb, acell, acode, avb, tofs, sw.segMask, sw.segBase, pos, nrm, tris, st + 2);
        check_launch(ctx, "k_tmc_emit_amb");
    }
    mark(ctx, "end");
    unsigned long long hs[3];
    {
        HostReads rd(ctx);
        rd.add(hs, st, sizeof(hs));
        rd.wait();
    }
    finish_timing(ctx);
    if (hs[2] != 0) fail(ctx, DMC_ERR_STATE, "%llu cut edges without an active owner cell", hs[2]);
    tmc_mesh& m = ctx->tmesh;
    m = tmc_mesh{};
    m.nv = (int64_t)V;
    m.nt = (int64_t)Tn;
    m.d_positions = pos;
    m.d_normals = nrm;
    m.d_tris = tris;
    m.active_cells = (int64_t)A;
    m.ambiguous_cells = (int64_t)hs[0];
    m.face_failures = (int64_t)hs[1];
    ctx->have_tmesh = true;
}

// A pass (everything after the upload, for one volume or a batch) is queued
// without reading anything back (enqueue_pass); then the sizes block and
// the per-volume bookkeeping come back in one transfer (finish_pass).
// The pass's launches, in order; nothing is read back.
void enqueue_pass(dmc_context* ctx, const SweepPlan& P, int field_type, uint32_t flags) {
    const int n = ctx->nvol;
    ctx->commit_pending = false;
    if (field_type == DMC_FIELD_F32) run_cells<float>(ctx, P);
    else run_cells<double>(ctx, P);
    const bool iso = (flags & DMC_SIMPLIFY_3X3Y_ISOLATED) != 0, p3333 = (flags & DMC_SIMPLIFY_3333) != 0;
    if (flags & DMC_SIMPLIFY_3X3Y_COLOR) {
        mark(ctx, "p3x3y_color");
        simplify3x3y(ctx, true, !iso && !p3333);
    }
    if (iso) {
        mark(ctx, "p3x3y_isolated");
        simplify3x3y(ctx, false, !p3333);
    }
    if (p3333) {
        mark(ctx, "p3333");
        simplify3333(ctx, true);
    }
    mark(ctx, "halfedges");
    build_halfedges(ctx, true);
    gather_vertices(ctx);
    Sizes* S = dsizes(ctx);
    int32_t* tris = nullptr;
    if (flags & DMC_TRIANGLES) {
        mark(ctx, "triangles");
        tris = ensure<int32_t>(ctx, ctx->tris, 6 * ctx->qcap);
        launch(ctx, "k_triangles", k_triangles, dim3(grid_for(ctx, ctx->qcap)), dim3(kBlock),
               static_cast<int32_t*>(ctx->quads[ctx->cur].p), (const Sizes*)S,
               static_cast<double*>(ctx->pos[ctx->pcur].p), tris);
        check_launch(ctx, "k_triangles");
    }
    int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    int32_t* hev = static_cast<int32_t*>(ctx->hev.p);
    if (n > 1) {
        // each volume's ids start at 0
        launch(ctx, "k_rebase", k_rebase, dim3(grid_for(ctx, std::max(ctx->qcap, ctx->vcap))), dim3(kBlock),
               static_cast<const int64_t*>(ctx->vbound.p), static_cast<const int64_t*>(ctx->qbound.p), n, quads, twin,
               tris, (const Sizes*)S, hev);
        check_launch(ctx, "k_rebase");
    }
    mark(ctx, "end");
}

// The pass's sizes and per-volume bookkeeping back to the host (one
// transfer), and the meshes.  Returns false when a capacity was exceeded.
bool finish_pass(dmc_context* ctx, const std::vector<HostVol>& hv, uint32_t flags) {
    const int n = ctx->nvol;
    Sizes* S = dsizes(ctx);
    int32_t* tris = (flags & DMC_TRIANGLES) ? static_cast<int32_t*>(ctx->tris.p) : nullptr;
    int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    int32_t* hev = static_cast<int32_t*>(ctx->hev.p);
    // sizes and per-volume bookkeeping back to the host
    std::vector<int64_t> vb(2 * (n + 1)), qb(2 * (n + 1)), ab(n + 1), vhe(n > 1 ? 3 * n : 0);
    std::vector<VolStat> vs(n);
    Sizes& hs = ctx->hsz;
    {
        HostReads r(ctx);
        r.add(&hs, S, sizeof(Sizes));
        r.add(vb.data(), ctx->vbound.p, vb.size() * sizeof(int64_t));
        r.add(qb.data(), ctx->qbound.p, qb.size() * sizeof(int64_t));
        r.add(ab.data(), vol_ptr(ctx, 2), ab.size() * sizeof(int64_t));
        r.add(vs.data(), ctx->vstat.p, vs.size() * sizeof(VolStat));
        if (n > 1) r.add(vhe.data(), ctx->vhe.p, vhe.size() * sizeof(unsigned long long));
        r.wait();
    }
    finish_timing(ctx);
    if (hs.A > 0xFFFFFFFFull) fail(ctx, DMC_ERR_CAPACITY, "too many active cells (%llu)", (unsigned long long)hs.A);
    // the build's flagged-halfedge counts from their slots
    hs.he.nm4 = hs.he.nm3 = hs.he.errors = 0;
    for (const auto& part : hs.hep) {
        hs.he.nm4 += part[0];
        hs.he.nm3 += part[1];
        hs.he.errors += part[2];
    }
    if (hs.V > 0x7FFFFFFFull || hs.Q > 0x1FFFFFFFull)
        fail(ctx, DMC_ERR_CAPACITY, "mesh too large (V=%llu, Q=%llu)", (unsigned long long)hs.V,
             (unsigned long long)hs.Q);
    // what this context has seen, for the next pass's capacities: each
    // volume's cell-stage counts (exact whatever the capacities)
    for (int v = 0; v < n; v++) {
        const dmc_grid& g = hv[v].g;
        const double c = (double)(g.nx - 1) * (double)(g.ny - 1) * (double)(g.nz - 1);
        if (c <= 0) continue;
        const double cnt[3] = {n == 1 ? (double)hs.A : (double)(ab[v + 1] - ab[v]),
                               n == 1 ? (double)hs.V : (double)(vb[n + 1 + v + 1] - vb[n + 1 + v]),
                               n == 1 ? (double)hs.Q : (double)(qb[n + 1 + v + 1] - qb[n + 1 + v])};
        for (int k = 0; k < 3; k++) {
            ctx->rcell[k] = std::max(ctx->rcell[k], cnt[k] / c);
            ctx->rarea[k] = std::max(ctx->rarea[k], cnt[k] / area_of(c));
        }
    }
    if (hs.need_generic) {
        // a build needed generic rows: the same pass again with them
        ctx->force_generic = true;
        if (!hs.ovf) return false;
    }
    if (hs.ovf) {
        // exceeded a capacity: the same pass again with room for the exact
        // sizes (the counts above are exact whatever the capacities)
        ctx->acap_min = std::max<int64_t>(ctx->acap_min, (int64_t)hs.A);
        ctx->vcap_min = std::max<int64_t>(ctx->vcap_min, (int64_t)hs.V);
        ctx->qcap_min = std::max<int64_t>(ctx->qcap_min, (int64_t)hs.Q);
        if (hs.ovf & 8u) ctx->reccap = std::max<int64_t>(ctx->reccap, (int64_t)std::max(hs.nrec, hs.recmax));
        return false;
    }
    ctx->rows_live = hs.pair_rows != 0u;
    if (hs.rows != 4 * hs.Q)
        fail(ctx, DMC_ERR_STATE, "halfedge rows %llu != 4 Q %llu", (unsigned long long)hs.rows,
             (unsigned long long)(4 * hs.Q));
    // getBndVertices: the 3-face records, in key order
    ctx->bnd.clear();
    if (hs.nrec > 0) {
        std::vector<BndRec> h(hs.nrec);
        HostReads r(ctx);
        r.add(h.data(), ctx->bndrec.p, h.size() * sizeof(BndRec));
        r.wait();
        std::sort(h.begin(), h.end(), [](const BndRec& a, const BndRec& b) { return a.key < b.key; });
        for (const BndRec& x : h) {
            ctx->bnd.push_back(x.a);
            ctx->bnd.push_back(x.b);
        }
    }
    const int64_t nv = (int64_t)hs.nv, nq = (int64_t)hs.nq;
    dmc_mesh& M = ctx->mesh;
    M.nv = nv;
    M.nq = nq;
    M.nt = (flags & DMC_TRIANGLES) ? 2 * nq : 0;
    M.d_tris = tris;
    // the build counts flagged halfedges: 4 per non-manifold edge, 3 per
    // 3-face edge
    M.nonmanifold_edges = (int64_t)(hs.he.nm4 / 4 + hs.he.nm3 / 3);
    M.edge_errors = (int64_t)hs.he.errors;
    M.invalid_refs = (int64_t)hs.he.bad + (int64_t)hs.unresolved;
    M.valence_zero = hs.zero;
    M.nbnd = (int64_t)ctx->bnd.size();
    M.active_cells = (int64_t)hs.A;
    M.nv_cells = (int64_t)hs.V;
    M.nq_cells = (int64_t)hs.Q;
    M.d_positions = static_cast<double*>(ctx->pos[ctx->pcur].p);
    M.d_normals = static_cast<double*>(ctx->nrm[ctx->pcur].p);
    M.d_quads = quads;
    M.d_twins = twin;
    M.d_he_flags = static_cast<uint8_t*>(ctx->flag.p);
    M.d_he_v = hev;
    M.host_syncs = ctx->syncs - ctx->syncs_at_start;
    M.pass_mode = ctx->pass_mode;
    uint64_t f1 = 0, f2 = 0, nc = 0;
    for (const VolStat& x : vs) {
        f1 += x.fail1;
        f2 += x.fail2;
        nc += x.contours;
    }
    if (n == 1)
        for (unsigned long long c : hs.cpart) nc += c;
    M.failed_projections_l1 = (uint32_t)f1;
    M.failed_projections_l2 = (uint32_t)f2;
    M.nv_emitted = (int64_t)nc;
    // one view per volume
    ctx->vmesh.assign(n, dmc_mesh{});
    ctx->vbnd.assign(n, std::vector<int32_t>());
    if (n == 1) {
        ctx->vmesh[0] = M;
        ctx->vbnd[0] = ctx->bnd;
    } else {
        for (size_t x = 0; x + 1 < ctx->bnd.size(); x += 2) {
            const int64_t a = ctx->bnd[x];
            const int v = (int)(std::upper_bound(vb.begin(), vb.begin() + n, a) - vb.begin()) - 1;
            ctx->vbnd[v].push_back((int32_t)(ctx->bnd[x] - vb[v]));
            ctx->vbnd[v].push_back((int32_t)(ctx->bnd[x + 1] - vb[v]));
        }
        for (int v = 0; v < n; v++) {
            dmc_mesh& m = ctx->vmesh[v];
            m.nv = vb[v + 1] - vb[v];
            m.nq = qb[v + 1] - qb[v];
            m.nt = (flags & DMC_TRIANGLES) ? 2 * m.nq : 0;
            m.d_positions = M.d_positions + 3 * vb[v];
            m.d_normals = M.d_normals + 3 * vb[v];
            m.d_quads = quads + 4 * qb[v];
            m.d_tris = tris ? tris + 6 * qb[v] : nullptr;
            m.d_twins = twin + 4 * qb[v];
            m.d_he_flags = M.d_he_flags + 4 * qb[v];
            m.d_he_v = hev + vb[v];
            m.nonmanifold_edges = vhe[3 * v] / 4 + vhe[3 * v + 1] / 3;
            m.nbnd = (int64_t)ctx->vbnd[v].size();
            m.edge_errors = vhe[3 * v + 2];
            m.invalid_refs = M.invalid_refs;   // batch-wide
            m.valence_zero = M.valence_zero;   // batch-wide
            m.failed_projections_l1 = vs[v].fail1;
            m.failed_projections_l2 = vs[v].fail2;
            m.active_cells = ab[v + 1] - ab[v];
            m.nv_emitted = (int64_t)vs[v].contours;
            m.nv_cells = vb[n + 1 + v + 1] - vb[n + 1 + v];
            m.nq_cells = qb[n + 1 + v + 1] - qb[n + 1 + v];
            m.host_syncs = M.host_syncs;   // batch-wide
            m.pass_mode = M.pass_mode;
        }
    }
    ctx->have_mesh = true;
    return true;
}

// ---------------------------------------------------------------------------
// Captured passes.  A pass's launches depend only on its shape: the volumes'
// dimensions, the element type and alignment, the flags, the launch shape,
// the capacities and the buffers (every buffer allocation bumps buf_epoch);
// the fields, iso-values and geometry are read from the volume table in
// device memory.  The third time a shape comes, its pass is captured into a
// HIP graph; later passes of that shape replay it.  Host-side state the
// launches leave (which double buffers hold the mesh) is recorded with the
// graph and restored after a replay.
// ---------------------------------------------------------------------------
std::string pass_key(dmc_context* ctx, const std::vector<HostVol>& hv, const SweepPlan& P, int field_type,
                     uint32_t flags) {
    std::string k;
    auto put = [&](int64_t x) { k.append(reinterpret_cast<const char*>(&x), sizeof(x)); };
    put(field_type);
    put(flags & (DMC_SIMPLIFY_ALL | DMC_TRIANGLES));
    put(P.vec);
    put(ctx->grid_cap);
    put(ctx->sweep_waves);
    put(ctx->acap);
    put(ctx->vcap);
    put(ctx->qcap);
    put(ctx->reccap);
    put(ctx->force_generic);
    put((int64_t)ctx->buf_epoch);
    put((int64_t)hv.size());
    for (const HostVol& h : hv) put(((int64_t)h.g.nx << 42) | ((int64_t)h.g.ny << 21) | h.g.nz);
    return k;
}

void save_post(dmc_context* ctx, dmc_context::PassGraph& g) {
    g.cur = ctx->cur;
    g.pcur = ctx->pcur;
    g.scur = ctx->scur;
    g.slots = ctx->slots;
    g.ready = ctx->ready;
    g.swapped = pair_order(ctx);
}

void load_post(dmc_context* ctx, const dmc_context::PassGraph& g) {
    ctx->cur = g.cur;
    ctx->pcur = g.pcur;
    ctx->scur = g.scur;
    ctx->slots = g.slots;
    ctx->ready = g.ready;
    set_pair_order(ctx, g.swapped);
}

void drop_graphs(dmc_context* ctx) {
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
}

// One pass: eager, captured and launched, or replayed.
void run_pass(dmc_context* ctx, const SweepPlan& P, int field_type, uint32_t flags) {
    ctx->pass_mode = 0;
    const bool graphs = ctx->use_graphs && !ctx->timing && !ctx->arena;
    std::string key;
    if (graphs) {
        // hv is only needed for the key: rebuild it from the volume table
        std::vector<HostVol> hv(ctx->nvol);
        const VolDesc* d = reinterpret_cast<const VolDesc*>(ctx->vtab_host.data());
        for (int v = 0; v < ctx->nvol; v++) hv[v].g = dmc_grid{d[v].d.nx, d[v].d.ny, d[v].d.nz, 0, 0, 0, 0, 0, 0};
        key = pass_key(ctx, hv, P, field_type, flags);
        for (auto& g : ctx->graphs) {
            if (g.key != key) continue;
            upload_vtab(ctx, P);
            if (ctx->off.p > ctx->off2.p) std::swap(ctx->off, ctx->off2);
            if (ctx->adj.p > ctx->adj2.p) std::swap(ctx->adj, ctx->adj2);
            HIPC(ctx, hipGraphLaunch(g.exec, ctx->stream));
            load_post(ctx, g);
            g.use = ++ctx->graph_clock;
            ctx->pass_mode = 2;
            return;
        }
    }
    upload_vtab(ctx, P);
    if (graphs && ctx->seen.size() >= kMaxSeen) ctx->seen.clear();
    if (!graphs || ctx->seen[key]++ < kCaptureAfter) {
        enqueue_pass(ctx, P, field_type, flags);
        return;
    }
    // a shape that keeps coming: capture its pass, then launch the graph
    const uint64_t epoch = ctx->buf_epoch;
    HIPC(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    hipGraph_t graph = nullptr;
    try {
        enqueue_pass(ctx, P, field_type, flags);
    } catch (...) {
        ctx->capturing = false;
        (void)hipStreamEndCapture(ctx->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    ctx->capturing = false;
    HIPC(ctx, hipStreamEndCapture(ctx->stream, &graph));
    dmc_context::PassGraph g{};
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) fail(ctx, DMC_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    if (ctx->buf_epoch != epoch) {
        (void)hipGraphExecDestroy(g.exec);
        fail(ctx, DMC_ERR_STATE, "buffers changed while a pass was captured");
    }
    g.key = key;
    save_post(ctx, g);
    // the capture recorded the launches without running them: launch now,
    // from the pass's starting state
    HIPC(ctx, hipGraphLaunch(g.exec, ctx->stream));
    g.use = ++ctx->graph_clock;
    if (ctx->graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                    [](const auto& x, const auto& y) { return x.use < y.use; });
        (void)hipGraphExecDestroy(lru->exec);
        ctx->graphs.erase(lru);
    }
    ctx->graphs.push_back(g);
    ctx->pass_mode = 1;
}

struct ExtractArgs {
    const dmc_volume* vols;
    int32_t nvol;
    int field_type;
    uint32_t flags;
    dmc_mesh* out;
    tmc_mesh* tout = nullptr;   // set: the topologically correct MC of vols[0]
};

void do_extract(dmc_context* ctx, void* a_) {
    ExtractArgs& a = *static_cast<ExtractArgs*>(a_);
    if (a.nvol < 1) fail(ctx, DMC_ERR_INVALID_ARG, "no volumes");
    if (a.field_type != DMC_FIELD_F32 && a.field_type != DMC_FIELD_F64)
        fail(ctx, DMC_ERR_INVALID_ARG, "unknown field type %d", a.field_type);
    const bool supplied = (a.flags & DMC_GRADIENTS_SUPPLIED) != 0;
    const size_t esz = a.field_type == DMC_FIELD_F32 ? 4 : 8;
    // layout of host fields packed into one device buffer (256-B aligned)
    std::vector<size_t> foff(a.nvol), goff(a.nvol);
    size_t fbytes = 0, gbytes = 0;
    for (int v = 0; v < a.nvol; v++) {
        const dmc_volume& x = a.vols[v];
        check_grid(ctx, x.grid);
        if (!x.field) fail(ctx, DMC_ERR_INVALID_ARG, "field of volume %d is NULL", v);
        if (supplied && !x.gradients) fail(ctx, DMC_ERR_INVALID_ARG, "gradients of volume %d are NULL", v);
        const size_t npts = (size_t)x.grid.nx * x.grid.ny * x.grid.nz;
        foff[v] = fbytes;
        fbytes += (npts * esz + 255) & ~(size_t)255;
        goff[v] = gbytes;
        if (supplied) gbytes += (3 * npts * sizeof(double) + 255) & ~(size_t)255;
    }
    ctx->timing = (a.flags & DMC_STAGE_TIMING) != 0;
    ctx->have_mesh = false;
    ctx->have_tmesh = false;
    ctx->mesh = dmc_mesh{};
    ctx->syncs_at_start = ctx->syncs;
    mark(ctx, "upload");
    std::vector<HostVol> hv(a.nvol);
    uint8_t* fdev = nullptr;
    uint8_t* gdev = nullptr;
    if (!(a.flags & DMC_FIELD_ON_DEVICE)) {
        fdev = ensure<uint8_t>(ctx, ctx->field, fbytes);
        if (supplied) gdev = ensure<uint8_t>(ctx, ctx->grads, gbytes);
    }
    for (int v = 0; v < a.nvol; v++) {
        const dmc_volume& x = a.vols[v];
        const size_t npts = (size_t)x.grid.nx * x.grid.ny * x.grid.nz;
        hv[v] = HostVol{x.grid, x.field, supplied ? x.gradients : nullptr, x.iso};
        if (fdev) {
            HIPC(ctx, hipMemcpyAsync(fdev + foff[v], x.field, npts * esz, hipMemcpyHostToDevice, ctx->stream));
            hv[v].field = fdev + foff[v];
            if (supplied) {
                HIPC(ctx, hipMemcpyAsync(gdev + goff[v], x.gradients, 3 * npts * sizeof(double),
                                         hipMemcpyHostToDevice, ctx->stream));
                hv[v].grads = reinterpret_cast<const double*>(gdev + goff[v]);
            }
        }
    }
    if (a.tout) {
        if (a.field_type == DMC_FIELD_F32) run_tmc<float>(ctx, hv);
        else run_tmc<double>(ctx, hv);
        *a.tout = ctx->tmesh;
        return;
    }
    // a pass that outgrew a capacity learned the exact sizes, and one that
    // needed generic halfedge rows asks for them: once more
    ctx->force_generic = ctx->always_generic;
    // exact sizes learned from an overflow hold for that extraction's rerun
    // only: later passes plan from the densities, which the overflow raised
    ctx->acap_min = ctx->vcap_min = ctx->qcap_min = 0;
    bool done = false;
    for (int attempt = 0; attempt < 3 && !done; attempt++) {
        if (attempt > 0 && ctx->arena) {
            // no-alloc call: the failed attempt's pieces are dead; carve anew
            for (Buf* b : ctx->bufs) *b = Buf{};
            ctx->arena_used = 0;
        }
        const SweepPlan P = a.field_type == DMC_FIELD_F32 ? plan_sweep<float>(ctx, hv) : plan_sweep<double>(ctx, hv);
        plan_caps(ctx, hv, P.cells);
        run_pass(ctx, P, a.field_type, a.flags);
        done = finish_pass(ctx, hv, a.flags);
    }
    if (!done) fail(ctx, DMC_ERR_STATE, "capacities still exceeded after resizing");
    if (a.out)
        for (int v = 0; v < a.nvol; v++) a.out[v] = ctx->vmesh[v];
}

}  // namespace

extern "C" {

const char* dmc_status_str(dmc_status s) {
    switch (s) {
    case DMC_OK: return "ok";
    case DMC_ERR_INVALID_ARG: return "invalid argument";
    case DMC_ERR_OUT_OF_MEMORY: return "out of device memory";
    case DMC_ERR_CAPACITY: return "capacity exceeded";
    case DMC_ERR_HIP: return "HIP error";
    case DMC_ERR_NO_DEVICE: return "no device";
    case DMC_ERR_BUDGET: return "memory budget exceeded";
    case DMC_ERR_STATE: return "invalid state";
    }
    return "unknown status";
}

const char* dmc_last_error(dmc_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

dmc_status dmc_create(dmc_context** out, int device) {
    if (!out) return DMC_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return DMC_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return DMC_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return DMC_ERR_HIP;
    dmc_context* ctx = new dmc_context();
    ctx->device = device;
    if (const char* cap = getenv("TMC_GRID_CAP")) {
        const long v = strtol(cap, nullptr, 10);
        if (v >= kGridCapMin && v <= kGridCapMax) ctx->grid_cap = (int)v;
    }
    if (const char* gr = getenv("TMC_GRAPHS")) ctx->use_graphs = strtol(gr, nullptr, 10) != 0;
    // every halfedge build through generic rows (the path a pass takes when
    // it is run again for them), for checking that path against the others
    if (const char* g = getenv("TMC_GENERIC_ROWS")) ctx->always_generic = strtol(g, nullptr, 10) != 0;
    if (const char* sp = getenv("TMC_SCAN_SPIN")) {
        const long v = strtol(sp, nullptr, 10);
        if (v >= 1 && v <= kScanSpin) ctx->scan_spin = (int)v;
    }
    if (const char* w = getenv("TMC_SWEEP_WAVES")) {
        const long v = strtol(w, nullptr, 10);
        if (v >= kSweepWavesMin && v <= kSweepWavesMax) ctx->sweep_waves = (int)v;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return DMC_ERR_HIP;
    }
    // pinned staging of count reads, and this device's no-alloc scratch
    // address (once per device)
    if (hipHostMalloc(&ctx->pin, 64 << 10) != hipSuccess || scratch_init(device) != hipSuccess) {
        (void)hipGetLastError();
        dmc_destroy(ctx);
        return DMC_ERR_HIP;
    }
    ctx->pin_cap = 64 << 10;
    *out = ctx;
    return DMC_OK;
}

dmc_status dmc_destroy(dmc_context* ctx) {
    if (!ctx) return DMC_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (Buf* b : ctx->bufs)
        if (b->p) (void)hipFree(b->p);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    for (hipEvent_t e : ctx->evpool) (void)hipEventDestroy(e);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DMC_OK;
}

dmc_status dmc_extract(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type, double iso,
                       uint32_t flags, const double* gradients, dmc_mesh* out) {
    if (!ctx || !grid) return DMC_ERR_INVALID_ARG;
    const dmc_volume vol{*grid, field, gradients, iso};
    ExtractArgs a{&vol, 1, field_type, flags, out};
    return guard(ctx, do_extract, &a);
}

dmc_status dmc_extract_batch(dmc_context* ctx, const dmc_volume* vols, int32_t nvol, int field_type, uint32_t flags,
                             dmc_mesh* out) {
    if (!ctx || !vols || nvol < 1) return DMC_ERR_INVALID_ARG;
    ExtractArgs a{vols, nvol, field_type, flags, out};
    return guard(ctx, do_extract, &a);
}

struct NoAllocArgs {
    ExtractArgs x;
    const dmc_out_buffers* ob;
};

// dmc_extract with every working buffer carved from this device's scratch
// array and the mesh copied into the caller's buffers.  The context's own
// buffers are set aside for the call and put back after it, so its pools
// are neither used nor grown.
static void do_extract_noalloc(dmc_context* ctx, void* a_) {
    NoAllocArgs& a = *static_cast<NoAllocArgs*>(a_);
    const dmc_grid& g = a.x.vols[0].grid;
    check_grid(ctx, g);
    const int64_t npts = (int64_t)g.nx * g.ny * g.nz;
    if (npts > kNoAllocMaxPoints)
        fail(ctx, DMC_ERR_CAPACITY, "dmc_extract_noalloc takes at most %lld grid points (got %lld)",
             (long long)kNoAllocMaxPoints, (long long)npts);
    const dmc_out_buffers& ob = *a.ob;
    const bool tris = (a.x.flags & DMC_TRIANGLES) != 0;
    if (!ob.positions || !ob.normals || !ob.quads || !ob.twins || !ob.he_flags || !ob.he_v || (tris && !ob.tris) ||
        ob.cap_v < 0 || ob.cap_q < 0)
        fail(ctx, DMC_ERR_INVALID_ARG, "output buffers missing");
    void* base = g_scratch_ptr[ctx->device];
    if (!base) fail(ctx, DMC_ERR_STATE, "no scratch address for device %d", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu[ctx->device]);
    std::vector<Buf> saved;
    saved.reserve(ctx->bufs.size());
    for (Buf* b : ctx->bufs) {
        saved.push_back(*b);
        *b = Buf{};
    }
    ctx->arena = static_cast<char*>(base);
    ctx->arena_cap = kScratchBytes;
    ctx->arena_used = 0;
    auto restore = [&] {
        // nothing of the call may still run on the scratch when it returns
        (void)hipStreamSynchronize(ctx->stream);
        for (size_t i = 0; i < saved.size(); i++) *ctx->bufs[i] = saved[i];
        ctx->arena = nullptr;
        ctx->arena_cap = ctx->arena_used = 0;
        ctx->have_mesh = false;
        ctx->have_tmesh = false;
        ctx->vmesh.clear();
        ctx->vbnd.clear();
    };
    try {
        do_extract(ctx, &a.x);
        dmc_mesh m = ctx->vmesh[0];
        if (m.nv > ob.cap_v || m.nq > ob.cap_q) {
            *a.x.out = m;
            fail(ctx, DMC_ERR_CAPACITY, "mesh of %lld vertices, %lld quadrilaterals exceeds the output buffers "
                 "(%lld, %lld)", (long long)m.nv, (long long)m.nq, (long long)ob.cap_v, (long long)ob.cap_q);
        }
        auto d2d = [&](void* dst, const void* src, size_t bytes) {
            if (bytes) HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        };
        d2d(ob.positions, m.d_positions, 24 * (size_t)m.nv);
        d2d(ob.normals, m.d_normals, 24 * (size_t)m.nv);
        d2d(ob.quads, m.d_quads, 16 * (size_t)m.nq);
        if (tris) d2d(ob.tris, m.d_tris, 12 * (size_t)m.nt);
        d2d(ob.twins, m.d_twins, 16 * (size_t)m.nq);
        d2d(ob.he_flags, m.d_he_flags, 4 * (size_t)m.nq);
        d2d(ob.he_v, m.d_he_v, 4 * (size_t)m.nv);
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
        m.d_positions = ob.positions;
        m.d_normals = ob.normals;
        m.d_quads = ob.quads;
        m.d_tris = tris ? ob.tris : nullptr;
        m.d_twins = ob.twins;
        m.d_he_flags = ob.he_flags;
        m.d_he_v = ob.he_v;
        *a.x.out = m;
        if (ob.bnd) {
            if ((int64_t)ctx->vbnd[0].size() > ob.cap_bnd)
                fail(ctx, DMC_ERR_CAPACITY, "%lld boundary vertices exceed bnd's %lld", (long long)ctx->vbnd[0].size(),
                     (long long)ob.cap_bnd);
            if (!ctx->vbnd[0].empty()) memcpy(ob.bnd, ctx->vbnd[0].data(), ctx->vbnd[0].size() * sizeof(int32_t));
        }
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

dmc_status dmc_extract_noalloc(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type, double iso,
                               uint32_t flags, const double* gradients, const dmc_out_buffers* out_buffers,
                               dmc_mesh* out) {
    if (!ctx || !grid || !out_buffers || !out) return DMC_ERR_INVALID_ARG;
    const dmc_volume vol{*grid, field, gradients, iso};
    // device field, no stage events (they would be created)
    const uint32_t f = (flags | DMC_FIELD_ON_DEVICE) & ~(uint32_t)DMC_STAGE_TIMING;
    NoAllocArgs a{ExtractArgs{&vol, 1, field_type, f, out}, out_buffers};
    return guard(ctx, do_extract_noalloc, &a);
}

dmc_status tmc_extract(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type, double iso,
                       uint32_t flags, tmc_mesh* out) {
    if (!ctx || !grid) return DMC_ERR_INVALID_ARG;
    if (flags & DMC_GRADIENTS_SUPPLIED) return DMC_ERR_INVALID_ARG;
    const dmc_volume vol{*grid, field, nullptr, iso};
    ExtractArgs a{&vol, 1, field_type, flags & (DMC_FIELD_ON_DEVICE | DMC_STAGE_TIMING), nullptr, out};
    return guard(ctx, do_extract, &a);
}

struct TmcDownloadArgs {
    double *pos, *nrm;
    int32_t* tris;
};

static void do_tmc_download(dmc_context* ctx, void* a_) {
    TmcDownloadArgs& a = *static_cast<TmcDownloadArgs*>(a_);
    if (!ctx->have_tmesh) fail(ctx, DMC_ERR_STATE, "no mesh: call tmc_extract first");
    const tmc_mesh& m = ctx->tmesh;
    HostReads r(ctx);
    if (a.pos) r.add(a.pos, m.d_positions, 24 * (size_t)m.nv);
    if (a.nrm) r.add(a.nrm, m.d_normals, 24 * (size_t)m.nv);
    if (a.tris) r.add(a.tris, m.d_tris, 12 * (size_t)m.nt);
    r.wait();
}

dmc_status tmc_download(dmc_context* ctx, double* positions, double* normals, int32_t* triangles) {
    if (!ctx) return DMC_ERR_INVALID_ARG;
    TmcDownloadArgs a{positions, normals, triangles};
    return guard(ctx, do_tmc_download, &a);
}

struct DownloadArgs {
    int32_t index;   // volume of the last extraction; -1: the only one
    double *pos, *nrm;
    int32_t *quads, *tris, *he, *he_v, *he_f, *bnd;
};

static int32_t volume_index(dmc_context* ctx, int32_t index) {
    if (!ctx->have_mesh) fail(ctx, DMC_ERR_STATE, "no mesh: call dmc_extract first");
    if (index < 0) {
        if (ctx->vmesh.size() != 1)
            fail(ctx, DMC_ERR_STATE, "the last extraction was a batch: use dmc_batch_download");
        return 0;
    }
    if (index >= (int32_t)ctx->vmesh.size()) fail(ctx, DMC_ERR_INVALID_ARG, "volume index %d out of range", index);
    return index;
}

static void do_download(dmc_context* ctx, void* a_) {
    DownloadArgs& a = *static_cast<DownloadArgs*>(a_);
    const int32_t vi = volume_index(ctx, a.index);
    const dmc_mesh& m = ctx->vmesh[vi];
    const std::vector<int32_t>& bnd = ctx->vbnd[vi];
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    d2h(a.pos, m.d_positions, 3 * m.nv * sizeof(double));
    d2h(a.nrm, m.d_normals, 3 * m.nv * sizeof(double));
    d2h(a.quads, m.d_quads, 4 * m.nq * sizeof(int32_t));
    d2h(a.tris, m.d_tris, 3 * m.nt * sizeof(int32_t));
    d2h(a.he_v, m.d_he_v, m.nv * sizeof(int32_t));
    std::vector<int32_t> q, tw;
    std::vector<uint8_t> fl;
    if (a.he) {
        q.resize(4 * m.nq);
        tw.resize(4 * m.nq);
        fl.resize(4 * m.nq);
        d2h(q.data(), m.d_quads, 4 * m.nq * sizeof(int32_t));
        d2h(tw.data(), m.d_twins, 4 * m.nq * sizeof(int32_t));
        d2h(fl.data(), m.d_he_flags, 4 * m.nq);
    }
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    if (a.he) {
        for (int64_t h = 0; h < 4 * m.nq; h++) {
            int32_t* o = a.he + 5 * h;
            o[0] = q[h];
            o[1] = (int32_t)(h / 4);
            o[2] = (int32_t)((h & ~3) | ((h + 1) & 3));
            o[3] = tw[h];
            o[4] = fl[h];
        }
    }
    if (a.he_f)
        for (int64_t f = 0; f < m.nq; f++) a.he_f[f] = (int32_t)(4 * f);
    if (a.bnd && !bnd.empty()) memcpy(a.bnd, bnd.data(), bnd.size() * sizeof(int32_t));
}

dmc_status dmc_mesh_download(dmc_context* ctx, double* positions, double* normals, int32_t* quads, int32_t* tris,
                             int32_t* halfedges, int32_t* he_v, int32_t* he_f, int32_t* bnd) {
    DownloadArgs a{-1, positions, normals, quads, tris, halfedges, he_v, he_f, bnd};
    return guard(ctx, do_download, &a);
}

dmc_status dmc_batch_download(dmc_context* ctx, int32_t index, double* positions, double* normals, int32_t* quads,
                              int32_t* tris, int32_t* halfedges, int32_t* he_v, int32_t* he_f, int32_t* bnd) {
    if (index < 0) return DMC_ERR_INVALID_ARG;
    DownloadArgs a{index, positions, normals, quads, tris, halfedges, he_v, he_f, bnd};
    return guard(ctx, do_download, &a);
}

struct F32Args {
    float *pos, *nrm;
    int negate;
};

static void do_download_f32(dmc_context* ctx, void* a_) {
    F32Args& a = *static_cast<F32Args*>(a_);
    const dmc_mesh& m = ctx->vmesh[volume_index(ctx, -1)];
    const int64_t n = 3 * m.nv;
    std::vector<double> tmp(n);
    if (a.pos) {
        HIPC(ctx, hipMemcpy(tmp.data(), m.d_positions, n * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) a.pos[i] = (float)tmp[i];
    }
    if (a.nrm) {
        HIPC(ctx, hipMemcpy(tmp.data(), m.d_normals, n * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) a.nrm[i] = a.negate ? -(float)tmp[i] : (float)tmp[i];
    }
}

dmc_status dmc_mesh_download_f32(dmc_context* ctx, float* positions, float* normals, int negate_normals) {
    F32Args a{positions, normals, negate_normals};
    return guard(ctx, do_download_f32, &a);
}

// ---------------------------------------------------------------------------
// diagnostics of the last mesh
// ---------------------------------------------------------------------------
// Diagnostics work on the mesh of a single-volume extraction (a batch's
// arrays hold per-volume ids).
static void diag_mesh_ok(dmc_context* ctx) {
    if (!ctx->have_mesh) fail(ctx, DMC_ERR_STATE, "no mesh: call dmc_extract first");
    if (ctx->vmesh.size() != 1) fail(ctx, DMC_ERR_STATE, "diagnostics need a single-volume extraction");
}

// The final mesh's halfedge rows (off/adj), counted from the quadrilaterals
// when the final build was carried over from a compaction and paired
// without them.
static void ensure_rows(dmc_context* ctx) {
    if (ctx->rows_live) return;
    Sizes* S = dsizes(ctx);
    const int64_t vcap = ctx->vcap, qcap = ctx->qcap, hcap = 4 * qcap;
    const int32_t* quads = static_cast<const int32_t*>(ctx->quads[ctx->cur].p);
    uint32_t* deg = ensure<uint32_t>(ctx, ctx->deg2, vcap);
    uint32_t* off = ensure<uint32_t>(ctx, ctx->off, vcap + 1);
    uint32_t* cursor = ensure<uint32_t>(ctx, ctx->cursor, vcap);
    int2* adj = ensure<int2>(ctx, ctx->adj, hcap);
    const DevN nv = dn(&S->nv, vcap);
    fill_u32(ctx, deg, nv, 0u);
    fill_u32(ctx, cursor, nv, 0u);
    launch(ctx, "k_he_degree", k_he_degree, dim3(grid_for(ctx, hcap)), dim3(kBlock), quads, (const Sizes*)S, deg,
           &S->he.pad, (const uint32_t*)nullptr);
    scan_to(ctx, ScanCounts{deg}, off, nv, &S->tot, "k_scan", nullptr, off);
    launch(ctx, "k_he_scatter", k_he_scatter, dim3(grid_for(ctx, qcap)), dim3(kBlock), quads, (const Sizes*)S, off,
           cursor, adj, (const uint32_t*)nullptr);
    check_launch(ctx, "rows");
    ctx->rows_live = true;
}

struct EdgeArgs {
    int32_t* edges;
    int64_t* count;
};

static void do_edges(dmc_context* ctx, void* a_) {
    EdgeArgs& a = *static_cast<EdgeArgs*>(a_);
    diag_mesh_ok(ctx);
    const dmc_mesh& m = ctx->mesh;
    const int64_t nv = m.nv, nhe = 4 * m.nq;
    uint32_t* rep = ensure<uint32_t>(ctx, ctx->drep, nhe);
    uint32_t* nid = ensure<uint32_t>(ctx, ctx->dnew, nhe);
    ensure_rows(ctx);
    if (nhe > 0)
        launch(ctx, "k_edge_reps", k_edge_reps, dim3(grid_for(ctx, nhe)), dim3(kBlock),
                           static_cast<const int32_t*>(m.d_quads), nhe, nv,
                           static_cast<const uint32_t*>(ctx->off.p), static_cast<const int2*>(ctx->adj.p),
                           static_cast<const int32_t*>(m.d_twins), static_cast<const uint8_t*>(m.d_he_flags), rep);
    check_launch(ctx, "k_edge_reps");
    const int64_t ne = (int64_t)scan(ctx, rep, nid, nhe);
    *a.count = ne;
    if (a.edges && ne > 0) {
        int2* e = ensure<int2>(ctx, ctx->dedges, ne);
        launch(ctx, "k_edge_emit", k_edge_emit, dim3(grid_for(ctx, nhe)), dim3(kBlock),
                           static_cast<const int32_t*>(m.d_quads), nhe, rep, nid, e);
        check_launch(ctx, "k_edge_emit");
        HIPC(ctx, hipMemcpyAsync(a.edges, e, ne * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
    }
}

dmc_status dmc_mesh_edges(dmc_context* ctx, int32_t* edges, int64_t* count) {
    if (!count) return DMC_ERR_INVALID_ARG;
    EdgeArgs a{edges, count};
    return guard(ctx, do_edges, &a);
}

struct ValenceArgs {
    int triangles;
    uint32_t* hist;
    int32_t* len;
    int64_t* beyond;
};

static void do_valence_hist(dmc_context* ctx, void* a_) {
    ValenceArgs& a = *static_cast<ValenceArgs*>(a_);
    diag_mesh_ok(ctx);
    const dmc_mesh& m = ctx->mesh;
    if (a.triangles && !m.d_tris) fail(ctx, DMC_ERR_STATE, "no triangles: extract with DMC_TRIANGLES");
    const int64_t nv = m.nv;
    uint32_t* vval = ensure<uint32_t>(ctx, ctx->dval, nv);
    uint32_t* hist = ensure<uint32_t>(ctx, ctx->dhist, kValenceBins + 1);
    ensure_rows(ctx);
    HIPC(ctx, hipMemsetAsync(hist, 0, (kValenceBins + 1) * sizeof(uint32_t), ctx->stream));
    if (nv > 0) {
        launch(ctx, "k_vertex_valence", k_vertex_valence, dim3(grid_for(ctx, nv)), dim3(kBlock),
                           static_cast<const int32_t*>(m.d_quads), a.triangles ? m.d_tris : nullptr, nv,
                           static_cast<const uint32_t*>(ctx->off.p), static_cast<const int2*>(ctx->adj.p), vval);
        launch(ctx, "k_valence_hist", k_valence_hist, dim3(grid_for(ctx, nv, kBlock, 1024)), dim3(kBlock), vval, nv,
                           hist);
    }
    check_launch(ctx, "valence histogram");
    uint32_t h[kValenceBins + 1];
    HIPC(ctx, hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    int32_t len = 1;
    for (int k = 1; k < kValenceBins; k++)
        if (h[k] > 0) len = k + 1;
    memcpy(a.hist, h, kValenceBins * sizeof(uint32_t));
    *a.len = len;
    if (a.beyond) *a.beyond = h[kValenceBins];
}

dmc_status dmc_valence_histogram(dmc_context* ctx, int triangles, uint32_t* hist, int32_t* len, int64_t* beyond) {
    if (!hist || !len) return DMC_ERR_INVALID_ARG;
    ValenceArgs a{triangles, hist, len, beyond};
    return guard(ctx, do_valence_hist, &a);
}

struct QualityArgs {
    int which;
    float* quality;
};

static void do_quality(dmc_context* ctx, void* a_) {
    QualityArgs& a = *static_cast<QualityArgs*>(a_);
    diag_mesh_ok(ctx);
    const dmc_mesh& m = ctx->mesh;
    if (a.which == 1 && !m.d_tris) fail(ctx, DMC_ERR_STATE, "no triangles: extract with DMC_TRIANGLES");
    const int64_t n = a.which == 0 ? m.nq : m.nt;
    if (n == 0) return;
    float* q = ensure<float>(ctx, ctx->dqual, n);
    if (a.which == 0)
        launch(ctx, "k_quality_quads", k_quality_quads, dim3(grid_for(ctx, n)), dim3(kBlock),
                           static_cast<const int32_t*>(m.d_quads), n, static_cast<const double*>(m.d_positions), q);
    else
        launch(ctx, "k_quality_tris", k_quality_tris, dim3(grid_for(ctx, n)), dim3(kBlock),
                           static_cast<const int32_t*>(m.d_tris), n, static_cast<const double*>(m.d_positions), q);
    check_launch(ctx, "element quality");
    HIPC(ctx, hipMemcpyAsync(a.quality, q, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
}

dmc_status dmc_element_quality(dmc_context* ctx, int which, float* quality) {
    if (!quality || (which != 0 && which != 1)) return DMC_ERR_INVALID_ARG;
    QualityArgs a{which, quality};
    return guard(ctx, do_quality, &a);
}

struct ObjArgs {
    const char* path;
    int triangles;
};

// cpp_dmc/main.cpp:35-87 writes doubles with the stream's default format,
// which is printf's "%g" (precision 6), and 1-based "i//i" face corners.
static void do_write_obj(dmc_context* ctx, void* a_) {
    ObjArgs& a = *static_cast<ObjArgs*>(a_);
    diag_mesh_ok(ctx);
    const dmc_mesh& m = ctx->mesh;
    if (a.triangles && !m.d_tris) fail(ctx, DMC_ERR_STATE, "no triangles: extract with DMC_TRIANGLES");
    const int corners = a.triangles ? 3 : 4;
    const int64_t nf = a.triangles ? m.nt : m.nq;
    std::vector<double> pos(3 * m.nv), nrm(3 * m.nv);
    std::vector<int32_t> faces(corners * nf);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    d2h(pos.data(), m.d_positions, pos.size() * sizeof(double));
    d2h(nrm.data(), m.d_normals, nrm.size() * sizeof(double));
    d2h(faces.data(), a.triangles ? m.d_tris : m.d_quads, faces.size() * sizeof(int32_t));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    FILE* f = fopen(a.path, "wb");
    if (!f) fail(ctx, DMC_ERR_INVALID_ARG, "cannot open %s for writing", a.path);
    std::string out;
    out.reserve(1 << 20);
    char buf[160];
    auto flush = [&](bool force) {
        if (force || out.size() > (1u << 20)) {
            if (!out.empty() && fwrite(out.data(), 1, out.size(), f) != out.size()) {
                fclose(f);
                fail(ctx, DMC_ERR_INVALID_ARG, "short write to %s", a.path);
            }
            out.clear();
        }
    };
    out += a.triangles ? "#Dual Marching Cubes, triangle mesh\n" : "#Dual Marching Cubes, quad only mesh\n";
    for (int64_t i = 0; i < m.nv; i++) {
        snprintf(buf, sizeof(buf), "v %g %g %g\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        out += buf;
        flush(false);
    }
    for (int64_t i = 0; i < m.nv; i++) {
        snprintf(buf, sizeof(buf), "vn %g %g %g\n", nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        out += buf;
        flush(false);
    }
    for (int64_t i = 0; i < nf; i++) {
        out += "f";
        for (int c = 0; c < corners; c++) {
            const long long x = (long long)faces[corners * i + c] + 1;
            snprintf(buf, sizeof(buf), " %lld//%lld", x, x);
            out += buf;
        }
        out += "\n";
        flush(false);
    }
    flush(true);
    if (fclose(f) != 0) fail(ctx, DMC_ERR_INVALID_ARG, "cannot close %s", a.path);
}

dmc_status dmc_write_obj(dmc_context* ctx, const char* path, int triangles) {
    if (!path) return DMC_ERR_INVALID_ARG;
    ObjArgs a{path, triangles};
    return guard(ctx, do_write_obj, &a);
}

dmc_status dmc_context_bytes(dmc_context* ctx, size_t* bytes) {
    if (!ctx || !bytes) return DMC_ERR_INVALID_ARG;
    size_t n = 0;
    for (const Buf* b : ctx->bufs) n += b->cap;
    *bytes = n;
    return DMC_OK;
}

dmc_status dmc_context_trim(dmc_context* ctx) {
    if (!ctx) return DMC_ERR_INVALID_ARG;
    if (ctx->arena) return DMC_ERR_STATE;
    (void)hipSetDevice(ctx->device);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    for (Buf* b : ctx->bufs) {
        if (b->p) (void)hipFree(b->p);
        *b = Buf{};
    }
    ctx->buf_epoch++;
    ctx->have_mesh = false;
    ctx->have_tmesh = false;
    ctx->vmesh.clear();
    ctx->vbnd.clear();
    if (e != hipSuccess) {
        ctx->err = hipGetErrorString(e);
        return DMC_ERR_HIP;
    }
    return DMC_OK;
}

dmc_status dmc_get_launch(dmc_context* ctx, dmc_launch* out) {
    if (!ctx || !out) return DMC_ERR_INVALID_ARG;
    out->grid_cap = ctx->grid_cap;
    out->sweep_waves = ctx->sweep_waves;
    return DMC_OK;
}

dmc_status dmc_set_launch(dmc_context* ctx, const dmc_launch* l) {
    if (!ctx || !l) return DMC_ERR_INVALID_ARG;
    if (l->grid_cap < kGridCapMin || l->grid_cap > kGridCapMax || l->sweep_waves < kSweepWavesMin ||
        l->sweep_waves > kSweepWavesMax) {
        ctx->err = "launch shape out of range";
        return DMC_ERR_INVALID_ARG;
    }
    ctx->grid_cap = l->grid_cap;
    ctx->sweep_waves = l->sweep_waves;
    return DMC_OK;
}

dmc_status dmc_stage_times(dmc_context* ctx, const char** names, float* ms, int* count) {
    if (!ctx || !count) return DMC_ERR_INVALID_ARG;
    const int n = (int)ctx->stage_ms.size();
    const int m = std::min(n, *count);
    for (int i = 0; i < m; i++) {
        if (names) names[i] = ctx->stage_ms[i].first;
        if (ms) ms[i] = ctx->stage_ms[i].second;
    }
    *count = n;
    return DMC_OK;
}

struct MemArgs {
    void* dst;
    const void* src;
    size_t bytes;
    void** out;
    int kind;
};

static void do_mem(dmc_context* ctx, void* a_) {
    MemArgs& a = *static_cast<MemArgs*>(a_);
    switch (a.kind) {
    case 0: {
        hipError_t e = hipMalloc(a.out, a.bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(ctx, DMC_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", a.bytes, hipGetErrorString(e));
        }
        break;
    }
    case 1: HIPC(ctx, hipFree(a.dst)); break;
    case 2: HIPC(ctx, hipMemcpy(a.dst, a.src, a.bytes, hipMemcpyHostToDevice)); break;
    case 3: HIPC(ctx, hipMemcpy(a.dst, a.src, a.bytes, hipMemcpyDeviceToHost)); break;
    case 4: HIPC(ctx, hipStreamSynchronize(ctx->stream)); break;
    }
}

dmc_status dmc_device_alloc(dmc_context* ctx, size_t bytes, void** ptr) {
    if (!ptr) return DMC_ERR_INVALID_ARG;
    MemArgs a{nullptr, nullptr, bytes, ptr, 0};
    return guard(ctx, do_mem, &a);
}
dmc_status dmc_device_free(dmc_context* ctx, void* ptr) {
    MemArgs a{ptr, nullptr, 0, nullptr, 1};
    return guard(ctx, do_mem, &a);
}
dmc_status dmc_copy_to_device(dmc_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return DMC_ERR_INVALID_ARG;
    MemArgs a{dst, src, bytes, nullptr, 2};
    return guard(ctx, do_mem, &a);
}
dmc_status dmc_copy_to_host(dmc_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return DMC_ERR_INVALID_ARG;
    MemArgs a{dst, src, bytes, nullptr, 3};
    return guard(ctx, do_mem, &a);
}
dmc_status dmc_synchronize(dmc_context* ctx) {
    MemArgs a{nullptr, nullptr, 0, nullptr, 4};
    return guard(ctx, do_mem, &a);
}

}  // extern "C"
