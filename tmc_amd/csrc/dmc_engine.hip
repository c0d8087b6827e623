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
#include <string>
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
// style): enough to fill every CU several times over, few enough that waves
// live long and keep many loads in flight.  The field sweep aims for
// kSweepWaves waves, which sets its planes per wave.
constexpr int kGridCap = 16384;
constexpr int kSweepWaves = 32768;
constexpr int64_t kGridCapMin = 64, kGridCapMax = 1 << 24;
constexpr int64_t kSweepWavesMin = 256, kSweepWavesMax = 1 << 24;

struct StageMark {
    const char* name;
    hipEvent_t ev;
};

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
    Buf field, grads, fails, scanPartial, scanTotal;
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
    int csr_ready = 0;  // off/adj already hold the live mesh's halfedge rows: 1 from the cells, 2 remapped
    // halfedges / simplification scratch
    Buf deg, off, cursor, adj, twin, flag, hev, ovr, bndrec, nbnd, val, p, em, vfirst, vsecond, lastT, lastR,
        newV, newQ, clist, ccnt, tris, stats, off2, adj2, rbad, qmap, vmap, rlist, mlist, slow;
    // diagnostics of the last mesh
    Buf drep, dnew, dval, dhist, dedges, dqual;
    // topologically correct MC: per active cell vertex / triangle counts and
    // offsets, the ambiguous-cell list, owner words; stats; the mesh
    Buf tcntv, tcntt, tvofs, ttofs, tamb, tvb, tstats, tpos, tnrm, ttris;
    tmc_mesh tmesh{};
    bool have_tmesh = false;
    uint32_t last_bad = 0;   // invalid vertex references seen by the last halfedge build
    // pinned host staging of device-to-host reads (counts, totals, tables)
    void* pin = nullptr;
    size_t pin_cap = 0;
    // deferred reads: device values copied into pinned slots without
    // waiting, valid once a later stream synchronisation has returned
    void* dpin = nullptr;
    int64_t syncs = 0;         // stream synchronisations of HostReads so far
    int64_t bad_epoch = -1;    // halfedge stats of a non-final build pending in slot 0
    int64_t rows_epoch = -1;   // first CSR's row total pending in slot 1 ...
    uint64_t rows_expect = 0;  // ... and the value it must have (4 Q)
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
    std::vector<std::pair<const char*, float>> stage_ms;

    dmc_context() {
        Buf* all[] = {&segCount, &segBase, &segMask, &lists, &acell, &acode, &aaux, &avloc, &aqloc, &rowV, &rowQ, &rowBaseV, &rowBaseQ,
                      &field, &grads, &fails, &scanPartial, &scanTotal, &pos[0], &pos[1], &nrm[0], &nrm[1], &vslot[0], &vslot[1],
                      &quads[0], &quads[1], &colors, &deg, &off, &cursor, &adj, &twin, &flag, &hev, &ovr,
                      &bndrec, &nbnd, &val, &p, &em, &vfirst, &vsecond, &lastT, &lastR, &newV,
                      &newQ, &clist, &ccnt, &tris, &stats, &vtab, &vstat, &vbound, &qbound, &vhe,
                      &off2, &adj2, &rbad, &qmap, &vmap, &rlist, &mlist, &slow, &drep, &dnew, &dval,
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
    if (b.cap < bytes) {
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

void mark(dmc_context* ctx, const char* name) {
    if (!ctx->timing) return;
    StageMark m{name, nullptr};
    HIPC(ctx, hipEventCreate(&m.ev));
    HIPC(ctx, hipEventRecord(m.ev, ctx->stream));
    ctx->marks.push_back(m);
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

// Deferred reads: `bytes` (at most kDeferSlot) of device memory copied into
// pinned slot `slot` without waiting.  Returns the synchronisation count at
// the copy; the slot is valid once a later synchronisation has returned.
constexpr size_t kDeferSlot = 256;
int64_t defer_read(dmc_context* ctx, int slot, const void* src, size_t bytes) {
    if (!ctx->dpin) {
        if (hipHostMalloc(&ctx->dpin, 4 * kDeferSlot) != hipSuccess) {
            (void)hipGetLastError();
            ctx->dpin = nullptr;
            fail(ctx, DMC_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu) failed", 4 * kDeferSlot);
        }
    }
    HIPC(ctx, hipMemcpyAsync(static_cast<char*>(ctx->dpin) + slot * kDeferSlot, src, bytes, hipMemcpyDeviceToHost,
                             ctx->stream));
    return ctx->syncs;
}

// A deferred slot's contents, synchronising first if nothing has since
// the copy.
const void* deferred(dmc_context* ctx, int slot, int64_t epoch) {
    if (ctx->syncs <= epoch) {
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
        ctx->syncs++;
    }
    return static_cast<const char*>(ctx->dpin) + slot * kDeferSlot;
}

// Device slots for scan totals: 0 for the scans whose total is read at
// once, the others for totals kept on the device or read later.
enum TotalSlot { kTotRead = 0, kTotV, kTotQ, kTotRows, kTotRemap, kTotSlots };
uint64_t* totals(dmc_context* ctx) { return ensure<uint64_t>(ctx, ctx->scanTotal, kTotSlots); }

// Exclusive scan of n uint32 counts into out; the 64-bit total goes to
// *dtotal (device memory) and the host does not wait for it.
template <class In>
void scan_to(dmc_context* ctx, In in, uint32_t* out, int64_t n, uint64_t* dtotal) {
    if (n <= 0) {
        HIPC(ctx, hipMemsetAsync(dtotal, 0, sizeof(uint64_t), ctx->stream));
        return;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    uint64_t* part = ensure<uint64_t>(ctx, ctx->scanPartial, tiles);
    hipLaunchKernelGGL(k_scan_reduce<In>, dim3((unsigned)tiles), dim3(kBlock), 0, ctx->stream, in, n, part);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, part, tiles, dtotal);
    hipLaunchKernelGGL(k_scan_down<In>, dim3((unsigned)tiles), dim3(kBlock), 0, ctx->stream, in, out, n, part);
    check_launch(ctx, "scan");
}

// Exclusive scan of n uint32 counts into out; returns the 64-bit total.
template <class In>
uint64_t scan_any(dmc_context* ctx, In in, uint32_t* out, int64_t n) {
    if (n <= 0) return 0;
    uint64_t* tot = totals(ctx) + kTotRead;
    scan_to(ctx, in, out, n, tot);
    return read_u64(ctx, tot);
}

// off[n] = total of a scan held on the device (low 32 bits).
void set_row_end(dmc_context* ctx, uint32_t* off, int64_t n, const uint64_t* dtotal) {
    HIPC(ctx, hipMemcpyAsync(off + n, dtotal, sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
}

// Exclusive scan of uint32 counts; returns the total.
uint64_t scan(dmc_context* ctx, const uint32_t* in, uint32_t* out, int64_t n) {
    return scan_any(ctx, ScanCounts{in}, out, n);
}

// New indices of the kept items (exclusive scan of removed[i] == 0);
// returns the number kept.
uint64_t scan_keep(dmc_context* ctx, const uint8_t* removed, uint32_t* out, int64_t n) {
    return scan_any(ctx, ScanKeep{removed}, out, n);
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
    uint64_t A;     // active cells
    int64_t R;      // cell rows (all volumes)
    VolTable vt;
    VolStat* vstat;
    uint32_t *lists, *segCount, *segBase;
    unsigned long long* segMask;
};

// Volume table, classification sweep and the scan of the active cells --
// shared by Dual Marching Cubes and the topologically correct MC.
template <class T>
Swept sweep_cells(dmc_context* ctx, const std::vector<HostVol>& hv) {
    const int n = (int)hv.size();
    ctx->nvol = n;
    // sweep tiling: J rows per wave; one plane count K for the whole
    // extraction, chosen so that about ctx->sweep_waves waves fill the chip
    std::vector<VolDesc> desc(n);
    int64_t units = 0;
    for (int v = 0; v < n; v++) {
        const dmc_grid& g = hv[v].g;
        VolDesc& D = desc[v];
        D.field = hv[v].field;
        D.grads = hv[v].grads;
        D.d = Dims{g.nx, g.ny, g.nz, (int64_t)g.nx - 1, (int64_t)(g.ny - 1) * (g.nz - 1)};
        D.geo = GridGeom{g.x0, g.y0, g.z0, g.dx, g.dy, g.dz};
        D.iso = hv[v].iso;
        D.iso_t = sizeof(T) == 4 ? (double)smallest_float_not_below(hv[v].iso) : hv[v].iso;
        D.S = (int)((D.d.C + kSeg - 1) / kSeg);
        D.ntj = (g.ny - 1 + kSweepRows - 1) / kSweepRows;
        units += (int64_t)(g.nz - 1) * D.S * D.ntj;
    }
    const int Kg = (int)std::max<int64_t>(8, (units + ctx->sweep_waves - 1) / ctx->sweep_waves);
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
    // pageable source: the copy has been staged when the call returns
    HIPC(ctx, hipMemcpyAsync(ctx->vtab.p, ctx->vtab_host.data(), tab_bytes, hipMemcpyHostToDevice, ctx->stream));
    const VolTable vt = vol_table(ctx);
    VolStat* vstat = ensure<VolStat>(ctx, ctx->vstat, n);
    HIPC(ctx, hipMemsetAsync(vstat, 0, (size_t)n * sizeof(VolStat), ctx->stream));

    uint32_t* segCount = ensure<uint32_t>(ctx, ctx->segCount, RS);
    uint32_t* segBase = ensure<uint32_t>(ctx, ctx->segBase, RS);
    uint32_t* lists = ensure<uint32_t>(ctx, ctx->lists, (size_t)RS * kSeg);
    unsigned long long* segMask = ensure<unsigned long long>(ctx, ctx->segMask, 4 * (size_t)RS);
    mark(ctx, "sweep");
    const int64_t sblocks = (W + kBlock / kWave - 1) / (kBlock / kWave);
    const dim3 sgrid((unsigned)((sblocks + kXcds - 1) / kXcds * kXcds));
    if (vec)
        hipLaunchKernelGGL((k_sweep<T, true>), sgrid, dim3(kBlock), 0, ctx->stream, vt, W, lists, segCount, segMask);
    else
        hipLaunchKernelGGL((k_sweep<T, false>), sgrid, dim3(kBlock), 0, ctx->stream, vt, W, lists, segCount,
                           segMask);
    check_launch(ctx, "k_sweep");
    mark(ctx, "rows");
    const uint64_t A = scan(ctx, segCount, segBase, RS);
    if (A > 0xFFFFFFFFull) fail(ctx, DMC_ERR_CAPACITY, "too many active cells (%llu)", (unsigned long long)A);
    ctx->mesh.active_cells = (int64_t)A;
    hipLaunchKernelGGL(k_vol_abase, dim3(grid_for(ctx, n + 1)), dim3(kBlock), 0, ctx->stream, vt, segBase, (int64_t)A,
                       vol_ptr(ctx, 2));
    return Swept{A, R, vt, vstat, lists, segCount, segBase, segMask};
}

template <class T>
void run_cells(dmc_context* ctx, const std::vector<HostVol>& hv) {
    const Swept sw = sweep_cells<T>(ctx, hv);
    const int n = (int)hv.size();
    const uint64_t A = sw.A;
    const int64_t R = sw.R;
    const VolTable vt = sw.vt;
    VolStat* vstat = sw.vstat;
    uint32_t *lists = sw.lists, *segCount = sw.segCount, *segBase = sw.segBase;
    unsigned long long* segMask = sw.segMask;
    uint32_t* acell = ensure<uint32_t>(ctx, ctx->acell, A);
    uint32_t* acode = ensure<uint32_t>(ctx, ctx->acode, A);
    uint32_t* aaux = ensure<uint32_t>(ctx, ctx->aaux, A);
    uint32_t* avloc = ensure<uint32_t>(ctx, ctx->avloc, A);
    uint32_t* aqloc = ensure<uint32_t>(ctx, ctx->aqloc, A);
    uint32_t* rowV = ensure<uint32_t>(ctx, ctx->rowV, R);
    uint32_t* rowQ = ensure<uint32_t>(ctx, ctx->rowQ, R);
    uint32_t* rowBaseV = ensure<uint32_t>(ctx, ctx->rowBaseV, R);
    uint32_t* rowBaseQ = ensure<uint32_t>(ctx, ctx->rowBaseQ, R);
    // fails[4]: vertices of valence 0 (all valence passes); fails[5]: quads
    // with an unresolved neighbour cell
    uint32_t* fails = ensure<uint32_t>(ctx, ctx->fails, 6);
    HIPC(ctx, hipMemsetAsync(fails, 0, 6 * sizeof(uint32_t), ctx->stream));
    hipLaunchKernelGGL(k_rows<T>, dim3(grid_for(ctx, R, kBlock / kWave)), dim3(kBlock), 0, ctx->stream, vt, R, lists,
                       segCount, segBase, acell, acode, aaux, avloc, aqloc, rowV, rowQ, vstat);
    check_launch(ctx, "k_rows");
    uint64_t* tot = totals(ctx);
    scan_to(ctx, ScanCounts{rowV}, rowBaseV, R, tot + kTotV);
    scan_to(ctx, ScanCounts{rowQ}, rowBaseQ, R, tot + kTotQ);
    uint64_t V = 0, Q = 0;
    {
        HostReads rd(ctx);
        rd.add(&V, tot + kTotV, sizeof(V));
        rd.add(&Q, tot + kTotQ, sizeof(Q));
        rd.wait();
    }
    if (V > 0x7FFFFFFFull || Q > 0x1FFFFFFFull)
        fail(ctx, DMC_ERR_CAPACITY, "mesh too large (V=%llu, Q=%llu)", (unsigned long long)V, (unsigned long long)Q);
    // [0, n]: live bounds; [n+1, 2n+1]: the cell stage's (before simplification)
    int64_t* vb = ensure<int64_t>(ctx, ctx->vbound, 2 * (size_t)(n + 1));
    int64_t* qb = ensure<int64_t>(ctx, ctx->qbound, 2 * (size_t)(n + 1));
    hipLaunchKernelGGL(k_vol_bounds, dim3(grid_for(ctx, n + 1)), dim3(kBlock), 0, ctx->stream, vt, rowBaseV, rowBaseQ,
                       (int64_t)V, (int64_t)Q, vb, qb);
    HIPC(ctx, hipMemcpyAsync(vb + n + 1, vb, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(ctx, hipMemcpyAsync(qb + n + 1, qb, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->cur = 0;
    ctx->pcur = 0;
    ctx->scur = 0;
    ctx->slots = false;
    double* pos = ensure<double>(ctx, ctx->pos[0], 3 * V);
    double* nrm = ensure<double>(ctx, ctx->nrm[0], 3 * V);
    int32_t* quads = ensure<int32_t>(ctx, ctx->quads[0], 4 * Q);
    int32_t* colors = ensure<int32_t>(ctx, ctx->colors, Q);
    // first halfedge CSR straight from the cells: degree of each vertex from
    // k_vertices, rows filled by k_quads
    uint32_t* deg = ensure<uint32_t>(ctx, ctx->deg, V);
    uint32_t* off = ensure<uint32_t>(ctx, ctx->off, V + 1);
    int2* adj = ensure<int2>(ctx, ctx->adj, 4 * Q);
    // cells deferred to the general sampling: slow[0] = count, then cell ids
    uint32_t* slow = ensure<uint32_t>(ctx, ctx->slow, A + 1);
    mark(ctx, "vertices");
    if (A > 0) {
        HIPC(ctx, hipMemsetAsync(slow, 0, sizeof(uint32_t), ctx->stream));
        hipLaunchKernelGGL(k_vertices<T>, dim3(grid_for(ctx, A)), dim3(kBlock), 0, ctx->stream, vt, acell, acode, aaux,
                           avloc, rowBaseV, (int64_t)A, pos, nrm, deg, slow);
        check_launch(ctx, "k_vertices");
        hipLaunchKernelGGL(k_vertices_slow<T>, dim3(grid_for(ctx, A, kBlock, 1024)), dim3(kBlock), 0,
                           ctx->stream, vt, acell, acode, aaux, avloc, rowBaseV, slow, pos, nrm, deg, vstat);
        check_launch(ctx, "k_vertices_slow");
    }
    mark(ctx, "quads");
    // the row total must be 4 Q; checked at the next synchronisation
    scan_to(ctx, ScanCounts{deg}, off, (int64_t)V, tot + kTotRows);
    set_row_end(ctx, off, (int64_t)V, tot + kTotRows);
    ctx->rows_epoch = defer_read(ctx, 1, tot + kTotRows, sizeof(uint64_t));
    ctx->rows_expect = 4 * Q;
    if (A > 0) {
        hipLaunchKernelGGL(k_quads, dim3(grid_for(ctx, A)), dim3(kBlock), 0, ctx->stream, vt, acell, acode, aaux, avloc,
                           aqloc,
                           rowBaseV, rowBaseQ, (int64_t)A, segMask, segBase, quads, colors, off, adj, fails + 5);
        check_launch(ctx, "k_quads");
    }
    ctx->csr_ready = 1;
    ctx->mesh.nv = (int64_t)V;
    ctx->mesh.nq = (int64_t)Q;
    ctx->mesh.nv_cells = (int64_t)V;
    ctx->mesh.nq_cells = (int64_t)Q;
}

// ---------------------------------------------------------------------------
// halfedges
// ---------------------------------------------------------------------------
// Final pass extras shared by both halfedge paths: 3-face edges override
// he_v and give the boundary vertices (sorted by key on the host).
template <class ThreeLaunch>
void three_face_edges(dmc_context* ctx, const HeStats& hs, int64_t nv, ThreeLaunch launch) {
    ctx->bnd.clear();
    if (hs.three == 0) return;
    unsigned long long* ovr = ensure<unsigned long long>(ctx, ctx->ovr, nv);
    BndRec* recs = ensure<BndRec>(ctx, ctx->bndrec, hs.three);
    unsigned int* nrec = ensure<unsigned int>(ctx, ctx->nbnd, 1);
    HIPC(ctx, hipMemsetAsync(ovr, 0, nv * sizeof(unsigned long long), ctx->stream));
    HIPC(ctx, hipMemsetAsync(nrec, 0, sizeof(unsigned int), ctx->stream));
    launch(ovr, recs, nrec, 0);
    launch(ovr, recs, nrec, 1);
    check_launch(ctx, "three-face edges");
    std::vector<BndRec> h(hs.three);
    HostReads r(ctx);
    r.add(h.data(), recs, hs.three * sizeof(BndRec));
    r.wait();
    std::sort(h.begin(), h.end(), [](const BndRec& a, const BndRec& b) { return a.key < b.key; });
    for (const BndRec& r : h) {
        ctx->bnd.push_back(r.a);
        ctx->bnd.push_back(r.b);
    }
}

// The first CSR's row total against 4 Q, once it has arrived.
void check_rows(dmc_context* ctx) {
    if (ctx->rows_epoch < 0) return;
    uint64_t rows = 0;
    std::memcpy(&rows, deferred(ctx, 1, ctx->rows_epoch), sizeof(rows));
    ctx->rows_epoch = -1;
    if (rows != ctx->rows_expect)
        fail(ctx, DMC_ERR_STATE, "halfedge rows %llu != 4 Q %llu", (unsigned long long)rows,
             (unsigned long long)ctx->rows_expect);
}

// Invalid vertex references seen by the last halfedge build (a non-final
// build's stats arrive deferred).
uint32_t last_bad(dmc_context* ctx) {
    if (ctx->bad_epoch >= 0) {
        HeStats hs{};
        std::memcpy(&hs, deferred(ctx, 0, ctx->bad_epoch), sizeof(hs));
        ctx->bad_epoch = -1;
        ctx->last_bad = hs.bad;
    }
    return ctx->last_bad;
}

// Halfedges through a CSR of outgoing (halfedge, destination) pairs: rows
// off[v] .. off[v+1] (off has nv+1 entries).  A non-final build leaves its
// stats on the device (read when the next stage needs them) and returns
// zeros.
HeStats build_halfedges(dmc_context* ctx, bool final_pass) {
    const int64_t nv = ctx->mesh.nv, nq = ctx->mesh.nq, nhe = 4 * nq;
    const int32_t* quads = static_cast<const int32_t*>(ctx->quads[ctx->cur].p);
    uint32_t* deg = ensure<uint32_t>(ctx, ctx->deg, nv);
    uint32_t* off = ensure<uint32_t>(ctx, ctx->off, nv + 1);
    uint32_t* cursor = ensure<uint32_t>(ctx, ctx->cursor, nv);
    int2* adj = ensure<int2>(ctx, ctx->adj, nhe);
    int32_t* twin = ensure<int32_t>(ctx, ctx->twin, nhe);
    uint8_t* flag = ensure<uint8_t>(ctx, ctx->flag, nhe);
    int32_t* hev = final_pass ? ensure<int32_t>(ctx, ctx->hev, nv) : nullptr;
    HeStats* st = ensure<HeStats>(ctx, ctx->stats, 1);
    HIPC(ctx, hipMemsetAsync(st, 0, sizeof(HeStats), ctx->stream));
    HeStats hs{};
    const int ready = ctx->csr_ready;
    ctx->csr_ready = 0;
    bool fresh = ready != 0;
    if (ready == 1) {
        // rows from the cell stage are complete unless a quadrilateral had an
        // unresolved neighbour cell
        uint32_t unresolved = 0;
        HostReads r(ctx);
        r.add(&unresolved, static_cast<uint32_t*>(ctx->fails.p) + 5, sizeof(unresolved));
        r.wait();
        check_rows(ctx);
        fresh = unresolved == 0;
    } else if (ready == 2) {
        // remapped rows: halfedges left without a valid origin were counted
        HIPC(ctx, hipMemcpyAsync(&st->bad, ctx->rbad.p, sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (!fresh) {
        if (nv > 0) {
            HIPC(ctx, hipMemsetAsync(deg, 0, nv * sizeof(uint32_t), ctx->stream));
            HIPC(ctx, hipMemsetAsync(cursor, 0, nv * sizeof(uint32_t), ctx->stream));
        }
        if (nhe > 0) {
            hipLaunchKernelGGL(k_he_degree, dim3(grid_for(ctx, nhe)), dim3(kBlock), 0, ctx->stream, quads, nhe, nv, deg,
                               &st->bad);
            check_launch(ctx, "k_he_degree");
        }
        const uint64_t rows = scan(ctx, deg, off, nv);
        HIPC(ctx, hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(off + nv), (int)(uint32_t)rows, 1, ctx->stream));
        if (nhe > 0) {
            hipLaunchKernelGGL(k_he_scatter, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, nv, off,
                               cursor, adj);
            check_launch(ctx, "k_he_scatter");
        }
    }
    unsigned long long* vst = nullptr;
    if (final_pass && ctx->nvol > 1) {
        vst = ensure<unsigned long long>(ctx, ctx->vhe, 2 * (size_t)ctx->nvol);
        HIPC(ctx, hipMemsetAsync(vst, 0, 2 * (size_t)ctx->nvol * sizeof(unsigned long long), ctx->stream));
    }
    if (nhe > 0) {
        // deg holds every vertex's out-degree here; the twins add boundary
        // arrivals, turning it into the valence
        hipLaunchKernelGGL(k_he_twins, dim3(grid_for(ctx, nhe)), dim3(kBlock), 0, ctx->stream, quads, nhe, nv, off, adj,
                           twin, flag, st, static_cast<const int64_t*>(ctx->qbound.p), ctx->nvol, vst,
                           final_pass ? nullptr : static_cast<uint32_t*>(ctx->deg.p));
        check_launch(ctx, "k_he_twins");
    }
    if (final_pass && nv > 0) {
        hipLaunchKernelGGL(k_he_v, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream, off, adj, nv, hev);
        check_launch(ctx, "k_he_v");
    }
    if (!final_pass) {
        ctx->bad_epoch = defer_read(ctx, 0, st, sizeof(HeStats));
        return hs;
    }
    {
        HostReads r(ctx);
        r.add(&hs, st, sizeof(hs));
        r.wait();
    }
    ctx->bad_epoch = -1;
    ctx->last_bad = hs.bad;
    if (final_pass)
        three_face_edges(ctx, hs, nv, [&](unsigned long long* ovr, BndRec* recs, unsigned int* nrec, int pass) {
            hipLaunchKernelGGL(k_he_three, dim3(grid_for(ctx, nhe)), dim3(kBlock), 0, ctx->stream, quads, nhe, nv, off,
                               adj, flag, ovr, recs, nrec, pass, hev);
        });
    return hs;
}

// Valence of the live mesh after a non-final halfedge build: the degree
// array, completed by the twins pass (out-degree + boundary arrivals).
int32_t* valence(dmc_context* ctx) {
    const int64_t nv = ctx->mesh.nv;
    uint32_t* val = static_cast<uint32_t*>(ctx->deg.p);
    unsigned int* zero = static_cast<unsigned int*>(ctx->fails.p) + 4;
    if (nv > 0)
        hipLaunchKernelGGL(k_count_zero, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream, val, nv, zero);
    check_launch(ctx, "valence");
    return reinterpret_cast<int32_t*>(val);
}

void color_faces(dmc_context* ctx) {
    const int64_t nq = ctx->mesh.nq;
    if (nq == 0) return;
    int32_t* colors = static_cast<int32_t*>(ctx->colors.p);
    const int nb = (int)std::min<int64_t>(kColorBlocks, (nq + kBlock - 1) / kBlock);
    uint32_t* hist = ensure<uint32_t>(ctx, ctx->ccnt, (size_t)kColorBlocks * 24 + 48);
    uint32_t* totals = hist + (size_t)kColorBlocks * 24;
    int32_t* list = ensure<int32_t>(ctx, ctx->clist, nq);
    hipLaunchKernelGGL(k_color_hist, dim3(nb), dim3(kBlock), 0, ctx->stream, colors, nq, hist);
    hipLaunchKernelGGL(k_color_offsets, dim3(1), dim3(kColorBlocks), 0, ctx->stream, hist, nb, totals);
    hipLaunchKernelGGL(k_color_scatter, dim3(nb), dim3(kBlock), 0, ctx->stream, colors, nq, hist, list);
    uint32_t h[48];
    HostReads r(ctx);
    r.add(h, totals, sizeof(h));
    r.wait();
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    for (int c = 5; c < 24; c++)
        if (h[c])
            hipLaunchKernelGGL(k_color_class, dim3(grid_for(ctx, h[c])), dim3(kBlock), 0, ctx->stream, list + h[24 + c],
                               (int64_t)h[c], twin, colors, 5);
    if (h[4])
        hipLaunchKernelGGL(k_color_class, dim3(grid_for(ctx, h[4])), dim3(kBlock), 0, ctx->stream, list + h[24 + 4],
                           (int64_t)h[4], twin, colors, 4);
    check_launch(ctx, "colour");
}

// Per-volume first vertex / quadrilateral after an order-preserving
// compaction (newidx = exclusive scan of the keep flags).
void remap_bounds(dmc_context* ctx, Buf& bounds, const uint32_t* newidx, int64_t nold, int64_t nnew) {
    hipLaunchKernelGGL(k_vol_remap, dim3(grid_for(ctx, ctx->nvol + 1)), dim3(kBlock), 0, ctx->stream,
                       static_cast<int64_t*>(bounds.p), ctx->nvol + 1, newidx, nold, nnew);
    check_launch(ctx, "k_vol_remap");
}

// Compact vertices by `removed`; vsecond (if non-null) receives new indices of
// kept vertices.  Returns the new count.  Positions and normals stay where
// they are; the slot map follows the vertices (gather_vertices lays them out
// at the end).
int64_t compact_vertices(dmc_context* ctx, const uint8_t* removed, int32_t* vsecond, uint32_t** newidx_out) {
    const int64_t nv = ctx->mesh.nv;
    uint32_t* newidx = ensure<uint32_t>(ctx, ctx->newV, nv);
    const int64_t nn = (int64_t)scan_keep(ctx, removed, newidx, nv);
    const uint32_t* slot = ctx->slots ? static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p) : nullptr;
    uint32_t* slot2 = ensure<uint32_t>(ctx, ctx->vslot[1 - ctx->scur], nn);
    int32_t* rlist = ensure<int32_t>(ctx, ctx->rlist, nv - nn);
    if (nv > 0)
        hipLaunchKernelGGL(k_compact_slots, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream, removed, newidx, nv, slot,
                           slot2, vsecond, rlist);
    check_launch(ctx, "k_compact_slots");
    ctx->scur = 1 - ctx->scur;
    ctx->slots = true;
    remap_bounds(ctx, ctx->vbound, newidx, nv, nn);
    if (newidx_out) *newidx_out = newidx;
    return nn;
}

// Lay the live vertices' positions and normals out densely (after the last
// compaction).
void gather_vertices(dmc_context* ctx) {
    if (!ctx->slots) return;
    const int64_t nv = ctx->mesh.nv;
    const int src = ctx->pcur, dst = 1 - ctx->pcur;
    double* pos2 = ensure<double>(ctx, ctx->pos[dst], 3 * nv);
    double* nrm2 = ensure<double>(ctx, ctx->nrm[dst], 3 * nv);
    if (nv > 0)
        hipLaunchKernelGGL(k_gather_vertices, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream,
                           static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p), nv,
                           static_cast<const double*>(ctx->pos[src].p), static_cast<const double*>(ctx->nrm[src].p),
                           pos2, nrm2);
    check_launch(ctx, "k_gather_vertices");
    ctx->pcur = dst;
    ctx->slots = false;
}

// Halfedge rows of the compacted mesh (nn vertices, nq2 quadrilaterals)
// from the current rows, into the spare CSR buffers, which then become the
// current ones.  Skipped (the next build counts from scratch) when the
// current rows left halfedges out.
void remap_rows(dmc_context* ctx, const Remap& m, int64_t nn, int64_t nq2, const int32_t* dlist, int64_t ndrop) {
    if (last_bad(ctx) != 0) return;
    const uint32_t* off = static_cast<const uint32_t*>(ctx->off.p);
    const int2* adj = static_cast<const int2*>(ctx->adj.p);
    uint32_t* deg2 = ensure<uint32_t>(ctx, ctx->deg, nn);
    uint32_t* off2 = ensure<uint32_t>(ctx, ctx->off2, nn + 1);
    uint32_t* cur2 = ensure<uint32_t>(ctx, ctx->cursor, nn);
    int2* adj2 = ensure<int2>(ctx, ctx->adj2, 4 * nq2);
    unsigned int* bad = ensure<unsigned int>(ctx, ctx->rbad, 1);
    HIPC(ctx, hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream));
    if (m.nv > 0) {
        // kept vertices: old row length less the corners in dropped quads
        hipLaunchKernelGGL(k_remap_degree_kept, dim3(grid_for(ctx, m.nv)), dim3(kBlock), 0, ctx->stream, m, off, deg2);
        if (ndrop > 0)
            hipLaunchKernelGGL(k_remap_drop, dim3(grid_for(ctx, ndrop)), dim3(kBlock), 0, ctx->stream, m,
                               static_cast<const int32_t*>(ctx->quads[ctx->cur].p), dlist, ndrop, deg2);
        if (m.nrem > 0)
            hipLaunchKernelGGL(k_remap_degree, dim3(grid_for(ctx, m.nrem)), dim3(kBlock), 0, ctx->stream, m, off, adj, nn,
                               deg2, 1, bad);
        check_launch(ctx, "k_remap_degree");
    }
    uint64_t* tot = totals(ctx) + kTotRemap;
    scan_to(ctx, ScanCounts{deg2}, off2, nn, tot);
    set_row_end(ctx, off2, nn, tot);
    if (m.nv > 0) {
        for (int pass = 0; pass < 2; pass++)
            hipLaunchKernelGGL(k_remap_rows, dim3(grid_for(ctx, pass == 0 ? m.nv : m.nrem)), dim3(kBlock), 0, ctx->stream,
                               m, off, adj, nn, off2, cur2, adj2, pass);
        check_launch(ctx, "k_remap_rows");
    }
    std::swap(ctx->off, ctx->off2);
    std::swap(ctx->adj, ctx->adj2);
    ctx->csr_ready = 2;
}

void simplify3x3y(dmc_context* ctx, bool colored) {
    build_halfedges(ctx, false);
    if (colored) color_faces(ctx);
    int32_t* val = valence(ctx);
    const int64_t nv = ctx->mesh.nv, nq = ctx->mesh.nq;
    const int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    uint8_t* p = ensure<uint8_t>(ctx, ctx->p, nq);
    uint8_t* em = ensure<uint8_t>(ctx, ctx->em, nq);
    uint8_t* vfirst = ensure<uint8_t>(ctx, ctx->vfirst, nv);
    int32_t* vsecond = ensure<int32_t>(ctx, ctx->vsecond, nv);
    int32_t* lastT = ensure<int32_t>(ctx, ctx->lastT, nv);
    int32_t* lastR = ensure<int32_t>(ctx, ctx->lastR, nv);
    if (nv > 0) {
        HIPC(ctx, hipMemsetAsync(vfirst, 0, nv, ctx->stream));
        HIPC(ctx, hipMemsetAsync(vsecond, 0xFF, nv * sizeof(int32_t), ctx->stream));
        HIPC(ctx, hipMemsetAsync(lastT, 0xFF, nv * sizeof(int32_t), ctx->stream));
        HIPC(ctx, hipMemsetAsync(lastR, 0xFF, nv * sizeof(int32_t), ctx->stream));
    }
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    const uint8_t* flag = static_cast<uint8_t*>(ctx->flag.p);
    const int32_t* colors = colored ? static_cast<int32_t*>(ctx->colors.p) : nullptr;
    uint32_t* newQ = ensure<uint32_t>(ctx, ctx->newQ, nq);
    if (nq > 0) {
        hipLaunchKernelGGL(k_mark3x3y, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, nv, val, p);
        hipLaunchKernelGGL(k_merge_claim, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, nv, val, p,
                           colors, twin, flag, lastT, lastR, em, vfirst);
        check_launch(ctx, "merge3x3y");
    }
    // the quadrilateral compaction's scan also lists the merging faces
    const int64_t qn = (int64_t)scan_keep(ctx, em, newQ, nq);
    if (nq > qn) {
        int32_t* mlist = ensure<int32_t>(ctx, ctx->mlist, nq - qn);
        hipLaunchKernelGGL(k_list_flagged, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, em, newQ, nq, mlist);
        hipLaunchKernelGGL(k_merge_apply, dim3(grid_for(ctx, nq - qn)), dim3(kBlock), 0, ctx->stream, quads, mlist, nq - qn,
                           em, val, lastT, lastR, static_cast<double*>(ctx->pos[ctx->pcur].p),
                           static_cast<double*>(ctx->nrm[ctx->pcur].p),
                           ctx->slots ? static_cast<const uint32_t*>(ctx->vslot[ctx->scur].p) : nullptr, vsecond);
        check_launch(ctx, "k_merge_apply");
    }
    const int64_t nn = compact_vertices(ctx, vfirst, vsecond, nullptr);
    const int dst = 1 - ctx->cur;
    int32_t* q2 = ensure<int32_t>(ctx, ctx->quads[dst], 4 * qn);
    int32_t* qmap = ensure<int32_t>(ctx, ctx->qmap, nq);
    int32_t* vmap = ensure<int32_t>(ctx, ctx->vmap, nv);
    HeStats* st = static_cast<HeStats*>(ctx->stats.p);
    if (nv > 0)
        hipLaunchKernelGGL(k_vmap3x3y, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream, vfirst, vsecond, nv, vmap);
    if (nq > 0)
        hipLaunchKernelGGL(k_compact_quads3x3y, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, em, newQ,
                           vmap, nv, q2, qmap, &st->bad);
    check_launch(ctx, "k_compact_quads3x3y");
    remap_bounds(ctx, ctx->qbound, newQ, nq, qn);
    remap_rows(ctx, Remap{qmap, vmap, vfirst, nullptr, static_cast<const int32_t*>(ctx->rlist.p), nv, nv - nn}, nn,
               qn, static_cast<const int32_t*>(ctx->mlist.p), nq - qn);
    ctx->cur = dst;
    ctx->mesh.nv = nn;
    ctx->mesh.nq = qn;
}

void simplify3333(dmc_context* ctx) {
    build_halfedges(ctx, false);
    int32_t* val = valence(ctx);
    const int64_t nv = ctx->mesh.nv, nq = ctx->mesh.nq;
    const int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    uint8_t* vflag = ensure<uint8_t>(ctx, ctx->vfirst, nv);
    int32_t* vtwin = ensure<int32_t>(ctx, ctx->vsecond, nv);
    int32_t* lastF = ensure<int32_t>(ctx, ctx->lastT, nv);
    uint8_t* p3333 = ensure<uint8_t>(ctx, ctx->p, nq);
    uint8_t* rflag = ensure<uint8_t>(ctx, ctx->em, nq);
    if (nv > 0) {
        HIPC(ctx, hipMemsetAsync(vflag, 0, nv, ctx->stream));
        HIPC(ctx, hipMemsetAsync(vtwin, 0xFF, nv * sizeof(int32_t), ctx->stream));
        HIPC(ctx, hipMemsetAsync(lastF, 0xFF, nv * sizeof(int32_t), ctx->stream));
    }
    if (nq > 0) {
        HIPC(ctx, hipMemsetAsync(p3333, 0, nq, ctx->stream));
        HIPC(ctx, hipMemsetAsync(rflag, 0, nq, ctx->stream));
    }
    const int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    const uint8_t* flag = static_cast<uint8_t*>(ctx->flag.p);
    if (nq > 0) {
        hipLaunchKernelGGL(k_mark3333, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, nv, twin, flag, val,
                           lastF, vflag, p3333, rflag);
        hipLaunchKernelGGL(k_twin3333, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, nv, twin, flag,
                           val, p3333, lastF, vtwin);
        check_launch(ctx, "mark3333");
    }
    uint32_t* newv = nullptr;
    const int64_t nn = compact_vertices(ctx, vflag, nullptr, &newv);
    uint32_t* newQ = ensure<uint32_t>(ctx, ctx->newQ, nq);
    const int64_t qn = (int64_t)scan_keep(ctx, rflag, newQ, nq);
    const int dst = 1 - ctx->cur;
    int32_t* q2 = ensure<int32_t>(ctx, ctx->quads[dst], 4 * qn);
    int32_t* qmap = ensure<int32_t>(ctx, ctx->qmap, nq);
    int32_t* vmap = ensure<int32_t>(ctx, ctx->vmap, nv);
    HeStats* st = static_cast<HeStats*>(ctx->stats.p);
    if (nv > 0)
        hipLaunchKernelGGL(k_vmap3333, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream, vflag, newv, nv, vmap);
    if (nq > 0)
        hipLaunchKernelGGL(k_compact_quads3333, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, quads, nq, rflag,
                           newQ, p3333, vtwin, vmap, nv, q2, qmap, &st->bad);
    check_launch(ctx, "k_compact_quads3333");
    remap_bounds(ctx, ctx->qbound, newQ, nq, qn);
    // the dropped quadrilaterals, for the row degrees
    int32_t* dlist = ensure<int32_t>(ctx, ctx->mlist, std::max<int64_t>(nq - qn, 1));
    if (nq > qn)
        hipLaunchKernelGGL(k_list_flagged, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream, rflag, newQ, nq, dlist);
    remap_rows(ctx, Remap{qmap, vmap, vflag, vtwin, static_cast<const int32_t*>(ctx->rlist.p), nv, nv - nn}, nn,
               qn, dlist, nq - qn);
    ctx->cur = dst;
    ctx->mesh.nv = nn;
    ctx->mesh.nq = qn;
}

void finish_timing(dmc_context* ctx) {
    ctx->stage_ms.clear();
    if (!ctx->timing || ctx->marks.empty()) return;
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i + 1 < ctx->marks.size(); i++) {
        float ms = 0;
        HIPC(ctx, hipEventElapsedTime(&ms, ctx->marks[i].ev, ctx->marks[i + 1].ev));
        ctx->stage_ms.emplace_back(ctx->marks[i].name, ms);
    }
    for (auto& m : ctx->marks) (void)hipEventDestroy(m.ev);
    ctx->marks.clear();
}

dmc_status guard(dmc_context* ctx, void (*fn)(dmc_context*, void*), void* arg) {
    if (!ctx) return DMC_ERR_INVALID_ARG;
    ctx->err.clear();
    try {
        HIPC(ctx, hipSetDevice(ctx->device));
        fn(ctx, arg);
        return DMC_OK;
    } catch (const Fail& f) {
        for (auto& m : ctx->marks) (void)hipEventDestroy(m.ev);
        ctx->marks.clear();
        return f.s;
    }
}

void check_grid(dmc_context* ctx, const dmc_grid& g) {
    if (g.nx < 2 || g.ny < 2 || g.nz < 2) fail(ctx, DMC_ERR_INVALID_ARG, "grid needs at least 2 points per axis");
    if (g.nx > (1 << 24)) fail(ctx, DMC_ERR_INVALID_ARG, "nx above 16777216");
    const uint64_t ncells = (uint64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    if (ncells > 0xFFFFFFFFull) fail(ctx, DMC_ERR_CAPACITY, "more than 2^32 cells");
}

// Topologically correct MC of one volume (mc/MarchingCubes.cpp t_mc): the
// sweep's active cells, their vertex and triangle counts, scans over the
// active cells (whose order is the reference's cell order), then vertices
// and triangles.
template <class T>
void run_tmc(dmc_context* ctx, const std::vector<HostVol>& hv) {
    const Swept sw = sweep_cells<T>(ctx, hv);
    const uint64_t A = sw.A;
    const int64_t R = sw.R;
    uint32_t* acell = ensure<uint32_t>(ctx, ctx->acell, A);
    uint32_t* acode = ensure<uint32_t>(ctx, ctx->acode, A);
    uint32_t* cnt_v = ensure<uint32_t>(ctx, ctx->tcntv, A);
    uint32_t* cnt_t = ensure<uint32_t>(ctx, ctx->tcntt, A);
    uint32_t* vofs = ensure<uint32_t>(ctx, ctx->tvofs, A);
    uint32_t* tofs = ensure<uint32_t>(ctx, ctx->ttofs, A);
    unsigned long long* st = ensure<unsigned long long>(ctx, ctx->tstats, 3);
    uint32_t* amb = ensure<uint32_t>(ctx, ctx->tamb, A + 1);
    HIPC(ctx, hipMemsetAsync(st, 0, 3 * sizeof(unsigned long long), ctx->stream));
    HIPC(ctx, hipMemsetAsync(amb, 0, sizeof(uint32_t), ctx->stream));
    hipLaunchKernelGGL(k_tmc_cells, dim3(grid_for(ctx, R, kBlock / kWave)), dim3(kBlock), 0, ctx->stream, sw.vt, R,
                       sw.lists, sw.segCount, sw.segBase, acell, acode, cnt_v, cnt_t, amb, st);
    check_launch(ctx, "k_tmc_cells");
    if (A > 0) {
        hipLaunchKernelGGL(k_tmc_amb_count<T>, dim3(grid_for(ctx, A, kBlock, 1024)), dim3(kBlock), 0, ctx->stream, sw.vt,
                           amb, acell, acode, cnt_v, cnt_t, st);
        check_launch(ctx, "k_tmc_amb_count");
    }
    uint64_t* tot = totals(ctx);
    scan_to(ctx, ScanCounts{cnt_v}, vofs, (int64_t)A, tot + kTotV);
    scan_to(ctx, ScanCounts{cnt_t}, tofs, (int64_t)A, tot + kTotQ);
    uint64_t V = 0, Tn = 0;
    {
        HostReads rd(ctx);
        rd.add(&V, tot + kTotV, sizeof(V));
        rd.add(&Tn, tot + kTotQ, sizeof(Tn));
        rd.wait();
    }
    if (V > 0x7FFFFFFFull || Tn > 0x7FFFFFFFull)
        fail(ctx, DMC_ERR_CAPACITY, "mesh too large (V=%llu, T=%llu)", (unsigned long long)V,
             (unsigned long long)Tn);
    double* pos = ensure<double>(ctx, ctx->tpos, 3 * V);
    double* nrm = ensure<double>(ctx, ctx->tnrm, 3 * V);
    int32_t* tris = ensure<int32_t>(ctx, ctx->ttris, 3 * Tn);
    uint2* avb = ensure<uint2>(ctx, ctx->tvb, A);
    mark(ctx, "emit");
    if (A > 0) {
        hipLaunchKernelGGL(k_tmc_vbase, dim3(grid_for(ctx, A)), dim3(kBlock), 0, ctx->stream, sw.vt, acell, acode, vofs,
                           (int64_t)A, avb);
        check_launch(ctx, "k_tmc_vbase");
        hipLaunchKernelGGL(k_tmc_emit<T>, dim3(grid_for(ctx, A)), dim3(kBlock), 0, ctx->stream, sw.vt, acell, acode, avb,
                           tofs, sw.segMask, sw.segBase, (int64_t)A, pos, nrm, tris, st + 2);
        check_launch(ctx, "k_tmc_emit");
        hipLaunchKernelGGL(k_tmc_emit_amb<T>, dim3(grid_for(ctx, A, kBlock, 1024)), dim3(kBlock), 0, ctx->stream, sw.vt,
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

// Everything after the upload, for one volume or a batch.
void run(dmc_context* ctx, const std::vector<HostVol>& hv, int field_type, uint32_t flags) {
    const int n = (int)hv.size();
    if (field_type == DMC_FIELD_F32) run_cells<float>(ctx, hv);
    else run_cells<double>(ctx, hv);
    if (flags & DMC_SIMPLIFY_3X3Y_COLOR) {
        mark(ctx, "p3x3y_color");
        simplify3x3y(ctx, true);
    }
    if (flags & DMC_SIMPLIFY_3X3Y_ISOLATED) {
        mark(ctx, "p3x3y_isolated");
        simplify3x3y(ctx, false);
    }
    if (flags & DMC_SIMPLIFY_3333) {
        mark(ctx, "p3333");
        simplify3333(ctx);
    }
    mark(ctx, "halfedges");
    const HeStats hs = build_halfedges(ctx, true);
    ctx->mesh.nonmanifold_edges = (int64_t)hs.nonmanifold;
    ctx->mesh.edge_errors = (int64_t)hs.errors;
    ctx->mesh.invalid_refs = (int64_t)hs.bad;
    ctx->mesh.nbnd = (int64_t)ctx->bnd.size();
    ctx->mesh.nt = 0;
    const int64_t nq = ctx->mesh.nq, nv = ctx->mesh.nv;
    gather_vertices(ctx);
    int32_t* tris = nullptr;
    if (flags & DMC_TRIANGLES) {
        mark(ctx, "triangles");
        tris = ensure<int32_t>(ctx, ctx->tris, 6 * nq);
        if (nq > 0)
            hipLaunchKernelGGL(k_triangles, dim3(grid_for(ctx, nq)), dim3(kBlock), 0, ctx->stream,
                               static_cast<int32_t*>(ctx->quads[ctx->cur].p), nq, nv,
                               static_cast<double*>(ctx->pos[ctx->pcur].p), tris);
        check_launch(ctx, "k_triangles");
        ctx->mesh.nt = 2 * nq;
        ctx->mesh.d_tris = tris;
    }
    int32_t* quads = static_cast<int32_t*>(ctx->quads[ctx->cur].p);
    int32_t* twin = static_cast<int32_t*>(ctx->twin.p);
    int32_t* hev = static_cast<int32_t*>(ctx->hev.p);
    if (n > 1 && (nq > 0 || nv > 0)) {
        // each volume's ids start at 0
        hipLaunchKernelGGL(k_rebase, dim3(grid_for(ctx, std::max(nq, nv))), dim3(kBlock), 0, ctx->stream,
                           static_cast<const int64_t*>(ctx->vbound.p), static_cast<const int64_t*>(ctx->qbound.p), n,
                           quads, twin, tris, nq, hev, nv);
        check_launch(ctx, "k_rebase");
    }
    mark(ctx, "end");
    // per-volume bookkeeping back to the host
    std::vector<int64_t> vb(2 * (n + 1)), qb(2 * (n + 1)), ab(n + 1), vhe(n > 1 ? 2 * n : 0);
    std::vector<VolStat> vs(n);
    uint32_t diag[6];
    HostReads r(ctx);
    r.add(vb.data(), ctx->vbound.p, vb.size() * sizeof(int64_t));
    r.add(qb.data(), ctx->qbound.p, qb.size() * sizeof(int64_t));
    r.add(ab.data(), vol_ptr(ctx, 2), ab.size() * sizeof(int64_t));
    r.add(vs.data(), ctx->vstat.p, vs.size() * sizeof(VolStat));
    if (n > 1) r.add(vhe.data(), ctx->vhe.p, vhe.size() * sizeof(unsigned long long));
    r.add(diag, ctx->fails.p, sizeof(diag));
    r.wait();
    check_rows(ctx);
    ctx->mesh.valence_zero = diag[4];
    ctx->mesh.invalid_refs += diag[5];
    finish_timing(ctx);
    ctx->mesh.d_positions = static_cast<double*>(ctx->pos[ctx->pcur].p);
    ctx->mesh.d_normals = static_cast<double*>(ctx->nrm[ctx->pcur].p);
    ctx->mesh.d_quads = quads;
    ctx->mesh.d_twins = twin;
    ctx->mesh.d_he_flags = static_cast<uint8_t*>(ctx->flag.p);
    ctx->mesh.d_he_v = hev;
    uint64_t f1 = 0, f2 = 0, nc = 0;
    for (const VolStat& x : vs) {
        f1 += x.fail1;
        f2 += x.fail2;
        nc += x.contours;
    }
    ctx->mesh.failed_projections_l1 = (uint32_t)f1;
    ctx->mesh.failed_projections_l2 = (uint32_t)f2;
    ctx->mesh.nv_emitted = (int64_t)nc;
    // one view per volume
    ctx->vmesh.assign(n, dmc_mesh{});
    ctx->vbnd.assign(n, std::vector<int32_t>());
    if (n == 1) {
        ctx->vmesh[0] = ctx->mesh;
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
            m.d_positions = ctx->mesh.d_positions + 3 * vb[v];
            m.d_normals = ctx->mesh.d_normals + 3 * vb[v];
            m.d_quads = quads + 4 * qb[v];
            m.d_tris = tris ? tris + 6 * qb[v] : nullptr;
            m.d_twins = twin + 4 * qb[v];
            m.d_he_flags = ctx->mesh.d_he_flags + 4 * qb[v];
            m.d_he_v = hev + vb[v];
            m.nonmanifold_edges = (int64_t)vhe[2 * v];
            m.nbnd = (int64_t)ctx->vbnd[v].size();
            m.edge_errors = (int64_t)vhe[2 * v + 1];
            m.invalid_refs = ctx->mesh.invalid_refs;   // batch-wide
            m.valence_zero = ctx->mesh.valence_zero;   // batch-wide
            m.failed_projections_l1 = vs[v].fail1;
            m.failed_projections_l2 = vs[v].fail2;
            m.active_cells = ab[v + 1] - ab[v];
            m.nv_emitted = (int64_t)vs[v].contours;
            m.nv_cells = vb[n + 1 + v + 1] - vb[n + 1 + v];
            m.nq_cells = qb[n + 1 + v + 1] - qb[n + 1 + v];
        }
    }
    ctx->have_mesh = true;
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
    ctx->bad_epoch = ctx->rows_epoch = -1;
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
    run(ctx, hv, a.field_type, a.flags);
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
    if (const char* w = getenv("TMC_SWEEP_WAVES")) {
        const long v = strtol(w, nullptr, 10);
        if (v >= kSweepWavesMin && v <= kSweepWavesMax) ctx->sweep_waves = (int)v;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return DMC_ERR_HIP;
    }
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
    if (ctx->dpin) (void)hipHostFree(ctx->dpin);
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
    // the halfedge rows and twins are those of the final build
    if (nhe > 0)
        hipLaunchKernelGGL(k_edge_reps, dim3(grid_for(ctx, nhe)), dim3(kBlock), 0, ctx->stream,
                           static_cast<const int32_t*>(m.d_quads), nhe, nv,
                           static_cast<const uint32_t*>(ctx->off.p), static_cast<const int2*>(ctx->adj.p),
                           static_cast<const int32_t*>(m.d_twins), static_cast<const uint8_t*>(m.d_he_flags), rep);
    check_launch(ctx, "k_edge_reps");
    const int64_t ne = (int64_t)scan(ctx, rep, nid, nhe);
    *a.count = ne;
    if (a.edges && ne > 0) {
        int2* e = ensure<int2>(ctx, ctx->dedges, ne);
        hipLaunchKernelGGL(k_edge_emit, dim3(grid_for(ctx, nhe)), dim3(kBlock), 0, ctx->stream,
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
    HIPC(ctx, hipMemsetAsync(hist, 0, (kValenceBins + 1) * sizeof(uint32_t), ctx->stream));
    if (nv > 0) {
        hipLaunchKernelGGL(k_vertex_valence, dim3(grid_for(ctx, nv)), dim3(kBlock), 0, ctx->stream,
                           static_cast<const int32_t*>(m.d_quads), a.triangles ? m.d_tris : nullptr, nv,
                           static_cast<const uint32_t*>(ctx->off.p), static_cast<const int2*>(ctx->adj.p), vval);
        hipLaunchKernelGGL(k_valence_hist, dim3(grid_for(ctx, nv, kBlock, 1024)), dim3(kBlock), 0, ctx->stream, vval, nv,
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
        hipLaunchKernelGGL(k_quality_quads, dim3(grid_for(ctx, n)), dim3(kBlock), 0, ctx->stream,
                           static_cast<const int32_t*>(m.d_quads), n, static_cast<const double*>(m.d_positions), q);
    else
        hipLaunchKernelGGL(k_quality_tris, dim3(grid_for(ctx, n)), dim3(kBlock), 0, ctx->stream,
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
