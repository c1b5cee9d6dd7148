// MI355X (gfx950) curvelet-construction pipeline and its C ABI.
//
// Stages and kernels (one map per call; everything stays in HBM between
// kernels, one HIP stream per context):
//
//   graph   k_prep_edges      row-major edges -> SoA x/y/theta, pixel cells, bbox
//           k_cell_count      counting sort of edges into a dense pixel-cell grid
//           scan              exclusive prefix sum over the grid (cell starts)
//           k_cell_scatter    edges into cell order
//           k_cell_order      within-cell order by edge id; cell-sorted id,
//                             cell and (x, y) arrays
//           k_discover_count  per anchor (a 16-lane segment): in-radius
//                             candidates in the 7x7 window
//           scan              CSR offsets
//           k_fill_csr        per anchor (a 16-lane segment): gather in the
//                             reference's scan order, distance sort
//                             (tie-exact std::sort replay), CSR.  On a
//                             guessed grid and buffer sizes (a map after
//                             another) it counts the lists itself and finds
//                             the offsets by a decoupled look-back, in place
//                             of the counting pass and the scan
//   curvelets k_edge_trig     per edge: sin/cos of theta, theta+-dt, theta+grid
//           k_pair_prep       per pair: direction-filter class, close/branch
//                             flags, the bundle's cell-independent terms
//           k_pair_trig       per (pair, grid angle): the transport's two sines
//                             (at Kt = 7 with k_pair_prep's work per pair first;
//                             queued with k_edge_trig behind the graph stage's
//                             count read where the pair buffers hold the
//                             guessed capacity)
//           k_bucket          anchors binned by their largest run (LDS capacity)
//           k_curvelets       one wave per anchor, bucket by bucket: bundles
//                             (lane = cell) into the wave's LDS slice, chain
//                             growth with ballot validity, dedup, staging rows
//           scan              per-anchor accepted counts -> output row offsets
//           k_row_slots       staging slot -> row, and the row count into the
//                             scalars read back with the stream
//           k_record          chain rows (1-based) and the 10 info columns
//
// The reference path this replaces is cpu_curvelet (preprocess.hpp,
// cpu_curvelet.hpp); each kernel cites the lines it follows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <type_traits>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <tuple>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/curvelet_gpu.h"
#include "cvg_device.hpp"
#include "cvg_math.hpp"
#include "cvg_stlsort.hpp"

using cvg::Cand;
using cvg::PairRec;

__constant__ double c_sincos_tab[CVG_SINCOS_TABLE_SIZE] = CVG_SINCOS_TABLE_INIT;
__constant__ double c_atan_tab[2 * CVG_ATAN_ROWS] = CVG_ATAN_TABLE_INIT;

// Device memory of cvg_run_device_noalloc: every intermediate buffer of a
// no-alloc run is carved out of this array.  Host code never uses the name
// itself as a device address: each device's copy is resolved once with
// hipGetSymbolAddress (scratch_address) and that pointer is what kernels get.
static __device__ __align__(256) unsigned char scratch[static_cast<size_t>(CVG_NOALLOC_SCRATCH_MIB) << 20];

namespace {

constexpr int kBlock = 256;
// k_fill_csr's look-back: polls of one predecessor's state before a
// workgroup gives up and raises S_MISS (the host then builds the graph on the
// counting path); a context created with CVG_LOOKBACK_SPINS set uses that
// limit instead (tests force the fallback with 0)
constexpr unsigned kLookbackSpins = 1u << 20;
constexpr int kInfoWidth = 10;
constexpr long long kMaxGridCells = 1LL << 28;  // larger spans take the sorted-key cells
constexpr double kMaxCoord = 2147475456.0;        // 2^31 - 2^13
// the static scratch array behind cvg_run_device_noalloc (one per device)
constexpr size_t kScratchBytes = static_cast<size_t>(CVG_NOALLOC_SCRATCH_MIB) << 20;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// A dense pixel-cell grid: origin and size in cells.
struct GridDesc {
    int x0, y0, w, h;
};

}  // namespace

struct cvg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // extra streams so the bucket launches of the fused kernel run concurrently
    hipStream_t aux[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
    // timing events: ev[0..2] run start, graph end, run end; ev[3..4] around
    // the fused kernel's launches
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // the fused kernel's start and end events of the last kFusedRing timed runs
    static constexpr int kFusedRing = 64;
    hipEvent_t fused_ev[kFusedRing][2] = {};
    long long fused_runs = 0;
    // events without timing for ordering and host waits: the graph stage's
    // count read, the fork to the auxiliary streams, the end of an untimed run
    hipEvent_t ev_read = nullptr, ev_fork = nullptr, ev_end = nullptr;
    // Runs with stats record the timing events (their stage times come from
    // them); runs without stats only while event_timing is set
    // (cvg_set_event_timing).  Each recorded timing event is a marker that
    // holds the stream for several microseconds.
    bool event_timing = false;
    bool timed_run = false;  // the current run records the timing events
    std::string err;
    size_t lds_limit = 65536;
    // fused-kernel launch shape: buckets whose key is at least coop_min_cap
    // run one anchor per workgroup of coop_waves waves (cvg_set_tuning, or
    // CVG_COOP_MIN_CAP / CVG_COOP_WAVES at creation; results never depend on it)
    int coop_min_cap = 8;
    int coop_waves = 4;
    int num_cu = 1;
    // lanes per anchor in the neighbour-window scans (8, 16, 32 or 64;
    // cvg_set_tuning, or CVG_FILL_LANES at creation; results never depend on it)
    int fill_lanes = 16;
    // grid of the light buckets' launches as a percentage of their resident
    // capacity (5..100), or 0: automatic (cvg_set_tuning, or CVG_GRID_PCT)
    int grid_pct = 0;
    // use the sorted-key cell layout even where the dense grid fits
    // (CVG_SPARSE_CELLS=1 at creation; results never depend on it)
    bool force_sparse = false;
    // run double maps on the general curvelet kernel even where the fused one
    // applies (CVG_GENERIC_KERNEL=1 at creation; same results)
    bool force_generic = false;

    unsigned *mailbox = nullptr;  // pinned host scalars read back after syncs
    // page-locked staging blocks for downloads into pageable caller memory
    void *h_stage[2] = {nullptr, nullptr};
    hipEvent_t ev_stage[2] = {nullptr, nullptr};

    // graph
    int n = 0;
    int sz = 4;
    double rad = 0;
    unsigned nbr_radius = 3;
    unsigned max_cand = 128;
    bool graph_ready = false;
    int pairs = 0;
    unsigned max_deg = 0;
    unsigned truncated = 0;
    unsigned ambiguous = 0;
    int gx0 = 0, gy0 = 0, gw = 0, gh = 0;
    // d_grid's leading cell counts known to be zero (while d_grid is grid_zero_ptr)
    const void *grid_zero_ptr = nullptr;
    long long grid_zero_cells = 0;
    // the grid, window radius and longest list a single map tries first (the
    // last single map's; spec_valid once there is one)
    bool spec_valid = false;
    bool spec_off = false;  // CVG_SPECULATE=0 at creation: every graph reads its sizes first
    unsigned to_host_rows = 0;  // rows of the last cvg_run_to_host (its early copy's size)
    unsigned long long rebuilds = 0;  // runs built again in the reference's within-cell order
    // a run's parameters while its graph is built (fused double maps): the
    // graph stage queues k_edge_trig behind its last host read
    const cvg_params *early_params = nullptr;
    bool edge_trig_queued = false;  // k_edge_trig of the current run is queued
    bool pair_setup_queued = false;  // and its per-pair records and sines
    GridDesc spec_g{};
    unsigned spec_R = 0;
    int spec_fill_cap = 1;
    // launch shapes already set up (prepare_launch)
    std::map<const void *, unsigned> shmem_set;
    std::map<std::tuple<const void *, int, unsigned>, int> occupancy;
    // the bytes last copied to d_tables (while d_tables is tables_ptr)
    const void *tables_ptr = nullptr;
    std::vector<char> tables_blob;
    DevBuf d_edges_up, d_x, d_y, d_th, d_cx, d_cy, d_scal, d_grid, d_tmp_ids, d_sorted, d_sxy, d_scell, d_part, d_deg,
        d_off, d_tiles, d_cand, d_ids, d_d2, d_pair_anchor;
    // the last graph's edge array (device) and, once the reference's
    // within-cell order has been replayed for it, each edge's rank in its cell
    const double *last_edges = nullptr;
    bool use_cellrank = false;
    // the run's map came in as float (the reference's float instantiation):
    // distances, bundles and info in its float arithmetic
    bool f32 = false;
    DevBuf d_cellrank, d_skey, d_sorttmp;

    // the maps of a batch (a single map outside batches): edge id ranges,
    // per-edge map ids, per-map cell boxes and shifts, output row ranges
    int num_maps = 1;
    std::vector<long long> map_eoff;
    std::vector<unsigned> map_rows;
    unsigned *h_maprow = nullptr;  // page-locked
    size_t h_maprow_bytes = 0;
    bool counts_pending = false;
    DevBuf d_map, d_eoff, d_mapbox, d_maprow;

    // curvelets
    bool curvelets_ready = false;
    cvg_params cur_params{};
    unsigned valid_pairs = 0;
    unsigned long long intersections = 0;
    unsigned num_curvelets = 0;
    unsigned width = 0;
    DevBuf d_trig, d_tdd, d_cls, d_prec, d_psin, d_lists, d_tables, d_st_ids, d_st_meta, d_st_k, d_st_rank, d_acc, d_rowoff, d_rowslot, d_gbundle,
        d_chains, d_info;

    // diagnostics
    DevBuf d_dbg;

    // look-back states of the fill's single-pass offsets (speculative graphs):
    // cleared when allocated (while d_fillstat is fillstat_zero_ptr), then
    // told apart by the launch's epoch
    DevBuf d_fillstat;
    unsigned lookback_spins = kLookbackSpins;  // CVG_LOOKBACK_SPINS at creation
    int gate_naps = 40;  // the start gate's bound in naps (CVG_GATE_NAPS at creation)
    // the same for the single-pass scans (k_scan_1pass)
    DevBuf d_scanstat;
    const void *scanstat_zero_ptr = nullptr;
    size_t scanstat_zero_bytes = 0;
    unsigned scan_epoch = 0;
    const void *fillstat_zero_ptr = nullptr;
    size_t fillstat_zero_bytes = 0;
    unsigned fill_epoch = 0;

    // this device's copy of the static scratch array (cvg_run_device_noalloc);
    // while `arena` is set, ensure() carves buffers out of it instead of
    // allocating
    char *scratch_base = nullptr;
    bool arena = false;
    size_t arena_used = 0;
    // caller-owned device outputs of a no-alloc run (null otherwise)
    unsigned *caller_chains = nullptr;
    double *caller_info = nullptr;
    unsigned caller_rows = 0;

    // every device pool of the context, for release and for the no-alloc swap
    template <typename F>
    void for_each_pool(F f)
    {
        DevBuf *bufs[] = {&d_edges_up, &d_x,      &d_y,      &d_th,     &d_cx,     &d_cy,          &d_scal,
                          &d_grid,     &d_tmp_ids, &d_sorted, &d_sxy,    &d_scell,  &d_part,        &d_deg,
                          &d_off,      &d_tiles,  &d_cand,   &d_ids,    &d_d2,     &d_pair_anchor, &d_trig,
                          &d_tdd,      &d_cls,    &d_prec,   &d_psin,   &d_lists,  &d_tables,      &d_st_ids,
                          &d_st_meta,  &d_st_k,   &d_st_rank, &d_acc,   &d_rowoff, &d_rowslot, &d_gbundle, &d_chains,
                          &d_info,     &d_dbg,    &d_map,    &d_eoff,   &d_mapbox, &d_maprow,       &d_cellrank,
                          &d_skey,     &d_sorttmp, &d_fillstat, &d_scanstat};
        for (DevBuf *b : bufs) f(*b);
    }
};

namespace {

int fail(cvg_context *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define CVG_HIP(ctx, call)                                                                               \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail((ctx), CVG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define CVG_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != CVG_OK) return rc_; \
    } while (0)

// Records timing event e on st in a timed run; nothing otherwise.
int record_timing(cvg_context *ctx, hipEvent_t e, hipStream_t st)
{
    if (ctx->timed_run) CVG_HIP(ctx, hipEventRecord(e, st));
    return CVG_OK;
}

// The event that marks the end of the current run's work on the context
// stream: the timing event ev[2] in a timed run, ev_end otherwise.
hipEvent_t end_event(const cvg_context *ctx)
{
    return ctx->timed_run ? ctx->ev[2] : ctx->ev_end;
}

int ensure(cvg_context *ctx, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return CVG_OK;
    if (ctx->arena) {
        // no-alloc mode: carve from the static scratch array, 256-byte aligned
        const size_t need = (bytes + 255) & ~static_cast<size_t>(255);
        if (ctx->arena_used + need > kScratchBytes)
            return fail(ctx, CVG_ERR_UNSUPPORTED,
                        "cvg_run_device_noalloc: this map needs more than the " +
                            std::to_string(kScratchBytes >> 20) +
                            " MiB static scratch array (use cvg_run_device for larger maps)");
        b.ptr = ctx->scratch_base + ctx->arena_used;
        b.bytes = bytes;
        ctx->arena_used += need;
        return CVG_OK;
    }
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    const size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return fail(ctx, CVG_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    b.bytes = want;
    return CVG_OK;
}

void release(DevBuf &b)
{
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

template <typename T>
T *as(DevBuf &b)
{
    return static_cast<T *>(b.ptr);
}

inline int blocks_for(long long n, int block = kBlock)
{
    return static_cast<int>((n + block - 1) / block);
}

// ---------------------------------------------------------------------------
// Exclusive prefix sum over uint32 (in place allowed), out[n] = total.
// ---------------------------------------------------------------------------

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ unsigned block_exclusive_scan(unsigned v, unsigned *lds_waves, unsigned *total)
{
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_waves[wid] = x;
    __syncthreads();
    unsigned pre = 0, tot = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        const unsigned t = lds_waves[w];
        if (w < wid) pre += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Workgroup max and sum of (max, count) pairs; the result is valid in thread 0.
__device__ void block_max_sum(unsigned &mx, unsigned &cnt, unsigned *lds_max, unsigned *lds_cnt)
{
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, static_cast<unsigned>(__shfl_xor(static_cast<int>(mx), o, 64)));
        cnt += static_cast<unsigned>(__shfl_xor(static_cast<int>(cnt), o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        lds_max[threadIdx.x >> 6] = mx;
        lds_cnt[threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
            mx = max(mx, lds_max[w]);
            cnt += lds_cnt[w];
        }
    }
    __syncthreads();
}

// Tile sums; with STATS also each tile's largest value and its count of values
// >= thr (tile_max / tile_cnt), so a scan over counts also yields their
// maximum and how many reached a cap, without contended atomics.
// A thread's kScanItems consecutive inputs (zeros past n): two 16-byte loads
// when the run lies inside the array and the array is 16-byte aligned.
__device__ inline void load_items(const unsigned *in, long long base, long long n, unsigned (&v)[kScanItems])
{
    static_assert(kScanItems == 8, "two 16-byte loads per thread");
    if (base + kScanItems <= n && (reinterpret_cast<size_t>(in) & 15) == 0) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
        const uint4 a = p[0], b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
}

template <bool STATS>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const unsigned *__restrict__ in, long long n,
                                                       unsigned *__restrict__ tile_sums, unsigned thr,
                                                       unsigned *__restrict__ tile_max, unsigned *__restrict__ tile_cnt)
{
    __shared__ unsigned lw[kBlock / 64];
    __shared__ unsigned lm[kBlock / 64], lc[kBlock / 64];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    unsigned s = 0, mx = 0, c = 0;
    unsigned v[kScanItems];
    load_items(in, base, n, v);
    for (int k = 0; k < kScanItems; ++k) {
        s += v[k];
        if constexpr (STATS) {
            // items past n are zeros, which leave the maximum as it is
            mx = max(mx, v[k]);
            c += (base + k < n && v[k] >= thr) ? 1u : 0u;
        }
    }
    unsigned tot;
    block_exclusive_scan(s, lw, &tot);
    if constexpr (STATS) block_max_sum(mx, c, lm, lc);
    if (threadIdx.x == 0) {
        tile_sums[blockIdx.x] = tot;
        if constexpr (STATS) {
            tile_max[blockIdx.x] = mx;
            tile_cnt[blockIdx.x] = c;
        }
    }
}

// Single workgroup: exclusive scan of the tile sums in place, total at
// [ntiles]; with STATS also the maximum and count over the tiles into
// *stat_max and *stat_cnt.
template <bool STATS>
__global__ __launch_bounds__(1024) void k_scan_tile_sums(unsigned *tile_sums, int ntiles,
                                                         const unsigned *__restrict__ tile_max,
                                                         const unsigned *__restrict__ tile_cnt,
                                                         unsigned *__restrict__ stat_max,
                                                         unsigned *__restrict__ stat_cnt,
                                                         unsigned *__restrict__ stat_total)
{
    __shared__ unsigned lw[1024 / 64];
    __shared__ unsigned lm[1024 / 64], lc[1024 / 64];
    unsigned carry = 0, mx = 0, c = 0;
    for (int base = 0; base < ntiles; base += 1024) {
        const int i = base + threadIdx.x;
        const unsigned v = i < ntiles ? tile_sums[i] : 0u;
        if constexpr (STATS) {
            if (i < ntiles) {
                mx = max(mx, tile_max[i]);
                c += tile_cnt[i];
            }
        }
        unsigned tot;
        const unsigned ex = block_exclusive_scan(v, lw, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if constexpr (STATS) block_max_sum(mx, c, lm, lc);
    if (threadIdx.x == 0) {
        tile_sums[ntiles] = carry;
        if constexpr (STATS) {
            *stat_max = mx;
            *stat_cnt = c;
            if (stat_total) *stat_total = carry;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const unsigned *in, long long n, const unsigned *tile_prefix,
                                                       int ntiles, unsigned *out)
{
    __shared__ unsigned lw[kBlock / 64];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    unsigned v[kScanItems];
    load_items(in, base, n, v);
    unsigned s = 0;
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    unsigned tot;
    unsigned run = block_exclusive_scan(s, lw, &tot) + tile_prefix[blockIdx.x];
    unsigned o[kScanItems];
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (base + kScanItems <= n && (reinterpret_cast<size_t>(out) & 15) == 0) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
        p[0] = make_uint4(o[0], o[1], o[2], o[3]);
        p[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_prefix[ntiles];
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor(static_cast<unsigned>(v), o, 64);
        const unsigned hi = __shfl_xor(static_cast<unsigned>(v >> 32), o, 64);
        v += (static_cast<unsigned long long>(hi) << 32) | lo;
    }
    return v;
}

// Single-pass exclusive prefix sum (out[i] = in[0] + ... + in[i-1], out[n] =
// total, *total too when given; in and out distinct) with a decoupled
// look-back over the tiles: each tile publishes its own sum, then its
// inclusive prefix, as one word flag (1: own sum, 2: inclusive) << 62 |
// epoch << 32 | sum, with relaxed device-scope atomics (no fence, which
// would write back or invalidate the whole L2 of the XCD).  The epoch, one
// per launch, makes words of earlier launches read as not ready.  Every
// predecessor tile is dispatched before its successor, so every wait ends; a
// tile that still sees no word after max_spins polls sums the inputs before
// it directly, so the result is exact either way.
__device__ inline unsigned scan_lookback(unsigned long long *st, unsigned epoch, int tile, unsigned agg,
                                         unsigned max_spins, const unsigned *__restrict__ in)
{
    const int lane = threadIdx.x & 63;
    if (lane == 0 && tile > 0)
        __hip_atomic_store(&st[tile], 1ULL << 62 | static_cast<unsigned long long>(epoch) << 32 | agg,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned tag_a = 1u << 30 | epoch, tag_p = 2u << 30 | epoch;
    unsigned long long sum = 0;
    bool gave_up = false;
    for (int base = tile - 1; base >= 0; base -= 64) {
        const int j = base - lane;
        unsigned tag = tag_p;  // before the first tile: an empty inclusive prefix
        unsigned long long w = 0;
        if (j >= 0) {
            w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            while (static_cast<unsigned>(w >> 32) != tag_a && static_cast<unsigned>(w >> 32) != tag_p) {
                if (++spins > max_spins) {
                    gave_up = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            tag = static_cast<unsigned>(w >> 32);
        }
        if (__ballot(gave_up) != 0ULL) break;
        const unsigned long long pm = __ballot(tag == tag_p);
        const int stop = pm != 0ULL ? __ffsll(static_cast<long long>(pm)) - 1 : 64;  // nearest inclusive prefix
        sum += wave_sum_u64(lane <= stop ? (w & 0xffffffffULL) : 0ULL);
        if (pm != 0ULL) break;
    }
    if (__ballot(gave_up) != 0ULL) {
        sum = 0;
        for (long long i = lane; i < static_cast<long long>(tile) * kScanTile; i += 64) sum += in[i];
        sum = wave_sum_u64(sum);
    }
    const unsigned excl = static_cast<unsigned>(sum);
    if (lane == 0)
        __hip_atomic_store(&st[tile], 2ULL << 62 | static_cast<unsigned long long>(epoch) << 32 | (excl + agg),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

__global__ __launch_bounds__(kBlock) void k_scan_1pass(const unsigned *__restrict__ in, long long n,
                                                       unsigned *__restrict__ out, unsigned *__restrict__ total,
                                                       unsigned long long *__restrict__ lb, unsigned epoch,
                                                       unsigned max_spins)
{
    __shared__ unsigned lw[kBlock / 64];
    __shared__ unsigned s_base;
    const int tile = static_cast<int>(blockIdx.x);
    const long long base = static_cast<long long>(tile) * kScanTile + threadIdx.x * kScanItems;
    unsigned v[kScanItems];
    load_items(in, base, n, v);
    unsigned s = 0;
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    unsigned agg;
    const unsigned ex = block_exclusive_scan(s, lw, &agg);
    if (threadIdx.x < 64) {
        const unsigned excl = scan_lookback(lb, epoch, tile, agg, max_spins, in);
        if (threadIdx.x == 0) s_base = excl;
    }
    __syncthreads();
    unsigned run = s_base + ex;
    unsigned o[kScanItems];
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (base + kScanItems <= n && (reinterpret_cast<size_t>(out) & 15) == 0) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
        p[0] = make_uint4(o[0], o[1], o[2], o[3]);
        p[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
    if (tile == static_cast<int>(gridDim.x) - 1 && threadIdx.x == 0) {
        out[n] = s_base + agg;
        if (total) *total = s_base + agg;
    }
}

// out has n + 1 entries (the tile buffer is sized here).  With stat_max set,
// the maximum of the inputs goes to *stat_max and the number of inputs >= thr
// to *stat_cnt.
int exclusive_scan(cvg_context *ctx, const unsigned *in, long long n, unsigned *out, DevBuf &tiles,
                   unsigned *stat_max = nullptr, unsigned *stat_cnt = nullptr, unsigned thr = 0,
                   unsigned *stat_total = nullptr)
{
    const int ntiles = static_cast<int>((n + kScanTile - 1) / kScanTile);
    CVG_TRY(ensure(ctx, tiles, static_cast<size_t>(3) * (ntiles + 1) * sizeof(unsigned)));
    unsigned *t = as<unsigned>(tiles);
    unsigned *tmax = t + (ntiles + 1), *tcnt = t + 2 * (ntiles + 1);
    if (stat_max) {
        if (ntiles > 0)
            hipLaunchKernelGGL(k_scan_tiles<true>, dim3(ntiles), dim3(kBlock), 0, ctx->stream, in, n, t, thr, tmax,
                               tcnt);
        hipLaunchKernelGGL(k_scan_tile_sums<true>, dim3(1), dim3(1024), 0, ctx->stream, t, ntiles, tmax, tcnt,
                           stat_max, stat_cnt, stat_total);
    } else {
        if (ntiles > 0)
            hipLaunchKernelGGL(k_scan_tiles<false>, dim3(ntiles), dim3(kBlock), 0, ctx->stream, in, n, t, thr, tmax,
                               tcnt);
        hipLaunchKernelGGL(k_scan_tile_sums<false>, dim3(1), dim3(1024), 0, ctx->stream, t, ntiles, tmax, tcnt,
                           stat_max, stat_cnt, static_cast<unsigned *>(nullptr));
    }
    if (ntiles > 0)
        hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kBlock), 0, ctx->stream, in, n, t, ntiles, out);
    else
        CVG_HIP(ctx, hipMemcpyAsync(out, t, sizeof(unsigned), hipMemcpyDeviceToDevice, ctx->stream));
    CVG_HIP(ctx, hipGetLastError());
    return CVG_OK;
}

// The single-pass scan (k_scan_1pass) for up to kScan1PassMaxTiles tiles,
// the three-kernel scan above for more: out has n + 1 entries, in and out
// are distinct.  One launch instead of three saves ~8 us on a scan of ~100
// tiles (the row offsets of a 1080p map).  When ~1,000 tiles all start at
// once (the cell counts of a 1080p grid), every tile reads back over many
// tiles that have only their own sums yet, and the single pass is slower
// (17.8 against ~14 us).
constexpr long long kScan1PassMaxTiles = 256;

int exclusive_scan_1pass(cvg_context *ctx, const unsigned *in, long long n, unsigned *out, DevBuf &tiles)
{
    const long long ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles > kScan1PassMaxTiles || ntiles == 0) return exclusive_scan(ctx, in, n, out, tiles);
    CVG_TRY(ensure(ctx, ctx->d_scanstat, static_cast<size_t>(ntiles) * sizeof(unsigned long long)));
    if (++ctx->scan_epoch >= (1u << 30)) ctx->scan_epoch = 1;  // words hold epoch << 32 below the 2-bit flag
    if (ctx->d_scanstat.ptr != ctx->scanstat_zero_ptr || ctx->d_scanstat.bytes != ctx->scanstat_zero_bytes ||
        ctx->scan_epoch == 1) {
        CVG_HIP(ctx, hipMemsetAsync(ctx->d_scanstat.ptr, 0, ctx->d_scanstat.bytes, ctx->stream));
        ctx->scanstat_zero_ptr = ctx->arena ? nullptr : ctx->d_scanstat.ptr;
        ctx->scanstat_zero_bytes = ctx->d_scanstat.bytes;
    }
    hipLaunchKernelGGL(k_scan_1pass, dim3(static_cast<unsigned>(ntiles)), dim3(kBlock), 0, ctx->stream, in, n, out,
                       static_cast<unsigned *>(nullptr), as<unsigned long long>(ctx->d_scanstat), ctx->scan_epoch,
                       ctx->lookback_spins);
    CVG_HIP(ctx, hipGetLastError());
    return CVG_OK;
}

// ---------------------------------------------------------------------------
// Graph stage: build_neighbor_csr_graph (preprocess.hpp:62-274)
// ---------------------------------------------------------------------------


// Where the edges of a window row lie in the cell-sorted arrays.  The dense
// grid keeps every cell's start (one load per row end); maps whose cells span
// more than the dense grid allows keep the edges' sorted linear cell keys
// instead, and a row's ends are binary searches over them.  Both give the
// same sorted positions, so the scans below are the same for either.
struct DenseCells {
    const unsigned *start;
    int x0, y0, w;
    // edges of cells (cx - R .. cx + R, py), py absolute
    __device__ void row(int cx, int py, int R, int &rs, int &len) const
    {
        const int rr = (py - y0) * w + (cx - x0);
        rs = static_cast<int>(start[rr - R]);
        len = static_cast<int>(start[rr + R + 1]) - rs;
    }
};

struct SparseCells {
    // cell key ((cy - y0) << 32) + (cx - x0) of each sorted position, ascending
    // (row-major; both offsets are below 2^32)
    const unsigned long long *key;
    int n;
    int x0, y0;
    __device__ int lower(unsigned long long k) const
    {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = lo + ((hi - lo) >> 1);
            if (key[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    }
    __device__ void row(int cx, int py, int R, int &rs, int &len) const
    {
        // cx - R - x0 >= 0; cx + R + 1 - x0 <= 2^32 carries into the next
        // row's first key, which is still this row's end
        const unsigned long long b = (static_cast<unsigned long long>(static_cast<long long>(py) - y0) << 32) +
                                     static_cast<unsigned long long>(static_cast<long long>(cx) - x0);
        rs = lower(b - static_cast<unsigned>(R));
        len = lower(b + static_cast<unsigned>(R) + 1ULL) - rs;
    }
};

// Scalars block layout (d_scal, uint32): [0..3] bbox (minx, maxx, miny, maxy as
// int), [4] bad-input flag, [5] max degree, [6] truncated anchors,
// [7] order-ambiguous anchors, [8] valid pairs, [9] anchors needing the
// std::sort replay, [10..11] the 64-bit count of bundle intersections, [12]
// CSR pairs, [13] a size guess that did not hold, [14] output rows.  The
// fused kernel's bucket list lengths follow, one 128-byte line each (atomics
// from many workgroups on one line serialize in L2).
enum {
    S_MINX = 0,
    S_MAXX = 1,
    S_MINY = 2,
    S_MAXY = 3,
    S_BAD = 4,
    S_MAXDEG = 5,
    S_TRUNC = 6,
    S_AMBIG = 7,
    S_VALID = 8,
    S_TIES = 9,
    S_COUNT = 10,
    S_PAIRS = 12,
    S_MISS = 13,
    S_ROWS = 14
};
constexpr int kScalarWords = S_ROWS + 1;  // the words read back to the host
constexpr int kMaxBuckets = 14;
constexpr int kBucketStride = 32;  // words
constexpr int kBucketBase = 32;    // bucket b's list length at [kBucketBase + b * kBucketStride]
// heavy-bucket anchor streams that have started (own 128-byte line)
constexpr int kStartedWord = kBucketBase + kMaxBuckets * kBucketStride;
constexpr int kScalarBufWords = kStartedWord + 32;

// Cells are (lrint(x), lrint(y)) (preprocess.hpp:85-86).  Coordinates must be
// finite with |x|, |y| < 2^31 - 2^13 (so that every window cell cx +- R of the
// reference's int arithmetic exists), and orientations finite; others raise
// S_BAD.
// Each workgroup leaves its cell box and bad flag in part[5 * block]
// (k_bbox_finish reduces them).
// With cnt (a map on the guessed grid g): each edge whose window (cell +- R)
// lies inside g is also counted into its cell (k_cell_count's work); an
// edge outside it makes the bounding box miss the grid, so the counts are
// never used and the host rebuilds them.
__global__ __launch_bounds__(kBlock) void k_prep_edges(int n, int sz, const double *__restrict__ e,
                                                       double *__restrict__ x, double *__restrict__ y,
                                                       double *__restrict__ th, int *__restrict__ cx,
                                                       int *__restrict__ cy, int *__restrict__ part, GridDesc g,
                                                       unsigned *__restrict__ cnt, int R)
{
    __shared__ int smin_x, smax_x, smin_y, smax_y, sbad;
    if (threadIdx.x == 0) {
        smin_x = INT_MAX;
        smax_x = INT_MIN;
        smin_y = INT_MAX;
        smax_y = INT_MIN;
        sbad = 0;
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double xv = e[static_cast<long long>(i) * sz];
        const double yv = e[static_cast<long long>(i) * sz + 1];
        const double tv = e[static_cast<long long>(i) * sz + 2];
        const bool ok = fabs(xv) < kMaxCoord && fabs(yv) < kMaxCoord && fabs(tv) <= DBL_MAX;
        if (!ok) atomicOr(&sbad, 1);
        const int ix = ok ? static_cast<int>(rint(xv)) : 0;
        const int iy = ok ? static_cast<int>(rint(yv)) : 0;
        x[i] = xv;
        y[i] = yv;
        th[i] = tv;
        cx[i] = ix;
        cy[i] = iy;
        if (cnt && ok && ix >= g.x0 + R && ix <= g.x0 + g.w - 1 - R && iy >= g.y0 + R && iy <= g.y0 + g.h - 1 - R)
            atomicAdd(&cnt[(iy - g.y0) * g.w + (ix - g.x0)], 1u);
    }
    // wave min/max, then one LDS atomic per wave
    int mnx = i < n ? cx[i] : INT_MAX, mxx = i < n ? cx[i] : INT_MIN;
    int mny = i < n ? cy[i] : INT_MAX, mxy = i < n ? cy[i] : INT_MIN;
    for (int o = 32; o > 0; o >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o, 64));
        mxx = max(mxx, __shfl_xor(mxx, o, 64));
        mny = min(mny, __shfl_xor(mny, o, 64));
        mxy = max(mxy, __shfl_xor(mxy, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&smin_x, mnx);
        atomicMax(&smax_x, mxx);
        atomicMin(&smin_y, mny);
        atomicMax(&smax_y, mxy);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int *o = part + 5 * static_cast<long long>(blockIdx.x);
        o[0] = smin_x;
        o[1] = smax_x;
        o[2] = smin_y;
        o[3] = smax_y;
        o[4] = sbad;
    }
}

// One workgroup: the cell box and bad flag over k_prep_edges' nparts partials
// into the scalars (a tree, not same-address atomics, which serialize).
// With spec_on, the map is about to be built on the guessed grid `spec`:
// S_MISS is raised when some edge's window (cell +- R) leaves it, and the
// speculative kernels then return at once.
__global__ __launch_bounds__(1024) void k_bbox_finish(int nparts, const int *__restrict__ part,
                                                      unsigned *__restrict__ scal, int spec_on, GridDesc spec, int R)
{
    __shared__ int l[5][1024 / 64];
    int v[5] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN, 0};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        const int *p = part + 5 * static_cast<long long>(i);
        v[0] = min(v[0], p[0]);
        v[1] = max(v[1], p[1]);
        v[2] = min(v[2], p[2]);
        v[3] = max(v[3], p[3]);
        v[4] |= p[4];
    }
    for (int o = 32; o > 0; o >>= 1) {
        v[0] = min(v[0], __shfl_xor(v[0], o, 64));
        v[1] = max(v[1], __shfl_xor(v[1], o, 64));
        v[2] = min(v[2], __shfl_xor(v[2], o, 64));
        v[3] = max(v[3], __shfl_xor(v[3], o, 64));
        v[4] |= __shfl_xor(v[4], o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) l[k][w] = v[k];
    __syncthreads();
    // every other scalar word starts at zero (this is the run's first write to them)
    if (threadIdx.x > S_BAD && threadIdx.x < kScalarWords) scal[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        for (int u = 1; u < static_cast<int>(blockDim.x >> 6); ++u) {
            v[0] = min(v[0], l[0][u]);
            v[1] = max(v[1], l[1][u]);
            v[2] = min(v[2], l[2][u]);
            v[3] = max(v[3], l[3][u]);
            v[4] |= l[4][u];
        }
        scal[S_MINX] = static_cast<unsigned>(v[0]);
        scal[S_MAXX] = static_cast<unsigned>(v[1]);
        scal[S_MINY] = static_cast<unsigned>(v[2]);
        scal[S_MAXY] = static_cast<unsigned>(v[3]);
        scal[S_BAD] = v[4] ? 1u : 0u;
        if (spec_on) {
            const bool fits = static_cast<long long>(v[0]) - R >= spec.x0 &&
                              static_cast<long long>(v[1]) + R <= static_cast<long long>(spec.x0) + spec.w - 1 &&
                              static_cast<long long>(v[2]) - R >= spec.y0 &&
                              static_cast<long long>(v[3]) + R <= static_cast<long long>(spec.y0) + spec.h - 1;
            if (!fits || v[4]) scal[S_MISS] = 1u;
        }
    }
}

__device__ inline int cell_index(const GridDesc &g, int cx, int cy) { return (cy - g.y0) * g.w + (cx - g.x0); }

// Batches: edge i belongs to map m with eoff[m] <= i < eoff[m+1].
__global__ __launch_bounds__(kBlock) void k_map_ids(int n, int num_maps, const long long *__restrict__ eoff,
                                                    int *__restrict__ map_of)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = num_maps - 1;  // largest m with eoff[m] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (eoff[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    map_of[i] = lo;
}

// Per-map cell boxes (minx, maxx, miny, maxy); box must start at
// (INT_MAX, INT_MIN, INT_MAX, INT_MIN).  A wave whose edges share one map
// reduces first and updates once.
__global__ __launch_bounds__(kBlock) void k_map_box(int n, const int *__restrict__ map_of,
                                                    const int *__restrict__ cx, const int *__restrict__ cy,
                                                    int *__restrict__ box)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    const int m = in ? map_of[i] : -1;
    int mnx = in ? cx[i] : INT_MAX, mxx = in ? cx[i] : INT_MIN;
    int mny = in ? cy[i] : INT_MAX, mxy = in ? cy[i] : INT_MIN;
    const int m0 = __shfl(m, 0, 64);
    if (__ballot(in && m != m0) == 0ULL) {
        for (int o = 32; o > 0; o >>= 1) {
            mnx = min(mnx, __shfl_xor(mnx, o, 64));
            mxx = max(mxx, __shfl_xor(mxx, o, 64));
            mny = min(mny, __shfl_xor(mny, o, 64));
            mxy = max(mxy, __shfl_xor(mxy, o, 64));
        }
        if ((threadIdx.x & 63) == 0 && m0 >= 0) {
            atomicMin(&box[4 * m0], mnx);
            atomicMax(&box[4 * m0 + 1], mxx);
            atomicMin(&box[4 * m0 + 2], mny);
            atomicMax(&box[4 * m0 + 3], mxy);
        }
    } else if (in) {
        atomicMin(&box[4 * m], mnx);
        atomicMax(&box[4 * m + 1], mxx);
        atomicMin(&box[4 * m + 2], mny);
        atomicMax(&box[4 * m + 3], mxy);
    }
}

__global__ void k_map_box_init(int num_maps, int *box)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * num_maps) box[i] = (i & 1) ? INT_MIN : INT_MAX;
}

// Each map's cells move into its own region of one dense grid: only the
// integer cell coordinates shift; x, y and theta are untouched, so every
// distance and bundle is computed from the map's own values, and the regions
// are separated by more than the window radius, so no window reaches another map.
__global__ __launch_bounds__(kBlock) void k_shift_cells(int n, const int *__restrict__ map_of,
                                                        const int *__restrict__ shift, int *__restrict__ cx,
                                                        int *__restrict__ cy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = map_of[i];
    cx[i] += shift[2 * m];
    cy[i] += shift[2 * m + 1];
}

// First output row of every map (rows are anchor-major, so map-major).
__global__ void k_map_rows(int num_maps, const long long *__restrict__ eoff, const unsigned *__restrict__ row_off,
                           unsigned *__restrict__ rows)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m <= num_maps) rows[m] = row_off[eoff[m]];
}

// With miss set, the grid is a guess (the last map's): an edge whose window
// (its cell +- R) leaves the grid is moved inside it, so that every later
// access stays in bounds, and *miss is raised; the host then builds the graph
// again on the map's own grid.
__global__ __launch_bounds__(kBlock) void k_cell_count(int n, int *__restrict__ cx, int *__restrict__ cy, GridDesc g,
                                                       unsigned *__restrict__ cnt, int R, unsigned *__restrict__ miss)
{
    // the bounding box already showed the guessed grid does not hold: count
    // nothing (k_cell_scatter then moves nothing, and the counts stay zero)
    if (miss && *miss) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false;
    if (i < n) {
        int x = cx[i], y = cy[i];
        if (miss) {
            const int lox = g.x0 + R, hix = g.x0 + g.w - 1 - R, loy = g.y0 + R, hiy = g.y0 + g.h - 1 - R;
            out = x < lox || x > hix || y < loy || y > hiy;
            if (out) {
                x = min(max(x, lox), hix);
                y = min(max(y, loy), hiy);
                cx[i] = x;
                cy[i] = y;
            }
        }
        atomicAdd(&cnt[cell_index(g, x, y)], 1u);
    }
    if (__ballot(out) != 0ULL && (threadIdx.x & 63) == 0) atomicOr(miss, 1u);
}

// cnt holds the per-cell counts on entry and is decremented to zero.
__global__ __launch_bounds__(kBlock) void k_cell_scatter(int n, const int *__restrict__ cx, const int *__restrict__ cy,
                                                         GridDesc g, const unsigned *__restrict__ start,
                                                         unsigned *__restrict__ cnt, int *__restrict__ tmp,
                                                         const unsigned *__restrict__ abort)
{
    if (abort && *abort) return;  // k_cell_count counted nothing either
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int c = cell_index(g, cx[i], cy[i]);
        const unsigned pos = start[c] + atomicSub(&cnt[c], 1u) - 1u;
        tmp[pos] = i;
    }
}

// Within-cell order: ascending edge id, or the given per-edge rank.  The
// reference's order is whatever its unstable std::sort of the index
// permutation yields (preprocess.hpp:90-92); the two agree for maps of up to
// 16 edges, and the order only matters for the anchors k_fill_csr counts as
// order-ambiguous, for which the host replays that sort and passes the
// resulting ranks.  The cell-sorted arrays are the edge
// id, its cell index and its (x, y) pair, so one window candidate costs one
// int, one int and one 16-byte load, all independent.
__global__ __launch_bounds__(kBlock) void k_cell_order(int n, const int *__restrict__ cx, const int *__restrict__ cy,
                                                       GridDesc g, const unsigned *__restrict__ start,
                                                       const int *__restrict__ tmp, const int *__restrict__ rank,
                                                       const double *__restrict__ x, const double *__restrict__ y,
                                                       int *__restrict__ sorted, int *__restrict__ scell,
                                                       double2 *__restrict__ sxy, const unsigned *__restrict__ abort)
{
    // a guessed grid that does not hold: the build is repeated, skip the work
    if (abort && *abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_index(g, cx[i], cy[i]);
    const unsigned s = start[c], e = start[c + 1];
    unsigned r = 0;
    if (rank) {
        // the reference's own within-cell order, replayed on the host
        r = min(static_cast<unsigned>(rank[i]), e - s - 1u);
    } else {
        for (unsigned k = s; k < e; ++k) r += (tmp[k] < i) ? 1u : 0u;
    }
    const unsigned p = s + r;
    sorted[p] = i;
    scell[p] = c;
    sxy[p] = make_double2(x[i], y[i]);
}

// Maps whose cells span more than the dense grid allows: every edge's cell
// key in the bounding box (row offset in the high 32 bits, column offset in
// the low ones, so keys order row-major), and its id (or, when the
// reference's within-cell order is replayed, its rank in its cell) for the
// stable sorts that follow.
__global__ __launch_bounds__(kBlock) void k_sparse_keys(int n, const int *__restrict__ cx, const int *__restrict__ cy,
                                                        int x0, int y0, const int *__restrict__ rank,
                                                        unsigned long long *__restrict__ key, int *__restrict__ rkey,
                                                        int *__restrict__ id)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (static_cast<unsigned long long>(static_cast<long long>(cy[i]) - y0) << 32) +
             static_cast<unsigned long long>(static_cast<long long>(cx[i]) - x0);
    if (rank) rkey[i] = rank[i];
    id[i] = i;
}

// The keys of edges already ordered by their within-cell rank.
__global__ __launch_bounds__(kBlock) void k_gather_keys(int n, const unsigned long long *__restrict__ key,
                                                        const int *__restrict__ id, unsigned long long *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = key[id[i]];
}

// The cell-sorted arrays from the sorted (key, id) pairs; a cell's identity
// is its first sorted position.
__global__ __launch_bounds__(kBlock) void k_sparse_fill(int n, const unsigned long long *__restrict__ skey,
                                                        const int *__restrict__ sid, const double *__restrict__ x,
                                                        const double *__restrict__ y, int *__restrict__ sorted,
                                                        int *__restrict__ scell, double2 *__restrict__ sxy)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long k = skey[p];
    int lo = 0, hi = p;
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (skey[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    const int i = sid[p];
    sorted[p] = i;
    scell[p] = lo;
    sxy[p] = make_double2(x[i], y[i]);
}

// sq_dist<T> (curvelet_utils.hpp:13-17) of edge coordinates held as doubles:
// in double, or for maps read as float (the float instantiation, --float) in
// float arithmetic on the (exact) float values.
__device__ inline double pair_dist2(double ax, double ay, double px, double py, bool f32)
{
    if (f32) {
        const float dx = static_cast<float>(ax) - static_cast<float>(px);
        const float dy = static_cast<float>(ay) - static_cast<float>(py);
        return static_cast<double>(dx * dx + dy * dy);
    }
    return (ax - px) * (ax - px) + (ay - py) * (ay - py);
}

// LDS or global writes by some lanes made visible to the other lanes of the
// same wave.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Broadcast a lane's value to the whole wave (lane index uniform).
__device__ inline double bcast(double v, int lane)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(u), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(u >> 32), lane);
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ inline int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// The neighbour-window scans give each anchor a segment of G lanes (64 / G
// anchors per wave).  The window is rows py = cy-R..cy+R; in each row the cells
// cx-R..cx+R are one contiguous range of the cell-sorted arrays
// (preprocess.hpp:141-150).  Rows rb .. rb+G-1 are loaded at once, lane s of
// the segment holding row rb+s: its first sorted position `rs` and the
// exclusive prefix `pre` of the row lengths; `tot` is the rows' total.  Row
// loads are independent, so the whole window costs one load latency.
template <int G, class Cells>
__device__ inline void window_rows(const Cells &cells, int cx, int cy, int R, int rb, int sl, bool valid, int &rs,
                                   int &pre, int &tot)
{
    const int r = rb + sl;
    int len = 0;
    rs = 0;
    if (valid && r <= 2 * R) cells.row(cx, cy - R + r, R, rs, len);
    int inc = len;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        const int v = __shfl_up(inc, o, G);
        if (sl >= o) inc += v;
    }
    pre = inc - len;
    tot = __shfl(inc, G - 1, G);
}

// Sorted position of the t-th candidate (0 <= t < tot) of rows rb..rb+nrows-1,
// or -1.  All lanes must call it (nrows is wave-uniform).
template <int G>
__device__ inline int window_pos(int rs, int pre, int tot, int nrows, int t)
{
    int k = -1;
#pragma unroll 8
    for (int j = 0; j < nrows; ++j) {
        const int pj = __shfl(pre, j, G);
        const int sj = __shfl(rs, j, G);
        if (t >= pj) k = sj + (t - pj);
    }
    return t < tot ? k : -1;
}

template <int G>
__device__ inline unsigned long long segment_mask(int lane)
{
    if constexpr (G == 64) return ~0ULL;
    else return ((1ULL << G) - 1ULL) << (lane & ~(G - 1));
}

// Largest value of a segment-uniform quantity over the wave's segments.
template <int G>
__device__ inline int wave_max_over_segments(int v)
{
    int m = bcast(v, 0);
#pragma unroll
    for (int s = 1; s < 64 / G; ++s) m = max(m, bcast(v, s * G));
    return m;
}

// discover_neighbors_csr counting pass (preprocess.hpp:114-184): in-radius
// candidates in the reference's form `!(d2 > rad2)`, the anchor itself
// excluded, at most max_candidates kept.  One G-lane segment per anchor.
template <int G, class Cells>
__global__ __launch_bounds__(kBlock) void k_discover_count(int n, const double *__restrict__ x,
                                                           const double *__restrict__ y, const int *__restrict__ cx,
                                                           const int *__restrict__ cy, Cells cells, int R, double rad2,
                                                           unsigned max_cand, const int *__restrict__ sorted,
                                                           const double2 *__restrict__ sxy,
                                                           unsigned *__restrict__ deg, bool f32,
                                                           const unsigned *__restrict__ abort)
{
    if (abort && *abort) return;  // a guessed grid that does not hold: the build is repeated
    const int lane = threadIdx.x & 63;
    const int sl = lane & (G - 1);
    const int a = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool valid = a < n;
    if (__ballot(valid) == 0ULL) return;  // wave-uniform
    const unsigned long long segmask = segment_mask<G>(lane);
    double ax = 0.0, ay = 0.0;
    int ix = 0, iy = 0;
    if (valid) {
        ax = x[a];
        ay = y[a];
        ix = cx[a];
        iy = cy[a];
    }
    unsigned cnt = 0;
    for (int rb = 0; rb <= 2 * R; rb += G) {
        int rs, pre, tot;
        window_rows<G>(cells, ix, iy, R, rb, sl, valid, rs, pre, tot);
        const int nrows = min(G, 2 * R + 1 - rb);
        for (int t0 = 0; __ballot(t0 < tot) != 0ULL; t0 += G) {
            const int k = window_pos<G>(rs, pre, tot, nrows, t0 + sl);
            bool pass = false;
            if (k >= 0) {
                const double2 p = sxy[k];
                const double d = pair_dist2(ax, ay, p.x, p.y, f32);
                pass = sorted[k] != a && !(d > rad2);
            }
            cnt += static_cast<unsigned>(__popcll(__ballot(pass) & segmask));
        }
    }
    // the largest degree and the anchors that reached max_candidates
    // (preprocess.hpp:224, a >= test) come out of the offsets scan
    if (valid && sl == 0) deg[a] = cnt < max_cand ? cnt : max_cand;
}

constexpr int kFillWaves = 4;  // waves per workgroup in k_fill_csr
constexpr int kFillMaxAnchors = 32;  // anchors per k_fill_csr workgroup: (64 / G) x waves, G >= 8

// Single-pass CSR offsets in k_fill_csr (decoupled look-back): each
// workgroup publishes its anchors' degree sum, then its inclusive prefix, as
// one 64-bit word: flag (1: own sums, 2: inclusive prefix) << 62 | epoch << 45
// | largest degree << 31 | degree sum.  The epoch, one per launch, makes the
// words of earlier launches read as not ready, so the array is cleared only
// when it is allocated and when the epoch wraps.  The words are read and
// written with relaxed device-scope atomics: one word carries all a reader
// needs, so no fence (which would write back or invalidate the whole L2 of
// the XCD) is required.  Sums saturate at 2^31 - 1, above any pair capacity,
// and degrees at 2^14 - 1, above any LDS list capacity (both raise S_MISS).
constexpr unsigned kFillEpochBits = 17;
constexpr unsigned kFillSumMax = 0x7fffffffu;
constexpr unsigned kFillDegMax = 0x3fffu;

__device__ inline unsigned long long fill_word(unsigned flag, unsigned epoch, unsigned mx, unsigned long long sum)
{
    return static_cast<unsigned long long>(flag) << 62 | static_cast<unsigned long long>(epoch) << 45 |
           static_cast<unsigned long long>(min(mx, kFillDegMax)) << 31 |
           min(sum, static_cast<unsigned long long>(kFillSumMax));
}

__device__ inline unsigned wave_max_u32(unsigned v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, static_cast<unsigned>(__shfl_xor(v, o, 64)));
    return v;
}

// Workgroup wg's exclusive prefix over its predecessors (degree sum and
// largest degree), by one whole wave, after lane 0 has published the
// workgroup's own sums (flag 1): it reads the words of 64 predecessors at a
// time (lane l: wg - 1 - l - 64 k), adding own sums back to the nearest
// inclusive prefix, then publishes the inclusive prefix.  Every predecessor
// was dispatched before wg, so every wait ends; a wait that does not end
// within max_spins polls raises S_MISS instead of hanging.
__device__ inline void fill_lookback(unsigned long long *st, unsigned epoch, int wg, unsigned agg, unsigned agg_max,
                                     unsigned max_spins, unsigned *scal, unsigned &excl, unsigned &excl_max)
{
    const int lane = threadIdx.x & 63;
    const unsigned tag_a = 1u << kFillEpochBits | epoch, tag_p = 2u << kFillEpochBits | epoch;
    unsigned long long sum = 0;
    unsigned mx = 0;
    bool miss = false;
    for (int base = wg - 1; base >= 0; base -= 64) {
        const int j = base - lane;
        unsigned tag = tag_p;  // before the first workgroup: an empty inclusive prefix
        unsigned long long w = 0;
        if (j >= 0) {
            w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            while (static_cast<unsigned>(w >> 45) != tag_a && static_cast<unsigned>(w >> 45) != tag_p) {
                if (++spins > max_spins) {
                    miss = true;
                    w = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            tag = miss ? tag_p : static_cast<unsigned>(w >> 45);
        }
        const unsigned long long pm = __ballot(tag == tag_p);
        const int stop = pm != 0ULL ? __ffsll(static_cast<long long>(pm)) - 1 : 64;  // nearest inclusive prefix
        const bool take = lane <= stop;
        sum += wave_sum_u64(take ? (w & kFillSumMax) : 0ULL);
        mx = max(mx, wave_max_u32(take ? static_cast<unsigned>(w >> 31) & kFillDegMax : 0u));
        if (pm != 0ULL) break;
    }
    if (__ballot(miss) != 0ULL && lane == 0) atomicOr(&scal[S_MISS], 1u);
    excl = static_cast<unsigned>(min(sum, static_cast<unsigned long long>(kFillSumMax)));
    excl_max = mx;
    if (lane == 0)
        __hip_atomic_store(&st[wg], fill_word(2u, epoch, max(mx, agg_max), sum + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Rank of list entry i (distance d, cell c) among the segment's dg entries
// in LDS: by d2, ties by scan position, which is the insertion-sort order
// std::sort yields up to 16 entries.  Also flags exact ties, two entries of
// one cell, and exact ties within one cell.  Entries j < dmax lie inside the
// segment's cap-sized area, so they are read unconditionally (branch-free,
// unrollable) and masked.
__device__ inline int list_rank(const double *ld, const int *lcell, int dg, int dmax, int i, bool active, double di,
                                int celli, bool &tie, bool &shared, bool &ambiguous)
{
    int rank = 0;
#pragma unroll 4
    for (int j = 0; j < dmax; ++j) {
        const double dj = ld[j];
        const int cellj = lcell[j];
        const bool in = j < dg;
        rank += (in && (dj < di || (dj == di && j < i))) ? 1 : 0;
        const bool eq = in && active && dj == di && j != i;
        const bool same = in && active && cellj == celli && j != i;
        tie = tie || eq;
        shared = shared || same;
        ambiguous = ambiguous || (eq && same);
    }
    return rank;
}

// LDS doubles per k_fill_csr segment of list capacity cap: d2 (8 B), id and
// cell (4 B each), and with the counting variant the sorted order (4 B).
__host__ __device__ inline int fill_seg_doubles(int cap, bool count)
{
    return count ? (cap * 5 + 1) / 2 : cap * 2;
}

// discover_neighbors_csr gathering pass + the distance sort (preprocess.hpp:141-183)
// + the CSR fill (:240-262), one G-lane segment per anchor.  The segment walks
// the window's candidates in the reference's scan order (py outer, px inner,
// then within-cell order), G at a time; a ballot prefix keeps the first
// max_candidates in-radius ones, in scan order, in the segment's LDS.  Then
// every lane ranks its candidates by d2 against the whole list (ties by scan
// position, which is the insertion-sort order std::sort yields up to 16
// entries); lists with ties and more than 16 entries get their scan-order
// copy in `cand` and are queued for the std::sort replay (k_tie_sort).
// COUNT: no counting pass ran.  Each segment counts its in-radius candidates
// over the whole window (the degree is the count capped at max_candidates,
// preprocess.hpp:165), the workgroup finds its anchors' offsets with the
// look-back above and writes off[] (and off[n]), and the last workgroup
// writes the pair count, the largest degree and the anchors at
// max_candidates to the scalars (what the offsets scan gives otherwise).
template <int G, class Cells, bool COUNT>
__global__ __launch_bounds__(64 * kFillWaves) void k_fill_csr(int n, const double *__restrict__ x,
                                                              const double *__restrict__ y, const int *__restrict__ cx,
                                                              const int *__restrict__ cy, Cells cells, int R,
                                                              double rad2, const int *__restrict__ sorted,
                                                              const int *__restrict__ scell,
                                                              const double2 *__restrict__ sxy,
                                                              unsigned *__restrict__ off, Cand *__restrict__ cand,
                                                              int *__restrict__ ids, double *__restrict__ d2,
                                                              int *__restrict__ pair_anchor, int *__restrict__ tie_list,
                                                              unsigned max_cand, int cap,
                                                              const int *__restrict__ map_of,
                                                              const long long *__restrict__ map_eoff,
                                                              unsigned *__restrict__ scal, bool f32,
                                                              unsigned long long pair_cap,
                                                              unsigned long long *__restrict__ lb, unsigned epoch,
                                                              unsigned max_spins)
{
    __shared__ unsigned s_deg[COUNT ? kFillMaxAnchors : 1];
    __shared__ unsigned s_off[COUNT ? kFillMaxAnchors : 1];
    __shared__ unsigned s_flag[1];
    if constexpr (COUNT) {
        // one reading per workgroup: its waves must agree (block barriers follow)
        if (threadIdx.x == 0) s_flag[0] = scal[S_MISS];
        __syncthreads();
        if (s_flag[0]) {
            // a guess that did not hold: publish an empty prefix so that no
            // successor waits, and do nothing else
            if (threadIdx.x == 0)
                __hip_atomic_store(&lb[blockIdx.x], fill_word(2u, epoch, 0u, 0ULL), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    } else {
        // a guess that did not hold (grid or sizes): the build is repeated
        if (scal[S_MISS]) return;
    }
    // per segment: cap candidates (d2, id, cell; COUNT: and the sorted order)
    // in LDS for the ranking
    extern __shared__ double fill_lds[];
    const int lane = threadIdx.x & 63;
    const int sl = lane & (G - 1);
    const int seg = static_cast<int>(threadIdx.x) / G;
    double *ld = fill_lds + static_cast<size_t>(seg) * fill_seg_doubles(cap, COUNT);
    int *lid = reinterpret_cast<int *>(ld + cap);
    int *lcell = lid + cap;
    int *lord = lcell + cap;  // COUNT only: lord[rank] = scan position
    const int a = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x / G) + seg;
    const bool valid = a < n;
    if (!COUNT && __ballot(valid) == 0ULL) return;  // wave-uniform (COUNT: every wave reaches the barriers)
    const unsigned long long segmask = segment_mask<G>(lane);
    const unsigned long long lt_mask = (1ULL << lane) - 1ULL;
    double ax = 0.0, ay = 0.0;
    int ix = 0, iy = 0, dg = 0;
    unsigned o = 0;
    if (valid) {
        ax = x[a];
        ay = y[a];
        ix = cx[a];
        iy = cy[a];
        if (!COUNT) {
            o = off[a];
            dg = static_cast<int>(off[a + 1] - o);
            // buffers sized by a guess (the last map's): a list beyond them is
            // skipped and S_MISS raised, and the host builds the graph again
            if (dg > cap || static_cast<unsigned long long>(o) + static_cast<unsigned>(dg) > pair_cap) {
                if (sl == 0) atomicOr(&scal[S_MISS], 1u);
                dg = 0;
            }
        }
    }
    // the cut position: the degree when it is known, max_candidates when the
    // segment counts (a cut happens exactly when the count exceeds it)
    const int lim = COUNT ? static_cast<int>(min(max_cand, static_cast<unsigned>(INT_MAX))) : dg;
    const int keep = COUNT ? min(lim, cap) : dg;  // candidates stored in LDS
    int kept = 0;
    int last_cell = -1;  // cell of the last kept candidate
    bool ambiguous = false;
    bool done = !valid;
    for (int rb = 0; rb <= 2 * R && __ballot(!done) != 0ULL; rb += G) {
        int rs, pre, tot;
        window_rows<G>(cells, ix, iy, R, rb, sl, valid, rs, pre, tot);
        const int nrows = min(G, 2 * R + 1 - rb);
        for (int t0 = 0; __ballot(!done && t0 < tot) != 0ULL; t0 += G) {
            int k = window_pos<G>(rs, pre, tot, nrows, t0 + sl);
            if (done) k = -1;
            bool pass = false;
            double d = 0.0;
            int b = -1, cell = -1;
            if (k >= 0) {
                b = sorted[k];
                cell = scell[k];
                const double2 p = sxy[k];
                d = pair_dist2(ax, ay, p.x, p.y, f32);
                pass = b != a && !(d > rad2);
            }
            const unsigned long long m = __ballot(pass) & segmask;
            const int pos = kept + __popcll(m & lt_mask);
            if (pass && pos < keep) {
                ld[pos] = d;
                lid[pos] = b;
                lcell[pos] = cell;
            }
            // a max_candidates cut inside one cell makes the kept set depend on
            // the within-cell order: compare the first dropped candidate's cell
            // with the last kept one's
            const unsigned long long bm = __ballot(pass && pos == lim) & segmask;
            const unsigned long long pm = __ballot(pass && pos == lim - 1) & segmask;
            const int bcell = __shfl(cell, bm != 0ULL ? __ffsll(static_cast<long long>(bm)) - 1 : lane);
            const int pcell = __shfl(cell, pm != 0ULL ? __ffsll(static_cast<long long>(pm)) - 1 : lane);
            const int hcell = __shfl(cell, m != 0ULL ? 63 - __clzll(static_cast<long long>(m)) : lane);
            if (bm != 0ULL) {
                const int prev = pm != 0ULL ? pcell : last_cell;
                if (lim > 0 && prev == bcell) ambiguous = true;
                done = true;
            } else if (m != 0ULL) {
                if (kept + __popcll(m) <= lim) last_cell = hcell;
            }
            kept += __popcll(m);
            // without a cut (dg < max_candidates) nothing passes after the dg-th
            if (!COUNT && kept >= dg && static_cast<unsigned>(dg) < max_cand) done = true;
        }
    }
    if constexpr (COUNT) {
        // The degrees; the workgroup's own sums published at once, the ranks
        // computed while its predecessors finish, then the offsets
        dg = valid ? min(kept, lim) : 0;
        const int nseg = static_cast<int>(blockDim.x) / G;
        if (sl == 0) s_deg[seg] = static_cast<unsigned>(dg);
        __syncthreads();
        unsigned agg = 0, agg_max = 0;
        if (threadIdx.x < 64) {
            unsigned agg_trunc = 0;
            if (lane == 0) {
                for (int s2 = 0; s2 < nseg; ++s2) {
                    const unsigned v = s_deg[s2];
                    s_off[s2] = agg;
                    agg += v;
                    agg_max = max(agg_max, v);
                    agg_trunc += v >= max_cand ? 1u : 0u;  // preprocess.hpp:224, a >= test
                }
                if (blockIdx.x > 0)
                    __hip_atomic_store(&lb[blockIdx.x], fill_word(1u, epoch, agg_max, agg), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                // anchors at max_candidates: rare, one atomic per workgroup that has any
                if (agg_trunc > 0) atomicAdd(&scal[S_TRUNC], agg_trunc);
            }
            agg = bcast(static_cast<int>(agg), 0);
            agg_max = bcast(static_cast<int>(agg_max), 0);
        }
        // a list longer than the LDS capacity was not kept whole: the guess
        // did not hold
        const bool fits = dg <= cap;
        if (valid && !fits && sl == 0) atomicOr(&scal[S_MISS], 1u);
        const int dr = fits ? dg : 0;
        wave_sync();
        bool tie = false, shared = false;
        const int dmax = wave_max_over_segments<G>(dr);
        for (int i0 = 0; i0 < dmax; i0 += G) {
            const int i = i0 + sl;
            const bool active = i < dr;
            double di = 0.0;
            int celli = -1;
            if (active) {
                di = ld[i];
                celli = lcell[i];
            }
            const int rank = list_rank(ld, lcell, dr, dmax, i, active, di, celli, tie, shared, ambiguous);
            if (active) lord[rank] = i;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            unsigned excl, excl_max;
            fill_lookback(lb, epoch, static_cast<int>(blockIdx.x), agg, agg_max, max_spins, scal, excl, excl_max);
            if (lane == 0) {
                for (int s2 = 0; s2 < nseg; ++s2) {
                    const unsigned long long v = static_cast<unsigned long long>(excl) + s_off[s2];
                    s_off[s2] = static_cast<unsigned>(min(v, static_cast<unsigned long long>(kFillSumMax)));
                }
                if (blockIdx.x == gridDim.x - 1) {
                    const unsigned long long t = static_cast<unsigned long long>(excl) + agg;
                    const unsigned total = static_cast<unsigned>(min(t, static_cast<unsigned long long>(kFillSumMax)));
                    off[n] = total;
                    scal[S_PAIRS] = total;
                    scal[S_MAXDEG] = max(excl_max, agg_max);
                }
            }
        }
        __syncthreads();
        o = s_off[seg];
        if (valid && sl == 0) off[a] = o;
        // buffers sized by a guess (the last map's): a list beyond them is
        // skipped and S_MISS raised, and the host builds the graph again
        const bool room = static_cast<unsigned long long>(o) + static_cast<unsigned>(dg) <= pair_cap;
        if (valid && !room && sl == 0) atomicOr(&scal[S_MISS], 1u);
        const int dw = room ? dr : 0;
        for (int r = sl; r < dw; r += G) {
            const int i = lord[r];
            ids[o + r] = lid[i];
            d2[o + r] = ld[i];
            pair_anchor[o + r] = a;
        }
        const bool any_tie = (__ballot(tie) & segmask) != 0ULL;
        const bool any_shared = (__ballot(shared) & segmask) != 0ULL;
        const bool any_amb = (__ballot(ambiguous) & segmask) != 0ULL;
        if (any_tie && dw > 16) {
            for (int i = sl; i < dw; i += G) cand[o + i] = Cand{ld[i], lid[i]};  // scan order
            if (sl == 0) tie_list[atomicAdd(&scal[S_TIES], 1u)] = a;
        }
        // (the within-cell order cases of the counting path below)
        if (valid && sl == 0 && (any_amb || (dw > 16 && any_tie && any_shared))) {
            const long long map_n = map_of ? map_eoff[map_of[a] + 1] - map_eoff[map_of[a]] : n;
            if (map_n > 16) atomicAdd(&scal[S_AMBIG], 1u);
        }
        return;
    }
    wave_sync();
    // rank sort by d2 with ties by scan position; detect ties, same-cell ties
    // and cells that give two or more candidates
    bool tie = false;
    bool shared = false;
    const int dmax = wave_max_over_segments<G>(dg);
    for (int i0 = 0; i0 < dmax; i0 += G) {
        const int i = i0 + sl;
        const bool active = i < dg;
        double di = 0.0;
        int idi = 0, celli = -1;
        if (active) {
            di = ld[i];
            idi = lid[i];
            celli = lcell[i];
        }
        const int rank = list_rank(ld, lcell, dg, dmax, i, active, di, celli, tie, shared, ambiguous);
        if (active) {
            ids[o + rank] = idi;
            d2[o + rank] = di;
        }
    }
    const bool any_tie = (__ballot(tie) & segmask) != 0ULL;
    const bool any_shared = (__ballot(shared) & segmask) != 0ULL;
    const bool any_amb = (__ballot(ambiguous) & segmask) != 0ULL;
    for (int i = sl; i < dg; i += G) pair_anchor[o + i] = a;
    if (any_tie && dg > 16) {
        for (int i = sl; i < dg; i += G) cand[o + i] = Cand{ld[i], lid[i]};  // scan order
        if (sl == 0) tie_list[atomicAdd(&scal[S_TIES], 1u)] = a;
    }
    // The list depends on the within-cell order when two candidates of one
    // cell tie (any length), when a max_candidates cut falls inside a cell, or
    // when a list of more than 16 (introsort, not insertion sort) has ties
    // and any cell gives it two candidates: introsort places tied entries by
    // the positions of all entries.  The map's own spatial-index sort leaves
    // cells in edge-id order (as here) when the map has at most 16 edges.
    if (valid && sl == 0 && (any_amb || (dg > 16 && any_tie && any_shared))) {
        const long long map_n = map_of ? map_eoff[map_of[a] + 1] - map_eoff[map_of[a]] : n;
        if (map_n > 16) atomicAdd(&scal[S_AMBIG], 1u);
    }
}

// Anchors whose (more than 16) candidates contain exact distance ties: replay
// libstdc++'s introsort on the scan-order list k_fill_csr left in cand.
// Grid-stride over the device-side count, so no host round trip is needed.
__global__ __launch_bounds__(kBlock) void k_tie_sort(const int *__restrict__ tie_list,
                                                     const unsigned *__restrict__ scal,
                                                     const unsigned *__restrict__ off, Cand *__restrict__ cand,
                                                     int *__restrict__ ids, double *__restrict__ d2)
{
    const unsigned count = scal[S_TIES];
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const int a = tie_list[t];
        const unsigned o = off[a];
        const int dg = static_cast<int>(off[a + 1] - o);
        cvg::stl_sort(cand + o, dg);
        for (int i = 0; i < dg; ++i) {
            ids[o + i] = cand[o + i].id;
            d2[o + i] = cand[o + i].d;
        }
    }
}

// ---------------------------------------------------------------------------
// Curvelet stage
// ---------------------------------------------------------------------------

// The rows of trig (cvg_device.hpp), and in tdd the same values as
// double-doubles for the per-pair sines: tdd[8i .. 8i+7] = sin, cos of
// theta_i + dt, then of theta_i - dt (hi, lo each); from tdd + 8n, edge i's
// grid angles, 4 per j: sin hi, sin lo, cos hi, cos lo of theta_i + off_t[j].
// One thread per (edge i, angle k), k fastest: k < Kt the grid angle j = k,
// k = Kt theta + dt, Kt + 1 theta - dt, Kt + 2 theta itself.
__global__ __launch_bounds__(kBlock) void k_edge_trig(int n, const double *__restrict__ th, double dt, int kt,
                                                      double st, double *__restrict__ trig, double *__restrict__ tdd,
                                                      unsigned *__restrict__ scal)
{
    // the curvelet stage's counters start at zero (block 0 clears them; the
    // graph stage's words are left alone, so this may run before its last
    // host read)
    if (blockIdx.x == 0) {
        const int t = static_cast<int>(threadIdx.x);
        if (t == 0) scal[S_VALID] = 0u;
        if (t == 1) scal[S_COUNT] = 0u;
        if (t == 2) scal[S_COUNT + 1] = 0u;
        for (int w = kBucketBase + t; w < kScalarBufWords; w += static_cast<int>(blockDim.x)) scal[w] = 0u;
    }
    const unsigned na = static_cast<unsigned>(kt) + 3u;
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= static_cast<unsigned>(n) * na) return;
    const int i = static_cast<int>(g / na);
    const int k = static_cast<int>(g - static_cast<unsigned>(i) * na);
    const double t = th[i];
    // delta_t2 of grid column k (cpu_curvelet.hpp:698), the host table's expression
    const double off_t = k < kt ? st * (static_cast<double>(k) - (static_cast<double>(kt) - 1.0) / 2.0) : 0.0;
    const double ang = k < kt ? t + off_t         // T2t = T2 + delta_t2, cpu_curvelet.hpp:702
                       : k == kt ? t + dt         // T0_p_dt, :655
                       : k == kt + 1 ? t - dt     // T0_m_dt, :656
                                     : t;
    cvg::dd sd, cd;
    cvg::cr_sincos_dd(ang, c_sincos_tab, &sd, &cd);
    if (k < kt) {
        const long long e = static_cast<long long>(i) * kt + k;
        reinterpret_cast<double2 *>(trig + static_cast<long long>(cvg::TRIG_GRID) * n)[e] = make_double2(sd.hi, cd.hi);
        reinterpret_cast<double4 *>(tdd + static_cast<long long>(n) * 8)[e] = make_double4(sd.hi, sd.lo, cd.hi, cd.lo);
    } else if (k < kt + 2) {
        const bool plus = k == kt;
        trig[(plus ? cvg::TRIG_SIN_P : cvg::TRIG_SIN_M) * n + i] = sd.hi;
        trig[(plus ? cvg::TRIG_COS_P : cvg::TRIG_COS_M) * n + i] = cd.hi;
        reinterpret_cast<double4 *>(tdd)[2 * static_cast<long long>(i) + (plus ? 0 : 1)] =
            make_double4(sd.hi, sd.lo, cd.hi, cd.lo);
    } else {
        trig[cvg::TRIG_COS * n + i] = cd.hi;
        trig[cvg::TRIG_SIN * n + i] = sd.hi;
    }
}

// Per CSR pair q = (a, b): the direction filter (cpu_curvelet.hpp:459-461),
// and the cell-independent part of compute_curve_bundle (:646-656, 682, 690).
//   pflag[q] = class | close << 2 | branch_a << 3
//   prec[8q .. 8q+7] = DX, DY, DX2, DY2, DX cos(T0-dt), DY sin(T0-dt),
//                      DX cos(T0+dt), DY sin(T0+dt)
struct PrepArgs {
    const double *x, *y, *trig;
    double close_thr;
    unsigned char *pflag;
    double *prec;
};

// Pair q = (a, b)'s flags and record; returns whether the pair is close.
__device__ __forceinline__ bool pair_prep_one(int q, int n, int a, int b, const PrepArgs &P)
{
    const double ax = P.x[a], ay = P.y[a], bx = P.x[b], by = P.y[b];
    const double cos_at = P.trig[cvg::TRIG_COS * n + a];
    const double sin_at = P.trig[cvg::TRIG_SIN * n + a];
    const int k = cvg::direction_class(ax, ay, cos_at, sin_at, bx, by, c_sincos_tab, c_atan_tab);
    const cvg::PairRec p = cvg::pair_record(ax, ay, bx, by, P.trig[cvg::TRIG_SIN_P * n + b],
                                            P.trig[cvg::TRIG_COS_P * n + b], P.trig[cvg::TRIG_SIN_M * n + b],
                                            P.trig[cvg::TRIG_COS_M * n + b]);
    const bool close = p.DX2 + p.DY2 < P.close_thr;
    const bool branch_a = (p.DX * cos_at + p.DY * sin_at) < 0;
    P.pflag[q] = static_cast<unsigned char>(k | (close ? 4 : 0) | (branch_a ? 8 : 0));
    double2 *o = reinterpret_cast<double2 *>(P.prec + static_cast<long long>(q) * 8);
    o[0] = make_double2(p.DX, p.DY);
    o[1] = make_double2(p.DX2, p.DY2);
    o[2] = make_double2(p.cT0m_DX, p.sT0m_DY);
    o[3] = make_double2(p.cT0p_DX, p.sT0p_DY);
    return close;
}

__global__ __launch_bounds__(kBlock) void k_pair_prep(int pairs, int n, const int *__restrict__ pair_anchor,
                                                      const int *__restrict__ ids, PrepArgs P)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= pairs) return;
    (void)pair_prep_one(q, n, pair_anchor[q], ids[q], P);
}

// Anchors in spatial (cell) order, binned by max(forward, backward) + neither
// of their class counts (read from the anchor's run of pair flags): that key
// bounds every run's slot count and half the degree (the third class is
// almost always empty), which size the fused kernel's LDS slots and staging.
// Each list entry carries (anchor, CSR base, degree).
constexpr int kBucketBlock = 1024;  // few workgroups: one list-length atomic each per bucket
constexpr int kHeavyCap = 32;       // buckets from this cap start ahead of the light bucket 0
// Automatic light-bucket grids: kGridAutoMinPct percent of the resident grid
// plus one percent per kGridAutoEdgesPerPct edges of the run, at most all of
// it (measured best on single maps from 320x240 to 3840x2160: ~10% at 6k
// edges, 20-25% at 30-50k, 33% at 90k, 50% at 150k, 50-75% at 200k, 100%
// from ~350k).
constexpr long long kGridAutoEdgesPerPct = 4000;
constexpr int kGridAutoMinPct = 10;

int auto_grid_pct(long long n)
{
    return static_cast<int>(std::min<long long>(100, kGridAutoMinPct + n / kGridAutoEdgesPerPct));
}

__global__ __launch_bounds__(kBucketBlock) void k_bucket(int n, const int *__restrict__ sorted,
                                                   const unsigned char *__restrict__ pflag,
                                                   const unsigned *__restrict__ off,
                                                   int nb, const int *caps, int4 *__restrict__ lists,
                                                   unsigned *__restrict__ scal)
{
    __shared__ unsigned s_count[kMaxBuckets];
    __shared__ unsigned s_base[kMaxBuckets];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < kMaxBuckets) s_count[threadIdx.x] = 0;
    __syncthreads();
    int a = -1, b = -1;
    unsigned o0 = 0, o1 = 0;
    if (p < n) {
        a = sorted[p];
        o0 = off[a];
        o1 = off[a + 1];
        // the anchor's run of flags, four per aligned 32-bit load (the flag
        // buffer is padded to whole words)
        unsigned c[3] = {0u, 0u, 0u};
        for (unsigned q4 = o0 & ~3u; q4 < o1; q4 += 4) {
            const unsigned w = *reinterpret_cast<const unsigned *>(pflag + q4);
#pragma unroll
            for (unsigned k = 0; k < 4; ++k) {
                const unsigned q = q4 + k;
                const unsigned f = (w >> (8 * k)) & 3u;
                const bool in = q >= o0 && q < o1;
                c[0] += in && f == 0u ? 1u : 0u;
                c[1] += in && f == 1u ? 1u : 0u;
                c[2] += in && f == 2u ? 1u : 0u;
            }
        }
        const unsigned m = max(c[1], c[2]) + c[0];
        b = 0;
        while (b < nb - 1 && m > static_cast<unsigned>(caps[b])) ++b;
    }
    // per wave and bucket: the group's leader reserves slots in the block's
    // count, then one global atomic per block and bucket
    unsigned my_off = 0;
    for (int k = 0; k < nb; ++k) {
        const unsigned long long grp = __ballot(b == k);
        if (grp == 0ULL) continue;
        const int leader = __ffsll(static_cast<long long>(grp)) - 1;
        unsigned off = 0;
        if (lane == leader) off = atomicAdd(&s_count[k], static_cast<unsigned>(__popcll(grp)));
        off = __shfl(off, leader, 64);
        if (b == k) my_off = off + static_cast<unsigned>(__popcll(grp & ((1ULL << lane) - 1ULL)));
    }
    __syncthreads();
    if (threadIdx.x < nb && s_count[threadIdx.x] > 0)
        s_base[threadIdx.x] = atomicAdd(&scal[kBucketBase + threadIdx.x * kBucketStride], s_count[threadIdx.x]);
    __syncthreads();
    if (b >= 0)
        lists[static_cast<long long>(b) * n + s_base[b] + my_off] =
            make_int4(a, static_cast<int>(o0), static_cast<int>(o1 - o0), 0);
}

// Start gate for the light buckets' first launch: the heavy buckets' few
// anchors are the fused stage's critical path, and their workgroups need far
// more LDS than light ones, so they are placed best before light workgroups
// fill the CUs.  One wave waits (bounded: at most max_naps sleeps of ~3 us)
// until every heavy-bucket stream with work has started (each counts itself
// in scal[kStartedWord]); the light launch queued behind it then starts.
struct GateArgs {
    int nheavy;
    int bucket[kMaxBuckets];
    int streams[kMaxBuckets];
};

__global__ void k_gate(GateArgs G, const unsigned *__restrict__ scal, int max_naps)
{
    if (threadIdx.x != 0) return;
    unsigned want = 0;
    for (int i = 0; i < G.nheavy; ++i)
        want += min(scal[kBucketBase + G.bucket[i] * kBucketStride], static_cast<unsigned>(G.streams[i]));
    for (int k = 0; k < max_naps; ++k) {
        if (__hip_atomic_load(scal + kStartedWord, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) return;
        __builtin_amdgcn_s_sleep(127);
    }
}

// Per-pair sines of the transport (compute_curve_bundle, cpu_curvelet.hpp:728-729):
// psin[q*2Kt + 2j] = sin(T2t_j - (theta_n - dt)), psin[q*2Kt + 2j + 1] =
// sin(T2t_j - (theta_n + dt)) with T2t_j = theta_a + delta_t2[j], for every
// pair that is not "close" (pair flag bit 2; close pairs keep the default
// bundle and never read them, and get zeros).  One thread per pair; KT > 0
// fixes Kt at compile time so that every load is issued before the arithmetic
// (KT = 0: any Kt up to kPairTrigMaxKt).  Each sine is the correctly rounded
// sine of the rounded difference, formed from the double-double sin/cos of
// both angles in tdd by the subtraction formula; the rare lane whose rounding
// that does not settle evaluates the sine directly.  For Kt up to
// kPairTrigMaxKt the block's results go through LDS (dynamic, kBlock x Kt
// double2) so that its output span is written contiguously.  PREP: the
// thread first forms its pair's flags and record (k_pair_prep's work) and
// takes the close bit from there instead of reading pflag.  With scal (a
// launch queued before the host has read the graph's counts, on buffers of
// the guessed capacity) the pair count comes from the device, and a guess
// that did not hold (or bad input) leaves the launch with nothing to do.
constexpr int kPairTrigMaxKt = 16;
template <int KT, bool PREP>
__global__ __launch_bounds__(kBlock) void k_pair_trig(int pairs, int n, int kt_rt, const int *__restrict__ pair_anchor,
                                                      const int *__restrict__ ids,
                                                      const unsigned char *__restrict__ pflag,
                                                      const double *__restrict__ th, double st_step, double dt,
                                                      const double *__restrict__ tdd, double *__restrict__ psin,
                                                      PrepArgs P, const unsigned *__restrict__ scal)
{
    extern __shared__ double2 stage[];
    if (scal) pairs = (scal[S_MISS] | scal[S_BAD]) ? 0 : static_cast<int>(scal[S_PAIRS]);
    if (static_cast<int>(blockIdx.x * blockDim.x) >= pairs) return;  // the whole block
    const int kt = KT > 0 ? KT : kt_rt;
    // delta_t2 of grid column j (cpu_curvelet.hpp:698), the host table's expression
    const double half = (static_cast<double>(kt) - 1.0) / 2.0;
    const bool staged = kt <= kPairTrigMaxKt;
    const int q0 = blockIdx.x * blockDim.x;
    const int q = q0 + threadIdx.x;
    double2 *out = reinterpret_cast<double2 *>(psin) + static_cast<long long>(q0) * kt;
    double2 *row = (staged ? stage : out) + threadIdx.x * kt;
    if (q < pairs) {
        const int a = pair_anchor[q];
        const int b = ids[q];
        const bool close = PREP ? pair_prep_one(q, n, a, b, P) : (pflag[q] & 4) != 0;
        if (close) {  // :682
            for (int j = 0; j < kt; ++j) row[j] = make_double2(0.0, 0.0);
        } else {
            const double ta = th[a];
            const double nt = th[b];
            const double tm = nt - dt, tp = nt + dt;
            const double4 *nr = reinterpret_cast<const double4 *>(tdd) + 2 * static_cast<long long>(b);
            const double4 p = nr[0], m = nr[1];
            const cvg::dd sp = {p.x, p.y}, cp = {p.z, p.w}, sm = {m.x, m.y}, cm = {m.z, m.w};
            const double4 *gr =
                reinterpret_cast<const double4 *>(tdd + static_cast<long long>(n) * 8) + static_cast<long long>(a) * kt;
            if (KT > 0) {
                double4 g[KT > 0 ? KT : 1];
#pragma unroll
                for (int j = 0; j < (KT > 0 ? KT : 1); ++j) g[j] = gr[j];
#pragma unroll
                for (int j = 0; j < (KT > 0 ? KT : 1); ++j) {
                    const double T2t = ta + st_step * (static_cast<double>(j) - half);
                    const cvg::dd st = {g[j].x, g[j].y}, ct = {g[j].z, g[j].w};
                    double2 v;
                    v.x = cvg::sin_difference(T2t, tm, st, ct, sm, cm, c_sincos_tab);
                    v.y = cvg::sin_difference(T2t, tp, st, ct, sp, cp, c_sincos_tab);
                    row[j] = v;
                }
            } else {
                for (int j = 0; j < kt; ++j) {
                    const double T2t = ta + st_step * (static_cast<double>(j) - half);
                    const double4 g = gr[j];
                    const cvg::dd st = {g.x, g.y}, ct = {g.z, g.w};
                    double2 v;
                    v.x = cvg::sin_difference(T2t, tm, st, ct, sm, cm, c_sincos_tab);
                    v.y = cvg::sin_difference(T2t, tp, st, ct, sp, cp, c_sincos_tab);
                    row[j] = v;
                }
            }
        }
    }
    if (!staged) return;
    __syncthreads();
    const int cnt = min(static_cast<int>(blockDim.x), pairs - q0) * kt;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) out[e] = stage[e];
}

// Host-computed per-parameter tables, laid out in one device block.  Lanes of
// the fused kernel own bundle cells in "rank" order: rank = position of cell
// c = i*Kt + j when cells are sorted by (ox^2 + ot^2, c), so the cell that
// fill_curvelet_info picks (smallest distance, first in row-major order,
// cpu_curvelet.hpp:202-219) is the lowest lane whose cell is valid.
struct Tables {
    double *rank_ox;       // [K]
    int *rank_j;           // [K]
    int *rank_cell;        // [K]
    int *rank_elig;        // [K] d < 100 (the initial best distance)
    double *off_x;         // [Kx]
    double *off_t;         // [Kt]
    int *caps;             // [kMaxBuckets]
};

struct FusedArgs {
    int n;
    const double *trig, *psin, *prec;
    const int *ids;
    const unsigned char *pflag;
    const int4 *list;  // (anchor, CSR base, degree, -)
    const unsigned *list_len;  // the bucket's entry count (device; written by k_bucket)
    int nb_cap;        // bundle slots per anchor: the bucket's longest run
    int deg_cap;       // staged neighbour records per anchor
    int chain_stride;  // ids per seed record: max(G, 1)
    int st_stride;     // staging ids per row: max(G, 1)
    double dx, dx2, max_k;
    int kt, K, G;
    const double *rank_ox;
    const int *rank_j, *rank_cell, *rank_elig;
    int rank_of_cell0;
    int *st_ids;
    unsigned *st_meta;
    double *st_k;
    int *st_rank;         // per CSR slot: row among the anchor's accepted chains, or -1
    unsigned *acc_count;
    unsigned *valid_pairs;
    unsigned long long *intersections;
    double *g_bundle;  // global-bundle kernels: nb_cap*K*2 doubles per anchor stream
    unsigned *started; // heavy buckets: each stream with work counts itself here at its start
};

// One anchor stream's LDS slice (byte offsets): the staging buffer (the
// anchor's pair records and sines, copied from the pre-passes' contiguous
// blocks), the run's bundles by slot (lo, hi; absent
// when they live in global memory), the seeds' results (length, hash, best
// cell, chain), the slots' hypothesis flags, a list area for compacting the
// run's slots and hypotheses (2 x 64 x nch ints) and the anchor's per-seed
// accept flags (dc ints).
struct AnchorCarve {
    size_t stg, stg_bytes, rec, sin, blo, bhi, ski, hyp, lst, acf, bytes;
};

__host__ __device__ inline AnchorCarve anchor_carve(int nb, int dc, int K, int kt, int cs, bool lds_bundles, int nch)
{
    AnchorCarve c{};
    c.rec = 0;  // within a staging buffer
    c.sin = static_cast<size_t>(dc) * 8 * sizeof(double);
    c.stg_bytes = c.sin + static_cast<size_t>(dc) * kt * 2 * sizeof(double);
    size_t o = 0;
    c.stg = o;
    o += c.stg_bytes;
    c.blo = o;
    if (lds_bundles) o += static_cast<size_t>(nb) * K * sizeof(double);
    c.bhi = o;
    if (lds_bundles) o += static_cast<size_t>(nb) * K * sizeof(double);
    c.ski = o;
    o += static_cast<size_t>(nb) * (cs + 3) * sizeof(int);
    c.hyp = o;
    o += static_cast<size_t>(nb) * sizeof(int);
    c.lst = o;
    o += static_cast<size_t>(2) * 64 * nch * sizeof(int);
    c.acf = o;
    o += static_cast<size_t>(dc) * sizeof(int);
    c.bytes = (o + 15) & ~static_cast<size_t>(15);
    return c;
}

// Load from a wave-uniform address through the scalar data cache (read only).
template <typename T>
__device__ __forceinline__ T uload(const T *p)
{
    return *(const __attribute__((address_space(4))) T *)p;
}

// Asynchronous copy of cnt 16-byte entries from global memory into LDS
// (global_load_lds_dwordx4: lane i of a chunk writes M0 + 16 i); waves t of W
// split the chunks.  Completion: s_waitcnt vmcnt(0), then a barrier.  Issued
// as inline assembly: through the builtin the compiler cannot tell the
// staging buffer from the rest of the LDS and adds vmcnt(0) waits before
// unrelated LDS reads.
__device__ __forceinline__ void copy_to_lds(const double2 *src, char *dst, int cnt, int t, int W, int lane)
{
    for (int c0 = t * 64; c0 < cnt; c0 += W * 64) {
        const int i = c0 + lane;
        const unsigned lds = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(
            reinterpret_cast<size_t>((__attribute__((address_space(3))) char *)(dst + 16 * c0))));
        if (i < cnt)
            asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
                         :
                         : "v"(src + i), "s"(lds)
                         : "memory", "m0");
    }
}

// Lane-indexed arrays: entry k of a wave-wide list lives in lane k % 64 of
// register v[k / 64]; k is wave-uniform.
template <int NCH>
__device__ __forceinline__ int lane_get(const int (&v)[NCH], int k)
{
    int r = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int t = __builtin_amdgcn_readlane(v[c], k & 63);
        if ((k >> 6) == c) r = t;
    }
    return r;
}

template <int NCH>
__device__ __forceinline__ void lane_set(int (&v)[NCH], int k, int val, int lane)
{
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        if ((k >> 6) == c && lane == (k & 63)) v[c] = val;
}

// Lowest remaining index of a multi-word lane mask (uniform), -1 when empty.
template <int NCH>
__device__ __forceinline__ int pop_next(unsigned long long (&m)[NCH])
{
    int k = -1;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const bool take = k < 0 && m[c] != 0ULL;
        const int r = __builtin_ctzll(m[c] | (1ULL << 63));
        k = take ? 64 * c + r : k;
        m[c] = take ? (m[c] & (m[c] - 1ULL)) : m[c];
    }
    return k;
}

__device__ __forceinline__ unsigned chain_hash(unsigned h, int id)
{
    return (h ^ static_cast<unsigned>(id)) * 16777619u;
}

// build_curvelets_greedy (cpu_curvelet.hpp:437-591) plus the
// form_pairwise_bundles valid-pair count (:305-361), in one pass.
//
// Persistent: the grid holds as many anchor streams as the GPU keeps resident;
// stream g processes list entries g, g + streams, ...  A stream is one wave
// (COOP = false, blockDim/64 streams per workgroup) or a whole workgroup whose
// W = blockDim/64 waves split each anchor's bundles (slot i on wave i % W) and
// seeds (hypothesis j on wave j % W) (COOP = true).  An anchor's pair records
// and sines are copied into the stream's LDS staging buffer (global_load_lds)
// when it starts; its list entry, flags, ids and per-lane trig values were
// loaded while the previous anchor was processed.  One staging buffer, not
// two: the copy's latency is covered by the other resident streams, and the
// smaller slice keeps more streams resident.
//
// Per anchor, runs 0 = forward neighbours (class 1), 1 = backward (class 2),
// 2 = neither (class 0: validity only, for the count).  Phase A: the run's
// bundles (lane = cell in rank order, CPL cells per lane) into LDS slots in CSR
// order, with hypothesis flags.  Phase C: each hypothesis seeds a chain grown
// by intersecting the hypotheses' bundles in CSR order (validity = one
// ballot); the seed's chain, hash and best cell (fill_curvelet_info,
// :202-226: the first valid cell in rank order) go to LDS.  Phase D (wave 0):
// in seed order, a chain longer than 2 is accepted unless an accepted chain of
// this (anchor, run) is identical (hash match, then element check), and
// written to the anchor's staging rows.  Control state is wave-uniform;
// per-neighbour flags and ids, slot ids and accepted hashes are lane-indexed
// registers.
template <int CPL, int NCH, bool GB, bool COOP>
__global__ __launch_bounds__(256) void k_curvelets(FusedArgs A)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int wpb = static_cast<int>(blockDim.x >> 6);
    const int W = COOP ? wpb : 1;      // waves per anchor
    const int t = COOP ? wave : 0;     // this wave's rank within its anchor
    const int grp = COOP ? 0 : wave;   // anchor stream within the workgroup
    const int apb = COOP ? 1 : wpb;    // anchor streams per workgroup
    const int streams = static_cast<int>(gridDim.x) * apb;
    const int g = static_cast<int>(blockIdx.x) * apb + grp;
    const int len_list = static_cast<int>(uload(A.list_len));
    if (A.started && g < len_list && lane == 0 && t == 0) atomicAdd(A.started, 1u);
    if (g >= len_list) return;  // COOP: the whole workgroup (same g)

    const int n = A.n;
    const int K = A.K;
    const int kt = A.kt;
    const int G = A.G;
    const int CS = A.chain_stride;
    const int SR = CS + 3;  // ints per seed record
    const int NB = A.nb_cap;
    const AnchorCarve cv = anchor_carve(NB, A.deg_cap, K, kt, CS, !GB, NCH);
    char *sl = reinterpret_cast<char *>(lds) + static_cast<size_t>(grp) * cv.bytes;
    int *ski = reinterpret_cast<int *>(sl + cv.ski);
    int *hypf = reinterpret_cast<int *>(sl + cv.hyp);
    int *lst = reinterpret_cast<int *>(sl + cv.lst);
    int *accf = reinterpret_cast<int *>(sl + cv.acf);
    const unsigned long long lt_mask = (1ULL << lane) - 1ULL;
    double *blo;
    double *bhi;
    if constexpr (GB) {
        blo = A.g_bundle + static_cast<size_t>(g) * NB * K * 2;
        bhi = blo + static_cast<size_t>(NB) * K;
    } else {
        blo = reinterpret_cast<double *>(sl + cv.blo);
        bhi = reinterpret_cast<double *>(sl + cv.bhi);
    }
    auto sync = [&]() {
        if constexpr (COOP)
            __syncthreads();
        else
            wave_sync();
    };

    // per-lane cells (rank order): anchor-independent
    double ox[CPL];
    int cj[CPL], ccell[CPL], coff[CPL];
    bool live[CPL], elig[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) {
        const int rho = lane + 64 * m;
        live[m] = rho < K;
        const int r = live[m] ? rho : 0;
        coff[m] = r;
        ox[m] = A.rank_ox[r];
        cj[m] = A.rank_j[r];
        ccell[m] = A.rank_cell[r];
        elig[m] = live[m] && A.rank_elig[r] != 0;
    }
    // the same as wave masks: a validity test is then one compare ballot
    unsigned long long lmask[CPL], emask[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) {
        lmask[m] = __ballot(live[m]);
        emask[m] = __ballot(elig[m]);
    }

    // the pipeline: entry, staging copy and per-lane values of the next anchor
    const double2 *prec2 = reinterpret_cast<const double2 *>(A.prec);
    const double2 *psin2 = reinterpret_cast<const double2 *>(A.psin);
    auto entry = [&](int L, int &ea, unsigned &eb, int &ed) {
        const int *e = reinterpret_cast<const int *>(A.list + L);
        ea = uload(e);
        eb = static_cast<unsigned>(uload(e + 1));
        ed = uload(e + 2);
    };
    auto stage = [&](int buf, unsigned b0, int d0) {
        char *sb = sl + cv.stg + static_cast<size_t>(buf) * cv.stg_bytes;
        copy_to_lds(prec2 + static_cast<size_t>(b0) * 4, sb + cv.rec, d0 * 4, t, W, lane);
        copy_to_lds(psin2 + static_cast<size_t>(b0) * kt, sb + cv.sin, d0 * kt, t, W, lane);
    };
    double nsT2t[CPL], ncT2t[CPL];
    int nnfl[NCH], nnid[NCH];
    auto lane_vals = [&](int a0, unsigned b0, int d0) {
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            const double2 sc = reinterpret_cast<const double2 *>(A.trig + static_cast<long long>(cvg::TRIG_GRID) * n)
                [static_cast<long long>(a0) * kt + cj[m]];
            nsT2t[m] = sc.x;
            ncT2t[m] = sc.y;
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int r = 64 * c + lane;
            nnfl[c] = r < d0 ? static_cast<int>(A.pflag[b0 + r]) : 3;
            nnid[c] = r < d0 ? A.ids[b0 + r] : 0;
        }
    };

    int L = g;
    int a, deg;
    unsigned base;
    entry(L, a, base, deg);
    lane_vals(a, base, deg);
    int Ln = L + streams;
    int na = 0, ndeg = 0;
    unsigned nbase = 0;
    if (Ln < len_list) entry(Ln, na, nbase, ndeg);

    unsigned vpw = 0;     // valid or close pairs computed by this wave
    unsigned isect = 0;   // intersections tested by this wave
    while (L < len_list) {
        const int buf = 0;
        double sT2t[CPL], cT2t[CPL];
        int nfl[NCH], nid[NCH];
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            sT2t[m] = nsT2t[m];
            cT2t[m] = ncT2t[m];
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            nfl[c] = nnfl[c];
            nid[c] = nnid[c];
        }
        sync();  // every reader of the previous anchor's staging is done
        stage(0, base, deg);
        __builtin_amdgcn_s_waitcnt(0);  // this anchor's staging copy and values have landed
        sync();
        // start the next anchor's loads
        const int Lnn = Ln + streams;
        int nna = 0, nndeg = 0;
        unsigned nnbase = 0;
        if (Ln < len_list) {
            lane_vals(na, nbase, ndeg);
            if (Lnn < len_list) entry(Lnn, nna, nnbase, nndeg);
        }

        const char *sb = sl + cv.stg + static_cast<size_t>(buf) * cv.stg_bytes;
        const double2 *srec = reinterpret_cast<const double2 *>(sb + cv.rec);
        const double2 *ssin = reinterpret_cast<const double2 *>(sb + cv.sin);
        int nseed = 0;  // hypotheses of the earlier runs (their seeds' slots come first)
        for (int run = 0; run < 3; ++run) {
            const int want = run == 0 ? 1 : (run == 1 ? 2 : 0);
            unsigned long long mk[NCH];
            int nL = 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                mk[c] = __ballot((nfl[c] & 3) == want);
                nL += __popcll(mk[c]);
            }
            if (nL == 0) continue;

            // phase A: the run's bundles (compute_curve_bundle, :636-819), slot
            // i = i-th neighbour of the run in CSR order.  The slots' neighbour
            // indices are compacted through the list area (every wave writes
            // the same values) into lane-indexed rr, their ids into rid.
            {
                int pb = 0;
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    if ((mk[c] >> lane) & 1ULL) lst[pb + __popcll(mk[c] & lt_mask)] = 64 * c + lane;
                    pb += __popcll(mk[c]);
                }
            }
            wave_sync();
            int rr[NCH], rid[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) rr[c] = 64 * c + lane < nL ? lst[64 * c + lane] : 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                int v = 0;
#pragma unroll
                for (int cc = 0; cc < NCH; ++cc) {
                    const int g = __shfl(nid[cc], rr[c] & 63, 64);
                    if ((rr[c] >> 6) == cc) v = g;
                }
                rid[c] = v;
            }
            for (int i = COOP ? t : 0; i < nL; i += W) {
                const int r = lane_get<NCH>(rr, i);
                const int fl = lane_get<NCH>(nfl, r);
                const double2 *rr = srec + r * 4;
                const double2 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3];
                cvg::PairRec q;
                q.DX = r0.x;
                q.DY = r0.y;
                q.DX2 = r1.x;
                q.DY2 = r1.y;
                q.cT0m_DX = r2.x;
                q.sT0m_DY = r2.y;
                q.cT0p_DX = r3.x;
                q.sT0p_DY = r3.y;
                const bool qclose = (fl & 4) != 0;
                const bool qbr = (fl & 8) != 0;
                double lo[CPL], hi[CPL];
                unsigned long long anym = 0ULL;
#pragma unroll
                for (int m = 0; m < CPL; ++m) {
                    lo[m] = -A.max_k;
                    hi[m] = A.max_k;
                    if (!qclose) {
                        const double2 s = ssin[r * kt + cj[m]];
                        cvg::bundle_cell(q, qbr, ox[m], sT2t[m], cT2t[m], s.x, s.y, A.dx, A.dx2, A.max_k, lo[m],
                                         hi[m]);
                    }
                    anym |= __builtin_amdgcn_ballot_w64(hi[m] > lo[m]) & lmask[m];
                }
                const bool hyp = qclose || anym != 0ULL;  // bundle_valid_check, :821-830
                vpw += hyp ? 1u : 0u;
                if (run == 2) continue;
                if (hyp) {
#pragma unroll
                    for (int m = 0; m < CPL; ++m) {
                        if (live[m]) {
                            blo[static_cast<size_t>(i) * K + coff[m]] = lo[m];
                            bhi[static_cast<size_t>(i) * K + coff[m]] = hi[m];
                        }
                    }
                }
                if (lane == 0) hypf[i] = hyp ? 1 : 0;
            }
            if (run == 2) continue;
            sync();

            // phase B: hypotheses = flagged slots, in CSR order; lane h holds
            // hypothesis h's bundle offset (hso, bytes) and neighbour id (hid)
            unsigned long long hm[NCH];
            int nH = 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int i = 64 * c + lane;
                hm[c] = __ballot(i < nL && hypf[i] != 0);
                nH += __popcll(hm[c]);
            }
            // compacted through the list area like the slots (the barrier above
            // ends every wave's reads of the slot list)
            {
                int pb = 0;
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    if ((hm[c] >> lane) & 1ULL) {
                        const int h = pb + __popcll(hm[c] & lt_mask);
                        lst[h] = (64 * c + lane) * K * 8;
                        lst[64 * NCH + h] = rid[c];
                    }
                    pb += __popcll(hm[c]);
                }
            }
            wave_sync();
            int hso[NCH], hid[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int h = 64 * c + lane;
                hso[c] = h < nH ? lst[h] : 0;
                hid[c] = h < nH ? lst[64 * NCH + h] : 0;
            }

            if (nH > 0 && G >= 3) {  // no chain longer than 2 when G < 3
                const char *blo_b = reinterpret_cast<const char *>(blo);
                const char *bhi_b = reinterpret_cast<const char *>(bhi);
                // phase C: growth (:479-536), seed j on wave j % W; the walk
                // visits the hypotheses in CSR order, the seed at its own rank
                for (int j = t; j < nH; j += W) {
                    double lo[CPL], hi[CPL];
                    const int soff = lane_get<NCH>(hso, j);
#pragma unroll
                    for (int m = 0; m < CPL; ++m) {
                        lo[m] = *reinterpret_cast<const double *>(blo_b + soff + coff[m] * 8);
                        hi[m] = *reinterpret_cast<const double *>(bhi_b + soff + coff[m] * 8);
                    }
                    int len = 1;
                    int cur = lane == 0 ? a : 0;  // chain position l in lane l
                    unsigned hsh = chain_hash(2166136261u, a);
                    // The seed joins at its own distance rank (:496-510); its step
                    // is an intersection with its own bundle, which leaves the
                    // running bundle (inside the seed's) bit for bit unchanged.
                    int h = 0;
                    for (; h < nH && len < G; ++h) {
                        double bl[CPL], bh[CPL];
                        const int noff = lane_get<NCH>(hso, h);
#pragma unroll
                        for (int m = 0; m < CPL; ++m) {
                            bl[m] = *reinterpret_cast<const double *>(blo_b + noff + coff[m] * 8);
                            bh[m] = *reinterpret_cast<const double *>(bhi_b + noff + coff[m] * 8);
                        }
                        double nlo[CPL], nhi[CPL];
                        unsigned long long okm = 0ULL;
#pragma unroll
                        for (int m = 0; m < CPL; ++m) {
                            nlo[m] = (lo[m] < bl[m]) ? bl[m] : lo[m];  // bundle_intersection, :838-841
                            nhi[m] = (hi[m] > bh[m]) ? bh[m] : hi[m];
                            okm |= __builtin_amdgcn_ballot_w64(nhi[m] > nlo[m]) & lmask[m];
                        }
                        if (h == j || okm != 0ULL) {
                            const int id = lane_get<NCH>(hid, h);
                            cur = lane == len ? id : cur;
                            hsh = chain_hash(hsh, id);
#pragma unroll
                            for (int m = 0; m < CPL; ++m) {
                                lo[m] = nlo[m];
                                hi[m] = nhi[m];
                            }
                            ++len;
                        }
                    }
                    isect += static_cast<unsigned>(h - (j < h ? 1 : 0));  // steps other than the seed's
                    int *rec = ski + j * SR;
                    if (lane == 0) rec[0] = len;
                    if (len <= 2) continue;  // :549
                    // best cell of the running bundle (fill_curvelet_info, :202-226)
                    int best = -1;
#pragma unroll
                    for (int m = 0; m < CPL; ++m) {
                        const unsigned long long mask = __builtin_amdgcn_ballot_w64(hi[m] > lo[m]) & emask[m];
                        if (best < 0 && mask != 0ULL) best = 64 * m + __builtin_ctzll(mask);
                    }
                    if (best < 0) best = A.rank_of_cell0;
                    const int bm = best >> 6;
                    const int bl = best & 63;
                    double kmax = 0.0, kmin = 0.0;
                    int best_cell = 0;
#pragma unroll
                    for (int m = 0; m < CPL; ++m) {
                        if (m == bm) {
                            kmax = bcast(hi[m], bl);
                            kmin = bcast(lo[m], bl);
                            best_cell = bcast(ccell[m], bl);
                        }
                    }
                    if (lane < len) rec[3 + lane] = cur;
                    if (lane == 0) {
                        rec[1] = static_cast<int>(hsh);
                        rec[2] = best_cell;
                    }
                    // the seed's row, staged at its own slot of the anchor; it is
                    // kept unless an earlier seed has the same chain
                    const long long q = static_cast<long long>(base) + nseed + j;
                    if (lane < len) A.st_ids[q * A.st_stride + lane] = cur;
                    if (lane == 0) {
                        A.st_meta[q] = static_cast<unsigned>(len) | (run == 0 ? 0x8000u : 0u) |
                                       (static_cast<unsigned>(best_cell) << 16);
                        A.st_k[2 * q] = kmax;
                        A.st_k[2 * q + 1] = kmin;
                    }
                }
                sync();

                // dedup (:546-560, 876-916), seed j on wave j % W: seed j is
                // dropped when an earlier seed j' (length > 2) has the identical
                // chain; the first of equal chains is the one the reference
                // accepted, so "identical to an earlier accepted chain" and
                // "identical to an earlier seed's chain" coincide.  Earlier seeds
                // with the same length and hash are found in one ballot (lane =
                // seed), then their chains compared element by element.
                {
                    int slen[NCH], shash[NCH];
#pragma unroll
                    for (int c = 0; c < NCH; ++c) {
                        const int jj = 64 * c + lane;
                        slen[c] = jj < nH ? ski[jj * SR] : 0;
                        shash[c] = (jj < nH && slen[c] > 2) ? ski[jj * SR + 1] : 0;
                    }
                    for (int j = t; j < nH; j += W) {
                        const int lj = lane_get<NCH>(slen, j);
                        bool accept = lj > 2;
                        if (accept) {
                            const int hj = lane_get<NCH>(shash, j);
#pragma unroll
                            for (int c = 0; c < NCH; ++c) {
                                unsigned long long cm =
                                    __ballot(64 * c + lane < j && slen[c] == lj && shash[c] == hj);
                                while (accept && cm != 0ULL) {
                                    const int jp = 64 * c + __builtin_ctzll(cm);
                                    cm &= cm - 1ULL;
                                    const bool ne = lane < lj && ski[j * SR + 3 + lane] != ski[jp * SR + 3 + lane];
                                    if (__ballot(ne) == 0ULL) accept = false;
                                }
                            }
                        }
                        if (lane == 0) accf[nseed + j] = accept ? 1 : 0;
                    }
                }
            }
            sync();  // the next run reuses the slots and seed records
            nseed += nH;
        }
        // every slot of the anchor: its row among the anchor's accepted chains
        // (run-0 seeds, then run-1 seeds, each in seed order), or -1
        if (t == 0) {
            int accn = 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int k = 64 * c + lane;
                const bool ac = G >= 3 && k < nseed && accf[k] != 0;
                const unsigned long long am = __ballot(ac);
                if (k < deg) A.st_rank[base + k] = ac ? accn + __popcll(am & lt_mask) : -1;
                accn += __popcll(am);
            }
            if (lane == 0) A.acc_count[a] = static_cast<unsigned>(accn);
        }

        // advance the pipeline
        L = Ln;
        a = na;
        base = nbase;
        deg = ndeg;
        Ln = Lnn;
        na = nna;
        nbase = nnbase;
        ndeg = nndeg;
    }
    if (lane == 0) {
        if (vpw) atomicAdd(A.valid_pairs, vpw);
        if (isect) atomicAdd(A.intersections, static_cast<unsigned long long>(isect));
    }
}

using FusedKernel = void (*)(FusedArgs);

// k_curvelets<CPL, NCH, GB, COOP> by (cells per lane, 64-neighbour chunks,
// bundles in global memory, waves sharing one anchor)
FusedKernel fused_kernel(int cpl, int nch, bool gb, bool coop)
{
#define CVG_FK(C, N)                                                                                         \
    if (cpl == C && nch == N)                                                                                \
        return gb ? k_curvelets<C, N, true, false> : (coop ? k_curvelets<C, N, false, true>                  \
                                                           : k_curvelets<C, N, false, false>);
    CVG_FK(1, 1) CVG_FK(1, 2) CVG_FK(1, 4)
    CVG_FK(2, 1) CVG_FK(2, 2) CVG_FK(2, 4)
    CVG_FK(4, 1) CVG_FK(4, 2) CVG_FK(4, 4)
#undef CVG_FK
    return nullptr;
}

struct RecordArgs {
    int pairs;
    const double *x, *y, *th;
    const int *pair_anchor;
    const int *map_of;        // per edge: its map (batches), or null
    const long long *map_eoff;  // per map: first edge id
    const unsigned *row_off;
    const int *st_rank;
    const int *st_ids;
    const unsigned *st_meta;
    const double *st_k;
    int st_stride;
    int G, kt;
    double nrad;
    const double *off_x, *off_t;
    unsigned *chains;
    double *info;
    unsigned max_rows;  // rows the output buffers hold; rows beyond are not written
    // float instantiation: info in float arithmetic from the float grid
    // offsets of fill_curvelet_info (cpu_curvelet.hpp:209-210, 221-222)
    bool f32;
    const float *inf_x, *inf_t;
    float nrad_f;
};

// record_curvelet_chain + fill_curvelet_info (cpu_curvelet.hpp:197-280) for
// every accepted staging row, written at its compacted position
// (compact_curvelet_output, :282-303).  One thread per staging slot q (one per
// seed of the anchor, the rest unused): an accepted row carries its rank among
// the anchor's accepted rows.  Chain ids are written 1-based and relative to
// the anchor's map (batches concatenate maps).  F32: the float instantiation
// (R.f32); each kernel instantiation holds only its own arithmetic.
constexpr int kRecordChunk = 8;  // chain ids (and their coordinates) loaded together

template <bool F32>
__device__ inline void record_row(const RecordArgs &R, unsigned o, int q)
{
    using T = typename std::conditional<F32, float, double>::type;
    const int a = R.pair_anchor[q];
    const int id0 = R.map_of ? static_cast<int>(R.map_eoff[R.map_of[a]]) : 0;
    const double ax = R.x[a], ay = R.y[a], at = R.th[a];
    const int W = R.G + 1;
    const unsigned meta = R.st_meta[q];
    const int len = static_cast<int>(meta & 0x7fffu);
    const bool fwd = (meta >> 15) & 1u;
    const int c = static_cast<int>(meta >> 16);
    const int *chain = R.st_ids + static_cast<long long>(q) * R.st_stride;
    unsigned *row = R.chains + static_cast<long long>(o) * W;
    row[0] = static_cast<unsigned>(a - id0) + 1;
    // The chain row, and the chain's length: the distances between
    // consecutive members in chain order (:231-240, sqrt in T).  Ids and
    // coordinates come kRecordChunk at a time, so their loads overlap.
    T length = 0;
    T x1 = 0, y1 = 0;
    for (int l0 = 0; l0 < R.G; l0 += kRecordChunk) {
        int id[kRecordChunk];
#pragma unroll
        for (int u = 0; u < kRecordChunk; ++u) id[u] = l0 + u < len ? chain[l0 + u] : 0;
#pragma unroll
        for (int u = 0; u < kRecordChunk; ++u)
            if (l0 + u < R.G) row[l0 + u + 1] = l0 + u < len ? static_cast<unsigned>(id[u] - id0) + 1 : 0u;
        T xs[kRecordChunk], ys[kRecordChunk];
#pragma unroll
        for (int u = 0; u < kRecordChunk; ++u) {
            xs[u] = l0 + u < len ? static_cast<T>(R.x[id[u]]) : T(0);
            ys[u] = l0 + u < len ? static_cast<T>(R.y[id[u]]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < kRecordChunk; ++u) {
            const int l = l0 + u;
            if (l >= len) continue;
            if (l > 0) {
                const T d2 = (x1 - xs[u]) * (x1 - xs[u]) + (y1 - ys[u]) * (y1 - ys[u]);
                if constexpr (F32)
                    length += cvg::sqrt_f(d2);
                else
                    length += sqrt(d2);
            }
            x1 = xs[u];
            y1 = ys[u];
        }
    }

    double *inf = R.info + static_cast<long long>(o) * kInfoWidth;
    if constexpr (F32) {
        // fill_curvelet_info<float> (:197-256): every value a float, std::sin,
        // std::cos and std::sqrt the float functions
        const float fx = static_cast<float>(ax), fy = static_cast<float>(ay), ft = static_cast<float>(at);
        const float ox = R.inf_x[c / R.kt];
        const float ot = R.inf_t[c % R.kt];
        const float k_max = static_cast<float>(R.st_k[2 * static_cast<long long>(q)]);
        const float k_min = static_cast<float>(R.st_k[2 * static_cast<long long>(q) + 1]);
        const float kk = (k_max + k_min) / 2.0f;
        const float theta = cvg::wrap_two_pi_f(ft + ot);
        float s, cth;
        cvg::cr_sincosf<false>(theta, c_sincos_tab, &s, &cth);  // theta in [0, 2 pi)
        const float px = fx - ox * s;
        const float py = fy + ox * cth;
        const float quality = (length > 0.0f && len > 0)
                                  ? 2.0f / (1.0f * R.nrad_f / length + 1.0f * length / static_cast<float>(len))
                                  : 0.0f;
        inf[0] = fwd ? 1.0 : 0.0;
        inf[1] = k_max;
        inf[2] = k_min;
        inf[3] = ft;
        inf[4] = px;
        inf[5] = py;
        inf[6] = theta;
        inf[7] = kk;
        inf[8] = length;
        inf[9] = quality;
    } else {
        const double ox = R.off_x[c / R.kt];
        const double ot = R.off_t[c % R.kt];
        const double k_max = R.st_k[2 * static_cast<long long>(q)];
        const double k_min = R.st_k[2 * static_cast<long long>(q) + 1];
        const double kk = (k_max + k_min) / 2.0;
        const double theta = cvg::wrap_two_pi(at + ot);
        double s, cth;
        cvg::cr_sincos<false>(theta, c_sincos_tab, &s, &cth);  // theta in [0, 2 pi)
        const double px = ax - ox * s;
        const double py = ay + ox * cth;
        const double quality = (length > 0.0 && len > 0) ? 2.0 / (R.nrad / length + length / double(len)) : 0.0;
        inf[0] = fwd ? 1.0 : 0.0;
        inf[1] = k_max;
        inf[2] = k_min;
        inf[3] = at;
        inf[4] = px;
        inf[5] = py;
        inf[6] = theta;
        inf[7] = kk;
        inf[8] = length;
        inf[9] = quality;
    }
}

// compact_curvelet_output (cpu_curvelet.hpp:282-303) in two passes.  First,
// one thread per staging slot: an accepted slot (its rank among the anchor's
// accepted rows, or -1) names its output row in slot_of_row.  Then one thread
// per output row records it; the row count (the last row offset) stays on
// the device, and the rows beyond the caller's buffers are left out (the host
// reports them).
__global__ __launch_bounds__(kBlock) void k_row_slots(RecordArgs R, int *__restrict__ slot_of_row,
                                                      const unsigned *__restrict__ num_rows, unsigned *__restrict__ scal)
{
    const int q = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q == 0) scal[S_ROWS] = *num_rows;  // read back with the other scalars
    if (q >= R.pairs) return;
    const int rk = R.st_rank[q];
    if (rk < 0) return;
    const unsigned o = R.row_off[R.pair_anchor[q]] + static_cast<unsigned>(rk);
    if (o < R.max_rows) slot_of_row[o] = q;
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void k_record(RecordArgs R, const int *__restrict__ slot_of_row,
                                                   const unsigned *__restrict__ num_rows)
{
    const unsigned rows = min(*num_rows, R.max_rows);
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < rows; o += stride)
        record_row<F32>(R, o, slot_of_row[o]);
}

// ---------------------------------------------------------------------------
// The general curvelet kernel: the float instantiation (run_curvelet<float>,
// main.cpp:110 with --float), and double maps outside the fused kernel's range
// (more than 256 neighbours per anchor, group_max_sz above 64, neighbour lists
// too long for one workgroup's LDS)
// ---------------------------------------------------------------------------

// Direction class of a neighbour in float: dot(To2Pi(atan2f(dy, dx)), theta_a)
// with atan2f, cosf and sinf (curvelet_utils.hpp:58-69 on T = float).  Its
// evaluation errors (atan2f, To2Pi's 2 pi, cosf, sinf, two products and a
// sum in float) stay below 2e-6, so where the projection of the neighbour on
// the anchor's direction exceeds 4e-6 |v| its sign is the class; elsewhere
// the float expression itself decides.
__device__ inline int direction_class_f(float ax, float ay, float at, float nx, float ny, float cos_at,
                                        float sin_at)
{
    const float vx = nx - ax, vy = ny - ay;
    const double p = static_cast<double>(vx) * cos_at + static_cast<double>(vy) * sin_at;
    const double r = sqrt(static_cast<double>(vx) * vx + static_cast<double>(vy) * vy);
    if (fabs(p) > 4e-6 * r) return p > 0 ? 1 : 2;
    const float dir = cvg::wrap_two_pi_f(cvg::cr_atan2f(vy, vx, c_atan_tab));
    float sd, cd;
    cvg::cr_sincosf(dir, c_sincos_tab, &sd, &cd);
    const float dot = cd * cos_at + sd * sin_at;
    return dot > 0 ? 1 : (dot < 0 ? 2 : 0);
}

// The arithmetic of one instantiation: T = double (the reference's default,
// and the same device functions as the fused kernel's pre-passes) or T =
// float (main.cpp includes <math.h>, whose float overloads the unqualified
// sin and cos calls of cpu_curvelet.hpp resolve to; the double literals of
// :697-698, 711-715 and 774 promote those expressions to double until they
// are stored in T).
template <typename T>
struct GenMath;

template <>
struct GenMath<double> {
    __device__ static void sincos(double a, double *s, double *c) { cvg::cr_sincos(a, c_sincos_tab, s, c); }
    __device__ static int direction(double ax, double ay, double at, double nx, double ny, double cat, double sat)
    {
        (void)at;
        return cvg::direction_class(ax, ay, cat, sat, nx, ny, c_sincos_tab, c_atan_tab);
    }
    // compute_curve_bundle<double> (:636-830) for the lane's cells; false
    // when the pair is close (its bundle stays the full default)
    template <int M>
    __device__ static bool bundle(double ax, double ay, double nx, double ny, double nt, double cat, double sat,
                                  const bool *live, const double *ox, const double *T2t, const double *sT2t,
                                  const double *cT2t, double dt, double dx, double dx2, double close_thr,
                                  double max_k, double *lo, double *hi)
    {
        double sp0, cp0, sm0, cm0;
        cvg::cr_sincos(nt + dt, c_sincos_tab, &sp0, &cp0);
        cvg::cr_sincos(nt - dt, c_sincos_tab, &sm0, &cm0);
        const PairRec pr = cvg::pair_record(ax, ay, nx, ny, sp0, cp0, sm0, cm0);
        const bool close = pr.DX2 + pr.DY2 < close_thr;  // :682
        const bool branch_a = (pr.DX * cat + pr.DY * sat) < 0;  // :690
#pragma unroll
        for (int m = 0; m < M; ++m) {
            lo[m] = -max_k;
            hi[m] = max_k;
            if (close || !live[m]) continue;
            const double sin_m = cvg::cr_sin(T2t[m] - (nt - dt), c_sincos_tab);
            const double sin_p = cvg::cr_sin(T2t[m] - (nt + dt), c_sincos_tab);
            cvg::bundle_cell(pr, branch_a, ox[m], sT2t[m], cT2t[m], sin_m, sin_p, dx, dx2, max_k, lo[m], hi[m]);
        }
        return !close;
    }
};

template <>
struct GenMath<float> {
    __device__ static void sincos(float a, float *s, float *c) { cvg::cr_sincosf(a, c_sincos_tab, s, c); }
    __device__ static int direction(float ax, float ay, float at, float nx, float ny, float cat, float sat)
    {
        return direction_class_f(ax, ay, at, nx, ny, cat, sat);
    }
    template <int M>
    __device__ static bool bundle(float ax, float ay, float nx, float ny, float nt, float cat, float sat,
                                  const bool *live, const float *ox, const float *T2t, const float *sT2t,
                                  const float *cT2t, float dt, float dx, float dx2, float close_thr, float max_k,
                                  float *lo, float *hi)
    {
        (void)close_thr;
        const float DX = ax - nx, DY = ay - ny;
        const float DX2 = DX * DX, DY2 = DY * DY;
        const float T0p = nt + dt, T0m = nt - dt;
        const bool close = DX2 + DY2 < 4 * dx * dx;  // :682
#pragma unroll
        for (int m = 0; m < M; ++m) {
            lo[m] = -max_k;
            hi[m] = max_k;
        }
        if (close) return false;
        const bool branch_a = (DX * cat + DY * sat) < 0;  // :690
        float sp0, cp0, sm0, cm0;
        cvg::cr_sincosf(T0p, c_sincos_tab, &sp0, &cp0);
        cvg::cr_sincosf(T0m, c_sincos_tab, &sm0, &cm0);
        const float sT0p_DY = DY * sp0;  // :730-733
        const float cT0p_DX = DX * cp0;
        const float sT0m_DY = DY * sm0;
        const float cT0m_DX = DX * cm0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (!live[m]) continue;
            const float o = ox[m];
            const float tt = sT2t[m] * DX - cT2t[m] * DY;  // :705
            const float ttm = tt - o;
            const float sm = cvg::cr_sinf(T2t[m] - T0m, c_sincos_tab);  // :728-729
            const float sp = cvg::cr_sinf(T2t[m] - T0p, c_sincos_tab);
            if (branch_a) {
                const float den = static_cast<float>(-2.0 * o * tt + DX2 + DY2 - dx2 + o * o);  // :711
                const float kx_min = static_cast<float>(2.0 * (ttm - dx) / den);
                const float kx_max = static_cast<float>(2.0 * (ttm + dx) / den);
                if (lo[m] < kx_min) lo[m] = kx_min;
                if (hi[m] > kx_max) hi[m] = kx_max;
                const float kt_min = sm / (-sm * o + cT0m_DX + sT0m_DY);
                const float kt_max = sp / (-sp * o + cT0p_DX + sT0p_DY);
                if (lo[m] < kt_min) lo[m] = kt_min;
                if (hi[m] > kt_max) hi[m] = kt_max;
            } else {
                const float den = static_cast<float>(-(2.0 * o * tt - DX2 - DY2 + dx2 - o * o) / 2.0);  // :774
                const float kx_min = (ttm - dx) / den;
                const float kx_max = (ttm + dx) / den;
                if (lo[m] < kx_min) lo[m] = kx_min;
                if (hi[m] > kx_max) hi[m] = kx_max;
                const float kt_max = sm / (-sm * o + cT0m_DX + sT0m_DY);
                if (hi[m] > kt_max) hi[m] = kt_max;
                const float kt_min = sp / (-sp * o + cT0p_DX + sT0p_DY);
                if (lo[m] < kt_min) lo[m] = kt_min;
            }
        }
        return true;
    }
};

template <typename T>
struct GenArgs {
    int n;
    const double *x, *y, *th;   // edge values (a float map's values, held exactly as doubles)
    const unsigned *off;
    const int *ids;
    int kt, K, G, cap, max_deg;  // cap: longest chain the map allows, min(G, max_deg + 1)
    T dx, dt, max_k, dx2, close_thr;
    const T *off_x, *off_t;      // delta_x2, delta_t2 (:697-698)
    const int *cell_rank;        // fill_curvelet_info's order of a cell (d, then index), K where d >= 100
    const int *rank_cell;        // the cell of each rank
    T *sc_bundle;                // per wave: max_deg x K lo, then hi
    int *sc_flag;                // per wave: max_deg neighbour flags
    int *sc_acc;                 // per wave: max_deg x (cap + 1) accepted chains of the run
    int *sc_chain;               // per wave: cap, the chain being grown
    int *st_ids;
    unsigned *st_meta;
    double *st_k;
    int *st_rank;
    unsigned *acc_count;
    unsigned *valid_pairs;
    unsigned long long *intersections;
};

constexpr int kGenWaves = 4;  // waves per workgroup of k_curvelets_generic

// build_curvelets_greedy<T> (cpu_curvelet.hpp:437-591) with the
// form_pairwise_bundles<T> count (:305-361), one wave per anchor (a
// persistent grid: wave w takes anchors w, w + waves, ...), lane = bundle
// cell (up to 4 per lane).  Each neighbour's bundle goes to the wave's slice
// of device memory once (it does not depend on the run); the runs'
// hypotheses are the neighbours of their direction class with a valid (or
// close) bundle.  Chains grow in the wave's chain slice (any length the map
// allows), and accepted rows are staged as k_curvelets stages them, at the
// anchor's slots base, base + 1, ...
template <typename T>
__global__ __launch_bounds__(64 * kGenWaves) void k_curvelets_generic(GenArgs<T> A)
{
    using M = GenMath<T>;
    const int lane = threadIdx.x & 63;
    const int w = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = static_cast<int>((gridDim.x * blockDim.x) >> 6);
    const int K = A.K;
    const int G = A.G;
    const int cw = A.cap + 1;
    T *blo = A.sc_bundle + static_cast<size_t>(w) * A.max_deg * K * 2;
    T *bhi = blo + static_cast<size_t>(A.max_deg) * K;
    int *flag = A.sc_flag + static_cast<size_t>(w) * A.max_deg;
    int *acc = A.sc_acc + static_cast<size_t>(w) * A.max_deg * cw;
    int *chn = A.sc_chain + static_cast<size_t>(w) * A.cap;
    // per-lane cells
    bool live[4];
    int cell[4], crank[4];
    T ox[4], ot[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int c = lane + 64 * m;
        live[m] = c < K;
        cell[m] = live[m] ? c : 0;
        ox[m] = A.off_x[cell[m] / A.kt];
        ot[m] = A.off_t[cell[m] % A.kt];
        crank[m] = live[m] ? A.cell_rank[cell[m]] : K;
    }
    unsigned vpw = 0;
    unsigned long long isect = 0;
    for (int a = w; a < A.n; a += nwaves) {
        const unsigned base = A.off[a];
        const int deg = static_cast<int>(A.off[a + 1] - base);
        const T ax = static_cast<T>(A.x[a]), ay = static_cast<T>(A.y[a]);
        const T at = static_cast<T>(A.th[a]);
        T sat, cat;
        M::sincos(at, &sat, &cat);
        T T2t[4], sT2t[4], cT2t[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            T2t[m] = at + ot[m];  // :702
            M::sincos(T2t[m], &sT2t[m], &cT2t[m]);
        }
        // every neighbour: direction class and bundle (compute_curve_bundle<T>)
        for (int r = 0; r < deg; ++r) {
            const int b = A.ids[base + r];
            const T nx = static_cast<T>(A.x[b]), ny = static_cast<T>(A.y[b]);
            const T nt = static_cast<T>(A.th[b]);
            const int cls = M::direction(ax, ay, at, nx, ny, cat, sat);
            T lo[4], hi[4];
            const bool far = M::template bundle<4>(ax, ay, nx, ny, nt, cat, sat, live, ox, T2t, sT2t, cT2t, A.dt,
                                                   A.dx, A.dx2, A.close_thr, A.max_k, lo, hi);
            unsigned long long anym = 0ULL;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                anym |= __ballot(live[m] && hi[m] > lo[m]);
                if (live[m]) {
                    blo[static_cast<size_t>(r) * K + cell[m]] = lo[m];
                    bhi[static_cast<size_t>(r) * K + cell[m]] = hi[m];
                }
            }
            const bool valid = !far || anym != 0ULL;  // bundle_valid_check, :821-830
            vpw += valid ? 1u : 0u;
            if (lane == 0) flag[r] = cls | (valid ? 4 : 0);
        }
        wave_sync();  // flags written by lane 0 are read by every lane
        unsigned nacc = 0;  // accepted rows of the anchor so far
        for (int run = 0; run < 2; ++run) {
            const int want = run == 0 ? 1 : 2;
            unsigned nrun = 0;  // accepted rows of this run (the dedup table)
            for (int sidx = 0; sidx < deg; ++sidx) {
                const int fs = flag[sidx];
                if ((fs & 3) != want || !(fs & 4)) continue;
                T lo[4], hi[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    lo[m] = live[m] ? blo[static_cast<size_t>(sidx) * K + cell[m]] : T(0);
                    hi[m] = live[m] ? bhi[static_cast<size_t>(sidx) * K + cell[m]] : T(0);
                }
                int len = 1;
                if (lane == 0) chn[0] = a;
                for (int r = 0; r < deg; ++r) {  // :479-536
                    const int fr = flag[r];
                    if ((fr & 3) != want || !(fr & 4)) continue;
                    const int nb = A.ids[base + r];
                    if (r == sidx) {
                        if (lane == 0) chn[len] = nb;
                        ++len;
                        if (len >= G) break;
                        continue;
                    }
                    ++isect;
                    T nlo[4], nhi[4];
                    unsigned long long okm = 0ULL;
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const T bl = live[m] ? blo[static_cast<size_t>(r) * K + cell[m]] : T(0);
                        const T bh = live[m] ? bhi[static_cast<size_t>(r) * K + cell[m]] : T(0);
                        nlo[m] = (lo[m] < bl) ? bl : lo[m];  // :838-841
                        nhi[m] = (hi[m] > bh) ? bh : hi[m];
                        okm |= __ballot(live[m] && nhi[m] > nlo[m]);
                    }
                    if (okm != 0ULL) {
                        if (lane == 0) chn[len] = nb;
                        ++len;
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            lo[m] = nlo[m];
                            hi[m] = nhi[m];
                        }
                    }
                    if (len >= G) break;
                }
                if (len <= 2) continue;  // :549
                wave_sync();  // the chain written by lane 0 is read by every lane
                // check_curvelet_exist (:876-916) against this run's accepted chains
                bool dup = false;
                for (unsigned k = 0; k < nrun && !dup; ++k) {
                    const int *row = acc + static_cast<size_t>(k) * cw;
                    if (row[0] != len) continue;
                    bool ne = false;
                    for (int j = lane; j < len; j += 64) ne = ne || row[1 + j] != chn[j];
                    dup = __ballot(ne) == 0ULL;
                }
                if (dup) continue;
                int *row = acc + static_cast<size_t>(nrun) * cw;
                if (lane == 0) row[0] = len;
                for (int j = lane; j < len; j += 64) row[1 + j] = chn[j];
                ++nrun;
                // fill_curvelet_info's cell (:202-226): smallest d among the
                // running bundle's valid cells, first in row-major order
                int key = K;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int k2 = (live[m] && hi[m] > lo[m]) ? crank[m] : K;
                    key = k2 < key ? k2 : key;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const int v = __shfl_xor(key, o, 64);
                    key = v < key ? v : key;
                }
                const int best = key < K ? A.rank_cell[key] : 0;
                const int bl = best & 63, bm = best >> 6;
                T kmax = T(0), kmin = T(0);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const T h = __shfl(hi[m], bl, 64), l = __shfl(lo[m], bl, 64);
                    if (m == bm) {
                        kmax = h;
                        kmin = l;
                    }
                }
                const long long q = static_cast<long long>(base) + nacc;
                for (int j = lane; j < len; j += 64) A.st_ids[q * A.cap + j] = chn[j];
                if (lane == 0) {
                    A.st_meta[q] = static_cast<unsigned>(len) | (run == 0 ? 0x8000u : 0u) |
                                   (static_cast<unsigned>(best) << 16);
                    A.st_k[2 * q] = static_cast<double>(kmax);
                    A.st_k[2 * q + 1] = static_cast<double>(kmin);
                }
                ++nacc;
                wave_sync();  // the dedup row is read by every lane
            }
        }
        for (int k = lane; k < deg; k += 64) A.st_rank[base + k] = k < static_cast<int>(nacc) ? k : -1;
        if (lane == 0) A.acc_count[a] = nacc;
        wave_sync();  // the scratch is rewritten by the next anchor
    }
    if (lane == 0 && vpw) atomicAdd(A.valid_pairs, vpw);
    if (lane == 0 && isect) atomicAdd(A.intersections, isect);
}

// ---------------------------------------------------------------------------
// Diagnostics kernels
// ---------------------------------------------------------------------------

// One anchor: per neighbour r and cell c (reference cell order) the bundle,
// with every per-pair quantity computed from scratch by the shared device
// functions.
__global__ void k_debug_anchor(int a, int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ th, const unsigned *__restrict__ off,
                               const int *__restrict__ ids, double dt, double dx, double dx2, double close_thr,
                               double max_k, int kx, int kt, const double *__restrict__ off_x,
                               const double *__restrict__ off_t, double *__restrict__ out /* deg*K*2 + deg*2 */)
{
    const int K = kx * kt;
    const unsigned base = off[a];
    const int deg = static_cast<int>(off[a + 1] - base);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= deg * K) return;
    const int r = t / K, c = t % K;
    const int i = c / kt, j = c % kt;
    const int b = ids[base + r];
    const double ax = x[a], ay = y[a], at = th[a];
    const double nx = x[b], ny = y[b], nt = th[b];
    double sa, ca;
    cvg::cr_sincos(at, c_sincos_tab, &sa, &ca);
    double sp, cp, sm, cm;
    cvg::cr_sincos(nt + dt, c_sincos_tab, &sp, &cp);
    cvg::cr_sincos(nt - dt, c_sincos_tab, &sm, &cm);
    const PairRec p = cvg::pair_record(ax, ay, nx, ny, sp, cp, sm, cm);
    const bool close = p.DX2 + p.DY2 < close_thr;
    const bool branch_a = (p.DX * ca + p.DY * sa) < 0;
    const double T2t = at + off_t[j];
    double s2, c2;
    cvg::cr_sincos(T2t, c_sincos_tab, &s2, &c2);
    const double sin_m = cvg::cr_sin(T2t - (nt - dt), c_sincos_tab);
    const double sin_p = cvg::cr_sin(T2t - (nt + dt), c_sincos_tab);
    double lo = -max_k, hi = max_k;
    if (!close) cvg::bundle_cell(p, branch_a, off_x[i], s2, c2, sin_m, sin_p, dx, dx2, max_k, lo, hi);
    out[static_cast<size_t>(r) * K + c] = lo;
    out[static_cast<size_t>(deg) * K + static_cast<size_t>(r) * K + c] = hi;
    if (c == 0) {
        out[static_cast<size_t>(2) * deg * K + r] =
            static_cast<double>(cvg::direction_class(ax, ay, ca, sa, nx, ny, c_sincos_tab, c_atan_tab));
        out[static_cast<size_t>(2) * deg * K + deg + r] = close ? 1.0 : 0.0;
    }
}

__host__ __device__ inline double math_op(int op, double x, double y, const double *sc, const double *atb)
{
    switch (op) {
    case 0: return cvg::cr_sin(x, sc);
    case 1: return cvg::cr_cos(x, sc);
    case 2: return cvg::cr_atan2(y, x, atb);
    case 3: return sqrt(x);
    case 4: return cvg::wrap_two_pi(x);
    case 5: return fmod(x, y);
    case 6:
    case 7: {
        double sv, cv;
        cvg::cr_sincos(x, sc, &sv, &cv);
        return op == 6 ? sv : cv;
    }
    case 8:
    case 9: {
        cvg::dd sa, ca, sb, cb;
        cvg::cr_sincos_dd(x, sc, &sa, &ca);
        cvg::cr_sincos_dd(y, sc, &sb, &cb);
        if (op == 8) return cvg::sin_difference(x, y, sa, ca, sb, cb, sc);
        double v;
        return cvg::sin_of_difference(x, y, sa, ca, sb, cb, &v) ? 1.0 : 0.0;
    }
    case 10:
    case 11:
    case 12:
    case 13: {
        cvg::dd sv, cv;
        cvg::cr_sincos_dd(x, sc, &sv, &cv);
        return op == 10 ? sv.hi : op == 11 ? sv.lo : op == 12 ? cv.hi : cv.lo;
    }
    // binary32: arguments rounded to float first, results widened (exactly)
    case 14: return cvg::cr_sinf(static_cast<float>(x), sc);
    case 15: return cvg::cr_cosf(static_cast<float>(x), sc);
    case 16: return cvg::cr_atan2f(static_cast<float>(y), static_cast<float>(x), atb);
    case 17: return cvg::wrap_two_pi_f(static_cast<float>(x));
    case 18: return cvg::sqrt_f(static_cast<float>(x));
    default: return 0.0;
    }
}

__global__ void k_math_eval(int op, int n, const double *__restrict__ x, const double *__restrict__ y,
                            double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = math_op(op, x[i], y ? y[i] : 0.0, c_sincos_tab, c_atan_tab);
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------

// A kernel's dynamic LDS limit raised to shmem (once per kernel and size) and,
// with occ, its resident workgroups per CU at that block and LDS size (the
// runtime's occupancy query, answered once per context and shape): both are
// host calls that otherwise cost tens of microseconds per run.
int prepare_launch(cvg_context *ctx, const void *fn, int block, unsigned shmem, int *occ)
{
    unsigned &lim = ctx->shmem_set[fn];
    if (lim < shmem) {
        CVG_HIP(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
        lim = shmem;
    }
    if (occ) {
        const auto key = std::make_tuple(fn, block, shmem);
        const auto it = ctx->occupancy.find(key);
        if (it != ctx->occupancy.end()) {
            *occ = it->second;
        } else {
            CVG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, fn, block, shmem));
            ctx->occupancy[key] = *occ;
        }
    }
    return CVG_OK;
}

int read_scalars(cvg_context *ctx)
{
    CVG_HIP(ctx, hipMemcpyAsync(ctx->mailbox, ctx->d_scal.ptr, kScalarWords * sizeof(unsigned), hipMemcpyDeviceToHost,
                                ctx->stream));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CVG_OK;
}

constexpr int kSpecMiss = 1000;  // internal: a guessed size did not hold (never returned to callers)

int queue_edge_trig(cvg_context *ctx, const cvg_params &p);
int queue_pair_setup(cvg_context *ctx, const cvg_params &p, unsigned long long pair_cap);
// Bytes of the pair-flag buffer for pairs pairs: whole 32-bit words, since
// k_bucket reads each anchor's run of flags a word at a time.
inline size_t cls_bytes(unsigned long long pairs)
{
    return static_cast<size_t>((pairs + 4) & ~3ULL);
}

// Neighbour counts, CSR offsets, then the gathering pass, the distance sort
// and the CSR fill (preprocess.hpp:186-274) over either cell layout.  The
// pair count and largest degree size the fill; with spec they are not read
// first: the fill runs on the buffers and list capacity of the context's last
// graph (guarded on the device, S_MISS when they do not hold), and the host
// reads the counts once, after it (kSpecMiss: build again with the real sizes).
template <class Cells>
int graph_scans(cvg_context *ctx, const Cells &cells, int n, unsigned R, unsigned max_cand, double rad2, int num_maps,
                bool spec)
{
    hipStream_t st = ctx->stream;
    unsigned *scal = as<unsigned>(ctx->d_scal);
    const size_t ni = static_cast<size_t>(n) * sizeof(int);
    CVG_TRY(ensure(ctx, ctx->d_off, static_cast<size_t>(n + 1) * sizeof(unsigned)));
    // on guessed buffers the fill counts the lists itself and finds the
    // offsets in the same pass; otherwise a counting pass and the offsets
    // scan size the buffers first
    const bool count_in_fill = spec;
    if (!count_in_fill) {
        CVG_TRY(ensure(ctx, ctx->d_deg, ni));
        using CountFn = void (*)(int, const double *, const double *, const int *, const int *, Cells, int, double,
                                 unsigned, const int *, const double2 *, unsigned *, bool, const unsigned *);
        const int G = ctx->fill_lanes;
        const CountFn fn = G == 8    ? k_discover_count<8, Cells>
                           : G == 32 ? k_discover_count<32, Cells>
                           : G == 64 ? k_discover_count<64, Cells>
                                     : k_discover_count<16, Cells>;
        hipLaunchKernelGGL(fn, dim3(blocks_for(static_cast<long long>(n) * G)), dim3(kBlock), 0, st, n,
                           as<double>(ctx->d_x), as<double>(ctx->d_y), as<int>(ctx->d_cx), as<int>(ctx->d_cy), cells,
                           static_cast<int>(R), rad2, max_cand, as<int>(ctx->d_sorted), as<double2>(ctx->d_sxy),
                           as<unsigned>(ctx->d_deg), ctx->f32, spec ? static_cast<const unsigned *>(scal + S_MISS) : nullptr);
        CVG_HIP(ctx, hipGetLastError());
        CVG_TRY(exclusive_scan(ctx, as<unsigned>(ctx->d_deg), n, as<unsigned>(ctx->d_off), ctx->d_tiles,
                               scal + S_MAXDEG, scal + S_TRUNC, max_cand, scal + S_PAIRS));
    }
    auto take_counts = [&]() -> int {
        const unsigned pairs_u = ctx->mailbox[S_PAIRS];
        if (pairs_u > static_cast<unsigned>(INT_MAX))
            return fail(ctx, CVG_ERR_UNSUPPORTED, "more than 2^31-1 neighbour pairs");
        ctx->pairs = static_cast<int>(pairs_u);
        ctx->max_deg = ctx->mailbox[S_MAXDEG];
        if (ctx->max_deg >= (1u << 23))
            return fail(ctx, CVG_ERR_UNSUPPORTED, "an anchor has 2^23 or more neighbours");
        ctx->truncated = ctx->mailbox[S_TRUNC];
        return CVG_OK;
    };
    int fill_cap;
    unsigned long long pair_cap;
    if (spec) {
        fill_cap = ctx->spec_fill_cap;
        pair_cap = std::min({ctx->d_cand.bytes / sizeof(Cand), ctx->d_ids.bytes / sizeof(int),
                             ctx->d_d2.bytes / sizeof(double), ctx->d_pair_anchor.bytes / sizeof(int)});
        // below the fill's saturated offset, so that any saturated sum misses
        pair_cap = std::min<unsigned long long>(pair_cap, kFillSumMax - 1);
    } else {
        CVG_TRY(read_scalars(ctx));
        CVG_TRY(take_counts());
        const size_t np = static_cast<size_t>(ctx->pairs);
        CVG_TRY(ensure(ctx, ctx->d_cand, np * sizeof(Cand)));
        CVG_TRY(ensure(ctx, ctx->d_ids, np * sizeof(int)));
        CVG_TRY(ensure(ctx, ctx->d_d2, np * sizeof(double)));
        CVG_TRY(ensure(ctx, ctx->d_pair_anchor, np * sizeof(int)));
        fill_cap = static_cast<int>(std::max(ctx->max_deg, 1u));
        pair_cap = np;
    }

    // gather, sort, fill: one G-lane segment per anchor; each segment keeps
    // its candidates in LDS
    using FillFn = void (*)(int, const double *, const double *, const int *, const int *, Cells, int, double,
                            const int *, const int *, const double2 *, unsigned *, Cand *, int *, double *,
                            int *, int *, unsigned, int, const int *, const long long *, unsigned *, bool,
                            unsigned long long, unsigned long long *, unsigned, unsigned);
    const size_t per_seg = static_cast<size_t>(fill_seg_doubles(fill_cap, count_in_fill)) * sizeof(double);
    int G = ctx->fill_lanes;
    // fewer anchors per wave when the lists are long
    while (G < 64 && per_seg * (64 / G) * kFillWaves > ctx->lds_limit) G *= 2;
    int fill_waves = kFillWaves;
    while (fill_waves > 1 && per_seg * (64 / G) * fill_waves > ctx->lds_limit) fill_waves /= 2;
    if (per_seg * (64 / G) * fill_waves > ctx->lds_limit)
        return fail(ctx, CVG_ERR_UNSUPPORTED, "an anchor has too many neighbours for one workgroup's LDS");
    const FillFn fill = count_in_fill ? (G == 8    ? k_fill_csr<8, Cells, true>
                                         : G == 16 ? k_fill_csr<16, Cells, true>
                                         : G == 32 ? k_fill_csr<32, Cells, true>
                                                   : k_fill_csr<64, Cells, true>)
                                      : (G == 8    ? k_fill_csr<8, Cells, false>
                                         : G == 16 ? k_fill_csr<16, Cells, false>
                                         : G == 32 ? k_fill_csr<32, Cells, false>
                                                   : k_fill_csr<64, Cells, false>);
    const int anchors_per_block = (64 / G) * fill_waves;
    const unsigned fill_shmem = static_cast<unsigned>(per_seg * anchors_per_block);
    const int fill_blocks = (n + anchors_per_block - 1) / anchors_per_block;
    unsigned long long *lb = nullptr;
    unsigned epoch = 0;
    if (count_in_fill) {
        const size_t lb_bytes = static_cast<size_t>(fill_blocks) * sizeof(unsigned long long);
        CVG_TRY(ensure(ctx, ctx->d_fillstat, lb_bytes));
        if (++ctx->fill_epoch >= (1u << kFillEpochBits)) ctx->fill_epoch = 1;
        if (ctx->d_fillstat.ptr != ctx->fillstat_zero_ptr || ctx->d_fillstat.bytes != ctx->fillstat_zero_bytes ||
            ctx->fill_epoch == 1) {
            CVG_HIP(ctx, hipMemsetAsync(ctx->d_fillstat.ptr, 0, ctx->d_fillstat.bytes, st));
            ctx->fillstat_zero_ptr = ctx->d_fillstat.ptr;
            ctx->fillstat_zero_bytes = ctx->d_fillstat.bytes;
        }
        lb = as<unsigned long long>(ctx->d_fillstat);
        epoch = ctx->fill_epoch;
    }
    CVG_TRY(prepare_launch(ctx, reinterpret_cast<const void *>(fill), 64 * fill_waves, fill_shmem, nullptr));
    hipLaunchKernelGGL(fill, dim3(fill_blocks), dim3(64 * fill_waves), fill_shmem,
                       st, n, as<double>(ctx->d_x), as<double>(ctx->d_y), as<int>(ctx->d_cx), as<int>(ctx->d_cy),
                       cells, static_cast<int>(R), rad2, as<int>(ctx->d_sorted), as<int>(ctx->d_scell),
                       as<double2>(ctx->d_sxy), as<unsigned>(ctx->d_off), as<Cand>(ctx->d_cand),
                       as<int>(ctx->d_ids), as<double>(ctx->d_d2), as<int>(ctx->d_pair_anchor),
                       as<int>(ctx->d_tmp_ids), max_cand, fill_cap, num_maps > 1 ? as<int>(ctx->d_map) : nullptr,
                       as<long long>(ctx->d_eoff), scal, ctx->f32, pair_cap, lb, epoch, ctx->lookback_spins);
    hipLaunchKernelGGL(k_tie_sort, dim3(std::min(blocks_for(n), 64)), dim3(kBlock), 0, st, as<int>(ctx->d_tmp_ids),
                       scal, as<unsigned>(ctx->d_off), as<Cand>(ctx->d_cand), as<int>(ctx->d_ids),
                       as<double>(ctx->d_d2));
    CVG_HIP(ctx, hipGetLastError());
    if (spec) {
        // the sizes come back while the device already runs the curvelet
        // stage's per-edge trig
        CVG_HIP(ctx, hipMemcpyAsync(ctx->mailbox, ctx->d_scal.ptr, kScalarWords * sizeof(unsigned),
                                    hipMemcpyDeviceToHost, st));
        CVG_HIP(ctx, hipEventRecord(ctx->ev_read, st));
        if (ctx->early_params && !ctx->edge_trig_queued) {
            CVG_TRY(queue_edge_trig(ctx, *ctx->early_params));
            CVG_TRY(queue_pair_setup(ctx, *ctx->early_params, pair_cap));
        }
        CVG_HIP(ctx, hipEventSynchronize(ctx->ev_read));
        if (ctx->mailbox[S_BAD] || ctx->mailbox[S_MISS]) {
            // the queued per-pair launch found the miss and did nothing: the
            // graph built again queues it again (k_edge_trig's rows still hold)
            ctx->pair_setup_queued = false;
            return kSpecMiss;  // the exact path reports bad input
        }
        const int rc = take_counts();
        if (rc != CVG_OK) {
            ctx->pair_setup_queued = false;
            return rc;
        }
    }
    return CVG_OK;
}

// The dense cell grid g (counting sort into cells), then the scans.  With
// spec, g is a guess (the last map's grid): k_cell_count raises S_MISS for an
// edge whose window leaves it, and graph_scans runs on guessed sizes;
// kSpecMiss then asks for a build on the map's own grid.
int graph_dense(cvg_context *ctx, int n, GridDesc g, unsigned R, unsigned max_cand, double rad2, int num_maps, bool spec)
{
    hipStream_t st = ctx->stream;
    const size_t nd = static_cast<size_t>(n) * sizeof(double), ni = static_cast<size_t>(n) * sizeof(int);
    CVG_TRY(ensure(ctx, ctx->d_tmp_ids, ni));
    CVG_TRY(ensure(ctx, ctx->d_sorted, ni));
    CVG_TRY(ensure(ctx, ctx->d_sxy, 2 * nd));
    CVG_TRY(ensure(ctx, ctx->d_scell, ni));
    const int *rank = ctx->use_cellrank ? as<int>(ctx->d_cellrank) : nullptr;
    unsigned *scal = as<unsigned>(ctx->d_scal);
    ctx->gx0 = g.x0;
    ctx->gy0 = g.y0;
    ctx->gw = g.w;
    ctx->gh = g.h;
    const long long cells = static_cast<long long>(g.w) * g.h;
    CVG_TRY(ensure(ctx, ctx->d_grid, static_cast<size_t>(2 * cells + 1) * sizeof(unsigned)));
    unsigned *cnt = as<unsigned>(ctx->d_grid);
    unsigned *start = cnt + cells;
    // k_cell_scatter leaves every count it decrements at zero, so the
    // counts of a grid no larger than the last one in the same context
    // buffer start zeroed (not tracked for the no-alloc scratch, whose
    // regions other buffers reuse)
    if (!spec) {  // (on the guessed grid k_prep_edges counted the cells)
        const bool zeroed =
            !ctx->arena && ctx->d_grid.ptr == ctx->grid_zero_ptr && cells <= ctx->grid_zero_cells;
        ctx->grid_zero_ptr = nullptr;
        if (!zeroed) CVG_HIP(ctx, hipMemsetAsync(cnt, 0, static_cast<size_t>(cells) * sizeof(unsigned), st));
        hipLaunchKernelGGL(k_cell_count, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_cx),
                           as<int>(ctx->d_cy), g, cnt, static_cast<int>(R), static_cast<unsigned *>(nullptr));
    }
    CVG_TRY(exclusive_scan_1pass(ctx, cnt, cells, start, ctx->d_tiles));
    hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_cx),
                       as<int>(ctx->d_cy), g, start, cnt, as<int>(ctx->d_tmp_ids),
                       spec ? static_cast<const unsigned *>(scal + S_MISS) : nullptr);
    if (!ctx->arena) {
        ctx->grid_zero_ptr = ctx->d_grid.ptr;
        ctx->grid_zero_cells = cells;
    }
    hipLaunchKernelGGL(k_cell_order, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_cx),
                       as<int>(ctx->d_cy), g, start, as<int>(ctx->d_tmp_ids), rank, as<double>(ctx->d_x),
                       as<double>(ctx->d_y), as<int>(ctx->d_sorted), as<int>(ctx->d_scell),
                       as<double2>(ctx->d_sxy), spec ? static_cast<const unsigned *>(scal + S_MISS) : nullptr);
    CVG_HIP(ctx, hipGetLastError());
    DenseCells cells_d{start, g.x0, g.y0, g.w};
    return graph_scans(ctx, cells_d, n, R, max_cand, rad2, num_maps, spec);
}

int build_graph(cvg_context *ctx, int n, int sz, const double *d_edges, double rad, unsigned R, unsigned max_cand,
                int num_maps = 1, const long long *eoff = nullptr)
{
    ctx->graph_ready = false;
    ctx->curvelets_ready = false;
    ctx->counts_pending = false;
    ctx->edge_trig_queued = false;  // a new graph: its per-edge trig is queued again
    ctx->pair_setup_queued = false;
    if (n <= 0 || sz < 4 || d_edges == nullptr)  // preprocess.hpp:197
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "build_neighbor_csr_graph: need num_edges > 0, sz_edge_data >= 4");
    if (R > 4096) return fail(ctx, CVG_ERR_UNSUPPORTED, "neighbor_radius above 4096 cells");
    if (max_cand == 0) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "max_candidates must be positive");
    if (num_maps < 1) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "a batch needs at least one map");
    ctx->num_maps = num_maps;
    ctx->map_eoff.assign(static_cast<size_t>(num_maps) + 1, 0);
    if (eoff) {
        for (int m = 0; m <= num_maps; ++m) ctx->map_eoff[m] = eoff[m];
    } else {
        ctx->map_eoff[1] = n;
    }
    ctx->n = n;
    ctx->sz = sz;
    ctx->last_edges = d_edges;
    ctx->rad = rad;
    ctx->nbr_radius = R;
    ctx->max_cand = max_cand;
    hipStream_t st = ctx->stream;
    const size_t nd = static_cast<size_t>(n) * sizeof(double), ni = static_cast<size_t>(n) * sizeof(int);
    CVG_TRY(ensure(ctx, ctx->d_x, nd));
    CVG_TRY(ensure(ctx, ctx->d_y, nd));
    CVG_TRY(ensure(ctx, ctx->d_th, nd));
    CVG_TRY(ensure(ctx, ctx->d_cx, ni));
    CVG_TRY(ensure(ctx, ctx->d_cy, ni));
    CVG_TRY(ensure(ctx, ctx->d_scal, kScalarBufWords * sizeof(unsigned)));
    unsigned *scal = as<unsigned>(ctx->d_scal);
    const int prep_blocks = blocks_for(n);
    CVG_TRY(ensure(ctx, ctx->d_part, static_cast<size_t>(prep_blocks) * 5 * sizeof(int)));
    // A single map first runs on the last single map's grid and buffer sizes
    // with no host read until the fill is done (one round trip instead of
    // three); when they do not hold, it is built again on its own grid.
    const bool spec = num_maps == 1 && !ctx->force_sparse && !ctx->arena && !ctx->spec_off && ctx->spec_valid &&
                      ctx->spec_R == R;
    auto prep = [&](bool spec_on) -> int {
        // on the guessed grid the cells are counted here (no k_cell_count)
        unsigned *cnt = nullptr;
        if (spec_on) {
            const long long cells = static_cast<long long>(ctx->spec_g.w) * ctx->spec_g.h;
            CVG_TRY(ensure(ctx, ctx->d_grid, static_cast<size_t>(2 * cells + 1) * sizeof(unsigned)));
            cnt = as<unsigned>(ctx->d_grid);
            const bool zeroed = ctx->d_grid.ptr == ctx->grid_zero_ptr && cells <= ctx->grid_zero_cells;
            ctx->grid_zero_ptr = nullptr;
            if (!zeroed) CVG_HIP(ctx, hipMemsetAsync(cnt, 0, static_cast<size_t>(cells) * sizeof(unsigned), st));
        }
        hipLaunchKernelGGL(k_prep_edges, dim3(prep_blocks), dim3(kBlock), 0, st, n, sz, d_edges,
                           as<double>(ctx->d_x), as<double>(ctx->d_y), as<double>(ctx->d_th), as<int>(ctx->d_cx),
                           as<int>(ctx->d_cy), as<int>(ctx->d_part), ctx->spec_g, cnt, static_cast<int>(R));
        hipLaunchKernelGGL(k_bbox_finish, dim3(1), dim3(1024), 0, st, prep_blocks, as<int>(ctx->d_part), scal,
                           spec_on ? 1 : 0, ctx->spec_g, static_cast<int>(R));
        CVG_HIP(ctx, hipGetLastError());
        return CVG_OK;
    };
    CVG_TRY(prep(spec));
    // preprocess.hpp:221: rad * rad in T
    const double rad2 = ctx->f32 ? static_cast<double>(static_cast<float>(rad) * static_cast<float>(rad)) : rad * rad;
    if (spec) {
        CVG_TRY(ensure(ctx, ctx->d_eoff, 2 * sizeof(long long)));
        const int rc = graph_dense(ctx, n, ctx->spec_g, R, max_cand, rad2, 1, true);
        if (rc == CVG_OK) {
            ctx->graph_ready = true;
            return CVG_OK;
        }
        if (rc != kSpecMiss) return rc;
        ctx->spec_valid = false;
        ctx->grid_zero_ptr = nullptr;  // the guessed grid's counts may be partial
        CVG_TRY(prep(false));  // the guess may have moved some cells: the map's own cells again
    }
    CVG_TRY(read_scalars(ctx));
    const unsigned *mb = ctx->mailbox;
    if (mb[S_BAD])
        return fail(ctx, CVG_ERR_BAD_INPUT,
                    "edge values out of range: need finite x, y with |x|,|y| < 2^31 - 2^13 and a finite theta");
    int minx = static_cast<int>(mb[S_MINX]), maxx = static_cast<int>(mb[S_MAXX]);
    int miny = static_cast<int>(mb[S_MINY]), maxy = static_cast<int>(mb[S_MAXY]);
    // the per-map edge ranges (kernels read them only for batches)
    CVG_TRY(ensure(ctx, ctx->d_eoff, (static_cast<size_t>(num_maps) + 1) * sizeof(long long)));
    if (num_maps > 1)
        CVG_HIP(ctx, hipMemcpyAsync(ctx->d_eoff.ptr, ctx->map_eoff.data(), (static_cast<size_t>(num_maps) + 1) *
                                    sizeof(long long), hipMemcpyHostToDevice, st));
    if (num_maps > 1) {
        // per-map boxes, then a shelf packing of the boxes into one grid
        CVG_TRY(ensure(ctx, ctx->d_map, ni));
        CVG_TRY(ensure(ctx, ctx->d_mapbox, static_cast<size_t>(num_maps) * 4 * sizeof(int)));
        hipLaunchKernelGGL(k_map_ids, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, num_maps,
                           as<long long>(ctx->d_eoff), as<int>(ctx->d_map));
        hipLaunchKernelGGL(k_map_box_init, dim3(blocks_for(4LL * num_maps)), dim3(kBlock), 0, st, num_maps,
                           as<int>(ctx->d_mapbox));
        hipLaunchKernelGGL(k_map_box, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_map),
                           as<int>(ctx->d_cx), as<int>(ctx->d_cy), as<int>(ctx->d_mapbox));
        std::vector<int> box(static_cast<size_t>(num_maps) * 4);
        CVG_HIP(ctx, hipMemcpyAsync(box.data(), ctx->d_mapbox.ptr, box.size() * sizeof(int), hipMemcpyDeviceToHost,
                                    st));
        CVG_HIP(ctx, hipStreamSynchronize(st));
        std::vector<int> shift(static_cast<size_t>(num_maps) * 2, 0);
        const long long gap = 2LL * R + 1;
        long long shelf_w = 8192;
        for (int m = 0; m < num_maps; ++m) {
            if (box[4 * m] > box[4 * m + 1]) continue;  // empty map
            shelf_w = std::max(shelf_w, static_cast<long long>(box[4 * m + 1]) - box[4 * m] + 1);
        }
        std::vector<int> order(num_maps);
        for (int m = 0; m < num_maps; ++m) order[m] = m;
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
            return static_cast<long long>(box[4 * p + 3]) - box[4 * p + 2] >
                   static_cast<long long>(box[4 * q + 3]) - box[4 * q + 2];
        });
        long long cx0 = 0, cy0 = 0, shelf_h = 0, used_w = 0;
        for (int m : order) {
            if (box[4 * m] > box[4 * m + 1]) continue;
            const long long w_m = static_cast<long long>(box[4 * m + 1]) - box[4 * m] + 1;
            const long long h_m = static_cast<long long>(box[4 * m + 3]) - box[4 * m + 2] + 1;
            if (cx0 > 0 && cx0 + w_m > shelf_w) {
                cy0 += shelf_h + gap;
                cx0 = 0;
                shelf_h = 0;
            }
            if (cy0 + h_m > (1LL << 30) || cx0 + w_m > (1LL << 30))
                return fail(ctx, CVG_ERR_UNSUPPORTED, "batch too large for one cell grid");
            const long long sx_m = cx0 - box[4 * m], sy_m = cy0 - box[4 * m + 2];
            if (sx_m > INT_MAX || sx_m < INT_MIN || sy_m > INT_MAX || sy_m < INT_MIN)
                return fail(ctx, CVG_ERR_UNSUPPORTED, "batch too large for one cell grid");
            shift[2 * m] = static_cast<int>(sx_m);
            shift[2 * m + 1] = static_cast<int>(sy_m);
            used_w = std::max(used_w, cx0 + w_m);
            shelf_h = std::max(shelf_h, h_m);
            cx0 += w_m + gap;
        }
        minx = 0;
        miny = 0;
        maxx = static_cast<int>(used_w - 1);
        maxy = static_cast<int>(cy0 + shelf_h - 1);
        CVG_HIP(ctx, hipMemcpyAsync(ctx->d_mapbox.ptr, shift.data(), shift.size() * sizeof(int),
                                    hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_shift_cells, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_map),
                           as<int>(ctx->d_mapbox), as<int>(ctx->d_cx), as<int>(ctx->d_cy));
        CVG_HIP(ctx, hipGetLastError());
    }
    const int x0 = minx - static_cast<int>(R);
    const int y0 = miny - static_cast<int>(R);
    const long long w = static_cast<long long>(maxx) - minx + 1 + 2LL * R;
    const long long h = static_cast<long long>(maxy) - miny + 1 + 2LL * R;
    ctx->gx0 = x0;
    ctx->gy0 = y0;
    CVG_TRY(ensure(ctx, ctx->d_tmp_ids, ni));
    CVG_TRY(ensure(ctx, ctx->d_sorted, ni));
    CVG_TRY(ensure(ctx, ctx->d_sxy, 2 * nd));
    CVG_TRY(ensure(ctx, ctx->d_scell, ni));
    const int *rank = ctx->use_cellrank ? as<int>(ctx->d_cellrank) : nullptr;
    if (w <= kMaxGridCells && h <= kMaxGridCells && w * h <= kMaxGridCells && !ctx->force_sparse) {
        GridDesc g;
        g.x0 = x0;
        g.y0 = y0;
        g.w = static_cast<int>(w);
        g.h = static_cast<int>(h);
        CVG_TRY(graph_dense(ctx, n, g, R, max_cand, rad2, num_maps, false));
        if (num_maps == 1 && !ctx->arena) {
            // the next single map first tries this grid (with a margin) and
            // these buffers, without reading its sizes first
            const int pad = 16;
            GridDesc sg{g.x0 - pad, g.y0 - pad, g.w + 2 * pad, g.h + 2 * pad};
            ctx->spec_valid = static_cast<long long>(sg.w) * sg.h <= kMaxGridCells;
            ctx->spec_g = sg;
            ctx->spec_R = R;
            ctx->spec_fill_cap = static_cast<int>(std::max(ctx->max_deg, 1u));
        }
    } else {
        // sorted linear cell keys: stable radix sorts of (key, id), by the
        // within-cell rank first when the reference's order is replayed, so
        // each cell's edges follow edge id (or rank) order as in the dense grid
        ctx->gw = 0;
        ctx->gh = 0;
        CVG_TRY(ensure(ctx, ctx->d_skey, 4 * static_cast<size_t>(n) * sizeof(unsigned long long)));
        unsigned long long *key = as<unsigned long long>(ctx->d_skey);
        unsigned long long *key2 = key + n;
        int *rkey = reinterpret_cast<int *>(key + 2 * n);
        int *rkey2 = rkey + n;
        int *id = reinterpret_cast<int *>(key + 3 * n);
        int *id2 = as<int>(ctx->d_tmp_ids);
        hipLaunchKernelGGL(k_sparse_keys, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, as<int>(ctx->d_cx),
                           as<int>(ctx->d_cy), x0, y0, rank, key, rkey, id);
        CVG_HIP(ctx, hipGetLastError());
        int key_bits = 33;  // the column offsets, then as many row bits as the rows need
        while (key_bits < 64 && (1LL << (key_bits - 32)) < h) ++key_bits;
        size_t tmp_bytes = 0, t2 = 0;
        CVG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2, id, id2, n, 0, key_bits, st));
        CVG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t2, rkey, rkey2, id, id2, n, 0, 32, st));
        tmp_bytes = std::max(tmp_bytes, t2);
        CVG_TRY(ensure(ctx, ctx->d_sorttmp, tmp_bytes));
        if (rank) {
            CVG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->d_sorttmp.ptr, tmp_bytes, rkey, rkey2, id, id2, n, 0,
                                                            32, st));
            hipLaunchKernelGGL(k_gather_keys, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, key, id2, key2);
            CVG_HIP(ctx, hipGetLastError());
            // (key2, id2) in rank order -> (key, id) sorted by key, stable
            CVG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->d_sorttmp.ptr, tmp_bytes, key2, key, id2, id, n, 0,
                                                            key_bits, st));
        } else {
            CVG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->d_sorttmp.ptr, tmp_bytes, key, key2, id, id2, n, 0,
                                                            key_bits, st));
            CVG_HIP(ctx, hipMemcpyAsync(key, key2, static_cast<size_t>(n) * sizeof(unsigned long long),
                                        hipMemcpyDeviceToDevice, st));
            CVG_HIP(ctx, hipMemcpyAsync(id, id2, ni, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(k_sparse_fill, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, key, id, as<double>(ctx->d_x),
                           as<double>(ctx->d_y), as<int>(ctx->d_sorted), as<int>(ctx->d_scell),
                           as<double2>(ctx->d_sxy));
        CVG_HIP(ctx, hipGetLastError());
        SparseCells cells_s{key, n, x0, y0};
        CVG_TRY(graph_scans(ctx, cells_s, n, R, max_cand, rad2, num_maps, false));
    }
    ctx->graph_ready = true;
    return CVG_OK;
}

struct Derived {
    double dt, dx, dx2, close_thr, max_k, nrad;
    unsigned kx, kt, K, G;
    int cpl;
    std::vector<double> off_x, off_t, rank_ox;
    std::vector<int> rank_j, rank_cell, rank_elig;
    int rank_of_cell0;
};

int derive(cvg_context *ctx, const cvg_params &p, Derived &d)
{
    if (!(p.sx > 0) || !(p.st > 0))
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "sx and st must be positive");
    d.dt = (p.dt_deg / 180.0) * 3.14159265358979323846;  // main.cpp:24
    d.dx = p.dx;
    d.dx2 = p.dx * p.dx;                                   // cpu_curvelet.hpp:654
    d.close_thr = 4 * p.dx * p.dx;                         // :682, (4*dx)*dx
    d.max_k = p.max_k;
    d.nrad = p.nrad;
    const double fx = 2 * std::floor(p.dx / p.sx + 0.5) + 1;  // :73-74
    const double ft = 2 * std::floor(d.dt / p.st + 0.5) + 1;
    if (!(fx >= 1 && ft >= 1) || fx * ft > 256)
        return fail(ctx, CVG_ERR_UNSUPPORTED, "bundle grid must have between 1 and 256 cells (dx/sx, dt/st)");
    d.kx = static_cast<unsigned>(fx);
    d.kt = static_cast<unsigned>(ft);
    d.K = d.kx * d.kt;
    d.cpl = d.K <= 64 ? 1 : (d.K <= 128 ? 2 : 4);
    d.G = p.group_max_sz;
    d.off_x.resize(d.kx);
    d.off_t.resize(d.kt);
    for (unsigned i = 0; i < d.kx; ++i) d.off_x[i] = p.sx * ((double)i - (double(d.kx) - 1.0) / 2.0);  // :697
    for (unsigned j = 0; j < d.kt; ++j) d.off_t[j] = p.st * ((double)j - (double(d.kt) - 1.0) / 2.0);  // :698
    std::vector<int> order(d.K);
    std::vector<double> dist(d.K);
    for (unsigned c = 0; c < d.K; ++c) {
        const double ox = d.off_x[c / d.kt], ot = d.off_t[c % d.kt];
        dist[c] = ox * ox + ot * ot;  // fill_curvelet_info's d, :211
        order[c] = static_cast<int>(c);
    }
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return dist[u] < dist[v]; });
    d.rank_ox.resize(d.K);
    d.rank_j.resize(d.K);
    d.rank_cell.resize(d.K);
    d.rank_elig.resize(d.K);
    for (unsigned r = 0; r < d.K; ++r) {
        const int c = order[r];
        d.rank_cell[r] = c;
        d.rank_ox[r] = d.off_x[c / d.kt];
        d.rank_j[r] = c % d.kt;
        d.rank_elig[r] = dist[c] < 100.0 ? 1 : 0;
        if (c == 0) d.rank_of_cell0 = static_cast<int>(r);
    }
    return CVG_OK;
}

// Per-edge trig of the fused stage (k_edge_trig), which needs only the
// edges' orientations and the parameters: queued by build_curvelets, or
// earlier, while the graph stage's last host read is in flight.
int queue_edge_trig(cvg_context *ctx, const cvg_params &p)
{
    Derived d;
    CVG_TRY(derive(ctx, p, d));
    const int n = ctx->n;
    if (static_cast<unsigned long long>(n) * (d.kt + 3) > 0xffffffffull)
        return fail(ctx, CVG_ERR_UNSUPPORTED, "more than 2^32 (edge, angle) items in one run");
    const int ntrig = cvg::TRIG_GRID + 2 * static_cast<int>(d.kt);
    CVG_TRY(ensure(ctx, ctx->d_trig, static_cast<size_t>(ntrig) * n * sizeof(double)));
    CVG_TRY(ensure(ctx, ctx->d_tdd, static_cast<size_t>(8 + 4 * d.kt) * n * sizeof(double)));
    hipLaunchKernelGGL(k_edge_trig, dim3(blocks_for(static_cast<long long>(n) * (d.kt + 3))), dim3(kBlock), 0,
                       ctx->stream, n, as<double>(ctx->d_th), d.dt, static_cast<int>(d.kt), p.st,
                       as<double>(ctx->d_trig), as<double>(ctx->d_tdd), as<unsigned>(ctx->d_scal));
    CVG_HIP(ctx, hipGetLastError());
    ctx->edge_trig_queued = true;
    return CVG_OK;
}

// The per-pair records and sines (k_pair_trig with the record, Kt = 7),
// queued behind k_edge_trig while the graph stage's count read is in flight:
// the launch covers the CSR buffers' capacity pair_cap and takes the pair
// count from the device.  Only where the pair buffers already hold that
// capacity: nothing is allocated before the counts are known.
int queue_pair_setup(cvg_context *ctx, const cvg_params &p, unsigned long long pair_cap)
{
    Derived d;
    CVG_TRY(derive(ctx, p, d));
    constexpr int kt = 7;
    if (d.kt != kt || ctx->arena || pair_cap == 0 || pair_cap > static_cast<unsigned long long>(INT_MAX))
        return CVG_OK;
    // after a much larger map the capacity would launch mostly idle
    // workgroups: then the launch waits for the count (ctx->pairs still holds
    // the last graph's)
    if (pair_cap > 2ULL * static_cast<unsigned long long>(std::max(ctx->pairs, 0)) + 65536ULL) return CVG_OK;
    if (ctx->d_cls.bytes < cls_bytes(pair_cap) || ctx->d_prec.bytes < pair_cap * 8 * sizeof(double) ||
        ctx->d_psin.bytes < pair_cap * kt * 2 * sizeof(double))
        return CVG_OK;
    const int cap = static_cast<int>(pair_cap);
    const int n = ctx->n;
    const PrepArgs P{as<double>(ctx->d_x), as<double>(ctx->d_y), as<double>(ctx->d_trig), d.close_thr,
                     as<unsigned char>(ctx->d_cls), as<double>(ctx->d_prec)};
    const auto fn = k_pair_trig<kt, true>;
    hipLaunchKernelGGL(fn, dim3(blocks_for(cap)), dim3(kBlock), kBlock * kt * sizeof(double2),
                       ctx->stream, cap, n, kt, as<int>(ctx->d_pair_anchor), as<int>(ctx->d_ids),
                       static_cast<const unsigned char *>(as<unsigned char>(ctx->d_cls)), as<double>(ctx->d_th), p.st,
                       d.dt, as<double>(ctx->d_tdd), as<double>(ctx->d_psin), P,
                       static_cast<const unsigned *>(as<unsigned>(ctx->d_scal)));
    CVG_HIP(ctx, hipGetLastError());
    ctx->pair_setup_queued = true;
    return CVG_OK;
}

// The parameter tables to d_tables, unless the same bytes are already there
// (repeated runs with the same parameters skip the copy; the no-alloc
// scratch's regions are reused by other buffers, so there it always copies).
int upload_tables(cvg_context *ctx, const std::vector<char> &blob)
{
    CVG_TRY(ensure(ctx, ctx->d_tables, blob.size()));
    if (!ctx->arena && ctx->d_tables.ptr == ctx->tables_ptr && blob == ctx->tables_blob) return CVG_OK;
    CVG_HIP(ctx, hipMemcpyAsync(ctx->d_tables.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream));
    ctx->tables_ptr = ctx->arena ? nullptr : ctx->d_tables.ptr;
    ctx->tables_blob = blob;
    return CVG_OK;
}

template <typename T>
size_t put(std::vector<char> &blob, const std::vector<T> &v)
{
    const size_t align = 16;
    size_t off = (blob.size() + align - 1) / align * align;
    blob.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// Page-locked host buffer for per-map row offsets.
int ensure_host(cvg_context *ctx, size_t bytes)
{
    if (ctx->h_maprow_bytes >= bytes) return CVG_OK;
    if (ctx->h_maprow) (void)hipHostFree(ctx->h_maprow);
    ctx->h_maprow = nullptr;
    ctx->h_maprow_bytes = 0;
    if (hipHostMalloc(reinterpret_cast<void **>(&ctx->h_maprow), bytes, hipHostMallocDefault) != hipSuccess) {
        ctx->h_maprow = nullptr;
        return fail(ctx, CVG_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
    }
    ctx->h_maprow_bytes = bytes;
    return CVG_OK;
}

// Comparator of the reference's spatial-index sort (preprocess.hpp:90-92):
// edge ids by packed cell key.
struct CellKeyLess {
    const long long *key;
    __host__ __device__ bool operator()(int a, int b) const { return key[a] < key[b]; }
};

// The reference's within-cell order (build_spatial_index_cpu,
// preprocess.hpp:78-92) for one map: its edge ids sorted by
// pack_cell_key(lrint(x), lrint(y)) (:41-45) with libstdc++'s introsort; each
// edge's position inside its cell's run goes to rank[i].
void cell_ranks(int n, const double *x, const double *y, int *rank)
{
    std::vector<long long> key(static_cast<size_t>(std::max(n, 0)));
    std::vector<int> perm(key.size());
    for (int i = 0; i < n; ++i) {
        const int cx = static_cast<int>(std::lrint(x[i]));
        const int cy = static_cast<int>(std::lrint(y[i]));
        key[i] = static_cast<long long>((static_cast<unsigned long long>(static_cast<long long>(cy)) << 32) |
                                        static_cast<unsigned long long>(static_cast<unsigned>(cx)));
        perm[i] = i;
    }
    cvg::stl_sort(perm.data(), n, CellKeyLess{key.data()});
    int run = 0;
    for (int q = 0; q < n; ++q) {
        if (q > 0 && key[perm[q]] != key[perm[q - 1]]) run = q;
        rank[perm[q]] = q - run;
    }
}

// Replays that order for every map of the last graph into d_cellrank; the
// next graph build orders cells by it.
int replay_cell_order(cvg_context *ctx)
{
    const int n = ctx->n;
    hipStream_t st = ctx->stream;
    std::vector<double> hx(static_cast<size_t>(n)), hy(static_cast<size_t>(n));
    CVG_HIP(ctx, hipMemcpyAsync(hx.data(), ctx->d_x.ptr, hx.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CVG_HIP(ctx, hipMemcpyAsync(hy.data(), ctx->d_y.ptr, hy.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CVG_HIP(ctx, hipStreamSynchronize(st));
    std::vector<int> rank(static_cast<size_t>(n), 0);
    for (int m = 0; m < ctx->num_maps; ++m) {
        const long long lo = ctx->map_eoff[m], hi = ctx->map_eoff[m + 1];
        cell_ranks(static_cast<int>(hi - lo), hx.data() + lo, hy.data() + lo, rank.data() + lo);
    }
    CVG_TRY(ensure(ctx, ctx->d_cellrank, rank.size() * sizeof(int)));
    CVG_HIP(ctx, hipMemcpyAsync(ctx->d_cellrank.ptr, rank.data(), rank.size() * sizeof(int), hipMemcpyHostToDevice, st));
    CVG_HIP(ctx, hipStreamSynchronize(st));
    ctx->use_cellrank = true;
    return CVG_OK;
}

// Rebuild the last graph with the reference's within-cell order.
int rebuild_graph_exact(cvg_context *ctx)
{
    CVG_TRY(replay_cell_order(ctx));
    const std::vector<long long> eoff = ctx->map_eoff;
    return build_graph(ctx, ctx->n, ctx->sz, ctx->last_edges, ctx->rad, ctx->nbr_radius, ctx->max_cand,
                       ctx->num_maps, ctx->num_maps > 1 ? eoff.data() : nullptr);
}

int build_curvelets(cvg_context *ctx, const cvg_params &p);

// The counts build_curvelets left in flight: waits for the stream once.  When
// some anchor's neighbour list depended on the within-cell order, the graph
// and curvelets are built again with the reference's order first.
int finish_curvelets(cvg_context *ctx)
{
    if (!ctx->counts_pending) return CVG_OK;
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->counts_pending = false;
    if (ctx->mailbox[S_AMBIG] > 0 && !ctx->use_cellrank) {
        const cvg_params p = ctx->cur_params;
        ++ctx->rebuilds;
        CVG_TRY(rebuild_graph_exact(ctx));
        CVG_TRY(build_curvelets(ctx, p));
        CVG_TRY(record_timing(ctx, ctx->ev[2], ctx->stream));
        CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->counts_pending = false;
    }
    const unsigned C = ctx->pairs > 0 ? ctx->mailbox[S_ROWS] : 0u;
    ctx->valid_pairs = ctx->mailbox[S_VALID];
    ctx->intersections = static_cast<unsigned long long>(ctx->mailbox[S_COUNT]) |
                         (static_cast<unsigned long long>(ctx->mailbox[S_COUNT + 1]) << 32);
    ctx->ambiguous = ctx->mailbox[S_AMBIG];
    ctx->num_curvelets = C;
    if (ctx->num_maps > 1) {
        ctx->map_rows.assign(ctx->h_maprow, ctx->h_maprow + ctx->num_maps + 1);
    } else {
        ctx->map_rows.assign(2, 0);
        ctx->map_rows[1] = C;
    }
    return CVG_OK;
}

// compact_curvelet_output + record_curvelet_chain (cpu_curvelet.hpp:258-303):
// row offsets from the per-anchor accepted counts, then every accepted
// staging row to its row (R carries the grid and info parameters), and the
// counts back with the stream.  Outputs are sized by their bound (one row per
// CSR pair at most), so the record runs without a host round trip; the
// counts are read at the first query (finish_curvelets).
int record_rows(cvg_context *ctx, const cvg_params &p, RecordArgs R, int st_stride)
{
    hipStream_t st = ctx->stream;
    const int n = ctx->n;
    const int pairs = ctx->pairs;
    const size_t np = static_cast<size_t>(std::max(pairs, 1));
    CVG_TRY(exclusive_scan_1pass(ctx, as<unsigned>(ctx->d_acc), n, as<unsigned>(ctx->d_rowoff), ctx->d_tiles));
    ctx->width = static_cast<unsigned>(R.G) + 1;
    unsigned *out_chains = ctx->caller_chains;
    double *out_info = ctx->caller_info;
    if (!out_chains) {
        CVG_TRY(ensure(ctx, ctx->d_chains, np * ctx->width * sizeof(unsigned)));
        CVG_TRY(ensure(ctx, ctx->d_info, np * kInfoWidth * sizeof(double)));
        out_chains = as<unsigned>(ctx->d_chains);
        out_info = as<double>(ctx->d_info);
    }
    R.pairs = pairs;
    R.pair_anchor = as<int>(ctx->d_pair_anchor);
    R.map_of = ctx->num_maps > 1 ? as<int>(ctx->d_map) : nullptr;
    R.map_eoff = as<long long>(ctx->d_eoff);
    R.x = as<double>(ctx->d_x);
    R.y = as<double>(ctx->d_y);
    R.th = as<double>(ctx->d_th);
    R.st_rank = as<int>(ctx->d_st_rank);
    R.row_off = as<unsigned>(ctx->d_rowoff);
    R.st_ids = as<int>(ctx->d_st_ids);
    R.st_meta = as<unsigned>(ctx->d_st_meta);
    R.st_k = as<double>(ctx->d_st_k);
    R.st_stride = st_stride;
    R.chains = out_chains;
    R.info = out_info;
    R.max_rows = ctx->caller_chains ? ctx->caller_rows : UINT_MAX;
    if (pairs > 0) {
        CVG_TRY(ensure(ctx, ctx->d_rowslot, np * sizeof(int)));
        int *slots = as<int>(ctx->d_rowslot);
        hipLaunchKernelGGL(k_row_slots, dim3(blocks_for(pairs)), dim3(kBlock), 0, st, R, slots,
                           static_cast<const unsigned *>(as<unsigned>(ctx->d_rowoff) + n), as<unsigned>(ctx->d_scal));
        // a resident grid strides over the rows (their count is on the device)
        const auto rec = R.f32 ? k_record<true> : k_record<false>;
        int occ = 0;
        CVG_TRY(prepare_launch(ctx, reinterpret_cast<const void *>(rec), kBlock, 0, &occ));
        const long long resident = static_cast<long long>(std::max(occ, 1)) * ctx->num_cu;
        const int grid = static_cast<int>(std::min(resident, static_cast<long long>(blocks_for(pairs))));
        hipLaunchKernelGGL(rec, dim3(grid), dim3(kBlock), 0, st, R, static_cast<const int *>(slots),
                           static_cast<const unsigned *>(as<unsigned>(ctx->d_rowoff) + n));
    }
    CVG_HIP(ctx, hipGetLastError());
    // the counts (the row count in S_ROWS when there are pairs; none otherwise)
    CVG_HIP(ctx, hipMemcpyAsync(ctx->mailbox, ctx->d_scal.ptr, kScalarWords * sizeof(unsigned),
                                hipMemcpyDeviceToHost, st));
    if (ctx->num_maps > 1) {
        const int M = ctx->num_maps;
        CVG_TRY(ensure(ctx, ctx->d_maprow, (static_cast<size_t>(M) + 1) * sizeof(unsigned)));
        CVG_TRY(ensure_host(ctx, (static_cast<size_t>(M) + 1) * sizeof(unsigned)));
        hipLaunchKernelGGL(k_map_rows, dim3(blocks_for(M + 1LL)), dim3(kBlock), 0, st, M, as<long long>(ctx->d_eoff),
                           as<unsigned>(ctx->d_rowoff), as<unsigned>(ctx->d_maprow));
        CVG_HIP(ctx, hipMemcpyAsync(ctx->h_maprow, ctx->d_maprow.ptr, (static_cast<size_t>(M) + 1) * sizeof(unsigned),
                                    hipMemcpyDeviceToHost, st));
    }
    ctx->counts_pending = true;
    ctx->cur_params = p;
    ctx->curvelets_ready = true;
    return CVG_OK;
}

// The curvelet stage on k_curvelets_generic: the float instantiation
// (run_curvelet<float>: the parameters cast to float as main.cpp:22-30 casts
// them, the bundle grid of init_curvelet_storage (:73-74) from float dx / sx
// and dt / st), or a double map outside the fused kernel's range; then the
// common record.
template <typename T>
int build_curvelets_generic(cvg_context *ctx, const cvg_params &p)
{
    constexpr bool kFloat = std::is_same<T, float>::value;
    if (!ctx->graph_ready) return fail(ctx, CVG_ERR_STATE, "build the neighbour graph first");
    ctx->curvelets_ready = false;
    if (!(p.sx > 0) || !(p.st > 0)) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "sx and st must be positive");
    hipStream_t st = ctx->stream;
    const int n = ctx->n;
    const int pairs = ctx->pairs;
    const T dx = static_cast<T>(p.dx), max_k = static_cast<T>(p.max_k), sx = static_cast<T>(p.sx),
            st_ = static_cast<T>(p.st);
    const T dt = kFloat ? static_cast<T>(static_cast<float>(p.dt_deg / 180.0) * static_cast<float>(3.14159265358979323846))
                        : static_cast<T>((p.dt_deg / 180.0) * 3.14159265358979323846);  // main.cpp:24
    const double fx = 2 * std::floor(dx / sx + 0.5) + 1;  // :73-74
    const double ft = 2 * std::floor(dt / st_ + 0.5) + 1;
    if (!(fx >= 1 && ft >= 1) || fx * ft > 256)
        return fail(ctx, CVG_ERR_UNSUPPORTED, "bundle grid must have between 1 and 256 cells (dx/sx, dt/st)");
    const int kx = static_cast<int>(fx), kt = static_cast<int>(ft), K = kx * kt;
    const int G = static_cast<int>(p.group_max_sz);
    const int max_deg = static_cast<int>(std::max(ctx->max_deg, 1u));
    // no chain is longer than its anchor plus all of its neighbours
    const int cap = std::max(1, static_cast<int>(std::min<long long>(G, static_cast<long long>(max_deg) + 1)));
    if (cap > 0x7fff) return fail(ctx, CVG_ERR_UNSUPPORTED, "chains of more than 32767 edges");
    std::vector<T> off_x(kx), off_t(kt);
    std::vector<float> inf_x(kx), inf_t(kt);
    std::vector<double> dist(K);
    for (int i = 0; i < kx; ++i) {
        off_x[i] = static_cast<T>(sx * (static_cast<T>(i) - (static_cast<T>(kx) - 1.0) / 2.0));  // :697
        inf_x[i] = static_cast<float>(sx) * (static_cast<float>(i) - (static_cast<float>(kx) - 1.0f) / 2.0f);  // :209
    }
    for (int j = 0; j < kt; ++j) {
        off_t[j] = static_cast<T>(st_ * (static_cast<T>(j) - (static_cast<T>(kt) - 1.0) / 2.0));  // :698
        inf_t[j] = static_cast<float>(st_) * (static_cast<float>(j) - (static_cast<float>(kt) - 1.0f) / 2.0f);  // :210
    }
    // fill_curvelet_info's d per cell (:211): in T from the info offsets (the
    // float ones for T = float; for T = double they equal off_x, off_t)
    std::vector<int> order(K), cell_rank(K), rank_cell(K);
    for (int c = 0; c < K; ++c) {
        if (kFloat) {
            const float a = inf_x[c / kt], b = inf_t[c % kt];
            dist[c] = static_cast<double>(a * a + b * b);
        } else {
            const double a = static_cast<double>(off_x[c / kt]), b = static_cast<double>(off_t[c % kt]);
            dist[c] = a * a + b * b;
        }
        order[c] = c;
    }
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return dist[u] < dist[v]; });
    for (int r = 0; r < K; ++r) {
        const int c = order[r];
        rank_cell[r] = c;
        cell_rank[c] = dist[c] < 100.0 ? r : K;
    }
    std::vector<char> blob;
    const size_t o_off_x = put(blob, off_x), o_off_t = put(blob, off_t), o_inf_x = put(blob, inf_x),
                 o_inf_t = put(blob, inf_t), o_crank = put(blob, cell_rank), o_rcell = put(blob, rank_cell);
    CVG_TRY(upload_tables(ctx, blob));
    char *tb = static_cast<char *>(ctx->d_tables.ptr);

    const size_t np = static_cast<size_t>(std::max(pairs, 1));
    CVG_TRY(ensure(ctx, ctx->d_st_ids, np * cap * sizeof(int)));
    CVG_TRY(ensure(ctx, ctx->d_st_meta, np * sizeof(unsigned)));
    CVG_TRY(ensure(ctx, ctx->d_st_k, np * 2 * sizeof(double)));
    CVG_TRY(ensure(ctx, ctx->d_st_rank, np * sizeof(int)));
    CVG_TRY(ensure(ctx, ctx->d_acc, static_cast<size_t>(n) * sizeof(unsigned)));
    CVG_TRY(ensure(ctx, ctx->d_rowoff, static_cast<size_t>(n + 1) * sizeof(unsigned)));
    unsigned *scal = as<unsigned>(ctx->d_scal);
    CVG_HIP(ctx, hipMemsetAsync(scal + S_VALID, 0, (kScalarBufWords - S_VALID) * sizeof(unsigned), st));

    // one wave per anchor, as many waves as the per-wave scratch allows
    const size_t per_wave = static_cast<size_t>(max_deg) * K * 2 * sizeof(T) +
                            static_cast<size_t>(max_deg) * sizeof(int) +
                            static_cast<size_t>(max_deg) * (cap + 1) * sizeof(int) + static_cast<size_t>(cap) * sizeof(int) +
                            256;
    long long waves = std::min<long long>(n, static_cast<long long>(ctx->num_cu) * 16);
    waves = std::max(1LL, std::min<long long>(waves, static_cast<long long>((size_t(512) << 20) / per_wave)));
    const int blocks = static_cast<int>((waves + kGenWaves - 1) / kGenWaves);
    const size_t nwaves = static_cast<size_t>(blocks) * kGenWaves;
    const size_t sb_elems = nwaves * max_deg * K * 2;
    const size_t sf_ints = nwaves * max_deg;
    const size_t sa_ints = nwaves * max_deg * (cap + 1);
    const size_t sc_ints = nwaves * cap;
    CVG_TRY(ensure(ctx, ctx->d_gbundle, sb_elems * sizeof(T) + (sf_ints + sa_ints + sc_ints) * sizeof(int) + 512));
    GenArgs<T> A{};
    A.n = n;
    A.x = as<double>(ctx->d_x);
    A.y = as<double>(ctx->d_y);
    A.th = as<double>(ctx->d_th);
    A.off = as<unsigned>(ctx->d_off);
    A.ids = as<int>(ctx->d_ids);
    A.kt = kt;
    A.K = K;
    A.G = G;
    A.cap = cap;
    A.max_deg = max_deg;
    A.dx = dx;
    A.dt = dt;
    A.max_k = max_k;
    A.dx2 = dx * dx;                      // cpu_curvelet.hpp:654
    A.close_thr = static_cast<T>(4) * dx * dx;  // :682, (4*dx)*dx
    A.off_x = reinterpret_cast<const T *>(tb + o_off_x);
    A.off_t = reinterpret_cast<const T *>(tb + o_off_t);
    A.cell_rank = reinterpret_cast<const int *>(tb + o_crank);
    A.rank_cell = reinterpret_cast<const int *>(tb + o_rcell);
    A.sc_bundle = as<T>(ctx->d_gbundle);
    A.sc_flag = reinterpret_cast<int *>(A.sc_bundle + sb_elems);
    A.sc_acc = A.sc_flag + sf_ints;
    A.sc_chain = A.sc_acc + sa_ints;
    A.st_ids = as<int>(ctx->d_st_ids);
    A.st_meta = as<unsigned>(ctx->d_st_meta);
    A.st_k = as<double>(ctx->d_st_k);
    A.st_rank = as<int>(ctx->d_st_rank);
    A.acc_count = as<unsigned>(ctx->d_acc);
    A.valid_pairs = scal + S_VALID;
    A.intersections = reinterpret_cast<unsigned long long *>(scal + S_COUNT);
    const int ring = static_cast<int>(ctx->fused_runs % cvg_context::kFusedRing);
    CVG_TRY(record_timing(ctx, ctx->ev[3], st));
    CVG_TRY(record_timing(ctx, ctx->fused_ev[ring][0], st));
    hipLaunchKernelGGL(k_curvelets_generic<T>, dim3(blocks), dim3(64 * kGenWaves), 0, st, A);
    CVG_HIP(ctx, hipGetLastError());
    CVG_TRY(record_timing(ctx, ctx->ev[4], st));
    CVG_TRY(record_timing(ctx, ctx->fused_ev[ring][1], st));
    if (ctx->timed_run) ++ctx->fused_runs;

    RecordArgs R{};
    R.G = G;
    R.kt = kt;
    R.nrad = p.nrad;
    if (kFloat) {
        R.f32 = true;
        R.inf_x = reinterpret_cast<const float *>(tb + o_inf_x);
        R.inf_t = reinterpret_cast<const float *>(tb + o_inf_t);
        R.nrad_f = static_cast<float>(p.nrad);
    } else {
        R.off_x = reinterpret_cast<const double *>(tb + o_off_x);
        R.off_t = reinterpret_cast<const double *>(tb + o_off_t);
    }
    return record_rows(ctx, p, R, cap);
}

int build_curvelets(cvg_context *ctx, const cvg_params &p)
{
    if (ctx->f32) return build_curvelets_generic<float>(ctx, p);
    if (!ctx->graph_ready) return fail(ctx, CVG_ERR_STATE, "build the neighbour graph first");
    ctx->curvelets_ready = false;
    Derived d;
    CVG_TRY(derive(ctx, p, d));
    // the fused kernel holds a chain in the lanes of a wave and up to four
    // 64-neighbour chunks per anchor; other maps take the general kernel
    if (d.G > 64 || ctx->max_deg > 256 || ctx->force_generic) {
        ctx->edge_trig_queued = false;  // the general kernel computes its own trig
        ctx->pair_setup_queued = false;
        return build_curvelets_generic<double>(ctx, p);
    }
    hipStream_t st = ctx->stream;
    const int n = ctx->n;
    const int pairs = ctx->pairs;

    // bucket capacities (neighbour slots per workgroup)
    std::vector<int> caps = {4, 6, 8, 11, 16, 24, 32, 48, 64};
    const int maxcap = static_cast<int>(std::max<unsigned>(ctx->max_deg, 1u));
    while (!caps.empty() && caps.back() >= maxcap) caps.pop_back();
    caps.push_back(maxcap);
    const int nb = static_cast<int>(caps.size());
    if (nb > kMaxBuckets) return fail(ctx, CVG_ERR_STATE, "internal: too many buckets");

    // parameter tables
    std::vector<char> blob;
    const size_t o_rank_ox = put(blob, d.rank_ox);
    const size_t o_rank_j = put(blob, d.rank_j);
    const size_t o_rank_cell = put(blob, d.rank_cell);
    const size_t o_rank_elig = put(blob, d.rank_elig);
    const size_t o_off_x = put(blob, d.off_x);
    const size_t o_off_t = put(blob, d.off_t);
    const size_t o_caps = put(blob, caps);
    CVG_TRY(upload_tables(ctx, blob));
    char *tb = static_cast<char *>(ctx->d_tables.ptr);
    Tables T;
    T.rank_ox = reinterpret_cast<double *>(tb + o_rank_ox);
    T.rank_j = reinterpret_cast<int *>(tb + o_rank_j);
    T.rank_cell = reinterpret_cast<int *>(tb + o_rank_cell);
    T.rank_elig = reinterpret_cast<int *>(tb + o_rank_elig);
    T.off_x = reinterpret_cast<double *>(tb + o_off_x);
    T.off_t = reinterpret_cast<double *>(tb + o_off_t);
    T.caps = reinterpret_cast<int *>(tb + o_caps);

    // per-edge trig, per-pair classes and sines, buckets
    if (!ctx->edge_trig_queued) CVG_TRY(queue_edge_trig(ctx, p));
    ctx->edge_trig_queued = false;
    const bool pair_setup_queued = ctx->pair_setup_queued;
    ctx->pair_setup_queued = false;
    CVG_TRY(ensure(ctx, ctx->d_cls, cls_bytes(pairs)));
    CVG_TRY(ensure(ctx, ctx->d_prec, static_cast<size_t>(std::max(pairs, 1)) * 8 * sizeof(double)));
    CVG_TRY(ensure(ctx, ctx->d_psin, static_cast<size_t>(std::max(pairs, 1)) * d.kt * 2 * sizeof(double)));
    CVG_TRY(ensure(ctx, ctx->d_lists, static_cast<size_t>(nb) * n * sizeof(int4)));
    unsigned *scal = as<unsigned>(ctx->d_scal);
    if (static_cast<unsigned long long>(pairs) * d.kt > 0xffffffffull)
        return fail(ctx, CVG_ERR_UNSUPPORTED, "more than 2^32 (pair, angle) items in one run");
    if (pairs > 0 && !pair_setup_queued) {
        const PrepArgs P{as<double>(ctx->d_x), as<double>(ctx->d_y), as<double>(ctx->d_trig), d.close_thr,
                         as<unsigned char>(ctx->d_cls), as<double>(ctx->d_prec)};
        // Kt = 7 (the default grid): one kernel forms the pairs' records and
        // sines; other grids run the record first
        const bool fused_prep = d.kt == 7;
        if (!fused_prep)
            hipLaunchKernelGGL(k_pair_prep, dim3(blocks_for(pairs)), dim3(kBlock), 0, st, pairs, n,
                               as<int>(ctx->d_pair_anchor), as<int>(ctx->d_ids), P);
        const size_t trig_lds = d.kt <= static_cast<unsigned>(kPairTrigMaxKt) ? kBlock * d.kt * sizeof(double2) : 0;
        const auto trig_fn = fused_prep ? k_pair_trig<7, true> : k_pair_trig<0, false>;
        hipLaunchKernelGGL(trig_fn, dim3(blocks_for(pairs)), dim3(kBlock), trig_lds, st,
                           pairs, n, static_cast<int>(d.kt), as<int>(ctx->d_pair_anchor), as<int>(ctx->d_ids),
                           static_cast<const unsigned char *>(as<unsigned char>(ctx->d_cls)), as<double>(ctx->d_th),
                           p.st, d.dt, as<double>(ctx->d_tdd), as<double>(ctx->d_psin), P,
                           static_cast<const unsigned *>(nullptr));
    }
    hipLaunchKernelGGL(k_bucket, dim3(blocks_for(n, kBucketBlock)), dim3(kBucketBlock), 0, st, n, as<int>(ctx->d_sorted),
                       as<unsigned char>(ctx->d_cls), as<unsigned>(ctx->d_off), nb, T.caps, as<int4>(ctx->d_lists),
                       scal);
    CVG_HIP(ctx, hipGetLastError());

    // staging
    const int st_stride = static_cast<int>(std::max(d.G, 1u));
    const int chain_stride = static_cast<int>(std::max(d.G, 1u));
    const size_t np = static_cast<size_t>(std::max(pairs, 1));
    CVG_TRY(ensure(ctx, ctx->d_st_ids, np * st_stride * sizeof(int)));
    CVG_TRY(ensure(ctx, ctx->d_st_meta, np * sizeof(unsigned)));
    CVG_TRY(ensure(ctx, ctx->d_st_k, np * 2 * sizeof(double)));
    CVG_TRY(ensure(ctx, ctx->d_st_rank, np * sizeof(int)));
    CVG_TRY(ensure(ctx, ctx->d_acc, static_cast<size_t>(n) * sizeof(unsigned)));
    CVG_TRY(ensure(ctx, ctx->d_rowoff, static_cast<size_t>(n + 1) * sizeof(unsigned)));

    FusedArgs A{};
    A.n = n;
    A.prec = as<double>(ctx->d_prec);
    A.trig = as<double>(ctx->d_trig);
    A.psin = as<double>(ctx->d_psin);
    A.ids = as<int>(ctx->d_ids);
    A.pflag = as<unsigned char>(ctx->d_cls);
    A.chain_stride = chain_stride;
    A.st_stride = st_stride;
    A.dx = d.dx;
    A.dx2 = d.dx2;
    A.max_k = d.max_k;
    A.kt = static_cast<int>(d.kt);
    A.K = static_cast<int>(d.K);
    A.G = static_cast<int>(d.G);
    A.rank_ox = T.rank_ox;
    A.rank_j = T.rank_j;
    A.rank_cell = T.rank_cell;
    A.rank_elig = T.rank_elig;
    A.rank_of_cell0 = d.rank_of_cell0;
    A.st_ids = as<int>(ctx->d_st_ids);
    A.st_meta = as<unsigned>(ctx->d_st_meta);
    A.st_k = as<double>(ctx->d_st_k);
    A.st_rank = as<int>(ctx->d_st_rank);
    A.acc_count = as<unsigned>(ctx->d_acc);
    A.valid_pairs = scal + S_VALID;
    A.intersections = reinterpret_cast<unsigned long long *>(scal + S_COUNT);

    // per bucket: 64-neighbour chunks of the lane-indexed lists (degree <= 2 x
    // the bucket's key and <= the map's largest); heavy buckets share one anchor
    // among the waves of a workgroup, light ones pack one anchor per wave;
    // bundles go to global memory only when one anchor's slice alone exceeds
    // LDS.  Each launch is persistent: as many anchor streams as stay resident
    // (the occupancy query accounts for registers and LDS), each looping over
    // its share of the bucket's list.
    struct Launch {
        int nch, wpb, apb, dcap, grid;
        bool gb, coop;
        size_t slice;
        unsigned shmem;
        FusedKernel fn;
    };
    std::vector<Launch> plan(nb);
    std::vector<size_t> gb_off(nb, 0);
    size_t gb_total = 0;
    for (int b = 0; b < nb; ++b) {
        Launch &l = plan[b];
        l.dcap = std::min(2 * caps[b], static_cast<int>(std::max(ctx->max_deg, 1u)));
        const int need = (l.dcap + 63) / 64;
        if (need > 4) return build_curvelets_generic<double>(ctx, p);
        l.nch = need <= 1 ? 1 : (need <= 2 ? 2 : 4);
        l.gb = false;
        l.coop = caps[b] >= ctx->coop_min_cap;
        l.slice = anchor_carve(caps[b], l.dcap, A.K, A.kt, chain_stride, true, l.nch).bytes;
        if (l.coop) {
            l.wpb = ctx->coop_waves;
        } else {
            l.wpb = 4;
            while (l.wpb > 1 && l.slice * l.wpb > ctx->lds_limit) l.wpb >>= 1;
        }
        if (l.slice > ctx->lds_limit) {
            l.gb = true;
            l.coop = false;
            l.wpb = 1;
            l.slice = anchor_carve(caps[b], l.dcap, A.K, A.kt, chain_stride, false, l.nch).bytes;
            if (l.slice > ctx->lds_limit) return build_curvelets_generic<double>(ctx, p);
        }
        l.apb = l.coop ? 1 : l.wpb;
        l.shmem = static_cast<unsigned>(l.slice * l.apb);
        l.fn = fused_kernel(d.cpl, l.nch, l.gb, l.coop);
        if (!l.fn) return fail(ctx, CVG_ERR_STATE, "internal: no fused kernel for this grid");
        int occ = 0;
        CVG_TRY(prepare_launch(ctx, reinterpret_cast<const void *>(l.fn), 64 * l.wpb, l.shmem, &occ));
        // the bucket sizes stay on the device: every launch gets the resident
        // grid (streams beyond the bucket's length exit at once), at most one
        // stream per anchor of the map
        long long resident = static_cast<long long>(std::max(occ, 1)) * ctx->num_cu;
        // Launches run four at a time: a light bucket's persistent grid sized
        // for the whole GPU has workgroups that are dispatched late and still
        // own a full share of the list, which stretches the stage's tail.  A
        // grid in proportion to the run's edges fits its lists better; large
        // maps and batches, whose launches run longer, keep the full one.
        if (caps[b] < kHeavyCap) {
            const int pct = ctx->grid_pct > 0 ? ctx->grid_pct : auto_grid_pct(n);
            resident = std::max(1LL, resident * pct / 100);
        }
        const long long max_blocks = (static_cast<long long>(n) + l.apb - 1) / l.apb;
        l.grid = static_cast<int>(std::min(resident, max_blocks));
        if (l.gb) {
            gb_off[b] = gb_total;
            gb_total += static_cast<size_t>(l.grid) * l.apb * caps[b] * A.K * 2;
        }
    }
    if (gb_total > 0) CVG_TRY(ensure(ctx, ctx->d_gbundle, gb_total * sizeof(double)));

    // bucket launches: the short-list bucket 0 (most anchors) alone on the main
    // stream, the long-list buckets round-robin over the auxiliary streams,
    // longest lists first, so their long-latency work starts at once and
    // overlaps the bulk
    const int ring = static_cast<int>(ctx->fused_runs % cvg_context::kFusedRing);
    CVG_TRY(record_timing(ctx, ctx->ev[3], st));
    CVG_TRY(record_timing(ctx, ctx->fused_ev[ring][0], st));
    const hipEvent_t fork = ctx->timed_run ? ctx->ev[3] : ctx->ev_fork;
    if (!ctx->timed_run) CVG_HIP(ctx, hipEventRecord(fork, st));
    for (int i = 0; i < 3; ++i) CVG_HIP(ctx, hipStreamWaitEvent(ctx->aux[i], fork, 0));
    bool aux_used[3] = {false, false, false};
    int rr = 0;
    GateArgs gate{};
    for (int b = nb - 1; b >= 0; --b) {
        const int slot = b == 0 ? 0 : 1 + (rr++ % 3);
        hipStream_t ls = slot == 0 ? st : ctx->aux[slot - 1];
        if (slot > 0) aux_used[slot - 1] = true;
        const Launch &l = plan[b];
        A.list = as<int4>(ctx->d_lists) + static_cast<size_t>(b) * n;
        A.list_len = scal + kBucketBase + b * kBucketStride;
        A.nb_cap = caps[b];
        const bool heavy = b > 0 && caps[b] >= kHeavyCap;  // on an auxiliary stream
        A.started = heavy ? scal + kStartedWord : nullptr;
        if (heavy) {
            gate.bucket[gate.nheavy] = b;
            gate.streams[gate.nheavy] = l.grid * l.apb;
            ++gate.nheavy;
        }
        if (b == 0 && gate.nheavy > 0) {
            hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, st, gate, static_cast<const unsigned *>(scal),
                               ctx->gate_naps);
            CVG_HIP(ctx, hipGetLastError());
        }
        A.deg_cap = l.dcap;
        A.g_bundle = l.gb ? as<double>(ctx->d_gbundle) + gb_off[b] : nullptr;
        hipLaunchKernelGGL(l.fn, dim3(l.grid), dim3(64 * l.wpb), l.shmem, ls, A);
        CVG_HIP(ctx, hipGetLastError());
    }
    for (int i = 0; i < 3; ++i) {
        if (!aux_used[i]) continue;
        CVG_HIP(ctx, hipEventRecord(ctx->ev_join[i], ctx->aux[i]));
        CVG_HIP(ctx, hipStreamWaitEvent(st, ctx->ev_join[i], 0));
    }
    CVG_TRY(record_timing(ctx, ctx->ev[4], st));
    CVG_TRY(record_timing(ctx, ctx->fused_ev[ring][1], st));
    if (ctx->timed_run) ++ctx->fused_runs;
    // every anchor is in exactly one bucket, so every acc_count entry is written
    RecordArgs R{};
    R.G = static_cast<int>(d.G);
    R.kt = static_cast<int>(d.kt);
    R.nrad = d.nrad;
    R.off_x = T.off_x;
    R.off_t = T.off_t;
    return record_rows(ctx, p, R, st_stride);
}

bool same_params(const cvg_params &a, const cvg_params &b)
{
    return a.nrad == b.nrad && a.dx == b.dx && a.dt_deg == b.dt_deg && a.max_k == b.max_k &&
           a.group_max_sz == b.group_max_sz && a.sx == b.sx && a.st == b.st;
}

int check_ctx(cvg_context *ctx)
{
    if (!ctx) return CVG_ERR_INVALID_ARGUMENT;
    CVG_HIP(ctx, hipSetDevice(ctx->device));
    return CVG_OK;
}

void fill_stats(cvg_context *ctx, cvg_run_stats *stats);

// timed: record the timing events although stats is null (the caller fills
// stats itself later)
int run_common(cvg_context *ctx, int n, int sz, const double *d_edges, const cvg_params *p, cvg_run_stats *stats,
               int num_maps = 1, const long long *eoff = nullptr, bool f32 = false, bool timed = false)
{
    ctx->f32 = f32;
    ctx->timed_run = timed || stats || ctx->event_timing;
    CVG_TRY(record_timing(ctx, ctx->ev[0], ctx->stream));
    ctx->use_cellrank = false;
    ctx->edge_trig_queued = false;
    ctx->early_params = (!f32 && p->group_max_sz <= 64 && !ctx->force_generic) ? p : nullptr;
    const int rc = build_graph(ctx, n, sz, d_edges, p->nrad, 3u, p->max_candidates, num_maps, eoff);  // main.cpp:48: 3
    ctx->early_params = nullptr;
    if (rc != CVG_OK) return rc;
    CVG_TRY(record_timing(ctx, ctx->ev[1], ctx->stream));
    CVG_TRY(build_curvelets(ctx, *p));
    CVG_TRY(record_timing(ctx, ctx->ev[2], ctx->stream));
    // without stats the call returns with the work queued: the counts come
    // back with the stream and are read at the first query of a result
    if (!stats) return CVG_OK;
    CVG_HIP(ctx, hipEventSynchronize(ctx->ev[2]));
    CVG_TRY(finish_curvelets(ctx));
    fill_stats(ctx, stats);
    return CVG_OK;
}

// The run's counts and stage times (the stream has passed ev[2]).
void fill_stats(cvg_context *ctx, cvg_run_stats *stats)
{
    {
        float t01 = 0, t12 = 0, t02 = 0;
        (void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
        (void)hipEventElapsedTime(&t02, ctx->ev[0], ctx->ev[2]);
        stats->valid_pairs = ctx->valid_pairs;
        stats->num_curvelets = ctx->num_curvelets;
        stats->chain_width = ctx->width;
        stats->total_neighbor_pairs = ctx->pairs;
        stats->max_num_of_neighbors = ctx->max_deg;
        stats->ms_graph = t01;
        stats->ms_curvelets = t12;
        stats->ms_total = t02;
        float tf = 0;
        (void)hipEventElapsedTime(&tf, ctx->ev[3], ctx->ev[4]);
        stats->ms_fused_kernel = tf;
        stats->intersections = ctx->intersections;
        stats->order_ambiguous_anchors = ctx->ambiguous;
        stats->lds_bytes_per_workgroup = static_cast<unsigned>(ctx->lds_limit);
    }
}

int upload(cvg_context *ctx, int n, int sz, const double *host, const double **dptr)
{
    if (!host || n <= 0 || sz < 1) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "upload: bad edge array");
    const size_t bytes = static_cast<size_t>(n) * sz * sizeof(double);
    CVG_TRY(ensure(ctx, ctx->d_edges_up, bytes));
    CVG_HIP(ctx, hipMemcpyAsync(ctx->d_edges_up.ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    *dptr = as<double>(ctx->d_edges_up);
    return CVG_OK;
}

// This device's copy of `scratch`, resolved once per device with that device
// current (each GPU holds its own copy at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(cvg_context *ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return CVG_OK;  // no-alloc runs unavailable there
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *ptr = nullptr;
        CVG_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(ctx, CVG_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[ctx->device] = static_cast<char *>(ptr);
    }
    ctx->scratch_base = g_scratch_ptr[ctx->device];
    return CVG_OK;
}

// Device-to-host copies into caller memory, complete on return.  Page-locked
// or small destinations take one DMA each.  A pageable destination would make
// the runtime stage the copy itself at about a third of the link's rate, so
// the pageable segments go, as one stream of kStageBytes chunks, through two
// page-locked blocks: while chunk k is copied out on host threads, chunk
// k + 1's DMA runs.
constexpr size_t kStageBytes = static_cast<size_t>(4) << 20;
constexpr size_t kStageMinBytes = static_cast<size_t>(1) << 20;

struct Segment {
    void *dst;
    const void *src;
    size_t bytes;
};

// Memory the runtime knows (page-locked host, device or managed): one DMA.
bool dma_direct(const void *p)
{
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();  // an unregistered pointer reports an error: it must not stay pending
    return e == hipSuccess && a.type != hipMemoryTypeUnregistered;
}

int download(cvg_context *ctx, std::initializer_list<Segment> segs)
{
    hipStream_t st = ctx->stream;
    struct Chunk {
        char *dst;
        const char *src;
        size_t len;
    };
    std::vector<Chunk> chunks;
    for (const Segment &g : segs) {
        if (g.bytes == 0 || !g.dst) continue;
        if (g.bytes < kStageMinBytes || dma_direct(g.dst)) {
            CVG_HIP(ctx, hipMemcpyAsync(g.dst, g.src, g.bytes, hipMemcpyDefault, st));
            continue;
        }
        for (size_t o = 0; o < g.bytes; o += kStageBytes)
            chunks.push_back({static_cast<char *>(g.dst) + o, static_cast<const char *>(g.src) + o,
                              std::min(kStageBytes, g.bytes - o)});
    }
    if (chunks.empty()) {
        CVG_HIP(ctx, hipStreamSynchronize(st));
        return CVG_OK;
    }
    for (int b = 0; b < 2; ++b) {
        if (!ctx->h_stage[b]) {
            if (hipHostMalloc(&ctx->h_stage[b], kStageBytes, hipHostMallocDefault) != hipSuccess) {
                ctx->h_stage[b] = nullptr;
                return fail(ctx, CVG_ERR_OUT_OF_MEMORY, "hipHostMalloc (download staging) failed");
            }
        }
        if (!ctx->ev_stage[b]) CVG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_stage[b], hipEventDisableTiming));
    }
    const size_t nchunk = chunks.size();
    auto issue = [&](size_t k) -> int {
        const int b = static_cast<int>(k & 1);
        CVG_HIP(ctx, hipMemcpyAsync(ctx->h_stage[b], chunks[k].src, chunks[k].len, hipMemcpyDeviceToHost, st));
        CVG_HIP(ctx, hipEventRecord(ctx->ev_stage[b], st));
        return CVG_OK;
    };
    // host copy-out: nw helper threads and this one split each chunk; the
    // chunk a helper may copy is published in `ready`, and `done` counts
    // finished slices (every helper does one slice per chunk)
    const unsigned hw = std::thread::hardware_concurrency();
    const int nw = static_cast<int>(std::min<unsigned>(hw > 1 ? hw - 1 : 0, 11));
    const int parts = nw + 1;
    std::atomic<long long> ready{-1};
    std::atomic<long long> done{0};
    std::atomic<bool> quit{false};
    auto slice = [&](size_t k, int part) {
        const size_t len = chunks[k].len;
        const size_t per = ((len + parts - 1) / parts + 63) & ~static_cast<size_t>(63);
        const size_t o = std::min(len, per * part), e = std::min(len, o + per);
        if (e > o) std::memcpy(chunks[k].dst + o, static_cast<const char *>(ctx->h_stage[k & 1]) + o, e - o);
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < nw; ++w) {
        try {
            pool.emplace_back([&, w] {
                for (size_t k = 0; k < nchunk; ++k) {
                    while (ready.load(std::memory_order_acquire) < static_cast<long long>(k)) {
                        if (quit.load(std::memory_order_relaxed)) return;
                        std::this_thread::yield();
                    }
                    slice(k, w + 1);
                    done.fetch_add(1, std::memory_order_release);
                }
            });
        } catch (...) {
            break;  // fewer helpers: the slices of missing ones are copied below
        }
    }
    const int started = static_cast<int>(pool.size());
    int rc = issue(0);
    for (size_t k = 0; rc == CVG_OK && k < nchunk; ++k) {
        // block (k + 1) & 1 last held chunk k - 1: every slice of it is copied
        while (done.load(std::memory_order_acquire) < static_cast<long long>(k) * started) std::this_thread::yield();
        if (k + 1 < nchunk) rc = issue(k + 1);
        if (rc != CVG_OK) break;
        if (hipEventSynchronize(ctx->ev_stage[k & 1]) != hipSuccess) {
            rc = fail(ctx, CVG_ERR_DEVICE, "download: staging copy failed");
            break;
        }
        ready.store(static_cast<long long>(k), std::memory_order_release);
        slice(k, 0);
        for (int w = started + 1; w < parts; ++w) slice(k, w);
    }
    if (rc == CVG_OK)
        while (done.load(std::memory_order_acquire) < static_cast<long long>(nchunk) * started)
            std::this_thread::yield();
    quit.store(true, std::memory_order_relaxed);
    for (auto &t : pool) t.join();
    // the direct copies, and after a failure any DMA still aimed at a block
    if (hipStreamSynchronize(st) != hipSuccess && rc == CVG_OK)
        rc = fail(ctx, CVG_ERR_DEVICE, "download: device-to-host copy failed");
    return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

void cvg_params_default(cvg_params *p)
{
    if (!p) return;
    p->nrad = 3.5;
    p->gap = 1.5;
    p->dx = 0.4;
    p->dt_deg = 15.0;
    p->max_k = 0.3;
    p->curvelet_style = 2;
    p->group_max_sz = 4;
    p->out_type = 0;
    p->sx = 0.1;
    p->st = 0.08;
    p->edge_data_sz = 4;
    p->max_candidates = 128;
}

const char *cvg_status_string(int status)
{
    switch (status) {
    case CVG_OK: return "ok";
    case CVG_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CVG_ERR_BAD_INPUT: return "bad input";
    case CVG_ERR_IO: return "i/o error";
    case CVG_ERR_OUT_OF_MEMORY: return "out of memory";
    case CVG_ERR_DEVICE: return "device error";
    case CVG_ERR_UNSUPPORTED: return "unsupported";
    case CVG_ERR_STATE: return "bad call order";
    default: return "unknown status";
    }
}

int cvg_create(int device, cvg_context **out_ctx)
{
    if (!out_ctx) return CVG_ERR_INVALID_ARGUMENT;
    *out_ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CVG_ERR_DEVICE;
    if (device < 0 || device >= count) return CVG_ERR_INVALID_ARGUMENT;
    cvg_context *ctx = new (std::nothrow) cvg_context();
    if (!ctx) return CVG_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    int rc = CVG_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = CVG_ERR_DEVICE; break; }
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { rc = CVG_ERR_DEVICE; break; }
        for (int i = 0; i < 3 && rc == CVG_OK; ++i) {
            if (hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming) != hipSuccess)
                rc = CVG_ERR_DEVICE;
        }
        if (rc != CVG_OK) break;
        for (auto &e : ctx->ev)
            if (hipEventCreate(&e) != hipSuccess) rc = CVG_ERR_DEVICE;
        for (hipEvent_t *e : {&ctx->ev_read, &ctx->ev_fork, &ctx->ev_end})
            if (rc == CVG_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rc = CVG_ERR_DEVICE;
        for (auto &pr : ctx->fused_ev)
            for (auto &e : pr)
                if (rc == CVG_OK && hipEventCreate(&e) != hipSuccess) rc = CVG_ERR_DEVICE;
        if (rc != CVG_OK) break;
        if (hipHostMalloc(reinterpret_cast<void **>(&ctx->mailbox), 64 * sizeof(unsigned), hipHostMallocDefault) !=
            hipSuccess) {
            rc = CVG_ERR_OUT_OF_MEMORY;
            break;
        }
        rc = scratch_address(ctx);
        if (rc != CVG_OK) break;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            size_t lim = prop.sharedMemPerBlock;
            if (prop.maxSharedMemoryPerMultiProcessor > 0)
                lim = std::min(lim, static_cast<size_t>(prop.maxSharedMemoryPerMultiProcessor));
            if (lim > 0) ctx->lds_limit = lim;
            if (prop.multiProcessorCount > 0) ctx->num_cu = prop.multiProcessorCount;
        }
        if (const char *v = std::getenv("CVG_COOP_MIN_CAP")) ctx->coop_min_cap = std::max(1, std::atoi(v));
        if (const char *v = std::getenv("CVG_COOP_WAVES")) {
            const int w = std::atoi(v);
            ctx->coop_waves = (w == 1 || w == 2 || w == 4) ? w : 4;
        }
        if (const char *v = std::getenv("CVG_FILL_LANES")) {
            const int f = std::atoi(v);
            ctx->fill_lanes = (f == 8 || f == 16 || f == 32 || f == 64) ? f : 16;
        }
        if (const char *v = std::getenv("CVG_GRID_PCT")) {
            const int g = std::atoi(v);
            ctx->grid_pct = (g >= 5 && g <= 100) ? g : 0;
        }
        if (const char *v = std::getenv("CVG_SPARSE_CELLS")) ctx->force_sparse = std::atoi(v) != 0;
        if (const char *v = std::getenv("CVG_GENERIC_KERNEL")) ctx->force_generic = std::atoi(v) != 0;
        if (const char *v = std::getenv("CVG_SPECULATE")) ctx->spec_off = std::atoi(v) == 0;
        if (const char *v = std::getenv("CVG_GATE_NAPS")) ctx->gate_naps = std::min(std::max(0, std::atoi(v)), 400);
        if (const char *v = std::getenv("CVG_LOOKBACK_SPINS"))
            ctx->lookback_spins = static_cast<unsigned>(std::min<long long>(std::max(0LL, std::atoll(v)), kLookbackSpins));
    } while (false);
    if (rc != CVG_OK) {
        cvg_destroy(ctx);
        return rc;
    }
    *out_ctx = ctx;
    return CVG_OK;
}

void cvg_destroy(cvg_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->for_each_pool([](DevBuf &b) { release(b); });
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ctx->ev_read, ctx->ev_fork, ctx->ev_end})
        if (e) (void)hipEventDestroy(e);
    for (auto &pr : ctx->fused_ev)
        for (auto &e : pr)
            if (e) (void)hipEventDestroy(e);
    if (ctx->mailbox) (void)hipHostFree(ctx->mailbox);
    if (ctx->h_maprow) (void)hipHostFree(ctx->h_maprow);
    for (int b = 0; b < 2; ++b) {
        if (ctx->h_stage[b]) (void)hipHostFree(ctx->h_stage[b]);
        if (ctx->ev_stage[b]) (void)hipEventDestroy(ctx->ev_stage[b]);
    }
    for (int i = 0; i < 3; ++i) {
        if (ctx->aux[i]) (void)hipStreamSynchronize(ctx->aux[i]);
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
        if (ctx->aux[i]) (void)hipStreamDestroy(ctx->aux[i]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *cvg_last_error(const cvg_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cvg_build_neighbor_csr_graph(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges,
                                 double rad, unsigned neighbor_radius, unsigned max_candidates)
{
    CVG_TRY(check_ctx(ctx));
    if (!edges || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "build_neighbor_csr_graph: need num_edges > 0, sz_edge_data >= 4");
    const double *d = nullptr;
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, edges, &d));
    ctx->use_cellrank = false;
    ctx->f32 = false;
    CVG_TRY(build_graph(ctx, num_edges, sz_edge_data, d, rad, neighbor_radius, max_candidates));
    CVG_TRY(read_scalars(ctx));
    if (ctx->mailbox[S_AMBIG] > 0) {
        CVG_TRY(rebuild_graph_exact(ctx));
        CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CVG_OK;
}

int cvg_build_neighbor_csr_graph_f32(cvg_context *ctx, int num_edges, int sz_edge_data, const float *edges,
                                     double rad, unsigned neighbor_radius, unsigned max_candidates)
{
    CVG_TRY(check_ctx(ctx));
    if (!edges || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "build_neighbor_csr_graph: need num_edges > 0, sz_edge_data >= 4");
    std::vector<double> wide;
    try {
        wide.assign(edges, edges + static_cast<size_t>(num_edges) * sz_edge_data);
    } catch (...) {
        return fail(ctx, CVG_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    const double *d = nullptr;
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, wide.data(), &d));
    ctx->use_cellrank = false;
    ctx->f32 = true;
    CVG_TRY(build_graph(ctx, num_edges, sz_edge_data, d, rad, neighbor_radius, max_candidates));
    CVG_TRY(read_scalars(ctx));
    if (ctx->mailbox[S_AMBIG] > 0) {
        CVG_TRY(rebuild_graph_exact(ctx));
        CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CVG_OK;
}

int cvg_get_graph_info(cvg_context *ctx, cvg_graph_info *info)
{
    CVG_TRY(check_ctx(ctx));
    if (!info) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "null info");
    if (!ctx->graph_ready) return fail(ctx, CVG_ERR_STATE, "no neighbour graph");
    CVG_TRY(finish_curvelets(ctx));
    info->num_edges = ctx->n;
    info->total_neighbor_pairs = ctx->pairs;
    info->max_num_of_neighbors = ctx->max_deg;
    info->truncated_anchors = ctx->truncated;
    // the ambiguity count is produced by the fill kernel; read it now
    CVG_TRY(read_scalars(ctx));
    ctx->ambiguous = ctx->mailbox[S_AMBIG];
    info->order_ambiguous_anchors = ctx->ambiguous;
    return CVG_OK;
}

int cvg_get_csr(cvg_context *ctx, int *offsets, int *ids, double *dist2)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->graph_ready) return fail(ctx, CVG_ERR_STATE, "no neighbour graph");
    CVG_TRY(finish_curvelets(ctx));
    const size_t P = ctx->pairs > 0 ? static_cast<size_t>(ctx->pairs) : 0;
    return download(ctx, {{offsets, ctx->d_off.ptr, static_cast<size_t>(ctx->n + 1) * sizeof(int)},
                          {ids, ctx->d_ids.ptr, P * sizeof(int)},
                          {dist2, ctx->d_d2.ptr, P * sizeof(double)}});
}

int cvg_form_pairwise_bundles(cvg_context *ctx, const cvg_params *params, unsigned *valid_pairs)
{
    CVG_TRY(check_ctx(ctx));
    if (!params) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "null params");
    ctx->timed_run = ctx->event_timing;
    CVG_TRY(build_curvelets(ctx, *params));
    CVG_TRY(finish_curvelets(ctx));
    if (valid_pairs) *valid_pairs = ctx->valid_pairs;
    return CVG_OK;
}

int cvg_build_curvelets_greedy(cvg_context *ctx, const cvg_params *params, unsigned *num_curvelets)
{
    CVG_TRY(check_ctx(ctx));
    if (!params) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "null params");
    if (!(ctx->curvelets_ready && same_params(ctx->cur_params, *params))) {
        ctx->timed_run = ctx->event_timing;
        CVG_TRY(build_curvelets(ctx, *params));
    }
    CVG_TRY(finish_curvelets(ctx));
    if (num_curvelets) *num_curvelets = ctx->num_curvelets;
    return CVG_OK;
}

int cvg_curvelet_shape(cvg_context *ctx, unsigned *num_curvelets, unsigned *chain_width)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->curvelets_ready) return fail(ctx, CVG_ERR_STATE, "no curvelets");
    CVG_TRY(finish_curvelets(ctx));
    if (num_curvelets) *num_curvelets = ctx->num_curvelets;
    if (chain_width) *chain_width = ctx->width;
    return CVG_OK;
}

int cvg_get_curvelets(cvg_context *ctx, unsigned *chains, double *info)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->curvelets_ready) return fail(ctx, CVG_ERR_STATE, "no curvelets");
    CVG_TRY(finish_curvelets(ctx));
    const size_t C = ctx->num_curvelets;
    return download(ctx, {{chains, ctx->d_chains.ptr, C * ctx->width * sizeof(unsigned)},
                          {info, ctx->d_info.ptr, C * kInfoWidth * sizeof(double)}});
}

int cvg_run(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges, const cvg_params *params,
            cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !edges || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run: bad arguments");
    const double *d = nullptr;
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, edges, &d));
    return run_common(ctx, num_edges, sz_edge_data, d, params, stats);
}

int cvg_run_f32(cvg_context *ctx, int num_edges, int sz_edge_data, const float *edges, const cvg_params *params,
                cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !edges || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_f32: bad arguments");
    // the float values held exactly as doubles on the device
    std::vector<double> wide;
    try {
        wide.assign(edges, edges + static_cast<size_t>(num_edges) * sz_edge_data);
    } catch (...) {
        return fail(ctx, CVG_ERR_OUT_OF_MEMORY, "cvg_run_f32: host allocation failed");
    }
    const double *d = nullptr;
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, wide.data(), &d));
    // the upload must have left the host vector before it goes away
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return run_common(ctx, num_edges, sz_edge_data, d, params, stats, 1, nullptr, true);
}

int cvg_run_to_host(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges, const cvg_params *params,
                    unsigned *chains, double *info, unsigned max_rows, unsigned *num_rows, cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !edges || !chains || !info || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_to_host: bad arguments");
    const double *d = nullptr;
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, edges, &d));
    CVG_TRY(run_common(ctx, num_edges, sz_edge_data, d, params, nullptr, 1, nullptr, false, stats != nullptr));
    // the host waits for the counts (the end of the run's kernels) while the
    // early copies below still run
    if (!ctx->timed_run) CVG_HIP(ctx, hipEventRecord(ctx->ev_end, ctx->stream));
    // As many rows as the last call had are copied as soon as the record
    // ends, before the count comes back: a map with no more rows (a
    // repeated or similar map) then needs no further round trip.  Only
    // page-locked outputs (one DMA each); rows past the count may be
    // overwritten.
    size_t early = 0;
    const size_t w = ctx->width;
    if (ctx->to_host_rows > 0 && dma_direct(chains) && dma_direct(info)) {
        // (within what the row buffers hold at this run's width)
        early = std::min({static_cast<size_t>(ctx->to_host_rows), static_cast<size_t>(max_rows),
                          ctx->d_chains.bytes / (w * sizeof(unsigned)), ctx->d_info.bytes / (kInfoWidth * sizeof(double))});
        CVG_HIP(ctx, hipMemcpyAsync(chains, ctx->d_chains.ptr, early * w * sizeof(unsigned), hipMemcpyDeviceToHost,
                                    ctx->stream));
        CVG_HIP(ctx, hipMemcpyAsync(info, ctx->d_info.ptr, early * kInfoWidth * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
    }
    const unsigned long long rebuilds = ctx->rebuilds;
    CVG_HIP(ctx, hipEventSynchronize(end_event(ctx)));
    CVG_TRY(finish_curvelets(ctx));
    if (ctx->rebuilds != rebuilds) early = 0;  // built again after the early copy
    const unsigned rows = ctx->num_curvelets;
    if (num_rows) *num_rows = rows;
    ctx->to_host_rows = rows;
    const size_t C = std::min(rows, max_rows);
    if (C > early)
        CVG_TRY(download(ctx, {{chains + early * w, static_cast<const unsigned *>(ctx->d_chains.ptr) + early * w,
                                (C - early) * w * sizeof(unsigned)},
                               {info + early * kInfoWidth,
                                static_cast<const double *>(ctx->d_info.ptr) + early * kInfoWidth,
                                (C - early) * kInfoWidth * sizeof(double)}}));
    if (stats) fill_stats(ctx, stats);
    if (rows > max_rows)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT,
                    "cvg_run_to_host: the outputs hold " + std::to_string(max_rows) + " rows, the run has " +
                        std::to_string(rows));
    return CVG_OK;
}

int cvg_run_device(cvg_context *ctx, int num_edges, int sz_edge_data, const double *device_edges,
                   const cvg_params *params, cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !device_edges || num_edges <= 0 || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_device: bad arguments");
    return run_common(ctx, num_edges, sz_edge_data, device_edges, params, stats);
}

int cvg_run_device_noalloc(cvg_context *ctx, int num_edges, int sz_edge_data, const double *device_edges,
                           const cvg_params *params, unsigned *device_chains, double *device_info, unsigned max_rows,
                           unsigned *num_curvelets, unsigned *valid_pairs)
{
    CVG_TRY(check_ctx(ctx));
    if (num_curvelets) *num_curvelets = 0;
    if (!params || !device_edges || num_edges <= 0 || sz_edge_data < 4 || !device_chains || !device_info ||
        !num_curvelets)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_device_noalloc: bad arguments");
    if (!ctx->scratch_base) return fail(ctx, CVG_ERR_DEVICE, "cvg_run_device_noalloc: no scratch array on this device");
    // counts of an earlier run still in flight come back before the mailbox is reused
    CVG_TRY(finish_curvelets(ctx));
    // the context's own pools step aside (nothing is freed) and every buffer of
    // this run is carved from the scratch array; the outputs are the caller's
    constexpr int kMaxPools = 64;
    DevBuf saved[kMaxPools];
    int npools = 0;
    ctx->for_each_pool([&](DevBuf &b) {
        if (npools < kMaxPools) {
            saved[npools++] = b;
            b = DevBuf{};
        }
    });
    ctx->arena = true;
    ctx->arena_used = 0;
    ctx->caller_chains = device_chains;
    ctx->caller_info = device_info;
    ctx->caller_rows = max_rows;
    int rc = run_common(ctx, num_edges, sz_edge_data, device_edges, params, nullptr);
    if (rc == CVG_OK) rc = finish_curvelets(ctx);
    else (void)hipStreamSynchronize(ctx->stream);
    ctx->arena = false;
    ctx->caller_chains = nullptr;
    ctx->caller_info = nullptr;
    ctx->caller_rows = 0;
    npools = 0;
    ctx->for_each_pool([&](DevBuf &b) {
        if (npools < kMaxPools) b = saved[npools++];
    });
    // the results live in the caller's buffers, not in the context
    ctx->graph_ready = false;
    ctx->curvelets_ready = false;
    if (rc != CVG_OK) return rc;
    *num_curvelets = ctx->num_curvelets;
    if (valid_pairs) *valid_pairs = ctx->valid_pairs;
    if (ctx->num_curvelets > max_rows)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT,
                    "cvg_run_device_noalloc: the output buffers hold " + std::to_string(max_rows) +
                        " rows; this map has " + std::to_string(ctx->num_curvelets) + " curvelets");
    return CVG_OK;
}

// Batches: maps concatenated in edge order; eoff[m] = first edge of map m.
static int batch_offsets(cvg_context *ctx, int num_maps, const int *num_edges, std::vector<long long> &eoff)
{
    if (num_maps <= 0 || !num_edges) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "batch: need num_maps > 0");
    eoff.assign(static_cast<size_t>(num_maps) + 1, 0);
    for (int m = 0; m < num_maps; ++m) {
        if (num_edges[m] < 0) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "batch: negative edge count");
        eoff[m + 1] = eoff[m] + num_edges[m];
    }
    if (eoff[num_maps] <= 0) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "batch: no edges");
    if (eoff[num_maps] > INT_MAX) return fail(ctx, CVG_ERR_UNSUPPORTED, "batch: more than 2^31-1 edges");
    return CVG_OK;
}

int cvg_run_batch(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data, const double *const *edges,
                  const cvg_params *params, cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !edges || sz_edge_data < 4) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_batch: bad arguments");
    std::vector<long long> eoff;
    CVG_TRY(batch_offsets(ctx, num_maps, num_edges, eoff));
    const int n = static_cast<int>(eoff[num_maps]);
    CVG_TRY(ensure(ctx, ctx->d_edges_up, static_cast<size_t>(n) * sz_edge_data * sizeof(double)));
    double *d = as<double>(ctx->d_edges_up);
    for (int m = 0; m < num_maps; ++m) {
        if (num_edges[m] == 0) continue;
        if (!edges[m]) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_batch: null edge array");
        CVG_HIP(ctx, hipMemcpyAsync(d + eoff[m] * sz_edge_data, edges[m],
                                    static_cast<size_t>(num_edges[m]) * sz_edge_data * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    return run_common(ctx, n, sz_edge_data, d, params, stats, num_maps, eoff.data());
}

int cvg_run_batch_device(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data,
                         const double *device_edges, const cvg_params *params, cvg_run_stats *stats)
{
    CVG_TRY(check_ctx(ctx));
    if (!params || !device_edges || sz_edge_data < 4)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_run_batch_device: bad arguments");
    std::vector<long long> eoff;
    CVG_TRY(batch_offsets(ctx, num_maps, num_edges, eoff));
    return run_common(ctx, static_cast<int>(eoff[num_maps]), sz_edge_data, device_edges, params, stats, num_maps,
                      eoff.data());
}

int cvg_get_tuning(cvg_context *ctx, cvg_tuning *tuning)
{
    CVG_TRY(check_ctx(ctx));
    if (!tuning) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_get_tuning: null tuning");
    tuning->coop_min_key = ctx->coop_min_cap;
    tuning->coop_waves = ctx->coop_waves;
    tuning->window_lanes = ctx->fill_lanes;
    tuning->grid_pct = ctx->grid_pct;
    return CVG_OK;
}

int cvg_set_tuning(cvg_context *ctx, const cvg_tuning *tuning)
{
    CVG_TRY(check_ctx(ctx));
    if (!tuning) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_set_tuning: null tuning");
    const int w = tuning->coop_waves, f = tuning->window_lanes, g = tuning->grid_pct;
    if (tuning->coop_min_key < 1 || !(w == 1 || w == 2 || w == 4) || !(f == 8 || f == 16 || f == 32 || f == 64) ||
        !(g == 0 || (g >= 5 && g <= 100)))
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT,
                    "cvg_set_tuning: need coop_min_key >= 1, coop_waves in {1, 2, 4}, window_lanes in {8, 16, 32, 64}, "
                    "grid_pct 0 or 5..100");
    ctx->coop_min_cap = tuning->coop_min_key;
    ctx->coop_waves = w;
    ctx->fill_lanes = f;
    ctx->grid_pct = g;
    return CVG_OK;
}

int cvg_autotune_batch(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data,
                       const double *device_edges, const cvg_params *params, int reps, cvg_tuning *best,
                       double *best_ms, cvg_tuning_trial *trials, int max_trials, int *num_trials)
{
    CVG_TRY(check_ctx(ctx));
    if (reps < 1 || (max_trials > 0 && !trials))
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "cvg_autotune_batch: need reps >= 1 and room for the trials");
    int count = 0;
    // mean wall time of `reps` runs (after one warm-up) under tuning t
    auto time_of = [&](const cvg_tuning &t, double *ms) -> int {
        CVG_TRY(cvg_set_tuning(ctx, &t));
        CVG_TRY(cvg_run_batch_device(ctx, num_maps, num_edges, sz_edge_data, device_edges, params, nullptr));
        CVG_TRY(cvg_synchronize(ctx));
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) {
            CVG_TRY(cvg_run_batch_device(ctx, num_maps, num_edges, sz_edge_data, device_edges, params, nullptr));
            CVG_TRY(finish_curvelets(ctx));
        }
        const auto t1 = std::chrono::steady_clock::now();
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
        if (count < max_trials) trials[count] = cvg_tuning_trial{t, *ms};
        ++count;
        return CVG_OK;
    };
    cvg_tuning cur{};
    CVG_TRY(cvg_get_tuning(ctx, &cur));
    double cur_ms = 0.0;
    CVG_TRY(time_of(cur, &cur_ms));
    // coordinate descent: the cooperative threshold, its waves, the scan
    // lanes, the light buckets' grid
    const int keys[] = {4, 6, 8, 11, 16, 1 << 30};
    const int waves[] = {2, 4};
    const int lanes[] = {8, 16, 32};
    const int grids[] = {25, 50, 75, 100};
    for (int field = 0; field < 4; ++field) {
        const int *vals = field == 0 ? keys : (field == 1 ? waves : (field == 2 ? lanes : grids));
        const int nv = field == 0 ? 6 : (field == 1 ? 2 : (field == 2 ? 3 : 4));
        cvg_tuning base = cur;
        for (int k = 0; k < nv; ++k) {
            cvg_tuning t = base;
            int &slot = field == 0 ? t.coop_min_key
                                   : (field == 1 ? t.coop_waves : (field == 2 ? t.window_lanes : t.grid_pct));
            const int old = field == 0 ? base.coop_min_key
                                       : (field == 1 ? base.coop_waves : (field == 2 ? base.window_lanes : base.grid_pct));
            if (vals[k] == old) continue;
            slot = vals[k];
            double ms = 0.0;
            CVG_TRY(time_of(t, &ms));
            if (ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    }
    CVG_TRY(cvg_set_tuning(ctx, &cur));
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (num_trials) *num_trials = count;
    return CVG_OK;
}

int cvg_batch_map_shape(cvg_context *ctx, int map, unsigned *num_curvelets, unsigned *chain_width)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->curvelets_ready) return fail(ctx, CVG_ERR_STATE, "no curvelets");
    CVG_TRY(finish_curvelets(ctx));
    if (map < 0 || map >= static_cast<int>(ctx->map_rows.size()) - 1)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "map index out of range");
    if (num_curvelets) *num_curvelets = ctx->map_rows[map + 1] - ctx->map_rows[map];
    if (chain_width) *chain_width = ctx->width;
    return CVG_OK;
}

int cvg_get_batch_map_curvelets(cvg_context *ctx, int map, unsigned *chains, double *info)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->curvelets_ready) return fail(ctx, CVG_ERR_STATE, "no curvelets");
    CVG_TRY(finish_curvelets(ctx));
    if (map < 0 || map >= static_cast<int>(ctx->map_rows.size()) - 1)
        return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "map index out of range");
    const size_t r0 = ctx->map_rows[map], rows = ctx->map_rows[map + 1] - r0;
    CVG_TRY(download(ctx, {{chains, as<unsigned>(ctx->d_chains) + r0 * ctx->width, rows * ctx->width * sizeof(unsigned)},
                           {info, as<double>(ctx->d_info) + r0 * kInfoWidth, rows * kInfoWidth * sizeof(double)}}));
    return CVG_OK;
}

int cvg_upload_edges(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges,
                     const double **device_edges)
{
    CVG_TRY(check_ctx(ctx));
    if (!device_edges) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "null output pointer");
    CVG_TRY(upload(ctx, num_edges, sz_edge_data, edges, device_edges));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CVG_OK;
}

int cvg_device_curvelets(cvg_context *ctx, const unsigned **device_chains, const double **device_info)
{
    CVG_TRY(check_ctx(ctx));
    if (!ctx->curvelets_ready) return fail(ctx, CVG_ERR_STATE, "no curvelets");
    CVG_TRY(finish_curvelets(ctx));
    if (device_chains) *device_chains = as<unsigned>(ctx->d_chains);
    if (device_info) *device_info = as<double>(ctx->d_info);
    return CVG_OK;
}

int cvg_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr || bytes == 0) return CVG_ERR_INVALID_ARGUMENT;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return CVG_ERR_OUT_OF_MEMORY;
    }
    return CVG_OK;
}

void cvg_host_free(void *ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

int cvg_get_stream(cvg_context *ctx, void **hip_stream)
{
    if (!ctx || !hip_stream) return CVG_ERR_INVALID_ARGUMENT;
    *hip_stream = ctx->stream;
    return CVG_OK;
}

int cvg_recent_fused_ms(cvg_context *ctx, int max, float *ms, int *count)
{
    CVG_TRY(check_ctx(ctx));
    if (!ms || !count || max < 0) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "recent_fused_ms: bad arguments");
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const long long have = std::min<long long>(ctx->fused_runs, cvg_context::kFusedRing);
    const int k = static_cast<int>(std::min<long long>(have, max));
    for (int i = 0; i < k; ++i) {
        const int r = static_cast<int>((ctx->fused_runs - k + i) % cvg_context::kFusedRing);
        float t = 0.0f;
        CVG_HIP(ctx, hipEventElapsedTime(&t, ctx->fused_ev[r][0], ctx->fused_ev[r][1]));
        ms[i] = t;
    }
    *count = k;
    return CVG_OK;
}

int cvg_set_event_timing(cvg_context *ctx, int on)
{
    CVG_TRY(check_ctx(ctx));
    ctx->event_timing = on != 0;
    return CVG_OK;
}

int cvg_synchronize(cvg_context *ctx)
{
    CVG_TRY(check_ctx(ctx));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CVG_OK;
}

int cvg_debug_anchor_bundles(cvg_context *ctx, const cvg_params *params, int anchor, unsigned char *dir,
                             unsigned char *hyp, double *bmin, double *bmax, int *deg_out, int *cells_out)
{
    CVG_TRY(check_ctx(ctx));
    if (!params) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "null params");
    if (!ctx->graph_ready) return fail(ctx, CVG_ERR_STATE, "no neighbour graph");
    if (anchor < 0 || anchor >= ctx->n) return fail(ctx, CVG_ERR_INVALID_ARGUMENT, "anchor out of range");
    Derived d;
    CVG_TRY(derive(ctx, *params, d));
    unsigned off2[2];
    CVG_HIP(ctx, hipMemcpyAsync(off2, as<unsigned>(ctx->d_off) + anchor, 2 * sizeof(unsigned), hipMemcpyDeviceToHost,
                                ctx->stream));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int deg = static_cast<int>(off2[1] - off2[0]);
    const int K = static_cast<int>(d.K);
    if (deg_out) *deg_out = deg;
    if (cells_out) *cells_out = K;
    if (deg == 0) return CVG_OK;
    std::vector<char> blob;
    const size_t o_x = put(blob, d.off_x);
    const size_t o_t = put(blob, d.off_t);
    const size_t o_out = (blob.size() + 15) / 16 * 16;
    const size_t out_doubles = static_cast<size_t>(2) * deg * K + 2 * deg;
    CVG_TRY(ensure(ctx, ctx->d_dbg, o_out + out_doubles * sizeof(double)));
    char *base = static_cast<char *>(ctx->d_dbg.ptr);
    CVG_HIP(ctx, hipMemcpyAsync(base, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream));
    double *out = reinterpret_cast<double *>(base + o_out);
    hipLaunchKernelGGL(k_debug_anchor, dim3(blocks_for(static_cast<long long>(deg) * K)), dim3(kBlock), 0, ctx->stream,
                       anchor, ctx->n, as<double>(ctx->d_x), as<double>(ctx->d_y), as<double>(ctx->d_th),
                       as<unsigned>(ctx->d_off), as<int>(ctx->d_ids), d.dt, d.dx, d.dx2, d.close_thr, d.max_k,
                       static_cast<int>(d.kx), static_cast<int>(d.kt), reinterpret_cast<double *>(base + o_x),
                       reinterpret_cast<double *>(base + o_t), out);
    CVG_HIP(ctx, hipGetLastError());
    std::vector<double> h(out_doubles);
    CVG_HIP(ctx, hipMemcpyAsync(h.data(), out, out_doubles * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < deg; ++r) {
        bool valid = false;
        for (int c = 0; c < K; ++c) {
            const double lo = h[static_cast<size_t>(r) * K + c];
            const double hi = h[static_cast<size_t>(deg) * K + static_cast<size_t>(r) * K + c];
            if (bmin) bmin[static_cast<size_t>(r) * K + c] = lo;
            if (bmax) bmax[static_cast<size_t>(r) * K + c] = hi;
            valid = valid || (hi > lo);
        }
        const bool close = h[static_cast<size_t>(2) * deg * K + deg + r] != 0.0;
        if (dir) dir[r] = static_cast<unsigned char>(h[static_cast<size_t>(2) * deg * K + r]);
        if (hyp) hyp[r] = (close || valid) ? 1 : 0;
    }
    return CVG_OK;
}

int cvg_debug_cell_order(int num_edges, const double *x, const double *y, int *rank)
{
    if (num_edges < 0 || (num_edges > 0 && (!x || !y || !rank))) return CVG_ERR_INVALID_ARGUMENT;
    try {
        cell_ranks(num_edges, x, y, rank);
    } catch (...) {
        return CVG_ERR_OUT_OF_MEMORY;
    }
    return CVG_OK;
}

int cvg_math_eval(cvg_context *ctx, int op, int n, const double *x, const double *y, double *out)
{
    if (!x || !out || n < 0 || op < 0 || op > 18) return CVG_ERR_INVALID_ARGUMENT;
    if ((op == 2 || op == 5 || op == 8 || op == 9 || op == 16) && !y) return CVG_ERR_INVALID_ARGUMENT;
    if (!ctx) {
        for (int i = 0; i < n; ++i)
            out[i] = math_op(op, x[i], y ? y[i] : 0.0, cvg::kSinCosTableHost, cvg::kAtanTableHost);
        return CVG_OK;
    }
    CVG_TRY(check_ctx(ctx));
    if (n == 0) return CVG_OK;
    const size_t bytes = static_cast<size_t>(n) * sizeof(double);
    CVG_TRY(ensure(ctx, ctx->d_dbg, 3 * bytes));
    double *dx = as<double>(ctx->d_dbg);
    double *dy = dx + n;
    double *dout = dy + n;
    CVG_HIP(ctx, hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, ctx->stream));
    if (y) CVG_HIP(ctx, hipMemcpyAsync(dy, y, bytes, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_math_eval, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, op, n, dx, y ? dy : nullptr,
                       dout);
    CVG_HIP(ctx, hipGetLastError());
    CVG_HIP(ctx, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CVG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CVG_OK;
}

}  // extern "C"
