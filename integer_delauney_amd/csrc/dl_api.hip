// C ABI of libdelauney_hip.so (include/delauney_hip.h) and the host-side
// orchestration of the batch pipeline:
//
//   seeds -> (id order, rank map) -> flood fill on the detection canvas
//         -> triangle extraction -> pixel assignment -> (H, W, 3) output
//
// Reference: src/cuda/bindings/*.hpp (the Python-facing calls this replaces)
// and src/delauney/reference/ (the semantics computed).
#include "../../include/delauney_hip.h"
#include "dl_internal.h"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>

namespace dl {

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line)
{
    const char* base = std::strrchr(file, '/');
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e),
                  base ? base + 1 : file, line);
    (void)hipGetLastError();  // clear the error so the process stays usable
    throw runtime_error(buf);
}

namespace {
std::atomic<uint64_t> g_alloc_epoch{0};
thread_local CaptureLog* g_capture_log = nullptr;

// A buffer is about to be reallocated or freed.
void note_realloc()
{
    if (g_capture_log) throw capture_abort("buffer reallocation during stream capture");
    g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
}

thread_local Arena* g_arena = nullptr;

// Bytes [used, used + n) of an arena region, 256-byte aligned.
void* carve(char* base, size_t cap, size_t& used, size_t n, const char* what)
{
    const size_t at = (used + 255) & ~size_t(255);
    if (at > cap || n > cap - at)
        throw runtime_error(std::string("the no-allocation scratch is exhausted (") + what + ")");
    used = at + n;
    return base + at;
}
}  // namespace

Arena*& current_arena() { return g_arena; }

DevBuf::~DevBuf()
{
    if (!ptr || arena) return;
    // No throw from a destructor: a capture that outlives a buffer is dropped.
    if (g_capture_log) g_capture_log->poisoned = true;
    g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
    (void)hipFree(ptr);
}

uint64_t alloc_epoch() { return g_alloc_epoch.load(std::memory_order_relaxed); }

CaptureLog*& capture_log() { return g_capture_log; }

void stage_h2d(PinnedBuf& stage, size_t offset, void* dst, const void* src, size_t n,
               hipStream_t s)
{
    if (n == 0) return;
    if (!stage.ptr || offset + n > stage.bytes) throw runtime_error("staging buffer too small");
    char* h = static_cast<char*>(stage.ptr) + offset;
    std::memcpy(h, src, n);
    if (g_capture_log) {   // issued before each replay instead (CaptureLog)
        const char* b = static_cast<const char*>(src);
        g_capture_log->staged.push_back(StagedBytes{h, dst, std::vector<char>(b, b + n)});
        return;
    }
    DL_HIP(hipMemcpyAsync(dst, h, n, hipMemcpyHostToDevice, s));
}

void copy_back_async(void* host, const void* dev, size_t n, hipStream_t s)
{
    if (n == 0) return;
    if (g_capture_log) {   // issued after each replay instead (CaptureLog)
        g_capture_log->copy_back.push_back(CopyBack{host, dev, n});
        return;
    }
    DL_HIP(hipMemcpyAsync(host, dev, n, hipMemcpyDeviceToHost, s));
}

void DevBuf::release()
{
    if (ptr && !arena) {
        note_realloc();
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    arena = false;
}

void DevBuf::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    if (g_arena) {
        if (ptr && !arena) throw runtime_error("an allocated buffer grew inside an arena");
        const size_t want = std::max<size_t>(n, 256);
        ptr = carve(g_arena->dev, g_arena->dev_cap, g_arena->dev_used, want, "device");
        bytes = want;
        arena = true;
        return;
    }
    if (g_capture_log) throw capture_abort("buffer growth during stream capture");
    const size_t old = ptr && !arena ? bytes : 0;
    release();
    g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
    size_t want = std::max<size_t>(n, 256);
    void* p = nullptr;
    // A buffer that grows takes half its old size again (at most 1 GiB more),
    // so that sizes creeping up from call to call do not reallocate each time.
    const size_t grown = std::max(want, old + std::min<size_t>(old / 2, size_t(1) << 30));
    hipError_t e = hipMalloc(&p, grown);
    if (e == hipSuccess) {
        want = grown;
    } else if (grown > want) {
        (void)hipGetLastError();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        char buf[256];
        std::snprintf(buf, sizeof(buf), "hipMalloc of %zu bytes failed: %s (dl_api.hip)", want,
                      hipGetErrorString(e));
        throw runtime_error(buf);
    }
    ptr = p;
    bytes = want;
}

PinnedBuf::~PinnedBuf()
{
    if (ptr && !arena) {
        if (g_capture_log) g_capture_log->poisoned = true;
        g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
        (void)hipHostFree(ptr);
    }
}

void PinnedBuf::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    if (g_arena) {
        if (ptr && !arena) throw runtime_error("an allocated buffer grew inside an arena");
        const size_t want = std::max<size_t>(n, 256);
        ptr = carve(g_arena->host, g_arena->host_cap, g_arena->host_used, want, "host");
        bytes = want;
        arena = true;
        return;
    }
    note_realloc();
    const size_t old = ptr && !arena ? bytes : 0;
    if (ptr && !arena) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    arena = false;
    // (grows like DevBuf)
    const size_t want =
        std::max(std::max<size_t>(n, 256), old + std::min<size_t>(old / 2, size_t(1) << 30));
    DL_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    bytes = want;
}

namespace {
thread_local KernelProfiler* g_profiler = nullptr;
}

KernelProfiler* current_profiler() { return g_profiler; }
void set_current_profiler(KernelProfiler* p) { g_profiler = p; }

hipEvent_t KernelProfiler::take()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    DL_HIP(hipEventCreate(&e));
    return e;
}

void KernelProfiler::collect()
{
    for (const auto& r : pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            count[r.kind] += 1;
            ms[r.kind] += t;
        }
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}

void KernelProfiler::reset()
{
    for (int i = 0; i < kKinds; ++i) {
        count[i] = 0;
        ms[i] = 0.0;
    }
}

KernelProfiler::~KernelProfiler()
{
    for (auto& r : pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(int kind, hipStream_t stream) : p(g_profiler), s(stream)
{
    if (!p) return;
    rec.kind = kind;
    rec.a = p->take();
    rec.b = p->take();
    DL_HIP(hipEventRecord(rec.a, s));
}

ProfScope::~ProfScope()
{
    if (!p) return;
    if (hipEventRecord(rec.b, s) == hipSuccess) p->pending.push_back(rec);
}

namespace {

__global__ void extract_ids_kernel(const int32_t* __restrict__ hw2, uint32_t* __restrict__ ids,
                                   int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = static_cast<uint32_t>(hw2[2 * i]);
}

}  // namespace
}  // namespace dl

// Device scratch of dl_triangulate_device_noalloc (see g_scratch_dev).
static __device__ __align__(256) unsigned char scratch[DL_NOALLOC_SCRATCH_BYTES];

using namespace dl;

namespace {

// A padded fill issued as a fixed number of launches without waiting; the
// caller settles it after its next synchronisation (finish_predicted_fill).
struct FillPending {
    bool active = false;
    int n = 0;                // launches issued (their n changed flags and the visit
                              // total come back with the final status read)
    BfsConfig cfg{};
    std::tuple<int, int, int, int, int, int> key;
};

// What the second half of the pipeline (finish_pipeline) needs from the first.
struct Issued {
    FillPending pend;
    int plane = 0;          // label plane holding the padded fill's result
    bool any_pad = false;   // some image's padding is watched
    int C = 3;              // output channels
    bool fixed_bins = true; // the speculative assignment may bin into fixed slices
};

// The first half of the pipeline captured as a graph, with the host state its
// replay restores.
struct PipelineGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> sig;     // call signature (pipeline_sig)
    uint64_t epoch = 0;           // allocation epoch at capture
    BatchPlan plan;               // the context's plan after issuing
    Issued iss;
    std::vector<BfsCanvas> fill_cvs;
    BfsConfig last_cfg{};
    uint32_t rows_cap = 0;
    int64_t list_cap = 0;
    std::vector<StagedBytes> staged;    // uploads issued before each replay
    std::vector<CopyBack> copy_back;    // read-backs issued after it

    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        sig.clear();
        staged.clear();
        copy_back.clear();
    }
};

}  // namespace

struct dl_context {
    int device = 0;
    hipStream_t stream = nullptr;
    BfsWork bfs;
    // Set by a session's finalise that keeps a fill record (see FillRecord):
    // the record and the session's seed bitmap; the pipeline's padded fill
    // then records, or replays from the record.  last_escaped: the padding
    // watch of the last pipeline fired for some image.
    FillRecord* rec = nullptr;
    const uint32_t* rec_bitmap = nullptr;
    int rec_w = 0, rec_h = 0;
    bool last_escaped = false;
    int64_t record_fills = 0, replay_fills = 0;
    // padA holds the final labels of the fill recorded as generation
    // plane_gen (FillRecord::gen; 0: no claim) at address plane_ptr.  Every
    // fill and graph replay on the context drops the claim before it writes
    // the planes; a recorded or replayed fill makes it again.
    uint64_t plane_gen = 0;
    const void* plane_ptr = nullptr;
    int64_t retained_fills = 0;
    TriWork tri;
    BatchPlan plan;   // images of the last call
    DevBuf xy, sorted;
    DevBuf padA, padB, unA, unB;
    DevBuf out, vgrid, ids, esc;
    PinnedBuf host_small;

    // state of the last triangulation
    int64_t n_tri = 0;   // all images
    bool have_padded = false;
    int pad_plane = 0;

    bool timing = false;
    double timings[9] = {0};
    // Launches of the last call's fills up to and including the first that
    // changed nothing (launches issued past convergence return at once).
    int64_t fill_converged = 0;
    // Predicted fills on this context that fell short and were continued.
    int64_t fill_shortfalls = 0;
    BfsConfig cfg{8, 8, 8};
    bool cfg_auto = true;
    bool autotune = true;   // auto mode may time the candidate shapes on a new canvas class
    bool predict = true;    // padded fills of known classes issued without waiting
    // Launches the padded fill of a canvas class and shape needed last time
    // (plus a margin): the next fill of the class issues that many without
    // waiting (see fill_detection).
    std::map<std::tuple<int, int, int, int, int, int>, int> fill_pred;
    std::vector<BfsCanvas> fill_cvs;   // canvases of the last padded fill
    BfsConfig last_cfg{0, 0, 0};
    bool force_unpadded = false;
    bool force_full_replay = false;   // (test hook) replays re-initialise the planes
    std::vector<hipEvent_t> events;
    bool profiling = false;
    KernelProfiler prof;
    // Canvas classes (as for the fill autotuner) whose fixed bin slices
    // overflowed once: their assignment bins with count + scan + fill.
    std::set<std::tuple<int, int, int>> bins_exact;
    // hipGraph replay of the pipeline (see run_pipeline)
    // 0 off, 1 on, 2 auto (the default): capture pipelines of fewer than
    // kGraphAutoMaxCells canvas cells only.  A replay saves the host's issue
    // work (small pipelines are bound by it) but costs about 0.5 us of GPU
    // time per kernel node more than launches issued from the host.
    int graph_mode = 2;
    PipelineGraph graph;
    std::vector<int64_t> last_sig;   // signature of the last uncaptured predicted run
    std::vector<int64_t> refused_sig;   // the last signature whose capture failed
    uint64_t last_epoch = ~0ull;
    int64_t graph_captures = 0, graph_replays = 0;
};

struct dl_session {
    int device = 0;
    int W = 0, H = 0, P = 0;
    SessionState st;
    DevBuf stage;        // the call's uploaded (x, y) pairs
    DevBuf stage_values; // and values
    DevBuf out;          // (H, W, 3) grid of the last finalise
    DevBuf rows;         // its triangle rows
    DevBuf sorted;       // its seeds' packed labels in sorted id order
    int64_t n_tri = 0;
    bool done = false;   // out/rows/sorted hold the current seed set's result
    DevBuf edges;        // distinct edges of the current result
    int64_t n_edges = -1;   // -1: not built since the last change
    DevBuf pix_tids, pix_sids, pix_mask;   // per-pixel arrays of dl_session_finalise_device
    DevBuf q_xy, q_t, q_tid, q_mask, q_out;   // query staging
    QueryWork qw;
    // dl_session_set_fill_record: keep the fill's trajectory between
    // finalises, so that the next one computes only what the changes reach
    bool record = false;
    FillRecord rec;
};

namespace {

thread_local std::string g_last_error;

// Installs the context's kernel profiler for the duration of one call.
struct ProfInstall {
    dl_context* ctx;
    explicit ProfInstall(dl_context* c) : ctx(c)
    {
        if (ctx && ctx->profiling) set_current_profiler(&ctx->prof);
    }
    ~ProfInstall()
    {
        if (ctx && ctx->profiling) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->prof.collect();
        }
        set_current_profiler(nullptr);
    }
};

template <class F>
int guarded(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return DL_OK;
    } catch (const value_error& e) {
        g_last_error = e.what();
        return DL_EVALUE;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return DL_EVALUE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return DL_ERUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return DL_ERUNTIME;
    }
}

int resolve_padding(int p)
{
    if (p == -1) return DL_DEFAULT_BORDER_PADDING;
    return p < 0 ? 0 : p;
}

// A canvas beyond the library's limit is a RuntimeError, as the reference's
// failing device allocation of an oversized canvas is
// (tests/test_cuda_error_checking.py:28-58); non-positive sides are a
// ValueError.
void check_dims(int W, int H, int P)
{
    if (W <= 0 || H <= 0) throw value_error("width and height must be positive");
    if (static_cast<int64_t>(W) + 2LL * P > DL_MAX_CANVAS ||
        static_cast<int64_t>(H) + 2LL * P > DL_MAX_CANVAS)
        throw runtime_error("canvas (image plus twice the border padding) exceeds " +
                            std::to_string(DL_MAX_CANVAS) +
                            " cells per side, the limit of 16-bit packed seed positions");
}

void check_seeds(const void* p, int64_t n)
{
    if (n <= 0) throw value_error("seeds must not be empty");
    if (!p) throw value_error("seeds pointer is null");
}

// ---------------------------------------------------------------------------
// Flood-fill autotuner.  Every tile shape computes the same labels (the tiled
// schedule is exact), so the shape is free to choose per input.  A canvas
// class is (device, log2 of the fill's total cells, log4 of cells per seed);
// the first fill of a class large enough to matter times each candidate shape
// on the live input (re-initialising the canvases between runs; the last
// run's labels are the result) and keeps the fastest for the process.
// Smaller fills use the default shape.
// ---------------------------------------------------------------------------
constexpr BfsConfig kDefaultCfg{8, 8, 8};
constexpr BfsConfig kTuneCandidates[] = {{8, 8, 8}, {16, 8, 8}, {16, 8, 12}, {16, 8, 16},
                                         {16, 4, 8}, {8, 8, 4}};
constexpr int64_t kTuneMinCells = int64_t(1) << 24;

using TuneKey = std::tuple<int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, BfsConfig> g_tuned;

TuneKey tune_key(int device, int64_t cells, int64_t seeds)
{
    const double c = static_cast<double>(std::max<int64_t>(cells, 1));
    const double per = c / static_cast<double>(std::max<int64_t>(seeds, 1));
    return TuneKey{device, static_cast<int>(std::floor(std::log2(c))),
                   static_cast<int>(std::floor(std::log2(std::max(per, 1.0)) / 2.0))};
}

constexpr int kMaxPredicted = 512;

// Launches to issue for a class whose last fill converged at launch index c:
// a margin absorbs small changes of the input (launches after convergence find
// no active tile and cost a few microseconds; a shortfall costs a second
// detection and assignment).
int predicted_launches(int c) { return c + 1 + 2 + c / 16; }

std::tuple<int, int, int, int, int, int> pred_key(int device, int64_t cells, int64_t seeds,
                                                  const BfsConfig& c)
{
    const TuneKey k = tune_key(device, cells, seeds);
    return {std::get<0>(k), std::get<1>(k), std::get<2>(k), c.waves, c.rows, c.iters};
}

// The shape for a fill of `cells` cells and `seeds` seeds; *tune is set when
// the class should be timed now.
BfsConfig choose_config(const dl_context* ctx, int64_t cells, int64_t seeds, bool* tune)
{
    *tune = false;
    if (!ctx->cfg_auto) return ctx->cfg;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        const auto it = g_tuned.find(tune_key(ctx->device, cells, seeds));
        if (it != g_tuned.end()) return it->second;
    }
    // (a session keeping a fill record fills with the default shape: the
    // shape is part of the record)
    *tune = ctx->autotune && cells >= kTuneMinCells && !ctx->rec;
    return kDefaultCfg;
}

// Stage timing.  With ctx->timing on, every mark(idx) records an event on the
// context's stream and the interval since the previous mark is added to
// stage idx (0 seeds, 1 fill, 2 detect, 3 dedup/rank, 4 assign, 5 output), so
// a stage run twice (a continued fill, an exact rerun) accumulates.
struct StageClock {
    dl_context* ctx;
    std::chrono::steady_clock::time_point t0;
    std::vector<int> stage;   // stage of the interval ending at event k + 1
    explicit StageClock(dl_context* c) : ctx(c), t0(std::chrono::steady_clock::now())
    {
        for (double& v : ctx->timings) v = 0.0;
        ctx->fill_converged = 0;
        if (ctx->timing) {
            event(0);
            DL_HIP(hipEventRecord(ctx->events[0], ctx->stream));
        }
    }
    hipEvent_t event(size_t k)
    {
        while (ctx->events.size() <= k) {
            hipEvent_t e;
            DL_HIP(hipEventCreate(&e));
            ctx->events.push_back(e);
        }
        return ctx->events[k];
    }
    void mark(int idx)
    {
        if (!ctx->timing) return;
        DL_HIP(hipEventRecord(event(stage.size() + 1), ctx->stream));
        stage.push_back(idx);
    }
    void finish(hipStream_t s)
    {
        DL_HIP(hipStreamSynchronize(s));
        if (ctx->timing) {
            for (size_t k = 0; k < stage.size(); ++k) {
                float ms = 0.f;
                DL_HIP(hipEventElapsedTime(&ms, ctx->events[k], ctx->events[k + 1]));
                ctx->timings[stage[k]] += ms;
            }
        }
        ctx->timings[6] =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Plans ctx->plan (images already listed), uploads it and orders the seeds on
// the device into ctx->sorted (id order per image) and the seed index.  Images
// without their own seed / output pointers read seeds from d_xy and write
// into d_out at their offsets in the concatenated layout (C channels).
void prepare(dl_context* ctx, const int32_t* d_xy, int32_t* d_out, int C, hipStream_t s,
             SeedCheck check = SeedCheck::kNow)
{
    BatchPlan& plan = ctx->plan;
    plan_batch(plan);
    for (ImageDesc& d : plan.img) {
        if (!d.xy) d.xy = d_xy + 2 * d.seed_off;
        if (!d.out && d_out) d.out = d_out + static_cast<int64_t>(C) * d.pix_off;
    }
    upload_plan(plan, ctx->tri, s);
    ctx->sorted.reserve(sizeof(uint32_t) * static_cast<size_t>(std::max<int64_t>(plan.n_seeds, 1)));
    normalise_seeds(plan, ctx->sorted.as<uint32_t>(), ctx->tri, s, check);
}

// Launches a padded fill of ctx->plan with shape cfg is issued as without
// waiting (0: none known, or the shape is being tuned; the fill then runs to
// convergence with waits).
int fill_prediction(const dl_context* ctx, const BfsConfig& cfg, bool tune)
{
    if (tune || !ctx->predict || ctx->rec) return 0;
    const auto it =
        ctx->fill_pred.find(pred_key(ctx->device, ctx->plan.n_cells, ctx->plan.n_seeds, cfg));
    return it != ctx->fill_pred.end() && it->second <= kMaxPredicted ? it->second : 0;
}

// Flood fill of every image's detection canvas ((W + 2P) x (H + 2P), seeds
// shifted by P).  watch[i] >= 0 records in ctx->esc[i] whether a cell of
// [m, Wc-m) x [m, Hc-m) ever received an update from outside it.  Returns the
// plane (0: padA, 1: padB) holding every canvas's result.
// visits_async: see bfs_run (the autotuning path always synchronises).
//
// pend: when given and an earlier fill of the same class and shape converged,
// the fill is issued as that many launches plus a margin (predicted_launches)
// without any wait, and *pend describes what finish_predicted_fill must check.
int fill_detection(dl_context* ctx, const std::vector<int>& watch, hipStream_t s,
                   int* visits_async = nullptr, FillPending* pend = nullptr)
{
    const BatchPlan& plan = ctx->plan;
    const int n_img = static_cast<int>(plan.img.size());
    const size_t cells = static_cast<size_t>(std::max<int64_t>(plan.n_cells, 1));
    ctx->padA.reserve(sizeof(uint32_t) * cells);
    ctx->padB.reserve(sizeof(uint32_t) * cells);
    ctx->esc.reserve(sizeof(int) * static_cast<size_t>(n_img));
    const FillSeg esc_clear{ctx->esc.as<uint32_t>(), static_cast<size_t>(n_img), 0u};
    BfsStats st;
    bool tune = false;
    BfsConfig cfg = choose_config(ctx, plan.n_cells, plan.n_seeds, &tune);
    int plane = 0;
    const auto pkey = pred_key(ctx->device, plan.n_cells, plan.n_seeds, cfg);
    const int pred = pend ? fill_prediction(ctx, cfg, tune) : 0;
    // A session's replay may start from the first plane as its last fill left
    // it (same record generation, same buffer); anything else re-initialises.
    BfsCanvas probe{};
    if (n_img == 1) {
        probe.W = plan.img[0].Wc;
        probe.H = plan.img[0].Hc;
        probe.P = watch[0] >= 0 ? watch[0] : 0;
    }
    const bool recorded = ctx->rec && !tune && n_img == 1 && bfs_record_supported(cfg, probe);
    const bool replay = recorded && bfs_record_usable(*ctx->rec, probe, cfg) &&
                        ctx->rec->prev_bitmap.ptr;
    const bool retained = replay && !ctx->force_full_replay && ctx->plane_gen != 0 &&
                          ctx->plane_gen == ctx->rec->gen && ctx->plane_ptr == ctx->padA.ptr &&
                          plan.img[0].cv_off == 0;
    ctx->plane_gen = 0;
    ctx->plane_ptr = nullptr;
    // A predicted fill's launch counters are cleared in the canvases' clear
    // launch.
    FillSeg cnt_clear{};
    if (pred > 0) {
        ctx->bfs.counters.reserve(sizeof(int) * bfs_fixed_counter_words(pred));
        cnt_clear = FillSeg{ctx->bfs.counters.as<uint32_t>(), bfs_fixed_counter_words(pred), 0u};
    }
    if (retained)
        fill_u32(esc_clear.p, 0u, esc_clear.n, s);
    else
        init_canvases(plan, ctx->sorted.as<uint32_t>(), ctx->padA.as<uint32_t>(),
                      ctx->padB.as<uint32_t>(), ctx->tri, s, &esc_clear,
                      pred > 0 ? &cnt_clear : nullptr);
    std::vector<BfsCanvas> cvs(n_img);
    for (int i = 0; i < n_img; ++i) {
        const ImageDesc& d = plan.img[i];
        BfsCanvas& cv = cvs[i];
        cv = BfsCanvas{};
        cv.lab0 = ctx->padA.as<uint32_t>() + d.cv_off;
        cv.lab1 = ctx->padB.as<uint32_t>() + d.cv_off;
        cv.W = d.Wc;
        cv.H = d.Hc;
        cv.P = watch[i] >= 0 ? watch[i] : 0;
        cv.escape = watch[i] >= 0 ? ctx->esc.as<int>() + i : nullptr;
    }
    if (recorded) {
        // A session keeping its fill record: replay from it when it describes
        // this canvas, otherwise record this fill (host waits, no prediction).
        FillRecord& rec = *ctx->rec;
        if (replay) {
            plane = bfs_run_incremental(cvs, cfg, ctx->bfs, s, &st, rec,
                                        rec.prev_bitmap.as<uint32_t>(), ctx->rec_bitmap,
                                        ctx->rec_w, ctx->rec_h, plan.img[0].P, retained);
            ++ctx->replay_fills;
            if (retained) ++ctx->retained_fills;
        } else {
            plane = bfs_run_recorded(cvs, cfg, ctx->bfs, s, &st, rec);
            ++ctx->record_fills;
        }
        // After convergence both planes hold the final labels; the next
        // replay of this record may start from the first one.
        if (rec.valid && plan.img[0].cv_off == 0) {
            ctx->plane_gen = rec.gen;
            ctx->plane_ptr = ctx->padA.ptr;
        }
        ctx->fill_converged += st.converged + 1;
        if (visits_async) *visits_async = st.tile_visits;
        st.tile_visits = 0;
    } else if (pred > 0) {
        // change flags and visit total stay on the device for the final status
        // read; visits are counted only for statistics
        ctx->bfs.count_visits = ctx->timing || ctx->profiling;
        plane = bfs_launch_fixed(cvs, cfg, ctx->bfs, s, pred, nullptr, true);
        pend->active = true;
        pend->n = pred;
        pend->cfg = cfg;
        pend->key = pkey;
        ctx->fill_cvs = cvs;
        st.launches = pred;
    } else if (!tune) {
        plane = bfs_run(cvs, cfg, ctx->bfs, s, &st, visits_async);
        ctx->fill_converged += st.converged + 1;
        if (pend) ctx->fill_pred[pkey] = predicted_launches(st.converged);
    } else {
        hipEvent_t ev[2];
        DL_HIP(hipEventCreate(&ev[0]));
        DL_HIP(hipEventCreate(&ev[1]));
        // Two timed runs per candidate (the faster counts); a shape replaces
        // the default only when it is clearly faster, so timing noise does
        // not move the choice.
        float best = 0.f, dflt = 0.f;
        bool first = true;
        for (const BfsConfig& cand : kTuneCandidates) {
            if (!bfs_config_supported(cand)) continue;
            float t_min = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                if (!first) {   // start again from the seeds
                    init_canvases(plan, ctx->sorted.as<uint32_t>(), ctx->padA.as<uint32_t>(),
                                  ctx->padB.as<uint32_t>(), ctx->tri, s, &esc_clear);
                }
                first = false;
                st = BfsStats{};
                DL_HIP(hipEventRecord(ev[0], s));
                plane = bfs_run(cvs, cand, ctx->bfs, s, &st);
                DL_HIP(hipEventRecord(ev[1], s));
                DL_HIP(hipEventSynchronize(ev[1]));
                float ms = 0.f;
                DL_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
                t_min = rep == 0 ? ms : std::min(t_min, ms);
            }
            const bool is_default = cand.waves == kDefaultCfg.waves && cand.rows == kDefaultCfg.rows &&
                                    cand.iters == kDefaultCfg.iters;
            if (is_default) {
                dflt = t_min;
                if (best == 0.f || t_min < best) {
                    best = t_min;
                    cfg = cand;
                }
            } else if (t_min < 0.97f * dflt && (best == 0.f || t_min < best)) {
                best = t_min;
                cfg = cand;
            }
        }
        DL_HIP(hipEventDestroy(ev[0]));
        DL_HIP(hipEventDestroy(ev[1]));
        if (visits_async) *visits_async = st.tile_visits;
        st.tile_visits = 0;
        ctx->fill_converged += st.converged + 1;
        if (pend) ctx->fill_pred[pred_key(ctx->device, plan.n_cells, plan.n_seeds, cfg)] =
            predicted_launches(st.converged);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(ctx->device, plan.n_cells, plan.n_seeds)] = cfg;
    }
    ctx->last_cfg = cfg;
    ctx->timings[7] += st.launches;
    ctx->timings[8] += st.tile_visits;
    ctx->have_padded = true;
    ctx->pad_plane = plane;
    return plane;
}

const uint32_t* pad_plane_ptr(const dl_context* ctx, int plane)
{
    return plane ? ctx->padB.as<uint32_t>() : ctx->padA.as<uint32_t>();
}

// After the stream's next synchronisation, with the launches' read-back change
// flags and visit total: did the predicted launches reach convergence?  If not, the same fill continues to convergence here (both
// planes then hold the final labels) and false is returned: everything
// computed from the labels must be computed again.
bool finish_predicted_fill(dl_context* ctx, const FillPending& pend, const int* changed,
                           hipStream_t s, int* visits)
{
    int conv = -1;
    for (int i = 0; i < pend.n; ++i)
        if (changed[i] == 0) {
            conv = i;
            break;
        }
    *visits += changed[pend.n];
    if (conv >= 0) {
        ctx->fill_converged += conv + 1;
        // Keep the count while it still leaves the margin, so that inputs of
        // one class whose fills converge a launch apart keep one signature
        // (and a captured graph); shrink it once it is far too generous.
        const int want = predicted_launches(conv);
        int& cur = ctx->fill_pred[pend.key];
        if (want > cur || cur > want + 2 + want / 8) cur = want;
        return true;
    }
    BfsStats st;
    ++ctx->fill_shortfalls;
    bfs_run(ctx->fill_cvs, pend.cfg, ctx->bfs, s, &st, nullptr, pend.n);
    ctx->timings[7] += st.launches - pend.n;
    ctx->fill_converged += st.converged + 1;
    *visits += st.tile_visits;
    ctx->fill_pred[pend.key] = predicted_launches(st.converged);
    return false;
}

// Unpadded fills (image-sized canvases) for the listed images; sets their
// output label source.
void fill_unpadded(dl_context* ctx, const std::vector<int>& which, hipStream_t s)
{
    BatchPlan& plan = ctx->plan;
    std::vector<int64_t> off(which.size());
    int64_t total = 0;
    for (size_t k = 0; k < which.size(); ++k) {
        off[k] = total;
        const ImageDesc& d = plan.img[which[k]];
        total += static_cast<int64_t>(d.W) * d.H;
    }
    ctx->unA.reserve(sizeof(uint32_t) * static_cast<size_t>(total));
    ctx->unB.reserve(sizeof(uint32_t) * static_cast<size_t>(total));
    std::vector<BfsCanvas> cvs(which.size());
    for (size_t k = 0; k < which.size(); ++k) {
        const ImageDesc& d = plan.img[which[k]];
        BfsCanvas& cv = cvs[k];
        cv = BfsCanvas{};
        cv.lab0 = ctx->unA.as<uint32_t>() + off[k];
        cv.lab1 = ctx->unB.as<uint32_t>() + off[k];
        cv.W = d.W;
        cv.H = d.H;
        bfs_init_canvas(cv, ctx->sorted.as<uint32_t>() + d.seed_off, d.n, s);
    }
    BfsStats st;
    const int plane = bfs_run(cvs, ctx->last_cfg, ctx->bfs, s, &st);
    ctx->timings[7] += st.launches;
    ctx->timings[8] += st.tile_visits;
    ctx->fill_converged += st.converged + 1;
    for (size_t k = 0; k < which.size(); ++k) {
        ImageDesc& d = plan.img[which[k]];
        d.out_lab = (plane ? ctx->unB.as<uint32_t>() : ctx->unA.as<uint32_t>()) + off[k];
        d.out_W = d.W;
        d.out_off = 0;
        d.out_shift = 0;
    }
}

// The batch pipeline on device-resident seeds (ctx->plan lists the images):
// seeds -> padded flood fill -> detection -> assignment fused with the
// (H, W, 3) output (or the bare tid raster when C == 1).
//
// After the flood fill nothing waits for the host: detection and assignment
// are sized from the seed count, and a single read-back at the end checks
// what the sizes and the padding watch assumed.  A capacity overflow or hash
// collision (rare) reruns detection and assignment with exact sizes; an image
// whose padding reached its interior (never seen in practice) gets its
// unpadded fill and its assignment again.
//
// issue_pipeline enqueues everything up to that read-back; finish_pipeline
// waits for it and settles.  When the fill is predicted (a canvas class seen
// before) the first half has no host wait at all, which is what lets
// run_pipeline capture it as a graph.
Issued issue_pipeline(dl_context* ctx, const int32_t* d_xy, int32_t* d_out, int C, hipStream_t s,
                      StageClock& clk)
{
    Issued iss;
    iss.C = C;
    BatchPlan& plan = ctx->plan;
    const int n_img = static_cast<int>(plan.img.size());
    // A fill of a class seen before is issued without waiting (predicted
    // launch count): its result plane is known now, so the descriptors carry
    // their output label source from the first upload, and seed validation
    // comes back with the final status read.  Otherwise the seed flags are
    // read back with the fill's first wait.
    plan_batch(plan);
    bool tune = false;
    const int pred =
        fill_prediction(ctx, choose_config(ctx, plan.n_cells, plan.n_seeds, &tune), tune);
    const size_t cells = static_cast<size_t>(std::max<int64_t>(plan.n_cells, 1));
    ctx->padA.reserve(sizeof(uint32_t) * cells);
    ctx->padB.reserve(sizeof(uint32_t) * cells);
    auto set_out_lab = [&](int plane) {
        const uint32_t* base = pad_plane_ptr(ctx, plane);
        for (ImageDesc& d : plan.img) {
            d.out_lab = base + d.cv_off;
            d.out_W = d.Wc;
            d.out_off = d.P;
            d.out_shift = d.P;
        }
    };
    if (pred > 0) set_out_lab(pred & 1);
    iss.fixed_bins = !ctx->bins_exact.count(tune_key(ctx->device, plan.n_cells, plan.n_seeds));
    prepare(ctx, d_xy, d_out, C, s, pred > 0 ? SeedCheck::kWithStatus : SeedCheck::kAfterWait);
    clk.mark(0);
    std::vector<int> watch(n_img, -1);
    for (int i = 0; i < n_img; ++i)
        if (C == 3 && plan.img[i].P > 0) {
            watch[i] = plan.img[i].P;
            iss.any_pad = true;
        }
    // The fill's visit total arrives with the final status read (a predicted
    // fill is settled after that read).
    int* visits_async = ctx->host_small.as<int>() + 64;
    *visits_async = 0;
    iss.plane = fill_detection(ctx, watch, s, visits_async, &iss.pend);
    if (!iss.pend.active || iss.plane != (pred & 1)) {
        // the fill has waited at least once
        if (!iss.pend.active) check_seed_flags(plan, ctx->tri.host_seed_flags.as<uint32_t>());
        set_out_lab(iss.plane);
        upload_desc(plan, ctx->tri, s);
    }
    const uint32_t* base = pad_plane_ptr(ctx, iss.plane);
    clk.mark(1);
    // (the status words were cleared with the seed flags)
    detect_triangles(plan, base, true, ctx->sorted.as<uint32_t>(), ctx->tri, s, true, true,
                     [&] { clk.mark(2); });
    clk.mark(3);
    // (the bin counts were cleared with the seed bitmap, in prepare)
    assign_tiled(plan, ctx->tri.rows.as<int32_t>(), -1, ctx->sorted.as<uint32_t>(), C, ctx->tri, s,
                 true, iss.fixed_bins, true);
    clk.mark(4);
    issue_status(plan, ctx->tri, iss.any_pad ? ctx->esc.as<int>() : nullptr,
                 iss.any_pad ? n_img : 0, iss.pend.active,
                 iss.pend.active ? ctx->bfs.counters.as<int>() : nullptr,
                 iss.pend.active ? iss.pend.n + 1 : 0, s);
    return iss;
}

void finish_pipeline(dl_context* ctx, const Issued& iss, hipStream_t s, StageClock& clk)
{
    BatchPlan& plan = ctx->plan;
    const int n_img = static_cast<int>(plan.img.size());
    const int C = iss.C;
    const uint32_t* base = pad_plane_ptr(ctx, iss.plane);
    int* visits_async = ctx->host_small.as<int>() + 64;
    DL_HIP(hipStreamSynchronize(s));
    const uint32_t* seed_flags = nullptr;
    const int* fill_changed = nullptr;
    SpeculativeStatus st =
        parse_status(plan, ctx->tri, iss.any_pad ? n_img : 0, iss.pend.active,
                     iss.pend.active ? iss.pend.n + 1 : 0, &seed_flags, &fill_changed);
    const int32_t* rows = ctx->tri.rows.as<int32_t>();
    auto detect_and_assign = [&]() {
        detect_triangles(plan, base, true, ctx->sorted.as<uint32_t>(), ctx->tri, s, true, false,
                         [&] { clk.mark(2); });
        clk.mark(3);
        rows = ctx->tri.rows.as<int32_t>();
        assign_tiled(plan, rows, -1, ctx->sorted.as<uint32_t>(), C, ctx->tri, s, true,
                     iss.fixed_bins);
        clk.mark(4);
        return read_status(plan, ctx->tri, iss.any_pad ? ctx->esc.as<int>() : nullptr,
                           iss.any_pad ? n_img : 0, s);
    };
    if (iss.pend.active) {
        check_seed_flags(plan, seed_flags);
        // (copied: a status read below reuses the buffer)
        const std::vector<int> changed(fill_changed, fill_changed + iss.pend.n + 1);
        if (!finish_predicted_fill(ctx, iss.pend, changed.data(), s, visits_async)) {
            clk.mark(1);
            st = detect_and_assign();
        }
    }
    ctx->timings[8] += *visits_async;
    if (st.bad_id) throw runtime_error("flood fill left an undefined cell");
    if (st.bin_redo)   // this class bins exactly from now on
        ctx->bins_exact.insert(tune_key(ctx->device, plan.n_cells, plan.n_seeds));
    if (st.redo) {
        const DetectResult dr =
            detect_triangles(plan, base, true, ctx->sorted.as<uint32_t>(), ctx->tri, s, false,
                             false, [&] { clk.mark(2); });
        clk.mark(3);
        rows = ctx->tri.rows.as<int32_t>();
        assign_tiled(plan, rows, dr.n_tri, ctx->sorted.as<uint32_t>(), C, ctx->tri, s, false);
        clk.mark(4);
        st.n_tri = dr.n_tri;
    } else if (st.bin_redo) {   // detection stands; bin and assign exactly
        assign_tiled(plan, rows, st.n_tri, ctx->sorted.as<uint32_t>(), C, ctx->tri, s, false);
        clk.mark(4);
    }
    ctx->last_escaped = false;
    for (size_t i = 0; i < st.escaped.size(); ++i) ctx->last_escaped |= st.escaped[i] != 0;
    if (C == 3) {
        // Did a padding cell ever update an image cell?  If not, the image's
        // part of the padded canvas evolved exactly as the unpadded fill.
        std::vector<int> which;
        for (int i = 0; i < n_img; ++i)
            if (plan.img[i].P > 0 &&
                ((!st.escaped.empty() && st.escaped[i]) || ctx->force_unpadded))
                which.push_back(i);
        if (!which.empty()) {
            fill_unpadded(ctx, which, s);
            clk.mark(1);
            upload_desc(plan, ctx->tri, s);
            assign_tiled(plan, rows, st.n_tri, ctx->sorted.as<uint32_t>(), C, ctx->tri, s, false);
            clk.mark(4);
        }
    }
    ctx->n_tri = st.n_tri;
}

// Everything the first half of the pipeline depends on besides device memory
// contents: the images (sizes, seed counts, seed and output addresses), the
// fill's shape and predicted launch count, the test hooks, and the stream.
std::vector<int64_t> pipeline_sig(const dl_context* ctx, const int32_t* d_xy, const int32_t* d_out,
                                  int C, hipStream_t s, const BfsConfig& cfg, int pred)
{
    auto a = [](const void* p) { return static_cast<int64_t>(reinterpret_cast<intptr_t>(p)); };
    std::vector<int64_t> v{a(d_xy),
                           a(d_out),
                           C,
                           a(s),
                           ctx->device,
                           cfg.waves,
                           cfg.rows,
                           cfg.iters,
                           pred,
                           ctx->bfs.list_min_tiles,
                           ctx->tri.spec_tiny,
                           ctx->force_unpadded ? 1 : 0,
                           ctx->bins_exact.count(tune_key(ctx->device, ctx->plan.n_cells,
                                                          ctx->plan.n_seeds))
                               ? 0
                               : 1};
    v.reserve(v.size() + 6 * ctx->plan.img.size());
    for (const ImageDesc& d : ctx->plan.img) {
        v.push_back(d.W);
        v.push_back(d.H);
        v.push_back(d.P);
        v.push_back(d.n);
        v.push_back(a(d.xy));
        v.push_back(a(d.out));
    }
    return v;
}

// Captures issue_pipeline on s into ctx->graph (nothing runs while capturing).
// Returns false, with nothing captured, when the first half would allocate,
// free or wait, or the runtime refuses; the caller then runs uncaptured.
bool capture_pipeline(dl_context* ctx, const int32_t* d_xy, int32_t* d_out, int C, hipStream_t s,
                      StageClock& clk, std::vector<int64_t>&& sig)
{
    PipelineGraph& g = ctx->graph;
    g.reset();
    CaptureLog log;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    capture_log() = &log;
    Issued iss;
    bool ok = true;
    try {
        iss = issue_pipeline(ctx, d_xy, d_out, C, s, clk);
    } catch (...) {
        ok = false;
    }
    capture_log() = nullptr;
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(s, &graph) != hipSuccess) {
        (void)hipGetLastError();
        ok = false;
    }
    ok = ok && graph && iss.pend.active && !log.poisoned;
    hipGraphExec_t exec = nullptr;
    if (ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        exec = nullptr;
        ok = false;
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) return false;
    g.exec = exec;
    g.sig = std::move(sig);
    g.epoch = alloc_epoch();
    g.plan = ctx->plan;
    g.iss = iss;
    g.fill_cvs = ctx->fill_cvs;
    g.last_cfg = ctx->last_cfg;
    g.rows_cap = ctx->tri.rows_cap;
    g.list_cap = ctx->tri.list_cap;
    g.staged = std::move(log.staged);
    g.copy_back = std::move(log.copy_back);
    ++ctx->graph_captures;
    return true;
}

// Launches the captured first half and restores the host state its issue left.
void replay_pipeline(dl_context* ctx, hipStream_t s)
{
    PipelineGraph& g = ctx->graph;
    // The previous call on this context ended with a synchronisation of its
    // stream, so no copy still reads the staging buffers.
    for (const StagedBytes& b : g.staged) {
        std::memcpy(b.host, b.bytes.data(), b.bytes.size());
        DL_HIP(hipMemcpyAsync(b.dev, b.host, b.bytes.size(), hipMemcpyHostToDevice, s));
    }
    ctx->plan = g.plan;
    ctx->fill_cvs = g.fill_cvs;
    ctx->last_cfg = g.last_cfg;
    ctx->tri.rows_cap = g.rows_cap;
    ctx->tri.list_cap = g.list_cap;
    ctx->have_padded = true;
    ctx->pad_plane = g.iss.plane;
    ctx->host_small.as<int>()[64] = 0;   // visit total (see issue_pipeline)
    ctx->timings[7] += g.iss.pend.n;
    ctx->plane_gen = 0;   // the graph's fill writes the planes
    ctx->plane_ptr = nullptr;
    DL_HIP(hipGraphLaunch(g.exec, s));
    for (const CopyBack& c : g.copy_back)
        DL_HIP(hipMemcpyAsync(c.host, c.dev, c.n, hipMemcpyDeviceToHost, s));
    ++ctx->graph_replays;
}

// The pipeline, with its first half replayed as a hipGraph where possible.
// A call whose signature matches the last predicted run's (and no buffer has
// been reallocated since) is captured; later calls with that signature
// replay the capture.  Anything else runs uncaptured.  Timing and kernel
// profiling, which record events between the kernels, run uncaptured too.
// Auto graph mode: pipelines below this many canvas cells are captured (at
// 256^2 and 1024^2 replays took 8 % and 4.5 % less time per call than host
// issue; at 4096^2, 1.1 % more).
constexpr int64_t kGraphAutoMaxCells = int64_t(1) << 22;

void run_pipeline(dl_context* ctx, const int32_t* d_xy, int32_t* d_out, int C, hipStream_t s,
                  StageClock& clk)
{
    plan_batch(ctx->plan);
    const bool graphs = ctx->graph_mode == 1 ||
                        (ctx->graph_mode == 2 && ctx->plan.n_cells < kGraphAutoMaxCells);
    if (!graphs || ctx->timing || ctx->profiling || ctx->rec) {
        const Issued iss = issue_pipeline(ctx, d_xy, d_out, C, s, clk);
        finish_pipeline(ctx, iss, s, clk);
        return;
    }
    bool tune = false;
    const BfsConfig cfg = choose_config(ctx, ctx->plan.n_cells, ctx->plan.n_seeds, &tune);
    const int pred = fill_prediction(ctx, cfg, tune);
    std::vector<int64_t> sig = pipeline_sig(ctx, d_xy, d_out, C, s, cfg, pred);
    const uint64_t epoch = alloc_epoch();
    PipelineGraph& g = ctx->graph;
    if (pred > 0 && g.exec && g.epoch == epoch && g.sig == sig) {
        replay_pipeline(ctx, s);
        finish_pipeline(ctx, g.iss, s, clk);
        return;
    }
    if (pred > 0 && ctx->last_epoch == epoch && ctx->last_sig == sig && ctx->refused_sig != sig) {
        ctx->last_sig.clear();
        const double launches = ctx->timings[7];
        if (capture_pipeline(ctx, d_xy, d_out, C, s, clk, std::vector<int64_t>(sig))) {
            replay_pipeline(ctx, s);
            finish_pipeline(ctx, g.iss, s, clk);
            return;
        }
        // Not capturable: run uncaptured, and do not try this signature again.
        ctx->timings[7] = launches;
        ctx->refused_sig = sig;
    }
    const Issued iss = issue_pipeline(ctx, d_xy, d_out, C, s, clk);
    finish_pipeline(ctx, iss, s, clk);
    if (iss.pend.active && pred > 0) {
        ctx->last_sig = std::move(sig);
        ctx->last_epoch = alloc_epoch();
    } else {
        ctx->last_sig.clear();
    }
}

void upload_seeds(dl_context* ctx, const int32_t* xy, int64_t n)
{
    ctx->xy.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(n));
    DL_HIP(hipMemcpyAsync(ctx->xy.ptr, xy, sizeof(int32_t) * 2 * static_cast<size_t>(n),
                          hipMemcpyHostToDevice, ctx->stream));
}

void plan_single(dl_context* ctx, int W, int H, int P, int64_t n)
{
    ctx->plan = BatchPlan{};
    ImageDesc d{};
    d.W = W;
    d.H = H;
    d.P = P;
    d.n = n;
    ctx->plan.img.push_back(d);
}

// Batch description from the C arrays; validates every image.
int64_t plan_many(dl_context* ctx, int n_images, const int32_t* widths, const int32_t* heights,
                  const int64_t* seed_counts, const int32_t* paddings)
{
    if (n_images <= 0) throw value_error("n_images must be positive");
    if (!widths || !heights || !seed_counts) throw value_error("null argument");
    ctx->plan = BatchPlan{};
    int64_t total = 0;
    for (int i = 0; i < n_images; ++i) {
        ImageDesc d{};
        d.W = widths[i];
        d.H = heights[i];
        d.P = resolve_padding(paddings ? paddings[i] : -1);
        d.n = seed_counts[i];
        try {
            check_dims(d.W, d.H, d.P);
            if (d.n <= 0) throw value_error("seeds must not be empty");
        } catch (const value_error& e) {
            throw value_error("image " + std::to_string(i) + ": " + e.what());
        }
        total += d.n;
        ctx->plan.img.push_back(d);
    }
    return total;
}

struct ProfDrop {
    dl_context* c;
    hipStream_t st;
    ProfDrop(dl_context* ctx, hipStream_t s) : c(ctx), st(s)
    {
        if (c->profiling) set_current_profiler(&c->prof);
    }
    ~ProfDrop()
    {
        if (c->profiling) {
            (void)hipStreamSynchronize(st);
            c->prof.collect();
        }
        set_current_profiler(nullptr);
    }
};


// Keeps the rows and sorted labels of one image of the last pipeline run
// (seeds from seed_off, rows from row_off) as the session's result.
void keep_session_result(dl_context* ctx, dl_session* s, int64_t seed_off, int64_t row_off,
                         hipStream_t st)
{
    s->rows.reserve(sizeof(int32_t) * 5 * static_cast<size_t>(std::max<int64_t>(s->n_tri, 1)));
    if (s->n_tri)
        DL_HIP(hipMemcpyAsync(s->rows.ptr, ctx->tri.rows.as<int32_t>() + 5 * row_off,
                              sizeof(int32_t) * 5 * static_cast<size_t>(s->n_tri),
                              hipMemcpyDeviceToDevice, st));
    s->sorted.reserve(sizeof(uint32_t) * static_cast<size_t>(std::max<int64_t>(s->st.n_live, 1)));
    if (s->st.n_live)
        DL_HIP(hipMemcpyAsync(s->sorted.ptr, ctx->sorted.as<uint32_t>() + seed_off,
                              sizeof(uint32_t) * static_cast<size_t>(s->st.n_live),
                              hipMemcpyDeviceToDevice, st));
    s->n_edges = -1;
}

// Recomputes the session's result (grid, rows, sorted labels) on the context.
void session_finalise(dl_context* ctx, dl_session* s)
{
    hipStream_t st = ctx->stream;
    const size_t px = static_cast<size_t>(s->W) * s->H;
    s->out.reserve(sizeof(int32_t) * 3 * px);
    if (s->st.n_live == 0) {
        StageClock clk(ctx);
        fill_u32(s->out.as<uint32_t>(), 0xFFFFFFFFu, 3 * px, st);
        clk.mark(5);
        clk.finish(st);
        s->n_tri = 0;
        s->n_edges = -1;
    } else {
        ProfInstall pi(ctx);
        StageClock clk(ctx);
        plan_single(ctx, s->W, s->H, s->P, s->st.n_live);
        // A session keeping its fill record hands it to the pipeline's fill
        // for this call; the bitmap then becomes the record's.
        struct RecScope {
            dl_context* c;
            ~RecScope() { c->rec = nullptr; c->rec_bitmap = nullptr; }
        } rs{ctx};
        if (s->record) {
            ctx->rec = &s->rec;
            ctx->rec_bitmap = s->st.bitmap.as<uint32_t>();
            ctx->rec_w = s->W;
            ctx->rec_h = s->H;
        }
        run_pipeline(ctx, s->st.list.as<int32_t>(), s->out.as<int32_t>(), 3, st, clk);
        if (s->record) {
            FillRecord& r = s->rec;
            r.prev_bitmap.reserve(sizeof(uint32_t) * static_cast<size_t>(s->st.words));
            DL_HIP(hipMemcpyAsync(r.prev_bitmap.ptr, s->st.bitmap.ptr,
                                  sizeof(uint32_t) * static_cast<size_t>(s->st.words),
                                  hipMemcpyDeviceToDevice, st));
            // a padding crossing anywhere: the record cannot vouch for the
            // tiles it did not recompute next time
            r.escaped = ctx->last_escaped;
            r.synced = true;
        }
        s->n_tri = ctx->n_tri;
        keep_session_result(ctx, s, 0, 0, st);
        clk.mark(5);
        clk.finish(st);
    }
    s->done = true;
}

SessionView view_of(const dl_session* s)
{
    SessionView v;
    v.W = s->W;
    v.H = s->H;
    v.P = s->P;
    v.n_live = s->st.n_live;
    v.list = s->st.list.as<int32_t>();
    v.values = s->st.values.as<float>();
    v.bitmap = s->st.bitmap.as<uint32_t>();
    v.words = s->st.words;
    v.sorted = s->sorted.as<uint32_t>();
    v.grid = s->out.as<int32_t>();
    v.rows = s->rows.as<int32_t>();
    v.n_tri = s->n_tri;
    return v;
}

void ensure_finalised(dl_context* ctx, dl_session* s)
{
    if (!s->done) session_finalise(ctx, s);
}

void ensure_edges(dl_context* ctx, dl_session* s)
{
    ensure_finalised(ctx, s);
    if (s->n_edges < 0) s->n_edges = build_edges(view_of(s), s->edges, s->qw, ctx->stream);
}

// Every device workspace of a context / session (for the byte counts and
// the trim).
template <class F>
void for_each_devbuf(dl_context* c, F&& f)
{
    BfsWork& b = c->bfs;
    for (DevBuf* d : {&b.flags, &b.counters, &b.visits, &b.list, &b.canvases, &b.tile_map}) f(*d);
    TriWork& t = c->tri;
    for (DevBuf* d : {&t.seed_bm, &t.seed_idx, &t.bitmap, &t.scan, &t.scan_tmp, &t.flags,
                      &t.table_tag, &t.table_tri, &t.cand, &t.counts, &t.offsets, &t.positions,
                      &t.route_ent, &t.route_idx, &t.route_aux,
                      &t.rows, &t.tidgrid, &t.bin_counts, &t.bin_offsets, &t.bin_list, &t.desc,
                      &t.status_dev})
        f(*d);
    for (DevBuf* d : {&c->xy, &c->sorted, &c->padA, &c->padB, &c->unA, &c->unB, &c->out,
                      &c->vgrid, &c->ids, &c->esc})
        f(*d);
}

template <class F>
void for_each_devbuf(dl_session* s, F&& f)
{
    SessionState& st = s->st;
    for (DevBuf* d : {&st.bitmap, &st.list, &st.spare, &st.values, &st.values_spare,
                      &st.backup_list, &st.backup_values, &st.mine, &st.keep, &st.flags})
        f(*d);
    for (DevBuf* d : {&s->stage, &s->stage_values, &s->out, &s->rows, &s->sorted, &s->edges,
                      &s->pix_tids, &s->pix_sids, &s->pix_mask, &s->q_xy, &s->q_t, &s->q_tid,
                      &s->q_mask, &s->q_out, &s->qw.a, &s->qw.b, &s->qw.c, &s->qw.d, &s->qw.e,
                      &s->qw.sort_tmp, &s->qw.scan_tmp})
        f(*d);
    FillRecord& r = s->rec;
    for (DevBuf* d : {&r.pool, &r.csr[0], &r.csr[1], &r.tile_off[0], &r.tile_off[1], &r.app,
                      &r.rank, &r.tstate, &r.ctl, &r.prev_bitmap, &r.list})
        f(*d);
}

void require_values(const dl_session* s, const char* what)
{
    if (!s->st.have_values)
        throw runtime_error(std::string(what) +
                            " needs a value per seed; pass values to insert_deferred");
}

// ---------------------------------------------------------------------------
// The no-allocation entry point (dl_triangulate_device_noalloc).  Its device
// workspace is one static array per GPU, `scratch`; host code never uses the
// array's name as an address (in host code it is only a stand-in for the
// symbol): the device address comes from hipGetSymbolAddress, once per device
// with that device current, and is kept in g_scratch_dev.  A call carves every
// workspace of a call-local context from that address (an Arena), and its
// small status read-backs from a static host block per device, so it calls no
// hipMalloc, hipHostMalloc, hipFree or event / stream creation.  The
// pipeline clears everything it reads before reading it, so the previous
// call's contents do not matter; two calls that overlap on one GPU share the
// array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = DL_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoAllocHostBytes = size_t(1) << 16;
constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kNoAllocDevices] = {};
alignas(256) char g_noalloc_host[kNoAllocDevices][kNoAllocHostBytes];

// An upper bound on the bytes one image's pipeline carves: per padded canvas
// cell the two label planes (8 B), the registration bitmap and its scan
// (1 B) and the detection counts; per pixel the seed index, the bins' fixed
// slices (at most 4 B) and the planes of the unpadded fill run after a
// padding crossing (8 B); per seed the candidate, hash-table and row
// capacities (about 560 B); and the capacities' fixed terms.
size_t noalloc_bound(int W, int H, int P, int64_t n)
{
    const size_t C = static_cast<size_t>(W + 2 * P) * static_cast<size_t>(H + 2 * P);
    const size_t px = static_cast<size_t>(W) * static_cast<size_t>(H);
    return 10 * C + 13 * px + 740 * static_cast<size_t>(n) + (size_t(2) << 20);
}

// The device address of `scratch` on the current device (set up on first use).
char* scratch_on(int device)
{
    if (device < 0 || device >= kNoAllocDevices)
        throw runtime_error("device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) throw_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) throw runtime_error("hipGetSymbolAddress(scratch) returned no address");
        g_scratch_dev[device] = static_cast<char*>(p);
    }
    return g_scratch_dev[device];
}

}  // namespace

extern "C" {

const char* dl_version(void) { return "integer_delauney_amd 0.2.0 (gfx950)"; }

const char* dl_last_error(void) { return g_last_error.c_str(); }

int dl_context_create(int device, dl_context** out)
{
    return guarded([&] {
        if (!out) throw value_error("out is null");
        int count = 0;
        DL_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw value_error("no such HIP device");
        DL_HIP(hipSetDevice(device));
        auto* ctx = new dl_context();
        ctx->device = device;
        const hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            throw_hip(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        }
        ctx->host_small.reserve(4096);
        *out = ctx;
    });
}

void dl_context_destroy(dl_context* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->graph.reset();
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int dl_set_graphs(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (enabled < 0 || enabled > 2) throw value_error("graph mode must be 0, 1 or 2");
        ctx->graph_mode = enabled;
        if (enabled == 0) {
            DL_HIP(hipSetDevice(ctx->device));
            ctx->graph.reset();
            ctx->last_sig.clear();
        }
    });
}

int dl_get_graph_stats(dl_context* ctx, int64_t* captures, int64_t* replays)
{
    return guarded([&] {
        if (!ctx || !captures || !replays) throw value_error("null argument");
        *captures = ctx->graph_captures;
        *replays = ctx->graph_replays;
    });
}

int dl_set_timing(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->timing = enabled != 0;
    });
}

int dl_get_timings(dl_context* ctx, double* out9)
{
    return guarded([&] {
        if (!ctx || !out9) throw value_error("null argument");
        for (int i = 0; i < 9; ++i) out9[i] = ctx->timings[i];
    });
}

int dl_context_device_bytes(dl_context* ctx, size_t* bytes)
{
    return guarded([&] {
        if (!ctx || !bytes) throw value_error("null argument");
        size_t n = 0;
        for_each_devbuf(ctx, [&](DevBuf& d) { n += d.bytes; });
        *bytes = n;
    });
}

int dl_context_trim(dl_context* ctx)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        DL_HIP(hipSetDevice(ctx->device));
        DL_HIP(hipStreamSynchronize(ctx->stream));
        ctx->graph.reset();
        for_each_devbuf(ctx, [](DevBuf& d) { d.release(); });
        ctx->plane_gen = 0;
        ctx->plane_ptr = nullptr;
        ctx->n_tri = 0;
        ctx->have_padded = false;
        ctx->tri.tri_base.clear();
    });
}

int dl_session_device_bytes(dl_session* s, size_t* bytes)
{
    return guarded([&] {
        if (!s || !bytes) throw value_error("null argument");
        size_t n = 0;
        for_each_devbuf(s, [&](DevBuf& d) { n += d.bytes; });
        *bytes = n;
    });
}

int dl_device_memory(int device, size_t* free_bytes, size_t* total_bytes)
{
    return guarded([&] {
        if (!free_bytes || !total_bytes) throw value_error("null argument");
        DL_HIP(hipSetDevice(device));
        DL_HIP(hipMemGetInfo(free_bytes, total_bytes));
    });
}

int dl_get_fill_stats(dl_context* ctx, int64_t* launches, int64_t* converged,
                      int64_t* tile_visits, int* iters_per_launch, int64_t* shortfalls)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (launches) *launches = static_cast<int64_t>(ctx->timings[7]);
        if (converged) *converged = ctx->fill_converged;
        if (tile_visits) *tile_visits = static_cast<int64_t>(ctx->timings[8]);
        if (iters_per_launch) *iters_per_launch = ctx->last_cfg.iters;
        if (shortfalls) *shortfalls = ctx->fill_shortfalls;
    });
}

int dl_set_bfs_config(dl_context* ctx, int waves, int rows, int iters)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (waves == 0 && rows == 0 && iters == 0) {
            ctx->cfg_auto = true;
            return;
        }
        BfsConfig c{waves, rows, iters};
        if (!bfs_config_supported(c)) throw value_error("unsupported flood-fill configuration");
        ctx->cfg = c;
        ctx->cfg_auto = false;
    });
}

int dl_set_autotune(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->autotune = enabled != 0;
    });
}

int dl_set_fill_min_encoding(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->bfs.allow_min_encoding = enabled != 0;
    });
}

int dl_set_fill_aligned_watch(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->bfs.allow_aligned_watch = enabled != 0;
    });
}

int dl_get_fill_aligned_watch(dl_context* ctx, int* aligned)
{
    return guarded([&] {
        if (!ctx || !aligned) throw value_error("null argument");
        *aligned = ctx->bfs.aligned ? 1 : 0;
    });
}

int dl_set_fill_prediction(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->predict = enabled != 0;
    });
}

int dl_autotune_reset(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.clear();
    });
}

int dl_set_kernel_profiling(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->profiling = enabled != 0;
    });
}

int dl_get_kernel_stats(dl_context* ctx, int64_t* counts, double* ms)
{
    return guarded([&] {
        if (!ctx || !counts || !ms) throw value_error("null argument");
        for (int i = 0; i < kKinds; ++i) {
            counts[i] = ctx->prof.count[i];
            ms[i] = ctx->prof.ms[i];
        }
    });
}

int dl_reset_kernel_stats(dl_context* ctx)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->prof.reset();
    });
}

int dl_get_bfs_config(dl_context* ctx, int* waves, int* rows, int* iters)
{
    return guarded([&] {
        if (!ctx || !waves || !rows || !iters) throw value_error("null argument");
        *waves = ctx->last_cfg.waves;
        *rows = ctx->last_cfg.rows;
        *iters = ctx->last_cfg.iters;
    });
}

int dl_set_force_unpadded(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->force_unpadded = enabled != 0;
    });
}

int dl_set_replay_plane_reuse(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->force_full_replay = enabled == 0;
    });
}

int dl_get_record_stats(dl_context* ctx, int64_t* recorded, int64_t* replayed, int64_t* retained)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (recorded) *recorded = ctx->record_fills;
        if (replayed) *replayed = ctx->replay_fills;
        if (retained) *retained = ctx->retained_fills;
    });
}

int dl_set_fill_list_threshold(dl_context* ctx, int min_tiles)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        ctx->bfs.list_min_tiles = min_tiles < 0 ? 8192 : min_tiles;
    });
}

int dl_set_tiny_capacities(dl_context* ctx, int enabled)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (enabled < 0 || enabled > 2) throw value_error("tiny-capacity mode must be 0, 1 or 2");
        ctx->tri.spec_tiny = enabled;
    });
}

int dl_voronoi(dl_context* ctx, int width, int height, const int32_t* seeds_xy, int64_t n,
               int32_t* out_hw2)
{
    return guarded([&] {
        if (!ctx || !out_hw2) throw value_error("null argument");
        check_dims(width, height, 0);
        check_seeds(seeds_xy, n);
        DL_HIP(hipSetDevice(ctx->device));
        ProfInstall pi(ctx);
        hipStream_t s = ctx->stream;
        StageClock clk(ctx);
        upload_seeds(ctx, seeds_xy, n);
        plan_single(ctx, width, height, 0, n);
        prepare(ctx, ctx->xy.as<int32_t>(), nullptr, 0, s);
        clk.mark(0);
        const int plane = fill_detection(ctx, std::vector<int>{-1}, s);
        clk.mark(1);
        const size_t px = static_cast<size_t>(width) * height;
        ctx->out.reserve(sizeof(int32_t) * 2 * px);
        write_output(width, height, pad_plane_ptr(ctx, plane), width, 0, 0,
                     seed_index(ctx->tri), 0, 0, height, nullptr, ctx->out.as<int32_t>(), 2, s);
        DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * px,
                              hipMemcpyDeviceToHost, s));
        ctx->have_padded = false;
        clk.finish(s);
    });
}

int dl_debug_fill_watch(dl_context* ctx, int width, int height, const int32_t* seeds_xy,
                        int64_t n, int margin, int32_t* out_hw2, int* escaped)
{
    return guarded([&] {
        if (!ctx || !escaped) throw value_error("null argument");
        check_dims(width, height, 0);
        check_seeds(seeds_xy, n);
        if (margin < 0 || 2 * margin >= width || 2 * margin >= height)
            throw value_error("watch margin must leave a non-empty box");
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        upload_seeds(ctx, seeds_xy, n);
        plan_single(ctx, width, height, 0, n);
        prepare(ctx, ctx->xy.as<int32_t>(), nullptr, 0, s);
        const int plane = fill_detection(ctx, std::vector<int>{margin}, s);
        DL_HIP(hipMemcpyAsync(ctx->host_small.ptr, ctx->esc.ptr, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        if (out_hw2) {
            const size_t px = static_cast<size_t>(width) * height;
            ctx->out.reserve(sizeof(int32_t) * 2 * px);
            write_output(width, height, pad_plane_ptr(ctx, plane), width, 0, 0,
                         seed_index(ctx->tri), 0, 0, height, nullptr, ctx->out.as<int32_t>(), 2, s);
            DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * px,
                                  hipMemcpyDeviceToHost, s));
        }
        DL_HIP(hipStreamSynchronize(s));
        *escaped = ctx->host_small.as<int>()[0] != 0;
        ctx->have_padded = false;
    });
}

int dl_voronoi_simple(dl_context* ctx, int width, int height, const int32_t* seeds_xy,
                      int64_t n, int32_t* out_hw2, int64_t* iters)
{
    return guarded([&] {
        if (!ctx || !out_hw2) throw value_error("null argument");
        check_dims(width, height, 0);
        check_seeds(seeds_xy, n);
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        upload_seeds(ctx, seeds_xy, n);
        plan_single(ctx, width, height, 0, n);
        prepare(ctx, ctx->xy.as<int32_t>(), nullptr, 0, s);
        const size_t cells = static_cast<size_t>(width) * height;
        ctx->unA.reserve(sizeof(uint32_t) * cells);
        ctx->unB.reserve(sizeof(uint32_t) * cells);
        BfsCanvas cv{};
        cv.lab0 = ctx->unA.as<uint32_t>();
        cv.lab1 = ctx->unB.as<uint32_t>();
        cv.W = width;
        cv.H = height;
        bfs_init_canvas(cv, ctx->sorted.as<uint32_t>(), n, s);
        const int plane = bfs_run_simple(cv, ctx->bfs, s, iters);
        ctx->out.reserve(sizeof(int32_t) * 2 * cells);
        write_output(width, height, plane ? cv.lab1 : cv.lab0, width, 0, 0,
                     seed_index(ctx->tri), 0, 0, height, nullptr, ctx->out.as<int32_t>(), 2, s);
        DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * cells,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        ctx->have_padded = false;
    });
}

int dl_triangulate(dl_context* ctx, int width, int height, const int32_t* seeds_xy, int64_t n,
                   int border_padding, int32_t* out_hw3, int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !out_hw3) throw value_error("null argument");
        const int P = resolve_padding(border_padding);
        check_dims(width, height, P);
        check_seeds(seeds_xy, n);
        DL_HIP(hipSetDevice(ctx->device));
        ProfInstall pi(ctx);
        hipStream_t s = ctx->stream;
        StageClock clk(ctx);
        upload_seeds(ctx, seeds_xy, n);
        plan_single(ctx, width, height, P, n);
        const size_t px = static_cast<size_t>(width) * height;
        ctx->out.reserve(sizeof(int32_t) * 3 * px);
        run_pipeline(ctx, ctx->xy.as<int32_t>(), ctx->out.as<int32_t>(), 3, s, clk);
        DL_HIP(hipMemcpyAsync(out_hw3, ctx->out.ptr, sizeof(int32_t) * 3 * px,
                              hipMemcpyDeviceToHost, s));
        clk.mark(5);
        clk.finish(s);
        if (n_tri) *n_tri = ctx->n_tri;
    });
}

int dl_triangulate_device(dl_context* ctx, int width, int height, const int32_t* d_seeds_xy,
                          int64_t n, int border_padding, int32_t* d_out_hw3, void* stream,
                          int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !d_out_hw3) throw value_error("null argument");
        const int P = resolve_padding(border_padding);
        check_dims(width, height, P);
        check_seeds(d_seeds_xy, n);
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        ProfDrop drop(ctx, s);
        StageClock clk(ctx);
        plan_single(ctx, width, height, P, n);
        run_pipeline(ctx, d_seeds_xy, d_out_hw3, 3, s, clk);
        clk.finish(s);
        if (n_tri) *n_tri = ctx->n_tri;
    });
}

int dl_triangulate_device_noalloc(dl_context* ctx, int width, int height,
                                  const int32_t* d_seeds_xy, int64_t n, int border_padding,
                                  int32_t* d_out_hw3, int32_t* d_rows, int64_t rows_cap,
                                  void* stream, int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !d_out_hw3) throw value_error("null argument");
        const int P = resolve_padding(border_padding);
        check_dims(width, height, P);
        check_seeds(d_seeds_xy, n);
        if (rows_cap < 0 || (rows_cap > 0 && !d_rows)) throw value_error("bad rows buffer");
        const size_t need = noalloc_bound(width, height, P, n);
        if (need > kNoAllocBytes)
            throw value_error("a " + std::to_string(width) + "x" + std::to_string(height) +
                              " image with " + std::to_string(n) + " seeds at padding " +
                              std::to_string(P) + " may need " + std::to_string(need) +
                              " bytes of scratch; the no-allocation entry point has " +
                              std::to_string(kNoAllocBytes) +
                              " (10 per canvas cell + 13 per pixel + 740 per seed + 2 MiB)");
        DL_HIP(hipSetDevice(ctx->device));
        char* dev = scratch_on(ctx->device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        Arena arena;
        arena.dev = dev;
        arena.dev_cap = kNoAllocBytes;
        arena.host = g_noalloc_host[ctx->device];
        arena.host_cap = kNoAllocHostBytes;
        struct Install {
            explicit Install(Arena* a) { current_arena() = a; }
            ~Install() { current_arena() = nullptr; }
        } install(&arena);
        // A call-local context: every workspace it reserves is carved from
        // the arena; no graphs, prediction, tuning or timing (they keep state
        // or create events).
        dl_context nc;
        nc.device = ctx->device;
        nc.stream = s;
        nc.graph_mode = 0;
        nc.predict = false;
        nc.autotune = false;
        nc.bfs.use_events = false;
        try {
            nc.host_small.reserve(4096);
            StageClock clk(&nc);
            plan_single(&nc, width, height, P, n);
            run_pipeline(&nc, d_seeds_xy, d_out_hw3, 3, s, clk);
            if (n_tri) *n_tri = nc.n_tri;
            if (d_rows && nc.n_tri > rows_cap)
                throw value_error("rows_cap " + std::to_string(rows_cap) + " is below the " +
                                  std::to_string(nc.n_tri) + " triangles");
            if (d_rows && nc.n_tri)
                DL_HIP(hipMemcpyAsync(d_rows, nc.tri.rows.ptr,
                                      sizeof(int32_t) * 5 * static_cast<size_t>(nc.n_tri),
                                      hipMemcpyDeviceToDevice, s));
            DL_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);   // nothing may still use the scratch
            throw;
        }
    });
}

int dl_triangulate_batch(dl_context* ctx, int n_images, const int32_t* widths,
                         const int32_t* heights, const int64_t* seed_counts,
                         const int32_t* seeds_xy, const int32_t* paddings, int32_t* out,
                         int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !out || !seeds_xy) throw value_error("null argument");
        const int64_t total = plan_many(ctx, n_images, widths, heights, seed_counts, paddings);
        DL_HIP(hipSetDevice(ctx->device));
        ProfInstall pi(ctx);
        hipStream_t s = ctx->stream;
        StageClock clk(ctx);
        upload_seeds(ctx, seeds_xy, total);
        int64_t px = 0;
        for (const auto& d : ctx->plan.img) px += static_cast<int64_t>(d.W) * d.H;
        ctx->out.reserve(sizeof(int32_t) * 3 * static_cast<size_t>(px));
        run_pipeline(ctx, ctx->xy.as<int32_t>(), ctx->out.as<int32_t>(), 3, s, clk);
        DL_HIP(hipMemcpyAsync(out, ctx->out.ptr, sizeof(int32_t) * 3 * static_cast<size_t>(px),
                              hipMemcpyDeviceToHost, s));
        clk.mark(5);
        clk.finish(s);
        if (n_tri)
            for (int i = 0; i < n_images; ++i) n_tri[i] = ctx->tri.tri_base[i + 1] - ctx->tri.tri_base[i];
    });
}

int dl_triangulate_batch_device(dl_context* ctx, int n_images, const int32_t* widths,
                                const int32_t* heights, const int64_t* seed_counts,
                                const int32_t* d_seeds_xy, const int32_t* paddings,
                                int32_t* d_out, void* stream, int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !d_out || !d_seeds_xy) throw value_error("null argument");
        plan_many(ctx, n_images, widths, heights, seed_counts, paddings);
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        ProfDrop drop(ctx, s);
        StageClock clk(ctx);
        run_pipeline(ctx, d_seeds_xy, d_out, 3, s, clk);
        clk.finish(s);
        if (n_tri)
            for (int i = 0; i < n_images; ++i) n_tri[i] = ctx->tri.tri_base[i + 1] - ctx->tri.tri_base[i];
    });
}

const int32_t* dl_device_triangles(dl_context* ctx)
{
    return ctx ? ctx->tri.rows.as<int32_t>() : nullptr;
}

int dl_grid_triangulation(dl_context* ctx, const int32_t* voronoi_grid_hw2, int width,
                          int height, const int32_t* seeds_xy, int64_t n, int border_padding,
                          int32_t* out_tid_hw, int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !out_tid_hw) throw value_error("null argument");
        const int P = resolve_padding(border_padding);
        check_dims(width, height, P);
        check_seeds(seeds_xy, n);
        DL_HIP(hipSetDevice(ctx->device));
        ProfInstall pi(ctx);
        hipStream_t s = ctx->stream;
        StageClock clk(ctx);
        upload_seeds(ctx, seeds_xy, n);
        plan_single(ctx, width, height, P, n);
        const size_t px = static_cast<size_t>(width) * height;
        ctx->tri.tidgrid.reserve(sizeof(int32_t) * px);
        if (P > 0) {
            run_pipeline(ctx, ctx->xy.as<int32_t>(), ctx->tri.tidgrid.as<int32_t>(), 1, s, clk);
        } else {
            if (!voronoi_grid_hw2) throw value_error("voronoi_grid is required at padding 0");
            prepare(ctx, ctx->xy.as<int32_t>(), nullptr, 0, s);
            clk.mark(0);
            ctx->have_padded = false;
            ctx->vgrid.reserve(sizeof(int32_t) * 2 * px);
            ctx->ids.reserve(sizeof(uint32_t) * px);
            DL_HIP(hipMemcpyAsync(ctx->vgrid.ptr, voronoi_grid_hw2, sizeof(int32_t) * 2 * px,
                                  hipMemcpyHostToDevice, s));
            extract_ids_kernel<<<static_cast<unsigned>((px + 255) / 256), 256, 0, s>>>(
                ctx->vgrid.as<int32_t>(), ctx->ids.as<uint32_t>(), static_cast<int64_t>(px));
            DL_LAUNCH_CHECK();
            clk.mark(1);
            const DetectResult dr = detect_triangles(ctx->plan, ctx->ids.as<uint32_t>(), false,
                                                     ctx->sorted.as<uint32_t>(), ctx->tri, s, false,
                                                     false, [&] { clk.mark(2); });
            clk.mark(3);
            ctx->n_tri = dr.n_tri;
            ctx->plan.img[0].out = ctx->tri.tidgrid.as<int32_t>();
            upload_desc(ctx->plan, ctx->tri, s);
            assign_tiled(ctx->plan, ctx->tri.rows.as<int32_t>(), dr.n_tri,
                         ctx->sorted.as<uint32_t>(), 1, ctx->tri, s, false);
            clk.mark(4);
        }
        DL_HIP(hipMemcpyAsync(out_tid_hw, ctx->tri.tidgrid.ptr, sizeof(int32_t) * px,
                              hipMemcpyDeviceToHost, s));
        clk.mark(5);
        clk.finish(s);
        if (n_tri) *n_tri = ctx->n_tri;
    });
}

int dl_copy_triangles(dl_context* ctx, int32_t* out_rows)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (ctx->n_tri == 0) return;
        if (!out_rows) throw value_error("null output");
        DL_HIP(hipSetDevice(ctx->device));
        DL_HIP(hipMemcpyAsync(out_rows, ctx->tri.rows.ptr,
                              sizeof(int32_t) * 5 * static_cast<size_t>(ctx->n_tri),
                              hipMemcpyDeviceToHost, ctx->stream));
        DL_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int dl_copy_padded_voronoi(dl_context* ctx, int32_t* out, int* out_w, int* out_h)
{
    return guarded([&] {
        if (!ctx) throw value_error("null context");
        if (ctx->plan.img.size() != 1) throw value_error("the last call was not a single image");
        const ImageDesc& d = ctx->plan.img[0];
        const int P = d.P;
        const int Wc = d.W + 2 * P, Hc = d.H + 2 * P;
        if (out_w) *out_w = Wc;
        if (out_h) *out_h = Hc;
        if (!out) return;
        if (!ctx->have_padded || P == 0)
            throw value_error("no padded canvas was computed by the last call");
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const size_t cells = static_cast<size_t>(Wc) * Hc;
        ctx->out.reserve(sizeof(int32_t) * 2 * cells);
        write_output(Wc, Hc, pad_plane_ptr(ctx, ctx->pad_plane), Wc, 0, P,
                     seed_index(ctx->tri), d.bm_off, d.seed_off, d.H, nullptr, ctx->out.as<int32_t>(),
                     2, s);
        DL_HIP(hipMemcpyAsync(out, ctx->out.ptr, sizeof(int32_t) * 2 * cells,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
    });
}

int dl_host_alloc(size_t bytes, void** out)
{
    return guarded([&] {
        if (!out) throw value_error("null argument");
        *out = nullptr;
        DL_HIP(hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    });
}

void dl_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

int dl_run_batch(dl_context* ctx, int n, dl_image* images, void* stream)
{
    return guarded([&] {
        if (!ctx || !images || n <= 0) throw value_error("null argument");
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        ProfDrop drop(ctx, s);
        ctx->plan = BatchPlan{};
        std::vector<int> idx;   // images with seeds, in plan order
        // Host images are staged through ctx->xy (seeds) and ctx->out (grids).
        int64_t host_seeds = 0, host_px = 0;
        for (int i = 0; i < n; ++i)
            if (!images[i].session && images[i].host && images[i].n_seeds > 0 &&
                images[i].width > 0 && images[i].height > 0) {
                host_seeds += images[i].n_seeds;
                host_px += static_cast<int64_t>(images[i].width) * images[i].height;
            }
        ctx->xy.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(std::max<int64_t>(host_seeds, 1)));
        ctx->out.reserve(sizeof(int32_t) * 3 * static_cast<size_t>(std::max<int64_t>(host_px, 1)));
        int64_t so = 0, po = 0;
        std::vector<int32_t*> dev_out(n, nullptr);
        for (int i = 0; i < n; ++i) {
            dl_image& im = images[i];
            im.n_tri = 0;
            im.tri_offset = 0;
            ImageDesc d{};
            if (im.session) {
                dl_session* ss = im.session;
                if (ss->device != ctx->device) throw value_error("session lives on another device");
                d.W = ss->W;
                d.H = ss->H;
                d.P = ss->P;
                d.n = ss->st.n_live;
                d.xy = ss->st.list.as<int32_t>();
                ss->out.reserve(sizeof(int32_t) * 3 * static_cast<size_t>(d.W) * d.H);
                d.out = ss->out.as<int32_t>();
                if (d.n == 0) {   // nothing live: all -1
                    fill_u32(ss->out.as<uint32_t>(), 0xFFFFFFFFu, 3 * static_cast<size_t>(d.W) * d.H, s);
                    ss->n_tri = 0;
                    ss->done = true;
                    continue;
                }
            } else {
                d.W = im.width;
                d.H = im.height;
                d.P = resolve_padding(im.border_padding);
                d.n = im.n_seeds;
                try {
                    check_dims(d.W, d.H, d.P);
                    if (d.n <= 0) throw value_error("seeds must not be empty");
                    if (!im.seeds_xy || !im.out) throw value_error("null seeds or output");
                } catch (const value_error& e) {
                    throw value_error("image " + std::to_string(i) + ": " + e.what());
                }
                if (im.host) {
                    int32_t* sx = ctx->xy.as<int32_t>() + 2 * so;
                    DL_HIP(hipMemcpyAsync(sx, im.seeds_xy, sizeof(int32_t) * 2 * static_cast<size_t>(d.n),
                                          hipMemcpyHostToDevice, s));
                    d.xy = sx;
                    d.out = ctx->out.as<int32_t>() + 3 * po;
                    dev_out[i] = d.out;
                    so += d.n;
                    po += static_cast<int64_t>(d.W) * d.H;
                } else {
                    d.xy = im.seeds_xy;
                    d.out = im.out;
                }
            }
            ctx->plan.img.push_back(d);
            idx.push_back(i);
        }
        if (idx.empty()) {
            DL_HIP(hipStreamSynchronize(s));
            ctx->n_tri = 0;
            return;
        }
        StageClock clk(ctx);
        run_pipeline(ctx, nullptr, nullptr, 3, s, clk);
        const std::vector<int64_t>& tb = ctx->tri.tri_base;
        for (size_t k = 0; k < idx.size(); ++k) {
            dl_image& im = images[idx[k]];
            im.tri_offset = tb[k];
            im.n_tri = tb[k + 1] - tb[k];
            if (im.session) {
                dl_session* ss = im.session;
                ss->n_tri = im.n_tri;
                keep_session_result(ctx, ss, ctx->plan.img[k].seed_off, tb[k], s);
                ss->done = true;
            } else if (im.host) {
                DL_HIP(hipMemcpyAsync(im.out, dev_out[idx[k]],
                                      sizeof(int32_t) * 3 * static_cast<size_t>(im.width) * im.height,
                                      hipMemcpyDeviceToHost, s));
            }
        }
        clk.mark(5);
        clk.finish(s);
    });
}

int dl_session_create(dl_context* ctx, int width, int height, int64_t max_seeds,
                      int border_padding, dl_session** out)
{
    return guarded([&] {
        if (!ctx || !out) throw value_error("null argument");
        if (max_seeds <= 0) throw value_error("max_seeds must be positive");
        // any negative padding selects the default (delaunay.cu:72)
        const int P = border_padding < 0 ? DL_DEFAULT_BORDER_PADDING : border_padding;
        check_dims(width, height, P);
        DL_HIP(hipSetDevice(ctx->device));
        auto* s = new dl_session();
        try {
            s->device = ctx->device;
            s->W = width;
            s->H = height;
            s->P = P;
            session_alloc(s->st, width, height, max_seeds, ctx->stream);
            DL_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

void dl_session_destroy(dl_session* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    delete s;
}

int dl_session_insert(dl_context* ctx, dl_session* s, const int32_t* seeds_xy, const float* values,
                      int64_t n)
{
    return guarded([&] {
        if (!ctx || !s) throw value_error("null argument");
        if (n <= 0) return;
        if (!seeds_xy) throw value_error("null seeds");
        DL_HIP(hipSetDevice(ctx->device));
        if (s->st.n_live + n > s->st.max_seeds)
            throw value_error("insert would exceed max_seeds capacity");
        s->stage.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(n));
        DL_HIP(hipMemcpyAsync(s->stage.ptr, seeds_xy, sizeof(int32_t) * 2 * static_cast<size_t>(n),
                              hipMemcpyHostToDevice, ctx->stream));
        const float* dv = nullptr;
        if (values) {
            s->stage_values.reserve(sizeof(float) * static_cast<size_t>(n));
            DL_HIP(hipMemcpyAsync(s->stage_values.ptr, values, sizeof(float) * static_cast<size_t>(n),
                                  hipMemcpyHostToDevice, ctx->stream));
            dv = s->stage_values.as<float>();
        }
        session_insert(s->st, s->stage.as<int32_t>(), dv, n, ctx->stream);
        DL_HIP(hipStreamSynchronize(ctx->stream));
        s->done = false;
        s->n_edges = -1;
    });
}

int dl_session_remove(dl_context* ctx, dl_session* s, const int32_t* seeds_xy, int64_t n)
{
    return guarded([&] {
        if (!ctx || !s) throw value_error("null argument");
        if (n <= 0) return;
        if (!seeds_xy) throw value_error("null seeds");
        DL_HIP(hipSetDevice(ctx->device));
        s->stage.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(n));
        DL_HIP(hipMemcpyAsync(s->stage.ptr, seeds_xy, sizeof(int32_t) * 2 * static_cast<size_t>(n),
                              hipMemcpyHostToDevice, ctx->stream));
        session_remove(s->st, s->stage.as<int32_t>(), n, ctx->tri.scan_tmp, ctx->stream);
        DL_HIP(hipStreamSynchronize(ctx->stream));
        s->done = false;
        s->n_edges = -1;
    });
}

int dl_session_update(dl_context* ctx, dl_session* s, const int32_t* remove_xy, int64_t n_remove,
                      const int32_t* insert_xy, const float* values, int64_t n_insert)
{
    return guarded([&] {
        if (!ctx || !s) throw value_error("null argument");
        if (n_remove < 0 || n_insert < 0) throw value_error("negative count");
        if ((n_remove && !remove_xy) || (n_insert && !insert_xy)) throw value_error("null seeds");
        if (!n_remove && !n_insert) return;
        DL_HIP(hipSetDevice(ctx->device));
        if (s->st.n_live - n_remove + n_insert > s->st.max_seeds)
            throw value_error("insert would exceed max_seeds capacity");
        const size_t nr = static_cast<size_t>(n_remove), ni = static_cast<size_t>(n_insert);
        s->stage.reserve(sizeof(int32_t) * 2 * (nr + ni));
        int32_t* d_rm = s->stage.as<int32_t>();
        int32_t* d_ins = d_rm + 2 * nr;
        if (nr)
            DL_HIP(hipMemcpyAsync(d_rm, remove_xy, sizeof(int32_t) * 2 * nr, hipMemcpyHostToDevice,
                                  ctx->stream));
        if (ni)
            DL_HIP(hipMemcpyAsync(d_ins, insert_xy, sizeof(int32_t) * 2 * ni, hipMemcpyHostToDevice,
                                  ctx->stream));
        const float* dv = nullptr;
        if (values && ni) {
            s->stage_values.reserve(sizeof(float) * ni);
            DL_HIP(hipMemcpyAsync(s->stage_values.ptr, values, sizeof(float) * ni,
                                  hipMemcpyHostToDevice, ctx->stream));
            dv = s->stage_values.as<float>();
        }
        session_update(s->st, d_rm, n_remove, d_ins, dv, n_insert, ctx->tri.scan_tmp, ctx->stream);
        DL_HIP(hipStreamSynchronize(ctx->stream));
        s->done = false;
        s->n_edges = -1;
    });
}

int dl_session_seed_count(dl_session* s, int64_t* n)
{
    return guarded([&] {
        if (!s || !n) throw value_error("null argument");
        *n = s->st.n_live;
    });
}

int dl_session_set_fill_record(dl_session* s, int enabled)
{
    return guarded([&] {
        if (!s) throw value_error("null session");
        DL_HIP(hipSetDevice(s->device));
        s->record = enabled != 0;
        if (!s->record) {
            FillRecord& r = s->rec;
            for (DevBuf* d : {&r.pool, &r.csr[0], &r.csr[1], &r.tile_off[0], &r.tile_off[1],
                              &r.app, &r.rank, &r.tstate, &r.ctl, &r.prev_bitmap, &r.list,
                              &r.freelist, &r.mark})
                d->release();
            r.valid = r.synced = r.escaped = false;
            r.pool_cap = r.ent_cap = r.entries = r.last_computed = 0;
            r.states = r.cur = 0;
            r.hw = 1;
            r.free_n = 0;
        }
    });
}

int dl_session_fill_record_stats(dl_session* s, int* valid, int64_t* entries,
                                 int64_t* tiles_computed)
{
    return guarded([&] {
        if (!s) throw value_error("null session");
        const FillRecord& r = s->rec;
        if (valid) *valid = (r.valid && r.synced && !r.escaped) ? 1 : 0;
        if (entries) *entries = r.entries;
        if (tiles_computed) *tiles_computed = r.last_computed;
    });
}

int dl_session_finalise(dl_context* ctx, dl_session* s, int64_t* n_tri)
{
    return guarded([&] {
        if (!ctx || !s) throw value_error("null argument");
        DL_HIP(hipSetDevice(ctx->device));
        session_finalise(ctx, s);
        if (n_tri) *n_tri = s->n_tri;
    });
}

int dl_session_copy_result(dl_session* s, int32_t* out_hw3, int32_t* out_rows)
{
    return guarded([&] {
        if (!s) throw value_error("null session");
        if (!s->done) throw value_error("the session has pending changes: finalise first");
        DL_HIP(hipSetDevice(s->device));
        if (out_hw3)
            DL_HIP(hipMemcpy(out_hw3, s->out.ptr, sizeof(int32_t) * 3 * static_cast<size_t>(s->W) * s->H,
                             hipMemcpyDeviceToHost));
        if (out_rows && s->n_tri)
            DL_HIP(hipMemcpy(out_rows, s->rows.ptr, sizeof(int32_t) * 5 * static_cast<size_t>(s->n_tri),
                             hipMemcpyDeviceToHost));
    });
}

int dl_session_copy_seeds(dl_session* s, int32_t* out_xy)
{
    return guarded([&] {
        if (!s || !out_xy) throw value_error("null argument");
        if (!s->st.n_live) return;
        DL_HIP(hipSetDevice(s->device));
        DL_HIP(hipMemcpy(out_xy, s->st.list.ptr, sizeof(int32_t) * 2 * static_cast<size_t>(s->st.n_live),
                         hipMemcpyDeviceToHost));
    });
}

const int32_t* dl_session_device_grid(dl_session* s)
{
    return s && s->done ? s->out.as<int32_t>() : nullptr;
}

int dl_device_to_host(void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        if (!bytes) return;
        if (!dst || !src) throw value_error("null argument");
        DL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int dl_session_has_values(dl_session* s, int* out)
{
    return guarded([&] {
        if (!s || !out) throw value_error("null argument");
        *out = s->st.have_values ? 1 : 0;
    });
}

int dl_session_copy_values(dl_session* s, float* out)
{
    return guarded([&] {
        if (!s || !out) throw value_error("null argument");
        if (!s->st.n_live) return;
        DL_HIP(hipSetDevice(s->device));
        DL_HIP(hipMemcpy(out, s->st.values.ptr, sizeof(float) * static_cast<size_t>(s->st.n_live),
                         hipMemcpyDeviceToHost));
    });
}

int dl_session_locate(dl_context* ctx, dl_session* s, const int32_t* points_xy, int64_t n,
                      int32_t* out_tid)
{
    return dl_session_in_circumsphere(ctx, s, points_xy, nullptr, n, nullptr, out_tid);
}

int dl_session_in_circumsphere(dl_context* ctx, dl_session* s, const int32_t* points_xy,
                               const double* t, int64_t n, uint8_t* out_mask, int32_t* out_tid)
{
    return guarded([&] {
        if (!ctx || !s) throw value_error("null argument");
        if (n < 0) throw value_error("negative point count");
        if (n == 0) return;
        if (!points_xy || !out_tid || (t && !out_mask)) throw value_error("null points or output");
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        ensure_finalised(ctx, s);
        const size_t nn = static_cast<size_t>(n);
        s->q_xy.reserve(sizeof(int32_t) * 2 * nn);
        s->q_tid.reserve(sizeof(int32_t) * nn);
        DL_HIP(hipMemcpyAsync(s->q_xy.ptr, points_xy, sizeof(int32_t) * 2 * nn, hipMemcpyHostToDevice,
                              st));
        const double* dt = nullptr;
        uint8_t* dm = nullptr;
        if (t) {
            s->q_t.reserve(sizeof(double) * nn);
            s->q_mask.reserve(nn);
            DL_HIP(hipMemcpyAsync(s->q_t.ptr, t, sizeof(double) * nn, hipMemcpyHostToDevice, st));
            dt = s->q_t.as<double>();
            dm = s->q_mask.as<uint8_t>();
        }
        query_points(view_of(s), s->q_xy.as<int32_t>(), dt, n, s->q_tid.as<int32_t>(), dm, s->qw, st);
        DL_HIP(hipMemcpyAsync(out_tid, s->q_tid.ptr, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
        if (t) DL_HIP(hipMemcpyAsync(out_mask, dm, nn, hipMemcpyDeviceToHost, st));
        DL_HIP(hipStreamSynchronize(st));
    });
}

int dl_session_edges(dl_context* ctx, dl_session* s, int64_t* n_edges)
{
    return guarded([&] {
        if (!ctx || !s || !n_edges) throw value_error("null argument");
        DL_HIP(hipSetDevice(ctx->device));
        ensure_edges(ctx, s);
        *n_edges = s->n_edges;
    });
}

int dl_session_copy_edges(dl_session* s, int32_t* out)
{
    return guarded([&] {
        if (!s || !out) throw value_error("null argument");
        if (s->n_edges < 0) throw value_error("the edge list is not built: call dl_session_edges");
        if (!s->n_edges) return;
        DL_HIP(hipSetDevice(s->device));
        DL_HIP(hipMemcpy(out, s->edges.ptr, sizeof(int32_t) * 2 * static_cast<size_t>(s->n_edges),
                         hipMemcpyDeviceToHost));
    });
}

int dl_session_edge_scores(dl_context* ctx, dl_session* s, double min_length, float* out)
{
    return guarded([&] {
        if (!ctx || !s || !out) throw value_error("null argument");
        DL_HIP(hipSetDevice(ctx->device));
        ensure_edges(ctx, s);
        if (!s->n_edges) return;
        require_values(s, "edge_scores");
        const size_t ne = static_cast<size_t>(s->n_edges);
        s->q_out.reserve(sizeof(float) * ne);
        edge_scores(view_of(s), s->edges.as<int32_t>(), s->n_edges, min_length, s->q_out.as<float>(),
                    ctx->stream);
        DL_HIP(hipMemcpyAsync(out, s->q_out.ptr, sizeof(float) * ne, hipMemcpyDeviceToHost,
                              ctx->stream));
        DL_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int dl_session_select_midpoints(dl_context* ctx, dl_session* s, double min_length, int64_t count,
                                float threshold, int32_t* out_xy, int64_t out_cap, int64_t* n_out)
{
    return guarded([&] {
        if (!ctx || !s || !n_out) throw value_error("null argument");
        *n_out = 0;
        DL_HIP(hipSetDevice(ctx->device));
        ensure_edges(ctx, s);
        if (!s->n_edges || count <= 0) return;
        require_values(s, "select_midpoints");
        const int64_t cap = std::min(count, s->n_edges);
        if (!out_xy || out_cap < cap)
            throw value_error("output capacity below min(count, edge count)");
        s->q_out.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(cap));
        const int64_t m = select_midpoints(view_of(s), s->edges.as<int32_t>(), s->n_edges, min_length,
                                           count, threshold, s->q_out.as<int32_t>(), s->qw,
                                           ctx->stream);
        if (m)
            DL_HIP(hipMemcpyAsync(out_xy, s->q_out.ptr, sizeof(int32_t) * 2 * static_cast<size_t>(m),
                                  hipMemcpyDeviceToHost, ctx->stream));
        DL_HIP(hipStreamSynchronize(ctx->stream));
        *n_out = m;
    });
}

int dl_session_finalise_device(dl_context* ctx, dl_session* s, int64_t* n_tri,
                               int32_t** pixel_tids, int32_t** pixel_seed_ids,
                               uint8_t** outside_mask)
{
    return guarded([&] {
        if (!ctx || !s || !pixel_tids || !pixel_seed_ids || !outside_mask)
            throw value_error("null argument");
        DL_HIP(hipSetDevice(ctx->device));
        ensure_finalised(ctx, s);
        const size_t px = static_cast<size_t>(s->W) * s->H;
        s->pix_tids.reserve(sizeof(int32_t) * px);
        s->pix_sids.reserve(sizeof(int32_t) * px);
        s->pix_mask.reserve(px);
        split_grid(s->out.as<int32_t>(), static_cast<int64_t>(px), s->pix_tids.as<int32_t>(),
                   s->pix_sids.as<int32_t>(), s->pix_mask.as<uint8_t>(), ctx->stream);
        DL_HIP(hipStreamSynchronize(ctx->stream));
        *pixel_tids = s->pix_tids.as<int32_t>();
        *pixel_seed_ids = s->pix_sids.as<int32_t>();
        *outside_mask = s->pix_mask.as<uint8_t>();
        if (n_tri) *n_tri = s->n_tri;
    });
}

}  // extern "C"
