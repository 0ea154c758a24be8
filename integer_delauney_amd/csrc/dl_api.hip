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

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

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

DevBuf::~DevBuf() { release(); }

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DevBuf::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    release();
    const size_t want = std::max<size_t>(n, 256);
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, want);
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
    if (ptr) (void)hipHostFree(ptr);
}

void PinnedBuf::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    const size_t want = std::max<size_t>(n, 256);
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

__global__ void shift_labels_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                    int64_t n, uint32_t add)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] + add;
}

__global__ void extract_ids_kernel(const int32_t* __restrict__ hw2, uint32_t* __restrict__ ids,
                                   int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = static_cast<uint32_t>(hw2[2 * i]);
}

}  // namespace
}  // namespace dl

using namespace dl;

struct dl_context {
    int device = 0;
    hipStream_t stream = nullptr;
    BfsWork bfs;
    TriWork tri;
    DevBuf xy, sorted, canvas_labels, rank_at;
    DevBuf padA, padB, unA, unB;
    DevBuf out, vgrid, ids, esc;
    PinnedBuf host_small;

    // state of the last triangulation
    int64_t n_tri = 0;
    int last_W = 0, last_H = 0, last_P = 0;
    bool have_padded = false;
    int pad_plane = 0;

    bool timing = false;
    double timings[9] = {0};
    BfsConfig cfg{8, 8, 8};
    bool cfg_auto = true;
    BfsConfig last_cfg{0, 0, 0};
    bool force_unpadded = false;
    std::vector<hipEvent_t> events;
    bool profiling = false;
    KernelProfiler prof;
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

void check_dims(int W, int H, int P)
{
    if (W <= 0 || H <= 0) throw value_error("width and height must be positive");
    if (static_cast<int64_t>(W) + 2LL * P > DL_MAX_CANVAS ||
        static_cast<int64_t>(H) + 2LL * P > DL_MAX_CANVAS)
        throw value_error("canvas (image plus twice the border padding) exceeds " +
                          std::to_string(DL_MAX_CANVAS) + " cells per side");
}

void check_seeds(const void* p, int64_t n)
{
    if (n <= 0) throw value_error("seeds must not be empty");
    if (!p) throw value_error("seeds pointer is null");
}

BfsConfig auto_config(const dl_context* ctx, int Wc, int Hc)
{
    if (!ctx->cfg_auto) return ctx->cfg;
    (void)Wc;
    (void)Hc;
    return BfsConfig{8, 8, 8};
}

struct StageClock {
    dl_context* ctx;
    std::chrono::steady_clock::time_point t0;
    int next = 0;
    explicit StageClock(dl_context* c) : ctx(c), t0(std::chrono::steady_clock::now())
    {
        if (ctx->timing) {
            while (ctx->events.size() < 8) {
                hipEvent_t e;
                DL_HIP(hipEventCreate(&e));
                ctx->events.push_back(e);
            }
            DL_HIP(hipEventRecord(ctx->events[0], ctx->stream));
            next = 1;
        }
        for (double& v : ctx->timings) v = 0.0;
    }
    // Marks the end of stage `idx` (0 seeds, 1 fill, 2 detect, 3 dedup, 4 assign, 5 output).
    void mark(int /*idx*/)
    {
        if (!ctx->timing) return;
        DL_HIP(hipEventRecord(ctx->events[next], ctx->stream));
        ++next;
    }
    void finish(hipStream_t s)
    {
        DL_HIP(hipStreamSynchronize(s));
        if (ctx->timing) {
            for (int i = 1; i < next && i <= 6; ++i) {
                float ms = 0.f;
                DL_HIP(hipEventElapsedTime(&ms, ctx->events[i - 1], ctx->events[i]));
                ctx->timings[i - 1] = ms;
            }
        }
        ctx->timings[6] =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Seeds on the device (d_xy) -> ctx->sorted (id order) and ctx->rank_at.
void prepare_seeds(dl_context* ctx, int W, int H, const int32_t* d_xy, int64_t n, hipStream_t s)
{
    ctx->sorted.reserve(sizeof(uint32_t) * static_cast<size_t>(n));
    ctx->rank_at.reserve(sizeof(int32_t) * static_cast<size_t>(W) * H);
    normalise_seeds(W, H, d_xy, n, ctx->sorted.as<uint32_t>(), ctx->rank_at.as<int32_t>(),
                    ctx->tri, s);
}

// Flood fill of the canvas (W + 2P) x (H + 2P) with the seeds shifted by P.
// With watch_margin >= 0 the fill records in ctx->esc whether a cell inside
// [m, Wc-m) x [m, Hc-m) ever received an update from a cell outside it.
// Returns the canvas descriptor; *plane_out names the plane holding the result.
BfsCanvas fill_canvas(dl_context* ctx, int W, int H, int P, int64_t n, int watch_margin,
                      DevBuf& A, DevBuf& B, hipStream_t s, int* plane_out)
{
    const bool watch = watch_margin >= 0;
    const int Wc = W + 2 * P, Hc = H + 2 * P;
    const size_t cells = static_cast<size_t>(Wc) * Hc;
    A.reserve(sizeof(uint32_t) * cells);
    B.reserve(sizeof(uint32_t) * cells);
    ctx->canvas_labels.reserve(sizeof(uint32_t) * static_cast<size_t>(n));
    const uint32_t add = (static_cast<uint32_t>(P) << 16) | static_cast<uint32_t>(P);
    shift_labels_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
        ctx->sorted.as<uint32_t>(), ctx->canvas_labels.as<uint32_t>(), n, add);
    DL_LAUNCH_CHECK();
    BfsCanvas cv{};
    cv.lab0 = A.as<uint32_t>();
    cv.lab1 = B.as<uint32_t>();
    cv.W = Wc;
    cv.H = Hc;
    cv.P = watch ? watch_margin : 0;
    cv.escape = nullptr;
    if (watch) {
        ctx->esc.reserve(sizeof(int) * 4);
        fill_u32(ctx->esc.as<uint32_t>(), 0u, 4, s);
        cv.escape = ctx->esc.as<int>();
    }
    bfs_init_canvas(cv, ctx->canvas_labels.as<uint32_t>(), n, s);
    std::vector<BfsCanvas> cvs{cv};
    BfsStats st;
    const BfsConfig cfg = auto_config(ctx, Wc, Hc);
    ctx->last_cfg = cfg;
    const int plane = bfs_run(cvs, cfg, ctx->bfs, s, &st);
    ctx->timings[7] += st.launches;
    ctx->timings[8] += st.tile_visits;
    *plane_out = plane;
    return cvs[0];
}

const uint32_t* plane_ptr(const BfsCanvas& cv, int plane) { return plane ? cv.lab1 : cv.lab0; }

// The batch pipeline on device-resident seeds.  Writes the (H, W, 3) grid to
// d_out when d_out != nullptr, else only the tid raster (ctx->tri.tidgrid).
void run_triangulate(dl_context* ctx, int W, int H, const int32_t* d_xy, int64_t n, int P,
                     int32_t* d_out, hipStream_t s, StageClock& clk)
{
    prepare_seeds(ctx, W, H, d_xy, n, s);
    clk.mark(0);
    const uint32_t* det_grid;
    int det_W, det_H;
    const uint32_t* out_labels;
    int out_Wc, out_off;
    ctx->have_padded = false;
    if (P > 0) {
        int plane = 0;
        BfsCanvas cv = fill_canvas(ctx, W, H, P, n, P, ctx->padA, ctx->padB, s, &plane);
        det_grid = plane_ptr(cv, plane);
        det_W = cv.W;
        det_H = cv.H;
        ctx->have_padded = true;
        ctx->pad_plane = plane;
        // Did any padding cell ever update an interior cell?  If not, the
        // interior's trajectory is that of the unpadded fill, pass by pass.
        int escaped = 0;
        DL_HIP(hipMemcpyAsync(ctx->host_small.ptr, ctx->esc.ptr, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        escaped = ctx->host_small.as<int>()[0];
        if (escaped || ctx->force_unpadded) {
            int uplane = 0;
            BfsCanvas ucv = fill_canvas(ctx, W, H, 0, n, -1, ctx->unA, ctx->unB, s, &uplane);
            out_labels = plane_ptr(ucv, uplane);
            out_Wc = W;
            out_off = 0;
        } else {
            out_labels = det_grid;
            out_Wc = det_W;
            out_off = P;
        }
    } else {
        int plane = 0;
        BfsCanvas cv = fill_canvas(ctx, W, H, 0, n, -1, ctx->unA, ctx->unB, s, &plane);
        det_grid = plane_ptr(cv, plane);
        det_W = W;
        det_H = H;
        out_labels = det_grid;
        out_Wc = W;
        out_off = 0;
    }
    clk.mark(1);
    const DetectResult dr = detect_triangles(det_grid, true, det_W, det_H, P,
                                             ctx->rank_at.as<int32_t>(), W,
                                             ctx->sorted.as<uint32_t>(), n, ctx->tri, s);
    clk.mark(2);
    clk.mark(3);
    ctx->n_tri = dr.n_tri;
    clk.mark(4);
    assign_tiled(W, H, ctx->tri.rows.as<int32_t>(), dr.n_tri, ctx->sorted.as<uint32_t>(), out_labels,
                 out_Wc, out_off, out_off, ctx->rank_at.as<int32_t>(), W, d_out, 3, ctx->tri, s);
    clk.mark(5);
    ctx->last_W = W;
    ctx->last_H = H;
    ctx->last_P = P;
}

void upload_seeds(dl_context* ctx, const int32_t* xy, int64_t n)
{
    ctx->xy.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(n));
    DL_HIP(hipMemcpyAsync(ctx->xy.ptr, xy, sizeof(int32_t) * 2 * static_cast<size_t>(n),
                          hipMemcpyHostToDevice, ctx->stream));
}

}  // namespace

extern "C" {

const char* dl_version(void) { return "integer_delauney_amd 0.1.0 (gfx950)"; }

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
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
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
        prepare_seeds(ctx, width, height, ctx->xy.as<int32_t>(), n, s);
        clk.mark(0);
        int plane = 0;
        BfsCanvas cv = fill_canvas(ctx, width, height, 0, n, -1, ctx->unA, ctx->unB, s, &plane);
        clk.mark(1);
        const size_t px = static_cast<size_t>(width) * height;
        ctx->out.reserve(sizeof(int32_t) * 2 * px);
        write_output(width, height, plane_ptr(cv, plane), width, 0, 0, ctx->rank_at.as<int32_t>(),
                     width, nullptr, ctx->out.as<int32_t>(), 2, s);
        DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * px,
                              hipMemcpyDeviceToHost, s));
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
        prepare_seeds(ctx, width, height, ctx->xy.as<int32_t>(), n, s);
        int plane = 0;
        BfsCanvas cv = fill_canvas(ctx, width, height, 0, n, margin, ctx->unA, ctx->unB, s, &plane);
        DL_HIP(hipMemcpyAsync(ctx->host_small.ptr, ctx->esc.ptr, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        if (out_hw2) {
            const size_t px = static_cast<size_t>(width) * height;
            ctx->out.reserve(sizeof(int32_t) * 2 * px);
            write_output(width, height, plane_ptr(cv, plane), width, 0, 0,
                         ctx->rank_at.as<int32_t>(), width, nullptr, ctx->out.as<int32_t>(), 2, s);
            DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * px,
                                  hipMemcpyDeviceToHost, s));
        }
        DL_HIP(hipStreamSynchronize(s));
        *escaped = ctx->host_small.as<int>()[0] != 0;
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
        prepare_seeds(ctx, width, height, ctx->xy.as<int32_t>(), n, s);
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
        write_output(width, height, plane_ptr(cv, plane), width, 0, 0, ctx->rank_at.as<int32_t>(),
                     width, nullptr, ctx->out.as<int32_t>(), 2, s);
        DL_HIP(hipMemcpyAsync(out_hw2, ctx->out.ptr, sizeof(int32_t) * 2 * cells,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
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
        const size_t px = static_cast<size_t>(width) * height;
        ctx->out.reserve(sizeof(int32_t) * 3 * px);
        run_triangulate(ctx, width, height, ctx->xy.as<int32_t>(), n, P, ctx->out.as<int32_t>(), s,
                        clk);
        DL_HIP(hipMemcpyAsync(out_hw3, ctx->out.ptr, sizeof(int32_t) * 3 * px,
                              hipMemcpyDeviceToHost, s));
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
        if (ctx->profiling) set_current_profiler(&ctx->prof);
        struct Drop {
            dl_context* c;
            hipStream_t st;
            ~Drop()
            {
                if (c->profiling) {
                    (void)hipStreamSynchronize(st);
                    c->prof.collect();
                }
                set_current_profiler(nullptr);
            }
        } drop{ctx, s};
        StageClock clk(ctx);
        run_triangulate(ctx, width, height, d_seeds_xy, n, P, d_out_hw3, s, clk);
        clk.finish(s);
        if (n_tri) *n_tri = ctx->n_tri;
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
        prepare_seeds(ctx, width, height, ctx->xy.as<int32_t>(), n, s);
        clk.mark(0);
        const size_t px = static_cast<size_t>(width) * height;
        DetectResult dr;
        ctx->have_padded = false;
        if (P > 0) {
            int plane = 0;
            BfsCanvas cv = fill_canvas(ctx, width, height, P, n, -1, ctx->padA, ctx->padB, s,
                                       &plane);
            ctx->have_padded = true;
            ctx->pad_plane = plane;
            clk.mark(1);
            dr = detect_triangles(plane_ptr(cv, plane), true, cv.W, cv.H, P,
                                  ctx->rank_at.as<int32_t>(), width, ctx->sorted.as<uint32_t>(), n,
                                  ctx->tri, s);
        } else {
            if (!voronoi_grid_hw2) throw value_error("voronoi_grid is required at padding 0");
            ctx->vgrid.reserve(sizeof(int32_t) * 2 * px);
            ctx->ids.reserve(sizeof(uint32_t) * px);
            DL_HIP(hipMemcpyAsync(ctx->vgrid.ptr, voronoi_grid_hw2, sizeof(int32_t) * 2 * px,
                                  hipMemcpyHostToDevice, s));
            extract_ids_kernel<<<static_cast<unsigned>((px + 255) / 256), 256, 0, s>>>(
                ctx->vgrid.as<int32_t>(), ctx->ids.as<uint32_t>(), static_cast<int64_t>(px));
            DL_LAUNCH_CHECK();
            clk.mark(1);
            dr = detect_triangles(ctx->ids.as<uint32_t>(), false, width, height, 0,
                                  ctx->rank_at.as<int32_t>(), width, ctx->sorted.as<uint32_t>(), n,
                                  ctx->tri, s);
        }
        clk.mark(2);
        clk.mark(3);
        ctx->n_tri = dr.n_tri;
        ctx->tri.tidgrid.reserve(sizeof(int32_t) * px);
        assign_tiled(width, height, ctx->tri.rows.as<int32_t>(), dr.n_tri,
                     ctx->sorted.as<uint32_t>(), nullptr, 0, 0, 0, nullptr, width,
                     ctx->tri.tidgrid.as<int32_t>(), 1, ctx->tri, s);
        clk.mark(4);
        DL_HIP(hipMemcpyAsync(out_tid_hw, ctx->tri.tidgrid.ptr, sizeof(int32_t) * px,
                              hipMemcpyDeviceToHost, s));
        clk.mark(5);
        clk.finish(s);
        ctx->last_W = width;
        ctx->last_H = height;
        ctx->last_P = P;
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
        const int P = ctx->last_P;
        const int Wc = ctx->last_W + 2 * P, Hc = ctx->last_H + 2 * P;
        if (out_w) *out_w = Wc;
        if (out_h) *out_h = Hc;
        if (!out) return;
        if (!ctx->have_padded) throw value_error("no padded canvas was computed by the last call");
        DL_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const size_t cells = static_cast<size_t>(Wc) * Hc;
        ctx->out.reserve(sizeof(int32_t) * 2 * cells);
        const uint32_t* lab = ctx->pad_plane ? ctx->padB.as<uint32_t>() : ctx->padA.as<uint32_t>();
        write_output(Wc, Hc, lab, Wc, 0, P, ctx->rank_at.as<int32_t>(), ctx->last_W, nullptr,
                     ctx->out.as<int32_t>(), 2, s);
        DL_HIP(hipMemcpyAsync(out, ctx->out.ptr, sizeof(int32_t) * 2 * cells,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
