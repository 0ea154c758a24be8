// Internal declarations shared by the HIP translation units of
// libdelauney_hip.so.  Nothing here is part of the C ABI (include/delauney_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dl {

// ---------------------------------------------------------------------------
// Errors.  ValueError-class failures (bad input) and RuntimeError-class
// failures (device, allocation) travel as C++ exceptions inside the library and
// become status codes at the C ABI.
// ---------------------------------------------------------------------------
struct value_error : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct runtime_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define DL_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) ::dl::throw_hip(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define DL_LAUNCH_CHECK() DL_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Label encoding.  A cell's state is the packed canvas position of its owning
// seed, (x << 16) | y.  Unsigned order on labels is (x, y) lexicographic order,
// which is the reference's seed-id order (np.lexsort((y, x)),
// src/delauney/reference/voronoi.py:64-66), so "higher id wins a tie"
// (voronoi.py:181-186) is "higher label wins".
// ---------------------------------------------------------------------------
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kMaxSide = 32768;  // keeps d^2 < 2^31-1 and detection keys < 2^32

__host__ __device__ inline uint32_t pack_label(int x, int y)
{
    return (static_cast<uint32_t>(x) << 16) | static_cast<uint32_t>(y);
}

// ---------------------------------------------------------------------------
// Device memory: a grow-only buffer owned by a context or session.
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void reserve(size_t n);  // keeps nothing on growth
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct PinnedBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf();
    void reserve(size_t n);
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// BFS (Voronoi flood fill) over one or more canvases.
// ---------------------------------------------------------------------------
struct BfsCanvas {
    uint32_t* lab0;     // ping-pong label planes, row-major W x H
    uint32_t* lab1;
    int W, H;           // canvas size
    int P;              // > 0: image interior is [P, W-P) x [P, H-P), watched for
                        //      updates that arrive from the padding
    int tiles_x, tiles_y;
    int tile_base;      // first tile index of this canvas in the flag arrays
    int* escape;        // set to 1 when a padding cell updates an interior cell
};

struct BfsConfig {
    int waves;  // waves per tile (tile height = waves * rows)
    int rows;   // rows held per thread
    int iters;  // flood-fill iterations per launch (halo width)
};

struct BfsStats {
    int launches = 0;
    int tile_visits = 0;  // active tiles summed over launches
};

// Workspace that persists across BFS calls of one context (flags, counters).
struct BfsWork {
    DevBuf flags;     // 2 * total_tiles ints
    DevBuf counters;  // per-launch changed-tile counters
    DevBuf canvases;  // device copy of the canvas table
    DevBuf tile_map;  // tile -> canvas index (multi-canvas launches)
    PinnedBuf host_counters;
};

// Interior tile size for a configuration.
inline int bfs_tile_iw(const BfsConfig& c) { return 128 - 2 * c.iters; }
inline int bfs_tile_ih(const BfsConfig& c) { return c.waves * c.rows - 2 * c.iters; }

bool bfs_config_supported(const BfsConfig& c);

// Writes kInvalid everywhere and each seed's own label at its cell, in both
// label planes of every canvas.  seeds_canvas: packed labels already in canvas
// coordinates.
void bfs_init_canvas(const BfsCanvas& cv, const uint32_t* d_labels, int64_t n,
                     hipStream_t s);

// Runs the flood fill to convergence on the given canvases (all tiles start
// active).  Returns the plane index (0 or 1) holding the final labels.
int bfs_run(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
            hipStream_t s, BfsStats* stats);

// One-thread-per-cell, one-pass-per-launch version of the same schedule.  Kept
// as an independent second implementation for cross-checking the tiled kernel.
int bfs_run_simple(const BfsCanvas& cv, BfsWork& work, hipStream_t s, int64_t* iters);

// ---------------------------------------------------------------------------
// Seeds, detection, assignment, output.
// ---------------------------------------------------------------------------
struct SeedSet {
    // Outputs of normalise_seeds, all on the device.
    uint32_t* sorted;   // n packed image-space labels in id order
    int32_t* rank_at;   // W x H: id of the seed at (x, y); only seed cells written
};

struct TriWork {
    DevBuf bitmap;      // seed presence / registration bitmaps
    DevBuf scan;        // word prefix sums
    DevBuf scan_tmp;
    DevBuf flags;       // device status flags
    PinnedBuf host_flags;
    DevBuf table_tag;   // registration hash table slots
    DevBuf cand;        // raw candidate records
    DevBuf counts;      // per-thread / per-block candidate counts
    DevBuf offsets;     // their prefix sums
    DevBuf positions;   // compact list of block positions holding candidates
    DevBuf rows;        // (T, 5) triangle rows
    DevBuf tidgrid;     // W x H tid raster
    DevBuf bin_counts, bin_offsets, bin_list;   // triangle -> tile bins
};

// Validates (bounds, duplicates) and orders seeds on the device.
// d_xy: n (x, y) int32 pairs in caller order.  Throws value_error on
// duplicates or out-of-bounds coordinates.
void normalise_seeds(int W, int H, const int32_t* d_xy, int64_t n, uint32_t* d_sorted,
                     int32_t* d_rank_at, TriWork& work, hipStream_t s);

// Exclusive prefix sum of popcount(words[i]) into out[i]; returns nothing on the
// host (total is out[n] written by the kernel).  out has n+1 entries.
void popcount_scan(const uint32_t* words, int64_t n, uint32_t* out, DevBuf& tmp,
                   hipStream_t s);

// Exclusive prefix sum of in[0..n) into out[0..n); out[n] receives the total.
void exclusive_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, DevBuf& tmp, hipStream_t s);

struct DetectResult {
    int64_t n_tri = 0;
    int64_t n_quads = 0;
    int64_t n_cand = 0;
};

// Triangle extraction on a canvas of ids or labels.  If labels is true, grid
// holds packed canvas labels and ids are looked up through rank_at (image
// space, shifted by P); otherwise grid holds seed ids directly.
// Writes (T, 5) rows (x - P, y - P, a, b, c) into work.rows.
DetectResult detect_triangles(const uint32_t* grid, bool labels, int Wp, int Hp, int P,
                              const int32_t* rank_at, int W_img, const uint32_t* sorted,
                              int64_t n_seeds, TriWork& work, hipStream_t s);

// (H, W, C) int32 output.  Output pixel (x, y) reads canvas cell
// (x + off, y + off); the owner's image position is its canvas position minus
// shift.  channels 0/1 = (rank(owner), d^2), channel 2 (C == 3) = tidgrid.
void write_output(int W, int H, const uint32_t* canvas_labels, int Wc, int off, int shift,
                  const int32_t* rank_at, int W_img, const int32_t* tidgrid, int32_t* out, int C,
                  hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// Pixel assignment binned by 64x64 tiles and fused with the output write.
// C == 3: out is the (H, W, 3) grid, channels 0/1 from the canvas labels as in
// write_output; C == 1: out is the (H, W) tid raster.
void assign_tiled(int W, int H, const int32_t* rows, int64_t n_tri, const uint32_t* sorted,
                  const uint32_t* labels, int Wc, int off, int shift, const int32_t* rank_at,
                  int W_img, int32_t* out, int C, struct TriWork& work, hipStream_t s);

// ---------------------------------------------------------------------------
// Per-kernel timing.  When a profiler is installed for the calling thread,
// launch sites bracket their kernels with events; collect() turns the pairs
// into per-kind counts and milliseconds after the stream has been synchronised.
// ---------------------------------------------------------------------------
enum KernelKind {
    kKFill = 0,      // flood-fill tile kernel
    kKSeeds = 1,     // seed marking / ranking / scan
    kKDetect = 2,    // triangle candidate detection
    kKRegister = 3,  // first-registration marking, scan, row emission
    kKAssign = 4,    // triangle binning for the pixel assignment
    kKOutput = 5,    // per-tile assignment fused with the (H, W, C) output
    kKInit = 6,      // canvas initialisation and fills
    kKinds = 7
};

struct KernelProfiler {
    int64_t count[kKinds] = {0};
    double ms[kKinds] = {0};
    std::vector<hipEvent_t> pool;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    hipEvent_t take();
    void collect();   // the stream of every pending event must be synchronised
    void reset();
    ~KernelProfiler();
};

KernelProfiler* current_profiler();
void set_current_profiler(KernelProfiler* p);

// Brackets one kernel launch: construct before the launch, destroy after it.
struct ProfScope {
    KernelProfiler* p;
    hipStream_t s;
    KernelProfiler::Pending rec;
    ProfScope(int kind, hipStream_t stream);
    ~ProfScope();
};

}  // namespace dl
