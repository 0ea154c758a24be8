// Internal declarations shared by the HIP translation units of
// libdelauney_hip.so.  Nothing here is part of the C ABI (include/delauney_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
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
    bool arena = false;   // carved from the calling thread's arena (never freed)
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
    bool arena = false;   // carved from the calling thread's host arena
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf();
    void reserve(size_t n);
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// Arenas.  While an arena is installed for the calling thread, every DevBuf /
// PinnedBuf that grows takes its bytes from the arena's device / host region
// with a bump pointer (256-byte aligned) instead of allocating, and a carved
// buffer is never freed.  The no-allocation entry point installs one over its
// static scratch for the duration of a call; running out is a runtime_error.
// ---------------------------------------------------------------------------
struct Arena {
    char* dev = nullptr;
    size_t dev_cap = 0, dev_used = 0;
    char* host = nullptr;
    size_t host_cap = 0, host_used = 0;
};
Arena*& current_arena();

// ---------------------------------------------------------------------------
// Stream capture support.  A captured graph holds raw device and page-locked
// host addresses, so it stays valid only while no buffer has been reallocated
// or freed: every reallocation and release of a DevBuf or PinnedBuf advances a
// process-wide allocation epoch, and a graph records the epoch it was captured
// at.  While the calling thread captures, a reserve that would reallocate
// throws capture_abort instead (the pipeline then runs uncaptured).
// ---------------------------------------------------------------------------
struct capture_abort : std::runtime_error {
    using std::runtime_error::runtime_error;
};
uint64_t alloc_epoch();

// A captured graph holds kernels only.  The host -> device uploads of the
// captured stretch are recorded instead (staging address, device address,
// the bytes seen at capture) and issued as ordinary stream copies before
// every replay, and its device -> host copies are issued after it: no kernel
// of the stretch writes an uploaded region or reads a copied-back one
// before the end, so stream order makes this equivalent.
struct StagedBytes {
    void* host;
    void* dev;
    std::vector<char> bytes;
};
struct CopyBack {
    void* host;
    const void* dev;
    size_t n;
};
struct CaptureLog {
    std::vector<StagedBytes> staged;
    std::vector<CopyBack> copy_back;
    bool poisoned = false;   // a buffer was freed while capturing
};
// Non-null while the calling thread captures a stream.
CaptureLog*& capture_log();

// Device -> host copy on s; recorded for after the replay while capturing.
void copy_back_async(void* host, const void* dev, size_t n, hipStream_t s);

// Host -> device copy of n bytes through bytes [offset, offset + n) of the
// page-locked buffer `stage`, which the caller has reserved.  One buffer per
// upload site: a site is rewritten only after the stream it copied on has been
// synchronised.
void stage_h2d(PinnedBuf& stage, size_t offset, void* dst, const void* src, size_t n,
               hipStream_t s);

// Several fills (p[0 .. n) = v) in one launch.
struct FillSeg {
    uint32_t* p;
    size_t n;
    uint32_t v;
};
struct FillSegs {
    FillSeg seg[4];
    int k;
};
struct FillBlocks {
    unsigned b[5];
};
void fill_u32_multi(const FillSegs& segs, hipStream_t s);

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
    int converged = -1;   // index of the first launch that changed nothing
};

// Workspace that persists across BFS calls of one context (flags, counters).
struct BfsWork {
    DevBuf flags;     // 2 * total_tiles ints
    DevBuf counters;  // per-launch "some tile changed" flags, then the visit total
    DevBuf visits;    // per tile: launches that processed it
    DevBuf list;      // active tiles of the coming launch (large fills)
    int list_min_tiles = 8192;   // fills with fewer tiles launch a block per tile
    DevBuf canvases;  // device copy of the canvas table
    DevBuf tile_map;  // tile -> canvas index (multi-canvas launches)
    PinnedBuf host_counters;
    PinnedBuf host_canvases;   // staging of the canvas table
    // bfs_launch_fixed: count tile visits (statistics only) into the visit total
    bool count_visits = true;
    // bfs_run waits for a chunk's counters with an event; false: with a stream
    // synchronisation (no event objects are created)
    bool use_events = true;
    // Key encoding of the fill's tile kernels (set per fill from the canvas
    // sizes): 1 = min encoding, 0 = compare and select.  allow_min_encoding
    // false (a test hook) keeps every fill on the compare-and-select one.
    int enc = 0;
    bool allow_min_encoding = true;
    // Padding watch of the fill (set per fill): true when every watched
    // canvas's interior rows begin and end on wave-row boundaries
    // (bfs_watch_aligned), so the tile kernels test one row per y pass and no
    // row condition in the x passes.  allow_aligned_watch false (a test hook)
    // keeps every fill on the per-row form.
    bool aligned = false;
    bool allow_aligned_watch = true;
    DevBuf scan_tmp;  // scratch of the fill record's prefix scan
};

// Interior tile size for a configuration.
inline int bfs_tile_iw(const BfsConfig& c) { return 128 - 2 * c.iters; }
inline int bfs_tile_ih(const BfsConfig& c) { return c.waves * c.rows - 2 * c.iters; }

// A wave of a tile holds rows [y0, y0 + rows) with y0 = ty * IH - iters +
// w * rows.  When IH is a multiple of rows every y0 is -iters modulo rows, and
// the watched interior [P, H - P) of a canvas of height H starts and ends on
// wave boundaries when P and H - P are such rows.  (P <= 0: nothing watched.)
inline bool bfs_watch_aligned(int P, int H, const BfsConfig& c)
{
    if (P <= 0) return true;
    const int R = c.rows, K = c.iters, IH = bfs_tile_ih(c);
    if (R <= 0 || IH % R != 0) return false;
    auto boundary = [&](int y) { return ((y + K) % R + R) % R == 0; };
    return boundary(P) && boundary(H - P);
}

bool bfs_config_supported(const BfsConfig& c);

// Writes kInvalid everywhere and each seed's own label at its cell in the
// first label plane (the second needs none, as for init_canvases).  d_labels:
// packed labels already in canvas coordinates.
void bfs_init_canvas(const BfsCanvas& cv, const uint32_t* d_labels, int64_t n,
                     hipStream_t s);

// Runs the flood fill to convergence on the given canvases (all tiles start
// active).  Returns the plane index (0 or 1) holding the final labels.
// With visits_async (page-locked host memory) the visit total is copied there
// asynchronously and the call returns without waiting for the stream; the
// caller reads it after its next synchronisation (stats->tile_visits is then
// left at 0).
//
// start_launch > 0 resumes a fill whose launches 0 .. start_launch - 1 already
// ran (bfs_launch_fixed): the label planes and tile flags carry over.
int bfs_run(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
            hipStream_t s, BfsStats* stats, int* visits_async = nullptr, int start_launch = 0);

// Issues exactly n_launches tile launches of a fresh fill without waiting.
// Each launch's "some tile changed" flag, then the visit total, end up in
// work.counters (n_launches + 1 ints); with host_changed (page-locked) they
// are also copied there asynchronously.  The fill has converged iff some
// launch changed nothing; the labels are then final in both planes.  Returns
// the plane the last launch wrote.  counters_zeroed: the caller already
// cleared the first bfs_fixed_counter_words(n_launches) words of
// work.counters on the stream.
int bfs_launch_fixed(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                     hipStream_t s, int n_launches, int* host_changed,
                     bool counters_zeroed = false);
// Counter words bfs_launch_fixed uses (n change flags, the visit total, n list
// counts and one spare).
inline size_t bfs_fixed_counter_words(int n_launches)
{
    return 2 * static_cast<size_t>(n_launches + 1);
}

// ---------------------------------------------------------------------------
// Fill records: replaying only what changed (sessions).
//
// A record holds one fill's trajectory: for every tile and every launch that
// changed the tile's interior, the interior after that launch (a pool slot)
// and the launch's change mask.  The next fill of the same canvas (the
// session's next seed set) computes only the tiles whose 3 x 3 neighbourhood
// differs from the recorded trajectory, launch by launch, and takes every
// other tile's state from the record: bit for bit the launch-by-launch fill.
// States are numbered by launches done: state s = after s launches.
// An event created on first use and destroyed with its owner.
struct HipEvent {
    hipEvent_t e = nullptr;
    HipEvent() = default;
    HipEvent(const HipEvent&) = delete;
    HipEvent& operator=(const HipEvent&) = delete;
    ~HipEvent()
    {
        if (e) (void)hipEventDestroy(e);
    }
    hipEvent_t get()
    {
        if (!e) DL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }
};

struct FillRecord {
    bool valid = false;      // csr/tile_off describe the last fill of this canvas
    bool escaped = false;    // that fill's padding watch fired (no reuse then)
    bool synced = false;     // prev_bitmap is the seed set of the recorded fill
    int W = 0, H = 0, P = 0, tiles = 0;
    BfsConfig cfg{};
    int states = 0;          // states of the recorded fill (its launches + 1)
    int64_t entries = 0;     // its entries
    DevBuf pool;             // slots of IW x IH labels; slot 0 = all undefined
    int64_t pool_cap = 0;    // slots
    DevBuf csr[2];           // entries by tile, state order: int4 (state, mask, slot, 0)
    DevBuf tile_off[2];      // tiles + 1 offsets into csr
    int cur = 0;             // csr[cur] / tile_off[cur]: the record
    DevBuf app;              // entries of the fill being recorded, append order:
                             // int4 (tile, state, mask, slot)
    DevBuf rank;             // their per-tile ranks
    int64_t ent_cap = 0;
    DevBuf tstate;           // per tile: entry cursors, dirty and mask by parity, counts
    DevBuf ctl;              // counters (slots used, entries used, overflow, list count,
                             // per-launch changes)
    DevBuf prev_bitmap;      // seed presence bitmap of the recorded fill
    DevBuf list;             // tiles computed in the coming launch
    DevBuf freelist;         // pool slots below hw that the record does not use
    DevBuf mark;             // scratch: slots the record uses
    int64_t hw = 1;          // slots [0, hw) have been handed out
    int64_t free_n = 0;      // entries of freelist
    PinnedBuf host;          // read-back of the counters
    HipEvent polled;         // replay: the change flags read back so far have landed
    int64_t last_computed = 0;   // tile runs of the last fill (statistics)
    uint64_t gen = 0;            // a process-wide number given to each record when it is made
};

// Device view of a record's buffers for the fill kernels.
struct RecArgs {
    uint32_t* pool;
    int64_t pool_cap;            // slots
    int4* app;                   // entries of the fill being recorded
    int* rank;
    int* nent;                   // per tile: entries appended so far
    unsigned long long* used;    // [0] slots handed out by this fill, [1] entries
    int* overflow;               // set when a slot or an entry did not fit
    int64_t ent_cap;
    // the j-th slot handed out is freelist[j] for j < free_n, then
    // fresh + (j - free_n)
    const int* freelist;
    long long free_n;
    long long fresh;
};

// The padded fill of a session's canvas, launch by launch, recording its
// trajectory into rec (from scratch: the pool is reused from slot 1).  Same
// result and plane convention as bfs_run.  The record is valid afterwards
// unless it overflowed its capacities.
int bfs_run_recorded(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                     hipStream_t s, BfsStats* stats, FillRecord& rec);
// The same fill computed from rec, the record of the previous fill of this
// canvas: prev_bitmap and bitmap are the seed-presence bitmaps (bit
// x * img_h + y of the img_w x img_h image, which sits at offset img_pad in
// the canvas) of the recorded fill and of this one, and only tiles whose
// neighbourhood differs from the record are computed.  Records this fill in
// turn.  Returns the plane holding the labels.  retained: the canvas's first
// plane still holds the recorded fill's final labels (nothing has written it
// since), so only the tiles near the changed seeds are seeded and only the
// tiles the replay touches are written back.
int bfs_run_incremental(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                        hipStream_t s, BfsStats* stats, FillRecord& rec,
                        const uint32_t* prev_bitmap, const uint32_t* bitmap, int img_w,
                        int img_h, int img_pad, bool retained = false);
// Can rec serve the next fill of this canvas with this shape?
bool bfs_record_usable(const FillRecord& rec, const BfsCanvas& cv, const BfsConfig& cfg);
// Shapes a record can be made with.
bool bfs_record_supported(const BfsConfig& cfg, const BfsCanvas& cv);

// One-thread-per-cell, one-pass-per-launch version of the same schedule.  Kept
// as an independent second implementation for cross-checking the tiled kernel.
int bfs_run_simple(const BfsCanvas& cv, BfsWork& work, hipStream_t s, int64_t* iters);

// ---------------------------------------------------------------------------
// Seeds, detection, assignment, output -- over a batch of images.
//
// Every stage runs once per batch: per-image buffers are concatenated, each
// image's slice starts at the offsets below, and a kernel maps its global work
// index to the image by binary search over an offset table.  A single image is
// a batch of one.
// ---------------------------------------------------------------------------
struct ImageDesc {
    int W, H, P;          // image size and border padding
    int Wc, Hc;           // detection canvas (W + 2P) x (H + 2P)
    int64_t n;            // seeds
    int64_t seed_off;     // first seed: xy pairs, sorted labels
    int64_t bm_off;       // first word of the seed presence bitmap (W*H bits)
    int64_t pix_off;      // first pixel of the output (times channels)
    int64_t cv_off;       // first cell of the detection canvas planes
    int64_t nb;           // 2x2 blocks of the canvas, (Wc-1)*(Hc-1)
    int64_t det_off;      // first detection thread block
    int bx_per_row;       // detection thread blocks per row of 2x2 blocks
    int tiles_x;          // 64x64 assignment tiles per row
    int64_t key_off;      // first word of the registration bitmap (4 scans x rows of
                          // blocks, each row padded to whole detection strips)
    int64_t tile_off;     // first assignment tile
    // Source of output channels 0/1: cell (x + out_off, y + out_off) of a
    // plane out_W wide; an owner's image position is its label's position
    // minus (out_shift, out_shift).
    const uint32_t* out_lab;
    int out_W, out_off, out_shift;
    const int32_t* xy;    // the image's n (x, y) seed pairs (device)
    int32_t* out;         // its output: (H, W, 3) grid or (H, W) tid raster (device)
};

struct BatchPlan {
    std::vector<ImageDesc> img;
    // Offset tables (n_img + 1 entries) for the binary searches.
    std::vector<int64_t> seed_seg, det_seg, tile_seg;
    int64_t n_seeds = 0, n_words = 0, n_pix = 0, n_cells = 0;
    int64_t n_det_blocks = 0, n_key_words = 0, n_tiles = 0;
    int max_seed_bits = 1;   // bits of the largest per-image seed id
};

// Fills the offsets of plan.img (W, H, P and n set by the caller).
void plan_batch(BatchPlan& plan);

struct TriWork;

// Sorted seed id of a seed position: the position's rank among the set bits
// of the presence bitmap (bit x*H + y of an image, words from bm_off on).
// Each word is stored with the exclusive popcount prefix of all words before
// it (the seeds of earlier images included), so a lookup is one 8-byte load
// from an array of W*H/4 bytes that stays in cache, where an image-sized id
// map would be a scattered HBM read.
struct SeedIndex {
    const uint2* wp = nullptr;   // per bitmap word: (prefix, word)
};

// seed_off: the image's first seed (= the prefix at its first word).  -1
// where no seed sits (only for a seed set that failed validation).
__device__ __forceinline__ int32_t seed_id_at(const SeedIndex& si, int64_t bm_off,
                                             int64_t seed_off, int H, int x, int y)
{
    const uint64_t bit = static_cast<uint64_t>(x) * static_cast<uint64_t>(H) +
                         static_cast<uint64_t>(y);
    const uint2 e = si.wp[bm_off + static_cast<int64_t>(bit >> 5)];
    const uint32_t m = 1u << (bit & 31);
    if (!(e.y & m)) return -1;
    return static_cast<int32_t>(static_cast<int64_t>(e.x) - seed_off +
                                __popc(e.y & (m - 1u)));
}

struct TriWork {
    DevBuf seed_bm;     // seed presence bitmaps
    DevBuf seed_idx;    // per word: (popcount prefix, word) (SeedIndex)
    DevBuf bitmap;      // registration bitmaps
    DevBuf scan;        // word prefix sums
    DevBuf scan_tmp;
    DevBuf flags;       // device status flags (per image, then batch-wide)
    PinnedBuf host_flags;
    DevBuf table_tag;   // registration hash table slots
    DevBuf table_tri;   // their triplets (hashed tags only)
    DevBuf cand;        // raw candidate records
    DevBuf counts;      // per-thread / per-block candidate counts
    DevBuf offsets;     // their prefix sums
    DevBuf positions;   // compact list of block positions holding candidates
    DevBuf route_ent, route_idx;   // owner-bucket registration: entries, their slots
    DevBuf route_aux;   // bucket counters, per-slot first flags, per-position scans
    DevBuf rows;        // (T, 5) triangle rows of all images, image after image
    DevBuf tidgrid;     // W x H tid raster
    DevBuf bin_counts, bin_offsets, bin_list;   // triangle -> tile bins
    DevBuf desc;        // device copy of the plan's ImageDesc array, then (from
                        // segs_off) the offset tables + tri_base
    size_t segs_off = 0;
    DevBuf status_dev;  // status words gathered for the read-back
    PinnedBuf host_plan;    // staging of upload_plan (descriptors + offset tables)
    PinnedBuf host_stage;   // staging of upload_desc
    PinnedBuf host_seed_flags;   // per-image seed validation flags (read back late)
    std::vector<int64_t> tri_base;   // host copy: first row of each image, then the total
    uint32_t rows_cap = 0;           // capacity of rows (detection)
    int64_t list_cap = 0;            // capacity of bin_list (assignment)
    int spec_tiny = 0;               // test hook: 1 every speculative capacity of one,
                                     // 2 only the registration buckets'
};

// Uploads the plan's descriptors and offset tables (the output sources may be
// re-uploaded later with upload_desc).
void upload_plan(const BatchPlan& plan, TriWork& work, hipStream_t s);
void upload_desc(const BatchPlan& plan, TriWork& work, hipStream_t s);

// Per-image seed validation flags live in work.flags from this word on (the
// words before it are detection's status).
constexpr int kSeedFlagBase = 8;

enum class SeedCheck {
    kNow,          // read the flags back and check them before returning
    kAfterWait,    // copy them to work.host_seed_flags; checked after the next wait
    kWithStatus,   // leave them on the device for the final status read
};

// Validates (bounds, duplicates) and orders the seeds of every image (read
// from each image's xy).  Throws value_error naming the first failing image
// when the batch has more than one.
// Unless checked now, the call returns at once and check_seed_flags must run
// on the read-back flags after the stream's next synchronisation point.
// Until then invalid seeds only produce garbage labels (every kernel in
// between stays in bounds).
void normalise_seeds(const BatchPlan& plan, uint32_t* d_sorted, TriWork& work, hipStream_t s,
                     SeedCheck check = SeedCheck::kNow);
inline SeedIndex seed_index(const TriWork& w)
{
    return SeedIndex{w.seed_idx.as<uint2>()};
}
// hf: the images' read-back seed flags.
void check_seed_flags(const BatchPlan& plan, const uint32_t* hf);

// Writes kInvalid over the first plane of every canvas, then each seed's
// canvas label (image label + (P, P)) at its cell.  plane_b needs no
// initialisation: the flood fill's first launch visits every tile and writes
// every cell of the second plane.  plane_a/plane_b: n_cells each.  extra: one
// more fill issued in the same launch as the plane's.
void init_canvases(const BatchPlan& plan, const uint32_t* d_sorted, uint32_t* plane_a,
                   uint32_t* plane_b, struct TriWork& work, hipStream_t s,
                   const FillSeg* extra = nullptr, const FillSeg* extra2 = nullptr);

// Exclusive prefix sum of popcount(words[i]) into out[i]; out[n] receives the
// total.  out has n+1 entries.
void popcount_scan(const uint32_t* words, int64_t n, uint32_t* out, DevBuf& tmp,
                   hipStream_t s);

// The same scan written as pairs[i] = (prefix before word i, word i).
void popcount_scan_pairs(const uint32_t* words, int64_t n, uint2* pairs, DevBuf& tmp,
                         hipStream_t s);

// Exclusive prefix sum of in[0..n) into out[0..n); out[n] receives the total.
void exclusive_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, DevBuf& tmp, hipStream_t s);

struct DetectResult {
    int64_t n_tri = 0;     // triangles of the whole batch
    int64_t n_quads = 0;
    int64_t n_cand = 0;
};

// Triangle extraction on every image's canvas.  grid_base + cv_off is image
// i's canvas; it holds packed canvas labels (labels = true; ids through
// the seed index) or seed ids (labels = false).  Writes rows
// (x - P, y - P, a, b, c) of all images, image after image, into work.rows and
// the first row of each image into work.tri_base (host) and the device table.
//
// speculative = false sizes every buffer exactly, reading counts back (host
// syncs), and fills work.tri_base.  speculative = true sizes them from the
// seed count without any sync; capacity overflows, hash collisions and the
// triangle counts are then checked by read_status after the assignment.
// flags_zeroed: the detection status words (work.flags[0 .. 8)) are known to
// be zero already (normalise_seeds clears them), so no clearing launch.
// after_detect (optional) runs on the host once the candidate kernels (count,
// scan, write) are issued and before the registration kernels (insert, mark,
// scan, emit): the stage clock's boundary between detection and
// deduplication.
DetectResult detect_triangles(const BatchPlan& plan, const uint32_t* grid_base, bool labels,
                              const uint32_t* sorted, TriWork& work,
                              hipStream_t s, bool speculative, bool flags_zeroed = false,
                              const std::function<void()>& after_detect = {});

struct SpeculativeStatus {
    bool bad_id = false;   // a canvas cell without owner / an id out of range
    bool redo = false;     // a speculative capacity was exceeded or a hash collided
    bool bin_redo = false; // a tile's fixed bin slice overflowed (assignment only)
    int64_t n_tri = 0;
    std::vector<int> escaped;   // the padding watch's flags, per image
};

// One sync: status word, triangle offsets (into work.tri_base), bin-list
// total, and n_esc padding-watch flags from d_esc (may be null).
SpeculativeStatus read_status(const BatchPlan& plan, TriWork& work, const int* d_esc, int n_esc,
                              hipStream_t s);
// read_status in two halves: one kernel gathers the status words on the
// device and one copy brings them into work.host_flags; after the stream has
// been synchronised parse_status reads them.  Optionally the gather also takes
// the images' seed flags (seed_flags) and n_fill words from d_fill (a
// predicted fill's change flags and visit total); parse_status then points
// *seed_out and *fill_out at their copies (valid until the next status read).
void issue_status(const BatchPlan& plan, TriWork& work, const int* d_esc, int n_esc,
                  bool seed_flags, const int* d_fill, int n_fill, hipStream_t s);
SpeculativeStatus parse_status(const BatchPlan& plan, TriWork& work, int n_esc, bool seed_flags,
                               int n_fill, const uint32_t** seed_out, const int** fill_out);

// Pixel assignment binned by 64x64 tiles and fused with the output write into
// each image's out: C == 3, the (H, W, 3) grid with channels 0/1 from its
// output label source; C == 1, the (H, W) tid raster.
// n_tri < 0: the count is on the device (after a speculative detection);
// speculative = true sizes the bin list without a sync (checked by read_status);
// with allow_fixed it may bin in one pass into fixed per-tile slices, whose
// overflow the status reports as bin_redo (the assignment alone runs again).
// counts_zeroed: work.bin_counts (2 * n_tiles + 1 words) is already clear on
// the stream (normalise_seeds clears it with the seed bitmap).
void assign_tiled(const BatchPlan& plan, const int32_t* rows, int64_t n_tri,
                  const uint32_t* sorted, int C, TriWork& work,
                  hipStream_t s, bool speculative, bool allow_fixed = true,
                  bool counts_zeroed = false);

// (H, W, C) int32 output of one image without triangles (C == 2) or with the
// tid raster (C == 3).  Output pixel (x, y) reads canvas cell (x + off, y + off);
// the owner's image position is its canvas position minus shift.
void write_output(int W, int H, const uint32_t* canvas_labels, int Wc, int off, int shift,
                  const SeedIndex& si, int64_t bm_off, int64_t seed_off, int H_img,
                  const int32_t* tidgrid, int32_t* out, int C,
                  hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// ---------------------------------------------------------------------------
// Incremental sessions: the live seed set on the device.
// ---------------------------------------------------------------------------
struct SessionState {
    int W = 0, H = 0;
    int64_t max_seeds = 0;
    int64_t words = 0;     // presence bitmap words
    int64_t n_live = 0;
    bool have_values = false;   // some insert carried per-seed values
    DevBuf bitmap;         // bit x*H + y set for every live seed
    DevBuf list, spare;    // live (x, y) pairs in insertion order; compaction target
    DevBuf values, values_spare;   // one float per live seed, in list order
    DevBuf backup_list, backup_values;   // undo copies of an update
    DevBuf mine;           // per-call: which bits this call changed (rollback)
    DevBuf keep;           // compaction flags and their prefix sums
    DevBuf flags;
    PinnedBuf host_flags;
};

void session_alloc(SessionState& st, int W, int H, int64_t max_seeds, hipStream_t s);
void session_clear(SessionState& st, hipStream_t s);
// d_xy: n device (x, y) pairs; d_values: n device floats or null (zeros).
// Throws value_error (state unchanged) on an out-of-bounds, already-live or
// repeated position, or on overflow.
void session_insert(SessionState& st, const int32_t* d_xy, const float* d_values, int64_t n,
                    hipStream_t s);
// Throws value_error (state unchanged) on a position that is not live or repeats.
void session_remove(SessionState& st, const int32_t* d_xy, int64_t n, DevBuf& tmp, hipStream_t s);
// Removal then insertion as one step: if either fails, the state (bitmap,
// list order, values) is as before the call.
void session_update(SessionState& st, const int32_t* d_rm, int64_t n_rm, const int32_t* d_ins,
                    const float* d_values, int64_t n_ins, DevBuf& tmp, hipStream_t s);

// ---------------------------------------------------------------------------
// Queries on a finalised session (dl_query.hip).  All pointers are device
// pointers; every function runs on stream s and leaves results on the device
// unless it returns a count.
// ---------------------------------------------------------------------------
struct SessionView {
    int W = 0, H = 0, P = 0;
    int64_t n_live = 0;
    const int32_t* list = nullptr;      // live (x, y) pairs, insertion order
    const float* values = nullptr;      // per live seed, insertion order
    const uint32_t* bitmap = nullptr;   // live-seed presence bits (x*H + y)
    int64_t words = 0;
    const uint32_t* sorted = nullptr;   // packed labels x<<16|y in sorted id order
    const int32_t* grid = nullptr;      // (H, W, 3) result of the last finalise
    const int32_t* rows = nullptr;      // (n_tri, 5) triangle rows, sorted ids
    int64_t n_tri = 0;
};

struct QueryWork {
    DevBuf a, b, c, d, e;   // per-query / per-edge scratch
    DevBuf sort_tmp;        // radix-sort temporary storage
    DevBuf scan_tmp;
    PinnedBuf host;
};

// tid of the closed triangle containing each query point (image coordinates),
// max tid on overlaps; -1 where none does or outside the padded canvas.  With
// d_t given, also the lifted in-circumsphere mask.
void query_points(const SessionView& v, const int32_t* d_q, const double* d_t, int64_t n,
                  int32_t* d_tid, uint8_t* d_mask, QueryWork& w, hipStream_t s);
// Distinct undirected edges (a < b, insertion-order ids) in ascending (a, b)
// order, written to edges (2 int32 each, capacity 3 * n_tri); returns their count.
int64_t build_edges(const SessionView& v, DevBuf& edges, QueryWork& w, hipStream_t s);
// |value_a - value_b| * |ab| per edge, 0 below min_length.
void edge_scores(const SessionView& v, const int32_t* d_edges, int64_t n_edges, double min_length,
                 float* d_scores, hipStream_t s);
// Midpoints of the best `count` edges (score descending, index ascending),
// without positions that hold a seed, one per pixel (the better edge), in
// raster order.  Writes (x, y) pairs to d_out (capacity min(count, n_edges))
// and returns their number.
int64_t select_midpoints(const SessionView& v, const int32_t* d_edges, int64_t n_edges,
                         double min_length, int64_t count, float threshold, int32_t* d_out,
                         QueryWork& w, hipStream_t s);
// (H, W, 3) grid -> per-pixel tid (0 outside every triangle), seed id, and
// outside mask.
void split_grid(const int32_t* grid, int64_t n_px, int32_t* tids, int32_t* sids, uint8_t* mask,
                hipStream_t s);

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
