/*
 * delauney_hip.h -- C ABI of libdelauney_hip.so, the MI355X (gfx950) backend of
 * the `delauney` integer Voronoi / Delaunay pipeline.
 *
 * The entry points replace the pybind11 module `_delauney_cuda` of the
 * reference (src/cuda/bindings.cpp) and compute the semantics of its NumPy
 * reference path (src/delauney/reference/), bit for bit.  Each declaration
 * cites the reference interface it stands in for.
 *
 * Conventions
 *   - Every function returns a status: DL_OK, DL_EVALUE (invalid input; the
 *     Python layer raises ValueError) or DL_ERUNTIME (device or allocation
 *     failure; RuntimeError).  dl_last_error() returns the message of the last
 *     failing call on the calling thread.
 *   - Seeds are n (x, y) int32 pairs in any order.  Seed ids are the
 *     (x asc, y asc) order (reference voronoi.py:64-66).
 *   - Host pointers are plain caller-owned C arrays; *_device entry points take
 *     device pointers and a hipStream_t passed as void* (NULL: the context's
 *     own stream).
 *   - border_padding: -1 selects DL_DEFAULT_BORDER_PADDING (the reference's
 *     border_padding=None / AUTO_BORDER_PADDING, binding_helpers.hpp:18-20);
 *     other negative values clamp to 0 (triangulation.py:75-78).
 *   - Canvas sides (image side + 2 * padding) are limited to DL_MAX_CANVAS;
 *     a larger canvas is DL_ERUNTIME, as the reference's failing allocation
 *     of an oversized canvas is (tests/test_cuda_error_checking.py:28-58).
 *     A failing device allocation is DL_ERUNTIME naming hipMalloc, and the
 *     context stays usable.
 *   - Calls on one context are serialised by the caller (one context per
 *     thread); they block until their results are ready and may be made with
 *     the Python GIL released.
 */
#ifndef DELAUNEY_HIP_H
#define DELAUNEY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DL_OK 0
#define DL_EVALUE 1
#define DL_ERUNTIME 2

#define DL_DEFAULT_BORDER_PADDING 16 /* src/delauney/__init__.py:16 */
#define DL_MAX_CANVAS 32000

typedef struct dl_context dl_context;

/* Library version string. */
const char* dl_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* dl_last_error(void);

/* A context owns a HIP stream on `device` and grow-only device workspaces.
 * (The reference has no context; its calls allocate per call, e.g.
 * src/cuda/triangulation.cu:143-365.) */
int dl_context_create(int device, dl_context** out);
void dl_context_destroy(dl_context* ctx);

/* Voronoi().compute(width, height, seeds) -> (H, W, 2) int32 (seed_id, d^2).
 * Replaces PyVoronoi::compute, src/cuda/bindings/py_voronoi.hpp:11-24;
 * semantics of src/delauney/reference/voronoi.py:34-163. */
int dl_voronoi(dl_context* ctx, int width, int height, const int32_t* seeds_xy, int64_t n,
               int32_t* out_hw2);

/* GridTriangulation().compute(voronoi_grid, seeds, border_padding).
 * voronoi_grid: (H, W, 2) int32 host array; only its id channel is read, and
 * only when the resolved padding is 0 (triangulation.py:79-84).  Writes the
 * (H, W) triangle-id raster to out_tid_hw (channel 2 of the reference's grid;
 * channels 0/1 are the caller's grid, triangulation.py:206-209) and the
 * triangle count to *n_tri.  The triangles are kept in the context for
 * dl_copy_triangles.  Replaces PyGridTriangulation::compute,
 * src/cuda/bindings/py_grid_triangulation.hpp:10-24. */
int dl_grid_triangulation(dl_context* ctx, const int32_t* voronoi_grid_hw2, int width,
                          int height, const int32_t* seeds_xy, int64_t n, int border_padding,
                          int32_t* out_tid_hw, int64_t* n_tri);

/* triangulate(width, height, seeds, border_padding) -> (H, W, 3) int32 grid,
 * equal to Voronoi().compute followed by GridTriangulation().compute.
 * Replaces triangulate(), src/cuda/bindings/triangulate.hpp:15-43. */
int dl_triangulate(dl_context* ctx, int width, int height, const int32_t* seeds_xy, int64_t n,
                   int border_padding, int32_t* out_hw3, int64_t* n_tri);

/* Batch of n_images independent triangulations in one pipeline run: every
 * stage (seed ordering, flood fill, detection, assignment) runs once for the
 * whole batch.  Image i is widths[i] x heights[i] with seed_counts[i] seeds;
 * seeds_xy holds the images' seeds one image after another (any order within
 * an image), paddings[i] its border padding (NULL: all -1, the default).  out
 * receives the (H_i, W_i, 3) grids one after another, n_tri[i] image i's
 * triangle count; dl_copy_triangles then returns all rows, image after image,
 * each image's tids starting at 0.  Every image's result equals dl_triangulate
 * on that image alone.  Errors name the failing image.  No single reference
 * counterpart: the reference triangulates one image per call
 * (triangulate.hpp:15-43). */
int dl_triangulate_batch(dl_context* ctx, int n_images, const int32_t* widths,
                         const int32_t* heights, const int64_t* seed_counts,
                         const int32_t* seeds_xy, const int32_t* paddings, int32_t* out,
                         int64_t* n_tri);

/* dl_triangulate_batch on device buffers: d_seeds_xy and d_out are device
 * pointers laid out as above; the size arrays stay on the host. */
int dl_triangulate_batch_device(dl_context* ctx, int n_images, const int32_t* widths,
                                const int32_t* heights, const int64_t* seed_counts,
                                const int32_t* d_seeds_xy, const int32_t* paddings,
                                int32_t* d_out, void* stream, int64_t* n_tri);

/* Incremental sessions (the reference's Delaunay class, bindings.cpp:86-227,
 * delaunay.cuh:23-378), with one change of contract: after any sequence of
 * inserts and removals, finalise gives bit for bit what dl_triangulate gives
 * for the live seed set (the reference's incremental path follows insertion
 * order and drifts from its batch result, test_incremental_deferred.py:333-355).
 * The live seeds stay on the device between calls (presence bitmap + list in
 * insertion order), and so do the last result grid and rows.  A call that
 * fails leaves the session unchanged.  Removal has no reference counterpart.
 *   create:   border_padding < 0 selects the default (delaunay.cu:72).
 *   insert:   ValueError on a position out of bounds, already live or repeated
 *             in the batch, or when max_seeds would be exceeded
 *             (delaunay.cu:191-211).  values: one float per seed (the scalar
 *             field of insert_deferred, delaunay.cu:274-309) or NULL (zeros).
 *   remove:   ValueError on a position that is not live or repeats.
 *   finalise: runs the pipeline on the live seeds (finalise,
 *             delaunay.cu:323-345); with none live the grid is all -1.
 *   copy_result: the last finalise's (H, W, 3) grid and/or rows (either may
 *             be NULL); ValueError while changes are pending.
 *   copy_seeds: live seeds in insertion order (seed_count pairs). */
typedef struct dl_session dl_session;
int dl_session_create(dl_context* ctx, int width, int height, int64_t max_seeds,
                      int border_padding, dl_session** out);
void dl_session_destroy(dl_session* s);
int dl_session_insert(dl_context* ctx, dl_session* s, const int32_t* seeds_xy, const float* values,
                      int64_t n);
int dl_session_remove(dl_context* ctx, dl_session* s, const int32_t* seeds_xy, int64_t n);
/* remove then insert as one step: on any failure the session (live set,
 * insertion order, values) is as before the call.  The serving loop's update
 * request; no reference counterpart (the reference has no removal). */
int dl_session_update(dl_context* ctx, dl_session* s, const int32_t* remove_xy, int64_t n_remove,
                      const int32_t* insert_xy, const float* values, int64_t n_insert);
int dl_session_seed_count(dl_session* s, int64_t* n);
int dl_session_finalise(dl_context* ctx, dl_session* s, int64_t* n_tri);
int dl_session_copy_result(dl_session* s, int32_t* out_hw3, int32_t* out_rows);
int dl_session_copy_seeds(dl_session* s, int32_t* out_xy);
/* Device pointer of the last finalise's (H, W, 3) grid (NULL while pending). */
const int32_t* dl_session_device_grid(dl_session* s);

/* Fill records (no reference counterpart: the reference recomputes, or
 * patches inexactly).  enabled != 0: the session keeps its last flood fill's
 * trajectory on the device (per tile and launch, the tile's labels after
 * every launch that changed it), and the next finalise recomputes only the
 * tiles whose neighbourhood the inserted and removed seeds reach, launch by
 * launch, taking every other tile from the record; the result is the same as
 * recomputing from scratch, bit for bit.  The record costs device memory
 * (a pool of tile states, grown on demand; dl_session_device_bytes counts
 * it); enabled = 0 frees it.  Applies to canvases of the default tile shape
 * with sides up to about 23,170 (others always recompute).
 * dl_session_fill_record_stats: whether a record is held, its entries, and
 * the tile computations of the last finalise's fill. */
int dl_session_set_fill_record(dl_session* s, int enabled);
int dl_session_fill_record_stats(dl_session* s, int* valid, int64_t* entries,
                                 int64_t* tiles_computed);

/* The scalar field: whether any insert carried values (edge_scores and
 * select_midpoints need it, delaunay_scalar_field.cu:218-220), and the values
 * of the live seeds in insertion order (get_values, delaunay_output.cu:221-224). */
int dl_session_has_values(dl_session* s, int* out);
int dl_session_copy_values(dl_session* s, float* out);

/* Queries.  Each finalises first when changes are pending.
 *   locate: tid of the triangle containing each of n (x, y) points in image
 *     coordinates -- the max tid over closed containing triangles, which is the
 *     grid's tid channel inside the image -- or -1; points beyond the padded
 *     canvas get -1 (Delaunay::locate, delaunay_query.cu:113-133, 196-227).
 *   in_circumsphere: also mask[i] = 1 iff dx^2 + dy^2 + t[i]^2 < R^2 about the
 *     containing triangle's circumcircle, in the reference's double operation
 *     order on padded-canvas coordinates (delaunay_query.cu:33-100, 135-194). */
int dl_session_locate(dl_context* ctx, dl_session* s, const int32_t* points_xy, int64_t n,
                      int32_t* out_tid);
int dl_session_in_circumsphere(dl_context* ctx, dl_session* s, const int32_t* points_xy,
                               const double* t, int64_t n, uint8_t* out_mask, int32_t* out_tid);

/* Edges and the scalar field over them (delaunay_scalar_field.cu:35-312).
 *   edges: builds the distinct undirected edges, (a, b) with a < b in
 *     insertion-order ids, ascending; *n_edges receives their count.
 *   copy_edges: the (n_edges, 2) int32 list of the last dl_session_edges.
 *   edge_scores: n_edges floats |v_a - v_b| * |ab| (0 when |ab| < min_length);
 *     RuntimeError when no insert carried values.
 *   select_midpoints: midpoints of the best `count` edges (score descending,
 *     edge index ascending; only scores > threshold and > 0), positions holding
 *     a seed dropped, one per pixel (the better edge), in raster order;
 *     out_xy holds out_cap >= min(count, n_edges) pairs, *n_out receives the
 *     number written. */
int dl_session_edges(dl_context* ctx, dl_session* s, int64_t* n_edges);
int dl_session_copy_edges(dl_session* s, int32_t* out);
int dl_session_edge_scores(dl_context* ctx, dl_session* s, double min_length, float* out);
int dl_session_select_midpoints(dl_context* ctx, dl_session* s, double min_length, int64_t count,
                                float threshold, int32_t* out_xy, int64_t out_cap, int64_t* n_out);

/* finalise with the per-pixel result left on the device
 * (Delaunay::finalise_device, delaunay.cu:347-360, delaunay_output.cu:119-163):
 * three H*W arrays owned by the session -- tid (0 where no triangle contains
 * the pixel), sorted seed id, and the outside mask (1 there) -- valid until the
 * session's next insert, remove, finalise_device or destroy. */
int dl_session_finalise_device(dl_context* ctx, dl_session* s, int64_t* n_tri,
                               int32_t** pixel_tids, int32_t** pixel_seed_ids,
                               uint8_t** outside_mask);
/* Blocking copy of device memory to the host (the views' to_host,
 * py_device_array_view.hpp:49-65). */
int dl_device_to_host(void* dst, const void* src, size_t bytes);

/* Page-locked host memory (hipHostMalloc) for results that are copied back
 * at full link speed; the serving loop pools it.  No reference counterpart. */
int dl_host_alloc(size_t bytes, void** out);
void dl_host_free(void* p);

/* One image of dl_run_batch: seeds and an output buffer (device pointers, or
 * host pointers when host != 0: the library stages them), or a session (its
 * live seeds are used and its result stays in the session). */
typedef struct dl_image {
    int width, height, border_padding;  /* ignored for sessions */
    int host;                           /* seeds_xy / out are host pointers */
    int64_t n_seeds;                    /* ignored for sessions */
    const int32_t* seeds_xy;            /* n_seeds (x, y) pairs */
    int32_t* out;                       /* (height, width, 3) int32 */
    dl_session* session;                /* non-NULL: finalise this session */
    int64_t n_tri;                      /* out: the image's triangle count */
    int64_t tri_offset;                 /* out: its first row in dl_copy_triangles */
} dl_image;

/* Runs the pipeline once for a mixed batch of fresh images and sessions (the
 * serving loop's unit of work); every image's result equals its single-image
 * result.  Sessions with no live seed get an all -1 grid.  With timing on
 * (dl_set_timing), dl_get_timings' output stage [5] is the copy of host
 * images' grids back to the host.  No reference counterpart. */
int dl_run_batch(dl_context* ctx, int n, dl_image* images, void* stream);

/* Device memory accounting for admission control (the serving loop); no
 * reference counterpart.
 *   context_device_bytes: bytes of device workspace the context holds now
 *     (grow-only between trims).
 *   context_trim: releases every device workspace of the context (the next
 *     call allocates again); results of earlier calls are no longer
 *     available from it (dl_copy_triangles, dl_device_triangles).
 *   session_device_bytes: bytes of device memory the session holds.
 *   device_memory: hipMemGetInfo of a device. */
int dl_context_device_bytes(dl_context* ctx, size_t* bytes);
int dl_context_trim(dl_context* ctx);
int dl_session_device_bytes(dl_session* s, size_t* bytes);
int dl_device_memory(int device, size_t* free_bytes, size_t* total_bytes);

/* Rows (x, y, id_a, id_b, id_c) in tid order of the last triangulation on this
 * context: the reference's triangle_map values (triangulation.py:157;
 * binding_helpers.hpp:61-87).  out_rows holds n_tri * 5 int32. */
int dl_copy_triangles(dl_context* ctx, int32_t* out_rows);

/* The padded Voronoi canvas of the last dl_grid_triangulation /
 * dl_triangulate call at padding > 0, as (H+2P, W+2P, 2) int32 (seed_id, d^2).
 * Replaces the padded_voronoi_out of compute_debug,
 * src/cuda/bindings/py_grid_triangulation.hpp:32-56.  *out_w / *out_h receive
 * the canvas size; pass out == NULL to query it. */
int dl_copy_padded_voronoi(dl_context* ctx, int32_t* out, int* out_w, int* out_h);

/* Stage timings of the last call in ms: [0] seeds, [1] flood fill, [2] detect,
 * [3] dedup/rank, [4] assign, [5] output, [6] total, [7] flood-fill launches,
 * [8] flood-fill tile visits.  Mirrors the *_timed dictionaries
 * (py_grid_triangulation.hpp:39-46, triangulation.cuh:29-34). */
int dl_get_timings(dl_context* ctx, double* out9);

/* Flood-fill statistics of the last call: launches issued, launches up to
 * and including the first that changed nothing (later ones return at once),
 * tile visits (counted while timing or kernel profiling is on, else 0), and
 * iterations per launch; and, over the context's life, the predicted fills
 * (see dl_set_fill_prediction) that fell short and were continued.
 * bfs_iters of the session *_timed dictionaries is converged *
 * iters_per_launch, the iterations the fill ran (the reference counts its
 * BFS passes the same way, delaunay_voronoi.cu:57-108,
 * py_delaunay.hpp:268-277).  Any pointer may be NULL. */
int dl_get_fill_stats(dl_context* ctx, int64_t* launches, int64_t* converged,
                      int64_t* tile_visits, int* iters_per_launch, int64_t* shortfalls);

/* Enables event timing of stages for subsequent calls (adds synchronisation). */
int dl_set_timing(dl_context* ctx, int enabled);

/* Flood-fill tile configuration: waves per tile, rows per thread, iterations
 * per launch.  (0, 0, 0) restores automatic selection.  No reference
 * counterpart (the reference uses 16x16 blocks, voronoi.cu:95-113). */
int dl_set_bfs_config(dl_context* ctx, int waves, int rows, int iters);

/* The configuration the last flood fill on this context ran with. */
int dl_get_bfs_config(dl_context* ctx, int* waves, int* rows, int* iters);

/* Autotuning of the automatic selection (default on).  Every tile shape gives
 * the same labels, so the shape is chosen per canvas class (device, log2 of
 * the fill's cells, log4 of cells per seed): the first fill of a class with at
 * least 2^24 cells times the candidate shapes on its own input and the
 * fastest is kept for the process; smaller fills use 8,8,8.  dl_autotune_reset
 * forgets every class.  No reference counterpart. */
int dl_set_autotune(dl_context* ctx, int enabled);
int dl_autotune_reset(void);

/* Launch-count prediction (default on).  A padded fill of a canvas class
 * (as above) and shape seen before on this context is issued as the launch
 * count that fill needed plus a margin, without waiting for the host; the
 * pipeline checks convergence with its final status read and, on a shortfall,
 * continues the fill and recomputes detection and assignment.  Results are
 * the same either way; off trades the saved host wait for a steady latency
 * when the inputs of a class vary (the serving loop turns it off).  No
 * reference counterpart. */
int dl_set_fill_prediction(dl_context* ctx, int enabled);

/* hipGraph replay of the pipeline: 0 off, 1 on, 2 auto (the default: on for
 * pipelines of fewer than 2^22 canvas cells).  Everything a triangulation
 * enqueues before its final status read (seed ordering, the predicted padded
 * fill, detection, registration, assignment and output, and the status
 * copies) is captured once per call signature: the images' sizes, seed
 * counts, padding, seed and output addresses, the stream and the fill's shape
 * and launch count.  A call matching the last predicted run's signature is
 * captured; later matching calls replay the graph, with the host planning and
 * the ~30 launches replaced by one graph launch.  Any buffer reallocation in
 * the process drops the graph.  Results are identical with and without.
 * Timing and kernel profiling run uncaptured.  A replay saves the host's
 * issue work, which bounds small pipelines, but on ROCm 7 costs about 0.5 us
 * of GPU time per kernel node more than host-issued launches, which shows on
 * large ones; hence the size threshold.  Other values: ValueError.
 * No reference counterpart. */
int dl_set_graphs(dl_context* ctx, int enabled);
/* Graphs captured and replayed on this context so far. */
int dl_get_graph_stats(dl_context* ctx, int64_t* captures, int64_t* replays);

/* Device-resident triangulate: seeds, output and triangles stay on the device.
 * d_seeds_xy: n (x, y) int32 pairs; d_out_hw3: (H, W, 3) int32.  The triangle
 * rows stay in the context (dl_device_triangles).  No reference counterpart;
 * it is the device-resident form of triangulate() for callers whose data are
 * already in HBM. */
int dl_triangulate_device(dl_context* ctx, int width, int height, const int32_t* d_seeds_xy,
                          int64_t n, int border_padding, int32_t* d_out_hw3, void* stream,
                          int64_t* n_tri);

/* dl_triangulate_device that allocates nothing: no hipMalloc, hipHostMalloc,
 * hipFree, or event or stream creation, for callers that cannot allocate
 * (tight serving loops, code that must not touch the allocator).  No
 * counterpart in the reference.  Its workspaces are carved from one static
 * device array of DL_NOALLOC_SCRATCH_BYTES per GPU (held from the library's
 * load on that GPU until the process ends; its address is looked up once per
 * device) and a static host block.  Limit: images for which
 * 10 * (W + 2P)(H + 2P) + 13 * W * H + 740 * n + 2 MiB <= DL_NOALLOC_SCRATCH_BYTES
 * (e.g. 1024 x 1024 at padding 16 with up to 60,000 seeds); larger inputs are
 * DL_EVALUE before anything runs.  The grid goes to d_out_hw3 and the
 * (n_tri, 5) rows to d_rows (capacity rows_cap rows; NULL: not copied;
 * DL_EVALUE if too small, with *n_tri set), both device memory.  The call
 * waits on the stream before it returns.  Every call on a GPU shares that
 * GPU's scratch: calls that overlap on one GPU (another thread, another
 * stream) race.  Results are those of dl_triangulate_device. */
#define DL_NOALLOC_SCRATCH_BYTES (64u << 20)
int dl_triangulate_device_noalloc(dl_context* ctx, int width, int height,
                                  const int32_t* d_seeds_xy, int64_t n, int border_padding,
                                  int32_t* d_out_hw3, int32_t* d_rows, int64_t rows_cap,
                                  void* stream, int64_t* n_tri);

/* Device pointer to the (n_tri, 5) rows of the last triangulation. */
const int32_t* dl_device_triangles(dl_context* ctx);

/* Voronoi flood fill through the one-pass-per-launch kernel instead of the
 * tiled one (cross-checking only).  Same output as dl_voronoi. */
int dl_voronoi_simple(dl_context* ctx, int width, int height, const int32_t* seeds_xy,
                      int64_t n, int32_t* out_hw2, int64_t* iters);

/* Per-kernel timing with HIP events on the stream each kernel runs on.
 * While enabled, every call brackets its kernels with events and accumulates
 * counts and milliseconds per kind: [0] flood-fill tile kernel, [1] seed
 * ordering, [2] triangle detection, [3] registration (first-registration
 * marking, prefix scan, row emission), [4] triangle binning, [5] per-tile
 * pixel assignment fused with the output
 * assembly, [6] initialisation fills.  counts and ms hold 7 entries each.
 * No reference counterpart (the reference's PhaseTimer, phase_timer.cuh:17-62,
 * times phases, not kernels). */
int dl_set_kernel_profiling(dl_context* ctx, int enabled);
int dl_get_kernel_stats(dl_context* ctx, int64_t* counts, double* ms);
int dl_reset_kernel_stats(dl_context* ctx);

/* Forces the second (unpadded) flood fill in dl_triangulate even when the
 * padded fill shows that the image interior never received an update from the
 * padding (cross-checking only). */
int dl_set_force_unpadded(dl_context* ctx, int enabled);

/* Replays of a session's fill record (dl_session_set_fill_record) start from
 * the context's label plane as the previous finalise left it when nothing
 * has written it since, seeding and writing back only the tiles the replay
 * touches.  enabled = 0 makes every replay re-initialise the whole canvas
 * (cross-checking only; the default is 1).  No counterpart in the reference. */
int dl_set_replay_plane_reuse(dl_context* ctx, int enabled);

/* Fills of this context, over its life: recorded (dl_session_set_fill_record,
 * from scratch), replayed from a record, and replayed from the retained
 * plane (a subset of the replayed).  Any pointer may be NULL.  No counterpart
 * in the reference. */
int dl_get_record_stats(dl_context* ctx, int64_t* recorded, int64_t* replayed, int64_t* retained);

/* Flood fill of the width x height canvas (seeds as given, no padding) that
 * also reports in *escaped whether any cell of the box
 * [margin, width-margin) x [margin, height-margin) ever received an update
 * from a cell outside the box.  The pipeline runs the same watch on its padded
 * canvas with the box = the image; this entry point exposes it on arbitrary
 * boxes (cross-checking only; no reference counterpart).  The flag may be set
 * without a crossing (tiles recompute halo cells) but is never clear when one
 * happened.  out_hw2 may be null. */
int dl_debug_fill_watch(dl_context* ctx, int width, int height, const int32_t* seeds_xy,
                        int64_t n, int margin, int32_t* out_hw2, int* escaped);

/* Test hook: fills with at least min_tiles tiles (default 8192; negative
 * restores it) run their launches after the first over a list of the active
 * tiles; 0 forces that path on every fill.  Results are the same either way. */
int dl_set_fill_list_threshold(dl_context* ctx, int min_tiles);

/* Test hook: enabled = 0 keeps every flood fill on the compare-and-select
 * key update that canvases with (W-1)^2 + (H-1)^2 > 32767^2 use; 1 (the
 * default) lets smaller canvases use the min-based update.  Results are the
 * same either way.  No reference counterpart. */
int dl_set_fill_min_encoding(dl_context* ctx, int enabled);

/* Test hook: enabled = 0 keeps every flood fill's padding watch on its
 * per-row form; 1 (the default) lets fills whose watched interiors begin and
 * end on the tile kernel's wave-row boundaries (the default tile shape, the
 * min-based update, P and H - P both multiples of 8) test one row per y pass
 * instead.  Labels are the same either way, and so is the watch's flag.
 * dl_get_fill_aligned_watch reports which form the context's last fill used.
 * No reference counterpart. */
int dl_set_fill_aligned_watch(dl_context* ctx, int enabled);
int dl_get_fill_aligned_watch(dl_context* ctx, int* aligned);

/* Cross-checking only: makes speculative sizes too small on purpose, so that
 * every call takes the exact-size rerun.  enabled: 0 off; 1 every speculative
 * size of detection and assignment; 2 only the capacity of the registration
 * buckets (the rerun then registers through the hash table).  Results are
 * unchanged.  No reference counterpart. */
int dl_set_tiny_capacities(dl_context* ctx, int enabled);

#ifdef __cplusplus
}
#endif

#endif /* DELAUNEY_HIP_H */
