// Seed normalisation, triangle extraction, pixel assignment and output
// assembly on gfx950.
//
// Reference semantics (src/delauney/reference/triangulation.py):
//  * seed ids are the (x asc, y asc) order (voronoi.py:64-72);
//  * a 2x2 block with four distinct regions blocks two triplets globally, chosen
//    by the cyclic order about the corner centroid (triangulation.py:109-144);
//  * four L-shape scans, each in row-major order, register pairwise-distinct,
//    unblocked, unseen triplets; the tid is the order of first registration and
//    the first registration fixes (x, y) and the vertex order (:146-183);
//  * a pixel takes the largest tid over triangles containing its integer point
//    in the closed sense, else -1 (:193-204, 226-253).
//
// Design: registration order is reproduced without sorting.  Every candidate
// (scan s, block (rx, ry)) has a unique key s*NB + ry*(Wp-1) + rx, which is its
// position in the reference's scan order.  A hash table keeps the minimum key
// of each distinct triplet; a candidate whose key is its triplet's minimum (and
// whose triplet is not blocked) sets bit `key` of a bitmap, and the tid of that
// registration is the number of set bits below it -- a popcount prefix scan.
#include "dl_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>

namespace dl {

// ===========================================================================
// Popcount prefix scan
// ===========================================================================
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanBlock = kScanThreads * kScanItems;

template <bool POPC>
__device__ __forceinline__ uint32_t scan_item(uint32_t v)
{
    return POPC ? static_cast<uint32_t>(__popc(v)) : v;
}

template <bool POPC>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(
    const uint32_t* __restrict__ w, int64_t n, uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t red[kScanThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = base + j * kScanThreads + threadIdx.x;
        if (i < n) s += scan_item<POPC>(w[i]);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < kScanThreads / 64; ++k) t += red[k];
        bsum[blockIdx.x] = t;
    }
}

// Exclusive scan of bsum[0..nb) in place with one 1024-thread block; writes the
// grand total to *total (when given).
__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t* __restrict__ bsum, int64_t nb,
                                                           uint32_t* __restrict__ total)
{
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t incl = part[threadIdx.x];
        if (i < nb) bsum[i] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// The scan of each block's kScanBlock items, offset by the block's prefix.
// SELF: the prefix is the sum of the unscanned partials bsum[0 .. block)
// (none when bsum is null: a one-block scan), and the last block writes the
// total to out[n]; otherwise bsum holds the scanned prefixes.  INTER writes
// (prefix, raw word) pairs to inter instead of prefixes to out (no total).
template <bool POPC, bool INTER, bool SELF>
__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(
    const uint32_t* __restrict__ w, int64_t n, const uint32_t* __restrict__ bsum,
    uint32_t* __restrict__ out, uint2* __restrict__ inter)
{
    __shared__ uint32_t part[kScanThreads];
    __shared__ uint32_t red[kScanThreads / 64];
    uint32_t prefix;
    if (SELF) {
        uint32_t pf = 0;
        if (bsum)
            for (int64_t k = threadIdx.x; k < blockIdx.x; k += kScanThreads) pf += bsum[k];
        for (int off = 32; off > 0; off >>= 1) pf += __shfl_down(pf, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pf;
        __syncthreads();
        prefix = 0;
#pragma unroll
        for (int k = 0; k < kScanThreads / 64; ++k) prefix += red[k];
    } else {
        prefix = bsum[blockIdx.x];
    }
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock +
                         static_cast<int64_t>(threadIdx.x) * kScanItems;
    // A thread's kScanItems consecutive items move as 16-byte vectors when
    // they are all inside and aligned (every thread but the tail's).
    const bool full = base + kScanItems <= n;
    uint32_t raw[kScanItems], c[kScanItems];
    if (full && (reinterpret_cast<uintptr_t>(w + base) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(w + base)[q];
            raw[4 * q] = v.x;
            raw[4 * q + 1] = v.y;
            raw[4 * q + 2] = v.z;
            raw[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) raw[j] = base + j < n ? w[base + j] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        c[j] = scan_item<POPC>(raw[j]);
        s += c[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = prefix + part[threadIdx.x] - s;
    uint32_t pre[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        pre[j] = run;
        run += c[j];
    }
    if (INTER) {
        if (full && (reinterpret_cast<uintptr_t>(inter + base) & 15) == 0) {
#pragma unroll
            for (int q = 0; q < kScanItems / 2; ++q)
                reinterpret_cast<uint4*>(inter + base)[q] =
                    make_uint4(pre[2 * q], raw[2 * q], pre[2 * q + 1], raw[2 * q + 1]);
        } else {
#pragma unroll
            for (int j = 0; j < kScanItems; ++j)
                if (base + j < n) inter[base + j] = make_uint2(pre[j], raw[j]);
        }
    } else {
        if (full && (reinterpret_cast<uintptr_t>(out + base) & 15) == 0) {
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q)
                reinterpret_cast<uint4*>(out + base)[q] =
                    make_uint4(pre[4 * q], pre[4 * q + 1], pre[4 * q + 2], pre[4 * q + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < kScanItems; ++j)
                if (base + j < n) out[base + j] = pre[j];
        }
    }
    if (!INTER && SELF && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1)
        out[n] = run;
}

}  // namespace

// Up to this many partial sums, each block of the down pass adds up the ones
// before it itself (no separate pass over the partials).
constexpr int64_t kScanSelfMax = 2048;

template <bool POPC, bool INTER>
static void scan_impl(const uint32_t* words, int64_t n, uint32_t* out, uint2* inter, DevBuf& tmp,
                      hipStream_t s)
{
    const int64_t nb = std::max<int64_t>(1, (n + kScanBlock - 1) / kScanBlock);
    if (nb == 1) {   // one block: no partial sums at all
        scan_down_kernel<POPC, INTER, true><<<1, kScanThreads, 0, s>>>(words, n, nullptr, out, inter);
        DL_LAUNCH_CHECK();
        return;
    }
    tmp.reserve(sizeof(uint32_t) * static_cast<size_t>(nb));
    uint32_t* bsum = tmp.as<uint32_t>();
    scan_reduce_kernel<POPC><<<static_cast<unsigned>(nb), kScanThreads, 0, s>>>(words, n, bsum);
    DL_LAUNCH_CHECK();
    if (nb <= kScanSelfMax) {
        scan_down_kernel<POPC, INTER, true>
            <<<static_cast<unsigned>(nb), kScanThreads, 0, s>>>(words, n, bsum, out, inter);
    } else {
        scan_blocks_kernel<<<1, 1024, 0, s>>>(bsum, nb, INTER ? nullptr : out + n);
        DL_LAUNCH_CHECK();
        scan_down_kernel<POPC, INTER, false>
            <<<static_cast<unsigned>(nb), kScanThreads, 0, s>>>(words, n, bsum, out, inter);
    }
    DL_LAUNCH_CHECK();
}

void popcount_scan(const uint32_t* words, int64_t n, uint32_t* out, DevBuf& tmp, hipStream_t s)
{
    scan_impl<true, false>(words, n, out, nullptr, tmp, s);
}

void popcount_scan_pairs(const uint32_t* words, int64_t n, uint2* pairs, DevBuf& tmp, hipStream_t s)
{
    scan_impl<true, true>(words, n, nullptr, pairs, tmp, s);
}

void exclusive_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, DevBuf& tmp, hipStream_t s)
{
    scan_impl<false, false>(in, n, out, nullptr, tmp, s);
}


// ===========================================================================
// Batch plan
// ===========================================================================
namespace {

// Detection thread blocks cover kDetChunk consecutive 2x2 blocks in each of
// kDetRows rows of blocks; each thread owns kDetItems of them per row, strided
// by the thread count so that a wave reads consecutive cells.  Thread blocks
// are numbered strip-major (a strip is kDetChunk columns wide, top to bottom),
// so the candidate list runs through the canvas in x order, as seed ids do.
constexpr int kDetThreads = 256;
constexpr int kDetItems = 1;
constexpr int kDetChunk = kDetThreads * kDetItems;
constexpr int kDetRows = 32;
static_assert(kDetItems * kDetRows <= 32, "one mask bit per 2x2 block of a thread");
static_assert(kDetChunk % 32 == 0, "a detection strip is whole registration bitmap words");
constexpr int kTileSide = 64;   // assignment tile side in pixels

// Largest i in [0, n) with off[i] <= g (off is non-decreasing, off[0] = 0).
// Empty segments (off[i] == off[i + 1]) are skipped by construction.
__device__ __forceinline__ int find_seg(const int64_t* __restrict__ off, int n, int64_t g)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

int bits_for(int64_t v)   // bits needed to hold values in [0, v]
{
    int b = 1;
    while (b < 62 && (v >> b) != 0) ++b;
    return b;
}

}  // namespace

void plan_batch(BatchPlan& plan)
{
    const size_t n = plan.img.size();
    plan.seed_seg.assign(n + 1, 0);
    plan.det_seg.assign(n + 1, 0);
    plan.tile_seg.assign(n + 1, 0);
    int64_t seed = 0, words = 0, pix = 0, cells = 0, det = 0, keyw = 0, tiles = 0, maxn = 1;
    for (size_t i = 0; i < n; ++i) {
        ImageDesc& d = plan.img[i];
        d.Wc = d.W + 2 * d.P;
        d.Hc = d.H + 2 * d.P;
        d.seed_off = seed;
        seed += d.n;
        d.bm_off = words;
        words += (static_cast<int64_t>(d.W) * d.H + 31) / 32;
        d.pix_off = pix;
        pix += static_cast<int64_t>(d.W) * d.H;
        d.cv_off = cells;
        cells += static_cast<int64_t>(d.Wc) * d.Hc;
        const int bw = d.Wc - 1, bh = d.Hc - 1;
        d.nb = (bw > 0 && bh > 0) ? static_cast<int64_t>(bw) * bh : 0;
        d.bx_per_row = d.nb ? (bw + kDetChunk - 1) / kDetChunk : 0;
        d.det_off = det;
        det += d.nb ? static_cast<int64_t>(d.bx_per_row) * ((bh + kDetRows - 1) / kDetRows) : 0;
        // registration bitmap: 4 scans x bh rows x whole detection strips
        // (8 words of 32 blocks each per strip and row; see reg_word)
        d.key_off = keyw;
        keyw += d.nb ? 4 * static_cast<int64_t>(bh) * d.bx_per_row * (kDetChunk / 32) : 0;
        d.tiles_x = (d.W + kTileSide - 1) / kTileSide;
        d.tile_off = tiles;
        tiles += static_cast<int64_t>(d.tiles_x) * ((d.H + kTileSide - 1) / kTileSide);
        maxn = std::max<int64_t>(maxn, d.n);
        plan.seed_seg[i + 1] = seed;
        plan.det_seg[i + 1] = det;
        plan.tile_seg[i + 1] = tiles;
    }
    plan.n_seeds = seed;
    plan.n_words = words;
    plan.n_pix = pix;
    plan.n_cells = cells;
    plan.n_det_blocks = det;
    plan.n_key_words = keyw;
    plan.n_tiles = tiles;
    plan.max_seed_bits = bits_for(maxn - 1);
}

// Device layout of work.segs: seed_seg, det_seg, tile_seg, tri_base, each
// n_img + 1 int64.
namespace {
inline int64_t* seg_ptr(const TriWork& w, int which, size_t n_img)
{
    return reinterpret_cast<int64_t*>(static_cast<char*>(w.desc.ptr) + w.segs_off) +
           static_cast<size_t>(which) * (n_img + 1);
}

// work.desc holds the descriptors, then (from a 256-byte boundary) the offset
// tables, so one copy uploads a plan.
size_t plan_bytes(size_t n_img, size_t* segs_off)
{
    *segs_off = (sizeof(ImageDesc) * n_img + 255) / 256 * 256;
    return *segs_off + sizeof(int64_t) * 4 * (n_img + 1);
}
}  // namespace

void upload_desc(const BatchPlan& plan, TriWork& work, hipStream_t s)
{
    const size_t n = plan.img.size();
    const size_t db = sizeof(ImageDesc) * n;
    work.desc.reserve(plan_bytes(n, &work.segs_off));
    work.host_stage.reserve(db);
    stage_h2d(work.host_stage, 0, work.desc.ptr, plan.img.data(), db, s);
}

void upload_plan(const BatchPlan& plan, TriWork& work, hipStream_t s)
{
    // Descriptors and the three host-made offset tables in one copy, from a
    // staging buffer of their own (upload_desc may rewrite its buffer before
    // this copy has run).
    const size_t n = plan.img.size();
    const size_t db = sizeof(ImageDesc) * n;
    const size_t sb = sizeof(int64_t) * (n + 1);
    work.desc.reserve(plan_bytes(n, &work.segs_off));
    const size_t up = work.segs_off + 3 * sb;
    std::vector<char> h(up, 0);
    std::memcpy(h.data(), plan.img.data(), db);
    const std::vector<int64_t>* tabs[3] = {&plan.seed_seg, &plan.det_seg, &plan.tile_seg};
    for (int k = 0; k < 3; ++k) std::memcpy(h.data() + work.segs_off + k * sb, tabs[k]->data(), sb);
    work.host_plan.reserve(up);
    stage_h2d(work.host_plan, 0, work.desc.ptr, h.data(), up, s);
}

// ===========================================================================
// Seeds
// ===========================================================================
namespace {

// Status flag bits shared with the host: per image (seeds) and batch-wide.
constexpr uint32_t kFlagXRange = 1u;
constexpr uint32_t kFlagYRange = 2u;
constexpr uint32_t kFlagDuplicate = 4u;
constexpr uint32_t kFlagTableOverflow = 16u;
constexpr uint32_t kFlagCollision = 32u;
constexpr uint32_t kFlagBadId = 64u;
constexpr uint32_t kFlagOverflow = 128u;   // a speculative capacity was too small
constexpr uint32_t kFlagBinOverflow = 256u;   // a tile's fixed bin slice was too small

// Presence bit of (x, y) in its image's bitmap is x*H + y: bit order is the id
// order.
__global__ void seeds_mark_kernel(int64_t n_total, const int64_t* __restrict__ seed_seg, int n_img,
                                  const ImageDesc* __restrict__ D, uint32_t* __restrict__ bm,
                                  uint32_t* __restrict__ flags)
{
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const int img = n_img > 1 ? find_seg(seed_seg, n_img, g) : 0;
    const ImageDesc& d = D[img];
    const int32_t* p = d.xy + 2 * (g - d.seed_off);
    const int x = p[0], y = p[1];
    if (x < 0 || x >= d.W) { atomicOr(&flags[img], kFlagXRange); return; }
    if (y < 0 || y >= d.H) { atomicOr(&flags[img], kFlagYRange); return; }
    const uint64_t bit = static_cast<uint64_t>(x) * d.H + y;
    const uint32_t m = 1u << (bit & 31);
    const uint32_t old = atomicOr(&bm[d.bm_off + static_cast<int64_t>(bit >> 5)], m);
    if (old & m) atomicOr(&flags[img], kFlagDuplicate);
}

__global__ void seeds_rank_kernel(int64_t n_total, const int64_t* __restrict__ seed_seg, int n_img,
                                  const ImageDesc* __restrict__ D, const uint2* __restrict__ idx,
                                  uint32_t* __restrict__ sorted)
{
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const int img = n_img > 1 ? find_seg(seed_seg, n_img, g) : 0;
    const ImageDesc& d = D[img];
    const int32_t* p = d.xy + 2 * (g - d.seed_off);
    const int x = p[0], y = p[1];
    if (x < 0 || x >= d.W || y < 0 || y >= d.H) return;   // flagged by seeds_mark_kernel
    const uint64_t bit = static_cast<uint64_t>(x) * d.H + y;
    const int64_t w = d.bm_off + static_cast<int64_t>(bit >> 5);
    const uint2 e = idx[w];   // (prefix, word)
    const uint32_t below = e.y & ((1u << (bit & 31)) - 1u);
    // rank among the image's set bits (< n even when an earlier image holds
    // duplicates, which are flagged and leave fewer bits than seeds)
    const uint32_t rank = e.x - idx[d.bm_off].x + static_cast<uint32_t>(__popc(below));
    sorted[d.seed_off + rank] = pack_label(x, y);
}

// Seed cells of every canvas: the image label shifted by (P, P).
__global__ void canvas_seed_kernel(const uint32_t* __restrict__ sorted, int64_t n_total,
                                   const int64_t* __restrict__ seed_seg, int n_img,
                                   const ImageDesc* __restrict__ D, uint32_t* __restrict__ a)
{
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const int img = n_img > 1 ? find_seg(seed_seg, n_img, g) : 0;
    const ImageDesc& d = D[img];
    const uint32_t l0 = sorted[g];
    // Entries of a seed set that failed validation (checked by the host
    // later) may be stale: keep the write inside the canvas.
    if ((l0 >> 16) >= static_cast<uint32_t>(d.W) || (l0 & 0xFFFFu) >= static_cast<uint32_t>(d.H))
        return;
    const uint32_t l = l0 + ((static_cast<uint32_t>(d.P) << 16) | static_cast<uint32_t>(d.P));
    const int64_t c = d.cv_off + static_cast<int64_t>(l & 0xFFFFu) * d.Wc + (l >> 16);
    a[c] = l;
}

}  // namespace

void check_seed_flags(const BatchPlan& plan, const uint32_t* hf)
{
    const int n_img = static_cast<int>(plan.img.size());
    for (int i = 0; i < n_img; ++i) {
        const uint32_t f = hf[i];
        if (!f) continue;
        const std::string who = n_img > 1 ? "image " + std::to_string(i) + ": " : "";
        if (f & kFlagXRange) throw value_error(who + "seed x coordinate out of bounds");
        if (f & kFlagYRange) throw value_error(who + "seed y coordinate out of bounds");
        if (f & kFlagDuplicate) throw value_error(who + "duplicate seed positions are not allowed");
    }
}

void normalise_seeds(const BatchPlan& plan, uint32_t* d_sorted, TriWork& work, hipStream_t s,
                     SeedCheck check)
{
    const int n_img = static_cast<int>(plan.img.size());
    const int64_t nw = plan.n_words;
    work.seed_bm.reserve(sizeof(uint32_t) * static_cast<size_t>(std::max<int64_t>(nw, 1)));
    work.flags.reserve(sizeof(uint32_t) * (n_img + 8));
    work.host_flags.reserve(sizeof(uint32_t) * (n_img + 8));
    uint32_t* bm = work.seed_bm.as<uint32_t>();
    uint32_t* flags = work.flags.as<uint32_t>();
    const ImageDesc* D = work.desc.as<ImageDesc>();
    const int64_t* sseg = seg_ptr(work, 0, n_img);
    const unsigned blocks = static_cast<unsigned>((plan.n_seeds + 255) / 256);
    {
        ProfScope ps(kKSeeds, s);
        FillSegs fs{};
        fs.seg[0] = FillSeg{bm, static_cast<size_t>(nw), 0u};
        fs.seg[1] = FillSeg{flags, static_cast<size_t>(n_img + 8), 0u};
        fs.k = 2;
        // and the assignment's bin counts (see assign_tiled's counts_zeroed)
        if (plan.n_tiles > 0) {
            const size_t nb = static_cast<size_t>(2 * plan.n_tiles + 1);
            work.bin_counts.reserve(sizeof(uint32_t) * nb);
            fs.seg[fs.k++] = FillSeg{work.bin_counts.as<uint32_t>(), nb, 0u};
        }
        fill_u32_multi(fs, s);
        seeds_mark_kernel<<<blocks, 256, 0, s>>>(plan.n_seeds, sseg, n_img, D, bm,
                                                 flags + kSeedFlagBase);
        DL_LAUNCH_CHECK();
    }
    if (check != SeedCheck::kWithStatus) {
        // A buffer of its own: the final status read reuses host_flags.
        work.host_seed_flags.reserve(sizeof(uint32_t) * static_cast<size_t>(std::max(n_img, 1)));
        DL_HIP(hipMemcpyAsync(work.host_seed_flags.ptr, flags + kSeedFlagBase,
                              sizeof(uint32_t) * n_img, hipMemcpyDeviceToHost, s));
    }
    if (check == SeedCheck::kNow) {
        DL_HIP(hipStreamSynchronize(s));
        check_seed_flags(plan, work.host_seed_flags.as<uint32_t>());
    }
    ProfScope ps(kKSeeds, s);
    work.seed_idx.reserve(sizeof(uint2) * static_cast<size_t>(std::max<int64_t>(nw, 1)));
    popcount_scan_pairs(bm, nw, work.seed_idx.as<uint2>(), work.scan_tmp, s);
    seeds_rank_kernel<<<blocks, 256, 0, s>>>(plan.n_seeds, sseg, n_img, D,
                                             work.seed_idx.as<uint2>(), d_sorted);
    DL_LAUNCH_CHECK();
}

void init_canvases(const BatchPlan& plan, const uint32_t* d_sorted, uint32_t* plane_a,
                   uint32_t* plane_b, TriWork& work, hipStream_t s, const FillSeg* extra,
                   const FillSeg* extra2)
{
    const int n_img = static_cast<int>(plan.img.size());
    ProfScope ps(kKInit, s);
    // Only the first plane: the flood fill's first launch visits every tile
    // and writes every cell of the second plane.
    (void)plane_b;
    FillSegs fs{};
    fs.seg[0] = FillSeg{plane_a, static_cast<size_t>(plan.n_cells), kInvalid};
    fs.k = 1;
    if (extra) fs.seg[fs.k++] = *extra;
    if (extra2) fs.seg[fs.k++] = *extra2;
    fill_u32_multi(fs, s);
    canvas_seed_kernel<<<static_cast<unsigned>((plan.n_seeds + 255) / 256), 256, 0, s>>>(
        d_sorted, plan.n_seeds, seg_ptr(work, 0, n_img), n_img, work.desc.as<ImageDesc>(),
        plane_a);
    DL_LAUNCH_CHECK();
}

// ===========================================================================
// Detection
// ===========================================================================
namespace {

struct Cand {
    uint32_t key;    // registration key scan * nb + block, 0xFFFFFFFF = none
    uint32_t slot;
    int32_t a, b, c;
    uint32_t img;
};

constexpr uint64_t kEmptyTag = ~0ull;

__device__ __forceinline__ void sort3(int32_t& a, int32_t& b, int32_t& c)
{
    int32_t t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One 16-byte slot per distinct (image, triplet).
struct alignas(16) Slot {
    unsigned long long tag;
    uint32_t minkey;   // smallest registration key seen for the triplet
    uint32_t blk;      // 1: blocked by a cocircular quad
};

// The sorted triplet of a slot, kept only when tags are hashes (the
// collision check of mark_one).
struct SlotTri {
    int32_t a, b, c;
    uint32_t img;
};

// Open addressing with linear probing.  A triplet's probe starts at
// stride * (batch-wide id of its smallest seed), so the table is laid out in
// seed-id order, which is x order on the canvas: detection's candidate list
// also runs in x order, and an XCD working through a contiguous range of it
// touches a contiguous part of the table, which stays in its L2.
struct Table {
    Slot* slot;
    SlotTri* tri;     // null when tags are exact
    uint64_t mask;
    int exact;        // tags are exact packings of (image, sorted triplet)
    int bits;         // bits per seed id in an exact tag
    uint64_t salt;
    uint64_t stride;  // slots per seed (>= 1)
};

__device__ __forceinline__ uint64_t triplet_tag(const Table& t, uint32_t img, int32_t a, int32_t b,
                                                int32_t c)
{
    // a < b < c
    if (t.exact)
        return static_cast<uint64_t>(a) | (static_cast<uint64_t>(b) << t.bits) |
               (static_cast<uint64_t>(c) << (2 * t.bits)) |
               (static_cast<uint64_t>(img) << (3 * t.bits));
    uint64_t h = mix64(static_cast<uint64_t>(static_cast<uint32_t>(a)) ^ t.salt);
    h = mix64(h ^ (static_cast<uint64_t>(static_cast<uint32_t>(b)) << 1));
    h = mix64(h ^ (static_cast<uint64_t>(static_cast<uint32_t>(c)) << 2));
    h = mix64(h ^ (static_cast<uint64_t>(img) << 3));
    return h & 0x7FFFFFFFFFFFFFFFull;
}

// Returns the slot of (img, sorted triplet), inserting it if absent; -1 (flag
// raised) if the table is full, which its sizing rules out.
// a < b < c; seed_off: the image's first seed in the batch.
__device__ int64_t table_insert(const Table& t, uint32_t img, int64_t seed_off, int32_t a,
                                int32_t b, int32_t c, uint32_t* flags)
{
    const uint64_t tg = triplet_tag(t, img, a, b, c);
    uint64_t h = (static_cast<uint64_t>(seed_off + a) * t.stride) & t.mask;
    // One compare-and-swap per probed slot (its return value says whether the
    // slot was empty, holds this tag, or holds another): device-scope atomics
    // and loads are executed past the L2, so a load before the swap would
    // cost as much as the swap.
    for (uint64_t probe = 0; probe <= t.mask; ++probe) {
        Slot& sl = t.slot[h];
        const unsigned long long prev = atomicCAS(&sl.tag, static_cast<unsigned long long>(kEmptyTag),
                                                  static_cast<unsigned long long>(tg));
        if (prev == kEmptyTag) {
            if (t.tri) t.tri[h] = SlotTri{a, b, c, img};
            return static_cast<int64_t>(h);
        }
        if (prev == tg) return static_cast<int64_t>(h);
        h = (h + 1) & t.mask;
    }
    atomicOr(flags, kFlagTableOverflow);
    return -1;
}

// Empties the table and clears the registration bitmap (bm: nw words).
__global__ void table_init_kernel(Slot* __restrict__ slots, uint64_t n, uint32_t* __restrict__ bm,
                                  uint64_t nw)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // tag empty, minkey none, not blocked
    if (i < n) *reinterpret_cast<uint4*>(slots + i) = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u);
    for (uint64_t w = i; w < nw; w += static_cast<uint64_t>(gridDim.x) * blockDim.x) bm[w] = 0u;
}

// Block b of a grid that is a multiple of 8 runs on XCD b % 8 (dispatch is
// round-robin over the XCDs).  Each XCD takes a contiguous eighth of [0, n);
// the XCD's threads stride over it.
struct XcdRange {
    uint32_t first, end, step;
};
__device__ __forceinline__ XcdRange xcd_range(uint32_t n)
{
    const uint32_t xcd = blockIdx.x & 7u, per = gridDim.x >> 3, rank = blockIdx.x >> 3;
    const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(n) * xcd / 8);
    const uint32_t hi = static_cast<uint32_t>(static_cast<uint64_t>(n) * (xcd + 1) / 8);
    return XcdRange{lo + rank * blockDim.x + threadIdx.x, hi, per * blockDim.x};
}

// Orientation class of a direction for the reference's arctan2 ordering:
// y < 0 -> (-pi, 0); y == 0, x >= 0 -> 0; y > 0 -> (0, pi); y == 0, x < 0 -> pi.
__device__ __forceinline__ int angle_class(int64_t x, int64_t y)
{
    if (y < 0) return 0;
    if (y == 0) return x >= 0 ? 1 : 3;
    return 2;
}

// a sorts strictly before b in np.arctan2 order (exact for integer inputs).
__device__ __forceinline__ bool angle_less(int64_t ax, int64_t ay, int64_t bx, int64_t by)
{
    const int ca = angle_class(ax, ay), cb = angle_class(bx, by);
    if (ca != cb) return ca < cb;
    if (ca == 1 || ca == 3) return false;
    return ax * by - ay * bx > 0;
}

// triangulation.py:115-144 for one quad block.  ids: TL, TR, BL, BR;
// (px, py): their seed positions.  Writes the two blocked triplets.
__device__ void quad_rule(const int32_t ids[4], const int64_t px[4], const int64_t py[4],
                          int32_t out[2][3])
{
    const int64_t sx = px[0] + px[1] + px[2] + px[3];
    const int64_t sy = py[0] + py[1] + py[2] + py[3];
    int64_t vx[4], vy[4];
    int order[4];
    for (int k = 0; k < 4; ++k) {
        vx[k] = 4 * px[k] - sx;   // 4 * (p - centroid), exact
        vy[k] = 4 * py[k] - sy;
        order[k] = k;
    }
    for (int i = 1; i < 4; ++i) {   // stable insertion sort
        const int v = order[i];
        int j = i - 1;
        while (j >= 0 && angle_less(vx[v], vy[v], vx[order[j]], vy[order[j]])) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = v;
    }
    const int d0a = order[0], d0b = order[2], d1a = order[1], d1b = order[3];
    const int64_t l0 = (px[d0a] - px[d0b]) * (px[d0a] - px[d0b]) + (py[d0a] - py[d0b]) * (py[d0a] - py[d0b]);
    const int64_t l1 = (px[d1a] - px[d1b]) * (px[d1a] - px[d1b]) + (py[d1a] - py[d1b]) * (py[d1a] - py[d1b]);
    bool choose0;
    if (l0 < l1) choose0 = true;
    else if (l1 < l0) choose0 = false;
    else {
        int lowest = 0;
        for (int k = 1; k < 4; ++k)
            if (ids[k] < ids[lowest]) lowest = k;
        choose0 = !(lowest == d0a || lowest == d0b);
    }
    const int ca = choose0 ? d0a : d1a, cb = choose0 ? d0b : d1b;
    const int ra = choose0 ? d1a : d0a, rb = choose0 ? d1b : d0b;
    out[0][0] = ids[ra]; out[0][1] = ids[rb]; out[0][2] = ids[ca];
    out[1][0] = ids[ra]; out[1][1] = ids[rb]; out[1][2] = ids[cb];
}

struct DetectArgs {
    const uint32_t* grid;         // canvas planes of all images (image i at D[i].cv_off)
    int labels;                   // grid holds labels (else ids)
    const ImageDesc* D;
    const int64_t* det_seg;       // first detection thread block of each image
    int n_img;
    SeedIndex si;                 // position -> sorted id (labels mode)
    const uint32_t* sorted;       // id -> packed image position, per image
    uint2* positions;             // list of (image, block) of blocks holding candidates
    uint32_t* masks;              // per thread: its candidate blocks (count pass)
    uint32_t* block_counts;       // per thread block: candidate blocks / four-region
    uint32_t* block_quads;        //   blocks (count pass)
    const uint32_t* block_offsets;   // exclusive scan of block_counts
    Cand* cand;
    uint32_t* flags;              // batch-wide status word
    Table table;
    const uint32_t* npos;         // candidate positions actually listed (device)
    const uint32_t* n_listed;     // candidate positions found (scan total)
    const uint32_t* n_quads;      // four-region blocks found (scan total)
    uint64_t table_cap;
    uint32_t pos_cap;             // capacity of positions (cand holds 2 per position)
    uint32_t rows_cap;            // capacity of rows
    // Registration by owner buckets (ent null: the hash table registers).
    uint4* ent;                   // own_cap entries per bucket: (a, b, c) sorted
                                  // batch-wide ids, key (kBlockedKey: blocked)
    uint32_t* eidx;               // each entry's candidate slot
    uint32_t* fill;               // entries routed to each bucket
    uint8_t* first;               // per candidate slot: 1 = a first registration
    uint8_t* scode;               // per position: scans of its two candidates
    uint32_t n_buckets;
    uint32_t own_cap;
};

__device__ __forceinline__ int32_t to_id(const DetectArgs& A, const ImageDesc& d, uint32_t v)
{
    if (A.labels) {
        if (v == kInvalid) {   // an undefined cell: impossible after a converged fill
            atomicOr(A.flags, kFlagBadId);
            return -1;
        }
        const int lx = static_cast<int>(v >> 16) - d.P, ly = static_cast<int>(v & 0xFFFFu) - d.P;
        const int32_t id = seed_id_at(A.si, d.bm_off, d.seed_off, d.H, lx, ly);
        // No seed there only when the seeds failed validation (read back late).
        if (id < 0 || id >= d.n) {
            atomicOr(A.flags, kFlagBadId);
            return -1;
        }
        return id;
    }
    const int32_t id = static_cast<int32_t>(v);
    if (id < 0 || id >= d.n) {
        atomicOr(A.flags, kFlagBadId);
        return -1;
    }
    return id;
}

// The four corners of a 2x2 block (TL, TR, BL, BR) and which of the four
// L-scans (triangulation.py:174,177,180,183) see three distinct regions there.
struct BlockCorners {
    uint32_t tl, tr, bl, br;
    bool v[4];
    bool quad;
};

__device__ __forceinline__ BlockCorners corners_of(uint32_t tl, uint32_t tr, uint32_t bl, uint32_t br)
{
    BlockCorners c;
    c.tl = tl;
    c.tr = tr;
    c.bl = bl;
    c.br = br;
    const bool e_tl_tr = c.tl == c.tr, e_tl_bl = c.tl == c.bl, e_tl_br = c.tl == c.br;
    const bool e_tr_bl = c.tr == c.bl, e_tr_br = c.tr == c.br, e_bl_br = c.bl == c.br;
    c.v[0] = !e_tl_tr && !e_tl_br && !e_tr_br;   // (TL, TR, BR)
    c.v[1] = !e_tl_tr && !e_tl_bl && !e_tr_bl;   // (TL, TR, BL)
    c.v[2] = !e_bl_br && !e_tl_bl && !e_tl_br;   // (BL, BR, TL)
    c.v[3] = !e_bl_br && !e_tr_bl && !e_tr_br;   // (BL, BR, TR)
    c.quad = c.v[0] && c.v[1] && c.v[2] && c.v[3];
    return c;
}

__device__ __forceinline__ BlockCorners block_corners(const uint32_t* g, int Wc, int rx, int ry)
{
    const size_t o = static_cast<size_t>(ry) * Wc + rx;
    return corners_of(g[o], g[o + 1], g[o + Wc], g[o + Wc + 1]);
}

// Candidates a block contributes.  With three distinct regions every valid
// scan names the same triplet, and only the lowest scan's key can be that
// triplet's first registration: one candidate.  A four-region block blocks
// two of its four triplets globally; they can never register, so only the
// two triangles of the chosen diagonal are candidates.
__device__ __forceinline__ uint32_t block_candidates(const BlockCorners& c)
{
    if (c.quad) return 2;
    return (c.v[0] | c.v[1] | c.v[2] | c.v[3]) ? 1 : 0;
}

// Exclusive scans of two short arrays in one launch of one 1024-thread block:
// thread t owns a contiguous chunk of each; out_*[n] receives the totals.
__global__ __launch_bounds__(1024) void scan_two_small_kernel(const uint32_t* __restrict__ a,
                                                              const uint32_t* __restrict__ b,
                                                              int64_t n, uint32_t* __restrict__ oa,
                                                              uint32_t* __restrict__ ob)
{
    __shared__ uint32_t pa[1024], pb[1024];
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = threadIdx.x * per, hi = min(n, lo + per);
    uint32_t sa = 0, sb = 0;
    for (int64_t i = lo; i < hi; ++i) {
        sa += a[i];
        sb += b[i];
    }
    pa[threadIdx.x] = sa;
    pb[threadIdx.x] = sb;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t xa = threadIdx.x >= off ? pa[threadIdx.x - off] : 0u;
        const uint32_t xb = threadIdx.x >= off ? pb[threadIdx.x - off] : 0u;
        __syncthreads();
        pa[threadIdx.x] += xa;
        pb[threadIdx.x] += xb;
        __syncthreads();
    }
    uint32_t ra = pa[threadIdx.x] - sa, rb = pb[threadIdx.x] - sb;
    for (int64_t i = lo; i < hi; ++i) {
        oa[i] = ra;
        ob[i] = rb;
        ra += a[i];
        rb += b[i];
    }
    if (threadIdx.x == 1023) {
        oa[n] = pa[1023];
        ob[n] = pb[1023];
    }
}

// Candidate blocks in three steps without contended atomics: a count pass
// keeps each thread's mask of candidate blocks and each thread block's total;
// the block totals are scanned; a write pass lists the positions at the
// scanned offsets.  A thread block covers kDetRows x kDetChunk 2x2 blocks.
__device__ __forceinline__ int64_t det_block_origin(const DetectArgs& A, int64_t gb, int& img,
                                                    int& ry0, int& rx0)
{
    img = A.n_img > 1 ? find_seg(A.det_seg, A.n_img, gb) : 0;
    const ImageDesc& d = A.D[img];
    const int64_t lb = gb - A.det_seg[img];
    const int n_bands = (d.Hc - 1 + kDetRows - 1) / kDetRows;
    const int bx = static_cast<int>(lb / n_bands);
    const int band = static_cast<int>(lb - static_cast<int64_t>(bx) * n_bands);
    ry0 = band * kDetRows;
    rx0 = bx * kDetChunk + static_cast<int>(threadIdx.x);
    return lb;
}

__global__ __launch_bounds__(kDetThreads) void detect_count_kernel(DetectArgs A)
{
    __shared__ uint32_t red[2][kDetThreads / 64];
    const int64_t gb = blockIdx.x;
    int img, ry0, rx0;
    det_block_origin(A, gb, img, ry0, rx0);
    const ImageDesc& d = A.D[img];
    const int bw = d.Wc - 1, bh = d.Hc - 1;
    const uint32_t* g = A.grid + d.cv_off;
    const int rows = min(kDetRows, bh - ry0);   // block-uniform
    uint32_t mask = 0, quads = 0;   // bit r * kDetItems + j
#pragma unroll
    for (int j = 0; j < kDetItems; ++j) {
        // One column of blocks down the rows: each label row is the bottom
        // of one block and the top of the next, two loads per block.  All
        // of the column's loads are issued before the first block is
        // classified, so a wave waits for memory once, not once per row.
        // Rows past the canvas (last band) read its last row again and count
        // nothing, so the loads need no branch.
        const int rx = rx0 + j * kDetThreads;
        const bool col_ok = rx < bw;
        const uint32_t* col = g + min(rx, bw - 1);
        uint32_t lv[kDetRows + 1], rv[kDetRows + 1];
#pragma unroll
        for (int r = 0; r <= kDetRows; ++r) {
            const size_t o = static_cast<size_t>(min(ry0 + r, bh)) * d.Wc;
            lv[r] = col[o];
            rv[r] = col[o + 1];
        }
#pragma unroll
        for (int r = 0; r < kDetRows; ++r) {
            const BlockCorners c = corners_of(lv[r], rv[r], lv[r + 1], rv[r + 1]);
            if (col_ok && r < rows) {
                if (block_candidates(c)) mask |= 1u << (r * kDetItems + j);
                quads += c.quad;
            }
        }
    }
    A.masks[gb * kDetThreads + threadIdx.x] = mask;
    uint32_t cnt = static_cast<uint32_t>(__popc(mask));
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_down(cnt, off, 64);
        quads += __shfl_down(quads, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = cnt;
        red[1][threadIdx.x >> 6] = quads;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, q = 0;
        for (int w = 0; w < kDetThreads / 64; ++w) {
            c += red[0][w];
            q += red[1][w];
        }
        A.block_counts[gb] = c;
        A.block_quads[gb] = q;
    }
}

__global__ __launch_bounds__(kDetThreads) void detect_write_kernel(DetectArgs A)
{
    __shared__ uint32_t part[kDetThreads];
    const int64_t gb = blockIdx.x;
    if (gb == 0 && threadIdx.x == 0) {
        // Positions listed (capped) and the capacity checks of a speculatively
        // sized detection: more positions than the list holds, or more
        // distinct triplets than three quarters of the table, raise
        // kFlagOverflow (read by the later kernels and the host).
        const uint32_t n = *A.n_listed, q = *A.n_quads;
        *const_cast<uint32_t*>(A.npos) = min(n, A.pos_cap);
        if (n > A.pos_cap || 4 * (static_cast<uint64_t>(n) + 2ull * q) > 3 * A.table_cap)
            atomicOr(A.flags, kFlagOverflow);
    }
    if (A.ent)   // the owner buckets' counters, for route_kernel
        for (uint64_t k = static_cast<uint64_t>(gb) * kDetThreads + threadIdx.x; k < A.n_buckets;
             k += static_cast<uint64_t>(gridDim.x) * kDetThreads)
            A.fill[k] = 0u;
    int img, ry0, rx0;
    det_block_origin(A, gb, img, ry0, rx0);
    const ImageDesc& d = A.D[img];
    uint32_t mask = A.masks[gb * kDetThreads + threadIdx.x];
    const uint32_t cnt = static_cast<uint32_t>(__popc(mask));
    part[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kDetThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t out = A.block_offsets[gb] + part[threadIdx.x] - cnt;
    const uint32_t bwu = static_cast<uint32_t>(d.Wc - 1);
    while (mask) {
        const int b = __ffs(mask) - 1;
        mask &= mask - 1;
        if (out >= A.pos_cap) {   // capacity too small: flagged, the host reruns
            atomicOr(A.flags, kFlagOverflow);
            return;
        }
        const uint32_t ry = static_cast<uint32_t>(ry0 + b / kDetItems);
        const uint32_t rx = static_cast<uint32_t>(rx0 + (b % kDetItems) * kDetThreads);
        A.positions[out++] = make_uint2(static_cast<uint32_t>(img), ry * bwu + rx);
    }
}

__device__ __forceinline__ void add_candidate(const DetectArgs& A, const ImageDesc& d, uint32_t img,
                                              uint32_t& out, int scan, uint32_t pos, int32_t a,
                                              int32_t b, int32_t c)
{
    int32_t sa = a, sb = b, sc = c;
    sort3(sa, sb, sc);
    const int64_t sl = table_insert(A.table, img, d.seed_off, sa, sb, sc, A.flags);
    if (sl < 0) {
        A.cand[out++].key = 0xFFFFFFFFu;
        return;
    }
    const uint32_t key = static_cast<uint32_t>(scan) * static_cast<uint32_t>(d.nb) + pos;
    atomicMin(&A.table.slot[sl].minkey, key);
    Cand cd;
    cd.key = key;
    cd.slot = static_cast<uint32_t>(sl);
    cd.a = a;
    cd.b = b;
    cd.c = c;
    cd.img = img;
    A.cand[out++] = cd;
}

// Listed positions, grid-stride over the device count.  A position's
// candidates go to slots 2i and 2i+1 of the candidate array; an unused slot
// keeps key 0xFFFFFFFF.
__device__ void insert_position(const DetectArgs& A, uint32_t i);

__global__ __launch_bounds__(kDetThreads) void detect_insert_kernel(DetectArgs A)
{
    // Sizes already known to be too small: the host reruns with exact ones,
    // so skip the (then long) probing.
    if (*A.flags & kFlagOverflow) return;
    const XcdRange r = xcd_range(*A.npos);
    for (uint32_t i = r.first; i < r.end; i += r.step) insert_position(A, i);
}

__device__ void insert_position(const DetectArgs& A, uint32_t i)
{
    const uint2 ip = A.positions[i];
    const uint32_t img = ip.x, pos = ip.y;
    const ImageDesc& d = A.D[img];
    const uint32_t bw = static_cast<uint32_t>(d.Wc - 1);
    const int ry = static_cast<int>(pos / bw);
    const int rx = static_cast<int>(pos - static_cast<uint32_t>(ry) * bw);
    const BlockCorners c = block_corners(A.grid + d.cv_off, d.Wc, rx, ry);
    uint32_t out = 2 * i;
    A.cand[out + 1].key = 0xFFFFFFFFu;
    const int32_t tl = to_id(A, d, c.tl), tr = to_id(A, d, c.tr);
    const int32_t bl = to_id(A, d, c.bl), br = to_id(A, d, c.br);
    if (tl < 0 || tr < 0 || bl < 0 || br < 0) {
        A.cand[out].key = 0xFFFFFFFFu;
        return;
    }
    const int32_t tri[4][3] = {{tl, tr, br}, {tl, tr, bl}, {bl, br, tl}, {bl, br, tr}};
    if (!c.quad) {
        int k = 0;
        while (!c.v[k]) ++k;
        add_candidate(A, d, img, out, k, pos, tri[k][0], tri[k][1], tri[k][2]);
        return;
    }
    const int32_t ids[4] = {tl, tr, bl, br};
    const uint32_t* srt = A.sorted + d.seed_off;
    int64_t px[4], py[4];
    for (int k = 0; k < 4; ++k) {
        const uint32_t l = srt[ids[k]];
        px[k] = static_cast<int64_t>(l >> 16);
        py[k] = static_cast<int64_t>(l & 0xFFFFu);
    }
    int32_t blk[2][3];
    quad_rule(ids, px, py, blk);
    bool blocked[4] = {false, false, false, false};
    for (int q = 0; q < 2; ++q) {
        int32_t a = blk[q][0], b = blk[q][1], cc = blk[q][2];
        sort3(a, b, cc);
        const int64_t sl = table_insert(A.table, img, d.seed_off, a, b, cc, A.flags);
        if (sl >= 0) A.table.slot[sl].blk = 1u;
        for (int k = 0; k < 4; ++k) {
            int32_t x = tri[k][0], y = tri[k][1], z = tri[k][2];
            sort3(x, y, z);
            if (x == a && y == b && z == cc) blocked[k] = true;
        }
    }
    for (int k = 0; k < 4; ++k)
        if (!blocked[k]) add_candidate(A, d, img, out, k, pos, tri[k][0], tri[k][1], tri[k][2]);
}

__device__ __forceinline__ bool is_first(const Cand& cd, const Table& t)
{
    const Slot& sl = t.slot[cd.slot];
    return sl.minkey == cd.key && sl.blk == 0u;
}

// The registration bitmap holds one bit per (scan, row of blocks, column):
// word key_off + (scan * bh + ry) * (8 * bx_per_row) + rx / 32, bit rx % 32.
// Rows are padded to whole detection strips of kDetChunk columns, so every
// word belongs to the blocks of one detection thread block, and word order is
// key order (scan, then row-major position): the reference's registration
// order.
struct RegBit {
    int64_t word;
    uint32_t bit;
};

__device__ __forceinline__ RegBit reg_bit(const ImageDesc& d, uint32_t scan, int ry, int rx)
{
    const int64_t row = static_cast<int64_t>(scan) * (d.Hc - 1) + ry;
    return RegBit{d.key_off + row * (d.bx_per_row * (kDetChunk / 32)) + (rx >> 5),
                  static_cast<uint32_t>(rx & 31)};
}

// Scan and block position of a registration key scan * nb + ry * bw + rx.
__device__ __forceinline__ RegBit key_bit(const ImageDesc& d, uint32_t key, uint32_t& scan,
                                          int& ry, int& rx)
{
    const uint32_t nb = static_cast<uint32_t>(d.nb);
    scan = key / nb;
    const uint32_t rem = key - scan * nb;
    const uint32_t bw = static_cast<uint32_t>(d.Wc - 1);
    ry = static_cast<int>(rem / bw);
    rx = static_cast<int>(rem - static_cast<uint32_t>(ry) * bw);
    return reg_bit(d, scan, ry, rx);
}

// A set bit of the registration bitmap marks a first registration.
__device__ void mark_one(const Cand* __restrict__ cand, uint32_t i, const Table& t,
                         const ImageDesc* __restrict__ D, uint32_t* __restrict__ bm,
                         uint32_t* __restrict__ flags);

// Candidates, grid-stride over twice the device position count.
__global__ void mark_kernel(const Cand* __restrict__ cand, const uint32_t* __restrict__ npos,
                            Table t, const ImageDesc* __restrict__ D, uint32_t* __restrict__ bm,
                            uint32_t* __restrict__ flags)
{
    if (*flags & kFlagOverflow) return;   // the insert did not run: candidates unset
    const XcdRange r = xcd_range(2 * *npos);
    for (uint32_t i = r.first; i < r.end; i += r.step) mark_one(cand, i, t, D, bm, flags);
}

__device__ void mark_one(const Cand* __restrict__ cand, uint32_t i, const Table& t,
                         const ImageDesc* __restrict__ D, uint32_t* __restrict__ bm,
                         uint32_t* __restrict__ flags)
{
    const Cand cd = cand[i];
    if (cd.key == 0xFFFFFFFFu) return;
    if (t.tri) {   // hashed tags: a different triplet with the same tag shares the slot
        int32_t a = cd.a, b = cd.b, c = cd.c;
        sort3(a, b, c);
        const SlotTri st = t.tri[cd.slot];
        if (st.a != a || st.b != b || st.c != c || st.img != cd.img) {
            atomicOr(flags, kFlagCollision);
            return;
        }
    }
    if (is_first(cd, t)) {
        uint32_t scan;
        int ry, rx;
        const RegBit rb = key_bit(D[cd.img], cd.key, scan, ry, rx);
        atomicOr(&bm[rb.word], 1u << rb.bit);
    }
}

// First row of each image = number of first registrations before it.
__global__ void tri_base_kernel(const ImageDesc* __restrict__ D, int n_img,
                                const uint32_t* __restrict__ pre, int64_t n_words,
                                int64_t* __restrict__ tri_base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_img) return;
    tri_base[i] = i < n_img ? pre[D[i].key_off] : pre[n_words];
}

__device__ void emit_one(const Cand* __restrict__ cand, uint32_t i, const Table& t,
                         const uint8_t* __restrict__ first, const ImageDesc* __restrict__ D,
                         const uint32_t* __restrict__ bm, const uint32_t* __restrict__ pre,
                         int32_t* __restrict__ rows, uint32_t rows_cap,
                         uint32_t* __restrict__ flags);

// first: the owner buckets' per-slot decisions; null: the hash table's.
__global__ void emit_kernel(const Cand* __restrict__ cand, const uint32_t* __restrict__ npos,
                            Table t, const uint8_t* __restrict__ first,
                            const ImageDesc* __restrict__ D, const uint32_t* __restrict__ bm,
                            const uint32_t* __restrict__ pre, int32_t* __restrict__ rows,
                            uint32_t rows_cap, uint32_t* __restrict__ flags, int n_img,
                            int64_t n_words, int64_t* __restrict__ tri_base)
{
    // With tri_base given, block 0 also writes each image's first row (as
    // tri_base_kernel does).
    if (tri_base && blockIdx.x == 0)
        for (int i = threadIdx.x; i <= n_img; i += blockDim.x)
            tri_base[i] = i < n_img ? pre[D[i].key_off] : pre[n_words];
    if (*flags & kFlagOverflow) return;   // rerun with exact sizes follows
    const XcdRange r = xcd_range(2 * *npos);
    for (uint32_t i = r.first; i < r.end; i += r.step)
        emit_one(cand, i, t, first, D, bm, pre, rows, rows_cap, flags);
}

__device__ void emit_one(const Cand* __restrict__ cand, uint32_t i, const Table& t,
                         const uint8_t* __restrict__ first, const ImageDesc* __restrict__ D,
                         const uint32_t* __restrict__ bm, const uint32_t* __restrict__ pre,
                         int32_t* __restrict__ rows, uint32_t rows_cap,
                         uint32_t* __restrict__ flags)
{
    Cand cd;
    if (first) {
        if (!first[i]) return;
        cd = cand[i];
    } else {
        // Most positions fill one of their two slots: read the key alone first.
        if (cand[i].key == 0xFFFFFFFFu) return;
        cd = cand[i];
        if (!is_first(cd, t)) return;
    }
    const ImageDesc& d = D[cd.img];
    uint32_t scan;
    int ry, rx;
    const RegBit rb = key_bit(d, cd.key, scan, ry, rx);
    // global row = first registrations of earlier images and earlier keys
    const uint32_t row = pre[rb.word] +
                         static_cast<uint32_t>(__popc(bm[rb.word] & ((1u << rb.bit) - 1u)));
    if (row >= rows_cap) {
        atomicOr(flags, kFlagOverflow);
        return;
    }
    // detection pixel offsets of the four scans (triangulation.py:174-183)
    const int gx = rx + ((scan == 0 || scan == 3) ? 1 : 0);
    const int gy = ry + ((scan >= 2) ? 1 : 0);
    int32_t* r = rows + 5 * static_cast<size_t>(row);
    r[0] = gx - d.P;
    r[1] = gy - d.P;
    r[2] = cd.a;
    r[3] = cd.b;
    r[4] = cd.c;
}

// ---------------------------------------------------------------------------
// Registration by owner buckets (speculative runs).  Every occurrence of a
// triplet -- a candidate, or a triplet that a four-region block blocks -- has
// the same smallest seed id, so routing each occurrence to the bucket of its
// smallest batch-wide id brings all occurrences of a triplet to one bucket.
// There one workgroup decides in LDS which candidate is the first
// registration (triangulation.py:146-183): the one with the smallest key
// scan * nb + block whose triplet no occurrence blocks.  No global atomic is
// made per candidate: a routing workgroup reserves each bucket's slots with
// one atomicAdd, and each detection thread block writes its own words of the
// registration bitmap with plain stores.  A bucket whose fixed capacity
// overflows raises kFlagOverflow, and the exact rerun uses the hash table.
constexpr int kOwnShift = 8;                    // 256 consecutive seed ids per bucket
constexpr uint32_t kOwnCap = 1536;              // entries per bucket: 6 per seed
constexpr uint32_t kBlockedKey = 0xFFFFFFFFu;   // key of a blocked-triplet entry
constexpr int kRouteSpan = 512;                 // buckets a routing block counts in LDS

// What route_position produces for listed position i besides its entries:
// the candidate records of slots 2i and 2i+1 and the scans they came from.
// route_kernel stores them after its atomics: on this GPU a wait for a load or
// a returning atomic also waits for every store issued before it.
struct RouteOut {
    Cand o0, o1;   // (two fields, not an array: a runtime index would put it in scratch)
    uint32_t code;
};

// The entries of listed position i: bit k of the result marks entry ev[k]
// valid; entries 0 and 1 are the candidates in slots 2i and 2i+1, entries 2
// and 3 the triplets a four-region block blocks.  The candidate records and
// the scan code go to ro.
__device__ uint32_t route_position(const DetectArgs& A, uint32_t i, uint4 ev[4], RouteOut& ro)
{
    const uint2 ip = A.positions[i];
    const uint32_t img = ip.x, pos = ip.y;
    const ImageDesc& d = A.D[img];
    const uint32_t bw = static_cast<uint32_t>(d.Wc - 1);
    const int ry = static_cast<int>(pos / bw);
    const int rx = static_cast<int>(pos - static_cast<uint32_t>(ry) * bw);
    const BlockCorners c = block_corners(A.grid + d.cv_off, d.Wc, rx, ry);
    auto put = [&ro](int slot, const Cand& v) {
        if (slot == 0) ro.o0 = v;
        else ro.o1 = v;
    };
    const int32_t tl = to_id(A, d, c.tl), tr = to_id(A, d, c.tr);
    const int32_t bl = to_id(A, d, c.bl), br = to_id(A, d, c.br);
    Cand cd;
    cd.slot = 0;
    cd.img = img;
    if (tl < 0 || tr < 0 || bl < 0 || br < 0) {   // flagged: the host raises
        cd.key = 0xFFFFFFFFu;
        cd.a = cd.b = cd.c = 0;
        ro.o0 = cd;
        ro.o1 = cd;
        ro.code = 0;
        return 0u;
    }
    const uint32_t off = static_cast<uint32_t>(d.seed_off);
    const uint32_t nb = static_cast<uint32_t>(d.nb);
    const int32_t tri[4][3] = {{tl, tr, br}, {tl, tr, bl}, {bl, br, tl}, {bl, br, tr}};
    auto entry = [&](int32_t a, int32_t b, int32_t cc, uint32_t key) {
        sort3(a, b, cc);
        return make_uint4(off + static_cast<uint32_t>(a), off + static_cast<uint32_t>(b),
                          off + static_cast<uint32_t>(cc), key);
    };
    if (!c.quad) {
        int k = 0;
        while (!c.v[k]) ++k;
        cd.key = static_cast<uint32_t>(k) * nb + pos;
        cd.a = tri[k][0];
        cd.b = tri[k][1];
        cd.c = tri[k][2];
        ro.o0 = cd;
        ev[0] = entry(cd.a, cd.b, cd.c, cd.key);
        cd.key = 0xFFFFFFFFu;
        ro.o1 = cd;
        ro.code = static_cast<uint32_t>(k);
        return 1u;
    }
    const int32_t ids[4] = {tl, tr, bl, br};
    const uint32_t* srt = A.sorted + d.seed_off;
    int64_t px[4], py[4];
    for (int k = 0; k < 4; ++k) {
        const uint32_t l = srt[ids[k]];
        px[k] = static_cast<int64_t>(l >> 16);
        py[k] = static_cast<int64_t>(l & 0xFFFFu);
    }
    int32_t blk[2][3];
    quad_rule(ids, px, py, blk);
    ev[2] = entry(blk[0][0], blk[0][1], blk[0][2], kBlockedKey);
    ev[3] = entry(blk[1][0], blk[1][1], blk[1][2], kBlockedKey);
    // the two triplets of the chosen diagonal, in scan order, to slots 0 and 1
    int slot = 0;
    uint32_t code = 0;
    for (int k = 0; k < 4; ++k) {
        const uint4 e = entry(tri[k][0], tri[k][1], tri[k][2], 0u);
        const bool blocked = (e.x == ev[2].x && e.y == ev[2].y && e.z == ev[2].z) ||
                             (e.x == ev[3].x && e.y == ev[3].y && e.z == ev[3].z);
        if (blocked || slot == 2) continue;
        cd.key = static_cast<uint32_t>(k) * nb + pos;
        cd.a = tri[k][0];
        cd.b = tri[k][1];
        cd.c = tri[k][2];
        put(slot, cd);
        code |= static_cast<uint32_t>(k) << (2 * slot);
        if (slot == 0) ev[0] = make_uint4(e.x, e.y, e.z, cd.key);
        else ev[1] = make_uint4(e.x, e.y, e.z, cd.key);
        ++slot;
    }
    ro.code = code;
    // (a four-region block leaves exactly two; the rest is defensive)
    const uint32_t m = slot == 2 ? 0xFu : (slot == 1 ? 0xDu : 0xCu);
    for (cd.key = 0xFFFFFFFFu; slot < 2; ++slot) put(slot, cd);
    return m;
}

// One thread per listed position.  The block's entries are counted per
// bucket in LDS (buckets within kRouteSpan of the block's smallest; a block's
// positions are neighbours in x, so its buckets are too), each bucket's slots
// are reserved with one global atomicAdd, and the entries are written there.
__global__ __launch_bounds__(kDetThreads) void route_kernel(DetectArgs A)
{
    __shared__ uint32_t s_cnt[kRouteSpan];
    __shared__ uint32_t s_base[kRouteSpan];
    __shared__ uint32_t s_min[kDetThreads / 64];
    __shared__ uint32_t s_skip;
    const uint32_t npos = *A.npos;
    // Block b runs on XCD b % 8 (round-robin dispatch; the grid is a multiple
    // of 8, sized for the capacity): each XCD takes a contiguous eighth of the
    // listed positions, so the label lines and seed-index words of its x
    // range stay in its L2.
    const uint32_t per = ((npos + kDetThreads - 1) / kDetThreads + 7) / 8;
    if ((blockIdx.x >> 3) >= per) return;   // block-uniform
    const uint32_t p0 = ((blockIdx.x & 7u) * per + (blockIdx.x >> 3)) * kDetThreads;
    if (p0 >= npos) return;   // block-uniform
    if (threadIdx.x == 0) s_skip = *A.flags & kFlagOverflow;
    for (int t = threadIdx.x; t < kRouteSpan; t += kDetThreads) s_cnt[t] = 0u;
    __syncthreads();
    // Sizes already known to be too small: the host reruns with exact ones.
    if (s_skip) return;       // block-uniform (read once above)
    const uint32_t i = p0 + threadIdx.x;
    uint4 ev[4];
    RouteOut ro;
    const uint32_t m = i < npos ? route_position(A, i, ev, ro) : 0u;
    uint32_t bmin = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (m & (1u << k)) bmin = min(bmin, ev[k].x >> kOwnShift);
    for (int o = 32; o > 0; o >>= 1) bmin = min(bmin, static_cast<uint32_t>(__shfl_xor(bmin, o, 64)));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = bmin;
    __syncthreads();
    uint32_t b0 = s_min[0];
#pragma unroll
    for (int w = 1; w < kDetThreads / 64; ++w) b0 = min(b0, s_min[w]);
    uint32_t rank[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rank[k] = 0xFFFFFFFFu;
        if (m & (1u << k)) {
            const uint32_t o = (ev[k].x >> kOwnShift) - b0;
            if (o < static_cast<uint32_t>(kRouteSpan)) rank[k] = atomicAdd(&s_cnt[o], 1u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kRouteSpan; t += kDetThreads) {
        const uint32_t n = s_cnt[t];
        if (n) s_base[t] = atomicAdd(&A.fill[b0 + t], n);
    }
    __syncthreads();
    // Every slot first (an entry outside the block's span takes its slot
    // with an atomic), then every store.
    uint32_t sl[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sl[k] = 0xFFFFFFFFu;
        if (!(m & (1u << k))) continue;
        const uint32_t b = ev[k].x >> kOwnShift;
        sl[k] = rank[k] != 0xFFFFFFFFu ? s_base[b - b0] + rank[k] : atomicAdd(&A.fill[b], 1u);
    }
    asm volatile("" ::"v"(sl[0]), "v"(sl[1]), "v"(sl[2]), "v"(sl[3]));   // (waited for here)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!(m & (1u << k))) continue;
        if (sl[k] >= A.own_cap) {
            atomicOr(A.flags, kFlagOverflow);
            continue;
        }
        const size_t e = static_cast<size_t>(ev[k].x >> kOwnShift) * A.own_cap + sl[k];
        A.ent[e] = ev[k];
        A.eidx[e] = k < 2 ? 2 * i + static_cast<uint32_t>(k) : 0xFFFFFFFFu;
    }
    if (i < npos) {   // the candidate records, their scans, their first flags cleared
        Cand* out = A.cand + 2 * static_cast<size_t>(i);
        out[0] = ro.o0;
        out[1] = ro.o1;
        A.scode[i] = static_cast<uint8_t>(ro.code);
        *reinterpret_cast<uint16_t*>(A.first + 2 * static_cast<size_t>(i)) = 0;
    }
}

// One workgroup per bucket: its entries are grouped by smallest id in LDS (a
// counting sort over the bucket's 256 ids), and each candidate compares
// itself with the entries of its group.  It is a first registration unless an
// entry of the same triplet has a smaller key or blocks the triplet.
__global__ __launch_bounds__(256) void owner_kernel(DetectArgs A)
{
    __shared__ uint4 E[kOwnCap];
    __shared__ uint16_t ord[kOwnCap];
    __shared__ uint32_t start[257];
    __shared__ uint32_t cur[256];
    __shared__ uint32_t wsum[4];
    if (*A.flags & kFlagOverflow) return;   // set before this launch: uniform
    const uint32_t b = blockIdx.x;
    const uint32_t n = min(A.fill[b], A.own_cap);
    const uint32_t base = b << kOwnShift;
    const int t = threadIdx.x;
    start[t] = 0u;
    cur[t] = 0u;
    __syncthreads();
    const uint4* ent = A.ent + static_cast<size_t>(b) * A.own_cap;
    for (uint32_t j = t; j < n; j += 256) {
        const uint4 e = ent[j];
        E[j] = e;
        atomicAdd(&start[e.x - base], 1u);
    }
    __syncthreads();
    // exclusive scan of the 256 group sizes
    const uint32_t v = start[t];
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if ((t & 63) >= o) incl += u;
    }
    if ((t & 63) == 63) wsum[t >> 6] = incl;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < (t >> 6); ++w) add += wsum[w];
    __syncthreads();
    start[t] = add + incl - v;
    if (t == 255) start[256] = add + incl;
    __syncthreads();
    for (uint32_t j = t; j < n; j += 256) {
        const uint32_t g = E[j].x - base;
        ord[start[g] + atomicAdd(&cur[g], 1u)] = static_cast<uint16_t>(j);
    }
    __syncthreads();
    for (uint32_t j = t; j < n; j += 256) {
        const uint4 e = E[j];
        if (e.w == kBlockedKey) continue;
        const uint32_t g = e.x - base;
        bool first = true;
        for (uint32_t q = start[g]; q < start[g + 1]; ++q) {
            const uint4 f = E[ord[q]];
            if (f.y == e.y && f.z == e.z && (f.w < e.w || f.w == kBlockedKey)) {
                first = false;
                break;
            }
        }
        if (first) A.first[A.eidx[static_cast<size_t>(b) * A.own_cap + j]] = 1u;
    }
}

// One workgroup per detection thread block (its 256 x 32 window of blocks):
// the first registrations among its listed positions set bits of its LDS copy
// of its registration bitmap words, which it then stores, zeros included.
__global__ __launch_bounds__(256) void mark_blocks_kernel(DetectArgs A, uint32_t* __restrict__ bm)
{
    constexpr int kWords = kDetChunk / 32;          // words per strip row
    __shared__ uint32_t W[4 * kDetRows * kWords];   // [scan][row][word]
    if (*A.flags & kFlagOverflow) return;   // set before this launch: uniform
    const int64_t gb = blockIdx.x;
    const int img = A.n_img > 1 ? find_seg(A.det_seg, A.n_img, gb) : 0;
    const ImageDesc& d = A.D[img];
    const int64_t lb = gb - A.det_seg[img];
    const int bh = d.Hc - 1;
    const int n_bands = (bh + kDetRows - 1) / kDetRows;
    const int bx = static_cast<int>(lb / n_bands);
    const int ry0 = static_cast<int>(lb - static_cast<int64_t>(bx) * n_bands) * kDetRows;
    const int rx0 = bx * kDetChunk;
    for (int t = threadIdx.x; t < 4 * kDetRows * kWords; t += 256) W[t] = 0u;
    __syncthreads();
    const uint32_t npos = *A.npos;
    const uint32_t lo = min(A.block_offsets[gb], npos), hi = min(A.block_offsets[gb + 1], npos);
    const uint32_t bw = static_cast<uint32_t>(d.Wc - 1);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
        const uint32_t f = *reinterpret_cast<const uint16_t*>(A.first + 2 * static_cast<size_t>(i));
        if (!f) continue;
        const uint32_t pos = A.positions[i].y;
        const int ry = static_cast<int>(pos / bw);
        const int rx = static_cast<int>(pos - static_cast<uint32_t>(ry) * bw);
        const int lr = ry - ry0, lc = rx - rx0;
        const uint32_t code = A.scode[i];
        const uint32_t bit = 1u << (lc & 31);
        if (f & 0xFFu) atomicOr(&W[((code & 3u) * kDetRows + lr) * kWords + (lc >> 5)], bit);
        if (f >> 8) atomicOr(&W[(((code >> 2) & 3u) * kDetRows + lr) * kWords + (lc >> 5)], bit);
    }
    __syncthreads();
    const int rows = min(kDetRows, bh - ry0);
    for (int t = threadIdx.x; t < 4 * kDetRows * kWords; t += 256) {
        const int sc = t / (kDetRows * kWords), r = (t / kWords) % kDetRows, w = t % kWords;
        if (r < rows) {
            const RegBit rb = reg_bit(d, static_cast<uint32_t>(sc), ry0 + r, rx0);
            bm[rb.word + w] = W[t];
        }
    }
}

__global__ void zero_u32_kernel(uint32_t* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

DetectResult detect_triangles(const BatchPlan& plan, const uint32_t* grid_base, bool labels,
                              const uint32_t* sorted, TriWork& work,
                              hipStream_t s, bool speculative, bool flags_zeroed,
                              const std::function<void()>& after_detect)
{
    DetectResult res;
    const int n_img = static_cast<int>(plan.img.size());
    work.flags.reserve(sizeof(uint32_t) * (n_img + 8));
    work.host_flags.reserve(sizeof(uint32_t) * 8 + sizeof(int64_t) * (n_img + 1));
    uint32_t* dflags = work.flags.as<uint32_t>();
    uint32_t* hflags = work.host_flags.as<uint32_t>();
    int64_t* tri_base = seg_ptr(work, 3, n_img);
    work.tri_base.assign(n_img + 1, 0);
    work.rows_cap = 0;
    const int64_t nblocks = plan.n_det_blocks;
    // flags layout: [0] status bits, [2] four-region blocks (counter),
    // [4] candidate positions found (counter), [6] positions in the list
    // (min of [4] and the capacity; read by the later kernels).
    if (!flags_zeroed) {
        ProfScope ps(kKDetect, s);
        zero_u32_kernel<<<1, 64, 0, s>>>(dflags, 8);
        DL_LAUNCH_CHECK();
    }
    if (nblocks == 0) {
        work.rows.reserve(sizeof(int32_t) * 5);
        fill_u32(reinterpret_cast<uint32_t*>(tri_base), 0u, 2 * static_cast<size_t>(n_img + 1), s);
        return res;
    }
    DetectArgs A{};
    A.grid = grid_base;
    A.labels = labels ? 1 : 0;
    A.D = work.desc.as<ImageDesc>();
    A.det_seg = seg_ptr(work, 1, n_img);
    A.n_img = n_img;
    A.si = seed_index(work);
    A.sorted = sorted;
    A.flags = dflags;
    A.npos = dflags + 6;
    const int64_t nthr = nblocks * kDetThreads;
    work.counts.reserve(sizeof(uint32_t) * static_cast<size_t>(nthr + 4 * (nblocks + 1)));
    A.masks = work.counts.as<uint32_t>();
    A.block_counts = A.masks + nthr;
    A.block_quads = A.block_counts + (nblocks + 1);
    uint32_t* block_offsets = A.block_quads + (nblocks + 1);
    uint32_t* quad_scan = block_offsets + (nblocks + 1);
    A.block_offsets = block_offsets;
    {
        ProfScope ps(kKDetect, s);
        detect_count_kernel<<<static_cast<unsigned>(nblocks), kDetThreads, 0, s>>>(A);
        DL_LAUNCH_CHECK();
        // One 1024-thread block walks its chunks serially (≈ 5 us for the
        // 2,200 blocks of a 4096^2 image, 440 us for 77,000); beyond 8 per
        // thread the multi-block scans are faster.
        if (nblocks <= 1024 * 8) {
            scan_two_small_kernel<<<1, 1024, 0, s>>>(A.block_counts, A.block_quads, nblocks,
                                                     block_offsets, quad_scan);
            DL_LAUNCH_CHECK();
        } else {
            exclusive_scan_u32(A.block_counts, nblocks, block_offsets, work.scan_tmp, s);
            exclusive_scan_u32(A.block_quads, nblocks, quad_scan, work.scan_tmp, s);
        }
    }
    const uint32_t* n_listed = block_offsets + nblocks;   // scan totals
    const uint32_t* n_quads = quad_scan + nblocks;
    // Capacities: from the seed count (positions ~ triangles ~ 2 per seed),
    // checked on the device; or exact, from the totals read back.
    const uint64_t n_seeds = static_cast<uint64_t>(plan.n_seeds);
    uint64_t pos_cap = 3 * n_seeds + 4096;
    uint64_t table_cap = next_pow2(4 * n_seeds + 4096);
    if (speculative) {
        if (work.spec_tiny == 1) {   // test hook: capacities certain to overflow
            pos_cap = 1;
            table_cap = 1024;
        }
    } else {
        DL_HIP(hipMemcpyAsync(hflags + 2, n_quads, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        DL_HIP(hipMemcpyAsync(hflags + 4, n_listed, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        const uint64_t npos = hflags[4], nquads = hflags[2];
        pos_cap = std::max<uint64_t>(npos, 1);
        // distinct triplets <= 2 per position (a quad adds two blocked ones):
        // at most half full
        table_cap = next_pow2(2 * (npos + 2 * nquads) + 1024);
        res.n_quads = static_cast<int64_t>(nquads);
    }
    A.pos_cap = static_cast<uint32_t>(std::min<uint64_t>(pos_cap, 0x7FFFFFF0ull));
    A.n_listed = n_listed;
    A.n_quads = n_quads;
    A.table_cap = table_cap;
    work.positions.reserve(sizeof(uint2) * static_cast<size_t>(A.pos_cap));
    A.positions = work.positions.as<uint2>();
    A.rows_cap = 2 * A.pos_cap;
    work.rows_cap = A.rows_cap;
    work.cand.reserve(sizeof(Cand) * 2 * static_cast<size_t>(A.pos_cap));
    A.cand = work.cand.as<Cand>();
    work.rows.reserve(sizeof(int32_t) * 5 * static_cast<size_t>(A.rows_cap));
    // Speculative runs register by owner buckets (batch-wide seed ids must
    // fit 31 bits); the exact rerun after any overflow uses the hash table.
    const bool route = speculative && n_seeds < (1ull << 31);
    if (route) {
        const uint64_t nbk = std::max<uint64_t>(1, (n_seeds + (1u << kOwnShift) - 1) >> kOwnShift);
        A.n_buckets = static_cast<uint32_t>(nbk);
        A.own_cap = work.spec_tiny ? 1u : kOwnCap;   // (test hook: buckets of one entry)
        work.route_ent.reserve(sizeof(uint4) * nbk * kOwnCap);
        work.route_idx.reserve(sizeof(uint32_t) * nbk * kOwnCap);
        work.route_aux.reserve(sizeof(uint32_t) * nbk + 3 * static_cast<size_t>(A.pos_cap) + 16);
        A.ent = work.route_ent.as<uint4>();
        A.eidx = work.route_idx.as<uint32_t>();
        A.fill = work.route_aux.as<uint32_t>();
        A.first = reinterpret_cast<uint8_t*>(A.fill + nbk);
        A.scode = A.first + 2 * static_cast<size_t>(A.pos_cap);
    } else {
        work.table_tag.reserve(sizeof(Slot) * table_cap);
    }
    {
        ProfScope ps(kKDetect, s);
        detect_write_kernel<<<static_cast<unsigned>(nblocks), kDetThreads, 0, s>>>(A);
        DL_LAUNCH_CHECK();
    }
    if (after_detect) after_detect();
    const int64_t nwords = plan.n_key_words;
    work.bitmap.reserve(sizeof(uint32_t) * static_cast<size_t>(std::max<int64_t>(nwords, 1)));
    work.scan.reserve(sizeof(uint32_t) * static_cast<size_t>(nwords + 1));
    uint32_t* bm = work.bitmap.as<uint32_t>();
    uint32_t* pre = work.scan.as<uint32_t>();
    // grid-stride launches sized for the capacity, at most 16 blocks per CU
    // (multiples of 8: the kernels split their range over the XCDs)
    const unsigned cblocks = static_cast<unsigned>(
        (std::min<uint64_t>((2ull * A.pos_cap + 255) / 256, 4096) + 7) / 8 * 8);
    if (route) {
        {
            ProfScope ps(kKDetect, s);
            route_kernel<<<static_cast<unsigned>(
                               ((A.pos_cap + kDetThreads - 1) / kDetThreads + 7) / 8 * 8),
                           kDetThreads, 0, s>>>(A);
            DL_LAUNCH_CHECK();
        }
        ProfScope ps(kKRegister, s);
        owner_kernel<<<A.n_buckets, 256, 0, s>>>(A);
        DL_LAUNCH_CHECK();
        // (every word of the bitmap is stored by its detection thread block)
        mark_blocks_kernel<<<static_cast<unsigned>(nblocks), 256, 0, s>>>(A, bm);
        DL_LAUNCH_CHECK();
        popcount_scan(bm, nwords, pre, work.scan_tmp, s);
        emit_kernel<<<cblocks, 256, 0, s>>>(A.cand, A.npos, Table{}, A.first, A.D, bm, pre,
                                            work.rows.as<int32_t>(), A.rows_cap, dflags, n_img,
                                            nwords, tri_base);
        DL_LAUNCH_CHECK();
        res.n_tri = -1;   // known on the device only (tri_base[n_img])
        return res;
    }
    uint64_t salt = 0x243F6A8885A308D3ull;
    const int img_bits = n_img > 1 ? bits_for(n_img - 1) : 0;
    const int exact = 3 * plan.max_seed_bits + img_bits <= 63 ? 1 : 0;
    const unsigned pblocks = static_cast<unsigned>(
        (std::min<uint64_t>((A.pos_cap + kDetThreads - 1) / kDetThreads, 4096) + 7) / 8 * 8);
    if (!exact) work.table_tri.reserve(sizeof(SlotTri) * table_cap);
    for (int attempt = 0; attempt < 4; ++attempt) {
        Table t;
        t.slot = work.table_tag.as<Slot>();
        t.tri = exact ? nullptr : work.table_tri.as<SlotTri>();
        t.mask = table_cap - 1;
        t.stride = std::max<uint64_t>(1, table_cap / std::max<uint64_t>(n_seeds, 1));
        t.exact = exact;
        t.bits = plan.max_seed_bits;
        t.salt = salt;
        A.cand = work.cand.as<Cand>();
        A.table = t;
        {
            ProfScope ps(kKInit, s);
            table_init_kernel<<<static_cast<unsigned>((table_cap + 255) / 256), 256, 0, s>>>(
                t.slot, table_cap, bm, static_cast<uint64_t>(nwords));
            DL_LAUNCH_CHECK();
        }
        {
            ProfScope ps(kKDetect, s);
            detect_insert_kernel<<<pblocks, kDetThreads, 0, s>>>(A);
            DL_LAUNCH_CHECK();
        }
        {
            ProfScope ps(kKRegister, s);
            mark_kernel<<<cblocks, 256, 0, s>>>(A.cand, A.npos, t, A.D, bm, dflags);
            DL_LAUNCH_CHECK();
            popcount_scan(bm, nwords, pre, work.scan_tmp, s);
            if (!speculative) {   // read back below; otherwise emit_kernel writes it
                tri_base_kernel<<<static_cast<unsigned>((n_img + 256) / 256), 256, 0, s>>>(
                    A.D, n_img, pre, nwords, tri_base);
                DL_LAUNCH_CHECK();
            }
        }
        if (!speculative) {
            int64_t* htb = reinterpret_cast<int64_t*>(hflags + 8);
            DL_HIP(hipMemcpyAsync(hflags, dflags, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            DL_HIP(hipMemcpyAsync(htb, tri_base, sizeof(int64_t) * (n_img + 1),
                                  hipMemcpyDeviceToHost, s));
            DL_HIP(hipStreamSynchronize(s));
            const uint32_t f = hflags[0];
            if (f & kFlagBadId) {
                if (labels) throw runtime_error("flood fill left an undefined cell");
                throw value_error("voronoi_grid holds a seed id outside [0, n_seeds)");
            }
            if (f & (kFlagTableOverflow | kFlagOverflow))
                throw runtime_error("triangle registration table overflow");
            if (f & kFlagCollision) {
                salt = salt * 6364136223846793005ull + 1442695040888963407ull;
                continue;
            }
            work.tri_base.assign(htb, htb + n_img + 1);
            res.n_tri = htb[n_img];
        } else {
            res.n_tri = -1;   // known on the device only (tri_base[n_img])
        }
        {
            ProfScope ps(kKRegister, s);
            emit_kernel<<<cblocks, 256, 0, s>>>(A.cand, A.npos, t, nullptr, A.D, bm, pre,
                                                work.rows.as<int32_t>(), A.rows_cap, dflags,
                                                n_img, nwords, speculative ? tri_base : nullptr);
            DL_LAUNCH_CHECK();
        }
        return res;
    }
    throw runtime_error("triangle registration hash collided repeatedly");
}

namespace {

// Words of the status read-back, gathered on the device and copied in one
// piece into work.host_flags: [0] status bits, [1] bin-list entries needed,
// [2..) tri_base (n_img + 1 int64), then the padding-watch flags, the seed
// flags and the fill's per-launch change flags with its visit total.
struct StatusLayout {
    uint32_t tb, esc, seed, fill, words;
};
StatusLayout status_layout(int n_img, int n_esc, bool seed, int n_fill)
{
    StatusLayout L;
    L.tb = 2;
    L.esc = L.tb + 2 * static_cast<uint32_t>(n_img + 1);
    L.seed = L.esc + static_cast<uint32_t>(n_esc);
    L.fill = L.seed + (seed ? static_cast<uint32_t>(n_img) : 0u);
    L.words = L.fill + static_cast<uint32_t>(n_fill);
    return L;
}

struct GatherArgs {
    const uint32_t* src[6];
    uint32_t words[6];
    uint32_t dst[6];
    int n;
};

__global__ __launch_bounds__(256) void gather_words_kernel(GatherArgs g, uint32_t* __restrict__ out)
{
    for (int k = 0; k < g.n; ++k)
        for (uint32_t i = threadIdx.x; i < g.words[k]; i += blockDim.x)
            out[g.dst[k] + i] = g.src[k][i];
}

}  // namespace

void issue_status(const BatchPlan& plan, TriWork& work, const int* d_esc, int n_esc,
                  bool seed_flags, const int* d_fill, int n_fill, hipStream_t s)
{
    const int n_img = static_cast<int>(plan.img.size());
    if (!d_esc) n_esc = 0;
    if (!d_fill) n_fill = 0;
    const StatusLayout L = status_layout(n_img, n_esc, seed_flags, n_fill);
    work.status_dev.reserve(sizeof(uint32_t) * L.words);
    work.host_flags.reserve(sizeof(uint32_t) * L.words);
    GatherArgs g{};
    auto add = [&](const void* src, uint32_t words, uint32_t dst) {
        if (!words) return;
        g.src[g.n] = static_cast<const uint32_t*>(src);
        g.words[g.n] = words;
        g.dst[g.n] = dst;
        ++g.n;
    };
    const uint32_t* flags = work.flags.as<uint32_t>();
    add(flags, 1, 0);
    // bin-list entries needed: copied only when the assignment ran
    if (plan.n_tiles > 0 && work.list_cap > 0)
        add(work.bin_offsets.as<uint32_t>() + plan.n_tiles, 1, 1);
    add(seg_ptr(work, 3, n_img), 2 * static_cast<uint32_t>(n_img + 1), L.tb);
    add(d_esc, static_cast<uint32_t>(n_esc), L.esc);
    if (seed_flags) add(flags + kSeedFlagBase, static_cast<uint32_t>(n_img), L.seed);
    add(d_fill, static_cast<uint32_t>(n_fill), L.fill);
    uint32_t* dev = work.status_dev.as<uint32_t>();
    {
        ProfScope ps(kKRegister, s);
        gather_words_kernel<<<1, 256, 0, s>>>(g, dev);
        DL_LAUNCH_CHECK();
    }
    copy_back_async(work.host_flags.ptr, dev, sizeof(uint32_t) * L.words, s);
}

SpeculativeStatus parse_status(const BatchPlan& plan, TriWork& work, int n_esc, bool seed_flags,
                               int n_fill, const uint32_t** seed_out, const int** fill_out)
{
    const int n_img = static_cast<int>(plan.img.size());
    const StatusLayout L = status_layout(n_img, n_esc, seed_flags, n_fill);
    const uint32_t* h = work.host_flags.as<uint32_t>();
    const int64_t* htb = reinterpret_cast<const int64_t*>(h + L.tb);
    SpeculativeStatus st;
    const uint32_t f = h[0];
    st.bad_id = (f & kFlagBadId) != 0;
    const int64_t listed = plan.n_tiles > 0 && work.list_cap > 0 ? h[1] : 0;
    st.redo = (f & (kFlagOverflow | kFlagTableOverflow | kFlagCollision)) != 0 ||
              listed > work.list_cap;
    st.bin_redo = (f & kFlagBinOverflow) != 0;
    st.n_tri = htb[n_img];
    work.tri_base.assign(htb, htb + n_img + 1);
    const int* hesc = reinterpret_cast<const int*>(h + L.esc);
    st.escaped.assign(hesc, hesc + n_esc);
    if (seed_out) *seed_out = h + L.seed;
    if (fill_out) *fill_out = reinterpret_cast<const int*>(h + L.fill);
    return st;
}

SpeculativeStatus read_status(const BatchPlan& plan, TriWork& work, const int* d_esc, int n_esc,
                              hipStream_t s)
{
    issue_status(plan, work, d_esc, n_esc, false, nullptr, 0, s);
    DL_HIP(hipStreamSynchronize(s));
    return parse_status(plan, work, d_esc ? n_esc : 0, false, 0, nullptr, nullptr);
}

// ===========================================================================
// Voronoi output (one image, no triangles)
// ===========================================================================
namespace {

// channels: (rank(owner), d^2[, tid]).  Output pixel (x, y) reads canvas cell
// (x + off, y + off); the owner's image position is its canvas position - shift.
template <int C>
__global__ void output_kernel(int W, int H, const uint32_t* __restrict__ lab, int Wc, int off,
                              int shift, SeedIndex si, int64_t bm_off, int64_t seed_off,
                              int H_img, const int32_t* __restrict__ tid,
                              int32_t* __restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const uint32_t l = lab[static_cast<size_t>(y + off) * Wc + (x + off)];
    const int lx = static_cast<int>(l >> 16), ly = static_cast<int>(l & 0xFFFFu);
    const int dx = x + off - lx, dy = y + off - ly;
    const size_t p = static_cast<size_t>(y) * W + x;
    int32_t* o = out + C * p;
    if (l == kInvalid) {   // undefined cell (cannot occur after a converged fill)
        o[0] = -1;
        o[1] = -1;
    } else {
        o[0] = seed_id_at(si, bm_off, seed_off, H_img, lx - shift, ly - shift);
        o[1] = dx * dx + dy * dy;
    }
    if (C == 3) o[2] = tid[p];
}

}  // namespace

void write_output(int W, int H, const uint32_t* canvas_labels, int Wc, int off, int shift,
                  const SeedIndex& si, int64_t bm_off, int64_t seed_off, int H_img,
                  const int32_t* tidgrid, int32_t* out, int C, hipStream_t s)
{
    ProfScope ps(kKOutput, s);
    dim3 grid((W + 255) / 256, H);
    if (C == 3)
        output_kernel<3><<<grid, 256, 0, s>>>(W, H, canvas_labels, Wc, off, shift, si, bm_off,
                                              seed_off, H_img, tidgrid, out);
    else
        output_kernel<2><<<grid, 256, 0, s>>>(W, H, canvas_labels, Wc, off, shift, si, bm_off,
                                              seed_off, H_img, tidgrid, out);
    DL_LAUNCH_CHECK();
}

// ===========================================================================
// Tiled assignment fused with output assembly
//
// Triangles are binned to 64x64 pixel tiles by their clipped bounding box; one
// workgroup per tile rasterises its triangles into an LDS tile with LDS max
// (the reduction is order independent) and then writes the tile's output:
// channel 2 of the (H, W, 3) grid together with channels 0/1 read from the
// image's output label source, or the bare tid raster.
// ===========================================================================
namespace {

constexpr int kTS = 64;   // tile side in pixels

struct TriVerts {
    int64_t x[3], y[3];
    int64_t S;   // doubled signed area
};

__device__ __forceinline__ TriVerts load_tri(const int32_t* rows, const uint32_t* sorted, int64_t t)
{
    TriVerts v;
    const int32_t* r = rows + 5 * t;
    for (int k = 0; k < 3; ++k) {
        const uint32_t l = sorted[r[2 + k]];
        v.x[k] = static_cast<int64_t>(l >> 16);
        v.y[k] = static_cast<int64_t>(l & 0xFFFFu);
    }
    v.S = (v.x[1] - v.x[0]) * (v.y[2] - v.y[0]) - (v.y[1] - v.y[0]) * (v.x[2] - v.x[0]);
    return v;
}

// Pixel bounding box of the triangle's closed region clipped to the image;
// false if empty.  A collinear triple contains its whole line, so its box is
// the line's extent across the image.
__device__ __forceinline__ bool tri_box(const TriVerts& v, int W, int H, int& x0, int& y0, int& x1,
                                        int& y1)
{
    if (v.S != 0) {
        int64_t a = min(v.x[0], min(v.x[1], v.x[2])), b = max(v.x[0], max(v.x[1], v.x[2]));
        int64_t c = min(v.y[0], min(v.y[1], v.y[2])), d = max(v.y[0], max(v.y[1], v.y[2]));
        a = max<int64_t>(a, 0); b = min<int64_t>(b, W - 1);
        c = max<int64_t>(c, 0); d = min<int64_t>(d, H - 1);
        x0 = static_cast<int>(a); x1 = static_cast<int>(b);
        y0 = static_cast<int>(c); y1 = static_cast<int>(d);
        return a <= b && c <= d;
    }
    const int64_t dx = v.x[1] - v.x[0], dy = v.y[1] - v.y[0];
    if (dy == 0) {
        if (v.y[0] < 0 || v.y[0] >= H) return false;
        x0 = 0; x1 = W - 1;
        y0 = y1 = static_cast<int>(v.y[0]);
        return true;
    }
    if (dx == 0) {
        if (v.x[0] < 0 || v.x[0] >= W) return false;
        y0 = 0; y1 = H - 1;
        x0 = x1 = static_cast<int>(v.x[0]);
        return true;
    }
    x0 = 0; x1 = W - 1; y0 = 0; y1 = H - 1;
    return true;
}

// Row spans of a non-degenerate triangle.  Edge e's function at (px, y),
// multiplied by sign(S), is al*px + k0 + k1*y; the pixel is inside iff all
// three are >= 0 (the three sum to |S| > 0, so "all <= 0" cannot hold).  All
// quantities are integers below 2^34 and exact in double; the boundary from
// the reciprocal can be off by one next to an integer, which the exact test
// of the candidate and its neighbour corrects.
struct EdgeRows {
    double al[3], k0[3], k1[3], inv[3];
};

__device__ __forceinline__ EdgeRows edge_rows(const TriVerts& v)
{
    EdgeRows E;
    const double sg = v.S > 0 ? 1.0 : -1.0;
    for (int e = 0; e < 3; ++e) {
        const int n = e == 2 ? 0 : e + 1;
        const double ax = static_cast<double>(v.x[e]), ay = static_cast<double>(v.y[e]);
        const double bx = static_cast<double>(v.x[n]), by = static_cast<double>(v.y[n]);
        E.al[e] = sg * (ay - by);
        E.k0[e] = sg * (ax * by - bx * ay);
        E.k1[e] = sg * (bx - ax);
        // Hardware reciprocal estimate refined by two Newton steps (each
        // squares the relative error); row_span corrects the rounded bound
        // with exact integer tests, so only its estimate depends on this.
        if (E.al[e] != 0.0) {
            double r = __builtin_amdgcn_rcp(E.al[e]);
            r = fma(r, fma(-E.al[e], r, 1.0), r);
            E.inv[e] = fma(r, fma(-E.al[e], r, 1.0), r);
        } else {
            E.inv[e] = 0.0;
        }
    }
    return E;
}

__device__ __forceinline__ void row_span(const EdgeRows& E, double y, double& lo, double& hi)
{
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double al = E.al[e];
        const double be = fma(E.k1[e], y, E.k0[e]);
        if (al > 0.0) {   // px >= -be/al
            double x = ceil(-be * E.inv[e]);
            if (fma(al, x, be) < 0.0) x += 1.0;
            else if (fma(al, x - 1.0, be) >= 0.0) x -= 1.0;
            lo = fmax(lo, x);
        } else if (al < 0.0) {   // px <= -be/al
            double x = floor(-be * E.inv[e]);
            if (fma(al, x, be) < 0.0) x -= 1.0;
            else if (fma(al, x + 1.0, be) >= 0.0) x += 1.0;
            hi = fmin(hi, x);
        } else if (be < 0.0) {
            lo = 1.0;
            hi = 0.0;
        }
    }
}

// A triangle of a tile's raster round (see tile_assign_kernel).
struct RasterEntry {
    uint16_t x[3], y[3];   // vertices (image positions, below 2^15)
    int32_t ya;            // first row inside the tile
    int32_t id;            // the image's own tid
};

// Exclusive scan of one value per thread of a 256-thread block into
// pref[0 .. 256); pref[256] and the return value are the total.  tmp: 4 words.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* pref, uint32_t* tmp)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) tmp[wv] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t tw = tmp[w];
        if (w < wv) before += tw;
        total += tw;
    }
    pref[threadIdx.x] = before + incl - v;
    if (threadIdx.x == 0) pref[256] = total;
    __syncthreads();
    return total;
}

// Binning of triangles to 64x64 tiles.  Triangles are taken in chunks of
// 256 consecutive tids; consecutive tids are neighbours along a scan row, so a
// chunk touches a few dozen tiles.  A chunk first aggregates its (tile, count)
// pairs in an LDS hash table and then makes one global atomic per distinct
// tile, instead of one per (triangle, tile) pair from thousands of threads.
// A triangle whose box covers more than kBinDirect tiles (long, thin or
// collinear ones) goes straight to the global counters.
constexpr int kBinSlots = 1024;   // LDS table: at most 256 triangles x 4 tiles
constexpr int kBinDirect = 4;
constexpr uint32_t kBinEmpty = 0xFFFFFFFFu;

constexpr int kBinCount = 0, kBinFill = 1, kBinFixed = 2;   // bin_kernel modes (below)

// A triangle whose box covers more than kBinWide tiles (the long edges
// between clusters of a clustered seed set reach thousands) is queued, and the
// whole block then shares out its tiles: one thread walking 2,000 tiles with
// an atomic whose result it waits for took most of a 0.8 ms fill pass.
constexpr int kBinWide = 64;
constexpr int kBinQueue = 64;

struct BinLds {
    uint32_t key[kBinSlots];     // global tile index, kBinEmpty = free
    uint32_t count[kBinSlots];   // pairs of this chunk in the tile
    uint32_t base[kBinSlots];    // (fill) first list position of the chunk's pairs
    uint32_t wide_t[kBinQueue];  // wide triangles of the chunk
    int4 wide_box[kBinQueue];    // their tile boxes (tx0, ty0, tx1, ty1)
    uint32_t wide_off[kBinQueue];   // their image's first tile
    int wide_tx[kBinQueue];         // ... and tiles per row
    int wide_n;
};

// Slot of `tile` in the chunk table (inserting it); its count is incremented
// and the previous count (the pair's rank within the chunk) returned in rank.
__device__ __forceinline__ int bin_slot(BinLds& L, uint32_t tile, uint32_t& rank)
{
    uint32_t h = (tile * 2654435761u) & (kBinSlots - 1);
    for (int probe = 0; probe < kBinSlots; ++probe) {
        const uint32_t k = atomicCAS(&L.key[h], kBinEmpty, tile);
        if (k == kBinEmpty || k == tile) {
            rank = atomicAdd(&L.count[h], 1u);
            return static_cast<int>(h);
        }
        h = (h + 1) & (kBinSlots - 1);
    }
    return -1;   // unreachable: at most kBinSlots distinct tiles per chunk
}

__device__ __forceinline__ void bin_clear(BinLds& L)
{
    for (int i = threadIdx.x; i < kBinSlots; i += blockDim.x) {
        L.key[i] = kBinEmpty;
        L.count[i] = 0u;
    }
    if (threadIdx.x == 0) L.wide_n = 0;
}

// One (triangle, tile) pair written straight to the global bins.
template <int MODE>
__device__ __forceinline__ void bin_direct(uint32_t tile, uint32_t t, uint32_t* __restrict__ counts,
                                           const uint32_t* __restrict__ offsets,
                                           uint32_t* __restrict__ list, uint32_t cap,
                                           uint32_t* __restrict__ flags)
{
    if (MODE == kBinCount) {
        atomicAdd(&counts[tile], 1u);
    } else if (MODE == kBinFill) {
        const uint32_t k = offsets[tile] + atomicAdd(&counts[tile], 1u);
        if (k < cap) list[k] = t;
        else atomicOr(flags, kFlagOverflow);
    } else {
        const uint32_t k = atomicAdd(&counts[tile], 1u);
        if (k < cap) list[static_cast<size_t>(tile) * cap + k] = t;
        else atomicOr(flags, kFlagBinOverflow);
    }
}

// Triangles t < min(device count, rows_cap) (the rows buffer's capacity).
// kBinCount: counts[tile] += pairs.  kBinFill: list positions from
// offsets[tile] + cursor[tile] (counts is the cursor); entries beyond cap are
// dropped and flagged (speculative sizing).  kBinFixed: tile t owns list
// entries [t * cap, (t + 1) * cap) and counts[t] is its cursor; entries past
// a tile's capacity are dropped and flagged with kFlagBinOverflow (only the
// assignment needs to run again).

template <int MODE>
__global__ __launch_bounds__(256) void bin_kernel(
    const int32_t* __restrict__ rows, int64_t rows_cap, const int64_t* __restrict__ tri_base,
    int n_img, const ImageDesc* __restrict__ D, const uint32_t* __restrict__ sorted,
    uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets,
    uint32_t* __restrict__ list, uint32_t cap, uint32_t* __restrict__ flags)
{
    __shared__ BinLds L;
    const int64_t T = min(tri_base[n_img], rows_cap);
    bin_clear(L);
    __syncthreads();
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < T;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {   // uniform across the block
        const int64_t t = base + threadIdx.x;
        int slot[kBinDirect];
        uint32_t rank[kBinDirect];
        int n_local = 0;
        if (t < T) {
            const int img = n_img > 1 ? find_seg(tri_base, n_img, t) : 0;
            const ImageDesc& d = D[img];
            const TriVerts v = load_tri(rows, sorted + d.seed_off, t);
            int x0, y0, x1, y1;
            if (tri_box(v, d.W, d.H, x0, y0, x1, y1)) {
                const int tx0 = x0 / kTS, tx1 = x1 / kTS, ty0 = y0 / kTS, ty1 = y1 / kTS;
                const int nbox = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
                int q = -1;
                if (nbox > kBinWide) {
                    q = atomicAdd(&L.wide_n, 1);
                    if (q < kBinQueue) {
                        L.wide_t[q] = static_cast<uint32_t>(t);
                        L.wide_box[q] = make_int4(tx0, ty0, tx1, ty1);
                        L.wide_off[q] = static_cast<uint32_t>(d.tile_off);
                        L.wide_tx[q] = d.tiles_x;
                    }
                }
                const bool direct = nbox > kBinDirect;
                if (q < 0 || q >= kBinQueue)   // (a full queue: this thread walks them)
                    for (int ty = ty0; ty <= ty1; ++ty)
                        for (int tx = tx0; tx <= tx1; ++tx) {
                            const uint32_t tile =
                                static_cast<uint32_t>(d.tile_off + ty * d.tiles_x + tx);
                            if (direct) {
                                bin_direct<MODE>(tile, static_cast<uint32_t>(t), counts, offsets,
                                                 list, cap, flags);
                            } else {
                                slot[n_local] = bin_slot(L, tile, rank[n_local]);
                                ++n_local;
                            }
                        }
            }
        }
        __syncthreads();
        // the chunk's wide triangles, their tiles shared out over the block
        const int nw = min(L.wide_n, kBinQueue);
        for (int q = 0; q < nw; ++q) {
            const int4 bx = L.wide_box[q];
            const int bw = bx.z - bx.x + 1;
            const int nbox = bw * (bx.w - bx.y + 1);
            for (int k = threadIdx.x; k < nbox; k += blockDim.x) {
                const uint32_t tile = L.wide_off[q] +
                                      static_cast<uint32_t>((bx.y + k / bw) * L.wide_tx[q] + bx.x + k % bw);
                bin_direct<MODE>(tile, L.wide_t[q], counts, offsets, list, cap, flags);
            }
        }
        // one global atomic per distinct tile of the chunk
        for (int i = threadIdx.x; i < kBinSlots; i += blockDim.x) {
            const uint32_t tile = L.key[i];
            if (tile == kBinEmpty) continue;
            const uint32_t c = atomicAdd(&counts[tile], L.count[i]);
            if (MODE == kBinFill) L.base[i] = offsets[tile] + c;
            if (MODE == kBinFixed) L.base[i] = c;
        }
        __syncthreads();
        if (MODE != kBinCount) {
            for (int j = 0; j < n_local; ++j) {
                if (slot[j] < 0) continue;
                const uint32_t k = L.base[slot[j]] + rank[j];
                if (MODE == kBinFill) {
                    if (k < cap) list[k] = static_cast<uint32_t>(t);
                    else atomicOr(flags, kFlagOverflow);
                } else {
                    if (k < cap) list[static_cast<size_t>(L.key[slot[j]]) * cap + k] = static_cast<uint32_t>(t);
                    else atomicOr(flags, kFlagBinOverflow);
                }
            }
        }
        __syncthreads();
        bin_clear(L);
        __syncthreads();
    }
}

template <int C>
__global__ __launch_bounds__(256) void tile_assign_kernel(
    const int64_t* __restrict__ tile_seg, const int64_t* __restrict__ tri_base, int n_img,
    const ImageDesc* __restrict__ D, const int32_t* __restrict__ rows,
    const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ list, SeedIndex si, uint32_t cap, uint32_t n_tiles,
    const uint32_t* __restrict__ fixed_counts)
{
    __shared__ __align__(16) int32_t tid_tile[kTS * kTS];
    __shared__ __align__(16) RasterEntry ent[256];   // (also the output row buffers)
    static_assert(sizeof(RasterEntry) * 256 >= 4 * 3 * kTS * sizeof(int32_t), "row buffers");
    __shared__ uint32_t pref[257];
    __shared__ uint32_t scan_tmp[4];
    // Each XCD takes a contiguous range of tiles (tile rows of one image):
    // the seed-index words and label lines a tile row reads are then shared
    // within one L2.  The grid is a multiple of 8 (see xcd_range).
    const uint32_t per = gridDim.x >> 3;
    const int64_t gt = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (gt >= n_tiles) return;
    const int img = n_img > 1 ? find_seg(tile_seg, n_img, gt) : 0;
    const ImageDesc& d = D[img];
    const int W = d.W, H = d.H;
    const int tile = static_cast<int>(gt - d.tile_off);
    const int tx = tile % d.tiles_x, ty = tile / d.tiles_x;
    const int X0 = tx * kTS, Y0 = ty * kTS;
    const int X1 = min(X0 + kTS, W) - 1, Y1 = min(Y0 + kTS, H) - 1;
    const uint32_t* srt = sorted + d.seed_off;
    const int64_t t0 = tri_base[img];
    for (int i = threadIdx.x; i < kTS * kTS; i += blockDim.x) tid_tile[i] = -1;
    __syncthreads();
    // This tile's list entries: its fixed slice (fixed_counts given: cap per
    // tile) or its range from the offsets.  An overflowed list (speculative
    // sizing) is redone by the host.
    uint32_t b, e;
    if (fixed_counts) {
        b = static_cast<uint32_t>(gt) * cap;
        e = b + min(fixed_counts[gt], cap);
    } else {
        b = min(offsets[gt], cap);
        e = min(offsets[gt + 1], cap);
    }
    // The tile's triangles in rounds of one per thread.  A thread loads its
    // triangle; a collinear one (rare) it rasterises itself, otherwise it
    // records the vertices, first row and row count.  After a block scan of
    // the row counts every thread takes (triangle, row) pairs, so long and
    // short triangles spread evenly over the block.
    for (uint32_t base = b; base < e; base += blockDim.x) {   // block-uniform
        const uint32_t k = base + threadIdx.x;
        int nrows = 0;
        if (k < e) {
            const uint32_t t = list[k];
            const TriVerts v = load_tri(rows, srt, t);
            const int32_t id = static_cast<int32_t>(t - t0);   // the image's own tid
            if (v.S == 0) {
                const int64_t dx = v.x[1] - v.x[0], dy = v.y[1] - v.y[0];
                for (int y = Y0; y <= Y1; ++y) {
                    if (dy == 0) {
                        if (y != v.y[0]) continue;
                        for (int x = X0; x <= X1; ++x) atomicMax(&tid_tile[(y - Y0) * kTS + (x - X0)], id);
                        continue;
                    }
                    const int64_t num = (y - v.y[0]) * dx;
                    if (num % dy != 0) continue;
                    const int64_t x = v.x[0] + num / dy;
                    if (x >= X0 && x <= X1) atomicMax(&tid_tile[(y - Y0) * kTS + (x - X0)], id);
                }
            } else {
                const int ya = static_cast<int>(max<int64_t>(Y0, min(v.y[0], min(v.y[1], v.y[2]))));
                const int yb = static_cast<int>(min<int64_t>(Y1, max(v.y[0], max(v.y[1], v.y[2]))));
                nrows = max(0, yb - ya + 1);
                RasterEntry& re = ent[threadIdx.x];
                for (int c = 0; c < 3; ++c) {
                    re.x[c] = static_cast<uint16_t>(v.x[c]);
                    re.y[c] = static_cast<uint16_t>(v.y[c]);
                }
                re.ya = ya;
                re.id = id;
            }
        }
        const uint32_t total = block_exclusive_scan(static_cast<uint32_t>(nrows), pref, scan_tmp);
        // pref[i]: first pair of entry i (pref[blockDim.x] = total)
        for (uint32_t p = threadIdx.x; p < total; p += blockDim.x) {
            int lo_e = 0, hi_e = static_cast<int>(blockDim.x) - 1;   // last entry with pref <= p
            while (lo_e < hi_e) {
                const int mid = (lo_e + hi_e + 1) >> 1;
                if (pref[mid] <= p) lo_e = mid; else hi_e = mid - 1;
            }
            const RasterEntry& re = ent[lo_e];
            TriVerts v;
            for (int c = 0; c < 3; ++c) {
                v.x[c] = re.x[c];
                v.y[c] = re.y[c];
            }
            v.S = (v.x[1] - v.x[0]) * (v.y[2] - v.y[0]) - (v.y[1] - v.y[0]) * (v.x[2] - v.x[0]);
            const int y = re.ya + static_cast<int>(p - pref[lo_e]);
            const EdgeRows E = edge_rows(v);
            double lo = X0, hi = X1;
            row_span(E, static_cast<double>(y), lo, hi);
            const int rb = (y - Y0) * kTS - X0;
            const int l = static_cast<int>(lo), h = static_cast<int>(hi);
            for (int x = l; x <= h; ++x) atomicMax(&tid_tile[rb + x], re.id);
        }
        __syncthreads();   // the next round rewrites ent and pref
    }
    __syncthreads();
    const int Xw = X1 - X0 + 1;   // pixels per tile row
    if (C == 1) {
        // Each thread writes four consecutive pixels of a tile row (16 bytes).
        const int q = threadIdx.x & 15;
        const int x = X0 + 4 * q;
        for (int r = threadIdx.x >> 4; r < kTS; r += blockDim.x / 16) {
            const int y = Y0 + r;
            if (y > Y1 || x > X1) continue;
            const int64_t p = static_cast<int64_t>(y) * W + x;
            const int4 t4 = *reinterpret_cast<const int4*>(&tid_tile[r * kTS + 4 * q]);
            const int32_t tv[4] = {t4.x, t4.y, t4.z, t4.w};
            const int n = min(4, X1 - x + 1);
            // (named as global memory: a FLAT store would make the next row's
            // LDS read wait for it)
            __attribute__((address_space(1))) int32_t* dst =
                (__attribute__((address_space(1))) int32_t*)d.out + p;
            if (n == 4 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
                typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
                *reinterpret_cast<__attribute__((address_space(1))) i32x4*>(dst) =
                    i32x4{t4.x, t4.y, t4.z, t4.w};
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k < n) dst[k] = tv[k];
            }
        }
        return;
    }
    // C == 3.  Each wave takes kOutRows tile rows at a time, one pixel per
    // lane: the rows' label loads are issued together, then their owner-id
    // gathers, so a wave waits for memory twice per batch (occupancy here is
    // bound by LDS, not registers).  A row's 3 * Xw output ints are then staged in the wave's own
    // LDS row buffer (each lane writes its pixel's three channels) and written
    // as chunks of four by lanes 0..47 (chunk L = ints 4L .. 4L + 3), so every
    // store instruction covers one contiguous run of the row.  The buffer
    // reuses the raster entries, which are dead by now.
    // The label source and the output are named as global memory and the row
    // buffer as LDS: through generic pointers (the descriptor's, or a cast of
    // the shared array) every access would be a FLAT one, which counts as a
    // memory operation too, so each row's wait for its LDS buffer would also
    // wait for every output store issued before it.
    using GI32 = __attribute__((address_space(1))) int32_t;
    using GCU32 = const __attribute__((address_space(1))) uint32_t;
    using LI32 = __attribute__((address_space(3))) int32_t;
    typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int off = d.out_off, shift = d.out_shift, Wl = d.out_W;
    const int n_rows = Y1 - Y0 + 1;
    const int n_out = min(4, 3 * Xw - 4 * lane);            // ints of the chunk inside the row
    GCU32* out_lab = (GCU32*)d.out_lab;
    GI32* out_g = (GI32*)d.out;
    LI32* rowbuf = (LI32*)reinterpret_cast<int32_t*>(ent) + wv * (3 * kTS);   // 3 ints per pixel
    constexpr int kOutRows = kTS / 4;   // four waves cover the tile in one batch each
    for (int r0 = kOutRows * wv; r0 < n_rows; r0 += 4 * kOutRows) {
        uint32_t lab[kOutRows];
#pragma unroll
        for (int i = 0; i < kOutRows; ++i) {
            const int r = r0 + i;
            lab[i] = (r < n_rows && lane < Xw)
                         ? out_lab[static_cast<size_t>(Y0 + r + off) * Wl + off + X0 + lane]
                         : kInvalid;
        }
        int32_t id[kOutRows], dd[kOutRows];
#pragma unroll
        for (int i = 0; i < kOutRows; ++i) {
            id[i] = -1;
            dd[i] = -1;
            if (lab[i] != kInvalid) {
                const int lx = static_cast<int>(lab[i] >> 16), ly = static_cast<int>(lab[i] & 0xFFFFu);
                const int dx = X0 + lane + off - lx, dy = Y0 + r0 + i + off - ly;
                dd[i] = dx * dx + dy * dy;
                id[i] = seed_id_at(si, d.bm_off, d.seed_off, H, lx - shift, ly - shift);
            }
        }
#pragma unroll
        for (int i = 0; i < kOutRows; ++i) {
            const int r = r0 + i;
            if (r >= n_rows) break;   // uniform across the wave
            const int32_t tid = tid_tile[r * kTS + lane];
            // LDS operations of one wave complete in order: the chunk reads
            // below see these writes, and the next row's writes come after
            // the reads.
            rowbuf[3 * lane] = id[i];
            rowbuf[3 * lane + 1] = dd[i];
            rowbuf[3 * lane + 2] = tid;
            __builtin_amdgcn_wave_barrier();
            if (n_out <= 0) continue;
            const i32x4 o = *reinterpret_cast<const __attribute__((address_space(3))) i32x4*>(rowbuf + 4 * lane);
            GI32* dst = out_g + 3 * (static_cast<int64_t>(Y0 + r) * W + X0) + 4 * lane;
            if (n_out == 4 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
                *reinterpret_cast<__attribute__((address_space(1))) i32x4*>(dst) = o;
            } else {
                dst[0] = o.x;
                if (n_out > 1) dst[1] = o.y;
                if (n_out > 2) dst[2] = o.z;
                if (n_out > 3) dst[3] = o.w;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

}  // namespace

namespace {
// Bin-list entries per tile for the speculative fixed-slice layout: three
// times the triangles a tile of the densest image expects (2 per seed, by
// area), plus a margin.  0 when that layout would take much more memory than
// the exact one (very dense or tiny images): those use count + scan + fill.
uint32_t fixed_bin_capacity(const BatchPlan& plan, int64_t t_bound)
{
    double per_tile = 0.0;
    for (const ImageDesc& d : plan.img) {
        const double area = static_cast<double>(d.W) * d.H;
        const double tile = static_cast<double>(std::min(d.W, kTS)) * std::min(d.H, kTS);
        per_tile = std::max(per_tile, 2.0 * static_cast<double>(d.n) * tile / area);
    }
    const double capd = std::ceil((3.0 * per_tile + 128.0) / 32.0) * 32.0;
    if (capd > 4096.0) return 0;
    const uint32_t cap = static_cast<uint32_t>(capd);
    const double entries = static_cast<double>(cap) * static_cast<double>(plan.n_tiles);
    // (slice offsets are 32-bit in the kernels)
    if (entries > 16.0 * static_cast<double>(t_bound) + (1 << 22) || entries >= 2147483648.0)
        return 0;
    return cap;
}
}  // namespace

void assign_tiled(const BatchPlan& plan, const int32_t* rows, int64_t n_tri,
                  const uint32_t* sorted, int C, TriWork& work,
                  hipStream_t s, bool speculative, bool allow_fixed, bool counts_zeroed)
{
    const int n_img = static_cast<int>(plan.img.size());
    const int64_t ntiles = plan.n_tiles;
    work.list_cap = 0;
    if (ntiles == 0) return;
    work.bin_counts.reserve(sizeof(uint32_t) * static_cast<size_t>(2 * ntiles + 1));
    work.bin_offsets.reserve(sizeof(uint32_t) * static_cast<size_t>(ntiles + 1));
    work.host_flags.reserve(sizeof(uint32_t) * 8);
    uint32_t* counts = work.bin_counts.as<uint32_t>();
    uint32_t* cursor = counts + ntiles;
    uint32_t* offs = work.bin_offsets.as<uint32_t>();
    const ImageDesc* D = work.desc.as<ImageDesc>();
    const int64_t* tri_base = seg_ptr(work, 3, n_img);
    const int64_t* tile_seg = seg_ptr(work, 2, n_img);
    // Triangle count: exact, or on the device (speculative: bounded by rows_cap).
    const int64_t rows_cap = static_cast<int64_t>(work.rows_cap);
    const int64_t t_bound = n_tri >= 0 ? std::min(n_tri, rows_cap) : rows_cap;
    const unsigned tblocks = static_cast<unsigned>(std::max<int64_t>(
        1, std::min<int64_t>((t_bound + 255) / 256, 8192)));
    // Speculative: one pass into fixed per-tile slices when they fit (no
    // count, no scan); an overflowing tile raises kFlagOverflow and the host
    // reruns exactly.
    const uint32_t fixed = speculative && allow_fixed
                               ? (work.spec_tiny == 1 ? 1u : fixed_bin_capacity(plan, t_bound))
                               : 0u;
    int64_t cap;
    if (fixed) {
        cap = fixed;
        work.bin_list.reserve(sizeof(uint32_t) * static_cast<size_t>(ntiles) * fixed);
        ProfScope ps(kKAssign, s);
        if (!counts_zeroed) fill_u32(counts, 0u, static_cast<size_t>(ntiles), s);
        if (t_bound > 0) {
            bin_kernel<kBinFixed><<<tblocks, 256, 0, s>>>(rows, rows_cap, tri_base, n_img, D, sorted,
                                                          counts, nullptr, work.bin_list.as<uint32_t>(),
                                                          fixed, work.flags.as<uint32_t>());
            DL_LAUNCH_CHECK();
        }
    } else {
        {
            ProfScope ps(kKAssign, s);
            if (!counts_zeroed) fill_u32(counts, 0u, static_cast<size_t>(2 * ntiles), s);
            if (t_bound > 0) {
                bin_kernel<kBinCount><<<tblocks, 256, 0, s>>>(rows, rows_cap, tri_base, n_img, D,
                                                              sorted, counts, nullptr, nullptr, 0u,
                                                              nullptr);
                DL_LAUNCH_CHECK();
            }
            exclusive_scan_u32(counts, ntiles, offs, work.scan_tmp, s);
        }
        if (speculative) {
            // a small triangle's box covers up to four tiles
            cap = work.spec_tiny == 1 ? 1 : 4 * t_bound + ntiles + 1024;
        } else {
            DL_HIP(hipMemcpyAsync(work.host_flags.ptr, offs + ntiles, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
            DL_HIP(hipStreamSynchronize(s));
            cap = std::max<int64_t>(work.host_flags.as<uint32_t>()[0], 1);
        }
        cap = std::min<int64_t>(cap, 0xFFFFFFF0ll);
        work.list_cap = cap;   // the status read checks the list's total against it
        work.bin_list.reserve(sizeof(uint32_t) * static_cast<size_t>(cap));
        ProfScope ps(kKAssign, s);
        if (t_bound > 0) {
            bin_kernel<kBinFill><<<tblocks, 256, 0, s>>>(rows, rows_cap, tri_base, n_img, D, sorted,
                                                         cursor, offs, work.bin_list.as<uint32_t>(),
                                                         static_cast<uint32_t>(cap),
                                                         work.flags.as<uint32_t>());
            DL_LAUNCH_CHECK();
        }
    }
    ProfScope ps(kKOutput, s);
    const unsigned tgrid = static_cast<unsigned>((ntiles + 7) / 8 * 8);
    const uint32_t* fixed_counts = fixed ? counts : nullptr;
    if (C == 3)
        tile_assign_kernel<3><<<tgrid, 256, 0, s>>>(
            tile_seg, tri_base, n_img, D, rows, sorted, offs, work.bin_list.as<uint32_t>(),
            seed_index(work), static_cast<uint32_t>(cap), static_cast<uint32_t>(ntiles),
            fixed_counts);
    else
        tile_assign_kernel<1><<<tgrid, 256, 0, s>>>(
            tile_seg, tri_base, n_img, D, rows, sorted, offs, work.bin_list.as<uint32_t>(),
            seed_index(work), static_cast<uint32_t>(cap), static_cast<uint32_t>(ntiles),
            fixed_counts);
    DL_LAUNCH_CHECK();
}

}  // namespace dl
