// Queries on a finalised session: point location, the in-circumsphere
// predicate lifted by t, the undirected edge list, edge scores over the
// per-seed values, best-k midpoint selection, and the split of the result grid
// into per-pixel device arrays.
//
// Reference semantics (src/cuda): locate / in_circumsphere follow
// delaunay_query.cu:33-133; get_edges / edge_scores / select_midpoints follow
// delaunay_scalar_field.cu:35-312; the grid split follows
// delaunay_output.cu:119-151.  The differences of layout: triangles are the
// session's rows (x, y, a, b, c) in sorted seed ids, coordinates come from the
// session's sorted labels (x << 16 | y), and the per-pixel result is the
// (H, W, 3) grid of the last finalise, so location inside the image is a
// gather of its tid channel -- the closed-triangle max-tid rule of the batch
// pipeline (src/delauney/reference/triangulation.py:193-204), which the
// reference's locate is tested against (tests/test_scalar_field.py:252-258).
#include "dl_internal.h"

#include <algorithm>

namespace dl {
namespace {

constexpr int32_t kNoTri = -1;
constexpr int32_t kRing = -2;   // in the padding ring: resolved by a scan over triangles
constexpr uint64_t kDeadKey = ~0ull;

unsigned blocks_of(int64_t n, int t = 256)
{
    return static_cast<unsigned>(std::max<int64_t>(1, (n + t - 1) / t));
}

int bits_for(uint64_t v)   // bits needed to hold values 0..v
{
    int b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

__device__ __forceinline__ void tri_xy(const int32_t* __restrict__ rows,
                                       const uint32_t* __restrict__ sorted, int64_t t,
                                       int64_t X[3], int64_t Y[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t l = sorted[rows[5 * t + 2 + k]];
        X[k] = static_cast<int64_t>(l >> 16);
        Y[k] = static_cast<int64_t>(l & 0xFFFFu);
    }
}

// Closed containment of (px, py) in triangle (X, Y), exact in integers
// (triangulation.py:226-253): not both a negative and a positive edge cross
// product.  A collinear triple contains exactly the points of its line.
__device__ __forceinline__ bool contains(int64_t px, int64_t py, const int64_t X[3],
                                         const int64_t Y[3])
{
    bool neg = false, pos = false;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int n = e == 2 ? 0 : e + 1;
        const int64_t d = (X[e] - px) * (Y[n] - py) - (Y[e] - py) * (X[n] - px);
        neg |= d < 0;
        pos |= d > 0;
    }
    return !(neg && pos);
}

// In-image points read the grid's tid channel; ring points (inside the padded
// canvas, outside the image) are flagged for the triangle scan; anything
// farther out has no triangle (delaunay_query.cu:128-131).
__global__ void locate_gather_kernel(const int32_t* __restrict__ q, int64_t n, int W, int H, int P,
                                     const int32_t* __restrict__ grid, int32_t* __restrict__ tid,
                                     uint32_t* __restrict__ ring)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = q[2 * i], y = q[2 * i + 1];
    int32_t t = kNoTri;
    uint32_t r = 0;
    if (x >= 0 && x < W && y >= 0 && y < H) {
        t = grid ? grid[(static_cast<int64_t>(y) * W + x) * 3 + 2] : kNoTri;
    } else if (x >= -P && x < W + P && y >= -P && y < H + P) {
        t = kRing;
        r = 1;
    }
    tid[i] = t;
    ring[i] = r;
}

__global__ void compact_index_kernel(const uint32_t* __restrict__ flag,
                                     const uint32_t* __restrict__ pos, int64_t n,
                                     uint32_t* __restrict__ list)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) list[pos[i]] = static_cast<uint32_t>(i);
}

// One block per ring point at a time: the max tid over all triangles that
// contain it.  *count (device) is the number of ring points.
__global__ __launch_bounds__(256) void locate_ring_kernel(
    const int32_t* __restrict__ q, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count, const int32_t* __restrict__ rows, int64_t T,
    const uint32_t* __restrict__ sorted, int32_t* __restrict__ tid)
{
    __shared__ int32_t red[4];
    const uint32_t m = *count;
    for (uint32_t k = blockIdx.x; k < m; k += gridDim.x) {
        const uint32_t i = list[k];
        const int64_t px = q[2 * i], py = q[2 * i + 1];
        int32_t best = kNoTri;
        for (int64_t t = threadIdx.x; t < T; t += blockDim.x) {
            int64_t X[3], Y[3];
            tri_xy(rows, sorted, t, X, Y);
            if (contains(px, py, X, Y)) best = static_cast<int32_t>(t);   // t ascends per thread
        }
        for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_down(best, off, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
        __syncthreads();
        if (threadIdx.x == 0) tid[i] = max(max(red[0], red[1]), max(red[2], red[3]));
        __syncthreads();
    }
}

// dx^2 + dy^2 + t^2 < R^2 about the circumcentre of the containing triangle,
// in double, in the operation order of delaunay_query.cu:58-100 (the
// circumcentre as A + (|AC|^2 (N x AB) + |AB|^2 (AC x N)) / (2 |N|^2), the
// radius as the farthest vertex, distances compared after the square root).
// Collinear triangles (nz^2 < 1e-12) use the centroid.  Coordinates are taken
// on the padded canvas (image + P), as the reference holds its seeds, since the
// rounding depends on the origin.  No contraction into fused multiply-adds, so
// each operation rounds as written.
__global__ void circumsphere_kernel(const int32_t* __restrict__ q, const double* __restrict__ tq,
                                    int64_t n, int P, const int32_t* __restrict__ tid,
                                    const int32_t* __restrict__ rows,
                                    const uint32_t* __restrict__ sorted,
                                    uint8_t* __restrict__ mask)
{
#pragma clang fp contract(off)
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t t = tid[i];
    if (t < 0) {
        mask[i] = 0;
        return;
    }
    int64_t X[3], Y[3];
    tri_xy(rows, sorted, t, X, Y);
    const double ax = static_cast<double>(X[0] + P), ay = static_cast<double>(Y[0] + P);
    const double bx = static_cast<double>(X[1] + P), by = static_cast<double>(Y[1] + P);
    const double cx = static_cast<double>(X[2] + P), cy = static_cast<double>(Y[2] + P);
    const double abx = bx - ax, aby = by - ay;
    const double acx = cx - ax, acy = cy - ay;
    const double nz = abx * acy - aby * acx;
    double ox, oy;
    if (nz * nz < 1e-12) {
        ox = (ax + bx + cx) / 3.0;
        oy = (ay + by + cy) / 3.0;
    } else {
        const double ab2 = abx * abx + aby * aby;
        const double ac2 = acx * acx + acy * acy;
        const double den = 2.0 * (nz * nz);
        const double num_x = ac2 * (-(nz * aby)) + ab2 * (acy * nz);
        const double num_y = ac2 * (nz * abx) + ab2 * (-(acx * nz));
        ox = ax + num_x / den;
        oy = ay + num_y / den;
    }
    const double ra = (ax - ox) * (ax - ox) + (ay - oy) * (ay - oy);
    const double rb = (bx - ox) * (bx - ox) + (by - oy) * (by - oy);
    const double rc = (cx - ox) * (cx - ox) + (cy - oy) * (cy - oy);
    const double r2 = fmax(fmax(ra, rb), rc);
    const double dx = static_cast<double>(q[2 * i] + P) - ox;
    const double dy = static_cast<double>(q[2 * i + 1] + P) - oy;
    const double tt = tq[i];
    mask[i] = sqrt(dx * dx + dy * dy + tt * tt) < sqrt(r2) ? 1 : 0;
}

// sorted id -> insertion index of each live seed: its rank in the presence
// bitmap is its sorted id.
__global__ void insertion_of_sorted_kernel(const int32_t* __restrict__ list, int64_t n, int H,
                                           const uint32_t* __restrict__ bm,
                                           const uint32_t* __restrict__ pre,
                                           int32_t* __restrict__ ins)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t bit = static_cast<uint64_t>(list[2 * i]) * H + list[2 * i + 1];
    const uint32_t below = bm[bit >> 5] & ((1u << (bit & 31)) - 1u);
    ins[pre[bit >> 5] + __popc(below)] = static_cast<int32_t>(i);
}

// Three (min, max) edge keys per triangle, in insertion ids, packed as
// min << b | max so that key order is (min, max) order
// (delaunay_scalar_field.cu:35-60).
__global__ void edge_keys_kernel(const int32_t* __restrict__ rows, int64_t T,
                                 const int32_t* __restrict__ ins, int b,
                                 uint64_t* __restrict__ keys)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int32_t v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = ins[rows[5 * t + 2 + k]];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int32_t p = v[e], r = v[e == 2 ? 0 : e + 1];
        const uint64_t lo = static_cast<uint64_t>(min(p, r)), hi = static_cast<uint64_t>(max(p, r));
        keys[3 * t + e] = (lo << b) | hi;
    }
}

// flag[i] = 1 for the first of each run of equal keys.
__global__ void first_of_run_kernel(const uint64_t* __restrict__ k, int64_t n,
                                    uint32_t* __restrict__ flag)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

__global__ void edge_unpack_kernel(const uint64_t* __restrict__ k, const uint32_t* __restrict__ flag,
                                   const uint32_t* __restrict__ pos, int64_t n, int b,
                                   int32_t* __restrict__ edges)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint64_t key = k[i];
    const uint32_t p = pos[i];
    edges[2 * static_cast<int64_t>(p)] = static_cast<int32_t>(key >> b);
    edges[2 * static_cast<int64_t>(p) + 1] = static_cast<int32_t>(key & ((1ull << b) - 1));
}

// |v_a - v_b| * |ab| in double, rounded once to float; 0 when
// |ab|^2 < min_len2 (delaunay_scalar_field.cu:95-113).
__global__ void edge_score_kernel(const int32_t* __restrict__ edges, int64_t E,
                                  const int32_t* __restrict__ list, const float* __restrict__ val,
                                  double min_len2, float* __restrict__ score)
{
#pragma clang fp contract(off)
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int32_t a = edges[2 * e], b = edges[2 * e + 1];
    const double dx = static_cast<double>(list[2 * a]) - static_cast<double>(list[2 * b]);
    const double dy = static_cast<double>(list[2 * a + 1]) - static_cast<double>(list[2 * b + 1]);
    const double len2 = dx * dx + dy * dy;
    if (len2 < min_len2) {
        score[e] = 0.0f;
        return;
    }
    const double dv = fabs(static_cast<double>(val[a]) - static_cast<double>(val[b]));
    score[e] = static_cast<float>(dv * sqrt(len2));
}

// Best-first key (delaunay_scalar_field.cu:124-139): the complemented bits of
// a positive float order descending, the edge index breaks ties ascending;
// ineligible edges (score <= threshold or <= 0) get the high half kIneligible,
// which no eligible edge has (it would need a score of +0), so after sorting
// the eligible edges are exactly the keys before the first such one.
constexpr uint32_t kIneligible = 0xFFFFFFFFu;

__global__ void score_key_kernel(const float* __restrict__ score, int64_t E, float threshold,
                                 uint64_t* __restrict__ keys)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float s = score[e];
    const uint32_t hi = (s > threshold && s > 0.0f) ? ~__float_as_uint(s) : kIneligible;
    keys[e] = (static_cast<uint64_t>(hi) << 32) | static_cast<uint32_t>(e);
}

// Midpoint of the rank-r edge for r < take = min(count, eligible), rounded
// half to even on the padded canvas (as the reference holds its seeds: with an
// odd P a tie rounds the other way than in image coordinates) and clamped to
// the image, keyed (pixel << 32 | r); dropped (kDeadKey) where a live seed sits
// (delaunay_scalar_field.cu:145-171).
__global__ void midpoint_kernel(const uint64_t* __restrict__ sorted_keys, int64_t cap,
                                const int32_t* __restrict__ edges, const int32_t* __restrict__ list,
                                int W, int H, int P, const uint32_t* __restrict__ bm,
                                uint64_t* __restrict__ out)
{
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= cap) return;   // cap = min(count, edges)
    const uint64_t key = sorted_keys[r];
    if ((key >> 32) == kIneligible) {   // r >= number of eligible edges
        out[r] = kDeadKey;
        return;
    }
    const int64_t e = static_cast<int64_t>(key & 0xFFFFFFFFu);
    const int32_t a = edges[2 * e], b = edges[2 * e + 1];
    int mx = static_cast<int>(
        rint(0.5 * (static_cast<double>(list[2 * a] + P) + (list[2 * b] + P))));
    int my = static_cast<int>(
        rint(0.5 * (static_cast<double>(list[2 * a + 1] + P) + (list[2 * b + 1] + P))));
    mx = min(max(mx, P), W - 1 + P) - P;
    my = min(max(my, P), H - 1 + P) - P;
    const uint64_t bit = static_cast<uint64_t>(mx) * H + my;
    if ((bm[bit >> 5] >> (bit & 31)) & 1u) {
        out[r] = kDeadKey;
        return;
    }
    const uint64_t pixel = static_cast<uint64_t>(my) * W + mx;
    out[r] = (pixel << 32) | static_cast<uint64_t>(r);
}

// First key of each pixel (the lowest rank), dead keys excluded.
__global__ void first_of_pixel_kernel(const uint64_t* __restrict__ k, int64_t n,
                                      uint32_t* __restrict__ flag)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = k[i];
    flag[i] = (key != kDeadKey && (i == 0 || (key >> 32) != (k[i - 1] >> 32))) ? 1u : 0u;
}

__global__ void midpoint_out_kernel(const uint64_t* __restrict__ k, const uint32_t* __restrict__ flag,
                                    const uint32_t* __restrict__ pos, int64_t n, int W,
                                    int32_t* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint64_t pixel = k[i] >> 32;
    const uint32_t p = pos[i];
    out[2 * static_cast<int64_t>(p)] = static_cast<int32_t>(pixel % static_cast<uint64_t>(W));
    out[2 * static_cast<int64_t>(p) + 1] = static_cast<int32_t>(pixel / static_cast<uint64_t>(W));
}

// (H, W, 3) grid -> tid (0 outside every triangle), seed id, outside flag
// (delaunay_output.cu:119-151).  Four pixels per thread: three 16-byte loads,
// two 16-byte stores and one 4-byte store.
__device__ __forceinline__ void split_one(int32_t sid, int32_t t, int32_t& tid, int32_t& sd,
                                          uint32_t& m)
{
    sd = sid;
    tid = t == kNoTri ? 0 : t;
    m = t == kNoTri ? 1u : 0u;
}

__global__ void split_kernel(const int32_t* __restrict__ grid, int64_t n, int32_t* __restrict__ tids,
                             int32_t* __restrict__ sids, uint8_t* __restrict__ mask)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t i0 = 4 * q;
    if (i0 >= n) return;
    if (i0 + 4 <= n) {
        const int4* g = reinterpret_cast<const int4*>(grid + 3 * i0);
        const int4 a = g[0], b = g[1], c = g[2];
        // pixels: (a.x a.y a.z) (a.w b.x b.y) (b.z b.w c.x) (c.y c.z c.w)
        int4 t, sd;
        uint32_t m0, m1, m2, m3;
        split_one(a.x, a.z, t.x, sd.x, m0);
        split_one(a.w, b.y, t.y, sd.y, m1);
        split_one(b.z, c.x, t.z, sd.z, m2);
        split_one(c.y, c.w, t.w, sd.w, m3);
        *reinterpret_cast<int4*>(tids + i0) = t;
        *reinterpret_cast<int4*>(sids + i0) = sd;
        *reinterpret_cast<uint32_t*>(mask + i0) = m0 | (m1 << 8) | (m2 << 16) | (m3 << 24);
        return;
    }
    for (int64_t i = i0; i < n; ++i) {
        int32_t t, sd;
        uint32_t m;
        split_one(grid[3 * i], grid[3 * i + 2], t, sd, m);
        tids[i] = t;
        sids[i] = sd;
        mask[i] = static_cast<uint8_t>(m);
    }
}

// ---------------------------------------------------------------------------
// LSD radix sort of 64-bit keys, 8 bits per pass.  A pass is a digit
// histogram per tile of kSortTile keys (digit-major, so that one exclusive
// scan gives every tile's start for every digit), the scan, and a stable
// scatter: a tile's keys are taken in rounds of 256 in their original order,
// and within a round a key's place among equal digits is its wave's match
// rank (eight ballots) plus the counts of the waves before it.
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 4;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kSortWaves = kSortThreads / 64;

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                                  int64_t n, int shift,
                                                                  uint32_t dmask, int tiles,
                                                                  uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0u;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortRounds; ++k) {
        const int64_t i = base + k * kSortThreads + tid;
        if (i < n) atomicAdd(&h[static_cast<uint32_t>(keys[i] >> shift) & dmask], 1u);
    }
    __syncthreads();
    hist[static_cast<int64_t>(tid) * tiles + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n, int shift,
    uint32_t dmask, int tiles, const uint32_t* __restrict__ start)
{
    __shared__ uint32_t run[256];                 // next output position per digit
    __shared__ uint32_t wcnt[kSortWaves][256];    // this round's count per wave and digit
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = start[static_cast<int64_t>(tid) * tiles + blockIdx.x];
    const uint64_t lt = (1ull << lane) - 1ull;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int k = 0; k < kSortRounds; ++k) {
        const int64_t i = base + k * kSortThreads + tid;
        const bool valid = i < n;
        const uint64_t key = valid ? in[i] : 0ull;
        const uint32_t d = static_cast<uint32_t>(key >> shift) & dmask;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t v = __ballot(bit);
            m &= bit ? v : ~v;
        }
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) wcnt[w][tid] = 0u;
        __syncthreads();
        const uint32_t rank = static_cast<uint32_t>(__popcll(m & lt));
        if (valid && rank == 0u) wcnt[wave][d] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            out[pos] = key;
        }
        __syncthreads();
        uint32_t add = 0u;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) add += wcnt[w][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// Ascending stable radix sort of n 64-bit keys on their bits [0, low_end)
// and [high_begin, high_end): in -> out (in is left as it was).  Keys whose
// other bits differ may come out in either order of those bits; callers pass
// the ranges their keys use.  tmp holds the second key buffer and the
// histograms.
void sort_keys(const uint64_t* in, uint64_t* out, int64_t n, int low_end, int high_begin,
               int high_end, DevBuf& tmp, DevBuf& scan_tmp, hipStream_t s)
{
    if (n <= 0) return;
    if (n > static_cast<int64_t>(INT32_MAX)) throw value_error("too many keys to sort");
    low_end = std::max(0, std::min(low_end, 64));
    high_begin = std::max(low_end, std::min(high_begin, 64));
    high_end = std::max(high_begin, std::min(high_end, 64));
    // passes: (shift, bits), up to 8 bits each, low range first
    int shifts[16], widths[16], passes = 0;
    for (int b = 0; b < low_end; b += 8) {
        shifts[passes] = b;
        widths[passes++] = std::min(8, low_end - b);
    }
    for (int b = high_begin; b < high_end; b += 8) {
        shifts[passes] = b;
        widths[passes++] = std::min(8, high_end - b);
    }
    if (passes == 0) {
        DL_HIP(hipMemcpyAsync(out, in, sizeof(uint64_t) * static_cast<size_t>(n),
                              hipMemcpyDeviceToDevice, s));
        return;
    }
    const int tiles = static_cast<int>((n + kSortTile - 1) / kSortTile);
    const size_t key_bytes = (sizeof(uint64_t) * static_cast<size_t>(n) + 255) / 256 * 256;
    const size_t hist_words = 256 * static_cast<size_t>(tiles) + 1;
    tmp.reserve(key_bytes + 2 * sizeof(uint32_t) * hist_words);
    uint64_t* alt = tmp.as<uint64_t>();
    uint32_t* hist = reinterpret_cast<uint32_t*>(static_cast<char*>(tmp.ptr) + key_bytes);
    uint32_t* start = hist + hist_words;
    // the last pass writes `out`
    const uint64_t* src = in;
    for (int p = 0; p < passes; ++p) {
        uint64_t* dst = ((passes - 1 - p) % 2 == 0) ? out : alt;
        const int shift = shifts[p];
        const uint32_t dmask = (1u << widths[p]) - 1u;
        radix_hist_kernel<<<tiles, kSortThreads, 0, s>>>(src, n, shift, dmask, tiles, hist);
        DL_LAUNCH_CHECK();
        exclusive_scan_u32(hist, 256 * static_cast<int64_t>(tiles), start, scan_tmp, s);
        radix_scatter_kernel<<<tiles, kSortThreads, 0, s>>>(src, dst, n, shift, dmask, tiles,
                                                            start);
        DL_LAUNCH_CHECK();
        src = dst;
    }
}

uint32_t read_u32(const uint32_t* d, QueryWork& w, hipStream_t s)
{
    w.host.reserve(64);
    DL_HIP(hipMemcpyAsync(w.host.ptr, d, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    return w.host.as<uint32_t>()[0];
}

}  // namespace

void query_points(const SessionView& v, const int32_t* d_q, const double* d_t, int64_t n,
                  int32_t* d_tid, uint8_t* d_mask, QueryWork& w, hipStream_t s)
{
    if (n <= 0) return;
    if (n >= static_cast<int64_t>(UINT32_MAX)) throw value_error("too many query points");
    w.c.reserve(sizeof(uint32_t) * static_cast<size_t>(n));
    w.d.reserve(sizeof(uint32_t) * static_cast<size_t>(n + 1));
    w.e.reserve(sizeof(uint32_t) * static_cast<size_t>(n));
    uint32_t* ring = w.c.as<uint32_t>();
    uint32_t* pos = w.d.as<uint32_t>();
    uint32_t* list = w.e.as<uint32_t>();
    const bool tris = v.n_tri > 0;
    // Without triangles every point, ring or not, has none.
    const int ring_width = tris ? v.P : 0;
    locate_gather_kernel<<<blocks_of(n), 256, 0, s>>>(d_q, n, v.W, v.H, ring_width,
                                                      tris ? v.grid : nullptr, d_tid, ring);
    DL_LAUNCH_CHECK();
    if (ring_width > 0) {
        exclusive_scan_u32(ring, n, pos, w.scan_tmp, s);
        compact_index_kernel<<<blocks_of(n), 256, 0, s>>>(ring, pos, n, list);
        DL_LAUNCH_CHECK();
        const unsigned g = static_cast<unsigned>(std::min<int64_t>(n, 2048));
        locate_ring_kernel<<<g, 256, 0, s>>>(d_q, list, pos + n, v.rows, v.n_tri, v.sorted, d_tid);
        DL_LAUNCH_CHECK();
    }
    if (d_t && d_mask) {
        circumsphere_kernel<<<blocks_of(n), 256, 0, s>>>(d_q, d_t, n, v.P, d_tid, v.rows, v.sorted,
                                                         d_mask);
        DL_LAUNCH_CHECK();
    }
}

int64_t build_edges(const SessionView& v, DevBuf& edges, QueryWork& w, hipStream_t s)
{
    const int64_t T = v.n_tri;
    if (T <= 0 || v.n_live <= 0) return 0;
    const int64_t nk = 3 * T;
    if (nk >= static_cast<int64_t>(UINT32_MAX)) throw value_error("too many triangles");
    // sorted id -> insertion index
    w.d.reserve(sizeof(uint32_t) * static_cast<size_t>(v.words + 1));
    w.e.reserve(sizeof(int32_t) * static_cast<size_t>(v.n_live));
    uint32_t* pre = w.d.as<uint32_t>();
    int32_t* ins = w.e.as<int32_t>();
    popcount_scan(v.bitmap, v.words, pre, w.scan_tmp, s);
    insertion_of_sorted_kernel<<<blocks_of(v.n_live), 256, 0, s>>>(v.list, v.n_live, v.H, v.bitmap,
                                                                    pre, ins);
    DL_LAUNCH_CHECK();
    // keys, sort, first of each run
    const int b = bits_for(static_cast<uint64_t>(v.n_live - 1));
    w.a.reserve(sizeof(uint64_t) * static_cast<size_t>(nk));
    w.b.reserve(sizeof(uint64_t) * static_cast<size_t>(nk));
    uint64_t* k0 = w.a.as<uint64_t>();
    uint64_t* k1 = w.b.as<uint64_t>();
    edge_keys_kernel<<<blocks_of(T), 256, 0, s>>>(v.rows, T, ins, b, k0);
    DL_LAUNCH_CHECK();
    sort_keys(k0, k1, nk, 2 * b, 64, 64, w.sort_tmp, w.scan_tmp, s);
    w.c.reserve(sizeof(uint32_t) * static_cast<size_t>(nk));
    w.d.reserve(sizeof(uint32_t) * static_cast<size_t>(nk + 1));
    uint32_t* flag = w.c.as<uint32_t>();
    pre = w.d.as<uint32_t>();
    first_of_run_kernel<<<blocks_of(nk), 256, 0, s>>>(k1, nk, flag);
    DL_LAUNCH_CHECK();
    exclusive_scan_u32(flag, nk, pre, w.scan_tmp, s);
    edges.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(nk));
    edge_unpack_kernel<<<blocks_of(nk), 256, 0, s>>>(k1, flag, pre, nk, b, edges.as<int32_t>());
    DL_LAUNCH_CHECK();
    return static_cast<int64_t>(read_u32(pre + nk, w, s));
}

void edge_scores(const SessionView& v, const int32_t* d_edges, int64_t n_edges, double min_length,
                 float* d_scores, hipStream_t s)
{
    if (n_edges <= 0) return;
    edge_score_kernel<<<blocks_of(n_edges), 256, 0, s>>>(d_edges, n_edges, v.list, v.values,
                                                         min_length * min_length, d_scores);
    DL_LAUNCH_CHECK();
}

int64_t select_midpoints(const SessionView& v, const int32_t* d_edges, int64_t n_edges,
                         double min_length, int64_t count, float threshold, int32_t* d_out,
                         QueryWork& w, hipStream_t s)
{
    if (n_edges <= 0 || count <= 0) return 0;
    const int64_t cap = std::min(count, n_edges);
    // scores, best-first keys, eligible count
    w.e.reserve(sizeof(float) * static_cast<size_t>(n_edges));
    float* score = w.e.as<float>();
    edge_scores(v, d_edges, n_edges, min_length, score, s);
    w.a.reserve(sizeof(uint64_t) * static_cast<size_t>(n_edges));
    w.b.reserve(sizeof(uint64_t) * static_cast<size_t>(n_edges));
    uint64_t* k0 = w.a.as<uint64_t>();
    uint64_t* k1 = w.b.as<uint64_t>();
    score_key_kernel<<<blocks_of(n_edges), 256, 0, s>>>(score, n_edges, threshold, k0);
    DL_LAUNCH_CHECK();
    // (key: score word above, edge index below)
    sort_keys(k0, k1, n_edges, bits_for(static_cast<uint64_t>(n_edges)), 32, 64, w.sort_tmp,
              w.scan_tmp, s);
    // midpoints of the first `take`, keyed by (pixel, rank), sorted
    midpoint_kernel<<<blocks_of(cap), 256, 0, s>>>(k1, cap, d_edges, v.list, v.W, v.H, v.P, v.bitmap,
                                                   k0);
    DL_LAUNCH_CHECK();
    // (key: pixel above, rank below; a dead key has every bit set, so it stays
    // above every live key on these bits too)
    sort_keys(k0, k1, cap, bits_for(static_cast<uint64_t>(cap)), 32,
              32 + bits_for(static_cast<uint64_t>(v.W) * static_cast<uint64_t>(v.H)), w.sort_tmp,
              w.scan_tmp, s);
    w.c.reserve(sizeof(uint32_t) * static_cast<size_t>(cap));
    w.d.reserve(sizeof(uint32_t) * static_cast<size_t>(cap + 1));
    uint32_t* flag = w.c.as<uint32_t>();
    uint32_t* pos = w.d.as<uint32_t>();
    first_of_pixel_kernel<<<blocks_of(cap), 256, 0, s>>>(k1, cap, flag);
    DL_LAUNCH_CHECK();
    exclusive_scan_u32(flag, cap, pos, w.scan_tmp, s);
    midpoint_out_kernel<<<blocks_of(cap), 256, 0, s>>>(k1, flag, pos, cap, v.W, d_out);
    DL_LAUNCH_CHECK();
    return static_cast<int64_t>(read_u32(pos + cap, w, s));
}

void split_grid(const int32_t* grid, int64_t n_px, int32_t* tids, int32_t* sids, uint8_t* mask,
                hipStream_t s)
{
    if (n_px <= 0) return;
    split_kernel<<<blocks_of((n_px + 3) / 4), 256, 0, s>>>(grid, n_px, tids, sids, mask);
    DL_LAUNCH_CHECK();
}

}  // namespace dl
