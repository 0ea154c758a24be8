// Voronoi flood fill on gfx950: an exact replay of the reference schedule.
//
// Reference: src/delauney/reference/voronoi.py:110-163.  Every iteration runs
// four passes with sources x-1, x+1, y-1, y+1 in that order; inside a pass each
// cell reads its source's pre-pass value (Jacobi), and the update is in place so
// later passes see it.  A cell takes the source's owner if it is undefined, or
// the owner is strictly closer, or equally close with a higher id.  The fill
// stops after the first iteration with no change.
//
// Why the tiled replay below is exact:
//  * One iteration reads a Chebyshev-radius-1 neighbourhood, so k iterations of
//    a tile with a halo of k cells reproduce the interior exactly.
//  * Every update strictly improves the cell in the order (defined, lower d,
//    higher id), so the state never revisits a value: a tile whose interior did
//    not change over a launch is a fixed point of its neighbourhood, and running
//    iterations past the reference's stop iteration changes nothing.
//  * A tile is recomputed only if a tile within one tile of it changed in the
//    previous launch; otherwise both label planes already hold its state.
//
// Layout: a tile is 128 columns wide (two columns per lane of a 64-wide wave)
// and waves*rows tall, with `iters` halo cells on each side.  Passes along x use
// DPP whole-wave lane shifts (wave_shr / wave_shl); passes along y run down a
// thread's registers and exchange one boundary row per wave through LDS.  The
// state per cell is its owner label plus the cached squared distance.
#include "dl_internal.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>

namespace dl {
namespace {

__device__ __forceinline__ int dist2_lab(int cx, int cy, uint32_t lab)
{
    const int dx = cx - static_cast<int>(lab >> 16);
    const int dy = cy - static_cast<int>(lab & 0xFFFFu);
    return dx * dx + dy * dy;
}

// voronoi.py:174-196 for the one-pass kernel: take the source's owner if this
// cell is undefined, or the owner is strictly closer, or equally close with a
// higher id (= higher label).
__device__ __forceinline__ bool relax(uint32_t& lab, int& d, uint32_t nl, int cx, int cy)
{
    const int c = dist2_lab(cx, cy, nl);
    const bool take = (nl != kInvalid) & ((c < d) | ((c == d) & (nl > lab)));
    lab = take ? nl : lab;
    d = take ? c : d;
    return take;
}

// The tiled kernel keeps each cell as a 64-bit key (d, ~label): the update
// rule above is then "take the candidate iff its key is smaller", one 64-bit
// compare.
//
// The squared distance comes from packed 16-bit arithmetic: a label is the
// pair (x, y) in its two halves, so with inv = ~label and
// cxy1 = ((cx + 1) << 16) | (cy + 1), the half-wise sum inv + cxy1 is
// (cx - x, cy - y), exact as int16 because canvas sides (with a tile halo) stay
// below 32767, and one dot2 gives dx^2 + dy^2 (below 2^31).
//
// Undefined and outside cells carry the far label inv = kFarInvWide =
// 0x7FFF7FFF, which no real label has (it would be x = 32768).  The half-wise
// sum saturates, so at any cell with cx, cy >= 0 that label is at
// (32767, 32767) relative offset: distance kFarD = 2 * 32767^2, above every
// real distance.  An undefined cell is the key (kFarD, kFarInvWide): a real
// candidate is smaller, and the far label itself ties, so it never updates
// anything.  A cell outside the canvas is (0, far label): no candidate is
// smaller (d = 0 would need a seed there), so it never changes and only ever
// offers the far label.  No per-update validity test is needed.
typedef short s16x2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kFarInvWide = 0x7FFF7FFFu;

// Second encoding, for canvases with (W-1)^2 + (H-1)^2 <= 32767^2 (sides up
// to 23,170): the key's smaller-of-two is one v_min_f64.  A key (d, inv) with
// d < 0x7FF00000 read as an IEEE double is a positive finite number (or a
// subnormal, which gfx950 keeps for f64), and positive finite doubles order as
// their bit patterns, so min_f64 of two keys is the smaller key, bit for bit.
// Every key and candidate must then stay finite: the far label becomes
// kFarInvNarrow = 0x7FFF0000 (x half saturates as before, y half adds cy + 1
// without saturating), whose distance at a canvas cell is
// 32767^2 + (cy + 1)^2: finite (below 0x7FF00000 for cy < 32752) and above
// every real distance on such canvases.  An undefined cell holds exactly
// that far candidate's key, so the far label still ties and never updates.
constexpr uint32_t kFarInvNarrow = 0x7FFF0000u;

template <bool F>
__device__ __forceinline__ constexpr uint32_t far_inv() { return F ? kFarInvNarrow : kFarInvWide; }

// Largest canvas side pair for the min encoding.
inline bool bfs_min_encoding_ok(int W, int H)
{
    const int64_t w = W - 1, h = H - 1;
    return w * w + h * h <= 32767LL * 32767LL;
}

__device__ __forceinline__ uint32_t key_dist(uint32_t ninv, uint32_t cxy1)
{
    const s16x2 dd = __builtin_elementwise_add_sat(__builtin_bit_cast(s16x2, ninv),
                                                   __builtin_bit_cast(s16x2, cxy1));
    // Clamping cannot change the sum (it stays below 2^31); it selects the
    // three-operand dot2 with an inline zero accumulator.
    return static_cast<uint32_t>(__builtin_amdgcn_sdot2(dd, dd, 0, true));
}

// A cell's key lives in one 64-bit register pair, so the compare needs no
// operand assembly; only the candidate pair (ninv, c) is formed.  F: the min
// encoding (compare, then one v_min_f64 instead of two selects).
template <bool F>
__device__ __forceinline__ bool relax_key(uint64_t& key, uint32_t ninv, uint32_t cxy1)
{
    const uint32_t c = key_dist(ninv, cxy1);
    const uint64_t ck = (static_cast<uint64_t>(c) << 32) | ninv;
    const bool take = ck < key;
    if constexpr (F) {
        double r;
        asm("v_min_f64 %0, %1, %2"
            : "=v"(r)
            : "v"(__builtin_bit_cast(double, key)), "v"(__builtin_bit_cast(double, ck)));
        key = __builtin_bit_cast(uint64_t, r);
    } else {
        key = take ? ck : key;
    }
    return take;
}

__device__ __forceinline__ uint32_t inv_of(uint64_t key) { return static_cast<uint32_t>(key); }

// Whole-wave shifts.  The lane without a source (0 for a right shift, 63 for
// a left shift) keeps `old`.  Callers chain the shifts of a pass through one
// variable that starts at the far label: that lane never receives data, so it
// holds the far label for every row and no per-row re-initialisation is needed.
__device__ __forceinline__ uint32_t from_left(uint32_t old, uint32_t v)   // lane l <- lane l-1
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
        static_cast<int>(old), static_cast<int>(v), 0x138, 0xF, 0xF, false));
}

__device__ __forceinline__ uint32_t from_right(uint32_t old, uint32_t v)  // lane l <- lane l+1
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
        static_cast<int>(old), static_cast<int>(v), 0x130, 0xF, 0xF, false));
}

// Escape watch for canvases whose image interior must not receive updates
// from the padding (see BfsCanvas::P).  In pass 1 the only source of a cell is
// its left neighbour, so a cell of the interior's first column changes in
// pass 1 exactly when a padding cell updates it; likewise the last column in
// pass 2, the first row in pass 3 and the last row in pass 4.  A watched tile
// tests the update flags of those cells in those passes.  The flags are lane
// masks already (one bit per lane), so the test is scalar work: the ballot of
// a pass's update flag ANDed with a mask of the lanes holding watched cells.
// Every mask is zero outside watched tiles.
using GU32 = __attribute__((address_space(1))) uint32_t;

// Tile change-mask bits: the interior changed somewhere / in its K-wide left,
// right, top or bottom band / in one of its K x K corners.
constexpr int kChAny = 1, kChLeft = 2, kChRight = 4, kChTop = 8, kChBottom = 16;
constexpr int kChTL = 32, kChTR = 64, kChBL = 128, kChBR = 256;
constexpr int kChMask = 511;
constexpr int kDiffers = 512;   // (replay) the result differs from the record

struct Watch {
    uint64_t l0, l1;     // lanes whose column 0 / 1 is the interior's first column
    uint64_t r0, r1;     // ... last column
    uint64_t i0, i1;     // lanes whose column 0 / 1 lies inside the interior
    uint32_t rows_in;    // bit r: row r of the wave lies inside the interior
    int r_top, r_bot;    // wave row of the interior's first / last row (-1: none)
};

// K iterations of the four passes on the tile held in registers.
//
// Work skipping, per wave (a band of R rows across the tile), exact by
// monotonicity:
//  * passes 1-2 (along x) read only the wave's own rows, so they cannot change
//    anything if the wave's state has not changed since they last ran without
//    change (chx and chy of the previous iteration both false);
//  * passes 3-4 (along y) read the wave's rows and one boundary row from each
//    neighbouring wave; they cannot change anything if, for this wave and both
//    neighbours, nothing changed since this wave's last passes 3-4 ran without
//    change (q = chy of the previous iteration or chx of this one, all false).
// A tile where no wave has q set is a fixed point: the remaining iterations
// are skipped.
//
// AW (aligned watch): every wave's rows lie wholly inside or wholly outside
// the watched interior's rows (bfs_watch_aligned), so the watch of passes 1-2
// needs no row condition (the lane masks of waves outside it are zero), and
// the interior's first / last row is a wave's first / last row: passes 3-4
// test that one row.  The general form keeps a word of row bits per pass.
template <int NW, int R, int K, bool F, bool AW = false>
__device__ __forceinline__ void iterate(uint64_t (&K0)[R], uint64_t (&K1)[R], int c0, int y0,
                                        const Watch& wt,
                                        uint32_t (*ldsA)[64][2], uint32_t (*ldsB)[64][2],
                                        int (*lds_q)[NW], int wave, int lane, int& esc)
{
    // packed (cx + 1, cy + 1) is (px | ((y0 + r + 1) & 0xFFFF)) for row r: the
    // halves are formed separately so that a negative halo row cannot carry
    // into the x half.
    const uint32_t px0 = static_cast<uint32_t>(c0 + 1) << 16;
    const uint32_t px1 = static_cast<uint32_t>(c0 + 2) << 16;
#define DL_P0(r) (px0 | (static_cast<uint32_t>(y0 + (r) + 1) & 0xFFFFu))
#define DL_P1(r) (px1 | (static_cast<uint32_t>(y0 + (r) + 1) & 0xFFFFu))
    // The watch records, per pass, which rows had a watched update (bit r + 1
    // of a 32-bit word: a constant select for every row, row 0 included) and
    // applies the row condition once per pass, so no per-row lane masks are
    // kept live across the loop.
    bool chx = true, chy = true;   // changes of the previous iteration (forced on at entry)
    for (int k = 0; k < K; ++k) {
        bool cx = false;
        if (chx || chy) {   // wave-uniform
            bool ch = false;
            uint32_t hr = 0u;    // rows with a watched update in pass 1 or 2
            uint64_t wm = 0ull;  // (AW) watched lanes updated in pass 1 or 2
            // pass 1: source x-1
            uint32_t left = far_inv<F>();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                left = from_left(left, inv_of(K1[r]));
                const uint32_t own0 = inv_of(K0[r]);
                const bool t1 = relax_key<F>(K1[r], own0, DL_P1(r));
                const bool t0 = relax_key<F>(K0[r], left, DL_P0(r));
                ch |= t0 | t1;
                if constexpr (AW)
                    wm |= (__ballot(t0) & wt.l0) | (__ballot(t1) & wt.l1);
                else
                    hr |= ((__ballot(t0) & wt.l0) | (__ballot(t1) & wt.l1)) ? (2u << r) : 0u;
            }
            // pass 2: source x+1
            uint32_t right = far_inv<F>();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                right = from_right(right, inv_of(K0[r]));
                const uint32_t own1 = inv_of(K1[r]);
                const bool t0 = relax_key<F>(K0[r], own1, DL_P0(r));
                const bool t1 = relax_key<F>(K1[r], right, DL_P1(r));
                ch |= t0 | t1;
                if constexpr (AW)
                    wm |= (__ballot(t0) & wt.r0) | (__ballot(t1) & wt.r1);
                else
                    hr |= ((__ballot(t0) & wt.r0) | (__ballot(t1) & wt.r1)) ? (2u << r) : 0u;
            }
            if constexpr (AW)
                esc |= wm != 0ull;
            else
                esc |= (hr & (wt.rows_in << 1)) != 0u;
            cx = __ballot(ch) != 0ull;
        }
        // pass 3: source y-1.  Publish this wave's last row and its q flag.
        ldsA[wave][lane][0] = inv_of(K0[R - 1]);
        ldsA[wave][lane][1] = inv_of(K1[R - 1]);
        if (lane == 0) lds_q[k & 1][wave] = (chy || cx) ? 1 : 0;
        __syncthreads();
        int any = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) any |= lds_q[k & 1][w];
        if (!any) break;   // uniform across the block
        const bool run_y = lds_q[k & 1][wave] || (wave > 0 && lds_q[k & 1][wave - 1]) ||
                           (wave < NW - 1 && lds_q[k & 1][wave + 1]);
        bool ch = false;
        uint32_t h3 = 0u, h4 = 0u;   // rows with an update of a watched column, pass 3 / 4
        uint64_t w3 = 0ull, w4 = 0ull;   // (AW) watched lanes of row 0 (pass 3) / R - 1 (pass 4)
        if (run_y) {
            const uint32_t top0 = wave > 0 ? ldsA[wave - 1][lane][0] : far_inv<F>();
            const uint32_t top1 = wave > 0 ? ldsA[wave - 1][lane][1] : far_inv<F>();
#pragma unroll
            for (int r = R - 1; r >= 0; --r) {
                const uint32_t a0 = r > 0 ? inv_of(K0[r > 0 ? r - 1 : 0]) : top0;
                const uint32_t a1 = r > 0 ? inv_of(K1[r > 0 ? r - 1 : 0]) : top1;
                const bool t0 = relax_key<F>(K0[r], a0, DL_P0(r));
                const bool t1 = relax_key<F>(K1[r], a1, DL_P1(r));
                ch |= t0 | t1;
                if constexpr (AW) {
                    if (r == 0) w3 = (__ballot(t0) & wt.i0) | (__ballot(t1) & wt.i1);
                } else {
                    h3 |= ((__ballot(t0) & wt.i0) | (__ballot(t1) & wt.i1)) ? (2u << r) : 0u;
                }
            }
        }
        // pass 4: source y+1
        ldsB[wave][lane][0] = inv_of(K0[0]);
        ldsB[wave][lane][1] = inv_of(K1[0]);
        __syncthreads();
        if (run_y) {
            const uint32_t bot0 = wave < NW - 1 ? ldsB[wave + 1][lane][0] : far_inv<F>();
            const uint32_t bot1 = wave < NW - 1 ? ldsB[wave + 1][lane][1] : far_inv<F>();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t b0 = r < R - 1 ? inv_of(K0[r < R - 1 ? r + 1 : r]) : bot0;
                const uint32_t b1 = r < R - 1 ? inv_of(K1[r < R - 1 ? r + 1 : r]) : bot1;
                const bool t0 = relax_key<F>(K0[r], b0, DL_P0(r));
                const bool t1 = relax_key<F>(K1[r], b1, DL_P1(r));
                ch |= t0 | t1;
                if constexpr (AW) {
                    if (r == R - 1) w4 = (__ballot(t0) & wt.i0) | (__ballot(t1) & wt.i1);
                } else {
                    h4 |= ((__ballot(t0) & wt.i0) | (__ballot(t1) & wt.i1)) ? (2u << r) : 0u;
                }
            }
        }
        // only the interior's first row (pass 3) and last row (pass 4) are watched
        if constexpr (AW)
            esc |= (wt.r_top == 0 && w3 != 0ull) || (wt.r_bot == R - 1 && w4 != 0ull);
        else
            esc |= (wt.r_top >= 0 && wt.r_top < R && ((h3 >> (wt.r_top + 1)) & 1u)) ||
                   (wt.r_bot >= 0 && wt.r_bot < R && ((h4 >> (wt.r_bot + 1)) & 1u));
        chx = cx;
        chy = __ballot(ch) != 0ull;
    }
#undef DL_P0
#undef DL_P1
}

// Does tile (tx, ty) of a watched canvas touch the image interior's boundary
// (within its halo)?  Such tiles run the variant that records updates reaching
// the interior from the padding.
__host__ __device__ inline bool tile_watched(int P, int W, int H, int tx, int ty, int IW, int IH,
                                             int K)
{
    if (P <= 0) return false;
    const int ox = tx * IW, oy = ty * IH;
    return (ox - K <= P && P < ox + IW + K) || (ox - K <= W - P - 1 && W - P - 1 < ox + IW + K) ||
           (oy - K <= P && P < oy + IH + K) || (oy - K <= H - P - 1 && H - P - 1 < oy + IH + K);
}

// What a recorded (MODE 1) or replayed (MODE 2) launch needs beyond the plain
// one (see FillRecord).  launch: the launch's index, i.e. the state it reads.
struct TileExtra {
    RecArgs rec;
    int launch;
    // MODE 2: the per-tile replay state (see ReplayTs)
    const int* slot_cur;          // per tile: the record's pool slot at state launch
    const int* slot_next;         // ... at state launch + 1 (its latest entry of a
                                  // state <= launch + 1; every tile has one of state 1)
    const int* dirty_cur;         // per tile: differs from the record at state launch
    int* dirty_next;              // ... at state launch + 1 (written for this tile)
    int* mask_next;               // this tile's change mask of this launch
};

// One launch of one tile.  MODE 0: the plain fill.  MODE 1: also records the
// tile's interior after the launch (when it changed, and always in launch 0).
// MODE 2: replays from a record: the 3 x 3 tiles' states come from the label
// plane where they differ from the record and from the record's pool
// elsewhere, the result is compared with the record's next state, and the
// tile is recorded as in MODE 1.
template <int NW, int R, int K, bool F, int MODE = 0, bool AW = false>
__device__ __forceinline__ void tile_body(
    const BfsCanvas* __restrict__ cvs, int n_canvas, int tile, int src_plane,
    const int* __restrict__ flags_prev, int* __restrict__ flags_cur, int first_launch,
    int* __restrict__ changed_any, int* __restrict__ visits, const TileExtra& ex)
{
    constexpr int IW = 128 - 2 * K;
    constexpr int IH = NW * R - 2 * K;
    static_assert(MODE != 2 || NW >= 4, "the replay's set-up uses one lane of waves 0-3");
    __shared__ uint32_t ldsA[NW][64][2];
    __shared__ uint32_t ldsB[NW][64][2];
    __shared__ int lds_any[2][NW];   // per-wave flags (see iterate)
    // MODE 2: source of each of the 3 x 3 tiles (dy * 3 + dx): base address,
    // row pitch and the canvas origin of that base; MODE 1, 2: record slot
    __shared__ const uint32_t* lds_src[9];
    __shared__ int lds_pitch[9], lds_col[9], lds_row[9];
    __shared__ int lds_slot;
    // MODE 2: the record's slot at state launch + 1, this tile's entry count,
    // and the slot counter value taken for this tile's result
    __shared__ int lds_cmp, lds_nent;
    __shared__ unsigned long long lds_take;

    int ci = 0;
    if (n_canvas > 1) {
        int lo = 0, hi = n_canvas - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cvs[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
        }
        ci = lo;
    }
    const BfsCanvas& cv = cvs[ci];
    const int W = cv.W, H = cv.H;
    const int tiles_x = cv.tiles_x, tiles_y = cv.tiles_y, tile_base = cv.tile_base;
    const int local = tile - tile_base;
    const int tx = local % tiles_x, ty = local / tiles_x;
    const int tid = threadIdx.x;
    // The wave index is uniform; readfirstlane lets the compiler keep every
    // row-dependent quantity in scalar registers.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;

    // Activity: a launch reads the tile's interior and a K-cell halo, so it
    // can change something only if last launch changed a cell of the interior
    // or of that halo -- the band (or corner) of a neighbour's interior that
    // faces this tile.  Tile flags are change masks (see the write-back).
    // Every wave reads the nine flags itself, so no barrier is needed.
    if (MODE != 2 && !first_launch) {
        int f = 0;
        if (lane < 9) {
            const int nx = tx + lane % 3 - 1, ny = ty + lane / 3 - 1;
            // bit of neighbour (dx, dy) = (lane % 3 - 1, lane / 3 - 1) that faces us
            constexpr int kFacing[9] = {kChBR, kChBottom, kChBL, kChRight, kChAny,
                                        kChLeft, kChTR, kChTop, kChTL};
            if (nx >= 0 && ny >= 0 && nx < tiles_x && ny < tiles_y)
                f = flags_prev[tile_base + ny * tiles_x + nx] & kFacing[lane];
        }
        if (__ballot(f != 0) == 0ull) {
            if (tid == 0) flags_cur[tile] = 0;
            return;
        }
    }

    const int ox = tx * IW, oy = ty * IH;   // interior origin
    const int c0 = ox - K + 2 * lane;       // canvas column of this thread's first column
    const int y0 = oy - K + wave * R;       // canvas row of this thread's first row
    const GU32* src = (const GU32*)(src_plane ? cv.lab1 : cv.lab0);
    GU32* dst = (GU32*)(src_plane ? cv.lab0 : cv.lab1);
    if constexpr (MODE == 2) {
        // Sources: the plane for tiles that differ from the record (and for
        // every tile in launch 0, whose input is the seeds), the record's
        // pool slot of state `launch` otherwise.  The same round trip fetches
        // what the end of the launch needs: the record's slot to compare
        // with, the tile's entry count, and a pool slot for the result (taken
        // whether or not it will be used: a slot that no entry references is
        // free again when the record is finished).
        if (tid < 9) {
            const int nx = tx + tid % 3 - 1, ny = ty + tid / 3 - 1;
            const uint32_t* base = src_plane ? cv.lab1 : cv.lab0;
            int pitch = W, col = 0, row = 0;
            if (nx >= 0 && ny >= 0 && nx < tiles_x && ny < tiles_y) {
                const int u = tile_base + ny * tiles_x + nx;
                const int dirty = ex.dirty_cur[u];
                const int slot = ex.slot_cur[u];
                if (ex.launch > 0 && !dirty) {
                    base = ex.rec.pool + static_cast<size_t>(slot) * (IW * IH);
                    pitch = IW;
                    col = nx * IW;
                    row = ny * IH;
                }
            }
            lds_src[tid] = base;
            lds_pitch[tid] = pitch;
            lds_col[tid] = col;
            lds_row[tid] = row;
        } else if (tid == 64) {   // (one lane of three other waves: the loads overlap)
            lds_cmp = ex.slot_next[tile];
        } else if (tid == 128) {
            lds_nent = ex.rec.nent[tile];
        } else if (tid == 192) {
            lds_take = atomicAdd(ex.rec.used, 1ull);
        }
        __syncthreads();
    }

    // Per-cell state: ~label and squared distance (see relax_key).  Loads use
    // clamped addresses and selects, so no per-row lane masks stay live.
    uint64_t K0[R], K1[R];
    const bool in0 = c0 >= 0 && c0 < W;
    const bool in1 = c0 + 1 >= 0 && c0 + 1 < W;
    const int cc0 = min(max(c0, 0), W - 1);
    const int cc1 = min(max(c0 + 1, 0), W - 1);
    // All loads first, then the keys.
    uint32_t la[R], lc[R];
    // (MODE 2: the source tile of each loaded column: 0, 1, 2 = left, own, right)
    const int sx0 = cc0 < ox ? 0 : cc0 >= ox + IW ? 2 : 1;
    const int sx1 = cc1 < ox ? 0 : cc1 >= ox + IW ? 2 : 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        la[r] = lc[r] = kInvalid;
        if (y >= 0 && y < H) {   // wave-uniform
            if constexpr (MODE == 2) {
                const int sy = y < oy ? 0 : y >= oy + IH ? 6 : 3;   // row of sources, wave-uniform
                const int k0 = sy + sx0, k1 = sy + sx1;
                // (plane or pool: global memory, named so to avoid FLAT loads)
                la[r] = ((const GU32*)lds_src[k0])[static_cast<size_t>(y - lds_row[k0]) *
                                                       lds_pitch[k0] + (cc0 - lds_col[k0])];
                lc[r] = ((const GU32*)lds_src[k1])[static_cast<size_t>(y - lds_row[k1]) *
                                                       lds_pitch[k1] + (cc1 - lds_col[k1])];
            } else {
                const size_t rowoff = static_cast<size_t>(y) * static_cast<size_t>(W);
                la[r] = src[rowoff + cc0];
                lc[r] = src[rowoff + cc1];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        if (y >= 0 && y < H) {   // wave-uniform
            const uint32_t a = la[r];
            const uint32_t c = lc[r];
            const uint32_t ia = (in0 && a != kInvalid) ? ~a : far_inv<F>();
            const uint32_t ic = (in1 && c != kInvalid) ? ~c : far_inv<F>();
            // A defined cell's distance, or at an undefined one the far
            // label's (the same packed arithmetic as the update; on a canvas
            // cell the wide far label saturates to kFarD = 2 * 32767^2).
            const uint32_t yy = static_cast<uint32_t>(y + 1) & 0xFFFFu;
            const uint32_t da = key_dist(ia, (static_cast<uint32_t>(c0 + 1) << 16) | yy);
            const uint32_t dc = key_dist(ic, (static_cast<uint32_t>(c0 + 2) << 16) | yy);
            K0[r] = (static_cast<uint64_t>(in0 ? da : 0u) << 32) | ia;
            K1[r] = (static_cast<uint64_t>(in1 ? dc : 0u) << 32) | ic;
        } else {
            K0[r] = far_inv<F>();   // outside: distance 0
            K1[r] = far_inv<F>();
        }
    }

    int esc = 0;
    const int P = cv.P;
    Watch wt;
    const bool on = tile_watched(P, W, H, tx, ty, IW, IH, K);   // block-uniform
    const uint64_t onm = on ? ~0ull : 0ull;
    // (AW: the column masks of passes 1-2 only in waves whose rows are inside)
    const uint64_t xm = (!AW || (y0 >= P && y0 < H - P)) ? onm : 0ull;
    wt.l0 = __ballot(c0 == P) & xm;
    wt.l1 = __ballot(c0 + 1 == P) & xm;
    wt.r0 = __ballot(c0 == W - P - 1) & xm;
    wt.r1 = __ballot(c0 + 1 == W - P - 1) & xm;
    wt.i0 = __ballot(c0 >= P && c0 < W - P) & onm;
    wt.i1 = __ballot(c0 + 1 >= P && c0 + 1 < W - P) & onm;
    wt.r_top = on ? P - y0 : -1;
    wt.r_bot = on ? H - P - 1 - y0 : -1;
    wt.rows_in = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (on && y0 + r >= P && y0 + r < H - P) wt.rows_in |= 1u << r;
    iterate<NW, R, K, F, AW>(K0, K1, c0, y0, wt, ldsA, ldsB, lds_any, wave, lane, esc);

    // Write the interior back and note whether it changed.  The row origin is
    // passed through an empty asm statement so that the compiler recomputes
    // the row addresses here instead of keeping the load-time ones live in
    // scalar registers across the whole fill loop.
    int yb = y0;
    int Wb = W;
    asm volatile("" : "+s"(yb), "+s"(Wb));
    // Change mask: which of the interior, its four K-wide border bands and
    // its four K x K corners changed (the parts that neighbours' halos read).
    const int lx0 = 2 * lane - K;   // interior-relative column of column 0
    const bool w0 = lx0 >= 0 && lx0 < IW && c0 < Wb;
    const bool w1 = lx0 + 1 >= 0 && lx0 + 1 < IW && c0 + 1 < Wb;
    bool ch0 = false, ch1 = false;      // column 0 / 1 changed in some row
    bool top0 = false, top1 = false;    // ... in a row of the top band
    bool bot0 = false, bot1 = false;    // ... in a row of the bottom band
    // The rows' launch-start labels, all loaded before the first store.
    // MODE 2: with them, the record's state launch + 1 to compare with (its
    // entry of that state, or the state before when the recorded launch
    // changed nothing) and the pool slot this tile's result would take.
    uint32_t old0[R], old1[R];
    uint32_t cmp0[MODE == 2 ? R : 1], cmp1[MODE == 2 ? R : 1];
    if constexpr (MODE == 2) {
        // (kept in LDS, not in a register that could be spilled and then
        // reloaded after the stores)
        if (tid == 0) {
            const unsigned long long j = lds_take;
            const unsigned long long fn = static_cast<unsigned long long>(ex.rec.free_n);
            const unsigned long long sl =
                j < fn ? static_cast<unsigned long long>(ex.rec.freelist[j])
                       : static_cast<unsigned long long>(ex.rec.fresh) + (j - fn);
            lds_take = sl < static_cast<unsigned long long>(ex.rec.pool_cap) ? sl : ~0ull;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ly = wave * R + r - K;
        const int y = yb + r;
        old0[r] = old1[r] = 0u;
        if constexpr (MODE == 2) cmp0[r] = cmp1[r] = 0u;
        if (ly < 0 || ly >= IH || y >= H) continue;   // wave-uniform
        const size_t rowoff = static_cast<size_t>(y) * static_cast<size_t>(Wb);
        if constexpr (MODE == 2) {
            // the tile's own source (the plane or its record slot)
            const GU32* own = ((const GU32*)lds_src[4]) +
                              static_cast<size_t>(y - lds_row[4]) * lds_pitch[4] - lds_col[4];
            const GU32* cmp =
                ((const GU32*)ex.rec.pool) + static_cast<size_t>(lds_cmp) * (IW * IH);
            if (w0) old0[r] = own[c0];
            if (w1) old1[r] = own[c0 + 1];
            if (w0) cmp0[r] = cmp[ly * IW + lx0];
            if (w1) cmp1[r] = cmp[ly * IW + lx0 + 1];
        } else {
            if (w0) old0[r] = src[rowoff + c0];
            if (w1) old1[r] = src[rowoff + c0 + 1];
        }
    }
    // Labels and change flags, then the stores: nothing waits for a load
    // after the first store (on this GPU a wait for a load also waits for the
    // stores issued before it).
    uint32_t nv0[R], nv1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        nv0[r] = inv_of(K0[r]) == far_inv<F>() ? kInvalid : ~inv_of(K0[r]);
        nv1[r] = inv_of(K1[r]) == far_inv<F>() ? kInvalid : ~inv_of(K1[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ly = wave * R + r - K;
        const int y = yb + r;
        if (ly < 0 || ly >= IH || y >= H) continue;   // wave-uniform
        const bool c0r = w0 && old0[r] != nv0[r];
        const bool c1r = w1 && old1[r] != nv1[r];
        ch0 |= c0r;
        ch1 |= c1r;
        if (ly < K) {   // wave-uniform
            top0 |= c0r;
            top1 |= c1r;
        }
        if (ly >= IH - K) {
            bot0 |= c0r;
            bot1 |= c1r;
        }
    }
    // (MODE 2 stores at the end of the tile, after its record bookkeeping.)
    auto store_rows = [&]() {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ly = wave * R + r - K;
            const int y = yb + r;
            if (ly < 0 || ly >= IH || y >= H) continue;   // wave-uniform
            const size_t rowoff = static_cast<size_t>(y) * static_cast<size_t>(Wb);
            if (w0) dst[rowoff + c0] = nv0[r];
            if (w1) dst[rowoff + c0 + 1] = nv1[r];
        }
    };
    if constexpr (MODE != 2) store_rows();
    if (esc && lane == 0 && cv.escape) atomicOr(cv.escape, 1);   // esc is wave-uniform
    const bool l0 = lx0 < K, l1 = lx0 + 1 < K;                  // left band columns
    const bool r0 = lx0 >= IW - K, r1 = lx0 + 1 >= IW - K;      // right band columns
    int m = 0;
    if (__ballot(ch0 || ch1)) m |= kChAny;
    if (__ballot((ch0 && l0) || (ch1 && l1))) m |= kChLeft;
    if (__ballot((ch0 && r0) || (ch1 && r1))) m |= kChRight;
    if (__ballot(top0 || top1)) m |= kChTop;
    if (__ballot(bot0 || bot1)) m |= kChBottom;
    if (__ballot((top0 && l0) || (top1 && l1))) m |= kChTL;
    if (__ballot((top0 && r0) || (top1 && r1))) m |= kChTR;
    if (__ballot((bot0 && l0) || (bot1 && l1))) m |= kChBL;
    if (__ballot((bot0 && r0) || (bot1 && r1))) m |= kChBR;
    if constexpr (MODE == 2) {
        // Compare with the record's state launch + 1.
        bool diff = false;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ly = wave * R + r - K;
            if (ly < 0 || ly >= IH || yb + r >= H) continue;   // wave-uniform
            if (w0) diff |= cmp0[r] != nv0[r];
            if (w1) diff |= cmp1[r] != nv1[r];
        }
        if (__ballot(diff)) m |= kDiffers;
    }
    __syncthreads();   // every wave is past its last read of lds_any
    if (lane == 0) lds_any[0][wave] = m;
    __syncthreads();
    if constexpr (MODE < 2) {
        if (tid == 0) {
            int any = 0;
            for (int w = 0; w < NW; ++w) any |= lds_any[0][w];
            flags_cur[tile] = any;
            // No atomics here: thousands of blocks would contend for one
            // address.  Only this block owns the tile's visit count in this
            // launch, and the change flag is an idempotent store.
            visits[tile] += 1;
            if (any & kChAny) *changed_any = 1;
        }
    }
    if constexpr (MODE >= 1) {
        // Record the interior after this launch when it changed (and every
        // tile in launch 0, so that a record has every tile's state 1).
        int any = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) any |= lds_any[0][w];
        const int mask = any & kChMask;
        if (tid == 0) {
            int slot = -1;
            if constexpr (MODE == 2) {
                // The entry's atomic is waited for before this lane's first
                // store: on this GPU a wait for a returning memory operation
                // also waits for every store issued before it.
                const bool keep = (mask & kChAny) || ex.launch == 0;
                const unsigned long long id = keep ? atomicAdd(ex.rec.used + 1, 1ull) : 0ull;
                asm volatile("" ::"v"(id));
                const unsigned long long take = lds_take;   // ~0: beyond the pool
                ex.dirty_next[tile] = (any & kDiffers) ? 1 : 0;
                ex.mask_next[tile] = mask;
                atomicAdd(visits + tile, 1);   // (statistics; no value returned)
                if (mask & kChAny) *changed_any = 1;
                if (keep) {
                    if (take != ~0ull && id < static_cast<unsigned long long>(ex.rec.ent_cap)) {
                        slot = static_cast<int>(take);
                        ex.rec.app[id] = make_int4(tile, ex.launch + 1, mask, slot);
                        ex.rec.rank[id] = lds_nent;
                        ex.rec.nent[tile] = lds_nent + 1;
                    } else {
                        *ex.rec.overflow = 1;
                    }
                }
            } else if ((mask & kChAny) || ex.launch == 0) {
                const unsigned long long j = atomicAdd(ex.rec.used, 1ull);
                const unsigned long long id = atomicAdd(ex.rec.used + 1, 1ull);
                const unsigned long long fn = static_cast<unsigned long long>(ex.rec.free_n);
                const unsigned long long sl =
                    j < fn ? static_cast<unsigned long long>(ex.rec.freelist[j])
                           : static_cast<unsigned long long>(ex.rec.fresh) + (j - fn);
                if (sl < static_cast<unsigned long long>(ex.rec.pool_cap) &&
                    id < static_cast<unsigned long long>(ex.rec.ent_cap)) {
                    slot = static_cast<int>(sl);
                    ex.rec.app[id] = make_int4(tile, ex.launch + 1, mask, slot);
                    ex.rec.rank[id] = ex.rec.nent[tile]++;
                } else {
                    *ex.rec.overflow = 1;
                }
            }
            lds_slot = slot;
        }
        __syncthreads();
        const int slot = lds_slot;
        if constexpr (MODE == 2) store_rows();
        if (slot >= 0) {
            uint32_t* out = ex.rec.pool + static_cast<size_t>(slot) * (IW * IH);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ly = wave * R + r - K;
                if (ly < 0 || ly >= IH || yb + r >= H) continue;   // wave-uniform
                if (w0) out[ly * IW + lx0] = nv0[r];
                if (w1) out[ly * IW + lx0 + 1] = nv1[r];
            }
        }
    }
}

// One block per tile.  Contiguous tile ranges per XCD: consecutive dispatches
// go round-robin over the 8 XCDs, so give each XCD a contiguous slice of tiles
// (neighbouring tiles share halo lines in the same L2).  The grid is launched
// as a multiple of 8 blocks, which makes this a bijection on [0, gridDim.x).
// Tile-kernel shapes of at most 16 waves x 8 rows are compiled for 64 VGPRs
// (8 waves per SIMD): four resident 8-wave blocks per CU instead of three, or
// two 16-wave blocks instead of one.  The list kernel takes the same bound:
// its spills (a few words kept across its loop over tiles) sit outside the
// iteration loop, and the extra resident block is worth more (16384^2
// clustered round 48.9 -> 47.6 ms).
template <int NW, int R>
constexpr int kFillWavesPerEu = (NW <= 16 && R <= 8) ? 8 : 1;

template <int NW, int R, int K, bool F, int MODE = 0, bool AW = false>
//
// prev_changed (nullable): the previous launch's change counter.  When it is
// 0 the fill has converged (no cell changed anywhere, so both planes hold the
// final labels) and every block returns at once, writing nothing: the next
// launch sees this launch's counter still 0 and returns too.
// ex: MODE 1 (recorded fill) only.
__global__ __launch_bounds__(NW * 64, (kFillWavesPerEu<NW, R>)) void bfs_tile_kernel(
    const BfsCanvas* __restrict__ cvs, int n_canvas, int total_tiles, int src_plane,
    const int* __restrict__ flags_prev, int* __restrict__ flags_cur, int first_launch,
    int* __restrict__ changed_any, int* __restrict__ visits, const int* __restrict__ prev_changed,
    TileExtra ex)
{
    if (prev_changed && *prev_changed == 0) return;   // uniform across the grid
    const int per_xcd = gridDim.x / 8;
    const int b = blockIdx.x;
    const int tile = (b % 8) * per_xcd + b / 8;
    if (tile >= total_tiles) return;
    tile_body<NW, R, K, F, MODE, AW>(cvs, n_canvas, tile, src_plane, flags_prev, flags_cur,
                                 first_launch, changed_any, visits, ex);
}

// The same launch over a list of the active tiles (bfs_active_list_kernel):
// a fixed grid strides over the list, each XCD over a contiguous eighth of it.
// On large canvases late launches have few active tiles, and a block per tile
// would spend most of the launch starting and retiring idle blocks.
// MODE 2 (replay): the list is the plan kernel's and ex describes the record.
template <int NW, int R, int K, bool F, int MODE = 0, bool AW = false>
__global__ __launch_bounds__(NW * 64, (kFillWavesPerEu<NW, R>)) void bfs_list_kernel(
    const BfsCanvas* __restrict__ cvs, int n_canvas, const int* __restrict__ list,
    const int* __restrict__ count, int src_plane, const int* __restrict__ flags_prev,
    int* __restrict__ flags_cur, int* __restrict__ changed_any, int* __restrict__ visits,
    TileExtra ex)
{
    const int n = *count;
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, rank = blockIdx.x >> 3;
    const int lo = static_cast<int>(static_cast<int64_t>(n) * xcd / 8);
    const int hi = static_cast<int>(static_cast<int64_t>(n) * (xcd + 1) / 8);
    for (int i = lo + rank; i < hi; i += per) {
        tile_body<NW, R, K, F, MODE, AW>(cvs, n_canvas, list[i], src_plane, flags_prev, flags_cur,
                                         0, changed_any, visits, ex);
        __syncthreads();   // the next tile reuses the LDS exchange buffers
    }
}

// Tiles active in the coming launch (some neighbour's last change faces them,
// as in tile_body's own test), appended to `list` (one atomic per block; the
// order is irrelevant: tiles of a launch are independent).  Also clears the
// flags the coming launch writes, since it visits only listed tiles.
__global__ __launch_bounds__(256) void bfs_active_list_kernel(
    const BfsCanvas* __restrict__ cvs, int n_canvas, int total_tiles,
    const int* __restrict__ flags_prev, int* __restrict__ flags_next, int* __restrict__ list,
    int* __restrict__ count, const int* __restrict__ prev_changed)
{
    __shared__ int wave_n[4];
    __shared__ int base;
    const int tile = blockIdx.x * 256 + threadIdx.x;
    // After a launch that changed nothing the fill has converged: list no
    // tile (the flags are still cleared, for the launch after this one).
    const bool live = !prev_changed || *prev_changed != 0;
    bool act = false;
    if (tile < total_tiles) {
        flags_next[tile] = 0;
    }
    if (tile < total_tiles && live) {
        int ci = 0;
        if (n_canvas > 1) {
            int lo = 0, hi = n_canvas - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (cvs[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
            }
            ci = lo;
        }
        const BfsCanvas& cv = cvs[ci];
        const int local = tile - cv.tile_base;
        const int tx = local % cv.tiles_x, ty = local / cv.tiles_x;
        constexpr int kFacing[9] = {kChBR, kChBottom, kChBL, kChRight, kChAny,
                                    kChLeft, kChTR, kChTop, kChTL};
        int f = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int nx = tx + k % 3 - 1, ny = ty + k / 3 - 1;
            if (nx >= 0 && ny >= 0 && nx < cv.tiles_x && ny < cv.tiles_y)
                f |= flags_prev[cv.tile_base + ny * cv.tiles_x + nx] & kFacing[k];
        }
        act = f != 0;
    }
    const uint64_t m = __ballot(act);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wave_n[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
        base = t ? atomicAdd(count, t) : 0;
    }
    __syncthreads();
    if (act) {
        int off = base;
        for (int k = 0; k < w; ++k) off += wave_n[k];
        off += __popcll(m & ((1ull << lane) - 1ull));
        list[off] = tile;
    }
}

// ---------------------------------------------------------------------------
// Simple version: one thread per cell, one pass per launch, ping-pong planes.
// ---------------------------------------------------------------------------
__global__ void bfs_pass_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                int W, int H, int dir, int* __restrict__ changed)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const size_t c = static_cast<size_t>(y) * W + x;
    uint32_t lab = src[c];
    int d = lab == kInvalid ? INT_MAX : dist2_lab(x, y, lab);
    int sx = x, sy = y;
    if (dir == 0) sx = x - 1;
    else if (dir == 1) sx = x + 1;
    else if (dir == 2) sy = y - 1;
    else sy = y + 1;
    if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
        const uint32_t nl = src[static_cast<size_t>(sy) * W + sx];
        if (relax(lab, d, nl, x, y)) *changed = 1;
    }
    dst[c] = lab;
}

__global__ void bfs_seed_kernel(uint32_t* __restrict__ a, uint32_t* __restrict__ b, int W,
                                const uint32_t* __restrict__ labels, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = labels[i];
    const size_t c = static_cast<size_t>(l & 0xFFFFu) * W + (l >> 16);
    a[c] = l;
    b[c] = l;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n)
{
    size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 4;
    for (; i < n; i += stride) {
        if (i + 4 <= n && (reinterpret_cast<uintptr_t>(p + i) & 15) == 0) {
            *reinterpret_cast<uint4*>(p + i) = make_uint4(v, v, v, v);
        } else {
            for (size_t j = i; j < n && j < i + 4; ++j) p[j] = v;
        }
    }
}

// out[0] = sum of p[0..n) (one block).
// Up to four fills in one launch: block b works on the segment whose block
// range holds b (ranges in `first`, one per segment).
__global__ void fill_multi_kernel(FillSegs f, FillBlocks first)
{
    int k = 0;
    while (k + 1 < f.k && blockIdx.x >= first.b[k + 1]) ++k;
    uint32_t* p = f.seg[k].p;
    const uint32_t v = f.seg[k].v;
    const size_t n = f.seg[k].n;
    const size_t stride = static_cast<size_t>(first.b[k + 1] - first.b[k]) * blockDim.x * 4;
    size_t i = (static_cast<size_t>(blockIdx.x - first.b[k]) * blockDim.x + threadIdx.x) * 4;
    for (; i < n; i += stride) {
        if (i + 4 <= n && (reinterpret_cast<uintptr_t>(p + i) & 15) == 0) {
            *reinterpret_cast<uint4*>(p + i) = make_uint4(v, v, v, v);
        } else {
            for (size_t j = i; j < n && j < i + 4; ++j) p[j] = v;
        }
    }
}

__global__ __launch_bounds__(1024) void sum_ints_kernel(const int* __restrict__ p, int n,
                                                        int* __restrict__ out)
{
    __shared__ int red[16];
    int v = 0;
    for (int i = threadIdx.x; i < n; i += 1024) v += p[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; ++w) t += red[w];
        out[0] = t;
    }
}

__global__ void zero_ints_kernel(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

using TileKernel = void (*)(const BfsCanvas*, int, int, int, const int*, int*, int, int*, int*,
                           const int*, TileExtra);
using ListKernel = void (*)(const BfsCanvas*, int, const int*, const int*, int, const int*, int*,
                            int*, int*, TileExtra);

// fn / list_fn[e]: e = 1 for the min encoding (canvases within
// bfs_min_encoding_ok), 0 for the compare-and-select one.  fn_aw / list_fn_aw:
// the min encoding with the aligned watch (the default shape only; null
// elsewhere).
struct Variant {
    int waves, rows, iters;
    TileKernel fn[2];
    ListKernel list_fn[2];
    TileKernel fn_aw;
    ListKernel list_fn_aw;
};

#define DL_VARIANT(W_, R_, K_)                                                       \
    {W_, R_, K_, {bfs_tile_kernel<W_, R_, K_, false>, bfs_tile_kernel<W_, R_, K_, true>}, \
     {bfs_list_kernel<W_, R_, K_, false>, bfs_list_kernel<W_, R_, K_, true>}, nullptr, nullptr}
#define DL_VARIANT_AW(W_, R_, K_)                                                    \
    {W_, R_, K_, {bfs_tile_kernel<W_, R_, K_, false>, bfs_tile_kernel<W_, R_, K_, true>}, \
     {bfs_list_kernel<W_, R_, K_, false>, bfs_list_kernel<W_, R_, K_, true>},              \
     bfs_tile_kernel<W_, R_, K_, true, 0, true>, bfs_list_kernel<W_, R_, K_, true, 0, true>}

// Instantiated configurations.  The tile is 128 wide; height waves*rows.
const Variant kVariants[] = {
    DL_VARIANT(16, 8, 8), DL_VARIANT(16, 8, 7),  DL_VARIANT(16, 8, 4),  DL_VARIANT(16, 8, 12),
    DL_VARIANT(16, 8, 16),
    DL_VARIANT(8, 8, 4),  DL_VARIANT(8, 8, 5),   DL_VARIANT(8, 8, 6),   DL_VARIANT(8, 8, 7),
    DL_VARIANT_AW(8, 8, 8), DL_VARIANT(8, 16, 8), DL_VARIANT(16, 4, 4), DL_VARIANT(16, 4, 8),
};

const Variant* find_variant(const BfsConfig& c)
{
    for (const auto& v : kVariants)
        if (v.waves == c.waves && v.rows == c.rows && v.iters == c.iters) return &v;
    return nullptr;
}

// Launches after the first go through the active-tile list on fills of at
// least BfsWork::list_min_tiles tiles; the list launch uses a fixed grid.
// (Smaller fills launch a block per tile throughout: switching them to the
// list for their sparse late launches measured slower at 4096^2, since each
// list launch needs the listing kernel first.)
constexpr int kListGrid = 2048;

// One tile launch of the fill: launch 0 (and every launch of a small fill)
// visits every tile; later launches of a large fill list the active tiles
// first.  changed / listcnt: this launch's zeroed counters.
// prev_changed: see bfs_tile_kernel (nullptr: always run).
// The recording (MODE 1) kernels: built for the default shape and the min
// encoding only (the shapes and canvases a session records; see
// bfs_record_supported).
struct RecKernels {
    TileKernel fn;
    ListKernel list_fn;
};
const RecKernels kRecKernels[2] = {
    {bfs_tile_kernel<8, 8, 8, true, 1>, bfs_list_kernel<8, 8, 8, true, 1>},
    {bfs_tile_kernel<8, 8, 8, true, 1, true>, bfs_list_kernel<8, 8, 8, true, 1, true>}};

// rec (nullable): record this launch into the record ex describes.
void launch_tiles(const Variant* v, const BfsConfig& cfg, BfsWork& work, int n_canvas, int total,
                  int launch, int* changed, int* listcnt, hipStream_t s,
                  const int* prev_changed = nullptr, const TileExtra* rec = nullptr)
{
    TileExtra ex{};
    if (rec) {
        ex = *rec;
        ex.launch = launch;
    }
    int* flags = work.flags.as<int>();
    const int src_plane = launch & 1;
    int* fprev = flags + ((launch + 1) & 1) * total;
    int* fcur = flags + (launch & 1) * total;
    const BfsCanvas* cv = reinterpret_cast<const BfsCanvas*>(work.canvases.ptr);
    int* visits = work.visits.as<int>();
    ProfScope ps(kKFill, s);
    if (launch > 0 && total >= work.list_min_tiles) {
        int* list = work.list.as<int>();
        bfs_active_list_kernel<<<(total + 255) / 256, 256, 0, s>>>(cv, n_canvas, total, fprev, fcur,
                                                                   list, listcnt, prev_changed);
        DL_LAUNCH_CHECK();
        const ListKernel lk = rec ? kRecKernels[work.aligned ? 1 : 0].list_fn
                                  : (work.aligned && v->list_fn_aw ? v->list_fn_aw
                                                                   : v->list_fn[work.enc]);
        lk<<<kListGrid, v->waves * 64, 0, s>>>(
            cv, n_canvas, list, listcnt, src_plane, fprev, fcur, changed, visits, ex);
    } else {
        const int grid_blocks = (total + 7) / 8 * 8;
        const TileKernel tk = rec ? kRecKernels[work.aligned ? 1 : 0].fn
                                  : (work.aligned && v->fn_aw ? v->fn_aw : v->fn[work.enc]);
        tk<<<grid_blocks, v->waves * 64, 0, s>>>(
            cv, n_canvas, total, src_plane, fprev, fcur, launch == 0 ? 1 : 0, changed, visits,
            launch > 0 ? prev_changed : nullptr, ex);
    }
    DL_LAUNCH_CHECK();
}

}  // namespace

bool bfs_config_supported(const BfsConfig& c) { return find_variant(c) != nullptr; }

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return;
    const size_t threads = (n + 3) / 4;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((threads + 255) / 256, 65536));
    fill_u32_kernel<<<blocks, 256, 0, s>>>(p, v, n);
    DL_LAUNCH_CHECK();
}

void fill_u32_multi(const FillSegs& segs, hipStream_t s)
{
    FillSegs f{};
    FillBlocks first{};
    unsigned blocks = 0;
    for (int k = 0; k < segs.k; ++k) {
        if (segs.seg[k].n == 0) continue;
        first.b[f.k] = blocks;
        f.seg[f.k++] = segs.seg[k];
        const size_t threads = (segs.seg[k].n + 3) / 4;
        blocks += static_cast<unsigned>(std::min<size_t>((threads + 255) / 256, 16384));
    }
    if (f.k == 0) return;
    first.b[f.k] = blocks;
    fill_multi_kernel<<<blocks, 256, 0, s>>>(f, first);
    DL_LAUNCH_CHECK();
}

void bfs_init_canvas(const BfsCanvas& cv, const uint32_t* d_labels, int64_t n, hipStream_t s)
{
    ProfScope ps(kKInit, s);
    const size_t cells = static_cast<size_t>(cv.W) * static_cast<size_t>(cv.H);
    fill_u32(cv.lab0, kInvalid, cells, s);
    if (n > 0) {
        bfs_seed_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(cv.lab0, cv.lab1,
                                                                              cv.W, d_labels, n);
        DL_LAUNCH_CHECK();
    }
}

// Tiles of every canvas and the device copy of the canvas table; returns the
// tile count and the launch bound.
static int bfs_setup(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                     hipStream_t s, int64_t* max_iters_out)
{
    const int iw = bfs_tile_iw(cfg), ih = bfs_tile_ih(cfg);
    int total = 0;
    int64_t max_iters = 64;
    // The planes hold labels, not keys, so the key encoding may differ from
    // one fill (or launch) to the next.
    bool min_ok = work.allow_min_encoding;
    bool aligned = work.allow_aligned_watch;
    for (auto& c : canvases) {
        min_ok = min_ok && bfs_min_encoding_ok(c.W, c.H);
        aligned = aligned && bfs_watch_aligned(c.escape ? c.P : 0, c.H, cfg);
        c.tiles_x = (c.W + iw - 1) / iw;
        c.tiles_y = (c.H + ih - 1) / ih;
        c.tile_base = total;
        total += c.tiles_x * c.tiles_y;
        max_iters = std::max<int64_t>(max_iters, 8LL * (c.W + c.H) + 64);
    }
    if (total == 0) return 0;
    work.enc = min_ok ? 1 : 0;
    work.aligned = aligned && min_ok;   // (aligned kernels use the min encoding)
    const int n_canvas = static_cast<int>(canvases.size());
    work.canvases.reserve(sizeof(BfsCanvas) * n_canvas);
    work.host_canvases.reserve(sizeof(BfsCanvas) * n_canvas);
    stage_h2d(work.host_canvases, 0, work.canvases.ptr, canvases.data(),
              sizeof(BfsCanvas) * n_canvas, s);
    work.flags.reserve(sizeof(int) * 2 * static_cast<size_t>(total));
    work.list.reserve(sizeof(int) * static_cast<size_t>(total));
    work.visits.reserve(sizeof(int) * static_cast<size_t>(total));
    *max_iters_out = max_iters;
    return total;
}

int bfs_launch_fixed(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                     hipStream_t s, int n_launches, int* host_changed, bool counters_zeroed)
{
    const Variant* v = find_variant(cfg);
    if (!v) throw runtime_error("unsupported flood-fill tile configuration");
    int64_t max_iters = 0;
    const int total = bfs_setup(canvases, cfg, work, s, &max_iters);
    if (total == 0 || n_launches <= 0) return 0;
    const int n_canvas = static_cast<int>(canvases.size());
    // counters: n changed flags, the visit total, then n list counts
    work.counters.reserve(sizeof(int) * bfs_fixed_counter_words(n_launches));
    int* dcnt = work.counters.as<int>();
    int* visits = work.visits.as<int>();
    // Without visit counting the per-tile counts are left as they are (never
    // read) and the total stays 0.
    if (work.count_visits)
        fill_u32(reinterpret_cast<uint32_t*>(visits), 0u, static_cast<size_t>(total), s);
    if (!counters_zeroed)
        fill_u32(reinterpret_cast<uint32_t*>(dcnt), 0u, bfs_fixed_counter_words(n_launches), s);
    for (int launch = 0; launch < n_launches; ++launch)
        // (the counters were zeroed above, so a launch's counter stays 0
        // unless it changed something)
        launch_tiles(v, cfg, work, n_canvas, total, launch, dcnt + launch,
                     dcnt + n_launches + 1 + launch, s, launch > 0 ? dcnt + launch - 1 : nullptr);
    if (work.count_visits) {
        sum_ints_kernel<<<1, 1024, 0, s>>>(visits, total, dcnt + n_launches);
        DL_LAUNCH_CHECK();
    }
    if (host_changed)
        DL_HIP(hipMemcpyAsync(host_changed, dcnt,
                              sizeof(int) * static_cast<size_t>(n_launches + 1),
                              hipMemcpyDeviceToHost, s));
    return n_launches & 1;
}

namespace {
int run_impl(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
             hipStream_t s, BfsStats* stats, int* visits_async, int start_launch,
             const TileExtra* rec);
}  // namespace

int bfs_run(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
            hipStream_t s, BfsStats* stats, int* visits_async, int start_launch)
{
    return run_impl(canvases, cfg, work, s, stats, visits_async, start_launch, nullptr);
}

namespace {
// bfs_run, recording every launch into rec when given (see launch_tiles).
int run_impl(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
             hipStream_t s, BfsStats* stats, int* visits_async, int start_launch,
             const TileExtra* rec)
{
    const Variant* v = find_variant(cfg);
    if (!v) throw runtime_error("unsupported flood-fill tile configuration");
    int64_t max_iters = 0;
    const int total = bfs_setup(canvases, cfg, work, s, &max_iters);
    const int n_canvas = static_cast<int>(canvases.size());
    if (total == 0) return 0;

    // Launches are issued in chunks; the host reads the changed-tile counter of
    // a chunk while the next one runs.  Launches after convergence find no
    // active tile and return at once.
    constexpr int kChunk = 4;
    constexpr int kSlots = 2 * kChunk;  // counters for two chunks in flight
    work.counters.reserve(sizeof(int) * (2 * kSlots + 2));
    work.host_counters.reserve(sizeof(int) * (2 * kSlots + 2));
    work.visits.reserve(sizeof(int) * static_cast<size_t>(total));
    int* dcnt = work.counters.as<int>();
    int* hcnt = work.host_counters.as<int>();
    int* visits = work.visits.as<int>();
    fill_u32(reinterpret_cast<uint32_t*>(visits), 0u, static_cast<size_t>(total), s);
    zero_ints_kernel<<<1, 64, 0, s>>>(dcnt, 2 * kSlots);
    DL_LAUNCH_CHECK();

    // Without events (work.use_events false) a chunk's counters are waited for
    // with a stream synchronisation, which also waits for the chunk behind it.
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (work.use_events) {
        DL_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        DL_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    }

    const int max_launches = static_cast<int>((max_iters + cfg.iters - 1) / cfg.iters) + 2;
    int launch = start_launch;
    int chunk = 0;
    int final_plane = 0;
    bool done = false;
    int pending_chunk = 0;
    BfsStats st;

    auto issue_chunk = [&](int ch) {
        const int slot0 = (ch % 2) * kChunk;
        zero_ints_kernel<<<1, 64, 0, s>>>(dcnt + 2 * slot0, 2 * kChunk);
        for (int j = 0; j < kChunk; ++j) {
            launch_tiles(v, cfg, work, n_canvas, total, launch, dcnt + 2 * (slot0 + j),
                         dcnt + 2 * (slot0 + j) + 1, s, nullptr, rec);
            ++launch;
        }
        DL_HIP(hipMemcpyAsync(hcnt + 2 * slot0, dcnt + 2 * slot0, sizeof(int) * 2 * kChunk,
                              hipMemcpyDeviceToHost, s));
        if (ev[ch % 2]) DL_HIP(hipEventRecord(ev[ch % 2], s));
    };

    issue_chunk(chunk++);
    pending_chunk = 0;
    while (!done) {
        if (launch < max_launches) issue_chunk(chunk++);  // keep one chunk in flight
        if (ev[pending_chunk % 2])
            DL_HIP(hipEventSynchronize(ev[pending_chunk % 2]));
        else
            DL_HIP(hipStreamSynchronize(s));
        const int slot0 = (pending_chunk % 2) * kChunk;
        for (int j = 0; j < kChunk; ++j) {
            const int changed = hcnt[2 * (slot0 + j)];
            const int launch_idx = start_launch + pending_chunk * kChunk + j;
            if (changed == 0) {
                // Launch launch_idx wrote its (unchanged) result into plane
                // (launch_idx + 1) & 1; both planes agree from here on.
                final_plane = (launch_idx + 1) & 1;
                st.converged = launch_idx;
                done = true;
                break;
            }
        }
        if (done) break;
        if (pending_chunk + 1 >= chunk) {
            throw runtime_error("flood fill did not converge within the iteration bound");
        }
        pending_chunk = pending_chunk + 1;
    }
    // Sum the tiles' visit counts.  Launches still queued behind the converged
    // one find no active tile; later work on the stream simply follows them
    // (both label planes hold the converged labels from the converged launch
    // on), so nothing here needs to wait unless the caller wants the total now.
    sum_ints_kernel<<<1, 1024, 0, s>>>(visits, total, dcnt + 2 * kSlots);
    DL_LAUNCH_CHECK();
    if (visits_async) {
        DL_HIP(hipMemcpyAsync(visits_async, dcnt + 2 * kSlots, sizeof(int), hipMemcpyDeviceToHost,
                              s));
    } else {
        DL_HIP(hipMemcpyAsync(hcnt + 2 * kSlots, dcnt + 2 * kSlots, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        st.tile_visits = hcnt[2 * kSlots];
    }
    if (ev[0]) DL_HIP(hipEventDestroy(ev[0]));
    if (ev[1]) DL_HIP(hipEventDestroy(ev[1]));
    st.launches = launch;
    if (stats) *stats = st;
    return final_plane;
}
}  // namespace

int bfs_run_simple(const BfsCanvas& cv, BfsWork& work, hipStream_t s, int64_t* iters_out)
{
    work.counters.reserve(sizeof(int) * 2);
    work.host_counters.reserve(sizeof(int) * 2);
    int* dc = work.counters.as<int>();
    int* hc = work.host_counters.as<int>();
    dim3 grid((cv.W + 255) / 256, cv.H);
    int plane = 0;
    int64_t iters = 0;
    const int64_t bound = 8LL * (cv.W + cv.H) + 64;
    for (;;) {
        ++iters;
        zero_ints_kernel<<<1, 64, 0, s>>>(dc, 1);
        for (int dir = 0; dir < 4; ++dir) {
            const uint32_t* src = plane ? cv.lab1 : cv.lab0;
            uint32_t* dst = plane ? cv.lab0 : cv.lab1;
            bfs_pass_kernel<<<grid, 256, 0, s>>>(src, dst, cv.W, cv.H, dir, dc);
            DL_LAUNCH_CHECK();
            plane ^= 1;
        }
        DL_HIP(hipMemcpyAsync(hc, dc, sizeof(int), hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        if (hc[0] == 0) break;
        if (iters > bound) throw runtime_error("flood fill did not converge");
    }
    if (iters_out) *iters_out = iters;
    return plane;
}

namespace {

// ---------------------------------------------------------------------------
// Fill records (see FillRecord): recording, and replaying only what changed.
//
// Replay of launch L (computing state L + 1), exactly as the launch-by-launch
// fill would on the new seeds:
//  * dirty[t]: the tile's state differs from the record's at state L.  Seeded
//    from the seed-presence bitmaps (tiles holding an inserted or removed
//    seed differ at state 0).
//  * A tile outside the 3 x 3 neighbourhood of every dirty tile reads exactly
//    what the recorded launch read, so its state L + 1 is the record's.
//  * Inside it, a tile is computed when a neighbour's change of the last
//    launch faces it in either fill (otherwise neither fill changes it, and
//    its difference from the record stays what it was), from the plane for
//    dirty tiles and from the record for the others; its result is compared
//    with the record's state L + 1.
//  * A dirty tile holds its state in both label planes whenever it is not
//    computed: it becomes dirty only as the result of a launch that changed
//    it (or whose recorded counterpart did), so it is computed in the next
//    launch too, and a launch that changes nothing writes what it read.
// The fill has converged after the first launch that changes nothing in the
// new fill (computed tiles, and record changes of the tiles that follow it).
// Clean tiles then take their record state, and the new fill is recorded in
// turn: computed tiles store their results, the others reuse record slots.

// Per-tile replay state words (FillRecord::tstate), by parity where noted.
constexpr int kTsNext = 0;    // [2][tiles] first record entry of state >= L (parity L)
constexpr int kTsDirty = 2;   // [2][tiles]
constexpr int kTsMask = 4;    // [2][tiles] the new fill's change mask of the last launch
constexpr int kTsNent = 6;    // [tiles] entries of the fill being recorded
// The record at the tile itself, carried from launch to launch so that no
// kernel of the replay walks from a cursor to an entry before it can load:
constexpr int kTsSlot = 7;    // [2][tiles] the record's pool slot at state L
constexpr int kTsRecm = 9;    // [2][tiles] the record's change mask of launch L - 1 (0: none)
constexpr int kTsLast = 11;   // [tiles] state of the tile's last record entry
constexpr int kTsSeeded = 12; // [tiles] (retained plane) seeded for launch 0 by this replay
constexpr int kTsWords = 13;

// Control words (FillRecord::ctl): 64-bit slot / entry counters, overflow,
// two scratch words, then one change flag per launch and one list count per
// launch (both cleared once per fill).
constexpr int kCtlUsed = 0;     // 2 x uint64
constexpr int kCtlOverflow = 4;
constexpr int kCtlCount = 5;    // scratch: visit total, free-slot count
constexpr int kCtlChanged = 8;

__global__ void rec_reset_kernel(int* nent, int tiles, uint32_t* pool0, int slot_cells,
                                 unsigned long long* used, int* overflow)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tiles) nent[i] = 0;
    if (i < slot_cells) pool0[i] = kInvalid;   // slot 0: every cell undefined
    if (i == 0) {
        used[0] = 0ull;
        used[1] = 0ull;
        *overflow = 0;
    }
}

// Entries (append order) -> by tile: csr[tile_off[tile] + rank] = (state, mask, slot, 0).
__global__ void rec_scatter_kernel(const int4* __restrict__ app, const int* __restrict__ rank,
                                   int64_t n, const uint32_t* __restrict__ tile_off,
                                   int4* __restrict__ csr)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 a = app[i];
    csr[tile_off[a.x] + rank[i]] = make_int4(a.y, a.z, a.w, 0);
}

// Slots used by a record.
__global__ void rec_mark_kernel(const int4* __restrict__ csr, int64_t n, uint32_t* mark)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int slot = csr[i].z;
    atomicOr(mark + (slot >> 5), 1u << (slot & 31));
}

// Slots in [1, hw) the record does not use, appended to freelist (any order).
__global__ __launch_bounds__(256) void rec_free_kernel(const uint32_t* __restrict__ mark,
                                                       int64_t hw, int* __restrict__ freelist,
                                                       int* count)
{
    __shared__ int wave_n[4], base;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool fr = i >= 1 && i < hw && !((mark[i >> 5] >> (i & 31)) & 1u);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(fr);
    if (lane == 0) wave_n[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
        base = t ? atomicAdd(count, t) : 0;
    }
    __syncthreads();
    if (fr) {
        int off = base;
        for (int k = 0; k < w; ++k) off += wave_n[k];
        freelist[off + __popcll(m & ((1ull << lane) - 1ull))] = static_cast<int>(i);
    }
}

// Replay set-up: cursors at the first entries, nothing dirty or changed yet.
__global__ void incr_init_kernel(const int4* __restrict__ csr, const int* __restrict__ tile_off,
                                 int* ts, int tiles, int* ctl, int n_flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tiles) {
        const int beg = tile_off[i], end = tile_off[i + 1];
        ts[kTsNext * tiles + i] = beg;
        ts[kTsDirty * tiles + i] = 0;
        ts[kTsMask * tiles + i] = 0;
        ts[kTsNent * tiles + i] = 0;
        ts[kTsSlot * tiles + i] = 0;   // (state 0 is read from the plane)
        ts[kTsRecm * tiles + i] = 0;
        ts[kTsLast * tiles + i] = end > beg ? csr[end - 1].x : 0;
        ts[kTsSeeded * tiles + i] = 0;
    }
    if (i < 3 * n_flags) ctl[kCtlChanged + i] = 0;   // change flags, list counts, busy flags
}

// Tiles holding a seed that only one of the two bitmaps has differ at state 0.
__global__ void incr_seed_diff_kernel(const uint32_t* __restrict__ prev,
                                      const uint32_t* __restrict__ cur, int64_t words, int img_w,
                                      int img_h,
                                      int P, int IW, int IH, int tiles_x, int* dirty)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= words) return;
    uint32_t d = prev[i] ^ cur[i];
    while (d) {
        const int b = __builtin_ctz(d);
        d &= d - 1u;
        const int64_t bit = i * 32 + b;   // bit x * img_h + y
        const int x = static_cast<int>(bit / img_h), y = static_cast<int>(bit % img_h);
        if (x >= img_w) continue;         // (bits past the image are never set)
        dirty[((y + P) / IH) * tiles_x + (x + P) / IW] = 1;
    }
}

// Replay onto a retained plane (the plane already holds the recorded fill's
// final labels): launch 0 reads the plane only in the 3 x 3 neighbourhoods
// of the tiles it computes, which lie next to tiles differing at state 0, so
// only tiles within two tiles of those need the seeded canvas (undefined
// cells and this fill's seeds, from the seed bitmap: bit x * img_h + y of the
// image, which sits at (P, P) in the canvas).  Every other tile keeps its
// final labels, which are this fill's too unless the replay touches it.
__global__ __launch_bounds__(256) void incr_seed_region_kernel(
    int W, int H, int tiles_x, int tiles_y, int IW, int IH, const int* __restrict__ dirty0,
    int* __restrict__ seeded, const uint32_t* __restrict__ bitmap, int img_w, int img_h, int P,
    uint32_t* __restrict__ plane)
{
    // The tile's bitmap words, per interior column: its rows [oy - P, oy - P + IH)
    // of the image are at most three consecutive words of the column's bits.
    constexpr int kMaxIW = 128, kWords = 3;
    __shared__ uint32_t words[kMaxIW][kWords];
    const int t = blockIdx.x;
    const int tx = t % tiles_x, ty = t / tiles_x;
    int near = 0;
    if (threadIdx.x < 25) {
        const int nx = tx + static_cast<int>(threadIdx.x) % 5 - 2;
        const int ny = ty + static_cast<int>(threadIdx.x) / 5 - 2;
        if (nx >= 0 && ny >= 0 && nx < tiles_x && ny < tiles_y) near = dirty0[ny * tiles_x + nx];
    }
    if (!__syncthreads_or(near)) return;   // block-uniform
    if (threadIdx.x == 0) seeded[t] = 1;
    const int ox = tx * IW, oy = ty * IH;
    const int y_lo = max(oy - P, 0);                          // first image row of the tile
    const int64_t last_word = (static_cast<int64_t>(img_w) * img_h - 1) >> 5;
    for (int j = threadIdx.x; j < IW * kWords; j += 256) {
        const int c = j / kWords, k = j % kWords;
        const int x = ox + c - P;
        uint32_t w = 0u;
        if (x >= 0 && x < img_w && y_lo < img_h) {
            const int64_t wi = ((static_cast<int64_t>(x) * img_h + y_lo) >> 5) + k;
            if (wi <= last_word) w = bitmap[wi];
        }
        words[c][k] = w;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < IW * IH; i += 256) {
        const int c = i % IW;
        const int cx = ox + c, cy = oy + i / IW;
        if (cx >= W || cy >= H) continue;
        uint32_t v = kInvalid;
        const int x = cx - P, y = cy - P;
        if (x >= 0 && y >= 0 && x < img_w && y < img_h) {
            const int64_t bit = static_cast<int64_t>(x) * img_h + y;
            const int64_t w0 = (static_cast<int64_t>(x) * img_h + y_lo) >> 5;
            if ((words[c][(bit >> 5) - w0] >> (bit & 31)) & 1u)
                v = (static_cast<uint32_t>(cx) << 16) | static_cast<uint32_t>(cy);
        }
        plane[static_cast<size_t>(cy) * W + cx] = v;
    }
}

// Per-tile replay state of one launch: parity L (cur, state L) and L + 1 (nxt).
struct ReplayTs {
    const int* next_cur;   // first record entry of state >= L
    int* next_nxt;
    const int* dirty_cur;
    int* dirty_nxt;
    const int* mask_cur;   // the new fill's change mask of launch L - 1
    int* mask_nxt;
    const int* slot_cur;   // the record's pool slot at state L
    int* slot_nxt;
    const int* recm_cur;   // the record's change mask of launch L - 1
    int* recm_nxt;
    const int* last;       // state of the tile's last record entry
};

// Launch L's plan: which tiles to compute; the others' next dirty flag and
// change mask; record entries of the tiles that follow the record; every
// tile's record slot and change mask at state L + 1.
// busy[L] is set unless the fill is quiet: nothing to compute in this launch
// and no record tile next to a tile near a dirty one changes in a later
// launch, so that no later launch computes anything either.  Nothing after
// convergence (the previous launch changed nothing) or a quiet launch.
//
// The kernel is a chain of dependent memory round trips on a few thousand
// tiles, so it is built for latency: every load that depends only on the tile
// index is issued before the uniform exit test, an entry and its successor are
// loaded together, neighbours' record masks come from the carried per-tile
// arrays, and the block's two atomics are issued back to back.
__global__ __launch_bounds__(256) void incr_plan_kernel(
    int tiles_x, int tiles_y, int L, const int4* __restrict__ csr,
    const int* __restrict__ tile_off, ReplayTs ts, int* __restrict__ list, int* count, int* ctl,
    int* busy, RecArgs rec)
{
    __shared__ int wave_c[4], wave_f[4], base_c;
    __shared__ unsigned long long base_f;
    int* changed = ctl + kCtlChanged;
    const int tiles = tiles_x * tiles_y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool own = t < tiles;
    const int tx = t % tiles_x, ty = t / tiles_x;
    // bit of neighbour (dx, dy) = (k % 3 - 1, k / 3 - 1) that faces this tile
    constexpr int kFacing[9] = {kChBR, kChBottom, kChBL, kChRight, kChAny,
                                kChLeft, kChTR, kChTop, kChTL};
    const int prev_changed = L > 0 ? changed[L - 1] : 1;
    const int prev_busy = L > 0 ? busy[L - 1] : 1;
    int e = 0, end = 0, dc = 0, slot_l = 0, nent = 0, nd = 0;
    if (own) {
        e = ts.next_cur[t];
        end = tile_off[t + 1];
        dc = ts.dirty_cur[t];
        slot_l = ts.slot_cur[t];
        nent = rec.nent[t];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int nx = tx + k % 3 - 1, ny = ty + k / 3 - 1;
            if (nx < 0 || ny < 0 || nx >= tiles_x || ny >= tiles_y) continue;
            nd |= ts.dirty_cur[ny * tiles_x + nx];
        }
    }
    if (prev_changed == 0 || prev_busy == 0) return;   // uniform
    // The entries at the cursor and after it (state -1: none).
    int4 c0 = make_int4(-1, 0, 0, 0), c1 = make_int4(-1, 0, 0, 0);
    if (own && e < end) c0 = csr[e];
    if (own && e + 1 < end) c1 = csr[e + 1];
    // (activity only matters next to a tile that differs from the record)
    int act = L == 0 ? 1 : 0;
    bool later = false;   // a neighbour's record changes in a later launch
    if (own && nd && L > 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int nx = tx + k % 3 - 1, ny = ty + k / 3 - 1;
            if (nx < 0 || ny < 0 || nx >= tiles_x || ny >= tiles_y) continue;
            const int u = ny * tiles_x + nx;
            act |= ((ts.mask_cur[u] | ts.recm_cur[u]) & kFacing[k]) ? 1 : 0;
            later |= ts.last[u] > L;
        }
    }
    const bool at_l = c0.x == L;          // an entry of state L at the cursor
    const int4 f1 = at_l ? c1 : c0;       // the entry after it
    const bool has1 = f1.x == L + 1;      // ... is of state L + 1
    const bool comp = own && nd && act;
    bool follow = false;                  // a record entry of state L + 1 to reuse
    if (own) {
        ts.next_nxt[t] = e + (at_l ? 1 : 0);
        ts.slot_nxt[t] = has1 ? f1.z : slot_l;
        ts.recm_nxt[t] = has1 ? f1.y : 0;
        if (comp || later) busy[L] = 1;
        if (!comp) {
            ts.dirty_nxt[t] = dc;
            int m = 0;
            if (!dc && has1) {
                m = f1.y;
                follow = true;
            }
            ts.mask_nxt[t] = m;
            if (m & kChAny) changed[L] = 1;
        }
    }
    // The list of tiles to compute and the followed entries: one atomic each
    // per block.
    const uint64_t mc = __ballot(comp), mf = __ballot(follow);
    if (lane == 0) {
        wave_c[w] = __popcll(mc);
        wave_f[w] = __popcll(mf);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nc = wave_c[0] + wave_c[1] + wave_c[2] + wave_c[3];
        const int nf = wave_f[0] + wave_f[1] + wave_f[2] + wave_f[3];
        const int bc = atomicAdd(count, nc);
        const unsigned long long bf = atomicAdd(rec.used + 1, static_cast<unsigned long long>(nf));
        base_c = bc;
        base_f = bf;
    }
    __syncthreads();
    int oc = base_c;
    unsigned long long of = base_f;
    for (int k = 0; k < w; ++k) {
        oc += wave_c[k];
        of += static_cast<unsigned long long>(wave_f[k]);
    }
    const uint64_t below = (1ull << lane) - 1ull;
    if (comp) list[oc + __popcll(mc & below)] = t;
    if (follow) {
        const unsigned long long id = of + static_cast<unsigned long long>(__popcll(mf & below));
        if (id < static_cast<unsigned long long>(rec.ent_cap)) {
            rec.app[id] = make_int4(t, L + 1, f1.y, f1.z);
            rec.rank[id] = nent;
            rec.nent[t] = nent + 1;
        } else {
            *rec.overflow = 1;
        }
    }
}

// After a quiet launch (state st), the rest of the fill is the record's
// where it agrees with the record and the current state elsewhere: each
// tile's record entries of a later state join the new record.
__global__ __launch_bounds__(256) void incr_tail_kernel(int tiles, int st,
                                                        const int4* __restrict__ csr,
                                                        const int* __restrict__ tile_off,
                                                        const int* __restrict__ next_ent, RecArgs rec)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int e = 0, n = 0;
    if (t < tiles) {
        e = next_ent[t];   // first entry of a state >= st
        const int end = tile_off[t + 1];
        if (e < end && csr[e].x == st) ++e;   // (state st joined in the last launch)
        n = end - e;
    }
    int incl = n;   // inclusive prefix over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const int tot = __shfl(incl, 63, 64);
    if (tot == 0) return;
    unsigned long long b0 = 0ull;
    if (lane == 63) b0 = atomicAdd(rec.used + 1, static_cast<unsigned long long>(tot));
    b0 = __shfl(b0, 63, 64);
    if (n == 0) return;
    const unsigned long long id0 = b0 + static_cast<unsigned long long>(incl - n);
    const int r0 = rec.nent[t];
    for (int k = 0; k < n; ++k) {
        const unsigned long long id = id0 + k;
        if (id < static_cast<unsigned long long>(rec.ent_cap)) {
            const int4 c = csr[e + k];
            rec.app[id] = make_int4(t, c.x, c.y, c.z);
            rec.rank[id] = r0 + k;
        } else {
            *rec.overflow = 1;
        }
    }
    rec.nent[t] = r0 + n;
}

// Clean tiles take the record's state at state `st` (into `plane`).  With
// next_ent = tile_off + 1 they take the record's last (final) state.
// seeded (retained plane, else null): only tiles the replay seeded or computed
// (visits) are written; the others hold that state already.
__global__ __launch_bounds__(256) void incr_materialise_kernel(
    int W, int H, int tiles_x, int IW, int IH, int st, const int4* __restrict__ csr,
    const int* __restrict__ tile_off, const int* __restrict__ next_ent,
    const int* __restrict__ dirty, const int* __restrict__ seeded, const int* __restrict__ visits,
    const uint32_t* __restrict__ pool, uint32_t* __restrict__ plane)
{
    const int t = blockIdx.x;
    if (dirty[t]) return;
    if (seeded && !seeded[t] && !visits[t]) return;
    const int e = next_ent[t], beg = tile_off[t], end = tile_off[t + 1];
    const int slot = (e < end && csr[e].x == st) ? csr[e].z : (e > beg ? csr[e - 1].z : 0);
    const uint32_t* src = pool + static_cast<size_t>(slot) * IW * IH;
    const int ox = (t % tiles_x) * IW, oy = (t / tiles_x) * IH;
    for (int i = threadIdx.x; i < IW * IH; i += 256) {
        const int x = ox + i % IW, y = oy + i / IW;
        if (x < W && y < H) plane[static_cast<size_t>(y) * W + x] = src[i];
    }
}

const ListKernel kIncrKernel[2] = {bfs_list_kernel<8, 8, 8, true, 2>,
                                   bfs_list_kernel<8, 8, 8, true, 2, true>};

int tiles_of(const BfsCanvas& cv, const BfsConfig& cfg, int* tx, int* ty)
{
    *tx = (cv.W + bfs_tile_iw(cfg) - 1) / bfs_tile_iw(cfg);
    *ty = (cv.H + bfs_tile_ih(cfg) - 1) / bfs_tile_ih(cfg);
    return *tx * *ty;
}

// Capacities and reset of the recording buffers for a fill of `tiles` tiles.
// reset_slots: a new record from scratch (the pool is reused from slot 1).
RecArgs rec_prepare(FillRecord& rec, const BfsConfig& cfg, int tiles, int max_launches,
                    bool reset_slots, hipStream_t s)
{
    const int cells = bfs_tile_iw(cfg) * bfs_tile_ih(cfg);
    if (rec.pool_cap == 0) rec.pool_cap = std::max<int64_t>(4096, 24LL * tiles + 1);
    rec.pool.reserve(sizeof(uint32_t) * static_cast<size_t>(cells) *
                     static_cast<size_t>(rec.pool_cap));
    rec.ent_cap = 2 * rec.pool_cap;
    rec.app.reserve(sizeof(int4) * static_cast<size_t>(rec.ent_cap));
    rec.rank.reserve(sizeof(int) * static_cast<size_t>(rec.ent_cap));
    rec.tstate.reserve(sizeof(int) * kTsWords * static_cast<size_t>(tiles));
    rec.ctl.reserve(sizeof(int) * (kCtlChanged + 3 * static_cast<size_t>(max_launches) + 8));
    rec.list.reserve(sizeof(int) * static_cast<size_t>(tiles));
    rec.host.reserve(sizeof(int) * 64);
    RecArgs a{};
    a.pool = rec.pool.as<uint32_t>();
    a.pool_cap = rec.pool_cap;
    a.app = rec.app.as<int4>();
    a.rank = rec.rank.as<int>();
    a.nent = rec.tstate.as<int>() + kTsNent * tiles;
    a.used = reinterpret_cast<unsigned long long*>(rec.ctl.as<int>() + kCtlUsed);
    a.overflow = rec.ctl.as<int>() + kCtlOverflow;
    a.ent_cap = rec.ent_cap;
    if (reset_slots) {
        rec.hw = 1;
        rec.free_n = 0;
    }
    a.freelist = rec.freelist.as<int>();
    a.free_n = rec.free_n;
    a.fresh = rec.hw;
    const int n = std::max(tiles, cells);
    rec_reset_kernel<<<(n + 255) / 256, 256, 0, s>>>(a.nent, tiles, a.pool, cells, a.used,
                                                     a.overflow);
    DL_LAUNCH_CHECK();
    return a;
}

// Entries recorded by the fill just run -> the record (by tile).  Waits for
// the stream.  False when something overflowed (no valid record then).
bool rec_finish(FillRecord& rec, const RecArgs& a, int tiles, DevBuf& tmp, hipStream_t s)
{
    int* h = rec.host.as<int>();
    DL_HIP(hipMemcpyAsync(h, rec.ctl.ptr, sizeof(int) * (kCtlOverflow + 1), hipMemcpyDeviceToHost,
                          s));
    DL_HIP(hipStreamSynchronize(s));
    const unsigned long long* used = reinterpret_cast<const unsigned long long*>(h + kCtlUsed);
    if (h[kCtlOverflow] || used[1] > static_cast<unsigned long long>(rec.ent_cap)) {
        rec.valid = false;
        // the next record gets twice the room
        rec.pool_cap *= 2;
        return false;
    }
    const int nx = rec.cur ^ 1;
    const int64_t n = static_cast<int64_t>(used[1]);
    rec.tile_off[nx].reserve(sizeof(uint32_t) * (static_cast<size_t>(tiles) + 1));
    rec.csr[nx].reserve(sizeof(int4) * static_cast<size_t>(std::max<int64_t>(n, 1)));
    exclusive_scan_u32(reinterpret_cast<const uint32_t*>(a.nent), tiles,
                       rec.tile_off[nx].as<uint32_t>(), tmp, s);
    if (n > 0) {
        rec_scatter_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
            a.app, a.rank, n, rec.tile_off[nx].as<uint32_t>(), rec.csr[nx].as<int4>());
        DL_LAUNCH_CHECK();
    }
    rec.cur = nx;
    rec.entries = n;
    rec.valid = true;
    static std::atomic<uint64_t> g_rec_gen{0};
    rec.gen = ++g_rec_gen;
    // Slots this fill took beyond the free list extend the pool; every slot
    // below the new high-water mark that the new record does not use is
    // free for the next fill.
    // (A replay takes a slot for every tile it computes, used or not; slots
    // at or past the capacity were never written.)
    const int64_t taken = static_cast<int64_t>(used[0]);
    rec.hw = std::min<int64_t>(a.fresh + std::max<int64_t>(0, taken - a.free_n), rec.pool_cap);
    const size_t words = static_cast<size_t>((rec.hw + 31) / 32);
    rec.mark.reserve(sizeof(uint32_t) * words);
    rec.freelist.reserve(sizeof(int) * static_cast<size_t>(rec.pool_cap));
    fill_u32(rec.mark.as<uint32_t>(), 0u, words, s);
    if (n > 0) {
        rec_mark_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
            rec.csr[nx].as<int4>(), n, rec.mark.as<uint32_t>());
        DL_LAUNCH_CHECK();
    }
    int* cnt = rec.ctl.as<int>() + kCtlCount + 2;
    zero_ints_kernel<<<1, 64, 0, s>>>(cnt, 1);
    DL_LAUNCH_CHECK();
    rec_free_kernel<<<static_cast<unsigned>((rec.hw + 255) / 256), 256, 0, s>>>(
        rec.mark.as<uint32_t>(), rec.hw, rec.freelist.as<int>(), cnt);
    DL_LAUNCH_CHECK();
    DL_HIP(hipMemcpyAsync(h + 40, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    rec.free_n = h[40];
    return true;
}

}  // namespace

bool bfs_record_supported(const BfsConfig& cfg, const BfsCanvas& cv)
{
    return cfg.waves == 8 && cfg.rows == 8 && cfg.iters == 8 && bfs_min_encoding_ok(cv.W, cv.H);
}

bool bfs_record_usable(const FillRecord& rec, const BfsCanvas& cv, const BfsConfig& cfg)
{
    return rec.valid && rec.synced && !rec.escaped && rec.W == cv.W && rec.H == cv.H &&
           rec.P == cv.P &&
           rec.cfg.waves == cfg.waves && rec.cfg.rows == cfg.rows && rec.cfg.iters == cfg.iters;
}

int bfs_run_recorded(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                     hipStream_t s, BfsStats* stats, FillRecord& rec)
{
    if (canvases.size() != 1 || !bfs_record_supported(cfg, canvases[0]))
        throw runtime_error("fill records need one canvas and the default tile shape");
    int tx = 0, ty = 0;
    const int tiles = tiles_of(canvases[0], cfg, &tx, &ty);
    const int max_launches =
        static_cast<int>((8LL * (canvases[0].W + canvases[0].H) + 64 + cfg.iters - 1) / cfg.iters) +
        2;
    rec.valid = false;
    rec.synced = false;
    work.enc = 1;   // (the recording kernels use the min encoding)
    TileExtra ex{};
    ex.rec = rec_prepare(rec, cfg, tiles, max_launches, true, s);
    BfsStats st;
    const bool allow = work.allow_min_encoding;
    work.allow_min_encoding = true;
    int plane = 0;
    try {
        plane = run_impl(canvases, cfg, work, s, &st, nullptr, 0, &ex);
    } catch (...) {
        work.allow_min_encoding = allow;
        throw;
    }
    work.allow_min_encoding = allow;
    if (rec_finish(rec, ex.rec, tiles, work.scan_tmp, s)) {
        rec.W = canvases[0].W;
        rec.H = canvases[0].H;
        rec.P = canvases[0].P;
        rec.tiles = tiles;
        rec.cfg = cfg;
        rec.states = st.converged + 2;
        rec.escaped = false;
    }
    rec.last_computed = st.tile_visits;
    if (stats) *stats = st;
    return plane;
}

int bfs_run_incremental(std::vector<BfsCanvas>& canvases, const BfsConfig& cfg, BfsWork& work,
                        hipStream_t s, BfsStats* stats, FillRecord& rec,
                        const uint32_t* prev_bitmap, const uint32_t* bitmap, int img_w, int img_h,
                        int img_pad, bool retained)
{
    if (canvases.size() != 1 || !bfs_record_usable(rec, canvases[0], cfg))
        throw runtime_error("no usable fill record for this canvas");
    const Variant* v = find_variant(cfg);
    int64_t max_iters = 0;
    const int total = bfs_setup(canvases, cfg, work, s, &max_iters);
    const int tiles = rec.tiles;
    if (total != tiles) throw runtime_error("fill record and canvas disagree");
    const int max_launches = static_cast<int>((max_iters + cfg.iters - 1) / cfg.iters) + 2;
    const BfsCanvas& cv = canvases[0];
    const int IW = bfs_tile_iw(cfg), IH = bfs_tile_ih(cfg);
    const int tiles_x = cv.tiles_x, tiles_y = cv.tiles_y;
    // The record being replayed stays in csr[cur] / tile_off[cur]; this fill's
    // entries are appended after its pool slots.  Until the caller has
    // stored this fill's seed bitmap, no later fill may replay.
    rec.synced = false;
    const int cur = rec.cur;
    RecArgs ra = rec_prepare(rec, cfg, tiles, max_launches, false, s);
    const int4* csr = rec.csr[cur].as<int4>();
    const int* toff = rec.tile_off[cur].as<int>();
    int* ts = rec.tstate.as<int>();
    int* ctl = rec.ctl.as<int>();
    int* list = rec.list.as<int>();
    work.visits.reserve(sizeof(int) * static_cast<size_t>(tiles));
    fill_u32(work.visits.as<uint32_t>(), 0u, static_cast<size_t>(tiles), s);
    incr_init_kernel<<<(std::max(tiles, 3 * max_launches) + 255) / 256, 256, 0, s>>>(
        csr, toff, ts, tiles, ctl, max_launches);
    DL_LAUNCH_CHECK();
    {
        const int64_t words = (static_cast<int64_t>(img_w) * img_h + 31) / 32;
        incr_seed_diff_kernel<<<static_cast<unsigned>((words + 255) / 256), 256, 0, s>>>(
            prev_bitmap, bitmap, words, img_w, img_h, img_pad, IW, IH, tiles_x,
            ts + kTsDirty * tiles);
        DL_LAUNCH_CHECK();
    }
    const int* seeded = retained ? ts + kTsSeeded * tiles : nullptr;
    if (retained) {
        incr_seed_region_kernel<<<tiles, 256, 0, s>>>(cv.W, cv.H, tiles_x, tiles_y, IW, IH,
                                                      ts + kTsDirty * tiles, ts + kTsSeeded * tiles,
                                                      bitmap, img_w, img_h, img_pad, cv.lab0);
        DL_LAUNCH_CHECK();
    }
    const BfsCanvas* dcv = reinterpret_cast<const BfsCanvas*>(work.canvases.ptr);
    int* changed = ctl + kCtlChanged;                  // per launch, cleared above
    int* counts = changed + max_launches;
    int* busy = changed + 2 * max_launches;
    auto issue = [&](int L) {
        const int p = L & 1, q = p ^ 1;
        ProfScope ps(kKFill, s);
        ReplayTs rt{};
        rt.next_cur = ts + (kTsNext + p) * tiles;
        rt.next_nxt = ts + (kTsNext + q) * tiles;
        rt.dirty_cur = ts + (kTsDirty + p) * tiles;
        rt.dirty_nxt = ts + (kTsDirty + q) * tiles;
        rt.mask_cur = ts + (kTsMask + p) * tiles;
        rt.mask_nxt = ts + (kTsMask + q) * tiles;
        rt.slot_cur = ts + (kTsSlot + p) * tiles;
        rt.slot_nxt = ts + (kTsSlot + q) * tiles;
        rt.recm_cur = ts + (kTsRecm + p) * tiles;
        rt.recm_nxt = ts + (kTsRecm + q) * tiles;
        rt.last = ts + kTsLast * tiles;
        incr_plan_kernel<<<(tiles + 255) / 256, 256, 0, s>>>(tiles_x, tiles_y, L, csr, toff, rt,
                                                            list, counts + L, ctl, busy, ra);
        DL_LAUNCH_CHECK();
        TileExtra ex{};
        ex.rec = ra;
        ex.launch = L;
        ex.slot_cur = rt.slot_cur;
        ex.slot_next = rt.slot_nxt;
        ex.dirty_cur = ts + (kTsDirty + p) * tiles;
        ex.dirty_next = ts + (kTsDirty + q) * tiles;
        ex.mask_next = ts + (kTsMask + q) * tiles;
        kIncrKernel[work.aligned ? 1 : 0]<<<kListGrid, v->waves * 64, 0, s>>>(
            dcv, 1, list, counts + L, p, nullptr, nullptr, changed + L, work.visits.as<int>(), ex);
        DL_LAUNCH_CHECK();
    };
    // Launches are issued a window ahead of the flags read back: the host
    // waits for the flags of launches [checked, upto) while the window runs.
    // The replay ends at the first launch that changed nothing (converged) or
    // that was quiet; launches issued past that point do nothing.
    constexpr int kWindow = 16;
    rec.host.reserve(sizeof(int) * (2 * static_cast<size_t>(max_launches) + 64));
    int* h = rec.host.as<int>();
    int* h_busy = h + max_launches;
    const hipEvent_t polled = rec.polled.get();
    int issued = 0, checked = 0, stop = -1;
    bool converged = false;
    while (issued < std::min(max_launches, kWindow)) issue(issued++);
    while (stop < 0) {
        const int upto = issued, n = upto - checked;
        DL_HIP(hipMemcpyAsync(h + checked, changed + checked, sizeof(int) * n,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipMemcpyAsync(h_busy + checked, busy + checked, sizeof(int) * n,
                              hipMemcpyDeviceToHost, s));
        DL_HIP(hipEventRecord(polled, s));
        for (int j = 0; j < kWindow && issued < max_launches; ++j) issue(issued++);
        DL_HIP(hipEventSynchronize(polled));
        for (int L = checked; L < upto && stop < 0; ++L) {
            if (h[L] == 0) {
                stop = L;
                converged = true;
            } else if (h_busy[L] == 0) {
                stop = L;
            }
        }
        checked = upto;
        if (stop < 0 && checked >= max_launches)
            throw runtime_error("flood fill did not converge within the iteration bound");
    }
    // The state arrays of parity (stop + 1) & 1 describe state stop + 1.
    const int q = (stop + 1) & 1;
    uint32_t* out_plane = cv.lab0;
    int conv = stop;
    if (converged) {
        incr_materialise_kernel<<<tiles, 256, 0, s>>>(
            cv.W, cv.H, tiles_x, IW, IH, conv + 1, csr, toff, ts + (kTsNext + q) * tiles,
            ts + (kTsDirty + q) * tiles, seeded, work.visits.as<int>(), ra.pool, out_plane);
        DL_LAUNCH_CHECK();
    } else {
        // No tile is computed after the quiet launch `stop`: a dirty tile's
        // state (in both planes) is final, a clean one follows the record,
        // which changes until its own convergence.  The record's later entries
        // join the new record and its final states are the clean tiles'.
        conv = std::max(rec.states - 2, stop + 1);
        incr_tail_kernel<<<(tiles + 255) / 256, 256, 0, s>>>(tiles, stop + 1, csr, toff,
                                                              ts + (kTsNext + q) * tiles, ra);
        DL_LAUNCH_CHECK();
        incr_materialise_kernel<<<tiles, 256, 0, s>>>(cv.W, cv.H, tiles_x, IW, IH, conv + 1, csr,
                                                       toff, toff + 1, ts + (kTsDirty + q) * tiles,
                                                       seeded, work.visits.as<int>(), ra.pool,
                                                       out_plane);
        DL_LAUNCH_CHECK();
    }
    // visit total (statistics)
    BfsStats st;
    st.converged = conv;
    st.launches = issued;
    sum_ints_kernel<<<1, 1024, 0, s>>>(work.visits.as<int>(), tiles, ctl + kCtlCount + 1);
    DL_LAUNCH_CHECK();
    DL_HIP(hipMemcpyAsync(h, ctl + kCtlCount + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    st.tile_visits = h[0];
    if (rec_finish(rec, ra, tiles, work.scan_tmp, s)) rec.states = conv + 2;
    rec.last_computed = st.tile_visits;
    if (stats) *stats = st;
    return 0;
}

}  // namespace dl
