// Device-resident seed sets of incremental sessions.
//
// A session keeps its live seeds on the device as a presence bitmap (bit
// x*H + y, the same layout as the seed ordering uses) and a compact (x, y) list
// in insertion order.  Inserts and removals are validated on the device with
// atomic bit updates; a failing batch is rolled back bit by bit, so a call that
// raises leaves the session unchanged (the reference checks its inputs before
// touching state, delaunay.cu:191-211).  The pipeline then runs on the live
// list exactly as on a fresh seed array, so the session's results are, bit for
// bit, the batch results on its current seed set.
#include "dl_internal.h"

#include <algorithm>
#include <utility>

namespace dl {
namespace {

constexpr uint32_t kSessRange = 1u;     // a coordinate outside the image
constexpr uint32_t kSessConflict = 2u;  // insert: already live / twice in the batch;
                                        // remove: not live / twice in the batch

// Insert: set each seed's bit; note in `mine` which bits this call set.
__global__ void sess_insert_kernel(const int32_t* __restrict__ xy, int64_t n, int W, int H,
                                   uint32_t* __restrict__ bm, uint8_t* __restrict__ mine,
                                   uint32_t* __restrict__ flags)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mine[i] = 0;
    const int x = xy[2 * i], y = xy[2 * i + 1];
    if (x < 0 || x >= W || y < 0 || y >= H) {
        atomicOr(flags, kSessRange);
        return;
    }
    const uint64_t bit = static_cast<uint64_t>(x) * H + y;
    const uint32_t m = 1u << (bit & 31);
    const uint32_t old = atomicOr(&bm[bit >> 5], m);
    if (old & m) atomicOr(flags, kSessConflict);
    else mine[i] = 1;
}

// Remove: clear each seed's bit; note in `mine` which bits this call cleared.
__global__ void sess_remove_kernel(const int32_t* __restrict__ xy, int64_t n, int W, int H,
                                   uint32_t* __restrict__ bm, uint8_t* __restrict__ mine,
                                   uint32_t* __restrict__ flags)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mine[i] = 0;
    const int x = xy[2 * i], y = xy[2 * i + 1];
    if (x < 0 || x >= W || y < 0 || y >= H) {
        atomicOr(flags, kSessConflict);
        return;
    }
    const uint64_t bit = static_cast<uint64_t>(x) * H + y;
    const uint32_t m = 1u << (bit & 31);
    const uint32_t old = atomicAnd(&bm[bit >> 5], ~m);
    if (old & m) mine[i] = 1;
    else atomicOr(flags, kSessConflict);
}

// Undo the bit updates of a failed call: restore = 1 sets the bits back
// (failed removal), restore = 0 clears them (failed insert).
__global__ void sess_rollback_kernel(const int32_t* __restrict__ xy, int64_t n, int H,
                                     uint32_t* __restrict__ bm, const uint8_t* __restrict__ mine,
                                     int restore)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !mine[i]) return;
    const uint64_t bit = static_cast<uint64_t>(xy[2 * i]) * H + xy[2 * i + 1];
    const uint32_t m = 1u << (bit & 31);
    if (restore) atomicOr(&bm[bit >> 5], m);
    else atomicAnd(&bm[bit >> 5], ~m);
}

// Sets the bits of n positions (restores a removal that succeeded).
__global__ void sess_set_bits_kernel(const int32_t* __restrict__ xy, int64_t n, int H,
                                     uint32_t* __restrict__ bm)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t bit = static_cast<uint64_t>(xy[2 * i]) * H + xy[2 * i + 1];
    atomicOr(&bm[bit >> 5], 1u << (bit & 31));
}

// keep[j] = 1 if list entry j is still live.
__global__ void sess_keep_kernel(const int32_t* __restrict__ xy, int64_t n, int H,
                                 const uint32_t* __restrict__ bm, uint32_t* __restrict__ keep)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t bit = static_cast<uint64_t>(xy[2 * j]) * H + xy[2 * j + 1];
    keep[j] = (bm[bit >> 5] >> (bit & 31)) & 1u;
}

// Moves the kept entries (and their values) to their compacted positions.
__global__ void sess_scatter_kernel(const int32_t* __restrict__ xy, const float* __restrict__ val,
                                    int64_t n, const uint32_t* __restrict__ keep,
                                    const uint32_t* __restrict__ pos, int32_t* __restrict__ out,
                                    float* __restrict__ val_out)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n || !keep[j]) return;
    const uint32_t p = pos[j];
    out[2 * p] = xy[2 * j];
    out[2 * p + 1] = xy[2 * j + 1];
    val_out[p] = val[j];
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, (n + 255) / 256)); }

}  // namespace

void session_clear(SessionState& st, hipStream_t s)
{
    fill_u32(st.bitmap.as<uint32_t>(), 0u, static_cast<size_t>(st.words), s);
    st.n_live = 0;
    st.have_values = false;
}

void session_alloc(SessionState& st, int W, int H, int64_t max_seeds, hipStream_t s)
{
    st.W = W;
    st.H = H;
    st.max_seeds = max_seeds;
    st.words = (static_cast<int64_t>(W) * H + 31) / 32;
    st.bitmap.reserve(sizeof(uint32_t) * static_cast<size_t>(st.words));
    st.list.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(max_seeds));
    st.spare.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(max_seeds));
    st.values.reserve(sizeof(float) * static_cast<size_t>(max_seeds));
    st.values_spare.reserve(sizeof(float) * static_cast<size_t>(max_seeds));
    session_clear(st, s);
}

void session_insert(SessionState& st, const int32_t* d_xy, const float* d_values, int64_t n,
                    hipStream_t s)
{
    if (n <= 0) return;
    if (st.n_live + n > st.max_seeds) throw value_error("insert would exceed max_seeds capacity");
    st.mine.reserve(static_cast<size_t>(n));
    st.flags.reserve(sizeof(uint32_t) * 2);
    st.host_flags.reserve(sizeof(uint32_t) * 2);
    uint32_t* flags = st.flags.as<uint32_t>();
    fill_u32(flags, 0u, 2, s);
    sess_insert_kernel<<<blocks_for(n), 256, 0, s>>>(d_xy, n, st.W, st.H, st.bitmap.as<uint32_t>(),
                                                     st.mine.as<uint8_t>(), flags);
    DL_LAUNCH_CHECK();
    DL_HIP(hipMemcpyAsync(st.host_flags.ptr, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    const uint32_t f = st.host_flags.as<uint32_t>()[0];
    if (f) {
        sess_rollback_kernel<<<blocks_for(n), 256, 0, s>>>(d_xy, n, st.H, st.bitmap.as<uint32_t>(),
                                                           st.mine.as<uint8_t>(), 0);
        DL_LAUNCH_CHECK();
        DL_HIP(hipStreamSynchronize(s));
        if (f & kSessRange) throw value_error("seed coordinate out of bounds");
        throw value_error("seed position already exists or repeats within the batch");
    }
    DL_HIP(hipMemcpyAsync(st.list.as<int32_t>() + 2 * st.n_live, d_xy,
                          sizeof(int32_t) * 2 * static_cast<size_t>(n), hipMemcpyDeviceToDevice, s));
    float* v = st.values.as<float>() + st.n_live;
    if (d_values) {
        DL_HIP(hipMemcpyAsync(v, d_values, sizeof(float) * static_cast<size_t>(n),
                              hipMemcpyDeviceToDevice, s));
        st.have_values = true;
    } else {
        fill_u32(reinterpret_cast<uint32_t*>(v), 0u, static_cast<size_t>(n), s);   // 0.0f
    }
    st.n_live += n;
}

void session_remove(SessionState& st, const int32_t* d_xy, int64_t n, DevBuf& tmp, hipStream_t s)
{
    if (n <= 0) return;
    st.mine.reserve(static_cast<size_t>(n));
    st.flags.reserve(sizeof(uint32_t) * 2);
    st.host_flags.reserve(sizeof(uint32_t) * 2);
    uint32_t* flags = st.flags.as<uint32_t>();
    fill_u32(flags, 0u, 2, s);
    sess_remove_kernel<<<blocks_for(n), 256, 0, s>>>(d_xy, n, st.W, st.H, st.bitmap.as<uint32_t>(),
                                                     st.mine.as<uint8_t>(), flags);
    DL_LAUNCH_CHECK();
    DL_HIP(hipMemcpyAsync(st.host_flags.ptr, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    if (st.host_flags.as<uint32_t>()[0]) {
        sess_rollback_kernel<<<blocks_for(n), 256, 0, s>>>(d_xy, n, st.H, st.bitmap.as<uint32_t>(),
                                                           st.mine.as<uint8_t>(), 1);
        DL_LAUNCH_CHECK();
        DL_HIP(hipStreamSynchronize(s));
        throw value_error("seed position is not present or repeats within the batch");
    }
    // Compact the list in order: keep the entries whose bit is still set.
    const int64_t m = st.n_live;
    st.keep.reserve(sizeof(uint32_t) * 2 * static_cast<size_t>(m + 1));
    uint32_t* keep = st.keep.as<uint32_t>();
    uint32_t* pos = keep + m;   // m + 1 entries: exclusive scan and total
    sess_keep_kernel<<<blocks_for(m), 256, 0, s>>>(st.list.as<int32_t>(), m, st.H,
                                                   st.bitmap.as<uint32_t>(), keep);
    DL_LAUNCH_CHECK();
    exclusive_scan_u32(keep, m, pos, tmp, s);
    sess_scatter_kernel<<<blocks_for(m), 256, 0, s>>>(st.list.as<int32_t>(), st.values.as<float>(),
                                                      m, keep, pos, st.spare.as<int32_t>(),
                                                      st.values_spare.as<float>());
    DL_LAUNCH_CHECK();
    std::swap(st.list.ptr, st.spare.ptr);
    std::swap(st.list.bytes, st.spare.bytes);
    std::swap(st.values.ptr, st.values_spare.ptr);
    std::swap(st.values.bytes, st.values_spare.bytes);
    st.n_live -= n;
}

void session_update(SessionState& st, const int32_t* d_rm, int64_t n_rm, const int32_t* d_ins,
                    const float* d_values, int64_t n_ins, DevBuf& tmp, hipStream_t s)
{
    if (n_rm <= 0) {
        session_insert(st, d_ins, d_values, n_ins, s);
        return;
    }
    // Copies of the list and values, to undo the removal if the insert fails.
    const int64_t n0 = st.n_live;
    st.backup_list.reserve(sizeof(int32_t) * 2 * static_cast<size_t>(std::max<int64_t>(n0, 1)));
    st.backup_values.reserve(sizeof(float) * static_cast<size_t>(std::max<int64_t>(n0, 1)));
    DL_HIP(hipMemcpyAsync(st.backup_list.ptr, st.list.ptr, sizeof(int32_t) * 2 * static_cast<size_t>(n0),
                          hipMemcpyDeviceToDevice, s));
    DL_HIP(hipMemcpyAsync(st.backup_values.ptr, st.values.ptr, sizeof(float) * static_cast<size_t>(n0),
                          hipMemcpyDeviceToDevice, s));
    session_remove(st, d_rm, n_rm, tmp, s);
    if (n_ins <= 0) return;
    try {
        session_insert(st, d_ins, d_values, n_ins, s);
    } catch (...) {
        sess_set_bits_kernel<<<blocks_for(n_rm), 256, 0, s>>>(d_rm, n_rm, st.H, st.bitmap.as<uint32_t>());
        DL_LAUNCH_CHECK();
        DL_HIP(hipMemcpyAsync(st.list.ptr, st.backup_list.ptr, sizeof(int32_t) * 2 * static_cast<size_t>(n0),
                              hipMemcpyDeviceToDevice, s));
        DL_HIP(hipMemcpyAsync(st.values.ptr, st.backup_values.ptr,
                              sizeof(float) * static_cast<size_t>(n0), hipMemcpyDeviceToDevice, s));
        DL_HIP(hipStreamSynchronize(s));
        st.n_live = n0;
        throw;
    }
}

}  // namespace dl
