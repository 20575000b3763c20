// vdb_gridbuild.hip -- the GridBuild function table (nanovdb_editor/putil/GridBuild.h:60-152) on
// this library's sorted-key node2 builder.
//
// The reference inserts each fragment window's points into the tree in place (four levels of mask
// atomics, new-child counts, a scan and node allocation per window; GridBuild.cpp:361-506). Here a
// window's points are compacted to 51-bit voxel keys, sorted and deduplicated, and appended with
// the window's index; grid_build_finalize sorts all keys once more, keeps each voxel's first window
// and hands the list to the node2 builder of the raster path, which reproduces the reference's
// allocation order (window, level, parent, slot) from those first windows. The fanout state
// machine (GridBuild.cpp:654-942) maps window slots to (prim, footprint index) with one binary
// search of the inclusive footprint scan per slot instead of a scatter + max-scan pass.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "vdb_device.h"
#include "vdb_internal.h"

namespace vdbr
{

constexpr int32_t GB_SKIP = 0x7FFFFFFF;    // grid_build_set_masks.slang:52
constexpr int32_t GB_LO = -65536, GB_HI = 65535; // node2 root range of this library
constexpr uint32_t GB_KEY_BITS = 51;        // 17 bits per axis

__device__ __forceinline__ uint64_t gb_key(int32_t i, int32_t j, int32_t k)
{
    return ((uint64_t)(uint32_t)(i - GB_LO) << 34) | ((uint64_t)(uint32_t)(j - GB_LO) << 17) | (uint64_t)(uint32_t)(k - GB_LO);
}

__device__ __forceinline__ void gb_unkey(uint64_t key, int32_t* i, int32_t* j, int32_t* k)
{
    *i = (int32_t)((key >> 34) & 0x1FFFFu) + GB_LO;
    *j = (int32_t)((key >> 17) & 0x1FFFFu) + GB_LO;
    *k = (int32_t)(key & 0x1FFFFu) + GB_LO;
}

// valid points of a window -> keys, compacted per wave (one counter reservation per wave)
__global__ __launch_bounds__(256) void k_gb_collect(const int32_t* __restrict__ ijk, uint64_t n,
                                                    uint64_t* __restrict__ keys, unsigned long long* __restrict__ counters)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const uint64_t p = base + threadIdx.x;
        bool keep = false;
        bool bad = false;
        uint64_t key = 0;
        if (p < n)
        {
            const int32_t i = ijk[3 * p + 0], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
            if (k != GB_SKIP)
            {
                bad = i < GB_LO || i > GB_HI || j < GB_LO || j > GB_HI || k < GB_LO || k > GB_HI;
                keep = !bad;
                key = keep ? gb_key(i, j, k) : 0ull;
            }
        }
        const uint64_t mask = __ballot(keep);
        const uint64_t badmask = __ballot(bad);
        unsigned long long wbase = 0;
        if (lane == 0u)
        {
            if (mask)
                wbase = atomicAdd(&counters[0], (unsigned long long)__popcll(mask));
            if (badmask)
                atomicAdd(&counters[1], (unsigned long long)__popcll(badmask));
        }
        wbase = __shfl(wbase, 0, 64);
        if (keep)
            keys[wbase + __popcll(mask & ((1ull << lane) - 1ull))] = key;
    }
}

// run heads of sorted keys: flag[i] = key[i] differs from key[i - 1]
__global__ __launch_bounds__(256) void k_gb_heads(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flag)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// first entry of each run -> out[excl[i]] (with its payload, or the constant t without one)
__global__ __launch_bounds__(256) void k_gb_unique(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ t_in,
                                                   uint32_t t_const, const uint32_t* __restrict__ flag,
                                                   const uint32_t* __restrict__ excl, uint64_t n,
                                                   uint64_t* __restrict__ out_keys, uint32_t* __restrict__ out_t)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i])
    {
        out_keys[excl[i]] = keys[i];
        out_t[excl[i]] = t_in ? t_in[i] : t_const;
    }
}

__global__ __launch_bounds__(256) void k_gb_bounds(const uint64_t* __restrict__ keys, uint64_t n, int32_t* __restrict__ bounds)
{
    int32_t mn[3] = { 2147483647, 2147483647, 2147483647 }, mx[3] = { -2147483647, -2147483647, -2147483647 };
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
    {
        int32_t c[3];
        gb_unkey(keys[v], &c[0], &c[1], &c[2]);
        for (int a = 0; a < 3; a++)
        {
            mn[a] = min(mn[a], c[a]);
            mx[a] = max(mx[a], c[a]);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++)
        {
            mn[a] = min(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], off, 64));
        }
    if ((threadIdx.x & 63u) == 0u)
        for (int a = 0; a < 3; a++)
        {
            atomicMin(&bounds[a], mn[a]);
            atomicMax(&bounds[3 + a], mx[a]);
        }
}

__global__ void k_gb_init_bounds(int32_t* b)
{
    if (threadIdx.x < 3)
        b[threadIdx.x] = 2147483647;
    else if (threadIdx.x < 6)
        b[threadIdx.x] = -2147483647;
}

// absolute keys -> the builder's voxel key layout (leaf-major)
__global__ __launch_bounds__(256) void k_gb_encode(const uint64_t* __restrict__ keys, uint64_t n, KeyLayout kl,
                                                   uint64_t* __restrict__ out)
{
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    int32_t i, j, k;
    gb_unkey(keys[v], &i, &j, &k);
    out[v] = encode_voxel(kl, i, j, k);
}

__global__ __launch_bounds__(256) void k_gb_fill_i32(int32_t* __restrict__ p, uint64_t n, int32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// fanout_compute_fragment_idxs.slang semantics: slot s of window [w0, w0 + W) belongs to the prim
// whose inclusive footprint range contains w0 + s (the first prim with scan > w0 + s), at index
// w0 + s - (scan - count) of its footprint; slots at or past the batch total get ~0.
__global__ __launch_bounds__(256) void k_gb_fanout(const uint64_t* __restrict__ counts, const uint64_t* __restrict__ scan,
                                                   uint32_t prims, uint64_t w0, uint32_t W, uint32_t* __restrict__ prim_idx,
                                                   uint32_t* __restrict__ raster_idx)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= W)
        return;
    const uint64_t g = w0 + s;
    const uint64_t total = scan[prims - 1];
    uint32_t lo = 0, hi = prims; // first p in [0, prims) with scan[p] > g, prims when none
    while (lo < hi)
    {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (scan[mid] > g)
            hi = mid;
        else
            lo = mid + 1;
    }
    if (g < total && lo < prims)
    {
        prim_idx[s] = lo;
        raster_idx[s] = (uint32_t)(g - (scan[lo] - counts[lo]));
    }
    else
    {
        prim_idx[s] = prims - 1;
        raster_idx[s] = 0xFFFFFFFFu;
    }
}

static inline uint32_t gb_blocks(uint64_t n)
{
    return (uint32_t)std::max<uint64_t>(1, (n + 255) / 256);
}

// ------------------------------------------------------------------------------ state objects
struct GridBuildState
{
    uint64_t* keys = nullptr; // inserted voxels (absolute keys), deduplicated per window
    uint32_t* t = nullptr;    // their window
    uint64_t count = 0, cap = 0;
    uint32_t windows = 0;
    uint64_t rejected = 0; // points outside the root range
    int device = 0;
    ~GridBuildState()
    {
        (void)hipSetDevice(device);
        if (keys)
            (void)hipFree(keys);
        if (t)
            (void)hipFree(t);
    }
};

static Context* gb_ctx(pnanovdb_grid_build_context_t* c)
{
    return (Context*)c;
}

static GridBuildState* gb_state(pnanovdb_grid_build_state_t* s)
{
    return s ? (GridBuildState*)s->node_addresses_buffer : nullptr;
}

static void* gb_ptr(pnanovdb_compute_buffer_t* b)
{
    return b ? ((vdbr_buffer_t*)b)->device_ptr : nullptr;
}

static uint64_t gb_size(pnanovdb_compute_buffer_t* b)
{
    return b ? ((vdbr_buffer_t*)b)->size_bytes : 0;
}

// keeps the first `keep` entries while growing both arrays to hold `want`
static int gb_reserve(Context* ctx, GridBuildState* gs, uint64_t want)
{
    if (want <= gs->cap)
        return 0;
    const uint64_t cap = std::max<uint64_t>(want, gs->cap + gs->cap / 2 + 4096);
    void *k = nullptr, *t = nullptr;
    if (hipMalloc(&k, cap * 8) != hipSuccess || hipMalloc(&t, cap * 4) != hipSuccess)
    {
        (void)hipGetLastError();
        if (k)
            (void)hipFree(k);
        return ctx->fail("grid_build: device allocation of the point keys failed");
    }
    if (gs->count)
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(k, gs->keys, gs->count * 8, hipMemcpyDeviceToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(t, gs->t, gs->count * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (gs->keys)
        (void)hipFree(gs->keys);
    if (gs->t)
        (void)hipFree(gs->t);
    gs->keys = (uint64_t*)k;
    gs->t = (uint32_t*)t;
    gs->cap = cap;
    return 0;
}

// unique runs of sorted keys[n] appended to (out_keys, out_t) at `base`; returns the run count
static int gb_unique_append(Context* ctx, const uint64_t* keys, const uint32_t* t_in, uint32_t t_const, uint64_t n,
                            uint64_t* out_keys, uint32_t* out_t, uint64_t* runs)
{
    if (n == 0)
    {
        *runs = 0;
        return 0;
    }
    if (n >= 0xFFFFFFFFull)
        return ctx->fail("grid_build: more than 2^32-1 points in one call");
    uint32_t* flag = (uint32_t*)ctx->ws(WS_GB_FLAGS, (2 * n + 1) * sizeof(uint32_t));
    if (!flag)
        return -1;
    uint32_t* excl = flag + n;
    hipLaunchKernelGGL(k_gb_heads, dim3(gb_blocks(n)), dim3(256), 0, ctx->stream, keys, n, flag);
    if (scan_u32(ctx, flag, excl, n, false, excl + n))
        return -1;
    hipLaunchKernelGGL(k_gb_unique, dim3(gb_blocks(n)), dim3(256), 0, ctx->stream, keys, t_in, t_const,
                       (const uint32_t*)flag, (const uint32_t*)excl, n, out_keys, out_t);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, excl + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *runs = *(const uint32_t*)ctx->pinned;
    return 0;
}

// one window of points -> appended to the state
static int gb_insert(Context* ctx, GridBuildState* gs, const int32_t* ijk, uint64_t n)
{
    if (n > 0)
    {
        uint64_t* tmp = (uint64_t*)ctx->ws(WS_GB_KEYS, n * sizeof(uint64_t));
        unsigned long long* counters = (unsigned long long*)ctx->ws(WS_MISC, 4096);
        if (!tmp || !counters)
            return -1;
        VDBR_CHECK(ctx, hipMemsetAsync(counters, 0, 16, ctx->stream));
        const uint32_t blocks = (uint32_t)std::min<uint64_t>(gb_blocks(n), 8192u);
        hipLaunchKernelGGL(k_gb_collect, dim3(blocks), dim3(256), 0, ctx->stream, ijk, n, tmp, counters);
        VDBR_CHECK(ctx, hipGetLastError());
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, counters, 16, hipMemcpyDeviceToHost, ctx->stream));
        VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        const uint64_t c = ctx->pinned[0];
        gs->rejected += ctx->pinned[1];
        if (c > 0)
        {
            if (radix_sort_u64(ctx, tmp, nullptr, c, 0, GB_KEY_BITS, nullptr))
                return -1;
            // at most c new entries
            if (gb_reserve(ctx, gs, gs->count + c))
                return -1;
            uint64_t runs = 0;
            if (gb_unique_append(ctx, tmp, nullptr, gs->windows, c, gs->keys + gs->count, gs->t + gs->count, &runs))
                return -1;
            gs->count += runs;
        }
    }
    gs->windows++;
    return 0;
}

static int gb_write_out(Context* ctx, const uint8_t* grid, uint64_t size, pnanovdb_compute_buffer_t* out,
                        uint64_t word_count)
{
    void* dst = gb_ptr(out);
    if (!dst)
        return ctx->fail("grid_build: null output buffer");
    uint64_t cap = std::min<uint64_t>(word_count * 4u, gb_size(out));
    if (size > cap)
        return ctx->fail("grid_build: the grid needs " + std::to_string(size) + " bytes, the output buffer holds " +
                         std::to_string(cap));
    VDBR_CHECK(ctx, hipMemcpyAsync(dst, grid, size, hipMemcpyDeviceToDevice, ctx->stream));
    if (cap > size)
        VDBR_CHECK(ctx, hipMemsetAsync((uint8_t*)dst + size, 0, cap - size, ctx->stream));
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// the grid of every point since the reset (none: the empty grid)
static int gb_finalize(Context* ctx, GridBuildState* gs, float vs, pnanovdb_compute_buffer_t* out, uint64_t word_count)
{
    if (!(vs > 0.f) || !(vs < 3.0e38f))
        return ctx->fail("grid_build: voxel_size must be a positive finite number");
    if (gs->rejected)
        return ctx->fail("grid_build: " + std::to_string(gs->rejected) +
                         " points lie outside the node2 root range [-65536, 65535]");
    const uint64_t n = gs->count;
    int32_t* bounds = (int32_t*)ctx->ws(WS_MISC, 4096);
    if (!bounds)
        return -1;
    hipLaunchKernelGGL(k_gb_init_bounds, dim3(1), dim3(64), 0, ctx->stream, bounds);
    KeyLayout kl;
    memset(&kl, 0, sizeof(kl));
    uint64_t nvox = 0;
    uint64_t* vox_key = nullptr;
    uint32_t* vox_t = nullptr;
    if (n > 0)
    {
        hipLaunchKernelGGL(k_gb_bounds, dim3((uint32_t)std::min<uint64_t>(gb_blocks(n), 1024u)), dim3(256), 0,
                           ctx->stream, (const uint64_t*)gs->keys, n, bounds);
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, bounds, 24, hipMemcpyDeviceToHost, ctx->stream));
        VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        int32_t hb[6];
        memcpy(hb, ctx->pinned, 24);
        for (int a = 0; a < 3; a++)
        {
            kl.lmin[a] = hb[a] >> 3;
            const uint64_t range = (uint64_t)((hb[3 + a] >> 3) - kl.lmin[a] + 1);
            uint32_t b = 0;
            while ((1ull << b) < range)
                b++;
            kl.bits[a] = b;
        }
        kl.voxel_bits = kl.bits[0] + kl.bits[1] + kl.bits[2] + 9u;
        uint64_t* enc = (uint64_t*)ctx->ws(WS_GB_ENC, n * sizeof(uint64_t));
        uint32_t* tt = (uint32_t*)ctx->ws(WS_GB_T, n * sizeof(uint32_t));
        if (!enc || !tt)
            return -1;
        hipLaunchKernelGGL(k_gb_encode, dim3(gb_blocks(n)), dim3(256), 0, ctx->stream, (const uint64_t*)gs->keys, n,
                           kl, enc);
        VDBR_CHECK(ctx, hipMemcpyAsync(tt, gs->t, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
        // stable: entries of one voxel keep their insertion (window) order, the first is the earliest
        if (radix_sort_u64(ctx, enc, tt, n, 0, kl.voxel_bits, nullptr))
            return -1;
        vox_key = (uint64_t*)ctx->ws(WS_VOX_KEY, n * sizeof(uint64_t));
        vox_t = (uint32_t*)ctx->ws(WS_VOX_T, n * sizeof(uint32_t));
        if (!vox_key || !vox_t)
            return -1;
        if (gb_unique_append(ctx, enc, tt, 0u, n, vox_key, vox_t, &nvox))
            return -1;
    }
    else
    {
        vox_key = (uint64_t*)ctx->ws(WS_VOX_KEY, 8);
        vox_t = (uint32_t*)ctx->ws(WS_VOX_T, 8);
        if (!vox_key || !vox_t)
            return -1;
    }
    uint32_t* vox_val = (uint32_t*)ctx->ws(WS_VOX_VAL, std::max<uint64_t>(nvox, 1) * sizeof(uint32_t));
    if (!vox_val)
        return -1;
    // the reference leaves the values to its colour pass: zero here
    VDBR_CHECK(ctx, hipMemsetAsync(vox_val, 0, std::max<uint64_t>(nvox, 1) * sizeof(uint32_t), ctx->stream));
    Node2BuildInput in;
    in.kl = kl;
    in.vox_key = vox_key;
    in.vox_val = vox_val;
    in.vox_t = vox_t;
    in.nvox = nvox;
    in.vbounds = bounds;
    in.T = std::max<uint32_t>(gs->windows, 1u);
    in.max_nodes = 0xFFFFFFFFu;
    in.spatial_order = false;
    in.voxel_coords = false;
    in.value_type = VDB_GRID_TYPE_RGBA8;
    in.voxel_size = vs;
    Node2BuildOutput bo;
    if (build_node2_grid(ctx, in, &bo))
        return -1;
    return gb_write_out(ctx, bo.grid, bo.grid_size, out, word_count);
}

} // namespace vdbr

using namespace vdbr;

// ================================================================================= table entries
static pnanovdb_grid_build_context_t* PNANOVDB_ABI gb_create_context(const pnanovdb_compute_t* compute,
                                                                     pnanovdb_compute_queue_t* queue)
{
    vdbr_context_t* c = vdbr_context_create(0);
    if (!c)
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build: no HIP device");
    return (pnanovdb_grid_build_context_t*)c;
}

static void PNANOVDB_ABI gb_destroy_context(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                            pnanovdb_grid_build_context_t* context)
{
    (void)compute;
    (void)queue;
    vdbr_context_destroy((vdbr_context_t*)context);
}

static void PNANOVDB_ABI gb_init(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                 pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                 pnanovdb_compute_buffer_t* nanovdb_out, pnanovdb_uint64_t nanovdb_word_count,
                                 float voxel_size, pnanovdb_uint32_t dispatch_count)
{
    (void)nanovdb_out;
    (void)nanovdb_word_count;
    (void)voxel_size;
    (void)dispatch_count;
    if (!state || !context)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build_init: null state or context");
        return;
    }
    memset(state, 0, sizeof(*state));
    GridBuildState* gs = new (std::nothrow) GridBuildState();
    if (gs)
        gs->device = gb_ctx(context)->device;
    state->node_addresses_buffer = (pnanovdb_compute_buffer_t*)gs;
}

static void PNANOVDB_ABI gb_reset(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                  pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                  pnanovdb_compute_buffer_t* nanovdb_out, pnanovdb_uint64_t nanovdb_word_count,
                                  float voxel_size, pnanovdb_uint32_t dispatch_count)
{
    (void)dispatch_count;
    GridBuildState* gs = gb_state(state);
    Context* ctx = gb_ctx(context);
    if (!gs || !ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build_reset: state not initialised");
        return;
    }
    gs->count = 0;
    gs->windows = 0;
    gs->rejected = 0;
    if (gb_finalize(ctx, gs, voxel_size, nanovdb_out, nanovdb_word_count))
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build_reset: %s", ctx->error.c_str());
}

static void PNANOVDB_ABI gb_destroy(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                    pnanovdb_grid_build_state_t* state)
{
    (void)compute;
    (void)queue;
    if (!state)
        return;
    delete gb_state(state);
    memset(state, 0, sizeof(*state));
}

static void PNANOVDB_ABI gb_build(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                  pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                  pnanovdb_compute_buffer_t* point_ijk_in, pnanovdb_compute_buffer_t* nanovdb_out,
                                  pnanovdb_uint64_t point_count, pnanovdb_uint64_t nanovdb_word_count, float voxel_size,
                                  pnanovdb_uint32_t dispatch_count)
{
    (void)nanovdb_out;
    (void)nanovdb_word_count;
    (void)voxel_size;
    (void)dispatch_count;
    GridBuildState* gs = gb_state(state);
    Context* ctx = gb_ctx(context);
    if (!gs || !ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build: state not initialised");
        return;
    }
    const int32_t* ijk = (const int32_t*)gb_ptr(point_ijk_in);
    if (point_count > 0 && (!ijk || gb_size(point_ijk_in) < point_count * 12u))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build: the point buffer holds fewer than %llu points",
                    (unsigned long long)point_count);
        return;
    }
    if (gb_insert(ctx, gs, ijk, point_count))
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build: %s", ctx->error.c_str());
}

static void PNANOVDB_ABI gb_finalize_entry(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                           pnanovdb_compute_buffer_t* point_ijk_in,
                                           pnanovdb_compute_buffer_t* nanovdb_out, pnanovdb_uint64_t point_count,
                                           pnanovdb_uint64_t nanovdb_word_count, float voxel_size,
                                           pnanovdb_uint32_t dispatch_count)
{
    (void)point_ijk_in;
    (void)point_count;
    (void)dispatch_count;
    GridBuildState* gs = gb_state(state);
    Context* ctx = gb_ctx(context);
    if (!gs || !ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build_finalize: state not initialised");
        return;
    }
    if (gb_finalize(ctx, gs, voxel_size, nanovdb_out, nanovdb_word_count))
        log_message(compute, queue, LOG_LEVEL_ERROR, "grid_build_finalize: %s", ctx->error.c_str());
}

// ------------------------------------------------------------------------------------ fanout
static vdbr_buffer_t* gb_buffer(uint64_t bytes)
{
    vdbr_buffer_t* b = new (std::nothrow) vdbr_buffer_t();
    if (!b)
        return nullptr;
    b->size_bytes = bytes;
    b->device_ptr = nullptr;
    if (hipMalloc(&b->device_ptr, bytes ? bytes : 8) != hipSuccess)
    {
        (void)hipGetLastError();
        delete b;
        return nullptr;
    }
    return b;
}

static void gb_free_buffer(pnanovdb_compute_buffer_t*& b)
{
    vdbr_buffer_t* v = (vdbr_buffer_t*)b;
    if (v)
    {
        (void)hipFree(v->device_ptr);
        delete v;
    }
    b = nullptr;
}

// GridBuild.cpp:656-660
static void PNANOVDB_ABI fo_reset(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                  pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_fanout_state_t* state)
{
    (void)compute;
    (void)queue;
    (void)context;
    if (!state)
        return;
    state->prim_batch_idx = 0u;
    state->ijk_batch_idx = 0u;
    state->total_fragment_count = 0u;
}

// GridBuild.cpp:663-669
static pnanovdb_bool_t PNANOVDB_ABI fo_valid(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                             pnanovdb_grid_build_context_t* context,
                                             pnanovdb_grid_build_fanout_state_t* state)
{
    (void)compute;
    (void)queue;
    (void)context;
    return state && state->prim_batch_idx < state->prim_batch_count ? PNANOVDB_TRUE : PNANOVDB_FALSE;
}

// GridBuild.cpp:671-683: the next window, or the first window of the next prim batch
static void PNANOVDB_ABI fo_increment(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                      pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_fanout_state_t* state)
{
    (void)compute;
    (void)queue;
    (void)context;
    if (!state)
        return;
    state->ijk_batch_idx++;
    if (state->ijk_batch_idx >= state->ijk_batch_count)
    {
        state->prim_batch_idx++;
        state->ijk_batch_idx = 0u;
        state->ijk_batch_count = state->ijk_batch_count_max;
    }
}

static void PNANOVDB_ABI fo_destroy(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                    pnanovdb_grid_build_fanout_state_t* state)
{
    (void)compute;
    (void)queue;
    if (!state)
        return;
    gb_free_buffer(state->fragment_counts_buffer);
    gb_free_buffer(state->scan_fragment_counts_buffer);
    gb_free_buffer(state->prim_bbox_buffer);
    gb_free_buffer(state->scatter_prim_idxs_buffer);
    gb_free_buffer(state->prim_idxs_buffer);
    gb_free_buffer(state->prim_raster_idxs_buffer);
    gb_free_buffer(state->prim_raster_ijks_buffer);
    vdbr_buffer_t* rb = (vdbr_buffer_t*)state->readback_buffer;
    if (rb)
    {
        delete[](uint64_t*) rb->device_ptr;
        delete rb;
    }
    state->readback_buffer = nullptr;
    delete[] state->readback_frames;
    state->readback_frames = nullptr;
    state->readback_frame_count = 0u;
}

// GridBuild.cpp:685-752: parameters, buffers of the reference's sizes
static void PNANOVDB_ABI fo_init(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                 pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_fanout_state_t* state,
                                 pnanovdb_uint32_t prim_batch_size, pnanovdb_uint32_t ijk_batch_size,
                                 pnanovdb_uint32_t ijk_batch_count_max, pnanovdb_uint32_t prim_count, float voxel_size)
{
    Context* ctx = gb_ctx(context);
    if (!state || !ctx || prim_batch_size == 0 || ijk_batch_size == 0)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "fanout_state_init: null state or context, or a zero batch size");
        return;
    }
    memset(state, 0, sizeof(*state));
    state->prim_batch_size = prim_batch_size;
    state->ijk_batch_size = ijk_batch_size;
    state->prim_count = prim_count;
    state->voxel_size = voxel_size;
    state->prim_batch_count = (uint32_t)(((uint64_t)prim_count + prim_batch_size - 1u) / prim_batch_size);
    state->ijk_batch_count = ijk_batch_count_max;
    state->ijk_batch_count_max = ijk_batch_count_max;
    fo_reset(compute, queue, context, state);
    (void)hipSetDevice(ctx->device);
    const uint64_t PB = prim_batch_size, W = ijk_batch_size;
    state->fragment_counts_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(PB * 8u);
    state->scan_fragment_counts_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(PB * 8u);
    state->prim_bbox_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(PB * 24u);
    state->scatter_prim_idxs_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(W * 4u);
    state->prim_idxs_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(W * 4u);
    state->prim_raster_idxs_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(W * 4u);
    state->prim_raster_ijks_buffer = (pnanovdb_compute_buffer_t*)gb_buffer(W * 12u);
    state->readback_frame_count = state->prim_batch_count;
    state->readback_frames = new (std::nothrow) uint64_t[std::max<uint32_t>(state->prim_batch_count, 1u)];
    vdbr_buffer_t* rb = new (std::nothrow) vdbr_buffer_t();
    if (rb)
    {
        rb->size_bytes = (uint64_t)std::max<uint32_t>(state->prim_batch_count, 1u) * 8u;
        rb->device_ptr = new (std::nothrow) uint64_t[std::max<uint32_t>(state->prim_batch_count, 1u)];
    }
    state->readback_buffer = (pnanovdb_compute_buffer_t*)rb;
    if (!state->fragment_counts_buffer || !state->scan_fragment_counts_buffer || !state->prim_bbox_buffer ||
        !state->scatter_prim_idxs_buffer || !state->prim_idxs_buffer || !state->prim_raster_idxs_buffer ||
        !state->prim_raster_ijks_buffer || !state->readback_frames || !rb || !rb->device_ptr)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "fanout_state_init: allocation failed");
        fo_destroy(compute, queue, state);
        return;
    }
    for (uint64_t i = 0; i < state->readback_frame_count; i++)
        state->readback_frames[i] = ~0ull;
    // footprints start at 0; unwritten voxel slots are skipped by grid_build
    (void)hipMemsetAsync(gb_ptr(state->fragment_counts_buffer), 0, PB * 8u, ctx->stream);
    (void)hipMemsetAsync(gb_ptr(state->scatter_prim_idxs_buffer), 0, W * 4u, ctx->stream);
    hipLaunchKernelGGL(k_gb_fill_i32, dim3((uint32_t)std::min<uint64_t>(gb_blocks(3 * W), 4096u)), dim3(256), 0,
                       ctx->stream, (int32_t*)gb_ptr(state->prim_raster_ijks_buffer), 3 * W, GB_SKIP);
    (void)hipStreamSynchronize(ctx->stream);
}

// GridBuild.cpp:840-941: scan of the batch's footprints, the batch total on its first window, the
// slot -> (prim, footprint index) map of the current window
static void PNANOVDB_ABI fo_fanout(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                   pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_fanout_state_t* state,
                                   pnanovdb_uint32_t dispatch_count)
{
    (void)dispatch_count;
    Context* ctx = gb_ctx(context);
    if (!state || !ctx || !state->fragment_counts_buffer || state->prim_batch_idx >= state->prim_batch_count)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "fanout: state not initialised or past its last batch");
        return;
    }
    const uint32_t PB = state->prim_batch_size;
    const uint64_t* counts = (const uint64_t*)gb_ptr(state->fragment_counts_buffer);
    uint64_t* scan = (uint64_t*)gb_ptr(state->scan_fragment_counts_buffer);
    if (scan_u64(ctx, counts, scan, PB, true, nullptr))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "fanout: %s", ctx->error.c_str());
        return;
    }
    uint64_t* totals = (uint64_t*)((vdbr_buffer_t*)state->readback_buffer)->device_ptr;
    const uint32_t b = state->prim_batch_idx;
    if (state->readback_frames[b] == ~0ull)
    {
        if (hipMemcpyAsync(&totals[b], scan + PB - 1, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
        {
            (void)hipGetLastError();
            log_message(compute, queue, LOG_LEVEL_ERROR, "fanout: readback of the batch total failed");
            return;
        }
        state->readback_frames[b] = 0u;
    }
    if (state->ijk_batch_count == state->ijk_batch_count_max)
    {
        state->total_fragment_count += totals[b];
        state->ijk_batch_count = (uint32_t)((totals[b] + state->ijk_batch_size - 1u) / state->ijk_batch_size);
    }
    const uint64_t w0 = (uint64_t)state->ijk_batch_idx * state->ijk_batch_size;
    hipLaunchKernelGGL(k_gb_fanout, dim3(gb_blocks(state->ijk_batch_size)), dim3(256), 0, ctx->stream, counts,
                       (const uint64_t*)scan, PB, w0, state->ijk_batch_size,
                       (uint32_t*)gb_ptr(state->prim_idxs_buffer), (uint32_t*)gb_ptr(state->prim_raster_idxs_buffer));
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    {
        (void)hipGetLastError();
        log_message(compute, queue, LOG_LEVEL_ERROR, "fanout: kernel failed");
    }
}

static pnanovdb_grid_build_t g_grid_build;
static bool g_grid_build_init = false;

PNANOVDB_API pnanovdb_grid_build_t* pnanovdb_get_grid_build(void)
{
    if (!g_grid_build_init)
    {
        memset(&g_grid_build, 0, sizeof(g_grid_build));
        g_grid_build.create_context = gb_create_context;
        g_grid_build.destroy_context = gb_destroy_context;
        g_grid_build.grid_build_init = gb_init;
        g_grid_build.grid_build_reset = gb_reset;
        g_grid_build.grid_build_destroy = gb_destroy;
        g_grid_build.grid_build = gb_build;
        g_grid_build.grid_build_finalize = gb_finalize_entry;
        g_grid_build.fanout_state_reset = fo_reset;
        g_grid_build.fanout_state_valid = fo_valid;
        g_grid_build.fanout_state_increment = fo_increment;
        g_grid_build.fanout_state_init = fo_init;
        g_grid_build.fanout_state_destroy = fo_destroy;
        g_grid_build.fanout = fo_fanout;
        g_grid_build.compute = compute_table();
        g_grid_build_init = true;
    }
    return &g_grid_build;
}
