// vdb_prims.hip -- device-wide scans and the onesweep LSD radix sort (gfx950, wave64).
//
// Scans (ParallelPrimitives.cpp:151-310 semantics: inclusive, u32 add / u64 add / u32 max):
// reduce-then-scan over 2048-element tiles (tile reduce, one-block scan of tile partials,
// tile scan + offset).
//
// Radix sort (ParallelPrimitives.cpp:685-869 semantics: stable, ascending, LSD over a bit
// range): one histogram pass over the keys builds the digit counts of every 8-bit pass; each
// pass is then a single onesweep kernel over 8192-key tiles (512 threads x 16 keys):
//   - wave-local stable ranking with ballot peer masks (8 ballots per digit bit set),
//   - per-tile digit counts published to a status array and resolved by decoupled look-back
//     that inspects four predecessor tiles per step
//     (epoch-tagged 64-bit status words, so the array is never cleared between passes),
//   - keys staged in LDS in tile-sorted order and written out in digit runs.
// Tiles take their index from an atomic counter, so every tile a block waits on has already
// been claimed by a running block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <stdint.h>

#include "vdb_internal.h"

namespace vdbr
{

// ============================================================================ fills and copies
// Kernels instead of hipMemsetAsync / hipMemcpyAsync where the length may be a device count, and
// in work that a caller may capture into a graph.
template <typename T>
__global__ void k_fill(T* __restrict__ p, DCount cn, T v)
{
    const uint64_t n = cn.get();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void k_copy(const T* __restrict__ src, T* __restrict__ dst, DCount cn)
{
    const uint64_t n = cn.get();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static uint32_t stride_blocks(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (uint32_t)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

template <typename T>
static int fill_impl(Context* ctx, T* p, DCount n, T v)
{
    if (!n.dev && n.cap == 0)
        return 0;
    hipLaunchKernelGGL((k_fill<T>), dim3(stride_blocks(n.cap)), dim3(256), 0, ctx->stream, p, n, v);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

int fill_words(Context* ctx, uint32_t* p, DCount n, uint32_t v)
{
    return fill_impl<uint32_t>(ctx, p, n, v);
}
int fill_words(Context* ctx, uint64_t* p, DCount n, uint64_t v)
{
    return fill_impl<uint64_t>(ctx, p, n, v);
}

template <typename T>
static int copy_impl(Context* ctx, const T* src, T* dst, DCount n)
{
    if (!n.dev && n.cap == 0)
        return 0;
    hipLaunchKernelGGL((k_copy<T>), dim3(stride_blocks(n.cap)), dim3(256), 0, ctx->stream, src, dst, n);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

int copy_words(Context* ctx, const uint32_t* src, uint32_t* dst, DCount n)
{
    return copy_impl<uint32_t>(ctx, src, dst, n);
}
int copy_words(Context* ctx, const uint64_t* src, uint64_t* dst, DCount n)
{
    return copy_impl<uint64_t>(ctx, src, dst, n);
}

// ============================================================================ scans
template <typename T, int OP>
struct ScanOp;
template <>
struct ScanOp<uint32_t, 0>
{
    static __device__ __forceinline__ uint32_t apply(uint32_t a, uint32_t b)
    {
        return a + b;
    }
    static __device__ __forceinline__ uint32_t identity()
    {
        return 0u;
    }
};
template <>
struct ScanOp<uint64_t, 1>
{
    static __device__ __forceinline__ uint64_t apply(uint64_t a, uint64_t b)
    {
        return a + b;
    }
    static __device__ __forceinline__ uint64_t identity()
    {
        return 0ull;
    }
};
template <>
struct ScanOp<uint32_t, 2>
{
    static __device__ __forceinline__ uint32_t apply(uint32_t a, uint32_t b)
    {
        return a > b ? a : b;
    }
    static __device__ __forceinline__ uint32_t identity()
    {
        return 0u;
    }
};

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// inclusive scan of one value per thread across a 256-thread block; returns the inclusive
// prefix, *total receives the block aggregate
template <typename T, int OP>
__device__ __forceinline__ T block_scan_inclusive(T v, T* smem_waves, T* total)
{
    using Op = ScanOp<T, OP>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        T o = __shfl_up(v, off, 64);
        if (lane >= off)
            v = Op::apply(o, v);
    }
    if (lane == 63)
        smem_waves[wave] = v;
    __syncthreads();
    T prefix = Op::identity();
    for (int w = 0; w < wave; w++)
        prefix = Op::apply(prefix, smem_waves[w]);
    T tot = Op::identity();
    for (int w = 0; w < (SCAN_THREADS >> 6); w++)
        tot = Op::apply(tot, smem_waves[w]);
    *total = tot;
    __syncthreads();
    return Op::apply(prefix, v);
}

template <typename T, int OP>
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_reduce(const T* __restrict__ in, DCount cn, T* __restrict__ partials)
{
    using Op = ScanOp<T, OP>;
    __shared__ T smem[SCAN_THREADS / 64];
    const uint64_t n = cn.get();
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    if (base >= n)
        return; // block-uniform: the grid may be sized by a capacity
    T acc = Op::identity();
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++)
    {
        uint64_t idx = base + (uint64_t)m * SCAN_THREADS + threadIdx.x;
        if (idx < n)
            acc = Op::apply(acc, in[idx]);
    }
    T total;
    block_scan_inclusive<T, OP>(acc, smem, &total);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = total;
}

// exclusive scan of the tile partials in place, one block
template <typename T, int OP>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(T* __restrict__ partials, DCount cn, T* __restrict__ grand_total)
{
    using Op = ScanOp<T, OP>;
    __shared__ T smem[SCAN_THREADS / 64];
    const uint64_t count = (cn.get() + SCAN_TILE - 1) / SCAN_TILE; // tiles
    T carry = Op::identity();
    for (uint64_t base = 0; base < count; base += SCAN_THREADS)
    {
        uint64_t idx = base + threadIdx.x;
        T v = idx < count ? partials[idx] : Op::identity();
        T total;
        T incl = block_scan_inclusive<T, OP>(v, smem, &total);
        T excl_local = Op::identity();
        // exclusive value: inclusive of the previous lane (shift through shared memory)
        __shared__ T sh[SCAN_THREADS];
        sh[threadIdx.x] = incl;
        __syncthreads();
        excl_local = threadIdx.x == 0 ? Op::identity() : sh[threadIdx.x - 1];
        __syncthreads();
        if (idx < count)
            partials[idx] = Op::apply(carry, excl_local);
        carry = Op::apply(carry, total);
        __syncthreads();
    }
    if (threadIdx.x == 0 && grand_total)
        *grand_total = carry;
}

template <typename T, int OP>
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_scan(const T* __restrict__ in, T* __restrict__ out, DCount cn,
                                                         const T* __restrict__ partials, int inclusive)
{
    using Op = ScanOp<T, OP>;
    __shared__ T tile[SCAN_TILE];
    __shared__ T smem[SCAN_THREADS / 64];
    const uint64_t n = cn.get();
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    if (base >= n)
        return; // block-uniform
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++)
    {
        uint64_t idx = base + (uint64_t)m * SCAN_THREADS + threadIdx.x;
        tile[m * SCAN_THREADS + threadIdx.x] = idx < n ? in[idx] : Op::identity();
    }
    __syncthreads();
    T items[SCAN_ITEMS];
    T acc = Op::identity();
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++)
    {
        items[m] = tile[threadIdx.x * SCAN_ITEMS + m];
        acc = Op::apply(acc, items[m]);
    }
    T total;
    T incl = block_scan_inclusive<T, OP>(acc, smem, &total);
    // exclusive prefix of this thread = inclusive prefix of the previous thread
    __shared__ T sh[SCAN_THREADS];
    sh[threadIdx.x] = incl;
    __syncthreads();
    T prefix = Op::apply(partials[blockIdx.x], threadIdx.x == 0 ? Op::identity() : sh[threadIdx.x - 1]);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++)
    {
        T before = prefix;
        prefix = Op::apply(prefix, items[m]);
        tile[threadIdx.x * SCAN_ITEMS + m] = inclusive ? prefix : before;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++)
    {
        uint64_t idx = base + (uint64_t)m * SCAN_THREADS + threadIdx.x;
        if (idx < n)
            out[idx] = tile[m * SCAN_THREADS + threadIdx.x];
    }
}

// A count held on the device (cn.dev) sizes the launches by its bound and always runs the three
// kernels (the partials pass writes the total, 0 for an empty input); a host count of 0 only
// clears the total.
template <typename T, int OP>
int scan_impl(Context* ctx, const T* in, T* out, DCount cn, bool inclusive, T* total_dev)
{
    if (!cn.dev && cn.cap == 0)
    {
        if (total_dev)
            return fill_words(ctx, total_dev, host_count(1), 0);
        return 0;
    }
    uint64_t tiles = (cn.cap + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 0)
        tiles = 1;
    if (tiles > 0xFFFFFFFFull)
        return ctx->fail("scan: too many tiles");
    T* partials = (T*)ctx->scratch_scan(tiles * sizeof(T));
    if (!partials)
        return -1;
    hipLaunchKernelGGL((k_tile_reduce<T, OP>), dim3((uint32_t)tiles), dim3(SCAN_THREADS), 0, ctx->stream, in, cn, partials);
    hipLaunchKernelGGL((k_scan_partials<T, OP>), dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, partials, cn, total_dev);
    hipLaunchKernelGGL((k_tile_scan<T, OP>), dim3((uint32_t)tiles), dim3(SCAN_THREADS), 0, ctx->stream, in, out, cn,
                       (const T*)partials, inclusive ? 1 : 0);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

int scan_u32(Context* ctx, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive, uint32_t* total_dev)
{
    return scan_impl<uint32_t, 0>(ctx, in, out, host_count(n), inclusive, total_dev);
}
int scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, uint64_t n, bool inclusive, uint64_t* total_dev)
{
    return scan_impl<uint64_t, 1>(ctx, in, out, host_count(n), inclusive, total_dev);
}
int scan_u32(Context* ctx, const uint32_t* in, uint32_t* out, DCount n, bool inclusive, uint32_t* total_dev)
{
    return scan_impl<uint32_t, 0>(ctx, in, out, n, inclusive, total_dev);
}
int scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, DCount n, bool inclusive, uint64_t* total_dev)
{
    return scan_impl<uint64_t, 1>(ctx, in, out, n, inclusive, total_dev);
}
int scan_max_u32(Context* ctx, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive)
{
    return scan_impl<uint32_t, 2>(ctx, in, out, host_count(n), inclusive, nullptr);
}

// ============================================================================ radix sort
// onesweep tile shapes (threads per block x keys per thread); SORT_VARIANTS[0] is the default
struct SortShape
{
    int threads, items;
    bool with_vals; // the shape also serves sorts with a 32-bit payload (LDS room for it)
};
// the last two shapes stage 16384 / 12288 keys (128 / 96 KB of LDS, one block per CU): keys-only sorts
constexpr SortShape SORT_VARIANTS[] = { { 512, 16, true },  { 256, 16, true },   { 1024, 8, true }, { 256, 24, true },
                                        { 512, 12, true },  { 1024, 16, false }, { 768, 16, false } };
constexpr uint32_t SORT_VARIANT_COUNT = sizeof(SORT_VARIANTS) / sizeof(SORT_VARIANTS[0]);
constexpr int SORT_MAX_PASSES = 11; // 64 bits in digits of at least 6 bits

// status word: [63:48] epoch, [47:46] flag (1 = tile aggregate, 2 = inclusive prefix), [45:0] count
#define ST_AGG 1ull
#define ST_INCL 2ull
__device__ __forceinline__ uint64_t st_pack(uint32_t epoch, uint64_t flag, uint64_t value)
{
    return ((uint64_t)epoch << 48) | (flag << 46) | (value & ((1ull << 46) - 1ull));
}

// digit of pass p covers key bits [bit_begin + W p, min(bit_end, bit_begin + W p + W)) for the
// digit width W <= SORT_MAX_DIGIT; the same mask is applied here and in k_sort_pass so the per-pass counts
// always agree.
// Each thread counts 16 consecutive keys and adds runs of equal digits at once: the leading
// digits of neighbouring pair keys are mostly equal, so per-key LDS atomics would serialise.
constexpr int HIST_RUN = 16;
__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t* __restrict__ keys, DCount cn, uint32_t bit_begin,
                                                   uint32_t bit_count, uint32_t W, unsigned long long* __restrict__ hist)
{
    const uint64_t n = cn.get();
    const uint32_t passes = (bit_count + W - 1u) / W;
    __shared__ uint32_t h[SORT_MAX_PASSES][SORT_BINS];
    for (uint32_t x = threadIdx.x; x < SORT_MAX_PASSES * SORT_BINS; x += 256u)
        (&h[0][0])[x] = 0u;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * HIST_RUN;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * HIST_RUN; c < n; c += stride)
    {
        uint64_t k[HIST_RUN];
        uint32_t cnt = (n - c) < (uint64_t)HIST_RUN ? (uint32_t)(n - c) : (uint32_t)HIST_RUN;
        if (cnt == HIST_RUN)
        {
            const ulonglong2* src = (const ulonglong2*)(keys + c);
#pragma unroll
            for (int r = 0; r < HIST_RUN / 2; r++)
            {
                ulonglong2 v = src[r];
                k[2 * r] = v.x;
                k[2 * r + 1] = v.y;
            }
        }
        else
        {
#pragma unroll
            for (int r = 0; r < HIST_RUN; r++)
                k[r] = (uint32_t)r < cnt ? keys[c + r] : 0ull;
        }
        for (uint32_t p = 0; p < passes; p++)
        {
            const uint32_t shift = bit_begin + W * p;
            const uint32_t rem = bit_count - W * p;
            const uint32_t dm = rem >= W ? ((1u << W) - 1u) : ((1u << rem) - 1u);
            uint32_t cur = (uint32_t)(k[0] >> shift) & dm;
            uint32_t run = 1u;
#pragma unroll
            for (int r = 1; r < HIST_RUN; r++)
            {
                if ((uint32_t)r >= cnt)
                    break;
                uint32_t d = (uint32_t)(k[r] >> shift) & dm;
                if (d != cur)
                {
                    atomicAdd(&h[p][cur], run);
                    cur = d;
                    run = 0u;
                }
                run++;
            }
            atomicAdd(&h[p][cur], run);
        }
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < passes * SORT_BINS; x += 256u)
    {
        uint32_t c = (&h[0][0])[x];
        if (c)
            atomicAdd(&hist[x], (unsigned long long)c);
    }
}

// exclusive scan of each pass's SORT_BINS digit counts (one block, a thread per digit)
__global__ __launch_bounds__(SORT_BINS) void k_sort_digit_starts(const unsigned long long* __restrict__ hist,
                                                                 uint32_t passes, uint64_t* __restrict__ starts)
{
    __shared__ uint64_t wsum[SORT_BINS / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t p = 0; p < passes; p++)
    {
        const uint64_t c = hist[p * SORT_BINS + threadIdx.x];
        uint64_t v = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            const uint64_t o = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off)
                v += o;
        }
        if (lane == 63u)
            wsum[wave] = v;
        __syncthreads();
        uint64_t pre = 0ull;
        for (uint32_t w = 0; w < wave; w++)
            pre += wsum[w];
        starts[p * SORT_BINS + threadIdx.x] = pre + v - c;
        __syncthreads();
    }
}

// keys only, tiles of up to 8192 keys: at most 128 VGPRs (4 waves per SIMD), so that two
// 512-thread tiles share a CU; the other shapes are held to one or two tiles per CU by their LDS
// STABLE = false: keys of one digit may leave the tile in any order (each lane takes its rank from
// an LDS atomic; no peer masks)
template <int SORT_THREADS, int SORT_ITEMS, bool HAS_VALS, bool STABLE = true>
__global__ __launch_bounds__(SORT_THREADS) __attribute__((
    amdgpu_waves_per_eu(!HAS_VALS && SORT_ITEMS <= 16 && SORT_THREADS * SORT_ITEMS <= 8192 ? 4 : 1))) void k_sort_pass(const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout,
                                                            const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                            DCount cn, uint32_t shift, uint32_t digit_bits,
                                                            const uint64_t* __restrict__ digit_start,
                                                            uint64_t* __restrict__ status, uint32_t* __restrict__ tile_counter,
                                                            uint32_t epoch)
{
    constexpr int WAVES = SORT_THREADS / 64;
    constexpr int SORT_WAVE_TILE = 64 * SORT_ITEMS;
    constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
    static_assert(WAVES >= 4, "digit bookkeeping: a thread per digit of up to 8 bits");
    __shared__ uint32_t s_tile;
    // per wave and digit: 16-bit counts (at most 64 x SORT_ITEMS), two digits per word, later the
    // wave's exclusive offset within the tile's digit
    __shared__ uint32_t wcnt[WAVES][SORT_BINS / 2];
    __shared__ uint32_t s_hist_excl[SORT_BINS];
    __shared__ uint64_t s_base[SORT_BINS];
    __shared__ uint32_t s_wsum[SORT_BINS / 64];
    static_assert(64 * SORT_ITEMS < 65536, "16-bit wave counts");
    __shared__ uint64_t s_keys[SORT_TILE];
    __shared__ uint32_t s_vals[HAS_VALS ? SORT_TILE : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t lane = tid & 63u;
    if (tid == 0)
        s_tile = atomicAdd(tile_counter, 1u);
    for (uint32_t x = tid; x < WAVES * (SORT_BINS / 2); x += SORT_THREADS)
        (&wcnt[0][0])[x] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t n = cn.get();
    if ((uint64_t)tile * SORT_TILE >= n)
        return; // block-uniform: a grid sized by a capacity has tiles past the keys
    const uint64_t base = (uint64_t)tile * SORT_TILE + (uint64_t)wave * SORT_WAVE_TILE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint32_t dmask = (1u << digit_bits) - 1u;

    uint64_t keys[SORT_ITEMS];
    uint32_t vals[SORT_ITEMS];
    uint32_t rank[SORT_ITEMS];
#pragma unroll
    for (int m = 0; m < SORT_ITEMS; m++)
    {
        uint64_t idx = base + (uint64_t)m * 64u + lane;
        bool valid = idx < n;
        keys[m] = valid ? kin[idx] : ~0ull;
        if (HAS_VALS)
            vals[m] = valid ? vin[idx] : 0u;
    }
    // ranks in groups of G keys: the lowest lane of each peer group (lanes holding the same digit)
    // adds the group's size to the wave's count of the digit and gets the count before it. The
    // LDS unit applies one wave's atomics in issue order, so a group's G additions are issued
    // back to back without waiting for each other's result; the leaders' old counts are then
    // broadcast to their peers.
    constexpr int G = 4;
    static_assert(SORT_ITEMS % G == 0, "whole groups");
    if (!STABLE)
    {
        // every lane adds itself to its digit's count; the returned counts are distinct
#pragma unroll
        for (int m = 0; m < SORT_ITEMS; m++)
        {
            const uint64_t idx = base + (uint64_t)m * 64u + lane;
            const uint32_t d = (uint32_t)(keys[m] >> shift) & dmask;
            const uint32_t hs = (d & 1u) << 4;
            rank[m] = idx < n ? (atomicAdd(&wcnt[wave][d >> 1], 1u << hs) >> hs) & 0xFFFFu : 0u;
        }
    }
#pragma unroll
    for (int m0 = 0; m0 < SORT_ITEMS && STABLE; m0 += G)
    {
        uint32_t lead[G]; // old count (at the leader, < 2^26) and the leader's lane
#pragma unroll
        for (int q = 0; q < G; q++)
        {
            const int m = m0 + q;
            uint64_t idx = base + (uint64_t)m * 64u + lane;
            bool valid = idx < n;
            uint32_t d = (uint32_t)(keys[m] >> shift) & dmask;
            // peers: one step per digit bit (a uniform exit ends the unrolled steps at the pass's
            // width); per step the ballot of the bit is combined as
            // peers &= ~(ballot ^ sign-extended bit): ballot where the bit is set, ~ballot where not
            const uint64_t vb = __ballot(valid);
            uint32_t plo = (uint32_t)vb, phi = (uint32_t)(vb >> 32);
#pragma unroll
            for (uint32_t b = 0; b < SORT_MAX_DIGIT; b++)
            {
                if (b >= digit_bits)
                    break;
                const int32_t sbit = (int32_t)(d << (31u - b)) >> 31; // 0 or -1
                const uint64_t bal = __ballot(sbit != 0);
                plo &= ~((uint32_t)bal ^ (uint32_t)sbit);
                phi &= ~((uint32_t)(bal >> 32) ^ (uint32_t)sbit);
            }
            const uint64_t peers = ((uint64_t)phi << 32) | plo;
            const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
            uint32_t old = 0u;
            if (valid && below == 0u)
            {
                const uint32_t hs = (d & 1u) << 4;
                old = (atomicAdd(&wcnt[wave][d >> 1], (uint32_t)__popcll(peers) << hs) >> hs) & 0xFFFFu;
            }
            rank[m] = below;
            lead[q] = (old << 6) | (valid ? (uint32_t)__builtin_ctzll(peers) : lane);
        }
#pragma unroll
        for (int q = 0; q < G; q++)
            rank[m0 + q] += (uint32_t)__shfl((int)(lead[q] >> 6), (int)(lead[q] & 63u), 64);
    }
    __syncthreads();

    // per digit (thread tid < 2^digit_bits handles digit tid): wave offsets, tile count, look-back
    const uint32_t nbins = 1u << digit_bits;
    uint16_t* woff = (uint16_t*)&wcnt[0][0];
    uint64_t excl = 0ull;
    uint32_t total = 0u;
    const uint32_t d = tid;
    if (tid < nbins)
    {
        uint32_t c[WAVES];
#pragma unroll
        for (int w = 0; w < WAVES; w++)
            c[w] = (wcnt[w][d >> 1] >> ((d & 1u) << 4)) & 0xFFFFu;
        // each thread rewrites only its own 16-bit half: the wave's exclusive offset
#pragma unroll
        for (int w = 0; w < WAVES; w++)
        {
            woff[w * SORT_BINS + d] = (uint16_t)total;
            total += c[w];
        }
        uint64_t* my_status = status + (uint64_t)tile * SORT_BINS + d;
        // publish the tile aggregate first (look-back chains never wait on a later tile)
        __hip_atomic_store(my_status, st_pack(epoch, tile == 0 ? ST_INCL : ST_AGG, total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // exclusive scan of the tile digit counts (threads 0 .. nbins - 1)
    if (tid < nbins)
    {
        uint32_t v = total;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            uint32_t o = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off)
                v += o;
        }
        if (lane == 63u)
            s_wsum[wave] = v;
    }
    __syncthreads();
    if (tid < nbins)
    {
        uint32_t pre = 0u;
        for (uint32_t w = 0; w < wave; w++)
            pre += s_wsum[w];
        uint32_t v = total;
        // recompute the in-wave inclusive prefix (cheap) to avoid keeping it live across the barrier
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            uint32_t o = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off)
                v += o;
        }
        s_hist_excl[d] = pre + v - total;
        // decoupled look-back, four predecessors per step
        if (tile > 0)
        {
            int64_t t = (int64_t)tile - 1;
            while (true)
            {
                uint64_t st4[4];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    st4[q] = (t - q >= 0) ? __hip_atomic_load(status + (uint64_t)(t - q) * SORT_BINS + d,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                          : 0ull;
                int consumed = 0;
                bool done = false;
#pragma unroll
                for (int q = 0; q < 4; q++)
                {
                    if (done || consumed != q || t - q < 0)
                        continue;
                    uint64_t sq = st4[q];
                    uint32_t ep = (uint32_t)(sq >> 48);
                    uint64_t flag = (sq >> 46) & 3ull;
                    if (ep != epoch || flag == 0ull)
                        continue; // not ready: stop consuming at this predecessor
                    excl += sq & ((1ull << 46) - 1ull);
                    consumed = q + 1;
                    if (flag == ST_INCL)
                        done = true;
                }
                if (done)
                    break;
                t -= consumed;
                if (consumed == 0)
                    __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(status + (uint64_t)tile * SORT_BINS + d, st_pack(epoch, ST_INCL, excl + total),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_base[d] = digit_start[d] + excl - (uint64_t)s_hist_excl[d];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SORT_ITEMS; m++)
    {
        uint64_t idx = base + (uint64_t)m * 64u + lane;
        if (idx < n)
        {
            uint32_t dd = (uint32_t)(keys[m] >> shift) & dmask;
            uint32_t pos = s_hist_excl[dd] + (uint32_t)woff[wave * SORT_BINS + dd] + rank[m];
            s_keys[pos] = keys[m];
            if (HAS_VALS)
                s_vals[pos] = vals[m];
        }
    }
    __syncthreads();
    uint64_t tile_begin = (uint64_t)tile * SORT_TILE;
    uint32_t tile_n = (uint32_t)((n - tile_begin) < (uint64_t)SORT_TILE ? (n - tile_begin) : (uint64_t)SORT_TILE);
    // 8 keys per thread at a time: their LDS reads (key, then its digit's base) are issued
    // together instead of each store waiting for two dependent reads of its own
    constexpr uint32_t U = 8;
    for (uint32_t p0 = tid; p0 < tile_n; p0 += SORT_THREADS * U)
    {
        uint64_t k[U];
        uint64_t g[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
        {
            const uint32_t p = p0 + u * SORT_THREADS;
            k[u] = p < tile_n ? s_keys[p] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            g[u] = s_base[(uint32_t)(k[u] >> shift) & dmask] + p0 + u * SORT_THREADS;
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
        {
            const uint32_t p = p0 + u * SORT_THREADS;
            if (p < tile_n)
            {
                kout[g[u]] = k[u];
                if (HAS_VALS)
                    vout[g[u]] = s_vals[p];
            }
        }
    }
}

// chosen shape and digit width for a sort of n keys (autotuned per size bucket when available)
static uint32_t size_bucket(uint64_t n)
{
    uint32_t b = 0;
    while (b < 63u && (2ull << b) <= n)
        b++;
    return b < SORT_BUCKETS ? b : SORT_BUCKETS - 1;
}

// Untuned (or an unusable entry): the default shape with the fewest passes of digits up to
// SORT_MAX_DIGIT bits, split evenly (21 bits: 7 + 7 + 7; 18 bits: 9 + 9; 64 bits: 8 x 8).
static SortConfig sort_config(const Context* ctx, uint64_t n, bool with_vals, uint32_t bits)
{
    SortConfig c = ctx->sort_tuned[size_bucket(n)];
    if (c.variant >= SORT_VARIANT_COUNT || c.digit_bits < SORT_MIN_DIGIT || c.digit_bits > SORT_MAX_DIGIT)
    {
        const uint32_t b = bits ? bits : 1u;
        const uint32_t passes = (b + SORT_MAX_DIGIT - 1u) / SORT_MAX_DIGIT;
        const uint32_t w = (b + passes - 1u) / passes;
        c = SortConfig{ 0u, std::max<uint32_t>(SORT_MIN_DIGIT, w) };
    }
    if (with_vals && !SORT_VARIANTS[c.variant].with_vals)
        c.variant = 0u;
    // 9-bit digits need a thread per digit
    if (c.digit_bits > 8u && SORT_VARIANTS[c.variant].threads < (int)SORT_BINS)
        c.digit_bits = 8u;
    return c;
}

static uint64_t sort_tile(uint32_t variant)
{
    return (uint64_t)SORT_VARIANTS[variant].threads * (uint64_t)SORT_VARIANTS[variant].items;
}

template <bool V>
static void launch_sort_pass(uint32_t variant, uint64_t tiles, hipStream_t st, const uint64_t* kin, uint64_t* kout,
                             const uint32_t* vin, uint32_t* vout, DCount n, uint32_t shift, uint32_t dbits,
                             const uint64_t* starts, uint64_t* status, uint32_t* counter, uint32_t epoch,
                             bool stable = true)
{
    if (!stable && !V && variant == 0u)
    {
        // with ranks from plain atomics the tile can double: 16384 keys per 1024-thread block (one
        // per CU) halve the look-back chain (18.9 -> 18.1 ms per 1.67 G-key pair sort); the status
        // words sized for 8192-key tiles cover half as many tiles
        hipLaunchKernelGGL((k_sort_pass<1024, 16, false, false>), dim3((uint32_t)((tiles + 1) / 2)), dim3(1024), 0, st,
                           kin, kout, nullptr, nullptr, n, shift, dbits, starts, status, counter, epoch);
        return;
    }
    switch (variant)
    {
    case 1:
        hipLaunchKernelGGL((k_sort_pass<256, 16, V>), dim3((uint32_t)tiles), dim3(256), 0, st, kin, kout, vin, vout, n,
                           shift, dbits, starts, status, counter, epoch);
        break;
    case 2:
        hipLaunchKernelGGL((k_sort_pass<1024, 8, V>), dim3((uint32_t)tiles), dim3(1024), 0, st, kin, kout, vin, vout, n,
                           shift, dbits, starts, status, counter, epoch);
        break;
    case 3:
        hipLaunchKernelGGL((k_sort_pass<256, 24, V>), dim3((uint32_t)tiles), dim3(256), 0, st, kin, kout, vin, vout, n,
                           shift, dbits, starts, status, counter, epoch);
        break;
    case 4:
        hipLaunchKernelGGL((k_sort_pass<512, 12, V>), dim3((uint32_t)tiles), dim3(512), 0, st, kin, kout, vin, vout, n,
                           shift, dbits, starts, status, counter, epoch);
        break;
    case 5: // keys only (sort_config never selects 5 or 6 with a payload)
        if (!V)
            hipLaunchKernelGGL((k_sort_pass<1024, 16, false>), dim3((uint32_t)tiles), dim3(1024), 0, st, kin, kout,
                               nullptr, nullptr, n, shift, dbits, starts, status, counter, epoch);
        break;
    case 6:
        if (!V)
            hipLaunchKernelGGL((k_sort_pass<768, 16, false>), dim3((uint32_t)tiles), dim3(768), 0, st, kin, kout,
                               nullptr, nullptr, n, shift, dbits, starts, status, counter, epoch);
        break;
    default:
        hipLaunchKernelGGL((k_sort_pass<512, 16, V>), dim3((uint32_t)tiles), dim3(512), 0, st, kin, kout, vin, vout, n,
                           shift, dbits, starts, status, counter, epoch);
        break;
    }
}

// Stable LSD sort of keys (and a u32 payload) over [bit_begin, bit_end). With a device count the
// launches are sized by its bound and the kernels sort the actual count; no host synchronisation.
int radix_sort_u64(Context* ctx, uint64_t* keys, uint32_t* vals, DCount cn, uint32_t bit_begin, uint32_t bit_end)
{
    if ((!cn.dev && cn.cap <= 1) || bit_end <= bit_begin || cn.cap == 0)
        return 0;
    const uint64_t n = cn.cap;
    uint32_t bits = bit_end - bit_begin;
    const SortConfig cfg = sort_config(ctx, n, vals != nullptr, bits);
    const uint32_t W = cfg.digit_bits;
    uint32_t passes = (bits + W - 1u) / W;
    if (passes > SORT_MAX_PASSES)
        return ctx->fail("radix sort: more than 64 key bits");
    uint64_t tiles = (n + sort_tile(cfg.variant) - 1) / sort_tile(cfg.variant);
    if (tiles > 0xFFFFFFFFull)
        return ctx->fail("radix sort: too many tiles");
    SortWorkspace ws;
    if (ctx->sort_workspace(n, vals != nullptr, tiles, true, &ws))
        return -1;
    if (fill_words(ctx, ws.hist, host_count(SORT_MAX_PASSES * SORT_BINS), 0ull) ||
        fill_words(ctx, ws.tile_counters, host_count(SORT_MAX_PASSES), 0u))
        return -1;
    uint32_t hist_blocks = (uint32_t)((n + 256 * HIST_RUN - 1) / (256 * HIST_RUN));
    if (hist_blocks > 2048u)
        hist_blocks = 2048u;
    hipLaunchKernelGGL(k_sort_hist, dim3(hist_blocks), dim3(256), 0, ctx->stream, keys, cn, bit_begin, bits, W,
                       (unsigned long long*)ws.hist);
    hipLaunchKernelGGL(k_sort_digit_starts, dim3(1), dim3(SORT_BINS), 0, ctx->stream,
                       (const unsigned long long*)ws.hist, passes, ws.starts);
    uint64_t* kin = keys;
    uint64_t* kout = ws.keys_alt;
    uint32_t* vin = vals;
    uint32_t* vout = ws.vals_alt;
    for (uint32_t p = 0; p < passes; p++)
    {
        uint32_t shift = bit_begin + W * p;
        uint32_t dbits = (bits - W * p) < W ? (bits - W * p) : W;
        uint32_t epoch = ctx->next_sort_epoch(tiles);
        if (vals)
            launch_sort_pass<true>(cfg.variant, tiles, ctx->stream, kin, kout, vin, vout, cn, shift, dbits,
                                   ws.starts + p * SORT_BINS, ws.status, ws.tile_counters + p, epoch);
        else
            launch_sort_pass<false>(cfg.variant, tiles, ctx->stream, kin, kout, nullptr, nullptr, cn, shift, dbits,
                                    ws.starts + p * SORT_BINS, ws.status, ws.tile_counters + p, epoch);
        uint64_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    if (kin != keys)
    {
        if (copy_words(ctx, (const uint64_t*)kin, keys, cn))
            return -1;
        if (vals && copy_words(ctx, (const uint32_t*)vin, vals, cn))
            return -1;
    }
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

int radix_sort_u64(Context* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin, uint32_t bit_end,
                   SortTiming* timing)
{
    (void)timing;
    return radix_sort_u64(ctx, keys, vals, host_count(n), bit_begin, bit_end);
}

__global__ void k_widen_u32(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ out)
{
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n)
        out[x] = in[x];
}

__global__ void k_narrow_u64(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out)
{
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n)
        out[x] = (uint32_t)in[x];
}

// ParallelPrimitives.cpp:347-525 (u32 keys, u32 payload): keys are widened to 64 bits
int radix_sort_u32(Context* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits)
{
    if (n <= 1 || key_bits == 0)
        return 0;
    uint64_t* wide = (uint64_t*)ctx->ws(WS_LEAF_SORT_KEYS, n * sizeof(uint64_t));
    if (!wide)
        return -1;
    uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_widen_u32, dim3(blocks), dim3(256), 0, ctx->stream, keys, n, wide);
    if (radix_sort_u64(ctx, wide, vals, n, 0, key_bits, nullptr))
        return -1;
    hipLaunchKernelGGL(k_narrow_u64, dim3(blocks), dim3(256), 0, ctx->stream, wide, n, keys);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------- dual-key sort
// Stable ascending sort by (key_high over its low high_bits bits, key_low over its low low_bits
// bits) with both keys and the payload permuted (the order of ParallelPrimitives.cpp:527-683,
// which sorts twice on 32-bit keys). Here one 64-bit composite key (high << low_bits | low) is
// sorted once with the element index as payload, then the three arrays are gathered.
__global__ void k_dual_pack(const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, uint64_t n,
                            uint32_t low_bits, uint32_t high_bits, uint64_t* __restrict__ wide, uint32_t* __restrict__ idx)
{
    const uint32_t ml = low_bits >= 32u ? 0xFFFFFFFFu : ((1u << low_bits) - 1u);
    const uint32_t mh = high_bits >= 32u ? 0xFFFFFFFFu : ((1u << high_bits) - 1u);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        wide[i] = ((uint64_t)(hi[i] & mh) << low_bits) | (uint64_t)(lo[i] & ml);
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_dual_gather(const uint32_t* __restrict__ idx, uint64_t n, const uint32_t* __restrict__ lo_src,
                              const uint32_t* __restrict__ hi_src, const uint32_t* __restrict__ val_src,
                              uint32_t* __restrict__ lo, uint32_t* __restrict__ hi, uint32_t* __restrict__ val)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        uint32_t s = idx[i];
        lo[i] = lo_src[s];
        hi[i] = hi_src[s];
        if (val)
            val[i] = val_src[s];
    }
}

int radix_sort_dual_key(Context* ctx, uint32_t* key_low, uint32_t* key_high, uint32_t* vals, uint64_t n,
                        uint32_t low_bits, uint32_t high_bits)
{
    if (n <= 1)
        return 0;
    if (n >= 0xFFFFFFFFull)
        return ctx->fail("dual-key sort: more than 2^32-1 keys");
    if (!key_low || !key_high)
        return ctx->fail("dual-key sort: null key array");
    low_bits = low_bits > 32u ? 32u : low_bits;
    high_bits = high_bits > 32u ? 32u : high_bits;
    if (low_bits + high_bits == 0u)
        return 0;
    uint64_t* wide = (uint64_t*)ctx->ws(WS_DUAL_WIDE, n * sizeof(uint64_t));
    uint32_t* tmp = (uint32_t*)ctx->ws(WS_DUAL_TMP, 4 * n * sizeof(uint32_t));
    if (!wide || !tmp)
        return -1;
    uint32_t* idx = tmp;
    uint32_t* lo_src = tmp + n;
    uint32_t* hi_src = tmp + 2 * n;
    uint32_t* val_src = tmp + 3 * n;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_dual_pack, dim3(blocks), dim3(256), 0, ctx->stream, key_low, key_high, n, low_bits, high_bits,
                       wide, idx);
    VDBR_CHECK(ctx, hipGetLastError());
    if (radix_sort_u64(ctx, wide, idx, n, 0, low_bits + high_bits, nullptr))
        return -1;
    VDBR_CHECK(ctx, hipMemcpyAsync(lo_src, key_low, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    VDBR_CHECK(ctx, hipMemcpyAsync(hi_src, key_high, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    if (vals)
        VDBR_CHECK(ctx, hipMemcpyAsync(val_src, vals, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_dual_gather, dim3(blocks), dim3(256), 0, ctx->stream, (const uint32_t*)idx, n,
                       (const uint32_t*)lo_src, (const uint32_t*)hi_src, (const uint32_t*)val_src, key_low, key_high,
                       vals);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// sort returning the buffer that holds the result (no copy back): *result_keys is either
// keys or the workspace alternate
uint32_t sort_digit_bits(const Context* ctx, uint64_t n, uint32_t bits)
{
    return sort_config(ctx, n, false, bits).digit_bits;
}

unsigned long long* sort_hist_buffer(Context* ctx)
{
    // the first part of the small sort workspace (same size as sort_workspace asks for)
    return (unsigned long long*)ctx->ws(WS_SORT_SMALL, 16 * SORT_BINS * 8 * 2 + 64);
}

int radix_sort_u64_noswap(Context* ctx, uint64_t* keys, DCount cn, uint32_t bit_begin, uint32_t bit_end,
                          uint64_t** result_keys, hipEvent_t passes_begin, hipEvent_t passes_end, uint32_t* passes_out,
                          bool hist_ready, bool free_order)
{
    *result_keys = keys;
    if (passes_out)
        *passes_out = 0;
    if ((!cn.dev && cn.cap <= 1) || bit_end <= bit_begin || cn.cap == 0)
        return 0;
    const uint64_t n = cn.cap;
    uint32_t bits = bit_end - bit_begin;
    const SortConfig cfg = sort_config(ctx, n, false, bits);
    const uint32_t W = cfg.digit_bits;
    uint32_t passes = (bits + W - 1u) / W;
    if (passes > SORT_MAX_PASSES)
        return ctx->fail("radix sort: more than 64 key bits");
    uint64_t tiles = (n + sort_tile(cfg.variant) - 1) / sort_tile(cfg.variant);
    if (tiles > 0xFFFFFFFFull)
        return ctx->fail("radix sort: too many tiles");
    SortWorkspace ws;
    if (ctx->sort_workspace(n, false, tiles, false, &ws))
        return -1;
    if (fill_words(ctx, ws.tile_counters, host_count(SORT_MAX_PASSES), 0u))
        return -1;
    if (!hist_ready)
    {
        if (fill_words(ctx, ws.hist, host_count(SORT_MAX_PASSES * SORT_BINS), 0ull))
            return -1;
        uint32_t hist_blocks = (uint32_t)((n + 256 * HIST_RUN - 1) / (256 * HIST_RUN));
        if (hist_blocks > 2048u)
            hist_blocks = 2048u;
        hipLaunchKernelGGL(k_sort_hist, dim3(hist_blocks), dim3(256), 0, ctx->stream, keys, cn, bit_begin, bits, W,
                           (unsigned long long*)ws.hist);
    }
    hipLaunchKernelGGL(k_sort_digit_starts, dim3(1), dim3(SORT_BINS), 0, ctx->stream,
                       (const unsigned long long*)ws.hist, passes, ws.starts);
    uint64_t* kin = keys;
    uint64_t* kout = ws.keys_alt;
    if (passes_begin && !ctx->capturing)
        VDBR_CHECK(ctx, hipEventRecord(passes_begin, ctx->stream));
    for (uint32_t p = 0; p < passes; p++)
    {
        uint32_t shift = bit_begin + W * p;
        uint32_t dbits = (bits - W * p) < W ? (bits - W * p) : W;
        uint32_t epoch = ctx->next_sort_epoch(tiles);
        // LSD: only the first pass may scramble keys of one digit (the later stable passes keep
        // its order only where every sorted bit is equal)
        launch_sort_pass<false>(cfg.variant, tiles, ctx->stream, kin, kout, nullptr, nullptr, cn, shift, dbits,
                                ws.starts + p * SORT_BINS, ws.status, ws.tile_counters + p, epoch,
                                !(free_order && p == 0u));
        uint64_t* tk = kin;
        kin = kout;
        kout = tk;
    }
    if (passes_end && !ctx->capturing)
        VDBR_CHECK(ctx, hipEventRecord(passes_end, ctx->stream));
    if (passes_out)
        *passes_out = passes;
    *result_keys = kin;
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// ============================================================================ sort autotuner
// Times every tile shape x digit width (9, 8 and 7 bits; 9 on the shapes of at least 512
// threads) on hashed keys of the requested size and bit count and records the fastest for that
// size bucket and all larger ones; every candidate produces the same sorted order.
__global__ void k_tune_keys(uint64_t* __restrict__ keys, uint64_t n, uint32_t key_bits, uint64_t seed)
{
    const uint64_t mask = key_bits >= 64u ? ~0ull : ((1ull << key_bits) - 1ull);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        uint64_t x = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 29;
        keys[i] = x & mask;
    }
}

int set_sort_config(Context* ctx, uint64_t from_count, uint32_t variant, uint32_t digit_bits)
{
    if (variant >= SORT_VARIANT_COUNT || digit_bits < SORT_MIN_DIGIT || digit_bits > SORT_MAX_DIGIT)
        return ctx->fail("sort config: unknown variant or digit width outside [6, 9]");
    if (digit_bits > 8u && SORT_VARIANTS[variant].threads < (int)SORT_BINS)
        return ctx->fail("sort config: 9-bit digits need a shape of at least 512 threads");
    for (uint32_t b = size_bucket(from_count); b < SORT_BUCKETS; b++)
        ctx->sort_tuned[b] = SortConfig{ variant, digit_bits };
    return 0;
}

int autotune_sort(Context* ctx, uint64_t n, uint32_t key_bits, vdbr_tune_result_t* results, uint32_t cap,
                  uint32_t* count, uint32_t* best_variant, uint32_t* best_digit_bits)
{
    if (n < 2 || key_bits == 0 || key_bits > 64)
        return ctx->fail("autotune: need at least 2 keys and 1..64 key bits");
    ctx->drop_pairs(); // the timed sorts reuse the pair sort's buffers
    uint64_t* pristine = (uint64_t*)ctx->ws(WS_TUNE_KEYS, 2 * n * sizeof(uint64_t));
    if (!pristine)
        return -1;
    uint64_t* work = pristine + n;
    hipLaunchKernelGGL(k_tune_keys, dim3(2048), dim3(256), 0, ctx->stream, pristine, n, key_bits, 0x5DEECE66Dull);
    VDBR_CHECK(ctx, hipGetLastError());
    hipEvent_t e0, e1;
    VDBR_CHECK(ctx, hipEventCreate(&e0));
    VDBR_CHECK(ctx, hipEventCreate(&e1));
    const uint32_t bucket = size_bucket(n);
    const SortConfig saved = ctx->sort_tuned[bucket];
    float best = 3.0e38f;
    SortConfig best_cfg{ 0u, 8u };
    uint32_t k = 0;
    int rc = 0;
    const uint32_t widths[3] = { 9u, 8u, 7u };
    for (uint32_t v = 0; v < SORT_VARIANT_COUNT && rc == 0; v++)
        for (uint32_t wi = 0; wi < 3 && rc == 0; wi++)
        {
            if (widths[wi] > 8u && SORT_VARIANTS[v].threads < (int)SORT_BINS)
                continue; // 9-bit digits need a thread per digit
            ctx->sort_tuned[bucket] = SortConfig{ v, widths[wi] };
            float ms_best = 3.0e38f;
            for (int rep = 0; rep < 2 && rc == 0; rep++)
            {
                uint64_t* out = nullptr;
                if (hipMemcpyAsync(work, pristine, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream) !=
                        hipSuccess ||
                    hipEventRecord(e0, ctx->stream) != hipSuccess)
                {
                    rc = ctx->fail("autotune: copy or event failed");
                    break;
                }
                if (radix_sort_u64_noswap(ctx, work, host_count(n), 0, key_bits, &out, nullptr, nullptr, nullptr))
                {
                    rc = -1;
                    break;
                }
                float ms = 0.f;
                if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                {
                    rc = ctx->fail("autotune: timing failed");
                    break;
                }
                ms_best = ms < ms_best ? ms : ms_best;
            }
            if (rc)
                break;
            if (results && k < cap)
            {
                results[k].variant = v;
                results[k].threads = (uint32_t)SORT_VARIANTS[v].threads;
                results[k].keys_per_thread = (uint32_t)SORT_VARIANTS[v].items;
                results[k].digit_bits = widths[wi];
                results[k].ms = ms_best;
            }
            k++;
            if (ms_best < best)
            {
                best = ms_best;
                best_cfg = ctx->sort_tuned[bucket];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->ws_free(WS_TUNE_KEYS); // the hashed keys are needed only while tuning
    if (rc)
    {
        ctx->sort_tuned[bucket] = saved;
        return rc;
    }
    for (uint32_t b = bucket; b < SORT_BUCKETS; b++)
        ctx->sort_tuned[b] = best_cfg;
    if (count)
        *count = k;
    if (best_variant)
        *best_variant = best_cfg.variant;
    if (best_digit_bits)
        *best_digit_bits = best_cfg.digit_bits;
    return 0;
}

// ============================================================================ allocation-free sort
// Scratch of vdbr_radix_sort_u64_noalloc: the ping-pong keys and payload, the look-back status
// words and the digit histograms of one sort of at most NOALLOC_MAX_KEYS keys, in one
// file-scope device array. The loader places a copy in the memory of every device the code
// object is loaded on; it stays there until the process ends. Calls on one device share it.
constexpr uint64_t NOALLOC_MAX_KEYS = 1ull << 22;
constexpr uint32_t NOALLOC_VARIANT = 0u; // 512 threads x 16 keys, 8-bit digits
constexpr uint64_t NOALLOC_TILE = 512ull * 16ull;
constexpr uint64_t NOALLOC_MAX_TILES = (NOALLOC_MAX_KEYS + NOALLOC_TILE - 1) / NOALLOC_TILE;
constexpr size_t NA_KEYS_OFF = 0;
constexpr size_t NA_VALS_OFF = NA_KEYS_OFF + NOALLOC_MAX_KEYS * sizeof(uint64_t);
constexpr size_t NA_STATUS_OFF = NA_VALS_OFF + NOALLOC_MAX_KEYS * sizeof(uint32_t);
constexpr size_t NA_HIST_OFF = NA_STATUS_OFF + NOALLOC_MAX_TILES * SORT_BINS * sizeof(uint64_t);
constexpr size_t NA_STARTS_OFF = NA_HIST_OFF + SORT_MAX_PASSES * SORT_BINS * sizeof(uint64_t);
constexpr size_t NA_COUNTERS_OFF = NA_STARTS_OFF + SORT_MAX_PASSES * SORT_BINS * sizeof(uint64_t);
constexpr size_t NA_BYTES = NA_COUNTERS_OFF + 16 * sizeof(uint32_t);
static_assert(NOALLOC_MAX_KEYS % NOALLOC_TILE == 0, "whole tiles");
static_assert(NA_STATUS_OFF % 256 == 0 && NA_HIST_OFF % 8 == 0, "aligned regions");

static __device__ __align__(256) uint8_t scratch[NA_BYTES];

// zero fill by a kernel rather than hipMemsetAsync, so the call can sit inside captured work
__global__ void k_fill_zero_u64(uint64_t* __restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

namespace
{
constexpr int NOALLOC_DEVICES = 64;
struct NoAllocDevice
{
    uint8_t* ptr = nullptr; // device address of `scratch` on this device (from hipGetSymbolAddress)
    uint32_t epoch = 0;     // last look-back epoch used on this device's status words
};
std::mutex g_noalloc_mutex;
NoAllocDevice g_noalloc[NOALLOC_DEVICES];
} // namespace

uint64_t noalloc_max_keys()
{
    return NOALLOC_MAX_KEYS;
}

static int noalloc_sort_on_device(int dev, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin,
                                  uint32_t bit_end, hipStream_t st, std::string* err);

int radix_sort_u64_noalloc(int dev, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin, uint32_t bit_end,
                           hipStream_t st, std::string* err)
{
    // run with dev current (its copy of the scratch array), then restore the caller's device
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess)
    {
        (void)hipGetLastError();
        return *err = "radix_sort_u64_noalloc: no current device", -1;
    }
    if (cur != dev && hipSetDevice(dev) != hipSuccess)
    {
        (void)hipGetLastError();
        return *err = "radix_sort_u64_noalloc: cannot select device " + std::to_string(dev), -1;
    }
    const int rc = noalloc_sort_on_device(dev, keys, vals, n, bit_begin, bit_end, st, err);
    if (cur != dev)
        (void)hipSetDevice(cur);
    return rc;
}

static int noalloc_sort_on_device(int dev, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin,
                                  uint32_t bit_end, hipStream_t st, std::string* err)
{
    if (bit_begin > bit_end || bit_end > 64u)
        return *err = "radix_sort_u64_noalloc: the bit range must satisfy bit_begin <= bit_end <= 64", -1;
    if (n > NOALLOC_MAX_KEYS)
        return *err = "radix_sort_u64_noalloc: " + std::to_string(n) + " keys exceed the scratch maximum of " +
                      std::to_string(NOALLOC_MAX_KEYS),
               -1;
    if (n <= 1 || bit_end == bit_begin)
        return 0;
    if (!keys)
        return *err = "radix_sort_u64_noalloc: null key array", -1;
    if (dev < 0 || dev >= NOALLOC_DEVICES)
        return *err = "radix_sort_u64_noalloc: device index out of range", -1;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    NoAllocDevice& d = g_noalloc[dev];
    const uint32_t bits = bit_end - bit_begin;
    const uint32_t W = 8u;
    const uint32_t passes = (bits + W - 1u) / W;
    const uint64_t tiles = (n + NOALLOC_TILE - 1) / NOALLOC_TILE;
    if (!d.ptr)
    {
        // first call on this device: its own copy of the array, then clear the status words
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            (void)hipGetLastError();
            return *err = std::string("radix_sort_u64_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e), -1;
        }
        d.ptr = (uint8_t*)p;
        d.epoch = 0u;
        hipLaunchKernelGGL(k_fill_zero_u64, dim3(256), dim3(256), 0, st, (uint64_t*)(d.ptr + NA_STATUS_OFF),
                           (uint64_t)NOALLOC_MAX_TILES * SORT_BINS);
    }
    if (d.epoch + passes >= 65535u)
    {
        // epochs are 16-bit: clear the status words before they could repeat
        hipLaunchKernelGGL(k_fill_zero_u64, dim3(256), dim3(256), 0, st, (uint64_t*)(d.ptr + NA_STATUS_OFF),
                           (uint64_t)NOALLOC_MAX_TILES * SORT_BINS);
        d.epoch = 0u;
    }
    uint64_t* keys_alt = (uint64_t*)(d.ptr + NA_KEYS_OFF);
    uint32_t* vals_alt = (uint32_t*)(d.ptr + NA_VALS_OFF);
    uint64_t* status = (uint64_t*)(d.ptr + NA_STATUS_OFF);
    unsigned long long* hist = (unsigned long long*)(d.ptr + NA_HIST_OFF);
    uint64_t* starts = (uint64_t*)(d.ptr + NA_STARTS_OFF);
    uint32_t* counters = (uint32_t*)(d.ptr + NA_COUNTERS_OFF);
    // histograms, digit starts and tile counters start at zero on every call
    hipLaunchKernelGGL(k_fill_zero_u64, dim3(8), dim3(256), 0, st, (uint64_t*)hist,
                       (uint64_t)((NA_BYTES - NA_HIST_OFF) / sizeof(uint64_t)));
    uint32_t hist_blocks = (uint32_t)((n + 256 * HIST_RUN - 1) / (256 * HIST_RUN));
    if (hist_blocks > 2048u)
        hist_blocks = 2048u;
    hipLaunchKernelGGL(k_sort_hist, dim3(hist_blocks), dim3(256), 0, st, keys, host_count(n), bit_begin, bits, W, hist);
    hipLaunchKernelGGL(k_sort_digit_starts, dim3(1), dim3(SORT_BINS), 0, st, (const unsigned long long*)hist, passes,
                       starts);
    uint64_t* kin = keys;
    uint64_t* kout = keys_alt;
    uint32_t* vin = vals;
    uint32_t* vout = vals_alt;
    for (uint32_t p = 0; p < passes; p++)
    {
        const uint32_t shift = bit_begin + W * p;
        const uint32_t dbits = (bits - W * p) < W ? (bits - W * p) : W;
        const uint32_t epoch = ++d.epoch;
        if (vals)
            launch_sort_pass<true>(NOALLOC_VARIANT, tiles, st, kin, kout, vin, vout, host_count(n), shift, dbits,
                                   starts + p * SORT_BINS, status, counters + p, epoch);
        else
            launch_sort_pass<false>(NOALLOC_VARIANT, tiles, st, kin, kout, nullptr, nullptr, host_count(n), shift, dbits,
                                    starts + p * SORT_BINS, status, counters + p, epoch);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys)
    {
        hipError_t e = hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && vals)
            e = hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess)
            return *err = std::string("radix_sort_u64_noalloc: copy back: ") + hipGetErrorString(e), -1;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return *err = std::string("radix_sort_u64_noalloc: launch: ") + hipGetErrorString(e), -1;
    return 0;
}

} // namespace vdbr
