// vdb_tiles.hip -- tile binning of projected Gaussians: the scan/sort consumer of Raster2D.
//
// Restates raster/Raster2D.cpp:347-535: per-Gaussian tile counts (gaussian_count_tiles.slang:14-46),
// inclusive scan, intersection records (gaussian_tile_intersections.slang:18-65), the dual-key sort
// by (tile index, depth bits) (Raster2D.cpp:495-507) and the per-tile offsets
// (gaussian_tile_offsets.slang:13-56). Records are written once as 64-bit composite keys
// (tile << 32 | depth bits) and sorted in one onesweep sort over 32 + tile-id bits; the order is
// the stable (high, low) order of the reference's two-sort dual-key primitive. Intersections are
// generated one thread per record (binary search of the count scan), so a Gaussian covering
// many tiles does not serialise a lane.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vdb_internal.h"

namespace vdbr
{

// float -> u32 as v_cvt_u32_f32 converts: NaN and negatives to 0, saturating at 2^32-1 (the
// shaders write (uint32_t)floor(x), undefined below zero in C; both GPU vendors saturate)
__device__ __forceinline__ uint32_t sat_u32(float x)
{
    if (!(x > 0.f))
        return 0u;
    if (x >= 4294967295.f)
        return 0xFFFFFFFFu;
    return (uint32_t)x;
}

struct TileBox
{
    uint32_t x0, y0, x1, y1;
};

// gaussian_count_tiles.slang:27-44
__device__ __forceinline__ TileBox tile_box(float mx, float my, int32_t radius, uint32_t tile_size, uint32_t tiles_w,
                                            uint32_t tiles_h)
{
    const float ts = (float)tile_size;
    const float tile_radius = (float)radius / ts;
    const float u = mx / ts;
    const float v = my / ts;
    TileBox b;
    b.x0 = min(sat_u32(floorf(u - tile_radius)), tiles_w);
    b.y0 = min(sat_u32(floorf(v - tile_radius)), tiles_h);
    b.x1 = min(sat_u32(ceilf(u + tile_radius)), tiles_w);
    b.y1 = min(sat_u32(ceilf(v + tile_radius)), tiles_h);
    return b;
}

__global__ void k_tile_count(const float2* __restrict__ means2d, const int32_t* __restrict__ radii, uint32_t n,
                             uint32_t tile_size, uint32_t tiles_w, uint32_t tiles_h, uint32_t* __restrict__ counts)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x)
    {
        int32_t r = radii[g];
        uint32_t c = 0u;
        if (r > 0)
        {
            float2 m = means2d[g];
            TileBox b = tile_box(m.x, m.y, r, tile_size, tiles_w, tiles_h);
            c = (b.y1 - b.y0) * (b.x1 - b.x0);
        }
        counts[g] = c;
    }
}

// one thread per intersection record e: the owning Gaussian is the first g with scan[g] > e
__global__ void k_tile_records(const float2* __restrict__ means2d, const int32_t* __restrict__ radii,
                               const float* __restrict__ depths, const uint32_t* __restrict__ scan, uint32_t n,
                               uint32_t total, uint32_t tile_size, uint32_t tiles_w, uint32_t tiles_h,
                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x)
    {
        uint32_t lo = 0u, hi = n;
        while (lo < hi)
        {
            uint32_t mid = lo + ((hi - lo) >> 1);
            if (scan[mid] > e)
                hi = mid;
            else
                lo = mid + 1u;
        }
        const uint32_t g = lo;
        const uint32_t local = e - (g ? scan[g - 1] : 0u);
        float2 m = means2d[g];
        TileBox b = tile_box(m.x, m.y, radii[g], tile_size, tiles_w, tiles_h);
        const uint32_t nx = b.x1 - b.x0;
        const uint32_t ty = b.y0 + local / nx;
        const uint32_t tx = b.x0 + local % nx;
        keys[e] = ((uint64_t)(ty * tiles_w + tx) << 32) | (uint64_t)__float_as_uint(depths[g]);
        vals[e] = g;
    }
}

// split the sorted composite keys and write the tile offsets (gaussian_tile_offsets.slang:24-56)
__global__ void k_tile_split(const uint64_t* __restrict__ keys, uint32_t total, uint32_t num_tiles,
                             uint32_t* __restrict__ key_low, uint32_t* __restrict__ key_high,
                             uint32_t* __restrict__ offsets)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x)
    {
        const uint64_t k = keys[e];
        const uint32_t t = (uint32_t)(k >> 32);
        key_low[e] = (uint32_t)k;
        key_high[e] = t;
        if (!offsets)
            continue;
        if (e == 0u)
            for (uint32_t i = 0; i <= t && i < num_tiles; i++)
                offsets[i] = 0u;
        if (e == total - 1u)
            for (uint32_t i = t + 1u; i < num_tiles; i++)
                offsets[i] = total;
        if (e > 0u)
        {
            const uint32_t p = (uint32_t)(keys[e - 1] >> 32);
            if (p != t)
                for (uint32_t i = p + 1u; i <= t && i < num_tiles; i++)
                    offsets[i] = e;
        }
    }
}

int tile_bin(Context* ctx, const float* means2d, const int32_t* radii, const float* depths, uint32_t n,
             uint32_t width, uint32_t height, uint32_t tile_size, uint32_t* offsets, uint32_t** keys_low_out,
             uint32_t** keys_high_out, uint32_t** vals_out, uint64_t* count_out)
{
    if (!count_out)
        return ctx->fail("tile_bin: count is NULL");
    *count_out = 0;
    if (tile_size == 0 || width == 0 || height == 0)
        return ctx->fail("tile_bin: image size and tile size must be positive");
    if (n > 0 && (!means2d || !radii || !depths))
        return ctx->fail("tile_bin: null input array");
    const uint32_t tiles_w = (width + tile_size - 1u) / tile_size; // Raster2D.cpp:202-205
    const uint32_t tiles_h = (height + tile_size - 1u) / tile_size;
    const uint64_t tiles64 = (uint64_t)tiles_w * tiles_h;
    if (tiles64 > 0xFFFFFFFFull)
        return ctx->fail("tile_bin: more than 2^32-1 tiles");
    const uint32_t num_tiles = (uint32_t)tiles64;
    hipStream_t st = ctx->stream;
    uint64_t total = 0;
    uint32_t* scan = nullptr;
    if (n > 0)
    {
        scan = (uint32_t*)ctx->ws(WS_TILE_SCAN, (uint64_t)(n + 1) * sizeof(uint32_t));
        if (!scan)
            return -1;
        const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 4096u);
        hipLaunchKernelGGL(k_tile_count, dim3(blocks), dim3(256), 0, st, (const float2*)means2d, radii, n, tile_size,
                           tiles_w, tiles_h, scan);
        VDBR_CHECK(ctx, hipGetLastError());
        if (scan_u32(ctx, scan, scan, n, true, scan + n))
            return -1;
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, scan + n, 4, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        total = (uint32_t)ctx->pinned[0];
    }
    if (total == 0)
    {
        if (offsets && num_tiles)
            VDBR_CHECK(ctx, hipMemsetAsync(offsets, 0, (size_t)num_tiles * 4, st));
        if (keys_low_out)
            *keys_low_out = nullptr;
        if (keys_high_out)
            *keys_high_out = nullptr;
        if (vals_out)
            *vals_out = nullptr;
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        return 0;
    }
    uint64_t* keys = (uint64_t*)ctx->ws(WS_TILE_KEYS, total * sizeof(uint64_t));
    uint32_t* out = (uint32_t*)ctx->ws(WS_TILE_OUT, 3 * total * sizeof(uint32_t));
    if (!keys || !out)
        return -1;
    uint32_t* vals = out;
    uint32_t* klo = out + total;
    uint32_t* khi = out + 2 * total;
    const uint32_t rblocks = (uint32_t)std::min<uint64_t>((total + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_tile_records, dim3(rblocks), dim3(256), 0, st, (const float2*)means2d, radii, depths,
                       (const uint32_t*)scan, n, (uint32_t)total, tile_size, tiles_w, tiles_h, keys, vals);
    VDBR_CHECK(ctx, hipGetLastError());
    uint32_t tile_bits = 0;
    while (tile_bits < 32u && (1ull << tile_bits) < num_tiles)
        tile_bits++;
    if (radix_sort_u64(ctx, keys, vals, total, 0, 32u + tile_bits, nullptr))
        return -1;
    hipLaunchKernelGGL(k_tile_split, dim3(rblocks), dim3(256), 0, st, (const uint64_t*)keys, (uint32_t)total, num_tiles,
                       klo, khi, offsets);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    if (keys_low_out)
        *keys_low_out = klo;
    if (keys_high_out)
        *keys_high_out = khi;
    if (vals_out)
        *vals_out = vals;
    *count_out = total;
    return 0;
}

} // namespace vdbr
