// vdb_bvh.hip -- VoxelBVH Gaussian keys: the 64-bit sort consumer of the VoxelBVH build.
//
// Restates raster/VoxelBVH.cpp:1225-1435 (ijkl_from_gaussians): the world box of all Gaussian
// bounding cubes (voxelbvh_gaussians_bbox_reduce1/2.slang), up to eight (i, j, k, level) cell keys
// per Gaussian at the finest level whose overlapping cells still number at most eight
// (voxelbvh_gaussians_to_ijkl.slang), the stable 64-bit key sort with the Gaussian index as payload
// (radix_sort_key64), and the [start, end) header of every run of equal keys
// (voxelbvh_find_range_starts.slang, global_scan, voxelbvh_scatter_range_headers.slang).
//
// Arithmetic follows the shaders' FP32 expressions in written order without FMA contraction;
// log/division/square root are correctly rounded (vdb_fpmath.h). The optional 4x4 transform is
// applied as a row vector times the matrix, summed left to right.
//
// Sorting: when every key's cell coordinates fit 12 bits (always the case inside the world box at
// resolution <= 4096), keys are packed to 40 bits (i:12 j:12 k:12 level:4, unused slots
// 2^40 - 1), which orders exactly as the full keys, and sorted in 5 onesweep passes instead of 8.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vdb_device.h"
#include "vdb_fpmath.h"
#include "vdb_internal.h"

namespace vdbr
{

struct BvhTransform
{
    float m[16];   // pnanovdb_camera_mat_t rows x, y, z, w: p' = (p, 1) * m
    float inv[16]; // its inverse (pnanovdb_camera_mat_inverse)
};

struct BvhGaussian
{
    float3 mean;
    float4 q; // (x, y, z, w)
    float3 scale;
    float local_scale;
};

__device__ __forceinline__ float3 bvh_xf(const float* m, float3 p)
{
    float3 r;
    r.x = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    r.y = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    r.z = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    return r;
}

__device__ __forceinline__ BvhGaussian bvh_load(const vdbr_gaussians_t& g, uint32_t idx)
{
    BvhGaussian b;
    b.mean = make_float3(g.means[3 * idx + 0], g.means[3 * idx + 1], g.means[3 * idx + 2]);
    b.q = make_float4(g.quats[4 * idx + 1], g.quats[4 * idx + 2], g.quats[4 * idx + 3], g.quats[4 * idx + 0]);
    b.scale = make_float3(g.scales[3 * idx + 0], g.scales[3 * idx + 1], g.scales[3 * idx + 2]);
    // voxelbvh_gaussians_to_ijkl.slang:49
    b.local_scale = fsqrt_rn(2.f * vdb_logf(fdiv_rn(fmaxf(0.01f, g.opacities[idx]), 0.01f)));
    return b;
}

// world-space box of the Gaussian's oriented bounding cube (voxelbvh_gaussians_bbox_reduce1.slang:56-89)
__device__ __forceinline__ void bvh_aabb(const BvhGaussian& g, const BvhTransform& xf, float3& mn, float3& mx)
{
    const float x = g.q.x, y = g.q.y, z = g.q.z, w = g.q.w;
    const float3 r0 = make_float3(1.f - 2.f * (y * y + z * z), 2.f * (x * y - z * w), 2.f * (x * z + y * w));
    const float3 r1 = make_float3(2.f * (x * y + z * w), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - x * w));
    const float3 r2 = make_float3(2.f * (x * z - y * w), 2.f * (y * z + x * w), 1.f - 2.f * (x * x + y * y));
    const float3 ext = make_float3(g.scale.x * g.local_scale, g.scale.y * g.local_scale, g.scale.z * g.local_scale);
    mn = make_float3(__int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000));
    mx = make_float3(__int_as_float(0xFF800000), __int_as_float(0xFF800000), __int_as_float(0xFF800000));
#pragma unroll
    for (uint32_t v = 0; v < 8u; v++)
    {
        const float3 c = make_float3((v & 1u) ? ext.x : -ext.x, (v & 2u) ? ext.y : -ext.y, (v & 4u) ? ext.z : -ext.z);
        float3 p;
        p.x = c.x * r0.x + c.y * r0.y + c.z * r0.z + g.mean.x;
        p.y = c.x * r1.x + c.y * r1.y + c.z * r1.z + g.mean.y;
        p.z = c.x * r2.x + c.y * r2.y + c.z * r2.z + g.mean.z;
        p = bvh_xf(xf.m, p);
        mn = make_float3(fminf(mn.x, p.x), fminf(mn.y, p.y), fminf(mn.z, p.z));
        mx = make_float3(fmaxf(mx.x, p.x), fmaxf(mx.y, p.y), fmaxf(mx.z, p.z));
    }
}

// per-block box of the Gaussians' boxes (voxelbvh_gaussians_bbox_reduce1.slang)
__global__ __launch_bounds__(256) void k_bvh_bbox1(vdbr_gaussians_t g, uint32_t n, BvhTransform xf,
                                                   float* __restrict__ partial)
{
    __shared__ float s[6][4];
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    float3 mn = make_float3(__int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000));
    float3 mx = make_float3(__int_as_float(0xFF800000), __int_as_float(0xFF800000), __int_as_float(0xFF800000));
    if (idx < n)
        bvh_aabb(bvh_load(g, idx), xf, mn, mx);
    float v[6] = { mn.x, mn.y, mn.z, mx.x, mx.y, mx.z };
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
    {
#pragma unroll
        for (int a = 0; a < 6; a++)
        {
            float o = __shfl_xor(v[a], off, 64);
            v[a] = a < 3 ? fminf(v[a], o) : fmaxf(v[a], o);
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u)
        for (int a = 0; a < 6; a++)
            s[a][wave] = v[a];
    __syncthreads();
    if (threadIdx.x < 6u)
    {
        const uint32_t a = threadIdx.x;
        float r = s[a][0];
        for (int w = 1; w < 4; w++)
            r = a < 3 ? fminf(r, s[a][w]) : fmaxf(r, s[a][w]);
        partial[6u * blockIdx.x + a] = r;
    }
}

// box of the per-block boxes (voxelbvh_gaussians_bbox_reduce2.slang), one block
__global__ __launch_bounds__(256) void k_bvh_bbox2(const float* __restrict__ partial, uint32_t blocks,
                                                   float* __restrict__ bbox)
{
    __shared__ float s[6][4];
    float v[6] = { __int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000),
                   __int_as_float(0xFF800000), __int_as_float(0xFF800000), __int_as_float(0xFF800000) };
    for (uint32_t b = threadIdx.x; b < blocks; b += 256u)
        for (int a = 0; a < 6; a++)
            v[a] = a < 3 ? fminf(v[a], partial[6u * b + a]) : fmaxf(v[a], partial[6u * b + a]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
    {
#pragma unroll
        for (int a = 0; a < 6; a++)
        {
            float o = __shfl_xor(v[a], off, 64);
            v[a] = a < 3 ? fminf(v[a], o) : fmaxf(v[a], o);
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u)
        for (int a = 0; a < 6; a++)
            s[a][wave] = v[a];
    __syncthreads();
    if (threadIdx.x < 6u)
    {
        const uint32_t a = threadIdx.x;
        float r = s[a][0];
        for (int w = 1; w < 4; w++)
            r = a < 3 ? fminf(r, s[a][w]) : fmaxf(r, s[a][w]);
        bbox[a] = r;
    }
}

// float -> int32 as v_cvt_i32_f32 converts (NaN to 0, saturating)
__device__ __forceinline__ int32_t sat_i32(float x)
{
    if (!(x == x))
        return 0;
    if (x <= -2147483648.f)
        return INT32_MIN;
    if (x >= 2147483648.f)
        return INT32_MAX;
    return (int32_t)x;
}

struct BvhFrame
{
    float3 rinv0, rinv1, rinv2; // rows of R^T (voxelbvh_gaussians_to_ijkl.slang:66-77)
    float3 inv_ext;             // 1 / (local_scale * scale)
};

// voxelbvh_gaussians_to_ijkl.slang:29-105: does the level cell at ijk meet the Gaussian's unit ellipsoid?
__device__ __forceinline__ bool bvh_overlap(const BvhGaussian& g, const BvhFrame& f, const BvhTransform& xf,
                                            float4 w2v, int32_t i, int32_t j, int32_t k, uint32_t level)
{
    float3 vmin = make_float3(0.f, 0.f, 0.f), vmax = vmin;
    const float step = (float)(1u << level);
#pragma unroll
    for (uint32_t v = 0; v < 8u; v++)
    {
        float3 vp = make_float3((float)i, (float)j, (float)k);
        if (v & 1u)
            vp.x += step;
        if (v & 2u)
            vp.y += step;
        if (v & 4u)
            vp.z += step;
        float3 lp = make_float3(fdiv_rn(vp.x - w2v.x, w2v.w), fdiv_rn(vp.y - w2v.y, w2v.w), fdiv_rn(vp.z - w2v.z, w2v.w));
        lp = bvh_xf(xf.inv, lp);
        lp = make_float3(lp.x - g.mean.x, lp.y - g.mean.y, lp.z - g.mean.z);
        float3 r;
        r.x = lp.x * f.rinv0.x + lp.y * f.rinv0.y + lp.z * f.rinv0.z;
        r.y = lp.x * f.rinv1.x + lp.y * f.rinv1.y + lp.z * f.rinv1.z;
        r.z = lp.x * f.rinv2.x + lp.y * f.rinv2.y + lp.z * f.rinv2.z;
        r = make_float3(r.x * f.inv_ext.x, r.y * f.inv_ext.y, r.z * f.inv_ext.z);
        if (v == 0u)
        {
            vmin = r;
            vmax = r;
        }
        else
        {
            vmin = make_float3(fminf(vmin.x, r.x), fminf(vmin.y, r.y), fminf(vmin.z, r.z));
            vmax = make_float3(fmaxf(vmax.x, r.x), fmaxf(vmax.y, r.y), fmaxf(vmax.z, r.z));
        }
    }
    float d2 = 1.f;
    if (0.f < vmin.x)
        d2 -= vmin.x * vmin.x;
    else if (0.f > vmax.x)
        d2 -= vmax.x * vmax.x;
    if (0.f < vmin.y)
        d2 -= vmin.y * vmin.y;
    else if (0.f > vmax.y)
        d2 -= vmax.y * vmax.y;
    if (0.f < vmin.z)
        d2 -= vmin.z * vmin.z;
    else if (0.f > vmax.z)
        d2 -= vmax.z * vmax.z;
    return d2 > 0.f;
}

constexpr uint64_t BVH_COMPACT_NONE = (1ull << 40) - 1ull;

// voxelbvh_gaussians_to_ijkl.slang:117-258, one thread per Gaussian; keys[8 idx + s], ids[8 idx + s]
__global__ __launch_bounds__(256) void k_bvh_keys(vdbr_gaussians_t g, uint32_t n, uint32_t resolution, BvhTransform xf,
                                                  const float* __restrict__ bbox, uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ ids, uint32_t* __restrict__ not_compact)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= n)
        return;
    // compute_world_to_voxel (voxelbvh_gaussians_to_ijkl.slang:107-114)
    const float3 wmin = make_float3(bbox[0], bbox[1], bbox[2]);
    const float3 wmax = make_float3(bbox[3], bbox[4], bbox[5]);
    const float3 wd = make_float3(wmax.x - wmin.x, wmax.y - wmin.y, wmax.z - wmin.z);
    const float wdm = fmaxf(wd.x, fmaxf(wd.y, wd.z));
    const float sc = fdiv_rn((float)(resolution - 1u), wdm);
    const float4 w2v = make_float4(-wmin.x * sc + 0.5f, -wmin.y * sc + 0.5f, -wmin.z * sc + 0.5f, sc);

    const BvhGaussian gs = bvh_load(g, idx);
    float3 amin, amax;
    bvh_aabb(gs, xf, amin, amax);
    const int32_t imin[3] = { sat_i32(floorf(w2v.w * amin.x + w2v.x)), sat_i32(floorf(w2v.w * amin.y + w2v.y)),
                              sat_i32(floorf(w2v.w * amin.z + w2v.z)) };
    const int32_t imax[3] = { sat_i32(floorf(w2v.w * amax.x + w2v.x)), sat_i32(floorf(w2v.w * amax.y + w2v.y)),
                              sat_i32(floorf(w2v.w * amax.z + w2v.z)) };
    // cells of a finite Gaussian lie in [0, resolution) (the world box contains every Gaussian box);
    // non-finite parameters would make the cell loops unbounded: flag them and emit no cell
    const int32_t lim = (int32_t)resolution + 1;
    bool bounded = true;
    for (int a = 0; a < 3; a++)
        bounded = bounded && imin[a] >= -2 && imax[a] <= lim && imin[a] <= imax[a];
    if (!bounded)
    {
        for (uint32_t sub = 0; sub < 8u; sub++)
        {
            keys[8ull * idx + sub] = ~0ull;
            ids[8ull * idx + sub] = ~0u;
        }
        atomicOr(not_compact, 2u);
        return;
    }
    BvhFrame f;
    {
        const float x = gs.q.x, y = gs.q.y, z = gs.q.z, w = gs.q.w;
        f.rinv0 = make_float3(1.f - 2.f * (y * y + z * z), 2.f * (x * y + z * w), 2.f * (x * z - y * w));
        f.rinv1 = make_float3(2.f * (x * y - z * w), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + x * w));
        f.rinv2 = make_float3(2.f * (x * z + y * w), 2.f * (y * z - x * w), 1.f - 2.f * (x * x + y * y));
        f.inv_ext = make_float3(fdiv_rn(1.f, gs.local_scale * gs.scale.x), fdiv_rn(1.f, gs.local_scale * gs.scale.y),
                                fdiv_rn(1.f, gs.local_scale * gs.scale.z));
    }
    // finest level (from 11 down) whose overlapping cells number at most 8; a level with more
    // stops the search at the previous one (the count only matters up to 9)
    uint32_t level = 11u, last_good = 11u;
    while (level < 12u)
    {
        const int32_t mask = ~((1 << level) - 1);
        uint32_t count = 0u;
        for (int32_t i = imin[0] & mask; i <= (imax[0] & mask) && count <= 8u; i += (1 << level))
            for (int32_t j = imin[1] & mask; j <= (imax[1] & mask) && count <= 8u; j += (1 << level))
                for (int32_t k = imin[2] & mask; k <= (imax[2] & mask) && count <= 8u; k += (1 << level))
                    count += bvh_overlap(gs, f, xf, w2v, i, j, k, level) ? 1u : 0u;
        if (count > 8u)
        {
            level = last_good;
            break;
        }
        if (level == 0u)
            break;
        last_good = level;
        level -= 1u;
    }
    uint32_t sub = 0u;
    bool compact = true;
    const int32_t mask = ~((1 << level) - 1);
    for (int32_t i = imin[0] & mask; i <= (imax[0] & mask) && sub < 8u; i += (1 << level))
        for (int32_t j = imin[1] & mask; j <= (imax[1] & mask) && sub < 8u; j += (1 << level))
            for (int32_t k = imin[2] & mask; k <= (imax[2] & mask) && sub < 8u; k += (1 << level))
                if (bvh_overlap(gs, f, xf, w2v, i, j, k, level))
                {
                    keys[8ull * idx + sub] = ((uint64_t)(int64_t)i << 48) | ((uint64_t)(int64_t)j << 32) |
                                             ((uint64_t)(int64_t)k << 16) | (uint64_t)level;
                    ids[8ull * idx + sub] = idx;
                    compact = compact && (uint32_t)i < 4096u && (uint32_t)j < 4096u && (uint32_t)k < 4096u;
                    sub++;
                }
    for (; sub < 8u; sub++)
    {
        keys[8ull * idx + sub] = ~0ull;
        ids[8ull * idx + sub] = ~0u;
    }
    if (!compact)
        atomicOr(not_compact, 1u);
}

// full key -> 40-bit key (i:12 j:12 k:12 level:4); unused slots -> 2^40 - 1
__global__ void k_bvh_pack(uint64_t* __restrict__ keys, uint64_t count)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint64_t k = keys[e];
        keys[e] = k == ~0ull ? BVH_COMPACT_NONE
                             : (((k >> 48) & 0xFFFull) << 28) | (((k >> 32) & 0xFFFull) << 16) |
                                   (((k >> 16) & 0xFFFull) << 4) | (k & 0xFull);
    }
}

// run-start flags (voxelbvh_find_range_starts.slang:28-40); packed keys compare as the full keys
__global__ void k_bvh_starts(const uint64_t* __restrict__ keys, uint64_t count, uint32_t* __restrict__ starts)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (uint64_t)gridDim.x * blockDim.x)
        starts[e] = (e == 0 || keys[e - 1] != keys[e]) ? 1u : 0u;
}

// 40-bit key -> full key
__global__ void k_bvh_unpack(uint64_t* __restrict__ keys, uint64_t count)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint64_t k = keys[e];
        keys[e] = k == BVH_COMPACT_NONE ? ~0ull
                                        : (((k >> 28) & 0xFFFull) << 48) | (((k >> 16) & 0xFFFull) << 32) |
                                              (((k >> 4) & 0xFFFull) << 16) | (k & 0xFull);
    }
}

// [start, end) header of every run (voxelbvh_scatter_range_headers.slang:29-54); scan is inclusive
__global__ void k_bvh_headers(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ starts,
                              const uint32_t* __restrict__ scan, uint64_t count, uint32_t* __restrict__ headers)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint32_t slot = scan[e] - 1u;
        if (starts[e])
            headers[2ull * slot] = (uint32_t)e;
        if (e + 1 == count || keys[e + 1] != keys[e])
            headers[2ull * slot + 1] = (uint32_t)(e + 1);
    }
}

// pnanovdb_camera_mat_inverse (putil/Camera.h:316-375): the cofactor expansion with the terms in the
// same order, so the inverse is bit-identical; term t of a sum is the product of m[r][perm[t][r]]
// over the rows in the sum, with its sign
static void mat_inverse(const float* m, float* out)
{
    auto at = [&](int r, int c) { return m[4 * r + c]; };
    // determinant: 24 products m[0][a] m[1][b] m[2][c] m[3][d], positive terms first
    static const int det_terms[24][4] = {
        { 0, 1, 2, 3 }, { 0, 2, 3, 1 }, { 0, 3, 1, 2 }, { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 3, 2, 0 },
        { 2, 0, 1, 3 }, { 2, 1, 3, 0 }, { 2, 3, 0, 1 }, { 3, 0, 2, 1 }, { 3, 1, 0, 2 }, { 3, 2, 1, 0 },
        { 0, 1, 3, 2 }, { 0, 2, 1, 3 }, { 0, 3, 2, 1 }, { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 },
        { 2, 0, 3, 1 }, { 2, 1, 0, 3 }, { 2, 3, 1, 0 }, { 3, 0, 1, 2 }, { 3, 1, 2, 0 }, { 3, 2, 0, 1 },
    };
    float det = 0.f;
    for (int t = 0; t < 24; t++)
    {
        float p = at(0, det_terms[t][0]) * at(1, det_terms[t][1]) * at(2, det_terms[t][2]) * at(3, det_terms[t][3]);
        det = t == 0 ? p : det + (t < 12 ? p : -p);
    }
    const float f = 1.f / det;
    // cofactor a_rc (out[4 r + c] = a_rc * f): six triple products over the three rows other than
    // the excluded one; each entry lists (sign, column of each remaining row in row order)
    struct Cof
    {
        int rows[3];
        int terms[6][4]; // sign, col0, col1, col2
    };
    static const Cof cof[16] = {
        // a00
        { { 1, 2, 3 }, { { 1, 1, 2, 3 }, { 1, 2, 3, 1 }, { 1, 3, 1, 2 }, { -1, 1, 3, 2 }, { -1, 2, 1, 3 }, { -1, 3, 2, 1 } } },
        // a10
        { { 0, 2, 3 }, { { 1, 1, 3, 2 }, { 1, 2, 1, 3 }, { 1, 3, 2, 1 }, { -1, 1, 2, 3 }, { -1, 2, 3, 1 }, { -1, 3, 1, 2 } } },
        // a20
        { { 0, 1, 3 }, { { 1, 1, 2, 3 }, { 1, 2, 3, 1 }, { 1, 3, 1, 2 }, { -1, 1, 3, 2 }, { -1, 2, 1, 3 }, { -1, 3, 2, 1 } } },
        // a30
        { { 0, 1, 2 }, { { 1, 1, 3, 2 }, { 1, 2, 1, 3 }, { 1, 3, 2, 1 }, { -1, 1, 2, 3 }, { -1, 2, 3, 1 }, { -1, 3, 1, 2 } } },
        // a01
        { { 1, 2, 3 }, { { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 3, 2, 0 }, { -1, 0, 2, 3 }, { -1, 2, 3, 0 }, { -1, 3, 0, 2 } } },
        // a11
        { { 0, 2, 3 }, { { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 }, { -1, 0, 3, 2 }, { -1, 2, 0, 3 }, { -1, 3, 2, 0 } } },
        // a21
        { { 0, 1, 3 }, { { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 3, 2, 0 }, { -1, 0, 2, 3 }, { -1, 2, 3, 0 }, { -1, 3, 0, 2 } } },
        // a31
        { { 0, 1, 2 }, { { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 }, { -1, 0, 3, 2 }, { -1, 2, 0, 3 }, { -1, 3, 2, 0 } } },
        // a02
        { { 1, 2, 3 }, { { 1, 0, 1, 3 }, { 1, 1, 3, 0 }, { 1, 3, 0, 1 }, { -1, 0, 3, 1 }, { -1, 1, 0, 3 }, { -1, 3, 1, 0 } } },
        // a12
        { { 0, 2, 3 }, { { -1, 0, 1, 3 }, { -1, 1, 3, 0 }, { -1, 3, 0, 1 }, { 1, 0, 3, 1 }, { 1, 1, 0, 3 }, { 1, 3, 1, 0 } } },
        // a22
        { { 0, 1, 3 }, { { 1, 0, 1, 3 }, { 1, 1, 3, 0 }, { 1, 3, 0, 1 }, { -1, 0, 3, 1 }, { -1, 1, 0, 3 }, { -1, 3, 1, 0 } } },
        // a32
        { { 0, 1, 2 }, { { 1, 0, 3, 1 }, { 1, 1, 0, 3 }, { 1, 3, 1, 0 }, { -1, 1, 3, 0 }, { -1, 3, 0, 1 }, { -1, 0, 1, 3 } } },
        // a03
        { { 1, 2, 3 }, { { 1, 0, 2, 1 }, { 1, 1, 0, 2 }, { 1, 2, 1, 0 }, { -1, 0, 1, 2 }, { -1, 1, 2, 0 }, { -1, 2, 0, 1 } } },
        // a13
        { { 0, 2, 3 }, { { 1, 0, 1, 2 }, { 1, 1, 2, 0 }, { 1, 2, 0, 1 }, { -1, 0, 2, 1 }, { -1, 1, 0, 2 }, { -1, 2, 1, 0 } } },
        // a23
        { { 0, 1, 3 }, { { 1, 0, 2, 1 }, { 1, 1, 0, 2 }, { 1, 2, 1, 0 }, { -1, 0, 1, 2 }, { -1, 1, 2, 0 }, { -1, 2, 0, 1 } } },
        // a33
        { { 0, 1, 2 }, { { 1, 0, 1, 2 }, { 1, 1, 2, 0 }, { 1, 2, 0, 1 }, { -1, 0, 2, 1 }, { -1, 1, 0, 2 }, { -1, 2, 1, 0 } } },
    };
    // output order of the reference's return value: a00 a10 a20 a30, a01 a11 a21 a31, ...
    for (int e = 0; e < 16; e++)
    {
        const Cof& c = cof[e];
        float s = 0.f;
        for (int t = 0; t < 6; t++)
        {
            float p = at(c.rows[0], c.terms[t][1]);
            p = (c.terms[t][0] < 0 ? -p : p) * at(c.rows[1], c.terms[t][2]) * at(c.rows[2], c.terms[t][3]);
            s = t == 0 ? p : s + p;
        }
        out[e] = s * f;
    }
}

int voxelbvh_ijkl(Context* ctx, const vdbr_gaussians_t* gs, uint32_t resolution, const float* transform,
                  uint64_t* ijkl, uint32_t* prim_ids, uint32_t* ranges, float* world_bbox)
{
    if (!gs)
        return ctx->fail("voxelbvh_ijkl: gaussians is NULL");
    if (resolution < 1u || resolution > 4096u)
        return ctx->fail("voxelbvh_ijkl: resolution must be in [1, 4096]");
    const uint64_t n = gs->count;
    if (n == 0)
        return 0;
    if (n * 8u >= 0xFFFFFFFFull)
        return ctx->fail("voxelbvh_ijkl: more than (2^32-1)/8 Gaussians");
    if (!gs->means || !gs->opacities || !gs->quats || !gs->scales || !ijkl || !prim_ids || !ranges || !world_bbox)
        return ctx->fail("voxelbvh_ijkl: null array");
    BvhTransform xf;
    for (int e = 0; e < 16; e++)
        xf.m[e] = xf.inv[e] = (e % 5 == 0) ? 1.f : 0.f;
    if (transform)
    {
        for (int e = 0; e < 16; e++)
            xf.m[e] = transform[e];
        mat_inverse(xf.m, xf.inv);
    }
    hipStream_t st = ctx->stream;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    const uint64_t count = 8 * n;
    float* partial = (float*)ctx->ws(WS_BVH_PARTIAL, (size_t)blocks * 6 * sizeof(float));
    uint32_t* starts = (uint32_t*)ctx->ws(WS_BVH_STARTS, (count + 1) * 2 * sizeof(uint32_t));
    if (!partial || !starts)
        return -1;
    uint32_t* scan = starts + (count + 1);
    uint32_t* flag = starts + count; // one word past the flags: the compact-domain flag
    hipLaunchKernelGGL(k_bvh_bbox1, dim3(blocks), dim3(256), 0, st, *gs, (uint32_t)n, xf, partial);
    hipLaunchKernelGGL(k_bvh_bbox2, dim3(1), dim3(256), 0, st, (const float*)partial, blocks, world_bbox);
    VDBR_CHECK(ctx, hipMemsetAsync(flag, 0, 4, st));
    hipLaunchKernelGGL(k_bvh_keys, dim3(blocks), dim3(256), 0, st, *gs, (uint32_t)n, resolution, xf,
                       (const float*)world_bbox, ijkl, prim_ids, flag);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, flag, 4, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint32_t flags = (uint32_t)ctx->pinned[0];
    if (flags & 2u)
        return ctx->fail("voxelbvh_ijkl: non-finite Gaussian parameters (a bounding box outside the world box)");
    const bool packed = flags == 0u;
    const uint32_t eblocks = (uint32_t)std::min<uint64_t>((count + 255) / 256, 16384);
    if (packed)
    {
        hipLaunchKernelGGL(k_bvh_pack, dim3(eblocks), dim3(256), 0, st, ijkl, count);
        if (radix_sort_u64(ctx, ijkl, prim_ids, count, 0, 40, nullptr))
            return -1;
    }
    else if (radix_sort_u64(ctx, ijkl, prim_ids, count, 0, 64, nullptr))
        return -1;
    hipLaunchKernelGGL(k_bvh_starts, dim3(eblocks), dim3(256), 0, st, (const uint64_t*)ijkl, count, starts);
    if (packed)
        hipLaunchKernelGGL(k_bvh_unpack, dim3(eblocks), dim3(256), 0, st, ijkl, count);
    VDBR_CHECK(ctx, hipGetLastError());
    if (scan_u32(ctx, starts, scan, count, true, nullptr))
        return -1;
    VDBR_CHECK(ctx, hipMemsetAsync(ranges, 0, count * 2 * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_bvh_headers, dim3(eblocks), dim3(256), 0, st, (const uint64_t*)ijkl, (const uint32_t*)starts,
                       (const uint32_t*)scan, count, ranges);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    return 0;
}

} // namespace vdbr
