// vdb_device.h -- device-side records and arithmetic of the raster3d pipeline (gfx950, wave64).
//
// Every FP32 expression that decides topology or the RGBA8 value is written in the operation
// order of the reference shaders (gaussian_prim.slang, gaussian_frag_alloc.slang,
// gaussian_frag_color.slang); the library is compiled with -ffp-contract=off, divisions and
// square roots are IEEE correctly rounded, and exp/log/pow come from vdb_fpmath.h.
#ifndef VDB_DEVICE_H
#define VDB_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vdb_fpmath.h"
#include "vdb_keylayout.h"
#include "vdb_layout.h"

#define VDB_DEV __device__ __forceinline__

// ----------------------------------------------------------------------------- per-Gaussian
// Predicate frame of one Gaussian (gaussian_frag_alloc.slang:62-121): position, the rows of the
// inverse rotation, 1/r and 1/s (s floored to voxel_size/8).
struct __align__(16) PrimFrame
{
    float mu[3];
    float inv_r;
    float rinv[9]; // rot0_inv, rot1_inv, rot2_inv (x, y, z each)
    float inv_s[3];
};

// Voxel AABB of one Gaussian (gaussian_prim.slang:70-139); rng = max + 1 - min, 0 when dropped.
struct __align__(16) PrimBox
{
    int32_t mn[3];
    int32_t rng_x;
    int32_t rng_y;
    int32_t rng_z;
    uint32_t pad0;
    uint32_t pad1;
};

// Value-pass constants of one Gaussian (gaussian_frag_color.slang:184-212).
struct __align__(16) PrimVal
{
    float opacity;
    float local_scale; // r
    float pow_exp;     // 1 / voxelRelativeScale
    uint32_t rgb_cull; // r8 | g8 << 8 | b8 << 16 | culled << 24
};

// Compact voxel key: leaf coordinates relative to the global leaf minimum, packed
// lexicographically (x, y, z) with per-axis bit widths, followed by the 9-bit in-leaf index
// n = (i&7)<<6 | (j&7)<<3 | (k&7). Voxels of one leaf are contiguous and in leaf order.
// Pair key = voxel << (gid_bits + alpha_bits) | gid << alpha_bits | alpha8, alpha_bits 8 when
// the RGBA8 alpha of the pair travels in the key (0 otherwise).

// Correctly rounded FP32 division and square root, independent of compiler flags: the exact
// result of a float operation rounded to double (53 bits) and then to float (24 bits) equals the
// direct float rounding because 53 >= 2*24 + 2, and f64 division / square root are IEEE
// correctly rounded on gfx950.
VDB_DEV float fdiv_rn(float a, float b)
{
    return (float)((double)a / (double)b);
}
VDB_DEV float fsqrt_rn(float a)
{
    return (float)__builtin_sqrt((double)a);
}

VDB_DEV float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    float s = ax * bx;
    s = s + ay * by;
    s = s + az * bz;
    return s;
}

VDB_DEV uint64_t encode_voxel(const KeyLayout& kl, int32_t i, int32_t j, int32_t k)
{
    uint64_t lx = (uint32_t)((i >> 3) - kl.lmin[0]);
    uint64_t ly = (uint32_t)((j >> 3) - kl.lmin[1]);
    uint64_t lz = (uint32_t)((k >> 3) - kl.lmin[2]);
    uint64_t leaf = (lx << (kl.bits[1] + kl.bits[2])) | (ly << kl.bits[2]) | lz;
    uint32_t n = ((uint32_t)(i & 7) << 6) | ((uint32_t)(j & 7) << 3) | (uint32_t)(k & 7);
    return (leaf << 9) | n;
}

VDB_DEV void decode_voxel(const KeyLayout& kl, uint64_t v, int32_t* i, int32_t* j, int32_t* k)
{
    uint32_t n = (uint32_t)(v & 511u);
    uint64_t leaf = v >> 9;
    uint32_t lz = (uint32_t)(leaf & ((1ull << kl.bits[2]) - 1ull));
    uint32_t ly = (uint32_t)((leaf >> kl.bits[2]) & ((1ull << kl.bits[1]) - 1ull));
    uint32_t lx = (uint32_t)(leaf >> (kl.bits[1] + kl.bits[2]));
    *i = ((int32_t)lx + kl.lmin[0]) * 8 + (int32_t)(n >> 6);
    *j = ((int32_t)ly + kl.lmin[1]) * 8 + (int32_t)((n >> 3) & 7u);
    *k = ((int32_t)lz + kl.lmin[2]) * 8 + (int32_t)(n & 7u);
}

// leaf index of a pair key's leaf field
VDB_DEV uint64_t leaf_of_field(const KeyLayout& kl, uint64_t field)
{
    return kl.rank_bits ? (uint64_t)kl.leaf_of_rank[field] : field;
}

// voxel coordinates of a pair key's voxel part (leaf field << 9 | slot)
VDB_DEV void decode_pair_voxel(const KeyLayout& kl, uint64_t v, int32_t* i, int32_t* j, int32_t* k)
{
    decode_voxel(kl, (leaf_of_field(kl, v >> 9) << 9) | (v & 511u), i, j, k);
}

// gaussian_frag_alloc.slang:79-121 -- AABB of the voxel's 8 corners in the Gaussian frame,
// each corner computed independently in the reference operation order.
VDB_DEV void voxel_box(const PrimFrame& f, int32_t i, int32_t j, int32_t k, float vs, float vmin[3], float vmax[3])
{
#pragma unroll
    for (uint32_t v = 0; v < 8u; v++)
    {
        float px = (float)i, py = (float)j, pz = (float)k;
        if (v & 1u)
            px += 1.f;
        if (v & 2u)
            py += 1.f;
        if (v & 4u)
            pz += 1.f;
        float lx = px * vs;
        float ly = py * vs;
        float lz = pz * vs;
        lx = lx - f.mu[0];
        ly = ly - f.mu[1];
        lz = lz - f.mu[2];
        float q0 = dot3(lx, ly, lz, f.rinv[0], f.rinv[1], f.rinv[2]);
        float q1 = dot3(lx, ly, lz, f.rinv[3], f.rinv[4], f.rinv[5]);
        float q2 = dot3(lx, ly, lz, f.rinv[6], f.rinv[7], f.rinv[8]);
        q0 = q0 * f.inv_r;
        q1 = q1 * f.inv_r;
        q2 = q2 * f.inv_r;
        q0 = q0 * f.inv_s[0];
        q1 = q1 * f.inv_s[1];
        q2 = q2 * f.inv_s[2];
        if (v == 0u)
        {
            vmin[0] = vmax[0] = q0;
            vmin[1] = vmax[1] = q1;
            vmin[2] = vmax[2] = q2;
        }
        else
        {
            vmin[0] = fminf(vmin[0], q0);
            vmin[1] = fminf(vmin[1], q1);
            vmin[2] = fminf(vmin[2], q2);
            vmax[0] = fmaxf(vmax[0], q0);
            vmax[1] = fmaxf(vmax[1], q1);
            vmax[2] = fmaxf(vmax[2], q2);
        }
    }
}

// gaussian_frag_alloc.slang:139-162 -- voxel kept iff !(dist2 <= 0)
VDB_DEV bool box_kept(const float vmin[3], const float vmax[3])
{
    float d2 = 1.f;
#pragma unroll
    for (int a = 0; a < 3; a++)
    {
        if (0.f < vmin[a])
            d2 = d2 - vmin[a] * vmin[a];
        else if (0.f > vmax[a])
            d2 = d2 - vmax[a] * vmax[a];
    }
    return !(d2 <= 0.f);
}

VDB_DEV uint32_t u8_of(float v)
{
    return (uint32_t)(255.f * fmaxf(0.f, fminf(1.f, v)));
}

// gaussian_frag_color.slang:184-212 -- 8-bit alpha of one surviving pair. The 64-sample
// maximum of o*exp(-(r*|p|^2)) equals o*exp(-(r*min|p|^2)) because every step is monotone in
// |p|^2 (correctly rounded exp, multiplication by positive r and o); the minimum of the
// separable sum over the 4x4x4 lattice is the sum of per-axis minima for the same reason
// (rounded addition is monotone in each operand).
// minimum over the 4x4x4 sample lattice of |lerp(vmin, vmax, w)|^2 (gaussian_frag_color.slang:192-199)
// The end samples are exact: w = 0 gives 0*vmax + 1*vmin = vmin and w = fl(3 fl(1/3)) = 1 gives
// vmax + 0*vmin = vmax for finite box bounds, so only the two interior samples are formed.
VDB_DEV float min_sample_dist2(const float vmin[3], const float vmax[3])
{
    const float third = 1.f / 3.f;
    const float w1 = third * 1.f, ow1 = 1.f - w1;
    const float w2 = third * 2.f, ow2 = 1.f - w2;
    float msq[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        // the reference's minimum of the 4 rounded squares: fl(s * s) depends on |s| only and is
        // monotone in it, so that minimum is the rounded square of the sample of least magnitude
        float l1 = w1 * vmax[c] + ow1 * vmin[c];
        float l2 = w2 * vmax[c] + ow2 * vmin[c];
        // all four are finite: the plain min instructions with |.| operands give the least magnitude
        // (fminf would first quiet signalling NaNs, one more instruction per operand from memory)
        float m;
        asm("v_min3_f32 %0, |%1|, |%2|, |%3|" : "=v"(m) : "v"(vmin[c]), "v"(l1), "v"(l2));
        asm("v_min_f32_e64 %0, %1, |%2|" : "=v"(m) : "v"(m), "v"(vmax[c]));
        msq[c] = m * m;
    }
    float d = msq[0] + msq[1];
    d = d + msq[2];
    return d;
}

// alpha from the minimum sample distance (gaussian_frag_color.slang:197-212)
VDB_DEV uint32_t alpha8_from_dist2(const PrimVal& pv, float d)
{
    if (pv.rgb_cull >> 24)
        return 0u; // culled: alpha forced to 0 (the colour channels are per Gaussian)
    float e = pv.opacity * vdb_expf(-(pv.local_scale * d));
    float a = fmaxf(0.f, e);
    a = 1.f - vdb_powf(1.f - a, pv.pow_exp);
    return u8_of(a);
}

// Same result as alpha8_from_dist2, decided cheaply when possible. The reference chain is
//   x = -(r d), t = exp(x), e = o t, a = max(0, e), b = 1 - a, p = pow(b, k), a2 = 1 - p,
//   y = 255 clamp01(a2), alpha = trunc(y)
// with exp and pow correctly rounded. Here exp and pow come from the hardware base-2 exp / log
// (v_exp_f32, v_log_f32; 1 ulp each by the ISA, assumed 4 ulp) and every other step is the same
// FP32 operation. |y_ref - y| is bounded by propagating the exp / pow errors and one rounding
// difference per step (u = 2^-24 relative; dp/db = k p / b); the bound is doubled. When y lies
// farther than the bound from an integer, trunc(y) = trunc(y_ref); otherwise (and for any
// non-finite intermediate) the correctly rounded path decides.
VDB_DEV bool alpha8_try_fast(const PrimVal& pv, float d, uint32_t* out)
{
    if (pv.rgb_cull >> 24)
    {
        *out = 0u;
        return true;
    }
    const float u = 5.9604644775390625e-08f;
    const float x = -(pv.local_scale * d);
    const float t = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    const float e = pv.opacity * t;
    const float a = fmaxf(0.f, e);
    const float b = 1.f - a;
    const float m = pv.pow_exp * __builtin_amdgcn_logf(b);
    const float p = __builtin_amdgcn_exp2f(m);
    const float a2 = 1.f - p; // <= 1 (p >= 0): only the lower clamp can act
    const float y = 255.f * fmaxf(0.f, a2);
    // In units of u: eps_t = (|x| + 8) u is the relative error of t, db = (eps_t + 3u) a + 2u =
    // u ((|x| + 11) a + 2) bounds |b_ref - b|, eps_p = (4 |m| + 8) u the relative error of p at
    // fixed b, and the bound on |y_ref - y| is 255 (k (p / b) db + eps_p p + 2u) + 3.1e-5, doubled.
    // The bound's own roundings (fused here) are relative errors of a few u, far inside the
    // factor 2. p / b through the hardware reciprocal (1 ulp) likewise; b = 0 or a denormal b
    // gives a non-finite or huge bound, which leaves the decision to the correctly rounded path as
    // the exact quotient would.
    const float dbu = __builtin_fmaf(fabsf(x) + 11.f, a, 2.f);
    const float pbu = __builtin_fmaf(__builtin_fmaf(4.f, fabsf(m), 8.f), p, 2.f);
    const float inner = __builtin_fmaf(pv.pow_exp * (p * __builtin_amdgcn_rcpf(b)), dbu, pbu);
    const float bound = __builtin_fmaf(inner, 2.f * 255.f * u, 2.f * 3.1e-5f + 1e-6f);
    const float fl = floorf(y);
    if ((y - fl > bound) & ((fl + 1.f) - y > bound)) // both sides in one branch
    {
        *out = (uint32_t)fl;
        return true;
    }
    return false;
}

VDB_DEV uint32_t alpha8_fast(const PrimVal& pv, float d)
{
    uint32_t a8;
    if (alpha8_try_fast(pv, d, &a8))
        return a8;
    return alpha8_from_dist2(pv, d);
}

// ----------------------------------------------------------------------------- wave helpers
VDB_DEV uint32_t lane_id()
{
    return __lane_id();
}

// DPP lane moves (GFX9 data-parallel primitives; no LDS traffic). Lanes whose source is outside
// the row, or whose row is masked off, receive `old`.
template <int CTRL, int ROW_MASK>
VDB_DEV uint32_t dpp_u32(uint32_t old, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, ROW_MASK, 0xF, false);
}

// value of lane - 1 across the whole wave (lane 0 receives old)
VDB_DEV uint64_t wave_shr1_u64(uint64_t old, uint64_t x)
{
    uint32_t lo = dpp_u32<0x138, 0xF>((uint32_t)old, (uint32_t)x);
    uint32_t hi = dpp_u32<0x138, 0xF>((uint32_t)(old >> 32), (uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

VDB_DEV uint64_t readlane_u64(uint64_t x, int lane)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// exclusive add scan over the wave; *total receives the wave sum
VDB_DEV uint32_t wave_excl_add(uint32_t x, uint32_t lane, uint32_t* total)
{
    // inclusive add scan over the wave (DPP), then shifted to exclusive
    uint32_t y;
    y = dpp_u32<0x111, 0xF>(0u, x);
    x += y;
    y = dpp_u32<0x112, 0xF>(0u, x);
    x += y;
    y = dpp_u32<0x114, 0xF>(0u, x);
    x += y;
    y = dpp_u32<0x118, 0xF>(0u, x);
    x += y;
    y = dpp_u32<0x142, 0xA>(0u, x);
    x += y;
    y = dpp_u32<0x143, 0xC>(0u, x);
    x += y;
    *total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    uint32_t ex = dpp_u32<0x138, 0xF>(0u, x);
    return lane == 0u ? 0u : ex;
}

// Inclusive segmented scan over the wave with an idempotent-or-associative op: lane L combines
// the values of lanes [seg_start, L] (seg_start <= L is the first lane of L's run). Hillis-Steele
// inside each row of 16 lanes (row_shr 1, 2, 4, 8), then the last lane of rows 0 / 2 into rows
// 1 / 3 (row_bcast:15) and lane 31 into rows 2 and 3 (row_bcast:31).
template <typename Op>
VDB_DEV uint32_t seg_scan_u32(uint32_t x, uint32_t identity, uint32_t lane, uint32_t seg_start, Op op)
{
    uint32_t y;
    y = dpp_u32<0x111, 0xF>(identity, x);
    if (lane >= seg_start + 1u)
        x = op(x, y);
    y = dpp_u32<0x112, 0xF>(identity, x);
    if (lane >= seg_start + 2u)
        x = op(x, y);
    y = dpp_u32<0x114, 0xF>(identity, x);
    if (lane >= seg_start + 4u)
        x = op(x, y);
    y = dpp_u32<0x118, 0xF>(identity, x);
    if (lane >= seg_start + 8u)
        x = op(x, y);
    y = dpp_u32<0x142, 0xA>(identity, x);
    if ((lane & 16u) && seg_start < (lane & ~15u))
        x = op(x, y);
    y = dpp_u32<0x143, 0xC>(identity, x);
    if (lane >= 32u && seg_start < 32u)
        x = op(x, y);
    return x;
}

#endif
