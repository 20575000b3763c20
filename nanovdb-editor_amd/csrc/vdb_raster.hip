// vdb_raster.hip -- the raster3d pipeline on gfx950: Gaussian arrays -> node2 NanoVDB grid.
//
//   preprocess  (RasterUtils.cpp:31-136)          N threads, in place
//   setup       (gaussian_prim.slang:25-145)      per-Gaussian frame/AABB/footprint/row count
//   emit        (gaussian_frag_alloc.slang:28-177) one thread per AABB row (i, j); the row's
//               8-corner transforms are shared along k (each lattice point computed once, in
//               the reference operation order), survivors compacted with wave ballots into an
//               LDS stage and flushed in 4 KB bursts -> 64-bit keys (voxel << gid_bits | gid)
//   sort        stable onesweep radix sort of the keys (vdb_prims.hip)
//   reduce      one thread per voxel run: RGBA8 max (gaussian_frag_color.slang:184-247) or
//               density terms in double summed exactly in fixed point (any pair order); no float atomics
//   build       leaves/lowers/uppers from the sorted voxel keys (ballot-free: runs of key
//               prefixes), node order of the reference's windowed allocation
//               (grid_build_scatter_new.slang:110-174, Raster.cpp:441-535), packed prefix
//               sums, node2 write (GridBuild.cpp:118-187, grid_build_blindmetadata.slang,
//               grid_build_compute_bbox.slang)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "vdb_device.h"
#include "vdb_internal.h"
#include "vdb_node2.h"

namespace vdbr
{

// ============================================================================ block reductions
// min/max of 6 ints (xyz min, xyz max) over a 256-thread block, committed with one set of
// atomics per block: per-wave atomics on one cache line serialise at the L2.
__device__ __forceinline__ void commit_bounds(int32_t mn[3], int32_t mx[3], int32_t* bounds)
{
    __shared__ int32_t sb[4][6];
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; c++)
        {
            mn[c] = min(mn[c], __shfl_xor(mn[c], off, 64));
            mx[c] = max(mx[c], __shfl_xor(mx[c], off, 64));
        }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0u)
        for (int c = 0; c < 3; c++)
        {
            sb[wave][c] = mn[c];
            sb[wave][3 + c] = mx[c];
        }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (uint32_t w = 1; w < (blockDim.x >> 6); w++)
            for (int c = 0; c < 3; c++)
            {
                sb[0][c] = min(sb[0][c], sb[w][c]);
                sb[0][3 + c] = max(sb[0][3 + c], sb[w][3 + c]);
            }
        if (sb[0][0] <= sb[0][3])
        {
            for (int c = 0; c < 3; c++)
            {
                atomicMin(&bounds[c], sb[0][c]);
                atomicMax(&bounds[3 + c], sb[0][3 + c]);
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void commit_sum(unsigned long long v, unsigned long long* target)
{
    __shared__ unsigned long long ss[4];
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0u)
        ss[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++)
            t += ss[w];
        if (t)
            atomicAdd(target, t);
    }
    __syncthreads();
}

// ============================================================================ preprocess
// RasterUtils.cpp:59-117 (quaternion normalisation, exp scales, SH0 colour, sigmoid)
__global__ __launch_bounds__(256) void k_preprocess(uint64_t n, float* __restrict__ opac, float* __restrict__ quats,
                                                    float* __restrict__ scales, const float* sh0, float* colors)
{
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    float x = quats[4 * g + 1], y = quats[4 * g + 2], z = quats[4 * g + 3], w = quats[4 * g + 0];
    float sum = x * x;
    sum = sum + y * y;
    sum = sum + z * z;
    sum = sum + w * w;
    float magn_inv = fdiv_rn(1.f, fsqrt_rn(sum));
    quats[4 * g + 1] = x * magn_inv;
    quats[4 * g + 2] = y * magn_inv;
    quats[4 * g + 3] = z * magn_inv;
    quats[4 * g + 0] = w * magn_inv;
    for (int a = 0; a < 3; a++)
        scales[3 * g + a] = vdb_expf(scales[3 * g + a]);
    const float c0 = 0.28209479177387814f;
    for (int a = 0; a < 3; a++)
    {
        float s = sh0[3 * g + a];
        float c = c0 * s;
        colors[3 * g + a] = c + 0.5f;
    }
    float o = opac[g];
    opac[g] = fdiv_rn(1.f, 1.f + vdb_expf(-o));
}

// ============================================================================ setup
struct SetupOut
{
    PrimFrame* frames;
    PrimBox* boxes;
    PrimVal* vals;
    uint64_t* foot;
    uint64_t* rows;
    int32_t* bounds;             // [6] min xyz, max xyz over valid Gaussian AABBs
    unsigned long long* est;     // pair-count bound accumulator
};

__device__ __forceinline__ void setup_one(uint64_t g, const float* __restrict__ means, const float* __restrict__ quats,
                                          const float* __restrict__ scales, const float* __restrict__ opac,
                                          const float* __restrict__ colors, float vs, float vs_inv, float cull_scale,
                                          const SetupOut& o, int32_t bmn[3], int32_t bmx[3],
                                          unsigned long long* est_sum)
{
    float pos[3] = { means[3 * g + 0], means[3 * g + 1], means[3 * g + 2] };
    float x = quats[4 * g + 1], y = quats[4 * g + 2], z = quats[4 * g + 3], w = quats[4 * g + 0];
    float opacity = opac[g];
    // gaussian_prim.slang:65-68
    float local_scale = fsqrt_rn(2.f * vdb_logf(fdiv_rn(opacity, 0.01f)));
    const float min_scale = 1.f / 8.f;
    float floor_s = min_scale * vs;
    float s[3];
    for (int a = 0; a < 3; a++)
    {
        float v = scales[3 * g + a];
        s[a] = fmaxf(v, floor_s);
    }
    // drop Gaussians whose radius is not a positive finite number (opacity <= 0.01)
    bool valid = (local_scale > 0.f) && (local_scale <= 3.402823466e38f);

    int32_t mn[3] = { 0, 0, 0 }, mx[3] = { -1, -1, -1 };
    if (valid)
    {
        // gaussian_prim.slang:70-111
        float r0x = 1.f - 2.f * (y * y + z * z), r0y = 2.f * (x * y - z * w), r0z = 2.f * (x * z + y * w);
        float r1x = 2.f * (x * y + z * w), r1y = 1.f - 2.f * (x * x + z * z), r1z = 2.f * (y * z - x * w);
        float r2x = 2.f * (x * z - y * w), r2y = 2.f * (y * z + x * w), r2z = 1.f - 2.f * (x * x + y * y);
        for (uint32_t v = 0; v < 8u; v++)
        {
            float c[3] = { (v & 1u) ? 1.f : -1.f, (v & 2u) ? 1.f : -1.f, (v & 4u) ? 1.f : -1.f };
            for (int a = 0; a < 3; a++)
            {
                c[a] = c[a] * s[a];
                c[a] = c[a] * local_scale;
            }
            float q[3] = { dot3(c[0], c[1], c[2], r0x, r0y, r0z), dot3(c[0], c[1], c[2], r1x, r1y, r1z),
                           dot3(c[0], c[1], c[2], r2x, r2y, r2z) };
            for (int a = 0; a < 3; a++)
            {
                q[a] = q[a] + pos[a];
                int32_t ijk = (int32_t)floorf(q[a] * vs_inv);
                if (v == 0u)
                {
                    mn[a] = ijk;
                    mx[a] = ijk;
                }
                else
                {
                    mn[a] = min(mn[a], ijk);
                    mx[a] = max(mx[a], ijk);
                }
            }
        }
    }
    PrimBox box;
    box.mn[0] = mn[0];
    box.mn[1] = mn[1];
    box.mn[2] = mn[2];
    box.rng_x = valid ? mx[0] + 1 - mn[0] : 0;
    box.rng_y = valid ? mx[1] + 1 - mn[1] : 0;
    box.rng_z = valid ? mx[2] + 1 - mn[2] : 0;
    box.pad0 = 0u;
    box.pad1 = 0u;
    o.boxes[g] = box;
    uint64_t rows = valid ? (uint64_t)box.rng_x * (uint64_t)box.rng_y : 0ull;
    o.rows[g] = rows;
    o.foot[g] = rows * (uint64_t)(valid ? box.rng_z : 0);

    // gaussian_frag_alloc.slang:103-120 quantities
    PrimFrame f;
    f.mu[0] = pos[0];
    f.mu[1] = pos[1];
    f.mu[2] = pos[2];
    f.inv_r = valid ? fdiv_rn(1.f, local_scale) : 0.f;
    f.rinv[0] = 1.f - 2.f * (y * y + z * z);
    f.rinv[1] = 2.f * (x * y + z * w);
    f.rinv[2] = 2.f * (x * z - y * w);
    f.rinv[3] = 2.f * (x * y - z * w);
    f.rinv[4] = 1.f - 2.f * (x * x + z * z);
    f.rinv[5] = 2.f * (y * z + x * w);
    f.rinv[6] = 2.f * (x * z + y * w);
    f.rinv[7] = 2.f * (y * z - x * w);
    f.rinv[8] = 1.f - 2.f * (x * x + y * y);
    for (int a = 0; a < 3; a++)
        f.inv_s[a] = fdiv_rn(1.f, s[a]);
    o.frames[g] = f;

    // gaussian_frag_color.slang:184-212 per-Gaussian constants
    PrimVal pv;
    pv.opacity = opacity;
    pv.local_scale = local_scale;
    float scale1d = fsqrt_rn(dot3(s[0], s[1], s[2], s[0], s[1], s[2]));
    float vrs = fdiv_rn(0.1f * scale1d, vs);
    pv.pow_exp = fdiv_rn(1.f, vrs);
    const float default_cull = 0.05f;
    float cs = cull_scale == 0.f ? default_cull : cull_scale;
    uint32_t culled = (scale1d > cs) ? 1u : 0u;
    pv.rgb_cull = u8_of(colors[3 * g + 0]) | (u8_of(colors[3 * g + 1]) << 8) | (u8_of(colors[3 * g + 2]) << 16) |
                  (culled << 24);
    o.vals[g] = pv;

    if (valid)
    {
        for (int c = 0; c < 3; c++)
        {
            bmn[c] = min(bmn[c], mn[c]);
            bmx[c] = max(bmx[c], mx[c]);
        }
        // bound on survivors: voxel centres of kept voxels satisfy |A(c - mu)| <= 1 + |h|, h the
        // local half extent of a voxel; volume (4/3) pi (1+|h|)^3 r^3 s_x s_y s_z / vs^3
        float h2 = 0.f;
        for (int a = 0; a < 3; a++)
        {
            float ha = 0.5f * vs * (fabsf(f.rinv[3 * a]) + fabsf(f.rinv[3 * a + 1]) + fabsf(f.rinv[3 * a + 2])) * f.inv_r *
                       f.inv_s[a];
            h2 += ha * ha;
        }
        double rad = 1.0 + sqrt((double)h2);
        double vol = 4.18879020478639 * rad * rad * rad * (double)local_scale * local_scale * local_scale *
                     ((double)s[0] / vs) * ((double)s[1] / vs) * ((double)s[2] / vs);
        double est = 1.25 * vol + 16.0;
        double foot = (double)o.foot[g];
        if (est > foot)
            est = foot;
        *est_sum += (unsigned long long)est;
    }
}

constexpr uint32_t SETUP_BLOCKS = 2048;
__global__ __launch_bounds__(256) void k_setup(uint64_t n, const float* __restrict__ means, const float* __restrict__ quats,
                                               const float* __restrict__ scales, const float* __restrict__ opac,
                                               const float* __restrict__ colors, float vs, float vs_inv,
                                               float cull_scale, SetupOut o)
{
    // grid-stride on a bounded grid (SETUP_BLOCKS): every block adds its bounds and estimate with
    // atomics on one cache line, which serialize, so fewer blocks each take more Gaussians
    int32_t bmn[3] = { 2147483647, 2147483647, 2147483647 }, bmx[3] = { -2147483647, -2147483647, -2147483647 };
    unsigned long long est_sum = 0ull;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x)
        setup_one(g, means, quats, scales, opac, colors, vs, vs_inv, cull_scale, o, bmn, bmx, &est_sum);
    commit_bounds(bmn, bmx, o.bounds);
    commit_sum(est_sum, o.est);
}

// ============================================================================ emit
constexpr int EMIT_THREADS = 256;
constexpr int EMIT_ROWS = 1024;       // rows per block
constexpr int EMIT_WAVE_STAGE = 1792; // staged survivors per wave
constexpr int EMIT_HIST_PASSES = 4;   // sort passes whose digit histograms the emission can count
constexpr int EMIT_HIST_WORDS = 1024; // LDS counters: passes x 2^digit width of them (pass-major)
constexpr int EMIT_RING = 128;        // staged survivors whose alpha is pending, per wave
constexpr int EMIT_RING_WORDS = 10;   // per survivor: the box extremes (6), its Gaussian's PrimVal (4)

struct EmitArgs
{
    const PrimFrame* frames;
    const PrimBox* boxes;
    const PrimVal* vals; // alpha of each pair computed here when kl.alpha_bits == 8
    const uint64_t* row_scan; // exclusive scan of rows, n + 1 entries
    const struct RowFrame* rframes; // per-Gaussian row-interval constants (row pass)
    uint64_t n;
    uint64_t total_rows;
    float vs;
    KeyLayout kl;
    uint64_t* out;
    uint64_t cap;
    unsigned long long* counter;
    const void* segs;  // candidate segments (k_row_segments), in SEG_PARTS partitions
    uint32_t seg_compact; // 8-byte segment records (SegCompact) instead of 16-byte ones (SegWide)
    uint64_t seg_part_cap;          // capacity of one partition
    const uint64_t* seg_part_count; // segments written per partition
    const uint64_t* seg_chunk_scan; // SEG_PARTS + 1 exclusive scan of 64-segment chunks per partition
    int32_t slab_i0;                // first voxel column of the slab (INT32_MIN: no slab)
    unsigned long long* hist;       // digit histograms of the emitted keys for their sort (or null)
    uint32_t hist_shift;            // first sorted key bit
    uint32_t hist_bits;             // sorted key bits
    uint32_t hist_w;                // digit width
};

// whether the emission can count the digit histograms of a sort of `bits` key bits in digits of
// `w` bits (its LDS counters hold passes x 2^w of them)
static inline bool emit_counts_hist(uint32_t bits, uint32_t w)
{
    const uint32_t passes = (bits + w - 1u) / w;
    return passes <= (uint32_t)EMIT_HIST_PASSES && (passes << w) <= (uint32_t)EMIT_HIST_WORDS;
}

// rows of every Gaussian inside the voxel columns [i0, i1) of a slab
__global__ void k_slab_rows(const PrimBox* __restrict__ boxes, uint64_t n, int32_t i0, int32_t i1,
                            uint64_t* __restrict__ rows)
{
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x)
    {
        const PrimBox b = boxes[g];
        const int64_t lo = max((int64_t)b.mn[0], (int64_t)i0);
        const int64_t hi = min((int64_t)b.mn[0] + (int64_t)b.rng_x, (int64_t)i1);
        rows[g] = hi > lo ? (uint64_t)(hi - lo) * (uint64_t)b.rng_y : 0ull;
    }
}

// candidate fragments per leaf column (x / 8 - lmin_x), each Gaussian's footprint spread evenly
// over the columns its AABB spans: the slab planner's load estimate. Few columns receive many
// Gaussians, so each block sums into LDS first (up to SLAB_LDS_COLS columns) and adds one value
// per column to the global histogram.
constexpr uint32_t SLAB_LDS_COLS = 4096;
__global__ __launch_bounds__(256) void k_slab_hist(const PrimBox* __restrict__ boxes, const uint64_t* __restrict__ foot,
                                                   uint64_t n, int32_t lmin_x, uint32_t cols,
                                                   unsigned long long* __restrict__ hist)
{
    __shared__ unsigned long long sh[SLAB_LDS_COLS];
    const bool local = cols <= SLAB_LDS_COLS;
    if (local)
        for (uint32_t c = threadIdx.x; c < cols; c += blockDim.x)
            sh[c] = 0ull;
    __syncthreads();
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint64_t fp = foot[g];
        if (!fp)
            continue;
        const PrimBox b = boxes[g];
        const int32_t c0 = (b.mn[0] >> 3) - lmin_x;
        const int32_t c1 = ((b.mn[0] + (int32_t)b.rng_x - 1) >> 3) - lmin_x;
        const uint64_t share = fp / (uint64_t)(c1 - c0 + 1) + 1ull;
        for (int32_t c = max(c0, 0); c <= c1 && c < (int32_t)cols; c++)
        {
            if (local)
                atomicAdd(&sh[c], (unsigned long long)share);
            else
                atomicAdd(&hist[c], (unsigned long long)share);
        }
    }
    __syncthreads();
    if (local)
        for (uint32_t c = threadIdx.x; c < cols; c += blockDim.x)
            if (sh[c])
                atomicAdd(&hist[c], sh[c]);
}

// segment buffer partitions: block b reserves in partition b % SEG_PARTS (separate counters on
// separate cache lines, so reservations do not serialize on one address)
constexpr uint32_t SEG_PARTS = 64;
constexpr uint32_t SEG_COUNTER_STRIDE = 16; // u64 words between counters (128 bytes)

// A candidate segment: up to EMIT_SEG consecutive voxels (i, j, k .. k + count - 1) of one
// Gaussian's pruned row interval.
// (measured at 2M clustered: 6 -> 36.7 ms emission, 8 -> 35.1, 10 -> 33.9, 14 -> 33.5, 16 -> 33.3,
// 20 -> 33.1; longer caps let a long row hold its wave)
constexpr int32_t EMIT_SEG = 16;

// Segment records. SegCompact, when the voxel key of the segment's first voxel (leaf index and
// slot), the Gaussian id and the count fit 64 bits: voxel key << (gid_bits + 4) | gaussian id << 4
// | (count - 1). SegWide otherwise: x = gaussian id, y = i, z = j, w = (k + 2^17) | count << 20.
// The walk keeps records in flight as they are; gid() and leaf() are a few bit operations.
struct SegCompact
{
    using Rec = uint64_t;
    static __device__ __forceinline__ Rec make(const KeyLayout& kl, uint64_t g, int32_t i, int32_t j, int32_t k,
                                               int32_t cnt)
    {
        return (encode_voxel(kl, i, j, k) << (kl.gid_bits + 4u)) | (g << 4) | (uint64_t)(cnt - 1);
    }
    static __device__ __forceinline__ uint64_t gid(const KeyLayout& kl, Rec r)
    {
        return (r >> 4) & ((1ull << kl.gid_bits) - 1ull);
    }
    static __device__ __forceinline__ uint64_t vkey(const KeyLayout& kl, Rec r)
    {
        return r >> (kl.gid_bits + 4u);
    }
    static __device__ __forceinline__ void decode(const KeyLayout& kl, Rec r, int32_t* i, int32_t* j, int32_t* k,
                                                  int32_t* len)
    {
        decode_voxel(kl, vkey(kl, r), i, j, k);
        *len = (int32_t)(r & 15u) + 1;
    }
};
struct SegWide
{
    using Rec = uint4;
    static __device__ __forceinline__ Rec make(const KeyLayout&, uint64_t g, int32_t i, int32_t j, int32_t k,
                                               int32_t cnt)
    {
        return make_uint4((uint32_t)g, (uint32_t)i, (uint32_t)j, ((uint32_t)(k + 131072) & 0xFFFFFu) | ((uint32_t)cnt << 20));
    }
    static __device__ __forceinline__ uint64_t gid(const KeyLayout&, Rec r)
    {
        return r.x;
    }
    static __device__ __forceinline__ uint64_t vkey(const KeyLayout& kl, Rec r)
    {
        return encode_voxel(kl, (int32_t)r.y, (int32_t)r.z, (int32_t)(r.w & 0xFFFFFu) - 131072);
    }
    static __device__ __forceinline__ void decode(const KeyLayout&, Rec r, int32_t* i, int32_t* j, int32_t* k,
                                                  int32_t* len)
    {
        *i = (int32_t)r.y;
        *j = (int32_t)r.z;
        *k = (int32_t)(r.w & 0xFFFFFu) - 131072;
        *len = (int32_t)(r.w >> 20);
    }
};
static_assert(EMIT_SEG <= 16, "SegCompact holds count - 1 in 4 bits");

__device__ __forceinline__ uint64_t find_prim(const uint64_t* __restrict__ row_scan, uint64_t lo, uint64_t hi, uint64_t row)
{
    // largest g in [lo, hi] with row_scan[g] <= row
    while (lo < hi)
    {
        uint64_t mid = (lo + hi + 1) >> 1;
        if (row_scan[mid] <= row)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// largest g in [0, n) with scan[g] <= row for a non-decreasing scan with scan[0] = 0, by the whole
// wave (row wave-uniform): every round probes 64 evenly spaced entries of [lo, hi] and keeps the
// gap where the scan passes the target, so 2M entries take 4 rounds of one load per lane instead of
// 21 dependent loads
template <typename T>
__device__ __forceinline__ uint64_t wave_find_prim(const T* __restrict__ row_scan, uint64_t n, uint64_t row,
                                                   uint32_t lane)
{
    uint64_t lo = 0, hi = n - 1; // row_scan[lo] <= row; the answer lies in [lo, hi]
    while (hi > lo)
    {
        const uint64_t span = hi - lo;
        auto probe = [&](uint64_t l) { return lo + (span * (l + 1u) + 63u) / 64u; }; // in (lo, hi]
        const bool le = (uint64_t)row_scan[probe(lane)] <= row;
        // the probes are non-decreasing in the lane and row_scan is non-decreasing: a prefix holds
        const uint32_t c = (uint32_t)__popcll(__ballot(le));
        if (c == 0u)
            hi = probe(0) - 1u;
        else if (c == 64u)
            lo = hi;
        else
        {
            const uint64_t nlo = probe(c - 1u), nhi = probe(c) - 1u;
            lo = nlo;
            hi = nhi;
        }
    }
    return lo;
}

// orders a wave's LDS accesses before and after it (lanes exchange data through LDS)
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Alpha of the staged survivors [s0, s0 + n), n <= 64, one per lane: the minimum sample distance
// of the voxel box held in the wave's ring (slot s at s mod EMIT_RING, with the value constants
// of its Gaussian), then the 8-bit alpha ORed into the staged key. The walk defers this work so
// that it runs on 64 survivors at once; evaluated inside the walk it issues for the whole wave on
// nearly every step while only the lanes that keep their voxel (about one in seven) use it. It
// reads nothing but LDS, so it never waits on the wave's outstanding global loads and stores.
__device__ __forceinline__ void emit_alpha_batch(uint64_t* stage, const uint32_t (*ring)[EMIT_RING], uint32_t s0,
                                                 uint32_t n, uint32_t lane)
{
    if (lane < n)
    {
        const uint32_t s = s0 + lane;
        const uint32_t r = s & (uint32_t)(EMIT_RING - 1);
        const uint64_t key = stage[s];
        PrimVal pv;
        pv.opacity = __uint_as_float(ring[6][r]);
        pv.local_scale = __uint_as_float(ring[7][r]);
        pv.pow_exp = __uint_as_float(ring[8][r]);
        pv.rgb_cull = ring[9][r];
        // the ring holds each axis' two box extremes in either order; they are finite, so the
        // plain min / max instructions order them (fminf / fmaxf would first quiet signalling
        // NaNs of values loaded from memory, two more instructions per pair)
        float vmin[3], vmax[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const float ea = __uint_as_float(ring[c][r]), eb = __uint_as_float(ring[3 + c][r]);
            asm("v_min_f32 %0, %1, %2" : "=v"(vmin[c]) : "v"(ea), "v"(eb));
            asm("v_max_f32 %0, %1, %2" : "=v"(vmax[c]) : "v"(ea), "v"(eb));
        }
        stage[s] = key | (uint64_t)alpha8_fast(pv, min_sample_dist2(vmin, vmax));
    }
}

// Writes a wave's staged pairs (complete keys) at an atomically reserved position.
__device__ __forceinline__ void emit_flush_wave(const EmitArgs& a, const uint64_t* stage, uint32_t cnt, uint32_t lane,
                                                uint32_t* s_hist)
{
    if (cnt == 0u)
        return;
    if (a.hist)
    {
        // digit counts of the keys for every sort pass (the alpha bits lie below the sorted bits),
        // each lane over a contiguous slice, run-length compressed, into the block's LDS histograms.
        // The slice is read 8 keys at a time (the reads overlap instead of each waiting for the
        // previous one) and every key read once for all passes.
        const uint32_t passes = (a.hist_bits + a.hist_w - 1u) / a.hist_w;
        const uint32_t per = (cnt + 63u) >> 6;
        const uint32_t e0 = lane * per;
        const uint32_t e1 = min(cnt, e0 + per);
        uint32_t shift[EMIT_HIST_PASSES], dm[EMIT_HIST_PASSES], cur[EMIT_HIST_PASSES], run[EMIT_HIST_PASSES];
        const uint64_t k0 = e0 < e1 ? stage[e0] : 0ull;
#pragma unroll
        for (uint32_t p = 0; p < (uint32_t)EMIT_HIST_PASSES; p++)
        {
            shift[p] = a.hist_shift + a.hist_w * p;
            const uint32_t rem = p < passes ? a.hist_bits - a.hist_w * p : 0u;
            dm[p] = rem >= a.hist_w ? ((1u << a.hist_w) - 1u) : ((1u << rem) - 1u);
            cur[p] = (uint32_t)(k0 >> shift[p]) & dm[p];
            run[p] = 0u;
        }
        constexpr uint32_t U = 8;
        for (uint32_t e = e0; e < e1; e += U)
        {
            uint64_t k[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                k[u] = e + u < e1 ? stage[e + u] : 0ull;
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
            {
                if (e + u < e1)
                {
#pragma unroll
                    for (uint32_t p = 0; p < (uint32_t)EMIT_HIST_PASSES; p++)
                    {
                        const uint32_t d = (uint32_t)(k[u] >> shift[p]) & dm[p];
                        if (p < passes && d != cur[p])
                        {
                            atomicAdd(&s_hist[(p << a.hist_w) + cur[p]], run[p]);
                            cur[p] = d;
                            run[p] = 0u;
                        }
                        run[p]++;
                    }
                }
            }
        }
        if (e0 < e1)
        {
#pragma unroll
            for (uint32_t p = 0; p < (uint32_t)EMIT_HIST_PASSES; p++)
                if (p < passes)
                    atomicAdd(&s_hist[(p << a.hist_w) + cur[p]], run[p]);
        }
    }
    unsigned long long base = 0ull;
    if (lane == 0u)
        base = atomicAdd(a.counter, (unsigned long long)cnt);
    base = __shfl(base, 0, 64);
    // 4 keys per lane at a time: their LDS reads are issued together before the stores
    constexpr uint32_t V = 4;
    for (uint32_t e = lane; e < cnt; e += 64u * V)
    {
        uint64_t k[V];
#pragma unroll
        for (uint32_t u = 0; u < V; u++)
            k[u] = e + 64u * u < cnt ? stage[e + 64u * u] : 0ull;
#pragma unroll
        for (uint32_t u = 0; u < V; u++)
        {
            const uint64_t pos = base + e + 64u * u;
            if (e + 64u * u < cnt && pos < a.cap)
                a.out[pos] = k[u];
        }
    }
}

// Conservative candidate interval of one AABB row (i, j): in exact arithmetic the local-frame
// corners of voxel k form a parallelepiped centred at C(k) = C0 + k B with AABB half extent H,
// so the voxel's box is farther than |C(k)| - |H| from the origin. The float corner transforms
// differ from the exact ones by at most delta (operation-count bound, doubled), so every voxel
// with |C(k)| > 1 + |H| + delta is culled by the exact FP32 predicate. Solving
// |C0 + k B|^2 <= R^2 in double gives [ka, kb]; the interval is widened by one voxel each side
// and clipped to the row.
__device__ __forceinline__ void row_interval(const PrimFrame& f, const PrimBox& b, int32_t i, int32_t j, float vs,
                                             int32_t* k0, int32_t* len)
{
    const double vsd = (double)vs;
    double A[3][3];
    double H2 = 0.0, hmax = 0.0;
    double w = (double)f.inv_r;
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        double sc = w * (double)f.inv_s[c];
        double hs = 0.0;
#pragma unroll
        for (int bb = 0; bb < 3; bb++)
        {
            A[c][bb] = (double)f.rinv[3 * c + bb] * sc;
            hs += fabs(A[c][bb]);
        }
        double h = 0.5 * vsd * hs;
        H2 += h * h;
        hmax = fmax(hmax, sc);
    }
    // magnitude of the world-space operands of the corner transforms
    double M = fmax(fmax(fabs((double)f.mu[0]), fabs((double)f.mu[1])), fabs((double)f.mu[2])) +
               vsd * (double)max(max(b.rng_x, b.rng_y), b.rng_z) + vsd;
    const double u = 5.9604644775390625e-08; // 2^-24
    double H = sqrt(H2);
    double delta = 2.0 * (hmax * 10.4 * u * M + 2.0 * u * (1.0 + 2.0 * H)) * 1.7320508075688772 + 1e-6;
    double R = 1.0 + H + delta;
    double dx = ((double)i + 0.5) * vsd - (double)f.mu[0];
    double dy = ((double)j + 0.5) * vsd - (double)f.mu[1];
    double dz = 0.5 * vsd - (double)f.mu[2];
    double qa = 0.0, qb = 0.0, qc = -R * R;
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        double c0 = A[c][0] * dx + A[c][1] * dy + A[c][2] * dz;
        double bc = vsd * A[c][2];
        qa += bc * bc;
        qb += 2.0 * c0 * bc;
        qc += c0 * c0;
    }
    int32_t lo = *k0, hi = *k0 + *len - 1;
    if (qa > 0.0)
    {
        double disc = qb * qb - 4.0 * qa * qc;
        if (disc < 0.0)
        {
            *len = 0;
            return;
        }
        double sq = sqrt(disc);
        double r1 = (-qb - sq) / (2.0 * qa);
        double r2 = (-qb + sq) / (2.0 * qa);
        if (r1 > (double)hi + 1.0 || r2 < (double)lo - 1.0)
        {
            *len = 0;
            return;
        }
        double fl = floor(r1) - 1.0, ce = ceil(r2) + 1.0;
        if (fl > (double)lo)
            lo = (int32_t)fl;
        if (ce < (double)hi)
            hi = (int32_t)ce;
    }
    else if (qc > 0.0)
    {
        *len = 0;
        return;
    }
    *k0 = lo;
    *len = hi >= lo ? hi - lo + 1 : 0;
}

// The candidate voxels (i, j, k0 + st), st < len, of one AABB row, walked in lock step by the
// wave: steps is the wave maximum of len, and on_step(st, keep, ext_a, ext_b) runs on every lane
// for every st < steps (keep false past the lane's own interval). The per-corner operation order
// is that of gaussian_frag_alloc.slang:79-121: corner (col, kz) of axis c is
// q = ((part[c][col] + tz(kz)) * inv_r) * inv_s[c], with part the row-constant (x, y) terms and
// tz(kz) = ((kz * vs) - mu_z) * rinv[c][2]. Every step of that expression is a correctly rounded
// operation that is monotone non-decreasing in part (a sum with a fixed operand, then products
// with the positive finite factors inv_r and inv_s[c]) and in tz, and tz is monotone in kz with
// the sign of rinv[c][2]. So the minimum of axis c over a voxel's 8 corners is the corner of the
// column with the smallest part on the plane where tz is smaller, and the maximum is the corner of
// the column with the largest part on the other plane: per axis, one value is carried from the
// voxel's lower plane and one is new on its upper plane, and they are the box's extremes with the
// same bits as a min / max over all 8 corners (up to the sign of a zero, which neither the
// predicate nor the sample distances see). ext_a[c] and ext_b[c] are those two extremes in either
// order (fminf / fmaxf order them); the predicate uses med3(a, 0, b), which is the value that
// gaussian_frag_alloc.slang:139-162 subtracts squared (vmin if 0 < vmin, vmax if vmax < 0, else 0).
template <typename OnStep>
__device__ __forceinline__ void walk_row(const PrimFrame& f, int32_t i, int32_t j, int32_t k0, int32_t len,
                                         int32_t steps, float vs, OnStep&& on_step)
{
    // row-constant partial dot products of the 4 corner columns (dx, dy): their extremes per axis
    float part_lo[3] = { 0.f, 0.f, 0.f }, part_hi[3] = { 0.f, 0.f, 0.f };
    if (len > 0)
    {
#pragma unroll
        for (int col = 0; col < 4; col++)
        {
            float px = (float)i;
            float py = (float)j;
            if (col & 1)
                px += 1.f;
            if (col & 2)
                py += 1.f;
            float lx = px * vs;
            float ly = py * vs;
            lx = lx - f.mu[0];
            ly = ly - f.mu[1];
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
                float t = lx * f.rinv[3 * c + 0];
                t = t + ly * f.rinv[3 * c + 1];
                part_lo[c] = col == 0 ? t : fminf(part_lo[c], t);
                part_hi[c] = col == 0 ? t : fmaxf(part_hi[c], t);
            }
        }
    }
    // per axis: the part whose corner is carried from a voxel's lower plane (pc) and the part whose
    // corner is taken on its upper plane (pn)
    float pc[3], pn[3], rz[3], sc[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        rz[c] = f.rinv[3 * c + 2];
        sc[c] = f.inv_s[c];
        const bool up = !(rz[c] < 0.f);
        pc[c] = up ? part_lo[c] : part_hi[c];
        pn[c] = up ? part_hi[c] : part_lo[c];
    }
    const float inv_r = f.inv_r, mu_z = f.mu[2];
    float carry[3];
    {
        float lz = (float)k0 * vs;
        lz = lz - mu_z;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const float tz = lz * rz[c];
            float q = pc[c] + tz;
            q = q * inv_r;
            carry[c] = q * sc[c];
        }
    }
    for (int32_t st = 0; st < steps; st++)
    {
        float lz = (float)(k0 + st + 1) * vs;
        lz = lz - mu_z;
        float ea[3], eb[3];
        float d2 = 1.f;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const float tz = lz * rz[c];
            float qn = pn[c] + tz;
            qn = qn * inv_r;
            qn = qn * sc[c];
            float qc = pc[c] + tz;
            qc = qc * inv_r;
            qc = qc * sc[c];
            ea[c] = carry[c];
            eb[c] = qn;
            carry[c] = qc;
            const float t = __builtin_amdgcn_fmed3f(ea[c], 0.f, eb[c]);
            d2 = d2 - t * t;
        }
        const bool keep = st < len && !(d2 <= 0.f);
        on_step(st, keep, ea, eb);
    }
}

// Per-Gaussian constants of row_interval: with c0(i, j) = P + i Ai + j Aj the local-frame centre
// of voxel k = 0 of row (i, j) (the terms of row_interval's C0 regrouped), B = bc the step per
// voxel along k, R^2 the squared culling radius (1 + |H| + delta)^2, qa = |B|^2.
struct __align__(16) RowFrame
{
    double P[3], Ai[3], Aj[3], bc[3];
    double R2, qa, inv2qa, pad;
};
static_assert(sizeof(RowFrame) == 128, "one cache line per Gaussian");

__global__ __launch_bounds__(256) void k_row_frames(uint64_t n, const PrimFrame* __restrict__ frames,
                                                    const PrimBox* __restrict__ boxes, float vs,
                                                    RowFrame* __restrict__ out)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const PrimFrame f = frames[g];
    const PrimBox b = boxes[g];
    const double vsd = (double)vs;
    double A[3][3];
    double H2 = 0.0, hmax = 0.0;
    const double w = (double)f.inv_r;
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        const double sc = w * (double)f.inv_s[c];
        double hs = 0.0;
#pragma unroll
        for (int bb = 0; bb < 3; bb++)
        {
            A[c][bb] = (double)f.rinv[3 * c + bb] * sc;
            hs += fabs(A[c][bb]);
        }
        const double h = 0.5 * vsd * hs;
        H2 += h * h;
        hmax = fmax(hmax, sc);
    }
    // the same bound as row_interval: magnitude of the corner transforms' world-space operands and
    // the operation-count error bound of the FP32 predicate
    const double M = fmax(fmax(fabs((double)f.mu[0]), fabs((double)f.mu[1])), fabs((double)f.mu[2])) +
                     vsd * (double)max(max(b.rng_x, b.rng_y), b.rng_z) + vsd;
    const double u = 5.9604644775390625e-08; // 2^-24
    const double H = sqrt(H2);
    const double delta = 2.0 * (hmax * 10.4 * u * M + 2.0 * u * (1.0 + 2.0 * H)) * 1.7320508075688772 + 1e-6;
    const double R = 1.0 + H + delta;
    RowFrame r;
    const double dx0 = 0.5 * vsd - (double)f.mu[0];
    const double dy0 = 0.5 * vsd - (double)f.mu[1];
    const double dz = 0.5 * vsd - (double)f.mu[2];
    double qa = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        r.P[c] = A[c][0] * dx0 + A[c][1] * dy0 + A[c][2] * dz;
        r.Ai[c] = A[c][0] * vsd;
        r.Aj[c] = A[c][1] * vsd;
        r.bc[c] = vsd * A[c][2];
        qa += r.bc[c] * r.bc[c];
    }
    r.R2 = R * R;
    r.qa = qa;
    r.inv2qa = qa > 0.0 ? 0.5 / qa : 0.0;
    r.pad = 0.0;
    out[g] = r;
}

// row_interval from the Gaussian's RowFrame: the quadratic |c0 + k B|^2 <= R^2 in k solved in
// double, the roots widened by one voxel each side and clipped to the row. The regrouped terms
// round differently from row_interval by a few ulp of double, far inside that widening and the
// error bound in R, so every voxel the FP32 predicate keeps stays a candidate.
__device__ __forceinline__ void row_interval_frame(const RowFrame& rf, int32_t i, int32_t j, int32_t* k0, int32_t* len)
{
    const double di = (double)i, dj = (double)j;
    double qb = 0.0, qc = -rf.R2;
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        const double c0 = rf.P[c] + di * rf.Ai[c] + dj * rf.Aj[c];
        qb += 2.0 * c0 * rf.bc[c];
        qc += c0 * c0;
    }
    int32_t lo = *k0, hi = *k0 + *len - 1;
    if (rf.qa > 0.0)
    {
        const double disc = qb * qb - 4.0 * rf.qa * qc;
        if (disc < 0.0)
        {
            *len = 0;
            return;
        }
        const double sq = sqrt(disc);
        const double r1 = (-qb - sq) * rf.inv2qa;
        const double r2 = (-qb + sq) * rf.inv2qa;
        if (r1 > (double)hi + 1.0 || r2 < (double)lo - 1.0)
        {
            *len = 0;
            return;
        }
        const double fl = floor(r1) - 1.0, ce = ceil(r2) + 1.0;
        if (fl > (double)lo)
            lo = (int32_t)fl;
        if (ce < (double)hi)
            hi = (int32_t)ce;
    }
    else if (qc > 0.0)
    {
        *len = 0;
        return;
    }
    *k0 = lo;
    *len = hi >= lo ? hi - lo + 1 : 0;
}

// Row -> (Gaussian, i, j, candidate interval): the Gaussian owning the row (binary search of the
// row scan between the block's first and last Gaussian), its frame, and the pruned k interval.
// rows of a Gaussian are its AABB's (i, j) pairs with i >= i_lo (the first voxel column of the
// slab being emitted; INT32_MIN without slabs)
__device__ __forceinline__ bool row_candidates(const uint64_t* __restrict__ row_scan, const PrimFrame* __restrict__ frames,
                                               const PrimBox* __restrict__ boxes, uint64_t row, uint64_t total_rows,
                                               uint64_t glo, uint64_t ghi, float vs, uint64_t* g, PrimFrame* f,
                                               int32_t* i, int32_t* j, int32_t* k0, int32_t* len,
                                               int32_t i_lo = INT32_MIN)
{
    *g = 0;
    *i = *j = *k0 = *len = 0;
    if (row >= total_rows)
        return false;
    *g = find_prim(row_scan, glo, ghi, row);
    PrimBox b = boxes[*g];
    *f = frames[*g];
    uint64_t lr = row - row_scan[*g];
    *i = max(b.mn[0], i_lo) + (int32_t)(lr / (uint64_t)b.rng_y);
    *j = b.mn[1] + (int32_t)(lr % (uint64_t)b.rng_y);
    *k0 = b.mn[2];
    *len = b.rng_z;
    row_interval(*f, b, *i, *j, vs, k0, len);
    return true;
}

__device__ __forceinline__ int32_t wave_max(int32_t v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// wave maximum of v in [0, 2^bits) by ballots (no cross-lane data movement)
__device__ __forceinline__ int32_t wave_max_small(int32_t v, int bits)
{
    int32_t m = 0;
    for (int b = bits - 1; b >= 0; b--)
    {
        const int32_t t = m | (1 << b);
        if (__ballot(v >= t))
            m = t;
    }
    return m;
}

// value of lane src (wave-uniform) of a 64-bit register
__device__ __forceinline__ uint64_t read_lane_u64(uint64_t v, uint32_t src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

// Row pass of the segment-driven emission: one lane per AABB row computes the pruned candidate
// interval and writes it as segments of at most EMIT_SEG voxels (wave-compacted, one atomic
// reservation per wave). The emission pass then walks segments, so the lanes of a wave walk at
// most EMIT_SEG steps each instead of the longest row of their 64.
__global__ __launch_bounds__(EMIT_THREADS) void k_row_segments(EmitArgs a, void* __restrict__ segs, uint64_t part_cap,
                                                               unsigned long long* __restrict__ seg_counters)
{
    const uint32_t part = blockIdx.x % SEG_PARTS;
    unsigned long long* seg_counter = seg_counters + part * SEG_COUNTER_STRIDE;
    SegCompact::Rec* part_compact = (SegCompact::Rec*)segs + (uint64_t)part * part_cap;
    SegWide::Rec* part_wide = (SegWide::Rec*)segs + (uint64_t)part * part_cap;
    __shared__ uint64_t g_range[2];
    // Gaussian of each of the block's rows, relative to the first one: the inclusive prefix count
    // of the Gaussians in (glo, ghi] starting at or before the row
    __shared__ uint32_t s_owner[EMIT_ROWS];
    __shared__ uint32_t s_wtot[EMIT_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t row0 = (uint64_t)blockIdx.x * EMIT_ROWS;
    if (wave < 2u)
    {
        uint64_t target = row0;
        if (wave == 1u)
        {
            target = row0 + EMIT_ROWS - 1;
            if (target >= a.total_rows)
                target = a.total_rows - 1;
        }
        const uint64_t g = wave_find_prim(a.row_scan, a.n, target, lane);
        if (lane == 0u)
            g_range[wave] = g;
    }
    for (uint32_t x = threadIdx.x; x < (uint32_t)EMIT_ROWS; x += EMIT_THREADS)
        s_owner[x] = 0u;
    __syncthreads();
    const uint64_t glo = g_range[0], ghi = g_range[1];
    // row_scan[glo] <= row0 < row_scan[g] <= last row for g in (glo, ghi]
    for (uint64_t g = glo + 1u + threadIdx.x; g <= ghi; g += EMIT_THREADS)
        atomicAdd(&s_owner[(uint32_t)(a.row_scan[g] - row0)], 1u);
    __syncthreads();
    {
        static_assert(EMIT_ROWS == 4 * EMIT_THREADS, "four rows per thread in the owner scan");
        const uint32_t t4 = threadIdx.x * 4u;
        const uint32_t v0 = s_owner[t4], v1 = s_owner[t4 + 1u], v2 = s_owner[t4 + 2u], v3 = s_owner[t4 + 3u];
        const uint32_t sum = v0 + v1 + v2 + v3;
        uint32_t x = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
            if (lane >= (uint32_t)off)
                x += y;
        }
        if (lane == 63u)
            s_wtot[wave] = x;
        __syncthreads();
        uint32_t pre = 0u;
        for (uint32_t w = 0; w < wave; w++)
            pre += s_wtot[w];
        const uint32_t e = pre + x - sum;
        s_owner[t4] = e + v0;
        s_owner[t4 + 1u] = e + v0 + v1;
        s_owner[t4 + 2u] = e + v0 + v1 + v2;
        s_owner[t4 + 3u] = e + sum;
    }
    __syncthreads();
    for (uint32_t chunk = wave; chunk < EMIT_ROWS / 64u; chunk += EMIT_THREADS / 64u)
    {
        const uint64_t row = row0 + (uint64_t)chunk * 64u + lane;
        if (row0 + (uint64_t)chunk * 64u >= a.total_rows)
            break;
        uint64_t g = 0;
        int32_t i = 0, j = 0, k0 = 0, len = 0;
        if (row < a.total_rows)
        {
            g = glo + s_owner[(uint32_t)(row - row0)];
            const PrimBox b = a.boxes[g];
            const uint64_t lr = row - a.row_scan[g];
            i = max(b.mn[0], a.slab_i0) + (int32_t)(lr / (uint64_t)b.rng_y);
            j = b.mn[1] + (int32_t)(lr % (uint64_t)b.rng_y);
            k0 = b.mn[2];
            len = b.rng_z;
            row_interval_frame(a.rframes[g], i, j, &k0, &len);
        }
        const uint32_t nseg = len > 0 ? (uint32_t)((len + EMIT_SEG - 1) / EMIT_SEG) : 0u;
        uint32_t total, cand_total;
        const uint32_t excl = wave_excl_add(nseg, lane, &total);
        (void)wave_excl_add(len > 0 ? (uint32_t)len : 0u, lane, &cand_total);
        unsigned long long base = 0ull;
        if (lane == 0u && total)
        {
            base = atomicAdd(seg_counter, (unsigned long long)total);
            // candidate voxels handed to the walk, in the padding word after the partition's counter
            // (one address per partition: a single shared counter would serialize the waves)
            atomicAdd(seg_counter + 1, (unsigned long long)cand_total);
        }
        base = __shfl(base, 0, 64);
        // each lane writes its own segments at base + excl + q: in the first round the lanes with a
        // segment write neighbouring records, later rounds only the lanes of long rows. A compact
        // record is the first segment's (count 1) plus q times two z leaves and the count - 1: the
        // segments of a row start 16 voxels apart, in the same slot of the leaf two further along z
        static_assert(EMIT_SEG == 16, "two z leaves per segment");
        const uint64_t rec0 = a.seg_compact && nseg ? SegCompact::make(a.kl, g, i, j, k0, 1) : 0ull;
        const uint32_t leaf_shift = a.kl.gid_bits + 4u + 9u;
        for (uint32_t q = 0; q < nseg; q++)
        {
            const uint64_t pos = base + excl + q;
            if (pos < part_cap)
            {
                const int32_t cnt = min(EMIT_SEG, len - (int32_t)q * EMIT_SEG);
                if (a.seg_compact)
                    part_compact[pos] = rec0 + ((uint64_t)(2u * q) << leaf_shift) + (uint64_t)(cnt - 1);
                else
                    part_wide[pos] = SegWide::make(a.kl, g, i, j, k0 + (int32_t)q * EMIT_SEG, cnt);
            }
        }
    }
}

// Segment-driven emission: each lane takes one candidate segment, walks its voxels with the row
// walker (same FP32 predicate, lattice planes shared within the segment) and stages survivors as
// in the row-per-lane form.
// 2 workgroups (80 KB of LDS each) per CU. Every flush reserves its output range with one atomic
// on the single pair counter, and those same-address atomics are served at a fixed rate (about
// 78 M/s measured), so the stage is as large as the LDS allows: at 1.67 G pairs, 896-survivor
// flushes bound the kernel at ~25 ms, 1728-survivor flushes at ~12 ms.
template <typename SEG>
__global__ __launch_bounds__(EMIT_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_emit_segments(EmitArgs a)
{
    using Rec = typename SEG::Rec;
    const Rec* __restrict__ seg_recs = (const Rec*)a.segs;
    __shared__ uint64_t stage_all[EMIT_THREADS / 64][EMIT_WAVE_STAGE];
    __shared__ uint32_t ring_all[EMIT_THREADS / 64][EMIT_RING_WORDS][EMIT_RING];
    __shared__ uint32_t s_hist[EMIT_HIST_WORDS];
    for (uint32_t x = threadIdx.x; x < (uint32_t)EMIT_HIST_WORDS; x += EMIT_THREADS)
        s_hist[x] = 0u;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t* stage = stage_all[wave];
    uint32_t(*ring)[EMIT_RING] = ring_all[wave];
    const bool with_alpha = a.kl.alpha_bits != 0u;
    const float vs = a.vs;
    uint32_t count = 0u; // staged survivors
    uint32_t done = 0u;  // staged survivors whose key holds its alpha: [done, count) wait in the ring
    // the alpha of every staged survivor (before a flush)
    auto finish_alpha = [&]() {
        if (!with_alpha)
            return;
        emit_alpha_batch(stage, ring, done, count - done, lane);
        wave_sync_lds();
        done = count;
    };
    const uint64_t waves_total = (uint64_t)gridDim.x * (EMIT_THREADS / 64);
    const uint64_t chunks_total = a.seg_chunk_scan[SEG_PARTS];
    // partition table in registers (lane p: partition p's first chunk, end chunk, segment count),
    // so that a chunk's partition is one ballot and no memory access
    static_assert(SEG_PARTS == 64, "one partition per lane");
    const uint64_t p_first = a.seg_chunk_scan[lane];
    const uint64_t p_end = a.seg_chunk_scan[lane + 1u];
    const uint64_t p_count = a.seg_part_count[lane];
    // wait for them here: left pending into the chunk loop, their first use there would also wait
    // for the prefetches issued after them (vmcnt: vector memory counter 0, others unchanged)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // this lane's segment of chunk c (false past the chunk's partition count or the last chunk)
    auto chunk_segment = [&](uint64_t c, Rec& sg) -> bool {
        if (c >= chunks_total)
            return false;
        const uint32_t part = (uint32_t)__popcll(__ballot(p_end <= c));
        const uint64_t first = read_lane_u64(p_first, part);
        const uint64_t cnt = read_lane_u64(p_count, part);
        const uint64_t local = (c - first) * 64u + lane;
        if (local >= cnt)
            return false;
        sg = seg_recs[(uint64_t)part * a.seg_part_cap + local];
        return true;
    };
    // software pipeline over this wave's chunks c, c + W, c + 2W, ...: while chunk c is walked,
    // the frames of chunk c + W and the segments of chunk c + 2W are in flight
    const uint64_t c_first =
        (uint64_t)blockIdx.x * (EMIT_THREADS / 64) + (uint64_t)__builtin_amdgcn_readfirstlane(wave);
    Rec sg_cur = {}, sg_next = {};
    bool has_cur = chunk_segment(c_first, sg_cur);
    bool has_next = chunk_segment(c_first + waves_total, sg_next);
    // with candidate-leaf ranks, the bitmap word and rank prefix of the segment's first leaf
    const bool ranked = a.kl.rank_bits != 0u;
    auto rank_fetch = [&](bool has, const Rec& sg, uint64_t& word, uint32_t& pre, uint32_t& bit) {
        word = 0ull;
        pre = 0u;
        bit = 0u;
        if (ranked && has)
        {
            const uint64_t leaf = SEG::vkey(a.kl, sg) >> 9;
            word = a.kl.leaf_map[leaf >> 6];
            pre = a.kl.leaf_prefix[leaf >> 6];
            bit = (uint32_t)(leaf & 63u);
        }
    };
    // lanes without a segment load Gaussian 0's frame (a zero record; there are Gaussians
    // whenever there are chunks) and walk no step
    PrimFrame f_cur = {};
    PrimVal v_cur = {};
    uint64_t rw_cur = 0ull;
    uint32_t rp_cur = 0u, rb_cur = 0u;
    if (c_first < chunks_total)
    {
        const uint64_t g0 = SEG::gid(a.kl, sg_cur);
        f_cur = a.frames[g0];
        if (with_alpha)
            v_cur = a.vals[g0];
        rank_fetch(has_cur, sg_cur, rw_cur, rp_cur, rb_cur);
    }
    for (uint64_t c = c_first; c < chunks_total; c += waves_total)
    {
        const uint64_t g_next = SEG::gid(a.kl, sg_next);
        const PrimFrame f_next = a.frames[g_next];
        PrimVal v_next = {};
        if (with_alpha)
            v_next = a.vals[g_next];
        uint64_t rw_next;
        uint32_t rp_next, rb_next;
        rank_fetch(has_next, sg_next, rw_next, rp_next, rb_next);
        Rec sg_after = {};
        const bool has_after = chunk_segment(c + 2u * waves_total, sg_after);

        int32_t len = 0, i = 0, j = 0, k0 = 0;
        uint64_t g = 0;
        const PrimFrame f = f_cur;
        const PrimVal pv = v_cur;
        const uint64_t rw = rw_cur;
        const uint32_t rp = rp_cur, rb = rb_cur;
        uint64_t vkey0 = 0ull; // voxel key of the first voxel (leaf index and slot)
        if (has_cur)
        {
            g = SEG::gid(a.kl, sg_cur);
            SEG::decode(a.kl, sg_cur, &i, &j, &k0, &len);
            vkey0 = SEG::vkey(a.kl, sg_cur);
        }
        sg_cur = sg_next;
        has_cur = has_next;
        f_cur = f_next;
        v_cur = v_next;
        rw_cur = rw_next;
        rp_cur = rp_next;
        rb_cur = rb_next;
        sg_next = sg_after;
        has_next = has_after;
        static_assert(EMIT_SEG < (1 << 5), "segment lengths fit 5 bits");
        const int32_t steps = wave_max_small(len, 5);
        // key of voxel (i, j, k0 + st) along the segment: the (i, j) part is fixed, so from one k
        // to the next the key moves by one slot, or past the z leaf's last slot (k = 7 mod 8) to
        // slot 0 of the next leaf (+ 2^9 - 7 in the voxel field)
        const uint32_t kshift = a.kl.gid_bits + a.kl.alpha_bits;
        const uint64_t key_step = 1ull << kshift;
        const uint64_t key_leaf = (512ull - 7ull) << kshift;
        uint64_t key = 0ull;
        if (len > 0)
        {
            uint64_t vkey;
            if (ranked)
            {
                // rank of the first leaf; the next leaf along z is the next candidate (the
                // segment lies in the Gaussian's AABB, whose leaves are all marked)
                const uint64_t rank = (uint64_t)rp + (uint64_t)__popcll(rw & ((1ull << rb) - 1ull));
                vkey = (rank << 9) | (vkey0 & 511u);
            }
            else
                vkey = vkey0;
            key = (vkey << kshift) | (g << a.kl.alpha_bits);
        }
        walk_row(f, i, j, k0, len, steps, vs, [&](int32_t st, bool keep, const float* ea, const float* eb) {
            const uint64_t m = __ballot(keep);
            if (keep)
            {
                const uint32_t slot =
                    count + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                stage[slot] = key;
                if (with_alpha)
                {
                    // the box and value constants for the alpha; [done, count) spans fewer than
                    // 64 + 64 slots
                    const uint32_t r = slot & (uint32_t)(EMIT_RING - 1);
#pragma unroll
                    for (int c = 0; c < 3; c++)
                    {
                        ring[c][r] = __float_as_uint(ea[c]);
                        ring[3 + c][r] = __float_as_uint(eb[c]);
                    }
                    ring[6][r] = __float_as_uint(pv.opacity);
                    ring[7][r] = __float_as_uint(pv.local_scale);
                    ring[8][r] = __float_as_uint(pv.pow_exp);
                    ring[9][r] = pv.rgb_cull;
                }
            }
            key += ((k0 + st) & 7) == 7 ? key_leaf : key_step;
            count += (uint32_t)__popcll(m);
            if (with_alpha && count - done >= 64u)
            {
                wave_sync_lds();
                emit_alpha_batch(stage, ring, done, 64u, lane);
                wave_sync_lds();
                done += 64u;
            }
            if (count > (uint32_t)(EMIT_WAVE_STAGE - 64))
            {
                wave_sync_lds();
                finish_alpha();
                emit_flush_wave(a, stage, count, lane, s_hist);
                wave_sync_lds();
                count = 0u;
                done = 0u;
            }
        });
    }
    wave_sync_lds();
    finish_alpha();
    emit_flush_wave(a, stage, count, lane, s_hist);
    if (a.hist)
    {
        __syncthreads();
        const uint32_t passes = (a.hist_bits + a.hist_w - 1u) / a.hist_w;
        for (uint32_t x = threadIdx.x; x < (passes << a.hist_w); x += EMIT_THREADS)
        {
            const uint32_t c = s_hist[x];
            if (c)
                atomicAdd(&a.hist[(x >> a.hist_w) * SORT_BINS + (x & ((1u << a.hist_w) - 1u))],
                          (unsigned long long)c);
        }
    }
}

uint32_t emit_blocks_per_cu()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_emit_segments<SegCompact>, EMIT_THREADS, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 0u;
    }
    return blocks > 0 ? (uint32_t)blocks : 0u;
}

// whether the segments of a grid with this key layout fit SegCompact records
static bool seg_compact_fits(const KeyLayout& kl)
{
    return kl.bits[0] + kl.bits[1] + kl.bits[2] + 9u + kl.gid_bits + 4u <= 64u;
}

static void launch_emit_segments(const EmitArgs& ea, uint32_t blocks, hipStream_t st)
{
    if (ea.seg_compact)
        hipLaunchKernelGGL(k_emit_segments<SegCompact>, dim3(blocks), dim3(EMIT_THREADS), 0, st, ea);
    else
        hipLaunchKernelGGL(k_emit_segments<SegWide>, dim3(blocks), dim3(EMIT_THREADS), 0, st, ea);
}

// 2^(j / 64), j = 0..63, as the nearest doubles
__constant__ double EXP2_J64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// exp(x) for -745 <= x <= 0 with its 2^(j/64) table in LDS (t64): x = k ln2/64 + r, |r| <= ln2/128
// (Cody-Waite reduction, ln2/64 split into a 36-bit head, exact times k, and a tail), exp(r) by
// the degree-5 Taylor polynomial (truncation below 4e-17 relative), scaled by 2^(k >> 6) and the
// table entry k & 63. Within 2 ulp of exp(x). Six double constants instead of a library exp's
// thirteen, which the register allocator otherwise keeps live across the pair loops.
__device__ __forceinline__ double exp_neg(double x, const double* __restrict__ t64)
{
    const double kd = rint(x * 0x1.71547652b82fep+6); // 64 / ln2
    const double r0 = fma(-kd, 0x1.62e42fee00000p-7, x);
    const double r = fma(-kd, 0x1.a39ef35793c76p-39, r0);
    double p = 1.0 / 120.0;
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const int32_t k = (int32_t)kd;
    return ldexp(t64[k & 63] * p, k >> 6);
}

// ============================================================================ density backward
// Adjoint of the density value mode: rho(v) = sum over pairs (v, g) of o_g exp(-m_g(v) / 2), with
// m = |diag(1/s) R^T (x_v - mu)|^2 at the voxel centre, differentiated with respect to the raw
// parameters (means, log-scales, unnormalised quaternions (w, x, y, z), opacity logits) in
// double precision. The pairs are enumerated again exactly as the forward emits them (same row
// intervals and predicate) and every Gaussian accumulates 10 sums:
//   E = sum r e,  P_a = sum W v_a,  S_ab = sum W v_a v_b (a <= b)
// with r = dL/drho(v), e = exp(-m/2), W = r o e, d = x_v - mu, v = diag(1/s) R^T d, m = |v|^2.
// The finalize pass forms from them the sums of u = R^T d = diag(s) v and d = R u:
//   Q1_a = sum W u_a = s_a P_a,  Q2_a = sum W u_a^2 = s_a^2 S_aa,
//   D_ab = sum W u_a d_b = s_a sum_c R_bc s_c S_ac.
//
// dL/drho comes from a leaf table built from the grid first: every leaf is found by descending
// the root, upper and lower child masks (PNanoVDBExt.h:225-318 run over every slot), entered in
// an open-addressing hash on its leaf coordinates and expanded to 512 per-voxel values (0 for
// inactive voxels), so a pair costs one load, plus one hash probe per leaf along its row, instead
// of a tree descent and a mask lookup.
//
// Each wave takes 64 consecutive AABB rows and cuts their candidate intervals into segments of
// at most BWD_SEG voxels. Per round it walks 64 segments with the FP32 predicate (its lanes step
// through at most BWD_SEG voxels instead of the longest of 64 rows) and stages the kept voxels in
// LDS in segment order; the double-precision terms then run over the staged pairs with every lane
// busy, each lane summing a contiguous slice sequentially. Segments keep row order, so Gaussian ids
// are non-decreasing over the staged pairs: runs complete inside a lane are stored directly, one
// segmented scan over the lanes per round joins the runs that cross lanes, and the round's last
// run carries into the next round. A Gaussian owning the first or last row of a 64-row chunk
// leaves its chunk sum in that chunk's record (the finalize pass sums records in chunk order); one
// strictly inside a chunk stores its sum directly. Every sum runs in an order fixed by the rows,
// so the result does not depend on scheduling.
constexpr int BWD_ACC = 10;
constexpr int32_t BWD_SEG = 8;
constexpr uint32_t LT_NONE = 0xFFFFFFFFu;

// double-precision parameters of one Gaussian as the pair terms use them (96 bytes)
struct BwdPrim
{
    double M[9]; // diag(1/s) R^T, row-major (floored scales)
    double o;
    float mu[3]; // the float means, exact in the double arithmetic of the terms
    float pad;
};

struct BwdArgs
{
    const PrimFrame* frames;
    const PrimBox* boxes;
    const uint64_t* row_scan;
    uint64_t n;
    uint64_t total_rows;
    float vs;
    const float* means;
    const float* log_scales;
    const float* quats; // raw
    const float* logits;
    BwdPrim* prims;
    const uint4* lt_slots;
    uint32_t lt_mask;
    const float* leaf_grad;
    double* acc;       // [n][16] Gaussians whose rows lie inside one chunk, away from its ends (zeroed first)
    double* rec;       // [chunks][2][16] first and last Gaussian of each chunk
    uint64_t* rec_gid; // [chunks][2] (~0: no second Gaussian)
    float* grad_means;
    float* grad_logits;
    float* grad_quats;
    float* grad_log_scales;
};

// Leaf table of a node2 grid (built by k_lt_root, k_lt_upper, k_lt_lower, k_lt_grad in order)
struct NodeRef
{
    uint64_t word; // node2 word index of the node
    int32_t o[3];  // index-space origin
    int32_t pad;
};

struct LeafTable
{
    const uint64_t* grid; // node2 grid words
    uint64_t words;       // grid size in words
    NodeRef* uppers;
    NodeRef* lowers;
    uint64_t* leaves;     // node2 word index by leaf ordinal
    uint4* slots;         // (i >> 3, j >> 3, k >> 3, ordinal + 1); w == 0: empty
    uint32_t mask;        // slot count - 1 (a power of two, at least twice the leaf capacity)
    uint32_t cap[3];      // upper, lower, leaf list capacities
    unsigned int* counters; // [0..2] nodes found per level, [3] inconsistency flag
    const float* grad_values;
    uint64_t value_count;
    float* leaf_grad;     // [ordinal][512] dL/drho by voxel slot n = (i & 7) << 6 | (j & 7) << 3 | (k & 7)
};

__device__ __forceinline__ uint32_t lt_hash(int32_t li, int32_t lj, int32_t lk)
{
    uint32_t h = (uint32_t)li * 0x9E3779B1u ^ (uint32_t)lj * 0x85EBCA77u ^ (uint32_t)lk * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

// leaf ordinal of leaf coordinates (li, lj, lk), or LT_NONE
__device__ __forceinline__ uint32_t lt_find(const uint4* __restrict__ slots, uint32_t mask, int32_t li, int32_t lj,
                                            int32_t lk)
{
    uint32_t h = lt_hash(li, lj, lk) & mask;
    for (uint32_t p = 0; p <= mask; p++)
    {
        const uint4 s = slots[h];
        if (s.w == 0u)
            return LT_NONE;
        if (s.x == (uint32_t)li && s.y == (uint32_t)lj && s.z == (uint32_t)lk)
            return s.w - 1u;
        h = (h + 1u) & mask;
    }
    return LT_NONE;
}

// child node of a parent from its table entry (byte offset from the parent), bounded by the grid
__device__ __forceinline__ bool lt_child(const LeafTable& t, uint64_t parent, uint64_t entry, uint64_t child_words,
                                         uint64_t* child)
{
    const uint64_t rel = t.grid[entry] >> 3;
    if (rel == 0 || rel > t.words || parent + rel + child_words > t.words)
        return false;
    *child = parent + rel;
    return true;
}

__device__ __forceinline__ bool lt_push(const LeafTable& t, int level, uint32_t* slot)
{
    const uint32_t s = atomicAdd(&t.counters[level], 1u);
    if (s >= t.cap[level])
    {
        atomicOr(&t.counters[3], 1u);
        return false;
    }
    *slot = s;
    return true;
}

// root slots -> upper nodes
__global__ __launch_bounds__(256) void k_lt_root(LeafTable t)
{
    const uint32_t rn = blockIdx.x * 256u + threadIdx.x;
    if (rn >= 32768u)
        return;
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    bool act = false, kact = false;
    const uint32_t rank =
        n2_mask_n_to_idx(t.grid, root + VDB_OFF_CMP_ROOT, VDB_OFF_CM_ROOT - VDB_OFF_CMP_ROOT, rn, &act);
    if (!act)
        return;
    const uint32_t kidx =
        n2_mask_n_to_idx(t.grid, root + VDB_OFF_KMP_ROOT, VDB_OFF_KM_ROOT - VDB_OFF_KMP_ROOT, rn, &kact);
    if (!kact)
        return;
    const uint32_t child_count = (uint32_t)(t.grid[root + VDB_OFF_CMP_ROOT] & 0xFFFFu);
    // the key holds (u >> 12) of the unsigned coordinates u: x at bit 42, y at bit 21, z at bit 0
    const uint64_t key = t.grid[root + VDB_OFF_CHILDREN_ROOT + child_count + kidx];
    const uint32_t ku[3] = { (uint32_t)(key >> 42) & 0xFFFFFu, (uint32_t)(key >> 21) & 0xFFFFFu,
                             (uint32_t)key & 0xFFFFFu };
    // a lookup reaches this slot only with coordinates whose low five root bits name the slot
    if ((ku[0] & 31u) != (rn >> 10) || (ku[1] & 31u) != ((rn >> 5) & 31u) || (ku[2] & 31u) != (rn & 31u))
        return;
    uint64_t upper;
    uint32_t s;
    if (!lt_child(t, root, root + VDB_OFF_CHILDREN_ROOT + rank, VDB_SIZE_UPPER >> 3, &upper))
    {
        atomicOr(&t.counters[3], 1u);
        return;
    }
    if (!lt_push(t, 0, &s))
        return;
    NodeRef r;
    r.word = upper;
    r.pad = 0;
    for (int c = 0; c < 3; c++)
        r.o[c] = (int32_t)(ku[c] << 12);
    t.uppers[s] = r;
}

// (upper, slot) -> lower nodes
__global__ __launch_bounds__(256) void k_lt_upper(LeafTable t)
{
    const uint64_t x = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t u = x >> 15;
    const uint32_t un = (uint32_t)x & 32767u;
    if (u >= (uint64_t)min(t.counters[0], t.cap[0]))
        return;
    const NodeRef p = t.uppers[u];
    bool act = false;
    const uint32_t rank =
        n2_mask_n_to_idx(t.grid, p.word + VDB_OFF_CMP_UPPER, VDB_OFF_CM_UPPER - VDB_OFF_CMP_UPPER, un, &act);
    if (!act)
        return;
    uint64_t lower;
    uint32_t s;
    if (!lt_child(t, p.word, p.word + VDB_OFF_CHILDREN_UPPER + rank, VDB_SIZE_LOWER >> 3, &lower))
    {
        atomicOr(&t.counters[3], 1u);
        return;
    }
    if (!lt_push(t, 1, &s))
        return;
    NodeRef r;
    r.word = lower;
    r.pad = 0;
    r.o[0] = p.o[0] + (int32_t)((un >> 10) << 7);
    r.o[1] = p.o[1] + (int32_t)(((un >> 5) & 31u) << 7);
    r.o[2] = p.o[2] + (int32_t)((un & 31u) << 7);
    t.lowers[s] = r;
}

// (lower, slot) -> leaves: ordinal, hash entry
__global__ __launch_bounds__(256) void k_lt_lower(LeafTable t)
{
    const uint64_t x = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t l = x >> 12;
    const uint32_t ln = (uint32_t)x & 4095u;
    if (l >= (uint64_t)min(t.counters[1], t.cap[1]))
        return;
    const NodeRef p = t.lowers[l];
    bool act = false;
    const uint32_t rank =
        n2_mask_n_to_idx(t.grid, p.word + VDB_OFF_CMP_LOWER, VDB_OFF_CM_LOWER - VDB_OFF_CMP_LOWER, ln, &act);
    if (!act)
        return;
    uint64_t leaf;
    uint32_t s;
    if (!lt_child(t, p.word, p.word + VDB_OFF_CHILDREN_LOWER + rank, VDB_SIZE_LEAF >> 3, &leaf))
    {
        atomicOr(&t.counters[3], 1u);
        return;
    }
    if (!lt_push(t, 2, &s))
        return;
    t.leaves[s] = leaf;
    const int32_t li = (p.o[0] >> 3) + (int32_t)(ln >> 8);
    const int32_t lj = (p.o[1] >> 3) + (int32_t)((ln >> 4) & 15u);
    const int32_t lk = (p.o[2] >> 3) + (int32_t)(ln & 15u);
    uint32_t h = lt_hash(li, lj, lk) & t.mask;
    for (uint32_t q = 0; q <= t.mask; q++)
    {
        if (atomicCAS(&t.slots[h].w, 0u, s + 1u) == 0u)
        {
            t.slots[h].x = (uint32_t)li;
            t.slots[h].y = (uint32_t)lj;
            t.slots[h].z = (uint32_t)lk;
            return;
        }
        h = (h + 1u) & t.mask;
    }
    atomicOr(&t.counters[3], 1u);
}

// (leaf ordinal, voxel slot) -> dL/drho, 0 for inactive voxels
__global__ __launch_bounds__(256) void k_lt_grad(LeafTable t)
{
    const uint64_t x = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t l = x >> 9;
    const uint32_t n = (uint32_t)x & 511u;
    if (l >= (uint64_t)min(t.counters[2], t.cap[2]))
        return;
    const uint64_t leaf = t.leaves[l];
    bool act = false;
    const uint32_t local =
        n2_mask_n_to_idx(t.grid, leaf + VDB_OFF_VMP_LEAF, VDB_OFF_VM_LEAF - VDB_OFF_VMP_LEAF, n, &act);
    const uint64_t vi = t.grid[leaf + VDB_OFF_VALUE_IDX] + local;
    t.leaf_grad[x] = act && vi < t.value_count ? t.grad_values[vi] : 0.f;
}

// double-precision parameters of one Gaussian from its raw parameters (RasterUtils.cpp:59-117
// preprocessing, gaussian_prim.slang:67-68 scale floor)
struct PrimD
{
    double R[3][3];
    double q[4]; // normalised (w, x, y, z)
    double qn;
    double s[3];
    double is2[3]; // 1 / s^2
    bool floored[3];
    double o;
    double mu[3];
};

__device__ __forceinline__ void load_prim_d(const BwdArgs& a, uint64_t g, PrimD* p)
{
    double qw = a.quats[4 * g + 0], qx = a.quats[4 * g + 1], qy = a.quats[4 * g + 2], qz = a.quats[4 * g + 3];
    double qn = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    double w = qw / qn, x = qx / qn, y = qy / qn, z = qz / qn;
    p->q[0] = w;
    p->q[1] = x;
    p->q[2] = y;
    p->q[3] = z;
    p->qn = qn;
    p->R[0][0] = 1 - 2 * (y * y + z * z);
    p->R[0][1] = 2 * (x * y - z * w);
    p->R[0][2] = 2 * (x * z + y * w);
    p->R[1][0] = 2 * (x * y + z * w);
    p->R[1][1] = 1 - 2 * (x * x + z * z);
    p->R[1][2] = 2 * (y * z - x * w);
    p->R[2][0] = 2 * (x * z - y * w);
    p->R[2][1] = 2 * (y * z + x * w);
    p->R[2][2] = 1 - 2 * (x * x + y * y);
    const double fl = (double)(0.125f * a.vs);
    for (int c = 0; c < 3; c++)
    {
        double e = exp((double)a.log_scales[3 * g + c]);
        p->floored[c] = !(e > fl);
        p->s[c] = p->floored[c] ? fl : e;
        p->is2[c] = 1.0 / (p->s[c] * p->s[c]);
        p->mu[c] = a.means[3 * g + c];
    }
    p->o = 1.0 / (1.0 + exp(-(double)a.logits[g]));
}

__global__ __launch_bounds__(256) void k_bwd_prims(BwdArgs a)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n)
        return;
    PrimD p;
    load_prim_d(a, g, &p);
    BwdPrim q;
    for (int c = 0; c < 3; c++)
    {
        for (int b = 0; b < 3; b++)
            q.M[3 * c + b] = p.R[b][c] / p.s[c];
        q.mu[c] = a.means[3 * g + c];
    }
    q.o = p.o;
    q.pad = 0.f;
    a.prims[g] = q;
}

__device__ __forceinline__ double readlane_f64(double x, int l)
{
    return __longlong_as_double((long long)readlane_u64((uint64_t)__double_as_longlong(x), l));
}

// last lane L whose (non-decreasing) exclusive count excl(L) is <= idx
__device__ __forceinline__ uint32_t wave_owner(uint32_t excl, uint32_t idx)
{
    uint32_t lo = 0, hi = 63;
#pragma unroll
    for (int step = 0; step < 6; step++)
    {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        const uint32_t ex_mid = (uint32_t)__shfl((int)excl, (int)mid, 64);
        if (ex_mid <= idx)
            lo = mid;
        else
            hi = mid - 1u;
    }
    return lo;
}

// At least 4 waves per SIMD (128 VGPRs, a few spills): the pass waits on memory, and 4 waves beat
// the 3 that the unconstrained allocation (144 VGPRs) allows (74.5 vs 86 ms on 2M Gaussians).
__global__ __launch_bounds__(EMIT_THREADS, 4) void k_density_bwd(BwdArgs a)
{
    __shared__ uint64_t g_range[2];
    __shared__ double carry_all[EMIT_THREADS / 64][BWD_ACC];
    __shared__ uint32_t staged_all[EMIT_THREADS / 64][64 * BWD_SEG]; // leaf ordinal << 3 | step
    __shared__ uint32_t seg_all[EMIT_THREADS / 64][5][64];           // per segment: id, i, j, k0, pair offset
    __shared__ double fpiece_all[EMIT_THREADS / 64][BWD_ACC * 64];    // per lane: sums before its first change
    __shared__ double s_exp2[64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    double* carry = carry_all[wave];
    uint32_t* staged = staged_all[wave];
    uint32_t* seg_gid = seg_all[wave][0];
    int32_t* seg_i = (int32_t*)seg_all[wave][1];
    int32_t* seg_j = (int32_t*)seg_all[wave][2];
    int32_t* seg_k0 = (int32_t*)seg_all[wave][3];
    uint32_t* seg_ex = seg_all[wave][4];
    double* fpiece = fpiece_all[wave];
    const uint64_t row0 = (uint64_t)blockIdx.x * EMIT_ROWS;
    if (threadIdx.x == 0)
    {
        g_range[0] = find_prim(a.row_scan, 0, a.n - 1, row0);
        uint64_t last = row0 + EMIT_ROWS - 1;
        if (last >= a.total_rows)
            last = a.total_rows - 1;
        g_range[1] = find_prim(a.row_scan, 0, a.n - 1, last);
    }
    if (threadIdx.x < 64u)
        s_exp2[threadIdx.x] = EXP2_J64[threadIdx.x];
    __syncthreads();
    const uint64_t glo = g_range[0], ghi = g_range[1];
    const float vs = a.vs;
    const double vsd = (double)vs;
    for (uint32_t chunk = wave; chunk < EMIT_ROWS / 64u; chunk += EMIT_THREADS / 64u)
    {
        const uint64_t chunk_row = row0 + (uint64_t)chunk * 64u;
        if (chunk_row >= a.total_rows)
            break;
        const uint64_t row = chunk_row + lane;
        uint64_t g;
        int32_t i, j, k0, len;
        PrimFrame f;
        const bool active =
            row_candidates(a.row_scan, a.frames, a.boxes, row, a.total_rows, glo, ghi, vs, &g, &f, &i, &j, &k0, &len);
        const uint64_t key = active ? g : ~0ull;
        const uint64_t amask = __ballot(active);
        const uint32_t last_lane = 63u - (uint32_t)__clzll(amask);
        const uint64_t first_key = readlane_u64(key, 0);
        const uint64_t last_key = readlane_u64(key, (int)last_lane);
        const uint64_t c = chunk_row >> 6;
        const uint32_t nseg = len > 0 ? (uint32_t)((len + BWD_SEG - 1) / BWD_SEG) : 0u;
        uint32_t total;
        const uint32_t excl = wave_excl_add(nseg, lane, &total);
        auto dest = [&](uint64_t gg) -> double* {
            if (gg == first_key)
                return a.rec + (2 * c) * BWD_ACC;
            if (gg == last_key)
                return a.rec + (2 * c + 1) * BWD_ACC;
            return a.acc + gg * BWD_ACC;
        };
        bool first_done = false, last_done = false; // wave-uniform
        uint64_t carry_g = ~0ull;                   // wave-uniform
        for (uint32_t s0 = 0; s0 < total; s0 += 64u)
        {
            // walk: segment sidx belongs to the last lane (row) whose exclusive count is <= sidx
            const uint32_t sidx = s0 + lane;
            const uint32_t owner = wave_owner(excl, sidx);
            const uint32_t o_excl = (uint32_t)__shfl((int)excl, (int)owner, 64);
            const uint64_t o_g = __shfl(key, (int)owner, 64);
            const int32_t si = __shfl(i, (int)owner, 64);
            const int32_t sj = __shfl(j, (int)owner, 64);
            const int32_t o_k0 = __shfl(k0, (int)owner, 64);
            const int32_t o_len = __shfl(len, (int)owner, 64);
            const bool has = sidx < total;
            const int32_t q = (int32_t)(sidx - o_excl);
            uint64_t sg = ~0ull;
            int32_t sk0 = 0, slen = 0;
            PrimFrame sf;
            if (has)
            {
                sg = o_g;
                sk0 = o_k0 + q * BWD_SEG;
                slen = min(BWD_SEG, o_len - q * BWD_SEG);
                sf = a.frames[sg];
            }
            const int32_t steps = wave_max(slen);
            uint32_t lord = LT_NONE, kept = 0u;
            int32_t leaf_k = INT32_MIN;
            walk_row(sf, si, sj, sk0, slen, steps, vs, [&](int32_t st, bool keep, const float*, const float*) {
                if (!keep)
                    return;
                const int32_t k = sk0 + st;
                if ((k >> 3) != leaf_k)
                {
                    lord = lt_find(a.lt_slots, a.lt_mask, si >> 3, sj >> 3, k >> 3);
                    leaf_k = k >> 3;
                }
                if (lord != LT_NONE)
                    staged[lane * BWD_SEG + kept++] = (lord << 3) | (uint32_t)st;
            });
            uint32_t kept_total;
            const uint32_t kexcl = wave_excl_add(kept, lane, &kept_total);
            if (kept_total == 0u)
            {
                wave_sync_lds();
                continue;
            }
            seg_gid[lane] = (uint32_t)sg;
            seg_i[lane] = si;
            seg_j[lane] = sj;
            seg_k0[lane] = sk0;
            seg_ex[lane] = kexcl;
            wave_sync_lds();
            // terms: lane L takes the staged pairs [L K, L K + K) in segment order and sums them
            // sequentially; F = the piece before its first Gaussian change (it continues the run of
            // the pair before), runs between changes are complete, A = the piece after the last change
            const uint32_t K = (kept_total + 63u) >> 6; // at most BWD_SEG
            const uint32_t q0 = lane * K;
            const uint32_t q1 = min(q0 + K, kept_total);
            uint32_t sl = wave_owner(kexcl, q0 < kept_total ? q0 : 0u);
            const uint32_t sprev = wave_owner(kexcl, q0 > 0u ? min(q0 - 1u, kept_total - 1u) : 0u);
            const uint32_t carry32 = carry_g == ~0ull ? 0xFFFFFFFFu : (uint32_t)carry_g;
            const uint32_t gprev = q0 == 0u ? carry32 : seg_gid[sprev];
            double A[BWD_ACC];
#pragma unroll
            for (int t = 0; t < BWD_ACC; t++)
                A[t] = 0.0;
            uint32_t cur = gprev, heads = 0u;
            bool hit_first = false, hit_last = false;
#pragma unroll 1
            for (uint32_t q = q0; q < q1; q++)
            {
                while (sl < 63u && q >= seg_ex[sl + 1u])
                    sl++;
                const uint32_t g = seg_gid[sl];
                if (g != cur)
                {
                    if (heads == 0u)
                    {
                        // F waits in LDS for the cross-lane join
                        for (int t = 0; t < BWD_ACC; t++)
                            fpiece[t * 64 + lane] = A[t];
                    }
                    else
                    {
                        double* dst = dest(cur);
                        for (int t = 0; t < BWD_ACC; t++)
                            dst[t] = A[t];
                        hit_first = hit_first || cur == first_key;
                        hit_last = hit_last || cur == last_key;
                    }
                    heads++;
                    cur = g;
#pragma unroll
                    for (int t = 0; t < BWD_ACC; t++)
                        A[t] = 0.0;
                }
                const uint32_t ent = staged[sl * BWD_SEG + (q - seg_ex[sl])];
                const int32_t pi = seg_i[sl], pj = seg_j[sl];
                const int32_t k = seg_k0[sl] + (int32_t)(ent & 7u);
                const uint32_t nv = (((uint32_t)pi & 7u) << 6) | (((uint32_t)pj & 7u) << 3) | ((uint32_t)k & 7u);
                const double r = (double)a.leaf_grad[(uint64_t)(ent >> 3) * 512u + nv];
                const BwdPrim& spm = a.prims[g];
                const double d[3] = { ((double)pi + 0.5) * vsd - spm.mu[0], ((double)pj + 0.5) * vsd - spm.mu[1],
                                      ((double)k + 0.5) * vsd - spm.mu[2] };
                double v[3];
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
                    v[cc] = spm.M[3 * cc + 0] * d[0] + spm.M[3 * cc + 1] * d[1] + spm.M[3 * cc + 2] * d[2];
                const double m = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                const double e = exp_neg(-0.5 * m, s_exp2);
                const double W = r * spm.o * e;
                const double w0 = W * v[0], w1 = W * v[1], w2 = W * v[2];
                A[0] += r * e;
                A[1] += w0;
                A[2] += w1;
                A[3] += w2;
                A[4] += w0 * v[0];
                A[5] += w1 * v[1];
                A[6] += w2 * v[2];
                A[7] += w0 * v[1];
                A[8] += w0 * v[2];
                A[9] += w1 * v[2];
            }
            // cross-lane join: segmented inclusive scan over the lanes of A (the piece after the
            // lane's last change, or the whole lane inside a run), segments starting at lanes with
            // a change
            const bool hh = heads > 0u;
            const uint64_t hb = __ballot(hh);
            const uint64_t hle = hb & (lane == 63u ? ~0ull : ((2ull << lane) - 1ull));
            const uint32_t segid = hle ? 63u - (uint32_t)__clzll(hle) : 0u;
            double* S = A;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1)
            {
                const uint32_t sid = (uint32_t)__shfl_up((int)segid, off, 64);
                const bool take = lane >= (uint32_t)off && sid == segid;
#pragma unroll
                for (int t = 0; t < BWD_ACC; t++)
                {
                    const double u = __shfl_up(S[t], off, 64);
                    if (take)
                        S[t] += u;
                }
            }
            // the run coming into a lane with a change ends at that change: the scan of the lanes
            // before it, the lane's F and, for the round's first run, the carried piece
            const bool round_first = (hb & ((1ull << lane) - 1ull)) == 0ull;
            const bool add_carry = round_first && carry32 != 0xFFFFFFFFu && gprev == carry32;
            double inc[BWD_ACC];
#pragma unroll
            for (int t = 0; t < BWD_ACC; t++)
            {
                const double u = __shfl_up(S[t], 1, 64);
                inc[t] = (lane > 0u ? u : 0.0) + (hh ? fpiece[t * 64 + lane] : 0.0);
            }
            if (hh && gprev != 0xFFFFFFFFu)
            {
                double* dst = dest(gprev);
                for (int t = 0; t < BWD_ACC; t++)
                    dst[t] = add_carry ? carry[t] + inc[t] : inc[t];
                hit_first = hit_first || gprev == first_key;
                hit_last = hit_last || gprev == last_key;
            }
            first_done = first_done || __ballot(hit_first) != 0ull;
            last_done = last_done || __ballot(hit_last) != 0ull;
            // the run open at the round's last pair carries into the next round
            const uint32_t lastl = (kept_total - 1u) / K;
            const uint32_t last_g = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)lastl);
            double nc[BWD_ACC];
#pragma unroll
            for (int t = 0; t < BWD_ACC; t++)
                nc[t] = readlane_f64(S[t], (int)lastl);
            wave_sync_lds();
            if (lane == 0u)
            {
                for (int t = 0; t < BWD_ACC; t++)
                    carry[t] = hb == 0ull ? carry[t] + nc[t] : nc[t];
            }
            carry_g = last_g;
            wave_sync_lds();
        }
        if (carry_g != ~0ull)
        {
            if (lane == 0u)
            {
                double* dst = dest(carry_g);
                for (int t = 0; t < BWD_ACC; t++)
                    dst[t] = carry[t];
            }
            first_done = first_done || carry_g == first_key;
            last_done = last_done || carry_g == last_key;
        }
        // chunk records: ids, and zero sums for end Gaussians without pairs in this chunk
        if (lane == 0u)
        {
            a.rec_gid[2 * c] = first_key;
            a.rec_gid[2 * c + 1] = last_key != first_key ? last_key : ~0ull;
            if (!first_done)
            {
                for (int t = 0; t < BWD_ACC; t++)
                    a.rec[(2 * c) * BWD_ACC + t] = 0.0;
            }
            if (last_key != first_key && !last_done)
            {
                for (int t = 0; t < BWD_ACC; t++)
                    a.rec[(2 * c + 1) * BWD_ACC + t] = 0.0;
            }
        }
        wave_sync_lds();
    }
}

// gradients of Gaussian g from its 10 sums A (see above) through the chain rule of the
// preprocessing (RasterUtils.cpp:59-117, gaussian_prim.slang:67-68 scale floor)
__device__ __forceinline__ void bwd_write_grads(const BwdArgs& a, uint64_t g, const double* A)
{
    PrimD p;
    load_prim_d(a, g, &p);
    // sums of v = diag(1/s) R^T d: P_a = A[1 + a], S (symmetric) from A[4..9]
    double S[3][3];
    S[0][0] = A[4];
    S[1][1] = A[5];
    S[2][2] = A[6];
    S[0][1] = S[1][0] = A[7];
    S[0][2] = S[2][0] = A[8];
    S[1][2] = S[2][1] = A[9];
    a.grad_logits[g] = (float)(A[0] * p.o * (1.0 - p.o));
    // dL/dmu_b = sum_c R_bc Q1_c / s_c^2 = sum_c R_bc P_c / s_c; dL/dlog s_b = Q2_b / s_b^2 = S_bb
    for (int b = 0; b < 3; b++)
    {
        double gm = 0.0;
        for (int c = 0; c < 3; c++)
            gm += p.R[b][c] * A[1 + c] / p.s[c];
        a.grad_means[3 * g + b] = (float)gm;
        a.grad_log_scales[3 * g + b] = p.floored[b] ? 0.f : (float)S[b][b];
    }
    // dL/dR[b][c] = -D_cb / s_c^2 with D_cb = s_c sum_e R_be s_e S_ce, then through R(q) and
    // q = q_raw / |q_raw|
    double G[3][3];
    for (int b = 0; b < 3; b++)
        for (int c = 0; c < 3; c++)
        {
            double dcb = 0.0;
            for (int e = 0; e < 3; e++)
                dcb += p.R[b][e] * p.s[e] * S[c][e];
            G[b][c] = -dcb / p.s[c];
        }
    const double w = p.q[0], x = p.q[1], y = p.q[2], z = p.q[3];
    const double dRw[3][3] = { { 0, -2 * z, 2 * y }, { 2 * z, 0, -2 * x }, { -2 * y, 2 * x, 0 } };
    const double dRx[3][3] = { { 0, 2 * y, 2 * z }, { 2 * y, -4 * x, -2 * w }, { 2 * z, 2 * w, -4 * x } };
    const double dRy[3][3] = { { -4 * y, 2 * x, 2 * w }, { 2 * x, 0, 2 * z }, { -2 * w, 2 * z, -4 * y } };
    const double dRz[3][3] = { { -4 * z, -2 * w, 2 * x }, { 2 * w, -4 * z, 2 * y }, { 2 * x, 2 * y, 0 } };
    double gq[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (int b = 0; b < 3; b++)
        for (int c = 0; c < 3; c++)
        {
            gq[0] += G[b][c] * dRw[b][c];
            gq[1] += G[b][c] * dRx[b][c];
            gq[2] += G[b][c] * dRy[b][c];
            gq[3] += G[b][c] * dRz[b][c];
        }
    const double dot = gq[0] * w + gq[1] * x + gq[2] * y + gq[3] * z;
    for (int k = 0; k < 4; k++)
        a.grad_quats[4 * g + k] = (float)((gq[k] - dot * p.q[k]) / p.qn);
}

// per Gaussian: sum its partials in chunk order and apply the chain rule of the preprocessing
__global__ __launch_bounds__(256) void k_density_bwd_finalize(BwdArgs a)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n)
        return;
    double A[BWD_ACC];
#pragma unroll
    for (int q = 0; q < BWD_ACC; q++)
        A[q] = 0.0;
    const uint64_t r0 = a.row_scan[g], r1 = a.row_scan[g + 1];
    if (r1 > r0)
    {
        const uint64_t fc = r0 >> 6, lc = (r1 - 1) >> 6;
        const bool direct = fc == lc && a.rec_gid[2 * fc] != g && a.rec_gid[2 * fc + 1] != g;
        if (direct)
        {
#pragma unroll
            for (int q = 0; q < BWD_ACC; q++)
                A[q] = a.acc[g * BWD_ACC + q];
        }
        else
        {
            for (uint64_t c = fc; c <= lc; c++)
            {
                const double* src = nullptr;
                if (a.rec_gid[2 * c] == g)
                    src = a.rec + (2 * c) * BWD_ACC;
                else if (a.rec_gid[2 * c + 1] == g)
                    src = a.rec + (2 * c + 1) * BWD_ACC;
                if (src)
                {
#pragma unroll
                    for (int q = 0; q < BWD_ACC; q++)
                        A[q] += src[q];
                }
            }
        }
    }
    bwd_write_grads(a, g, A);
}

// ============================================================================ density backward from the pair list
// The same 10 sums per Gaussian, taken over the sorted pair list that the density forward kept
// (grouped by leaf; inside a leaf in emission order, where neighbouring pairs mostly belong to
// one Gaussian). Every term is bounded by |dL/drho|: e <= 1, |v_a| e <= sqrt(m) e^(-m/2) <= e^-1/2
// and |v_a v_b| e <= m e^(-m/2) <= 2/e. With 2^scale = 2^BP_TERM_BITS / 2^ceil(log2 max|dL/drho|)
// every scaled term rounds to nearest to an integer of magnitude at most 2^48, and the sums are
// integer sums: exact, so their value does not depend on the order in which pairs, lanes or
// blocks add them. Each lane takes BP_PER_LANE consecutive pairs and sums the run of one Gaussian
// in registers; a finished run goes to the block's LDS table (BP_SLOTS Gaussians, 64-bit sums: a
// tile holds at most 2^12 terms of magnitude 2^48), or straight to the global sums when its probe
// finds no free slot; at the end of the tile the table goes to the global 128-bit sums, Gaussian
// by Gaussian (the 10 sums of one Gaussian are adjacent, so a wave's atomics touch few lines).
constexpr int BP_THREADS = 256;
constexpr int BP_PER_LANE = 16;
constexpr uint64_t BP_TILE = (uint64_t)BP_THREADS * BP_PER_LANE;
constexpr uint32_t BP_SLOTS = 256;
constexpr uint32_t BP_PROBES = 32;
constexpr uint32_t BP_EMPTY = 0xFFFFFFFFu;
constexpr int BP_TERM_BITS = 48;

struct PairBwdArgs
{
    const uint64_t* keys; // sorted pair keys: voxel key << gid_bits | Gaussian id
    uint64_t npairs;
    KeyLayout kl;
    double vs;
    const BwdPrim* prims;
    const uint4* lt_slots;
    uint32_t lt_mask;
    const float* leaf_grad;
    const unsigned int* rmax_bits; // bits of max |dL/drho| (float)
    unsigned long long* sums;      // [n][BWD_ACC] x (low word, high word) of 128-bit sums, zeroed first
};

// 2^s with max|dL/drho| * 2^s < 2^BP_TERM_BITS (s = 0 when every dL/drho is 0)
__device__ __forceinline__ int bp_scale_exp(unsigned int rmax_bits)
{
    const double r = (double)__uint_as_float(rmax_bits);
    int ex = 0;
    if (r > 0.0)
        (void)frexp(r, &ex); // r < 2^ex
    return r > 0.0 ? BP_TERM_BITS - ex : 0;
}

// x rounded to the nearest integer (ties to even), |x| < 2^51: x + 1.5 * 2^52 has an ulp of 1 and
// keeps the integer in its low mantissa bits
__device__ __forceinline__ int64_t bp_round(double x)
{
    const double big = 0x1.8p52;
    return (int64_t)((uint64_t)__double_as_longlong(x + big) - (uint64_t)__double_as_longlong(big));
}

// v added to the 128-bit two's complement sum (w[0] low word, w[1] high word). The carry out of
// the low word is taken against the value this atomic replaced, so concurrent adds compose exactly.
__device__ __forceinline__ void atomic_add_i128(unsigned long long* w, int64_t v)
{
    const unsigned long long lo = (unsigned long long)v;
    const unsigned long long old = atomicAdd(&w[0], lo);
    const unsigned long long hi = (v < 0 ? ~0ull : 0ull) + (old + lo < old ? 1ull : 0ull);
    if (hi)
        atomicAdd(&w[1], hi);
}

__device__ __forceinline__ void bp_flush(const PairBwdArgs& a, uint32_t* s_gid,
                                         unsigned long long (*s_acc)[BP_SLOTS], uint32_t g, const int64_t* A)
{
    uint32_t h = (g * 0x9E3779B1u) >> 24;
    for (uint32_t p = 0; p < BP_PROBES; p++)
    {
        const uint32_t old = atomicCAS(&s_gid[h], BP_EMPTY, g);
        if (old == BP_EMPTY || old == g)
        {
#pragma unroll
            for (int t = 0; t < BWD_ACC; t++)
                if (A[t] != 0)
                    atomicAdd(&s_acc[t][h], (unsigned long long)A[t]);
            return;
        }
        h = (h + 1u) & (BP_SLOTS - 1u);
    }
#pragma unroll
    for (int t = 0; t < BWD_ACC; t++)
        if (A[t] != 0)
            atomic_add_i128(a.sums + ((uint64_t)g * BWD_ACC + t) * 2u, A[t]);
}

__global__ __launch_bounds__(BP_THREADS) void k_density_bwd_pairs(PairBwdArgs a)
{
    __shared__ uint32_t s_gid[BP_SLOTS];
    __shared__ unsigned long long s_acc[BWD_ACC][BP_SLOTS];
    __shared__ double s_exp2[64];
    for (uint32_t e = threadIdx.x; e < BP_SLOTS; e += BP_THREADS)
    {
        s_gid[e] = BP_EMPTY;
#pragma unroll
        for (int t = 0; t < BWD_ACC; t++)
            s_acc[t][e] = 0ull;
    }
    if (threadIdx.x < 64u)
        s_exp2[threadIdx.x] = EXP2_J64[threadIdx.x];
    __syncthreads();
    const double scale = ldexp(1.0, bp_scale_exp(*a.rmax_bits));
    const uint32_t gb = a.kl.gid_bits;
    const uint64_t gmask = (1ull << gb) - 1ull;
    const double vs = a.vs;
    const uint64_t q0 = (uint64_t)blockIdx.x * BP_TILE + (uint64_t)threadIdx.x * BP_PER_LANE;
    const uint64_t q1 = q0 + BP_PER_LANE < a.npairs ? q0 + BP_PER_LANE : a.npairs;
    int64_t A[BWD_ACC];
#pragma unroll
    for (int t = 0; t < BWD_ACC; t++)
        A[t] = 0;
    uint32_t cur = BP_EMPTY;
    BwdPrim p;
    uint64_t leaf_cur = ~0ull;
    uint32_t lord = LT_NONE;
    int32_t li = 0, lj = 0, lk = 0;
#pragma unroll 1
    for (uint64_t q = q0; q < q1; q++)
    {
        const uint64_t key = a.keys[q];
        const uint32_t g = (uint32_t)(key & gmask);
        const uint64_t vox = key >> gb;
        if (g != cur)
        {
            if (cur != BP_EMPTY)
            {
                bp_flush(a, s_gid, s_acc, cur, A);
#pragma unroll
                for (int t = 0; t < BWD_ACC; t++)
                    A[t] = 0;
            }
            cur = g;
            p = a.prims[g];
        }
        const uint64_t leaf = vox >> 9;
        if (leaf != leaf_cur)
        {
            int32_t i0, j0, k0;
            decode_voxel(a.kl, leaf_of_field(a.kl, leaf) << 9, &i0, &j0, &k0);
            li = i0 >> 3;
            lj = j0 >> 3;
            lk = k0 >> 3;
            lord = lt_find(a.lt_slots, a.lt_mask, li, lj, lk);
            leaf_cur = leaf;
        }
        const uint32_t nv = (uint32_t)vox & 511u;
        // dL/drho pre-multiplied by the power-of-two scale (exact)
        const double r = lord != LT_NONE ? (double)a.leaf_grad[(uint64_t)lord * 512u + nv] * scale : 0.0;
        const double d[3] = { ((double)(li * 8 + (int32_t)(nv >> 6)) + 0.5) * vs - p.mu[0],
                              ((double)(lj * 8 + (int32_t)((nv >> 3) & 7u)) + 0.5) * vs - p.mu[1],
                              ((double)(lk * 8 + (int32_t)(nv & 7u)) + 0.5) * vs - p.mu[2] };
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            v[c] = p.M[3 * c + 0] * d[0] + p.M[3 * c + 1] * d[1] + p.M[3 * c + 2] * d[2];
        const double m = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        const double e = exp_neg(-0.5 * m, s_exp2);
        const double W = r * p.o * e;
        const double w0 = W * v[0], w1 = W * v[1], w2 = W * v[2];
        A[0] += bp_round(r * e);
        A[1] += bp_round(w0);
        A[2] += bp_round(w1);
        A[3] += bp_round(w2);
        A[4] += bp_round(w0 * v[0]);
        A[5] += bp_round(w1 * v[1]);
        A[6] += bp_round(w2 * v[2]);
        A[7] += bp_round(w0 * v[1]);
        A[8] += bp_round(w0 * v[2]);
        A[9] += bp_round(w1 * v[2]);
    }
    if (cur != BP_EMPTY)
        bp_flush(a, s_gid, s_acc, cur, A);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < BP_SLOTS * BWD_ACC; e += BP_THREADS)
    {
        const uint32_t slot = e / BWD_ACC, t = e - slot * BWD_ACC;
        const uint32_t g = s_gid[slot];
        const int64_t v = (int64_t)s_acc[t][slot];
        if (g != BP_EMPTY && v != 0)
            atomic_add_i128(a.sums + ((uint64_t)g * BWD_ACC + t) * 2u, v);
    }
}

// per Gaussian: the 128-bit sums back to doubles, then the chain rule
__global__ __launch_bounds__(256) void k_density_bwd_pairs_finalize(BwdArgs a, const unsigned long long* __restrict__ sums,
                                                                     const unsigned int* __restrict__ rmax_bits)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n)
        return;
    const int s = bp_scale_exp(*rmax_bits);
    double A[BWD_ACC];
#pragma unroll
    for (int t = 0; t < BWD_ACC; t++)
    {
        const uint64_t lo = sums[(g * BWD_ACC + t) * 2u];
        const uint64_t hi = sums[(g * BWD_ACC + t) * 2u + 1u];
        // |sum| < 2^80 (at most 2^32 terms of magnitude 2^48): sum >> 40 has at most 41 bits, and
        // the double of top * 2^40 + low rounds once
        const int64_t top = (int64_t)((hi << 24) + (lo >> 40));
        const double d = (double)top * 0x1p40 + (double)(lo & ((1ull << 40) - 1ull));
        A[t] = ldexp(d, -s);
    }
    bwd_write_grads(a, g, A);
}

// max over x[0..n) of |x| as float bits (a NaN compares above infinity)
__global__ __launch_bounds__(256) void k_abs_max_bits(const float* __restrict__ x, uint64_t n, unsigned int* out)
{
    __shared__ uint32_t s_max[4];
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        m = max(m, __float_as_uint(x[i]) & 0x7FFFFFFFu);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63u) == 0u)
        s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(out, max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3])));
}

// ============================================================================ alpha self-test
// Compares alpha8_fast with the correctly rounded alpha8_from_dist2 on hashed inputs that cover
// the ranges the value pass produces: opacity in (0.01, 1), r = sqrt(2 log(o / 0.01)) as in
// setup, squared sample distance in [0, 4), pow exponent in [0.02, 48], about 1/8 culled.
__device__ __forceinline__ uint32_t hash32(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

__global__ void k_selftest_alpha8(uint64_t count, uint64_t seed, unsigned long long* mism,
                                  unsigned long long* fast_taken)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long bad = 0ull, fast = 0ull;
    for (; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint64_t h = seed * 0x9E3779B97F4A7C15ull + i;
        const uint32_t h0 = hash32(h), h1 = hash32(h + 0x1234567ull), h2 = hash32(h + 0x89abcdefull);
        PrimVal pv;
        pv.opacity = 0.0100001f + 0.9899999f * ((float)(h0 >> 8) * (1.f / 16777216.f));
        pv.local_scale = fsqrt_rn(2.f * vdb_logf(fdiv_rn(pv.opacity, 0.01f)));
        pv.pow_exp = 0.02f + 48.f * ((float)(h1 >> 8) * (1.f / 16777216.f));
        pv.rgb_cull = (h2 & 0x00FFFFFFu) | (((h2 >> 29) == 0u) ? (1u << 24) : 0u);
        const float d = 4.f * ((float)(hash32(h + 0x5555ull) >> 8) * (1.f / 16777216.f));
        const uint32_t ref = alpha8_from_dist2(pv, d);
        uint32_t got;
        if (alpha8_try_fast(pv, d, &got))
        {
            fast += 1ull;
            bad += got != ref ? 1ull : 0ull;
        }
    }
    atomicAdd(mism, bad);
    atomicAdd(fast_taken, fast);
}

int selftest_alpha8(Context* ctx, uint64_t count, uint64_t seed, uint64_t* mismatches, uint64_t* decided)
{
    unsigned long long* d = (unsigned long long*)ctx->ws(WS_MISC, 4096);
    if (!d)
        return -1;
    unsigned long long* cnt = d + 32;
    VDBR_CHECK(ctx, hipMemsetAsync(cnt, 0, 16, ctx->stream));
    hipLaunchKernelGGL(k_selftest_alpha8, dim3(4096), dim3(256), 0, ctx->stream, count, seed, cnt, cnt + 1);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, cnt, 16, hipMemcpyDeviceToHost, ctx->stream));
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *mismatches = ctx->pinned[0];
    if (decided)
        *decided = ctx->pinned[1];
    return 0;
}

// ============================================================================ run heads
constexpr int HEAD_THREADS = 256;
constexpr int HEAD_ITEMS = 16;
constexpr int HEAD_TILE = HEAD_THREADS * HEAD_ITEMS;

// counts positions i where (key[i] >> shift) differs from (key[i-1] >> shift)
__global__ __launch_bounds__(HEAD_THREADS) void k_head_count(const uint64_t* __restrict__ keys, DCount cn, uint32_t shift,
                                                             uint32_t* __restrict__ tile_counts)
{
    const uint64_t n = cn.get();
    const uint64_t tiles = (n + HEAD_TILE - 1) / HEAD_TILE;
    __shared__ uint32_t cnt;
    // blocks stride over the tiles (a grid sized by a capacity stays bounded)
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        if (threadIdx.x == 0)
            cnt = 0u;
        __syncthreads();
        const uint64_t base = tile * HEAD_TILE;
        uint32_t local = 0u;
        for (int m = 0; m < HEAD_ITEMS; m++)
        {
            uint64_t idx = base + (uint64_t)m * HEAD_THREADS + threadIdx.x;
            if (idx < n)
            {
                uint64_t v = keys[idx] >> shift;
                bool head = idx == 0 || (keys[idx - 1] >> shift) != v;
                local += head ? 1u : 0u;
            }
        }
        atomicAdd(&cnt, local);
        __syncthreads();
        if (threadIdx.x == 0)
            tile_counts[tile] = cnt;
    }
}

// Same tile counts as k_head_count when key >> shift is non-decreasing (keys sorted over every
// bit from shift up): a 64-key segment whose last value equals the value just before it holds no
// head, so each lane reads one segment end and the wave reads in full only the segments whose ends
// differ, after a first check of the whole tile. One wave per HEAD_TILE tile.
constexpr int HEAD_SEG = 64;
static_assert(HEAD_TILE == 64 * HEAD_SEG, "one segment per lane");
__global__ __launch_bounds__(256) void k_head_count_sorted(const uint64_t* __restrict__ keys, DCount cn, uint32_t shift,
                                                           uint32_t* __restrict__ tile_counts)
{
    const uint64_t n = cn.get();
    const uint64_t tiles = (n + HEAD_TILE - 1) / HEAD_TILE;
    const uint32_t lane = threadIdx.x & 63u;
    // waves stride over the tiles (a grid sized by a capacity stays bounded)
    for (uint64_t tile = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); tile < tiles; tile += (uint64_t)gridDim.x * 4u)
    {
        const uint64_t t0 = tile * HEAD_TILE;
        const uint64_t t1 = t0 + HEAD_TILE < n ? t0 + HEAD_TILE : n;
        // a tile whose last value equals the value before it holds no head (most tiles of long runs)
        if (t0 > 0 && (keys[t0 - 1] >> shift) == (keys[t1 - 1] >> shift))
        {
            if (lane == 0u)
                tile_counts[tile] = 0u;
            continue; // wave-uniform
        }
        const uint64_t s = t0 + (uint64_t)lane * HEAD_SEG;
        const bool live = s < t1;
        const uint64_t e = s + HEAD_SEG < t1 ? s + HEAD_SEG : t1;
        const uint64_t vend = live ? keys[e - 1] >> shift : 0ull;
        // value just before each segment: the previous lane's segment end, key[t0 - 1] for lane 0
        uint64_t vbefore = wave_shr1_u64(0ull, vend);
        if (lane == 0u && t0 > 0)
            vbefore = keys[t0 - 1] >> shift;
        const bool first = lane == 0u && t0 == 0; // position 0 is always a head
        uint64_t todo = __ballot(live && (first || vbefore != vend));
        uint32_t count = 0u;
        while (todo)
        {
            const int l = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            const uint64_t ss = t0 + (uint64_t)l * HEAD_SEG;
            const uint64_t idx = ss + lane;
            const bool valid = idx < t1;
            const uint64_t v = valid ? keys[idx] >> shift : 0ull;
            uint64_t pv = wave_shr1_u64(0ull, v);
            const uint64_t before_l = readlane_u64(vbefore, l);
            if (lane == 0u)
                pv = before_l;
            const bool head = valid && (idx == 0 || v != pv);
            count += (uint32_t)__popcll(__ballot(head));
        }
        if (lane == 0u)
            tile_counts[tile] = count;
    }
}

// run starts of sorted values in order (starts[rank] = index of the run head), visiting only the
// 64-key segments that hold a head, as k_head_count_sorted counts them; tile_offsets is the
// exclusive scan of those counts
__global__ __launch_bounds__(256) void k_head_write_sorted(const uint64_t* __restrict__ keys, DCount cn, uint32_t shift,
                                                           const uint32_t* __restrict__ tile_offsets,
                                                           uint32_t* __restrict__ starts)
{
    const uint64_t n = cn.get();
    const uint64_t tiles = (n + HEAD_TILE - 1) / HEAD_TILE;
    const uint32_t lane = threadIdx.x & 63u;
    // waves stride over the tiles (a grid sized by a capacity stays bounded)
    for (uint64_t tile = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); tile < tiles; tile += (uint64_t)gridDim.x * 4u)
    {
        const uint64_t t0 = tile * HEAD_TILE;
        const uint64_t t1 = t0 + HEAD_TILE < n ? t0 + HEAD_TILE : n;
        if (t0 > 0 && (keys[t0 - 1] >> shift) == (keys[t1 - 1] >> shift))
            continue; // wave-uniform: no head in this tile
        const uint64_t s = t0 + (uint64_t)lane * HEAD_SEG;
        const bool live = s < t1;
        const uint64_t e = s + HEAD_SEG < t1 ? s + HEAD_SEG : t1;
        const uint64_t vend = live ? keys[e - 1] >> shift : 0ull;
        uint64_t vbefore = wave_shr1_u64(0ull, vend);
        if (lane == 0u && t0 > 0)
            vbefore = keys[t0 - 1] >> shift;
        const bool first = lane == 0u && t0 == 0;
        uint64_t todo = __ballot(live && (first || vbefore != vend));
        uint32_t rank = tile_offsets[tile];
        const uint64_t lt_mask = (1ull << lane) - 1ull;
        while (todo)
        {
            const int l = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            const uint64_t idx = t0 + (uint64_t)l * HEAD_SEG + lane;
            const bool valid = idx < t1;
            const uint64_t v = valid ? keys[idx] >> shift : 0ull;
            uint64_t pv = wave_shr1_u64(0ull, v);
            const uint64_t before_l = readlane_u64(vbefore, l);
            if (lane == 0u)
                pv = before_l;
            const bool head = valid && (idx == 0 || v != pv);
            const uint64_t hb = __ballot(head);
            if (head)
                starts[rank + (uint32_t)__popcll(hb & lt_mask)] = (uint32_t)idx;
            rank += (uint32_t)__popcll(hb);
        }
    }
}

// writes run starts in order: starts[rank] = index of the run head
__global__ __launch_bounds__(HEAD_THREADS) void k_head_write(const uint64_t* __restrict__ keys, DCount cn, uint32_t shift,
                                                             const uint32_t* __restrict__ tile_offsets,
                                                             uint32_t* __restrict__ starts)
{
    const uint64_t n = cn.get();
    const uint64_t tiles = (n + HEAD_TILE - 1) / HEAD_TILE;
    __shared__ uint32_t wave_tot[HEAD_THREADS / 64];
    __shared__ uint32_t run_base;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const uint64_t base = tile * HEAD_TILE;
        if (threadIdx.x == 0)
            run_base = tile_offsets[tile];
        __syncthreads();
        for (int m = 0; m < HEAD_ITEMS; m++)
        {
            uint64_t idx = base + (uint64_t)m * HEAD_THREADS + threadIdx.x;
            bool head = false;
            if (idx < n)
            {
                uint64_t v = keys[idx] >> shift;
                head = idx == 0 || (keys[idx - 1] >> shift) != v;
            }
            uint64_t bal = __ballot(head);
            if (lane == 0)
                wave_tot[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t pre = run_base;
            for (uint32_t w = 0; w < wave; w++)
                pre += wave_tot[w];
            if (head)
                starts[pre + (uint32_t)__popcll(bal & lt_mask)] = (uint32_t)idx;
            __syncthreads();
            if (threadIdx.x == 0)
            {
                uint32_t t = 0u;
                for (uint32_t w = 0; w < HEAD_THREADS / 64; w++)
                    t += wave_tot[w];
                run_base += t;
            }
            __syncthreads();
        }
    }
}

// Block shape of the tile reduces: one run-head tile of HEAD_TILE pairs per block, RED_ITEMS
// consecutive pairs per lane.
constexpr int RED_THREADS = 256;
constexpr int RED_ITEMS = HEAD_TILE / RED_THREADS; // one run-head tile per block

// ============================================================================ RGBA8 reduce by leaf
// With the RGBA8 alpha in the key, the pairs are sorted by their leaf bits only (the 9 in-leaf
// slot bits and the Gaussian id stay unsorted): each voxel's value, the maximum packed
// r8 | g8 << 8 | b8 << 16 | a8 << 24 over its pairs, and its lowest Gaussian id are independent
// of the pair order. One block per tile of HEAD_TILE pairs folds its pairs into LDS slots of up to
// LR_LEAVES leaves (pass 1: maximum alpha and lowest id per slot; pass 2: colour gathers only
// for pairs holding their slot's maximum alpha), then merges the slots into the leaf staging
// arrays with atomicMax / atomicMin (order independent). Leaves past LR_LEAVES in one tile go
// straight to the staging arrays. A count pass, a scan and a write pass then list the active
// voxels in key order.
constexpr int LR_LEAVES = 8;

template <bool ALPHA_IN_KEY>
__global__ __launch_bounds__(RED_THREADS) void k_leaf_reduce_rgba(const uint64_t* __restrict__ keys, DCount cn,
                                                                  KeyLayout kl, float vs,
                                                                  const uint32_t* __restrict__ tile_offsets,
                                                                  const PrimFrame* __restrict__ frames,
                                                                  const PrimVal* __restrict__ vals,
                                                                  uint32_t* __restrict__ lval, uint32_t* __restrict__ lg0,
                                                                  uint64_t* __restrict__ lkey)
{
    const uint32_t gid_bits = kl.gid_bits;
    const uint32_t ab = ALPHA_IN_KEY ? 8u : 0u;
    const uint32_t lshift = gid_bits + ab + 9u;
    __shared__ uint32_t s_val[LR_LEAVES * 512];
    __shared__ uint32_t s_g0[LR_LEAVES * 512];
    __shared__ uint32_t wave_heads[RED_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t gmask = (1ull << gid_bits) - 1ull;
    const uint64_t n = cn.get();
    const uint64_t t0 = (uint64_t)blockIdx.x * HEAD_TILE;
    if (t0 >= n)
        return; // block-uniform: the grid may be sized by a capacity
    const uint64_t pb = t0 + (uint64_t)threadIdx.x * RED_ITEMS;
    const uint32_t tile_base = tile_offsets[blockIdx.x];
    // the tile's first pair continues a leaf of the previous tile
    const bool first_cont = t0 > 0 && (keys[t0] >> lshift) == (keys[t0 - 1] >> lshift);
    for (uint32_t x = threadIdx.x; x < (uint32_t)(LR_LEAVES * 512); x += RED_THREADS)
    {
        s_val[x] = 0u;
        s_g0[x] = 0xFFFFFFFFu;
    }
    uint64_t key[RED_ITEMS];
    if (pb + RED_ITEMS <= n)
    {
        const ulonglong2* src = (const ulonglong2*)(keys + pb);
#pragma unroll
        for (int i = 0; i < RED_ITEMS / 2; i++)
        {
            ulonglong2 v = src[i];
            key[2 * i] = v.x;
            key[2 * i + 1] = v.y;
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < RED_ITEMS; i++)
            key[i] = pb + i < n ? keys[pb + i] : ~0ull;
    }
    // leaf heads: positions whose leaf differs from the previous pair's
    uint64_t prev = wave_shr1_u64(0ull, key[RED_ITEMS - 1]);
    if (lane == 0u)
        prev = (pb > 0 && pb < n) ? keys[pb - 1] : ~0ull;
    prev >>= lshift;
    uint32_t headbits = 0u;
    {
        uint64_t pv = prev;
#pragma unroll
        for (int i = 0; i < RED_ITEMS; i++)
        {
            const uint64_t lk = key[i] >> lshift;
            if (pb + i < n && (pb + i == 0 || lk != pv))
                headbits |= 1u << i;
            pv = lk;
        }
    }
    uint32_t wtot;
    const uint32_t excl = wave_excl_add((uint32_t)__popc(headbits), lane, &wtot);
    if (lane == 0u)
        wave_heads[wave] = wtot;
    __syncthreads();
    uint32_t before = excl, tile_heads = 0u;
    for (uint32_t w = 0; w < RED_THREADS / 64u; w++)
    {
        before += w < wave ? wave_heads[w] : 0u;
        tile_heads += wave_heads[w];
    }
    const uint32_t cont = first_cont ? 1u : 0u;
    // pass 1: maximum alpha (alpha in the key) or maximum packed value (evaluated here from the
    // voxel box, as the emission evaluates it) and lowest Gaussian id per slot; leaves past
    // LR_LEAVES go to the staging arrays directly (with their colours)
    uint32_t h = before;
#pragma unroll
    for (int i = 0; i < RED_ITEMS; i++)
    {
        if ((headbits >> i) & 1u)
            h++;
        if (pb + i >= n)
            continue;
        const uint32_t gl = tile_base + h - 1u; // leaf index of the pair
        const uint32_t ord = h - 1u + cont;     // leaf ordinal within the tile
        const uint32_t slot = (uint32_t)(key[i] >> (lshift - 9u)) & 511u;
        const uint32_t gid = (uint32_t)((key[i] >> ab) & gmask);
        uint32_t a8 = 0u, rgba = 0u;
        if (ALPHA_IN_KEY)
            a8 = (uint32_t)(key[i] & 255u);
        else
        {
            int32_t vi, vj, vk;
            decode_pair_voxel(kl, key[i] >> gid_bits, &vi, &vj, &vk);
            float vmin[3], vmax[3];
            voxel_box(frames[gid], vi, vj, vk, vs, vmin, vmax);
            const PrimVal pv = vals[gid];
            rgba = (pv.rgb_cull & 0x00FFFFFFu) | (alpha8_fast(pv, min_sample_dist2(vmin, vmax)) << 24);
        }
        if ((headbits >> i) & 1u)
            lkey[gl] = leaf_of_field(kl, key[i] >> lshift);
        if (ord < (uint32_t)LR_LEAVES)
        {
            atomicMax(&s_val[ord * 512u + slot], ALPHA_IN_KEY ? a8 << 24 : rgba);
            atomicMin(&s_g0[ord * 512u + slot], gid);
        }
        else
        {
            atomicMax(&lval[(uint64_t)gl * 512u + slot],
                      ALPHA_IN_KEY ? (vals[gid].rgb_cull & 0x00FFFFFFu) | (a8 << 24) : rgba);
            atomicMin(&lg0[(uint64_t)gl * 512u + slot], gid);
        }
    }
    __syncthreads();
    // pass 2 (alpha in the key): colours of the pairs that hold their slot's maximum alpha
    h = before;
#pragma unroll
    for (int i = 0; ALPHA_IN_KEY && i < RED_ITEMS; i++)
    {
        if ((headbits >> i) & 1u)
            h++;
        const uint32_t ord = h - 1u + cont;
        if (pb + i >= n || ord >= (uint32_t)LR_LEAVES)
            continue;
        const uint32_t slot = (uint32_t)(key[i] >> (lshift - 9u)) & 511u;
        const uint32_t a8 = (uint32_t)(key[i] & 255u);
        if ((s_val[ord * 512u + slot] >> 24) == a8)
        {
            const uint32_t gid = (uint32_t)((key[i] >> 8) & gmask);
            atomicMax(&s_val[ord * 512u + slot], (vals[gid].rgb_cull & 0x00FFFFFFu) | (a8 << 24));
        }
    }
    __syncthreads();
    // merge the LDS leaves into the staging arrays
    const uint32_t leaves = tile_heads + cont;
    const uint32_t nl = leaves < (uint32_t)LR_LEAVES ? leaves : (uint32_t)LR_LEAVES;
    const uint32_t gbase = tile_base - cont;
    for (uint32_t x = threadIdx.x; x < nl * 512u; x += RED_THREADS)
    {
        const uint32_t g = s_g0[x];
        if (g != 0xFFFFFFFFu)
        {
            const uint64_t at = (uint64_t)(gbase + x / 512u) * 512u + (x & 511u);
            atomicMax(&lval[at], s_val[x]);
            atomicMin(&lg0[at], g);
        }
    }
}

// RGBA8 by leaf chunks (the alpha in the key): one block per chunk of at most LRC_CHUNK pairs of
// one leaf, the chunks listed from the leaf starts as for the density reduce. Pass 1 folds the
// maximum alpha and the lowest Gaussian id of every slot into LDS; pass 2 takes the same keys
// from registers (16 per thread) and gathers colours only for pairs holding their slot's maximum
// alpha. A leaf of
// one chunk (most leaves) stores its 512 slots with plain stores; the chunks of a longer leaf
// merge into the prefilled staging arrays with atomicMax / atomicMin (order independent).
// 4096 pairs per 256-thread block: 3.65 ms at 2M, against 4.10 ms for 16384-pair chunks read twice
// and 5.43 ms for 16384 held by 1024 threads (small leaves leave large blocks idle)
constexpr uint32_t LRC_CHUNK = 4096;
constexpr int LRC_THREADS = 256;
constexpr int LRC_KEYS = (int)LRC_CHUNK / LRC_THREADS; // per thread, held in registers for both passes

__global__ __launch_bounds__(LRC_THREADS) void k_leaf_reduce_rgba_chunks(const uint64_t* __restrict__ keys, KeyLayout kl,
                                                                         const uint32_t* __restrict__ starts,
                                                                         const uint32_t* __restrict__ item_scan,
                                                                         DCount cnleaf, const PrimVal* __restrict__ vals,
                                                                         uint32_t* __restrict__ lval,
                                                                         uint32_t* __restrict__ lg0,
                                                                         uint64_t* __restrict__ lkey)
{
    __shared__ uint32_t s_val[512];
    __shared__ uint32_t s_g0[512];
    __shared__ uint32_t s_leaf;
    const uint64_t nleaf = cnleaf.get();
    const uint32_t item = blockIdx.x;
    if (item >= item_scan[nleaf])
        return; // block-uniform: the grid is an upper bound of the chunk count
    if (threadIdx.x < 64u)
    {
        // the leaf of this chunk: largest l with item_scan[l] <= item
        const uint32_t l = (uint32_t)wave_find_prim(item_scan, nleaf, item, threadIdx.x);
        if (threadIdx.x == 0)
            s_leaf = l;
    }
    for (uint32_t x = threadIdx.x; x < 512u; x += LRC_THREADS)
    {
        s_val[x] = 0u;
        s_g0[x] = 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t l = s_leaf;
    const uint32_t chunk = item - item_scan[l];
    const uint32_t nchunks = item_scan[l + 1] - item_scan[l];
    const uint64_t first = starts[l];
    const uint64_t p0 = first + (uint64_t)chunk * LRC_CHUNK;
    const uint64_t pend = starts[l + 1];
    const uint64_t p1 = p0 + LRC_CHUNK < pend ? p0 + LRC_CHUNK : pend;
    const uint32_t gid_shift = 8u;
    const uint32_t slot_shift = kl.gid_bits + 8u;
    const uint64_t gmask = (1ull << kl.gid_bits) - 1ull;
    if (chunk == 0u && threadIdx.x == 0)
        lkey[l] = leaf_of_field(kl, keys[first] >> (slot_shift + 9u));
    // the chunk's keys, LRC_KEYS per thread, read once
    uint64_t k[LRC_KEYS];
#pragma unroll
    for (int u = 0; u < LRC_KEYS; u++)
    {
        const uint64_t q = p0 + (uint64_t)u * LRC_THREADS + threadIdx.x;
        k[u] = q < p1 ? keys[q] : ~0ull;
    }
    // pass 1: maximum alpha and lowest id per slot
#pragma unroll
    for (int u = 0; u < LRC_KEYS; u++)
    {
        if (p0 + (uint64_t)u * LRC_THREADS + threadIdx.x < p1)
        {
            const uint32_t slot = (uint32_t)(k[u] >> slot_shift) & 511u;
            atomicMax(&s_val[slot], (uint32_t)(k[u] & 255u) << 24);
            atomicMin(&s_g0[slot], (uint32_t)((k[u] >> gid_shift) & gmask));
        }
    }
    __syncthreads();
    // pass 2 (same keys, from registers): colours of the pairs holding their slot's maximum alpha
#pragma unroll
    for (int u = 0; u < LRC_KEYS; u++)
    {
        if (p0 + (uint64_t)u * LRC_THREADS + threadIdx.x < p1)
        {
            const uint32_t slot = (uint32_t)(k[u] >> slot_shift) & 511u;
            const uint32_t a8 = (uint32_t)(k[u] & 255u);
            if ((s_val[slot] >> 24) == a8)
            {
                const uint32_t gid = (uint32_t)((k[u] >> gid_shift) & gmask);
                atomicMax(&s_val[slot], (vals[gid].rgb_cull & 0x00FFFFFFu) | (a8 << 24));
            }
        }
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < 512u; x += LRC_THREADS)
    {
        const uint64_t at = (uint64_t)l * 512u + x;
        if (nchunks == 1u)
        {
            lval[at] = s_val[x];
            lg0[at] = s_g0[x];
        }
        else if (s_g0[x] != 0xFFFFFFFFu)
        {
            atomicMax(&lval[at], s_val[x]);
            atomicMin(&lg0[at], s_g0[x]);
        }
    }
}

// one wave per leaf: number of voxels with a pair
__global__ __launch_bounds__(256) void k_leaf_count(const uint32_t* __restrict__ lg0, DCount cnleaf,
                                                    uint32_t* __restrict__ cnt)
{
    const uint64_t nleaf = cnleaf.get();
    const uint64_t leaf = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
    if (leaf >= nleaf)
        return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint4 a = *(const uint4*)(lg0 + leaf * 512u + lane * 8u);
    const uint4 b = *(const uint4*)(lg0 + leaf * 512u + lane * 8u + 4u);
    uint32_t c = (a.x != 0xFFFFFFFFu) + (a.y != 0xFFFFFFFFu) + (a.z != 0xFFFFFFFFu) + (a.w != 0xFFFFFFFFu) +
                 (b.x != 0xFFFFFFFFu) + (b.y != 0xFFFFFFFFu) + (b.z != 0xFFFFFFFFu) + (b.w != 0xFFFFFFFFu);
    for (int off = 32; off > 0; off >>= 1)
        c += (uint32_t)__shfl_xor((int)c, off, 64);
    if (lane == 0u)
        cnt[leaf] = c;
}

// one wave per leaf: its voxels with a pair, in slot (= key) order, at the leaf's offset
// Writes one leaf's active voxels (slots holding a lowest Gaussian id) at base in slot order. Each
// lane stages its up to 8 slots in the wave's LDS rows at its prefix position lpos, then the wave
// stores 64 consecutive entries per instruction: written from the lanes' own runs, every store
// would spread over many partly written lines.
__device__ __forceinline__ void leaf_write_staged(uint32_t lane, const uint32_t g[8], const uint32_t v[8], uint64_t kbase,
                                                  uint64_t base, uint32_t lpos, uint32_t wtot, uint64_t* s_key,
                                                  uint32_t* s_val, uint32_t* s_g0, uint64_t* __restrict__ vox_key,
                                                  uint32_t* __restrict__ vox_val, uint32_t* __restrict__ vox_g0)
{
#pragma unroll
    for (int t = 0; t < 8; t++)
    {
        if (g[t] != 0xFFFFFFFFu)
        {
            s_key[lpos] = kbase | (uint64_t)(lane * 8u + (uint32_t)t);
            s_val[lpos] = v[t];
            s_g0[lpos] = g[t];
            lpos++;
        }
    }
    wave_sync_lds();
    for (uint32_t i = lane; i < wtot; i += 64u)
    {
        vox_key[base + i] = s_key[i];
        vox_val[base + i] = s_val[i];
        vox_g0[base + i] = s_g0[i];
    }
}

__global__ __launch_bounds__(256) void k_leaf_write(const uint32_t* __restrict__ lval, const uint32_t* __restrict__ lg0,
                                                    const uint64_t* __restrict__ lkey,
                                                    const uint32_t* __restrict__ lscan, DCount cnleaf,
                                                    uint64_t* __restrict__ vox_key, uint32_t* __restrict__ vox_val,
                                                    uint32_t* __restrict__ vox_g0)
{
    __shared__ uint64_t s_key[4][512];
    __shared__ uint32_t s_val[4][512];
    __shared__ uint32_t s_g0[4][512];
    const uint64_t nleaf = cnleaf.get();
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t leaf = (uint64_t)blockIdx.x * 4u + wave;
    if (leaf >= nleaf)
        return; // wave-uniform; the kernel has no block barrier
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t at = leaf * 512u + lane * 8u;
    uint32_t g[8], v[8];
    {
        const uint4 a = *(const uint4*)(lg0 + at);
        const uint4 b = *(const uint4*)(lg0 + at + 4u);
        g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
        g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
        const uint4 c = *(const uint4*)(lval + at);
        const uint4 d = *(const uint4*)(lval + at + 4u);
        v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w;
        v[4] = d.x; v[5] = d.y; v[6] = d.z; v[7] = d.w;
    }
    uint32_t cnt = 0u;
#pragma unroll
    for (int t = 0; t < 8; t++)
        cnt += g[t] != 0xFFFFFFFFu ? 1u : 0u;
    uint32_t wtot;
    const uint32_t lpos = wave_excl_add(cnt, lane, &wtot);
    leaf_write_staged(lane, g, v, lkey[leaf] << 9, lscan[leaf], lpos, wtot, s_key[wave], s_val[wave], s_g0[wave],
                      vox_key, vox_val, vox_g0);
}

// ============================================================================ density reduce
// Density values per voxel: rho(v) = sum over the voxel's pairs of o * exp(-m / 2), m the squared
// Mahalanobis distance of the voxel centre, each term in double from the preprocessed float
// parameters (rotation from the normalised quaternion, scales floored at vs / 8 as in
// gaussian_prim.slang:67-68). Terms are rounded once to multiples of 2^-120 and summed as
// integers (exact, so the sum does not depend on the pair order), then rounded to float.
struct DensPrim
{
    double M[9]; // diag(1/s) R^T, row-major
    double o;
    float mu[3];
    float pad;
    double pad2[4]; // 128 bytes: one cache line per gather
};

// per Gaussian: the double frame of its density terms
__global__ __launch_bounds__(256) void k_dens_prims(uint64_t n, const float* __restrict__ means,
                                                    const float* __restrict__ quats, const float* __restrict__ scales,
                                                    const float* __restrict__ opacities, float vs,
                                                    DensPrim* __restrict__ out)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const double w = quats[4 * g + 0], x = quats[4 * g + 1], y = quats[4 * g + 2], z = quats[4 * g + 3];
    const double R[3][3] = { { 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w) },
                             { 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w) },
                             { 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y) } };
    const double fl = (double)(0.125f * vs);
    DensPrim p;
    for (int c = 0; c < 3; c++)
    {
        const double sc = fmax((double)scales[3 * g + c], fl);
        for (int b = 0; b < 3; b++)
            p.M[3 * c + b] = R[b][c] / sc;
        p.mu[c] = means[3 * g + c];
    }
    p.o = (double)opacities[g];
    p.pad = 0.f;
    for (int q = 0; q < 4; q++)
        p.pad2[q] = 0.0;
    out[g] = p;
}

// t in [0, 2^8) -> round(t * 2^120) (half up) as a 128-bit integer
__device__ __forceinline__ void fix120(double t, uint64_t* lo, uint64_t* hi)
{
    *lo = 0ull;
    *hi = 0ull;
    const uint64_t bits = (uint64_t)__double_as_longlong(t);
    const int32_t be = (int32_t)((bits >> 52) & 0x7FFu);
    if (be == 0 || (bits >> 63)) // zero, subnormal or negative: nothing to add
        return;
    const uint64_t m = (bits & ((1ull << 52) - 1ull)) | (1ull << 52); // t = m * 2^(be - 1075)
    const int32_t sh = be - 1075 + 120;                               // position of m's bit 0
    if (sh < 0)
    {
        if (sh >= -53)
        {
            const int32_t r = -sh;
            *lo = (m + (1ull << (r - 1))) >> r;
        }
    }
    else if (sh < 64)
    {
        *lo = m << sh;
        *hi = sh ? (m >> (64 - sh)) : 0ull;
    }
    else if (sh < 76)
        *hi = m << (sh - 64);
}

// w += (a2:a1:a0) with the carries taken from the returned old words: every wrap of a word is
// seen by exactly one add, so the total is exact whatever the order of concurrent adds
__device__ __forceinline__ void atomic_add192(unsigned long long* w, uint64_t a0, uint64_t a1, uint64_t a2)
{
    const uint64_t o0 = atomicAdd(&w[0], (unsigned long long)a0);
    const uint64_t k0 = o0 + a0 < o0 ? 1ull : 0ull;
    const uint64_t b1 = a1 + k0;
    uint64_t k1 = b1 < a1 ? 1ull : 0ull;
    const uint64_t o1 = atomicAdd(&w[1], (unsigned long long)b1);
    k1 += o1 + b1 < o1 ? 1ull : 0ull;
    atomicAdd(&w[2], (unsigned long long)(a2 + k1));
}

// (w2:w1:w0) * 2^-120 -> nearest float of its double approximation (deterministic)
__device__ __forceinline__ float fix120_to_float(uint64_t w0, uint64_t w1, uint64_t w2)
{
    const double v = (double)w2 * 0x1p8 + (double)w1 * 0x1p-56 + (double)w0 * 0x1p-120;
    return (float)v;
}

// Density values by leaf: the pairs are sorted by their leaf bits only and each voxel's sum of
// fixed-point terms is exact integer arithmetic, so the pair order inside a leaf does not matter.
// One block per chunk of at most LRD_CHUNK pairs of one leaf adds its terms into LDS slots
// (192-bit accumulators, 64-bit atomic adds carrying between words). A leaf of one chunk (most of
// them) is stored with plain stores; the chunks of a longer leaf merge into the zeroed staging
// arrays with the same carrying atomic adds.
constexpr uint32_t LRD_CHUNK = 16384;

__device__ __forceinline__ void add192(unsigned long long* w, uint64_t a0, uint64_t a1, uint64_t a2)
{
    atomic_add192(w, a0, a1, a2);
}

// chunks of every leaf (starts holds nleaf + 1 entries, the last one the pair count)
__global__ __launch_bounds__(256) void k_leaf_chunks(const uint32_t* __restrict__ starts, DCount cnleaf,
                                                     uint32_t chunk, uint32_t* __restrict__ items)
{
    const uint64_t nleaf = cnleaf.get();
    const uint64_t l = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (l < nleaf)
        items[l] = (starts[l + 1] - starts[l] + chunk - 1u) / chunk;
}

__global__ void k_store_u32(uint32_t* p, uint32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

__global__ __launch_bounds__(RED_THREADS) void k_leaf_reduce_density(const uint64_t* __restrict__ keys, KeyLayout kl,
                                                                     float vs, const uint32_t* __restrict__ starts,
                                                                     const uint32_t* __restrict__ item_scan,
                                                                     DCount cnleaf, const DensPrim* __restrict__ prims,
                                                                     unsigned long long* __restrict__ lw,
                                                                     uint32_t* __restrict__ lg0,
                                                                     uint64_t* __restrict__ lkey)
{
    const uint64_t nleaf = cnleaf.get();
    __shared__ unsigned long long s_w[512 * 3];
    __shared__ uint32_t s_g0[512];
    __shared__ double s_exp2[64];
    __shared__ uint32_t s_leaf;
    const uint32_t item = blockIdx.x;
    if (item >= item_scan[nleaf])
        return; // block-uniform: the grid is an upper bound of the chunk count
    if (threadIdx.x == 0)
    {
        // the leaf of this chunk: largest l with item_scan[l] <= item
        uint32_t lo = 0, hi = (uint32_t)nleaf - 1u;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (item_scan[mid] <= item)
                lo = mid;
            else
                hi = mid - 1u;
        }
        s_leaf = lo;
    }
    for (uint32_t x = threadIdx.x; x < 512u * 3u; x += RED_THREADS)
        s_w[x] = 0ull;
    for (uint32_t x = threadIdx.x; x < 512u; x += RED_THREADS)
        s_g0[x] = 0xFFFFFFFFu;
    if (threadIdx.x < 64u)
        s_exp2[threadIdx.x] = EXP2_J64[threadIdx.x];
    __syncthreads();
    const uint32_t l = s_leaf;
    const uint32_t chunk = item - item_scan[l];
    const uint32_t nchunks = item_scan[l + 1] - item_scan[l];
    const uint64_t first = starts[l];
    const uint64_t p0 = first + (uint64_t)chunk * LRD_CHUNK;
    const uint64_t pend = starts[l + 1];
    const uint64_t p1 = p0 + LRD_CHUNK < pend ? p0 + LRD_CHUNK : pend;
    const uint32_t lshift = kl.gid_bits + 9u;
    const uint64_t gmask = (1ull << kl.gid_bits) - 1ull;
    const uint64_t leaf_bits = leaf_of_field(kl, keys[first] >> lshift);
    int32_t i0, j0, k0;
    decode_voxel(kl, leaf_bits << 9, &i0, &j0, &k0);
    if (chunk == 0u && threadIdx.x == 0)
        lkey[l] = leaf_bits;
    const double vsd = (double)vs;
    for (uint64_t q = p0 + threadIdx.x; q < p1; q += RED_THREADS)
    {
        const uint64_t key = keys[q];
        const uint32_t gid = (uint32_t)(key & gmask);
        const uint32_t slot = (uint32_t)(key >> kl.gid_bits) & 511u;
        const DensPrim& pr = prims[gid];
        const double d[3] = { ((double)(i0 + (int32_t)(slot >> 6)) + 0.5) * vsd - (double)pr.mu[0],
                              ((double)(j0 + (int32_t)((slot >> 3) & 7u)) + 0.5) * vsd - (double)pr.mu[1],
                              ((double)(k0 + (int32_t)(slot & 7u)) + 0.5) * vsd - (double)pr.mu[2] };
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const double u = pr.M[3 * c + 0] * d[0] + pr.M[3 * c + 1] * d[1] + pr.M[3 * c + 2] * d[2];
            m += u * u;
        }
        uint64_t lo, hi;
        fix120(pr.o * exp_neg(-0.5 * m, s_exp2), &lo, &hi);
        add192(&s_w[slot * 3u], lo, hi, 0ull);
        atomicMin(&s_g0[slot], gid);
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < 512u; x += RED_THREADS)
    {
        const uint64_t at = (uint64_t)l * 512u + x;
        if (nchunks == 1u)
        {
            lw[at * 3u + 0u] = s_w[3u * x + 0u];
            lw[at * 3u + 1u] = s_w[3u * x + 1u];
            lw[at * 3u + 2u] = s_w[3u * x + 2u];
            lg0[at] = s_g0[x];
        }
        else if (s_g0[x] != 0xFFFFFFFFu)
        {
            add192(&lw[at * 3u], s_w[3u * x], s_w[3u * x + 1u], s_w[3u * x + 2u]);
            atomicMin(&lg0[at], s_g0[x]);
        }
    }
}

// one wave per leaf: its voxels with a pair, in slot (= key) order, at the leaf's offset, with the
// float of each fixed-point sum
__global__ __launch_bounds__(256) void k_leaf_write_density(const unsigned long long* __restrict__ lw,
                                                            const uint32_t* __restrict__ lg0,
                                                            const uint64_t* __restrict__ lkey,
                                                            const uint32_t* __restrict__ lscan, DCount cnleaf,
                                                            uint64_t* __restrict__ vox_key, uint32_t* __restrict__ vox_val,
                                                            uint32_t* __restrict__ vox_g0)
{
    __shared__ uint64_t s_key[4][512];
    __shared__ uint32_t s_val[4][512];
    __shared__ uint32_t s_g0[4][512];
    const uint64_t nleaf = cnleaf.get();
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t leaf = (uint64_t)blockIdx.x * 4u + wave;
    if (leaf >= nleaf)
        return; // wave-uniform; the kernel has no block barrier
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t at = leaf * 512u + lane * 8u;
    uint32_t g[8];
    {
        const uint4 a = *(const uint4*)(lg0 + at);
        const uint4 b = *(const uint4*)(lg0 + at + 4u);
        g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
        g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
    }
    uint32_t cnt = 0u;
#pragma unroll
    for (int t = 0; t < 8; t++)
        cnt += g[t] != 0xFFFFFFFFu ? 1u : 0u;
    uint32_t wtot;
    const uint32_t lpos = wave_excl_add(cnt, lane, &wtot);
    uint32_t v[8];
#pragma unroll
    for (int t = 0; t < 8; t++)
    {
        v[t] = 0u;
        if (g[t] != 0xFFFFFFFFu)
        {
            const unsigned long long* w = lw + (at + (uint64_t)t) * 3u;
            v[t] = __float_as_uint(fix120_to_float(w[0], w[1], w[2]));
        }
    }
    leaf_write_staged(lane, g, v, lkey[leaf] << 9, lscan[leaf], lpos, wtot, s_key[wave], s_val[wave], s_g0[wave],
                      vox_key, vox_val, vox_g0);
}

// floor(x / d) for d >= 1. Below 2^52, x and d are exact doubles and x * fl(1/d) is within a
// relative 2^-52 of x / d, so less than 1 from it: its floor is the quotient or one of its
// neighbours, and one integer check each way settles it. Larger x take the integer division.
__device__ __forceinline__ uint64_t div_u64_by_u32(uint64_t x, uint32_t d, double inv_d)
{
    if (x >= (1ull << 52))
        return x / d;
    uint64_t q = (uint64_t)floor((double)x * inv_d);
    if (q * d > x)
        q--;
    else if ((q + 1u) * d <= x)
        q++;
    return q;
}

// per voxel: first-touch window from the lowest gaussian id, and voxel bounds
__global__ __launch_bounds__(256) void k_voxel_finish(DCount cnvox, KeyLayout kl, const uint64_t* __restrict__ vox_key,
                                                      const uint32_t* __restrict__ vox_g0, const PrimBox* __restrict__ boxes,
                                                      const uint64_t* __restrict__ foot_scan,
                                                      const uint32_t* __restrict__ win_base, uint32_t prim_batch,
                                                      uint32_t ijk_batch, uint32_t* __restrict__ vox_t,
                                                      int32_t* __restrict__ bounds)
{
    const uint64_t nvox = cnvox.get();
    const double inv_ijk = 1.0 / (double)ijk_batch; // ijk_batch >= 1
    int32_t mn[3] = { 2147483647, 2147483647, 2147483647 }, mx[3] = { -2147483647, -2147483647, -2147483647 };
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += stride)
    {
        int32_t i, j, k;
        decode_voxel(kl, vox_key[v], &i, &j, &k);
        uint32_t g0 = vox_g0[v];
        PrimBox bx = boxes[g0];
        uint64_t ri = ((uint64_t)(i - bx.mn[0]) * (uint64_t)bx.rng_y + (uint64_t)(j - bx.mn[1])) * (uint64_t)bx.rng_z +
                      (uint64_t)(k - bx.mn[2]);
        uint32_t batch = g0 / prim_batch;
        uint64_t slot = foot_scan[g0] - foot_scan[(uint64_t)batch * prim_batch] + ri;
        vox_t[v] = win_base[batch] + (uint32_t)div_u64_by_u32(slot, ijk_batch, inv_ijk);
        mn[0] = min(mn[0], i);
        mn[1] = min(mn[1], j);
        mn[2] = min(mn[2], k);
        mx[0] = max(mx[0], i);
        mx[1] = max(mx[1], j);
        mx[2] = max(mx[2], k);
    }
    commit_bounds(mn, mx, bounds);
}

// ============================================================================ leaves
// per leaf (run of voxels sharing key >> 9): value mask, first-touch window, tree key. One wave
// per leaf reads the leaf's voxel run coalesced; each lane ORs its voxels' slot bits into 16
// 32-bit halves of the 512-bit mask, and shuffles join the lanes.
__global__ __launch_bounds__(256) void k_leaf_info(const uint64_t* __restrict__ vox_key, const uint32_t* __restrict__ vox_t,
                                                   const uint32_t* __restrict__ leaf_first, DCount cnleaf, KeyLayout kl,
                                                   uint64_t* __restrict__ leaf_mask, uint32_t* __restrict__ leaf_t,
                                                   uint64_t* __restrict__ sort_keys, uint32_t* __restrict__ sort_vals)
{
    const uint64_t nleaf = cnleaf.get();
    const uint64_t l = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
    if (l >= nleaf)
        return; // wave-uniform
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b = leaf_first[l], e = leaf_first[l + 1];
    uint32_t m[16];
#pragma unroll
    for (int h = 0; h < 16; h++)
        m[h] = 0u;
    uint32_t t = 0xFFFFFFFFu;
    for (uint32_t v = b + lane; v < e; v += 64u)
    {
        const uint32_t n = (uint32_t)(vox_key[v] & 511u);
        const uint32_t bit = 1u << (n & 31u);
#pragma unroll
        for (int h = 0; h < 16; h++)
            m[h] |= (n >> 5) == (uint32_t)h ? bit : 0u;
        t = min(t, vox_t[v]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
#pragma unroll
        for (int h = 0; h < 16; h++)
            m[h] |= (uint32_t)__shfl_xor((int)m[h], off, 64);
        t = min(t, (uint32_t)__shfl_xor((int)t, off, 64));
    }
    if (lane != 0u)
        return;
#pragma unroll
    for (int w = 0; w < 8; w++)
        leaf_mask[8 * l + w] = (uint64_t)m[2 * w] | ((uint64_t)m[2 * w + 1] << 32);
    leaf_t[l] = t;
    int32_t i, j, k;
    decode_voxel(kl, vox_key[b] & ~511ull, &i, &j, &k);
    uint32_t ui = (uint32_t)i & 0x1FFFFu, uj = (uint32_t)j & 0x1FFFFu, uk = (uint32_t)k & 0x1FFFFu;
    uint64_t root_n = ((uint64_t)(ui >> 12) << 10) | ((uint64_t)(uj >> 12) << 5) | (uint64_t)(uk >> 12);
    uint64_t upper_n = ((uint64_t)((ui & 0xFFFu) >> 7) << 10) | ((uint64_t)((uj & 0xFFFu) >> 7) << 5) |
                       (uint64_t)((uk & 0xFFFu) >> 7);
    uint64_t lower_n = ((uint64_t)((ui & 0x7Fu) >> 3) << 8) | ((uint64_t)((uj & 0x7Fu) >> 3) << 4) |
                       (uint64_t)((uk & 0x7Fu) >> 3);
    sort_keys[l] = (root_n << 27) | (upper_n << 12) | lower_n;
    sort_vals[l] = (uint32_t)l;
}

// gather per tree-ordered leaf: key, t; and inverse permutation
__global__ __launch_bounds__(256) void k_leaf_gather(const uint64_t* __restrict__ sorted_keys,
                                                     const uint32_t* __restrict__ sorted_idx, DCount cnleaf,
                                                     const uint32_t* __restrict__ leaf_t, uint32_t* __restrict__ tree_t,
                                                     uint32_t* __restrict__ leaf_pos)
{
    const uint64_t nleaf = cnleaf.get();
    uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf)
        return;
    uint32_t src = sorted_idx[l];
    tree_t[l] = leaf_t[src];
    leaf_pos[src] = (uint32_t)l;
}

// min of a u32 array over runs [first[r], first[r+1])
// out[r] = minimum of vals over run r; one wave per run
__global__ __launch_bounds__(256) void k_run_min(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ first,
                                                 DCount cnrun, uint32_t* __restrict__ out)
{
    const uint64_t nrun = cnrun.get();
    const uint64_t r = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
    if (r >= nrun)
        return; // wave-uniform
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t e = first[r + 1];
    uint32_t t = 0xFFFFFFFFu;
    for (uint32_t x = first[r] + lane; x < e; x += 64u)
        t = min(t, vals[x]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        t = min(t, (uint32_t)__shfl_xor((int)t, off, 64));
    if (lane == 0u)
        out[r] = t;
}

// key of each run head: out[r] = keys[first[r]] >> shift
__global__ __launch_bounds__(256) void k_run_keys(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ first,
                                                  DCount cnrun, uint32_t shift, uint64_t* __restrict__ out)
{
    const uint64_t nrun = cnrun.get();
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrun)
        return;
    out[r] = keys[first[r]] >> shift;
}

// ============================================================================ node order
// window histogram per level: hist[level * T + t]. Grid-stride over the device count: the grid is
// bounded (NODE_GRID_BLOCKS), not sized by the count's capacity
constexpr uint32_t NODE_GRID_BLOCKS = 2048;
__global__ __launch_bounds__(256) void k_window_hist(const uint32_t* __restrict__ t, DCount ccount, uint32_t T,
                                                     uint32_t* __restrict__ hist_level)
{
    // nodes of one window cluster (few windows, many leaves): each wave adds one count per distinct
    // window among its 64 nodes, so same-address atomics do not serialize per node
    const uint64_t count = ccount.get();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < count; base += stride)
    {
        const uint64_t x = base + lane;
        const bool valid = x < count;
        const uint32_t tv = valid ? t[x] : 0u;
        uint64_t pending = __ballot(valid); // wave-uniform
        while (pending)
        {
            const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
            const uint32_t tl = (uint32_t)__shfl((int)tv, (int)leader, 64);
            const uint64_t same = __ballot(valid && tv == tl);
            if (lane == leader)
                atomicAdd(&hist_level[tl], (uint32_t)__popcll(same));
            pending &= ~same;
        }
    }
}

__global__ __launch_bounds__(256) void k_window_totals(const uint32_t* __restrict__ hist, uint32_t T,
                                                       uint32_t* __restrict__ tot)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T)
        return;
    tot[t] = hist[t] + hist[T + t] + hist[2 * T + t];
}

// sort key of a node: (t << tshift) | (parent node idx << nbits) | slot
__global__ __launch_bounds__(256) void k_node_keys(const uint32_t* __restrict__ t, const uint32_t* __restrict__ parent_idx,
                                                   const uint64_t* __restrict__ slot_key, DCount ccount, uint32_t slot_bits,
                                                   uint32_t parent_bits, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals)
{
    const uint64_t count = ccount.get();
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count)
        return;
    uint64_t slot = slot_key[x] & ((1ull << slot_bits) - 1ull);
    uint64_t par = parent_idx ? (uint64_t)parent_idx[x] : 0ull;
    keys[x] = ((uint64_t)t[x] << (slot_bits + parent_bits)) | (par << slot_bits) | slot;
    vals[x] = (uint32_t)x;
}

// node_idx from the sorted order: 1 + tot_before[t] + level_offset_in_t + (pos - level_before[t])
__global__ __launch_bounds__(256) void k_node_assign(const uint64_t* __restrict__ sorted_keys,
                                                     const uint32_t* __restrict__ sorted_vals, DCount ccount,
                                                     uint32_t tshift, const uint32_t* __restrict__ tot_before,
                                                     const uint32_t* __restrict__ level_before,
                                                     const uint32_t* __restrict__ hist, uint32_t T, uint32_t level,
                                                     uint32_t* __restrict__ node_idx)
{
    const uint64_t count = ccount.get();
    uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= count)
        return;
    uint32_t t = (uint32_t)(sorted_keys[pos] >> tshift);
    uint32_t within = 0u;
    for (uint32_t l = 0; l < level; l++)
        within += hist[l * T + t];
    node_idx[sorted_vals[pos]] = 1u + tot_before[t] + within + (uint32_t)(pos - level_before[t]);
}

// node_idx = base + (counts of the levels before) + position, in tree order
__global__ __launch_bounds__(256) void k_node_bfs(DCount ccount, uint32_t base, DCount before1, DCount before2,
                                                  uint32_t* __restrict__ node_idx)
{
    const uint64_t count = ccount.get();
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count)
        return;
    node_idx[x] = base + (uint32_t)before1.get() + (uint32_t)before2.get() + (uint32_t)x;
}

// parent node idx of each child run member: child c in [first[p], first[p+1]) -> parent_idx[p]
__global__ __launch_bounds__(256) void k_parent_idx(const uint32_t* __restrict__ first, DCount cnparent,
                                                    const uint32_t* __restrict__ parent_node_idx,
                                                    uint32_t* __restrict__ child_parent)
{
    // one wave per parent (its children written 64 at a time), grid-stride over the device count
    const uint64_t nparent = cnparent.get();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64u);
    for (uint64_t p = (uint64_t)blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6); p < nparent; p += waves)
    {
        const uint32_t idx = parent_node_idx ? parent_node_idx[p] : 0u;
        const uint32_t c1 = first[p + 1];
        for (uint32_t c = first[p] + lane; c < c1; c += 64u)
            child_parent[c] = idx;
    }
}

// per node idx: byte size and value count
__global__ __launch_bounds__(256) void k_node_sizes(const uint32_t* __restrict__ node_idx, DCount ccount, uint32_t nmax,
                                                    uint64_t size,
                                                    const uint64_t* __restrict__ leaf_mask,
                                                    const uint32_t* __restrict__ tree_to_compact,
                                                    uint64_t* __restrict__ size_by_idx, uint64_t* __restrict__ vals_by_idx)
{
    const uint64_t count = ccount.get();
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count)
        return;
    uint32_t ni = node_idx[x];
    if (ni == 0u || ni >= nmax)
        return; // unreachable for consistent orders; keeps a faulty order from writing out of bounds
    size_by_idx[ni] = size;
    uint64_t pc = 0;
    if (leaf_mask)
    {
        uint32_t l = tree_to_compact[x];
        for (int w = 0; w < 8; w++)
            pc += (uint64_t)__popcll(leaf_mask[8 * (uint64_t)l + w]);
    }
    vals_by_idx[ni] = pc;
}

// ============================================================================ node2 write
struct GridInfo
{
    uint64_t root_addr;
    uint64_t nodes_begin; // byte offset of node idx 1
    uint64_t end;         // end of the node segment
    uint64_t meta0;
    uint64_t meta1;
    uint64_t bboxes;
    uint64_t values;
    uint64_t grid_size;
    uint64_t value_count;
    uint32_t upper_count;
    uint32_t lower_count;
    uint32_t leaf_count;
    uint32_t value_type;
    uint32_t node_count;
    float voxel_size;
    float voxel_size_inv;
};

__device__ __forceinline__ void put_u32(uint8_t* g, uint64_t off, uint32_t v)
{
    *(uint32_t*)(g + off) = v;
}
__device__ __forceinline__ void put_u64(uint8_t* g, uint64_t off, uint64_t v)
{
    *(uint64_t*)(g + off) = v;
}
__device__ __forceinline__ void put_f32(uint8_t* g, uint64_t off, float v)
{
    *(float*)(g + off) = v;
}
__device__ __forceinline__ void put_f64(uint8_t* g, uint64_t off, double v)
{
    *(double*)(g + off) = v;
}

// node byte address of node idx ni (ni >= 1): nodes_begin + exclusive size scan
__device__ __forceinline__ uint64_t node_addr(const GridInfo& gi, const uint64_t* size_scan, uint32_t ni)
{
    // indices outside [1, node_count) cannot come from a consistent order; they map to the root
    // so that no write leaves the grid
    return (ni == 0u || ni >= gi.node_count) ? gi.root_addr : gi.nodes_begin + size_scan[ni];
}

__global__ void k_write_header(uint8_t* __restrict__ g, GridInfo gi_in, const GridInfo* __restrict__ gip, const int32_t* __restrict__ bounds)
{
    const GridInfo gi = gip ? *gip : gi_in;
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    // GridBuild.cpp:126-187
    put_u64(g, VDB_GRID_OFF_MAGIC, VDB_MAGIC_GRID);
    put_u32(g, VDB_GRID_OFF_VERSION, VDB_VERSION);
    put_u32(g, VDB_GRID_OFF_FLAGS, 0u);
    put_u32(g, VDB_GRID_OFF_GRID_INDEX, 0u);
    put_u32(g, VDB_GRID_OFF_GRID_COUNT, 1u);
    put_u64(g, VDB_GRID_OFF_GRID_SIZE, gi.grid_size);
    put_u32(g, VDB_GRID_OFF_NAME + 0, 0x65646f6eu); // "node2"
    put_u32(g, VDB_GRID_OFF_NAME + 4, 0x00000032u);
    for (int a = 0; a < 3; a++)
        put_f64(g, VDB_GRID_OFF_VOXEL_SIZE + 8 * a, (double)gi.voxel_size);
    put_u32(g, VDB_GRID_OFF_CLASS, 0u);
    put_u32(g, VDB_GRID_OFF_TYPE, VDB_GRID_TYPE_NODE2);
    const uint64_t map = VDB_GRID_OFF_MAP;
    for (int a = 0; a < 3; a++)
    {
        put_f32(g, map + VDB_MAP_OFF_MATF + 4 * (4 * a), gi.voxel_size);
        put_f32(g, map + VDB_MAP_OFF_INVMATF + 4 * (4 * a), gi.voxel_size_inv);
        put_f64(g, map + VDB_MAP_OFF_MATD + 8 * (4 * a), (double)gi.voxel_size);
        put_f64(g, map + VDB_MAP_OFF_INVMATD + 8 * (4 * a), (double)gi.voxel_size_inv);
    }
    put_u64(g, VDB_GRID_OFF_BLIND_OFFSET, gi.meta0);
    put_u32(g, VDB_GRID_OFF_BLIND_COUNT, 2u);
    // tree
    put_u64(g, VDB_TREE_ADDR + VDB_TREE_OFF_ROOT, gi.root_addr - VDB_TREE_ADDR);
    put_u32(g, VDB_TREE_ADDR + VDB_TREE_OFF_COUNT_LEAF, gi.leaf_count);
    put_u32(g, VDB_TREE_ADDR + VDB_TREE_OFF_COUNT_LOWER, gi.lower_count);
    put_u32(g, VDB_TREE_ADDR + VDB_TREE_OFF_COUNT_UPPER, gi.upper_count);
    // blind metadata (grid_build_blindmetadata.slang:85-110)
    put_u64(g, gi.meta0 + VDB_BLIND_OFF_DATA, gi.bboxes - gi.meta0);
    put_u64(g, gi.meta0 + VDB_BLIND_OFF_COUNT, 2u);
    put_u32(g, gi.meta0 + VDB_BLIND_OFF_SIZE, 12u);
    put_u32(g, gi.meta0 + VDB_BLIND_OFF_TYPE, VDB_GRID_TYPE_VEC3F);
    put_u64(g, gi.meta1 + VDB_BLIND_OFF_DATA, gi.values - gi.meta1);
    put_u64(g, gi.meta1 + VDB_BLIND_OFF_COUNT, gi.value_count);
    put_u32(g, gi.meta1 + VDB_BLIND_OFF_SIZE, 4u);
    put_u32(g, gi.meta1 + VDB_BLIND_OFF_TYPE, gi.value_type);
    // root bbox (grid_build_compute_bbox.slang:150-158): min over voxels, or the empty sentinel
    for (int a = 0; a < 3; a++)
    {
        put_u32(g, gi.bboxes + 4 * a, (uint32_t)bounds[a]);
        put_u32(g, gi.bboxes + 12 + 4 * a, (uint32_t)bounds[3 + a]);
    }
    // root node idx 0, value idx 2
    put_u64(g, gi.root_addr + 8ull * VDB_OFF_VALUE_IDX, 2ull);
}

// packed prefix lanes of a mask held in LDS (count words, count % 4 == 0) -> out words
__device__ void write_prefix_words(const uint64_t* mask_lds, uint32_t count, uint32_t* pc_lds, uint64_t* out)
{
    // popcounts and exclusive scan (single thread scan is enough for <= 512 words per node,
    // done by wave 0 with a shuffle scan over 8-word chunks)
    for (uint32_t w = threadIdx.x; w < count; w += blockDim.x)
        pc_lds[w] = (uint32_t)__popcll(mask_lds[w]);
    __syncthreads();
    if (threadIdx.x < 64)
    {
        uint32_t lane = threadIdx.x;
        uint32_t per = (count + 63) / 64; // words per lane
        uint32_t s = 0;
        for (uint32_t q = 0; q < per; q++)
        {
            uint32_t w = lane * per + q;
            if (w < count)
                s += pc_lds[w];
        }
        uint32_t incl = s;
        for (int off = 1; off < 64; off <<= 1)
        {
            uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off)
                incl += o;
        }
        uint32_t run = incl - s;
        for (uint32_t q = 0; q < per; q++)
        {
            uint32_t w = lane * per + q;
            if (w < count)
            {
                uint32_t c = pc_lds[w];
                pc_lds[w] = run; // exclusive
                run += c;
            }
        }
        if (lane == 63)
            pc_lds[count] = incl; // total
    }
    __syncthreads();
    uint32_t total = pc_lds[count];
    for (uint32_t q = threadIdx.x; q < count / 4; q += blockDim.x)
    {
        uint64_t l0 = (q == 0) ? total : pc_lds[4 * q];
        uint64_t v = (l0 & 0xFFFFull) | ((uint64_t)(pc_lds[4 * q + 1] & 0xFFFFu) << 16) |
                     ((uint64_t)(pc_lds[4 * q + 2] & 0xFFFFu) << 32) | ((uint64_t)(pc_lds[4 * q + 3] & 0xFFFFu) << 48);
        out[q] = v;
    }
}

// root: child mask + key mask over upper root slots, prefixes, children offsets and keys
__global__ __launch_bounds__(1024) void k_write_root(uint8_t* __restrict__ g, GridInfo gi_in, const GridInfo* __restrict__ gip,
                                                     const uint64_t* __restrict__ upper_key, DCount cnupper,
                                                     const uint32_t* __restrict__ upper_node_idx,
                                                     const uint64_t* __restrict__ size_scan)
{
    const GridInfo gi = gip ? *gip : gi_in;
    const uint64_t nupper = cnupper.get();
    __shared__ uint64_t mask[512];
    __shared__ uint32_t pc[513];
    for (uint32_t w = threadIdx.x; w < 512; w += blockDim.x)
        mask[w] = 0ull;
    __syncthreads();
    for (uint64_t u = threadIdx.x; u < nupper; u += blockDim.x)
    {
        uint32_t n = (uint32_t)upper_key[u];
        atomicOr((unsigned long long*)&mask[n >> 6], 1ull << (n & 63u));
    }
    __syncthreads();
    uint64_t* root = (uint64_t*)(g + gi.root_addr);
    for (uint32_t w = threadIdx.x; w < 512; w += blockDim.x)
    {
        root[VDB_OFF_CM_ROOT + w] = mask[w];
        root[VDB_OFF_KM_ROOT + w] = mask[w];
    }
    write_prefix_words(mask, 512, pc, root + VDB_OFF_CMP_ROOT);
    __syncthreads();
    write_prefix_words(mask, 512, pc, root + VDB_OFF_KMP_ROOT);
    // children (rank order == root slot order) then keys, rest END_KEY
    for (uint64_t x = threadIdx.x; x < 65536ull; x += blockDim.x)
    {
        uint64_t v;
        if (x < nupper)
        {
            v = node_addr(gi, size_scan, upper_node_idx[x]) - gi.root_addr;
        }
        else if (x < 2 * nupper)
        {
            uint32_t n = (uint32_t)upper_key[x - nupper];
            int32_t ci = (int32_t)(((n >> 10) & 0x1Fu) << 12), cj = (int32_t)(((n >> 5) & 0x1Fu) << 12),
                    ck = (int32_t)((n & 0x1Fu) << 12);
            const int32_t half = 65536;
            ci = ci >= half ? ci - half - half : ci;
            cj = cj >= half ? cj - half - half : cj;
            ck = ck >= half ? ck - half - half : ck;
            v = (uint64_t)((uint32_t)ck >> 12) | ((uint64_t)((uint32_t)cj >> 12) << 21) |
                ((uint64_t)((uint32_t)ci >> 12) << 42);
        }
        else
        {
            v = VDB_END_KEY;
        }
        root[VDB_OFF_CHILDREN_ROOT + x] = v;
    }
}

// internal node (upper or lower): one block per node
template <int WORDS>
__global__ __launch_bounds__(256) void k_write_internal(uint8_t* __restrict__ g, GridInfo gi_in, const GridInfo* __restrict__ gip,
                                                        const uint32_t* __restrict__ node_first, // children runs
                                                        const uint64_t* __restrict__ child_slot_key,
                                                        uint32_t slot_bits,
                                                        const uint32_t* __restrict__ node_idx,
                                                        const uint32_t* __restrict__ child_node_idx,
                                                        const uint64_t* __restrict__ size_scan,
                                                        const uint64_t* __restrict__ vals_scan,
                                                        uint32_t off_cmp, uint32_t off_cm, uint32_t off_children,
                                                        DCount ccount)
{
    if (blockIdx.x >= ccount.get())
        return; // block-uniform: the grid may be sized by a capacity
    const GridInfo gi = gip ? *gip : gi_in;
    __shared__ uint64_t mask[WORDS];
    __shared__ uint32_t pc[WORDS + 1];
    const uint32_t node = blockIdx.x;
    for (uint32_t w = threadIdx.x; w < WORDS; w += blockDim.x)
        mask[w] = 0ull;
    __syncthreads();
    uint32_t b = node_first[node], e = node_first[node + 1];
    const uint64_t smask = (1ull << slot_bits) - 1ull;
    for (uint32_t c = b + threadIdx.x; c < e; c += blockDim.x)
    {
        uint32_t n = (uint32_t)(child_slot_key[c] & smask);
        atomicOr((unsigned long long*)&mask[n >> 6], 1ull << (n & 63u));
    }
    __syncthreads();
    uint32_t ni = node_idx[node];
    uint64_t addr = node_addr(gi, size_scan, ni);
    uint64_t* nd = (uint64_t*)(g + addr);
    for (uint32_t w = threadIdx.x; w < WORDS; w += blockDim.x)
        nd[off_cm + w] = mask[w];
    write_prefix_words(mask, WORDS, pc, nd + off_cmp);
    for (uint32_t c = b + threadIdx.x; c < e; c += blockDim.x)
        nd[off_children + (c - b)] = node_addr(gi, size_scan, child_node_idx[c]) - addr;
    if (threadIdx.x == 0)
    {
        nd[VDB_OFF_NODE_IDX] = ni;
        nd[VDB_OFF_VALUE_IDX] = 2ull + vals_scan[ni];
    }
}

__global__ __launch_bounds__(256) void k_write_leaves(uint8_t* __restrict__ g, GridInfo gi_in, const GridInfo* __restrict__ gip, DCount cnleaf,
                                                      const uint32_t* __restrict__ tree_to_compact,
                                                      const uint64_t* __restrict__ leaf_mask,
                                                      const uint32_t* __restrict__ leaf_node_idx,
                                                      const uint64_t* __restrict__ size_scan,
                                                      const uint64_t* __restrict__ vals_scan)
{
    const GridInfo gi = gip ? *gip : gi_in;
    const uint64_t nleaf = cnleaf.get();
    uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf)
        return;
    uint32_t src = tree_to_compact[l];
    uint32_t ni = leaf_node_idx[l];
    uint64_t* nd = (uint64_t*)(g + node_addr(gi, size_scan, ni));
    uint64_t m[8];
    uint32_t ex[9];
    ex[0] = 0;
    for (int w = 0; w < 8; w++)
    {
        m[w] = leaf_mask[8 * (uint64_t)src + w];
        ex[w + 1] = ex[w] + (uint32_t)__popcll(m[w]);
    }
    nd[VDB_OFF_NODE_IDX] = ni;
    nd[VDB_OFF_VALUE_IDX] = 2ull + vals_scan[ni];
    for (int w = 0; w < 8; w++)
        nd[VDB_OFF_VM_LEAF + w] = m[w];
    nd[VDB_OFF_VMP_LEAF + 0] = (uint64_t)(ex[8] & 0xFFFFu) | ((uint64_t)ex[1] << 16) | ((uint64_t)ex[2] << 32) |
                               ((uint64_t)ex[3] << 48);
    nd[VDB_OFF_VMP_LEAF + 1] = (uint64_t)ex[4] | ((uint64_t)ex[5] << 16) | ((uint64_t)ex[6] << 32) |
                               ((uint64_t)ex[7] << 48);
}

__global__ __launch_bounds__(256) void k_write_values(uint8_t* __restrict__ g, GridInfo gi_in, const GridInfo* __restrict__ gip,
                                                      DCount cnleaf, const uint32_t* __restrict__ vox_val,
                                                      const uint64_t* __restrict__ vox_key,
                                                      const uint32_t* __restrict__ leaf_first,
                                                      const uint32_t* __restrict__ leaf_pos,
                                                      const uint32_t* __restrict__ leaf_node_idx,
                                                      const uint64_t* __restrict__ vals_scan, KeyLayout kl,
                                                      int32_t* __restrict__ ijk_out)
{
    // one wave per leaf (grid-stride): its node index and first value index are read once, and its
    // voxels [leaf_first[l], leaf_first[l + 1]) go to consecutive value slots
    const GridInfo gi = gip ? *gip : gi_in;
    const uint64_t nleaf = cnleaf.get();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64u);
    for (uint64_t l = (uint64_t)blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6); l < nleaf; l += waves)
    {
        const uint32_t b = leaf_first[l], e = leaf_first[l + 1];
        const uint32_t ni = leaf_node_idx[leaf_pos[l]];
        if (ni >= gi.node_count)
            continue;
        const uint64_t base = 2ull + vals_scan[ni];
        for (uint32_t v = b + lane; v < e; v += 64u)
        {
            const uint64_t idx = base + (v - b);
            if (idx < gi.value_count)
            {
                *(uint32_t*)(g + gi.values + 4ull * idx) = vox_val[v];
                if (ijk_out) // voxel coordinates in value-index order (VDBR_FLAG_VOXEL_COORDS)
                {
                    int32_t i, j, k;
                    decode_voxel(kl, vox_key[v], &i, &j, &k);
                    ijk_out[3 * idx + 0] = i;
                    ijk_out[3 * idx + 1] = j;
                    ijk_out[3 * idx + 2] = k;
                }
            }
        }
    }
}

// voxel -> leaf run index
// member_of[x] = r for every x of run r; one wave per run (coalesced writes)
__global__ void k_gather_u64(const uint64_t* __restrict__ src, uint64_t stride, uint64_t count,
                             uint64_t* __restrict__ dst)
{
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count)
        return;
    dst[x] = src[x * stride];
}

__global__ void k_init_bounds(int32_t* b)
{
    if (threadIdx.x < 3)
        b[threadIdx.x] = 2147483647;
    else if (threadIdx.x < 6)
        b[threadIdx.x] = -2147483647;
}

__global__ void k_pairs_decode(const uint64_t* __restrict__ keys, uint64_t n, KeyLayout kl, int32_t* __restrict__ ijk,
                               uint32_t* __restrict__ gid)
{
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n)
        return;
    uint64_t k = keys[x];
    if (gid)
        gid[x] = (uint32_t)((k >> kl.alpha_bits) & ((1ull << kl.gid_bits) - 1ull));
    if (ijk)
    {
        int32_t i, j, kk;
        decode_pair_voxel(kl, k >> (kl.gid_bits + kl.alpha_bits), &i, &j, &kk);
        ijk[3 * x + 0] = i;
        ijk[3 * x + 1] = j;
        ijk[3 * x + 2] = kk;
    }
}

// ============================================================================ host side
static inline uint32_t blocks_for(uint64_t n, uint32_t t = 256)
{
    return (uint32_t)((n + t - 1) / t);
}

// ---------------------------------------------------------------- candidate-leaf ranks
// Marks the leaves of every Gaussian's voxel AABB in a bitmap over the leaf index (x, y, z leaf
// coordinates relative to lmin, z least significant): a leaf column (x, y) of the AABB is one run
// of consecutive bits. A word already holding the run's bits is not written again.
// marks the leaves z0..z1 (inclusive) of one leaf column in the candidate bitmap
__device__ __forceinline__ void mark_leaf_run(unsigned long long* __restrict__ map, uint64_t col, uint32_t z0, uint32_t z1)
{
    uint64_t lo = col + z0;
    const uint64_t hi = col + z1; // inclusive
    while (lo <= hi)
    {
        const uint64_t w = lo >> 6;
        const uint32_t b0 = (uint32_t)(lo & 63u);
        const uint32_t b1 = (hi >> 6) == w ? (uint32_t)(hi & 63u) : 63u;
        const unsigned long long m = (~0ull >> (63u - (b1 - b0))) << b0;
        if ((map[w] & m) != m)
            atomicOr(&map[w], m);
        lo = (w + 1) << 6;
    }
}

// columns of one Gaussian above which it marks its own columns (keeps a wave's column total
// below 2^26)
constexpr uint64_t LEAF_MARK_SPREAD_MAX = 1ull << 20;

__global__ __launch_bounds__(256) void k_leaf_mark(const PrimBox* __restrict__ boxes, uint64_t n, KeyLayout kl,
                                                   unsigned long long* __restrict__ map)
{
    // a wave takes 64 Gaussians at a time and spreads their leaf columns (x, y) over its lanes:
    // column t of the group belongs to the lane whose exclusive column prefix is the largest one
    // <= t (an LDS search), so a large AABB does not hold the other 63 lanes
    __shared__ uint32_t s_pre[4][64], s_ny[4][64], s_z0[4][64], s_z1[4][64];
    __shared__ int32_t s_x0[4][64], s_y0[4][64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t waves_total = (uint64_t)gridDim.x * 4u;
    for (uint64_t g0 = ((uint64_t)blockIdx.x * 4u + wave) * 64u; g0 < n; g0 += waves_total * 64u)
    {
        const uint64_t g = g0 + lane;
        uint32_t ncol = 0u, ny = 1u, z0 = 0u, z1 = 0u;
        int32_t x0 = 0, y0 = 0;
        uint64_t own_nx = 0u; // a Gaussian of more than LEAF_MARK_SPREAD_MAX columns: marked by its lane
        if (g < n)
        {
            const PrimBox b = boxes[g];
            if (b.rng_x > 0 && b.rng_y > 0 && b.rng_z > 0)
            {
                x0 = (b.mn[0] >> 3) - kl.lmin[0];
                y0 = (b.mn[1] >> 3) - kl.lmin[1];
                const int32_t x1 = ((b.mn[0] + (int32_t)b.rng_x - 1) >> 3) - kl.lmin[0];
                const int32_t y1 = ((b.mn[1] + (int32_t)b.rng_y - 1) >> 3) - kl.lmin[1];
                z0 = (uint32_t)((b.mn[2] >> 3) - kl.lmin[2]);
                z1 = (uint32_t)(((b.mn[2] + (int32_t)b.rng_z - 1) >> 3) - kl.lmin[2]);
                const uint64_t nx = (uint64_t)(x1 - x0 + 1);
                ny = (uint32_t)(y1 - y0 + 1);
                if (nx * ny > LEAF_MARK_SPREAD_MAX)
                    own_nx = nx;
                else
                    ncol = (uint32_t)(nx * ny);
            }
        }
        uint32_t total;
        const uint32_t pre = wave_excl_add(ncol, lane, &total);
        s_pre[wave][lane] = pre;
        s_ny[wave][lane] = ny;
        s_z0[wave][lane] = z0;
        s_z1[wave][lane] = z1;
        s_x0[wave][lane] = x0;
        s_y0[wave][lane] = y0;
        wave_sync_lds();
        for (uint32_t t0 = 0; t0 < total; t0 += 64u)
        {
            const uint32_t t = t0 + lane;
            if (t < total)
            {
                uint32_t o = 0u; // largest lane with s_pre <= t (s_pre[0] = 0)
#pragma unroll
                for (uint32_t step = 32u; step > 0u; step >>= 1)
                    if (s_pre[wave][o + step] <= t)
                        o += step;
                const uint32_t local = t - s_pre[wave][o];
                const uint32_t cy = s_ny[wave][o];
                const uint32_t dx = local / cy;
                const uint64_t x = (uint64_t)(uint32_t)(s_x0[wave][o] + (int32_t)dx);
                const uint64_t y = (uint64_t)(uint32_t)(s_y0[wave][o] + (int32_t)(local - dx * cy));
                mark_leaf_run(map, (x << (kl.bits[1] + kl.bits[2])) | (y << kl.bits[2]), s_z0[wave][o],
                              s_z1[wave][o]);
            }
        }
        for (uint64_t dx = 0; dx < own_nx; dx++)
            for (uint32_t dy = 0; dy < ny; dy++)
            {
                const uint64_t x = (uint64_t)(uint32_t)(x0 + (int32_t)dx), y = (uint64_t)(uint32_t)(y0 + (int32_t)dy);
                mark_leaf_run(map, (x << (kl.bits[1] + kl.bits[2])) | (y << kl.bits[2]), z0, z1);
            }
        wave_sync_lds(); // the next group overwrites the wave's rows
    }
}

__global__ __launch_bounds__(256) void k_leaf_popc(const unsigned long long* __restrict__ map, uint64_t words,
                                                   uint32_t* __restrict__ cnt)
{
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        cnt[w] = (uint32_t)__popcll(map[w]);
}

// rank -> leaf index for every marked leaf
__global__ __launch_bounds__(256) void k_leaf_unrank(const unsigned long long* __restrict__ map, uint64_t words,
                                                     const uint32_t* __restrict__ prefix, uint32_t* __restrict__ leaf_of_rank)
{
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
    {
        unsigned long long m = map[w];
        uint32_t r = prefix[w];
        while (m)
        {
            const uint32_t bit = (uint32_t)__builtin_ctzll(m);
            leaf_of_rank[r++] = (uint32_t)((w << 6) + bit);
            m &= m - 1ull;
        }
    }
}

static inline uint32_t bits_for(uint64_t range) // bits to represent values in [0, range)
{
    uint32_t b = 0;
    while (b < 64 && (1ull << b) < range)
        b++;
    return b;
}

// starts[runs] = n after the run starts (the end of the last run)
__global__ void k_store_end(uint32_t* __restrict__ starts, const uint32_t* __restrict__ runs, DCount cn)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        starts[*runs] = (uint32_t)cn.get();
}

// Run heads of keys >> shift: writes starts[0..runs) and starts[runs] = n when starts is not
// null; the run count goes to *runs_out (read back) and, through tile_offsets_out, the exclusive
// head counts of the HEAD_TILE-sized tiles (valid until the next run_heads call), whose total
// (the run count) sits in the device word *runs_dev_out. sorted: the keys are sorted over every
// bit from shift up, so the count reads only the segments that hold a head.
// With a device count n (graph capture) nothing is read back: *runs_out receives runs_cap, the
// bound the caller sizes by, and the count stays in the device word.
// runs_word, when given, receives the run count instead of the tile-count buffer (which the next
// run_heads call reuses): the device-count form keeps counts there that later stages read.
static int run_heads(Context* ctx, const uint64_t* keys, DCount n, uint32_t shift, uint32_t* starts,
                     uint64_t* runs_out, const uint32_t** runs_dev_out = nullptr,
                     const uint32_t** tile_offsets_out = nullptr, bool sorted = false, uint64_t runs_cap = 0,
                     uint32_t* runs_word = nullptr)
{
    const bool dev = n.dev != nullptr;
    if (!dev && n.cap == 0)
    {
        *runs_out = 0;
        uint32_t* tc = (uint32_t*)ctx->ws(WS_SEG_TILES, 2 * sizeof(uint32_t));
        if (!tc)
            return -1;
        if (fill_words(ctx, tc, host_count(2), 0u))
            return -1;
        if (starts && fill_words(ctx, starts, host_count(1), 0u))
            return -1;
        if (runs_word && fill_words(ctx, runs_word, host_count(1), 0u))
            return -1;
        if (runs_dev_out)
            *runs_dev_out = runs_word ? runs_word : tc;
        if (tile_offsets_out)
            *tile_offsets_out = tc + 1;
        return 0;
    }
    if (n.cap >= 0xFFFFFFFFull)
        return ctx->fail("run heads: more than 2^32-1 elements");
    const DCount tiles = count_div(n, HEAD_TILE);
    const uint64_t tcap = tiles.cap > 0 ? tiles.cap : 1;
    uint32_t* tile_counts = (uint32_t*)ctx->ws(WS_SEG_TILES, (tcap + 2) * sizeof(uint32_t));
    if (!tile_counts)
        return -1;
    uint32_t* total = runs_word ? runs_word : tile_counts + tcap + 1; // the run count
    // one wave (sorted) or one block per tile, at most a few thousand blocks striding over the rest
    const uint32_t wave_blocks = (uint32_t)std::min<uint64_t>((tcap + 3) / 4, 8192);
    const uint32_t tile_blocks = (uint32_t)std::min<uint64_t>(tcap, 8192);
    if (sorted)
        hipLaunchKernelGGL(k_head_count_sorted, dim3(wave_blocks), dim3(256), 0, ctx->stream, keys, n, shift,
                           tile_counts);
    else
        hipLaunchKernelGGL(k_head_count, dim3(tile_blocks), dim3(HEAD_THREADS), 0, ctx->stream, keys, n, shift,
                           tile_counts);
    if (scan_u32(ctx, tile_counts, tile_counts, tiles, false, total))
        return -1;
    if (starts && sorted)
        hipLaunchKernelGGL(k_head_write_sorted, dim3(wave_blocks), dim3(256), 0, ctx->stream, keys, n, shift,
                           (const uint32_t*)tile_counts, starts);
    else if (starts)
        hipLaunchKernelGGL(k_head_write, dim3(tile_blocks), dim3(HEAD_THREADS), 0, ctx->stream, keys, n, shift,
                           (const uint32_t*)tile_counts, starts);
    if (starts)
        hipLaunchKernelGGL(k_store_end, dim3(1), dim3(64), 0, ctx->stream, starts, (const uint32_t*)total, n);
    VDBR_CHECK(ctx, hipGetLastError());
    if (dev)
        *runs_out = runs_cap;
    else
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, total, 4, hipMemcpyDeviceToHost, ctx->stream));
        VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *runs_out = *(uint32_t*)ctx->pinned;
    }
    if (runs_dev_out)
        *runs_dev_out = total;
    if (tile_offsets_out)
        *tile_offsets_out = tile_counts;
    return 0;
}

struct Timer
{
    hipEvent_t ev[8];
    int n = 0;
};

// node count 1 + uppers + lowers + leaves (device-count form), checked against the node bound
__global__ void k_node_total(const uint32_t* __restrict__ cnt, uint32_t max_nodes, uint64_t leaf_cap, uint64_t lower_cap,
                             uint64_t upper_cap, uint32_t* __restrict__ node_count, uint32_t* __restrict__ flags)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const uint64_t nu = cnt[0], nl = cnt[1], nf = cnt[2];
    const uint64_t total = 1 + nu + nl + nf;
    *node_count = (uint32_t)(total < 0xFFFFFFFFull ? total : 0xFFFFFFFFull);
    if (nf > leaf_cap || nl > lower_cap || nu > upper_cap)
        atomicOr(flags, GF_LEAVES);
    if (total > max_nodes)
        atomicOr(flags, GF_NODES);
}

// grid layout from the node and voxel counts (device-count form of the host layout below)
__global__ void k_grid_info(GridInfo base, const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ nvox,
                            uint64_t nvox_cap, GridInfo* __restrict__ gi)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    GridInfo g = base;
    const uint64_t nu = cnt[0], nl = cnt[1], nf = cnt[2];
    const uint64_t nv = *nvox < nvox_cap ? *nvox : nvox_cap;
    g.end = g.nodes_begin + nu * VDB_SIZE_UPPER + nl * VDB_SIZE_LOWER + nf * VDB_SIZE_LEAF;
    g.meta0 = g.end;
    g.meta1 = g.end + VDB_BLINDMETADATA_SIZE;
    g.bboxes = g.meta1 + VDB_BLINDMETADATA_SIZE;
    g.values = g.bboxes + 32;
    g.value_count = 2 + nv;
    g.grid_size = g.values + 4 * g.value_count;
    g.upper_count = (uint32_t)nu;
    g.lower_count = (uint32_t)nl;
    g.leaf_count = (uint32_t)nf;
    g.node_count = (uint32_t)(1 + nu + nl + nf);
    *gi = g;
}

// Node2 grid from sorted unique voxel keys (the GridBuild stage of the raster path,
// GridBuild.cpp:361-650): leaves, lowers and uppers from key runs, node order by first-touch
// window (or tree order), node sizes and value indices by scans, then the grid written by one
// pass per level. The grid lives in the WS_GRID slot of the context.
// Host form: every count is read back and the buffers fit them exactly. Device-count form
// (in.nvox_dev set, for graph capture): no count leaves the device; buffers and launches are
// sized by the bounds of the input, kernels read the counts, and the layout is computed on the
// device (out->grid_size_dev).
int build_node2_grid(Context* ctx, const Node2BuildInput& in, Node2BuildOutput* out)
{
    hipStream_t st = ctx->stream;
    const bool devc = in.nvox_dev != nullptr;
    const KeyLayout kl = in.kl;
    const DCount nvox = devc ? dev_count(in.nvox_dev, in.nvox) : host_count(in.nvox);
    uint64_t* vox_key = in.vox_key;
    uint32_t* vox_val = in.vox_val;
    uint32_t* vox_t = in.vox_t;
    const int32_t* vbounds = in.vbounds;
    const uint32_t T = in.T;
    const uint32_t max_nodes = in.max_nodes;
    const float vs = in.voxel_size;
    const float vs_inv = 1.f / vs;
    const uint64_t nv1 = nvox.cap > 0 ? nvox.cap : 1;
    // device-count form: [0..2] upper, lower, leaf counts, [3] node count
    uint32_t* cnt = devc ? (uint32_t*)ctx->ws(WS_GB_COUNTS, 64) : nullptr;
    if (devc && !cnt)
        return -1;

    // ---------------------------------------------------------------- leaves / lowers / uppers
    uint32_t* leaf_first = (uint32_t*)ctx->ws(WS_LEAF_FIRST, (nv1 + 2) * sizeof(uint32_t));
    if (!leaf_first)
        return -1;
    uint64_t nleaf_c = 0;
    const uint32_t* nleaf_dev = nullptr;
    if (run_heads(ctx, vox_key, nvox, 9, leaf_first, &nleaf_c, &nleaf_dev, nullptr, false, in.leaf_cap,
                  devc ? cnt + 2 : nullptr))
        return -1;
    const DCount nleaf = devc ? dev_count(nleaf_dev, nleaf_c) : host_count(nleaf_c);
    const uint64_t nl1 = nleaf.cap > 0 ? nleaf.cap : 1;
    uint64_t* leaf_mask = (uint64_t*)ctx->ws(WS_LEAF_MASK, nl1 * 8 * sizeof(uint64_t));
    uint32_t* leaf_info = (uint32_t*)ctx->ws(WS_LEAF_INFO, nl1 * 6 * sizeof(uint32_t));
    uint32_t* leaf_t = leaf_info;               // compact order
    uint32_t* tree_t = leaf_info + nl1;         // tree order
    uint32_t* leaf_pos = leaf_info + 2 * nl1;   // compact -> tree position
    uint32_t* leaf_node_idx = leaf_info + 3 * nl1;
    uint32_t* leaf_parent = leaf_info + 4 * nl1;
    uint64_t* lkeys = (uint64_t*)ctx->ws(WS_LEAF_SORT_KEYS, nl1 * sizeof(uint64_t));
    uint32_t* lvals = (uint32_t*)ctx->ws(WS_LEAF_SORT_VALS, nl1 * sizeof(uint32_t));
    if (!leaf_mask || !leaf_info || !lkeys || !lvals)
        return -1;
    if (nleaf.cap > 0)
    {
        hipLaunchKernelGGL(k_leaf_info, dim3((uint32_t)((nleaf.cap + 3) / 4)), dim3(256), 0, st, vox_key, vox_t,
                           leaf_first, nleaf, kl, leaf_mask, leaf_t, lkeys, lvals);
        if (radix_sort_u64(ctx, lkeys, lvals, nleaf, 0, 42))
            return -1;
        hipLaunchKernelGGL(k_leaf_gather, dim3(blocks_for(nleaf.cap)), dim3(256), 0, st, lkeys, lvals, nleaf, leaf_t,
                           tree_t, leaf_pos);
    }
    // lowers: runs of tree key >> 12 over the tree-ordered leaves
    uint32_t* lower_first = (uint32_t*)ctx->ws(WS_LOWER, (nl1 + 2) * sizeof(uint32_t) * 4);
    if (!lower_first)
        return -1;
    uint64_t nlower_c = 0;
    const uint32_t* nlower_dev = nullptr;
    if (run_heads(ctx, lkeys, nleaf, 12, lower_first, &nlower_c, &nlower_dev, nullptr, false, in.lower_cap,
                  devc ? cnt + 1 : nullptr))
        return -1;
    const DCount nlower = devc ? dev_count(nlower_dev, nlower_c) : host_count(nlower_c);
    const uint64_t nlo1 = nlower.cap > 0 ? nlower.cap : 1;
    uint64_t* lower_key = (uint64_t*)ctx->ws(WS_LOWER_MASK, nlo1 * sizeof(uint64_t) * 2);
    uint32_t* lower_t = lower_first + (nl1 + 2);
    uint32_t* lower_node_idx = lower_t + nlo1;
    uint32_t* lower_parent = lower_node_idx + nlo1;
    if (!lower_key)
        return -1;
    if (nlower.cap > 0)
    {
        hipLaunchKernelGGL(k_run_keys, dim3(blocks_for(nlower.cap)), dim3(256), 0, st, lkeys, lower_first, nlower, 12u,
                           lower_key);
        hipLaunchKernelGGL(k_run_min, dim3((uint32_t)((nlower.cap + 3) / 4)), dim3(256), 0, st, tree_t, lower_first,
                           nlower, lower_t);
    }
    uint32_t* upper_first = (uint32_t*)ctx->ws(WS_UPPER, (nlo1 + 2) * sizeof(uint32_t) * 4);
    if (!upper_first)
        return -1;
    uint64_t nupper_c = 0;
    const uint32_t* nupper_dev = nullptr;
    if (run_heads(ctx, lower_key, nlower, 15, upper_first, &nupper_c, &nupper_dev, nullptr, false, in.upper_cap,
                  devc ? cnt : nullptr))
        return -1;
    const DCount nupper = devc ? dev_count(nupper_dev, nupper_c) : host_count(nupper_c);
    const uint64_t nu1 = nupper.cap > 0 ? nupper.cap : 1;
    uint64_t* upper_key = (uint64_t*)ctx->ws(WS_UPPER_MASK, nu1 * sizeof(uint64_t));
    uint32_t* upper_t = upper_first + (nlo1 + 2);
    uint32_t* upper_node_idx = upper_t + nu1;
    if (!upper_key)
        return -1;
    if (nupper.cap > 0)
    {
        hipLaunchKernelGGL(k_run_keys, dim3(blocks_for(nupper.cap)), dim3(256), 0, st, lower_key, upper_first, nupper,
                           15u, upper_key);
        hipLaunchKernelGGL(k_run_min, dim3((uint32_t)((nupper.cap + 3) / 4)), dim3(256), 0, st, lower_t, upper_first,
                           nupper, upper_t);
    }
    const uint64_t node_count = 1 + nupper.cap + nlower.cap + nleaf.cap; // exact, or the bound
    if (!devc && node_count > max_nodes)
        return ctx->fail("raster_node2: node count exceeds max_node_count");
    DCount nodes = host_count(node_count);
    if (devc)
    {
        // the three run counts sit side by side in cnt: the node total and its checks
        hipLaunchKernelGGL(k_node_total, dim3(1), dim3(64), 0, st, (const uint32_t*)cnt, max_nodes, nleaf.cap,
                           nlower.cap, nupper.cap, cnt + 3, in.flags_dev);
        nodes = dev_count((const uint32_t*)(cnt + 3), node_count);
    }

    // ---------------------------------------------------------------- node order
    bool spatial = in.spatial_order || T <= 1;
    if (!spatial && node_count > 1)
    {
        // histogram of first-touch windows per level
        uint32_t* hist = (uint32_t*)ctx->ws(WS_HIST, (size_t)(3 * T + 4 * (T + 1)) * sizeof(uint32_t));
        if (!hist)
            return -1;
        uint32_t* tot = hist + 3 * T;
        uint32_t* tot_before = tot + (T + 1);
        uint32_t* lev_before = tot_before + (T + 1); // 2 arrays of T+1 used per level in turn
        if (fill_words(ctx, hist, host_count((uint64_t)3 * T), 0u))
            return -1;
        if (nupper.cap)
            hipLaunchKernelGGL(k_window_hist, dim3(std::min(blocks_for(nupper.cap), NODE_GRID_BLOCKS)), dim3(256), 0, st,
                               upper_t, nupper, T, hist);
        if (nlower.cap)
            hipLaunchKernelGGL(k_window_hist, dim3(std::min(blocks_for(nlower.cap), NODE_GRID_BLOCKS)), dim3(256), 0, st,
                               lower_t, nlower, T,
                               hist + T);
        if (nleaf.cap)
            hipLaunchKernelGGL(k_window_hist, dim3(std::min(blocks_for(nleaf.cap), NODE_GRID_BLOCKS)), dim3(256), 0, st,
                               tree_t, nleaf, T,
                               hist + 2 * T);
        hipLaunchKernelGGL(k_window_totals, dim3(blocks_for(T)), dim3(256), 0, st, hist, T, tot);
        if (scan_u32(ctx, tot, tot_before, (uint64_t)T, false, nullptr))
            return -1;
        uint64_t* nkeys = (uint64_t*)ctx->ws(WS_NODE_KEYS, nl1 * sizeof(uint64_t));
        uint32_t* nvals = (uint32_t*)ctx->ws(WS_NODE_VALS, nl1 * sizeof(uint32_t));
        if (!nkeys || !nvals)
            return -1;
        struct Level
        {
            DCount count;
            const uint32_t* t;
            const uint32_t* parent;
            const uint64_t* slot_key;
            uint32_t slot_bits;
            uint32_t* node_idx;
        };
        // parents: upper -> root (0); lower -> its upper; leaf -> its lower
        Level levels[3] = { { nupper, upper_t, nullptr, upper_key, 15u, upper_node_idx },
                            { nlower, lower_t, lower_parent, lower_key, 15u, lower_node_idx },
                            { nleaf, tree_t, leaf_parent, lkeys, 12u, leaf_node_idx } };
        // key fields as wide as their values: windows t < T, parents < node_count + 1 (none for the
        // uppers, whose parent is the root), so the sorts take no pass over bits that are all zero
        const uint32_t tbits = bits_for(T);
        const uint32_t parent_bits = bits_for(node_count + 1);
        if (tbits + parent_bits + 15 > 64)
            return ctx->fail("raster_node2: node order keys exceed 64 bits");
        for (uint32_t lv = 0; lv < 3; lv++)
        {
            Level& L = levels[lv];
            if (L.count.cap == 0)
                continue;
            if (lv == 1)
                hipLaunchKernelGGL(k_parent_idx, dim3(std::min(blocks_for(nupper.cap, 4u), NODE_GRID_BLOCKS)), dim3(256), 0,
                                   st, upper_first, nupper,
                                   upper_node_idx, lower_parent);
            if (lv == 2)
                hipLaunchKernelGGL(k_parent_idx, dim3(std::min(blocks_for(nlower.cap, 4u), NODE_GRID_BLOCKS)), dim3(256), 0,
                                   st, lower_first, nlower,
                                   lower_node_idx, leaf_parent);
            if (scan_u32(ctx, hist + lv * T, lev_before, (uint64_t)T, false, nullptr))
                return -1;
            const uint32_t pbits = lv == 0 ? 0u : parent_bits;
            hipLaunchKernelGGL(k_node_keys, dim3(blocks_for(L.count.cap)), dim3(256), 0, st, L.t,
                               lv == 0 ? nullptr : L.parent, L.slot_key, L.count, L.slot_bits, pbits, nkeys, nvals);
            if (radix_sort_u64(ctx, nkeys, nvals, L.count, 0, tbits + pbits + L.slot_bits))
                return -1;
            hipLaunchKernelGGL(k_node_assign, dim3(blocks_for(L.count.cap)), dim3(256), 0, st, nkeys, nvals, L.count,
                               pbits + L.slot_bits, tot_before, lev_before, hist, T, lv, L.node_idx);
        }
    }
    else
    {
        const DCount none = host_count(0);
        if (nupper.cap)
            hipLaunchKernelGGL(k_node_bfs, dim3(blocks_for(nupper.cap)), dim3(256), 0, st, nupper, 1u, none, none,
                               upper_node_idx);
        if (nlower.cap)
            hipLaunchKernelGGL(k_node_bfs, dim3(blocks_for(nlower.cap)), dim3(256), 0, st, nlower, 1u, nupper, none,
                               lower_node_idx);
        if (nleaf.cap)
            hipLaunchKernelGGL(k_node_bfs, dim3(blocks_for(nleaf.cap)), dim3(256), 0, st, nleaf, 1u, nupper, nlower,
                               leaf_node_idx);
    }
    // sizes and value counts by node idx -> exclusive scans
    uint64_t* byidx = (uint64_t*)ctx->ws(WS_NODE_BYIDX, (node_count + 1) * 2 * sizeof(uint64_t));
    uint64_t* nscan = (uint64_t*)ctx->ws(WS_NODE_SCAN, (node_count + 1) * 2 * sizeof(uint64_t));
    if (!byidx || !nscan)
        return -1;
    uint64_t* size_by_idx = byidx;
    uint64_t* vals_by_idx = byidx + (node_count + 1);
    uint64_t* size_scan = nscan;
    uint64_t* vals_scan = nscan + (node_count + 1);
    if (fill_words(ctx, byidx, host_count((node_count + 1) * 2), 0ull))
        return -1;
    if (nupper.cap)
        hipLaunchKernelGGL(k_node_sizes, dim3(blocks_for(nupper.cap)), dim3(256), 0, st, upper_node_idx, nupper,
                           (uint32_t)node_count, VDB_SIZE_UPPER, nullptr, nullptr, size_by_idx, vals_by_idx);
    if (nlower.cap)
        hipLaunchKernelGGL(k_node_sizes, dim3(blocks_for(nlower.cap)), dim3(256), 0, st, lower_node_idx, nlower,
                           (uint32_t)node_count, VDB_SIZE_LOWER, nullptr, nullptr, size_by_idx, vals_by_idx);
    if (nleaf.cap)
        hipLaunchKernelGGL(k_node_sizes, dim3(blocks_for(nleaf.cap)), dim3(256), 0, st, leaf_node_idx, nleaf,
                           (uint32_t)node_count, VDB_SIZE_LEAF, leaf_mask, lvals, size_by_idx, vals_by_idx);
    // node idx 0 is the root: its slot holds size 0 so the scan at idx 1 starts at 0
    if (scan_u64(ctx, size_by_idx, size_scan, nodes, false, size_scan + node_count))
        return -1;
    if (scan_u64(ctx, vals_by_idx, vals_scan, nodes, false, vals_scan + node_count))
        return -1;

    // ---------------------------------------------------------------- grid layout
    GridInfo gi;
    memset(&gi, 0, sizeof(gi));
    gi.root_addr = VDB_ROOT_ADDR;
    gi.nodes_begin = VDB_ROOT_ADDR + VDB_SIZE_ROOT;
    gi.end = gi.nodes_begin + nupper.cap * VDB_SIZE_UPPER + nlower.cap * VDB_SIZE_LOWER + nleaf.cap * VDB_SIZE_LEAF;
    gi.meta0 = gi.end;
    gi.meta1 = gi.end + VDB_BLINDMETADATA_SIZE;
    gi.bboxes = gi.meta1 + VDB_BLINDMETADATA_SIZE;
    gi.values = gi.bboxes + 32;
    gi.value_count = 2 + nvox.cap;
    gi.grid_size = gi.values + 4 * gi.value_count; // exact, or the bound
    gi.upper_count = (uint32_t)nupper.cap;
    gi.lower_count = (uint32_t)nlower.cap;
    gi.leaf_count = (uint32_t)nleaf.cap;
    gi.value_type = in.value_type;
    gi.node_count = (uint32_t)node_count;
    gi.voxel_size = vs;
    gi.voxel_size_inv = vs_inv;
    GridInfo* gi_dev = nullptr;
    if (devc)
    {
        gi_dev = (GridInfo*)ctx->ws(WS_GB_INFO, sizeof(GridInfo));
        if (!gi_dev)
            return -1;
        hipLaunchKernelGGL(k_grid_info, dim3(1), dim3(64), 0, st, gi, (const uint32_t*)cnt, in.nvox_dev, nvox.cap, gi_dev);
    }
    uint8_t* grid = (uint8_t*)ctx->ws(WS_GRID, (gi.grid_size + 7) & ~7ull);
    if (!grid)
        return -1;
    if (devc)
    {
        DCount gw = count_div(dev_count(&gi_dev->grid_size, gi.grid_size), 8u);
        if (fill_words(ctx, (uint64_t*)grid, gw, 0ull))
            return -1;
    }
    else if (fill_words(ctx, (uint64_t*)grid, host_count((gi.grid_size + 7) / 8), 0ull))
        return -1;
    hipLaunchKernelGGL(k_write_header, dim3(1), dim3(64), 0, st, grid, gi, (const GridInfo*)gi_dev,
                       (const int32_t*)vbounds);
    hipLaunchKernelGGL(k_write_root, dim3(1), dim3(1024), 0, st, grid, gi, (const GridInfo*)gi_dev,
                       (const uint64_t*)upper_key, nupper, (const uint32_t*)upper_node_idx, (const uint64_t*)size_scan);
    if (nupper.cap)
        hipLaunchKernelGGL((k_write_internal<512>), dim3((uint32_t)nupper.cap), dim3(256), 0, st, grid, gi,
                           (const GridInfo*)gi_dev, (const uint32_t*)upper_first, (const uint64_t*)lower_key, 15u,
                           (const uint32_t*)upper_node_idx, (const uint32_t*)lower_node_idx,
                           (const uint64_t*)size_scan, (const uint64_t*)vals_scan, VDB_OFF_CMP_UPPER, VDB_OFF_CM_UPPER,
                           VDB_OFF_CHILDREN_UPPER, nupper);
    if (nlower.cap)
        hipLaunchKernelGGL((k_write_internal<64>), dim3((uint32_t)nlower.cap), dim3(256), 0, st, grid, gi,
                           (const GridInfo*)gi_dev, (const uint32_t*)lower_first, (const uint64_t*)lkeys, 12u,
                           (const uint32_t*)lower_node_idx, (const uint32_t*)leaf_node_idx, (const uint64_t*)size_scan,
                           (const uint64_t*)vals_scan, VDB_OFF_CMP_LOWER, VDB_OFF_CM_LOWER, VDB_OFF_CHILDREN_LOWER,
                           nlower);
    if (nleaf.cap)
        hipLaunchKernelGGL(k_write_leaves, dim3(blocks_for(nleaf.cap)), dim3(256), 0, st, grid, gi,
                           (const GridInfo*)gi_dev, nleaf, (const uint32_t*)lvals, (const uint64_t*)leaf_mask,
                           (const uint32_t*)leaf_node_idx, (const uint64_t*)size_scan, (const uint64_t*)vals_scan);
    int32_t* ijk_out = nullptr;
    ctx->last_voxel_ijk = nullptr;
    ctx->last_value_count = 0;
    if (in.voxel_coords)
    {
        if (devc)
            return ctx->fail("raster_node2: voxel coordinates are not available in the device-count form");
        ijk_out = (int32_t*)ctx->ws(WS_VOX_IJK, 3 * (2 + nvox.cap) * sizeof(int32_t));
        if (!ijk_out)
            return -1;
        // the two background slots have no voxel
        VDBR_CHECK(ctx, hipMemsetAsync(ijk_out, 0x80, 6 * sizeof(int32_t), st));
        ctx->last_voxel_ijk = ijk_out;
        ctx->last_value_count = 2 + nvox.cap;
    }
    if (nvox.cap && nleaf.cap)
        hipLaunchKernelGGL(k_write_values, dim3(std::min<uint32_t>(blocks_for(nleaf.cap, 4u), 16384u)), dim3(256), 0, st,
                           grid, gi, (const GridInfo*)gi_dev, nleaf, (const uint32_t*)vox_val, (const uint64_t*)vox_key,
                           (const uint32_t*)leaf_first, (const uint32_t*)leaf_pos, (const uint32_t*)leaf_node_idx,
                           (const uint64_t*)vals_scan, kl, ijk_out);
    VDBR_CHECK(ctx, hipGetLastError());

    out->grid = grid;
    out->grid_size = gi.grid_size;
    out->upper_count = nupper.cap;
    out->lower_count = nlower.cap;
    out->leaf_count = nleaf.cap;
    out->grid_size_dev = devc ? &gi_dev->grid_size : nullptr;
    out->counts_dev = devc ? cnt : nullptr;
    return 0;
}

// ============================================================================ setup stage
__global__ void k_init_setup(int32_t* __restrict__ bounds, int32_t* __restrict__ vbounds, unsigned long long* __restrict__ est)
{
    if (threadIdx.x < 3)
    {
        bounds[threadIdx.x] = 2147483647;
        vbounds[threadIdx.x] = 2147483647;
    }
    else if (threadIdx.x < 6)
    {
        bounds[threadIdx.x] = -2147483647;
        vbounds[threadIdx.x] = -2147483647;
    }
    else if (threadIdx.x < 8)
        est[threadIdx.x - 6] = 0ull;
}

// Per-Gaussian buffers of one raster call and the words its setup reduces into.
struct SetupBufs
{
    PrimFrame* frames;
    PrimBox* boxes;
    PrimVal* pvals;
    uint64_t* foot;
    uint64_t* foot_scan; // n + 1 entries (exclusive scan, total last)
    uint64_t* rows;
    uint64_t* row_scan; // n + 1 entries
    int32_t* bounds;    // AABB voxel bounds (min xyz, max xyz)
    int32_t* vbounds;   // active-voxel bounds, filled by the reduce
    unsigned long long* est; // [0] pair estimate, [1] cleared
};

// optional in-place preprocessing (RasterUtils.cpp:59-117), prim setup (gaussian_prim.slang),
// footprint and row scans; no host synchronisation
static int launch_setup(Context* ctx, const vdbr_gaussians_t* gs, const vdbr_raster_params_t* prm, const SetupBufs& b,
                        bool preprocess)
{
    hipStream_t st = ctx->stream;
    const uint64_t n = gs->count;
    const float vs = prm->voxel_size;
    if (preprocess && n > 0)
        hipLaunchKernelGGL(k_preprocess, dim3(blocks_for(n)), dim3(256), 0, st, n, gs->opacities, gs->quats, gs->scales,
                           gs->sh0, gs->colors);
    hipLaunchKernelGGL(k_init_setup, dim3(1), dim3(64), 0, st, b.bounds, b.vbounds, b.est);
    if (n > 0)
    {
        SetupOut so{ b.frames, b.boxes, b.pvals, b.foot, b.rows, b.bounds, b.est };
        hipLaunchKernelGGL(k_setup, dim3(std::min(blocks_for(n), SETUP_BLOCKS)), dim3(256), 0, st, n, gs->means,
                           gs->quats, gs->scales,
                           gs->opacities, gs->colors, vs, 1.f / vs, prm->cull_scale, so);
        if (scan_u64(ctx, b.foot, b.foot_scan, n, false, b.foot_scan + n))
            return -1;
        if (scan_u64(ctx, b.rows, b.row_scan, n, false, b.row_scan + n))
            return -1;
    }
    else if (fill_words(ctx, b.foot_scan, host_count(1), 0ull) || fill_words(ctx, b.row_scan, host_count(1), 0ull))
        return -1;
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// ============================================================================ reduce stage
// A device count past the bound its buffers were sized by sets a flag bit and zeroes the counts
// that later kernels of the stage derive write positions from (kill32: a u32 count, kill64: a
// u64 count), so the rest of the captured work runs empty instead of writing out of bounds.
__global__ void k_check_bound(uint32_t* __restrict__ word, uint64_t cap, uint32_t bit, uint32_t* __restrict__ flags,
                              uint32_t* __restrict__ kill32, uint64_t* __restrict__ kill64)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && (uint64_t)*word > cap)
    {
        atomicOr(flags, bit);
        *word = 0u;
        if (kill32)
            *kill32 = 0u;
        if (kill64)
            *kill64 = 0ull;
    }
}

// item_scan[*index] = item_scan[at] (moves a scan total to the position a device count names)
__global__ void k_move_total(uint32_t* __restrict__ item_scan, const uint32_t* __restrict__ index, uint64_t at)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        item_scan[*index] = item_scan[at];
}

// Per-leaf reduce of one slab's sorted pairs (the pairs are sorted over their leaf bits): leaf
// runs, staging arrays of 512 slots per leaf (maximum packed RGBA8 value or 192-bit density sum,
// lowest Gaussian id), the tile or chunk reduce, and the active-voxel count per leaf with its
// scan. Host form (pairs a host count): the leaf and voxel counts are read back and the buffers
// fit them. Device form: the counts stay on the device (leaves, svox) and buffers are sized by
// leaf_cap and vox_cap; a count past its bound sets a bit in *flags.
struct LeafReduce
{
    // in
    const uint64_t* sorted = nullptr;
    DCount pairs = host_count(0);
    KeyLayout kl{};
    float vs = 0.f;
    bool rgba_wave = false, density = false;
    const PrimFrame* frames = nullptr;
    const PrimVal* pvals = nullptr;
    const DensPrim* dprims = nullptr;
    uint64_t leaf_cap = 0, vox_cap = 0; // device form
    uint32_t* flags = nullptr;          // device form
    uint32_t* leaves_word = nullptr;    // device form: where the leaf count is kept
    hipEvent_t ev_begin = nullptr, ev_end = nullptr; // host form: reduce kernel timing
    // out
    DCount leaves = host_count(0);
    DCount svox = host_count(0);
    uint32_t* lval = nullptr;
    unsigned long long* lw = nullptr;
    uint32_t* lg0 = nullptr;
    uint64_t* lkey = nullptr;
    uint32_t* lcnt = nullptr;
};

static int reduce_leaves(Context* ctx, LeafReduce* r)
{
    hipStream_t st = ctx->stream;
    const bool devc = r->pairs.dev != nullptr;
    const KeyLayout kl = r->kl;
    const uint32_t lshift = kl.gid_bits + kl.alpha_bits + 9u;
    const uint32_t* tile_offsets = nullptr;
    const uint32_t* leaves_dev = nullptr;
    uint64_t sleaves = 0;
    if (run_heads(ctx, r->sorted, r->pairs, lshift, nullptr, &sleaves, &leaves_dev, &tile_offsets, true, r->leaf_cap,
                  r->leaves_word))
        return -1;
    r->leaves = devc ? dev_count(leaves_dev, sleaves) : host_count(sleaves);
    r->svox = host_count(0);
    if (sleaves == 0)
        return 0;
    if (devc)
        hipLaunchKernelGGL(k_check_bound, dim3(1), dim3(64), 0, st, (uint32_t*)leaves_dev, sleaves, GF_LEAVES, r->flags,
                           (uint32_t*)nullptr, (uint64_t*)r->pairs.dev);
    const DCount leaves = r->leaves;
    if (!r->density)
        r->lval = (uint32_t*)ctx->ws(WS_LR_VAL, sleaves * 512u * sizeof(uint32_t));
    else
        r->lw = (unsigned long long*)ctx->ws(WS_LR_VAL, sleaves * 512u * 3u * sizeof(uint64_t));
    r->lg0 = (uint32_t*)ctx->ws(WS_LR_G0, sleaves * 512u * sizeof(uint32_t));
    r->lkey = (uint64_t*)ctx->ws(WS_LR_KEY, sleaves * sizeof(uint64_t));
    r->lcnt = (uint32_t*)ctx->ws(WS_LR_CNT, (sleaves + 1) * sizeof(uint32_t));
    if ((!r->lval && !r->lw) || !r->lg0 || !r->lkey || !r->lcnt)
        return -1;
    if (!r->density)
    {
        if (fill_words(ctx, r->lval, count_mul(leaves, 512u), 0u))
            return -1;
    }
    else if (fill_words(ctx, (uint64_t*)r->lw, count_mul(leaves, 512u * 3u), 0ull))
        return -1;
    if (fill_words(ctx, r->lg0, count_mul(leaves, 512u), 0xFFFFFFFFu))
        return -1;
    const uint64_t tiles = (r->pairs.cap + HEAD_TILE - 1) / HEAD_TILE;
    if (r->ev_begin && !ctx->capturing)
        VDBR_CHECK(ctx, hipEventRecord(r->ev_begin, st));
    if (!r->rgba_wave && !r->density)
        hipLaunchKernelGGL(k_leaf_reduce_rgba<false>, dim3((uint32_t)tiles), dim3(RED_THREADS), 0, st, r->sorted,
                           r->pairs, kl, r->vs, tile_offsets, r->frames, r->pvals, r->lval, r->lg0, r->lkey);
    else
    {
        // leaf starts (the sparse head pass over the sorted keys), chunks per leaf, one block per
        // chunk (RGBA8 with the alpha in the key, or density)
        const uint32_t chunk = r->density ? LRD_CHUNK : LRC_CHUNK;
        uint32_t* lstarts = (uint32_t*)ctx->ws(WS_LR_STARTS, (sleaves + 1) * sizeof(uint32_t));
        uint32_t* litems = (uint32_t*)ctx->ws(WS_LR_ITEMS, (sleaves + 1) * sizeof(uint32_t));
        if (!lstarts || !litems)
            return -1;
        hipLaunchKernelGGL(k_head_write_sorted, dim3((uint32_t)std::min<uint64_t>((tiles + 3) / 4, 8192)), dim3(256), 0, st,
                           r->sorted, r->pairs, lshift, tile_offsets, lstarts);
        if (devc)
            hipLaunchKernelGGL(k_store_end, dim3(1), dim3(64), 0, st, lstarts, leaves_dev, r->pairs);
        else
            hipLaunchKernelGGL(k_store_u32, dim3(1), dim3(64), 0, st, lstarts + sleaves, (uint32_t)r->pairs.cap);
        hipLaunchKernelGGL(k_leaf_chunks, dim3(blocks_for(sleaves)), dim3(256), 0, st, (const uint32_t*)lstarts, leaves,
                           chunk, litems);
        // the item scan's total goes after the last leaf (the chunk reduces read item_scan[nleaf])
        if (scan_u32(ctx, litems, litems, leaves, false, litems + sleaves))
            return -1;
        if (devc)
            hipLaunchKernelGGL(k_move_total, dim3(1), dim3(64), 0, st, litems, leaves_dev, sleaves);
        const uint64_t items_max = sleaves + r->pairs.cap / chunk + 1;
        if (items_max > 0x7FFFFFFFull)
            return ctx->fail("raster_node2: too many reduce chunks");
        if (r->density)
            hipLaunchKernelGGL(k_leaf_reduce_density, dim3((uint32_t)items_max), dim3(RED_THREADS), 0, st, r->sorted, kl,
                               r->vs, (const uint32_t*)lstarts, (const uint32_t*)litems, leaves, r->dprims, r->lw,
                               r->lg0, r->lkey);
        else
            hipLaunchKernelGGL(k_leaf_reduce_rgba_chunks, dim3((uint32_t)items_max), dim3(LRC_THREADS), 0, st, r->sorted,
                               kl, (const uint32_t*)lstarts, (const uint32_t*)litems, leaves, r->pvals, r->lval, r->lg0,
                               r->lkey);
    }
    VDBR_CHECK(ctx, hipGetLastError());
    if (r->ev_end && !ctx->capturing)
        VDBR_CHECK(ctx, hipEventRecord(r->ev_end, st));
    const uint32_t lblocks = (uint32_t)((sleaves + 3) / 4);
    hipLaunchKernelGGL(k_leaf_count, dim3(lblocks), dim3(256), 0, st, (const uint32_t*)r->lg0, leaves, r->lcnt);
    if (scan_u32(ctx, r->lcnt, r->lcnt, leaves, false, r->lcnt + sleaves))
        return -1;
    if (devc)
    {
        hipLaunchKernelGGL(k_check_bound, dim3(1), dim3(64), 0, st, r->lcnt + sleaves, r->vox_cap, GF_VOXELS, r->flags,
                           (uint32_t*)leaves_dev, (uint64_t*)nullptr);
        r->svox = dev_count((const uint32_t*)(r->lcnt + sleaves), r->vox_cap);
    }
    else
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, r->lcnt + sleaves, 4, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        r->svox = host_count(*(const uint32_t*)ctx->pinned);
    }
    return 0;
}

struct VoxelLists
{
    uint64_t* key;
    uint32_t* val;
    uint32_t* t;
    uint32_t* g0;
};

// the active voxels of the reduced leaves in key order, their first-touch windows and bounds
static int write_voxels(Context* ctx, const LeafReduce& r, const VoxelLists& out, const PrimBox* boxes,
                        const uint64_t* foot_scan, const uint32_t* win_base_dev, uint32_t PB, uint32_t W, int32_t* vbounds)
{
    hipStream_t st = ctx->stream;
    const uint32_t lblocks = (uint32_t)((r.leaves.cap + 3) / 4);
    if (!r.density)
        hipLaunchKernelGGL(k_leaf_write, dim3(lblocks), dim3(256), 0, st, (const uint32_t*)r.lval, (const uint32_t*)r.lg0,
                           (const uint64_t*)r.lkey, (const uint32_t*)r.lcnt, r.leaves, out.key, out.val, out.g0);
    else
        hipLaunchKernelGGL(k_leaf_write_density, dim3(lblocks), dim3(256), 0, st, (const unsigned long long*)r.lw,
                           (const uint32_t*)r.lg0, (const uint64_t*)r.lkey, (const uint32_t*)r.lcnt, r.leaves, out.key,
                           out.val, out.g0);
    VDBR_CHECK(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_voxel_finish, dim3(std::min<uint32_t>(blocks_for(r.svox.cap), 2048u)), dim3(256), 0, st, r.svox,
                       r.kl, (const uint64_t*)out.key, (const uint32_t*)out.g0, boxes, foot_scan, win_base_dev, PB, W,
                       out.t, vbounds);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// Candidate-leaf ranks for the pair keys: the leaves of every Gaussian's voxel AABB are marked in
// a bitmap over the leaf index, and a pair's leaf field becomes the leaf's rank among them (ranks
// follow the leaf index order, so the sort, the reduce and the voxel lists see the same order).
// Used when the ranks need fewer bits than the leaf index (at 2M clustered Gaussians 191,382
// candidates: 18 bits instead of 21, two 9-bit sort passes instead of three) and the bitmap stays
// within 2^28 bits. On return kl holds the rank layout, or is unchanged.
static int leaf_ranks(Context* ctx, hipStream_t st, const PrimBox* boxes, uint64_t n, KeyLayout* kl)
{
    const uint32_t index_bits = kl->bits[0] + kl->bits[1] + kl->bits[2];
    if (index_bits < 10u || index_bits > 28u)
        return 0;
    const uint64_t words = ((1ull << index_bits) + 63u) / 64u;
    unsigned long long* map = (unsigned long long*)ctx->ws(WS_LEAF_MAP, words * sizeof(uint64_t));
    // prefix (words + 1), per-word counts (words), total (1)
    uint32_t* prefix = (uint32_t*)ctx->ws(WS_LEAF_PREFIX, (2 * words + 2) * sizeof(uint32_t));
    if (!map || !prefix)
        return -1;
    uint32_t* counts = prefix + words + 1;
    uint32_t* total = counts + words;
    if (fill_words(ctx, (uint64_t*)map, host_count(words), 0ull))
        return -1;
    hipLaunchKernelGGL(k_leaf_mark, dim3((uint32_t)std::min<uint64_t>(blocks_for(n), 4096u)), dim3(256), 0, st, boxes,
                       n, *kl, map);
    hipLaunchKernelGGL(k_leaf_popc, dim3((uint32_t)std::min<uint64_t>(blocks_for(words), 4096u)), dim3(256), 0, st,
                       (const unsigned long long*)map, words, counts);
    VDBR_CHECK(ctx, hipGetLastError());
    if (scan_u32(ctx, counts, prefix, words, false, total))
        return -1;
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, total, 4, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint32_t cand = (uint32_t)ctx->pinned[0];
    const uint32_t rank_bits = std::max<uint32_t>(1u, bits_for(cand));
    if (cand == 0u || rank_bits >= index_bits)
        return 0;
    uint32_t* of_rank = (uint32_t*)ctx->ws(WS_LEAF_RANK, (uint64_t)cand * sizeof(uint32_t));
    if (!of_rank)
        return -1;
    hipLaunchKernelGGL(k_leaf_unrank, dim3((uint32_t)std::min<uint64_t>(blocks_for(words), 4096u)), dim3(256), 0, st,
                       (const unsigned long long*)map, words, (const uint32_t*)prefix, of_rank);
    VDBR_CHECK(ctx, hipGetLastError());
    kl->rank_bits = rank_bits;
    kl->leaf_map = (const uint64_t*)map;
    kl->leaf_prefix = prefix;
    kl->leaf_of_rank = of_rank;
    kl->voxel_bits = rank_bits + 9u;
    return 0;
}

int raster_node2(Context* ctx, const vdbr_gaussians_t* gs, const vdbr_raster_params_t* prm, void** grid_dev,
                 uint64_t* grid_size_out, vdbr_raster_stats_t* stats)
{
    if (!gs || !prm || !grid_dev || !grid_size_out)
        return ctx->fail("raster_node2: null argument");
    const uint64_t n = gs->count;
    const float vs = prm->voxel_size;
    if (!(vs > 0.f) || !(vs < 3.0e38f))
        return ctx->fail("raster_node2: voxel_size must be a positive finite number");
    if (n >= (1ull << 32))
        return ctx->fail("raster_node2: more than 2^32-1 Gaussians");
    if (n > 0 && (!gs->means || !gs->opacities || !gs->quats || !gs->scales || !gs->colors))
        return ctx->fail("raster_node2: null Gaussian array");
    if (prm->value_mode != VDBR_VALUE_RGBA8 && prm->value_mode != VDBR_VALUE_DENSITY)
        return ctx->fail("raster_node2: unknown value_mode");
    ctx->last_voxel_ijk = nullptr;
    ctx->last_value_count = 0;
    ctx->drop_pairs();
    const uint32_t PB = prm->prim_batch_size ? prm->prim_batch_size : 262144u;
    const uint32_t W = prm->ijk_batch_size ? prm->ijk_batch_size : 8388608u;
    const uint32_t max_nodes = prm->max_node_count ? prm->max_node_count : 2097152u;
    hipStream_t st = ctx->stream;

    const int NEV = 16;
    hipEvent_t ev[NEV];
    for (int e = 0; e < NEV; e++)
        VDBR_CHECK(ctx, hipEventCreate(&ev[e]));
    struct EvGuard
    {
        hipEvent_t* e;
        ~EvGuard()
        {
            for (int i = 0; i < 16; i++)
                (void)hipEventDestroy(e[i]);
        }
    } evg{ ev };
    uint32_t sort_passes = 0;
    VDBR_CHECK(ctx, hipEventRecord(ev[0], st));

    // ---------------------------------------------------------------- preprocess + setup
    if ((prm->flags & VDBR_FLAG_RAW_INPUTS) && n > 0 && !gs->sh0)
        return ctx->fail("raster_node2: raw inputs need sh0");
    const uint64_t nn = n > 0 ? n : 1;
    SetupBufs sb;
    sb.frames = (PrimFrame*)ctx->ws(WS_FRAMES, nn * sizeof(PrimFrame));
    sb.boxes = (PrimBox*)ctx->ws(WS_BOXES, nn * sizeof(PrimBox));
    sb.pvals = (PrimVal*)ctx->ws(WS_VALS, nn * sizeof(PrimVal));
    sb.foot = (uint64_t*)ctx->ws(WS_FOOT, nn * sizeof(uint64_t));
    sb.foot_scan = (uint64_t*)ctx->ws(WS_FOOT_SCAN, (nn + 1) * sizeof(uint64_t));
    sb.rows = (uint64_t*)ctx->ws(WS_ROWS, nn * sizeof(uint64_t));
    sb.row_scan = (uint64_t*)ctx->ws(WS_ROW_SCAN, (nn + 1) * sizeof(uint64_t));
    uint8_t* misc = (uint8_t*)ctx->ws(WS_MISC, 4096);
    if (!sb.frames || !sb.boxes || !sb.pvals || !sb.foot || !sb.foot_scan || !sb.rows || !sb.row_scan || !misc)
        return -1;
    sb.bounds = (int32_t*)misc;                                     // 6 ints: AABB bounds
    sb.vbounds = (int32_t*)(misc + 32);                             // 6 ints: voxel bounds
    sb.est = (unsigned long long*)(misc + 64);                      // pair estimate
    unsigned long long* counter = (unsigned long long*)(misc + 72); // emitted pairs
    if (launch_setup(ctx, gs, prm, sb, (prm->flags & VDBR_FLAG_RAW_INPUTS) != 0))
        return -1;
    PrimFrame* frames = sb.frames;
    PrimBox* boxes = sb.boxes;
    PrimVal* pvals = sb.pvals;
    uint64_t* foot = sb.foot;
    uint64_t* foot_scan = sb.foot_scan;
    uint64_t* row_scan = sb.row_scan;
    int32_t* bounds = sb.bounds;
    int32_t* vbounds = sb.vbounds;
    unsigned long long* est = sb.est;
    // batch starts of the footprint scan (window model)
    const uint64_t nbatch = n > 0 ? (n + PB - 1) / PB : 0;
    std::vector<uint64_t> batch_start(nbatch + 1, 0);
    uint64_t* bs_dev = nullptr;
    {
        uint64_t* host = ctx->pinned;
        VDBR_CHECK(ctx, hipMemcpyAsync(host, bounds, 24, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(host + 3, est, 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(host + 4, foot_scan + n, 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(host + 5, row_scan + n, 8, hipMemcpyDeviceToHost, st));
        bs_dev = (uint64_t*)ctx->ws(WS_BATCH, (nbatch + 2) * sizeof(uint64_t) + (nbatch + 2) * sizeof(uint32_t));
        if (!bs_dev)
            return -1;
        if (nbatch > 0)
        {
            hipLaunchKernelGGL(k_gather_u64, dim3(blocks_for(nbatch)), dim3(256), 0, st, foot_scan, (uint64_t)PB, nbatch,
                               bs_dev);
        }
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        std::vector<uint64_t> tmp(nbatch);
        if (nbatch > 0)
            VDBR_CHECK(ctx, hipMemcpy(tmp.data(), bs_dev, nbatch * 8, hipMemcpyDeviceToHost));
        for (uint64_t b = 0; b < nbatch; b++)
            batch_start[b] = tmp[b];
        batch_start[nbatch] = host[4];
    }
    int32_t hb[6];
    memcpy(hb, ctx->pinned, 24);
    const uint64_t est_pairs = ctx->pinned[3];
    const uint64_t frag_total = ctx->pinned[4];
    const uint64_t total_rows = ctx->pinned[5];
    // window bases per prim batch
    std::vector<uint32_t> win_base(nbatch + 1, 0);
    uint64_t wtot = 0;
    for (uint64_t b = 0; b < nbatch; b++)
    {
        win_base[b] = (uint32_t)wtot;
        uint64_t tb = batch_start[b + 1] - batch_start[b];
        wtot += (tb + W - 1) / W;
    }
    if (wtot >= (1ull << 20))
        return ctx->fail("raster_node2: more than 2^20 fragment windows in the node-order model");
    const uint32_t T = (uint32_t)(wtot > 0 ? wtot : 1);
    uint32_t* win_base_dev = (uint32_t*)(bs_dev + (nbatch + 2));
    if (nbatch > 0)
        VDBR_CHECK(ctx, hipMemcpy(win_base_dev, win_base.data(), nbatch * 4, hipMemcpyHostToDevice));
    VDBR_CHECK(ctx, hipEventRecord(ev[1], st));

    // ---------------------------------------------------------------- key layout
    KeyLayout kl;
    memset(&kl, 0, sizeof(kl));
    bool any = frag_total > 0;
    if (any)
    {
        for (int a = 0; a < 3; a++)
        {
            if (hb[a] < -65536 || hb[3 + a] > 65535)
                return ctx->fail("raster_node2: voxel coordinates outside [-65536, 65535] (node2 root range)");
            kl.lmin[a] = hb[a] >> 3;
            int32_t lmax = hb[3 + a] >> 3;
            kl.bits[a] = bits_for((uint64_t)(lmax - kl.lmin[a] + 1));
        }
    }
    kl.gid_bits = bits_for(n > 1 ? n : 2);
    kl.voxel_bits = kl.bits[0] + kl.bits[1] + kl.bits[2] + 9u;
    if (any && leaf_ranks(ctx, st, (const PrimBox*)boxes, n, &kl))
        return -1;
    if (kl.voxel_bits + kl.gid_bits > 64)
        return ctx->fail("raster_node2: pair keys need more than 64 bits");
    // Only the voxel bits are sorted: RGBA8 (the pair alpha rides in the key) keeps the max and
    // the lowest gaussian id per voxel, both independent of the pair order; density sums its
    // terms exactly in fixed point, so the sum does not depend on the pair order either.
    const bool alpha_in_key = prm->value_mode == VDBR_VALUE_RGBA8 && kl.voxel_bits + kl.gid_bits + 8u <= 64u;
    kl.alpha_bits = alpha_in_key ? 8u : 0u;
    const uint32_t key_bits = kl.voxel_bits + kl.gid_bits + kl.alpha_bits;
    // both value modes are reduced per leaf: only the leaf bits are sorted
    const uint32_t sort_begin = kl.gid_bits + kl.alpha_bits + 9u;

    // ---------------------------------------------------------------- slab plan
    // Pairs are emitted, sorted and reduced per slab of leaf-aligned voxel columns x in [i0, i1);
    // x is the most significant voxel-key axis, so each slab's sorted voxels form one contiguous
    // key range and the slabs' voxel lists concatenate into the global sorted list. One slab
    // unless the estimated pairs exceed the budget; a slab that still overflows is split.
    const uint64_t slab_budget = prm->slab_pairs ? (uint64_t)prm->slab_pairs : (1ull << 31);
    const uint64_t pair_limit = std::min<uint64_t>(0xFFFFFFFEull, 2 * slab_budget);
    struct Slab
    {
        int32_t i0, i1;
        uint64_t est;
    };
    std::vector<Slab> todo; // processed from the back
    // the pair estimate is approximate: one slab up to 1.5x the budget (an actual overflow of the
    // hard limit still splits the slab)
    if (!any || est_pairs <= slab_budget + slab_budget / 2)
        todo.push_back(Slab{ INT32_MIN, INT32_MAX, est_pairs });
    else
    {
        const int32_t lmin_x = kl.lmin[0];
        const uint32_t cols = (uint32_t)((hb[3] >> 3) - lmin_x + 1);
        unsigned long long* hist = (unsigned long long*)ctx->ws(WS_SLAB_HIST, (size_t)cols * 8);
        if (!hist)
            return -1;
        VDBR_CHECK(ctx, hipMemsetAsync(hist, 0, (size_t)cols * 8, st));
        hipLaunchKernelGGL(k_slab_hist, dim3(std::min<uint32_t>(blocks_for(n), 1024u)), dim3(256), 0, st,
                           (const PrimBox*)boxes, (const uint64_t*)foot, n, lmin_x, cols, hist);
        VDBR_CHECK(ctx, hipGetLastError());
        std::vector<unsigned long long> h(cols);
        VDBR_CHECK(ctx, hipMemcpyAsync(h.data(), hist, (size_t)cols * 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        double tot = 0.0;
        for (uint32_t c = 0; c < cols; c++)
            tot += (double)h[c];
        const double scale = tot > 0.0 ? (double)est_pairs / tot : 0.0;
        std::vector<Slab> plan;
        uint32_t start = 0;
        double acc = 0.0;
        for (uint32_t c = 0; c < cols; c++)
        {
            const double w = (double)h[c] * scale;
            if (c > start && acc + w > (double)slab_budget)
            {
                plan.push_back(Slab{ (lmin_x + (int32_t)start) * 8, (lmin_x + (int32_t)c) * 8, (uint64_t)acc });
                start = c;
                acc = 0.0;
            }
            acc += w;
        }
        plan.push_back(Slab{ (lmin_x + (int32_t)start) * 8, (lmin_x + (int32_t)cols) * 8, (uint64_t)acc });
        plan.front().i0 = INT32_MIN; // every voxel belongs to exactly one slab
        plan.back().i1 = INT32_MAX;
        todo.assign(plan.rbegin(), plan.rend());
    }

    // ---------------------------------------------------------------- emit + sort + reduce per slab
    const bool rgba_wave = kl.alpha_bits != 0u;
    const bool density = prm->value_mode == VDBR_VALUE_DENSITY;
    // row-interval constants of every Gaussian (the row pass of every slab reads them)
    RowFrame* rframes = nullptr;
    if (any && n > 0)
    {
        rframes = (RowFrame*)ctx->ws(WS_ROWFRAMES, n * sizeof(RowFrame));
        if (!rframes)
            return -1;
        hipLaunchKernelGGL(k_row_frames, dim3(blocks_for(n)), dim3(256), 0, st, n, (const PrimFrame*)frames,
                           (const PrimBox*)boxes, vs, rframes);
    }
    DensPrim* dprims = nullptr;
    if (density && n > 0)
    {
        dprims = (DensPrim*)ctx->ws(WS_DENS_PRIMS, n * sizeof(DensPrim));
        if (!dprims)
            return -1;
        hipLaunchKernelGGL(k_dens_prims, dim3(blocks_for(n)), dim3(256), 0, st, n, gs->means, gs->quats, gs->scales,
                           gs->opacities, vs, dprims);
    }
    uint64_t pair_count = 0; // all slabs
    uint64_t nvox = 0;       // all slabs
    uint64_t vcap = 0;       // capacity of the voxel arrays
    uint64_t* vox_key = nullptr;
    uint32_t* vox_val = nullptr;
    uint32_t* vox_t = nullptr;
    uint32_t* vox_g0 = nullptr;
    uint32_t attempts = 0;
    uint32_t slab_count = 0;
    float acc_emit = 0.f, acc_sort = 0.f, acc_reduce = 0.f, acc_passes = 0.f, acc_emit_k = 0.f, acc_reduce_k = 0.f;
    float acc_row_k = 0.f, acc_walk_k = 0.f;
    uint64_t acc_cand = 0;
    uint64_t* last_sorted = nullptr;
    uint64_t last_slab_pairs = 0;
    // grows one voxel array, keeping its first `keep` bytes
    auto grow_keep = [&](WsSlot slot, size_t keep, size_t bytes) -> void* {
        if (ctx->ws_size[slot] >= bytes && ctx->ws_ptr[slot])
            return ctx->ws_ptr[slot];
        void* tmp = nullptr;
        if (keep)
        {
            tmp = ctx->ws(WS_VOX_GROW, keep);
            if (!tmp || hipMemcpyAsync(tmp, ctx->ws_ptr[slot], keep, hipMemcpyDeviceToDevice, st) != hipSuccess)
                return nullptr;
        }
        void* p = ctx->ws(slot, bytes);
        if (p && keep && hipMemcpyAsync(p, tmp, keep, hipMemcpyDeviceToDevice, st) != hipSuccess)
            return nullptr;
        return p;
    };
    while (any && !todo.empty())
    {
        const Slab sl = todo.back();
        todo.pop_back();
        const bool whole = sl.i0 == INT32_MIN && sl.i1 == INT32_MAX;
        VDBR_CHECK(ctx, hipEventRecord(ev[12], st));
        // rows of this slab
        const uint64_t* srow_scan = row_scan;
        uint64_t srows = total_rows;
        if (!whole)
        {
            uint64_t* rows_s = (uint64_t*)ctx->ws(WS_SLAB_ROWS, 2 * (nn + 1) * sizeof(uint64_t));
            if (!rows_s)
                return -1;
            uint64_t* scan_s = rows_s + (nn + 1);
            hipLaunchKernelGGL(k_slab_rows, dim3(std::min<uint32_t>(blocks_for(n), 4096u)), dim3(256), 0, st,
                               (const PrimBox*)boxes, n, sl.i0, sl.i1, rows_s);
            VDBR_CHECK(ctx, hipGetLastError());
            if (scan_u64(ctx, rows_s, scan_s, n, false, scan_s + n))
                return -1;
            VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, scan_s + n, 8, hipMemcpyDeviceToHost, st));
            VDBR_CHECK(ctx, hipStreamSynchronize(st));
            srow_scan = scan_s;
            srows = ctx->pinned[0];
        }
        const uint64_t sest = whole ? est_pairs : sl.est;
        // a slab can be split when it spans more than one leaf column of the voxel bounds
        const int64_t span_lo = whole ? (int64_t)hb[0] : std::max<int64_t>((int64_t)sl.i0, (int64_t)hb[0]);
        const int64_t span_hi = whole ? (int64_t)hb[3] + 1 : std::min<int64_t>((int64_t)sl.i1, (int64_t)hb[3] + 1);
        const int64_t split_at = ((span_lo + span_hi) / 2) & ~7ll;
        const bool splittable = split_at > span_lo && split_at < span_hi;
        uint64_t spairs = 0;
        uint64_t* pairs = nullptr;
        unsigned long long* hist_buf = nullptr;
        uint32_t hist_w = 8;
        void* segs = nullptr;
        const bool seg_compact = seg_compact_fits(kl) && !(prm->flags & VDBR_FLAG_WIDE_SEGMENTS);
        const size_t seg_rec_bytes = seg_compact ? sizeof(SegCompact::Rec) : sizeof(SegWide::Rec);
        uint64_t seg_count = 0;
        if (srows > 0)
        {
            // candidate segments of the pruned row intervals
            uint64_t blocks_rows = (srows + EMIT_ROWS - 1) / EMIT_ROWS;
            if (blocks_rows > 0x7FFFFFFFull)
                return ctx->fail("raster_node2: too many AABB rows for one launch");
            uint64_t seg_total = srows + (2 * sest) / EMIT_SEG + 1024;
            if (seg_total > srows + frag_total / EMIT_SEG + 1024)
                seg_total = srows + frag_total / EMIT_SEG + 1024;
            // partitions receive blocks round robin; a quarter of slack absorbs their imbalance
            uint64_t part_cap = seg_total / SEG_PARTS + seg_total / (4 * SEG_PARTS) + 4096;
            uint64_t* seg_meta =
                (uint64_t*)ctx->ws(WS_SEG_META, (SEG_PARTS * SEG_COUNTER_STRIDE + 2 * (SEG_PARTS + 1)) * 8);
            if (!seg_meta)
                return -1;
            unsigned long long* seg_counters = (unsigned long long*)seg_meta;
            uint64_t* part_count_dev = seg_meta + SEG_PARTS * SEG_COUNTER_STRIDE;
            uint64_t* chunk_scan_dev = part_count_dev + (SEG_PARTS + 1);
            VDBR_CHECK(ctx, hipEventRecord(ev[8], st));
            for (uint32_t sa = 0;; sa++)
            {
                segs = ctx->ws(WS_SEGS, SEG_PARTS * part_cap * seg_rec_bytes);
                if (!segs)
                    return -1;
                VDBR_CHECK(ctx, hipMemsetAsync(seg_counters, 0, SEG_PARTS * SEG_COUNTER_STRIDE * 8, st));
                EmitArgs ra;
                memset(&ra, 0, sizeof(ra));
                ra.frames = frames;
                ra.boxes = boxes;
                ra.row_scan = srow_scan;
                ra.n = n;
                ra.total_rows = srows;
                ra.vs = vs;
                ra.slab_i0 = whole ? INT32_MIN : sl.i0;
                ra.rframes = rframes;
                ra.kl = kl;
                ra.seg_compact = seg_compact ? 1u : 0u;
                hipLaunchKernelGGL(k_row_segments, dim3((uint32_t)blocks_rows), dim3(EMIT_THREADS), 0, st, ra, segs,
                                   part_cap, seg_counters);
                VDBR_CHECK(ctx, hipGetLastError());
                VDBR_CHECK(ctx, hipEventRecord(ev[3], st));
                VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, seg_counters, SEG_PARTS * SEG_COUNTER_STRIDE * 8,
                                               hipMemcpyDeviceToHost, st));
                VDBR_CHECK(ctx, hipStreamSynchronize(st));
                uint64_t maxc = 0;
                seg_count = 0;
                for (uint32_t q = 0; q < SEG_PARTS; q++)
                {
                    uint64_t cq = ctx->pinned[q * SEG_COUNTER_STRIDE];
                    maxc = cq > maxc ? cq : maxc;
                    seg_count += cq;
                }
                if (maxc <= part_cap)
                {
                    for (uint32_t q = 0; q < SEG_PARTS; q++)
                        acc_cand += ctx->pinned[q * SEG_COUNTER_STRIDE + 1];
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, ev[8], ev[3]) == hipSuccess)
                        acc_row_k += ms;
                    break;
                }
                if (sa >= 1)
                    return ctx->fail("raster_node2: segment buffer overflow after resize");
                part_cap = maxc;
            }
            {
                uint64_t host[2 * (SEG_PARTS + 1)];
                uint64_t acc = 0;
                for (uint32_t q = 0; q < SEG_PARTS; q++)
                {
                    uint64_t cq = ctx->pinned[q * SEG_COUNTER_STRIDE];
                    host[q] = cq;
                    host[SEG_PARTS + 1 + q] = acc;
                    acc += (cq + 63) / 64;
                }
                host[SEG_PARTS] = 0;
                host[2 * SEG_PARTS + 1] = acc;
                VDBR_CHECK(ctx, hipMemcpy(part_count_dev, host, sizeof(host), hipMemcpyHostToDevice));
            }
            uint64_t cap = sest + sest / 4 + 1024;
            if (cap > frag_total)
                cap = frag_total;
            if (cap > pair_limit + 1)
                cap = pair_limit + 1;
            // the emission counts the sort's digit histograms (for the digit width of the
            // estimated pair count; the sort recounts when the final count picks another width)
            const uint32_t hbits = key_bits - sort_begin;
            hist_w = sort_digit_bits(ctx, sest > 1 ? sest : 2, hbits);
            hist_buf = emit_counts_hist(hbits, hist_w) ? sort_hist_buffer(ctx) : nullptr;
            uint32_t sattempts = 0;
            for (;;)
            {
                sattempts++;
                pairs = (uint64_t*)ctx->ws(WS_PAIRS, cap * sizeof(uint64_t));
                if (!pairs)
                    return -1;
                VDBR_CHECK(ctx, hipMemsetAsync(counter, 0, 8, st));
                if (hist_buf)
                    VDBR_CHECK(ctx, hipMemsetAsync(hist_buf, 0, EMIT_HIST_PASSES * SORT_BINS * sizeof(uint64_t), st));
                EmitArgs ea;
                memset(&ea, 0, sizeof(ea));
                ea.frames = frames;
                ea.boxes = boxes;
                ea.vals = pvals;
                ea.row_scan = srow_scan;
                ea.n = n;
                ea.total_rows = srows;
                ea.vs = vs;
                ea.kl = kl;
                ea.out = pairs;
                ea.cap = cap;
                ea.counter = counter;
                ea.hist = hist_buf;
                ea.hist_shift = sort_begin;
                ea.hist_bits = key_bits - sort_begin;
                ea.hist_w = hist_w;
                ea.segs = segs;
                ea.seg_compact = seg_compact ? 1u : 0u;
                ea.seg_part_cap = part_cap;
                ea.seg_part_count = part_count_dev;
                ea.seg_chunk_scan = chunk_scan_dev;
                ea.slab_i0 = whole ? INT32_MIN : sl.i0;
                // persistent waves: each walks many chunks, so its LDS stage flushes only when full
                // (one reservation per ~EMIT_WAVE_STAGE pairs)
                uint64_t chunks = (seg_count + 63) / 64 + SEG_PARTS;
                uint64_t blocks = (chunks + EMIT_THREADS / 64 - 1) / (EMIT_THREADS / 64);
                if (blocks > ctx->emit_blocks)
                    blocks = ctx->emit_blocks;
                if (blocks == 0)
                    blocks = 1;
                VDBR_CHECK(ctx, hipEventRecord(ev[2], st));
                launch_emit_segments(ea, (uint32_t)blocks, st);
                VDBR_CHECK(ctx, hipGetLastError());
                VDBR_CHECK(ctx, hipEventRecord(ev[9], st));
                VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, counter, 8, hipMemcpyDeviceToHost, st));
                VDBR_CHECK(ctx, hipStreamSynchronize(st));
                spairs = ctx->pinned[0];
                if (spairs <= cap)
                    break;
                if (spairs > pair_limit && splittable)
                    break; // the slab is split below
                if (spairs > 0xFFFFFFFEull)
                    return ctx->fail("raster_node2: more than 2^32-2 pairs in one leaf column");
                if (sattempts >= 2)
                    return ctx->fail("raster_node2: pair buffer overflow after resize");
                cap = spairs;
            }
            attempts = std::max(attempts, sattempts);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[8], ev[9]) == hipSuccess)
                acc_emit_k += ms;
            if (hipEventElapsedTime(&ms, ev[2], ev[9]) == hipSuccess)
                acc_walk_k += ms;
        }
        if (spairs > pair_limit && splittable)
        {
            // too many pairs for one slab: split at a leaf column boundary and redo both halves
            todo.push_back(Slab{ (int32_t)split_at, sl.i1, spairs / 2 + 1 });
            todo.push_back(Slab{ sl.i0, (int32_t)split_at, spairs / 2 + 1 });
            continue;
        }
        slab_count++;
        pair_count += spairs;
        VDBR_CHECK(ctx, hipEventRecord(ev[13], st));

        // sort
        uint64_t* sorted = pairs;
        uint32_t passes = 0;
        if (spairs > 1)
        {
            const bool hist_ready = hist_buf && sort_digit_bits(ctx, spairs, key_bits - sort_begin) == hist_w;
            // the reduces do not depend on the order of the pairs inside a leaf
            if (radix_sort_u64_noswap(ctx, pairs, host_count(spairs), sort_begin, key_bits, &sorted, ev[6], ev[7], &passes,
                                      hist_ready, true))
                return -1;
            sort_passes = passes;
        }
        VDBR_CHECK(ctx, hipEventRecord(ev[14], st));
        last_sorted = sorted;
        last_slab_pairs = spairs;

        // reduce per leaf, voxels appended after those of the previous slabs
        LeafReduce lr;
        lr.sorted = sorted;
        lr.pairs = host_count(spairs);
        lr.kl = kl;
        lr.vs = vs;
        lr.rgba_wave = rgba_wave;
        lr.density = density;
        lr.frames = frames;
        lr.pvals = pvals;
        lr.dprims = dprims;
        lr.ev_begin = ev[10];
        lr.ev_end = ev[11];
        if (reduce_leaves(ctx, &lr))
            return -1;
        const uint64_t svox = lr.svox.cap;
        const uint64_t sleaves = lr.leaves.cap;
        if (nvox + svox > vcap || !vox_key)
        {
            const uint64_t want = whole ? std::max<uint64_t>(svox, 1) : std::max<uint64_t>(nvox + svox, vcap + vcap / 2);
            vox_key = (uint64_t*)grow_keep(WS_VOX_KEY, nvox * sizeof(uint64_t), want * sizeof(uint64_t));
            vox_val = (uint32_t*)grow_keep(WS_VOX_VAL, nvox * sizeof(uint32_t), want * sizeof(uint32_t));
            vox_t = (uint32_t*)grow_keep(WS_VOX_T, nvox * sizeof(uint32_t), want * sizeof(uint32_t));
            vox_g0 = (uint32_t*)grow_keep(WS_VOX_G0, nvox * sizeof(uint32_t), want * sizeof(uint32_t));
            if (!vox_key || !vox_val || !vox_t || !vox_g0)
                return -1;
            vcap = want;
        }
        if (svox > 0)
        {
            VoxelLists vl{ vox_key + nvox, vox_val + nvox, vox_t + nvox, vox_g0 + nvox };
            if (write_voxels(ctx, lr, vl, boxes, foot_scan, win_base_dev, PB, W, vbounds))
                return -1;
        }
        VDBR_CHECK(ctx, hipEventRecord(ev[15], st));
        VDBR_CHECK(ctx, hipEventSynchronize(ev[15]));
        nvox += svox;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[12], ev[13]) == hipSuccess)
            acc_emit += ms;
        if (hipEventElapsedTime(&ms, ev[13], ev[14]) == hipSuccess)
            acc_sort += ms;
        if (hipEventElapsedTime(&ms, ev[14], ev[15]) == hipSuccess)
            acc_reduce += ms;
        if (passes && hipEventElapsedTime(&ms, ev[6], ev[7]) == hipSuccess)
            acc_passes += ms;
        if (sleaves && hipEventElapsedTime(&ms, ev[10], ev[11]) == hipSuccess)
            acc_reduce_k += ms;
    }
    if (nvox >= 0xFFFFFFFFull)
        return ctx->fail("raster_node2: more than 2^32-1 active voxels");
    const uint64_t nv1 = nvox > 0 ? nvox : 1;
    if (!vox_key)
    {
        vox_key = (uint64_t*)ctx->ws(WS_VOX_KEY, nv1 * sizeof(uint64_t));
        vox_val = (uint32_t*)ctx->ws(WS_VOX_VAL, nv1 * sizeof(uint32_t));
        vox_t = (uint32_t*)ctx->ws(WS_VOX_T, nv1 * sizeof(uint32_t));
        if (!vox_key || !vox_val || !vox_t)
            return -1;
    }
    // the pair list of a single-slab build stays available (vdbr_last_pairs)
    ctx->last_pairs = slab_count == 1 ? last_sorted : nullptr;
    ctx->last_pair_count = slab_count == 1 ? last_slab_pairs : 0;
    if (slab_count == 1 && density)
    {
        // the density adjoint may reuse this list (vdbr_density_backward_pairs)
        ctx->last_pair_id = ++ctx->pair_gen;
        ctx->last_pair_n = n;
        ctx->last_pair_vs = vs;
    }
    ctx->last_kl = kl;
    VDBR_CHECK(ctx, hipEventRecord(ev[4], st));
    Node2BuildInput bin;
    bin.kl = kl;
    bin.vox_key = vox_key;
    bin.vox_val = vox_val;
    bin.vox_t = vox_t;
    bin.nvox = nvox;
    bin.vbounds = vbounds;
    bin.T = T;
    bin.max_nodes = max_nodes;
    bin.spatial_order = (prm->flags & VDBR_FLAG_SPATIAL_ORDER) != 0;
    bin.voxel_coords = (prm->flags & VDBR_FLAG_VOXEL_COORDS) != 0;
    bin.value_type = prm->value_mode == VDBR_VALUE_RGBA8 ? VDB_GRID_TYPE_RGBA8 : VDB_GRID_TYPE_FLOAT;
    bin.voxel_size = vs;
    Node2BuildOutput bout;
    if (build_node2_grid(ctx, bin, &bout))
        return -1;
    VDBR_CHECK(ctx, hipEventRecord(ev[5], st));
    VDBR_CHECK(ctx, hipEventSynchronize(ev[5]));
    uint8_t* grid = bout.grid;
    GridInfo gi;
    gi.grid_size = bout.grid_size;
    const uint64_t nupper = bout.upper_count, nlower = bout.lower_count, nleaf = bout.leaf_count;
    *grid_dev = grid;
    *grid_size_out = gi.grid_size;
    if (stats)
    {
        memset(stats, 0, sizeof(*stats));
        stats->gaussian_count = n;
        stats->fragment_count = frag_total;
        stats->pair_count = pair_count;
        stats->voxel_count = nvox;
        stats->grid_size = gi.grid_size;
        stats->upper_count = (uint32_t)nupper;
        stats->lower_count = (uint32_t)nlower;
        stats->leaf_count = (uint32_t)nleaf;
        stats->window_count = (uint32_t)wtot;
        stats->key_bits = key_bits;
        stats->sort_bit_begin = sort_begin;
        stats->emit_attempts = attempts;
        stats->slab_count = slab_count;
        stats->pair_list_id = ctx->last_pair_id;
        float ms;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        stats->ms_setup = ms;
        (void)hipEventElapsedTime(&ms, ev[4], ev[5]);
        stats->ms_build = ms;
        (void)hipEventElapsedTime(&ms, ev[0], ev[5]);
        stats->ms_total = ms;
        // per-stage sums over the slabs
        stats->ms_emit = acc_emit;
        stats->ms_sort = acc_sort;
        stats->ms_reduce = acc_reduce;
        stats->sort_passes = sort_passes;
        stats->ms_sort_passes = acc_passes;
        stats->ms_emit_kernel = acc_emit_k;
        stats->ms_row_kernel = acc_row_k;
        stats->ms_walk_kernel = acc_walk_k;
        stats->candidate_voxels = acc_cand;
        stats->ms_reduce_kernel = acc_reduce_k;
        (void)hipGetLastError();
    }
    return 0;
}

int preprocess_device(Context* ctx, uint64_t n, float* opacities, float* quats, float* scales, const float* sh0,
                      float* colors)
{
    if (n == 0)
        return 0;
    hipLaunchKernelGGL(k_preprocess, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, n, opacities, quats, scales, sh0,
                       colors);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int last_pairs(Context* ctx, int32_t* ijk_dev, uint32_t* gid_dev, uint64_t cap, uint64_t* count)
{
    if (count)
        *count = ctx->last_pair_count;
    if (!ijk_dev && !gid_dev)
        return 0;
    uint64_t n = ctx->last_pair_count < cap ? ctx->last_pair_count : cap;
    if (n == 0)
        return 0;
    const KeyLayout kl = ctx->last_kl;
    hipLaunchKernelGGL(k_pairs_decode, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, ctx->last_pairs, n, kl, ijk_dev,
                       gid_dev);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}


int last_voxels(Context* ctx, int32_t* ijk_dev, uint64_t cap, uint64_t* count)
{
    if (count)
        *count = ctx->last_value_count;
    if (!ijk_dev || !ctx->last_voxel_ijk)
        return 0;
    const uint64_t n = ctx->last_value_count < cap ? ctx->last_value_count : cap;
    if (n)
        VDBR_CHECK(ctx, hipMemcpyAsync(ijk_dev, ctx->last_voxel_ijk, 3 * n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                       ctx->stream));
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// value words of a node2 grid at voxel coordinates (0 where the voxel is inactive or outside)
__global__ void k_node2_sample(const uint64_t* __restrict__ w, uint64_t nwords, const uint32_t* __restrict__ values,
                               uint64_t value_count, const int32_t* __restrict__ ijk, uint64_t n,
                               uint32_t* __restrict__ out, uint32_t* __restrict__ bad_flag)
{
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (uint64_t)gridDim.x * blockDim.x)
    {
        const int32_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        uint32_t v = 0u;
        const bool in_range = i >= -65536 && i < 65536 && j >= -65536 && j < 65536 && k >= -65536 && k < 65536;
        bool bad = false;
        const uint64_t leaf = in_range ? n2_find_leaf(w, nwords, i, j, k, &bad) : 0ull;
        if (bad)
            atomicOr(bad_flag, 1u);
        if (leaf)
        {
            const uint64_t idx = n2_leaf_value_index(w, leaf, i, j, k);
            if (idx < value_count)
                v = values[idx];
        }
        out[q] = v;
    }
}

int node2_sample(Context* ctx, const void* grid, uint64_t grid_size, const uint32_t* values, uint64_t value_count,
                 const int32_t* ijk, uint64_t n, uint32_t* out)
{
    if (!grid || !values || (n && (!ijk || !out)))
        return ctx->fail("node2_sample: null argument");
    if (grid_size < VDB_ROOT_ADDR + VDB_SIZE_ROOT)
        return ctx->fail("node2_sample: grid is not a node2 grid");
    if (n == 0)
        return 0;
    uint32_t* bad_flag = (uint32_t*)ctx->ws(WS_MISC, 4096);
    if (!bad_flag)
        return -1;
    VDBR_CHECK(ctx, hipMemsetAsync(bad_flag, 0, 4, ctx->stream));
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_node2_sample, dim3(blocks), dim3(256), 0, ctx->stream, (const uint64_t*)grid, grid_size / 8u,
                       values, value_count, ijk, n, out, bad_flag);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, bad_flag, 4, hipMemcpyDeviceToHost, ctx->stream));
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (*(const uint32_t*)ctx->pinned)
        return ctx->fail("node2_sample: child offsets outside the grid (corrupt or truncated node2 grid)");
    return 0;
}

// leaf table of a node2 grid with dL/drho per voxel slot; the node counts of the tree header bound
// the node lists (checked against the grid size here, against the nodes found by the caller
// through lt->counters[3] after the stream has run)
static int bwd_leaf_table(Context* ctx, const void* grid, uint64_t grid_size, const float* grad_values,
                          uint64_t value_count, LeafTable* lt)
{
    hipStream_t st = ctx->stream;
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, (const uint8_t*)grid + VDB_TREE_ADDR + VDB_TREE_OFF_COUNT_LEAF, 12,
                                   hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint32_t* tree_counts = (const uint32_t*)ctx->pinned;
    const uint64_t n_leaf = tree_counts[0], n_lower = tree_counts[1], n_upper = tree_counts[2];
    if (n_upper > 32768u || n_lower > n_upper * 32768u || n_leaf > n_lower * 4096u ||
        n_upper * VDB_SIZE_UPPER + n_lower * VDB_SIZE_LOWER + n_leaf * VDB_SIZE_LEAF > grid_size)
        return ctx->fail("density_backward: grid tree node counts are inconsistent with its size");
    uint64_t lt_cap = 2;
    while (lt_cap < 2 * n_leaf)
        lt_cap <<= 1;
    if (lt_cap > (1ull << 31))
        return ctx->fail("density_backward: more than 2^30 leaves");
    uint8_t* lt_ws = (uint8_t*)ctx->ws(WS_BWD_LEAVES, 64 + (n_upper + n_lower) * sizeof(NodeRef) +
                                                          n_leaf * sizeof(uint64_t) + lt_cap * sizeof(uint4));
    float* leaf_grad = (float*)ctx->ws(WS_BWD_LEAFGRAD, (n_leaf * 512 + 1) * sizeof(float));
    if (!lt_ws || !leaf_grad)
        return -1;
    LeafTable& t = *lt;
    t.grid = (const uint64_t*)grid;
    t.words = grid_size >> 3;
    t.counters = (unsigned int*)lt_ws;
    t.slots = (uint4*)(lt_ws + 64);
    t.uppers = (NodeRef*)(lt_ws + 64 + lt_cap * sizeof(uint4));
    t.lowers = t.uppers + n_upper;
    t.leaves = (uint64_t*)(t.lowers + n_lower);
    t.mask = (uint32_t)(lt_cap - 1);
    t.cap[0] = (uint32_t)n_upper;
    t.cap[1] = (uint32_t)n_lower;
    t.cap[2] = (uint32_t)n_leaf;
    t.grad_values = grad_values;
    t.value_count = value_count;
    t.leaf_grad = leaf_grad;
    VDBR_CHECK(ctx, hipMemsetAsync(lt_ws, 0, 64 + lt_cap * sizeof(uint4), st));
    hipLaunchKernelGGL(k_lt_root, dim3(32768 / 256), dim3(256), 0, st, t);
    if (n_upper)
        hipLaunchKernelGGL(k_lt_upper, dim3(blocks_for(n_upper << 15)), dim3(256), 0, st, t);
    if (n_lower)
        hipLaunchKernelGGL(k_lt_lower, dim3(blocks_for(n_lower << 12)), dim3(256), 0, st, t);
    if (n_leaf)
        hipLaunchKernelGGL(k_lt_grad, dim3(blocks_for(n_leaf << 9)), dim3(256), 0, st, t);
    VDBR_CHECK(ctx, hipGetLastError());
    return 0;
}

// argument checks shared by both adjoints; *empty when there is nothing to compute
static int bwd_check_args(Context* ctx, const vdbr_gaussians_t* raw, float vs, const void* grid, uint64_t grid_size,
                          const float* grad_values, const vdbr_gaussian_grads_t* grads, bool* empty)
{
    *empty = false;
    if (!raw || !grads || !grid)
        return ctx->fail("density_backward: null argument");
    const uint64_t n = raw->count;
    if (!(vs > 0.f) || !(vs < 3.0e38f))
        return ctx->fail("density_backward: voxel_size must be a positive finite number");
    if (n >= (1ull << 32))
        return ctx->fail("density_backward: more than 2^32-1 Gaussians");
    if (grid_size < VDB_ROOT_ADDR + VDB_SIZE_ROOT)
        return ctx->fail("density_backward: grid is not a node2 grid");
    if (n == 0)
    {
        *empty = true;
        return 0;
    }
    if (!raw->means || !raw->opacities || !raw->quats || !raw->scales || !grad_values || !grads->means ||
        !grads->opacities || !grads->quats || !grads->scales)
        return ctx->fail("density_backward: null array");
    return 0;
}

int density_backward(Context* ctx, const vdbr_gaussians_t* raw, float vs, const void* grid, uint64_t grid_size,
                     const float* grad_values, uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                     float* ms_out)
{
    bool empty = false;
    if (bwd_check_args(ctx, raw, vs, grid, grid_size, grad_values, grads, &empty))
        return -1;
    if (empty)
        return 0;
    const uint64_t n = raw->count;
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1;
    VDBR_CHECK(ctx, hipEventCreate(&e0));
    VDBR_CHECK(ctx, hipEventCreate(&e1));
    struct EvGuard
    {
        hipEvent_t a, b;
        ~EvGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{ e0, e1 };
    VDBR_CHECK(ctx, hipEventRecord(e0, st));
    // preprocessed copies of the parameters (the forward's topology is recomputed from them)
    float* pre = (float*)ctx->ws(WS_RAW_COPY, n * (4 + 3 + 1 + 3) * sizeof(float));
    if (!pre)
        return -1;
    float* q_pre = pre;
    float* s_pre = q_pre + 4 * n;
    float* o_pre = s_pre + 3 * n;
    float* c_pre = o_pre + n;
    VDBR_CHECK(ctx, hipMemcpyAsync(q_pre, raw->quats, 4 * n * 4, hipMemcpyDeviceToDevice, st));
    VDBR_CHECK(ctx, hipMemcpyAsync(s_pre, raw->scales, 3 * n * 4, hipMemcpyDeviceToDevice, st));
    VDBR_CHECK(ctx, hipMemcpyAsync(o_pre, raw->opacities, n * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_preprocess, dim3(blocks_for(n)), dim3(256), 0, st, n, o_pre, q_pre, s_pre, raw->means, c_pre);
    PrimFrame* frames = (PrimFrame*)ctx->ws(WS_FRAMES, n * sizeof(PrimFrame));
    PrimBox* boxes = (PrimBox*)ctx->ws(WS_BOXES, n * sizeof(PrimBox));
    PrimVal* pvals = (PrimVal*)ctx->ws(WS_VALS, n * sizeof(PrimVal));
    uint64_t* foot = (uint64_t*)ctx->ws(WS_FOOT, n * sizeof(uint64_t));
    uint64_t* rows = (uint64_t*)ctx->ws(WS_ROWS, n * sizeof(uint64_t));
    uint64_t* row_scan = (uint64_t*)ctx->ws(WS_ROW_SCAN, (n + 1) * sizeof(uint64_t));
    uint8_t* misc = (uint8_t*)ctx->ws(WS_MISC, 4096);
    if (!frames || !boxes || !pvals || !foot || !rows || !row_scan || !misc)
        return -1;
    int32_t* bounds = (int32_t*)misc;
    unsigned long long* est = (unsigned long long*)(misc + 64);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, bounds);
    VDBR_CHECK(ctx, hipMemsetAsync(est, 0, 8, st));
    SetupOut so{ frames, boxes, pvals, foot, rows, bounds, est };
    hipLaunchKernelGGL(k_setup, dim3(std::min(blocks_for(n), SETUP_BLOCKS)), dim3(256), 0, st, n, raw->means, q_pre,
                       s_pre, o_pre, c_pre, vs,
                       1.f / vs, 0.f, so);
    if (scan_u64(ctx, rows, row_scan, n, false, row_scan + n))
        return -1;
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, row_scan + n, 8, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint64_t total_rows = ctx->pinned[0];
    LeafTable lt;
    if (bwd_leaf_table(ctx, grid, grid_size, grad_values, value_count, &lt))
        return -1;
    BwdPrim* prims = (BwdPrim*)ctx->ws(WS_BWD_PRIMS, n * sizeof(BwdPrim));
    if (!prims)
        return -1;
    const uint64_t chunks = (total_rows + 63) / 64;
    double* acc = (double*)ctx->ws(WS_BWD_ACC, n * BWD_ACC * sizeof(double));
    double* rec = (double*)ctx->ws(WS_BWD_REC, (chunks + 1) * 2 * BWD_ACC * sizeof(double));
    uint64_t* rec_gid = (uint64_t*)ctx->ws(WS_BWD_RECGID, (chunks + 1) * 2 * sizeof(uint64_t));
    if (!acc || !rec || !rec_gid)
        return -1;
    VDBR_CHECK(ctx, hipMemsetAsync(acc, 0, n * BWD_ACC * sizeof(double), st));
    BwdArgs a;
    a.frames = frames;
    a.boxes = boxes;
    a.row_scan = row_scan;
    a.n = n;
    a.total_rows = total_rows;
    a.vs = vs;
    a.means = raw->means;
    a.log_scales = raw->scales;
    a.quats = raw->quats;
    a.logits = raw->opacities;
    a.prims = prims;
    a.lt_slots = lt.slots;
    a.lt_mask = lt.mask;
    a.leaf_grad = lt.leaf_grad;
    a.acc = acc;
    a.rec = rec;
    a.rec_gid = rec_gid;
    a.grad_means = grads->means;
    a.grad_logits = grads->opacities;
    a.grad_quats = grads->quats;
    a.grad_log_scales = grads->scales;
    hipLaunchKernelGGL(k_bwd_prims, dim3(blocks_for(n)), dim3(256), 0, st, a);
    if (total_rows > 0)
    {
        uint64_t blocks = (total_rows + EMIT_ROWS - 1) / EMIT_ROWS;
        if (blocks > 0x7FFFFFFFull)
            return ctx->fail("density_backward: too many AABB rows for one launch");
        hipLaunchKernelGGL(k_density_bwd, dim3((uint32_t)blocks), dim3(EMIT_THREADS), 0, st, a);
    }
    hipLaunchKernelGGL(k_density_bwd_finalize, dim3(blocks_for(n)), dim3(256), 0, st, a);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipEventRecord(e1, st));
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, lt.counters, 16, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    if (((const uint32_t*)ctx->pinned)[3])
        return ctx->fail("density_backward: the grid's nodes do not match its tree header");
    if (ms_out)
        VDBR_CHECK(ctx, hipEventElapsedTime(ms_out, e0, e1));
    return 0;
}

int density_backward_pairs(Context* ctx, const vdbr_gaussians_t* raw, float vs, const void* grid, uint64_t grid_size,
                           const float* grad_values, uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                           uint64_t pair_list_id, float* ms_out, int* used_pairs)
{
    *used_pairs = 0;
    const bool have = pair_list_id != 0 && pair_list_id == ctx->last_pair_id && raw && raw->count == ctx->last_pair_n &&
                      vs == ctx->last_pair_vs && ctx->last_kl.alpha_bits == 0u;
    if (!have)
        return density_backward(ctx, raw, vs, grid, grid_size, grad_values, value_count, grads, ms_out);
    bool empty = false;
    if (bwd_check_args(ctx, raw, vs, grid, grid_size, grad_values, grads, &empty))
        return -1;
    if (empty)
    {
        *used_pairs = 1;
        return 0;
    }
    const uint64_t n = raw->count;
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1;
    VDBR_CHECK(ctx, hipEventCreate(&e0));
    VDBR_CHECK(ctx, hipEventCreate(&e1));
    struct EvGuard
    {
        hipEvent_t a, b;
        ~EvGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{ e0, e1 };
    VDBR_CHECK(ctx, hipEventRecord(e0, st));
    // max |dL/drho| sets the fixed-point scale; a non-finite gradient takes the re-walk, which
    // carries it into the results
    uint8_t* misc = (uint8_t*)ctx->ws(WS_MISC, 4096);
    if (!misc)
        return -1;
    unsigned int* rmax = (unsigned int*)(misc + 256);
    VDBR_CHECK(ctx, hipMemsetAsync(rmax, 0, 4, st));
    if (value_count)
        hipLaunchKernelGGL(k_abs_max_bits, dim3(std::min<uint32_t>(blocks_for(value_count), 1024u)), dim3(256), 0, st,
                           grad_values, value_count, rmax);
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, rmax, 4, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    if (*(const uint32_t*)ctx->pinned >= 0x7F800000u)
        return density_backward(ctx, raw, vs, grid, grid_size, grad_values, value_count, grads, ms_out);
    LeafTable lt;
    if (bwd_leaf_table(ctx, grid, grid_size, grad_values, value_count, &lt))
        return -1;
    BwdPrim* prims = (BwdPrim*)ctx->ws(WS_BWD_PRIMS, n * sizeof(BwdPrim));
    unsigned long long* sums = (unsigned long long*)ctx->ws(WS_BWD_ACC, n * BWD_ACC * 2 * sizeof(uint64_t));
    if (!prims || !sums)
        return -1;
    VDBR_CHECK(ctx, hipMemsetAsync(sums, 0, n * BWD_ACC * 2 * sizeof(uint64_t), st));
    BwdArgs a;
    memset(&a, 0, sizeof(a));
    a.n = n;
    a.vs = vs;
    a.means = raw->means;
    a.log_scales = raw->scales;
    a.quats = raw->quats;
    a.logits = raw->opacities;
    a.prims = prims;
    a.grad_means = grads->means;
    a.grad_logits = grads->opacities;
    a.grad_quats = grads->quats;
    a.grad_log_scales = grads->scales;
    hipLaunchKernelGGL(k_bwd_prims, dim3(blocks_for(n)), dim3(256), 0, st, a);
    PairBwdArgs pa;
    memset(&pa, 0, sizeof(pa));
    pa.keys = ctx->last_pairs;
    pa.npairs = ctx->last_pair_count;
    pa.kl = ctx->last_kl; // density pair keys: no alpha bits
    pa.vs = (double)vs;
    pa.prims = prims;
    pa.lt_slots = lt.slots;
    pa.lt_mask = lt.mask;
    pa.leaf_grad = lt.leaf_grad;
    pa.rmax_bits = rmax;
    pa.sums = sums;
    const uint64_t blocks = (pa.npairs + BP_TILE - 1) / BP_TILE;
    if (blocks > 0x7FFFFFFFull)
        return ctx->fail("density_backward: too many pairs for one launch");
    if (blocks && pa.keys)
        hipLaunchKernelGGL(k_density_bwd_pairs, dim3((uint32_t)blocks), dim3(BP_THREADS), 0, st, pa);
    hipLaunchKernelGGL(k_density_bwd_pairs_finalize, dim3(blocks_for(n)), dim3(256), 0, st, a,
                       (const unsigned long long*)sums, (const unsigned int*)rmax);
    VDBR_CHECK(ctx, hipGetLastError());
    VDBR_CHECK(ctx, hipEventRecord(e1, st));
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, lt.counters, 16, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    if (((const uint32_t*)ctx->pinned)[3])
        return ctx->fail("density_backward: the grid's nodes do not match its tree header");
    if (ms_out)
        VDBR_CHECK(ctx, hipEventElapsedTime(ms_out, e0, e1));
    *used_pairs = 1;
    return 0;
}


// ============================================================================ graph-captured batch
// Segment counts per partition -> part counts and the exclusive scan of 64-segment chunks that
// k_emit_segments walks (the host form computes them from a read back); a partition past its
// capacity sets GF_SEGMENTS and keeps the segments that fit.
__global__ void k_seg_plan(const unsigned long long* __restrict__ counters, uint64_t part_cap,
                           uint64_t* __restrict__ part_count, uint64_t* __restrict__ chunk_scan,
                           uint32_t* __restrict__ flags)
{
    if (blockIdx.x != 0 || threadIdx.x >= 64u)
        return;
    const uint32_t q = threadIdx.x;
    uint64_t c = counters[q * SEG_COUNTER_STRIDE];
    if (c > part_cap)
    {
        atomicOr(flags, GF_SEGMENTS);
        c = part_cap;
    }
    part_count[q] = c;
    uint64_t chunks = (c + 63) / 64, incl = chunks;
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint64_t o = __shfl_up(incl, off, 64);
        if (q >= (uint32_t)off)
            incl += o;
    }
    chunk_scan[q] = incl - chunks;
    if (q == 63u)
    {
        part_count[SEG_PARTS] = 0;
        chunk_scan[SEG_PARTS] = incl;
    }
}

// emitted pairs -> the sorted count; more than the buffer holds sets GF_PAIRS and empties the rest
__global__ void k_pairs_final(const unsigned long long* __restrict__ counter, uint64_t cap, uint64_t* __restrict__ npairs,
                              uint32_t* __restrict__ flags)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const uint64_t c = *counter;
    if (c > cap)
    {
        atomicOr(flags, GF_PAIRS);
        *npairs = 0;
    }
    else
        *npairs = c;
}

// Output record of one cloud of a captured batch (device memory, read back after the launch).
struct CloudRecord
{
    uint64_t offset;   // byte offset of the grid in the arena (~0: none)
    uint64_t size;     // grid bytes
    uint64_t pairs;    // emitted pairs
    uint64_t voxels;   // active voxels
    uint32_t flags;    // GF_* bits: a bound was exceeded, the cloud must run again uncaptured
    uint32_t upper, lower, leaf;
};

// reserves the cloud's grid in the arena (a running offset shared by the batch) and records its
// counts; a flagged cloud or an arena too small gets no grid
__global__ void k_out_reserve(const uint64_t* __restrict__ grid_size, const uint32_t* __restrict__ node_counts,
                              const uint32_t* __restrict__ nvox, const unsigned long long* __restrict__ pairs,
                              unsigned long long* __restrict__ cursor, uint64_t arena_cap, CloudRecord* __restrict__ rec)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const uint64_t size = *grid_size;
    rec->pairs = *pairs;
    rec->voxels = *nvox;
    rec->upper = node_counts[0];
    rec->lower = node_counts[1];
    rec->leaf = node_counts[2];
    rec->size = size;
    rec->offset = ~0ull;
    if (rec->flags != 0u)
        return;
    const uint64_t at = atomicAdd(cursor, (unsigned long long)((size + 255) & ~255ull));
    if (at + size > arena_cap)
    {
        rec->flags |= GF_ARENA;
        return;
    }
    rec->offset = at;
}

__global__ void k_out_copy(const uint64_t* __restrict__ grid, uint8_t* __restrict__ arena,
                           const CloudRecord* __restrict__ rec)
{
    const uint64_t off = rec->offset;
    if (off == ~0ull)
        return;
    const uint64_t words = (rec->size + 7) / 8;
    uint64_t* dst = (uint64_t*)(arena + off);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = grid[i];
}

// One cloud of a captured batch: the setup counts read back before capture and the bounds the
// captured buffers are sized by.
struct CloudPlan
{
    uint32_t item = 0;
    vdbr_gaussians_t gs{};
    vdbr_raster_params_t prm{};
    SetupBufs sb{};
    uint64_t* batch_starts_dev = nullptr; // foot_scan at every prim-batch start (gathered before capture)
    bool graph = false;
    // read back
    int32_t hb[6] = { 0, 0, 0, 0, 0, 0 };
    uint64_t est = 0, frag_total = 0, total_rows = 0;
    // derived
    uint32_t PB = 0, W = 0, T = 1, max_nodes = 0;
    uint32_t win_off = 0; // first entry of the cloud's window bases in the batch table
    KeyLayout kl{};
    uint32_t key_bits = 0, sort_begin = 0;
    bool rgba_wave = false, density = false;
    uint64_t part_cap = 0, pair_cap = 0, leaf_cap = 0, vox_cap = 0, lower_cap = 0, upper_cap = 0;
};

// the captured stages of one cloud on worker context w (emission to the grid copied into the arena)
static int capture_cloud(Context* w, const CloudPlan& p, const uint32_t* win_base_dev, CloudRecord* rec,
                         uint8_t* arena, uint64_t arena_cap, unsigned long long* cursor)
{
    hipStream_t st = w->stream;
    uint32_t* flags = &rec->flags;
    const uint64_t n = p.gs.count;
    const float vs = p.prm.voxel_size;
    if (fill_words(w, flags, host_count(1), 0u))
        return -1;
    // counter (u64) and the sorted pair count (u64)
    uint64_t* words = (uint64_t*)w->ws(WS_GRAPH, 64);
    if (!words)
        return -1;
    unsigned long long* counter = (unsigned long long*)words;
    uint64_t* npairs = words + 1;
    DensPrim* dprims = nullptr;
    if (p.density)
    {
        dprims = (DensPrim*)w->ws(WS_DENS_PRIMS, n * sizeof(DensPrim));
        if (!dprims)
            return -1;
        hipLaunchKernelGGL(k_dens_prims, dim3(blocks_for(n)), dim3(256), 0, st, n, p.gs.means, p.gs.quats,
                           p.gs.scales, p.gs.opacities, vs, dprims);
    }
    // ---- emission: candidate segments, then the persistent predicate walk
    uint64_t* seg_meta = (uint64_t*)w->ws(WS_SEG_META, (SEG_PARTS * SEG_COUNTER_STRIDE + 2 * (SEG_PARTS + 1)) * 8);
    const bool seg_compact = seg_compact_fits(p.kl);
    void* segs = w->ws(WS_SEGS, SEG_PARTS * p.part_cap * (seg_compact ? sizeof(SegCompact::Rec) : sizeof(SegWide::Rec)));
    uint64_t* pairs = (uint64_t*)w->ws(WS_PAIRS, p.pair_cap * sizeof(uint64_t));
    if (!seg_meta || !segs || !pairs)
        return -1;
    unsigned long long* seg_counters = (unsigned long long*)seg_meta;
    uint64_t* part_count_dev = seg_meta + SEG_PARTS * SEG_COUNTER_STRIDE;
    uint64_t* chunk_scan_dev = part_count_dev + (SEG_PARTS + 1);
    if (fill_words(w, seg_meta, host_count(SEG_PARTS * SEG_COUNTER_STRIDE), 0ull))
        return -1;
    EmitArgs ra;
    memset(&ra, 0, sizeof(ra));
    ra.frames = p.sb.frames;
    ra.boxes = p.sb.boxes;
    ra.row_scan = p.sb.row_scan;
    ra.n = n;
    ra.total_rows = p.total_rows;
    ra.vs = vs;
    ra.slab_i0 = INT32_MIN;
    ra.kl = p.kl;
    ra.seg_compact = seg_compact ? 1u : 0u;
    RowFrame* rframes = (RowFrame*)w->ws(WS_ROWFRAMES, n * sizeof(RowFrame));
    if (!rframes)
        return -1;
    hipLaunchKernelGGL(k_row_frames, dim3(blocks_for(n)), dim3(256), 0, st, n, (const PrimFrame*)p.sb.frames,
                       (const PrimBox*)p.sb.boxes, vs, rframes);
    ra.rframes = rframes;
    const uint64_t blocks_rows = (p.total_rows + EMIT_ROWS - 1) / EMIT_ROWS;
    hipLaunchKernelGGL(k_row_segments, dim3((uint32_t)blocks_rows), dim3(EMIT_THREADS), 0, st, ra, segs, p.part_cap,
                       seg_counters);
    hipLaunchKernelGGL(k_seg_plan, dim3(1), dim3(64), 0, st, (const unsigned long long*)seg_counters, p.part_cap,
                       part_count_dev, chunk_scan_dev, flags);
    const uint32_t hbits = p.key_bits - p.sort_begin;
    const uint32_t hist_w = sort_digit_bits(w, p.pair_cap, hbits);
    unsigned long long* hist_buf =
        emit_counts_hist(hbits, hist_w) ? sort_hist_buffer(w) : nullptr;
    if (fill_words(w, words, host_count(2), 0ull))
        return -1;
    if (hist_buf && fill_words(w, (uint64_t*)hist_buf, host_count(EMIT_HIST_PASSES * SORT_BINS), 0ull))
        return -1;
    EmitArgs ea = ra;
    ea.vals = p.sb.pvals;
    ea.kl = p.kl;
    ea.out = pairs;
    ea.cap = p.pair_cap;
    ea.counter = counter;
    ea.hist = hist_buf;
    ea.hist_shift = p.sort_begin;
    ea.hist_bits = hbits;
    ea.hist_w = hist_w;
    ea.segs = segs;
    ea.seg_part_cap = p.part_cap;
    ea.seg_part_count = part_count_dev;
    ea.seg_chunk_scan = chunk_scan_dev;
    launch_emit_segments(ea, (uint32_t)w->emit_blocks, st);
    hipLaunchKernelGGL(k_pairs_final, dim3(1), dim3(64), 0, st, (const unsigned long long*)counter, p.pair_cap, npairs,
                       flags);
    VDBR_CHECK(w, hipGetLastError());
    // ---- sort over the leaf bits
    const DCount pc = dev_count((const uint64_t*)npairs, p.pair_cap);
    uint64_t* sorted = pairs;
    uint32_t passes = 0;
    if (radix_sort_u64_noswap(w, pairs, pc, p.sort_begin, p.key_bits, &sorted, nullptr, nullptr, &passes,
                              hist_buf != nullptr, true))
        return -1;
    // ---- reduce by leaf, voxel lists
    LeafReduce lr;
    lr.sorted = sorted;
    lr.pairs = pc;
    lr.kl = p.kl;
    lr.vs = vs;
    lr.rgba_wave = p.rgba_wave;
    lr.density = p.density;
    lr.frames = p.sb.frames;
    lr.pvals = p.sb.pvals;
    lr.dprims = dprims;
    lr.leaf_cap = p.leaf_cap;
    lr.vox_cap = p.vox_cap;
    lr.flags = flags;
    lr.leaves_word = (uint32_t*)(words + 3);
    if (reduce_leaves(w, &lr))
        return -1;
    const uint64_t vc = p.vox_cap > 0 ? p.vox_cap : 1;
    VoxelLists vl;
    vl.key = (uint64_t*)w->ws(WS_VOX_KEY, vc * sizeof(uint64_t));
    vl.val = (uint32_t*)w->ws(WS_VOX_VAL, vc * sizeof(uint32_t));
    vl.t = (uint32_t*)w->ws(WS_VOX_T, vc * sizeof(uint32_t));
    vl.g0 = (uint32_t*)w->ws(WS_VOX_G0, vc * sizeof(uint32_t));
    if (!vl.key || !vl.val || !vl.t || !vl.g0)
        return -1;
    // no leaves: the voxel count word is the host 0 of reduce_leaves; an empty device word stands in
    const uint32_t* nvox_dev = (const uint32_t*)lr.svox.dev;
    if (!nvox_dev)
    {
        nvox_dev = (const uint32_t*)(words + 2);
        if (fill_words(w, (uint32_t*)(words + 2), host_count(1), 0u))
            return -1;
    }
    else if (write_voxels(w, lr, vl, p.sb.boxes, p.sb.foot_scan, win_base_dev, p.PB, p.W, p.sb.vbounds))
        return -1;
    // ---- grid build
    Node2BuildInput bin;
    bin.kl = p.kl;
    bin.vox_key = vl.key;
    bin.vox_val = vl.val;
    bin.vox_t = vl.t;
    bin.nvox = p.vox_cap;
    bin.vbounds = p.sb.vbounds;
    bin.T = p.T;
    bin.max_nodes = p.max_nodes;
    bin.spatial_order = (p.prm.flags & VDBR_FLAG_SPATIAL_ORDER) != 0;
    bin.voxel_coords = false;
    bin.value_type = p.density ? VDB_GRID_TYPE_FLOAT : VDB_GRID_TYPE_RGBA8;
    bin.voxel_size = vs;
    bin.nvox_dev = nvox_dev;
    bin.leaf_cap = p.leaf_cap;
    bin.lower_cap = p.lower_cap;
    bin.upper_cap = p.upper_cap;
    bin.flags_dev = flags;
    Node2BuildOutput bout;
    if (build_node2_grid(w, bin, &bout))
        return -1;
    // ---- into the arena
    hipLaunchKernelGGL(k_out_reserve, dim3(1), dim3(64), 0, st, bout.grid_size_dev, bout.counts_dev, nvox_dev,
                       (const unsigned long long*)counter, cursor, arena_cap, rec);
    const uint64_t gw = (bout.grid_size + 7) / 8;
    hipLaunchKernelGGL(k_out_copy, dim3(std::min<uint64_t>((gw + 255) / 256, 2048)), dim3(256), 0, st,
                       (const uint64_t*)bout.grid, arena, (const CloudRecord*)rec);
    VDBR_CHECK(w, hipGetLastError());
    return 0;
}

// the per-cloud part of raster_node2's set-up decisions (key layout, windows, bounds); false when
// the cloud is left to raster_node2
static bool plan_cloud(Context* ctx, CloudPlan& p, const uint64_t* host_starts)
{
    const uint64_t n = p.gs.count;
    p.PB = p.prm.prim_batch_size ? p.prm.prim_batch_size : 262144u;
    p.W = p.prm.ijk_batch_size ? p.prm.ijk_batch_size : 8388608u;
    p.max_nodes = p.prm.max_node_count ? p.prm.max_node_count : 2097152u;
    if (p.frag_total == 0 || (p.prm.flags & VDBR_FLAG_VOXEL_COORDS))
        return false;
    for (int a = 0; a < 3; a++)
        if (p.hb[a] < -65536 || p.hb[3 + a] > 65535)
            return false; // raster_node2 reports the error
    // windows of the node-order model (host_starts: foot_scan at every prim-batch start, then the total)
    const uint64_t nbatch = (n + p.PB - 1) / p.PB;
    uint64_t wtot = 0;
    for (uint64_t b = 0; b < nbatch; b++)
        wtot += (host_starts[b + 1] - host_starts[b] + p.W - 1) / p.W;
    if (wtot >= (1ull << 20))
        return false;
    p.T = (uint32_t)(wtot > 0 ? wtot : 1);
    KeyLayout& kl = p.kl;
    memset(&kl, 0, sizeof(kl));
    uint64_t leaves_bb = 1, lowers_bb = 1, uppers_bb = 1;
    for (int a = 0; a < 3; a++)
    {
        kl.lmin[a] = p.hb[a] >> 3;
        const int32_t lmax = p.hb[3 + a] >> 3;
        kl.bits[a] = bits_for((uint64_t)(lmax - kl.lmin[a] + 1));
        leaves_bb *= (uint64_t)(lmax - kl.lmin[a] + 1);
        lowers_bb *= (uint64_t)((p.hb[3 + a] >> 7) - (p.hb[a] >> 7) + 1);
        uppers_bb *= (uint64_t)((p.hb[3 + a] >> 12) - (p.hb[a] >> 12) + 1);
    }
    kl.gid_bits = bits_for(n > 1 ? n : 2);
    kl.voxel_bits = kl.bits[0] + kl.bits[1] + kl.bits[2] + 9u;
    if (kl.voxel_bits + kl.gid_bits > 64)
        return false;
    p.density = p.prm.value_mode == VDBR_VALUE_DENSITY;
    const bool alpha_in_key = !p.density && kl.voxel_bits + kl.gid_bits + 8u <= 64u;
    kl.alpha_bits = alpha_in_key ? 8u : 0u;
    p.rgba_wave = alpha_in_key;
    p.key_bits = kl.voxel_bits + kl.gid_bits + kl.alpha_bits;
    p.sort_begin = kl.gid_bits + kl.alpha_bits + 9u;
    // one slab only (raster_node2 splits larger clouds)
    const uint64_t slab_budget = p.prm.slab_pairs ? (uint64_t)p.prm.slab_pairs : (1ull << 31);
    if (p.est > slab_budget + slab_budget / 2)
        return false;
    // the bounds of raster_node2's first attempt
    uint64_t seg_total = p.total_rows + (2 * p.est) / EMIT_SEG + 1024;
    if (seg_total > p.total_rows + p.frag_total / EMIT_SEG + 1024)
        seg_total = p.total_rows + p.frag_total / EMIT_SEG + 1024;
    p.part_cap = seg_total / SEG_PARTS + seg_total / (4 * SEG_PARTS) + 4096;
    uint64_t cap = p.est + p.est / 4 + 1024;
    if (cap > p.frag_total)
        cap = p.frag_total;
    const uint64_t pair_limit = std::min<uint64_t>(0xFFFFFFFEull, 2 * slab_budget);
    if (cap > pair_limit)
        cap = pair_limit;
    p.pair_cap = cap;
    p.leaf_cap = std::min<uint64_t>(leaves_bb, cap);
    p.vox_cap = std::min<uint64_t>(cap, p.leaf_cap * 512u);
    p.lower_cap = std::min<uint64_t>(lowers_bb, p.leaf_cap);
    p.upper_cap = std::min<uint64_t>(uppers_bb, p.lower_cap);
    if (1 + p.upper_cap + p.lower_cap + p.leaf_cap > 0xFFFFFFF0ull)
        return false;
    (void)ctx;
    return true;
}

// bound of a cloud's grid bytes (the layout of build_node2_grid at the plan's bounds)
static uint64_t grid_bound(const CloudPlan& p)
{
    return VDB_ROOT_ADDR + VDB_SIZE_ROOT + p.upper_cap * VDB_SIZE_UPPER + p.lower_cap * VDB_SIZE_LOWER +
           p.leaf_cap * VDB_SIZE_LEAF + 2 * VDB_BLINDMETADATA_SIZE + 32 + 4 * (2 + p.vox_cap) + 256;
}

static size_t align256(size_t x)
{
    return (x + 255) & ~(size_t)255;
}

// the batch runs on the context's own stream (a capturable stream), ordered after the caller's
struct OwnStream
{
    Context* ctx;
    hipStream_t caller;
    bool ok = true;
    OwnStream(Context* c) : ctx(c), caller(c->stream)
    {
        if (caller == c->own_stream)
            return;
        hipEvent_t e;
        ok = hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess && hipEventRecord(e, caller) == hipSuccess &&
             hipStreamWaitEvent(c->own_stream, e, 0) == hipSuccess;
        (void)hipEventDestroy(e);
        c->stream = c->own_stream;
    }
    ~OwnStream()
    {
        if (caller == ctx->own_stream)
            return;
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess)
        {
            (void)hipEventRecord(e, ctx->own_stream);
            (void)hipStreamWaitEvent(caller, e, 0);
            (void)hipEventDestroy(e);
        }
        ctx->stream = caller;
    }
};

int raster_batch_graph(Context* ctx, const std::vector<Context*>& workers, vdbr_batch_item_t* items, uint32_t count,
                       std::vector<uint32_t>* rerun, std::vector<uint8_t>* setup_done)
{
    rerun->clear();
    setup_done->assign(count, 0);
    if (count == 0)
        return 0;
    if (workers.empty())
        return ctx->fail("raster batch: no worker contexts");
    const auto t_begin = std::chrono::steady_clock::now();
    OwnStream own(ctx);
    if (!own.ok)
        return ctx->fail("raster batch: could not order the context stream after the caller's");
    hipStream_t st = ctx->stream;
    // ---- validation (raster_node2's checks; invalid clouds fail there with its message)
    std::vector<CloudPlan> plans;
    plans.reserve(count);
    for (uint32_t i = 0; i < count; i++)
    {
        const vdbr_gaussians_t& g = items[i].gaussians;
        const vdbr_raster_params_t& prm = items[i].params;
        items[i].grid_dev = nullptr;
        items[i].grid_size = 0;
        memset(&items[i].stats, 0, sizeof(items[i].stats));
        const bool ok = prm.voxel_size > 0.f && prm.voxel_size < 3.0e38f && g.count > 0 && g.count < (1ull << 32) &&
                        g.means && g.opacities && g.quats && g.scales && g.colors &&
                        (prm.value_mode == VDBR_VALUE_RGBA8 || prm.value_mode == VDBR_VALUE_DENSITY) &&
                        (!(prm.flags & VDBR_FLAG_RAW_INPUTS) || g.sh0);
        if (!ok)
        {
            rerun->push_back(i);
            continue;
        }
        CloudPlan p;
        p.item = i;
        p.gs = g;
        p.prm = prm;
        plans.push_back(p);
    }
    // ---- setup of every cloud into its own buffers, one read back for all
    std::vector<size_t> off(plans.size() + 1, 0);
    std::vector<size_t> starts_at(plans.size() + 1, 0); // u64 entries of the pinned read back
    for (size_t c = 0; c < plans.size(); c++)
    {
        const uint64_t n = plans[c].gs.count;
        const uint64_t PB = plans[c].prm.prim_batch_size ? plans[c].prm.prim_batch_size : 262144u;
        const uint64_t nbatch = (n + PB - 1) / PB;
        off[c + 1] = off[c] + align256(n * sizeof(PrimFrame)) + align256(n * sizeof(PrimBox)) +
                     align256(n * sizeof(PrimVal)) + 2 * align256(n * 8) + 2 * align256((n + 1) * 8) + 256 +
                     align256((nbatch + 1) * 8);
        starts_at[c + 1] = starts_at[c] + 8 + nbatch + 1;
    }
    uint8_t* setup = (uint8_t*)ctx->ws(WS_BATCH_SETUP, off[plans.size()] + 256);
    if (!setup)
        return -1;
    if (ctx->batch_pinned_words < starts_at[plans.size()])
    {
        if (ctx->batch_pinned)
            (void)hipHostFree(ctx->batch_pinned);
        ctx->batch_pinned = nullptr;
        ctx->batch_pinned_words = 0;
        if (hipHostMalloc((void**)&ctx->batch_pinned, starts_at[plans.size()] * 8, hipHostMallocDefault) != hipSuccess)
        {
            (void)hipGetLastError();
            return ctx->fail("raster batch: pinned read-back buffer");
        }
        ctx->batch_pinned_words = starts_at[plans.size()];
    }
    for (size_t c = 0; c < plans.size(); c++)
    {
        CloudPlan& p = plans[c];
        const uint64_t n = p.gs.count;
        uint8_t* b = setup + off[c];
        auto take = [&](size_t bytes) {
            uint8_t* r = b;
            b += align256(bytes);
            return r;
        };
        p.sb.frames = (PrimFrame*)take(n * sizeof(PrimFrame));
        p.sb.boxes = (PrimBox*)take(n * sizeof(PrimBox));
        p.sb.pvals = (PrimVal*)take(n * sizeof(PrimVal));
        p.sb.foot = (uint64_t*)take(n * 8);
        p.sb.rows = (uint64_t*)take(n * 8);
        p.sb.foot_scan = (uint64_t*)take((n + 1) * 8);
        p.sb.row_scan = (uint64_t*)take((n + 1) * 8);
        uint8_t* wordsp = take(256);
        p.sb.bounds = (int32_t*)wordsp;
        p.sb.vbounds = (int32_t*)(wordsp + 32);
        p.sb.est = (unsigned long long*)(wordsp + 64);
        const uint64_t PB = p.prm.prim_batch_size ? p.prm.prim_batch_size : 262144u;
        const uint64_t nbatch = (n + PB - 1) / PB;
        p.batch_starts_dev = (uint64_t*)take((nbatch + 1) * 8);
        if (launch_setup(ctx, &p.gs, &p.prm, p.sb, (p.prm.flags & VDBR_FLAG_RAW_INPUTS) != 0))
            return -1;
        (*setup_done)[p.item] = 1; // raw inputs are preprocessed in place from here on
        hipLaunchKernelGGL(k_gather_u64, dim3(blocks_for(nbatch + 1)), dim3(256), 0, st, p.sb.foot_scan, PB,
                           nbatch, p.batch_starts_dev);
        uint64_t* h = ctx->batch_pinned + starts_at[c];
        VDBR_CHECK(ctx, hipMemcpyAsync(h, p.sb.bounds, 24, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(h + 3, p.sb.est, 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(h + 4, p.sb.foot_scan + n, 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(h + 5, p.sb.row_scan + n, 8, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipMemcpyAsync(h + 8, p.batch_starts_dev, nbatch * 8, hipMemcpyDeviceToHost, st));
    }
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    // ---- plans, window tables, output bounds
    std::vector<uint32_t> win_table;
    uint64_t arena_need = 0;
    std::vector<size_t> graph_clouds;
    for (size_t c = 0; c < plans.size(); c++)
    {
        CloudPlan& p = plans[c];
        const uint64_t* h = ctx->batch_pinned + starts_at[c];
        memcpy(p.hb, h, 24);
        p.est = h[3];
        p.frag_total = h[4];
        p.total_rows = h[5];
        const uint64_t n = p.gs.count;
        const uint64_t PB = p.prm.prim_batch_size ? p.prm.prim_batch_size : 262144u;
        const uint64_t nbatch = (n + PB - 1) / PB;
        std::vector<uint64_t> bs(nbatch + 1);
        for (uint64_t b = 0; b < nbatch; b++)
            bs[b] = h[8 + b];
        bs[nbatch] = p.frag_total;
        p.graph = plan_cloud(ctx, p, bs.data());
        if (!p.graph)
        {
            rerun->push_back(p.item);
            continue;
        }
        p.win_off = (uint32_t)win_table.size();
        const uint64_t W = p.W;
        uint64_t wtot = 0;
        for (uint64_t b = 0; b < nbatch; b++)
        {
            win_table.push_back((uint32_t)wtot);
            wtot += (bs[b + 1] - bs[b] + W - 1) / W;
        }
        win_table.push_back((uint32_t)wtot);
        arena_need += align256(grid_bound(p));
        graph_clouds.push_back(c);
    }
    if (graph_clouds.empty())
        return 0;
    // ---- per-cloud records and window tables (uploaded before capture)
    const size_t rec_bytes = align256(graph_clouds.size() * sizeof(CloudRecord));
    uint8_t* table = (uint8_t*)ctx->ws(WS_BATCH_TABLE, rec_bytes + 256 + align256(win_table.size() * 4));
    if (!table)
        return -1;
    CloudRecord* recs = (CloudRecord*)table;
    unsigned long long* cursor = (unsigned long long*)(table + rec_bytes);
    uint32_t* win_dev = (uint32_t*)(table + rec_bytes + 256);
    VDBR_CHECK(ctx, hipMemcpyAsync(win_dev, win_table.data(), win_table.size() * 4, hipMemcpyHostToDevice, st));
    // The arena: every grid at its bound would often be ten times the grids' real size, so it is
    // sized from the grids of the last batch on this context (8 GiB before the first) and at most a
    // quarter of the free device memory; a grid that does not fit is flagged and its cloud runs
    // again outside the graph.
    uint64_t arena_want = std::min<uint64_t>(arena_need, ctx->batch_arena_hint ? ctx->batch_arena_hint : (8ull << 30));
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            arena_want = std::min<uint64_t>(arena_want, std::max<uint64_t>((uint64_t)free_b / 4 + ctx->ws_size[WS_BATCH_ARENA],
                                                                          256ull << 20));
        (void)hipGetLastError();
    }
    uint8_t* arena = (uint8_t*)ctx->ws(WS_BATCH_ARENA, std::max<uint64_t>(arena_want, 256));
    if (!arena)
        return -1;
    const uint64_t arena_cap = ctx->ws_size[WS_BATCH_ARENA];
    // ---- clouds to worker streams, largest estimate first onto the least loaded
    const size_t K = std::min<size_t>(workers.size(), graph_clouds.size());
    std::vector<std::vector<size_t>> assign(K);
    {
        std::vector<size_t> order = graph_clouds;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return plans[a].est > plans[b].est; });
        std::vector<uint64_t> load(K, 0);
        for (size_t c : order)
        {
            size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            assign[k].push_back(c);
            load[k] += plans[c].est + plans[c].gs.count * 64;
        }
    }
    std::vector<size_t> rec_of(plans.size(), 0);
    for (size_t r = 0; r < graph_clouds.size(); r++)
        rec_of[graph_clouds[r]] = r;
    for (size_t k = 0; k < K; k++)
        std::copy(ctx->sort_tuned, ctx->sort_tuned + SORT_BUCKETS, workers[k]->sort_tuned);

    // ---- every host value the capture depends on: pointers and sizes of the workspaces, the
    // clouds and their plans, the split over workers. An unchanged signature relaunches the graph
    // captured by an earlier call instead of capturing again.
    auto signature = [&]() {
        std::vector<uint64_t> sig;
        auto put = [&](const void* p, size_t bytes) {
            const uint8_t* b = (const uint8_t*)p;
            for (size_t x = 0; x < bytes; x += 8)
            {
                uint64_t v = 0;
                memcpy(&v, b + x, std::min<size_t>(8, bytes - x));
                sig.push_back(v);
            }
        };
        sig.push_back(K);
        for (WsSlot sl : { WS_BATCH_SETUP, WS_BATCH_TABLE, WS_BATCH_ARENA })
        {
            sig.push_back((uint64_t)(uintptr_t)ctx->ws_ptr[sl]);
            sig.push_back(ctx->ws_size[sl]);
        }
        for (size_t k = 0; k < K; k++)
        {
            const Context* w = workers[k];
            put(w->ws_ptr, sizeof(w->ws_ptr));
            put(w->ws_size, sizeof(w->ws_size));
            put(w->sort_tuned, sizeof(w->sort_tuned));
            sig.push_back(w->emit_blocks);
            sig.push_back((uint64_t)(uintptr_t)w->stream);
            sig.push_back(assign[k].size());
            for (size_t c : assign[k])
            {
                const CloudPlan& p = plans[c];
                sig.push_back(p.item);
                put(&p.gs, sizeof(p.gs));
                put(&p.prm, sizeof(p.prm));
                put(&p.sb, sizeof(p.sb));
                put(p.hb, sizeof(p.hb));
                sig.push_back(p.est);
                sig.push_back(p.frag_total);
                sig.push_back(p.total_rows);
                sig.push_back(p.T);
                sig.push_back(p.win_off);
            }
        }
        return sig;
    };
    const auto t_plan = std::chrono::steady_clock::now();
    // One graph per worker stream, launched together: a single graph holding the workers as
    // parallel branches measured 964 ms per 256-cloud batch with 4 workers against 877 ms for
    // per-stream graphs (the graph executor overlaps a graph's branches less than separate
    // streams overlap).
    std::vector<hipGraphExec_t>& execs = ctx->batch_execs;
    bool captured = false;
    if (execs.size() != K || ctx->batch_sig != signature())
    {
        ctx->drop_batch_graphs();
        // a sizing pass per worker (discarded), the workspaces grown to it, then the real capture
        for (int pass = 0; pass < 2; pass++)
        {
            const bool probe = pass == 0;
            for (size_t k = 0; k < K; k++)
            {
                Context* w = workers[k];
                if (!probe)
                {
                    for (int sl = 0; sl < WS_COUNT; sl++)
                        if (w->probe_size[sl] && !w->ws((WsSlot)sl, w->probe_size[sl]))
                            return ctx->fail("raster batch: worker workspace: " + w->error);
                    // the captured sorts use look-back epochs from 1 (the status words are cleared
                    // before every launch)
                    w->sort_epoch = 0;
                    VDBR_CHECK(ctx, hipStreamSynchronize(w->stream));
                }
                VDBR_CHECK(ctx, hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
                w->probe = probe;
                w->capturing = true;
                if (probe)
                    std::fill(w->probe_size, w->probe_size + WS_COUNT, (size_t)0);
                int rc = 0;
                std::string err;
                for (size_t c : assign[k])
                {
                    const CloudPlan& p = plans[c];
                    if (capture_cloud(w, p, win_dev + p.win_off, recs + rec_of[c], arena, arena_cap, cursor))
                    {
                        rc = -1;
                        err = w->error;
                        break;
                    }
                }
                w->probe = false;
                w->capturing = false;
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(w->stream, &g);
                if (rc || e != hipSuccess)
                {
                    if (g)
                        (void)hipGraphDestroy(g);
                    (void)hipGetLastError();
                    ctx->drop_batch_graphs();
                    return ctx->fail("raster batch: capture failed: " +
                                     (err.empty() ? std::string(hipGetErrorString(e)) : err));
                }
                if (probe)
                {
                    (void)hipGraphDestroy(g);
                    continue;
                }
                hipGraphExec_t x = nullptr;
                e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess)
                {
                    (void)hipGetLastError();
                    ctx->drop_batch_graphs();
                    return ctx->fail(std::string("raster batch: graph instantiation: ") + hipGetErrorString(e));
                }
                execs.push_back(x);
            }
        }
        ctx->batch_sig = signature();
        captured = true;
    }
    const auto t_graph = std::chrono::steady_clock::now();
    // before every launch: the output cursor and the workers' look-back status words cleared
    if (fill_words(ctx, (uint64_t*)cursor, host_count(1), 0ull))
        return -1;
    for (size_t k = 0; k < K; k++)
    {
        Context* w = workers[k];
        if (w->ws_ptr[WS_SORT_STATUS] &&
            fill_words(ctx, (uint64_t*)w->ws_ptr[WS_SORT_STATUS], host_count(w->ws_size[WS_SORT_STATUS] / 8), 0ull))
            return -1;
    }
    // the worker streams wait for the context's stream, run their graphs, and the context's stream
    // waits for all of them
    hipEvent_t go = nullptr;
    VDBR_CHECK(ctx, hipEventCreateWithFlags(&go, hipEventDisableTiming));
    std::vector<hipEvent_t> done(K, nullptr);
    hipError_t e = hipEventRecord(go, st);
    for (size_t k = 0; k < K && e == hipSuccess; k++)
    {
        e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(workers[k]->stream, go, 0);
        if (e == hipSuccess)
            e = hipGraphLaunch(execs[k], workers[k]->stream);
        if (e == hipSuccess)
            e = hipEventRecord(done[k], workers[k]->stream);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(st, done[k], 0);
    }
    std::vector<CloudRecord> hrec(graph_clouds.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(hrec.data(), recs, graph_clouds.size() * sizeof(CloudRecord), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    (void)hipEventDestroy(go);
    for (hipEvent_t d : done)
        if (d)
            (void)hipEventDestroy(d);
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return ctx->fail(std::string("raster batch: graph launch: ") + hipGetErrorString(e));
    }
    const auto t_run = std::chrono::steady_clock::now();
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return (float)std::chrono::duration<double, std::milli>(b - a).count();
    };
    ctx->batch_ms[0] = ms(t_begin, t_plan);
    ctx->batch_ms[1] = ms(t_plan, t_graph);
    ctx->batch_ms[2] = ms(t_graph, t_run);
    ctx->batch_ms[3] = captured ? 1.f : 0.f;
    // the next batch sizes its arena from these grids (with a quarter of headroom)
    uint64_t grids_total = 0;
    for (const CloudRecord& cr : hrec)
        grids_total += (cr.size + 255) & ~255ull;
    ctx->batch_arena_hint = std::max<uint64_t>(grids_total + grids_total / 4, 256ull << 20);
    for (size_t r = 0; r < graph_clouds.size(); r++)
    {
        const CloudPlan& p = plans[graph_clouds[r]];
        const CloudRecord& cr = hrec[r];
        vdbr_batch_item_t& it = items[p.item];
        if (cr.flags != 0u || cr.offset == ~0ull)
        {
            rerun->push_back(p.item);
            continue;
        }
        it.grid_dev = arena + cr.offset;
        it.grid_size = cr.size;
        vdbr_raster_stats_t& s = it.stats;
        s.gaussian_count = p.gs.count;
        s.fragment_count = p.frag_total;
        s.pair_count = cr.pairs;
        s.voxel_count = cr.voxels;
        s.grid_size = cr.size;
        s.upper_count = cr.upper;
        s.lower_count = cr.lower;
        s.leaf_count = cr.leaf;
        s.window_count = p.T;
        s.key_bits = p.key_bits;
        s.sort_bit_begin = p.sort_begin;
        s.emit_attempts = 1;
        s.slab_count = 1;
    }
    std::sort(rerun->begin(), rerun->end());
    return 0;
}

} // namespace vdbr
