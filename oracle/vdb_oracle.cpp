/*
 * vdb_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle; never linked into the product).
 *
 * A sequential CPU restatement of the reference raster3d path of openvdb/nanovdb-editor.
 * The node2 build is emulated window by window exactly as the reference GPU passes run
 * (set_masks -> count_new -> scan -> scatter_new -> finalize per tree level, per fragment
 * window), on a byte buffer with the reference node2 layout. The output grid is therefore
 * the reference's grid, including its allocation-order-dependent node offsets.
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math: IEEE FP32 order of
 * operations as written in the shaders; x86-64 SSE has no implicit FMA).
 *
 * Transcendentals: the reference host code uses libm expf/sqrtf (RasterUtils.cpp:67-116);
 * the shader log/exp/pow/sqrt precision is implementation defined on Vulkan. The contract
 * of this build is "correctly rounded" for all of them. Here they are evaluated in x87
 * extended precision (expl/logl/powl, 64-bit significand) and rounded once to float;
 * glibc's expf/logf are not correctly rounded (about 1 in 10^4 inputs differs by one ulp).
 *
 * Parity pins (see tests/test_oracle.py):
 *   vdbo_preprocess        pinned: closed forms (unit quats, exp, sigmoid, SH0 constant).
 *   vdbo_prim_aabb         pinned: axis-aligned closed form floor((mu +- r s)/vs) and brute force.
 *   vdbo_pairs             pinned: brute force over a bounding cube (every voxel tested),
 *                          analytic inclusion of the mean's voxel, rotation invariance cases.
 *   vdbo_raster_node2      pinned: structural invariants (masks == OR of children, prefix
 *                          lanes == brute popcounts, value idx == exclusive scan from 2,
 *                          round trip of active voxels == pair voxel set), window-order
 *                          invariance of topology, single-window node order == BFS order.
 *                          Byte layout of NanoVDB headers (PNanoVDB.h, not in the reference
 *                          tree) is restated from the public format: parity unpinned.
 *   vdbo_density_f64       pinned: single axis-aligned Gaussian closed form.
 *   vdbo_density_grad_f64  pinned: FP64 central finite differences.
 *   vdbo_node2_active_voxels / vdbo_node2_to_onindex
 *                          pinned: the OnIndex grid's active voxels and value indices equal
 *                          the node2 grid's (the Convert.h validation); OnIndex header bytes
 *                          are restated from the public NanoVDB format: parity unpinned.
 *   vdbo_tile_bin          pinned: intersections equal a brute-force half-open interval test
 *                          per tile, the order equals a stable sort by (tile, depth bits), and
 *                          the offsets equal a left binary search of every tile index.
 */
#include "vdb_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

namespace
{

// correctly rounded FP32 transcendentals (extended-precision evaluation, single rounding)
inline float cr_expf(float x)
{
    return (float)expl((long double)x);
}
inline float cr_logf(float x)
{
    return (float)logl((long double)x);
}
inline float cr_powf(float x, float y)
{
    return (float)powl((long double)x, (long double)y);
}

// ------------------------------------------------------------------------------------------
// node2 layout (PNanoVDBExt.h:50-83), u64 word offsets
const uint32_t OFF_NODE_IDX = 0u;
const uint32_t OFF_VALUE_IDX = 1u;
const uint32_t OFF_VALUE_LEVEL = 2u;
const uint32_t OFF_VMP[4] = { 258, 258, 34, 6 };
const uint32_t OFF_VM[4] = { 258 + 128, 258 + 128, 34 + 16, 6 + 2 };
const uint32_t OFF_CMP[4] = { 898, 898, 114, 6 };
const uint32_t OFF_CM[4] = { 898 + 128, 898 + 128, 114 + 16, 6 + 2 };
const uint32_t OFF_KMP[4] = { 1538, 898, 114, 6 };
const uint32_t OFF_KM[4] = { 1538 + 128, 898 + 128, 114 + 16, 6 + 2 };
const uint32_t OFF_CHILDREN[4] = { 2180, 1538, 194, 16 };
const uint32_t PREFIX_OFF_MASK[4] = { 128, 128, 16, 2 };
const uint32_t MAX_SIZE[4] = { 541728, 274464, 34336, 128 };
const uint32_t IJK_MASK[4] = { 0x1FFFF, 0xFFF, 0x7F, 0x7 };
const uint32_t FANOUT_BITS2[4] = { 10, 10, 8, 6 };
const uint32_t FANOUT_BITS[4] = { 5, 5, 4, 3 };
const uint32_t TILEDIM_BITS[4] = { 12, 7, 3, 0 };
const uint32_t FANOUT_1D[4] = { 32768, 32768, 4096, 512 };
const uint32_t LEVEL_MASK_I[4] = { 0x7800, 0x7800, 0xE00, 0x180 };
const uint32_t LEVEL_MASK_J[4] = { 0x03C0, 0x03C0, 0x0E0, 0x030 };
const uint32_t LEVEL_MASK_K[4] = { 0x001E, 0x001E, 0x00E, 0x006 };
const uint64_t END_KEY = 1ull << 63u;
enum
{
    ROOT = 0,
    UPPER = 1,
    LOWER = 2,
    LEAF = 3
};

// NanoVDB header layout (public NanoVDB/PNanoVDB format; not present in the reference tree)
const uint32_t GRID_SIZE = 672u;
const uint32_t TREE_SIZE = 64u;
const uint32_t BLINDMETADATA_SIZE = 288u;
const uint64_t MAGIC_GRID = 0x314244566f6e614eull; // "NanoVDB1", FileHeaderInfo.cpp:184
const uint32_t VERSION = (32u << 21u) | (7u << 10u) | 0u;
const uint32_t GRID_TYPE_VEC3F = 6u; // FileHeaderInfo.cpp:33
const uint32_t GRID_TYPE_FLOAT = 1u;
const uint32_t GRID_TYPE_RGBA8 = 12u;
const uint32_t GRID_TYPE_NODE2 = 27u; // PNanoVDBExt.h:21

// ------------------------------------------------------------------------------------------
struct Buf
{
    std::vector<uint64_t> w;
    uint64_t r(uint64_t idx) const
    {
        return idx < w.size() ? w[idx] : 0ull;
    }
    void wr(uint64_t idx, uint64_t v)
    {
        if (idx >= w.size())
            w.resize(std::max<uint64_t>(idx + 1, w.size() * 2 + 1024), 0ull);
        w[idx] = v;
    }
    void ensure_bytes(uint64_t nbytes)
    {
        uint64_t words = (nbytes + 7) / 8;
        if (words > w.size())
            w.resize(words, 0ull);
    }
    void put(uint64_t byte_off, const void* src, uint32_t n)
    {
        ensure_bytes(byte_off + n);
        memcpy(reinterpret_cast<uint8_t*>(w.data()) + byte_off, src, n);
    }
    void get(uint64_t byte_off, void* dst, uint32_t n) const
    {
        if (byte_off + n > w.size() * 8)
        {
            memset(dst, 0, n);
            return;
        }
        memcpy(dst, reinterpret_cast<const uint8_t*>(w.data()) + byte_off, n);
    }
    void u32(uint64_t off, uint32_t v)
    {
        put(off, &v, 4);
    }
    void u64(uint64_t off, uint64_t v)
    {
        put(off, &v, 8);
    }
    void f32(uint64_t off, float v)
    {
        put(off, &v, 4);
    }
    void f64(uint64_t off, double v)
    {
        put(off, &v, 8);
    }
    uint32_t gu32(uint64_t off) const
    {
        uint32_t v;
        get(off, &v, 4);
        return v;
    }
    uint64_t gu64(uint64_t off) const
    {
        uint64_t v;
        get(off, &v, 8);
        return v;
    }
};

struct Coord
{
    int32_t x, y, z;
};

const uint64_t TREE_ADDR = GRID_SIZE;
const uint64_t ROOT_ADDR = GRID_SIZE + TREE_SIZE;
const uint32_t ROOT_IDX64 = (uint32_t)(ROOT_ADDR >> 3);

uint32_t tree_count(const Buf& b, int type)
{
    // tree: node_count_leaf @32, lower @36, upper @40
    if (type == LEAF)
        return b.gu32(TREE_ADDR + 32);
    if (type == LOWER)
        return b.gu32(TREE_ADDR + 36);
    return b.gu32(TREE_ADDR + 40);
}
void set_tree_count(Buf& b, int type, uint32_t v)
{
    if (type == LEAF)
        b.u32(TREE_ADDR + 32, v);
    else if (type == LOWER)
        b.u32(TREE_ADDR + 36, v);
    else
        b.u32(TREE_ADDR + 40, v);
}

// PNanoVDBExt.h:94-103
uint32_t coord_to_n(uint32_t t, Coord ijk)
{
    return ((((uint32_t)ijk.x & IJK_MASK[t]) >> TILEDIM_BITS[t]) << FANOUT_BITS2[t]) |
           ((((uint32_t)ijk.y & IJK_MASK[t]) >> TILEDIM_BITS[t]) << FANOUT_BITS[t]) |
           (((uint32_t)ijk.z & IJK_MASK[t]) >> TILEDIM_BITS[t]);
}

// PNanoVDBExt.h:105-126
uint32_t adaptive_n(const Buf& b, uint32_t node, uint32_t t, uint32_t n, uint32_t* out_level)
{
    uint32_t level_idx = ((n & LEVEL_MASK_I[t]) >> 3) | ((n & LEVEL_MASK_J[t]) >> 2) | ((n & LEVEL_MASK_K[t]) >> 1);
    uint32_t level = (uint32_t)(b.r(node + OFF_VALUE_LEVEL + (level_idx >> 4u)) >> ((level_idx << 2u) & 63u)) & 15u;
    uint32_t lm = (1u << level) - 1u;
    uint32_t lm3 = (lm << FANOUT_BITS2[t]) | (lm << FANOUT_BITS[t]) | lm;
    *out_level = level + TILEDIM_BITS[t];
    return n & ~lm3;
}

// PNanoVDBExt.h:128-143
uint32_t mask_n_to_idx(const Buf& b, uint32_t prefix_idx64, uint32_t t, uint32_t n, bool* active)
{
    uint64_t mask = b.r(prefix_idx64 + PREFIX_OFF_MASK[t] + (n >> 6u));
    uint32_t prefix = (uint32_t)(b.r(prefix_idx64 + (n >> 8u)) >> ((n >> 2u) & 0x30u)) & 0xFFFFu;
    uint64_t cmp = 1ull << (n & 63u);
    uint32_t idx = (uint32_t)__builtin_popcountll(mask & (cmp - 1ull)) + (n >= 64u ? prefix : 0u);
    if (active)
        *active = (mask & cmp) != 0ull;
    return idx;
}

uint64_t root_key(Coord ijk)
{
    return (uint64_t)((uint32_t)ijk.z >> 12) | ((uint64_t)((uint32_t)ijk.y >> 12) << 21) |
           ((uint64_t)((uint32_t)ijk.x >> 12) << 42);
}

// PNanoVDBExt.h:145-182
int32_t find_n_offset_to_key(const Buf& b, uint32_t node, uint32_t n_off, uint32_t node_n, Coord ijk)
{
    int32_t match = 0;
    uint64_t key = root_key(ijk);
    uint32_t child_count = (uint32_t)b.r(node + OFF_CMP[ROOT]) & 0xFFFFu;
    for (; n_off < 32768u; n_off++)
    {
        uint32_t cmp_n = (node_n + n_off) & 32767u;
        bool key_active = false;
        uint32_t key_idx = mask_n_to_idx(b, node + OFF_KMP[ROOT], ROOT, cmp_n, &key_active);
        if (key_active)
        {
            uint64_t cmp_key = b.r(node + OFF_CHILDREN[ROOT] + child_count + key_idx);
            if (key == cmp_key || END_KEY == cmp_key)
            {
                match = (int32_t)(cmp_n - node_n);
                break;
            }
        }
    }
    return match;
}

// PNanoVDBExt.h:195-259
void find_node_pass(const Buf& b, uint32_t node, int32_t root_n_off, uint32_t* out_node, uint32_t* out_type,
                    uint32_t* out_n, bool* key_is_match, Coord ijk)
{
    *key_is_match = true;
    uint32_t t = 0, n = 0, level = 0;
    for (; t < 4u; t++)
    {
        n = coord_to_n(t, ijk);
        n = adaptive_n(b, node, t, n, &level);
        if (t == ROOT)
        {
            n += (uint32_t)root_n_off;
            uint64_t key = root_key(ijk);
            uint32_t child_count = (uint32_t)b.r(node + OFF_CMP[ROOT]) & 0xFFFFu;
            bool key_active = false;
            uint32_t key_idx = mask_n_to_idx(b, node + OFF_KMP[ROOT], ROOT, n, &key_active);
            uint64_t cmp_key = b.r(node + OFF_CHILDREN[ROOT] + child_count + key_idx);
            *key_is_match = !key_active || (key_active && (key == cmp_key || END_KEY == cmp_key));
        }
        if (t == LEAF)
            break;
        bool is_active = false;
        uint32_t child_idx = mask_n_to_idx(b, node + OFF_CMP[t], t, n, &is_active);
        uint32_t child = 0;
        if (is_active)
            child = (uint32_t)(b.r(node + OFF_CHILDREN[t] + child_idx) >> 3u);
        if (child == 0u)
            break;
        node = node + child;
    }
    *out_node = node;
    *out_type = t;
    *out_n = n;
}

// PNanoVDBExt.h:261-281
void find_node(const Buf& b, Coord ijk, uint32_t* out_node, uint32_t* out_type, uint32_t* out_n)
{
    bool match = true;
    find_node_pass(b, ROOT_IDX64, 0, out_node, out_type, out_n, &match, ijk);
    if (!match)
    {
        uint32_t level = 0;
        uint32_t node_n = coord_to_n(ROOT, ijk);
        node_n = adaptive_n(b, ROOT_IDX64, ROOT, node_n, &level);
        int32_t n_off = find_n_offset_to_key(b, ROOT_IDX64, 1u, node_n, ijk);
        find_node_pass(b, ROOT_IDX64, n_off, out_node, out_type, out_n, &match, ijk);
    }
}

// grid_build_count_new.slang:58-99 -- packed 16-bit inclusive prefix lanes, lane 0 of word 0
// holds the total (the wrap-around write). Returns the total popcount.
uint32_t write_prefix(Buf& b, uint32_t node, uint32_t t, uint32_t mask_off, uint32_t prefix_off)
{
    uint32_t words = FANOUT_1D[t] >> 6u;
    std::vector<uint32_t> incl(words);
    uint32_t acc = 0;
    for (uint32_t wi = 0; wi < words; wi++)
    {
        acc += (uint32_t)__builtin_popcountll(b.r(node + mask_off + wi));
        incl[wi] = acc;
    }
    for (uint32_t q = 0; q < words / 4u; q++)
    {
        uint64_t lane0 = (q == 0u) ? acc : incl[4u * q - 1u];
        uint64_t v = (lane0 & 0xFFFFull) | ((uint64_t)(incl[4u * q + 0u] & 0xFFFFu) << 16) |
                     ((uint64_t)(incl[4u * q + 1u] & 0xFFFFu) << 32) | ((uint64_t)(incl[4u * q + 2u] & 0xFFFFu) << 48);
        b.wr(node + prefix_off + q, v);
    }
    return acc;
}

// GridBuild.cpp:118-187 -- empty node2 grid
void empty_grid(Buf& b, float voxel_size)
{
    float vs_inv = 1.f / voxel_size;
    b.w.assign((ROOT_ADDR + MAX_SIZE[ROOT]) / 8, 0ull);
    b.u64(0, MAGIC_GRID);
    b.u32(16, VERSION);
    b.u32(20, 0u);
    b.u32(24, 0u);
    b.u32(28, 1u);
    b.u64(32, ROOT_ADDR + MAX_SIZE[ROOT]);
    b.u32(40, 0x65646f6eu); // "node2"
    b.u32(44, 0x00000032u);
    b.f64(608 + 0, (double)voxel_size);
    b.f64(608 + 8, (double)voxel_size);
    b.f64(608 + 16, (double)voxel_size);
    b.u32(632, 0u);
    b.u32(636, GRID_TYPE_NODE2);
    const uint64_t map = 296;
    b.f32(map + 0 + 4 * 0, voxel_size);
    b.f32(map + 0 + 4 * 4, voxel_size);
    b.f32(map + 0 + 4 * 8, voxel_size);
    b.f32(map + 36 + 4 * 0, vs_inv);
    b.f32(map + 36 + 4 * 4, vs_inv);
    b.f32(map + 36 + 4 * 8, vs_inv);
    b.f64(map + 88 + 8 * 0, (double)voxel_size);
    b.f64(map + 88 + 8 * 4, (double)voxel_size);
    b.f64(map + 88 + 8 * 8, (double)voxel_size);
    b.f64(map + 160 + 8 * 0, (double)vs_inv);
    b.f64(map + 160 + 8 * 4, (double)vs_inv);
    b.f64(map + 160 + 8 * 8, (double)vs_inv);
    for (uint32_t i = 0; i < 2u * FANOUT_1D[ROOT]; i++)
        b.wr(ROOT_IDX64 + OFF_CHILDREN[ROOT] + i, END_KEY);
    b.u64(TREE_ADDR + 24, ROOT_ADDR - TREE_ADDR);
}

struct Builder
{
    Buf b;
    std::vector<uint32_t> node_addr;
    std::vector<uint32_t> node_type;
    std::vector<uint32_t> last_value_scan; // scan of the last value pass
    uint32_t max_node_count;
    bool overflow = false;

    uint32_t node_count() const
    {
        return 1u + tree_count(b, UPPER) + tree_count(b, LOWER) + tree_count(b, LEAF);
    }
    uint64_t end_addr() const
    {
        return ROOT_ADDR + MAX_SIZE[ROOT] + (uint64_t)tree_count(b, UPPER) * MAX_SIZE[UPPER] +
               (uint64_t)tree_count(b, LOWER) * MAX_SIZE[LOWER] + (uint64_t)tree_count(b, LEAF) * MAX_SIZE[LEAF];
    }

    // GridBuild.cpp:361-506 -- one window: 4 levels of set_masks/count_new/scan/scatter_new/finalize
    void build_window(const std::vector<Coord>& ijks)
    {
        for (int level = 0; level < 4; level++)
        {
            int A = level, C = level < 3 ? level + 1 : LEAF;
            // grid_build_set_masks.slang:28-79
            for (const Coord& ijk : ijks)
            {
                uint32_t node, t, n;
                find_node(b, ijk, &node, &t, &n);
                if ((int)t != A)
                    continue;
                uint32_t base = (t == LEAF) ? OFF_VM[t] : OFF_CM[t];
                uint64_t bit = 1ull << (n & 63u);
                uint64_t idx = node + base + (n >> 6u);
                uint64_t old = b.r(idx);
                b.wr(idx, old | bit);
                if (t != LEAF && (old & bit) == 0ull)
                {
                    uint64_t vidx = node + OFF_VM[t] + (n >> 6u);
                    b.wr(vidx, b.r(vidx) | bit);
                }
            }
            uint32_t nc = node_count();
            if (nc > max_node_count)
            {
                overflow = true;
                return;
            }
            // grid_build_count_new.slang:30-183
            std::vector<uint32_t> counts(nc, 0u), scan(nc, 0u);
            for (uint32_t i = 0; i < nc; i++)
            {
                uint32_t node = node_addr[i] >> 3u, t = node_type[i];
                if (A == C)
                {
                    counts[i] = write_prefix(b, node, t, OFF_VM[t], OFF_VMP[t]) + (t == ROOT ? 2u : 0u);
                }
                else if ((int)t == A && t != LEAF)
                {
                    write_prefix(b, node, t, OFF_CM[t], OFF_CMP[t]);
                    uint32_t cnt = 0;
                    for (uint32_t wi = 0; wi < (FANOUT_1D[t] >> 6u); wi++)
                        cnt += (uint32_t)__builtin_popcountll(b.r(node + OFF_CM[t] + wi) & b.r(node + OFF_VM[t] + wi));
                    counts[i] = cnt;
                }
            }
            // ParallelPrimitives.cpp:276 global_scan (inclusive)
            uint32_t acc = 0;
            for (uint32_t i = 0; i < nc; i++)
            {
                acc += counts[i];
                scan[i] = acc;
            }
            // grid_build_scatter_new.slang:31-193
            uint64_t end = end_addr();
            for (uint32_t i = 0; i < nc; i++)
            {
                uint32_t node = node_addr[i] >> 3u, t = node_type[i];
                if (A == C)
                {
                    uint32_t value_idx = scan[i] - counts[i] + (t == ROOT ? 2u : 0u);
                    b.wr(node + OFF_VALUE_IDX, (uint64_t)value_idx);
                    continue;
                }
                if ((int)t != A || t == LEAF)
                    continue;
                uint32_t fan = FANOUT_1D[t];
                // old children in old rank order
                std::vector<uint64_t> old_children;
                uint32_t total_children = 0;
                for (uint32_t n = 0; n < fan; n++)
                {
                    uint64_t cm = b.r(node + OFF_CM[t] + (n >> 6u)), vm = b.r(node + OFF_VM[t] + (n >> 6u));
                    uint64_t bit = 1ull << (n & 63u);
                    if (cm & bit)
                    {
                        total_children++;
                        if (!(vm & bit))
                            old_children.push_back(0ull);
                    }
                }
                for (uint64_t k = 0; k < old_children.size(); k++)
                    old_children[k] = b.r(node + OFF_CHILDREN[t] + k);
                uint32_t rank = 0, old_rank = 0, new_rank = 0;
                uint32_t new_base = scan[i] - counts[i];
                for (uint32_t n = 0; n < fan; n++)
                {
                    uint64_t cm = b.r(node + OFF_CM[t] + (n >> 6u)), vm = b.r(node + OFF_VM[t] + (n >> 6u));
                    uint64_t bit = 1ull << (n & 63u);
                    if (!(cm & bit))
                        continue;
                    if (vm & bit)
                    {
                        uint32_t new_child_idx = new_base + new_rank++;
                        uint64_t addr = end + (uint64_t)new_child_idx * MAX_SIZE[C];
                        uint32_t new_node_idx = new_child_idx + nc;
                        if (new_node_idx < max_node_count)
                        {
                            if (node_addr.size() <= new_node_idx)
                            {
                                node_addr.resize(new_node_idx + 1, 0u);
                                node_type.resize(new_node_idx + 1, 0u);
                            }
                            node_addr[new_node_idx] = (uint32_t)addr;
                            node_type[new_node_idx] = (uint32_t)C;
                            b.wr(node + OFF_CHILDREN[t] + rank, addr - ((uint64_t)node << 3u));
                            b.wr((addr >> 3u) + OFF_NODE_IDX, (uint64_t)new_node_idx);
                        }
                        else
                        {
                            overflow = true;
                        }
                    }
                    else
                    {
                        b.wr(node + OFF_CHILDREN[t] + rank, old_children[old_rank++]);
                    }
                    rank++;
                }
                (void)total_children;
                for (uint32_t wi = 0; wi < (fan >> 6u); wi++)
                {
                    uint64_t cm = b.r(node + OFF_CM[t] + wi), vm = b.r(node + OFF_VM[t] + wi);
                    b.wr(node + OFF_VM[t] + wi, vm & ~cm);
                }
            }
            // grid_build_finalize.slang:29-74
            if (A != C && nc > 0u)
            {
                uint32_t total_new = scan[nc - 1u];
                set_tree_count(b, C, tree_count(b, C) + total_new);
            }
            if (A == C)
                last_value_scan = scan;
        }
    }

    // grid_build_blindmetadata.slang:29-203 + grid_build_compute_bbox.slang:30-167
    void finalize(bool density_values)
    {
        uint32_t nc = node_count();
        uint32_t value_count = last_value_scan.empty() ? 2u : last_value_scan[nc - 1u];
        uint64_t end = end_addr();
        uint64_t meta0 = end, meta1 = end + BLINDMETADATA_SIZE;
        uint64_t bboxes = meta1 + BLINDMETADATA_SIZE;
        uint64_t values = bboxes + 32u;
        uint64_t fin = values + 4ull * value_count;
        b.ensure_bytes(fin);
        b.u64(meta0 + 0, bboxes - meta0);
        b.u64(meta0 + 8, 2u);
        b.u32(meta0 + 16, 12u);
        b.u32(meta0 + 20, 0u);
        b.u32(meta0 + 24, 0u);
        b.u32(meta0 + 28, GRID_TYPE_VEC3F);
        b.u32(meta0 + 32, 0u);
        b.u64(meta1 + 0, values - meta1);
        b.u64(meta1 + 8, value_count);
        b.u32(meta1 + 16, 4u);
        b.u32(meta1 + 20, 0u);
        b.u32(meta1 + 24, 0u);
        b.u32(meta1 + 28, density_values ? GRID_TYPE_FLOAT : GRID_TYPE_RGBA8);
        b.u32(meta1 + 32, 0u);
        b.u64(640, meta0);
        b.u32(648, 2u);
        b.u64(32, fin);

        // root keys
        const uint32_t root = ROOT_IDX64;
        for (uint32_t n = 0; n < 32768u; n++)
        {
            uint64_t bit = 1ull << (n & 63u);
            bool is_child = (b.r(root + OFF_CM[ROOT] + (n >> 6u)) & bit) != 0ull;
            bool is_value = (b.r(root + OFF_VM[ROOT] + (n >> 6u)) & bit) != 0ull;
            if (is_child || is_value)
                b.wr(root + OFF_KM[ROOT] + (n >> 6u), b.r(root + OFF_KM[ROOT] + (n >> 6u)) | bit);
        }
        write_prefix(b, root, ROOT, OFF_KM[ROOT], OFF_KMP[ROOT]);
        uint32_t child_count = (uint32_t)b.r(root + OFF_CMP[ROOT]) & 0xFFFFu;
        for (uint32_t n = 0; n < 32768u; n++)
        {
            bool active = false;
            uint32_t key_idx = mask_n_to_idx(b, root + OFF_KMP[ROOT], ROOT, n, &active);
            if (!active)
                continue;
            Coord ijk;
            ijk.x = (int32_t)(((n >> 10u) & 0x1Fu) << 12u);
            ijk.y = (int32_t)(((n >> 5u) & 0x1Fu) << 12u);
            ijk.z = (int32_t)((n & 0x1Fu) << 12u);
            const int32_t half = (1 << 17) >> 1;
            ijk.x = ijk.x >= half ? ijk.x - half - half : ijk.x;
            ijk.y = ijk.y >= half ? ijk.y - half - half : ijk.y;
            ijk.z = ijk.z >= half ? ijk.z - half - half : ijk.z;
            b.wr(root + OFF_CHILDREN[ROOT] + child_count + key_idx, root_key(ijk));
        }

        // bboxes, leaf -> root
        std::vector<int32_t> bb((size_t)nc * 6u, 0);
        const int32_t IMAX = 2147483647, IMIN = -2147483647;
        for (int level = 3; level >= 0; level--)
        {
            for (uint32_t i = 0; i < nc; i++)
            {
                if ((int)node_type[i] != level)
                    continue;
                uint32_t node = node_addr[i] >> 3u, t = node_type[i];
                int32_t mn[3] = { IMAX, IMAX, IMAX }, mx[3] = { IMIN, IMIN, IMIN };
                uint32_t fb = FANOUT_BITS[t], tb = TILEDIM_BITS[t], m = (1u << fb) - 1u;
                for (uint32_t n = 0; n < FANOUT_1D[t]; n++)
                {
                    int32_t o[3] = { (int32_t)(((n >> (fb + fb)) & m) << tb), (int32_t)(((n >> fb) & m) << tb),
                                     (int32_t)((n & m) << tb) };
                    if (t == LEAF)
                    {
                        if (!(b.r(node + OFF_VM[t] + (n >> 6u)) & (1ull << (n & 63u))))
                            continue;
                        for (int a = 0; a < 3; a++)
                        {
                            mn[a] = std::min(mn[a], o[a]);
                            mx[a] = std::max(mx[a], o[a]);
                        }
                        continue;
                    }
                    bool active = false;
                    uint32_t ci = mask_n_to_idx(b, node + OFF_CMP[t], t, n, &active);
                    uint32_t child = active ? (uint32_t)(b.r(node + OFF_CHILDREN[t] + ci) >> 3u) : 0u;
                    if (child == 0u)
                        continue;
                    uint32_t cidx = (uint32_t)b.r(node + child + OFF_NODE_IDX);
                    if (t == ROOT)
                    {
                        const int32_t off = (int32_t)((((1u << fb) >> 1u)) << tb);
                        const int32_t rng = (int32_t)((((1u << fb) >> 1u) * 2u) << tb);
                        for (int a = 0; a < 3; a++)
                            if (o[a] >= off)
                                o[a] -= rng;
                    }
                    for (int a = 0; a < 3; a++)
                    {
                        mn[a] = std::min(mn[a], bb[6u * cidx + a] + o[a]);
                        mx[a] = std::max(mx[a], bb[6u * cidx + 3 + a] + o[a]);
                    }
                }
                for (int a = 0; a < 3; a++)
                {
                    bb[6u * i + a] = mn[a];
                    bb[6u * i + 3 + a] = mx[a];
                }
                if (t == ROOT)
                {
                    for (int a = 0; a < 3; a++)
                    {
                        b.u32(bboxes + 4u * a, (uint32_t)mn[a]);
                        b.u32(bboxes + 12u + 4u * a, (uint32_t)mx[a]);
                    }
                }
            }
        }
    }

    uint64_t values_addr() const
    {
        uint64_t meta1 = b.gu64(640) + BLINDMETADATA_SIZE;
        return meta1 + b.gu64(meta1);
    }

    // value index of an active voxel (gaussian_frag_color.slang:219-236); ~0 if inactive
    uint64_t voxel_value_idx(Coord ijk) const
    {
        uint32_t node, t, n;
        find_node(b, ijk, &node, &t, &n);
        if (t != LEAF)
            return ~0ull;
        if (!(b.r(node + OFF_VM[LEAF] + (n >> 6u)) & (1ull << (n & 63u))))
            return ~0ull;
        bool active = false;
        uint32_t local = mask_n_to_idx(b, node + OFF_VMP[LEAF], LEAF, n, &active);
        return b.r(node + OFF_VALUE_IDX) + local;
    }
};

// ------------------------------------------------------------------------------------------
// Per-Gaussian shader quantities (gaussian_prim.slang:48-68, gaussian_frag_alloc.slang:62-121)
struct Prim
{
    float pos[3];
    float qx, qy, qz, qw;
    float scale[3]; // floored
    float opacity;
    float local_scale; // r = sqrt(2 log(o / 0.01))
    bool valid;
};

Prim load_prim(uint64_t g, const float* means, const float* quats, const float* scales, const float* opac,
               float voxel_size)
{
    Prim p;
    p.pos[0] = means[3 * g + 0];
    p.pos[1] = means[3 * g + 1];
    p.pos[2] = means[3 * g + 2];
    p.qx = quats[4 * g + 1];
    p.qy = quats[4 * g + 2];
    p.qz = quats[4 * g + 3];
    p.qw = quats[4 * g + 0];
    p.opacity = opac[g];
    p.local_scale = sqrtf(2.f * cr_logf(p.opacity / 0.01f));
    const float min_scale = 1.f / 8.f;
    float floor_s = min_scale * voxel_size;
    for (int a = 0; a < 3; a++)
    {
        float s = scales[3 * g + a];
        p.scale[a] = std::max(s, floor_s); // HLSL max: s if s > floor
    }
    // contract: a Gaussian whose radius is not a positive finite number (opacity <= 0.01 or
    // non-finite inputs) is dropped (footprint 0).
    p.valid = std::isfinite(p.local_scale) && p.local_scale > 0.f;
    return p;
}

inline float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    float s = ax * bx;
    s = s + ay * by;
    s = s + az * bz;
    return s;
}

// gaussian_prim.slang:70-139
uint64_t prim_aabb(const Prim& p, float voxel_size_inv, int32_t mn[3], int32_t mx[3])
{
    if (!p.valid)
    {
        mn[0] = mn[1] = mn[2] = 0;
        mx[0] = mx[1] = mx[2] = -1;
        return 0ull;
    }
    const float x = p.qx, y = p.qy, z = p.qz, w = p.qw;
    float r0x = 1.f - 2.f * (y * y + z * z), r0y = 2.f * (x * y - z * w), r0z = 2.f * (x * z + y * w);
    float r1x = 2.f * (x * y + z * w), r1y = 1.f - 2.f * (x * x + z * z), r1z = 2.f * (y * z - x * w);
    float r2x = 2.f * (x * z - y * w), r2y = 2.f * (y * z + x * w), r2z = 1.f - 2.f * (x * x + y * y);
    for (uint32_t v = 0; v < 8u; v++)
    {
        float c[3] = { (v & 1u) ? 1.f : -1.f, (v & 2u) ? 1.f : -1.f, (v & 4u) ? 1.f : -1.f };
        for (int a = 0; a < 3; a++)
        {
            c[a] = c[a] * p.scale[a];
            c[a] = c[a] * p.local_scale;
        }
        float q[3] = { dot3(c[0], c[1], c[2], r0x, r0y, r0z), dot3(c[0], c[1], c[2], r1x, r1y, r1z),
                       dot3(c[0], c[1], c[2], r2x, r2y, r2z) };
        int32_t ijk[3];
        for (int a = 0; a < 3; a++)
        {
            q[a] = q[a] + p.pos[a];
            ijk[a] = (int32_t)floorf(q[a] * voxel_size_inv);
        }
        for (int a = 0; a < 3; a++)
        {
            if (v == 0u)
            {
                mn[a] = ijk[a];
                mx[a] = ijk[a];
            }
            else
            {
                mn[a] = std::min(mn[a], ijk[a]);
                mx[a] = std::max(mx[a], ijk[a]);
            }
        }
    }
    return (uint64_t)(mx[0] + 1 - mn[0]) * (uint64_t)(mx[1] + 1 - mn[1]) * (uint64_t)(mx[2] + 1 - mn[2]);
}

// gaussian_frag_alloc.slang:79-162: voxel box corners in the Gaussian frame; returns
// whether the voxel is kept (dist2 > 0), and the local box for the color pass.
bool voxel_test(const Prim& p, const int32_t ijk[3], float voxel_size, float vmin[3], float vmax[3])
{
    const float x = p.qx, y = p.qy, z = p.qz, w = p.qw;
    float i0x = 1.f - 2.f * (y * y + z * z), i0y = 2.f * (x * y + z * w), i0z = 2.f * (x * z - y * w);
    float i1x = 2.f * (x * y - z * w), i1y = 1.f - 2.f * (x * x + z * z), i1z = 2.f * (y * z + x * w);
    float i2x = 2.f * (x * z + y * w), i2y = 2.f * (y * z - x * w), i2z = 1.f - 2.f * (x * x + y * y);
    float inv_r = 1.f / p.local_scale;
    float inv_s[3] = { 1.f / p.scale[0], 1.f / p.scale[1], 1.f / p.scale[2] };
    for (uint32_t v = 0; v < 8u; v++)
    {
        float vp[3] = { (float)ijk[0], (float)ijk[1], (float)ijk[2] };
        if (v & 1u)
            vp[0] += 1.f;
        if (v & 2u)
            vp[1] += 1.f;
        if (v & 4u)
            vp[2] += 1.f;
        float lp[3];
        for (int a = 0; a < 3; a++)
        {
            lp[a] = vp[a] * voxel_size;
            lp[a] = lp[a] - p.pos[a];
        }
        float q[3] = { dot3(lp[0], lp[1], lp[2], i0x, i0y, i0z), dot3(lp[0], lp[1], lp[2], i1x, i1y, i1z),
                       dot3(lp[0], lp[1], lp[2], i2x, i2y, i2z) };
        for (int a = 0; a < 3; a++)
        {
            q[a] = q[a] * inv_r;
            q[a] = q[a] * inv_s[a];
        }
        for (int a = 0; a < 3; a++)
        {
            if (v == 0u)
            {
                vmin[a] = q[a];
                vmax[a] = q[a];
            }
            else
            {
                vmin[a] = fminf(vmin[a], q[a]);
                vmax[a] = fmaxf(vmax[a], q[a]);
            }
        }
    }
    float d2 = 1.f;
    for (int a = 0; a < 3; a++)
    {
        if (0.f < vmin[a])
            d2 = d2 - vmin[a] * vmin[a];
        else if (0.f > vmax[a])
            d2 = d2 - vmax[a] * vmax[a];
    }
    return !(d2 <= 0.f);
}

inline uint32_t u8_of(float v)
{
    return (uint32_t)(255.f * std::max(0.f, std::min(1.f, v)));
}

// gaussian_frag_color.slang:184-212 -- packed RGBA8 of one pair
uint32_t pair_rgba(const Prim& p, const float* color, const float vmin[3], const float vmax[3], float voxel_size,
                   float cull_scale)
{
    float a = 0.f;
    const float third = 1.f / 3.f;
    for (uint32_t s = 0; s < 64u; s++)
    {
        float wv[3] = { third * (float)((s >> 4) & 3u), third * (float)((s >> 2) & 3u), third * (float)(s & 3u) };
        float lp[3];
        for (int k = 0; k < 3; k++)
        {
            float ow = 1.f - wv[k];
            lp[k] = wv[k] * vmax[k] + ow * vmin[k];
        }
        float d = dot3(lp[0], lp[1], lp[2], lp[0], lp[1], lp[2]);
        float e = p.opacity * cr_expf(-(p.local_scale * d));
        a = std::max(a, e);
    }
    float scale1d = sqrtf(dot3(p.scale[0], p.scale[1], p.scale[2], p.scale[0], p.scale[1], p.scale[2]));
    float vrs = (0.1f * scale1d) / voxel_size;
    a = 1.f - cr_powf(1.f - a, 1.f / vrs);
    const float default_cull = 0.05f;
    float cs = cull_scale == 0.f ? default_cull : cull_scale;
    if (scale1d > cs)
        a = 0.f;
    return u8_of(color[0]) | (u8_of(color[1]) << 8) | (u8_of(color[2]) << 16) | (u8_of(a) << 24);
}

struct PairRec
{
    uint64_t slot; // fragment slot within the prim batch
    uint32_t gid;
    int32_t ijk[3];
    uint32_t rgba;
};

// fanout_*.slang + gaussian_frag_alloc.slang: surviving fragments of one prim batch in slot order
void batch_pairs(uint64_t b0, uint64_t b1, const float* means, const float* quats, const float* scales,
                 const float* colors, const float* opac, float voxel_size, float cull_scale, bool want_rgba,
                 int nthreads, std::vector<PairRec>& out, uint64_t* frag_total)
{
    float vs_inv = 1.f / voxel_size;
    uint64_t np = b1 - b0;
    std::vector<uint64_t> fp(np), excl(np);
    std::vector<int32_t> bbox(np * 6);
    std::vector<Prim> prims(np);
    for (uint64_t i = 0; i < np; i++)
    {
        prims[i] = load_prim(b0 + i, means, quats, scales, opac, voxel_size);
        fp[i] = prim_aabb(prims[i], vs_inv, &bbox[6 * i], &bbox[6 * i + 3]);
    }
    uint64_t acc = 0;
    for (uint64_t i = 0; i < np; i++)
    {
        excl[i] = acc;
        acc += fp[i];
    }
    *frag_total = acc;
    int T = std::max(1, nthreads);
    std::vector<std::vector<PairRec>> parts(T);
    auto work = [&](int tid) {
        uint64_t lo = np * tid / T, hi = np * (tid + 1) / T;
        for (uint64_t i = lo; i < hi; i++)
        {
            if (fp[i] == 0ull)
                continue;
            const int32_t* mn = &bbox[6 * i];
            const int32_t* mx = &bbox[6 * i + 3];
            int64_t rx = mx[0] + 1 - mn[0], ry = mx[1] + 1 - mn[1], rz = mx[2] + 1 - mn[2];
            for (int64_t ri = 0; ri < rx * ry * rz; ri++)
            {
                int32_t ijk[3] = { (int32_t)(ri / (rz * ry)) + mn[0], (int32_t)((ri / rz) % ry) + mn[1],
                                   (int32_t)(ri % rz) + mn[2] };
                float vmin[3], vmax[3];
                if (!voxel_test(prims[i], ijk, voxel_size, vmin, vmax))
                    continue;
                PairRec r;
                r.slot = excl[i] + (uint64_t)ri;
                r.gid = (uint32_t)(b0 + i);
                r.ijk[0] = ijk[0];
                r.ijk[1] = ijk[1];
                r.ijk[2] = ijk[2];
                r.rgba = want_rgba ? pair_rgba(prims[i], &colors[3 * (b0 + i)], vmin, vmax, voxel_size, cull_scale) : 0u;
                parts[tid].push_back(r);
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++)
        th.emplace_back(work, t);
    work(0);
    for (auto& t : th)
        t.join();
    for (int t = 0; t < T; t++)
        out.insert(out.end(), parts[t].begin(), parts[t].end());
}

int resolve_threads(int nthreads)
{
    if (nthreads > 0)
        return nthreads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// FP64 density quantities of one Gaussian from preprocessed fp32 parameters
struct PrimD
{
    double mu[3];
    double R[3][3]; // rows rot0..rot2 (gaussian_prim.slang:82-89)
    double s[3];
    double o;
};

PrimD load_prim_d(uint64_t g, const float* means, const float* quats, const float* scales, const float* opac,
                  float voxel_size)
{
    PrimD p;
    double w = quats[4 * g + 0], x = quats[4 * g + 1], y = quats[4 * g + 2], z = quats[4 * g + 3];
    p.R[0][0] = 1 - 2 * (y * y + z * z);
    p.R[0][1] = 2 * (x * y - z * w);
    p.R[0][2] = 2 * (x * z + y * w);
    p.R[1][0] = 2 * (x * y + z * w);
    p.R[1][1] = 1 - 2 * (x * x + z * z);
    p.R[1][2] = 2 * (y * z - x * w);
    p.R[2][0] = 2 * (x * z - y * w);
    p.R[2][1] = 2 * (y * z + x * w);
    p.R[2][2] = 1 - 2 * (x * x + y * y);
    double fl = (double)(0.125f * voxel_size);
    for (int a = 0; a < 3; a++)
    {
        p.mu[a] = means[3 * g + a];
        p.s[a] = std::max((double)scales[3 * g + a], fl);
    }
    p.o = opac[g];
    return p;
}

double density_contrib(const PrimD& p, const int32_t ijk[3], double vs)
{
    double d[3], m = 0.0;
    for (int a = 0; a < 3; a++)
        d[a] = ((double)ijk[a] + 0.5) * vs - p.mu[a];
    for (int a = 0; a < 3; a++)
    {
        double u = p.R[0][a] * d[0] + p.R[1][a] * d[1] + p.R[2][a] * d[2];
        double q = u / p.s[a];
        m += q * q;
    }
    return p.o * exp(-0.5 * m);
}

} // namespace

// ==========================================================================================
extern "C" {

// RasterUtils.cpp:31-136
void vdbo_preprocess(uint64_t n, float* means, float* opacities, float* quats, float* scales, const float* sh0,
                     float* colors_out)
{
    (void)means;
    for (uint64_t i = 0; i < n; i++)
    {
        float x = quats[4 * i + 1], y = quats[4 * i + 2], z = quats[4 * i + 3], w = quats[4 * i + 0];
        float sum = x * x;
        sum = sum + y * y;
        sum = sum + z * z;
        sum = sum + w * w;
        float magn_inv = 1.f / sqrtf(sum);
        quats[4 * i + 1] = x * magn_inv;
        quats[4 * i + 2] = y * magn_inv;
        quats[4 * i + 3] = z * magn_inv;
        quats[4 * i + 0] = w * magn_inv;
    }
    for (uint64_t i = 0; i < 3 * n; i++)
        scales[i] = cr_expf(scales[i]);
    const float c0 = 0.28209479177387814f;
    for (uint64_t i = 0; i < 3 * n; i++)
        colors_out[i] = c0 * sh0[i] + 0.5f;
    for (uint64_t i = 0; i < n; i++)
        opacities[i] = 1.f / (1.f + cr_expf(-opacities[i]));
}

void vdbo_prim_aabb(uint64_t n, const float* means, const float* quats, const float* scales, const float* opacities,
                    float voxel_size, int32_t* bbox6_out, uint64_t* footprint_out)
{
    float vs_inv = 1.f / voxel_size;
    for (uint64_t g = 0; g < n; g++)
    {
        Prim p = load_prim(g, means, quats, scales, opacities, voxel_size);
        footprint_out[g] = prim_aabb(p, vs_inv, &bbox6_out[6 * g], &bbox6_out[6 * g + 3]);
    }
}

uint64_t vdbo_pairs(uint64_t n, const float* means, const float* quats, const float* scales, const float* colors,
                    const float* opacities, float voxel_size, float cull_scale, uint32_t* gid_out, int32_t* ijk_out,
                    uint32_t* rgba_out, uint64_t cap, int nthreads)
{
    std::vector<PairRec> pairs;
    uint64_t total = 0;
    const uint64_t PB = 262144u;
    for (uint64_t b0 = 0; b0 < n; b0 += PB)
    {
        uint64_t ft = 0;
        std::vector<PairRec> part;
        batch_pairs(b0, std::min(n, b0 + PB), means, quats, scales, colors, opacities, voxel_size, cull_scale,
                    colors != nullptr && rgba_out != nullptr, resolve_threads(nthreads), part, &ft);
        for (const PairRec& r : part)
        {
            if (total < cap)
            {
                if (gid_out)
                    gid_out[total] = r.gid;
                if (ijk_out)
                {
                    ijk_out[3 * total + 0] = r.ijk[0];
                    ijk_out[3 * total + 1] = r.ijk[1];
                    ijk_out[3 * total + 2] = r.ijk[2];
                }
                if (rgba_out)
                    rgba_out[total] = r.rgba;
            }
            total++;
        }
    }
    return total;
}

uint64_t vdbo_raster_node2(uint64_t n, const float* means, const float* quats, const float* scales,
                           const float* colors, const float* opacities, float voxel_size, float cull_scale,
                           uint32_t prim_batch_size, uint32_t ijk_batch_size, uint32_t max_node_count, int value_mode,
                           uint8_t* out, uint64_t cap, int nthreads, vdbo_stats_t* stats)
{
    if (prim_batch_size == 0u)
        prim_batch_size = 262144u;
    if (ijk_batch_size == 0u)
        ijk_batch_size = 8388608u;
    if (max_node_count == 0u)
        max_node_count = 2u * 1024u * 1024u;
    int T = resolve_threads(nthreads);

    Builder bld;
    bld.max_node_count = max_node_count;
    empty_grid(bld.b, voxel_size);
    bld.node_addr.assign(1, (uint32_t)ROOT_ADDR);
    bld.node_type.assign(1, (uint32_t)ROOT);

    std::vector<PairRec> all_pairs;
    uint64_t frag_total_all = 0, windows = 0;
    bool ran_any_window = false;
    // Raster.cpp:481-503 -- topology pass, prim batches x fragment windows
    for (uint64_t b0 = 0; b0 < n; b0 += prim_batch_size)
    {
        uint64_t b1 = std::min<uint64_t>(n, b0 + prim_batch_size);
        std::vector<PairRec> part;
        uint64_t ft = 0;
        batch_pairs(b0, b1, means, quats, scales, colors, opacities, voxel_size, cull_scale, value_mode == 0, T,
                    part, &ft);
        frag_total_all += ft;
        uint64_t nwin = (ft + ijk_batch_size - 1) / ijk_batch_size;
        size_t pi = 0;
        for (uint64_t wdx = 0; wdx < nwin; wdx++)
        {
            uint64_t hi = (wdx + 1) * (uint64_t)ijk_batch_size;
            std::vector<Coord> ijks;
            while (pi < part.size() && part[pi].slot < hi)
            {
                ijks.push_back(Coord{ part[pi].ijk[0], part[pi].ijk[1], part[pi].ijk[2] });
                pi++;
            }
            // OR-ing a mask bit is idempotent and the "new child" flag depends only on the
            // state before the pass, so a window's voxel multiset acts as its voxel set.
            std::sort(ijks.begin(), ijks.end(), [](const Coord& a, const Coord& b) {
                if (a.x != b.x)
                    return a.x < b.x;
                if (a.y != b.y)
                    return a.y < b.y;
                return a.z < b.z;
            });
            ijks.erase(std::unique(ijks.begin(), ijks.end(),
                                   [](const Coord& a, const Coord& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }),
                       ijks.end());
            bld.build_window(ijks);
            ran_any_window = true;
            windows++;
            if (bld.overflow)
                return 0ull;
        }
        all_pairs.insert(all_pairs.end(), part.begin(), part.end());
    }
    if (!ran_any_window)
    {
        std::vector<Coord> none;
        bld.build_window(none);
    }
    // GridBuild.cpp:508-650
    bld.finalize(value_mode != 0);
    uint64_t values = bld.values_addr();
    uint64_t value_count = bld.b.gu64(bld.b.gu64(640) + BLINDMETADATA_SIZE + 8);
    // Raster.cpp:514-535 -- value pass
    if (value_mode == 0)
    {
        for (const PairRec& r : all_pairs)
        {
            uint64_t vi = bld.voxel_value_idx(Coord{ r.ijk[0], r.ijk[1], r.ijk[2] });
            if (vi == ~0ull)
                continue;
            uint64_t a = values + 4ull * vi;
            uint32_t cur = bld.b.gu32(a);
            if (r.rgba > cur)
                bld.b.u32(a, r.rgba);
        }
    }
    else
    {
        // density: sum over pairs in ascending gaussian id (all_pairs is gid-major)
        std::vector<double> acc(value_count, 0.0);
        std::vector<PrimD> cache;
        for (const PairRec& r : all_pairs)
        {
            uint64_t vi = bld.voxel_value_idx(Coord{ r.ijk[0], r.ijk[1], r.ijk[2] });
            if (vi == ~0ull)
                continue;
            PrimD p = load_prim_d(r.gid, means, quats, scales, opacities, voxel_size);
            acc[vi] += density_contrib(p, r.ijk, (double)voxel_size);
        }
        for (uint64_t vi = 2; vi < value_count; vi++)
            bld.b.f32(values + 4ull * vi, (float)acc[vi]);
    }
    uint64_t grid_size = bld.b.gu64(32);
    if (stats)
    {
        stats->fragment_count = frag_total_all;
        stats->pair_count = all_pairs.size();
        stats->window_count = windows;
        stats->grid_size = grid_size;
        stats->upper_count = tree_count(bld.b, UPPER);
        stats->lower_count = tree_count(bld.b, LOWER);
        stats->leaf_count = tree_count(bld.b, LEAF);
        stats->value_count = value_count;
        uint64_t bb = bld.b.gu64(640) + bld.b.gu64(bld.b.gu64(640));
        for (int a = 0; a < 3; a++)
        {
            stats->bbox_min[a] = (int32_t)bld.b.gu32(bb + 4 * a);
            stats->bbox_max[a] = (int32_t)bld.b.gu32(bb + 12 + 4 * a);
        }
    }
    if (out && cap >= grid_size)
    {
        bld.b.ensure_bytes(grid_size);
        memcpy(out, bld.b.w.data(), grid_size);
    }
    return grid_size;
}

uint64_t vdbo_density_f64(uint64_t pair_count, const uint32_t* gid, const int32_t* ijk, const float* means,
                          const float* quats, const float* scales, const float* opacities, float voxel_size,
                          int32_t* voxel_ijk_out, double* density_out)
{
    std::vector<uint64_t> order(pair_count);
    for (uint64_t i = 0; i < pair_count; i++)
        order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
        for (int k = 0; k < 3; k++)
            if (ijk[3 * a + k] != ijk[3 * b + k])
                return ijk[3 * a + k] < ijk[3 * b + k];
        return gid[a] < gid[b];
    });
    uint64_t nv = 0;
    for (uint64_t idx = 0; idx < pair_count; idx++)
    {
        uint64_t i = order[idx];
        bool head = idx == 0 || memcmp(&ijk[3 * i], &ijk[3 * order[idx - 1]], 12) != 0;
        if (head)
        {
            voxel_ijk_out[3 * nv + 0] = ijk[3 * i + 0];
            voxel_ijk_out[3 * nv + 1] = ijk[3 * i + 1];
            voxel_ijk_out[3 * nv + 2] = ijk[3 * i + 2];
            density_out[nv] = 0.0;
            nv++;
        }
        PrimD p = load_prim_d(gid[i], means, quats, scales, opacities, voxel_size);
        density_out[nv - 1] += density_contrib(p, &ijk[3 * i], (double)voxel_size);
    }
    return nv;
}

void vdbo_density_grad_f64(uint64_t n, uint64_t pair_count, const uint32_t* gid, const int32_t* ijk,
                           const float* means, const float* log_scales, const float* quats_raw,
                           const float* opacity_logits, float voxel_size, const double* residual_per_pair,
                           double* grad_means, double* grad_log_scales, double* grad_quats,
                           double* grad_opacity_logits)
{
    for (uint64_t i = 0; i < 3 * n; i++)
    {
        grad_means[i] = 0.0;
        grad_log_scales[i] = 0.0;
    }
    for (uint64_t i = 0; i < 4 * n; i++)
        grad_quats[i] = 0.0;
    for (uint64_t i = 0; i < n; i++)
        grad_opacity_logits[i] = 0.0;
    const double vs = voxel_size;
    const double fl = (double)(0.125f * voxel_size);
    for (uint64_t pi = 0; pi < pair_count; pi++)
    {
        uint64_t g = gid[pi];
        double qw = quats_raw[4 * g + 0], qx = quats_raw[4 * g + 1], qy = quats_raw[4 * g + 2],
               qz = quats_raw[4 * g + 3];
        double qn = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
        double w = qw / qn, x = qx / qn, y = qy / qn, z = qz / qn;
        double R[3][3] = { { 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w) },
                           { 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w) },
                           { 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y) } };
        double s[3];
        bool floored[3];
        for (int a = 0; a < 3; a++)
        {
            double e = exp((double)log_scales[3 * g + a]);
            floored[a] = !(e > fl);
            s[a] = floored[a] ? fl : e;
        }
        double th = opacity_logits[g];
        double o = 1.0 / (1.0 + exp(-th));
        double d[3], u[3], m = 0.0;
        for (int a = 0; a < 3; a++)
            d[a] = ((double)ijk[3 * pi + a] + 0.5) * vs - (double)means[3 * g + a];
        for (int a = 0; a < 3; a++)
        {
            u[a] = R[0][a] * d[0] + R[1][a] * d[1] + R[2][a] * d[2];
            m += (u[a] / s[a]) * (u[a] / s[a]);
        }
        double e = exp(-0.5 * m);
        double res = residual_per_pair[pi];
        double contrib = o * e;
        grad_opacity_logits[g] += res * e * o * (1.0 - o);
        double dm = -0.5 * contrib * res; // dL/dm
        double gu[3];
        for (int a = 0; a < 3; a++)
        {
            gu[a] = dm * 2.0 * u[a] / (s[a] * s[a]);
            if (!floored[a])
                grad_log_scales[3 * g + a] += dm * (-2.0 * u[a] * u[a] / (s[a] * s[a]));
        }
        // u_a = sum_b R[b][a] d_b ; dd/dmu = -I
        for (int b = 0; b < 3; b++)
        {
            double gd = R[b][0] * gu[0] + R[b][1] * gu[1] + R[b][2] * gu[2];
            grad_means[3 * g + b] -= gd;
        }
        // dL/dR[b][a] = gu[a] * d[b]
        double G[3][3];
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++)
                G[b][a] = gu[a] * d[b];
        double dRw[3][3] = { { 0, -2 * z, 2 * y }, { 2 * z, 0, -2 * x }, { -2 * y, 2 * x, 0 } };
        double dRx[3][3] = { { 0, 2 * y, 2 * z }, { 2 * y, -4 * x, -2 * w }, { 2 * z, 2 * w, -4 * x } };
        double dRy[3][3] = { { -4 * y, 2 * x, 2 * w }, { 2 * x, 0, 2 * z }, { -2 * w, 2 * z, -4 * y } };
        double dRz[3][3] = { { -4 * z, -2 * w, 2 * x }, { 2 * w, -4 * z, 2 * y }, { 2 * x, 2 * y, 0 } };
        double gq[4] = { 0, 0, 0, 0 }; // w, x, y, z of the normalised quaternion
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++)
            {
                gq[0] += G[b][a] * dRw[b][a];
                gq[1] += G[b][a] * dRx[b][a];
                gq[2] += G[b][a] * dRy[b][a];
                gq[3] += G[b][a] * dRz[b][a];
            }
        double qh[4] = { w, x, y, z };
        double dot = gq[0] * qh[0] + gq[1] * qh[1] + gq[2] * qh[2] + gq[3] * qh[3];
        for (int k = 0; k < 4; k++)
            grad_quats[4 * g + k] += (gq[k] - dot * qh[k]) / qn;
    }
}

uint64_t vdbo_node2_active_voxels(const uint8_t* grid, uint64_t grid_size, int32_t* ijk_out, uint64_t* value_idx_out,
                                  uint64_t cap)
{
    Buf b;
    b.w.assign((grid_size + 7) / 8, 0ull);
    memcpy(b.w.data(), grid, grid_size);
    uint64_t count = 0;
    const uint32_t root = ROOT_IDX64;
    for (uint32_t rn = 0; rn < 32768u; rn++)
    {
        bool a0 = false;
        uint32_t c0 = mask_n_to_idx(b, root + OFF_CMP[ROOT], ROOT, rn, &a0);
        if (!a0)
            continue;
        uint32_t up = root + (uint32_t)(b.r(root + OFF_CHILDREN[ROOT] + c0) >> 3u);
        int32_t rx = (int32_t)(((rn >> 10) & 31u) << 12), ry = (int32_t)(((rn >> 5) & 31u) << 12),
                rz = (int32_t)((rn & 31u) << 12);
        if (rx >= 65536)
            rx -= 131072;
        if (ry >= 65536)
            ry -= 131072;
        if (rz >= 65536)
            rz -= 131072;
        for (uint32_t un = 0; un < 32768u; un++)
        {
            bool a1 = false;
            uint32_t c1 = mask_n_to_idx(b, up + OFF_CMP[UPPER], UPPER, un, &a1);
            if (!a1)
                continue;
            uint32_t lo = up + (uint32_t)(b.r(up + OFF_CHILDREN[UPPER] + c1) >> 3u);
            int32_t ux = rx + (int32_t)(((un >> 10) & 31u) << 7), uy = ry + (int32_t)(((un >> 5) & 31u) << 7),
                    uz = rz + (int32_t)((un & 31u) << 7);
            for (uint32_t ln = 0; ln < 4096u; ln++)
            {
                bool a2 = false;
                uint32_t c2 = mask_n_to_idx(b, lo + OFF_CMP[LOWER], LOWER, ln, &a2);
                if (!a2)
                    continue;
                uint32_t lf = lo + (uint32_t)(b.r(lo + OFF_CHILDREN[LOWER] + c2) >> 3u);
                int32_t lx = ux + (int32_t)(((ln >> 8) & 15u) << 3), ly = uy + (int32_t)(((ln >> 4) & 15u) << 3),
                        lz = uz + (int32_t)((ln & 15u) << 3);
                uint64_t vbase = b.r(lf + OFF_VALUE_IDX);
                for (uint32_t vn = 0; vn < 512u; vn++)
                {
                    bool a3 = false;
                    uint32_t li = mask_n_to_idx(b, lf + OFF_VMP[LEAF], LEAF, vn, &a3);
                    if (!a3)
                        continue;
                    if (count < cap)
                    {
                        ijk_out[3 * count + 0] = lx + (int32_t)((vn >> 6) & 7u);
                        ijk_out[3 * count + 1] = ly + (int32_t)((vn >> 3) & 7u);
                        ijk_out[3 * count + 2] = lz + (int32_t)(vn & 7u);
                        value_idx_out[count] = vbase + li;
                    }
                    count++;
                }
            }
        }
    }
    return count;
}


// putil/Convert.h:40-313 (pnanovdb_node2_convert_to_grid_type, dst type ONINDEX) -- node2 grid to a
// NanoVDB OnIndex grid. Five passes over the node2 tree in slot order (root tiles, uppers,
// lowers, leaves, values), then the value blind data. The OnIndex node layout (struct offsets of
// RootData/Tile/InternalData/LeafData<ValueOnIndex>) comes from the public NanoVDB format, which
// PNanoVDB.h encodes and which is not part of the reference tree.
uint64_t vdbo_node2_to_onindex(const uint8_t* grid, uint64_t grid_size, uint8_t* out, uint64_t cap)
{
    const uint32_t ROOT_SZ = 96u, TILE_SZ = 32u, UPPER_SZ = 270400u, LOWER_SZ = 33856u, LEAF_SZ = 96u;
    const uint32_t UPPER_VM = 32u, UPPER_CM = 4128u, UPPER_TABLE = 8256u;
    const uint32_t LOWER_VM = 32u, LOWER_CM = 544u, LOWER_TABLE = 1088u;
    const uint32_t LEAF_VM = 16u, LEAF_TABLE = 80u;
    const uint32_t GRID_TYPE_ONINDEX = 20u;
    Buf b;
    b.w.assign((grid_size + 7) / 8, 0ull);
    memcpy(b.w.data(), grid, grid_size);
    Buf d;
    const uint32_t root = ROOT_IDX64;
    auto get_key = [&](uint32_t n) -> uint64_t {
        bool act = false;
        uint32_t ki = mask_n_to_idx(b, root + OFF_KMP[ROOT], ROOT, n, &act);
        if (!act)
            return END_KEY;
        uint32_t child_count = (uint32_t)(b.r(root + OFF_CMP[ROOT]) & 0xFFFFu);
        return b.r(root + OFF_CHILDREN[ROOT] + child_count + ki);
    };
    auto get_child = [&](uint32_t node, uint32_t t, uint32_t n) -> uint32_t {
        bool act = false;
        uint32_t ci = mask_n_to_idx(b, node + OFF_CMP[t], t, n, &act);
        if (!act)
            return 0u;
        return node + (uint32_t)(b.r(node + OFF_CHILDREN[t] + ci) >> 3u);
    };
    auto value_bit = [&](uint32_t node, uint32_t t, uint32_t n) -> bool {
        return (b.r(node + OFF_VM[t] + (n >> 6u)) >> (n & 63u)) & 1ull;
    };
    auto copy_bytes = [&](uint64_t dst, uint64_t src_byte, uint32_t n) {
        std::vector<uint8_t> tmp(n);
        b.get(src_byte, tmp.data(), n);
        d.put(dst, tmp.data(), n);
    };
    auto clear = [&](uint64_t dst, uint64_t n) { d.ensure_bytes(dst + n); };
    const uint64_t dst_root = GRID_SIZE + TREE_SIZE;
    auto find_tile = [&](uint64_t key) -> uint64_t {
        uint32_t count = d.gu32(dst_root + 24u);
        for (uint32_t t = 0; t < count; t++)
        {
            uint64_t tile = dst_root + ROOT_SZ + (uint64_t)t * TILE_SZ;
            if (d.gu64(tile) == key)
                return tile;
        }
        return 0ull;
    };
    uint64_t end = 0;
    // grid header copied with the type changed; tree header cleared
    copy_bytes(0, 0, GRID_SIZE);
    d.u32(636, GRID_TYPE_ONINDEX);
    clear(GRID_SIZE, TREE_SIZE);
    end = GRID_SIZE + TREE_SIZE;
    for (uint32_t pass = 0; pass < 5u; pass++)
    {
        if (pass == 0u)
        {
            clear(end, ROOT_SZ);
            end += ROOT_SZ;
            for (uint32_t rn = 0; rn < 32768u; rn++)
            {
                uint64_t key = get_key(rn);
                if (key == END_KEY)
                    continue;
                uint64_t tile = end;
                clear(end, TILE_SZ);
                end += TILE_SZ;
                d.u64(tile + 0u, key);
                d.u32(tile + 16u, value_bit(root, ROOT, rn) ? 1u : 0u);
                d.u32(dst_root + 24u, d.gu32(dst_root + 24u) + 1u);
            }
            // bbox from blind data [0]
            uint64_t meta0 = b.gu64(640);
            uint64_t bboxes = meta0 + b.gu64(meta0);
            for (uint32_t a = 0; a < 6u; a++)
                d.u32(dst_root + 4u * a, b.gu32(bboxes + 4u * a));
            d.u64(GRID_SIZE + 24u, dst_root - GRID_SIZE);
            continue;
        }
        uint64_t root_value_idx = b.r(root + OFF_VALUE_IDX);
        for (uint32_t rn = 0; rn < 32768u; rn++)
        {
            uint64_t key = get_key(rn);
            uint64_t tile = key != END_KEY ? find_tile(key) : 0ull;
            if (pass == 4u)
            {
                bool act = false;
                uint32_t local = mask_n_to_idx(b, root + OFF_VMP[ROOT], ROOT, rn, &act);
                if (act)
                    d.u64(tile + 24u, root_value_idx + local);
            }
            uint32_t up = get_child(root, ROOT, rn);
            if (!up)
                continue;
            if (pass == 1u)
            {
                uint64_t du = end;
                clear(end, UPPER_SZ);
                end += UPPER_SZ;
                copy_bytes(du + UPPER_CM, 8ull * (up + OFF_CM[UPPER]), 32768u / 8u);
                copy_bytes(du + UPPER_VM, 8ull * (up + OFF_VM[UPPER]), 32768u / 8u);
                d.u64(tile + 8u, du - dst_root);
                continue;
            }
            uint64_t upper_value_idx = b.r(up + OFF_VALUE_IDX);
            uint64_t du = dst_root + d.gu64(tile + 8u);
            for (uint32_t un = 0; un < 32768u; un++)
            {
                if (pass == 4u)
                {
                    bool act = false;
                    uint32_t local = mask_n_to_idx(b, up + OFF_VMP[UPPER], UPPER, un, &act);
                    if (act)
                        d.u64(du + UPPER_TABLE + 8ull * un, upper_value_idx + local);
                }
                uint32_t lo = get_child(up, UPPER, un);
                if (!lo)
                    continue;
                if (pass == 2u)
                {
                    uint64_t dl = end;
                    clear(end, LOWER_SZ);
                    end += LOWER_SZ;
                    copy_bytes(dl + LOWER_CM, 8ull * (lo + OFF_CM[LOWER]), 4096u / 8u);
                    copy_bytes(dl + LOWER_VM, 8ull * (lo + OFF_VM[LOWER]), 4096u / 8u);
                    d.u64(du + UPPER_TABLE + 8ull * un, dl - du);
                    continue;
                }
                uint64_t lower_value_idx = b.r(lo + OFF_VALUE_IDX);
                uint64_t dl = du + d.gu64(du + UPPER_TABLE + 8ull * un);
                for (uint32_t ln = 0; ln < 4096u; ln++)
                {
                    if (pass == 4u)
                    {
                        bool act = false;
                        uint32_t local = mask_n_to_idx(b, lo + OFF_VMP[LOWER], LOWER, ln, &act);
                        if (act)
                            d.u64(dl + LOWER_TABLE + 8ull * ln, lower_value_idx + local);
                    }
                    uint32_t lf = get_child(lo, LOWER, ln);
                    if (!lf)
                        continue;
                    if (pass == 3u)
                    {
                        uint64_t dfl = end;
                        clear(end, LEAF_SZ);
                        end += LEAF_SZ;
                        copy_bytes(dfl + LEAF_VM, 8ull * (lf + OFF_VM[LEAF]), 512u / 8u);
                        d.u64(dl + LOWER_TABLE + 8ull * ln, dfl - dl);
                        continue;
                    }
                    if (pass == 4u)
                    {
                        uint64_t dfl = dl + d.gu64(dl + LOWER_TABLE + 8ull * ln);
                        d.u64(dfl + LEAF_TABLE, b.r(lf + OFF_VALUE_IDX));
                        uint64_t sum = 0ull;
                        uint32_t accum = 0u;
                        for (uint32_t w = 0; w < 7u; w++)
                        {
                            accum += (uint32_t)__builtin_popcountll(d.gu64(dfl + LEAF_VM + 8u * w));
                            sum |= (uint64_t)accum << (9u * w);
                        }
                        d.u64(dfl + LEAF_TABLE + 8u, sum);
                    }
                }
            }
        }
    }
    // value blind data (entry [1] of the node2 grid) as the only blind data
    uint64_t meta1 = b.gu64(640) + BLINDMETADATA_SIZE;
    uint64_t src_data = meta1 + b.gu64(meta1);
    uint64_t dmeta = end;
    end += BLINDMETADATA_SIZE;
    copy_bytes(dmeta, meta1, BLINDMETADATA_SIZE);
    uint64_t count = b.gu64(meta1 + 8u);
    uint32_t vsize = b.gu32(meta1 + 16u);
    uint64_t ddata = end;
    uint64_t nbytes = count * vsize;
    end += nbytes;
    {
        std::vector<uint8_t> tmp(nbytes);
        b.get(src_data, tmp.data(), (uint32_t)nbytes);
        d.put(ddata, tmp.data(), (uint32_t)nbytes);
    }
    d.u64(dmeta, ddata - dmeta);
    d.u64(640, dmeta);
    d.u32(648, 1u);
    d.ensure_bytes(end);
    if (out && cap >= end)
        memcpy(out, d.w.data(), end);
    return end;
}

} // extern "C"

// ============================================================================ Raster2D tile binning
namespace
{
// float -> u32 as the GPUs' conversion instruction does it (NaN and negatives to 0, saturating
// at 2^32-1); the shaders write (uint32_t)floor(x), which C leaves undefined below zero
uint32_t sat_u32(float x)
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

// gaussian_count_tiles.slang:27-44 (the same box as gaussian_tile_intersections.slang:30-47)
TileBox tile_box(float mx, float my, int32_t radius, uint32_t tile_size, uint32_t tiles_w, uint32_t tiles_h)
{
    float radiusf = (float)radius;
    float tile_radius = radiusf / (float)tile_size;
    float u = mx / (float)tile_size;
    float v = my / (float)tile_size;
    TileBox b;
    b.x0 = std::min(sat_u32(std::floor(u - tile_radius)), tiles_w);
    b.y0 = std::min(sat_u32(std::floor(v - tile_radius)), tiles_h);
    b.x1 = std::min(sat_u32(std::ceil(u + tile_radius)), tiles_w);
    b.y1 = std::min(sat_u32(std::ceil(v + tile_radius)), tiles_h);
    return b;
}
} // namespace

extern "C" {

uint64_t vdbo_tile_bin(uint32_t n, const float* means2d, const int32_t* radii, const float* depths, uint32_t width,
                       uint32_t height, uint32_t tile_size, uint32_t* keys_low, uint32_t* keys_high, uint32_t* vals,
                       uint64_t cap, uint32_t* tile_offsets)
{
    const uint32_t tiles_w = (width + tile_size - 1u) / tile_size; // Raster2D.cpp:202-205
    const uint32_t tiles_h = (height + tile_size - 1u) / tile_size;
    const uint32_t num_tiles = tiles_w * tiles_h;
    // counts and their inclusive scan (Raster2D.cpp:347-366)
    std::vector<uint32_t> scan(n);
    uint32_t run = 0;
    for (uint32_t g = 0; g < n; g++)
    {
        uint32_t c = 0;
        if (radii[g] > 0)
        {
            TileBox b = tile_box(means2d[2 * g], means2d[2 * g + 1], radii[g], tile_size, tiles_w, tiles_h);
            c = (b.y1 - b.y0) * (b.x1 - b.x0);
        }
        run += c;
        scan[g] = run;
    }
    const uint64_t total = n ? scan[n - 1] : 0;
    if (tile_offsets)
        std::fill(tile_offsets, tile_offsets + num_tiles, 0u);
    if (total > cap || !keys_low || !keys_high || !vals)
        return total;
    // intersections in Gaussian-major, row-major tile order (gaussian_tile_intersections.slang:49-64)
    std::vector<uint32_t> lo(total), hi(total), va(total);
    for (uint32_t g = 0; g < n; g++)
    {
        if (radii[g] <= 0)
            continue;
        TileBox b = tile_box(means2d[2 * g], means2d[2 * g + 1], radii[g], tile_size, tiles_w, tiles_h);
        uint32_t depth_bits;
        std::memcpy(&depth_bits, &depths[g], 4);
        uint32_t cur = g ? scan[g - 1] : 0u;
        for (uint32_t i = b.y0; i < b.y1; i++)
            for (uint32_t j = b.x0; j < b.x1; j++)
            {
                lo[cur] = depth_bits;
                hi[cur] = i * tiles_w + j;
                va[cur] = g;
                cur++;
            }
    }
    // radix_sort_dual_key(low 32 bits, high tile-id bits): stable by (high, low) (Raster2D.cpp:495-507)
    uint32_t tile_bits = 0;
    while (tile_bits < 32 && (1u << tile_bits) < num_tiles)
        tile_bits++;
    const uint32_t mh = tile_bits >= 32 ? 0xFFFFFFFFu : ((1u << tile_bits) - 1u);
    std::vector<uint32_t> order(total);
    for (uint64_t e = 0; e < total; e++)
        order[e] = (uint32_t)e;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        uint32_t ha = hi[a] & mh, hb = hi[b] & mh;
        return ha != hb ? ha < hb : lo[a] < lo[b];
    });
    for (uint64_t e = 0; e < total; e++)
    {
        keys_low[e] = lo[order[e]];
        keys_high[e] = hi[order[e]];
        vals[e] = va[order[e]];
    }
    // tile offsets (gaussian_tile_offsets.slang:24-56); no intersections leaves every offset 0
    if (tile_offsets)
    {
        for (uint64_t e = 0; e < total; e++)
        {
            uint32_t t = keys_high[e];
            if (e == 0)
                for (uint32_t i = 0; i < t + 1u && i < num_tiles; i++)
                    tile_offsets[i] = 0u;
            if (e == total - 1)
                for (uint32_t i = t + 1u; i < num_tiles; i++)
                    tile_offsets[i] = (uint32_t)total;
            if (e > 0 && keys_high[e - 1] != t)
                for (uint32_t i = keys_high[e - 1] + 1u; i < t + 1u && i < num_tiles; i++)
                    tile_offsets[i] = (uint32_t)e;
        }
    }
    return total;
}

} // extern "C"

// ============================================================================ VoxelBVH Gaussian keys
namespace
{
struct V3
{
    float x, y, z;
};

// (p, 1) * m with m given as rows (the shaders' mul(float4(p, 1), transform) under the
// column-major constant layout of the transposed host matrix), summed left to right
V3 row_times(const float* m, V3 p)
{
    V3 r;
    r.x = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    r.y = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    r.z = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    return r;
}

struct BvhPrim
{
    V3 mean, scale;
    float qx, qy, qz, qw;
    float local_scale;
};

BvhPrim bvh_prim(const float* means, const float* opacities, const float* quats, const float* scales, uint64_t g)
{
    BvhPrim p;
    p.mean = { means[3 * g], means[3 * g + 1], means[3 * g + 2] };
    p.scale = { scales[3 * g], scales[3 * g + 1], scales[3 * g + 2] };
    p.qw = quats[4 * g];
    p.qx = quats[4 * g + 1];
    p.qy = quats[4 * g + 2];
    p.qz = quats[4 * g + 3];
    float o = std::fmax(0.01f, opacities[g]); // voxelbvh_gaussians_to_ijkl.slang:49
    p.local_scale = sqrtf(2.f * cr_logf(o / 0.01f));
    return p;
}

// voxelbvh_gaussians_to_ijkl.slang:150-182 (and bbox_reduce1.slang:56-89)
void bvh_box(const BvhPrim& p, const float* xf, V3& mn, V3& mx)
{
    const float x = p.qx, y = p.qy, z = p.qz, w = p.qw;
    const float R[3][3] = { { 1.f - 2.f * (y * y + z * z), 2.f * (x * y - z * w), 2.f * (x * z + y * w) },
                            { 2.f * (x * y + z * w), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - x * w) },
                            { 2.f * (x * z - y * w), 2.f * (y * z + x * w), 1.f - 2.f * (x * x + y * y) } };
    const float inf = INFINITY;
    mn = { inf, inf, inf };
    mx = { -inf, -inf, -inf };
    for (int v = 0; v < 8; v++)
    {
        float c[3] = { (v & 1) ? 1.f : -1.f, (v & 2) ? 1.f : -1.f, (v & 4) ? 1.f : -1.f };
        c[0] *= p.scale.x * p.local_scale;
        c[1] *= p.scale.y * p.local_scale;
        c[2] *= p.scale.z * p.local_scale;
        V3 q = { c[0] * R[0][0] + c[1] * R[0][1] + c[2] * R[0][2] + p.mean.x,
                 c[0] * R[1][0] + c[1] * R[1][1] + c[2] * R[1][2] + p.mean.y,
                 c[0] * R[2][0] + c[1] * R[2][1] + c[2] * R[2][2] + p.mean.z };
        q = row_times(xf, q);
        mn = { std::fmin(mn.x, q.x), std::fmin(mn.y, q.y), std::fmin(mn.z, q.z) };
        mx = { std::fmax(mx.x, q.x), std::fmax(mx.y, q.y), std::fmax(mx.z, q.z) };
    }
}

// voxelbvh_gaussians_to_ijkl.slang:29-105
bool bvh_overlap(const BvhPrim& p, const float* xf_inv, const float w2v[4], int32_t i, int32_t j, int32_t k,
                 uint32_t level)
{
    const float x = p.qx, y = p.qy, z = p.qz, w = p.qw;
    const float Ri[3][3] = { { 1.f - 2.f * (y * y + z * z), 2.f * (x * y + z * w), 2.f * (x * z - y * w) },
                             { 2.f * (x * y - z * w), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + x * w) },
                             { 2.f * (x * z + y * w), 2.f * (y * z - x * w), 1.f - 2.f * (x * x + y * y) } };
    const float inv_ext[3] = { 1.f / (p.local_scale * p.scale.x), 1.f / (p.local_scale * p.scale.y),
                               1.f / (p.local_scale * p.scale.z) };
    float lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
    for (int v = 0; v < 8; v++)
    {
        float vp[3] = { (float)i, (float)j, (float)k };
        for (int a = 0; a < 3; a++)
            if (v & (1 << a))
                vp[a] += (float)(1u << level);
        V3 l = { (vp[0] - w2v[0]) / w2v[3], (vp[1] - w2v[1]) / w2v[3], (vp[2] - w2v[2]) / w2v[3] };
        l = row_times(xf_inv, l);
        l = { l.x - p.mean.x, l.y - p.mean.y, l.z - p.mean.z };
        float r[3];
        for (int a = 0; a < 3; a++)
            r[a] = (l.x * Ri[a][0] + l.y * Ri[a][1] + l.z * Ri[a][2]) * inv_ext[a];
        for (int a = 0; a < 3; a++)
        {
            lo[a] = v == 0 ? r[a] : std::fmin(lo[a], r[a]);
            hi[a] = v == 0 ? r[a] : std::fmax(hi[a], r[a]);
        }
    }
    float d2 = 1.f;
    for (int a = 0; a < 3; a++)
    {
        if (0.f < lo[a])
            d2 -= lo[a] * lo[a];
        else if (0.f > hi[a])
            d2 -= hi[a] * hi[a];
    }
    return d2 > 0.f;
}

int32_t cvt_i32(float x) // GPU float -> int conversion (NaN 0, saturating)
{
    if (x != x)
        return 0;
    if (x <= -2147483648.f)
        return INT32_MIN;
    if (x >= 2147483648.f)
        return INT32_MAX;
    return (int32_t)x;
}
} // namespace

extern "C" {

void vdbo_mat_inverse(const float* m, float* out)
{
    // Camera.h:318-374: 1 / (sum of 24 signed products of one entry per row), then 16 cofactors of
    // 6 signed triple products each; products are formed left to right over rows, sums left to right
    auto M = [&](int r, int c) { return m[4 * r + c]; };
    static const int det_cols[24][4] = { { 0, 1, 2, 3 }, { 0, 2, 3, 1 }, { 0, 3, 1, 2 }, { 1, 0, 3, 2 }, { 1, 2, 0, 3 },
                                         { 1, 3, 2, 0 }, { 2, 0, 1, 3 }, { 2, 1, 3, 0 }, { 2, 3, 0, 1 }, { 3, 0, 2, 1 },
                                         { 3, 1, 0, 2 }, { 3, 2, 1, 0 }, { 0, 1, 3, 2 }, { 0, 2, 1, 3 }, { 0, 3, 2, 1 },
                                         { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 }, { 2, 0, 3, 1 }, { 2, 1, 0, 3 },
                                         { 2, 3, 1, 0 }, { 3, 0, 1, 2 }, { 3, 1, 2, 0 }, { 3, 2, 0, 1 } };
    float det = 0.f;
    for (int t = 0; t < 24; t++)
    {
        float a = M(0, det_cols[t][0]);
        if (t >= 12)
            a = -a;
        float prod = a * M(1, det_cols[t][1]) * M(2, det_cols[t][2]) * M(3, det_cols[t][3]);
        det = t ? det + prod : prod;
    }
    const float f = float(1.0) / det;
    // entry e of the result (rows of the returned matrix): excluded row / column, and the six terms
    // (sign, columns of the three remaining rows in ascending row order)
    static const int ex_row[16] = { 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3 };
    static const int terms[16][6][4] = {
        { { 1, 1, 2, 3 }, { 1, 2, 3, 1 }, { 1, 3, 1, 2 }, { -1, 1, 3, 2 }, { -1, 2, 1, 3 }, { -1, 3, 2, 1 } },
        { { 1, 1, 3, 2 }, { 1, 2, 1, 3 }, { 1, 3, 2, 1 }, { -1, 1, 2, 3 }, { -1, 2, 3, 1 }, { -1, 3, 1, 2 } },
        { { 1, 1, 2, 3 }, { 1, 2, 3, 1 }, { 1, 3, 1, 2 }, { -1, 1, 3, 2 }, { -1, 2, 1, 3 }, { -1, 3, 2, 1 } },
        { { 1, 1, 3, 2 }, { 1, 2, 1, 3 }, { 1, 3, 2, 1 }, { -1, 1, 2, 3 }, { -1, 2, 3, 1 }, { -1, 3, 1, 2 } },
        { { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 3, 2, 0 }, { -1, 0, 2, 3 }, { -1, 2, 3, 0 }, { -1, 3, 0, 2 } },
        { { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 }, { -1, 0, 3, 2 }, { -1, 2, 0, 3 }, { -1, 3, 2, 0 } },
        { { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 3, 2, 0 }, { -1, 0, 2, 3 }, { -1, 2, 3, 0 }, { -1, 3, 0, 2 } },
        { { 1, 0, 2, 3 }, { 1, 2, 3, 0 }, { 1, 3, 0, 2 }, { -1, 0, 3, 2 }, { -1, 2, 0, 3 }, { -1, 3, 2, 0 } },
        { { 1, 0, 1, 3 }, { 1, 1, 3, 0 }, { 1, 3, 0, 1 }, { -1, 0, 3, 1 }, { -1, 1, 0, 3 }, { -1, 3, 1, 0 } },
        { { -1, 0, 1, 3 }, { -1, 1, 3, 0 }, { -1, 3, 0, 1 }, { 1, 0, 3, 1 }, { 1, 1, 0, 3 }, { 1, 3, 1, 0 } },
        { { 1, 0, 1, 3 }, { 1, 1, 3, 0 }, { 1, 3, 0, 1 }, { -1, 0, 3, 1 }, { -1, 1, 0, 3 }, { -1, 3, 1, 0 } },
        { { 1, 0, 3, 1 }, { 1, 1, 0, 3 }, { 1, 3, 1, 0 }, { -1, 1, 3, 0 }, { -1, 3, 0, 1 }, { -1, 0, 1, 3 } },
        { { 1, 0, 2, 1 }, { 1, 1, 0, 2 }, { 1, 2, 1, 0 }, { -1, 0, 1, 2 }, { -1, 1, 2, 0 }, { -1, 2, 0, 1 } },
        { { 1, 0, 1, 2 }, { 1, 1, 2, 0 }, { 1, 2, 0, 1 }, { -1, 0, 2, 1 }, { -1, 1, 0, 2 }, { -1, 2, 1, 0 } },
        { { 1, 0, 2, 1 }, { 1, 1, 0, 2 }, { 1, 2, 1, 0 }, { -1, 0, 1, 2 }, { -1, 1, 2, 0 }, { -1, 2, 0, 1 } },
        { { 1, 0, 1, 2 }, { 1, 1, 2, 0 }, { 1, 2, 0, 1 }, { -1, 0, 2, 1 }, { -1, 1, 0, 2 }, { -1, 2, 1, 0 } },
    };
    for (int e = 0; e < 16; e++)
    {
        int rows[3], n = 0;
        for (int r = 0; r < 4; r++)
            if (r != ex_row[e])
                rows[n++] = r;
        float s = 0.f;
        for (int t = 0; t < 6; t++)
        {
            float a = M(rows[0], terms[e][t][1]);
            if (terms[e][t][0] < 0)
                a = -a;
            float prod = a * M(rows[1], terms[e][t][2]) * M(rows[2], terms[e][t][3]);
            s = t ? s + prod : prod;
        }
        out[e] = s * f;
    }
}

int vdbo_voxelbvh_ijkl(uint64_t n, const float* means, const float* opacities, const float* quats,
                       const float* scales, uint32_t resolution, const float* transform, uint64_t* keys,
                       uint32_t* ids, uint32_t* ranges, float* bbox)
{
    if (resolution < 1 || resolution > 4096) // VoxelBVH.cpp:30-33
        return -1;
    float xf[16], xf_inv[16];
    for (int e = 0; e < 16; e++)
        xf[e] = xf_inv[e] = (e % 5 == 0) ? 1.f : 0.f;
    if (transform)
    {
        std::memcpy(xf, transform, sizeof(xf));
        vdbo_mat_inverse(xf, xf_inv);
    }
    // world box (bbox_reduce1/2.slang): min / max over every corner of every Gaussian
    V3 wmn = { INFINITY, INFINITY, INFINITY }, wmx = { -INFINITY, -INFINITY, -INFINITY };
    for (uint64_t g = 0; g < n; g++)
    {
        V3 mn, mx;
        bvh_box(bvh_prim(means, opacities, quats, scales, g), xf, mn, mx);
        wmn = { std::fmin(wmn.x, mn.x), std::fmin(wmn.y, mn.y), std::fmin(wmn.z, mn.z) };
        wmx = { std::fmax(wmx.x, mx.x), std::fmax(wmx.y, mx.y), std::fmax(wmx.z, mx.z) };
    }
    bbox[0] = wmn.x, bbox[1] = wmn.y, bbox[2] = wmn.z, bbox[3] = wmx.x, bbox[4] = wmx.y, bbox[5] = wmx.z;
    if (n == 0)
        return 0;
    // compute_world_to_voxel (voxelbvh_gaussians_to_ijkl.slang:107-114)
    const float dx = wmx.x - wmn.x, dy = wmx.y - wmn.y, dz = wmx.z - wmn.z;
    const float dmax = std::fmax(dx, std::fmax(dy, dz));
    const float sc = float(resolution - 1u) / dmax;
    const float w2v[4] = { -wmn.x * sc + 0.5f, -wmn.y * sc + 0.5f, -wmn.z * sc + 0.5f, sc };
    std::vector<uint64_t> k(8 * n);
    std::vector<uint32_t> id(8 * n);
    for (uint64_t g = 0; g < n; g++)
    {
        const BvhPrim p = bvh_prim(means, opacities, quats, scales, g);
        V3 mn, mx;
        bvh_box(p, xf, mn, mx);
        int32_t lo[3] = { cvt_i32(std::floor(sc * mn.x + w2v[0])), cvt_i32(std::floor(sc * mn.y + w2v[1])),
                          cvt_i32(std::floor(sc * mn.z + w2v[2])) };
        int32_t hi[3] = { cvt_i32(std::floor(sc * mx.x + w2v[0])), cvt_i32(std::floor(sc * mx.y + w2v[1])),
                          cvt_i32(std::floor(sc * mx.z + w2v[2])) };
        auto overlaps = [&](uint32_t level, uint32_t cap) {
            const int32_t msk = ~((1 << level) - 1);
            uint32_t c = 0;
            for (int32_t i = lo[0] & msk; i <= (hi[0] & msk); i += (1 << level))
                for (int32_t j = lo[1] & msk; j <= (hi[1] & msk); j += (1 << level))
                    for (int32_t kk = lo[2] & msk; kk <= (hi[2] & msk); kk += (1 << level))
                    {
                        if (bvh_overlap(p, xf_inv, w2v, i, j, kk, level))
                        {
                            if (c < cap)
                            {
                                k[8 * g + c] = ((uint64_t)(int64_t)i << 48) | ((uint64_t)(int64_t)j << 32) |
                                               ((uint64_t)(int64_t)kk << 16) | (uint64_t)level;
                                id[8 * g + c] = (uint32_t)g;
                            }
                            c++;
                        }
                    }
            return c;
        };
        // voxelbvh_gaussians_to_ijkl.slang:186-221: search from level 11 towards 0
        uint32_t level = 11, last_good = 11;
        while (level < 12)
        {
            if (overlaps(level, 0) > 8)
            {
                level = last_good;
                break;
            }
            if (level == 0)
                break;
            last_good = level;
            level -= 1;
        }
        uint32_t written = std::min<uint32_t>(overlaps(level, 8), 8);
        for (uint32_t s2 = written; s2 < 8; s2++)
        {
            k[8 * g + s2] = ~0ull;
            id[8 * g + s2] = ~0u;
        }
    }
    // stable sort of (key, id) by key (radix_sort_key64 over 64 bits)
    std::vector<uint32_t> order(8 * n);
    for (uint64_t e = 0; e < 8 * n; e++)
        order[e] = (uint32_t)e;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    for (uint64_t e = 0; e < 8 * n; e++)
    {
        keys[e] = k[order[e]];
        ids[e] = id[order[e]];
    }
    // run headers (find_range_starts.slang, global_scan, scatter_range_headers.slang)
    std::fill(ranges, ranges + 16 * n, 0u);
    uint32_t run = 0;
    for (uint64_t e = 0; e < 8 * n; e++)
    {
        if (e == 0 || keys[e] != keys[e - 1])
        {
            ranges[2 * run] = (uint32_t)e;
            run++;
        }
        if (e + 1 == 8 * n || keys[e + 1] != keys[e])
            ranges[2 * (run - 1) + 1] = (uint32_t)(e + 1);
    }
    return 0;
}

} // extern "C"
