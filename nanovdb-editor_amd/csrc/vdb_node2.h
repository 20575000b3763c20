// vdb_node2.h -- node2 accessors on a grid held as 64-bit words (device kernels and host checks).
//
// Restate nanovdb_editor/PNanoVDBExt.h:94-143 (mask_n_to_idx over the packed 16-bit prefix
// lanes) and the descent of pnanovdb_node2_find_node (PNanoVDBExt.h:225-318) for the grids this
// library writes: root keys sit at their direct slot (coordinates within the 17-bit root range,
// no probing) and every node has value level 0. The grid may come from a caller, so every child
// step is bounded: the child's rank must lie below its parent's fanout and the child node must lie
// inside the grid's nwords words; a step that fails reports the grid as corrupt (*bad) instead of
// reading outside it.
#ifndef VDB_NODE2_H
#define VDB_NODE2_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vdb_layout.h"

#define VDB_N2 __host__ __device__ __forceinline__

// node sizes in 64-bit words
constexpr uint64_t N2_WORDS_ROOT = VDB_SIZE_ROOT / 8u;
constexpr uint64_t N2_WORDS_UPPER = VDB_SIZE_UPPER / 8u;
constexpr uint64_t N2_WORDS_LOWER = VDB_SIZE_LOWER / 8u;
constexpr uint64_t N2_WORDS_LEAF = VDB_SIZE_LEAF / 8u;

VDB_N2 uint32_t n2_mask_n_to_idx(const uint64_t* w, uint64_t prefix_idx64, uint32_t prefix_to_mask, uint32_t n,
                                 bool* active)
{
    uint64_t mask = w[prefix_idx64 + prefix_to_mask + (n >> 6u)];
    uint32_t prefix = (uint32_t)(w[prefix_idx64 + (n >> 8u)] >> ((n >> 2u) & 0x30u)) & 0xFFFFu;
    uint64_t cmp = 1ull << (n & 63u);
    *active = (mask & cmp) != 0ull;
    return (uint32_t)__builtin_popcountll(mask & (cmp - 1ull)) + (n >= 64u ? prefix : 0u);
}

// word index of child `rank` of `node` (children table at children_off, `fanout` slots), or 0 with
// *bad set when the rank or the child node (child_words long) falls outside the grid
VDB_N2 uint64_t n2_child(const uint64_t* w, uint64_t nwords, uint64_t node, uint32_t children_off, uint32_t rank,
                         uint32_t fanout, uint64_t child_words, bool* bad)
{
    if (rank >= fanout)
    {
        *bad = true;
        return 0;
    }
    const uint64_t rel = w[node + children_off + rank] >> 3;
    if (rel == 0 || rel > nwords || node + rel > nwords - child_words)
    {
        *bad = true;
        return 0;
    }
    return node + rel;
}

// the root's key of slot rn (its rank kidx among the keyed slots), or ~0 with *bad set when the
// key table index leaves the root node
VDB_N2 uint64_t n2_root_key(const uint64_t* w, uint32_t kidx, bool* bad)
{
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    const uint32_t child_count = (uint32_t)(w[root + VDB_OFF_CMP_ROOT] & 0xFFFFu);
    if (child_count > 32768u || kidx >= 32768u)
    {
        *bad = true;
        return ~0ull;
    }
    return w[root + VDB_OFF_CHILDREN_ROOT + child_count + kidx];
}

// word index of the leaf holding voxel (i, j, k), or 0 when the voxel lies in no leaf (or the
// descent leaves the grid: then *bad is set too). nwords >= the root's end is the caller's check.
VDB_N2 uint64_t n2_find_leaf(const uint64_t* w, uint64_t nwords, int32_t i, int32_t j, int32_t k, bool* bad)
{
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    const uint32_t ui = (uint32_t)i, uj = (uint32_t)j, uk = (uint32_t)k;
    uint32_t rn = (((ui & 0x1FFFFu) >> 12) << 10) | (((uj & 0x1FFFFu) >> 12) << 5) | ((uk & 0x1FFFFu) >> 12);
    bool act = false;
    uint32_t rank = n2_mask_n_to_idx(w, root + VDB_OFF_CMP_ROOT, VDB_OFF_CM_ROOT - VDB_OFF_CMP_ROOT, rn, &act);
    if (!act)
        return 0;
    bool kact = false;
    uint32_t kidx = n2_mask_n_to_idx(w, root + VDB_OFF_KMP_ROOT, VDB_OFF_KM_ROOT - VDB_OFF_KMP_ROOT, rn, &kact);
    uint64_t key = (uint64_t)(uk >> 12) | ((uint64_t)(uj >> 12) << 21) | ((uint64_t)(ui >> 12) << 42);
    if (!kact || n2_root_key(w, kidx, bad) != key)
        return 0;
    uint64_t upper = n2_child(w, nwords, root, VDB_OFF_CHILDREN_ROOT, rank, 32768u, N2_WORDS_UPPER, bad);
    if (!upper)
        return 0;
    uint32_t un = (((ui & 0xFFFu) >> 7) << 10) | (((uj & 0xFFFu) >> 7) << 5) | ((uk & 0xFFFu) >> 7);
    rank = n2_mask_n_to_idx(w, upper + VDB_OFF_CMP_UPPER, VDB_OFF_CM_UPPER - VDB_OFF_CMP_UPPER, un, &act);
    if (!act)
        return 0;
    uint64_t lower = n2_child(w, nwords, upper, VDB_OFF_CHILDREN_UPPER, rank, 32768u, N2_WORDS_LOWER, bad);
    if (!lower)
        return 0;
    uint32_t ln = (((ui & 0x7Fu) >> 3) << 8) | (((uj & 0x7Fu) >> 3) << 4) | ((uk & 0x7Fu) >> 3);
    rank = n2_mask_n_to_idx(w, lower + VDB_OFF_CMP_LOWER, VDB_OFF_CM_LOWER - VDB_OFF_CMP_LOWER, ln, &act);
    if (!act)
        return 0;
    return n2_child(w, nwords, lower, VDB_OFF_CHILDREN_LOWER, rank, 4096u, N2_WORDS_LEAF, bad);
}

// value index of voxel (i, j, k) in a leaf (word index), or ~0 when the voxel is inactive
VDB_N2 uint64_t n2_leaf_value_index(const uint64_t* w, uint64_t leaf, int32_t i, int32_t j, int32_t k)
{
    uint32_t n = (((uint32_t)i & 7u) << 6) | (((uint32_t)j & 7u) << 3) | ((uint32_t)k & 7u);
    bool act = false;
    uint32_t local = n2_mask_n_to_idx(w, leaf + VDB_OFF_VMP_LEAF, VDB_OFF_VM_LEAF - VDB_OFF_VMP_LEAF, n, &act);
    return act ? w[leaf + VDB_OFF_VALUE_IDX] + local : ~0ull;
}

#endif
