// vdb_node2.h -- device-side node2 accessors on a grid held as 64-bit words.
//
// Restate nanovdb_editor/PNanoVDBExt.h:94-143 (mask_n_to_idx over the packed 16-bit prefix
// lanes) and the descent of pnanovdb_node2_find_node (PNanoVDBExt.h:225-318) for the grids this
// library writes: root keys sit at their direct slot (coordinates within the 17-bit root range,
// no probing) and every node has value level 0.
#ifndef VDB_NODE2_H
#define VDB_NODE2_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vdb_layout.h"

__device__ __forceinline__ uint32_t n2_mask_n_to_idx(const uint64_t* w, uint64_t prefix_idx64, uint32_t prefix_to_mask,
                                                     uint32_t n, bool* active)
{
    uint64_t mask = w[prefix_idx64 + prefix_to_mask + (n >> 6u)];
    uint32_t prefix = (uint32_t)(w[prefix_idx64 + (n >> 8u)] >> ((n >> 2u) & 0x30u)) & 0xFFFFu;
    uint64_t cmp = 1ull << (n & 63u);
    *active = (mask & cmp) != 0ull;
    return (uint32_t)__popcll(mask & (cmp - 1ull)) + (n >= 64u ? prefix : 0u);
}

// word index of the leaf holding voxel (i, j, k), or 0 when the voxel lies in no leaf
__device__ __forceinline__ uint64_t n2_find_leaf(const uint64_t* w, int32_t i, int32_t j, int32_t k)
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
    uint32_t child_count = (uint32_t)(w[root + VDB_OFF_CMP_ROOT] & 0xFFFFu);
    uint64_t key = (uint64_t)(uk >> 12) | ((uint64_t)(uj >> 12) << 21) | ((uint64_t)(ui >> 12) << 42);
    if (!kact || w[root + VDB_OFF_CHILDREN_ROOT + child_count + kidx] != key)
        return 0;
    uint64_t upper = root + (w[root + VDB_OFF_CHILDREN_ROOT + rank] >> 3);
    uint32_t un = (((ui & 0xFFFu) >> 7) << 10) | (((uj & 0xFFFu) >> 7) << 5) | ((uk & 0xFFFu) >> 7);
    rank = n2_mask_n_to_idx(w, upper + VDB_OFF_CMP_UPPER, VDB_OFF_CM_UPPER - VDB_OFF_CMP_UPPER, un, &act);
    if (!act)
        return 0;
    uint64_t lower = upper + (w[upper + VDB_OFF_CHILDREN_UPPER + rank] >> 3);
    uint32_t ln = (((ui & 0x7Fu) >> 3) << 8) | (((uj & 0x7Fu) >> 3) << 4) | ((uk & 0x7Fu) >> 3);
    rank = n2_mask_n_to_idx(w, lower + VDB_OFF_CMP_LOWER, VDB_OFF_CM_LOWER - VDB_OFF_CMP_LOWER, ln, &act);
    if (!act)
        return 0;
    return lower + (w[lower + VDB_OFF_CHILDREN_LOWER + rank] >> 3);
}

// value index of voxel (i, j, k) in a leaf (word index), or ~0 when the voxel is inactive
__device__ __forceinline__ uint64_t n2_leaf_value_index(const uint64_t* w, uint64_t leaf, int32_t i, int32_t j,
                                                        int32_t k)
{
    uint32_t n = (((uint32_t)i & 7u) << 6) | (((uint32_t)j & 7u) << 3) | ((uint32_t)k & 7u);
    bool act = false;
    uint32_t local = n2_mask_n_to_idx(w, leaf + VDB_OFF_VMP_LEAF, VDB_OFF_VM_LEAF - VDB_OFF_VMP_LEAF, n, &act);
    return act ? w[leaf + VDB_OFF_VALUE_IDX] + local : ~0ull;
}

#endif
