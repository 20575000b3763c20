// vdb_keylayout.h -- bit layout of the 64-bit pair keys and voxel keys (host and device).
#ifndef VDB_KEYLAYOUT_H
#define VDB_KEYLAYOUT_H

#include <stdint.h>

// voxel key = (leaf index << 9) | slot; leaf index = per-axis leaf coordinates relative to lmin,
// x most significant (bits[0] + bits[1] + bits[2] bits); slot = (i & 7) << 6 | (j & 7) << 3 | (k & 7).
// pair key = leaf field << (9 + gid_bits + alpha_bits) | slot << (gid_bits + alpha_bits)
//            | gaussian id << alpha_bits | alpha8,
// where the leaf field is the leaf index, or with rank_bits > 0 the leaf's rank among the
// candidate leaves (the leaves of the Gaussians' AABBs, marked in leaf_map): ranks follow the leaf
// index order, so pair keys sort the same way in rank_bits instead of bits[0..2] bits. Voxel keys
// always carry the leaf index.
struct KeyLayout
{
    int32_t lmin[3];
    uint32_t bits[3];
    uint32_t gid_bits;
    uint32_t voxel_bits; // bits of a pair key's voxel part: leaf field + 9
    uint32_t alpha_bits; // 8 or 0
    uint32_t rank_bits;  // 0: the leaf field is the leaf index
    const uint64_t* leaf_map;     // candidate-leaf bitmap over the leaf index (rank_bits > 0)
    const uint32_t* leaf_prefix;  // candidate leaves before each bitmap word
    const uint32_t* leaf_of_rank; // leaf index of each rank
};

#endif
