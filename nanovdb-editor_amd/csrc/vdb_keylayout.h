// vdb_keylayout.h -- bit layout of the 64-bit pair keys and voxel keys (host and device).
#ifndef VDB_KEYLAYOUT_H
#define VDB_KEYLAYOUT_H

#include <stdint.h>

// voxel key = (leaf index << 9) | slot; leaf index = per-axis leaf coordinates relative to lmin,
// x most significant (bits[0] + bits[1] + bits[2] bits); slot = (i & 7) << 6 | (j & 7) << 3 | (k & 7).
// pair key = voxel key << (gid_bits + alpha_bits) | gaussian id << alpha_bits | alpha8.
struct KeyLayout
{
    int32_t lmin[3];
    uint32_t bits[3];
    uint32_t gid_bits;
    uint32_t voxel_bits; // bits[0] + bits[1] + bits[2] + 9
    uint32_t alpha_bits; // 8 or 0
    uint32_t pad;
};

#endif
