// vdb_layout.h -- node2 and NanoVDB header layout constants used by the grid writer.
//
// node2 (nanovdb_editor/PNanoVDBExt.h:21-83): 4 node types ROOT/UPPER/LOWER/LEAF with fanouts
// 32768/32768/4096/512 and child spans 4096/128/8/1 voxels; offsets below are in 64-bit words
// from the node start. Packed prefix sums: each word holds four 16-bit lanes; lane l of word q
// is the popcount of mask words [0, 4q + l), except lane 0 of word 0 which holds the node's
// total (grid_build_count_new.slang:74-99).
//
// NanoVDB header layout (GridData 672 B, TreeData 64 B, GridBlindMetaData 288 B) follows the
// public NanoVDB format (PNanoVDB.h, fetched by the reference build and not part of its tree).
#ifndef VDB_LAYOUT_H
#define VDB_LAYOUT_H

#include <stdint.h>

#define VDB_NODE_ROOT 0u
#define VDB_NODE_UPPER 1u
#define VDB_NODE_LOWER 2u
#define VDB_NODE_LEAF 3u

// word offsets
#define VDB_OFF_NODE_IDX 0u
#define VDB_OFF_VALUE_IDX 1u
#define VDB_OFF_VALUE_LEVEL 2u
// value mask prefix / value mask / child mask prefix / child mask / key mask prefix / key mask / children
#define VDB_OFF_VMP_ROOT 258u
#define VDB_OFF_VM_ROOT 386u
#define VDB_OFF_CMP_ROOT 898u
#define VDB_OFF_CM_ROOT 1026u
#define VDB_OFF_KMP_ROOT 1538u
#define VDB_OFF_KM_ROOT 1666u
#define VDB_OFF_CHILDREN_ROOT 2180u
#define VDB_OFF_VMP_UPPER 258u
#define VDB_OFF_VM_UPPER 386u
#define VDB_OFF_CMP_UPPER 898u
#define VDB_OFF_CM_UPPER 1026u
#define VDB_OFF_CHILDREN_UPPER 1538u
#define VDB_OFF_VMP_LOWER 34u
#define VDB_OFF_VM_LOWER 50u
#define VDB_OFF_CMP_LOWER 114u
#define VDB_OFF_CM_LOWER 130u
#define VDB_OFF_CHILDREN_LOWER 194u
#define VDB_OFF_VMP_LEAF 6u
#define VDB_OFF_VM_LEAF 8u

// node byte sizes (PNanoVDBExt.h:63)
#define VDB_SIZE_ROOT 541728ull
#define VDB_SIZE_UPPER 274464ull
#define VDB_SIZE_LOWER 34336ull
#define VDB_SIZE_LEAF 128ull

#define VDB_END_KEY (1ull << 63)

// NanoVDB headers
#define VDB_GRID_SIZE 672u
#define VDB_TREE_SIZE 64u
#define VDB_BLINDMETADATA_SIZE 288u
#define VDB_TREE_ADDR 672ull
#define VDB_ROOT_ADDR 736ull
#define VDB_MAGIC_GRID 0x314244566f6e614eull /* "NanoVDB1" */
#define VDB_VERSION ((32u << 21) | (7u << 10) | 0u)
#define VDB_GRID_TYPE_FLOAT 1u
#define VDB_GRID_TYPE_VEC3F 6u
#define VDB_GRID_TYPE_RGBA8 12u
#define VDB_GRID_TYPE_ONINDEX 20u
#define VDB_GRID_TYPE_NODE2 27u

// GridData byte offsets
#define VDB_GRID_OFF_MAGIC 0u
#define VDB_GRID_OFF_VERSION 16u
#define VDB_GRID_OFF_FLAGS 20u
#define VDB_GRID_OFF_GRID_INDEX 24u
#define VDB_GRID_OFF_GRID_COUNT 28u
#define VDB_GRID_OFF_GRID_SIZE 32u
#define VDB_GRID_OFF_NAME 40u
#define VDB_GRID_OFF_MAP 296u
#define VDB_GRID_OFF_WORLD_BBOX 560u
#define VDB_GRID_OFF_VOXEL_SIZE 608u
#define VDB_GRID_OFF_CLASS 632u
#define VDB_GRID_OFF_TYPE 636u
#define VDB_GRID_OFF_BLIND_OFFSET 640u
#define VDB_GRID_OFF_BLIND_COUNT 648u
// Map (264 B): matf[9] @0, invmatf[9] @36, vecf[3] @72, taperf @84, matd[9] @88,
// invmatd[9] @160, vecd[3] @232, taperd @256
#define VDB_MAP_OFF_MATF 0u
#define VDB_MAP_OFF_INVMATF 36u
#define VDB_MAP_OFF_MATD 88u
#define VDB_MAP_OFF_INVMATD 160u
// TreeData: node_offset_leaf/lower/upper/root @0/8/16/24, node_count_leaf/lower/upper @32/36/40
#define VDB_TREE_OFF_ROOT 24u
#define VDB_TREE_OFF_COUNT_LEAF 32u
#define VDB_TREE_OFF_COUNT_LOWER 36u
#define VDB_TREE_OFF_COUNT_UPPER 40u
// GridBlindMetaData: data_offset @0, value_count @8, value_size @16, semantic @20,
// data_class @24, data_type @28, name @32
#define VDB_BLIND_OFF_DATA 0u
#define VDB_BLIND_OFF_COUNT 8u
#define VDB_BLIND_OFF_SIZE 16u
#define VDB_BLIND_OFF_TYPE 28u

#endif
