/*
 * vdb_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference raster3d path (Gaussian array -> node2 NanoVDB grid).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker. The product (libvdbraster) never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * openvdb/nanovdb-editor source tree). Parity status of each function is listed in
 * the header comment of vdb_oracle.cpp.
 */
#ifndef VDB_ORACLE_H
#define VDB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vdbo_stats_t
{
    uint64_t fragment_count;  /* AABB candidates (F) over all prim batches */
    uint64_t pair_count;      /* surviving (voxel, gaussian) pairs (P) */
    uint64_t window_count;    /* non-empty fragment windows processed */
    uint64_t grid_size;       /* bytes of the node2 grid */
    uint32_t upper_count;
    uint32_t lower_count;
    uint32_t leaf_count;
    uint32_t pad;
    uint64_t value_count;     /* including the 2 background slots */
    int32_t bbox_min[3];
    int32_t bbox_max[3];
} vdbo_stats_t;

/* RasterUtils.cpp:31-136 -- in-place preprocessing of raw arrays + color derivation. */
void vdbo_preprocess(uint64_t n, float* means, float* opacities, float* quats, float* scales,
                     const float* sh0, float* colors_out);

/* gaussian_prim.slang:25-145 -- per-Gaussian voxel AABB and footprint. */
void vdbo_prim_aabb(uint64_t n, const float* means, const float* quats, const float* scales,
                    const float* opacities, float voxel_size, int32_t* bbox6_out, uint64_t* footprint_out);

/* gaussian_frag_alloc.slang:28-177 (+ gaussian_frag_color.slang:184-212 for rgba) --
 * every surviving pair in reference fragment order (prim, then raster index).
 * Writes min(count, cap) entries; returns the total count. */
uint64_t vdbo_pairs(uint64_t n, const float* means, const float* quats, const float* scales,
                    const float* colors, const float* opacities, float voxel_size, float cull_scale,
                    uint32_t* gid_out, int32_t* ijk_out, uint32_t* rgba_out, uint64_t cap, int nthreads);

/* Raster.cpp:420-558 + GridBuild.cpp:361-650 + grid_build_*.slang -- full windowed build.
 * value_mode 0: RGBA8 compat values (gaussian_frag_color.slang); 1: float density (FP64
 * oracle rounded to float). Returns the grid size in bytes; writes the grid to out when
 * cap >= size. Returns 0 on capacity errors (node count above max_node_count). */
uint64_t vdbo_raster_node2(uint64_t n, const float* means, const float* quats, const float* scales,
                           const float* colors, const float* opacities, float voxel_size, float cull_scale,
                           uint32_t prim_batch_size, uint32_t ijk_batch_size, uint32_t max_node_count,
                           int value_mode, uint8_t* out, uint64_t cap, int nthreads, vdbo_stats_t* stats);

/* FP64 density at voxel centres for a pair list; pairs need not be sorted.
 * Output: unique voxels (ijk, sorted lexicographically i,j,k) and density sums taken in
 * ascending gaussian id. Returns the unique voxel count. */
uint64_t vdbo_density_f64(uint64_t pair_count, const uint32_t* gid, const int32_t* ijk, const float* means,
                          const float* quats, const float* scales, const float* opacities, float voxel_size,
                          int32_t* voxel_ijk_out, double* density_out);

/* FP64 adjoint of L = sum_v (rho(v) - target(v))^2 over a pair list, w.r.t. the raw
 * parameters (means, log-scales, unnormalised quats, opacity logits). */
void vdbo_density_grad_f64(uint64_t n, uint64_t pair_count, const uint32_t* gid, const int32_t* ijk,
                           const float* means, const float* log_scales, const float* quats_raw,
                           const float* opacity_logits, float voxel_size, const double* residual_per_pair,
                           double* grad_means, double* grad_log_scales, double* grad_quats,
                           double* grad_opacity_logits);

/* node2 reader (PNanoVDBExt.h:94-318): list active voxels with their value index, in
 * tree order (root slot, upper slot, lower slot, leaf slot). Returns the count. */
uint64_t vdbo_node2_active_voxels(const uint8_t* grid, uint64_t grid_size, int32_t* ijk_out,
                                  uint64_t* value_idx_out, uint64_t cap);

/* putil/Convert.h:40-313 -- node2 grid to NanoVDB OnIndex (value blind data carried over).
 * Returns the OnIndex size in bytes; writes it to out when cap >= size. */
uint64_t vdbo_node2_to_onindex(const uint8_t* grid, uint64_t grid_size, uint8_t* out, uint64_t cap);

/* Raster2D tile binning (raster/Raster2D.cpp:347-535): per-Gaussian tile counts
 * (gaussian_count_tiles.slang:14-46), inclusive scan, intersections (gaussian_tile_intersections.slang:18-65),
 * radix_sort_dual_key(32, tile-id bits) (Raster2D.cpp:495-507, ParallelPrimitives.cpp:527-683) and
 * tile offsets (gaussian_tile_offsets.slang:13-56). Returns the intersection count; fills the
 * sorted keys/vals when cap >= count and tile_offsets[num_tiles] when not NULL. */
uint64_t vdbo_tile_bin(uint32_t n, const float* means2d, const int32_t* radii, const float* depths, uint32_t width,
                       uint32_t height, uint32_t tile_size, uint32_t* keys_low, uint32_t* keys_high, uint32_t* vals,
                       uint64_t cap, uint32_t* tile_offsets);

/* VoxelBVH Gaussian keys (raster/VoxelBVH.cpp:1225-1435): world box, up to 8 cell keys per
 * Gaussian (voxelbvh_gaussians_to_ijkl.slang), stable sort, run headers. keys/ids hold 8 n
 * entries, ranges 16 n, bbox 6. transform: 16 floats (rows) or NULL. Returns 0, or -1 for a
 * resolution outside [1, 4096]. */
int vdbo_voxelbvh_ijkl(uint64_t n, const float* means, const float* opacities, const float* quats,
                       const float* scales, uint32_t resolution, const float* transform, uint64_t* keys,
                       uint32_t* ids, uint32_t* ranges, float* bbox);

/* pnanovdb_camera_mat_inverse (putil/Camera.h:316-375) in the reference's term order. */
void vdbo_mat_inverse(const float* m, float* out);

#ifdef __cplusplus
}
#endif

#endif
