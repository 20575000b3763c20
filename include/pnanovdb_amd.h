/*
 * pnanovdb_amd.h -- C ABI of libvdbraster.so, the MI355X (gfx950) Gaussian -> NanoVDB voxelizer.
 *
 * Two layers:
 *   1. The reference's function-table ABI for the raster3d path (pnanovdb_get_raster,
 *      pnanovdb_get_compute, pnanovdb_get_parallel_primitives, pnanovdb_get_grid_build),
 *      with struct layouts identical to openvdb/nanovdb-editor so existing callers (the
 *      ctypes pymodule, C++ hosts copying the tables by value) bind without change.
 *   2. Direct entry points (vdbr_*) on device pointers, used by the Python package on torch
 *      tensors, by the tests and by the benchmark.
 * No torch or HIP types appear in any signature; streams travel as void*.
 */
#ifndef PNANOVDB_AMD_H
#define PNANOVDB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- basic types
 * nanovdb_editor/putil/Reflect.h:30-61, PNanoVDB.h scalar typedefs */
#ifndef PNANOVDB_ABI
#define PNANOVDB_ABI
#endif
#ifndef PNANOVDB_API
#ifdef __cplusplus
#define PNANOVDB_API extern "C" __attribute__((visibility("default")))
#else
#define PNANOVDB_API __attribute__((visibility("default")))
#endif
#endif

typedef uint32_t pnanovdb_uint32_t;
typedef int32_t pnanovdb_int32_t;
typedef uint64_t pnanovdb_uint64_t;
typedef int64_t pnanovdb_int64_t;
typedef int32_t pnanovdb_bool_t; /* 0 / 1 */
#define PNANOVDB_FALSE 0
#define PNANOVDB_TRUE 1

typedef struct pnanovdb_reflect_data_type_t pnanovdb_reflect_data_type_t;

/* nanovdb_editor/putil/Compute.h:765-771 */
typedef struct pnanovdb_compute_array_t
{
    void* data;
    pnanovdb_uint64_t element_size;
    pnanovdb_uint64_t element_count;
    const char* filepath;
} pnanovdb_compute_array_t;

/* Opaque handles of the reference ABI. In this library a pnanovdb_compute_buffer_t* is a
 * vdbr_buffer_t* (device allocation owned by the caller); queues and contexts are opaque. */
typedef struct pnanovdb_compute_queue_t pnanovdb_compute_queue_t;
typedef struct pnanovdb_compute_buffer_t pnanovdb_compute_buffer_t;
typedef struct pnanovdb_compute_texture_t pnanovdb_compute_texture_t;
typedef struct pnanovdb_compiler_t pnanovdb_compiler_t;
typedef struct pnanovdb_raster_context_t pnanovdb_raster_context_t;
typedef struct pnanovdb_raster_gaussian_data_t pnanovdb_raster_gaussian_data_t;
typedef struct pnanovdb_grid_build_context_t pnanovdb_grid_build_context_t;
typedef struct pnanovdb_parallel_primitives_context_t pnanovdb_parallel_primitives_context_t;
typedef struct pnanovdb_camera_mat_t pnanovdb_camera_mat_t;
typedef struct pnanovdb_raster_shader_params_t pnanovdb_raster_shader_params_t;
/* nanovdb_editor/putil/Compute.h:546-551: one timed entry of a profiler report; times in seconds.
 * The raster entry points that take a pnanovdb_profiler_report_t call it once per raster with one
 * entry per pipeline stage (GPU time from HIP events) and a total (host wall time and GPU time),
 * passing "raster" as userdata as Raster.cpp:583 arranges. */
typedef struct pnanovdb_compute_profiler_entry_t
{
    const char* label;
    float cpu_delta_time;
    float gpu_delta_time;
} pnanovdb_compute_profiler_entry_t;
typedef void(PNANOVDB_ABI* pnanovdb_profiler_report_t)(void* userdata, pnanovdb_uint64_t capture_id,
                                                       pnanovdb_uint32_t num_entries,
                                                       pnanovdb_compute_profiler_entry_t* entries);
typedef void(PNANOVDB_ABI* pnanovdb_compute_log_print_t)(pnanovdb_uint32_t level, const char* format, ...);

/* nanovdb_editor/putil/Compute.h:734-747 (4 pointers) */
typedef struct pnanovdb_compute_shader_interface_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    void* create_shader;
    void* map_shader_build;
    void* destroy_shader;
} pnanovdb_compute_shader_interface_t;

/* nanovdb_editor/putil/Compute.h:566-664 (1 + 33 pointers). This library keeps every slot
 * NULL: there is no Vulkan device manager; the HIP device is chosen by vdbr_context_create. */
typedef struct pnanovdb_compute_device_interface_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    void* create_device_manager;
    void* destroy_device_manager;
    void* enumerate_devices;
    void* create_device;
    void* destroy_device;
    void* get_memory_stats;
    void* create_semaphore;
    void* destroy_semaphore;
    void* get_semaphore_external_handle;
    void* close_semaphore_external_handle;
    void* get_device_queue;
    void* get_compute_queue;
    void* flush;
    void* get_frame_global_completed;
    void* wait_for_frame;
    void* wait_idle;
    void* get_compute_interface;
    void* get_compute_context;
    void* create_swapchain;
    void* destroy_swapchain;
    void* resize_swapchain;
    void* present_swapchain;
    void* get_swapchain_front_texture;
    void* create_encoder;
    void* destroy_encoder;
    void* present_encoder;
    void* get_encoder_front_texture;
    void* map_encoder_data;
    void* unmap_encoder_data;
    void* enable_profiler;
    void* disable_profiler;
    void* set_resource_min_lifetime;
    void* get_device_index;
} pnanovdb_compute_device_interface_t;

typedef struct pnanovdb_compute_t pnanovdb_compute_t;

/* nanovdb_editor/putil/Compute.h:784-853. Implemented: create_array, destroy_array,
 * map_array, unmap_array, duplicate_array, compute_array_print_range (Compute.cpp:707-760
 * semantics: zero-filled host storage of max(size, 64 KiB)). Shader and Vulkan entries are NULL. */
struct pnanovdb_compute_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    void* module;
    const pnanovdb_compiler_t* compiler;
    pnanovdb_compute_shader_interface_t shader_interface;
    pnanovdb_compute_device_interface_t device_interface;
    void* load_nanovdb;
    void* save_nanovdb;
    void* create_shader_context;
    void* destroy_shader_context;
    void* init_shader;
    void* destroy_shader;
    void* dispatch_shader;
    void* dispatch_shader_on_array;
    void* dispatch_shader_on_nanovdb_array;
    pnanovdb_compute_array_t*(PNANOVDB_ABI* create_array)(size_t element_size, pnanovdb_uint64_t element_count,
                                                          const void* data);
    void(PNANOVDB_ABI* destroy_array)(pnanovdb_compute_array_t* array);
    void*(PNANOVDB_ABI* map_array)(pnanovdb_compute_array_t* array);
    void(PNANOVDB_ABI* unmap_array)(pnanovdb_compute_array_t* array);
    void(PNANOVDB_ABI* compute_array_print_range)(const pnanovdb_compute_t* compute,
                                                  pnanovdb_compute_log_print_t log_print, const char* name,
                                                  pnanovdb_compute_array_t* arr, pnanovdb_uint32_t channel_count);
    void* nanovdb_from_image_rgba8;
    pnanovdb_compute_array_t*(PNANOVDB_ABI* duplicate_array)(pnanovdb_compute_array_t* array);
};

/* nanovdb_editor/putil/Raster.h:58-174 (field order is the ABI). raster_gaussian_2d and
 * create_gaussian_data_from_desc are NULL (outside the raster3d path). raster_file
 * (RasterUtils.cpp:295-395) reads a Gaussian PLY or NPZ (see pnanovdb_fileformat_t) and rasterizes
 * it as raster_to_nanovdb_from_arrays does, or builds gaussian data when gaussian_data and
 * raster_context are given; the sparse-voxel file layout the reference leaves unimplemented fails. */
typedef struct pnanovdb_raster_t pnanovdb_raster_t;
struct pnanovdb_raster_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    const pnanovdb_compute_t* compute;
    pnanovdb_raster_context_t*(PNANOVDB_ABI* create_context)(const pnanovdb_compute_t* compute,
                                                             pnanovdb_compute_queue_t* queue);
    void(PNANOVDB_ABI* destroy_context)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_raster_context_t* context);
    pnanovdb_raster_gaussian_data_t*(PNANOVDB_ABI* create_gaussian_data)(
        const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, pnanovdb_raster_context_t* context,
        pnanovdb_compute_array_t* means, pnanovdb_compute_array_t* quaternions, pnanovdb_compute_array_t* scales,
        pnanovdb_compute_array_t* colors, pnanovdb_compute_array_t* sh_0, pnanovdb_compute_array_t* sh_n,
        pnanovdb_compute_array_t* opacities, pnanovdb_compute_array_t** shader_params_arrays,
        pnanovdb_raster_shader_params_t* raster_params);
    void(PNANOVDB_ABI* upload_gaussian_data)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                             pnanovdb_raster_context_t* context, pnanovdb_raster_gaussian_data_t* data);
    void(PNANOVDB_ABI* destroy_gaussian_data)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                              pnanovdb_raster_gaussian_data_t* data);
    void* raster_gaussian_2d;
    void(PNANOVDB_ABI* raster_gaussian_3d)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_raster_context_t* context, float voxel_size,
                                           pnanovdb_raster_gaussian_data_t* data, pnanovdb_compute_buffer_t* nanovdb_out,
                                           pnanovdb_uint64_t nanovdb_word_count, void* userdata);
    pnanovdb_compute_array_t*(PNANOVDB_ABI* raster_to_nanovdb)(
        const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, float voxel_size,
        pnanovdb_compute_array_t* means, pnanovdb_compute_array_t* quaternions, pnanovdb_compute_array_t* scales,
        pnanovdb_compute_array_t* colors, pnanovdb_compute_array_t* sh_0, pnanovdb_compute_array_t* sh_n,
        pnanovdb_compute_array_t* opacities, pnanovdb_compute_array_t** shader_params_arrays,
        pnanovdb_profiler_report_t profiler_report, void* userdata);
    pnanovdb_bool_t(PNANOVDB_ABI* raster_file)(pnanovdb_raster_t* raster, const pnanovdb_compute_t* compute,
                                               pnanovdb_compute_queue_t* queue, const char* filename, float voxel_size,
                                               pnanovdb_compute_array_t** nanovdb_arr,
                                               pnanovdb_raster_gaussian_data_t** gaussian_data,
                                               pnanovdb_raster_context_t** raster_context,
                                               pnanovdb_compute_array_t** shader_params_arrays,
                                               pnanovdb_raster_shader_params_t* raster_params,
                                               pnanovdb_profiler_report_t profiler_report, void* userdata);
    pnanovdb_bool_t(PNANOVDB_ABI* raster_to_nanovdb_from_arrays)(pnanovdb_raster_t* raster,
                                                                 const pnanovdb_compute_t* compute,
                                                                 pnanovdb_compute_queue_t* queue, float voxel_size,
                                                                 pnanovdb_compute_array_t** arrays_gaussian,
                                                                 pnanovdb_uint32_t array_count,
                                                                 pnanovdb_compute_array_t** out_nanovdb_arr);
    pnanovdb_bool_t(PNANOVDB_ABI* create_gaussian_data_from_arrays)(
        pnanovdb_raster_t* raster, const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
        pnanovdb_compute_array_t** arrays_gaussian, pnanovdb_uint32_t array_count,
        pnanovdb_raster_gaussian_data_t** gaussian_data, pnanovdb_raster_shader_params_t* raster_params,
        pnanovdb_raster_context_t** raster_context);
    void* create_gaussian_data_from_desc;
};

/* nanovdb_editor/putil/ParallelPrimitives.h:22-88. Buffers are vdbr_buffer_t*. Scans are
 * inclusive; sorts are stable, ascending, LSD over the low key_bit_count bits; key_count 0 is
 * a no-op. radix_sort_dual_key orders by (high key, low key), each over its own bit count. */
typedef struct pnanovdb_parallel_primitives_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    pnanovdb_parallel_primitives_context_t*(PNANOVDB_ABI* create_context)(const pnanovdb_compute_t* compute,
                                                                          pnanovdb_compute_queue_t* queue);
    void(PNANOVDB_ABI* destroy_context)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_parallel_primitives_context_t* context);
    void(PNANOVDB_ABI* global_scan)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                    pnanovdb_parallel_primitives_context_t* context, pnanovdb_compute_buffer_t* val_in,
                                    pnanovdb_compute_buffer_t* val_out, pnanovdb_uint64_t val_count,
                                    pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* global_scan_uint64)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_parallel_primitives_context_t* context,
                                           pnanovdb_compute_buffer_t* val_in, pnanovdb_compute_buffer_t* val_out,
                                           pnanovdb_uint64_t val_count, pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* global_scan_max)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_parallel_primitives_context_t* context,
                                        pnanovdb_compute_buffer_t* val_in, pnanovdb_compute_buffer_t* val_out,
                                        pnanovdb_uint64_t val_count, pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* radix_sort)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                   pnanovdb_parallel_primitives_context_t* context_in,
                                   pnanovdb_compute_buffer_t* key_inout, pnanovdb_compute_buffer_t* val_inout,
                                   pnanovdb_uint64_t key_count, pnanovdb_uint64_t buffer_key_count,
                                   pnanovdb_uint32_t key_bit_count);
    void(PNANOVDB_ABI* radix_sort_dual_key)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                            pnanovdb_parallel_primitives_context_t* context_in,
                                            pnanovdb_compute_buffer_t* key_low_inout,
                                            pnanovdb_compute_buffer_t* key_high_inout,
                                            pnanovdb_compute_buffer_t* val_inout, pnanovdb_uint64_t key_count,
                                            pnanovdb_uint64_t buffer_key_count, pnanovdb_uint32_t key_low_bit_count,
                                            pnanovdb_uint32_t key_high_bit_count);
    void(PNANOVDB_ABI* radix_sort_key64)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                         pnanovdb_parallel_primitives_context_t* context_in,
                                         pnanovdb_compute_buffer_t* key_inout, pnanovdb_compute_buffer_t* val_inout,
                                         pnanovdb_uint64_t key_count, pnanovdb_uint64_t buffer_key_count,
                                         pnanovdb_uint32_t key_bit_count);
    const pnanovdb_compute_t* compute;
} pnanovdb_parallel_primitives_t;

/* Getters: nanovdb_editor/putil/Raster.h:197, Compute.h:883, ParallelPrimitives.h:107,
 * GridBuild.h:176, FileFormat.h. Each returns a writable static table that lives as long as the
 * library; the compute and fileformat tables carry the library's own handle in `module`, so hosts
 * resolve the other getters through it (putil/Raster.h:199-210, GridBuild.h:178-181). */
/* nanovdb_editor/putil/FileFormat.h:19-35 (without the optional e57 entry). Loads Gaussian
 * arrays by name from binary PLY (3DGS properties: x y z, opacity, rot_0..3, scale_0..2,
 * f_dc_0..2, f_rest_*, red green blue) or NPZ (members <name>.npy; "sh" [N, K, 3] stands for sh_0
 * and sh_n) into arrays created with the compute table's create_array (free them with its
 * destroy_array). Names: means/positions, opacities, quaternions, scales, sh_0, sh_n, colors.
 * can_load_file accepts .ply and .npz. */
typedef struct pnanovdb_fileformat_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    pnanovdb_bool_t(PNANOVDB_ABI* can_load_file)(const char* filename);
    pnanovdb_bool_t(PNANOVDB_ABI* load_file)(const char* filename, pnanovdb_uint32_t array_count,
                                             const char** array_names, pnanovdb_compute_array_t** out_arrays);
    void* module;
} pnanovdb_fileformat_t;

/* nanovdb_editor/putil/GridBuild.h:22-29. The buffers are this library's device buffers
 * (vdbr_buffer_t*); node_addresses_buffer holds the build's own accumulation state between
 * grid_build_init and grid_build_destroy, the other slots stay NULL. */
typedef struct pnanovdb_grid_build_state_t
{
    pnanovdb_compute_buffer_t* node_addresses_buffer;
    pnanovdb_compute_buffer_t* node_types_buffer;
    pnanovdb_compute_buffer_t* new_child_counts_buffer;
    pnanovdb_compute_buffer_t* scan_new_child_counts_buffer;
    pnanovdb_compute_buffer_t* bboxes_buffer;
} pnanovdb_grid_build_state_t;

/* nanovdb_editor/putil/GridBuild.h:31-58 (field order is the ABI). fanout_state_init allocates the
 * buffers as vdbr_buffer_t device buffers of the reference's sizes: fragment_counts_buffer and
 * scan_fragment_counts_buffer (u64 per prim of a batch; the caller writes each prim's footprint,
 * 0 past the prim count), prim_bbox_buffer (6 int32 per prim, for the caller), prim_idxs_buffer and
 * prim_raster_idxs_buffer (u32 per window slot: the slot's prim within the batch and its index in
 * that prim's footprint, ~0 for slots past the batch's fragments), prim_raster_ijks_buffer (3 int32
 * per slot for the caller's voxel coordinates, initialised to 0x7FFFFFFF = skip) and
 * scatter_prim_idxs_buffer (unused here). readback_buffer is host memory holding each batch's
 * fragment total (read back synchronously: ijk_batch_count is exact from a batch's first window). */
typedef struct pnanovdb_grid_build_fanout_state_t
{
    pnanovdb_uint32_t prim_batch_size;
    pnanovdb_uint32_t ijk_batch_size;
    pnanovdb_uint32_t prim_count;
    float voxel_size;

    pnanovdb_uint32_t prim_batch_idx;
    pnanovdb_uint32_t prim_batch_count;
    pnanovdb_uint32_t ijk_batch_idx;
    pnanovdb_uint32_t ijk_batch_count;
    pnanovdb_uint32_t ijk_batch_count_max;

    pnanovdb_uint64_t total_fragment_count;

    pnanovdb_compute_buffer_t* fragment_counts_buffer;
    pnanovdb_compute_buffer_t* scan_fragment_counts_buffer;
    pnanovdb_compute_buffer_t* prim_bbox_buffer;

    pnanovdb_compute_buffer_t* scatter_prim_idxs_buffer;
    pnanovdb_compute_buffer_t* prim_idxs_buffer;
    pnanovdb_compute_buffer_t* prim_raster_idxs_buffer;
    pnanovdb_compute_buffer_t* prim_raster_ijks_buffer;

    pnanovdb_compute_buffer_t* readback_buffer;
    pnanovdb_uint64_t* readback_frames;
    pnanovdb_uint64_t readback_frame_count;
} pnanovdb_grid_build_fanout_state_t;

/* nanovdb_editor/putil/GridBuild.h:60-152 (field order is the ABI). Semantics as GridBuild.cpp:
 *  - grid_build_init / grid_build_destroy: the build state; grid_build_reset writes an empty node2
 *    grid (header, root, blind metadata) into nanovdb_out and forgets every inserted point.
 *  - grid_build: one fragment window of point_count int32 (i, j, k) points (point_ijk_in); points
 *    whose k is 0x7FFFFFFF are skipped (grid_build_set_masks.slang:52). Coordinates must lie in
 *    [-65536, 65535]. Each call is one window of the node allocation order (GridBuild.cpp:361-506:
 *    window, level, parent, slot). nanovdb_out keeps the empty grid until grid_build_finalize.
 *  - grid_build_finalize: writes the grid of every point inserted since the reset into nanovdb_out
 *    (nanovdb_word_count 32-bit words; a larger grid is an error, nothing is written), with value
 *    indices in node order and blind metadata (bboxes, RGBA8 values, all values 0 as the reference
 *    leaves them for its colour pass) as GridBuild.cpp:508-650.
 *  - fanout_*: the prim batch x fragment window loop (GridBuild.cpp:654-942): after the caller has
 *    written the batch's fragment counts, fanout maps every slot of the current window to (prim,
 *    footprint index). Calls are synchronous. */
typedef struct pnanovdb_grid_build_t
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    pnanovdb_grid_build_context_t*(PNANOVDB_ABI* create_context)(const pnanovdb_compute_t* compute,
                                                                 pnanovdb_compute_queue_t* queue);
    void(PNANOVDB_ABI* destroy_context)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_grid_build_context_t* context);
    void(PNANOVDB_ABI* grid_build_init)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                        pnanovdb_compute_buffer_t* nanovdb_out, pnanovdb_uint64_t nanovdb_word_count,
                                        float voxel_size, pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* grid_build_reset)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                         pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                         pnanovdb_compute_buffer_t* nanovdb_out, pnanovdb_uint64_t nanovdb_word_count,
                                         float voxel_size, pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* grid_build_destroy)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_grid_build_state_t* state);
    void(PNANOVDB_ABI* grid_build)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                   pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                   pnanovdb_compute_buffer_t* point_ijk_in, pnanovdb_compute_buffer_t* nanovdb_out,
                                   pnanovdb_uint64_t point_count, pnanovdb_uint64_t nanovdb_word_count, float voxel_size,
                                   pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* grid_build_finalize)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                            pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_state_t* state,
                                            pnanovdb_compute_buffer_t* point_ijk_in, pnanovdb_compute_buffer_t* nanovdb_out,
                                            pnanovdb_uint64_t point_count, pnanovdb_uint64_t nanovdb_word_count,
                                            float voxel_size, pnanovdb_uint32_t dispatch_count);
    void(PNANOVDB_ABI* fanout_state_reset)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_grid_build_context_t* context,
                                           pnanovdb_grid_build_fanout_state_t* state);
    pnanovdb_bool_t(PNANOVDB_ABI* fanout_state_valid)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                                      pnanovdb_grid_build_context_t* context,
                                                      pnanovdb_grid_build_fanout_state_t* state);
    void(PNANOVDB_ABI* fanout_state_increment)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                               pnanovdb_grid_build_context_t* context,
                                               pnanovdb_grid_build_fanout_state_t* state);
    void(PNANOVDB_ABI* fanout_state_init)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                          pnanovdb_grid_build_context_t* context,
                                          pnanovdb_grid_build_fanout_state_t* state, pnanovdb_uint32_t prim_batch_size,
                                          pnanovdb_uint32_t ijk_batch_size, pnanovdb_uint32_t ijk_batch_count_max,
                                          pnanovdb_uint32_t prim_count, float voxel_size);
    void(PNANOVDB_ABI* fanout_state_destroy)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                             pnanovdb_grid_build_fanout_state_t* state);
    void(PNANOVDB_ABI* fanout)(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                               pnanovdb_grid_build_context_t* context, pnanovdb_grid_build_fanout_state_t* state,
                               pnanovdb_uint32_t dispatch_count);
    const pnanovdb_compute_t* compute;
} pnanovdb_grid_build_t;

PNANOVDB_API pnanovdb_fileformat_t* pnanovdb_get_fileformat(void);
/* nanovdb_editor/putil/GridBuild.h:176 */
PNANOVDB_API pnanovdb_grid_build_t* pnanovdb_get_grid_build(void);
PNANOVDB_API pnanovdb_raster_t* pnanovdb_get_raster(void);
PNANOVDB_API pnanovdb_compute_t* pnanovdb_get_compute(void);
PNANOVDB_API pnanovdb_parallel_primitives_t* pnanovdb_get_parallel_primitives(void);

/* ================================================================ direct device API
 * Return codes: 0 = success, negative = error (message from vdbr_last_error). Calls on one
 * context are synchronous with respect to the host unless stated otherwise, and run on the
 * context's stream. One context per host thread at a time. */
typedef struct vdbr_context_t vdbr_context_t;

/* device buffer handle; also the concrete type behind pnanovdb_compute_buffer_t */
typedef struct vdbr_buffer_t
{
    void* device_ptr;
    uint64_t size_bytes;
} vdbr_buffer_t;

PNANOVDB_API vdbr_context_t* vdbr_context_create(int device);
PNANOVDB_API void vdbr_context_destroy(vdbr_context_t* ctx);
PNANOVDB_API const char* vdbr_last_error(vdbr_context_t* ctx);
/* Run the context's work on a caller stream (hipStream_t passed as void*). NULL selects the
 * device's null stream, which orders with torch's default stream; a new context starts on a
 * non-blocking stream of its own (vdbr_context_get_stream returns it). Work already queued on
 * the previous stream is completed first. */
PNANOVDB_API int vdbr_context_set_stream(vdbr_context_t* ctx, void* stream);
PNANOVDB_API void* vdbr_context_get_stream(vdbr_context_t* ctx);
PNANOVDB_API int vdbr_synchronize(vdbr_context_t* ctx);
/* kind: 0 host->device, 1 device->host, 2 device->device */
PNANOVDB_API int vdbr_copy(vdbr_context_t* ctx, void* dst, const void* src, uint64_t bytes, int kind);
/* bytes of device memory currently held by the context: its workspace, its batch workers' and the
 * batch output arena chunks */
PNANOVDB_API uint64_t vdbr_workspace_bytes(vdbr_context_t* ctx);
/* Releases the context's workspace (and its batch workers'); device buffers handed out by earlier
 * calls become invalid. The next call allocates again. */
PNANOVDB_API int vdbr_context_trim(vdbr_context_t* ctx);
PNANOVDB_API int vdbr_release_workspace(vdbr_context_t* ctx);

/* Gaussian arrays on the device, float32, reference layouts: means[3N], opacities[N],
 * quats[4N] (w,x,y,z), scales[3N], colors[3N]. When VDBR_FLAG_RAW_INPUTS is set the arrays
 * hold opacity logits, unnormalised quaternions, log-scales and sh_0, and are preprocessed in
 * place exactly as RasterUtils.cpp:31-136 (colors receives 0.28209479*sh0 + 0.5; sh0 may
 * alias colors). */
typedef struct vdbr_gaussians_t
{
    float* means;
    float* opacities;
    float* quats;
    float* scales;
    float* colors;
    const float* sh0;
    uint64_t count;
} vdbr_gaussians_t;

/* Gradient outputs of vdbr_density_backward (device, float32, the layouts of vdbr_gaussians_t):
 * d/d means[3N], d/d opacity logits[N], d/d unnormalised quats[4N], d/d log-scales[3N]. */
typedef struct vdbr_gaussian_grads_t
{
    float* means;
    float* opacities;
    float* quats;
    float* scales;
} vdbr_gaussian_grads_t;

#define VDBR_VALUE_RGBA8 0u   /* gaussian_frag_color.slang semantics (packed RGBA8, max) */
#define VDBR_VALUE_DENSITY 1u /* sum_g o_g exp(-m_g/2) at voxel centres (exact sum of double terms), float32 */

#define VDBR_FLAG_RAW_INPUTS 1u
#define VDBR_FLAG_SPATIAL_ORDER 2u /* nodes in BFS tree order instead of the windowed order */
#define VDBR_FLAG_VOXEL_COORDS 4u  /* also record every value slot's voxel (vdbr_last_voxels) */
#define VDBR_FLAG_WIDE_SEGMENTS 8u /* 16-byte candidate-segment records even where 8-byte ones fit */

typedef struct vdbr_raster_params_t
{
    float voxel_size;
    float cull_scale;           /* 0 -> 0.05 (gaussian_frag_color.slang:205-208) */
    uint32_t value_mode;        /* VDBR_VALUE_* */
    uint32_t flags;             /* VDBR_FLAG_* */
    uint32_t prim_batch_size;   /* node-order model, 0 -> 262144 (Raster.cpp:443) */
    uint32_t ijk_batch_size;    /* 0 -> 8388608 (Raster.cpp:444) */
    uint32_t max_node_count;    /* 0 -> 2097152 (GridBuild.cpp:222) */
    uint32_t slab_pairs;        /* pairs per slab before the build is windowed over voxel x; 0 -> 2^31 */
} vdbr_raster_params_t;

typedef struct vdbr_raster_stats_t
{
    uint64_t gaussian_count;
    uint64_t fragment_count;   /* AABB candidates */
    uint64_t pair_count;       /* surviving (voxel, gaussian) pairs */
    uint64_t voxel_count;      /* active voxels */
    uint64_t grid_size;        /* bytes */
    uint32_t upper_count;
    uint32_t lower_count;
    uint32_t leaf_count;
    uint32_t window_count;     /* fragment windows of the node-order model */
    uint32_t key_bits;         /* bits of the sorted pair keys */
    uint32_t emit_attempts;    /* 1, or 2 when the pair buffer had to grow */
    float ms_setup, ms_emit, ms_sort, ms_reduce, ms_build, ms_total;
    uint32_t sort_passes;      /* onesweep passes of the pair sort */
    float ms_sort_passes;      /* time of the onesweep pass kernels only (HIP events) */
    float ms_emit_kernel;      /* emission: row pass, segment count readback and walk (HIP events) */
    float ms_reduce_kernel;    /* time of the per-voxel reduction kernel (HIP events) */
    uint32_t sort_bit_begin;   /* the pair sort orders key bits [sort_bit_begin, key_bits) */
    uint32_t slab_count;       /* voxel-x slabs the pairs were emitted, sorted and reduced in */
    uint64_t pair_list_id;     /* density mode, one slab: id of the pair list the context keeps for
                                  vdbr_density_backward_pairs; 0 otherwise */
    float ms_row_kernel;       /* k_row_segments alone: AABB rows -> candidate segments (HIP events) */
    float ms_walk_kernel;      /* k_emit_segments alone: predicate walk, staging, key writes (HIP events) */
    uint64_t candidate_voxels; /* voxels of the candidate segments the walk evaluates the predicate on */
} vdbr_raster_stats_t;

/* Full raster3d pipeline on device inputs. The node2 grid is written to a context-owned
 * device buffer (*grid_dev, valid until the next raster call on the context); *grid_size
 * receives its size in bytes. */
PNANOVDB_API int vdbr_raster_node2(vdbr_context_t* ctx, const vdbr_gaussians_t* gaussians,
                                   const vdbr_raster_params_t* params, void** grid_dev, uint64_t* grid_size,
                                   vdbr_raster_stats_t* stats);

/* One cloud of a batched call: inputs and parameters as vdbr_raster_node2 takes them; the call
 * fills grid_dev/grid_size (a node2 grid in context-owned device memory, valid until the next
 * batched call on the same context) and stats. */
typedef struct vdbr_batch_item_t
{
    vdbr_gaussians_t gaussians;
    vdbr_raster_params_t params;
    void* grid_dev;
    uint64_t grid_size;
    vdbr_raster_stats_t stats;
} vdbr_batch_item_t;

/* Batched raster3d (no counterpart in the reference, whose entry point takes one cloud): runs
 * every cloud of items[count] through the full pipeline, each result identical to a
 * vdbr_raster_node2 call on that cloud alone. The per-Gaussian setup of every cloud runs first
 * with one read back of its counts; from those the call bounds each cloud's buffers and captures
 * the remaining stages of all clouds (emission, sort, reduce, grid build, the grid copied into a
 * context-owned output arena) as HIP graphs, one per worker stream (`workers`: 0 -> 4, at most
 * 16; each with its own workspace, clouds balanced by their pair estimates), launched together:
 * no count is read back between stages. A later call with the same clouds, plans and workspaces
 * relaunches the captured graphs. Clouds the capture does not take (more than one slab, no
 * fragments, invalid input) and clouds whose counts exceeded their bounds run after the graphs
 * through vdbr_raster_node2. The work is ordered after the context's stream and the context's
 * stream after it. On failure the message names the first failing cloud. */
PNANOVDB_API int vdbr_raster_node2_batch(vdbr_context_t* ctx, vdbr_batch_item_t* items, uint32_t count,
                                         uint32_t workers);

/* Clouds of the last vdbr_raster_node2_batch call on ctx that ran inside its captured graphs
 * (no counterpart in the reference). */
PNANOVDB_API uint32_t vdbr_last_batch_graph_clouds(vdbr_context_t* ctx);

/* Host-clock phases of the last vdbr_raster_node2_batch call, in milliseconds (no counterpart in
 * the reference): ms[0] set-up of every cloud and its read back, ms[1] capture and instantiation
 * of the graphs (close to 0 when the graphs of an earlier call with the same signature -- the
 * same clouds, plans and workspaces -- were launched again), ms[2] graph launches to completion,
 * ms[3] 1 when the graphs were captured by this call, ms[4] clouds run outside the graphs. */
PNANOVDB_API int vdbr_last_batch_timing(vdbr_context_t* ctx, float* ms);

/* Adjoint of the density value mode (no counterpart in the reference, which has no backward
 * pass; the forward it differentiates is vdbr_raster_node2 with VDBR_VALUE_DENSITY):
 * rho(v) = sum_g o_g exp(-|diag(1/s_g) R_g^T (x_v - mu_g)|^2 / 2) over the pairs of voxel v.
 * raw holds the raw parameters of that forward call (means, opacity logits, unnormalised
 * quaternions, log-scales; colors and sh0 unused), grid_dev/grid_size its node2 grid, and
 * grad_values_dev[value_count] the gradient of the loss with respect to the grid's value array
 * (indexed by value index; slots 0 and 1 are the background). Gradients are computed in double
 * precision, summed in a fixed order (bitwise reproducible) and written as float32. */
PNANOVDB_API int vdbr_density_backward(vdbr_context_t* ctx, const vdbr_gaussians_t* raw, float voxel_size,
                                       const void* grid_dev, uint64_t grid_size, const float* grad_values_dev,
                                       uint64_t value_count, const vdbr_gaussian_grads_t* grads);

/* vdbr_density_backward from the sorted pair list that the density forward kept on this context
 * (no counterpart in the reference): pair_list_id is the forward's stats.pair_list_id. The pair
 * terms are evaluated in double precision, rounded to 48-bit fixed point under a power-of-two
 * scale taken from max |grad_values| and summed as integers (bitwise reproducible, independent of
 * the pair order); the gradients agree with vdbr_density_backward to double rounding. Returns 0
 * when the pair list was used and 1 when it was gone (another raster call or sort autotune on the
 * context since the forward, a forward of more than one slab, another Gaussian count or voxel
 * size) or grad_values holds a non-finite value: then vdbr_density_backward ran instead. The
 * grid must be the forward's. */
PNANOVDB_API int vdbr_density_backward_pairs(vdbr_context_t* ctx, const vdbr_gaussians_t* raw, float voxel_size,
                                             const void* grid_dev, uint64_t grid_size, const float* grad_values_dev,
                                             uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                                             uint64_t pair_list_id);

/* The surviving pair list of the last vdbr_raster_node2 call (its last slab), in the order the
 * value pass consumed it: grouped by leaf (leaves in key order, pairs inside a leaf unordered);
 * voxel ijk (int32 x3) and gaussian id per pair, copied to device buffers of the caller. Returns the pair count through *count (buffers may be NULL
 * to query it). */
PNANOVDB_API int vdbr_last_pairs(vdbr_context_t* ctx, int32_t* ijk_dev, uint32_t* gid_dev, uint64_t cap,
                                 uint64_t* count);

/* Voxel coordinates (int32 i, j, k per value slot, value-index order; slots 0 and 1, the
 * background, hold INT32_MIN-like sentinels) of the last vdbr_raster_node2 call made with
 * VDBR_FLAG_VOXEL_COORDS, copied to ijk_dev[3 * cap]; *count receives the value count. No
 * counterpart in the reference (its consumers walk the tree); used to evaluate losses between grids. */
PNANOVDB_API int vdbr_last_voxels(vdbr_context_t* ctx, int32_t* ijk_dev, uint64_t cap, uint64_t* count);

/* Value words of a node2 grid at voxel coordinates: out[q] = values[value index of ijk[3q..3q+2]]
 * (PNanoVDBExt.h:225-318 descent), 0 where the voxel is inactive. values is the grid's value array
 * (value_count 32-bit words, e.g. the trailing blind data of a grid this library built). */
PNANOVDB_API int vdbr_node2_sample(vdbr_context_t* ctx, const void* grid_dev, uint64_t grid_size, const uint32_t* values,
                                   uint64_t value_count, const int32_t* ijk_dev, uint64_t count, uint32_t* out_dev);

/* Stable LSD radix sort of 64-bit keys over bits [bit_begin, bit_end), optional 32-bit
 * payload, in place (ParallelPrimitives.cpp:685-869 semantics). */
PNANOVDB_API int vdbr_radix_sort_u64(vdbr_context_t* ctx, uint64_t* keys, uint32_t* vals, uint64_t count,
                                     uint32_t bit_begin, uint32_t bit_end);
/* Allocation-free variant of vdbr_radix_sort_u64. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same sort, same result; the ping-pong buffers, look-back
 * status words and histograms live in a file-scope device array of the library (about 50 MiB,
 * one copy on every device the library runs on, held until the process ends), so the call makes
 * no hipMalloc/hipFree and allocates nothing else. count is at most 4194304 (2^22) keys; larger
 * counts fail with a message in vdbr_last_error. The work is queued on the context's stream and
 * the call returns without synchronizing. All calls on one device share the array: calls that
 * overlap on one device (two streams, two host threads, two contexts) race and must be ordered by
 * the caller. */
PNANOVDB_API int vdbr_radix_sort_u64_noalloc(vdbr_context_t* ctx, uint64_t* keys, uint32_t* vals, uint64_t count,
                                             uint32_t bit_begin, uint32_t bit_end);
/* Stable LSD radix sort of 32-bit keys with optional 32-bit payload, in place. */
PNANOVDB_API int vdbr_radix_sort_u32(vdbr_context_t* ctx, uint32_t* keys, uint32_t* vals, uint64_t count,
                                     uint32_t key_bit_count);
/* VoxelBVH Gaussian keys (raster/VoxelBVH.cpp:1225-1435, ijkl_from_gaussians): preprocessed device
 * arrays (means, opacities, quats (w,x,y,z), scales; colors and sh0 unused) -> world box of the
 * Gaussians' bounding cubes (world_bbox[6]: min xyz, max xyz), up to 8 cell keys per Gaussian
 * ((i << 48) | (j << 32) | (k << 16) | level, unused slots ~0) sorted ascending into
 * ijkl[8 N] with the Gaussian index (unused ~0) in prim_ids[8 N], and the [start, end) key index
 * of every run of equal keys in ranges[2 * 8 N] (entries past the run count are 0). All output
 * buffers are caller device memory. resolution is the voxel count along the box's longest axis
 * (1..4096, VoxelBVH.cpp:30-33); transform (optional, 16 floats, rows of a 4x4 matrix applied as
 * (p, 1) * M) maps Gaussian space to world space as the reference's transform_floats. */
PNANOVDB_API int vdbr_voxelbvh_ijkl(vdbr_context_t* ctx, const vdbr_gaussians_t* gaussians, uint32_t resolution,
                                    const float* transform, uint64_t* ijkl, uint32_t* prim_ids, uint32_t* ranges,
                                    float* world_bbox);
/* Stable sort by (key_high over its low key_high_bit_count bits, then key_low over its low
 * key_low_bit_count bits), both keys and the optional payload permuted in place
 * (ParallelPrimitives.cpp:527-683, radix_sort_dual_key; bit counts clamp to 32). */
PNANOVDB_API int vdbr_radix_sort_dual_key(vdbr_context_t* ctx, uint32_t* key_low, uint32_t* key_high, uint32_t* vals,
                                          uint64_t count, uint32_t key_low_bit_count, uint32_t key_high_bit_count);
/* Tile binning of projected Gaussians, the scan/sort stage of Raster2D (raster/Raster2D.cpp:347-535:
 * gaussian_count_tiles.slang, global_scan, gaussian_tile_intersections.slang,
 * radix_sort_dual_key(32, tile-id bits), gaussian_tile_offsets.slang). Device inputs: means2d[2N]
 * (pixels), radii[N] (pixels; <= 0 culls), depths[N]. Tiles are tile_size pixels square over a
 * width x height image (row-major tile index). Outputs: *count intersections; *keys_low (depth
 * bits), *keys_high (tile index) and *vals (Gaussian index) point to context-owned device arrays
 * sorted by (tile, depth bits) in Gaussian order within ties (valid until the next call);
 * tile_offsets (optional caller device buffer of tiles_w * tiles_h entries) receives the first
 * record of each tile (all 0 when there is no intersection). Float-to-tile conversions saturate
 * (NaN and negatives to 0), as the GPUs' conversion instruction does. */
PNANOVDB_API int vdbr_tile_bin(vdbr_context_t* ctx, const float* means2d, const int32_t* radii, const float* depths,
                               uint32_t count_in, uint32_t image_width, uint32_t image_height, uint32_t tile_size,
                               uint32_t* tile_offsets, uint32_t** keys_low, uint32_t** keys_high, uint32_t** vals,
                               uint64_t* count);
/* Inclusive scans (ParallelPrimitives.cpp:151-310): op 0 = add u32, 1 = add u64, 2 = max u32. */
PNANOVDB_API int vdbr_inclusive_scan(vdbr_context_t* ctx, const void* in, void* out, uint64_t count, int op);

/* putil/Convert.h:40-313 (pnanovdb_node2_convert_to_grid_type with PNANOVDB_GRID_TYPE_ONINDEX, as
 * called by RasterUtils.cpp:187-204): converts a node2 grid in device memory into a NanoVDB
 * OnIndex grid whose only blind data is the node2 value array. *onindex_dev points into the
 * context workspace (valid until the next call on ctx); *onindex_size is the OnIndex layout size
 * and *buffer_size (optional) the zero-padded length that also covers the grid size field the
 * OnIndex header copies from the node2 header. Returns 0 on success. */
PNANOVDB_API int vdbr_node2_to_onindex(vdbr_context_t* ctx, const void* node2_dev, uint64_t node2_size,
                                       void** onindex_dev, uint64_t* onindex_size, uint64_t* buffer_size);

/* Self-test of the RGBA8 alpha evaluation (no counterpart in the reference): `count` hashed
 * (opacity, radius, sample distance, pow exponent) inputs from `seed`; *mismatches receives the
 * number of inputs where the fast decision differs from the correctly rounded chain of
 * gaussian_frag_color.slang:197-212, *fast_decided (optional) how many the fast path decided. */
PNANOVDB_API int vdbr_selftest_alpha8(vdbr_context_t* ctx, uint64_t count, uint64_t seed, uint64_t* mismatches,
                                      uint64_t* fast_decided);

/* One autotuner measurement: onesweep tile shape (variant, threads per block, keys per thread),
 * digit width in bits, and the best of two full sorts in milliseconds. */
typedef struct vdbr_tune_result_t
{
    uint32_t variant;
    uint32_t threads;
    uint32_t keys_per_thread;
    uint32_t digit_bits;
    float ms;
} vdbr_tune_result_t;

/* Sort autotuner (no counterpart in the reference, whose radix sort has fixed shapes): times
 * every tile shape x digit width on key_count hashed keys of key_bits bits, keeps the fastest
 * for sorts of key_count keys or more on this context (every configuration produces the same
 * stable order), and reports all measurements (up to cap) and the choice. */
PNANOVDB_API int vdbr_autotune_sort(vdbr_context_t* ctx, uint64_t key_count, uint32_t key_bits,
                                    vdbr_tune_result_t* results, uint32_t cap, uint32_t* count,
                                    uint32_t* best_variant, uint32_t* best_digit_bits);

/* Fixes the sort configuration for sorts of from_count keys or more (variant index as reported
 * by vdbr_autotune_sort, digit width 6..9; 9-bit digits need a shape of at least 512 threads). */
PNANOVDB_API int vdbr_set_sort_config(vdbr_context_t* ctx, uint64_t from_count, uint32_t variant,
                                      uint32_t digit_bits);

#ifdef __cplusplus
}
#endif

#endif
