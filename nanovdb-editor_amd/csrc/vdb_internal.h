// vdb_internal.h -- host-side context, workspace and internal entry points of libvdbraster.
#ifndef VDB_INTERNAL_H
#define VDB_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/pnanovdb_amd.h"
#include "vdb_count.h"
#include "vdb_keylayout.h"

namespace vdbr
{

// workspace slots: each slot is one growable device allocation reused across calls
enum WsSlot
{
    WS_FRAMES,
    WS_BOXES,
    WS_VALS,
    WS_FOOT,
    WS_FOOT_SCAN,
    WS_ROWS,
    WS_ROW_SCAN,
    WS_MISC,
    WS_PAIRS,
    WS_SORT_ALT,
    WS_SORT_VALS_ALT,
    WS_SORT_STATUS,
    WS_SORT_SMALL,
    WS_SCAN_PARTIALS,
    WS_SEG_TILES,
    WS_VOX_KEY,
    WS_VOX_VAL,
    WS_VOX_T,
    WS_LEAF_FIRST,
    WS_LEAF_INFO,
    WS_LEAF_MASK,
    WS_LEAF_SORT_KEYS,
    WS_LEAF_SORT_VALS,
    WS_LOWER,
    WS_LOWER_MASK,
    WS_UPPER,
    WS_UPPER_MASK,
    WS_NODE_KEYS,
    WS_NODE_VALS,
    WS_NODE_BYIDX,
    WS_NODE_SCAN,
    WS_HIST,
    WS_GRID,
    WS_BATCH,
    WS_RAW_COPY,
    WS_AUX_ALT,
    WS_AUX_VALS_ALT,
    WS_CONV_ROOT,
    WS_CONV_UPPERS,
    WS_CONV_FLAGS,
    WS_CONV_LOWERS,
    WS_CONV_FLAGS2,
    WS_CONV_LEAVES,
    WS_ONINDEX,
    WS_SEGS,
    WS_SEG_META,
    WS_TUNE_KEYS,
    WS_DUAL_WIDE,
    WS_TILE_SCAN,
    WS_VOX_IJK,
    WS_VOX_G0,
    WS_VOX_GROW,
    WS_SLAB_HIST,
    WS_SLAB_ROWS,
    WS_BVH_PARTIAL,
    WS_BVH_STARTS,
    WS_TILE_KEYS,
    WS_TILE_OUT,
    WS_DUAL_TMP,
    WS_BWD_ACC,
    WS_BWD_REC,
    WS_BWD_RECGID,
    WS_BWD_LEAVES,
    WS_BWD_LEAFGRAD,
    WS_BWD_PRIMS,
    WS_DENS_PRIMS,
    WS_LR_VAL,
    WS_LR_G0,
    WS_LR_KEY,
    WS_LR_CNT,
    WS_LR_STARTS, // first pair of every leaf of the sorted pairs (density reduce)
    WS_LR_ITEMS,  // exclusive scan of the density reduce's chunks per leaf
    WS_GB_KEYS,
    WS_GB_FLAGS,
    WS_GB_ENC,
    WS_GB_T,
    WS_GB_COUNTS, // node counts of the device-count build
    WS_GB_INFO,   // grid layout of the device-count build
    WS_GRAPH,     // count words and flags of the device-count raster pipeline
    WS_ROWFRAMES, // per-Gaussian row-interval constants of the row pass
    WS_BATCH_SETUP, // per-cloud setup buffers of a graph-captured batch
    WS_BATCH_TABLE, // per-cloud window bases and output records of a graph-captured batch
    WS_BATCH_ARENA, // output grids of a graph-captured batch
    WS_LEAF_MAP,    // candidate-leaf bitmap of the last raster call (rank-keyed pairs)
    WS_LEAF_PREFIX, // candidate leaves before each bitmap word (and the per-word counts)
    WS_LEAF_RANK,   // leaf index of each candidate-leaf rank
    WS_COUNT
};

struct SortWorkspace
{
    uint64_t* keys_alt;
    uint32_t* vals_alt;
    uint64_t* status;
    uint64_t* hist;
    uint64_t* starts;
    uint32_t* tile_counters;
};

struct SortTiming
{
    float ms;
};

// onesweep configuration: tile shape variant (threads x keys per thread) and digit width
struct SortConfig
{
    uint32_t variant;
    uint32_t digit_bits;
};
constexpr uint32_t SORT_BUCKETS = 40; // per floor(log2(key count))
// radix digits of 6 to 9 bits; histograms, digit starts and look-back status words are laid out
// with SORT_BINS entries per pass (per tile)
constexpr uint32_t SORT_MIN_DIGIT = 6;
constexpr uint32_t SORT_MAX_DIGIT = 9;
constexpr uint32_t SORT_BINS = 1u << SORT_MAX_DIGIT;

struct Context
{
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string error;
    void* ws_ptr[WS_COUNT] = {};
    size_t ws_size[WS_COUNT] = {};
    uint64_t* pinned = nullptr; // small host readback area (pinned)
    // set while the context's stream is being captured into a graph: no allocation, no event,
    // no synchronisation may happen (Context::ws fails instead of growing a slot)
    bool capturing = false;
    // sizing pass of a capture (the recorded work is discarded, never launched): Context::ws only
    // records the largest request per slot and hands out placeholder addresses
    bool probe = false;
    size_t probe_size[WS_COUNT] = {};
    // pinned read back of a captured batch's setup counts (grown as needed)
    uint64_t* batch_pinned = nullptr;
    size_t batch_pinned_words = 0;
    uint32_t emit_blocks = 2048; // resident blocks of the persistent emission kernel (CUs x blocks per CU)
    SortConfig sort_tuned[SORT_BUCKETS] = {}; // autotuned sort configuration per size bucket
    uint32_t sort_epoch = 0;
    uint64_t sort_status_tiles = 0;

    // state of the last raster call (for vdbr_last_pairs)
    uint64_t last_pair_count = 0;
    uint64_t* last_pairs = nullptr;
    // pair-key layout of the last raster call (its rank tables live in the WS_LEAF_* slots)
    KeyLayout last_kl = {};
    // the retained pair list of the last density forward (single slab): its id (0: none; ids are
    // never reused on a context), Gaussian count and voxel size, checked by the pair-list adjoint
    uint64_t pair_gen = 0;
    uint64_t last_pair_id = 0;
    uint64_t last_pair_n = 0;
    float last_pair_vs = 0.f;
    int32_t* last_voxel_ijk = nullptr; // value-index-ordered voxel coordinates (VDBR_FLAG_VOXEL_COORDS)
    uint64_t last_value_count = 0;

    // batched launch: worker contexts (own streams and workspaces) and the output arena chunks
    std::vector<Context*> batch_workers;
    std::vector<std::pair<void*, size_t>> batch_chunks;
    uint32_t last_batch_graph_clouds = 0; // clouds of the last batched call that ran in its graph
    // instantiated graphs (one per worker stream) of the last captured batch and the signature
    // they were captured under
    std::vector<hipGraphExec_t> batch_execs;
    uint64_t batch_arena_hint = 0; // arena bytes the last batch's grids needed, with headroom
    std::vector<uint64_t> batch_sig;
    // host milliseconds of the last batched call: set-up and read back, capture and
    // instantiation, graph run, rerun clouds; [3] = 1 when the graph was captured anew
    float batch_ms[5] = { 0.f, 0.f, 0.f, 0.f, 0.f };

    // forget the retained pair list (its buffers are about to be reused)
    void drop_pairs()
    {
        last_pairs = nullptr;
        last_pair_count = 0;
        last_pair_id = 0;
        last_kl = KeyLayout{};
    }
    int fail(const std::string& msg)
    {
        error = msg;
        return -1;
    }
    void drop_batch_graphs();
    void* ws(WsSlot slot, size_t bytes);
    void ws_free(WsSlot slot); // releases one slot (after the stream has drained)
    void* scratch_scan(size_t bytes)
    {
        return ws(WS_SCAN_PARTIALS, bytes);
    }
    int sort_workspace(uint64_t n, bool with_vals, uint64_t tiles, bool aux, SortWorkspace* out);
    uint32_t next_sort_epoch(uint64_t tiles);
    uint64_t workspace_bytes() const;
    void release();
};

#define VDBR_CHECK(ctx, expr)                                                                                          \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t _e = (expr);                                                                                        \
        if (_e != hipSuccess)                                                                                          \
            return (ctx)->fail(std::string(#expr) + ": " + hipGetErrorString(_e));                                     \
    } while (0)

// vdb_prims.hip
int scan_u32(Context* ctx, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive, uint32_t* total_dev);
int scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, uint64_t n, bool inclusive, uint64_t* total_dev);
int scan_max_u32(Context* ctx, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive);
// the same scans over a count that may live on the device (launches sized by its bound)
int scan_u32(Context* ctx, const uint32_t* in, uint32_t* out, DCount n, bool inclusive, uint32_t* total_dev);
int scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, DCount n, bool inclusive, uint64_t* total_dev);
int radix_sort_u64(Context* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin, uint32_t bit_end,
                   SortTiming* timing);
int radix_sort_u64(Context* ctx, uint64_t* keys, uint32_t* vals, DCount n, uint32_t bit_begin, uint32_t bit_end);
// fills and copies by kernels (lengths may be device counts; safe inside captured work)
int fill_words(Context* ctx, uint32_t* p, DCount n, uint32_t v);
int fill_words(Context* ctx, uint64_t* p, DCount n, uint64_t v);
int copy_words(Context* ctx, const uint32_t* src, uint32_t* dst, DCount n);
int copy_words(Context* ctx, const uint64_t* src, uint64_t* dst, DCount n);
int radix_sort_dual_key(Context* ctx, uint32_t* key_low, uint32_t* key_high, uint32_t* vals, uint64_t n,
                        uint32_t low_bits, uint32_t high_bits);
int last_voxels(Context* ctx, int32_t* ijk_dev, uint64_t cap, uint64_t* count);
int node2_sample(Context* ctx, const void* grid, uint64_t grid_size, const uint32_t* values, uint64_t value_count,
                 const int32_t* ijk, uint64_t n, uint32_t* out);
// vdb_bvh.hip
int voxelbvh_ijkl(Context* ctx, const vdbr_gaussians_t* gs, uint32_t resolution, const float* transform,
                  uint64_t* ijkl, uint32_t* prim_ids, uint32_t* ranges, float* world_bbox);
// vdb_tiles.hip
int tile_bin(Context* ctx, const float* means2d, const int32_t* radii, const float* depths, uint32_t n,
             uint32_t width, uint32_t height, uint32_t tile_size, uint32_t* offsets, uint32_t** keys_low_out,
             uint32_t** keys_high_out, uint32_t** vals_out, uint64_t* count_out);
// sort autotuner (vdb_prims.hip)
int autotune_sort(Context* ctx, uint64_t n, uint32_t key_bits, vdbr_tune_result_t* results, uint32_t cap,
                  uint32_t* count, uint32_t* best_variant, uint32_t* best_digit_bits);
int set_sort_config(Context* ctx, uint64_t from_count, uint32_t variant, uint32_t digit_bits);
// free_order: keys equal in every sorted bit may end in any order (the first pass then ranks
// with per-lane LDS atomics instead of stable peer ranking)
int radix_sort_u64_noswap(Context* ctx, uint64_t* keys, DCount n, uint32_t bit_begin, uint32_t bit_end,
                          uint64_t** result_keys, hipEvent_t passes_begin, hipEvent_t passes_end, uint32_t* passes_out,
                          bool hist_ready = false, bool free_order = false);
// allocation-free sort (vdb_prims.hip): the stable LSD sort of radix_sort_u64 over a file-scope
// device scratch array of the library (at most noalloc_max_keys() keys); queued on st on device
// dev without synchronizing; calls that overlap on one device race on the shared scratch
int radix_sort_u64_noalloc(int dev, uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bit_begin, uint32_t bit_end,
                           hipStream_t st, std::string* err);
uint64_t noalloc_max_keys();
// digit width the sort of n keys will use, and the buffer that holds its digit histograms
// (passes x 256 counts, u64): a producer of the keys may fill it and pass hist_ready
// digit width the sort of n keys over `bits` key bits uses
uint32_t sort_digit_bits(const Context* ctx, uint64_t n, uint32_t bits);
unsigned long long* sort_hist_buffer(Context* ctx);

// vdb_raster.hip: the GridBuild stage shared by the raster path and the grid build table.
// vox_key holds nvox sorted unique voxel keys in the layout kl (leaf index << 9 | slot), vox_val
// their value words, vox_t their first-touch fragment windows (T windows); vbounds (device, 6
// int32: min xyz, max xyz) the voxel bounding box.
struct Node2BuildInput
{
    KeyLayout kl;
    uint64_t* vox_key;
    uint32_t* vox_val;
    uint32_t* vox_t;
    uint64_t nvox;
    const int32_t* vbounds;
    uint32_t T;
    uint32_t max_nodes;
    bool spatial_order;
    bool voxel_coords;
    uint32_t value_type;
    float voxel_size;
    // device-count form (graph capture): the voxel count lives in nvox_dev (nvox is its bound),
    // node counts stay on the device and every buffer is sized by the bounds below; a count past
    // its bound or past max_nodes sets a bit in *flags_dev and the grid is not valid
    const uint32_t* nvox_dev = nullptr;
    uint64_t leaf_cap = 0, lower_cap = 0, upper_cap = 0;
    uint32_t* flags_dev = nullptr;
};
struct Node2BuildOutput
{
    uint8_t* grid; // context workspace (WS_GRID), valid until the next build on the context
    uint64_t grid_size; // bound of the grid size in the device-count form
    uint64_t upper_count, lower_count, leaf_count; // bounds in the device-count form
    const uint64_t* grid_size_dev = nullptr;       // device-count form: the grid size word
    const uint32_t* counts_dev = nullptr;          // device-count form: upper, lower, leaf counts
};
// flag bits of the device-count pipeline (a count exceeded the bound its buffers were sized by)
constexpr uint32_t GF_SEGMENTS = 1u, GF_PAIRS = 2u, GF_LEAVES = 4u, GF_VOXELS = 8u, GF_NODES = 16u, GF_ARENA = 32u,
                   GF_RANGE = 64u;
int build_node2_grid(Context* ctx, const Node2BuildInput& in, Node2BuildOutput* out);
// workgroups of the persistent emission kernel resident on one CU (occupancy query; 0 on failure)
uint32_t emit_blocks_per_cu();

// vdb_raster.hip
int raster_node2(Context* ctx, const vdbr_gaussians_t* g, const vdbr_raster_params_t* p, void** grid_dev,
                 uint64_t* grid_size, vdbr_raster_stats_t* stats);
// Batched raster3d as one captured HIP graph: the setup of every cloud runs first with one read
// back of its counts; the remaining stages of all clouds (emission to grid, the grids copied into
// an output arena of the context) are captured on the worker contexts' streams and launched as
// one graph. Clouds that need several slabs or have no fragments, and clouds whose counts exceed
// the bounds their captured buffers were sized by, are listed in *rerun (grid_dev left null) for
// the caller to run through raster_node2; (*setup_done)[i] tells that cloud i's raw inputs were
// preprocessed in place already.
int raster_batch_graph(Context* ctx, const std::vector<Context*>& workers, vdbr_batch_item_t* items, uint32_t count,
                       std::vector<uint32_t>* rerun, std::vector<uint8_t>* setup_done);
int last_pairs(Context* ctx, int32_t* ijk_dev, uint32_t* gid_dev, uint64_t cap, uint64_t* count);
int preprocess_device(Context* ctx, uint64_t n, float* opacities, float* quats, float* scales, const float* sh0,
                      float* colors);
int radix_sort_u32(Context* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits);

// density adjoint (vdb_raster.hip): gradients of sum_v grad_values[value_idx(v)] * rho(v) with
// respect to the raw parameters; raw->scales are log-scales, raw->opacities logits, raw->quats
// unnormalised. grid is the node2 density grid of the forward pass on the same parameters.
int density_backward(Context* ctx, const vdbr_gaussians_t* raw, float vs, const void* grid, uint64_t grid_size,
                     const float* grad_values, uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                     float* ms_out);
// the same gradients from the sorted pair list that the density forward with id pair_list_id
// retained on this context; when that list is gone (another raster call, a multi-slab build, a
// different Gaussian count or voxel size) or grad_values holds a non-finite value, runs
// density_backward instead. *used_pairs tells which path ran.
int density_backward_pairs(Context* ctx, const vdbr_gaussians_t* raw, float vs, const void* grid, uint64_t grid_size,
                           const float* grad_values, uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                           uint64_t pair_list_id, float* ms_out, int* used_pairs);

// vdb_raster.hip: fast RGBA8 alpha decision against the correctly rounded path on hashed inputs
int selftest_alpha8(Context* ctx, uint64_t count, uint64_t seed, uint64_t* mismatches, uint64_t* decided);

// vdb_convert.hip: node2 -> NanoVDB OnIndex (putil/Convert.h:40-313). *out_size is the OnIndex
// layout size; *alloc_size (>= out_size and >= the grid size field copied from the node2 header,
// a multiple of 8) is the zero-padded buffer length. The result lives in the context workspace.
int node2_to_onindex(Context* ctx, const void* node2, uint64_t node2_size, void** out_dev, uint64_t* out_size,
                     uint64_t* alloc_size);

// vdb_abi.hip: the library's compute table, and messages routed as RasterUtils.cpp:163-186 does
// (the host compute interface's log_print when the compute table carries one, else stderr)
constexpr uint32_t LOG_LEVEL_ERROR = 0u; // PNANOVDB_COMPUTE_LOG_LEVEL_ERROR (putil/Compute.h)
constexpr uint32_t LOG_LEVEL_WARNING = 1u;
const pnanovdb_compute_t* compute_table();
void log_message(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, uint32_t level, const char* fmt,
                 ...);

} // namespace vdbr

#endif
