// vdb_convert.hip -- node2 grid -> NanoVDB OnIndex grid on the GPU (gfx950).
//
// Restates putil/Convert.h:40-313 (pnanovdb_node2_convert_to_grid_type with the OnIndex target,
// called by RasterUtils.cpp:187-204). The reference walks the node2 tree five times on the CPU
// and appends nodes in slot order: all root tiles, then the uppers in root-slot order, the lowers
// in (upper, slot) order and the leaves in (lower, slot) order. Here each level is one flag pass
// over the slots of the previous level's nodes, an exclusive scan that gives every child its
// allocation rank (and thus its address), and one write pass; the output layout is therefore the
// reference's, byte for byte.
//
// OnIndex node layouts (RootData / Tile / InternalData / LeafData<ValueOnIndex> of the public
// NanoVDB format, 32-byte data alignment), byte offsets:
//   root  : bbox min @0, bbox max @12, tile count @24, size 96; tiles follow (32 B each:
//           key @0, child offset from the root @8, state @16, value @24)
//   upper : value mask @32, child mask @4128, table @8256 (8 B per slot), size 270400
//   lower : value mask @32, child mask @544, table @1088, size 33856
//   leaf  : value mask @16, first value index @80, 9-bit packed prefix sums @88, size 96
// Internal table entries hold the child's byte offset from its parent (or a tile's value index).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vdb_internal.h"
#include "vdb_layout.h"
#include "vdb_node2.h"

namespace vdbr
{

namespace
{

constexpr uint32_t ONI_ROOT_SIZE = 96u;
constexpr uint32_t ONI_TILE_SIZE = 32u;
constexpr uint32_t ONI_UPPER_VM = 32u, ONI_UPPER_CM = 4128u, ONI_UPPER_TABLE = 8256u, ONI_UPPER_SIZE = 270400u;
constexpr uint32_t ONI_LOWER_VM = 32u, ONI_LOWER_CM = 544u, ONI_LOWER_TABLE = 1088u, ONI_LOWER_SIZE = 33856u;
constexpr uint32_t ONI_LEAF_VM = 16u, ONI_LEAF_TABLE = 80u, ONI_LEAF_SIZE = 96u;
constexpr uint32_t GRID_TYPE_ONINDEX = 20u;

constexpr int CONV_THREADS = 256;

struct ConvLevel
{
    const uint64_t* src;      // node2 grid words
    uint8_t* dst;             // OnIndex grid bytes
    const uint64_t* parents;  // node2 word index of each parent node (level order)
    uint64_t parent_count;
    uint32_t fanout;          // slots per parent
    uint32_t cmp_off, vmp_off, children_off; // node2 word offsets of the child / value prefix sums and children
    uint32_t prefix_to_mask;  // words from a prefix-sum array to its mask (fanout / 256)
    uint64_t dst_first;       // byte address of the first parent in the output
    uint32_t dst_parent_size; // parent node size in the output
    uint32_t dst_table;       // table offset within the parent
    uint64_t child_first;     // byte address of the first child in the output
    uint32_t child_size;      // child node size in the output
    const uint32_t* child_rank; // exclusive scan of the child flags
    uint32_t* flags;          // child flags (flag pass)
    uint64_t* children;       // out: node2 word index of each child, in allocation order
    uint64_t nwords;          // node2 grid length in 64-bit words
    uint64_t child_words;     // node2 size of a child node in words
    uint32_t* bad;            // set when a child offset leaves the grid
};

// flag pass: one thread per (parent, slot)
__global__ void k_conv_flags(ConvLevel L)
{
    uint64_t t = (uint64_t)blockIdx.x * CONV_THREADS + threadIdx.x;
    if (t >= L.parent_count * L.fanout)
        return;
    uint64_t p = t / L.fanout;
    uint32_t n = (uint32_t)(t % L.fanout);
    uint64_t node = L.parents[p];
    uint64_t cm = L.src[node + L.cmp_off + L.prefix_to_mask + (n >> 6u)];
    L.flags[t] = (uint32_t)((cm >> (n & 63u)) & 1ull);
}

// child list pass: node2 word index of every child, in allocation (parent, slot) order
__global__ void k_conv_children(ConvLevel L)
{
    uint64_t t = (uint64_t)blockIdx.x * CONV_THREADS + threadIdx.x;
    if (t >= L.parent_count * L.fanout || !L.flags[t])
        return;
    uint64_t p = t / L.fanout;
    uint32_t n = (uint32_t)(t % L.fanout);
    uint64_t node = L.parents[p];
    bool active = false, bad = false;
    uint32_t rank = n2_mask_n_to_idx(L.src, node + L.cmp_off, L.prefix_to_mask, n, &active);
    const uint64_t child = n2_child(L.src, L.nwords, node, L.children_off, rank, L.fanout, L.child_words, &bad);
    if (bad)
        atomicOr(L.bad, 1u);
    // a rejected child points at the grid's first node words (in bounds); the host stops before use
    L.children[L.child_rank[t]] = bad ? 0ull : child;
}

// table pass: child byte offsets and tile value indices of the parents (Convert.h:171-233)
__global__ void k_conv_tables(ConvLevel L)
{
    uint64_t t = (uint64_t)blockIdx.x * CONV_THREADS + threadIdx.x;
    if (t >= L.parent_count * L.fanout)
        return;
    uint64_t p = t / L.fanout;
    uint32_t n = (uint32_t)(t % L.fanout);
    uint64_t node = L.parents[p];
    uint64_t dparent = L.dst_first + p * (uint64_t)L.dst_parent_size;
    uint64_t* entry = (uint64_t*)(L.dst + dparent + L.dst_table + 8ull * n);
    if (L.flags[t])
        *entry = (L.child_first + (uint64_t)L.child_rank[t] * L.child_size) - dparent;
    bool value_active = false;
    uint32_t local = n2_mask_n_to_idx(L.src, node + L.vmp_off, L.prefix_to_mask, n, &value_active);
    if (value_active) // tile values: the value pass of the reference runs after the child passes
        *entry = L.src[node + VDB_OFF_VALUE_IDX] + local;
}

// value and child masks of internal nodes (byte copies of mask_words words each)
__global__ void k_conv_masks(const uint64_t* __restrict__ src, uint8_t* __restrict__ dst, const uint64_t* nodes,
                             uint64_t count, uint32_t mask_words, uint32_t vm_off, uint32_t cm_off, uint64_t dst_first,
                             uint32_t dst_size, uint32_t dst_vm, uint32_t dst_cm)
{
    uint64_t t = (uint64_t)blockIdx.x * CONV_THREADS + threadIdx.x;
    if (t >= count * mask_words)
        return;
    uint64_t p = t / mask_words;
    uint32_t wi = (uint32_t)(t % mask_words);
    uint64_t node = nodes[p];
    uint8_t* d = dst + dst_first + p * (uint64_t)dst_size;
    ((uint64_t*)(d + dst_vm))[wi] = src[node + vm_off + wi];
    ((uint64_t*)(d + dst_cm))[wi] = src[node + cm_off + wi];
}

// leaves: value mask, first value index and the packed prefix sums (Convert.h:255-279)
__global__ void k_conv_leaves(const uint64_t* __restrict__ src, uint8_t* __restrict__ dst, const uint64_t* leaves,
                              uint64_t count, uint64_t dst_first)
{
    uint64_t t = (uint64_t)blockIdx.x * CONV_THREADS + threadIdx.x;
    if (t >= count)
        return;
    uint64_t node = leaves[t];
    uint64_t* d = (uint64_t*)(dst + dst_first + t * (uint64_t)ONI_LEAF_SIZE);
    uint64_t sum = 0ull;
    uint32_t accum = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 8u; w++)
    {
        uint64_t m = src[node + VDB_OFF_VM_LEAF + w];
        d[ONI_LEAF_VM / 8u + w] = m;
        if (w < 7u)
        {
            accum += (uint32_t)__popcll(m);
            sum |= (uint64_t)accum << (9u * w);
        }
    }
    d[ONI_LEAF_TABLE / 8u] = src[node + VDB_OFF_VALUE_IDX];
    d[ONI_LEAF_TABLE / 8u + 1u] = sum;
}

// root slots: tile flags (an active key) and upper flags (a child)
__global__ void k_conv_root_flags(const uint64_t* __restrict__ src, uint32_t* tile_flag, uint32_t* child_flag,
                                  uint32_t* bad_flag)
{
    uint32_t n = blockIdx.x * CONV_THREADS + threadIdx.x;
    if (n >= 32768u)
        return;
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    bool key_active = false, bad = false;
    uint32_t kidx = n2_mask_n_to_idx(src, root + VDB_OFF_KMP_ROOT, VDB_OFF_KM_ROOT - VDB_OFF_KMP_ROOT, n, &key_active);
    bool has_tile = key_active && n2_root_key(src, kidx, &bad) != VDB_END_KEY && !bad;
    if (bad)
        atomicOr(bad_flag, 1u);
    tile_flag[n] = has_tile ? 1u : 0u;
    child_flag[n] = (has_tile && ((src[root + VDB_OFF_CM_ROOT + (n >> 6u)] >> (n & 63u)) & 1ull)) ? 1u : 0u;
}

__global__ void k_conv_root_uppers(const uint64_t* __restrict__ src, uint64_t nwords, const uint32_t* child_flag,
                                   const uint32_t* upper_rank, uint64_t* uppers, uint32_t* bad_flag)
{
    uint32_t n = blockIdx.x * CONV_THREADS + threadIdx.x;
    if (n >= 32768u || !child_flag[n])
        return;
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    bool a = false, bad = false;
    uint32_t rank = n2_mask_n_to_idx(src, root + VDB_OFF_CMP_ROOT, VDB_OFF_CM_ROOT - VDB_OFF_CMP_ROOT, n, &a);
    const uint64_t upper = n2_child(src, nwords, root, VDB_OFF_CHILDREN_ROOT, rank, 32768u, N2_WORDS_UPPER, &bad);
    if (bad)
        atomicOr(bad_flag, 1u);
    uppers[upper_rank[n]] = bad ? 0ull : upper;
}

// root tiles in slot order (Convert.h:76-99, 116-127, 155-157)
__global__ void k_conv_root_tiles(const uint64_t* __restrict__ src, uint8_t* __restrict__ dst,
                                  const uint32_t* tile_rank, const uint32_t* tile_flag, const uint32_t* upper_rank,
                                  const uint32_t* child_flag, uint64_t upper_first)
{
    uint32_t n = blockIdx.x * CONV_THREADS + threadIdx.x;
    if (n >= 32768u || !tile_flag[n])
        return;
    const uint64_t root = VDB_ROOT_ADDR >> 3;
    const uint64_t droot = VDB_ROOT_ADDR;
    bool key_active = false, bad = false;
    uint32_t kidx = n2_mask_n_to_idx(src, root + VDB_OFF_KMP_ROOT, VDB_OFF_KM_ROOT - VDB_OFF_KMP_ROOT, n, &key_active);
    uint8_t* tile = dst + droot + ONI_ROOT_SIZE + (uint64_t)tile_rank[n] * ONI_TILE_SIZE;
    *(uint64_t*)(tile + 0) = n2_root_key(src, kidx, &bad); // checked by the flag pass
    bool value_active = false;
    uint32_t local = n2_mask_n_to_idx(src, root + VDB_OFF_VMP_ROOT, VDB_OFF_VM_ROOT - VDB_OFF_VMP_ROOT, n, &value_active);
    *(uint32_t*)(tile + 16) = value_active ? 1u : 0u;
    if (value_active)
        *(uint64_t*)(tile + 24) = src[root + VDB_OFF_VALUE_IDX] + local;
    if (child_flag[n])
        *(int64_t*)(tile + 8) = (int64_t)(upper_first + (uint64_t)upper_rank[n] * ONI_UPPER_SIZE) - (int64_t)droot;
}

// grid header, tree, root header and the single blind metadata entry
__global__ void k_conv_header(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t src_bbox,
                              uint64_t src_meta1, uint32_t tile_count, uint64_t dst_meta, uint64_t dst_data)
{
    uint32_t t = threadIdx.x;
    // grid header (672 B) and blind metadata (288 B) as 32-bit words
    for (uint32_t w = t; w < VDB_GRID_SIZE / 4u; w += blockDim.x)
        ((uint32_t*)dst)[w] = ((const uint32_t*)src)[w];
    for (uint32_t w = t; w < VDB_BLINDMETADATA_SIZE / 4u; w += blockDim.x)
        ((uint32_t*)(dst + dst_meta))[w] = ((const uint32_t*)(src + src_meta1))[w];
    __syncthreads();
    if (t == 0)
    {
        *(uint32_t*)(dst + VDB_GRID_OFF_TYPE) = GRID_TYPE_ONINDEX;
        *(uint64_t*)(dst + VDB_GRID_OFF_BLIND_OFFSET) = dst_meta;
        *(uint32_t*)(dst + VDB_GRID_OFF_BLIND_COUNT) = 1u;
        *(uint64_t*)(dst + VDB_TREE_ADDR + VDB_TREE_OFF_ROOT) = VDB_ROOT_ADDR - VDB_TREE_ADDR;
        for (uint32_t a = 0; a < 6u; a++)
            *(uint32_t*)(dst + VDB_ROOT_ADDR + 4u * a) = *(const uint32_t*)(src + src_bbox + 4u * a);
        *(uint32_t*)(dst + VDB_ROOT_ADDR + 24u) = tile_count;
        *(uint64_t*)(dst + dst_meta + VDB_BLIND_OFF_DATA) = dst_data - dst_meta;
    }
}

uint32_t blocks_for(uint64_t n)
{
    return (uint32_t)((n + CONV_THREADS - 1) / CONV_THREADS);
}

} // namespace

int node2_to_onindex(Context* ctx, const void* node2, uint64_t node2_size, void** out_dev, uint64_t* out_size,
                     uint64_t* alloc_size)
{
    if (!node2 || node2_size < VDB_ROOT_ADDR + VDB_SIZE_ROOT)
        return ctx->fail("node2_to_onindex: input is not a node2 grid");
    hipStream_t st = ctx->stream;
    const uint64_t* src = (const uint64_t*)node2;
    uint8_t* hdr = (uint8_t*)ctx->pinned;
    VDBR_CHECK(ctx, hipMemcpyAsync(hdr, node2, VDB_ROOT_ADDR, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    if (*(uint64_t*)(hdr + VDB_GRID_OFF_MAGIC) != VDB_MAGIC_GRID ||
        *(uint32_t*)(hdr + VDB_GRID_OFF_TYPE) != VDB_GRID_TYPE_NODE2)
        return ctx->fail("node2_to_onindex: input is not a node2 grid");
    uint64_t meta0 = *(uint64_t*)(hdr + VDB_GRID_OFF_BLIND_OFFSET);
    uint32_t blind_count = *(uint32_t*)(hdr + VDB_GRID_OFF_BLIND_COUNT);
    uint64_t header_grid_size = *(uint64_t*)(hdr + VDB_GRID_OFF_GRID_SIZE);
    if (blind_count < 2u || meta0 + 2ull * VDB_BLINDMETADATA_SIZE > node2_size)
        return ctx->fail("node2_to_onindex: node2 grid without value blind data");
    uint8_t* meta = hdr + 1024;
    VDBR_CHECK(ctx, hipMemcpyAsync(meta, (const uint8_t*)node2 + meta0, 2 * VDB_BLINDMETADATA_SIZE,
                                   hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint64_t meta1 = meta0 + VDB_BLINDMETADATA_SIZE;
    const uint64_t src_bbox = meta0 + *(uint64_t*)(meta + VDB_BLIND_OFF_DATA);
    const uint64_t src_values = meta1 + *(uint64_t*)(meta + VDB_BLINDMETADATA_SIZE + VDB_BLIND_OFF_DATA);
    const uint64_t value_count = *(uint64_t*)(meta + VDB_BLINDMETADATA_SIZE + VDB_BLIND_OFF_COUNT);
    const uint32_t value_size = *(uint32_t*)(meta + VDB_BLINDMETADATA_SIZE + VDB_BLIND_OFF_SIZE);
    const uint64_t value_bytes = value_count * value_size;
    if (src_values + value_bytes > node2_size || src_bbox + 24 > node2_size)
        return ctx->fail("node2_to_onindex: blind data outside the grid");

    // root level: tile and upper flags, ranks, the upper list
    const uint64_t nwords = node2_size / 8u;
    uint32_t* root_ws = (uint32_t*)ctx->ws(WS_CONV_ROOT, (4 * 32768 + 4) * sizeof(uint32_t));
    if (!root_ws)
        return -1;
    uint32_t* bad_flag = root_ws + 4 * 32768 + 2;
    VDBR_CHECK(ctx, hipMemsetAsync(bad_flag, 0, 4, st));
    // a child offset outside the grid (a corrupt or truncated grid) stops the conversion before the
    // rejected nodes are read
    auto grid_ok = [&]() -> bool {
        if (hipMemcpyAsync(ctx->pinned, bad_flag, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return false;
        return *(const uint32_t*)ctx->pinned == 0u;
    };
    uint32_t* tile_flag = root_ws;
    uint32_t* child_flag = root_ws + 32768;
    uint32_t* tile_rank = root_ws + 2 * 32768;
    uint32_t* upper_rank = root_ws + 3 * 32768;
    uint32_t* totals = root_ws + 4 * 32768;
    hipLaunchKernelGGL(k_conv_root_flags, dim3(blocks_for(32768)), dim3(CONV_THREADS), 0, st, src, tile_flag,
                       child_flag, bad_flag);
    if (scan_u32(ctx, tile_flag, tile_rank, 32768, false, totals) ||
        scan_u32(ctx, child_flag, upper_rank, 32768, false, totals + 1))
        return -1;
    VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, totals, 8, hipMemcpyDeviceToHost, st));
    VDBR_CHECK(ctx, hipStreamSynchronize(st));
    const uint32_t tiles = ((uint32_t*)ctx->pinned)[0];
    const uint64_t U = ((uint32_t*)ctx->pinned)[1];
    uint64_t* uppers = (uint64_t*)ctx->ws(WS_CONV_UPPERS, (U + 1) * sizeof(uint64_t));
    if (!uppers)
        return -1;
    hipLaunchKernelGGL(k_conv_root_uppers, dim3(blocks_for(32768)), dim3(CONV_THREADS), 0, st, src, nwords,
                       child_flag, upper_rank, uppers, bad_flag);
    if (!grid_ok())
        return ctx->fail("node2_to_onindex: root keys or upper offsets outside the grid");

    // upper level: U x 32768 slots -> lower ranks and the lower list
    const uint64_t un = U * 32768ull;
    uint32_t* uflags = (uint32_t*)ctx->ws(WS_CONV_FLAGS, (2 * un + 2) * sizeof(uint32_t));
    if (!uflags)
        return -1;
    ConvLevel LU{};
    LU.src = src;
    LU.parents = uppers;
    LU.parent_count = U;
    LU.fanout = 32768u;
    LU.cmp_off = VDB_OFF_CMP_UPPER;
    LU.vmp_off = VDB_OFF_VMP_UPPER;
    LU.children_off = VDB_OFF_CHILDREN_UPPER;
    LU.prefix_to_mask = VDB_OFF_CM_UPPER - VDB_OFF_CMP_UPPER;
    LU.dst_parent_size = ONI_UPPER_SIZE;
    LU.dst_table = ONI_UPPER_TABLE;
    LU.child_size = ONI_LOWER_SIZE;
    LU.flags = uflags;
    LU.child_rank = uflags + un;
    LU.nwords = nwords;
    LU.child_words = N2_WORDS_LOWER;
    LU.bad = bad_flag;
    uint64_t L = 0;
    if (U)
    {
        hipLaunchKernelGGL(k_conv_flags, dim3(blocks_for(un)), dim3(CONV_THREADS), 0, st, LU);
        if (scan_u32(ctx, uflags, uflags + un, un, false, uflags + 2 * un))
            return -1;
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, uflags + 2 * un, 4, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        L = ((uint32_t*)ctx->pinned)[0];
    }
    uint64_t* lowers = (uint64_t*)ctx->ws(WS_CONV_LOWERS, (L + 1) * sizeof(uint64_t));
    if (!lowers)
        return -1;
    LU.children = lowers;
    if (U)
        hipLaunchKernelGGL(k_conv_children, dim3(blocks_for(un)), dim3(CONV_THREADS), 0, st, LU);
    if (!grid_ok())
        return ctx->fail("node2_to_onindex: lower node offsets outside the grid");

    // lower level: L x 4096 slots -> leaf ranks and the leaf list
    const uint64_t ln = L * 4096ull;
    uint32_t* lflags = (uint32_t*)ctx->ws(WS_CONV_FLAGS2, (2 * ln + 2) * sizeof(uint32_t));
    if (!lflags)
        return -1;
    ConvLevel LL{};
    LL.src = src;
    LL.parents = lowers;
    LL.parent_count = L;
    LL.fanout = 4096u;
    LL.cmp_off = VDB_OFF_CMP_LOWER;
    LL.vmp_off = VDB_OFF_VMP_LOWER;
    LL.children_off = VDB_OFF_CHILDREN_LOWER;
    LL.prefix_to_mask = VDB_OFF_CM_LOWER - VDB_OFF_CMP_LOWER;
    LL.dst_parent_size = ONI_LOWER_SIZE;
    LL.dst_table = ONI_LOWER_TABLE;
    LL.child_size = ONI_LEAF_SIZE;
    LL.flags = lflags;
    LL.child_rank = lflags + ln;
    LL.nwords = nwords;
    LL.child_words = N2_WORDS_LEAF;
    LL.bad = bad_flag;
    uint64_t F = 0;
    if (L)
    {
        hipLaunchKernelGGL(k_conv_flags, dim3(blocks_for(ln)), dim3(CONV_THREADS), 0, st, LL);
        if (scan_u32(ctx, lflags, lflags + ln, ln, false, lflags + 2 * ln))
            return -1;
        VDBR_CHECK(ctx, hipMemcpyAsync(ctx->pinned, lflags + 2 * ln, 4, hipMemcpyDeviceToHost, st));
        VDBR_CHECK(ctx, hipStreamSynchronize(st));
        F = ((uint32_t*)ctx->pinned)[0];
    }
    uint64_t* leaves = (uint64_t*)ctx->ws(WS_CONV_LEAVES, (F + 1) * sizeof(uint64_t));
    if (!leaves)
        return -1;
    LL.children = leaves;
    if (L)
        hipLaunchKernelGGL(k_conv_children, dim3(blocks_for(ln)), dim3(CONV_THREADS), 0, st, LL);
    if (!grid_ok())
        return ctx->fail("node2_to_onindex: leaf offsets outside the grid");

    // output layout: header, root + tiles, uppers, lowers, leaves, blind metadata, values
    const uint64_t upper_first = VDB_ROOT_ADDR + ONI_ROOT_SIZE + (uint64_t)tiles * ONI_TILE_SIZE;
    const uint64_t lower_first = upper_first + U * ONI_UPPER_SIZE;
    const uint64_t leaf_first = lower_first + L * ONI_LOWER_SIZE;
    const uint64_t dst_meta = leaf_first + F * ONI_LEAF_SIZE;
    const uint64_t dst_data = dst_meta + VDB_BLINDMETADATA_SIZE;
    const uint64_t end = dst_data + value_bytes;
    // the copied grid header keeps the node2 grid size: the buffer covers both
    uint64_t alloc = end > header_grid_size ? end : header_grid_size;
    alloc = (alloc + 7) & ~7ull;
    uint8_t* dst = (uint8_t*)ctx->ws(WS_ONINDEX, alloc);
    if (!dst)
        return -1;
    VDBR_CHECK(ctx, hipMemsetAsync(dst, 0, alloc, st));
    hipLaunchKernelGGL(k_conv_header, dim3(1), dim3(256), 0, st, (const uint8_t*)node2, dst, src_bbox, meta1, tiles,
                       dst_meta, dst_data);
    hipLaunchKernelGGL(k_conv_root_tiles, dim3(blocks_for(32768)), dim3(CONV_THREADS), 0, st, src, dst, tile_rank,
                       tile_flag, upper_rank, child_flag, upper_first);
    if (U)
    {
        LU.dst = dst;
        LU.dst_first = upper_first;
        LU.child_first = lower_first;
        hipLaunchKernelGGL(k_conv_tables, dim3(blocks_for(un)), dim3(CONV_THREADS), 0, st, LU);
        const uint32_t mw = 32768u / 64u;
        hipLaunchKernelGGL(k_conv_masks, dim3(blocks_for(U * mw)), dim3(CONV_THREADS), 0, st, src, dst, uppers, U, mw,
                           VDB_OFF_VM_UPPER, VDB_OFF_CM_UPPER, upper_first, ONI_UPPER_SIZE, ONI_UPPER_VM, ONI_UPPER_CM);
    }
    if (L)
    {
        LL.dst = dst;
        LL.dst_first = lower_first;
        LL.child_first = leaf_first;
        hipLaunchKernelGGL(k_conv_tables, dim3(blocks_for(ln)), dim3(CONV_THREADS), 0, st, LL);
        const uint32_t mw = 4096u / 64u;
        hipLaunchKernelGGL(k_conv_masks, dim3(blocks_for(L * mw)), dim3(CONV_THREADS), 0, st, src, dst, lowers, L, mw,
                           VDB_OFF_VM_LOWER, VDB_OFF_CM_LOWER, lower_first, ONI_LOWER_SIZE, ONI_LOWER_VM, ONI_LOWER_CM);
    }
    if (F)
        hipLaunchKernelGGL(k_conv_leaves, dim3(blocks_for(F)), dim3(CONV_THREADS), 0, st, src, dst, leaves, F,
                           leaf_first);
    if (value_bytes)
        VDBR_CHECK(ctx, hipMemcpyAsync(dst + dst_data, (const uint8_t*)node2 + src_values, value_bytes,
                                       hipMemcpyDeviceToDevice, st));
    VDBR_CHECK(ctx, hipGetLastError());
    *out_dev = dst;
    *out_size = end;
    *alloc_size = alloc;
    return 0;
}

} // namespace vdbr
