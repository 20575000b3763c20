// vdb_abi.hip -- context/workspace management and the exported C ABI of libvdbraster.so:
// the reference function tables (pnanovdb_get_raster / _compute / _parallel_primitives) and
// the direct vdbr_* device entry points. No C++ exception crosses the ABI.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "vdb_fileformat.h"
#include "vdb_internal.h"

namespace vdbr
{

static const size_t PINNED_BYTES = 1u << 20;

void Context::drop_batch_graphs()
{
    for (hipGraphExec_t x : batch_execs)
        (void)hipGraphExecDestroy(x);
    batch_execs.clear();
    batch_sig.clear();
}

void Context::ws_free(WsSlot slot)
{
    if (!ws_ptr[slot])
        return;
    (void)hipStreamSynchronize(stream);
    (void)hipFree(ws_ptr[slot]);
    ws_ptr[slot] = nullptr;
    ws_size[slot] = 0;
}

void* Context::ws(WsSlot slot, size_t bytes)
{
    if (bytes == 0)
        bytes = 8;
    if (probe)
    {
        // sizing pass: nothing recorded now is ever launched
        probe_size[slot] = std::max(probe_size[slot], bytes);
        return (void*)(uintptr_t)((uint64_t)(slot + 1) << 40);
    }
    if (bytes <= ws_size[slot] && ws_ptr[slot])
        return ws_ptr[slot];
    if (capturing)
    {
        fail("workspace slot " + std::to_string((int)slot) + " needs " + std::to_string(bytes) +
             " bytes during graph capture (it was not reserved)");
        return nullptr;
    }
    if (ws_ptr[slot])
    {
        // earlier kernels on the stream may still read the old allocation
        (void)hipStreamSynchronize(stream);
        (void)hipFree(ws_ptr[slot]);
        ws_ptr[slot] = nullptr;
        ws_size[slot] = 0;
    }
    size_t alloc = bytes + bytes / 8; // headroom for growing workloads
    alloc = (alloc + 255) & ~(size_t)255;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, alloc);
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        e = hipMalloc(&p, (bytes + 255) & ~(size_t)255);
        alloc = (bytes + 255) & ~(size_t)255;
    }
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        fail("device allocation of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    ws_ptr[slot] = p;
    ws_size[slot] = alloc;
    if (slot == WS_SORT_STATUS)
    {
        (void)hipMemsetAsync(p, 0, alloc, stream);
        sort_epoch = 0;
    }
    return p;
}

int Context::sort_workspace(uint64_t n, bool with_vals, uint64_t tiles, bool aux, SortWorkspace* out)
{
    out->keys_alt = (uint64_t*)ws(aux ? WS_AUX_ALT : WS_SORT_ALT, n * sizeof(uint64_t));
    out->vals_alt = with_vals ? (uint32_t*)ws(aux ? WS_AUX_VALS_ALT : WS_SORT_VALS_ALT, n * sizeof(uint32_t)) : nullptr;
    out->status = (uint64_t*)ws(WS_SORT_STATUS, tiles * SORT_BINS * sizeof(uint64_t));
    // histograms and digit starts of up to 16 passes, then 16 tile counters
    uint8_t* small = (uint8_t*)ws(WS_SORT_SMALL, 16 * SORT_BINS * 8 * 2 + 64);
    if (!out->keys_alt || (with_vals && !out->vals_alt) || !out->status || !small)
        return -1;
    out->hist = (uint64_t*)small;
    out->starts = (uint64_t*)(small + 16 * SORT_BINS * 8);
    out->tile_counters = (uint32_t*)(small + 16 * SORT_BINS * 8 * 2);
    sort_status_tiles = ws_size[WS_SORT_STATUS] / (SORT_BINS * sizeof(uint64_t));
    return 0;
}

uint32_t Context::next_sort_epoch(uint64_t tiles)
{
    (void)tiles;
    sort_epoch++;
    if (sort_epoch >= 65535u)
    {
        // epochs are 16-bit: clear the status words before one could repeat (a fill kernel, so
        // captured work may contain it)
        (void)fill_words(this, (uint64_t*)ws_ptr[WS_SORT_STATUS], host_count(ws_size[WS_SORT_STATUS] / 8), 0ull);
        sort_epoch = 1u;
    }
    return sort_epoch;
}

uint64_t Context::workspace_bytes() const
{
    uint64_t s = 0;
    for (int i = 0; i < WS_COUNT; i++)
        s += ws_size[i];
    return s;
}

void Context::release()
{
    (void)hipStreamSynchronize(stream);
    // a captured batch graph refers to the workspace being released
    drop_batch_graphs();
    for (int i = 0; i < WS_COUNT; i++)
    {
        if (ws_ptr[i])
            (void)hipFree(ws_ptr[i]);
        ws_ptr[i] = nullptr;
        ws_size[i] = 0;
    }
    drop_pairs();
}

static Context* create_context(int device)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    {
        (void)hipGetLastError();
        return nullptr;
    }
    if (device < 0 || device >= count)
        return nullptr;
    if (hipSetDevice(device) != hipSuccess)
        return nullptr;
    Context* ctx = new (std::nothrow) Context();
    if (!ctx)
        return nullptr;
    ctx->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    {
        const uint32_t per_cu = emit_blocks_per_cu();
        ctx->emit_blocks = (uint32_t)cus * (per_cu ? per_cu : 3u);
    }
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess)
    {
        delete ctx;
        return nullptr;
    }
    ctx->stream = ctx->own_stream;
    if (hipHostMalloc((void**)&ctx->pinned, PINNED_BYTES, hipHostMallocDefault) != hipSuccess)
    {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return nullptr;
    }
    return ctx;
}

static void destroy_context(Context* ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    for (Context* w : ctx->batch_workers)
        destroy_context(w);
    ctx->batch_workers.clear();
    for (auto& c : ctx->batch_chunks)
        (void)hipFree(c.first);
    ctx->batch_chunks.clear();
    ctx->release();
    if (ctx->pinned)
        (void)hipHostFree(ctx->pinned);
    if (ctx->batch_pinned)
        (void)hipHostFree(ctx->batch_pinned);
    if (ctx->own_stream)
        (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

// ----------------------------------------------------------------------------- batched launch
constexpr size_t BATCH_CHUNK_BYTES = size_t(1) << 30; // output arena grows in chunks of at least 1 GiB

// Output arena: grids of one batched call are packed into chunks that persist on the context and
// are reused by the next call (no device allocation once the chunks cover a batch).
struct BatchArena
{
    Context* owner;
    std::mutex lock;
    size_t chunk = 0;  // current chunk
    size_t offset = 0; // first free byte in it

    void* reserve(size_t bytes, std::string* err)
    {
        bytes = (bytes + 255u) & ~size_t(255u);
        std::lock_guard<std::mutex> g(lock);
        while (chunk < owner->batch_chunks.size() && offset + bytes > owner->batch_chunks[chunk].second)
        {
            chunk++;
            offset = 0;
        }
        if (chunk == owner->batch_chunks.size())
        {
            size_t sz = std::max(bytes, BATCH_CHUNK_BYTES);
            void* p = nullptr;
            hipError_t e = hipMalloc(&p, sz);
            if (e != hipSuccess)
            {
                *err = std::string("batch output arena: hipMalloc: ") + hipGetErrorString(e);
                return nullptr;
            }
            owner->batch_chunks.emplace_back(p, sz);
            offset = 0;
        }
        void* p = (uint8_t*)owner->batch_chunks[chunk].first + offset;
        offset += bytes;
        return p;
    }
};

// Batched raster3d: every cloud that fits the captured pipeline runs in one HIP graph
// (raster_batch_graph); the rest (several slabs, no fragments, invalid input, or counts past the
// bounds of the capture) run one by one through raster_node2, their grids packed into the arena.
static int raster_batch(Context* ctx, vdbr_batch_item_t* items, uint32_t count, uint32_t workers)
{
    if (count == 0)
        return 0;
    if (!items)
        return ctx->fail("raster batch: items is NULL");
    if (workers == 0)
        workers = 4;
    workers = std::min<uint32_t>(std::min<uint32_t>(workers, 16u), count);
    while (ctx->batch_workers.size() < workers)
    {
        Context* w = create_context(ctx->device);
        if (!w)
            return ctx->fail("raster batch: could not create a worker context");
        ctx->batch_workers.push_back(w);
    }
    std::vector<Context*> ws(ctx->batch_workers.begin(), ctx->batch_workers.begin() + workers);
    std::vector<uint32_t> rerun;
    std::vector<uint8_t> setup_done;
    ctx->last_batch_graph_clouds = 0;
    if (raster_batch_graph(ctx, ws, items, count, &rerun, &setup_done))
        return -1;
    ctx->last_batch_graph_clouds = count - (uint32_t)rerun.size();
    const auto t0 = std::chrono::steady_clock::now();
    BatchArena arena;
    arena.owner = ctx;
    for (uint32_t i : rerun)
    {
        vdbr_batch_item_t& it = items[i];
        vdbr_raster_params_t prm = it.params;
        if (setup_done[i])
            prm.flags &= ~VDBR_FLAG_RAW_INPUTS;
        void* grid = nullptr;
        uint64_t size = 0;
        if (raster_node2(ctx, &it.gaussians, &prm, &grid, &size, &it.stats) != 0)
            return ctx->fail("raster batch: cloud " + std::to_string(i) + ": " + ctx->error);
        it.grid_dev = nullptr;
        it.grid_size = size;
        if (size > 0)
        {
            std::string err;
            void* dst = arena.reserve(size, &err);
            if (!dst)
                return ctx->fail("raster batch: cloud " + std::to_string(i) + ": " + err);
            VDBR_CHECK(ctx, hipMemcpyAsync(dst, grid, size, hipMemcpyDeviceToDevice, ctx->stream));
            it.grid_dev = dst;
        }
    }
    ctx->batch_ms[4] = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// ----------------------------------------------------------------------------- default context
// The fn-table entry points carry no HIP state of their own: they run on one library context
// (device 0), serialised by a mutex (the reference raster entry points are synchronous too).
static std::mutex g_default_mutex;
static Context* g_default = nullptr;

static Context* default_context()
{
    if (!g_default)
        g_default = create_context(0);
    return g_default;
}

} // namespace vdbr

using namespace vdbr;

// ============================================================================= compute arrays
// Compute.cpp:707-760 semantics
static pnanovdb_compute_array_t* PNANOVDB_ABI arr_create(size_t element_size, pnanovdb_uint64_t element_count,
                                                         const void* data)
{
    pnanovdb_compute_array_t* a = new (std::nothrow) pnanovdb_compute_array_t();
    if (!a)
        return nullptr;
    a->element_size = element_size;
    a->element_count = element_count;
    a->filepath = nullptr;
    size_t bytes = element_size * element_count;
    size_t alloc = bytes < 65536u ? 65536u : bytes;
    a->data = new (std::nothrow) char[alloc];
    if (!a->data)
    {
        delete a;
        return nullptr;
    }
    memset(a->data, 0, 65536u < alloc ? 65536u : alloc);
    if (data)
        memcpy(a->data, data, bytes);
    else
        memset(a->data, 0, bytes);
    return a;
}

static void PNANOVDB_ABI arr_destroy(pnanovdb_compute_array_t* a)
{
    if (!a)
        return;
    delete[](char*) a->data;
    a->data = nullptr;
    delete a;
}

static void* PNANOVDB_ABI arr_map(pnanovdb_compute_array_t* a)
{
    return a ? a->data : nullptr;
}

static void PNANOVDB_ABI arr_unmap(pnanovdb_compute_array_t* a)
{
    (void)a;
}

static pnanovdb_compute_array_t* PNANOVDB_ABI arr_duplicate(pnanovdb_compute_array_t* a)
{
    return a ? arr_create(a->element_size, a->element_count, a->data) : nullptr;
}

static void PNANOVDB_ABI arr_print_range(const pnanovdb_compute_t* compute, pnanovdb_compute_log_print_t log_print,
                                         const char* name, pnanovdb_compute_array_t* arr,
                                         pnanovdb_uint32_t channel_count)
{
    (void)compute;
    if (!arr || !log_print || channel_count == 0)
        return;
    if (channel_count > 4u)
        channel_count = 4u;
    float mn[4] = { INFINITY, INFINITY, INFINITY, INFINITY }, mx[4] = { -INFINITY, -INFINITY, -INFINITY, -INFINITY };
    const float* v = (const float*)arr->data;
    for (uint64_t i = 0; i < arr->element_count; i++)
    {
        uint32_t c = (uint32_t)(i % channel_count);
        mn[c] = fminf(mn[c], v[i]);
        mx[c] = fmaxf(mx[c], v[i]);
    }
    for (uint32_t c = 0; c < channel_count; c++)
        log_print(1u, "%s[%u] min(%f) max(%f)", name ? name : "array", c, mn[c], mx[c]);
}

static pnanovdb_compute_t g_compute;
static bool g_compute_init = false;

// This library's own handle (dlopen of the file that holds this function, no new load), which the
// tables carry as `module` so that hosts resolve the other getters through it as
// pnanovdb_get_proc_address(compute->module, ...) does (putil/Raster.h:199-210, GridBuild.h:178-181).
static void* own_module()
{
    static void* handle = nullptr;
    if (!handle)
    {
        Dl_info info;
        if (dladdr((void*)&own_module, &info) && info.dli_fname)
            handle = dlopen(info.dli_fname, RTLD_NOW | RTLD_NOLOAD);
    }
    return handle;
}

const pnanovdb_compute_t* vdbr::compute_table()
{
    if (!g_compute_init)
    {
        memset(&g_compute, 0, sizeof(g_compute));
        g_compute.module = own_module();
        g_compute.create_array = arr_create;
        g_compute.destroy_array = arr_destroy;
        g_compute.map_array = arr_map;
        g_compute.unmap_array = arr_unmap;
        g_compute.compute_array_print_range = arr_print_range;
        g_compute.duplicate_array = arr_duplicate;
        g_compute_init = true;
    }
    return &g_compute;
}

// ============================================================================= raster table
// Log routing as RasterUtils.cpp:163-186: a host that gives the compute table a device interface
// (get_compute_interface / get_compute_context) receives the messages through the compute
// interface's get_log_print; otherwise they go to stderr.
struct ComputeInterfacePrefix // leading fields of pnanovdb_compute_interface_t (putil/Compute.h:338-349)
{
    const pnanovdb_reflect_data_type_t* interface_pnanovdb_reflect_data_type;
    void* get_api;
    void* is_feature_supported;
    void* get_frame_info;
    pnanovdb_compute_log_print_t(PNANOVDB_ABI* get_log_print)(void* context);
};

static pnanovdb_compute_log_print_t host_log_print(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue)
{
    if (!compute || !queue)
        return nullptr;
    typedef void*(PNANOVDB_ABI * get_iface_t)(pnanovdb_compute_queue_t*);
    get_iface_t get_iface = (get_iface_t)compute->device_interface.get_compute_interface;
    get_iface_t get_ctx = (get_iface_t)compute->device_interface.get_compute_context;
    if (!get_iface || !get_ctx)
        return nullptr;
    ComputeInterfacePrefix* ci = (ComputeInterfacePrefix*)get_iface(queue);
    void* context = get_ctx(queue);
    if (!ci || !ci->get_log_print)
        return nullptr;
    return ci->get_log_print(context);
}

void vdbr::log_message(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, uint32_t level,
                       const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    pnanovdb_compute_log_print_t lp = host_log_print(compute, queue);
    if (lp)
        lp(level, "%s", buf);
    else
        fprintf(stderr, "libvdbraster: %s\n", buf);
}

// Profiler report as Raster.cpp:583 arranges it (enable_profiler with userdata "raster"): one entry
// per pipeline stage with its GPU time (HIP events on the library's stream) in seconds, then the
// whole call with its host wall time.
static void report_profile(pnanovdb_profiler_report_t report, const vdbr_raster_stats_t& st, float ms_convert,
                           double host_seconds)
{
    if (!report)
        return;
    static std::atomic<uint64_t> capture{ 0 };
    pnanovdb_compute_profiler_entry_t entries[7] = { { "raster_preprocess_setup", 0.f, st.ms_setup * 1e-3f },
                         { "raster_emit_pairs", 0.f, st.ms_emit * 1e-3f },
                         { "raster_sort_pairs", 0.f, st.ms_sort * 1e-3f },
                         { "raster_reduce_values", 0.f, st.ms_reduce * 1e-3f },
                         { "raster_grid_build", 0.f, st.ms_build * 1e-3f },
                         { "convert_onindex", 0.f, ms_convert * 1e-3f },
                         { "raster_total", (float)host_seconds, (st.ms_total + ms_convert) * 1e-3f } };
    report((void*)"raster", capture.fetch_add(1), 7u, entries);
}

struct RasterContextImpl
{
    int device;
};

// Gaussian data (RasterContext.cpp:92-229): host copies at creation; one device allocation holding
// means[3N] | quats[4N] | scales[3N] | colors[3N] | opacities[N], filled once by upload (or by the
// first raster) and rasterized from there by raster_gaussian_3d without another transfer.
struct GaussianDataImpl
{
    uint64_t count = 0;
    std::vector<float> means, quats, scales, colors, opacities;
    float cull_scale = 0.f;
    float* dev = nullptr;
    int device = 0;
    bool uploaded = false;

    float* d_means() const { return dev; }
    float* d_quats() const { return dev + 3 * count; }
    float* d_scales() const { return dev + 7 * count; }
    float* d_colors() const { return dev + 10 * count; }
    float* d_opac() const { return dev + 13 * count; }
    ~GaussianDataImpl()
    {
        if (dev)
        {
            (void)hipSetDevice(device);
            (void)hipFree(dev);
        }
    }
};

static pnanovdb_raster_context_t* PNANOVDB_ABI r_create_context(const pnanovdb_compute_t* compute,
                                                                pnanovdb_compute_queue_t* queue)
{
    (void)compute;
    (void)queue;
    RasterContextImpl* c = new (std::nothrow) RasterContextImpl();
    if (c)
        c->device = 0;
    return (pnanovdb_raster_context_t*)c;
}

static void PNANOVDB_ABI r_destroy_context(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                           pnanovdb_raster_context_t* context)
{
    (void)compute;
    (void)queue;
    delete (RasterContextImpl*)context;
}

// a host compute array holds at least `count` floats
static bool floats_ok(const pnanovdb_compute_array_t* a, uint64_t count)
{
    return a && a->element_size == 4 && a->element_count >= count && (count == 0 || a->data);
}

static void copy_floats(std::vector<float>& dst, const pnanovdb_compute_array_t* a, uint64_t count)
{
    dst.resize(count);
    if (count)
        memcpy(dst.data(), a->data, count * sizeof(float));
}

// shader_params_arrays[gaussian_frag_color] holds {float cull_scale; bool use_default; ...}
// (gaussian_frag_color.slang:21-28); index 1 in the reference shader enum (Raster.h:212-214)
static float cull_scale_of(pnanovdb_compute_array_t** shader_params_arrays)
{
    if (!shader_params_arrays || !shader_params_arrays[1] || !shader_params_arrays[1]->data)
        return 0.f;
    if (shader_params_arrays[1]->element_size * shader_params_arrays[1]->element_count < 8)
        return 0.f;
    const uint8_t* p = (const uint8_t*)shader_params_arrays[1]->data;
    float cs;
    memcpy(&cs, p, 4);
    uint32_t use_default;
    memcpy(&use_default, p + 4, 4);
    return use_default ? 0.f : cs;
}

static pnanovdb_raster_gaussian_data_t* PNANOVDB_ABI r_create_gaussian_data(
    const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, pnanovdb_raster_context_t* context,
    pnanovdb_compute_array_t* means, pnanovdb_compute_array_t* quaternions, pnanovdb_compute_array_t* scales,
    pnanovdb_compute_array_t* colors, pnanovdb_compute_array_t* sh_0, pnanovdb_compute_array_t* sh_n,
    pnanovdb_compute_array_t* opacities, pnanovdb_compute_array_t** shader_params_arrays,
    pnanovdb_raster_shader_params_t* raster_params)
{
    (void)context;
    (void)sh_0;
    (void)sh_n;
    (void)raster_params;
    if (!means || !(means->element_count == 0 || means->data))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "create_gaussian_data: missing means array");
        return nullptr;
    }
    const uint64_t n = means->element_count / 3u;
    if (!floats_ok(means, 3 * n) || !floats_ok(quaternions, 4 * n) || !floats_ok(scales, 3 * n) ||
        !floats_ok(colors, 3 * n) || !floats_ok(opacities, n))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR,
                    "create_gaussian_data: arrays must be float32 with 3N means, 4N quaternions, 3N scales, 3N "
                    "colors and N opacities (N = %llu)",
                    (unsigned long long)n);
        return nullptr;
    }
    GaussianDataImpl* d = new (std::nothrow) GaussianDataImpl();
    if (!d)
        return nullptr;
    try
    {
        d->count = n;
        copy_floats(d->means, means, 3 * n);
        copy_floats(d->quats, quaternions, 4 * n);
        copy_floats(d->scales, scales, 3 * n);
        copy_floats(d->colors, colors, 3 * n);
        copy_floats(d->opacities, opacities, n);
    }
    catch (...)
    {
        delete d;
        return nullptr;
    }
    d->cull_scale = cull_scale_of(shader_params_arrays);
    return (pnanovdb_raster_gaussian_data_t*)d;
}

// one upload per data object (RasterContext.cpp:201-229, has_uploaded)
static int upload_data(GaussianDataImpl* d, Context* ctx)
{
    if (d->uploaded)
        return 0;
    d->device = ctx->device;
    const uint64_t n = d->count;
    if (!d->dev)
    {
        void* p = nullptr;
        if (hipMalloc(&p, (n > 0 ? n : 1) * 16 * sizeof(float)) != hipSuccess)
        {
            (void)hipGetLastError();
            return ctx->fail("upload_gaussian_data: device allocation failed");
        }
        d->dev = (float*)p;
    }
    if (n > 0)
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(d->d_means(), d->means.data(), 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d->d_quats(), d->quats.data(), 4 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d->d_scales(), d->scales.data(), 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d->d_colors(), d->colors.data(), 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d->d_opac(), d->opacities.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    d->uploaded = true;
    return 0;
}

static void PNANOVDB_ABI r_upload_gaussian_data(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                                pnanovdb_raster_context_t* context, pnanovdb_raster_gaussian_data_t* data)
{
    (void)context;
    GaussianDataImpl* d = (GaussianDataImpl*)data;
    if (!d)
        return;
    std::lock_guard<std::mutex> lock(g_default_mutex);
    Context* ctx = default_context();
    if (!ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "upload_gaussian_data: no HIP device");
        return;
    }
    if (upload_data(d, ctx))
        log_message(compute, queue, LOG_LEVEL_ERROR, "%s", ctx->error.c_str());
}

static void PNANOVDB_ABI r_destroy_gaussian_data(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                                 pnanovdb_raster_gaussian_data_t* data)
{
    (void)compute;
    (void)queue;
    delete (GaussianDataImpl*)data;
}

// uploads host arrays (preprocessed or raw) into one device region and runs the pipeline;
// returns the grid on the host in a vector
static int run_host(Context* ctx, uint64_t n, const float* means, const float* quats, const float* scales,
                    const float* colors_or_sh0, const float* opacities, float voxel_size, float cull_scale,
                    uint32_t flags, std::vector<uint8_t>* grid_out, float* quats_back, float* scales_back,
                    float* opac_back, bool onindex = false, vdbr_raster_stats_t* stats = nullptr,
                    float* ms_convert = nullptr)
{
    uint64_t nn = n > 0 ? n : 1;
    size_t bytes = nn * (3 + 4 + 3 + 3 + 1 + 3) * sizeof(float);
    float* dev = (float*)ctx->ws(WS_RAW_COPY, bytes);
    if (!dev)
        return -1;
    float* d_means = dev;
    float* d_quats = d_means + 3 * nn;
    float* d_scales = d_quats + 4 * nn;
    float* d_colors = d_scales + 3 * nn;
    float* d_opac = d_colors + 3 * nn;
    float* d_sh0 = d_opac + nn;
    if (n > 0)
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(d_means, means, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d_quats, quats, 4 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d_scales, scales, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync((flags & VDBR_FLAG_RAW_INPUTS) ? d_sh0 : d_colors, colors_or_sh0, 3 * n * 4,
                                       hipMemcpyHostToDevice, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(d_opac, opacities, n * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    vdbr_gaussians_t g;
    g.means = d_means;
    g.opacities = d_opac;
    g.quats = d_quats;
    g.scales = d_scales;
    g.colors = d_colors;
    g.sh0 = d_sh0;
    g.count = n;
    vdbr_raster_params_t p;
    memset(&p, 0, sizeof(p));
    p.voxel_size = voxel_size;
    p.cull_scale = cull_scale;
    p.value_mode = VDBR_VALUE_RGBA8;
    p.flags = flags;
    void* grid = nullptr;
    uint64_t size = 0;
    vdbr_raster_stats_t local;
    if (raster_node2(ctx, &g, &p, &grid, &size, stats ? stats : &local))
        return -1;
    if (onindex)
    {
        // RasterUtils.cpp:187-204: the node2 grid converted to OnIndex (putil/Convert.h)
        void* oni = nullptr;
        uint64_t oni_size = 0, oni_alloc = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool timed = ms_convert && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        if (timed)
            (void)hipEventRecord(e0, ctx->stream);
        const int rc = node2_to_onindex(ctx, grid, size, &oni, &oni_size, &oni_alloc);
        if (timed)
        {
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(ms_convert, e0, e1);
        }
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        if (rc)
            return -1;
        grid = oni;
        size = oni_alloc;
    }
    grid_out->resize(size);
    VDBR_CHECK(ctx, hipMemcpyAsync(grid_out->data(), grid, size, hipMemcpyDeviceToHost, ctx->stream));
    if (n > 0 && quats_back)
    {
        VDBR_CHECK(ctx, hipMemcpyAsync(quats_back, d_quats, 4 * n * 4, hipMemcpyDeviceToHost, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(scales_back, d_scales, 3 * n * 4, hipMemcpyDeviceToHost, ctx->stream));
        VDBR_CHECK(ctx, hipMemcpyAsync(opac_back, d_opac, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    VDBR_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

static pnanovdb_compute_array_t* make_output(const pnanovdb_compute_t* compute, const std::vector<uint8_t>& grid)
{
    const pnanovdb_compute_t* c = (compute && compute->create_array) ? compute : compute_table();
    uint64_t words = (grid.size() + 3) / 4;
    pnanovdb_compute_array_t* out = c->create_array(4u, words, nullptr);
    if (out && !grid.empty())
        memcpy(out->data, grid.data(), grid.size());
    return out;
}

static double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Raster.cpp:560-672: preprocessed arrays -> node2 grid array, trimmed to the grid size
static pnanovdb_compute_array_t* PNANOVDB_ABI r_raster_to_nanovdb(
    const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue, float voxel_size,
    pnanovdb_compute_array_t* means, pnanovdb_compute_array_t* quaternions, pnanovdb_compute_array_t* scales,
    pnanovdb_compute_array_t* colors, pnanovdb_compute_array_t* sh_0, pnanovdb_compute_array_t* sh_n,
    pnanovdb_compute_array_t* opacities, pnanovdb_compute_array_t** shader_params_arrays,
    pnanovdb_profiler_report_t profiler_report, void* userdata)
{
    (void)sh_0;
    (void)sh_n;
    (void)userdata;
    const auto t0 = std::chrono::steady_clock::now();
    if (!means)
        return nullptr;
    const uint64_t n = means->element_count / 3u;
    if (!floats_ok(means, 3 * n) || !floats_ok(quaternions, 4 * n) || !floats_ok(scales, 3 * n) ||
        !floats_ok(colors, 3 * n) || !floats_ok(opacities, n))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_to_nanovdb: array lengths disagree with %llu Gaussians",
                    (unsigned long long)n);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_default_mutex);
    Context* ctx = default_context();
    if (!ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_to_nanovdb: no HIP device");
        return nullptr;
    }
    std::vector<uint8_t> grid;
    vdbr_raster_stats_t st;
    memset(&st, 0, sizeof(st));
    if (run_host(ctx, n, (const float*)means->data, (const float*)quaternions->data, (const float*)scales->data,
                 (const float*)colors->data, (const float*)opacities->data, voxel_size,
                 cull_scale_of(shader_params_arrays), 0u, &grid, nullptr, nullptr, nullptr, false, &st))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_to_nanovdb: %s", ctx->error.c_str());
        return nullptr;
    }
    pnanovdb_compute_array_t* out = make_output(compute, grid);
    report_profile(profiler_report, st, 0.f, seconds_since(t0));
    return out;
}

// Raster.cpp:420-558: the Gaussian data's device arrays (uploaded once) -> node2 grid written into
// the caller's device buffer of nanovdb_word_count words (zero-filled past the grid). A grid larger
// than the buffer is an error: nothing is written and the message is logged.
static void PNANOVDB_ABI r_raster_gaussian_3d(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                              pnanovdb_raster_context_t* context, float voxel_size,
                                              pnanovdb_raster_gaussian_data_t* data, pnanovdb_compute_buffer_t* nanovdb_out,
                                              pnanovdb_uint64_t nanovdb_word_count, void* userdata)
{
    (void)context;
    (void)userdata;
    GaussianDataImpl* d = (GaussianDataImpl*)data;
    vdbr_buffer_t* out = (vdbr_buffer_t*)nanovdb_out;
    if (!d || !out || !out->device_ptr)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_gaussian_3d: null gaussian data or output buffer");
        return;
    }
    std::lock_guard<std::mutex> lock(g_default_mutex);
    Context* ctx = default_context();
    if (!ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_gaussian_3d: no HIP device");
        return;
    }
    if (upload_data(d, ctx))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_gaussian_3d: %s", ctx->error.c_str());
        return;
    }
    vdbr_gaussians_t g;
    g.means = d->d_means();
    g.opacities = d->d_opac();
    g.quats = d->d_quats();
    g.scales = d->d_scales();
    g.colors = d->d_colors();
    g.sh0 = nullptr;
    g.count = d->count;
    vdbr_raster_params_t p;
    memset(&p, 0, sizeof(p));
    p.voxel_size = voxel_size;
    p.cull_scale = d->cull_scale;
    p.value_mode = VDBR_VALUE_RGBA8;
    void* grid = nullptr;
    uint64_t size = 0;
    vdbr_raster_stats_t st;
    if (raster_node2(ctx, &g, &p, &grid, &size, &st))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_gaussian_3d: %s", ctx->error.c_str());
        return;
    }
    uint64_t cap = nanovdb_word_count * 4u;
    if (out->size_bytes < cap)
        cap = out->size_bytes;
    if (size > cap)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR,
                    "raster_gaussian_3d: the grid needs %llu bytes, the output buffer holds %llu",
                    (unsigned long long)size, (unsigned long long)cap);
        return;
    }
    hipError_t e = hipMemcpyAsync(out->device_ptr, grid, size, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess && cap > size)
        e = hipMemsetAsync((uint8_t*)out->device_ptr + size, 0, cap - size, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        log_message(compute, queue, LOG_LEVEL_ERROR, "raster_gaussian_3d: %s", hipGetErrorString(e));
}

// raw Gaussian arrays (means, opacities, quats, scales, sh_0[, sh_n]) of a table call: present,
// float32 and long enough for the Gaussian count of the means
static bool raw_arrays_ok(pnanovdb_compute_array_t** arrays, uint32_t count, uint64_t* n_out)
{
    if (!arrays || count < 5u || !arrays[0])
        return false;
    const uint64_t n = arrays[0]->element_count / 3u;
    const uint64_t need[5] = { 3 * n, n, 4 * n, 3 * n, 3 * n };
    for (int i = 0; i < 5; i++)
        if (!floats_ok(arrays[i], need[i]))
            return false;
    *n_out = n;
    return true;
}

// RasterUtils.cpp:143-207, 400-415: raw arrays (means, opacities, quats, scales, sh_0, sh_n);
// the caller's opacity/quaternion/scale arrays receive their preprocessed values in place
static pnanovdb_bool_t PNANOVDB_ABI r_raster_to_nanovdb_from_arrays(pnanovdb_raster_t* raster,
                                                                    const pnanovdb_compute_t* compute,
                                                                    pnanovdb_compute_queue_t* queue, float voxel_size,
                                                                    pnanovdb_compute_array_t** arrays,
                                                                    pnanovdb_uint32_t array_count,
                                                                    pnanovdb_compute_array_t** out)
{
    (void)raster;
    uint64_t n = 0;
    if (!out || !raw_arrays_ok(arrays, array_count, &n))
        return PNANOVDB_FALSE;
    std::lock_guard<std::mutex> lock(g_default_mutex);
    Context* ctx = default_context();
    if (!ctx)
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "Failed to rasterize Gaussian data: no HIP device");
        return PNANOVDB_FALSE;
    }
    std::vector<uint8_t> grid;
    if (run_host(ctx, n, (const float*)arrays[0]->data, (const float*)arrays[2]->data, (const float*)arrays[3]->data,
                 (const float*)arrays[4]->data, (const float*)arrays[1]->data, voxel_size, 0.f, VDBR_FLAG_RAW_INPUTS,
                 &grid, (float*)arrays[2]->data, (float*)arrays[3]->data, (float*)arrays[1]->data, true))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "Failed to rasterize Gaussian data: %s", ctx->error.c_str());
        return PNANOVDB_FALSE;
    }
    *out = make_output(compute, grid);
    return *out ? PNANOVDB_TRUE : PNANOVDB_FALSE;
}

// RasterUtils.cpp:209-290, 417-441 (process_arrays_to_create_gaussian_data): the caller's raw arrays
// are preprocessed in place (quaternions normalised, scales exponentiated, opacities through the
// sigmoid; colours from sh_0), then Gaussian data is created from them. The preprocessing runs on
// the device and its results stay there as the data's uploaded arrays.
static pnanovdb_bool_t PNANOVDB_ABI r_create_gaussian_data_from_arrays(
    pnanovdb_raster_t* raster, const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
    pnanovdb_compute_array_t** arrays, pnanovdb_uint32_t array_count, pnanovdb_raster_gaussian_data_t** gaussian_data,
    pnanovdb_raster_shader_params_t* raster_params, pnanovdb_raster_context_t** raster_context)
{
    (void)raster;
    (void)raster_params;
    if (pnanovdb_compute_log_print_t lp = host_log_print(compute, queue))
        lp(LOG_LEVEL_WARNING, "pnanovdb_raster::create_gaussian_data_from_arrays is deprecated and will be removed in "
                              "version 0.2.0. Please use pnanovdb_raster::create_gaussian_data_from_desc instead.");
    uint64_t n = 0;
    if (!gaussian_data || !raw_arrays_ok(arrays, array_count, &n))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR,
                    "create_gaussian_data_from_arrays: need float32 means[3N], opacities[N], quaternions[4N], "
                    "scales[3N] and sh_0[3N]");
        return PNANOVDB_FALSE;
    }
    GaussianDataImpl* d = new (std::nothrow) GaussianDataImpl();
    if (!d)
        return PNANOVDB_FALSE;
    d->count = n;
    std::lock_guard<std::mutex> lock(g_default_mutex);
    Context* ctx = default_context();
    auto fail = [&](const char* what) {
        log_message(compute, queue, LOG_LEVEL_ERROR, "create_gaussian_data_from_arrays: %s", what);
        delete d;
        return PNANOVDB_FALSE;
    };
    if (!ctx)
        return fail("no HIP device");
    try
    {
        copy_floats(d->means, arrays[0], 3 * n);
        d->quats.resize(4 * n);
        d->scales.resize(3 * n);
        d->colors.resize(3 * n);
        d->opacities.resize(n);
    }
    catch (...)
    {
        return fail("host allocation failed");
    }
    void* p = nullptr;
    if (hipMalloc(&p, (n > 0 ? n : 1) * 19 * sizeof(float)) != hipSuccess) // 16 data + 3 sh_0 floats per Gaussian
    {
        (void)hipGetLastError();
        return fail("device allocation failed");
    }
    d->dev = (float*)p;
    d->device = ctx->device;
    float* d_sh0 = d->dev + 16 * n;
    if (n > 0)
    {
        hipError_t e = hipMemcpyAsync(d->d_means(), arrays[0]->data, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->d_opac(), arrays[1]->data, n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->d_quats(), arrays[2]->data, 4 * n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->d_scales(), arrays[3]->data, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_sh0, arrays[4]->data, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess)
            return fail(hipGetErrorString(e));
        if (preprocess_device(ctx, n, d->d_opac(), d->d_quats(), d->d_scales(), d_sh0, d->d_colors()))
            return fail(ctx->error.c_str());
        e = hipMemcpyAsync(d->quats.data(), d->d_quats(), 4 * n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->scales.data(), d->d_scales(), 3 * n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->opacities.data(), d->d_opac(), n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->colors.data(), d->d_colors(), 3 * n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess)
            return fail(hipGetErrorString(e));
        // RasterUtils.cpp:59-117 preprocesses the caller's arrays in place
        memcpy(arrays[1]->data, d->opacities.data(), n * 4);
        memcpy(arrays[2]->data, d->quats.data(), 4 * n * 4);
        memcpy(arrays[3]->data, d->scales.data(), 3 * n * 4);
    }
    d->uploaded = true;
    *gaussian_data = (pnanovdb_raster_gaussian_data_t*)d;
    if (raster_context && !*raster_context)
        *raster_context = r_create_context(compute, queue);
    return PNANOVDB_TRUE;
}

// RasterUtils.cpp:295-395: Gaussian file -> OnIndex grid (or gaussian data)
static pnanovdb_bool_t PNANOVDB_ABI r_raster_file(pnanovdb_raster_t* raster, const pnanovdb_compute_t* compute,
                                                  pnanovdb_compute_queue_t* queue, const char* filename,
                                                  float voxel_size, pnanovdb_compute_array_t** nanovdb_arr,
                                                  pnanovdb_raster_gaussian_data_t** gaussian_data,
                                                  pnanovdb_raster_context_t** raster_context,
                                                  pnanovdb_compute_array_t** shader_params_arrays,
                                                  pnanovdb_raster_shader_params_t* raster_params,
                                                  pnanovdb_profiler_report_t profiler_report, void* userdata)
{
    (void)userdata;
    const auto t0 = std::chrono::steady_clock::now();
    const pnanovdb_compute_t* c = (compute && compute->create_array) ? compute : compute_table();
    const char* names[6] = { "means", "opacities", "quaternions", "scales", "sh_0", "sh_n" };
    pnanovdb_compute_array_t* arrays[6] = {};
    std::string err;
    FileArrayAlloc alloc = [&](size_t es, uint64_t count, const void* data) { return c->create_array(es, count, data); };
    if (!ff_load(filename, 6, names, arrays, alloc, &err))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR, "Failed to load Gaussian data from file '%s': %s",
                    filename ? filename : "", err.c_str());
        for (int i = 0; i < 6; i++)
            if (arrays[i])
                c->destroy_array(arrays[i]);
        return PNANOVDB_FALSE;
    }
    pnanovdb_bool_t ok = PNANOVDB_TRUE;
    uint64_t n = 0;
    if (!raw_arrays_ok(arrays, 6, &n))
    {
        log_message(compute, queue, LOG_LEVEL_ERROR,
                    "'%s': Gaussian arrays missing, not float32 or of disagreeing lengths", filename);
        ok = PNANOVDB_FALSE;
    }
    else if (gaussian_data && raster_context)
        ok = r_create_gaussian_data_from_arrays(raster, compute, queue, arrays, 6, gaussian_data, raster_params,
                                                raster_context);
    else if (!nanovdb_arr)
        ok = PNANOVDB_FALSE;
    else
    {
        // RasterUtils.cpp:143-207: preprocess, rasterize with the call's shader parameters, OnIndex
        std::lock_guard<std::mutex> lock(g_default_mutex);
        Context* ctx = default_context();
        std::vector<uint8_t> grid;
        vdbr_raster_stats_t st;
        memset(&st, 0, sizeof(st));
        float ms_convert = 0.f;
        if (!ctx || run_host(ctx, n, (const float*)arrays[0]->data, (const float*)arrays[2]->data,
                             (const float*)arrays[3]->data, (const float*)arrays[4]->data,
                             (const float*)arrays[1]->data, voxel_size, cull_scale_of(shader_params_arrays),
                             VDBR_FLAG_RAW_INPUTS, &grid, nullptr, nullptr, nullptr, true, &st, &ms_convert))
        {
            log_message(compute, queue, LOG_LEVEL_ERROR, "Failed to rasterize Gaussian data: %s",
                        ctx ? ctx->error.c_str() : "no HIP device");
            ok = PNANOVDB_FALSE;
        }
        else
        {
            *nanovdb_arr = make_output(compute, grid);
            ok = *nanovdb_arr ? PNANOVDB_TRUE : PNANOVDB_FALSE;
            report_profile(profiler_report, st, ms_convert, seconds_since(t0));
        }
    }
    for (int i = 0; i < 6; i++)
        if (arrays[i])
            c->destroy_array(arrays[i]);
    return ok;
}

static pnanovdb_raster_t g_raster;
static bool g_raster_init = false;

// FileFormat.cpp:795-806
static pnanovdb_bool_t PNANOVDB_ABI ff_can_load_file(const char* filename)
{
    return ff_can_load(filename) ? PNANOVDB_TRUE : PNANOVDB_FALSE;
}

static pnanovdb_bool_t PNANOVDB_ABI ff_load_file(const char* filename, pnanovdb_uint32_t array_count,
                                                 const char** array_names, pnanovdb_compute_array_t** out_arrays)
{
    const pnanovdb_compute_t* c = compute_table();
    std::string err;
    FileArrayAlloc alloc = [&](size_t es, uint64_t count, const void* data) { return c->create_array(es, count, data); };
    if (!ff_load(filename, array_count, array_names, out_arrays, alloc, &err))
    {
        fprintf(stderr, "Error: Failed to load file: %s (%s)\n", filename ? filename : "", err.c_str());
        return PNANOVDB_FALSE;
    }
    return PNANOVDB_TRUE;
}

static pnanovdb_fileformat_t g_fileformat;
static bool g_fileformat_init = false;

// ============================================================================= parallel primitives
static pnanovdb_parallel_primitives_context_t* PNANOVDB_ABI pp_create_context(const pnanovdb_compute_t* compute,
                                                                             pnanovdb_compute_queue_t* queue)
{
    (void)compute;
    (void)queue;
    return (pnanovdb_parallel_primitives_context_t*)create_context(0);
}

static void PNANOVDB_ABI pp_destroy_context(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                            pnanovdb_parallel_primitives_context_t* context)
{
    (void)compute;
    (void)queue;
    destroy_context((Context*)context);
}

static void* buf_ptr(pnanovdb_compute_buffer_t* b)
{
    return b ? ((vdbr_buffer_t*)b)->device_ptr : nullptr;
}

static void PNANOVDB_ABI pp_global_scan(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                        pnanovdb_parallel_primitives_context_t* context, pnanovdb_compute_buffer_t* val_in,
                                        pnanovdb_compute_buffer_t* val_out, pnanovdb_uint64_t val_count,
                                        pnanovdb_uint32_t dispatch_count)
{
    (void)compute;
    (void)queue;
    (void)dispatch_count;
    Context* ctx = (Context*)context;
    if (!ctx || val_count == 0)
        return;
    scan_u32(ctx, (const uint32_t*)buf_ptr(val_in), (uint32_t*)buf_ptr(val_out), val_count, true, nullptr);
    (void)hipStreamSynchronize(ctx->stream);
}

static void PNANOVDB_ABI pp_global_scan_uint64(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                               pnanovdb_parallel_primitives_context_t* context,
                                               pnanovdb_compute_buffer_t* val_in, pnanovdb_compute_buffer_t* val_out,
                                               pnanovdb_uint64_t val_count, pnanovdb_uint32_t dispatch_count)
{
    (void)compute;
    (void)queue;
    (void)dispatch_count;
    Context* ctx = (Context*)context;
    if (!ctx || val_count == 0)
        return;
    scan_u64(ctx, (const uint64_t*)buf_ptr(val_in), (uint64_t*)buf_ptr(val_out), val_count, true, nullptr);
    (void)hipStreamSynchronize(ctx->stream);
}

static void PNANOVDB_ABI pp_global_scan_max(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                            pnanovdb_parallel_primitives_context_t* context,
                                            pnanovdb_compute_buffer_t* val_in, pnanovdb_compute_buffer_t* val_out,
                                            pnanovdb_uint64_t val_count, pnanovdb_uint32_t dispatch_count)
{
    (void)compute;
    (void)queue;
    (void)dispatch_count;
    Context* ctx = (Context*)context;
    if (!ctx || val_count == 0)
        return;
    scan_max_u32(ctx, (const uint32_t*)buf_ptr(val_in), (uint32_t*)buf_ptr(val_out), val_count, true);
    (void)hipStreamSynchronize(ctx->stream);
}

static void PNANOVDB_ABI pp_radix_sort(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                       pnanovdb_parallel_primitives_context_t* context,
                                       pnanovdb_compute_buffer_t* key_inout, pnanovdb_compute_buffer_t* val_inout,
                                       pnanovdb_uint64_t key_count, pnanovdb_uint64_t buffer_key_count,
                                       pnanovdb_uint32_t key_bit_count)
{
    (void)compute;
    (void)queue;
    (void)buffer_key_count;
    if (!context || key_count == 0)
        return;
    Context* ctx = (Context*)context;
    radix_sort_u32(ctx, (uint32_t*)buf_ptr(key_inout), (uint32_t*)buf_ptr(val_inout), key_count,
                   key_bit_count > 32 ? 32 : key_bit_count);
    (void)hipStreamSynchronize(ctx->stream);
}

static void PNANOVDB_ABI pp_radix_sort_key64(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                             pnanovdb_parallel_primitives_context_t* context,
                                             pnanovdb_compute_buffer_t* key_inout, pnanovdb_compute_buffer_t* val_inout,
                                             pnanovdb_uint64_t key_count, pnanovdb_uint64_t buffer_key_count,
                                             pnanovdb_uint32_t key_bit_count)
{
    (void)compute;
    (void)queue;
    (void)buffer_key_count;
    Context* ctx = (Context*)context;
    if (!ctx || key_count == 0)
        return;
    radix_sort_u64(ctx, (uint64_t*)buf_ptr(key_inout), (uint32_t*)buf_ptr(val_inout), key_count, 0,
                   key_bit_count > 64 ? 64 : key_bit_count, nullptr);
    (void)hipStreamSynchronize(ctx->stream);
}

static void PNANOVDB_ABI pp_radix_sort_dual_key(const pnanovdb_compute_t* compute, pnanovdb_compute_queue_t* queue,
                                                pnanovdb_parallel_primitives_context_t* context,
                                                pnanovdb_compute_buffer_t* key_low_inout,
                                                pnanovdb_compute_buffer_t* key_high_inout,
                                                pnanovdb_compute_buffer_t* val_inout, pnanovdb_uint64_t key_count,
                                                pnanovdb_uint64_t buffer_key_count, pnanovdb_uint32_t key_low_bit_count,
                                                pnanovdb_uint32_t key_high_bit_count)
{
    (void)compute;
    (void)queue;
    (void)buffer_key_count;
    Context* ctx = (Context*)context;
    if (!ctx || key_count == 0)
        return;
    if (radix_sort_dual_key(ctx, (uint32_t*)buf_ptr(key_low_inout), (uint32_t*)buf_ptr(key_high_inout),
                            (uint32_t*)buf_ptr(val_inout), key_count, key_low_bit_count, key_high_bit_count))
        fprintf(stderr, "radix_sort_dual_key: %s\n", ctx->error.c_str());
    (void)hipStreamSynchronize(ctx->stream);
}

static pnanovdb_parallel_primitives_t g_pp;
static bool g_pp_init = false;

// ============================================================================= exports
extern "C" {

PNANOVDB_API pnanovdb_compute_t* pnanovdb_get_compute(void)
{
    return (pnanovdb_compute_t*)compute_table();
}

PNANOVDB_API pnanovdb_raster_t* pnanovdb_get_raster(void)
{
    if (!g_raster_init)
    {
        memset(&g_raster, 0, sizeof(g_raster));
        g_raster.compute = compute_table();
        g_raster.create_context = r_create_context;
        g_raster.destroy_context = r_destroy_context;
        g_raster.create_gaussian_data = r_create_gaussian_data;
        g_raster.upload_gaussian_data = r_upload_gaussian_data;
        g_raster.destroy_gaussian_data = r_destroy_gaussian_data;
        g_raster.raster_gaussian_3d = r_raster_gaussian_3d;
        g_raster.raster_to_nanovdb = r_raster_to_nanovdb;
        g_raster.raster_file = r_raster_file;
        g_raster.raster_to_nanovdb_from_arrays = r_raster_to_nanovdb_from_arrays;
        g_raster.create_gaussian_data_from_arrays = r_create_gaussian_data_from_arrays;
        g_raster_init = true;
    }
    return &g_raster;
}

PNANOVDB_API pnanovdb_fileformat_t* pnanovdb_get_fileformat(void)
{
    if (!g_fileformat_init)
    {
        memset(&g_fileformat, 0, sizeof(g_fileformat));
        g_fileformat.can_load_file = ff_can_load_file;
        g_fileformat.load_file = ff_load_file;
        g_fileformat.module = own_module();
        g_fileformat_init = true;
    }
    return &g_fileformat;
}

PNANOVDB_API pnanovdb_parallel_primitives_t* pnanovdb_get_parallel_primitives(void)
{
    if (!g_pp_init)
    {
        memset(&g_pp, 0, sizeof(g_pp));
        g_pp.create_context = pp_create_context;
        g_pp.destroy_context = pp_destroy_context;
        g_pp.global_scan = pp_global_scan;
        g_pp.global_scan_uint64 = pp_global_scan_uint64;
        g_pp.global_scan_max = pp_global_scan_max;
        g_pp.radix_sort = pp_radix_sort;
        g_pp.radix_sort_dual_key = pp_radix_sort_dual_key;
        g_pp.radix_sort_key64 = pp_radix_sort_key64;
        g_pp.compute = compute_table();
        g_pp_init = true;
    }
    return &g_pp;
}

PNANOVDB_API vdbr_context_t* vdbr_context_create(int device)
{
    return (vdbr_context_t*)create_context(device);
}

PNANOVDB_API void vdbr_context_destroy(vdbr_context_t* ctx)
{
    destroy_context((Context*)ctx);
}

PNANOVDB_API const char* vdbr_last_error(vdbr_context_t* ctx)
{
    return ctx ? ((Context*)ctx)->error.c_str() : "null context";
}

PNANOVDB_API int vdbr_context_set_stream(vdbr_context_t* ctx, void* stream)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipStreamSynchronize(c->stream);
    c->stream = (hipStream_t)stream;
    return 0;
}

PNANOVDB_API void* vdbr_context_get_stream(vdbr_context_t* ctx)
{
    return ctx ? (void*)((Context*)ctx)->stream : nullptr;
}

PNANOVDB_API int vdbr_synchronize(vdbr_context_t* ctx)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

PNANOVDB_API int vdbr_copy(vdbr_context_t* ctx, void* dst, const void* src, uint64_t bytes, int kind)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    if (bytes == 0)
        return 0;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    VDBR_CHECK(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

PNANOVDB_API int vdbr_context_trim(vdbr_context_t* ctx)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipSetDevice(c->device);
    for (Context* w : c->batch_workers)
        w->release();
    c->release();
    c->last_voxel_ijk = nullptr;
    c->last_value_count = 0;
    return 0;
}

PNANOVDB_API uint64_t vdbr_workspace_bytes(vdbr_context_t* ctx)
{
    if (!ctx)
        return 0;
    // the context's slots, its batch workers' slots and the batch output chunks
    const Context* c = (const Context*)ctx;
    uint64_t b = c->workspace_bytes();
    for (const Context* w : c->batch_workers)
        b += w->workspace_bytes();
    for (const auto& ch : c->batch_chunks)
        b += ch.second;
    return b;
}

PNANOVDB_API int vdbr_release_workspace(vdbr_context_t* ctx)
{
    if (!ctx)
        return -1;
    ((Context*)ctx)->release();
    return 0;
}

PNANOVDB_API int vdbr_raster_node2(vdbr_context_t* ctx, const vdbr_gaussians_t* gaussians,
                                   const vdbr_raster_params_t* params, void** grid_dev, uint64_t* grid_size,
                                   vdbr_raster_stats_t* stats)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipSetDevice(c->device);
    return raster_node2(c, gaussians, params, grid_dev, grid_size, stats);
}

PNANOVDB_API int vdbr_raster_node2_batch(vdbr_context_t* ctx, vdbr_batch_item_t* items, uint32_t count,
                                         uint32_t workers)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipSetDevice(c->device);
    return raster_batch(c, items, count, workers);
}

PNANOVDB_API uint32_t vdbr_last_batch_graph_clouds(vdbr_context_t* ctx)
{
    Context* c = (Context*)ctx;
    return c ? c->last_batch_graph_clouds : 0u;
}

PNANOVDB_API int vdbr_last_batch_timing(vdbr_context_t* ctx, float* ms)
{
    Context* c = (Context*)ctx;
    if (!c || !ms)
        return -1;
    for (int i = 0; i < 5; i++)
        ms[i] = c->batch_ms[i];
    return 0;
}

PNANOVDB_API int vdbr_last_voxels(vdbr_context_t* ctx, int32_t* ijk_dev, uint64_t cap, uint64_t* count)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return last_voxels(c, ijk_dev, cap, count);
}

PNANOVDB_API int vdbr_node2_sample(vdbr_context_t* ctx, const void* grid_dev, uint64_t grid_size, const uint32_t* values,
                                   uint64_t value_count, const int32_t* ijk_dev, uint64_t count, uint32_t* out_dev)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return node2_sample(c, grid_dev, grid_size, values, value_count, ijk_dev, count, out_dev);
}

PNANOVDB_API int vdbr_last_pairs(vdbr_context_t* ctx, int32_t* ijk_dev, uint32_t* gid_dev, uint64_t cap,
                                 uint64_t* count)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return last_pairs(c, ijk_dev, gid_dev, cap, count);
}

PNANOVDB_API int vdbr_voxelbvh_ijkl(vdbr_context_t* ctx, const vdbr_gaussians_t* gaussians, uint32_t resolution,
                                    const float* transform, uint64_t* ijkl, uint32_t* prim_ids, uint32_t* ranges,
                                    float* world_bbox)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipSetDevice(c->device);
    return voxelbvh_ijkl(c, gaussians, resolution, transform, ijkl, prim_ids, ranges, world_bbox);
}

PNANOVDB_API int vdbr_tile_bin(vdbr_context_t* ctx, const float* means2d, const int32_t* radii, const float* depths,
                               uint32_t count_in, uint32_t image_width, uint32_t image_height, uint32_t tile_size,
                               uint32_t* tile_offsets, uint32_t** keys_low, uint32_t** keys_high, uint32_t** vals,
                               uint64_t* count)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    (void)hipSetDevice(c->device);
    return tile_bin(c, means2d, radii, depths, count_in, image_width, image_height, tile_size, tile_offsets, keys_low,
                    keys_high, vals, count);
}

PNANOVDB_API int vdbr_radix_sort_dual_key(vdbr_context_t* ctx, uint32_t* key_low, uint32_t* key_high, uint32_t* vals,
                                          uint64_t count, uint32_t key_low_bit_count, uint32_t key_high_bit_count)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    if (radix_sort_dual_key(c, key_low, key_high, vals, count, key_low_bit_count, key_high_bit_count))
        return -1;
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

PNANOVDB_API int vdbr_radix_sort_u64(vdbr_context_t* ctx, uint64_t* keys, uint32_t* vals, uint64_t count,
                                     uint32_t bit_begin, uint32_t bit_end)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    if (bit_end > 64 || bit_begin > bit_end)
        return c->fail("radix sort: bit range must satisfy bit_begin <= bit_end <= 64");
    if (radix_sort_u64(c, keys, vals, count, bit_begin, bit_end, nullptr))
        return -1;
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

PNANOVDB_API int vdbr_radix_sort_u64_noalloc(vdbr_context_t* ctx, uint64_t* keys, uint32_t* vals, uint64_t count,
                                             uint32_t bit_begin, uint32_t bit_end)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    std::string err;
    if (radix_sort_u64_noalloc(c->device, keys, vals, count, bit_begin, bit_end, c->stream, &err))
        return c->fail(err);
    return 0;
}

PNANOVDB_API int vdbr_radix_sort_u32(vdbr_context_t* ctx, uint32_t* keys, uint32_t* vals, uint64_t count,
                                     uint32_t key_bit_count)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    if (key_bit_count > 32)
        return c->fail("radix sort: key_bit_count must be <= 32");
    if (radix_sort_u32(c, keys, vals, count, key_bit_count))
        return -1;
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

PNANOVDB_API int vdbr_density_backward(vdbr_context_t* ctx, const vdbr_gaussians_t* raw, float voxel_size,
                                       const void* grid_dev, uint64_t grid_size, const float* grad_values_dev,
                                       uint64_t value_count, const vdbr_gaussian_grads_t* grads)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return density_backward(c, raw, voxel_size, grid_dev, grid_size, grad_values_dev, value_count, grads, nullptr);
}

PNANOVDB_API int vdbr_density_backward_pairs(vdbr_context_t* ctx, const vdbr_gaussians_t* raw, float voxel_size,
                                             const void* grid_dev, uint64_t grid_size, const float* grad_values_dev,
                                             uint64_t value_count, const vdbr_gaussian_grads_t* grads,
                                             uint64_t pair_list_id)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    int used = 0;
    if (density_backward_pairs(c, raw, voxel_size, grid_dev, grid_size, grad_values_dev, value_count, grads,
                               pair_list_id, nullptr, &used))
        return -1;
    return used ? 0 : 1;
}

PNANOVDB_API int vdbr_selftest_alpha8(vdbr_context_t* ctx, uint64_t count, uint64_t seed, uint64_t* mismatches,
                                      uint64_t* fast_decided)
{
    Context* c = (Context*)ctx;
    if (!c || !mismatches)
        return -1;
    return selftest_alpha8(c, count, seed, mismatches, fast_decided);
}

PNANOVDB_API int vdbr_autotune_sort(vdbr_context_t* ctx, uint64_t key_count, uint32_t key_bits,
                                    vdbr_tune_result_t* results, uint32_t cap, uint32_t* count,
                                    uint32_t* best_variant, uint32_t* best_digit_bits)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return autotune_sort(c, key_count, key_bits, results, cap, count, best_variant, best_digit_bits);
}

PNANOVDB_API int vdbr_set_sort_config(vdbr_context_t* ctx, uint64_t from_count, uint32_t variant,
                                      uint32_t digit_bits)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    return set_sort_config(c, from_count, variant, digit_bits);
}

PNANOVDB_API int vdbr_node2_to_onindex(vdbr_context_t* ctx, const void* node2_dev, uint64_t node2_size,
                                       void** onindex_dev, uint64_t* onindex_size, uint64_t* buffer_size)
{
    Context* c = (Context*)ctx;
    if (!c || !onindex_dev || !onindex_size)
        return -1;
    uint64_t alloc = 0;
    int r = node2_to_onindex(c, node2_dev, node2_size, onindex_dev, onindex_size, &alloc);
    if (buffer_size)
        *buffer_size = alloc;
    return r;
}

PNANOVDB_API int vdbr_inclusive_scan(vdbr_context_t* ctx, const void* in, void* out, uint64_t count, int op)
{
    Context* c = (Context*)ctx;
    if (!c)
        return -1;
    int r;
    if (op == 0)
        r = scan_u32(c, (const uint32_t*)in, (uint32_t*)out, count, true, nullptr);
    else if (op == 1)
        r = scan_u64(c, (const uint64_t*)in, (uint64_t*)out, count, true, nullptr);
    else if (op == 2)
        r = scan_max_u32(c, (const uint32_t*)in, (uint32_t*)out, count, true);
    else
        return c->fail("scan: unknown op");
    if (r)
        return r;
    VDBR_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

} // extern "C"
