// vx_api.hip -- host layer of the C ABI (include/vx/vx_api.h).
//
// Owns device volumes and models, validates arguments, plans the rebuild
// batches and launches the gfx950 kernels of vx_kernels.hip. No CPU fallback:
// every pass runs on the HIP device or returns an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "vx_internal.hpp"

using namespace vx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static vx_status fail(vx_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// Runs an entry point's body; C++ exceptions (host allocation) do not cross the C ABI.
template <class F>
static vx_status guarded(const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(VX_ERR_OUT_OF_MEMORY, "%s: host allocation failed", what);
    } catch (const std::exception& e) {
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: %s", what, e.what());
    } catch (...) {
        return fail(VX_ERR_INVALID_ARGUMENT, "%s failed", what);
    }
}

// The host-only sources (asset loaders) report through the same per-thread message.
vx_status vx_set_error(vx_status code, const char* msg) {
    g_last_error = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Failure injection for tests of the error paths: the environment variables
// VX_FAIL_ALLOC=n and VX_FAIL_LAUNCH=n (read when the library loads), or
// vx_debug_inject_failure, make the n-th device allocation / kernel-launch
// check from then on (counting from 0) fail once, with hipErrorOutOfMemory /
// hipErrorLaunchFailure. -1 (the default) injects nothing.
// ---------------------------------------------------------------------------
static int64_t env_count(const char* name) {
    const char* v = std::getenv(name);
    if (!v || !*v) return -1;
    char* end = nullptr;
    const long long n = std::strtoll(v, &end, 10);
    return (end && *end == 0 && n >= 0) ? (int64_t)n : -1;
}
static std::atomic<int64_t> g_fail_alloc{env_count("VX_FAIL_ALLOC")};
static std::atomic<int64_t> g_fail_launch{env_count("VX_FAIL_LAUNCH")};

// True for the one event the counter selects (it then stays at -1).
static bool injected(std::atomic<int64_t>& c) {
    int64_t v = c.load();
    while (v >= 0)
        if (c.compare_exchange_weak(v, v - 1)) return v == 0;
    return false;
}

template <class T>
static hipError_t dev_malloc(T** p, size_t n) {
    if (injected(g_fail_alloc)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, n);
}

// Status of the launches just enqueued (the last launch error), or the
// injected launch failure.
static hipError_t launch_status() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && injected(g_fail_launch)) e = hipErrorLaunchFailure;
    return e;
}

#define VX_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError();                                                          \
            return fail(e_ == hipErrorOutOfMemory ? VX_ERR_OUT_OF_MEMORY : VX_ERR_DEVICE,     \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,      \
                        __LINE__);                                                            \
        }                                                                                     \
    } while (0)

#define VX_REQUIRE(cond, ...)                                                                 \
    do {                                                                                      \
        if (!(cond)) return fail(VX_ERR_INVALID_ARGUMENT, __VA_ARGS__);                       \
    } while (0)

#define VX_LAUNCHED(name)                                                                     \
    do {                                                                                      \
        hipError_t e_ = launch_status();                                                      \
        if (e_ != hipSuccess) return fail(VX_ERR_DEVICE, "launch of %s failed: %s", name,     \
                                          hipGetErrorString(e_));                             \
    } while (0)

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------
// Scratch of the ray-parallel trace: per-ray results, the list of rays left
// to the literal path and the queue counters, sized in ray slots.
struct MarchBuffers {
    float* res;
    uint16_t* fres;
    uint32_t* deferred;
    uint32_t* counters;      // kCounterWords
    size_t slots;            // capacity of res / deferred
    size_t fslots;           // capacity of fres
};

static void march_buffers_free(MarchBuffers& m) {
    (void)hipFree(m.res);
    (void)hipFree(m.fres);
    (void)hipFree(m.deferred);
    (void)hipFree(m.counters);
    std::memset(&m, 0, sizeof(m));
}

// ---------------------------------------------------------------------------
// Scratch of vx_trace_noalloc: one file-scope device array per GPU, sized for
// kNoAllocSlots ray slots (result, listed-ray and fetch-count arrays) plus the
// queue counters. Host code never names the array in a launch: its device
// address is taken once per device with hipGetSymbolAddress (vx_ctx_create)
// and kept in g_noalloc_ptr.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocSlots = size_t(1) << 22;
constexpr size_t kNoAllocCounterBytes = 4096;
constexpr size_t kNoAllocBytes =
    kNoAllocCounterBytes + kNoAllocSlots * (sizeof(float) + sizeof(uint32_t) + sizeof(uint16_t));
static_assert(kCounterWords * sizeof(uint32_t) <= kNoAllocCounterBytes, "queue counters fit their region");
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static uint8_t* g_noalloc_ptr[kMaxDevices];

// Device address of this device's scratch (the device must be current).
static hipError_t noalloc_scratch(int device, uint8_t** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[device] = static_cast<uint8_t*>(p);
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

// Persistent-grid size of one march kernel, cached per context: the device's
// CU count and the kernel's occupancy do not change between launches.
struct GridCache {
    int cus;
    uint32_t grid[2][2][4];   // [single][block 256/512][steps 2/4/8/16]: 0 = not yet asked
};

struct vx_ctx {
    int32_t device;
    GridCache gcache;
    uint8_t* noalloc;        // this device's vx_trace_noalloc scratch (device address)
    long long* grad_acc;     // 256*6+6 fixed-point gradient accumulator
    ShadeArgs* shade_items;  // batched adjoint: per-view arguments (device)
    uint32_t* shade_block0;  // and each view's first block (n + 1 entries)
    uint32_t shade_cap;
    MarchBuffers march;
    vx_trace_tuning tuning;
    bool profile;            // record events around march kernels
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof;
    InstConst* inst;         // multi-instance primary: per-draw constants
    int32_t* inst_rects;     // and screen rectangles (device)
    uint32_t inst_cap;
    uint16_t* tile_lists;    // per-tile instance lists
    uint32_t* tile_counts;
    uint32_t tile_cap;
    void* taa_rec;           // LightTAA: packed per-pixel attributes (64 B each)
    uint32_t taa_cap;
    // light pass ray list: per pixel first record and count, the total, the records
    uint32_t* lt_first;
    uint32_t* lt_count;
    uint32_t* lt_total;
    size_t lt_px_cap;
    LightRec* lt_rec;
    float* lt_shadow;
    uint16_t* lt_slot_light;
    size_t lt_ray_cap;
    uint16_t* lt_rfetch;
    size_t lt_fetch_cap;
    int32_t* rb_aabb;        // rebuild: per-instance AABB slots
    size_t rb_aabb_cap;
    FootprintOp* rb_ops;     // rebuild: footprint operations
    size_t rb_ops_cap;
};

static void profile_clear(vx_ctx* c) {
    for (auto& p : c->prof) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    c->prof.clear();
}

struct vx_model {
    vx_ctx* ctx;
    int32_t dims[3];
    uint8_t* mip[3];
};

struct vx_volume {
    vx_ctx* ctx;
    int32_t sx, sy, sz;
    int32_t nlevels;
    int32_t bx, by, bz;
    int32_t wx[3], wy[3], wz[3];    // level word grids
    int32_t cx[3], cy[3], cz[3];    // level cell grids
    uint64_t* cells;
    uint64_t* lvl[3];
    size_t ncells;
    // March grid of the ray-parallel trace: the cells in a linear grid with a
    // border cell per side and out-of-bounds voxel bits set; kept in step with
    // cells by every write (clear, upload, rebuild).
    uint64_t* grid;
    int32_t gx, gy, gz;
    size_t ngrid;
};

// Refreshes the march grid over texel box [t0, t1] (inclusive), clamped to the
// volume, plus the border cells next to it.
static hipError_t mgrid_update(const vx_volume* v, const int32_t t0[3], const int32_t t1[3], hipStream_t s) {
    const int32_t gdim[3] = {v->gx, v->gy, v->gz};
    const int32_t tdim[3] = {v->sx, v->sy, v->sz};
    int32_t g0[3], n[3];
    for (int k = 0; k < 3; ++k) {
        int32_t lo = std::max(t0[k], 0), hi = std::min(t1[k], tdim[k] - 1);
        if (hi < lo) return hipSuccess;
        // fine voxels [2lo, 2hi+1] -> grid cells ((f >> 2) + 1), widened by one cell
        int32_t c0 = std::max(((2 * lo) >> 2) + 1 - 1, 0);
        int32_t c1 = std::min(((2 * hi + 1) >> 2) + 1 + 1, gdim[k] - 1);
        g0[k] = c0;
        n[k] = c1 - c0 + 1;
    }
    size_t total = (size_t)n[0] * n[1] * n[2];
    OccView V;
    std::memset(&V, 0, sizeof(V));
    V.cells = v->cells;
    V.sx = v->sx; V.sy = v->sy; V.sz = v->sz;
    V.bx = v->bx; V.by = v->by; V.bz = v->bz;
    hipLaunchKernelGGL(k_mgrid, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, v->grid, V, v->gx, v->gy,
                       g0[0], g0[1], g0[2], n[0], n[1], n[2]);
    return hipGetLastError();
}

// The whole march grid (every cell, border included).
static hipError_t mgrid_full(const vx_volume* v, hipStream_t s) {
    OccView V;
    std::memset(&V, 0, sizeof(V));
    V.cells = v->cells;
    V.sx = v->sx; V.sy = v->sy; V.sz = v->sz;
    V.bx = v->bx; V.by = v->by; V.bz = v->bz;
    hipLaunchKernelGGL(k_mgrid, dim3((uint32_t)((v->ngrid + 255) / 256)), dim3(256), 0, s, v->grid, V, v->gx, v->gy,
                       0, 0, 0, v->gx, v->gy, v->gz);
    return hipGetLastError();
}

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

static vx_status use_device(const vx_ctx* ctx) {
    VX_HIP(hipSetDevice(ctx->device));
    return VX_OK;
}

static OccView occ_view(const vx_volume* v) {
    OccView V;
    V.cells = v->cells;
    for (int l = 0; l < 3; ++l) {
        V.lvl[l] = v->lvl[l];
        V.wx[l] = v->wx[l]; V.wy[l] = v->wy[l]; V.wz[l] = v->wz[l];
    }
    V.sx = v->sx; V.sy = v->sy; V.sz = v->sz;
    V.bx = v->bx; V.by = v->by; V.bz = v->bz;
    V.nlevels = v->nlevels;
    return V;
}

static TileMap tile_map(int32_t w, int32_t h) {
    TileMap t;
    t.width = w;
    t.height = h;
    t.tiles_x = cdiv((uint32_t)w, 16);
    t.ntiles = t.tiles_x * cdiv((uint32_t)h, 16);
    t.per_xcd = cdiv(t.ntiles, 8);
    return t;
}

static vx_status check_view(const vx_view* view) {
    VX_REQUIRE(view, "view is NULL");
    VX_REQUIRE(view->width > 0 && view->height > 0 && view->width <= 16384 && view->height <= 16384,
               "view size %dx%d outside 1..16384", view->width, view->height);
    return VX_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int32_t vx_api_version(void) { return VX_API_VERSION; }

extern "C" vx_status vx_debug_inject_failure(int32_t kind, int64_t after) {
    VX_REQUIRE(kind == VX_FAIL_ALLOC || kind == VX_FAIL_LAUNCH, "kind %d is not VX_FAIL_ALLOC/VX_FAIL_LAUNCH", kind);
    VX_REQUIRE(after >= -1, "after %lld below -1", (long long)after);
    (kind == VX_FAIL_ALLOC ? g_fail_alloc : g_fail_launch).store(after);
    return VX_OK;
}
extern "C" const char* vx_last_error(void) { return g_last_error.c_str(); }

extern "C" vx_status vx_ctx_create(int32_t device, vx_ctx** out) {
    VX_REQUIRE(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(VX_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    }
    VX_REQUIRE(device >= 0 && device < n, "device %d out of range (0..%d)", device, n - 1);
    VX_REQUIRE(device < kMaxDevices, "device %d beyond the library's %d-device limit", device, kMaxDevices);
    VX_HIP(hipSetDevice(device));
    uint8_t* noalloc = nullptr;
    VX_HIP(noalloc_scratch(device, &noalloc));
    int cus = 0;
    VX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    vx_ctx* c = new (std::nothrow) vx_ctx();
    if (!c) return fail(VX_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->device = device;
    std::memset(&c->gcache, 0, sizeof(c->gcache));
    c->gcache.cus = cus;
    c->noalloc = noalloc;
    c->grad_acc = nullptr;
    c->shade_items = nullptr;
    c->shade_block0 = nullptr;
    c->shade_cap = 0;
    std::memset(&c->march, 0, sizeof(c->march));
    std::memset(&c->tuning, 0, sizeof(c->tuning));
    c->profile = false;
    c->inst = nullptr;
    c->inst_rects = nullptr;
    c->inst_cap = 0;
    c->tile_lists = nullptr;
    c->tile_counts = nullptr;
    c->tile_cap = 0;
    c->taa_rec = nullptr;
    c->taa_cap = 0;
    c->lt_first = c->lt_count = c->lt_total = nullptr;
    c->lt_px_cap = 0;
    c->lt_rec = nullptr;
    c->lt_shadow = nullptr;
    c->lt_slot_light = nullptr;
    c->lt_ray_cap = 0;
    c->lt_rfetch = nullptr;
    c->lt_fetch_cap = 0;
    c->rb_aabb = nullptr;
    c->rb_aabb_cap = 0;
    c->rb_ops = nullptr;
    c->rb_ops_cap = 0;
    hipError_t me = dev_malloc(&c->grad_acc, (256 * 6 + 6) * sizeof(long long));
    if (me == hipSuccess) me = dev_malloc(&c->march.counters, kCounterWords * sizeof(uint32_t));
    if (me != hipSuccess) {
        (void)hipFree(c->grad_acc);
        march_buffers_free(c->march);
        delete c;
        return fail(VX_ERR_OUT_OF_MEMORY, "context scratch allocation failed");
    }
    *out = c;
    return VX_OK;
}

extern "C" vx_status vx_ctx_destroy(vx_ctx* ctx) {
    if (!ctx) return VX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipFree(ctx->grad_acc);
    (void)hipFree(ctx->shade_items);
    (void)hipFree(ctx->shade_block0);
    march_buffers_free(ctx->march);
    (void)hipFree(ctx->inst);
    (void)hipFree(ctx->inst_rects);
    (void)hipFree(ctx->tile_lists);
    (void)hipFree(ctx->tile_counts);
    (void)hipFree(ctx->taa_rec);
    (void)hipFree(ctx->lt_first);
    (void)hipFree(ctx->lt_count);
    (void)hipFree(ctx->lt_total);
    (void)hipFree(ctx->lt_rec);
    (void)hipFree(ctx->lt_shadow);
    (void)hipFree(ctx->lt_slot_light);
    (void)hipFree(ctx->lt_rfetch);
    (void)hipFree(ctx->rb_aabb);
    (void)hipFree(ctx->rb_ops);
    profile_clear(ctx);
    delete ctx;
    return VX_OK;
}

extern "C" vx_status vx_ctx_set_trace_tuning(vx_ctx* ctx, const vx_trace_tuning* t) {
    VX_REQUIRE(ctx, "ctx is NULL");
    if (!t) {
        std::memset(&ctx->tuning, 0, sizeof(ctx->tuning));
        return VX_OK;
    }
    VX_REQUIRE(t->kernel >= 0 && t->kernel <= 2, "tuning.kernel %d outside 0..2", t->kernel);
    VX_REQUIRE(t->block == 0 || t->block == 256 || t->block == 512, "tuning.block %d not 256/512", t->block);
    VX_REQUIRE(t->blocks_per_cu >= 0 && t->blocks_per_cu <= 64, "tuning.blocks_per_cu outside 0..64");
    VX_REQUIRE(t->refill >= 0 && t->refill <= 64, "tuning.refill outside 0..64");
    VX_REQUIRE(t->steps == 0 || t->steps == 2 || t->steps == 4 || t->steps == 8 || t->steps == 16,
               "tuning.steps %d not 2/4/8/16", t->steps);
    VX_REQUIRE(t->tile >= 0 && t->tile <= 2, "tuning.tile %d outside 0..2", t->tile);
    VX_REQUIRE(t->order == 0 || t->order == 1, "tuning.order %d not 0/1", t->order);
    VX_REQUIRE(t->lights == 0 || t->lights == 1, "tuning.lights %d not 0/1", t->lights);
    VX_REQUIRE(t->chunks >= 0 && t->chunks <= 64, "tuning.chunks %d outside 0..64", t->chunks);
    ctx->tuning = *t;
    return VX_OK;
}

extern "C" vx_status vx_ctx_get_trace_tuning(const vx_ctx* ctx, vx_trace_tuning* t) {
    VX_REQUIRE(ctx && t, "ctx/tuning is NULL");
    *t = ctx->tuning;
    return VX_OK;
}

extern "C" vx_status vx_ctx_profile(vx_ctx* ctx, int32_t enable) {
    VX_REQUIRE(ctx, "ctx is NULL");
    if (vx_status st = use_device(ctx)) return st;
    if (enable) {
        VX_HIP(hipDeviceSynchronize());
        profile_clear(ctx);
    }
    ctx->profile = enable != 0;
    return VX_OK;
}

extern "C" vx_status vx_ctx_profile_read(vx_ctx* ctx, double* march_ms, uint32_t* launches) {
    VX_REQUIRE(ctx && march_ms && launches, "ctx/march_ms/launches is NULL");
    if (vx_status st = use_device(ctx)) return st;
    double total = 0.0;
    for (auto& p : ctx->prof) {
        VX_HIP(hipEventSynchronize(p.second));
        float ms = 0.0f;
        VX_HIP(hipEventElapsedTime(&ms, p.first, p.second));
        total += ms;
    }
    *march_ms = total;
    *launches = (uint32_t)ctx->prof.size();
    return VX_OK;
}

// ---------------------------------------------------------------------------
// models (VoxAsset)
// ---------------------------------------------------------------------------
extern "C" vx_status vx_model_create(vx_ctx* ctx, int32_t sx, int32_t sy, int32_t sz, vx_model** out) {
    VX_REQUIRE(ctx && out, "ctx/out is NULL");
    VX_REQUIRE(sx > 0 && sy > 0 && sz > 0 && sx <= 4096 && sy <= 4096 && sz <= 4096,
               "model dims %dx%dx%d outside 1..4096", sx, sy, sz);
    *out = nullptr;
    if (vx_status st = use_device(ctx)) return st;
    vx_model* m = new (std::nothrow) vx_model();
    if (!m) return fail(VX_ERR_OUT_OF_MEMORY, "host allocation failed");
    m->ctx = ctx;
    // VoxAsset.h:27-29: round each dim up to a multiple of 4.
    m->dims[0] = ((sx - 1) & ~3) + 4;
    m->dims[1] = ((sy - 1) & ~3) + 4;
    m->dims[2] = ((sz - 1) & ~3) + 4;
    for (int k = 0; k < 3; ++k) m->mip[k] = nullptr;
    for (int k = 0; k < 3; ++k) {
        size_t n = (size_t)(m->dims[0] >> k) * (m->dims[1] >> k) * (m->dims[2] >> k);
        hipError_t e = dev_malloc(&m->mip[k], n);
        if (e == hipSuccess) e = hipMemset(m->mip[k], 0, n);
        if (e != hipSuccess) {
            for (int j = 0; j < 3; ++j) (void)hipFree(m->mip[j]);
            delete m;
            return fail(VX_ERR_OUT_OF_MEMORY, "model allocation failed: %s", hipGetErrorString(e));
        }
    }
    *out = m;
    return VX_OK;
}

extern "C" vx_status vx_model_destroy(vx_model* m) {
    if (!m) return VX_OK;
    (void)hipSetDevice(m->ctx->device);
    for (int k = 0; k < 3; ++k) (void)hipFree(m->mip[k]);
    delete m;
    return VX_OK;
}

extern "C" vx_status vx_model_dims(const vx_model* m, int32_t dims[3]) {
    VX_REQUIRE(m && dims, "model/dims is NULL");
    for (int k = 0; k < 3; ++k) dims[k] = m->dims[k];
    return VX_OK;
}

extern "C" vx_status vx_model_mip_ptr(const vx_model* m, int32_t mip, uint8_t** dev_ptr) {
    VX_REQUIRE(m && dev_ptr, "model/dev_ptr is NULL");
    VX_REQUIRE(mip >= 0 && mip < 3, "mip %d outside 0..2", mip);
    *dev_ptr = m->mip[mip];
    return VX_OK;
}

extern "C" vx_status vx_model_upload(vx_model* m, const uint8_t* host, size_t nbytes, void* stream) {
    VX_REQUIRE(m && host, "model/host is NULL");
    size_t n = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
    VX_REQUIRE(nbytes == n, "upload of %zu bytes, model holds %zu", nbytes, n);
    if (vx_status st = use_device(m->ctx)) return st;
    VX_HIP(hipMemcpyAsync(m->mip[0], host, n, hipMemcpyHostToDevice, S(stream)));
    VX_HIP(hipStreamSynchronize(S(stream)));
    return VX_OK;
}

extern "C" vx_status vx_model_download(const vx_model* m, int32_t mip, uint8_t* host, size_t nbytes, void* stream) {
    VX_REQUIRE(m && host, "model/host is NULL");
    VX_REQUIRE(mip >= 0 && mip < 3, "mip %d outside 0..2", mip);
    size_t n = (size_t)(m->dims[0] >> mip) * (m->dims[1] >> mip) * (m->dims[2] >> mip);
    VX_REQUIRE(nbytes == n, "download buffer holds %zu bytes, mip %d has %zu", nbytes, mip, n);
    if (vx_status st = use_device(m->ctx)) return st;
    VX_HIP(hipMemcpyAsync(host, m->mip[mip], n, hipMemcpyDeviceToHost, S(stream)));
    VX_HIP(hipStreamSynchronize(S(stream)));
    return VX_OK;
}

extern "C" vx_status vx_model_build_mips(vx_model* m, void* stream) {
    VX_REQUIRE(m, "model is NULL");
    if (vx_status st = use_device(m->ctx)) return st;
    for (int k = 1; k < 3; ++k) {
        int32_t px = m->dims[0] >> (k - 1), py = m->dims[1] >> (k - 1), pz = m->dims[2] >> (k - 1);
        size_t n = (size_t)(px / 2) * (py / 2) * (pz / 2);
        hipLaunchKernelGGL(k_build_mip, dim3(cdiv(n, 256)), dim3(256), 0, S(stream), m->mip[k - 1], px, py, pz,
                           m->mip[k]);
        VX_LAUNCHED("k_build_mip");
    }
    return VX_OK;
}

extern "C" vx_status vx_model_generate(vx_model* m, const vx_scene_desc* d, void* stream) {
    VX_REQUIRE(m && d, "model/desc is NULL");
    VX_REQUIRE(d->nboxes >= 0 && (d->nboxes == 0 || (d->box_min >= 1 && d->box_max >= d->box_min)),
               "bad box parameters");
    if (vx_status st = use_device(m->ctx)) return st;
    const int32_t nx = m->dims[0], ny = m->dims[1], nz = m->dims[2];
    float* h = nullptr;
    VX_HIP(dev_malloc(&h, (size_t)nx * nz * sizeof(float)));
    hipLaunchKernelGGL(k_terrain_heights, dim3(cdiv((size_t)nx * nz, 256)), dim3(256), 0, S(stream), h, nx, ny, nz,
                       d->seed, d->terrain_fill);
    size_t nq = (size_t)nx * ny * nz / 4;
    hipLaunchKernelGGL(k_terrain_fill, dim3(cdiv(nq, 256)), dim3(256), 0, S(stream), m->mip[0], h, nx, ny, nz,
                       d->seed, d->terrain_fill > 0.0f ? 1 : 0);
    for (int32_t b0 = 0; b0 < d->nboxes; b0 += 65535) {
        int32_t nb = std::min(65535, d->nboxes - b0);
        uint32_t maxvox = (uint32_t)d->box_max * d->box_max * d->box_max;
        hipLaunchKernelGGL(k_boxes, dim3(std::min<uint32_t>(cdiv(maxvox, 256), 64), nb), dim3(256), 0, S(stream),
                           m->mip[0], nx, ny, nz, d->seed, b0, d->box_min, d->box_max, d->boxes_grounded);
    }
    // the height map is released on every path once the kernels are done with it
    hipError_t e = launch_status();
    const hipError_t se = hipStreamSynchronize(S(stream));
    (void)hipFree(h);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "launch of scene generation failed: %s", hipGetErrorString(e));
    VX_HIP(se);
    return VX_OK;
}

// ---------------------------------------------------------------------------
// occupancy volume (ShadowVoxSystem)
// ---------------------------------------------------------------------------
extern "C" vx_status vx_volume_create(vx_ctx* ctx, int32_t sx, int32_t sy, int32_t sz, int32_t nlevels,
                                      vx_volume** out) {
    VX_REQUIRE(ctx && out, "ctx/out is NULL");
    VX_REQUIRE(sx > 0 && sy > 0 && sz > 0 && sx <= 4096 && sy <= 4096 && sz <= 4096,
               "volume texel dims %dx%dx%d outside 1..4096", sx, sy, sz);
    VX_REQUIRE(nlevels >= 0 && nlevels <= 3, "nlevels %d outside 0..3", nlevels);
    *out = nullptr;
    if (vx_status st = use_device(ctx)) return st;
    vx_volume* v = new (std::nothrow) vx_volume();
    if (!v) return fail(VX_ERR_OUT_OF_MEMORY, "host allocation failed");
    v->ctx = ctx;
    v->sx = sx; v->sy = sy; v->sz = sz;
    v->nlevels = nlevels;
    // One extra brick per axis keeps texel index s (fine 2s, 2s+1) inside the grid.
    v->bx = sx / 4 + 1; v->by = sy / 4 + 1; v->bz = sz / 4 + 1;
    v->ncells = (size_t)v->bx * v->by * v->bz * 8;
    v->cells = nullptr;
    for (int l = 0; l < 3; ++l) {
        v->lvl[l] = nullptr;
        int32_t shift = 3 * l;   // level l+1 cell = brick >> 3l
        v->cx[l] = (int32_t)cdiv(v->bx, 1u << shift);
        v->cy[l] = (int32_t)cdiv(v->by, 1u << shift);
        v->cz[l] = (int32_t)cdiv(v->bz, 1u << shift);
        v->wx[l] = (int32_t)cdiv(v->cx[l], 4);
        v->wy[l] = (int32_t)cdiv(v->cy[l], 4);
        v->wz[l] = (int32_t)cdiv(v->cz[l], 4);
    }
    // March grid: fine voxels -1 .. 2s+2 per axis, cell = (f >> 2) + 1.
    v->gx = ((2 * sx + 2) >> 2) + 2;
    v->gy = ((2 * sy + 2) >> 2) + 2;
    v->gz = ((2 * sz + 2) >> 2) + 2;
    v->ngrid = (size_t)v->gx * v->gy * v->gz;
    v->grid = nullptr;
    // The ray-parallel trace addresses the march grid with 32-bit word indices.
    if (v->ngrid >= (size_t(1) << 32)) {
        const size_t words = v->ngrid;
        delete v;
        return fail(VX_ERR_RANGE,
                    "volume %dx%dx%d needs a march grid of %zu words; the trace indexes it with 32-bit word "
                    "indices (at most 2^32 - 1 words: with 4096 x 4096 texel slices, at most 2040 slices)",
                    sx, sy, sz, words);
    }
    hipError_t e = dev_malloc(&v->cells, v->ncells * 8);
    if (e == hipSuccess) e = hipMemset(v->cells, 0, v->ncells * 8);
    for (int l = 0; l < nlevels && e == hipSuccess; ++l) {
        size_t nw = (size_t)v->wx[l] * v->wy[l] * v->wz[l];
        e = dev_malloc(&v->lvl[l], nw * 8);
        if (e == hipSuccess) e = hipMemset(v->lvl[l], 0, nw * 8);
    }
    if (e == hipSuccess) e = dev_malloc(&v->grid, v->ngrid * 8);
    if (e == hipSuccess) e = mgrid_full(v, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        (void)hipFree(v->cells);
        (void)hipFree(v->grid);
        for (int l = 0; l < 3; ++l) (void)hipFree(v->lvl[l]);
        delete v;
        return fail(VX_ERR_OUT_OF_MEMORY, "volume allocation failed: %s", hipGetErrorString(e));
    }
    *out = v;
    return VX_OK;
}

extern "C" vx_status vx_volume_destroy(vx_volume* v) {
    if (!v) return VX_OK;
    (void)hipSetDevice(v->ctx->device);
    (void)hipFree(v->cells);
    (void)hipFree(v->grid);
    for (int l = 0; l < 3; ++l) (void)hipFree(v->lvl[l]);
    delete v;
    return VX_OK;
}

extern "C" vx_status vx_volume_dims(const vx_volume* v, int32_t dims[3], int32_t* nlevels) {
    VX_REQUIRE(v && dims, "volume/dims is NULL");
    dims[0] = v->sx; dims[1] = v->sy; dims[2] = v->sz;
    if (nlevels) *nlevels = v->nlevels;
    return VX_OK;
}

extern "C" vx_status vx_volume_clear(vx_volume* v, void* stream) {
    VX_REQUIRE(v, "volume is NULL");
    if (vx_status st = use_device(v->ctx)) return st;
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(v->ncells, 256)), dim3(256), 0, S(stream), v->cells, v->ncells, 0ull);
    for (int l = 0; l < v->nlevels; ++l) {
        size_t nw = (size_t)v->wx[l] * v->wy[l] * v->wz[l];
        hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(nw, 256)), dim3(256), 0, S(stream), v->lvl[l], nw, 0ull);
    }
    VX_LAUNCHED("k_fill_u64");
    if (hipError_t e = mgrid_full(v, S(stream))) return fail(VX_ERR_DEVICE, "march grid: %s", hipGetErrorString(e));
    return VX_OK;
}

static vx_status check_region(const vx_volume* v, const vx_region& r) {
    VX_REQUIRE(r.x >= 0 && r.y >= 0 && r.z >= 0 && r.w > 0 && r.h > 0 && r.d > 0 &&
                   (int64_t)r.x + r.w <= v->sx && (int64_t)r.y + r.h <= v->sy && (int64_t)r.z + r.d <= v->sz,
               "region (%d,%d,%d)+(%u,%u,%u) outside the %dx%dx%d volume", r.x, r.y, r.z, r.w, r.h, r.d, v->sx,
               v->sy, v->sz);
    return VX_OK;
}

extern "C" vx_status vx_volume_upload_region(vx_volume* v, const uint8_t* host, const vx_region* regions,
                                             uint32_t nregions, void* stream) {
    VX_REQUIRE(v && host, "volume/host is NULL");
    VX_REQUIRE(nregions == 0 || regions, "regions is NULL");
    for (uint32_t i = 0; i < nregions; ++i)
        if (vx_status st = check_region(v, regions[i])) return st;
    if (nregions == 0) return VX_OK;
    if (vx_status st = use_device(v->ctx)) return st;
    size_t nb = (size_t)v->sx * v->sy * v->sz;
    uint8_t* stage = nullptr;
    VX_HIP(dev_malloc(&stage, nb));
    hipError_t e = hipMemcpyAsync(stage, host, nb, hipMemcpyHostToDevice, S(stream));
    OccView V = occ_view(v);
    for (uint32_t i = 0; i < nregions && e == hipSuccess; ++i) {
        const vx_region& r = regions[i];
        size_t n = (size_t)r.w * r.h * r.d;
        hipLaunchKernelGGL(k_bytes_to_cells, dim3(cdiv(n, 256)), dim3(256), 0, S(stream), v->cells, stage, V, r.x,
                           r.y, r.z, (int32_t)r.w, (int32_t)r.h, (int32_t)r.d);
        e = hipGetLastError();
        if (e == hipSuccess) {
            const int32_t t0[3] = {r.x, r.y, r.z};
            const int32_t t1[3] = {r.x + (int32_t)r.w - 1, r.y + (int32_t)r.h - 1, r.z + (int32_t)r.d - 1};
            e = mgrid_update(v, t0, t1, S(stream));
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "region upload failed: %s", hipGetErrorString(e));
    return VX_OK;
}

extern "C" vx_status vx_volume_download(const vx_volume* v, uint8_t* host, size_t nbytes, void* stream) {
    VX_REQUIRE(v && host, "volume/host is NULL");
    size_t nb = (size_t)v->sx * v->sy * v->sz;
    VX_REQUIRE(nbytes == nb, "download buffer holds %zu bytes, volume has %zu", nbytes, nb);
    if (vx_status st = use_device(v->ctx)) return st;
    uint8_t* stage = nullptr;
    VX_HIP(dev_malloc(&stage, nb));
    hipLaunchKernelGGL(k_cells_to_bytes, dim3(cdiv(nb, 256)), dim3(256), 0, S(stream), stage, v->cells, occ_view(v));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, stage, nb, hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "volume download failed: %s", hipGetErrorString(e));
    return VX_OK;
}

// glm::translate(m, v) (Vendor/glm/ext/matrix_transform.inl:10-15):
// Result[3] = ((m[0]*v[0] + m[1]*v[1]) + m[2]*v[2]) + m[3].
static void translate(const float m[16], const float v[3], float out[16]) {
    std::memcpy(out, m, 16 * sizeof(float));
    for (int r = 0; r < 4; ++r) {
        float a = m[0 * 4 + r] * v[0];
        float b = m[1 * 4 + r] * v[1];
        float c = m[2 * 4 + r] * v[2];
        out[3 * 4 + r] = ((a + b) + c) + m[3 * 4 + r];
    }
}

struct PlannedOp {
    FootprintOp op;
    int64_t lo[3], hi[3];   // conservative fine-voxel footprint box
};

static PlannedOp make_op(const vx_instance& in, const float m[16], int32_t value, int32_t slot) {
    PlannedOp p;
    FootprintOp& o = p.op;
    o.model = in.model;
    o.sx = in.sx; o.sy = in.sy; o.sz = in.sz;
    o.value = value;
    o.inst = slot;
    for (int k = 0; k < 3; ++k) {
        o.o[k] = m[12 + k];
        o.dx[k] = m[0 + k] * 0.1f;
        o.dy[k] = m[4 + k] * 0.1f;
        o.dz[k] = m[8 + k] * 0.1f;
    }
    // Corners of the voxel index box mapped exactly, widened by two fine
    // voxels for fp32 rounding and truncation.
    for (int k = 0; k < 3; ++k) {
        double mn = 1e300, mx = -1e300;
        for (int c = 0; c < 8; ++c) {
            double x = (c & 1) ? in.sx - 1 : 0, y = (c & 2) ? in.sy - 1 : 0, z = (c & 4) ? in.sz - 1 : 0;
            double w = ((double)o.o[k] + x * o.dx[k] + y * o.dy[k] + z * o.dz[k]) * 10.0;
            mn = std::min(mn, w);
            mx = std::max(mx, w);
        }
        double span = std::fabs((double)o.dx[k]) * in.sx + std::fabs((double)o.dy[k]) * in.sy +
                      std::fabs((double)o.dz[k]) * in.sz + std::fabs((double)o.o[k]);
        double slack = 2.0 + span * 10.0 * 1e-6;
        p.lo[k] = (int64_t)std::floor(std::max(mn - slack, -9.0e18));
        p.hi[k] = (int64_t)std::ceil(std::min(mx + slack, 9.0e18));
    }
    return p;
}

static bool disjoint(const PlannedOp& a, const PlannedOp& b) {
    for (int k = 0; k < 3; ++k)
        if (a.hi[k] < b.lo[k] || b.hi[k] < a.lo[k]) return true;
    return false;
}

static vx_status vx_volume_rebuild_impl(vx_volume* v, const vx_instance* inst, uint32_t n, int32_t thr,
                                       vx_region* dirty, uint32_t* ndirty, void* stream) {
    VX_REQUIRE(v, "volume is NULL");
    VX_REQUIRE(n == 0 || inst, "instances is NULL");
    VX_REQUIRE(dirty || n == 0, "dirty is NULL");
    VX_REQUIRE(ndirty, "ndirty is NULL");
    VX_REQUIRE(thr >= 0 && thr <= 256, "occupancy threshold %d outside 0..256", thr);
    *ndirty = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vx_instance& in = inst[i];
        VX_REQUIRE(in.flags == VX_INST_UPDATE || in.flags == VX_INST_DESTROY, "instance %u: bad flags %d", i,
                   in.flags);
        if (in.model)
            VX_REQUIRE(in.sx > 0 && in.sy > 0 && in.sz > 0 && (int64_t)in.sx * in.sy * in.sz < (1ll << 40),
                       "instance %u: bad model dims", i);
    }
    if (n == 0) return VX_OK;
    if (vx_status st = use_device(v->ctx)) return st;

    // Operation list in reference order, then greedy batching: an operation
    // joins the latest compatible batch it can be hoisted into. It may move
    // past a batch of the other value only when its footprint box is disjoint
    // from every operation in that batch; same-value operations commute.
    std::vector<PlannedOp> ops;
    for (uint32_t i = 0; i < n; ++i) {
        const vx_instance& in = inst[i];
        if (!in.model) continue;   // ShadowVoxSystem.cpp:120 (invalid Vox asset)
        if (in.flags == VX_INST_DESTROY) {
            ops.push_back(make_op(in, in.cur, 0, (int32_t)i));
        } else {
            float np[3] = {-in.pivot[0], -in.pivot[1], -in.pivot[2]};
            float m[16];
            translate(in.prev, np, m);
            ops.push_back(make_op(in, m, 0, (int32_t)i));
            translate(in.cur, np, m);
            ops.push_back(make_op(in, m, 1, (int32_t)i));
        }
    }
    std::vector<std::vector<int>> batches;
    std::vector<int32_t> batch_value;
    for (int j = 0; j < (int)ops.size(); ++j) {
        int target = -1;
        for (int b = (int)batches.size() - 1; b >= 0; --b) {
            if (batch_value[b] == ops[j].op.value) { target = b; continue; }
            bool ok = true;
            for (int q : batches[b])
                if (!disjoint(ops[q], ops[j])) { ok = false; break; }
            if (!ok) break;
        }
        if (target < 0) {
            batches.push_back({j});
            batch_value.push_back(ops[j].op.value);
        } else {
            batches[target].push_back(j);
        }
    }

    // Per-instance AABB slots start at (startmin, startmax) (ShadowVoxSystem.cpp:128-132).
    std::vector<int32_t> aabb_h((size_t)n * 6);
    for (uint32_t i = 0; i < n; ++i) {
        aabb_h[6 * i + 0] = v->sx - 1; aabb_h[6 * i + 1] = v->sy - 1; aabb_h[6 * i + 2] = v->sz - 1;
        aabb_h[6 * i + 3] = 0; aabb_h[6 * i + 4] = 0; aabb_h[6 * i + 5] = 0;
    }
    // per-context scratch, grown on demand (the rebuild is synchronous)
    vx_ctx* cx = v->ctx;
    if (cx->rb_aabb_cap < aabb_h.size() || cx->rb_ops_cap < ops.size()) {
        VX_HIP(hipDeviceSynchronize());
        if (cx->rb_aabb_cap < aabb_h.size()) {
            (void)hipFree(cx->rb_aabb);
            cx->rb_aabb = nullptr;
            cx->rb_aabb_cap = 0;
            const size_t cap = std::max<size_t>(aabb_h.size(), 6 * 256);
            VX_HIP(dev_malloc(&cx->rb_aabb, cap * sizeof(int32_t)));
            cx->rb_aabb_cap = cap;
        }
        if (cx->rb_ops_cap < ops.size()) {
            (void)hipFree(cx->rb_ops);
            cx->rb_ops = nullptr;
            cx->rb_ops_cap = 0;
            const size_t cap = std::max<size_t>(ops.size(), 512);
            VX_HIP(dev_malloc(&cx->rb_ops, cap * sizeof(FootprintOp)));
            cx->rb_ops_cap = cap;
        }
    }
    int32_t* aabb_d = cx->rb_aabb;
    FootprintOp* ops_d = cx->rb_ops;
    hipError_t e = hipSuccess;
    std::vector<FootprintOp> flat;
    std::vector<size_t> first;
    for (auto& b : batches) {
        first.push_back(flat.size());
        for (int q : b) flat.push_back(ops[q].op);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(aabb_d, aabb_h.data(), aabb_h.size() * sizeof(int32_t), hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess)
        e = hipMemcpyAsync(ops_d, flat.data(), flat.size() * sizeof(FootprintOp), hipMemcpyHostToDevice, S(stream));
    OccView V = occ_view(v);
    for (size_t b = 0; b < batches.size() && e == hipSuccess; ++b) {
        size_t maxq = 0;   // quads of four x voxels
        for (int q : batches[b])
            maxq = std::max(maxq, (size_t)((ops[q].op.sx + 3) / 4) * ops[q].op.sy * ops[q].op.sz);
        uint32_t gx = std::min<uint32_t>(cdiv(maxq, 256), 4096);
        for (size_t c0 = 0; c0 < batches[b].size() && e == hipSuccess; c0 += 65535) {
            uint32_t cnt = (uint32_t)std::min<size_t>(65535, batches[b].size() - c0);
            hipLaunchKernelGGL(k_footprint, dim3(gx, cnt), dim3(256), 0, S(stream), v->cells, V,
                               ops_d + first[b] + c0, aabb_d, thr);
            e = launch_status();
        }
    }
    // March grid over the conservative footprint boxes (fine voxels -> texels):
    // one refresh over their union when it is no larger than the boxes together
    // (an instance's erase and write boxes usually overlap) or the list is
    // long, else one refresh per operation.
    if (e == hipSuccess) {
        auto texel_box = [&](const PlannedOp& p, int32_t t0[3], int32_t t1[3]) {
            for (int k = 0; k < 3; ++k) {
                int64_t lo = std::max<int64_t>(p.lo[k], -8), hi = std::min<int64_t>(p.hi[k], 1ll << 20);
                t0[k] = (int32_t)(lo >> 1) - 1;
                t1[k] = (int32_t)(hi >> 1) + 1;
            }
        };
        // texels of a box once clipped to the volume (the refresh clips the same way)
        auto clipped = [&](const int32_t t0[3], const int32_t t1[3]) {
            const int32_t dims[3] = {v->sx, v->sy, v->sz};
            double n = 1.0;
            for (int k = 0; k < 3; ++k)
                n *= (double)std::max<int64_t>(0, (int64_t)std::min(t1[k], dims[k] - 1) - std::max(t0[k], 0) + 1);
            return n;
        };
        int32_t u0[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, u1[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
        double sum = 0.0;
        for (const PlannedOp& p : ops) {
            int32_t t0[3], t1[3];
            texel_box(p, t0, t1);
            sum += clipped(t0, t1);
            for (int k = 0; k < 3; ++k) { u0[k] = std::min(u0[k], t0[k]); u1[k] = std::max(u1[k], t1[k]); }
        }
        if (ops.empty()) {
            // nothing was written
        } else if (ops.size() > 64 || clipped(u0, u1) <= sum) {
            e = mgrid_update(v, u0, u1, S(stream));
        } else {
            for (size_t j = 0; j < ops.size() && e == hipSuccess; ++j) {
                int32_t t0[3], t1[3];
                texel_box(ops[j], t0, t1);
                e = mgrid_update(v, t0, t1, S(stream));
            }
        }
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(aabb_h.data(), aabb_d, aabb_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "rebuild failed: %s", hipGetErrorString(e));

    // Dirty regions, ShadowVoxSystem.cpp:180-190.
    uint32_t nd = 0;
    const int32_t smin[3] = {v->sx - 1, v->sy - 1, v->sz - 1};
    for (uint32_t i = 0; i < n; ++i) {
        if (!inst[i].model) continue;
        const int32_t* a = &aabb_h[6 * i];
        if (a[3] == 0 && a[4] == 0 && a[5] == 0) continue;
        int32_t mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::max(a[k] / 2, 0);
            mx[k] = std::min(a[3 + k] / 2, smin[k]);
        }
        vx_region& r = dirty[nd++];
        r.x = mn[0]; r.y = mn[1]; r.z = mn[2];
        r.w = (uint32_t)(mx[0] - mn[0] + 1);
        r.h = (uint32_t)(mx[1] - mn[1] + 1);
        r.d = (uint32_t)(mx[2] - mn[2] + 1);
    }
    *ndirty = nd;
    return VX_OK;
}

static vx_status vx_volume_build_levels_impl(vx_volume* v, const vx_region* regions, uint32_t nregions, void* stream) {
    VX_REQUIRE(v, "volume is NULL");
    VX_REQUIRE(nregions == 0 || regions, "regions is NULL");
    if (v->nlevels == 0) return VX_OK;
    if (vx_status st = use_device(v->ctx)) return st;
    // Brick ranges to refresh ([lo, hi] inclusive), whole grid when no regions.
    std::vector<std::array<int32_t, 6>> ranges;
    if (nregions == 0) {
        ranges.push_back({0, 0, 0, v->bx - 1, v->by - 1, v->bz - 1});
    } else {
        for (uint32_t i = 0; i < nregions; ++i) {
            const vx_region& r = regions[i];
            if (r.w == 0 || r.h == 0 || r.d == 0) continue;
            int64_t lo[3] = {r.x, r.y, r.z};
            int64_t hi[3] = {(int64_t)r.x + r.w - 1, (int64_t)r.y + r.h - 1, (int64_t)r.z + r.d - 1};
            const int32_t tdim[3] = {v->sx, v->sy, v->sz};
            bool empty = false;
            std::array<int32_t, 6> br;
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::max<int64_t>(lo[k], 0);
                hi[k] = std::min<int64_t>(hi[k], tdim[k] - 1);
                if (hi[k] < lo[k]) empty = true;
                br[k] = (int32_t)(lo[k] >> 2);
                br[3 + k] = (int32_t)(hi[k] >> 2);
            }
            if (!empty) ranges.push_back(br);
        }
    }
    OccView V = occ_view(v);
    for (auto& br : ranges) {
        for (int l = 0; l < v->nlevels; ++l) {
            // level l+1 words covering the brick range: word = brick >> (3l + 2)
            int32_t sh = 3 * l + 2;
            int32_t w0[3], w1[3];
            for (int k = 0; k < 3; ++k) { w0[k] = br[k] >> sh; w1[k] = br[3 + k] >> sh; }
            int32_t nw[3] = {w1[0] - w0[0] + 1, w1[1] - w0[1] + 1, w1[2] - w0[2] + 1};
            size_t nthreads = (size_t)nw[0] * nw[1] * nw[2] * 64;
            if (l == 0) {
                hipLaunchKernelGGL(k_level1, dim3(cdiv(nthreads, 256)), dim3(256), 0, S(stream), v->lvl[0], v->cells,
                                   V, w0[0], w0[1], w0[2], nw[0], nw[1], nw[2]);
            } else {
                hipLaunchKernelGGL(k_level_up, dim3(cdiv(nthreads, 256)), dim3(256), 0, S(stream), v->lvl[l],
                                   v->lvl[l - 1], v->wx[l - 1], v->wy[l - 1], v->wz[l - 1], v->wx[l], v->wy[l],
                                   v->wz[l], w0[0], w0[1], w0[2], nw[0], nw[1], nw[2]);
            }
            VX_LAUNCHED("level build");
        }
    }
    return VX_OK;
}

extern "C" vx_status vx_volume_download_level(const vx_volume* v, int32_t level, uint64_t* host, size_t nwords,
                                              size_t* words_total, void* stream) {
    VX_REQUIRE(v, "volume is NULL");
    VX_REQUIRE(level >= 1 && level <= v->nlevels, "level %d outside 1..%d", level, v->nlevels);
    size_t nw = (size_t)v->wx[level - 1] * v->wy[level - 1] * v->wz[level - 1];
    if (words_total) *words_total = nw;
    if (!host) return VX_OK;
    VX_REQUIRE(nwords >= nw, "buffer holds %zu words, level has %zu", nwords, nw);
    if (vx_status st = use_device(v->ctx)) return st;
    VX_HIP(hipMemcpyAsync(host, v->lvl[level - 1], nw * 8, hipMemcpyDeviceToHost, S(stream)));
    VX_HIP(hipStreamSynchronize(S(stream)));
    return VX_OK;
}

// ---------------------------------------------------------------------------
// passes
// ---------------------------------------------------------------------------
// a * b in fp32, column j = ((a0*bj.x + a1*bj.y) + a2*bj.z) + a3*bj.w.
static void mat_mul_f32(const float* a, const float* b, float* out) {
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            out[j * 4 + r] = ((a[0 + r] * b[j * 4 + 0] + a[4 + r] * b[j * 4 + 1]) + a[8 + r] * b[j * 4 + 2]) +
                             a[12 + r] * b[j * 4 + 3];
}

extern "C" vx_status vx_primary(vx_ctx* ctx, const vx_model* model, const vx_view* view, const vx_gbuf* g,
                                void* stream) {
    VX_REQUIRE(ctx && model && g, "ctx/model/gbuf is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->pos && g->normal && g->material, "gbuf has a NULL pointer");
    if (vx_status st = use_device(ctx)) return st;
    PrimaryArgs A;
    A.view = *view;
    A.tiles = tile_map(view->width, view->height);
    mat_mul_f32(view->proj, view->view, A.pv);
    for (int k = 0; k < 3; ++k) { A.mips[k] = model->mip[k]; A.dims[k] = model->dims[k]; }
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    hipLaunchKernelGGL(k_primary, dim3(8 * A.tiles.per_xcd), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_primary");
    return VX_OK;
}

// ---- multi-instance primary: host-side per-draw constants --------------------
// 4x4 inverse (column-major) in float64 by cofactors and the determinant.
static bool inverse4(const double* m, double* inv) {
    double c[16];
    c[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    c[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    c[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    c[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    c[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    c[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    c[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    c[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    c[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    c[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    c[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    c[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    c[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    c[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    c[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    c[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) inv[i] = c[i] / det;
    return true;
}

// Conservative pixel rectangle of a draw's box (x0 > x1: nothing visible).
static void inst_rect(const vx_view& v, const float* world, const int32_t dims[3], int32_t r[4]) {
    double pv[16];
    for (int j = 0; j < 4; ++j)
        for (int q = 0; q < 4; ++q) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += (double)v.proj[k * 4 + q] * (double)v.view[j * 4 + k];
            pv[j * 4 + q] = acc;
        }
    bool near_or_behind = false, all_behind = true;
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (int c = 0; c < 8; ++c) {
        const double l[3] = {(c & 1) ? dims[0] * 0.1 : 0.0, (c & 2) ? dims[1] * 0.1 : 0.0,
                             (c & 4) ? dims[2] * 0.1 : 0.0};
        double w[4], cl[4];
        for (int q = 0; q < 4; ++q)
            w[q] = (double)world[q] * l[0] + (double)world[4 + q] * l[1] + (double)world[8 + q] * l[2] +
                   (double)world[12 + q];
        for (int q = 0; q < 4; ++q) cl[q] = pv[q] * w[0] + pv[4 + q] * w[1] + pv[8 + q] * w[2] + pv[12 + q] * w[3];
        if (cl[3] > -1e-3) all_behind = false;
        if (cl[3] < 1e-3) {
            near_or_behind = true;
            continue;
        }
        const double px = (cl[0] / cl[3] + 1.0) * 0.5 * v.res[0] - 0.5 - v.jitter[0];
        const double py = (1.0 - cl[1] / cl[3]) * 0.5 * v.res[1] - 0.5 - v.jitter[1];
        x0 = std::min(x0, px); x1 = std::max(x1, px);
        y0 = std::min(y0, py); y1 = std::max(y1, py);
    }
    if (all_behind) {
        r[0] = 1; r[1] = 1; r[2] = 0; r[3] = 0;
        return;
    }
    if (near_or_behind) {
        r[0] = 0; r[1] = 0; r[2] = v.width - 1; r[3] = v.height - 1;
        return;
    }
    const double lim = 1e9;
    x0 = std::max(std::floor(x0) - 2.0, -lim); y0 = std::max(std::floor(y0) - 2.0, -lim);
    x1 = std::min(std::ceil(x1) + 2.0, lim); y1 = std::min(std::ceil(y1) + 2.0, lim);
    r[0] = (int32_t)std::max(x0, 0.0); r[1] = (int32_t)std::max(y0, 0.0);
    r[2] = (int32_t)std::min(x1, (double)(v.width - 1)); r[3] = (int32_t)std::min(y1, (double)(v.height - 1));
}

static vx_status vx_primary_instances_impl(vx_ctx* ctx, const vx_vox_cmd* cmds, uint32_t n, const vx_view* view,
                                          const vx_gbuf* g, const vx_gbuf_ext* ext, void* stream) {
    VX_REQUIRE(ctx && g, "ctx/gbuf is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->pos && g->normal && g->material, "gbuf has a NULL pointer");
    if (n > 4096) return fail(VX_ERR_RANGE, "%u voxel draws, at most 4096", n);
    VX_REQUIRE(n == 0 || cmds, "cmds is NULL");
    std::vector<InstConst> h(n);
    std::vector<int32_t> rects(4 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) {
        const vx_vox_cmd& c = cmds[k];
        VX_REQUIRE(c.model, "draw %u: model is NULL", k);
        VX_REQUIRE(c.palette >= 0 && c.palette <= 65535, "draw %u: palette %d outside 0..65535", k, c.palette);
        InstConst& I = h[k];
        std::memset(&I, 0, sizeof(I));
        double wd[16], wi[16];
        for (int q = 0; q < 16; ++q) {
            I.w[q] = c.world[q];
            I.lw[q] = c.last_world[q];
            wd[q] = c.world[q];
        }
        if (!inverse4(wd, wi)) return fail(VX_ERR_INVALID_ARGUMENT, "draw %u: world matrix is singular", k);
        for (int q = 0; q < 16; ++q) I.winv[q] = (float)wi[q];
        const double cx = view->cam_pos[0], cy = view->cam_pos[1], cz = view->cam_pos[2];
        for (int q = 0; q < 3; ++q)
            I.lc[q] = (float)((((wi[q] * cx + wi[4 + q] * cy) + wi[8 + q] * cz) + wi[12 + q]) * 10.0);
        I.rid = c.volume_rid;
        I.palette = c.palette;
        for (int q = 0; q < 3; ++q) {
            I.m.mips[q] = c.model->mip[q];
            I.m.dims[q] = c.model->dims[q];
        }
        inst_rect(*view, c.world, c.model->dims, &rects[4 * (size_t)k]);
    }
    if (vx_status st = use_device(ctx)) return st;
    const TileMap tiles = tile_map(view->width, view->height);
    if (ctx->inst_cap < std::max(n, 1u) || ctx->tile_cap < tiles.ntiles) {
        VX_HIP(hipDeviceSynchronize());
        if (ctx->inst_cap < std::max(n, 1u)) {
            (void)hipFree(ctx->inst);
            (void)hipFree(ctx->inst_rects);
            ctx->inst = nullptr;
            ctx->inst_rects = nullptr;
            ctx->inst_cap = 0;
            const uint32_t cap = std::max(n, 64u);
            VX_HIP(dev_malloc(&ctx->inst, cap * sizeof(InstConst)));
            VX_HIP(dev_malloc(&ctx->inst_rects, cap * 4 * sizeof(int32_t)));
            ctx->inst_cap = cap;
        }
        if (ctx->tile_cap < tiles.ntiles) {
            (void)hipFree(ctx->tile_lists);
            (void)hipFree(ctx->tile_counts);
            ctx->tile_lists = nullptr;
            ctx->tile_counts = nullptr;
            ctx->tile_cap = 0;
            VX_HIP(dev_malloc(&ctx->tile_lists, (size_t)tiles.ntiles * kTileInstCap * sizeof(uint16_t)));
            VX_HIP(dev_malloc(&ctx->tile_counts, (size_t)tiles.ntiles * sizeof(uint32_t)));
            ctx->tile_cap = tiles.ntiles;
        }
    }
    if (n) {
        VX_HIP(hipMemcpyAsync(ctx->inst, h.data(), n * sizeof(InstConst), hipMemcpyHostToDevice, S(stream)));
        VX_HIP(hipMemcpyAsync(ctx->inst_rects, rects.data(), rects.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              S(stream)));
    }
    BinArgs B;
    B.rects = ctx->inst_rects;
    B.ninst = n;
    B.tiles_x = tiles.tiles_x;
    B.ntiles = tiles.ntiles;
    B.lists = ctx->tile_lists;
    B.counts = ctx->tile_counts;
    hipLaunchKernelGGL(k_bin_instances, dim3(cdiv(tiles.ntiles, 256)), dim3(256), 0, S(stream), B);
    VX_LAUNCHED("k_bin_instances");
    PrimaryInstArgs A;
    A.view = *view;
    A.tiles = tiles;
    mat_mul_f32(view->proj, view->view, A.pv);
    mat_mul_f32(view->proj, view->last_view, A.plv);
    A.inst = ctx->inst;
    A.ninst = n;
    A.lists = ctx->tile_lists;
    A.counts = ctx->tile_counts;
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    A.motion = ext ? ext->motion : nullptr;
    A.depth = ext ? ext->depth : nullptr;
    A.palette = ext ? ext->palette : nullptr;
    hipLaunchKernelGGL(k_primary_inst, dim3(8 * tiles.per_xcd), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_primary_inst");
    return VX_OK;
}

extern "C" vx_status vx_light_taa(vx_ctx* ctx, const vx_view* view, const vx_taa_io* io, uint32_t seed,
                                  void* stream) {
    VX_REQUIRE(ctx && io, "ctx/io is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(io->light && io->last && io->depth && io->normal && io->material && io->palette && io->out,
               "io has a NULL pointer (only motion may be NULL)");
    VX_REQUIRE((const void*)io->out != (const void*)io->last, "out must not alias last");
    VX_REQUIRE(((uintptr_t)io->last & 15u) == 0, "last must be 16-byte aligned");
    if (vx_status st = use_device(ctx)) return st;
    TaaArgs A;
    A.width = view->width;
    A.height = view->height;
    A.frame = view->frame;
    A.seed = seed;
    A.light = io->light; A.last = io->last; A.motion = io->motion; A.depth = io->depth;
    A.normal = io->normal; A.mat = io->material; A.palette = io->palette; A.out = io->out;
    const uint32_t npx = (uint32_t)view->width * (uint32_t)view->height;
    if (ctx->taa_cap < npx) {
        VX_HIP(hipDeviceSynchronize());
        (void)hipFree(ctx->taa_rec);
        ctx->taa_rec = nullptr;
        ctx->taa_cap = 0;
        VX_HIP(dev_malloc(&ctx->taa_rec, (size_t)npx * 64));
        ctx->taa_cap = npx;
    }
    A.rec = reinterpret_cast<float4*>(ctx->taa_rec);
    hipLaunchKernelGGL(k_taa_pack, dim3(cdiv(npx, 256)), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_taa_pack");
    hipLaunchKernelGGL(k_light_taa, dim3(cdiv(npx, 256)), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_light_taa");
    return VX_OK;
}

// Persistent grid of a march kernel: CUs x resident blocks per CU (at most
// want_per_cu when that is set), a multiple of 8 so every XCD band has blocks.
// The occupancy query runs once per kernel and context.
static vx_status persistent_grid(vx_ctx* ctx, bool single, int block, int steps, int want_per_cu, uint32_t& grid) {
    uint32_t& per_cu_c = ctx->gcache.grid[single ? 1 : 0][block == 512 ? 1 : 0]
                                                  [steps == 2 ? 0 : (steps == 4 ? 1 : (steps == 8 ? 2 : 3))];
    if (per_cu_c == 0) {
        int per_cu = 0;
        VX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, march_kernel(block, steps, single), block, 0));
        per_cu_c = (uint32_t)std::max(per_cu, 1);
    }
    uint32_t per_cu = per_cu_c;
    if (want_per_cu > 0 && (uint32_t)want_per_cu < per_cu) per_cu = (uint32_t)want_per_cu;
    grid = (uint32_t)ctx->gcache.cus * per_cu;
    grid = (grid + 7u) & ~7u;
    return VX_OK;
}

// Grows the ray-parallel scratch to `slots` ray slots (and fetch counts when
// asked); buffers in use by queued work are released after a device sync.
static vx_status ensure_march(MarchBuffers& m, size_t slots, bool fetches) {
    if (!m.counters) VX_HIP(dev_malloc(&m.counters, kCounterWords * sizeof(uint32_t)));
    const bool grow = m.slots < slots, fgrow = fetches && m.fslots < slots;
    if (!grow && !fgrow) return VX_OK;
    VX_HIP(hipDeviceSynchronize());
    if (grow) {
        (void)hipFree(m.res);
        (void)hipFree(m.deferred);
        m.res = nullptr;
        m.deferred = nullptr;
        m.slots = 0;
        VX_HIP(dev_malloc(&m.res, slots * sizeof(float)));
        VX_HIP(dev_malloc(&m.deferred, slots * sizeof(uint32_t)));
        m.slots = slots;
    }
    if (fgrow) {
        (void)hipFree(m.fres);
        m.fres = nullptr;
        m.fslots = 0;
        VX_HIP(dev_malloc(&m.fres, slots * sizeof(uint16_t)));
        m.fslots = slots;
    }
    return VX_OK;
}

static TileGeom tile_geom(const vx_trace_tuning& tu, int32_t width, int32_t height) {
    TileGeom g;
    g.tw_shift = tu.tile == 1 ? 4u : (tu.tile == 2 ? 2u : 3u);
    g.order = tu.order == 1 ? 1u : 0u;
    g.tiles_x = cdiv((uint32_t)width, 1u << g.tw_shift);
    g.tiles_y = cdiv((uint32_t)height, 64u >> g.tw_shift);
    g.nbx = cdiv(g.tiles_x, 8);
    g.ntiles = g.order == 0 ? g.tiles_x * g.tiles_y : g.nbx * cdiv(g.tiles_y, 8) * 64u;
    return g;
}

static MarchItem make_march_item(const vx_volume* vol, const TraceArgs& T, uint32_t tile0, const TileGeom& g) {
    MarchItem I;
    std::memset(&I, 0, sizeof(I));
    I.V = occ_view(vol);
    I.T = T;
    I.grid = vol->grid;
    I.gx = vol->gx;
    I.gxy = vol->gx * vol->gy;
    I.lim[0] = 2 * vol->sx + 1; I.lim[1] = 2 * vol->sy + 1; I.lim[2] = 2 * vol->sz + 1;
    I.tiles_x = g.tiles_x;
    I.ntiles = g.ntiles;
    I.tile0 = tile0;
    I.tw_shift = g.tw_shift;
    I.order = g.order;
    I.nbx = g.nbx;
    return I;
}

struct MarchShape {
    int block, steps;
    uint32_t grid, rest_grid;
};

static vx_status march_shape(vx_ctx* ctx, bool single, MarchShape& sh) {
    const vx_trace_tuning& tu = ctx->tuning;
    sh.block = tu.block == 512 ? 512 : 256;
    sh.steps = tu.steps ? tu.steps : 16;
    sh.rest_grid = 64;
    return persistent_grid(ctx, single, sh.block, sh.steps, tu.blocks_per_cu, sh.grid);
}

static void march_common(MarchArgs& A, const vx_trace_params& p, const vx_trace_tuning& tu) {
    A.nrays = p.n_shadow + p.n_ao + p.n_spec;
    // chunks (0: set_bands picks) | shadow rays << 8 | spec rays << 16
    A.gen = (uint32_t)std::max(tu.chunks, 0) | ((uint32_t)p.n_shadow << 8) | ((uint32_t)p.n_spec << 16);
    A.refill = (tu.refill >= 1 && tu.refill <= 64) ? tu.refill : 16;

    A.maxt[0] = p.shadow_maxt;
    A.maxt[1] = p.ao_maxt;
    A.maxt[2] = p.spec_maxt;
}

// The tile queue's eight XCD bands over A.ntiles tiles: each band holds a
// contiguous range of tiles, once per chunk of their rays. Chunks (queue
// items per tile) unless tuned: a tile's rays are a few hundred microseconds
// of a wave's work and the launch ends when its last wave finishes its last
// item, so a launch takes chunks of about six rays, and more (down to one ray)
// when that leaves fewer than two items per wave of the grid. The band
// counters stay below 2^31.
static void set_bands(MarchArgs& A, uint32_t waves) {
    A.band_tiles = cdiv(A.ntiles, 8);
    uint64_t want = A.gen & 0xffu;
    if (want == 0) {
        const uint64_t spread = cdiv(2ull * waves, std::max<uint64_t>(A.ntiles, 1));
        want = std::max<uint64_t>((uint64_t)A.nrays / 6, spread);
    }
    const uint64_t cap = (0x7fffffffull) / (8ull * std::max<uint64_t>(A.band_tiles, 1));
    const uint32_t chunks = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(std::min<uint64_t>(want, std::max<int32_t>(A.nrays, 1)), cap));
    A.band = A.band_tiles * chunks;
    A.gen = (A.gen & ~0xffu) | chunks;
}

// Scratch of one single-view march launch: per-slot results, listed rays,
// fetch counts (may be null) and the queue counters.
struct MarchScratch {
    float* res;
    uint32_t* deferred;
    uint16_t* fres;
    uint32_t* counters;
};

// Enqueues the ray-parallel trace of the view's 8x8 tiles [tile_begin,
// tile_begin + ntiles) (counter clear, march, listed rays, resolve); slots
// count from the first of those tiles. e0 / e1 (optional) bracket the march.
static vx_status enqueue_trace_march(vx_ctx* ctx, const vx_volume* vol, const TraceArgs& T, uint32_t tile_begin,
                                     uint32_t ntiles, const MarchScratch& M, hipStream_t stream, hipEvent_t e0,
                                     hipEvent_t e1, vx_trace_hit* hits = nullptr) {
    MarchArgs A;
    std::memset(&A, 0, sizeof(A));
    // tile0 = -tile_begin (mod 2^32): launch tile t is view tile t + tile_begin
    A.item0 = make_march_item(vol, T, 0u - tile_begin, tile_geom(ctx->tuning, T.view.width, T.view.height));
    march_common(A, T.params, ctx->tuning);
    A.items = nullptr;
    A.nitems = 1;
    A.ntiles = ntiles;
    const size_t slots = (size_t)A.ntiles * (size_t)A.nrays * 64;
    A.counters = M.counters;
    A.deferred = M.deferred;
    A.deferred_cap = (uint32_t)slots;
    A.res = M.res;
    A.fres = T.fetches ? M.fres : nullptr;
    A.hits = hits;
    MarchShape sh;
    if (vx_status st = march_shape(ctx, true, sh)) return st;
    set_bands(A, sh.grid * (uint32_t)(sh.block / 64));
    hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(kCounterWords, 256)), dim3(256), 0, stream, A.counters, (int)kCounterWords);
    hipError_t e = launch_march(sh.block, sh.steps, true, sh.grid, sh.rest_grid, stream, A, e0, e1);
    if (e == hipSuccess) e = launch_status();
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "launch of the ray-parallel trace failed: %s", hipGetErrorString(e));
    return VX_OK;
}

static vx_status launch_trace_march(vx_ctx* ctx, const vx_volume* vol, const TraceArgs& T, hipStream_t stream,
                                    vx_trace_hit* hits = nullptr) {
    const uint32_t ntiles = tile_geom(ctx->tuning, T.view.width, T.view.height).ntiles;
    const int32_t nrays = T.params.n_shadow + T.params.n_ao + T.params.n_spec;
    const size_t slots = (size_t)ntiles * (size_t)nrays * 64;
    VX_REQUIRE(slots < (1ull << 32), "view needs %zu ray slots, more than 2^32", slots);
    if (vx_status st = ensure_march(ctx->march, slots, T.fetches != nullptr)) return st;
    MarchScratch M = {ctx->march.res, ctx->march.deferred, ctx->march.fres, ctx->march.counters};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profile) {
        VX_HIP(hipEventCreate(&e0));
        if (hipError_t ce = hipEventCreate(&e1)) {
            (void)hipEventDestroy(e0);
            return fail(VX_ERR_DEVICE, "hipEventCreate failed: %s", hipGetErrorString(ce));
        }
        ctx->prof.emplace_back(e0, e1);
    }
    return enqueue_trace_march(ctx, vol, T, 0, ntiles, M, stream, e0, e1, hits);
}

// vx_trace_noalloc's march: the view's tiles in consecutive ranges whose ray
// slots fit the device's scratch, one after another on the stream (a range
// reuses the scratch once the previous range's resolve has read it).
static vx_status launch_trace_march_noalloc(vx_ctx* ctx, const vx_volume* vol, const TraceArgs& T,
                                            hipStream_t stream) {
    const uint32_t ntiles = tile_geom(ctx->tuning, T.view.width, T.view.height).ntiles;
    const size_t per_tile = (size_t)(T.params.n_shadow + T.params.n_ao + T.params.n_spec) * 64;
    if (per_tile > kNoAllocSlots)
        return fail(VX_ERR_RANGE, "an 8x8 tile needs %zu ray slots; the no-allocation scratch holds %zu", per_tile,
                    kNoAllocSlots);
    const uint32_t chunk = (uint32_t)std::max<size_t>(kNoAllocSlots / std::max<size_t>(per_tile, 1), 1);
    uint8_t* base = ctx->noalloc;
    MarchScratch M;
    M.counters = reinterpret_cast<uint32_t*>(base);
    M.res = reinterpret_cast<float*>(base + kNoAllocCounterBytes);
    M.deferred = reinterpret_cast<uint32_t*>(base + kNoAllocCounterBytes + kNoAllocSlots * sizeof(float));
    M.fres = reinterpret_cast<uint16_t*>(base + kNoAllocCounterBytes + kNoAllocSlots * (sizeof(float) + sizeof(uint32_t)));
    for (uint32_t t0 = 0; t0 < ntiles; t0 += chunk) {
        const uint32_t n = std::min(chunk, ntiles - t0);
        if (vx_status st = enqueue_trace_march(ctx, vol, T, t0, n, M, stream, nullptr, nullptr)) return st;
    }
    return VX_OK;
}

static vx_status check_params(const vx_trace_params* p) {
    VX_REQUIRE(p, "params is NULL");
    auto okm = [](int32_t m) { return m == VX_MARCH_DDA || m == VX_MARCH_SPARSE || m == VX_MARCH_SUPERSPARSE; };
    VX_REQUIRE(okm(p->shadow_march) && okm(p->ao_march) && okm(p->spec_march), "unknown marcher");
    VX_REQUIRE(p->n_shadow >= 0 && p->n_shadow <= 1 && p->n_spec >= 0 && p->n_spec <= 1 && p->n_ao >= 0 &&
                   p->n_ao <= 64,
               "ray counts outside shadow 0..1, ao 0..64, spec 0..1");
    VX_REQUIRE(p->shadow_maxt > 0.0f && p->ao_maxt > 0.0f && p->spec_maxt > 0.0f, "maxt must be positive");
    return VX_OK;
}

static vx_status trace_impl(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                            const float* palette, const vx_trace_params* params, const vx_trace_out* out,
                            void* stream, bool noalloc, vx_trace_hit* hits = nullptr) {
    VX_REQUIRE(ctx && vol && g && out && palette, "ctx/volume/gbuf/out/palette is NULL");
    if (vx_status st = check_view(view)) return st;
    if (vx_status st = check_params(params)) return st;
    VX_REQUIRE(g->pos && g->normal && g->material, "gbuf has a NULL pointer");
    VX_REQUIRE(out->shadow && out->ao && out->spec, "trace output has a NULL pointer");
    if (vx_status st = use_device(ctx)) return st;
    TraceArgs A;
    A.view = *view;
    A.params = *params;
    A.tiles = tile_map(view->width, view->height);
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    A.palette = palette;
    A.shadow = out->shadow; A.ao = out->ao; A.spec = out->spec; A.fetches = out->fetches;
    const bool all_dda = (params->n_shadow == 0 || params->shadow_march == VX_MARCH_DDA) &&
                         (params->n_ao == 0 || params->ao_march == VX_MARCH_DDA) &&
                         (params->n_spec == 0 || params->spec_march == VX_MARCH_DDA);
    if (hits) {
        VX_REQUIRE(all_dda, "hit records are written by the DDA trace: every ray class must use VX_MARCH_DDA");
        VX_REQUIRE(((uintptr_t)hits & 7u) == 0, "hits must be 8-byte aligned");
        const size_t n64 = (size_t)view->width * view->height * (size_t)(params->n_shadow + params->n_ao +
                                                                         params->n_spec) * (sizeof(vx_trace_hit) / 8);
        if (n64) {
            hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(n64, 256)), dim3(256), 0, S(stream), reinterpret_cast<uint64_t*>(hits),
                               n64, 0ull);
            VX_LAUNCHED("k_fill_u64");
        }
        return launch_trace_march(ctx, vol, A, S(stream), hits);
    }
    if (all_dda && ctx->tuning.kernel != 1)
        return noalloc ? launch_trace_march_noalloc(ctx, vol, A, S(stream)) : launch_trace_march(ctx, vol, A, S(stream));
    hipLaunchKernelGGL(k_trace, dim3(8 * A.tiles.per_xcd), dim3(256), 0, S(stream), occ_view(vol), A);
    VX_LAUNCHED("k_trace");
    return VX_OK;
}

extern "C" vx_status vx_trace(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                              const float* palette, const vx_trace_params* params, const vx_trace_out* out,
                              void* stream) {
    return trace_impl(ctx, vol, view, g, palette, params, out, stream, false);
}

extern "C" vx_status vx_trace_hits(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                                   const float* palette, const vx_trace_params* params, const vx_trace_out* out,
                                   vx_trace_hit* hits, void* stream) {
    VX_REQUIRE(hits, "hits is NULL");
    return trace_impl(ctx, vol, view, g, palette, params, out, stream, false, hits);
}

extern "C" vx_status vx_trace_noalloc(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                                      const float* palette, const vx_trace_params* params, const vx_trace_out* out,
                                      void* stream) {
    return trace_impl(ctx, vol, view, g, palette, params, out, stream, true);
}

extern "C" vx_status vx_trace_rays(vx_ctx* ctx, const vx_volume* vol, const float* rays, uint32_t n, int32_t march,
                                   vx_ray_result* out, void* stream) {
    VX_REQUIRE(ctx && vol, "ctx/volume is NULL");
    VX_REQUIRE(n == 0 || (rays && out), "rays/out is NULL");
    VX_REQUIRE(march == VX_MARCH_DDA || march == VX_MARCH_SPARSE || march == VX_MARCH_SUPERSPARSE, "unknown marcher %d",
               march);
    if (n == 0) return VX_OK;
    if (vx_status st = use_device(ctx)) return st;
    hipLaunchKernelGGL(k_trace_rays, dim3(cdiv(n, 256)), dim3(256), 0, S(stream), occ_view(vol), rays, n, march, out);
    VX_LAUNCHED("k_trace_rays");
    return VX_OK;
}

// Grows a device array to n elements (contents not kept); work queued on the
// device that may still read the old one is waited for first.
template <class T>
static vx_status grow(T*& p, size_t& cap, size_t n) {
    if (cap >= n) return VX_OK;
    VX_HIP(hipDeviceSynchronize());
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
    VX_HIP(dev_malloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    cap = n;
    return VX_OK;
}

// The light pass as a list of (pixel, light) shadow rays: count (one atomic
// per 16x16 tile), read the total back, list, march one ray per lane, fold
// each pixel's terms in draw order. Blocks until the total is known.
static vx_status light_pass_list(vx_ctx* ctx, const vx_volume* vol, const LightArgs& A, hipStream_t stream) {
    const size_t npx = (size_t)A.view.width * A.view.height;
    size_t one = ctx->lt_total ? 1 : 0;
    if (vx_status st = grow(ctx->lt_total, one, 1)) return st;
    size_t pxc = ctx->lt_px_cap;
    if (vx_status st = grow(ctx->lt_first, pxc, npx)) return st;
    pxc = ctx->lt_px_cap;
    if (vx_status st = grow(ctx->lt_count, pxc, npx)) return st;
    ctx->lt_px_cap = std::max(ctx->lt_px_cap, npx);
    LightListArgs B;
    B.total = ctx->lt_total;
    B.first = ctx->lt_first;
    B.count = ctx->lt_count;
    B.rec = nullptr;
    B.shadow = nullptr;
    B.slot_light = nullptr;
    B.rfetch = nullptr;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, B.total, 1);
    hipLaunchKernelGGL(k_light_count, dim3(8 * A.tiles.per_xcd), dim3(256), 0, stream, A, B);
    VX_LAUNCHED("k_light_count");
    uint32_t total = 0;
    VX_HIP(hipMemcpyAsync(&total, B.total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    VX_HIP(hipStreamSynchronize(stream));
    if (total == 0) {
        hipLaunchKernelGGL(k_light_fold, dim3(cdiv(npx, 256)), dim3(256), 0, stream, A, B);
        VX_LAUNCHED("k_light_fold");
        return VX_OK;
    }
    size_t rc = ctx->lt_ray_cap;
    if (vx_status st = grow(ctx->lt_rec, rc, total)) return st;
    rc = ctx->lt_ray_cap;
    if (vx_status st = grow(ctx->lt_shadow, rc, total)) return st;
    rc = ctx->lt_ray_cap;
    if (vx_status st = grow(ctx->lt_slot_light, rc, total)) return st;
    ctx->lt_ray_cap = std::max<size_t>(ctx->lt_ray_cap, total);
    if (A.fetches) {
        if (vx_status st = grow(ctx->lt_rfetch, ctx->lt_fetch_cap, total)) return st;
    }
    B.rec = ctx->lt_rec;
    B.shadow = ctx->lt_shadow;
    B.slot_light = ctx->lt_slot_light;
    B.rfetch = A.fetches ? ctx->lt_rfetch : nullptr;
    hipLaunchKernelGGL(k_light_records, dim3(8 * A.tiles.per_xcd), dim3(256), 0, stream, A, B);
    VX_LAUNCHED("k_light_records");
    const uint32_t grid = std::min<uint32_t>(cdiv(total, 256), (uint32_t)ctx->gcache.cus * 8u);
    hipLaunchKernelGGL(k_light_march, dim3(grid), dim3(256), 0, stream, occ_view(vol), A, B);
    VX_LAUNCHED("k_light_march");
    hipLaunchKernelGGL(k_light_fold, dim3(cdiv(npx, 256)), dim3(256), 0, stream, A, B);
    VX_LAUNCHED("k_light_fold");
    return VX_OK;
}

extern "C" vx_status vx_light_pass(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                                   const float* palette, const vx_light* lights, uint32_t nlights,
                                   const vx_light_params* params, float* rgb, uint32_t* fetches, void* stream) {
    VX_REQUIRE(ctx && vol && g && palette && params && rgb, "ctx/volume/gbuf/palette/params/rgb is NULL");
    VX_REQUIRE(nlights <= 4096 && (nlights == 0 || lights), "lights: 0..4096 with a device array");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->pos && g->normal && g->material, "gbuf has a NULL pointer");
    auto okm = [](int32_t m) { return m == VX_MARCH_DDA || m == VX_MARCH_SPARSE || m == VX_MARCH_SUPERSPARSE; };
    VX_REQUIRE(okm(params->point_march) && okm(params->spot_march), "unknown marcher");
    if (vx_status st = use_device(ctx)) return st;
    LightArgs A;
    std::memset(&A, 0, sizeof(A));
    A.view = *view;
    A.tiles = tile_map(view->width, view->height);
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    A.palette = palette;
    A.lights = lights;
    A.nlights = nlights;
    A.params = *params;
    A.rgb = rgb;
    A.fetches = fetches;
    if (ctx->tuning.lights == 0 || nlights == 0) {
        hipLaunchKernelGGL(k_lights, dim3(8 * A.tiles.per_xcd), dim3(256), 0, S(stream), occ_view(vol), A);
        VX_LAUNCHED("k_lights");
        return VX_OK;
    }
    return light_pass_list(ctx, vol, A, S(stream));
}

extern "C" vx_status vx_reflection(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* g,
                                   const float* palette, const vx_reflection_io* io, const vx_skybox* sky,
                                   const vx_reflection_params* params, void* stream) {
    VX_REQUIRE(ctx && vol && g && palette && io && params, "ctx/volume/gbuf/palette/io/params is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->pos && g->normal && g->material, "gbuf has a NULL pointer");
    VX_REQUIRE(io->light && io->depth && io->out, "io: light, depth and out are required");
    VX_REQUIRE(((uintptr_t)io->out & 15u) == 0 && ((uintptr_t)io->light & 3u) == 0, "io.out must be 16-byte aligned");
    VX_REQUIRE(params->march == VX_MARCH_DDA || params->march == VX_MARCH_SPARSE ||
                   params->march == VX_MARCH_SUPERSPARSE, "unknown marcher %d", params->march);
    VX_REQUIRE(params->maxt > 0.0f, "maxt must be positive");
    VX_REQUIRE(!sky || !sky->faces || (sky->size >= 1 && sky->size <= 16384 && ((uintptr_t)sky->faces & 15u) == 0),
               "skybox: size 1..16384 and 16-byte aligned faces");
    if (vx_status st = use_device(ctx)) return st;
    ReflectionArgs A;
    std::memset(&A, 0, sizeof(A));
    A.view = *view;
    A.tiles = tile_map(view->width, view->height);
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    A.palette = palette;
    A.depth = io->depth;
    A.light = io->light;
    A.sky = (sky && sky->faces) ? sky->faces : nullptr;
    A.sky_size = A.sky ? sky->size : 0;
    A.params = *params;
    A.out = io->out;
    A.t = io->t;
    hipLaunchKernelGGL(k_reflection, dim3(8 * A.tiles.per_xcd), dim3(256), 0, S(stream), occ_view(vol), A);
    VX_LAUNCHED("k_reflection");
    return VX_OK;
}

extern "C" vx_status vx_ssao(vx_ctx* ctx, const vx_view* view, const vx_gbuf* g, const float* depth, uint32_t seed,
                             float* ssao, void* stream) {
    VX_REQUIRE(ctx && g && depth && ssao, "ctx/gbuf/depth/ssao is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->normal && g->material, "gbuf normal/material is NULL");
    if (vx_status st = use_device(ctx)) return st;
    SsaoArgs A;
    A.view = *view;
    A.nrm = g->normal;
    A.mat = g->material;
    A.depth = depth;
    A.seed = seed;
    A.out = ssao;
    const uint32_t npx = (uint32_t)view->width * (uint32_t)view->height;
    hipLaunchKernelGGL(k_ssao, dim3(cdiv(npx, 256)), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_ssao");
    return VX_OK;
}

extern "C" vx_status vx_gbuf_depth(vx_ctx* ctx, const vx_view* view, const vx_gbuf* g, int32_t volume_rid,
                                   float* depth, void* stream) {
    VX_REQUIRE(ctx && g && depth, "ctx/gbuf/depth is NULL");
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g->pos && g->material, "gbuf pos/material is NULL");
    if (vx_status st = use_device(ctx)) return st;
    DepthArgs A;
    mat_mul_f32(view->proj, view->view, A.pv);
    A.pos = g->pos;
    A.mat = g->material;
    A.depth = depth;
    A.npx = (uint32_t)view->width * (uint32_t)view->height;
    A.rid_scale = 1.0f - 1e-7f * (float)volume_rid;
    hipLaunchKernelGGL(k_gbuf_depth, dim3(cdiv(A.npx, 256)), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_gbuf_depth");
    return VX_OK;
}

static vx_status shade_args(const vx_view* view, const vx_gbuf* g, const vx_trace_out* vis, const float* ssao,
                            const float* palette, const float light[6], ShadeArgs& A) {
    if (vx_status st = check_view(view)) return st;
    VX_REQUIRE(g && vis && palette && light, "gbuf/vis/palette/light is NULL");
    VX_REQUIRE(g->pos && g->normal && g->material && vis->shadow && vis->ao, "NULL buffer in gbuf/vis");
    A.view = *view;
    A.npx = (size_t)view->width * view->height;
    A.pos = g->pos; A.nrm = g->normal; A.mat = g->material;
    A.shadow = vis->shadow; A.ao = vis->ao; A.ssao = ssao;
    A.palette = palette;
    for (int k = 0; k < 6; ++k) A.light[k] = light[k];
    A.rgb = nullptr;
    A.dl_drgb = nullptr;
    return VX_OK;
}

extern "C" vx_status vx_shade_forward(vx_ctx* ctx, const vx_view* view, const vx_gbuf* g, const vx_trace_out* vis,
                                      const float* ssao, const float* palette, const float light[6], float* rgb,
                                      void* stream) {
    VX_REQUIRE(ctx && rgb, "ctx/rgb is NULL");
    ShadeArgs A;
    if (vx_status st = shade_args(view, g, vis, ssao, palette, light, A)) return st;
    if (vx_status st = use_device(ctx)) return st;
    A.rgb = rgb;
    hipLaunchKernelGGL(k_shade_fwd, dim3(cdiv(A.npx, 256)), dim3(256), 0, S(stream), A);
    VX_LAUNCHED("k_shade_fwd");
    return VX_OK;
}

extern "C" vx_status vx_shade_backward(vx_ctx* ctx, const vx_view* view, const vx_gbuf* g, const vx_trace_out* vis,
                                       const float* ssao, const float* palette, const float light[6],
                                       const float* dl_drgb, float* dpalette, float* dlight, void* stream) {
    VX_REQUIRE(ctx && dl_drgb && dpalette && dlight, "ctx/dl_drgb/dpalette/dlight is NULL");
    ShadeArgs A;
    if (vx_status st = shade_args(view, g, vis, ssao, palette, light, A)) return st;
    if (vx_status st = use_device(ctx)) return st;
    A.dl_drgb = dl_drgb;
    const int nacc = 256 * 6 + 6;
    hipLaunchKernelGGL(k_zero_i64, dim3(cdiv(nacc, 256)), dim3(256), 0, S(stream), ctx->grad_acc, nacc);
    uint32_t grid = std::min<uint32_t>(cdiv(A.npx, 256), 1024);
    hipLaunchKernelGGL(k_shade_bwd, dim3(grid), dim3(256), 0, S(stream), A, ctx->grad_acc);
    hipLaunchKernelGGL(k_shade_bwd_finish, dim3(cdiv(nacc, 256)), dim3(256), 0, S(stream), ctx->grad_acc, dpalette,
                       dlight);
    VX_LAUNCHED("k_shade_bwd");
    return VX_OK;
}

static vx_status vx_shade_backward_batch_impl(vx_ctx* ctx, const vx_shade_item* items, uint32_t n,
                                             const float* palette, const float light[6], float* dpalette,
                                             float* dlight, void* stream) {
    VX_REQUIRE(ctx && dpalette && dlight, "ctx/dpalette/dlight is NULL");
    VX_REQUIRE(n >= 1 && n <= 65536 && items, "batched adjoint needs 1..65536 views");
    std::vector<ShadeArgs> h(n);
    std::vector<uint32_t> b0(n + 1);
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vx_shade_item& it = items[i];
        VX_REQUIRE(it.dl_drgb, "item %u: dl_drgb is NULL", i);
        if (vx_status st = shade_args(&it.view, &it.gbuf, &it.vis, it.ssao, palette, light, h[i])) return st;
        h[i].dl_drgb = it.dl_drgb;
        b0[i] = (uint32_t)blocks;
        blocks += std::min<uint32_t>(cdiv(h[i].npx, 256), 1024);
    }
    b0[n] = (uint32_t)blocks;
    if (vx_status st = use_device(ctx)) return st;
    if (ctx->shade_cap < n) {
        VX_HIP(hipDeviceSynchronize());
        (void)hipFree(ctx->shade_items);
        (void)hipFree(ctx->shade_block0);
        ctx->shade_items = nullptr;
        ctx->shade_block0 = nullptr;
        ctx->shade_cap = 0;
        VX_HIP(dev_malloc(&ctx->shade_items, n * sizeof(ShadeArgs)));
        VX_HIP(dev_malloc(&ctx->shade_block0, (n + 1) * sizeof(uint32_t)));
        ctx->shade_cap = n;
    }
    // pageable sources: the copies are complete when the calls return
    VX_HIP(hipMemcpyAsync(ctx->shade_items, h.data(), n * sizeof(ShadeArgs), hipMemcpyHostToDevice, S(stream)));
    VX_HIP(hipMemcpyAsync(ctx->shade_block0, b0.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          S(stream)));
    const int nacc = 256 * 6 + 6;
    hipLaunchKernelGGL(k_zero_i64, dim3(cdiv(nacc, 256)), dim3(256), 0, S(stream), ctx->grad_acc, nacc);
    hipLaunchKernelGGL(k_shade_bwd_batch, dim3((uint32_t)blocks), dim3(256), 0, S(stream), ctx->shade_items,
                       ctx->shade_block0, n, ctx->grad_acc);
    hipLaunchKernelGGL(k_shade_bwd_finish, dim3(cdiv(nacc, 256)), dim3(256), 0, S(stream), ctx->grad_acc, dpalette,
                       dlight);
    VX_LAUNCHED("k_shade_bwd_batch");
    return VX_OK;
}

// ---------------------------------------------------------------------------
// batched trace
// ---------------------------------------------------------------------------
struct vx_batch {
    vx_ctx* ctx;
    uint32_t n;
    const float* palette;
    vx_trace_params params;
    MarchItem* items;            // device
    size_t items_cap;
    uint16_t* tile_item;         // device: item of each global tile
    size_t tile_cap;
    MarchBuffers march;
    MarchArgs args;
    MarchShape shape;
    hipStream_t cap;             // capture stream
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint32_t captures;           // graphs captured (1 after creation, +1 per vx_batch_update)
    bool valid;                  // false after an update that failed part way
};

static void batch_free(vx_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    (void)hipFree(b->items);
    (void)hipFree(b->tile_item);
    march_buffers_free(b->march);
    delete b;
}

// The kernels of one batched launch: clear the queue counters (a fill kernel,
// no memset node), march, listed rays, resolve.
static hipError_t batch_enqueue(vx_batch* b, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(kCounterWords, 256)), dim3(256), 0, s, b->march.counters, (int)kCounterWords);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_march(b->shape.block, b->shape.steps, false, b->shape.grid, b->shape.rest_grid, s, b->args, e0,
                        e1);
}

// Plans the batch for items (tile queue, descriptor table, scratch, launch
// arguments) and captures its launch, replacing a previous graph. Arguments
// are checked before anything is changed; a failure after that leaves the
// batch invalid until the next successful plan.
static vx_status batch_plan(vx_batch* b, const vx_batch_item* items, uint32_t n) {
    vx_ctx* ctx = b->ctx;
    const vx_trace_params* params = &b->params;
    VX_REQUIRE(n >= 1 && n <= 65536 && items, "batch needs 1..65536 items");
    std::vector<MarchItem> h(n);
    uint64_t tiles = 0;
    bool fetches = false;
    for (uint32_t i = 0; i < n; ++i) {
        const vx_batch_item& it = items[i];
        VX_REQUIRE(it.volume, "item %u: volume is NULL", i);
        if (vx_status st = check_view(&it.view)) return st;
        VX_REQUIRE(it.gbuf.pos && it.gbuf.normal && it.gbuf.material && it.out.shadow && it.out.ao && it.out.spec,
                   "item %u: NULL buffer", i);
        TraceArgs T;
        std::memset(&T, 0, sizeof(T));
        T.view = it.view;
        T.params = *params;
        T.tiles = tile_map(it.view.width, it.view.height);
        T.pos = it.gbuf.pos; T.nrm = it.gbuf.normal; T.mat = it.gbuf.material;
        T.palette = b->palette;
        T.shadow = it.out.shadow; T.ao = it.out.ao; T.spec = it.out.spec; T.fetches = it.out.fetches;
        fetches = fetches || it.out.fetches;
        h[i] = make_march_item(it.volume, T, (uint32_t)tiles, tile_geom(ctx->tuning, it.view.width, it.view.height));
        tiles += h[i].ntiles;
        VX_REQUIRE(tiles < (1ull << 31), "batch too large");
    }
    const int32_t nrays = params->n_shadow + params->n_ao + params->n_spec;
    const size_t slots = (size_t)tiles * (size_t)nrays * 64;
    VX_REQUIRE(slots < (1ull << 32), "batch needs %zu ray slots, more than 2^32", slots);
    if (vx_status st = use_device(ctx)) return st;
    // from here on the batch changes: a previous graph's launches finish first
    VX_HIP(hipDeviceSynchronize());
    b->valid = false;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->n = n;
    std::vector<uint16_t> tile_item((size_t)tiles);
    for (uint32_t i = 0; i < n; ++i)
        std::fill(tile_item.begin() + h[i].tile0, tile_item.begin() + h[i].tile0 + h[i].ntiles, (uint16_t)i);
    hipError_t e = hipSuccess;
    if (b->items_cap < n) {
        (void)hipFree(b->items);
        b->items = nullptr;
        b->items_cap = 0;
        e = dev_malloc(&b->items, n * sizeof(MarchItem));
        if (e == hipSuccess) b->items_cap = n;
    }
    if (e == hipSuccess) e = hipMemcpy(b->items, h.data(), n * sizeof(MarchItem), hipMemcpyHostToDevice);
    if (e == hipSuccess && b->tile_cap < tiles) {
        (void)hipFree(b->tile_item);
        b->tile_item = nullptr;
        b->tile_cap = 0;
        e = dev_malloc(&b->tile_item, tile_item.size() * sizeof(uint16_t));
        if (e == hipSuccess) b->tile_cap = tiles;
    }
    if (e == hipSuccess && tiles)
        e = hipMemcpy(b->tile_item, tile_item.data(), tile_item.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? VX_ERR_OUT_OF_MEMORY : VX_ERR_DEVICE, "batch allocation failed: %s",
                    hipGetErrorString(e));
    }
    if (vx_status st = ensure_march(b->march, slots, fetches)) return st;
    MarchArgs& A = b->args;
    std::memset(&A, 0, sizeof(A));
    A.item0 = h[0];
    march_common(A, *params, ctx->tuning);
    A.items = b->items;
    A.tile_item = b->tile_item;
    A.nitems = n;
    A.ntiles = (uint32_t)tiles;
    A.counters = b->march.counters;
    A.deferred = b->march.deferred;
    A.deferred_cap = (uint32_t)slots;
    A.res = b->march.res;
    A.fres = fetches ? b->march.fres : nullptr;
    if (vx_status st = march_shape(ctx, false, b->shape)) return st;
    // a batch takes one queue item per tile (its launches hold many tiles)
    A.gen = (A.gen & ~0xffu) | 1u;
    set_bands(A, b->shape.grid * (uint32_t)(b->shape.block / 64));
    // capture the launch; vx_batch_launch replays it
    if (!b->cap) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = batch_enqueue(b, b->cap);
        if (le == hipSuccess) le = launch_status();
        e = hipStreamEndCapture(b->cap, &b->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(VX_ERR_DEVICE, "batch graph capture failed: %s", hipGetErrorString(e));
    }
    b->captures++;
    b->valid = true;
    return VX_OK;
}

static vx_status vx_batch_create_impl(vx_ctx* ctx, const vx_batch_item* items, uint32_t n, const float* palette,
                                     const vx_trace_params* params, vx_batch** out) {
    VX_REQUIRE(ctx && out && palette, "ctx/out/palette is NULL");
    if (vx_status st = check_params(params)) return st;
    VX_REQUIRE((params->n_shadow == 0 || params->shadow_march == VX_MARCH_DDA) &&
                   (params->n_ao == 0 || params->ao_march == VX_MARCH_DDA) &&
                   (params->n_spec == 0 || params->spec_march == VX_MARCH_DDA),
               "the batched trace runs the DDA marcher for every ray class");
    *out = nullptr;
    vx_batch* b = new (std::nothrow) vx_batch();
    if (!b) return fail(VX_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::memset(b, 0, sizeof(*b));
    b->ctx = ctx;
    b->palette = palette;
    b->params = *params;
    if (vx_status st = batch_plan(b, items, n)) {
        batch_free(b);
        return st;
    }
    *out = b;
    return VX_OK;
}

extern "C" vx_status vx_batch_launch(vx_batch* b, void* stream) {
    VX_REQUIRE(b, "batch is NULL");
    if (!b->valid) return fail(VX_ERR_STATE, "the batch's last update failed; update it again or destroy it");
    if (vx_status st = use_device(b->ctx)) return st;
    VX_HIP(hipGraphLaunch(b->exec, S(stream)));
    return VX_OK;
}

static vx_status vx_batch_launch_direct_impl(vx_batch* b, void* stream) {
    VX_REQUIRE(b, "batch is NULL");
    if (!b->valid) return fail(VX_ERR_STATE, "the batch's last update failed; update it again or destroy it");
    if (vx_status st = use_device(b->ctx)) return st;
    vx_ctx* ctx = b->ctx;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profile) {
        VX_HIP(hipEventCreate(&e0));
        if (hipError_t ce = hipEventCreate(&e1)) {
            (void)hipEventDestroy(e0);
            return fail(VX_ERR_DEVICE, "hipEventCreate failed: %s", hipGetErrorString(ce));
        }
        ctx->prof.emplace_back(e0, e1);
    }
    hipError_t e = batch_enqueue(b, S(stream), e0, e1);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "batch launch failed: %s", hipGetErrorString(e));
    return VX_OK;
}

extern "C" vx_status vx_batch_destroy(vx_batch* b) {
    if (!b) return VX_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    return VX_OK;
}

// ---------------------------------------------------------------------------
// entry points whose bodies allocate host containers
// ---------------------------------------------------------------------------
extern "C" vx_status vx_volume_rebuild(vx_volume* v, const vx_instance* inst, uint32_t n, int32_t thr,
                                       vx_region* dirty, uint32_t* ndirty, void* stream) {
    return guarded("vx_volume_rebuild",
                   [&] { return vx_volume_rebuild_impl(v, inst, n, thr, dirty, ndirty, stream); });
}

extern "C" vx_status vx_volume_build_levels(vx_volume* v, const vx_region* regions, uint32_t nregions, void* stream) {
    return guarded("vx_volume_build_levels",
                   [&] { return vx_volume_build_levels_impl(v, regions, nregions, stream); });
}

extern "C" vx_status vx_primary_instances(vx_ctx* ctx, const vx_vox_cmd* cmds, uint32_t n, const vx_view* view,
                                          const vx_gbuf* g, const vx_gbuf_ext* ext, void* stream) {
    return guarded("vx_primary_instances",
                   [&] { return vx_primary_instances_impl(ctx, cmds, n, view, g, ext, stream); });
}

extern "C" vx_status vx_shade_backward_batch(vx_ctx* ctx, const vx_shade_item* items, uint32_t n,
                                             const float* palette, const float light[6], float* dpalette,
                                             float* dlight, void* stream) {
    return guarded("vx_shade_backward_batch", [&] {
        return vx_shade_backward_batch_impl(ctx, items, n, palette, light, dpalette, dlight, stream);
    });
}

extern "C" vx_status vx_batch_create(vx_ctx* ctx, const vx_batch_item* items, uint32_t n, const float* palette,
                                     const vx_trace_params* params, vx_batch** out) {
    return guarded("vx_batch_create", [&] { return vx_batch_create_impl(ctx, items, n, palette, params, out); });
}

extern "C" vx_status vx_batch_update(vx_batch* b, const vx_batch_item* items, uint32_t n) {
    VX_REQUIRE(b, "batch is NULL");
    return guarded("vx_batch_update", [&] { return batch_plan(b, items, n); });
}

extern "C" vx_status vx_batch_info(const vx_batch* b, uint32_t* nitems, uint32_t* captures) {
    VX_REQUIRE(b, "batch is NULL");
    if (nitems) *nitems = b->n;
    if (captures) *captures = b->captures;
    return VX_OK;
}

extern "C" vx_status vx_batch_launch_direct(vx_batch* b, void* stream) {
    return guarded("vx_batch_launch_direct", [&] { return vx_batch_launch_direct_impl(b, stream); });
}
