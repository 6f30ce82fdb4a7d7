// vx_internal.hpp -- structures shared by the host layer and the kernels.
#pragma once

#include "vx_device.hpp"
#include "../../include/vx/vx_api.h"

namespace vx {

// One footprint loop of the rebuild (ShadowVoxSystem.cpp:137-177 / 22-40):
// origin and axis steps of M' = translate(M, -pivot), already scaled by 0.1.
struct FootprintOp {
    const uint8_t* model;
    int32_t sx, sy, sz;
    int32_t value;     // 0 erase, 1 write
    int32_t inst;      // instance slot for the dirty AABB
    float o[3], dx[3], dy[3], dz[3];
};

// Block -> pixel tile mapping (16x16 pixels per 256-thread block).
struct TileMap {
    int32_t width, height;
    uint32_t tiles_x, ntiles, per_xcd;
};

struct TraceArgs {
    vx_view view;
    vx_trace_params params;
    TileMap tiles;
    const float* pos;
    const float* nrm;
    const uint8_t* mat;
    const float* palette;
    uint8_t* shadow;
    float* ao;
    float* spec;
    uint32_t* fetches;
};

// One item of the ray-parallel trace: a volume and a view.
struct MarchItem {
    OccView V;                 // occupancy (literal path of listed rays)
    TraceArgs T;               // view, params, G-buffer, outputs
    const uint64_t* grid;      // march grid (border cells included, words bit-reversed)
    int32_t gx, gxy;           // march grid strides in words
    int32_t lim[3];            // 2*s+1: last in-bounds fine voxel per axis
    uint32_t tiles_x, ntiles;  // tiles of the view per row; queue entries (padded to whole blocks in order 1)
    uint32_t tile0;            // first global tile
    uint32_t tw_shift;         // tile width 1 << tw_shift pixels (3: 8x8, 4: 16x4, 2: 4x16)
    uint32_t order;            // 0: rows of tiles; 1: rows of 8x8-tile blocks, row-major inside a block
    uint32_t nbx;              // blocks per row (order 1)
};

// Tile geometry of a view for the ray-parallel trace under a tuning.
struct TileGeom {
    uint32_t tw_shift, order, tiles_x, tiles_y, nbx, ntiles;
};

// Queue counters of the ray-parallel trace, one 128-byte line each: 8 band
// counters, then the listed-ray count.
constexpr uint32_t kCounterStride = 32;
constexpr uint32_t kCounterWords = 9 * kCounterStride;

struct MarchArgs {
    MarchItem item0;           // the item of a single-view launch (a copy of items[0] otherwise)
    const MarchItem* items;    // device array of a batched launch
    const uint16_t* tile_item; // item of each global tile (batched launches; null: search)
    uint32_t nitems, ntiles;   // items, global tiles
    uint32_t band_tiles;       // tiles per XCD band
    uint32_t band;             // queue items per XCD band: band_tiles x chunks
    uint32_t gen;              // chunks (queue items per tile) | n_shadow << 8 | n_spec << 16
    uint32_t* counters;        // band counters, listed-ray count (kCounterStride apart)
    uint32_t* deferred;        // slots of rays left to the literal path
    uint32_t deferred_cap;     // capacity of deferred (every slot fits)
    float* res;                // per-ray best_t of a hit, a marker NaN otherwise; slot = ((tile * nrays + r) << 6) | pixel-in-tile
    uint16_t* fres;            // per-ray fetch count, or null
    vx_trace_hit* hits;        // per (pixel, ray) hit records of a single-view launch, or null
    int32_t nrays;
    int32_t refill;            // not-running lanes that trigger a refill
    float maxt[3];             // shadow, AO, spec
};

struct LightArgs {
    vx_view view;
    TileMap tiles;
    const float* pos;
    const float* nrm;
    const uint8_t* mat;
    const float* palette;
    const vx_light* lights;
    uint32_t nlights;
    vx_light_params params;
    float* rgb;
    uint32_t* fetches;
};

// LightReflection (LightReflection.frag:60-139).
struct ReflectionArgs {
    vx_view view;
    TileMap tiles;
    const float* pos;
    const float* nrm;
    const uint8_t* mat;
    const float* palette;
    const float* depth;        // D24 depth per pixel (1 = sky)
    const float* light;        // 4 per pixel (lightDiffuse)
    const float* sky;          // cube faces (6 * size^2 * 4) or null: the procedural sky
    int32_t sky_size;
    vx_reflection_params params;
    float* out;                // 4 per pixel
    float* t;                  // per pixel or null
};

// Screen-space occlusion (LightAmbient.frag:59-109).
struct SsaoArgs {
    vx_view view;
    const float* nrm;
    const uint8_t* mat;
    const float* depth;
    uint32_t seed;
    float* out;
};

// Per-pixel D24 depth of a G-buffer (vx_gbuf_depth).
struct DepthArgs {
    float pv[16];              // Projection * View
    const float* pos;
    const uint8_t* mat;
    float* depth;
    uint32_t npx;
    float rid_scale;           // 1 - 1e-7 * VolumeRID
};

// The light pass's ray list (k_light_count / _records / _march / _fold).
// A listed shadow ray: origin, direction and length from the light's draw
// (light_ray), its marcher, and its slot (its pixel's first slot + the
// light's rank among the pixel's lights).
struct LightRec {
    float o[3];
    float d[3];
    float hit_dist;
    int32_t march;
    uint32_t slot;
};

struct LightListArgs {
    uint32_t* total;           // rays listed (one atomic per tile)
    uint32_t* first;           // per pixel: its first record
    uint32_t* count;           // per pixel: its in-range lights
    LightRec* rec;             // the records, per tile light by light (pixels in order)
    float* shadow;             // per slot: the march's shadow factor (0 or 1)
    uint16_t* slot_light;      // per slot: the light's index
    uint16_t* rfetch;          // per slot: the marcher's fetch count, or null
};

// A model's three mips (X-fastest bytes) and mip-0 dims.
struct ModelMips {
    const uint8_t* mips[3];
    int32_t dims[3];
};

// One VoxCmd with its per-instance constants (host-computed).
struct InstConst {
    float w[16];      // WorldMatrix (pivot applied)
    float lw[16];     // LastWorldMatrix
    float winv[16];   // inverse(WorldMatrix), from a float64 inverse
    float lc[3];      // localCameraPos: inverse(WorldMatrix) * camera * 10
    int32_t rid;      // VolumeRID
    int32_t palette;  // PalleteIndex
    int32_t pad;
    ModelMips m;
};

constexpr uint32_t kTileInstCap = 64;

struct BinArgs {
    const int32_t* rects;   // per instance x0, y0, x1, y1 (pixels, inclusive); x0 > x1 = culled
    uint32_t ninst, tiles_x, ntiles;
    uint16_t* lists;        // ntiles * kTileInstCap
    uint32_t* counts;       // ntiles
};

struct PrimaryInstArgs {
    vx_view view;
    TileMap tiles;
    float pv[16];           // Projection * View
    float plv[16];          // Projection * LastView
    const InstConst* inst;
    uint32_t ninst;
    const uint16_t* lists;
    const uint32_t* counts;
    float* pos;
    float* nrm;
    uint8_t* mat;
    float* motion;
    float* depth;
    uint16_t* palette;
};

// LightTAA (LightTAA.frag): per-pixel attributes of this frame and the last.
struct TaaArgs {
    int32_t width, height;
    int32_t frame;
    uint32_t seed;
    const float* light;       // 3 per pixel
    const float* last;        // 4 per pixel
    const float* motion;      // 2 per pixel or NULL
    const float* depth;       // 1 per pixel
    const float* normal;      // 3 per pixel
    const uint8_t* mat;       // material index per pixel
    const float* palette;     // 256 x 6
    float* out;               // 4 per pixel
    float4* rec;              // 4 float4 per pixel: packed attributes (k_taa_pack)
};

struct PrimaryArgs {
    vx_view view;
    TileMap tiles;
    float pv[16];             // projection * view (the box coverage test)
    const uint8_t* mips[3];
    int32_t dims[3];
    float* pos;
    float* nrm;
    uint8_t* mat;
};

struct ShadeArgs {
    vx_view view;
    size_t npx;
    const float* pos;
    const float* nrm;
    const uint8_t* mat;
    const uint8_t* shadow;
    const float* ao;
    const float* ssao;
    const float* palette;
    float light[6];
    float* rgb;
    const float* dl_drgb;
};

__global__ void k_terrain_heights(float* h, int32_t nx, int32_t ny, int32_t nz, uint32_t seed, float fill);
__global__ void k_terrain_fill(uint8_t* vol, const float* h, int32_t nx, int32_t ny, int32_t nz, uint32_t seed,
                               int32_t use_terrain);
__global__ void k_boxes(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed, int32_t b0, int32_t emin,
                        int32_t emax, int32_t grounded);
__global__ void k_build_mip(const uint8_t* parent, int32_t px, int32_t py, int32_t pz, uint8_t* out);
__global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v);
__global__ void k_bytes_to_cells(uint64_t* cells, const uint8_t* bytes, OccView V, int32_t rx, int32_t ry, int32_t rz,
                                 int32_t rw, int32_t rh, int32_t rd);
__global__ void k_cells_to_bytes(uint8_t* bytes, const uint64_t* cells, OccView V);
__global__ void k_footprint(uint64_t* cells, OccView V, const FootprintOp* ops, int32_t* aabb, int32_t thr);
__global__ void k_level1(uint64_t* words, const uint64_t* cells, OccView V, int32_t w0x, int32_t w0y, int32_t w0z,
                         int32_t nwx, int32_t nwy, int32_t nwz);
__global__ void k_level_up(uint64_t* words, const uint64_t* lower, int32_t lwx, int32_t lwy, int32_t lwz, int32_t uwx,
                           int32_t uwy, int32_t uwz, int32_t w0x, int32_t w0y, int32_t w0z, int32_t nwx, int32_t nwy,
                           int32_t nwz);
__global__ void k_trace_rays(OccView V, const float* rays, uint32_t n, int32_t march, vx_ray_result* out);
__global__ void __launch_bounds__(256) k_trace(OccView V, TraceArgs A);
__global__ void k_zero_u32(uint32_t* p, int n);
__global__ void __launch_bounds__(256) k_lights(OccView V, LightArgs A);
__global__ void __launch_bounds__(256) k_light_count(LightArgs A, LightListArgs B);
__global__ void __launch_bounds__(256) k_light_records(LightArgs A, LightListArgs B);
__global__ void __launch_bounds__(256) k_light_march(OccView V, LightArgs A, LightListArgs B);
__global__ void __launch_bounds__(256) k_light_fold(LightArgs A, LightListArgs B);
__global__ void __launch_bounds__(256) k_primary(PrimaryArgs A);
__global__ void __launch_bounds__(256) k_reflection(OccView V, ReflectionArgs A);
__global__ void __launch_bounds__(256) k_gbuf_depth(DepthArgs A);
__global__ void __launch_bounds__(256) k_ssao(SsaoArgs A);
__global__ void __launch_bounds__(256) k_bin_instances(BinArgs A);
__global__ void __launch_bounds__(256) k_primary_inst(PrimaryInstArgs A);
__global__ void __launch_bounds__(256) k_taa_pack(TaaArgs A);
__global__ void __launch_bounds__(256) k_light_taa(TaaArgs A);
__global__ void k_shade_fwd(ShadeArgs A);
__global__ void __launch_bounds__(256) k_shade_bwd(ShadeArgs A, long long* acc);
__global__ void __launch_bounds__(256) k_shade_bwd_batch(const ShadeArgs* items, const uint32_t* block0, uint32_t n,
                                                         long long* acc);
__global__ void k_shade_bwd_finish(const long long* acc, float* dpal, float* dlight);
__global__ void k_zero_i64(long long* p, int n);

__global__ void k_mgrid(uint64_t* grid, OccView V, int32_t gx, int32_t gy, int32_t g0x, int32_t g0y, int32_t g0z,
                        int32_t nx, int32_t ny, int32_t nz);

// Ray-parallel trace, instantiated for 256/512-thread blocks and 2/4/8 DDA
// steps between refills; launch_march enqueues march, listed rays and resolve
// (e0 / e1, when given, are recorded around the march kernel).
const void* march_kernel(int block, int steps, bool single);
hipError_t launch_march(int block, int steps, bool single, uint32_t grid, uint32_t rest_grid, hipStream_t s,
                        const MarchArgs& A, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

}  // namespace vx
