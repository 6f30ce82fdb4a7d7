/*
 * vx_api.h -- C ABI of the MI355X (gfx950) voxel-traversal lighting path.
 *
 * Drop-in replacement for VoxelEngine's voxelized-scene rebuild and the voxel
 * ray marching behind its shadow / ambient-occlusion / specular-occlusion
 * passes. Each entry point names the reference interface it replaces
 * (paths relative to the VoxelEngine source tree).
 *
 * Conventions
 *  - Plain C, POD structs, no exceptions across the ABI. Every call returns
 *    vx_status: 0 on success, a negative VX_ERR_* code on failure; the message
 *    of the last failure on the calling thread is vx_last_error().
 *  - Device pointers (suffix _dev or documented as device) are HIP device
 *    memory owned by the caller; host pointers are owned by the caller.
 *    Volumes and models are opaque handles owned by the library.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream). Work
 *    is stream-ordered; calls return once the work is enqueued unless noted.
 *  - A context is used by one host thread at a time.
 *  - Matrices are column-major 4x4 (glm layout). World units are the
 *    reference's (one world unit = ten fine voxels).
 */
#ifndef VX_API_H
#define VX_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VX_API_VERSION 1

typedef int32_t vx_status;
enum {
    VX_OK = 0,
    VX_ERR_INVALID_ARGUMENT = -1,
    VX_ERR_OUT_OF_MEMORY = -2,
    VX_ERR_DEVICE = -3,       /* no usable device or a HIP runtime failure */
    VX_ERR_UNSUPPORTED = -4,
    VX_ERR_RANGE = -5,        /* a size beyond a documented limit */
    VX_ERR_STATE = -6
};

/* Marchers. DDA = raycastShadowVolume (Light.frag:29-81); SPARSE =
 * raycastShadowVolumeSparse (Light.frag:131-173); SUPERSPARSE =
 * raycastShadowVolumeSuperSparse (Light.frag:175-217). */
enum { VX_MARCH_DDA = 0, VX_MARCH_SPARSE = 1, VX_MARCH_SUPERSPARSE = 2 };

/* Instance change kinds. UPDATE = ShadowVoxSystem::OnUpdate for an entity with
 * the Changed tag (ShadowVoxSystem.cpp:119-191); DESTROY =
 * ShadowVoxSystem::OnVoxDestroyed (ShadowVoxSystem.cpp:7-53). */
enum { VX_INST_UPDATE = 1, VX_INST_DESTROY = 2 };

typedef struct vx_ctx vx_ctx;
typedef struct vx_volume vx_volume;
typedef struct vx_model vx_model;
typedef struct vx_batch vx_batch;

/* evk ImageRegion (Vendor/evk/evk.h:152-155) in texels of the occupancy volume. */
typedef struct {
    int32_t x, y, z;
    uint32_t w, h, d;
} vx_region;

/* One changed entity: VoxRenderer {Vox, Pivot} + Transform {PreviousWorldMatrix,
 * WorldMatrix} (Components.h:30-35,60-67). model is a DEVICE pointer to the
 * model's X-fastest material bytes (VoxAsset::PixelAt, VoxAsset.h:52-56). */
typedef struct {
    const uint8_t* model;
    int32_t sx, sy, sz;
    float prev[16];
    float cur[16];
    float pivot[3];
    int32_t flags;     /* VX_INST_UPDATE or VX_INST_DESTROY */
} vx_instance;

/* ViewData (Sources/Graphics/Renderer/View.h:16-31; Common.frag:45-61) plus the
 * pixel grid it is rendered at. */
typedef struct {
    float last_view[16];
    float view[16];
    float inv_view[16];
    float proj[16];
    float inv_proj[16];
    float res[2];
    float ires[2];
    float cam_pos[3];
    float jitter[2];
    int32_t frame;
    int32_t width, height;
} vx_view;

/* Ray set-up of the lighting passes. Reference defaults: shadow SPARSE to 128
 * (LightAmbient.frag:167), AO SUPERSPARSE (LightAmbient.frag:121), spec-occ
 * SPARSE to 256 (LightReflection.frag:113). */
typedef struct {
    int32_t shadow_march, ao_march, spec_march;
    int32_t n_shadow, n_ao, n_spec;     /* rays per pixel: 0/1, 0..64, 0/1 */
    float shadow_maxt, ao_maxt, spec_maxt;
    float light_dir[3];                 /* direction toward the light (SUN_DIR) */
    uint32_t seed;
    int32_t frame;
    int32_t jitter_shadow;              /* 1: soft-shadow jitter of LightAmbient.frag:151-156 */
    int32_t pad;
} vx_trace_params;

/* G-buffer (device). Replaces the Geometry pass attachments (Graphics.h:46-60)
 * with fp32 hit points in fine-voxel units. material 0 = no hit. */
typedef struct {
    float* pos;          /* width*height*3 */
    float* normal;       /* width*height*3 */
    uint8_t* material;   /* width*height */
} vx_gbuf;

/* Per-pixel visibility (device): shadow 1 = lit; ao = mean (t/cap)^2 over the
 * AO rays; spec = reflection-ray distance / spec_maxt (1 = unoccluded).
 * fetches (optional, may be NULL): occupancy-byte fetches the reference
 * algorithm performs for the pixel's rays. */
typedef struct {
    uint8_t* shadow;
    float* ao;
    float* spec;
    uint32_t* fetches;
} vx_trace_out;

/* Explicit-ray result, 48 bytes. */
typedef struct {
    int32_t hit;       /* 1 hit */
    int32_t term;      /* DDA: 0 loop exhausted, 1 hit, 2 out of bounds; sparse: 0 miss, 1 hit */
    int32_t steps;     /* DDA: reference step count nt (x4 restart included); sparse: iterations */
    int32_t fetches;   /* occupancy-byte fetches of the reference algorithm */
    int32_t voxel[3];  /* DDA hit voxel (fine units) */
    float t;           /* DDA: best_t at the hit; sparse: returned distance */
    float hitpos[3];   /* DDA: origin + dir*best_t */
    float normal[3];   /* DDA: -sign(dir)*select */
} vx_ray_result;

/* One ray of the trace as raycastShadowVolume reports it (Light.frag:57-63,
 * 73-79), written by vx_trace_hits; 40 bytes. Fields other than term, steps
 * and fetches are 0 unless the ray hit (term 1). */
typedef struct {
    int32_t term;      /* 0 loop exhausted (n == 256 or totalt >= maxT), 1 hit, 2 out of bounds */
    int32_t steps;     /* the reference's step count nt (the x4 restart of a term-0 ray included) */
    int32_t fetches;   /* occupancy-byte fetches of the reference algorithm */
    int32_t voxel[3];  /* hit voxel (fine units) */
    float t;           /* best_t = dot(t_max, select) at the hit */
    float normal[3];   /* -sign(dir) * select at the hit: the face the ray leaves the voxel by */
} vx_trace_hit;

/* Synthetic scene description (build-defined inputs for tests and benchmarks). */
typedef struct {
    uint32_t seed;
    float terrain_fill;     /* target fill of the fBm heightfield, 0 = no terrain */
    int32_t nboxes;
    int32_t box_min, box_max;
    int32_t boxes_grounded;
} vx_scene_desc;

/* Launch tuning of the trace kernels (no counterpart in the reference, whose
 * launch shape is one full-screen draw per pass). 0 / -1 fields mean "default";
 * the autotuner (voxelengine_amd/autotune.py) picks values per workload. */
typedef struct {
    int32_t kernel;          /* 0 auto, 1 per-pixel 16x16 tiles, 2 ray-parallel DDA (all-DDA rays only) */
    int32_t block;           /* ray-parallel kernel threads per block: 256 or 512 (0 = 256) */
    int32_t blocks_per_cu;   /* persistent grid = CUs x this (0 = occupancy limit) */
    int32_t refill;          /* lanes without a running ray that trigger a refill, 1..64 (0 = 16) */
    int32_t steps;           /* DDA steps between refill checks: 2, 4, 8 or 16 (0 = 16) */
    int32_t tile;            /* pixels of a ray-parallel tile: 0 = 8x8, 1 = 16x4, 2 = 4x16 */
    int32_t order;           /* tile queue order: 0 = rows of tiles, 1 = rows of 8x8-tile blocks */
    int32_t lights;          /* vx_light_pass: 0 = a loop per pixel; 1 = list of (pixel, light) rays, one per lane */
    int32_t chunks;          /* ray-parallel queue items per tile of a single-view trace: its rays in this
                                many chunks, 1..64 (0 = auto); a batch takes one item per tile */
} vx_trace_tuning;

/* ---------------------------------------------------------------- context */
int32_t vx_api_version(void);
const char* vx_last_error(void);
/* Failure injection for tests of error handling (no counterpart in the
 * reference): the `after`-th device allocation (VX_FAIL_ALLOC) or kernel
 * launch check (VX_FAIL_LAUNCH) from now on, counting from 0, fails once
 * (the call returns VX_ERR_OUT_OF_MEMORY / VX_ERR_DEVICE and releases what it
 * allocated); -1 injects nothing. The environment variables VX_FAIL_ALLOC=n
 * and VX_FAIL_LAUNCH=n set the same counters when the library loads. */
enum { VX_FAIL_ALLOC = 0, VX_FAIL_LAUNCH = 1 };
vx_status vx_debug_inject_failure(int32_t kind, int64_t after);
/* Replaces Graphics module start-up (Graphics.cpp) for this path: selects the
 * HIP device and allocates per-device state. */
vx_status vx_ctx_create(int32_t device, vx_ctx** out);
vx_status vx_ctx_destroy(vx_ctx* ctx);
/* Set (NULL resets) / read the context's trace launch tuning. */
vx_status vx_ctx_set_trace_tuning(vx_ctx* ctx, const vx_trace_tuning* tuning);
vx_status vx_ctx_get_trace_tuning(const vx_ctx* ctx, vx_trace_tuning* tuning);
/* Kernel timing of the ray-parallel trace (no counterpart in the reference):
 * while enabled, every vx_trace / vx_batch_launch_direct records HIP events
 * around its march kernel on the launch stream. vx_ctx_profile(ctx, 1) starts
 * a new collection, 0 stops it; vx_ctx_profile_read waits for the recorded
 * launches and returns their total march-kernel milliseconds and count. */
vx_status vx_ctx_profile(vx_ctx* ctx, int32_t enable);
vx_status vx_ctx_profile_read(vx_ctx* ctx, double* march_ms, uint32_t* launches);

/* ---------------------------------------------------------------- models
 * VoxAsset (VoxAsset.h:8-62): R8 material volume with 3 mips. Dims are rounded
 * up to multiples of 4 (VoxAsset.h:27-29). */
vx_status vx_model_create(vx_ctx* ctx, int32_t sx, int32_t sy, int32_t sz, vx_model** out);
vx_status vx_model_destroy(vx_model* m);
vx_status vx_model_dims(const vx_model* m, int32_t dims[3]);
/* Device pointer of mip level (0..2), X-fastest bytes of dims>>mip. */
vx_status vx_model_mip_ptr(const vx_model* m, int32_t mip, uint8_t** dev_ptr);
/* Host upload of mip 0 (VoxAsset::Upload's first-level copy, VoxAsset.cpp:9-13). */
vx_status vx_model_upload(vx_model* m, const uint8_t* host, size_t nbytes, void* stream);
/* Download mip level (0..2) to host, nbytes = product of dims>>mip. Synchronous. */
vx_status vx_model_download(const vx_model* m, int32_t mip, uint8_t* host, size_t nbytes, void* stream);
/* Mip build: first non-zero child, VoxAsset::Upload (VoxAsset.cpp:16-53). */
vx_status vx_model_build_mips(vx_model* m, void* stream);
/* Fill mip 0 with a synthetic terrain/box scene (build-defined). */
vx_status vx_model_generate(vx_model* m, const vx_scene_desc* desc, void* stream);

/* ---------------------------------------------------------------- occupancy volume
 * ShadowVoxSystem's volume (ShadowVoxSystem.cpp:55-79): texel dims sx,sy,sz;
 * each texel a 2x2x2 fine-voxel bitmask. nlevels coarse occupancy levels
 * (0..3) of 8^3, 64^3 and 512^3 fine-voxel cells. Limits: texel dims <= 4096. */
vx_status vx_volume_create(vx_ctx* ctx, int32_t sx, int32_t sy, int32_t sz, int32_t nlevels,
                           vx_volume** out);
vx_status vx_volume_destroy(vx_volume* v);
vx_status vx_volume_dims(const vx_volume* v, int32_t dims[3], int32_t* nlevels);
vx_status vx_volume_clear(vx_volume* v, void* stream);
/* Upload texel regions from a host copy of the whole volume in the reference
 * byte layout (X-fastest); replaces the staging copy of ShadowVoxSystem.cpp:193-200
 * (evk.cpp:759-780). Synchronous. */
vx_status vx_volume_upload_region(vx_volume* v, const uint8_t* host_bytes, const vx_region* regions,
                                  uint32_t nregions, void* stream);
/* Download the whole volume in the reference byte layout. Synchronous. */
vx_status vx_volume_download(const vx_volume* v, uint8_t* host_bytes, size_t nbytes, void* stream);
/* ShadowVoxSystem::OnUpdate / OnVoxDestroyed for n changed instances, in order.
 * Writes one dirty region per instance whose footprint reached the volume
 * (ShadowVoxSystem.cpp:180-190) to dirty (capacity n). Blocks until the
 * regions are known. */
vx_status vx_volume_rebuild(vx_volume* v, const vx_instance* inst, uint32_t n, int32_t occ_threshold,
                            vx_region* dirty, uint32_t* ndirty, void* stream);
/* Rebuild the coarse occupancy levels over the given texel regions (NULL = all). */
vx_status vx_volume_build_levels(vx_volume* v, const vx_region* regions, uint32_t nregions, void* stream);
/* Download level (1..nlevels) occupancy words (u64, 4x4x4 cells each). Synchronous. */
vx_status vx_volume_download_level(const vx_volume* v, int32_t level, uint64_t* host, size_t nwords,
                                   size_t* words_total, void* stream);

/* ---------------------------------------------------------------- passes
 * Primary hit: GeometryVoxelPipeline::Use (GeometryVoxelPipeline.h:48-70) +
 * GeometryVoxel.frag main/intersectVolume (GeometryVoxel.frag:63-183) for the
 * world model with an identity WorldMatrix. A pixel is marched where the
 * model's box is drawn (its far faces, GeometryVoxel.vert:5-42, at a view
 * depth >= NEAR); elsewhere, e.g. where the box is behind the camera, it stays
 * empty (position, normal and material 0). */
vx_status vx_primary(vx_ctx* ctx, const vx_model* model, const vx_view* view, const vx_gbuf* gbuf,
                     void* stream);
/* One voxel draw of the geometry pass (GeometryVoxelPipeline::Cmd,
 * GeometryVoxelPipeline.h:29-35, filled by WorldRenderer.cpp:210-220):
 * world is the entity's WorldMatrix translated by -Pivot, last_world the
 * PreviousWorldMatrix translated the same way. 144 bytes. */
typedef struct {
    const vx_model* model;
    float world[16];
    float last_world[16];
    int32_t volume_rid;    /* VolumeRID: scales the fragment depth by 1 - 1e-7 * rid */
    int32_t palette;       /* PalleteIndex, written to vx_gbuf_ext.palette */
} vx_vox_cmd;

/* Optional geometry-pass outputs beyond vx_gbuf (device, NULL = not written). */
typedef struct {
    float* motion;         /* width*height*2: out_Motion (GeometryVoxel.frag:171) */
    float* depth;          /* width*height: the D24 depth as read back, q / (2^24 - 1); 1 = no hit */
    uint16_t* palette;     /* width*height: the hit's PalleteIndex (0 = no hit) */
} vx_gbuf_ext;

/* The geometry pass over n voxel draws (GeometryVoxelPipeline::Use,
 * GeometryVoxelPipeline.h:48-70; GeometryVoxel.vert:62-78,
 * GeometryVoxel.frag:127-183): each pixel's camera ray is marched through
 * every draw whose box covers it, in model space with the model's mips, and
 * the D24 depth test (LESS, draws in array order) keeps the nearest hit. A box
 * covers a pixel when the pixel's forward ray leaves it at a view depth >= NEAR
 * (its inward-facing triangles and back-face culling draw the far faces,
 * GeometryVoxel.vert:5-42, evk.cpp:470-471).
 * gbuf.pos is the world hit point in fine-voxel units (world * 10), gbuf.normal
 * the world normal. n <= 4096 (MAX_INSTANCES, GeometryVoxelPipeline.h:17). */
vx_status vx_primary_instances(vx_ctx* ctx, const vx_vox_cmd* cmds, uint32_t n, const vx_view* view,
                               const vx_gbuf* gbuf, const vx_gbuf_ext* ext, void* stream);

/* Shadow + AO + spec-occ rays: LightAmbientPipeline::Use (LightAmbientPipeline.h:35)
 * and LightReflectionPipeline::Use (LightReflectionPipeline.h:34) ray set-up and
 * marching. palette_dev: 256*6 floats {albedo rgb, roughness, metallic, emissive}. */
vx_status vx_trace(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* gbuf,
                   const float* palette_dev, const vx_trace_params* params, const vx_trace_out* out,
                   void* stream);
/* The same trace as vx_trace, with no allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * ray-parallel path's scratch (per-ray results, listed rays, fetch counts,
 * queue counters) is a file-scope device array of the library, one per GPU,
 * of 2^22 ray slots (about 40 MiB), held until the process ends; the view is
 * traced in consecutive ranges of 8x8 tiles whose rays fit it (at most
 * 2^22 / (64 * rays per pixel) tiles per range), one range after another on
 * `stream`. Every call on a GPU shares that GPU's scratch: calls that overlap
 * on one GPU (other streams, other threads, other contexts) race and must be
 * serialized by the caller. No kernel timing (vx_ctx_profile) is recorded.
 * VX_ERR_RANGE if one tile's rays exceed the scratch. */
vx_status vx_trace_noalloc(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* gbuf,
                           const float* palette_dev, const vx_trace_params* params, const vx_trace_out* out,
                           void* stream);
/* vx_trace (all rays through the DDA marcher) that also writes every ray's
 * hit record: hits_dev[pixel * nrays + r] for ray r of a pixel (r in the
 * order shadow, AO, spec; nrays = n_shadow + n_ao + n_spec), all zero for sky
 * pixels. The same ray-parallel kernel as vx_trace runs, in its step-counting
 * form. VX_ERR_UNSUPPORTED unless every ray class uses VX_MARCH_DDA. */
vx_status vx_trace_hits(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* gbuf,
                        const float* palette_dev, const vx_trace_params* params, const vx_trace_out* out,
                        vx_trace_hit* hits_dev, void* stream);
/* Explicit rays (device, n*7 floats: origin3, dir3, maxt) through one marcher. */
vx_status vx_trace_rays(vx_ctx* ctx, const vx_volume* vol, const float* rays_dev, uint32_t n,
                        int32_t march, vx_ray_result* out_dev, void* stream);
/* ---------------------------------------------------------------- batched trace
 * Many scenes and views of uneven size in one persistent launch (no
 * counterpart in the reference, which issues one full-screen draw per pass and
 * view). All items share the palette and the ray set-up; each has its own
 * volume, view, G-buffer and outputs. vx_batch_create plans the tile queue,
 * uploads a device descriptor table and captures the launch in a hipGraph;
 * vx_batch_launch replays the graph on `stream`. The item buffers must stay
 * valid while the batch exists. Two launches of one batch must not overlap. */
typedef struct {
    const vx_volume* volume;
    vx_view view;
    vx_gbuf gbuf;
    vx_trace_out out;
} vx_batch_item;

vx_status vx_batch_create(vx_ctx* ctx, const vx_batch_item* items, uint32_t n, const float* palette_dev,
                          const vx_trace_params* params, vx_batch** out);
vx_status vx_batch_launch(vx_batch* batch, void* stream);
/* Replaces the batch's items (any number, any view sizes) keeping its palette
 * and ray set-up: waits for the device, re-plans the tile queue, re-uploads
 * the descriptor table, grows the scratch when needed and captures the launch
 * again (the launch's shape lives in the graph). Arguments are checked
 * first; if anything fails after that, the batch reports VX_ERR_STATE until a
 * later update succeeds. */
vx_status vx_batch_update(vx_batch* batch, const vx_batch_item* items, uint32_t n);
/* The batch's item count and how many times its launch has been captured. */
vx_status vx_batch_info(const vx_batch* batch, uint32_t* nitems, uint32_t* captures);
/* The same work as vx_batch_launch, launched directly instead of through the graph. */
vx_status vx_batch_launch_direct(vx_batch* batch, void* stream);
vx_status vx_batch_destroy(vx_batch* batch);

/* Shaded radiance: LightAmbient.frag:178-214 with PBR.frag:3-68, visibility
 * given. light: host {dir3, rgb3}. ssao_dev may be NULL (1.0). rgb_dev: npx*3. */
vx_status vx_shade_forward(vx_ctx* ctx, const vx_view* view, const vx_gbuf* gbuf,
                           const vx_trace_out* vis, const float* ssao_dev, const float* palette_dev,
                           const float light[6], float* rgb_dev, void* stream);
/* Adjoint of vx_shade_forward with visibility fixed. dl_drgb_dev: npx*3.
 * Writes dpalette_dev (256*6) and dlight_dev (6); deterministic. */
vx_status vx_shade_backward(vx_ctx* ctx, const vx_view* view, const vx_gbuf* gbuf,
                            const vx_trace_out* vis, const float* ssao_dev, const float* palette_dev,
                            const float light[6], const float* dl_drgb_dev, float* dpalette_dev,
                            float* dlight_dev, void* stream);

/* ---------------------------------------------------------------- lights
 * Point and spot lights: LightPointPipeline::Use + DrawLight
 * (LightPointPipeline.h:58-100, LightPoint.frag:85-152) and LightSpotPipeline
 * (LightSpotPipeline.h:60-104, LightSpot.frag:73-140). Each light's
 * full-screen draw becomes one iteration of a per-pixel light loop. */
enum { VX_LIGHT_POINT = 0, VX_LIGHT_SPOT = 1 };

/* PointLight / SpotLight (Common.frag:15-30), world units. color is
 * Color * Intensity and direction the light's WorldMatrix[2]
 * (WorldRenderer.cpp:247,255). 64 bytes. */
typedef struct {
    int32_t type;
    float position[3];
    float range;
    float color[3];
    float attenuation;
    float direction[3];    /* spot only */
    float angle;           /* spot only */
    float angle_attenuation;
    int32_t pad;
} vx_light;

typedef struct {
    int32_t point_march;   /* reference: VX_MARCH_SPARSE (LightPoint.frag:125) */
    int32_t spot_march;    /* reference: VX_MARCH_SUPERSPARSE (LightSpot.frag:113) */
    uint32_t seed;
    int32_t frame;
} vx_light_params;

/* Adds every light's shadowed direct term (PBRDirectLight, PBR.frag:48-68)
 * to rgb_dev (npx*3), lights in array order, as the reference's additive
 * blending of the light draws does. lights_dev: device array of nlights
 * (0..4096). fetches_dev (optional): the marchers' fetch counts per pixel. */
vx_status vx_light_pass(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* gbuf,
                        const float* palette_dev, const vx_light* lights_dev, uint32_t nlights,
                        const vx_light_params* params, float* rgb_dev, uint32_t* fetches_dev, void* stream);

/* LightTAA inputs and output (device; width*height pixels, rows from the top). */
typedef struct {
    const float* light;        /* 3 per pixel: this frame's light buffer (ambient + lights) */
    const float* last;         /* 4 per pixel: last frame's TAA output (rgb, variance); 16-byte aligned */
    const float* motion;       /* 2 per pixel (vx_gbuf_ext.motion); NULL = no motion */
    const float* depth;        /* 1 per pixel (vx_gbuf_ext.depth, 1 = sky) */
    const float* normal;       /* 3 per pixel (vx_gbuf.normal) */
    const uint8_t* material;   /* vx_gbuf.material: colour and material from the palette */
    const float* palette;      /* 256 x 6 */
    float* out;                /* 4 per pixel: rgb, variance; must not alias last */
} vx_taa_io;

/* Temporal accumulation of the light buffer (LightTAAPipeline::Use,
 * LightTAAPipeline.h:35-51; LightTAA.frag:37-141): a 12-sample spiral of
 * same-surface neighbours, the last frame's light (at uv + motion) clipped to
 * the neighbourhood's range and blended by the running variance; pixels whose
 * last position left the screen average a 19x19 same-surface neighbourhood.
 * Uses view width/height and frame; seed keys the per-pixel spiral angle. */
vx_status vx_light_taa(vx_ctx* ctx, const vx_view* view, const vx_taa_io* io, uint32_t seed, void* stream);

/* ---------------------------------------------------------------- reflection pass
 * LightReflectionPipeline::Use(cmd, viewBuffer, lightDiffuse, geometryFB,
 * shadowVox, skybox, blueNoise) (LightReflectionPipeline.h:34) with
 * LightReflection.frag:60-139. Per pixel whose depth is below 0.999: the
 * reflection ray (reflect(-V, N) mixed with a cosine sample by roughness *
 * 0.1 and scaled by 1 + noise.w * 0.5, from the hit + normal * noise.w) is
 * marched to maxt; on a miss (t == maxt) the sky is taken in the ray's
 * direction, otherwise the hit is reprojected and, when its linear depth lies
 * within +-0.001 of the depth attachment there, the light buffer is taken
 * there. Output (ambient * F * (1 - roughness), F) with F =
 * fresnelSchlickRoughness(max(N.V, 0), mix(0.04, 1, metallic), roughness) * 5
 * (albedo 1, LightReflection.frag:74-86). The light and depth images are
 * sampled bilinearly with 8-bit sub-texel weights and an opaque-black border.
 * Noise: the pixel's seeded counter hash (the blue-noise texel's stand-in). */
typedef struct {
    int32_t march;      /* reference: VX_MARCH_SPARSE (LightReflection.frag:113) */
    float maxt;         /* reference: 256 */
    uint32_t seed;
    int32_t frame;
} vx_reflection_params;

/* The skybox (SkyBoxAsset's cube image): faces +X, -X, +Y, -Y, +Z, -Z (Vulkan
 * layer order), each size x size RGBA fp32 texels, rows from t = 0 (device).
 * A lookup selects the face of the direction's major axis (x, then y, then z
 * on ties) and filters bilinearly within the face, the texel position
 * clamped to it. faces NULL: the procedural getSkyColor of
 * LightAmbient.frag:128-131, which LightReflection.frag:55-56 keeps as a
 * comment beside its skybox lookup. */
typedef struct {
    const float* faces;
    int32_t size;
    int32_t pad;
} vx_skybox;

typedef struct {
    const float* light;   /* lightDiffuse: 4 floats per pixel (the TAA light buffer), 16-byte aligned */
    const float* depth;   /* the depth attachment as read back (vx_gbuf_ext.depth or vx_gbuf_depth), 1 = sky */
    float* out;           /* 4 floats per pixel (the reflection buffer), 16-byte aligned */
    float* t;             /* optional: the ray's marched distance per pixel, NaN where no ray was marched */
} vx_reflection_io;

vx_status vx_reflection(vx_ctx* ctx, const vx_volume* vol, const vx_view* view, const vx_gbuf* gbuf,
                        const float* palette_dev, const vx_reflection_io* io, const vx_skybox* sky,
                        const vx_reflection_params* params, void* stream);

/* The depth attachment of a G-buffer written by vx_primary (which writes no
 * depth): per pixel the geometry pass's fragment depth (1 - 1e-7 volume_rid)
 * (w - NEAR) / (FAR - NEAR) of the hit (GeometryVoxel.frag:167-170), quantized
 * to D24 and read back as q / (2^24 - 1), 1 where nothing was hit; the same
 * values vx_primary_instances writes to vx_gbuf_ext.depth. depth_dev: npx. */
vx_status vx_gbuf_depth(vx_ctx* ctx, const vx_view* view, const vx_gbuf* gbuf, int32_t volume_rid, float* depth_dev,
                        void* stream);

/* Screen-space occlusion of the ambient pass: calculateOcclusion(N)
 * (LightAmbient.frag:89-109, with screenspaceOcclusion :59-79), the factor of
 * the ambient term that vx_shade_forward takes as ssao_dev. Per pixel with a
 * hit and depth below 0.999: the view-space position rebuilt from the depth
 * attachment (FarVec * (depth * (1 + 1/FAR) + NEAR/FAR)), four samples within
 * 0.2 world units in the cosine hemisphere of the view-space normal, each
 * occluded where the depth attachment at its projection lies within
 * 0.2 / FAR in front of it (weight smoothstep(max, min, depth) * distance);
 * ssao = clamp(1 - 3 * mean, 0, 1); 1 elsewhere. Noise: the pixel's seeded
 * counter hash (the blue-noise texel's stand-in). depth_dev: the depth
 * attachment (vx_gbuf_ext.depth or vx_gbuf_depth); ssao_dev: npx floats. */
vx_status vx_ssao(vx_ctx* ctx, const vx_view* view, const vx_gbuf* gbuf, const float* depth_dev, uint32_t seed,
                  float* ssao_dev, void* stream);

/* One view of a batched adjoint: its G-buffer, fixed visibility, optional
 * SSAO and the loss gradient w.r.t. its shaded image (all device). */
typedef struct {
    vx_view view;
    vx_gbuf gbuf;
    vx_trace_out vis;
    const float* ssao;
    const float* dl_drgb;
} vx_shade_item;

/* Adjoint over n views sharing one palette and light, gradients accumulated
 * across the views in one launch (the views of a scene rendered from several
 * cameras; no counterpart in the reference, which has no adjoint). Same
 * per-pixel partials as vx_shade_backward, summed exactly in 2^-30 fixed
 * point, so the result does not depend on the order of views or pixels. */
vx_status vx_shade_backward_batch(vx_ctx* ctx, const vx_shade_item* items, uint32_t n,
                                  const float* palette_dev, const float light[6], float* dpalette_dev,
                                  float* dlight_dev, void* stream);

/* ---------------------------------------------------------------- asset files (host)
 * Readers for the reference's on-disk formats. They read files and host
 * memory only; vx_model_load also uploads the model it reads. */

/* AssetGUID of a mod-relative path such as "default/FarmHouse/0.v": the hash
 * of the path string (Assets::Hash, Assets.h:207-210) over the identifiers
 * ModFolder maps (paths relative to Assets/Mods, Mod.h:86-94). The
 * reference's files were written by a standard library whose string hash is
 * 64-bit FNV-1a; this is that hash. */
uint64_t vx_asset_guid(const char* path, size_t len);

/* .v model (VoxAsset::Serialize, VoxAsset.h:42-50): u32 SizeX, SizeY, SizeZ
 * (little-endian), then SizeX*SizeY*SizeZ material bytes, X fastest. Checks
 * the header against nbytes (sizes 1..4096) and returns the dims and the
 * offset of the voxel bytes (12). */
vx_status vx_vox_parse(const uint8_t* bytes, size_t nbytes, int32_t dims[3], size_t* data_offset);
/* Reads a .v file into a new model, uploads mip 0 and builds the mips
 * (VoxAsset::OnLoad, VoxAsset.h:36-39). Sizes must be multiples of 4. */
vx_status vx_model_load(vx_ctx* ctx, const char* path, vx_model** out, void* stream);

/* .p palette (PalleteAsset::Serialize, PalleteAsset.h:63-69): 256 VoxMaterial
 * of 7 bytes {r, g, b, a, roughness, metallic, emit}. Writes the 256 x 6 fp32
 * palette the passes read, {albedo rgb, roughness, metallic, emissive}, each
 * value byte/255 as the RGBA8 UNORM palette textures of
 * PalleteCache::UploadPallete (PalleteCache.cpp:5-23) read. Host memory. */
vx_status vx_palette_parse(const uint8_t* bytes, size_t nbytes, float* palette);
vx_status vx_palette_load(const char* path, float* palette);

/* .pf prefab (PrefabAsset::Spawn, PrefabAsset.cpp:30-142) spawned into a flat
 * entity list, nested Instance prefabs expanded where they occur. GUIDs
 * resolve through every file under mods_dir (the Assets/Mods folder); path
 * is mod-relative ("default/FarmHouse.pf"). */
typedef struct vx_prefab vx_prefab;
enum { VX_ENT_TRANSFORM = 1, VX_ENT_VOX = 2, VX_ENT_LIGHT = 4, VX_ENT_INSTANCE = 8 };

/* One spawned entity (Components.h Transform / VoxRenderer / Light). Matrices
 * are column-major: matrix is Transform::Matrix, world the WorldMatrix
 * TransformSystem computes (TransformSystem.cpp:36-70,124-135) in the
 * reference's fp32 operation order. 256 bytes. */
typedef struct {
    int32_t id;                /* "Id" in the file that made the entity */
    int32_t parent;            /* index of the parent in the list, -1 at the root */
    int32_t flags;             /* VX_ENT_* */
    int32_t model;             /* VoxRenderer.Vox as a file index, -1 if absent or not found */
    int32_t palette;           /* VoxRenderer.Pallete as a file index, -1 if absent or not found */
    int32_t light_type;        /* Light::Type (0 point, 1 spot, 2 directional, 3 ambient), -1 */
    uint64_t vox_guid;
    uint64_t palette_guid;
    float position[3];
    float rotation[3];         /* radians, applied Z, then Y, then X */
    float scale[3];
    float pivot[3];
    float matrix[16];
    float world[16];
    float light_color[3];
    float light_intensity;
    float light_attenuation;
    float light_range;
    float light_angle;
    float light_angle_attenuation;
    int32_t pad[2];
} vx_prefab_entity;

vx_status vx_prefab_load(const char* mods_dir, const char* path, vx_prefab** out);
vx_status vx_prefab_destroy(vx_prefab* p);
/* Number of entities and of distinct referenced asset files (.v and .p). */
vx_status vx_prefab_counts(const vx_prefab* p, uint32_t* nentities, uint32_t* nfiles);
/* Entities in spawn order (the registry's creation order). cap >= nentities. */
vx_status vx_prefab_entities(const vx_prefab* p, vx_prefab_entity* out, uint32_t cap);
/* Full path (mods_dir + "/" + mod-relative path) of referenced file i. */
vx_status vx_prefab_file(const vx_prefab* p, uint32_t i, char* buf, size_t cap);
/* Entity indices of the VoxRenderer entities in the order
 * ShadowVoxSystem::OnUpdate visits them on the frame after the spawn
 * (ShadowVoxSystem.cpp:116-120): EnTT 3.6 views walk the smallest pool,
 * here VoxRenderer's, from its newest element to its oldest. *n is the
 * count even when it exceeds cap (then VX_ERR_RANGE). */
vx_status vx_prefab_rebuild_order(const vx_prefab* p, uint32_t* idx, uint32_t cap, uint32_t* n);
/* Point and spot lights as vx_light in the order the light passes draw them
 * (WorldRenderer.cpp:243-257): the point pipeline's draws, then the spot
 * pipeline's, each walking the Light pool newest first. position is
 * WorldMatrix[3], direction WorldMatrix[2], color Color * Intensity. */
vx_status vx_prefab_lights(const vx_prefab* p, vx_light* out, uint32_t cap, uint32_t* n);
/* The geometry pass's draws for the spawned prefab (WorldRenderer.cpp:210-220):
 * entities with a resolved Vox and Pallete, in the owning group's order
 * (newest VoxRenderer first). world = WorldMatrix translated by -Pivot;
 * last_world = PreviousWorldMatrix (prev[16*entity], or the identity a spawned
 * entity starts with when prev is NULL) translated likewise. models[i] is the
 * loaded model of file i (NULL for other files); volume_rid = file index + 1,
 * palette = the palette's file index. *n is the count even above cap. */
vx_status vx_prefab_vox_cmds(const vx_prefab* p, const vx_model* const* models, uint32_t nmodels,
                             const float* prev, vx_vox_cmd* out, uint32_t cap, uint32_t* n);

#ifdef __cplusplus
}
#endif
#endif
