/*
 * vx_oracle.h -- CPU restatement of VoxelEngine's voxel-traversal lighting path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product in
 * voxelengine_amd/. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the CPU side. The
 * product never links, loads or calls it.
 *
 * Every function states the reference file:line it restates (paths relative to
 * the VoxelEngine source tree). Arithmetic is IEEE fp32 in the reference's
 * evaluation order, compiled with -ffp-contract=off and without fast-math, so
 * each float op is a single correctly rounded +,-,*,/ or sqrt.
 *
 * GLSL builtins and reference conventions pinned here (the reference leaves
 * them to the Vulkan driver):
 *   sign(+-0) = +0 and sign(NaN) = 0;  step(e,x) = x < e ? 0 : 1;
 *   min(x,y) = y < x ? y : x;  max(x,y) = x < y ? y : x;
 *   mod(x,y) = x - y*floor(x/y);  ivec3(v) truncates toward zero, saturating
 *   out-of-range values and mapping NaN to 0;  integer '/' truncates;
 *   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z;  normalize(v) = v / sqrt(dot(v,v));
 *   out-of-range texelFetch returns 0 (empty);  pow(x,5) = x*x*x*x*x;
 *   sin/cos use the build's own Cody-Waite reduction and Cephes polynomials.
 */
#ifndef VX_ORACLE_H
#define VX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VXO_MARCH_DDA = 0, VXO_MARCH_SPARSE = 1, VXO_MARCH_SUPERSPARSE = 2 };
enum { VXO_INST_UPDATE = 1, VXO_INST_DESTROY = 2 };

/* Occupancy ("shadow vox") volume in the reference byte layout: texel dims
 * sx,sy,sz, X-fastest, each byte a 2x2x2 fine-voxel bitmask
 * (ShadowVoxSystem.cpp:56-63,85,92). */
typedef struct {
    const uint8_t* data;
    int32_t sx, sy, sz;
} vxo_occ;

/* Model instance for the rebuild (ShadowVoxSystem.cpp:116-201; Components.h:30-35,60-67). */
typedef struct {
    const uint8_t* model;     /* material volume, X-fastest (VoxAsset.h:52-56) */
    int32_t sx, sy, sz;       /* model dims */
    float prev[16];           /* PreviousWorldMatrix, column-major glm mat4 */
    float cur[16];            /* WorldMatrix */
    float pivot[3];           /* VoxRenderer::Pivot */
    int32_t flags;            /* VXO_INST_UPDATE (OnUpdate) or VXO_INST_DESTROY (OnVoxDestroyed) */
} vxo_instance;

/* Per-ray record for explicit-ray parity. */
typedef struct {
    int32_t hit;        /* 1 hit, 0 miss */
    int32_t term;       /* DDA: 0 loop exhausted (x4 restart), 1 hit, 2 out of bounds. Sparse: 0 miss, 1 hit */
    int32_t steps;      /* DDA: nt (reference step count); sparse: loop iterations */
    int32_t fetches;    /* occupancy-byte fetches the reference performs for this ray */
    int32_t voxel[3];   /* DDA: fine voxel at the hit */
    float t;            /* DDA: best_t at the hit; sparse: returned d */
    float hitpos[3];    /* DDA: origin + dir*best_t */
    float normal[3];    /* DDA: -sign(dir)*select */
} vxo_ray_result;

/* View: the reference ViewData (View.h:16-31) plus the pixel grid. Matrices
 * column-major. World units are the reference's (1 unit = 10 fine voxels). */
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
} vxo_view;

typedef struct {
    int32_t shadow_march, ao_march, spec_march;   /* VXO_MARCH_* */
    int32_t n_shadow, n_ao, n_spec;               /* rays per pixel per class */
    float shadow_maxt, ao_maxt, spec_maxt;
    float light_dir[3];                           /* SUN_DIR, LightAmbient.frag:15 */
    uint32_t seed;
    int32_t frame;
    int32_t jitter_shadow;                        /* 1: LightAmbient.frag:151-156 cone jitter */
    int32_t pad;
} vxo_trace_params;

/* ---- scalar building blocks (exported for unit tests) ---- */
int32_t vxo_f2i(float x);
float vxo_sign(float x);
float vxo_mod(float x, float y);
void vxo_sincos(float x, float* s, float* c);
uint32_t vxo_hash(uint32_t x);
float vxo_rand(uint32_t pixkey, uint32_t idx);
uint32_t vxo_pixkey(uint32_t seed, int32_t frame, uint32_t pixel);
void vxo_cosine_hemisphere(float u1, float u2, float out[3]);

/* ---- rebuild (R1-R4) ---- */
void vxo_set_volume_at(uint8_t* vol, int32_t sx, int32_t sy, int32_t sz,
                       int32_t x, int32_t y, int32_t z, int32_t value);
void vxo_translate(const float m[16], const float v[3], float out[16]);
int32_t vxo_rebuild(uint8_t* vol, int32_t sx, int32_t sy, int32_t sz,
                    const vxo_instance* inst, int32_t n, int32_t occ_threshold,
                    int32_t* regions /* n*6: x,y,z,w,h,d (w,h,d as uint32) */);

/* ---- model mips (R6) ---- */
void vxo_build_mip(const uint8_t* parent, int32_t px, int32_t py, int32_t pz, uint8_t* out);

/* ---- traversal (R7-R11) ---- */
int32_t vxo_get_volume_at(const vxo_occ* v, int32_t x, int32_t y, int32_t z, int32_t mip);
int32_t vxo_dda(const vxo_occ* v, const float o[3], const float d[3], float maxt, vxo_ray_result* r);
float vxo_sparse(const vxo_occ* v, const float o[3], const float d[3], float dist, vxo_ray_result* r);
float vxo_supersparse(const vxo_occ* v, const float o[3], const float d[3], float dist, vxo_ray_result* r);
void vxo_trace_rays(const vxo_occ* v, const float* rays /* n*7: o3,d3,maxt */, int32_t n,
                    int32_t march, vxo_ray_result* out, int32_t nthreads);
void vxo_clip_to_aabb(const float o[3], const float d[3], const float size[3], float out[3]);
int32_t vxo_intersect_volume(const uint8_t* const mips[3], const int32_t dims[3],
                             const float o[3], const float d[3], const float cam_local[3],
                             const float uv[2], const float res[2], int32_t frame,
                             float hit[3], float normal[3], uint8_t* material);

/* ---- full passes ---- */
void vxo_primary(const uint8_t* const mips[3], const int32_t dims[3], const vxo_view* view,
                 float* pos, float* nrm, uint8_t* mat, int32_t nthreads);
/* vxo_primary for the listed pixels only; outputs are compact (entry j for
 * pixel pixels[j]). */
void vxo_primary_pixels(const uint8_t* const mips[3], const int32_t dims[3], const vxo_view* view,
                        const int32_t* pixels, int32_t npixels, float* pos, float* nrm, uint8_t* mat,
                        int32_t nthreads);
/* One voxel draw (GeometryVoxelPipeline::Cmd, GeometryVoxelPipeline.h:29-35)
 * with its model's mips. */
typedef struct {
    const uint8_t* mips[3];
    int32_t dims[3];
    float world[16];
    float last_world[16];
    int32_t rid;
    int32_t palette;
} vxo_vox_cmd;
/* The geometry pass over n draws (GeometryVoxel.vert:62-78, GeometryVoxel.frag:
 * 127-183, D24 LESS depth test in draw order); every draw is tried at every
 * pixel. motion / depth / palette may be NULL. */
void vxo_primary_instances(const vxo_vox_cmd* cmds, int32_t n, const vxo_view* view, float* pos, float* nrm,
                           uint8_t* mat, float* motion, float* depth, uint16_t* palette, int32_t nthreads);
/* LightTAA (LightTAA.frag:37-141) over width x height pixels. light: 3 per
 * pixel, last / out: 4, motion: 2 or NULL, depth: 1, normal: 3, mat: index,
 * palette: 256 x 6. */
void vxo_light_taa(int32_t width, int32_t height, int32_t frame, uint32_t seed, const float* light,
                   const float* last, const float* motion, const float* depth, const float* normal,
                   const uint8_t* mat, const float* palette, float* out, int32_t nthreads);
/* The TAA sampler: bilinear, 8-bit sub-texel weights, opaque-black border,
 * over an n-channel float image (test hook). */
void vxo_taa_sample(const float* img, int32_t width, int32_t height, int32_t n, float u, float v, float* out);
/* The rays vxo_trace marches for each listed pixel, in ray order (n_shadow
 * shadow, n_ao AO, n_spec spec): npixels * nrays * 7 floats (origin3, dir3,
 * maxt); all zero for a sky pixel (material 0). */
void vxo_pixel_rays(const vxo_view* view, const vxo_trace_params* p, const float* pos, const float* nrm,
                    const uint8_t* mat, const float* palette, const int32_t* pixels, int32_t npixels,
                    float* rays);
/* pixels: optional list of pixel indices to process (NULL = all). */
void vxo_trace(const vxo_occ* v, const vxo_view* view, const vxo_trace_params* p,
               const float* pos, const float* nrm, const uint8_t* mat, const float* palette,
               uint8_t* shadow, float* ao, float* spec, uint32_t* fetches,
               const int32_t* pixels, int32_t npixels, int32_t nthreads);
void vxo_shade_f32(const vxo_view* view, const float* pos, const float* nrm, const uint8_t* mat,
                   const uint8_t* shadow, const float* ao, const float* ssao,
                   const float* palette, const float light[6], float* rgb, int32_t nthreads);
void vxo_shade_f64(const vxo_view* view, const float* pos, const float* nrm, const uint8_t* mat,
                   const uint8_t* shadow, const float* ao, const float* ssao,
                   const double* palette, const double light[6], double* rgb, int32_t nthreads);
void vxo_shade_backward_f64(const vxo_view* view, const float* pos, const float* nrm,
                            const uint8_t* mat, const uint8_t* shadow, const float* ao,
                            const float* ssao, const double* palette, const double light[6],
                            const double* dl_drgb, double* dpalette, double dlight[6]);

/* ---- point / spot lights (LightPoint.frag:85-152, LightSpot.frag:73-140) ---- */
/* pow(x, y) of the attenuation terms: exp2(y * log2(x)) from the build's own
 * log2 / exp2 (range reduction + Cephes single-precision polynomials, only
 * + - * /), replacing the driver's pow. Defined for x >= 0. */
float vxo_log2(float x);
float vxo_exp2(float t);
float vxo_pow(float x, float y);

enum { VXO_LIGHT_POINT = 0, VXO_LIGHT_SPOT = 1 };

/* PointLight / SpotLight (Common.frag:15-30); direction = the light's
 * WorldMatrix[2] (WorldRenderer.cpp:255), color = Color * Intensity. */
typedef struct {
    int32_t type;
    float position[3];
    float range;
    float color[3];
    float attenuation;
    float direction[3];
    float angle;
    float angle_attenuation;
    int32_t pad;
} vxo_light;

typedef struct {
    int32_t point_march;    /* reference: SPARSE (LightPoint.frag:125) */
    int32_t spot_march;     /* reference: SUPERSPARSE (LightSpot.frag:113) */
    uint32_t seed;
    int32_t frame;
} vxo_light_params;

/* Adds each light's direct term to rgb (npx*3), lights in array order. */
void vxo_light_pass(const vxo_occ* v, const vxo_view* view, const vxo_light* lights, int32_t nlights,
                    const vxo_light_params* p, const float* pos, const float* nrm, const uint8_t* mat,
                    const float* palette, float* rgb, uint32_t* fetches, int32_t nthreads);

/* ---- reflection pass (LightReflection.frag:60-139) ---- */
typedef struct {
    int32_t march;      /* reference: SPARSE (LightReflection.frag:113) */
    float maxt;         /* reference: 256 */
    uint32_t seed;
    int32_t frame;
} vxo_reflection_params;

/* getSkyColor (LightAmbient.frag:128-131) of direction e. */
void vxo_sky_procedural(const float e[3], float out[3]);
/* Cube-map lookup: faces +X, -X, +Y, -Y, +Z, -Z of size x size RGBA. */
void vxo_cube_sample(const float* faces, int32_t size, const float d[3], float out[4]);
/* The reflection ray of each listed pixel: wcp (fine units) and direction,
 * 6 floats; the march starts at wcp + normal. */
void vxo_reflection_rays(const vxo_view* view, const vxo_reflection_params* p, const float* pos, const float* nrm,
                         const uint8_t* mat, const float* palette, const int32_t* pixels, int32_t npixels,
                         float* rays);
/* out: 4 floats per pixel, (ambient * F * (1 - roughness), F). depth: the D24
 * depth per pixel (1 = sky); light: 4 floats per pixel; sky: NULL for the
 * procedural sky. t_out (optional): the marched distance, NaN where no ray. */
void vxo_reflection(const vxo_occ* v, const vxo_view* view, const vxo_reflection_params* p, const float* pos,
                    const float* nrm, const uint8_t* mat, const float* palette, const float* depth,
                    const float* light, const float* sky, int32_t sky_size, float* out, float* t_out,
                    int32_t nthreads);

/* ---- screen-space occlusion (LightAmbient.frag:59-109) ---- */
/* calculateOcclusion(N) per pixel with depth < 0.999 and a hit (1 elsewhere):
 * 4 samples within 0.2 world units of the view-space position rebuilt from
 * the depth attachment, occluded where the depth attachment at the sample's
 * projection lies within 0.2 / FAR in front of it; 1 - 3 * mean, clamped. */
void vxo_ssao(const vxo_view* view, const float* nrm, const uint8_t* mat, const float* depth, uint32_t seed,
              float* out, int32_t nthreads);

/* ---- synthetic scenes (inputs, not reference code) ---- */
void vxo_gen_terrain(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed,
                     float fill, int32_t nthreads);
void vxo_gen_boxes(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed,
                   int32_t nboxes, int32_t emin, int32_t emax, int32_t grounded);

#ifdef __cplusplus
}
#endif
#endif
