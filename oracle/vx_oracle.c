/*
 * vx_oracle.c -- CPU restatement of VoxelEngine's voxel-traversal lighting path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg. Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Reference paths below are relative to the VoxelEngine source tree.
 * Conventions for GLSL builtins and undefined behaviour are listed in
 * vx_oracle.h; the build-defined parts (RNG, sin/cos, scene generators, ray
 * set-up deviations) are marked "build-defined".
 */
#include "vx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(__FAST_MATH__)
#error "the oracle must be built without fast-math"
#endif

static const float NEAR_PLANE = 0.1f;      /* Common.frag:12 */
static const float FAR_PLANE = 4096.0f;    /* Common.frag:13 */

/* ------------------------------------------------------------------------ */
/* GLSL / glm scalar semantics                                              */
/* ------------------------------------------------------------------------ */

/* ivec3(v) / glm::ivec3(v): truncation toward zero; out-of-range saturates and
 * NaN gives 0 (the reference leaves these undefined). */
int32_t vxo_f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x < -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

/* GLSL sign(): +0 for +-0 and for NaN. */
float vxo_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* GLSL step(edge, x) = x < edge ? 0 : 1. */
static inline float g_step(float e, float x) { return x < e ? 0.0f : 1.0f; }
/* GLSL min/max as specified: min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x. */
static inline float g_min(float x, float y) { return y < x ? y : x; }
static inline float g_max(float x, float y) { return x < y ? y : x; }
static inline double g_maxd(double x, double y) { return x < y ? y : x; }
static inline float g_clamp(float x, float lo, float hi) { return g_min(g_max(x, lo), hi); }

/* GLSL mod(x,y) = x - y*floor(x/y). */
float vxo_mod(float x, float y) { return x - y * floorf(x / y); }

static inline float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline double dot3d(const double a[3], const double b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void normalize3(const float v[3], float out[3]) {
    float l = sqrtf(dot3(v, v));
    out[0] = v[0] / l; out[1] = v[1] / l; out[2] = v[2] / l;
}
static inline void normalize3d(const double v[3], double out[3]) {
    double l = sqrt(dot3d(v, v));
    out[0] = v[0] / l; out[1] = v[1] / l; out[2] = v[2] / l;
}

/* ------------------------------------------------------------------------ */
/* Build-defined sin/cos: Cody-Waite reduction by pi/2 with a three-part     */
/* constant and the Cephes (S. L. Moshier) single-precision polynomials for  */
/* [-pi/4, pi/4]. Only +,-,* so CPU and GPU evaluate identical bits.         */
/* ------------------------------------------------------------------------ */
void vxo_sincos(float x, float* s, float* c) {
    const float two_over_pi = 0.636619772367581343f;
    const float c1 = 1.5703125f;                       /* 201/128 */
    const float c2 = 4.837512969970703125e-4f;
    const float c3 = 7.54978995489188216e-8f;
    float kf = floorf(x * two_over_pi + 0.5f);
    int32_t k = vxo_f2i(kf);
    float r = ((x - kf * c1) - kf * c2) - kf * c3;
    float z = r * r;
    float sp = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float cp = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
               - 0.5f * z + 1.0f;
    switch (k & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

/* ------------------------------------------------------------------------ */
/* Build-defined counter-hash RNG (replaces the blue-noise texture of        */
/* LightAmbient.frag:44-52). lowbias32 integer hash.                          */
/* ------------------------------------------------------------------------ */
uint32_t vxo_hash(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
uint32_t vxo_pixkey(uint32_t seed, int32_t frame, uint32_t pixel) {
    uint32_t key = vxo_hash(seed ^ vxo_hash((uint32_t)frame + 0x9e3779b9u));
    return vxo_hash(key ^ vxo_hash(pixel));
}
float vxo_rand(uint32_t pixkey, uint32_t idx) {
    return (float)(vxo_hash(pixkey + idx * 0x9e3779b9u) >> 8) * (1.0f / 16777216.0f);
}

/* cosineSampleHemisphere, LightAmbient.frag:81-87 (sin/cos build-defined). */
void vxo_cosine_hemisphere(float u1, float u2, float out[3]) {
    float r = sqrtf(u1);
    float theta = 6.283f * u2;
    float sn, cs;
    vxo_sincos(theta, &sn, &cs);
    out[0] = r * cs;
    out[1] = r * sn;
    out[2] = sqrtf(g_max(0.0f, 1.0f - u1));
}

/* ------------------------------------------------------------------------ */
/* Rebuild: ShadowVoxSystem                                                  */
/* ------------------------------------------------------------------------ */

/* ShadowVoxSystem::SetVolumeAt, ShadowVoxSystem.cpp:82-94. Texel dims sx,sy,sz. */
void vxo_set_volume_at(uint8_t* vol, int32_t sx, int32_t sy, int32_t sz,
                       int32_t x, int32_t y, int32_t z, int32_t value) {
    if (x < 0 || y < 0 || z < 0 || x >= sx * 2 || y >= sy * 2 || z >= sz * 2) return;
    int32_t bit = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    int32_t mask = 1 << bit;
    x /= 2; y /= 2; z /= 2;
    uint8_t* d = vol + ((size_t)x + (size_t)y * (size_t)sx + (size_t)z * (size_t)sx * (size_t)sy);
    *d = (uint8_t)((*d & ~mask) | (value << bit));
}

/* glm::translate(m, v): Result[3] = m[0]*v[0] + m[1]*v[1] + m[2]*v[2] + m[3]
 * (Vendor/glm/ext/matrix_transform.inl:10-15), evaluated left to right. */
void vxo_translate(const float m[16], const float v[3], float out[16]) {
    memcpy(out, m, 16 * sizeof(float));
    for (int r = 0; r < 4; ++r) {
        float a = m[0 * 4 + r] * v[0];
        float b = m[1 * 4 + r] * v[1];
        float c = m[2 * 4 + r] * v[2];
        out[3 * 4 + r] = ((a + b) + c) + m[3 * 4 + r];
    }
}

typedef struct { int32_t mn[3], mx[3]; } aabb_i;

/* One footprint loop of ShadowVoxSystem::OnUpdate (ShadowVoxSystem.cpp:137-155
 * and 159-177) or OnVoxDestroyed (:22-40): for every model voxel with material
 * >= threshold, wp = ((o + x*dx) + y*dy) + z*dz, fwp = ivec3(wp*10). */
static void footprint(uint8_t* vol, int32_t sx, int32_t sy, int32_t sz, const vxo_instance* in,
                      const float m[16], int32_t value, int32_t thr, aabb_i* box) {
    float o[3], dx[3], dy[3], dz[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = m[12 + k];
        dx[k] = m[0 + k] * 0.1f;
        dy[k] = m[4 + k] * 0.1f;
        dz[k] = m[8 + k] * 0.1f;
    }
    for (int32_t z = 0; z < in->sz; ++z)
        for (int32_t y = 0; y < in->sy; ++y)
            for (int32_t x = 0; x < in->sx; ++x) {
                uint8_t mat = in->model[(size_t)x + (size_t)y * in->sx + (size_t)z * in->sx * in->sy];
                if (mat < thr) continue;
                int32_t f[3];
                for (int k = 0; k < 3; ++k) {
                    float wp = ((o[k] + (float)x * dx[k]) + (float)y * dy[k]) + (float)z * dz[k];
                    f[k] = vxo_f2i(wp * 10.0f);
                    if (f[k] < box->mn[k]) box->mn[k] = f[k];
                    if (f[k] > box->mx[k]) box->mx[k] = f[k];
                }
                vxo_set_volume_at(vol, sx, sy, sz, f[0], f[1], f[2], value);
            }
}

/* ShadowVoxSystem::OnUpdate (ShadowVoxSystem.cpp:116-201) for VXO_INST_UPDATE,
 * OnVoxDestroyed (:7-53) for VXO_INST_DESTROY. Instances run in order; per
 * instance the previous footprint is erased before the current one is written.
 * Returns the number of dirty regions written to regions (6 ints each, the
 * ImageRegion of :180-190 with extents stored as uint32 bit patterns). */
int32_t vxo_rebuild(uint8_t* vol, int32_t sx, int32_t sy, int32_t sz,
                    const vxo_instance* inst, int32_t n, int32_t thr, int32_t* regions) {
    int32_t nreg = 0;
    const float zero_pivot[3] = {0.0f, 0.0f, 0.0f};
    (void)zero_pivot;
    for (int32_t i = 0; i < n; ++i) {
        const vxo_instance* in = &inst[i];
        if (!in->model) continue;                                  /* :120 / :9 */
        aabb_i box;
        box.mn[0] = sx - 1; box.mn[1] = sy - 1; box.mn[2] = sz - 1; /* startmin, :128 */
        box.mx[0] = box.mx[1] = box.mx[2] = 0;                     /* startmax, :129 */
        if (in->flags == VXO_INST_DESTROY) {
            footprint(vol, sx, sy, sz, in, in->cur, 0, thr, &box);  /* WorldMatrix, no pivot :22-25 */
        } else {
            float npiv[3] = {-in->pivot[0], -in->pivot[1], -in->pivot[2]};
            float m[16];
            vxo_translate(in->prev, npiv, m);                       /* :134-135 */
            footprint(vol, sx, sy, sz, in, m, 0, thr, &box);
            vxo_translate(in->cur, npiv, m);                        /* :157-158 */
            footprint(vol, sx, sy, sz, in, m, 1, thr, &box);
        }
        if (box.mx[0] != 0 || box.mx[1] != 0 || box.mx[2] != 0) {  /* :181 */
            int32_t mn[3], mx[3];
            const int32_t smin[3] = {sx - 1, sy - 1, sz - 1};
            for (int k = 0; k < 3; ++k) {
                mn[k] = box.mn[k] / 2;
                mx[k] = box.mx[k] / 2;
                if (mn[k] < 0) mn[k] = 0;
                if (mx[k] > smin[k]) mx[k] = smin[k];
            }
            int32_t* r = regions + 6 * nreg++;
            r[0] = mn[0]; r[1] = mn[1]; r[2] = mn[2];
            for (int k = 0; k < 3; ++k) r[3 + k] = (int32_t)(uint32_t)(mx[k] - mn[k] + 1);
        }
    }
    return nreg;
}

/* ------------------------------------------------------------------------ */
/* Model mips: VoxAsset::Upload, VoxAsset.cpp:26-47. The child scan runs      */
/* vi = 0..8 (vi = 8 re-reads child 0) and keeps the first non-zero value.   */
/* ------------------------------------------------------------------------ */
void vxo_build_mip(const uint8_t* parent, int32_t px, int32_t py, int32_t pz, uint8_t* out) {
    int32_t mx = px > 1 ? px / 2 : 1, my = py > 1 ? py / 2 : 1, mz = pz > 1 ? pz / 2 : 1;
    for (int32_t x = 0; x < mx; ++x)
        for (int32_t y = 0; y < my; ++y)
            for (int32_t z = 0; z < mz; ++z) {
                uint8_t vox = 0;
                for (int vi = 0; vi < 9; ++vi) {
                    uint8_t u = parent[(size_t)(2 * x + ((vi >> 0) & 1)) +
                                       (size_t)(2 * y + ((vi >> 1) & 1)) * px +
                                       (size_t)(2 * z + ((vi >> 2) & 1)) * px * py];
                    if (u) { vox = u; break; }
                }
                out[(size_t)x + (size_t)y * mx + (size_t)z * mx * my] = vox;
            }
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                 */
/* ------------------------------------------------------------------------ */

static inline uint32_t occ_texel(const vxo_occ* v, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= v->sx || y >= v->sy || z >= v->sz) return 0;
    return v->data[(size_t)x + (size_t)y * v->sx + (size_t)z * v->sx * v->sy];
}

/* getVolumeAt, Light.frag:14-27. */
int32_t vxo_get_volume_at(const vxo_occ* v, int32_t x, int32_t y, int32_t z, int32_t mip) {
    int32_t bit = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    int32_t mask = 1 << bit;
    uint32_t vox = occ_texel(v, x / 2, y / 2, z / 2);   /* mip/2 == 0: the volume has no mips */
    if (mip % 2 == 1) return vox != 0;
    return (vox & (uint32_t)mask) != 0;
}

/* raycastShadowVolume, Light.frag:29-81, with mip starting (and staying) at 0.
 * The do{}while(++i<4) outer loop re-traces from the same origin with the same
 * mip whenever the inner loop ends by n==256 or totalt>=maxt, so passes 2..4
 * repeat pass 1 exactly; the step count nt keeps counting (Light.frag:38,74). */
int32_t vxo_dda(const vxo_occ* v, const float o[3], const float d[3], float maxt, vxo_ray_result* r) {
    float s[3], td[3], tm[3];
    int32_t cv[3], up[3];
    const int32_t dims[3] = {v->sx, v->sy, v->sz};
    for (int k = 0; k < 3; ++k) {
        s[k] = vxo_sign(d[k]);
        td[k] = 1.0f / (d[k] * s[k]);
        cv[k] = vxo_f2i(floorf(o[k]));
        float nb = (float)cv[k] + (s[k] * 0.5f + 0.5f);
        tm[k] = (nb - o[k]) / d[k];
        up[k] = vxo_f2i((float)dims[k] / 1.0f) * 2 + 1;
    }
    memset(r, 0, sizeof(*r));
    int32_t nt = 0, n = 0;
    float totalt = 0.0f;
    int32_t term = 0;
    do {
        float sel[3];
        sel[0] = g_step(tm[0], tm[2]) * g_step(tm[0], tm[1]);
        sel[1] = g_step(tm[1], tm[0]) * g_step(tm[1], tm[2]);
        sel[2] = g_step(tm[2], tm[1]) * g_step(tm[2], tm[0]);
        int oob = 0;
        for (int k = 0; k < 3; ++k) {
            int32_t c = cv[k] < 0 ? 0 : (cv[k] > up[k] ? up[k] : cv[k]);
            if (c != cv[k]) oob = 1;
        }
        if (oob) { term = 2; break; }
        int32_t vox = vxo_get_volume_at(v, cv[0], cv[1], cv[2], 0);
        float best_t = (tm[0] * sel[0] + tm[1] * sel[1]) + tm[2] * sel[2];
        if (vox) {
            term = 1;
            for (int k = 0; k < 3; ++k) {
                r->normal[k] = -s[k] * sel[k];
                r->hitpos[k] = o[k] + d[k] * best_t;
                r->voxel[k] = cv[k];
            }
            r->t = best_t;
            break;
        }
        for (int k = 0; k < 3; ++k) {
            cv[k] += vxo_f2i(sel[k] * s[k]);
            tm[k] += td[k] * sel[k];
        }
        totalt = best_t;
        nt++;
    } while (++n < 256 && totalt < maxt);
    if (term == 0) {
        r->fetches = 4 * nt;
        nt *= 4;
    } else {
        r->fetches = nt + (term == 1 ? 1 : 0);
    }
    r->hit = term == 1;
    r->term = term;
    r->steps = nt;
    return r->hit;
}

/* raycastShadowVolumeSparse (Light.frag:131-173) and ...SuperSparse
 * (Light.frag:175-217): identical bodies with first step 0.5 resp. 2.5. */
static float sparse_march(const vxo_occ* v, const float o[3], const float d[3], float dist,
                          float step_factor, vxo_ray_result* r) {
    float sf = step_factor;
    float sd[3] = {d[0] * sf, d[1] * sf, d[2] * sf};
    float pos[3] = {o[0], o[1], o[2]};
    float dd = sf;
    int32_t it = 0;
    memset(r, 0, sizeof(*r));
    while (dd < 16.0f) {
        it++;
        uint32_t vol = occ_texel(v, vxo_f2i(pos[0] / 2.0f), vxo_f2i(pos[1] / 2.0f), vxo_f2i(pos[2] / 2.0f));
        uint32_t bit = 0u;
        bit += vxo_mod(pos[0], 0.5f) > 0.25f ? 1u : 0u;
        bit += vxo_mod(pos[1], 0.5f) > 0.25f ? 2u : 0u;
        bit += vxo_mod(pos[2], 0.5f) > 0.25f ? 4u : 0u;
        uint32_t mask = 1u << bit;
        if ((mask & vol) != 0) { r->hit = 1; r->term = 1; r->steps = r->fetches = it; r->t = dd; return dd; }
        for (int k = 0; k < 3; ++k) pos[k] += sd[k];
        dd += sf;
    }
    sf *= 2.0f;
    for (int k = 0; k < 3; ++k) sd[k] *= 2.0f;
    float lim = g_min(dist, 164.0f);
    while (dd < lim) {
        it++;
        if (vxo_get_volume_at(v, vxo_f2i(pos[0]), vxo_f2i(pos[1]), vxo_f2i(pos[2]), 1)) {
            r->hit = 1; r->term = 1; r->steps = r->fetches = it; r->t = dd; return dd;
        }
        for (int k = 0; k < 3; ++k) pos[k] += sd[k];
        dd += sf;
    }
    r->steps = r->fetches = it;
    r->t = dist;
    return dist;
}
float vxo_sparse(const vxo_occ* v, const float o[3], const float d[3], float dist, vxo_ray_result* r) {
    return sparse_march(v, o, d, dist, 0.5f, r);
}
float vxo_supersparse(const vxo_occ* v, const float o[3], const float d[3], float dist, vxo_ray_result* r) {
    return sparse_march(v, o, d, dist, 2.5f, r);
}

static void march_any(const vxo_occ* v, int32_t march, const float o[3], const float d[3], float maxt,
                      vxo_ray_result* r) {
    if (march == VXO_MARCH_DDA) vxo_dda(v, o, d, maxt, r);
    else if (march == VXO_MARCH_SPARSE) vxo_sparse(v, o, d, maxt, r);
    else vxo_supersparse(v, o, d, maxt, r);
}

void vxo_trace_rays(const vxo_occ* v, const float* rays, int32_t n, int32_t march,
                    vxo_ray_result* out, int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < n; ++i) {
        const float* ry = rays + (size_t)7 * i;
        march_any(v, march, ry, ry + 3, ry[6], &out[i]);
    }
    (void)nthreads;
}

/* clipToAABB, GeometryVoxel.frag:49-60. */
void vxo_clip_to_aabb(const float o[3], const float d[3], const float size[3], float out[3]) {
    int inside = 1;
    for (int k = 0; k < 3; ++k)
        if (g_clamp(o[k], 0.0f, size[k]) != o[k]) inside = 0;
    if (inside) { out[0] = o[0]; out[1] = o[1]; out[2] = o[2]; return; }
    float t[3];
    for (int k = 0; k < 3; ++k) {
        float inv = 1.0f / d[k];
        float sgn = g_step(d[k], 0.0f);
        t[k] = (sgn * size[k] - o[k]) * inv;
    }
    float tmin = g_max(g_max(t[0], t[1]), t[2]);
    for (int k = 0; k < 3; ++k) out[k] = o[k] + d[k] * (tmin - 0.001f);
}

/* intersectVolume, GeometryVoxel.frag:63-125 (LOD 1, MIP_COUNT 2, TRANSPARENCY 1). */
int32_t vxo_intersect_volume(const uint8_t* const mips[3], const int32_t dims[3],
                             const float o_in[3], const float d[3], const float cam_local[3],
                             const float uv[2], const float res[2], int32_t frame,
                             float hit[3], float normal[3], uint8_t* material) {
    float s[3], td[3], o[3] = {o_in[0], o_in[1], o_in[2]};
    for (int k = 0; k < 3; ++k) {
        s[k] = vxo_sign(d[k]);
        td[k] = 1.0f / (d[k] * s[k]);
    }
    int32_t mip = 2;
    int32_t i = 0;
    do {
        float mip_size = (float)(1 << mip);
        int32_t cv[3], up[3], md[3];
        float tm[3];
        for (int k = 0; k < 3; ++k) {
            o[k] = o[k] / mip_size;
            cv[k] = vxo_f2i(floorf(o[k]));
            float nb = (float)cv[k] + (s[k] * 0.5f + 0.5f);
            tm[k] = (nb - o[k]) / d[k];
            up[k] = vxo_f2i((float)dims[k] / mip_size) + 1;
            md[k] = dims[k] >> mip;
        }
        int32_t n = 0;
        int descended = 0;
        do {
            float sel[3];
            sel[0] = g_step(tm[0], tm[2]) * g_step(tm[0], tm[1]);
            sel[1] = g_step(tm[1], tm[0]) * g_step(tm[1], tm[2]);
            sel[2] = g_step(tm[2], tm[1]) * g_step(tm[2], tm[0]);
            for (int k = 0; k < 3; ++k) cv[k] += vxo_f2i(sel[k] * s[k]);
            for (int k = 0; k < 3; ++k) {
                int32_t c = cv[k] < -1 ? -1 : (cv[k] > up[k] ? up[k] : cv[k]);
                if (c != cv[k]) return 0;
            }
            uint32_t vox = 0;
            if (cv[0] >= 0 && cv[1] >= 0 && cv[2] >= 0 && cv[0] < md[0] && cv[1] < md[1] && cv[2] < md[2])
                vox = mips[mip][(size_t)cv[0] + (size_t)cv[1] * md[0] + (size_t)cv[2] * md[0] * md[1]];
            if (vox != 0u) {
                float best_t = (tm[0] * sel[0] + tm[1] * sel[1]) + tm[2] * sel[2];
                int accept = (mip == 0);
                if (!accept) {
                    float pt[3], df[3];
                    for (int k = 0; k < 3; ++k) {
                        pt[k] = (o[k] + d[k] * best_t) * mip_size;
                        df[k] = cam_local[k] - pt[k];
                    }
                    accept = (float)mip < 0.001f * sqrtf(dot3(df, df));
                }
                if (accept) {
                    float cx = roundf((uv[0] * res[0]) * 0.5f);
                    float cy = roundf((uv[1] * res[1]) * 0.5f);
                    int glass = vox < 16u && vxo_mod(cy + cx, 2.0f) == (float)(frame % 2);
                    if (!glass) {
                        for (int k = 0; k < 3; ++k) {
                            normal[k] = -s[k] * sel[k];
                            hit[k] = (o[k] + d[k] * best_t) * mip_size;
                        }
                        *material = (uint8_t)vox;
                        return 1;
                    }
                } else {
                    mip--;
                    for (int k = 0; k < 3; ++k) o[k] = o[k] + d[k] * best_t / mip_size;
                    descended = 1;
                    break;
                }
            }
            for (int k = 0; k < 3; ++k) tm[k] += td[k] * sel[k];
        } while (++n < 512);
        (void)descended;
        for (int k = 0; k < 3; ++k) o[k] = o[k] * mip_size;
    } while (++i < 4);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Full passes                                                               */
/* ------------------------------------------------------------------------ */

/* mat4 * vec4, column-major, r_i = ((m0i*x + m1i*y) + m2i*z) + m3i*w. */
static inline void mat_vec(const float m[16], const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[0 + r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

/* a * b, column-major, each entry summed in the order of mat_vec. */
static void mat_mul_f(const float* a, const float* b, float* out) {
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            out[j * 4 + r] = ((a[0 + r] * b[j * 4 + 0] + a[4 + r] * b[j * 4 + 1]) + a[8 + r] * b[j * 4 + 2]) +
                             a[12 + r] * b[j * 4 + 3];
}

/* Build-defined pinhole ray through pixel (px,py) from the ViewData inverse
 * matrices (LightAmbient.vert:32-36 FarVec, then the inverse view). */
static void pixel_ray(const vxo_view* vw, int32_t px, int32_t py, float dir[3], float ndc[2]) {
    float u = ((float)px + 0.5f + vw->jitter[0]) / vw->res[0];
    float v = ((float)py + 0.5f + vw->jitter[1]) / vw->res[1];
    ndc[0] = u * 2.0f - 1.0f;
    ndc[1] = 1.0f - v * 2.0f;
    float c[4] = {ndc[0], ndc[1], 1.0f, 1.0f}, f[4];
    mat_vec(vw->inv_proj, c, f);
    float fv[4] = {f[0] / f[3], f[1] / f[3], f[2] / f[3], 0.0f}, w[4];
    mat_vec(vw->inv_view, fv, w);
    normalize3(w, dir);
}

/* A draw's box [0, dims] (model space, fine voxels) rasterizes at a pixel when
 * the pixel's forward ray (o, d) leaves it at a view depth >= NEAR: the CUBE
 * triangles (GeometryVoxel.vert:5-42) face inward and back faces are culled
 * with CCW fronts (evk.cpp:470-471), so the far faces are the ones drawn,
 * clipped by the near plane. world maps model space (in tenths) to world
 * space; NULL is the identity. pv = projection * view. */
static int box_far_face_drawn(const int32_t dims[3], const float o[3], const float d[3], const float* world,
                              const float* pv) {
    float tn = -INFINITY, tf = INFINITY;
    for (int k = 0; k < 3; ++k) {
        float inv = 1.0f / d[k];
        float t0 = (0.0f - o[k]) * inv, t1 = ((float)dims[k] - o[k]) * inv;
        tn = fmaxf(tn, fminf(t0, t1));
        tf = fminf(tf, fmaxf(t0, t1));
    }
    if (!(tn <= tf && tf > 0.0f)) return 0;
    float pe[4] = {(o[0] + d[0] * tf) * 0.1f, (o[1] + d[1] * tf) * 0.1f, (o[2] + d[2] * tf) * 0.1f, 1.0f}, we[4], ce[4];
    if (world) mat_vec(world, pe, we);
    else for (int k = 0; k < 4; ++k) we[k] = pe[k];
    mat_vec(pv, we, ce);
    return ce[3] >= 0.1f;
}

/* GeometryVoxel main (GeometryVoxel.frag:127-183) for the world model with an
 * identity WorldMatrix: localCameraPos = CameraPosition*10
 * (GeometryVoxel.vert:74); G-buffer gets the fp32 local hit point. A pixel
 * the model's box does not cover (box_far_face_drawn: no fragment, e.g. a
 * box behind the camera) stays empty. pv = projection * view. */
static void primary_pixel(const uint8_t* const mips[3], const int32_t dims[3], const vxo_view* vw, const float* pv,
                          int32_t px, int32_t py, float* pos, float* nrm, uint8_t* mat) {
    float dir[3], ndc[2], cam[3], start[3], hit[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    pixel_ray(vw, px, py, dir, ndc);
    for (int k = 0; k < 3; ++k) cam[k] = vw->cam_pos[k] * 10.0f;
    float size[3] = {(float)dims[0], (float)dims[1], (float)dims[2]};
    uint8_t m = 0;
    int h = 0;
    if (box_far_face_drawn(dims, cam, dir, NULL, pv)) {
        vxo_clip_to_aabb(cam, dir, size, start);
        h = vxo_intersect_volume(mips, dims, start, dir, cam, ndc, vw->res, vw->frame, hit, n, &m);
    }
    if (!h) m = 0;
    for (int k = 0; k < 3; ++k) { pos[k] = h ? hit[k] : 0.0f; nrm[k] = h ? n[k] : 0.0f; }
    *mat = m;
}

void vxo_primary(const uint8_t* const mips[3], const int32_t dims[3], const vxo_view* vw,
                 float* pos, float* nrm, uint8_t* mat, int32_t nthreads) {
    const int32_t W = vw->width, H = vw->height;
    float pv[16];
    mat_mul_f(vw->proj, vw->view, pv);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int32_t py = 0; py < H; ++py) {
        for (int32_t px = 0; px < W; ++px) {
            size_t i = (size_t)py * W + px;
            primary_pixel(mips, dims, vw, pv, px, py, pos + 3 * i, nrm + 3 * i, mat + i);
        }
    }
    (void)nthreads;
}

void vxo_primary_pixels(const uint8_t* const mips[3], const int32_t dims[3], const vxo_view* vw,
                        const int32_t* pixels, int32_t npixels, float* pos, float* nrm, uint8_t* mat,
                        int32_t nthreads) {
    float pv[16];
    mat_mul_f(vw->proj, vw->view, pv);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int32_t j = 0; j < npixels; ++j) {
        const int32_t i = pixels[j];
        primary_pixel(mips, dims, vw, pv, i % vw->width, i / vw->width, pos + 3 * (size_t)j,
                      nrm + 3 * (size_t)j, mat + j);
    }
    (void)nthreads;
}

/* ---- multi-instance geometry pass (next rank 3) ---- */
/* 4x4 inverse (column-major) in float64: cofactors over the determinant. */
static int inverse4(const double* m, double* inv) {
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
    double det = m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12];
    if (det == 0.0 || !isfinite(det)) return 0;
    for (int i = 0; i < 16; ++i) inv[i] = c[i] / det;
    return 1;
}

static int box_covers(const vxo_vox_cmd* c, const float o[3], const float d[3], const float* pv) {
    return box_far_face_drawn(c->dims, o, d, c->world, pv);
}

/* The geometry pass per pixel: GeometryVoxel.vert:72-75 gives the camera in
 * model space (inverse(WorldMatrix) * camera * 10) and the direction toward
 * the fragment; GeometryVoxel.frag:131-146 clips to the box and marches; the
 * hit's fragment depth (frag:176-177) goes through the D24 conversion and the
 * LESS test (evk.cpp:481) against the cleared 1.0, draws in order. */
void vxo_primary_instances(const vxo_vox_cmd* cmds, int32_t n, const vxo_view* vw, float* pos, float* nrm,
                           uint8_t* mat, float* motion, float* depth, uint16_t* palette, int32_t nthreads) {
    const int32_t W = vw->width, H = vw->height;
    float (*winv)[16] = (float (*)[16])malloc(sizeof(float) * 16 * (size_t)(n > 0 ? n : 1));
    float (*lc)[3] = (float (*)[3])malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    int* ok = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int32_t k = 0; k < n; ++k) {
        double wd[16], wi[16];
        for (int q = 0; q < 16; ++q) wd[q] = cmds[k].world[q];
        ok[k] = inverse4(wd, wi);
        for (int q = 0; q < 16; ++q) winv[k][q] = (float)wi[q];
        double cx = vw->cam_pos[0], cy = vw->cam_pos[1], cz = vw->cam_pos[2];
        for (int q = 0; q < 3; ++q) lc[k][q] = (float)((((wi[q] * cx + wi[4 + q] * cy) + wi[8 + q] * cz) + wi[12 + q]) * 10.0);
    }
    float pv[16], plv[16];
    mat_mul_f(vw->proj, vw->view, pv);
    mat_mul_f(vw->proj, vw->last_view, plv);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int32_t py = 0; py < H; ++py) {
        for (int32_t px = 0; px < W; ++px) {
            size_t i = (size_t)py * W + px;
            float dw[3], ndc[2];
            pixel_ray(vw, px, py, dw, ndc);
            uint32_t best_q = 16777215u;
            int32_t best = -1;
            float bhit[3] = {0, 0, 0}, bn[3] = {0, 0, 0};
            uint8_t bmat = 0;
            for (int32_t k = 0; k < n; ++k) {
                if (!ok[k]) continue;
                const float* Wi = winv[k];
                float dl[3], dr[3], start[3], hit[3], nn[3];
                dr[0] = (Wi[0] * dw[0] + Wi[4] * dw[1]) + Wi[8] * dw[2];
                dr[1] = (Wi[1] * dw[0] + Wi[5] * dw[1]) + Wi[9] * dw[2];
                dr[2] = (Wi[2] * dw[0] + Wi[6] * dw[1]) + Wi[10] * dw[2];
                normalize3(dr, dl);
                if (!box_covers(&cmds[k], lc[k], dl, pv)) continue;
                float size[3] = {(float)cmds[k].dims[0], (float)cmds[k].dims[1], (float)cmds[k].dims[2]};
                vxo_clip_to_aabb(lc[k], dl, size, start);
                uint8_t m = 0;
                if (!vxo_intersect_volume(cmds[k].mips, cmds[k].dims, start, dl, lc[k], ndc, vw->res, vw->frame,
                                          hit, nn, &m))
                    continue;
                float h4[4] = {hit[0] * 0.1f, hit[1] * 0.1f, hit[2] * 0.1f, 1.0f}, wh[4], cur[4];
                mat_vec(cmds[k].world, h4, wh);
                mat_vec(pv, wh, cur);
                float d = ((1.0f - 1e-7f * (float)cmds[k].rid) * (cur[3] - 0.1f)) / (4096.0f - 0.1f);
                d = d < 0.0f ? 0.0f : (d > 1.0f ? 1.0f : d);
                uint32_t q = (uint32_t)rintf(d * 16777215.0f);
                if (q < best_q) {
                    best_q = q; best = k; bmat = m;
                    for (int c = 0; c < 3; ++c) { bhit[c] = hit[c]; bn[c] = nn[c]; }
                }
            }
            float po[3] = {0, 0, 0}, no[3] = {0, 0, 0}, mo[2] = {0, 0};
            uint16_t pal = 0;
            if (best >= 0) {
                const vxo_vox_cmd* c = &cmds[best];
                float h4[4] = {bhit[0] * 0.1f, bhit[1] * 0.1f, bhit[2] * 0.1f, 1.0f}, wh[4], cur[4], lw[4], last[4];
                mat_vec(c->world, h4, wh);
                mat_vec(pv, wh, cur);
                mat_vec(c->last_world, h4, lw);
                mat_vec(plv, lw, last);
                float n4[4] = {bn[0], bn[1], bn[2], 0.0f}, wn[4];
                mat_vec(c->world, n4, wn);
                float l = sqrtf(((wn[0] * wn[0] + wn[1] * wn[1]) + wn[2] * wn[2]) + wn[3] * wn[3]);
                for (int k = 0; k < 3; ++k) { po[k] = wh[k] * 10.0f; no[k] = wn[k] / l; }
                mo[0] = -0.5f * (cur[0] / cur[3] - last[0] / last[3]);
                mo[1] = 0.5f * (cur[1] / cur[3] - last[1] / last[3]);
                pal = (uint16_t)c->palette;
            }
            for (int k = 0; k < 3; ++k) { pos[3 * i + k] = po[k]; nrm[3 * i + k] = no[k]; }
            mat[i] = bmat;
            if (motion) { motion[2 * i] = mo[0]; motion[2 * i + 1] = mo[1]; }
            if (depth) depth[i] = (float)best_q / 16777215.0f;
            if (palette) palette[i] = pal;
        }
    }
    free(winv); free(lc); free(ok);
    (void)nthreads;
}

/* ---- LightTAA (next rank 4) ---- */
/* A TAA input: per-pixel attribute with the border of an opaque-black
 * clamp-to-border sampler (rgb 0, alpha 1). kind selects how texels are made. */
typedef struct {
    int32_t w, h, kind, n;   /* kind 0: plain image of n channels; 1: light (+alpha from mat);
                                2: colour from palette; 3: material from palette */
    const float* img;
    const uint8_t* mat;
    const float* pal;
} taa_tex;

static float taa_texel(const taa_tex* t, int32_t x, int32_t y, int c) {
    if (x < 0 || y < 0 || x >= t->w || y >= t->h) return c == 3 ? 1.0f : 0.0f;
    size_t i = (size_t)y * t->w + x;
    switch (t->kind) {
        case 1: return c < 3 ? t->img[3 * i + c] : (t->mat[i] == 0 ? 1.0f : 0.0f);
        case 2: return t->mat[i] ? t->pal[6 * t->mat[i] + c] : 0.0f;
        case 3: return (t->mat[i] && c < 3) ? t->pal[6 * t->mat[i] + 3 + c] : 0.0f;
        default: return t->img ? t->img[(size_t)t->n * i + c] : 0.0f;
    }
}

/* Position x in texels to the left texel and the 8-bit weight of its right
 * neighbour. Positions left of -2 or right of size + 1 (and NaN) read border
 * texels only and are clamped first. */
static void taa_weight(float x, int32_t size, int32_t* x0, float* f) {
    if (!(x >= -2.0f)) x = -2.0f;
    if (x > (float)size + 1.0f) x = (float)size + 1.0f;
    float fl = floorf(x);
    float w = floorf((x - fl) * 256.0f + 0.5f) * (1.0f / 256.0f);
    if (w >= 1.0f) { fl = fl + 1.0f; w = 0.0f; }
    *x0 = (int32_t)fl;
    *f = w;
}

static void taa_sample(const taa_tex* t, float u, float v, int n, float* out) {
    int32_t x0, y0;
    float fx, fy;
    taa_weight(u * (float)t->w - 0.5f, t->w, &x0, &fx);
    taa_weight(v * (float)t->h - 0.5f, t->h, &y0, &fy);
    for (int c = 0; c < n; ++c) {
        float a = taa_texel(t, x0, y0, c), b = taa_texel(t, x0 + 1, y0, c);
        float cc = taa_texel(t, x0, y0 + 1, c), d = taa_texel(t, x0 + 1, y0 + 1, c);
        out[c] = (a * (1.0f - fx) + b * fx) * (1.0f - fy) + (cc * (1.0f - fx) + d * fx) * fy;
    }
}

void vxo_taa_sample(const float* img, int32_t width, int32_t height, int32_t n, float u, float v, float* out) {
    taa_tex t = {width, height, 0, n, img, NULL, NULL};
    taa_sample(&t, u, v, n, out);
}

static float taa_len(const float* d, int n) {
    float s = d[0] * d[0];
    for (int k = 1; k < n; ++k) s = s + d[k] * d[k];
    return sqrtf(s);
}

void vxo_light_taa(int32_t W, int32_t H, int32_t frame, uint32_t seed, const float* light, const float* last,
                   const float* motion, const float* depth, const float* normal, const uint8_t* mat,
                   const float* palette, float* out, int32_t nthreads) {
    const taa_tex T_light = {W, H, 1, 3, light, mat, palette}, T_last = {W, H, 0, 4, last, mat, palette};
    const taa_tex T_motion = {W, H, 0, 2, motion, mat, palette}, T_color = {W, H, 2, 3, NULL, mat, palette};
    const taa_tex T_matl = {W, H, 3, 4, NULL, mat, palette}, T_normal = {W, H, 0, 3, normal, mat, palette};
    const taa_tex T_depth = {W, H, 0, 1, depth, mat, palette};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t ii = 0; ii < (int64_t)W * H; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const int32_t px = (int32_t)(i % (uint32_t)W), py = (int32_t)(i / (uint32_t)W);
        const float u = ((float)px + 0.5f) / (float)W, v = ((float)py + 0.5f) / (float)H;
        const float irx = 1.0f / (float)W, iry = 1.0f / (float)H;
        float mo[2], cur[4], col[3], matl[4], nrm[3], dep, lst[4];
        float* o = out + 4 * (size_t)i;
        taa_sample(&T_motion, u, v, 2, mo);
        const float ou = u + mo[0], ov = v + mo[1];
        taa_sample(&T_light, u, v, 4, cur);
        taa_sample(&T_color, u, v, 3, col);
        taa_sample(&T_matl, u, v, 4, matl);
        taa_sample(&T_normal, u, v, 3, nrm);
        taa_sample(&T_depth, u, v, 1, &dep);
        if (dep == 1.0f) {                                       /* LightTAA.frag:50-53 */
            for (int c = 0; c < 4; ++c) o[c] = cur[c];
            continue;
        }
        taa_sample(&T_last, ou, ov, 4, lst);
        float lvar = lst[3];
        if (g_clamp(ou, 0.0f, 1.0f) != ou || g_clamp(ov, 0.0f, 1.0f) != ov) {   /* LightTAA.frag:58-83 */
            float count = 0.0f, acc[3] = {0.0f, 0.0f, 0.0f};
            for (int x = -9; x <= 9; ++x)
                for (int y = -9; y <= 9; ++y) {
                    const float su = g_clamp(u + (float)x * irx, 0.001f, 0.999f);
                    const float sv = g_clamp(v + (float)y * iry, 0.001f, 0.999f);
                    float nm[2], nc[3], nn[3], nmat[4], nd, nl[3], dd[4];
                    taa_sample(&T_motion, su, sv, 2, nm);
                    taa_sample(&T_color, su, sv, 3, nc);
                    taa_sample(&T_normal, su, sv, 3, nn);
                    taa_sample(&T_matl, su, sv, 4, nmat);
                    taa_sample(&T_depth, su, sv, 1, &nd);
                    float f = g_max((nrm[0] * nn[0] + nrm[1] * nn[1]) + nrm[2] * nn[2], 0.0f);
                    for (int k = 0; k < 4; ++k) dd[k] = matl[k] - nmat[k];
                    f *= g_step(0.8f, 1.0f - taa_len(dd, 4));
                    f *= 1.0f - g_clamp(fabsf(dep - nd) * 4096.0f, 0.0f, 1.0f);
                    for (int k = 0; k < 3; ++k) dd[k] = col[k] - nc[k];
                    f *= 1.0f - g_clamp(taa_len(dd, 3), 0.0f, 1.0f);
                    dd[0] = mo[0] - nm[0]; dd[1] = mo[1] - nm[1];
                    if (taa_len(dd, 2) > 0.1f) f = 0.0f;
                    taa_sample(&T_light, su, sv, 3, nl);
                    for (int k = 0; k < 3; ++k) acc[k] = acc[k] + f * nl[k];
                    count = count + f;
                }
            o[0] = acc[0] / count; o[1] = acc[1] / count; o[2] = acc[2] / count; o[3] = 1.0f;
            continue;
        }
        /* LightTAA.frag:87-137 */
        float nmin[3] = {10000.0f, 10000.0f, 10000.0f}, nmax[3] = {0.0f, 0.0f, 0.0f}, acc[3] = {0.0f, 0.0f, 0.0f};
        float diff = 1.0f, count = 0.0f, radius = 1.0f;
        const float size = 12.0f;
        const float noise = vxo_rand(vxo_pixkey(seed ^ 0x85ebca6bu, frame, i), 0u);
        const float scale = g_clamp(0.1f, 0.5f, 1.0f / dep);
        for (float angle = (noise * 3.1415f) * 2.118033988749895f; radius <= size; angle = angle + 2.39f) {
            radius = radius + 1.0f;
            float sn, cs, nm[2], nc[3], nn[3], nmat[4], nd, nl[3], dd[4];
            vxo_sincos(angle, &sn, &cs);
            const float rr = radius * (lvar + 1.0f);
            const float su = g_clamp(u + ((cs * irx) * rr) * scale, 0.001f, 0.999f);
            const float sv = g_clamp(v + ((sn * iry) * rr) * scale, 0.001f, 0.999f);
            taa_sample(&T_motion, su, sv, 2, nm);
            taa_sample(&T_color, su, sv, 3, nc);
            taa_sample(&T_normal, su, sv, 3, nn);
            taa_sample(&T_matl, su, sv, 4, nmat);
            taa_sample(&T_depth, su, sv, 1, &nd);
            float f = 1.212f - radius / size;
            for (int k = 0; k < 4; ++k) dd[k] = matl[k] - nmat[k];
            f *= g_step(0.8f, 1.0f - taa_len(dd, 4));
            f *= g_max((nrm[0] * nn[0] + nrm[1] * nn[1]) + nrm[2] * nn[2], 0.0f);
            f *= 1.0f - g_clamp(fabsf(dep - nd) * 4096.0f, 0.0f, 1.0f);
            for (int k = 0; k < 3; ++k) dd[k] = col[k] - nc[k];
            f *= 1.0f - g_clamp(taa_len(dd, 3) * 10000.0f, 0.0f, 1.0f);
            dd[0] = mo[0] - nm[0]; dd[1] = mo[1] - nm[1];
            if (taa_len(dd, 2) > 0.1f) f = 0.0f;
            taa_sample(&T_light, su, sv, 3, nl);
            for (int k = 0; k < 3; ++k) {
                nmin[k] = g_min(nmin[k], nl[k]);
                nmax[k] = g_max(nmax[k], nl[k]);
                acc[k] = acc[k] + f * nl[k];
                dd[k] = nl[k] - cur[k];
            }
            count = count + f;
            diff = diff + taa_len(dd, 3) * f;
        }
        float d[3];
        for (int k = 0; k < 3; ++k) cur[k] = (cur[k] + acc[k]) / (count + 1.0f);
        diff = diff / radius;
        for (int k = 0; k < 3; ++k) {
            lst[k] = g_clamp(lst[k], nmin[k], nmax[k]);
            d[k] = fabsf(cur[k] - lst[k]);
        }
        float var = (d[0] * 0.2125f + d[1] * 0.7154f) + d[2] * 0.0721f;   /* luminance, frag:31-36 */
        var = g_clamp(var * 5.5f, 0.0f, 1.0f);
        lvar = lvar + var;
        lvar = lvar - diff * 0.08f;
        {
            float m2[2] = {mo[0], mo[1]};
            lvar = lvar + taa_len(m2, 2) * 20.0f;
        }
        const float t = g_clamp(1.0f - lvar, 0.3f, 0.9f);
        for (int k = 0; k < 3; ++k) o[k] = cur[k] * (1.0f - t) + lst[k] * t;
        o[3] = g_clamp(lvar * 0.7f, 0.0f, 1.0f);
    }
    (void)nthreads;
}

/* Tangent frame of calculateAmbientIrradiance, LightAmbient.frag:112-113. */
static inline void tangent_frame(const float n[3], float t[3], float b[3]) {
    if (fabsf(n[2]) > 0.5f) { t[0] = 0.0f; t[1] = -n[2]; t[2] = n[1]; }
    else { t[0] = -n[1]; t[1] = n[0]; t[2] = 0.0f; }
    b[0] = n[1] * t[2] - n[2] * t[1];
    b[1] = n[2] * t[0] - n[0] * t[2];
    b[2] = n[0] * t[1] - n[1] * t[0];
}

static inline float march_dist(const vxo_occ* v, int32_t march, const float o[3], const float d[3],
                               float maxt, uint32_t* fetches, int32_t* hit) {
    vxo_ray_result r;
    march_any(v, march, o, d, maxt, &r);
    *fetches += (uint32_t)r.fetches;
    *hit = r.hit;
    if (march == VXO_MARCH_DDA) return r.hit ? g_min(r.t, maxt) : maxt;
    return r.t;
}

/* Rays of one non-sky pixel, in ray order: n_shadow shadow rays, n_ao AO rays,
 * n_spec specular-occlusion rays; rays[7*r..7*r+6] = origin xyz, direction
 * xyz, maxt. The ambient pass's ray set-up (LightAmbient.frag:134-175)
 * generalised to n_ao cosine AO rays with a length cap, and a GGX-sampled
 * reflection ray from the origin of LightReflection.frag:93-113
 * (build-defined direction). */
static void pixel_rays(const vxo_view* vw, const vxo_trace_params* p, const float* pos, const float* nrm,
                       const uint8_t* mat, const float* palette, size_t i, float* rays) {
    const float* pp = pos + 3 * i;
    const float* n = nrm + 3 * i;
    uint32_t pk = vxo_pixkey(p->seed, p->frame, (uint32_t)i);
    float r0[4];
    for (int k = 0; k < 4; ++k) r0[k] = vxo_rand(pk, (uint32_t)k);

    /* linear depth for the normal bias (LightAmbient.frag:158; Common.frag:12-13) */
    float wx = pp[0] * 0.1f, wy = pp[1] * 0.1f, wz = pp[2] * 0.1f;
    const float* V = vw->view;
    float vz = ((V[2] * wx + V[6] * wy) + V[10] * wz) + V[14];
    float depth = ((-vz) - NEAR_PLANE) / (FAR_PLANE - NEAR_PLANE);
    float st = g_clamp((depth - 0.0f) / (0.2f - 0.0f), 0.0f, 1.0f);
    float bias = (st * st) * (3.0f - 2.0f * st) * 50.0f + 1.5f;

    /* shadow ray set-up, LightAmbient.frag:149-158 */
    float wd[3] = {p->light_dir[0], p->light_dir[1], p->light_dir[2]};
    float wcp[3] = {pp[0], pp[1], pp[2]};
    if (p->jitter_shadow) {
        float rv[3];
        vxo_cosine_hemisphere(r0[0], r0[1], rv);
        for (int k = 0; k < 3; ++k) rv[k] = rv[k] * 0.1f;
        rv[2] = rv[2] * vxo_sign(r0[2] - 0.5f);
        float mx[3];
        for (int k = 0; k < 3; ++k) mx[k] = wd[k] * (1.0f - 0.5f) + rv[k] * 0.5f;
        normalize3(mx, wd);
        for (int k = 0; k < 3; ++k) wcp[k] = wcp[k] + wd[k] * (r0[3] * 1.0f);
        for (int k = 0; k < 3; ++k) wcp[k] = wcp[k] + rv[k] * 2.5f;
    }
    float org[3];
    for (int k = 0; k < 3; ++k) org[k] = wcp[k] + n[k] * bias;

    float* r = rays;
    for (int32_t s = 0; s < p->n_shadow; ++s, r += 7) {
        r[0] = org[0]; r[1] = org[1]; r[2] = org[2];
        r[3] = wd[0]; r[4] = wd[1]; r[5] = wd[2];
        r[6] = p->shadow_maxt;
    }

    /* AO rays, calculateAmbientIrradiance LightAmbient.frag:111-126 */
    if (p->n_ao > 0) {
        float T[3], B[3];
        tangent_frame(n, T, B);
        for (int32_t a = 0; a < p->n_ao; ++a, r += 7) {
            float u1 = vxo_rand(pk, (uint32_t)(4 * (1 + a) + 0));
            float u2 = vxo_rand(pk, (uint32_t)(4 * (1 + a) + 1));
            float rv[3];
            vxo_cosine_hemisphere(u1, u2, rv);
            r[0] = org[0]; r[1] = org[1]; r[2] = org[2];
            for (int k = 0; k < 3; ++k) r[3 + k] = (T[k] * rv[0] + B[k] * rv[1]) + n[k] * rv[2];
            r[6] = p->ao_maxt;
        }
    }

    /* specular-occlusion ray: LightReflection.frag:93,113 origin (wcp + normal
     * * noise.w, marched from + normal) with a GGX-sampled half vector
     * (build-defined; DistributionGGX's NDF, PBR.frag:12-23) */
    if (p->n_spec > 0) {
        uint32_t base = (uint32_t)(4 * (1 + p->n_ao));
        float u1 = vxo_rand(pk, base + 0), u2 = vxo_rand(pk, base + 1), u3 = vxo_rand(pk, base + 3);
        float rough = palette[6 * mat[i] + 3];
        float cam[3], vv[3], Vd[3];
        for (int k = 0; k < 3; ++k) cam[k] = vw->cam_pos[k] * 10.0f;
        for (int k = 0; k < 3; ++k) vv[k] = cam[k] - pp[k];
        normalize3(vv, Vd);
        float alpha = rough * rough;
        float a2 = alpha * alpha;
        float ct2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
        float ct = sqrtf(ct2);
        float stt = sqrtf(g_max(0.0f, 1.0f - ct2));
        float sp, cp;
        vxo_sincos(6.2831855f * u2, &sp, &cp);
        float T[3], B[3], h[3];
        tangent_frame(n, T, B);
        float hl[3] = {stt * cp, stt * sp, ct};
        for (int k = 0; k < 3; ++k) h[k] = (T[k] * hl[0] + B[k] * hl[1]) + n[k] * hl[2];
        float I[3] = {-Vd[0], -Vd[1], -Vd[2]};
        float dhi = dot3(h, I);
        for (int k = 0; k < 3; ++k) r[3 + k] = I[k] - (2.0f * dhi) * h[k];
        for (int k = 0; k < 3; ++k) r[k] = (pp[k] + n[k] * u3) + n[k];
        r[6] = p->spec_maxt;
    }
}

void vxo_pixel_rays(const vxo_view* vw, const vxo_trace_params* p, const float* pos, const float* nrm,
                    const uint8_t* mat, const float* palette, const int32_t* pixels, int32_t npixels,
                    float* rays) {
    const int32_t nr = p->n_shadow + p->n_ao + p->n_spec;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t j = 0; j < npixels; ++j) {
        const size_t i = (size_t)pixels[j];
        float* r = rays + (size_t)j * nr * 7;
        if (mat[i] == 0) memset(r, 0, (size_t)nr * 7 * sizeof(float));
        else pixel_rays(vw, p, pos, nrm, mat, palette, i, r);
    }
}

/* One pixel of the ambient pass (LightAmbient.frag:134-175) and the specular
 * occlusion ray (LightReflection.frag:86-113): the pixel's rays, each
 * marched, folded in ray order. Shadow: lit unless the ray hit (DDA) or the
 * marcher returned less than maxt (LightAmbient.frag:167); AO: mean of
 * (min(t, cap) / cap)^2 (the d*d of LightAmbient.frag:121-125 per ray);
 * spec: min(t, maxt) / maxt. */
static void trace_pixel(const vxo_occ* v, const vxo_view* vw, const vxo_trace_params* p,
                        const float* pos, const float* nrm, const uint8_t* mat, const float* palette,
                        size_t i, uint8_t* shadow, float* ao, float* spec, uint32_t* fetches) {
    uint32_t nf = 0;
    if (mat[i] == 0) {
        shadow[i] = 0; ao[i] = 0.0f; spec[i] = 0.0f;
        if (fetches) fetches[i] = 0;
        return;
    }
    float rays[7 * (1 + 64 + 1)];
    pixel_rays(vw, p, pos, nrm, mat, palette, i, rays);
    const float* r = rays;

    uint8_t lit = 1;
    for (int32_t s = 0; s < p->n_shadow; ++s, r += 7) {
        int32_t h;
        float t = march_dist(v, p->shadow_march, r, r + 3, r[6], &nf, &h);
        if (p->shadow_march == VXO_MARCH_DDA) lit = h ? 0 : 1;
        else lit = (t != p->shadow_maxt) ? 0 : 1;      /* LightAmbient.frag:167 */
    }
    shadow[i] = lit;

    if (p->n_ao > 0) {
        float aosum = 0.0f;
        for (int32_t a = 0; a < p->n_ao; ++a, r += 7) {
            int32_t h;
            float t = march_dist(v, p->ao_march, r, r + 3, r[6], &nf, &h);
            float q = g_min(t, p->ao_maxt) / p->ao_maxt;
            aosum = aosum + q * q;
        }
        ao[i] = aosum / (float)p->n_ao;
    } else {
        ao[i] = 1.0f;
    }

    if (p->n_spec > 0) {
        int32_t hh;
        float t = march_dist(v, p->spec_march, r, r + 3, r[6], &nf, &hh);
        spec[i] = g_min(t, p->spec_maxt) / p->spec_maxt;
    } else {
        spec[i] = 1.0f;
    }
    if (fetches) fetches[i] = nf;
}

void vxo_trace(const vxo_occ* v, const vxo_view* vw, const vxo_trace_params* p,
               const float* pos, const float* nrm, const uint8_t* mat, const float* palette,
               uint8_t* shadow, float* ao, float* spec, uint32_t* fetches,
               const int32_t* pixels, int32_t npixels, int32_t nthreads) {
    int32_t total = pixels ? npixels : vw->width * vw->height;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int32_t j = 0; j < total; ++j) {
        size_t i = pixels ? (size_t)pixels[j] : (size_t)j;
        trace_pixel(v, vw, p, pos, nrm, mat, palette, i, shadow, ao, spec, fetches);
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* Shading: LightAmbient.frag:178-214 with PBRDirectLight (PBR.frag:48-68)   */
/* and fresnelSchlickRoughness (PBR.frag:8-10), in world space. Palette entry */
/* m: {albedo r,g,b, roughness, metallic, emissive}; light: {dir3, rgb3}.    */
/* ------------------------------------------------------------------------ */
static const float PI_F = 3.14159265359f;       /* PBR.frag:1 */
static const double PI_D = 3.14159265359;

void vxo_shade_f32(const vxo_view* vw, const float* pos, const float* nrm, const uint8_t* mat,
                   const uint8_t* shadow, const float* ao, const float* ssao,
                   const float* pal, const float light[6], float* rgb, int32_t nthreads) {
    const int32_t npx = vw->width * vw->height;
    /* getSkyColor(vec3(1,0,0)), LightAmbient.frag:128-131 */
    float ey = (g_max(0.0f, 0.0f) * 0.8f + 0.2f) * 0.8f;
    float irr[3] = {((1.0f - ey) * (1.0f - ey)) * 1.1f, (1.0f - ey) * 1.1f, (0.6f + (1.0f - ey) * 0.4f) * 1.1f};
    float L[3];
    normalize3(light, L);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < npx; ++i) {
        int m = mat[i];
        if (m == 0) { rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = 0.0f; continue; }
        const float* e = pal + 6 * m;
        const float* pp = pos + 3 * (size_t)i;
        const float* N = nrm + 3 * (size_t)i;
        float cam[3], vv[3], V[3];
        for (int k = 0; k < 3; ++k) { cam[k] = vw->cam_pos[k] * 10.0f; vv[k] = cam[k] - pp[k]; }
        normalize3(vv, V);
        float sh = (float)shadow[i];
        float A = (ao[i] * 0.05f) * (ssao ? ssao[i] : 1.0f);
        float ndv = g_max(dot3(N, V), 0.0f);
        float x = g_max(1.0f - ndv, 0.0f);
        float s5 = (((x * x) * x) * x) * x;
        float rough = e[3], metal = e[4], emit = e[5];
        float f0d = 0.04f * (1.0f - metal) + 1.0f * metal;
        float F = f0d + (1.0f - f0d) * s5;
        float kD = (1.0f - F) * (1.0f - metal);
        float ndl = g_max(dot3(N, L), 0.0f);
        for (int c = 0; c < 3; ++c) {
            float rad = light[3 + c] * sh;
            float lo = (((kD * e[c]) / PI_F) * rad) * ndl;
            float f0a = 0.04f * (1.0f - metal) + e[c] * metal;
            float mx = g_max(1.0f - rough, f0a);
            float fr = f0a + (mx - f0a) * s5;
            float kda = 1.0f - fr;
            float amb = (irr[c] * e[c]) * (emit * 10.0f + (kda * A));
            rgb[3 * (size_t)i + c] = amb + lo;
        }
    }
    (void)nthreads;
}

void vxo_shade_f64(const vxo_view* vw, const float* pos, const float* nrm, const uint8_t* mat,
                   const uint8_t* shadow, const float* ao, const float* ssao,
                   const double* pal, const double light[6], double* rgb, int32_t nthreads) {
    const int32_t npx = vw->width * vw->height;
    double ey = (0.0 * 0.8 + 0.2) * 0.8;
    double irr[3] = {((1.0 - ey) * (1.0 - ey)) * 1.1, (1.0 - ey) * 1.1, (0.6 + (1.0 - ey) * 0.4) * 1.1};
    double L[3];
    normalize3d(light, L);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < npx; ++i) {
        int m = mat[i];
        if (m == 0) { rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = 0.0; continue; }
        const double* e = pal + 6 * m;
        double pp[3], N[3], vv[3], V[3];
        for (int k = 0; k < 3; ++k) {
            pp[k] = pos[3 * (size_t)i + k];
            N[k] = nrm[3 * (size_t)i + k];
            vv[k] = (double)(vw->cam_pos[k] * 10.0f) - pp[k];
        }
        normalize3d(vv, V);
        double sh = (double)shadow[i];
        double A = ((double)ao[i] * 0.05) * (ssao ? (double)ssao[i] : 1.0);
        double ndv = g_maxd(dot3d(N, V), 0.0);
        double x = g_maxd(1.0 - ndv, 0.0);
        double s5 = x * x * x * x * x;
        double rough = e[3], metal = e[4], emit = e[5];
        double f0d = 0.04 * (1.0 - metal) + metal;
        double F = f0d + (1.0 - f0d) * s5;
        double kD = (1.0 - F) * (1.0 - metal);
        double ndl = g_maxd(dot3d(N, L), 0.0);
        for (int c = 0; c < 3; ++c) {
            double rad = light[3 + c] * sh;
            double lo = kD * e[c] / PI_D * rad * ndl;
            double f0a = 0.04 * (1.0 - metal) + e[c] * metal;
            double mx = g_maxd(1.0 - rough, f0a);
            double fr = f0a + (mx - f0a) * s5;
            double amb = irr[c] * e[c] * (emit * 10.0 + (1.0 - fr) * A);
            rgb[3 * (size_t)i + c] = amb + lo;
        }
    }
    (void)nthreads;
}

/* Analytic adjoint of vxo_shade_f64 with visibility (shadow, ao, ssao) held
 * fixed: accumulates dL/dpalette (256*6) and dL/dlight (6) given dL/dI. */
void vxo_shade_backward_f64(const vxo_view* vw, const float* pos, const float* nrm,
                            const uint8_t* mat, const uint8_t* shadow, const float* ao,
                            const float* ssao, const double* pal, const double light[6],
                            const double* g, double* dpal, double dlight[6]) {
    const int32_t npx = vw->width * vw->height;
    double ey = (0.0 * 0.8 + 0.2) * 0.8;
    double irr[3] = {((1.0 - ey) * (1.0 - ey)) * 1.1, (1.0 - ey) * 1.1, (0.6 + (1.0 - ey) * 0.4) * 1.1};
    double ll = sqrt(dot3d(light, light));
    double L[3] = {light[0] / ll, light[1] / ll, light[2] / ll};
    memset(dpal, 0, 256 * 6 * sizeof(double));
    memset(dlight, 0, 6 * sizeof(double));
    for (int32_t i = 0; i < npx; ++i) {
        int m = mat[i];
        if (m == 0) continue;
        const double* e = pal + 6 * m;
        double* de = dpal + 6 * m;
        double pp[3], N[3], vv[3], V[3];
        for (int k = 0; k < 3; ++k) {
            pp[k] = pos[3 * (size_t)i + k];
            N[k] = nrm[3 * (size_t)i + k];
            vv[k] = (double)(vw->cam_pos[k] * 10.0f) - pp[k];
        }
        normalize3d(vv, V);
        double sh = (double)shadow[i];
        double A = ((double)ao[i] * 0.05) * (ssao ? (double)ssao[i] : 1.0);
        double ndv = g_maxd(dot3d(N, V), 0.0);
        double x = g_maxd(1.0 - ndv, 0.0);
        double s5 = x * x * x * x * x;
        double rough = e[3], metal = e[4], emit = e[5];
        double f0d = 0.04 * (1.0 - metal) + metal;
        double F = f0d + (1.0 - f0d) * s5;
        double kD = (1.0 - F) * (1.0 - metal);
        double nl = dot3d(N, L);
        double ndl = g_maxd(nl, 0.0);
        double dF_dme = 0.96 * (1.0 - s5);
        double dkD_dme = -dF_dme * (1.0 - metal) - (1.0 - F);
        for (int c = 0; c < 3; ++c) {
            double gc = g[3 * (size_t)i + c];
            double a = e[c];
            double rad = light[3 + c] * sh;
            double f0a = 0.04 * (1.0 - metal) + a * metal;
            int pick_f0 = (1.0 - rough) < f0a;
            double mx = pick_f0 ? f0a : (1.0 - rough);
            double fr = f0a + (mx - f0a) * s5;
            double E = emit * 10.0 + (1.0 - fr) * A;
            /* albedo */
            double dlo_da = kD / PI_D * rad * ndl;
            double dmx_da = pick_f0 ? metal : 0.0;
            double dfr_da = metal + (dmx_da - metal) * s5;
            double damb_da = irr[c] * E - irr[c] * a * A * dfr_da;
            de[c] += gc * (dlo_da + damb_da);
            /* roughness: only through max(1-rough, F0) */
            double dfr_dro = pick_f0 ? 0.0 : -s5;
            de[3] += gc * (-irr[c] * a * A * dfr_dro);
            /* metallic */
            double dlo_dme = dkD_dme * a / PI_D * rad * ndl;
            double df0a_dme = a - 0.04;
            double dmx_dme = pick_f0 ? df0a_dme : 0.0;
            double dfr_dme = df0a_dme + (dmx_dme - df0a_dme) * s5;
            de[4] += gc * (dlo_dme - irr[c] * a * A * dfr_dme);
            /* emissive */
            de[5] += gc * (irr[c] * a * 10.0);
            /* light colour */
            dlight[3 + c] += gc * (kD * a / PI_D * sh * ndl);
            /* light direction through max(dot(N, normalize(ld)), 0) */
            if (!(nl < 0.0)) {
                double dlo_dndl = kD * a / PI_D * rad;
                for (int k = 0; k < 3; ++k) dlight[k] += gc * dlo_dndl * (N[k] - L[k] * nl) / ll;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Build-defined synthetic scenes (inputs for the benchmarks and tests).     */
/* Terrain: occupied where y < H(x,z), H a 4-octave value-noise fBm with the */
/* octave structure of Noise::GetOctave (Noise.cpp:93-107). Materials are a  */
/* hash of the voxel position, uniform in 1..255, so box overlays commute.   */
/* ---------------------------------------------------------------------------
 * Point and spot lights.
 * ------------------------------------------------------------------------ */
static inline float bits_to_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t f_to_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* log2 of a positive finite x: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln(m) = z - z^2/2 + z^3 P(z), z = m - 1 (Cephes logf polynomial). */
float vxo_log2(float x) {
    int32_t e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }   /* denormal: scale by 2^23 */
    uint32_t b = f_to_bits(x);
    e += (int32_t)((b >> 23) & 0xffu) - 127;
    float m = bits_to_f((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float z = m - 1.0f;
    float z2 = z * z;
    float p = 7.0376836292e-2f;
    p = p * z - 1.1514610310e-1f;
    p = p * z + 1.1676998740e-1f;
    p = p * z - 1.2420140846e-1f;
    p = p * z + 1.4249322787e-1f;
    p = p * z - 1.6668057665e-1f;
    p = p * z + 2.0000714765e-1f;
    p = p * z - 2.4999993993e-1f;
    p = p * z + 3.3333331174e-1f;
    float ln = z + ((z * z2) * p - 0.5f * z2);
    return ln * 1.44269504088896341f + (float)e;
}

/* 2^t: t = n + f, n = floor(t + 1/2), 2^f = 1 + f P(f) (Cephes exp2f
 * polynomial), then two exact power-of-two scalings. */
float vxo_exp2(float t) {
    if (!(t == t)) return t;
    if (t >= 128.0f) return INFINITY;
    if (t < -150.0f) return 0.0f;
    float n = floorf(t + 0.5f);
    float f = t - n;
    float p = 1.535336188319500e-4f;
    p = p * f + 1.339887440266574e-3f;
    p = p * f + 9.618437357674640e-3f;
    p = p * f + 5.550332471162809e-2f;
    p = p * f + 2.402264791363012e-1f;
    p = p * f + 6.931472028550421e-1f;
    float r = 1.0f + f * p;
    int32_t k = (int32_t)n;
    int32_t k1 = k / 2, k2 = k - k1;
    r = r * bits_to_f((uint32_t)(k1 + 127) << 23);
    return r * bits_to_f((uint32_t)(k2 + 127) << 23);
}

float vxo_pow(float x, float y) {
    if (!(x == x) || !(y == y) || x < 0.0f) return NAN;
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == INFINITY) return y > 0.0f ? INFINITY : 0.0f;
    return vxo_exp2(y * vxo_log2(x));
}

/* One pixel of the light passes: each light's full-screen draw
 * (LightPointPipeline.h:58-100, LightSpotPipeline.h:60-104) blended additively
 * in draw order, evaluated in world space. */
static void light_pixel(const vxo_occ* v, const vxo_view* vw, const vxo_light* lights, int32_t nlights,
                        const vxo_light_params* p, const float* pos, const float* nrm, const uint8_t* mat,
                        const float* pal, float* rgb, uint32_t* fetches, int32_t i) {
    int m = mat[i];
    if (m == 0) return;
    const float* pp = pos + 3 * (size_t)i;
    const float* N = nrm + 3 * (size_t)i;
    uint32_t pk = vxo_pixkey(p->seed, p->frame, (uint32_t)i);
    float r0[4];
    for (int k = 0; k < 4; ++k) r0[k] = vxo_rand(pk, (uint32_t)k);
    float wpos[3], cam[3], vv[3], V[3];
    for (int k = 0; k < 3; ++k) {
        wpos[k] = pp[k] * 0.1f;                                  /* worldPos */
        cam[k] = vw->cam_pos[k] * 10.0f;
        vv[k] = cam[k] - pp[k];
    }
    normalize3(vv, V);
    const float* e = pal + 6 * m;
    float metal = e[4];
    float ndv = g_max(dot3(N, V), 0.0f);
    float x = g_max(1.0f - ndv, 0.0f);
    float s5 = (((x * x) * x) * x) * x;
    /* PBRDirectLight (PBR.frag:48-68): F0 = mix(0.04, 1, metallic) */
    float f0d = 0.04f * (1.0f - metal) + 1.0f * metal;
    float F = f0d + (1.0f - f0d) * s5;
    float kD = (1.0f - F) * (1.0f - metal);
    float rv[3];
    vxo_cosine_hemisphere(r0[0], r0[1], rv);
    for (int k = 0; k < 3; ++k) rv[k] = rv[k] * 0.1f;
    rv[2] = rv[2] * vxo_sign(r0[2] - 0.5f);
    uint32_t nf = 0;
    for (int32_t l = 0; l < nlights; ++l) {
        const vxo_light* L = &lights[l];
        float ld[3];
        for (int k = 0; k < 3; ++k) ld[k] = L->position[k] - wpos[k];      /* lightDir */
        float dist = sqrtf(dot3(ld, ld));                                  /* lightDistance */
        if (dist > L->range) continue;                                     /* discard */
        const int spot = L->type == VXO_LIGHT_SPOT;
        float wd[3], hit_dist;
        if (spot) {
            float nl[3];
            normalize3(ld, nl);
            for (int k = 0; k < 3; ++k) wd[k] = nl[k] * 10.0f;
            hit_dist = dist * 10.0f;
        } else {
            for (int k = 0; k < 3; ++k) wd[k] = ld[k];
            hit_dist = dist * 10.5f;
        }
        float mx[3], wdn[3], o[3];
        for (int k = 0; k < 3; ++k) mx[k] = wd[k] * (1.0f - 0.5f) + rv[k] * 0.5f;
        normalize3(mx, wdn);
        for (int k = 0; k < 3; ++k) {
            float wcp = wpos[k] * 10.0f;
            wcp = wcp + wdn[k] * (r0[3] * 1.0f);
            wcp = wcp + rv[k] * 2.5f;
            o[k] = wcp + N[k] * 0.5f;
        }
        int32_t hit;
        float d = march_dist(v, spot ? p->spot_march : p->point_march, o, wdn, hit_dist, &nf, &hit);
        float shadow = (d < hit_dist) ? 0.0f : 1.0f;
        float att = g_clamp(L->range - dist, 0.0f, 1.0f) / vxo_pow(dist, L->attenuation);
        if (spot) {
            float nl[3];
            normalize3(ld, nl);
            float ad = g_max(dot3(nl, L->direction) - (1.0f - L->angle), 0.0f) / L->angle;
            att = (vxo_pow(ad, L->angle_attenuation) * g_clamp(L->range - dist, 0.0f, 1.0f)) /
                  vxo_pow(dist, L->attenuation);
        }
        float ndl = g_max(dot3(N, ld), 0.0f);
        for (int c = 0; c < 3; ++c) {
            float rad = (L->color[c] * att) * shadow;
            float lo = (((kD * e[c]) / PI_F) * rad) * ndl;
            rgb[3 * (size_t)i + c] = rgb[3 * (size_t)i + c] + lo;
        }
    }
    if (fetches) fetches[i] = nf;
}

void vxo_light_pass(const vxo_occ* v, const vxo_view* vw, const vxo_light* lights, int32_t nlights,
                    const vxo_light_params* p, const float* pos, const float* nrm, const uint8_t* mat,
                    const float* palette, float* rgb, uint32_t* fetches, int32_t nthreads) {
    const int32_t npx = vw->width * vw->height;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < npx; ++i) light_pixel(v, vw, lights, nlights, p, pos, nrm, mat, palette, rgb, fetches, i);
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* Reflection pass: LightReflection.frag:60-139 (LightReflectionPipeline.h:34) */
/* ------------------------------------------------------------------------ */

/* getSkyColor of LightAmbient.frag:128-131 (the formula LightReflection.frag:
 * 55-56 keeps commented out beside its skybox lookup): pow(x, 2) = x*x. */
void vxo_sky_procedural(const float e[3], float out[3]) {
    float ey = (g_max(e[1], 0.0f) * 0.8f + 0.2f) * 0.8f;
    float a = 1.0f - ey;
    out[0] = (a * a) * 1.1f;
    out[1] = a * 1.1f;
    out[2] = (0.6f + a * 0.4f) * 1.1f;
}

/* texture(samplerCube, d): the face of the major axis (x, then y, then z on
 * ties) and its (s, t) per the Vulkan cube-map face selection table, then
 * the build's bilinear filter (8-bit sub-texel weights) with texel
 * coordinates clamped to the face (no filtering across face edges). Faces
 * +X, -X, +Y, -Y, +Z, -Z, each size x size RGBA, rows from t = 0. */
void vxo_cube_sample(const float* faces, int32_t size, const float d[3], float out[4]) {
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d[0] >= 0.0f) { face = 0; sc = -d[2]; tc = -d[1]; }
        else { face = 1; sc = d[2]; tc = -d[1]; }
    } else if (ay >= az) {
        ma = ay;
        if (d[1] >= 0.0f) { face = 2; sc = d[0]; tc = d[2]; }
        else { face = 3; sc = d[0]; tc = -d[2]; }
    } else {
        ma = az;
        if (d[2] >= 0.0f) { face = 4; sc = d[0]; tc = -d[1]; }
        else { face = 5; sc = -d[0]; tc = -d[1]; }
    }
    float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    const float fs = (float)size;
    float x = s * fs - 0.5f, y = t * fs - 0.5f;
    x = g_clamp(x, 0.0f, fs - 1.0f);
    y = g_clamp(y, 0.0f, fs - 1.0f);
    int32_t x0, y0;
    float fx, fy;
    taa_weight(x, size, &x0, &fx);
    taa_weight(y, size, &y0, &fy);
    int32_t x1 = x0 + 1 < size ? x0 + 1 : size - 1, y1 = y0 + 1 < size ? y0 + 1 : size - 1;
    const float* F = faces + (size_t)face * size * size * 4;
    for (int c = 0; c < 4; ++c) {
        float a = F[((size_t)y0 * size + x0) * 4 + c], b = F[((size_t)y0 * size + x1) * 4 + c];
        float cc = F[((size_t)y1 * size + x0) * 4 + c], dd = F[((size_t)y1 * size + x1) * 4 + c];
        out[c] = (a * (1.0f - fx) + b * fx) * (1.0f - fy) + (cc * (1.0f - fx) + dd * fx) * fy;
    }
}

/* The reflection ray of pixel i (LightReflection.frag:79-98), in world space:
 * wcp = the hit (fine units) + normal * noise.w, direction
 * mix(normalize(reflect(-V, N)), cosineSampleHemisphere(noise.xy) with its z
 * signed by noise.z, roughness * 0.1) * (1 + noise.w * 0.5); the march starts
 * at wcp + normal. noise = the pixel's four counter-hash values (the same
 * stream as the ambient pass's shadow jitter, as both passes read the same
 * blue-noise texel). */
static void reflection_ray(const vxo_view* vw, const vxo_reflection_params* p, const float* pos, const float* nrm,
                           const uint8_t* mat, const float* palette, size_t i, float wcp[3], float wd[3]) {
    const float* pp = pos + 3 * i;
    const float* n = nrm + 3 * i;
    float rough = palette[6 * mat[i] + 3];
    uint32_t pk = vxo_pixkey(p->seed, p->frame, (uint32_t)i);
    float nz[4];
    for (int k = 0; k < 4; ++k) nz[k] = vxo_rand(pk, (uint32_t)k);
    float cam[3], vv[3], V[3], I[3], R[3];
    for (int k = 0; k < 3; ++k) cam[k] = vw->cam_pos[k] * 10.0f;
    for (int k = 0; k < 3; ++k) vv[k] = cam[k] - pp[k];
    normalize3(vv, V);
    for (int k = 0; k < 3; ++k) I[k] = -V[k];
    float dni = dot3(n, I);
    for (int k = 0; k < 3; ++k) R[k] = I[k] - (2.0f * dni) * n[k];     /* reflect(-V, N) */
    normalize3(R, wd);
    float rv[3];
    vxo_cosine_hemisphere(nz[0], nz[1], rv);
    rv[2] = rv[2] * vxo_sign(nz[2] - 0.5f);
    const float a = rough * 0.1f;
    for (int k = 0; k < 3; ++k) wd[k] = wd[k] * (1.0f - a) + rv[k] * a;
    for (int k = 0; k < 3; ++k) wcp[k] = pp[k] + n[k] * nz[3];
    for (int k = 0; k < 3; ++k) wd[k] = wd[k] * (1.0f + nz[3] * 0.5f);
}

void vxo_reflection_rays(const vxo_view* vw, const vxo_reflection_params* p, const float* pos, const float* nrm,
                         const uint8_t* mat, const float* palette, const int32_t* pixels, int32_t npixels,
                         float* rays) {
    for (int32_t j = 0; j < npixels; ++j)
        reflection_ray(vw, p, pos, nrm, mat, palette, (size_t)pixels[j], rays + 6 * (size_t)j, rays + 6 * (size_t)j + 3);
}

static void reflection_pixel(const vxo_occ* v, const vxo_view* vw, const vxo_reflection_params* p,
                             const float* pos, const float* nrm, const uint8_t* mat, const float* palette,
                             const float* depth_img, const float* light, const float* sky, int32_t sky_size,
                             size_t i, float* out, float* t_out) {
    const int32_t W = vw->width, H = vw->height;
    const float* pp = pos + 3 * i;
    const float* n = nrm + 3 * i;
    const float depth = depth_img[i];
    const float rough = palette[6 * mat[i] + 3], metal = palette[6 * mat[i] + 4];
    /* F0 = mix(vec3(0.04), albedo = 1, metallic); F = fresnelSchlickRoughness(max(N.V, 0), F0, roughness) * 5
     * (LightReflection.frag:74-86, PBR.frag:8-10) */
    const float f0 = 0.04f * (1.0f - metal) + 1.0f * metal;
    float cam[3], vv[3], V[3];
    for (int k = 0; k < 3; ++k) cam[k] = vw->cam_pos[k] * 10.0f;
    for (int k = 0; k < 3; ++k) vv[k] = cam[k] - pp[k];
    normalize3(vv, V);
    const float ndv = g_max(dot3(n, V), 0.0f);
    const float x = g_max(1.0f - ndv, 0.0f);
    const float s5 = (((x * x) * x) * x) * x;
    const float F = (f0 + (g_max(1.0f - rough, f0) - f0) * s5) * 5.0f + 0.0f;
    float amb[3] = {0.0f, 0.0f, 0.0f};
    float tt = NAN;
    if (depth < 0.999f) {
        float wcp[3], wd[3], o[3];
        reflection_ray(vw, p, pos, nrm, mat, palette, i, wcp, wd);
        for (int k = 0; k < 3; ++k) o[k] = wcp[k] + n[k];
        uint32_t nf = 0;
        int32_t hit;
        const float t = march_dist(v, p->march, o, wd, p->maxt, &nf, &hit);
        tt = t;
        /* reprojection into the light buffer, LightReflection.frag:115-133 */
        float h4[4], pv[4], pc[4];
        for (int k = 0; k < 3; ++k) h4[k] = (wcp[k] + wd[k] * t) * 0.1f;
        h4[3] = 1.0f;
        mat_vec(vw->view, h4, pv);
        pv[3] = 1.0f;
        mat_vec(vw->proj, pv, pc);
        const float lin = (pc[3] - NEAR_PLANE) / (FAR_PLANE - NEAR_PLANE);
        const float u = ((pc[0] / pc[3]) * 1.0f) * 0.5f + 0.5f;
        const float vv2 = ((pc[1] / pc[3]) * -1.0f) * 0.5f + 0.5f;
        float ds;
        taa_tex dt = {W, H, 0, 1, depth_img, NULL, NULL};
        taa_sample(&dt, u, vv2, 1, &ds);
        if (t == p->maxt) {
            if (sky) {
                float c4[4];
                vxo_cube_sample(sky, sky_size, wd, c4);
                amb[0] = c4[0]; amb[1] = c4[1]; amb[2] = c4[2];
            } else {
                vxo_sky_procedural(wd, amb);
            }
        } else if (lin > ds - 0.001f && lin < ds + 0.001f) {
            float l4[4];
            taa_tex lt = {W, H, 0, 4, light, NULL, NULL};
            taa_sample(&lt, u, vv2, 4, l4);
            amb[0] = l4[0]; amb[1] = l4[1]; amb[2] = l4[2];
        }
    }
    for (int k = 0; k < 3; ++k) out[4 * i + k] = (amb[k] * F) * (1.0f - rough);
    out[4 * i + 3] = F;
    if (t_out) t_out[i] = tt;
}

void vxo_reflection(const vxo_occ* v, const vxo_view* vw, const vxo_reflection_params* p, const float* pos,
                    const float* nrm, const uint8_t* mat, const float* palette, const float* depth,
                    const float* light, const float* sky, int32_t sky_size, float* out, float* t_out,
                    int32_t nthreads) {
    const int32_t npx = vw->width * vw->height;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < npx; ++i)
        reflection_pixel(v, vw, p, pos, nrm, mat, palette, depth, light, sky, sky_size, (size_t)i, out, t_out);
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* Screen-space occlusion: calculateOcclusion / screenspaceOcclusion,        */
/* LightAmbient.frag:59-79,89-109                                            */
/* ------------------------------------------------------------------------ */
static const float OCC_STRENGTH = 3.0f;     /* LightAmbient.frag:9 */
static const float OCC_SIZE = 0.2f;         /* :10 */
static const float OCC_THICKNESS = 0.2f;    /* :11 */

/* The sample's noise (getNoise(i), LightAmbient.frag:44-47): the pixel's
 * counter hash under its own key, indices 4 i + {0, 1, 2}. */
static uint32_t ssao_key(uint32_t seed, int32_t frame, uint32_t pixel) {
    return vxo_pixkey(seed ^ 0x632be5abu, frame, pixel);
}

static float ssao_pixel(const vxo_view* vw, const float* nrm, const float* depth_img, uint32_t seed, size_t i) {
    const int32_t W = vw->width, H = vw->height;
    const int32_t px = (int32_t)(i % (size_t)W), py = (int32_t)(i / (size_t)W);
    /* the view-space normal: (View * vec4(normal, 0)).xyz (LightAmbient.frag:192) */
    float n4[4] = {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.0f}, N[4];
    mat_vec(vw->view, n4, N);
    float t0[3], T[3], B[3], cb[3];
    if (fabsf(N[2]) > 0.5f) { t0[0] = 0.0f; t0[1] = -N[2]; t0[2] = N[1]; }
    else { t0[0] = -N[1]; t0[1] = N[0]; t0[2] = 0.0f; }
    normalize3(t0, T);
    cb[0] = N[1] * T[2] - N[2] * T[1];
    cb[1] = N[2] * T[0] - N[0] * T[2];
    cb[2] = N[0] * T[1] - N[1] * T[0];
    normalize3(cb, B);
    /* In.FarVec at the pixel centre (LightAmbient.vert:32-36, 43) */
    const float u = ((float)px + 0.5f) / vw->res[0], v = ((float)py + 0.5f) / vw->res[1];
    float c4[4] = {u * 2.0f - 1.0f, 1.0f - v * 2.0f, 1.0f, 1.0f}, f4[4];
    mat_vec(vw->inv_proj, c4, f4);
    const float fv[3] = {f4[0] / f4[3], f4[1] / f4[3], f4[2] / f4[3]};
    const float depth = depth_img[i];
    const float sc = depth * (1.0f + 1.0f / FAR_PLANE) + NEAR_PLANE / FAR_PLANE;
    const float pos[3] = {fv[0] * sc, fv[1] * sc, fv[2] * sc};
    const float size = OCC_SIZE * (1.0f + depth * 0.0f);
    const uint32_t pk = ssao_key(seed, vw->frame, (uint32_t)i);
    taa_tex dt = {W, H, 0, 1, depth_img, NULL, NULL};
    float occ = 0.0f;
    for (int s = 0; s < 4; ++s) {
        const float nx = vxo_rand(pk, (uint32_t)(4 * s)), ny = vxo_rand(pk, (uint32_t)(4 * s + 1));
        const float dist = vxo_rand(pk, (uint32_t)(4 * s + 2));
        float rv[3], dir[3], dn[3];
        vxo_cosine_hemisphere(nx, ny, rv);
        for (int k = 0; k < 3; ++k) dir[k] = (T[k] * rv[0] + B[k] * rv[1]) + N[k] * rv[2];
        normalize3(dir, dn);
        /* screenspaceOcclusion(pos, normalize(dir) * size, dist) */
        float mid[4];
        for (int k = 0; k < 3; ++k) mid[k] = pos[k] + (dn[k] * size) * dist;
        mid[3] = 1.0f;
        float mp[4];
        mat_vec(vw->proj, mid, mp);
        const float sd = (mp[3] - NEAR_PLANE) / (FAR_PLANE - NEAR_PLANE);
        const float su = ((mp[0] / mp[3]) * 1.0f) * 0.5f + 0.5f;
        const float sv = ((mp[1] / mp[3]) * -1.0f) * 0.5f + 0.5f;
        float mn;
        taa_sample(&dt, su, sv, 1, &mn);
        const float mx = mn + OCC_THICKNESS / FAR_PLANE;
        float o = 0.0f;
        if (g_clamp(su, 0.0f, 1.0f) == su && g_clamp(sv, 0.0f, 1.0f) == sv && mn < sd && sd < mx) {
            /* smoothstep(maxDepth, minDepth, d) with its edges reversed, as written */
            const float tt = g_clamp((sd - mx) / (mn - mx), 0.0f, 1.0f);
            o = ((tt * tt) * (3.0f - 2.0f * tt)) * dist;
        }
        occ = occ + o;
    }
    occ = occ / 4.0f;
    occ = occ * OCC_STRENGTH;
    return g_clamp(1.0f - occ, 0.0f, 1.0f);
}

void vxo_ssao(const vxo_view* vw, const float* nrm, const uint8_t* mat, const float* depth, uint32_t seed,
              float* out, int32_t nthreads) {
    const int32_t npx = vw->width * vw->height;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int32_t i = 0; i < npx; ++i)
        out[i] = (mat[i] != 0 && depth[i] < 0.999f) ? ssao_pixel(vw, nrm, depth, seed, (size_t)i) : 1.0f;
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
static inline uint32_t vox_material(uint32_t seed, int32_t x, int32_t y, int32_t z) {
    uint32_t h = vxo_hash(seed ^ 0x68bc21ebu);
    h = vxo_hash(h + (uint32_t)x);
    h = vxo_hash(h + (uint32_t)y);
    h = vxo_hash(h + (uint32_t)z);
    return 1u + h % 255u;
}
static inline float lattice(uint32_t seed, int32_t ix, int32_t iz, int32_t octave) {
    uint32_t h = vxo_hash(seed + 0x9e3779b9u * (uint32_t)(octave + 1));
    h = vxo_hash(h + (uint32_t)ix);
    h = vxo_hash(h + (uint32_t)iz);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
static float value_noise(uint32_t seed, float fx, float fz, int32_t octave) {
    float x0 = floorf(fx), z0 = floorf(fz);
    int32_t ix = vxo_f2i(x0), iz = vxo_f2i(z0);
    float tx = fx - x0, tz = fz - z0;
    float sx = (tx * tx) * (3.0f - 2.0f * tx);
    float sz = (tz * tz) * (3.0f - 2.0f * tz);
    float v00 = lattice(seed, ix, iz, octave), v10 = lattice(seed, ix + 1, iz, octave);
    float v01 = lattice(seed, ix, iz + 1, octave), v11 = lattice(seed, ix + 1, iz + 1, octave);
    float a = v00 + (v10 - v00) * sx;
    float b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sz;
}
static float terrain_height(uint32_t seed, int32_t x, int32_t z, float base_freq, float fill, int32_t ny) {
    float fx = (float)x + 0.5f, fz = (float)z + 0.5f;
    float total = 0.0f, freq = base_freq, amp = 1.0f, maxv = 0.0f;
    for (int32_t o = 0; o < 4; ++o) {
        total = total + value_noise(seed, fx * freq, fz * freq, o) * amp;
        maxv = maxv + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return ((total / maxv) + (fill - 0.5f)) * (float)ny;
}
void vxo_gen_terrain(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed, float fill,
                     int32_t nthreads) {
    float base_freq = 4.0f / (float)(nx > nz ? nx : nz);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int32_t z = 0; z < nz; ++z)
        for (int32_t x = 0; x < nx; ++x) {
            float h = terrain_height(seed, x, z, base_freq, fill, ny);
            for (int32_t y = 0; y < ny; ++y) {
                size_t i = (size_t)x + (size_t)y * nx + (size_t)z * nx * ny;
                vol[i] = ((float)y < h) ? (uint8_t)vox_material(seed, x, y, z) : (uint8_t)0;
            }
        }
    (void)nthreads;
}
void vxo_gen_boxes(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed,
                   int32_t nboxes, int32_t emin, int32_t emax, int32_t grounded) {
    uint32_t base = vxo_hash(seed ^ 0xb5297a4du);
    const int32_t dims[3] = {nx, ny, nz};
    for (int32_t b = 0; b < nboxes; ++b) {
        uint32_t r[6];
        for (int k = 0; k < 6; ++k) r[k] = vxo_hash(base + (uint32_t)b * 8u + (uint32_t)k);
        int32_t e[3], o[3];
        for (int k = 0; k < 3; ++k) {
            e[k] = emin + (int32_t)(r[k] % (uint32_t)(emax - emin + 1));
            if (e[k] > dims[k]) e[k] = dims[k];
            o[k] = (int32_t)(r[3 + k] % (uint32_t)(dims[k] - e[k] + 1));
        }
        if (grounded) o[1] = 0;
        for (int32_t z = o[2]; z < o[2] + e[2]; ++z)
            for (int32_t y = o[1]; y < o[1] + e[1]; ++y)
                for (int32_t x = o[0]; x < o[0] + e[0]; ++x)
                    vol[(size_t)x + (size_t)y * nx + (size_t)z * nx * ny] = (uint8_t)vox_material(seed, x, y, z);
    }
}
