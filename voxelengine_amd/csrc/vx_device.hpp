// vx_device.hpp -- gfx950 device-side scalar semantics and occupancy access.
//
// Every float operation here is a single IEEE op (+,-,*,/,sqrt) in a fixed
// order; the library is compiled with -ffp-contract=off, correctly rounded
// fp32 divide/sqrt and fp32 denormals preserved, so results are identical
// bit for bit to a strict IEEE host evaluation of the same expressions.
// GLSL builtins follow the conventions pinned for this path:
//   sign(+-0) = +0, sign(NaN) = 0;  step(e,x) = x < e ? 0 : 1;
//   min(x,y) = y < x ? y : x;  max(x,y) = x < y ? y : x;
//   mod(x,y) = x - y*floor(x/y);  ivec3() truncates toward zero, saturates
//   out-of-range values, NaN -> 0;  out-of-range texel fetches read 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

#define VXD __device__ __forceinline__

// ivec3(x): truncation toward zero, saturating to the int32 range, NaN -> 0.
// That is v_cvt_i32_f32 exactly (checked on NaN, infinities, values at and
// past the range and denormals); a C++ cast would leave those cases undefined.
VXD int32_t f2i(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
VXD float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
VXD float gstep(float e, float x) { return x < e ? 0.0f : 1.0f; }
VXD float gmin(float x, float y) { return y < x ? y : x; }
VXD float gmax(float x, float y) { return x < y ? y : x; }
VXD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
VXD float gmod(float x, float y) { return x - y * floorf(x / y); }

struct f3 { float x, y, z; };
VXD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
VXD float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VXD f3 normalize3(f3 v) {
    float l = sqrtf(dot3(v, v));
    return mk3(v.x / l, v.y / l, v.z / l);
}
VXD f3 load3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Column-major mat4 * vec4: r_i = ((m0i*x + m1i*y) + m2i*z) + m3i*w.
VXD void mat_vec(const float* m, const float v[4], float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[0 + r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

// sin/cos: Cody-Waite reduction by pi/2 (three-part constant) and the Cephes
// single-precision polynomials on [-pi/4, pi/4]; only +,-,*.
VXD void sincos_vx(float x, float* s, float* c) {
    const float two_over_pi = 0.636619772367581343f;
    const float c1 = 1.5703125f;
    const float c2 = 4.837512969970703125e-4f;
    const float c3 = 7.54978995489188216e-8f;
    float kf = floorf(x * two_over_pi + 0.5f);
    int32_t k = f2i(kf);
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

// log2 / exp2 / pow for the light attenuation terms (the driver's pow is
// replaced): range reduction plus the Cephes single-precision log and exp2
// polynomials, only + - * /, the same expressions as the oracle's.
VXD float log2_vx(float x) {
    int32_t e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b = __float_as_uint(x);
    e += (int32_t)((b >> 23) & 0xffu) - 127;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
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
VXD float exp2_vx(float t) {
    if (!(t == t)) return t;
    if (t >= 128.0f) return __uint_as_float(0x7f800000u);
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
    r = r * __uint_as_float((uint32_t)(k1 + 127) << 23);
    return r * __uint_as_float((uint32_t)(k2 + 127) << 23);
}
VXD float pow_vx(float x, float y) {
    const float inf = __uint_as_float(0x7f800000u);
    if (!(x == x) || !(y == y) || x < 0.0f) return __uint_as_float(0x7fc00000u);
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : inf;
    if (x == inf) return y > 0.0f ? inf : 0.0f;
    return exp2_vx(y * log2_vx(x));
}

// Counter-hash RNG (lowbias32) replacing the blue-noise texture.
VXD uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
VXD uint32_t pixkey(uint32_t seed, int32_t frame, uint32_t pixel) {
    uint32_t key = hash32(seed ^ hash32((uint32_t)frame + 0x9e3779b9u));
    return hash32(key ^ hash32(pixel));
}
VXD float rand01(uint32_t pk, uint32_t idx) {
    return (float)(hash32(pk + idx * 0x9e3779b9u) >> 8) * (1.0f / 16777216.0f);
}

// cosineSampleHemisphere (LightAmbient.frag:81-87).
VXD f3 cosine_hemisphere(float u1, float u2) {
    float r = sqrtf(u1);
    float theta = 6.283f * u2;
    float sn, cs;
    sincos_vx(theta, &sn, &cs);
    return mk3(r * cs, r * sn, sqrtf(gmax(0.0f, 1.0f - u1)));
}

// Tangent frame of calculateAmbientIrradiance (LightAmbient.frag:112-113).
VXD void tangent_frame(f3 n, f3& t, f3& b) {
    if (fabsf(n.z) > 0.5f) t = mk3(0.0f, -n.z, n.y);
    else t = mk3(-n.y, n.x, 0.0f);
    b = mk3(n.y * t.z - n.z * t.y, n.z * t.x - n.x * t.z, n.x * t.y - n.y * t.x);
}

// ---------------------------------------------------------------------------
// Occupancy volume in brick layout.
//   fine voxel f -> cell f>>2 (u64 = 4x4x4 fine voxels, bit (x&3)|(y&3)<<2|
//   (z&3)<<4) -> brick f>>3 (64 B = 2x2x2 cells, cell (cx&1)|(cy&1)<<1|(cz&1)<<2).
//   Bricks X-fastest. A reference texel byte (2x2x2 fine voxels, bit
//   (x&1)|(y&1)<<1|(z&1)<<2, ShadowVoxSystem.cpp:85) occupies cell bits
//   base + {0,1,4,5,16,17,20,21} with base = 2(tx&1) + 8(ty&1) + 32(tz&1).
//   The brick grid is padded by one so texel index s (the first out-of-range
//   one, reachable by the DDA bound 2*s+1) stays inside and reads 0.
// Coarse level k (1..3): 1 bit per (8^k)^3 fine voxels, u64 words of 4x4x4
// bits, bit (cx&3)|(cy&3)<<2|(cz&3)<<4, words X-fastest.
// ---------------------------------------------------------------------------
struct OccView {
    const uint64_t* cells;
    const uint64_t* lvl[3];
    int32_t sx, sy, sz;        // texel dims
    int32_t bx, by, bz;        // brick grid (padded)
    int32_t wx[3], wy[3], wz[3];  // level word grids
    int32_t nlevels;
};

VXD uint32_t sub3(int32_t x, int32_t y, int32_t z) {
    return (uint32_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
}
VXD size_t brick_index(const OccView& V, int32_t bx, int32_t by, int32_t bz) {
    return ((size_t)bz * V.by + by) * V.bx + bx;
}
// Cell word index of fine voxel f (f inside the padded brick grid).
VXD size_t cell_index(const OccView& V, int32_t fx, int32_t fy, int32_t fz) {
    return brick_index(V, fx >> 3, fy >> 3, fz >> 3) * 8 + sub3(fx >> 2, fy >> 2, fz >> 2);
}
// cell_index in 32 bits, for volumes of fewer than 2^31 cell words.
VXD uint32_t cell_index32(const OccView& V, int32_t fx, int32_t fy, int32_t fz) {
    return ((((uint32_t)(fz >> 3) * (uint32_t)V.by + (uint32_t)(fy >> 3)) * (uint32_t)V.bx + (uint32_t)(fx >> 3)) << 3) +
           sub3(fx >> 2, fy >> 2, fz >> 2);
}
// Bit of fine voxel f inside its cell word.
VXD uint32_t cell_bit(int32_t fx, int32_t fy, int32_t fz) {
    return (uint32_t)((fx & 3) | ((fy & 3) << 2) | ((fz & 3) << 4));
}
// Cell-bit offset of texel t's byte (see layout above).
VXD uint32_t texel_base(int32_t tx, int32_t ty, int32_t tz) {
    return (uint32_t)(((tx & 1) << 1) | ((ty & 1) << 3) | ((tz & 1) << 5));
}
// Spread a reference texel byte to its cell-bit pattern (relative to texel_base).
VXD uint32_t spread_byte(uint32_t b) {
    return (b & 3u) | ((b & 0xcu) << 2) | ((b & 0x30u) << 12) | ((b & 0xc0u) << 14);
}
// Whether the volume's cell words are indexed in 32 bits (fewer than 2^32
// cell words: 2^29 bricks of 8).
VXD bool volume_narrow(const OccView& V) {
    return (uint64_t)(uint32_t)V.bx * (uint32_t)V.by * (uint32_t)V.bz < (1ull << 29);
}
// Cell-word index of texel t and whether t lies inside [0, s); texels
// outside read cell word 0 (the word is read unconditionally, so a march step
// carries no branch for it) and decode to 0. NARROW: volume_narrow(V).
template <bool NARROW>
VXD size_t texel_word_n(const OccView& V, int32_t tx, int32_t ty, int32_t tz, bool& in) {
    in = (uint32_t)tx < (uint32_t)V.sx && (uint32_t)ty < (uint32_t)V.sy && (uint32_t)tz < (uint32_t)V.sz;
    if (NARROW) {
        // brick (tx, ty, tz) >> 2 of a brick grid under 2^29 bricks of at most
        // 1025 per axis: 24-bit multiplies are exact for a texel inside (and
        // a texel outside reads word 0)
        // (two v_mad_u32_u24: the compiler's own choice for the same sums is a
        // quarter-rate v_mad_u64_u32)
        uint32_t a, b;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"((uint32_t)tz >> 2), "v"((uint32_t)V.by), "v"((uint32_t)ty >> 2));
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(b) : "v"(a), "v"((uint32_t)V.bx), "v"((uint32_t)tx >> 2));
        uint32_t c = (b << 3) | sub3(tx >> 1, ty >> 1, tz >> 1);
        // computed for every lane, then selected (kept out of a branch)
        asm volatile("" : "+v"(c));
        return in ? c : 0u;
    }
    size_t c = brick_index(V, tx >> 2, ty >> 2, tz >> 2) * 8 + sub3(tx >> 1, ty >> 1, tz >> 1);
    asm volatile("" : "+v"(c));
    return in ? c : 0;
}
VXD size_t texel_word(const OccView& V, int32_t tx, int32_t ty, int32_t tz, bool& in) {
    return volume_narrow(V) ? texel_word_n<true>(V, tx, ty, tz, in) : texel_word_n<false>(V, tx, ty, tz, in);
}
// Reference byte of texel t from its cell word w (texel_word's `in`).
VXD uint32_t texel_from_word(uint64_t w, int32_t tx, int32_t ty, int32_t tz, bool in) {
    const uint32_t x = (uint32_t)(w >> texel_base(tx, ty, tz));
    const uint32_t b = (x & 3u) | ((x >> 2) & 0xcu) | ((x >> 12) & 0x30u) | ((x >> 14) & 0xc0u);
    return in ? b : 0u;
}
// Reference byte of texel t (0 when outside [0, s)), texelFetch semantics.
VXD uint32_t texel_byte(const OccView& V, int32_t tx, int32_t ty, int32_t tz) {
    bool in;
    const size_t ci = texel_word(V, tx, ty, tz, in);
    return texel_from_word(V.cells[ci], tx, ty, tz, in);
}
VXD bool level_bit(const OccView& V, int32_t lvl, int32_t cx, int32_t cy, int32_t cz) {
    size_t wi = ((size_t)(cz >> 2) * V.wy[lvl] + (cy >> 2)) * V.wx[lvl] + (cx >> 2);
    uint32_t b = (uint32_t)((cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4));
    return (V.lvl[lvl][wi] >> b) & 1ull;
}

}  // namespace vx
