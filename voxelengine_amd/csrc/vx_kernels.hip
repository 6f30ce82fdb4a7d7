// vx_kernels.hip -- gfx950 kernels of the voxel-traversal lighting path.
//
// Kernels:
//   scene generation   k_terrain_heights, k_terrain_fill, k_boxes (build-defined inputs)
//   model mips         k_build_mip            VoxAsset.cpp:16-53
//   occupancy volume   k_fill_u64, k_bytes_to_cells, k_cells_to_bytes, k_footprint,
//                      k_level1, k_level_up   ShadowVoxSystem.cpp:82-201
//   primary hit        k_primary              GeometryVoxel.frag:49-183
//   traversal          k_trace, k_trace_rays  Light.frag:14-217, LightAmbient.frag:111-175,
//                                             LightReflection.frag:86-113
//   shading            k_shade_fwd, k_shade_bwd, k_shade_bwd_finish
//                                             LightAmbient.frag:178-214, PBR.frag:3-68
#include "vx_internal.hpp"

namespace vx {

// ===========================================================================
// Synthetic scenes. Terrain: y < H(x,z), H = 4-octave value-noise fBm with
// the octave structure of Noise::GetOctave (Noise.cpp:93-107). Materials are
// a hash of the voxel position (uniform 1..255), so box overlays commute.
// ===========================================================================
VXD uint32_t vox_material(uint32_t seed, int32_t x, int32_t y, int32_t z) {
    uint32_t h = hash32(seed ^ 0x68bc21ebu);
    h = hash32(h + (uint32_t)x);
    h = hash32(h + (uint32_t)y);
    h = hash32(h + (uint32_t)z);
    return 1u + h % 255u;
}
VXD float lattice(uint32_t seed, int32_t ix, int32_t iz, int32_t octave) {
    uint32_t h = hash32(seed + 0x9e3779b9u * (uint32_t)(octave + 1));
    h = hash32(h + (uint32_t)ix);
    h = hash32(h + (uint32_t)iz);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
VXD float value_noise(uint32_t seed, float fx, float fz, int32_t octave) {
    float x0 = floorf(fx), z0 = floorf(fz);
    int32_t ix = f2i(x0), iz = f2i(z0);
    float tx = fx - x0, tz = fz - z0;
    float sx = (tx * tx) * (3.0f - 2.0f * tx);
    float sz = (tz * tz) * (3.0f - 2.0f * tz);
    float v00 = lattice(seed, ix, iz, octave), v10 = lattice(seed, ix + 1, iz, octave);
    float v01 = lattice(seed, ix, iz + 1, octave), v11 = lattice(seed, ix + 1, iz + 1, octave);
    float a = v00 + (v10 - v00) * sx;
    float b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sz;
}

__global__ void k_terrain_heights(float* h, int32_t nx, int32_t ny, int32_t nz, uint32_t seed, float fill) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * nz) return;
    int32_t x = i % nx, z = i / nx;
    float base_freq = 4.0f / (float)(nx > nz ? nx : nz);
    float fx = (float)x + 0.5f, fz = (float)z + 0.5f;
    float total = 0.0f, freq = base_freq, amp = 1.0f, maxv = 0.0f;
    for (int32_t o = 0; o < 4; ++o) {
        total = total + value_noise(seed, fx * freq, fz * freq, o) * amp;
        maxv = maxv + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    h[i] = ((total / maxv) + (fill - 0.5f)) * (float)ny;
}

// One thread per 4 consecutive x voxels (one u32 store).
__global__ void k_terrain_fill(uint8_t* vol, const float* h, int32_t nx, int32_t ny, int32_t nz, uint32_t seed,
                               int32_t use_terrain) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nq = ((size_t)nx * ny * nz) / 4;
    if (q >= nq) return;
    size_t i0 = q * 4;
    int32_t x = (int32_t)(i0 % nx);
    size_t r = i0 / nx;
    int32_t y = (int32_t)(r % ny), z = (int32_t)(r / ny);
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t m = 0;
        if (use_terrain && (float)y < h[(size_t)z * nx + x + k]) m = vox_material(seed, x + k, y, z);
        word |= m << (8 * k);
    }
    reinterpret_cast<uint32_t*>(vol)[q] = word;
}

__global__ void k_boxes(uint8_t* vol, int32_t nx, int32_t ny, int32_t nz, uint32_t seed, int32_t b0,
                        int32_t emin, int32_t emax, int32_t grounded) {
    int32_t b = b0 + blockIdx.y;
    uint32_t base = hash32(seed ^ 0xb5297a4du);
    uint32_t r[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = hash32(base + (uint32_t)b * 8u + (uint32_t)k);
    const int32_t dims[3] = {nx, ny, nz};
    int32_t e[3], o[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e[k] = emin + (int32_t)(r[k] % (uint32_t)(emax - emin + 1));
        if (e[k] > dims[k]) e[k] = dims[k];
        o[k] = (int32_t)(r[3 + k] % (uint32_t)(dims[k] - e[k] + 1));
    }
    if (grounded) o[1] = 0;
    int32_t nvox = e[0] * e[1] * e[2];
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += gridDim.x * blockDim.x) {
        int32_t x = o[0] + v % e[0];
        int32_t y = o[1] + (v / e[0]) % e[1];
        int32_t z = o[2] + v / (e[0] * e[1]);
        vol[(size_t)x + (size_t)y * nx + (size_t)z * nx * ny] = (uint8_t)vox_material(seed, x, y, z);
    }
}

// ===========================================================================
// Model mips: VoxAsset::Upload (VoxAsset.cpp:26-47): child scan vi = 0..8
// (vi = 8 re-reads child 0), first non-zero material wins.
// ===========================================================================
__global__ void k_build_mip(const uint8_t* parent, int32_t px, int32_t py, int32_t pz, uint8_t* out) {
    int32_t mx = px > 1 ? px / 2 : 1, my = py > 1 ? py / 2 : 1, mz = pz > 1 ? pz / 2 : 1;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)mx * my * mz) return;
    int32_t x = (int32_t)(i % mx);
    int32_t y = (int32_t)((i / mx) % my);
    int32_t z = (int32_t)(i / ((size_t)mx * my));
    uint8_t vox = 0;
    for (int vi = 0; vi < 9; ++vi) {
        uint8_t u = parent[(size_t)(2 * x + ((vi >> 0) & 1)) + (size_t)(2 * y + ((vi >> 1) & 1)) * px +
                           (size_t)(2 * z + ((vi >> 2) & 1)) * px * py];
        if (u) { vox = u; break; }
    }
    out[i] = vox;
}

// ===========================================================================
// Occupancy volume: layout conversion, rebuild, coarse levels.
// ===========================================================================
__global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Reference byte layout -> brick layout over a texel region (one thread per texel).
__global__ void k_bytes_to_cells(uint64_t* cells, const uint8_t* bytes, OccView V, int32_t rx, int32_t ry,
                                 int32_t rz, int32_t rw, int32_t rh, int32_t rd) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rw * rh * rd) return;
    int32_t tx = rx + (int32_t)(i % rw);
    int32_t ty = ry + (int32_t)((i / rw) % rh);
    int32_t tz = rz + (int32_t)(i / ((size_t)rw * rh));
    uint8_t b = bytes[(size_t)tx + (size_t)ty * V.sx + (size_t)tz * V.sx * V.sy];
    size_t ci = brick_index(V, tx >> 2, ty >> 2, tz >> 2) * 8 + sub3(tx >> 1, ty >> 1, tz >> 1);
    uint32_t base = texel_base(tx, ty, tz);
    // Eight texels share a cell word; each texel's bits lie in one 32-bit half.
    uint32_t* w32 = reinterpret_cast<uint32_t*>(cells + ci) + (base >> 5);
    uint32_t s32 = base & 31u;
    atomicAnd(w32, ~(0x330033u << s32));
    atomicOr(w32, spread_byte(b) << s32);
}

// Brick layout -> reference byte layout (one thread per texel, whole volume).
__global__ void k_cells_to_bytes(uint8_t* bytes, const uint64_t* cells, OccView V) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)V.sx * V.sy * V.sz) return;
    int32_t tx = (int32_t)(i % V.sx);
    int32_t ty = (int32_t)((i / V.sx) % V.sy);
    int32_t tz = (int32_t)(i / ((size_t)V.sx * V.sy));
    OccView W = V;
    W.cells = cells;
    bytes[i] = (uint8_t)texel_byte(W, tx, ty, tz);
}

// One footprint loop of ShadowVoxSystem::OnUpdate / OnVoxDestroyed
// (ShadowVoxSystem.cpp:142-155, 164-177, 28-40) for a batch of operations that
// commute (same value, or disjoint footprints). blockIdx.y = operation.
// A thread takes four consecutive x voxels of a row; voxels that land in
// the same 32-bit cell word share one atomic. The per-voxel arithmetic is the
// reference's, voxel by voxel. A thread steps through its quads by a fixed
// stride, so their row and slice advance by carries instead of divisions.
__global__ void k_footprint(uint64_t* cells, OccView V, const FootprintOp* ops, int32_t* aabb, int32_t thr) {
    const FootprintOp op = ops[blockIdx.y];
    const uint32_t qx = ((uint32_t)op.sx + 3u) >> 2, sy = (uint32_t)op.sy;
    const uint64_t nq = (uint64_t)qx * sy * (uint32_t)op.sz;
    // rows of whole quads and a 4-byte aligned model: one dword per quad
    const bool dword = ((uintptr_t)op.model & 3u) == 0 && (op.sx & 3) == 0;
    const bool narrow_cells = (uint64_t)V.bx * (uint64_t)V.by * (uint64_t)V.bz * 16u < (1ull << 31);
    int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    uint32_t* words = reinterpret_cast<uint32_t*>(cells);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // (xq, y, z) of quad q, and the stride as whole rows and slices plus a remainder
    uint32_t xq, y, z;
    {
        const uint64_t r = q / qx;
        xq = (uint32_t)(q - r * qx);
        y = (uint32_t)(r % sy);
        z = (uint32_t)(r / sy);
    }
    const uint64_t srow = stride / qx;
    const uint32_t sxq = (uint32_t)(stride - srow * qx), sy_ = (uint32_t)(srow % sy);
    const uint32_t sz_ = (uint32_t)(srow / sy);
    for (; q < nq; q += stride) {
        const uint32_t cxq = xq, cy = y, cz = z;
        // the next quad of this thread
        xq += sxq;
        const uint32_t carry = xq >= qx ? 1u : 0u;
        xq -= carry ? qx : 0u;
        y += sy_ + carry;
        const uint32_t carry2 = y >= sy ? 1u : 0u;
        y -= carry2 ? sy : 0u;
        z += sz_ + carry2;

        uint32_t m4 = 0;
        const uint32_t nvalid = min(4u, (uint32_t)op.sx - 4u * cxq);
        if (dword) {
            m4 = reinterpret_cast<const uint32_t*>(op.model)[q];
            if ((int32_t)(m4 & 255u) < thr && (int32_t)((m4 >> 8) & 255u) < thr &&
                (int32_t)((m4 >> 16) & 255u) < thr && (int32_t)(m4 >> 24) < thr)
                continue;
        } else {
            const uint8_t* p = op.model + (((size_t)cz * sy + cy) * (uint32_t)op.sx + 4u * cxq);
            for (uint32_t j = 0; j < nvalid; ++j) m4 |= (uint32_t)p[j] << (8 * j);
        }
        const float fy = (float)cy, fz = (float)cz;
        size_t wa[4];
        uint32_t wm[4];
        bool has[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            has[j] = false;
            wa[j] = 0;
            wm[j] = 0;
            if ((uint32_t)j >= nvalid || (int32_t)((m4 >> (8 * j)) & 255u) < thr) continue;
            const float fx = (float)(int32_t)(4 * cxq + j);
            int32_t f[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float wp = ((op.o[k] + fx * op.dx[k]) + fy * op.dy[k]) + fz * op.dz[k];
                f[k] = f2i(wp * 10.0f);
                mn[k] = min(mn[k], f[k]);
                mx[k] = max(mx[k], f[k]);
            }
            // SetVolumeAt (ShadowVoxSystem.cpp:82-94): out-of-range writes are dropped.
            if ((uint32_t)f[0] >= 2u * (uint32_t)V.sx || (uint32_t)f[1] >= 2u * (uint32_t)V.sy ||
                (uint32_t)f[2] >= 2u * (uint32_t)V.sz)
                continue;
            const uint32_t bit = cell_bit(f[0], f[1], f[2]);
            wa[j] = (narrow_cells ? (size_t)cell_index32(V, f[0], f[1], f[2]) : cell_index(V, f[0], f[1], f[2])) * 2 +
                    (bit >> 5);
            wm[j] = 1u << (bit & 31u);
            has[j] = true;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!has[j]) continue;
#pragma unroll
            for (int k = j + 1; k < 4; ++k)
                if (has[k] && wa[k] == wa[j]) {
                    wm[j] |= wm[k];
                    has[k] = false;
                }
            if (op.value) atomicOr(words + wa[j], wm[j]);
            else atomicAnd(words + wa[j], ~wm[j]);
        }
    }
    // AABB of the footprint (ShadowVoxSystem.cpp:148-149): wave reduction, then
    // at most one atomic per bound and wave. Every wave targets the same six
    // words, whose atomics serialize, so a bound that the slot already holds is
    // skipped; the slot is read without an atomic, and a stale value is only
    // ever larger for a min (smaller for a max), which costs an atomic, never a bound.
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int32_t a = mn[k], b = mx[k];
        for (int off = 32; off > 0; off >>= 1) {
            a = min(a, __shfl_xor(a, off));
            b = max(b, __shfl_xor(b, off));
        }
        if ((threadIdx.x & 63) == 0) {
            int32_t* lo = &aabb[op.inst * 6 + k];
            int32_t* hi = &aabb[op.inst * 6 + 3 + k];
            if (a != INT32_MAX && a < __builtin_nontemporal_load(lo)) atomicMin(lo, a);
            if (b != INT32_MIN && b > __builtin_nontemporal_load(hi)) atomicMax(hi, b);
        }
    }
}

// Level 1: one bit per brick (64 B). One wave per level word: lane l handles
// brick (4*wx + (l&3), 4*wy + ((l>>2)&3), 4*wz + (l>>4)); ballot packs the bits.
__global__ void k_level1(uint64_t* words, const uint64_t* cells, OccView V, int32_t w0x, int32_t w0y, int32_t w0z,
                         int32_t nwx, int32_t nwy, int32_t nwz) {
    int32_t wave = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int32_t lane = threadIdx.x & 63;
    if (wave >= nwx * nwy * nwz) return;
    int32_t wx = w0x + wave % nwx, wy = w0y + (wave / nwx) % nwy, wz = w0z + wave / (nwx * nwy);
    int32_t bx = 4 * wx + (lane & 3), by = 4 * wy + ((lane >> 2) & 3), bz = 4 * wz + (lane >> 4);
    bool occ = false;
    if (bx < V.bx && by < V.by && bz < V.bz) {
        const uint64_t* b = cells + brick_index(V, bx, by, bz) * 8;
        uint64_t acc = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) acc |= b[c];
        occ = acc != 0;
    }
    uint64_t m = __ballot(occ);
    if (lane == 0) words[((size_t)wz * V.wy[0] + wy) * V.wx[0] + wx] = m;
}

// Level k >= 2 from level k-1: cell (cx,cy,cz) covers 2x2x2 words of level k-1.
__global__ void k_level_up(uint64_t* words, const uint64_t* lower, int32_t lwx, int32_t lwy, int32_t lwz,
                           int32_t uwx, int32_t uwy, int32_t uwz, int32_t w0x, int32_t w0y, int32_t w0z,
                           int32_t nwx, int32_t nwy, int32_t nwz) {
    int32_t wave = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int32_t lane = threadIdx.x & 63;
    if (wave >= nwx * nwy * nwz) return;
    int32_t wx = w0x + wave % nwx, wy = w0y + (wave / nwx) % nwy, wz = w0z + wave / (nwx * nwy);
    int32_t cx = 4 * wx + (lane & 3), cy = 4 * wy + ((lane >> 2) & 3), cz = 4 * wz + (lane >> 4);
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int32_t lx = 2 * cx + (j & 1), ly = 2 * cy + ((j >> 1) & 1), lz = 2 * cz + (j >> 2);
        if (lx < lwx && ly < lwy && lz < lwz) acc |= lower[((size_t)lz * lwy + ly) * lwx + lx];
    }
    uint64_t m = __ballot(acc != 0);
    if (lane == 0) words[((size_t)wz * uwy + wy) * uwx + wx] = m;
}

// ===========================================================================
// Traversal
// ===========================================================================

// Coarse emptiness of brick (bx,by,bz): any level bit clear means every fine
// voxel of the brick is empty (levels are OR-reductions), so the fetch of the
// fine bit can be skipped without changing the result.
VXD bool brick_is_empty(const OccView& V, int32_t bx, int32_t by, int32_t bz) {
    for (int32_t l = V.nlevels - 1; l >= 0; --l) {
        int32_t sh = 3 * l;
        if (!level_bit(V, l, bx >> sh, by >> sh, bz >> sh)) return true;
    }
    return false;
}

struct RayOut {
    int32_t hit, term, steps, fetches;
    int32_t vox[3];
    float t;
    f3 hitpos, normal;
};

// raycastShadowVolume (Light.frag:29-81), mip 0 only (the ShadowVox image has
// one level, evk.h:265). Passes 2..4 of the outer loop repeat pass 1 exactly,
// so a pass that ends by n==256 or totalt>=maxt is traced once and reported
// with nt*4 (Light.frag:38,74,79). Coarse levels only skip fine fetches.
VXD void dda_exact(const OccView& V, f3 o, f3 d, float maxt, RayOut& R) {
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float s[3], td[3], tm[3];
    int32_t cv[3], up[3];
    const int32_t dims[3] = {V.sx, V.sy, V.sz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s[k] = gsign(dd[k]);
        td[k] = 1.0f / (dd[k] * s[k]);
        cv[k] = f2i(floorf(oo[k]));
        float nb = (float)cv[k] + (s[k] * 0.5f + 0.5f);
        tm[k] = (nb - oo[k]) / dd[k];
        up[k] = dims[k] * 2 + 1;
    }
    int32_t nt = 0, n = 0, term = 0;
    float totalt = 0.0f;
    int32_t bkx = -1, bky = -1, bkz = -1;
    bool bempty = true;
    size_t ckey = ~(size_t)0;
    uint64_t cw = 0;
    R.hit = 0;
    do {
        float s0 = gstep(tm[0], tm[2]) * gstep(tm[0], tm[1]);
        float s1 = gstep(tm[1], tm[0]) * gstep(tm[1], tm[2]);
        float s2 = gstep(tm[2], tm[1]) * gstep(tm[2], tm[0]);
        if ((uint32_t)cv[0] > (uint32_t)up[0] || (uint32_t)cv[1] > (uint32_t)up[1] ||
            (uint32_t)cv[2] > (uint32_t)up[2]) {
            term = 2;
            break;
        }
        int32_t bx = cv[0] >> 3, by = cv[1] >> 3, bz = cv[2] >> 3;
        if (bx != bkx || by != bky || bz != bkz) {
            bkx = bx; bky = by; bkz = bz;
            bempty = brick_is_empty(V, bx, by, bz);
        }
        bool vox = false;
        if (!bempty) {
            size_t ci = cell_index(V, cv[0], cv[1], cv[2]);
            if (ci != ckey) { ckey = ci; cw = V.cells[ci]; }
            vox = (cw >> cell_bit(cv[0], cv[1], cv[2])) & 1ull;
        }
        float best_t = (tm[0] * s0 + tm[1] * s1) + tm[2] * s2;
        if (vox) {
            term = 1;
            R.normal = mk3(-s[0] * s0, -s[1] * s1, -s[2] * s2);
            R.hitpos = mk3(oo[0] + dd[0] * best_t, oo[1] + dd[1] * best_t, oo[2] + dd[2] * best_t);
            R.vox[0] = cv[0]; R.vox[1] = cv[1]; R.vox[2] = cv[2];
            R.t = best_t;
            break;
        }
        cv[0] += f2i(s0 * s[0]);
        cv[1] += f2i(s1 * s[1]);
        cv[2] += f2i(s2 * s[2]);
        tm[0] += td[0] * s0;
        tm[1] += td[1] * s1;
        tm[2] += td[2] * s2;
        totalt = best_t;
        nt++;
    } while (++n < 256 && totalt < maxt);
    if (term == 0) {
        R.fetches = 4 * nt;
        nt *= 4;
    } else {
        R.fetches = nt + (term == 1 ? 1 : 0);
    }
    R.hit = term == 1;
    R.term = term;
    R.steps = nt;
}

// Positions per loop of the sparse marchers in the light and reflection
// passes (1, 2, 4 and 8 measured; 2 was fastest on both).
constexpr int kSparsePrefetch = 2;

// raycastShadowVolumeSparse / ...SuperSparse (Light.frag:131-217). The march
// samples positions pos = o + sd, o + 2 sd, ... (float sums, pos += sd per
// position) at distances dd = sf, 2 sf, ...: while dd < 16 the step is the
// marcher's factor (0.5 Sparse, 2.5 SuperSparse) and a position tests the
// sub-voxel bit of texel ivec3(pos / 2) that mod(pos, 0.5) > 0.25 selects;
// from there the step doubles, up to min(dist, 164), and a position tests
// whether texel ivec3(pos) / 2 has any bit set. A hit returns its dd; a ray
// that reaches the limit returns dist. mod(p, 0.5) > 0.25 is tested as
// fract(2p) > 0.5: mod is p - 0.5 floor(2p), whose float value is half that
// of 2p - floor(2p) (doubling is exact); v_fract_f32 returns values that round
// to 1 as 1 - 2^-24, still > 0.5, and NaN for NaN.
//
// A lane takes K positions per loop. The positions do not depend on what the
// march reads, so it computes its next K positions, requests their K cell
// words together and then tests them in order: a ray waits for one fetch
// latency per K positions instead of per position. The words requested past
// the position that ends the march (at most K - 1 per loop) read an in-range
// word or word 0 and are dropped. K = 1 is the plain march, for kernels with
// many rays per lane.
template <int K, bool COARSE, bool NARROW>
VXD bool sparse_batch(const OccView& V, f3& pos, f3 sd, float& dd, float sf, float lim, int32_t& it, bool& hit) {
    f3 p[K];
    float dk[K];
    uint64_t w[K];
    uint32_t sh[K];
    bool in[K];
    p[0] = pos;
    dk[0] = dd;
#pragma unroll
    for (int k = 1; k < K; ++k) {
        p[k] = mk3(p[k - 1].x + sd.x, p[k - 1].y + sd.y, p[k - 1].z + sd.z);
        dk[k] = dk[k - 1] + sf;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        // texel of the position: ivec3(pos / 2) while the step is fine,
        // ivec3(pos) / 2 once it is coarse (getVolumeAt(ivec3(pos), 1)). The
        // two agree wherever either is a texel of the volume (truncating
        // twice truncates once; past 2^31 both saturate outside), so both
        // phases take ivec3(pos * 0.5).
        const int32_t tx = f2i(p[k].x * 0.5f), ty = f2i(p[k].y * 0.5f), tz = f2i(p[k].z * 0.5f);
        const size_t ci = texel_word_n<NARROW>(V, tx, ty, tz, in[k]);
        w[k] = V.cells[ci];
        sh[k] = texel_base(tx, ty, tz);
    }
    // the first position that hits or lies past the limit ends the batch
    int32_t end = K;
    bool endhit = false;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        // the texel byte's bits sit at cell bits base + {0, 1, 4, 5, 16, 17,
        // 20, 21}; byte bit fx + 2 fy + 4 fz is cell bit base + fx + 4 fy + 16 fz
        bool h;
        if (COARSE) {
            h = in[k] && ((uint32_t)(w[k] >> sh[k]) & 0x330033u) != 0u;
        } else {
            uint32_t b = sh[k];
            b += __builtin_amdgcn_fractf(p[k].x * 2.0f) > 0.5f ? 1u : 0u;
            b += __builtin_amdgcn_fractf(p[k].y * 2.0f) > 0.5f ? 4u : 0u;
            b += __builtin_amdgcn_fractf(p[k].z * 2.0f) > 0.5f ? 16u : 0u;
            h = in[k] && ((w[k] >> b) & 1u) != 0u;
        }
        const bool valid = dk[k] < lim;
        if (!valid || h) {
            end = k;
            endhit = valid;
        }
    }
    if (end == K) {
        it += K;
        pos = mk3(p[K - 1].x + sd.x, p[K - 1].y + sd.y, p[K - 1].z + sd.z);
        dd = dk[K - 1] + sf;
        return false;
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (k == end) {
            pos = p[k];
            dd = dk[k];
        }
    it += end + (endhit ? 1 : 0);
    hit = endhit;
    return true;
}
template <int K, bool NARROW>
VXD float sparse_march_n(const OccView& V, f3 o, f3 d, float dist, float step_factor, RayOut& R) {
    float sf = step_factor;
    f3 sd = mk3(d.x * sf, d.y * sf, d.z * sf);
    f3 pos = o;
    float dd = sf;
    int32_t it = 0;
    bool hit = false;
    R.hit = 0; R.term = 0;
    while (!sparse_batch<K, false, NARROW>(V, pos, sd, dd, sf, 16.0f, it, hit)) {
    }
    if (!hit) {
        sf *= 2.0f;
        sd.x *= 2.0f; sd.y *= 2.0f; sd.z *= 2.0f;
        const float lim = gmin(dist, 164.0f);
        if (dd < lim)
            while (!sparse_batch<K, true, NARROW>(V, pos, sd, dd, sf, lim, it, hit)) {
            }
    }
    R.steps = R.fetches = it;
    if (hit) {
        R.hit = 1; R.term = 1; R.t = dd;
        return dd;
    }
    R.t = dist;
    return dist;
}
template <int K>
VXD float sparse_march(const OccView& V, f3 o, f3 d, float dist, float step_factor, RayOut& R) {
    return volume_narrow(V) ? sparse_march_n<K, true>(V, o, d, dist, step_factor, R)
                            : sparse_march_n<K, false>(V, o, d, dist, step_factor, R);
}

VXD void march_any(const OccView& V, int32_t march, f3 o, f3 d, float maxt, RayOut& R) {
    if (march == VX_MARCH_DDA) dda_exact(V, o, d, maxt, R);
    else sparse_march<1>(V, o, d, maxt, march == VX_MARCH_SPARSE ? 0.5f : 2.5f, R);
}

// Distance a marcher reports for AO / spec-occ accumulation; the sparse
// marchers take K positions per loop.
template <int K = 1>
VXD float march_dist(const OccView& V, int32_t march, f3 o, f3 d, float maxt, uint32_t& fetches, int32_t& hit) {
    RayOut R;
    if (march == VX_MARCH_DDA)
        dda_exact(V, o, d, maxt, R);
    else
        sparse_march<K>(V, o, d, maxt, march == VX_MARCH_SPARSE ? 0.5f : 2.5f, R);
    fetches += (uint32_t)R.fetches;
    hit = R.hit;
    if (march == VX_MARCH_DDA) return R.hit ? gmin(R.t, maxt) : maxt;
    return R.t;
}

__global__ void k_trace_rays(OccView V, const float* rays, uint32_t n, int32_t march, vx_ray_result* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* ry = rays + (size_t)7 * i;
    RayOut R;
    R.vox[0] = R.vox[1] = R.vox[2] = 0;
    R.t = 0.0f;
    R.hitpos = mk3(0.0f, 0.0f, 0.0f);
    R.normal = mk3(0.0f, 0.0f, 0.0f);
    march_any(V, march, mk3(ry[0], ry[1], ry[2]), mk3(ry[3], ry[4], ry[5]), ry[6], R);
    vx_ray_result o;
    o.hit = R.hit; o.term = R.term; o.steps = R.steps; o.fetches = R.fetches;
    if (march == VX_MARCH_DDA && R.hit) {
        o.voxel[0] = R.vox[0]; o.voxel[1] = R.vox[1]; o.voxel[2] = R.vox[2];
        o.t = R.t;
        o.hitpos[0] = R.hitpos.x; o.hitpos[1] = R.hitpos.y; o.hitpos[2] = R.hitpos.z;
        o.normal[0] = R.normal.x; o.normal[1] = R.normal.y; o.normal[2] = R.normal.z;
    } else {
        o.voxel[0] = o.voxel[1] = o.voxel[2] = 0;
        o.t = (march == VX_MARCH_DDA) ? 0.0f : R.t;
        o.hitpos[0] = o.hitpos[1] = o.hitpos[2] = 0.0f;
        o.normal[0] = o.normal[1] = o.normal[2] = 0.0f;
    }
    out[i] = o;
}

// Pixel of a 16x16 block tile; blocks are remapped so each XCD (blockIdx % 8)
// walks a contiguous band of tiles.
VXD bool tile_pixel(const TileMap& tm, int32_t& px, int32_t& py) {
    uint32_t b = blockIdx.x;
    uint32_t logical = (b % 8u) * tm.per_xcd + b / 8u;
    if (logical >= tm.ntiles) return false;
    int32_t tx = (int32_t)(logical % tm.tiles_x), ty = (int32_t)(logical / tm.tiles_x);
    int32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    px = tx * 16 + (wave & 1) * 8 + (lane & 7);
    py = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    return px < tm.width && py < tm.height;
}

// Ambient-pass pixel (LightAmbient.frag:134-175) generalised to n_ao cosine AO
// rays with a length cap, plus a GGX-sampled specular-occlusion ray
// (LightReflection.frag:86-113 ray origin), with the literal marchers.
VXD void trace_pixel_exact(const OccView& V, const TraceArgs& A, size_t i);

__global__ void __launch_bounds__(256) k_trace(OccView V, TraceArgs A) {
    int32_t px, py;
    if (!tile_pixel(A.tiles, px, py)) return;
    trace_pixel_exact(V, A, (size_t)py * A.view.width + px);
}

VXD void trace_pixel_exact(const OccView& V, const TraceArgs& A, size_t i) {
    const vx_trace_params& P = A.params;
    uint8_t m = A.mat[i];
    if (m == 0) {
        A.shadow[i] = 0; A.ao[i] = 0.0f; A.spec[i] = 0.0f;
        if (A.fetches) A.fetches[i] = 0;
        return;
    }
    uint32_t nf = 0;
    f3 pp = load3(A.pos + 3 * i), n = load3(A.nrm + 3 * i);
    uint32_t pk = pixkey(P.seed, P.frame, (uint32_t)i);
    float r0[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r0[k] = rand01(pk, (uint32_t)k);

    float wx = pp.x * 0.1f, wy = pp.y * 0.1f, wz = pp.z * 0.1f;
    const float* Vm = A.view.view;
    float vz = ((Vm[2] * wx + Vm[6] * wy) + Vm[10] * wz) + Vm[14];
    float depth = ((-vz) - 0.1f) / (4096.0f - 0.1f);
    float st = gclamp((depth - 0.0f) / (0.2f - 0.0f), 0.0f, 1.0f);
    float bias = (st * st) * (3.0f - 2.0f * st) * 50.0f + 1.5f;

    f3 wd = mk3(P.light_dir[0], P.light_dir[1], P.light_dir[2]);
    f3 wcp = pp;
    if (P.jitter_shadow) {
        f3 rv = cosine_hemisphere(r0[0], r0[1]);
        rv = mk3(rv.x * 0.1f, rv.y * 0.1f, rv.z * 0.1f);
        rv.z = rv.z * gsign(r0[2] - 0.5f);
        f3 mx = mk3(wd.x * (1.0f - 0.5f) + rv.x * 0.5f, wd.y * (1.0f - 0.5f) + rv.y * 0.5f,
                    wd.z * (1.0f - 0.5f) + rv.z * 0.5f);
        wd = normalize3(mx);
        float w = r0[3] * 1.0f;
        wcp = mk3(wcp.x + wd.x * w, wcp.y + wd.y * w, wcp.z + wd.z * w);
        wcp = mk3(wcp.x + rv.x * 2.5f, wcp.y + rv.y * 2.5f, wcp.z + rv.z * 2.5f);
    }
    f3 org = mk3(wcp.x + n.x * bias, wcp.y + n.y * bias, wcp.z + n.z * bias);

    uint8_t lit = 1;
    if (P.n_shadow > 0) {
        int32_t h;
        float t = march_dist(V, P.shadow_march, org, wd, P.shadow_maxt, nf, h);
        if (P.shadow_march == VX_MARCH_DDA) lit = h ? 0 : 1;
        else lit = (t != P.shadow_maxt) ? 0 : 1;
    }
    A.shadow[i] = lit;

    float aov = 1.0f;
    if (P.n_ao > 0) {
        f3 T, B;
        tangent_frame(n, T, B);
        float aosum = 0.0f;
        for (int32_t a = 0; a < P.n_ao; ++a) {
            float u1 = rand01(pk, (uint32_t)(4 * (1 + a) + 0));
            float u2 = rand01(pk, (uint32_t)(4 * (1 + a) + 1));
            f3 rv = cosine_hemisphere(u1, u2);
            f3 dir = mk3((T.x * rv.x + B.x * rv.y) + n.x * rv.z, (T.y * rv.x + B.y * rv.y) + n.y * rv.z,
                         (T.z * rv.x + B.z * rv.y) + n.z * rv.z);
            int32_t h;
            float t = march_dist(V, P.ao_march, org, dir, P.ao_maxt, nf, h);
            float q = gmin(t, P.ao_maxt) / P.ao_maxt;
            aosum = aosum + q * q;
        }
        aov = aosum / (float)P.n_ao;
    }
    A.ao[i] = aov;

    float spv = 1.0f;
    if (P.n_spec > 0) {
        uint32_t base = (uint32_t)(4 * (1 + P.n_ao));
        float u1 = rand01(pk, base + 0), u2 = rand01(pk, base + 1), u3 = rand01(pk, base + 3);
        float rough = A.palette[6 * m + 3];
        f3 cam = mk3(A.view.cam_pos[0] * 10.0f, A.view.cam_pos[1] * 10.0f, A.view.cam_pos[2] * 10.0f);
        f3 Vd = normalize3(mk3(cam.x - pp.x, cam.y - pp.y, cam.z - pp.z));
        float alpha = rough * rough;
        float a2 = alpha * alpha;
        float ct2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
        float ct = sqrtf(ct2);
        float stt = sqrtf(gmax(0.0f, 1.0f - ct2));
        float sp, cp;
        sincos_vx(6.2831855f * u2, &sp, &cp);
        f3 T, B;
        tangent_frame(n, T, B);
        f3 hl = mk3(stt * cp, stt * sp, ct);
        f3 h = mk3((T.x * hl.x + B.x * hl.y) + n.x * hl.z, (T.y * hl.x + B.y * hl.y) + n.y * hl.z,
                   (T.z * hl.x + B.z * hl.y) + n.z * hl.z);
        f3 I = mk3(-Vd.x, -Vd.y, -Vd.z);
        float dhi = dot3(h, I);
        f3 wdir = mk3(I.x - (2.0f * dhi) * h.x, I.y - (2.0f * dhi) * h.y, I.z - (2.0f * dhi) * h.z);
        f3 so = mk3((pp.x + n.x * u3) + n.x, (pp.y + n.y * u3) + n.y, (pp.z + n.z * u3) + n.z);
        int32_t hh;
        float t = march_dist(V, P.spec_march, so, wdir, P.spec_maxt, nf, hh);
        spv = gmin(t, P.spec_maxt) / P.spec_maxt;
    }
    A.spec[i] = spv;
    if (A.fetches) A.fetches[i] = nf;
}

// ===========================================================================
// Ray-parallel DDA trace: every ray class through raycastShadowVolume
// semantics (Light.frag:29-81).
//
// Work is split by ray. A persistent wave takes 8x8-pixel tiles from its XCD's
// band and generates their rays 64 at a time into a ring in LDS: one ray index
// over the tile's valid pixels per batch, so the set-up code runs with one ray
// class and all lanes. A lane whose ray ended takes the next ray from the
// ring. Every ray owns one result slot (best_t of a hit, and the fetch count),
// slot = ((tile * nrays + r) << 6) | pixel-in-tile; k_resolve folds a pixel's
// slots in ray order, so the float sums keep the reference order. A slot holds
// best_t of a hit and a marker NaN otherwise; k_resolve derives the values.
//
// Exact fast path. For a ray with every |dir| in [2^-100, 3e38] and |origin| <
// 2^22, t_max and t_delta stay finite and t_max >= 0 for the whole march, so:
//   select_k = (t_max_k == min3(t_max))            (step(): no NaNs)
//   best_t   = dot(t_max, select) = t, t+t or (t+t)+t for 1, 2 or 3 selected,
//              plus +0 (an unselected t_max*0) unless all 3 are selected
//   t_max_k += t_delta_k * select_k  ==  select_k ? t_max_k + t_delta_k : t_max_k
//   voxel_k += ivec3(select_k * sign_k) == select_k ? sign_k : 0
// Other rays are listed and traced afterwards by the literal path (k_march_rest).
//
// Voxels come from the march grid: the occupancy cells (u64 = 4x4x4 fine
// voxels, bit (x&3)|(y&3)<<2|(z&3)<<4) in a linear X-fastest grid with one
// border cell per side, where every fine voxel outside [0, 2s+1]^3 has its bit
// set. A ray that starts inside the bounds stops at its first set bit, either
// an occupied voxel (hit) or the first out-of-bounds voxel (term 2), told
// apart once at its end, and every voxel it visits lies inside the grid.
// Lane coordinates carry a +4 bias (grid cell = (X + 4) >> 2 per axis).
// ===========================================================================

constexpr int kRing = 128;       // ring entries per wave
constexpr int kRingFields = 10;  // t_max[3], t_delta[3], x|y<<16, z|signs<<16|kind<<19, slot, grid word

VXD bool fast_ray_ok(f3 o, f3 d) {
    const float lo = 7.8886091e-31f;   // 2^-100
    bool ok = true;
    float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = fabsf(dd[k]);
        ok = ok && (a >= lo) && (a <= 3.0e38f) && (fabsf(oo[k]) < 4194304.0f);
    }
    return ok;
}

// Per-lane pixel state of the generation tile.
struct PixState {
    f3 org;
    uint32_t pk;
    uint32_t i;
};

enum { RAY_SHADOW = 0, RAY_AO = 1, RAY_SPEC = 2 };

// Pixel set-up shared by its rays (LightAmbient.frag:136-158) for a non-sky
// pixel i; wd receives the shadow direction.
VXD void pixel_setup(const TraceArgs& A, uint32_t i, PixState& S, f3& wd) {
    const vx_trace_params& P = A.params;
    f3 p = load3(A.pos + 3 * (size_t)i);
    f3 n = load3(A.nrm + 3 * (size_t)i);
    S.i = i;
    S.pk = pixkey(P.seed, P.frame, i);
    float r0[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r0[k] = rand01(S.pk, (uint32_t)k);
    float wx = p.x * 0.1f, wy = p.y * 0.1f, wz = p.z * 0.1f;
    const float* Vm = A.view.view;
    float vz = ((Vm[2] * wx + Vm[6] * wy) + Vm[10] * wz) + Vm[14];
    float depth = ((-vz) - 0.1f) / (4096.0f - 0.1f);
    float st = gclamp((depth - 0.0f) / (0.2f - 0.0f), 0.0f, 1.0f);
    float bias = (st * st) * (3.0f - 2.0f * st) * 50.0f + 1.5f;
    wd = mk3(P.light_dir[0], P.light_dir[1], P.light_dir[2]);
    f3 wcp = p;
    if (P.jitter_shadow) {
        f3 rv = cosine_hemisphere(r0[0], r0[1]);
        rv = mk3(rv.x * 0.1f, rv.y * 0.1f, rv.z * 0.1f);
        rv.z = rv.z * gsign(r0[2] - 0.5f);
        f3 mx = mk3(wd.x * (1.0f - 0.5f) + rv.x * 0.5f, wd.y * (1.0f - 0.5f) + rv.y * 0.5f,
                    wd.z * (1.0f - 0.5f) + rv.z * 0.5f);
        wd = normalize3(mx);
        float w = r0[3] * 1.0f;
        wcp = mk3(wcp.x + wd.x * w, wcp.y + wd.y * w, wcp.z + wd.z * w);
        wcp = mk3(wcp.x + rv.x * 2.5f, wcp.y + rv.y * 2.5f, wcp.z + rv.z * 2.5f);
    }
    S.org = mk3(wcp.x + n.x * bias, wcp.y + n.y * bias, wcp.z + n.z * bias);
}

// Origin, direction and maxt of ray r >= n_shadow of the pixel (AO or spec).
VXD int32_t ray_setup(const TraceArgs& A, const PixState& S, int32_t r, f3& o, f3& d, float& maxt) {
    const vx_trace_params& P = A.params;
    f3 n = load3(A.nrm + 3 * (size_t)S.i);
    f3 T, B;
    tangent_frame(n, T, B);
    int32_t a = r - P.n_shadow;
    if (a < P.n_ao) {
        float u1 = rand01(S.pk, (uint32_t)(4 * (1 + a) + 0));
        float u2 = rand01(S.pk, (uint32_t)(4 * (1 + a) + 1));
        f3 rv = cosine_hemisphere(u1, u2);
        d = mk3((T.x * rv.x + B.x * rv.y) + n.x * rv.z, (T.y * rv.x + B.y * rv.y) + n.y * rv.z,
                (T.z * rv.x + B.z * rv.y) + n.z * rv.z);
        o = S.org;
        maxt = P.ao_maxt;
        return RAY_AO;
    }
    f3 p = load3(A.pos + 3 * (size_t)S.i);
    uint32_t base = (uint32_t)(4 * (1 + P.n_ao));
    float u1 = rand01(S.pk, base + 0), u2 = rand01(S.pk, base + 1), u3 = rand01(S.pk, base + 3);
    float rough = A.palette[6 * (uint32_t)A.mat[S.i] + 3];
    f3 cam = mk3(A.view.cam_pos[0] * 10.0f, A.view.cam_pos[1] * 10.0f, A.view.cam_pos[2] * 10.0f);
    f3 Vd = normalize3(mk3(cam.x - p.x, cam.y - p.y, cam.z - p.z));
    float alpha = rough * rough;
    float a2 = alpha * alpha;
    float ct2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
    float ct = sqrtf(ct2);
    float stt = sqrtf(gmax(0.0f, 1.0f - ct2));
    float sp, cp;
    sincos_vx(6.2831855f * u2, &sp, &cp);
    f3 hl = mk3(stt * cp, stt * sp, ct);
    f3 h = mk3((T.x * hl.x + B.x * hl.y) + n.x * hl.z, (T.y * hl.x + B.y * hl.y) + n.y * hl.z,
               (T.z * hl.x + B.z * hl.y) + n.z * hl.z);
    f3 I = mk3(-Vd.x, -Vd.y, -Vd.z);
    float dhi = dot3(h, I);
    d = mk3(I.x - (2.0f * dhi) * h.x, I.y - (2.0f * dhi) * h.y, I.z - (2.0f * dhi) * h.z);
    o = mk3((p.x + n.x * u3) + n.x, (p.y + n.y * u3) + n.y, (p.z + n.z * u3) + n.z);
    maxt = P.spec_maxt;
    return RAY_SPEC;
}

// Ray r of pixel S (wd: the pixel's shadow direction).
VXD int32_t pixel_ray(const TraceArgs& A, const PixState& S, f3 wd, int32_t r, f3& o, f3& d, float& maxt) {
    if (r < A.params.n_shadow) {
        o = S.org; d = wd; maxt = A.params.shadow_maxt;
        return RAY_SHADOW;
    }
    return ray_setup(A, S, r, o, d, maxt);
}

// Value a pixel folds for one ray (shadow: lit, AO and spec: min(dist, maxt) /
// maxt with dist = maxt unless hit) and the ray's fetch count. The march
// stores per ray only best_t of a hit (kNoHitBits otherwise); k_resolve folds.
VXD float ray_value(int32_t kind, int32_t term, float t, float maxt) {
    bool hit = term == 1;
    if (kind == RAY_SHADOW) return hit ? 0.0f : 1.0f;
    float dist = hit ? gmin(t, maxt) : maxt;
    return gmin(dist, maxt) / maxt;
}
VXD uint32_t ray_fetches(int32_t term, int32_t nt) {
    return (uint32_t)(term == 0 ? 4 * nt : nt + (term == 1 ? 1 : 0));
}

// Slot value of a ray that did not hit: a NaN whose payload the path's
// arithmetic does not produce (it yields the default NaN), so every hit's
// best_t, including -inf or NaN from the literal path, stays distinguishable.
constexpr uint32_t kNoHitBits = 0x7fc0beefu;

// best_t of the current voxel from t_max (see the fast-path identities).
VXD float best_t_of(float tm0, float tm1, float tm2) {
    float t = fminf(fminf(tm0, tm1), tm2);
    int32_t ns = (tm0 == t ? 1 : 0) + (tm1 == t ? 1 : 0) + (tm2 == t ? 1 : 0);
    if (ns == 3) return (t + t) + t;
    return (ns == 2 ? t + t : t) + 0.0f;
}

template <bool SINGLE>
VXD const MarchItem& march_item(const MarchArgs& A, uint32_t k);
template <bool SINGLE>
VXD uint32_t item_of_tile(const MarchArgs& A, uint32_t gt);
VXD bool tile_lane_pixel(const MarchItem& I, uint32_t lt, uint32_t l, uint32_t& i);

// Hit record of the ray in `slot` (A.hits[pixel * nrays + r]).
template <bool SINGLE>
VXD void put_hit(const MarchArgs& A, uint32_t slot, int32_t term, int32_t steps, int32_t fetches,
                 const int32_t vox[3], float t, const float nrm[3]) {
    const uint32_t q = slot >> 6, l = slot & 63u;
    const uint32_t gt = q / (uint32_t)A.nrays, r = q - gt * (uint32_t)A.nrays;
    const MarchItem& I = march_item<SINGLE>(A, item_of_tile<SINGLE>(A, gt));
    uint32_t i = 0;
    if (!tile_lane_pixel(I, gt - I.tile0, l, i)) return;
    vx_trace_hit h;
    h.term = term;
    h.steps = steps;
    h.fetches = fetches;
    h.voxel[0] = vox[0]; h.voxel[1] = vox[1]; h.voxel[2] = vox[2];
    h.t = t;
    h.normal[0] = nrm[0]; h.normal[1] = nrm[1]; h.normal[2] = nrm[2];
    A.hits[(size_t)i * (uint32_t)A.nrays + r] = h;
}

template <bool SINGLE>
VXD const MarchItem& march_item(const MarchArgs& A, uint32_t k) {
    if (SINGLE) return A.item0;
    return A.items[k];
}

// The same for a wave-uniform item: the items are read-only for the whole
// launch, so they are read through the constant address space, which keeps
// the item's fields in scalar registers.
template <bool SINGLE, std::enable_if_t<SINGLE, int> = 0>
VXD const MarchItem& march_item_uniform(const MarchArgs& A, uint32_t) {
    return A.item0;
}
template <bool SINGLE, std::enable_if_t<!SINGLE, int> = 0>
VXD MarchItem march_item_uniform(const MarchArgs& A, uint32_t k) {
    static_assert(sizeof(MarchItem) % 4 == 0, "MarchItem is whole words");
    typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
    const ConstWords src = (ConstWords)(A.items + __builtin_amdgcn_readfirstlane(k));
    MarchItem I;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&I);
#pragma unroll
    for (uint32_t w = 0; w < sizeof(MarchItem) / 4; ++w) dst[w] = src[w];
    return I;
}

// Item of global tile gt (items ordered by tile0).
template <bool SINGLE>
VXD uint32_t item_of_tile(const MarchArgs& A, uint32_t gt) {
    if (SINGLE) return 0;
    if (A.tile_item) return A.tile_item[gt];
    uint32_t lo = 0, hi = A.nitems - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (A.items[mid].tile0 <= gt) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// Pixel index of lane l of queue tile lt, or false outside the view. A tile
// is (1 << tw_shift) x (64 >> tw_shift) pixels, lanes row-major inside it; in
// order 1 the queue walks 8x8-tile blocks row by row, tiles row-major inside
// a block (the padding tiles of the last block column / row lie outside).
VXD bool tile_lane_pixel(const MarchItem& I, uint32_t lt, uint32_t l, uint32_t& i) {
    uint32_t tx, ty;
    if (I.order == 0) {
        tx = lt % I.tiles_x;
        ty = lt / I.tiles_x;
    } else {
        const uint32_t b = lt >> 6, w = lt & 63u;
        tx = (b % I.nbx) * 8u + (w & 7u);
        ty = (b / I.nbx) * 8u + (w >> 3);
    }
    const uint32_t px = (tx << I.tw_shift) + (l & ((1u << I.tw_shift) - 1u));
    const uint32_t py = ty * (64u >> I.tw_shift) + (l >> I.tw_shift);
    i = py * (uint32_t)I.T.view.width + px;
    return tx < I.tiles_x && px < (uint32_t)I.T.view.width && py < (uint32_t)I.T.view.height;
}

// Origin voxel and DDA state of ray (o, d) at the origin, Light.frag:32-46
// (|o| < 2^22, so the int conversion is exact).
VXD void ray_origin_voxel(f3 o, int32_t cv[3]) {
    cv[0] = (int32_t)floorf(o.x);
    cv[1] = (int32_t)floorf(o.y);
    cv[2] = (int32_t)floorf(o.z);
}

// Writes the ring entry of a fast-path ray that starts inside the bounds.
VXD void ring_put_ray(uint32_t (*ring)[kRing], uint32_t e, f3 o, f3 d, const int32_t cv[3], int32_t kind,
                      uint32_t slot, uint32_t gx, uint32_t gxy) {
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    uint32_t signs = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        bool neg = !(dd[k] > 0.0f);
        float s = neg ? -1.0f : 1.0f;
        signs |= neg ? (1u << k) : 0u;
        float nb = (float)cv[k] + (neg ? 0.0f : 1.0f);   // vec3(cv) + (sign*0.5 + 0.5)
        ring[k][e] = __float_as_uint((nb - oo[k]) / dd[k]);
        ring[3 + k][e] = __float_as_uint(1.0f / (dd[k] * s));
    }
    ring[6][e] = (uint32_t)(cv[0] + 4) | ((uint32_t)(cv[1] + 4) << 16);
    ring[7][e] = (uint32_t)(cv[2] + 4) | (signs << 16) | ((uint32_t)kind << 19);
    ring[8][e] = slot;
    ring[9][e] = __umul24((uint32_t)((cv[2] + 4) >> 2), gxy) +
                  (__umul24((uint32_t)((cv[1] + 4) >> 2), gx) + (uint32_t)((cv[0] + 4) >> 2));
}

// The ray a lane marches (coordinates as below); meta = (n + 0x700) << 20 |
// kind << 16, so n++ is meta += 1 << 20 and n < 256 is the sign bit clear.
struct MarchRay {
    float tm0, tm1, tm2, td0, td1, td2;
    int32_t x, y, z, sx, sy, sz;
    uint32_t meta;
    float maxt;
    uint32_t slot;
    uint32_t ci;           // index of the grid word at the lane's position
    uint64_t cw;           // grid word of the lane's voxel
};

// Lane coordinates: x = X + 4, y = (Y + 4) << 2, z = (Z + 4) << 4 for fine
// voxel (X, Y, Z), so the in-cell fields sit at their cell-bit positions.

// Shift amount selecting the lane's voxel in its cell word: the low 6 bits are
// cell_bit(X, Y, Z) = (x & 3) | (y & 12) | (z & 48); the bits above are
// don't-care, since a 64-bit shift reads only the low 6 bits of its amount.
VXD uint32_t cell_shift(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t t, r;
    asm("v_bfi_b32 %0, 12, %1, %2" : "=v"(t) : "v"(y), "v"(z));
    asm("v_bfi_b32 %0, 3, %1, %2" : "=v"(r) : "v"(x), "v"(t));
    return r;
}

// Grid word index of the lane's voxel. It stays inside the grid: a ray starts
// inside the bounds and stops at the first border voxel it reaches.
// Two v_mad_u32_u24 with the pitches in scalar registers (the compiler's own
// choice, two multiplies and a three-way add, is one instruction longer).
VXD uint32_t grid_word(int32_t x, int32_t y, int32_t z, uint32_t gx, uint32_t gxy) {
    uint32_t a, r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"((uint32_t)(y >> 4)), "s"(gx), "v"((uint32_t)(x >> 2)));
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"((uint32_t)(z >> 6)), "s"(gxy), "v"(a));
    return r;
}

// Position and t_max steps along x / y / z of the lanes in e0 / e1 / e2 that
// are also in m. The adds run with those lanes as exec, which the block
// restores to `full` (the wave's exec where it is called) before it ends:
// each costs one add, instead of an add and a select.
VXD void move_axes(MarchRay& R, uint64_t e0, uint64_t e1, uint64_t e2, uint64_t m, uint64_t full) {
    asm volatile(
        "s_and_b64 exec, %[e0], %[m]\n\t"
        "v_add_u32 %[x], %[x], %[sx]\n\t"
        "s_and_b64 exec, %[e1], %[m]\n\t"
        "v_add_u32 %[y], %[y], %[sy]\n\t"
        "s_and_b64 exec, %[e2], %[m]\n\t"
        "v_add_u32 %[z], %[z], %[sz]\n\t"
        "s_mov_b64 exec, %[full]"
        : [x] "+v"(R.x), [y] "+v"(R.y), [z] "+v"(R.z)
        : [e0] "s"(e0), [e1] "s"(e1), [e2] "s"(e2), [m] "s"(m), [full] "s"(full), [sx] "v"(R.sx),
          [sy] "v"(R.sy), [sz] "v"(R.sz)
        : "scc");
}
VXD void advance_tmax(MarchRay& R, uint64_t e0, uint64_t e1, uint64_t e2, uint64_t m, uint64_t full) {
    asm volatile(
        "s_and_b64 exec, %[e0], %[m]\n\t"
        "v_add_f32 %[t0], %[t0], %[d0]\n\t"
        "s_and_b64 exec, %[e1], %[m]\n\t"
        "v_add_f32 %[t1], %[t1], %[d1]\n\t"
        "s_and_b64 exec, %[e2], %[m]\n\t"
        "v_add_f32 %[t2], %[t2], %[d2]\n\t"
        "s_mov_b64 exec, %[full]"
        : [t0] "+v"(R.tm0), [t1] "+v"(R.tm1), [t2] "+v"(R.tm2)
        : [e0] "s"(e0), [e1] "s"(e1), [e2] "s"(e2), [m] "s"(m), [full] "s"(full), [d0] "v"(R.td0),
          [d1] "v"(R.td1), [d2] "v"(R.td2)
        : "scc");
}

// A march grid as a global-memory pointer (its fetches are global loads, not
// flat ones), optionally held in vector registers (the same value in every
// lane).
typedef const __attribute__((address_space(1))) uint64_t* GridPtr;
VXD GridPtr grid_ptr(const uint64_t* p) {
    return (GridPtr)p;
}
VXD GridPtr in_vgpr(const uint64_t* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    asm volatile("" : "+v"(v));
    return (GridPtr)v;
}

// meta += inc in the lanes of m (exec restored before the block ends).
VXD void add_masked(uint32_t& meta, uint32_t inc, uint64_t m) {
    uint64_t sv;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_mov_b64 exec, %[m]\n\t"
        "v_add_u32 %[v], %[i], %[v]\n\t"
        "s_mov_b64 exec, %[sv]"
        : [v] "+v"(meta), [sv] "=&s"(sv)
        : [m] "s"(m), [i] "s"(inc));
}

// Occupancy: 7 waves per SIMD (72 VGPRs, 140 KB of LDS rings per CU). A wave
// marches one item at a time, so the item's grid address and pitches stay
// wave-uniform: in a batch, a wave that draws a tile of another item first
// lets its rays in flight finish, then switches.
// COUNT: every ray's step count is kept for the fetch-count output. Without
// it the count only serves the 256-step limit, which is tested at block ends,
// so it advances by STEPS once per block.
template <int BLOCK, int STEPS, bool SINGLE, bool COUNT>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_march(MarchArgs A) {
    static_assert(256 % STEPS == 0, "the step limit is tested at block ends");
    constexpr int WAVES = BLOCK / 64;
    __shared__ uint32_t ring_s[WAVES][kRingFields][kRing];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t(*ring)[kRing] = ring_s[threadIdx.x >> 6];
    const int32_t nrays = A.nrays;
    const uint64_t lane_lt = (1ull << lane) - 1ull;

    // generation cursor (wave-uniform): queue item = (tile, chunk of its
    // rays); gr runs over the chunk's generation indices [gr, gend)
    uint32_t q = blockIdx.x & 7u, misses = 0;
    bool gen_done = nrays <= 0;   // no rays per pixel: k_resolve writes the defaults
    uint32_t gitem = 0, gtile = 0;
    int32_t gr = SINGLE ? 0 : nrays, gend = 0;
    uint32_t head = 0, tail = 0;
    // this lane's pixel of the generation tile
    PixState P;
    P.org = mk3(0.0f, 0.0f, 0.0f); P.pk = 0; P.i = 0;
    bool pvalid = false;
    // this lane's ray
    // A lane without a ray keeps x = y = z = 0 and ci = 0 = the grid word of
    // that position: the march's steps leave it in place and load nothing.
    MarchRay R;
    R.tm0 = 0.0f; R.tm1 = 1.0f; R.tm2 = 2.0f;
    R.td0 = R.td1 = R.td2 = 0.0f;
    R.x = R.y = R.z = 0; R.sx = R.sy = R.sz = 1; R.meta = 0; R.maxt = 0.0f; R.slot = 0; R.ci = 0; R.cw = 0;
    bool hasray = false, running = false;
    const uint64_t lane_bit = 1ull << lane;
    // the wave's item (its rays in flight all belong to it)
    uint32_t citem = SINGLE ? 0u : ~0u, nitem = 0;
    bool pending = false;
    // the item's grid address, held in vector registers: the steps use it for
    // every fetch, and in scalar registers it competes with the launch's
    // uniform state (the compiler then re-reads it from a spill every step)
    GridPtr grid = SINGLE ? in_vgpr(A.item0.grid) : grid_ptr(A.item0.grid);
    uint32_t ggx = (uint32_t)A.item0.gx, ggxy = (uint32_t)A.item0.gxy;

    while (true) {
        const uint64_t runm = __ballot(running);
        if (runm == 0 || (uint32_t)__popcll(~runm) >= (uint32_t)A.refill) {
            // ---- finish ended rays
            // (set-up and refill run at the base wave priority, the steps one
            // above it: the SIMD then favours waves that are marching)
            __builtin_amdgcn_s_setprio(0);
            if (hasray && !running) {
                // The march moved the ray's position ahead in its last step
                // (its t_max did not advance): step it back to the voxel that
                // step tested. A ray whose step ran past maxt or the step limit
                // tested a free voxel, so the ray sits in an occupied voxel
                // exactly when it stopped there.
                {
                    const float tmin = fminf(fminf(R.tm0, R.tm1), R.tm2);
                    R.x = R.tm0 == tmin ? R.x - R.sx : R.x;
                    R.y = R.tm1 == tmin ? R.y - R.sy : R.y;
                    R.z = R.tm2 == tmin ? R.z - R.sz : R.z;
                }
                R.ci = grid_word(R.x, R.y, R.z, ggx, ggxy);
                R.cw = grid[R.ci];
                const bool stopped =
                    (int64_t)(R.cw << (cell_shift((uint32_t)R.x, (uint32_t)R.y, (uint32_t)R.z) & 63u)) < 0;
                const uint32_t kind = (R.meta >> 16) & 3u, n = (R.meta >> 20) - 0x700u - (stopped ? 1u : 0u);
                const MarchItem& I = march_item_uniform<SINGLE>(A, citem);
                int32_t term = 0;
                float v = __uint_as_float(kNoHitBits);
                if (stopped) {
                    bool oob = (uint32_t)(R.x - 4) > (uint32_t)I.lim[0] ||
                               (uint32_t)((R.y >> 2) - 4) > (uint32_t)I.lim[1] ||
                               (uint32_t)((R.z >> 4) - 4) > (uint32_t)I.lim[2];
                    term = oob ? 2 : 1;
                    if (!oob) v = best_t_of(R.tm0, R.tm1, R.tm2);
                }
                A.res[R.slot] = v;
                if (A.fres) A.fres[R.slot] = (uint16_t)ray_fetches(term, (int32_t)n);
                if (COUNT && A.hits) {
                    // Light.frag:57-63 at the hit voxel: normal = -stepSign * select
                    int32_t vox[3] = {0, 0, 0};
                    float nrm[3] = {0.0f, 0.0f, 0.0f}, tt = 0.0f;
                    if (term == 1) {
                        vox[0] = R.x - 4; vox[1] = (R.y >> 2) - 4; vox[2] = (R.z >> 4) - 4;
                        tt = v;
                        const float tmin = fminf(fminf(R.tm0, R.tm1), R.tm2);
                        nrm[0] = -(R.sx > 0 ? 1.0f : -1.0f) * (R.tm0 == tmin ? 1.0f : 0.0f);
                        nrm[1] = -(R.sy > 0 ? 1.0f : -1.0f) * (R.tm1 == tmin ? 1.0f : 0.0f);
                        nrm[2] = -(R.sz > 0 ? 1.0f : -1.0f) * (R.tm2 == tmin ? 1.0f : 0.0f);
                    }
                    put_hit<SINGLE>(A, R.slot, term, term == 0 ? 4 * (int32_t)n : (int32_t)n,
                                    (int32_t)ray_fetches(term, (int32_t)n), vox, tt, nrm);
                }
                hasray = false;
                R.x = R.y = R.z = 0;
                R.ci = 0;
                R.tm0 = 0.0f; R.tm1 = 1.0f; R.tm2 = 2.0f;
            }
            // ---- generate until the ring holds 64 rays or the work is out
            while (tail - head < 64u && !gen_done) {
                f3 wd = mk3(0.0f, 0.0f, 0.0f);
                if (gr >= (SINGLE ? gend : nrays)) {
                    if (!pending) {
                        // next queue item: own band first, then the others. A
                        // band is its tile range once per chunk; the last band's
                        // items past the last tile are skipped.
                        uint32_t tile = 0, chunk = 0;
                        bool got = false;
                        while (misses < 8u) {
                            uint32_t t = 0;
                            if (lane == 0) t = atomicAdd(&A.counters[q * kCounterStride], 1u);
                            t = __builtin_amdgcn_readfirstlane(t);
                            if (SINGLE) {
                                if (t < A.band) {
                                    chunk = t / A.band_tiles;
                                    tile = q * A.band_tiles + (t - chunk * A.band_tiles);
                                    if (tile < A.ntiles) { got = true; break; }
                                    continue;
                                }
                            } else {
                                // a batch takes one item per tile
                                tile = q * A.band + t;
                                if (t < A.band && tile < A.ntiles) { got = true; break; }
                            }
                            q = (q + 1u) & 7u;
                            misses++;
                        }
                        if (!got) { gen_done = true; break; }
                        gtile = tile;
                        if (SINGLE) {
                            const uint32_t chunks = A.gen & 0xffu;
                            gr = (int32_t)((chunk * (uint32_t)nrays) / chunks);
                            gend = (int32_t)(((chunk + 1u) * (uint32_t)nrays) / chunks);
                        }
                        nitem = __builtin_amdgcn_readfirstlane(item_of_tile<SINGLE>(A, gtile));
                    }
                    if (!SINGLE && nitem != citem) {
                        // another item: switch once nothing of this one is in flight
                        if (head != tail || __ballot(hasray) != 0) {
                            pending = true;
                            break;
                        }
                        citem = nitem;
                        const MarchItem& J = march_item_uniform<SINGLE>(A, citem);
                        grid = grid_ptr(J.grid);
                        ggx = (uint32_t)J.gx;
                        ggxy = (uint32_t)J.gxy;
                    }
                    pending = false;
                    gitem = nitem;
                    const MarchItem& I = march_item_uniform<SINGLE>(A, gitem);
                    uint32_t i = 0;
                    pvalid = tile_lane_pixel(I, gtile - I.tile0, lane, i) && I.T.mat[i] != 0;
                    if (pvalid) pixel_setup(I.T, i, P, wd);
                    if (!SINGLE) gr = 0;
                }
                // one batch: ray r of every valid pixel of the tile; generation
                // index gr runs shadow rays, then spec rays (the long ones,
                // early in the launch), then AO rays
                int32_t r = gr;
                if (SINGLE) {
                    const int32_t nsh = (int32_t)((A.gen >> 8) & 0xffu), nsp = (int32_t)(A.gen >> 16);
                    r = gr < nsh ? gr : (gr < nsh + nsp ? nrays - 1 - (gr - nsh) : gr - nsp);
                }
                const MarchItem& I = march_item_uniform<SINGLE>(A, gitem);
                f3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(1.0f, 1.0f, 1.0f);
                float maxt = 1.0f;
                int32_t kind = 0;
                int32_t cv[3] = {0, 0, 0};
                bool push = false;
                const uint32_t slot = ((gtile * (uint32_t)nrays + (uint32_t)r) << 6) | lane;
                if (pvalid) {
                    kind = pixel_ray(I.T, P, wd, r, o, d, maxt);
                    if (!fast_ray_ok(o, d)) {
                        uint32_t k = atomicAdd(&A.counters[8 * kCounterStride], 1u);
                        A.deferred[k] = slot;
                    } else {
                        ray_origin_voxel(o, cv);
                        if ((uint32_t)cv[0] <= (uint32_t)I.lim[0] && (uint32_t)cv[1] <= (uint32_t)I.lim[1] &&
                            (uint32_t)cv[2] <= (uint32_t)I.lim[2]) {
                            push = true;
                        } else {
                            // starts outside: term 2 at step 0
                            A.res[slot] = __uint_as_float(kNoHitBits);
                            if (A.fres) A.fres[slot] = 0;
                            if (COUNT && A.hits) {
                                const int32_t z3[3] = {0, 0, 0};
                                const float zf[3] = {0.0f, 0.0f, 0.0f};
                                put_hit<SINGLE>(A, slot, 2, 0, 0, z3, 0.0f, zf);
                            }
                        }
                    }
                }
                const uint64_t pm = __ballot(push);
                if (push)
                    ring_put_ray(ring, (tail + (uint32_t)__popcll(pm & lane_lt)) & (kRing - 1), o, d, cv, kind, slot,
                                 (uint32_t)I.gx, (uint32_t)I.gxy);
                tail += (uint32_t)__popcll(pm);
                gr++;
            }
            // ---- lanes without a ray take one from the ring
            const uint64_t needm = __ballot(!hasray);
            const uint32_t avail = tail - head;
            if (!hasray) {
                uint32_t rank = (uint32_t)__popcll(needm & lane_lt);
                if (rank < avail) {
                    uint32_t e = (head + rank) & (kRing - 1);
                    R.tm0 = __uint_as_float(ring[0][e]);
                    R.tm1 = __uint_as_float(ring[1][e]);
                    R.tm2 = __uint_as_float(ring[2][e]);
                    R.td0 = __uint_as_float(ring[3][e]);
                    R.td1 = __uint_as_float(ring[4][e]);
                    R.td2 = __uint_as_float(ring[5][e]);
                    uint32_t xy = ring[6][e], zm = ring[7][e];
                    R.slot = ring[8][e];
                    R.x = (int32_t)(xy & 0xffffu);
                    R.y = (int32_t)(xy >> 16) << 2;
                    R.z = (int32_t)(zm & 0xffffu) << 4;
                    R.sx = (zm & 0x10000u) ? -1 : 1;
                    R.sy = (zm & 0x20000u) ? -4 : 4;
                    R.sz = (zm & 0x40000u) ? -16 : 16;
                    uint32_t kind = (zm >> 19) & 3u;
                    R.meta = (0x700u << 20) | (kind << 16);
                    R.maxt = kind == RAY_SHADOW ? A.maxt[0] : (kind == RAY_AO ? A.maxt[1] : A.maxt[2]);
                    R.ci = ring[9][e];
                    R.cw = grid[R.ci];
                    hasray = true;
                    running = true;
                }
            }
            head += min((uint32_t)__popcll(needm), avail);
            if (__ballot(hasray) == 0 && gen_done && head == tail) break;
        }
        // ---- march: STEPS steps of the running lanes. The wave keeps which
        // lanes run as a lane mask; every lane evaluates the step and only the
        // masked lanes change. A running lane moves its position first, to
        // the voxel the step enters when its own voxel is free, tests the
        // current voxel's bit in the word fetched one step earlier, and then
        // fetches the entered voxel's word: the fetch is in flight through
        // the rest of the step, and the next step uses it. t_max advances
        // only when the voxel was free and the ray goes on; a ray that ends
        // keeps its t_max, and the finish steps its position back. A lane
        // that does not run keeps its position (a lane without a ray sits at
        // word 0).
        __builtin_amdgcn_s_setprio(1);
        uint64_t run = __ballot(running);
        const uint64_t full = __builtin_amdgcn_read_exec();
        uint64_t w = R.cw;
#pragma unroll
        for (int u = 0; u < STEPS; ++u) {
            const uint32_t b = cell_shift((uint32_t)R.x, (uint32_t)R.y, (uint32_t)R.z) & 63u;
            const float tmin = fminf(fminf(R.tm0, R.tm1), R.tm2);
            const uint64_t m0 = __ballot(R.tm0 == tmin), m1 = __ballot(R.tm1 == tmin), m2 = __ballot(R.tm2 == tmin);
            move_axes(R, m0, m1, m2, run, full);
            // the word is fetched only when the position enters another cell.
            // The current voxel's bit is tested first (grid words are
            // bit-reversed: voxel bit b sits at bit 63 - b), so the fetch
            // lands in the register of the word it replaces; the empty asm
            // keeps the test ahead of the fetch (otherwise the word is copied
            // to a second register at every step to keep it for the test).
            const uint32_t ci = grid_word(R.x, R.y, R.z, ggx, ggxy);
            uint64_t occm = __ballot((int64_t)(w << b) < 0);
            asm volatile("" : "+s"(occm) : : "memory");
            if (ci != R.ci) w = grid[ci];
            R.ci = ci;
            // two or three axes step together (rare): best_t sums their t. (A
            // lane without a ray holds distinct t_max values, so it never
            // counts here.)
            // (median == minimum: two vector instructions, where the masks'
            // scalar logic took five scalar ones)
            float bt = tmin;
            const bool tie = __builtin_amdgcn_fmed3f(R.tm0, R.tm1, R.tm2) == tmin;
            if (__builtin_expect(__ballot(tie) != 0, 0)) {
                if (tie) {
                    const float t2 = tmin + tmin;
                    bt = (R.tm0 == tmin && R.tm1 == tmin && R.tm2 == tmin) ? t2 + tmin : t2;
                }
            }
            // A ray enters the march between blocks of STEPS steps with a
            // count of 0 and steps once per iteration while it runs, so its
            // count reaches 256 (a multiple of STEPS) only at a block's last
            // step: the limit is tested there alone.
            if (COUNT) add_masked(R.meta, 1u << 20, run);
            else if (u == STEPS - 1) add_masked(R.meta, (uint32_t)STEPS << 20, run);
            uint64_t ok = __ballot(bt < R.maxt);
            if (u == STEPS - 1) ok &= __ballot((int32_t)R.meta >= 0);
            // An occupied voxel stops the ray there; its step count (counted
            // here) loses that step when it finishes.
            run = run & ~occm & ok;
            advance_tmax(R, m0, m1, m2, run, full);
        }
        R.cw = w;
        running = (run & lane_bit) != 0;
    }
}

// Rays listed by k_march (outside its fast path), one thread each, literal DDA.
template <bool SINGLE>
__global__ void __launch_bounds__(256) k_march_rest(MarchArgs A) {
    const uint32_t n = min(A.counters[8 * kCounterStride], A.deferred_cap);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t slot = A.deferred[j];
        uint32_t l = slot & 63u, tr = slot >> 6;
        uint32_t gt = tr / (uint32_t)A.nrays, r = tr % (uint32_t)A.nrays;
        const MarchItem& I = march_item<SINGLE>(A, item_of_tile<SINGLE>(A, gt));
        uint32_t i = 0;
        (void)tile_lane_pixel(I, gt - I.tile0, l, i);
        PixState P;
        f3 wd;
        pixel_setup(I.T, i, P, wd);
        f3 o, d;
        float maxt;
        (void)pixel_ray(I.T, P, wd, (int32_t)r, o, d, maxt);
        RayOut R;
        R.vox[0] = R.vox[1] = R.vox[2] = 0;
        R.t = 0.0f;
        R.normal = mk3(0.0f, 0.0f, 0.0f);
        dda_exact(I.V, o, d, maxt, R);
        A.res[slot] = R.hit ? R.t : __uint_as_float(kNoHitBits);
        if (A.fres) A.fres[slot] = (uint16_t)R.fetches;
        if (A.hits) {
            const bool h = R.hit != 0;
            const int32_t vox[3] = {h ? R.vox[0] : 0, h ? R.vox[1] : 0, h ? R.vox[2] : 0};
            const float nrm[3] = {h ? R.normal.x : 0.0f, h ? R.normal.y : 0.0f, h ? R.normal.z : 0.0f};
            put_hit<SINGLE>(A, slot, R.term, R.steps, R.fetches, vox, h ? R.t : 0.0f, nrm);
        }
    }
}

// Per-pixel fold of the ray slots in ray order (LightAmbient.frag:160-175:
// shadow lit, mean of the squared AO values, spec value) and sky pixels.
template <bool SINGLE>
__global__ void __launch_bounds__(256) k_resolve(MarchArgs A) {
    const uint32_t gt = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (gt >= A.ntiles) return;
    const uint32_t l = threadIdx.x & 63u;
    // the tile's slots exist for all 64 lanes: the first eight are loaded
    // while the item, the pixel and its material are looked up
    const size_t base = (((size_t)gt * (size_t)A.nrays) << 6) | l;
    float first[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) first[k] = k < A.nrays ? A.res[base + ((size_t)k << 6)] : 0.0f;
    const MarchItem& I = march_item_uniform<SINGLE>(A, item_of_tile<SINGLE>(A, gt));
    uint32_t i = 0;
    if (!tile_lane_pixel(I, gt - I.tile0, l, i)) return;
    const TraceArgs& T = I.T;
    const vx_trace_params& P = T.params;
    if (T.mat[i] == 0) {
        T.shadow[i] = 0; T.ao[i] = 0.0f; T.spec[i] = 0.0f;
        if (T.fetches) T.fetches[i] = 0;
        return;
    }
    uint32_t lit = 1, nf = 0;
    float aosum = 0.0f, spv = 1.0f;
    // eight slots in flight per pass, folded in ray order
    for (int32_t r0 = 0; r0 < A.nrays; r0 += 8) {
        float vv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            vv[k] = r0 == 0 ? first[k] : (r0 + k < A.nrays ? A.res[base + ((size_t)(r0 + k) << 6)] : 0.0f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t r = r0 + k;
            if (r >= A.nrays) break;
            const float v = vv[k];
            const bool hit = __float_as_uint(v) != kNoHitBits;
            if (r < P.n_shadow) {
                lit = hit ? 0u : 1u;
            } else {
                const float maxt = r < P.n_shadow + P.n_ao ? P.ao_maxt : P.spec_maxt;
                const float q = ray_value(RAY_AO, hit ? 1 : 0, v, maxt);
                if (r < P.n_shadow + P.n_ao) aosum = aosum + q * q;
                else spv = q;
            }
        }
        if (A.fres) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (r0 + k < A.nrays) nf += A.fres[base + ((size_t)(r0 + k) << 6)];
        }
    }
    T.shadow[i] = (uint8_t)lit;
    T.ao[i] = P.n_ao > 0 ? aosum / (float)P.n_ao : 1.0f;
    T.spec[i] = P.n_spec > 0 ? spv : 1.0f;
    if (T.fetches) T.fetches[i] = nf;
}

// Launchers (template instances live in this file).
template <int BLOCK, int STEPS>
static void launch_march_t(bool single, uint32_t grid, hipStream_t s, const MarchArgs& A) {
    // the per-step count only where fetch counts or hit records are asked for
    const bool count = A.fres != nullptr || A.hits != nullptr;
    if (single && count) hipLaunchKernelGGL((k_march<BLOCK, STEPS, true, true>), dim3(grid), dim3(BLOCK), 0, s, A);
    else if (single) hipLaunchKernelGGL((k_march<BLOCK, STEPS, true, false>), dim3(grid), dim3(BLOCK), 0, s, A);
    else if (count) hipLaunchKernelGGL((k_march<BLOCK, STEPS, false, true>), dim3(grid), dim3(BLOCK), 0, s, A);
    else hipLaunchKernelGGL((k_march<BLOCK, STEPS, false, false>), dim3(grid), dim3(BLOCK), 0, s, A);
}

template <int BLOCK, int STEPS>
static const void* march_kernel_t(bool single) {
    return single ? reinterpret_cast<const void*>(&k_march<BLOCK, STEPS, true, false>)
                  : reinterpret_cast<const void*>(&k_march<BLOCK, STEPS, false, false>);
}

const void* march_kernel(int block, int steps, bool single) {
    if (block == 256) {
        if (steps == 2) return march_kernel_t<256, 2>(single);
        if (steps == 8) return march_kernel_t<256, 8>(single);
        if (steps == 16) return march_kernel_t<256, 16>(single);
        return march_kernel_t<256, 4>(single);
    }
    if (steps == 2) return march_kernel_t<512, 2>(single);
    if (steps == 8) return march_kernel_t<512, 8>(single);
    if (steps == 16) return march_kernel_t<512, 16>(single);
    return march_kernel_t<512, 4>(single);
}

hipError_t launch_march(int block, int steps, bool single, uint32_t grid, uint32_t rest_grid, hipStream_t s,
                        const MarchArgs& A, hipEvent_t e0, hipEvent_t e1) {
    if (e0) {
        hipError_t er = hipEventRecord(e0, s);
        if (er != hipSuccess) return er;
    }
    if (block == 256) {
        if (steps == 2) launch_march_t<256, 2>(single, grid, s, A);
        else if (steps == 8) launch_march_t<256, 8>(single, grid, s, A);
        else if (steps == 16) launch_march_t<256, 16>(single, grid, s, A);
        else launch_march_t<256, 4>(single, grid, s, A);
    } else {
        if (steps == 2) launch_march_t<512, 2>(single, grid, s, A);
        else if (steps == 8) launch_march_t<512, 8>(single, grid, s, A);
        else if (steps == 16) launch_march_t<512, 16>(single, grid, s, A);
        else launch_march_t<512, 4>(single, grid, s, A);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (e1) {
        e = hipEventRecord(e1, s);
        if (e != hipSuccess) return e;
    }
    uint32_t rgrid = (A.ntiles + 3u) / 4u;
    if (single) {
        hipLaunchKernelGGL(k_march_rest<true>, dim3(rest_grid), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_resolve<true>, dim3(rgrid), dim3(256), 0, s, A);
    } else {
        hipLaunchKernelGGL(k_march_rest<false>, dim3(rest_grid), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_resolve<false>, dim3(rgrid), dim3(256), 0, s, A);
    }
    return hipGetLastError();
}

// March grid words for grid cells [g0, g0 + n) (grid coordinates, border
// cells included): the occupancy cell plus the out-of-bounds voxel bits, bit
// reversed (voxel bit b stored at bit 63 - b, so a left shift by b moves it to
// the sign bit).
__global__ void k_mgrid(uint64_t* grid, OccView V, int32_t gx, int32_t gy, int32_t g0x, int32_t g0y, int32_t g0z,
                        int32_t nx, int32_t ny, int32_t nz) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nx * ny * nz) return;
    int32_t cx = g0x + (int32_t)(idx % (size_t)nx);
    int32_t cy = g0y + (int32_t)((idx / (size_t)nx) % (size_t)ny);
    int32_t cz = g0z + (int32_t)(idx / ((size_t)nx * ny));
    const int32_t f0[3] = {4 * (cx - 1), 4 * (cy - 1), 4 * (cz - 1)};
    const int32_t lim[3] = {2 * V.sx + 1, 2 * V.sy + 1, 2 * V.sz + 1};
    uint32_t om[3];
    bool any_in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        om[k] = 0;
        for (int p = 0; p < 4; ++p)
            if ((uint32_t)(f0[k] + p) > (uint32_t)lim[k]) om[k] |= 1u << p;
        any_in = any_in && om[k] != 15u;
    }
    uint64_t w = 0;
    if (any_in) w = V.cells[cell_index(V, f0[0], f0[1], f0[2])];
    for (uint32_t b = 0; b < 64; ++b) {
        uint32_t px = b & 3u, py = (b >> 2) & 3u, pz = b >> 4;
        if (((om[0] >> px) | (om[1] >> py) | (om[2] >> pz)) & 1u) w |= 1ull << b;
    }
    grid[((size_t)cz * gy + cy) * gx + cx] = __builtin_bitreverse64(w);
}

// ===========================================================================
// Point and spot lights (LightPoint.frag:85-152, LightSpot.frag:73-140): one
// thread per pixel, the lights' draws as a loop in draw order, world space.
//
// A block (a 16x16 tile) first bounds its hit pixels' world positions and
// keeps, in draw order, the lights whose sphere (range plus a wide margin)
// reaches that box; every pixel still applies the reference's range test, so
// a light dropped here is one the test would skip at every pixel of the tile.
// A tile with a non-finite position keeps every light.
// ===========================================================================
constexpr uint32_t kMaxLights = 4096;

// The lights of a 16x16 block tile that can reach its hit pixels, in draw
// order, into list_s; returns their count. Every thread of the block calls it
// (it synchronizes the block). valid / wpos: this thread's pixel.
VXD uint32_t tile_light_list(const LightArgs& A, bool valid, f3 wpos, uint16_t* list_s, float (*box_s)[6],
                             uint32_t* wcount_s, uint32_t* bad_s) {
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    if (t == 0) *bad_s = 0;
    const bool finite = isfinite(wpos.x) && isfinite(wpos.y) && isfinite(wpos.z);
    float lo[3] = {valid ? wpos.x : INFINITY, valid ? wpos.y : INFINITY, valid ? wpos.z : INFINITY};
    float hi[3] = {valid ? wpos.x : -INFINITY, valid ? wpos.y : -INFINITY, valid ? wpos.z : -INFINITY};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
    }
    const bool wave_bad = __ballot(valid && !finite) != 0;
    __syncthreads();   // bad_s cleared
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { box_s[wave][k] = lo[k]; box_s[wave][3 + k] = hi[k]; }
        if (wave_bad) *bad_s = 1;
    }
    __syncthreads();
    float blo[3], bhi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        blo[k] = fminf(fminf(box_s[0][k], box_s[1][k]), fminf(box_s[2][k], box_s[3][k]));
        bhi[k] = fmaxf(fmaxf(box_s[0][3 + k], box_s[1][3 + k]), fmaxf(box_s[2][3 + k], box_s[3][3 + k]));
    }
    const bool keep_all = *bad_s != 0;
    // 256 lights at a time, each wave's survivors after those of the waves before it
    uint32_t n = 0;
    for (uint32_t c0 = 0; c0 < A.nlights; c0 += 256) {
        const uint32_t l = c0 + t;
        bool keep = false;
        if (l < A.nlights) {
            const vx_light& L = A.lights[l];
            float d2 = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float q = fmaxf(fmaxf(blo[k] - L.position[k], L.position[k] - bhi[k]), 0.0f);
                d2 = d2 + q * q;
            }
            const float r = L.range * 1.001f + 0.01f;
            keep = keep_all || !(d2 > r * r);   // NaN range or box: kept
        }
        const uint64_t km = __ballot(keep);
        if (lane == 0) wcount_s[wave] = (uint32_t)__popcll(km);
        __syncthreads();
        uint32_t base = n;
        for (uint32_t w = 0; w < wave; ++w) base += wcount_s[w];
        if (keep) list_s[base + (uint32_t)__popcll(km & ((1ull << lane) - 1ull))] = (uint16_t)l;
        n += (wcount_s[0] + wcount_s[1]) + (wcount_s[2] + wcount_s[3]);
        __syncthreads();   // counts read, list written
    }
    return n;
}

// A hit pixel's inputs to every light's term (LightPoint.frag:95-108,
// LightSpot.frag:83-96; PBRDirectLight's F and kD, PBR.frag:48-59).
struct LightPix {
    f3 wpos, N, rv;
    float r3, kD;
    const float* e;
};

VXD void light_pixel_setup(const LightArgs& A, size_t i, LightPix& P) {
    const f3 pp = load3(A.pos + 3 * i);
    P.wpos = mk3(pp.x * 0.1f, pp.y * 0.1f, pp.z * 0.1f);
    P.N = load3(A.nrm + 3 * i);
    const uint32_t pk = pixkey(A.params.seed, A.params.frame, (uint32_t)i);
    float r0[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r0[k] = rand01(pk, (uint32_t)k);
    const f3 cam = mk3(A.view.cam_pos[0] * 10.0f, A.view.cam_pos[1] * 10.0f, A.view.cam_pos[2] * 10.0f);
    const f3 Vd = normalize3(mk3(cam.x - pp.x, cam.y - pp.y, cam.z - pp.z));
    P.e = A.palette + 6 * (uint32_t)A.mat[i];
    const float metal = P.e[4];
    const float ndv = gmax(dot3(P.N, Vd), 0.0f);
    const float x = gmax(1.0f - ndv, 0.0f);
    const float s5 = (((x * x) * x) * x) * x;
    // PBRDirectLight (PBR.frag:48-68): F0 = mix(0.04, 1, metallic)
    const float f0d = 0.04f * (1.0f - metal) + 1.0f * metal;
    const float F = f0d + (1.0f - f0d) * s5;
    P.kD = (1.0f - F) * (1.0f - metal);
    f3 rv = cosine_hemisphere(r0[0], r0[1]);
    rv = mk3(rv.x * 0.1f, rv.y * 0.1f, rv.z * 0.1f);
    rv.z = rv.z * gsign(r0[2] - 0.5f);
    P.rv = rv;
    P.r3 = r0[3];
}

// The range test of a light draw (LightPoint.frag:99-100): false outside.
VXD bool light_in_range(const vx_light& L, f3 wpos, f3& ld, float& dist) {
    ld = mk3(L.position[0] - wpos.x, L.position[1] - wpos.y, L.position[2] - wpos.z);
    dist = sqrtf(dot3(ld, ld));
    return !(dist > L.range);
}

// The light's shadow ray (LightPoint.frag:110-125, LightSpot.frag:98-113).
VXD void light_ray(const vx_light& L, const LightPix& P, f3 ld, float dist, f3& o, f3& wdn, float& hit_dist) {
    const bool spot = L.type == VX_LIGHT_SPOT;
    f3 wd;
    if (spot) {
        const f3 nl = normalize3(ld);
        wd = mk3(nl.x * 10.0f, nl.y * 10.0f, nl.z * 10.0f);
        hit_dist = dist * 10.0f;
    } else {
        wd = ld;
        hit_dist = dist * 10.5f;
    }
    const f3 rv = P.rv;
    wdn = normalize3(mk3(wd.x * (1.0f - 0.5f) + rv.x * 0.5f, wd.y * (1.0f - 0.5f) + rv.y * 0.5f,
                         wd.z * (1.0f - 0.5f) + rv.z * 0.5f));
    float c0 = P.wpos.x * 10.0f, c1 = P.wpos.y * 10.0f, c2 = P.wpos.z * 10.0f;
    c0 = c0 + wdn.x * (P.r3 * 1.0f); c1 = c1 + wdn.y * (P.r3 * 1.0f); c2 = c2 + wdn.z * (P.r3 * 1.0f);
    c0 = c0 + rv.x * 2.5f; c1 = c1 + rv.y * 2.5f; c2 = c2 + rv.z * 2.5f;
    o = mk3(c0 + P.N.x * 0.5f, c1 + P.N.y * 0.5f, c2 + P.N.z * 0.5f);
}

// The light's shadowed direct term per channel (LightPoint.frag:127-150,
// LightSpot.frag:115-139 with PBRDirectLight): the value the reference's
// additive blending adds to the light buffer.
VXD void light_term(const vx_light& L, const LightPix& P, f3 ld, float dist, float shadow, float term[3]) {
    float att = gclamp(L.range - dist, 0.0f, 1.0f) / pow_vx(dist, L.attenuation);
    if (L.type == VX_LIGHT_SPOT) {
        const f3 nl = normalize3(ld);
        const f3 dir = mk3(L.direction[0], L.direction[1], L.direction[2]);
        const float ad = gmax(dot3(nl, dir) - (1.0f - L.angle), 0.0f) / L.angle;
        att = (pow_vx(ad, L.angle_attenuation) * gclamp(L.range - dist, 0.0f, 1.0f)) / pow_vx(dist, L.attenuation);
    }
    const float ndl = gmax(dot3(P.N, ld), 0.0f);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float rad = (L.color[c] * att) * shadow;
        term[c] = (((P.kD * P.e[c]) / 3.14159265359f) * rad) * ndl;
    }
}

__global__ void __launch_bounds__(256) k_lights(OccView V, LightArgs A) {
    __shared__ uint16_t list_s[kMaxLights];
    __shared__ float box_s[4][6];
    __shared__ uint32_t wcount_s[4];
    __shared__ uint32_t bad_s;
    const uint32_t logical = (blockIdx.x % 8u) * A.tiles.per_xcd + blockIdx.x / 8u;
    if (logical >= A.tiles.ntiles) return;   // block-uniform
    int32_t px, py;
    const bool inview = tile_pixel(A.tiles, px, py);
    const size_t i = inview ? (size_t)py * A.view.width + px : 0;
    const uint32_t m = inview ? A.mat[i] : 0u;
    const bool valid = m != 0;
    f3 pp = mk3(0.0f, 0.0f, 0.0f);
    if (valid) pp = load3(A.pos + 3 * i);
    const uint32_t n = tile_light_list(A, valid, mk3(pp.x * 0.1f, pp.y * 0.1f, pp.z * 0.1f), list_s, box_s,
                                       wcount_s, &bad_s);
    if (!inview) return;
    if (!valid) {
        if (A.fetches) A.fetches[i] = 0;
        return;
    }
    LightPix P;
    light_pixel_setup(A, i, P);
    float acc[3] = {A.rgb[3 * i + 0], A.rgb[3 * i + 1], A.rgb[3 * i + 2]};
    uint32_t nf = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const vx_light& L = A.lights[__builtin_amdgcn_readfirstlane(list_s[k])];
        f3 ld;
        float dist;
        if (!light_in_range(L, P.wpos, ld, dist)) continue;
        f3 o, wdn;
        float hit_dist;
        light_ray(L, P, ld, dist, o, wdn, hit_dist);
        const bool spot = L.type == VX_LIGHT_SPOT;
        int32_t hit;
        const float d = march_dist<kSparsePrefetch>(V, spot ? A.params.spot_march : A.params.point_march, o, wdn, hit_dist, nf, hit);
        float term[3];
        light_term(L, P, ld, dist, (d < hit_dist) ? 0.0f : 1.0f, term);
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + term[c];
    }
    A.rgb[3 * i + 0] = acc[0];
    A.rgb[3 * i + 1] = acc[1];
    A.rgb[3 * i + 2] = acc[2];
    if (A.fetches) A.fetches[i] = nf;
}

// ---------------------------------------------------------------------------
// The light pass as a ray list: a pixel's (pixel, light) shadow rays are
// listed, marched one per lane, and their terms folded per pixel in light
// order. Pass 1 counts each pixel's in-range lights and gives each 16x16
// tile a contiguous range of the list (one atomic per tile) and each pixel
// a contiguous range of slots in it; pass 2 sets each pixel up once and
// writes the tile's rays (origin, direction, length, marcher, slot) light by
// light, each light's in-range pixels in pixel order, so that neighbouring
// rays run from neighbouring pixels to one light; pass 3 marches every ray
// and stores its shadow factor in its slot; pass 4 sets each pixel up again
// and adds its lights' terms to its light buffer value in slot order, which
// is draw order. The arithmetic is k_lights' (the same device functions).
// ---------------------------------------------------------------------------
template <bool WRITE>
VXD void light_list_tile(const LightArgs& A, const LightListArgs& B) {
    __shared__ uint16_t list_s[kMaxLights];
    __shared__ float box_s[4][6];
    __shared__ uint32_t wcount_s[4];
    __shared__ uint32_t bad_s;
    __shared__ uint32_t wsum_s[4];
    __shared__ uint32_t tile_base_s;
    const uint32_t logical = (blockIdx.x % 8u) * A.tiles.per_xcd + blockIdx.x / 8u;
    if (logical >= A.tiles.ntiles) return;   // block-uniform
    int32_t px, py;
    const bool inview = tile_pixel(A.tiles, px, py);
    const size_t i = inview ? (size_t)py * A.view.width + px : 0;
    const bool valid = inview && A.mat[i] != 0;
    f3 wpos = mk3(0.0f, 0.0f, 0.0f);
    if (valid) {
        const f3 pp = load3(A.pos + 3 * i);
        wpos = mk3(pp.x * 0.1f, pp.y * 0.1f, pp.z * 0.1f);
    }
    const uint32_t n = tile_light_list(A, valid, wpos, list_s, box_s, wcount_s, &bad_s);
    if (!WRITE) {
        uint32_t c = 0;
        if (valid)
            for (uint32_t k = 0; k < n; ++k) {
                f3 ld;
                float dist;
                c += light_in_range(A.lights[__builtin_amdgcn_readfirstlane(list_s[k])], wpos, ld, dist) ? 1u : 0u;
            }
        // the tile's exclusive scan: in-wave, then over the four waves
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        uint32_t x = c;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)lane >= off) x += y;
        }
        if (lane == 63u) wsum_s[wave] = x;
        __syncthreads();
        uint32_t wbase = 0;
        for (uint32_t w = 0; w < wave; ++w) wbase += wsum_s[w];
        if (threadIdx.x == 0)
            tile_base_s = atomicAdd(B.total, (wsum_s[0] + wsum_s[1]) + (wsum_s[2] + wsum_s[3]));
        __syncthreads();
        if (inview) {
            B.first[i] = tile_base_s + wbase + (x - c);
            B.count[i] = c;
        }
    } else {
        // thread 0's pixel is the tile's first (always in view): its first
        // slot is the tile's first record
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        const uint64_t lane_lt = (1ull << lane) - 1ull;
        if (threadIdx.x == 0) tile_base_s = B.first[i];
        uint32_t slot = valid ? B.first[i] : 0u;
        LightPix P;
        if (valid) light_pixel_setup(A, i, P);
        __syncthreads();
        uint32_t off = tile_base_s;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t l = list_s[k];
            const vx_light& L = A.lights[__builtin_amdgcn_readfirstlane(l)];
            f3 ld;
            float dist;
            const bool in = valid && light_in_range(L, wpos, ld, dist);
            const uint64_t m = __ballot(in);
            if (lane == 0) wsum_s[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t at = off + (uint32_t)__popcll(m & lane_lt);
            for (uint32_t w = 0; w < wave; ++w) at += wsum_s[w];
            if (in) {
                f3 o, wdn;
                float hit_dist;
                light_ray(L, P, ld, dist, o, wdn, hit_dist);
                LightRec& r = B.rec[at];
                r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z;
                r.d[0] = wdn.x; r.d[1] = wdn.y; r.d[2] = wdn.z;
                r.hit_dist = hit_dist;
                r.march = L.type == VX_LIGHT_SPOT ? A.params.spot_march : A.params.point_march;
                r.slot = slot;
                B.slot_light[slot] = (uint16_t)l;
                ++slot;
            }
            off += (wsum_s[0] + wsum_s[1]) + (wsum_s[2] + wsum_s[3]);
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(256) k_light_count(LightArgs A, LightListArgs B) { light_list_tile<false>(A, B); }
__global__ void __launch_bounds__(256) k_light_records(LightArgs A, LightListArgs B) { light_list_tile<true>(A, B); }

// One shadow ray per thread over the list (grid-stride to the device total):
// its shadow factor and fetch count into its slot.
__global__ void __launch_bounds__(256) k_light_march(OccView V, LightArgs A, LightListArgs B) {
    const uint32_t total = B.total[0];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
        const LightRec r = B.rec[j];
        uint32_t nf = 0;
        int32_t hit;
        const float d = march_dist<kSparsePrefetch>(V, r.march, mk3(r.o[0], r.o[1], r.o[2]), mk3(r.d[0], r.d[1], r.d[2]), r.hit_dist,
                                   nf, hit);
        B.shadow[r.slot] = (d < r.hit_dist) ? 0.0f : 1.0f;
        if (B.rfetch) B.rfetch[r.slot] = (uint16_t)nf;
    }
}

// A pixel's terms added to its light buffer value in list (draw) order.
__global__ void __launch_bounds__(256) k_light_fold(LightArgs A, LightListArgs B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(A.view.width * A.view.height)) return;
    if (A.mat[i] == 0) {
        if (A.fetches) A.fetches[i] = 0;
        return;
    }
    const uint32_t j0 = B.first[i], n = B.count[i];
    float acc[3] = {A.rgb[3 * (size_t)i + 0], A.rgb[3 * (size_t)i + 1], A.rgb[3 * (size_t)i + 2]};
    uint32_t nf = 0;
    LightPix P;
    if (n) light_pixel_setup(A, i, P);
    for (uint32_t k = 0; k < n; ++k) {
        const vx_light& L = A.lights[B.slot_light[j0 + k]];
        f3 ld;
        float dist;
        (void)light_in_range(L, P.wpos, ld, dist);
        float term[3];
        light_term(L, P, ld, dist, B.shadow[j0 + k], term);
        acc[0] = acc[0] + term[0];
        acc[1] = acc[1] + term[1];
        acc[2] = acc[2] + term[2];
        if (A.fetches) nf += B.rfetch[j0 + k];
    }
    A.rgb[3 * (size_t)i + 0] = acc[0];
    A.rgb[3 * (size_t)i + 1] = acc[1];
    A.rgb[3 * (size_t)i + 2] = acc[2];
    if (A.fetches) A.fetches[i] = nf;
}

// ===========================================================================
// Primary hit: GeometryVoxel.frag (LOD 1, MIP_COUNT 2, TRANSPARENCY 1) for the
// world model with an identity WorldMatrix.
// ===========================================================================
// Byte index (c2 * my + c1) * mx + c0 of cell c of a level in 32 bits, as
// two v_mad_u32_u24 (the compiler's own form is three quarter-rate
// v_mad_u64_u32): exact for a cell inside the level, the levels being at most
// 4096 per axis (c2 * my + c1 < my * mz <= 2^24); a cell outside reads byte 0.
VXD uint32_t level_index32(int32_t c0, int32_t c1, int32_t c2, uint32_t mx, uint32_t my) {
    uint32_t a, r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"((uint32_t)c2), "v"(my), "v"((uint32_t)c1));
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(mx), "v"((uint32_t)c0));
    return r;
}
// c < -1 || c > up on any axis, without branches: c + 1 as unsigned exceeds
// up + 1 exactly then (c = INT_MAX wraps to 2^31).
VXD bool level_oob(int32_t c0, int32_t c1, int32_t c2, const int32_t up[3]) {
    return ((uint32_t)(c0 + 1) > (uint32_t)(up[0] + 1)) | ((uint32_t)(c1 + 1) > (uint32_t)(up[1] + 1)) |
           ((uint32_t)(c2 + 1) > (uint32_t)(up[2] + 1));
}

// intersectVolume, GeometryVoxel.frag:63-125, in the model's voxel space.
// res / frame: the view's resolution and frame (the glass dither).
// The select products step(a,b)*step(a,c) are 0 or 1 and sign(d) is -1, 0 or
// 1, so the voxel step ivec3(select * stepSign) is taken as an integer select;
// t_max, best_t and the normal keep the reference's float products (an
// infinite t_delta times a 0 select is NaN there too).
// WIDTH: 1 when every level's voxel index fits 32 bits (mip 0 holds fewer than
// 2^32 voxels), 2 when it does not, 0 to decide at each step. The single-model
// pass takes the fixed forms; the instance pass keeps the test in the loop,
// which leaves it fewer registers and more waves.
// K: steps per loop. The steps of one level do not depend on what they read,
// so with K > 1 a thread takes its next K steps, requests their K voxels
// together and then tests them in order (one fetch latency per K steps);
// the first step that leaves the volume or meets a voxel is handled as the
// reference handles it, and the march goes on from it.
template <int WIDTH, int K = 1>
VXD int32_t intersect_volume_t(const ModelMips& A, const float res[2], int32_t frame, f3 o3, f3 d3, f3 cam,
                               float uvx, float uvy, f3& hit, f3& nrm, uint32_t& material) {
    float o[3] = {o3.x, o3.y, o3.z}, d[3] = {d3.x, d3.y, d3.z}, c[3] = {cam.x, cam.y, cam.z};
    float s[3], td[3];
    int32_t is[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s[k] = gsign(d[k]);
        td[k] = 1.0f / (d[k] * s[k]);
        is[k] = (int32_t)s[k];
    }
    const bool narrow =
        WIDTH == 1 || (WIDTH == 0 && (uint64_t)(uint32_t)A.dims[0] * (uint32_t)A.dims[1] * (uint32_t)A.dims[2] < (1ull << 32));
    int32_t mip = 2;
    int32_t i = 0;
    do {
        float mip_size = (float)(1 << mip);
        const uint8_t* base = mip == 2 ? A.mips[2] : (mip == 1 ? A.mips[1] : A.mips[0]);
        const uint32_t mx = (uint32_t)(A.dims[0] >> mip), my = (uint32_t)(A.dims[1] >> mip),
                       mz = (uint32_t)(A.dims[2] >> mip);
        int32_t cv[3], up[3];
        float tm[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[k] = o[k] / mip_size;
            cv[k] = f2i(floorf(o[k]));
            float nb = (float)cv[k] + (s[k] * 0.5f + 0.5f);
            tm[k] = (nb - o[k]) / d[k];
            up[k] = f2i((float)A.dims[k] / mip_size) + 1;
        }
        int32_t n = 0;
        if constexpr (K > 1) {
            while (n < 512) {
                int32_t cb[K][3];
                float tb[K][3];
                uint32_t sb[K], vb[K];
                bool ob[K], ib[K];
                float t0 = tm[0], t1 = tm[1], t2 = tm[2];
                int32_t c0 = cv[0], c1 = cv[1], c2 = cv[2];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const bool b0 = !(t2 < t0) && !(t1 < t0);
                    const bool b1 = !(t0 < t1) && !(t2 < t1);
                    const bool b2 = !(t1 < t2) && !(t0 < t2);
                    tb[k][0] = t0; tb[k][1] = t1; tb[k][2] = t2;
                    sb[k] = (b0 ? 1u : 0u) | (b1 ? 2u : 0u) | (b2 ? 4u : 0u);
                    c0 += b0 ? is[0] : 0;
                    c1 += b1 ? is[1] : 0;
                    c2 += b2 ? is[2] : 0;
                    cb[k][0] = c0; cb[k][1] = c1; cb[k][2] = c2;
                    ob[k] = level_oob(c0, c1, c2, up);
                    ib[k] = (uint32_t)c0 < mx && (uint32_t)c1 < my && (uint32_t)c2 < mz;
                    if (narrow) {
                        const uint32_t idx = level_index32(c0, c1, c2, mx, my);
                        vb[k] = base[ib[k] ? idx : 0u];
                    } else {
                        const size_t idx = (size_t)(uint32_t)c0 + (size_t)(uint32_t)c1 * mx + (size_t)(uint32_t)c2 * mx * my;
                        vb[k] = base[ib[k] ? idx : (size_t)0];
                    }
                    t0 += td[0] * (b0 ? 1.0f : 0.0f);
                    t1 += td[1] * (b1 ? 1.0f : 0.0f);
                    t2 += td[2] * (b2 ? 1.0f : 0.0f);
                }
                // the first step of the level's 512 that leaves or meets a voxel
                int32_t e = K;
#pragma unroll
                for (int k = K - 1; k >= 0; --k)
                    if (n + k < 512 && (ob[k] || (ib[k] && vb[k] != 0u))) e = k;
                if (e == K) {
                    cv[0] = c0; cv[1] = c1; cv[2] = c2;
                    tm[0] = t0; tm[1] = t1; tm[2] = t2;
                    n += K;
                    continue;
                }
                uint32_t se = 0, vox = 0;
                bool oe = false;
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (k == e) {
                        cv[0] = cb[k][0]; cv[1] = cb[k][1]; cv[2] = cb[k][2];
                        tm[0] = tb[k][0]; tm[1] = tb[k][1]; tm[2] = tb[k][2];
                        se = sb[k];
                        vox = ib[k] ? vb[k] : 0u;
                        oe = ob[k];
                    }
                if (oe) return 0;
                const float s0 = (se & 1u) ? 1.0f : 0.0f, s1 = (se & 2u) ? 1.0f : 0.0f, s2 = (se & 4u) ? 1.0f : 0.0f;
                float best_t = (tm[0] * s0 + tm[1] * s1) + tm[2] * s2;
                bool accept = (mip == 0);
                if (!accept) {
                    f3 df = mk3(c[0] - (o[0] + d[0] * best_t) * mip_size, c[1] - (o[1] + d[1] * best_t) * mip_size,
                                c[2] - (o[2] + d[2] * best_t) * mip_size);
                    accept = (float)mip < 0.001f * sqrtf(dot3(df, df));
                }
                if (!accept) {
                    mip--;
#pragma unroll
                    for (int k = 0; k < 3; ++k) o[k] = o[k] + d[k] * best_t / mip_size;
                    break;
                }
                float cx = roundf((uvx * res[0]) * 0.5f);
                float cy = roundf((uvy * res[1]) * 0.5f);
                bool glass = vox < 16u && gmod(cy + cx, 2.0f) == (float)(frame % 2);
                if (!glass) {
                    nrm = mk3(-s[0] * s0, -s[1] * s1, -s[2] * s2);
                    hit = mk3((o[0] + d[0] * best_t) * mip_size, (o[1] + d[1] * best_t) * mip_size,
                              (o[2] + d[2] * best_t) * mip_size);
                    material = vox;
                    return 1;
                }
                // a glass voxel the dither lets through: the march goes on after it
                tm[0] += td[0] * s0;
                tm[1] += td[1] * s1;
                tm[2] += td[2] * s2;
                n += e + 1;
            }
        } else {
        do {
            const bool b0 = !(tm[2] < tm[0]) && !(tm[1] < tm[0]);
            const bool b1 = !(tm[0] < tm[1]) && !(tm[2] < tm[1]);
            const bool b2 = !(tm[1] < tm[2]) && !(tm[0] < tm[2]);
            const float s0 = b0 ? 1.0f : 0.0f, s1 = b1 ? 1.0f : 0.0f, s2 = b2 ? 1.0f : 0.0f;
            cv[0] += b0 ? is[0] : 0;
            cv[1] += b1 ? is[1] : 0;
            cv[2] += b2 ? is[2] : 0;
            const bool oob = level_oob(cv[0], cv[1], cv[2], up);
            // the texel is read unconditionally (byte 0 of the level for a cell
            // outside it, then taken as empty), so the step has one branch
            const bool inside = (uint32_t)cv[0] < mx && (uint32_t)cv[1] < my && (uint32_t)cv[2] < mz;
            uint32_t vox;
            if (narrow) {
                const uint32_t idx = level_index32(cv[0], cv[1], cv[2], mx, my);
                vox = base[inside ? idx : 0u];
            } else {
                const size_t idx = (size_t)(uint32_t)cv[0] + (size_t)(uint32_t)cv[1] * mx + (size_t)(uint32_t)cv[2] * mx * my;
                vox = base[inside ? idx : (size_t)0];
            }
            vox = inside ? vox : 0u;
            if (oob || vox != 0u) {
                if (oob) return 0;
                float best_t = (tm[0] * s0 + tm[1] * s1) + tm[2] * s2;
                bool accept = (mip == 0);
                if (!accept) {
                    f3 df = mk3(c[0] - (o[0] + d[0] * best_t) * mip_size, c[1] - (o[1] + d[1] * best_t) * mip_size,
                                c[2] - (o[2] + d[2] * best_t) * mip_size);
                    accept = (float)mip < 0.001f * sqrtf(dot3(df, df));
                }
                if (accept) {
                    float cx = roundf((uvx * res[0]) * 0.5f);
                    float cy = roundf((uvy * res[1]) * 0.5f);
                    bool glass = vox < 16u && gmod(cy + cx, 2.0f) == (float)(frame % 2);
                    if (!glass) {
                        nrm = mk3(-s[0] * s0, -s[1] * s1, -s[2] * s2);
                        hit = mk3((o[0] + d[0] * best_t) * mip_size, (o[1] + d[1] * best_t) * mip_size,
                                  (o[2] + d[2] * best_t) * mip_size);
                        material = vox;
                        return 1;
                    }
                } else {
                    mip--;
#pragma unroll
                    for (int k = 0; k < 3; ++k) o[k] = o[k] + d[k] * best_t / mip_size;
                    break;
                }
            }
            tm[0] += td[0] * s0;
            tm[1] += td[1] * s1;
            tm[2] += td[2] * s2;
        } while (++n < 512);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = o[k] * mip_size;
    } while (++i < 4);
    return 0;
}

// Steps per loop of the single-model primary pass (1, 2, 4 and 8 measured
// 1.25, 1.02, 1.00 and 1.00 ms on the F frame; the instance pass, with
// fewer registers to spare, measured 0.35, 0.37 and 0.43 ms at 1, 2 and 4 and
// keeps 1).
constexpr int kPrimarySteps = 4;

VXD int32_t intersect_volume(const ModelMips& A, const float res[2], int32_t frame, f3 o3, f3 d3, f3 cam,
                             float uvx, float uvy, f3& hit, f3& nrm, uint32_t& material) {
    if ((uint64_t)(uint32_t)A.dims[0] * (uint32_t)A.dims[1] * (uint32_t)A.dims[2] < (1ull << 32))
        return intersect_volume_t<1, kPrimarySteps>(A, res, frame, o3, d3, cam, uvx, uvy, hit, nrm, material);
    return intersect_volume_t<2, kPrimarySteps>(A, res, frame, o3, d3, cam, uvx, uvy, hit, nrm, material);
}

// Pinhole ray of pixel (px, py) from the inverse projection / view
// (LightAmbient.vert:32-36), normalized; ndc: the pixel's normalized coordinates.
VXD f3 pixel_dir(const vx_view& vw, int32_t px, int32_t py, float& ndc_x, float& ndc_y) {
    float u = ((float)px + 0.5f + vw.jitter[0]) / vw.res[0];
    float v = ((float)py + 0.5f + vw.jitter[1]) / vw.res[1];
    ndc_x = u * 2.0f - 1.0f;
    ndc_y = 1.0f - v * 2.0f;
    float c4[4] = {ndc_x, ndc_y, 1.0f, 1.0f}, f4[4], w4[4];
    mat_vec(vw.inv_proj, c4, f4);
    float fv[4] = {f4[0] / f4[3], f4[1] / f4[3], f4[2] / f4[3], 0.0f};
    mat_vec(vw.inv_view, fv, w4);
    return normalize3(mk3(w4[0], w4[1], w4[2]));
}

// clipToAABB, GeometryVoxel.frag:49-60.
VXD f3 clip_to_aabb(f3 o, f3 d, const int32_t dims[3]) {
    float size[3] = {(float)dims[0], (float)dims[1], (float)dims[2]};
    float co[3] = {o.x, o.y, o.z}, cd[3] = {d.x, d.y, d.z};
    bool inside = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (gclamp(co[k], 0.0f, size[k]) != co[k]) inside = false;
    if (inside) return o;
    float t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float inv = 1.0f / cd[k];
        float sgn = gstep(cd[k], 0.0f);
        t[k] = (sgn * size[k] - co[k]) * inv;
    }
    float tmin = gmax(gmax(t[0], t[1]), t[2]);
    return mk3(co[0] + cd[0] * (tmin - 0.001f), co[1] + cd[1] * (tmin - 0.001f), co[2] + cd[2] * (tmin - 0.001f));
}

// Whether a draw's box [0, dims] (model space, fine voxels) rasterizes at a
// pixel. The CUBE triangles (GeometryVoxel.vert:5-42) face inward and the
// pipeline culls back faces with CCW fronts (evk.cpp:470-471), so the far
// faces are drawn: the pixel is covered when its forward ray (o, d: slabs in
// model voxel units from the local camera) leaves the box at a view depth >=
// NEAR (the near-plane clip). world maps model space (in tenths) to world
// space, nullptr for the identity; pv = projection * view.
VXD bool box_far_face_drawn(const int32_t dims[3], f3 lc, f3 dl, const float* world, const float* pv) {
    const float o[3] = {lc.x, lc.y, lc.z}, d[3] = {dl.x, dl.y, dl.z};
    float tn = -INFINITY, tf = INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float inv = 1.0f / d[k];
        const float t0 = (0.0f - o[k]) * inv, t1 = ((float)dims[k] - o[k]) * inv;
        tn = fmaxf(tn, fminf(t0, t1));
        tf = fminf(tf, fmaxf(t0, t1));
    }
    if (!(tn <= tf && tf > 0.0f)) return false;
    float pe[4] = {(o[0] + d[0] * tf) * 0.1f, (o[1] + d[1] * tf) * 0.1f, (o[2] + d[2] * tf) * 0.1f, 1.0f}, we[4], ce[4];
    if (world) {
        mat_vec(world, pe, we);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) we[k] = pe[k];
    }
    mat_vec(pv, we, ce);
    return ce[3] >= 0.1f;
}

__global__ void __launch_bounds__(256) k_primary(PrimaryArgs A) {
    int32_t px, py;
    if (!tile_pixel(A.tiles, px, py)) return;
    const vx_view& vw = A.view;
    size_t i = (size_t)py * vw.width + px;
    float ndc_x, ndc_y;
    f3 dir = pixel_dir(vw, px, py, ndc_x, ndc_y);
    f3 cam = mk3(vw.cam_pos[0] * 10.0f, vw.cam_pos[1] * 10.0f, vw.cam_pos[2] * 10.0f);
    ModelMips M;
#pragma unroll
    for (int k = 0; k < 3; ++k) { M.mips[k] = A.mips[k]; M.dims[k] = A.dims[k]; }
    f3 hit = mk3(0.0f, 0.0f, 0.0f), nrm = mk3(0.0f, 0.0f, 0.0f);
    uint32_t mat = 0;
    int32_t h = 0;
    // a pixel the box does not cover (no fragment: e.g. the box is behind
    // the camera) stays empty
    if (box_far_face_drawn(M.dims, cam, dir, nullptr, A.pv)) {
        f3 start = clip_to_aabb(cam, dir, M.dims);
        h = intersect_volume(M, vw.res, vw.frame, start, dir, cam, ndc_x, ndc_y, hit, nrm, mat);
    }
    if (!h) {
        hit = mk3(0.0f, 0.0f, 0.0f);
        nrm = mk3(0.0f, 0.0f, 0.0f);
        mat = 0;
    }
    A.pos[3 * i + 0] = hit.x; A.pos[3 * i + 1] = hit.y; A.pos[3 * i + 2] = hit.z;
    A.nrm[3 * i + 0] = nrm.x; A.nrm[3 * i + 1] = nrm.y; A.nrm[3 * i + 2] = nrm.z;
    A.mat[i] = (uint8_t)mat;
}

// ===========================================================================
// Multi-instance primary hit: GeometryVoxelPipeline::Use (GeometryVoxelPipeline.h:48-70)
// draws each VoxCmd's box, GeometryVoxel.vert:62-78 / GeometryVoxel.frag:127-183
// march it in model space and the D24 depth test (LESS) keeps the nearest
// fragment. Here each pixel's camera ray is taken into every instance binned
// to its 16x16 tile, in draw order, and the same depth test picks the result.
// ===========================================================================
// One thread per tile: the instances whose conservative screen rectangle
// covers the tile, in draw order; more than kTileInstCap marks the tile as
// "every instance".
__global__ void __launch_bounds__(256) k_bin_instances(BinArgs A) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ntiles) return;
    int32_t x0 = (int32_t)(t % A.tiles_x) * 16, y0 = (int32_t)(t / A.tiles_x) * 16;
    int32_t x1 = x0 + 15, y1 = y0 + 15;
    uint32_t c = 0;
    for (uint32_t k = 0; k < A.ninst; ++k) {
        const int32_t* r = A.rects + 4 * k;
        if (r[0] > x1 || r[2] < x0 || r[1] > y1 || r[3] < y0) continue;
        if (c < kTileInstCap) A.lists[(size_t)t * kTileInstCap + c] = (uint16_t)k;
        ++c;
    }
    A.counts[t] = c;
}

// Whether the draw's box rasterizes at this pixel (box_far_face_drawn).
VXD bool box_covers(const InstConst& I, f3 lc, f3 dl, const float* pv) {
    return box_far_face_drawn(I.m.dims, lc, dl, I.w, pv);
}

// D24 UNORM conversion of a fragment depth (clamped to the depth range).
VXD uint32_t depth_d24(float d) {
    d = gclamp(d, 0.0f, 1.0f);
    return (uint32_t)rintf(d * 16777215.0f);
}

__global__ void __launch_bounds__(256) k_primary_inst(PrimaryInstArgs A) {
    __shared__ uint16_t list[kTileInstCap];
    __shared__ uint32_t count;
    uint32_t b = blockIdx.x;
    uint32_t logical = (b % 8u) * A.tiles.per_xcd + b / 8u;
    if (logical >= A.tiles.ntiles) return;   // uniform over the block
    if (threadIdx.x == 0) count = A.counts[logical];
    if (threadIdx.x < kTileInstCap) list[threadIdx.x] = A.lists[(size_t)logical * kTileInstCap + threadIdx.x];
    __syncthreads();
    int32_t px, py;
    if (!tile_pixel(A.tiles, px, py)) return;
    const vx_view& vw = A.view;
    const size_t i = (size_t)py * vw.width + px;
    float ndc_x, ndc_y;
    const f3 dw = pixel_dir(vw, px, py, ndc_x, ndc_y);
    const uint32_t c = count;
    const uint32_t n = c > kTileInstCap ? A.ninst : c;
    uint32_t best_q = 16777215u;      // cleared depth 1.0
    int32_t best = -1;
    f3 bhit = mk3(0.0f, 0.0f, 0.0f), bn = mk3(0.0f, 0.0f, 0.0f);
    uint32_t bmat = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t k = c > kTileInstCap ? j : (uint32_t)list[j];
        const InstConst& I = A.inst[k];
        f3 dl = mk3((I.winv[0] * dw.x + I.winv[4] * dw.y) + I.winv[8] * dw.z,
                    (I.winv[1] * dw.x + I.winv[5] * dw.y) + I.winv[9] * dw.z,
                    (I.winv[2] * dw.x + I.winv[6] * dw.y) + I.winv[10] * dw.z);
        dl = normalize3(dl);
        const f3 lc = mk3(I.lc[0], I.lc[1], I.lc[2]);
        if (!box_covers(I, lc, dl, A.pv)) continue;
        const f3 start = clip_to_aabb(lc, dl, I.m.dims);
        f3 hit, nrm;
        uint32_t mat = 0;
        if (!intersect_volume_t<0>(I.m, vw.res, vw.frame, start, dl, lc, ndc_x, ndc_y, hit, nrm, mat)) continue;
        float h4[4] = {hit.x * 0.1f, hit.y * 0.1f, hit.z * 0.1f, 1.0f}, wh[4], cur[4];
        mat_vec(I.w, h4, wh);
        mat_vec(A.pv, wh, cur);
        const float depth = ((1.0f - 1e-7f * (float)I.rid) * (cur[3] - 0.1f)) / (4096.0f - 0.1f);
        const uint32_t q = depth_d24(depth);
        if (q < best_q) {
            best_q = q;
            best = (int32_t)k;
            bhit = hit;
            bn = nrm;
            bmat = mat;
        }
    }
    float pos[3] = {0.0f, 0.0f, 0.0f}, nw[3] = {0.0f, 0.0f, 0.0f}, mo[2] = {0.0f, 0.0f};
    uint16_t pal = 0;
    if (best >= 0) {
        const InstConst& I = A.inst[best];
        float h4[4] = {bhit.x * 0.1f, bhit.y * 0.1f, bhit.z * 0.1f, 1.0f}, wh[4], cur[4], lw[4], last[4];
        mat_vec(I.w, h4, wh);
        mat_vec(A.pv, wh, cur);
        mat_vec(I.lw, h4, lw);
        mat_vec(A.plv, lw, last);
        // out_Normal = normalize(WorldMatrix * vec4(normal, 0))
        float n4[4] = {bn.x, bn.y, bn.z, 0.0f}, wn[4];
        mat_vec(I.w, n4, wn);
        const float l = sqrtf(((wn[0] * wn[0] + wn[1] * wn[1]) + wn[2] * wn[2]) + wn[3] * wn[3]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pos[k] = wh[k] * 10.0f;
            nw[k] = wn[k] / l;
        }
        // out_Motion = -vec2(0.5, -0.5) * (cur.xy / cur.w - last.xy / last.w)
        mo[0] = -0.5f * (cur[0] / cur[3] - last[0] / last[3]);
        mo[1] = 0.5f * (cur[1] / cur[3] - last[1] / last[3]);
        pal = (uint16_t)I.palette;
    }
    A.pos[3 * i + 0] = pos[0]; A.pos[3 * i + 1] = pos[1]; A.pos[3 * i + 2] = pos[2];
    A.nrm[3 * i + 0] = nw[0]; A.nrm[3 * i + 1] = nw[1]; A.nrm[3 * i + 2] = nw[2];
    A.mat[i] = (uint8_t)bmat;
    if (A.motion) { A.motion[2 * i + 0] = mo[0]; A.motion[2 * i + 1] = mo[1]; }
    if (A.depth) A.depth[i] = (float)best_q / 16777215.0f;
    if (A.palette) A.palette[i] = pal;
}

// ===========================================================================
// Temporal accumulation of the light buffer: LightTAA.frag:37-141
// (LightTAAPipeline.h:35-51, WorldRenderer.cpp:262-267).
//
// Textures are sampled as the reference's samplers do (linear filter, clamp
// to an opaque-black border, evk.cpp:278-292): bilinear over the four nearest
// texels with 8-bit sub-texel weights; border texels read 0 in rgb and 1 in
// alpha. Colour and material come from the palette by the pixel's material
// index (the G-buffer colour / material attachments).
// ===========================================================================
namespace taa {
// One pixel's attributes in one 64-byte record, so a bilinear sample of every
// attribute loads four records (k_taa_pack writes them):
//   v0 = light rgb, light alpha (1 on sky pixels)   v1 = colour rgb, depth
//   v2 = material rgb, motion x                      v3 = normal xyz, motion y
// The material's fourth channel is 0 inside the image and 1 on the border.
struct Texel {
    float4 v0, v1, v2, v3;
};

VXD Texel border_texel() {
    Texel t;
    t.v0 = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    t.v1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    t.v2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    t.v3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return t;
}

VXD Texel fetch(const TaaArgs& A, int32_t x, int32_t y) {
    if (x < 0 || y < 0 || x >= A.width || y >= A.height) return border_texel();
    const float4* r = A.rec + 4 * ((size_t)y * A.width + x);
    Texel t;
    t.v0 = r[0]; t.v1 = r[1]; t.v2 = r[2]; t.v3 = r[3];
    return t;
}

// Sub-texel weight: the fraction rounded to 1/256; 1 moves to the next texel.
// Positions left of -2 or right of size + 1 (and NaN) sample border texels
// only; they are clamped first so that x0 and x0 + 1 stay small integers.
VXD void weight(float x, int32_t size, int32_t& x0, float& f) {
    if (!(x >= -2.0f)) x = -2.0f;
    if (x > (float)size + 1.0f) x = (float)size + 1.0f;
    float fl = floorf(x);
    f = floorf((x - fl) * 256.0f + 0.5f) * (1.0f / 256.0f);
    if (f >= 1.0f) {
        fl = fl + 1.0f;
        f = 0.0f;
    }
    x0 = (int32_t)fl;
}

VXD float bl(float a, float b, float c, float d, float fx, float fy) {
    return (a * (1.0f - fx) + b * fx) * (1.0f - fy) + (c * (1.0f - fx) + d * fx) * fy;
}

// Every attribute of texture(*, uv).
struct Sample {
    float light[4], color[3], depth, matl[4], nrm[3], mo[2];
};

VXD void sample_all(const TaaArgs& A, float u, float v, Sample& s) {
    int32_t x0, y0;
    float fx, fy;
    weight(u * (float)A.width - 0.5f, A.width, x0, fx);
    weight(v * (float)A.height - 0.5f, A.height, y0, fy);
    const Texel a = fetch(A, x0, y0), b = fetch(A, x0 + 1, y0), c = fetch(A, x0, y0 + 1), d = fetch(A, x0 + 1, y0 + 1);
    s.light[0] = bl(a.v0.x, b.v0.x, c.v0.x, d.v0.x, fx, fy);
    s.light[1] = bl(a.v0.y, b.v0.y, c.v0.y, d.v0.y, fx, fy);
    s.light[2] = bl(a.v0.z, b.v0.z, c.v0.z, d.v0.z, fx, fy);
    s.light[3] = bl(a.v0.w, b.v0.w, c.v0.w, d.v0.w, fx, fy);
    s.color[0] = bl(a.v1.x, b.v1.x, c.v1.x, d.v1.x, fx, fy);
    s.color[1] = bl(a.v1.y, b.v1.y, c.v1.y, d.v1.y, fx, fy);
    s.color[2] = bl(a.v1.z, b.v1.z, c.v1.z, d.v1.z, fx, fy);
    s.depth = bl(a.v1.w, b.v1.w, c.v1.w, d.v1.w, fx, fy);
    s.matl[0] = bl(a.v2.x, b.v2.x, c.v2.x, d.v2.x, fx, fy);
    s.matl[1] = bl(a.v2.y, b.v2.y, c.v2.y, d.v2.y, fx, fy);
    s.matl[2] = bl(a.v2.z, b.v2.z, c.v2.z, d.v2.z, fx, fy);
    const auto out = [&](int32_t x, int32_t y) { return (x < 0 || y < 0 || x >= A.width || y >= A.height) ? 1.0f : 0.0f; };
    s.matl[3] = bl(out(x0, y0), out(x0 + 1, y0), out(x0, y0 + 1), out(x0 + 1, y0 + 1), fx, fy);
    s.mo[0] = bl(a.v2.w, b.v2.w, c.v2.w, d.v2.w, fx, fy);
    s.nrm[0] = bl(a.v3.x, b.v3.x, c.v3.x, d.v3.x, fx, fy);
    s.nrm[1] = bl(a.v3.y, b.v3.y, c.v3.y, d.v3.y, fx, fy);
    s.nrm[2] = bl(a.v3.z, b.v3.z, c.v3.z, d.v3.z, fx, fy);
    s.mo[1] = bl(a.v3.w, b.v3.w, c.v3.w, d.v3.w, fx, fy);
}

// texture(LAST_LIGHT_TEXTURE, uv): 4 channels, border (0, 0, 0, 1).
VXD void sample_last(const TaaArgs& A, float u, float v, float* o) {
    int32_t x0, y0;
    float fx, fy;
    weight(u * (float)A.width - 0.5f, A.width, x0, fx);
    weight(v * (float)A.height - 0.5f, A.height, y0, fy);
    float t[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t x = x0 + (k & 1), y = y0 + (k >> 1);
        const bool in = !(x < 0 || y < 0 || x >= A.width || y >= A.height);
        const float4 q = in ? reinterpret_cast<const float4*>(A.last)[(size_t)y * A.width + x]
                            : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        t[k][0] = q.x; t[k][1] = q.y; t[k][2] = q.z; t[k][3] = q.w;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = bl(t[0][c], t[1][c], t[2][c], t[3][c], fx, fy);
}

VXD float len2(float x, float y) { return sqrtf(x * x + y * y); }
VXD float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
VXD float len4(const float* a, const float* b) {
    const float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
    return sqrtf(((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3);
}
}  // namespace taa

// One record per pixel from the G-buffer, the palette and this frame's light.
__global__ void __launch_bounds__(256) k_taa_pack(TaaArgs A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(A.width * A.height)) return;
    const uint32_t m = A.mat[i];
    const float* pe = A.palette + 6 * m;
    const float mx = A.motion ? A.motion[2 * (size_t)i] : 0.0f, my = A.motion ? A.motion[2 * (size_t)i + 1] : 0.0f;
    float4* r = A.rec + 4 * (size_t)i;
    r[0] = make_float4(A.light[3 * (size_t)i], A.light[3 * (size_t)i + 1], A.light[3 * (size_t)i + 2],
                       m == 0 ? 1.0f : 0.0f);
    r[1] = make_float4(m ? pe[0] : 0.0f, m ? pe[1] : 0.0f, m ? pe[2] : 0.0f, A.depth[i]);
    r[2] = make_float4(m ? pe[3] : 0.0f, m ? pe[4] : 0.0f, m ? pe[5] : 0.0f, mx);
    r[3] = make_float4(A.normal[3 * (size_t)i], A.normal[3 * (size_t)i + 1], A.normal[3 * (size_t)i + 2], my);
}

__global__ void __launch_bounds__(256) k_light_taa(TaaArgs A) {
    using namespace taa;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(A.width * A.height)) return;
    const int32_t px = (int32_t)(i % (uint32_t)A.width), py = (int32_t)(i / (uint32_t)A.width);
    const float W = (float)A.width, H = (float)A.height;
    const float u = ((float)px + 0.5f) / W, v = ((float)py + 0.5f) / H;
    const float irx = 1.0f / W, iry = 1.0f / H;
    Sample c;
    sample_all(A, u, v, c);
    const float ou = u + c.mo[0], ov = v + c.mo[1];
    float* o = A.out + 4 * (size_t)i;
    if (c.depth == 1.0f) {
        o[0] = c.light[0]; o[1] = c.light[1]; o[2] = c.light[2]; o[3] = c.light[3];
        return;
    }
    float last[4];
    sample_last(A, ou, ov, last);
    float lvar = last[3];
    if (gclamp(ou, 0.0f, 1.0f) != ou || gclamp(ov, 0.0f, 1.0f) != ov) {
        // out of bounds: this frame's neighbourhood only
        float count = 0.0f, acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
        for (int x = -9; x <= 9; ++x) {
#pragma unroll 1
            for (int y = -9; y <= 9; ++y) {
                const float su = gclamp(u + (float)x * irx, 0.001f, 0.999f);
                const float sv = gclamp(v + (float)y * iry, 0.001f, 0.999f);
                Sample n;
                sample_all(A, su, sv, n);
                float f = gmax((c.nrm[0] * n.nrm[0] + c.nrm[1] * n.nrm[1]) + c.nrm[2] * n.nrm[2], 0.0f);
                f *= gstep(0.8f, 1.0f - len4(c.matl, n.matl));
                f *= 1.0f - gclamp(fabsf(c.depth - n.depth) * 4096.0f, 0.0f, 1.0f);
                f *= 1.0f - gclamp(len3(c.color[0] - n.color[0], c.color[1] - n.color[1], c.color[2] - n.color[2]),
                                   0.0f, 1.0f);
                if (len2(c.mo[0] - n.mo[0], c.mo[1] - n.mo[1]) > 0.1f) f = 0.0f;
                acc[0] = acc[0] + f * n.light[0];
                acc[1] = acc[1] + f * n.light[1];
                acc[2] = acc[2] + f * n.light[2];
                count = count + f;
            }
        }
        o[0] = acc[0] / count; o[1] = acc[1] / count; o[2] = acc[2] / count; o[3] = 1.0f;
        return;
    }
    float nmin[3] = {10000.0f, 10000.0f, 10000.0f}, nmax[3] = {0.0f, 0.0f, 0.0f}, acc[3] = {0.0f, 0.0f, 0.0f};
    float diff = 1.0f, count = 0.0f;
    const float size = 12.0f;
    float radius = 1.0f;
    const float noise = rand01(pixkey(A.seed ^ 0x85ebca6bu, A.frame, i), 0u);
    const float scale = gclamp(0.1f, 0.5f, 1.0f / c.depth);
#pragma unroll 1
    for (float angle = (noise * 3.1415f) * 2.118033988749895f; radius <= size; angle = angle + 2.39f) {
        radius = radius + 1.0f;
        float sn, cs;
        sincos_vx(angle, &sn, &cs);
        const float rr = radius * (lvar + 1.0f);
        const float su = gclamp(u + ((cs * irx) * rr) * scale, 0.001f, 0.999f);
        const float sv = gclamp(v + ((sn * iry) * rr) * scale, 0.001f, 0.999f);
        Sample n;
        sample_all(A, su, sv, n);
        float f = 1.212f - radius / size;
        f *= gstep(0.8f, 1.0f - len4(c.matl, n.matl));
        f *= gmax((c.nrm[0] * n.nrm[0] + c.nrm[1] * n.nrm[1]) + c.nrm[2] * n.nrm[2], 0.0f);
        f *= 1.0f - gclamp(fabsf(c.depth - n.depth) * 4096.0f, 0.0f, 1.0f);
        f *= 1.0f - gclamp(len3(c.color[0] - n.color[0], c.color[1] - n.color[1], c.color[2] - n.color[2]) * 10000.0f,
                           0.0f, 1.0f);
        if (len2(c.mo[0] - n.mo[0], c.mo[1] - n.mo[1]) > 0.1f) f = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            nmin[k] = gmin(nmin[k], n.light[k]);
            nmax[k] = gmax(nmax[k], n.light[k]);
            acc[k] = acc[k] + f * n.light[k];
        }
        count = count + f;
        diff = diff + len3(n.light[0] - c.light[0], n.light[1] - c.light[1], n.light[2] - c.light[2]) * f;
    }
    float cur[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cur[k] = (c.light[k] + acc[k]) / (count + 1.0f);
    diff = diff / radius;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        last[k] = gclamp(last[k], nmin[k], nmax[k]);
        d[k] = fabsf(cur[k] - last[k]);
    }
    float var = (d[0] * 0.2125f + d[1] * 0.7154f) + d[2] * 0.0721f;
    var = gclamp(var * 5.5f, 0.0f, 1.0f);
    lvar = lvar + var;
    lvar = lvar - diff * 0.08f;
    lvar = lvar + len2(c.mo[0], c.mo[1]) * 20.0f;
    const float t = gclamp(1.0f - lvar, 0.3f, 0.9f);
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = cur[k] * (1.0f - t) + last[k] * t;
    o[3] = gclamp(lvar * 0.7f, 0.0f, 1.0f);
}

// ===========================================================================
// Reflection pass: LightReflection.frag:60-139 (LightReflectionPipeline::Use,
// LightReflectionPipeline.h:34), one thread per pixel, world space. The
// reflection ray's march, its reprojection into the light buffer with the
// +-0.001 linear-depth window against the depth attachment, the sky on a
// miss, and (ambient * F * (1 - roughness), F) with F the Schlick-roughness
// Fresnel of albedo 1, times 5. Light and depth are sampled with the TAA
// pass's sampler (bilinear, 8-bit sub-texel weights, opaque-black border).
// ===========================================================================
namespace refl {
// getSkyColor (LightAmbient.frag:128-131).
VXD f3 sky_procedural(f3 e) {
    const float ey = (gmax(e.y, 0.0f) * 0.8f + 0.2f) * 0.8f;
    const float a = 1.0f - ey;
    return mk3((a * a) * 1.1f, a * 1.1f, (0.6f + a * 0.4f) * 1.1f);
}

// texture(samplerCube, d): Vulkan face selection (major axis x, then y, then
// z on ties), bilinear within the face with the texel position clamped to
// it. Faces +X, -X, +Y, -Y, +Z, -Z, size x size RGBA.
VXD f3 cube_sample(const float* faces, int32_t size, f3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int32_t face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; }
        else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; }
        else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; }
        else { face = 5; sc = -d.x; tc = -d.y; }
    }
    const float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    const float fs = (float)size;
    const float x = gclamp(s * fs - 0.5f, 0.0f, fs - 1.0f), y = gclamp(t * fs - 0.5f, 0.0f, fs - 1.0f);
    int32_t x0, y0;
    float fx, fy;
    taa::weight(x, size, x0, fx);
    taa::weight(y, size, y0, fy);
    const int32_t x1 = x0 + 1 < size ? x0 + 1 : size - 1, y1 = y0 + 1 < size ? y0 + 1 : size - 1;
    const float4* F = reinterpret_cast<const float4*>(faces) + (size_t)face * size * size;
    const float4 a = F[(size_t)y0 * size + x0], b = F[(size_t)y0 * size + x1];
    const float4 c = F[(size_t)y1 * size + x0], e = F[(size_t)y1 * size + x1];
    return mk3(taa::bl(a.x, b.x, c.x, e.x, fx, fy), taa::bl(a.y, b.y, c.y, e.y, fx, fy),
               taa::bl(a.z, b.z, c.z, e.z, fx, fy));
}

// Bilinear sample of an n-channel float image (border: 0, alpha 1).
template <int N>
VXD void sample(const float* img, int32_t w, int32_t h, float u, float v, float* out) {
    int32_t x0, y0;
    float fx, fy;
    taa::weight(u * (float)w - 0.5f, w, x0, fx);
    taa::weight(v * (float)h - 0.5f, h, y0, fy);
    float t[4][N];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t x = x0 + (k & 1), y = y0 + (k >> 1);
        const bool in = !(x < 0 || y < 0 || x >= w || y >= h);
#pragma unroll
        for (int c = 0; c < N; ++c) t[k][c] = in ? img[((size_t)y * w + x) * N + c] : (c == 3 ? 1.0f : 0.0f);
    }
#pragma unroll
    for (int c = 0; c < N; ++c) out[c] = taa::bl(t[0][c], t[1][c], t[2][c], t[3][c], fx, fy);
}
}  // namespace refl

__global__ void __launch_bounds__(256) k_reflection(OccView V, ReflectionArgs A) {
    int32_t px, py;
    if (!tile_pixel(A.tiles, px, py)) return;
    const vx_view& vw = A.view;
    const int32_t W = vw.width, H = vw.height;
    const size_t i = (size_t)py * W + px;
    const uint32_t m = A.mat[i];
    const f3 pp = load3(A.pos + 3 * i), n = load3(A.nrm + 3 * i);
    const float depth = A.depth[i];
    const float rough = A.palette[6 * m + 3], metal = A.palette[6 * m + 4];
    const float f0 = 0.04f * (1.0f - metal) + 1.0f * metal;
    const f3 cam = mk3(vw.cam_pos[0] * 10.0f, vw.cam_pos[1] * 10.0f, vw.cam_pos[2] * 10.0f);
    const f3 Vd = normalize3(mk3(cam.x - pp.x, cam.y - pp.y, cam.z - pp.z));
    const float ndv = gmax(dot3(n, Vd), 0.0f);
    const float x = gmax(1.0f - ndv, 0.0f);
    const float s5 = (((x * x) * x) * x) * x;
    const float F = (f0 + (gmax(1.0f - rough, f0) - f0) * s5) * 5.0f + 0.0f;
    f3 amb = mk3(0.0f, 0.0f, 0.0f);
    float tt = __uint_as_float(0x7fc00000u);
    if (depth < 0.999f) {
        // the ray (LightReflection.frag:79-98)
        const uint32_t pk = pixkey(A.params.seed, A.params.frame, (uint32_t)i);
        float nz[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) nz[k] = rand01(pk, (uint32_t)k);
        const f3 I = mk3(-Vd.x, -Vd.y, -Vd.z);
        const float dni = dot3(n, I);
        f3 wd = normalize3(mk3(I.x - (2.0f * dni) * n.x, I.y - (2.0f * dni) * n.y, I.z - (2.0f * dni) * n.z));
        f3 rv = cosine_hemisphere(nz[0], nz[1]);
        rv.z = rv.z * gsign(nz[2] - 0.5f);
        const float a = rough * 0.1f;
        wd = mk3(wd.x * (1.0f - a) + rv.x * a, wd.y * (1.0f - a) + rv.y * a, wd.z * (1.0f - a) + rv.z * a);
        const f3 wcp = mk3(pp.x + n.x * nz[3], pp.y + n.y * nz[3], pp.z + n.z * nz[3]);
        const float sc = 1.0f + nz[3] * 0.5f;
        wd = mk3(wd.x * sc, wd.y * sc, wd.z * sc);
        uint32_t nf = 0;
        int32_t hit;
        const float t = march_dist<kSparsePrefetch>(V, A.params.march, mk3(wcp.x + n.x, wcp.y + n.y, wcp.z + n.z), wd, A.params.maxt,
                                   nf, hit);
        tt = t;
        // reprojection (LightReflection.frag:115-133)
        float h4[4] = {(wcp.x + wd.x * t) * 0.1f, (wcp.y + wd.y * t) * 0.1f, (wcp.z + wd.z * t) * 0.1f, 1.0f};
        float pv[4], pc[4];
        mat_vec(vw.view, h4, pv);
        pv[3] = 1.0f;
        mat_vec(vw.proj, pv, pc);
        const float lin = (pc[3] - 0.1f) / (4096.0f - 0.1f);
        const float u = ((pc[0] / pc[3]) * 1.0f) * 0.5f + 0.5f;
        const float v = ((pc[1] / pc[3]) * -1.0f) * 0.5f + 0.5f;
        float ds;
        refl::sample<1>(A.depth, W, H, u, v, &ds);
        if (t == A.params.maxt) {
            amb = A.sky ? refl::cube_sample(A.sky, A.sky_size, wd) : refl::sky_procedural(wd);
        } else if (lin > ds - 0.001f && lin < ds + 0.001f) {
            float l4[4];
            refl::sample<4>(A.light, W, H, u, v, l4);
            amb = mk3(l4[0], l4[1], l4[2]);
        }
    }
    float4 o;
    o.x = (amb.x * F) * (1.0f - rough);
    o.y = (amb.y * F) * (1.0f - rough);
    o.z = (amb.z * F) * (1.0f - rough);
    o.w = F;
    reinterpret_cast<float4*>(A.out)[i] = o;
    if (A.t) A.t[i] = tt;
}

// Screen-space occlusion, calculateOcclusion (LightAmbient.frag:89-109) with
// screenspaceOcclusion (:59-79), one thread per pixel: the view-space
// position rebuilt from the depth attachment and FarVec, four samples within
// OCCLUSION_SIZE = 0.2 world units in the normal's cosine hemisphere (view
// space), each occluded where the depth attachment at its projection lies
// within OCCLUSION_TICKNESS / FAR in front of it, weighted by
// smoothstep(maxDepth, minDepth, depth) (edges reversed, evaluated as
// written) times its distance; 1 - 3 * mean, clamped. The noise is the
// pixel's counter hash under its own key. 1 for sky pixels.
__global__ void __launch_bounds__(256) k_ssao(SsaoArgs A) {
    const vx_view& vw = A.view;
    const int32_t W = vw.width, H = vw.height;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(W * H)) return;
    const float depth = A.depth[i];
    if (A.mat[i] == 0 || !(depth < 0.999f)) {
        A.out[i] = 1.0f;
        return;
    }
    const int32_t px = (int32_t)(i % (uint32_t)W), py = (int32_t)(i / (uint32_t)W);
    float n4[4] = {A.nrm[3 * (size_t)i], A.nrm[3 * (size_t)i + 1], A.nrm[3 * (size_t)i + 2], 0.0f}, N4[4];
    mat_vec(vw.view, n4, N4);
    const f3 N = mk3(N4[0], N4[1], N4[2]);
    const f3 T = normalize3(fabsf(N.z) > 0.5f ? mk3(0.0f, -N.z, N.y) : mk3(-N.y, N.x, 0.0f));
    const f3 B = normalize3(mk3(N.y * T.z - N.z * T.y, N.z * T.x - N.x * T.z, N.x * T.y - N.y * T.x));
    const float u = ((float)px + 0.5f) / vw.res[0], v = ((float)py + 0.5f) / vw.res[1];
    float c4[4] = {u * 2.0f - 1.0f, 1.0f - v * 2.0f, 1.0f, 1.0f}, f4[4];
    mat_vec(vw.inv_proj, c4, f4);
    const float fx = f4[0] / f4[3], fy = f4[1] / f4[3], fz = f4[2] / f4[3];
    const float sc = depth * (1.0f + 1.0f / 4096.0f) + 0.1f / 4096.0f;
    const f3 pos = mk3(fx * sc, fy * sc, fz * sc);
    const float size = 0.2f * (1.0f + depth * 0.0f);
    const uint32_t pk = pixkey(A.seed ^ 0x632be5abu, vw.frame, i);
    float occ = 0.0f;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const float nx = rand01(pk, (uint32_t)(4 * s)), ny = rand01(pk, (uint32_t)(4 * s + 1));
        const float dist = rand01(pk, (uint32_t)(4 * s + 2));
        const f3 rv = cosine_hemisphere(nx, ny);
        const f3 dn = normalize3(mk3((T.x * rv.x + B.x * rv.y) + N.x * rv.z, (T.y * rv.x + B.y * rv.y) + N.y * rv.z,
                                     (T.z * rv.x + B.z * rv.y) + N.z * rv.z));
        float mid[4] = {pos.x + (dn.x * size) * dist, pos.y + (dn.y * size) * dist, pos.z + (dn.z * size) * dist, 1.0f};
        float mp[4];
        mat_vec(vw.proj, mid, mp);
        const float sd = (mp[3] - 0.1f) / (4096.0f - 0.1f);
        const float su = ((mp[0] / mp[3]) * 1.0f) * 0.5f + 0.5f;
        const float sv = ((mp[1] / mp[3]) * -1.0f) * 0.5f + 0.5f;
        float mn;
        refl::sample<1>(A.depth, W, H, su, sv, &mn);
        const float mx = mn + 0.2f / 4096.0f;
        float o = 0.0f;
        if (gclamp(su, 0.0f, 1.0f) == su && gclamp(sv, 0.0f, 1.0f) == sv && mn < sd && sd < mx) {
            const float tt = gclamp((sd - mx) / (mn - mx), 0.0f, 1.0f);
            o = ((tt * tt) * (3.0f - 2.0f * tt)) * dist;
        }
        occ = occ + o;
    }
    occ = occ / 4.0f;
    occ = occ * 3.0f;
    A.out[i] = gclamp(1.0f - occ, 0.0f, 1.0f);
}

// D24 depth of a G-buffer's hits: the geometry pass's fragment depth// D24 depth of a G-buffer's hits: the geometry pass's fragment depth
// (1 - 1e-7 rid)(w - NEAR)/(FAR - NEAR) of the world hit, quantized to 24
// bits and read back as q / (2^24 - 1); 1 where nothing was hit.
__global__ void __launch_bounds__(256) k_gbuf_depth(DepthArgs A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.npx) return;
    uint32_t q = 16777215u;
    if (A.mat[i]) {
        float h4[4] = {A.pos[3 * (size_t)i] * 0.1f, A.pos[3 * (size_t)i + 1] * 0.1f, A.pos[3 * (size_t)i + 2] * 0.1f,
                       1.0f}, c[4];
        mat_vec(A.pv, h4, c);
        q = depth_d24((A.rid_scale * (c[3] - 0.1f)) / (4096.0f - 0.1f));
    }
    A.depth[i] = (float)q / 16777215.0f;
}

// ===========================================================================
// Shading (LightAmbient.frag:178-214, PBRDirectLight PBR.frag:48-68,
// fresnelSchlickRoughness PBR.frag:8-10) and its adjoint.
// ===========================================================================
struct ShadePix {
    f3 N, V;
    float sh, A, s5, ndl, nl;
};

VXD void sky_irradiance(float irr[3]) {
    // getSkyColor(vec3(1,0,0)), LightAmbient.frag:128-131.
    float ey = (gmax(0.0f, 0.0f) * 0.8f + 0.2f) * 0.8f;
    irr[0] = ((1.0f - ey) * (1.0f - ey)) * 1.1f;
    irr[1] = (1.0f - ey) * 1.1f;
    irr[2] = (0.6f + (1.0f - ey) * 0.4f) * 1.1f;
}

VXD void shade_setup(const ShadeArgs& A, size_t i, f3 L, ShadePix& S) {
    f3 pp = load3(A.pos + 3 * i);
    S.N = load3(A.nrm + 3 * i);
    f3 cam = mk3(A.view.cam_pos[0] * 10.0f, A.view.cam_pos[1] * 10.0f, A.view.cam_pos[2] * 10.0f);
    S.V = normalize3(mk3(cam.x - pp.x, cam.y - pp.y, cam.z - pp.z));
    S.sh = (float)A.shadow[i];
    S.A = (A.ao[i] * 0.05f) * (A.ssao ? A.ssao[i] : 1.0f);
    float ndv = gmax(dot3(S.N, S.V), 0.0f);
    float x = gmax(1.0f - ndv, 0.0f);
    S.s5 = (((x * x) * x) * x) * x;
    S.nl = dot3(S.N, L);
    S.ndl = gmax(S.nl, 0.0f);
}

__global__ void k_shade_fwd(ShadeArgs A) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.npx) return;
    uint32_t m = A.mat[i];
    if (m == 0) {
        A.rgb[3 * i] = 0.0f; A.rgb[3 * i + 1] = 0.0f; A.rgb[3 * i + 2] = 0.0f;
        return;
    }
    float irr[3];
    sky_irradiance(irr);
    f3 L = normalize3(mk3(A.light[0], A.light[1], A.light[2]));
    ShadePix S;
    shade_setup(A, i, L, S);
    const float* e = A.palette + 6 * m;
    float rough = e[3], metal = e[4], emit = e[5];
    float f0d = 0.04f * (1.0f - metal) + 1.0f * metal;
    float F = f0d + (1.0f - f0d) * S.s5;
    float kD = (1.0f - F) * (1.0f - metal);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float rad = A.light[3 + c] * S.sh;
        float lo = (((kD * e[c]) / 3.14159265359f) * rad) * S.ndl;
        float f0a = 0.04f * (1.0f - metal) + e[c] * metal;
        float mx = gmax(1.0f - rough, f0a);
        float fr = f0a + (mx - f0a) * S.s5;
        float kda = 1.0f - fr;
        float amb = (irr[c] * e[c]) * (emit * 10.0f + (kda * S.A));
        A.rgb[3 * i + c] = amb + lo;
    }
}

// Fixed-point accumulation: 2^30 units per 1.0 in int64, so the sum is
// independent of the order of the atomics (deterministic gradients).
VXD long long to_fixed(float v) { return __float2ll_rn(v * 1073741824.0f); }

// Adjoint of k_shade_fwd with visibility fixed. Per-pixel partials are summed
// into an LDS histogram of 256x6 (+6 light) int64 fixed-point cells, then
// flushed with int64 atomics into the global accumulator acc[256*6+6].
// Partials of one pixel: palette entry m into hist (LDS, fixed point), light
// into gl (per lane, summed per wave by the caller).
VXD void shade_bwd_pixel(const ShadeArgs& A, size_t i, f3 L, float ll, const float irr[3], long long* hist,
                         float gl[6]) {
    uint32_t m = A.mat[i];
    if (m == 0) return;
    ShadePix S;
    shade_setup(A, i, L, S);
    const float* e = A.palette + 6 * m;
    float rough = e[3], metal = e[4];
    float f0d = 0.04f * (1.0f - metal) + 1.0f * metal;
    float F = f0d + (1.0f - f0d) * S.s5;
    float kD = (1.0f - F) * (1.0f - metal);
    float dF_dme = 0.96f * (1.0f - S.s5);
    float dkD_dme = -dF_dme * (1.0f - metal) - (1.0f - F);
    float gp[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float gc = A.dl_drgb[3 * i + c];
        float a = e[c];
        float rad = A.light[3 + c] * S.sh;
        float f0a = 0.04f * (1.0f - metal) + a * metal;
        bool pick = (1.0f - rough) < f0a;
        float mx = pick ? f0a : (1.0f - rough);
        float fr = f0a + (mx - f0a) * S.s5;
        float E = e[5] * 10.0f + (1.0f - fr) * S.A;
        float dlo_da = ((kD / 3.14159265359f) * rad) * S.ndl;
        float dmx_da = pick ? metal : 0.0f;
        float dfr_da = metal + (dmx_da - metal) * S.s5;
        gp[c] = gc * (dlo_da + (irr[c] * E - ((irr[c] * a) * S.A) * dfr_da));
        float dfr_dro = pick ? 0.0f : -S.s5;
        gp[3] += gc * (-((irr[c] * a) * S.A) * dfr_dro);
        float dlo_dme = ((dkD_dme * a / 3.14159265359f) * rad) * S.ndl;
        float df0a_dme = a - 0.04f;
        float dmx_dme = pick ? df0a_dme : 0.0f;
        float dfr_dme = df0a_dme + (dmx_dme - df0a_dme) * S.s5;
        gp[4] += gc * (dlo_dme - ((irr[c] * a) * S.A) * dfr_dme);
        gp[5] += gc * ((irr[c] * a) * 10.0f);
        gl[3 + c] += gc * ((((kD * a) / 3.14159265359f) * S.sh) * S.ndl);
        if (!(S.nl < 0.0f)) {
            float k = gc * (((kD * a) / 3.14159265359f) * rad) / ll;
            gl[0] += k * (S.N.x - L.x * S.nl);
            gl[1] += k * (S.N.y - L.y * S.nl);
            gl[2] += k * (S.N.z - L.z * S.nl);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) atomicAdd((unsigned long long*)&hist[6 * m + k], (unsigned long long)to_fixed(gp[k]));
}

// Light partials: per-lane fixed point, wave sum in int64, one LDS atomic per
// wave; then the block's histogram into the global accumulator.
VXD void shade_bwd_flush(long long* hist, const float gl[6], long long* acc) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        long long v = to_fixed(gl[k]);
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)&hist[256 * 6 + k], (unsigned long long)v);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 256 * 6 + 6; j += blockDim.x)
        if (hist[j] != 0) atomicAdd((unsigned long long*)&acc[j], (unsigned long long)hist[j]);
}

VXD void shade_light(const float light[6], f3& L, float& ll) {
    f3 ld = mk3(light[0], light[1], light[2]);
    ll = sqrtf(dot3(ld, ld));
    L = mk3(ld.x / ll, ld.y / ll, ld.z / ll);
}

__global__ void __launch_bounds__(256) k_shade_bwd(ShadeArgs A, long long* acc) {
    __shared__ long long hist[256 * 6 + 6];
    for (int j = threadIdx.x; j < 256 * 6 + 6; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    float irr[3];
    sky_irradiance(irr);
    f3 L;
    float ll;
    shade_light(A.light, L, ll);
    float gl[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.npx; i += (size_t)gridDim.x * blockDim.x)
        shade_bwd_pixel(A, i, L, ll, irr, hist, gl);
    shade_bwd_flush(hist, gl, acc);
}

// Batched adjoint: block b belongs to the item with block0[item] <= b; the
// item's blocks stride over its pixels.
__global__ void __launch_bounds__(256) k_shade_bwd_batch(const ShadeArgs* items, const uint32_t* block0, uint32_t n,
                                                         long long* acc) {
    __shared__ long long hist[256 * 6 + 6];
    for (int j = threadIdx.x; j < 256 * 6 + 6; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (block0[mid] <= blockIdx.x) lo = mid;
        else hi = mid - 1u;
    }
    const ShadeArgs& A = items[lo];
    const uint32_t nb = block0[lo + 1] - block0[lo];
    float irr[3];
    sky_irradiance(irr);
    f3 L;
    float ll;
    shade_light(A.light, L, ll);
    float gl[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = (size_t)(blockIdx.x - block0[lo]) * blockDim.x + threadIdx.x; i < A.npx;
         i += (size_t)nb * blockDim.x)
        shade_bwd_pixel(A, i, L, ll, irr, hist, gl);
    shade_bwd_flush(hist, gl, acc);
}

__global__ void k_shade_bwd_finish(const long long* acc, float* dpal, float* dlight) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 256 * 6) dpal[j] = (float)((double)acc[j] * (1.0 / 1073741824.0));
    else if (j < 256 * 6 + 6) dlight[j - 256 * 6] = (float)((double)acc[j] * (1.0 / 1073741824.0));
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = 0;
}

__global__ void k_zero_i64(long long* p, int n) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = 0;
}

}  // namespace vx
