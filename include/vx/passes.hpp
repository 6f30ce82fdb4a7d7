// passes.hpp -- C++17 host wrappers over the C ABI that keep the reference's
// system and render-pass shapes, so engine code written against VoxelEngine's
// ShadowVoxSystem (Sources/World/Systems/ShadowVoxSystem.h:9-37) and
// Light*Pipeline::Use (Sources/Graphics/Pipelines/LightAmbientPipeline.h:35,
// LightReflectionPipeline.h:34, LightPointPipeline.h:58-100,
// LightSpotPipeline.h:60-104) drops onto the gfx950 path.
//
// Header-only; link with libvx_amd.so. Errors throw vx::Error, as the
// reference's CHECK throws (Sources/Core/Core.h:83-86).
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "vx_api.h"

namespace vx {

class Error : public std::runtime_error {
public:
    Error(vx_status code, const char* msg) : std::runtime_error(msg ? msg : "vx error"), code(code) {}
    vx_status code;
};

inline void check(vx_status s) {
    if (s != VX_OK) throw Error(s, vx_last_error());
}

// Graphics module start-up for this path: one HIP device.
class Context {
public:
    explicit Context(int device = 0) { check(vx_ctx_create(device, &h_)); }
    ~Context() { vx_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    vx_ctx* get() const { return h_; }

private:
    vx_ctx* h_ = nullptr;
};

// VoxAsset (Sources/Asset/VoxAsset.h:8-62): material volume with 3 mips.
class VoxAsset {
public:
    VoxAsset(Context& ctx, int sx, int sy, int sz) { check(vx_model_create(ctx.get(), sx, sy, sz, &h_)); }
    // A .v file (VoxAsset::Serialize + OnLoad, VoxAsset.h:36-50).
    VoxAsset(Context& ctx, const std::string& path, void* stream = nullptr) {
        check(vx_model_load(ctx.get(), path.c_str(), &h_, stream));
    }
    ~VoxAsset() { vx_model_destroy(h_); }
    VoxAsset(const VoxAsset&) = delete;
    VoxAsset& operator=(const VoxAsset&) = delete;
    // VoxAsset::Upload (VoxAsset.cpp:3-64): level 0 copy, then the mip build.
    void Upload(const std::vector<uint8_t>& data, void* stream = nullptr) {
        check(vx_model_upload(h_, data.data(), data.size(), stream));
        check(vx_model_build_mips(h_, stream));
    }
    void Generate(const vx_scene_desc& d, void* stream = nullptr) {
        check(vx_model_generate(h_, &d, stream));
        check(vx_model_build_mips(h_, stream));
    }
    std::array<int32_t, 3> Extent() const {
        std::array<int32_t, 3> d{};
        check(vx_model_dims(h_, d.data()));
        return d;
    }
    const uint8_t* Data() const {
        uint8_t* p = nullptr;
        check(vx_model_mip_ptr(h_, 0, &p));
        return p;
    }
    vx_model* GetImage() const { return h_; }

private:
    vx_model* h_ = nullptr;
};

// PalleteAsset (Sources/Asset/PalleteAsset.h:27-70) as the 256 x 6 fp32
// palette the passes read (host memory; copy it to the device to use it).
inline std::vector<float> LoadPallete(const std::string& path) {
    std::vector<float> pal(256 * 6);
    check(vx_palette_load(path.c_str(), pal.data()));
    return pal;
}

// PrefabAsset (Sources/Asset/PrefabAsset.h:14-40): the entities
// PrefabAsset::Spawn (PrefabAsset.cpp:30-142) creates, with their world
// matrices, the files their GUIDs name, the rebuild order and the light draws.
class PrefabAsset {
public:
    PrefabAsset(const std::string& modsDir, const std::string& path) {
        check(vx_prefab_load(modsDir.c_str(), path.c_str(), &h_));
        uint32_t ne = 0, nf = 0;
        check(vx_prefab_counts(h_, &ne, &nf));
        entities_.resize(ne);
        check(vx_prefab_entities(h_, entities_.data(), ne));
        std::vector<char> buf(4096);
        for (uint32_t i = 0; i < nf; ++i) {
            check(vx_prefab_file(h_, i, buf.data(), buf.size()));
            files_.emplace_back(buf.data());
        }
    }
    ~PrefabAsset() { vx_prefab_destroy(h_); }
    PrefabAsset(const PrefabAsset&) = delete;
    PrefabAsset& operator=(const PrefabAsset&) = delete;

    const std::vector<vx_prefab_entity>& Entities() const { return entities_; }
    const std::vector<std::string>& Files() const { return files_; }
    std::vector<uint32_t> RebuildOrder() const {
        std::vector<uint32_t> idx(entities_.size());
        uint32_t n = 0;
        check(vx_prefab_rebuild_order(h_, idx.data(), (uint32_t)idx.size(), &n));
        idx.resize(n);
        return idx;
    }
    // The geometry pass's draws (WorldRenderer.cpp:210-220); models[i]: loaded
    // model of file i or nullptr; prev: per-entity PreviousWorldMatrix or nullptr.
    std::vector<vx_vox_cmd> VoxCmds(const std::vector<const vx_model*>& models, const float* prev = nullptr) const {
        uint32_t n = 0;
        const vx_status s = vx_prefab_vox_cmds(h_, models.data(), (uint32_t)models.size(), prev, nullptr, 0, &n);
        if (s != VX_OK && s != VX_ERR_RANGE) check(s);
        std::vector<vx_vox_cmd> out(n);
        if (n) check(vx_prefab_vox_cmds(h_, models.data(), (uint32_t)models.size(), prev, out.data(), n, &n));
        return out;
    }
    // Point lights, then spot lights, in the reference's draw order.
    std::vector<vx_light> Lights() const {
        uint32_t n = 0;
        const vx_status s = vx_prefab_lights(h_, nullptr, 0, &n);  // count query: VX_ERR_RANGE when n > 0
        if (s != VX_OK && s != VX_ERR_RANGE) check(s);
        std::vector<vx_light> out(n);
        if (n) check(vx_prefab_lights(h_, out.data(), n, &n));
        return out;
    }

private:
    vx_prefab* h_ = nullptr;
    std::vector<vx_prefab_entity> entities_;
    std::vector<std::string> files_;
};

using Entity = uint32_t;

// Transform + VoxRenderer of one entity (Sources/World/Components.h:30-35,60-67).
struct VoxEntity {
    const VoxAsset* Vox = nullptr;
    float Pivot[3] = {0.0f, 0.0f, 0.0f};
    float PreviousWorldMatrix[16];
    float WorldMatrix[16];
    bool Changed = true;
};

// ShadowVoxSystem (ShadowVoxSystem.h:9-37): owns the occupancy volume and
// rebuilds it from the entities changed since the last update.
class ShadowVoxSystem {
public:
    // Reference volume: 524 x 188 x 524 texels (ShadowVoxSystem.cpp:56-58).
    explicit ShadowVoxSystem(Context& ctx, int sx = 524, int sy = 188, int sz = 524, int levels = 2) {
        check(vx_volume_create(ctx.get(), sx, sy, sz, levels, &vol_));
        check(vx_volume_clear(vol_, nullptr));
    }
    ~ShadowVoxSystem() { vx_volume_destroy(vol_); }
    ShadowVoxSystem(const ShadowVoxSystem&) = delete;
    ShadowVoxSystem& operator=(const ShadowVoxSystem&) = delete;

    void OnCreate() {}
    void OnDestroy() {}

    Entity CreateEntity(const VoxAsset* vox, const float world[16], const float pivot[3] = nullptr) {
        VoxEntity e;
        e.Vox = vox;
        if (pivot) std::memcpy(e.Pivot, pivot, sizeof(e.Pivot));
        static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        std::memcpy(e.PreviousWorldMatrix, I, sizeof(I));
        std::memcpy(e.WorldMatrix, world, sizeof(e.WorldMatrix));
        e.Changed = true;
        Entity id = next_++;
        entities_[id] = e;
        return id;
    }

    // TransformSystem: PreviousWorldMatrix <- WorldMatrix, new WorldMatrix, Changed tag
    // (TransformSystem.cpp:35-71).
    void SetWorldMatrix(Entity id, const float world[16]) {
        VoxEntity& e = entities_.at(id);
        if (!e.Changed) std::memcpy(e.PreviousWorldMatrix, e.WorldMatrix, sizeof(e.WorldMatrix));
        std::memcpy(e.WorldMatrix, world, sizeof(e.WorldMatrix));
        e.Changed = true;
    }

    // OnVoxDestroyed (ShadowVoxSystem.cpp:7-53): erase with the WorldMatrix and no
    // pivot; the erase runs before the next update's entities.
    void OnVoxDestroyed(Entity id) {
        auto it = entities_.find(id);
        if (it == entities_.end()) return;
        const VoxEntity& e = it->second;
        if (e.Vox) pending_.push_back(instance(e, VX_INST_DESTROY));
        entities_.erase(it);
    }

    // The VoxRenderer entities of a spawned prefab in creation order, so the
    // next OnUpdate visits them as the reference does. models maps a file
    // index of the prefab to its loaded asset; entities whose Vox did not
    // resolve are skipped (ShadowVoxSystem.cpp:120).
    std::vector<Entity> Spawn(const PrefabAsset& pf, const std::map<int32_t, const VoxAsset*>& models) {
        std::vector<Entity> ids;
        const auto& ents = pf.Entities();
        for (uint32_t i = 0; i < ents.size(); ++i) {
            const vx_prefab_entity& e = ents[i];
            if (!(e.flags & VX_ENT_VOX) || !(e.flags & VX_ENT_TRANSFORM) || e.model < 0) continue;
            auto it = models.find(e.model);
            if (it == models.end()) throw Error(VX_ERR_INVALID_ARGUMENT, "Spawn: a model of the prefab is not loaded");
            ids.push_back(CreateEntity(it->second, e.world, e.pivot));
        }
        return ids;
    }

    // OnUpdate (ShadowVoxSystem.cpp:116-201): every Changed entity, newest
    // first (EnTT 3.6 views walk their pools from the last element), then the
    // coarse levels over the dirty regions.
    void OnUpdate(float /*dt*/, void* stream = nullptr) {
        std::vector<vx_instance> inst = std::move(pending_);
        pending_.clear();
        for (auto it = entities_.rbegin(); it != entities_.rend(); ++it) {
            VoxEntity& e = it->second;
            if (!e.Changed) continue;
            if (e.Vox) inst.push_back(instance(e, VX_INST_UPDATE));
            e.Changed = false;
            std::memcpy(e.PreviousWorldMatrix, e.WorldMatrix, sizeof(e.WorldMatrix));
        }
        if (inst.empty()) return;
        regions_.assign(inst.size(), vx_region{});
        uint32_t nd = 0;
        check(vx_volume_rebuild(vol_, inst.data(), (uint32_t)inst.size(), 16, regions_.data(), &nd, stream));
        regions_.resize(nd);
        if (nd) check(vx_volume_build_levels(vol_, regions_.data(), nd, stream));
    }

    const std::vector<vx_region>& LastDirtyRegions() const { return regions_; }
    vx_volume* GetVolumeImage() const { return vol_; }

private:
    static vx_instance instance(const VoxEntity& e, int32_t flags) {
        vx_instance in{};
        std::array<int32_t, 3> d = e.Vox->Extent();
        in.model = e.Vox->Data();
        in.sx = d[0]; in.sy = d[1]; in.sz = d[2];
        std::memcpy(in.prev, e.PreviousWorldMatrix, sizeof(in.prev));
        std::memcpy(in.cur, e.WorldMatrix, sizeof(in.cur));
        std::memcpy(in.pivot, e.Pivot, sizeof(in.pivot));
        in.flags = flags;
        return in;
    }

    vx_volume* vol_ = nullptr;
    std::map<Entity, VoxEntity> entities_;
    std::vector<vx_instance> pending_;
    std::vector<vx_region> regions_;
    Entity next_ = 1;
};

// GeometryVoxelPipeline (GeometryVoxelPipeline.h:15-76) for the world model:
// fills the fp32 G-buffer.
class GeometryVoxelPipeline {
public:
    static GeometryVoxelPipeline& Get() {
        static GeometryVoxelPipeline p;
        return p;
    }
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const VoxAsset& world, const vx_gbuf& geometryFB) {
        check(vx_primary(ctx.get(), world.GetImage(), &viewBuffer, &geometryFB, cmd));
    }
    // Use(CmdBuffer&, Buffer& viewBuffer, glm::vec2 res, Image& blueNoise, std::vector<Cmd>& cmds)
    // (GeometryVoxelPipeline.h:48-70): every voxel draw, depth-tested in order.
    using Cmd = vx_vox_cmd;
    static constexpr int MAX_INSTANCES = 4096;
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const std::vector<Cmd>& cmds, const vx_gbuf& geometryFB,
             const vx_gbuf_ext* extra = nullptr) {
        check(vx_primary_instances(ctx.get(), cmds.data(), (uint32_t)cmds.size(), &viewBuffer, &geometryFB, extra, cmd));
    }
};

// LightAmbientPipeline (LightAmbientPipeline.h:12-58): sun shadow and voxel AO
// rays of the ambient pass; the shading itself is vx_shade_forward.
class LightAmbientPipeline {
public:
    static LightAmbientPipeline& Get() {
        static LightAmbientPipeline p;
        return p;
    }
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_gbuf& geometryFB, vx_volume* shadowVox,
             const float* palette, const vx_trace_params& params, const vx_trace_out& lightOut) {
        vx_trace_params p = params;
        p.n_spec = 0;
        check(vx_trace(ctx.get(), shadowVox, &viewBuffer, &geometryFB, palette, &p, &lightOut, cmd));
    }
    // calculateOcclusion (LightAmbient.frag:89-109) from the depth attachment:
    // the screen-space factor of the ambient term, for vx_shade_forward.
    void UseOcclusion(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_gbuf& geometryFB,
                      const float* geometryDepth, uint32_t seed, float* occlusionOut) {
        check(vx_ssao(ctx.get(), &viewBuffer, &geometryFB, geometryDepth, seed, occlusionOut, cmd));
    }
};

// LightReflectionPipeline (LightReflectionPipeline.h:12-57):
// Use(cmd, viewBuffer, lightDiffuse, geometryFB, shadowVox, skybox, blueNoise)
// draws the reflection buffer (LightReflection.frag:60-139): per pixel, the
// reflection ray marched through the ShadowVox volume, reprojected into
// lightDiffuse (the TAA light buffer, 4 floats per pixel) when it lands within
// the depth window, the skybox on a miss, out = (ambient * F * (1 -
// roughness), F). The geometry framebuffer is the G-buffer plus its depth
// attachment; the blue-noise texture is the seeded counter hash of params.
class LightReflectionPipeline {
public:
    static LightReflectionPipeline& Get() {
        static LightReflectionPipeline p;
        return p;
    }
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const float* lightDiffuse, const vx_gbuf& geometryFB,
             const float* geometryDepth, vx_volume* shadowVox, const vx_skybox& skybox, const float* palette,
             const vx_reflection_params& params, float* reflectionOut) {
        vx_reflection_io io;
        io.light = lightDiffuse;
        io.depth = geometryDepth;
        io.out = reflectionOut;
        io.t = nullptr;
        check(vx_reflection(ctx.get(), shadowVox, &viewBuffer, &geometryFB, palette, &io, &skybox, &params, cmd));
    }
    // The build's GGX specular-occlusion ray alone (spec = t / maxT per
    // pixel), the ray class vx_trace marches beside shadow and AO.
    void UseSpecularOcclusion(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_gbuf& geometryFB,
                              vx_volume* shadowVox, const float* palette, const vx_trace_params& params,
                              const vx_trace_out& reflectionOut) {
        vx_trace_params p = params;
        p.n_shadow = 0;
        p.n_ao = 0;
        p.n_spec = 1;
        check(vx_trace(ctx.get(), shadowVox, &viewBuffer, &geometryFB, palette, &p, &reflectionOut, cmd));
    }
};

// LightPointPipeline / LightSpotPipeline (LightPointPipeline.h:58-100,
// LightSpotPipeline.h:60-104): Use() opens the pass, the callback calls
// DrawLight once per light (at most 64 per type, as the reference's buffers),
// and the pass ends with one light-loop launch that adds every light's
// shadowed direct term to the light buffer. lightsDevice is caller-owned
// device memory for 64 vx_light entries; lightBufferRGB holds npx*3 floats.
class LightPointPipeline {
public:
    static constexpr int kMaxLights = 64;
    static LightPointPipeline& Get() {
        static LightPointPipeline p;
        return p;
    }
    template <class F>
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_gbuf& geometryFB, vx_volume* shadowVox,
             const float* palette, const vx_light_params& params, vx_light* lightsDevice, float* lightBufferRGB,
             F&& callback) {
        lights_.clear();
        callback(*this);
        if (lights_.empty()) return;
        check_hip(hipMemcpyAsync(lightsDevice, lights_.data(), lights_.size() * sizeof(vx_light),
                                 hipMemcpyHostToDevice, static_cast<hipStream_t>(cmd)));
        check(vx_light_pass(ctx.get(), shadowVox, &viewBuffer, &geometryFB, palette, lightsDevice,
                            (uint32_t)lights_.size(), &params, lightBufferRGB, nullptr, cmd));
    }
    // DrawLight(position, range, color * intensity, attenuation)
    void DrawLight(const float position[3], float range, const float color[3], float attenuation) {
        if ((int)lights_.size() >= kMaxLights) return;   // the reference drops lights past its buffer
        vx_light l;
        std::memset(&l, 0, sizeof(l));
        l.type = VX_LIGHT_POINT;
        std::memcpy(l.position, position, sizeof(l.position));
        l.range = range;
        std::memcpy(l.color, color, sizeof(l.color));
        l.attenuation = attenuation;
        lights_.push_back(l);
    }

private:
    static void check_hip(hipError_t e) {
        if (e != hipSuccess) throw Error(VX_ERR_DEVICE, hipGetErrorString(e));
    }
    std::vector<vx_light> lights_;
};

class LightSpotPipeline {
public:
    static constexpr int kMaxLights = 64;
    static LightSpotPipeline& Get() {
        static LightSpotPipeline p;
        return p;
    }
    template <class F>
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_gbuf& geometryFB, vx_volume* shadowVox,
             const float* palette, const vx_light_params& params, vx_light* lightsDevice, float* lightBufferRGB,
             F&& callback) {
        lights_.clear();
        callback(*this);
        if (lights_.empty()) return;
        hipError_t e = hipMemcpyAsync(lightsDevice, lights_.data(), lights_.size() * sizeof(vx_light),
                                      hipMemcpyHostToDevice, static_cast<hipStream_t>(cmd));
        if (e != hipSuccess) throw Error(VX_ERR_DEVICE, hipGetErrorString(e));
        check(vx_light_pass(ctx.get(), shadowVox, &viewBuffer, &geometryFB, palette, lightsDevice,
                            (uint32_t)lights_.size(), &params, lightBufferRGB, nullptr, cmd));
    }
    // DrawLight(position, range, color * intensity, attenuation, WorldMatrix[2], angle, angleAttenuation)
    void DrawLight(const float position[3], float range, const float color[3], float attenuation,
                   const float direction[3], float angle, float angleAttenuation) {
        if ((int)lights_.size() >= kMaxLights) return;
        vx_light l;
        std::memset(&l, 0, sizeof(l));
        l.type = VX_LIGHT_SPOT;
        std::memcpy(l.position, position, sizeof(l.position));
        l.range = range;
        std::memcpy(l.color, color, sizeof(l.color));
        l.attenuation = attenuation;
        std::memcpy(l.direction, direction, sizeof(l.direction));
        l.angle = angle;
        l.angle_attenuation = angleAttenuation;
        lights_.push_back(l);
    }

private:
    std::vector<vx_light> lights_;
};

// LightTAAPipeline (LightTAAPipeline.h:10-57): Use(cmd, viewBuffer, blueNoise,
// LastLightTAAFB, LightFB, GeometryFB) becomes one call over the device
// buffers; the caller swaps out and last between frames (WorldRenderer.cpp:140).
class LightTAAPipeline {
public:
    static LightTAAPipeline& Get() {
        static LightTAAPipeline p;
        return p;
    }
    void Use(Context& ctx, void* cmd, const vx_view& viewBuffer, const vx_taa_io& io, uint32_t seed = 0) {
        check(vx_light_taa(ctx.get(), &viewBuffer, &io, seed, cmd));
    }
};

}  // namespace vx
