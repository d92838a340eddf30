// Drop-in C++ facades for the engine's RayTracing::RayTracingWorld
// (src/engine/raytracing/RayTracingWorld.h:31-99) and
// Renderer::PathTracingRenderer (src/engine/renderer/PathTracingRenderer.h:16-104),
// backed by the atlas_rt C ABI (include/atlas_rt.h).  Header-only; link
// libatlas_rt.so.
//
// Public members and signatures follow the engine:
//   RayTracingWorld::Update(Scene::Subset<MeshComponent, TransformComponent>, bool updateTriangleLights)
//   RayTracingWorld::UpdateMaterials(), IsValid(), Clear(), includeObjectHistory
//   PathTracingRenderer::Render(renderTarget, scene, ivec2 imageSubdivisions, commandList)
//   PathTracingRenderer::ResetSampleCount(), GetSampleCount(), and the public
//   settings bounces, realTime, realTimeSamplesPerFrame, historyLengthMax,
//   historyClipMax, currentClipFactor, maxRadiance, sampleEmissives with the
//   engine's defaults.
// The engine's scene, mesh, camera and render-target classes are large; this
// header carries small stand-ins holding exactly what the two classes read:
//   Mesh::Mesh            triangles, vertex normals, material indices, materials,
//                         castShadow, BuildBVH() (Mesh::BuildBVH, Mesh.cpp:97-108)
//   MeshComponent         mesh
//   TransformComponent    globalMatrix, inverseGlobalMatrix, lastGlobalMatrix
//   Scene::Subset<M, T>   entities with both components; Get(entity)
//   Scene::Scene          the components, the RayTracingWorld, the main camera,
//                         the sun and the constant environment; Timestep()
//                         updates the world as Scene::Timestep does
//                         (Scene.cpp:285-297)
//   CameraComponent       location, rotation, field of view, aspect ratio,
//                         near plane, exposure, matrices; GetFrustumCorners
//   PathTracerRenderTarget the output images in device memory
// Behavioural notes:
//  * Update rebuilds the TLAS over every entity whose mesh has a BVH
//    (RayTracingWorld.cpp:32-162), one BLAS per distinct mesh, material
//    offsets in first-seen mesh order; with includeObjectHistory the
//    instances' last matrices go to the path tracer (velocity of realtime
//    frames).
//  * Errors throw std::runtime_error with atlas_rt_last_error()'s text (the
//    C ABI itself never throws).
//  * The command list argument is accepted for source compatibility: the
//    work is queued on the HIP stream given to the render target (NULL = the
//    null stream).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "atlas_rt.h"
#include "atlas_rt_volume.h"

namespace Atlas {

namespace detail {
inline void check(atlas_rt_status st) {
    if (st != ATLAS_RT_OK) throw std::runtime_error(atlas_rt_last_error());
}
}  // namespace detail

#ifdef ATLAS_RT_FACADE_USE_GLM
using vec2 = glm::vec2;
using ivec2 = glm::ivec2;
using mat4 = glm::mat4;
inline float mat_at(const mat4& m, int col, int row) { return m[col][row]; }
#else
struct vec2 {
    float x = 0.0f, y = 0.0f;
    vec2() = default;
    vec2(float a, float b) : x(a), y(b) {}
};
struct ivec2 {
    int32_t x = 0, y = 0;
    ivec2() = default;
    explicit ivec2(int32_t s) : x(s), y(s) {}
    ivec2(int32_t a, int32_t b) : x(a), y(b) {}
};
// Column-major 4x4 (m[column][row]), as glm stores it.
struct mat4 {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float* operator[](int c) { return m[c]; }
    const float* operator[](int c) const { return m[c]; }
};
inline float mat_at(const mat4& m, int col, int row) { return m.m[col][row]; }
#endif

namespace Mesh {

// The part of Mesh::Mesh / MeshData the ray tracer reads: triangles (three
// vertices each), per-vertex normals, a material index per triangle, the
// mesh's materials and castShadow.
class Mesh {
  public:
    Mesh(std::vector<vec3> positions, std::vector<vec3> normals, std::vector<int32_t> materialIndices,
         std::vector<atlas_rt_material> materials)
        : positions_(std::move(positions)),
          normals_(std::move(normals)),
          material_idx_(std::move(materialIndices)),
          materials(std::move(materials)) {
        if (positions_.size() % 3 || normals_.size() != positions_.size() ||
            material_idx_.size() * 3 != positions_.size())
            throw std::runtime_error("Mesh: need 3 positions and 3 normals per triangle and one material index");
    }

    // Mesh::BuildBVH (Mesh.cpp:97-108) -> MeshData::BuildBVH (MeshData.cpp:63-269):
    // the BLAS in the engine's layout on the device, with its shading data.
    void BuildBVH(bool parallelBuild = true) {
        const size_t n = material_idx_.size();
        std::vector<atlas_rt_bvh_triangle> tris(n);
        for (size_t i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++) {
                tris[i].v0[k] = positions_[3 * i][k];
                tris[i].v1[k] = positions_[3 * i + 1][k];
                tris[i].v2[k] = positions_[3 * i + 2][k];
            }
            tris[i].idx = uint32_t(i);
        }
        atlas_rt_blas* b = nullptr;
        detail::check(atlas_rt_blas_build(nullptr, tris.data(), uint32_t(n),
                                          parallelBuild ? ATLAS_RT_BUILD_PARALLEL : ATLAS_RT_BUILD_SERIAL, &b));
        blas_.reset(b);
        std::vector<float> nrm(9 * n);
        for (size_t i = 0; i < 3 * n; i++)
            for (int k = 0; k < 3; k++) nrm[3 * i + k] = normals_[i][k];
        detail::check(atlas_rt_blas_set_shading(b, nrm.data(), material_idx_.data()));
    }

    bool IsBVHBuilt() const { return blas_ != nullptr; }
    const atlas_rt_blas* blas() const { return blas_.get(); }

  private:
    struct Deleter {
        void operator()(atlas_rt_blas* b) const { atlas_rt_blas_destroy(b); }
    };
    std::vector<vec3> positions_;
    std::vector<vec3> normals_;
    std::vector<int32_t> material_idx_;
    std::unique_ptr<atlas_rt_blas, Deleter> blas_;

  public:
    bool castShadow = true;
    std::vector<atlas_rt_material> materials;
};

}  // namespace Mesh

struct MeshComponent {
    Mesh::Mesh* mesh = nullptr;
};

struct TransformComponent {
    mat4 globalMatrix;
    mat4 inverseGlobalMatrix;
    mat4 lastGlobalMatrix;
};

namespace Scene {

using Entity = uint32_t;

template <class... Components>
class Subset;

// Entities that have both a mesh and a transform component.
template <>
class Subset<MeshComponent, TransformComponent> {
  public:
    Subset() = default;
    explicit Subset(std::vector<std::pair<MeshComponent*, TransformComponent*>> items) : items_(std::move(items)) {}

    struct iterator {
        Entity e;
        Entity operator*() const { return e; }
        iterator& operator++() {
            ++e;
            return *this;
        }
        bool operator!=(const iterator& o) const { return e != o.e; }
    };
    iterator begin() const { return iterator{0}; }
    iterator end() const { return iterator{Entity(items_.size())}; }
    size_t Size() const { return items_.size(); }
    std::tuple<MeshComponent&, TransformComponent&> Get(Entity e) const {
        return {*items_[e].first, *items_[e].second};
    }

  private:
    std::vector<std::pair<MeshComponent*, TransformComponent*>> items_;
};

}  // namespace Scene

namespace RayTracing {

class RayTracingWorld {
  public:
    RayTracingWorld() {
        atlas_rt_world* w = nullptr;
        detail::check(atlas_rt_world_create(&w));
        world_.reset(w);
    }

    // RayTracingWorld::Update + UpdateForSoftwareRayTracing
    // (RayTracingWorld.cpp:32-162, 267-307).
    void Update(Scene::Subset<MeshComponent, TransformComponent> subset, bool updateTriangleLights) {
        std::vector<atlas_rt_instance_desc> inst;
        std::vector<float> last;
        std::unordered_map<const Mesh::Mesh*, uint32_t> index;
        std::vector<const atlas_rt_blas*> blases;
        std::vector<int32_t> mat_offset;
        std::vector<atlas_rt_material> materials;
        for (Scene::Entity e : subset) {
            const auto& [mc, tc] = subset.Get(e);
            // entities whose mesh has no BVH yet are skipped (RayTracingWorld.cpp:93-95)
            if (!mc.mesh || !mc.mesh->IsBVHBuilt()) continue;
            auto it = index.find(mc.mesh);
            if (it == index.end()) {
                it = index.emplace(mc.mesh, uint32_t(blases.size())).first;
                blases.push_back(mc.mesh->blas());
                mat_offset.push_back(int32_t(materials.size()));
                materials.insert(materials.end(), mc.mesh->materials.begin(), mc.mesh->materials.end());
            }
            atlas_rt_instance_desc d{};
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 4; c++) {
                    d.object_to_world[r][c] = mat_at(tc.globalMatrix, c, r);
                    d.world_to_object[r][c] = mat_at(tc.inverseGlobalMatrix, c, r);
                }
            d.blas = it->second;
            d.material_offset = mat_offset[it->second];
            d.mask = ATLAS_RT_MASK_ALL | (mc.mesh->castShadow ? ATLAS_RT_MASK_SHADOW : 0u);
            inst.push_back(d);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 4; c++) last.push_back(mat_at(tc.lastGlobalMatrix, c, r));
        }
        valid_ = !inst.empty();
        if (!valid_) return;  // the engine keeps an empty world invalid
        detail::check(atlas_rt_world_update(world_.get(), inst.data(), uint32_t(inst.size()), nullptr,
                                            blases.data(), uint32_t(blases.size())));
        if (includeObjectHistory)
            detail::check(atlas_rt_world_set_history(world_.get(), last.data(), uint32_t(inst.size())));
        materials_ = std::move(materials);
        triangle_lights_ = updateTriangleLights;
        generation_++;
    }

    // The material table is rebuilt by every Update here; UpdateMaterials
    // (RayTracingWorld.cpp:164-251) only marks it for the renderer to upload.
    void UpdateMaterials() { generation_++; }

    bool IsValid() { return valid_; }

    void Clear() {
        atlas_rt_world* w = nullptr;
        detail::check(atlas_rt_world_create(&w));
        world_.reset(w);
        materials_.clear();
        valid_ = false;
        generation_++;
    }

    bool includeObjectHistory = false;

    // What the engine's RayTracingHelper reads as a friend.
    atlas_rt_world* handle() const { return world_.get(); }
    const std::vector<atlas_rt_material>& materials() const { return materials_; }
    uint64_t generation() const { return generation_; }

  private:
    struct Deleter {
        void operator()(atlas_rt_world* w) const { atlas_rt_world_destroy(w); }
    };
    std::unique_ptr<atlas_rt_world, Deleter> world_;
    std::vector<atlas_rt_material> materials_;
    bool valid_ = false;
    bool triangle_lights_ = false;
    uint64_t generation_ = 0;
};

}  // namespace RayTracing

// The camera as PathTracingRenderer reads it: location, rotation
// (yaw, pitch in radians), vertical field of view (degrees), aspect ratio,
// near plane, exposure; GetFrustumCorners gives the near-plane corners
// 4 (top left), 5 (top right), 6 (bottom left) the ray generation uses
// (PathTracingRenderer.cpp:157-162).
struct CameraComponent {
    vec3 location = vec3(0.0f);
    vec2 rotation = vec2(0.0f, 0.0f);
    float fieldOfView = 45.0f;
    float aspectRatio = 1.0f;
    float nearPlane = 0.1f;
    float farPlane = 1000.0f;
    float exposure = 1.0f;
    mat4 viewMatrix;
    mat4 projectionMatrix;
    mat4 lastJitteredMatrix;  // last frame's jittered view-projection
    vec2 jitter = vec2(0.0f, 0.0f);

    vec3 GetLocation() const { return location; }

    std::vector<vec3> GetFrustumCorners(float nearDist, float farDist) const {
        const float yaw = rotation.x, pitch = rotation.y;
        const vec3 fwd(std::cos(pitch) * std::sin(yaw), -std::sin(pitch), std::cos(pitch) * std::cos(yaw));
        vec3 right(-fwd.z, 0.0f, fwd.x);  // fwd x (0, 1, 0)
        const float rl = std::sqrt(right.x * right.x + right.z * right.z);
        right = vec3(right.x / rl, 0.0f, right.z / rl);
        const vec3 up(right.y * fwd.z - right.z * fwd.y, right.z * fwd.x - right.x * fwd.z,
                      right.x * fwd.y - right.y * fwd.x);
        std::vector<vec3> c(8);
        const float dist[2] = {nearDist, farDist};
        for (int p = 0; p < 2; p++) {
            const float h = dist[p] * std::tan(fieldOfView * 3.14159265358979f / 360.0f);
            const float w = aspectRatio * h;
            const vec3 mid(location.x + fwd.x * dist[p], location.y + fwd.y * dist[p], location.z + fwd.z * dist[p]);
            const float sx[4] = {-1, 1, -1, 1}, sy[4] = {1, 1, -1, -1};
            for (int k = 0; k < 4; k++)
                c[4 * (1 - p) + k] = vec3(mid.x + sy[k] * h * up.x + sx[k] * w * right.x,
                                          mid.y + sy[k] * h * up.y + sx[k] * w * right.y,
                                          mid.z + sy[k] * h * up.z + sx[k] * w * right.z);
        }
        return c;  // 4..7: near plane (the corners the ray generation reads)
    }
};

namespace Scene {

class Scene {
  public:
    void Add(MeshComponent* mesh, TransformComponent* transform) { entities_.emplace_back(mesh, transform); }
    Subset<MeshComponent, TransformComponent> GetSubset() const { return Subset<MeshComponent, TransformComponent>(entities_); }
    // Scene::Timestep's ray-tracing part (Scene.cpp:285-297).
    void Timestep() { rayTracingWorld->Update(GetSubset(), true); }
    bool IsRtDataValid() { return rayTracingWorld && rayTracingWorld->IsValid(); }
    CameraComponent& GetMainCamera() { return mainCamera; }

    std::shared_ptr<RayTracing::RayTracingWorld> rayTracingWorld = std::make_shared<RayTracing::RayTracingWorld>();
    CameraComponent mainCamera;
    atlas_rt_light sun{{0.0f, -1.0f, 0.0f}, ATLAS_RT_LIGHT_DIRECTIONAL, {1.0f, 1.0f, 1.0f}, 0.0f};
    vec3 environment = vec3(0.0f);

  private:
    std::vector<std::pair<MeshComponent*, TransformComponent*>> entities_;
};

}  // namespace Scene

namespace Renderer {

// The path tracer's images (PathTracerRenderTarget): the progressive
// accumulation (rgba32f, sample count in w) or, in realtime mode, the
// resolved radiance and the G-buffer, read into caller device memory.
class PathTracerRenderTarget {
  public:
    PathTracerRenderTarget(int32_t width, int32_t height, void* stream = nullptr)
        : width_(width), height_(height), stream_(stream) {}
    int32_t GetWidth() const { return width_; }
    int32_t GetHeight() const { return height_; }
    void* stream() const { return stream_; }

    // Copies into d_out: the accumulation image (w * h float4), or in
    // realtime mode one of ATLAS_RT_RT_* (atlas_rt_pathtracer_read_realtime).
    void Read(void* d_out, uint32_t realtimeBuffer = ATLAS_RT_RT_RESOLVED) const {
        if (!tracer_) throw std::runtime_error("PathTracerRenderTarget: nothing rendered yet");
        if (realtime_)
            detail::check(atlas_rt_pathtracer_read_realtime(tracer_, realtimeBuffer, d_out, stream_));
        else
            detail::check(atlas_rt_pathtracer_read(tracer_, static_cast<float*>(d_out), stream_));
    }

  private:
    friend class PathTracingRenderer;
    int32_t width_, height_;
    void* stream_;
    atlas_rt_pathtracer* tracer_ = nullptr;
    bool realtime_ = false;
};

class PathTracingRenderer {
  public:
    PathTracingRenderer() = default;
    ~PathTracingRenderer() { release(); }
    PathTracingRenderer(const PathTracingRenderer&) = delete;
    PathTracingRenderer& operator=(const PathTracingRenderer&) = delete;

    // PathTracingRenderer::Render (PathTracingRenderer.cpp:29-292): a camera
    // move resets the sample count; realtime mode renders
    // realTimeSamplesPerFrame paths per pixel and resolves them against the
    // history, else one progressive sample per pixel is accumulated.
    // imageSubdivisions and commandList are accepted for source
    // compatibility (the whole image is one launch).
    void Render(PathTracerRenderTarget& renderTarget, Scene::Scene& scene, ivec2 imageSubdivisions,
                void* commandList) {
        (void)imageSubdivisions;
        (void)commandList;
        if (!scene.IsRtDataValid()) return;
        RayTracing::RayTracingWorld& world = *scene.rayTracingWorld;
        CameraComponent& camera = scene.GetMainCamera();
        const bool moved = dist3(camera.GetLocation(), cameraLocation_) > 1e-3f ||
                           std::hypot(camera.rotation.x - cameraRotation_.x, camera.rotation.y - cameraRotation_.y) >
                               1e-3f;
        const bool rebuild = !tracer_ || world.handle() != world_ || world.generation() != world_gen_ ||
                             renderTarget.GetWidth() != width_ || renderTarget.GetHeight() != height_ ||
                             bounces != bounces_ || realTime != realtime_ || realTimeSamplesPerFrame != spf_ ||
                             sampleEmissives != emissives_;
        if (rebuild) create(renderTarget, scene);
        if (moved) {
            cameraLocation_ = camera.GetLocation();
            cameraRotation_ = camera.rotation;
            if (!realTime) detail::check(atlas_rt_pathtracer_reset(tracer_));
        }
        renderTarget.tracer_ = tracer_;
        renderTarget.realtime_ = realTime;
        const std::vector<vec3> corners = camera.GetFrustumCorners(camera.nearPlane, camera.farPlane);
        atlas_rt_camera cam;
        for (int k = 0; k < 3; k++) {
            cam.location[k] = camera.location[k];
            cam.origin[k] = corners[4][k];
            cam.right[k] = corners[5][k] - corners[4][k];
            cam.bottom[k] = corners[6][k] - corners[4][k];
        }
        if (realTime) {
            atlas_rt_frame_matrices fm;
            for (int c = 0; c < 4; c++)
                for (int r = 0; r < 4; r++) {
                    fm.view[c][r] = mat_at(camera.viewMatrix, c, r);
                    fm.projection[c][r] = mat_at(camera.projectionMatrix, c, r);
                    fm.last_view_projection[c][r] = mat_at(camera.lastJitteredMatrix, c, r);
                }
            fm.jitter[0] = camera.jitter.x;
            fm.jitter[1] = camera.jitter.y;
            detail::check(atlas_rt_pathtracer_render_realtime(tracer_, &cam, &fm, renderTarget.stream()));
        } else {
            detail::check(atlas_rt_pathtracer_render(tracer_, &cam, 1, renderTarget.stream()));
        }
    }

    bool UpdateData(Scene::Scene* scene) { return scene && scene->IsRtDataValid(); }
    void UpdateMaterials(Scene::Scene* scene) {
        if (scene && scene->rayTracingWorld) scene->rayTracingWorld->UpdateMaterials();
    }

    void ResetSampleCount() {
        if (tracer_) detail::check(atlas_rt_pathtracer_reset(tracer_));
    }

    int32_t GetSampleCount() const {
        int32_t n = 0;
        if (tracer_) detail::check(atlas_rt_pathtracer_sample_count(tracer_, &n));
        return n;
    }

    int32_t bounces = 10;
    int32_t bvhDepth = 0;
    int32_t lightCount = 512;

    bool realTime = true;
    int32_t realTimeSamplesPerFrame = 4;

    int32_t historyLengthMax = 32;
    float historyClipMax = 0.2f;
    float currentClipFactor = 2.0f;

    float maxRadiance = 65535.0f;

    bool sampleEmissives = false;

  private:
    static float dist3(const vec3& a, const vec3& b) {
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }

    void release() {
        if (tracer_) atlas_rt_pathtracer_destroy(tracer_);
        tracer_ = nullptr;
    }

    void create(PathTracerRenderTarget& target, Scene::Scene& scene) {
        release();
        RayTracing::RayTracingWorld& world = *scene.rayTracingWorld;
        atlas_rt_pt_params p{};
        p.width = uint32_t(target.GetWidth());
        p.height = uint32_t(target.GetHeight());
        p.bounces = bounces;
        p.exposure = scene.GetMainCamera().exposure;
        for (int k = 0; k < 3; k++) p.env_radiance[k] = scene.environment[k];
        p.flags = 0;
        const std::vector<atlas_rt_material>& mats = world.materials();
        detail::check(atlas_rt_pathtracer_create(world.handle(), &p, mats.data(), uint32_t(mats.size()), &scene.sun, 1,
                                                 &tracer_));
        if (realTime) {
            atlas_rt_realtime_params rt{realTimeSamplesPerFrame, historyLengthMax, historyClipMax, currentClipFactor,
                                        maxRadiance};
            detail::check(atlas_rt_pathtracer_set_realtime(tracer_, &rt));
        }
        width_ = target.GetWidth();
        height_ = target.GetHeight();
        bounces_ = bounces;
        realtime_ = realTime;
        spf_ = realTimeSamplesPerFrame;
        emissives_ = sampleEmissives;
        world_ = world.handle();
        world_gen_ = world.generation();
    }

    atlas_rt_pathtracer* tracer_ = nullptr;
    const atlas_rt_world* world_ = nullptr;
    uint64_t world_gen_ = 0;
    int32_t width_ = 0, height_ = 0, bounces_ = 0, spf_ = 0;
    bool realtime_ = false, emissives_ = false;
    vec3 cameraLocation_ = vec3(0.0f);
    vec2 cameraRotation_ = vec2(0.0f, 0.0f);
};

}  // namespace Renderer
}  // namespace Atlas
