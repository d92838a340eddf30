// Drop-in check of the RayTracingWorld / PathTracingRenderer facades:
// engine-style call sites (Scene::Timestep's world update, then
// PathTracingRenderer::Render with a render target) compile unchanged
// against atlas_rt_raytracing.h.  With a GPU they also render: the progressive
// image after two Render calls must equal, bit for bit, what the C ABI gives
// for the same world, camera and settings; a camera move restarts the sample
// count; realtime mode resolves a frame with radiance.  Prints
// "raytracing facade ok ...".
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "atlas_rt_raytracing.h"

using namespace Atlas;

namespace {

// Latitude-longitude sphere: 3 positions, 3 normals, 1 material per triangle.
Mesh::Mesh* sphere(int rings, int segs, float r, atlas_rt_material mat) {
    std::vector<vec3> P, N;
    std::vector<int32_t> M;
    auto at = [&](int i, int j) {
        const float th = 3.14159265f * float(i) / float(rings), ph = 6.2831853f * float(j) / float(segs);
        return vec3(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
    };
    for (int i = 0; i < rings; i++)
        for (int j = 0; j < segs; j++) {
            const vec3 q[4] = {at(i, j), at(i + 1, j), at(i + 1, j + 1), at(i, j + 1)};
            const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
            for (const auto& t : tri) {
                for (int k : t) {
                    P.emplace_back(q[k].x * r, q[k].y * r, q[k].z * r);
                    N.push_back(q[k]);
                }
                M.push_back(0);
            }
        }
    return new Mesh::Mesh(P, N, M, {mat});
}

Mesh::Mesh* ground(float s, atlas_rt_material mat) {
    std::vector<vec3> P = {vec3(-s, 0, -s), vec3(s, 0, s), vec3(s, 0, -s), vec3(-s, 0, -s), vec3(-s, 0, s), vec3(s, 0, s)};
    std::vector<vec3> N(6, vec3(0, 1, 0));
    return new Mesh::Mesh(P, N, {0, 0}, {mat});
}

atlas_rt_material material(float r, float g, float b, float rough) {
    atlas_rt_material m;
    std::memset(&m, 0, sizeof(m));
    m.base_color[0] = r;
    m.base_color[1] = g;
    m.base_color[2] = b;
    m.opacity = 1.0f;
    m.roughness = rough;
    m.ao = 1.0f;
    m.reflectance = 0.5f;
    m.opacity_texture = m.base_color_texture = m.roughness_texture = -1;
    m.metalness_texture = m.ao_texture = m.normal_texture = -1;
    return m;
}

TransformComponent translate(float x, float y, float z) {
    TransformComponent t;
    t.globalMatrix[3][0] = x;
    t.globalMatrix[3][1] = y;
    t.globalMatrix[3][2] = z;
    t.inverseGlobalMatrix[3][0] = -x;
    t.inverseGlobalMatrix[3][1] = -y;
    t.inverseGlobalMatrix[3][2] = -z;
    t.lastGlobalMatrix = t.globalMatrix;
    return t;
}

std::vector<float> download(const void* d, size_t floats) {
    std::vector<float> h(floats);
    if (hipMemcpy(h.data(), d, floats * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return {};
    return h;
}

}  // namespace

int main() {
    Mesh::Mesh* ball = sphere(12, 24, 1.0f, material(0.8f, 0.3f, 0.2f, 0.4f));
    Mesh::Mesh* floor = ground(20.0f, material(0.6f, 0.6f, 0.6f, 0.9f));
    if (atlas_rt_device_count() == 0) {
        std::printf("raytracing facade ok (host only): compiled\n");
        return 0;
    }
    ball->BuildBVH();
    floor->BuildBVH();
    MeshComponent mc_ball{ball}, mc_floor{floor};
    TransformComponent tf[4] = {translate(-2.5f, 1.0f, 0.0f), translate(0.0f, 1.0f, 1.5f), translate(2.5f, 1.0f, 0.0f),
                                translate(0.0f, 0.0f, 0.0f)};
    Scene::Scene scene;
    scene.Add(&mc_ball, &tf[0]);
    scene.Add(&mc_ball, &tf[1]);
    scene.Add(&mc_ball, &tf[2]);
    scene.Add(&mc_floor, &tf[3]);
    scene.sun = atlas_rt_light{{0.3f, -1.0f, 0.2f}, ATLAS_RT_LIGHT_DIRECTIONAL, {3.0f, 2.9f, 2.7f}, 0.0f};
    scene.environment = vec3(0.25f, 0.3f, 0.4f);
    CameraComponent& cam = scene.GetMainCamera();
    cam.location = vec3(0.0f, 3.0f, -9.0f);
    cam.rotation = vec2(0.0f, 0.25f);
    cam.fieldOfView = 50.0f;
    const int W = 96, H = 64;
    cam.aspectRatio = float(W) / float(H);
    scene.Timestep();
    if (!scene.IsRtDataValid()) return 2;

    Renderer::PathTracingRenderer renderer;
    renderer.realTime = false;
    renderer.bounces = 3;
    Renderer::PathTracerRenderTarget target(W, H);
    renderer.Render(target, scene, ivec2(1), nullptr);
    renderer.Render(target, scene, ivec2(1), nullptr);
    if (renderer.GetSampleCount() != 2) return 3;
    void* d_img = nullptr;
    if (hipMalloc(&d_img, size_t(W) * H * 16) != hipSuccess) return 4;
    target.Read(d_img);
    const std::vector<float> img = download(d_img, size_t(W) * H * 4);

    // the same frames through the C ABI
    const std::vector<vec3> corners = cam.GetFrustumCorners(cam.nearPlane, cam.farPlane);
    atlas_rt_camera c;
    for (int k = 0; k < 3; k++) {
        c.location[k] = cam.location[k];
        c.origin[k] = corners[4][k];
        c.right[k] = corners[5][k] - corners[4][k];
        c.bottom[k] = corners[6][k] - corners[4][k];
    }
    atlas_rt_world* w = nullptr;
    if (atlas_rt_world_create(&w) != ATLAS_RT_OK) return 5;
    atlas_rt_instance_desc inst[4];
    std::memset(inst, 0, sizeof(inst));
    for (int i = 0; i < 4; i++) {
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 4; k++) {
                inst[i].object_to_world[r][k] = tf[i].globalMatrix[k][r];
                inst[i].world_to_object[r][k] = tf[i].inverseGlobalMatrix[k][r];
            }
        inst[i].blas = i < 3 ? 0u : 1u;
        inst[i].material_offset = i < 3 ? 0 : 1;
        inst[i].mask = ATLAS_RT_MASK_ALL | ATLAS_RT_MASK_SHADOW;
    }
    const atlas_rt_blas* blases[2] = {ball->blas(), floor->blas()};
    if (atlas_rt_world_update(w, inst, 4, nullptr, blases, 2) != ATLAS_RT_OK) return 6;
    atlas_rt_material mats[2] = {ball->materials[0], floor->materials[0]};
    atlas_rt_pt_params p{};
    p.width = W;
    p.height = H;
    p.bounces = 3;
    p.exposure = 1.0f;
    p.env_radiance[0] = 0.25f;
    p.env_radiance[1] = 0.3f;
    p.env_radiance[2] = 0.4f;
    atlas_rt_pathtracer* pt = nullptr;
    if (atlas_rt_pathtracer_create(w, &p, mats, 2, &scene.sun, 1, &pt) != ATLAS_RT_OK) return 7;
    if (atlas_rt_pathtracer_render(pt, &c, 2, nullptr) != ATLAS_RT_OK) return 8;
    void* d_ref = nullptr;
    if (hipMalloc(&d_ref, size_t(W) * H * 16) != hipSuccess) return 9;
    if (atlas_rt_pathtracer_read(pt, static_cast<float*>(d_ref), nullptr) != ATLAS_RT_OK) return 10;
    const std::vector<float> ref = download(d_ref, size_t(W) * H * 4);
    if (img.empty() || img.size() != ref.size() || std::memcmp(img.data(), ref.data(), img.size() * 4) != 0) {
        std::printf("facade image differs from the C ABI's\n");
        return 11;
    }
    double sum = 0.0;
    for (size_t i = 0; i < img.size(); i += 4) sum += img[i] + img[i + 1] + img[i + 2];
    if (!(sum > 0.0)) return 12;

    // a camera move restarts the accumulation
    cam.location = vec3(0.5f, 3.0f, -9.0f);
    renderer.Render(target, scene, ivec2(1), nullptr);
    if (renderer.GetSampleCount() != 1) return 13;

    // realtime mode: one frame of four samples per pixel, resolved
    Renderer::PathTracingRenderer rt;
    rt.bounces = 2;
    scene.rayTracingWorld->includeObjectHistory = true;
    scene.Timestep();
    Renderer::PathTracerRenderTarget rt_target(W, H);
    rt.Render(rt_target, scene, ivec2(1), nullptr);
    rt.Render(rt_target, scene, ivec2(1), nullptr);
    rt_target.Read(d_img, ATLAS_RT_RT_RESOLVED);
    const std::vector<float> res = download(d_img, size_t(W) * H * 4);
    double rsum = 0.0;
    for (size_t i = 0; i < res.size(); i += 4) rsum += res[i] + res[i + 1] + res[i + 2];
    if (!(rsum > 0.0)) return 14;

    atlas_rt_pathtracer_destroy(pt);
    atlas_rt_world_destroy(w);
    (void)hipFree(d_img);
    (void)hipFree(d_ref);
    std::printf("raytracing facade ok: progressive image equals the C ABI's (mean %.4f), realtime mean %.4f\n",
                sum / (3.0 * W * H), rsum / (3.0 * W * H));
    return 0;
}
