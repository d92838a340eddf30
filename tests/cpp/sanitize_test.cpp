// Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the C ABI's
// host paths (SAH/SBVH builder, flattening and packing, the AABB tree, the
// engine's CPU traversal, argument checks) and the oracle's builder and
// traversal, in one program built by atlas-engine_amd/Makefile.sanitize.
// Inputs cover random soups of 1 to 20,000 triangles, a heightfield with
// exact equal-t ties, degenerate (zero-area, collinear, repeated) triangles
// and rays with tmin > 0.  The product must equal the oracle: nodes and
// triangle records byte for byte, the R-CPU hits (caller index and FP32 t)
// bit for bit, any-hit flags exactly.  Runs with no device.  Prints
// "sanitize ok".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "atlas_rt.h"

extern "C" {
struct OracleBVH;
OracleBVH* oracle_bvh_build_tris(const float* aabbs, const float* verts, uint32_t n);
uint32_t oracle_bvh_node_count(const OracleBVH* o);
uint32_t oracle_bvh_ref_count(const OracleBVH* o);
void oracle_bvh_export(const OracleBVH* o, atlas_rt_gpu_bvh_node* nodes, uint32_t* ref_idx, uint8_t* end_of_node,
                       const float* verts, atlas_rt_gpu_bvh_triangle* tris, int32_t material_idx, float opacity);
void oracle_bvh_free(OracleBVH* o);
void oracle_trace_closest(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                          const atlas_rt_gpu_bvh_triangle* tris, const uint32_t* orig_idx, uint32_t tri_count,
                          const atlas_rt_ray* rays, uint32_t n, atlas_rt_hit* hits, int32_t* slots, float* raw_t,
                          uint32_t* node_visits, uint32_t* tri_tests, uint8_t* ties, int32_t mode, int32_t threads);
void oracle_trace_any(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                      const atlas_rt_gpu_bvh_triangle* tris, uint32_t tri_count, const atlas_rt_ray* rays,
                      uint32_t n, uint32_t* hit, uint32_t* node_visits, uint32_t* tri_tests, int32_t mode,
                      int32_t threads);
void oracle_lbvh_build(const float* verts, uint32_t n, atlas_rt_gpu_bvh_node* nodes, atlas_rt_gpu_bvh_triangle* tris,
                       uint32_t* orig, uint32_t* max_depth);
void oracle_sincos(const float* x, float* sn, float* cs, float* p5, uint32_t n);
}

namespace {

int g_fail = 0;

#define CHECK(c, ...)                              \
    do {                                           \
        if (!(c)) {                                \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);              \
            std::printf("\n");                     \
            g_fail++;                              \
        }                                          \
    } while (0)

std::vector<float> soup(uint32_t n, uint32_t seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-10.0f, 10.0f), E(-1.0f, 1.0f);
    std::vector<float> v(9 * size_t(n));
    for (uint32_t i = 0; i < n; i++) {
        const float c[3] = {U(rng), U(rng), U(rng)};
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) v[9 * size_t(i) + 3 * k + a] = c[a] + E(rng);
    }
    return v;
}

// A grid of quads (two triangles each) on y = h(x, z): shared edges give
// exact equal-t ties.
std::vector<float> heightfield(uint32_t g) {
    std::vector<float> v;
    auto h = [](uint32_t x, uint32_t z) { return float((x * 7 + z * 13) % 5) * 0.25f; };
    for (uint32_t z = 0; z < g; z++)
        for (uint32_t x = 0; x < g; x++) {
            const float p[4][3] = {{float(x), h(x, z), float(z)},
                                   {float(x + 1), h(x + 1, z), float(z)},
                                   {float(x), h(x, z + 1), float(z + 1)},
                                   {float(x + 1), h(x + 1, z + 1), float(z + 1)}};
            const int tri[2][3] = {{0, 1, 2}, {1, 3, 2}};
            for (auto& t : tri)
                for (int k : t)
                    for (int a = 0; a < 3; a++) v.push_back(p[k][a]);
        }
    return v;
}

std::vector<atlas_rt_ray> rays_for(const std::vector<float>& v, uint32_t n, uint32_t seed, float tmin) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-12.0f, 12.0f), D(-1.0f, 1.0f);
    std::uniform_int_distribution<uint32_t> T(0, uint32_t(v.size() / 9) - 1);
    std::vector<atlas_rt_ray> r(n);
    for (uint32_t i = 0; i < n; i++) {
        float o[3] = {U(rng), U(rng), U(rng)}, d[3];
        if (i % 2 == 0) {  // aimed at a triangle's centroid
            const float* t = &v[9 * size_t(T(rng))];
            for (int a = 0; a < 3; a++) d[a] = (t[a] + t[3 + a] + t[6 + a]) / 3.0f - o[a];
        } else {
            for (int a = 0; a < 3; a++) d[a] = D(rng);
        }
        const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; a++) {
            r[i].origin[a] = o[a];
            r[i].direction[a] = len > 0.0f ? d[a] / len : (a == 0 ? 1.0f : 0.0f);
        }
        r[i].tmin = tmin;
        r[i].tmax = 1e30f;
    }
    return r;
}

void check_mesh(const char* name, const std::vector<float>& v, uint32_t n_rays, float tmin) {
    const uint32_t n = uint32_t(v.size() / 9);
    std::vector<atlas_rt_bvh_triangle> tris(n);
    for (uint32_t i = 0; i < n; i++) {
        std::memcpy(tris[i].v0, &v[9 * size_t(i)], 12);
        std::memcpy(tris[i].v1, &v[9 * size_t(i) + 3], 12);
        std::memcpy(tris[i].v2, &v[9 * size_t(i) + 6], 12);
        tris[i].idx = i;
    }
    for (uint32_t flags : {ATLAS_RT_BUILD_HOST_ONLY, ATLAS_RT_BUILD_HOST_ONLY | ATLAS_RT_BUILD_PARALLEL}) {
        atlas_rt_blas* b = nullptr;
        CHECK(atlas_rt_blas_build(nullptr, tris.data(), n, flags, &b) == ATLAS_RT_OK, "%s build: %s", name,
              atlas_rt_last_error());
        if (!b) continue;
        uint32_t nn = 0, nt = 0, depth = 0;
        CHECK(atlas_rt_blas_info(b, &nn, &nt, &depth) == ATLAS_RT_OK, "%s info", name);
        std::vector<atlas_rt_gpu_bvh_node> nodes(nn);
        std::vector<atlas_rt_gpu_bvh_triangle> packed(nt);
        std::vector<uint32_t> orig(nt);
        CHECK(atlas_rt_blas_export(b, nn ? nodes.data() : nullptr, packed.data(), orig.data()) == ATLAS_RT_OK,
              "%s export", name);
        // the oracle's build of the same mesh, byte for byte
        OracleBVH* o = oracle_bvh_build_tris(nullptr, v.data(), n);
        CHECK(o != nullptr, "%s oracle build", name);
        if (o) {
            CHECK(oracle_bvh_node_count(o) == nn && oracle_bvh_ref_count(o) == nt, "%s sizes %u/%u vs %u/%u", name,
                  oracle_bvh_node_count(o), oracle_bvh_ref_count(o), nn, nt);
            if (oracle_bvh_node_count(o) == nn && oracle_bvh_ref_count(o) == nt) {
                std::vector<atlas_rt_gpu_bvh_node> on(nn);
                std::vector<atlas_rt_gpu_bvh_triangle> ot(nt);
                std::vector<uint32_t> oref(nt);
                std::vector<uint8_t> oend(nt);
                oracle_bvh_export(o, on.data(), oref.data(), oend.data(), v.data(), ot.data(), 0, 1.0f);
                CHECK(std::memcmp(on.data(), nodes.data(), nn * sizeof(atlas_rt_gpu_bvh_node)) == 0,
                      "%s nodes differ", name);
                CHECK(std::memcmp(ot.data(), packed.data(), nt * sizeof(atlas_rt_gpu_bvh_triangle)) == 0,
                      "%s triangle records differ", name);
                CHECK(oref == orig, "%s reference order differs", name);
            }
            oracle_bvh_free(o);
        }
        // the engine's CPU traversal against the oracle's restatement of it
        const std::vector<atlas_rt_ray> rays = rays_for(v, n_rays, n + flags, tmin);
        std::vector<atlas_rt_hit> got(n_rays), ref(n_rays);
        std::vector<float> raw_t(n_rays);
        std::vector<uint32_t> any_got(n_rays), any_ref(n_rays);
        CHECK(atlas_rt_trace_closest_host(b, rays.data(), got.data(), n_rays) == ATLAS_RT_OK, "%s closest", name);
        CHECK(atlas_rt_trace_any_host(b, rays.data(), any_got.data(), n_rays) == ATLAS_RT_OK, "%s any", name);
        oracle_trace_closest(nn ? nodes.data() : nullptr, nn, packed.data(), orig.data(), nt, rays.data(), n_rays,
                             ref.data(), nullptr, raw_t.data(), nullptr, nullptr, nullptr, 1, 2);
        oracle_trace_any(nn ? nodes.data() : nullptr, nn, packed.data(), nt, rays.data(), n_rays, any_ref.data(),
                         nullptr, nullptr, 1, 2);
        uint32_t bad = 0, hits = 0;
        for (uint32_t i = 0; i < n_rays; i++) {
            hits += got[i].prim_id >= 0;
            const bool same = got[i].prim_id == ref[i].prim_id &&
                              (got[i].prim_id < 0 || std::memcmp(&got[i].t, &raw_t[i], 4) == 0) &&
                              any_got[i] == any_ref[i];
            bad += same ? 0 : 1;
        }
        CHECK(bad == 0, "%s: %u of %u rays differ from the oracle", name, bad, n_rays);
        std::printf("%s: %u tris, %u nodes, depth %u, %u/%u hits\n", name, n, nn, depth, hits, n_rays);
        CHECK(atlas_rt_blas_destroy(b) == ATLAS_RT_OK, "%s destroy", name);
    }
}

void check_aabb_tree() {
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(-50.0f, 50.0f), S(0.1f, 4.0f);
    for (uint32_t count : {1u, 2u, 3u, 17u, 1000u}) {
        std::vector<atlas_rt_aabb> boxes(count);
        for (auto& b : boxes)
            for (int a = 0; a < 3; a++) {
                b.min[a] = U(rng);
                b.max[a] = b.min[a] + S(rng);
            }
        uint32_t nc = 0, rc = 0;
        // capacities of zero: refused with the sizes written
        CHECK(atlas_rt_aabb_tree_build(boxes.data(), count, 0, nullptr, &nc, nullptr, nullptr, &rc) ==
                  ATLAS_RT_E_INVALID_ARG, "aabb tree capacity check");
        std::vector<atlas_rt_gpu_bvh_node> nodes(nc);
        std::vector<uint32_t> ref(rc);
        std::vector<uint8_t> end(rc);
        CHECK(atlas_rt_aabb_tree_build(boxes.data(), count, ATLAS_RT_BUILD_PARALLEL, nodes.data(), &nc, ref.data(),
                                       end.data(), &rc) == ATLAS_RT_OK, "aabb tree %u: %s", count,
              atlas_rt_last_error());
        CHECK(count == 1 ? (nc == 1 && rc == 2) : (nc == count - 1 && rc == count), "aabb tree sizes %u: %u %u",
              count, nc, rc);
    }
}

void check_arguments() {
    atlas_rt_blas* b = nullptr;
    CHECK(atlas_rt_blas_build(nullptr, nullptr, 4, ATLAS_RT_BUILD_HOST_ONLY, &b) != ATLAS_RT_OK, "NULL triangles");
    CHECK(atlas_rt_blas_build(nullptr, nullptr, 0, ATLAS_RT_BUILD_HOST_ONLY, &b) != ATLAS_RT_OK || b != nullptr,
          "empty mesh");
    if (b) atlas_rt_blas_destroy(b);
    b = nullptr;
    atlas_rt_bvh_triangle t{};
    CHECK(atlas_rt_blas_build(nullptr, &t, 1, 0x80u, &b) != ATLAS_RT_OK, "unknown flag");
    CHECK(atlas_rt_trace_closest_host(nullptr, nullptr, nullptr, 1) != ATLAS_RT_OK, "NULL blas");
    CHECK(atlas_rt_blas_destroy(nullptr) == ATLAS_RT_OK, "destroy NULL");
    CHECK(std::strlen(atlas_rt_last_error()) > 0, "error text");
}

void check_oracle_extras() {
    // the oracle's linear BVH over a soup, and its sin / cos / pow5
    const std::vector<float> v = soup(513, 9);
    const uint32_t n = 513;
    std::vector<atlas_rt_gpu_bvh_node> nodes(n - 1);
    std::vector<atlas_rt_gpu_bvh_triangle> tris(n);
    std::vector<uint32_t> orig(n);
    uint32_t depth = 0;
    oracle_lbvh_build(v.data(), n, nodes.data(), tris.data(), orig.data(), &depth);
    CHECK(depth >= 9 && depth < n, "lbvh depth %u", depth);
    std::vector<float> x(1000), sn(1000), cs(1000), p5(1000);
    for (int i = 0; i < 1000; i++) x[i] = float(i) * 0.0063f - 0.1f;
    oracle_sincos(x.data(), sn.data(), cs.data(), p5.data(), 1000);
    for (int i = 0; i < 1000; i++)
        CHECK(std::fabs(sn[i] * sn[i] + cs[i] * cs[i] - 1.0f) < 1e-6f, "sin^2 + cos^2 at %g", x[i]);
}

}  // namespace

int main() {
#if defined(__SANITIZE_ADDRESS__)
    std::printf("address sanitizer on\n");
#endif
    check_arguments();
    for (uint32_t n : {1u, 2u, 7u, 100u, 3000u}) {
        char name[32];
        std::snprintf(name, sizeof(name), "soup %u", n);
        check_mesh(name, soup(n, n), 4096, 0.0f);
    }
    check_mesh("soup 20000, tmin 0.5", soup(20000, 11), 8192, 0.5f);
    check_mesh("heightfield 24x24", heightfield(24), 4096, 0.0f);
    // degenerate triangles: zero area, collinear, repeated
    std::vector<float> deg = soup(64, 4);
    for (int i = 0; i < 8; i++) {
        float* t = &deg[9 * size_t(i)];
        for (int a = 0; a < 3; a++) t[3 + a] = t[a], t[6 + a] = t[a];  // a point
    }
    for (int i = 8; i < 16; i++) {
        float* t = &deg[9 * size_t(i)];
        for (int a = 0; a < 3; a++) t[6 + a] = 2.0f * t[3 + a] - t[a];  // collinear
    }
    for (int i = 16; i < 24; i++) std::memcpy(&deg[9 * size_t(i)], &deg[9 * size_t(i + 24)], 36);  // repeated
    check_mesh("degenerate", deg, 2048, 0.0f);
    check_aabb_tree();
    check_oracle_extras();
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
