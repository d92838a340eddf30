"""Seeded synthetic workloads for the ray-tracing hot path.

Each generator is deterministic (numpy PCG64 seeded with 0xA71A5 + config id)
and returns float32 arrays in the layouts of include/atlas_rt.h.

  c0  random triangle soup: 65,536 triangles with centroids U[-10,10]^3, edge
      lengths U[0.05,0.5] at an angle U[20,160] deg, uniformly random rotation;
      1,048,576 rays with origins U[-10,10]^3 and directions uniform on the
      sphere, tmin 1e-4, tmax 1e12 (the shaders' INF, common.hsh:8).
  c1  1024 x 1024 heightfield grid (2,093,058 triangles), 6-octave value noise
      of amplitude 50, cell 1.0; 1920x1080 pinhole camera pitched 30 deg down;
      shadow rays toward a fixed directional light.
"""
import math

import numpy as np

from ._abi import RAY_DTYPE

SEED_BASE = 0xA71A5
INF = 1.0e12  # data/shader/raytracer/common.hsh:8


def rng(config_id, salt=0):
    return np.random.Generator(np.random.PCG64([SEED_BASE + config_id, salt]))


def _random_rotations(g, n):
    q = g.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def triangle_soup(n_tris, g, extent=10.0, edge=(0.05, 0.5)):
    """(n, 9) float32 random triangles (config 0 recipe)."""
    c = g.uniform(-extent, extent, (n_tris, 3))
    a = g.uniform(edge[0], edge[1], n_tris)
    b = g.uniform(edge[0], edge[1], n_tris)
    phi = np.radians(g.uniform(20.0, 160.0, n_tris))
    e1 = np.zeros((n_tris, 3))
    e1[:, 0] = a
    e2 = np.zeros((n_tris, 3))
    e2[:, 0] = b * np.cos(phi)
    e2[:, 1] = b * np.sin(phi)
    R = _random_rotations(g, n_tris)
    e1 = np.einsum("nij,nj->ni", R, e1)
    e2 = np.einsum("nij,nj->ni", R, e2)
    v0 = c - (e1 + e2) / 3.0
    v1 = v0 + e1
    v2 = v0 + e2
    return np.ascontiguousarray(np.concatenate([v0, v1, v2], axis=1).astype(np.float32))


def random_rays(n_rays, g, extent=10.0, tmin=1e-4, tmax=INF):
    """Incoherent rays: origins U[-extent, extent]^3, directions uniform on S^2."""
    rays = np.zeros(n_rays, dtype=RAY_DTYPE)
    rays["origin"] = g.uniform(-extent, extent, (n_rays, 3)).astype(np.float32)
    d = g.standard_normal((n_rays, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["direction"] = d.astype(np.float32)
    rays["tmin"] = np.float32(tmin)
    rays["tmax"] = np.float32(tmax)
    return rays


def c0(n_tris=65536, n_rays=1048576, seed_salt=0):
    """Plain closest-hit workload (config 0): (verts (n,9) float32, rays)."""
    g = rng(0, seed_salt)
    verts = triangle_soup(n_tris, g)
    rays = random_rays(n_rays, g)
    return verts, rays


# ---------------------------------------------------------------------------
# Config 1: heightfield + camera + shadow rays
# ---------------------------------------------------------------------------
def _lattice_hash(ix, iy, octave, seed):
    """Integer lattice hash -> [0,1) (PCG-style mixing on uint64)."""
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
    h ^= np.uint64(seed * 0x165667B19E3779F9 + octave * 0x27D4EB2F165667C5 & 0xFFFFFFFFFFFFFFFF)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def value_noise(x, y, octaves=6, seed=1, base_freq=1.0 / 64.0):
    total = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        fx, fy = x * freq, y * freq
        ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
        tx, ty = fx - ix, fy - iy
        sx, sy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        n00 = _lattice_hash(ix, iy, o, seed)
        n10 = _lattice_hash(ix + 1, iy, o, seed)
        n01 = _lattice_hash(ix, iy + 1, o, seed)
        n11 = _lattice_hash(ix + 1, iy + 1, o, seed)
        nx0 = n00 + sx * (n10 - n00)
        nx1 = n01 + sx * (n11 - n01)
        total += amp * (nx0 + sy * (nx1 - nx0))
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def heightfield(grid=1024, amplitude=50.0, cell=1.0, seed=1):
    """(2*(grid-1)^2, 9) float32 triangles of a value-noise heightfield in the xz plane."""
    xs = np.arange(grid, dtype=np.float64) * cell
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    H = amplitude * value_noise(X, Z, seed=seed)
    P = np.stack([X, H, Z], axis=-1).astype(np.float32)  # [z][x][3]
    p00 = P[:-1, :-1].reshape(-1, 3)
    p10 = P[:-1, 1:].reshape(-1, 3)
    p01 = P[1:, :-1].reshape(-1, 3)
    p11 = P[1:, 1:].reshape(-1, 3)
    t0 = np.concatenate([p00, p01, p10], axis=1)
    t1 = np.concatenate([p10, p01, p11], axis=1)
    tris = np.empty((2 * len(p00), 9), dtype=np.float32)
    tris[0::2] = t0
    tris[1::2] = t1
    return tris


def pinhole_rays(width, height, eye, yaw_deg, pitch_deg, fov_y_deg=60.0, tmin=1e-4, tmax=INF):
    """Primary camera rays, one per pixel centre, row-major from the top-left."""
    pitch, yaw = math.radians(pitch_deg), math.radians(yaw_deg)
    fwd = np.array([math.cos(pitch) * math.sin(yaw), -math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = math.tan(math.radians(fov_y_deg) / 2.0)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width * 2.0 - 1.0
    py = 1.0 - (np.arange(height) + 0.5) / height * 2.0
    PX, PY = np.meshgrid(px, py, indexing="xy")
    d = fwd[None, None, :] + PX[..., None] * (th * aspect) * right + PY[..., None] * th * up
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros(width * height, dtype=RAY_DTYPE)
    rays["origin"] = np.asarray(eye, dtype=np.float32)
    rays["direction"] = d.reshape(-1, 3).astype(np.float32)
    rays["tmin"] = np.float32(tmin)
    rays["tmax"] = np.float32(tmax)
    return rays


C1_LIGHT_DIR = np.array([0.4, 0.8, 0.3]) / np.linalg.norm([0.4, 0.8, 0.3])


# ---------------------------------------------------------------------------
# Config 2: instanced blob meshes (two-level)
# ---------------------------------------------------------------------------
def blob_mesh(g, stacks=32, slices=64, bumps=3, amplitude=0.25):
    """A unit sphere (stacks x slices lat-long, 2*stacks*slices - 2*slices
    triangles = 3,968 for 32 x 64) radially displaced by a few random
    spherical-harmonic-like bumps.  Returns (verts (n,9) float32, vertex
    normals (n,9) float32)."""
    th = np.linspace(0, np.pi, stacks + 1)
    ph = np.linspace(0, 2 * np.pi, slices + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    r = np.ones_like(T)
    for _ in range(bumps):
        k = g.standard_normal(3)
        k /= np.linalg.norm(k)
        f = g.uniform(2.0, 5.0)
        phase = g.uniform(0, 2 * np.pi)
        r += amplitude / bumps * np.sin(f * (dirs @ k) * np.pi + phase)
    pos = dirs * r[..., None]
    tris = []
    for i in range(stacks):
        for j in range(slices):
            a, b = pos[i, j], pos[i + 1, j]
            c, d = pos[i + 1, j + 1], pos[i, j + 1]
            # winding chosen so the engine's geometric normal
            # cross(v0 - v1, v0 - v2) points outwards
            if i != 0:
                tris.append(np.concatenate([a, d, b]))
            if i != stacks - 1:
                tris.append(np.concatenate([b, d, c]))
    verts = np.asarray(tris, dtype=np.float32)
    # smooth normals: average of the face normals around each lattice vertex
    # is approximated by the displaced sphere direction's gradient-free proxy
    nrm = verts.reshape(-1, 3, 3).astype(np.float64)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    return verts, nrm.reshape(-1, 9).astype(np.float32)


def random_affine(g, n, scale=(0.5, 3.0), extent=100.0, height=20.0):
    """(n, 3, 4) object-to-world matrices: uniform scale, uniform rotation,
    translation over [-extent, extent] in x/z and [-height, height] in y."""
    R = _random_rotations(g, n)
    s = g.uniform(scale[0], scale[1], n)
    t = np.column_stack([g.uniform(-extent, extent, n), g.uniform(-height, height, n), g.uniform(-extent, extent, n)])
    M = np.zeros((n, 3, 4))
    M[:, :, :3] = R * s[:, None, None]
    M[:, :, 3] = t
    return M


def c2_scene(n_meshes=8, n_instances=256, seed_salt=0):
    """Two-level scene: n_meshes blob meshes, n_instances random instances,
    per-mesh Lambert+GGX materials (albedo U[0.1,0.9], roughness U[0.05,1])."""
    g = rng(2, seed_salt)
    meshes = [blob_mesh(g) for _ in range(n_meshes)]
    mats = {
        "albedo": g.uniform(0.1, 0.9, (n_meshes, 3)).astype(np.float32),
        "roughness": g.uniform(0.05, 1.0, n_meshes).astype(np.float32),
        "metalness": np.zeros(n_meshes, np.float32),
    }
    xf = random_affine(g, n_instances)
    which = g.integers(0, n_meshes, n_instances)
    return meshes, mats, xf, which


def c4_scene(seed_salt=0, extent=15.0, height=15.0):
    """Radiance-cache workload: config 2's meshes, materials and 256
    instances, with the translations drawn over a 30 x 30 x 30 box instead of
    200 x 40 x 200 so that light inter-reflects between the blobs (in config
    2's layout nearly every second-bounce ray escapes to the sky and there is
    no path tail for a cache to replace)."""
    g = rng(2, seed_salt)
    meshes = [blob_mesh(g) for _ in range(8)]
    mats = {
        "albedo": g.uniform(0.1, 0.9, (8, 3)).astype(np.float32),
        "roughness": g.uniform(0.05, 1.0, 8).astype(np.float32),
        "metalness": np.zeros(8, np.float32),
    }
    xf = random_affine(g, 256, extent=extent, height=height)
    which = g.integers(0, 8, 256)
    return meshes, mats, xf, which


def c4_camera(width=1280, height=720, eye=(0.0, 0.0, -17.0), pitch_deg=0.0):
    """Eye at the edge of the dense instance cluster, looking in."""
    return dict(eye=eye, yaw_deg=0.0, pitch_deg=pitch_deg, fov_y_deg=55.0, width=width, height=height)


C2_SUN_DIR = np.array([-0.3, -1.0, -0.2]) / np.linalg.norm([-0.3, -1.0, -0.2])  # direction light travels


def c2_camera(width=1280, height=720):
    """Eye outside the instance cloud looking at its centre."""
    return dict(eye=(0.0, 60.0, -260.0), yaw_deg=0.0, pitch_deg=12.0, fov_y_deg=55.0, width=width, height=height)


def c1(grid=1024, width=1920, height=1080):
    """Coherent primary rays over a heightfield (config 1): (verts, primary rays)."""
    verts = heightfield(grid=grid)
    span = (grid - 1) * 1.0
    eye = (0.5 * span, 120.0, -0.15 * span)
    rays = pinhole_rays(width, height, eye, yaw_deg=0.0, pitch_deg=30.0)
    return verts, rays


def shadow_rays(primary, hits, verts, light_dir=C1_LIGHT_DIR, offset=1e-3, tmin=1e-4, tmax=INF):
    """One any-hit ray per primary hit toward a directional light.

    The origin is the hit point pushed `offset` along the geometric normal
    facing the light; rays for misses get NaN directions (the traversal's
    isnan3 early-out, bvh.hsh:361) so the batch keeps one ray per pixel.
    """
    n = len(primary)
    out = np.zeros(n, dtype=RAY_DTYPE)
    hit = hits["prim_id"] >= 0
    o = primary["origin"].astype(np.float64)
    d = primary["direction"].astype(np.float64)
    p = o + d * hits["t"].astype(np.float64)[:, None]
    tri = verts.reshape(-1, 3, 3).astype(np.float64)
    idx = np.where(hit, hits["prim_id"], 0)
    e0 = tri[idx, 1] - tri[idx, 0]
    e1 = tri[idx, 2] - tri[idx, 0]
    nrm = np.cross(e0, e1)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    L = np.asarray(light_dir, dtype=np.float64)
    nrm *= np.where((nrm @ L) < 0, -1.0, 1.0)[:, None]
    out["origin"] = (p + offset * nrm).astype(np.float32)
    dirs = np.broadcast_to(L, (n, 3)).astype(np.float32).copy()
    dirs[~hit] = np.nan
    out["direction"] = dirs
    out["tmin"] = np.float32(tmin)
    out["tmax"] = np.float32(tmax)
    return out


# ---------------------------------------------------------------------------
# Config 3: batched uneven workloads
# ---------------------------------------------------------------------------
def c3(n_batches=512, tri_range=(1024, 65536), ray_range=(4096, 262144), seed_salt=0):
    """512 independent triangle soups (sizes log-uniform in tri_range, built
    as config 0) each with its own random ray batch (log-uniform in
    ray_range).  Returns (list of (n,9) meshes, concatenated rays, batch
    table as (blas, ray_offset, ray_count) rows)."""
    g = rng(3, seed_salt)
    lt = np.log2(tri_range)
    lr = np.log2(ray_range)
    sizes = np.round(2.0 ** g.uniform(lt[0], lt[1], n_batches)).astype(np.int64)
    counts = np.round(2.0 ** g.uniform(lr[0], lr[1], n_batches)).astype(np.int64)
    meshes = [triangle_soup(int(s), g) for s in sizes]
    rays = random_rays(int(counts.sum()), g)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    table = np.stack([np.arange(n_batches), offs, counts], axis=1).astype(np.uint32)
    return meshes, rays, table


def c2_materials(mats):
    """atlas_rt_material records for the config-2 blob materials (Lambert +
    GGX: metalness 0, reflectance 0.5)."""
    from ._abi import MATERIAL_DTYPE
    n = len(mats["roughness"])
    out = np.zeros(n, dtype=MATERIAL_DTYPE)
    out["base_color"] = mats["albedo"]
    out["opacity"] = 1.0
    out["roughness"] = mats["roughness"]
    out["metalness"] = mats.get("metalness", np.zeros(n, np.float32))
    out["ao"] = 1.0
    out["reflectance"] = 0.5
    out["normal_scale"] = 1.0
    for f in ("base_color_texture", "opacity_texture", "roughness_texture", "metalness_texture", "ao_texture",
              "normal_texture"):
        out[f] = -1
    return out


def sun_light(direction=C2_SUN_DIR, radiance=(3.0, 2.9, 2.7)):
    from ._abi import LIGHT_DTYPE
    L = np.zeros(1, dtype=LIGHT_DTYPE)
    L["direction"] = np.asarray(direction, dtype=np.float32)
    L["radiance"] = np.asarray(radiance, dtype=np.float32)
    return L


C2_ENV = (0.25, 0.3, 0.4)
