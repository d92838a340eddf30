"""Shared inputs for the parity tests: seeded scenes, edge-case rays and
hand-built trees (no GPU needed to construct them)."""
import numpy as np

import atlas_rt
from atlas_rt import workloads as W

RAY = atlas_rt.RAY_DTYPE


def soup(n_tris, n_rays, salt):
    g = W.rng(0, salt)
    return W.triangle_soup(n_tris, g), W.random_rays(n_rays, g)


def make_rays(origins, dirs, tmin=1e-4, tmax=W.INF):
    origins = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
    dirs = np.asarray(dirs, dtype=np.float32).reshape(-1, 3)
    r = np.zeros(len(origins), dtype=RAY)
    r["origin"] = origins
    r["direction"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def edge_rays(verts, n_random=4096, salt=5):
    """Rays that stress the arithmetic: axis-parallel directions (exact zero
    components, infinite reciprocals), rays aimed exactly at vertices and edge
    midpoints (u = 0, v = 0, u + v = 1), origins on triangles, NaN and zero
    directions, tiny tmax and tmin above the hit."""
    g = W.rng(0, 1000 + salt)
    tri = verts.reshape(-1, 3, 3)
    pick = g.integers(0, len(tri), n_random)
    o = g.uniform(-12, 12, (n_random, 3)).astype(np.float32)
    targets = []
    # vertices, edge midpoints, centroids
    targets.append(tri[pick[: n_random // 4], 0])
    targets.append(0.5 * (tri[pick[n_random // 4: n_random // 2], 0] + tri[pick[n_random // 4: n_random // 2], 1]))
    targets.append(0.5 * (tri[pick[n_random // 2: 3 * n_random // 4], 1] + tri[pick[n_random // 2: 3 * n_random // 4], 2]))
    targets.append(tri[pick[3 * n_random // 4:]].mean(axis=1))
    tgt = np.concatenate(targets).astype(np.float64)
    d = tgt - o.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    aimed = make_rays(o, d.astype(np.float32))
    # axis-parallel rays from the box faces
    axis = []
    for k in range(3):
        for s in (-1.0, 1.0):
            oo = g.uniform(-10, 10, (256, 3)).astype(np.float32)
            oo[:, k] = -s * 11.0
            dd = np.zeros((256, 3), np.float32)
            dd[:, k] = s
            axis.append(make_rays(oo, dd))
    # origins exactly on triangle vertices / centroids leaving in random directions
    on = tri[pick[:512]].mean(axis=1).astype(np.float32)
    dd = g.standard_normal((512, 3))
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    onsurf = make_rays(on, dd.astype(np.float32))
    # NaN, zero and denormal directions; tmax limits
    special = make_rays(np.zeros((6, 3)), [[np.nan, 0, 1], [0, 0, 0], [1e-40, 1, 0], [0, 1e-30, 1],
                                            [1, 1, 1], [-1, 0, 0]])
    special["tmax"][4] = 1e-3
    special["tmin"][5] = 1e9
    return np.concatenate([aimed] + axis + [onsurf, special])


def chain_tree(depth):
    """A hand-built BLAS whose traversal stack grows by one per level.

    Node i's left child is node i+1 (a box spanning z in [0, 200]) and its
    right child is a one-triangle leaf in the plane z = 100 + i.  A ray along
    +z enters the left box first, pushes the leaf and descends, so the stack
    reaches `depth` entries before the leaves are popped (deepest first).
    """
    n_nodes = depth
    n_tris = depth + 1
    nodes = np.zeros(n_nodes, dtype=atlas_rt.NODE_DTYPE)
    tris = np.zeros(n_tris, dtype=atlas_rt.GPU_TRI_DTYPE)
    for k in range(n_tris):
        z = 100.0 + k
        tris[k]["v0"] = (-2.0, -2.0, z, 1.0)
        tris[k]["v1"] = (4.0, -2.0, z, 0.0)
        tris[k]["v2"] = (-2.0, 4.0, z, 1.0)
    for i in range(n_nodes):
        last = i == n_nodes - 1
        z = 100.0 + i
        nodes[i]["right_min"] = (-2.0, -2.0, z)
        nodes[i]["right_max"] = (4.0, 4.0, z)
        nodes[i]["right_ptr"] = ~i
        if last:
            nodes[i]["left_min"] = (-2.0, -2.0, z + 1)
            nodes[i]["left_max"] = (4.0, 4.0, z + 1)
            nodes[i]["left_ptr"] = ~(i + 1)
        else:
            nodes[i]["left_min"] = (-3.0, -3.0, 0.0)
            nodes[i]["left_max"] = (5.0, 5.0, 200.0)
            nodes[i]["left_ptr"] = i + 1
    g = np.random.default_rng(3)
    o = np.column_stack([g.uniform(-1, 1, 1024), g.uniform(-1, 1, 1024), np.full(1024, -5.0)])
    rays = make_rays(o, np.tile([0.0, 0.0, 1.0], (1024, 1)))
    return nodes, tris, rays


def hit_bits(h):
    """(t, u, v, prim) as integers for bit-exact comparison."""
    return np.stack([h["t"].view(np.int32), h["u"].view(np.int32), h["v"].view(np.int32), h["prim_id"]], axis=1)
