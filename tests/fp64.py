"""FP64 recompute of hit (t, u, v) for accuracy reporting.

For each hit the intersection of the ray with the hit triangle's plane is
recomputed in float64 from the same float32 inputs (the formula of
intersections.hsh:36-58 evaluated in double):
  rel_t = |t - t64| / |t64|,  err_u = |u - u64| / max(|u64|, 1), same for v.
Entries above 1e-5 come from near-degenerate determinants |q.e0| (grazing
rays); they are counted, and the conditioning |q.e0| / (|q||e0|) is returned
for them.
"""
import numpy as np


def recompute(verts, rays, prim):
    tri = verts.reshape(-1, 3, 3).astype(np.float64)[prim]
    o = rays["origin"].astype(np.float64)
    d = rays["direction"].astype(np.float64)
    v0, v1, v2 = tri[:, 0], tri[:, 1], tri[:, 2]
    e0, e1, s = v1 - v0, v2 - v0, o - v0
    p = np.cross(s, e0)
    q = np.cross(d, e1)
    den = np.einsum("ij,ij->i", q, e0)
    t = np.einsum("ij,ij->i", p, e1) / den
    u = np.einsum("ij,ij->i", q, s) / den
    v = np.einsum("ij,ij->i", p, d) / den
    cond = np.abs(den) / np.maximum(np.linalg.norm(q, axis=1) * np.linalg.norm(e0, axis=1), 1e-300)
    return t, u, v, cond


def tuv_errors(verts, rays, hits, tol=1e-5):
    m = hits["prim_id"] >= 0
    prim = hits["prim_id"][m]
    t64, u64, v64, cond = recompute(verts, rays[m], prim)
    rt = np.abs(hits["t"][m].astype(np.float64) - t64) / np.abs(t64)
    ru = np.abs(hits["u"][m].astype(np.float64) - u64) / np.maximum(np.abs(u64), 1.0)
    rv = np.abs(hits["v"][m].astype(np.float64) - v64) / np.maximum(np.abs(v64), 1.0)
    worst = np.maximum(rt, np.maximum(ru, rv))
    over = worst > tol
    return {
        "n_hits": int(m.sum()),
        "max_rel_t": float(rt.max()) if len(rt) else 0.0,
        "max_err_u": float(ru.max()) if len(ru) else 0.0,
        "max_err_v": float(rv.max()) if len(rv) else 0.0,
        "p9999": float(np.quantile(worst, 0.9999)) if len(worst) else 0.0,
        "n_over": int(over.sum()),
        "min_cond_over": float(cond[over].min()) if over.any() else None,
    }
