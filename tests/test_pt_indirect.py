"""Statistical pin of the oracle's indirect bounce (CPU only).

EvaluateIndirectLight (rayHit.csh:237-325) chooses refraction with chance
clamp(1 - opacity, 0.1, 0.9), else a specular (VNDF, brdfSample.hsh:36-90) or
diffuse (cosine, :8-34) lobe with the Fresnel-based chance, weights the
sample by reflectance * NdotL / pdf and plays Russian roulette.  Whatever the
lobe chances and roulette probabilities, the expected throughput after one
bounce is fixed by the BRDF alone.  For a flat plane under a constant
environment E, lit by no light it can see (the sun shines from below) and
bounces = 1, every pixel's expected
radiance is

    E * [(1 - a) + a * (S(V) + (1 - m)^2 * base * D(V))]

with a = opacity, m = metalness,
    S(V) = integral over the hemisphere of F(L.H) * D_GGX(N.H) * Vis(N.V, N.L) * N.L
    D(V) = integral of RenormalizedDisneyDiffuse(N.V, N.L, L.H, r^2) / pi * N.L
(brdf.hsh, brdfEval.hsh; F0 = mix(0.16 reflectance^2, base, m),
F90 = saturate(50 F0 . 0.333), surface.hsh:25-42).  The diffuse lobe's
(1 - m) enters twice: EvaluateDiffuseBRDF applies it and EvaluateIndirectLight
multiplies it in again (rayHit.csh:279).  The integrals are evaluated here in
float64 by quadrature from the BRDF formulas, independently of the oracle's
sampling code; the oracle's mean over independent sample batches must agree
within four standard errors, and the standard error is small enough (under
1% of the value) that a wrong VNDF pdf, a lost lobe weight, a missing or
extra (1 - m), a roulette that divides by the wrong probability or a wrong
refraction weight fails the test.  The dark material keeps the roulette
probability below one at the first bounce (3 x throughput < 1).
"""
import math

import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import workloads as W

ENV = np.array([0.5, 0.45, 0.4])


def _fresnel(f0, f90, c):
    return f0 + (f90 - f0) * (1.0 - c) ** 5


def _albedo_parts(theta_v, base, rough, metal, refl=0.5, n_t=384, n_p=384):
    """(S per colour channel, D) for a view direction at polar angle
    theta_v, by midpoint quadrature (float64)."""
    base = np.asarray(base, np.float64)
    f0 = 0.16 * refl * refl * (1.0 - metal) + base * metal
    f90 = min(max(50.0 * float(np.dot(f0, [0.333, 0.333, 0.333])), 0.0), 1.0)
    alpha = rough * rough
    a2 = alpha * alpha
    lr = max(rough * rough, 1e-5)
    V = np.array([math.sin(theta_v), 0.0, math.cos(theta_v)])  # N = +z
    nv = V[2]
    # specular: integrate over half vectors, dL = 4 (V.H) dH
    th = (np.arange(n_t) + 0.5) * (0.5 * math.pi / n_t)
    ph = (np.arange(n_p) + 0.5) * (2.0 * math.pi / n_p)
    T, P = np.meshgrid(th, ph, indexing="ij")
    H = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)
    dH = np.sin(T) * (0.5 * math.pi / n_t) * (2.0 * math.pi / n_p)
    vh = H @ V
    L = 2.0 * vh[..., None] * H - V
    nl = L[..., 2]
    ok = (vh > 0) & (nl > 0)
    nh = H[..., 2]
    D = a2 / (math.pi * ((nh * nh) * (a2 - 1.0) + 1.0) ** 2)
    ggxl = nv * np.sqrt(nl * nl * (1.0 - a2) + a2)
    ggxv = nl * np.sqrt(nv * nv * (1.0 - a2) + a2)
    vis = 0.5 / (ggxl + ggxv)
    w = np.where(ok, D * vis * nl * 4.0 * vh * dH, 0.0)
    spec = np.array([np.sum(w * _fresnel(f0[c], f90, np.clip(vh, 0, 1))) for c in range(3)])
    # diffuse: integrate over light directions (the same grid)
    Ld = H
    nl = Ld[..., 2]
    Hd = Ld + V
    Hd /= np.linalg.norm(Hd, axis=-1, keepdims=True)
    lh = np.clip(np.sum(Ld * Hd, axis=-1), 0, 1)
    fd90 = 0.5 * lr + 2.0 * lh * lh * lr
    ls = 1.0 + (fd90 - 1.0) * (1.0 - nl) ** 5
    vs = 1.0 + (fd90 - 1.0) * (1.0 - nv) ** 5
    ef = 1.0 + (1.0 / 1.51 - 1.0) * lr
    diff = float(np.sum(ls * vs * ef / math.pi * nl * dH))
    return spec, diff


def _expected_albedo(theta_v, base, rough, metal):
    """S + (1 - m)^2 base D per colour channel."""
    spec, diff = _albedo_parts(theta_v, base, rough, metal)
    return spec + (1.0 - metal) ** 2 * np.asarray(base, np.float64) * diff


def _scene(mat):
    v, n = pt_cases.quad_mesh(200.0)
    mats = W.c2_materials({"albedo": np.array([mat["base"]], np.float32),
                           "roughness": np.array([mat["rough"]], np.float32),
                           "metalness": np.array([mat.get("metal", 0.0)], np.float32)})
    mats["opacity"] = mat.get("opacity", 1.0)
    xf = np.zeros((1, 3, 4))
    xf[0, :, :3] = np.eye(3)
    return pt_cases.oracle_scene([v], [n], mats, xf, np.zeros(1, np.int64)), mats


MATERIALS = {
    "rough dielectric": {"base": (0.8, 0.5, 0.2), "rough": 0.9},
    "glossy dielectric": {"base": (0.6, 0.6, 0.6), "rough": 0.35},
    "half metal": {"base": (0.9, 0.6, 0.3), "rough": 0.5, "metal": 0.5},
    "dark (roulette at bounce 0)": {"base": (0.05, 0.05, 0.05), "rough": 0.7},
    "translucent": {"base": (0.7, 0.7, 0.5), "rough": 0.6, "opacity": 0.6},
}


@pytest.mark.parametrize("name", list(MATERIALS))
def test_oracle_one_bounce_mean_equals_quadrature(name):
    mat = MATERIALS[name]
    S, mats = _scene(mat)
    w = h = 12
    pitch = 40.0
    cam = atlas_rt.camera((0.0, 10.0, -10.0 / math.tan(math.radians(pitch))), 0.0, pitch, 2.0, 1.0)
    # the sun shines from below the plane: N.L = 0, no direct light
    sun = W.sun_light(direction=(0.0, 1.0, 0.0), radiance=(1.0, 1.0, 1.0))
    frames, batches, bounces = 256, 8, 1
    means = []
    for k in range(batches):
        acc = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                                  ENV.astype(np.float32), cam, w, h, bounces, frames, k * frames,
                                  k * frames * (bounces + 1))
        assert np.all(acc[..., 3] == frames)
        means.append(acc[..., :3].astype(np.float64) / frames)
    means = np.array(means)  # (batches, h, w, 3)
    # view angle per pixel centre (the jitter averages to it to second order)
    ox, oy, oz = np.array(cam.origin), np.array(cam.right), np.array(cam.bottom)
    loc = np.array(cam.location)
    u = (np.arange(w) + 0.5) / w
    vv = (np.arange(h) + 0.5) / h
    P = ox[None, None] + u[None, :, None] * oy[None, None] + vv[:, None, None] * oz[None, None]
    d = P - loc
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    theta = np.arccos(np.clip(-d[..., 1], -1, 1))  # angle between V = -d and N = +y
    t_nodes = np.linspace(theta.min(), theta.max(), 5)
    alb = np.array([_expected_albedo(t, mat["base"], mat["rough"], mat.get("metal", 0.0)) for t in t_nodes])
    A = np.stack([np.interp(theta, t_nodes, alb[:, c]) for c in range(3)], axis=-1)
    a = mat.get("opacity", 1.0)
    expect = ENV * ((1.0 - a) + a * A)
    dev = (means - expect[None]).mean(axis=(1, 2))  # per batch, per channel
    mean_dev = dev.mean(axis=0)
    se = dev.std(axis=0, ddof=1) / math.sqrt(batches)
    print(name, "expected", expect.mean(axis=(0, 1)), "deviation", mean_dev, "standard error", se)
    assert np.all(se < 0.01 * expect.mean(axis=(0, 1))), "too few samples to pin anything"
    assert np.all(np.abs(mean_dev) <= 4.0 * se + 1e-6), (mean_dev, se)


def test_quadrature_closed_forms():
    """The quadrature against closed forms of the same integrals.
    Roughness 1 (alpha = 1): D = 1/pi and Vis = 0.5 / (N.V + N.L), so with
    F0 = F90 = 1 (metal, white) S = 1 - mu ln(1 + 1/mu), mu = N.V.  Head-on
    (N.V = 1, L.H^2 = (1 + N.L) / 2) the renormalized Disney diffuse at
    roughness 1 integrates to (2 / 1.51) (1/2 + B(2,6) / 2 + B(3,6)) =
    (2 / 1.51) (1/2 + 1/56)."""
    for th in (0.0, 0.3, 1.0):
        mu = math.cos(th)
        spec, _ = _albedo_parts(th, (1.0, 1.0, 1.0), 1.0, 1.0)
        assert np.allclose(spec, 1.0 - mu * math.log(1.0 + 1.0 / mu), rtol=2e-3), (th, spec)
    _, diff = _albedo_parts(0.0, (1.0, 1.0, 1.0), 1.0, 0.0)
    assert abs(diff - (2.0 / 1.51) * (0.5 + 1.0 / 56.0)) < 2e-3 * diff, diff
    # smooth metal: nearly all energy comes back (the GGX albedo tends to 1)
    spec, _ = _albedo_parts(0.3, (1.0, 1.0, 1.0), 0.2, 1.0, n_t=1536, n_p=768)
    assert np.all(np.abs(spec - 1.0) < 0.01), spec


def test_oracle_sincos_and_pow5_are_accurate():
    """The restatement's sine and cosine of sample angles (Cephes' method)
    against float64 sin / cos over [0, 2 pi], within 2 ulp of the result plus
    1e-7 absolute (near the zeros the reduction's absolute error dominates),
    and pow(x, 5) as (x^2)^2 x within 3 ulp (three roundings) on [0, 1] (1e-37 absolute in the
    subnormal range); the octant signs are
    checked at the quarter turns."""
    x = np.linspace(0.0, 2.0 * np.pi, 200001, dtype=np.float32)
    sn, cs, p5 = oracle_rt.sincos(x)
    xd = x.astype(np.float64)
    for got, ref in ((sn, np.sin(xd)), (cs, np.cos(xd))):
        ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - ref)
        assert np.all(err <= 2.0 * ulp + 1e-7), (err.max(), x[np.argmax(err - 2.0 * ulp)])
    q = np.array([0.0, np.pi / 2, np.pi, 1.5 * np.pi], np.float32)
    sq, cq, _ = oracle_rt.sincos(q)
    assert np.allclose(sq, [0, 1, 0, -1], atol=1e-6) and np.allclose(cq, [1, 0, -1, 0], atol=1e-6)
    u = np.linspace(0.0, 1.0, 100001, dtype=np.float32)
    _, _, p5 = oracle_rt.sincos(u)
    ref = u.astype(np.float64) ** 5
    # three roundings; below ~1e-30 the result is subnormal and absolute error rules
    assert np.all(np.abs(p5 - ref) <= 3.0 * np.spacing(ref.astype(np.float32)) + 1e-37)
