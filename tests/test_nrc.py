"""Radiance cache: closed-form and finite-difference pins of the FP64
oracle (CPU), and the GPU kernels (fused encoding + BF16 MFMA MLP forward,
backward, Adam) against it.

The engine has no radiance cache, so the FP64 oracle is the only reference here."""
import math

import numpy as np
import pytest

import nrc_oracle as O


def rand_queries(g, n, lo=-1.2, hi=1.2):
    q = np.zeros((n, 12), np.float32)
    q[:, 0:3] = g.uniform(lo, hi, (n, 3))
    q[:, 3] = g.uniform(0.0, 1.0, n)
    d = g.normal(size=(n, 3))
    q[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    m = g.normal(size=(n, 3))
    q[:, 8:11] = m / np.linalg.norm(m, axis=1, keepdims=True)
    return q


def rand_master(g, scale=1.0):
    w = np.zeros(O.MASTER_SIZE, np.float32)
    for l in range(6):
        lim = scale * math.sqrt(6.0 / (O.K[l] + O.N[l]))
        w[O.MASTER_OFFSETS[l]:O.MASTER_OFFSETS[l + 1]] = g.uniform(-lim, lim, O.K[l] * O.N[l])
    return w


def rand_table(g, P, amp=1.0):
    return g.uniform(-amp, amp, (P.levels, P.T, 2)).astype(np.float16)


# ---------------------------------------------------------------------------
# Oracle pins (CPU)
# ---------------------------------------------------------------------------

def test_level_resolutions_closed_form():
    r = O.level_resolutions(12, 16, 2048)
    assert r[0] == 16 and r[-1] == 2048
    b = 2048 / 16
    for l in range(12):
        assert r[l] == math.floor(16 * b ** (l / 11) + 1e-9)
    assert np.all(np.diff(r) > 0)
    assert list(O.level_resolutions(1, 7, 7)) == [7]


def test_encode_reproduces_linear_function_on_dense_level():
    """Trilinear interpolation of a function linear in the lattice
    coordinates is that function: feature = a sx + b sy + c sz."""
    P = O.Params([-1, -1, -1], [1, 1, 1], levels=1, log2_table=10, base=8, maxres=8)
    assert P.dense[0] and P.res[0] == 8
    R1 = 9
    tab = np.zeros((1, P.T, 2))
    a, b, c = 0.3, -1.1, 0.7
    for z in range(R1):
        for y in range(R1):
            for x in range(R1):
                tab[0, x + R1 * (y + R1 * z), 0] = a * x + b * y + c * z
                tab[0, x + R1 * (y + R1 * z), 1] = 1.0
    g = np.random.default_rng(1)
    q = rand_queries(g, 500, -1.0, 1.0)
    X, _ = O.encode(P, tab, q)
    s = (((q[:, :3] - P.bmin) * P.binv).astype(np.float32) * np.float32(8)).astype(np.float64)
    expect = a * s[:, 0] + b * s[:, 1] + c * s[:, 2]
    assert np.allclose(X[:, 0], expect, atol=1e-9)
    assert np.allclose(X[:, 1], 1.0, atol=1e-12)  # partition of unity
    # direction, normal, roughness and the constant input
    assert np.array_equal(X[:, 24:27], q[:, 4:7].astype(np.float64))
    assert np.array_equal(X[:, 27:30], q[:, 8:11].astype(np.float64))
    assert np.array_equal(X[:, 30], q[:, 3].astype(np.float64)) and np.all(X[:, 31] == 1.0)
    assert np.all(X[:, 2:24] == 0.0)


def test_encode_lattice_vertex_reads_hashed_entry():
    """On a lattice vertex the feature is the entry of that vertex; the hash
    is (x * 1 ^ y * 2654435761 ^ z * 805459861) mod 2^32 mod T."""
    P = O.Params([0, 0, 0], [1, 1, 1], levels=1, log2_table=10, base=64, maxres=64)
    assert not P.dense[0]
    g = np.random.default_rng(2)
    tab = g.uniform(-1, 1, (1, P.T, 2))
    k = g.integers(0, 65, (300, 3))
    q = np.zeros((300, 12), np.float32)
    q[:, :3] = k / 64.0
    X, _ = O.encode(P, tab, q)
    for i in range(300):
        x, y, z = (int(v) for v in k[i])
        h = ((x ^ ((y * 2654435761) % 2 ** 32) ^ ((z * 805459861) % 2 ** 32)) % 2 ** 32) & (P.T - 1)
        assert X[i, 0] == pytest.approx(tab[0, h, 0], abs=1e-12)
        assert X[i, 1] == pytest.approx(tab[0, h, 1], abs=1e-12)


def test_forward_single_path_closed_form():
    """One chain of non-zero weights from the constant input to output 1:
    y1 = product of the weights while every partial product stays positive,
    else 0 (ReLU); weights are stored input-major ([K][N])."""
    P = O.Params([0, 0, 0], [1, 1, 1], levels=2, log2_table=10, base=4, maxres=8)
    chain = [(31, 5), (5, 7), (7, 0), (0, 63), (63, 2), (2, 1)]
    q = rand_queries(np.random.default_rng(3), 7, 0, 1)
    tab = np.zeros((2, P.T, 2))
    for vals, expect in (([0.5, 2.0, 1.5, 0.8, 1.25, -3.0], 0.5 * 2.0 * 1.5 * 0.8 * 1.25 * -3.0),
                         ([0.5, -2.0, 1.5, 0.8, 1.25, -3.0], 0.0)):
        m = np.zeros(O.MASTER_SIZE)
        for l, ((i, j), v) in enumerate(zip(chain, vals)):
            m[O.MASTER_OFFSETS[l] + i * O.N[l] + j] = v
        y = O.predict(P, m, tab, q)
        assert np.allclose(y[:, 1], expect, rtol=1e-15) and np.all(y[:, [0, 2]] == 0)


def test_gradients_match_finite_differences():
    """Analytic FP64 gradients (stop-gradient denominator) against central
    differences of the loss with that denominator frozen.  The loss sums
    ~100 terms of order 1, so its rounding noise (~1e-14) over 2h bounds the
    absolute tolerance."""
    g = np.random.default_rng(4)
    P = O.Params([-1, -1, -1], [1, 1, 1], levels=3, log2_table=10, base=4, maxres=24)
    m = rand_master(g).astype(np.float64)
    tab = g.uniform(-0.5, 0.5, (P.levels, P.T, 2))
    q = rand_queries(g, 40, -1.1, 1.1)
    t = np.abs(g.normal(size=(40, 4)))
    L, gm, gt, y = O.loss_and_grads(P, m, tab, q, t)
    den = y * y + O.LOSS_FLOOR
    h = 1e-5

    def lossf(mm, tt):
        return O.loss_and_grads(P, mm, tt, q, t, frozen_den=den)[0]

    for l in range(6):
        for j in g.choice(O.K[l] * O.N[l], 5, replace=False):
            k = O.MASTER_OFFSETS[l] + j
            mp, mn = m.copy(), m.copy()
            mp[k] += h
            mn[k] -= h
            fd = (lossf(mp, tab) - lossf(mn, tab)) / (2 * h)
            assert fd == pytest.approx(gm[k], rel=1e-4, abs=3e-9), (l, j)
    touched = np.argwhere(gt != 0)
    assert len(touched) > 100
    for lev, e, c in touched[g.choice(len(touched), 30, replace=False)]:
        tp, tn = tab.copy(), tab.copy()
        tp[lev, e, c] += h
        tn[lev, e, c] -= h
        fd = (lossf(m, tp) - lossf(m, tn)) / (2 * h)
        assert fd == pytest.approx(gt[lev, e, c], rel=1e-4, abs=3e-9)


def test_quantized_mode_stays_close_to_fp64():
    """BF16 operand rounding perturbs the outputs and weight gradients by a
    few per cent; the input (table) gradients move more (~10%) because
    pre-activations near zero change sign and flip their ReLU gates."""
    g = np.random.default_rng(5)
    P = O.Params([-1, -1, -1], [1, 1, 1], levels=4, log2_table=11, base=4, maxres=32)
    m = rand_master(g)
    tab = rand_table(g, P).astype(np.float64)
    q = rand_queries(g, 256)
    t = np.abs(g.normal(size=(256, 4)))
    L0, gm0, gt0, y0 = O.loss_and_grads(P, m, tab, q, t)
    L1, gm1, gt1, y1 = O.loss_and_grads(P, m, tab, q, t, quantize=True)
    assert np.linalg.norm(y1 - y0) <= 0.03 * np.linalg.norm(y0)
    assert np.linalg.norm(gm1 - gm0) <= 0.05 * np.linalg.norm(gm0)
    assert np.linalg.norm(gt1 - gt0) <= 0.2 * np.linalg.norm(gt0)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

GP = dict(bounds_min=(-1.0, -1.0, -1.0), bounds_max=(1.0, 1.0, 1.0), levels=12, log2_table=14, base_resolution=4,
          max_resolution=256)


def _cache(torch, seed=7, lr=1e-2):
    import atlas_rt
    return atlas_rt.RadianceCache(learning_rate=lr, seed=seed, **GP)


def _oracle_params():
    return O.Params(GP["bounds_min"], GP["bounds_max"], GP["levels"], GP["log2_table"], GP["base_resolution"],
                    GP["max_resolution"])


def _set_random(torch, rc, g, P, scale=1.0, amp=1.0):
    m = rand_master(g, scale)
    t16 = rand_table(g, P, amp)
    rc.set_params(torch.from_numpy(m).cuda(), torch.from_numpy(t16.view(np.int16).reshape(-1).copy()).cuda())
    m_back, t_back = rc.get_params()
    assert np.array_equal(m_back.cpu().numpy(), m)
    assert np.array_equal(t_back.cpu().numpy().view(np.float16).reshape(t16.shape), t16)
    return m, t16.astype(np.float64)


@pytest.mark.gpu
def test_gpu_nrc_infer_matches_oracle(torch_cuda):
    print("case nrc infer", flush=True)
    torch = torch_cuda
    g = np.random.default_rng(11)
    P = _oracle_params()
    rc = _cache(torch)
    m, tab = _set_random(torch, rc, g, P)
    for n in (1, 63, 5000):
        q = rand_queries(g, n)
        q[: n // 10, 0] = np.nan if n > 10 else q[: n // 10, 0]  # NaN positions clamp to the box corner
        out = rc.infer(torch.from_numpy(q).cuda()).cpu().numpy()
        yq = O.predict(P, m, tab, q, quantize=True)
        y64 = O.predict(P, m, tab, q)
        scale = np.abs(y64).max()
        err_q = np.abs(out[:, :3] - yq).max() / scale
        err_64 = np.linalg.norm(out[:, :3] - y64) / np.linalg.norm(y64)
        print(f"nrc infer n={n}: max err vs bf16-operand oracle {err_q:.2e}, rel L2 vs FP64 {err_64:.2e}")
        assert np.all(out[:, 3] == 0)
        assert err_q < 2e-3
        assert err_64 < 3e-2


@pytest.mark.gpu
def test_gpu_nrc_gradients_match_oracle(torch_cuda):
    print("case nrc gradients", flush=True)
    torch = torch_cuda
    g = np.random.default_rng(12)
    P = _oracle_params()
    rc = _cache(torch)
    m, tab = _set_random(torch, rc, g, P)
    n = 3001
    q = rand_queries(g, n)
    t = np.abs(g.normal(size=(n, 4))).astype(np.float32)
    gm, gt = rc.gradients(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda())
    gm, gt = gm.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64).reshape(P.levels, P.T, 2)
    Lq, gmq, gtq, _ = O.loss_and_grads(P, m, tab, q, t, quantize=True)
    L64, gm64, gt64, _ = O.loss_and_grads(P, m, tab, q, t)
    for l in range(6):
        sl = slice(O.MASTER_OFFSETS[l], O.MASTER_OFFSETS[l + 1])
        eq = np.linalg.norm(gm[sl] - gmq[sl]) / np.linalg.norm(gmq[sl])
        e64 = np.linalg.norm(gm[sl] - gm64[sl]) / np.linalg.norm(gm64[sl])
        print(f"nrc grad layer {l}: rel L2 vs bf16-operand oracle {eq:.2e}, vs FP64 {e64:.2e}")
        assert eq < 1e-2 and e64 < 8e-2
    eq = np.linalg.norm(gt - gtq) / np.linalg.norm(gtq)
    e64 = np.linalg.norm(gt - gt64) / np.linalg.norm(gt64)
    print(f"nrc grad table: rel L2 vs bf16-operand oracle {eq:.2e}, vs FP64 {e64:.2e}")
    assert eq < 1e-2 and e64 < 0.25
    assert np.array_equal(gt != 0, gtq != 0) or (np.sum((gt != 0) != (gtq != 0)) < 1e-3 * gt.size)


@pytest.mark.gpu
def test_gpu_nrc_first_adam_step_closed_form(torch_cuda):
    """From zero moments Adam's first step moves every parameter with a
    non-zero gradient by exactly -lr * sign(g); the loss matches the oracle."""
    print("case nrc adam", flush=True)
    torch = torch_cuda
    g = np.random.default_rng(13)
    P = _oracle_params()
    lr = 1e-3
    rc = _cache(torch, lr=lr)
    m, tab = _set_random(torch, rc, g, P, amp=0.5)
    n = 2048
    q = torch.from_numpy(rand_queries(g, n)).cuda()
    t = torch.from_numpy(np.abs(g.normal(size=(n, 4))).astype(np.float32)).cuda()
    gm, gt = rc.gradients(q, t)
    gm, gt = gm.cpu().numpy(), gt.cpu().numpy()
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    rc.train(q, t, loss)
    m1, t1 = rc.get_params()
    m1 = m1.cpu().numpy().astype(np.float64)
    L = O.loss_and_grads(P, m, tab, q.cpu().numpy(), t.cpu().numpy(), quantize=True)[0]
    assert loss.item() == pytest.approx(L, rel=2e-3)
    big = np.abs(gm) > 1e-4 * np.abs(gm).max()
    step = m1 - m.astype(np.float64)
    assert np.allclose(step[big], -lr * np.sign(gm[big]), rtol=1e-4, atol=2e-7)
    assert np.all(step[gm == 0] == 0)
    t0 = tab.reshape(-1)
    t1 = t1.cpu().numpy().view(np.float16).astype(np.float64)
    sel = np.abs(gt) > 1e-4 * np.abs(gt).max()
    expect = (t0 - lr * np.sign(gt)).astype(np.float16).astype(np.float64)
    ulp = np.abs(np.spacing(expect.astype(np.float16)).astype(np.float64))
    assert np.all(np.abs(t1[sel] - expect[sel]) <= ulp[sel])
    assert np.array_equal(t1[gt == 0], t0[gt == 0])


def smooth_radiance(q):
    """A smooth, view- and normal-dependent test function of the query."""
    p, d, nrm = q[:, 0:3].astype(np.float64), q[:, 4:7].astype(np.float64), q[:, 8:11].astype(np.float64)
    base = 0.6 + 0.4 * np.sin(2.5 * p[:, 0]) * np.cos(1.7 * p[:, 1]) + 0.2 * p[:, 2]
    out = np.stack([base * (0.8 + 0.2 * nrm[:, 1]), base * 0.5 + 0.1 * d[:, 0] ** 2, 0.3 + 0.2 * np.abs(nrm[:, 2])], 1)
    return np.concatenate([np.maximum(out, 0), np.zeros((len(q), 1))], 1).astype(np.float32)


@pytest.mark.gpu
def test_gpu_nrc_learns_smooth_function(torch_cuda):
    print("case nrc learn", flush=True)
    torch = torch_cuda
    g = np.random.default_rng(14)
    rc = _cache(torch, seed=3, lr=1e-2)
    qv = rand_queries(g, 20000, -1, 1)
    tv = smooth_radiance(qv)
    d_qv = torch.from_numpy(qv).cuda()

    def rel_mse():
        y = rc.infer(d_qv).cpu().numpy()[:, :3].astype(np.float64)
        ref = tv[:, :3].astype(np.float64)
        return float(np.mean((y - ref) ** 2 / (ref ** 2 + 0.01)))

    before = rel_mse()
    mean_pred = float(np.mean((tv[:, :3].mean(axis=0) - tv[:, :3]) ** 2 / (tv[:, :3] ** 2 + 0.01)))
    losses = []
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    for it in range(300):
        q = rand_queries(g, 16384, -1, 1)
        rc.train(torch.from_numpy(q).cuda(), torch.from_numpy(smooth_radiance(q)).cuda(), loss)
        if it % 50 == 0 or it == 299:
            losses.append(loss.item())
    after = rel_mse()
    print(f"nrc learn: rel-MSE before {before:.4f}, after {after:.5f}, constant-mean predictor {mean_pred:.4f}, "
          f"losses {losses}")
    assert after < 0.05 * before and after < 0.1 * mean_pred
    assert losses[-1] < 0.1 * losses[0]
