"""GPU tests of the learned flux correction, all through the C ABI.

* MLP kernels (BF16 MFMA, FP32 accumulation) against the numpy oracle with the same quantisation:
  forward, gradient, Adam step; training reduces the loss of a teacher-student problem.
* Data pipeline: the recorded fluxes reproduce the state change exactly (flux-form identity), and
  the collected features and targets equal a host recomputation from the downloaded states.
* The corrected FP32 stage: with a zero network it equals the plain FP32 solver; with a random
  network one RK3 step equals the oracle's corrected step, with the same fallback counts.
"""
import numpy as np
import pytest

from gaukuk_amd import problems
from oracle import learned as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from gaukuk_amd import hip as H
    H.lib()
    return H


def _rand_net(seed, bias=0.1):
    rng = np.random.default_rng(seed)
    p = L.init_params(seed)
    for k in ("b1", "b2"):
        p[L.OFF[k]:L.OFF[k] + 64] = rng.normal(0, bias, 64)
    p[L.OFF["b3"]:L.OFF["b3"] + 4] = rng.normal(0, bias, 4)
    return p.astype(np.float32).astype(np.float64)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_mlp_forward_matches_oracle(hip):
    rng = np.random.default_rng(0)
    p = _rand_net(1)
    s = rng.uniform(0.5, 2.0, 16).astype(np.float32)
    x = rng.normal(size=(1000, 16)).astype(np.float32)      # not a multiple of 16: masked tail
    net = hip.Net(p, s, np.ones(4))
    y = net.forward(x)
    ref = L.forward(p, x, s, quant=True)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(y - ref)) < 1e-5 * scale, np.max(np.abs(y - ref)) / scale
    # a transposed operand or a wrong lane layout would not be within bf16 rounding of the truth
    exact = L.forward(p, x, s, quant=False)
    assert np.max(np.abs(y - exact)) < 3e-2 * scale


@pytest.mark.parametrize("n", [1, 100, 4096])
def test_mlp_gradient_matches_oracle(hip, n):
    rng = np.random.default_rng(n)
    p = _rand_net(2)
    s = rng.uniform(0.5, 2.0, 16).astype(np.float32)
    out = np.array([0.5, 1.0, 2.0, 4.0], np.float32)
    x = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=(n, 4)).astype(np.float32)
    wt = rng.uniform(0.0, 1.0, n).astype(np.float32)
    net = hip.Net(p, s, out)
    loss, g = net.gradient(x, t, wt)
    lq, gq = L.backward(p, x, t / out, wt, s, quant=True)
    le, ge = L.backward(p, x, t / out, wt, s, quant=False)
    assert abs(loss - lq) <= 1e-4 * lq
    assert _rel(g, gq) < 2e-3, _rel(g, gq)
    assert _rel(g, ge) < 3e-2, _rel(g, ge)
    # every block of the parameter vector is populated (layout of the gradient write-back)
    for k, (a, b) in zip(L.OFF, zip(L.unpack(g).values(), L.unpack(gq).values())):
        assert _rel(np.asarray(a), np.asarray(b)) < 2e-2, k


def test_adam_matches_oracle(hip):
    rng = np.random.default_rng(3)
    p = _rand_net(3)
    net = hip.Net(p)
    m = np.zeros(L.NPARAM)
    v = np.zeros(L.NPARAM)
    ref = p.copy()
    for step in range(1, 4):
        g = rng.normal(size=L.NPARAM).astype(np.float32)
        net.adam(g, 1e-3)
        ref, m, v = L.adam(ref, m, v, g.astype(np.float64), step, 1e-3)
    got = net.params()
    assert np.max(np.abs(got - ref)) < 1e-6


def test_training_reduces_loss(hip):
    rng = np.random.default_rng(4)
    teacher = _rand_net(40, bias=0.3)
    x = rng.normal(size=(1 << 15, 16)).astype(np.float32)
    t = L.forward(teacher, x, np.ones(16), quant=False).astype(np.float32)
    net = hip.Net(_rand_net(41), np.ones(16), np.ones(4))
    net.reserve(len(x))
    net.add_samples(x, t)
    l0, _ = net.gradient(x[:4096], t[:4096])
    log = net.train(400, 3e-3, blocks=64, tiles_per_wave=1, seed=5)
    l1, _ = net.gradient(x[:4096], t[:4096])
    assert np.all(np.isfinite(log))
    assert l1 < 0.1 * l0, (l0, l1)
    # training is deterministic: same seed, same start -> same parameters
    net2 = hip.Net(_rand_net(41), np.ones(16), np.ones(4))
    net2.reserve(len(x))
    net2.add_samples(x, t)
    net2.train(400, 3e-3, blocks=64, tiles_per_wave=1, seed=5)
    assert np.array_equal(net.params(), net2.params())


def _div(fx, fy, dxr, dyr):
    """Flux-form change of (mass, m_x, m_y, E) from face sums fx [4, ny, nx+1], fy [4, ny+1, nx]."""
    return -(fx[:, :, 1:] - fx[:, :, :-1]) * dxr - (fy[:, 1:, :] - fy[:, :-1, :]) * dyr


@pytest.mark.parametrize("span", [None, 3.5])
def test_flux_record_identity(hip, span):
    prob = problems.kelvin_helmholtz(64, seed=3, integrator=3)
    s = hip.Solver(prob)
    u0 = s.download()
    s.flux_record(True)
    if span is None:
        s.run_steps(1)
    else:
        s.run_steps(1)            # learn dt, then span several steps through advance's clipping
        dt = s.time()[1]
        s.advance(span * dt)
    s.flux_record(False)
    u1 = s.download()
    fx, fy = s.flux_download()
    d = _div(fx, fy, prob.nx, prob.ny)
    for c, n in enumerate((0, 1, 2, 4)):
        err = np.max(np.abs((u1[n] - u0[n]) - d[c]))
        assert err < 1e-12 * np.max(np.abs(u0[n])) + 1e-14, (n, err)


def _restrict(u, f):
    q, ny, nx = u.shape
    return u.reshape(q, ny // f, f, nx // f, f).mean(axis=(2, 4))


def _host_faces(u, gamma):
    """Stencils (rho, v_n, v_t, p) of every face of a periodic grid, in the face order of the
    library: x faces [ny, nx+1] then y faces [ny+1, nx]."""
    w = L._prim(u, gamma)
    ny, nx = u.shape[1:]
    ii = np.arange(nx + 1)
    st_x = np.stack([np.stack([w[0], w[1], w[2], w[4]], -1)[:, (ii - 2 + k) % nx] for k in range(4)], axis=-2)
    jj = np.arange(ny + 1)
    st_y = np.stack([np.stack([w[0], w[2], w[1], w[4]], -1)[(jj - 2 + k) % ny, :] for k in range(4)], axis=-2)
    return st_x.reshape(-1, 4, 4), st_y.reshape(-1, 4, 4)


def test_collect_targets(hip):
    f = 4
    fine_p = problems.kelvin_helmholtz(64, seed=5, integrator=3)
    coarse_p = fine_p.with_(nx=16, ny=16, cons=_restrict(fine_p.cons, f))
    fine = hip.Solver(fine_p)
    coarse = hip.Solver(coarse_p)
    fine.advance(0.05)                       # a developed, non-trivial state
    uf0 = fine.download()
    net = hip.Net()
    nfc = coarse.n_faces()
    net.reserve(2 * nfc)
    D, added = net.collect(fine, coarse)
    assert added == nfc and net.size() == nfc
    uc0 = _restrict(uf0, f)
    uf1, uc1 = fine.download(), coarse.download()
    ffx, ffy = fine.flux_download()
    cfx, cfy = coarse.flux_download()
    # restricted fine face sums over f fine faces per coarse face, over f * D
    Ffx = ffx[:, :, ::f].reshape(4, 16, f, 17).sum(axis=2) / (f * D)
    Ffy = ffy[:, ::f, :].reshape(4, 17, 16, f).sum(axis=3) / (f * D)
    # flux-form identities: the coarse step and the restricted fine run over D
    d_c = _div(cfx, cfy, 16, 16)
    d_f = _div(Ffx * D, Ffy * D, 16, 16)
    for c, n in enumerate((0, 1, 2, 4)):
        sc = np.max(np.abs(uc0[n])) + 1e-300
        assert np.max(np.abs((uc1[n] - uc0[n]) - d_c[c])) < 1e-12 * sc + 1e-14
        assert np.max(np.abs((_restrict(uf1, f)[n] - uc0[n]) - d_f[c])) < 1e-12 * sc + 1e-14
    # targets and features against the host recomputation
    X, T, W = net.dataset()
    stx, sty = _host_faces(uc0, 1.4)
    st = np.concatenate([stx, sty])
    labx = (Ffx - cfx / D).reshape(4, -1).T
    laby = (Ffy - cfy / D).reshape(4, -1).T
    lab = np.concatenate([labx[:, [0, 1, 2, 3]], laby[:, [0, 2, 1, 3]]])   # (mass, m_n, m_t, E)
    feats = L.face_features(st, 1.4)
    # the bound is formed from the stored (float32) features, as on the device
    bound = net.kappa * L.dissipation_bound(X.astype(np.float64), 1.4)
    raw = L.to_network_units(lab, st, 1.4)
    tgt = np.clip(raw, -bound, bound)
    live = W > 0
    assert live.sum() == 2 * 16 * 16           # the wrap faces are listed twice and kept once
    assert np.max(np.abs(X - feats)) < 1e-6 * np.max(np.abs(feats)) + 1e-7
    scale = np.max(np.abs(tgt[live]), axis=0)
    assert np.all(np.max(np.abs(T[live] - tgt[live]), axis=0) < 1e-5 * scale + 1e-12)
    # fine and coarse fluxes differ (the target is not trivially zero), and the bound is active on
    # part of the faces only
    assert np.all(scale > 0)
    clipped = np.mean(np.abs(raw[live]) > bound[live])
    assert 0.0 < clipped < 0.9, clipped


def test_corrected_stage_zero_net_equals_plain(hip):
    prob = problems.kelvin_helmholtz(64, seed=6, integrator=3)
    plain = hip.Solver(prob, precision=32)
    corr = hip.Solver(prob, precision=32)
    net = hip.Net(np.zeros(L.NPARAM), np.ones(16), np.ones(4))
    corr.set_net(net)
    corr.net_stats(reset=True)
    plain.run_steps(10)
    corr.run_steps(10)
    a, b = plain.download(), corr.download()
    for n in (0, 1, 2, 4):
        assert np.sum(np.abs(a[n] - b[n])) / np.sum(np.abs(a[n])) < 1e-6
    corrected, fired, rejected = corr.net_stats()
    assert fired == 0 and rejected == 0
    assert corrected == 10 * 3 * 2 * 64 * 64


def test_attached_net_rejects_unsupported_setters(hip):
    """The corrected stage has no gravity kick and no saved-data boundaries: with a network
    attached, enrolling gravity, fixed data or a shell fails instead of being silently dropped;
    after detaching, the same calls succeed."""
    prob = problems.kelvin_helmholtz(64, seed=6, integrator=3)
    s = hip.Solver(prob, precision=32)
    net = hip.Net(np.zeros(L.NPARAM), np.ones(16), np.ones(4))
    s.set_net(net)
    with pytest.raises(hip.GaukukError, match="GK_EINVAL"):
        s._check(s.L.gk_set_point_gravity(s.h, 1.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.1))
    with pytest.raises(hip.GaukukError, match="GK_EINVAL"):
        s.set_fixed_boundary(np.zeros(s._ref_shape()))
    with pytest.raises(hip.GaukukError, match="GK_EINVAL"):
        s._check(s.L.gk_set_immersed_shell(s.h, 0.5, 0.5, 0.0, 0.1))
    s.run_steps(2)   # still runnable as configured
    s.synchronize()
    s.set_net(None)
    s._check(s.L.gk_set_point_gravity(s.h, 1.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.1))
    # a grid that gained gravity cannot take the network back
    with pytest.raises(hip.GaukukError, match="GK_EINVAL"):
        s.set_net(net)


@pytest.mark.parametrize("kappa", [0.5, 4.0])
def test_corrected_step_matches_oracle(hip, kappa):
    n = 32
    prob = problems.kelvin_helmholtz(n, seed=7, integrator=3)
    u0 = prob.cons.copy()
    # a pressure bump makes the shock sensor fire around it
    u0[4, 10:14, 20:24] *= 1.6
    prob = prob.with_(cons=u0)
    p = _rand_net(8)
    in_s = np.full(16, 2.0, np.float32)
    out_s = np.array([0.03, 0.03, 0.03, 0.03], np.float32)
    s = hip.Solver(prob, precision=32)
    net = hip.Net(p, in_s, out_s)
    net.kappa = kappa
    s.set_net(net)
    s.net_stats(reset=True)
    s.run_steps(1)
    got = s.download()
    dt = s.time()[1]
    u32 = u0.astype(np.float32).astype(np.float64)
    netd = (p, in_s, out_s, kappa)
    f13 = 1.0 / 3.0
    u1, f1, r1 = L.stage_update(u32, None, (1.0, 1.0), dt, 1 / n, 1 / n, 1.4, netd)
    u2, f2, r2 = L.stage_update(u1, u32, (0.75, 0.25, 0.25), dt, 1 / n, 1 / n, 1.4, netd)
    u3, f3, r3 = L.stage_update(u2, u32, (f13, 2 * f13, 2 * f13), dt, 1 / n, 1 / n, 1.4, netd)
    plain = L.rk3_step(u32, dt, 1 / n, 1 / n, 1.4)
    for q in (0, 1, 2, 4):
        den = np.sum(np.abs(u3[q]))
        err = np.sum(np.abs(got[q] - u3[q])) / den
        effect = np.sum(np.abs(plain[q] - u3[q])) / den
        assert err < 3e-5, (q, err)
        assert effect > 30 * err, (q, effect, err)   # the correction is visible above the tolerance
    corrected, fired, rejected = s.net_stats()
    assert abs(fired - (f1 + f2 + f3)) <= 2
    assert fired > 0
    assert rejected == r1 + r2 + r3
    assert corrected + fired + rejected == 3 * 2 * n * n
