"""CPU checks of the learned-correction oracle (oracle/learned.py).

The reference has no learned component, so nothing in it pins these functions; they are pinned by
mathematics instead: central finite differences for the analytic gradient, known bit patterns for
the BF16 rounding, exact inverses for the frame/scaling maps, the tangential Galilean invariance of
the PLM-HLLC flux (which the feature frame relies on), and Adam's closed-form first step.
"""
import numpy as np
import pytest

from oracle import learned as L


def test_bf16_rounding_known_values():
    one = 1.0
    assert L.bf16(one) == 1.0
    assert L.bf16(1.0 + 2.0 ** -8) == 1.0                      # tie -> even (down)
    assert L.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6      # tie -> even (up)
    assert L.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert L.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).normal(size=10000)
    q = L.bf16(x)
    assert np.all(np.abs(q - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-6))
    # idempotent
    assert np.array_equal(L.bf16(q), q)


def _problem(seed=1, n=37):
    rng = np.random.default_rng(seed)
    p = L.init_params(seed)
    p[L.OFF["b1"]:L.OFF["b1"] + 64] = rng.normal(0, 0.1, 64)
    p[L.OFF["b2"]:L.OFF["b2"] + 64] = rng.normal(0, 0.1, 64)
    x = rng.normal(size=(n, 16))
    t = rng.normal(size=(n, 4))
    wt = rng.uniform(0.5, 1.5, n)
    s = rng.uniform(0.5, 2.0, 16)
    return p, x, t, wt, s


def test_exact_gradient_matches_finite_differences():
    p, x, t, wt, s = _problem()
    _, g = L.backward(p, x, t, wt, s, quant=False)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(np.arange(k0, k0 + n), min(12, n), replace=False)
                          for k0, n in ((0, 1024), (1024, 64), (1088, 4096), (5184, 64), (5248, 256), (5504, 4))])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (L.loss(pp, x, t, wt, s) - L.loss(pm, x, t, wt, s)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_planted_gradient_mistake_is_caught():
    """A dropped ReLU mask in the backward (a plausible mistake) fails the finite-difference test."""
    p, x, t, wt, s = _problem(3)
    y, k = L.forward(p, x, s, quant=False, keep=True)
    n = len(x)
    dY = 2.0 * wt[:, None] * (y - t) / (4 * n)
    dZ2 = dY @ k["W3"]                         # mask omitted
    gW2_wrong = dZ2.T @ k["H1"]
    _, g = L.backward(p, x, t, wt, s, quant=False)
    gW2 = g[L.OFF["W2"]:L.OFF["W2"] + 4096].reshape(64, 64)
    assert np.max(np.abs(gW2_wrong - gW2)) > 1e-3 * np.max(np.abs(gW2))


def test_quantised_gradient_close_to_exact():
    p, x, t, wt, s = _problem(4, 256)
    le, ge = L.backward(p, x, t, wt, s, quant=False)
    lq, gq = L.backward(p, x, t, wt, s, quant=True)
    assert abs(lq - le) < 2e-2 * le
    assert np.linalg.norm(gq - ge) < 3e-2 * np.linalg.norm(ge)


def test_adam_first_step_closed_form():
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p1, m, v = L.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, 0.01)
    # m_hat = g, v_hat = g^2: step = lr * g / (|g| + eps)
    assert np.allclose(p1, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)


def _stencil(rng, n):
    st = np.empty((n, 4, 4))
    st[..., 0] = rng.uniform(0.5, 2.0, (n, 4))
    st[..., 1] = rng.normal(0, 0.5, (n, 4))
    st[..., 2] = rng.normal(0, 0.5, (n, 4))
    st[..., 3] = rng.uniform(0.5, 3.0, (n, 4))
    return st


def test_frame_maps_are_inverse():
    rng = np.random.default_rng(5)
    st = _stencil(rng, 1000)
    d = rng.normal(size=(1000, 4))
    back = L.correction_lab(L.to_network_units(d, st, 1.4), st, 1.4)
    assert np.allclose(back, d, rtol=1e-12, atol=1e-12)


def test_features_and_sensor():
    rng = np.random.default_rng(6)
    st = _stencil(rng, 50)
    f = L.face_features(st, 1.4).reshape(50, 4, 4)
    rb, pb, cb, vtb = L.face_means(st, 1.4)
    # the two cells next to the face average to the means
    assert np.allclose(f[:, 1, 0] + f[:, 2, 0], 0)
    assert np.allclose(f[:, 1, 2] + f[:, 2, 2], 0)
    assert np.allclose(f[:, 1, 3] + f[:, 2, 3], 0)
    assert np.allclose(f[:, 0, 1] * cb, st[:, 0, 1])
    flat = st.copy()
    flat[..., 3] = 2.0
    assert not L.sensor(flat).any()
    flat[:, 0, 3] = 2.0 * 1.11                 # 11 % jump on the outer pair fires
    assert L.sensor(flat).all()
    flat[:, 0, 3] = 2.0 * 1.09
    assert not L.sensor(flat).any()


def test_flux_tangential_galilean_invariance():
    """Boosting every cell by u along the face transforms the PLM-HLLC flux exactly as the
    correction's frame assumes: F_mt -> F_mt + u F_rho, F_E -> F_E + u F_mt + u^2/2 F_rho."""
    rng = np.random.default_rng(7)
    n, g, u = 200, 1.4, 0.37
    cells = np.empty((4, 5, n))
    cells[:, 0] = rng.uniform(0.5, 2.0, (4, n))
    cells[:, 1] = rng.normal(0, 0.4, (4, n))
    cells[:, 2] = rng.normal(0, 0.4, (4, n))
    cells[:, 3] = 0.0
    cells[:, 4] = rng.uniform(0.5, 3.0, (4, n))

    def flux(c):
        sL = L._vanleer(c[1] - c[0], c[2] - c[1])
        sR = L._vanleer(c[2] - c[1], c[3] - c[2])
        return L.hllc(c[1] + 0.5 * sL, c[2] - 0.5 * sR, g)

    F0 = flux(cells)
    boosted = cells.copy()
    boosted[:, 2] += u
    F1 = flux(boosted)
    assert np.allclose(F1[0], F0[0], rtol=1e-12, atol=1e-12)
    assert np.allclose(F1[1], F0[1], rtol=1e-12, atol=1e-12)
    assert np.allclose(F1[2], F0[2] + u * F0[0], rtol=1e-11, atol=1e-12)
    assert np.allclose(F1[4], F0[4] + u * F0[2] + 0.5 * u * u * F0[0], rtol=1e-11, atol=1e-12)


def test_hllc_restatement_matches_golden_vectors():
    """The numpy HLLC used by the corrected-stage oracle equals the C++ oracle on the reference's
    own vectors (tests/golden/vec_hllc.npz)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "vec_hllc.npz"))
    got = L.hllc(z["hllc_ul"], z["hllc_ur"], 1.4)     # x direction: states already in face order
    ref = z["hllc_dir1"]
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-12 * np.max(np.abs(ref)))


def test_positivity_test():
    u = np.array([[1.0, 0.1, 0.0, 0.0, 2.5]])
    zero = np.zeros((1, 5))
    assert L.positivity_ok(u, u, zero, np.array([0.1]), 1.4).all()
    big = np.array([[2.0, 0, 0, 0, 0]])        # removes all mass from the left cell
    assert not L.positivity_ok(u, u, big, np.array([0.2]), 1.4).any()


def test_stage_zero_net_equals_plain_scheme():
    """With a network whose output is zero the corrected stage equals the plain scheme."""
    rng = np.random.default_rng(8)
    u = np.empty((5, 12, 16))
    u[0] = rng.uniform(0.8, 1.2, (12, 16))
    u[1] = rng.normal(0, 0.1, (12, 16))
    u[2] = rng.normal(0, 0.1, (12, 16))
    u[3] = 0
    u[4] = 2.5 / 0.4 + 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0]
    zero = np.zeros(L.NPARAM)
    a, *_ = L.stage_update(u, None, (1.0, 1.0), 1e-3, 1 / 16, 1 / 12, 1.4)
    b, *_ = L.stage_update(u, None, (1.0, 1.0), 1e-3, 1 / 16, 1 / 12, 1.4, (zero, np.ones(16), np.ones(4)))
    assert np.array_equal(a, b)


def test_dissipation_bound():
    """Uniform stencils and contacts at rest have no upwind dissipation (bound 0); a contact moving
    at v_n = u gets 1/2 |u| |d rho| (1, |u|, 0, u^2/2); a pure acoustic jump gets both acoustic waves
    (planted mistake: a wave speed without |.| or a dropped wave changes these values)."""
    rng = np.random.default_rng(9)
    st = _stencil(rng, 100)
    flat = np.repeat(st[:, :1], 4, axis=1)
    assert np.all(L.dissipation_bound(L.face_features(flat, 1.4), 1.4) == 0)
    step = np.zeros((1, 4, 4))
    step[0, :, 0] = [1.0, 1.0, 2.0, 2.0]
    step[0, :, 3] = 1.0
    assert np.all(L.dissipation_bound(L.face_features(step, 1.4), 1.4) == 0)
    u = -0.3
    mov = step.copy()
    mov[0, :, 1] = u
    f = L.face_features(mov, 1.4)
    rb, pb, cb, vtb = L.face_means(mov, 1.4)
    drho = (2.0 - 1.0) / rb[0]                     # jump of rho / rho_bar (PLM slopes vanish)
    un = u / cb[0]
    b = L.dissipation_bound(f, 1.4)[0]
    assert np.allclose(b, 0.5 * abs(un) * drho * np.array([1.0, abs(un), 0.0, 0.5 * un * un]), rtol=1e-12)
    # acoustic jump: dp with rho c dvn = dp (a right-going simple wave), at rest: only lambda_4 = c
    ac = np.zeros((1, 4, 4))
    ac[0, :, 0] = 1.0
    ac[0, :, 3] = [1.0, 1.0, 1.1, 1.1]
    c = np.sqrt(1.4 * 1.05 / 1.0)
    ac[0, :, 1] = [0.0, 0.0, 0.1 / c, 0.1 / c]
    bac = L.dissipation_bound(L.face_features(ac, 1.4), 1.4)[0]
    assert bac[0] > 0 and bac[1] > 0 and bac[3] > 0
    assert bac[2] == 0                              # no tangential velocity: |r_mt| = |v_t| = 0

