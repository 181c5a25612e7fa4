"""CPU restatement of the learned flux correction -- TEST INFRASTRUCTURE ONLY.

The reference (ShunqH/Gaukuk) has no learned component: its coarse-grid flux is the HLLC flux of
src/flux/hllc.cpp:22-103 on PLM states (src/reconstruction/recon_plm.cpp, van Leer limiter of
reconstruction.hpp:36-42), advanced by the RK stages of src/evolution/rk3.cpp:11-28.  This module
states, in numpy, every piece the project adds on top of that scheme, so the HIP kernels can be
checked against it:

* the MLP 16 -> 64 -> 64 -> 4 (ReLU) with BF16 operands and FP32 accumulation: ``forward``,
  ``backward`` (the quantised analytic gradient the device computes) and the same functions with
  ``quant=False`` (exact float64, checked against finite differences), and ``adam``;
* the face features, the tangential Galilean frame, the output scaling, the dissipation bound,
  the shock sensor and the positivity test: ``face_features``, ``correction_lab``,
  ``dissipation_bound``, ``sensor``, ``positivity_ok``;
* one corrected RK stage on a small periodic grid: ``stage_update`` (and ``rk3_step``).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU baseline may import it.
"""
from __future__ import annotations

import numpy as np

NIN, NHID, NOUT = 16, 64, 4
OFF = {"W1": 0, "b1": 1024, "W2": 1088, "b2": 5184, "W3": 5248, "b3": 5504}
SHAPE = {"W1": (NHID, NIN), "b1": (NHID,), "W2": (NHID, NHID), "b2": (NHID,),
         "W3": (NOUT, NHID), "b3": (NOUT,)}
NPARAM = 5508
SENSOR = 0.1          # |dp| / p threshold of the shock sensor
POS_FACES = 4.0       # faces per cell in 2D: the convex-split factor of the positivity test


# ---------------------------------------------------------------------------------------------
# BF16
def bf16(x) -> np.ndarray:
    """Round to the nearest bfloat16, ties to even, via float32 (the device's conversion:
    u + 0x7FFF + ((u >> 16) & 1), keep the high 16 bits).  Returns float64 values."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p) -> dict:
    p = np.asarray(p, dtype=np.float64)
    return {k: p[OFF[k]:OFF[k] + int(np.prod(SHAPE[k]))].reshape(SHAPE[k]) for k in OFF}


def pack(d: dict) -> np.ndarray:
    p = np.zeros(NPARAM, dtype=np.float64)
    for k in OFF:
        p[OFF[k]:OFF[k] + int(np.prod(SHAPE[k]))] = np.asarray(d[k], dtype=np.float64).ravel()
    return p


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (deterministic numpy generator)."""
    rng = np.random.default_rng(seed)
    d = {}
    for k, fan_in in (("W1", NIN), ("W2", NHID), ("W3", NHID)):
        lim = np.sqrt(6.0 / fan_in)
        d[k] = rng.uniform(-lim, lim, SHAPE[k])
    for k in ("b1", "b2", "b3"):
        d[k] = np.zeros(SHAPE[k])
    d["W3"] *= 0.1
    return pack(d).astype(np.float32).astype(np.float64)


# ---------------------------------------------------------------------------------------------
# MLP
def forward(p, x, in_scale, quant: bool = True, keep: bool = False):
    """y = W3 relu(W2 relu(W1 q(x*in_scale) + b1) + b2) + b3 with q() = BF16 rounding of every
    MFMA operand (inputs, weights, hidden activations) and float32 bias/accumulation."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    xs = np.asarray(x, np.float32) * np.asarray(in_scale, np.float32) if quant else \
        np.asarray(x, np.float64) * np.asarray(in_scale, np.float64)
    X = q(xs)
    W1, W2, W3 = q(w["W1"]), q(w["W2"]), q(w["W3"])
    Z1 = X @ W1.T + w["b1"]
    H1 = q(np.maximum(Z1, 0.0))
    Z2 = H1 @ W2.T + w["b2"]
    H2 = q(np.maximum(Z2, 0.0))
    Y = H2 @ W3.T + w["b3"]
    if keep:
        return Y, dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, W3=W3)
    return Y


def loss(p, x, t, wt, in_scale, quant: bool = False) -> float:
    """Mean over samples and outputs of wt * (y - t)^2, divided by the batch size n."""
    y = forward(p, x, in_scale, quant)
    n = len(x)
    return float(np.sum(np.asarray(wt)[:, None] * (y - t) ** 2) / (NOUT * n))


def backward(p, x, t, wt, in_scale, quant: bool = True):
    """(loss, gradient) of ``loss``.  With quant=True the backward products take BF16 operands as
    on the device (dY, dZ2, dZ1 and the stored activations rounded before each product); bias
    gradients sum the unrounded values."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    y, k = forward(p, x, in_scale, quant, keep=True)
    n = len(x)
    wt = np.asarray(wt, dtype=np.float64)[:, None]
    L = float(np.sum(wt * (y - t) ** 2) / (NOUT * n))
    dY = 2.0 * wt * (y - t) / (NOUT * n)
    dYq = q(dY)
    dH2 = dYq @ k["W3"]
    dZ2 = dH2 * (k["Z2"] > 0)
    dZ2q = q(dZ2)
    dH1 = dZ2q @ k["W2"]
    dZ1 = dH1 * (k["Z1"] > 0)
    dZ1q = q(dZ1)
    g = {"W3": dYq.T @ k["H2"], "b3": dY.sum(0),
         "W2": dZ2q.T @ k["H1"], "b2": dZ2.sum(0),
         "W1": dZ1q.T @ k["X"], "b1": dZ1.sum(0)}
    return L, pack(g)


def adam(p, m, v, g, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) with bias correction; step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------------------------
# Face features and the correction in the lab frame
def face_means(st, gamma):
    """st: [..., 4 cells, 4] primitive stencil (rho, v_n, v_t, p) across a face (cells i-2..i+1).
    Returns rho_bar, p_bar, c_bar, vt_bar from the two cells next to the face."""
    rb = 0.5 * (st[..., 1, 0] + st[..., 2, 0])
    pb = 0.5 * (st[..., 1, 3] + st[..., 2, 3])
    cb = np.sqrt(gamma * pb / rb)
    vtb = 0.5 * (st[..., 1, 2] + st[..., 2, 2])
    return rb, pb, cb, vtb


def face_features(st, gamma):
    """16 features per face, cell-major: rho/rho_bar - 1, v_n/c_bar, (v_t - vt_bar)/c_bar, p/p_bar - 1."""
    rb, pb, cb, vtb = face_means(st, gamma)
    f = np.empty(st.shape[:-2] + (4, 4))
    f[..., 0] = st[..., 0] / rb[..., None] - 1.0
    f[..., 1] = st[..., 1] / cb[..., None]
    f[..., 2] = (st[..., 2] - vtb[..., None]) / cb[..., None]
    f[..., 3] = st[..., 3] / pb[..., None] - 1.0
    return f.reshape(st.shape[:-2] + (16,))


def sensor(st):
    """True where the shock sensor fires: max over the stencil's neighbour pairs of
    |p_{k+1} - p_k| / min(p_k, p_{k+1}) > 0.1."""
    p = st[..., 3]
    r = np.abs(np.diff(p, axis=-1)) / np.minimum(p[..., 1:], p[..., :-1])
    return np.max(r, axis=-1) > SENSOR


def dissipation_bound(feats, gamma):
    """D_q of the bound, in network units, from the features: cell k has rho = 1 + f[4k],
    v_n = f[4k+1], v_t = f[4k+2], p = (1 + f[4k+3]) / gamma in units of (rho_bar, c_bar).  The jump
    between the two PLM face states is split into the Euler waves (Toro, Riemann Solvers and
    Numerical Methods, ch. 3: acoustic v_n -+ c, contact and shear v_n) and
    D = 1/2 sum_k |lambda_k| |alpha_k| |r_k| component by component, with arithmetic averages."""
    f = np.asarray(feats, np.float64).reshape(feats.shape[:-1] + (4, 4))
    w = np.stack([1.0 + f[..., 0], f[..., 1], f[..., 2], (1.0 + f[..., 3]) / gamma], axis=-1)
    l = w[..., 1, :] + 0.5 * _vanleer(w[..., 1, :] - w[..., 0, :], w[..., 2, :] - w[..., 1, :])
    r = w[..., 2, :] - 0.5 * _vanleer(w[..., 2, :] - w[..., 1, :], w[..., 3, :] - w[..., 2, :])
    rho, vn, vt, p = [0.5 * (l[..., q] + r[..., q]) for q in range(4)]
    dr, dvn, dvt, dp = [r[..., q] - l[..., q] for q in range(4)]
    c = np.sqrt(gamma * p / rho)
    a1 = (dp - rho * c * dvn) / (2 * c * c)
    a4 = (dp + rho * c * dvn) / (2 * c * c)
    a2 = dr - dp / (c * c)
    a3 = rho * dvt
    H = c * c / (gamma - 1.0) + 0.5 * (vn * vn + vt * vt)
    # |lambda| |alpha| per wave, times |r_k| per component
    waves = [(np.abs(vn - c) * np.abs(a1), (1.0, np.abs(vn - c), np.abs(vt), np.abs(H - vn * c))),
             (np.abs(vn) * np.abs(a2), (1.0, np.abs(vn), np.abs(vt), 0.5 * (vn * vn + vt * vt))),
             (np.abs(vn) * np.abs(a3), (0.0, 0.0, 1.0, np.abs(vt))),
             (np.abs(vn + c) * np.abs(a4), (1.0, np.abs(vn + c), np.abs(vt), np.abs(H + vn * c)))]
    return np.stack([0.5 * sum(wk * rk[q] for wk, rk in waves) for q in range(4)], axis=-1)


def to_network_units(dF, st, gamma):
    """Lab-frame flux difference (mass, m_n, m_t, E) -> the network's target: the flux in the frame
    moving tangentially with vt_bar, over (rho c, rho c^2, rho c^2, rho c^3) of the face means."""
    rb, pb, cb, vtb = face_means(st, gamma)
    m, mn, mt, e = dF[..., 0], dF[..., 1], dF[..., 2], dF[..., 3]
    out = np.stack([m, mn, mt - vtb * m, e - vtb * mt + 0.5 * vtb * vtb * m], axis=-1)
    scale = np.stack([rb * cb, rb * cb ** 2, rb * cb ** 2, rb * cb ** 3], axis=-1)
    return out / scale


def correction_lab(yhat, st, gamma):
    """Inverse of ``to_network_units``: network output (already times out_scale) -> lab-frame
    (mass, m_n, m_t, E) flux correction."""
    rb, pb, cb, vtb = face_means(st, gamma)
    a = yhat[..., 0] * rb * cb
    b = yhat[..., 1] * rb * cb ** 2
    c = yhat[..., 2] * rb * cb ** 2
    d = yhat[..., 3] * rb * cb ** 3
    return np.stack([a, b, c + vtb * a, d + vtb * c + 0.5 * vtb * vtb * a], axis=-1)


def _pressure(u, gamma):
    ke = 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0]
    return (gamma - 1.0) * (u[..., 4] - ke)


def positivity_ok(uL, uR, dF5, nu, gamma):
    """The corrected face passes when both one-face shares U_L - 4 nu dF and U_R + 4 nu dF keep
    density and pressure positive (uL, uR, dF5: [..., 5] conserved lab vectors)."""
    a = uL - POS_FACES * nu[..., None] * dF5
    b = uR + POS_FACES * nu[..., None] * dF5
    return (a[..., 0] > 0) & (b[..., 0] > 0) & (_pressure(a, gamma) > 0) & (_pressure(b, gamma) > 0)


# ---------------------------------------------------------------------------------------------
# One corrected stage on a periodic grid, float64 (the scheme of gk_physics.hpp restated)
def _prim(u, gamma):
    d = np.maximum(u[0], 1e-16)
    v = u[1:4] / d
    p = (gamma - 1.0) * (u[4] - 0.5 * d * np.sum(v * v, axis=0))
    p = np.maximum(p, 1e-16)
    return np.stack([d, v[0], v[1], v[2], p])


def _vanleer(a, b):
    m = a * b
    with np.errstate(divide="ignore", invalid="ignore"):
        s = 2.0 * m / (a + b)
    return np.where(m <= 0, 0.0, s)


def hllc(wl, wr, gamma):
    """Vectorised HLLC flux (src/flux/hllc.cpp:22-103) on rotated primitive states [5, ...]."""
    dl, ul, vl, zl, pl = wl
    dr, ur, vr, zr, pr = wr
    al, ar = np.sqrt(gamma * pl / dl), np.sqrt(gamma * pr / dr)
    pv = np.maximum(0.5 * (pl + pr) - 0.125 * (ur - ul) * (dl + dr) * (al + ar), 0.0)
    qk = 0.5 * (1.0 + 1.0 / gamma)
    ql = np.sqrt(1.0 + qk * np.maximum(pv / pl - 1.0, 0.0))
    qr = np.sqrt(1.0 + qk * np.maximum(pv / pr - 1.0, 0.0))
    sl, sr = ul - al * ql, ur + ar * qr
    ss = (pr - pl + dl * ul * (sl - ul) - dr * ur * (sr - ur)) / (dl * (sl - ul) - dr * (sr - ur))
    left = ss >= 0
    dk, vk = np.where(left, dl, dr), np.where(left, ul, ur)
    v2, v3 = np.where(left, vl, vr), np.where(left, zl, zr)
    pk, sk = np.where(left, pl, pr), np.where(left, sl, sr)
    take = np.where(left, sk < 0, sk > 0)
    mk = np.where(take, sk, 0.0)
    ek = pk / (gamma - 1.0) + 0.5 * dk * (vk * vk + v2 * v2 + v3 * v3)
    ck = mk * (ss - vk) / (sk - ss)
    f1 = dk * (vk + ck)
    return np.stack([f1, dk * (vk * vk + ck * sk) + pk, f1 * v2, f1 * v3,
                     (ek + pk) * (vk + ck) + dk * ck * ss * (sk - vk)])


def _axis_fluxes(w, u, axis, gamma, net, dtdh):
    """Fluxes at faces i-1/2 (i = 0..n-1) along `axis` (2: x, 1: y) of a periodic grid [5, ny, nx],
    plus the correction; returns (F [5, ny, nx], fired, rejected) with F rotated back to the lab."""
    rot = [0, 1, 2, 3, 4] if axis == 2 else [0, 2, 3, 1, 4]
    back = [0, 1, 2, 3, 4] if axis == 2 else [0, 3, 1, 2, 4]
    W = w[rot]
    cells = [np.roll(W, s, axis=axis) for s in (2, 1, 0, -1)]   # i-2, i-1, i, i+1
    sL = _vanleer(cells[1] - cells[0], cells[2] - cells[1])
    sR = _vanleer(cells[2] - cells[1], cells[3] - cells[2])
    F = hllc(cells[1] + 0.5 * sL, cells[2] - 0.5 * sR, gamma)[back]
    fired = np.zeros(F.shape[1:], bool)
    rejected = np.zeros(F.shape[1:], bool)
    if net is not None:
        p, in_scale, out_scale = net[:3]
        kappa = net[3] if len(net) > 3 else 0.5
        # (rho, v_n, v_t, p) with v_t the in-plane tangential velocity (v_y on x faces, v_x on y
        # faces; the cyclic rotation above put v_x in slot 3 for y faces)
        vt = 2 if axis == 2 else 3
        st = np.stack([np.stack([c[0], c[1], c[vt], c[4]], axis=-1) for c in cells], axis=-2)
        feats = face_features(st, gamma)
        y = forward(p, feats.reshape(-1, 16), in_scale).reshape(feats.shape[:-1] + (4,))
        bound = kappa * dissipation_bound(feats, gamma)
        yh = np.clip(y * np.asarray(out_scale, np.float64), -bound, bound)
        dlab = correction_lab(yh, st, gamma)      # mass, m_n, m_t, E
        d5 = np.zeros((5,) + dlab.shape[:-1])
        d5[0] = dlab[..., 0]
        d5[1 if axis == 2 else 2] = dlab[..., 1]
        d5[2 if axis == 2 else 1] = dlab[..., 2]
        d5[4] = dlab[..., 3]
        fired = sensor(st)
        uL = np.moveaxis(np.roll(u, 1, axis=axis), 0, -1)
        uR = np.moveaxis(u, 0, -1)
        ok = positivity_ok(uL, uR, np.moveaxis(d5, 0, -1), np.full(fired.shape, dtdh), gamma)
        ok &= np.all(np.isfinite(d5), axis=0)
        rejected = ~fired & ~ok
        use = ~fired & ok
        F = F + np.where(use, d5, 0.0)
    return F, fired, rejected


def stage_update(u_in, u_base, c, dt, dx, dy, gamma, net=None):
    """One RK stage of the corrected scheme on a periodic grid: res = c1*base + c2*u_in - c3*dF
    (three coefficients) or c1*u_in - c2*dF (two: pass u_base=None).  Floors as UpdateCons
    (src/evolution/update_cons.cpp:55-67).  Returns (res, fired faces, rejected faces)."""
    w = _prim(u_in, gamma)
    Fx, fx, rx = _axis_fluxes(w, u_in, 2, gamma, net, dt / dx)
    Fy, fy, ry = _axis_fluxes(w, u_in, 1, gamma, net, dt / dy)
    dF = (np.roll(Fx, -1, axis=2) - Fx) * (dt / dx) + (np.roll(Fy, -1, axis=1) - Fy) * (dt / dy)
    if u_base is None:
        res = c[0] * u_in - c[1] * dF
    else:
        res = c[0] * u_base + c[1] * u_in - c[2] * dF
    low = res[0] < 1e-16
    ke = 0.5 * (res[1] ** 2 + res[2] ** 2 + res[3] ** 2) / res[0]
    press = (gamma - 1.0) * (res[4] - ke)
    res[4] = np.where(press < 1e-16, 1e-16 / (gamma - 1.0) + ke, res[4])
    return res, int(fx.sum() + fy.sum()), int(rx.sum() + ry.sum())


def rk3_step(u, dt, dx, dy, gamma, net=None):
    """SSP-RK3 in the coefficient form of src/evolution/rk3.cpp:11-28."""
    f13 = 1.0 / 3.0
    u1, *_ = stage_update(u, None, (1.0, 1.0), dt, dx, dy, gamma, net)
    u2, *_ = stage_update(u1, u, (0.75, 0.25, 0.25), dt, dx, dy, gamma, net)
    u3, *_ = stage_update(u2, u, (f13, 2 * f13, 2 * f13), dt, dx, dy, gamma, net)
    return u3


def cmax(u, gamma):
    w = _prim(u, gamma)
    cs = np.sqrt(gamma * w[4] / w[0])
    return float(np.max(np.maximum(np.abs(w[1]), np.abs(w[2])) + cs))
