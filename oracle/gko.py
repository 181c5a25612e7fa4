"""ctypes binding of the CPU oracle (gko_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

``OracleSim`` mirrors the reference's Sim (src/sim.hpp:23-58) over the oracle's C API, and
``run_to_end`` mirrors the reference main loop (src/main.cpp:44-52).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
                ("gamma", C.c_double), ("cfl", C.c_double),
                ("integrator", C.c_int), ("rc_order", C.c_int), ("stepmax", C.c_int),
                ("bc", C.c_int * 6), ("g", C.c_double * 3),
                ("npoint", C.c_int), ("pg", (C.c_double * 8) * 2), ("nz", C.c_int)]


_libs = {}


def lib(flavour: str = "strict") -> C.CDLL:
    """Load liboracle_<flavour>.so (strict: IEEE parity build; fast: timed baseline build)."""
    if flavour not in _libs:
        path = os.path.join(HERE, f"liboracle_{flavour}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.gko_create.restype = vp; L.gko_create.argtypes = [C.POINTER(Params), ip]
        L.gko_destroy.argtypes = [vp]
        L.gko_set_cons.argtypes = [vp, vp]
        L.gko_get_cons.argtypes = [vp, vp]
        L.gko_advance.restype = ip; L.gko_advance.argtypes = [vp, C.c_double]
        L.gko_time.restype = C.c_double; L.gko_time.argtypes = [vp]
        L.gko_step.restype = ip; L.gko_step.argtypes = [vp]
        L.gko_is_continue.restype = ip; L.gko_is_continue.argtypes = [vp]
        L.gko_last_dt.restype = C.c_double; L.gko_last_dt.argtypes = [vp]
        L.gko_update_bd.restype = ip
        L.gko_update_bd.argtypes = [vp]
        L.gko_enroll_fixed.argtypes = [vp, ip]
        L.gko_setup_immersed.restype = ip
        L.gko_setup_immersed.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.gko_dt_log.restype = ip; L.gko_dt_log.argtypes = [vp, dp, dp, ip]
        L.gko_hllc.argtypes = [ip, dp, dp, ip, C.c_double, dp]
        L.gko_plm_row.argtypes = [ip, dp, dp, dp]
        L.gko_cons_to_prim.argtypes = [ip, dp, C.c_double, dp]
        L.gko_prim_to_cons.argtypes = [ip, dp, C.c_double, dp]
        L.gko_cmax.restype = C.c_double; L.gko_cmax.argtypes = [ip, dp, C.c_double, C.c_double]
        _libs[flavour] = L
    return _libs[flavour]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(prob) -> Params:
    p = Params()
    p.nx, p.ny, p.ng = prob.nx, prob.ny, prob.ng
    (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax) = prob.extent
    p.gamma, p.cfl = prob.gamma, prob.cfl
    p.integrator, p.rc_order, p.stepmax = prob.integrator, prob.rc_order, prob.stepmax
    for k in range(4):
        p.bc[k] = prob.device_bc()[k]
    p.bc[4], p.bc[5] = prob.bcz
    p.nz = prob.nz
    for k in range(3):
        p.g[k] = prob.g[k]
    p.npoint = len(prob.point)
    for q, m in enumerate(prob.point):
        for k in range(8):
            p.pg[q][k] = m[k]
    return p


class OracleSim:
    """The oracle's Sim: state with ghosts, Advance(dtoutput), time/step accessors."""

    def __init__(self, prob, precision: int = 64, flavour: str = "strict"):
        self.L = lib(flavour)
        self.prob = prob
        self.dtype = np.float64 if precision == 64 else np.float32
        self.params = make_params(prob)
        self.h = self.L.gko_create(C.byref(self.params), precision)
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        ng = prob.ng
        self.shape = (5, prob.ny + 2 * ng, prob.nx + 2 * ng)
        if prob.nz > 1:
            self.shape = (5, prob.nz + 2 * ng) + self.shape[1:]
        full = np.zeros(self.shape, self.dtype)
        enrolled = prob.enrolled_state()
        if enrolled is not None:
            # a setup that fills the ghost cells, then enrolls the fixed data
            full[:] = enrolled
        full[self._active()] = prob.cons
        self.L.gko_set_cons(self.h, full.ctypes.data)
        if enrolled is not None:
            self.L.gko_enroll_fixed(self.h, int(bool(prob.fix_inner)))
        if prob.shell:
            if self.L.gko_setup_immersed(self.h, *[float(v) for v in prob.shell]) != 0:
                raise ValueError("the oracle rejected the immersed shell (out of domain or too close to a face)")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.gko_destroy(h)
            self.h = None

    def enroll_fixed(self, full: np.ndarray):
        """EnrollFixedBoundaryData with the ghost layers of `full` ([5][NY+2g][NX+2g]); the
        current state is kept."""
        cur = self.full()
        snap = np.ascontiguousarray(full, self.dtype)
        assert snap.shape == self.shape
        self.L.gko_set_cons(self.h, snap.ctypes.data)
        self.L.gko_enroll_fixed(self.h, 0)
        self.L.gko_set_cons(self.h, cur.ctypes.data)

    def _active(self):
        g, p = self.prob.ng, self.prob
        act = (slice(None), slice(g, g + p.ny), slice(g, g + p.nx))
        return act[:1] + (slice(g, g + p.nz),) + act[1:] if p.nz > 1 else act

    def full(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.L.gko_get_cons(self.h, out.ctypes.data)
        return out

    def cons(self) -> np.ndarray:
        return self.full()[self._active()].copy()

    def update_bd(self):
        """Boundary::UpdateBD on the current state, as main does once after Setup (src/main.cpp:30):
        ghost cells, and in place the immersed shell and the inner fixed region."""
        if self.L.gko_update_bd(self.h) != 0:
            raise ValueError("boundary update failed (fixed data not enrolled?)")
        self._bd_done = True

    def advance(self, dtoutput: float) -> int:
        if not getattr(self, "_bd_done", False):
            self.update_bd()
        return self.L.gko_advance(self.h, dtoutput)

    @property
    def t(self) -> float:
        return self.L.gko_time(self.h)

    @property
    def step(self) -> int:
        return self.L.gko_step(self.h)

    @property
    def is_continue(self) -> bool:
        return bool(self.L.gko_is_continue(self.h))

    def dt_log(self, cap: int = 1 << 16):
        t = np.zeros(cap); d = np.zeros(cap)
        n = self.L.gko_dt_log(self.h, _dp(t), _dp(d), cap)
        return t[:min(n, cap)], d[:min(n, cap)]

    def run_to_end(self):
        """The reference main loop (src/main.cpp:44-52): Advance by dtoutput up to tmax."""
        tmax, dtoutput = self.prob.tmax, self.prob.dtoutput
        tnow = self.t
        while abs(tmax - tnow) > 1e-12 and self.is_continue:
            dtoutput = min(dtoutput, tmax - tnow)
            self.advance(dtoutput)
            tnow = self.t
        return self


# ---- per-function helpers -------------------------------------------------------------------
def hllc(ul: np.ndarray, ur: np.ndarray, direction: int, gamma: float) -> np.ndarray:
    ul = np.ascontiguousarray(ul, np.float64); ur = np.ascontiguousarray(ur, np.float64)
    n = ul.shape[1]
    f = np.zeros((5, n))
    lib().gko_hllc(n, _dp(ul), _dp(ur), direction, gamma, _dp(f))
    return f


def plm_row(prim: np.ndarray):
    prim = np.ascontiguousarray(prim, np.float64)
    m = prim.shape[1]
    ul = np.zeros((5, m + 1)); ur = np.zeros((5, m + 1))
    lib().gko_plm_row(m, _dp(prim), _dp(ul), _dp(ur))
    return ul, ur


def cons_to_prim(cons: np.ndarray, gamma: float) -> np.ndarray:
    cons = np.ascontiguousarray(cons, np.float64)
    out = np.zeros_like(cons)
    lib().gko_cons_to_prim(cons.shape[1], _dp(cons), gamma, _dp(out))
    return out


def prim_to_cons(prim: np.ndarray, gamma: float) -> np.ndarray:
    prim = np.ascontiguousarray(prim, np.float64)
    out = np.zeros_like(prim)
    lib().gko_prim_to_cons(prim.shape[1], _dp(prim), gamma, _dp(out))
    return out


def cmax(cons: np.ndarray, gamma: float, cmax0: float = 1e-16) -> float:
    cons = np.ascontiguousarray(cons, np.float64)
    return lib().gko_cmax(cons.shape[1], _dp(cons), gamma, cmax0)
