"""Python binding of the HIP library's C ABI (include/gaukuk.h) via ctypes.

``Solver`` mirrors the reference's Sim surface (src/sim.hpp:23-58): upload a state, ``advance``
by an output interval, read time/step, download the state.  There is no fallback: if
``libgaukuk_hip.so`` is missing or a HIP call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GK_LIB_PATH") or os.path.join(HERE, "libgaukuk_hip.so")

GK_STATUS = {0: "GK_OK", 1: "GK_EINVAL", 2: "GK_EHIP", 3: "GK_ENOMEM", 4: "GK_ESTATE", 5: "GK_ENUMERIC"}

# Every symbol include/gaukuk.h declares (checked by tests/test_capi_symbols.py).
SYMBOLS = ["gk_create", "gk_destroy", "gk_upload_cons", "gk_download_cons", "gk_upload_active",
           "gk_download_active", "gk_advance", "gk_get_time", "gk_step_log", "gk_run_steps",
           "gk_synchronize", "gk_profile_steps", "gk_cons_sums", "gk_check_finite", "gk_set_tile",
           "gk_get_tile", "gk_selftest", "gk_last_error", "gk_version",
           "gk_ens_create", "gk_ens_destroy", "gk_ens_upload", "gk_ens_download", "gk_ens_set_tile",
           "gk_ens_get_tile", "gk_ens_run_steps", "gk_ens_capture", "gk_ens_synchronize", "gk_ens_time",
           "gk_ens_get_time", "gk_ens_last_error"]


class GkGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nghost", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double)]


class GkPhys(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("g", C.c_double * 3),
                ("integrator", C.c_int), ("rc_order", C.c_int), ("stepmax", C.c_int),
                ("bc", C.c_int * 6), ("precision", C.c_int)]


class GaukukError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{GK_STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C gaukuk_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, st = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int
        L.gk_create.restype = st
        L.gk_create.argtypes = [C.POINTER(GkGrid), C.POINTER(GkPhys), C.POINTER(vp)]
        L.gk_destroy.argtypes = [vp]
        for name in ("gk_upload_cons", "gk_download_cons"):
            getattr(L, name).restype = st
            getattr(L, name).argtypes = [vp, vp]
        for name in ("gk_upload_active", "gk_download_active"):
            getattr(L, name).restype = st
            getattr(L, name).argtypes = [vp, dp]
        L.gk_advance.restype = st
        L.gk_advance.argtypes = [vp, C.c_double, dp, C.POINTER(ip)]
        L.gk_get_time.restype = st
        L.gk_get_time.argtypes = [vp, dp, dp, C.POINTER(ip), C.POINTER(ip)]
        L.gk_step_log.restype = st
        L.gk_step_log.argtypes = [vp, C.POINTER(ip), dp, dp, ip, C.POINTER(ip)]
        L.gk_run_steps.restype = st
        L.gk_run_steps.argtypes = [vp, ip]
        L.gk_synchronize.restype = st
        L.gk_synchronize.argtypes = [vp]
        L.gk_profile_steps.restype = st
        L.gk_profile_steps.argtypes = [vp, ip, dp, dp, C.POINTER(ip)]
        L.gk_cons_sums.restype = st
        L.gk_cons_sums.argtypes = [vp, dp]
        L.gk_check_finite.restype = st
        L.gk_check_finite.argtypes = [vp]
        L.gk_set_tile.restype = st
        L.gk_set_tile.argtypes = [vp, ip]
        L.gk_get_tile.restype = st
        L.gk_get_tile.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
        L.gk_last_error.restype = C.c_char_p
        L.gk_last_error.argtypes = [vp]
        L.gk_selftest.restype = st
        L.gk_selftest.argtypes = [C.c_longlong, C.c_ulonglong, ip, C.POINTER(C.c_longlong)]
        L.gk_version.restype = C.c_char_p
        L.gk_version.argtypes = []
        L.gk_ens_create.restype = st
        L.gk_ens_create.argtypes = [ip, C.POINTER(GkGrid), C.POINTER(GkPhys), C.POINTER(vp)]
        L.gk_ens_destroy.argtypes = [vp]
        for name in ("gk_ens_upload", "gk_ens_download"):
            getattr(L, name).restype = st
            getattr(L, name).argtypes = [vp, ip, dp]
        L.gk_ens_set_tile.restype = st
        L.gk_ens_set_tile.argtypes = [vp, ip, ip]
        L.gk_ens_get_tile.restype = st
        L.gk_ens_get_tile.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
        L.gk_ens_run_steps.restype = st
        L.gk_ens_run_steps.argtypes = [vp, ip, ip]
        L.gk_ens_capture.restype = st
        L.gk_ens_capture.argtypes = [vp, ip]
        L.gk_ens_synchronize.restype = st
        L.gk_ens_synchronize.argtypes = [vp]
        L.gk_ens_time.restype = st
        L.gk_ens_time.argtypes = [vp, ip, ip, dp]
        L.gk_ens_get_time.restype = st
        L.gk_ens_get_time.argtypes = [vp, ip, dp, dp, C.POINTER(ip)]
        L.gk_ens_last_error.restype = C.c_char_p
        L.gk_ens_last_error.argtypes = [vp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_grid(prob) -> GkGrid:
    g = GkGrid()
    g.nx, g.ny, g.nz, g.nghost = prob.nx, prob.ny, 1, prob.ng
    (g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax) = prob.extent
    return g


def make_phys(prob, precision: int = 64) -> GkPhys:
    p = GkPhys()
    p.gamma, p.cfl = prob.gamma, prob.cfl
    for k in range(3):
        p.g[k] = prob.g[k]
    p.integrator, p.rc_order, p.stepmax = prob.integrator, prob.rc_order, prob.stepmax
    for k in range(4):
        p.bc[k] = prob.bc[k]
    p.bc[4] = p.bc[5] = 1
    p.precision = precision
    return p


class Solver:
    """One simulation on the GPU through the C ABI."""

    def __init__(self, prob, precision: int = 64, seg_rows: int = 0, upload: bool = True):
        self.L = lib()
        self.prob = prob
        self.precision = precision
        self.grid = make_grid(prob)
        self.phys = make_phys(prob, precision)
        h = C.c_void_p()
        st = self.L.gk_create(C.byref(self.grid), C.byref(self.phys), C.byref(h))
        if st != 0:
            raise GaukukError(st, self.L.gk_last_error(None).decode())
        self.h = h
        if seg_rows:
            self._check(self.L.gk_set_tile(self.h, seg_rows))
        if upload and prob.cons is not None:
            self.upload(prob.cons)

    def _check(self, st):
        if st != 0:
            raise GaukukError(st, self.L.gk_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.gk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, cons: np.ndarray):
        a = np.ascontiguousarray(cons, np.float64)
        assert a.shape == (5, self.prob.ny, self.prob.nx)
        self._check(self.L.gk_upload_active(self.h, _dp(a)))

    def download(self) -> np.ndarray:
        out = np.empty((5, self.prob.ny, self.prob.nx), np.float64)
        self._check(self.L.gk_download_active(self.h, _dp(out)))
        return out

    def download_ref_layout(self) -> np.ndarray:
        ng = self.prob.ng
        dt = np.float64 if self.precision == 64 else np.float32
        out = np.empty((5, 1 + 2 * ng, self.prob.ny + 2 * ng, self.prob.nx + 2 * ng), dt)
        self._check(self.L.gk_download_cons(self.h, out.ctypes.data))
        return out

    def advance(self, dtoutput: float):
        t = C.c_double()
        s = C.c_int()
        self._check(self.L.gk_advance(self.h, dtoutput, C.byref(t), C.byref(s)))
        return t.value, s.value

    def time(self):
        t, dt, s, ic = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self._check(self.L.gk_get_time(self.h, C.byref(t), C.byref(dt), C.byref(s), C.byref(ic)))
        return t.value, dt.value, s.value, bool(ic.value)

    @property
    def t(self):
        return self.time()[0]

    @property
    def step(self):
        return self.time()[2]

    def run_to_end(self):
        """The reference main loop (src/main.cpp:44-52)."""
        tmax, dtoutput = self.prob.tmax, self.prob.dtoutput
        tnow = self.t
        while abs(tmax - tnow) > 1e-12 and self.time()[3]:
            dtoutput = min(dtoutput, tmax - tnow)
            tnow, _ = self.advance(dtoutput)
        return self

    def step_log(self, cap: int = 4096):
        steps = np.zeros(cap, np.int32)
        t = np.zeros(cap)
        dt = np.zeros(cap)
        n = C.c_int()
        self._check(self.L.gk_step_log(self.h, steps.ctypes.data_as(C.POINTER(C.c_int)), _dp(t), _dp(dt), cap,
                                       C.byref(n)))
        return steps[:n.value], t[:n.value], dt[:n.value]

    def run_steps(self, n: int):
        self._check(self.L.gk_run_steps(self.h, n))

    def synchronize(self):
        self._check(self.L.gk_synchronize(self.h))

    def profile(self, n: int):
        tot = C.c_double()
        kms = (C.c_double * 4)()
        cnt = (C.c_int * 4)()
        self._check(self.L.gk_profile_steps(self.h, n, C.byref(tot), kms, cnt))
        return tot.value, list(kms), list(cnt)

    def sums(self) -> np.ndarray:
        out = np.zeros(5)
        self._check(self.L.gk_cons_sums(self.h, _dp(out)))
        return out

    def check_finite(self) -> bool:
        st = self.L.gk_check_finite(self.h)
        if st == 5:
            return False
        self._check(st)
        return True

    def tile(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.gk_get_tile(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


BATCHED, LOOPED, GRAPH = 0, 1, 2


class Ensemble:
    """Many independent problems advanced together on the GPU (gk_ens_* entry points).

    Problems share precision, integrator, RcOrder, gamma and gravity; each has its own size,
    extent, CFL, stepmax, boundaries and time step."""

    def __init__(self, probs, precision: int = 32, seg_rows: int = 0, waves_per_cu: int = 0):
        self.L = lib()
        self.probs = list(probs)
        P = len(self.probs)
        grids = (GkGrid * P)(*[make_grid(p) for p in self.probs])
        phys = (GkPhys * P)(*[make_phys(p, precision) for p in self.probs])
        h = C.c_void_p()
        st = self.L.gk_ens_create(P, grids, phys, C.byref(h))
        if st != 0:
            raise GaukukError(st, self.L.gk_ens_last_error(None).decode())
        self.h = h
        self.precision = precision
        if seg_rows or waves_per_cu:
            self.set_tile(seg_rows, waves_per_cu)
        for q, p in enumerate(self.probs):
            self.upload(q, p.cons)

    def _check(self, st):
        if st != 0:
            raise GaukukError(st, self.L.gk_ens_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.gk_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.probs)

    @property
    def cells(self) -> int:
        return sum(p.nx * p.ny for p in self.probs)

    def upload(self, q: int, cons):
        a = np.ascontiguousarray(cons, np.float64)
        self._check(self.L.gk_ens_upload(self.h, q, _dp(a)))

    def download(self, q: int) -> np.ndarray:
        p = self.probs[q]
        out = np.empty((5, p.ny, p.nx), np.float64)
        self._check(self.L.gk_ens_download(self.h, q, _dp(out)))
        return out

    def set_tile(self, seg_rows: int = 0, waves_per_cu: int = 0):
        self._check(self.L.gk_ens_set_tile(self.h, seg_rows, waves_per_cu))

    def tile(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.gk_ens_get_tile(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def run_steps(self, n: int, mode: int = BATCHED):
        self._check(self.L.gk_ens_run_steps(self.h, n, mode))

    def capture(self, n: int):
        self._check(self.L.gk_ens_capture(self.h, n))

    def synchronize(self):
        self._check(self.L.gk_ens_synchronize(self.h))

    def time(self, n: int, mode: int = BATCHED) -> float:
        ms = C.c_double()
        self._check(self.L.gk_ens_time(self.h, n, mode, C.byref(ms)))
        return ms.value

    def get_time(self, q: int):
        t, dt, s = C.c_double(), C.c_double(), C.c_int()
        self._check(self.L.gk_ens_get_time(self.h, q, C.byref(t), C.byref(dt), C.byref(s)))
        return t.value, dt.value, s.value


def selftest(count: int = 1 << 24, seed: int = 12345, emax: int = 500) -> int:
    """Bitwise check of the kernels' short FP64 quotient/root against IEEE; returns mismatches."""
    bad = C.c_longlong()
    st = lib().gk_selftest(count, seed, emax, C.byref(bad))
    if st != 0:
        raise GaukukError(st, lib().gk_last_error(None).decode())
    return bad.value


def version() -> str:
    return lib().gk_version().decode()
