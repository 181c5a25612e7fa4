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
SYMBOLS = ["gk_set_bd_hook", "gk_set_time", "gk_set_point_gravity", "gk_set_fixed_boundary", "gk_set_immersed_shell", "gk_update_bd", "gk_create", "gk_destroy", "gk_upload_cons", "gk_download_cons", "gk_upload_active",
           "gk_download_active", "gk_advance", "gk_get_time", "gk_step_log", "gk_run_steps",
           "gk_synchronize", "gk_profile_steps", "gk_cons_sums", "gk_check_finite", "gk_set_tile",
           "gk_get_tile", "gk_selftest", "gk_cons_sums_noalloc", "gk_last_error", "gk_version",
           "gk_ens_create", "gk_ens_destroy", "gk_ens_upload", "gk_ens_download", "gk_ens_set_tile",
           "gk_ens_get_tile", "gk_ens_run_steps", "gk_ens_capture", "gk_ens_synchronize", "gk_ens_time",
           "gk_ens_get_time", "gk_ens_last_error",
           "gk_net_create", "gk_net_destroy", "gk_net_set_params", "gk_net_get_params", "gk_net_set_scales",
           "gk_net_get_scales", "gk_net_set_kappa", "gk_net_get_kappa", "gk_net_forward", "gk_net_gradient", "gk_net_adam", "gk_net_reserve",
           "gk_net_add_samples", "gk_net_dataset", "gk_net_dataset_stats", "gk_net_train", "gk_net_last_error",
           "gk_net_collect", "gk_flux_record", "gk_flux_download", "gk_restrict", "gk_set_net", "gk_net_stats"]
NET_PARAMS = 5508


class GkGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nghost", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double)]


class GkPhys(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("g", C.c_double * 3),
                ("integrator", C.c_int), ("rc_order", C.c_int), ("stepmax", C.c_int),
                ("bc", C.c_int * 6), ("precision", C.c_int)]


# void hook(void* user, void* mirror): gk_set_bd_hook's host boundary function
BD_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


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
        L.gk_set_point_gravity.restype = st
        L.gk_set_point_gravity.argtypes = [vp] + [C.c_double] * 8
        L.gk_set_fixed_boundary.restype = st
        L.gk_set_fixed_boundary.argtypes = [vp, vp, ip]
        L.gk_set_immersed_shell.restype = st
        L.gk_set_immersed_shell.argtypes = [vp] + [C.c_double] * 4
        L.gk_set_time.restype = st
        L.gk_set_time.argtypes = [vp, C.c_double, ip]
        L.gk_set_bd_hook.restype = st
        L.gk_set_bd_hook.argtypes = [vp, ip, ip, BD_HOOK, vp, vp]
        L.gk_update_bd.restype = st
        L.gk_update_bd.argtypes = [vp]
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
        L.gk_cons_sums_noalloc.restype = st
        L.gk_cons_sums_noalloc.argtypes = [vp, ip, ip, ip, ip, vp, vp]
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
        fp, ll, llp = C.POINTER(C.c_float), C.c_longlong, C.POINTER(C.c_longlong)
        L.gk_net_create.restype = st
        L.gk_net_create.argtypes = [C.POINTER(vp)]
        L.gk_net_destroy.argtypes = [vp]
        for name in ("gk_net_set_params", "gk_net_get_params"):
            getattr(L, name).restype = st
            getattr(L, name).argtypes = [vp, fp]
        for name in ("gk_net_set_scales", "gk_net_get_scales"):
            getattr(L, name).restype = st
            getattr(L, name).argtypes = [vp, fp, fp]
        L.gk_net_set_kappa.restype = st
        L.gk_net_set_kappa.argtypes = [vp, C.c_float]
        L.gk_net_get_kappa.restype = st
        L.gk_net_get_kappa.argtypes = [vp, fp]
        L.gk_net_forward.restype = st
        L.gk_net_forward.argtypes = [vp, ll, fp, fp]
        L.gk_net_gradient.restype = st
        L.gk_net_gradient.argtypes = [vp, ll, fp, fp, fp, fp, dp]
        L.gk_net_adam.restype = st
        L.gk_net_adam.argtypes = [vp, fp, C.c_float]
        L.gk_net_reserve.restype = st
        L.gk_net_reserve.argtypes = [vp, ll]
        L.gk_net_add_samples.restype = st
        L.gk_net_add_samples.argtypes = [vp, ll, fp, fp, fp]
        L.gk_net_dataset.restype = st
        L.gk_net_dataset.argtypes = [vp, ll, ll, fp, fp, fp, llp]
        L.gk_net_dataset_stats.restype = st
        L.gk_net_dataset_stats.argtypes = [vp, dp, dp, dp]
        L.gk_net_train.restype = st
        L.gk_net_train.argtypes = [vp, ip, ip, ip, C.c_float, C.c_ulonglong, fp]
        L.gk_net_last_error.restype = C.c_char_p
        L.gk_net_last_error.argtypes = [vp]
        L.gk_net_collect.restype = st
        L.gk_net_collect.argtypes = [vp, vp, vp, ip, dp, llp]
        L.gk_flux_record.restype = st
        L.gk_flux_record.argtypes = [vp, ip]
        L.gk_flux_download.restype = st
        L.gk_flux_download.argtypes = [vp, dp]
        L.gk_restrict.restype = st
        L.gk_restrict.argtypes = [vp, vp]
        L.gk_set_net.restype = st
        L.gk_set_net.argtypes = [vp, vp]
        L.gk_net_stats.restype = st
        L.gk_net_stats.argtypes = [vp, llp, ip]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_grid(prob) -> GkGrid:
    g = GkGrid()
    g.nx, g.ny, g.nz, g.nghost = prob.nx, prob.ny, prob.nz, prob.ng
    (g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax) = prob.extent
    return g


def make_phys(prob, precision: int = 64) -> GkPhys:
    p = GkPhys()
    p.gamma, p.cfl = prob.gamma, prob.cfl
    for k in range(3):
        p.g[k] = prob.g[k]
    p.integrator, p.rc_order, p.stepmax = prob.integrator, prob.rc_order, prob.stepmax
    for k in range(4):
        p.bc[k] = prob.device_bc()[k]
    p.bc[4], p.bc[5] = prob.bcz
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
        for m in getattr(prob, "point", ()):
            self._check(self.L.gk_set_point_gravity(self.h, *[float(v) for v in m]))
        enrolled = prob.enrolled_state()
        if enrolled is not None:
            full = np.zeros(self._ref_shape())
            if prob.nz > 1:
                full[:] = enrolled
            else:
                full[:, prob.ng] = enrolled
            self.set_fixed_boundary(full, prob.fix_inner)
            self._enrolled_ref = full
        if prob.shell:
            self._check(self.L.gk_set_immersed_shell(self.h, *[float(v) for v in prob.shell]))
        if seg_rows:
            self._check(self.L.gk_set_tile(self.h, seg_rows))
        if upload and prob.cons is not None:
            if prob.nz > 1 and enrolled is not None:
                # the setup's array holds the enrolled ghost values too, and in 3D the cells of the
                # nearest z ghost planes that no face fills keep them
                self.upload_ref_layout(self._enrolled_ref)
            else:
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

    def _active_shape(self):
        p = self.prob
        return (5, p.nz, p.ny, p.nx) if p.nz > 1 else (5, p.ny, p.nx)

    def upload(self, cons: np.ndarray):
        a = np.ascontiguousarray(cons, np.float64)
        assert a.shape == self._active_shape()
        self._check(self.L.gk_upload_active(self.h, _dp(a)))
        self._bd_done = False

    def upload_ref_layout(self, ref: np.ndarray):
        """gk_upload_cons: a whole reference-layout state [5][NZ+2g][NY+2g][NX+2g] (ghost cells
        included) in the solver's precision."""
        dt = np.float64 if self.precision == 64 else np.float32
        a = np.ascontiguousarray(ref, dt)
        assert a.shape == self._ref_shape()
        self._check(self.L.gk_upload_cons(self.h, a.ctypes.data))
        self._bd_done = False

    def _begin(self):
        # main's UpdateBD after Setup (src/main.cpp:30) changes active cells only through the
        # immersed shell and the inner fixed region
        if not getattr(self, "_bd_done", True) and (self.prob.shell or self.prob.fix_inner):
            self.update_bd()

    def download(self) -> np.ndarray:
        out = np.empty(self._active_shape(), np.float64)
        self._check(self.L.gk_download_active(self.h, _dp(out)))
        return out

    def _ref_shape(self):
        ng = self.prob.ng
        return (5, max(self.prob.nz, 1) + 2 * ng, self.prob.ny + 2 * ng, self.prob.nx + 2 * ng)

    def download_ref_layout(self) -> np.ndarray:
        dt = np.float64 if self.precision == 64 else np.float32
        out = np.empty(self._ref_shape(), dt)
        self._check(self.L.gk_download_cons(self.h, out.ctypes.data))
        return out

    def set_fixed_boundary(self, ref: np.ndarray, fix_inner: bool = False):
        """Boundary::EnrollFixedBoundaryData: the ghost layers of `ref` (reference layout
        [5][NZ+2g][NY+2g][NX+2g]) become the values of the type-4 faces; with fix_inner its active
        cells in the inner region are restored before every stage."""
        dt = np.float64 if self.precision == 64 else np.float32
        a = np.ascontiguousarray(ref, dt)
        assert a.shape == self._ref_shape()
        self._check(self.L.gk_set_fixed_boundary(self.h, a.ctypes.data, int(bool(fix_inner))))

    def set_bd_hook(self, face_mask: int, band: int, fn):
        """gk_set_bd_hook: faces in face_mask (bit f: 0 x left .. 3 y right, and 4 z left, 5 z
        right on 3D grids) get their ghost layers from fn(mirror) before every stage, mirror being
        the reference-layout array [5][NZ+2g][NY+2g][NX+2g] in the solver's precision (NaN except
        the mirrored bands of `band` active layers next to those faces, which the library writes
        before each call)."""
        dt = np.float64 if self.precision == 64 else np.float32
        mirror = np.full(self._ref_shape(), np.nan, dt)
        hook = BD_HOOK(lambda user, ptr: fn(mirror))
        self._check(self.L.gk_set_bd_hook(self.h, face_mask, band, hook, None, mirror.ctypes.data))
        # the library holds both from now on (a failed call leaves the previous ones in place)
        self._mirror, self._hook_fn = mirror, hook

    def update_bd(self):
        """Boundary::UpdateBD's in-place part on the current state (immersed shell, inner region),
        as the reference's main does once after Setup; advance() does it before the first step."""
        self._check(self.L.gk_update_bd(self.h))
        self._bd_done = True

    def advance(self, dtoutput: float):
        self._begin()
        t = C.c_double()
        s = C.c_int()
        self._check(self.L.gk_advance(self.h, dtoutput, C.byref(t), C.byref(s)))
        return t.value, s.value

    def time(self):
        t, dt, s, ic = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self._check(self.L.gk_get_time(self.h, C.byref(t), C.byref(dt), C.byref(s), C.byref(ic)))
        return t.value, dt.value, s.value, bool(ic.value)

    def set_time(self, t: float, step: int):
        """Set the simulation time and step count (gk_set_time)."""
        self._check(self.L.gk_set_time(self.h, float(t), int(step)))

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
        self._begin()
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

    # learned flux correction
    def flux_record(self, on: bool = True):
        self._check(self.L.gk_flux_record(self.h, 1 if on else 0))

    def n_faces(self) -> int:
        nx, ny = self.prob.nx, self.prob.ny
        return (nx + 1) * ny + nx * (ny + 1)

    def flux_download(self):
        """Recorded sums as (fx [4, ny, nx+1], fy [4, ny+1, nx]): mass, m_x, m_y, E."""
        nx, ny = self.prob.nx, self.prob.ny
        acc = np.empty((4, self.n_faces()))
        self._check(self.L.gk_flux_download(self.h, _dp(acc)))
        nxf = (nx + 1) * ny
        return acc[:, :nxf].reshape(4, ny, nx + 1), acc[:, nxf:].reshape(4, ny + 1, nx)

    def restrict_from(self, fine: "Solver"):
        self._check(self.L.gk_restrict(fine.h, self.h))

    def set_net(self, net):
        self._check(self.L.gk_set_net(self.h, net.h if net is not None else None))
        self._net = net   # keep the network alive while attached

    def net_stats(self, reset: bool = False):
        out = (C.c_longlong * 3)()
        self._check(self.L.gk_net_stats(self.h, out, 1 if reset else 0))
        return list(out)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    """The flux-correction MLP on the GPU: parameters, Adam state and a device data set."""

    def __init__(self, params=None, in_scale=None, out_scale=None):
        self.L = lib()
        h = C.c_void_p()
        st = self.L.gk_net_create(C.byref(h))
        if st != 0:
            raise GaukukError(st, "gk_net_create failed")
        self.h = h
        if params is not None:
            self.set_params(params)
        if in_scale is not None or out_scale is not None:
            self.set_scales(np.ones(16) if in_scale is None else in_scale,
                            np.ones(4) if out_scale is None else out_scale)

    def _check(self, st):
        if st != 0:
            raise GaukukError(st, self.L.gk_net_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.gk_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        assert a.shape == (NET_PARAMS,)
        self._check(self.L.gk_net_set_params(self.h, _fp(a)))

    def params(self) -> np.ndarray:
        a = np.empty(NET_PARAMS, np.float32)
        self._check(self.L.gk_net_get_params(self.h, _fp(a)))
        return a

    def set_scales(self, in_scale, out_scale):
        i = np.ascontiguousarray(in_scale, np.float32)
        o = np.ascontiguousarray(out_scale, np.float32)
        self._check(self.L.gk_net_set_scales(self.h, _fp(i), _fp(o)))

    def scales(self):
        i, o = np.empty(16, np.float32), np.empty(4, np.float32)
        self._check(self.L.gk_net_get_scales(self.h, _fp(i), _fp(o)))
        return i, o

    @property
    def kappa(self) -> float:
        k = C.c_float()
        self._check(self.L.gk_net_get_kappa(self.h, C.byref(k)))
        return k.value

    @kappa.setter
    def kappa(self, value: float):
        self._check(self.L.gk_net_set_kappa(self.h, value))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((len(x), 4), np.float32)
        self._check(self.L.gk_net_forward(self.h, len(x), _fp(x), _fp(y)))
        return y

    def gradient(self, x, t, wt=None):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        wt = np.ascontiguousarray(np.ones(len(x)) if wt is None else wt, np.float32)
        g = np.empty(NET_PARAMS, np.float32)
        loss = C.c_double()
        self._check(self.L.gk_net_gradient(self.h, len(x), _fp(x), _fp(t), _fp(wt), _fp(g), C.byref(loss)))
        return loss.value, g

    def adam(self, grad, lr: float):
        g = np.ascontiguousarray(grad, np.float32)
        self._check(self.L.gk_net_adam(self.h, _fp(g), lr))

    def reserve(self, cap: int):
        self._check(self.L.gk_net_reserve(self.h, cap))

    def add_samples(self, x, t, wt=None):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        wt = np.ascontiguousarray(np.ones(len(x)) if wt is None else wt, np.float32)
        self._check(self.L.gk_net_add_samples(self.h, len(x), _fp(x), _fp(t), _fp(wt)))

    def size(self) -> int:
        n = C.c_longlong()
        self._check(self.L.gk_net_dataset(self.h, 0, 0, None, None, None, C.byref(n)))
        return n.value

    def dataset(self, start: int = 0, count: int = -1):
        n = self.size()
        count = n - start if count < 0 else count
        x = np.empty((count, 16), np.float32)
        t = np.empty((count, 4), np.float32)
        w = np.empty(count, np.float32)
        self._check(self.L.gk_net_dataset(self.h, start, count, _fp(x), _fp(t), _fp(w), None))
        return x, t, w

    def stats(self):
        mx, mt, w = np.zeros(16), np.zeros(4), C.c_double()
        self._check(self.L.gk_net_dataset_stats(self.h, _dp(mx), _dp(mt), C.byref(w)))
        return mx, mt, w.value

    def train(self, iters: int, lr: float, blocks: int = 256, tiles_per_wave: int = 1, seed: int = 0):
        log = np.empty(max(iters, 1), np.float32)
        self._check(self.L.gk_net_train(self.h, iters, blocks, tiles_per_wave, lr, seed, _fp(log)))
        return log[:iters]

    def collect(self, fine: Solver, coarse: Solver, steps: int = 1):
        """Append one sample per coarse face over `steps` coarse steps; returns (span, added)."""
        dt, added = C.c_double(), C.c_longlong()
        self._check(self.L.gk_net_collect(self.h, fine.h, coarse.h, steps, C.byref(dt), C.byref(added)))
        return dt.value, added.value

    def save(self, path: str, **meta):
        i, o = self.scales()
        np.savez(path, params=self.params(), in_scale=i, out_scale=o, kappa=self.kappa, **meta)

    @classmethod
    def load(cls, path: str) -> "Net":
        with np.load(path, allow_pickle=False) as z:
            net = cls(z["params"], z["in_scale"], z["out_scale"])
            if "kappa" in z:
                net.kappa = float(z["kappa"])
        return net


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
    """gk_selftest: the kernels' quotient/root cores against IEEE within their ulp bounds, and the
    DPP lane shifts against cross-lane shuffles; returns the number of failures."""
    bad = C.c_longlong()
    st = lib().gk_selftest(count, seed, emax, C.byref(bad))
    if st != 0:
        raise GaukukError(st, lib().gk_last_error(None).decode())
    return bad.value


def version() -> str:
    return lib().gk_version().decode()
