"""Seeded initial conditions for the benchmark problems, in conserved variables.

Every implementation (HIP path, CPU oracle, reference build) starts from the arrays these
functions return, so the initial data is bit-identical everywhere.  Primitive -> conserved
follows EquationOfState::PrimToCons (src/eos/adiabatic.cpp:162-198) operation for operation:
``E = p*(1/(gamma-1)) + 0.5*rho*(vx*vx + vy*vy + vz*vz)``.

Random perturbations use SplitMix64 (Steele, Lea & Flood 2014) keyed by (seed, cell index),
so a value depends only on the seed and the cell, never on array order or thread count.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .gkio import cell_centres

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 output function applied to state ``x`` (uint64, wraps mod 2**64)."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, idx: np.ndarray, stream: int = 0) -> np.ndarray:
    """Uniform in [-1, 1): 53 high bits of SplitMix64(seed*2**32 + stream*2**28 + idx)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(32)) + (np.uint64(stream) << np.uint64(28)) + idx.astype(np.uint64)
    u = (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


@dataclass
class Problem:
    """Grid, physics and boundary parameters plus the active-region conserved state."""
    name: str
    nx: int
    ny: int
    extent: tuple            # (xmin, xmax, ymin, ymax, zmin, zmax)
    gamma: float = 1.4
    cfl: float = 0.4
    integrator: int = 2      # 1 Euler, 2 RK2, 3 RK3
    rc_order: int = 2
    stepmax: int = -1
    tmax: float = 1e30
    dtoutput: float = 1e30
    bc: tuple = (1, 1, 1, 1)  # xleft, xright, yleft, yright
    nz: int = 1               # active z planes; > 1 is the 3D path (cons is then (5, nz, ny, nx))
    bcz: tuple = (1, 1)       # zleft, zright (read only in 3D)
    g: tuple = (0.0, 0.0, 0.0)
    # point masses, each (gm, x, y, z, vx, vy, vz, rs) as SourceTerm::EnrollPointGravity takes them
    # (src/source_term/source.cpp:33-67); two masses = binary gravity
    point: tuple = ()
    # fixed (type 4) boundaries: what the setup's ghost cells hold when it enrolls the fixed data
    # (Boundary::EnrollFixedBoundaryData, src/boundary/fixed.cpp:242-269): one conserved state for
    # every ghost cell (5 values) or a ghost-inclusive array [5][NY+2g][NX+2g] ([5][NZ+2g][NY+2g][NX+2g]
    # in 3D)
    fixed: tuple = ()
    # self-defined faces (Boundary::EnrollSelfDefineBD*, src/boundary/boundary.hpp:28-45) of the
    # state-independent kind: (face, conserved state) pairs, face 0..3 = xleft, xright, yleft,
    # yright, whose function writes that state into the face's ghost cells (the inflow of
    # src/setup/setup_solid_body.cpp)
    self_faces: tuple = ()
    # EnrollFixedBoundaryData(..., fixInner = true): cells of the inner region are reset to the
    # enrolled state before every stage (src/boundary/fixed.cpp:181-207)
    fix_inner: bool = False
    # immersed reflective sphere (x, y, z, r) (Boundary::SetupImmersedShell, src/boundary/immersed.cpp)
    shell: tuple = ()
    ng: int = 2
    cons: np.ndarray = field(default=None, repr=False)  # (5, ny, nx) float64

    def with_(self, **kw) -> "Problem":
        return replace(self, **kw)

    def ghost_shape(self) -> tuple:
        """The reference-layout state shape: [5][NY+2g][NX+2g] in 2D, [5][NZ+2g][NY+2g][NX+2g] in 3D."""
        g = self.ng
        plane = (self.ny + 2 * g, self.nx + 2 * g)
        return (5, self.nz + 2 * g) + plane if self.nz > 1 else (5,) + plane

    def fixed_full(self):
        """The ghost-inclusive state (ghost_shape()) the setup enrolls as fixed-face data, or None."""
        if len(self.fixed) == 0:
            return None
        shape = self.ghost_shape()
        f = np.asarray(self.fixed, np.float64)
        if f.ndim == 1:
            return np.broadcast_to(f.reshape((5,) + (1,) * (len(shape) - 1)), shape).copy()
        if f.shape != shape:
            raise ValueError(f"fixed-face data has shape {f.shape}, expected {shape}")
        return f.copy()

    def device_bc(self) -> tuple:
        """Face types as the device runs them: a state-independent self-defined face is a fixed face."""
        bc = list(self.bc)
        for f, _ in self.self_faces:
            bc[f] = 4
        return tuple(bc)

    def enrolled_state(self):
        """The ghost-inclusive state (ghost_shape()) the setup holds when it enrolls its fixed data
        (active cells = cons; ghosts of type-4 faces from `fixed`; ghosts of self-defined faces hold
        their state), or None when nothing is enrolled."""
        if not (len(self.fixed) or self.self_faces or self.fix_inner):
            return None
        g = self.ng
        full = self.fixed_full()
        if full is None:
            full = np.zeros(self.ghost_shape())
        if self.nz > 1:
            if self.self_faces:
                raise ValueError("self-defined faces are given for 2D problems only")
            full[:, g:g + self.nz, g:g + self.ny, g:g + self.nx] = self.cons
            return full
        for f, state in self.self_faces:
            st = np.asarray(state, np.float64)[:, None, None]
            if f == 0:
                full[:, g:g + self.ny, :g] = st
            elif f == 1:
                full[:, g:g + self.ny, g + self.nx:] = st
            elif f == 2:
                full[:, :g, g:g + self.nx] = st
            else:
                full[:, g + self.ny:, g:g + self.nx] = st
        full[:, g:g + self.ny, g:g + self.nx] = self.cons
        return full

    def _point_keys(self) -> dict:
        keys = {}
        for f, state in self.self_faces:
            keys[f"selfbd_{f}"] = 1
            for q, v in enumerate(state):
                keys[f"sd{f}_{q}"] = v
        if self.fix_inner:
            keys["fixinner"] = 1
        if self.shell:
            keys["immersed"] = 1
            for name, v in zip(("x", "y", "z", "r"), self.shell):
                keys[f"shell_{name}"] = v
        if len(self.fixed):
            if np.ndim(self.fixed) == 1:
                keys["fixedbd"] = 1
                for q, v in enumerate(self.fixed):
                    keys[f"fix_{q}"] = v
            else:
                # the array comes from the file named by $GK_FIXED_FILE: one plane (2D) or the
                # whole ghost-inclusive volume (3D)
                keys["fixedbd"] = 3 if np.ndim(self.fixed) == 4 else 2
        if not self.point:
            return keys
        keys["npoint"] = len(self.point)
        for q, m in enumerate(self.point):
            for name, v in zip(("gm", "x", "y", "z", "vx", "vy", "vz", "rs"), m):
                keys[f"pg{q}_{name}"] = v
        return keys

    def in_values(self) -> dict:
        """Keys of the reference's `.in` grammar describing this problem."""
        x0, x1, y0, y1, z0, z1 = self.extent
        return {
            "gamma": self.gamma, "tmax": self.tmax, "stepmax": self.stepmax, "CFL": self.cfl,
            "integrator": self.integrator, "RcOrder": self.rc_order, "dtoutput": self.dtoutput,
            "dataType": 0, "xmin": x0, "xmax": x1, "ymin": y0, "ymax": y1, "zmin": z0, "zmax": z1,
            "NX": self.nx, "NY": self.ny, "NZ": self.nz, "NGhost": self.ng,
            "xleft": self.bc[0], "xright": self.bc[1], "yleft": self.bc[2], "yright": self.bc[3],
            "zleft": self.bcz[0], "zright": self.bcz[1], "gx": self.g[0], "gy": self.g[1], "gz": self.g[2],
            **self._point_keys(),
        }


def prim_to_cons(rho, vx, vy, vz, p, gamma: float) -> np.ndarray:
    gm1rec = 1.0 / (gamma - 1.0)
    e = p * gm1rec + (0.5 * rho) * (vx * vx + vy * vy + vz * vz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, e]).astype(np.float64)


def _centres(extent, nx, ny, ng):
    x = cell_centres(extent[0], extent[1], nx, ng)[ng:ng + nx]
    y = cell_centres(extent[2], extent[3], ny, ng)[ng:ng + ny]
    return np.meshgrid(x, y, indexing="xy")  # arrays of shape (ny, nx)


def sod(nx: int = 256, **kw) -> Problem:
    """Sod shock tube on [0,1]: (rho, p) = (1, 1) for x < 0.5 else (0.125, 0.1), v = 0.

    Setup logic of src/setup/setup_shock_tube.cpp:10-48 with x0 = 0.5; y/z extents of 1 keep
    drmin = dx (src/sim.cpp:35)."""
    ext = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    x, _ = _centres(ext, nx, 1, 2)
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    base = dict(name="sod", nx=nx, ny=1, extent=ext, integrator=2, rc_order=2, cfl=0.4,
                tmax=0.2, dtoutput=0.2, bc=(0, 0, 0, 0))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, z, z, z, p, base.get("gamma", 1.4)), **base)


def kelvin_helmholtz(n: int = 2048, ny: int | None = None, seed: int = 0, amp: float = 0.01,
                     noise: float = 0.01, noise_cells: int | None = None, **kw) -> Problem:
    """Periodic KH on [0,1]^2: rho = 2, vx = +0.5 for |y-0.5| < 0.25, else rho = 1, vx = -0.5;
    p = 2.5; vy = amp*sin(4 pi x) + noise*U(-1,1) (SplitMix64 keyed by j*nx + i).  With
    noise_cells = m (dividing n and ny) the noise is drawn on an m-wide grid of the same domain and
    held constant over each of its cells, so grids of different resolution share one perturbation."""
    ny = n if ny is None else ny
    ext = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    x, y = _centres(ext, n, ny, 2)
    inner = np.abs(y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    if noise_cells is None or noise_cells == n:
        idx = np.arange(n * ny, dtype=np.uint64).reshape(ny, n)
    else:
        m = noise_cells
        mx, my = m, m * ny // n
        if n % m or ny % my or my * n != m * ny:
            raise ValueError("noise_cells must divide the grid")
        ci = np.arange(n) // (n // mx)
        cj = np.arange(ny) // (ny // my)
        idx = (cj[:, None] * mx + ci[None, :]).astype(np.uint64)
    vy = amp * np.sin(4.0 * np.pi * x) + noise * uniform_pm1(seed, idx)
    p = np.full_like(rho, 2.5)
    base = dict(name="kh", nx=n, ny=ny, extent=ext, integrator=3, rc_order=2, cfl=0.4,
                stepmax=200, bc=(1, 1, 1, 1))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, vx, vy, np.zeros_like(rho), p, base.get("gamma", 1.4)), **base)


def rayleigh_taylor(nx: int = 1024, ny: int = 3072, gy: float = -0.1, p0: float = 2.5,
                    amp: float = 0.01, **kw) -> Problem:
    """RT on [0,0.5]x[0,1.5], rho = 2 above y = 0.75 else 1, hydrostatic
    p = p0 + gy*rho*(y - 0.75) (the form of src/setup/setup_rt.cpp:54 with the interface at
    0.75), vy = amp*(1+cos 4 pi x)(1+cos 3 pi (y-0.75))/4; periodic x, reflecting y, gravity gy."""
    ext = (0.0, 0.5, 0.0, 1.5, 0.0, 1.0)
    x, y = _centres(ext, nx, ny, 2)
    rho = np.where(y > 0.75, 2.0, 1.0)
    p = p0 + gy * rho * (y - 0.75)
    vy = amp * (1.0 + np.cos(4.0 * np.pi * x)) * (1.0 + np.cos(3.0 * np.pi * (y - 0.75))) / 4.0
    z = np.zeros_like(rho)
    base = dict(name="rt", nx=nx, ny=ny, extent=ext, integrator=2, rc_order=2, cfl=0.4,
                bc=(1, 1, 2, 2), g=(0.0, gy, 0.0))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, z, vy, z, p, base.get("gamma", 1.4)), **base)


def blast(nx: int, ny: int, p_ratio: float = 50.0, radius: float = 0.1, center=None,
          bc: int = 1, rho0: float = 1.0, p_amb: float = 0.1, **kw) -> Problem:
    """2D circular overpressure: rho = rho0, p = p_amb*p_ratio inside r < radius, else p_amb,
    v = 0, on [0, nx/L] x [0, ny/L] with L = max(nx, ny) so dx = dy = 1/L."""
    L = float(max(nx, ny))
    ext = (0.0, nx / L, 0.0, ny / L, 0.0, 1.0)
    cx, cy = center if center is not None else (0.5 * nx / L, 0.5 * ny / L)
    x, y = _centres(ext, nx, ny, 2)
    r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    rho = np.full_like(x, rho0)
    p = np.where(r < radius, p_amb * p_ratio, p_amb)
    z = np.zeros_like(rho)
    bcs = (bc,) * 4 if isinstance(bc, int) else tuple(bc)
    base = dict(name="blast", nx=nx, ny=ny, extent=ext, integrator=2, rc_order=2, cfl=0.4, bc=bcs)
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, z, z, z, p, base.get("gamma", 1.4)), **base)


def solid_body(nx: int = 1024, ny: int | None = None, rho: float = 1.0, p: float = 1.0, vx_in: float = 4.0,
               x0: float = -3.0, y0: float = 0.0, z0: float = 0.0, r_solid: float = 1.0,
               extent=(-5.0, 5.0, -5.0, 5.0, -0.5, 0.5), **kw) -> Problem:
    """Supersonic flow past a reflective cylinder, as src/setup/setup_solid_body.cpp with
    input/solid_body.in: uniform (rho, vx_in, p), an x-left face that injects that state (the
    setup's self-defined face), outflow elsewhere, an immersed shell of radius r_solid at (x0, y0, z0);
    RK2, CFL 0.5."""
    ny = nx if ny is None else ny
    x, _ = _centres(extent, nx, ny, 2)
    one = np.ones_like(x)
    gamma = kw.get("gamma", 1.4)
    cons = prim_to_cons(rho * one, vx_in * one, 0 * one, 0 * one, p * one, gamma)
    inflow = (rho, rho * vx_in, 0.0, 0.0, 0.5 * rho * (vx_in * vx_in) + p * (1.0 / (gamma - 1.0)))
    base = dict(name="solid_body", nx=nx, ny=ny, extent=tuple(extent), integrator=2, rc_order=2, cfl=0.5,
                bc=(3, 3, 3, 3), self_faces=((0, inflow),), shell=(x0, y0, z0, r_solid))
    base.update(kw)
    return Problem(cons=cons, **base)


def blast3d(nx: int, ny: int, nz: int, p_ratio: float = 30.0, radius: float = 0.15, center=None,
            bc=1, bcz=(1, 1), rho0: float = 1.0, p_amb: float = 0.1, **kw) -> Problem:
    """3D spherical overpressure on [0, nx/L] x [0, ny/L] x [0, nz/L], L = max(nx, ny, nz), with a
    density gradient along x so that every axis carries structure."""
    L = float(max(nx, ny, nz))
    ext = (0.0, nx / L, 0.0, ny / L, 0.0, nz / L)
    cx, cy, cz = center if center is not None else (0.5 * nx / L, 0.5 * ny / L, 0.5 * nz / L)
    x = (np.arange(nx) + 0.5) / L
    y = (np.arange(ny) + 0.5) / L
    z = (np.arange(nz) + 0.5) / L
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2)
    rho = rho0 * (1.0 + 0.2 * X)
    p = np.where(r < radius, p_amb * p_ratio, p_amb)
    zero = np.zeros_like(rho)
    bcs = (bc,) * 4 if isinstance(bc, int) else tuple(bc)
    base = dict(name="blast3d", nx=nx, ny=ny, nz=nz, extent=ext, integrator=2, rc_order=2, cfl=0.4, bc=bcs,
                bcz=tuple(bcz))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, zero, zero, zero, p, base.get("gamma", 1.4)), **base)


def sphere_flow3d(nx: int, ny: int, nz: int, **kw) -> Problem:
    """An immersed sphere of radius 0.2 at (0.55, 0.45, 0.4) on [0, nx/L] x [0, ny/L] x [0, nz/L],
    L = max(nx, ny, nz), in the oblique flow (0.5, 0.1, -0.05) with an overpressure blob upstream;
    periodic x and z, reflective y left, outflow y right, RK3."""
    prob = blast3d(nx, ny, nz, bc=(1, 1, 2, 3), bcz=(1, 1))
    L = float(max(nx, ny, nz))
    Z, Y, X = np.meshgrid((np.arange(nz) + 0.5) / L, (np.arange(ny) + 0.5) / L, (np.arange(nx) + 0.5) / L,
                          indexing="ij")
    one = np.ones_like(X)
    p = np.where((X - 0.2) ** 2 + (Y - 0.45) ** 2 + (Z - 0.4) ** 2 < 0.01, 1.5, 0.6)
    cons = prim_to_cons(1.0 + 0.1 * X, 0.5 * one, 0.1 * one, -0.05 * one, p, prob.gamma)
    base = dict(name="sphere3d", integrator=3, cons=cons, shell=(0.55, 0.45, 0.4, 0.2))
    base.update(kw)
    return prob.with_(**base)


def shear3d(n: int, nz: int, seed: int = 0, amp: float = 0.02, **kw) -> Problem:
    """3D Kelvin-Helmholtz: the configuration's shear layer (rho 2 / 1, vx -+0.5, p 2.5) on
    [0,1]^2 x [0, nz/n], with a seeded perturbation in vy and vz."""
    ext = (0.0, 1.0, 0.0, 1.0, 0.0, nz / n)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid((np.arange(nz) + 0.5) / n, x, x, indexing="ij")
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    idx = np.arange(n * n * nz).reshape(nz, n, n)
    vy = amp * np.sin(4 * np.pi * X) + amp * uniform_pm1(seed, idx, 0)
    vz = amp * uniform_pm1(seed, idx, 1)
    p = np.full_like(rho, 2.5)
    base = dict(name="shear3d", nx=n, ny=n, nz=nz, extent=ext, integrator=3, rc_order=2, cfl=0.4, bc=(1, 1, 1, 1),
                bcz=(1, 1))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, vx, vy, vz, p, base.get("gamma", 1.4)), **base)


def riemann_1d(nx: int, left, right, x0: float = 0.5, **kw) -> Problem:
    """1D Riemann problem on [0,1]: left/right = (rho, vx, p)."""
    ext = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    x, _ = _centres(ext, nx, 1, 2)
    L = x < x0
    rho = np.where(L, left[0], right[0])
    vx = np.where(L, left[1], right[1])
    p = np.where(L, left[2], right[2])
    z = np.zeros_like(rho)
    base = dict(name="riemann", nx=nx, ny=1, extent=ext, integrator=2, rc_order=2, cfl=0.4, bc=(0, 0, 0, 0))
    base.update(kw)
    return Problem(cons=prim_to_cons(rho, vx, z, z, p, base.get("gamma", 1.4)), **base)


def ensemble(count: int = 512, seed: int = 0, nmin: int = 64, nmax: int = 512) -> list:
    """Seeded ensemble of 2D blast problems: nx, ny ~ U{nmin..nmax} independently; p ratio
    U(10, 100); radius U(0.05, 0.2); centre at the domain centre; BC periodic or outflow by a
    seeded coin.  All FP-free choices come from SplitMix64 streams of ``seed``."""
    k = np.arange(count, dtype=np.uint64)
    span = nmax - nmin + 1
    nxs = nmin + (splitmix64((np.uint64(seed) << np.uint64(32)) + k * np.uint64(8) + np.uint64(1)) % np.uint64(span)).astype(np.int64)
    nys = nmin + (splitmix64((np.uint64(seed) << np.uint64(32)) + k * np.uint64(8) + np.uint64(2)) % np.uint64(span)).astype(np.int64)
    ratio = 10.0 + 90.0 * (0.5 * (uniform_pm1(seed, k, stream=3) + 1.0))
    rad = 0.05 + 0.15 * (0.5 * (uniform_pm1(seed, k, stream=4) + 1.0))
    coin = (splitmix64((np.uint64(seed) << np.uint64(32)) + k * np.uint64(8) + np.uint64(5)) & np.uint64(1)).astype(np.int64)
    out = []
    for q in range(count):
        bc = 1 if coin[q] == 0 else 3
        out.append(blast(int(nxs[q]), int(nys[q]), p_ratio=float(ratio[q]), radius=float(rad[q]), bc=bc,
                         name=f"blast{q}"))
    return out
