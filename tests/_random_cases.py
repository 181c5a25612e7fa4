"""Seeded random configurations shared by the GPU random-configuration test and the oracle's
golden checksums.

Each case draws grid sizes (1D, 2D and 3D, sizes around the 60-column strip width), face types
0-3 per face (0-4 in 2D, type 4 with random saved ghost data), the integrator, the reconstruction
order, constant gravity, point and binary masses, and in 2D the inner fixed region and an immersed shell,
plus a smooth random state with a pressure bump; every case runs 8 steps.
"""
import numpy as np

from gaukuk_amd import problems as P

CASES = 1000


def random_problem(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = rng.choice(["1d", "2d", "2d", "2d", "3d"])
    if kind == "1d":
        nx, ny, nz = int(rng.integers(20, 200)), 1, 1
    elif kind == "2d":
        nx, ny, nz = int(rng.integers(17, 140)), int(rng.integers(17, 100)), 1
    else:
        nx, ny, nz = int(rng.integers(12, 70)), int(rng.integers(8, 30)), int(rng.integers(3, 12))
    L = float(max(nx, ny, nz))
    ext = (0.0, nx / L, 0.0, max(ny, 1) / L, 0.0, max(nz, 1) / L if nz > 1 else 1.0)
    x = (np.arange(nx) + 0.5) / L
    y = (np.arange(ny) + 0.5) / L
    z = (np.arange(nz) + 0.5) / L
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 6)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y + ph[1]) + 0.1 * np.sin(2 * np.pi * Z + ph[5])
    vx = 0.4 * np.sin(2 * np.pi * Y + ph[2])
    vy = (0.3 * np.cos(2 * np.pi * X + ph[3])) if ny > 1 else 0 * X
    vz = (0.2 * np.sin(2 * np.pi * X + ph[4])) if nz > 1 else 0 * X
    c = rng.uniform(0.2, 0.8, 3)
    r2 = (X - c[0] * nx / L) ** 2 + (Y - c[1] * ny / L) ** 2 + ((Z - c[2] * nz / L) ** 2 if nz > 1 else 0)
    p = 0.5 + rng.uniform(2, 15) * np.exp(-r2 / 0.01)
    cons = P.prim_to_cons(rho, vx, vy, vz, p, 1.4)
    if nz == 1:
        cons = cons[:, 0]
    plane = kind == "2d"
    bc = tuple(int(v) for v in rng.integers(0, 5 if plane else 4, 4))
    g = tuple(float(v) for v in (rng.uniform(-0.2, 0.2, 3) if rng.random() < 0.4 else (0, 0, 0)))
    point = ()
    if kind != "1d" and rng.random() < 0.3:
        m0 = (float(rng.uniform(0.05, 0.2)), float(rng.uniform(0.3, 0.7) * nx / L), float(rng.uniform(0.3, 0.7) * ny / L),
              float(rng.uniform(-0.3, 0.3)), 0, 0, 0, float(rng.uniform(0.05, 0.15)))
        point = (m0,)
        if rng.random() < 0.5:   # a second mass: binary gravity
            point = (m0, (float(rng.uniform(0.01, 0.05)), float(rng.uniform(0.1, 0.9) * nx / L),
                          float(rng.uniform(0.1, 0.9) * ny / L), m0[3], 0, 0, 0, float(rng.uniform(0.02, 0.08))))
    fixed, fix_inner, shell = (), False, ()
    if plane and rng.random() < 0.25:
        fix_inner = True
    if plane and (4 in bc or fix_inner):
        # saved ghost-inclusive state: a second smooth field, so every saved cell differs
        g2 = 2
        xs = (np.arange(nx + 2 * g2) - g2 + 0.5) / L
        ys = (np.arange(ny + 2 * g2) - g2 + 0.5) / L
        Xs, Ys = np.meshgrid(xs, ys, indexing="xy")
        zs = np.zeros_like(Xs)
        fixed = P.prim_to_cons(1.2 + 0.2 * np.cos(3 * Xs + ph[1]), 0.3 * np.sin(2 * Ys), 0.1 + zs, zs,
                               0.8 + 0.1 * np.sin(4 * Xs * Ys), 1.4)
    if plane and g == (0, 0, 0) and not point and min(nx, ny) >= 40 and rng.random() < 0.4:
        shell = (float(rng.uniform(0.4, 0.6) * nx / L), float(rng.uniform(0.4, 0.6) * ny / L), 0.0,
                 float(rng.uniform(0.08, 0.15) * min(nx, ny) / L))
    return P.Problem(name=f"random{seed}", nx=nx, ny=ny, nz=nz, extent=ext, integrator=int(rng.integers(1, 4)),
                     rc_order=int(rng.integers(1, 3)), cfl=float(rng.uniform(0.2, 0.5)), stepmax=8, bc=bc,
                     bcz=tuple(int(v) for v in rng.integers(0, 4, 2)), g=g, point=point, cons=cons,
                     fixed=fixed, fix_inner=fix_inner, shell=shell)


def describe(seed, prob):
    return (f"random {seed}: {prob.nx}x{prob.ny}x{prob.nz} int={prob.integrator} rc={prob.rc_order} "
            f"bc={prob.bc} bcz={prob.bcz} g={prob.g} point={bool(prob.point)} inner={prob.fix_inner} "
            f"shell={bool(prob.shell)}")


def checksums(cons: np.ndarray) -> np.ndarray:
    """Per variable: sum |U| and a fixed pseudo-random projection of U (10 numbers a wrong cell,
    sign or index would change)."""
    u = np.asarray(cons, np.float64).reshape(5, -1)
    w = np.random.default_rng(7).uniform(-1.0, 1.0, u.shape[1])
    return np.concatenate([np.abs(u).sum(axis=1), u @ w])


def with_outputs(prob, seed):
    """The same problem run to a random end time in 1-6 outputs, without or with a step limit
    (the loop test, dt clipped to each output time, several advance calls)."""
    rng = np.random.default_rng(99000 + seed)
    tmax = float(rng.uniform(0.005, 0.08))
    nout = int(rng.integers(1, 7))
    stepmax = int(rng.choice([-1, -1, int(rng.integers(3, 30))]))
    return prob.with_(tmax=tmax, dtoutput=tmax / nout, stepmax=stepmax)
