"""gk_set_bd_hook: self-defined faces whose ghost values depend on the state, refreshed from a
host function before every stage (Boundary::EnrollSelfDefineBD*, reference
src/boundary/boundary.hpp:28-45, called in every UpdateBD, src/boundary/boundary.cpp:127-135).

A hook that copies the edge cell into every ghost cell is the reference's copy-outflow face
(its documented template, src/setup/setup_example.cpp:128-351, and the built-in type 0,
src/boundary/outflow_copy.cpp): run through the hook on all four faces, the solver must give the
built-in type-0 result bit for bit (no gravity: the ghost values are the same copies), for every
integrator and both precisions.  With constant gravity the hook sees the state after the first
half-kick, kicked on the host, so the two agree to rounding (1e-12).  The reference-built fixture
of tests/golden/setup_statebd.cpp (state-dependent faces of four different kinds) is checked
through the host driver in test_gpu_reference_setups.py.

In 3D a hook that mirrors the planes nearest a z face with the normal momentum negated is the
built-in reflective z face (src/boundary/reflective.cpp:168-232): hooks on all six faces must give
the built-in result bit for bit.  With point or binary gravity the band is kicked on the host with
the stage kernel's formula, so hooked and built-in faces agree to rounding.  The reference-built
fixtures of tests/golden/setup_statebd3d.cpp (six state-dependent faces in 3D, and under a point
mass in 3D and 2D) are checked through the host driver in test_gpu_reference_setups.py.
"""
import numpy as np
import pytest

from gaukuk_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from gaukuk_amd import hip as H
    return H


def copy_outflow(g, mask):
    """Ghost cells of the faces in mask take the edge cell of their row/column (active range)."""
    def fn(m):
        a = m[:, g]   # the active plane of the 2D reference layout [5][ly][lx]
        ny, nx = a.shape[1] - 2 * g, a.shape[2] - 2 * g
        rows = slice(g, g + ny)
        cols = slice(g, g + nx)
        if mask & 1:
            a[:, rows, :g] = a[:, rows, g:g + 1]
        if mask & 2:
            a[:, rows, g + nx:] = a[:, rows, g + nx - 1:g + nx]
        if mask & 4:
            a[:, :g, cols] = a[:, g:g + 1, cols]
        if mask & 8:
            a[:, g + ny:, cols] = a[:, g + ny - 1:g + ny, cols]
    return fn


def faces3d(g, mask, yl_reflect=False):
    """3D mirror [5][nz+2g][ny+2g][nx+2g]: x and y faces copy the edge cell (type 0), z faces
    mirror the nearest planes with MTZ negated (type 2); with yl_reflect the y left face mirrors
    its rows with MTY negated instead."""
    A = slice(g, -g)

    def fn(m):
        if mask & 1:
            m[:, A, A, :g] = m[:, A, A, g:g + 1]
        if mask & 2:
            m[:, A, A, -g:] = m[:, A, A, -g - 1:-g]
        if mask & 4:
            if yl_reflect:
                for d in range(g):
                    m[:, A, g - 1 - d, A] = m[:, A, g + d, A]
                    m[2, A, g - 1 - d, A] *= -1
            else:
                m[:, A, :g, A] = m[:, A, g:g + 1, A]
        if mask & 8:
            m[:, A, -g:, A] = m[:, A, -g - 1:-g, A]
        nz = m.shape[1] - 2 * g
        for d in range(g):
            if mask & 16:
                m[:, g - 1 - d, A, A] = m[:, g + d, A, A]
                m[3, g - 1 - d, A, A] *= -1
            if mask & 32:
                m[:, g + nz + d, A, A] = m[:, g + nz - 1 - d, A, A]
                m[3, g + nz + d, A, A] *= -1
    return fn


def rel_l1(a, b):
    return [np.abs(a[n] - b[n]).sum() / max(np.abs(b[n]).sum(), 1e-300) for n in range(5)]


@pytest.mark.parametrize("integrator,precision", [(3, 64), (2, 32), (1, 64)])
def test_3d_hooks_equal_builtin_faces(hip, integrator, precision):
    """All six faces hooked (copy x/y, reflective z) equal the built-in types 0 and 2 bit for bit."""
    prob = P.blast3d(40, 36, 30, p_ratio=20.0, radius=0.2, center=(0.3, 0.45, 0.25), bc=(0, 0, 0, 0),
                     bcz=(2, 2), integrator=integrator, stepmax=20)
    print(f"blast3d 40x36x30 integrator={integrator} FP{precision}", flush=True)
    ref = hip.Solver(prob, precision=precision).run_to_end()
    s = hip.Solver(prob, precision=precision)
    s.set_bd_hook(63, 2, faces3d(prob.ng, 63))
    s.run_to_end()
    assert s.step == ref.step == 20 and s.t == ref.t
    assert np.array_equal(s.download(), ref.download())


def test_3d_hooked_y_face_beside_type0_z_face(hip):
    """A type-0 z face registers the y copy, run after the y faces: the y left hook (a mirror here)
    is overwritten by the copy, so the run equals the built-in run with a copy y left face."""
    prob = P.blast3d(36, 40, 24, p_ratio=20.0, radius=0.2, center=(0.4, 0.2, 0.3), bc=(0, 0, 0, 0), bcz=(0, 2),
                     integrator=3, stepmax=20)
    print("blast3d 36x40x24 bcz=(0, 2)", flush=True)
    ref = hip.Solver(prob).run_to_end()
    s = hip.Solver(prob)
    s.set_bd_hook(12 | 32, 2, faces3d(prob.ng, 12 | 32, yl_reflect=True))
    s.run_to_end()
    assert s.step == ref.step == 20
    assert np.array_equal(s.download(), ref.download())


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_hooks_with_point_gravity(hip, dims):
    """Point (3D) and binary (2D) gravity: the first stage's bands are kicked on the host before the
    hook runs, as the stage kernel kicks U^n."""
    if dims == "3d":
        prob = P.blast3d(36, 40, 28, p_ratio=10.0, radius=0.2, bc=(0, 0, 0, 0), bcz=(2, 2), integrator=3,
                         stepmax=30, point=((0.05, 0.45, 0.5, 0.4, 0.0, 0.0, 0.0, 0.1),), g=(0.0, -0.2, 0.1))
        mask, fn = 63, faces3d(2, 63)
    else:
        prob = P.blast(64, 72, p_ratio=10.0, radius=0.15, bc=(0, 0, 0, 0), integrator=2, stepmax=40,
                       point=((0.05, 0.4, 0.5, 0.0, 0.0, 0.0, 0.0, 0.08), (0.02, 0.65, 0.5, 0.0, 0.0, 0.0, 0.0, 0.05)))
        mask, fn = 15, copy_outflow(prob.ng, 15)
    print(f"{dims} point gravity", flush=True)
    ref = hip.Solver(prob).run_to_end()
    s = hip.Solver(prob)
    s.set_bd_hook(mask, 2, fn)
    s.run_to_end()
    assert s.step == ref.step
    err = rel_l1(s.download(), ref.download())
    print("rel L1", err, flush=True)
    assert max(err) <= 1e-12, err


def test_3d_hook_fills_output_ghosts(hip):
    """gk_download_cons fills a hooked z face's ghost planes by calling the hook on the final state."""
    prob = P.blast3d(32, 28, 20, bc=(0, 0, 0, 0), bcz=(1, 2), stepmax=5)
    s = hip.Solver(prob)
    s.set_bd_hook(32, 2, faces3d(prob.ng, 32))
    s.run_to_end()
    full = s.download_ref_layout()
    g, nz = prob.ng, prob.nz
    A = slice(g, -g)
    for d in range(g):
        ghost, src = full[:, g + nz + d, A, A], full[:, g + nz - 1 - d, A, A]
        assert np.array_equal(ghost[[0, 1, 2, 4]], src[[0, 1, 2, 4]])
        assert np.array_equal(ghost[3], -src[3])


@pytest.mark.parametrize("integrator,precision", [(3, 64), (2, 64), (1, 64), (3, 32)])
def test_copy_hook_equals_builtin_copy_faces(hip, integrator, precision):
    prob = P.blast(70, 62, p_ratio=30.0, radius=0.2, center=(0.2, 0.3), bc=(0, 0, 0, 0), integrator=integrator,
                   stepmax=40)
    print(f"blast 70x62 integrator={integrator} FP{precision}", flush=True)
    ref = hip.Solver(prob, precision=precision).run_to_end()
    s = hip.Solver(prob, precision=precision)
    s.set_bd_hook(15, 2, copy_outflow(prob.ng, 15))
    s.run_to_end()
    assert s.step == ref.step == 40 and s.t == ref.t
    assert np.array_equal(s.download(), ref.download())


def test_copy_hook_with_gravity(hip):
    """Constant gravity: the first stage's band is kicked on the host before the hook runs."""
    prob = P.blast(64, 96, p_ratio=10.0, radius=0.15, bc=(0, 0, 0, 0), integrator=3, stepmax=40)
    prob = prob.with_(g=(0.0, -0.5, 0.0))
    ref = hip.Solver(prob).run_to_end()
    s = hip.Solver(prob)
    s.set_bd_hook(12, 2, copy_outflow(prob.ng, 12))   # y faces hooked, x faces built-in copy
    s.run_to_end()
    assert s.step == ref.step
    a, b = s.download(), ref.download()
    err = [np.abs(a[n] - b[n]).sum() / max(np.abs(b[n]).sum(), 1e-300) for n in range(5)]
    print("rel L1", err, flush=True)
    assert max(err) <= 1e-12, err


def test_hook_fills_output_ghosts(hip):
    """gk_download_cons fills the hooked faces' ghost cells by calling the hook on the final state."""
    prob = P.blast(40, 36, p_ratio=10.0, radius=0.2, bc=(0, 2, 1, 1), stepmax=8)
    s = hip.Solver(prob)
    s.set_bd_hook(1, 2, copy_outflow(prob.ng, 1))
    s.run_to_end()
    full = s.download_ref_layout()
    g = prob.ng
    act = full[:, g, g:g + prob.ny, :]
    assert np.array_equal(act[:, :, 0], act[:, :, g]) and np.array_equal(act[:, :, 1], act[:, :, g])


def test_hook_reading_beyond_band_fails(hip):
    """A face function that reads deeper than the mirrored band sees NaN and the step fails."""
    prob = P.blast(40, 36, bc=(0, 0, 0, 0), stepmax=3)
    g = prob.ng

    def deep(m):
        a = m[:, g]
        a[:, g:-g, :g] = a[:, g:-g, g + 1:g + 2]   # reads the second active column
    s = hip.Solver(prob)
    s.set_bd_hook(1, 1, deep)    # band of one layer
    with pytest.raises(hip.GaukukError, match="non-finite ghost"):
        s.run_steps(1)


def test_hook_rejects_unsupported(hip):
    prob = P.blast(40, 36, bc=(0, 0, 0, 0), stepmax=3)
    s = hip.Solver(prob)
    with pytest.raises(hip.GaukukError, match="face_mask"):
        s.set_bd_hook(16, 2, lambda m: None)
    with pytest.raises(hip.GaukukError, match="band"):
        s.set_bd_hook(1, 0, lambda m: None)
    with pytest.raises(hip.GaukukError, match="NZ > 1"):
        s.set_bd_hook(1 | 16, 2, lambda m: None)   # z faces of a 2D grid
    s.set_bd_hook(1, 2, copy_outflow(prob.ng, 1))
    # a later call may add faces but not drop a hooked one (it would keep stale ghost layers);
    # the refused call leaves the first hook working
    with pytest.raises(hip.GaukukError, match="keep every face"):
        s.set_bd_hook(2, 2, copy_outflow(prob.ng, 2))
    s.run_to_end()
    ref = hip.Solver(prob).run_to_end()
    assert s.step == ref.step == 3
    assert np.array_equal(s.download(), ref.download())
