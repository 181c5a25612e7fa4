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
    s.set_bd_hook(1, 2, copy_outflow(prob.ng, 1))
    with pytest.raises(hip.GaukukError, match="point gravity"):
        s._check(s.L.gk_set_point_gravity(s.h, 1.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.1))
    # a later call may add faces but not drop a hooked one (it would keep stale ghost layers);
    # the refused call leaves the first hook working
    with pytest.raises(hip.GaukukError, match="keep every face"):
        s.set_bd_hook(2, 2, copy_outflow(prob.ng, 2))
    s.run_to_end()
    ref = hip.Solver(prob).run_to_end()
    assert s.step == ref.step == 3
    assert np.array_equal(s.download(), ref.download())
