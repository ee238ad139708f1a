"""The whole two-ring vortex step on the GPU (pmlib3d_amd.step.VortexStep:
P2M, order-6 solve, M2P, RK2 advance, P2M + remesh, all device-resident)
against the reference driver's own step (tests/golden/ref_step_*.npz) and
against the oracle composition of the same routines.  FP64 bar: same
particle set (positions bit-identical: remeshed cell centres), vorticity and
velocity within 1e-12 relative L-inf."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import oracle_vortex_step, rel_l2, rel_linf, ring_particles

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _run(n, bc, nsteps, min_vort, pos, vort):
    import pmlib3d_amd as pm
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, bound_cond=bc)
    st = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3,
                       capacity=max(pos.shape[1], n ** 3))
    st.load(pos, vort)
    for _ in range(nsteps):
        st.step(0.5 * h, min_vort)
    return st.fetch()


@pytest.mark.parametrize("name", ["free16", "mixed16"])
def test_step_matches_reference_golden(name):
    d = np.load(os.path.join(GOLD, "ref_step_%s.npz" % name), allow_pickle=False)
    n, bc, nsteps, min_vort = json.loads(str(d["spec"]))
    print("step", name, flush=True)
    pos, vort, vel = _run(n, tuple(bc), nsteps, min_vort, d["pos_in"], d["vort_in"])
    assert pos.shape == d["pos"].shape
    assert np.array_equal(pos, d["pos"])
    assert rel_linf(vort, d["vort"]) <= 1e-12
    assert rel_linf(vel, d["vel"]) <= 1e-12


def test_step_matches_oracle_32():
    n = 32
    h = 2.0 / n
    g = O.make_geom((n, n, n), 0, [-1.0] * 3, h, (0, 0, 0), 3)
    p = O.make_params(6, 1, 2.0)
    pos0, vort0 = ring_particles(n)
    ref = oracle_vortex_step(g, p, pos0, vort0, 0.5 * h, 1, 1e-3)
    got = _run(n, (0, 0, 0), 1, 1e-3, pos0, vort0)
    assert got[0].shape == ref[0].shape and np.array_equal(got[0], ref[0])
    assert rel_linf(got[1], ref[1]) <= 1e-12
    assert rel_linf(got[2], ref[2]) <= 1e-12


@pytest.mark.parametrize("fd_order", [2, 4])
def test_step_with_vorticity_rhs_matches_oracle(fd_order):
    """The step with stretching and diffusion (visc > 0): per stage the
    right-hand side on the mesh and its M2P into dvort, which the RK update
    of the vorticity then uses; a free patch with two ghost layers."""
    import pmlib3d_amd as pm
    n = 24
    h = 2.0 / n
    print("step with rhs, fd", fd_order, flush=True)
    g = O.make_geom((n, n, n), 2, [-1.0] * 3, h, (0, 0, 0), 3)
    p = O.make_params(6, 1, 2.0)
    pos0, vort0 = ring_particles(n)
    ref = oracle_vortex_step(g, p, pos0, vort0, 0.5 * h, 1, 1e-3, visc=0.05, fd_order=fd_order)
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, nghost=2)
    st = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3, capacity=(n + 4) ** 3,
                       visc=0.05, fd_order=fd_order)
    st.load(pos0, vort0)
    st.step(0.5 * h, 1e-3)
    got = st.fetch()
    assert got[0].shape == ref[0].shape and np.array_equal(got[0], ref[0])
    assert rel_linf(got[1], ref[1]) <= 1e-12
    assert rel_linf(got[2], ref[2]) <= 1e-12
    # the right-hand side changed the vorticity: the advection-only step differs
    adv = oracle_vortex_step(g, p, pos0, vort0, 0.5 * h, 1, 1e-3)
    assert adv[1].shape != ref[1].shape or rel_linf(adv[1], ref[1]) > 1e-6


def test_step_fp32_against_fp64_64():
    """The FP32 step (FP32 meshes, solve and interpolation with FP32
    weights, Kahan-added row partials in deposition, FP64 particles): the same particle set and velocities within 1e-5
    relative L2 of the FP64 step after one step."""
    import pmlib3d_amd as pm
    n = 64
    h = 2.0 / n
    pos0, vort0 = ring_particles(n)
    res = []
    for prec in (pm.FP64, pm.FP32):
        patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3)
        st = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3,
                           capacity=n ** 3, precision=prec)
        st.load(pos0, vort0)
        st.step(0.5 * h, 1e-3)
        res.append(st.fetch())
    (p64, w64, v64), (p32, w32, v32) = res
    print("particles fp64 %d fp32 %d" % (p64.shape[1], p32.shape[1]), flush=True)
    # remeshed particles sit on cell centres: match them by position (a cell
    # within rounding of the threshold may be kept by one precision only)
    assert abs(p64.shape[1] - p32.shape[1]) <= 2e-4 * p64.shape[1]
    k64 = {tuple(c): i for i, c in enumerate(p64.T)}
    pairs = [(k64[tuple(c)], j) for j, c in enumerate(p32.T) if tuple(c) in k64]
    assert len(pairs) >= 0.9998 * p64.shape[1]
    i64, i32 = (np.array(x) for x in zip(*pairs))
    assert rel_l2(v32[:, i32], v64[:, i64]) <= 1e-5
    assert rel_l2(w32[:, i32], w64[:, i64]) <= 1e-5
