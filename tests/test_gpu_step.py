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
from tests.helpers import oracle_vortex_step, rel_linf, ring_particles

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
