"""The graph-recorded vortex step (pmlib3d_amd.GraphVortexStep, pmgpu_vstep_*
through the C ABI: the particle count in device memory, capacity-sized
launches, the whole RK2 step plus remesh recorded once and replayed) against
the eager step (VortexStep, which the reference golden steps and the oracle
pin): the same particle set and bit-identical positions, vorticity and
velocity after every step, for the free and the mixed-BC two-ring cases of
the reference driver (tests/golden/ref_step_*.npz) and at 32^3, in FP64 and
FP32; and the recorded graph against its own eager enqueue.  References:
vortex_codes/vortex_structures/source/main.f:453-505 (the client loop)."""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_l2, ring_particles

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _patch(n, bc):
    import pmlib3d_amd as pm
    h = 2.0 / n
    return pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, bound_cond=bc)


def _compare(n, bc, nsteps, min_vort, pos, vort, precision, graph=True):
    import pmlib3d_amd as pm
    h = 2.0 / n
    patch = _patch(n, bc)
    cap = max(pos.shape[1], n ** 3)
    eager = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3, capacity=cap,
                          precision=precision)
    eager.load(pos, vort)
    gs = pm.GraphVortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3, capacity=cap,
                            precision=precision)
    gs.load(pos, vort)
    for k in range(nsteps):
        eager.step(0.5 * h, min_vort)
        gs.step(0.5 * h, min_vort, graph=graph)
        a, b = eager.fetch(), gs.fetch()
        print("step", k, "particles eager %d graph %d" % (a[0].shape[1], b[0].shape[1]), flush=True)
        assert a[0].shape == b[0].shape
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("name", ["free16", "mixed16"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_graph_step_matches_eager_step_on_reference_cases(name, precision):
    d = np.load(os.path.join(GOLD, "ref_step_%s.npz" % name), allow_pickle=False)
    n, bc, nsteps, min_vort = json.loads(str(d["spec"]))
    print("case", name, "precision", precision, flush=True)
    _compare(n, tuple(bc), nsteps + 2, min_vort, d["pos_in"], d["vort_in"], precision)


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "enqueue"])
def test_graph_step_matches_eager_step_32(graph):
    pos, vort = ring_particles(32)
    _compare(32, (0, 0, 0), 3, 1e-3, pos, vort, 0, graph=graph)


def test_graph_step_replays_after_eager_work_and_reload():
    """Replays interleaved with other eager work on the device and a reload
    of the particle set: every replay equals a fresh eager run."""
    import torch
    import pmlib3d_amd as pm
    n = 24
    h = 2.0 / n
    pos, vort = ring_particles(n)
    patch = _patch(n, (0, 0, 0))
    gs = pm.GraphVortexStep(patch, order=6, rk_order=2, capacity=n ** 3)
    eager = pm.VortexStep(patch, order=6, rk_order=2, capacity=n ** 3)
    for rep in range(3):
        gs.load(pos, vort)
        eager.load(pos, vort)
        x = torch.randn(1 << 20, device="cuda:0")
        for _ in range(2):
            gs.step(0.5 * h, 1e-3)
            x = x * 1.0001 + 0.5   # unrelated eager work between replays
            eager.step(0.5 * h, 1e-3)
        torch.cuda.synchronize()
        a, b = eager.fetch(), gs.fetch()
        assert a[0].shape == b[0].shape
        for u, v in zip(a, b):
            assert np.array_equal(u, v), rep


def test_graph_step_reports_a_remesh_above_capacity():
    import pmlib3d_amd as pm
    n = 16
    h = 2.0 / n
    pos, vort = ring_particles(n)
    cap = n ** 3 // 2
    gs = pm.GraphVortexStep(_patch(n, (0, 0, 0)), order=6, rk_order=2, capacity=cap)
    gs.load(pos[:, :cap], vort[:, :cap])
    gs.step(0.5 * h, 0.0)   # min_vort 0 keeps every cell the deposition reached: more than cap
    np_, status = gs.sync()
    print("remeshed count", np_, "status", status, flush=True)
    assert status & 1 and np_ == cap
    with pytest.raises(pm.PmgpuError, match="status"):
        gs.fetch()
