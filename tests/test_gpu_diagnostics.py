"""GPU parity of the FP64 global diagnostics (pmgpu_diagnostics_particles,
pmgpu_diagnostics_mesh through the C ABI) against the oracle's
or_diag_particles / or_diag_mesh, which the closed-form tests of
tests/test_oracle.py pin (discrete ring, uniform fields).  Every term is
formed in the reference's operation order; the sums run in a fixed tree
order instead of the serial loop, so each total is checked to 1e-12 of the
sum of the magnitudes of its terms.  FP32 meshes are accumulated in FP64
and checked against the FP64 totals of the same (rounded) values.
References: diagnostics_ring.f:72-226, diagnostics_energy_mesh.f:72-134
(vortex_codes/vortex_structures/source/diagnostics/)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field

pytestmark = pytest.mark.gpu


def _random_particles(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.asfortranarray(rng.uniform(-1, 1, (3, n)))
    vort = np.asfortranarray(rng.normal(size=(3, n)))
    vel = np.asfortranarray(rng.normal(size=(3, n)))
    return pos, vort, vel


def _term_scale(pos, vort, vel, h3):
    """Sum of |per-particle term| for each total (tolerance scale only)."""
    e = (vort ** 2).sum(0)
    m1 = np.cross(pos.T, vort.T).T
    m2 = np.cross(pos.T, m1.T).T
    r = np.abs
    return np.array([r(m1[0]).sum() / 2, r(m1[1]).sum() / 2, r(m1[2]).sum() / 2,
                     r(m2[0]).sum() / 3, r(m2[1]).sum() / 3, r(m2[2]).sum() / 3,
                     r((vel * m1).sum(0)).sum(), e.sum(), r((vel * vort).sum(0)).sum()]) * h3


@pytest.mark.parametrize("n", [1, 255, 4097, 300001])
def test_particle_diagnostics_match_oracle(n):
    import pmlib3d_amd as pm
    print("particles", n, flush=True)
    pos, vort, vel = _random_particles(n, 7 + n)
    h = (0.02, 0.025, 0.03)
    ref = O.diag_particles(h, pos, vort, vel)
    got = np.array(list(pm.diagnostics_ring(h, pos, vort, vel).values()))
    sc = _term_scale(pos, vort, vel, h[0] * h[1] * h[2])
    for q in range(9):
        assert abs(got[q] - ref[q]) <= 1e-12 * sc[q], (O.DIAG_PARTICLE_NAMES[q], got[q], ref[q])
    # centroid, radius, circulation, core: quotients of such sums
    for q in (9, 10, 11, 12, 13, 14):
        assert abs(got[q] - ref[q]) <= 1e-10 * max(1.0, abs(ref[q])), (O.DIAG_PARTICLE_NAMES[q], got[q], ref[q])
    assert got[15] == ref[15]


def test_particle_diagnostics_of_a_discrete_ring():
    """The closed-form ring of the oracle tests, on the GPU."""
    import pmlib3d_amd as pm
    n, R, x0, gamma, h = 720, 0.37, 0.11, 1.3, 0.02
    th = 2 * np.pi * (np.arange(n) + 0.25) / n
    pos = np.stack([np.full(n, x0), R * np.cos(th), R * np.sin(th)])
    s = gamma * (2 * np.pi * R / n) / h ** 3
    vort = np.stack([np.zeros(n), -s * np.sin(th), s * np.cos(th)])
    vel = np.tile(np.array([0.7, 0.2, -0.3])[:, None], (1, n))
    d = pm.diagnostics_ring((h, h, h), pos, vort, vel)
    assert abs(d["impulse_x"] - np.pi * R * R * gamma) <= 1e-13 * np.pi * R * R * gamma
    assert abs(d["circulation"] - gamma) <= 1e-13 * gamma
    assert abs(d["vort_radius"] - R) <= 1e-14
    assert abs(d["kin_energy"] - 2 * 0.7 * np.pi * R * R * gamma) <= 1e-13


def test_particle_diagnostics_are_run_to_run_identical():
    import torch
    import pmlib3d_amd as pm
    pos, vort, vel = _random_particles(123457, 3)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(a.T)).to(dev) for a in (pos, vort, vel)]
    d_out = torch.zeros(16, dtype=torch.float64, device=dev)
    a = pm.diagnostics_particles_device((0.01,) * 3, *[x.data_ptr() for x in t], pos.shape[1])
    pm.diagnostics_particles_device((0.01,) * 3, *[x.data_ptr() for x in t], pos.shape[1], d_out=d_out.data_ptr(),
                                    host=False)
    torch.cuda.synchronize()
    assert np.array_equal(a, d_out.cpu().numpy())


MESH_CASES = [((24, 20, 18), 2, (0, 0, 0)), ((16, 16, 16), 0, (1, 1, 1)), ((33, 9, 70), 1, (1, 0, 0))]


@pytest.mark.parametrize("case", MESH_CASES, ids=[str(i) for i in range(len(MESH_CASES))])
def test_mesh_diagnostics_match_oracle(case):
    import torch
    import pmlib3d_amd as pm
    ncell, ng, bc = case
    print("mesh", case, flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.4, -0.6], 1.0 / 32, bc)
    vel = blob_field(g, seed=5, width=0.2)
    vort = blob_field(g, seed=6, width=0.2) + 0.01   # no cell with zero vorticity
    ref = O.diag_mesh(g, vel, vort)
    patch = pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                     bound_cond=tuple(g.bound_cond), nghost=g.nghost)
    got = np.array(list(pm.diagnostics_energy_mesh(patch, vel, vort).values()))
    h3 = np.prod(g.dx)
    e = (vort ** 2).sum(0)
    n = np.prod(ncell)
    sc = [0.5 * (vel ** 2).sum() * h3, e.sum() * h3, (vel ** 2).sum() * h3, (vel ** 2).sum() * h3]
    for q in range(4):
        assert abs(got[q] - ref[q]) <= 1e-12 * max(sc[q], abs(ref[q])), (O.DIAG_MESH_NAMES[q], got[q], ref[q], n)
    assert got[4] == ref[4]
    # FP32 meshes accumulate in FP64: the totals of the rounded values
    dev = torch.device("cuda:0")
    v32, w32 = (torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev).float() for a in (vel, vort))
    got32 = pm.diagnostics_mesh_device(patch, pm.FP32, v32.data_ptr(), w32.data_ptr())
    ref32 = O.diag_mesh(g, v32.double().cpu().numpy().reshape(vel.shape, order="F"),
                        w32.double().cpu().numpy().reshape(vort.shape, order="F"))
    for q in range(4):
        assert abs(got32[q] - ref32[q]) <= 1e-12 * max(sc[q], abs(ref32[q]))
    assert got32[4] == ref32[4]
