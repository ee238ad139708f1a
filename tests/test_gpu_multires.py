"""GPU parity of the multiresolution coupling through the C ABI:
pmgpu_interp_mesh_mesh (interpolation/pmlib_interp_mesh_mesh.f:5-2078)
agrees with the compiled reference's golden vectors
(tests/golden/ref_m2m_*.npz) and with the oracle at larger sizes, in both
directions: to 1e-14 relative L-inf in its default separable form (three 1D
passes: the reference's term-by-term sum regrouped), bit for bit in the
PMGPU_M2M_EXACT=1 gathers.  The refinement-patch solve (pmgpu_poisson_set_level +
pmgpu_poisson_setup, poisson/pmlib_poisson_setup.f:446-528 and the other
orders' refinement blocks) matches the reference's velocities
(tests/golden/ref_refine_*.npz) to 1e-12 relative (FP64) and 1e-4 (FP32)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
MODES = ["separable", "exact"]


@pytest.fixture
def m2m_mode(request):
    """Selects the mesh-to-mesh form for one test (the library reads
    PMGPU_M2M_EXACT on every call)."""
    old = os.environ.get("PMGPU_M2M_EXACT")
    os.environ["PMGPU_M2M_EXACT"] = "1" if request.param == "exact" else "0"
    yield request.param
    if old is None:
        del os.environ["PMGPU_M2M_EXACT"]
    else:
        os.environ["PMGPU_M2M_EXACT"] = old


def _agree(got, want, mode):
    if mode == "exact":
        np.testing.assert_array_equal(got, want)
    else:
        assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max(), np.abs(got - want).max()


def _patch(ncell, ng, bc, dx=1.0 / 8, xmin=(-0.5, -0.6, -0.7), level=1, parent=(0, 0)):
    import pmlib3d_amd as pm
    xmax = tuple(xmin[d] + ncell[d] * dx for d in range(3))
    return pm.Patch(ncell=tuple(ncell), nghost=ng, bound_cond=tuple(bc), dx=(dx,) * 3, xmin=tuple(xmin),
                    xmax=xmax, level=level, parent=parent)


@pytest.mark.parametrize("m2m_mode", MODES, indirect=True)
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_m2m_*.npz"))),
                         ids=lambda p: os.path.basename(p)[8:-4])
def test_m2m_matches_reference_golden(path, m2m_mode):
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, io, level, clear = json.loads(str(d["spec"]))
    print("m2m", os.path.basename(path), m2m_mode, flush=True)
    patch = _patch(ncell, ng, bc, level=level, parent=(1, 0))
    parent = pm.pmlib_topology_parent(patch, interp_order=io)
    assert parent.ncell == tuple(d["parent_ncell"])
    down = np.asfortranarray(np.full(d["coarse"].shape, 7.0))
    pm.pmlib_interp_mesh_mesh(patch, parent, d["fine"], down, clear=True, interp_order=io)
    _agree(down, d["down"], m2m_mode)
    up = np.array(d["base"], order="F", copy=True)
    pm.pmlib_interp_mesh_mesh(parent, patch, d["coarse"], up, clear=clear, interp_order=io)
    _agree(up, d["up"], m2m_mode)


@pytest.mark.parametrize("m2m_mode", MODES, indirect=True)
@pytest.mark.parametrize("io,bc,level", [(3, (0, 0, 0), 2), (4, (0, 0, 0), 2), (4, (1, 1, 0), 2),
                                         (3, (1, 1, 1), 2), (4, (0, 1, 0), 3)])
def test_m2m_device_matches_oracle(io, bc, level, m2m_mode):
    """48^3-cell patches through the device entry point, both directions,
    clear and accumulate, against the oracle (FP64)."""
    import torch
    import pmlib3d_amd as pm
    print("m2m device", io, bc, level, m2m_mode, flush=True)
    n, ng = 48, 3
    patch = _patch((n, n, n), ng, bc, dx=1.0 / 24, level=level, parent=(1, 0))
    parent = pm.pmlib_topology_parent(patch, interp_order=io)
    g = O.make_geom((n, n, n), ng, patch.xmin, patch.dx[0], bc, io)
    gp = O.topology_parent(g, level, 1)
    rng = np.random.default_rng(17 + io + level)
    fine = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(g)))
    coarse = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(gp)))
    base = np.asfortranarray(rng.normal(size=fine.shape))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()
    d_fine, d_coarse, d_base = dev(fine), dev(coarse), dev(base)
    d_down = torch.full_like(d_coarse, 3.0)
    pm.interp_mesh_mesh_device(patch, parent, pm.FP64, d_fine.data_ptr(), d_down.data_ptr(), clear=True,
                               interp_order=io)
    d_up = d_base.clone()
    pm.interp_mesh_mesh_device(parent, patch, pm.FP64, d_coarse.data_ptr(), d_up.data_ptr(), clear=False,
                               interp_order=io)
    torch.cuda.synchronize()
    host = lambda t, shp: t.cpu().numpy().reshape(shp, order="F")
    _agree(host(d_down, coarse.shape), O.interp_mesh_mesh(g, gp, fine, clear=True), m2m_mode)
    _agree(host(d_up, base.shape), O.interp_mesh_mesh(gp, g, coarse, base, clear=False), m2m_mode)


@pytest.mark.parametrize("m2m_mode", MODES, indirect=True)
def test_m2m_fp32_matches_oracle(m2m_mode):
    import torch
    import pmlib3d_amd as pm
    print("m2m fp32", flush=True)
    io, bc = 4, (0, 1, 0)
    patch = _patch((32, 32, 32), 3, bc, dx=1.0 / 16, level=2, parent=(1, 0))
    parent = pm.pmlib_topology_parent(patch, interp_order=io)
    g = O.make_geom((32, 32, 32), 3, patch.xmin, patch.dx[0], bc, io)
    gp = O.topology_parent(g, 2, 1)
    rng = np.random.default_rng(5)
    fine = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(g))).astype(np.float32)
    coarse = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(gp))).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()
    d_fine, d_coarse = dev(fine), dev(coarse)
    d_down, d_up = torch.zeros_like(d_coarse), torch.zeros_like(d_fine)
    pm.interp_mesh_mesh_device(patch, parent, pm.FP32, d_fine.data_ptr(), d_down.data_ptr(), True, interp_order=io)
    pm.interp_mesh_mesh_device(parent, patch, pm.FP32, d_coarse.data_ptr(), d_up.data_ptr(), True, interp_order=io)
    torch.cuda.synchronize()
    down = d_down.cpu().numpy().reshape(coarse.shape, order="F")
    up = d_up.cpu().numpy().reshape(fine.shape, order="F")
    want_down = O.interp_mesh_mesh(g, gp, fine.astype(np.float64), clear=True)
    want_up = O.interp_mesh_mesh(gp, g, coarse.astype(np.float64), clear=True)
    # FP64 sums of FP32 inputs, rounded to FP32 once per pass (separable: three
    # passes, so three roundings)
    tol = 2e-7 if m2m_mode == "exact" else 5e-7
    assert np.abs(down - want_down).max() <= tol * np.abs(want_down).max() + 1e-7
    assert np.abs(up - want_up).max() <= tol * np.abs(want_up).max() + 1e-7


@pytest.mark.parametrize("m2m_mode", MODES, indirect=True)
def test_m2m_rejects_what_it_does_not_support(m2m_mode):
    import pmlib3d_amd as pm
    patch = _patch((32, 16, 16), 2, (0, 0, 0), level=5, parent=(1, 0))
    parent = pm.pmlib_topology_parent(patch, interp_order=3)   # ratio 16
    with pytest.raises(pm.PmgpuError, match="ratios above 8"):
        pm.pmlib_interp_mesh_mesh(patch, parent, np.zeros((3, 36, 20, 20), order="F"),
                                  np.zeros((3, 6, 5, 5), order="F"), True, interp_order=3)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_refine_*.npz"))),
                         ids=lambda p: os.path.basename(p)[11:-4])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_refinement_solve_matches_reference_golden(path, prec):
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, order, kernel, level, parent = json.loads(str(d["spec"]))
    print("refine", os.path.basename(path), prec, flush=True)
    patch = _patch(ncell, ng, bc, level=level, parent=(parent, 0))
    s = pm.PoissonSolver(patch, pm.FP64 if prec == "fp64" else pm.FP32, order=order, kernel=kernel,
                         eps_radius=2.0, interp_order=3)
    s.setup()
    vel, _, _ = s.solve(d["vort_in"])
    err = np.abs(vel - d["vel"]).max() / np.abs(d["vel"]).max()
    assert err <= (1e-12 if prec == "fp64" else 1e-4), err


def test_refinement_solve_at_size_matches_oracle():
    """A 60^3 free-space refinement patch (order 6, kernel 1, a = 2) against
    the oracle's solve."""
    import pmlib3d_amd as pm
    print("refine 60^3", flush=True)
    patch = _patch((60, 60, 60), 2, (0, 0, 0), dx=1.0 / 30, level=2, parent=(1, 0))
    g = O.make_geom((60, 60, 60), 2, patch.xmin, patch.dx[0])
    p = O.make_params(6, 1, 2.0, refine=2.0)
    w = blob_field(g, seed=9)
    want, _, _ = O.poisson_solve(g, p, O.poisson_setup(g, p), w)
    s = pm.PoissonSolver(patch, pm.FP64, order=6, kernel=1, eps_radius=2.0, interp_order=3)
    s.setup()
    vel, _, _ = s.solve(w)
    assert np.abs(vel - want).max() <= 1e-12 * np.abs(want).max()


def test_refinement_order_0_fails_in_setup():
    import pmlib3d_amd as pm
    patch = _patch((16, 16, 16), 0, (0, 0, 0), level=2, parent=(1, 0))
    s = pm.PoissonSolver(patch, pm.FP64, order=0, kernel=1, eps_radius=2.0, interp_order=3)
    with pytest.raises(pm.PmgpuError, match="multi-patch"):
        s.setup()
