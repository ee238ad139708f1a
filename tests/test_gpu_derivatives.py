"""Vorticity right-hand side (part_deriv_vorticity.f:57-208) on the GPU
against the oracle, which the closed-form tests in test_oracle.py and the
client routine's own outputs (tests/golden/ref_rhs_*.npz) pin: FP64
bit-identical (the kernel rounds every operation explicitly, in the
reference's order), FP32 within 1e-5 relative L2 of FP64; free-side ghost
cells of dvort keep their values, periodic ones are the interior's images."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOMS = [
    # ncell, nghost, bound_cond, interp_order
    ((20, 18, 16), 2, (0, 0, 0), 3),
    ((16, 12, 20), 0, (1, 1, 1), 3),
    ((16, 14, 12), 2, (1, 0, 1), 4),
    ((24, 20, 18), 1, (0, 0, 0), 3),   # one ghost layer: 2nd order only
    ((90, 11, 19), 2, (0, 0, 0), 3),   # rows of 270 words: two tiles, a cell across the tile edge
    ((172, 7, 5), 0, (1, 1, 1), 3),    # 516 words: three tiles; partial row groups
]


def _patch(g):
    import pmlib3d_amd as pm
    return pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                    bound_cond=tuple(g.bound_cond), nghost=g.nghost)


def _fields(g, seed):
    E = O.mesh_extent(g)
    rng = np.random.default_rng(seed)
    w = np.asfortranarray(rng.normal(size=(3,) + tuple(E)))
    u = np.asfortranarray(rng.normal(size=(3,) + tuple(E)))
    d0 = np.asfortranarray(rng.normal(size=(3,) + tuple(E)))
    return w, u, d0


# 4th-order differences need two ghost layers on free sides
RUNS = [(gi, fd) for gi in range(len(GEOMS)) for fd in (2, 4)
        if not (fd == 4 and GEOMS[gi][2] != (1, 1, 1) and GEOMS[gi][1] < 2)]


@pytest.mark.parametrize("gi,fd_order", RUNS)
def test_vorticity_rhs_fp64_bit_identical(gi, fd_order):
    import pmlib3d_amd as pm
    ncell, ng, bc, io = GEOMS[gi]
    print("geom", GEOMS[gi], "fd", fd_order, flush=True)
    g = O.make_geom(ncell, ng, [-0.4, -0.3, -0.5], 0.05, bc, io)
    g.dx[1] = 0.06
    g.xmax[1] = g.xmin[1] + g.ncell[1] * g.dx[1]
    w, u, d0 = _fields(g, 3 + gi)
    want = O.vort_rhs(g, w, u, 0.011, fd_order, dvort=d0)
    got = pm.vorticity_rhs(_patch(g), w, u, 0.011, fd_order=fd_order, dvort=d0, interp_order=io)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("gi", [0, 2])
def test_vorticity_rhs_device_fp32_against_fp64(gi):
    import torch
    import pmlib3d_amd as pm
    ncell, ng, bc, io = GEOMS[gi]
    print("geom", GEOMS[gi], flush=True)
    g = O.make_geom(ncell, ng, [-0.4, -0.3, -0.5], 0.05, bc, io)
    w, u, d0 = _fields(g, 21 + gi)
    want = O.vort_rhs(g, w, u, 0.02, 4, dvort=d0)
    dev = torch.device("cuda:0")
    for prec, dt, tol in ((pm.FP64, torch.float64, 0.0), (pm.FP32, torch.float32, 1e-5)):
        tw = torch.from_numpy(w.ravel(order="F")).to(dev).to(dt)
        tu = torch.from_numpy(u.ravel(order="F")).to(dev).to(dt)
        td = torch.from_numpy(d0.ravel(order="F")).to(dev).to(dt)
        pm.vorticity_rhs_device(_patch(g), prec, tw.data_ptr(), tu.data_ptr(), td.data_ptr(), 0.02, fd_order=4,
                                interp_order=io)
        torch.cuda.synchronize()
        got = td.double().cpu().numpy().reshape(want.shape, order="F")
        if tol == 0.0:
            assert np.array_equal(got, want)
        else:
            assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)


def test_vorticity_rhs_errors_and_ierr():
    import pmlib3d_amd as pm
    g = O.make_geom((12, 10, 8), 1, [-0.4, -0.3, -0.5], 0.05, (0, 0, 0), 3)
    w, u, d0 = _fields(g, 5)
    with pytest.raises(pm.PmgpuError):
        pm.vorticity_rhs(_patch(g), w, u, 0.01, fd_order=4, interp_order=3)   # one ghost layer
    with pytest.raises(pm.PmgpuError):
        pm.vorticity_rhs(_patch(g), w, u, 0.01, fd_order=3, interp_order=3)
    m = pm.Mesh()
    m.vort, m.vel = w, u
    pm.params.pmlib_fd_order = 2
    try:
        assert pm.part_deriv_vorticity(_patch(g), m, 0.01, interp_order=3) == 0
        assert np.array_equal(m.dvort, O.vort_rhs(g, w, u, 0.01, 2))
        pm.params.pmlib_fd_order = 4
        assert pm.part_deriv_vorticity(_patch(g), m, 0.01, interp_order=3) == -1
    finally:
        pm.params.pmlib_fd_order = -99


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_rhs_*.npz"))),
                         ids=lambda p: os.path.basename(p)[8:-4])
def test_vorticity_rhs_matches_client_routine_bitwise(path):
    """Against the dvort the vortex client's own part_deriv_vorticity gives
    (compiled with the reference library)."""
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, io, fd = json.loads(str(d["spec"]))
    print("golden", os.path.basename(path), flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc, io)
    got = pm.vorticity_rhs(_patch(g), d["vort"], d["vel"], float(d["visc"]), fd_order=fd, dvort=d["dvort_in"],
                           interp_order=io)
    assert np.array_equal(got, d["dvort"])
