"""GPU parity of the Brinkman penalisation step (pmgpu_penalise* through the
C ABI) against the oracle's or_penalise, which the polynomial-exactness tests
of tests/test_oracle.py pin.  FP64: the vorticity bit-identical, the sums
within 1e-12 relative (they are reduced in a different order).  FP32 against
FP64: relative L2 <= 1e-5."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field

pytestmark = pytest.mark.gpu


def _patch(g):
    import pmlib3d_amd as pm
    return pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                    bound_cond=tuple(g.bound_cond), nghost=g.nghost)


CASES = [
    # ncell, nghost, bound_cond, fd_order, solid velocity: None / "field" / "const"
    ((16, 12, 10), 2, (0, 0, 0), 4, None),
    ((70, 9, 8), 3, (0, 0, 0), 4, "field"),
    ((12, 12, 12), 0, (1, 1, 1), 4, "const"),
    ((10, 14, 9), 1, (0, 0, 0), 2, "field"),
    ((9, 10, 12), 2, (1, 0, 1), 2, "const"),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_penalise_fp64_matches_oracle(case):
    import pmlib3d_amd as pm
    ncell, ng, bc, fd, sv = case
    print("case", case, flush=True)
    g = O.make_geom(ncell, ng, [-0.4, -0.3, -0.5], 0.07, bc)
    E = O.mesh_extent(g)
    rng = np.random.default_rng(sum(ncell) + fd)
    vel = blob_field(g, seed=3)
    vort = np.asfortranarray(rng.normal(size=(3,) + E))
    patch = _patch(g)
    chi = pm.brinkman_factor(pm.sphere_mask(patch, 0.2, (0.05, 0.0, -0.02)), 1e4, 0.01)
    uc = None
    field = None
    if sv == "field":
        field = np.asfortranarray(rng.normal(size=(3,) + E))
    elif sv == "const":
        uc = (-1.0, 0.25, 0.5)
        field = np.zeros((3,) + E, order="F")
        field[0], field[1], field[2] = uc
    ref_w, ref_d = O.penalise(g, vel, chi, vort, solid_vel=field, relax=0.8, fd_order=fd)
    got_w, got_d = pm.penalise(patch, vel, chi, vort, solid_vel=field if sv == "field" else None,
                               us_const=uc, relax=0.8, fd_order=fd)
    assert np.array_equal(got_w, ref_w)
    assert np.abs(got_d - ref_d).max() <= 1e-12 * np.abs(ref_d).max()


def test_penalise_device_fp32_against_fp64():
    import torch
    import pmlib3d_amd as pm
    n = 64
    g = O.make_geom((n, n, n), 2, [-1, -1, -1], 2.0 / n, (0, 0, 0))
    patch = _patch(g)
    E = O.mesh_extent(g)
    vel = blob_field(g, seed=8, width=0.1)
    vort = blob_field(g, seed=9, width=0.1)
    chi = pm.brinkman_factor(pm.sphere_mask(patch, 0.25), 1e4, 0.5 * 2.0 / n)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev)
         for k, a in (("vel", vel), ("vort", vort), ("chi", chi))}
    w64 = t["vort"].clone()
    d64 = pm.penalise_device(patch, pm.FP64, t["vel"].data_ptr(), t["chi"].data_ptr(), w64.data_ptr(),
                             us_const=(-1.0, 0.0, 0.0), relax=1.0)
    t32 = {k: v.float() for k, v in t.items()}
    w32 = t32["vort"].clone()
    d32 = pm.penalise_device(patch, pm.FP32, t32["vel"].data_ptr(), t32["chi"].data_ptr(), w32.data_ptr(),
                             us_const=(-1.0, 0.0, 0.0), relax=1.0)
    torch.cuda.synchronize()
    a, b = w32.double(), w64
    err = float(torch.sqrt(torch.sum((a - b) ** 2) / torch.sum(b * b)))
    print("fp32 vs fp64 rel L2 %.3e" % err, flush=True)
    assert err <= 1e-5
    assert np.abs(d32 - d64).max() <= 1e-4 * np.abs(d64).max()
    # the host FP64 entry point gives the same bits as the device one
    ref_w, _ = pm.penalise(patch, vel, chi, vort, us_const=(-1.0, 0.0, 0.0), relax=1.0)
    assert np.array_equal(w64.cpu().numpy().reshape((3,) + E, order="F"), ref_w)


def test_penalise_rejects_bad_input():
    import pmlib3d_amd as pm
    g = O.make_geom((8, 8, 8), 1, [0, 0, 0], 0.1, (0, 0, 0))
    E = O.mesh_extent(g)
    z = np.zeros((3,) + E, order="F")
    with pytest.raises(pm.PmgpuError):
        pm.penalise(_patch(g), z, np.zeros(E), z, fd_order=4)   # needs 2 ghost layers
    with pytest.raises(pm.PmgpuError):
        pm.penalise(_patch(g), z, np.zeros(E), z, fd_order=3)
