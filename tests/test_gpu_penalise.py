"""GPU parity of the Brinkman penalisation step (pmgpu_penalise* through the
C ABI) against the oracle's or_penalise, which the polynomial-exactness tests
of tests/test_oracle.py pin.  FP64: the vorticity bit-identical, the sums
within 1e-12 relative (they are reduced in a different order).  FP32 against
FP64: relative L2 <= 1e-5."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field, rel_l2, rel_linf

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
    # several 16-plane z chunks of the z-marching kernel, the last one partial
    ((20, 11, 37), 2, (0, 0, 0), 4, "field"),
    ((13, 9, 33), 1, (1, 0, 1), 2, None),
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


# ------------------------------------------------------ penalisation_solve
SOLVE_CASES = [
    # ncell, nghost, bound_cond, fd_order, reproject, solid: "rest" / "field" / "const", max_iter, tolerance
    ((16, 16, 16), 2, (0, 0, 0), 4, False, "rest", 3, 0.0),
    ((16, 12, 20), 2, (0, 0, 0), 2, True, "const", 6, 0.05),
    ((12, 12, 12), 0, (1, 1, 1), 4, True, "field", 2, 0.0),
    ((12, 16, 10), 2, (1, 1, 0), 4, False, "rest", 4, 0.0),
]


@pytest.mark.parametrize("case", SOLVE_CASES, ids=[str(i) for i in range(len(SOLVE_CASES))])
def test_penalisation_solve_matches_oracle(case):
    """The whole iteration (penalise, solve, force residual test, map back,
    final solve, free stream with the reference's component quirk) against
    the oracle composition, FP64 within 1e-12; the same iteration count."""
    import torch
    import pmlib3d_amd as pm
    ncell, ng, bc, fd, reproj, solid, nmax, tol = case
    print("penalisation_solve case", case, flush=True)
    g = O.make_geom(ncell, ng, [-0.6, -0.5, -0.7], 1.2 / 16, bc)
    E = O.mesh_extent(g)
    p = O.make_params(4, 1, 2.0)
    patch = _patch(g)
    vel = blob_field(g, seed=31, width=0.15)
    vel[0] += 1.0
    vort = blob_field(g, seed=32, width=0.15)
    chi = pm.brinkman_factor(pm.sphere_mask(patch, 0.3, (0.05, -0.02, 0.0)), 1e4, 0.02)
    rng = np.random.default_rng(7)
    sv, uc = None, None
    if solid == "field":
        sv = np.asfortranarray(0.1 * rng.normal(size=(3,) + E))
    elif solid == "const":
        uc = (0.3, -0.1, 0.2)
        sv = np.zeros((3,) + E, order="F")
        sv[0], sv[1], sv[2] = uc
    kw = dict(relaxation=0.9, tolerance=tol, max_iter=nmax, fd_order=fd, dtime=0.02, vel_inf=(1.0, 0.5, 0.25),
              reproject=reproj)
    rv, rw, rf, rn, rres = O.penalisation_solve(g, p, O.poisson_setup(g, p), vel, vort, chi, solid_vel=sv, **kw)
    dev = torch.device("cuda:0")
    for prec, dt in ((pm.FP64, torch.float64), (pm.FP32, torch.float32)):
        s = pm.PoissonSolver(patch, prec, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        s.setup()
        t = {k: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev).to(dt)
             for k, a in (("vel", vel), ("vort", vort), ("chi", chi))}
        d_sv = None
        if solid == "field":
            tsv = torch.from_numpy(np.ascontiguousarray(sv.ravel(order="F"))).to(dev).to(dt)
            d_sv = tsv.data_ptr()
        f, n, res = pm.penalisation_solve_device(s, t["chi"].data_ptr(), t["vel"].data_ptr(), t["vort"].data_ptr(),
                                                 d_solid_vel=d_sv, us_const=uc if solid == "const" else None, **kw)
        gv = t["vel"].double().cpu().numpy().reshape(vel.shape, order="F")
        gw = t["vort"].double().cpu().numpy().reshape(vort.shape, order="F")
        if prec == pm.FP64:
            print("iterations gpu %d oracle %d; force %s / %s" % (n, rn, f, rf), flush=True)
            assert n == rn
            assert np.abs(f - rf).max() <= 1e-12 * np.abs(rf).max()
            ev, ew = rel_linf(gv, rv), rel_linf(gw, rw)
            print("fp64 vel %.3e vort %.3e" % (ev, ew), flush=True)
            assert ev <= 1e-12 and ew <= 1e-12
            v64, w64 = gv, gw
        else:
            ev, ew = rel_l2(gv, v64), rel_l2(gw, w64)
            print("fp32 vs fp64 vel %.3e vort %.3e, iterations %d" % (ev, ew, n), flush=True)
            assert ev <= 1e-5 and ew <= 1e-5
