"""GPU parity of the Poisson solve (pmgpu_poisson_* through the C ABI) against
the oracle (oracle/pmoracle.c), which restates poisson/pmlib_poisson_setup.f
and poisson/pmlib_poisson_solve.f and is itself pinned against the compiled
reference (tests/test_oracle.py).  FP64 bar: relative L-inf <= 1e-12."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field, gaussian_ring, rel_l2, rel_linf

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _solver(g, p, precision=0):
    import pmlib3d_amd as pm
    patch = pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                     bound_cond=tuple(g.bound_cond), nghost=g.nghost)
    s = pm.PoissonSolver(patch, precision, order=p.poisson_order, kernel=p.poisson_kernel,
                         eps_radius=p.regularisation_radius, interp_order=g.interp_order)
    s.setup()
    return s


CASES = [
    # ncell, nghost, bound_cond, order, (reg_vort, reproj, calc_vecpot)
    ((16, 16, 16), 0, (0, 0, 0), 4, (False, False, True)),
    ((12, 10, 8), 2, (0, 0, 0), 6, (True, True, True)),
    ((10, 8, 12), 0, (0, 0, 0), 8, (True, False, False)),
    ((8, 8, 10), 1, (0, 0, 0), 10, (False, True, True)),
    ((10, 10, 10), 0, (0, 0, 0), 0, (True, False, True)),
    ((20, 12, 18), 0, (0, 0, 0), 2, (False, False, True)),   # lengths 40, 24, 36: radix 3 and 5
    ((8, 8, 12), 2, (1, 1, 0), 4, (True, False, True)),      # mixed periodic / free
    ((12, 8, 8), 0, (1, 1, 1), 6, (True, True, True)),       # fully periodic
    ((16, 20, 10), 0, (0, 1, 0), 6, (False, True, True)),
    ((7, 11, 13), 0, (0, 0, 0), 4, (True, True, True)),      # lengths 14, 22, 26: radix 7, 11, 13
    ((14, 6, 10), 2, (1, 0, 0), 6, (False, False, True)),    # periodic x of length 14
    # order 100: the band-limited kernel Si(pi r / h) / (2 pi^2 r)
    ((12, 10, 8), 1, (0, 0, 0), 100, (False, False, True)),
    ((16, 20, 10), 0, (0, 1, 0), 100, (False, False, True)),
    ((12, 8, 8), 0, (1, 1, 1), 100, (False, False, True)),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_solve_fp64_matches_oracle(case):
    ncell, ng, bc, order, (rv, rp, cv) = case
    print("case", case, flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 1, 2.0)
    w = blob_field(g, seed=sum(ncell) + order)
    Gh = O.poisson_setup(g, p)
    ref = O.poisson_solve(g, p, Gh, w, rv, rp, cv)
    s = _solver(g, p)
    got = s.solve(w, rv, rp, cv)
    for name, a, b in zip(("vel", "vort", "psi"), got, ref):
        if b is None:
            assert a is None
            continue
        e = rel_linf(a, b)
        print(name, e)
        assert e <= TOL64, (name, e)


def test_kernel_octant_matches_oracle():
    g = O.make_geom((12, 10, 14), 1, [-0.5, -0.5, -0.5], 1.0 / 16, (0, 0, 0))
    p = O.make_params(6, 1, 2.0)
    Gh = O.poisson_setup(g, p)
    L = [2 * (n + 2) for n in g.ncell]
    Gh = Gh.reshape(L, order="F")
    s = _solver(g, p)
    oct_ = s.kernel_octant()          # [ky', kx, kz']
    Hx, Hy, Hz = L[0] // 2 + 1, L[1] // 2 + 1, L[2] // 2 + 1
    ref = np.real(Gh[:Hx, :Hy, :Hz]).transpose(1, 0, 2)
    assert rel_linf(oct_, ref) <= 1e-13
    # the spectrum of the even kernel is real: imaginary parts are roundoff
    assert np.abs(np.imag(Gh)).max() <= 1e-13 * np.abs(Gh).max()


def test_config1_ring_64cube_fp64():
    """64^3 cells on [-1,1]^3, free space, order 4, eps = 2h, vortex ring
    R=0.5, sigma=0.1, Gamma=1, stream function and velocity."""
    g = O.make_geom((64, 64, 64), 0, [-1.0, -1.0, -1.0], 2.0 / 64, (0, 0, 0))
    p = O.make_params(4, 1, 2.0)
    w = gaussian_ring(g)
    Gh = O.poisson_setup(g, p)
    ref_vel, _, ref_psi = O.poisson_solve(g, p, Gh, w, False, False, True)
    s = _solver(g, p)
    vel, _, psi = s.solve(w, False, False, True)
    ev, ep = rel_linf(vel, ref_vel), rel_linf(psi, ref_psi)
    print("64^3 fp64 rel Linf vel %.3e psi %.3e" % (ev, ep))
    assert ev <= TOL64 and ep <= TOL64
    s32 = _solver(g, p, precision=1)
    v32, _, p32 = s32.solve(w, False, False, True)
    e32v, e32p = rel_l2(v32, vel), rel_l2(p32, psi)
    print("64^3 fp32 vs fp64 rel L2 vel %.3e psi %.3e" % (e32v, e32p))
    assert e32v <= 1e-5 and e32p <= 1e-5


def test_ring_kat_against_reference_errors():
    """The closed-form ring of validation/poisson (order 6, eps = 2h,
    nghost = 4, reg_vort) on the grids that pmlib_patch_adjust produces.  The
    rms velocity errors must equal what the compiled reference gives on the
    same grids (golden values in tests/golden/kat_ring_ref.json)."""
    import json
    import kat_ring as K
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_ring_ref.json")))
    for (h, ncell, xmin, _stored), r in zip(K.KAT_CASES, ref["cases"]):
        g = O.make_geom(ncell, K.KAT_NGHOST, xmin, h, (0, 0, 0))
        p = O.make_params(K.KAT_ORDER, 1, K.KAT_EPS_RADIUS)
        w, u = K.ring_fields(xmin, h, ncell, K.KAT_NGHOST)
        s = _solver(g, p)
        vel, vo, _ = s.solve(w, True, False, False)
        err = K.rms_velocity_error(vel, u, K.KAT_NGHOST)
        print("h=%g gpu err %.12e reference %.12e" % (h, err, r["rms_error"]))
        assert abs(err - r["rms_error"]) <= 1e-9 * r["rms_error"]


def test_bad_inputs_report_errors():
    import pmlib3d_amd as pm
    from pmlib3d_amd import _lib
    g = O.make_geom((8, 8, 9), 0, [0, 0, 0], 0.1, (1, 1, 1))  # periodic odd length
    with pytest.raises(_lib.PmgpuError, match="odd"):
        _solver(g, O.make_params(4, 1, 2.0))
    g = O.make_geom((17, 8, 8), 0, [0, 0, 0], 0.1, (0, 0, 0))  # L = 34 = 2 * 17
    with pytest.raises(_lib.PmgpuError, match="prime factor"):
        _solver(g, O.make_params(4, 1, 2.0))
    g = O.make_geom((8, 8, 8), 0, [0, 0, 0], 0.1, (0, 0, 0))
    with pytest.raises(_lib.PmgpuError, match="poisson_order"):
        _solver(g, O.make_params(3, 1, 2.0))
    patch = pm.Patch(ncell=(8, 8, 8), dx=(0.1,) * 3, xmin=(0, 0, 0), xmax=(0.8,) * 3)
    mesh = pm.Mesh.allocate(patch)
    assert pm.pmlib_poisson_solve(patch, mesh) == -1     # setup not called


@pytest.mark.parametrize("n,precision", [(32, 0), (64, 1), (128, 0), (256, 1)])
def test_z_pass_tile_widths_agree(n, precision, monkeypatch):
    """Every column-tile width of the register-resident z pass (2, 4, 8
    columns per block) gives the LDS-resident z pass's velocity, on each of
    three repeated solves (white-noise vorticity, device pointers)."""
    import torch
    import pmlib3d_amd as pm
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3)
    dt = torch.float64 if precision == 0 else torch.float32
    w = torch.from_numpy(np.random.default_rng(n).normal(size=3 * n ** 3)).to("cuda:0", dt)

    def solve(zkernel, tc, reps):
        monkeypatch.setenv("PMGPU_Z_KERNEL", str(zkernel))
        monkeypatch.setenv("PMGPU_Z_TC", str(tc))
        s = pm.PoissonSolver(patch, precision, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        s.setup()
        outs = []
        for _ in range(reps):
            out = torch.full_like(w, float("nan"))
            s.solve_device(w.data_ptr(), out.data_ptr())
            torch.cuda.synchronize()
            outs.append(out)
        return outs

    ref = solve(1, 1, 1)[0]
    scale = float(ref.abs().max())
    tol = 1e-13 if precision == 0 else 2e-6
    for tc in (2, 4, 8):
        for r, out in enumerate(solve(0, tc, 3)):
            e = float((out - ref).abs().max()) / scale
            print("n %d precision %d tc %d rep %d rel max diff %.3e" % (n, precision, tc, r, e), flush=True)
            assert e <= tol, (tc, r, e)


@pytest.mark.parametrize("name", ["free_ng1_o100_psi", "periodic_o100_psi"])
def test_order100_matches_reference_golden(name):
    """Order 100 (poisson/pmlib_poisson_setup.f:1419-1505, the sine integral
    of :1739-1841) against the compiled reference's velocity and stream
    function."""
    import json
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_solve_%s.npz" % name), allow_pickle=False)
    ncell, ng, bc, order, (rv, rp, cv) = json.loads(str(d["spec"]))[:5]
    print("golden", name, flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 1, 2.0)
    vel, _, psi = _solver(g, p).solve(d["vort_in"], rv, rp, cv)
    assert rel_linf(vel, d["vel"]) <= TOL64
    assert rel_linf(psi, d["psi"]) <= TOL64
    v32, _, _ = _solver(g, p, 1).solve(d["vort_in"], rv, rp, cv)
    assert rel_l2(v32, vel) <= 1e-5


def test_order100_refuses_vorticity_outputs():
    """The reference's solve defines no vorticity filter for order 100."""
    import pmlib3d_amd as pm
    g = O.make_geom((8, 8, 8), 0, [-0.5] * 3, 1.0 / 8)
    s = _solver(g, O.make_params(100, 1, 2.0))
    with pytest.raises(pm.PmgpuError, match="order 100"):
        s.solve(np.zeros((3, 8, 8, 8), order="F"), True, False, False)
