"""GPU parity of the kernel-2 Poisson solve (grad G; pmgpu_poisson_* through
the C ABI): against the compiled reference's outputs (tests/golden/
ref_solve_*k2*.npz, written by tests/golden/make_golden.py) and against the
oracle (oracle/pmoracle.c, whose kernel 2 tests/test_oracle.py pins as the
gradient of kernel 1 and by the same golden vectors).  FP64 within 1e-12
relative L-inf, FP32 within 1e-5 relative L2 of FP64.
References: poisson/pmlib_poisson_setup.f:580-635 (and the kernel-2 block of
every order), poisson/pmlib_poisson_solve.f:873-913."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field, rel_l2, rel_linf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _solver(g, p, precision=0):
    import pmlib3d_amd as pm
    patch = pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                     bound_cond=tuple(g.bound_cond), nghost=g.nghost)
    s = pm.PoissonSolver(patch, precision, order=p.poisson_order, kernel=p.poisson_kernel,
                         eps_radius=p.regularisation_radius, interp_order=g.interp_order)
    s.setup()
    return s


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_solve_*k2*.npz"))),
                         ids=lambda p: os.path.basename(p)[10:-4])
def test_kernel2_matches_reference_golden(path):
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, order, (rv, rp, cv), kernel = json.loads(str(d["spec"]))
    assert kernel == 2
    print("golden", os.path.basename(path), flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 2, 2.0)
    vel, vo, _ = _solver(g, p).solve(d["vort_in"], rv, rp, False)
    e = rel_linf(vel, d["vel"])
    print("vel rel Linf %.3e" % e, flush=True)
    assert e <= 1e-12
    if rv or rp:
        e = rel_linf(vo, d["vort"])
        print("vort rel Linf %.3e" % e, flush=True)
        assert e <= 1e-12
    v32, vo32, _ = _solver(g, p, 1).solve(d["vort_in"], rv, rp, False)
    assert rel_l2(v32, vel) <= 1e-5
    if rv or rp:
        assert rel_l2(vo32, vo) <= 1e-5


CASES = [
    # ncell, nghost, bound_cond, order, (reg_vort, reproj)
    ((16, 16, 16), 0, (0, 0, 0), 4, (False, False)),
    ((12, 10, 8), 2, (0, 0, 0), 6, (True, True)),
    ((20, 12, 18), 0, (0, 0, 0), 2, (False, True)),     # lengths 40, 24, 36: radix 3 and 5
    ((8, 8, 12), 2, (1, 1, 0), 4, (True, True)),        # the mixed-BC form
    ((16, 20, 10), 0, (0, 1, 0), 8, (False, True)),
    ((7, 11, 13), 0, (0, 0, 0), 4, (True, True)),       # lengths 14, 22, 26: runtime-planned kernels
    ((12, 8, 8), 0, (1, 1, 1), 6, (True, True)),        # all periodic: the spectral form
    ((32, 32, 32), 0, (0, 0, 0), 10, (False, False)),   # length 64 (register-resident x and y passes)
]


@pytest.mark.parametrize("zkernel", [0, 1], ids=["zdefault", "zlds"])
@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_kernel2_matches_oracle(case, zkernel, monkeypatch):
    ncell, ng, bc, order, (rv, rp) = case
    monkeypatch.setenv("PMGPU_Z_KERNEL", str(zkernel))
    print("case", case, "z kernel", zkernel, flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 2, 2.0)
    w = blob_field(g, seed=sum(ncell) + order)
    ref_v, ref_w, _ = O.poisson_solve(g, p, O.poisson_setup(g, p), w, rv, rp, False)
    vel, vo, _ = _solver(g, p).solve(w, rv, rp, False)
    e = rel_linf(vel, ref_v)
    print("vel rel Linf %.3e" % e, flush=True)
    assert e <= 1e-12
    if rv or rp:
        e = rel_linf(vo, ref_w)
        print("vort rel Linf %.3e" % e, flush=True)
        assert e <= 1e-12


def test_kernel2_refuses_the_scalar_solve():
    import pmlib3d_amd as pm
    from pmlib3d_amd import _lib
    g = O.make_geom((16, 16, 16), 0, [-0.5] * 3, 1.0 / 16, (0, 0, 0))
    s = _solver(g, O.make_params(4, 2, 2.0))
    with pytest.raises(_lib.PmgpuError, match="poisson_kernel = 1"):
        s.solve_scalar(np.zeros((16, 16, 16), order="F"))
    assert isinstance(s, pm.PoissonSolver)
