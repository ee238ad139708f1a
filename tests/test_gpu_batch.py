"""Single-component solves (psi = G * w) and the graph-captured multipatch
batch.  The scalar solve is checked against the oracle's psi for a vorticity
with that one component (the solve is linear and component-wise for psi,
poisson/pmlib_poisson_solve.f:854-858): FP64 within 1e-12 relative L-inf,
FP32 within 1e-5 relative L2.  The batch must give exactly the results of
the patches' own solves, on every replay of its graph."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field, rel_l2, rel_linf

pytestmark = pytest.mark.gpu


def _solver(g, p, precision):
    import pmlib3d_amd as pm
    patch = pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                     bound_cond=tuple(g.bound_cond), nghost=g.nghost)
    s = pm.PoissonSolver(patch, precision, order=p.poisson_order, kernel=p.poisson_kernel,
                         eps_radius=p.regularisation_radius, interp_order=g.interp_order)
    s.setup()
    return s


SCALAR = [
    # ncell, nghost, bound_cond, order (transform lengths all specialised)
    ((16, 16, 16), 0, (0, 0, 0), 4),
    ((16, 24, 40), 0, (0, 0, 0), 6),
    ((32, 32, 12), 2, (1, 1, 0), 4),
    ((32, 16, 32), 0, (1, 1, 1), 2),
]


@pytest.mark.parametrize("case", SCALAR, ids=[str(i) for i in range(len(SCALAR))])
def test_scalar_solve_matches_oracle_psi(case):
    import pmlib3d_amd as pm
    ncell, ng, bc, order = case
    print("case", case, flush=True)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 1, 2.0)
    w3 = blob_field(g, seed=sum(ncell))
    w3[0] = 0.0
    w3[1] = 0.0
    _, _, psi = O.poisson_solve(g, p, O.poisson_setup(g, p), w3, False, False, True)
    got = _solver(g, p, pm.FP64).solve_scalar(w3[2])
    assert rel_linf(got, psi[2]) <= 1e-12
    got32 = _solver(g, p, pm.FP32).solve_scalar(w3[2])
    assert rel_l2(got32, psi[2]) <= 1e-5


# every scalar z-pass kernel, forced: 1..16 the LDS-resident one with that
# many columns per block, 100 + TC the register-resident one (transform
# lengths 32 and 64, both register-capable)
Z1_KERNELS = [1, 2, 4, 8, 16, 104, 108, 116]
Z1_CASES = [
    ((32, 24, 32), 0, (0, 0, 0), 4),   # free z, zero upper half skipped
    ((28, 24, 28), 2, (0, 0, 0), 6),   # ghosts: Mz = 32 of Lz = 64
    ((32, 16, 32), 0, (0, 0, 1), 4),   # periodic z, Mz = Lz
    ((32, 16, 32), 0, (1, 1, 1), 2),   # all periodic: G formed in the pass
]


@pytest.mark.parametrize("kern", Z1_KERNELS)
@pytest.mark.parametrize("case", Z1_CASES, ids=[str(i) for i in range(len(Z1_CASES))])
def test_scalar_z_kernels_match_oracle_psi(case, kern, monkeypatch):
    import pmlib3d_amd as pm
    ncell, ng, bc, order = case
    print("case", case, "z1 kernel", kern, flush=True)
    monkeypatch.setenv("PMGPU_Z1_KERNEL", str(kern))
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, 1, 2.0)
    w3 = blob_field(g, seed=7 + sum(ncell))
    w3[0] = 0.0
    w3[1] = 0.0
    _, _, psi = O.poisson_solve(g, p, O.poisson_setup(g, p), w3, False, False, True)
    for rep in range(2):
        got = _solver(g, p, pm.FP64).solve_scalar(w3[2])
        assert rel_linf(got, psi[2]) <= 1e-12
        got32 = _solver(g, p, pm.FP32).solve_scalar(w3[2])
        assert rel_l2(got32, psi[2]) <= 1e-5


@pytest.mark.parametrize("grouped", [1, 0], ids=["grouped", "per_patch"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_batch_graph_matches_individual_solves_on_every_replay(precision, grouped, monkeypatch):
    import torch
    import pmlib3d_amd as pm
    print("precision", precision, "grouped", grouped, flush=True)
    monkeypatch.setenv("PMGPU_BATCH_GROUPED", str(grouped))
    rng = np.random.default_rng(4)
    sizes = [tuple(int(v) for v in rng.choice([16, 32, 48], 3)) for _ in range(7)]
    dev = torch.device("cuda:0")
    dt = torch.float64 if precision == 0 else torch.float32
    solvers, ins, outs, refs = [], [], [], []
    p = O.make_params(4, 1, 2.0)
    for i, n in enumerate(sizes):
        g = O.make_geom(n, 0, [-0.5, -0.5, -0.5], 1.0 / 128, (0, 0, 0))
        s = _solver(g, p, precision)
        w = torch.from_numpy(blob_field(g, seed=i)[2].ravel(order="F")).to(dev).to(dt)
        ref = torch.zeros_like(w)
        s.solve_scalar_device(w.data_ptr(), ref.data_ptr())
        solvers.append(s)
        ins.append(w)
        outs.append(torch.zeros_like(w))
        refs.append(ref)
    torch.cuda.synchronize()
    batch = pm.PoissonBatch(solvers)
    for rep in range(4):
        for o in outs:
            o.fill_(float("nan"))
        batch.solve_scalar([t.data_ptr() for t in ins], [t.data_ptr() for t in outs])
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.equal(o, r), "replay %d" % rep
        # eager work between replays
        ins[0].mul_(1.0)
        torch.cuda.synchronize()
    # new output buffers: the batch captures a new graph
    outs2 = [torch.zeros_like(w) for w in ins]
    batch.solve_scalar([t.data_ptr() for t in ins], [t.data_ptr() for t in outs2])
    torch.cuda.synchronize()
    for o, r in zip(outs2, refs):
        assert torch.equal(o, r)


@pytest.mark.parametrize("grouped", [1, 0], ids=["grouped", "per_patch"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_batch_mixed_patches_match_oracle(precision, grouped, monkeypatch):
    """Batch over free, ghosted, mixed and periodic patches of several
    lengths, in both forms (grouped: groups of one and of several patches,
    periodic ghost fills after the group launches): each patch's psi against
    the oracle."""
    import torch
    import pmlib3d_amd as pm
    print("precision", precision, "grouped", grouped, flush=True)
    monkeypatch.setenv("PMGPU_BATCH_GROUPED", str(grouped))
    dev = torch.device("cuda:0")
    dt = torch.float64 if precision == 0 else torch.float32
    cases = [((32, 32, 32), 0, (0, 0, 0), 4), ((32, 24, 32), 0, (0, 0, 0), 4), ((28, 24, 28), 2, (0, 0, 0), 4),
             ((32, 32, 32), 0, (0, 0, 1), 4), ((32, 16, 32), 0, (1, 1, 1), 4), ((48, 32, 16), 0, (0, 0, 0), 4),
             ((32, 32, 32), 0, (0, 0, 0), 4)]
    solvers, ins, outs, refs = [], [], [], []
    for i, (ncell, ng, bc, order) in enumerate(cases):
        print("patch", i, ncell, ng, bc, flush=True)
        g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
        p = O.make_params(order, 1, 2.0)
        w3 = blob_field(g, seed=11 + i)
        w3[0] = 0.0
        w3[1] = 0.0
        _, _, psi = O.poisson_solve(g, p, O.poisson_setup(g, p), w3, False, False, True)
        solvers.append(_solver(g, p, precision))
        w = torch.from_numpy(np.asfortranarray(w3[2]).ravel(order="F")).to(dev).to(dt)
        ins.append(w)
        outs.append(torch.full_like(w, float("nan")))
        refs.append(psi[2])
    batch = pm.PoissonBatch(solvers)
    for rep in range(3):
        batch.solve_scalar([t.data_ptr() for t in ins], [t.data_ptr() for t in outs])
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(outs, refs)):
            got = o.cpu().numpy().astype(np.float64).reshape(r.shape, order="F")
            if precision == 0:
                assert rel_linf(got, r) <= 1e-12, (rep, i)
            else:
                assert rel_l2(got, r) <= 1e-5, (rep, i)
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
