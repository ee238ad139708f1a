"""The allocation-free penalisation entry point
(pmgpu_penalise_device_noalloc, through the C ABI) against the allocating
pmgpu_penalise_device on the same inputs: vorticity and the five FP64 sums
bit-identical (the same kernels, the same reduction order), on two calls in
a row, the second on a smaller mesh, in FP64 and FP32; a mesh beyond the
fixed scratch is refused before any launch."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field

pytestmark = pytest.mark.gpu


def _inputs(n, ng, seed, dt):
    import torch
    import pmlib3d_amd as pm
    g = O.make_geom((n, n + 3, n - 2), ng, [-1, -1, -1], 2.0 / n, (0, 0, 0))
    patch = pm.Patch(xmin=tuple(g.xmin), xmax=tuple(g.xmax), ncell=tuple(g.ncell), dx=tuple(g.dx),
                     bound_cond=tuple(g.bound_cond), nghost=g.nghost)
    vel = blob_field(g, seed=seed, width=0.1)
    vort = blob_field(g, seed=seed + 1, width=0.1)
    chi = pm.brinkman_factor(pm.sphere_mask(patch, 0.25), 1e4, 0.5 * 2.0 / n)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev).to(dt)
         for k, a in (("vel", vel), ("vort", vort), ("chi", chi))}
    return patch, t


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_noalloc_penalise_matches_allocating_entry_point(precision):
    import torch
    import pmlib3d_amd as pm
    dt = torch.float64 if precision == 0 else torch.float32
    dev = torch.device("cuda:0")
    d_diag = torch.full((5,), float("nan"), dtype=torch.float64, device=dev)
    for call, (n, fd) in enumerate(((72, 4), (40, 2))):   # the second call on a smaller mesh
        print("call", call, "n", n, "fd_order", fd, "precision", precision, flush=True)
        patch, t = _inputs(n, 2, 11 + call, dt)
        w_ref = t["vort"].clone()
        ref = pm.penalise_device(patch, precision, t["vel"].data_ptr(), t["chi"].data_ptr(), w_ref.data_ptr(),
                                 us_const=(-1.0, 0.0, 0.0), relax=1.0, fd_order=fd)
        w = t["vort"].clone()
        pm.penalise_device_noalloc(patch, precision, t["vel"].data_ptr(), t["chi"].data_ptr(), w.data_ptr(),
                                   d_diag=d_diag.data_ptr(), us_const=(-1.0, 0.0, 0.0), relax=1.0, fd_order=fd)
        torch.cuda.synchronize()
        assert torch.equal(w, w_ref)
        got = d_diag.cpu().numpy()
        print("sums", got, "reference", ref, flush=True)
        assert np.array_equal(got, ref)
        # without sums: vorticity only, d_diag untouched
        d_diag.fill_(-7.0)
        w2 = t["vort"].clone()
        pm.penalise_device_noalloc(patch, precision, t["vel"].data_ptr(), t["chi"].data_ptr(), w2.data_ptr(),
                                   us_const=(-1.0, 0.0, 0.0), relax=1.0, fd_order=fd)
        torch.cuda.synchronize()
        assert torch.equal(w2, w_ref)
        assert bool((d_diag == -7.0).all())


def test_noalloc_penalise_refuses_a_mesh_beyond_its_scratch():
    import pmlib3d_amd as pm
    from pmlib3d_amd import _lib
    # ceil(512/64) * ceil(512/4) * 512 = 524288 blocks > 262144; refused
    # before any pointer is touched
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(512,) * 3, dx=(2.0 / 512,) * 3, nghost=2)
    with pytest.raises(_lib.PmgpuError, match="fixed scratch"):
        pm.penalise_device_noalloc(patch, 0, 1, 1, 1, d_diag=1)
