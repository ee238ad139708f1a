"""The on-disk formats written from device-resident state
(pmgpu_output_mesh / pmgpu_output_particles with device pointers): the same
bytes as the compiled reference's files (tests/golden/ref_vtk_*.npz,
ref_checkpoint_*.npz), from FP64 and FP32 device meshes, and from the
reference harness rebuilt on the Fortran drop-in (pmlib_output_mesh of
pmlib3d_amd/fortran/pmgpu_pmlib.f90).
References: source/output/pmlib_output_mesh.f:5-188,
vortex_codes/vortex_structures/source/output/output_particles.f:50-69."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_formats import _patch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
SHIM = os.path.join(ROOT, "tests", "fortran", "_build", "shim_driver")
VTK = sorted(glob.glob(os.path.join(GOLD, "ref_vtk_*.npz")))


@pytest.mark.parametrize("path", VTK, ids=lambda p: os.path.basename(p)[8:-4])
def test_vtk_output_from_device_meshes(path, tmp_path):
    import torch
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, xmin, dx = json.loads(str(d["spec"]))
    patch = _patch(ncell, ng, bc, xmin, dx)
    dev = torch.device("cuda:0")
    for prec, dt in ((pm.FP64, torch.float64), (pm.FP32, torch.float32)):
        # an FP32 mesh holds the same float32 values the writer produces
        w, u = (torch.from_numpy(np.ascontiguousarray(d[k].ravel(order="F"))).to(dev).to(dt) for k in ("vort", "vel"))
        out = tmp_path / ("p%d" % prec)
        out.mkdir()
        pm.output_mesh_device(str(out) + "/", "out", patch, prec, w.data_ptr(), u.data_ptr())
        assert open(out / "out.pvti", "rb").read() == bytes(d["pvti"])
        assert open(out / "out_P00000.vti", "rb").read() == bytes(d["vti"])


def test_checkpoint_from_device_particles(tmp_path):
    import torch
    import pmlib3d_amd as pm
    d = np.load(sorted(glob.glob(os.path.join(GOLD, "ref_checkpoint_*.npz")))[0], allow_pickle=False)
    npart, itime, time, dtime = json.loads(str(d["spec"]))
    dev = torch.device("cuda:0")
    p, w = (torch.from_numpy(np.ascontiguousarray(d[k].T)).to(dev) for k in ("pos", "vort"))
    f = str(tmp_path / "part.dat")
    pm.output_particles_device(f, itime, time, dtime, p.data_ptr(), w.data_ptr(), npart)
    assert open(f, "rb").read() == bytes(d["dat"])


@pytest.mark.skipif(not os.path.exists(SHIM), reason="tests/fortran/_build/shim_driver not built")
@pytest.mark.parametrize("path", VTK, ids=lambda p: os.path.basename(p)[8:-4])
def test_shim_pmlib_output_mesh_is_byte_identical(path, tmp_path):
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, xmin, dx = json.loads(str(d["spec"]))
    g = O.make_geom(ncell, ng, xmin, dx, bc)
    job = dict(O.geom_job(g, O.make_params(4, 1, 2.0)), mode="vtk", outdir=str(tmp_path) + "/", tag="out")
    O.ref_run(job, [d["vort"], d["vel"]], driver=SHIM, timeout=120)
    assert open(tmp_path / "out.pvti", "rb").read() == bytes(d["pvti"])
    assert open(tmp_path / "out_P00000.vti", "rb").read() == bytes(d["vti"])
