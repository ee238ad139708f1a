"""On-disk formats through the C ABI (host I/O only, no device): the VTK
pair of pmlib_output_mesh and the particle checkpoint of output_particles
must be byte-identical to the files the compiled reference writes from the
same inputs (tests/golden/ref_vtk_*.npz, ref_checkpoint_*.npz, written by
tests/golden/make_golden.py); the checkpoint reader keeps the reference's
truncation rule and order.
References: source/output/pmlib_output_mesh.f:5-188,
vortex_codes/vortex_structures/source/output/output_particles.f:50-69,
.../input/input_particles.f:81-135, 200-230."""
import glob
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _patch(ncell, ng, bc, xmin, dx):
    import pmlib3d_amd as pm
    return pm.Patch(xmin=tuple(xmin), xmax=tuple(x + n * dx for x, n in zip(xmin, ncell)), ncell=tuple(ncell),
                    dx=(dx,) * 3, bound_cond=tuple(bc), nghost=ng)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_vtk_*.npz"))),
                         ids=lambda p: os.path.basename(p)[8:-4])
def test_vtk_output_is_byte_identical(path, tmp_path):
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, xmin, dx = json.loads(str(d["spec"]))
    assert pm.pmlib_output_mesh(str(tmp_path) + "/", "out", _patch(ncell, ng, bc, xmin, dx), d["vort"],
                                d["vel"]) == 0
    assert open(tmp_path / "out.pvti", "rb").read() == bytes(d["pvti"])
    assert open(tmp_path / "out_P00000.vti", "rb").read() == bytes(d["vti"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_checkpoint_*.npz"))),
                         ids=lambda p: os.path.basename(p)[15:-4])
def test_particle_checkpoint_round_trip(path, tmp_path):
    import pmlib3d_amd as pm
    d = np.load(path, allow_pickle=False)
    npart, itime, time, dtime = json.loads(str(d["spec"]))
    f = str(tmp_path / "part.dat")
    pm.output_particles(f, itime, time, dtime, d["pos"], d["vort"])
    assert open(f, "rb").read() == bytes(d["dat"])
    # the reader on the reference's own file: header, truncation, order
    ref = str(tmp_path / "ref.dat")
    open(ref, "wb").write(bytes(d["dat"]))
    it, t, pos, vort, vmax = pm.input_particles(ref, trunc=0.0)
    assert it == itime + 1 and t == time + dtime
    assert np.array_equal(pos, d["pos"]) and np.array_equal(vort, d["vort"])
    mag = np.sqrt((d["vort"] ** 2).sum(0))
    keep = mag >= 0.5
    assert 0 < keep.sum() < npart
    it, t, pos, vort, vmax = pm.input_particles(ref, trunc=0.5)
    assert np.array_equal(pos, d["pos"][:, keep]) and np.array_equal(vort, d["vort"][:, keep])
    assert vmax == mag[keep].max()


def test_checkpoint_reader_refuses_bad_files(tmp_path):
    import pmlib3d_amd as pm
    from pmlib3d_amd import _lib
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\x04\x00\x00\x00\x01\x00")
    with pytest.raises(_lib.PmgpuError, match="bad header"):
        pm.input_particles(str(bad))
    d = np.load(sorted(glob.glob(os.path.join(GOLD, "ref_checkpoint_*.npz")))[0], allow_pickle=False)
    cut = tmp_path / "cut.dat"
    cut.write_bytes(bytes(d["dat"])[:-20])
    with pytest.raises(_lib.PmgpuError, match="truncated"):
        pm.input_particles(str(cut))
    with pytest.raises(_lib.PmgpuError, match="No particles above truncation"):
        full = tmp_path / "full.dat"
        full.write_bytes(bytes(d["dat"]))
        pm.input_particles(str(full), trunc=1e9)
