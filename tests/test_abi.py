"""The C ABI library loads without a GPU and exports every entry point that
include/pmgpu.h declares; the Python binding covers all of them."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pmgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(pmgpu_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pmgpu_poisson_solve" in names and "pmgpu_poisson_setup" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from pmlib3d_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from pmlib3d_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)
    _lib.load()   # binds argtypes of every symbol


def test_no_oracle_in_the_product():
    """The package never imports or links the oracle (test infrastructure)."""
    pkg = os.path.join(ROOT, "pmlib3d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".f90", "Makefile")):
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in txt.replace("oracle/", "").lower() or f.endswith(".md"), f


def test_topology_parent_through_the_c_abi():
    """pmgpu_topology_parent is host-only: the parent cuboid of
    topology/pmlib_topology_parent.f:5-100 matches the oracle's, and its
    error paths (parent level 0, a cell count that is not a multiple of the
    ratio) return -1."""
    import ctypes as C
    from oracle import oracle as O
    import pmlib3d_amd as pm
    from pmlib3d_amd import _lib
    patch = pm.Patch(ncell=(16, 12, 8), dx=(0.125,) * 3, xmin=(-0.5, -0.6, -0.7), xmax=(1.5, 0.9, 0.3),
                     bound_cond=(0, 1, 0), nghost=2, level=3, parent=(1, 0))
    par = pm.pmlib_topology_parent(patch, interp_order=4)
    want = O.topology_parent(O.make_geom((16, 12, 8), 2, [-0.5, -0.6, -0.7], 0.125, (0, 1, 0), 4), 3, 1)
    assert par.ncell == tuple(want.ncell) and par.dx == tuple(want.dx) and par.nghost == want.nghost
    assert par.level == 3 and par.parent == (1, 0) and par.bound_cond == (0, 1, 0)
    g, out = patch.geom(4), _lib.Geom()
    L = _lib.load()
    assert L.pmgpu_topology_parent(C.byref(g), 3, 0, C.byref(out)) == -1
    g.ncell[1] = 10
    assert L.pmgpu_topology_parent(C.byref(g), 3, 1, C.byref(out)) == -1
