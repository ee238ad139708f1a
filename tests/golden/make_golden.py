"""Writes the golden fixtures of tests/golden/ from the compiled reference
(oracle/_ref/ref_driver, built from the reference sources by
oracle/Makefile.ref).  Run in the build container only:

    python tests/golden/make_golden.py

Outputs (small, committed):
  kat_ring_ref.json      rms velocity errors of the validation/poisson ring on
                         the pmlib_patch_adjust grids (order 6, eps 2h, ng 4)
  ref_solve_*.npz        inputs + reference outputs of small solves
  ref_particles_*.npz    inputs + reference outputs of P2M/M2P/remesh/advance
  ref_step_*.npz         inputs + reference outputs of whole two-ring vortex steps
  ref_vtk_*.npz          inputs + the .pvti / .vti bytes of pmlib_output_mesh
  ref_checkpoint_*.npz   inputs + the file bytes of output_particles
  ref_refine_*.npz       inputs + velocities of refinement-patch solves
  ref_m2m_*.npz          inputs + both directions of pmlib_interp_mesh_mesh
  ref_rhs_*.npz          inputs + dvort of the vortex client's part_deriv_vorticity
(`python tests/golden/make_golden.py formats` rewrites only the vtk and
checkpoint fixtures, `... multires` only the refinement and m2m ones, `...
rhs` only the right-hand-side ones.)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
import kat_ring as K  # noqa: E402
from tests.helpers import blob_field  # noqa: E402

SOLVE_CASES = {
    "free_o4_psi": ((16, 16, 16), 0, (0, 0, 0), 4, (False, False, True)),
    "free_ng2_o6_all": ((12, 10, 8), 2, (0, 0, 0), 6, (True, True, True)),
    "mixed_o4": ((8, 8, 12), 2, (1, 1, 0), 4, (True, False, True)),
    "periodic_o6": ((12, 8, 8), 0, (1, 1, 1), 6, (True, True, True)),
    "free_o2_r35": ((20, 12, 18), 0, (0, 0, 0), 2, (False, False, True)),
    # kernel 2 (the gradient of G, poisson/pmlib_poisson_setup.f:580-635 and
    # the kernel-2 branch of poisson/pmlib_poisson_solve.f:873-913)
    "free_k2_o4_all": ((16, 12, 10), 0, (0, 0, 0), 4, (True, True, False), 2),
    "free_ng2_k2_o6": ((10, 12, 8), 2, (0, 0, 0), 6, (False, False, False), 2),
    "mixed_k2_o8": ((8, 12, 10), 2, (1, 0, 1), 8, (True, True, False), 2),
    "periodic_k2_o2": ((12, 8, 10), 0, (1, 1, 1), 2, (True, True, False), 2),
    "free_k2_o10": ((12, 10, 14), 1, (0, 0, 0), 10, (False, True, False), 2),
    # order 100: the band-limited kernel Si(pi r / h) / (2 pi^2 r)
    # (poisson/pmlib_poisson_setup.f:1419-1505)
    "free_ng1_o100_psi": ((12, 10, 8), 1, (0, 0, 0), 100, (False, False, True)),
    "mixed_k2_o100": ((8, 10, 12), 0, (0, 1, 0), 100, (False, False, False), 2),
    "periodic_o100_psi": ((8, 12, 8), 0, (1, 1, 1), 100, (False, False, True)),
}


def solve_spec(spec):
    """(ncell, nghost, bound_cond, order, (reg_vort, reproj, calc_vecpot),
    kernel); kernel 1 when the spec has five fields."""
    return tuple(spec) if len(spec) == 6 else tuple(spec) + (1,)


def solve_case(name, spec):
    ncell, ng, bc, order, (rv, rp, cv), kernel = solve_spec(spec)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, kernel, 2.0)
    w = blob_field(g, seed=sum(ncell) + order)
    job = dict(O.geom_job(g, p), mode="solve", reg_vort=rv, reproj=rp, calc_vecpot=cv)
    raw, _ = O.ref_run(job, [w])
    ref = np.frombuffer(raw, dtype=np.float64)
    n = w.size
    outs = {"vel": ref[:n]}
    q = 1
    if rv or rp:
        outs["vort"] = ref[q * n:(q + 1) * n]
        q += 1
    if cv and kernel == 1:
        outs["psi"] = ref[q * n:(q + 1) * n]
    np.savez_compressed(os.path.join(HERE, "ref_solve_%s.npz" % name), vort_in=w,
                        spec=np.array(json.dumps(spec)),
                        **{k: v.reshape(w.shape, order="F") for k, v in outs.items()})


def particle_case(name, ncell, ng, bc, io, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / 8
    g = O.make_geom(ncell, ng, [-0.5, -0.5, -0.5], h, bc, io)
    xmin, xmax = np.array(g.xmin), np.array(g.xmax)
    npart = 300
    pos = rng.uniform(xmin - 2 * h, xmax + 2 * h, size=(npart, 3)).T.copy(order="F")
    pos[:, :40] = xmin[:, None] + ((rng.integers(0, 8, size=(3, 40)) - 1) + 0.5) * h
    pos[:, 40:60] = xmin[:, None] + rng.integers(0, 8, size=(3, 20)) * h
    if io == 4:   # keep M6* M2P inside its defined range (k3 + 2 >= 1 - ng)
        pos[2] = np.maximum(pos[2], xmin[2] - (ng + 0.5) * h)
    val = np.asfortranarray(rng.normal(size=(3, npart)))
    ext = O.mesh_extent(g)
    out = {"pos": pos, "val": val}
    raw, _ = O.ref_run(dict(O.geom_job(g), mode="p2m", npart=npart, interp_order=io), [pos, val])
    out["p2m"] = np.frombuffer(raw).reshape((3,) + ext, order="F")
    mesh = np.asfortranarray(rng.normal(size=(3,) + ext))
    out["mesh"] = mesh
    raw, _ = O.ref_run(dict(O.geom_job(g), mode="m2p", npart=npart, interp_order=io), [mesh, pos])
    out["m2p"] = np.frombuffer(raw).reshape((3, npart), order="F")
    vm = np.asfortranarray(rng.normal(size=(3,) + ext))
    dvm = np.asfortranarray(0.1 * rng.normal(size=vm.shape))
    velm = np.asfortranarray(rng.normal(size=vm.shape))
    out.update(rm_vort=vm, rm_dvort=dvm, rm_vel=velm)
    for ig in (0, 1):
        raw, _ = O.ref_run(dict(O.geom_job(g), mode="remesh", min_vort=1.2, min_dvort=0.15,
                                incl_ghost=bool(ig), interp_order=io), [vm, dvm, velm])
        n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
        r = np.frombuffer(raw[4:]).reshape((3, 3 * n), order="F")
        out["remesh%d" % ig] = r
    arrs = [pos] + [np.asfortranarray(rng.normal(size=(3, npart)) * s) for s in (1.0, 1.0, 0.5, 0.3, 0.2)]
    for k, a in zip(("a_pos", "a_vel", "a_vort", "a_dvort", "a_vel_rk", "a_dvort_rk"), arrs):
        out[k] = a
    for st in (1, 2):
        raw, _ = O.ref_run(dict(O.geom_job(g), mode="advance", npart=npart, stage=st, dtime=0.2,
                                rk_order=2, interp_order=io), arrs)
        n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
        out["advance%d" % st] = np.frombuffer(raw[4:]).reshape((3, 4 * n), order="F")
    out["geom"] = np.array(json.dumps([ncell, ng, bc, io]))
    np.savez_compressed(os.path.join(HERE, "ref_particles_%s.npz" % name), **out)


STEP_CASES = {
    # name: (n, bound_cond, nsteps, min_vort)
    "free16": (16, (0, 0, 0), 2, 1e-3),
    "mixed16": (16, (1, 1, 0), 1, 1e-3),
}


def step_case(name, spec):
    """The two-ring vortex step (P2M, order-6 solve, M2P, RK2 advance, then
    P2M + remesh) run by the reference's own routines, nsteps times."""
    from tests.helpers import ring_particles
    n, bc, nsteps, min_vort = spec
    h = 2.0 / n
    g = O.make_geom((n, n, n), 0, [-1.0] * 3, h, bc, 3)
    p = O.make_params(6, 1, 2.0)
    pos, vort = ring_particles(n)
    job = dict(O.geom_job(g, p), mode="step", npart=pos.shape[1], rk_order=2, dtime=0.5 * h,
               nsteps=nsteps, min_vort=min_vort, min_dvort=0.0)
    raw, _ = O.ref_run(job, [pos, vort])
    npart = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    out = np.frombuffer(raw[4:]).reshape((3, 3 * npart), order="F")
    np.savez_compressed(os.path.join(HERE, "ref_step_%s.npz" % name), pos_in=pos, vort_in=vort,
                        pos=out[:, :npart], vort=out[:, npart:2 * npart], vel=out[:, 2 * npart:],
                        spec=np.array(json.dumps(spec)))


def kat():
    cases = []
    for (h, ncell, xmin, stored) in K.KAT_CASES:
        g = O.make_geom(ncell, K.KAT_NGHOST, xmin, h, (0, 0, 0))
        p = O.make_params(K.KAT_ORDER, 1, K.KAT_EPS_RADIUS)
        w, u = K.ring_fields(xmin, h, ncell, K.KAT_NGHOST)
        raw, _ = O.ref_run(dict(O.geom_job(g, p), mode="solve", reg_vort=True), [w], timeout=1200)
        vel = np.frombuffer(raw, dtype=np.float64)[:w.size].reshape(w.shape, order="F")
        err = K.rms_velocity_error(vel, u, K.KAT_NGHOST)
        print("KAT h=%g reference rms error %.15e (convergence.dat stores %.12e)" % (h, err, stored))
        cases.append({"h": h, "ncell": list(ncell), "xmin": list(xmin), "rms_error": err,
                      "convergence_dat": stored})
    json.dump({"source": "compiled reference (oracle/_ref), order 6, eps 2h, nghost 4, reg_vort",
               "cases": cases}, open(os.path.join(HERE, "kat_ring_ref.json"), "w"), indent=1)


def vtk_case(name, ncell, ng, bc, xmin, dx, seed):
    """pmlib_output_mesh of a small mesh: the .pvti / .vti pair, byte for
    byte, with the input meshes."""
    import tempfile
    g = O.make_geom(ncell, ng, xmin, dx, bc)
    rng = np.random.default_rng(seed)
    ext = O.mesh_extent(g)
    w = np.asfortranarray(rng.normal(size=(3,) + ext))
    u = np.asfortranarray(rng.normal(size=(3,) + ext))
    with tempfile.TemporaryDirectory() as td:
        job = dict(O.geom_job(g, O.make_params(4, 1, 2.0)), mode="vtk", outdir=td + "/", tag="out")
        O.ref_run(job, [w, u])
        pvti = open(os.path.join(td, "out.pvti"), "rb").read()
        vti = open(os.path.join(td, "out_P00000.vti"), "rb").read()
    np.savez_compressed(os.path.join(HERE, "ref_vtk_%s.npz" % name), vort=w, vel=u,
                        spec=np.array(json.dumps([ncell, ng, bc, xmin, dx])),
                        pvti=np.frombuffer(pvti, dtype=np.uint8), vti=np.frombuffer(vti, dtype=np.uint8))


def checkpoint_case(name, npart, itime, time, dtime, seed):
    """output_particles of a particle set: the checkpoint file, byte for
    byte, with its inputs (some particles below a truncation of 0.5)."""
    import tempfile
    rng = np.random.default_rng(seed)
    pos = np.asfortranarray(rng.uniform(-1, 1, (3, npart)))
    vort = np.asfortranarray(rng.normal(size=(3, npart)))
    vort[:, ::5] *= 0.1
    g = O.make_geom((8, 8, 8), 0, [-1.0] * 3, 0.25, (0, 0, 0))
    with tempfile.TemporaryDirectory() as td:
        job = dict(O.geom_job(g), mode="partout", npart=npart, outdir=td + "/", tag="part", itime=itime,
                   time=time, dtime=dtime)
        O.ref_run(job, [pos, vort])
        dat = open(os.path.join(td, "part.dat"), "rb").read()
    np.savez_compressed(os.path.join(HERE, "ref_checkpoint_%s.npz" % name), pos=pos, vort=vort,
                        spec=np.array(json.dumps([npart, itime, time, dtime])),
                        dat=np.frombuffer(dat, dtype=np.uint8))


# refinement patches (level 2 on a level-1 parent): the difference kernels
# of poisson/pmlib_poisson_setup.f:446-528 and the other orders' blocks
REFINE_CASES = {
    # name: (ncell, nghost, bound_cond, order, kernel, level, parent)
    "free_o4_k1": ((12, 10, 8), 0, (0, 0, 0), 4, 1, 2, 1),
    "free_ng2_o6_k2": ((10, 12, 8), 2, (0, 0, 0), 6, 2, 2, 1),
    "mixed_o2_k1": ((8, 8, 12), 2, (1, 1, 0), 2, 1, 2, 1),
    "free_ng1_o10_k1_l3": ((12, 8, 10), 1, (0, 0, 0), 10, 1, 3, 1),
    "mixed_o8_k2": ((8, 10, 12), 0, (0, 1, 0), 8, 2, 2, 1),
}

# mesh-to-mesh between a patch and its parent (pmlib_topology_parent +
# pmlib_interp_mesh_mesh both ways)
M2M_CASES = {
    # name: (ncell, nghost, bound_cond, interp_order, level, clear)
    "m4_free": ((12, 8, 8), 2, (0, 0, 0), 3, 2, True),
    "m6_mixed": ((12, 12, 8), 3, (1, 1, 0), 4, 2, False),
    "m4_periodic": ((8, 8, 12), 0, (1, 1, 1), 3, 2, True),
    "m6_ratio4": ((16, 12, 12), 3, (0, 0, 0), 4, 3, True),
    "m4_ratio4_mixed": ((8, 16, 8), 1, (0, 1, 0), 3, 3, False),
}


def refine_case(name, spec):
    ncell, ng, bc, order, kernel, level, parent = spec
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, kernel, 2.0)
    w = blob_field(g, seed=sum(ncell) + order + level)
    job = dict(O.geom_job(g, p), mode="solve", level=level, parent_level=parent)
    raw, _ = O.ref_run(job, [w])
    vel = np.frombuffer(raw, dtype=np.float64).reshape(w.shape, order="F")
    np.savez_compressed(os.path.join(HERE, "ref_refine_%s.npz" % name), vort_in=w, vel=vel,
                        spec=np.array(json.dumps(spec)))


def m2m_case(name, spec, seed):
    ncell, ng, bc, io, level, clear = spec
    rng = np.random.default_rng(seed)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc, io)
    gp = O.topology_parent(g, level, 1)
    fine = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(g)))
    coarse = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(gp)))
    base = np.asfortranarray(rng.normal(size=(3,) + O.mesh_extent(g)))
    job = dict(O.geom_job(g), mode="m2m", level=level, parent_level=1, interp_order=io, clear=clear)
    raw, _ = O.ref_run(job, [fine, coarse, base])
    r = np.frombuffer(raw, dtype=np.float64)
    down = r[:coarse.size].reshape(coarse.shape, order="F")
    up = r[coarse.size:].reshape(base.shape, order="F")
    np.savez_compressed(os.path.join(HERE, "ref_m2m_%s.npz" % name), fine=fine, coarse=coarse, base=base,
                        down=down, up=up, parent_ncell=np.array(gp.ncell), spec=np.array(json.dumps(spec)))


# the vortex client's vorticity right-hand side (part_deriv_vorticity,
# vortex_codes/vortex_structures/source/particle_derivatives/part_deriv_vorticity.f)
RHS_CASES = {
    # name: (ncell, nghost, bound_cond, interp_order, fd_order)
    "free_fd4": ((12, 10, 8), 2, (0, 0, 0), 3, 4),
    "free_fd2": ((10, 8, 12), 1, (0, 0, 0), 3, 2),
    "mixed_fd4": ((8, 8, 12), 2, (1, 1, 0), 3, 4),
    "periodic_m6_fd4": ((8, 10, 8), 0, (1, 1, 1), 4, 4),
    "mixed_m6_fd2": ((12, 8, 10), 2, (0, 1, 0), 4, 2),
}


def rhs_case(name, spec, seed):
    ncell, ng, bc, io, fd = spec
    rng = np.random.default_rng(seed)
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc, io)
    E = (3,) + O.mesh_extent(g)
    w, u, dw = (np.asfortranarray(rng.normal(size=E)) for _ in range(3))
    raw, _ = O.ref_run(dict(O.geom_job(g), mode="rhs", fd_order=fd, visc=0.013, interp_order=io), [w, u, dw])
    out = np.frombuffer(raw, dtype=np.float64).reshape(E, order="F")
    np.savez_compressed(os.path.join(HERE, "ref_rhs_%s.npz" % name), vort=w, vel=u, dvort_in=dw, dvort=out,
                        visc=np.array(0.013), spec=np.array(json.dumps(spec)))


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: make -f oracle/Makefile.ref")
    if len(sys.argv) > 1 and sys.argv[1] == "formats":
        vtk_case("free_ng2", (5, 4, 3), 2, (0, 0, 0), [-0.5, -0.37, 0.1234], 1.0 / 7.0, 51)
        vtk_case("periodic", (6, 3, 4), 0, (1, 1, 1), [0.0, -1.0, -2.5e-3], 0.03125, 52)
        checkpoint_case("mixed", 37, 41, 0.37, 0.005, 53)
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "rhs":
        for q, (name, spec) in enumerate(RHS_CASES.items()):
            rhs_case(name, spec, 70 + q)
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "multires":
        for name, spec in REFINE_CASES.items():
            refine_case(name, spec)
        for q, (name, spec) in enumerate(M2M_CASES.items()):
            m2m_case(name, spec, 60 + q)
        sys.exit(0)
    if len(sys.argv) > 1:   # only the named solve cases
        for name in sys.argv[1:]:
            solve_case(name, SOLVE_CASES[name])
        sys.exit(0)
    for name, spec in SOLVE_CASES.items():
        solve_case(name, spec)
    for name, spec in STEP_CASES.items():
        step_case(name, spec)
    for name, spec in REFINE_CASES.items():
        refine_case(name, spec)
    for q, (name, spec) in enumerate(M2M_CASES.items()):
        m2m_case(name, spec, 60 + q)
    for q, (name, spec) in enumerate(RHS_CASES.items()):
        rhs_case(name, spec, 70 + q)
    particle_case("m4_free", (10, 8, 12), 0, (0, 0, 0), 3, 11)
    particle_case("m4_ng2", (8, 10, 8), 2, (0, 0, 0), 3, 12)
    particle_case("m4_mixed", (8, 8, 10), 0, (1, 1, 0), 3, 13)
    particle_case("m6_ng1", (10, 8, 8), 1, (0, 0, 0), 4, 14)
    particle_case("m6_periodic", (8, 8, 8), 0, (1, 1, 1), 4, 15)
    kat()
