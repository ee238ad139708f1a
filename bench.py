"""Benchmark of the particle-mesh hot path on one MI355X.

Headline (`value`, one "step" = one full velocity solve): 256^3 cells on
[-1,1]^3 (h = 1/128), free-space boundaries, doubled FFT grid 512^3 handled
with implicit zero padding, three vorticity components in one solve,
regularised Green's function of order 4 with eps = 2h built in FP64.
Vorticity: 32 seeded Gaussian blobs (centres U[-0.6,0.6]^3, sigma U[2h,6h],
strengths N(0,1)^3, seed 3), projected divergence-free with the solver's
spectral reprojection.  Inputs and outputs stay resident in HBM; the timed
region contains only solves.  FP32 (value) beside FP64.

Beside it, on the same GPU (each with its FP32-vs-FP64 error):
  ring64     64^3 Gaussian ring, FP64 stream function + velocity solve
  step128    two-ring vortex step at 128^3 on all 128^3 jittered particles
             (the particle set is re-loaded before every timed step)
  batch24    24 free-space patches of uneven sizes, one captured graph
  mixed384   384^2 x 384 periodic-x,y / free-z solve of a turbulence-like
             field, then remeshed FP32 vortex steps on that grid
  rhs256     the vorticity right-hand side at 256^3
  penalised256  penalisation_solve (iterated Brinkman penalisation and
             solves) on the 256^3 mesh around a sphere
  diagnostics   FP64 ring diagnostics of 2.1M particles, mesh energies
and the compiled reference on the host's cores (cpu_baseline).

Algorithmic bytes follow the 5-pass implicit-padding schedule (X R2C, Y,
fused Z, Y inverse, X inverse), G-hat counted as a real half-spectrum of
Hx * Ly * Lz values.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 256
ORDER = 4
EPS_RADIUS = 2.0
PEAK_HBM_GBS = 8000.0
STEP_MIN_VORT = 1e-6   # remesh threshold of the two-ring step (peak |omega| is ~41)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def blob_vorticity(n, seed=3):
    import numpy as np
    h = 2.0 / n
    rng = np.random.default_rng(seed)
    w = np.zeros((3, n, n, n), dtype=np.float64, order="F")
    x = -1.0 + (np.arange(n) + 0.5) * h
    for _ in range(32):
        c = rng.uniform(-0.6, 0.6, 3)
        sig = rng.uniform(2 * h, 6 * h)
        s = rng.normal(size=3)
        r = int(math.ceil(5 * sig / h))
        idx = [np.clip(np.arange(int((c[d] + 1) / h) - r, int((c[d] + 1) / h) + r + 1), 0, n - 1)
               for d in range(3)]
        idx = [np.unique(i) for i in idx]
        X, Y, Z = np.meshgrid(x[idx[0]], x[idx[1]], x[idx[2]], indexing="ij")
        gval = np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sig * sig))
        for q in range(3):
            w[q][np.ix_(*idx)] += s[q] * gval
    return w


def solve_bytes(M, L, cin, cout, s, s_g):
    """Bytes one solve must move with the 5-pass implicit-padding schedule:
    M, L solve-domain and transform lengths per axis, cin / cout components
    in and out, s the mesh word, s_g the G-hat word (Hx * Ly * Lz values).
    Returns (total, fused z pass)."""
    Mx, My, Mz = M
    Lx, Ly, Lz = L
    Hx = Lx // 2 + 1
    c = 2 * s
    x_f = cin * Mx * My * Mz * s + cin * Hx * My * Mz * c
    y_f = cin * Hx * My * Mz * c + cin * Hx * Ly * Mz * c
    z_p = cin * Hx * Ly * Mz * c + Hx * Ly * Lz * s_g + cout * Hx * Ly * Mz * c
    y_i = cout * Hx * Ly * Mz * c + cout * Hx * My * Mz * c
    x_i = cout * Hx * My * Mz * c + cout * Mx * My * Mz * s
    return x_f + y_f + z_p + y_i + x_i, z_p


def free_bytes(n, cin, cout, s, s_g):
    M = tuple(n) if isinstance(n, (tuple, list)) else (n,) * 3
    return solve_bytes(M, tuple(2 * m for m in M), cin, cout, s, s_g)


def rel_l2_t(a, b):
    return float(((a.double() - b.double()) ** 2).sum().sqrt() / (b.double() ** 2).sum().sqrt())


# --------------------------------------------------------------- CPU reference
def cpu_ranks():
    """MPI ranks of the CPU reference: the cores this process may run on,
    at most 16 (a one-GPU job's CPU share on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _reps(out):
    line = [l for l in out.splitlines() if l.startswith("REPS")][0]
    return [float(v) for v in line.split()[1:]]


def cpu_reference_solve(w64, n, order, calc_vecpot=False, nrep=6):
    """The compiled reference (oracle/_ref, FP64, native Glassman FFT) on a
    free-space n^3 solve under MPI with cpu_ranks() ranks: one warm-up
    repetition, then the median of the remaining nrep - 1; G-hat set-up
    timed separately and not counted.  Falls back to the C oracle (port,
    1 thread) when the reference is not built."""
    import numpy as np
    from oracle import oracle as O
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, 2.0 / n, (0, 0, 0))
    p = O.make_params(order, 1, EPS_RADIUS)
    ranks = cpu_ranks()
    what = "psi and u" if calc_vecpot else "u"
    if O.ref_available() and os.path.exists(O.MPIEXEC):
        try:
            t = time.time()
            _, out = O.ref_run(dict(O.geom_job(g, p), mode="solve", nrep=nrep, calc_vecpot=bool(calc_vecpot)),
                               [w64], nproc=ranks, timeout=400)
            reps = _reps(out)
            med = float(np.median(reps[1:]))
            setup_s = float([l for l in out.splitlines() if l.startswith("TIMING")][0]
                            .split("setup_s=")[1].split()[0])
            log("cpu reference %d^3 solve: ranks %d, reps %s s, setup %.2f s (wall %.1f s)" %
                (n, ranks, " ".join("%.3f" % r for r in reps), setup_s, time.time() - t))
            return {"value": 1.0 / med, "unit": "solves/s", "cores": ranks, "kind": "reference",
                    "sample": "%d^3 free-space solve (3 components, %s, order %d, FP64, native Glassman FFT; "
                              "FFTW absent) with mpiexec -n %d (the job's CPU cores); 1 warm-up + median "
                              "of %d; setup %.2fs not counted" % (n, what, order, ranks, nrep - 1, setup_s)}
        except Exception as e:  # pragma: no cover - reported, not hidden
            log("cpu reference failed (%s); timing the C oracle instead" % e)
    n2 = min(n, 128)
    g2 = O.make_geom((n2,) * 3, 0, [-1.0] * 3, 2.0 / n2, (0, 0, 0))
    step = n // n2
    w2 = np.asfortranarray(w64[:, ::step, ::step, ::step])
    Gh = O.poisson_setup(g2, p)
    t = time.time()
    O.poisson_solve(g2, p, Gh, w2, False, False, calc_vecpot)
    dt = time.time() - t
    scale = (2 * n) ** 3 * math.log(2 * n) / ((2 * n2) ** 3 * math.log(2 * n2))
    return {"value": 1.0 / (dt * scale), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "one %d^3 solve of the C oracle (1 thread), scaled by L^3 log L to %d^3" % (n2, n)}


def cpu_reference_step(n=128, nsteps=6):
    """The two-ring vortex step of the compiled reference at n^3 with all
    n^3 particles in every step (the harness re-loads them before each),
    under MPI with cpu_ranks() ranks (each keeps its subdomain's
    particles): one warm-up step, the median of the rest."""
    import numpy as np
    from oracle import oracle as O
    from tests.helpers import ring_particles
    if not (O.ref_available() and os.path.exists(O.MPIEXEC)):
        return None
    h = 2.0 / n
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, h, (0, 0, 0), 3)
    p = O.make_params(6, 1, 2.0)
    pos, vort = ring_particles(n)
    ranks = cpu_ranks()
    try:
        t = time.time()
        _, out = O.ref_run(dict(O.geom_job(g, p), mode="step", npart=pos.shape[1], rk_order=2,
                                dtime=0.5 * h, nsteps=nsteps, min_vort=STEP_MIN_VORT, min_dvort=0.0,
                                reload=True), [pos, vort], nproc=ranks, timeout=400)
    except Exception as e:  # pragma: no cover - reported, not hidden
        log("cpu reference step failed (%s)" % e)
        return None
    reps = _reps(out)
    med = float(np.median(reps[1:]))
    log("cpu reference step: ranks %d, steps %s s (wall %.1f s); %s" %
        (ranks, " ".join("%.3f" % r for r in reps), time.time() - t,
         [l for l in out.splitlines() if l.startswith("TIMING")][0]))
    return {"value": pos.shape[1] / med / 1e6, "unit": "Mparticle-steps/s", "cores": ranks,
            "kind": "reference",
            "sample": "two-ring vortex step at %d^3 on all %d particles (order-6 solve, M'4, RK2, remesh), "
                      "FP64, native Glassman FFT, mpiexec -n %d (the job's CPU cores); 1 warm-up + median "
                      "of %d" % (n, pos.shape[1], ranks, nsteps - 1)}


# ------------------------------------------------------------------ GPU parts
def time_solves(s, d_in, d_out, k, warm=2, **kw):
    """Seconds per solve over k stream-ordered solves (after warm-up)."""
    import torch
    for _ in range(warm):
        s.solve_device(d_in.data_ptr(), d_out.data_ptr(), **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        s.solve_device(d_in.data_ptr(), d_out.data_ptr(), **kw)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / k


def ring64_benchmark(reps):
    """64^3 Gaussian ring on [-1,1]^3 (R = 0.5, sigma = 0.1, Gamma = 1, axis
    z), order 4, eps = 2h, free space: stream function and velocity
    (calc_vecpot) per solve, FP64 (the plain form) and FP32."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    from oracle import oracle as O
    from tests.helpers import gaussian_ring
    n = 64
    h = 2.0 / n
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, h, (0, 0, 0))
    w = gaussian_ring(g)
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3)
    dev = torch.device("cuda:0")
    res = {"workload": "64^3 free-space solve of a Gaussian ring (R=0.5, sigma=0.1, Gamma=1, axis z) on "
                       "[-1,1]^3, order 4, eps 2h, stream function + velocity (calc_vecpot)"}
    outs = {}
    for name, prec, dt, sz in (("fp64", pm.FP64, torch.float64, 8), ("fp32", pm.FP32, torch.float32, 4)):
        s = pm.PoissonSolver(patch, prec, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        s.setup()
        d_w = torch.from_numpy(w.ravel(order="F")).to(dev).to(dt)
        d_v, d_p = torch.empty_like(d_w), torch.empty_like(d_w)
        sec = time_solves(s, d_w, d_v, reps, d_vecpot=d_p.data_ptr(), calc_vecpot=True)
        nbytes = free_bytes(n, 3, 6, sz, sz)[0]
        res[name] = {"solves_per_s": 1.0 / sec, "ms": 1e3 * sec, "algorithmic_bytes": nbytes,
                     "achieved_GBs": nbytes / sec / 1e9, "frac_hbm": nbytes / sec / 1e9 / PEAK_HBM_GBS}
        outs[name] = (d_v.clone(), d_p.clone())
        del s
    res["fp32"]["rel_l2_vs_fp64"] = {"vel": rel_l2_t(outs["fp32"][0], outs["fp64"][0]),
                                     "psi": rel_l2_t(outs["fp32"][1], outs["fp64"][1])}
    log("ring64: fp64 %.0f solves/s (%.3f ms), fp32 %.0f solves/s; fp32 err vel %.2e psi %.2e" %
        (res["fp64"]["solves_per_s"], res["fp64"]["ms"], res["fp32"]["solves_per_s"],
         res["fp32"]["rel_l2_vs_fp64"]["vel"], res["fp32"]["rel_l2_vs_fp64"]["psi"]))
    return res, np.asfortranarray(w)


def step_benchmark(steps, n=128):
    """The two-ring vortex step at n^3 (128^3 particles on the jittered
    cell-centre lattice, order-6 solve, M'4, RK2, dt = h/2, remesh with
    min_vort = 1e-6), every array resident in HBM.  Each timed step starts
    from the full lattice (re-loaded device to device outside the timed
    region), so every step advances n^3 particles.  FP64 throughout, and the
    FP32 mesh side (particles FP64); the FP32 error is that of one step's
    velocities and vorticities against FP64, particles matched by position."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    from tests.helpers import ring_particles
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3)
    pos, vort = ring_particles(n)
    npart = pos.shape[1]
    dev = torch.device("cuda:0")
    d_pos = torch.from_numpy(pos.ravel(order="F")).to(dev)
    d_vort = torch.from_numpy(vort.ravel(order="F")).to(dev)
    res = {"workload": "two-ring vortex step, 128^3 free, all 128^3 = %d jittered particles every step "
                       "(re-loaded before each), order 6, eps 2h, M'4, RK2 dt=h/2, remesh min_vort=1e-6; "
                       "FP32 = FP32 meshes, solve and interpolation (FP32 weights; deposition in per-row FP32 partials, Kahan-added), FP64 particles" % npart}
    fetched = {}
    for name, prec in (("fp64", pm.FP64), ("fp32", pm.FP32)):
        st = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3,
                           capacity=max(npart, n ** 3), precision=prec)
        st.load_device(d_pos, d_vort, npart)
        st.step(0.5 * h, STEP_MIN_VORT)  # warm-up
        fetched[name] = st.fetch()
        kept = st.np
        total = 0.0
        for _ in range(steps):
            st.load_device(d_pos, d_vort, npart)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st.step(0.5 * h, STEP_MIN_VORT)
            torch.cuda.synchronize()
            total += time.perf_counter() - t0
        sec = total / steps
        res[name] = {"Mparticle_steps_per_s": npart / sec / 1e6, "ms_per_step": 1e3 * sec, "particles": npart,
                     "particles_after_remesh": kept}
        del st
        # the same step with device-side counts, recorded once as a HIP graph
        gs = pm.GraphVortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3,
                                capacity=max(npart, n ** 3), precision=prec)
        gs.load_device(d_pos, d_vort, npart)
        gs.step(0.5 * h, STEP_MIN_VORT)  # records the graph
        gkept, status = gs.sync()
        total = 0.0
        for _ in range(steps):
            gs.load_device(d_pos, d_vort, npart)
            t0 = time.perf_counter()
            gs.step(0.5 * h, STEP_MIN_VORT)
            gs.sync()
            total += time.perf_counter() - t0
        gsec = total / steps
        res[name]["graph"] = {"Mparticle_steps_per_s": npart / gsec / 1e6, "ms_per_step": 1e3 * gsec,
                              "particles_after_remesh": gkept, "status": status,
                              "same_count_as_eager": gkept == kept}
        del gs
    (p64, w64, v64), (p32, w32, v32) = fetched["fp64"], fetched["fp32"]
    key = {tuple(c): i for i, c in enumerate(p64.T)}
    pairs = [(key[tuple(c)], j) for j, c in enumerate(p32.T) if tuple(c) in key]
    i64, i32 = (np.array(x) for x in zip(*pairs))
    rl2 = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    res["fp32"]["rel_l2_vs_fp64"] = {"vel": rl2(v32[:, i32], v64[:, i64]), "vort": rl2(w32[:, i32], w64[:, i64]),
                                     "matched_particles": len(pairs)}
    log("step128: fp64 %.1f, fp32 %.1f Mparticle-steps/s (%.3f / %.3f ms, %d particles); fp32 err vel %.2e" %
        (res["fp64"]["Mparticle_steps_per_s"], res["fp32"]["Mparticle_steps_per_s"], res["fp64"]["ms_per_step"],
         res["fp32"]["ms_per_step"], npart, res["fp32"]["rel_l2_vs_fp64"]["vel"]))
    log("step128 graph: fp64 %.1f, fp32 %.1f Mparticle-steps/s (%.3f / %.3f ms)" %
        (res["fp64"]["graph"]["Mparticle_steps_per_s"], res["fp32"]["graph"]["Mparticle_steps_per_s"],
         res["fp64"]["graph"]["ms_per_step"], res["fp32"]["graph"]["ms_per_step"]))
    return res


def ring_patch(n, h, rng):
    """x component of a Gaussian ring's vorticity on a patch centred at the
    origin: R = 0.3 min extent, sigma = 0.15 R, Gamma = 1, axis uniform on
    the sphere."""
    import numpy as np
    ext = [v * h for v in n]
    R = 0.3 * min(ext)
    sig = 0.15 * R
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    c = [-e / 2 + (np.arange(v) + 0.5) * h for e, v in zip(ext, n)]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    P = np.stack([X, Y, Z])
    z = np.tensordot(axis, P, axes=1)
    radial = P - axis[:, None, None, None] * z
    rho = np.sqrt((radial ** 2).sum(axis=0))
    wt = 1.0 / (np.pi * sig ** 2) * np.exp(-((rho - R) ** 2 + z ** 2) / sig ** 2)
    eth = np.cross(axis[:, None, None, None], radial, axis=0) / np.maximum(rho, 1e-30)
    return np.asfortranarray(wt * eth[0])


def batch_benchmark(steps, warmup, npatch=24, seed=4):
    """24 free-space patches, per-axis sizes drawn from {32,48,64,96,128,160}
    (seeded), h = 1/128, single-component solves of order 4 (eps = 2h), all
    patches replayed as one captured HIP graph; FP32 timed, FP64 beside it
    for the rate and the error."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    rng = np.random.default_rng(seed)
    h = 1.0 / 128
    sizes = [tuple(int(v) for v in rng.choice([32, 48, 64, 96, 128, 160], 3)) for _ in range(npatch)]
    fields = [ring_patch(n, h, rng) for n in sizes]
    dev = torch.device("cuda:0")
    res = {"workload": "24 free-space patches, sizes per axis from {32,48,64,96,128,160} (seed 4), h=1/128, "
                       "one Gaussian ring each (R=0.3 min extent, sigma=0.15R, random axis), single-component psi "
                       "solve, order 4, eps 2h, one captured HIP graph per batch", "patches": sizes}
    outs = {}
    for name, prec, dt, sz in (("fp32", pm.FP32, torch.float32, 4), ("fp64", pm.FP64, torch.float64, 8)):
        solvers, ins, o = [], [], []
        for n, f in zip(sizes, fields):
            patch = pm.Patch(xmin=tuple(-v * h / 2 for v in n), xmax=tuple(v * h / 2 for v in n), ncell=n,
                             dx=(h,) * 3)
            sv = pm.PoissonSolver(patch, prec, order=4, kernel=1, eps_radius=2.0, interp_order=3)
            sv.setup()
            solvers.append(sv)
            ins.append(torch.from_numpy(f.ravel(order="F")).to(dev).to(dt))
            o.append(torch.zeros_like(ins[-1]))
        batch = pm.PoissonBatch(solvers)
        pi, po = [t.data_ptr() for t in ins], [t.data_ptr() for t in o]
        for _ in range(warmup):
            batch.solve_scalar(pi, po)
        torch.cuda.synchronize()
        k = steps if name == "fp32" else max(2, steps // 4)
        t0 = time.perf_counter()
        for _ in range(k):
            batch.solve_scalar(pi, po)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / k
        nbytes = sum(free_bytes(n, 1, 1, sz, sz)[0] for n in sizes)
        res[name] = {"batches_per_s": 1.0 / sec, "ms_per_batch": 1e3 * sec, "algorithmic_bytes": nbytes,
                     "achieved_GBs": nbytes / sec / 1e9, "frac_hbm": nbytes / sec / 1e9 / PEAK_HBM_GBS}
        outs[name] = o
        del batch, solvers
    res["fp32"]["rel_l2_vs_fp64_max_over_patches"] = max(rel_l2_t(a, b) for a, b in zip(outs["fp32"], outs["fp64"]))
    log("batch24: fp32 %.1f batches/s (%.3f ms, %.0f GB/s), fp64 %.1f; fp32 err %.2e" %
        (res["fp32"]["batches_per_s"], res["fp32"]["ms_per_batch"], res["fp32"]["achieved_GBs"],
         res["fp64"]["batches_per_s"], res["fp32"]["rel_l2_vs_fp64_max_over_patches"]))
    return res


def mixed384_benchmark(reps, nsteps, n=384):
    """384^2 x 384 on [-1,1]^3 (h = 2/384), periodic x and y, free z (FFT
    384 x 384 x 768 with implicit padding in z), three components, order 4,
    eps 2h.  Vorticity: seeded random solenoidal Fourier modes with
    E(k) ~ k^4 exp(-2 (k/8)^2) times exp(-z^2 / 0.3^2)
    (tests.helpers.turbulence_vorticity).  FP32 and FP64 solve rates and
    the FP32 error against FP64; then nsteps remeshed FP32 vortex steps on
    that grid (particles seeded on every cell with |omega| > 1e-6 max|omega|,
    M'4, RK2, dt = h/2)."""
    import torch
    import pmlib3d_amd as pm
    from tests.helpers import turbulence_vorticity
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, bound_cond=(1, 1, 0))
    E = pm.mesh_extent(patch, 3)
    dev = torch.device("cuda:0")
    t = time.time()
    w_int = turbulence_vorticity(n, torch_device=dev)   # (3, n, n, n) flat, Fortran order
    w64 = torch.zeros(3 * E[0] * E[1] * E[2], dtype=torch.float64, device=dev)
    lo = [(E[d] - n) // 2 for d in range(3)]
    w64.view(E[2], E[1], E[0], 3)[lo[2]:lo[2] + n, lo[1]:lo[1] + n, lo[0]:lo[0] + n, :] = \
        w_int.view(n, n, n, 3)
    del w_int
    wmax = float(w64.abs().max())
    log("mixed384: field built in %.1f s, max |omega| %.3f" % (time.time() - t, wmax))
    res = {"workload": "384^2 x 384 on [-1,1]^3, periodic x,y, free z (FFT 384x384x768, implicit z padding), 3 "
                       "components, order 4, eps 2h; seeded solenoidal random Fourier modes E(k)~k^4 exp(-2(k/8)^2) "
                       "times exp(-z^2/0.3^2)"}
    M, L = (n, n, n), (n, n, 2 * n)
    vel = {}
    for name, prec, dt, sz in (("fp32", pm.FP32, torch.float32, 4), ("fp64", pm.FP64, torch.float64, 8)):
        s = pm.PoissonSolver(patch, prec, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        s.setup()
        d_w = w64.to(dt)
        d_v = torch.zeros_like(d_w)
        sec = time_solves(s, d_w, d_v, reps if name == "fp32" else max(2, reps // 2))
        nbytes = solve_bytes(M, L, 3, 3, sz, sz)[0]
        res[name] = {"solves_per_s": 1.0 / sec, "ms": 1e3 * sec, "algorithmic_bytes": nbytes,
                     "achieved_GBs": nbytes / sec / 1e9, "frac_hbm": nbytes / sec / 1e9 / PEAK_HBM_GBS}
        vel[name] = d_v
        del s, d_w
    res["fp32"]["rel_l2_vs_fp64"] = rel_l2_t(vel["fp32"], vel["fp64"])
    log("mixed384 solve: fp32 %.1f solves/s (%.3f ms), fp64 %.1f; fp32 err %.2e" %
        (res["fp32"]["solves_per_s"], res["fp32"]["ms"], res["fp64"]["solves_per_s"], res["fp32"]["rel_l2_vs_fp64"]))
    del vel
    if nsteps > 0:
        torch.cuda.empty_cache()
        st = pm.VortexStep(patch, order=4, eps_radius=2.0, rk_order=2, interp_order=3,
                           capacity=E[0] * E[1] * E[2], precision=pm.FP32)
        w32 = w64.float()
        npart0 = st.seed_from_mesh(w32.data_ptr(), 1e-6 * wmax)
        del w32
        torch.cuda.synchronize()
        counted = 0
        t0 = time.perf_counter()
        for _ in range(nsteps):
            counted += st.np
            st.step(0.5 * h, 1e-6 * wmax)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        res["steps"] = {"steps": nsteps, "seconds": sec, "ms_per_step": 1e3 * sec / nsteps,
                        "particles_initial": npart0, "particles_final": st.np,
                        "Mparticle_steps_per_s": counted / sec / 1e6, "dtype": "f32 mesh side, f64 particles",
                        "workload": "remeshed vortex steps on the mixed-BC grid: M'4, RK2, dt=h/2, order 4, "
                                    "remesh |omega| > 1e-6 max|omega|"}
        log("mixed384 steps: %d steps in %.2f s (%.1f ms/step), %d -> %d particles, %.1f Mparticle-steps/s" %
            (nsteps, sec, 1e3 * sec / nsteps, npart0, st.np, counted / sec / 1e6))
        del st
    return res


def rhs_benchmark(reps, n=256):
    """Vorticity right-hand side (stretching + diffusion, 4th-order
    differences) on a 256^3 free patch with two ghost layers, FP64 and FP32,
    timed with HIP events over `reps` launches on the null stream.
    Algorithmic bytes per launch: read w and u, write dw on the interior
    cells, 9 s bytes per cell (the stencil re-reads are cache traffic)."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, nghost=2)
    ext = (3,) + pm.mesh_extent(patch, 3)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(7)
    res = {"workload": "256^3 free patch, 2 ghost layers, 4th-order differences, visc 0.01, "
                       "seeded normal w and u"}
    outs = {}
    for name, prec, dt, s in (("fp64", pm.FP64, torch.float64, 8), ("fp32", pm.FP32, torch.float32, 4)):
        if name == "fp64":
            w = torch.randn(int(np.prod(ext)), dtype=dt, device=dev, generator=g)
            u = torch.randn_like(w)
        else:
            w, u = outs["w"].float(), outs["u"].float()
        dw = torch.zeros_like(w)
        run = lambda: pm.vorticity_rhs_device(patch, prec, w.data_ptr(), u.data_ptr(), dw.data_ptr(), 0.01,
                                              fd_order=4, interp_order=3)
        for _ in range(3):
            run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 9 * s * n ** 3
        res[name] = {"ms": ms, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / ms / 1e6,
                     "frac_hbm": nbytes / ms / 1e6 / PEAK_HBM_GBS}
        if name == "fp64":
            outs.update(w=w, u=u, dw=dw)
        else:
            res[name]["rel_l2_vs_fp64"] = rel_l2_t(dw, outs["dw"])
    return res


def penal256_benchmark(reps, n=256, ng=2):
    """The penalised solve of the 256^3 configuration: a 252^3 interior with
    the 2 ghost layers the 4th-order curl needs (a 256^3 mesh, FFT 512^3),
    free space, order 4, eps 2h; the 32-blob vorticity reprojected, its
    velocity plus the free stream U = (1, 0, 0); a sphere of radius 0.25 at
    the origin with a tanh mask smoothed over 2h and the finite-lambda factor
    lambda dt chi / (1 + lambda dt chi), lambda = 1e4, dt = h / 2; then
    penalisation_solve with the reference sphere case's settings
    (vortex_codes/vortex_penalisation/setup_sphere.in: tolerance 0.01,
    relaxation 1, at most 1000 iterations, reprojection on).  Each timed call
    starts from the same inputs (restored outside the events)."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    N_int = n - 2 * ng
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0 + ng * h,) * 3, xmax=(1.0 - ng * h,) * 3, ncell=(N_int,) * 3, dx=(h,) * 3,
                     nghost=ng)
    dev = torch.device("cuda:0")
    w = torch.from_numpy(blob_vorticity(n).ravel(order="F")).to(dev)
    chi = torch.from_numpy(np.ascontiguousarray(
        pm.brinkman_factor(pm.sphere_mask(patch, 0.25), 1e4, 0.5 * h).ravel(order="F"))).to(dev)
    res = {"workload": "penalised solve: 252^3 interior + 2 ghost layers (256^3 mesh, FFT 512^3), free, "
                       "order 4, eps 2h, 32 reprojected blobs + U=(1,0,0), sphere r=0.25 (tanh mask over 2h, "
                       "lambda=1e4, dt=h/2), penalisation_solve with tolerance 0.01, relaxation 1, "
                       "reprojection (the reference sphere case's settings)"}
    kw = dict(relaxation=1.0, tolerance=0.01, max_iter=1000, fd_order=4, dtime=0.5 * h, vel_inf=(1.0, 0.0, 0.0),
              reproject=True)
    outs = {}
    for name, prec, dt, sz in (("fp64", pm.FP64, torch.float64, 8), ("fp32", pm.FP32, torch.float32, 4)):
        s = pm.PoissonSolver(patch, prec, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        s.setup()
        w0 = w.to(dt)
        v0 = torch.empty_like(w0)
        wp = torch.empty_like(w0)
        s.solve_device(w0.data_ptr(), v0.data_ptr(), wp.data_ptr(), None, reg_vort=False, reproj=True)
        w0 = wp.clone()
        v0 = v0.view(-1, 3)
        v0[:, 0] += 1.0
        v0 = v0.view(-1)
        c = chi.to(dt)
        vel, vort = torch.empty_like(v0), torch.empty_like(w0)
        times, iters = [], []
        for r in range(reps + 1):
            vel.copy_(v0)
            vort.copy_(w0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f, it, rr = pm.penalisation_solve_device(s, c.data_ptr(), vel.data_ptr(), vort.data_ptr(), **kw)
            e1.record()
            torch.cuda.synchronize()
            if r > 0:
                times.append(e0.elapsed_time(e1))
                iters.append(it)
        ms = float(np.median(times))
        E3 = n ** 3
        solve_b = free_bytes((n,) * 3, 3, 3, sz, sz)[0]
        it = iters[-1]
        # per iteration: penalise (chi, vel_pen, u_pen read; vort_pen read and
        # written: 13 s per cell) + one solve; once: vel_pen (6 s), the add (9 s),
        # the final solve, the free stream (6 s)
        nbytes = it * (13 * sz * E3 + solve_b) + 21 * sz * E3 + solve_b
        res[name] = {"ms_per_penalised_solve": ms, "penalised_solves_per_s": 1e3 / ms, "iterations": it,
                     "ms_per_iteration": (ms / (it + 1)), "force": [float(x) for x in f],
                     "residual_force": rr, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / ms / 1e6,
                     "frac_hbm": nbytes / ms / 1e6 / PEAK_HBM_GBS}
        log("penalised 256^3 %s: %.3f ms, %d iterations, force %s" % (name, ms, it, f))
        if name == "fp64":
            outs.update(vel=vel, vort=vort)
        else:
            res[name]["rel_l2_vs_fp64"] = {"vel": rel_l2_t(vel, outs["vel"]), "vort": rel_l2_t(vort, outs["vort"])}
        del s
        torch.cuda.empty_cache()
    return res


def diagnostics_benchmark(reps, n=128):
    """FP64 global diagnostics: the ring diagnostics of the 128^3 two-ring
    particle set (2.1M particles; three sweeps reading 72 + 48 + 48 bytes
    per particle) and the mesh energy integrals of a 256^3 FP32 mesh pair
    (24 bytes per cell), timed with HIP events."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    from tests.helpers import ring_particles
    dev = torch.device("cuda:0")
    pos, vort = ring_particles(n)
    t = [torch.from_numpy(np.ascontiguousarray(a.T)).to(dev) for a in (pos, vort)]
    vel = torch.randn_like(t[0])
    d_out = torch.zeros(16, dtype=torch.float64, device=dev)
    h = 2.0 / n
    run = lambda: pm.diagnostics_particles_device((h,) * 3, t[0].data_ptr(), t[1].data_ptr(), vel.data_ptr(),
                                                  pos.shape[1], d_out=d_out.data_ptr(), host=False)
    res = {"workload": "ring diagnostics of 2097152 particles (128^3 two-ring set); mesh energy of a 256^3 "
                       "FP32 mesh pair"}
    for _ in range(3):
        run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nb = 168 * pos.shape[1]
    res["particles"] = {"ms": ms, "algorithmic_bytes": nb, "achieved_GBs": nb / ms / 1e6,
                        "frac_hbm": nb / ms / 1e6 / PEAK_HBM_GBS}
    m = 256
    mp = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(m,) * 3, dx=(2.0 / m,) * 3)
    u = torch.randn(3 * m ** 3, device=dev)
    wv = torch.randn(3 * m ** 3, device=dev)
    run = lambda: pm.diagnostics_mesh_device(mp, pm.FP32, u.data_ptr(), wv.data_ptr(), d_out=d_out.data_ptr(),
                                             host=False)
    for _ in range(3):
        run()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nb = 24 * m ** 3
    res["mesh"] = {"ms": ms, "algorithmic_bytes": nb, "achieved_GBs": nb / ms / 1e6,
                   "frac_hbm": nb / ms / 1e6 / PEAK_HBM_GBS}
    log("diagnostics: particles %.3f ms, mesh %.3f ms" % (res["particles"]["ms"], ms))
    return res


Z_KERNEL = "k_z3_r<float, 512, 4, false, true>"


def multires_benchmark(reps, n=256):
    """Multiresolution coupling on a 256^3 free refinement patch (level 2, two
    ghost layers, M'4) and its 128^3 parent: patch -> parent deposit (clear)
    and parent -> patch interpolation (added onto the patch field), FP64 and
    FP32, HIP events over `reps` launches on the null stream.  Algorithmic
    bytes: deposit 3 E_f^3 s read + 3 E_c^3 s written; interpolation 3 E_c^3 s
    read + 3 E_f^3 s read and written."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3, nghost=2, level=2,
                     parent=(1, 0))
    parent = pm.pmlib_topology_parent(patch, interp_order=3)
    ef, ec = pm.mesh_extent(patch, 3), pm.mesh_extent(parent, 3)
    nf, nc = 3 * int(np.prod(ef)), 3 * int(np.prod(ec))
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(11)
    res = {"workload": "256^3 free refinement patch (2 ghost layers) and its 128^3 parent, M'4, "
                       "seeded normal fields"}
    for name, prec, dt, sz in (("fp64", pm.FP64, torch.float64, 8), ("fp32", pm.FP32, torch.float32, 4)):
        fine = torch.randn(nf, dtype=dt, device=dev, generator=g)
        coarse = torch.randn(nc, dtype=dt, device=dev, generator=g)
        legs = {"deposit": (lambda: pm.interp_mesh_mesh_device(patch, parent, prec, fine.data_ptr(),
                                                               coarse.data_ptr(), True, interp_order=3),
                            (nf + nc) * sz),
                "interpolate": (lambda: pm.interp_mesh_mesh_device(parent, patch, prec, coarse.data_ptr(),
                                                                   fine.data_ptr(), False, interp_order=3),
                                (nc + 2 * nf) * sz)}
        res[name] = {}
        for leg, (run, nbytes) in legs.items():
            for _ in range(2):
                run()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[name][leg] = {"ms": ms, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / ms / 1e6,
                              "frac_hbm": nbytes / ms / 1e6 / PEAK_HBM_GBS}
        del fine, coarse
    log("multires 256^3 <-> 128^3: deposit fp64 %.3f / fp32 %.3f ms, interpolation fp64 %.3f / fp32 %.3f ms" %
        (res["fp64"]["deposit"]["ms"], res["fp32"]["deposit"]["ms"], res["fp64"]["interpolate"]["ms"],
         res["fp32"]["interpolate"]["ms"]))
    return res


def measured_traffic():
    """HBM bytes per launch of the z pass from the committed PMC summary
    (profiles/z_pass_traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, each scaled by the copy-calibration factor), or None when that
    summary was taken on another kernel than the one this build launches."""
    path = os.path.join(ROOT, "profiles", "z_pass_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("kernel") != Z_KERNEL:
        return None
    return (d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"]) * 1024.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--ring-steps", type=int, default=50, help="timed 64^3 ring solves (0 skips)")
    ap.add_argument("--step-steps", type=int, default=5,
                    help="timed two-ring vortex steps at 128^3 on 2.1M particles (0 skips)")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed 24-patch batched single-component solves (0 skips)")
    ap.add_argument("--mixed-steps", type=int, default=10, help="timed 384^2x384 mixed-BC solves (0 skips)")
    ap.add_argument("--mixed-vortex-steps", type=int, default=100,
                    help="remeshed FP32 vortex steps on the mixed-BC grid")
    ap.add_argument("--rhs-steps", type=int, default=20,
                    help="timed vorticity right-hand-side launches at 256^3 (0 skips)")
    ap.add_argument("--penal-steps", type=int, default=3,
                    help="timed penalised 256^3 solves (penalisation_solve; 0 skips)")
    ap.add_argument("--diag-steps", type=int, default=20, help="timed FP64 diagnostics (0 skips)")
    ap.add_argument("--full-steps", type=int, default=5,
                    help="timed full-output 256^3 solves (velocity, stream function, filtered vorticity; 0 skips)")
    ap.add_argument("--multires-steps", type=int, default=10,
                    help="timed patch <-> parent mesh interpolations at 256^3 / 128^3 (0 skips)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the single-GPU box runs --gpus 1 "
              "(multi-GPU pencil decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch
    import pmlib3d_amd as pm

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    h = 2.0 / N
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(N,) * 3, dx=(h,) * 3)
    log("building vorticity")
    w = blob_vorticity(N)
    s64 = pm.PoissonSolver(patch, pm.FP64, order=ORDER, kernel=1, eps_radius=EPS_RADIUS, interp_order=3)
    s64.setup()
    # spectral reprojection makes the input divergence-free (poisson_solve.f:861-872)
    d_w = torch.from_numpy(w.ravel(order="F")).to(dev)
    d_v = torch.empty_like(d_w)
    d_wp = torch.empty_like(d_w)
    s64.solve_device(d_w.data_ptr(), d_v.data_ptr(), d_wp.data_ptr(), None, reg_vort=False, reproj=True)
    torch.cuda.synchronize()
    del d_w, d_v
    w64 = d_wp
    w32 = w64.to(torch.float32)
    v32 = torch.empty_like(w32)
    v64 = torch.empty_like(w64)

    s32 = pm.PoissonSolver(patch, pm.FP32, order=ORDER, kernel=1, eps_radius=EPS_RADIUS, interp_order=3)
    s32.setup()
    torch.cuda.synchronize()

    def run(s, wi, vo, k):
        for _ in range(k):
            s.solve_device(wi.data_ptr(), vo.data_ptr())

    # FP32 timed region
    run(s32, w32, v32, args.warmup)
    torch.cuda.synchronize()
    s32.set_timing(True)
    t0 = time.perf_counter()
    run(s32, w32, v32, args.steps)
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    z32_ms, tot32_ms = s32.last_timing()
    s32.set_timing(False)

    # FP64 beside it
    run(s64, w64, v64, 2)
    torch.cuda.synchronize()
    s64.set_timing(True)
    t0 = time.perf_counter()
    run(s64, w64, v64, args.fp64_steps)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    z64_ms, tot64_ms = s64.last_timing()
    s64.set_timing(False)

    # the full-output form (velocity + stream function + filtered vorticity),
    # z pass and whole solve by HIP events, FP32 and FP64
    full = {}
    if args.full_steps > 0:
        for name, s_, wi, vo in (("fp32", s32, w32, v32), ("fp64", s64, w64, v64)):
            ps_, wo_ = torch.empty_like(wi), torch.empty_like(wi)
            kw = dict(d_vort_out=wo_.data_ptr(), d_vecpot=ps_.data_ptr(), reg_vort=True, calc_vecpot=True)
            for _ in range(2):
                s_.solve_device(wi.data_ptr(), vo.data_ptr(), **kw)
            torch.cuda.synchronize()
            s_.set_timing(True)
            for _ in range(args.full_steps):
                s_.solve_device(wi.data_ptr(), vo.data_ptr(), **kw)
            torch.cuda.synchronize()
            zf, tf = s_.last_timing()
            s_.set_timing(False)
            full[name] = {"zpass_ms": zf, "solve_ms": tf, "zpass_over_velocity_zpass": zf / (z32_ms if name == "fp32" else z64_ms)}
            del ps_, wo_
        log("full-output solve (u, psi, filtered w): fp32 %.3f ms (z %.3f), fp64 %.3f ms (z %.3f)" %
            (full["fp32"]["solve_ms"], full["fp32"]["zpass_ms"], full["fp64"]["solve_ms"], full["fp64"]["zpass_ms"]))

    if args.copy_calibration:
        # an elementwise kernel of known size: reads 1 GiB, writes 1 GiB
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 3 x torch.neg over 1 GiB (1 GiB read + 1 GiB written)")

    err = rel_l2_t(v32, v64)
    sps32 = args.steps / t32
    sps64 = args.fp64_steps / t64
    bytes32, z_bytes32 = free_bytes(N, 3, 3, 4, 4)
    bytes64, z_bytes64 = free_bytes(N, 3, 3, 8, 8)
    z_gbs = z_bytes32 / (z32_ms * 1e-3) / 1e9
    log("fp32: %.1f solves/s (%.3f ms, events %.3f ms/solve, z pass %.3f ms)" %
        (sps32, 1e3 / sps32, tot32_ms, z32_ms))
    log("fp64: %.1f solves/s (%.3f ms, events %.3f ms/solve, z pass %.3f ms)" %
        (sps64, 1e3 / sps64, tot64_ms, z64_ms))
    log("fp32 vs fp64 relative L2 velocity error %.3e" % err)
    w_cpu = w64.cpu().numpy().reshape((3, N, N, N), order="F")
    del s32, s64, w32, v32, v64, w64
    torch.cuda.empty_cache()

    ring = ring_w = None
    if args.ring_steps > 0:
        ring, ring_w = ring64_benchmark(args.ring_steps)
    step = step_benchmark(args.step_steps) if args.step_steps > 0 else None
    batch = batch_benchmark(args.batch_steps, 2) if args.batch_steps > 0 else None
    mixed = None
    if args.mixed_steps > 0:
        mixed = mixed384_benchmark(args.mixed_steps, args.mixed_vortex_steps)
        torch.cuda.empty_cache()
    rhs = None
    if args.rhs_steps > 0:
        rhs = rhs_benchmark(args.rhs_steps)
        log("vorticity rhs 256^3: fp64 %.3f ms (%.0f GB/s), fp32 %.3f ms (%.0f GB/s)" %
            (rhs["fp64"]["ms"], rhs["fp64"]["achieved_GBs"], rhs["fp32"]["ms"], rhs["fp32"]["achieved_GBs"]))

    penal = penal256_benchmark(args.penal_steps) if args.penal_steps > 0 else None
    diag = diagnostics_benchmark(args.diag_steps) if args.diag_steps > 0 else None
    multires = multires_benchmark(args.multires_steps) if args.multires_steps > 0 else None

    cpu = cpu_ring = cpu_step = None
    if not args.no_cpu_baseline:
        cpu = cpu_reference_solve(w_cpu, N, ORDER)
        if ring is not None:
            cpu_ring = cpu_reference_solve(ring_w, 64, 4, calc_vecpot=True)
            ring["cpu_baseline"] = cpu_ring
        if step is not None:
            cpu_step = cpu_reference_step()
            step["cpu_baseline"] = cpu_step

    out = {
        "metric": "unbounded-Poisson solves/s + Mparticle-steps/s at 256³, FP32 vs FP64 GPU",
        "value": sps32,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "256^3 free-space Poisson solve, 3 vorticity components, "
                               "doubled FFT grid 512^3 (implicit padding), order-4 regularised "
                               "G-hat (FP64-built), 32 seeded Gaussian blobs reprojected",
                   "grid": [N, N, N], "fft_grid": [2 * N] * 3, "components": 3, "order": ORDER,
                   "eps": "2h"},
        "fp64_solves_per_s": sps64,
        "fp64_ms_per_step": 1e3 / sps64,
        "fp32_over_fp64": sps32 / sps64,
        "fp32_rel_l2_vs_fp64": err,
        "solve_algorithmic_bytes": {"fp32": bytes32, "fp64": bytes64},
        "solve_achieved_GBs": {"fp32": bytes32 * sps32 / 1e9, "fp64": bytes64 * sps64 / 1e9},
        "solve_frac_hbm": {"fp32": bytes32 * sps32 / 1e9 / PEAK_HBM_GBS,
                           "fp64": bytes64 * sps64 / 1e9 / PEAK_HBM_GBS},
        "ring64": ring,
        "vortex_step": step,
        "batched_multipatch": batch,
        "mixed384": mixed,
        "vorticity_rhs": rhs,
        "penalised256": penal,
        "diagnostics": diag,
        "multires": multires,
        "full_output_solve256": full or None,
        "roofline": {"bound": "hbm", "kernel": Z_KERNEL + " (fused z pass)",
                     "achieved": z_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": z_gbs / PEAK_HBM_GBS, "traffic": measured_traffic(),
                     "algorithmic_bytes_per_launch": z_bytes32, "avg_launch_ms": z32_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
