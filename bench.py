"""Benchmark of the unbounded Poisson solve on one MI355X.

Workload (one "step" = one full velocity solve): 256^3 cells on [-1,1]^3
(h = 1/128), free-space boundaries, doubled FFT grid 512^3 handled with
implicit zero padding, three vorticity components in one solve, regularised
Green's function of order 4 with eps = 2h built in FP64.  Vorticity: 32
seeded Gaussian blobs (centres U[-0.6,0.6]^3, sigma U[2h,6h], strengths
N(0,1)^3, seed 3), projected divergence-free with the solver's spectral
reprojection.  Inputs and outputs stay resident in HBM; the timed region
contains only solves.

Prints one JSON line: FP32 solves/s (value), the FP64 rate beside it, the
roofline of the dominant kernel (the fused z pass) from HIP events on the
solver's stream, and the compiled reference CPU solve rate (cpu_baseline).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 256
ORDER = 4
EPS_RADIUS = 2.0
PEAK_HBM_GBS = 8000.0


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


def algorithmic_bytes(n, s, s_g):
    """Bytes one solve must move with the 5-pass implicit-padding schedule,
    3 components in and out, G-hat stored as its real octant."""
    M = n
    L = 2 * n
    Hx = L // 2 + 1
    Hy = Hz = L // 2 + 1
    c = 2 * s
    x_f = 3 * M ** 3 * s + 3 * Hx * M * M * c
    y_f = 3 * Hx * M * M * c + 3 * Hx * L * M * c
    z_p = 3 * Hx * L * M * c + Hx * Hy * Hz * s_g + 3 * Hx * L * M * c
    y_i = 3 * Hx * L * M * c + 3 * Hx * M * M * c
    x_i = 3 * Hx * M * M * c + 3 * M ** 3 * s
    return x_f + y_f + z_p + y_i + x_i, z_p


def cpu_reference_baseline(w64):
    """The compiled reference (oracle/_ref, FP64, native Glassman FFT) on one
    full 256^3 solve under MPI; falls back to the C oracle on 128^3."""
    from oracle import oracle as O
    g = O.make_geom((N,) * 3, 0, [-1.0] * 3, 2.0 / N, (0, 0, 0))
    p = O.make_params(ORDER, 1, EPS_RADIUS)
    ranks = max(1, min(16, os.cpu_count() or 1))
    if O.ref_available() and os.path.exists(O.MPIEXEC):
        try:
            t = time.time()
            _, out = O.ref_run(dict(O.geom_job(g, p), mode="solve", nrep=1), [w64], nproc=ranks,
                               timeout=240)
            line = [l for l in out.splitlines() if l.startswith("TIMING")][0]
            solve_s = float(line.split("solve_min_s=")[1].split()[0])
            setup_s = float(line.split("setup_s=")[1].split()[0])
            log("cpu reference: %s (wall %.1fs)" % (line, time.time() - t))
            return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": ranks, "kind": "reference",
                    "sample": "one 256^3 free-space solve (3 components, order 4, FP64, native "
                              "Glassman FFT; FFTW absent) with mpiexec -n %d; setup %.2fs not "
                              "counted" % (ranks, setup_s)}
        except Exception as e:  # pragma: no cover - reported, not hidden
            log("cpu reference failed (%s); timing the C oracle instead" % e)
    import numpy as np
    n2 = 128
    g2 = O.make_geom((n2,) * 3, 0, [-1.0] * 3, 2.0 / n2, (0, 0, 0))
    w2 = np.asfortranarray(w64[:, ::2, ::2, ::2])
    Gh = O.poisson_setup(g2, p)
    t = time.time()
    O.poisson_solve(g2, p, Gh, w2)
    dt = time.time() - t
    scale = (2 * N) ** 3 * math.log(2 * N) / ((2 * n2) ** 3 * math.log(2 * n2))
    return {"value": 1.0 / (dt * scale), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "one 128^3 solve of the C oracle (1 thread), scaled by L^3 log L to 256^3"}


STEP_MIN_VORT = 1e-6   # remesh threshold of the vortex step (peak |omega| is ~41)


def step_benchmark(steps, warmup, n=128, precision=0):
    """The two-ring vortex step (config of 128^3 cells, 128^3 particles on the
    jittered cell-centre lattice, order-6 solve, M'4, RK2, dt = h/2, remesh
    with min_vort = 1e-6) on the GPU, every array resident in HBM; precision
    0 is FP64 throughout, 1 the FP32 mesh side (particles FP64).
    Returns (Mparticle-steps/s, ms per step, particle count)."""
    import torch
    import pmlib3d_amd as pm
    from tests.helpers import ring_particles
    h = 2.0 / n
    patch = pm.Patch(xmin=(-1.0,) * 3, xmax=(1.0,) * 3, ncell=(n,) * 3, dx=(h,) * 3)
    pos, vort = ring_particles(n)
    st = pm.VortexStep(patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3,
                       capacity=max(pos.shape[1], n ** 3), precision=precision)
    st.load(pos, vort)
    for _ in range(warmup):
        st.step(0.5 * h, STEP_MIN_VORT)
    torch.cuda.synchronize()
    counted = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        counted += st.np
        st.step(0.5 * h, STEP_MIN_VORT)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return counted / dt / 1e6, 1e3 * dt / steps, st.np


def patch_bytes(n, s, s_g, comps=1):
    """Algorithmic bytes of one free-space solve of a patch n = (nx, ny, nz)
    with the 5-pass implicit-padding schedule (comps components in and out,
    G-hat as its real octant)."""
    Mx, My, Mz = n
    Lx, Ly, Lz = 2 * Mx, 2 * My, 2 * Mz
    Hx, Hy, Hz = Lx // 2 + 1, Ly // 2 + 1, Lz // 2 + 1
    c = 2 * s
    b = comps * (Mx * My * Mz * s + Hx * My * Mz * c)           # x forward
    b += comps * (Hx * My * Mz * c + Hx * Ly * Mz * c)          # y forward
    b += comps * 2 * Hx * Ly * Mz * c + Hx * Hy * Hz * s_g      # fused z
    b += comps * (Hx * Ly * Mz * c + Hx * My * Mz * c)          # y inverse
    b += comps * (Hx * My * Mz * c + Mx * My * Mz * s)          # x inverse
    return b


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
    (seeded), h = 1/128, single-component FP32 solves of order 4 (eps = 2h),
    all patches replayed as one captured HIP graph.  Returns batches/s, the
    algorithmic bytes per batch and the patch sizes."""
    import numpy as np
    import torch
    import pmlib3d_amd as pm
    rng = np.random.default_rng(seed)
    h = 1.0 / 128
    sizes = [tuple(int(v) for v in rng.choice([32, 48, 64, 96, 128, 160], 3)) for _ in range(npatch)]
    dev = torch.device("cuda:0")
    solvers, ins, outs = [], [], []
    for n in sizes:
        patch = pm.Patch(xmin=tuple(-v * h / 2 for v in n), xmax=tuple(v * h / 2 for v in n), ncell=n,
                         dx=(h,) * 3)
        sv = pm.PoissonSolver(patch, pm.FP32, order=4, kernel=1, eps_radius=2.0, interp_order=3)
        sv.setup()
        w = torch.from_numpy(ring_patch(n, h, rng).ravel(order="F")).to(dev).float()
        solvers.append(sv)
        ins.append(w)
        outs.append(torch.zeros_like(w))
    batch = pm.PoissonBatch(solvers)
    pi, po = [t.data_ptr() for t in ins], [t.data_ptr() for t in outs]
    for _ in range(warmup):
        batch.solve_scalar(pi, po)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.solve_scalar(pi, po)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    nbytes = sum(patch_bytes(n, 4, 4) for n in sizes)
    return 1.0 / dt, nbytes, sizes


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
    for name, prec, dt, s in (("fp64", pm.FP64, torch.float64, 8), ("fp32", pm.FP32, torch.float32, 4)):
        w = torch.randn(int(np.prod(ext)), dtype=dt, device=dev, generator=g)
        u = torch.randn_like(w)
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
    return res


def cpu_step_baseline(n=128):
    """One two-ring vortex step of the compiled reference (one rank: its step
    mode is single-rank), FP64."""
    from oracle import oracle as O
    from tests.helpers import ring_particles
    if not (O.ref_available() and os.path.exists(O.MPIEXEC)):
        return None
    h = 2.0 / n
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, h, (0, 0, 0), 3)
    p = O.make_params(6, 1, 2.0)
    pos, vort = ring_particles(n)
    try:
        _, out = O.ref_run(dict(O.geom_job(g, p), mode="step", npart=pos.shape[1], rk_order=2,
                                dtime=0.5 * h, nsteps=1, min_vort=STEP_MIN_VORT, min_dvort=0.0),
                           [pos, vort], timeout=400)
    except Exception as e:  # pragma: no cover - reported, not hidden
        log("cpu reference step failed (%s)" % e)
        return None
    line = [l for l in out.splitlines() if l.startswith("TIMING")][0]
    step_s = float(line.split("step_s=")[1].split()[0])
    log("cpu reference step: %s" % line)
    return {"value": pos.shape[1] / step_s / 1e6, "unit": "Mparticle-steps/s", "cores": 1,
            "kind": "reference",
            "sample": "one two-ring vortex step at 128^3 (2.1M particles, order-6 solve, M'4, RK2, "
                      "remesh), FP64, native Glassman FFT, 1 rank"}


Z_KERNEL = "k_z3_r<float, 512, 4, false, true>"


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
    ap.add_argument("--step-steps", type=int, default=5,
                    help="timed two-ring vortex steps at 128^3 (0 skips)")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed 24-patch batched single-component solves (0 skips)")
    ap.add_argument("--rhs-steps", type=int, default=20,
                    help="timed vorticity right-hand-side launches at 256^3 (0 skips)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the single-GPU box runs --gpus 1 "
              "(multi-GPU pencil decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
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

    if args.copy_calibration:
        # an elementwise kernel of known size: reads 1 GiB, writes 1 GiB
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log("copy calibration: 3 x torch.neg over 1 GiB (1 GiB read + 1 GiB written)")

    a = v32.double()
    b = v64
    err = float(torch.sqrt(torch.sum((a - b) ** 2) / torch.sum(b * b)))

    sps32 = args.steps / t32
    sps64 = args.fp64_steps / t64
    bytes32, z_bytes32 = algorithmic_bytes(N, 4, 4)
    bytes64, z_bytes64 = algorithmic_bytes(N, 8, 8)
    z_gbs = z_bytes32 / (z32_ms * 1e-3) / 1e9
    log("fp32: %.1f solves/s (%.3f ms, events %.3f ms/solve, z pass %.3f ms)" %
        (sps32, 1e3 / sps32, tot32_ms, z32_ms))
    log("fp64: %.1f solves/s (%.3f ms, events %.3f ms/solve, z pass %.3f ms)" %
        (sps64, 1e3 / sps64, tot64_ms, z64_ms))
    log("fp32 vs fp64 relative L2 velocity error %.3e" % err)

    step = None
    if args.step_steps > 0:
        del s32, s64
        step = {"workload": "two-ring vortex step, 128^3 free, 128^3 jittered particles, order 6, "
                            "eps 2h, M'4, RK2 dt=h/2, remesh min_vort=1e-6; FP32 = FP32 meshes, solve and "
                            "Kahan-summed interpolation, FP64 particles"}
        for name, prec in (("fp64", 0), ("fp32", 1)):
            mps, ms, npart = step_benchmark(args.step_steps, 1, precision=prec)
            log("vortex step %s: %.1f Mparticle-steps/s (%.3f ms/step, %d particles)" % (name, mps, ms, npart))
            step[name] = {"Mparticle_steps_per_s": mps, "ms_per_step": ms, "particles": npart}

    batch = None
    if args.batch_steps > 0:
        bps, nbytes, sizes = batch_benchmark(args.batch_steps, 2)
        log("batched multipatch: %.1f batches/s (%.3f ms/batch, %.2f GB algorithmic, %.0f GB/s)" %
            (bps, 1e3 / bps, nbytes / 1e9, nbytes * bps / 1e9))
        batch = {"batches_per_s": bps, "ms_per_batch": 1e3 / bps, "algorithmic_bytes": nbytes,
                 "achieved_GBs": nbytes * bps / 1e9, "frac_hbm": nbytes * bps / 1e9 / PEAK_HBM_GBS,
                 "dtype": "f32", "patches": sizes,
                 "workload": "24 free-space patches, sizes per axis from {32,48,64,96,128,160} (seed 4), "
                             "h=1/128, one Gaussian ring each (R=0.3 min extent, sigma=0.15R, random axis), "
                             "single-component psi solve, order 4, eps 2h, one captured HIP graph per batch"}

    rhs = None
    if args.rhs_steps > 0:
        rhs = rhs_benchmark(args.rhs_steps)
        log("vorticity rhs 256^3: fp64 %.3f ms (%.0f GB/s), fp32 %.3f ms (%.0f GB/s)" %
            (rhs["fp64"]["ms"], rhs["fp64"]["achieved_GBs"], rhs["fp32"]["ms"], rhs["fp32"]["achieved_GBs"]))

    cpu = None
    cpu_step = None
    if not args.no_cpu_baseline:
        cpu = cpu_reference_baseline(w64.cpu().numpy().reshape((3, N, N, N), order="F"))
        if step is not None:
            cpu_step = cpu_step_baseline()

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
        "solve_achieved_GBs": {"fp32": bytes32 / (1e-3 / sps32 * 1e3) / 1e9 * 1e3 / 1e3,
                               "fp64": bytes64 * sps64 / 1e9},
        "vortex_step": step,
        "batched_multipatch": batch,
        "vorticity_rhs": rhs,
        "cpu_baseline_step": cpu_step,
        "roofline": {"bound": "hbm", "kernel": Z_KERNEL + " (fused z pass)",
                     "achieved": z_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": z_gbs / PEAK_HBM_GBS, "traffic": measured_traffic(),
                     "algorithmic_bytes_per_launch": z_bytes32, "avg_launch_ms": z32_ms},
        "cpu_baseline": cpu,
    }
    out["solve_achieved_GBs"]["fp32"] = bytes32 * sps32 / 1e9
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
