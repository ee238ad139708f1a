"""Writes tests/golden/ref_fullsize_checksums.json from the compiled
reference (oracle/_ref/ref_driver, built from the reference sources by
oracle/Makefile.ref).  Run in the build container only:

    python tests/golden/make_fullsize.py

The full-size outputs are too large to commit (a 256^3 velocity is 400 MB),
so the fixture holds size-independent fingerprints of them
(tests/helpers.plane_checksums / particle_checksums):
  solve_256   the 256^3 free-space solve of the benchmark (order 4, eps 2h,
              32 seeded Gaussian blobs, bench.blob_vorticity), velocity
  step_128    one two-ring vortex step at 128^3 with 128^3 jittered
              particles (order 6, M'4, RK2, dt = h/2, remesh min_vort 1e-6)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.helpers import particle_checksums, plane_checksums, ring_particles  # noqa: E402

STEP_MIN_VORT = 1e-6


def solve_256():
    from bench import blob_vorticity
    n = 256
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, 2.0 / n, (0, 0, 0))
    p = O.make_params(4, 1, 2.0)
    w = blob_vorticity(n)
    t = time.time()
    raw, out = O.ref_run(dict(O.geom_job(g, p), mode="solve"), [w], timeout=3000)
    print("reference 256^3 solve: %.1f s; %s" % (time.time() - t, out.strip().splitlines()[-1]), flush=True)
    vel = np.frombuffer(raw, dtype=np.float64)[:w.size].reshape(w.shape, order="F")
    return dict(plane_checksums(vel), input="bench.blob_vorticity(256), seed 3", order=4, eps_radius=2.0)


def step_128():
    n = 128
    h = 2.0 / n
    g = O.make_geom((n,) * 3, 0, [-1.0] * 3, h, (0, 0, 0), 3)
    p = O.make_params(6, 1, 2.0)
    pos, vort = ring_particles(n)
    t = time.time()
    raw, out = O.ref_run(dict(O.geom_job(g, p), mode="step", npart=pos.shape[1], rk_order=2, dtime=0.5 * h,
                              nsteps=1, min_vort=STEP_MIN_VORT, min_dvort=0.0), [pos, vort], timeout=3000)
    print("reference 128^3 step: %.1f s" % (time.time() - t), flush=True)
    npart = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    arr = np.frombuffer(raw[4:]).reshape((3, 3 * npart), order="F")
    d = particle_checksums(arr[:, :npart], arr[:, npart:2 * npart], arr[:, 2 * npart:])
    d.update(np_in=int(pos.shape[1]), min_vort=STEP_MIN_VORT, order=6, rk_order=2, dtime=0.5 * h)
    return d


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: make -f oracle/Makefile.ref")
    res = {"source": "compiled reference (oracle/_ref, FP64, native Glassman FFT, 1 rank)",
           "solve_256": solve_256(), "step_128": step_128()}
    json.dump(res, open(os.path.join(HERE, "ref_fullsize_checksums.json"), "w"))
