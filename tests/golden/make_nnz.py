"""Writes tests/golden/nnz/nnz.json: the accepted (ray, voxel) pair counts of
the reference discretisation, counted by the CPU oracle over every ray and
every voxel of the plain workloads:

* forward: segments of every detector ray that CudaPerformProjection accepts
  (index inside the volume and weight > 1e-6, Projector.cu:356-366);
* backprojector: (pixel, voxel) pairs that CudaPerformBackProjection
  accumulates (window pixel whose box intersection passes the tests of
  Projector.cu:710-729).

These are the numerators of the roofline's algorithmic bytes.  The full-size
count walks 104.9 M rays and 67.1 M voxels on the host (several minutes).

    python tests/golden/make_nnz.py [small|full ...]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kevernals_amd import geometry as G  # noqa: E402
from oracle.oracle import Oracle, threads  # noqa: E402

WORKLOADS = {
    "small": ("128x128x32, 256^2 detector, 15 views on +-15 deg", G.small_plain),
    "full": ("1024x1024x64, 2048^2 detector, 25 views on +-20 deg", G.full_plain),
}
PATH = os.path.join(HERE, "nnz", "nnz.json")


def main(names):
    out = json.load(open(PATH)) if os.path.exists(PATH) else {}
    for name in names:
        desc, make = WORKLOADS[name]
        o = Oracle(make(), threads())
        t = time.time()
        fp = o.fp_nnz()
        t_fp = time.time() - t
        t = time.time()
        bp = o.bp_nnz()
        t_bp = time.time() - t
        print(f"{name}: forward {fp} ({t_fp:.1f} s), backprojector {bp} ({t_bp:.1f} s)", flush=True)
        out[name] = {"workload": desc, "mode": "R", "forward": fp, "backprojector": bp}
    os.makedirs(os.path.dirname(PATH), exist_ok=True)
    with open(PATH, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
