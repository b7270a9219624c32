"""Writes tests/golden/inputs.sha256: SHA-256 of the synthetic benchmark
inputs, so the generators (coupe_amd/workloads.c, which uses the host libm's
log / cos / exp) are pinned.  Run from the repository root:

    python tests/golden/make_input_hashes.py [--with-jgrid512]

Entries: the 2^20-point 2D workload, the 2^24-point Gaussian workload, the
batch header of the 4096-instance workload and its first 16 instances, and
(with --with-jgrid512, 4.3 GB) the 512^3 jittered grid.
"""
import hashlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from coupe_amd import workloads  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "inputs.sha256")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(a.tobytes())
    return h.hexdigest()


def entries(with_jgrid=False):
    out = {}
    p, w, _, _ = workloads.standard("uniform2d")
    out["uniform2d"] = sha(p, w)
    p, w, _, _ = workloads.standard("gauss3d")
    out["gauss3d"] = sha(p, w)
    n, dim, lo, ext, iters = workloads.batch_header(4096, seed=3)
    out["batch4096.header"] = sha(n, dim, lo, ext, iters)
    parts = []
    for q in range(16):
        d = int(dim[q])
        parts += list(workloads.box(int(n[q]), d, lo[q, :d], ext[q, :d], workloads.instance_seed(3, q), 16))
    out["batch4096.first16"] = sha(*parts)
    if with_jgrid:
        p, w, _, _ = workloads.standard("jgrid512")
        out["jgrid512"] = sha(p, w)
    return out


if __name__ == "__main__":
    e = entries("--with-jgrid512" in sys.argv)
    old = {}
    if os.path.exists(OUT):
        for line in open(OUT):
            if line.strip():
                h, k = line.split()
                old[k] = h
    old.update(e)
    with open(OUT, "w") as f:
        for k in sorted(old):
            f.write(f"{old[k]}  {k}\n")
    print(open(OUT).read())
