"""Run the reference Fortran hot path (oracle/_ref/<kernel>/ref_harness).

Test infrastructure only: used to write golden vectors under tests/golden/ and
as the "reference" CPU baseline.  The binary is built from the reference
sources by oracle/ref.mk; see oracle/ref_harness.f90 for the file format.
"""
import os
import re
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KERNEL_NAMES = {0: "m4", 1: "m6", 2: "gauss"}


def harness_path(kernel):
    return os.path.join(HERE, "_ref", KERNEL_NAMES[kernel], "ref_harness")


def available(kernel=1):
    return os.path.exists(harness_path(kernel))


def write_input(path, state, store, mode=0, sts_s=0, dtcap=0.0):
    """Write a problem in the ref_harness.f90 input format."""
    store = np.ascontiguousarray(store, dtype=np.float64)
    n_real = int(state[24])
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", mode, n_real, store.shape[0] - n_real, sts_s))
        f.write(struct.pack("<d", dtcap))
        f.write(np.asarray(state, dtype="<f8").tobytes())
        f.write(store.astype("<f8").tobytes())


def run(state, store, kernel=1, mode=0, sts_s=0, dtcap=0.0, threads=None):
    """Run the reference on ``(state, store)``.

    Returns dict(store=(n,32) array after the call, dt=float, neigh=list of
    (rj, code) int32 arrays per real particle (rj 0-based real index, code =
    (sx+1)+3(sy+1)+9(sz+1)), seconds=float wall time of the timed region)."""
    store = np.ascontiguousarray(store, dtype=np.float64)
    n = store.shape[0]
    n_real = int(state[24])
    n_fixed = n - n_real
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", mode, n_real, n_fixed, sts_s))
            f.write(struct.pack("<d", dtcap))
            f.write(np.asarray(state, dtype="<f8").tobytes())
            f.write(store.astype("<f8").tobytes())
        p = subprocess.run([harness_path(kernel), fin, fout], env=env,
                           capture_output=True, text=True)
        if p.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError("reference harness failed (rc=%d):\n%s\n%s"
                               % (p.returncode, p.stdout[-2000:], p.stderr[-2000:]))
        m = re.search(r"seconds\s+([0-9.Ee+-]+)", p.stderr)
        secs = float(m.group(1)) if m else float("nan")
        raw = open(fout, "rb").read()
    off = 0
    (ntot,) = struct.unpack_from("<i", raw, off)
    off += 4
    out = np.frombuffer(raw, dtype="<f8", count=32 * ntot, offset=off).reshape(ntot, 32).copy()
    off += 8 * 32 * ntot
    (dt,) = struct.unpack_from("<d", raw, off)
    off += 8
    (nr,) = struct.unpack_from("<i", raw, off)
    off += 4
    if nr == -1:
        (npairs,) = struct.unpack_from("<q", raw, off)
        return dict(store=out, dt=dt, neigh=None, n_pairs=npairs, seconds=secs, stdout=p.stdout)
    neigh = []
    for _ in range(nr):
        (cnt,) = struct.unpack_from("<i", raw, off)
        off += 4
        a = np.frombuffer(raw, dtype="<i4", count=2 * cnt, offset=off).reshape(cnt, 2).copy()
        off += 8 * cnt
        a[:, 0] -= 1
        neigh.append(a)
    return dict(store=out, dt=dt, neigh=neigh, n_pairs=sum(len(a) for a in neigh), seconds=secs,
                stdout=p.stdout)
