"""Runner for the compiled reference FDTD (oracle/_ref/ref_fdtd).

TEST INFRASTRUCTURE ONLY.  The binary is built by oracle/Makefile from the
reference sources (src/fdtd/fdtd_simulation.cpp) and a driver of our own
(oracle/ref_fdtd_driver.cpp).  It is used to write golden vectors and, on the
GPU host, as the reference CPU baseline.
"""
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_FDTD = os.path.join(_HERE, "_ref", "ref_fdtd")


def available():
    return os.path.exists(REF_FDTD)


def field_sizes(nx, ny, nz):
    return [
        (nx + 1) * (ny + 2) * (nz + 2),
        (nx + 2) * (ny + 1) * (nz + 2),
        (nx + 2) * (ny + 2) * (nz + 1),
        nx * (ny + 1) * (nz + 1),
        (nx + 1) * ny * (nz + 1),
        (nx + 1) * (ny + 1) * nz,
    ]


def _stdin(grid, h, dt, eps, mu, tf, nsteps, sources, observations):
    lines = ["%d %d %d" % tuple(grid), "%.17g %.17g %.17g %.17g %d %d" % (h, dt, eps, mu, tf, nsteps),
             str(len(sources))]
    for comp, p in sources:
        lines.append("%d %d %d %d" % (comp, p[0], p[1], p[2]))
    lines.append(str(len(observations)))
    for lists in observations:
        lines.append("%d %d %d" % tuple(len(l) for l in lists))
        for l in lists:
            for p in l:
                lines.append("%d %d %d" % tuple(p))
    return "\n".join(lines) + "\n"


def run_reference(grid, h, dt, eps, mu, tf, nsteps, sources, observations, want_output=True):
    """Run the reference class; returns (traces[nobs][nslot][3], fields[6], run_seconds)."""
    inp = _stdin(grid, h, dt, eps, mu, tf, nsteps, sources, observations)
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin") if want_output else "-"
        r = subprocess.run([REF_FDTD, out], input=inp.encode(), stdout=subprocess.PIPE,
                           stderr=subprocess.DEVNULL, check=True)
        secs = float(r.stdout.decode().split()[-1])
        if not want_output:
            return None, None, secs
        raw = open(out, "rb").read()
    nobs, nslot = np.frombuffer(raw[:8], dtype=np.int32)
    off = 8
    ntr = int(nobs) * int(nslot) * 3
    traces = np.frombuffer(raw[off:off + 8 * ntr], dtype=np.float64).reshape(int(nobs), int(nslot), 3)
    off += 8 * ntr
    fields = []
    for n in field_sizes(*grid):
        fields.append(np.frombuffer(raw[off:off + 8 * n], dtype=np.float64).copy())
        off += 8 * n
    return traces.copy(), fields, secs
