"""Write the FDTD golden vectors from the compiled reference.

Runs oracle/_ref/ref_fdtd (src/fdtd/fdtd_simulation.cpp compiled with g++
and a driver of our own, see oracle/Makefile) on small cases and stores the
full-precision traces and fields as .npz fixtures next to this script:

    make -C oracle && python tests/golden/gen_fdtd_golden.py

Only cubic or Nx == Ny grids are used: the reference writes Hy with a
k-stride that is correct only for Nx == Ny (src/fdtd/fdtd_simulation.cpp:144).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref  # noqa: E402

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5


def cfl_dt(h, frac):
    return frac * h / (C0 * 3 ** 0.5)


def main_like_case(n):
    """src/fdtd/main.cpp:12-65 at n^3 cells instead of 500^3."""
    h = 1.0 / n
    num_steps = int(round(780 * n / 500))
    dt = 3.0e-9 / num_steps
    sp = [0.49, 0.49, 0.49]
    op = [0.51, 0.51, 0.51]

    def r(v):  # std::round (half away from zero)
        return int(np.floor(v / h + 0.5))

    src = [(0, (r(sp[0]) - 1, r(sp[1]), r(sp[2])))]
    obs = [[
        [(r(op[0]) - 1, r(op[1]), r(op[2]))],
        [(r(op[0]) - 1, r(op[1]) - 1, r(op[2])), (r(op[0]) - 1, r(op[1]), r(op[2])),
         (r(op[0]), r(op[1]) - 1, r(op[2])), (r(op[0]), r(op[1]), r(op[2]))],
        [(r(op[0]) - 1, r(op[1]), r(op[2]) - 1), (r(op[0]) - 1, r(op[1]), r(op[2])),
         (r(op[0]), r(op[1]), r(op[2]) - 1), (r(op[0]), r(op[1]), r(op[2]))],
    ]]
    return dict(grid=(n, n, n), h=h, dt=dt, eps=EPS0, mu=MU0, tf=1, nsteps=num_steps,
                sources=src, observations=obs)


CASES = {
    # 16^3, every component sourced, two observation sets
    "fdtd_small": dict(grid=(16, 16, 16), h=1e-3, dt=cfl_dt(1e-3, 0.99), eps=EPS0, mu=MU0, tf=1,
                       nsteps=120,
                       sources=[(0, (5, 8, 8)), (1, (8, 5, 8)), (2, (8, 8, 5))],
                       observations=[[[(8, 8, 8)], [(8, 8, 8)], [(8, 8, 8)]],
                                     [[(2, 3, 4), (12, 11, 10)], [(3, 2, 13)], [(14, 14, 1), (1, 1, 14)]]],
                       full_fields=True),
    # Nx = Ny != Nz, tf = 3, a repeated source cell, ghost-cell observation
    "fdtd_rect": dict(grid=(20, 20, 28), h=2e-3, dt=cfl_dt(2e-3, 0.95), eps=EPS0 * 1.7, mu=MU0, tf=3,
                      nsteps=150,
                      sources=[(0, (5, 10, 14)), (0, (5, 10, 14)), (1, (9, 4, 20)), (2, (12, 12, 7))],
                      observations=[[[(10, 10, 14)], [(9, 10, 14), (10, 10, 14)],
                                     [(10, 10, 13), (10, 10, 14), (11, 11, 14)]],
                                    [[(3, 0, 3)], [(17, 17, 25)], [(5, 6, 7)]]],
                      full_fields=True),
    # the reference main.cpp setup at 50^3
    "fdtd_main50": dict(main_like_case(50), full_fields=False),
    # vacuum/PEC version of the 64^3 1000-step workload (receivers on y=z=32)
    "fdtd_vacuum64": dict(grid=(64, 64, 64), h=1e-3, dt=cfl_dt(1e-3, 0.99), eps=EPS0, mu=MU0, tf=1,
                          nsteps=1000, sources=[(2, (16, 32, 32))],
                          observations=[[[(10 + int(round((r + 0.5) * 44 / 8)), 32, 32)],
                                         [(10 + int(round((r + 0.5) * 44 / 8)), 32, 32)],
                                         [(10 + int(round((r + 0.5) * 44 / 8)), 32, 32)]]
                                        for r in range(8)],
                          full_fields=False),
}


def field_summary(fields):
    """Per-field sum, sum of squares and max |.|, plus the mid-z plane of each field."""
    out = {}
    for name, f in zip(("ex", "ey", "ez", "hx", "hy", "hz"), fields):
        out[name + "_stats"] = np.array([f.sum(), (f * f).sum(), np.abs(f).max()])
    return out


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/ref_fdtd missing: run make -C oracle")
    for name, c in CASES.items():
        tr, fl, secs = ref.run_reference(c["grid"], c["h"], c["dt"], c["eps"], c["mu"], c["tf"],
                                         c["nsteps"], c["sources"], c["observations"])
        arrays = {"traces": tr}
        arrays.update(field_summary(fl))
        if c["full_fields"]:
            for nm, f in zip(("ex", "ey", "ez", "hx", "hy", "hz"), fl):
                arrays[nm] = f
        meta = {k: v for k, v in c.items() if k != "full_fields"}
        arrays["config"] = np.array(json.dumps(meta))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print("%-16s %s  traces %s  ref run %.3fs  -> %s (%d KB)" % (
            name, c["grid"], tr.shape, secs, os.path.basename(path), os.path.getsize(path) // 1024))


if __name__ == "__main__":
    main()
