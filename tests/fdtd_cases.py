"""Shared FDTD case builders: the same configuration applied to the oracle and
to the GPU solver (both expose the reference's method names)."""
import numpy as np

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5


def cfl_dt(h, frac):
    return frac * h / (C0 * 3 ** 0.5)


def apply_sources_observations(sim, cfg):
    setters = (sim.setSource_x, sim.setSource_y, sim.setSource_z)
    for comp, p in cfg["sources"]:
        setters[comp](tuple(p))
    for lists in cfg["observations"]:
        sim.setObservationPoint([[tuple(p) for p in l] for l in lists])


def material_grid(grid, kind, seed=0):
    """eps_r / sigma on the (nx+1)(ny+1)(nz+1) material grid; value (i,j,k)
    sits at the cell centre ((i-1/2)h, (j-1/2)h, (k-1/2)h)."""
    nx, ny, nz = grid
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    x, y, z = i - 0.5, j - 0.5, k - 0.5
    rng = np.random.default_rng(seed)
    if kind == "sphere":
        cx, cy, cz = nx / 2, ny / 2, nz / 2
        r = min(grid) / 4
        er = np.where((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r, 4.0, 1.0)
        return er.ravel(), None
    if kind == "random":
        er = 1.0 + 3.0 * rng.random(k.shape)
        sg = 0.02 * rng.random(k.shape)
        return er.ravel(), sg.ravel()
    if kind == "lossless_random":
        return (1.0 + 3.0 * rng.random(k.shape)).ravel(), None
    raise ValueError(kind)


# GPU-vs-oracle parity cases (small enough for the oracle to finish in seconds)
PARITY_CASES = {
    "pec_cubic": dict(grid=(16, 16, 16), h=1e-3, frac=0.99, tf=1, nsteps=80,
                      sources=[(0, (5, 8, 8)), (1, (8, 5, 8)), (2, (8, 8, 5))],
                      observations=[[[(8, 8, 8)], [(8, 8, 8)], [(8, 8, 8)]]]),
    "pec_rect_nonsquare": dict(grid=(23, 17, 30), h=1e-3, frac=0.9, tf=2, nsteps=60,
                               sources=[(2, (11, 8, 15)), (0, (4, 4, 4))],
                               observations=[[[(11, 8, 16)], [(12, 8, 15), (11, 9, 15)], [(11, 8, 15)]],
                                             [[(1, 1, 1)], [(22, 16, 29)], []]]),
    "sphere_cpml_ricker": dict(grid=(32, 32, 32), h=1e-3, frac=0.99, tf=1, nsteps=100,
                               sources=[(2, (10, 16, 16))],
                               observations=[[[(22, 16, 16)], [(22, 16, 16)], [(22, 16, 16)]]],
                               material="sphere", npml=6, wavelet=(1, 10e9)),
    "lossy_cpml_rect": dict(grid=(24, 20, 28), h=2e-3, frac=0.95, tf=1, nsteps=80,
                            sources=[(0, (12, 10, 14)), (1, (9, 9, 9))],
                            observations=[[[(13, 11, 15)], [(13, 11, 15)], [(13, 11, 15)]]],
                            material="random", npml=5, wavelet=(2, 20e9), pml_kappa=3.0,
                            pml_alpha=0.02),
}


def build(cls, case, precision=None):
    grid = case["grid"]
    h = case["h"]
    dt = cfl_dt(h, case["frac"])
    kw = {} if precision is None else {"precision": precision}
    sim = cls(grid, h, dt, EPS0, MU0, case["tf"], **kw)
    return sim


def configure(sim, case, oracle_side):
    """Apply material, CPML and wavelet.  The oracle takes CPML/wavelet as
    constructor arguments, the GPU solver through setters."""
    apply_sources_observations(sim, case)
    if case.get("material"):
        er, sg = material_grid(case["grid"], case["material"], seed=7)
        sim.set_material(er, sg)
    if not oracle_side:
        if case.get("npml"):
            sim.set_cpml(case["npml"], 3.0, 1e-6, case.get("pml_kappa", 1.0), case.get("pml_alpha", 0.0))
        if case.get("wavelet"):
            kind, f0 = case["wavelet"]
            sim.set_wavelet(kind, f0, 2 ** 0.5 / f0, 1.0)


def oracle_kwargs(case):
    kw = {}
    if case.get("npml"):
        kw.update(npml=case["npml"], pml_order=3.0, pml_r0=1e-6,
                  pml_kappa_max=case.get("pml_kappa", 1.0), pml_alpha_max=case.get("pml_alpha", 0.0))
    if case.get("wavelet"):
        kind, f0 = case["wavelet"]
        kw.update(wavelet=kind, f0=f0, t0=2 ** 0.5 / f0, amp=1.0)
    return kw


def run_oracle(case):
    from oracle.fdtd import OracleFDTD
    dt = cfl_dt(case["h"], case["frac"])
    o = OracleFDTD(case["grid"], case["h"], dt, EPS0, MU0, tf=case["tf"], **oracle_kwargs(case))
    configure(o, case, oracle_side=True)
    o.run(case["nsteps"])
    return o
