"""Synthetic FDTD workloads (seeded, numpy) for benches and tests.

The models follow the benchmark definitions: eps_r from 3 to 6 random
ellipsoids (eps_r in [2, 8], semi-axes U[0.1, 0.3]*N) in vacuum, a 10-cell
CPML, a Ricker source with f0 in [5, 15] GHz at a cell uniform in the inner
50 % of the domain, and 32 receivers on a planar 8 x 4 grid.  Values live on
the (nx+1)(ny+1)(nz+1) material grid; value (i,j,k) is sampled at the cell
centre ((i-1/2)h, (j-1/2)h, (k-1/2)h).
"""
import numpy as np

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5


def cfl_dt(h, frac=0.99):
    return frac * h / (C0 * 3 ** 0.5)


def ellipsoid_model(n, seed):
    """eps_r on the material grid of an n^3 model."""
    rng = np.random.default_rng(seed)
    count = int(rng.integers(3, 7))
    er = np.ones((n + 1, n + 1, n + 1), dtype=np.float64)
    c1 = np.arange(n + 1, dtype=np.float64) - 0.5
    for _ in range(count):
        centre = rng.uniform(0.25 * n, 0.75 * n, 3)
        axes = rng.uniform(0.1, 0.3, 3) * n
        val = rng.uniform(2.0, 8.0)
        # bounding box to keep the evaluation cheap
        lo = np.maximum(np.floor(centre - axes).astype(int), 0)
        hi = np.minimum(np.ceil(centre + axes).astype(int) + 2, n + 1)
        z = ((c1[lo[2]:hi[2]] - centre[2]) / axes[2]) ** 2
        y = ((c1[lo[1]:hi[1]] - centre[1]) / axes[1]) ** 2
        x = ((c1[lo[0]:hi[0]] - centre[0]) / axes[0]) ** 2
        inside = (z[:, None, None] + y[None, :, None] + x[None, None, :]) <= 1.0
        sub = er[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[inside] = val
    return er.reshape(-1)


def sphere_model(n, radius=16.0, eps_r=4.0, centre=None):
    """eps_r on the material grid: a sphere of the given radius (cells) and
    eps_r in vacuum, centred at (n/2, n/2, n/2) unless given; sampled at the
    cell centres like the other models."""
    c = np.full(3, n / 2.0) if centre is None else np.asarray(centre, dtype=np.float64)
    c1 = np.arange(n + 1, dtype=np.float64) - 0.5
    d2 = ((c1 - c[2]) ** 2)[:, None, None] + ((c1 - c[1]) ** 2)[None, :, None] + ((c1 - c[0]) ** 2)[None, None, :]
    er = np.where(d2 <= radius * radius, eps_r, 1.0)
    return er.reshape(-1)


def blob_model(n, seed, background=2.0, count=3, d_eps=0.5, width=12.0, d_sigma=0.01):
    """(eps_r, sigma) on the material grid: background eps_r plus `count`
    seeded Gaussian blobs of amplitude d_eps (eps_r) and d_sigma (S/m) and
    standard deviation `width` cells."""
    rng = np.random.default_rng(seed)
    c1 = np.arange(n + 1, dtype=np.float64) - 0.5
    er = np.full((n + 1, n + 1, n + 1), background)
    sg = np.zeros((n + 1, n + 1, n + 1))
    for _ in range(count):
        c = rng.uniform(0.3 * n, 0.7 * n, 3)
        gz = np.exp(-0.5 * ((c1 - c[2]) / width) ** 2)
        gy = np.exp(-0.5 * ((c1 - c[1]) / width) ** 2)
        gx = np.exp(-0.5 * ((c1 - c[0]) / width) ** 2)
        blob = gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        er += d_eps * blob
        sg += d_sigma * blob
    return er.reshape(-1), sg.reshape(-1)


def shot(n, seed):
    """Source cell (Ez), Ricker f0 and 32 receivers (8 x 4 plane) of one shot."""
    rng = np.random.default_rng(seed + 1000003)
    src = tuple(int(v) for v in rng.integers(n // 4 + 1, 3 * n // 4, 3))
    f0 = float(rng.uniform(5e9, 15e9))
    zr = n - 20
    xs = np.linspace(n // 5, 4 * n // 5, 8).astype(int)
    ys = np.linspace(n // 3, 2 * n // 3, 4).astype(int)
    recs = [(int(x), int(y), int(zr)) for y in ys for x in xs]
    return src, f0, recs


def fdtd_cell_bytes(n, npml, word, material_words):
    """Algorithmic HBM bytes of one full Yee step on an n^3 grid:
    per cell w*(12 + m), plus w*8 per CPML axis the cell lies in
    (two E-psi and two H-psi, each read and written once)."""
    cells = n ** 3
    base = cells * word * (12 + material_words)
    if npml <= 0:
        return base, base / cells
    inpml = np.zeros(n, dtype=np.int64)
    idx = np.arange(n)
    inpml[(idx < npml) | (idx >= n - npml)] = 1
    d = inpml[:, None, None] + inpml[None, :, None] + inpml[None, None, :]
    extra = int(d.sum()) * word * 8
    return base + extra, (base + extra) / cells


def tet_model(n, seed, sigma_fraction=0.0, sigma_max=0.05):
    """(eps_r, sigma) per tetrahedron of an n^3-cube Kuhn mesh: eps_r per cube
    log-uniform in [1, 9], smoothed by a Gaussian of 3 cubes (reflecting
    edges), inherited by the cube's 6 tetrahedra (tetrahedron id
    6 (i + n (j + n k)) + t); sigma uniform in [0, sigma_max] S/m on a seeded
    `sigma_fraction` of the tetrahedra (None when 0)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    er = np.exp(rng.uniform(0.0, np.log(9.0), (n, n, n)))  # [k][j][i]
    er = gaussian_filter(er, 3.0, mode="reflect")
    er = np.repeat(er.reshape(-1), 6)
    sg = None
    if sigma_fraction > 0:
        sg = np.zeros(er.size)
        pick = rng.random(er.size) < sigma_fraction
        sg[pick] = rng.uniform(0.0, sigma_max, int(pick.sum()))
    return er, sg


def tet_receivers(n, h, count, seed):
    """`count` seeded random interior points (inner 80 % of the cube)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.1, 0.9, (count, 3)) * n * h
