import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libems.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    import json
    out["config"] = json.loads(str(out["config"]))
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


@pytest.fixture(scope="session")
def golden():
    return load_golden


def tf_misfits(traces, ref, dt):
    """Largest single-valued TF envelope and phase misfits (Kristekova et al.)
    over the observation components of traces [nobs][nslot][3] against the
    reference traces; dt is the slot spacing.  The traces must be finite
    wherever the reference is (a NaN or Inf there fails the assertion);
    components the reference leaves non-finite (the 0/0 average of an empty
    receiver list) are skipped, as are components whose peak is below 1e-2
    of the largest reference sample: rounding errors scale with the whole
    field, so on a weak component (or the cancellation noise of one the source does not
    drive) a relative criterion measures the strong components' rounding."""
    from emsolvers_amd import tfmisfit
    traces = np.asarray(traces, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    # a blown-up trace must fail, never fold away as a skipped component: the
    # only non-finite samples allowed are those of the reference itself (its
    # 0/0 average of an empty receiver list), at the same places
    bad = ~np.isfinite(ref)
    assert np.array_equal(~np.isfinite(traces), bad), "tf_misfits: non-finite samples where the reference is finite"
    nslot = ref.shape[1]
    fmin, fmax = 2.0 / (nslot * dt), 0.25 / dt
    finite = ref[~bad]
    floor = 1e-2 * np.abs(finite).max() if finite.size else 0.0
    em = pm = 0.0
    for o in range(ref.shape[0]):
        for c in range(ref.shape[2]):
            r, s = ref[o, :, c], traces[o, :, c]
            if bad[o, :, c].any() or not np.abs(r).max() > floor:
                continue
            m = tfmisfit.misfits(s, r, dt, fmin, fmax, nf=40)
            assert np.isfinite(m["EM"]) and np.isfinite(m["PM"]), \
                "tf_misfits: non-finite misfit at observation %d component %d" % (o, c)
            em, pm = float(np.fmax(em, m["EM"])), float(np.fmax(pm, m["PM"]))
    return em, pm
