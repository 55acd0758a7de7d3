"""Time-frequency envelope and phase misfit and goodness-of-fit criteria.

A re-implementation of the TF_MISFIT_GOF_CRITERIA tool the reference uses to
compare seismogram-like traces (src/TF_MISFIT_GOF_master: input namelist
HF_TF-MISFIT_GOF, signal files "TIME S(1..NC)" per line, output
MISFIT-GOF.DAT as described in its user's guide, pp. 3-8).  The shipped
program is a closed binary; this module follows the published method:

  Kristekova, Kristek, Moczo, Day (2006), Misfit criteria for quantitative
  comparison of seismograms, BSSA 96(5), 1836-1850;
  Kristekova, Kristek, Moczo (2009), Time-frequency misfit and goodness-of-fit
  criteria for quantitative comparison of time signals, GJI 178, 813-825.

Time-frequency representation: continuous wavelet transform with the Morlet
wavelet psi(t) = pi^(-1/4) exp(i w0 t) exp(-t^2/2), w0 = 6, at scale
a = w0 / (2 pi f) for NF_TF logarithmically equidistant frequencies in
[FMIN, FMAX]:  W(t, f) = a^(-1/2) int s(tau) psi*((tau - t)/a) dtau
(evaluated with the FFT).  With W_REF the reference signal's W (S2 when
IS_S2_REFERENCE, else the larger of the two moduli, see `_reference`):

  dE(t,f) = |W1| - |W_REF|,   dP(t,f) = |W_REF| * wrap(arg W1 - arg W_REF) / pi
  global:  TFEM = dE / max|W_REF|,  TFPM = dP / max|W_REF|
           TEM(t) = sum_f dE / max_t sum_f |W_REF|  (FEM(f) alike over t)
  local:   TFEM = dE / |W_REF|, TEM(t) = sum_f dE / sum_f |W_REF|, ...
  EM = sqrt(sum dE^2 / sum |W_REF|^2),  PM = sqrt(sum dP^2 / sum |W_REF|^2)
  goodness of fit (A = 10, k = 1): EG = A exp(-|EM|),  PG = A (1 - |PM|),
  and likewise pointwise for TFEG, TFPG, TEG, TPG, FEG, FPG.

Parity: pinned by the closed forms the papers give (identical signals give
EM = PM = 0 and EG = PG = 10; S1 = a S2 gives PM = 0 and EM = |a - 1|; a
delay gives a phase misfit ~ 2 f delay), tests/test_tfmisfit.py.  The
program's frequency count NF_TF is not documented; it is a parameter here.
"""
import os
import re
import sys

import numpy as np

W0 = 6.0
A_GOF = 10.0
K_GOF = 1.0


def frequencies(fmin, fmax, nf):
    return np.exp(np.linspace(np.log(fmin), np.log(fmax), nf)) if nf > 1 else np.array([fmin])


def cwt(s, dt, freqs, w0=W0):
    """Morlet CWT W[f, t] of a real signal sampled at dt (zero-padded FFT)."""
    s = np.asarray(s, dtype=np.float64)
    mt = s.size
    nfft = 1 << int(np.ceil(np.log2(2 * mt)))
    S = np.fft.fft(s, nfft)
    w = 2 * np.pi * np.fft.fftfreq(nfft, dt)
    out = np.empty((len(freqs), mt), dtype=np.complex128)
    for q, f in enumerate(freqs):
        a = w0 / (2 * np.pi * f)
        # Fourier transform of psi*((tau)/a) / sqrt(a) as a correlation kernel:
        # Psi_hat(a w) = pi^(-1/4) sqrt(2 pi) exp(-(a w - w0)^2 / 2), times sqrt(a)
        psi = np.pi ** -0.25 * np.sqrt(2 * np.pi) * np.sqrt(a) * np.exp(-0.5 * (a * w - w0) ** 2)
        psi[w < 0] = 0.0
        out[q] = np.fft.ifft(S * psi)[:mt]
    return out


def _reference(W1, W2, s2_reference):
    if s2_reference:
        return W2
    # no reference: the larger modulus of the two at every (t, f), W2's phase
    m = np.where(np.abs(W1) > np.abs(W2), np.abs(W1), np.abs(W2))
    return m * np.exp(1j * np.angle(W2))


def misfits(s1, s2, dt, fmin, fmax, nf=100, s2_reference=True, local_norm=False):
    """All criteria of one component; s2 is the reference when s2_reference."""
    freqs = frequencies(fmin, fmax, nf)
    W1 = cwt(s1, dt, freqs)
    W2 = cwt(s2, dt, freqs)
    Wr = _reference(W1, W2, s2_reference)
    ar = np.abs(Wr)
    dE = np.abs(W1) - ar
    dphi = np.angle(W1) - np.angle(Wr)
    dphi = (dphi + np.pi) % (2 * np.pi) - np.pi
    dP = ar * dphi / np.pi
    tiny = np.finfo(np.float64).tiny
    if local_norm:
        tfem = dE / np.maximum(ar, tiny)
        tfpm = dphi / np.pi
        tem = dE.sum(axis=0) / np.maximum(ar.sum(axis=0), tiny)
        tpm = dP.sum(axis=0) / np.maximum(ar.sum(axis=0), tiny)
        fem = dE.sum(axis=1) / np.maximum(ar.sum(axis=1), tiny)
        fpm = dP.sum(axis=1) / np.maximum(ar.sum(axis=1), tiny)
    else:
        mx = max(ar.max(), tiny)
        tfem = dE / mx
        tfpm = dP / mx
        tem = dE.sum(axis=0) / max(ar.sum(axis=0).max(), tiny)
        tpm = dP.sum(axis=0) / max(ar.sum(axis=0).max(), tiny)
        fem = dE.sum(axis=1) / max(ar.sum(axis=1).max(), tiny)
        fpm = dP.sum(axis=1) / max(ar.sum(axis=1).max(), tiny)
    den = max((ar ** 2).sum(), tiny)
    em = float(np.sqrt((dE ** 2).sum() / den))
    pm = float(np.sqrt((dP ** 2).sum() / den))
    eg = lambda x: A_GOF * np.exp(-np.abs(x) ** K_GOF)
    pg = lambda x: A_GOF * (1.0 - np.abs(x) ** K_GOF)
    return dict(freqs=freqs, EM=em, PM=pm, EG=float(eg(em)), PG=float(pg(pm)), TFEM=tfem, TFPM=tfpm, TEM=tem,
                TPM=tpm, FEM=fem, FPM=fpm, TFEG=eg(tfem), TFPG=pg(tfpm), TEG=eg(tem), TPG=pg(tpm), FEG=eg(fem),
                FPG=pg(fpm), TFRS1=np.abs(W1), TFRS2=np.abs(W2))


def compare(s1, s2, dt, fmin, fmax, nf=100, s2_reference=True, local_norm=False):
    """Per-component criteria of signals shaped [MT] or [MT][NC]."""
    a = np.asarray(s1, dtype=np.float64)
    b = np.asarray(s2, dtype=np.float64)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    return [misfits(a[:, j], b[:, j], dt, fmin, fmax, nf, s2_reference, local_norm) for j in range(a.shape[1])]


# ------------------------------------------------------------- tool files
def read_namelist(path):
    """HF_TF-MISFIT_GOF: &INPUT ... / with MT, DT, FMIN, FMAX, S1_NAME,
    S2_NAME, NC, IS_S2_REFERENCE, LOCAL_NORM."""
    text = open(path).read()
    body = text[text.index("&INPUT") + 6:text.rindex("/")]
    out = dict(NC=1, IS_S2_REFERENCE=False, LOCAL_NORM=False)
    for m in re.finditer(r"(\w+)\s*=\s*('[^']*'|[^,\n]+)", body):
        k, v = m.group(1).upper(), m.group(2).strip()
        if v.startswith("'"):
            out[k] = v.strip("'")
        elif v.upper().strip(".") in ("T", "TRUE"):
            out[k] = True
        elif v.upper().strip(".") in ("F", "FALSE"):
            out[k] = False
        elif k in ("MT", "NC"):
            out[k] = int(v)
        else:
            out[k] = float(v.replace("d", "e").replace("D", "e"))
    return out


def read_signal(path, mt, nc):
    data = np.loadtxt(path, ndmin=2)[:mt]
    return data[:, 0], data[:, 1:1 + nc]


def write_outputs(res, par, s1, s2, directory="."):
    """MISFIT-GOF.DAT in the layout of the user's guide (p. 5) plus the
    per-component TF / time / frequency files (pp. 5-8)."""
    nc = len(res)
    nf = len(res[0]["freqs"])
    mt = s1.shape[0]
    fmt = lambda *v: " ".join("%.7E" % x if isinstance(x, float) else str(x) for x in v) + "\n"
    with open(os.path.join(directory, "MISFIT-GOF.DAT"), "w") as f:
        f.write(fmt(float(par["FMIN"]), float(par["FMAX"])))
        f.write(fmt(nf, mt))
        f.write(fmt(float(par["DT"]), nc))
        f.write(fmt(float(max(np.abs(s1).max(), np.abs(s2).max()))))
        for r in res:
            f.write(fmt(r["EM"], r["PM"]))
        for r in res:
            f.write(fmt(r["EG"], r["PG"]))
        mx = lambda key: float(max(np.abs(r[key]).max() for r in res))
        f.write(fmt(mx("TFEM"), mx("TFPM")))
        f.write(fmt(mx("FEM"), mx("FPM")))
        f.write(fmt(mx("TEM"), mx("TPM")))
        f.write(fmt(mx("TFRS1"), mx("TFRS2")))
    for x, r in enumerate(res, start=1):
        for key in ("TFEM", "TFEG", "TFPM", "TFPG"):
            np.savetxt(os.path.join(directory, "%s%d.DAT" % (key, x)), r[key], fmt="%.7E")
        for key in ("TEM", "TEG", "TPM", "TPG", "FEM", "FEG", "FPM", "FPG"):
            np.savetxt(os.path.join(directory, "%s%d.DAT" % (key, x)), r[key], fmt="%.7E")
        np.savetxt(os.path.join(directory, "TFRS1_%d.DAT" % x), r["TFRS1"], fmt="%.7E")
        np.savetxt(os.path.join(directory, "TFRS2_%d.DAT" % x), r["TFRS2"], fmt="%.7E")


def run_tool(directory=".", nf=100):
    """The tool's run: HF_TF-MISFIT_GOF and the two signal files in
    `directory` -> MISFIT-GOF.DAT and the criteria files there."""
    par = read_namelist(os.path.join(directory, "HF_TF-MISFIT_GOF"))
    mt, nc = int(par["MT"]), int(par.get("NC", 1))
    _, s1 = read_signal(os.path.join(directory, par["S1_NAME"]), mt, nc)
    _, s2 = read_signal(os.path.join(directory, par["S2_NAME"]), mt, nc)
    res = compare(s1, s2, float(par["DT"]), float(par["FMIN"]), float(par["FMAX"]), nf,
                  bool(par.get("IS_S2_REFERENCE", False)), bool(par.get("LOCAL_NORM", False)))
    write_outputs(res, par, s1, s2, directory)
    return res


if __name__ == "__main__":
    run_tool(sys.argv[1] if len(sys.argv) > 1 else ".")
