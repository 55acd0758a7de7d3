"""Build libems.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m emsolvers_amd.build          # or emsolvers_amd.build.build()

Every HIP source under emsolvers_amd/csrc is compiled by hipcc for gfx950
into emsolvers_amd/libems.so; the reference-compatible command-line drivers
(emsolvers_amd/drivers) are compiled by the host C++ compiler against it into
emsolvers_amd/bin.  Packed FP32 instructions are disabled
(-packed-fp32-ops) and floating-point contraction is off, so that FP64
kernels keep the reference's operation order; FP32 kernels request fused
multiply-adds explicitly where they want them.
"""
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libems.so")
BIN = os.path.join(HERE, "bin")
DRIVERS = os.path.join(HERE, "drivers")
# program name -> (source, extra defines); src/<solver>/main.cpp of the reference
PROGRAMS = {
    "ems_fdtd": ("fdtd_main.cpp", []),
    "ems_fem": ("fem_main.cpp", []),
    "ems_fem_abm": ("fem_abm_main.cpp", []),
    "ems_fem_v2": ("hex_main.cpp", ["-DEMS_HEX_STEPPER=0"]),
    "ems_symplectic_2": ("hex_main.cpp", ["-DEMS_HEX_STEPPER=1"]),
    "ems_symplectic_4": ("hex_main.cpp", ["-DEMS_HEX_STEPPER=2"]),
}
ARCH = os.environ.get("EMS_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build libems.so")


def flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-I", CSRC,
        "-Wall", "-Wno-unused-function",
    ]


def build(verbose=False, jobs=8):
    """Compile every csrc/*.hip to an object and link libems.so."""
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [cc] + flags() + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join("== %s ==\n%s" % f for f in failed)
        raise RuntimeError("libems build failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    os.replace(tmp, LIB)
    build_drivers(verbose)
    return LIB


def host_cxx():
    for cand in (os.environ.get("CXX"), shutil.which("g++"), shutil.which("c++")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("no host C++ compiler for the drivers")


def build_drivers(verbose=False):
    """Compile the command-line drivers against libems.so (runpath
    $ORIGIN/.., so the tree can move as a whole)."""
    os.makedirs(BIN, exist_ok=True)
    cxx = host_cxx()
    procs = []
    for prog, (src, defs) in PROGRAMS.items():
        out = os.path.join(BIN, prog)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", DRIVERS] + defs + [
            os.path.join(DRIVERS, src), "-o", out + ".tmp", "-L", HERE, "-lems",
            "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((prog, out, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for prog, out, p in procs:
        text = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed.append((prog, text))
        else:
            os.replace(out + ".tmp", out)
    if failed:
        raise RuntimeError("driver build failed:\n" + "\n".join("== %s ==\n%s" % f for f in failed))
    return BIN


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
