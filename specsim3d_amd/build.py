"""Build recipe of the engine (run in-tree; the products ship with the tree).

  specsim3d_amd/lib/libspecsim3d.so   HIP kernels for gfx950 + the C ABI (include/specsim3d.h)
  specsim3d_amd/bin/specsim3d         the Fortran driver (amdflang, ISO_C_BINDING), linked to the
                                      library with an $ORIGIN-relative rpath

Usage: python -m specsim3d_amd.build   (or __graft_entry__.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FC = os.path.join(ROCM, "bin", "amdflang")
ARCH = "gfx950"

# -ffp-contract=off: FP64 kernels and host model arithmetic round every product and sum
# separately, as the reference build does (no FMA on its x86-64 baseline target).
# packed-fp32-ops disabled: packed FP32 arithmetic is not used anywhere in this engine.
COMMON = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]
UNITS = {
    "ss3d_k64.hip": ["-ffp-contract=off"],
    "ss3d_k32.hip": [],
    "ss3d_k32_r3.hip": [],
    "ss3d_k32_r4.hip": [],
    "ss3d_k32_r5.hip": [],
    "ss3d_k32_c2r.hip": [],
    "ss3d_host.hip": ["-ffp-contract=off"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = []
    for src, extra in UNITS.items():
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        jobs.append((obj, [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]))
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for out in ex.map(lambda j: _run(j[1]), jobs):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libspecsim3d.so")
    tmp = so + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *[j[0] for j in jobs]])
    os.replace(tmp, so)
    return so


def build_driver(verbose=False):
    srcs = [os.path.join(FSRC, f) for f in ("ss3d_capi.f90", "specsim3d.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(BIN, exist_ok=True)
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    exe = os.path.join(BIN, "specsim3d")
    tmp = exe + ".tmp"
    _run([FC, "-O2", "-module-dir", mod, *srcs, "-L", LIB, "-lspecsim3d", "-Wl,-rpath,$ORIGIN/../lib", "-o", tmp])
    os.replace(tmp, exe)
    return exe


def build(verbose=False):
    so = build_library(verbose)
    exe = build_driver(verbose)
    return so, exe


if __name__ == "__main__":
    so, exe = build(verbose="-v" in sys.argv)
    print(so)
    print(exe)
