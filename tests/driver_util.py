"""Helpers to drive the Fortran executable the way the reference's Python scripts do
(tests/test_cases.py:103-111, tests/test_gauss_amplitudes.py:27-33)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "specsim3d_amd", "bin", "specsim3d")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "specsim3d")


def par_text(origin, spacing, grid, nugget, structures, iseed, nreal, ipad, igauss=None, base="t", codes=None,
             header=True, icond=0):
    lines = []
    if header:
        lines += ["Parameters for SPECSIM3D", "************************", "", "START OF PARAMETERS:"]
    lines += ["%r %r %r" % tuple(float(v) for v in origin), "%r %r %r" % tuple(float(v) for v in spacing),
              "%d %d %d" % tuple(grid), "%r %d" % (float(nugget), len(structures))]
    for s in structures:
        lines.append("%d %r %r %r %r %r %r %r" % (s[0], *map(float, s[1:])))
    lines.append("%d %d %d" % (iseed, nreal, ipad) + ("" if igauss is None else " %d" % igauss))
    lines.append(base)
    if codes is not None:
        lines += [str(len(codes))] + [str(c) for c in codes] + [str(icond)]
    return "\n".join(lines) + "\n"


def run(exe, wd, par_name, *args, timeout=300):
    return subprocess.run([exe, par_name, *args], cwd=wd, capture_output=True, text=True, timeout=timeout)


def strip_times(text):
    return [ln for ln in text.splitlines() if "time" not in ln and ln.strip()]
