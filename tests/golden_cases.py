"""Loader of the reference fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "cases.json")) as fh:
        return json.load(fh)


def cases():
    return manifest()["cases"]


def field(name):
    return np.load(os.path.join(GOLDEN, name + ".npy"), allow_pickle=False)


def args(c):
    """Positional arguments of oracle_ffi.simulate / specsim3d_amd.simulate for a case."""
    return (c["structures"], c["nugget"], *c["spacing"], *c["grid"], c["ipad"], c["iseed"], c["nreal"], c["igauss"])


def stdout_stats(c):
    """(mean, variance) per realization as printed by the reference (f9.4)."""
    out = []
    for ln in c["stdout"]:
        if ln.strip().startswith("realization"):
            parts = ln.replace("=", " ").split()
            out.append((float(parts[3]), float(parts[5])))
    return out


VARIO = os.path.join(GOLDEN, "vario")


def vario_cases():
    with open(os.path.join(VARIO, "vario.json")) as fh:
        return json.load(fh)["cases"]


def vario_field(name):
    return np.load(os.path.join(VARIO, name + ".npy"), allow_pickle=False)


def vario_table(name):
    """The reference's <case>_vario.csv: (header, rows as lists of strings)."""
    with open(os.path.join(VARIO, name + "_vario.csv")) as fh:
        lines = fh.read().splitlines()
    return lines[0].split(","), [ln.split(",") for ln in lines[1:]]


COND = os.path.join(GOLDEN, "cond")


def cond_manifest():
    with open(os.path.join(COND, "cond.json")) as fh:
        return json.load(fh)


COND_KEYS = ["ixc", "iyc", "izc", "ivc", "iwc", "tmin", "tmax", "itrans", "skmean", "zlo", "zhi", "ltail", "ltpar",
             "utail", "utpar", "rmax", "rmed", "rvert", "sa1", "sa2", "sa3", "ndmax"]


def parse_cond_block(lines):
    """The conditioning block: data file, then 9 lines of parameters (src/specsim3d.f90:731-747)."""
    vals = [v for ln in lines[1:10] for v in ln.split()]
    out = {}
    for k, v in zip(COND_KEYS, vals):
        out[k] = float(v) if k in ("tmin", "tmax", "skmean", "zlo", "zhi", "ltpar", "utpar", "rmax", "rmed",
                                   "rvert", "sa1", "sa2", "sa3") else int(v)
    return lines[0].split()[0], out


def cond_case(name):
    """A conditioning case of tests/golden/cond: model, grid, conditioning parameters and data rows."""
    with open(os.path.join(COND, name + ".par")) as fh:
        lines = [ln for ln in fh.read().splitlines()]
    origin = tuple(map(float, lines[0].split()))
    spacing = tuple(map(float, lines[1].split()))
    grid = tuple(map(int, lines[2].split()))
    nugget, nst = float(lines[3].split()[0]), int(lines[3].split()[1])
    sts = [tuple([int(ln.split()[0])] + list(map(float, ln.split()[1:8]))) for ln in lines[4:4 + nst]]
    seed = lines[4 + nst].split()
    iseed, nreal, ipad = int(seed[0]), int(seed[1]), int(seed[2])
    igauss = int(seed[3]) if len(seed) > 3 else 0
    base = lines[5 + nst].strip()
    rest = lines[6 + nst:]
    if rest:  # new format: nout, codes, icond, conditioning block
        nout = int(rest[0])
        block = rest[2 + nout:2 + nout + 10]
    else:
        with open(os.path.join(COND, name + ".cpar")) as fh:
            block = fh.read().splitlines()
    datafile, params = parse_cond_block(block)
    with open(os.path.join(COND, datafile)) as fh:
        dl = fh.read().splitlines()
    nvar = int(dl[1])
    rows = np.array([list(map(float, ln.split())) for ln in dl[2 + nvar:] if ln.strip()])
    return dict(origin=origin, spacing=spacing, grid=grid, nugget=nugget, sts=sts, iseed=iseed, nreal=nreal,
                ipad=ipad, igauss=igauss, base=base, params=params, rows=rows, datafile=datafile)


def cond_field(name):
    return np.load(os.path.join(COND, name + ".npy"), allow_pickle=False)
