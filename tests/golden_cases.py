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
