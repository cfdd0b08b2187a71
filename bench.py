#!/usr/bin/env python3
"""Throughput of SPECSIM3D's unconditional spectral simulation on one MI355X.

Workload (the configuration the headline metric is quoted on): a 256^3 output grid in a 512^3
periodic grid; exponential model exp(-3r), sill 0.7 plus nugget 0.3, ranges 80/40/10, azimuth
30, dip 15, ipad=1, igauss=1 (Gaussian spectral coefficients); 64 realizations per step.

A step = one pass of the per-realization hot path over one batch: the 64 realizations (fill,
inverse FFT passes, crop, statistics) written to HBM, inputs (the FP64 spectral density, built
once before timing) already resident.  Timed: the FP32 batched path (the `value`) and the FP64
drop-in path, each bracketed by device synchronisation.  Per-pass HIP events on the library's
stream give the dominant kernel's duration for the roofline.  The CPU baseline is the
reference program (oracle/_ref, built from the reference source) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

STRUCTS = [(2, 0.7, 80.0, 40.0, 10.0, 30.0, 15.0, 0.0)]
NUGGET = 0.3
GRID = (256, 256, 256)
ISEED = 69069
IGAUSS = 1
HBM_PEAK_GBS = 8000.0


class Hip:
    """Device allocations for the benchmark's output buffers (plumbing)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = ctypes.c_void_p()
        rc = self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({n}) failed: {rc}")
        return p.value

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))

    def copy_d2h(self, dst, src, n):
        rc = self.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(n), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy d2h failed: {rc}")

    def copy_d2d(self, dst, src, n):
        rc = self.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(n), 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy d2d failed: {rc}")
        self.L.hipDeviceSynchronize()


REGION = "timed FP32 region"


class Regions:
    """A ROCTx range named REGION around the timed FP32 region (rocprofiler-sdk ROCTx library), so
    that `rocprofv3 --kernel-trace --marker-trace` of this script marks which dispatches were timed
    (see --trace-summary); the region's CLOCK_MONOTONIC bounds, the clock of rocprofv3's
    timestamps, also go into the JSON line.  Without the profiler the calls do nothing."""

    def __init__(self):
        self.L = None
        self.bounds = [None, None]
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.L is not None:
            self.L.roctxRangePushA.argtypes = [ctypes.c_char_p]
            self.L.roctxRangePushA.restype = ctypes.c_int
            self.L.roctxRangePop.restype = ctypes.c_int

    def resume(self):
        self.bounds[0] = time.monotonic_ns()
        if self.L is not None:
            self.L.roctxRangePushA(REGION.encode())

    def pause(self):
        if self.L is not None:
            self.L.roctxRangePop()
        self.bounds[1] = time.monotonic_ns()


# pass kind of the library's profile -> kernel of that pass (rocprofv3 names)
KIND_KERNELS = {"real_zy_fp32": r"fill_load_kernel<", "real_z_fp32": r"fast::strided_kernel<\d+, \d+, 3, 2,",
                "real_y_fp32": r"fast::strided_kernel<\d+, \d+, 2, 1,", "real_x_c2r_fp32": r"fast::c2r_(loop_|zr_)?kernel<"}


def trace_summary(trace_csv, marker_csv, bench_json, command):
    """Markdown summary of the kernels of the timed FP32 region: the dispatches of a rocprofv3
    kernel trace that lie inside the ROCTx range REGION (or the JSON line's monotonic bounds), with
    the dominant kernel's average duration beside the HIP-event figure of the same run."""
    import csv

    with open(bench_json) as fh:
        line = json.loads([ln for ln in fh.read().splitlines() if ln.startswith("{")][-1])
    t0, t1 = line.get("timed_region_monotonic_ns", [None, None])
    if marker_csv and os.path.exists(marker_csv):
        with open(marker_csv) as fh:
            for r in csv.DictReader(fh):
                if r["Function"] == REGION:
                    t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
    with open(trace_csv) as fh:
        rows = [r for r in csv.DictReader(fh) if t0 <= int(r["Start_Timestamp"]) and int(r["End_Timestamp"]) <= t1]
    stats = {}
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
        n, tot = stats.get(r["Kernel_Name"], (0, 0.0))
        stats[r["Kernel_Name"]] = (n + 1, tot + d)
    total = sum(v[1] for v in stats.values()) or 1.0
    rf = line["roofline"]
    out = ["# rocprofv3 kernel trace -- the timed FP32 region of bench.py (256^3 in 512^3, 64 realizations per step)", "",
           "Workload: 256^3 output grid in a 512^3 periodic grid, exponential sill 0.7 + nugget 0.3, ranges 80/40/10, "
           "az 30, dip 15, ipad=1, igauss=1; FP32 output in HBM.  Dispatches: those inside the ROCTx range "
           f"'{REGION}' that bench.py pushes around its timed steps ({line['steps']} steps after "
           f"{line['warmup']} warm-up steps); nothing else of the run is counted.", "",
           "Commands (repository root on the GPU host, after `cd /tmp && export TMPDIR=/tmp`):", "",
           "    " + command, f"    python3 bench.py --trace-summary {trace_csv} --marker-csv {marker_csv} "
           f"--bench-json {bench_json}", "",
           f"bench.py in the same run: {line['value']} Gcells/s, {line['ms_per_step']} ms per step.", "",
           "| kernel | calls | total ms | average ms | % |", "|---|---|---|---|---|"]
    for k, (n, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{k}` | {n} | {tot:.2f} | {tot / n:.4f} | {100 * tot / total:.2f} |")
    pat = re.compile(KIND_KERNELS.get(rf["kernel"], r"^$"))
    dom = [(k, v) for k, v in stats.items() if pat.search(k)]
    if dom:
        k, (n, tot) = max(dom, key=lambda kv: kv[1][1])
        avg = tot / n
        ach = rf["algorithmic_bytes_per_launch"] / (avg * 1e-3) / 1e9
        out += ["", f"Dominant kernel ({rf['kernel']}): {n} dispatches, average {avg:.4f} ms in the trace against "
                    f"{rf['ms_per_launch']:.4f} ms from the library's HIP events (difference "
                    f"{100 * (avg / rf['ms_per_launch'] - 1):+.1f}%).  Algorithmic bytes per launch "
                    f"{rf['algorithmic_bytes_per_launch'] / 1e9:.3f} GB: {ach:.0f} GB/s = {ach / rf['peak']:.4f} of "
                    f"{rf['peak']:.0f} GB/s from the trace, {rf['frac']} in the JSON line."]
    return "\n".join(out) + "\n"


# pass kind -> kernel, for the FP64 realization passes as well (hardware-counter summary)
KIND_KERNELS_ALL = dict(KIND_KERNELS, real_x_fp64=r"f64::rows_pair_kernel<", real_y_fp64=r"f64::strided_kernel<\d+, 0>",
                        real_z_fp64=r"f64::strided_kernel<\d+, 2>")


def traffic_summary(fetch_csv, write_csv, rdreq_csv, bench_json, command, out_json):
    """HBM bytes per launch of each realization pass from rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE
    and optionally the TCC_EA0_RDREQ request-size counters; one counter set per run, KiB or counts per
    dispatch) of `command`, each scaled by what the counter reads for the 1 GiB device-to-device copies
    of --copy-mib 1024.  Writes `out_json` (read by load_traffic) and returns a markdown table."""
    import csv

    def per_dispatch(path):
        d = {}
        with open(path) as fh:
            for r in csv.DictReader(fh):
                key = (int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size"]))
                d.setdefault(key, {})
                d[key][r["Counter_Name"]] = d[key].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        return d

    with open(bench_json) as fh:
        line = json.loads([ln for ln in fh.read().splitlines() if ln.startswith("{")][-1])
    runs = {"FETCH_SIZE": per_dispatch(fetch_csv), "WRITE_SIZE": per_dispatch(write_csv)}
    rd = per_dispatch(rdreq_csv) if rdreq_csv else {}
    copy_kib = 1024 * 1024  # one 1 GiB copy, in KiB
    factor = {}
    for name, d in runs.items():
        # the 1 GiB copies: the largest copy dispatches (smaller copies of the run are left out)
        vals = sorted((v[name] for (_, k, _), v in d.items() if "copyBuffer" in k and name in v), reverse=True)
        vals = [x for x in vals if x >= 0.5 * vals[0]] if vals else []
        factor[name] = copy_kib / (sum(vals) / len(vals)) if vals else 1.0
    kinds = {}
    for kind, pat in KIND_KERNELS_ALL.items():
        p = re.compile(pat)
        ks = {}
        for name, d in runs.items():
            for (_, k, grid), v in d.items():
                if p.search(k):
                    ks.setdefault((k, grid), {}).setdefault(name, []).append(v[name])
        if not ks or kind not in line["passes"]:
            continue
        # the kernel and grid the pass runs most often (the FP64 y kernel also serves the spectrum once)
        (k, grid), c = max(ks.items(), key=lambda kv: len(kv[1].get("FETCH_SIZE", [])))
        f = sum(c["FETCH_SIZE"]) / len(c["FETCH_SIZE"])
        w = sum(c["WRITE_SIZE"]) / len(c["WRITE_SIZE"])
        rq = [v for (_, kk, gg), v in rd.items() if kk == k and gg == grid]
        rbytes = None
        if rq:
            rbytes = sum(32 * v.get("TCC_EA0_RDREQ_32B_sum", 0) + 64 * v.get("TCC_EA0_RDREQ_64B_sum", 0) +
                         128 * v.get("TCC_EA0_RDREQ_128B_sum", 0) for v in rq) / len(rq)
        kinds[kind] = {"kernel": k, "grid_threads": grid, "dispatches": len(c["FETCH_SIZE"]), "fetch_kib": f,
                       "write_kib": w, "hbm_bytes_per_launch": (f * factor["FETCH_SIZE"] + w * factor["WRITE_SIZE"]) * 1024,
                       "read_bytes_from_request_sizes": rbytes,
                       "algorithmic_bytes_per_launch": line["passes"][kind]["GB_per_launch"] * 1e9}
    with open(out_json, "w") as fh:
        json.dump({"method": "rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch), each scaled by "
                             "what it reads for a 1 GiB hipMemcpy device-to-device copy",
                   "command": command, "fetch_factor": factor["FETCH_SIZE"], "write_factor": factor["WRITE_SIZE"],
                   "kinds": kinds}, fh, indent=1)
    out = [f"Calibration (1 GiB device-to-device copies): FETCH_SIZE x {factor['FETCH_SIZE']:.3f}, "
           f"WRITE_SIZE x {factor['WRITE_SIZE']:.3f}.", "",
           "| pass | kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM GB (scaled) | read GB from request sizes | "
           "algorithmic GB | ratio |", "|---|---|---|---|---|---|---|---|"]
    for kind, v in kinds.items():
        rb = f"{v['read_bytes_from_request_sizes'] / 1e9:.3f}" if v["read_bytes_from_request_sizes"] else "--"
        out.append(f"| {kind} | `{v['kernel']}` | {v['fetch_kib']:.0f} | {v['write_kib']:.0f} | "
                   f"{v['hbm_bytes_per_launch'] / 1e9:.3f} | {rb} | {v['algorithmic_bytes_per_launch'] / 1e9:.3f} | "
                   f"{v['hbm_bytes_per_launch'] / v['algorithmic_bytes_per_launch']:.2f} |")
    return "\n".join(out) + "\n"


def host_cpu():
    """(model name, physical cores) of the host, from lscpu."""
    model, cores = "unknown", os.cpu_count() or 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        m = re.search(r"Model name:\s+(.+)", out)
        model = m.group(1).strip() if m else model
        p = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        pairs = {ln for ln in p.splitlines() if ln and not ln.startswith("#")}
        cores = len(pairs) or cores
    except (OSError, subprocess.SubprocessError):
        pass
    return model, cores


def par_text(nreal, base):
    s = STRUCTS[0]
    return ("0 0 0\n1 1 1\n%d %d %d\n%r 1\n%d %r %r %r %r %r %r %r\n%d %d 1 %d\n%s\n1\n3\n0\n"
            % (*GRID, NUGGET, s[0], *s[1:], ISEED, nreal, IGAUSS, base))


def reference_writer(exe, threads):
    """The reference program's GSLIB writer: 8 FP64 realizations of 128x128x64 to .out (its elapsed
    time minus its spectrum and realization timers)."""
    g, nr = (128, 128, 64), 8
    s0 = STRUCTS[0]
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "g.par"), "w") as fh:
            fh.write("0 0 0\n1 1 1\n%d %d %d\n%r 1\n%d %r %r %r %r %r %r %r\n%d %d 1 %d\ng\n1\n1\n0\n"
                     % (*g, NUGGET, s0[0], *s0[1:], ISEED, nr, IGAUSS))
        r = subprocess.run([exe, "g.par"], cwd=wd, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=600)
    if r.returncode != 0:
        return None
    num = lambda pat: float(re.search(pat, r.stdout).group(1))
    tw = num(r"elapsed time:\s+([\d.]+)") - num(r"realizations\s+:\s+([\d.]+)") - \
        num(r"covariance\+spectrum:\s+([\d.]+)")
    return {"sample": "GSLIB .out of 8 FP64 realizations of 128x128x64", "seconds": round(tw, 2)}


def reference_conditioning(exe, threads, nd=2000, ndmax=16):
    """The reference program's own kriging ('time conditioning', kriging part) on 128x128x64 with the
    data density, search and ndmax of the GPU conditioning leg, 2 realizations."""
    import numpy as np

    g = (128, 128, 64)
    nodes = g[0] * g[1] * g[2]
    nds = max(1, int(round(nd * nodes / (GRID[0] * GRID[1] * GRID[2]))))
    rng = np.random.default_rng(12)
    idx = rng.choice(nodes, nds, replace=False)
    ci, cj, ck = np.unravel_index(idx, (g[2], g[1], g[0]))[::-1]
    s0 = STRUCTS[0]
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "d.dat"), "w") as fh:
            fh.write("d\n4\nx\ny\nz\nv\n" + "".join("%d %d %d %.8g\n" % (a, b, c, v) for a, b, c, v in
                                                         zip(ci, cj, ck, rng.normal(size=nds))))
        with open(os.path.join(wd, "c.par"), "w") as fh:
            fh.write("0 0 0\n1 1 1\n%d %d %d\n%r 1\n%d %r %r %r %r %r %r %r\n%d 2 1 %d\nc\n1\n3\n1\n"
                     % (*g, NUGGET, s0[0], *s0[1:], ISEED, IGAUSS) +
                     "d.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.0\n-5 5\n1 0\n1 0\n%r %r %r\n%r %r %r\n%d\n"
                     % (s0[2], s0[3], s0[4], s0[5], s0[6], s0[7], ndmax))
        r = subprocess.run([exe, "c.par"], cwd=wd, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=900)
    m = re.search(r"time conditioning:\s+([\d.]+) s \(kriging\)", r.stdout) if r.returncode == 0 else None
    if not m:
        return None
    tk = float(m.group(1))
    return {"sample": f"{g[0]}x{g[1]}x{g[2]}, {nds} data, ndmax {ndmax}, 2 realizations", "kriging_s": tk,
            "Mnodes_s": round(nodes / max(tk, 1e-9) / 1e6, 3)}


def cpu_threads():
    """Threads of the CPU legs: the host's physical cores, bounded by the process's CPU share
    (OMP_NUM_THREADS when set, the CPUs this process may run on)."""
    model, phys = host_cpu()
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    return max(1, min(phys, share)), model, phys


def cpu_baseline(nreal=2, runs=3):
    """Reference program on the host cores: 256^3 (512^3 periodic), `nreal` realizations, best of `runs`
    (its "time realizations" line, 0.01 s resolution); plus its GSLIB writer and its conditioning on
    bounded samples (the CPU sides of the output and conditioning legs)."""
    threads, model, phys = cpu_threads()
    exe = os.path.join(ROOT, "oracle", "_ref", "specsim3d")
    ncell = GRID[0] * GRID[1] * GRID[2]
    if os.path.exists(exe):
        best = None
        for _ in range(runs):
            with tempfile.TemporaryDirectory() as wd:
                with open(os.path.join(wd, "b.par"), "w") as fh:
                    fh.write(par_text(nreal, "b"))
                t0 = time.perf_counter()
                r = subprocess.run([exe, "b.par"], cwd=wd, capture_output=True, text=True,
                                   env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=600)
                wall = time.perf_counter() - t0
            if r.returncode != 0:
                best = None
                break
            tr = float(re.search(r"realizations\s+:\s+([\d.]+)", r.stdout).group(1))
            ts = float(re.search(r"covariance\+spectrum:\s+([\d.]+)", r.stdout).group(1))
            if best is None or tr < best[0]:
                best = (tr, ts, wall)
        if best is not None:
            tr, ts, wall = best
            return {"value": nreal * ncell / tr / 1e9, "unit": "Gcells/s", "cores": threads, "kind": "reference",
                    "sample": f"reference program (amdflang -O3 -fopenmp; gfortran is not installed), 256^3 grid / "
                              f"512^3 periodic, {nreal} realizations, best of {runs}: time realizations {tr:.2f} s, "
                              f"covariance+spectrum {ts:.2f} s, wall {wall:.1f} s",
                    "host": {"cpu_model": model, "physical_cores": phys, "threads_used": threads,
                             "why_threads": "physical cores bounded by this process's CPU share (OMP_NUM_THREADS, "
                                            "CPU affinity)"},
                    "gslib_writer": reference_writer(exe, threads),
                    "conditioning": reference_conditioning(exe, threads)}
    # fallback: the oracle's C restatement, one thread, 128^3 grid (256^3 periodic)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O

    t0 = time.perf_counter()
    O.simulate(STRUCTS, NUGGET, 1, 1, 1, 128, 128, 128, 1, ISEED, 1, IGAUSS)
    dt = time.perf_counter() - t0
    return {"value": 128 ** 3 / dt / 1e9, "unit": "Gcells/s", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement, 128^3 grid / 256^3 periodic, 1 realization incl. spectrum: {dt:.2f} s",
            "host": {"cpu_model": model, "physical_cores": phys, "threads_used": 1}}


def spectrum_build_ms(S, eng, grid):
    """Wall time of a spectrum build of the main workload's model once the process has built one (best
    of 3): a full build (FP64 and FP32 amplitudes) and an FP32-only one (SS3D_OPT_SPEC_FP32: the
    passes after the x pass over the columns kx <= npx/2, the way the request loop and the batch
    build the spectra of FP32 work)."""
    out = {}
    try:
        for name, opt in (("full", 0), ("fp32_only", 1)):
            eng.set_option(S.OPT_SPEC_FP32, opt)
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                sp = eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *grid)
                eng.synchronize()
                t = (time.perf_counter() - t0) * 1e3
                sp.close()
                best = t if best is None else min(best, t)
            out[name] = round(best, 2)
    finally:
        eng.set_option(S.OPT_SPEC_FP32, 0)
    return out


def serve(S, eng, rate, duration, seed, nreal=64, nslots=2, spot_frac=0.05, repeat_frac=0.3):
    """Poisson arrivals of 256^3 requests like the headline workload (igauss=1, FP32 output left in HBM): 30% repeat
    an earlier model (spectral-density cache), the rest draw ranges and angles anew; 5% carry an FP64 spot
    check of realization 0.  Returns latency percentiles and throughput."""
    import threading
    import numpy as np

    rng = np.random.default_rng(seed)
    n = max(1, int(rng.poisson(rate * duration)))
    arrivals = np.sort(rng.uniform(0.0, duration, n))
    models = []
    reqs = []
    for i in range(n):
        if models and rng.random() < repeat_frac:
            m = models[int(rng.integers(len(models)))]
        else:
            m = [(2, 0.7, float(rng.uniform(60, 100)), float(rng.uniform(30, 50)), float(rng.uniform(5, 15)),
                  float(rng.uniform(0, 180)), float(rng.uniform(-30, 30)), 0.0)]
            models.append(m)
        reqs.append(S.Server.request(i, m, NUGGET, 1, 1, 1, *GRID, 1, int(rng.integers(1, 2**31 - 1)), nreal, IGAUSS,
                                     S.FP32, spot_check=bool(rng.random() < spot_frac)))
    done = {}
    srv = S.Server(eng, nslots=nslots)
    try:
        def collect():
            while len(done) < n:
                for c in srv.poll(timeout=0.2):
                    done[c["id"]] = c
        th = threading.Thread(target=collect)
        th.start()
        t_start = srv.clock()
        t_note = time.perf_counter()
        for t, r in zip(arrivals, reqs):
            while srv.clock() - t_start < t:
                time.sleep(min(0.002, max(0.0, t - (srv.clock() - t_start))))
            srv.submit(r)
            if time.perf_counter() - t_note > 10.0:  # progress of a long run
                t_note = time.perf_counter()
                print(f"[serve {rate:g} req/s] {r.id + 1}/{n} submitted, {len(done)} done", file=sys.stderr, flush=True)
        while th.is_alive():
            th.join(timeout=10.0)
            if th.is_alive():
                print(f"[serve {rate:g} req/s] draining: {len(done)}/{n} done", file=sys.stderr, flush=True)
    finally:
        srv.close()
    order = sorted(done.values(), key=lambda c: c["t_submit"])
    lat = np.array([c["t_done"] - c["t_submit"] for c in order]) * 1e3
    third = max(1, len(lat) // 3)
    ok = all(c["status"] == 0 for c in done.values())
    t_first = min(c["t_submit"] for c in done.values())
    t_last = max(c["t_done"] for c in done.values())
    cells = n * nreal * GRID[0] * GRID[1] * GRID[2]
    spots = [c["spot_rel_rms"] for c in done.values() if c["spot_rel_rms"] >= 0]
    return {"offered_req_s": rate, "seconds_of_arrivals": duration, "requests": n, "all_ok": ok,
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "p99_ms_first_third": round(float(np.percentile(lat[:third], 99)), 2),
            "p99_ms_last_third": round(float(np.percentile(lat[-third:], 99)), 2),
            "throughput_gcells_s": round(cells / (t_last - t_first) / 1e9, 2),
            "completed_req_s": round(n / (t_last - t_first), 2),
            "cache_hit_frac": round(sum(c["cache_hit"] for c in done.values()) / n, 3),
            "spot_checks": len(spots), "spot_max_rel_rms": max(spots) if spots else None}


def uneven_batch(S, eng, seed, njobs=32, nlanes=8, reps=5):
    """32 FP32 jobs of uneven grids (nx, ny in {32,64,128,256}, nz in {1,16,64,128}; spherical,
    exponential or gaussian model, ranges U(5,100) per axis, azimuth U(0,180), dip U(-30,30),
    nugget U(0,0.3), sill 1 - nugget, nreal U{1..16}, ipad=1) captured into one HIP graph on
    `nlanes` concurrent branches.  Creation (spectral densities, per-grid autotuning, one eager run,
    capture) is untimed; timed: `reps` graph replays, then `reps` eager enqueues of the same passes."""
    import numpy as np

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(njobs):
        nx, ny = (int(v) for v in rng.choice([32, 64, 128, 256], 2))
        nz = int(rng.choice([1, 16, 64, 128]))
        nug = float(rng.uniform(0.0, 0.3))
        st = [(int(rng.integers(1, 4)), 1.0 - nug, *(float(v) for v in rng.uniform(5, 100, 3)),
               float(rng.uniform(0, 180)), float(rng.uniform(-30, 30)), 0.0)]
        reqs.append(S.Server.request(i, st, nug, 1, 1, 1, nx, ny, nz, 1, int(rng.integers(1, 2**31 - 1)),
                                     int(rng.integers(1, 17)), 1, S.FP32))
    # untuned: a fresh context (empty tuning table: largest radix 32, default tiles), no batch tuning
    with S.Engine(0) as e0:
        e0.set_option(S.OPT_BATCH_AUTOTUNE, 0)
        b0 = S.Batch(e0, reqs, nlanes=nlanes)
        try:
            b0.run(graph=True)
            e0.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                b0.run(graph=True)
            e0.synchronize()
            untuned = (time.perf_counter() - t0) / reps
        finally:
            b0.close()
    t0 = time.perf_counter()
    b = S.Batch(eng, reqs, nlanes=nlanes)
    eng.synchronize()
    t_create = time.perf_counter() - t0
    try:
        out = {}
        for mode, graph in (("graph", True), ("eager", False)):
            b.run(graph=graph)
            eng.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                b.run(graph=graph)
            eng.synchronize()
            out[mode] = (time.perf_counter() - t0) / reps
    finally:
        b.close()
    # the plans and tiles the batch's autotuning left for every pass length of its grids
    picks = {}
    for r in reqs:
        g = S.periodic_grid([(r.st[0].itype, r.st[0].sill, r.st[0].ahmax, r.st[0].ahmin, r.st[0].avert, r.st[0].az,
                              r.st[0].dip, r.st[0].plg)], 1, 1, 1, r.nx, r.ny, r.nz, 1)
        lx, ly, lz = (int(v).bit_length() - 1 for v in g)
        for name, kind, L in (("real_z_fp32", 6, lz), ("real_y_fp32", 7, ly), ("real_x_c2r_fp32", 8, max(lx - 1, 0))):
            t = eng.tuning_get_tile(kind, L)
            picks[f"{name} 2^{L}"] = f"radix 2^{t['rm']}, tile {t['tile']}{' padded' if t['lds_pad'] else ''}" + \
                (" (variant 1)" if t["variant"] else "")
    return {"jobs": njobs, "lanes": nlanes, "cells_per_batch": b.cells, "reps": reps,
            "graph_gcells_s": round(b.cells / out["graph"] / 1e9, 3), "graph_ms": round(out["graph"] * 1e3, 3),
            "eager_gcells_s": round(b.cells / out["eager"] / 1e9, 3), "eager_ms": round(out["eager"] * 1e3, 3),
            "create_s": round(t_create, 2),
            "untuned_graph_gcells_s": round(b.cells / untuned / 1e9, 3),
            "tuned_picks": dict(sorted(picks.items())),
            "distinct_periodic_grids": len({S.periodic_grid([(r.st[0].itype, r.st[0].sill, r.st[0].ahmax,
                                                              r.st[0].ahmin, r.st[0].avert, r.st[0].az, r.st[0].dip,
                                                              r.st[0].plg)], 1, 1, 1, r.nx, r.ny, r.nz, 1)
                                             for r in reqs})}


def large_job(S, eng, hip, nreal=32, checks=(0, 31)):
    """One large job: 512x512x256 output in a 1024x1024x512 periodic grid, gaussian model
    exp(-3r^2), sill 1, ranges 150/60/20, azimuth 120, dip -20, ipad=1, igauss=1, nreal FP32
    realizations kept in HBM.  Reports the realization throughput, the FP32-vs-FP64 relative RMS
    error of the checked realizations (FP64 drop-in path on the same draws) and the device
    variograms of every realization along the 3 grid axes."""
    import numpy as np

    sts = [(3, 1.0, 150.0, 60.0, 20.0, 120.0, -20.0, 0.0)]
    nx, ny, nz = 512, 512, 256
    grid = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    ncell = nx * ny * nz
    out = hip.malloc(nreal * ncell * 4)
    stats = hip.malloc(nreal * 3 * 8)
    # untuned: a fresh context (empty tuning table), same work
    with S.Engine(0) as e0:
        sp0 = e0.spectrum(sts, 0.0, 1.0, 1.0, 1.0, *grid)
        try:
            e0.realize_device(sp0, 4711, 1, 0, nreal, nx, ny, nz, S.FP32, out, stats)
            e0.synchronize()
            t0 = time.perf_counter()
            e0.realize_device(sp0, 4711, 1, 0, nreal, nx, ny, nz, S.FP32, out, stats)
            e0.synchronize()
            t_untuned = time.perf_counter() - t0
        finally:
            sp0.close()
    t0 = time.perf_counter()
    spec = eng.spectrum(sts, 0.0, 1.0, 1.0, 1.0, *grid)
    eng.synchronize()
    spec_s = time.perf_counter() - t0
    eng.autotune(spec, nx, ny, nz, 1, reps=1)
    plan = tile_plan(eng, grid)
    try:
        eng.realize_device(spec, 4711, 1, 0, nreal, nx, ny, nz, S.FP32, out, stats)  # warm-up (same work)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.realize_device(spec, 4711, 1, 0, nreal, nx, ny, nz, S.FP32, out, stats)
        eng.synchronize()
        t_real = time.perf_counter() - t0
        # FP32 vs FP64 on the same draws
        errs = []
        for r in checks:
            f64, _, _, _ = eng.realize(spec, 4711, 1, r, 1, nx, ny, nz, S.FP64)
            f32 = np.empty(ncell, dtype=np.float32)
            hip.copy_d2h(f32.ctypes.data, out + r * ncell * 4, ncell * 4)
            d = f32.astype(np.float64) - f64.ravel()
            errs.append(float(np.sqrt(np.mean(d * d) / np.mean(f64 * f64))))
        # variograms along x, y, z of every realization (device field)
        dirs = [(1, 0, 0, 128), (0, 1, 0, 64), (0, 0, 1, 32)]
        t0 = time.perf_counter()
        gam = eng.variogram(out, dirs, device_shape=(nreal, nz, ny, nx), precision=S.FP32)
        t_var = time.perf_counter() - t0
    finally:
        hip.free(out)
        hip.free(stats)
        spec.close()
    c0 = 1.0
    mean_g = gam.mean(axis=0)
    model = [[c0 - S.covariance(sts, 0.0, l * d[0], l * d[1], l * d[2]) for l in (1, d[3] // 2, d[3])] for d in dirs]
    return {"workload": "512x512x256 (1024x1024x512 periodic), gaussian sill 1, ranges 150/60/20, az 120, dip -20, "
                        f"ipad=1, igauss=1, {nreal} FP32 realizations in HBM",
            "periodic_grid": list(grid), "spectrum_s": round(spec_s, 3), "radix_plan": plan,
            "realizations_s": round(t_real, 4), "gcells_s": round(nreal * ncell / t_real / 1e9, 2),
            "untuned_gcells_s": round(nreal * ncell / t_untuned / 1e9, 2),
            "fp32_vs_fp64_rel_rms": {str(r): e for r, e in zip(checks, errs)},
            "variogram": {"dirs": dirs, "seconds": round(t_var, 4),
                          "pairs_per_s": float(sum(S.npairs(nx, ny, nz, l * d[0], l * d[1], l * d[2])
                                                   for d in dirs for l in range(1, d[3] + 1)) * nreal / t_var),
                          "mean_at_lags_1_half_max": [[round(float(mean_g[k, l - 1]), 4) for l in (1, d[3] // 2, d[3])]
                                                      for k, d in enumerate(dirs)],
                          "model_at_same_lags": [[round(v, 4) for v in m] for m in model]}}


def writer_leg(S, eng, spec, nreal):
    """Output files.  (1) The workload's nreal realizations to a raw float32 .bin file, FP32 and
    FP64 fields: streamed by the library's writer (device compute, device-to-host copy and file
    writes overlapped) vs the serial order of the reference (every realization to host memory,
    then the file).  (2) GSLIB text (.out, es15.7, every realization of a node per line) of 8
    FP64 realizations of 128x128x64 through the writer (the reference program's own writer on the
    same case is timed by the CPU baseline)."""
    import numpy as np

    nx, ny, nz = GRID
    ncell = nx * ny * nz
    wd = tempfile.mkdtemp(prefix="ss3d_out_")
    out = {"workload": f"{nx}x{ny}x{nz}, {nreal} realizations to a raw float32 .bin (files in {wd}, no fsync)"}
    try:
        for prec, name in ((S.FP32, "fp32"), (S.FP64, "fp64")):
            path = os.path.join(wd, name + ".bin")
            t0 = time.perf_counter()
            with S.Writer(eng, {"bin": path}, nx, ny, nz, nreal, prec) as w:
                done = 0
                while done < nreal:
                    c = min(w.slot, nreal - done)
                    w.realize(spec, ISEED, IGAUSS, c)
                    done += c
            t_stream = time.perf_counter() - t0
            size = os.path.getsize(path)
            head = np.fromfile(path, dtype="<f4", count=ncell)
            os.remove(path)
            t0 = time.perf_counter()
            f = eng.realize(spec, ISEED, IGAUSS, 0, nreal, nx, ny, nz, prec)[0]
            t_host = time.perf_counter() - t0
            (f if prec == S.FP32 else f.astype(np.float32)).tofile(path)
            t_serial = time.perf_counter() - t0
            same = bool(np.array_equal(head, f[0].astype(np.float32).ravel()))
            del f
            os.remove(path)
            out[name] = {"streamed_s": round(t_stream, 3), "streamed_GB_s": round(size / t_stream / 1e9, 2),
                         "serial_s": round(t_serial, 3), "serial_realize_to_host_s": round(t_host, 3),
                         "file_GB": round(size / 1e9, 3), "speedup": round(t_serial / t_stream, 2),
                         "first_realization_equal": same}
        # GSLIB text
        g = (128, 128, 64)
        nr = 8
        grid = S.periodic_grid(STRUCTS, 1.0, 1.0, 1.0, *g, 1)
        sp = eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *grid)
        path = os.path.join(wd, "g.out")
        t0 = time.perf_counter()
        with S.Writer(eng, {"gslib": path}, *g, nr, S.FP64) as w:
            w.realize(sp, ISEED, IGAUSS, nr)
        t_text = time.perf_counter() - t0
        size = os.path.getsize(path)
        os.remove(path)
        sp.close()
        text = {"case": "128x128x64, 8 FP64 realizations, es15.7", "seconds": round(t_text, 3),
                "MB_s": round(size / t_text / 1e6, 1), "file_MB": round(size / 1e6, 1)}
        out["gslib_text"] = text
    finally:
        import shutil

        shutil.rmtree(wd, ignore_errors=True)
    return out


def conditioning_leg(S, eng, hip, spec, nreal, nd=2000, ndmax=16):
    """Conditioning of the workload's realizations on nd data at random nodes (values N(0, 1), simple
    kriging, ndmax nearest data in the model's anisotropic ellipsoid): the kriging weights of all
    256^3 nodes once, then each step's 64 FP32 realizations conditioned on the device (the reference
    program's own kriging on a bounded sample is timed by the CPU baseline)."""
    import numpy as np

    nx, ny, nz = GRID
    ncell = nx * ny * nz
    rng = np.random.default_rng(11)
    idx = rng.choice(ncell, nd, replace=False)
    i, j, k = np.unravel_index(idx, (nz, ny, nx))[::-1]
    rows = np.c_[i, j, k, rng.normal(size=nd)].astype(float)
    s0 = STRUCTS[0]
    prep = S.cond_prepare(rows, GRID, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 2, 3, 4)
    t0 = time.perf_counter()
    cond = eng.conditioning(STRUCTS, NUGGET, GRID, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), prep["ijk"], prep["dg"],
                            s0[2], s0[3], s0[4], (s0[5], s0[6], s0[7]), ndmax)
    t_weights = time.perf_counter() - t0
    info = cond.info()
    out = hip.malloc(nreal * ncell * 4)
    try:
        eng.realize_device(spec, ISEED, IGAUSS, 0, nreal, nx, ny, nz, S.FP32, out)
        cond.apply(out, S.FP32, nreal)  # warm-up
        eng.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for r in range(reps):
            eng.realize_device(spec, ISEED, IGAUSS, r * nreal, nreal, nx, ny, nz, S.FP32, out)
            cond.apply(out, S.FP32, nreal)
        eng.synchronize()
        t_step = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for r in range(reps):
            cond.apply(out, S.FP32, nreal)
        eng.synchronize()
        t_apply = (time.perf_counter() - t0) / reps
    finally:
        hip.free(out)
        cond.close()
    res = {"workload": f"{nx}x{ny}x{nz}, {nd} data at random nodes, ndmax {ndmax}, search {s0[2]}/{s0[3]}/{s0[4]} "
                       f"az {s0[5]} dip {s0[6]}, simple kriging, {nreal} FP32 realizations per step",
           "weights_s": round(t_weights, 3), "weights_Mnodes_s": round(ncell / t_weights / 1e6, 1),
           "mean_neighbours": round(info["nsum"] / ncell, 2), "apply_ms_per_step": round(t_apply * 1e3, 2),
           "step_with_conditioning_ms": round(t_step * 1e3, 2),
           "gcells_s_conditioned": round(nreal * ncell / t_step / 1e9, 2)}
    return res


def spectrum_file_leg(S, eng, spec, grid, build_ms):
    """The workload's spectral density saved to and loaded from a file (ss3d_spectrum_save/load), timed
    against building it on the device."""
    wd = tempfile.mkdtemp(prefix="ss3d_spec_")
    path = os.path.join(wd, "c2.spec")
    try:
        t0 = time.perf_counter()
        spec.save(path)
        t_save = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        ld = eng.load_spectrum(path, STRUCTS, NUGGET, 1.0, 1.0, 1.0, grid)
        t_load = time.perf_counter() - t0
        ld.close()
    finally:
        import shutil

        shutil.rmtree(wd, ignore_errors=True)
    return {"file_GB": round(size / 1e9, 3), "save_s": round(t_save, 3), "load_s": round(t_load, 3),
            "build_on_device_ms": round(build_ms, 2)}


def tile_plan(eng, grid):
    """The tuned radix plan and tile of the three FP32 passes of a periodic grid."""
    lx, ly, lz = (int(v).bit_length() - 1 for v in grid)
    return {name: eng.tuning_get_tile(kind, L) for name, kind, L in
            (("real_z_fp32", 6, lz), ("real_y_fp32", 7, ly), ("real_x_c2r_fp32", 8, max(lx - 1, 0)))}


def load_traffic(kind):
    """HBM bytes per launch of a pass kind measured with rocprofv3 counters (profiles/traffic_c2.json:
    FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what they read for a 1 GiB device copy)."""
    path = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    v = t.get("kinds", {}).get(kind)
    return None if v is None else v.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nreal", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=45.0,
                    help="Poisson rate (req/s) of the sustained serving run")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="seconds of arrivals of the sustained run")
    ap.add_argument("--overload-seconds", type=float, default=5.0,
                    help="seconds of arrivals at 200 req/s (the overload run)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving legs and print their JSON (not the benchmark line)")
    ap.add_argument("--spectrum-only", type=int, default=0, metavar="REPS",
                    help="build the main workload's spectral density REPS times in full and REPS times FP32-only, "
                         "print the wall times as JSON and exit (for kernel traces of the spectrum passes)")
    ap.add_argument("--trace-summary", default="",
                    help="summarize a rocprofv3 kernel_trace.csv of this script (no GPU work) and exit")
    ap.add_argument("--marker-csv", default="", help="with --trace-summary: the run's marker_api_trace.csv")
    ap.add_argument("--bench-json", default="", help="with --trace-summary: the run's JSON line")
    ap.add_argument("--trace-command", default="rocprofv3 --kernel-trace --marker-trace --stats --output-format csv "
                                               "-d gpurun_out/prof_c2 -o c2 -- python3 -u bench.py",
                    help="with --trace-summary: the command that made the trace (quoted in the summary)")
    ap.add_argument("--traffic-summary", nargs=3, default=None, metavar=("FETCH_CSV", "WRITE_CSV", "RDREQ_CSV"),
                    help="summarize the counter_collection.csv files of three rocprofv3 --pmc runs of this script "
                         "(FETCH_SIZE; WRITE_SIZE; the TCC_EA0_RDREQ request sizes, or '-') into "
                         "profiles/traffic_c2.json and a table on stdout (no GPU work), then exit; --bench-json names "
                         "one run's JSON line, --trace-command the command the runs shared")
    ap.add_argument("--serve-slots", type=int, default=2,
                    help="lanes of the request loop (2: a spectrum build or a request overlaps another request; "
                         "more lanes thrash the caches shared by their amplitude tables)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-writer", action="store_true")
    ap.add_argument("--no-cond", action="store_true")
    ap.add_argument("--plan", default="",
                    help="radix plans Z,Y,X (log2 of the largest radix of the FP32 z fill, y and C2R x passes) "
                         "instead of autotuning, so that profiling runs launch the same kernels")
    ap.add_argument("--fuse", type=int, default=-1,
                    help="SS3D_OPT_FUSE of the engine (0: separate launches, 1: fused); default: the library's")
    ap.add_argument("--zrun", type=int, default=-1,
                    help="SS3D_OPT_ZRUN of the engine (0: row layouts, 1: z-run layout); default: the library's")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run device-to-device copies of this many MiB (hardware-counter calibration)")
    args = ap.parse_args()
    if args.trace_summary:
        sys.stdout.write(trace_summary(args.trace_summary, args.marker_csv, args.bench_json, args.trace_command))
        return
    if args.traffic_summary:
        f, w, r = args.traffic_summary
        sys.stdout.write(traffic_summary(f, w, None if r == "-" else r, args.bench_json, args.trace_command,
                                         os.path.join(ROOT, "profiles", "traffic_c2.json")))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this engine has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Realizations are independent: a multi-GPU run would give each GPU a range of them.", file=sys.stderr)
        sys.exit(2)

    import specsim3d_amd as S

    hip = Hip()
    regions = Regions()
    if args.copy_mib > 0:
        n = args.copy_mib << 20
        a, b = hip.malloc(n), hip.malloc(n)
        for _ in range(3):
            hip.copy_d2d(b, a, n)
        hip.free(a)
        hip.free(b)
    eng = S.Engine(0)
    if args.fuse >= 0:
        eng.set_option(S.OPT_FUSE, args.fuse)
    if args.zrun >= 0:
        eng.set_option(S.OPT_ZRUN, args.zrun)
    if args.spectrum_only > 0:
        grid = S.periodic_grid(STRUCTS, 1.0, 1.0, 1.0, *GRID, 1)
        eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *grid).close()  # loads the kernels
        times = {}
        try:
            for name, opt in (("full", 0), ("fp32_only", 1)):
                eng.set_option(S.OPT_SPEC_FP32, opt)
                times[name] = []
                for _ in range(args.spectrum_only):
                    t0 = time.perf_counter()
                    eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *grid).close()
                    times[name].append(round((time.perf_counter() - t0) * 1e3, 3))
        finally:
            eng.set_option(S.OPT_SPEC_FP32, 0)
        eng.close()
        print(json.dumps({"spectrum_build_ms": times, "periodic_grid": grid}), flush=True)
        return
    if args.serve_only:
        # the radix plans the full run would have tuned for this grid
        spec = eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *S.periodic_grid(STRUCTS, 1.0, 1.0, 1.0, *GRID, 1))
        eng.autotune(spec, *GRID, IGAUSS, reps=3)
        spec.close()
        sustained = serve(S, eng, args.serve_rate, args.serve_seconds, 1, nslots=args.serve_slots)
        overload = serve(S, eng, 200.0, args.overload_seconds, 2, nslots=args.serve_slots)
        eng.close()
        print(json.dumps({"serving": {"sustained": sustained, "overload": overload}}), flush=True)
        return
    nx, ny, nz = GRID
    grid = S.periodic_grid(STRUCTS, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    t0 = time.perf_counter()
    spec = eng.spectrum(STRUCTS, NUGGET, 1.0, 1.0, 1.0, *grid)
    eng.synchronize()
    spec_ms = (time.perf_counter() - t0) * 1e3
    spec_builds = spectrum_build_ms(S, eng, grid)
    nreal = args.nreal
    ncell = nx * ny * nz
    untuned = None
    if not args.plan:
        # the step with the untuned table (largest radix 32, default tiles): 1 warm-up + 2 steps
        o = hip.malloc(nreal * ncell * 4)
        st0 = hip.malloc(nreal * 3 * 8)
        for k in range(3):
            if k == 1:
                eng.synchronize()
                t0 = time.perf_counter()
            eng.realize_device(spec, ISEED, IGAUSS, k * nreal, nreal, nx, ny, nz, S.FP32, o, st0)
        eng.synchronize()
        untuned = 2 * nreal * ncell / (time.perf_counter() - t0) / 1e9
        hip.free(o)
        hip.free(st0)
    t0 = time.perf_counter()
    if args.plan:
        lx, ly, lz = (int(v).bit_length() - 1 for v in grid)
        kinds = {"real_z_fp32": (6, lz), "real_y_fp32": (7, ly), "real_x_c2r_fp32": (8, lx - 1)}
        for (kind, logn), rm in zip(kinds.values(), (int(v) for v in args.plan.split(","))):
            eng.tuning_set(kind, logn, rm)
        plan = {k: eng.tuning_get(*v) for k, v in kinds.items()}
    else:
        plan = eng.autotune(spec, nx, ny, nz, IGAUSS, reps=3)
    tune_s = time.perf_counter() - t0
    plans_tiles = tile_plan(eng, grid)
    out32 = hip.malloc(nreal * ncell * 4)
    stats = hip.malloc(nreal * 3 * 8)

    def run(prec, out, step, count):
        eng.realize_device(spec, ISEED, IGAUSS, step * nreal, count, nx, ny, nz, prec, out, stats)

    # ---- FP32 batched path
    for w in range(args.warmup):
        run(S.FP32, out32, w, nreal)
    eng.synchronize()
    eng.profile(True)
    regions.resume()
    t0 = time.perf_counter()
    for k in range(args.steps):
        run(S.FP32, out32, args.warmup + k, nreal)
    eng.synchronize()
    t32 = time.perf_counter() - t0
    regions.pause()
    eng.profile(False)
    prof32 = eng.profile_read()

    # ---- FP64 drop-in path (same batch, FP64 output); its last timed step realizes the draws of
    # the last FP32 step, so that realization 0 of both compares the FP32 field with the FP64 one
    t64, prof64, rel32 = None, {}, None
    if args.fp64_steps > 0:
        out64 = hip.malloc(nreal * ncell * 8)
        last = args.warmup + args.steps - 1
        run(S.FP64, out64, 0, nreal)
        eng.synchronize()
        eng.profile(True)
        t0 = time.perf_counter()
        for k in range(args.fp64_steps):
            run(S.FP64, out64, max(0, last - args.fp64_steps + 1 + k), nreal)
        eng.synchronize()
        t64 = time.perf_counter() - t0
        eng.profile(False)
        prof64 = eng.profile_read()
        import numpy as np

        f64, f32 = np.empty(ncell, np.float64), np.empty(ncell, np.float32)
        hip.copy_d2h(f64.ctypes.data, out64, ncell * 8)
        hip.copy_d2h(f32.ctypes.data, out32, ncell * 4)
        rel32 = float(np.sqrt(np.mean((f32.astype(np.float64) - f64) ** 2) / np.mean(f64 ** 2)))
        hip.free(out64)
    hip.free(out32)
    hip.free(stats)
    writer = None if args.no_writer else writer_leg(S, eng, spec, nreal)
    specfile = None if args.no_writer else spectrum_file_leg(S, eng, spec, grid, spec_ms)
    conditioning = None if args.no_cond else conditioning_leg(S, eng, hip, spec, nreal)
    spec.close()

    uneven = None if args.no_batch else uneven_batch(S, eng, 3)
    large = None if args.no_large else large_job(S, eng, hip)
    serving = None
    if not args.no_serve:
        sustained = serve(S, eng, args.serve_rate, args.serve_seconds, 1, nslots=args.serve_slots)
        overload = serve(S, eng, 200.0, args.overload_seconds, 2, nslots=args.serve_slots)
        serving = {"sustained": sustained, "overload": overload}
    eng.close()

    cells32 = nreal * ncell * args.steps
    value = cells32 / t32 / 1e9
    fp64_value = nreal * ncell * args.fp64_steps / t64 / 1e9 if t64 else None

    # dominant kernel of the measured (FP32) path
    kind = max(prof32, key=lambda k: prof32[k][0])
    ms, byt, nl = prof32[kind]
    achieved = byt / (ms * 1e-3) / 1e9
    passes = {k: {"ms_per_launch": v[0] / v[2], "GB_per_launch": v[1] / v[2] / 1e9,
                  "GB_s": v[1] / (v[0] * 1e-3) / 1e9, "launches": v[2]} for k, v in {**prof32, **prof64}.items()}

    line = {
        "metric": "Gcells/s at 256^3 (FP32 batched vs FP64 plain); p50/p99 request latency",
        "value": round(value, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "256^3 grid (512^3 periodic), exponential sill 0.7 + nugget 0.3, ranges 80/40/10, "
                               "az 30, dip 15, ipad=1, igauss=1, 64 realizations per step, FP32 output in HBM",
                   "global_batch": nreal, "grid": list(GRID), "periodic_grid": list(grid), "parallelism": "none"},
        "fp64_plain": ({"value": round(fp64_value, 3), "unit": "Gcells/s", "steps": args.fp64_steps,
                        "ms_per_realization": round(t64 / (args.fp64_steps * nreal) * 1e3, 4)} if t64 else None),
        "fp32_over_fp64": round(value / fp64_value, 2) if t64 else None,
        "fp32_vs_fp64_rel_rms": rel32,
        "spectrum_ms": round(spec_ms, 2),
        "spectrum_build_ms": spec_builds,
        "autotune": {"largest_radix_log2": plan, "plans_and_tiles": plans_tiles,
                     "untuned_gcells_s": round(untuned, 3) if untuned else None, "tuned_gcells_s": round(value, 3),
                     "seconds": round(tune_s, 2)},
        "uneven_batch": uneven,
        "large_job": large,
        "output_files": writer,
        "spectrum_file": specfile,
        "conditioning": conditioning,
        "latency_ms": ({"p50": serving["sustained"]["p50_ms"], "p99": serving["sustained"]["p99_ms"],
                        "rate_req_s": args.serve_rate, "serving": serving} if serving else None),
        "timed_region_monotonic_ns": regions.bounds,
        "roofline": {"kernel": kind, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(kind),
                     "algorithmic_bytes_per_launch": byt / nl, "ms_per_launch": ms / nl},
        "passes": passes,
    }
    # the stated targets, against this run
    frac = achieved / HBM_PEAK_GBS
    line["targets"] = {
        "fp32_over_fp64_plain": {"target": ">= 10", "value": line["fp32_over_fp64"],
                                 "met": bool(t64) and value / fp64_value >= 10},
        "dominant_pass_hbm_frac": {"target": ">= 0.60", "value": round(frac, 4), "met": frac >= 0.60},
        "fp32_rel_rms_vs_fp64": {"target": "<= 1e-5", "value": rel32, "met": rel32 is not None and rel32 <= 1e-5},
    }
    if not args.no_cpu:
        cb = cpu_baseline()
        line["cpu_baseline"] = cb
        if writer and cb.get("gslib_writer"):
            t = writer["gslib_text"]
            t["speedup_vs_reference_writer"] = round(cb["gslib_writer"]["seconds"] / t["seconds"], 1)
        if conditioning and cb.get("conditioning"):
            conditioning["weights_speedup_vs_reference"] = round(
                conditioning["weights_Mnodes_s"] / cb["conditioning"]["Mnodes_s"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
