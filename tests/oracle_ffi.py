"""ctypes binding of oracle/liboracle.so -- the CPU checker (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class Structure(ctypes.Structure):
    """Scalar members of the reference's `type structure` (src/specsim3d.f90:95-100)."""
    _fields_ = [("itype", ctypes.c_int32), ("pad_", ctypes.c_int32), ("sill", ctypes.c_double),
                ("ahmax", ctypes.c_double), ("ahmin", ctypes.c_double), ("avert", ctypes.c_double),
                ("az", ctypes.c_double), ("dip", ctypes.c_double), ("plg", ctypes.c_double)]


class SpecDiag(ctypes.Structure):
    _fields_ = [("max_imag", ctypes.c_double), ("max_real", ctypes.c_double), ("c0", ctypes.c_double),
                ("spos", ctypes.c_double), ("sneg", ctypes.c_double)]


class CondParams(ctypes.Structure):
    """The conditioning block without its data file name (src/specsim3d.f90:731-747)."""
    _fields_ = [("ixc", ctypes.c_int32), ("iyc", ctypes.c_int32), ("izc", ctypes.c_int32), ("ivc", ctypes.c_int32),
                ("iwc", ctypes.c_int32), ("itrans", ctypes.c_int32), ("ltail", ctypes.c_int32),
                ("utail", ctypes.c_int32), ("ndmax", ctypes.c_int32), ("pad_", ctypes.c_int32)] + \
               [(n, ctypes.c_double) for n in ("tmin", "tmax", "skmean", "zlo", "zhi", "ltpar", "utpar", "rmax",
                                                "rmed", "rvert", "sa1", "sa2", "sa3")]


class CondInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("nd", "nmerged", "nout", "nsbx", "nsby", "nsbz", "nnone", "nexact",
                                               "nfail", "nt")] + [("nsum", ctypes.c_int64)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle liboracle.so`)")
        L = ctypes.CDLL(LIB_PATH)
        L.or_cutoff.restype = ctypes.c_double
        L.or_covariance.restype = ctypes.c_double
        L.or_covariance.argtypes = [ctypes.POINTER(Structure), _dp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_setrot.argtypes = [ctypes.POINTER(Structure), _dp, _dp]
        L.or_periodic_grid.argtypes = [ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
        L.or_twiddles.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.or_fft1d.argtypes = [_dp, ctypes.c_int, _dp]
        L.or_fft3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_spectrum.argtypes = [ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp, _dp, _dp, ctypes.POINTER(SpecDiag)]
        L.or_seed_state.restype = ctypes.c_int64
        L.or_seed_state.argtypes = [ctypes.c_int]
        L.or_urand.restype = ctypes.c_double
        L.or_urand.argtypes = [_i64p]
        L.or_gauss_pair.argtypes = [_i64p, _dp, _dp]
        L.or_pm_pow.restype = ctypes.c_int64
        L.or_pm_pow.argtypes = [ctypes.c_int64, ctypes.c_uint64]
        L.or_draws_per_realization.restype = ctypes.c_int64
        L.or_draws_per_realization.argtypes = [ctypes.c_int] * 4
        L.or_point_draw_index.restype = ctypes.c_int64
        L.or_point_draw_index.argtypes = [ctypes.c_int] * 7
        L.or_fill.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, _i64p]
        L.or_simulate.restype = ctypes.c_int
        L.or_simulate.argtypes = [ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                  ctypes.POINTER(SpecDiag), ctypes.POINTER(ctypes.c_int)]
        L.or_npair_count.restype = ctypes.c_int64
        L.or_npair_count.argtypes = [ctypes.c_int] * 6
        L.or_vario_lag.restype = ctypes.c_double
        L.or_vario_lag.argtypes = [_dp] + [ctypes.c_int] * 6
        L.or_gcum.restype = ctypes.c_double
        L.or_gcum.argtypes = [ctypes.c_double]
        L.or_gauinv.restype = ctypes.c_double
        L.or_gauinv.argtypes = [ctypes.c_double]
        L.or_heapsort_idx.argtypes = [_dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_nscore.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_backtr.restype = ctypes.c_double
        L.or_backtr.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.or_condition.restype = ctypes.c_int
        L.or_condition.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double,
                                   ctypes.POINTER(CondParams), _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(CondInfo), _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def structures(sts):
    """sts: list of (itype, sill, ahmax, ahmin, avert, az, dip, plg)."""
    arr = (Structure * max(len(sts), 1))()
    for i, s in enumerate(sts):
        arr[i].itype = int(s[0])
        (arr[i].sill, arr[i].ahmax, arr[i].ahmin, arr[i].avert, arr[i].az, arr[i].dip, arr[i].plg) = map(float, s[1:8])
    return arr


def setrot(st):
    rot = np.zeros(9)
    axes = np.zeros(9)
    arr = structures([st])
    lib().or_setrot(arr, _p(rot), _p(axes))
    return rot.reshape(3, 3), axes.reshape(3, 3)


def periodic_grid(sts, dx, dy, dz, nx, ny, nz, ipad):
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib().or_periodic_grid(structures(sts), len(sts), dx, dy, dz, nx, ny, nz, ipad,
                                ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return (a.value, b.value, c.value), rc


def covariance(sts, nugget, hx, hy, hz):
    arr = structures(sts)
    rot = np.concatenate([setrot(s)[0].ravel() for s in sts])
    return lib().or_covariance(arr, _p(rot), len(sts), nugget, hx, hy, hz)


def twiddles(n, isign):
    tw = np.zeros(2 * max(n // 2, 1))
    lib().or_twiddles(n, isign, _p(tw))
    return tw[0::2] + 1j * tw[1::2]


def fft1d(x, isign):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    n = x.size
    tw = np.zeros(2 * max(n // 2, 1))
    lib().or_twiddles(n, isign, _p(tw))
    buf = x.view(np.float64)
    lib().or_fft1d(_p(buf), n, _p(tw))
    return x


def fft3d(a, isign):
    """a: complex array indexed [k, j, i] (C order == Fortran a(i,j,k))."""
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nz, ny, nx = a.shape
    lib().or_fft3d(_p(a.view(np.float64)), nx, ny, nz, isign)
    return a


def spectrum(sts, nugget, dx, dy, dz, npx, npy, npz):
    n = npx * npy * npz
    amp = np.zeros(n)
    sre = np.zeros(n)
    sim = np.zeros(n)
    d = SpecDiag()
    lib().or_spectrum(structures(sts), len(sts), nugget, dx, dy, dz, npx, npy, npz, _p(amp), _p(sre), _p(sim),
                      ctypes.byref(d))
    shp = (npz, npy, npx)
    diag = dict(max_imag=d.max_imag, max_real=d.max_real, c0=d.c0, spos=d.spos, sneg=d.sneg)
    return amp.reshape(shp), (sre + 1j * sim).reshape(shp), diag


def seed_state(iseed):
    return lib().or_seed_state(iseed)


def draw_index(i, j, k, npx, npy, npz, igauss):
    return lib().or_point_draw_index(i, j, k, npx, npy, npz, igauss)


def fill(amp, igauss, state):
    """One realization's serial fill; returns (c, new_state, draw_log)."""
    npz, npy, npx = amp.shape
    amp = np.ascontiguousarray(amp, dtype=np.float64)
    c = np.zeros(amp.shape, dtype=np.complex128)
    st = ctypes.c_int64(state)
    log = np.zeros(amp.size, dtype=np.int64)
    lib().or_fill(_p(amp), _p(c.view(np.float64)), npx, npy, npz, igauss, ctypes.byref(st),
                  log.ctypes.data_as(_i64p))
    return c, st.value, log.reshape(amp.shape)


def simulate(sts, nugget, dx, dy, dz, nx, ny, nz, ipad, iseed, nreal, igauss):
    """Whole unconditional run; returns dict(res [nreal,nz,ny,nx], mean, var, maximag, diag, np)."""
    res = np.zeros(nreal * nx * ny * nz)
    mean = np.zeros(nreal)
    var = np.zeros(nreal)
    mi = np.zeros(nreal)
    d = SpecDiag()
    npo = (ctypes.c_int * 3)()
    rc = lib().or_simulate(structures(sts), len(sts), nugget, dx, dy, dz, nx, ny, nz, ipad, iseed, nreal, igauss,
                           _p(res), _p(mean), _p(var), _p(mi), ctypes.byref(d), npo)
    if rc != 0:
        raise ValueError("periodic grid too large (>2^27 nodes)")
    return dict(res=res.reshape(nreal, nz, ny, nx), mean=mean, var=var, maximag=mi,
                diag=dict(max_imag=d.max_imag, max_real=d.max_real, c0=d.c0, spos=d.spos, sneg=d.sneg),
                np=tuple(npo))


def npair_count(nx, ny, nz, lx, ly, lz):
    return lib().or_npair_count(nx, ny, nz, lx, ly, lz)


def vario_lag(f, lx, ly, lz):
    """f: one field, shape (nz, ny, nx) (x fastest in memory)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, ny, nx = f.shape
    return lib().or_vario_lag(_p(f), nx, ny, nz, lx, ly, lz)


# ---- conditioning (src/specsim3d.f90:548-1083)
def gcum(x):
    return lib().or_gcum(float(x))


def gauinv(p):
    return lib().or_gauinv(float(p))


def heapsort_idx(key):
    """1-based permutation that sorts key ascending, in the reference's heapsort order."""
    key = np.ascontiguousarray(key, dtype=np.float64)
    idx = (ctypes.c_int * max(key.size, 1))(*range(1, key.size + 1))
    lib().or_heapsort_idx(_p(key), idx, key.size)
    return np.array(idx[:key.size])


def nscore(v, w):
    v = np.ascontiguousarray(v, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    g, vr, gt = np.zeros(v.size), np.zeros(v.size), np.zeros(v.size)
    lib().or_nscore(v.size, _p(v), _p(w), _p(g), _p(vr), _p(gt))
    return g, vr, gt


def cond_params(**kw):
    p = CondParams()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def condition(res, origin, spacing, sts, nugget, params, rows):
    """Conditions res (nreal, nz, ny, nx) FP64 in place; returns a dict with the used data nodes,
    their values, Gaussian values, the normal-score table and the counts the reference prints."""
    nreal, nz, ny, nx = res.shape
    assert res.dtype == np.float64 and res.flags.c_contiguous
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    nrow, ncol = rows.shape
    info = CondInfo()
    n = max(nrow, 1)
    ix, iy, iz = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_int * n)()
    dval, dg, vrt, gtt = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib().or_condition(_p(res), nx, ny, nz, nreal, *map(float, origin), *map(float, spacing), structures(sts),
                            len(sts), float(nugget), ctypes.byref(params), _p(rows), nrow, ncol, ctypes.byref(info),
                            ix, iy, iz, _p(dval), _p(dg), _p(vrt), _p(gtt))
    if rc != 0:
        raise ValueError(f"or_condition failed ({rc})")
    nd = info.nd
    return dict(rc=rc, info={f[0]: getattr(info, f[0]) for f in CondInfo._fields_},
                ijk=np.array([ix[:nd], iy[:nd], iz[:nd]]).T, dval=dval[:nd], dg=dg[:nd], vrt=vrt[:info.nt],
                gtt=gtt[:info.nt])
