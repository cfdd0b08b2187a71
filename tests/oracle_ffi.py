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
