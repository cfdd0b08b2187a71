"""specsim3d_amd -- MI355X engine for SPECSIM3D's unconditional spectral simulation.

The product is native: libspecsim3d.so (HIP kernels for gfx950 behind the C ABI of
include/specsim3d.h) and the Fortran driver bin/specsim3d that keeps the reference program's
command line, parameter file, stdout and output files.  This module is the Python binding of
the same C ABI (ctypes), used by tests, bench.py and Python callers.  It has no fallback:
every call runs the HIP library, and importing a missing library raises.

Reference: src/specsim3d.f90 of stebond/specsim3d (the hot block f90:1234-1345).
"""
import ctypes
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libspecsim3d.so")
EXE_PATH = os.path.join(PKG, "bin", "specsim3d")

FP64 = 0
FP32 = 1
OPT_PIPELINE = 1
OPT_FP64_GENERIC = 2
OPT_FUSE = 3
OPT_BATCH_AUTOTUNE = 4
OPT_ZRUN = 5
OPT_SPEC_FP32 = 6
SCRATCH_BYTES = 256 << 20  # SS3D_SCRATCH_BYTES: the scratch array of ss3d_realize_device_noalloc

ERRORS = {1: "invalid argument", 2: "too big", 3: "out of memory", 4: "HIP error", 5: "no device", 6: "I/O error"}


class SS3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"specsim3d error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Structure(ctypes.Structure):
    """ss3d_structure: the scalar members of the reference's `type structure` (f90:95-100)."""
    _fields_ = [("itype", ctypes.c_int32), ("pad_", ctypes.c_int32), ("sill", ctypes.c_double),
                ("ahmax", ctypes.c_double), ("ahmin", ctypes.c_double), ("avert", ctypes.c_double),
                ("az", ctypes.c_double), ("dip", ctypes.c_double), ("plg", ctypes.c_double)]


class Request(ctypes.Structure):
    """ss3d_request: one simulation request of the request loop."""
    _fields_ = [("id", ctypes.c_int64), ("nst", ctypes.c_int32), ("iseed", ctypes.c_int32),
                ("st", Structure * 10), ("nugget", ctypes.c_double), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("ipad", ctypes.c_int32), ("igauss", ctypes.c_int32),
                ("nreal", ctypes.c_int32), ("precision", ctypes.c_int32), ("spot_check", ctypes.c_int32),
                ("host_out", ctypes.c_void_p)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int32), ("cache_hit", ctypes.c_int32),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("spot_rel_rms", ctypes.c_double), ("mean0", ctypes.c_double), ("var0", ctypes.c_double),
                ("out_bytes", ctypes.c_int64)]


class SpecDiag(ctypes.Structure):
    _fields_ = [("max_imag", ctypes.c_double), ("max_real", ctypes.c_double), ("c0", ctypes.c_double),
                ("spos", ctypes.c_double), ("sneg", ctypes.c_double)]


class Outputs(ctypes.Structure):
    """ss3d_outputs: the files of one run (the reference's writers, f90:307-391, 1357-1414)."""
    _fields_ = [("bin_path", ctypes.c_char_p), ("vti_path", ctypes.c_char_p), ("vtk_base", ctypes.c_char_p),
                ("vtk_count", ctypes.c_int32), ("reserved", ctypes.c_int32), ("gslib_path", ctypes.c_char_p),
                ("vtk_ascii_path", ctypes.c_char_p), ("origin", ctypes.c_double * 3),
                ("spacing", ctypes.c_double * 3)]


class CondParams(ctypes.Structure):
    """ss3d_cond_params: search ellipsoid, neighbours and back transform (src/specsim3d.f90:731-747)."""
    _fields_ = [("rmax", ctypes.c_double), ("rmed", ctypes.c_double), ("rvert", ctypes.c_double),
                ("sa1", ctypes.c_double), ("sa2", ctypes.c_double), ("sa3", ctypes.c_double),
                ("ndmax", ctypes.c_int32), ("itrans", ctypes.c_int32), ("skmean", ctypes.c_double),
                ("nt", ctypes.c_int32), ("ltail", ctypes.c_int32), ("utail", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("vr", ctypes.c_void_p), ("gt", ctypes.c_void_p),
                ("zlo", ctypes.c_double), ("zhi", ctypes.c_double), ("ltpar", ctypes.c_double),
                ("utpar", ctypes.c_double)]


class CondColumns(ctypes.Structure):
    """ss3d_cond_columns: the data-file part of the conditioning block (src/specsim3d.f90:732-734)."""
    _fields_ = [(n, ctypes.c_int32) for n in ("ixc", "iyc", "izc", "ivc", "iwc", "itrans")] + \
               [("tmin", ctypes.c_double), ("tmax", ctypes.c_double), ("skmean", ctypes.c_double)]


class CondInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("nsbx", "nsby", "nsbz", "nnone", "nexact", "nfail")] + \
               [("nsum", ctypes.c_int64), ("weights_s", ctypes.c_double), ("apply_s", ctypes.c_double),
                ("backtr_s", ctypes.c_double)]


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# exported symbols and their signatures (must match include/specsim3d.h)
SIGNATURES = {
    "ss3d_version": (ctypes.c_int, []),
    "ss3d_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64]),
    "ss3d_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "ss3d_context_destroy": (ctypes.c_int, [_vp]),
    "ss3d_context_synchronize": (ctypes.c_int, [_vp]),
    "ss3d_periodic_grid": (ctypes.c_int, [ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _ip, _ip, _ip]),
    "ss3d_spectrum_create": (ctypes.c_int, [_vp, ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                            ctypes.POINTER(SpecDiag)]),
    "ss3d_spectrum_destroy": (ctypes.c_int, [_vp]),
    "ss3d_spectrum_bytes": (ctypes.c_int64, [_vp]),
    "ss3d_spectrum_amplitude": (ctypes.c_int, [_vp, _dp]),
    "ss3d_draws_per_realization": (ctypes.c_int64, [ctypes.c_int] * 4),
    "ss3d_realization_state": (ctypes.c_int64, [ctypes.c_int] * 5 + [ctypes.c_int64]),
    "ss3d_realize": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _dp, _dp, _dp]),
    "ss3d_realize_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "ss3d_realize_device_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "ss3d_context_profile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ss3d_autotune": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ss3d_tuning_get": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ss3d_tuning_set": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ss3d_tuning_get_tile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip]),
    "ss3d_tuning_set_tile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ss3d_context_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ss3d_server_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "ss3d_server_submit": (ctypes.c_int, [_vp, ctypes.c_void_p]),
    "ss3d_server_poll": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_double]),
    "ss3d_server_info": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int64)] * 4),
    "ss3d_server_clock": (ctypes.c_double, [_vp]),
    "ss3d_server_destroy": (ctypes.c_int, [_vp]),
    "ss3d_batch_create": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ss3d_batch_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ss3d_batch_output": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ss3d_batch_cells": (ctypes.c_int64, [_vp]),
    "ss3d_batch_destroy": (ctypes.c_int, [_vp]),
    "ss3d_variogram": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ss3d_npairs": (ctypes.c_int64, [ctypes.c_int] * 6),
    "ss3d_covariance": (ctypes.c_double, [ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double]),
    "ss3d_context_profile_read": (ctypes.c_int, [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]),
    "ss3d_writer_open": (ctypes.c_int, [_vp, ctypes.POINTER(Outputs), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ss3d_writer_slot": (ctypes.c_int, [_vp]),
    "ss3d_writer_realize": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                           ctypes.POINTER(_vp)]),
    "ss3d_writer_put": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int]),
    "ss3d_writer_close": (ctypes.c_int, [_vp]),
    "ss3d_spectrum_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "ss3d_spectrum_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(Structure), ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                          ctypes.POINTER(SpecDiag)]),
    "ss3d_cond_prepare": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(CondColumns)] +
                          [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_int32)] * 3 +
                          [ctypes.c_void_p] * 5),
    "ss3d_cond_create": (ctypes.c_int, [_vp, ctypes.POINTER(Structure), ctypes.c_int, ctypes.c_double] +
                         [ctypes.c_int] * 3 + [ctypes.c_double] * 6 +
                         [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CondParams),
                          ctypes.POINTER(_vp), ctypes.POINTER(CondInfo)]),
    "ss3d_cond_apply": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "ss3d_cond_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(CondInfo)]),
    "ss3d_cond_destroy": (ctypes.c_int, [_vp]),
    "ss3d_writer_condition": (ctypes.c_int, [_vp, _vp]),
    "ss3d_format_es": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
}

NPASS_KINDS = 10
PASS_KINDS = ["spectrum_x_fp64", "spectrum_y_fp64", "spectrum_z_fp64", "real_x_fp64", "real_y_fp64", "real_z_fp64",
              "real_z_fp32", "real_y_fp32", "real_x_c2r_fp32", "real_zy_fp32"]


def lib():
    """Load libspecsim3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libspecsim3d.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().ss3d_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != 0:
        raise SS3DError(rc, last_error())


def structures(sts):
    """sts: iterable of (itype, sill, ahmax, ahmin, avert, azimuth, dip, plunge)."""
    sts = list(sts)
    arr = (Structure * max(len(sts), 1))()
    for i, s in enumerate(sts):
        arr[i].itype = int(s[0])
        (arr[i].sill, arr[i].ahmax, arr[i].ahmin, arr[i].avert, arr[i].az, arr[i].dip,
         arr[i].plg) = map(float, s[1:8])
    return arr, len(sts)


def periodic_grid(sts, dx, dy, dz, nx, ny, nz, ipad):
    arr, n = structures(sts)
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().ss3d_periodic_grid(arr, n, dx, dy, dz, nx, ny, nz, ipad, ctypes.byref(a), ctypes.byref(b),
                                   ctypes.byref(c)))
    return a.value, b.value, c.value


def draws_per_realization(npx, npy, npz, igauss):
    return lib().ss3d_draws_per_realization(npx, npy, npz, igauss)


def realization_state(iseed, npx, npy, npz, igauss, r):
    return lib().ss3d_realization_state(iseed, npx, npy, npz, igauss, r)


class Spectrum:
    """Device-resident spectral density of one (model, periodic grid) (f90:1256-1297)."""

    def __init__(self, engine, handle, grid, diag):
        self.engine = engine
        self.handle = handle
        self.grid = grid
        self.diag = diag

    def amplitude(self):
        npx, npy, npz = self.grid
        out = np.empty(npx * npy * npz)
        check(lib().ss3d_spectrum_amplitude(self.handle, out.ctypes.data_as(_dp)))
        return out.reshape(npz, npy, npx)

    def nbytes(self):
        return lib().ss3d_spectrum_bytes(self.handle)

    def save(self, path):
        """Write the spectrum, its diagnostics and its model key to a file (ss3d_spectrum_save)."""
        check(lib().ss3d_spectrum_save(self.handle, os.fsencode(path)))

    def close(self):
        if self.handle:
            lib().ss3d_spectrum_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One HIP context (device, stream, twiddle tables, work buffers)."""

    def __init__(self, device=0):
        self.handle = _vp()
        self._deps = weakref.WeakSet()  # spectra, servers, batches, writers: closed before the context
        check(lib().ss3d_context_create(device, ctypes.byref(self.handle)))

    def spectrum(self, sts, nugget, dx, dy, dz, npx, npy, npz):
        arr, n = structures(sts)
        h = _vp()
        d = SpecDiag()
        check(lib().ss3d_spectrum_create(self.handle, arr, n, nugget, dx, dy, dz, npx, npy, npz, ctypes.byref(h),
                                         ctypes.byref(d)))
        diag = dict(max_imag=d.max_imag, max_real=d.max_real, c0=d.c0, spos=d.spos, sneg=d.sneg)
        sp = Spectrum(self, h, (npx, npy, npz), diag)
        self._deps.add(sp)
        return sp

    def load_spectrum(self, path, sts=None, nugget=0.0, dx=1.0, dy=1.0, dz=1.0, grid=(1, 1, 1)):
        """Spectrum from a file of Spectrum.save; with sts given, the file must match that model, spacing
        and periodic grid."""
        h = _vp()
        d = SpecDiag()
        if sts is None:
            arr, n = None, 0
        else:
            arr, n = structures(sts)
        check(lib().ss3d_spectrum_load(self.handle, os.fsencode(path), arr, n, nugget, dx, dy, dz, *grid,
                                       ctypes.byref(h), ctypes.byref(d)))
        with open(path, "rb") as fh:  # header: 8-byte magic, then int32 version, nst, npx, npy, npz, hxa
            grid = tuple(int(v) for v in np.frombuffer(fh.read(32)[8:], dtype="<i4")[2:5])
        diag = dict(max_imag=d.max_imag, max_real=d.max_real, c0=d.c0, spos=d.spos, sneg=d.sneg)
        sp = Spectrum(self, h, grid, diag)
        self._deps.add(sp)
        return sp

    def conditioning(self, sts, nugget, grid, origin, spacing, ijk, g, rmax, rmed, rvert, sa=(0.0, 0.0, 0.0),
                     ndmax=16, itrans=0, skmean=0.0, table=None, zlo=0.0, zhi=0.0, ltail=1, ltpar=1.0, utail=1,
                     utpar=1.0):
        """Kriging weights of every node for conditioning on data (ss3d_cond_create).  ijk: (nd, 3)
        1-based nodes; g: Gaussian values; table = (vr, gt) for the back transform (itrans = 1)."""
        return Conditioning(self, sts, nugget, grid, origin, spacing, ijk, g, rmax, rmed, rvert, sa, ndmax, itrans,
                            skmean, table, zlo, zhi, ltail, ltpar, utail, utpar)

    def realize(self, spec, iseed, igauss, r0, nreal, nx, ny, nz, precision=FP64):
        """Realizations r0..r0+nreal-1 -> (field[nreal, nz, ny, nx], mean, var, maximag)."""
        dt = np.float64 if precision == FP64 else np.float32
        out = np.empty((nreal, nz, ny, nx), dtype=dt)
        mean, var, mi = np.empty(nreal), np.empty(nreal), np.empty(nreal)
        check(lib().ss3d_realize(self.handle, spec.handle, iseed, igauss, r0, nreal, nx, ny, nz, precision,
                                 out.ctypes.data_as(_vp), mean.ctypes.data_as(_dp), var.ctypes.data_as(_dp),
                                 mi.ctypes.data_as(_dp)))
        return out, mean, var, mi

    def realize_device(self, spec, iseed, igauss, r0, nreal, nx, ny, nz, precision, dev_out, dev_stats=None):
        """Asynchronous variant: dev_out / dev_stats are device addresses (ints)."""
        check(lib().ss3d_realize_device(self.handle, spec.handle, iseed, igauss, r0, nreal, nx, ny, nz, precision,
                                        _vp(dev_out), _vp(dev_stats) if dev_stats else None))

    def realize_device_noalloc(self, spec, iseed, igauss, r0, nreal, nx, ny, nz, precision, dev_out, dev_stats=None):
        """realize_device with no device allocation: work buffers in the library's fixed scratch
        array (SCRATCH_BYTES per GPU); grids whose one realization needs more raise SS3DError."""
        check(lib().ss3d_realize_device_noalloc(self.handle, spec.handle, iseed, igauss, r0, nreal, nx, ny, nz,
                                                precision, _vp(dev_out), _vp(dev_stats) if dev_stats else None))

    def synchronize(self):
        check(lib().ss3d_context_synchronize(self.handle))

    def variogram(self, field, dirs, device_shape=None, precision=None):
        """Experimental semivariograms (ss3d_variogram).  field: numpy array (nreal, nz, ny, nx) of
        float64 or float32 in host memory, or a device address with device_shape=(nreal, nz, ny, nx)
        and precision given.  dirs: [(sx, sy, sz, nlag), ...].  Returns gam[r, d, l-1]."""
        dirs = [tuple(int(v) for v in d) for d in dirs]
        darr = np.array(dirs, dtype=np.int32).reshape(-1, 4)
        nlagmax = max(1, int(darr[:, 3].max()))
        if device_shape is None:
            field = np.ascontiguousarray(field)
            if field.ndim == 3:
                field = field[None]
            prec = FP64 if field.dtype == np.float64 else FP32
            if field.dtype not in (np.float64, np.float32):
                raise TypeError("field must be float64 or float32")
            nreal, nz, ny, nx = field.shape
            ptr, on_dev = field.ctypes.data, 0
        else:
            nreal, nz, ny, nx = device_shape
            prec, ptr, on_dev = precision, field, 1
        gam = np.zeros((nreal, len(dirs), nlagmax))
        check(lib().ss3d_variogram(self.handle, _vp(ptr), on_dev, prec, nx, ny, nz, nreal, len(dirs),
                                   darr.ctypes.data_as(_vp), gam.ctypes.data_as(_vp)))
        return gam

    def autotune(self, spec, nx, ny, nz, igauss, reps=3):
        """Pick the fastest radix plan of each FP32 pass for this grid; returns {kind: rm}."""
        check(lib().ss3d_autotune(self.handle, spec.handle, nx, ny, nz, igauss, reps))
        npx, npy, npz = spec.grid
        lx, ly, lz = (int(v).bit_length() - 1 for v in (npx, npy, npz))
        kinds = {"real_z_fp32": (6, lz), "real_y_fp32": (7, ly), "real_x_c2r_fp32": (8, max(lx - 1, 0))}
        return {k: lib().ss3d_tuning_get(self.handle, *v) for k, v in kinds.items()}

    def set_option(self, option, value):
        check(lib().ss3d_context_set_option(self.handle, option, value))

    def tuning_set(self, kind, logn, rm):
        check(lib().ss3d_tuning_set(self.handle, kind, logn, rm))

    def tuning_get_tile(self, kind, logn):
        """dict(rm, variant, tile, lds_pad) of a tuning entry (ss3d_tuning_get_tile)."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().ss3d_tuning_get_tile(self.handle, kind, logn, *[ctypes.byref(x) for x in v]))
        return dict(rm=v[0].value, variant=v[1].value, tile=v[2].value, lds_pad=bool(v[3].value))

    def tuning_set_tile(self, kind, logn, rm, variant):
        check(lib().ss3d_tuning_set_tile(self.handle, kind, logn, rm, variant))

    def tuning_get(self, kind, logn):
        """Radix plan (log2 of the largest radix) of FP32 pass kind 6..8 at length 2^logn."""
        return lib().ss3d_tuning_get(self.handle, kind, logn)

    def profile(self, enable):
        """Start (clears) or stop per-pass HIP-event timing on the context stream."""
        check(lib().ss3d_context_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """{kind: (ms, algorithmic bytes, launches)} for kinds with launches."""
        ms = (ctypes.c_double * NPASS_KINDS)()
        by = (ctypes.c_double * NPASS_KINDS)()
        ln = (ctypes.c_int64 * NPASS_KINDS)()
        check(lib().ss3d_context_profile_read(self.handle, ms, by, ln, NPASS_KINDS))
        return {PASS_KINDS[k]: (ms[k], by[k], ln[k]) for k in range(NPASS_KINDS) if ln[k]}

    def close(self):
        if self.handle:
            deps = list(self._deps)
            for d in [d for d in deps if not isinstance(d, Spectrum)] + [d for d in deps if isinstance(d, Spectrum)]:
                try:
                    d.close()
                except SS3DError:
                    pass
            lib().ss3d_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def cond_prepare(rows, grid, origin, spacing, ixc, iyc, izc, ivc, iwc=0, tmin=-1e21, tmax=1e21, itrans=0,
                 skmean=0.0):
    """The reference's data preparation (ss3d_cond_prepare): data rows -> dict(ijk (nd, 3) 1-based,
    dval, dg, vr, gt (itrans = 1), nd, nmerged, nout)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    nrow, ncol = rows.shape
    n = max(nrow, 1)
    ijk = np.zeros((n, 3), dtype=np.int32)
    dval, dg, vr, gt = (np.zeros(n) for _ in range(4))
    nd, nm, no = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    cc = CondColumns(ixc, iyc, izc, ivc, iwc, itrans, tmin, tmax, skmean)
    check(lib().ss3d_cond_prepare(rows.ctypes.data, nrow, ncol, ctypes.byref(cc), *grid, *origin, *spacing,
                                  ctypes.byref(nd), ctypes.byref(nm), ctypes.byref(no), ijk.ctypes.data,
                                  dval.ctypes.data, dg.ctypes.data, vr.ctypes.data, gt.ctypes.data))
    k = nd.value
    return dict(ijk=ijk[:k], dval=dval[:k], dg=dg[:k], vr=vr[:k] if itrans == 1 else None,
                gt=gt[:k] if itrans == 1 else None, nd=k, nmerged=nm.value, nout=no.value)


class Conditioning:
    """Conditioning of realizations on data (ss3d_cond_*): see Engine.conditioning."""

    def __init__(self, engine, sts, nugget, grid, origin, spacing, ijk, g, rmax, rmed, rvert, sa, ndmax, itrans,
                 skmean, table, zlo, zhi, ltail, ltpar, utail, utpar):
        self.engine = engine
        self.grid = tuple(grid)
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        g = np.ascontiguousarray(g, dtype=np.float64)
        self._table = None
        p = CondParams(rmax=rmax, rmed=rmed, rvert=rvert, sa1=sa[0], sa2=sa[1], sa3=sa[2], ndmax=ndmax,
                       itrans=itrans, skmean=skmean, ltail=ltail, utail=utail, zlo=zlo, zhi=zhi, ltpar=ltpar,
                       utpar=utpar)
        if itrans == 1:
            vr, gt = (np.ascontiguousarray(t, dtype=np.float64) for t in table)
            self._table = (vr, gt)
            p.nt, p.vr, p.gt = vr.size, vr.ctypes.data, gt.ctypes.data
        arr, n = structures(sts)
        self.handle = _vp()
        info = CondInfo()
        check(lib().ss3d_cond_create(engine.handle, arr, n, nugget, *grid, *origin, *spacing, len(g),
                                     ijk.ctypes.data, g.ctypes.data, ctypes.byref(p), ctypes.byref(self.handle),
                                     ctypes.byref(info)))
        engine._deps.add(self)

    def apply(self, dev_field, precision, nreal):
        """Conditions nreal realizations at device address dev_field in place (context stream)."""
        check(lib().ss3d_cond_apply(self.handle, _vp(dev_field), precision, nreal))

    def info(self):
        i = CondInfo()
        check(lib().ss3d_cond_info_get(self.handle, ctypes.byref(i)))
        return {f[0]: getattr(i, f[0]) for f in CondInfo._fields_}

    def close(self):
        if self.handle:
            lib().ss3d_cond_destroy(self.handle)
            self.handle = None


class Writer:
    """Streaming writer of the reference's output files (ss3d_writer_*).

    paths: dict with any of bin, vti, vtk_base (+ vtk_count), gslib, vtk_ascii; origin and
    spacing as in the parameter file.  engine=None gives a writer fed only by put()."""

    def __init__(self, engine, paths, nx, ny, nz, nreal, precision=FP64, origin=(0.0, 0.0, 0.0),
                 spacing=(1.0, 1.0, 1.0)):
        enc = lambda k: os.fsencode(paths[k]) if paths.get(k) else None
        self._o = Outputs(enc("bin"), enc("vti"), enc("vtk_base"), int(paths.get("vtk_count", 0)), 0, enc("gslib"),
                          enc("vtk_ascii"), (ctypes.c_double * 3)(*origin), (ctypes.c_double * 3)(*spacing))
        self.shape = (nx, ny, nz)
        self.nreal = nreal
        self.precision = precision
        self.handle = _vp()
        check(lib().ss3d_writer_open(engine.handle if engine else None, ctypes.byref(self._o), nx, ny, nz, nreal,
                                     precision, ctypes.byref(self.handle)))
        self.slot = lib().ss3d_writer_slot(self.handle)
        if engine:
            engine._deps.add(self)

    def realize(self, spec, iseed, igauss, count):
        """Next `count` realizations on the device; returns (mean, var, maximag, device address)."""
        mean, var, mi = np.empty(count), np.empty(count), np.empty(count)
        dev = _vp()
        check(lib().ss3d_writer_realize(self.handle, spec.handle, iseed, igauss, count, mean.ctypes.data_as(_dp),
                                        var.ctypes.data_as(_dp), mi.ctypes.data_as(_dp), ctypes.byref(dev)))
        return mean, var, mi, dev.value

    def condition(self, cond):
        """Condition every later realize() chunk with a Conditioning before it is written."""
        check(lib().ss3d_writer_condition(self.handle, cond.handle if cond else None))
        self._cond = cond

    def put(self, field):
        """Next realizations from a host array (nreal, nz, ny, nx) of the writer's precision."""
        dt = np.float64 if self.precision == FP64 else np.float32
        field = np.ascontiguousarray(field, dtype=dt)
        nx, ny, nz = self.shape
        n = field.size // (nx * ny * nz)
        check(lib().ss3d_writer_put(self.handle, field.ctypes.data_as(_vp), n))

    def close(self):
        if self.handle:
            h, self.handle = self.handle, None
            check(lib().ss3d_writer_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if exc[0] is None:
            self.close()
        elif self.handle:
            h, self.handle = self.handle, None
            lib().ss3d_writer_close(h)


def format_es(values, w, d):
    """Fortran ESw.d of each value, as the library's writers print them -> list of str."""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    buf = ctypes.create_string_buffer(max(1, v.size * w))
    check(lib().ss3d_format_es(v.ctypes.data_as(_dp), v.size, w, d, buf))
    raw = buf.raw[: v.size * w].decode("ascii")
    return [raw[i * w:(i + 1) * w] for i in range(v.size)]


def simulate(engine, sts, nugget, dx, dy, dz, nx, ny, nz, ipad, iseed, nreal, igauss, precision=FP64):
    """Whole unconditional run as the reference's main block (f90:1234-1345)."""
    grid = periodic_grid(sts, dx, dy, dz, nx, ny, nz, ipad)
    spec = engine.spectrum(sts, nugget, dx, dy, dz, *grid)
    try:
        res, mean, var, mi = engine.realize(spec, iseed, igauss, 0, nreal, nx, ny, nz, precision)
    finally:
        diag = spec.diag
        spec.close()
    return dict(res=res, mean=mean, var=var, maximag=mi, diag=diag, np=grid)


class Server:
    """Request loop over an Engine: continuous batching of simulation requests on `nslots` lanes,
    admission by HBM footprint, spectral-density cache (ss3d_server_* in include/specsim3d.h)."""

    def __init__(self, engine, nslots=4, hbm_budget=0):
        self.engine = engine
        self.handle = _vp()
        check(lib().ss3d_server_create(engine.handle, nslots, hbm_budget, ctypes.byref(self.handle)))
        engine._deps.add(self)
        self._keep = {}

    @staticmethod
    def request(rid, sts, nugget, dx, dy, dz, nx, ny, nz, ipad, iseed, nreal, igauss, precision=FP32,
                spot_check=False, host_out=None):
        r = Request()
        r.id = rid
        r.nst = len(sts)
        for i, t in enumerate(sts):
            r.st[i].itype = int(t[0])
            (r.st[i].sill, r.st[i].ahmax, r.st[i].ahmin, r.st[i].avert, r.st[i].az, r.st[i].dip,
             r.st[i].plg) = map(float, t[1:8])
        r.iseed, r.nugget, r.dx, r.dy, r.dz = iseed, nugget, dx, dy, dz
        r.nx, r.ny, r.nz, r.ipad, r.igauss, r.nreal = nx, ny, nz, ipad, igauss, nreal
        r.precision, r.spot_check = precision, 1 if spot_check else 0
        r.host_out = host_out.ctypes.data if host_out is not None else None
        return r

    def submit(self, req, keep=None):
        """Queue a request; `keep` (e.g. its host output array) stays referenced until completion."""
        if keep is not None:
            self._keep[req.id] = keep
        check(lib().ss3d_server_submit(self.handle, ctypes.byref(req)))

    def poll(self, max_n=64, timeout=0.1):
        buf = (Completion * max_n)()
        n = ctypes.c_int()
        check(lib().ss3d_server_poll(self.handle, buf, max_n, ctypes.byref(n), timeout))
        out = []
        for i in range(n.value):
            c = buf[i]
            self._keep.pop(c.id, None)
            out.append({f: getattr(c, f) for f, _ in Completion._fields_})
        return out

    def info(self):
        v = [ctypes.c_int64() for _ in range(4)]
        check(lib().ss3d_server_info(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(queued=v[0].value, running=v[1].value, cached=v[2].value, cache_bytes=v[3].value)

    def clock(self):
        return lib().ss3d_server_clock(self.handle)

    def close(self):
        if self.handle:
            lib().ss3d_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many FP32 jobs of uneven grids captured into one HIP graph (ss3d_batch_* in
    include/specsim3d.h).  `requests` are Server.request records; host_out is ignored."""

    def __init__(self, engine, requests, nlanes=4):
        self.engine = engine
        self.requests = list(requests)
        arr = (Request * len(self.requests))(*self.requests)
        self.handle = _vp()
        check(lib().ss3d_batch_create(engine.handle, arr, len(self.requests), nlanes, ctypes.byref(self.handle)))
        engine._deps.add(self)
        self.cells = lib().ss3d_batch_cells(self.handle)

    def run(self, graph=True):
        """Enqueue the whole batch (one graph launch, or every pass again); returns at once."""
        check(lib().ss3d_batch_run(self.handle, 1 if graph else 0))

    def output(self, i):
        """(field[r, z, y, x] float32, stats[r, 3]) of job i, after the work has ended."""
        r = self.requests[i]
        f = np.empty((r.nreal, r.nz, r.ny, r.nx), dtype=np.float32)
        st = np.empty((r.nreal, 3), dtype=np.float64)
        check(lib().ss3d_batch_output(self.handle, i, f.ctypes.data, st.ctypes.data))
        return f, st

    def close(self):
        if self.handle:
            lib().ss3d_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def npairs(nx, ny, nz, lx, ly, lz):
    """Node pairs at lag (lx, ly, lz) (ss3d_npairs)."""
    return lib().ss3d_npairs(nx, ny, nz, lx, ly, lz)


def covariance(sts, nugget, hx, hy, hz):
    """Model covariance C(h) (ss3d_covariance); the model variogram is C(0) - C(h)."""
    arr, n = structures(sts)
    return lib().ss3d_covariance(arr, n, nugget, hx, hy, hz)
