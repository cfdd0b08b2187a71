"""Spectral densities as files (ss3d_spectrum_save / ss3d_spectrum_load): a loaded spectrum has the
saved amplitudes and diagnostics bit for bit, and its FP64 and FP32 realizations are identical to
those of the spectrum that was saved (and to the oracle's within the FP64 tolerance); a file made for
another model, spacing or grid is refused, and so is a payload that no longer matches its
diagnostics."""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as O
import specsim3d_amd as S

pytestmark = pytest.mark.gpu

STS = [(1, 0.6, 12, 8, 5, 40, 10, 0), (2, 0.3, 25, 25, 10, 0, 0, 0)]
NUGGET = 0.1
SPC = (1.0, 1.0, 2.0)
GRID = (30, 20, 9)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_save_load_round_trip(tmp_path):
    say("[specfile] round trip")
    npg = S.periodic_grid(STS, *SPC, *GRID, 1)
    path = str(tmp_path / "s.spec")
    with S.Engine(0) as eng:
        sp = eng.spectrum(STS, NUGGET, *SPC, *npg)
        sp.save(path)
        ld = eng.load_spectrum(path, STS, NUGGET, *SPC, npg)
        assert ld.grid == tuple(npg) and ld.diag == sp.diag
        assert np.array_equal(ld.amplitude(), sp.amplitude())
        for prec in (S.FP64, S.FP32):
            a = eng.realize(sp, 777, 1, 3, 4, *GRID, precision=prec)
            b = eng.realize(ld, 777, 1, 3, 4, *GRID, precision=prec)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        anon = eng.load_spectrum(path)  # no key given: the file's
        assert anon.grid == tuple(npg)
        assert np.array_equal(eng.realize(anon, 5, 0, 0, 1, *GRID)[0], eng.realize(sp, 5, 0, 0, 1, *GRID)[0])
    o = O.simulate(STS, NUGGET, *SPC, *GRID, 1, 777, 7, 1)["res"][3:]
    with S.Engine(0) as eng:  # a fresh context loads the file made by another one
        ld = eng.load_spectrum(path, STS, NUGGET, *SPC, npg)
        f = eng.realize(ld, 777, 1, 3, 4, *GRID, precision=S.FP64)[0]
    assert np.abs(f - o).max() / np.abs(o).max() < 1e-10


def test_load_refuses_other_model_and_corruption(tmp_path):
    say("[specfile] refusals")
    npg = S.periodic_grid(STS, *SPC, *GRID, 1)
    path = str(tmp_path / "s.spec")
    with S.Engine(0) as eng:
        eng.spectrum(STS, NUGGET, *SPC, *npg).save(path)
        other = [(1, 0.6, 12, 8, 5, 41, 10, 0), STS[1]]
        for args in [(other, NUGGET, *SPC, npg), (STS, 0.2, *SPC, npg), (STS, NUGGET, 1.0, 1.0, 1.0, npg),
                     (STS, NUGGET, *SPC, (npg[0] * 2, npg[1], npg[2]))]:
            with pytest.raises(S.SS3DError) as e:
                eng.load_spectrum(path, *args)
            assert e.value.code == 1 and "another model" in str(e.value)
        raw = bytearray(open(path, "rb").read())
        bad = str(tmp_path / "bad.spec")
        # exponent bits of amp64[n - 13] (a low frequency, so a large amplitude): header 64 bytes,
        # structures 64 each, diagnostics 40, sizes 16
        n64 = int(np.prod(npg))
        i = 64 + 64 * len(STS) + 40 + 16 + 8 * (n64 - 13) + 6
        raw[i] ^= 0x40
        open(bad, "wb").write(bytes(raw))
        with pytest.raises(S.SS3DError) as e:
            eng.load_spectrum(bad, STS, NUGGET, *SPC, npg)
        assert "diagnostics" in str(e.value)
        open(bad, "wb").write(bytes(raw[:1000]))
        with pytest.raises(S.SS3DError) as e:
            eng.load_spectrum(bad)
        assert e.value.code in (1, 6)
        with pytest.raises(S.SS3DError) as e:
            eng.load_spectrum(str(tmp_path / "missing.spec"))
        assert e.value.code == 6
