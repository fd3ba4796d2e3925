"""mgpu_diagnostics_noalloc (partial results in the library's static device
array, no device allocation) against mgpu_diagnostics (partial results in a
buffer the handle allocates), through the C-ABI: two calls in a row, the
second on a smaller block, so that stale partial results of the first call
would show.  Sums within 1e-12 relative (only the order of the additions
differs); the maximum of |u| + c and the cell count exactly."""
import ctypes as C

import pytest

from musica2_newgrid_amd import _lib, cases

pytestmark = pytest.mark.gpu


def _both(case, steps):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(case) as s:
        s.step(steps)
        ref = s.diagnostics()
        d = _lib.MgpuDiag()
        rc = _lib.lib().mgpu_diagnostics_noalloc(s.h, C.byref(d))
        _lib.check(rc, s.h, "mgpu_diagnostics_noalloc")
        got = {"mass": d.mass, "kinetic": d.kinetic, "enstrophy": d.enstrophy, "umax": d.umax, "cells": d.cells}
    return ref, got


def _check(ref, got):
    for k in ("mass", "kinetic", "enstrophy"):
        assert got[k] == pytest.approx(ref[k], rel=1e-12, abs=0), k
    assert got["umax"] == ref["umax"]
    assert got["cells"] == ref["cells"]


def test_noalloc_diagnostics_match_two_calls_second_smaller():
    big = cases.vortex_wavy(48, 32, 24, filter=2)
    small = cases.tgv(16)
    print(f"noalloc diagnostics: c3 {big.nx}x{big.ny}x{big.nz}, then TGV 16^3", flush=True)
    ref1, got1 = _both(big, 2)
    print(f"call 1: {got1} vs {ref1}", flush=True)
    _check(ref1, got1)
    ref2, got2 = _both(small, 2)
    print(f"call 2: {got2} vs {ref2}", flush=True)
    _check(ref2, got2)
    assert got2["cells"] < got1["cells"]


def test_noalloc_diagnostics_walled_channel():
    c = cases.channel(32, 32, 12)
    print(f"noalloc diagnostics: channel {c.nx}x{c.ny}x{c.nz}", flush=True)
    ref, got = _both(c, 3)
    _check(ref, got)
