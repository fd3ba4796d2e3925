"""Python front end of the C-ABI: one :class:`GpuSolver` per handle.

Mirrors the reference's solver surface for the hot path -- set up the flow
state and metrics once, then call the time step repeatedly
(src/Compute_flow/solver.f90:394-535 calling runge_kutta) -- with the step
running in libmusica_gpu.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib
from ._lib import MgpuBlock, MgpuCase, check, dptr, lib


def make_case_struct(case, precision: int = _lib.PREC_FP64):
    blk = MgpuBlock()
    blk.nx, blk.ny, blk.nz = case.nx, case.ny, case.nz
    for f in range(6):
        blk.face_kind[f] = _lib.FACE_PERIODIC
        blk.nbr_block[f] = 0
        blk.nbr_face[f] = f ^ 1
        blk.orient[f] = 0
    blocks = (MgpuBlock * 1)(blk)
    cs = MgpuCase()
    cs.nblocks = 1
    cs.curvilinear = int(case.curvilinear)
    cs.iorder_visc = int(case.iorder_visc)
    cs.filter = int(case.filter)
    cs.is_ducros = int(case.ducros)
    cs.nrk = int(case.nrk)
    cs.precision = int(precision)
    cs.blocks = C.cast(blocks, C.POINTER(MgpuBlock))
    for q, v in enumerate(case.crk):
        cs.crk[q] = v
    cs.deltat = case.dt
    cs.gam, cs.rg = case.gam, case.rg
    cs.c0, cs.S0 = case.c0, case.S0
    cs.diffscale = case.diffscale
    cs.cok_ratio = case.cok_ratio
    cs.dissip_coeff = case.chi
    cs.dissip_shock = case.dissip_shock
    cs.csens = case.csens
    return cs, blocks


class GpuSolver:
    """Owns one mgpu handle (device state, metrics, scratch) for a case."""

    def __init__(self, case, precision: int = _lib.PREC_FP64):
        L = lib()
        self.case = case
        self._cs, self._blocks = make_case_struct(case, precision)
        h = C.c_void_p()
        rc = L.mgpu_create(C.byref(self._cs), C.byref(h))
        check(rc, None, "mgpu_create")
        self.h = h
        self._set_metrics()
        self.put_state(case.U)

    def _set_metrics(self):
        L, c = lib(), self.case
        if c.curvilinear:
            self._m = [np.ascontiguousarray(a) for a in c.m]
            tab = (_lib._dp * 10)(*[dptr(a) for a in self._m])
            if c.mv is not None:
                self._mv = [np.ascontiguousarray(a) for a in c.mv]
                tabv = (_lib._dp * 10)(*[dptr(a) for a in self._mv])
            else:
                tabv = None
            check(L.mgpu_set_metrics_c3(self.h, 0, tab, tabv), self.h, "mgpu_set_metrics_c3")
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in c.cart]
            check(L.mgpu_set_metrics_cart(self.h, 0, *[dptr(a) for a in arrs]), self.h,
                  "mgpu_set_metrics_cart")

    def put_state(self, U: Sequence[np.ndarray]) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in U]
        check(lib().mgpu_put_state(self.h, 0, *[dptr(a) for a in arrs]), self.h, "mgpu_put_state")

    def get_state(self) -> List[np.ndarray]:
        c = self.case
        out = [np.empty((c.nz + 10, c.ny + 10, c.nx + 10)) for _ in range(5)]
        check(lib().mgpu_get_state(self.h, 0, *[dptr(a) for a in out]), self.h, "mgpu_get_state")
        return out

    def step(self, nsteps: int = 1) -> None:
        check(lib().mgpu_step(self.h, int(nsteps)), self.h, "mgpu_step")

    def autotune(self, force: bool = False) -> None:
        check(lib().mgpu_autotune(self.h, int(force)), self.h, "mgpu_autotune")

    def stage_config(self, variant: int = 0):
        tx, ty, kc = C.c_int(), C.c_int(), C.c_int()
        check(lib().mgpu_stage_config(self.h, 0, variant, C.byref(tx), C.byref(ty), C.byref(kc)), self.h,
              "mgpu_stage_config")
        return tx.value, ty.value, kc.value

    def sync(self) -> None:
        check(lib().mgpu_sync(self.h), self.h, "mgpu_sync")

    def profile(self, enable: bool) -> None:
        check(lib().mgpu_profile(self.h, int(enable)), self.h, "mgpu_profile")

    def kernel_time(self, kind: int):
        n = C.c_int64()
        ms = C.c_double()
        check(lib().mgpu_kernel_time(self.h, kind, C.byref(n), C.byref(ms)), self.h, "mgpu_kernel_time")
        return int(n.value), float(ms.value)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().mgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
