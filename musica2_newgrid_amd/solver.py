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


def make_case_struct(case, precision: int = _lib.PREC_FP64, topology=None, launch_flags: int = 0):
    """C-ABI case description.  `topology` is a block table (see
    multiblock.py); None means one block with periodic faces."""
    if topology is None:
        faces = tuple(getattr(case, "faces", (0,) * 6))
        # the case's face codes: 0 periodic, 2 wall, 3 characteristic outflow, 4 supersonic inlet,
        # 5 wall with the slip flag
        code = {2: _lib.FACE_WALL, 3: _lib.FACE_CHARAC, 4: _lib.FACE_INLET, 5: _lib.FACE_SLIP}
        kinds = [code.get(k, _lib.FACE_PERIODIC) for k in faces]
        topology = [{"nx": case.nx, "ny": case.ny, "nz": case.nz, "kind": kinds,
                     "nbr_block": [0] * 6, "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}]
    blocks = (MgpuBlock * len(topology))()
    for b, d in enumerate(topology):
        blk = blocks[b]
        blk.nx, blk.ny, blk.nz = d["nx"], d["ny"], d["nz"]
        kinds = d.get("kind", [_lib.FACE_INTERFACE] * 6)
        for f in range(6):
            blk.face_kind[f] = kinds[f]
            blk.nbr_block[f] = d["nbr_block"][f]
            blk.nbr_face[f] = d["nbr_face"][f]
            blk.orient[f] = d["orient"][f]
    cs = MgpuCase()
    cs.nblocks = len(topology)
    cs.curvilinear = int(case.curvilinear)
    cs.iorder_visc = int(case.iorder_visc)
    cs.filter = int(case.filter)
    cs.is_ducros = int(case.ducros)
    cs.nrk = int(case.nrk)
    cs.precision = int(precision)
    cs.launch_flags = int(launch_flags)
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
    cs.is_adiab = int(getattr(case, "adiabatic", True))
    cs.is_sw_edoh = int(getattr(case, "sw_edoh", False))
    cs.t_wall = float(getattr(case, "t_wall", 0.0))
    cs.p_ref = float(getattr(case, "p_ref", 0.0))
    cs.is_wall2 = int(getattr(case, "wall2", False))
    return cs, blocks


def slip_face(case, kind: int) -> bool:
    """A face of kind `kind` (MGPU_FACE_*) is a slip wall: with is_wall2, every
    wall of an inviscid case and the flagged walls of an adiabatic one
    (assign_procedures.f90:415-470)."""
    if not getattr(case, "wall2", False) or kind not in (_lib.FACE_WALL, _lib.FACE_SLIP):
        return False
    if int(case.iorder_visc) == 0:
        return True
    return bool(getattr(case, "adiabatic", True)) and kind == _lib.FACE_SLIP


class GpuSolver:
    """Owns one mgpu handle (device state, metrics, scratch) for a case.

    One block: ``GpuSolver(case)``.  Several blocks: ``GpuSolver(case,
    topology=blocks, block_cases=cases)``, where `case` carries the flags and
    block_cases[b] the extents, metrics and initial state of block b."""

    def __init__(self, case, precision: int = _lib.PREC_FP64, topology=None, block_cases=None,
                 eager: bool = False, stage_kernel: str = None):
        """stage_kernel: None (the autotuner picks), "classic" or "pipelined"
        (the stage kernel family, MGPU_LAUNCH_STAGE_*)."""
        L = lib()
        self.case = case
        self.block_cases = [case] if block_cases is None else list(block_cases)
        if topology is not None and len(topology) != len(self.block_cases):
            raise ValueError("topology and block_cases differ in length")
        flags = _lib.LAUNCH_EAGER if eager else 0
        flags |= {None: 0, "classic": _lib.LAUNCH_STAGE_CLASSIC, "pipelined": _lib.LAUNCH_STAGE_PIPELINED}[stage_kernel]
        self._cs, self._blocks = make_case_struct(case, precision, topology, flags)
        h = C.c_void_p()
        rc = L.mgpu_create(C.byref(self._cs), C.byref(h))
        check(rc, None, "mgpu_create")
        self.h = h
        self._topology = topology
        for b, bc in enumerate(self.block_cases):
            self._set_metrics(b, bc)
            if self._has_wall(b) and int(case.curvilinear) == _lib.GEOM_C3:
                # curvilinear walls: the per-point coefficients of each wall face
                for f, slabs in bc.wall_c3.items():
                    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in slabs]
                    tab = (_lib._dp * 8)(*[dptr(a) for a in arrs[:8]])
                    check(L.mgpu_set_wall_c3(self.h, b, int(f), tab), self.h, "mgpu_set_wall_c3")
                    if slip_face(case, self._blocks[b].face_kind[int(f)]):
                        # a slip face: the dual tangents too
                        tab = (_lib._dp * 6)(*[dptr(a) for a in arrs[8:14]])
                        check(L.mgpu_set_wall_c3_slip(self.h, b, int(f), tab), self.h, "mgpu_set_wall_c3_slip")
            elif self._has_wall(b):
                c2 = np.ascontiguousarray(getattr(bc, "cextrp2", (0.0,) * 6), dtype=np.float64)
                c3 = np.ascontiguousarray(getattr(bc, "cextrp3", (0.0,) * 6), dtype=np.float64)
                check(L.mgpu_set_wall(self.h, b, dptr(c2), dptr(c3)), self.h, "mgpu_set_wall")
            self.put_state(bc.U, b)

    def _has_wall(self, b: int) -> bool:
        return any(self._blocks[b].face_kind[f] in (_lib.FACE_WALL, _lib.FACE_SLIP) for f in range(6))

    @property
    def nblocks(self) -> int:
        return len(self.block_cases)

    def _set_metrics(self, b, c):
        L = lib()
        if int(c.curvilinear) == _lib.GEOM_C2:
            m2 = [np.ascontiguousarray(a, dtype=np.float64) for a in c.m2]
            m2v = [np.ascontiguousarray(a, dtype=np.float64) for a in (c.m2v if c.m2v is not None else c.m2)]
            tab = (_lib._dp * 5)(*[dptr(a) for a in m2])
            tabv = (_lib._dp * 5)(*[dptr(a) for a in m2v])
            idz = np.ascontiguousarray(c.cart[2], dtype=np.float64)
            idzv = np.ascontiguousarray(c.cart[5], dtype=np.float64)
            check(L.mgpu_set_metrics_c2(self.h, b, tab, tabv, dptr(idz), dptr(idzv)), self.h, "mgpu_set_metrics_c2")
        elif c.curvilinear:
            m = [np.ascontiguousarray(a) for a in c.m]
            tab = (_lib._dp * 10)(*[dptr(a) for a in m])
            tabv = None
            if c.mv is not None:
                mv = [np.ascontiguousarray(a) for a in c.mv]
                tabv = (_lib._dp * 10)(*[dptr(a) for a in mv])
            check(L.mgpu_set_metrics_c3(self.h, b, tab, tabv), self.h, "mgpu_set_metrics_c3")
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in c.cart]
            check(L.mgpu_set_metrics_cart(self.h, b, *[dptr(a) for a in arrs]), self.h,
                  "mgpu_set_metrics_cart")

    def put_state(self, U: Sequence[np.ndarray], block: int = 0) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in U]
        check(lib().mgpu_put_state(self.h, block, *[dptr(a) for a in arrs]), self.h, "mgpu_put_state")

    def get_state(self, block: int = 0) -> List[np.ndarray]:
        c = self.block_cases[block]
        out = [np.empty((c.nz + 10, c.ny + 10, c.nx + 10)) for _ in range(5)]
        check(lib().mgpu_get_state(self.h, block, *[dptr(a) for a in out]), self.h, "mgpu_get_state")
        return out

    def get_states(self) -> List[List[np.ndarray]]:
        return [self.get_state(b) for b in range(self.nblocks)]

    def step(self, nsteps: int = 1) -> None:
        check(lib().mgpu_step(self.h, int(nsteps)), self.h, "mgpu_step")

    def autotune(self, force: bool = False) -> None:
        check(lib().mgpu_autotune(self.h, int(force)), self.h, "mgpu_autotune")

    def stage_config(self, variant: int = 0, block: int = 0):
        tx, ty, kc = C.c_int(), C.c_int(), C.c_int()
        check(lib().mgpu_stage_config(self.h, block, variant, C.byref(tx), C.byref(ty), C.byref(kc)), self.h,
              "mgpu_stage_config")
        return tx.value, ty.value, kc.value

    def stage_pipelined(self, variant: int = 0) -> bool:
        """Whether stage variant 0 (Euler stages) / 1 (last stage) runs the pipelined kernel."""
        r = lib().mgpu_stage_pipelined(self.h, variant)
        check(min(r, 0), self.h, "mgpu_stage_pipelined")
        return bool(r)

    def diagnostics(self) -> dict:
        """Sums of rho, 1/2 rho|u|^2, 1/2 rho|curl u|^2 over all interior cells
        (stats_tgv) and max(|u| + c) (timestep_new)."""
        d = _lib.MgpuDiag()
        check(lib().mgpu_diagnostics(self.h, C.byref(d)), self.h, "mgpu_diagnostics")
        return {"mass": d.mass, "kinetic": d.kinetic, "enstrophy": d.enstrophy, "umax": d.umax,
                "cells": d.cells}

    def max_cfl(self):
        """Maximum CFL number along i, j, k of the current state (check_max_cfl)."""
        out = np.zeros(3)
        check(lib().mgpu_max_cfl(self.h, dptr(out)), self.h, "mgpu_max_cfl")
        return out

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
