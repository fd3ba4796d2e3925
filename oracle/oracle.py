"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the results are compared with.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MUSICA_ORACLE_LIB: another build of the same sources (the planted-mistake test only)
LIB_PATH = os.environ.get("MUSICA_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("curv", C.c_int), ("visc", C.c_int), ("filter", C.c_int), ("ducros", C.c_int),
        ("nrk", C.c_int),
        ("face", (C.c_int * 2) * 3),
        ("pad_", C.c_int),
        ("crk", C.c_double * 8),
        ("dt", C.c_double),
        ("gam", C.c_double), ("rg", C.c_double),
        ("c0", C.c_double), ("S0", C.c_double), ("diffscale", C.c_double),
        ("cok_ratio", C.c_double),
        ("chi", C.c_double), ("dissip_shock", C.c_double), ("csens", C.c_double),
        ("adiabatic", C.c_int), ("sw_edoh", C.c_int), ("t_wall", C.c_double),
        ("p_ref", C.c_double), ("wall2", C.c_int), ("pad2_", C.c_int),
    ]


class OrcBlock(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nbr_block", C.c_int * 6), ("nbr_face", C.c_int * 6), ("orient", C.c_int * 6),
        ("wall", C.c_int * 6),
    ]


class OrcGrid(C.Structure):
    _fields_ = [
        ("idx", _dp), ("idy", _dp), ("idz", _dp),
        ("idx_v", _dp), ("idy_v", _dp), ("idz_v", _dp),
        ("m", _dp * 10),
        ("mv", _dp * 10),
        ("cextrp2", C.c_double * 6), ("cextrp3", C.c_double * 6),
        ("m2", _dp * 5),
        ("m2v", _dp * 5),
        ("wc3", (_dp * 14) * 6),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_step.argtypes = [C.POINTER(OrcCfg), C.POINTER(OrcGrid), C.POINTER(_dp), C.c_int]
        L.orc_step.restype = C.c_int
        for f in ("orc_euler_rhs", "orc_visc_rhs"):
            getattr(L, f).argtypes = [C.POINTER(OrcCfg), C.POINTER(OrcGrid), C.POINTER(_dp), C.POINTER(_dp)]
            getattr(L, f).restype = None
        L.orc_filter_rhs.argtypes = [C.POINTER(OrcCfg), C.POINTER(OrcGrid), C.POINTER(_dp),
                                     C.POINTER(_dp), C.POINTER(_dp)]
        L.orc_filter_rhs.restype = None
        L.orc_fill_periodic.argtypes = [C.POINTER(OrcCfg), _dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_set_num_threads.restype = None
        L.orc_step_multi.argtypes = [C.POINTER(OrcCfg), C.c_int, C.POINTER(OrcBlock), C.POINTER(OrcGrid),
                                     C.POINTER(_dp), C.c_int]
        L.orc_step_multi.restype = C.c_int
        L.orc_fill_interfaces.argtypes = [C.c_int, C.POINTER(OrcBlock), C.POINTER(_dp)]
        L.orc_fill_interfaces.restype = None
        L.orc_diag.argtypes = [C.POINTER(OrcCfg), C.POINTER(OrcGrid), C.POINTER(_dp), _dp]
        L.orc_diag.restype = C.c_int
        L.orc_max_cfl.argtypes = [C.POINTER(OrcCfg), C.POINTER(OrcGrid), C.POINTER(_dp), _dp]
        L.orc_max_cfl.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Oracle:
    """CPU restatement of the reference step for one block (periodic faces, or
    walls: then the step runs as a one-block multiblock case)."""

    def __init__(self, case):
        self.case = case
        c = OrcCfg()
        c.nx, c.ny, c.nz = case.nx, case.ny, case.nz
        c.curv = int(case.curvilinear)  # 0 Cartesian, 1 c3, 2 2.5-D
        c.visc = int(case.iorder_visc != 0)
        c.filter = int(case.filter)
        c.ducros = int(case.ducros)
        c.sw_edoh = int(getattr(case, "sw_edoh", False))
        c.nrk = int(case.nrk)
        faces = tuple(getattr(case, "faces", (0,) * 6))
        for d in range(3):
            for s in range(2):
                c.face[d][s] = faces[2 * d + s]
        # boundary faces (wall 2, characteristic outflow 3, supersonic inlet 4,
        # wall with the slip flag 5)
        self.walls = [int(k >= 2) for k in faces]
        c.wall2 = int(getattr(case, "wall2", False))
        c.adiabatic = int(getattr(case, "adiabatic", True))
        c.t_wall = float(getattr(case, "t_wall", 0.0))
        c.p_ref = float(getattr(case, "p_ref", 0.0))
        for q, v in enumerate(case.crk):
            c.crk[q] = v
        c.dt = case.dt
        c.gam, c.rg, c.c0, c.S0 = case.gam, case.rg, case.c0, case.S0
        c.diffscale, c.cok_ratio = case.diffscale, case.cok_ratio
        c.chi, c.dissip_shock, c.csens = case.chi, case.dissip_shock, case.csens
        self.cfg = c
        g = OrcGrid()
        self._keep = []
        if int(case.curvilinear) == 2:
            for q in range(5):
                a = np.ascontiguousarray(case.m2[q], dtype=np.float64)
                b = np.ascontiguousarray(case.m2v[q], dtype=np.float64)
                self._keep += [a, b]
                g.m2[q] = _p(a)
                g.m2v[q] = _p(b)
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in case.cart]
            self._keep += arrs
            g.idx, g.idy, g.idz, g.idx_v, g.idy_v, g.idz_v = [_p(a) for a in arrs]
        elif case.curvilinear:
            for q in range(10):
                a = np.ascontiguousarray(case.m[q])
                self._keep.append(a)
                g.m[q] = _p(a)
                if case.mv is not None:
                    b = np.ascontiguousarray(case.mv[q])
                    self._keep.append(b)
                    g.mv[q] = _p(b)
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in case.cart]
            self._keep += arrs
            g.idx, g.idy, g.idz, g.idx_v, g.idy_v, g.idz_v = [_p(a) for a in arrs]
        for f in range(6):
            g.cextrp2[f] = float(getattr(case, "cextrp2", (0.0,) * 6)[f])
            g.cextrp3[f] = float(getattr(case, "cextrp3", (0.0,) * 6)[f])
        # c3 walls: the per-point coefficients of each wall face (8 slabs, 14
        # with the dual tangents of a slip wall)
        for f, slabs in (getattr(case, "wall_c3", None) or {}).items():
            for q in range(len(slabs)):
                a = np.ascontiguousarray(slabs[q], dtype=np.float64)
                self._keep.append(a)
                g.wc3[f][q] = _p(a)
        self.grid = g

    @staticmethod
    def _tab(arrs):
        return (_dp * 5)(*[_p(a) for a in arrs])

    def step(self, U, nsteps: int = 1):
        """Advance the list of 5 padded arrays U in place."""
        if any(self.walls):
            blk = self_block(self.case)
            rc = lib().orc_step_multi(C.byref(self.cfg), 1, make_blocks([blk]), C.byref(self.grid),
                                      self._tab(U), int(nsteps))
            if rc != 0:
                raise RuntimeError(f"orc_step_multi failed rc={rc}")
            return U
        rc = lib().orc_step(C.byref(self.cfg), C.byref(self.grid), self._tab(U), int(nsteps))
        if rc != 0:
            raise RuntimeError(f"orc_step failed rc={rc}")
        return U

    def euler_rhs(self, U):
        K = [np.zeros_like(U[0]) for _ in range(5)]
        lib().orc_euler_rhs(C.byref(self.cfg), C.byref(self.grid), self._tab(U), self._tab(K))
        return K

    def visc_rhs(self, U, K=None):
        if K is None:
            K = [np.zeros_like(U[0]) for _ in range(5)]
        lib().orc_visc_rhs(C.byref(self.cfg), C.byref(self.grid), self._tab(U), self._tab(K))
        return K

    def diag(self, U):
        """(sum rho, sum 1/2 rho|u|^2, sum 1/2 rho|curl u|^2, max |u|+c) over the interior."""
        out = np.zeros(4)
        rc = lib().orc_diag(C.byref(self.cfg), C.byref(self.grid), self._tab(U), _p(out))
        if rc != 0:
            raise RuntimeError(f"orc_diag failed rc={rc}")
        return out

    def max_cfl(self, U):
        """Maximum CFL number along i, j, k over the interior (check_max_cfl)."""
        out = np.zeros(3)
        rc = lib().orc_max_cfl(C.byref(self.cfg), C.byref(self.grid), self._tab(U), _p(out))
        if rc != 0:
            raise RuntimeError(f"orc_max_cfl failed rc={rc}")
        return out

    def filter_rhs(self, Un, Uref=None):
        K = [np.zeros_like(Un[0]) for _ in range(5)]
        Uref = Un if Uref is None else Uref
        lib().orc_filter_rhs(C.byref(self.cfg), C.byref(self.grid), self._tab(Un), self._tab(Uref),
                             self._tab(K))
        return K


def self_block(case):
    """One-block table: periodic faces are interfaces with the block itself;
    faces of kind 2 (wall), 3 (characteristic outflow), 4 (supersonic inlet),
    5 (wall with the slip flag) are boundaries of that kind."""
    faces = tuple(getattr(case, "faces", (0,) * 6))
    return dict(nx=case.nx, ny=case.ny, nz=case.nz, nbr_block=[0] * 6,
                nbr_face=[f ^ 1 for f in range(6)], orient=[0] * 6,
                wall=[max(0, int(k) - 1) for k in faces])


def make_blocks(blocks):
    """blocks: list of dicts with nx, ny, nz, nbr_block[6], nbr_face[6], orient[6]
    and optionally wall[6]."""
    arr = (OrcBlock * len(blocks))()
    for q, b in enumerate(blocks):
        arr[q].nx, arr[q].ny, arr[q].nz = b["nx"], b["ny"], b["nz"]
        for f in range(6):
            arr[q].nbr_block[f] = b["nbr_block"][f]
            arr[q].nbr_face[f] = b["nbr_face"][f]
            arr[q].orient[f] = b["orient"][f]
            arr[q].wall[f] = int(b.get("wall", [0] * 6)[f])
    return arr


def step_multi(case, blocks, block_cases, Us, nsteps: int = 1):
    """Advance the multiblock state Us (list per block of 5 padded arrays) in
    place.  case gives the flags; block_cases[b] carries block b's metrics."""
    orcs = [Oracle(bc) for bc in block_cases]
    cfg = Oracle(case).cfg
    grids = (OrcGrid * len(blocks))(*[o.grid for o in orcs])
    tab = (_dp * (5 * len(blocks)))(*[_p(a) for U in Us for a in U])
    rc = lib().orc_step_multi(C.byref(cfg), len(blocks), make_blocks(blocks), grids, tab, int(nsteps))
    if rc != 0:
        raise RuntimeError(f"orc_step_multi failed rc={rc}")
    return Us


def fill_interfaces(blocks, arrs):
    tab = (_dp * len(arrs))(*[_p(a) for a in arrs])
    lib().orc_fill_interfaces(len(blocks), make_blocks(blocks), tab)


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))
