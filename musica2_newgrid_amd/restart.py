"""Restart files in the reference's format, so that GPU runs start from and
emit the reference's checkpoints.

* ``restart_bl<N>.bin`` (``restart<stamp>_bl<N>.bin`` with a file stamp): for
  each of rro, rou, rov, row, roe in that order, the block's interior
  nx*ny*nz doubles in Fortran order (i fastest), native byte order, no header
  -- what read_write_volume writes (src/File_IO/mod_io.f90:49-395, variable
  list :216-236) through write_tec's file view (src/File_IO/mod_tecplot.f90:
  91-124, 396-402).  Block numbers are 1-based, written without leading zeros
  (numchar, src/Utils/mod_utils.f90:13-44).
* ``info.ini``: the text of read_write_info (src/File_IO/read_write_info.f90:
  28-45); its reader takes the values after the ``=`` of the labelled lines
  (:80-100).
* ``time.ini``: one line per checkpoint, ``<stamp> = ntotal cputot time``
  (:47-57, read back :120-147).
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence

import numpy as np

from .grid import G

VARS = ("rro", "rou", "rov", "row", "roe")


def restart_name(block: int, filestamp: str = "0000_0000") -> str:
    if filestamp == "0000_0000":
        return f"restart_bl{block}.bin"
    return f"restart{filestamp}_bl{block}.bin"


def write_restart(path: str, U: Sequence[np.ndarray]) -> None:
    """U: 5 padded arrays (nz+10, ny+10, nx+10) -- i is the last numpy axis,
    so C order of the interior is the Fortran order of the file."""
    with open(path, "wb") as f:
        for a in U:
            np.ascontiguousarray(a[G:-G, G:-G, G:-G], dtype=np.float64).tofile(f)


def read_restart(path: str, nx: int, ny: int, nz: int) -> List[np.ndarray]:
    n = nx * ny * nz
    raw = np.fromfile(path, dtype=np.float64)
    if raw.size != 5 * n:
        raise ValueError(f"{path}: {raw.size} doubles, expected 5*{nx}*{ny}*{nz}")
    out = []
    for v in range(5):
        a = np.zeros((nz + 2 * G, ny + 2 * G, nx + 2 * G))
        a[G:-G, G:-G, G:-G] = raw[v * n:(v + 1) * n].reshape(nz, ny, nx)
        out.append(a)
    return out


def _e(x: float) -> str:
    return f" {x:19.12E}"


def write_info(path: str, dims: Sequence[Sequence[int]], is_curv: bool, deltat: float,
               ektot0: float = 0.0, mgtot0: float = 0.0, xmin=(0.0, 0.0, 0.0), xmax=(0.0, 0.0, 0.0),
               mach: float = 0.0, re: float = 0.0, muw_ref: float = 0.0, mu_ref: float = 0.0,
               rho_ref: float = 0.0, p_ref: float = 0.0, T_ref: float = 0.0,
               u_ref: float = 0.0, c_ref: float = 0.0, tscale: float = 0.0) -> None:
    lines = [f"nbloc & is_curv   = {len(dims):5d} {'T' if is_curv else 'F'}"]
    for n, (ni, nj, nk) in enumerate(dims, start=1):
        lines.append(f"NX NY NZ bl{n:04d}   = {ni:5d} {nj:5d} {nk:5d}")
    lines.append("Etot0 mgtot0      =" + _e(ektot0) + _e(mgtot0))
    lines.append("xmin ymin zmin    =" + "".join(_e(x) for x in xmin))
    lines.append("xmax ymax zmax    =" + "".join(_e(x) for x in xmax))
    lines.append("Mref Reref        =" + _e(mach) + _e(re))
    lines.append("Mupref Muref      =" + _e(muw_ref) + _e(mu_ref))
    lines.append("Roref Pref Tref   =" + _e(rho_ref) + _e(p_ref) + _e(T_ref))
    lines.append("Uref cref Tscale  =" + _e(u_ref) + _e(c_ref) + _e(tscale))
    lines.append("time step deltat  =" + _e(deltat))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_info(path: str) -> Dict:
    """Values of info.ini by label (the text before '=')."""
    out: Dict = {"blocks": []}
    with open(path) as f:
        for line in f:
            if "=" not in line:
                continue
            key, val = line.split("=", 1)
            key = key.strip()
            toks = val.split()
            if key == "nbloc & is_curv":
                out["nbloc"] = int(toks[0])
                out["is_curv"] = toks[1].upper().startswith("T")
            elif key.startswith("NX NY NZ bl"):
                out["blocks"].append(tuple(int(t) for t in toks[:3]))
            else:
                out[key] = [float(t) for t in toks]
    out["deltat"] = out.get("time step deltat", [None])[0]
    return out


def append_time(path: str, filestamp: str, ntotal: int, cputot: float, time: float) -> None:
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(f"{'Timestamp'}{' ' * 5}{'Ntotal'}{' ' * 6}{'cputot'}{' ' * 15}{'time'}\n")
        f.write(f"{filestamp} = {ntotal:8d} {cputot!r} {time!r}\n")


def read_last_time(path: str):
    """(filestamp, ntotal, cputot, time) of the last line of time.ini."""
    with open(path) as f:
        last = [ln for ln in f if ln.strip()][-1]
    stamp = last[:9]
    toks = last.split("=", 1)[1].split()
    return stamp, int(toks[0]), float(toks[1]), float(toks[2])
