"""Multiblock set-up: block tables and oriented splits of a periodic block.

A block table is a list of dicts {nx, ny, nz, nbr_block[6], nbr_face[6],
orient[6]}: the connectivity of param_blocks.ini
(src/Main/mod_block.f90:109-229) in the form the C-ABI takes.  Face f
(imin,imax,jmin,jmax,kmin,kmax = 0..5) joins face nbr_face[f] of block
nbr_block[f].  The tangential axes of a face are its two other axes in
increasing order; orient bit 0 swaps them, bits 1 and 2 reverse the
neighbour's first and second tangential axis (the swapped/reversed face
types of src/Parallel/mod_mpi_types_two_sided.f90:331-512 with the
directions of src/Grid/mod_grid_directions.f90:16-40).

:func:`split_oriented` cuts a fully periodic parent block along x into
sub-blocks whose local axes are signed permutations of the parent's.  Every
interface orientation code can be produced that way, and the parent, advanced
as one block, is the closed-form answer for the split case.
"""
from __future__ import annotations

import copy
import itertools
from typing import List, Sequence, Tuple

import numpy as np

from .grid import G


def periodic_block(nx, ny, nz, b=0):
    return {"nx": nx, "ny": ny, "nz": nz, "nbr_block": [b] * 6, "nbr_face": [1, 0, 3, 2, 5, 4],
            "orient": [0] * 6, "kind": [0] * 6}


def face_axes(f: int) -> Tuple[int, int, int]:
    a = f // 2
    t1 = 1 if a == 0 else 0
    t2 = 1 if a == 2 else 2
    return a, t1, t2


def inverse_orient(o: int) -> int:
    swap, r1, r2 = o & 1, (o >> 1) & 1, (o >> 2) & 1
    return swap | ((r2 if swap else r1) << 1) | ((r1 if swap else r2) << 2)


def perm_det(perm: Sequence[int], sign: Sequence[int]) -> int:
    inv = sum(1 for a, b in itertools.combinations(range(3), 2) if perm[a] > perm[b])
    return (-1) ** inv * int(np.prod(sign))


def _wrap_box(f: np.ndarray, x0: int, m: int) -> np.ndarray:
    """Parent-frame padded box (z, y, x) of interior x range [x0, x0+m), full
    y and z, ghosts taken periodically from the parent interior."""
    core = f[G:-G, G:-G, G:-G]
    nz, ny, nx = core.shape
    ix = np.arange(x0 - G, x0 + m + G) % nx
    iy = np.arange(-G, ny + G) % ny
    iz = np.arange(-G, nz + G) % nz
    return np.ascontiguousarray(core[np.ix_(iz, iy, ix)])


def _to_local(box: np.ndarray, perm, sign) -> np.ndarray:
    """Reorder a parent-frame (z, y, x) array into the local frame: local axis
    d runs along parent axis perm[d], reversed when sign[d] < 0."""
    # numpy axis 2-d holds axis d
    order = [2 - perm[2 - a] for a in range(3)]
    out = np.transpose(box, order)
    for a in range(3):
        if sign[2 - a] < 0:
            out = np.flip(out, axis=a)
    return np.ascontiguousarray(out)


def _to_parent(loc: np.ndarray, perm, sign) -> np.ndarray:
    out = loc
    for a in range(3):
        if sign[2 - a] < 0:
            out = np.flip(out, axis=a)
    order = [2 - perm[2 - a] for a in range(3)]
    inv = np.argsort(order)
    return np.ascontiguousarray(np.transpose(out, inv))


def split_oriented(parent, parts: int, orients: Sequence[Tuple[Sequence[int], Sequence[int]]]):
    """Split a fully periodic parent case along x into `parts` blocks.

    orients[b] = (perm, sign): local axis d of block b runs along parent axis
    perm[d], in the parent's direction when sign[d] > 0.  A Cartesian parent
    allows sign changes only (perm = identity).  Returns (blocks, cases): the
    block table and one Case per block (metrics and state in the block's own
    frame)."""
    assert parent.nx % parts == 0
    m = parent.nx // parts
    pext = (m, parent.ny, parent.nz)
    blocks, cases = [], []
    for b in range(parts):
        perm, sign = orients[b]
        if not parent.curvilinear:
            assert tuple(perm) == (0, 1, 2), "a Cartesian block keeps its axes (signs may flip)"
        n = [pext[perm[d]] for d in range(3)]
        c = copy.copy(parent)
        c.nx, c.ny, c.nz = n
        x0 = b * m
        c.U = [_to_local(_wrap_box(a, x0, m), perm, sign) for a in parent.U]
        if parent.curvilinear:
            det = perm_det(perm, sign)

            def tr(ms):
                out = []
                for d in range(3):
                    for comp in range(3):
                        src = ms[3 * perm[d] + comp]
                        out.append(det * sign[d] * _to_local(_wrap_box(src, x0, m), perm, sign))
                out.append(det * _to_local(_wrap_box(ms[9], x0, m), perm, sign))
                return out

            c.m = tr(parent.m)
            c.mv = tr(parent.mv) if parent.mv is not None else None
        else:
            cart = []
            for q in range(6):
                d = q % 3
                a = np.asarray(parent.cart[q])
                npar = a.size - 2 * G
                if d == 0:
                    idx = np.arange(x0 - G, x0 + m + G) % npar + G
                else:
                    idx = np.arange(-G, npar + G) % npar + G
                v = a[idx]
                if sign[d] < 0:
                    v = -v[::-1]
                cart.append(np.ascontiguousarray(v))
            c.cart = cart
        cases.append(c)
        blocks.append({"nx": n[0], "ny": n[1], "nz": n[2], "nbr_block": [0] * 6, "nbr_face": [0] * 6,
                       "orient": [0] * 6, "kind": [1] * 6})
    # connectivity from the geometry: outward normal of local face (d, s) is
    # parent axis perm[d], direction sign[d]*(+1 if s else -1)
    for b in range(parts):
        perm, sign = orients[b]
        for f in range(6):
            d, s = f // 2, f % 2
            pa = perm[d]
            out_dir = sign[d] * (1 if s else -1)
            nb = (b + out_dir) % parts if pa == 0 else b
            nperm, nsign = orients[nb]
            dn = list(nperm).index(pa)
            # neighbour's outward normal is the opposite direction
            sn = 1 if nsign[dn] * (-out_dir) > 0 else 0
            nf = 2 * dn + sn
            a, t1, t2 = face_axes(f)
            an, u1, u2 = face_axes(nf)
            swap = 0 if perm[t1] == nperm[u1] else 1
            src1 = t2 if swap else t1  # our axis feeding the neighbour's u1
            src2 = t1 if swap else t2
            r1 = 1 if sign[src1] != nsign[u1] else 0
            r2 = 1 if sign[src2] != nsign[u2] else 0
            blocks[b]["nbr_block"][f] = nb
            blocks[b]["nbr_face"][f] = nf
            blocks[b]["orient"][f] = swap | (r1 << 1) | (r2 << 2)
    return blocks, cases


def join_oriented(parent, Us: List[List[np.ndarray]], orients) -> List[np.ndarray]:
    """Parent-frame padded arrays (interior filled) from the blocks' states."""
    parts = len(Us)
    m = parent.nx // parts
    out = [np.zeros_like(a) for a in parent.U]
    for b in range(parts):
        perm, sign = orients[b]
        for v in range(5):
            box = _to_parent(Us[b][v], perm, sign)
            out[v][G:-G, G:-G, G + b * m:G + (b + 1) * m] = box[G:-G, G:-G, G:-G]
    return out


def check_table(blocks) -> None:
    """Every interface is declared from both sides with the inverse orientation."""
    for b, blk in enumerate(blocks):
        for f in range(6):
            nb, nf, o = blk["nbr_block"][f], blk["nbr_face"][f], blk["orient"][f]
            other = blocks[nb]
            assert other["nbr_block"][nf] == b and other["nbr_face"][nf] == f, (b, f)
            assert other["orient"][nf] == inverse_orient(o), (b, f, o, other["orient"][nf])


def _random_mapping(lengths, h, seed, amp_cells=0.02, modes=3):
    """Smooth periodic distortion: each coordinate displaced by a sum of
    low-mode sines of the other two, amplitude amp_cells*h, random phases."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = lengths
    tp = 2.0 * np.pi
    terms = []
    for _ in range(3):
        t = []
        for _ in range(modes):
            m1, m2 = rng.integers(1, 3, size=2)
            ph1, ph2 = rng.uniform(0, tp, size=2)
            t.append((m1, m2, ph1, ph2, rng.uniform(0.5, 1.0) / modes))
        terms.append(t)
    a = amp_cells * h

    def f(XI, ET, ZE):
        out = []
        base = (XI, ET, ZE)
        others = ((ET / Ly, ZE / Lz), (XI / Lx, ZE / Lz), (XI / Lx, ET / Ly))
        for c in range(3):
            u, v = others[c]
            d = 0.0
            for m1, m2, ph1, ph2, w in terms[c]:
                d = d + w * np.sin(tp * m1 * u + ph1) * np.sin(tp * m2 * v + ph2)
            out.append(base[c] + a * d)
        return tuple(out)

    return f


def pulse_case(nx, ny, nz, h, seed, mach=0.3, amp=1e-3, width_cells=4.0, filter=1, iorder_visc=4,
               T_inf=300.0, rho_inf=1.177, re_cell=50.0, cfl=0.8):
    """Periodic curvilinear block (random smooth distortion of 2 % of a
    cell) with a freestream at `mach` along x and a Gaussian pressure pulse
    of relative amplitude `amp` and half-width `width_cells` cells at its
    centre (ar = -ln2/(b dx)^2, the pulse of mod_init_flow.f90:2270)."""
    import math

    from . import cases
    from . import grid as gr
    lengths = (nx * h, ny * h, nz * h)
    m, mv, (xc, yc, zc) = gr.curvilinear_metrics_periodic(_random_mapping(lengths, h, seed), nx, ny, nz, lengths,
                                                            viscous=iorder_visc != 0)
    c_inf = math.sqrt(cases.GAM * cases.RG * T_inf)
    u_inf = mach * c_inf
    p_inf = rho_inf * cases.RG * T_inf
    xm, ym, zm = (L / 2.0 for L in lengths)
    ar = -math.log(2.0) / (width_cells * h) ** 2
    p = p_inf * (1.0 + amp * np.exp(ar * ((xc - xm) ** 2 + (yc - ym) ** 2 + (zc - zm) ** 2)))
    rho = p / (cases.RG * T_inf)
    u = np.full_like(p, u_inf)
    v = np.zeros_like(p)
    w = np.zeros_like(p)
    U = cases._to_padded(cases.conserved(rho, u, v, w, p), nx, ny, nz)
    # viscosity from a cell Reynolds number rho u h / mu = re_cell
    diffscale = (rho_inf * u_inf * h / re_cell) / float(cases.sutherland(np.float64(T_inf)))
    cs = cases.Case("pulse", nx, ny, nz, curvilinear=True, iorder_visc=iorder_visc, filter=filter, ducros=True,
                    diffscale=diffscale, m=m, mv=mv, U=U)
    cs.dt = cases.timestep(U, h, cfl)
    cs.info = dict(mach=mach, amp=amp, width_cells=width_cells, seed=seed)
    return cs


# Uneven multiblock workload: 7 blocks of the sizes 384x192x96,
# 2x(256x128x96), 3x(128x128x64), 96^3.  A 1-to-1 interface joins faces of
# equal extent, and only blocks of equal size have such faces here, so the
# blocks form four periodic islands: A alone, B1-B2 (a ring along i, B2
# reversed), C1-C2-C3 (a ring whose blocks' axes are permuted and reversed,
# joining i-faces to j-faces), D alone (axes permuted).
ISLANDS = [
    ((384, 192, 96), 1, [((0, 1, 2), (1, 1, 1))]),
    ((512, 128, 96), 2, [((0, 1, 2), (1, 1, 1)), ((0, 1, 2), (-1, 1, -1))]),
    ((384, 128, 64), 3, [((0, 1, 2), (1, 1, 1)), ((1, 0, 2), (1, -1, 1)), ((1, 0, 2), (-1, 1, -1))]),
    ((96, 96, 96), 1, [((2, 0, 1), (1, -1, 1))]),
]


def uneven_islands(seed: int = 20261202, scale: int = 1, filter: int = 1):
    """The 7-block workload (sizes divided by `scale`), every block with the
    same nominal spacing h and one time step.  Returns (flags case, block
    table, block cases)."""
    h = 0.01 * scale
    blocks, bcases = [], []
    for q, (ext, parts, orients) in enumerate(ISLANDS):
        nx, ny, nz = (e // scale for e in ext)
        parent = pulse_case(nx, ny, nz, h, seed + q, filter=filter)
        tab, bcs = split_oriented(parent, parts, orients)
        off = len(blocks)
        for blk in tab:
            blk["nbr_block"] = [b + off for b in blk["nbr_block"]]
        blocks += tab
        bcases += bcs
    flags = copy.copy(bcases[0])
    flags.dt = min(bc.dt for bc in bcases)
    for bc in bcases:
        bc.dt = flags.dt
    return flags, blocks, bcases
