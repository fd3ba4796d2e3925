"""Multiblock set-up helpers: block tables and splitting a periodic block.

A block table is a list of dicts {nx, ny, nz, nbr_block[6], nbr_face[6],
orient[6]} -- the connectivity of param_blocks.ini
(src/Main/mod_block.f90:109-229) in the form the C-ABI takes: face f
(imin,imax,jmin,jmax,kmin,kmax = 0..5) joins face nbr_face[f] of block
nbr_block[f]; orient bit 0 swaps the tangential axes, bits 1/2 reverse the
first/second tangential axis of the neighbour face.
"""
from __future__ import annotations

import numpy as np

from . import grid as gr
from .grid import G


def periodic_block(nx, ny, nz, b=0):
    return {"nx": nx, "ny": ny, "nz": nz, "nbr_block": [b] * 6, "nbr_face": [1, 0, 3, 2, 5, 4],
            "orient": [0] * 6}


def split_x(nx, ny, nz, parts, reverse=()):
    """Split a fully periodic nx*ny*nz block along i into `parts` equal
    blocks.  Blocks listed in `reverse` have their i axis reversed (i runs
    towards decreasing x).  Returns the block table."""
    assert nx % parts == 0
    m = nx // parts
    rev = set(reverse)
    blocks = []
    for b in range(parts):
        blk = periodic_block(m, ny, nz, b)
        # the physical low-x / high-x sides of block b
        lo_face, hi_face = (1, 0) if b in rev else (0, 1)
        left, right = (b - 1) % parts, (b + 1) % parts
        # neighbour's face touching us: the neighbour's physical high side (left) / low side (right)
        left_face = 0 if left in rev else 1
        right_face = 1 if right in rev else 0
        blk["nbr_block"][lo_face] = left
        blk["nbr_face"][lo_face] = left_face
        blk["nbr_block"][hi_face] = right
        blk["nbr_face"][hi_face] = right_face
        blocks.append(blk)
    return blocks


def split_x_state(U, parts, reverse=()):
    """Sub-block padded arrays (ghosts left empty) of a padded parent state."""
    nz, ny, nx = (s - 2 * G for s in U[0].shape)
    m = nx // parts
    out = []
    for b in range(parts):
        arrs = []
        for a in U:
            sub = gr.padded(m, ny, nz)
            blk = a[:, :, G + b * m:G + (b + 1) * m]
            if b in reverse:
                blk = blk[:, :, ::-1]
            sub[:, :, G:G + m] = blk
            arrs.append(sub)
        out.append(arrs)
    return out


def join_x_state(Us, reverse=()):
    """Inverse of split_x_state for the interior."""
    parts = len(Us)
    nz, ny, m = (s - 2 * G for s in Us[0][0].shape)
    out = []
    for v in range(5):
        a = gr.padded(m * parts, ny, nz)
        for b in range(parts):
            blk = Us[b][v][:, :, G:G + m]
            if b in reverse:
                blk = blk[:, :, ::-1]
            a[:, :, G + b * m:G + (b + 1) * m] = blk
        out.append(a)
    return out
