"""Multiblock interface copies of the oracle: a periodic block split into
blocks (aligned and reversed) advances exactly like the single block."""
import numpy as np

from musica2_newgrid_amd import cases, multiblock as mb
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc


def test_single_block_multi_path_equals_single_path():
    c = cases.tgv(16)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = c.copy_state()
    orc.step_multi(c, [mb.periodic_block(16, 16, 16)], [c], [B], 2)
    for a, b in zip(A, B):
        assert np.array_equal(gr.interior(a), gr.interior(b))


def _split_case(c, m, reverse):
    bc = cases.tgv(m)  # carries the metric layout of an m x n x n Cartesian block
    bc.nx, bc.ny, bc.nz = m, c.ny, c.nz
    idx = c.cart[0][0]
    sign = -1.0 if reverse else 1.0
    bc.cart = [np.full(m + 10, sign * idx), np.full(c.ny + 10, idx), np.full(c.nz + 10, idx),
               np.full(m + 10, sign * idx), np.full(c.ny + 10, idx), np.full(c.nz + 10, idx)]
    return bc


def test_split_blocks_equal_single_block():
    n, parts, rev = 24, 3, (1,)
    c = cases.tgv(n)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    blocks = mb.split_x(n, n, n, parts, reverse=rev)
    Us = mb.split_x_state(c.U, parts, reverse=rev)
    for v in range(5):
        orc.fill_interfaces(blocks, [U[v] for U in Us])
    bcs = [_split_case(c, n // parts, b in rev) for b in range(parts)]
    orc.step_multi(c, blocks, bcs, Us, 2)
    B = mb.join_x_state(Us, reverse=rev)
    for q, (a, b) in enumerate(zip(A, B)):
        ai, bi = gr.interior(a), gr.interior(b)
        # reversed blocks sum their i-stencils in mirrored order: round-off only
        assert np.max(np.abs(ai - bi)) <= 1e-12 * np.max(np.abs(ai)) + 1e-12, q
    # aligned split only: bitwise
    blocks = mb.split_x(n, n, n, parts)
    Us = mb.split_x_state(c.U, parts)
    for v in range(5):
        orc.fill_interfaces(blocks, [U[v] for U in Us])
    orc.step_multi(c, blocks, [_split_case(c, n // parts, False) for _ in range(parts)], Us, 2)
    B = mb.join_x_state(Us)
    for a, b in zip(A, B):
        assert np.array_equal(gr.interior(a), gr.interior(b))
