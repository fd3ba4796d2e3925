"""Multiblock interface exchange of the oracle, pinned by invariance: a
periodic block cut into sub-blocks whose axes are signed permutations of the
parent's (every orientation code of the face exchange) advances like the
uncut parent.  A wrong ghost index, a missed reversal or a swapped tangential
axis puts a wrong neighbour value into the stencils and fails these."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases, multiblock as mb
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc

# (perm, sign) per sub-block; the c3 sets produce all eight orientation codes
CART_ORIENTS = [((0, 1, 2), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)), ((0, 1, 2), (1, -1, -1))]
C3_ORIENTS = [((0, 1, 2), (1, 1, 1)), ((2, 1, 0), (-1, 1, 1)), ((2, 1, 0), (1, 1, -1)),
              ((1, 0, 2), (-1, 1, -1)), ((2, 0, 1), (-1, 1, -1)), ((0, 2, 1), (-1, 1, 1)),
              ((1, 2, 0), (1, -1, 1))]
C3_PARENT = (35, 10, 10)


def _rel(A, B):
    return max(float(np.max(np.abs(gr.interior(a) - gr.interior(b)))) / float(np.max(np.abs(gr.interior(a))))
               for a, b in zip(A, B))


def test_single_block_multi_path_equals_single_path():
    c = cases.tgv(16)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = c.copy_state()
    orc.step_multi(c, [mb.periodic_block(16, 16, 16)], [c], [B], 2)
    for a, b in zip(A, B):
        assert np.array_equal(gr.interior(a), gr.interior(b))


def test_orientation_table_is_consistent():
    c = cases.vortex_wavy(*C3_PARENT, filter=2)
    blocks, _ = mb.split_oriented(c, len(C3_ORIENTS), C3_ORIENTS)
    mb.check_table(blocks)
    codes = {blk["orient"][f] for blk in blocks for f in range(6)}
    assert codes == set(range(8)), codes


def test_cartesian_split_with_reversed_axes_equals_parent():
    n, parts = 24, 3
    c = cases.tgv(n)
    A = c.copy_state()
    orc.Oracle(c).step(A, 3)
    blocks, bcs = mb.split_oriented(c, parts, CART_ORIENTS)
    Us = [bc.copy_state() for bc in bcs]
    orc.step_multi(c, blocks, bcs, Us, 3)
    B = mb.join_oriented(c, Us, CART_ORIENTS)
    assert _rel(A, B) < 1e-13


@pytest.mark.parametrize("filt", [1, 2])
def test_c3_split_with_permuted_axes_equals_parent(filt):
    c = cases.vortex_wavy(*C3_PARENT, filter=filt)
    A = c.copy_state()
    orc.Oracle(c).step(A, 3)
    blocks, bcs = mb.split_oriented(c, len(C3_ORIENTS), C3_ORIENTS)
    Us = [bc.copy_state() for bc in bcs]
    orc.step_multi(c, blocks, bcs, Us, 3)
    B = mb.join_oriented(c, Us, C3_ORIENTS)
    # permuted axes sum the three directions in another order: round-off only
    assert _rel(A, B) < 1e-12


def test_uneven_islands_equal_their_parents():
    """The 7-block workload (scaled down 8x): each island advanced as split
    blocks equals its parent advanced as one block."""
    flags, blocks, bcs = mb.uneven_islands(scale=8)
    Us = [bc.copy_state() for bc in bcs]
    orc.step_multi(flags, blocks, bcs, Us, 2)
    first = 0
    for q, (ext, parts, orients) in enumerate(mb.ISLANDS):
        nx, ny, nz = (e // 8 for e in ext)
        parent = mb.pulse_case(nx, ny, nz, 0.08, 20261202 + q, filter=1)
        parent.dt = flags.dt
        A = parent.copy_state()
        orc.Oracle(parent).step(A, 2)
        B = mb.join_oriented(parent, Us[first:first + parts], orients)
        # compare the change over the two steps (the state is a 1e-3 pulse on a uniform field)
        for a, b, a0 in zip(A, B, parent.U):
            da, db = gr.interior(a) - gr.interior(a0), gr.interior(b) - gr.interior(a0)
            assert np.max(np.abs(da - db)) <= 1e-11 * np.max(np.abs(da)) + 1e-14 * np.max(np.abs(gr.interior(a0)))
        first += parts
