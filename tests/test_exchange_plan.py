"""The library's interface exchange, checked on the host against the oracle's
face copies (the reference's face types, mod_mpi_types_two_sided.f90:47-197,
restated by orc_fill_interfaces) on the caller's arrays.

A block with walls on its i- or k-faces is stored with that axis exchanged
with j, one with walls on its j- and k-faces with its axes cycled (j, k, i
stored as i, j, k), and its momentum, stress and spectral-radius components
with them.
mgpu_exchange_plan returns what the halo copies write, mapped back to the
caller's frame: every ghost value of every array must come from the element
and the array the oracle's brute-force face copy names (the same quantity in
the caller's frame: the caller's axes are the physical axes on a Cartesian
case).  Covered: two-block rings with walls on any axis, or on two axes, of either
block, all reversal codes; a c3 ring whose blocks' axes are signed permutations of each
other (all eight orientation codes), where the spectral radius along an
index axis comes from the neighbour's index axis it meets; c3 rings of blocks
walled on their i- or k-faces.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases, multiblock as mb
from musica2_newgrid_amd.solver import make_case_struct
from oracle import oracle as orc

G = 5
BIG = 1 << 40


def plan(case, blocks, set_):
    cs, keep = make_case_struct(case, topology=blocks)
    L = _lib.lib()
    n = C.c_int64()
    rc = L.mgpu_exchange_plan(C.byref(cs), set_, None, 0, C.byref(n))
    assert rc == 0, L.mgpu_last_error(None)
    m = np.zeros((n.value, 6), dtype=np.int64)
    rc = L.mgpu_exchange_plan(C.byref(cs), set_, m.ctypes.data_as(C.POINTER(C.c_int64)), n.value, C.byref(n))
    assert rc == 0
    return m


def oracle_sources(blocks, ng):
    """{(block, element): (src block, src element)} of the ghost elements the
    oracle's face copy writes within ng layers of the faces (periodic faces are
    interfaces of a block with itself)."""
    obl = []
    for b, d in enumerate(blocks):
        kinds = d.get("kind", [_lib.FACE_INTERFACE] * 6)
        e = dict(nx=d["nx"], ny=d["ny"], nz=d["nz"], nbr_block=list(d["nbr_block"]), nbr_face=list(d["nbr_face"]),
                 orient=list(d["orient"]), wall=[int(k == _lib.FACE_WALL) for k in kinds])
        for f in range(6):
            if kinds[f] == _lib.FACE_PERIODIC:
                e["nbr_block"][f], e["nbr_face"][f], e["orient"][f] = b, f ^ 1, 0
        obl.append(e)
    arrs = []
    for b, d in enumerate(obl):
        n = (d["nx"] + 2 * G) * (d["ny"] + 2 * G) * (d["nz"] + 2 * G)
        arrs.append(b * BIG + np.arange(n, dtype=np.float64))
    before = [a.copy() for a in arrs]
    orc.fill_interfaces(obl, arrs)
    out = {}
    for b, d in enumerate(obl):
        ext = (d["nz"] + 2 * G, d["ny"] + 2 * G, d["nx"] + 2 * G)
        k, j, i = np.unravel_index(np.arange(arrs[b].size), ext)
        # layer depth beyond the interior along each axis
        depth = np.maximum.reduce([np.maximum(G - i, i - (G + d["nx"] - 1)),
                                   np.maximum(G - j, j - (G + d["ny"] - 1)),
                                   np.maximum(G - k, k - (G + d["nz"] - 1))])
        changed = np.nonzero((arrs[b] != before[b]) & (depth <= ng))[0]
        for e in changed:
            v = int(arrs[b][e])
            out[(b, int(e))] = (v // BIG, v % BIG)
    return out


def check_plan(case, blocks, set_, radius_axes=None):
    ng = G if set_ == 0 else 2
    m = plan(case, blocks, set_)
    want = oracle_sources(blocks, ng)
    nvar = 5 if set_ == 0 else 14
    assert len(m) == len(want) * nvar, (len(m), len(want), nvar)
    got = {}
    for row in m:
        db, de, dv, sb, se, sv = (int(x) for x in row)
        key = (db, de)
        assert want.get(key) == (sb, se), (row, want.get(key))
        if set_ == 1 and dv >= 11 and radius_axes is not None:
            assert sv == 11 + radius_axes(db, de, dv - 11, sb), row
        else:
            assert sv == dv, row
        got.setdefault(key, set()).add(dv)
    assert set(got) == set(want)
    assert all(len(s) == nvar for s in got.values())


def ring(wa, wb, x, o):
    """Blocks A (10x11x12) and B joined along axis x in a ring (A's max face to
    B's min face and back, orientation code o); A walled on axis wa, B on wb
    (None: periodic; a tuple: walls on each axis in it), the other axes
    periodic."""
    ext_a = [10, 11, 12]
    ext_b = list(ext_a)
    ext_b[x] = 13
    out = []
    for b, (ext, w) in enumerate(((ext_a, wa), (ext_b, wb))):
        kind = [_lib.FACE_PERIODIC] * 6
        nbr_block, nbr_face, orient = [b] * 6, [f ^ 1 for f in range(6)], [0] * 6
        for a in (() if w is None else w if isinstance(w, tuple) else (w,)):
            kind[2 * a] = kind[2 * a + 1] = _lib.FACE_WALL
        for s in range(2):
            f = 2 * x + s
            kind[f] = _lib.FACE_INTERFACE
            nbr_block[f], nbr_face[f], orient[f] = 1 - b, f ^ 1, o
        out.append({"nx": ext[0], "ny": ext[1], "nz": ext[2], "kind": kind, "nbr_block": nbr_block,
                    "nbr_face": nbr_face, "orient": orient})
    return out


RINGS = [(wa, wb, x) for wa in (None, 0, 1, 2) for wb in (None, 0, 1, 2) for x in range(3)
         if x != wa and x != wb]
# walls on two axes: j and k (axes cycled) against every other relabeling, i
# and j (none), i and k (j and k exchanged) against i alone (i and j exchanged)
RINGS += [((1, 2), wb, 0) for wb in (None, 1, 2, (1, 2))]
RINGS += [((0, 1), 0, 2), ((0, 2), 0, 1), ((0, 2), (0, 2), 1)]


@pytest.mark.parametrize("wa,wb,x", RINGS)
@pytest.mark.parametrize("set_", [0, 1])
def test_walled_rings_match_the_oracle_face_copies(wa, wb, x, set_):
    c = cases.tgv(12)
    for o in (0, 2, 4, 6):
        check_plan(c, ring(wa, wb, x, o), set_)


@pytest.mark.parametrize("w,x", [(0, 1), (0, 2), (2, 0), (2, 1)])
@pytest.mark.parametrize("set_", [0, 1])
def test_c3_walled_rings_match_the_oracle_face_copies(w, x, set_):
    """Curvilinear blocks walled on their i- or k-faces (stored with that axis
    exchanged with j): momentum and tau keep the caller's x, y, z, only the
    index axes (and the spectral radii along them) move."""
    c = cases.vortex_wavy(12, 12, 12)
    for o in (0, 2, 4, 6):
        check_plan(c, ring(w, w, x, o), set_)


def test_c3_permuted_ring_matches_the_oracle_face_copies():
    from tests.test_multiblock_oracle import C3_ORIENTS, C3_PARENT
    c = cases.vortex_wavy(*C3_PARENT, filter=4)
    blocks, _ = mb.split_oriented(c, len(C3_ORIENTS), C3_ORIENTS)
    for blk in blocks:
        blk["kind"] = [_lib.FACE_INTERFACE] * 6

    def radius_axes(db, de, d, sb):
        # the neighbour's local axis that runs along the same parent axis
        perm_d = C3_ORIENTS[db][0]
        perm_s = C3_ORIENTS[sb][0]
        return list(perm_s).index(perm_d[d])

    check_plan(c, blocks, 0)
    check_plan(c, blocks, 1, radius_axes)


def test_cartesian_interfaces_must_keep_their_axes():
    """Cartesian blocks tie each momentum component and metric to an index
    axis: an interface that exchanges tangential axes or joins faces of
    different axes is refused."""
    c = cases.tgv(12)
    blocks = ring(None, None, 0, 0)
    for b in range(2):
        for s in range(2):
            blocks[b]["orient"][s] = 1
    cs, keep = make_case_struct(c, topology=blocks)
    n = C.c_int64()
    rc = _lib.lib().mgpu_exchange_plan(C.byref(cs), 0, None, 0, C.byref(n))
    assert rc == _lib.EINVAL and b"same axis" in _lib.lib().mgpu_last_error(None)
