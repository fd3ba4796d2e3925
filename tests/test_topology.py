"""mgpu_create's checks of the block table (they run before any device
call, so on any host): neighbours in range, matching face extents under the
orientation, interfaces declared from both sides with the inverse
orientation."""
import ctypes as C

import pytest

from musica2_newgrid_amd import _lib, cases, multiblock as mb
from musica2_newgrid_amd.solver import make_case_struct


def _create(parent, blocks):
    cs, keep = make_case_struct(parent, topology=blocks)
    h = C.c_void_p()
    rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
    msg = _lib.lib().mgpu_last_error(None)
    return rc, (msg.decode() if isinstance(msg, bytes) else msg)


def _table():
    parent = cases.tgv(30)
    blocks, _ = mb.split_oriented(parent, 3, [((0, 1, 2), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)),
                                              ((0, 1, 2), (1, -1, -1))])
    return parent, blocks


def test_back_link_is_checked():
    parent, blocks = _table()
    blocks[0]["nbr_face"][1] = 0 if blocks[0]["nbr_face"][1] == 1 else 1
    rc, msg = _create(parent, blocks)
    assert rc == _lib.EINVAL and "point back" in msg


def test_orientation_must_be_inverse_on_both_sides():
    parent, blocks = _table()
    blocks[1]["orient"][2] = 2
    rc, msg = _create(parent, blocks)
    assert rc == _lib.EINVAL and "orient" in msg


def test_face_extents_must_match():
    parent, blocks = _table()
    blocks[2]["ny"] = 29
    rc, msg = _create(parent, blocks)
    assert rc == _lib.EINVAL and "extents" in msg


def test_neighbour_out_of_range():
    parent, blocks = _table()
    blocks[0]["nbr_block"][0] = 7
    rc, msg = _create(parent, blocks)
    assert rc == _lib.EINVAL and "out of range" in msg


def test_block_limits_are_checked():
    """Extents below the 5 ghost layers and blocks above 2^28 padded points are
    refused before any device work."""
    parent, blocks = _table()
    blocks[0]["nx"] = 4
    rc, msg = _create(parent, blocks)
    assert rc == _lib.EINVAL and ">= 5" in msg
    big = [{"nx": 1024, "ny": 1024, "nz": 300, "kind": [0] * 6, "nbr_block": [0] * 6,
            "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}]
    rc, msg = _create(parent, big)
    assert rc == _lib.EINVAL and "2^28" in msg


def test_flag_checks():
    import copy
    parent, blocks = _table()
    c = copy.copy(parent)
    c.filter = 2
    c.iorder_visc = 0
    rc, msg = _create(c, blocks)
    assert rc == _lib.EINVAL and "shock filter needs" in msg
    c = copy.copy(parent)
    c.nrk = 9
    rc, msg = _create(c, blocks)
    assert rc == _lib.EINVAL and "nrk" in msg


def _channel_table(**kw):
    c = cases.channel(16, 20, 6, **kw)
    blk = {"nx": c.nx, "ny": c.ny, "nz": c.nz, "kind": [_lib.FACE_PERIODIC, _lib.FACE_PERIODIC, _lib.FACE_WALL,
                                                        _lib.FACE_WALL, _lib.FACE_PERIODIC, _lib.FACE_PERIODIC],
           "nbr_block": [0] * 6, "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}
    return c, [blk]


def _create_case(c, blocks, precision=_lib.PREC_FP64):
    cs, keep = make_case_struct(c, precision=precision, topology=blocks)
    h = C.c_void_p()
    rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
    msg = _lib.lib().mgpu_last_error(None)
    return rc, (msg.decode() if isinstance(msg, bytes) else msg)


def test_wall_limits_are_checked():
    """Walls: the faces of one or two axes per block (Cartesian), Cartesian
    or 3-D curvilinear (not 2.5-D), >= 10 cells between them; refused before
    any device work."""
    c, blocks = _channel_table()
    for f in range(6):
        blocks[0]["kind"][f] = _lib.FACE_WALL
    rc, msg = _create_case(c, blocks)
    assert rc == _lib.EINVAL and "three axes" in msg
    c, blocks = _channel_table(filter=4)
    blocks[0]["kind"][0] = blocks[0]["kind"][1] = _lib.FACE_WALL
    rc, msg = _create_case(c, blocks)
    assert rc == _lib.EINVAL and "artificial viscosities" in msg
    c, blocks = _channel_table()
    c.curvilinear = 2
    rc, msg = _create_case(c, blocks)
    assert rc == _lib.EINVAL and "Cartesian" in msg
    c, blocks = _channel_table()
    blocks[0]["ny"] = 9
    rc, msg = _create_case(c, blocks)
    assert rc == _lib.EINVAL and ">= 10" in msg


def test_artvisc_flags_are_checked():
    """Artificial viscosity: Cartesian or 2.5-D with the viscous terms."""
    c = cases.tgv(16, filter=3, iorder_visc=0)
    rc, msg = _create_case(c, None)
    assert rc == _lib.EINVAL and "sound speed" in msg
    c = cases.tgv(16, filter=3)
    c.curvilinear = True
    rc, msg = _create_case(c, None)
    assert rc == _lib.EINVAL and "Cartesian" in msg


def test_artvisc_shock_flags_are_checked():
    """Artificial viscosity with shock capturing: the viscous terms (any
    geometry, walls allowed); filter codes above 4 are refused."""
    c = cases.tgv(16, filter=4, iorder_visc=0)
    rc, msg = _create_case(c, None)
    assert rc == _lib.EINVAL and "sound speed" in msg
    c = cases.tgv(16, filter=5)
    rc, msg = _create_case(c, None)
    assert rc == _lib.EINVAL and "filter=5" in msg


@pytest.mark.parametrize("which", ["z_cut_x", "x_cut_y", "y_cut_x"])
def test_wall_row_meeting_a_row_that_is_no_wall_is_refused(which):
    """A channel cut in two blocks where only block 0 keeps the walls (block
    1 is periodic on that axis): the wall rows of block 0 face ordinary rows
    of block 1 across the interface.  The viscous exchange carries tau/mu on
    wall rows, so mgpu_create refuses the table with a message (on the walls
    of the k-, i- and j-faces: relabeled or not)."""
    from tests.test_gpu_walls_axes import _channel, _cut_mixed, _turned
    if which == "y_cut_x":
        t, axis, wall_axis = _channel(), 0, 1
    else:
        t = _turned("z" if which == "z_cut_x" else "x")
        axis, wall_axis = (0, 2) if which == "z_cut_x" else (1, 0)
    blocks, bcs = _cut_mixed(t, axis, wall_axis)
    cs, keep = make_case_struct(t, topology=blocks)
    h = C.c_void_p()
    rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
    msg = _lib.lib().mgpu_last_error(None).decode()
    assert rc == _lib.EINVAL, rc
    assert "wall row" in msg, msg


def test_walls_continuing_across_an_interface_are_accepted_by_the_check():
    """The same cut with the walls kept on both blocks passes the wall-row
    check (the create may still need a device after it)."""
    from tests.test_gpu_walls_axes import _channel, _cut_mixed
    t = _channel()
    blocks, bcs = _cut_mixed(t, 0, 1)
    blocks[1]["kind"][2] = blocks[1]["kind"][3] = _lib.FACE_WALL
    blocks[1]["wall"] = [int(k == _lib.FACE_WALL) for k in blocks[1]["kind"]]
    cs, keep = make_case_struct(t, topology=blocks)
    h = C.c_void_p()
    rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
    msg = _lib.lib().mgpu_last_error(None).decode()
    if rc == 0:
        _lib.lib().mgpu_destroy(h)
    assert "wall row" not in msg


def _shock_bl_table(**kw):
    c = cases.shock_bl(24, 20, 6, **kw)
    blk = {"nx": c.nx, "ny": c.ny, "nz": c.nz,
           "kind": [_lib.FACE_INLET, _lib.FACE_CHARAC, _lib.FACE_WALL, _lib.FACE_CHARAC, _lib.FACE_PERIODIC,
                    _lib.FACE_PERIODIC],
           "nbr_block": [0] * 6, "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}
    return c, [blk]


def test_characteristic_faces_are_checked():
    """Characteristic outflow at imax / jmax and the supersonic inlet at imin
    of Cartesian blocks whose other boundary faces are j-walls, with the
    selective or shock filter (or none), the classic stage kernel and >= 10
    cells along a bounded axis; anything else is refused before any device
    work, with a message."""
    def refused(c, blocks, text, flags=0):
        cs, keep = make_case_struct(c, topology=blocks, launch_flags=flags)
        h = C.c_void_p()
        rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
        msg = _lib.lib().mgpu_last_error(None).decode()
        assert rc == _lib.EINVAL, (rc, msg)
        assert text in msg, msg

    c, b = _shock_bl_table()
    b[0]["kind"][0] = _lib.FACE_CHARAC
    refused(c, b, "imax and jmax")
    c, b = _shock_bl_table()
    b[0]["kind"][1] = _lib.FACE_INLET
    refused(c, b, "imin")
    c, b = _shock_bl_table()
    b[0]["kind"][4] = b[0]["kind"][5] = _lib.FACE_WALL
    refused(c, b, "j-faces only")
    c, b = _shock_bl_table()
    c.curvilinear = True
    refused(c, b, "Cartesian")
    c, b = _shock_bl_table()
    c.filter = 4
    refused(c, b, "filter 4")
    c, b = _shock_bl_table()
    refused(c, b, "classic", flags=_lib.LAUNCH_STAGE_PIPELINED)
    c, b = _shock_bl_table()
    b[0]["nx"] = 9
    refused(c, b, ">= 10")


def test_curvilinear_walls_are_checked():
    """Walls on 3-D curvilinear blocks sit on the faces of one axis (the
    kernels close the stored j-faces only, where i- or k-walls are relabeled);
    2.5-D blocks take no walls; refused before any device work."""
    c = cases.vortex_channel(24, 20, 8)
    blk = {"nx": c.nx, "ny": c.ny, "nz": c.nz,
           "kind": [_lib.FACE_WALL, _lib.FACE_WALL, _lib.FACE_WALL, _lib.FACE_WALL, _lib.FACE_PERIODIC,
                    _lib.FACE_PERIODIC],
           "nbr_block": [0] * 6, "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}
    rc, msg = _create_case(c, [blk])
    assert rc == _lib.EINVAL and "faces of one axis" in msg, msg
    c.curvilinear = 2
    blk["kind"] = [_lib.FACE_PERIODIC, _lib.FACE_PERIODIC, _lib.FACE_WALL, _lib.FACE_WALL, _lib.FACE_PERIODIC,
                   _lib.FACE_PERIODIC]
    rc, msg = _create_case(c, [blk])
    assert rc == _lib.EINVAL and "3-D curvilinear" in msg, msg


def test_advanced_wall_flags_are_checked():
    """is_wall2 is 0 or 1, and advanced walls do not meet characteristic or
    inlet faces; refused before any device work."""
    c, b = _shock_bl_table()
    c.wall2 = True
    rc, msg = _create_case(c, b)
    assert rc == _lib.EINVAL and "is_wall2" in msg, msg
    c = cases.channel(24, 20, 8, slip=True)
    cs, keep = make_case_struct(c)
    cs.is_wall2 = 2
    h = C.c_void_p()
    rc = _lib.lib().mgpu_create(C.byref(cs), C.byref(h))
    msg = _lib.lib().mgpu_last_error(None).decode()
    assert rc == _lib.EINVAL and "is_wall2=2" in msg, msg
    # the slip flag is a face kind of its own, accepted where walls are
    blk = {"nx": c.nx, "ny": c.ny, "nz": c.nz,
           "kind": [_lib.FACE_PERIODIC, _lib.FACE_PERIODIC, _lib.FACE_SLIP, _lib.FACE_SLIP, _lib.FACE_SLIP,
                    _lib.FACE_SLIP],
           "nbr_block": [0] * 6, "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}
    rc, msg = _create_case(c, [blk])
    assert rc == _lib.EINVAL and "walls" in msg, msg
