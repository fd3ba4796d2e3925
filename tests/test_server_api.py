"""Argument checks of the serving entry points (no device use before the
first tick, so these run on any host)."""
import ctypes as C

import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd.serving import Server


def test_submit_queues_and_take_waits_for_the_result():
    c = cases.tgv(16)
    with Server(1 << 30) as s:
        rid = s.submit(c, 3)
        assert s.status(rid) == (0, 0)
        with pytest.raises(_lib.MgpuError, match="not finished"):
            s.take(rid)
        with pytest.raises(_lib.MgpuError, match="unknown"):
            s.status(rid + 100)


def test_request_must_be_one_periodic_block():
    from musica2_newgrid_amd.solver import make_case_struct
    c = cases.tgv(16)
    topo = [{"nx": 16, "ny": 16, "nz": 16, "kind": [1] * 6, "nbr_block": [0] * 6,
             "nbr_face": [1, 0, 3, 2, 5, 4], "orient": [0] * 6}]
    cs, keep = make_case_struct(c, topology=topo)
    arrs = [np.ascontiguousarray(a) for a in c.cart]
    mt = (_lib._dp * 6)(*[_lib.dptr(a) for a in arrs])
    ut = (_lib._dp * 5)(*[_lib.dptr(a) for a in c.U])
    with Server(1 << 30) as s:
        rid = C.c_int64()
        rc = _lib.lib().mgpu_server_submit(s.h, C.byref(cs), mt, ut, 3, C.byref(rid))
        assert rc == _lib.EUNSUP


def test_request_larger_than_budget_is_refused():
    with Server(1 << 20) as s:
        with pytest.raises(_lib.MgpuError, match="budget"):
            s.submit(cases.tgv(32), 3)
