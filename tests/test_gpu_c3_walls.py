"""No-slip walls on the j-faces of curvilinear (c3) blocks on the GPU against
the oracle, through the C-ABI (mgpu_set_wall_c3): the curvilinear channel of
the headline configuration (tanh clustering at both walls, a wavy lower
wall, periodic x and z, an isentropic vortex) at a reduced size, FP64 within
1e-12 relative; the shock filter (Ducros), the selective filter and none,
adiabatic and isothermal walls; graph replays bit-identical to eager
launches; the unit-spacing c3 block equal to the Cartesian walled channel; the
mixed path against FP64."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

pytestmark = pytest.mark.gpu


def _oracle(c, n):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


def _gpu(c, n, eager=False, precision=_lib.PREC_FP64, tune=False):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager, precision=precision) as s:
        if tune:
            s.autotune()
        s.step(n)
        return s.get_state()


@pytest.mark.parametrize("filt,adiabatic", [(2, True), (1, True), (0, True), (2, False)])
def test_curvilinear_channel_matches_oracle(filt, adiabatic):
    c = cases.vortex_channel(48, 40, 12, filter=filt, adiabatic=adiabatic)
    n = 20
    print(f"c3 channel 48x40x12 filter={filt} adiabatic={adiabatic}, {n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    for v in (1, 2, 3):
        assert np.all(Ug[v][5:-5, 5, 5:-5] == 0.0)


def test_curvilinear_channel_graph_equals_eager_and_tuned_matches():
    c = cases.vortex_channel(40, 32, 12)
    print("c3 channel 40x32x12: graph vs eager, tuned vs oracle, 8 steps", flush=True)
    Ua = _gpu(c, 8)
    Ub = _gpu(c, 8, eager=True)
    for a, b in zip(Ua, Ub):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    Ut = _gpu(c, 8, tune=True)
    e = relerr(Ut, _oracle(c, 8))
    print(f"tuned relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_unit_c3_walls_equal_cartesian_walls_on_the_gpu():
    from tests.test_c3_wall_oracle import _pair
    cc, c3c, _, _ = _pair(20, 16, 10, 2, True)
    print("unit-spacing walled box: c3 vs Cartesian on the GPU, 6 steps", flush=True)
    Ua = _gpu(cc, 6)
    Ub = _gpu(c3c, 6)
    e = relerr(Ub, Ua)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-13


def test_curvilinear_channel_mixed_close_to_fp64():
    c = cases.vortex_channel(64, 48, 16)
    n = 40
    print(f"c3 channel 64x48x16 mixed vs FP64, {n} steps", flush=True)
    Um = _gpu(c, n, precision=_lib.PREC_MIXED)
    Ud = _gpu(c, n)
    a, b = gr.interior(Um[0]), gr.interior(Ud[0])
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(f"mixed vs FP64 rho Linf {e:.3e}", flush=True)
    assert np.all(np.isfinite(a))
    assert e < 1e-6
