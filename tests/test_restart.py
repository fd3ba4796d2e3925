"""Restart files in the reference's layout (restart_bl<N>.bin, info.ini, time.ini)."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases, restart
from musica2_newgrid_amd import grid as gr


def test_restart_file_layout_is_fortran_order_without_header(tmp_path):
    nx, ny, nz = 6, 5, 4
    U = [gr.padded(nx, ny, nz) for _ in range(5)]
    for v in range(5):
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        U[v][5:-5, 5:-5, 5:-5] = 1000 * v + 100 * k + 10 * j + i
    p = tmp_path / restart.restart_name(3)
    assert p.name == "restart_bl3.bin"
    restart.write_restart(str(p), U)
    raw = np.fromfile(p, dtype=np.float64)
    assert raw.size == 5 * nx * ny * nz  # no header, no record markers
    # rho(1,1,1), rho(2,1,1): i fastest; then j; the 2nd variable starts after nx*ny*nz values
    assert raw[0] == 0 and raw[1] == 1 and raw[nx] == 10 and raw[nx * ny] == 100
    assert raw[nx * ny * nz] == 1000
    back = restart.read_restart(str(p), nx, ny, nz)
    for a, b in zip(U, back):
        assert np.array_equal(gr.interior(a), gr.interior(b))


def test_info_and_time_files_round_trip(tmp_path):
    p = tmp_path / "info.ini"
    restart.write_info(str(p), [(64, 32, 16), (8, 8, 8)], True, 1.25e-5, ektot0=3.5, mgtot0=1.5)
    text = p.read_text().splitlines()
    assert text[0] == "nbloc & is_curv   =     2 T"
    assert text[1] == "NX NY NZ bl0001   =    64    32    16"
    d = restart.read_info(str(p))
    assert d["nbloc"] == 2 and d["is_curv"] and d["blocks"] == [(64, 32, 16), (8, 8, 8)]
    assert d["deltat"] == 1.25e-5 and d["Etot0 mgtot0"] == [3.5, 1.5]
    t = tmp_path / "time.ini"
    restart.append_time(str(t), "0001_0000", 100, 1.5, 2.5e-3)
    restart.append_time(str(t), "0002_0000", 200, 3.0, 5.0e-3)
    assert restart.read_last_time(str(t)) == ("0002_0000", 200, 3.0, 5.0e-3)


@pytest.mark.gpu
def test_gpu_run_continues_exactly_from_a_restart_file(tmp_path):
    from musica2_newgrid_amd.solver import GpuSolver
    case = cases.vortex_wavy(40, 28, 20, filter=2)
    print("restart continuity: 6 + 6 steps vs 12", flush=True)
    with GpuSolver(case) as s:
        s.step(12)
        U12 = s.get_state()
    with GpuSolver(case) as s:
        s.step(6)
        restart.write_restart(str(tmp_path / "restart_bl1.bin"), s.get_state())
    case2 = cases.vortex_wavy(40, 28, 20, filter=2)
    case2.U = restart.read_restart(str(tmp_path / "restart_bl1.bin"), 40, 28, 20)
    with GpuSolver(case2) as s:
        s.step(6)
        U = s.get_state()
    for a, b in zip(U12, U):
        assert np.array_equal(gr.interior(a), gr.interior(b))
