"""Serving loop on the GPU: requests of different sizes, cases and
precisions, admitted by a device-memory budget and advanced together, give
bit-for-bit the result each gives alone through mgpu_step."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr

pytestmark = pytest.mark.gpu


def _requests():
    return [(cases.tgv(24), 7, _lib.PREC_FP64), (cases.vortex_periodic(32, re=900.0), 5, _lib.PREC_MIXED),
            (cases.tgv(32, re=3000.0), 9, _lib.PREC_MIXED), (cases.vortex_periodic(24), 6, _lib.PREC_FP64),
            (cases.vortex_wavy(24, 16, 12, filter=2), 4, _lib.PREC_FP64),
            (cases.vortex_wavy_c2(24, 16, 12, filter=2), 5, _lib.PREC_FP64),
            (cases.vortex_wavy_c2(28, 20, 12, filter=4), 4, _lib.PREC_MIXED)]


def test_each_result_is_the_same_alone_and_in_the_mix():
    from musica2_newgrid_amd.serving import DONE, Server
    from musica2_newgrid_amd.solver import GpuSolver
    reqs = _requests()
    alone = []
    for case, n, prec in reqs:
        with GpuSolver(case, precision=prec) as s:
            s.step(n)
            alone.append(s.get_state())
    # budget for about two resident requests: the others queue and are admitted as memory frees up
    budget = 30 * 2 ** 20
    print(f"serving mix of {len(reqs)} requests, budget {budget / 2**20:.1f} MiB", flush=True)
    with Server(budget) as srv:
        ids = [srv.submit(c, n, p) for c, n, p in reqs]
        seen_queued = False
        for _ in range(200):
            nres, nq = srv.tick(3)
            seen_queued |= nq > 0
            if all(srv.status(r)[0] == DONE for r in ids):
                break
        assert all(srv.status(r)[0] == DONE for r in ids)
        assert seen_queued
        print(f"co-batching: {srv.batching()}", flush=True)
        mixed = [srv.take(r) for r in ids]
    for q, (a, b) in enumerate(zip(alone, mixed)):
        for x, y in zip(a, b):
            assert np.array_equal(gr.interior(x), gr.interior(y)), q


def test_poisson_stream_latencies():
    from musica2_newgrid_amd.serving import make_requests, run_stream
    reqs = make_requests(10, seed=5, edges=(24, 32, 48), nsteps=40, rate=50.0)
    lat, wall, work = run_stream(reqs, budget_bytes=1 << 30, quantum=8)
    p50, p99 = np.percentile(lat, [50, 99])
    print(f"stream: {len(reqs)} requests, wall {wall:.3f} s, {work / wall / 1e6:.1f} Mcell-RK-stages/s, "
          f"p50 {p50 * 1e3:.1f} ms, p99 {p99 * 1e3:.1f} ms", flush=True)
    assert all(x > 0 for x in lat)


def test_requests_of_one_kind_share_batched_launches():
    """Requests of the same kernel variant (Cartesian, FP64) and different
    sizes, all resident at once: each group step advances all of them in one
    batched launch per kernel (co-batching), and each result is still bit for
    bit the one it gives alone."""
    from musica2_newgrid_amd.serving import DONE, Server
    from musica2_newgrid_amd.solver import GpuSolver
    reqs = [(cases.tgv(24), 7, _lib.PREC_FP64), (cases.tgv(16), 5, _lib.PREC_FP64),
            (cases.tgv(20, re=800.0), 6, _lib.PREC_FP64)]
    alone = []
    for case, n, prec in reqs:
        with GpuSolver(case, precision=prec) as s:
            s.step(n)
            alone.append(s.get_state())
    print(f"serving {len(reqs)} requests of one kind", flush=True)
    with Server(256 * 2 ** 20) as srv:
        ids = [srv.submit(c, n, p) for c, n, p in reqs]
        for _ in range(50):
            srv.tick(2)
            if all(srv.status(r)[0] == DONE for r in ids):
                break
        assert all(srv.status(r)[0] == DONE for r in ids)
        bat = srv.batching()
        print(f"co-batching: {bat}", flush=True)
        assert bat["max_members"] == 3
        assert bat["request_steps"] == 7 + 5 + 6
        assert bat["group_steps"] < bat["request_steps"]
        mixed = [srv.take(r) for r in ids]
    for q, (a, b) in enumerate(zip(alone, mixed)):
        for x, y in zip(a, b):
            assert np.array_equal(gr.interior(x), gr.interior(y)), q
