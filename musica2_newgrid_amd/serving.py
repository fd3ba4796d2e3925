"""Serving loop front end: a stream of solve requests of different sizes on
one GPU (libmusica_gpu.so's mgpu_server_*).

The server admits requests in arrival order while their device memory fits
the budget, keeps their state resident and advances all resident requests
together (one batched launch per kernel and RK stage for every group of
requests sharing a kernel variant).  :func:`run_stream` replays a timed
arrival sequence against it and reports per-request latency (completion
minus arrival) and throughput.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import sys
import time
from typing import List, Optional, Sequence

import numpy as np

from . import _lib, cases
from ._lib import MgpuError, dptr, lib
from .solver import make_case_struct

QUEUED, RESIDENT, DONE = 0, 1, 2


def _check(rc, srv, what):
    if rc != 0:
        msg = lib().mgpu_server_last_error(srv).decode(errors="replace")
        raise MgpuError(f"{what} failed (rc={rc}): {msg}")


class Server:
    def __init__(self, budget_bytes: int):
        self.h = C.c_void_p()
        _check(lib().mgpu_server_create(int(budget_bytes), C.byref(self.h)), None, "mgpu_server_create")
        self._shapes = {}

    def submit(self, case, nsteps: int, precision: int = _lib.PREC_FP64) -> int:
        cs, blocks = make_case_struct(case, precision)
        if case.curvilinear == 2:
            arrs = ([np.ascontiguousarray(a) for a in case.m2] + [np.ascontiguousarray(a) for a in case.m2v] +
                    [np.ascontiguousarray(case.cart[2], dtype=np.float64),
                     np.ascontiguousarray(case.cart[5], dtype=np.float64)])
        elif case.curvilinear:
            arrs = [np.ascontiguousarray(a) for a in case.m] + [np.ascontiguousarray(a) for a in case.mv]
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in case.cart]
        mt = (_lib._dp * len(arrs))(*[dptr(a) for a in arrs])
        U = [np.ascontiguousarray(a, dtype=np.float64) for a in case.U]
        ut = (_lib._dp * 5)(*[dptr(a) for a in U])
        rid = C.c_int64()
        _check(lib().mgpu_server_submit(self.h, C.byref(cs), mt, ut, int(nsteps), C.byref(rid)), self.h,
               "mgpu_server_submit")
        self._shapes[rid.value] = (case.nz + 10, case.ny + 10, case.nx + 10)
        return rid.value

    def tick(self, max_steps: int = 1):
        nr, nq = C.c_int(), C.c_int()
        _check(lib().mgpu_server_tick(self.h, int(max_steps), C.byref(nr), C.byref(nq)), self.h,
               "mgpu_server_tick")
        return nr.value, nq.value

    def status(self, rid: int):
        st, sd = C.c_int(), C.c_int()
        _check(lib().mgpu_server_status(self.h, int(rid), C.byref(st), C.byref(sd)), self.h, "mgpu_server_status")
        return st.value, sd.value

    def take(self, rid: int) -> List[np.ndarray]:
        shape = self._shapes[rid]
        out = [np.empty(shape) for _ in range(5)]
        _check(lib().mgpu_server_take(self.h, int(rid), *[dptr(a) for a in out]), self.h, "mgpu_server_take")
        del self._shapes[rid]
        return out

    def memory(self) -> int:
        return int(lib().mgpu_server_memory(self.h))

    def batching(self) -> dict:
        """Group steps launched, request steps advanced, their ratio (requests
        per batched launch) and the largest group."""
        g, r, m = C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib().mgpu_server_batching(self.h, C.byref(g), C.byref(r), C.byref(m)), self.h,
               "mgpu_server_batching")
        return {"group_steps": g.value, "request_steps": r.value, "max_members": m.value,
                "requests_per_launch": (r.value / g.value) if g.value else 0.0}

    def close(self):
        if getattr(self, "h", None):
            lib().mgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


@dataclasses.dataclass
class Request:
    case: object
    nsteps: int
    precision: int
    arrival: float  # seconds after the start of the stream
    kind: str = ""


def make_requests(n: int, seed: int, edges=(64, 96, 128, 192, 256), re_range=(400.0, 4000.0),
                  nsteps: int = 500, mixed_fraction: float = 0.5, rate: float = 1.0) -> List[Request]:
    """Synthetic request stream: TGV or isentropic vortex on a periodic cube
    with an edge drawn from `edges`, Re log-uniform in re_range, FP64 or mixed,
    Poisson arrivals of the given rate (requests per second)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        edge = int(rng.choice(edges))
        re = float(math.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]))))
        prec = _lib.PREC_MIXED if rng.random() < mixed_fraction else _lib.PREC_FP64
        if rng.random() < 0.5:
            case, kind = cases.tgv(edge, re=re), "tgv"
        else:
            case, kind = cases.vortex_periodic(edge, re=re), "vortex"
        out.append(Request(case, nsteps, prec, t, f"{kind}{edge}"))
    return out


def run_stream(reqs: Sequence[Request], budget_bytes: int, quantum: int = 10,
               results: Optional[dict] = None, stats: Optional[dict] = None, progress: bool = False):
    """Replay the arrival sequence in wall-clock time against a server.
    Returns (latencies in s, wall time in s, cell-stage count); results, if
    given, receives the final state of each request keyed by its index, and
    stats the server's co-batching record.  progress: one line on stderr per
    completed request."""
    order = sorted(range(len(reqs)), key=lambda q: reqs[q].arrival)
    lat = [0.0] * len(reqs)
    ids = {}
    pending = list(order)
    with Server(budget_bytes) as srv:
        t0 = time.perf_counter()
        while pending or ids:
            now = time.perf_counter() - t0
            while pending and reqs[pending[0]].arrival <= now:
                q = pending.pop(0)
                ids[srv.submit(reqs[q].case, reqs[q].nsteps, reqs[q].precision)] = q
            if not ids:
                time.sleep(max(0.0, min(0.01, reqs[pending[0]].arrival - now)))
                continue
            srv.tick(quantum)
            for rid in list(ids):
                st, _ = srv.status(rid)
                if st == DONE:
                    q = ids.pop(rid)
                    U = srv.take(rid)
                    lat[q] = time.perf_counter() - t0 - reqs[q].arrival
                    if results is not None:
                        results[q] = U
                    if progress:
                        done = len(reqs) - len(pending) - len(ids)
                        print(f"serving: {done}/{len(reqs)} done at {time.perf_counter() - t0:.1f} s "
                              f"({reqs[q].kind}, latency {lat[q]:.2f} s)", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        if stats is not None:
            stats.update(srv.batching())
    work = sum(r.case.ncells * r.case.nrk * r.nsteps for r in reqs)
    return lat, wall, work


def service_rate(edges=(64, 96, 128, 192, 256), nsteps: int = 100, budget_bytes: int = 64 << 30,
                 seed: int = 11) -> float:
    """Measured throughput of the server when saturated (cell-stages/s): one
    TGV and one vortex request of every edge, FP64 and mixed, all submitted at
    once and run to completion (enough steps that admission, upload and
    retrieval do not dominate)."""
    reqs = []
    for q, e in enumerate(edges):
        reqs.append(Request(cases.tgv(e), nsteps, _lib.PREC_FP64, 0.0, f"tgv{e}"))
        reqs.append(Request(cases.vortex_periodic(e), nsteps, _lib.PREC_MIXED, 0.0, f"vortex{e}"))
    # warm-up: the first graphs and tables of each variant
    run_stream(reqs[:2], budget_bytes)
    lat, wall, work = run_stream(reqs, budget_bytes)
    return work / wall
