"""GPU tests of the served form (gpucad_server_*, gpucad_plain_run_trace).

Each request's field and mesh must be the same bits alone (one-grid entry
points, or a server that sees only that request) and in the mix (a server
that batches it with others, under memory pressure or not).  Admission must
keep the arena's peak use within the budget; a request larger than the
budget is rejected; the trace drivers complete every request.
"""
import time

import numpy as np
import pytest
import torch

from gpucadforam_amd import configs as CF
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

REQS = CF.served_requests(24, seed=0x5E7E)


def _alone(gc, p):
    f = torch.empty(p.npoints(), dtype=torch.float32, device="cuda")
    s = torch.zeros(gc.sign_words(p.nx, p.ny, p.nz), dtype=torch.int64, device="cuda")
    gc.lattice_field(p, out=f, signs=s, iso=0.0)
    m = gc.extract(f, p.nx, p.ny, p.nz, signs=s)
    torch.cuda.synchronize()
    return to_np(f), to_np(m.vertices), to_np(m.triangles)


@pytest.fixture(scope="module")
def alone(gc, dev):
    return [_alone(gc, p) for p in REQS]


def _drain(srv, want, timeout_s=60.0):
    got = {}
    t0 = time.time()
    while len(got) < want:
        for r in srv.poll():
            got[r.ticket] = r
        assert time.time() - t0 < timeout_s, "server did not finish"
    return got


def _check(srv, results, tickets, reqs, refs, label):
    for t, p, (f_ref, v_ref, t_ref) in zip(tickets, reqs, refs):
        r = results[t]
        assert r.status == 0, "%s: status %d" % (label, r.status)
        f, v, tr = srv.result_arrays(r, p)
        assert np.array_equal(f.view(np.uint32), f_ref.view(np.uint32)), "%s field" % label
        assert (len(v), len(tr)) == (len(v_ref), len(t_ref)), "%s counts" % label
        assert np.array_equal(v.view(np.uint32), v_ref.view(np.uint32)), "%s vertices" % label
        assert np.array_equal(tr, t_ref), "%s triangles" % label


@pytest.mark.parametrize("bx", [64, 16])
def test_mixed_equals_alone(gc, dev, alone, bx):
    srv = gc.Server(8 << 30, brick_x=bx, brick_lz=16)
    try:
        tickets = [srv.submit(p) for p in REQS]
        res = _drain(srv, len(REQS))
        _check(srv, res, tickets, REQS, alone, "mixed bx%d" % bx)
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(REQS) and st["groups"] >= 1
        for t in tickets:
            srv.release(t)
        assert srv.stats()["bytes_in_use"] == 0
    finally:
        srv.close()


def test_request_alone_in_a_server(gc, dev, alone):
    srv = gc.Server(2 << 30)
    try:
        for i in (0, 5, 11):
            t = srv.submit(REQS[i])
            res = _drain(srv, 1)
            _check(srv, res, [t], [REQS[i]], [alone[i]], "alone %d" % i)
            srv.release(t)
    finally:
        srv.close()


def test_admission_under_budget_pressure(gc, dev, alone):
    """A budget of about three of the largest requests: groups are small,
    requests wait for memory, results stay exact and the peak stays within
    the budget."""
    largest = max(p.npoints() for p in REQS)
    budget = int(largest * 4 * 3.5) + (64 << 20)
    srv = gc.Server(budget, slots=3)
    try:
        pending = {}
        done = {}
        order = []
        for i, p in enumerate(REQS):
            pending[srv.submit(p)] = i
            order.append(i)
        t0 = time.time()
        while len(done) < len(REQS):
            for r in srv.poll():
                i = pending[r.ticket]
                _check(srv, {r.ticket: r}, [r.ticket], [REQS[i]], [alone[i]], "budget %d" % i)
                done[r.ticket] = i
                srv.release(r.ticket)  # free memory so that waiting requests get in
            assert time.time() - t0 < 120
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["peak_bytes_in_use"] <= st["budget_bytes"]
        assert st["groups"] > 1
    finally:
        srv.close()


def test_oversized_request_rejected(gc, dev):
    srv = gc.Server(16 << 20)
    try:
        big = CF.make(256, 256, 256, 0, 20.0, sheet=True, t0=0.3)
        small = CF.make(24, 24, 24, 0, 8.0, sheet=True, t0=0.3)
        tb = srv.submit(big)
        ts = srv.submit(small)
        res = _drain(srv, 2)
        assert res[tb].status == gc.E_OOM
        assert res[ts].status == gc.OK
        assert srv.stats()["rejected"] == 1
        srv.release(tb)
        srv.release(ts)
    finally:
        srv.close()


def test_continuous_batching_admits_arrivals(gc, dev):
    """Requests submitted while groups run join later groups; all finish."""
    srv = gc.Server(8 << 30, slots=2, max_batch=4)
    try:
        tickets = [srv.submit(p) for p in REQS[:6]]
        got = {}
        for p in REQS[6:]:
            for r in srv.poll():
                got[r.ticket] = r
            tickets.append(srv.submit(p))
        got.update(_drain(srv, len(REQS) - len(got)))
        assert set(got) == set(tickets)
        st = srv.stats()
        assert st["groups"] >= len(REQS) // 4
        for t in tickets:
            srv.release(t)
    finally:
        srv.close()


def test_trace_drivers(gc, dev):
    reqs = CF.served_requests(40, seed=0x99)
    arr = CF.poisson_arrivals(40, 2000.0)
    srv = gc.Server(8 << 30)
    try:
        lat, wall = srv.run_trace(reqs, arr)
    finally:
        srv.close()
    assert all(x > 0 for x in lat) and wall >= arr[-1]
    lat_p, wall_p = gc.plain_run_trace(reqs, arr)
    assert all(x > 0 for x in lat_p) and wall_p >= arr[-1]
    print("served p50 %.3f ms, plain p50 %.3f ms" % (1e3 * np.median(lat), 1e3 * np.median(lat_p)),
          flush=True)


def test_server_bad_arguments(gc, dev):
    with pytest.raises(gc.GpucadError):
        gc.Server(1024)
    with pytest.raises(gc.GpucadError):
        gc.Server(1 << 30, brick_x=12)
    srv = gc.Server(1 << 30)
    try:
        bad = CF.plain_gyroid(32)
        bad.tpms = 9
        with pytest.raises(gc.GpucadError):
            srv.submit(bad)
        with pytest.raises(gc.GpucadError):
            srv.release(12345)
    finally:
        srv.close()


def test_served_workload_poisson_matches_alone(gc, dev):
    """The served workload's own settings: a 64 GB admission budget and
    requests from seed 0x5E7E (cubic edge lognormal median 96 capped at 384,
    gyroid / Schwarz-P / diamond sheet with a sphere or box boolean)
    submitted at their Poisson arrival times, at a rate that keeps several
    groups in flight.  Every request's field and mesh equal its alone run
    (one-grid entry points) bit for bit, compared on the device."""
    n = 300
    reqs = CF.served_requests(n, seed=0x5E7E)
    arr = CF.poisson_arrivals(n, 20000.0)
    srv = gc.Server(64 << 30)
    try:
        ticket_req, got, i = {}, {}, 0
        t0 = time.perf_counter()
        while len(got) < n:
            now = time.perf_counter() - t0
            while i < n and arr[i] <= now:
                ticket_req[srv.submit(reqs[i])] = i
                i += 1
            for r in srv.poll():
                got[r.ticket] = r
            assert time.perf_counter() - t0 < 90, "server did not finish"
        st = srv.stats()
        print("served workload: %s" % st, flush=True)
        assert st["completed"] == n and st["mean_group_size"] > 1.0
        assert st["peak_bytes_in_use"] <= st["budget_bytes"]
        iv = lambda t: t.contiguous().view(torch.int32)
        for tk, r in got.items():
            p = reqs[ticket_req[tk]]
            assert r.status == 0
            f, v, t = srv.result_tensors(r, p)
            fa = torch.empty(p.npoints(), dtype=torch.float32, device=dev)
            sa = torch.zeros(gc.sign_words(p.nx, p.ny, p.nz), dtype=torch.int64, device=dev)
            gc.lattice_field(p, out=fa, signs=sa, iso=0.0)
            m = gc.extract(fa, p.nx, p.ny, p.nz, signs=sa)
            assert torch.equal(iv(f), iv(fa)), "request %d field" % ticket_req[tk]
            assert (len(v), len(t)) == (m.n_vertices, m.n_triangles)
            assert torch.equal(iv(v), iv(m.vertices)), "request %d vertices" % ticket_req[tk]
            assert torch.equal(t, m.triangles), "request %d triangles" % ticket_req[tk]
            srv.release(tk)
    finally:
        srv.close()
