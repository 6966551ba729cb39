"""Serving scheduler logic on the host (no GPU): admission by footprint,
first-come-first-served with bounded bypass, continuous admission while
other jobs run, failures isolated to their job."""
import threading
import time

import numpy as np
import pytest

from tnbp_amd import serving
from tnbp_amd._lib import HBP_PREC_FP64, HBP_PREC_MIXED


class FakeCtx:
    def trim(self):
        pass

    def close(self):
        pass


def _job(name, fp, dur, fail=False):
    j = serving.EstimatorJob(edges=[(0, 1)], tpos=[], chi=1, terms=[], name=name)
    j._fp, j._dur, j._fail = fp, dur, fail
    return j


def _server(budget, slots, log, max_bypass=8):
    lock = threading.Lock()
    live = [0]

    def runner(ctx, job):
        with lock:
            live[0] += 1
            log.append(("start", job.name, live[0]))
        time.sleep(job._dur)
        with lock:
            live[0] -= 1
            log.append(("end", job.name, live[0]))
        if job._fail:
            raise RuntimeError("job failed")
        return {"value": job.name, "n_values": 1}

    srv = serving.Server(slots=slots, memory_budget=budget, max_bypass=max_bypass, runner=runner,
                         make_context=FakeCtx)
    return srv


@pytest.fixture(autouse=True)
def fake_footprint(monkeypatch):
    monkeypatch.setattr(serving, "footprint_bytes", lambda job: job._fp)


def test_jobs_overlap_and_all_finish():
    log = []
    with _server(100, 4, log) as srv:
        futs = [srv.submit(_job(f"j{i}", 10, 0.05)) for i in range(6)]
        res = [f.result(timeout=10) for f in futs]
    assert [r["value"] for r in res] == [f"j{i}" for i in range(6)]
    assert max(n for kind, _, n in log if kind == "start") >= 2
    st = srv.stats()
    assert st["jobs"] == 6 and st["p99_s"] >= st["p50_s"] > 0 and st["max_concurrent"] >= 2


def test_admission_respects_memory_budget():
    log = []
    with _server(100, 4, log) as srv:
        futs = [srv.submit(_job(f"big{i}", 60, 0.03)) for i in range(3)]
        for f in futs:
            f.result(timeout=10)
    # two 60-byte jobs never fit a 100-byte budget together
    assert max(n for kind, _, n in log if kind == "start") == 1
    with _server(100, 2, log) as srv:
        with pytest.raises(MemoryError):
            srv.submit(_job("huge", 101, 0.0))


def test_small_jobs_pass_a_blocked_head_but_not_forever():
    log = []
    with _server(100, 4, log, max_bypass=2) as srv:
        first = srv.submit(_job("running", 70, 0.3))
        time.sleep(0.05)
        head = srv.submit(_job("head", 60, 0.01))       # waits for "running"
        smalls = [srv.submit(_job(f"s{i}", 10, 0.01)) for i in range(5)]
        for f in [first, head] + smalls:
            f.result(timeout=10)
    starts = [name for kind, name, _ in log if kind == "start"]
    # two small jobs pass the head, then the head runs before the others
    assert starts.index("s0") < starts.index("head")
    assert starts.index("head") < starts.index("s4")


def test_failure_is_isolated():
    log = []
    with _server(100, 2, log) as srv:
        bad = srv.submit(_job("bad", 10, 0.01, fail=True))
        good = srv.submit(_job("good", 10, 0.01))
        with pytest.raises(RuntimeError):
            bad.result(timeout=10)
        assert good.result(timeout=10)["value"] == "good"


def test_footprint_model_scales(monkeypatch):
    monkeypatch.undo()
    rng = np.random.default_rng(0)
    small = serving.random_circuit_job(rng, 4, 2, 8, 4)
    big = serving.random_circuit_job(rng, 8, 2, 32, 4)
    fp64 = serving.random_circuit_job(rng, 8, 2, 32, 4, precision=HBP_PREC_FP64)
    assert serving.footprint_bytes(small) < serving.footprint_bytes(big) < serving.footprint_bytes(fp64)
    # 16x16 chi=32 mixed kicked-Ising job at its steady state (bonds 32, absorbed
    # 64 from the sixth step): resident state ~3.3 GB plus D=64 work buffers
    k = serving.kicked_job(16, 32, 7, 0.7, precision=HBP_PREC_MIXED)
    assert 4e9 < serving.footprint_bytes(k) < 10e9
    # one step from |0...0> never leaves bonds of 2
    k1 = serving.kicked_job(16, 32, 1, 0.7, precision=HBP_PREC_MIXED)
    assert serving.footprint_bytes(k1) < 0.5e9
    # the small-bond sweep's scratch at bond 16 on 16x16: 1024 workgroups (the
    # largest grid) of two 16^4 * 2-element complex64 buffers
    assert serving._small_scratch_bytes(k, [16] * len(k.edges)) == 1024 * 2 * 16 ** 4 * 2 * 8
    assert serving._small_scratch_bytes(k, [17] * len(k.edges)) == 0
    # a lazily absorbing PEPS also keeps its small sites' absorbed copies
    assert serving._small_lazy_bytes(k, [16] * len(k.edges)) == sum(
        2 * 16 ** len(ms) * 8 for ms in serving._incident(k.edges).values())


def test_small_jobs_absorb_eagerly(monkeypatch):
    monkeypatch.undo()
    rng = np.random.default_rng(1)
    small = serving.random_circuit_job(rng, 4, 2, 8, 4)
    assert serving.absorbs_eagerly(small)
    # the eager footprint holds the absorbed lattice on top of the lazy one
    assert serving.footprint_bytes(small) > serving.absorbed_bytes(small)
    # 16x16 chi=32: the absorbed lattice (~52 GB mixed) is absorbed site by site
    k = serving.kicked_job(16, 32, 7, 0.7, precision=HBP_PREC_MIXED)
    assert serving.absorbed_bytes(k) > 40e9
    assert not serving.absorbs_eagerly(k)


def test_bond_bounds_follow_the_gates(monkeypatch):
    """Bond bounds: a bond grows by the TPO bond dimension of each layer that
    puts a gate on it and is cut to chi after the layer."""
    monkeypatch.undo()
    from tnbp_amd import circuits as pc
    edges = pc.square_lattice(3, 3)
    n = 9
    rzz = pc.rzz(0.4)
    # layer 1: one gate on edge 0; layer 2: edge 0 again and edge 5; layer 3: edge 0
    layers = [[(rzz, edges[0])], [(rzz, edges[0]), (rzz, edges[5])], [(rzz, edges[0])]]
    job = serving.EstimatorJob(edges, [pc.layer_tpo(edges, g) for g in layers], chi=2,
                               terms=[(1.0, "Z" + "I" * (n - 1))])
    b = serving.bond_bounds(job)
    assert [r[0] for r, a in b] == [1, 2, 2]          # resident bond 0 per layer
    assert [a[0] for r, a in b] == [2, 4, 4]          # absorbed bond 0 per layer
    assert [a[5] for r, a in b] == [1, 2, 2]
    assert all(a[1] == 1 for r, a in b)
    # the uniform chi * t model would have put every bond at 4
    assert serving.footprint_bytes(job) < 200 << 20


def test_no_device_fails_jobs_instead_of_hanging():
    def broken():
        raise RuntimeError("no device")
    srv = serving.Server(slots=2, memory_budget=100, runner=lambda c, j: None, make_context=broken)
    fut = srv.submit(_job("x", 10, 0.0))
    with pytest.raises(RuntimeError):
        fut.result(timeout=10)
    srv.close()


def test_packed_tpo_layout(monkeypatch):
    """PackedTPO (the form submit() prepares): pointer i reads site i's
    operator in column-major order, shape row i holds its shape padded with
    zeros; the server's job keeps the packed layers."""
    import ctypes as C

    from tnbp_amd import circuits
    from tnbp_amd._lib import HBP_MAX_ORDER
    from tnbp_amd.peps import PackedTPO
    rng = np.random.default_rng(3)
    edges = circuits.square_lattice(3, 3)
    layer = circuits.layer_tpo(edges, circuits.random_layer(rng, edges, 9))
    pk = PackedTPO(layer)
    assert pk.n == len(layer)
    for i, o in enumerate(layer):
        a = np.asarray(o, dtype=np.complex128)
        addr = C.cast(pk.ptrs[i], C.c_void_p).value
        got = np.frombuffer((C.c_double * (2 * a.size)).from_address(addr), dtype=np.complex128)
        np.testing.assert_array_equal(got, a.reshape(-1, order="F"))
        row = list(pk.shapes[i * HBP_MAX_ORDER:(i + 1) * HBP_MAX_ORDER])
        assert row[:a.ndim] == list(a.shape) and not any(row[a.ndim:])
    job = serving.EstimatorJob(edges=edges, tpos=[layer, layer], chi=4, terms=[], name="p")
    monkeypatch.setattr(serving, "footprint_bytes", lambda j: 0)
    srv = serving.Server.__new__(serving.Server)
    srv.budget, srv._closing, srv._pending = 1 << 62, False, []
    srv._cv = threading.Condition()
    srv.slots = 1
    serving.Server.submit(srv, job)
    assert len(job.packed) == 2 and all(isinstance(p, PackedTPO) for p in job.packed)
    assert serving._layers(job) is job.packed


def test_batched_server_trims_only_when_idle():
    """BatchedServer._admit with nothing resident: the idle hook (the
    context's trim) runs once after IDLE_TRIM_S with nothing queued, the
    scheduler keeps waiting, and a job that then arrives is admitted."""
    srv = serving.BatchedServer.__new__(serving.BatchedServer)
    srv.budget, srv.max_bypass, srv._in_use, srv._closing = 1 << 40, 8, 0, False
    srv._pending = serving.deque()
    srv._cv = threading.Condition()
    srv.IDLE_TRIM_S = 0.05
    calls, got = [], []
    th = threading.Thread(target=lambda: got.append(srv._admit(True, idle=lambda: calls.append(1))))
    th.start()
    time.sleep(0.3)
    assert calls == [1] and not got
    job = serving.EstimatorJob(edges=[(0, 1)], tpos=[], chi=1, terms=[], name="late")
    with srv._cv:
        srv._pending.append((job, 10, None))
        srv._cv.notify_all()
    th.join(2)
    assert not th.is_alive() and got and [j.name for j, _, _ in got[0]] == ["late"]
    assert calls == [1]
