"""Estimator serving on one GPU: a request loop with continuous batching.

No reference counterpart: the reference runs one job per process
(apps/estimator/main.cc:42-239: QASM circuit -> TPO slices -> AbsorbTPO, BP,
Truncation per slice -> sparse-Pauli measurement).  Here a ``Server`` accepts
such jobs as they arrive and runs several at once:

* every job's PEPS and messages stay resident in HBM from admission to result;
* admission is by memory footprint (``footprint_bytes``, the resident state
  plus the work buffers of the largest absorbed site) against a budget of the
  device's free memory, first come first served, with later jobs allowed past
  a head job that does not fit yet (at most ``max_bypass`` times, so the head
  cannot starve);
* each of ``slots`` worker threads owns an hbp context (its own HIP stream and
  memory pool); the library calls release the interpreter lock, so admitted
  jobs' kernels overlap on the device, and a job admitted while others run
  starts at once (continuous batching, not a queue of batches);
* latency is measured from arrival to result; ``stats`` gives p50/p99 and
  expectation values/s.

Every job runs on its own stream with deterministic kernels (fixed-order
reductions, no atomics), so its result does not depend on what else is
resident.
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import circuits, pauli
from ._lib import (HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, HBP_BATCH_GRAPH, HBP_BATCH_SMALL, HBP_PREC_FP64,
                   HBP_PREC_MIXED)
from .peps import (AbsorbTPO, AbsorbTPOBatched, InitTensorProductState, PackedTPO, Truncation,
                   TruncationBatched, bp_loop)
from .tci import Context, device_memory


@dataclass
class EstimatorJob:
    """One estimator request: a circuit as TPO slices on a lattice and a sparse
    Pauli observable [(coefficient, label)] (qubit 0 = last character)."""
    edges: list
    tpos: list
    chi: int
    terms: list
    precision: int = HBP_PREC_MIXED
    max_bp_iterations: int = 50
    bp_tolerance: float = 1.0e-8      # apps/estimator/option.h:17
    sv_min: float = 1.0e-8
    mode: int = 0                     # HBP_MODE_FUSED; 1 = the reference's op sequence
    name: str = ""
    arrival: float = field(default=0.0, repr=False)
    bypassed: int = field(default=0, repr=False)
    # pauli.classify(edges, terms) and the layers as PackedTPOs, made by
    # submit() on the caller's thread
    plan: object = field(default=None, repr=False)
    packed: list = field(default=None, repr=False)


def _layers(job: EstimatorJob):
    """The job's circuit layers in the form AbsorbTPO takes: packed when
    submit() packed them."""
    return job.packed if job.packed is not None else job.tpos


def _incident(edges):
    """Incident edges of every site in ascending edge order (graph.h:40-55)."""
    inc = {}
    for m, (a, b) in enumerate(edges):
        inc.setdefault(a, []).append(m)
        inc.setdefault(b, []).append(m)
    return inc


def bond_bounds(job: EstimatorJob):
    """Upper bounds of every bond's dimension per circuit layer, from
    |0...0> (bonds of 1) and each layer's TPO bond dimensions t: while the
    layer's BP loop runs a bond is at most D t (absorbed), after its
    truncation at most min(chi, D t) (sv_min may cut further, never less).
    Returns [(resident, absorbed)] per layer (lists over edges)."""
    inc = _incident(job.edges)
    sites = sorted(inc)
    D = [1] * len(job.edges)
    out = []
    for tpo in job.tpos:
        t = [1] * len(job.edges)
        for s, O in zip(sites, tpo):
            if O.ndim > 2:
                for k, m in enumerate(inc[s]):
                    t[m] = max(t[m], O.shape[k])
        absorbed = [d * x for d, x in zip(D, t)]
        out.append((D, absorbed))
        D = [min(job.chi, x) for x in absorbed]
    return out


# A job whose absorbed lattice takes at most this many bytes absorbs each
# layer eagerly: its sites are then absorbed once per layer instead of once per
# BP sweep, which at small bond dimensions saves more host time (launches)
# than the memory is worth.
EAGER_ABSORB_MAX_BYTES = 512 << 20


def _site_bytes(job, dims):
    """Bytes of every site tensor with bond dimensions `dims` (physical 2)."""
    elem = 16 if job.precision == HBP_PREC_FP64 else 8
    out = {}
    for s, ms in _incident(job.edges).items():
        n = 2
        for m in ms:
            n *= dims[m]
        out[s] = n * elem
    return out


# The small-bond kernels (sites whose bonds are all <= 16, degree <= 4) take
# scratch of two site-sized buffers per workgroup: up to four work items per
# site and at most four workgroups per CU on a 256-CU device.
SMALL_D_MAX = 16
SMALL_WORKGROUPS_MAX = 4 * 256


def _small_sites(job, dims):
    """Element counts of the sites the small-bond sweep takes at bond dims `dims`."""
    out = []
    for s, ms in _incident(job.edges).items():
        if len(ms) > 4 or any(dims[m] > SMALL_D_MAX for m in ms):
            continue
        n = 2
        for m in ms:
            n *= dims[m]
        out.append(n)
    return out


def _small_scratch_bytes(job, dims):
    """Scratch of the small-bond sweep over the sites with bond dims `dims`."""
    elem = 16 if job.precision == HBP_PREC_FP64 else 8
    sizes = _small_sites(job, dims)
    return min(SMALL_WORKGROUPS_MAX, 4 * len(sizes)) * 2 * max(sizes, default=0) * elem


def _small_lazy_bytes(job, dims):
    """The absorbed copies of the small sites that a lazily absorbing PEPS
    keeps during a BP loop (the small-bond sweep reads them)."""
    elem = 16 if job.precision == HBP_PREC_FP64 else 8
    return sum(_small_sites(job, dims)) * elem


def absorbed_bytes(job: EstimatorJob) -> int:
    """Bytes of the largest absorbed lattice over the job's layers."""
    return max((sum(_site_bytes(job, ab).values()) for _, ab in bond_bounds(job)), default=0)


def absorbs_eagerly(job: EstimatorJob) -> bool:
    return job.mode == 1 or absorbed_bytes(job) <= EAGER_ABSORB_MAX_BYTES


def pair_cache_bytes(job: EstimatorJob) -> int:
    """Bytes of the merged sweep's pair-front site copies (sites of degree 3
    and 4 whose target pair, bonds h and h + 1 with h = z // 2, both lie in
    (32, 64] after absorption, when the job absorbs eagerly), at the largest
    layer; 0 otherwise."""
    if job.mode == 1 or not absorbs_eagerly(job):
        return 0
    inc = _incident(job.edges)
    best = 0
    for _, ab in bond_bounds(job):
        sb = _site_bytes(job, ab)
        tot = 0
        for s, ms in inc.items():
            z = len(ms)
            if z not in (3, 4):
                continue
            h = z // 2
            if all(32 < ab[ms[k]] <= 64 for k in (h, h + 1)):
                tot += sb[s]
        best = max(best, tot)
    return best


def footprint_bytes(job: EstimatorJob) -> int:
    """Device bytes a job holds at its peak, from the bond bounds of every
    layer: resident sites, the absorbed lattice when it absorbs eagerly (the
    plain mode always does) and then also the pair-front copies of the merged
    sweep, the messages (three sets during a sweep), the work buffers of the
    per-site sweep (four of the largest absorbed site), the small-bond
    sweep's scratch (and, absorbing lazily, its absorbed copies of the small
    sites), split-K partials, and slack."""
    eager = absorbs_eagerly(job)
    peak = 0
    for res, ab in bond_bounds(job):
        resident = max(sum(_site_bytes(job, res).values()),
                       sum(_site_bytes(job, [min(job.chi, x) for x in ab]).values()))
        work = 4 * max(_site_bytes(job, ab).values())
        msgs = 3 * 2 * sum(x * x for x in ab) * 16
        dmax = max(ab)
        partials = 1024 * min(dmax, 128) ** 2 * 16 * 2
        small = _small_scratch_bytes(job, ab)
        if not eager:
            small += _small_lazy_bytes(job, ab)
        peak = max(peak, resident + work + msgs + partials + small)
    if eager:
        peak += absorbed_bytes(job) + pair_cache_bytes(job)
    # + the context's fixed arenas (deferred split-K partials: 64 MiB) and tables
    return int(1.15 * peak) + (128 << 20)


def run_job(ctx: Context, job: EstimatorJob):
    """The estimator loop (apps/estimator/main.cc:119-178) and its measurement
    (:181-240, extended to every term and the coefficient-weighted sum)."""
    sites = sorted({s for e in job.edges for s in e})
    peps = InitTensorProductState(ctx, job.edges, [2] * len(sites), job.precision)
    # the footprint model counts chi-sized resident sites, plus the absorbed
    # lattice when that is small enough to absorb eagerly
    peps.set_absorb_mode(HBP_ABSORB_EAGER if absorbs_eagerly(job) else HBP_ABSORB_LAZY)
    peps.set_mode(job.mode)
    # the pair-front copies only where the admitted footprint counted them:
    # the device's free memory is shared with the other admitted jobs
    peps.set_pair_cache(1 if pair_cache_bytes(job) > 0 else 0)
    iters = []
    try:
        for tpo in _layers(job):
            AbsorbTPO(ctx, tpo, peps)
            it, _ = bp_loop(ctx, peps, job.max_bp_iterations, job.bp_tolerance)
            iters.append(it)
            Truncation(ctx, peps, job.chi, job.sv_min)
        total, vals = pauli.estimate(ctx, peps, job.edges, job.terms, job.plan)
    finally:
        peps.close()
    return {"value": total, "values": vals, "bp_iterations": iters, "n_values": len(vals)}


class Server:
    """Continuous-batching estimator server on one device (see module doc)."""

    def __init__(self, device=0, slots=4, memory_budget=None, max_bypass=8, runner=None,
                 make_context=None):
        self.device = device
        self._runner = runner or run_job
        self._make_context = make_context or (lambda: Context(device))
        self.slots = slots
        if memory_budget is None:
            free, _ = device_memory(device)
            memory_budget = int(0.9 * free)
        self.budget = memory_budget
        self.max_bypass = max_bypass
        self._pending = deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._running = 0
        self._closing = False
        self.records = []
        self._t0 = time.perf_counter()
        self._threads = [threading.Thread(target=self._worker, name=f"hbp-slot{i}", daemon=True)
                         for i in range(slots)]
        for th in self._threads:
            th.start()

    # ---------------------------------------------------------------- API
    def submit(self, job: EstimatorJob) -> Future:
        arrival = time.perf_counter()
        fp = footprint_bytes(job)
        if fp > self.budget:
            raise MemoryError(f"job {job.name!r} needs {fp / 2**30:.1f} GiB, budget "
                              f"{self.budget / 2**30:.1f} GiB")
        # the observable's host work here, off the device-feeding threads (a
        # term that cannot be measured raises again there, into the future)
        if job.plan is None:
            try:
                job.plan = pauli.classify(job.edges, job.terms)
            except Exception:
                job.plan = None
        if job.packed is None:
            try:
                job.packed = [PackedTPO(O) for O in job.tpos]
            except Exception:
                job.packed = None
        fut = Future()
        job.arrival = arrival
        with self._cv:
            if self._closing:
                raise RuntimeError("server is closed")
            if getattr(self, "_dead", 0) == self.slots:
                fut.set_exception(self._dead_reason)
                return fut
            self._pending.append((job, fp, fut))
            self._cv.notify_all()
        return fut

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for th in self._threads:
            th.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self):
        """p50/p99 latency (arrival -> result), mean queue wait, expectation
        values/s over the span from the first arrival to the last result."""
        recs = list(self.records)
        if not recs:
            return {}
        lat = np.array([r["latency"] for r in recs])
        span = max(r["done"] for r in recs) - min(r["arrival"] for r in recs)
        nv = sum(r["n_values"] for r in recs)
        return {"jobs": len(recs), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max()),
                "mean_wait_s": float(np.mean([r["wait"] for r in recs])),
                "values_per_s": nv / span if span > 0 else float("inf"),
                "span_s": span, "max_concurrent": max(r["concurrent"] for r in recs)}

    # ------------------------------------------------------------ workers
    def _take(self):
        """Next admissible job (FIFO; later jobs may pass a head that does not
        fit, up to max_bypass times), or None when closing and drained."""
        with self._cv:
            while True:
                free = self.budget - self._in_use
                for i, (job, fp, fut) in enumerate(self._pending):
                    if fp <= free:
                        del self._pending[i]
                        for j in range(i):
                            self._pending[j][0].bypassed += 1
                        self._in_use += fp
                        self._running += 1
                        return job, fp, fut, self._running
                    if job.bypassed >= self.max_bypass:
                        break  # the head has waited long enough: nobody passes it
                if self._closing and not self._pending:
                    return None
                self._cv.wait()

    def _release(self, fp):
        with self._cv:
            self._in_use -= fp
            self._running -= 1
            self._cv.notify_all()

    def _worker(self):
        try:
            ctx = self._make_context()
        except BaseException as e:  # no device: fail what is queued rather than hang
            with self._cv:
                self._dead = getattr(self, "_dead", 0) + 1
                self._dead_reason = e
                if self._dead == self.slots:
                    while self._pending:
                        _, _, fut = self._pending.popleft()
                        fut.set_exception(e)
                self._cv.notify_all()
            return
        try:
            while True:
                item = self._take()
                if item is None:
                    return
                job, fp, fut, conc = item
                start = time.perf_counter()
                try:
                    res = self._runner(ctx, job)
                except BaseException as e:  # the job fails, the server goes on
                    self._release(fp)
                    fut.set_exception(e)
                    continue
                ctx.trim()
                done = time.perf_counter()
                self._release(fp)
                rec = {"name": job.name, "arrival": job.arrival, "start": start, "done": done,
                       "latency": done - job.arrival, "wait": start - job.arrival,
                       "n_values": res["n_values"], "concurrent": conc, "footprint": fp}
                res.update(rec)
                self.records.append(rec)
                fut.set_result(res)
        finally:
            ctx.close()


class BatchedServer:
    """Continuous batching at the granularity of circuit layers, on one
    context (no reference counterpart).

    One scheduler thread keeps every admitted job's PEPS resident and runs
    rounds: every resident job absorbs its next TPO slice (one
    ``hbp_peps_absorb_tpo_batched`` call), then the BP loops of
    all resident jobs run as one ``hbp_peps_bp_run_batched`` call (one
    replay of a HIP graph per BP iteration for all of them: jobs whose bonds
    are <= 16 in one persistent launch over every site of every such job,
    larger ones with their own fused sweeps recorded in the same graph), then
    all resident jobs truncate in one ``hbp_peps_truncate_batched`` call; a
    job whose last slice is done is measured and leaves.  Queued
    jobs are admitted by footprint at every round boundary, so a job that
    arrives while others run joins the next round.  A job's result does not
    depend on the others in the batch (every site's messages are computed by
    one workgroup in a fixed order).  Same interface as ``Server``.
    """

    # device scratch of the batched sweep (per precision: up to 4 workgroups
    # per CU, two site-sized buffers each at bond 16), kept out of the budget
    SCRATCH_RESERVE = 6 << 30
    # The blocks a retired job frees stay with the one context and serve the
    # next jobs' allocations (an allocation that fails first gives the kept
    # blocks back and retries); they go back to the device only once nothing
    # has been resident or queued for this long.
    IDLE_TRIM_S = 0.5

    def __init__(self, device=0, memory_budget=None, max_bypass=8, make_context=None):
        self.device = device
        self._make_context = make_context or (lambda: Context(device))
        if memory_budget is None:
            free, _ = device_memory(device)
            memory_budget = int(0.9 * free) - self.SCRATCH_RESERVE
        self.budget = memory_budget
        self.max_bypass = max_bypass
        self._pending = deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._closing = False
        self.records = []
        self.rounds = []
        self._thread = threading.Thread(target=self._loop, name="hbp-batch", daemon=True)
        self._thread.start()

    submit = Server.submit
    stats = Server.stats

    @property
    def slots(self):
        return 1

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _admit(self, block, idle=None):
        """Jobs that fit now (FIFO with bounded bypass, like Server._take);
        waits for arrivals when `block` and nothing is resident.  ``idle``
        runs once (outside the lock) when nothing has been queued for
        IDLE_TRIM_S seconds while waiting."""
        out = []
        with self._cv:
            while True:
                i = 0
                while i < len(self._pending):
                    job, fp, fut = self._pending[i]
                    if fp <= self.budget - self._in_use:
                        del self._pending[i]
                        for j in range(i):
                            self._pending[j][0].bypassed += 1
                        self._in_use += fp
                        out.append((job, fp, fut))
                        continue
                    if job.bypassed >= self.max_bypass:
                        break
                    i += 1
                if out or not block:
                    return out
                if self._closing and not self._pending:
                    return None
                if idle is not None and not self._pending:
                    if not self._cv.wait(timeout=self.IDLE_TRIM_S) and not self._pending:
                        self._cv.release()
                        try:
                            idle()
                        finally:
                            self._cv.acquire()
                        idle = None
                    continue
                self._cv.wait()

    def _loop(self):
        from .peps import bp_loop_batched
        try:
            ctx = self._make_context()
        except BaseException as e:
            with self._cv:
                self._dead, self._dead_reason = 1, e
                while self._pending:
                    self._pending.popleft()[2].set_exception(e)
            return
        active = []
        try:
            while True:
                new = self._admit(block=not active, idle=ctx.trim)
                if new is None:
                    return
                for job, fp, fut in new:
                    start = time.perf_counter()
                    try:
                        sites = sorted({x for e in job.edges for x in e})
                        peps = InitTensorProductState(ctx, job.edges, [2] * len(sites), job.precision)
                        peps.set_absorb_mode(HBP_ABSORB_EAGER if absorbs_eagerly(job) else HBP_ABSORB_LAZY)
                        peps.set_mode(job.mode)
                        peps.set_pair_cache(1 if pair_cache_bytes(job) > 0 else 0)
                    except BaseException as e:
                        self._finish_fp(fp)
                        fut.set_exception(e)
                        continue
                    active.append({"job": job, "fp": fp, "fut": fut, "peps": peps, "layer": 0,
                                   "start": start, "iters": []})
                if not active:
                    continue
                t0 = time.perf_counter()
                AbsorbTPOBatched(ctx, [_layers(r["job"])[r["layer"]] for r in active],
                                 [r["peps"] for r in active])
                t1 = time.perf_counter()
                groups = {}
                for r in active:
                    groups.setdefault((r["job"].max_bp_iterations, r["job"].bp_tolerance), []).append(r)
                nb = ng = 0
                for (mi, tol), rs in groups.items():
                    out = bp_loop_batched(ctx, [r["peps"] for r in rs], mi, tol)
                    for r, (it, res, conv, bat) in zip(rs, out):
                        r["iters"].append(it)
                        nb += int(bat == HBP_BATCH_SMALL)
                        ng += int(bat == HBP_BATCH_GRAPH)
                t2 = time.perf_counter()
                TruncationBatched(ctx, [r["peps"] for r in active],
                                  [r["job"].chi for r in active], [r["job"].sv_min for r in active])
                for r in active:
                    r["layer"] += 1
                t3 = time.perf_counter()
                still = []
                for r in active:
                    if r["layer"] < len(r["job"].tpos):
                        still.append(r)
                        continue
                    self._retire(ctx, r, len(active))
                t4 = time.perf_counter()
                # wall seconds of the round's phases (each call returns once its
                # work is done on the device)
                self.rounds.append({"jobs": len(active), "batched": nb, "graph": ng,
                                    "seconds": t4 - t0, "absorb_s": t1 - t0, "bp_s": t2 - t1,
                                    "truncate_s": t3 - t2, "measure_s": t4 - t3})
                active = still
        except BaseException as e:  # fail every resident, queued and later job
            for r in active:
                if not r["fut"].done():
                    r["fut"].set_exception(e)
            with self._cv:
                self._dead, self._dead_reason = 1, e
                while self._pending:
                    self._pending.popleft()[2].set_exception(e)
        finally:
            for r in active:
                try:
                    r["peps"].close()
                except Exception:
                    pass
            ctx.close()

    def _finish_fp(self, fp):
        with self._cv:
            self._in_use -= fp
            self._cv.notify_all()

    def _retire(self, ctx, r, concurrent):
        job = r["job"]
        try:
            total, vals = pauli.estimate(ctx, r["peps"], job.edges, job.terms, job.plan)
        except BaseException as e:
            r["peps"].close()
            self._finish_fp(r["fp"])
            r["fut"].set_exception(e)
            return
        r["peps"].close()
        done = time.perf_counter()
        self._finish_fp(r["fp"])
        res = {"value": total, "values": vals, "bp_iterations": r["iters"], "n_values": len(vals)}
        rec = {"name": job.name, "arrival": job.arrival, "start": r["start"], "done": done,
               "latency": done - job.arrival, "wait": r["start"] - job.arrival,
               "n_values": len(vals), "concurrent": concurrent, "footprint": r["fp"]}
        res.update(rec)
        self.records.append(rec)
        r["fut"].set_result(res)


def run_job_batched(ctx, job: EstimatorJob):
    """``run_job`` through the batched BP loop (a batch of one): the result a
    job gets from ``BatchedServer`` alone."""
    from .peps import bp_loop_batched
    sites = sorted({s for e in job.edges for s in e})
    peps = InitTensorProductState(ctx, job.edges, [2] * len(sites), job.precision)
    peps.set_absorb_mode(HBP_ABSORB_EAGER if absorbs_eagerly(job) else HBP_ABSORB_LAZY)
    peps.set_mode(job.mode)
    peps.set_pair_cache(1 if pair_cache_bytes(job) > 0 else 0)
    iters = []
    try:
        for tpo in _layers(job):
            AbsorbTPO(ctx, tpo, peps)
            (it, _, _, _), = bp_loop_batched(ctx, [peps], job.max_bp_iterations, job.bp_tolerance)
            iters.append(it)
            Truncation(ctx, peps, job.chi, job.sv_min)
        total, vals = pauli.estimate(ctx, peps, job.edges, job.terms, job.plan)
    finally:
        peps.close()
    return {"value": total, "values": vals, "bp_iterations": iters, "n_values": len(vals)}


# ------------------------------------------------------------- workloads
def random_terms(rng, L, n_terms, w_min, w_max):
    """Pauli strings of weight w ~ U{w_min..w_max} on a contiguous row run of
    an L x L lattice (site x + L y), coefficients ~ N(0, 1)."""
    terms = []
    n = L * L
    for _ in range(n_terms):
        w = int(rng.integers(w_min, min(w_max, L) + 1))
        y = int(rng.integers(0, L))
        x0 = int(rng.integers(0, L - w + 1))
        label = ["I"] * n
        for x in range(x0, x0 + w):
            label[n - 1 - (x + L * y)] = "XYZ"[rng.integers(3)]
        terms.append((complex(rng.standard_normal()), "".join(label)))
    return terms


def random_circuit_job(rng, L, depth, chi, n_terms, w_max=4, precision=HBP_PREC_MIXED,
                       name=""):
    """A ragged-batch job: depth layers of RZZ(theta ~ U[0, pi]) on a random
    matching then U3 (angles ~ U[0, 2 pi]) on every site, one TPO slice per
    layer, and n_terms random Pauli strings of weight 1..w_max."""
    edges = circuits.square_lattice(L, L)
    tpos = [circuits.layer_tpo(edges, circuits.random_layer(rng, edges, L * L))
            for _ in range(depth)]
    return EstimatorJob(edges=edges, tpos=tpos, chi=chi, terms=random_terms(rng, L, n_terms, 1, w_max),
                        precision=precision, name=name or f"L{L}-d{depth}-chi{chi}")


def kicked_job(L, chi, steps, h, J=math.pi / 4, precision=HBP_PREC_MIXED, tol=1e-10, name=""):
    """A kicked-Ising job: `steps` Floquet steps RZZ(J) on every bond then
    RX(h) on every site, observable = all <Z_i> and nearest-neighbour <Z_i Z_j>
    with unit coefficients."""
    edges = circuits.square_lattice(L, L)
    tpo = circuits.kicked_ising_tpo(edges, J / (2 * math.pi), 0.0, h / (2 * math.pi), math.pi)
    n = L * L
    terms = []
    for s in range(n):
        lab = ["I"] * n
        lab[n - 1 - s] = "Z"
        terms.append((1.0 + 0j, "".join(lab)))
    for a, b in edges:
        lab = ["I"] * n
        lab[n - 1 - a] = "Z"
        lab[n - 1 - b] = "Z"
        terms.append((1.0 + 0j, "".join(lab)))
    return EstimatorJob(edges=edges, tpos=[tpo] * steps, chi=chi, terms=terms, precision=precision,
                        bp_tolerance=tol, name=name or f"kicked-L{L}-chi{chi}")


def poisson_submit(server, jobs, rate, rng):
    """Submit jobs with exponential inter-arrival times of mean 1/rate (s);
    returns their futures."""
    futs = []
    for job in jobs:
        futs.append(server.submit(job))
        time.sleep(float(rng.exponential(1.0 / rate)))
    return futs
