"""Serving loop: admission against the HBM budget, continuous batching,
failure isolation (host-side executor, no GPU), and on the GPU: every
request's result is the same alone and in a mix."""
import collections
import threading
import time

import numpy as np
import pytest

import clk_oracle as O

Vol = collections.namedtuple("Vol", "width height depth data")


def oracle_chain(data, chain):
    cur = data
    for st in chain:
        dt = {0: np.int8, 1: np.uint8, 2: np.int16, 3: np.uint16, 4: np.float32}[st.dst_dtype]
        if st.kind == "gaussian_blur_3d":
            cur = O.gaussian_blur_3d(cur, dt, *st.sigma, nthreads=2)
        elif st.kind == "median_box_3d":
            cur = O.median_box_3d(cur, dt, *(2 * r + 1 for r in st.radius), nthreads=2)
        else:
            op = {"mean_box_3d": O.OP_MEAN, "minimum_box_3d": O.OP_MIN, "maximum_box_3d": O.OP_MAX}[st.kind]
            cur = O.box_3d(op, cur, dt, *st.radius, nthreads=2)
    return cur


class HostExecutor:
    """Runs batches on the host with the oracle; records what the server
    does (batch sizes, budget in use, concurrency) and can fail on demand."""

    def __init__(self, server_ref, delay=0.01, poison=None):
        self.server_ref = server_ref
        self.delay = delay
        self.poison = poison or set()
        self.max_used = 0
        self.sizes = []
        self.lock = threading.Lock()

    def new_stream(self):
        return object()

    def run(self, batch, stream):
        srv = self.server_ref[0]
        with self.lock:
            self.max_used = max(self.max_used, srv.used)
            self.sizes.append(len(batch))
        if any(r.id in self.poison for r in batch):
            raise RuntimeError("injected failure")
        for r in batch:
            r.result = oracle_chain(r.src.data, r.chain)
        return time.perf_counter() + self.delay

    def wait(self, ev):
        dt = ev - time.perf_counter()
        if dt > 0:
            time.sleep(dt)


def make_server(budget, max_batch=8, n_streams=2, delay=0.01, poison=None):
    from clkmi.serve import FilterServer
    ref = [None]
    ex = HostExecutor(ref, delay=delay, poison=poison)
    srv = FilterServer(ex, budget_bytes=budget, max_batch=max_batch, n_streams=n_streams)
    ref[0] = srv
    return srv, ex


def small_request(seed, shape=(9, 10, 11)):
    from clkmi.serve import Request, DEFAULT_CHAIN
    rng = np.random.default_rng(seed)
    data = rng.poisson(150, size=shape).astype(np.uint16)
    return Request(Vol(shape[2], shape[1], shape[0], data), DEFAULT_CHAIN)


def test_all_requests_complete_with_oracle_results():
    srv, ex = make_server(budget=10**9)
    reqs = [srv.submit(small_request(i)) for i in range(20)]
    assert srv.drain(timeout=60)
    srv.close()
    for r in reqs:
        assert r.error is None and r.latency is not None
        assert np.array_equal(r.result, oracle_chain(r.src.data, r.chain))
    assert srv.used == 0
    assert sum(ex.sizes) == 20


def test_budget_is_never_exceeded_and_requests_wait():
    one = small_request(0).footprint()
    srv, ex = make_server(budget=3 * one, max_batch=8, delay=0.05)
    reqs = [srv.submit(small_request(i)) for i in range(12)]
    assert srv.drain(timeout=60)
    srv.close()
    assert ex.max_used <= 3 * one
    assert max(ex.sizes) <= 3 and len(ex.sizes) >= 4  # admission held requests back
    assert all(r.error is None for r in reqs) and srv.used == 0


def test_oversized_request_is_rejected_alone():
    from clkmi.serve import Request, DEFAULT_CHAIN
    srv, ex = make_server(budget=small_request(0).footprint() * 2)
    big = Request(Vol(64, 64, 64, np.zeros((64, 64, 64), np.uint16)), DEFAULT_CHAIN)
    r1 = srv.submit(small_request(1))
    rb = srv.submit(big)
    r2 = srv.submit(small_request(2))
    assert srv.drain(timeout=60)
    srv.close()
    assert rb.error and "budget" in rb.error
    assert r1.error is None and r2.error is None


def test_failed_batch_is_isolated():
    reqs = [small_request(i) for i in range(10)]
    poison = {reqs[3].id}
    srv, ex = make_server(budget=10**9, max_batch=2, poison=poison)
    for r in reqs:
        srv.submit(r)
        time.sleep(0.002)
    assert srv.drain(timeout=60)
    srv.close()
    failed = [r for r in reqs if r.error]
    assert reqs[3] in failed and len(failed) <= 2  # only its own batch
    for r in reqs:
        if r.error is None:
            assert np.array_equal(r.result, oracle_chain(r.src.data, r.chain))
    assert srv.used == 0


class TempExecutor(HostExecutor):
    """A host executor that keeps per-stream temporaries the way the library
    does (grown to the largest batch call run on the stream, held until
    released) and records the HBM in use at every batch."""

    def __init__(self, server_ref, **kw):
        super().__init__(server_ref, **kw)
        self.kept = {}
        self.peaks = []
        self.n_release = 0

    def run(self, batch, stream):
        from clkmi.serve import batch_temp_bytes
        srv = self.server_ref[0]
        with self.lock:
            self.kept[id(stream)] = max(self.kept.get(id(stream), 0), batch_temp_bytes(batch))
            self.peaks.append(srv.used + sum(self.kept.values()))
        return super().run(batch, stream)

    def held(self, stream):
        with self.lock:
            return self.kept.get(id(stream), 0), sum(self.kept.values())

    def release(self, stream):
        with self.lock:
            self.kept[id(stream)] = 0
            self.n_release += 1


def test_kept_temporaries_count_against_the_budget():
    """Outputs in flight plus the temporaries kept per stream never exceed
    the budget; idle streams give theirs back when the head request would
    not fit otherwise, and close() releases every stream's."""
    from clkmi.serve import FilterServer, batch_temp_bytes
    one = small_request(0)
    per = one.footprint() + batch_temp_bytes([one])
    budget = int(2.5 * per)  # two streams' temporaries leave no room for a third's
    ref = [None]
    ex = TempExecutor(ref, delay=0.02)
    # at least four batches: the third stream's block only fits once an
    # idle stream has given its block back
    srv = FilterServer(ex, budget_bytes=budget, max_batch=4, n_streams=3)
    ref[0] = srv
    reqs = [srv.submit(small_request(i)) for i in range(16)]
    assert srv.drain(timeout=60)
    srv.close()
    assert all(r.error is None for r in reqs)
    assert max(ex.peaks) <= budget
    assert srv.releases > 0
    assert sum(ex.kept.values()) == 0
    for r in reqs:
        assert np.array_equal(r.result, oracle_chain(r.src.data, r.chain))


def test_size_groups_complete_small_requests_first():
    from clkmi.serve import size_groups
    reqs = [small_request(i, shape=s) for i, s in enumerate([(40, 40, 40), (4, 4, 4), (9, 9, 9), (5, 5, 5)])]
    g = size_groups(reqs)
    assert [[r.voxels for r in grp] for grp in g] == [[64, 125], [729], [64000]]


def test_latency_stats():
    from clkmi.serve import latency_stats
    srv, ex = make_server(budget=10**9)
    reqs = [srv.submit(small_request(i)) for i in range(8)]
    srv.drain(timeout=60)
    srv.close()
    st = latency_stats(reqs)
    assert st["n"] == 8 and st["p99_ms"] >= st["p50_ms"] > 0


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(gpu):
    """Requests of different sizes, served alone (one at a time) and then all
    at once with continuous batching on three streams: identical outputs,
    equal to the oracle chain."""
    import torch
    import clkmi
    from clkmi.serve import FilterServer, GpuExecutor, Request, DEFAULT_CHAIN
    rng = np.random.default_rng(5)
    vols = []
    for i in range(10):
        L = int(rng.integers(12, 40))
        data = rng.poisson(150, size=(L, L + 3, L + 5)).astype(np.uint16)
        t = torch.from_numpy(data.view(np.int16).copy()).cuda()
        vols.append((data, t, clkmi.Image(t.data_ptr(), 3, L + 5, L + 3, L)))
    srv = FilterServer(GpuExecutor(torch), budget_bytes=10**10, max_batch=16, n_streams=3)
    alone = []
    for data, t, im in vols:
        r = srv.submit(Request(im, DEFAULT_CHAIN))
        r.wait(60)
        assert r.error is None, r.error
        alone.append(r.outputs[-1].cpu().numpy().view(np.uint16).copy())
    mix = [srv.submit(Request(im, DEFAULT_CHAIN)) for data, t, im in vols]
    assert srv.drain(timeout=120)
    srv.close()
    for (data, t, im), a, r in zip(vols, alone, mix):
        assert r.error is None, r.error
        m = r.outputs[-1].cpu().numpy().view(np.uint16)
        assert np.array_equal(a, m)
        assert np.array_equal(m, oracle_chain(data, DEFAULT_CHAIN))
