"""Serving loop for chains of 3D filters with continuous batching.

Requests arrive at any time, each a device-resident uint16 volume and a
chain of 3D filters (for instance Gaussian -> maximum box -> median box).
A scheduler thread admits waiting requests in arrival order while their HBM
footprint fits the budget, packs everything admitted into one batch and
runs the chain as one `clk_batch_filter_3d` call per chain step on the next
of several HIP streams -- so batches that are formed while earlier ones are
still on the GPU overlap with them.  One completion thread per stream waits
for that stream's batches in order, stamps their requests, releases their
budget and wakes the scheduler.  A request larger than the whole budget is
rejected; a batch that fails marks its own requests failed and leaves the
others untouched.

Each request's outputs are computed by the same kernels with the same
parameters whether it runs alone or in a batch, so results are identical
(tests check this bit for bit).
"""
import collections
import queue
import threading
import time

import numpy as np

from . import FilterJob, batch_filter_3d, Image

# bytes per pixel of the pixel type codes
_BPP = {0: 1, 1: 1, 2: 2, 3: 2, 4: 4}


class Step(collections.namedtuple("Step", "kind sigma radius dst_dtype")):
    """One 3D filter of a chain: kind is a clkmi.JOB_KINDS name; dst_dtype a
    pixel type code (0 char .. 4 float)."""

    def __new__(cls, kind, sigma=(0.0, 0.0, 0.0), radius=(0, 0, 0), dst_dtype=4):
        return super().__new__(cls, kind, tuple(sigma), tuple(radius), int(dst_dtype))


# the served chain of the benchmark: Gaussian (2,2,1) -> max r=2 -> median r=1
DEFAULT_CHAIN = (
    Step("gaussian_blur_3d", sigma=(2.0, 2.0, 1.0), dst_dtype=4),
    Step("maximum_box_3d", radius=(2, 2, 2), dst_dtype=4),
    Step("median_box_3d", radius=(1, 1, 1), dst_dtype=3),
)


class Request:
    """A volume (src: clkmi.Image on the device) and its chain."""

    _ids = iter(range(1 << 62))

    def __init__(self, src, chain=DEFAULT_CHAIN):
        self.id = next(Request._ids)
        self.src = src
        self.chain = tuple(chain)
        self.voxels = int(src.width * src.height * src.depth)
        self.t_arrival = None
        self.t_done = None
        self.error = None
        self.outputs = []      # device buffers of every step (kept alive until released)
        self.result = None     # Image of the last step's output
        self._done = threading.Event()

    def footprint(self):
        """HBM bytes the request holds while in flight: every step's output
        plus the float32 temporaries of a three-pass filter (two volumes)."""
        per = sum(_BPP[s.dst_dtype] for s in self.chain) + 8
        return self.voxels * per

    def wait(self, timeout=None):
        return self._done.wait(timeout)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class GpuExecutor:
    """Runs a batch of requests' chains on a HIP stream with libclkmi."""

    def __init__(self, torch):
        self.torch = torch

    def new_stream(self):
        return self.torch.cuda.Stream()

    def run(self, batch, stream):
        torch = self.torch
        with torch.cuda.stream(stream):
            cur = {r.id: r.src for r in batch}
            for k in range(max(len(r.chain) for r in batch)):
                jobs = []
                for r in batch:
                    if k >= len(r.chain):
                        continue
                    st = r.chain[k]
                    src = cur[r.id]
                    dt = {0: torch.int8, 1: torch.uint8, 2: torch.int16, 3: torch.int16, 4: torch.float32}[st.dst_dtype]
                    buf = torch.empty((src.depth, src.height, src.width), dtype=dt, device="cuda")
                    r.outputs.append(buf)
                    dst = Image(buf.data_ptr(), st.dst_dtype, src.width, src.height, src.depth)
                    jobs.append(FilterJob(st.kind, dst, src, st.sigma, st.radius))
                    cur[r.id] = dst
                batch_filter_3d(jobs, stream=stream)
            for r in batch:
                r.result = cur[r.id]
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev

    def wait(self, ev):
        ev.synchronize()


class FilterServer:
    """Continuous-batching server.  `executor` runs batches (GpuExecutor by
    default; tests inject a host-side one)."""

    def __init__(self, executor, budget_bytes=200 * 10**9, max_batch=64, n_streams=3):
        self.ex = executor
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.used = 0
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.running = True
        self.streams = [executor.new_stream() for _ in range(n_streams)]
        self.done_q = [queue.Queue() for _ in range(n_streams)]
        self.in_flight = [0] * n_streams
        self.next_stream = 0
        self.completed = []
        self.batches = 0
        self.max_in_flight = 0
        self._threads = [threading.Thread(target=self._schedule, daemon=True)]
        self._threads += [threading.Thread(target=self._complete, args=(i,), daemon=True)
                          for i in range(n_streams)]
        for t in self._threads:
            t.start()

    # ---- client side --------------------------------------------------------
    def submit(self, req):
        req.t_arrival = time.perf_counter()
        if req.footprint() > self.budget:
            req.error = "request needs %d bytes, more than the %d byte budget" % (req.footprint(), self.budget)
            req.t_done = req.t_arrival
            req._done.set()
            return req
        with self.cv:
            self.pending.append(req)
            self.cv.notify_all()
        return req

    def drain(self, timeout=None):
        """Wait until nothing is pending or in flight."""
        t0 = time.perf_counter()
        with self.cv:
            while self.pending or any(self.in_flight):
                left = None if timeout is None else timeout - (time.perf_counter() - t0)
                if left is not None and left <= 0:
                    return False
                self.cv.wait(left if left is not None else 0.05)
        return True

    def close(self):
        with self.cv:
            self.running = False
            self.cv.notify_all()
        for q in self.done_q:
            q.put(None)
        for t in self._threads:
            t.join(timeout=10)

    # ---- scheduler ----------------------------------------------------------
    def _admit(self):
        """Requests that fit the budget, in arrival order (head of line
        first: a large request is not overtaken indefinitely)."""
        batch = []
        while self.pending and len(batch) < self.max_batch:
            r = self.pending[0]
            if self.used + r.footprint() > self.budget:
                break
            self.pending.popleft()
            self.used += r.footprint()
            batch.append(r)
        return batch

    def _schedule(self):
        while True:
            with self.cv:
                while self.running and not (self.pending and self._admissible()):
                    self.cv.wait(0.05)
                if not self.running:
                    return
                batch = self._admit()
                si = self.next_stream
                self.next_stream = (si + 1) % len(self.streams)
                self.in_flight[si] += 1
                self.batches += 1
                self.max_in_flight = max(self.max_in_flight, sum(self.in_flight))
            try:
                ev = self.ex.run(batch, self.streams[si])
                self.done_q[si].put((batch, ev, None))
            except Exception as e:  # the batch fails alone
                self.done_q[si].put((batch, None, e))

    def _admissible(self):
        return self.used + self.pending[0].footprint() <= self.budget

    def _complete(self, si):
        q = self.done_q[si]
        while True:
            item = q.get()
            if item is None:
                return
            batch, ev, err = item
            if err is None:
                try:
                    self.ex.wait(ev)
                except Exception as e:
                    err = e
            t = time.perf_counter()
            with self.cv:
                for r in batch:
                    r.t_done = t
                    if err is not None:
                        r.error = repr(err)
                    del r.outputs[:-1]  # intermediates go; the client keeps the result
                    self.used -= r.footprint()
                    self.completed.append(r)
                self.in_flight[si] -= 1
                self.cv.notify_all()
            for r in batch:
                r._done.set()


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs if r.latency is not None and r.error is None])
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3)}
