"""Serving loop for chains of 3D filters with continuous batching.

Requests arrive at any time, each a device-resident volume and a chain of
3D filters (for instance Gaussian -> maximum box -> median box).  A
scheduler thread admits waiting requests in arrival order while their HBM
footprint fits the budget, packs everything admitted into one batch and
hands it to the next of several HIP streams -- so batches that are formed
while earlier ones are still on the GPU overlap with them.  Within a batch
the requests are grouped by size (volumes within a factor of 8 of each
other), smallest group first; each group runs its chain as one
`clk_batch_filter_3d` call per chain step and is completed on its own, so
a small request does not wait for the largest volume of its batch.  One
completion thread per stream waits for that stream's groups in order,
stamps their requests, releases their budget and wakes the scheduler.  A
request larger than the whole budget is rejected; a batch that fails marks
its own requests failed and leaves the others untouched.

HBM accounting: a request holds its chain's outputs while in flight; the
float32 temporaries of the batched passes are device blocks the library
keeps per stream (grown to the largest batch run there), which count
against the same budget for as long as they are held -- the pool's
reservation as the library reports it, plus the growth a new batch would
cause.  When the head request does not fit because of them, the blocks of
idle streams are released.  close() releases every stream's blocks.

Each request's outputs are computed by the same kernels with the same
parameters whether it runs alone or in a batch, so results are identical
(tests check this bit for bit).
"""
import collections
import queue
import threading
import time

import numpy as np

from . import FilterJob, batch_filter_3d, Image, release_temporaries, temporaries_held

# bytes per pixel of the pixel type codes
_BPP = {0: 1, 1: 1, 2: 2, 3: 2, 4: 4}
# launch tables of one batch call (a few KB per job), rounded up generously
_TABLE_SLACK = 1 << 20


def _align256(n):
    return (int(n) + 255) & ~255


def size_groups(batch, ratio=8):
    """Requests sorted by volume, split where a volume exceeds `ratio` times
    the smallest of its group."""
    groups = []
    for r in sorted(batch, key=lambda r: r.voxels):
        if groups and r.voxels <= ratio * groups[-1][0].voxels:
            groups[-1].append(r)
        else:
            groups.append([r])
    return groups


def batch_temp_bytes(batch):
    """Temporaries the largest batch call of `batch` needs (per size group
    and chain step, the sum over its requests)."""
    need = 0
    for g in size_groups(batch):
        per = [r.temp_bytes() for r in g]
        for k in range(max(len(p) for p in per)):
            need = max(need, sum(p[k] for p in per if k < len(p)) + _TABLE_SLACK)
    return need


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
        """HBM bytes the request holds while in flight: every step's output."""
        return self.voxels * sum(_BPP[s.dst_dtype] for s in self.chain)

    def temp_bytes(self):
        """Per chain step, the temporaries its 3D filter needs inside a
        batch call (as the library lays them out): one or two images of
        float32 (Gaussian, mean) or of the step's source type (minimum,
        maximum) when two or three axes are filtered; none for a median."""
        out = []
        src_dt = getattr(self.src, "dtype", 4)
        for st in self.chain:
            if st.kind == "median_box_3d":
                out.append(0)
            else:
                act = [v > 0 for v in (st.sigma if st.kind == "gaussian_blur_3d" else st.radius)]
                ntmp = min(max(sum(act) - 1, 0), 2)
                tsz = _BPP[src_dt] if st.kind in ("minimum_box_3d", "maximum_box_3d") else 4
                out.append(ntmp * _align256(self.voxels * tsz))
            src_dt = st.dst_dtype
        return out

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
        """Issue the batch on `stream`, one size group after the other;
        returns [(requests of a group, event recorded after its last step)]."""
        torch = self.torch
        done = []
        with torch.cuda.stream(stream):
            for group in size_groups(batch):
                cur = {r.id: r.src for r in group}
                for k in range(max(len(r.chain) for r in group)):
                    jobs = []
                    for r in group:
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
                for r in group:
                    r.result = cur[r.id]
                ev = torch.cuda.Event()
                ev.record(stream)
                done.append((group, ev))
        return done

    def wait(self, ev):
        ev.synchronize()

    def held(self, stream):
        """(bytes of temporaries kept for `stream`, bytes the library's pool
        holds on the device)."""
        return temporaries_held(stream)

    def release(self, stream):
        release_temporaries(stream)


class FilterServer:
    """Continuous-batching server.  `executor` runs batches (GpuExecutor by
    default; tests inject a host-side one)."""

    def __init__(self, executor, budget_bytes=200 * 10**9, max_batch=64, n_streams=3):
        self.ex = executor
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.used = 0                    # outputs of the requests in flight
        self.kept = [0] * n_streams      # temporaries the library keeps per stream
        self.pool_reserved = 0           # what the library's pool holds in all
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.running = True
        self.streams = [executor.new_stream() for _ in range(n_streams)]
        self.done_q = [queue.Queue() for _ in range(n_streams)]
        self.in_flight = [0] * n_streams
        self.next_stream = 0
        self.completed = []
        self.batches = 0
        self.groups = 0
        self.max_in_flight = 0
        self.releases = 0
        # executors that report held temporaries (the GPU one) have them
        # counted; a host-side executor holds none
        self.track_temps = hasattr(executor, "held")
        self._threads = [threading.Thread(target=self._schedule, daemon=True)]
        self._threads += [threading.Thread(target=self._complete, args=(i,), daemon=True)
                          for i in range(n_streams)]
        for t in self._threads:
            t.start()

    # ---- client side --------------------------------------------------------
    def submit(self, req):
        req.t_arrival = time.perf_counter()
        need = req.footprint() + (batch_temp_bytes([req]) if self.track_temps else 0)
        if need > self.budget:
            req.error = "request needs %d bytes, more than the %d byte budget" % (need, self.budget)
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
        """Stop the threads and release every stream's temporaries."""
        with self.cv:
            self.running = False
            self.cv.notify_all()
        for q in self.done_q:
            q.put(None)
        for t in self._threads:
            t.join(timeout=10)
        release = getattr(self.ex, "release", None)
        if release is not None:
            for st in self.streams:
                release(st)

    # ---- scheduler ----------------------------------------------------------
    def _held(self):
        return max(self.pool_reserved, sum(self.kept))

    def _fits(self, batch, si):
        """Outputs in flight + temporaries held + the growth of stream si's
        temporaries that `batch` would cause (the whole new block: the old
        one may stay reserved in the pool) within the budget."""
        T = batch_temp_bytes(batch) if self.track_temps else 0
        growth = T if T > self.kept[si] else 0
        return self.used + sum(r.footprint() for r in batch) + self._held() + growth <= self.budget

    def _admit(self, si):
        """Requests that fit the budget, in arrival order (head of line
        first: a large request is not overtaken indefinitely)."""
        batch = []
        while self.pending and len(batch) < self.max_batch:
            if not self._fits(batch + [self.pending[0]], si):
                break
            batch.append(self.pending.popleft())
        self.used += sum(r.footprint() for r in batch)
        return batch

    def _refresh(self, si):
        held = getattr(self.ex, "held", None)
        if held is not None:
            self.kept[si], self.pool_reserved = held(self.streams[si])

    def _release_idle(self):
        """Release the temporaries of streams with nothing in flight; True if
        any were held."""
        release = getattr(self.ex, "release", None)
        if release is None:
            return False
        freed = False
        for i, st in enumerate(self.streams):
            if self.in_flight[i] == 0 and self.kept[i] > 0:
                release(st)
                self.releases += 1
                self._refresh(i)
                freed = True
        return freed

    def _schedule(self):
        while True:
            with self.cv:
                while True:
                    if not self.running:
                        return
                    si = self.next_stream
                    batch = self._admit(si) if self.pending else []
                    if batch:
                        break
                    if self.pending and self._release_idle():
                        continue
                    self.cv.wait(0.05)
                self.next_stream = (si + 1) % len(self.streams)
                self.in_flight[si] += 1
                self.batches += 1
                self.max_in_flight = max(self.max_in_flight, sum(self.in_flight))
            try:
                done = self.ex.run(batch, self.streams[si])
                if not isinstance(done, list):
                    done = [(batch, done)]
                with self.cv:
                    self._refresh(si)
                    self.groups += len(done)
                self.done_q[si].put((done, None))
            except Exception as e:  # the batch fails alone
                self.done_q[si].put(([(batch, None)], e))

    def _complete(self, si):
        q = self.done_q[si]
        while True:
            item = q.get()
            if item is None:
                return
            groups, err = item
            for gi, (group, ev) in enumerate(groups):
                e = err
                if e is None:
                    try:
                        self.ex.wait(ev)
                    except Exception as x:
                        e = x
                t = time.perf_counter()
                with self.cv:
                    for r in group:
                        r.t_done = t
                        if e is not None:
                            r.error = repr(e)
                        del r.outputs[:-1]  # intermediates go; the client keeps the result
                        self.used -= r.footprint()
                        self.completed.append(r)
                    if gi == len(groups) - 1:
                        self.in_flight[si] -= 1
                    self.cv.notify_all()
                for r in group:
                    r._done.set()


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs if r.latency is not None and r.error is None])
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3)}
