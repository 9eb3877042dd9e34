"""Serving loop for suffix-array build requests on one MI355X.

Continuous batching: requests arrive at any time and wait in a FIFO queue.  Each worker owns an engine handle
(its own HIP stream and device-memory cache).  A free worker admits queued requests in arrival order for as long
as their estimated HBM footprint fits the budget that is still free, joins their texts on the device and builds
all their suffix arrays with one ``dcx_build_sa_batch`` call.  Requests that arrive while batches run are
admitted by the next free worker; with two workers the batches of both overlap on the device.  A large request
(it fills the GPU by itself) is built alone, and at most one large build runs at a time, so it is as fast as
on an idle GPU while the other worker serves the small requests.  With a batch window, a small request that
finds the GPU busy waits up to that long for others to join its batch.  Results are the per-text suffix arrays
(32- or 64-bit positions), bit-identical to building each text alone (the batched call is exact); a request
whose result the client does not keep releases it as soon as it completes.

The reference has no serving layer (it builds one SA per ``mpiexec`` run, lightweight_dcx/src/executables/cli.cpp);
this module serves callers with a stream of requests.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import DCX_E_ADMISSION, DCX_E_CANCELLED, DCX_E_HIP, DCXError, Engine

# per device: [stream, bytes reserved in torch's allocator for it] of the worker streams, shared by servers
_STREAMS: dict = {}

# device bytes per text character while a request is in flight: engine working set (see the engine's
# estimate) plus the joined copy of the text, and the SA (4 or 8 bytes per character)
BYTES_PER_CHAR = 64 + 1
FIXED_BYTES = 64 << 20


def footprint(n: int, sa_width: int = 4) -> int:
    return FIXED_BYTES + (BYTES_PER_CHAR + sa_width) * int(n)


@dataclass(eq=False)  # identity semantics: requests are compared and removed by object
class Request:
    text: "object"                  # torch.uint8 CUDA tensor
    rid: int = 0
    arrival: float = 0.0
    admitted: float = math.nan
    done: float = math.nan
    sa: "object" = None             # CUDA tensor of positions when finished: int32 (sa_width 4) or int64 (8)
    error: int = 0
    sigma_class: int = -1           # packing width class of the alphabet (3/5/8/9-bit characters)
    batch: int = -1
    large: bool = False             # built alone as a large request
    keep: bool = True               # False: the result is released as soon as the request completes
    event: threading.Event = field(default_factory=threading.Event)

    @property
    def n(self) -> int:
        return int(self.text.numel())

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class SAServer:
    """Continuous-batching server over ``workers`` engine handles."""

    def __init__(self, workers: int = 2, hbm_budget: int = 80 * 10 ** 9, max_batch_texts: int = 1024,
                 max_batch_chars: int = 64 << 20, solo_chars: int = 4 << 20, device: int = 0,
                 batching: bool = True, warm_chars: int = 0, reserve_bytes: int = 0, sa_width: int = 4,
                 batch_window_ms: float = 0.0, **build_options):
        import torch
        if sa_width not in (4, 8):
            raise ValueError("sa_width must be 4 (uint32 positions) or 8 (uint64 positions)")
        self.torch = torch
        self.sa_width = int(sa_width)
        # while another build runs, a small queue head waits up to this long for more small requests to
        # join its batch (0: admitted at once)
        self.window = float(batch_window_ms) / 1e3
        self.worker_errors: list[str] = []  # unexpected exceptions of failed builds (their requests get DCX_E_HIP)
        self.running = 0  # builds in flight
        self.budget = int(hbm_budget)
        self.free = int(hbm_budget)
        self.max_texts = max_batch_texts if batching else 1
        self.max_chars = int(max_batch_chars)
        self.solo_chars = int(solo_chars)  # requests this large already fill the GPU: built alone
        self.batching = batching
        self.opts = dict(build_options)
        self.queue: list[Request] = []
        self.cv = threading.Condition()
        self.stop = False
        self.next_id = 0
        self.batches = 0
        self.large_running = 0
        self.engines = [Engine(device) for _ in range(workers)]
        # worker streams are kept per process and reused by later servers: torch's caching allocator hands
        # a freed block only to allocations on the stream that freed it, so the blocks cached for these
        # streams (outputs, joined texts) serve every later server too
        pool = _STREAMS.setdefault(device, [])
        while len(pool) < workers:
            pool.append([torch.cuda.Stream(device=device), 0])
        self.streams = [pool[i][0] for i in range(workers)]
        if warm_chars:
            self._warm_up(int(warm_chars))
        if reserve_bytes:
            # one large block per worker stream (reserve_bytes each, once per process and stream): the
            # per-request outputs and joined texts are carved from it instead of calling hipMalloc while
            # requests are in flight
            for i in range(workers):
                if pool[i][1] == 0:
                    with torch.cuda.stream(pool[i][0]):
                        torch.empty(int(reserve_bytes), dtype=torch.uint8, device="cuda")
                    pool[i][1] = int(reserve_bytes)
            torch.cuda.synchronize()
        self.threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(workers)]
        for t in self.threads:
            t.start()
        if warm_chars:
            self._prime(int(warm_chars))

    def _warm_up(self, max_chars: int):
        """Builds throw-away suffix arrays of every power-of-two size up to max_chars (4- and 256-letter
        random texts, alone and batched) on every engine, so that the engines' device caches already hold
        blocks of every size class before the first request: a build that has to grow its cache calls
        hipMalloc and hipFree, which stall the other worker's work too."""
        torch = self.torch
        g = torch.Generator(device="cpu").manual_seed(12345)
        sizes = []
        n = 1 << 16
        while n <= max_chars:
            sizes.append(n)
            n <<= 1
        for sigma in (4, 256):
            big = torch.randint(0, sigma, (max(sizes),), generator=g, dtype=torch.uint8).cuda()
            out = torch.empty(max(sizes), dtype=torch.int32, device="cuda")
            for eng in self.engines:
                for n in reversed(sizes):
                    eng.build_sa_device(big.data_ptr(), n, out.data_ptr(), 4, **self.opts)
                if self.batching:
                    for n in reversed([x for x in sizes if x <= self.max_chars]):
                        k = max(1, min(self.max_texts, n // (1 << 16)))
                        off = np.linspace(0, n, k + 1).astype(np.uint64)
                        eng.build_sa_batch_device(big.data_ptr(), off, out.data_ptr(), **self.opts)
            del big, out
        torch.cuda.synchronize()

    def _prime(self, max_chars: int):
        """Runs a few throw-away requests through the worker threads (small ones of both alphabet classes,
        batched, and large ones alone) so that their first-use costs are paid before real requests."""
        torch = self.torch
        g = torch.Generator(device="cpu").manual_seed(54321)
        texts = []
        for sigma in (4, 256):
            texts += [torch.randint(0, sigma, (1 << 16,), generator=g, dtype=torch.uint8).cuda() for _ in range(4)]
            big = min(max_chars, max(self.solo_chars, 1 << 16))
            texts += [torch.randint(0, sigma, (big,), generator=g, dtype=torch.uint8).cuda() for _ in range(2)]
        reqs = [self.submit(t) for t in texts]
        for r in reqs:
            r.event.wait()
        with self.cv:
            self.batches = 0

    # -- client side ---------------------------------------------------------------------------------
    def submit(self, text) -> Request:
        r = text if isinstance(text, Request) else Request(text=text)
        with self.cv:
            r.rid = self.next_id
            self.next_id += 1
            if not r.arrival:
                r.arrival = time.perf_counter()
            if self.batching and r.sigma_class < 0 and r.n < self.solo_chars:
                r.sigma_class = self._sigma_class(r.text)
            if footprint(r.n, self.sa_width) > self.budget:
                r.error = DCX_E_ADMISSION
                r.done = time.perf_counter()
                r.event.set()
                return r
            self.queue.append(r)
            self.cv.notify_all()
        return r

    def _sigma_class(self, text) -> int:
        # texts are only batched with texts of the same packing width: the joined alphabet decides the
        # width of every character of the batch.  Counting the alphabet is GPU work and a device-to-host
        # read on the submitting thread; clients that know their alphabet pass the class instead.
        torch = self.torch
        if text.numel() == 0:
            return 0
        sigma = int((torch.bincount(text, minlength=256) > 0).sum())
        return 0 if sigma <= 6 else 1 if sigma <= 30 else 2 if sigma <= 254 else 3

    def cancel(self, r: Request) -> bool:
        """Cancels a request that has not been admitted yet."""
        with self.cv:
            if r in self.queue:
                self.queue.remove(r)
                r.error = DCX_E_CANCELLED
                r.done = time.perf_counter()
                r.event.set()
                return True
        return False

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        for e in self.engines:
            e.close()

    # -- workers -------------------------------------------------------------------------------------
    def _take(self) -> list[Request]:
        """Admits work in arrival order.  A large request (solo_chars or more) is built alone, and with
        batching at most one large build runs at a time: it gets the whole GPU, as in the plain path, while
        the other workers take small requests.  A small request is joined by later small requests of the
        same alphabet class while their footprint fits the free budget and the batch limits; small requests
        may pass a large queue head that waits for the running large build.  With a batch window, a small
        head that is not yet a full batch waits while another build runs, until it has waited the window:
        the requests arriving meanwhile join it.  Returns the batch (empty: nothing to admit) and how long
        the worker may sleep before looking again."""
        if not self.queue:
            return [], 0.05
        head = self.queue[0]
        if head.n >= self.solo_chars:
            if not (self.batching and self.large_running):
                fp = footprint(head.n, self.sa_width)
                if fp > self.free:
                    return [], 0.05
                self.queue.remove(head)
                self.free -= fp
                self.large_running += 1
                head.large = True
                return [head], 0.0
            smalls = [r for r in self.queue if r.n < self.solo_chars]
            if not smalls:
                return [], 0.05
            head = smalls[0]
        fp = footprint(head.n, self.sa_width)
        if fp > self.free:
            return [], 0.05
        batch, chars, need = [head], head.n + 1, fp
        if self.batching:
            for r in self.queue[self.queue.index(head) + 1:]:
                if len(batch) >= self.max_texts:
                    break
                if r.n >= self.solo_chars or r.sigma_class != head.sigma_class:
                    continue
                fr = footprint(r.n, self.sa_width)
                if chars + r.n + 1 > self.max_chars or need + fr > self.free:
                    continue
                batch.append(r)
                chars += r.n + 1
                need += fr
            if self.window > 0 and self.running > 0 and not self.stop and len(batch) < self.max_texts:
                wait = head.arrival + self.window - time.perf_counter()
                if wait > 0 and 2 * chars < self.max_chars:
                    return [], wait
        for r in batch:
            self.queue.remove(r)
        self.free -= need
        return batch, 0.0

    def _worker(self, wid: int):
        torch = self.torch
        eng = self.engines[wid]
        stream = self.streams[wid]
        while True:
            with self.cv:
                while True:
                    if self.stop and not self.queue:
                        return
                    batch, wait = self._take()
                    if batch:
                        break
                    self.cv.wait(wait)
                bid = self.batches
                self.batches += 1
                self.running += 1
            now = time.perf_counter()
            for r in batch:
                r.admitted = now
                r.batch = bid
            err = 0
            try:
                self._build(eng, stream, batch)
            except DCXError as e:
                err = e.code
            except Exception as e:  # torch OOM, ctypes errors ...: the request fails, the server lives on
                err = DCX_E_HIP
                self.worker_errors.append(repr(e))
            finally:
                # the budget, the large-build slot and the waiting clients are released whatever happened
                done = time.perf_counter()
                with self.cv:
                    self.free += sum(footprint(r.n, self.sa_width) for r in batch)
                    self.running -= 1
                    if batch[0].large:
                        self.large_running -= 1
                    self.cv.notify_all()
                for r in batch:
                    if err or not r.keep:
                        r.sa = None
                    r.error = err
                    r.done = done
                    r.event.set()

    def _build(self, eng: Engine, stream, batch: list[Request]):
        """Builds the suffix arrays of one admitted batch on the worker's engine and stream."""
        torch = self.torch
        with torch.cuda.stream(stream):
            if self.batching and len(batch) > 1:
                joined = torch.cat([r.text for r in batch])
                total = int(joined.numel())
                out = torch.empty(max(total, 1), dtype=torch.int32, device=joined.device)
                offsets = np.zeros(len(batch) + 1, dtype=np.uint64)
                offsets[1:] = np.cumsum([r.n for r in batch])
                stream.synchronize()
                eng.build_sa_batch_device(joined.data_ptr(), offsets, out.data_ptr(), **self.opts)
                if self.sa_width == 8:
                    # per-text positions are below 2^32: zero extension on the worker's stream
                    wide = out[:total].to(torch.int64) & 0xFFFFFFFF
                    for k, r in enumerate(batch):
                        r.sa = wide[int(offsets[k]):int(offsets[k + 1])]
                    stream.synchronize()
                else:
                    for k, r in enumerate(batch):
                        r.sa = out[int(offsets[k]):int(offsets[k + 1])]
            else:
                r = batch[0]
                dt = torch.int64 if self.sa_width == 8 else torch.int32
                out = torch.empty(max(r.n, 1), dtype=dt, device=r.text.device)
                stream.synchronize()
                eng.build_sa_device(r.text.data_ptr(), r.n, out.data_ptr(), self.sa_width, **self.opts)
                r.sa = out[:r.n]


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed + 1)
    return np.exp(rng.uniform(math.log(lo), math.log(hi + 1), count)).astype(np.int64).clip(lo, hi)


def run_open_loop(server: SAServer, texts: list, arrivals: np.ndarray, keep_results: bool = True,
                  sigma_classes=None, keep_every: int = 0) -> dict:
    """Submits texts at the given arrival offsets (seconds) and waits for all; returns latency statistics.
    sigma_classes (optional): the alphabet class of every text as the client knows it (see
    SAServer._sigma_class), so that submitting does not have to count the alphabet on the GPU."""
    t0 = time.perf_counter()
    reqs = []
    for k, (text, at) in enumerate(zip(texts, arrivals)):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        r = Request(text=text, arrival=t0 + at,
                    keep=keep_results or (keep_every > 0 and k % keep_every == 0))
        if sigma_classes is not None:
            r.sigma_class = int(sigma_classes[k])
        reqs.append(server.submit(r))
    for r in reqs:
        r.event.wait()
    t1 = time.perf_counter()
    ok = [r for r in reqs if r.error == 0]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    chars = sum(r.n for r in ok)
    res = {
        "requests": len(reqs), "completed": len(ok), "rejected": sum(1 for r in reqs if r.error),
        "wall_s": t1 - t0, "texts_per_s": len(ok) / (t1 - t0), "mb_per_s": chars / 1e6 / (t1 - t0),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "batches": server.batches,
    }
    for r in reqs:
        if not r.keep:
            r.sa = None
    res["requests_list"] = reqs
    return res
