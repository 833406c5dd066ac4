"""Continuous-batching RCWA server on one GPU.

Requests (a ``Problem`` with its ``Numerics``) arrive over time. Each is resolved into plane-wave
points; points wait in per-window queues keyed by (order window, solve path), and a launch
packs queued points of one window -- from any number of requests, whatever their layer counts --
into one launch sequence: the layer eigensolves of all its (point, layer) pairs as one batch
(``wb_layer_modes``), then one cascade over every point whatever its layer count
(``wb_solve_from_modes_ragged``: step l takes layer l of the points that have it). Launches run on
several HIP streams at once, so work admitted
while others are on the device overlaps with them. Admission is by HBM footprint: a launch
is started only while the workspaces of all launches in flight plus the device-resident
convolution matrices fit the memory budget; a request's points are admitted as slots free
up. Convolution matrices of each (raster stack, window) stay on the device in an LRU cache
and are reused by every later request of the same design. A request completes when its
last point has been solved; its latency is completion time minus arrival time.

A request may also ask for the gradient of its total transmitted efficiency (the sum of T over
its wavelengths and orders) with respect to every raster pixel and layer thickness. Such a
request runs as launches of its own (its points in chunks that fit half the HBM budget; the
chunks' gradients add up) through the hand-written reverse pass (forward with
per-layer checkpoints, then the adjoint), on a server stream and under the same HBM
admission, driven by a worker thread so the loop keeps admitting and launching forward work
meanwhile; its result carries ``diagnostics['gradient']``.

The reference has no serving loop (one solver per wavelength, ``_lifecycle.py:712-727``);
this is the batching layer around the same per-point arithmetic, so a request's result is the
same alone and in a mix.
"""
import collections
import dataclasses
import concurrent.futures
import hashlib
import heapq
import itertools
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _native
from . import _engine
from ._lifecycle import Execution, _resolve, _runtime_checks, flag_error, flag_notice
from .results import pack
from ..errors import NumericalError, UnsupportedFeatureError, ValidationError


@dataclass
class Request:
    problem: object
    numerics: object
    arrival: float = 0.0              # seconds on the server's clock
    id: int = -1
    gradient: bool = False            # also d(sum of T)/d(rasters, thicknesses)

    @classmethod
    def from_job(cls, text, arrival=0.0, id=-1, gradient=False):
        """A request from a ``wavebit.job/v2`` document (``wavebit_amd.jobs.Job``). The server
        runs every request on its own device, so the job's execution settings (the reference's
        default is ``backend='numpy'``) are not consulted."""
        from ..jobs import Job
        job = text if isinstance(text, Job) else Job.from_json(text)
        return cls(problem=job.problem, numerics=job.numerics, arrival=arrival, id=id, gradient=gradient)


@dataclass
class _Work:
    req: Request
    resolved: object
    n_points: int
    results: dict = field(default_factory=dict)
    done: int = 0
    t_admit: float = 0.0
    t_done: float = 0.0
    flags: int = 0               # WB_FLAG_* bits over the request's points and layers


class ConvLRU:
    """Device-resident convolution matrices, least-recently-used eviction under a byte budget."""

    def __init__(self, device, budget_bytes):
        self.device = device
        self.budget = budget_bytes
        self.bytes = 0
        self._d = collections.OrderedDict()
        self.hits = 0
        self.misses = 0

    def get(self, rasters, hu, hv, factorization='li', period=None):
        key = (hashlib.blake2b(np.ascontiguousarray(rasters).tobytes(), digest_size=16).hexdigest(),
               rasters.shape, hu, hv, factorization,
               _engine.cell_of(period) if factorization == 'nvf' else None)
        hit = self._d.get(key)
        if hit is not None:
            self._d.move_to_end(key)
            self.hits += 1
            return hit
        self.misses += 1
        eps = torch.from_numpy(np.ascontiguousarray(rasters)).to(self.device)
        try:
            entry = _engine.assemble_set(eps, hu, hv, factorization, period, rasters)
        except FloatingPointError:
            raise NumericalError('a layer convolution matrix is singular', code='rcwa.singular',
                                 path='stack.layers') from None
        self._d[key] = entry
        self.bytes += self._size(entry)
        while self.bytes > self.budget and len(self._d) > 1:
            _, old = self._d.popitem(last=False)
            self.bytes -= self._size(old)
        return entry

    @staticmethod
    def _size(entry):
        return sum(t.numel() * 16 for t in entry if t is not None)


class Server:
    """Continuous batching with HBM-footprint admission and overlapping streams. Forward points
    queue per order window and path; a layer count with ``window_below`` or more queued points
    launches alone, smaller counts share window launches (one eigensolve batch, one ragged
    cascade)."""

    def __init__(self, device='cuda:0', streams=3, memory_fraction=0.6, max_points_per_launch=64,
                 conv_cache_fraction=0.1, window_below=16):
        self.window_below = max(1, window_below)
        self.device = torch.device(device)
        _native.context(self.device)
        free, total = torch.cuda.mem_get_info(self.device)
        self.budget = int(free * memory_fraction)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, streams))]
        self.max_points = max_points_per_launch
        self.conv = ConvLRU(self.device, int(free * conv_cache_fraction))
        self.execution = Execution(backend='torch', device=str(self.device))
        self.stats = collections.Counter()
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)

    def close(self):
        """Stop the gradient worker thread (a later ``serve`` call starts a new one)."""
        self._pool.shutdown(wait=True)
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self._pool.shutdown(wait=True)
        return False

    # ------------------------------------------------------------------ request intake
    def _admit(self, req, now):
        r = _resolve(req.problem, req.numerics, self.execution)
        if r.scattering:
            raise UnsupportedFeatureError(
                'The server returns diffraction outputs; solve a scattering-matrix request with '
                'rcwa.run.', code='rcwa.unsupported', path='outputs.scattering')
        if r.fields:
            raise UnsupportedFeatureError(
                'The server returns diffraction outputs; solve a field request with rcwa.run.',
                code='rcwa.unsupported', path='outputs.fields')
        if req.gradient and r.factorization == 'nvf' and not bool(
                np.all((r.rasters.real > 0) & (r.rasters.imag >= 0))):
            raise UnsupportedFeatureError(
                "Gradient requests under factorization='nvf' are served for passive dielectric "
                'layers (Re eps > 0, Im eps >= 0 everywhere).', code='rcwa.gradient',
                path='numerics.factorization')
        if req.numerics.detuning:
            raise UnsupportedFeatureError(
                'The server solves the requested points; run a request with detuning through '
                'rcwa.run.', code='rcwa.unsupported', path='numerics.detuning')
        checks = _runtime_checks(r)
        bad = [k for k, ok in checks.items() if not ok]
        if bad:
            raise ValidationError(f'runtime guard {bad[0]} failed', code=f'prepared.{bad[0]}',
                                  path='problem')
        conv = self.conv.get(r.rasters, r.hu, r.hv, r.factorization, r.period)
        work = _Work(req=req, resolved=r, n_points=len(r.wavelengths), t_admit=now)
        work.checks = checks
        pts = []
        for i, wl in enumerate(r.wavelengths):
            pts.append((work, i, _engine.Point(
                stack=0, wavelength=float(wl), theta=r.source.theta, phi=r.source.phi,
                polarization=r.polarization, eps_top=r.eps_top, eps_bottom=r.eps_bottom,
                period=r.period, thickness=tuple(float(t) for t in r.thickness), k_direct=r.k_direct), conv))
        # forward points queue per order window and path, whatever their layer count; _take
        # decides per launch whether one layer count goes alone or several share a window
        key = (r.hu, r.hv, r.path)
        work.layers = r.rasters.shape[0]
        return work, key, pts

    def _launch_bytes(self, key, items):
        """Device bytes of a window launch over ``items``: the modes of every pair and the
        workspace that computes them, and the ragged cascade's workspace plus the outputs."""
        hu, hv, path = key[:3]
        mode = _engine._MODE[path]
        N = (2 * hu + 1) * (2 * hv + 1)
        n = 2 * N if mode == _native.MODE_2D else N
        groups = collections.Counter(int(conv[0].shape[0]) for _, _, _, conv in items)
        if len(groups) == 1:                 # one layer count: the fused solve
            (L, P), = groups.items()
            out = P * (3 * 4 * N * 8 + 3 * 4 * N * 16 + 3 * 8 + L * 4)
            return _native.solve_workspace_bytes(P, L, hu, hv, mode) + out
        pairs = sum(L * c for L, c in groups.items())
        modes = pairs * (2 * n * n + n) * 16 + _native.layer_modes_bytes(pairs, 1, hu, hv, mode)
        if max(groups) <= _native.RAGGED_MAX:
            cascade = _native.solve_ragged_workspace_bytes(
                [L for L, c in groups.items() for _ in range(c)], hu, hv, mode)
        else:
            cascade = max(_native.solve_workspace_bytes(c, L, hu, hv, mode) for L, c in groups.items())
        out = len(items) * (3 * 4 * N * 8 + 3 * 4 * N * 16 + 3 * 8) + pairs * 8
        return modes + cascade + out

    def _gradient_bytes(self, key, P):
        hu, hv, path, L = key
        mode = _engine._MODE[path]
        return (2 * _native.solve_workspace_bytes(P, L, hu, hv, mode)
                + _native.checkpoint_bytes(P, L, hu, hv, mode))

    def _gradient_chunk(self, key, n_points):
        """Points per forward-and-reverse launch of a gradient request: all of them when they fit
        half the HBM budget, else the most that do (at least one). The objective, a sum over the
        points, is additive, so the chunks' gradients add up to the request's."""
        P = n_points
        while P > 1 and self._gradient_bytes(key, P) > self.budget // 2:
            P = (P + 1) // 2
        return P

    def _gradient(self, work, stream):
        """Forward with checkpoints and the reverse pass for one request, on ``stream``: fills
        the request's per-wavelength results and returns d(sum of T)/d(eps), d/d(thickness)."""
        from . import adjoint
        r = work.resolved
        if r.uniform or r.path != _engine.PATH_2D:
            raise UnsupportedFeatureError(
                'Gradients are implemented for raster stacks on the 2D coupled path; this request '
                f'routes to {r.path!r}.', code='rcwa.gradient', path='problem')
        points = [_engine.Point(stack=0, wavelength=float(wl), theta=r.source.theta, phi=r.source.phi,
                                polarization=r.polarization, eps_top=r.eps_top, eps_bottom=r.eps_bottom,
                                period=r.period, thickness=tuple(float(t) for t in r.thickness), k_direct=r.k_direct)
                  for wl in r.wavelengths]
        chunk = self._gradient_chunk((r.hu, r.hv, r.path, r.rasters.shape[0]), len(points))
        lattice = _engine.lattice_of(r.period) if r.factorization == 'nvf' else None
        parts = []
        with torch.cuda.stream(stream):
            eps = torch.tensor(np.ascontiguousarray(r.rasters), dtype=torch.complex128, device=self.device,
                               requires_grad=True)
            thick = torch.tensor(np.asarray(r.thickness, dtype=np.float64), device=self.device,
                                 requires_grad=True)
            for c0 in range(0, len(points), chunk):
                eff, amp, pe, flags = adjoint.solve(eps, thick, points[c0:c0 + chunk], r.hu, r.hv,
                                                    self.device, factorization=r.factorization,
                                                    lattice=lattice)
                (eff[:, 0, 2] + eff[:, 0, 3]).sum().backward()      # accumulates into .grad
                parts.append(tuple(t.detach().cpu().numpy() for t in (eff, amp, pe, flags)))
                del eff, amp, pe, flags
            d_eps, d_t = eps.grad.detach().cpu().numpy(), thick.grad.detach().cpu().numpy()
            if r.side == 'bottom':               # solved on the mirrored stack: back to the declared order
                d_eps, d_t = d_eps[::-1].copy(), d_t[::-1].copy()
            eff_h, amp_h, pe_h, fl = (np.concatenate([p_[i] for p_ in parts]) for i in range(4))
        for i in range(work.n_points):
            work.flags |= int(np.bitwise_or.reduce(fl[i]))
            work.results[i] = _engine.unpack(eff_h, amp_h, pe_h, i, r.hu, r.hv, r.path)
        work.done = work.n_points
        return d_eps, d_t

    def _gradient_result(self, work, grads):
        from dataclasses import replace
        result = self._result(work)
        if isinstance(result, Exception):
            return result
        diag = dict(result.diagnostics)
        diag['gradient'] = {'objective': 'sum of T over the wavelengths and orders',
                            'd_epsilon': grads[0], 'd_thickness': grads[1]}
        return replace(result, diagnostics=diag)

    def _take(self, q):
        """The items of the next launch from one order window's queue (not yet removed). A
        layer count with at least ``window_below`` queued points goes alone, as one fused solve
        of up to ``max_points_per_launch`` points (the count queued earliest first); otherwise
        the oldest points share a window launch whatever their layer counts. On the uneven
        2048-problem batch (orders 3..12, 1-8 layers, ~25 points per count and window) this
        policy runs 43.5 solves/s (80 launches) against 39.4 for windows of up to 128 mixed
        points (33 launches); on its first 512 problems (~6 points per count) it makes every
        launch a window: 46.3 solves/s in 10 launches, one per order window, as windows alone
        (46.9)."""
        counts = collections.Counter(int(item[3][0].shape[0]) for item in q)
        full = [L for L, c in counts.items() if c >= min(self.window_below, self.max_points)]
        if full:
            first = {}
            for n, item in enumerate(q):
                first.setdefault(int(item[3][0].shape[0]), n)
            L = min(full, key=first.__getitem__)
            return [item for item in q if int(item[3][0].shape[0]) == L][:self.max_points]
        return list(itertools.islice(q, min(len(q), self.max_points)))

    def _launch(self, key, items, stream):
        """One window launch: the items grouped by layer count into buckets (each with its own
        conv stacks), their modes in one batch, one ragged cascade over all of them."""
        hu, hv, path = key[:3]
        groups = collections.OrderedDict()
        for item in items:
            groups.setdefault(int(item[3][0].shape[0]), []).append(item)
        buckets, convs, members = [], [], []
        for L, group in groups.items():
            cs, index, pts = [], {}, []
            for work, i, pt, conv in group:
                cid = id(conv[0])
                if cid not in index:
                    index[cid] = len(cs)
                    cs.append(conv)
                pts.append(dataclasses.replace(pt, stack=index[cid]))   # every other field as admitted
            buckets.append(_engine.Bucket(hu=hu, hv=hv, path=path, n_layers=L, points=pts))
            convs.append(cs)
            members.append(group)
        with torch.cuda.stream(stream):
            if len(buckets) == 1:
                outs = [_engine.run_bucket(buckets[0], convs[0], self.device)]
            else:
                # no wait on the device here: grazing points are settled when the launch completes
                outs, _ = _engine.run_window(buckets, convs, self.device, rebuild=False)
            done = torch.cuda.Event()
            done.record(stream)
        return dict(key=key, items=items, outs=outs, members=members, buckets=buckets, convs=convs,
                    event=done, bytes=self._launch_bytes(key, items), stream=stream)

    def _finish(self, launch, now, finished):
        hu, hv, path = launch['key'][:3]
        for out, bucket, convs, group in zip(launch['outs'], launch['buckets'], launch['convs'],
                                             launch['members']):
            eff, amp, pe, flags = (t.cpu().numpy() for t in out[:4])
            # a grazing cluster the window's rebuild did not settle is tried once more through
            # the fused rebuild, on the launch's stream (its outputs are complete)
            with torch.cuda.stream(launch['stream']):
                _engine.rebuild_grazing(bucket, convs, self.device, eff, amp, pe, flags)
            for p, (work, i, _, _) in enumerate(group):
                work.flags |= int(np.bitwise_or.reduce(flags[p]))
                work.results[i] = _engine.unpack(eff, amp, pe, p, hu, hv, path)
                work.done += 1
                if work.done == work.n_points:
                    work.t_done = now
                    finished.append(work)

    def _result(self, work):
        r = work.resolved
        error = flag_error(work.flags, work.req.numerics)
        if error is not None:
            return error
        per = [work.results[i] for i in range(work.n_points)]
        diag = {'solve_path': tuple(r.path for _ in r.wavelengths), 'served': True,
                'latency_s': work.t_done - work.req.arrival,
                'exceptional_points': flag_notice(work.flags, work.req.numerics)[1]}
        prov = {'solver': 'rcwa', 'engine': 'wavebit_amd/hip-gfx950 serve',
                'length_units': work.req.problem.units.length}
        return pack(per, r.hu, r.hv, r.wavelengths, work.checks, diag, prov, side=r.side)

    # ------------------------------------------------------------------ the loop
    def serve(self, requests, realtime=True, progress=None):
        """Run a request stream to completion. ``arrival`` times are seconds after the start
        (``realtime=False`` admits everything at once); ``progress(id, latency_s)``, when given,
        is called as each request completes. Returns (results by id, report)."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        for i, q in enumerate(reqs):
            if q.id < 0:
                q.id = i
        pending = collections.deque(reqs)
        grads = collections.deque()                           # admitted gradient requests
        queues = collections.defaultdict(collections.deque)   # bucket key -> deque of items
        order = collections.OrderedDict()                     # bucket key -> oldest arrival
        inflight = []
        results, latencies = {}, []
        # device bytes held by launches in flight; the resident conv matrices are charged as
        # they stand at each admission (the cache grows as requests bring new rasters)
        used = 0
        free_streams = list(self.streams)
        t0 = time.perf_counter()
        solved_points = 0
        grad_job = None
        while pending or grads or grad_job is not None or any(queues.values()) or inflight:
            now = time.perf_counter() - t0
            # admit arrived requests
            while pending and (not realtime or pending[0].arrival <= now):
                q = pending.popleft()
                try:
                    work, key, items = self._admit(q, now)
                except Exception as error:           # a bad request fails alone
                    results[q.id] = error
                    latencies.append(now - q.arrival)
                    continue
                if q.gradient:
                    grads.append((work, key))
                    continue
                queues[key].extend(items)
                order.setdefault(key, q.arrival)
            # retire finished launches
            finished = []
            still = []
            for ln in inflight:
                if ln['event'].query():
                    self._finish(ln, time.perf_counter() - t0, finished)
                    used -= ln['bytes']
                    free_streams.append(ln['stream'])
                    solved_points += len(ln['items'])
                else:
                    still.append(ln)
            inflight = still
            for work in finished:
                results[work.req.id] = self._result(work)
                latencies.append(work.t_done - work.req.arrival)
                if progress is not None:
                    progress(work.req.id, latencies[-1])
            # gradient requests: retire a finished one, start the next when a stream and its
            # memory are free (one at a time, on the worker thread)
            started = False
            if grad_job is not None and grad_job[0].done():
                fut, work, stream, need = grad_job
                try:
                    res = self._gradient_result(work, fut.result())
                except Exception as error:           # fails alone
                    res = error
                work.t_done = getattr(work, 't_grad_done', time.perf_counter() - t0)
                results[work.req.id] = res
                latencies.append(work.t_done - work.req.arrival)
                if progress is not None:
                    progress(work.req.id, latencies[-1])
                if not isinstance(res, Exception):
                    solved_points += work.n_points
                free_streams.append(stream)
                used -= need
                grad_job = None
                self.stats['gradient_requests'] += 1
            if grads and grad_job is None and free_streams:
                work, key = grads[0]
                gkey = key[:3] + (work.layers,)
                need = self._gradient_bytes(gkey, self._gradient_chunk(gkey, work.n_points))
                if self.conv.bytes + used + need <= self.budget or not inflight:
                    grads.popleft()
                    stream = free_streams.pop()
                    used += need

                    def job(work=work, stream=stream):
                        out = self._gradient(work, stream)
                        work.t_grad_done = time.perf_counter() - t0
                        return out
                    grad_job = (self._pool.submit(job), work, stream, need)
                    started = True
            # start launches while streams and memory allow: oldest bucket first
            for key in sorted([k for k in queues if queues[k]], key=lambda k: order.get(k, 0.0)):
                if not free_streams:
                    break
                q = queues[key]
                items = self._take(q)
                while len(items) > 1 and self.conv.bytes + used + self._launch_bytes(key, items) > self.budget:
                    items = items[:len(items) // 2]
                if self.conv.bytes + used + self._launch_bytes(key, items) > self.budget and inflight:
                    continue
                taken = set(map(id, items))
                queues[key] = q = collections.deque(it for it in q if id(it) not in taken)
                P = len(items)
                if not q:
                    order.pop(key, None)
                stream = free_streams.pop()
                ln = self._launch(key, items, stream)
                used += ln['bytes']
                inflight.append(ln)
                self.stats['launches'] += 1
                self.stats['launched_points'] += P
                started = True
            if not started and not finished:
                if inflight or grad_job is not None:
                    time.sleep(0.0005)
                elif pending and realtime:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        # hand the stream's cached workspaces back: launches of other sizes later (another
        # server, a prepared run) would otherwise find the device full of blocks they cannot
        # reuse, and kernels whose scratch the HIP runtime allocates on demand would fail
        torch.cuda.synchronize(self.device)
        torch.cuda.empty_cache()
        lat = np.array(latencies) if latencies else np.zeros(1)
        report = {'requests': len(reqs), 'points': solved_points, 'wall_s': wall,
                  'solves_per_s': solved_points / wall if wall > 0 else 0.0,
                  'p50_latency_s': float(np.percentile(lat, 50)),
                  'p99_latency_s': float(np.percentile(lat, 99)),
                  'launches': self.stats['launches'],
                  'mean_points_per_launch': self.stats['launched_points'] / max(1, self.stats['launches']),
                  'gradient_requests': self.stats['gradient_requests'],
                  'conv_cache_hits': self.conv.hits, 'conv_cache_misses': self.conv.misses,
                  'hbm_budget_bytes': self.budget}
        return results, report


def poisson_arrivals(count, rate, seed=0):
    """Arrival times (s) of a Poisson process of ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))
