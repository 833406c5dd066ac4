#!/usr/bin/env python3
"""Benchmark: RCWA solves/s at orders (9,9), complex FP64, batched vs plain launch, on one MI355X.

Workload (synthetic, seeded): 4 layers of 128x128 continuous-valued permittivity rasters
(eps U[1,12]), 200 nm each, square period 1000 nm, air over an eps 2.25 substrate; orders
(9,9) = 361 harmonics, 722x722 complex128 layer eigenproblems; 64 normal-incidence p-polarized
points at wavelengths evenly spread over [500, 900] nm.

A step is one batched solve of all 64 points from the device-resident rasters: Li-rule
convolution-matrix assembly, 256 layer eigensolves, gap-embedded layer S-matrices, the cascade
and the per-order efficiencies, in one launch sequence. ``value`` is points solved per second
(one solve = one (problem, wavelength, angle, polarization) point). The plain form solves the
same points one point per launch sequence; its rate and the ratio are reported beside it.

The gradient leg differentiates the sum of T over the same 64 points with respect to every
raster pixel and thickness (forward with checkpoint, then the hand-written reverse pass).

The serving leg drives the continuous-batching server with a Poisson stream of requests at
the same orders (1-4 layers, 1-8 wavelengths each, arriving while earlier ones are on the
device; 30% also ask for the raster gradient) and reports its p50/p99 request latency and
served solves/s.

Two more workloads run as legs of their own: the crossed grating (3 layers of 256x256 binary
rasters at orders (7,7), 512 points over wavelength, angle and polarization in one launch
sequence, with every point's energy balance) and, on request (--uneven-problems), a batch of
problems of different order windows and layer counts served at once, size-bucketed.

Also reported: the dominant kernel's roofline (algorithmic flops per launch / average launch
time, from HIP events recorded on the launching stream during the timed steps) and the NumPy
FP64 oracle timed on the host cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


KERNEL_SYMBOLS = {'hqr': 'k_hqr_ms', 'gehrd': 'k_gehrd', 'unghr': 'k_unghr', 'trevc': 'k_trevc',
                  'zgemm': 'k_zgemm', 'getrf_panel': 'k_getrf_panel'}


def _traffic(name, orders, layers, points):
    """HBM bytes per launch of kernel ``name`` from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE and WRITE_SIZE passes, calibrated by a known fill and scale-copy), or
    None when no summary matches the workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        'rcwa_orders9_64pts_traffic.json')
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None, None
    w = data.get('workload', {})
    if w.get('orders') != [orders, orders] or w.get('layers') != layers or w.get('points') != points:
        return None, None
    k = data.get('kernels', {}).get(KERNEL_SYMBOLS.get(name, name))
    if not k:
        return None, None
    return float(k['bytes_per_launch']), 'profiles/rcwa_orders9_64pts_hbm_traffic.md'


def _blas():
    """(vendor string, threads) of the BLAS NumPy calls on this host."""
    try:
        from threadpoolctl import threadpool_info
        info = [t for t in threadpool_info() if t.get('user_api') == 'blas']
    except Exception:
        info = []
    if not info:
        return 'unknown', os.cpu_count()
    t = info[0]
    return f"{t.get('internal_api')} {t.get('version')} ({t.get('threading_layer')})", int(t['num_threads'])


def _cpu_baseline(fn, calls, unit_per_call, sample):
    """The CPU baseline of one workload: one untimed warm-up call, then the median of ``calls``
    timed calls (time.perf_counter); ``fn(i)`` runs call i and returns the number of solves."""
    fn(0)
    times, solves = [], []
    for i in range(calls):
        t0 = time.perf_counter()
        solves.append(fn(i))
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    vendor, threads = _blas()
    return {'value': round(float(np.median(solves)) / med, 6), 'unit': unit_per_call,
            'cores': threads, 'kind': 'port', 'blas': vendor, 'host_cpus': os.cpu_count(),
            'sample': f'{sample}; median of {calls} calls after one warm-up, {med * 1e3:.4g} ms per '
                      f'call, {sum(times):.2f} s total'}


def _measured_peaks(torch, dev):
    """Measured FP64 matrix and HBM copy rates of this device beside the datasheet peaks: the
    vendor DGEMM (torch.matmul on 8192 x 8192 float64), and a 4 GiB device-to-device copy."""
    out = {}
    a = torch.randn(8192, 8192, dtype=torch.float64, device=dev)
    b = torch.randn(8192, 8192, dtype=torch.float64, device=dev)
    torch.matmul(a, b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        c = torch.matmul(a, b)
    torch.cuda.synchronize()
    out['dgemm_vendor_tflops'] = round(5 * 2 * 8192 ** 3 / (time.perf_counter() - t0) / 1e12, 2)
    del a, b, c
    n = (4 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    out['copy_gbs'] = round(5 * 2 * n * 8 / (time.perf_counter() - t0) / 1e9, 1)
    del src, dst
    out['datasheet_fp64_matrix_tflops'] = PEAK_FP64_TFLOPS
    out['datasheet_hbm_gbs'] = PEAK_HBM_GBS
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--points', type=int, default=64)
    ap.add_argument('--plain-points', type=int, default=2)
    ap.add_argument('--cpu-points', type=int, default=5)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--orders', type=int, default=9)
    ap.add_argument('--grad-points', type=int, default=64,
                    help='points of the gradient leg (0 skips it)')
    ap.add_argument('--serve-requests', type=int, default=48)
    ap.add_argument('--serve-rate', type=float, default=4.0, help='request arrivals per second')
    ap.add_argument('--crossed-points', type=int, default=512,
                    help='points of the crossed-grating leg (orders (7,7), 3 layers of 256x256; 0 skips)')
    ap.add_argument('--uneven-problems', type=int, default=512,
                    help='problems of the uneven-batch leg (first of the seeded 2048; 0 skips)')
    ap.add_argument('--sweep-runs', type=int, default=5,
                    help='thickness-sweep runs of the prepared (9,9) problem on resident modes (0 skips)')
    ap.add_argument('--cfg5-requests', type=int, default=12,
                    help='prefix of the seeded serving stream (orders 5..15, 1-10 layers, 64^2-256^2, '
                         '1-64 wavelengths, 30%% gradient) served in the cfg5 leg (0 skips)')
    ap.add_argument('--cfg5-rate', type=float, default=0.25,
                    help='cfg5 request arrivals per second (about what the device sustains)')
    ap.add_argument('--serve-grad-fraction', type=float, default=0.3,
                    help='fraction of served requests that also ask for the raster gradient')
    ap.add_argument('--autotune', type=int, default=1,
                    help='tune the GEMM tile shapes of the benchmark steps before timing (0: defaults)')
    ap.add_argument('--small-runs', type=int, default=50,
                    help='runs of the 1D orders-(10,0) grating, eager vs captured graph (0 skips)')
    ap.add_argument('--copy-calibration-mib', type=int, default=0,
                    help='run one device copy of this size first (HBM counter calibration)')
    ap.add_argument('--no-peaks', action='store_true', help='skip the measured DGEMM / copy peaks')
    ap.add_argument('--field-samples', type=int, default=64,
                    help='fields leg: x and y samples per cell (0 skips the leg); 16 heights per layer')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); the points of a batch are '
              'independent, so a multi-GPU run would shard them with no exchange -- not built '
              'in this engine.', file=sys.stderr)
        sys.exit(2)

    import torch
    from wavebit_amd import _native, workloads
    from wavebit_amd.rcwa import _engine as eng

    dev = torch.device('cuda:0')
    M = args.orders
    wl = workloads.continuous_stack(wavelengths=args.points, orders=(M, M))
    hu, hv = wl['orders']
    L = wl['rasters'].shape[0]
    P = len(wl['points'])
    N = (2 * hu + 1) * (2 * hv + 1)

    def records(points):
        recs = []
        for p in points:
            pt = eng.Point(stack=0, wavelength=p['wavelength'], theta=p['theta'], phi=p['phi'],
                           polarization=p['polarization'], eps_top=wl['eps_top'],
                           eps_bottom=wl['eps_bottom'], period=wl['period'],
                           thickness=tuple(wl['thickness']))
            recs.append(eng.point_record(pt, eng.PATH_2D))
        return torch.tensor(np.array(recs), dtype=torch.float64, device=dev)

    if args.copy_calibration_mib > 0:
        # a fill (writes the buffer once) and a scale-copy (reads it once, writes it once) of
        # known size: the HBM counters of a profiler run are scaled by what they read here
        n_el = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n_el, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    eps_d = torch.from_numpy(wl['rasters']).to(dev)
    pts_d = records(wl['points'])
    thick_d = torch.tensor(np.tile(wl['thickness'], (P, 1)), dtype=torch.float64, device=dev)
    idx_d = torch.tensor(np.tile(np.arange(L, dtype=np.int32), (P, 1)), device=dev)
    ws = torch.empty(_native.solve_workspace_bytes(P, L, hu, hv, _native.MODE_2D), dtype=torch.uint8,
                     device=dev)

    def step(pts, thick, idx, workspace):
        epx, epy, epz, info = _native.assemble_li(eps_d, hu, hv)
        return _native.solve(pts, thick, idx, epx, epy, epz, hu, hv, _native.MODE_2D, workspace)

    for _ in range(args.warmup):
        step(pts_d, thick_d, idx_d, ws)
    torch.cuda.synchronize()
    # GEMM tile shapes: record the shapes of one batched step and one plain (one-point) step,
    # then time every tile configuration on each and keep the fastest (results are bitwise the
    # same for every configuration)
    tuning = None
    ws1 = None
    if args.autotune:
        _native.gemm_tuning_record(True)
        step(pts_d, thick_d, idx_d, ws)
        if args.plain_points > 0:
            ws1 = torch.empty(_native.solve_workspace_bytes(1, L, hu, hv, _native.MODE_2D),
                              dtype=torch.uint8, device=dev)
            step(pts_d[:1].contiguous(), thick_d[:1].contiguous(), idx_d[:1].contiguous(), ws1)
        _native.gemm_tuning_record(False)
        t_tune = time.perf_counter()
        n_tuned = _native.gemm_autotune(dev, reps=2)
        t_tune = time.perf_counter() - t_tune
        cfg_names = _native.gemm_tuning_configs()
        counts = {}
        for ln in _native.gemm_tuning_table().splitlines():
            nm_ = cfg_names[int(ln.split()[-1])]
            counts[nm_] = counts.get(nm_, 0) + 1
        tuning = {'shapes_tuned': n_tuned, 'configs_chosen': counts, 'tuning_s': round(t_tune, 2)}
        step(pts_d, thick_d, idx_d, ws)
        torch.cuda.synchronize()
    names = list(_native.TIMER_IDS)
    _native.timing_enable(True)
    for n_ in names:
        _native.timing_read(n_, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step(pts_d, thick_d, idx_d, ws)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _native.timing_enable(False)
    kernels = {}
    for n_ in names:
        ms, cnt, fl = _native.timing_read(n_, reset=True)
        if cnt:
            kernels[n_] = dict(total_ms=ms, launches=cnt, flops=fl)
    eff, amp, pe, flags = out
    n_flagged = int(((flags & 3) != 0).sum())          # failed eigensolves or factorizations
    batched_rate = P * args.steps / dt

    # per-stage rates from one more (untimed-for-value) step with whole stages bracketed by HIP
    # events: the layer eigensolves and the S-matrix work (Omega^2, V, layer S, cascade) against
    # the FP64 peak, the assembly gathers (raster -> Toeplitz blocks) against the HBM peak
    _native.timing_enable(True, stages=True)
    for n_ in names:
        _native.timing_read(n_, reset=True)
    step(pts_d, thick_d, idx_d, ws)
    torch.cuda.synchronize()
    _native.timing_enable(False)
    stages = {}
    for key, nm in (('eig', 'stage_eig'), ('smatrix', 'stage_smat')):
        ms, cnt, fl = _native.timing_read(nm, reset=True)
        if cnt:
            tf = fl / (ms * 1e-3) / 1e12
            stages[key] = {'ms': round(ms, 3), 'flops': fl, 'achieved_tflops': round(tf, 3),
                           'peak_tflops': PEAK_FP64_TFLOPS, 'frac': round(tf / PEAK_FP64_TFLOPS, 4)}
    ms, cnt, _ = _native.timing_read('assemble', reset=True)
    if cnt:
        Ny, Nx = wl['rasters'].shape[1:]
        # algorithmic bytes: the complex rasters read once, three N x N complex blocks written
        asm_bytes = L * Ny * Nx * 16 + 3 * L * N * N * 16
        gbs = asm_bytes / (ms * 1e-3) / 1e9
        stages['assembly_gather'] = {'ms': round(ms, 4), 'bytes': asm_bytes,
                                     'achieved_gbs': round(gbs, 1), 'peak_gbs': PEAK_HBM_GBS,
                                     'frac': round(gbs / PEAK_HBM_GBS, 4)}

    # plain: one point per launch sequence (--plain-points 0 skips it, e.g. under a profiler)
    plain_rate = None
    if args.plain_points > 0:
        plain_n = args.plain_points
        if ws1 is None:
            ws1 = torch.empty(_native.solve_workspace_bytes(1, L, hu, hv, _native.MODE_2D),
                              dtype=torch.uint8, device=dev)
        step(pts_d[:1].contiguous(), thick_d[:1].contiguous(), idx_d[:1].contiguous(), ws1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(plain_n):
            step(pts_d[i:i + 1].contiguous(), thick_d[i:i + 1].contiguous(),
                 idx_d[i:i + 1].contiguous(), ws1)
        torch.cuda.synchronize()
        plain_rate = plain_n / (time.perf_counter() - t0)

    # dominant kernel roofline
    dom = max(kernels.items(), key=lambda kv: kv[1]['total_ms']) if kernels else None
    roofline = None
    if dom is not None:
        name, k = dom
        avg_s = k['total_ms'] / k['launches'] / 1e3
        flops_per_launch = k['flops'] / k['launches']
        achieved = flops_per_launch / avg_s / 1e12 if avg_s > 0 else 0.0
        traffic, traffic_src = _traffic(name, M, L, P)
        roofline = {'bound': 'mfma', 'kernel': name, 'achieved': round(achieved, 4),
                    'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
                    'frac': round(achieved / PEAK_FP64_TFLOPS, 5), 'traffic': traffic,
                    'traffic_source': traffic_src,
                    'avg_launch_ms': round(k['total_ms'] / k['launches'], 4),
                    'flops_per_launch': flops_per_launch,
                    'share_of_step': round(k['total_ms'] / 1e3 / dt, 4)}
    # every timed kernel's rate against the FP64 peak (algorithmic flops / HIP-event time)
    kernel_rates = {nm: {'avg_launch_ms': round(k['total_ms'] / k['launches'], 4),
                         'achieved_tflops': round(k['flops'] / (k['total_ms'] * 1e-3) / 1e12, 3),
                         'frac': round(k['flops'] / (k['total_ms'] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
                         'share_of_step': round(k['total_ms'] / 1e3 / dt, 4)}
                    for nm, k in kernels.items() if k['flops'] > 0 and k['total_ms'] > 0}

    # CPU baseline (oracle, NumPy FP64) on a bounded sample; parity of the same points.
    # One untimed warm-up solve, then the median per-point time of the sample.
    cpu = None
    oracle_err = None
    if not args.no_cpu and args.cpu_points > 0:
        from oracle import rcwa_oracle as orc
        n_cpu = min(args.cpu_points, P)
        effh, amph, peh = eff.cpu().numpy(), amp.cpu().numpy(), pe.cpu().numpy()

        def oracle_point(i):
            p = wl['points'][i]
            return orc.solve_raster(wl['rasters'], wl['thickness'], wl['eps_top'], wl['eps_bottom'],
                                    wl['period'], p['wavelength'], p['theta'], p['phi'],
                                    p['polarization'], (hu, hv))
        oracle_point(0)
        times, errs = [], []
        stride = max(1, P // n_cpu)
        for i in range(0, stride * n_cpu, stride):
            t0 = time.perf_counter()
            ref = oracle_point(i)
            times.append(time.perf_counter() - t0)
            got = eng.unpack(effh, amph, peh, i, hu, hv, eng.PATH_2D)
            errs.append(max(float(np.abs(got['incident'][k] - ref['incident'][k]).max())
                            for k in ('R', 'T')))
        med = float(np.median(times))
        vendor, threads = _blas()
        cpu = {'value': round(1.0 / med, 6), 'unit': 'solves/s', 'cores': threads,
               'kind': 'port', 'blas': vendor, 'host_cpus': os.cpu_count(),
               'sample': f'{n_cpu} points of the same workload (4 layers, orders (9,9), every '
                         f'{stride}th wavelength) through the NumPy FP64 oracle after one warm-up; '
                         f'median {med:.2f} s per point, {sum(times):.1f} s total'}
        oracle_err = max(errs)

    # gradient: sum of T over the points, differentiated with respect to every raster pixel and
    # thickness by the hand-written reverse pass (forward with checkpoint + adjoint + assembly
    # adjoint), the same 4-layer (9,9) stack
    grad = None
    if args.grad_points > 0:
        from wavebit_amd.rcwa import adjoint
        Pg = min(args.grad_points, P)
        pts_g = [eng.Point(stack=0, wavelength=p['wavelength'], theta=p['theta'], phi=p['phi'],
                           polarization=p['polarization'], eps_top=wl['eps_top'],
                           eps_bottom=wl['eps_bottom'], period=wl['period'],
                           thickness=tuple(wl['thickness'])) for p in wl['points'][:Pg]]
        eps_g = torch.tensor(wl['rasters'], device=dev, requires_grad=True)
        thick_g = torch.tensor(wl['thickness'], device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eff_g, _, _, flags_g = adjoint.solve(eps_g, thick_g, pts_g, hu, hv, dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (eff_g[:, 0, 2] + eff_g[:, 0, 3]).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grad = {'points': Pg, 'forward_with_checkpoint_s': round(t1 - t0, 4),
                'backward_s': round(t2 - t1, 4),
                'gradient_solves_per_s': round(Pg / (t2 - t0), 4),
                'backward_over_forward': round((t2 - t1) / (t1 - t0), 3),
                'grad_norm_eps': float(eps_g.grad.norm()), 'grad_thickness': thick_g.grad.tolist(),
                'flagged_points': int(((flags_g & 3) != 0).sum()),
                'objective': 'sum of T over the points; d/d every pixel of 4 x 128 x 128 and the '
                             '4 thicknesses'}
        del eff_g, eps_g, thick_g

    # serving: a Poisson request stream through the continuous-batching server
    serve_report = None
    if args.serve_requests > 0:
        from wavebit_amd.rcwa import serve
        specs = workloads.serving_requests(args.serve_requests, seed=11, m_range=(M, M),
                                           l_range=(1, 4), sizes=(64, 128), wl_range=(1, 8),
                                           grad_fraction=args.serve_grad_fraction)
        arrivals = serve.poisson_arrivals(len(specs), rate=args.serve_rate, seed=12)
        reqs = []
        for i, (s, t) in enumerate(zip(specs, arrivals)):
            prob, num = workloads.to_problem(s)
            reqs.append(serve.Request(problem=prob, numerics=num, arrival=float(t), id=i,
                                      gradient=bool(s['gradient'])))
        srv = serve.Server(device=dev, streams=3, max_points_per_launch=64)
        results, rep = srv.serve(reqs, realtime=True)
        failed = sum(1 for r in results.values() if isinstance(r, Exception))
        serve_report = {'requests': rep['requests'], 'arrival_rate_per_s': args.serve_rate,
                        'points': rep['points'], 'solves_per_s': round(rep['solves_per_s'], 4),
                        'p50_latency_s': round(rep['p50_latency_s'], 4),
                        'p99_latency_s': round(rep['p99_latency_s'], 4),
                        'launches': rep['launches'],
                        'gradient_requests': rep['gradient_requests'],
                        'mean_points_per_launch': round(rep['mean_points_per_launch'], 2),
                        'failed_requests': failed,
                        'stream': f'orders ({M},{M}), 1-4 layers of 64x64/128x128 rasters eps '
                                  'U[1,16]+iU[0,1], 1-8 wavelengths in [400,1000] nm, Poisson arrivals, '
                                  f'{args.serve_grad_fraction:.0%} also asking for the raster gradient'}

    # resident eigensystems: the prepared (9,9) problem run again with new thicknesses takes its
    # layer modes from the last run (no eigensolve), the S-matrix work and cascade only
    sweep = None
    if args.sweep_runs > 0:
        import wavebit_amd
        from wavebit_amd import rcwa

        def prob4(thick):
            layers = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(thick, wl['rasters'])]
            return rcwa.Problem(
                stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(wl['eps_top']),
                                 bottom=wavebit_amd.Material(wl['eps_bottom']), period=wl['period']),
                excitation=rcwa.Excitation(wavelength=[p['wavelength'] for p in wl['points']],
                                           sources=(rcwa.PlaneWave(polarization='p', theta=0.0, phi=0.0),)),
                units=wavebit_amd.Units(length='nm'))
        num4 = rcwa.Numerics(orders=(hu, hv))
        prep4 = rcwa.prepare(prob4(wl['thickness']), numerics=num4, backend='torch', device='cuda:0')
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prep4.run(prob4(wl['thickness']))
        first_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for i in range(args.sweep_runs):
            res4 = prep4.run(prob4(np.asarray(wl['thickness']) * (1.0 + 0.05 * (i + 1))))
        sweep_s = (time.perf_counter() - t0) / args.sweep_runs
        fresh4 = rcwa.run(prob4(np.asarray(wl['thickness']) * (1.0 + 0.05 * args.sweep_runs)),
                          numerics=num4, backend='torch', device='cuda:0')
        sweep = {'workload': 'the benchmark problem through rcwa.prepare / Prepared.run (64 '
                             'wavelengths, host results), then runs with every thickness scaled',
                 'first_run_s': round(first_s, 4), 'runs': args.sweep_runs,
                 'thickness_run_s': round(sweep_s, 4),
                 'thickness_sweep_solves_per_s': round(P / sweep_s, 3),
                 'equal_to_fresh_run': bool(np.array_equal(res4.R, fresh4.R)
                                            and np.array_equal(res4.T, fresh4.T))}
        del prep4, res4, fresh4

    # in-layer fields of the benchmark problem: the run that keeps the field state (layer modes and
    # mode amplitudes), then Result.field on a field_samples^2 x (16 per layer) grid of every
    # wavelength; the flux through every sampled height checked against T (the stack is lossless)
    fields_leg = None
    if args.field_samples > 0:
        import wavebit_amd
        from wavebit_amd import rcwa
        from wavebit_amd.units import Z0
        layers_f = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(wl['thickness'], wl['rasters'])]
        prob_f = rcwa.Problem(
            stack=rcwa.Stack(layers=layers_f, top=wavebit_amd.Material(wl['eps_top']),
                             bottom=wavebit_amd.Material(wl['eps_bottom']), period=wl['period']),
            excitation=rcwa.Excitation(wavelength=[p['wavelength'] for p in wl['points']],
                                       sources=(rcwa.PlaneWave(polarization='p', theta=0.0, phi=0.0),)),
            outputs=rcwa.Outputs(fields=True), units=wavebit_amd.Units(length='nm'))
        num_f = rcwa.Numerics(orders=(hu, hv))
        rcwa.run(prob_f, numerics=num_f, backend='torch', device='cuda:0')     # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_f = rcwa.run(prob_f, numerics=num_f, backend='torch', device='cuda:0')
        field_run_s = time.perf_counter() - t0
        ns, nz_l = args.field_samples, 16
        res_f.field(x=ns, y=ns, z=2)                                             # warm
        torch.cuda.synchronize()
        _native.timing_enable(True)
        _native.timing_read('zgemm', reset=True)
        t0 = time.perf_counter()
        vals_dev, _ = res_f._fields.sample(ns, ns, nz_l)          # the sampling, values on the device
        torch.cuda.synchronize()
        device_s = time.perf_counter() - t0
        _native.timing_enable(False)
        g_ms, g_cnt, g_fl = _native.timing_read('zgemm', reset=True)
        del vals_dev
        t0 = time.perf_counter()
        fs = res_f.field(x=ns, y=ns, z=nz_l)                        # the public call, host arrays
        sample_s = time.perf_counter() - t0
        n_samples = P * L * nz_l * ns * ns
        v = fs.values[:, :, :-1, :-1]
        flux = 0.5 * np.real(v[..., 0] * np.conj(v[..., 4]) - v[..., 1] * np.conj(v[..., 3])).mean(axis=(2, 3))
        n_top = np.sqrt(complex(wl['eps_top'])).real
        incident = 0.5 * n_top / Z0
        T_tot = res_f.T.sum(axis=(1, 2))
        lossless = all(np.isrealobj(r) or not np.any(np.imag(r)) for r in wl['rasters'])
        fields_leg = {
            'workload': f'the benchmark problem with Outputs(fields=True), then Result.field on '
                        f'{ns} x {ns} points x {nz_l} heights per layer for all {P} wavelengths',
            'field_run_s': round(field_run_s, 4),
            'field_run_solves_per_s': round(P / field_run_s, 3),
            'sample_device_s': round(device_s, 4), 'samples': n_samples,
            'device_samples_per_s': round(n_samples / device_s, 1),
            'field_call_s': round(sample_s, 4),
            'field_call_note': 'Result.field including the copy of every value to host memory',
            'zgemm_ms': round(g_ms, 3), 'zgemm_launches': g_cnt,
            'zgemm_tflops_per_s': round(g_fl / (g_ms * 1e-3) / 1e12, 3) if g_ms > 0 else None,
            'zgemm_frac_of_fp64_peak': round(g_fl / (g_ms * 1e-3) / 78.6e12, 4) if g_ms > 0 else None,
            'flux_vs_T_max_rel_err': (float(np.max(np.abs(flux + T_tot[:, None] * incident)) / incident)
                                      if lossless else None)}
        del res_f, fs, v

    # cfg5: a prefix of the seeded serving stream at its stated distribution
    cfg5 = None
    if args.cfg5_requests > 0:
        from wavebit_amd.rcwa import serve
        specs5 = workloads.serving_requests(args.cfg5_requests, seed=0)
        arrivals5 = serve.poisson_arrivals(len(specs5), rate=args.cfg5_rate, seed=1)
        reqs5 = []
        for i, (s5, t5) in enumerate(zip(specs5, arrivals5)):
            prob5, num5 = workloads.to_problem(s5)
            reqs5.append(serve.Request(problem=prob5, numerics=num5, arrival=float(t5), id=i,
                                       gradient=bool(s5['gradient'])))
        srv5 = serve.Server(device=dev, streams=3, max_points_per_launch=64)
        res5, rep5 = srv5.serve(reqs5, realtime=True, progress=lambda i, lat: print(
            f'cfg5 request {i} done, latency {lat:.2f} s', file=sys.stderr, flush=True))
        srv5.close()
        cfg5 = {'stream': 'workloads.serving_requests(seed=0): orders per axis U{5..15}, 1-10 layers '
                          'of 64^2/128^2/256^2 rasters eps U[1,16]+iU[0,1], 1-64 wavelengths in '
                          '[400,1000] nm, 30% also asking for the raster gradient; Poisson arrivals',
                'requests': rep5['requests'], 'arrival_rate_per_s': args.cfg5_rate,
                'prefix_of': 10000,
                'orders': [s5['orders'][0] for s5 in specs5],
                'layers': [len(s5['thickness']) for s5 in specs5],
                'wavelengths': [len(s5['wavelengths']) for s5 in specs5],
                'gradient': [bool(s5['gradient']) for s5 in specs5],
                'points': rep5['points'], 'wall_s': round(rep5['wall_s'], 3),
                'solves_per_s': round(rep5['solves_per_s'], 4),
                'p50_latency_s': round(rep5['p50_latency_s'], 4),
                'p99_latency_s': round(rep5['p99_latency_s'], 4),
                'launches': rep5['launches'], 'gradient_requests': rep5['gradient_requests'],
                'failed_requests': sum(1 for v in res5.values() if isinstance(v, Exception)),
                'errors': sorted({type(v).__name__ for v in res5.values() if isinstance(v, Exception)})}
        del srv5, res5

    # the launch-bound end: the 1D binary grating at orders (10,0) through Prepared.run, eager
    # launch sequence vs the captured HIP graph it replays from its second run on
    small = None
    if args.small_runs > 0:
        import wavebit_amd
        from wavebit_amd import rcwa
        g1 = workloads.binary_grating_1d()
        p1 = g1['points'][0]

        def prob1(wl_nm):
            layers = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(g1['thickness'], g1['rasters'])]
            return rcwa.Problem(
                stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(g1['eps_top']),
                                 bottom=wavebit_amd.Material(g1['eps_bottom']), period=g1['period']),
                excitation=rcwa.Excitation(wavelength=wl_nm, sources=(rcwa.PlaneWave(
                    polarization='s', theta=p1['theta'], phi=None),)),
                units=wavebit_amd.Units(length='nm'))
        num1 = rcwa.Numerics(orders=g1['orders'])
        timings = {}
        for mode in ('eager', 'graph'):
            prep = rcwa.prepare(prob1(p1['wavelength']), numerics=num1, backend='torch', device='cuda:0')
            if mode == 'eager':
                prep._cache.graph = None
            for _ in range(3):
                prep.run(prob1(p1['wavelength']))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(args.small_runs):
                res1 = prep.run(prob1(p1['wavelength'] + 0.01 * i))
            torch.cuda.synchronize()
            timings[mode] = (time.perf_counter() - t0) / args.small_runs
        # the device solve alone (engine level: inputs already on the device, one sync per solve)
        e_conv = eng.ConvCache(dev).get('g1', torch.from_numpy(np.ascontiguousarray(g1['rasters'])).to(dev),
                                        g1['orders'][0], g1['orders'][1])
        bucket = eng.Bucket(hu=g1['orders'][0], hv=g1['orders'][1], path=eng.PATH_TE, n_layers=1)
        bucket.points.append(eng.Point(stack=0, wavelength=p1['wavelength'], theta=p1['theta'], phi=None,
                                       polarization='s', eps_top=g1['eps_top'], eps_bottom=g1['eps_bottom'],
                                       period=g1['period'], thickness=tuple(g1['thickness'])))
        dev_t = {}
        for mode in ('eager', 'graph'):
            slot = None if mode == 'eager' else [None, None]
            for _ in range(3):
                eng.run_bucket(bucket, [e_conv], dev, graph_slot=slot)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.small_runs):
                eng.run_bucket(bucket, [e_conv], dev, graph_slot=slot)
                torch.cuda.synchronize()
            dev_t[mode] = (time.perf_counter() - t0) / args.small_runs
        # the configuration itself: exactly 900 nm, against the oracle
        res900 = rcwa.run(prob1(p1['wavelength']), numerics=num1, backend='torch', device='cuda:0')
        from oracle import rcwa_oracle as orc

        def oracle1(i=0):
            orc.solve_raster(g1['rasters'], g1['thickness'], g1['eps_top'], g1['eps_bottom'],
                             g1['period'], p1['wavelength'], p1['theta'], None, 's', g1['orders'])
            return 1
        ref1 = orc.solve_raster(g1['rasters'], g1['thickness'], g1['eps_top'], g1['eps_bottom'],
                                g1['period'], p1['wavelength'], p1['theta'], None, 's',
                                g1['orders'])['incident']
        centre = g1['orders'][0]
        small = {'workload': '1D binary grating, period 700 nm, one 300 nm layer, orders (10,0) = 21 '
                             'harmonics, one wavelength per run (timed runs step the wavelength by '
                             '0.01 nm; the values below are at exactly 900 nm)',
                 'runs': args.small_runs,
                 'lifecycle_eager_ms': round(timings['eager'] * 1e3, 4),
                 'lifecycle_graph_ms': round(timings['graph'] * 1e3, 4),
                 'solve_eager_ms': round(dev_t['eager'] * 1e3, 4),
                 'solve_graph_ms': round(dev_t['graph'] * 1e3, 4),
                 'wavelength_nm': p1['wavelength'],
                 'R_sum': float(res900.R.sum()), 'T_sum': float(res900.T.sum()),
                 'T_00': float(res900.T[0, 0, centre]),
                 'solve_path': res900.diagnostics['solve_path'][0],
                 'oracle_max_abs_err_RT': float(max(np.abs(res900.R[0] - ref1['R']).max(),
                                                    np.abs(res900.T[0] - ref1['T']).max())),
                 'cpu_baseline': None if args.no_cpu else _cpu_baseline(
                     oracle1, 500, 'solves/s', 'the NumPy FP64 oracle on the same point')}

    # the crossed grating: 3 layers of 256x256 binary rasters, orders (7,7) = 225 harmonics,
    # points over wavelength, angle and polarization in one batched launch sequence (assembly
    # included); lossless, so every point's R + T - 1 is reported
    crossed = None
    if args.crossed_points > 0:
        w2 = workloads.crossed_grating_2d(points=args.crossed_points)
        h2 = w2['orders']
        r2 = torch.from_numpy(w2['rasters']).to(dev)

        def crossed_step(points):
            b2 = eng.Bucket(hu=h2[0], hv=h2[1], path=eng.PATH_2D, n_layers=r2.shape[0])
            for p in points:
                b2.points.append(eng.Point(stack=0, wavelength=p['wavelength'], theta=p['theta'],
                                           phi=p['phi'], polarization=p['polarization'],
                                           eps_top=w2['eps_top'], eps_bottom=w2['eps_bottom'],
                                           period=w2['period'], thickness=tuple(w2['thickness'])))
            conv2 = [_native.assemble_li(r2, h2[0], h2[1])[:3]]
            return eng.run_bucket(b2, conv2, dev)
        crossed_step(w2['points'][:8])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o2 = crossed_step(w2['points'])
        torch.cuda.synchronize()
        dt2 = time.perf_counter() - t0
        e2, a2, p2, f2 = (t.cpu().numpy() for t in o2[:4])
        err2 = max(abs(float(u['R'].sum() + u['T'].sum()) - 1.0)
                   for u in (eng.unpack(e2, a2, p2, i, h2[0], h2[1], eng.PATH_2D)['incident']
                             for i in range(len(w2['points']))))
        crossed = {'workload': '3 layers of 256x256 binary {1, 12.25} rasters, period 1000 nm, orders '
                               '(7,7) = 225 harmonics (n=450), wavelength U[400,800] nm, theta U[0,30] '
                               'deg, phi U[0,360) deg, s/p, one batched launch sequence',
                   'points': len(w2['points']), 's': round(dt2, 4),
                   'solves_per_s': round(len(w2['points']) / dt2, 3),
                   'max_abs_R_plus_T_minus_1': err2, 'flagged_points': int(((f2 & 3) != 0).sum()),
                   'grazing_layer_solves': int(((f2 & _native.FLAG_GRAZING) != 0).sum()),
                   'near_defective_pair_layer_solves': int(((f2 & _native.FLAG_EP) != 0).sum()),
                   'loud_pair_layer_solves': int(((f2 & _native.FLAG_EP_LOUD) != 0).sum())}
        if not args.no_cpu:
            from oracle import rcwa_oracle as orc
            sample2 = list(range(0, len(w2['points']), max(1, len(w2['points']) // 5)))[:5]

            def oracle2(i):
                p = w2['points'][sample2[i % len(sample2)]]
                return orc.solve_raster(w2['rasters'], w2['thickness'], w2['eps_top'],
                                        w2['eps_bottom'], w2['period'], p['wavelength'],
                                        p['theta'], p['phi'], p['polarization'], h2)
            errs2 = []
            for j, i in enumerate(sample2):
                ref2 = oracle2(j)['incident']
                got2 = eng.unpack(e2, a2, p2, i, h2[0], h2[1], eng.PATH_2D)['incident']
                errs2.append(max(float(np.abs(got2[k] - ref2[k]).max()) for k in ('R', 'T')))
            crossed['oracle_max_abs_err_RT'] = max(errs2)
            crossed['oracle_points'] = sample2
            crossed['cpu_baseline'] = _cpu_baseline(
                lambda i: (oracle2(i), 1)[1], len(sample2), 'solves/s',
                f'{len(sample2)} points of the crossed grating (every '
                f'{max(1, len(w2["points"]) // 5)}th) through the NumPy FP64 oracle')

    # the uneven batch: problems of different order windows and layer counts served at once,
    # size-bucketed by (orders, layers), one launch sequence per bucket on three streams
    uneven = None
    if args.uneven_problems > 0:
        from wavebit_amd.rcwa import serve
        specs3 = workloads.uneven_problems()[:args.uneven_problems]
        reqs3 = []
        for i, s3 in enumerate(specs3):
            prob3, num3 = workloads.to_problem(s3)
            reqs3.append(serve.Request(problem=prob3, numerics=num3, arrival=0.0, id=i))
        srv3 = serve.Server(device=dev, streams=3, max_points_per_launch=64)
        res3, rep3 = srv3.serve(reqs3, realtime=False)
        uneven = {'workload': 'orders per axis U{3..12}, 1-8 layers of 128x128 two-material rasters '
                              '(eps\' U[1,16], eps\'\' U[0,2]), substrate U[1.5,4], one point each, '
                              'the first problems of the seeded 2048-problem set',
                  'problems': len(specs3), 'points': rep3['points'], 's': round(rep3['wall_s'], 3),
                  'solves_per_s': round(rep3['solves_per_s'], 3), 'launches': rep3['launches'],
                  'failed': sum(1 for v in res3.values() if isinstance(v, Exception))}
        if not args.no_cpu:
            from oracle import rcwa_oracle as orc
            small3 = [i for i, s3 in enumerate(specs3) if s3['orders'][0] <= 6][:5]

            def oracle3(j):
                s3 = specs3[small3[j % len(small3)]]
                p = s3['points'][0]
                return orc.solve_raster(s3['rasters'], s3['thickness'], s3['eps_top'],
                                        s3['eps_bottom'], s3['period'], p['wavelength'], p['theta'],
                                        p['phi'], p['polarization'], s3['orders'])
            errs3 = []
            for j, i in enumerate(small3):
                if isinstance(res3[i], Exception):
                    continue
                ref3 = oracle3(j)['incident']
                errs3.append(max(float(np.abs(res3[i].R[0] - ref3['R']).max()),
                                 float(np.abs(res3[i].T[0] - ref3['T']).max())))
            uneven['oracle_max_abs_err_RT'] = max(errs3) if errs3 else None
            uneven['oracle_problems'] = small3
            uneven['cpu_baseline'] = _cpu_baseline(
                lambda j: (oracle3(j), 1)[1], len(small3), 'solves/s',
                f'the first {len(small3)} problems with orders <= 6 per axis through the NumPy '
                'FP64 oracle (the set mean is heavier: orders up to 12)')

    peaks = None if args.no_peaks else _measured_peaks(torch, dev)
    fl_pt = workloads.flops_per_point((hu, hv), L)
    line = {
        'metric': 'RCWA solves/s @ (9,9) orders FP64 vs plain launch; p50/p99 serve latency',
        'value': round(batched_rate, 4), 'unit': 'solves/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(dt / args.steps * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'c128',
        'data': 'synthetic',
        'config': {'workload': f'{L} layers 128x128 continuous eps U[1,12], 200 nm, period 1000 nm, '
                               f'orders ({hu},{hv}) = {N} harmonics (n={2 * N}), {P} wavelengths '
                               '[500,900] nm, normal incidence, p-pol, one batched launch sequence',
                   'points_per_step': P, 'layers': L, 'orders': [hu, hv],
                   'parallelism': 'single GPU, batched points'},
        'plain_solves_per_s': round(plain_rate, 4) if plain_rate else None,
        'batched_vs_plain': round(batched_rate / plain_rate, 3) if plain_rate else None,
        'algorithmic_tflops_per_s': round(fl_pt * batched_rate / 1e12, 4),
        'roofline': roofline, 'kernel_rates': kernel_rates, 'cpu_baseline': cpu, 'stages': stages,
        'measured_peaks': peaks,
        'gradient': grad,
        'serve': serve_report,
        'thickness_sweep': sweep,
        'fields': fields_leg,
        'serve_cfg5': cfg5,
        'launch_bound': small,
        'crossed_grating': crossed,
        'uneven_batch': uneven,
        'gemm_autotune': tuning,
        'p50_latency_s': serve_report['p50_latency_s'] if serve_report else None,
        'p99_latency_s': serve_report['p99_latency_s'] if serve_report else None,
        'oracle_max_abs_err_RT': oracle_err, 'flagged_points': n_flagged,
        'kernels_ms': {k: round(v['total_ms'] / args.steps, 3) for k, v in kernels.items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
