"""Benchmark of the MI355X block-sparse backend (one GPU).

Headline: tensordot useful TFLOP/s of the DMRG two-site effective-Hamiltonian apply (four tensordots
of Env3.Heff2) at chi = 1024, U(1) spin-1/2, XXZ MPO of bond 5, float64.  A step is one Heff2 apply
(every merge, grouped GEMM and unmerge of the four contractions) on inputs resident in HBM.
Useful flops = sum over the sector GEMMs of 2 M N K (8 M N K for complex); padding and zero fill
do not count.

The same JSON line also carries: the plain per-block GPU path on the same inputs, the truncated
block SVD back to chi = 1024, the U(1)xU(1) complex transfer contraction (grouped vs per-block), the
roofline of the dominant kernel measured with HIP events, and the CPU baseline (the numpy oracle on
the host cores, a bounded sample).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-ctmrg]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def _count_flops(backend, fn, *args):
    """Run fn once with backend.dot wrapped, returning the useful flops of all sector GEMMs."""
    import torch
    flops = [0]
    orig = backend.dot

    def counting(A, B, meta, Dsize):
        cplx = A.is_complex() or B.is_complex()
        flops[0] += sum((8 if cplx else 2) * Da[0] * Da[1] * Db[1] for _, _, _, Da, _, Db in meta)
        return orig(A, B, meta, Dsize)
    backend.dot = counting
    try:
        with torch.no_grad():
            fn(*args)
    finally:
        backend.dot = orig
    return flops[0]


def _timeit(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def capture_graph(fn, warm=3):
    """Capture fn (a forward pass with static block structure, no host synchronisation) into a HIP
    graph; plans are compiled and cached by the warm-up passes so that capture records launches only."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warm):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = fn()
    torch.cuda.synchronize()
    return g, out


def _pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE, scaled by a copy of known size), or None when the summary has no such launch."""
    path = os.path.join(ROOT, 'profiles', 'heff2_dominant_gemm_pmc.txt')
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith('traffic | '):
            _, k, v = (x.strip() for x in line.split(' | '))
            if k == kernel:
                return int(v)
    return None


def _dot_launch_stats(backend, workload_fn, reps=50, top=None):
    """Per-launch duration (HIP events on the launch stream) of every grouped-GEMM launch of one
    workload pass (with ``top``: of the ``top`` launches with the most flops); returns
    [(flops, avg_ms, plan, dtype, bytes, args)] in launch order."""
    import torch
    from yastn_amd import plans as P
    calls = []
    orig = backend.dot

    def capture(A, B, meta, Dsize):
        calls.append((A, B, meta, Dsize))
        return orig(A, B, meta, Dsize)
    backend.dot = capture
    try:
        with torch.no_grad():
            workload_fn()
    finally:
        backend.dot = orig
    out = []
    stream = torch.cuda.current_stream()
    if top is not None:
        def _fl(c):
            return sum(Da[0] * Da[1] * Db[1] for _, _, _, Da, _, Db in c[2])
        keep = set(sorted(range(len(calls)), key=lambda i: -_fl(calls[i]))[:top])
        calls = [c for i, c in enumerate(calls) if i in keep]
    for A, B, meta, Dsize in calls:
        cplx = A.is_complex() or B.is_complex()
        fl = sum((8 if cplx else 2) * Da[0] * Da[1] * Db[1] for _, _, _, Da, _, Db in meta)
        isz = 16 if cplx else 8
        by = sum(isz * (Da[0] * Da[1] + Db[0] * Db[1] + Da[0] * Db[1]) for _, _, _, Da, _, Db in meta)
        for _ in range(3):
            backend.dot(A, B, meta, Dsize)
        # `reps` launches back to back, captured as one HIP graph and replayed between two events on the
        # launch stream: the average is the GPU's time per launch, free of the host's per-call cost
        # (~13 us of Python and ctypes per eager call, more than a small launch takes on the GPU)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(reps):
                backend.dot(A, B, meta, Dsize)
        g.replay()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        g.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del g
        dt = torch.promote_types(A.dtype, B.dtype)
        plan = P.dot_plan(meta, A.numel(), B.numel(), Dsize, A.device)
        out.append((fl, ms, plan, dt, by, (A, B, meta, Dsize)))
    return out


KNAME = {0: '32x32', 1: '64x64', 2: '128x64', 3: '16x64', 4: '64x16', 5: '128x128', 6: 'small-M streaming',
         7: '32x32 register-direct'}


def _kernel_desc(plan):
    return f'grouped GEMM, {KNAME.get(plan.cfg, plan.cfg)} tiles ({plan.ngemms} sector GEMMs, {plan.ntiles} tiles)'


def _pmc_sweep(backend, stats, reps=3):
    """For counter runs: after a marker dispatch, launch every dot of the step with every applicable
    tile shape (reps launches each, in the printed order), so the HBM counters of whichever shape the
    autotuner picks in a timing run are on record."""
    import torch
    from yastn_amd import _lib as L
    torch.cuda.synchronize()
    torch.cuda._sleep(1000)  # marker: the sweep's dispatches follow it
    order = []
    for _, _, plan, _, _, (A, B, meta, Dsize) in stats:
        if getattr(plan, 'part_buf', None) is not None:
            continue  # split-K launches are two kernels; not swept
        tuned = plan.cfg
        for cfg in L.TILE_DIMS:
            if not plan.applicable(cfg):
                continue
            plan.set_cfg(cfg)
            for _ in range(reps):
                backend.dot(A, B, meta, Dsize)
            order.append(_kernel_desc(plan))
        plan.set_cfg(tuned)
    torch.cuda.synchronize()
    print('pmc-sweep: ' + json.dumps({'reps': reps, 'launches': order}), flush=True)


def bench_heff2(args, out):
    import torch
    from yastn_amd import backend_hip, backend_plain
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    sync = torch.cuda.synchronize
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    AA, FL, FR, Wm = W.dmrg_heff2_setup(cfg, chi=1024, seed=0)
    flops = _count_flops(backend_hip, W.heff2, AA, FL, FR, Wm)
    with torch.no_grad():
        t_eager = _timeit(lambda: W.heff2(AA, FL, FR, Wm), args.steps, args.warmup, sync)
        g, res_g = capture_graph(lambda: W.heff2(AA, FL, FR, Wm))
        t = _timeit(g.replay, args.steps, args.warmup, sync)
        ref = W.heff2(AA, FL, FR, Wm)
        sync()
        graph_err = float((res_g.data - ref.data).abs().max())
    out['value'] = flops / t / 1e12
    out['ms_per_step'] = t * 1e3
    out['config']['useful_gflop_per_step'] = flops / 1e9
    out['config']['execution'] = 'HIP graph replay of the 4-tensordot step (captured once; same kernels per replay)'
    out['eager'] = {'ms_per_step': t_eager * 1e3, 'value': flops / t_eager / 1e12, 'graph_vs_eager_maxdiff': graph_err}
    cfgp = bs.make_config(backend_plain, 'U1', device='cuda:0')
    if not args.no_plain:  # plain per-block path on the same inputs
        AAp, FLp, FRp, Wp = (x._replace(config=cfgp) for x in (AA, FL, FR, Wm))
        with torch.no_grad():
            tp = _timeit(lambda: W.heff2(AAp, FLp, FRp, Wp), max(3, args.steps // 4), 1, sync)
        out['plain_gpu'] = {'ms_per_step': tp * 1e3, 'value': flops / tp / 1e12, 'speedup_batched': tp / t}
    # truncated block SVD back to chi (two-site update)
    with torch.no_grad():
        res = W.heff2(AA, FL, FR, Wm)
        ts = _timeit(lambda: W.two_site_truncation(res, 1024), 3, 1, sync)
        out['svd_truncation'] = {'workload': 'svd_with_truncation of the Heff2 output to D_total=1024 '
                                             '(27 blocks, max 198x198)', 'ms_batched': ts * 1e3}
        if not args.no_plain:
            resp = res._replace(config=cfgp)
            tsp = _timeit(lambda: W.two_site_truncation(resp, 1024), 2, 1, sync)
            out['svd_truncation']['ms_plain'] = tsp * 1e3
        if not args.no_cpu:  # the plain per-block path on the host cores (torch CPU; LAPACK per block)
            cfgh = bs.make_config(backend_plain, 'U1', device='cpu')
            resh = res._replace(config=cfgh, data=res.data.cpu())
            for _ in range(3):
                W.two_site_truncation(resh, 1024)
            th = []
            for _ in range(10):
                t0 = time.perf_counter()
                W.two_site_truncation(resh, 1024)
                th.append(time.perf_counter() - t0)
            out['svd_truncation']['ms_host'] = float(np.median(th)) * 1e3
            out['svd_truncation']['host'] = f'median of 10 after 3 warm-ups, torch CPU, {torch.get_num_threads()} threads'
            out['svd_truncation']['speedup_vs_host'] = float(np.median(th)) / ts
    # roofline of the dominant kernel (the launch with the largest duration)
    stats = _dot_launch_stats(backend_hip, lambda: W.heff2(AA, FL, FR, Wm))
    fl, ms, plan, dt, by, _ = max(stats, key=lambda x: x[1])
    # the bound follows the launch's arithmetic intensity against the ridge point of FP64 MFMA / HBM
    ridge = PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    common = {'kernel': _kernel_desc(plan),
              'avg_launch_ms': ms, 'flops_per_launch': fl, 'algorithmic_bytes_per_launch': by,
              'arithmetic_intensity': fl / by, 'all_dot_launches_ms': [round(x[1], 5) for x in stats]}
    common['traffic'] = _pmc_traffic(common['kernel'])
    if args.copy_probe:
        _pmc_sweep(backend_hip, stats)
    if fl / by < ridge:
        ach = by / (ms * 1e-3) / 1e9
        out['roofline'] = {'bound': 'hbm', 'achieved': ach, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                           'frac': ach / PEAK_HBM_GBS, **common}
    else:
        ach = fl / (ms * 1e-3) / 1e12
        out['roofline'] = {'bound': 'mfma', 'achieved': ach, 'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
                           'frac': ach / PEAK_FP64_TFLOPS, **common}
    return AA, FL, FR, Wm, flops


def bench_ctmrg_gemm():
    """The largest grouped-GEMM launches of the D=6, chi=64 CTMRG step's forward (the enlarged-corner
    products, sectors up to 440 x 440 x 644): TFLOP/s and roofline of the one with the most flops."""
    import torch
    from yastn_amd import backend_hip
    from yastn_amd import blocksparse as bs
    from yastn_amd import ctmrg as C
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A, env, sz = C.setup(cfg, D=6, chi=64, seed=0)
    with torch.no_grad():
        C.loss_fn(A, env, sz, 64)  # plans and tile choices
    stats = _dot_launch_stats(backend_hip, lambda: C.loss_fn(A, env, sz, 64), reps=20, top=4)
    fl, ms, plan, dt, by, _ = max(stats, key=lambda x: x[0])
    ach = fl / (ms * 1e-3) / 1e12
    ridge = PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    bound = 'mfma' if fl / by >= ridge else 'hbm'
    roof = ({'bound': 'mfma', 'achieved': ach, 'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
             'frac': ach / PEAK_FP64_TFLOPS} if bound == 'mfma' else
            {'bound': 'hbm', 'achieved': by / (ms * 1e-3) / 1e9, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
             'frac': by / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS})
    return {'workload': 'CTMRG D=6 chi=64 forward: the grouped-GEMM launch with the most flops',
            'kernel': _kernel_desc(plan), 'flops_per_launch': fl, 'avg_launch_ms': ms, 'tflops': ach,
            'algorithmic_bytes_per_launch': by, 'arithmetic_intensity': fl / by, 'roofline': roof,
            'top_launches': [{'gflop': x[0] / 1e9, 'ms': round(x[1], 5), 'tflops': x[0] / x[1] / 1e9}
                             for x in stats]}


def bench_u1xu1(args, out):
    import torch
    from yastn_amd import backend_hip, backend_plain
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    sync = torch.cuda.synchronize
    cfg = bs.make_config(backend_hip, 'U1xU1', device='cuda:0')
    A = W.u1xu1_transfer_setup(cfg, seed=0)
    flops = _count_flops(backend_hip, W.u1xu1_transfer, A)
    with torch.no_grad():
        t = _timeit(lambda: W.u1xu1_transfer(A), args.steps, args.warmup, sync)
        cfgp = bs.make_config(backend_plain, 'U1xU1', device='cuda:0')
        Ap = A._replace(config=cfgp)
        tp = _timeit(lambda: W.u1xu1_transfer(Ap), max(3, args.steps // 4), 1, sync)
    out['u1xu1_transfer'] = {'workload': 'tensordot(A, A*, ((0,1),(0,1))), U(1)xU(1), chi=2048 over 150 sectors, '
                                         'complex128', 'useful_gflop': flops / 1e9,
                             'tflops_batched': flops / t / 1e12, 'tflops_plain': flops / tp / 1e12,
                             'ms_batched': t * 1e3, 'ms_plain': tp * 1e3, 'speedup': tp / t}


def bench_qr(args, out):
    """backend.qr (batched Householder QR, hand-written adjoint) against the plain per-block
    torch.linalg.qr path of the same GPU, on 13 square blocks with the CTMRG chi D^2 sector sizes."""
    import numpy as np
    import torch
    from yastn_amd import backend_hip, backend_plain
    sync = torch.cuda.synchronize
    dims = (4, 20, 68, 152, 280, 396, 440, 396, 280, 152, 68, 20, 4)
    meta, off = [], 0
    for d in dims:
        meta.append(((off, off + d * d), (d, d), (off, off + d * d), (d, d), (off, off + d * d), (d, d)))
        off += d * d
    meta, sizes = tuple(meta), (off, off)
    a = torch.as_tensor(np.random.default_rng(0).standard_normal(off), device='cuda:0')
    w = torch.as_tensor(np.random.default_rng(1).standard_normal(off), device='cuda:0')

    def fwd(B):
        with torch.no_grad():
            B.qr(a, meta, sizes)

    def fwd_bwd(B):
        x = a.detach().requires_grad_(True)
        Q, R = B.qr(x, meta, sizes)
        ((Q * w).sum() + (R * w).sum()).backward()

    res = {'workload': 'backend.qr of 13 square float64 blocks (4 ... 440, the CTMRG chi D^2 sector sizes)'}
    for name, fn in (('fwd', fwd), ('fwd_bwd', fwd_bwd)):
        tb = _timeit(lambda: fn(backend_hip), 5, 2, sync)
        tp = _timeit(lambda: fn(backend_plain), 3, 1, sync)
        res[f'ms_{name}_batched'] = tb * 1e3
        res[f'ms_{name}_plain'] = tp * 1e3
        res[f'speedup_{name}'] = tp / tb
    out['qr'] = res


def bench_tds(args, out):
    """backend.transpose_dot_sum (no-fusion tensordot) against the reference torch kernel's per-block loop
    (_backend_torch_backwards.py:141-166) on the same GPU: 64 output blocks, each the sum of three
    products of permuted rank-3 blocks, sides 8 ... 64, float64."""
    import numpy as np
    import torch
    from yastn_amd import backend_hip
    sync = torch.cuda.synchronize
    rng = np.random.default_rng(0)
    Ar, Br, md, ao, bo, co, flops = [], [], [], 0, 0, 0, 0
    for _ in range(64):
        m, n = (int(x) for x in rng.integers(8, 65, 2))
        tab = []
        for _ in range(3):
            c0, c2 = (int(x) for x in rng.integers(2, 9, 2))
            Ar.append(((ao, ao + c0 * m * c2), (c0, m, c2), m, c0 * c2))
            Br.append(((bo, bo + c2 * c0 * n), (c2, c0, n), c0 * c2, n))
            ao, bo = ao + c0 * m * c2, bo + c2 * c0 * n
            tab.append((len(Ar) - 1, len(Br) - 1))
            flops += 2 * m * n * c0 * c2
        md.append(((co, co + m * n), (m, n), tuple(tab)))
        co += m * n
    Ar, Br, md = tuple(Ar), tuple(Br), tuple(md)
    A = torch.as_tensor(rng.standard_normal(ao), device='cuda:0')
    B = torch.as_tensor(rng.standard_normal(bo), device='cuda:0')
    order = (1, 0, 2)

    def plain():
        At = [A[sl[0]:sl[1]].view(Di).permute(order).reshape(Dl, Dr) for sl, Di, Dl, Dr in Ar]
        Bt = [B[sl[0]:sl[1]].view(Di).permute(order).reshape(Dl, Dr) for sl, Di, Dl, Dr in Br]
        C = torch.zeros(co, dtype=A.dtype, device=A.device)
        for sl, Dslc, tab in md:
            blk = C[sl[0]:sl[1]].view(Dslc)
            for ta, tb in tab:
                blk += At[ta] @ Bt[tb]
        return C

    with torch.no_grad():
        tb_ = _timeit(lambda: backend_hip.transpose_dot_sum(A, B, md, Ar, Br, order, order, co), 20, 3, sync)
        tp = _timeit(plain, 5, 1, sync)
    out['transpose_dot_sum'] = {'workload': '64 output blocks x 3 products of permuted rank-3 float64 blocks, sides '
                                            '8...64 (no-fusion tensordot)', 'useful_gflop': flops / 1e9,
                                'ms_batched': tb_ * 1e3, 'ms_plain': tp * 1e3, 'speedup': tp / tb_}


def bench_cpu(args, out, flops):
    """The numpy oracle on the host (per-block numpy.dot through OpenBLAS), bounded sample of the same
    Heff2 workload."""
    from oracle import np_module as NB
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    cores = os.cpu_count()
    try:
        from threadpoolctl import threadpool_info
        cores = max((x.get('num_threads', 0) for x in threadpool_info() if x.get('user_api') == 'blas'),
                    default=cores)
    except Exception:
        pass
    cfg = bs.make_config(NB, 'U1')
    AA, FL, FR, Wm = W.dmrg_heff2_setup(cfg, chi=1024, seed=0)
    W.heff2(AA, FL, FR, Wm)  # warm-up (planner caches)
    n, t0 = 0, time.perf_counter()
    while n < 3 or (time.perf_counter() - t0 < 10.0 and n < 50):
        W.heff2(AA, FL, FR, Wm)
        n += 1
    dt = (time.perf_counter() - t0) / n
    out['cpu_baseline'] = {'value': flops / dt / 1e12, 'unit': 'TFLOP/s', 'cores': int(cores), 'kind': 'port',
                           'sample': f'{n} Heff2 applies (same chi=1024 workload) with the numpy oracle, '
                                     f'{dt * 1e3:.1f} ms each', 'ms_per_step': dt * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-ctmrg', action='store_true')
    ap.add_argument('--only-heff2', action='store_true', help='headline workload only (for profiling)')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain per-block baselines (for profiling)')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run one 256 MiB device copy, the reference dispatch for scaling HBM counters')
    ap.add_argument('--unrolled', action='store_true',
                    help='add the unrolled CTMRG workload (D=8, chi=128, complex128, 20 steps + energy)')
    ap.add_argument('--only-unrolled', action='store_true', help='run the unrolled CTMRG workload alone')
    ap.add_argument('--only-ctmrg', action='store_true', help='run the D=6, chi=64 CTMRG step comparison alone')
    ap.add_argument('--unrolled-steps', type=int, default=20)
    ap.add_argument('--unrolled-plain-steps', type=int, default=1,
                    help='CTMRG steps of the short batched-vs-plain comparison (0 skips the plain path)')
    ap.add_argument('--unrolled-no-checkpoint', action='store_true',
                    help='also run the unrolled workload once without checkpointing (peak memory comparison)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU sharding of "
              f"this path is a design note only)", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    out = {'metric': 'tensordot useful TFLOP/s', 'value': None, 'unit': 'TFLOP/s', 'n_gpus': 1,
           'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': None, 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
           'config': {'workload': 'DMRG two-site Heff2 apply (4 tensordots), U(1) spin-1/2, chi=1024 over 25 '
                                  'sectors (D(q) ~ exp(-q^2/32)), XXZ MPO bond 5, float64 N(0,1) seed 0',
                      'tensordot_policy': 'fuse_contracted', 'device': torch.cuda.get_device_name(0)}}
    if args.copy_probe:
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device='cuda:0').fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only_unrolled:
        from yastn_amd import ctmrg
        print('bench: unrolled ctmrg', flush=True)
        out['ctmrg_unrolled'] = ctmrg.bench_unrolled(n_iter=args.unrolled_steps,
                                                     plain_iter=args.unrolled_plain_steps,
                                                     no_checkpoint_memory=args.unrolled_no_checkpoint)
        print(json.dumps(out), flush=True)
        return
    if args.only_ctmrg:
        from yastn_amd import ctmrg
        print('bench: ctmrg', flush=True)
        out['ctmrg'] = ctmrg.bench(steps=max(2, args.steps // 5))
        out['ctmrg']['gemm'] = bench_ctmrg_gemm()
        print(json.dumps(out), flush=True)
        return
    print('bench: heff2', flush=True)
    AA, FL, FR, Wm, flops = bench_heff2(args, out)
    if not args.only_heff2:
        print('bench: u1xu1', flush=True)
        bench_u1xu1(args, out)
        print('bench: qr', flush=True)
        bench_qr(args, out)
        print('bench: transpose_dot_sum', flush=True)
        bench_tds(args, out)
        if not args.no_ctmrg:
            from yastn_amd import ctmrg
            print('bench: ctmrg', flush=True)
            out['ctmrg'] = ctmrg.bench(steps=max(2, args.steps // 5))
            out['ctmrg']['gemm'] = bench_ctmrg_gemm()
            if args.unrolled:
                print('bench: unrolled ctmrg', flush=True)
                out['ctmrg_unrolled'] = ctmrg.bench_unrolled(n_iter=args.unrolled_steps,
                                                             plain_iter=args.unrolled_plain_steps)
    if not args.no_cpu:
        print('bench: cpu baseline', flush=True)
        bench_cpu(args, out, flops)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
