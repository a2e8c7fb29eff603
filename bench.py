'''
Seconds per two-site DMRG sweep of the U(1) Heisenberg chain on one MI355X: the batched form
(one composed leg move and one grouped MFMA launch per contraction step, head/tail moves hoisted out
of the Krylov loop) against the plain per-block FP64 form (the reference's per-step, per-block
launches), at L = 64, chi = 1024, n_iter = 5, seed 0.  `--precision fp32` runs the batched form's
Lanczos block products in FP32 on MFMA, followed by a two-step FP64 Lanczos restart from the Ritz
vector (defect correction); `fp64` keeps every product in FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--length L] [--bond CHI] [--precision P]

A step is one sweep.  W warmup sweeps ramp the bond (the state starts with one multiplet per
sector), then K sweeps are timed between synchronisations, with no instrumentation.  One further
sweep, outside the timed region, times every block-product, leg-move and tridiagonalisation-step
launch with HIP events attached to the kernel dispatch itself (hipExtLaunchKernelGGL), so that each
interval is the kernel alone.  The step of the split's batched eigensolver is the sweep's dominant
kernel: the roofline line is its algorithmic bytes (trailing block read and written once, plus the
step's vectors) over its average launch time, with the matvec's grouped GEMM (algorithmic 2mnk of its
descriptor table) and leg moves beside it.  From the state the timed sweeps started from, the
batched form then sweeps once in FP64 throughout (`fp64_form`), and the plain FP64 form once with
each split: one torch.linalg.svd per block as the reference factorises (`speedup_vs_plain`), and the
batched form's device eigensolver, so that its ratio isolates the contraction path (`plain_sweeps`,
with the plain block products' TFLOP/s).  |dE| between the forms' first sweep from that state is
reported (for the FP32 form against FP64 this is |E - E_FP64|), with the truncated weight of each.
The CPU baseline times the reference's own engine (its C++ sources built into oracle/_ref) under a
restatement of its DMRG driver on the reference's CPU case, L = 16, chi = 32, 3 sweeps, on one
thread and on all the process's threads, with the oracle (numpy, one thread) beside it.

One JSON line goes to stdout; progress goes to stderr.
'''

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(message: str) -> None:
    print(message, file=sys.stderr, flush=True)


def measured_traffic(compute):
    '''HBM bytes per block-product step of the matvec, from the hardware byte counters of two
    rocprofv3 passes over this bench (FETCH_SIZE, then WRITE_SIZE), each scaled by what the counter
    reads for a copy of known size in the same run; the summary is committed beside this file's
    profiles.  None when no summary for this compute type is there.'''
    path = os.path.join(HERE, 'profiles', 'dmrg_matvec_gemm_L64_chi1024_pmc.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(compute)
    if not entry:
        return None
    return {'bytes_per_step': entry['hbm_bytes_per_step'],
            'source': 'profiles/dmrg_matvec_gemm_L64_chi1024_pmc.json (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f, '
                      'calibrated on a %d MiB copy)' % (data['fetch_scale'], data['write_scale'], data['copy_mib'])}


def _cpu_model() -> str:
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(length=16, bond=32, sweeps=3):
    '''The reference's CPU case (L = 16, chi = 32, n_iter = 5, 3 sweeps from seed 0, FP64) on the host:
    the reference's own engine (built from its sources into oracle/_ref) under the oracle's
    restatement of its DMRG driver, on one ATen thread and on every thread the process may use; and
    the oracle (numpy) on one thread beside it.  Returns the baseline object of the bench line.'''
    import torch
    from threadpoolctl import threadpool_limits
    import oracle.symmetry as osym
    from oracle import ref_engine
    from oracle.dmrg import DMRG as OracleDMRG, Model as OracleModel
    usable = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    threads_all = torch.get_num_threads()
    out = {'unit': 's/sweep', 'nproc': os.cpu_count(), 'usable_cpus': usable, 'cpu_model': _cpu_model()}
    with threadpool_limits(limits=1):
        driver = OracleDMRG(OracleModel(length, cat=osym.u1), max_bond_dim=bond, n_iter=5, seed=0)
        times = []
        for sweep in range(sweeps):
            start = time.perf_counter()
            energy, _ = driver.sweep()
            times.append(time.perf_counter() - start)
        log(f'[cpu] oracle (numpy), 1 thread: {sum(times) / len(times):.3f} s/sweep, energy {energy:.12f}')
    port = {'value': round(sum(times) / len(times), 4), 'cores': 1, 'kind': 'port', 'energy': energy}
    if ref_engine.engine() is None:
        out.update(port)
        out['sample'] = ('oracle (numpy restatement of the reference engine), U(1) Heisenberg L=16 chi=32 n_iter=5, '
                         '3 sweeps from seed 0, one thread (the reference engine is not built here)')
        return out
    runs = {}
    for threads in (1, threads_all):
        with threadpool_limits(limits=threads):
            seconds, energies = ref_engine.sweeps(length, bond, sweeps, threads=threads)
        runs[threads] = (sum(seconds) / len(seconds), energies[-1])
        log(f'[cpu] reference engine, {threads} thread(s): {runs[threads][0]:.3f} s/sweep, energy {energies[-1]:.12f}')
    out.update({'value': round(runs[1][0], 4), 'cores': 1, 'kind': 'reference', 'energy': runs[1][1],
                'all_cores': {'value': round(runs[threads_all][0], 4), 'cores': threads_all},
                'port': port,
                'sample': 'the reference engine (its C++ sources built with g++ into oracle/_ref, ATen per block on '
                          'the CPU) under a restatement of its DMRG driver: U(1) Heisenberg L=16 chi=32 n_iter=5, '
                          f'{sweeps} sweeps from seed 0, FP64, on 1 and on {threads_all} ATen threads; the '
                          'reference\'s CPU case, not L=64 chi=1024'})
    return out


def time_breakdown(driver):
    '''One sweep with the main phases wrapped in synchronised timers (diagnostic only); whatever the
    wrapped phases leave out (head/tail einsums, Python) is the sweep minus their sum.'''
    import torch
    import tentorch_amd.algorithms as alg
    totals = {}

    def timed(name, fn):
        def wrapper(*a, **k):
            torch.cuda.synchronize()
            start = time.perf_counter()
            out = fn(*a, **k)
            torch.cuda.synchronize()
            totals[name] = totals.get(name, 0.0) + time.perf_counter() - start
            return out
        return wrapper

    envs, state = driver.environments, driver.state
    patched = [(alg, '_split', 'split'), (alg, 'contract', 'two-site contract'), (driver, '_solve', 'lanczos'),
               (envs, 'left_of', 'environments'), (envs, 'right_of', 'environments'),
               (state, '_gauge_step', 'canonical moves'), (state, 'normalize_on_device', 'normalize')]
    saved = [(owner, name, getattr(owner, name)) for owner, name, _ in patched]
    for owner, name, label in patched:
        setattr(owner, name, timed(label, getattr(owner, name)))
    try:
        torch.cuda.synchronize()
        start = time.perf_counter()
        driver.sweep()
        torch.cuda.synchronize()
        totals['sweep'] = time.perf_counter() - start
    finally:
        for owner, name, original in saved:
            if owner is alg:
                setattr(owner, name, original)
            else:
                delattr(owner, name)    # the instance attribute shadowed the class's method
    totals['rest'] = totals['sweep'] - sum(v for k, v in totals.items() if k != 'sweep')
    return {k: round(v, 3) for k, v in totals.items()}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--gpus', type=int, default=1)
    parser.add_argument('--steps', type=int, default=2, help='timed sweeps')
    parser.add_argument('--warmup', type=int, default=4,
                        help='untimed sweeps that ramp the bond and converge the state')
    parser.add_argument('--length', type=int, default=64)
    parser.add_argument('--bond', type=int, default=1024)
    parser.add_argument('--iterations', type=int, default=5)
    parser.add_argument('--precision', choices=('fp64', 'fp32'), default='fp32',
                        help='block products of the batched form: fp32 (with FP64 defect correction) or fp64')
    parser.add_argument('--refine', type=int, default=3, help='FP64 Lanczos steps after the FP32 solve')
    parser.add_argument('--no-plain', action='store_true', help='skip the plain per-block sweeps')
    parser.add_argument('--plain-split', choices=('reference', 'shared', 'both'), default='both',
                        help="plain sweeps' split: per-block torch SVD as the reference, the batched "
                             "form's device eigensolver, or one sweep with each")
    parser.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    parser.add_argument('--no-profile', action='store_true', help='do not time launches with events')
    parser.add_argument('--breakdown', action='store_true',
                        help='after the timed sweeps, one extra sweep with a synchronised time breakdown')
    parser.add_argument('--copy-mib', type=int, default=0,
                        help='at the end, one plain FP64 copy of this size (the calibration kernel of the '
                             'hardware byte counters)')
    args = parser.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: every box here has one, and this path has no multi-GPU form '
              f'(asked for {args.gpus}).', file=sys.stderr)
        return 2

    compute = 'f32' if args.precision == 'fp32' else 'f64'
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a GPU', file=sys.stderr)
        return 1
    import tentorch_amd as tt
    from tentorch_amd.tensor import Profiler, Settings, workspaces
    from tentorch_amd.linalg import SvdSettings

    log(f'[bench] device {torch.cuda.get_device_name(0)}; L={args.length} chi={args.bond} '
        f'n_iter={args.iterations} precision={args.precision} warmup={args.warmup} steps={args.steps}')
    driver = tt.DMRG(tt.Model(args.length, cat=tt.u1), max_bond_dim=args.bond, n_iter=args.iterations, seed=0,
                     precision=args.precision, refine=args.refine)
    # A progress line every few windows, so a long sweep is visibly alive.
    update = driver.update
    windows = {'count': 0, 'start': time.perf_counter()}

    def logged_update(direction):
        result = update(direction)
        windows['count'] += 1
        if windows['count'] % 16 == 0:
            log(f'[bench]   {windows["count"]} windows, {time.perf_counter() - windows["start"]:.1f} s')
        return result
    driver.update = logged_update
    for sweep in range(args.warmup):
        start = time.perf_counter()
        energy, lost = driver.sweep()
        torch.cuda.synchronize()
        log(f'[bench] warmup sweep {sweep + 1}: {time.perf_counter() - start:.2f} s energy {energy:.12f} '
            f'max bond {max(driver.state.bond_dims)}')
    snapshot = ([t.clone() for t in driver.state.tensors], driver.state.center)

    times, energies, truncated = [], [], []
    torch.cuda.synchronize()
    total_start = time.perf_counter()
    for sweep in range(args.steps):
        start = time.perf_counter()
        energy, lost = driver.sweep()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - start)
        energies.append(energy)
        truncated.append(lost)
        log(f'[bench] timed sweep {sweep + 1}: {times[-1]:.2f} s energy {energy:.12f} truncated {lost:.2e} '
            f'max bond {max(driver.state.bond_dims)}')
    total = time.perf_counter() - total_start
    per_sweep = total / args.steps

    after = ([t.clone() for t in driver.state.tensors], driver.state.center)

    def from_snapshot(label, mode, precision, method, profiler=None):
        '''One sweep from the state the timed sweeps started from, in another form: mode 1 is the
        plain per-block form (per-block launches for every contraction), `precision` the block
        products' type, `method` the split: 'exact' per block by torch.linalg.svd, as the reference
        factorises (svd.cpp:59), or 'gram', the batched device eigensolver the batched form uses.'''
        state = driver.state
        state.tensors = [t.clone() for t in snapshot[0]]
        state.center = snapshot[1]
        workspaces.clear()
        Settings.mode = mode
        saved = driver.precision, SvdSettings.method
        driver.precision, SvdSettings.method = precision, method
        try:
            torch.cuda.synchronize()
            start = time.perf_counter()
            if profiler is None:
                energy, lost = driver.sweep()
            else:
                with profiler:
                    energy, lost = driver.sweep()
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - start
        finally:
            Settings.mode = 0
            driver.precision, SvdSettings.method = saved
        log(f'[bench] {label} sweep: {elapsed:.2f} s energy {energy:.12f} |dE| {abs(energy - energies[0]):.2e} '
            f'truncated {lost:.2e}')
        return elapsed, energy, lost

    # the batched form in FP64 throughout, from the same state
    fp64_form = None
    if compute == 'f32':
        t64, e64, lost64 = from_snapshot('batched FP64', 0, 'fp64', SvdSettings.method)
        fp64_form = {'s_per_sweep': round(t64, 4), 'dE_vs_fp32_form': abs(e64 - energies[0]), 'truncated': lost64}

    # the plain form, its block products timed program by program
    plain, plain_tflops = {}, {}
    if not args.no_plain:
        for split in ('reference', 'shared'):
            if args.plain_split in (split, 'both'):
                timer = Profiler()
                plain[split] = from_snapshot(f'plain ({split} split)', 1, 'fp64',
                                             'exact' if split == 'reference' else 'gram', timer)
                part = timer.summary().get('plain')
                if part and part['ms'] > 0:
                    plain_tflops[split] = round(part['flops'] / (part['ms'] * 1e-3) / 1e12, 3)
                    log(f'[bench] plain ({split} split) block products: {part["ops"]} programs, '
                        f'{part["ms"]:.1f} ms, {plain_tflops[split]:.2f} TFLOP/s')
    first_plain = plain.get('reference') or plain.get('shared')
    plain_per_sweep, plain_energy, plain_lost = first_plain if first_plain else (None, None, None)
    delta = None if plain_energy is None else abs(plain_energy - energies[0])

    # One more sweep, after the timed region and outside it, from the state the timed sweeps left
    # and last on the device (so a profiler's last dispatches are this sweep's), with every leg move,
    # block product and tridiagonalisation step timed by events attached to the kernel dispatch.
    driver.state.tensors, driver.state.center = after
    workspaces.clear()
    profiler = Profiler()
    if not args.no_profile:
        start = time.perf_counter()
        with profiler:
            driver.sweep()
        torch.cuda.synchronize()
        log(f'[bench] instrumented sweep: {time.perf_counter() - start:.2f} s')

    breakdown = None
    if args.breakdown:
        breakdown = time_breakdown(driver)
        log(f'[bench] breakdown of one sweep: {json.dumps(breakdown)}')

    roofline = None
    matvec = 'EPaqj,JRSFb,AQsb->EPrf'
    peak = PEAK_FP32_TFLOPS if compute == 'f32' else PEAK_FP64_TFLOPS
    summary = profiler.summary(matvec, compute) if not args.no_profile else {}
    overall = profiler.summary() if not args.no_profile else {}
    refine = profiler.summary(matvec, 'f64') if not args.no_profile and compute == 'f32' else {}
    for name, part in (('all', overall), ('matvec', summary), ('matvec fp64 refine', refine)):
        for kind in ('gemm', 'move', 'tridiag'):
            if kind in part and part[kind]['ms'] > 0:
                k = part[kind]
                log(f'[bench] {name} {kind}: {k["launches"]} launches, {k["ms"]:.1f} ms, '
                    f'{k["flops"] / (k["ms"] * 1e-3) / 1e12:.2f} TFLOP/s, {k["bytes"] / (k["ms"] * 1e-3) / 1e9:.1f} GB/s')
    if 'tridiag' in overall and overall['tridiag']['ms'] > 0:
        # the dominant kernel of the sweep: a Householder step of the split's batched eigensolver
        tri = overall['tridiag']
        gbs = tri['bytes'] / (tri['ms'] * 1e-3) / 1e9
        traffic = measured_traffic('tridiag')
        roofline = {'bound': 'hbm',
                    'kernel': 'tridiag_step_kernel<4|8|16|32> (one Householder step of the split\'s batched '
                              'symmetric eigensolver over every Gram matrix of a pass, one wave per trailing '
                              'column)',
                    'achieved': round(gbs, 1), 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                    'frac': round(gbs / PEAK_HBM_GBS, 4),
                    'traffic': None if traffic is None else traffic['bytes_per_step'],
                    'traffic_source': None if traffic is None else traffic['source'],
                    'algorithmic_bytes_per_launch': tri['bytes'] / tri['launches'],
                    'launches': tri['launches'], 'avg_launch_us': round(tri['ms'] * 1e3 / tri['launches'], 3),
                    'timing': 'kernel-dispatch events (hipExtLaunchKernelGGL) on every step launch of one '
                              'instrumented sweep after the timed region; algorithmic bytes per step: the '
                              'trailing block read and written once plus the step vectors, 8 (2 m^2 + 6 m) per '
                              'matrix with m trailing columns'}
    gemm_entry = None
    if 'gemm' in summary and summary['gemm']['ms'] > 0:
        g = summary['gemm']
        achieved = g['flops'] / (g['ms'] * 1e-3) / 1e12
        kernel = ('ggemm_kernel<float,double,float,float> and <float,float,double,float> (the matvec\'s two '
                  'FP32-MFMA block-product steps: FP64 operand converted on load, FP64 result), with '
                  'ggemm_reduce_kernel<double,float> after a step whose products are split along their depth'
                  if compute == 'f32' else 'ggemm_kernel<double,double,double,double> (matvec block products), with '
                  'ggemm_reduce_kernel<double,double> after a step whose products are split along their depth')
        traffic = measured_traffic(compute)
        gemm_entry = {'bound': 'mfma', 'kernel': kernel,
                      'achieved': round(achieved, 3), 'peak': peak, 'unit': 'TFLOP/s',
                      'frac': round(achieved / peak, 4),
                      'traffic': None if traffic is None else traffic['bytes_per_step'],
                      'traffic_source': None if traffic is None else traffic['source'],
                      'algorithmic_bytes_per_step': g['bytes'] / g['ops'],
                      'steps': g['ops'], 'launches': g['launches'],
                      'avg_step_us': round(g['ms'] * 1e3 / g['ops'], 2),
                      'flops_per_step': g['flops'] / g['ops'],
                      'timing': 'kernel-dispatch events (hipExtLaunchKernelGGL) over one instrumented sweep '
                                'after the timed region; a step is one grouped launch plus, when its products '
                                'are split along their depth, the launch that adds their parts'}
    moves_entry = None
    if 'move' in summary and summary['move']['ms'] > 0:
        m = summary['move']
        gbs = m['bytes'] / (m['ms'] * 1e-3) / 1e9
        moves_entry = {'bound': 'hbm', 'kernel': 'gss_kernel (matvec leg moves into the FP32 form\'s workspace)'
                       if compute == 'f32' else 'gss_kernel (matvec leg moves)', 'achieved': round(gbs, 1),
                       'peak': PEAK_HBM_GBS, 'unit': 'GB/s', 'frac': round(gbs / PEAK_HBM_GBS, 4),
                       'launches': m['launches'], 'avg_launch_us': round(m['ms'] * 1e3 / m['launches'], 2),
                       'bytes_per_launch': m['bytes'] / m['launches']}
    if roofline is None and gemm_entry is not None:
        roofline = dict(gemm_entry)
    elif roofline is not None:
        roofline['matvec_gemm'] = gemm_entry
        roofline['matvec_moves'] = moves_entry
    if roofline is not None and 'gemm' in overall and overall['gemm']['ms'] > 0:
        roofline['all_block_products_tflops'] = round(overall['gemm']['flops'] / (overall['gemm']['ms'] * 1e-3) / 1e12, 3)

    if args.copy_mib:
        from tentorch_amd import _native
        n = args.copy_mib * (1 << 20) // 8
        x = torch.randn(n, dtype=torch.float64, device='cuda')
        y = torch.empty_like(x)
        _native.axpby(torch.float64, 1.0, x, 0.0, y)
        torch.cuda.synchronize()
        log(f'[bench] calibration copy: {args.copy_mib} MiB read, {args.copy_mib} MiB written (axpby_kernel<double>)')

    baseline = None
    if not args.no_cpu:
        baseline = cpu_baseline()

    line = {
        'metric': 's/sweep U(1) Heisenberg L=64 chi=1024, batched vs plain, one GPU, |dE|<=1e-10',
        'value': round(per_sweep, 4), 'unit': 's/sweep', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(per_sweep * 1e3, 1), 'higher_is_better': False, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': compute, 'data': 'synthetic (random initial MPS, seed 0)',
        'config': {'workload': f'two-site DMRG sweep, U(1) Heisenberg L={args.length} chi={args.bond} '
                               f'n_iter={args.iterations}, batched form, '
                               + ('FP32 block products + FP64 Lanczos/SVD/defect correction'
                                  if compute == 'f32' else 'FP64 throughout'),
                   'length': args.length, 'bond': args.bond, 'n_iter': args.iterations,
                   'precision': args.precision, 'refine': args.refine if compute == 'f32' else 0},
        'sweeps_s': [round(t, 4) for t in times], 'energies': energies, 'truncated': truncated,
        'plain_truncated': plain_lost,
        'plain_split': 'reference' if 'reference' in plain else ('shared' if plain else None),
        'plain_s_per_sweep': None if plain_per_sweep is None else round(plain_per_sweep, 4),
        'speedup_vs_plain': None if plain_per_sweep is None else round(plain_per_sweep / per_sweep, 3),
        'dE_plain_vs_batched': delta,
        'plain_sweeps': {split: {'s_per_sweep': round(t, 4), 'speedup': round(t / per_sweep, 3),
                                 'dE': abs(e - energies[0]), 'truncated': lost,
                                 'block_product_tflops': plain_tflops.get(split)}
                         for split, (t, e, lost) in plain.items()},
        'fp64_form': fp64_form,
        'roofline': roofline, 'cpu_baseline': baseline, 'breakdown_s': breakdown,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
