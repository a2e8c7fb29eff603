'''
The split's eigendecomposition on one MI355X: the batched device eigensolver (`tt_eigh_batched`)
against `torch.linalg.eigh` per matrix, on synthetic Gram matrices shaped like the split of one
DMRG window (the orders the U(1) Heisenberg window meets at chi = 1024 and chi = 2048; geometric
spectra down to 1e-16 of the largest, like the Gram matrices of a two-site tensor).

    python -m tentorch_amd.splitbench [--window chi1024|chi2048|both] [--reps R]

Times are wall clock between synchronisations around whole calls (every launch of the call), after
one warm call.  The device results are checked against torch's: eigenvalues within 1e-12 of the
largest.  One JSON line per window goes to stdout.
'''

import argparse
import json
import sys
import time

import numpy as np
import torch

# Gram orders of one window's blocks (the largest windows of each sweep, from the sweeps' splits)
WINDOWS = {
    'chi1024': [530, 500, 470, 420, 380, 330, 280, 230, 180, 140, 100, 70, 40, 20, 8, 2],
    'chi2048': [1054, 1020, 980, 900, 850, 760, 700, 640, 600, 560, 500, 430, 380, 320, 250, 190, 130, 80, 40,
                16, 4],
}


def gram_matrices(sizes, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        b = rng.standard_normal((n, n)) * np.geomspace(1.0, 1e-8, n)[None, :]
        out.append(b @ b.T)
    return out


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--window', choices=('chi1024', 'chi2048', 'both'), default='both')
    parser.add_argument('--reps', type=int, default=5)
    args = parser.parse_args(argv)
    if not torch.cuda.is_available():
        print('splitbench needs a GPU', file=sys.stderr)
        return 1
    from . import _native
    for name in (('chi1024', 'chi2048') if args.window == 'both' else (args.window,)):
        sizes = WINDOWS[name]
        grams = gram_matrices(sizes)
        packed = torch.tensor(np.concatenate([g.reshape(-1) for g in grams]), device='cuda')
        work = torch.empty(_native.eigh_workspace(sizes), dtype=torch.float64, device='cuda')
        device_grams = [torch.tensor(g, device='cuda') for g in grams]

        def device_call():
            return _native.eigh_batched(packed.clone(), sizes, work=work)

        def torch_call():
            values = []
            for g in device_grams:
                values.append(torch.linalg.eigh(g)[0])
            return values

        timings = {}
        for label, call in (('device', device_call), ('torch', torch_call)):
            call()
            torch.cuda.synchronize()
            times = []
            for _ in range(args.reps):
                start = time.perf_counter()
                result = call()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - start)
            timings[label] = (min(times), sorted(times)[len(times) // 2], result)
        values = timings['device'][2][0].cpu().numpy()
        at, worst = 0, 0.0
        for n, ref in zip(sizes, timings['torch'][2]):
            ref = ref.cpu().numpy()
            worst = max(worst, float(np.abs(values[at:at + n] - ref).max() / max(1.0, np.abs(ref).max())))
            at += n
        line = {'window': name, 'orders': sizes, 'device_ms': round(timings['device'][0] * 1e3, 3),
                'device_median_ms': round(timings['device'][1] * 1e3, 3),
                'torch_eigh_ms': round(timings['torch'][0] * 1e3, 3),
                'speedup': round(timings['torch'][0] / timings['device'][0], 2), 'max_eigenvalue_error': worst}
        print(json.dumps(line), flush=True)
        if worst > 1e-12:
            print(f'splitbench: device eigenvalues differ from torch by {worst:.2e}', file=sys.stderr)
            return 1
    return 0


if __name__ == '__main__':
    from tentorch_amd.splitbench import main as package_main
    sys.exit(package_main())
