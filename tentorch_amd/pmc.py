'''
Reduction of rocprofv3 output of `bench.py` to the figures its roofline line uses.

Three profiled runs of the same bench command (the FP32 form, with its instrumented sweep last and a
256 MiB calibration copy after it):

    rocprofv3 --kernel-trace --stats --output-format csv -d T -o bench -- python3 -u bench.py --no-plain --no-cpu > T/bench.json
    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex "ggemm|axpby" --output-format csv -d F -o fetch -- python3 -u bench.py --no-plain --no-cpu --copy-mib 256
    rocprofv3 --pmc WRITE_SIZE --kernel-include-regex "ggemm|axpby" --output-format csv -d W -o write -- python3 -u bench.py --no-plain --no-cpu --copy-mib 256
    python -m tentorch_amd.pmc --trace T --fetch F --write W --bench-line T/bench.json --out OUT

`--per-kernel` prints the mean counters of every kernel of two counter passes instead (used for
the block-einsum microbenchmark, `python -m tentorch_amd.microbench ... --copy-mib 256`).

The runs are deterministic (seed 0), so the last sweep of each is the same instrumented sweep.  Of
it, the matvec's FP32 block-product steps are taken: the last `steps / 2` dispatches of each of
the two kernel families (`ggemm_kernel<float, double, float, float, ...>` for the first product,
`<float, float, double, float, ...>` for the second) and the last `launches - steps` partial-sum
launches (`ggemm_reduce_kernel<double, float>`), the counts the bench line reports for that sweep.
The byte counters are scaled by what they read for the copy (read 256 MiB, write 256 MiB), the last
`axpby_kernel<double>` dispatch.  Writes `dmrg_matvec_gemm_L64_chi1024_pmc.json` (read by bench.py)
and `reduced.json` (kernel-trace averages and raw counter sums) into OUT.
'''

import argparse
import csv
import glob
import json
import os
import shutil
import sys

STEP1 = 'ggemm_kernel<float, double, float, float'
STEP2 = 'ggemm_kernel<float, float, double, float'
REDUCE = 'ggemm_reduce_kernel<double, float>'
COPY = 'axpby_kernel<double>'


def _rows(pattern):
    paths = glob.glob(pattern, recursive=True)
    if not paths:
        raise FileNotFoundError(pattern)
    with open(paths[0]) as f:
        return list(csv.DictReader(f))


def _last(rows, name, count):
    chosen = [r for r in rows if name in r['Kernel_Name']]
    return chosen[-count:] if count else []


def per_kernel(fetch_dir, write_dir, copy_mib):
    '''Mean FETCH_SIZE and WRITE_SIZE per dispatch of every kernel of two counter passes, with the
    copy's calibration (the last `axpby_kernel<double>` dispatch is the copy of `copy_mib` MiB).'''
    table = {}
    copies = {}
    for counter, directory in (('fetch', fetch_dir), ('write', write_dir)):
        rows = _rows(os.path.join(directory, '**', '*counter_collection.csv'))
        rows.sort(key=lambda r: int(r['Dispatch_Id']))
        copies[counter] = float([r for r in rows if COPY in r['Kernel_Name']][-1]['Counter_Value'])
        for r in rows:
            name = r['Kernel_Name'].replace('(anonymous namespace)::', '').split('(')[0]
            entry = table.setdefault(name, {'fetch': [], 'write': []})
            entry[counter].append(float(r['Counter_Value']))
    scale = {k: copy_mib * 1024.0 / v for k, v in copies.items()}
    print(f'calibration: FETCH_SIZE x {scale["fetch"]:.3f}, WRITE_SIZE x {scale["write"]:.3f}')
    print('| kernel | dispatches | FETCH_SIZE KiB (mean) | WRITE_SIZE KiB (mean) | HBM bytes per dispatch, calibrated |')
    print('|---|---:|---:|---:|---:|')
    for name, e in sorted(table.items()):
        f = sum(e['fetch']) / max(len(e['fetch']), 1)
        w = sum(e['write']) / max(len(e['write']), 1)
        total = (f * scale['fetch'] + w * scale['write']) * 1024.0
        print(f'| `{name}` | {len(e["fetch"])} | {f:.0f} | {w:.0f} | {total / 2**20:.1f} MiB |')
    return 0


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--per-kernel', action='store_true',
                        help='print a per-kernel table of the two counter passes instead (no --trace needed)')
    for flag in ('--trace', '--fetch', '--write', '--bench-line', '--out'):
        parser.add_argument(flag)
    parser.add_argument('--copy-mib', type=int, default=256)
    args = parser.parse_args(argv)
    if args.per_kernel:
        return per_kernel(args.fetch, args.write, args.copy_mib)
    if not all((args.trace, args.fetch, args.write, args.bench_line, args.out)):
        parser.error('--trace, --fetch, --write, --bench-line and --out are needed')
    os.makedirs(args.out, exist_ok=True)
    bench = json.loads(open(args.bench_line).read().strip().splitlines()[-1])
    roof = bench['roofline']
    steps, reduces = roof['steps'], roof['launches'] - roof['steps']
    families = {'step1': (STEP1, steps // 2), 'step2': (STEP2, steps // 2), 'reduce': (REDUCE, reduces)}

    trace = _rows(os.path.join(args.trace, '**', '*kernel_trace.csv'))
    trace.sort(key=lambda r: int(r['Start_Timestamp']))
    out = {'bench_roofline': roof, 'steps_per_sweep': steps, 'reduce_launches_per_sweep': reduces}
    total_ns = 0
    for key, (name, count) in families.items():
        d = [int(r['End_Timestamp']) - int(r['Start_Timestamp']) for r in _last(trace, name, count)]
        total_ns += sum(d)
        out['trace_' + key] = {'dispatches': len(d), 'avg_us': sum(d) / max(len(d), 1) / 1e3}
    out['trace_avg_step_us'] = total_ns / steps / 1e3

    def counters(directory):
        rows = _rows(os.path.join(directory, '**', '*counter_collection.csv'))
        rows.sort(key=lambda r: int(r['Dispatch_Id']))
        sums = {key: sum(float(r['Counter_Value']) for r in _last(rows, name, count))
                for key, (name, count) in families.items()}
        copy = [r for r in rows if COPY in r['Kernel_Name']][-1]
        return sums, float(copy['Counter_Value'])

    fetch, fetch_copy = counters(args.fetch)
    write, write_copy = counters(args.write)
    kib = args.copy_mib * 1024.0
    fetch_scale, write_scale = kib / fetch_copy, kib / write_copy    # the counters read KiB
    fetch_bytes = sum(fetch.values()) * fetch_scale * 1024.0 / steps
    write_bytes = sum(write.values()) * write_scale * 1024.0 / steps
    out['pmc'] = {'fetch_kib': fetch, 'write_kib': write, 'copy_fetch_kib': fetch_copy, 'copy_write_kib': write_copy}
    summary = {'f32': {'hbm_bytes_per_step': fetch_bytes + write_bytes, 'fetch_bytes_per_step': fetch_bytes,
                       'write_bytes_per_step': write_bytes,
                       'algorithmic_bytes_per_step': roof['algorithmic_bytes_per_step'], 'steps': steps},
               'fetch_scale': fetch_scale, 'write_scale': write_scale, 'copy_mib': args.copy_mib}
    with open(os.path.join(args.out, 'reduced.json'), 'w') as f:
        json.dump(out, f, indent=1)
    with open(os.path.join(args.out, 'dmrg_matvec_gemm_L64_chi1024_pmc.json'), 'w') as f:
        json.dump(summary, f, indent=1)
    for path in glob.glob(os.path.join(args.trace, '**', '*kernel_stats.csv'), recursive=True):
        shutil.copy(path, os.path.join(args.out, 'kernel_stats.csv'))
    print(json.dumps(summary))
    return 0


if __name__ == '__main__':
    sys.exit(main())
