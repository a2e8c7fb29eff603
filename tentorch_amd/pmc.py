'''
Reduction of rocprofv3 output of `bench.py` to the figures its roofline line uses.

Profiled runs of the bench command (the FP32 form, with its instrumented sweep last and a 256 MiB
calibration copy after it), a kernel trace and two counter passes:

    rocprofv3 --kernel-trace --stats --output-format csv -d T -o bench -- python3 -u bench.py ARGS > T.json
    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex "tridiag_step|ggemm|axpby" --output-format csv -d F -o fetch -- python3 -u bench.py ARGS --copy-mib 256 > F.json
    rocprofv3 --pmc WRITE_SIZE --kernel-include-regex "tridiag_step|ggemm|axpby" --output-format csv -d W -o write -- python3 -u bench.py ARGS --copy-mib 256 > W.json
    python -m tentorch_amd.pmc --reduce T --bench-line T.json --out trace.json     (and F, W alike)
    python -m tentorch_amd.pmc --combine --fetch fetch.json --write write.json --out OUT

Each run is reduced where it ran (its raw files are large); the reductions are small JSON files.

`--per-kernel` prints the mean counters of every kernel of two counter passes instead (used for
the block-einsum microbenchmark, `python -m tentorch_amd.microbench ... --copy-mib 256`).

The runs are deterministic (seed 0), so the last sweep of each is the same instrumented sweep.  Of
it, the step launches of the split's tridiagonalisation are taken (the last `launches` dispatches of
`tridiag_step_kernel<...>`, the count the bench line's roofline reports), and the matvec's FP32
block-product steps: the last `steps / 2` dispatches of each of
the two kernel families (`ggemm_kernel<float, double, float, float, ...>` for the first product,
`<float, float, double, float, ...>` for the second) and the last `launches - steps` partial-sum
launches (`ggemm_reduce_kernel<double, float>`), the counts the bench line reports for that sweep.
The byte counters are scaled by what they read for the copy (read 256 MiB, write 256 MiB), the last
`axpby_kernel<double>` dispatch.  `--combine` writes `dmrg_matvec_gemm_L64_chi1024_pmc.json` (read
by bench.py: HBM bytes per tridiagonalisation launch under `tridiag`, per block-product step under
`f32`) into OUT.
'''

import argparse
import csv
import glob
import json
import os
import sys

TRIDIAG = 'tridiag_step_kernel'
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


def _groups(roof):
    '''The kernel families of the instrumented sweep the bench line reports: name fragment and
    dispatch count, per group (the tridiagonalisation steps; the matvec's FP32 block products).'''
    gemm = roof.get('matvec_gemm') if roof.get('bound') == 'hbm' else roof
    groups = {}
    if roof.get('bound') == 'hbm':
        groups['tridiag'] = {'families': {'tridiag': (TRIDIAG, roof['launches'])}, 'units': roof['launches'],
                             'algorithmic': roof['algorithmic_bytes_per_launch'], 'unit': 'launch'}
    if gemm:
        steps = gemm['steps']
        groups['f32'] = {'families': {'step1': (STEP1, steps // 2), 'step2': (STEP2, steps // 2),
                                      'reduce': (REDUCE, gemm['launches'] - steps)},
                         'units': steps, 'algorithmic': gemm['algorithmic_bytes_per_step'],
                         'unit': 'block-product step'}
    return groups


def _stream_last(path, families, value):
    '''The last `count` values of each family in dispatch order, and the last copy dispatch's value.
    The file is read row by row and only matching dispatches are kept (a trace or counter file of a
    whole bench run holds about a million dispatches, not always in dispatch order).'''
    kept = {key: [] for key in families}
    copies = []
    with open(path) as f:
        for row in csv.DictReader(f):
            name = row['Kernel_Name']
            if COPY in name:
                copies.append((int(row['Dispatch_Id']), value(row)))
            for key, (fragment, count) in families.items():
                if count and fragment in name:
                    kept[key].append((int(row['Dispatch_Id']), value(row)))
    last = {}
    for key, (_, count) in families.items():
        ordered = sorted(kept[key])
        last[key] = [v for _, v in ordered[-count:]] if count else []
    copy = max(copies)[1] if copies else None
    return last, copy


def reduce_pass(directory, bench_line, out):
    '''One profiled run (kernel trace, or one counter) reduced to its instrumented sweep's figures.'''
    roof = json.loads(open(bench_line).read().strip().splitlines()[-1])['roofline']
    groups = _groups(roof)
    families = {f'{g}:{k}': v for g, spec in groups.items() for k, v in spec['families'].items()}
    traces = glob.glob(os.path.join(directory, '**', '*kernel_trace.csv'), recursive=True)
    result = {'roofline': roof}
    if traces:
        last, _ = _stream_last(traces[0], families,
                               lambda r: int(r['End_Timestamp']) - int(r['Start_Timestamp']))
        result['kind'] = 'trace'
        result['avg_us'] = {k: sum(v) / max(len(v), 1) / 1e3 for k, v in last.items()}
        result['dispatches'] = {k: len(v) for k, v in last.items()}
        for g, spec in groups.items():
            total = sum(sum(last[f'{g}:{k}']) for k in spec['families'])
            result.setdefault('avg_us_per_unit', {})[g] = total / spec['units'] / 1e3
        stats = glob.glob(os.path.join(directory, '**', '*kernel_stats.csv'), recursive=True)
        if stats:
            with open(stats[0]) as f:
                result['stats'] = [r for r in csv.DictReader(f)
                                   if any(fragment in r['Name'] for fragment, _ in families.values())]
    else:
        counters = glob.glob(os.path.join(directory, '**', '*counter_collection.csv'), recursive=True)
        if not counters:
            raise FileNotFoundError(f'no kernel trace or counter file under {directory}')
        names = set()

        def value(row):
            names.add(row['Counter_Name'])
            return float(row['Counter_Value'])
        last, copy = _stream_last(counters[0], families, value)
        result['kind'] = 'counter'
        result['counter'] = sorted(names)
        result['copy_kib'] = copy
        result['sum_kib'] = {g: sum(sum(last[f'{g}:{k}']) for k in spec['families']) for g, spec in groups.items()}
        result['dispatches'] = {k: len(v) for k, v in last.items()}
    with open(out, 'w') as f:
        json.dump(result, f, indent=1)
    print(json.dumps({k: v for k, v in result.items() if k not in ('roofline', 'stats')}))
    return result


def combine(fetch_json, write_json, copy_mib, out_dir):
    '''The two counter passes' figures as HBM bytes per unit of each group, calibrated on the copy.'''
    fetch, write = json.load(open(fetch_json)), json.load(open(write_json))
    kib = copy_mib * 1024.0
    fetch_scale, write_scale = kib / fetch['copy_kib'], kib / write['copy_kib']     # the counters read KiB
    summary = {'fetch_scale': fetch_scale, 'write_scale': write_scale, 'copy_mib': copy_mib}
    for g, spec in _groups(fetch['roofline']).items():
        fetch_bytes = fetch['sum_kib'][g] * fetch_scale * 1024.0 / spec['units']
        write_bytes = write['sum_kib'][g] * write_scale * 1024.0 / spec['units']
        summary[g] = {'hbm_bytes_per_step': fetch_bytes + write_bytes, 'fetch_bytes_per_step': fetch_bytes,
                      'write_bytes_per_step': write_bytes, 'algorithmic_bytes_per_step': spec['algorithmic'],
                      'units': spec['units'], 'unit': spec['unit']}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, 'dmrg_matvec_gemm_L64_chi1024_pmc.json'), 'w') as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary))
    return summary


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--per-kernel', action='store_true',
                        help='print a per-kernel table of two counter passes (--fetch, --write directories)')
    parser.add_argument('--reduce', help='one profiled run\'s output directory, reduced to --out (JSON)')
    parser.add_argument('--combine', action='store_true',
                        help='the reduced FETCH_SIZE and WRITE_SIZE passes (--fetch, --write JSON) to --out (a directory)')
    for flag in ('--fetch', '--write', '--bench-line', '--out'):
        parser.add_argument(flag)
    parser.add_argument('--copy-mib', type=int, default=256)
    args = parser.parse_args(argv)
    if args.per_kernel:
        return per_kernel(args.fetch, args.write, args.copy_mib)
    if args.reduce:
        if not (args.bench_line and args.out):
            parser.error('--reduce needs --bench-line and --out')
        reduce_pass(args.reduce, args.bench_line, args.out)
        return 0
    if args.combine:
        if not (args.fetch and args.write and args.out):
            parser.error('--combine needs --fetch, --write and --out')
        combine(args.fetch, args.write, args.copy_mib, args.out)
        return 0
    parser.error('one of --per-kernel, --reduce, --combine')
    return 2


if __name__ == '__main__':
    sys.exit(main())
