'''
The reduction of profiled bench runs (`tentorch_amd.pmc`): of a kernel trace and of one counter
pass, the last dispatches of each kernel family the bench line's roofline counts (its instrumented
sweep is the last on the device), and the calibration on the copy's counter values.  Synthetic
rocprofv3 files written here stand in for a run.
'''

import csv
import json

import pytest

from tentorch_amd import pmc


def _bench_line(path, launches=3, steps=2, gemm_launches=3):
    roof = {'bound': 'hbm', 'launches': launches, 'algorithmic_bytes_per_launch': 1000.0,
            'matvec_gemm': {'steps': steps, 'launches': gemm_launches, 'algorithmic_bytes_per_step': 500.0}}
    path.write_text(json.dumps({'roofline': roof}) + '\n')


def _write(path, fields, rows):
    with open(path, 'w', newline='') as f:
        writer = csv.DictWriter(f, fieldnames=fields)
        writer.writeheader()
        writer.writerows(rows)


def _kernels():
    # an early sweep's step, then the instrumented sweep's three steps, two products, one reduce,
    # and the copy last
    return (['void tridiag_step_kernel<4>(Batch const*)'] + ['void tridiag_step_kernel<8>(Batch const*)'] * 3
            + ['ggemm_kernel<float, double, float, float, 64>(x)', 'ggemm_kernel<float, float, double, float, 64>(x)',
               'ggemm_reduce_kernel<double, float>(x)', 'axpby_kernel<double>(long)'])


def test_counter_passes_reduce_to_the_last_sweep_and_calibrate(tmp_path):
    _bench_line(tmp_path / 'bench.json')
    for counter, scale in (('FETCH_SIZE', 2.0), ('WRITE_SIZE', 1.0)):
        d = tmp_path / counter
        d.mkdir()
        rows = [{'Dispatch_Id': i, 'Kernel_Name': k, 'Counter_Name': counter,
                 'Counter_Value': (1000.0 if i == 0 else 10.0 * (i + 1)) if 'axpby' not in k else 256 * 1024 / scale}
                for i, k in enumerate(_kernels())]
        # the profiler does not always write dispatches in order
        _write(d / 'run_counter_collection.csv', ['Dispatch_Id', 'Kernel_Name', 'Counter_Name', 'Counter_Value'],
               rows[::-1])
        pmc.reduce_pass(str(d), str(tmp_path / 'bench.json'), str(tmp_path / f'{counter}.json'))
    reduced = json.load(open(tmp_path / 'FETCH_SIZE.json'))
    assert reduced['sum_kib']['tridiag'] == 20.0 + 30.0 + 40.0      # the early step (1000 KiB) is left out
    assert reduced['sum_kib']['f32'] == 50.0 + 60.0 + 70.0
    out = pmc.combine(str(tmp_path / 'FETCH_SIZE.json'), str(tmp_path / 'WRITE_SIZE.json'), 256, str(tmp_path / 'out'))
    assert out['fetch_scale'] == pytest.approx(2.0) and out['write_scale'] == pytest.approx(1.0)
    # per launch: (90 KiB x 2 + 90 KiB x 1) / 3
    assert out['tridiag']['hbm_bytes_per_step'] == pytest.approx(90 * 1024 * 3 / 3)
    assert out['f32']['hbm_bytes_per_step'] == pytest.approx(180 * 1024 * 3 / 2)


def test_trace_reduces_to_the_last_sweep(tmp_path):
    _bench_line(tmp_path / 'bench.json')
    d = tmp_path / 'trace'
    d.mkdir()
    rows = [{'Dispatch_Id': i, 'Kernel_Name': k, 'Start_Timestamp': 1000 * i, 'End_Timestamp': 1000 * i + 100 * (i + 1)}
            for i, k in enumerate(_kernels())]
    _write(d / 'run_kernel_trace.csv', ['Dispatch_Id', 'Kernel_Name', 'Start_Timestamp', 'End_Timestamp'], rows)
    reduced = pmc.reduce_pass(str(d), str(tmp_path / 'bench.json'), str(tmp_path / 'trace.json'))
    assert reduced['avg_us']['tridiag:tridiag'] == pytest.approx((0.2 + 0.3 + 0.4) / 3)
    assert reduced['avg_us_per_unit']['f32'] == pytest.approx((0.5 + 0.6 + 0.7) / 2)
