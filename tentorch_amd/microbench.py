'''
Block-einsum microbenchmark: the U(1) two-site effective-Hamiltonian matvec
L[a,w,a'] theta[a,s1,s2,b] W[w,s1,s1',v] W[v,s2,s2',w'] R[b,w',b'] on one GPU.

Bond legs carry chi = 1024 over charges q in [-16, 16] with multiplicity proportional to
exp(-q^2/32), rounded by largest remainder to a total of 1024; the physical legs are spin-1/2
(charges +-1), and the operator leg is the Heisenberg automaton's D = 5 leg (charges 0 x3, +-2).
Entries are N(0, 1) from numpy's generator with a fixed seed, drawn block by block in key order.

Two forms of the same product:
- `3`: the form the Krylov loop runs, `EPaqj,JRSFb,AQsb->EPrf` with head = L W and tail = W R
  contracted beforehand (their layout moves are hoisted after the first call, as inside Lanczos);
- `5`: the five-operand einsum `AIe,IPqj,JRsk,Dkf,EQSf->APrd` in one call.

Every call hands the product a new vector (the last operand), as each Krylov step does, so its
layout move runs every time while the moves of the fixed operands are hoisted after the first
call.  Each runs in the per-block form (one launch per block product and per target block of every
elementary leg move, the reference's execution) and in the grouped form (one composed leg move and
one grouped MFMA launch per contraction step), and, for the grouped 3-operand form, replayed from a
captured hipGraph.  Times are HIP events on the launch stream around `--reps` calls after one warm
call; launches are counted by the device library.

    python -m tentorch_amd.microbench [--precision fp64 fp32 c128] [--reps 50] [--form 3 5]
                                      [--copy-mib 256]

`--copy-mib` adds a plain copy of that size (the calibration kernel for hardware byte counters).
One JSON line per case goes to stdout.
'''

import argparse
import json
import math
import sys

import numpy as np
import torch

from . import _native
from .models import heisenberg
from .symmetry import Leg, u1
from .tensor import Settings, Tensor, einsum, run, workspaces

EQ3 = 'EPaqj,JRSFb,AQsb->EPrf'
EQ5 = 'AIe,IPqj,JRsk,Dkf,EQSf->APrd'


def microbench_bond(total: int = 1024, span: int = 16, width: float = 32.0):
    '''Charges q in [-span, span] with multiplicities proportional to exp(-q^2/width), rounded by
    largest remainder to `total`; sectors that round to zero are dropped.'''
    charges = list(range(-span, span + 1))
    weights = np.array([math.exp(-q * q / width) for q in charges])
    quota = total * weights / weights.sum()
    counts = np.floor(quota).astype(int)
    remainder = quota - counts
    # ties go to the smaller |q|, then to the negative charge (a stable, symmetric-first order)
    order = sorted(range(len(charges)), key=lambda i: (-round(remainder[i], 12), abs(charges[i]), charges[i]))
    for i in order[:total - int(counts.sum())]:
        counts[i] += 1
    kept = [(q, int(c)) for q, c in zip(charges, counts) if c > 0]
    return [(q,) for q, _ in kept], [c for _, c in kept]


def operands(dtype=torch.float64, seed: int = 0, bond=None):
    '''left (E,I | a), two sites of the operator (I,P | q,j), right (F | k,b), theta (A,Q | s,b).'''
    sectors, mults = bond if bond is not None else microbench_bond()
    rng = np.random.default_rng(seed)
    b = Leg(u1, sectors, mults)
    site = heisenberg(u1, 3).tensors[1]
    m_leg, p_leg = site.codomain[0], site.codomain[1]
    left = Tensor.random(u1, [b, m_leg], [b], (0,), dtype=dtype, generator=rng)
    first = Tensor.random(u1, [m_leg, p_leg], [p_leg, m_leg], (0,), dtype=dtype, generator=rng)
    second = Tensor.random(u1, [m_leg, p_leg], [p_leg, m_leg], (0,), dtype=dtype, generator=rng)
    right = Tensor.random(u1, [b], [m_leg, b], (0,), dtype=dtype, generator=rng)
    theta = Tensor.random(u1, [b, p_leg], [p_leg, b], (0,), dtype=dtype, generator=rng)
    return left, first, second, right, theta


def _timed(fn, reps: int):
    '''Milliseconds per call (HIP events on the current stream) and library launches per call.'''
    fn()
    torch.cuda.synchronize()
    before = _native.launch_count()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps, (_native.launch_count() - before) / reps


def fresh(ops):
    '''The operands with the last one (the Krylov vector) wrapped anew, as every Lanczos step hands
    the matvec a new vector: its layout move runs on every call, while the moves of the fixed
    operands stay hoisted.'''
    *fixed, vector = ops
    return (*fixed, Tensor._wrap(vector.structure, vector.data))


def case(form: int, precision: str, mode: str, reps: int, seed: int = 0) -> dict:
    '''One measured configuration; `mode` is plain, grouped or graph.'''
    dtype = {'fp32': torch.float32, 'fp64': torch.float64, 'c128': torch.complex128}[precision]
    left, first, second, right, theta = operands(dtype, seed)
    workspaces.clear()
    Settings.mode = 1 if mode == 'plain' else 0
    try:
        if form == 3:
            head = einsum('EIa,IPqj->EPaqj', left, first)
            tail = einsum('JRsk,Fkb->JRSFb', second, right)
            from ._native import engine
            program = engine.plan_einsum(EQ3, [head.structure, tail.structure, theta.structure], None)
            ops = (head, tail, theta)
            equation = EQ3
        else:
            from ._native import engine
            ops = (left, first, second, right, theta)
            program = engine.plan_einsum(EQ5, [t.structure for t in ops], None)
            equation = EQ5
        out = torch.empty(program.out.size, dtype=dtype, device='cuda')
        if mode == 'graph':
            torch.cuda.synchronize()
            with torch.cuda.stream(torch.cuda.Stream()):
                run(program, ops, out=out)          # uploads the tables, fills the hoisted moves
                before = _native.launch_count()
                graph = _native.capture(lambda: run(program, fresh(ops), out=out))
                launches = _native.launch_count() - before   # kernels recorded into the graph
                ms, _ = _timed(graph.launch, reps)
            torch.cuda.synchronize()
        else:
            ms, launches = _timed(lambda: run(program, fresh(ops), out=out), reps)
        result = (out.real if out.is_complex() else out).double().cpu().numpy()
    finally:
        Settings.mode = 0
    gemm_flops = sum(cost[1] for cost in program.op_costs() if cost[0] == 'gemm')
    return {'case': f'{form}-operand {precision} {mode}', 'equation': equation, 'precision': precision,
            'mode': mode, 'ms': round(ms, 4), 'launches_per_call': launches,
            'gflop': round(gemm_flops / 1e9, 4), 'tflops': round(gemm_flops / (ms * 1e-3) / 1e12, 3),
            'blocks': len(program.out.blocks()), 'checksum': float(np.sum(result * np.cos(np.arange(result.size))))}


def copy_case(mib: int) -> dict:
    '''A plain FP64 copy (read n, write n) of the given size, timed like the other cases.'''
    n = mib * (1 << 20) // 8
    x = torch.randn(n, dtype=torch.float64, device='cuda')
    y = torch.empty_like(x)
    ms, _ = _timed(lambda: _native.axpby(torch.float64, 1.0, x, 0.0, y), 20)
    return {'case': f'copy {mib} MiB', 'ms': round(ms, 4), 'bytes': 2 * n * 8,
            'gbs': round(2 * n * 8 / (ms * 1e-3) / 1e9, 1)}


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--precision', nargs='+', default=['fp64', 'fp32'], choices=('fp64', 'fp32', 'c128'))
    parser.add_argument('--form', nargs='+', type=int, default=[3, 5], choices=(3, 5))
    parser.add_argument('--mode', nargs='+', default=['plain', 'grouped', 'graph'],
                        choices=('plain', 'grouped', 'graph'))
    parser.add_argument('--reps', type=int, default=50)
    parser.add_argument('--copy-mib', type=int, default=0)
    args = parser.parse_args(argv)
    if not torch.cuda.is_available():
        print('the microbenchmark needs a GPU', file=sys.stderr)
        return 1
    sectors, mults = microbench_bond()
    print(f'[microbench] bond: {len(sectors)} sectors, chi = {sum(mults)}, '
          f'multiplicities {dict((s[0], m) for s, m in zip(sectors, mults))}', file=sys.stderr, flush=True)
    for precision in args.precision:
        for form in args.form:
            for mode in args.mode:
                if mode == 'graph' and form != 3:
                    continue
                line = case(form, precision, mode, args.reps)
                print(json.dumps(line), flush=True)
    if args.copy_mib:
        print(json.dumps(copy_case(args.copy_mib)), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
