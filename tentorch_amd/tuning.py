'''
The learned launch schedule of the grouped block products.

Every grouped GEMM launch has a configuration: a tile shape (BM, BN, BK) and a packing of its tiles
into the launch (longest K first, or each block product's tiles on one XCD).  `tt_ggemm_num_configs`
lists them.  Two ways of choosing one per block-product op:

- the exhaustive autotuner runs the op under every configuration, on the program's own data inside
  the program that needs it, and keeps the fastest.  Its timings are also the training data and the
  yardstick of the cost model;
- the cost model is a multilayer perceptron (40 inputs, two hidden layers of 64, ReLU).  It predicts
  log kernel time from a plan's block shapes and a configuration (`Program.gemm_features`) and picks
  the configuration with the least prediction, for plans it has never seen.  Training (mean squared
  error, Adam) and inference run on the GPU through the device library's own kernels
  (`tt_mlp_gradient`, `tt_mlp_adam`, `tt_mlp_forward`).

`Settings.tiles` selects 'default' (the library's rule from block shapes), 'model' or 'exhaustive'.

    python -m tentorch_amd.tuning harvest --out samples.npz    # exhaustive timings of configs 1-3
                                                               # (FP64, FP32 and complex128)
    python -m tentorch_amd.tuning train --data samples.npz --out cost_model.npz
    python -m tentorch_amd.tuning regret --model cost_model.npz  # SU(2) and Fibonacci contractions
'''

import argparse
import json
import math
import os
import sys
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _native

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, 'cost_model.npz')
HIDDEN = 64
NUM_CONFIGS = int(_native.library().tt_ggemm_num_configs())


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class CostModel:
    '''The perceptron, its parameters on the device, and the input/output scaling it was fit with.'''

    def __init__(self, features: int, params: torch.Tensor, x_mean: np.ndarray, x_std: np.ndarray,
                 y_mean: float, y_std: float):
        self.features = features
        self.params = params
        self.x_mean, self.x_std = x_mean.astype(np.float32), x_std.astype(np.float32)
        self.y_mean, self.y_std = float(y_mean), float(y_std)

    @staticmethod
    def param_count(features: int) -> int:
        n = int(_native.library().tt_mlp_param_count(features, HIDDEN))
        if n < 0:
            raise ValueError(_native.last_error())
        return n

    @classmethod
    def initial(cls, features: int, seed: int = 0, device='cuda') -> 'CostModel':
        '''He-scaled normal weights, zero biases, identity scaling.'''
        rng = np.random.default_rng(seed)
        parts = [rng.standard_normal(features * HIDDEN) * math.sqrt(2.0 / features), np.zeros(HIDDEN),
                 rng.standard_normal(HIDDEN * HIDDEN) * math.sqrt(2.0 / HIDDEN), np.zeros(HIDDEN),
                 rng.standard_normal(HIDDEN) * math.sqrt(1.0 / HIDDEN), np.zeros(1)]
        flat = np.concatenate(parts).astype(np.float32)
        assert flat.size == cls.param_count(features)
        return cls(features, torch.tensor(flat, device=device), np.zeros(features), np.ones(features), 0.0, 1.0)

    def _scaled(self, x: np.ndarray) -> torch.Tensor:
        x = (np.asarray(x, dtype=np.float32) - self.x_mean) / self.x_std
        return torch.tensor(np.ascontiguousarray(x), device=self.params.device)

    def predict(self, x: np.ndarray) -> np.ndarray:
        '''Predicted log kernel time (log microseconds) of each row of `x`, on the device.'''
        x = np.atleast_2d(x)
        xs = self._scaled(x)
        y = torch.empty(x.shape[0], dtype=torch.float32, device=self.params.device)
        _native.check(_native.library().tt_mlp_forward(self.features, HIDDEN, x.shape[0], _ptr(xs), _ptr(self.params),
                                                       _ptr(y), _native.stream()), 'cost model')
        return y.cpu().numpy().astype(np.float64) * self.y_std + self.y_mean

    def gradient(self, x: torch.Tensor, target: torch.Tensor):
        '''(gradient, loss) of the mean squared error on scaled device inputs and targets.'''
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.params.device)
        _native.check(_native.library().tt_mlp_gradient(self.features, HIDDEN, x.shape[0], _ptr(x), _ptr(target),
                                                        _ptr(self.params), _ptr(grad), _ptr(loss), _native.stream()),
                      'cost model gradient')
        return grad, loss

    def fit(self, x: np.ndarray, y: np.ndarray, epochs: int = 400, lr: float = 2e-3, batch: int = 256,
            seed: int = 0) -> List[float]:
        '''Fits log time `y` (log microseconds) to features `x` with Adam; returns the loss per epoch.'''
        x = np.asarray(x, dtype=np.float32)
        y = np.asarray(y, dtype=np.float64)
        self.x_mean, self.x_std = x.mean(0), np.maximum(x.std(0), 1e-3)
        self.y_mean, self.y_std = float(y.mean()), float(max(y.std(), 1e-3))
        xs = self._scaled(x)
        ys = torch.tensor(((y - self.y_mean) / self.y_std).astype(np.float32), device=self.params.device)
        m, v = torch.zeros_like(self.params), torch.zeros_like(self.params)
        rng = np.random.default_rng(seed)
        losses, step = [], 0
        lib = _native.library()
        for _ in range(epochs):
            order = torch.tensor(rng.permutation(x.shape[0]), device=self.params.device)
            total = torch.zeros(1, dtype=torch.float32, device=self.params.device)
            for at in range(0, x.shape[0], batch):
                index = order[at:at + batch]
                xb, yb = xs[index].contiguous(), ys[index].contiguous()
                grad, loss = self.gradient(xb, yb)
                step += 1
                _native.check(lib.tt_mlp_adam(self.params.numel(), _ptr(self.params), _ptr(grad), _ptr(m), _ptr(v),
                                              lr, 0.9, 0.999, 1e-8, step, _native.stream()), 'cost model step')
                total += loss * index.numel()
            losses.append(float(total.item()) / x.shape[0])
        return losses

    def choose(self, features: np.ndarray) -> int:
        '''The configuration (row of `features`) with the least predicted time.'''
        return int(np.argmin(self.predict(features)))

    def save(self, path: str) -> None:
        np.savez(path, params=self.params.cpu().numpy(), x_mean=self.x_mean, x_std=self.x_std,
                 y=np.array([self.y_mean, self.y_std]), features=np.array([self.features]))

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL, device='cuda') -> 'CostModel':
        data = np.load(path)
        return cls(int(data['features'][0]), torch.tensor(data['params'], device=device), data['x_mean'],
                   data['x_std'], float(data['y'][0]), float(data['y'][1]))


def element_size(args) -> int:
    '''Bytes per element of the block products of a program run with these arguments.'''
    compute = args[10] if len(args) > 10 else args[0]
    return {_native.TT_F64: 8, _native.TT_F32: 4, _native.TT_C128: 16}[compute]


def measure(program, args, op: int, configs=None, reps: int = 3) -> np.ndarray:
    '''Kernel time (microseconds, least of `reps`) of block-product op `op` under each configuration,
    run on the data the program's earlier ops left in its workspace.  Restores the op's configuration.'''
    configs = range(NUM_CONFIGS) if configs is None else configs
    keep = program.gemm_config(op)
    base, extra = args[:7], args[9:] if len(args) > 7 else ()
    times = np.full(NUM_CONFIGS, np.inf)
    try:
        for c in configs:
            program.set_gemm_config(op, c)
            program.prepare()
            program.run(*base, op, op + 1, *extra)          # warm
            before = _native.launch_count()
            _native.timing(True)
            try:
                for _ in range(reps):
                    program.run(*base, op, op + 1, *extra)
            finally:
                _native.timing(False)
            durations = _native.timing_collect()
            # a product split along its depth runs two launches (parts, then their sum): a run's
            # time is the sum of its launches
            per_run = (_native.launch_count() - before) // reps
            times[c] = min(sum(durations[r * per_run:(r + 1) * per_run]) for r in range(reps)) * 1e3
    finally:
        program.set_gemm_config(op, keep)
    return times


class Tuner:
    '''Picks the configuration of every block product of each program the first time it runs.'''

    def __init__(self):
        self.done: Dict[int, List[int]] = {}
        self.samples: List[tuple] = []      # (features [configs x F], times [configs], default config)
        self.model: Optional[CostModel] = None
        self.seconds = 0.0

    def reset(self) -> None:
        self.done.clear()

    def before(self, program, args, mode: str) -> bool:
        '''Called before a program runs; returns True when the program must run once more after
        measurement (exhaustive mode measures on the data of a first run).'''
        if program.id in self.done or mode == 'default':
            return False
        if mode == 'model':
            start = time.perf_counter()
            model = self.model or CostModel.load()
            self.model = model
            choices = []
            size = element_size(args)
            for op, kind in enumerate(program.op_kinds()):
                if kind == 'gemm':
                    c = model.choose(program.gemm_features(op, size))
                    program.set_gemm_config(op, c)
                    choices.append(c)
            self.done[program.id] = choices
            self.seconds += time.perf_counter() - start
            return False
        return True

    def after_first_run(self, program, args) -> None:
        '''Exhaustive mode: the program ran once with its current configurations; time every block
        product under every configuration and keep the fastest.'''
        start = time.perf_counter()
        size = element_size(args)
        choices = []
        for op, kind in enumerate(program.op_kinds()):
            if kind != 'gemm':
                continue
            default = program.gemm_config(op)
            times = measure(program, args, op)
            self.samples.append((program.gemm_features(op, size), times, default))
            best = int(np.argmin(times))
            program.set_gemm_config(op, best)
            choices.append(best)
            if len(self.samples) % 250 == 0:
                print(f'[tuning] {len(self.samples)} block products timed under {NUM_CONFIGS} configurations, '
                      f'{self.seconds + time.perf_counter() - start:.1f} s', file=sys.stderr, flush=True)
        self.done[program.id] = choices
        self.seconds += time.perf_counter() - start


tuner = Tuner()


def samples_arrays(samples):
    '''Flattens (features, times, default) per op into training rows: X [rows x F], log-microsecond y,
    op index.'''
    xs, ys, ops = [], [], []
    for k, (features, times, _) in enumerate(samples):
        for c in range(len(times)):
            if np.isfinite(times[c]) and times[c] > 0:
                xs.append(features[c])
                ys.append(math.log(times[c]))
                ops.append(k)
    return np.array(xs, dtype=np.float32), np.array(ys), np.array(ops)


def regret(samples, choose) -> dict:
    '''Mean and worst ratio t(choice) / t(best) - 1 over ops, and the summed-time ratio; `choose`
    maps (features, default config) to a configuration.'''
    ratios, chosen, best = [], 0.0, 0.0
    for features, times, default in samples:
        c = choose(features, default)
        ratios.append(times[c] / times.min() - 1.0)
        chosen += times[c]
        best += times.min()
    return {'ops': len(samples), 'mean_regret': float(np.mean(ratios)), 'max_regret': float(np.max(ratios)),
            'total_time_ratio': chosen / best}


def default_choice(features, default) -> int:
    '''The device library's own rule from block shapes (what the plan runs with when nobody chooses).'''
    return int(default)


def model_choice(model: 'CostModel'):
    return lambda features, default: model.choose(features)


def save_samples(path: str, samples) -> None:
    np.savez(path, features=np.stack([f for f, _, _ in samples]), times=np.stack([t for _, t, _ in samples]),
             defaults=np.array([d for _, _, d in samples]))


def load_samples(path: str):
    data = np.load(path)
    return list(zip(data['features'], data['times'], data['defaults']))


# -- workloads -------------------------------------------------------------------------------------

def _dmrg_sweeps(length, bond, sweeps, precision, dtype=None):
    from .algorithms import DMRG
    from .models import Model
    from .symmetry import u1
    driver = DMRG(Model(length, cat=u1), max_bond_dim=bond, n_iter=5, seed=0, precision=precision, dtype=dtype)
    for s in range(sweeps):
        start = time.perf_counter()
        driver.sweep()
        torch.cuda.synchronize()
        print(f'[tuning] DMRG L={length} chi={bond} {precision} sweep {s + 1}: {time.perf_counter() - start:.1f} s, '
              f'{len(tuner.samples)} ops measured', file=sys.stderr, flush=True)


def _microbench():
    from . import microbench
    for precision in ('fp64', 'fp32', 'c128'):
        for form in (3, 5):
            microbench.case(form, precision, 'grouped', 2)
    print(f'[tuning] microbench: {len(tuner.samples)} ops measured', file=sys.stderr, flush=True)


def uneven_cases(seed: int = 0, trials: int = 4):
    '''Contractions over uneven blocks: random SU(2) tensors whose legs carry spins j = 0..4
    (labels 2j) with multiplicities ~ Poisson(12 exp(-j/2)), and Fibonacci tensors with 1 and tau
    multiplicities ~ Poisson(40), N(0, 1) entries.  Each contraction takes the middle leg of a
    three-leg codomain, so the operand's move recouples two legs through F and R symbols before the
    block products.  Fibonacci symbols carry phases, so those tensors are complex128.  Yields
    (name, operands, spec): `operands()` draws the tensors on the device.'''
    from .symmetry import Leg, fibonacci, su2
    from .tensor import Tensor
    rng = np.random.default_rng(seed)

    def su2_leg():
        sectors, mults = [], []
        for j in range(5):
            m = int(rng.poisson(12.0 * math.exp(-j / 2.0)))
            if m > 0:
                sectors.append((2 * j,))
                mults.append(m)
        return Leg(su2, sectors, mults)

    def fib_leg():
        return Leg(fibonacci, [(0,), (1,)], [max(1, int(rng.poisson(40))), max(1, int(rng.poisson(40)))])

    def drawn(cat, legs, shapes, generator_seed, dtype=torch.float64):
        def operands():
            g = np.random.default_rng(generator_seed)
            return [Tensor.random(cat, [legs[i] for i in cod], [legs[i] for i in dom], cat.trivial(), dtype=dtype,
                                  generator=g) for cod, dom in shapes]
        return operands

    for trial in range(trials):
        legs = [su2_leg() for _ in range(6)]
        yield (f'SU(2) {trial} ABCd,Be->ACde',
               drawn(su2, legs, [((0, 1, 2), (3,)), ((1,), (4,))], int(rng.integers(1 << 30))), 'ABCd,Be->ACde')
        yield (f'SU(2) {trial} ABCd,CEf->AEBdf',
               drawn(su2, legs, [((0, 1, 2), (3,)), ((2, 4), (5,))], int(rng.integers(1 << 30))), 'ABCd,CEf->AEBdf')
        legs = [fib_leg() for _ in range(5)]
        yield (f'Fibonacci {trial} ABCd,Be->ACde',
               drawn(fibonacci, legs, [((0, 1, 2), (3,)), ((1,), (4,))], int(rng.integers(1 << 30)),
                     torch.complex128), 'ABCd,Be->ACde')


def uneven_contractions(seed: int = 0, trials: int = 4):
    '''`uneven_cases` as (name, thunk): a thunk draws its tensors and runs the einsum.'''
    from .tensor import einsum

    def thunk_of(operands, spec):
        def thunk():
            tensors = operands()
            out = einsum(spec, *tensors)
            torch.cuda.synchronize()
            del out, tensors
        return thunk

    for name, operands, spec in uneven_cases(seed, trials):
        yield name, thunk_of(operands, spec)


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = parser.add_subparsers(dest='command', required=True)
    h = sub.add_parser('harvest')
    h.add_argument('--out', required=True)
    h.add_argument('--sweeps', type=int, default=3)
    t = sub.add_parser('train')
    t.add_argument('--data', required=True)
    t.add_argument('--out', required=True)
    t.add_argument('--epochs', type=int, default=150)
    r = sub.add_parser('regret')
    r.add_argument('--model', default=DEFAULT_MODEL)
    r.add_argument('--out', default=None)
    args = parser.parse_args(argv)
    if not torch.cuda.is_available():
        print('tuning needs a GPU', file=sys.stderr)
        return 1
    if args.command == 'harvest':
        from .tensor import Settings
        with Settings.using(tiles='exhaustive'):
            _dmrg_sweeps(16, 32, 2, 'fp64')
            # the reference engine's default scalar type
            _dmrg_sweeps(16, 32, 2, 'fp64', dtype=torch.complex128)
            _microbench()
            _dmrg_sweeps(64, 1024, args.sweeps, 'fp64')
            tuner.reset()
            _dmrg_sweeps(64, 1024, 1, 'fp32')
        save_samples(args.out, tuner.samples)
        print(json.dumps({'harvested_ops': len(tuner.samples), 'tuning_seconds': round(tuner.seconds, 1)}))
        return 0
    if args.command == 'train':
        samples = load_samples(args.data)
        x, y, ops = samples_arrays(samples)
        model = CostModel.initial(x.shape[1], seed=0)
        start = time.perf_counter()
        losses = model.fit(x, y, epochs=args.epochs)
        model.save(args.out)
        fitted = regret(samples, model_choice(model))
        print(json.dumps({'rows': int(x.shape[0]), 'ops': len(samples), 'first_loss': losses[0],
                          'final_loss': losses[-1], 'train_seconds': round(time.perf_counter() - start, 1),
                          'train_regret_model': fitted, 'train_regret_default': regret(samples, default_choice)}))
        return 0
    if args.command == 'regret':
        tuner.reset()
        tuner.samples.clear()
        with_model = CostModel.load(args.model)
        from .tensor import Settings
        for name, work in uneven_contractions():
            before = len(tuner.samples)
            start = time.perf_counter()
            with Settings.using(tiles='exhaustive'):
                work()
            name += f' ({time.perf_counter() - start:.1f} s)'
            print(f'[tuning] {name}: {len(tuner.samples) - before} block products', file=sys.stderr, flush=True)
        result = {'model': regret(tuner.samples, model_choice(with_model)),
                  'default': regret(tuner.samples, default_choice)}
        if args.out:
            save_samples(args.out, tuner.samples)
        print(json.dumps(result))
        return 0
    return 2


if __name__ == '__main__':
    # Run from the package's own module: tensor.run reports to `tentorch_amd.tuning.tuner`, which
    # under `python -m` is not this __main__ module's copy.
    from tentorch_amd.tuning import main as package_main
    sys.exit(package_main())
