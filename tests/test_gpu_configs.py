'''
The benchmark configurations on the HIP path, at their full sizes (the truncating L = 16, chi = 32
case is in test_gpu_dmrg.py):

- the chi = 1024 block-einsum microbenchmark (27-sector bond), three-operand Krylov form and
  five-operand form, in FP64 within 1e-12 and in FP32 within 2e-6 (largest error over largest
  entry) of the dense restatement of the same operands (`oracle/dense.py`) contracted by numpy;
- the L = 64, chi = 1024 and L = 100, chi = 2048 chains: from a converged state, every sweep of the
  FP32 form (FP32 MFMA block products, FP64 Lanczos scalars, split and a two-step FP64 defect
  correction) within 1e-10 of the FP64 form's sweep, with both forms' dropped weight reported;
- the uneven SU(2) and Fibonacci contractions the learned schedule is judged on: every launch
  configuration of every block product gives the same result bit for bit, and the exhaustive
  autotuner and the cost model run on them.
'''

import math

import numpy as np
import pytest
import torch

import tentorch_amd as tt
from tentorch_amd import microbench
from tentorch_amd.tensor import Settings, run
import oracle.symmetry as osym
import oracle.tensor as ot
from oracle.dense import dense, flip
from bridge import to_oracle

pytestmark = pytest.mark.gpu


# -- the chi = 1024 block einsum -----------------------------------------------------------------

def _dense_operands(tensors, spec):
    '''Dense arrays of the einsum's operands in one frame: a label's later occurrences are flipped
    to the dual sector's index, as a contraction pairs sector c with dual(c).'''
    seen, out = set(), []
    for t, group in zip(tensors, spec.split('->')[0].split(',')):
        o = to_oracle(t, osym.u1) if isinstance(t, tt.Tensor) else t
        array = dense(o)
        for axis, letter in enumerate(group.lower()):
            if letter in seen:
                array = flip(osym.u1, array, axis, o.legs()[axis])
            seen.add(letter)
        out.append(array)
    return out


@pytest.fixture(scope='module')
def microbench_reference():
    '''The five-operand product of the microbenchmark's FP64 operands, densely on the host:
    L[a,i,e] theta[e,q,s,f] W1[i,p,q,j] W2[j,r,s,k] R[d,k,f] -> [a,p,r,d].'''
    left, first, second, right, theta = microbench.operands(torch.float64, seed=0)
    spec = microbench.EQ5
    l, w1, w2, r, th = _dense_operands([left, first, second, right, theta], spec)
    x = np.tensordot(l, th, axes=([2], [0]))                    # a i q s f
    x = np.einsum('aiqsf,ipqj->asfpj', x, w1, optimize=True)
    x = np.einsum('asfpj,jrsk->afprk', x, w2, optimize=True)
    ref = np.tensordot(x, r, axes=([1, 4], [2, 1]))             # a p r d
    return ref


def _check_against(result: tt.Tensor, ref: np.ndarray, bound: float) -> float:
    got = dense(to_oracle(result, osym.u1))
    error = float(np.abs(got - ref).max() / np.abs(ref).max())
    assert error <= bound, error
    return error


@pytest.mark.parametrize('precision,bound', [('fp64', 1e-12), ('fp32', 2e-6)])
@pytest.mark.parametrize('form', [3, 5])
def test_block_einsum_at_chi_1024_matches_the_dense_restatement(microbench_reference, precision, bound, form):
    dtype = torch.float64 if precision == 'fp64' else torch.float32
    left, first, second, right, theta = microbench.operands(dtype, seed=0)
    if form == 3:
        head = tt.einsum('EIa,IPqj->EPaqj', left, first)
        tail = tt.einsum('JRsk,Fkb->JRSFb', second, right)
        result = tt.einsum(microbench.EQ3, head, tail, theta)
    else:
        result = tt.einsum(microbench.EQ5, left, first, second, right, theta)
    if dtype == torch.float32:
        result = tt.Tensor._wrap(result.structure, result.data.double())
    # EQ3 names the output legs E,P | r,f and EQ5 A,P | r,d: both are theta's legs in theta's order
    _check_against(result, microbench_reference, bound)


def test_block_einsum_per_block_form_at_chi_1024_matches_the_grouped_form():
    left, first, second, right, theta = microbench.operands(torch.float64, seed=0)
    grouped = tt.einsum(microbench.EQ5, left, first, second, right, theta).data
    with Settings.using(mode=1):
        plain = tt.einsum(microbench.EQ5, left, first, second, right, theta).data
    assert (plain - grouped).abs().max() <= 1e-12 * grouped.abs().max()


# -- the long chains: FP32 with defect correction against FP64 ------------------------------------

def _forms_from_one_state(length, bond, warmup, sweeps):
    driver = tt.DMRG(tt.Model(length, cat=tt.u1), max_bond_dim=bond, n_iter=5, seed=0, precision='fp64')
    for _ in range(warmup):
        driver.sweep()
    snapshot = [t.clone() for t in driver.state.tensors], driver.state.center
    out = {}
    for precision in ('fp64', 'fp32'):
        driver.state.tensors = [t.clone() for t in snapshot[0]]
        driver.state.center = snapshot[1]
        driver.precision = precision
        out[precision] = [driver.sweep() for _ in range(sweeps)]
    return out


@pytest.mark.parametrize('length,bond,warmup', [(64, 1024, 4), (100, 2048, 5)])
def test_fp32_form_with_defect_correction_tracks_fp64_per_sweep(length, bond, warmup):
    forms = _forms_from_one_state(length, bond, warmup, 2)
    for sweep, ((e64, lost64), (e32, lost32)) in enumerate(zip(forms['fp64'], forms['fp32'])):
        print(f'L={length} chi={bond} sweep {sweep + 1}: E_fp64 {e64:.13f} E_fp32 {e32:.13f} '
              f'|dE| {abs(e32 - e64):.2e} dropped {lost64:.1e} / {lost32:.1e}')
        assert abs(e32 - e64) <= 1e-10, f'sweep {sweep + 1}'
        # the energy per site is the chain's bulk value to the finite-size correction
        assert abs(e64 / length - (0.25 - math.log(2))) < 5e-3
        assert lost32 < 1e-12 and lost64 < 1e-12


# -- uneven SU(2) and Fibonacci contractions -------------------------------------------------------

def test_uneven_contractions_give_identical_products_under_every_launch_configuration():
    from tentorch_amd._native import engine
    from tentorch_amd.tuning import NUM_CONFIGS, uneven_cases
    for name, operands, spec in uneven_cases(seed=0, trials=1):
        tensors = operands()
        program = engine.plan_einsum(spec, [t.structure for t in tensors], None)
        gemms = [op for op, kind in enumerate(program.op_kinds()) if kind == 'gemm']
        assert gemms, name
        reference = run(program, tensors).data.clone()
        for config in range(NUM_CONFIGS):
            for op in gemms:
                program.set_gemm_config(op, config)
            got = run(program, [tt.Tensor._wrap(t.structure, t.data) for t in tensors]).data
            assert torch.equal(got, reference), f'{name}: configuration {config}'


def test_uneven_contractions_under_the_autotuner_and_the_cost_model():
    '''The exhaustive autotuner times every block product under every configuration and keeps the
    fastest, the cost model predicts one: both results equal the default configuration's, and the
    model's choices cost at most a quarter more than the autotuner's in total.'''
    from tentorch_amd.tuning import CostModel, model_choice, regret, tuner, uneven_cases, default_choice
    tuner.reset()
    tuner.samples.clear()
    for name, operands, spec in uneven_cases(seed=0, trials=1):
        tensors = operands()
        reference = tt.einsum(spec, *tensors).data.clone()
        with Settings.using(tiles='exhaustive'):
            tuned = tt.einsum(spec, *[tt.Tensor._wrap(t.structure, t.data) for t in tensors]).data
        tuner.reset()
        with Settings.using(tiles='model'):
            modelled = tt.einsum(spec, *[tt.Tensor._wrap(t.structure, t.data) for t in tensors]).data
        tuner.reset()
        assert torch.equal(tuned, reference) and torch.equal(modelled, reference), name
    model = regret(tuner.samples, model_choice(CostModel.load()))
    rule = regret(tuner.samples, default_choice)
    print(f'regret over {model["ops"]} block products: model {model}, default rule {rule}')
    assert model['total_time_ratio'] <= 1.25
