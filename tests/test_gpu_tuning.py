'''
The learned launch schedule: the cost model's forward and gradient kernels against torch's fp32
autograd on the CPU, training through the Adam kernel, and the autotuner inside a program (the
result never depends on the configuration chosen).
'''

import ctypes

import numpy as np
import pytest
import torch

import cabi

pytestmark = pytest.mark.gpu

HIDDEN = 64


@pytest.fixture(scope='module')
def lib():
    return cabi.load()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _unpack(params, f):
    '''The packed parameter vector as torch tensors: W1 [f, H], b1, W2 [H, H], b2, w3, b3.'''
    at, out = 0, []
    for shape in ((f, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN,), (1,)):
        n = int(np.prod(shape))
        out.append(params[at:at + n].reshape(shape))
        at += n
    return out


def _torch_forward(x, params, f):
    w1, b1, w2, b2, w3, b3 = _unpack(params, f)
    h1 = torch.relu(x @ w1 + b1)
    h2 = torch.relu(h1 @ w2 + b2)
    return h2 @ w3 + b3


def _random_model(f, seed):
    rng = np.random.default_rng(seed)
    n = f * HIDDEN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
    params = (rng.standard_normal(n) * 0.3).astype(np.float32)
    return params


@pytest.mark.parametrize('f,batch', [(40, 1), (40, 517), (7, 1000), (64, 64)])
def test_cost_model_forward_matches_torch(lib, f, batch):
    assert lib.tt_mlp_param_count(f, HIDDEN) == f * HIDDEN + 3 * HIDDEN + HIDDEN * HIDDEN + 1
    rng = np.random.default_rng(f + batch)
    params = _random_model(f, 1)
    x = rng.standard_normal((batch, f)).astype(np.float32)
    dp, dx = torch.tensor(params, device='cuda'), torch.tensor(x, device='cuda')
    y = torch.empty(batch, dtype=torch.float32, device='cuda')
    assert lib.tt_mlp_forward(f, HIDDEN, batch, dx.data_ptr(), dp.data_ptr(), y.data_ptr(), _stream()) == 0, cabi.error(lib)
    ref = _torch_forward(torch.tensor(x, dtype=torch.float64), torch.tensor(params, dtype=torch.float64), f)
    assert np.abs(y.cpu().numpy() - ref.numpy()).max() <= 1e-4 * max(1.0, float(ref.abs().max()))


@pytest.mark.parametrize('f,batch', [(40, 3), (40, 700), (13, 257)])
def test_cost_model_gradient_matches_torch_autograd(lib, f, batch):
    '''The fused forward-backward kernel's gradient of mean((y - t)^2) against torch autograd in
    float64 on the CPU, for every parameter block, and the loss it reports.'''
    rng = np.random.default_rng(100 + batch)
    params = _random_model(f, 2)
    x = rng.standard_normal((batch, f)).astype(np.float32)
    t = rng.standard_normal(batch).astype(np.float32)
    dp, dx, dt = (torch.tensor(v, device='cuda') for v in (params, x, t))
    grad = torch.empty_like(dp)
    loss = torch.empty(1, dtype=torch.float32, device='cuda')
    assert lib.tt_mlp_gradient(f, HIDDEN, batch, dx.data_ptr(), dt.data_ptr(), dp.data_ptr(), grad.data_ptr(),
                               loss.data_ptr(), _stream()) == 0, cabi.error(lib)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    ref_loss = ((_torch_forward(torch.tensor(x, dtype=torch.float64), p, f) - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    ref_loss.backward()
    got = grad.cpu().numpy().astype(np.float64)
    ref = p.grad.numpy()
    for name, g, r in zip(('W1', 'b1', 'W2', 'b2', 'w3', 'b3'), _unpack(torch.tensor(got), f), _unpack(torch.tensor(ref), f)):
        scale = max(1e-3, float(r.abs().max()))
        assert float((g - r).abs().max()) <= 2e-4 * scale, name
    assert abs(float(loss.item()) - float(ref_loss)) <= 1e-4 * float(ref_loss)


def test_cost_model_training_fits_a_smooth_function():
    '''Adam on the device kernels fits y = sin(x0) + x1 x2 from 2000 samples: the loss falls by two
    orders of magnitude and held-out predictions follow.'''
    from tentorch_amd.tuning import CostModel
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, (2500, 6)).astype(np.float32)
    y = np.sin(x[:, 0]) + x[:, 1] * x[:, 2]
    model = CostModel.initial(6, seed=0)
    losses = model.fit(x[:2000], y[:2000], epochs=150, lr=3e-3, batch=128)
    assert losses[-1] < losses[0] / 100
    pred = model.predict(x[2000:])
    assert np.sqrt(np.mean((pred - y[2000:]) ** 2)) < 0.1 * np.std(y)


def test_autotuned_einsum_equals_the_default_one():
    '''The exhaustive autotuner times every configuration of every block product inside the program
    and keeps the fastest; the einsum's result is bit-for-bit the default configuration's.'''
    import tentorch_amd as tt
    from tentorch_amd.tensor import Settings
    from tentorch_amd.tuning import tuner
    rng = np.random.default_rng(9)
    bond = tt.Leg(tt.u1, [(-2,), (0,), (2,)], [20, 33, 20])
    phys = tt.Leg(tt.u1, [(-1,), (1,)], [1, 1])
    a = tt.Tensor.random(tt.u1, [bond, phys], [phys, bond], (0,), generator=rng)
    b = tt.Tensor.random(tt.u1, [phys, bond], [phys, bond], (0,), generator=rng)
    ref = tt.einsum('ABcd,CDef->AEBf', a, b).data.clone()
    tuner.reset()
    before = len(tuner.samples)
    with Settings.using(tiles='exhaustive'):
        got = tt.einsum('ABcd,CDef->AEBf', a, b).data
        again = tt.einsum('ABcd,CDef->AEBf', a, b).data
    assert len(tuner.samples) > before
    features, times, default = tuner.samples[-1]
    assert features.shape[0] == times.shape[0] and np.all(np.isfinite(times)) and np.all(times > 0)
    assert torch.equal(got, ref) and torch.equal(again, ref)


def test_device_inference_makes_the_cpu_forward_choices():
    '''The committed cost model run by `tt_mlp_forward` agrees with a float64 CPU forward pass on the
    unseen contractions' feature rows (predictions within 1e-3 in log time).'''
    import os
    from tentorch_amd.tuning import CostModel, DEFAULT_MODEL
    from test_tuning import model_forward
    data = np.load(os.path.join(os.path.dirname(__file__), 'data', 'uneven_contraction_timings.npz'))
    model = CostModel.load(DEFAULT_MODEL)
    for rows in data['features']:
        got, ref = model.predict(rows), model_forward(DEFAULT_MODEL, rows)
        assert np.abs(got - ref).max() <= 1e-3 * max(1.0, np.abs(ref).max())
