'''
The learned launch schedule without a GPU: the committed cost model, evaluated by a float64 numpy
restatement of its forward pass, against the exhaustive autotuner's timings of contractions it was
not trained on (`tests/data/uneven_contraction_timings.npz`: every block product of SU(2) and
Fibonacci contractions that recouple two legs, timed under every launch configuration on an MI355X
by `python -m tentorch_amd.tuning regret`).  The model must choose better than the device library's
rule, which is what every block product runs with when nobody chooses.
'''

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIDDEN = 64


def model_forward(path, x):
    '''Predicted log time of rows `x` [n, F] under the packed perceptron stored at `path`.'''
    data = np.load(path)
    f = int(data['features'][0])
    p = data['params'].astype(np.float64)
    at, parts = 0, []
    for shape in ((f, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN,), (1,)):
        n = int(np.prod(shape))
        parts.append(p[at:at + n].reshape(shape))
        at += n
    assert at == p.size
    w1, b1, w2, b2, w3, b3 = parts
    xs = (np.asarray(x, dtype=np.float64) - data['x_mean']) / data['x_std']
    h1 = np.maximum(xs @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return (h2 @ w3 + b3[0]) * data['y'][1] + data['y'][0]


def regrets(times, choices):
    return np.array([t[c] / t.min() - 1.0 for t, c in zip(times, choices)])


def test_learned_schedule_beats_the_default_rule_on_unseen_contractions():
    data = np.load(os.path.join(HERE, 'data', 'uneven_contraction_timings.npz'))
    features, times, defaults = data['features'], data['times'], data['defaults']
    assert features.shape[0] == times.shape[0] == defaults.shape[0] >= 12
    assert np.all(np.isfinite(times)) and np.all(times > 0)
    model = os.path.join(ROOT, 'tentorch_amd', 'cost_model.npz')
    chosen = [int(np.argmin(model_forward(model, f))) for f in features]
    learned, rule = regrets(times, chosen), regrets(times, defaults)
    assert learned.mean() < rule.mean()
    assert learned.max() < rule.max()
    assert learned.mean() < 0.10


def test_model_file_matches_the_feature_count_of_the_engine():
    import tentorch_amd  # noqa: F401
    from tentorch_amd._native import engine
    data = np.load(os.path.join(ROOT, 'tentorch_amd', 'cost_model.npz'))
    f = int(data['features'][0])
    assert data['params'].size == f * HIDDEN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
    assert data['x_mean'].shape == (f,) and data['x_std'].shape == (f,)
    # the engine's feature rows have the same width
    from tentorch_amd.symmetry import Leg, u1
    bond = Leg(u1, [(-1,), (1,)], [3, 4])
    s = engine.Structure(u1, [bond], [bond], (0,))
    program = engine.plan_contract(s, s, [1], [0], None)
    op = program.op_kinds().index('gemm')
    rows = program.gemm_features(op, 8)
    assert rows.shape[1] == f
