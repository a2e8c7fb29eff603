'''
The split's ranking (`split_over`, svd.cpp:72-136) without a GPU: the engine's `rank_split` against
a literal walk over the ranked candidates, as the reference writes it, on random rankings -- U(1)
sectors (a sector and its negation pair up), ties, tolerances and caps, blocks without values -- and
the planted mistake of a re-pairing that skips the rewalk.
'''

import numpy as np
import pytest

import tentorch_amd as tt
from tentorch_amd._native import engine


def _walk(sectors, values, at, max_bond, tol, rewalk=True):
    '''svd.cpp:72-136 step by step: rank, cut, keep prefixes, pair each sector with its dual, and
    when the pairing changes a count walk the candidates again, a sector keeping its first ones.'''
    cands = [(float(values[j]), b, j - at[b]) for b in range(len(sectors)) for j in range(at[b], at[b + 1])]
    cands.sort(key=lambda c: -c[0])
    cands = cands[:next((i for i, c in enumerate(cands) if c[0] <= tol), len(cands))]
    if max_bond is not None:
        cands = cands[:max_bond]
    kept = [0] * len(sectors)
    for _, b, i in cands:
        kept[b] = max(kept[b], i + 1)
    mults = {}
    for b, s in enumerate(sectors):
        if kept[b]:
            mults[s] = mults.get(s, 0) + kept[b]
    paired = {s: min(m, mults[(-s[0],)]) for s, m in mults.items() if (-s[0],) in mults}
    if paired and paired != mults and rewalk:
        kept, used = [0] * len(sectors), {}
        for _, b, i in cands:
            s = sectors[b]
            if s not in paired or used.get(s, 0) >= paired[s]:
                continue
            used[s] = used.get(s, 0) + 1
            kept[b] = max(kept[b], i + 1)
        mults = paired
    nxt, offset = {}, []
    for b, s in enumerate(sectors):
        offset.append(nxt.get(s, 0))
        nxt[s] = offset[-1] + kept[b]
    bond = sorted(mults) or [(0,)]
    mult = [mults[s] for s in sorted(mults)] or [1]
    placed = dict(zip(bond, np.concatenate([[0], np.cumsum(mult)[:-1]]).tolist()))
    kept_values = np.zeros(sum(mult))
    for b, s in enumerate(sectors):
        if kept[b]:
            kept_values[placed[s] + offset[b]:placed[s] + offset[b] + kept[b]] = values[at[b]:at[b] + kept[b]]
    return kept, offset, bond, mult, kept_values


def _random_case(rng):
    blocks = int(rng.integers(1, 12))
    sizes = rng.integers(0, 30, size=blocks)
    sectors = [(int(q),) for q in rng.integers(-4, 5, size=blocks)]
    # a coarse grid of values makes ties common
    values = np.concatenate([np.sort(np.round(rng.random(n), 1 if rng.random() < 0.3 else 12))[::-1] for n in sizes])
    at = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64).tolist()
    max_bond = None if rng.random() < 0.3 else int(rng.integers(1, max(2, values.size + 2)))
    tol = float(rng.choice([0.0, 0.05, 0.5]))
    return sectors, values, at, max_bond, tol


def test_the_engine_ranking_equals_the_candidate_walk():
    rng = np.random.default_rng(21)
    repaired = 0
    for trial in range(600):
        sectors, values, at, max_bond, tol = _random_case(rng)
        got = engine.rank_split(tt.u1, sectors, values, at, -1 if max_bond is None else max_bond, tol)
        want = _walk(sectors, values, at, max_bond, tol)
        assert list(got[0]) == want[0] and list(got[1]) == want[1], trial
        assert list(got[2]) == want[2] and list(got[3]) == want[3], trial
        assert np.array_equal(got[4], want[4]), trial
        repaired += want[0] != _walk(sectors, values, at, max_bond, tol, rewalk=False)[0]
    # the planted mistake (re-pairing without the rewalk) shows in many of the trials
    assert repaired > 50


def test_the_engine_ranking_refuses_a_nan():
    with pytest.raises(ValueError, match='not a number'):
        engine.rank_split(tt.u1, [(0,)], np.array([1.0, np.nan]), [0, 2], -1, 0.0)


def test_svd_call_forms_bind_as_the_reference_overloads():
    '''`bindings.cpp:367-382`: the first overload (tensor, max_bond, tol, absorb) wins whenever it
    fits; a sequence in second place, or a `left` keyword, selects (tensor, left, max_bond, tol,
    absorb); anything else is a TypeError.'''
    from tentorch_amd.linalg import _bind_split_arguments as bind
    assert bind((), {}) == (None, None, 0.0, 0)
    assert bind((20, 1e-14, 2), {}) == (None, 20, 1e-14, 2)
    assert bind((None, 1e-14), {}) == (None, None, 1e-14, 0)
    assert bind(([0, 1], 20, 1e-14, 2), {}) == ([0, 1], 20, 1e-14, 2)
    assert bind((), {'left': (1, 0), 'absorb': 1}) == ([1, 0], None, 0.0, 1)
    assert bind((), {'max_bond': 3}) == (None, 3, 0.0, 0)
    for args, kwargs in (((None, 20, 1e-14, 2), {}), (('01',), {}), ((1.5,), {}), ((), {'rank': 3}),
                         ((20,), {'max_bond': 3})):
        with pytest.raises(TypeError):
            bind(args, kwargs)
