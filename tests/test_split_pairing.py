'''
The split's re-pairing of sectors and their duals (`split_over`, svd.cpp:97-119) without a GPU: the
vectorised form in `tentorch_amd.linalg` against a literal walk over the ranked candidates, as the
reference writes it, on random rankings.
'''

import numpy as np
import pytest

from tentorch_amd.linalg import _paired_prefixes


def _walk(cand_owner, cand_position, sectors, room):
    kept = [0] * len(sectors)
    used = {}
    for fi, index in zip(cand_owner.tolist(), cand_position.tolist()):
        s = sectors[fi]
        if s not in room or used.get(s, 0) >= room[s]:
            continue
        used[s] = used.get(s, 0) + 1
        kept[fi] = max(kept[fi], index + 1)
    return np.array(kept, dtype=np.int64)


def test_vectorised_pairing_equals_the_candidate_walk():
    rng = np.random.default_rng(21)
    for trial in range(400):
        blocks = int(rng.integers(1, 12))
        sizes = rng.integers(0, 30, size=blocks)
        # sectors: U(1)-like charges, several blocks may share one
        sectors = [(int(q),) for q in rng.integers(-4, 5, size=blocks)]
        values = np.concatenate([np.sort(rng.random(n))[::-1] for n in sizes]) if sizes.sum() else np.zeros(0)
        owner = np.repeat(np.arange(blocks), sizes)
        position = np.concatenate([np.arange(n) for n in sizes]) if sizes.sum() else np.zeros(0, dtype=np.int64)
        order = np.argsort(-values, kind='stable')[:int(rng.integers(0, values.size + 1))]
        room = {s: int(rng.integers(0, 20)) for s in set(sectors) if rng.random() < 0.8}
        got = _paired_prefixes(owner[order], position[order], sectors, room)
        assert np.array_equal(got, _walk(owner[order], position[order], sectors, room)), trial


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
