'''
The numpy oracle against the reference's own engine, built from its sources by
`oracle/build_ref.py` (skipped where it is not built): the two-site DMRG of both, from the same
starting state, agrees sweep by sweep, and the matvec einsum and the truncating split agree on
random operands.  This pins the oracle's restatement of `einsum.cpp`, `place.cpp`, `contract.cpp`
and `svd.cpp` on the U(1) path against the code it restates.
'''

import numpy as np
import pytest
import torch

import oracle.dmrg as od
import oracle.symmetry as osym
import oracle.tensor as ot
from oracle import ref_engine

ext = ref_engine.engine()
pytestmark = pytest.mark.skipif(ext is None, reason='the reference engine is not built (python oracle/build_ref.py)')


def _blocks(t):
    return {(tuple(k.codomain), tuple(k.domain), k.copy): b.numpy() for k, b in t.to_tuple()[3].items()}


def _same(reference, oracle_tensor, tol):
    blocks = _blocks(reference)
    assert set(blocks) == set(oracle_tensor.blocks), 'block keys'
    for key, b in oracle_tensor.blocks.items():
        assert np.abs(blocks[key] - b).max() <= tol * max(1.0, np.abs(b).max()), key


@pytest.mark.parametrize('length,bond', [(8, 16), (16, 32)])
def test_dmrg_sweeps_agree_with_the_reference_engine(length, bond):
    _, reference = ref_engine.sweeps(length, bond, 3, threads=1)
    driver = od.DMRG(od.Model(length), bond)
    mine = [driver.sweep()[0] for _ in range(3)]
    assert np.abs(np.array(mine) - np.array(reference)).max() <= 1e-10, (mine, reference)
    if length == 8:
        assert abs(reference[-1] - od.exact_energy(8)) <= 1e-11


def _random(codomain, domain, rng):
    return ot.random_tensor(osym.u1, codomain, domain, (0,), rng, np.float64)


def test_matvec_einsum_and_split_agree_with_the_reference_engine():
    rng = np.random.default_rng(3)
    u1 = osym.u1
    bond = osym.Leg(u1, [(-2,), (0,), (2,)], [2, 3, 2])
    phys = osym.Leg(u1, [(-1,), (1,)], [1, 1])
    mpo = osym.Leg(u1, [(-2,), (0,), (2,)], [1, 3, 1])
    head = _random([bond, phys], [bond, phys, mpo], rng)           # EPaqj
    tail = _random([mpo, phys, phys, bond], [bond], rng)          # JRSFb
    theta = _random([bond, phys], [phys, bond], rng)              # AQsb
    spec = 'EPaqj,JRSFb,AQsb->EPrf'
    mine = ot.einsum(spec, head, tail, theta)
    theirs = ext.einsum(spec, *[ref_engine.to_reference(ext, t) for t in (head, tail, theta)])
    _same(theirs, mine, 1e-12)
    for absorb in (1, 2):
        u, s, v = ot.svd(mine, None, 5, 1e-14, absorb)
        ru, rs, rv = ext.svd(theirs, 5, 1e-14, absorb)
        assert np.abs(np.sort(np.asarray(s).ravel()) - np.sort(rs.numpy().ravel())).max() <= 1e-12
        product = ot.contract(u, v, ot.domain_legs(u), [0])
        reference_product = ext.contract(ru, rv, ext.domain_legs(ru), [0])
        _same(reference_product, product, 1e-12)


# the legs of tests/test_engine.py's placement test, per category
_LEGS = {
    'u1': [([(-2,), (0,), (2,)], [2, 3, 2]), ([(-1,), (1,)], [1, 1])],
    'z3': [([(0,), (1,), (2,)], [2, 1, 1]), ([(0,), (1,), (2,)], [1, 1, 1])],
    'fermion_parity': [([(0,), (1,)], [2, 2]), ([(0,), (1,)], [1, 1])],
    'su2': [([(0,), (1,), (2,)], [2, 1, 2]), ([(1,)], [1])],
    'fibonacci': [([(0,), (1,)], [2, 3]), ([(0,), (1,)], [1, 1])],
}
_CATS = {'u1': osym.u1, 'z3': osym.z_n(3), 'fermion_parity': osym.fermion_parity, 'su2': osym.su2,
         'fibonacci': osym.fibonacci}


def _pair(name, cod, dom, seed=0):
    cat = _CATS[name]
    dtype = np.complex128 if name == 'fibonacci' else np.float64
    rng = np.random.default_rng(seed)
    o = ot.random_tensor(cat, [osym.Leg(cat, *l) for l in cod], [osym.Leg(cat, *l) for l in dom], cat.trivial(), rng,
                         dtype)
    return o, ref_engine.to_reference(ext, o)


@pytest.mark.parametrize('name', sorted(_LEGS))
def test_every_placement_agrees_with_the_reference_engine(name):
    '''All 5 x 24 placements of a 4-leg tensor: exchanges (braiding, `regroup.cpp:25-258`) and
    bends (`regroup.cpp:260-417`, the dual convention and the F-entry read included) of the oracle
    against the reference's own `Tensor.place`, for abelian, fermionic, SU(2) and Fibonacci
    categories: the recoupling gauge of the oracle is the reference's.'''
    import itertools
    a, b = _LEGS[name]
    o, r = _pair(name, [a, b], [b, a])
    for n_above in range(5):
        for perm in itertools.permutations(range(4)):
            above, below = list(perm[:n_above]), list(perm[n_above:])
            _same(r.place(above, below), o.place(above, below), 1e-12)


@pytest.mark.parametrize('name', ['su2', 'fibonacci'])
def test_recoupling_einsums_agree_with_the_reference_engine(name):
    '''Contractions whose first move recouples two legs (the learned schedule's uneven cases).'''
    a, b = _LEGS[name]
    first, rfirst = _pair(name, [a, b, a], [b], seed=1)
    second, rsecond = _pair(name, [b], [a], seed=2)
    for spec, operands, reference in (('ABCd,Be->ACde', (first, second), (rfirst, rsecond)),):
        _same(ext.einsum(spec, *reference), ot.einsum(spec, *operands), 1e-12)
