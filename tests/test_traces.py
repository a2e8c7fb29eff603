'''
Einsum traces inside one operand (`tentorch/linalg/graded/einsum.cpp:95-159,386-439`) on the CPU:
the oracle's restatement against the reference's own engine (where it is built) and, for the bosonic
abelian categories, against the dense array, where a trace is a numpy diagonal sum; then the host
engine's planned tables, interpreted by tests/emulate.py, against the oracle, for every category.
'''

import numpy as np
import pytest

import oracle.symmetry as osym
import oracle.tensor as ot
from oracle import ref_engine
from oracle.dense import dense, flip
from emulate import run_program
import tentorch_amd as tt
from tentorch_amd._native import engine

CATS = {'u1': (tt.u1, osym.u1), 'no_symmetry': (tt.no_symmetry, osym.no_symmetry),
        'z3': (tt.z_n(3), osym.z_n(3)), 'fermion_parity': (tt.fermion_parity, osym.fermion_parity),
        'su2': (tt.su2, osym.su2), 'fibonacci': (tt.fibonacci, osym.fibonacci)}

# two legs per category; the first is traced, the second kept
LEGS = {
    'u1': [([(-2,), (0,), (2,)], [2, 3, 2]), ([(-1,), (1,)], [2, 2])],
    'no_symmetry': [([(0,)], [3]), ([(0,)], [2])],
    'z3': [([(0,), (1,), (2,)], [2, 1, 1]), ([(0,), (1,), (2,)], [1, 1, 1])],
    'fermion_parity': [([(0,), (1,)], [2, 2]), ([(0,), (1,)], [1, 1])],
    'su2': [([(0,), (1,), (2,)], [2, 1, 2]), ([(1,)], [1])],
    'fibonacci': [([(0,), (1,)], [2, 3]), ([(0,), (1,)], [1, 1])],
}

# (equation, codomain legs, domain legs) by index into LEGS[name]: a leg traced within one side and
# across the two sides, with the kept legs bent, two traces in one operand, a full trace, and a
# trace followed by a contraction
ONE_OPERAND = [
    ('abac->bc', (0, 1), (0, 1)),
    ('ABaC->BC', (0, 1), (0, 1)),
    ('aBCa->BC', (0, 1), (1, 0)),
    ('aAbB->', (0, 0), (1, 1)),
    ('abAB->', (0, 1), (0, 1)),
]


def _pair(name, cod, dom, rng, q=None):
    cat_p, cat_o = CATS[name]
    legs = LEGS[name]
    dtype = np.complex128 if name == 'fibonacci' else np.float64
    q = cat_o.trivial() if q is None else q
    o = ot.random_tensor(cat_o, [osym.Leg(cat_o, *legs[i]) for i in cod], [osym.Leg(cat_o, *legs[i]) for i in dom], q,
                         rng, dtype)
    s = engine.Structure(cat_p, [tt.Leg(cat_p, *legs[i]) for i in cod], [tt.Leg(cat_p, *legs[i]) for i in dom], q)
    return o, s


def _cases(name):
    '''(equation, oracle operands, engine structures) of the trace equations for one category.'''
    rng = np.random.default_rng(11)
    out = []
    for spec, cod, dom in ONE_OPERAND:
        o, s = _pair(name, cod, dom, rng)
        out.append((spec, [o], [s]))
    # a traced operand contracted with a second one; the abelian U(1) legs need a second operand
    # whose legs close (odd and even sectors do not)
    first = _pair(name, (0, 1), (0, 1), rng)
    second = _pair(name, (1,), (1,), rng)
    out.append(('abac,cd->bd', [first[0], second[0]], [first[1], second[1]]))
    return out


@pytest.mark.skipif(ref_engine.engine() is None, reason='the reference engine is not built (python oracle/build_ref.py)')
@pytest.mark.parametrize('name', sorted(LEGS))
def test_oracle_traces_agree_with_the_reference_engine(name):
    ext = ref_engine.engine()
    for spec, operands, _ in _cases(name):
        mine = ot.einsum(spec, *operands)
        theirs = ext.einsum(spec, *[ref_engine.to_reference(ext, t) for t in operands])
        blocks = {(tuple(k.codomain), tuple(k.domain), k.copy): b.numpy() for k, b in theirs.to_tuple()[3].items()}
        assert set(blocks) == set(mine.blocks), spec
        for key, b in mine.blocks.items():
            assert np.abs(blocks[key] - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), (spec, key)


@pytest.mark.parametrize('name', ['u1', 'no_symmetry', 'z3'])
def test_oracle_traces_of_bosonic_abelian_tensors_are_dense_diagonal_sums(name):
    '''For U(1), Z3 and no symmetry every F and R symbol is 1, so in the dense frame (every leg read
    as a codomain leg) a trace pairs index c of the upper leg with index dual(c) of the lower one:
    flipping the lower axis makes it np.einsum's diagonal sum.'''
    _, cat_o = CATS[name]
    rng = np.random.default_rng(3)
    o, _ = _pair(name, (0, 1), (0, 1), rng)
    a = o.legs()[0]
    d = flip(cat_o, dense(o), 2, a)
    got = ot.einsum('abac->bc', o)
    assert np.abs(dense(got) - np.einsum('ijik->jk', d)).max() <= 1e-13
    o2, _ = _pair(name, (0, 1), (1, 0), rng)
    d2 = flip(cat_o, dense(o2), 3, a)
    got2 = ot.einsum('aBCa->BC', o2)
    assert np.abs(dense(got2) - np.einsum('ijki->jk', d2)).max() <= 1e-13
    # a planted mistake: the lower axis read without its flip differs wherever a sector is not
    # its own dual
    if name != 'no_symmetry':
        assert np.abs(dense(got) - np.einsum('ijik->jk', dense(o))).max() > 1e-3


@pytest.mark.parametrize('name', sorted(LEGS))
def test_trace_programs_match_the_oracle(name):
    for spec, operands, structures in _cases(name):
        program = engine.plan_einsum(spec, structures, None)
        dtype = operands[0].dtype
        got = run_program(program, [o.flat() for o in operands], dtype)
        ref = ot.einsum(spec, *operands)
        assert program.out.size == ref.numel(), spec
        assert [k for k, _, _, _ in program.out.blocks()] == list(ref.blocks.keys()), spec
        assert np.abs(got - ref.flat()).max() <= 1e-12 * max(1.0, np.abs(ref.flat()).max()), spec


def test_a_trace_is_a_hoistable_move():
    '''The trace reads one operand and writes the workspace, so a rerun with that operand unchanged
    skips it, like the placements before it.'''
    rng = np.random.default_rng(0)
    first = _pair('u1', (0, 1), (0, 1), rng)
    second = _pair('u1', (1,), (1,), rng)
    program = engine.plan_einsum('abac,cd->bd', [first[1], second[1]], None)
    assert program.hoistable & 1


def test_trace_errors_follow_the_reference():
    rng = np.random.default_rng(0)
    _, s = _pair('u1', (0, 1), (1, 0), rng)
    with pytest.raises(ValueError, match='names two different spaces in the operand 0'):
        engine.plan_einsum('abac->bc', [s], None)
    with pytest.raises(ValueError, match='names more than two legs'):
        engine.plan_einsum('aaab->b', [s], None)
