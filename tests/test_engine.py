'''
The host engine without a GPU: its symbols, trees, layouts and block keys against the oracle, the
leg-move and block-product tables it plans (interpreted on the CPU by tests/emulate.py) against the
oracle's per-entry algorithm, its error messages, and the C ABI library's exported symbols.
'''

import ctypes
import itertools

import numpy as np
import pytest

import cabi
import oracle.symmetry as osym
import oracle.tensor as ot
from emulate import run_program
import tentorch_amd as tt
from tentorch_amd._native import engine

CATS = {'u1': (tt.u1, osym.u1), 'no_symmetry': (tt.no_symmetry, osym.no_symmetry),
        'z3': (tt.z_n(3), osym.z_n(3)), 'fermion_parity': (tt.fermion_parity, osym.fermion_parity),
        'su2': (tt.su2, osym.su2), 'fibonacci': (tt.fibonacci, osym.fibonacci)}

LEGS = {
    'u1': [([(-2,), (0,), (2,)], [2, 3, 2]), ([(-1,), (1,)], [1, 1]), ([(-2,), (0,), (2,)], [1, 3, 1])],
    'no_symmetry': [([(0,)], [3]), ([(0,)], [2]), ([(0,)], [4])],
    'z3': [([(0,), (1,), (2,)], [2, 1, 1]), ([(0,), (1,), (2,)], [1, 1, 1]), ([(0,), (1,), (2,)], [1, 2, 1])],
    'fermion_parity': [([(0,), (1,)], [2, 2]), ([(0,), (1,)], [1, 1]), ([(0,), (1,)], [2, 1])],
    'su2': [([(0,), (1,), (2,)], [2, 1, 2]), ([(1,)], [1]), ([(0,), (2,)], [1, 1])],
    'fibonacci': [([(0,), (1,)], [2, 3]), ([(0,), (1,)], [1, 1]), ([(1,)], [2])],
}


def both(name, cod, dom, q=None, seed=0, dtype=np.float64):
    cat_p, cat_o = CATS[name]
    q = cat_o.trivial() if q is None else q
    rng = np.random.default_rng(seed)
    o = ot.random_tensor(cat_o, [osym.Leg(cat_o, *l) for l in cod], [osym.Leg(cat_o, *l) for l in dom], q, rng, dtype)
    s = engine.Structure(cat_p, [tt.Leg(cat_p, *l) for l in cod], [tt.Leg(cat_p, *l) for l in dom], q)
    return s, o


def test_c_abi_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(cabi.LIB_PATH)
    declared = cabi.declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.tt_version() == 1


@pytest.mark.parametrize('name', sorted(CATS))
def test_structures_and_layouts_match_the_oracle(name):
    cat_p, cat_o = CATS[name]
    legs = LEGS[name]
    for cod, dom in (([0, 1], [1, 0]), ([0, 1, 2], [0]), ([2], [1, 1, 0])):
        s, o = both(name, [legs[i] for i in cod], [legs[i] for i in dom])
        assert s.size == o.numel()
        keys = [k for k, _, _, _ in s.blocks()]
        assert keys == list(o.blocks.keys())
        for (k, rows, cols, off), ob in zip(s.blocks(), o.blocks.values()):
            assert (rows, cols) == ob.shape
        for root in osym.reachable_charges(cat_o, [osym.Leg(cat_o, *legs[i]) for i in cod]):
            trees_p = engine.fusion_trees(cat_p, [tt.Leg(cat_p, *legs[i]) for i in cod], root)
            trees_o = osym.fusion_trees(cat_o, [osym.Leg(cat_o, *legs[i]) for i in cod], root)
            assert [(tuple(a), tuple(b), tuple(c), d) for a, b, c, d in trees_p] == [t.key() for t in trees_o]


@pytest.mark.parametrize('name', sorted(CATS))
def test_legs_with_the_same_sectors_and_other_multiplicities_lay_out_alike(name):
    '''Legs that differ only in multiplicities share the planner's fusion-tree sets and reachable
    charges; each structure must still get its own sizes and offsets, and its moves the oracle's.'''
    cat_p, cat_o = CATS[name]
    dtype = np.complex128 if name == 'fibonacci' else np.float64
    exact = cat_o.abelian
    for bump in (0, 1, 3, 0, 2):
        legs = [(sectors, [m + bump for m in mults]) for sectors, mults in LEGS[name]]
        s, o = both(name, [legs[0], legs[1]], [legs[1], legs[0]], seed=bump, dtype=dtype)
        assert s.size == o.numel()
        assert [k for k, _, _, _ in s.blocks()] == list(o.blocks.keys())
        offset = 0
        for (k, rows, cols, off), ob in zip(s.blocks(), o.blocks.values()):
            assert (rows, cols) == ob.shape and off == offset
            offset += rows * cols
        for above, below in (([3, 1], [0, 2]), ([0, 1, 2], [3]), ([2], [3, 0, 1])):
            got = run_program(engine.plan_place(s, above, below), [o.flat()], dtype)
            ref = o.place(above, below).flat()
            if exact:
                assert np.array_equal(got, ref), (bump, above, below)
            else:
                assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (bump, above, below)


@pytest.mark.parametrize('name', ['su2', 'fibonacci', 'fermion_parity', 'u1'])
def test_symbols_match_the_oracle(name):
    cat_p, cat_o = CATS[name]
    labels = sorted({s for leg in LEGS[name] for s in leg[0]})
    for a, b, c in itertools.product(labels, repeat=3):
        for d in {x for y, _ in cat_o.outcomes(a, b) for x, _ in cat_o.outcomes(y, c)}:
            assert np.allclose(cat_p.F_symbol(a, b, c, d), cat_o.F_symbol(a, b, c, d), atol=1e-13)
        assert np.allclose(cat_p.R_symbol(b, c), cat_o.R_symbol(b, c), atol=1e-13)


@pytest.mark.parametrize('name', sorted(CATS))
def test_every_placement_matches_the_oracle(name):
    '''All placements of a 4-leg tensor: bit-exact for abelian categories, 1e-12 for recoupling.'''
    legs = LEGS[name]
    dtype = np.complex128 if name == 'fibonacci' else np.float64
    s, o = both(name, [legs[0], legs[1]], [legs[1], legs[0]], dtype=dtype)
    flat = o.flat()
    exact = CATS[name][1].abelian
    for n_above in range(5):
        for perm in itertools.permutations(range(4)):
            above, below = list(perm[:n_above]), list(perm[n_above:])
            program = engine.plan_place(s, above, below)
            got = run_program(program, [flat], dtype)
            ref = o.place(above, below).flat()
            if exact:
                assert np.array_equal(got, ref), (above, below)
            else:
                assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (above, below)


def test_the_per_step_tables_compose_to_the_same_move():
    '''The plain form's elementary tables, applied one after another, give the composed move.'''
    s, o = both('u1', [LEGS['u1'][0], LEGS['u1'][1]], [LEGS['u1'][1], LEGS['u1'][0]])
    program = engine.plan_place(s, [3, 1], [0, 2])
    assert program.op_kinds() == ['move']
    assert np.array_equal(run_program(program, [o.flat()]), o.place([3, 1], [0, 2]).flat())


BOND, PHYS, MPOL = LEGS['u1']


@pytest.mark.parametrize('spec,shapes', [
    ('EPaqj,JRSFb,AQsb->EPrf', [([BOND, PHYS], [BOND, PHYS, MPOL]), ([MPOL, PHYS, PHYS, BOND], [BOND]),
                                ([BOND, PHYS], [PHYS, BOND])]),
    ('EIa,IPqj->EPaqj', [([BOND, MPOL], [BOND]), ([MPOL, PHYS], [PHYS, MPOL])]),
    ('JRsk,Fkb->JRSFb', [([MPOL, PHYS], [PHYS, MPOL]), ([BOND, MPOL], [BOND])]),
    ('EIa,Feq,aPb,IqPj->FJb', [([BOND, MPOL], [BOND]), ([BOND], [BOND, PHYS]), ([BOND, PHYS], [BOND]),
                               ([MPOL, PHYS], [PHYS, MPOL])]),
    ('Gfq,aPb,IqPk,Gkb->FIa', [([BOND], [BOND, PHYS]), ([BOND, PHYS], [BOND]), ([MPOL, PHYS], [PHYS, MPOL]),
                               ([BOND, MPOL], [BOND])]),
    ('AIe,IPqj,JRsk,Dkf,EQSf->APrd', [([BOND, MPOL], [BOND]), ([MPOL, PHYS], [PHYS, MPOL]),
                                      ([MPOL, PHYS], [PHYS, MPOL]), ([BOND, MPOL], [BOND]),
                                      ([BOND, PHYS], [PHYS, BOND])]),
])
def test_dmrg_einsums_match_the_oracle(spec, shapes):
    rng_seed = 7
    pairs = [both('u1', c, d, seed=rng_seed + i) for i, (c, d) in enumerate(shapes)]
    program = engine.plan_einsum(spec, [p[0] for p in pairs], None)
    got = run_program(program, [p[1].flat() for p in pairs])
    ref = ot.einsum(spec, *[p[1] for p in pairs])
    assert program.out.size == ref.numel()
    assert [k for k, _, _, _ in program.out.blocks()] == list(ref.blocks.keys())
    assert np.abs(got - ref.flat()).max() <= 1e-12 * max(1.0, np.abs(ref.flat()).max())


def test_the_matvec_program_is_one_move_and_two_grouped_products_per_call():
    shapes = [([BOND, PHYS], [BOND, PHYS, MPOL]), ([MPOL, PHYS, PHYS, BOND], [BOND]), ([BOND, PHYS], [PHYS, BOND])]
    program = engine.plan_einsum('EPaqj,JRSFb,AQsb->EPrf', [both('u1', c, d)[0] for c, d in shapes], None)
    assert program.op_kinds() == ['move', 'gemm', 'move', 'move', 'gemm']
    # head (operand 0) and tail (operand 1) moves are hoistable; the state's is not
    assert program.hoistable == 0b011


@pytest.mark.parametrize('name', ['u1', 'su2'])
def test_contract_matches_the_oracle(name):
    legs = LEGS[name]
    a = both(name, [legs[0], legs[1]], [legs[0]], seed=1)
    b = both(name, [legs[0], legs[1]], [legs[0]], seed=2)
    program = engine.plan_contract(a[0], b[0], [2], [0], None)
    got = run_program(program, [a[1].flat(), b[1].flat()])
    ref = ot.contract(a[1], b[1], [2], [0]).flat()
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_einsum_errors_follow_the_reference():
    s, _ = both('u1', [BOND, PHYS], [BOND])
    with pytest.raises(ValueError, match='does not say what it leaves'):
        engine.plan_einsum('abc', [s], None)
    with pytest.raises(ValueError, match='names 2 operands but got 1'):
        engine.plan_einsum('abc,abc->', [s], None)
    with pytest.raises(ValueError, match='names one leg but is not kept'):
        engine.plan_einsum('abc->ab', [s], None)
    with pytest.raises(ValueError, match='is summed over and still kept'):
        engine.plan_einsum('abc,abd->abcd', [s, s], None)
    with pytest.raises(ValueError, match='names 2 legs but the tensor has 3'):
        engine.plan_einsum('ab->ab', [s], None)
    with pytest.raises(ValueError, match='cannot close to'):
        engine.plan_einsum('ABc->ABc', [s], (7,))


def test_place_errors_follow_the_reference():
    s, _ = both('u1', [BOND, PHYS], [BOND])
    with pytest.raises(ValueError, match='does not exist'):
        engine.plan_place(s, [0, 5], [1])
    with pytest.raises(ValueError, match='is named twice'):
        engine.plan_place(s, [0, 0], [1, 2])
    with pytest.raises(ValueError, match='is named on neither side'):
        engine.plan_place(s, [0], [1])


def test_contract_errors_follow_the_reference():
    a, _ = both('u1', [BOND, PHYS], [BOND])
    b, _ = both('u1', [PHYS], [PHYS], q=(2,))
    with pytest.raises(ValueError, match='must carry the trivial charge'):
        engine.plan_contract(a, b, [1], [0], None)
    with pytest.raises(ValueError, match='do not name one space'):
        engine.plan_contract(a, a, [0], [1], None)
    with pytest.raises(ValueError, match='summed twice'):
        engine.plan_contract(a, a, [0, 0], [0, 2], None)


def test_microbench_matvec_plan_matches_the_independent_counts():
    '''The block-einsum microbenchmark's bond (27 sectors, chi = 1024) and its matvec plan: 29 block
    products per contraction step, the largest (946, 198, 198), 0.898 GFLOP in all, theta of 286,650
    entries and head/tail of 1,400,610 -- figures counted independently of this engine from the
    sector arithmetic alone.'''
    from tentorch_amd.microbench import EQ3, microbench_bond
    from tentorch_amd.symmetry import Leg, u1
    sectors, mults = microbench_bond()
    assert len(sectors) == 27 and sum(mults) == 1024
    assert dict((s[0], m) for s, m in zip(sectors, mults))[0] == 102
    assert [dict((s[0], m) for s, m in zip(sectors, mults))[q] for q in (4, -4, 8, -8, 12, -12)] == [62, 62, 14, 14, 1, 1]
    b = Leg(u1, sectors, mults)
    M, P = Leg(u1, [(0,), (2,), (-2,)], (3, 1, 1)), Leg(u1, [(1,), (-1,)], (1, 1))

    def S(c, d):
        return engine.Structure(u1, c, d, (0,))

    W = S([M, P], [P, M])
    head = engine.plan_einsum('EIa,IPqj->EPaqj', [S([b, M], [b]), W], None).out
    tail = engine.plan_einsum('JRsk,Fkb->JRSFb', [W, S([b], [M, b])], None).out
    theta = S([b, P], [P, b])
    program = engine.plan_einsum(EQ3, [head, tail, theta], None)
    assert (theta.size, head.size, tail.size) == (286650, 1400610, 1400610)
    assert program.op_kinds().count('gemm') == 2
    shapes = []
    for i in range(2):
        groups, terms = program.gemm_tables(i)
        assert len(groups) == 29 and len(terms) == 29
        shapes += [(g[1], g[2], t[2]) for g in groups for t in terms[g[3]:g[3] + g[4]]]
    assert max(shapes, key=lambda s: s[0] * s[1] * s[2]) in ((946, 198, 198), (198, 198, 946))
    flops = sum(c[1] for c in program.op_costs() if c[0] == 'gemm')
    assert abs(flops - 0.898e9) < 0.001e9


@pytest.mark.parametrize('name', ['su2', 'fibonacci'])
def test_recoupling_einsum_program_matches_the_oracle(name):
    '''`ABCd,Be->ACde` takes the middle leg of a three-leg codomain, so its first move recouples two
    legs (multi-source tiles with F and R coefficients): the planned program, interpreted on the CPU,
    against the oracle.'''
    legs = {'su2': [([(0,), (2,), (4,)], [2, 2, 1]), ([(0,), (2,)], [1, 2]), ([(2,), (4,)], [2, 1]),
                    ([(0,), (2,)], [2, 1]), ([(0,), (2,), (4,)], [1, 1, 1])],
            'fibonacci': [([(0,), (1,)], [2, 3]), ([(0,), (1,)], [1, 2]), ([(0,), (1,)], [2, 1]),
                          ([(0,), (1,)], [1, 1]), ([(0,), (1,)], [2, 2])]}[name]
    dtype = np.complex128 if name == 'fibonacci' else np.float64
    cat_p, cat_o = CATS[name]
    rng = np.random.default_rng(7)
    operands, structures = [], []
    for cod, dom in (((0, 1, 2), (3,)), ((1,), (4,))):
        operands.append(ot.random_tensor(cat_o, [osym.Leg(cat_o, *legs[i]) for i in cod],
                                         [osym.Leg(cat_o, *legs[i]) for i in dom], cat_o.trivial(), rng, dtype))
        structures.append(engine.Structure(cat_p, [tt.Leg(cat_p, *legs[i]) for i in cod],
                                           [tt.Leg(cat_p, *legs[i]) for i in dom], cat_p.trivial()))
    program = engine.plan_einsum('ABCd,Be->ACde', structures, None)
    tasks, _ = program.move_tables(0)
    assert any(t[4] > 1 for t in tasks), 'the first move should recouple'
    got = run_program(program, [o.flat() for o in operands], dtype)
    ref = ot.einsum('ABCd,Be->ACde', *operands).flat()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
