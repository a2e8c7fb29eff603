'''
Pinning the oracle (the CPU restatement of the reference) against what fixes it independently:
exact diagonalisation and its closed form at L = 4, the dense restatement of abelian tensors
(place = transpose, contract = tensordot, einsum = numpy.einsum), and the categorical identities of
the F and R symbols (unitarity, pentagon, hexagon, the closed forms of SU(2) and Fibonacci).
'''

import itertools
import math

import numpy as np
import pytest

import oracle.symmetry as osym
import oracle.tensor as ot
from oracle.dense import dense, flip
from oracle.dmrg import DMRG, MPS, Model, exact_energy, heisenberg, lanczos


def test_exact_energy_known_answers():
    assert abs(exact_energy(4) - (-(3 + 2 * math.sqrt(3)) / 4)) < 1e-14
    assert abs(exact_energy(6) - (-2.493577133887925)) < 1e-12
    assert abs(exact_energy(8) - (-3.374932598687896)) < 1e-12


@pytest.mark.parametrize('cat', [osym.u1, osym.no_symmetry])
@pytest.mark.parametrize('length', [4, 6, 8])
def test_oracle_dmrg_reaches_exact_diagonalisation(cat, length):
    driver = DMRG(Model(length, cat=cat), max_bond_dim=64, n_iter=5)
    driver.run(num_sweeps=6)
    assert abs(driver.energy().real - exact_energy(length)) < 1e-10


def test_oracle_mpo_is_the_heisenberg_hamiltonian():
    '''The D = 5 automaton contracted to a dense 2^L matrix equals the Hamiltonian (L = 4).'''
    length = 4
    mpo = heisenberg(osym.no_symmetry, length)
    dense_sites = [dense(t) for t in mpo.tensors]   # (w_left, out, in, w_right)
    # contract by hand: chain of (w, out, in, w)
    full = dense_sites[0][0]  # (out, in, wr)
    for site in dense_sites[1:]:
        full = np.einsum('...b,bpjc->...pjc', full, site)
    full = full[..., 0]
    outs = full.transpose([0, 2, 4, 6, 1, 3, 5, 7]).reshape(16, 16)
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0, 1.0], [0, 0]])
    sm = sp.T
    ham = np.zeros((16, 16))
    for i in range(length - 1):
        for a, b, w in ((sz, sz, 1.0), (sp, sm, 0.5), (sm, sp, 0.5)):
            ops = [np.eye(2)] * length
            ops[i], ops[i + 1] = a, b
            term = ops[0]
            for o in ops[1:]:
                term = np.kron(term, o)
            ham += w * term
    assert np.abs(outs - ham).max() < 1e-14


def _random_u1(rng, cod, dom, q=(0,)):
    return ot.random_tensor(osym.u1, [osym.Leg(osym.u1, *l) for l in cod], [osym.Leg(osym.u1, *l) for l in dom],
                            q, rng)


BOND = ([(-2,), (0,), (2,)], [2, 3, 2])
PHYS = ([(-1,), (1,)], [1, 1])
MPOL = ([(-2,), (0,), (2,)], [1, 3, 1])


def test_place_is_a_transpose_of_the_dense_array():
    rng = np.random.default_rng(0)
    t = _random_u1(rng, [BOND, PHYS], [PHYS, BOND])
    d = dense(t)
    for n_above in range(5):
        for perm in itertools.permutations(range(4)):
            above, below = list(perm[:n_above]), list(perm[n_above:])
            assert np.array_equal(dense(t.place(above, below)), np.transpose(d, above + below))


def test_contract_is_a_tensordot_in_the_codomain_frame():
    rng = np.random.default_rng(1)
    a = _random_u1(rng, [BOND, PHYS], [PHYS, BOND])
    b = _random_u1(rng, [BOND, PHYS], [BOND])
    for la, lb in (([3], [0]), ([0], [0]), ([3, 2], [0, 1])):
        fb = dense(b)
        for axis in lb:
            fb = flip(osym.u1, fb, axis, b.legs()[axis])
        ref = np.tensordot(dense(a), fb, (la, lb))
        assert np.abs(dense(ot.contract(a, b, la, lb)) - ref).max() < 1e-13


@pytest.mark.parametrize('spec,shapes', [
    ('EPaqj,JRSFb,AQsb->EPrf', [([BOND, PHYS], [BOND, PHYS, MPOL]), ([MPOL, PHYS, PHYS, BOND], [BOND]),
                                ([BOND, PHYS], [PHYS, BOND])]),
    ('EIa,IPqj->EPaqj', [([BOND, MPOL], [BOND]), ([MPOL, PHYS], [PHYS, MPOL])]),
    ('EIa,Feq,aPb,IqPj->FJb', [([BOND, MPOL], [BOND]), ([BOND], [BOND, PHYS]), ([BOND, PHYS], [BOND]),
                               ([MPOL, PHYS], [PHYS, MPOL])]),
])
def test_einsum_is_numpy_einsum_in_the_codomain_frame(spec, shapes):
    rng = np.random.default_rng(2)
    ops = [_random_u1(rng, c, d) for c, d in shapes]
    result = ot.einsum(spec, *ops)
    lhs, rhs = spec.split('->')
    groups = lhs.split(',')
    dense_ops = []
    seen = {}
    for op, group in zip(ops, groups):
        arr = dense(op)
        for axis, letter in enumerate(group.lower()):
            if letter in seen:
                arr = flip(osym.u1, arr, axis, op.legs()[axis])
            seen[letter] = True
        dense_ops.append(arr)
    ref = np.einsum(','.join(g.lower() for g in groups) + '->' + rhs.lower(), *dense_ops)
    assert np.abs(dense(result) - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_abelian_closed_forms_equal_the_clebsch_gordan_overlaps():
    cat = osym.u1
    for a, b, c in itertools.product([(-1,), (0,), (2,)], repeat=3):
        d = (a[0] + b[0] + c[0],)
        assert np.allclose(cat.compute_F(a, b, c, d), osym.Group.compute_F(cat, a, b, c, d))
        assert np.allclose(cat.compute_R(b, c), osym.Group.compute_R(cat, b, c))


def _pentagon_holds(cat, labels):
    '''F^{e}_{f c d} F^{e}_{a b g} = Σ_h F^{d}_{a b c} F^{e}_{a h d} F^{g}_{b c d}, multiplicity-free form.'''
    def f(a, b, c, d, e, ff):
        rows = cat.f_rows(a, b, c, d)
        cols = cat.f_columns(a, b, c, d)
        m = cat.F_symbol(a, b, c, d)
        for i, r in enumerate(rows):
            for j, col in enumerate(cols):
                if r[0] == e and col[0] == ff:
                    return m[i, j]
        return 0.0
    worst = 0.0
    for a, b, c, d in itertools.product(labels, repeat=4):
        outs = set()
        for x, _ in cat.outcomes(a, b):
            for y, _ in cat.outcomes(x, c):
                for e, _ in cat.outcomes(y, d):
                    outs.add(e)
        for e in outs:
            for fx, _ in cat.outcomes(a, b):
                for gx, _ in cat.outcomes(c, d):
                    for p_, _ in cat.outcomes(fx, c):
                        for q_, _ in cat.outcomes(b, gx):
                            lhs = f(fx, c, d, e, p_, gx) * f(a, b, gx, e, fx, q_)
                            rhs = 0.0
                            for h, _ in cat.outcomes(b, c):
                                rhs += f(a, b, c, p_, fx, h) * f(a, h, d, e, p_, q_) * f(b, c, d, q_, h, gx)
                            worst = max(worst, abs(lhs - rhs))
    return worst


def test_su2_symbols_are_unitary_and_satisfy_the_pentagon():
    cat = osym.su2
    labels = [(0,), (1,), (2,), (3,)]
    for a, b, c in itertools.product(labels, repeat=3):
        for d, _ in cat.outcomes(*cat.outcomes(a, b)[0][0:1], c):
            m = cat.F_symbol(a, b, c, d)
            assert np.allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-12)
    for b, c in itertools.product(labels, repeat=2):
        r = cat.R_symbol(b, c)
        assert np.allclose(r @ r.conj().T, np.eye(r.shape[0]), atol=1e-12)
        # R^{e}_{bc} = (-1)^{j_b + j_c - j_e} for SU(2) in the Condon-Shortley gauge
        for i, (e, _) in enumerate(cat.r_rows(b, c)):
            assert abs(r[i, i] - (-1) ** ((b[0] + c[0] - e[0]) // 2)) < 1e-12
    assert _pentagon_holds(cat, [(0,), (1,), (2,)]) < 1e-12


def test_fibonacci_symbols_satisfy_pentagon_and_hexagon():
    cat = osym.fibonacci
    labels = [(0,), (1,)]
    assert _pentagon_holds(cat, labels) < 1e-12
    phi = (1 + math.sqrt(5)) / 2
    m = cat.F_symbol((1,), (1,), (1,), (1,))
    assert np.allclose(m, [[1 / phi, 1 / math.sqrt(phi)], [1 / math.sqrt(phi), -1 / phi]])
    # hexagon at a = b = c = d = tau: R^{tt}_e F_{eg} R^{tt}_g = sum_f F_{ef} R^{t f}_t F_{fg}, where
    # R^{t 1}_t = 1 and R^{t t}_t is the tau entry of R_{tau tau}
    rtt = cat.R_symbol((1,), (1,))
    r = np.diag([rtt[0, 0], rtt[1, 1]])
    middle = np.diag([cat.R_symbol((1,), (0,))[0, 0], rtt[1, 1]])
    assert np.allclose(r @ m @ r, m @ middle @ m, atol=1e-12)


def test_su2_place_preserves_the_norm_and_round_trips():
    rng = np.random.default_rng(4)
    cat = osym.su2
    spin = osym.Leg(cat, [(0,), (1,), (2,)], [2, 1, 2])
    half = osym.Leg(cat, [(1,)], [1])
    t = ot.random_tensor(cat, [spin, half], [half, spin], (0,), rng)
    base = ot.norm(t)
    for above, below in (([1, 0], [2, 3]), ([0, 2, 1], [3]), ([3], [0, 1, 2]), ([2, 0], [3, 1])):
        moved = t.place(above, below)
        assert abs(ot.norm(moved) - base) < 1e-12 * base
        inverse_above = [(above + below).index(leg) for leg in range(2)]
        inverse_below = [(above + below).index(leg) for leg in range(2, 4)]
        back = moved.place(inverse_above, inverse_below)
        assert ot.is_close(back, t, rtol=1e-12, atol=1e-12)


def test_fibonacci_place_preserves_the_norm():
    rng = np.random.default_rng(5)
    cat = osym.fibonacci
    leg = osym.Leg(cat, [(0,), (1,)], [2, 3])
    t = ot.random_tensor(cat, [leg, leg], [leg, leg], (0,), rng, dtype=np.complex128)
    base = ot.norm(t)
    for above, below in (([1, 0], [2, 3]), ([0, 2, 1], [3]), ([3], [0, 1, 2])):
        assert abs(ot.norm(t.place(above, below)) - base) < 1e-12 * base


def test_lanczos_matches_dense_projection():
    '''A 5-step Krylov space of a small symmetric matrix: the Ritz value equals the lowest eigenvalue of
    the matrix projected on the Krylov basis, computed densely.'''
    rng = np.random.default_rng(6)
    leg = osym.Leg(osym.no_symmetry, [(0,)], [12])
    v = ot.random_tensor(osym.no_symmetry, [leg], [osym.Leg(osym.no_symmetry, [(0,)], [1])], (0,), rng)
    m = rng.standard_normal((12, 12))
    m = m + m.T

    def action(x):
        out = x.clone()
        out.set_flat(m @ x.flat())
        return out

    value, vector = lanczos(action, v, n_iter=5)
    k = [v.flat() / np.linalg.norm(v.flat())]
    for _ in range(4):
        k.append(m @ k[-1])
    q, _ = np.linalg.qr(np.stack(k, axis=1))
    assert abs(value - np.linalg.eigvalsh(q.T @ m @ q)[0]) < 1e-10
    assert abs(np.linalg.norm(vector.flat()) - 1.0) < 1e-12


def test_dense_boxes_equal_the_entry_by_entry_expansion():
    from oracle.dense import dense_by_entry
    rng = np.random.default_rng(11)
    for cod, dom in (([BOND, PHYS], [PHYS, BOND]), ([BOND, MPOL], [BOND]), ([MPOL, PHYS, PHYS, BOND], [BOND]),
                     ([BOND], [BOND, PHYS, MPOL])):
        t = _random_u1(rng, cod, dom)
        assert np.array_equal(dense(t), dense_by_entry(t))


# -- the splittings (`svd.cpp:41-257`), pinned by the dense matrix each abelian tensor is -------------

def _matrix(t):
    '''The dense array of a tensor as (codomain legs) x (domain legs).'''
    d = dense(t)
    rows = int(np.prod([leg.multiplets() for leg in t.codomain]))
    return d.reshape(rows, -1)


def _tensor_with_spectra(rng, cat, codomain, domain, spectra):
    '''A tensor whose block with domain root s has singular values spectra[s] (padded with zeros).'''
    t = ot.Tensor(cat, codomain, domain, cat.trivial())
    for (r, s, c), block in t.blocks.items():
        m, n = block.shape
        k = min(m, n)
        values = np.zeros(k)
        given = np.asarray(spectra.get(s, []), dtype=np.float64)[:k]
        values[:given.size] = given
        u, _ = np.linalg.qr(rng.standard_normal((m, m)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        block[...] = (u[:, :k] * values) @ v[:, :k].T
    return t


@pytest.mark.parametrize('absorb', [0, 1, 2, 3])
def test_oracle_svd_is_the_truncated_svd_of_the_dense_matrix(absorb):
    '''No symmetry (one self-dual sector, so no re-pairing): the values are the dense matrix's
    singular values, the kept count is min(cap, #values > tol), and U V reconstructs the tensor up to
    exactly the dropped weight (Eckart-Young), whatever factor takes the values.'''
    rng = np.random.default_rng(12)
    cat = osym.no_symmetry
    a, b = osym.Leg(cat, [(0,)], [6]), osym.Leg(cat, [(0,)], [2])
    t = ot.random_tensor(cat, [a, b], [b, a], (0,), rng)
    dense_values = np.linalg.svd(_matrix(t), compute_uv=False)
    all_values = np.sort(ot.svdvals(t))[::-1]
    assert np.abs(all_values - dense_values).max() < 1e-12
    tol = dense_values[9] * 0.999            # the tenth value survives the tolerance, the eleventh not
    for cap, expected in ((None, 10), (4, 4), (40, 10)):
        u, s, v = ot.svd(t, None, cap, tol, absorb)
        assert s.size == expected == u.domain[0].multiplets()
        assert np.abs(np.sort(s)[::-1] - dense_values[:expected]).max() < 1e-12
        if absorb == 0:       # both factors isometric: the values sit between them
            rebuilt = _matrix(u) @ np.diag(s) @ _matrix(v)
        else:
            rebuilt = _matrix(ot.contract(u, v, [2], [0]))
        dropped = float(np.sum(dense_values[expected:] ** 2))
        assert abs(np.linalg.norm(_matrix(t) - rebuilt) ** 2 - dropped) < 1e-10
        if absorb in (0, 2):
            m = _matrix(u)
            assert np.abs(m.T @ m - np.eye(expected)).max() < 1e-12
        if absorb in (0, 1):
            m = _matrix(v)
            assert np.abs(m @ m.T - np.eye(expected)).max() < 1e-12


def test_oracle_svd_ranks_across_sectors_and_re_pairs_duals():
    '''U(1) blocks with hand-chosen values: the cap keeps the largest values over all sectors, then
    a sector and its dual keep the same count (the smaller), refilled in rank order
    (`svd.cpp:97-119`); exact zeros never survive, whatever the cap.'''
    rng = np.random.default_rng(13)
    cat = osym.u1
    bond = osym.Leg(cat, [(-2,), (0,), (2,)], [3, 3, 3])
    one = osym.Leg(cat, [(0,)], [1])
    spectra = {(-2,): [0.9, 0.5, 0.1], (0,): [0.8, 0.4, 0.05], (2,): [0.7, 0.45, 0.0]}
    t = _tensor_with_spectra(rng, cat, [bond, one], [bond], spectra)
    cases = {
        # cap: (multiplicities by sector, kept values by sector)
        4: ({(-2,): 1, (0,): 1, (2,): 1}, [0.9, 0.8, 0.7]),     # 0.5 of sector -2 loses its pair
        5: ({(-2,): 2, (0,): 1, (2,): 2}, [0.9, 0.5, 0.8, 0.7, 0.45]),   # 0.45 pairs 0.5
        6: ({(-2,): 2, (0,): 2, (2,): 2}, [0.9, 0.5, 0.8, 0.4, 0.7, 0.45]),
        9: ({(-2,): 2, (0,): 3, (2,): 2}, [0.9, 0.5, 0.8, 0.4, 0.05, 0.7, 0.45]),  # 0.1 unpaired, 0 below tol
    }
    for cap, (mults, values) in cases.items():
        u, s, v = ot.svd(t, None, cap, 1e-14, 2)
        bond_out = u.domain[0]
        assert dict(zip(bond_out.sectors, bond_out.multiplicities)) == mults, cap
        assert np.abs(s - values).max() < 1e-12, cap
    # rank deficiency: a block of zeros keeps nothing, and the bond still exists
    zero = _tensor_with_spectra(rng, cat, [bond, one], [bond], {})
    u, s, v = ot.svd(zero, None, None, 1e-14, 2)
    assert u.domain[0].sectors == [cat.trivial()] and u.domain[0].multiplets() == 1
    assert s.size == 1 and s[0] == 0.0


@pytest.mark.parametrize('split', ['qr', 'rq'])
def test_oracle_qr_and_rq_factor_every_block(split):
    '''qr: Q isometric (Q^T Q = 1) over (codomain | bond), R upper triangular per block, Q R the
    tensor; rq: the mirror, with the isometry on the right.  The bond carries min(rows, cols) per
    block of each domain root.'''
    rng = np.random.default_rng(14)
    cat = osym.u1
    bond = osym.Leg(cat, [(-1,), (1,)], [3, 2])
    phys = osym.Leg(cat, [(-1,), (1,)], [1, 1])
    out = osym.Leg(cat, [(-2,), (0,), (2,)], [2, 3, 2])
    t = ot.random_tensor(cat, [bond, phys], [out], (0,), rng)
    assert len(t.blocks) == 3
    left, right = (ot.qr if split == 'qr' else ot.rq)(t)
    assert np.abs(_matrix(ot.contract(left, right, [2], [0])) - _matrix(t)).max() < 1e-12
    iso = left if split == 'qr' else right
    m = _matrix(iso)
    if split == 'qr':
        assert np.abs(m.T @ m - np.eye(m.shape[1])).max() < 1e-12
        for block in right.blocks.values():
            assert np.abs(np.tril(block, -1)).max(initial=0.0) < 1e-15
    else:
        assert np.abs(m @ m.T - np.eye(m.shape[0])).max() < 1e-12
        for block in left.blocks.values():
            assert np.abs(np.triu(block, 1)).max(initial=0.0) < 1e-15
    expected = {}
    for (r, s, c), block in t.blocks.items():
        expected[s] = expected.get(s, 0) + min(block.shape)
    assert dict(zip(left.domain[0].sectors, left.domain[0].multiplicities)) == expected
