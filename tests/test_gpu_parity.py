'''
The HIP path end to end against the oracle: leg moves bit-exact, einsum/contract within 1e-12
relative (FP64 block products sum in another order than numpy), the plain per-block form agreeing
with the batched one, the device Lanczos and split against their host and exact counterparts.  The
DMRG scenarios of the reference's tests are in test_gpu_dmrg.py.
'''

import itertools

import numpy as np
import pytest
import torch

import tentorch_amd as tt
from tentorch_amd.tensor import Settings
import oracle.symmetry as osym
import oracle.tensor as ot
from oracle.dmrg import exact_energy

pytestmark = pytest.mark.gpu


def _pair(rng, cod, dom, q=(0,)):
    '''An oracle tensor and the same tensor on the GPU.'''
    o = ot.random_tensor(osym.u1, [osym.Leg(osym.u1, *l) for l in cod], [osym.Leg(osym.u1, *l) for l in dom], q, rng)
    d = tt.Tensor(tt.u1, [tt.Leg(tt.u1, *l) for l in cod], [tt.Leg(tt.u1, *l) for l in dom], q)
    d.data.copy_(torch.from_numpy(o.flat()))
    return o, d


BOND = ([(-2,), (0,), (2,)], [2, 3, 2])
PHYS = ([(-1,), (1,)], [1, 1])
MPO = ([(-2,), (0,), (2,)], [1, 3, 1])


def test_place_is_bit_exact():
    rng = np.random.default_rng(1)
    o, d = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    for n_above in range(5):
        for perm in itertools.permutations(range(4)):
            above, below = list(perm[:n_above]), list(perm[n_above:])
            assert np.array_equal(d.place(above, below).data.cpu().numpy(), o.place(above, below).flat())


@pytest.mark.parametrize('mode', [0, 1])
def test_matvec_einsum_matches_the_oracle(mode):
    rng = np.random.default_rng(2)
    ho, hd = _pair(rng, [BOND, PHYS], [BOND, PHYS, MPO])
    to, td = _pair(rng, [MPO, PHYS, PHYS, BOND], [BOND])
    vo, vd = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    ref = ot.einsum('EPaqj,JRSFb,AQsb->EPrf', ho, to, vo).flat()
    Settings.mode = mode
    try:
        for _ in range(2):  # the second call reuses the hoisted moves of head and tail
            got = tt.einsum('EPaqj,JRSFb,AQsb->EPrf', hd, td, vd).data.cpu().numpy()
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    finally:
        Settings.mode = 0


def test_contract_and_conj_match_the_oracle():
    rng = np.random.default_rng(3)
    ao, ad = _pair(rng, [BOND, PHYS], [BOND])
    bo, bd = _pair(rng, [BOND, PHYS], [BOND])
    ref = ot.contract(ao, bo, [2], [0]).flat()
    got = tt.contract(ad, bd, [2], [0]).data.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(ad.conj().data.cpu().numpy(), ao.conj().flat())
    assert abs(tt.inner(ad, bd).real - ot.inner(ao, bo).real) <= 1e-12 * ot.norm(ao) * ot.norm(bo)


def test_plain_and_batched_forms_agree_on_a_sweep():
    energies = []
    for mode in (0, 1):
        Settings.mode = mode
        try:
            driver = tt.DMRG(tt.Model(8, cat=tt.u1), max_bond_dim=16, n_iter=5)
            energies.append(driver.sweep()[0])
        finally:
            Settings.mode = 0
    assert abs(energies[0] - energies[1]) < 1e-10


def test_device_lanczos_matches_the_host_loop():
    '''The device-resident Lanczos and the reference's host loop give the same Ritz value and vector
    on one DMRG window (classical vs modified Gram-Schmidt: equal to rounding).'''
    from tentorch_amd.linalg import LanczosSettings
    results = []
    for mode in ('device', 'host'):
        LanczosSettings.mode = mode
        try:
            driver = tt.DMRG(tt.Model(8, cat=tt.u1), max_bond_dim=16, n_iter=5)
            driver.state.move_center(0)
            driver.cache_environments()
            at = driver.center
            first, second = driver.state.tensors[at], driver.state.tensors[at + 1]
            two = tt.contract(first, second, tt.domain_legs(first), [0])
            head = tt.einsum('EIa,IPqj->EPaqj', driver.left[at], driver.operator.tensors[at])
            tail = tt.einsum('JRsk,Fkb->JRSFb', driver.operator.tensors[at + 1], driver.right[at + 2])
            results.append(tt.lanczos(lambda v: tt.einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, v), two, 5))
        finally:
            LanczosSettings.mode = 'device'
    (e1, v1), (e2, v2) = results
    assert abs(e1 - e2) < 1e-12
    assert abs(abs(tt.inner(v1, v2)) - 1.0) < 1e-10


def test_deferred_lanczos_matches_lanczos():
    '''The window's Lanczos with the tridiagonal solved on the device (no host synchronisation) gives
    the Ritz value and vector of `lanczos` on one DMRG window, the vector's sign fixed by a
    non-negative overlap with the start; a zero start is refused when the value is read.'''
    from tentorch_amd.linalg import lanczos_deferred
    driver = tt.DMRG(tt.Model(8, cat=tt.u1), max_bond_dim=16, n_iter=5)
    driver.state.move_center(0)
    driver.cache_environments()
    at = driver.center
    first, second = driver.state.tensors[at], driver.state.tensors[at + 1]
    two = tt.contract(first, second, tt.domain_legs(first), [0])
    head = tt.einsum('EIa,IPqj->EPaqj', driver.left[at], driver.operator.tensors[at])
    tail = tt.einsum('JRsk,Fkb->JRSFb', driver.operator.tensors[at + 1], driver.right[at + 2])
    action = lambda v: tt.einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, v)
    e1, v1 = tt.lanczos(action, two, 5)
    value, v2 = lanczos_deferred(action, two, 5)
    assert abs(value.get() - e1) < 1e-12
    assert abs(abs(tt.inner(v1, v2)) - 1.0) < 1e-10
    assert tt.inner(two, v2).real >= 0.0
    zero = tt.Tensor._wrap(two.structure, torch.zeros_like(two.data))
    value, _ = lanczos_deferred(action, zero, 5)
    with pytest.raises(ValueError, match='non-zero norm'):
        value.get()


def test_fp32_form_with_defect_correction_tracks_fp64():
    '''The FP32 form (FP32 block products in Lanczos, FP64 two-step correction) stays within 1e-10 of
    the FP64 sweep energy once the window problem is converged, and reaches ED at L = 8.'''
    energies = {}
    for precision in ('fp64', 'fp32'):
        driver = tt.DMRG(tt.Model(8, cat=tt.u1), max_bond_dim=64, n_iter=5, precision=precision)
        driver.run(num_sweeps=6, verbose=False)
        energies[precision] = driver.energy().real
    assert abs(energies['fp32'] - exact_energy(8)) < 1e-10
    assert abs(energies['fp32'] - energies['fp64']) < 1e-10


def test_captured_matvec_replays_match_eager_runs():
    '''The grouped matvec captured into a hipGraph, replayed three times with new Krylov vectors
    written into its input between replays (and eager work and synchronisations in between): every
    replay equals an eager einsum of the same operands bit for bit.'''
    from tentorch_amd import _native
    from tentorch_amd.tensor import run
    from tentorch_amd._native import engine
    rng = np.random.default_rng(5)
    _, head = _pair(rng, [BOND, PHYS], [BOND, PHYS, MPO])
    _, tail = _pair(rng, [MPO, PHYS, PHYS, BOND], [BOND])
    _, vec = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    eq = 'EPaqj,JRSFb,AQsb->EPrf'
    program = engine.plan_einsum(eq, [head.structure, tail.structure, vec.structure], None)
    out = torch.empty(program.out.size, dtype=torch.float64, device='cuda')
    torch.cuda.synchronize()
    with torch.cuda.stream(torch.cuda.Stream()):
        run(program, (head, tail, vec), out=out)
        # a fresh wrap of the vector keeps its layout move in the graph (moves of unchanged
        # operands are hoisted out of repeated runs)
        graph = _native.capture(lambda: run(program, (head, tail, tt.Tensor._wrap(vec.structure, vec.data)), out=out))
        for replay in range(3):
            vec.data.copy_(torch.from_numpy(rng.standard_normal(vec.data.numel())))
            graph.launch()
            torch.cuda.synchronize()
            got = out.clone()
            eager = tt.einsum(eq, head, tail, tt.Tensor._wrap(vec.structure, vec.data)).data
            torch.cuda.synchronize()
            assert torch.equal(got, eager), f'replay {replay}'


RECOUPLING = {
    'su2': (tt.su2, osym.su2, [([(0,), (2,), (4,)], [2, 2, 1]), ([(0,), (2,)], [1, 2]), ([(2,), (4,)], [2, 1]),
                              ([(0,), (2,)], [2, 1]), ([(0,), (2,), (4,)], [1, 1, 1])], np.float64),
    'fibonacci': (tt.fibonacci, osym.fibonacci, [([(0,), (1,)], [2, 3]), ([(0,), (1,)], [1, 2]), ([(0,), (1,)], [2, 1]),
                                                 ([(0,), (1,)], [1, 1]), ([(0,), (1,)], [2, 2])], np.complex128),
}


@pytest.mark.parametrize('name', ['su2', 'fibonacci'])
def test_recoupling_einsum_matches_the_oracle(name):
    '''An einsum that takes the middle leg of a three-leg codomain (F and R recoupling of two legs
    before the block products), SU(2) in float64 and Fibonacci in complex128: within 1e-12 of the
    oracle's per-entry restatement.'''
    cat_p, cat_o, legs, dtype = RECOUPLING[name]
    rng = np.random.default_rng(7)
    ops_o, ops_d = [], []
    for cod, dom in (((0, 1, 2), (3,)), ((1,), (4,))):
        o = ot.random_tensor(cat_o, [osym.Leg(cat_o, *legs[i]) for i in cod], [osym.Leg(cat_o, *legs[i]) for i in dom],
                             cat_o.trivial(), rng, dtype)
        d = tt.Tensor(cat_p, [tt.Leg(cat_p, *legs[i]) for i in cod], [tt.Leg(cat_p, *legs[i]) for i in dom],
                      cat_p.trivial(), dtype=torch.complex128 if dtype == np.complex128 else torch.float64)
        d.data.copy_(torch.from_numpy(o.flat()))
        ops_o.append(o)
        ops_d.append(d)
    ref = ot.einsum('ABCd,Be->ACde', *ops_o).flat()
    got = tt.einsum('ABCd,Be->ACde', *ops_d).data.cpu().numpy()
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_complex_recoupling_on_a_real_tensor_is_refused():
    '''Fibonacci moves carry complex coefficients: a float64 tensor cannot take them, and the engine
    says so instead of dropping the imaginary parts.'''
    _, _, legs, _ = RECOUPLING['fibonacci']
    rng = np.random.default_rng(8)
    x = tt.Tensor.random(tt.fibonacci, [tt.Leg(tt.fibonacci, *legs[i]) for i in (0, 1, 2)],
                         [tt.Leg(tt.fibonacci, *legs[3])], tt.fibonacci.trivial(), generator=rng)
    y = tt.Tensor.random(tt.fibonacci, [tt.Leg(tt.fibonacci, *legs[1])], [tt.Leg(tt.fibonacci, *legs[4])],
                         tt.fibonacci.trivial(), generator=rng)
    with pytest.raises(ValueError, match='complex'):
        tt.einsum('ABCd,Be->ACde', x, y)


@pytest.mark.parametrize('absorb', [2, 1])
def test_split_by_the_device_eigensolver_matches_the_exact_svd(absorb):
    '''The DMRG split (absorb 2: U isometric, values into V; absorb 1: V isometric, values into U)
    through the batched device eigensolver and through one exact SVD per block: the same singular
    values, the same reconstruction, the isometric factor orthonormal, and the kept values the
    split hands back on the host equal to the device ones.'''
    from tentorch_amd.linalg import SvdSettings, _split
    rng = np.random.default_rng(31)
    bond = tt.Leg(tt.u1, [(-2,), (0,), (2,)], [9, 14, 9])
    phys = tt.Leg(tt.u1, [(-1,), (1,)], [1, 1])
    two = tt.Tensor.random(tt.u1, [bond, phys], [phys, bond], (0,), generator=rng)
    results = {}
    for method, eigh in (('gram', 'device'), ('gram', 'torch'), ('exact', 'device')):
        SvdSettings.method, SvdSettings.eigh = method, eigh
        try:
            results[(method, eigh)] = _split(two, None, 20, 1e-14, absorb)
        finally:
            SvdSettings.method, SvdSettings.eigh = 'gram', 'device'
    ref_u, ref_s, ref_v, _ = results[('exact', 'device')]
    ref_rec = tt.contract(ref_u, ref_v, tt.domain_legs(ref_u), [0]).data
    for key in (('gram', 'device'), ('gram', 'torch'), ('exact', 'device')):
        u, s, v, kept = results[key]
        assert np.array_equal(kept, s.cpu().numpy()), key
        assert torch.allclose(s.sort().values, ref_s.sort().values, rtol=1e-10, atol=1e-13), key
        rec = tt.contract(u, v, tt.domain_legs(u), [0]).data
        assert (rec - ref_rec).abs().max() <= 1e-12 * ref_rec.abs().max(), key
        # the isometric factor: orthonormal columns of U's blocks (absorb 2) or rows of V's (absorb 1)
        for block in (u if absorb == 2 else v).blocks().values():
            m = block if absorb == 2 else block.transpose(0, 1)
            eye = torch.eye(m.shape[1], dtype=m.dtype, device=m.device)
            assert (m.transpose(0, 1) @ m - eye).abs().max() <= 1e-12, key


def test_exact_split_reorders_unsorted_singular_values(monkeypatch):
    '''torch.linalg.svd on ROCm can return a block's smallest singular values slightly out of order;
    the split re-sorts every block with its singular vectors before ranking (the ranking keeps a
    leading run of each block).  A per-block SVD whose values come back shuffled gives the same
    values and the same reconstruction as the sorted one.'''
    from tentorch_amd.linalg import SvdSettings
    rng = np.random.default_rng(37)
    bond = tt.Leg(tt.u1, [(-2,), (0,), (2,)], [9, 14, 9])
    phys = tt.Leg(tt.u1, [(-1,), (1,)], [1, 1])
    two = tt.Tensor.random(tt.u1, [bond, phys], [phys, bond], (0,), generator=rng)
    SvdSettings.method = 'exact'
    try:
        ref_u, ref_s, ref_v = tt.svd(two, 20, 1e-14, 2)
        original = torch.linalg.svd
        shuffle = torch.Generator().manual_seed(5)

        def shuffled(block, full_matrices=False):
            u, s, vh = original(block, full_matrices=full_matrices)
            order = torch.randperm(s.numel(), generator=shuffle).to(s.device)
            return u[:, order], s[order], vh[order]
        monkeypatch.setattr(torch.linalg, 'svd', shuffled)
        u, s, v = tt.svd(two, 20, 1e-14, 2)
    finally:
        SvdSettings.method = 'gram'
    assert torch.equal(s, ref_s)
    rec = tt.contract(u, v, tt.domain_legs(u), [0]).data
    ref_rec = tt.contract(ref_u, ref_v, tt.domain_legs(ref_u), [0]).data
    assert (rec - ref_rec).abs().max() <= 1e-12 * ref_rec.abs().max()


def _deep_spectrum_tensor(seed):
    '''A U(1) two-site tensor (bond, phys | phys, bond) whose blocks hold hand-set singular values:
    one geometric sequence 10^-((k + 1/2) / 4), k = 0, 1, ..., from about 1 down past 1e-16, dealt
    round-robin to the blocks (so the ranking runs across sectors), each block padded with exact
    zeros up to its rank, between random orthogonal factors.  Consecutive values differ by a factor
    1.78, and none lies within a factor 1.3 of the tolerance 1e-14.'''
    rng = np.random.default_rng(seed)
    cat = osym.u1
    bond = osym.Leg(cat, [(-2,), (0,), (2,)], [9, 14, 9])
    phys = osym.Leg(cat, [(-1,), (1,)], [1, 1])
    t = ot.Tensor(cat, [bond, phys], [phys, bond], (0,))
    keys = list(t.blocks)
    sequence = [10.0 ** (-(k + 0.5) / 4) for k in range(68)]
    spectra = {key: sequence[i::len(keys)] for i, key in enumerate(keys)}
    for key, block in t.blocks.items():
        m, n = block.shape
        k = min(m, n)
        values = np.zeros(k)
        given = np.asarray(spectra[key])[:k]
        values[:given.size] = given
        u, _ = np.linalg.qr(rng.standard_normal((m, m)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        block[...] = (u[:, :k] * values) @ v[:, :k].T
    return t


@pytest.mark.parametrize('absorb', [2, 1])
@pytest.mark.parametrize('cap', [None, 30, 45])
def test_split_keeps_the_oracles_multiplets_on_a_deep_spectrum(absorb, cap):
    '''The DMRG split (Gram eigendecompositions on the device, projected row norms, decoupled tail
    passes) against the oracle's per-block SVD on singular values spanning sixteen decades with exact
    zeros: the same kept multiplets per sector (bond legs bit-equal) with the cap high in the
    spectrum, in its deep tail (about 1e-11) or absent (the tolerance then cuts, and the exact zeros
    must go), the same kept values to 1e-14 of the largest, and the same truncated tensor.'''
    from tentorch_amd.linalg import _split
    from bridge import same_legs, to_device
    ref_u, ref_s, ref_v = ot.svd(_deep_spectrum_tensor(17), None, cap, 1e-14, absorb)
    two = to_device(_deep_spectrum_tensor(17))
    u, s, v, kept = _split(two, None, cap, 1e-14, absorb)
    assert same_legs(u.domain[0], ref_u.domain[0]), (u.domain[0], ref_u.domain[0])
    assert np.abs(kept - ref_s).max() <= 1e-14 * np.abs(ref_s).max()
    assert np.all(kept > 1e-14)
    got = tt.contract(u, v, tt.domain_legs(u), [0]).data.cpu().numpy()
    ref = ot.contract(ref_u, ref_v, [2], [0]).flat()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the isometric factor stays orthonormal
    for block in (u if absorb == 2 else v).blocks().values():
        m = block if absorb == 2 else block.transpose(0, 1)
        eye = torch.eye(m.shape[1], dtype=m.dtype, device=m.device)
        assert (m.transpose(0, 1) @ m - eye).abs().max() <= 1e-12


def test_editing_a_block_between_two_identical_contractions_is_seen():
    '''A program run again on an operand whose buffer was written through a block view, a to_tuple
    block or `.data` (torch's in-place counter moves, the library's own counter does not) must
    redo that operand's layout move instead of reusing the hoisted copy.'''
    rng = np.random.default_rng(61)
    _, a = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    _, b = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    spec = 'ABcd,DCef->ABef'
    first = tt.einsum(spec, a, b).data.clone()
    key = a.keys()[0]
    a.block(key.codomain, key.domain, key.copy).mul_(-3.0)       # through a block view
    second = tt.einsum(spec, a, b).data.clone()
    a.to_tuple()[3][key].zero_()                                   # through a to_tuple block
    third = tt.einsum(spec, a, b).data.clone()
    a.data[:5] = 7.0                                               # through .data
    fourth = tt.einsum(spec, a, b).data.clone()
    for before, after in ((first, second), (second, third), (third, fourth)):
        assert not torch.equal(before, after)
    fresh = tt.einsum(spec, tt.Tensor._wrap(a.structure, a.data.clone()), b).data
    assert torch.equal(fresh, fourth)


def test_svd_takes_both_call_forms_of_the_reference():
    '''`svd(tensor, max_bond, tol, absorb)` (the form `dmrg.py:124` uses) and
    `svd(tensor, left, max_bond, tol, absorb)`, positionally and by keyword.'''
    rng = np.random.default_rng(62)
    _, two = _pair(rng, [BOND, PHYS], [PHYS, BOND])
    u1, s1, v1 = tt.svd(two, 5, 1e-14, 2)
    u2, s2, v2 = tt.svd(two, [0, 1], 5, 1e-14, 2)
    u3, s3, v3 = tt.svd(two, max_bond=5, tol=1e-14, absorb=2)
    u4, s4, v4 = tt.svd(two, left=[0, 1], max_bond=5, absorb=2, tol=1e-14)
    for s in (s2, s3, s4):
        assert torch.equal(s, s1)
    assert u1.domain[0].multiplets() <= 5      # the cap, then a sector and its dual keep one count
    full = tt.svd(two)[1]
    assert full.numel() == tt.svdvals(two).numel() == tt.svdvals(two, [0, 1]).numel()
    with pytest.raises(TypeError):
        tt.svd(two, None, 5, 1e-14, 2)            # neither form: the reference's binding refuses it too


@pytest.mark.parametrize('split', ['qr', 'rq'])
@pytest.mark.parametrize('bond,out', [
    (([(-1,), (1,)], [3, 2]), ([(-2,), (0,), (2,)], [2, 3, 2])),
    (([(-2,), (0,), (2,)], [20, 35, 20]), ([(-3,), (-1,), (1,), (3,)], [30, 41, 29, 7])),
])
def test_qr_and_rq_match_the_oracle(split, bond, out):
    '''The batched Householder QR against the oracle's per-block LAPACK QR (`svd.cpp:225-249`): the
    same reflector convention gives the same Q and R, signs included, on tall, wide and square
    blocks; the bond carries min(rows, cols) multiplets per block.'''
    rng = np.random.default_rng(21)
    o, d = _pair(rng, [bond, PHYS], [out], (0,))
    ref_left, ref_right = (ot.qr if split == 'qr' else ot.rq)(o)
    left, right = (tt.qr if split == 'qr' else tt.rq)(d)
    assert list(left.domain[0].sectors) == [tuple(s) for s in ref_left.domain[0].sectors]
    assert list(left.domain[0].multiplicities) == list(ref_left.domain[0].multiplicities)
    for mine, ref in ((left, ref_left), (right, ref_right)):
        got, want = mine.data.cpu().numpy(), ref.flat()
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_qr_of_a_zero_block_gives_unit_columns_and_a_zero_r():
    '''A column with nothing to reflect keeps its reflector the identity (LAPACK's dlarfg): a zero
    block factors as unit columns times zero, as numpy's QR gives.'''
    rng = np.random.default_rng(22)
    o, d = _pair(rng, [BOND, PHYS], [BOND])
    d.data.zero_()
    o = ot.Tensor(o.cat, o.codomain, o.domain, o.qtotal, o.dtype)
    ref_left, ref_right = ot.qr(o)
    left, right = tt.qr(d)
    assert np.array_equal(left.data.cpu().numpy(), ref_left.flat())
    assert np.array_equal(right.data.cpu().numpy(), ref_right.flat())
