'''
The two-site DMRG and the matrix product state on the HIP path.

- The reference's own scenarios (`tests/test_dmrg.py:16-104`, `tests/test_mps.py:41-78`), with the
  energy against exact diagonalisation tightened to 1e-11.
- The truncating case the reference's CPU numbers are quoted on (L = 16, chi = 32, n_iter = 5,
  3 sweeps, seed 0) against the oracle's DMRG from the same state: per-sweep energies within 1e-10
  and, after every sweep, every bond's sectors and multiplicities equal.
- The environment, head, tail, window-energy and expectation einsums of a real DMRG state against
  the oracle's per-entry restatement, within 1e-12 relative.
'''

import math

import numpy as np
import pytest
import torch

import tentorch_amd as tt
from tentorch_amd.algorithms import GROW_LEFT, GROW_RIGHT, HEAD, MATVEC, TAIL, WINDOW_ENERGY
from tentorch_amd.networks import MPS, entered
from tentorch_amd.symmetry import is_no_symmetry
import oracle.dmrg as od
import oracle.symmetry as osym
import oracle.tensor as ot
from oracle.dmrg import exact_energy
from bridge import same_legs, to_device

pytestmark = pytest.mark.gpu

CATEGORIES = [tt.no_symmetry, tt.u1]
IDS = ['no_symmetry', 'u1']


# -- the reference's scenarios --------------------------------------------------------------------

@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
@pytest.mark.parametrize('length', [4, 6, 8])
def test_the_energy_matches_exact_diagonalisation(cat, length):
    driver = tt.DMRG(tt.Model(length, cat=cat), max_bond_dim=64, n_iter=5)
    driver.run(num_sweeps=6, verbose=False)
    assert abs(driver.energy().real - exact_energy(length)) < 1e-11


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_the_driver_writes_the_state_and_the_energy_onto_the_model(cat):
    model = tt.Model(6, cat=cat)
    result = tt.dmrg(model, max_bond_dim=64, num_sweeps=4, verbose=False)
    assert result is model
    assert result.state is not None
    assert abs(result.energy - exact_energy(6)) < 1e-10


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_a_sweep_leaves_a_normalised_state(cat):
    driver = tt.DMRG(tt.Model(6, cat=cat), max_bond_dim=64, n_iter=5)
    driver.sweep()
    assert abs(driver.state.norm() - 1.0) < 1e-12
    assert driver.center == driver.state.center


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_the_measured_energy_is_the_expectation_of_the_state(cat):
    driver = tt.DMRG(tt.Model(6, cat=cat), max_bond_dim=64, n_iter=5)
    driver.run(num_sweeps=6, verbose=False)
    assert abs(driver.energy() - driver.state.expectation(driver.operator)) < 1e-10


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_a_bigger_bond_reaches_a_lower_energy(cat):
    small = tt.DMRG(tt.Model(8, cat=cat), max_bond_dim=2, n_iter=5)
    small.run(num_sweeps=6, verbose=False)
    large = tt.DMRG(tt.Model(8, cat=cat), max_bond_dim=64, n_iter=5)
    large.run(num_sweeps=6, verbose=False)
    assert small.energy().real > large.energy().real - 1e-10
    assert abs(large.energy().real - exact_energy(8)) < 1e-10


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_the_bond_stays_within_the_limit(cat):
    driver = tt.DMRG(tt.Model(12, cat=cat), max_bond_dim=4, n_iter=5)
    driver.run(num_sweeps=4, verbose=False)
    assert max(driver.state.bond_dims) <= 4


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_a_capped_bond_throws_weight_away(cat):
    capped = tt.DMRG(tt.Model(10, cat=cat), max_bond_dim=2, n_iter=5)
    _, capped_lost = capped.sweep()
    full = tt.DMRG(tt.Model(10, cat=cat), max_bond_dim=64, n_iter=5)
    full.sweep()
    _, full_lost = full.sweep()
    assert capped_lost > 1e-3
    assert full_lost < 1e-12


def test_the_energy_needs_a_sweep():
    driver = tt.DMRG(tt.Model(4, cat=tt.u1), max_bond_dim=16)
    with pytest.raises(RuntimeError):
        driver.energy()


def test_a_run_needs_a_sweep():
    with pytest.raises(ValueError):
        tt.DMRG(tt.Model(4, cat=tt.u1), max_bond_dim=16).run(num_sweeps=0)


def test_the_energy_density_of_a_long_chain_approaches_the_known_limit():
    '''The reference keeps this case behind an environment switch because it is slow on the CPU
    (`test_dmrg.py:97-104`); on the GPU it takes seconds.'''
    length, bond = 64, 64
    energy = tt.dmrg(tt.Model(length, cat=tt.u1), max_bond_dim=bond, num_sweeps=5, verbose=False).energy
    assert abs(energy / length - (0.25 - math.log(2))) < 5e-3


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_the_norm_does_not_depend_on_the_center(cat):
    operator = tt.heisenberg(cat, 4)
    state = MPS.random(cat, operator.physical, 4, seed=1, multiplets=2)
    state.normalize()
    for center in list(range(state.L)) + list(range(state.L - 1, -1, -1)):
        state.move_center(center)
        assert state.center == center
        assert abs(state.norm() - 1.0) < 1e-12


@pytest.mark.parametrize('cat', CATEGORIES, ids=IDS)
def test_the_expectation_does_not_depend_on_the_center(cat):
    operator = tt.heisenberg(cat, 4)
    state = MPS.random(cat, operator.physical, 4, seed=2, multiplets=2)
    state.normalize()
    values = []
    for center in range(state.L):
        state.move_center(center)
        values.append(complex(state.expectation(operator)))
    assert max(abs(value - values[0]) for value in values) < 1e-12


def _product_state(cat, physical, spins):
    '''One spin per site (`test_mps.py:13-38`): bonds of one multiplet, the charge summed so far.'''
    if is_no_symmetry(cat):
        bonds = [tt.Leg(cat, [(0,)], (1,)) for _ in range(len(spins) + 1)]
    else:
        bonds = [tt.Leg(cat, [(sum(spins[:at]),)], (1,)) for at in range(len(spins) + 1)]

    def fill(at):
        def entry(tree_out, copy_out, tree_in, copy_in):
            if is_no_symmetry(cat):
                return 1.0 if 1 - 2 * copy_out[1] == spins[at] else 0.0
            return 1.0 if tuple(tree_out.sectors[1]) == (spins[at],) else 0.0
        return entry

    return MPS([entered(cat, [bonds[at], physical], [bonds[at + 1]], fill(at)) for at in range(len(spins))],
               physical)


def test_a_product_state_has_the_energy_of_its_spins():
    cat = tt.no_symmetry
    operator = tt.heisenberg(cat, 4)
    spins = [1, -1, 1, -1]
    state = _product_state(cat, operator.physical, spins)
    assert abs(state.norm() - 1.0) < 1e-12
    pairs = sum(0.25 if a == b else -0.25 for a, b in zip(spins, spins[1:]))
    assert abs(complex(state.expectation(operator)) - pairs) < 1e-12
    assert abs(pairs - (-0.75)) < 1e-15


# -- the reference's CPU case, truncating, against the oracle -------------------------------------

SWEEPS = 3


@pytest.fixture(scope='module')
def oracle_run():
    '''The oracle's DMRG of the L = 16, chi = 32 chain: energy, dropped weight and every bond leg
    after each sweep.'''
    driver = od.DMRG(od.Model(16, cat=osym.u1), max_bond_dim=32, svd_tol=1e-14, n_iter=5, seed=0)
    out = []
    for _ in range(SWEEPS):
        energy, lost = driver.sweep()
        out.append((energy, lost, [t.domain[0] for t in driver.state.tensors[:-1]]))
    return out


def test_truncating_dmrg_matches_the_oracle_sweep_by_sweep(oracle_run):
    '''Config 1 of the benchmark set: the HIP path's sweeps from the same random state keep the same
    multiplets on every bond (sectors and multiplicities bit-equal) and reach the same energies.'''
    driver = tt.DMRG(tt.Model(16, cat=tt.u1), max_bond_dim=32, svd_tol=1e-14, n_iter=5, seed=0)
    for sweep, (ref_energy, ref_lost, ref_bonds) in enumerate(oracle_run):
        energy, lost = driver.sweep()
        bonds = [t.domain[0] for t in driver.state.tensors[:-1]]
        for cut, (mine, ref) in enumerate(zip(bonds, ref_bonds)):
            assert same_legs(mine, ref), f'sweep {sweep + 1}, bond {cut + 1}: {mine} against {ref}'
        assert abs(energy - ref_energy) < 1e-10, f'sweep {sweep + 1}'
        # the dropped weight is a difference of two numbers near 1: compared to its rounding
        assert abs(lost - ref_lost) < 1e-12, f'sweep {sweep + 1}'
    assert max(driver.state.bond_dims) == 32


def test_truncating_dmrg_in_the_plain_form_matches_the_oracle(oracle_run):
    '''The per-block form with the reference's split (one SVD per block) over the same case.'''
    from tentorch_amd.linalg import SvdSettings
    from tentorch_amd.tensor import Settings
    driver = tt.DMRG(tt.Model(16, cat=tt.u1), max_bond_dim=32, svd_tol=1e-14, n_iter=5, seed=0)
    with Settings.using(mode=1):
        SvdSettings.method = 'exact'
        try:
            energy, _ = driver.sweep()
        finally:
            SvdSettings.method = 'gram'
    ref_energy, _, ref_bonds = oracle_run[0]
    assert abs(energy - ref_energy) < 1e-10
    for mine, ref in zip([t.domain[0] for t in driver.state.tensors[:-1]], ref_bonds):
        assert same_legs(mine, ref)


# -- the sweep's einsums on a real state ----------------------------------------------------------

@pytest.fixture(scope='module')
def oracle_chain():
    '''An oracle DMRG state of the L = 8 chain after one sweep, with every left and right block
    recomputed from it, and the same tensors on the device.'''
    driver = od.DMRG(od.Model(8, cat=osym.u1), max_bond_dim=16, svd_tol=1e-14, n_iter=5, seed=3)
    driver.sweep()
    driver.cache_environments()
    for at in range(driver.L - 2):
        driver.left[at + 1] = driver._advance_left(driver.left[at], at)
    return driver


def _compare(got: tt.Tensor, ref: ot.Tensor):
    keys = [(tuple(k.codomain), tuple(k.domain), k.copy) for k in got.keys()]
    assert keys == [(tuple(r), tuple(s), c) for r, s, c in ref.keys()]
    g, r = got.data.cpu().numpy(), ref.flat()
    assert g.shape == r.shape
    assert np.abs(g - r).max() <= 1e-12 * max(1.0, np.abs(r).max())


@pytest.mark.parametrize('at', [1, 3, 5])
def test_environment_einsums_match_the_oracle(oracle_chain, at):
    d = oracle_chain
    site, op = d.state.tensors[at], d.operator.tensors[at]
    ket, op_dev = to_device(site), to_device(op)
    left = ot.einsum(GROW_LEFT, d.left[at], site.conj(), site, op)
    _compare(tt.einsum(GROW_LEFT, to_device(d.left[at]), ket.conj(), ket, op_dev), left)
    right = ot.einsum(GROW_RIGHT, site.conj(), site, op, d.right[at + 1])
    _compare(tt.einsum(GROW_RIGHT, ket.conj(), ket, op_dev, to_device(d.right[at + 1])), right)


@pytest.mark.parametrize('at', [0, 2, 5])
def test_window_einsums_match_the_oracle(oracle_chain, at):
    d = oracle_chain
    left, right = d.left[at], d.right[at + 2]
    w1, w2 = d.operator.tensors[at], d.operator.tensors[at + 1]
    two = ot.contract(d.state.tensors[at], d.state.tensors[at + 1], ot.domain_legs(d.state.tensors[at]), [0])
    head, tail = ot.einsum(HEAD, left, w1), ot.einsum(TAIL, w2, right)
    head_d, tail_d = tt.einsum(HEAD, to_device(left), to_device(w1)), tt.einsum(TAIL, to_device(w2), to_device(right))
    _compare(head_d, head)
    _compare(tail_d, tail)
    two_d = to_device(two)
    _compare(tt.einsum(MATVEC, head_d, tail_d, two_d), ot.einsum(MATVEC, head, tail, two))
    energy = ot.einsum(WINDOW_ENERGY, left, w1, w2, right, two)
    _compare(tt.einsum(WINDOW_ENERGY, to_device(left), to_device(w1), to_device(w2), to_device(right), two_d), energy)


def test_expectation_transfer_matches_the_oracle(oracle_chain):
    d = oracle_chain
    ref = d.state.expectation(d.operator)
    sites = [to_device(t) for t in d.state.tensors]
    state = MPS.__new__(MPS)
    state.tensors, state.physical, state.center = sites, None, d.state.center
    got = state.expectation(tt.heisenberg(tt.u1, d.L))
    assert abs(complex(got) - complex(ref)) <= 1e-12 * max(1.0, abs(ref))
