'''
Oracle: the Lanczos solver, the matrix product state and operator, the Heisenberg automaton and the
two-site DMRG driver of the reference, restated over the oracle's numpy tensors.

TEST INFRASTRUCTURE ONLY: imported by `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline`
leg of `bench.py`, never by the product.

Restates `tentorch/linalg/lanczos.py`, `tentorch/tensors/tree.py`, `tentorch/networks/mps.py`,
`tentorch/networks/mpo.py`, `tentorch/models/heisenberg.py`, `tentorch/models/model.py` and
`tentorch/algorithms/dmrg.py`.  One deliberate difference: the random starting state is drawn from a
host numpy generator (`numpy.random.default_rng(seed)`, N(0, 1) block by block in key order) rather
than torch's generator on the category's device, so that the product and the oracle start from the
same state.
'''

import itertools
import math
from typing import Callable, Dict, Optional, Sequence

import numpy as np

from .symmetry import Category, Leg, NoSymmetry, AbelianGroup, block_keys, fusion_trees, reachable_charges
from .tensor import (Tensor, contract, domain_legs, einsum, from_blocks, inner, norm, qr, random_tensor,
                     rq, svd)


def lanczos(action: Callable[[Tensor], Tensor], start: Tensor, n_iter: int = 5, tol: float = 1e-12):
    '''`lanczos` (`tentorch/linalg/lanczos.py:8-64`).'''
    if n_iter < 1:
        raise ValueError(f'Lanczos needs at least one step, not {n_iter}.')
    scale = norm(start)
    if scale == 0.0:
        raise ValueError('Lanczos needs a starting vector of non-zero norm.')
    basis = [start * (1.0 / scale)]
    alphas, betas = [], []
    for step in range(n_iter):
        image = action(basis[-1])
        alphas.append(inner(basis[-1], image).real)
        image -= basis[-1] * alphas[-1]
        if step > 0:
            image -= basis[-2] * betas[-1]
        for _ in range(2):
            for vector in basis:
                image -= vector * inner(vector, image)
        beta = norm(image)
        if beta < tol:
            break
        betas.append(beta)
        basis.append(image * (1.0 / beta))
    size = len(alphas)
    matrix = np.zeros((size, size))
    for step in range(size):
        matrix[step, step] = alphas[step]
    for step, beta in enumerate(betas[:size - 1]):
        matrix[step, step + 1] = beta
        matrix[step + 1, step] = beta
    values, vectors = np.linalg.eigh(matrix)
    coefficients = vectors[:, 0]
    vector = basis[0] * float(coefficients[0])
    for step in range(1, size):
        vector += basis[step] * float(coefficients[step])
    return float(values[0]), vector * (1.0 / norm(vector))


def rows(cat: Category, legs: Sequence[Leg], root):
    '''`rows` (`tentorch/tensors/tree.py:35-47`).'''
    for tree in fusion_trees(cat, legs, root):
        mults = [leg.mult_of(s) for leg, s in zip(legs, tree.sectors)]
        for copy in itertools.product(*[range(m) for m in mults]):
            yield tree, copy


def entered(cat, codomain, domain, entry, root=None, dtype=np.float64) -> Tensor:
    '''`entered` (`tentorch/tensors/tree.py:50-75`).'''
    charge = cat.trivial() if root is None else root
    blocks = {}
    for key in block_keys(cat, list(codomain), list(domain), charge):
        out_rows = list(rows(cat, codomain, key[0]))
        in_rows = list(rows(cat, domain, key[1]))
        data = np.zeros((len(out_rows), len(in_rows)), dtype=dtype)
        for r, (t_out, c_out) in enumerate(out_rows):
            for c, (t_in, c_in) in enumerate(in_rows):
                data[r, c] = entry(t_out, c_out, t_in, c_in)
        blocks[key] = data
    return from_blocks(cat, list(codomain), list(domain), charge, blocks, dtype)


class MPO:
    '''`MPO` (`tentorch/networks/mpo.py:7-86`).'''

    def __init__(self, sites, physical: Leg, cat: Category, sector=None):
        self.tensors = list(sites)
        if not self.tensors:
            raise ValueError('an operator needs at least one site')
        for at, site in enumerate(self.tensors):
            if site.nlegs() != 4:
                raise ValueError(f'site {at} has {site.nlegs()} legs, not the four of an operator')
        self.L = len(self.tensors)
        self.cat = cat
        self.physical = physical
        self.sector = cat.trivial() if sector is None else sector


# `heisenberg.py:15-37`: the automaton states and the operator each transition writes.
START, CLOSED = 0, 4
_STATE_OF_CHARGED = {(0,): {0: START, 1: 1, 2: CLOSED}, (2,): {0: 2}, (-2,): {0: 3}}
_OPERATOR = {
    (START, START): lambda out, into: 1.0 if out == into else 0.0,
    (CLOSED, CLOSED): lambda out, into: 1.0 if out == into else 0.0,
    (START, 1): lambda out, into: 0.5 if out == into == 1 else (-0.5 if out == into else 0.0),
    (1, CLOSED): lambda out, into: 0.5 if out == into == 1 else (-0.5 if out == into else 0.0),
    (START, 2): lambda out, into: 1.0 if out == into + 2 else 0.0,
    (2, CLOSED): lambda out, into: 0.5 if out == into - 2 else 0.0,
    (START, 3): lambda out, into: 1.0 if out == into - 2 else 0.0,
    (3, CLOSED): lambda out, into: 0.5 if out == into + 2 else 0.0,
}


def heisenberg(cat: Category, length: int, coupling: float = 1.0, dtype=np.float64) -> MPO:
    '''`heisenberg` (`tentorch/models/heisenberg.py:58-115`).'''
    if not isinstance(cat, AbelianGroup):
        raise ValueError(f'the rows of {cat.name} would need the recoupling to dress them')
    if isinstance(cat, NoSymmetry):
        phys, bond = Leg(cat, [(0,)], (2,)), Leg(cat, [(0,)], (5,))
        state_of = lambda sector, copy: copy
        spin_of = lambda sector, copy: 1 - 2 * copy
    else:
        phys = Leg(cat, [(1,), (-1,)], (1, 1))
        bond = Leg(cat, [(0,), (2,), (-2,)], (3, 1, 1))
        state_of = lambda sector, copy: _STATE_OF_CHARGED[sector][copy]
        spin_of = lambda sector, copy: sector[0]
    edge = Leg(cat, [cat.trivial()], (1,))

    def site(w_left, w_right, first, last):
        def entry(t_out, c_out, t_in, c_in):
            before = START if first else state_of(t_out.sectors[0], c_out[0])
            after = CLOSED if last else state_of(t_in.sectors[1], c_in[1])
            write = _OPERATOR.get((before, after))
            if write is None:
                return 0.0
            weight = coupling if before not in (START, CLOSED) and after == CLOSED else 1.0
            return weight * write(spin_of(t_out.sectors[1], c_out[1]), spin_of(t_in.sectors[0], c_in[0]))
        return entered(cat, [w_left, phys], [phys, w_right], entry, dtype=dtype)

    sites = [site(edge if at == 0 else bond, edge if at == length - 1 else bond, at == 0, at == length - 1)
             for at in range(length)]
    return MPO(sites, phys, cat)


def identity(cat, bond: Leg, middle: Optional[Leg] = None, dtype=np.float64) -> Tensor:
    '''`_identity` (`tentorch/networks/mps.py:12-28`).'''
    def entry(t_out, c_out, t_in, c_in):
        same = t_out.sectors[0] == t_in.sectors[0] and c_out[0] == c_in[0]
        return 1.0 if same and (middle is None or t_out.sectors[1] == cat.trivial()) else 0.0
    codomain = [bond] if middle is None else [bond, middle]
    return entered(cat, codomain, [bond], entry, dtype=dtype)


class MPS:
    '''`MPS` (`tentorch/networks/mps.py:31-204`).'''

    def __init__(self, tensors, physical: Leg, center: int = 0):
        if not tensors:
            raise ValueError('A state needs at least one site.')
        for at, t in enumerate(tensors):
            if t.nlegs() != 3:
                raise ValueError(f'site {at} needs three legs, not {t.nlegs()}')
        if not 0 <= center < len(tensors):
            raise IndexError(f'the centre {center} is outside the chain of {len(tensors)} sites')
        self.tensors = list(tensors)
        self.physical = physical
        self.center = center
        self.canonicalize()

    @property
    def L(self):
        return len(self.tensors)

    @property
    def cat(self):
        return self.tensors[0].cat

    @property
    def bond_dims(self):
        return [self.tensors[at].domain[0].multiplets() for at in range(self.L - 1)]

    @classmethod
    def random(cls, cat, physical, length, sector=None, seed=0, multiplets=1, dtype=np.float64):
        charge = cat.trivial() if sector is None else sector
        left = [reachable_charges(cat, [physical] * at) for at in range(length + 1)]
        tail = [reachable_charges(cat, [physical] * (length - at)) for at in range(length + 1)]
        bonds = []
        for at in range(length + 1):
            sectors = [b for b in left[at] if any(cat.can_fuse_to(b, r, charge) for r in tail[at])]
            bonds.append(Leg(cat, sectors, [multiplets] * len(sectors)))
        bonds[0] = bonds[-1] = Leg(cat, [charge], (1,))
        rng = np.random.default_rng(seed)
        tensors = [random_tensor(cat, [bonds[at], physical], [bonds[at + 1]], charge, rng, dtype)
                   for at in range(length)]
        return cls(tensors, physical)

    def _left_canonicalize(self, at):
        if at < 0 or at >= self.L - 1:
            raise IndexError(f'index {at} has no site to its right in a chain of {self.L} sites')
        factor, rest = qr(self.tensors[at], [0, 1])
        self.tensors[at] = factor
        joined = contract(rest, self.tensors[at + 1], domain_legs(rest), [0])
        self.tensors[at + 1] = joined.place([0, 1])
        self.center = at + 1

    def _right_canonicalize(self, at):
        if at <= 0 or at >= self.L:
            raise IndexError(f'index {at} has no site to its left in a chain of {self.L} sites')
        factor, rest = rq(self.tensors[at], [0])
        self.tensors[at] = rest.place([0, 1])
        self.tensors[at - 1] = contract(self.tensors[at - 1], factor, domain_legs(self.tensors[at - 1]), [0])
        self.center = at - 1

    def move_center(self, at):
        if not 0 <= at < self.L:
            raise IndexError(f'the centre {at} is outside the chain of {self.L} sites')
        while self.center < at:
            self._left_canonicalize(self.center)
        while self.center > at:
            self._right_canonicalize(self.center)

    def canonicalize(self):
        for at in range(self.center):
            self._left_canonicalize(at)
        for at in range(self.L - 1, self.center, -1):
            self._right_canonicalize(at)

    def norm(self):
        return math.sqrt(sum(float(np.sum(np.abs(b) ** 2)) for b in self.tensors[self.center].blocks.values()))

    def normalize(self):
        self.tensors[self.center] /= self.norm()

    def expectation(self, operator: MPO) -> complex:
        left = identity(self.cat, self.tensors[0].codomain[0], operator.tensors[0].codomain[0],
                        self.tensors[0].dtype)
        for at, t in enumerate(self.tensors):
            left = einsum('AIe,aPb,Feq,IPqj->BJf', left, t, t.conj(), operator.tensors[at])
        return inner(identity(self.cat, self.tensors[-1].domain[0], operator.tensors[-1].domain[1],
                              self.tensors[0].dtype), left)


class Model:
    '''`Model` (`tentorch/models/model.py:7-34`).'''

    def __init__(self, length, coupling=1.0, cat=None, physical=None, sector=None):
        from .symmetry import u1
        if length < 1:
            raise ValueError(f'a chain needs at least one site, not {length}')
        self.length = length
        self.coupling = coupling
        self.cat = u1 if cat is None else cat
        self.physical = physical
        self.sector = self.cat.trivial() if sector is None else sector
        self.state = None
        self.energy = None


class DMRG:
    '''`DMRG` (`tentorch/algorithms/dmrg.py:13-212`).'''

    def __init__(self, model: Model, max_bond_dim=50, svd_tol=1e-14, n_iter=5, seed=0, dtype=np.float64):
        self.model = model
        self.max_bond_dim = max_bond_dim
        self.svd_tol = svd_tol
        self.n_iter = n_iter
        self.operator = heisenberg(model.cat, model.length, model.coupling, dtype)
        physical = model.physical if model.physical is not None else self.operator.physical
        self.state = MPS.random(model.cat, physical, model.length, model.sector, seed=seed, dtype=dtype)
        self.left: Dict[int, Tensor] = {}
        self.right: Dict[int, Tensor] = {}

    @property
    def L(self):
        return self.model.length

    @property
    def center(self):
        return self.state.center

    @center.setter
    def center(self, at):
        self.state.center = at

    def _advance_left(self, left, at):
        t = self.state.tensors[at]
        return einsum('EIa,Feq,aPb,IqPj->FJb', left, t.conj(), t, self.operator.tensors[at])

    def _advance_right(self, right, at):
        t = self.state.tensors[at]
        return einsum('Gfq,aPb,IqPk,Gkb->FIa', t.conj(), t, self.operator.tensors[at], right)

    def cache_environments(self):
        dtype = self.state.tensors[0].dtype
        self.left = {0: identity(self.model.cat, self.state.tensors[0].codomain[0],
                                 self.operator.tensors[0].codomain[0], dtype)}
        self.right = {self.L: identity(self.model.cat, self.state.tensors[-1].domain[0],
                                       self.operator.tensors[-1].domain[1], dtype)}
        for at in range(self.L - 1, 1, -1):
            self.right[at] = self._advance_right(self.right[at + 1], at)

    def update(self, direction):
        at = self.center
        first, second = self.state.tensors[at], self.state.tensors[at + 1]
        two = contract(first, second, domain_legs(first), [0])
        site, following = self.operator.tensors[at], self.operator.tensors[at + 1]
        head = einsum('EIa,IPqj->EPaqj', self.left[at], site)
        tail = einsum('JRsk,Fkb->JRSFb', following, self.right[at + 2])

        def action(vector):
            return einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, vector)

        energy, vector = lanczos(action, two, self.n_iter)
        factor = 2 if direction == 'right' else 1
        left, values, right = svd(vector, None, self.max_bond_dim, self.svd_tol, factor)
        self.state.tensors[at], self.state.tensors[at + 1] = left, right.place([0, 1])
        self.center = at + 1 if direction == 'right' else at
        self.state.normalize()
        return energy, 1.0 - float(np.sum(np.abs(values) ** 2))

    def sweep(self):
        self.state.move_center(0)
        self.cache_environments()
        energy, discarded = 0.0, 0.0
        for at in range(self.L - 1):
            energy, lost = self.update('right')
            discarded = max(discarded, lost)
            if at < self.L - 2:
                self.left[at + 1] = self._advance_left(self.left[at], at)
        self.state._right_canonicalize(self.L - 1)
        self.right[self.L - 1] = self._advance_right(self.right[self.L], self.L - 1)
        for at in range(self.L - 3, -1, -1):
            self.center = at
            energy, lost = self.update('left')
            discarded = max(discarded, lost)
            if at > 0:
                self.right[at + 1] = self._advance_right(self.right[at + 2], at + 1)
        return energy, discarded

    def energy(self) -> complex:
        if not self.left:
            raise RuntimeError('the energy needs the environments of a sweep, and none has run')
        at = min(self.center, self.L - 2)
        two = contract(self.state.tensors[at], self.state.tensors[at + 1], domain_legs(self.state.tensors[at]), [0])
        image = einsum('AIe,IPqj,JRsk,Dkf,EQSf->APrd', self.left[at], self.operator.tensors[at],
                       self.operator.tensors[at + 1], self.right[at + 2], two)
        return complex(inner(two, image) / inner(two, two))

    def run(self, num_sweeps=5, energy_tol=1e-10, patience=2, verbose=False):
        if num_sweeps < 1:
            raise ValueError(f'a run needs at least one sweep, not {num_sweeps}')
        energy, best, stale = float('inf'), float('inf'), 0
        for sweep in range(1, num_sweeps + 1):
            energy, discarded = self.sweep()
            if verbose:
                print(f'sweep {sweep}: energy = {energy:.12f}, truncated = {discarded:.2e}, '
                      f'bond = {self.state.bond_dims}')
            if abs(energy - best) <= energy_tol:
                stale += 1
                if stale >= patience:
                    break
            else:
                stale, best = 0, energy
        self.model.state = self.state
        self.model.energy = energy
        return self.model


def exact_energy(length: int, coupling: float = 1.0) -> float:
    '''`exact_energy` (`tests/exact.py:15-54`): dense ED of the open spin-1/2 chain in Sz = 0.'''
    if length % 2:
        raise ValueError(f'a chain of {length} sites has no sector of zero magnetisation')
    states = [bits for bits in itertools.product((0, 1), repeat=length) if sum(bits) == length // 2]
    index = {bits: at for at, bits in enumerate(states)}
    matrix = np.zeros((len(states), len(states)))
    for at, bits in enumerate(states):
        for site in range(length - 1):
            if bits[site] == bits[site + 1]:
                matrix[at, at] += 0.25 * coupling
            else:
                matrix[at, at] -= 0.25 * coupling
                flipped = list(bits)
                flipped[site], flipped[site + 1] = flipped[site + 1], flipped[site]
                matrix[at, index[tuple(flipped)]] += 0.5 * coupling
    return float(np.linalg.eigvalsh(matrix)[0])
