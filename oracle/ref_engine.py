'''
The reference's own engine, built from its sources by `oracle/build_ref.py` into `oracle/_ref/`,
under a restatement of its two-site DMRG driver (`tentorch/algorithms/dmrg.py:96-161`, with the
Lanczos of `tentorch/linalg/lanczos.py:8-64` and the canonical moves of `tentorch/networks/mps.py`).

TEST INFRASTRUCTURE ONLY: the CPU baseline of `bench.py` times it, and `tests/` checks the numpy
oracle against it; the product never loads it.

Every contraction, split and scalar product here is the reference engine's (`einsum`, `contract`,
`svd`, `qr`, `rq`, `inner`, `norm`: ATen per block, on the CPU).  The operator and the starting
state are built by the oracle (the Heisenberg automaton, `MPS.random` from numpy's generator, so
all three engines start from the same state) and handed over block by block.  Floating point is
FP64 throughout.
'''

import importlib.util
import os
from typing import Optional

import numpy as np
import torch

from . import dmrg as odmrg
from . import symmetry as osym

LIBRARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref', 'symmetry_ext.so')
_module = None


def engine():
    '''The compiled reference engine, or None where it has not been built.  Its categories are made
    at import on the device `default_device` picks (`cat.cpp:13-20`): the GPU where torch sees one,
    unless TENTORCH_DEVICE is "cpu"; the import sets that, so the engine runs its CPU path on a GPU
    box too.'''
    global _module
    if _module is None and os.path.exists(LIBRARY):
        saved = os.environ.get('TENTORCH_DEVICE')
        os.environ['TENTORCH_DEVICE'] = 'cpu'
        try:
            spec = importlib.util.spec_from_file_location('symmetry_ext', LIBRARY)
            module = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(module)
        finally:
            if saved is None:
                del os.environ['TENTORCH_DEVICE']
            else:
                os.environ['TENTORCH_DEVICE'] = saved
        if str(module.u1.device) != 'cpu':
            raise RuntimeError(f'the reference engine made its categories on {module.u1.device}, not the CPU')
        _module = module
    return _module


def _category(ext, cat):
    '''The reference engine's category of an oracle category.'''
    if isinstance(cat, osym.U1Symmetry):
        return ext.u1
    if isinstance(cat, osym.NoSymmetry):
        return ext.no_symmetry
    if isinstance(cat, osym.SU2Symmetry):
        return ext.su2
    if isinstance(cat, osym.FibonacciCategory):
        return ext.fibonacci
    if isinstance(cat, osym.FermionParity):
        return ext.fermion_parity
    if isinstance(cat, osym.ZNSymmetry):
        return ext.z_n(cat.modulus)
    raise ValueError(f'no reference category for {cat.name}')


def _leg(ext, leg):
    return ext.Leg(_category(ext, leg.cat), list(leg.sectors), list(leg.multiplicities))


def to_reference(ext, t):
    '''An oracle tensor as a reference-engine tensor with the same blocks.'''
    blocks = {ext.BlockKey(r, s, c): torch.from_numpy(np.ascontiguousarray(b)) for (r, s, c), b in t.blocks.items()}
    return ext.Tensor.from_tuple(_category(ext, t.cat), [_leg(ext, l) for l in t.codomain],
                                 [_leg(ext, l) for l in t.domain], tuple(t.qtotal), blocks)


def _weight(values) -> float:
    '''The kept weight: the split hands back its kept singular values as one torch vector.'''
    return float(torch.sum(torch.abs(values) ** 2))


def lanczos(ext, action, start, n_iter: int = 5, tol: float = 1e-12):
    '''`lanczos` (`tentorch/linalg/lanczos.py:8-64`) over reference tensors.'''
    scale = ext.norm(start)
    if scale == 0.0:
        raise ValueError('Lanczos needs a starting vector of non-zero norm.')
    basis = [start * (1.0 / scale)]
    alphas, betas = [], []
    for step in range(n_iter):
        image = action(basis[-1])
        alphas.append(complex(ext.inner(basis[-1], image)).real)
        image -= basis[-1] * alphas[-1]
        if step > 0:
            image -= basis[-2] * betas[-1]
        for _ in range(2):
            for vector in basis:
                image -= vector * complex(ext.inner(vector, image)).real
        beta = float(ext.norm(image))
        if beta < tol:
            break
        betas.append(beta)
        basis.append(image * (1.0 / beta))
    size = len(alphas)
    matrix = np.diag(alphas) + np.diag(betas[:size - 1], 1) + np.diag(betas[:size - 1], -1)
    values, vectors = np.linalg.eigh(matrix)
    vector = basis[0] * float(vectors[0, 0])
    for step in range(1, size):
        vector += basis[step] * float(vectors[step, 0])
    return float(values[0]), vector * (1.0 / float(ext.norm(vector)))


class ReferenceDMRG:
    '''`DMRG` (`tentorch/algorithms/dmrg.py:13-212`) on the reference engine, from the oracle's
    operator and starting state (FP64).'''

    def __init__(self, length: int, max_bond_dim: int, svd_tol: float = 1e-14, n_iter: int = 5, seed: int = 0):
        ext = engine()
        if ext is None:
            raise FileNotFoundError(f'{LIBRARY} is not built (python oracle/build_ref.py)')
        self.ext = ext
        start = odmrg.DMRG(odmrg.Model(length), max_bond_dim, svd_tol, n_iter, seed)
        self.L, self.max_bond_dim, self.svd_tol, self.n_iter = length, max_bond_dim, svd_tol, n_iter
        self.operator = [to_reference(ext, t) for t in start.operator.tensors]
        self.tensors = [to_reference(ext, t) for t in start.state.tensors]
        self.center = start.state.center
        first, last = start.state.tensors[0], start.state.tensors[-1]
        self.left_edge = to_reference(ext, odmrg.identity(first.cat, first.codomain[0],
                                                           start.operator.tensors[0].codomain[0]))
        self.right_edge = to_reference(ext, odmrg.identity(last.cat, last.domain[0],
                                                            start.operator.tensors[-1].domain[1]))
        self.left, self.right = {}, {}

    # `mps.py` canonical moves
    def _left_canonicalize(self, at):
        ext = self.ext
        factor, rest = ext.qr(self.tensors[at], (0, 1))
        self.tensors[at] = factor
        self.tensors[at + 1] = ext.contract(rest, self.tensors[at + 1], ext.domain_legs(rest), [0]).place([0, 1])
        self.center = at + 1

    def _right_canonicalize(self, at):
        ext = self.ext
        factor, rest = ext.rq(self.tensors[at], (0,))
        self.tensors[at] = rest.place([0, 1])
        self.tensors[at - 1] = ext.contract(self.tensors[at - 1], factor, ext.domain_legs(self.tensors[at - 1]), [0])
        self.center = at - 1

    def move_center(self, at):
        while self.center < at:
            self._left_canonicalize(self.center)
        while self.center > at:
            self._right_canonicalize(self.center)

    def _advance_left(self, left, at):
        t = self.tensors[at]
        return self.ext.einsum('EIa,Feq,aPb,IqPj->FJb', left, t.conj(), t, self.operator[at])

    def _advance_right(self, right, at):
        t = self.tensors[at]
        return self.ext.einsum('Gfq,aPb,IqPk,Gkb->FIa', t.conj(), t, self.operator[at], right)

    def cache_environments(self):
        self.left = {0: self.left_edge}
        self.right = {self.L: self.right_edge}
        for at in range(self.L - 1, 1, -1):
            self.right[at] = self._advance_right(self.right[at + 1], at)

    def update(self, direction):
        ext, at = self.ext, self.center
        first, second = self.tensors[at], self.tensors[at + 1]
        two = ext.contract(first, second, ext.domain_legs(first), [0])
        head = ext.einsum('EIa,IPqj->EPaqj', self.left[at], self.operator[at])
        tail = ext.einsum('JRsk,Fkb->JRSFb', self.operator[at + 1], self.right[at + 2])
        energy, vector = lanczos(ext, lambda v: ext.einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, v), two, self.n_iter)
        left, values, right = ext.svd(vector, self.max_bond_dim, self.svd_tol, 2 if direction == 'right' else 1)
        self.tensors[at], self.tensors[at + 1] = left, right.place([0, 1])
        self.center = at + 1 if direction == 'right' else at
        self.tensors[self.center] /= float(ext.norm(self.tensors[self.center]))
        return energy, 1.0 - _weight(values)

    def sweep(self):
        self.move_center(0)
        self.cache_environments()
        energy, discarded = 0.0, 0.0
        for at in range(self.L - 1):
            energy, lost = self.update('right')
            discarded = max(discarded, lost)
            if at < self.L - 2:
                self.left[at + 1] = self._advance_left(self.left[at], at)
        self._right_canonicalize(self.L - 1)
        self.right[self.L - 1] = self._advance_right(self.right[self.L], self.L - 1)
        for at in range(self.L - 3, -1, -1):
            self.center = at
            energy, lost = self.update('left')
            discarded = max(discarded, lost)
            if at > 0:
                self.right[at + 1] = self._advance_right(self.right[at + 2], at + 1)
        return energy, discarded

    @property
    def bond_dims(self):
        return [self.tensors[at].domain[0].multiplets() for at in range(self.L - 1)]


def sweeps(length: int, bond: int, count: int, threads: Optional[int] = None, seed: int = 0):
    '''`count` sweeps of the reference engine's DMRG on `threads` ATen threads: (seconds per sweep,
    energies).  Building the operator and the canonical starting state is not timed.'''
    import time
    previous = torch.get_num_threads()
    if threads is not None:
        torch.set_num_threads(threads)
    try:
        driver = ReferenceDMRG(length, bond, seed=seed)
        times, energies = [], []
        for _ in range(count):
            start = time.perf_counter()
            energy, _ = driver.sweep()
            times.append(time.perf_counter() - start)
            energies.append(energy)
    finally:
        torch.set_num_threads(previous)
    return times, energies
