'''
The two-site DMRG driver (`tentorch/algorithms/dmrg.py`) over the backend's tensors.

`DMRG.update` keeps the reference's window: the two-site tensor, the `head`/`tail` halves of the
effective Hamiltonian, Lanczos with the 3-operand einsum as the action, and the truncating SVD.  The
action is the same einsum call; the backend plans it once per window, hoists the moves of `head`
and `tail` out of the Krylov loop, and runs each matvec as one composed move and two grouped GEMMs.
'''

from typing import Dict, Literal

import numpy as np
import torch

from .linalg import _split, lanczos_deferred
from .models import Model, heisenberg
from .networks import MPS, identity
from .tensor import Settings, Tensor, contract, domain_legs, einsum, inner


class DMRG:
    '''`DMRG` (`tentorch/algorithms/dmrg.py:13-212`).'''

    def __init__(self, model: Model, max_bond_dim: int = 50, svd_tol: float = 1e-14, n_iter: int = 5,
                 seed: int = 0, dtype=None, device=None, precision: str = 'fp64', refine: int = 2) -> None:
        if precision not in ('fp64', 'fp32'):
            raise ValueError(f"precision has to be 'fp64' or 'fp32', not {precision!r}")
        self.precision = precision
        self.refine = refine
        self.model = model
        self.max_bond_dim = max_bond_dim
        self.svd_tol = svd_tol
        self.n_iter = n_iter
        self.operator = heisenberg(model.cat, model.length, model.coupling, dtype, device)
        physical = model.physical if model.physical is not None else self.operator.physical
        self.state = MPS.random(model.cat, physical, model.length, model.sector, seed=seed, dtype=dtype,
                                device=device)
        self.left: Dict[int, Tensor] = {}
        self.right: Dict[int, Tensor] = {}

    @property
    def L(self) -> int:
        return self.model.length

    @property
    def center(self) -> int:
        return self.state.center

    @center.setter
    def center(self, at: int) -> None:
        self.state.center = at

    def _advance_left(self, left: Tensor, at: int) -> Tensor:
        tensor = self.state.tensors[at]
        return einsum('EIa,Feq,aPb,IqPj->FJb', left, tensor.conj(), tensor, self.operator.tensors[at])

    def _advance_right(self, right: Tensor, at: int) -> Tensor:
        tensor = self.state.tensors[at]
        return einsum('Gfq,aPb,IqPk,Gkb->FIa', tensor.conj(), tensor, self.operator.tensors[at], right)

    def cache_environments(self) -> None:
        dtype, device = self.state.dtype, self.state.device
        self.left = {0: identity(self.model.cat, self.state.tensors[0].codomain[0],
                                 self.operator.tensors[0].codomain[0], dtype, device)}
        self.right = {self.L: identity(self.model.cat, self.state.tensors[-1].domain[0],
                                       self.operator.tensors[-1].domain[1], dtype, device)}
        for at in range(self.L - 1, 1, -1):
            self.right[at] = self._advance_right(self.right[at + 1], at)

    def update(self, direction: Literal['left', 'right']):
        at = self.center
        first, second = self.state.tensors[at], self.state.tensors[at + 1]
        two = contract(first, second, domain_legs(first), [0])
        site, following = self.operator.tensors[at], self.operator.tensors[at + 1]
        head = einsum('EIa,IPqj->EPaqj', self.left[at], site)
        tail = einsum('JRsk,Fkb->JRSFb', following, self.right[at + 2])

        def action(vector: Tensor) -> Tensor:
            return einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, vector)

        # Lanczos with its tridiagonal solved on the device: no host synchronisation before the
        # split, whose readback also brings the Ritz values staged here (read once it is done).
        if self.precision == 'fp32':
            # The Krylov steps run their block products on FP32 MFMA; the Ritz vector is then
            # corrected in FP64 by a short Lanczos restart from it with FP64 block products.
            with Settings.using(precision='fp32'):
                first_value, vector = lanczos_deferred(action, two, self.n_iter)
            first_value.stage()
            if self.refine:
                value, vector = lanczos_deferred(action, vector, self.refine)
            else:
                value = first_value
        else:
            first_value, vector = lanczos_deferred(action, two, self.n_iter)
            value = first_value
        value.stage()
        factor = 2 if direction == 'right' else 1
        # the split's kept singular values come back on the host with it: the truncated weight and
        # the zero-norm check need no further synchronisation, nor does the normalisation
        left, values, right, kept = _split(vector, None, self.max_bond_dim, self.svd_tol, factor)
        first_value.get()
        energy = value.get()
        self.state.tensors[at], self.state.tensors[at + 1] = left, right.place([0, 1])
        self.center = at + 1 if direction == 'right' else at
        weight = float(np.sum(kept ** 2))
        if weight == 0.0:
            raise ValueError('the split kept no weight: the state cannot be normalised')
        self.state.normalize_on_device()
        return energy, 1.0 - weight

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

    def run(self, num_sweeps: int = 5, energy_tol: float = 1e-10, patience: int = 2, verbose: bool = True) -> Model:
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

    def __repr__(self) -> str:
        return f'DMRG(L={self.L}, max_bond_dim={self.max_bond_dim}, bond_dim={self.state.bond_dims})'


def dmrg(model: Model, max_bond_dim: int = 50, svd_tol: float = 1e-14, num_sweeps: int = 5,
         energy_tol: float = 1e-10, patience: int = 2, n_iter: int = 5, seed: int = 0,
         verbose: bool = True) -> Model:
    '''`dmrg` (`dmrg.py:220-240`).'''
    driver = DMRG(model, max_bond_dim, svd_tol, n_iter, seed)
    return driver.run(num_sweeps, energy_tol, patience, verbose)
