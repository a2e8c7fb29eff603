'''
Two-site DMRG of a chain over the backend's tensors.

The public surface is the reference's `DMRG`/`dmrg` (`tentorch/algorithms/dmrg.py:13-240`): the
same constructor arguments, the same window (the two-site tensor, the `head`/`tail` halves of the
effective Hamiltonian, Lanczos on the three-operand matvec einsum, the truncating split), the same
window order within a sweep, the same per-sweep report and convergence rule, and the same errors.
The organisation is this module's own, in three parts:

- `Environments` holds the left and right blocks of the effective Hamiltonian by cut and knows
  which are current.  A window asks for the two it needs and gets them grown from the nearest
  current one; a change to sites invalidates exactly the blocks that contain them.  So the sweep
  never schedules environment updates itself, and a measurement after the centre was moved by hand
  reads current blocks.
- `sweep_schedule` lists a sweep's steps: the windows of the right pass, the turn that hands the
  centre back from the last site, and the windows of the left pass.
- `DMRG.update` solves one window.  The matvec is the same einsum call on every Krylov step; the
  backend plans it once per window, hoists the moves of `head` and `tail` out of the Krylov loop, and
  runs each step as one composed move and two grouped block-product launches.  In the FP32 form the
  Krylov products run on FP32 MFMA and a short FP64 Lanczos restart from the Ritz vector corrects it.

Cut c sits between sites c - 1 and c: the left block at c contracts sites < c, the right block at c
contracts sites >= c.
'''

from typing import Dict, Iterator, Literal, Tuple

import numpy as np

from .linalg import EinsumAction, _split, lanczos_deferred
from .models import Model, heisenberg
from .networks import MPS, identity
from .tensor import Settings, Tensor, contract, domain_legs, einsum, inner

# The einsum equations of the window and of the environments (`dmrg.py:65,80,113-118,176-179`);
# labels in upper case sit in the codomain.
GROW_LEFT = 'EIa,Feq,aPb,IqPj->FJb'        # (left block, bra site, ket site, operator site)
GROW_RIGHT = 'Gfq,aPb,IqPk,Gkb->FIa'       # (bra site, ket site, operator site, right block)
HEAD = 'EIa,IPqj->EPaqj'                   # (left block, operator site)
TAIL = 'JRsk,Fkb->JRSFb'                   # (operator site, right block)
MATVEC = 'EPaqj,JRSFb,AQsb->EPrf'          # (head, tail, two-site tensor)
WINDOW_ENERGY = 'AIe,IPqj,JRsk,Dkf,EQSf->APrd'


class Environments:
    '''The effective-Hamiltonian blocks of a chain, by cut, grown on demand.'''

    def __init__(self, state: MPS, operator) -> None:
        self.state, self.operator = state, operator
        self.left: Dict[int, Tensor] = {}
        self.right: Dict[int, Tensor] = {}

    @property
    def built(self) -> bool:
        return bool(self.left)

    def reset(self) -> None:
        '''Only the two edge blocks: identities over the open bonds and the operator's edge bonds.'''
        state, operator = self.state, self.operator
        self.left = {0: identity(state.cat, state.tensors[0].codomain[0], operator.tensors[0].codomain[0],
                                 state.dtype, state.device)}
        self.right = {state.L: identity(state.cat, state.tensors[-1].domain[0], operator.tensors[-1].domain[1],
                                        state.dtype, state.device)}

    def changed(self, first: int, last: int) -> None:
        '''Sites first..last were replaced: drop every block that contracts one of them.'''
        for cut in [c for c in self.left if c > first]:
            del self.left[cut]
        for cut in [c for c in self.right if c <= last]:
            del self.right[cut]

    def left_of(self, cut: int) -> Tensor:
        '''The left block at `cut`, grown site by site from the nearest current one below it.'''
        start = max(c for c in self.left if c <= cut)
        for site in range(start, cut):
            ket = self.state.tensors[site]
            self.left[site + 1] = einsum(GROW_LEFT, self.left[site], ket.conj(), ket, self.operator.tensors[site])
        return self.left[cut]

    def right_of(self, cut: int) -> Tensor:
        '''The right block at `cut`, grown site by site from the nearest current one above it.'''
        start = min(c for c in self.right if c >= cut)
        for site in range(start - 1, cut - 1, -1):
            ket = self.state.tensors[site]
            self.right[site] = einsum(GROW_RIGHT, ket.conj(), ket, self.operator.tensors[site], self.right[site + 1])
        return self.right[cut]


def sweep_schedule(length: int) -> Iterator[Tuple[str, int]]:
    '''One sweep's steps (`dmrg.py:141-161`): ('right', at) solves the window (at, at + 1) and leaves
    the centre on at + 1, for at = 0..L-2; ('turn', L-1) gauges the last site right isometric,
    handing the centre to L-2; ('left', at) solves the window (at, at + 1) and leaves the centre on
    at, for at = L-3..0.'''
    for at in range(length - 1):
        yield 'right', at
    yield 'turn', length - 1
    for at in range(length - 3, -1, -1):
        yield 'left', at


def _settled(energies, tol: float, patience: int) -> bool:
    '''Whether the run may stop (`dmrg.py:201-209`): the last `patience` sweeps all stayed within
    `tol` of the anchor, the energy of the last sweep that moved by more than `tol` from its own
    anchor (the first sweep anchors against +inf, so it always moves).'''
    anchor, quiet = float('inf'), 0
    for energy in energies:
        if abs(energy - anchor) <= tol:
            quiet += 1
        else:
            anchor, quiet = energy, 0
    return quiet >= patience


class DMRG:
    '''`DMRG` (`tentorch/algorithms/dmrg.py:13-212`): the two-site density matrix renormalisation
    group of `model`, keeping at most `max_bond_dim` multiplets per bond and dropping singular values
    at or below `svd_tol`, with `n_iter` Krylov steps per window from a random state of `seed`.
    `precision` 'fp32' runs the Krylov block products on FP32 MFMA followed by `refine` FP64 Lanczos
    steps from the Ritz vector; 'fp64' keeps every product in FP64.'''

    def __init__(self, model: Model, max_bond_dim: int = 50, svd_tol: float = 1e-14, n_iter: int = 5,
                 seed: int = 0, dtype=None, device=None, precision: str = 'fp64', refine: int = 3) -> None:
        if precision not in ('fp64', 'fp32'):
            raise ValueError(f"precision has to be 'fp64' or 'fp32', not {precision!r}")
        self.model = model
        self.max_bond_dim, self.svd_tol, self.n_iter = max_bond_dim, svd_tol, n_iter
        self.precision, self.refine = precision, refine
        self.operator = heisenberg(model.cat, model.length, model.coupling, dtype, device)
        physical = self.operator.physical if model.physical is None else model.physical
        self.state = MPS.random(model.cat, physical, model.length, model.sector, seed=seed, dtype=dtype,
                                device=device)
        self.environments = Environments(self.state, self.operator)

    # -- the reference's attribute surface
    @property
    def L(self) -> int:
        return self.model.length

    @property
    def center(self) -> int:
        return self.state.center

    @center.setter
    def center(self, at: int) -> None:
        self.state.center = at

    @property
    def left(self) -> Dict[int, Tensor]:
        return self.environments.left

    @property
    def right(self) -> Dict[int, Tensor]:
        return self.environments.right

    def cache_environments(self) -> None:
        '''Rebuild the blocks from the current state: the edges, and every right block down to cut 2.'''
        self.environments.reset()
        self.environments.right_of(2)

    # -- one window
    def _solve(self, head: Tensor, tail: Tensor, start: Tensor):
        '''The window's ground state: Ritz value (still on the device) and Ritz vector.'''
        matvec = EinsumAction(MATVEC, [head, tail], 2)

        if self.precision == 'fp64':
            value, vector = lanczos_deferred(matvec, start, self.n_iter)
            return value.stage(), vector
        with Settings.using(precision='fp32'):
            rough, vector = lanczos_deferred(matvec, start, self.n_iter)
        rough.stage()
        if not self.refine:
            return rough, vector
        # defect correction: a short FP64 restart from the FP32 Ritz vector
        value, vector = lanczos_deferred(matvec, vector, self.refine)
        return value.stage(), vector

    def update(self, direction: Literal['left', 'right']) -> Tuple[float, float]:
        '''Solve the window (center, center + 1) and split it, carrying the singular values toward
        `direction` (`dmrg.py:96-139`).  Returns the window's energy and the weight the split dropped.
        Nothing is read back before the split: its one readback also brings the Ritz value.'''
        at = self.center
        envs = self.environments
        first, second = self.state.tensors[at], self.state.tensors[at + 1]
        start = contract(first, second, domain_legs(first), [0])
        head = einsum(HEAD, envs.left_of(at), self.operator.tensors[at])
        tail = einsum(TAIL, self.operator.tensors[at + 1], envs.right_of(at + 2))
        value, vector = self._solve(head, tail, start)
        # absorb 2 leaves the left factor isometric (the centre moves right), absorb 1 the right one
        absorb = 2 if direction == 'right' else 1
        isometry_left, _, rest, kept = _split(vector, None, self.max_bond_dim, self.svd_tol, absorb,
                                              device_values=False)
        energy = value.get()
        self.state.tensors[at], self.state.tensors[at + 1] = isometry_left, rest.place([0, 1])
        envs.changed(at, at + 1)
        self.center = at + 1 if direction == 'right' else at
        weight = float(np.sum(kept ** 2))
        if weight == 0.0:
            raise ValueError('the split kept no weight: the state cannot be normalised')
        self.state.normalize_on_device()
        return energy, 1.0 - weight

    # -- sweeps
    def sweep(self) -> Tuple[float, float]:
        '''One right pass and one left pass (`dmrg.py:141-161`).  Returns the last window's energy and
        the largest weight any window dropped.'''
        self.state.move_center(0)
        self.environments.reset()
        energy, dropped = 0.0, 0.0
        for kind, at in sweep_schedule(self.L):
            if kind == 'turn':
                self.state._gauge_step(-1)
                self.environments.changed(at - 1, at)
                continue
            self.center = at
            energy, lost = self.update(kind)
            dropped = max(dropped, lost)
        return energy, dropped

    def energy(self) -> complex:
        '''The Rayleigh quotient of the window on the centre (`dmrg.py:163-180`).'''
        if not self.environments.built:
            raise RuntimeError('the energy needs the environments of a sweep, and none has run')
        at = min(self.center, self.L - 2)
        first, second = self.state.tensors[at], self.state.tensors[at + 1]
        two = contract(first, second, domain_legs(first), [0])
        image = einsum(WINDOW_ENERGY, self.environments.left_of(at), self.operator.tensors[at],
                       self.operator.tensors[at + 1], self.environments.right_of(at + 2), two)
        return complex(inner(two, image) / inner(two, two))

    def run(self, num_sweeps: int = 5, energy_tol: float = 1e-10, patience: int = 2, verbose: bool = True) -> Model:
        '''Sweep until `patience` sweeps in a row leave the energy within `energy_tol`, at most
        `num_sweeps`, then write the state and the last energy onto the model (`dmrg.py:182-212`).'''
        if num_sweeps < 1:
            raise ValueError(f'a run needs at least one sweep, not {num_sweeps}')
        history = []
        for number in range(1, num_sweeps + 1):
            energy, dropped = self.sweep()
            history.append(energy)
            if verbose:
                print(f'sweep {number}: energy = {energy:.12f}, truncated = {dropped:.2e}, '
                      f'bond = {self.state.bond_dims}')
            if _settled(history, energy_tol, patience):
                break
        self.model.state, self.model.energy = self.state, history[-1]
        return self.model

    def __repr__(self) -> str:
        return f'DMRG(L={self.L}, max_bond_dim={self.max_bond_dim}, bond_dim={self.state.bond_dims})'


def dmrg(model: Model, max_bond_dim: int = 50, svd_tol: float = 1e-14, num_sweeps: int = 5,
         energy_tol: float = 1e-10, patience: int = 2, n_iter: int = 5, seed: int = 0,
         verbose: bool = True) -> Model:
    '''The ground state of `model` by two-site DMRG (`dmrg.py:220-240`).'''
    return DMRG(model, max_bond_dim, svd_tol, n_iter, seed).run(num_sweeps, energy_tol, patience, verbose)
