'''
Matrix product states and operators over the backend's tensors.

The public surface is the reference's: `MPS` (`tentorch/networks/mps.py:31-204`), `MPO`
(`tentorch/networks/mpo.py:7-86`) and the entry-wise constructors of `tentorch/tensors/tree.py`
(`rows`, `entered`), under the same names, arguments, errors and results.  The organisation is this
module's own:

- a chain's sites are held in mixed canonical form around `center`, and the only operation that
  moves the centre is `_gauge_step`, one site to the left or right (QR to the right, RQ to the
  left, the small factor folded into the neighbour).  `move_center`, `canonicalize` and the DMRG
  turn are sequences of that step;
- the random state's bonds come from `chain_bonds`, which intersects the charges reachable from the
  two ends of the chain;
- tensors written entry by entry enumerate each side's rows once per root (`side_rows`) and fill a
  numpy block per key before one upload.

The random state is drawn on the host from `numpy.random.default_rng(seed)`, block by block in key
order, and uploaded (the reference draws from torch's generator on its device, `mps.py:115-116`),
so that a run is reproducible across devices and the oracle starts from the same state.
'''

import math
from itertools import product
from typing import Callable, List, Optional, Sequence

import numpy as np
import torch

from . import _native
from .linalg import qr, rq
from .symmetry import Leg, fusion_trees, reachable_charges
from .tensor import Tensor, block_keys, contract, domain_legs, einsum, inner


# ------------------------------------------------------------------------------------------------
# Tensors written entry by entry


def side_rows(cat, legs: Sequence, root) -> list:
    '''Every row of one side of a block, in the block's row order: the fusion trees of `legs` onto
    `root` in tree order, each expanded over the multiplicity copies of its sectors with the last
    leg's copy running fastest (`tentorch/tensors/tree.py:35-47`).  Returns (tree, copies) pairs.'''
    counts = [dict(zip(map(tuple, leg.sectors), leg.multiplicities)) for leg in legs]
    out = []
    for tree in fusion_trees(cat, legs, root):
        extents = [count[tuple(sector)] for count, sector in zip(counts, tree.sectors)]
        out.extend((tree, copies) for copies in product(*map(range, extents)))
    return out


# `rows` is the reference's name for the same enumeration.
rows = side_rows


def entered(cat, codomain, domain, entry: Callable, root=None, dtype=None, device=None) -> Tensor:
    '''`entered` (`tentorch/tensors/tree.py:50-75`): the tensor whose entry at (codomain row, domain
    row) of every block is `entry(tree_out, copies_out, tree_in, copies_in)`.'''
    charge = cat.trivial() if root is None else root
    codomain, domain = list(codomain), list(domain)
    cache = {}

    def rows_of(side, legs, sector):
        key = (side, tuple(sector))
        if key not in cache:
            cache[key] = side_rows(cat, legs, sector)
        return cache[key]

    blocks = {}
    for key in block_keys(cat, codomain, domain, charge):
        above, below = rows_of(0, codomain, key.codomain), rows_of(1, domain, key.domain)
        values = [entry(t_out, c_out, t_in, c_in) for (t_out, c_out) in above for (t_in, c_in) in below]
        blocks[key] = np.asarray(values, dtype=np.complex128 if any(isinstance(v, complex) for v in values)
                                 else np.float64).reshape(len(above), len(below))
    return Tensor.from_tuple(cat, codomain, domain, charge, blocks, dtype, device)


def identity(cat, bond, middle=None, dtype=None, device=None) -> Tensor:
    '''The edge environment of a chain (`tentorch/networks/mps.py:12-28`): the identity on `bond`,
    over (bond | bond), or over (bond, middle | bond) with the identity placed on the trivial sector
    of `middle`.'''
    trivial = cat.trivial()

    def entry(t_out, c_out, t_in, c_in):
        if t_out.sectors[0] != t_in.sectors[0] or c_out[0] != c_in[0]:
            return 0.0
        return 1.0 if middle is None or tuple(t_out.sectors[1]) == tuple(trivial) else 0.0

    return entered(cat, [bond] if middle is None else [bond, middle], [bond], entry, dtype=dtype, device=device)


# ------------------------------------------------------------------------------------------------
# Operators


class MPO:
    '''A matrix product operator (`tentorch/networks/mpo.py:7-86`): site tensors over
    (bond left, out | in, bond right), one physical leg shared by every site.'''

    def __init__(self, sites, physical, cat, sector=None, dtype=None, device=None):
        tensors = list(sites)
        if not tensors:
            raise ValueError('an operator needs at least one site')
        wrong = next(((at, t.nlegs()) for at, t in enumerate(tensors) if t.nlegs() != 4), None)
        if wrong is not None:
            raise ValueError(f'site {wrong[0]} has {wrong[1]} legs, not the four of an operator')
        self.tensors = tensors
        self.L = len(tensors)
        self.cat, self.physical = cat, physical
        self.sector = cat.trivial() if sector is None else sector
        self.dtype = tensors[0].dtype if dtype is None else dtype
        self.device = tensors[0].device if device is None else device
        # the virtual bond at every cut, the two open ends included
        self.bond_dim = [t.codomain[0].multiplets() for t in tensors] + [tensors[-1].domain[1].multiplets()]

    @property
    def physical_dim(self) -> int:
        return sum(map(self.physical.multiplicity, range(self.physical.nblocks())))

    def __len__(self):
        return self.L

    def __iter__(self):
        return iter(self.tensors)

    def __getitem__(self, at):
        return self.tensors[at]

    def __repr__(self):
        return f'MPO(L={self.L}, physical_dim={self.physical_dim}, bond_dim={self.bond_dim[1:self.L]})'


# ------------------------------------------------------------------------------------------------
# States


def chain_bonds(cat, physical, length: int, charge, multiplets: int = 1) -> List:
    '''The bond legs of a random state of `length` sites and total `charge`
    (`tentorch/networks/mps.py:103-111`): at cut c, every charge reachable by c physical legs from
    the left end that still fuses with some charge reachable by the remaining `length - c` legs to
    `charge`, each with `multiplets` copies; both open ends carry `charge` once.'''
    from_left = [reachable_charges(cat, [physical] * cut) for cut in range(length + 1)]
    bonds = []
    for cut in range(length + 1):
        closing = from_left[length - cut]     # the same enumeration from the right end
        allowed = [b for b in from_left[cut] if any(cat.can_fuse_to(b, r, charge) for r in closing)]
        bonds.append(Leg(cat, allowed, [multiplets] * len(allowed)))
    end = Leg(cat, [charge], (1,))
    if not end.nblocks():
        raise ValueError(f'the charge {charge} admits no state of {length} sites')
    bonds[0] = bonds[-1] = end
    return bonds


class MPS:
    '''A matrix product state in mixed canonical form around `center` (`tentorch/networks/mps.py`):
    sites over (bond, physical | bond), left isometric before the centre and right isometric after
    it, so the state's norm lives in the centre site alone.'''

    def __init__(self, tensors, physical, center: int = 0):
        sites = list(tensors)
        if not sites:
            raise ValueError('A state needs at least one site.')
        wrong = next(((at, t.nlegs()) for at, t in enumerate(sites) if t.nlegs() != 3), None)
        if wrong is not None:
            raise ValueError(f'site {wrong[0]} needs three legs, not {wrong[1]}')
        if center < 0 or center >= len(sites):
            raise IndexError(f'the centre {center} is outside the chain of {len(sites)} sites')
        self.tensors = sites
        self.physical = physical
        self.center = center
        self.canonicalize()

    # -- shape
    @property
    def L(self) -> int:
        return len(self.tensors)

    @property
    def cat(self):
        return self.tensors[0].cat

    @property
    def dtype(self):
        return self.tensors[0].dtype

    @property
    def device(self):
        return self.tensors[0].device

    @property
    def bond_dims(self) -> List[int]:
        return [site.domain[0].multiplets() for site in self.tensors[:-1]]

    @classmethod
    def random(cls, cat, physical, length, sector=None, seed=0, multiplets=1, dtype=None, device=None) -> 'MPS':
        '''`MPS.random` (`mps.py:86-118`): N(0, 1) sites over every bond charge the total charge
        admits, centre at the left end.'''
        charge = cat.trivial() if sector is None else sector
        bonds = chain_bonds(cat, physical, length, charge, multiplets)
        draw = np.random.default_rng(seed)
        sites = [Tensor.random(cat, [left, physical], [right], charge, dtype, device, generator=draw)
                 for left, right in zip(bonds[:-1], bonds[1:])]
        return cls(sites, physical)

    # -- gauge
    def _gauge_step(self, toward: int) -> None:
        '''Move the centre one site: `toward` +1 splits the centre site by QR over (bond, physical)
        and folds R into the next site; -1 splits it by RQ over its left bond and folds R into the
        previous site (`mps.py:120-137`).'''
        at = self.center
        if toward > 0:
            if at < 0 or at >= self.L - 1:
                raise IndexError(f'index {at} has no site to its right in a chain of {self.L} sites')
            isometry, carried = qr(self.tensors[at], [0, 1])
            joined = contract(carried, self.tensors[at + 1], domain_legs(carried), [0])
            self.tensors[at], self.tensors[at + 1] = isometry, joined.place([0, 1])
        else:
            if at <= 0 or at >= self.L:
                raise IndexError(f'index {at} has no site to its left in a chain of {self.L} sites')
            carried, isometry = rq(self.tensors[at], [0])
            previous = self.tensors[at - 1]
            self.tensors[at - 1] = contract(previous, carried, domain_legs(previous), [0])
            self.tensors[at] = isometry.place([0, 1])
        self.center = at + (1 if toward > 0 else -1)

    def move_center(self, at: int) -> None:
        '''Carry the centre to site `at` one gauge step at a time.'''
        if at < 0 or at >= self.L:
            raise IndexError(f'the centre {at} is outside the chain of {self.L} sites')
        while self.center != at:
            self._gauge_step(1 if at > self.center else -1)

    def canonicalize(self) -> None:
        '''Bring every site into the mixed canonical form of the current centre: walk a gauge from the
        left end up to the centre, then from the right end down to it.'''
        target = self.center
        self.center = 0
        self.move_center(target)
        self.center = self.L - 1
        self.move_center(target)

    # -- norm
    def norm(self) -> float:
        return math.sqrt(float(torch.sum(self.tensors[self.center].data.abs() ** 2)))

    def normalize(self) -> None:
        self.tensors[self.center] /= self.norm()

    def normalize_on_device(self) -> None:
        '''`normalize` with the norm left on the device: the same sum of squares, square root and
        reciprocal (correctly rounded) and the same scaling kernel, without a synchronisation.  A
        zero norm is not checked here: the caller knows the weight it kept.'''
        site = self.tensors[self.center]
        if site.dtype.is_complex:
            self.normalize()
            return
        scale = torch.reciprocal(torch.sqrt(torch.sum(site.data.abs() ** 2)))
        _native.axpby(site.dtype, 0.0, None, 0.0, site.data, beta_dev=scale)
        site._version += 1

    # -- measurement
    def expectation(self, operator) -> complex:
        '''<psi|O|psi> by the transfer from the left end (`mps.py:176-195`): the edge identity over
        (ket bond, operator bond | bra bond), absorbing one site, its conjugate and the operator site
        per step, closed by the right edge's identity.'''
        edge = identity(self.cat, self.tensors[0].codomain[0], operator.tensors[0].codomain[0], self.dtype,
                        self.device)
        for site, op in zip(self.tensors, operator.tensors):
            edge = einsum('AIe,aPb,Feq,IPqj->BJf', edge, site, site.conj(), op)
        closing = identity(self.cat, self.tensors[-1].domain[0], operator.tensors[-1].domain[1], self.dtype,
                           self.device)
        return inner(closing, edge)

    def __len__(self):
        return self.L

    def __repr__(self):
        return (f'MPS(L={self.L}, blocks={self.physical.nblocks()}, bond_dim={self.bond_dims}, '
                f'center={self.center})')
