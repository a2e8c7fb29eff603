'''
Matrix product states and operators over the backend's tensors: restates
`tentorch/tensors/tree.py` (`rows`, `entered`), `tentorch/networks/mps.py` and
`tentorch/networks/mpo.py`.  The random state is drawn on the host from `numpy.random.default_rng(seed)`
block by block in key order and uploaded (the reference draws from torch's generator on its device,
`mps.py:115-116`), so a run is reproducible across devices.
'''

import itertools
import math
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native
from .symmetry import Leg, fusion_trees, reachable_charges
from .tensor import Tensor, block_keys, contract, domain_legs, einsum, inner
from .linalg import qr, rq


def rows(cat, legs: Sequence, root):
    '''`rows` (`tentorch/tensors/tree.py:35-47`): every row of a side in block order.'''
    for tree in fusion_trees(cat, legs, root):
        mults = [leg.multiplicity(leg.sectors.index(s)) for leg, s in zip(legs, tree.sectors)]
        for copy in itertools.product(*[range(m) for m in mults]):
            yield tree, copy


def entered(cat, codomain, domain, entry, root=None, dtype=None, device=None) -> Tensor:
    '''`entered` (`tentorch/tensors/tree.py:50-75`): a tensor written entry by entry on the host.'''
    charge = cat.trivial() if root is None else root
    blocks = {}
    for key in block_keys(cat, list(codomain), list(domain), charge):
        out_rows = list(rows(cat, codomain, key.codomain))
        in_rows = list(rows(cat, domain, key.domain))
        data = np.zeros((len(out_rows), len(in_rows)))
        for r, (t_out, c_out) in enumerate(out_rows):
            for c, (t_in, c_in) in enumerate(in_rows):
                data[r, c] = entry(t_out, c_out, t_in, c_in)
        blocks[key] = data
    return Tensor.from_tuple(cat, list(codomain), list(domain), charge, blocks, dtype, device)


def identity(cat, bond, middle=None, dtype=None, device=None) -> Tensor:
    '''`_identity` (`tentorch/networks/mps.py:12-28`).'''
    def entry(t_out, c_out, t_in, c_in):
        same = t_out.sectors[0] == t_in.sectors[0] and c_out[0] == c_in[0]
        return 1.0 if same and (middle is None or t_out.sectors[1] == cat.trivial()) else 0.0
    codomain = [bond] if middle is None else [bond, middle]
    return entered(cat, codomain, [bond], entry, dtype=dtype, device=device)


class MPO:
    '''`MPO` (`tentorch/networks/mpo.py:7-86`).'''

    def __init__(self, sites, physical, cat, sector=None, dtype=None, device=None):
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
        self.dtype = self.tensors[0].dtype if dtype is None else dtype
        self.device = self.tensors[0].device if device is None else device
        self.bond_dim = [site.codomain[0].multiplets() for site in self.tensors]
        self.bond_dim.append(self.tensors[-1].domain[1].multiplets())

    @property
    def physical_dim(self) -> int:
        return sum(self.physical.multiplicity(b) for b in range(self.physical.nblocks()))

    def __len__(self):
        return self.L

    def __iter__(self):
        return iter(self.tensors)

    def __getitem__(self, at):
        return self.tensors[at]

    def __repr__(self):
        return f'MPO(L={self.L}, physical_dim={self.physical_dim}, bond_dim={self.bond_dim[1:self.L]})'


class MPS:
    '''`MPS` (`tentorch/networks/mps.py:31-204`): mixed canonical form around `center`.'''

    def __init__(self, tensors, physical, center: int = 0):
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
    def dtype(self):
        return self.tensors[0].dtype

    @property
    def device(self):
        return self.tensors[0].device

    @property
    def bond_dims(self):
        return [self.tensors[at].domain[0].multiplets() for at in range(self.L - 1)]

    @classmethod
    def random(cls, cat, physical, length, sector=None, seed=0, multiplets=1, dtype=None, device=None):
        '''`MPS.random` (`mps.py:86-118`).'''
        charge = cat.trivial() if sector is None else sector
        left = [reachable_charges(cat, [physical] * at) for at in range(length + 1)]
        tail = [reachable_charges(cat, [physical] * (length - at)) for at in range(length + 1)]
        bonds = []
        for at in range(length + 1):
            sectors = [b for b in left[at] if any(cat.can_fuse_to(b, r, charge) for r in tail[at])]
            bonds.append(Leg(cat, sectors, [multiplets] * len(sectors)))
        bonds[0] = bonds[-1] = Leg(cat, [charge], (1,))
        if not bonds[0].nblocks() or not bonds[-1].nblocks():
            raise ValueError(f'the charge {charge} admits no state of {length} sites')
        rng = np.random.default_rng(seed)
        tensors = [Tensor.random(cat, [bonds[at], physical], [bonds[at + 1]], charge, dtype, device, generator=rng)
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

    def norm(self) -> float:
        t = self.tensors[self.center]
        return math.sqrt(float((t.data.abs() ** 2).sum()))

    def normalize(self):
        self.tensors[self.center] /= self.norm()

    def normalize_on_device(self):
        '''`normalize` with the norm left on the device: the same sum of squares, square root and
        reciprocal as `normalize` (both correctly rounded), then the same scaling kernel, without a
        synchronisation.  A zero norm is not checked here: the caller knows the weight it kept.'''
        t = self.tensors[self.center]
        if t.dtype.is_complex:
            self.normalize()
            return
        scale = torch.reciprocal(torch.sqrt((t.data.abs() ** 2).sum()))
        _native.axpby(t.dtype, 0.0, None, 0.0, t.data, beta_dev=scale)
        t._version += 1

    def expectation(self, operator) -> complex:
        '''`MPS.expectation` (`mps.py:176-195`).'''
        left = identity(self.cat, self.tensors[0].codomain[0], operator.tensors[0].codomain[0], self.dtype, self.device)
        for at, t in enumerate(self.tensors):
            left = einsum('AIe,aPb,Feq,IPqj->BJf', left, t, t.conj(), operator.tensors[at])
        return inner(identity(self.cat, self.tensors[-1].domain[0], operator.tensors[-1].domain[1], self.dtype,
                              self.device), left)

    def __len__(self):
        return self.L

    def __repr__(self):
        return (f'MPS(L={self.L}, blocks={self.physical.nblocks()}, bond_dim={self.bond_dims}, '
                f'center={self.center})')
