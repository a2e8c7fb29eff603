'''
tentorch_amd: an MI355X-native backend for TenTorch's block-sparse contraction path.

The public names follow the reference package (`tentorch/__init__.py:5-16`).  Tensors live in one
flat device buffer each; `einsum`, `contract` and `Tensor.place` run programs the C++ host engine
plans once per shape and the HIP device library (gfx950) executes.
'''

from .symmetry import Leg, Sector, fermion_parity, fibonacci, no_symmetry, su2, u1, z_n
from .tensor import (BlockKey, Tensor, block_keys, codomain_legs, contract, domain_legs, einsum, inner,
                     is_close, norm)
from .networks import MPO, MPS
from .models import Model, heisenberg
from .linalg import lanczos, qr, rq, svd, svdvals
from .algorithms import DMRG, dmrg

__all__ = [
    'BlockKey', 'DMRG', 'Leg', 'MPO', 'MPS', 'Model', 'Sector', 'Tensor', 'block_keys', 'codomain_legs',
    'contract', 'dmrg', 'domain_legs', 'einsum', 'fermion_parity', 'fibonacci', 'heisenberg', 'inner',
    'is_close', 'lanczos', 'no_symmetry', 'norm', 'qr', 'rq', 'su2', 'svd', 'svdvals', 'u1', 'z_n'
]
