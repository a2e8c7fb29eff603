'''
Writes tests/golden/u1_window.npz: the operands of one U(1) two-site DMRG matvec (head, tail and
the two-site tensor, random, seed 23) and the oracle's result of `EPaqj,JRSFb,AQsb->EPrf` on them,
all in the `to_tuple` layout of tensor_io.py.

    python tests/golden/make_fixtures.py
'''

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import oracle.symmetry as osym  # noqa: E402
import oracle.tensor as ot  # noqa: E402
from tensor_io import pack  # noqa: E402

BOND = ([(-3,), (-1,), (1,), (3,)], [2, 4, 4, 2])   # a sector and its dual carry one count (bends need it)
PHYS = ([(-1,), (1,)], [1, 1])
MPO = ([(-2,), (0,), (2,)], [1, 3, 1])
SHAPES = {'head': ([BOND, PHYS], [BOND, PHYS, MPO]), 'tail': ([MPO, PHYS, PHYS, BOND], [BOND]),
          'theta': ([BOND, PHYS], [PHYS, BOND])}
SPEC = 'EPaqj,JRSFb,AQsb->EPrf'


def legs_of(t):
    return [(list(l.sectors), list(l.multiplicities)) for l in t.codomain], \
           [(list(l.sectors), list(l.multiplicities)) for l in t.domain]


def main():
    rng = np.random.default_rng(23)
    tensors = {}
    for name, (cod, dom) in SHAPES.items():
        tensors[name] = ot.random_tensor(osym.u1, [osym.Leg(osym.u1, *l) for l in cod],
                                         [osym.Leg(osym.u1, *l) for l in dom], (0,), rng)
    tensors['result'] = ot.einsum(SPEC, tensors['head'], tensors['tail'], tensors['theta'])
    arrays = {}
    for name, t in tensors.items():
        codomain, domain = legs_of(t)
        arrays.update(pack(name, codomain, domain, t.qtotal, dict(t.blocks)))
    np.savez(os.path.join(HERE, 'u1_window.npz'), **arrays)


if __name__ == '__main__':
    main()
