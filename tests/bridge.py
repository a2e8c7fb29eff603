'''
Moving tensors between the oracle (numpy, host) and the product (flat device buffers): both keep
every block in key order, row-major, so a flat copy carries a tensor across once the legs are
rebuilt in the other library.
'''

import numpy as np
import torch

import oracle.symmetry as osym
import oracle.tensor as ot
import tentorch_amd as tt

CATEGORIES = {'U1': (tt.u1, osym.u1), 'NoSymmetry': (tt.no_symmetry, osym.no_symmetry)}


def _pair_of(cat_o):
    for product_cat, oracle_cat in CATEGORIES.values():
        if oracle_cat is cat_o:
            return product_cat
    raise KeyError(cat_o.name)


def device_leg(cat_p, leg):
    return tt.Leg(cat_p, [tuple(s) for s in leg.sectors], list(leg.multiplicities))


def oracle_leg(cat_o, leg):
    return osym.Leg(cat_o, [tuple(s) for s in leg.sectors], list(leg.multiplicities))


def to_device(o: ot.Tensor) -> tt.Tensor:
    cat_p = _pair_of(o.cat)
    dtype = torch.complex128 if np.iscomplexobj(np.zeros(0, dtype=o.dtype)) else torch.float64
    d = tt.Tensor(cat_p, [device_leg(cat_p, l) for l in o.codomain], [device_leg(cat_p, l) for l in o.domain],
                  o.qtotal, dtype=dtype)
    assert d.data.numel() == o.numel()
    d.data.copy_(torch.from_numpy(o.flat()))
    return d


def to_oracle(d: tt.Tensor, cat_o) -> ot.Tensor:
    o = ot.Tensor(cat_o, [oracle_leg(cat_o, l) for l in d.codomain], [oracle_leg(cat_o, l) for l in d.domain],
                  tuple(d.qtotal), np.complex128 if d.dtype.is_complex else np.float64)
    o.set_flat(d.data.cpu().numpy())
    return o


def same_legs(d_leg, o_leg) -> bool:
    return ([tuple(s) for s in d_leg.sectors] == [tuple(s) for s in o_leg.sectors]
            and list(d_leg.multiplicities) == list(o_leg.multiplicities))
