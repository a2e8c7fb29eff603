'''
The committed `to_tuple` fixture (tests/golden/u1_window.npz) through the HIP path: every operand
enters by `Tensor.from_tuple`, the matvec einsum runs on the GPU, and its `to_tuple` matches the
oracle's committed result (same keys in the same order, blocks within 1e-12 relative); a tensor
survives `to_tuple` -> `from_tuple` bit for bit, and blocks left out of `from_tuple` are zero.
'''

import os
import sys

import numpy as np
import pytest
import torch

import tentorch_amd as tt

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))
import tensor_io  # noqa: E402

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(HERE, 'golden', 'u1_window.npz')


def device_tensor(data, name):
    codomain, domain, qtotal, blocks = tensor_io.unpack(data, name)
    return tt.Tensor.from_tuple(tt.u1, [tt.Leg(tt.u1, *l) for l in codomain], [tt.Leg(tt.u1, *l) for l in domain],
                                qtotal, blocks)


def test_fixture_matvec_through_from_tuple_and_to_tuple():
    data = tensor_io.load(FIXTURE)
    head, tail, theta = (device_tensor(data, name) for name in ('head', 'tail', 'theta'))
    codomain, domain, qtotal, result = tt.einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, theta).to_tuple()
    ref_codomain, ref_domain, ref_qtotal, ref = tensor_io.unpack(data, 'result')
    assert tuple(qtotal) == ref_qtotal
    assert [(list(map(tuple, l.sectors)), list(l.multiplicities)) for l in codomain] == \
        [(s, list(m)) for s, m in ref_codomain]
    assert [(tuple(k.codomain), tuple(k.domain), k.copy) for k in result] == list(ref)
    scale = max(np.abs(b).max() for b in ref.values())
    for key, block in result.items():
        want = ref[(tuple(key.codomain), tuple(key.domain), key.copy)]
        assert np.abs(block.cpu().numpy() - want).max() <= 1e-12 * scale


def test_to_tuple_and_from_tuple_round_trip():
    data = tensor_io.load(FIXTURE)
    theta = device_tensor(data, 'theta')
    codomain, domain, qtotal, blocks = theta.to_tuple()
    again = tt.Tensor.from_tuple(tt.u1, codomain, domain, qtotal, {k: b.clone() for k, b in blocks.items()})
    assert again.structure == theta.structure
    assert torch.equal(again.data, theta.data)
    first = next(iter(blocks))
    partial = tt.Tensor.from_tuple(tt.u1, codomain, domain, qtotal, {first: blocks[first]})
    key = partial.keys()[0]
    assert torch.equal(partial.block(key.codomain, key.domain, key.copy), blocks[first])
    assert float(partial.data.abs().sum()) == float(blocks[first].abs().sum())
