'''
Tensors in the reference's `to_tuple` layout (`bindings.cpp:131-152,311-327`) as a committed golden
fixture (tests/golden/u1_window.npz, written by tests/golden/make_fixtures.py from the oracle): the
oracle reproduces the committed result from the committed operands, and the format reads back
what it wrote.  The GPU side (`from_tuple` -> einsum -> `to_tuple` against the same fixture) is
test_gpu_tensor_io.py.
'''

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

import oracle.symmetry as osym  # noqa: E402
import oracle.tensor as ot  # noqa: E402
import tensor_io  # noqa: E402

FIXTURE = os.path.join(HERE, 'golden', 'u1_window.npz')


def oracle_tensor(data, name):
    codomain, domain, qtotal, blocks = tensor_io.unpack(data, name)
    return ot.from_blocks(osym.u1, [osym.Leg(osym.u1, *l) for l in codomain],
                          [osym.Leg(osym.u1, *l) for l in domain], qtotal, blocks)


def test_the_oracle_reproduces_the_committed_matvec():
    data = tensor_io.load(FIXTURE)
    head, tail, theta = (oracle_tensor(data, name) for name in ('head', 'tail', 'theta'))
    ref = oracle_tensor(data, 'result')
    got = ot.einsum('EPaqj,JRSFb,AQsb->EPrf', head, tail, theta)
    assert list(got.blocks) == list(ref.blocks)
    assert np.abs(got.flat() - ref.flat()).max() <= 1e-13 * np.abs(ref.flat()).max()


def test_the_format_reads_back_what_it_wrote(tmp_path):
    data = tensor_io.load(FIXTURE)
    arrays = {}
    for name in ('head', 'tail', 'theta', 'result'):
        arrays.update(tensor_io.pack(name, *tensor_io.unpack(data, name)))
    np.savez(tmp_path / 'again.npz', **arrays)
    again = tensor_io.load(tmp_path / 'again.npz')
    for name in ('head', 'tail', 'theta', 'result'):
        first, second = tensor_io.unpack(data, name), tensor_io.unpack(again, name)
        assert first[:3] == second[:3]
        assert list(first[3]) == list(second[3])
        for key in first[3]:
            assert np.array_equal(first[3][key], second[3][key])
