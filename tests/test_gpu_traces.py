'''
Einsum traces inside one operand on the HIP path (`tentorch/linalg/graded/einsum.cpp:95-159,386-439`):
the device result of every trace equation of tests/test_traces.py, in every category, against the
oracle within 1e-12 relative; a rerun after the traced operand is edited sees the edit (the trace is
a hoisted move, skipped only while its operand is unchanged).
'''

import numpy as np
import pytest
import torch

import tentorch_amd as tt
import oracle.tensor as ot
from test_traces import CATS, LEGS, _cases

pytestmark = pytest.mark.gpu


def _device(name, o):
    cat_p, _ = CATS[name]
    legs = o.legs()
    ncod = len(o.codomain)
    dtype = torch.complex128 if o.dtype == np.complex128 else torch.float64
    d = tt.Tensor(cat_p, [tt.Leg(cat_p, l.sectors, l.multiplicities) for l in legs[:ncod]],
                  [tt.Leg(cat_p, l.sectors, l.multiplicities) for l in legs[ncod:]], o.qtotal, dtype=dtype)
    d.data.copy_(torch.from_numpy(o.flat()))
    return d


@pytest.mark.parametrize('name', sorted(LEGS))
def test_trace_einsums_match_the_oracle(name):
    for spec, operands, _ in _cases(name):
        ref = ot.einsum(spec, *operands).flat()
        got = tt.einsum(spec, *[_device(name, o) for o in operands]).data.cpu().numpy()
        assert got.shape == ref.shape, spec
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), spec


def test_a_rerun_after_editing_the_traced_operand_sees_the_edit():
    spec, operands, _ = _cases('su2')[-1]      # 'abac,cd->bd'
    first, second = (_device('su2', o) for o in operands)
    before = tt.einsum(spec, first, second).data.clone()
    first.data.mul_(2.0)
    after = tt.einsum(spec, first, second).data
    assert torch.allclose(after, 2.0 * before, rtol=1e-13, atol=1e-14)
