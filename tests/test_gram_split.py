'''
The split's device pipeline (`tentorch_amd.gram.GramSplit`: Gram products, eigendecompositions,
projections, decoupled tail passes, factor assembly) with its launches interpreted on the CPU
(tests/emulate_split.py), against the oracle's per-block SVD: the same kept multiplets per sector
and the same values on spectra spanning sixteen decades with exact zeros, whether the cap cuts high,
deep in the tail, or not at all (the tolerance then decides, and the zeros must go).  The GPU test
of the same cases runs the kernels themselves (test_gpu_parity.py).
'''

import numpy as np
import pytest
import torch

import oracle.tensor as ot
import tentorch_amd as tt
from tentorch_amd.linalg import SvdSettings, _split
from bridge import same_legs, to_oracle, device_leg
from emulate_split import emulated_split
from test_gpu_parity import _deep_spectrum_tensor


def _on_cpu(o):
    d = tt.Tensor(tt.u1, [device_leg(tt.u1, l) for l in o.codomain], [device_leg(tt.u1, l) for l in o.domain],
                  o.qtotal, device='cpu')
    d.data.copy_(torch.from_numpy(o.flat()))
    return d


@pytest.mark.parametrize('absorb', [2, 1])
@pytest.mark.parametrize('cap', [None, 30, 45])
def test_device_split_pipeline_keeps_the_oracles_multiplets(absorb, cap):
    import oracle.symmetry as osym
    ref_u, ref_s, ref_v = ot.svd(_deep_spectrum_tensor(17), None, cap, 1e-14, absorb)
    with emulated_split():
        u, s, v, kept = _split(_on_cpu(_deep_spectrum_tensor(17)), None, cap, 1e-14, absorb)
    assert same_legs(u.domain[0], ref_u.domain[0])
    assert np.abs(kept - ref_s).max() <= 1e-14 * np.abs(ref_s).max()
    got = ot.contract(to_oracle(u, osym.u1), to_oracle(v, osym.u1), [2], [0]).flat()
    ref = ot.contract(ref_u, ref_v, [2], [0]).flat()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    iso = (u if absorb == 2 else v)
    for block in iso.blocks().values():
        m = block.numpy() if absorb == 2 else block.numpy().T
        assert np.abs(m.T @ m - np.eye(m.shape[1])).max() <= 1e-12


def test_device_split_resolves_a_near_degenerate_pair_at_a_pass_boundary():
    '''Two singular values a relative 2e-7 apart straddle the first pass's reliable level: the tail
    starts at a wider gap, and both come out to the SVD's accuracy.'''
    from tentorch_amd.gram import GramSplit
    rng = np.random.default_rng(5)
    s = np.geomspace(1, 1e-16, 40)
    s[10], s[11] = 1.0000001e-5, 0.9999999e-5
    u, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    v, _ = np.linalg.qr(rng.standard_normal((55, 55)))
    block = torch.from_numpy((u * s) @ v[:, :40].T)
    with emulated_split():
        for left in (True, False):
            values = GramSplit([block], left).factor(lambda values: 0.0, SvdSettings.reliable, SvdSettings.resolved)[0]
            exact = np.linalg.svd(block.numpy(), compute_uv=False)
            assert np.abs(values - exact).max() <= 1e-15
