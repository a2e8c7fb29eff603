'''
The eigensolver's entry point that allocates nothing (`tt_eigh_batched_static`, its workspace a
static array the library holds per device) against the workspace form (`tt_eigh_batched`) through
the C ABI: the same eigenvalues and eigenvectors bit for bit on two calls in a row, the second with a
smaller batch (the scratch keeps the first call's contents), and a batch beyond the scratch refused
with an error before anything runs.
'''

import ctypes

import numpy as np
import pytest
import torch

import cabi

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 8 << 20


class EighProblem(ctypes.Structure):
    _fields_ = [('matrix_offset', ctypes.c_int64), ('value_offset', ctypes.c_int64), ('n', ctypes.c_int64)]


@pytest.fixture(scope='module')
def lib():
    lib = cabi.load()
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.tt_eigh_workspace.argtypes = [ctypes.POINTER(EighProblem), i64]
    lib.tt_eigh_workspace.restype = i64
    lib.tt_eigh_batched.argtypes = [ctypes.POINTER(EighProblem), i64, vp, vp, vp, vp, vp]
    lib.tt_eigh_batched_static.argtypes = [ctypes.POINTER(EighProblem), i64, vp, vp, vp, vp]
    return lib


def _problems(sizes):
    table = (EighProblem * len(sizes))()
    at = vat = 0
    for i, n in enumerate(sizes):
        table[i] = EighProblem(at, vat, n)
        at += n * n
        vat += n
    return table, at, vat


def _matrices(rng, sizes):
    parts = []
    for n in sizes:
        a = rng.standard_normal((n, n))
        parts.append((a + a.T).reshape(-1))
    return torch.tensor(np.concatenate(parts), device='cuda')


def _solve(lib, sizes, packed, static):
    table, elements, count = _problems(sizes)
    matrices = packed.clone()          # the solver overwrites its input with the reflectors
    values = torch.full((count,), float('nan'), dtype=torch.float64, device='cuda')
    vectors = torch.full((elements,), float('nan'), dtype=torch.float64, device='cuda')
    stream = torch.cuda.current_stream().cuda_stream
    if static:
        rc = lib.tt_eigh_batched_static(table, len(sizes), matrices.data_ptr(), values.data_ptr(),
                                        vectors.data_ptr(), stream)
    else:
        need = lib.tt_eigh_workspace(table, len(sizes))
        work = torch.full((need,), float('nan'), dtype=torch.float64, device='cuda')
        rc = lib.tt_eigh_batched(table, len(sizes), matrices.data_ptr(), values.data_ptr(), vectors.data_ptr(),
                                 work.data_ptr(), stream)
    assert rc == 0, cabi.error(lib)
    torch.cuda.synchronize()
    return values, vectors


def test_static_scratch_eigensolver_matches_the_workspace_form(lib):
    rng = np.random.default_rng(41)
    for sizes in ([300, 17, 129, 64, 1], [40, 5, 3]):       # the second call smaller than the first
        table, _, _ = _problems(sizes)
        assert lib.tt_eigh_workspace(table, len(sizes)) <= SCRATCH_DOUBLES
        packed = _matrices(rng, sizes)
        ref_values, ref_vectors = _solve(lib, sizes, packed, static=False)
        values, vectors = _solve(lib, sizes, packed, static=True)
        assert torch.equal(values, ref_values), sizes
        assert torch.equal(vectors, ref_vectors), sizes
        # and the solution itself: eigenvalues against numpy
        at = vat = 0
        for n in sizes:
            m = packed[at:at + n * n].view(n, n).cpu().numpy()
            ref = np.linalg.eigvalsh(m)
            assert np.abs(values[vat:vat + n].cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
            at += n * n
            vat += n


def test_a_batch_beyond_the_static_scratch_is_refused(lib):
    sizes = [2048, 2048, 2048]
    table, _, _ = _problems(sizes)
    assert lib.tt_eigh_workspace(table, len(sizes)) > SCRATCH_DOUBLES
    # refused while validating: no buffer is touched, so null pointers are never dereferenced
    rc = lib.tt_eigh_batched_static(table, len(sizes), None, None, None, None)
    assert rc != 0
    assert 'static scratch' in cabi.error(lib)
