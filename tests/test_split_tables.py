'''
The host checks every batched split launch passes before it reaches the device
(`tentorch_amd._native.check_bgemm`, `check_rows`, `check_lincomb`): a table whose operands reach
one element past their buffer, whose work items are misnumbered or whose row strides are shorter
than their rows is refused, and an exact fit is accepted with the right item count.  CPU tensors
stand in for device buffers (only their addresses are read).
'''

import numpy as np
import pytest
import torch

from tentorch_amd import _native
from tentorch_amd.gram import _bgemm as _bgemm_fields, _rowwise


def _buffers(*sizes):
    tensors = [torch.zeros(n, dtype=torch.float64) for n in sizes]
    extents = _native.Extents()
    for t in tensors:
        extents.add(t)
    return tensors, extents


GEMM = ('a', 'b', 'c', 'm', 'n', 'k', 'lda', 'ldb', 'ldc', 'trans_a', 'trans_b', 'alpha', 'beta', 'col_scale',
        'row_shift')
ROWS = ('src', 'dst', 'rows', 'cols', 'src_ld', 'dst_ld', 'transpose', 'index')
LINCOMB = ('x', 'y', 'dst', 'rows', 'cols', 'x_ld', 'y_ld', 'dst_ld', 'alpha', 'beta', 'gamma')


def _gemm(a, b, c, m, n, k, lda, ldb, ldc, trans_a=0, trans_b=0, alpha=1.0, beta=0.0, scale=0, shift=0):
    return (a, b, c, m, n, k, lda, ldb, ldc, trans_a, trans_b, alpha, beta, scale, shift)


def _columns(names, problems):
    return {name: [p[i] for p in problems] for i, name in enumerate(names)}


def _bgemm(problems):
    return _bgemm_fields(len(problems), **_columns(GEMM, problems))[0]


def _rows_table(*problems):
    padded = [p + (0,) * (len(ROWS) - len(p)) for p in problems]
    return _rowwise(_native.ROWS_PROBLEM, len(padded), **_columns(ROWS, padded))[0]


def _lincomb_table(*problems):
    return _rowwise(_native.LINCOMB_PROBLEM, len(problems), **_columns(LINCOMB, problems))[0]


def test_bgemm_table_exact_fit_and_one_past():
    (a, b, c), extents = _buffers(70 * 33, 33 * 129, 70 * 129)
    exact = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 70, 129, 33, 33, 129, 129)])
    assert _native.check_bgemm(exact, extents) == 2 * 3
    # transposed A stored 33 x 70 fills the same buffer
    turned = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 70, 129, 33, 70, 129, 129, trans_a=1)])
    assert _native.check_bgemm(turned, extents) == 6
    over = _bgemm([_gemm(a.data_ptr() + 8, b.data_ptr(), c.data_ptr(), 70, 129, 33, 33, 129, 129)])
    with pytest.raises(RuntimeError, match='bgemm A'):
        _native.check_bgemm(over, extents)
    wide = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 70, 130, 33, 33, 129, 129)])
    with pytest.raises(RuntimeError, match='row stride'):
        _native.check_bgemm(wide, extents)


def test_bgemm_scales_shifts_and_numbering():
    (a, b, c, s, r), extents = _buffers(16, 16, 16, 4, 3)
    ok = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 4, 4, 4, 4, 4, 4, scale=s.data_ptr()),
                 _gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 3, 4, 4, 4, 4, 4, scale=s.data_ptr(),
                       shift=r.data_ptr())])
    assert _native.check_bgemm(ok, extents) == 2
    no_scale = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 3, 4, 4, 4, 4, 4, shift=r.data_ptr())])
    with pytest.raises(RuntimeError, match='column shifts'):
        _native.check_bgemm(no_scale, extents)
    long_shift = _bgemm([_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), 4, 4, 4, 4, 4, 4, scale=s.data_ptr(),
                               shift=r.data_ptr())])
    with pytest.raises(RuntimeError, match='row shift'):
        _native.check_bgemm(long_shift, extents)
    renumbered = ok.copy()
    renumbered['first'][1] = 5
    with pytest.raises(RuntimeError, match='numbered'):
        _native.check_bgemm(renumbered, extents)


def test_overlapping_views_are_one_buffer_and_adjacent_buffers_are_not():
    base = torch.zeros(100, dtype=torch.float64)
    extents = _native.Extents().add(base[:60]).add(base[40:])
    extents.check(base.data_ptr(), 0, 8 * 99, 'whole')             # the union of two overlapping views
    first, second = torch.zeros(64, dtype=torch.float64), torch.zeros(64, dtype=torch.float64)
    apart = _native.Extents().add(first).add(second)
    lo, hi = sorted([first.data_ptr(), second.data_ptr()])
    with pytest.raises(RuntimeError):
        apart.check(lo, 0, 8 * 64, 'one past the first buffer')


def test_row_and_lincomb_tables():
    (src, dst, norms), extents = _buffers(10 * 12, 10 * 12, 10)
    index = torch.tensor([3, 1, 4], dtype=torch.int32)
    extents.add(index)
    rows = _rows_table((src.data_ptr(), norms.data_ptr(), 10, 11, 12, 1))
    assert _native.check_rows(rows, extents, vector_dst=True) == 10
    turned = _rows_table((src.data_ptr(), dst.data_ptr(), 3, 11, 12, 10, 1, index.data_ptr()))
    assert _native.check_rows(turned, extents, source_rows=[10]) == 3
    with pytest.raises(RuntimeError, match='row source'):
        _native.check_rows(turned, extents, source_rows=[11])
    comb = _lincomb_table((src.data_ptr(), dst.data_ptr(), dst.data_ptr(), 10, 12, 12, 12, 12, 1.0, 2.0, 0.0))
    assert _native.check_lincomb(comb, extents) == 10
    past = _lincomb_table((src.data_ptr(), 0, dst.data_ptr() + 8, 10, 12, 12, 0, 12, 1.0, 0.0, 0.0))
    with pytest.raises(RuntimeError, match='lincomb destination'):
        _native.check_lincomb(past, extents)


def test_tridiagonal_step_launches_and_bytes():
    '''The step launches tt_eigh_batched makes (one per step below the batch's largest order minus
    192, per batch of 64 matrices) and their algorithmic bytes, 8 (2 m^2 + 6 m) per matrix with
    m > 0 trailing columns.'''
    assert _native.tridiagonal_steps([192, 100, 5]) == (0, 0)
    launches, moved = _native.tridiagonal_steps([194, 193, 3])
    per = lambda m: 8 * (2 * m * m + 6 * m)
    assert launches == 2
    assert moved == per(193) + per(192) + per(2) + per(192) + per(191) + per(1)
    # 65 matrices make two batches: the second (one matrix of order 195) steps three times
    launches, _ = _native.tridiagonal_steps([264] + [10] * 63 + [195])
    assert launches == 72 + 3


def test_the_engine_chooses_the_same_tails_as_the_rule():
    '''The split's further passes are decided by the engine's port of `gram.choose_tails`: on random
    descending spectra (exact zeros, ties, passes already started, caps anywhere) both pick the same
    blocks and the same tail starts.'''
    import numpy as np
    from tentorch_amd import gram
    from tentorch_amd._native import engine
    rng = np.random.default_rng(1)
    for trial in range(2000):
        norms, ranks, first = [], [], []
        for _ in range(int(rng.integers(1, 12))):
            g = int(rng.integers(0, 40))
            v = np.sort(np.exp(-rng.uniform(0, 37, g)))[::-1] * rng.uniform(0.1, 2)
            if rng.random() < 0.2 and g:
                v[rng.integers(0, g):] = 0.0
            if rng.random() < 0.1 and g > 3:
                v[1] = v[2]
            r = int(rng.integers(0, g + 1))
            norms.append(v)
            ranks.append(r)
            first.append(int(rng.integers(0, r + 1)) if rng.random() < 0.3 else 0)
        floor = float(np.exp(-rng.uniform(0, 37)))
        want = gram.choose_tails(norms, ranks, first, floor, 1e-5, 1e-7)
        at = np.concatenate([[0], np.cumsum([v.size for v in norms])]).astype(np.int64)
        got = engine.choose_tails(np.concatenate(norms), at, np.array(ranks, dtype=np.int64),
                                  np.array(first, dtype=np.int64), floor, 1e-5, 1e-7)
        assert [tuple(x) for x in got] == want, trial
