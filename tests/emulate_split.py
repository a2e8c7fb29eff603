'''
The split's device launches interpreted on the CPU: `tt_bgemm`, `tt_row_norms`, `tt_copy_rows`,
`tt_lincomb_rows` and the batched eigensolver, run by numpy on the host buffers their problem tables
address (each table read from where it was staged).  With the tables pointing at CPU tensors, `tentorch_amd.gram.GramSplit` runs end to end without a GPU, so its
table construction, its passes and its bounds checks are tested on the CPU; the kernels themselves
are tested on the GPU against this reading of the header's semantics.
'''

import contextlib
import ctypes

import numpy as np
import torch

from tentorch_amd import _native


def _doubles(address: int, count: int) -> np.ndarray:
    return np.ctypeslib.as_array((ctypes.c_double * max(count, 1)).from_address(address))[:count]


def _int32s(address: int, count: int) -> np.ndarray:
    return np.ctypeslib.as_array((ctypes.c_int32 * max(count, 1)).from_address(address))[:count]


def _matrix(address: int, rows: int, cols: int, ld: int) -> np.ndarray:
    '''A rows x cols row-major view with row stride ld.'''
    flat = _doubles(address, (rows - 1) * ld + cols if rows and cols else 0)
    return np.lib.stride_tricks.as_strided(flat, shape=(rows, cols), strides=(8 * ld, 8)) if rows and cols \
        else np.zeros((rows, cols))


def _staged(address: int, table: np.ndarray) -> np.ndarray:
    '''The table as the launch reads it: len(table) records at `address` (the staged copy).'''
    if not len(table):
        return table
    raw = (ctypes.c_char * (len(table) * table.dtype.itemsize)).from_address(address)
    staged = np.frombuffer(raw, dtype=table.dtype).copy()
    assert staged.tobytes() == np.ascontiguousarray(table).tobytes(), 'staged table differs from the host table'
    return staged


@contextlib.contextmanager
def emulated_split():
    names = ('bgemm', 'row_norms', 'copy_rows', 'lincomb_rows', 'eigh_batched')
    saved = {name: getattr(_native, name) for name in names}

    def bgemm(table, address, tiles):
        for p in _staged(address, table):
            m, n, k = int(p['m']), int(p['n']), int(p['k'])
            if not (m and n):
                continue
            a = _matrix(int(p['a']), *((k, m) if p['trans_a'] else (m, k)), int(p['lda'])) if k else np.zeros((m, 0))
            b = _matrix(int(p['b']), *((n, k) if p['trans_b'] else (k, n)), int(p['ldb'])) if k else np.zeros((0, n))
            product = (a.T if p['trans_a'] else a) @ (b.T if p['trans_b'] else b)
            if int(p['row_shift']):
                product = product / (_doubles(int(p['col_scale']), n)[None, :] - _doubles(int(p['row_shift']), m)[:, None])
            elif int(p['col_scale']):
                product = product * _doubles(int(p['col_scale']), n)
            c = _matrix(int(p['c']), m, n, int(p['ldc']))
            c[...] = p['alpha'] * product + (p['beta'] * c if p['beta'] != 0.0 else 0.0)

    def row_norms(table, address, rows):
        for p in _staged(address, table):
            r, c = int(p['rows']), int(p['cols'])
            if r and c:
                _doubles(int(p['dst']), r)[...] = np.linalg.norm(_matrix(int(p['src']), r, c, int(p['src_ld'])), axis=1)

    def copy_rows(table, address, rows):
        for p in _staged(address, table):
            r, c = int(p['rows']), int(p['cols'])
            if not (r and c):
                continue
            index = _int32s(int(p['index']), r) if int(p['index']) else np.arange(r)
            held = int(index.max()) + 1
            source = _matrix(int(p['src']), held, c, int(p['src_ld']))[index]
            if p['transpose']:
                _matrix(int(p['dst']), c, r, int(p['dst_ld']))[...] = source.T
            else:
                _matrix(int(p['dst']), r, c, int(p['dst_ld']))[...] = source

    def lincomb_rows(table, address, rows):
        for p in _staged(address, table):
            r, c = int(p['rows']), int(p['cols'])
            if not (r and c):
                continue
            value = p['alpha'] * _matrix(int(p['x']), r, c, int(p['x_ld']))
            if int(p['y']):
                value = value + p['beta'] * _matrix(int(p['y']), r, c, int(p['y_ld']))
            value = value + p['gamma'] * np.eye(r, c)
            _matrix(int(p['dst']), r, c, int(p['dst_ld']))[...] = value

    def eigh_batched(matrices, sizes, offsets=None, work=None, vectors=None):
        values = torch.empty(max(sum(sizes), 1), dtype=torch.float64)
        vectors = torch.empty_like(matrices) if vectors is None else vectors
        at = 0
        for n, offset in zip(sizes, offsets):
            w, v = np.linalg.eigh(matrices[offset:offset + n * n].view(n, n).numpy())
            values[at:at + n] = torch.from_numpy(w)
            vectors[offset:offset + n * n] = torch.from_numpy(np.ascontiguousarray(v.T)).reshape(-1)
            at += n
        return values[:at], vectors

    replacements = dict(bgemm=bgemm, row_norms=row_norms, copy_rows=copy_rows,
                        lincomb_rows=lincomb_rows, eigh_batched=eigh_batched)
    for name, function in replacements.items():
        setattr(_native, name, function)
    try:
        yield
    finally:
        for name, function in saved.items():
            setattr(_native, name, function)
