'''
Loader of the two native libraries: the host engine (`_engine`, planning) and the device library
(`libtentorch_hip.so`, the C ABI of `include/tentorch_amd.h`).  There is no fallback: when either
library is missing, or no GPU is visible when device work is asked for, this module raises.
'''

import ctypes
import os

import numpy as np
from typing import Optional

import torch

HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _engine as engine  # noqa: F401
except ImportError as error:  # pragma: no cover - exercised only on a broken install
    raise ImportError('tentorch_amd: the host engine is not built; run `python tentorch_amd/build.py` '
                      f'({error})') from error

_lib = ctypes.CDLL(os.path.join(HERE, 'libtentorch_hip.so'))

TT_F64, TT_F32, TT_C128 = 0, 1, 2

_c_void_p, _i64, _int, _dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
_lib.tt_dot.argtypes = [_int, _i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_axpby.argtypes = [_int, _i64, _dbl, _c_void_p, _c_void_p, _dbl, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_convert.argtypes = [_int, _int, _i64, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
_lib.tt_version.argtypes = []
_lib.tt_launch_count.argtypes = []
_lib.tt_mlp_param_count.argtypes = [_int, _int]
_lib.tt_mlp_param_count.restype = ctypes.c_int64
_lib.tt_mlp_forward.argtypes = [_int, _int, _i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_mlp_gradient.argtypes = [_int, _int, _i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_mlp_adam.argtypes = [_i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, ctypes.c_float, _int, _c_void_p]
_lib.tt_ggemm_num_configs.argtypes = []


class EighProblem(ctypes.Structure):
    _fields_ = [('matrix_offset', ctypes.c_int64), ('value_offset', ctypes.c_int64), ('n', ctypes.c_int64)]


_lib.tt_eigh_max_order.argtypes = []
_lib.tt_eigh_max_order.restype = ctypes.c_int64
EIGH_MAX_ORDER = int(_lib.tt_eigh_max_order())
_lib.tt_eigh_workspace.argtypes = [ctypes.POINTER(EighProblem), _i64]
_lib.tt_eigh_workspace.restype = ctypes.c_int64
_lib.tt_eigh_batched.argtypes = [ctypes.POINTER(EighProblem), _i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                 _c_void_p]
_lib.tt_timing_enable.argtypes = [_int]
_lib.tt_timing_collect.argtypes = [ctypes.POINTER(_dbl), _i64, ctypes.POINTER(_i64)]
_lib.tt_launch_count.restype = ctypes.c_int64
_lib.tt_multidot.argtypes = [_i64, _int, _c_void_p, _i64, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_multiaxpy.argtypes = [_i64, _int, _c_void_p, _dbl, _c_void_p, _i64, _c_void_p, _c_void_p]
_lib.tt_multiaxpy_counted.argtypes = [_i64, _int, _c_void_p, _c_void_p, _dbl, _c_void_p, _i64, _c_void_p, _c_void_p]
_lib.tt_lanczos_ritz.argtypes = [_int, _c_void_p, _c_void_p, _dbl, _c_void_p, _c_void_p, _c_void_p, _c_void_p]
_lib.tt_lanczos_orthonormalize.argtypes = [_i64, _int, _c_void_p, _i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                           _c_void_p]
_lib.tt_bgemm.argtypes = [_c_void_p, _i64, _i64, _c_void_p]
_lib.tt_row_norms.argtypes = [_c_void_p, _i64, _i64, _c_void_p]
_lib.tt_copy_rows.argtypes = [_c_void_p, _i64, _i64, _c_void_p]
_lib.tt_lincomb_rows.argtypes = [_c_void_p, _i64, _i64, _c_void_p]
_lib.tt_qr_batched.argtypes = [_c_void_p, _i64, _c_void_p]
_lib.tt_qr_check.argtypes = [_c_void_p, _i64, _c_void_p, _c_void_p, _i64]
_lib.tt_bgemm_check.argtypes = [_c_void_p, _i64, _c_void_p, _c_void_p, _i64, ctypes.POINTER(_i64)]
_lib.tt_rows_check.argtypes = [_c_void_p, _i64, _c_void_p, _int, _c_void_p, _c_void_p, _i64, ctypes.POINTER(_i64)]
_lib.tt_lincomb_check.argtypes = [_c_void_p, _i64, _c_void_p, _c_void_p, _i64, ctypes.POINTER(_i64)]
_lib.tt_graph_begin.argtypes = [_c_void_p]
_lib.tt_graph_end.argtypes = [_c_void_p, ctypes.POINTER(_c_void_p)]
_lib.tt_graph_launch.argtypes = [_c_void_p, _c_void_p]
_lib.tt_graph_destroy.argtypes = [_c_void_p]

EXPORTED = ['tt_gss_plan_create', 'tt_gss_plan_destroy', 'tt_gss_plan_counts', 'tt_gss_execute',
            'tt_gss_execute_plain', 'tt_gss_execute_typed', 'tt_ggemm_plan_create', 'tt_ggemm_plan_destroy', 'tt_ggemm_plan_counts',
            'tt_ggemm_execute', 'tt_ggemm_execute_plain', 'tt_ggemm_plan_create_config', 'tt_ggemm_num_configs',
            'tt_ggemm_config_shape', 'tt_ggemm_plan_config', 'tt_dot', 'tt_axpby', 'tt_convert', 'tt_multidot',
            'tt_multiaxpy', 'tt_graph_begin',
            'tt_graph_end', 'tt_graph_launch', 'tt_graph_destroy', 'tt_last_error', 'tt_version',
            'tt_launch_count', 'tt_timing_enable', 'tt_timing_collect', 'tt_mlp_param_count', 'tt_mlp_forward',
            'tt_mlp_gradient', 'tt_mlp_adam', 'tt_eigh_max_order', 'tt_eigh_workspace', 'tt_eigh_batched',
            'tt_lanczos_orthonormalize', 'tt_lanczos_ritz', 'tt_multiaxpy_counted', 'tt_eigh_batched_static',
            'tt_bgemm', 'tt_row_norms', 'tt_copy_rows', 'tt_lincomb_rows', 'tt_bgemm_check', 'tt_rows_check',
            'tt_lincomb_check']


def library() -> ctypes.CDLL:
    return _lib


def launch_count() -> int:
    '''Leg-move and block-product kernel launches issued so far.'''
    return int(_lib.tt_launch_count())


def timing(on: bool) -> None:
    '''Attach kernel-dispatch events to every leg-move and block-product launch of this thread.'''
    check(_lib.tt_timing_enable(1 if on else 0), 'timing')


def timing_collect() -> list:
    '''Durations (ms) of the launches timed since the last collection, in launch order.'''
    count = _i64(0)
    check(_lib.tt_timing_collect(None, 0, ctypes.byref(count)), 'timing')
    buffer = (_dbl * max(1, count.value))()
    check(_lib.tt_timing_collect(buffer, count.value, ctypes.byref(count)), 'timing')
    return list(buffer[:count.value])


def last_error() -> str:
    buffer = ctypes.create_string_buffer(1024)
    _lib.tt_last_error(buffer, 1024)
    return buffer.value.decode(errors='replace')


def check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f'{what}: {last_error()}')


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return TT_F64
    if dtype == torch.float32:
        return TT_F32
    if dtype == torch.complex128:
        return TT_C128
    raise TypeError(f'the HIP backend does not run tensors of {dtype}')


def require_device(t: torch.Tensor) -> None:
    if t.device.type != 'cuda':
        raise RuntimeError('the HIP backend runs on the GPU only: this tensor lives on '
                           f'{t.device}, and there is no CPU path')


def stream() -> int:
    '''The current HIP stream of the current device, as a raw handle.'''
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


def dot(dtype: torch.dtype, x: torch.Tensor, y: torch.Tensor, out: torch.Tensor) -> None:
    '''out[0] (out[0:2] for complex) = sum conj(x) y, fixed summation order, on the device.'''
    check(_lib.tt_dot(dtype_code(dtype), x.numel(), x.data_ptr(), y.data_ptr(), out.data_ptr(), stream()), 'dot')


def axpby(dtype: torch.dtype, alpha, x, beta, y: torch.Tensor, alpha_dev=None, beta_dev=None) -> None:
    '''y = alpha x + beta y on the device (alpha/beta from device scalars when given).'''
    code = dtype_code(dtype)
    n = y.numel()
    if code == TT_C128:
        if alpha_dev is not None or beta_dev is not None or isinstance(alpha, complex) or isinstance(beta, complex):
            raise TypeError('complex scaling is not a real axpby')
        code, n = TT_F64, 2 * n
    check(_lib.tt_axpby(code, n, float(alpha), None if alpha_dev is None else alpha_dev.data_ptr(),
                        None if x is None else x.data_ptr(), float(beta),
                        None if beta_dev is None else beta_dev.data_ptr(), y.data_ptr(), stream()), 'axpby')


def multidot(V: torch.Tensor, k: int, w: torch.Tensor, out: torch.Tensor) -> None:
    '''out[:k] = V[:k] @ w (FP64 Krylov rows, fixed summation order, on the device).'''
    check(_lib.tt_multidot(w.numel(), k, V.data_ptr(), V.stride(0), w.data_ptr(), out.data_ptr(), stream()),
          'multidot')


def multiaxpy(V: torch.Tensor, k: int, coef: torch.Tensor, scale: float, w: torch.Tensor,
              count: Optional[torch.Tensor] = None) -> None:
    '''w += scale * coef[:k] @ V[:k] with the coefficients on the device; with `count` (a device
    int32), only the first min(k, count) rows of V are read.'''
    if count is None:
        check(_lib.tt_multiaxpy(w.numel(), k, coef.data_ptr(), float(scale), V.data_ptr(), V.stride(0), w.data_ptr(),
                                stream()), 'multiaxpy')
        return
    check(_lib.tt_multiaxpy_counted(w.numel(), k, count.data_ptr(), coef.data_ptr(), float(scale), V.data_ptr(),
                                    V.stride(0), w.data_ptr(), stream()), 'multiaxpy')


def lanczos_ritz(k: int, alphas: torch.Tensor, betas: torch.Tensor, tol: float, coef: torch.Tensor,
                 ritz: torch.Tensor, size: torch.Tensor) -> None:
    '''The Ritz pair of a Lanczos run from its device scalars, on the device (tt_lanczos_ritz).'''
    check(_lib.tt_lanczos_ritz(k, alphas.data_ptr(), betas.data_ptr(), float(tol), coef.data_ptr(), ritz.data_ptr(),
                               size.data_ptr(), stream()), 'lanczos ritz')


def lanczos_orthonormalize(V: torch.Tensor, k: int, w: torch.Tensor, coef: torch.Tensor, alpha: torch.Tensor,
                           beta: torch.Tensor) -> None:
    '''w (the image of V[k]) orthogonalised against V[:k+1] by two Gram-Schmidt passes; alpha and beta
    (device scalars) get its projection on V[k] and the norm left; V[k+1] = w / beta.  `coef` holds
    2 (k + 1) + 1 doubles of scratch.'''
    if coef.numel() < 2 * (k + 1) + 1 or V.shape[0] < k + 2:
        raise ValueError('lanczos_orthonormalize: scratch or basis too small')
    check(_lib.tt_lanczos_orthonormalize(w.numel(), k, V.data_ptr(), V.stride(0), w.data_ptr(), coef.data_ptr(),
                                         alpha.data_ptr(), beta.data_ptr(), stream()), 'lanczos orthonormalize')


def eigh_workspace(sizes) -> int:
    '''Doubles of workspace tt_eigh_batched needs for matrices of these orders.'''
    problems = (EighProblem * max(len(sizes), 1))()
    for b, n in enumerate(sizes):
        problems[b] = EighProblem(0, 0, n)
    size = _lib.tt_eigh_workspace(problems, len(sizes))
    if size < 0:
        raise RuntimeError(last_error())
    return size


# Called with the matrix orders of every tt_eigh_batched call while set (by a profiler); the
# call's tridiagonalisation step launches are then timed by events attached to their dispatch.
eigh_observer = None

# orders up to this are tridiagonalised in one launch per batch; larger ones step by step above it
EIGH_STEPPED_ABOVE = 192
EIGH_BATCH = 64


def tridiagonal_steps(sizes):
    '''(launches, algorithmic bytes) of the step-by-step tridiagonalisation tt_eigh_batched runs
    for matrices of these orders: per batch of 64, one launch for each step k below the batch's
    largest order minus 192.  Step k of a matrix with m = n - k - 1 > 0 trailing columns reads and
    writes its trailing block once and its vectors (column k, v and p of the previous step, v and p
    of this one): 8 (2 m^2 + 6 m) bytes.'''
    launches, moved = 0, 0
    for first in range(0, len(sizes), EIGH_BATCH):
        batch = np.asarray(sizes[first:first + EIGH_BATCH], dtype=np.int64)
        steps = max(0, int(batch.max()) - EIGH_STEPPED_ABOVE)
        launches += steps
        for k in range(steps):
            m = batch - k - 1
            m = m[m > 0]
            moved += int(8 * np.sum(2 * m * m + 6 * m))
    return launches, moved


def eigh_batched(matrices: torch.Tensor, sizes, offsets=None, work: Optional[torch.Tensor] = None,
                 vectors: Optional[torch.Tensor] = None):
    '''Eigenvalues (ascending) and eigenvectors of the symmetric FP64 matrices packed in `matrices`
    (matrix b is sizes[b] x sizes[b], at offsets[b], default back to back), on the device.  Returns
    (values, vectors): values packed back to back, vectors with eigenvector j of matrix b as row j
    of its n x n at the matrix's offset.  `matrices` is overwritten; `work` may supply the workspace
    (any contents) and `vectors` the eigenvector buffer.'''
    if matrices.dtype != torch.float64:
        raise TypeError('the batched eigensolver runs in float64')
    require_device(matrices)
    count = len(sizes)
    problems = (EighProblem * max(count, 1))()
    at, value_at = 0, 0
    for b, n in enumerate(sizes):
        offset = at if offsets is None else offsets[b]
        problems[b] = EighProblem(offset, value_at, n)
        at = offset + n * n
        value_at += n
    if offsets is None and at > matrices.numel():
        raise ValueError('the matrices do not fit the buffer')
    values = torch.empty(max(value_at, 1), dtype=torch.float64, device=matrices.device)
    if vectors is None:
        vectors = torch.empty_like(matrices)
    elif vectors.numel() < matrices.numel() or vectors.dtype != torch.float64:
        raise ValueError('the eigenvector buffer has to be as large as the matrix buffer')
    size = _lib.tt_eigh_workspace(problems, count)
    if size < 0:
        raise RuntimeError(last_error())
    workspace = torch.empty(max(size, 1), dtype=torch.float64, device=matrices.device) if work is None else work
    if workspace.numel() < size or workspace.dtype != torch.float64:
        raise ValueError(f'the eigensolver workspace holds {workspace.numel()} doubles, {size} needed')
    observer = eigh_observer
    if observer is not None:
        timing(True)
    try:
        check(_lib.tt_eigh_batched(problems, count, matrices.data_ptr(), values.data_ptr(), vectors.data_ptr(),
                                   workspace.data_ptr(), stream()), 'eigh')
    finally:
        if observer is not None:
            timing(False)
    if observer is not None:
        observer(list(sizes))
    return values[:value_at], vectors


def convert(x: torch.Tensor, y: torch.Tensor) -> None:
    check(_lib.tt_convert(dtype_code(x.dtype), dtype_code(y.dtype), x.numel(), x.data_ptr(), y.data_ptr(),
                          stream()), 'convert')


class Graph:
    '''A captured launch sequence (hipGraph) replayed on the current stream.'''

    def __init__(self, handle):
        self._handle = handle

    def launch(self) -> None:
        check(_lib.tt_graph_launch(self._handle, stream()), 'graph launch')

    def __del__(self):
        if getattr(self, '_handle', None):
            _lib.tt_graph_destroy(self._handle)
            self._handle = None


def capture(fn) -> Graph:
    '''Run `fn` under stream capture and return the instantiated graph.  Capture needs a stream of
    its own (`torch.cuda.stream(torch.cuda.Stream())`): the legacy default stream cannot be captured.
    Everything `fn` launches must already have its device tables uploaded (run it once first).'''
    if torch.cuda.current_stream() == torch.cuda.default_stream():
        raise RuntimeError('graph capture needs a non-default current stream')
    s = stream()
    check(_lib.tt_graph_begin(s), 'graph capture')
    try:
        fn()
    finally:
        handle = ctypes.c_void_p()
        status = _lib.tt_graph_end(s, ctypes.byref(handle))
    check(status, 'graph capture')
    return Graph(handle)


# -- the split's batched kernels: problem tables built on the host, uploaded, then launched --------

BGEMM_PROBLEM = np.dtype([('a', '<u8'), ('b', '<u8'), ('c', '<u8'), ('col_scale', '<u8'), ('row_shift', '<u8'),
                          ('first', '<i8'),
                          ('m', '<i4'), ('n', '<i4'), ('k', '<i4'), ('lda', '<i4'), ('ldb', '<i4'), ('ldc', '<i4'),
                          ('trans_a', '<i4'), ('trans_b', '<i4'), ('alpha', '<f8'), ('beta', '<f8')])
ROWS_PROBLEM = np.dtype([('src', '<u8'), ('dst', '<u8'), ('index', '<u8'), ('first', '<i8'), ('rows', '<i4'),
                         ('cols', '<i4'), ('src_ld', '<i4'), ('dst_ld', '<i4'), ('transpose', '<i4'),
                         ('reserved', '<i4')])
LINCOMB_PROBLEM = np.dtype([('x', '<u8'), ('y', '<u8'), ('dst', '<u8'), ('first', '<i8'), ('rows', '<i4'),
                            ('cols', '<i4'), ('x_ld', '<i4'), ('y_ld', '<i4'), ('dst_ld', '<i4'), ('reserved', '<i4'),
                            ('alpha', '<f8'), ('beta', '<f8'), ('gamma', '<f8')])
assert BGEMM_PROBLEM.itemsize == 96 and ROWS_PROBLEM.itemsize == 56 and LINCOMB_PROBLEM.itemsize == 80
QR_PROBLEM = np.dtype([('src', '<u8'), ('q', '<u8'), ('r', '<u8'), ('work', '<u8'), ('m', '<i4'), ('n', '<i4'),
                       ('src_ld', '<i4'), ('q_ld', '<i4'), ('r_ld', '<i4'), ('trans', '<i4'), ('q_trans', '<i4'),
                       ('r_trans', '<i4')])
assert QR_PROBLEM.itemsize == 64


class Extents:
    '''Device buffers a batch of problems may touch, as (first address, one past the last) byte
    ranges: every operand a table names is checked to lie inside one of them before the launch.
    Ranges that overlap are views of one allocation and are checked as their union.'''

    def __init__(self):
        self.ranges = []
        self._spans = None

    def add(self, t: torch.Tensor) -> 'Extents':
        start = t.data_ptr()
        return self.add_range(start, start + t.numel() * t.element_size())

    def add_range(self, start: int, end: int) -> 'Extents':
        '''The bytes [start, end) of a device buffer.'''
        self.ranges.append((start, end))
        self._spans = None
        return self

    def spans(self):
        '''The ranges sorted by start, overlapping ones merged: (starts, ends) as int64 arrays.'''
        if self._spans is None:
            merged = []
            for start, end in sorted(self.ranges):
                if merged and start < merged[-1][1]:
                    merged[-1][1] = max(merged[-1][1], end)
                else:
                    merged.append([start, end])
            spans = np.array(merged, dtype=np.int64).reshape(-1, 2)
            self._spans = (spans[:, 0].copy(), spans[:, 1].copy())
        return self._spans

    def check_many(self, pointers, last, mask, what: str) -> None:
        '''Bytes 0..last[i] (inclusive) from pointers[i] lie in one known range, for every i where
        mask[i] holds.'''
        pointers = np.asarray(pointers, dtype=np.int64)[mask]
        if not pointers.size:
            return
        hi = pointers + np.asarray(last, dtype=np.int64)[mask]
        starts, ends = self.spans()
        at = np.searchsorted(starts, pointers, side='right') - 1
        inside = (at >= 0) & (hi < ends[np.maximum(at, 0)])
        if not inside.all():
            bad = int(np.argmin(inside))
            raise RuntimeError(f'{what}: bytes {int(pointers[bad]):#x}..{int(hi[bad]):#x} lie outside every '
                               'buffer of the launch')

    def check(self, pointer: int, first: int, last: int, what: str) -> None:
        '''Bytes first..last (inclusive) of an operand at `pointer` lie in one known buffer.'''
        self.check_many([pointer + first], [last - first], [True], what)


def _checked(function, table, extents, *extra):
    '''Runs one of the library's table checks; returns the table's work-item count.'''
    starts, ends = extents.spans()
    items = _i64(0)
    check(function(table.ctypes.data if table.size else None, len(table), *extra, starts.ctypes.data, ends.ctypes.data,
                   len(starts), ctypes.byref(items)), 'table check')
    return items.value


def check_bgemm(table: np.ndarray, extents: Extents) -> int:
    '''Validates a bgemm table against the buffers it names (`tt_bgemm_check`); returns its tile count.'''
    return _checked(_lib.tt_bgemm_check, table, extents)


def check_rows(table: np.ndarray, extents: Extents, source_rows=None, vector_dst: bool = False) -> int:
    '''Validates a row-problem table (`tt_rows_check`); returns its row count.  With an index,
    `source_rows[i]` is the number of rows problem i's source holds, and the caller has checked that
    every index is below it.  `vector_dst`: dst receives one value per row (the norms).'''
    held = None if source_rows is None else np.ascontiguousarray(source_rows, dtype=np.int64)
    return _checked(_lib.tt_rows_check, table, extents, None if held is None else held.ctypes.data,
                    1 if vector_dst else 0)


def check_lincomb(table: np.ndarray, extents: Extents) -> int:
    '''Validates a lincomb table (`tt_lincomb_check`); returns its row count.'''
    return _checked(_lib.tt_lincomb_check, table, extents)


def check_qr(table: np.ndarray, extents: Extents) -> None:
    '''Validates a QR table against the buffers it names (`tt_qr_check`).'''
    starts, ends = extents.spans()
    check(_lib.tt_qr_check(table.ctypes.data if table.size else None, len(table), starts.ctypes.data, ends.ctypes.data,
                           len(starts)), 'table check')


def qr_batched(table: np.ndarray, address: int) -> None:
    check(_lib.tt_qr_batched(address, len(table), stream()), 'batched QR')


def lincomb_rows(table: np.ndarray, address: int, rows: int) -> None:
    check(_lib.tt_lincomb_rows(address, len(table), rows, stream()), 'lincomb rows')


def bgemm(table: np.ndarray, address: int, tiles: int) -> None:
    check(_lib.tt_bgemm(address, len(table), tiles, stream()), 'bgemm')


def row_norms(table: np.ndarray, address: int, rows: int) -> None:
    check(_lib.tt_row_norms(address, len(table), rows, stream()), 'row norms')


def copy_rows(table: np.ndarray, address: int, rows: int) -> None:
    check(_lib.tt_copy_rows(address, len(table), rows, stream()), 'copy rows')
