'''
The block-sparse tensor of the MI355X backend and the operations of the contraction path.

A tensor is a structure (the host engine's interned legs, charge and block layout) plus ONE flat
device buffer holding every block back to back in key order, each block row-major.  The operations
mirror the reference's surface (`tentorch/_ext/bindings.cpp:249-414`, `tentorch/tensors/__init__.py`):
`Tensor.place`, `conj`, `clone`, the in-place arithmetic, `einsum(spec, *t, outcharge=)`,
`contract`, `inner`, `norm`, `is_close`.  Every leg move and block product runs through a program
the host engine plans once per shape and the device library executes (include/tentorch_amd.h).

Differences from the reference, by design:
- the default scalar type is float64 (the reference's categories default to complex128,
  `tentorch/symmetry/cat/abelian/abelian.cpp:47`); complex128 and float32 are supported throughout;
- data must live on the GPU for any compute: there is no CPU path.
'''

import itertools
from collections import OrderedDict, namedtuple
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native
from ._native import engine

BlockKey = namedtuple('BlockKey', ['codomain', 'domain', 'copy'])
BlockKey.__new__.__defaults__ = (0,)

DEFAULT_DTYPE = torch.float64


def default_device() -> torch.device:
    return torch.device('cuda') if torch.cuda.is_available() else torch.device('cpu')


class _Counter:
    value = 0


def _next_uid() -> int:
    _Counter.value += 1
    return _Counter.value


class Tensor:
    '''
    A map between two graded spaces: codomain legs above, domain legs below, total charge `qtotal`
    (`tentorch/tensors/tensor.h:43-216`).  Every allowed block exists and starts at zero
    (`tensor.cpp:23-61`).
    '''

    __slots__ = ('structure', 'data', '_uid', '_version', '__weakref__')

    def __init__(self, cat, codomain, domain, qtotal, dtype=None, device=None, *, structure=None, data=None):
        if structure is None:
            structure = engine.Structure(cat, list(codomain), list(domain), tuple(qtotal))
        self.structure = structure
        if data is None:
            dtype = DEFAULT_DTYPE if dtype is None else dtype
            device = default_device() if device is None else torch.device(device)
            data = torch.zeros(structure.size, dtype=dtype, device=device)
        self.data = data
        self._uid = _next_uid()
        self._version = 0

    @classmethod
    def _wrap(cls, structure, data) -> 'Tensor':
        t = cls.__new__(cls)
        t.structure, t.data = structure, data
        t._uid, t._version = _next_uid(), 0
        return t

    # -- structure
    @property
    def cat(self):
        return self.structure.cat

    @property
    def codomain(self) -> List:
        return self.structure.codomain

    @property
    def domain(self) -> List:
        return self.structure.domain

    @property
    def qtotal(self) -> tuple:
        return self.structure.qtotal

    @property
    def dtype(self) -> torch.dtype:
        return self.data.dtype

    @property
    def device(self) -> torch.device:
        return self.data.device

    def legs(self) -> List:
        return self.codomain + self.domain

    def nlegs(self) -> int:
        return self.structure.nlegs()

    def nblocks(self) -> int:
        return self.structure.nblocks()

    def keys(self) -> List[BlockKey]:
        return [BlockKey(*key) for key, _, _, _ in self.structure.blocks()]

    def blocks(self) -> Dict[BlockKey, torch.Tensor]:
        '''Every block as a (rows, cols) view into the flat buffer, in key order.'''
        return OrderedDict((BlockKey(*key), self.data[off:off + rows * cols].view(rows, cols))
                           for key, rows, cols, off in self.structure.blocks())

    def block(self, codomain_root, domain_root, copy: int = 0) -> torch.Tensor:
        at = self.structure.find(tuple(codomain_root), tuple(domain_root), copy)
        if at < 0:
            raise IndexError(f'the tensor has no block at {codomain_root} x {domain_root} of copy {copy}')
        key, rows, cols, off = self.structure.blocks()[at]
        return self.data[off:off + rows * cols].view(rows, cols)

    def to_tuple(self):
        '''`to_tuple` (`bindings.cpp:311-323`): (codomain, domain, qtotal, {key: block}).'''
        return (self.codomain, self.domain, self.qtotal, dict(self.blocks()))

    @staticmethod
    def from_tuple(cat, codomain, domain, qtotal, blocks=None, dtype=None, device=None) -> 'Tensor':
        '''`from_tuple` (`bindings.cpp:131-152`): the given blocks, every other block zero.'''
        out = Tensor(cat, codomain, domain, qtotal, dtype, device)
        if blocks:
            host = torch.zeros(out.structure.size, dtype=out.dtype)
            layout = {BlockKey(*k): (r, c, o) for k, r, c, o in out.structure.blocks()}
            for key, value in blocks.items():
                key = BlockKey(*key)
                if key not in layout:
                    raise IndexError(f'the tensor has no block at {key}')
                rows, cols, off = layout[key]
                value = torch.as_tensor(np.asarray(value) if not isinstance(value, torch.Tensor) else value)
                if tuple(value.shape) != (rows, cols):
                    raise ValueError(f'the block at {key} has shape {tuple(value.shape)}, not {(rows, cols)}')
                value = value.detach().cpu()
                if not out.dtype.is_complex and value.is_complex():
                    value = value.real
                host[off:off + rows * cols] = value.reshape(-1).to(out.dtype)
            out.data.copy_(host)
        return out

    @staticmethod
    def random(cat, codomain, domain, qtotal, dtype=None, device=None, generator=None) -> 'Tensor':
        '''`Tensor.random` (`bindings.cpp:157-164`): N(0, 1) entries block by block in key order, drawn
        on the host (numpy `generator` when given, else torch's global CPU generator) and uploaded.'''
        out = Tensor(cat, codomain, domain, qtotal, dtype, device)
        host = np.zeros(out.structure.size, dtype=np.complex128 if out.dtype.is_complex else np.float64)
        for key, rows, cols, off in out.structure.blocks():
            n = rows * cols
            if generator is not None:
                if out.dtype.is_complex:
                    host[off:off + n] = (generator.standard_normal(n) + 1j * generator.standard_normal(n)) / np.sqrt(2)
                else:
                    host[off:off + n] = generator.standard_normal(n)
            else:
                sample = torch.randn(n, dtype=torch.complex128 if out.dtype.is_complex else torch.float64)
                host[off:off + n] = sample.numpy()
        out.data.copy_(torch.from_numpy(host).to(out.dtype))
        return out

    # -- moves
    def place(self, above: Sequence[int], below: Optional[Sequence[int]] = None) -> 'Tensor':
        '''`Tensor::place` (`tentorch/tensors/place.cpp:183-247`), as one composed move.'''
        program = engine.plan_place(self.structure, list(above), None if below is None else list(below))
        return run(program, [self])

    def conj(self) -> 'Tensor':
        '''`Tensor::conj` (`tensor.cpp:105-113`): swap sides, transpose every block, conjugate.'''
        out = run(engine.plan_conj(self.structure), [self])
        if out.dtype.is_complex:
            out.data = out.data.conj().resolve_conj()
        return out

    def clone(self) -> 'Tensor':
        return Tensor._wrap(self.structure, self.data.clone())

    # -- arithmetic (`tensor.cpp:138-181`), on the whole flat buffer
    def _check_same(self, other: 'Tensor', what: str) -> None:
        if self.structure != other.structure:
            raise ValueError(f'a {what} needs both tensors with the same legs and the same charge')

    def __iadd__(self, other: 'Tensor') -> 'Tensor':
        self._check_same(other, 'sum')
        _native.axpby(self.dtype, 1.0, other.data, 1.0, self.data)
        self._version += 1
        return self

    def __isub__(self, other: 'Tensor') -> 'Tensor':
        self._check_same(other, 'difference')
        _native.axpby(self.dtype, -1.0, other.data, 1.0, self.data)
        self._version += 1
        return self

    def __imul__(self, scalar) -> 'Tensor':
        scalar = complex(scalar)
        if scalar.imag == 0.0:
            _native.axpby(self.dtype, 0.0, None, scalar.real, self.data)
        else:
            if not self.dtype.is_complex:
                raise ValueError(f'a block of real entries cannot take the complex factor {scalar}')
            self.data.mul_(scalar)
        self._version += 1
        return self

    def __mul__(self, scalar) -> 'Tensor':
        out = self.clone()
        out *= scalar
        return out

    __rmul__ = __mul__

    def __itruediv__(self, scalar) -> 'Tensor':
        if scalar == 0.0:
            raise ValueError('a tensor cannot be divided by zero')
        return self.__imul__(1.0 / float(scalar))

    def __repr__(self) -> str:
        return (f'Tensor({self.cat.name}, codomain={len(self.codomain)}, domain={len(self.domain)}, '
                f'qtotal={self.qtotal}, blocks={self.nblocks()}, dtype={self.dtype})')


# ------------------------------------------------------------------------------------------------
# Running programs


class _Workspaces:
    '''Device workspaces of recently run programs, kept so that a program run again on the same
    window reuses the hoisted moves of operands that have not changed (head/tail of a Lanczos
    matvec).  Least recently used entries are dropped past a byte budget.'''

    budget = 8 << 30

    def __init__(self):
        self.entries: 'OrderedDict[tuple, list]' = OrderedDict()
        self.bytes = 0

    def get(self, program, dtype, device, elements: int):
        key = (program.id, dtype, str(device))
        entry = self.entries.get(key)
        if entry is not None and entry[0].numel() >= elements:
            self.entries.move_to_end(key)
            return entry
        if entry is not None:
            self.bytes -= entry[0].numel() * entry[0].element_size()
            del self.entries[key]
        buffer = torch.empty(max(elements, 1), dtype=dtype, device=device)
        entry = [buffer, None, program]
        self.entries[key] = entry
        self.bytes += buffer.numel() * buffer.element_size()
        while self.bytes > self.budget and len(self.entries) > 1:
            _, old = self.entries.popitem(last=False)
            self.bytes -= old[0].numel() * old[0].element_size()
        return entry

    def clear(self):
        self.entries.clear()
        self.bytes = 0


workspaces = _Workspaces()


class Settings:
    '''How programs run: `mode` 0 batched (composed moves, grouped GEMM, hoisting), 1 plain (the
    reference's per-step, per-block launches, nothing hoisted); `precision` 'fp64', or 'fp32' to run
    the block products of FP64 operands on FP32 MFMA with FP32 workspaces (the result stays FP64);
    `tiles` how each block product's launch configuration is picked: 'default' (the device
    library's rule from block shapes), 'model' (the learned cost model) or 'exhaustive' (time every
    configuration on the program's first run and keep the fastest).'''
    mode = 0
    precision = 'fp64'
    tiles = 'default'

    class using:
        '''Context manager that sets the fields given and restores them on exit.'''

        def __init__(self, **fields):
            self.fields = fields
            self.saved = {}

        def __enter__(self):
            for name, value in self.fields.items():
                self.saved[name] = getattr(Settings, name)
                setattr(Settings, name, value)

        def __exit__(self, *exc):
            for name, value in self.saved.items():
                setattr(Settings, name, value)


def run(program, operands: Sequence[Tensor], out: Optional[torch.Tensor] = None, label: str = '') -> Tensor:
    '''Execute a planned program on device operands; returns the result tensor.'''
    first = operands[0]
    dtype, device = first.dtype, first.device
    for t in operands:
        _native.require_device(t.data)
        if t.dtype != dtype:
            raise ValueError('a contraction needs both tensors of one scalar type')
    result = program.out
    if out is None:
        out = torch.empty(result.size, dtype=dtype, device=device)
    mode = Settings.mode
    mixed = Settings.precision == 'fp32' and dtype == torch.float64 and mode == 0
    ws_dtype = torch.float32 if mixed else dtype
    elements = program.workspace + (program.plain_workspace if mode == 1 else 0)
    skip = 0
    if elements:
        entry = workspaces.get(program, ws_dtype, device, elements)
        hoist = program.hoistable
        # torch's in-place counter (shared by every view of the buffer) catches writes through
        # `block()` views, `to_tuple()` blocks and `.data`; `_version` catches the library's own
        # writes, which torch does not see; the buffer address catches a replaced `.data`
        identity = tuple((t._uid, t._version, t.data._version, t.data.data_ptr()) for t in operands)
        if mode == 0 and hoist and entry[1] is not None:
            for i, (now, then) in enumerate(zip(identity, entry[1])):
                if (hoist >> i) & 1 and now == then:
                    skip |= 1 << i
        entry[1] = identity if mode == 0 else None
        ws = entry[0].data_ptr()
    else:
        ws = 0
    args = (_native.dtype_code(dtype), [t.data.data_ptr() for t in operands], out.data_ptr(), ws, _native.stream(),
            mode, skip)
    if mixed:
        args = args + (0, -1, _native.TT_F32, _native.TT_F32)
    measure = False
    if Settings.tiles != 'default' and mode == 0:
        from .tuning import tuner
        measure = tuner.before(program, args, Settings.tiles)
    if Profiler.active is None:
        program.run(*args)
    else:
        Profiler.active.run(program, args, skip, mode, label)
    if measure:
        # Every block product is timed under each configuration on the data of this run; the run is
        # then repeated with the chosen ones, so the result never depends on what the timing left.
        tuner.after_first_run(program, args)
        program.run(*args)
    return Tensor._wrap(result, out)


def krylov_steps(program, operands: Sequence[Tensor], slot: int, V: torch.Tensor, steps: int, image: torch.Tensor,
                 scratch: torch.Tensor, alphas: torch.Tensor, betas: torch.Tensor) -> bool:
    '''The Krylov steps of a Lanczos run whose action is `program` with the basis row in operand `slot`,
    in one engine call (`Program.krylov`): each step runs the program on V[k] into `image`, then the
    library's orthonormalisation makes V[k + 1], alpha and beta.  The workspace and the hoisting of the
    other operands' moves follow `run`.  Returns False, with nothing queued, when the run has to go
    step by step instead (the per-block form, a profiler, a tile autotuner, other than FP64 vectors).'''
    first = operands[0]
    if Settings.mode != 0 or Profiler.active is not None or Settings.tiles != 'default' or \
            any(t.dtype != torch.float64 for t in operands) or V.dtype != torch.float64:
        return False
    for t in operands:
        _native.require_device(t.data)
    mixed = Settings.precision == 'fp32'
    ws_dtype = torch.float32 if mixed else torch.float64
    elements = program.workspace
    skip, ws = 0, 0
    if elements:
        entry = workspaces.get(program, ws_dtype, first.device, elements)
        hoist = program.hoistable
        identity = tuple((t._uid, t._version, t.data._version, t.data.data_ptr()) for t in operands)
        if hoist and entry[1] is not None:
            for i, (now, then) in enumerate(zip(identity, entry[1])):
                if i != slot and (hoist >> i) & 1 and now == then:
                    skip |= 1 << i
        # the basis operand changes every step; the others keep their moves in this workspace
        entry[1] = tuple(x if i != slot else None for i, x in enumerate(identity))
        ws = entry[0].data_ptr()
    code = _native.TT_F32 if mixed else -1
    program.krylov([t.data.data_ptr() for t in operands], slot, V.data_ptr(), V.stride(0), steps, image.data_ptr(),
                   scratch.data_ptr(), alphas.data_ptr(), betas.data_ptr(), ws, _native.stream(), skip, code, code)
    return True


class Profiler:
    '''Times every block-product and leg-move launch of the grouped form while active, op by op,
    and every step launch of the eigensolver's tridiagonalisation (kind 'tridiag'), with a start and
    a stop event attached to the kernel dispatch itself (so the interval is the kernel's execution,
    not the host's launch latency); `summary()` synchronises and returns per-kind totals with their
    algorithmic work.  Programs of the per-block form (mode 1) are timed as a whole by events on the
    stream (kind 'plain', with the flops of their block products).'''

    active = None

    def __init__(self):
        self.records = []   # (kind, flops, bytes moved, compute type, label, launches)
        self.ms = []        # the kernel time of each record
        self.plain = []     # (flops, start event, stop event) of each per-block program

    def __enter__(self):
        Profiler.active = self
        _native.eigh_observer = self._eigh
        return self

    def __exit__(self, *exc):
        Profiler.active = None
        _native.eigh_observer = None
        self._collect()

    def _eigh(self, sizes):
        launches, moved = _native.tridiagonal_steps(sizes)
        if launches:
            self.records.append(('tridiag', 0.0, moved, 'f64', 'eigh', launches))

    def _collect(self):
        durations = _native.timing_collect()
        expected = sum(r[5] for r in self.records[len(self.ms):])
        if len(durations) != expected:
            raise RuntimeError(f'{expected} timed launches produced {len(durations)} durations')
        at = 0
        for record in self.records[len(self.ms):]:
            self.ms.append(sum(durations[at:at + record[5]]))
            at += record[5]

    def run(self, program, args, skip, mode, label=''):
        if mode != 0:
            # the per-block form is timed as a whole, not launch by launch
            import torch
            flops = sum(cost[1] for cost in program.op_costs() if cost[0] == 'gemm')
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            program.run(*args)
            stop.record()
            self.plain.append((flops, start, stop))
            return
        program.prepare()
        costs = program.op_costs()
        ops = program.ops()
        slots = program.slots()
        mixed = len(args) > 10
        size_of = {_native.TT_F64: 8, _native.TT_F32: 4, _native.TT_C128: 16}
        outer = size_of[args[0]]
        inner = size_of[args[9]] if mixed else outer
        compute = 'f32' if mixed and args[10] == _native.TT_F32 else {8: 'f64', 4: 'f32', 16: 'c128'}[outer]

        def size(value):
            return inner if value >= 0 and slots[value][0] == 'workspace' else outer

        for i, ((kind, flops, written, read), (_, a, b, dst, inputs)) in enumerate(zip(costs, ops)):
            single = inputs and not (inputs & (inputs - 1))
            hoisted = kind == 'move' and single and (inputs & skip)
            timed = kind in ('gemm', 'move') and not hoisted
            before = _native.launch_count()
            if timed:
                _native.timing(True)
            try:
                self._one(program, args, i)
            finally:
                if timed:
                    _native.timing(False)
            launches = _native.launch_count() - before
            if timed and launches:      # an op with no work launches nothing
                # a move is typed by the workspace it writes: FP32 in the FP32 form's programs
                moved = {4: 'f32', 8: 'f64', 16: 'c128'}[inner]
                self.records.append((kind, flops, written * size(dst) + read * size(a),
                                     compute if kind == 'gemm' else moved, label, launches))

    @staticmethod
    def _one(program, args, i):
        base, extra = args[:7], args[9:] if len(args) > 7 else ()
        program.run(*base, i, i + 1, *extra)

    def summary(self, label=None, compute=None):
        '''Totals per kind, over every record or over the records of one einsum equation and one
        compute type (a block product's MFMA type; a leg move's workspace type).'''
        out = {}
        if self.plain and label is None and compute is None:
            ms = sum(start.elapsed_time(stop) for _, start, stop in self.plain)
            out['plain'] = {'ops': len(self.plain), 'launches': 0, 'ms': ms,
                            'flops': float(sum(f for f, _, _ in self.plain)), 'bytes': 0.0}
        for (kind, flops, moved, kind_compute, lab, launches), ms in zip(self.records, self.ms):
            if label is not None and lab != label:
                continue
            if compute is not None and kind_compute != compute:
                continue
            entry = out.setdefault(kind, {'ops': 0, 'launches': 0, 'ms': 0.0, 'flops': 0.0, 'bytes': 0.0})
            entry['ops'] += 1
            entry['launches'] += launches
            entry['ms'] += ms
            entry['flops'] += flops
            entry['bytes'] += moved
        return out


class _Plans:
    '''Programs by (request, interned structure ids): a lookup that skips building the engine's key.
    The engine keeps its own cache; this one holds at most `limit` programs, least recent first out.'''
    limit = 4096
    table: 'OrderedDict' = OrderedDict()

    @classmethod
    def get(cls, key, make):
        program = cls.table.get(key)
        if program is None:
            program = make()
            cls.table[key] = program
            if len(cls.table) > cls.limit:
                cls.table.popitem(last=False)
        else:
            cls.table.move_to_end(key)
        return program

    @classmethod
    def clear(cls):
        cls.table.clear()


def clear_caches() -> None:
    '''Drop every cached plan and program (the engine's and this module's) and the workspaces.'''
    _Plans.clear()
    workspaces.clear()
    engine.clear_caches()


def einsum_program(spec: str, tensors: Sequence[Tensor], outcharge=None):
    '''The planned program of an einsum over these operands' structures (cached).'''
    if not tensors:
        raise ValueError('an einsum needs at least one operand')
    charge = None if outcharge is None else tuple(outcharge)
    key = ('einsum', spec, tuple(t.structure.id for t in tensors), charge)
    return _Plans.get(key, lambda: engine.plan_einsum(spec, [t.structure for t in tensors], charge))


def einsum(spec: str, *tensors: Tensor, outcharge=None) -> Tensor:
    '''`einsum` (`tentorch/linalg/graded/einsum.cpp:444-598`, binding `bindings.cpp:347-361`).'''
    return run(einsum_program(spec, tensors, outcharge), tensors, label=spec)


def contract(left: Tensor, right: Tensor, left_legs, right_legs, charge=None) -> Tensor:
    '''`contract` (`tentorch/linalg/graded/contract.cpp:117-175`).'''
    if left.dtype != right.dtype:
        raise ValueError('a contraction needs both tensors of one scalar type')
    charge = None if charge is None else tuple(charge)
    key = ('contract', left.structure.id, right.structure.id, tuple(left_legs), tuple(right_legs), charge)
    program = _Plans.get(key, lambda: engine.plan_contract(left.structure, right.structure, list(left_legs),
                                                           list(right_legs), charge))
    return run(program, [left, right])


def _scalar_buffer(device) -> torch.Tensor:
    return torch.empty(2, dtype=torch.float64, device=device)


def inner_device(first: Tensor, second: Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    '''The reduced scalar product left on the device (one double, two for complex).'''
    if first.cat != second.cat:
        raise ValueError('a scalar product needs both tensors over one category')
    if first.codomain != second.codomain or first.domain != second.domain:
        raise ValueError('a scalar product needs both tensors with the same legs')
    out = _scalar_buffer(first.device) if out is None else out
    if first.qtotal != second.qtotal:
        out.zero_()
        return out
    _native.dot(first.dtype, first.data, second.data, out)
    return out


def inner(first: Tensor, second: Tensor) -> complex:
    '''`inner` (`tentorch/linalg/graded/compare.cpp:31-43`): Σ conj(a) b over blocks, no q_dim.'''
    value = inner_device(first, second).cpu()
    if first.dtype.is_complex:
        return complex(float(value[0]), float(value[1]))
    return complex(float(value[0]), 0.0)


def norm(tensor: Tensor) -> float:
    '''`norm` (`compare.cpp:45`).'''
    return float(np.sqrt(abs(inner(tensor, tensor))))


def is_close(first: Tensor, second: Tensor, rtol: float = 1e-10, atol: float = 1e-12) -> bool:
    '''`is_close` (`compare.cpp:12-29`).'''
    if first.structure != second.structure:
        return False
    return bool(torch.allclose(first.data, second.data, rtol=rtol, atol=atol))


def block_keys(cat, codomain, domain, qtotal) -> List[BlockKey]:
    return [BlockKey(*k) for k in engine.block_keys(cat, list(codomain), list(domain), tuple(qtotal))]


def codomain_legs(tensor: Tensor) -> List[int]:
    return list(range(len(tensor.codomain)))


def domain_legs(tensor: Tensor) -> List[int]:
    return list(range(len(tensor.codomain), tensor.nlegs()))
