'''
Oracle: the block-sparse tensor of the reference engine and the operations on the contraction path
(`place`, `contract`, `einsum`, `inner`, `norm`, `is_close`, `svd`, `qr`, `rq`), restated on the CPU
with numpy blocks.

TEST INFRASTRUCTURE ONLY: imported by `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline`
leg of `bench.py`, never by the product.

The leg moves follow the reference's per-entry algorithm: a `place` is a chain of elementary steps
(`tentorch/tensors/place.cpp:41-107,199-247`), each step a regrouping made of strided runs
(`tentorch/tensors/regroup.cpp:25-417`), applied block by block through buffers
(`place.cpp:127-180`) with the host pointer walk `to[last + i*ts] += c * from[first + i*fs]`
(`tentorch/tensors/cache.cpp:66-93`).  The product composes a whole chain into one table at the
level of tree pairs instead, so the two are independent computations of the same map.
'''

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .symmetry import (Category, FusionTree, Leg, Sector, SideLayout, block_keys, layout_of, render,
                       trimmed)

BlockKey = Tuple[Sector, Sector, int]


def _same_legs(a: Sequence[Leg], b: Sequence[Leg]) -> bool:
    return len(a) == len(b) and all(x == y for x, y in zip(a, b))


class Tensor:
    '''`Tensor` (`tentorch/tensors/tensor.h:43-216`, `tensor.cpp:23-61`): every allowed key is
    allocated with zeros; blocks are kept in key order (a `std::map` in the reference).'''

    def __init__(self, cat: Category, codomain: Sequence[Leg], domain: Sequence[Leg], qtotal: Sector,
                 dtype=np.float64):
        self.cat = cat
        self.codomain = list(codomain)
        self.domain = list(domain)
        self.qtotal = tuple(qtotal)
        self.dtype = np.dtype(dtype)
        for leg in self.codomain + self.domain:
            if leg.cat is not cat:
                raise ValueError('a leg is over a different category')
        keys = block_keys(cat, self.codomain, self.domain, self.qtotal)
        self.blocks: Dict[BlockKey, np.ndarray] = {}
        for key in sorted(keys):
            rows = layout_of(cat, self.codomain, key[0]).rows()
            cols = layout_of(cat, self.domain, key[1]).rows()
            self.blocks[key] = np.zeros((rows, cols), dtype=self.dtype)

    # -- structure
    def legs(self) -> List[Leg]:
        return self.codomain + self.domain

    def nlegs(self) -> int:
        return len(self.codomain) + len(self.domain)

    def keys(self) -> List[BlockKey]:
        return list(self.blocks.keys())

    def block(self, r: Sector, s: Sector, copy: int = 0) -> np.ndarray:
        try:
            return self.blocks[(r, s, copy)]
        except KeyError:
            raise IndexError(f'the tensor has no block at {render(r)} x {render(s)} of copy {copy}')

    def numel(self) -> int:
        return sum(b.size for b in self.blocks.values())

    def flat(self) -> np.ndarray:
        '''All blocks concatenated in key order, row-major within a block.'''
        if not self.blocks:
            return np.zeros(0, dtype=self.dtype)
        return np.concatenate([b.reshape(-1) for b in self.blocks.values()])

    def set_flat(self, values: np.ndarray) -> 'Tensor':
        at = 0
        for key, b in self.blocks.items():
            b[...] = values[at:at + b.size].reshape(b.shape)
            at += b.size
        return self

    def _empty_like(self) -> 'Tensor':
        out = Tensor.__new__(Tensor)
        out.cat, out.codomain, out.domain = self.cat, list(self.codomain), list(self.domain)
        out.qtotal, out.dtype = self.qtotal, self.dtype
        out.blocks = {}
        return out

    def clone(self) -> 'Tensor':
        out = self._empty_like()
        out.blocks = {k: b.copy() for k, b in self.blocks.items()}
        return out

    def conj(self) -> 'Tensor':
        '''`Tensor::conj` (`tensor.cpp:105-113`).'''
        out = Tensor(self.cat, self.domain, self.codomain, self.cat.dual(self.qtotal), self.dtype)
        for (r, s, c), b in self.blocks.items():
            out.block(s, r, c)[...] = b.conj().T
        return out

    def _check_same(self, other: 'Tensor', what: str) -> None:
        if not (self.cat is other.cat and self.qtotal == other.qtotal
                and _same_legs(self.codomain, other.codomain) and _same_legs(self.domain, other.domain)):
            raise ValueError(f'a {what} needs both tensors with the same legs and the same charge')

    def __iadd__(self, other: 'Tensor') -> 'Tensor':
        self._check_same(other, 'sum')
        for k, b in other.blocks.items():
            self.blocks[k] += b
        return self

    def __isub__(self, other: 'Tensor') -> 'Tensor':
        self._check_same(other, 'difference')
        for k, b in other.blocks.items():
            self.blocks[k] -= b
        return self

    def __imul__(self, scalar) -> 'Tensor':
        '''`operator*=` (`tensor.cpp:159-175`): a complex factor with zero imaginary part is real.'''
        scalar = complex(scalar)
        if scalar.imag == 0.0:
            for b in self.blocks.values():
                b *= scalar.real
            return self
        for b in self.blocks.values():
            if not np.iscomplexobj(b):
                raise ValueError('a block of real entries cannot take a complex factor')
            b *= scalar
        return self

    def __mul__(self, scalar) -> 'Tensor':
        out = self.clone()
        out *= scalar
        return out

    __rmul__ = __mul__

    def __itruediv__(self, scalar) -> 'Tensor':
        if scalar == 0.0:
            raise ValueError('a tensor cannot be divided by zero')
        for b in self.blocks.values():
            b /= scalar
        return self

    # -- leg moves
    def place(self, above: Sequence[int], below: Optional[Sequence[int]] = None) -> 'Tensor':
        return place(self, list(above), None if below is None else list(below))

    def __repr__(self) -> str:
        return (f'Tensor({self.cat.name}, codomain={len(self.codomain)}, domain={len(self.domain)}, '
                f'q={self.qtotal}, blocks={len(self.blocks)})')


def from_blocks(cat, codomain, domain, qtotal, blocks=None, dtype=np.float64) -> Tensor:
    '''`Tensor.from_tuple` (`tentorch/_ext/bindings.cpp:131-152`).'''
    out = Tensor(cat, codomain, domain, qtotal, dtype)
    for key, value in (blocks or {}).items():
        mine = out.block(*key)
        value = np.asarray(value)
        if value.shape != mine.shape:
            raise ValueError(f'the block at {key} has shape {value.shape}, not {mine.shape}')
        mine[...] = value.astype(out.dtype) if np.iscomplexobj(mine) else value.real.astype(out.dtype)
    return out


def random_tensor(cat, codomain, domain, qtotal, rng: np.random.Generator, dtype=np.float64) -> Tensor:
    '''`Tensor.random` (`bindings.cpp:157-164`): N(0, 1) entries, drawn block by block in key order
    from a host generator (the reference draws from torch's generator on its device).'''
    out = Tensor(cat, codomain, domain, qtotal, dtype)
    for b in out.blocks.values():
        if np.iscomplexobj(b):
            b[...] = (rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)) / np.sqrt(2)
        else:
            b[...] = rng.standard_normal(b.shape)
    return out


# ----------------------------------------------------------------------------------------------
# Regroupings: one elementary step of a chain, for one source block.


class Shift:
    __slots__ = ('to', 'rows', 'cols', 'first', 'last', 'from_step', 'to_step', 'run', 'coef')

    def __init__(self, to, rows, cols):
        self.to, self.rows, self.cols = to, rows, cols
        self.first, self.last, self.from_step, self.to_step, self.run, self.coef = [], [], [], [], [], []

    def add(self, first, last, from_step, to_step, run, coef):
        self.first.append(first)
        self.last.append(last)
        self.from_step.append(from_step)
        self.to_step.append(to_step)
        self.run.append(run)
        self.coef.append(coef)


def _chain_at(cat: Category, tree: FusionTree, last: int) -> Sector:
    '''`chain_at` (`regroup.cpp:15-22`).'''
    if last < 0:
        return cat.trivial()
    if last + 1 == tree.size():
        return tree.coupled
    if last == 0:
        return tree.sectors[0]
    return tree.inner[last - 1]


def exchange_regrouping(cat, codomain, domain, key, qtotal, p) -> List[Shift]:
    '''`exchange_regrouping` (`regroup.cpp:25-258`): swap codomain legs p and p + 1.'''
    count = len(codomain)
    exchanged = list(codomain)
    exchanged[p], exchanged[p + 1] = exchanged[p + 1], exchanged[p]
    source = layout_of(cat, codomain, key[0])
    target = layout_of(cat, exchanged, key[0])
    stride = layout_of(cat, domain, key[1]).rows()
    shift = Shift(key, 0, 0)

    def signature(tree: FusionTree, exchanged_tree: bool):
        out = []
        for axis in range(count):
            frm = axis
            if exchanged_tree and axis in (p, p + 1):
                frm = p + 1 if axis == p else p
            out.extend(tree.sectors[frm])
        for at, charge in enumerate(tree.inner):
            if at == p - 1:
                continue
            out.extend(charge)
        for at, copy in enumerate(tree.copies):
            if at in (p - 1, p):
                continue
            out.append(copy)
        return tuple(out)

    def moved(axis):
        return p + 1 if axis == p else (p if axis == p + 1 else axis)

    def shape_of(tree, legs):
        return [legs[axis].mult_of(tree.sectors[axis]) for axis in range(count)]

    groups: Dict[tuple, List[int]] = {}
    for at, tree in enumerate(source.trees):
        groups.setdefault(signature(tree, False), []).append(at)
    partners: Dict[tuple, List[int]] = {}
    for at, tree in enumerate(target.trees):
        partners.setdefault(signature(tree, True), []).append(at)

    factors: Dict[Tuple[int, int], complex] = {}
    for marks in sorted(groups):
        olds = groups[marks]
        news = partners.get(marks)
        if news is None:
            raise RuntimeError('the exchange has no image for a tree of the codomain')
        sample = source.trees[olds[0]]
        y = _chain_at(cat, sample, p - 1)
        d = _chain_at(cat, sample, p + 1)
        b, c = sample.sectors[p], sample.sectors[p + 1]
        straight = cat.F_symbol(y, b, c, d)
        braided = cat.F_symbol(y, c, b, d)
        if straight.shape[0] != len(olds) or braided.shape[0] != len(news):
            raise RuntimeError('the exchange does not match the trees of the block')
        offset = []
        at = 0
        for sector, mult in cat.outcomes(b, c):
            offset.append((sector, at))
            at += mult
        frm = cat.f_columns(y, b, c, d)
        to = cat.f_columns(y, c, b, d)
        braid = cat.R_symbol(b, c)

        def braid_at(index):
            for sector, start in offset:
                if sector == index[0]:
                    return start + index[1]
            return -1

        rows = [braid_at(i) for i in to]
        cols = [braid_at(j) for j in frm]
        terms = []
        for i in range(len(to)):
            if rows[i] < 0:
                continue
            for j in range(len(frm)):
                if cols[j] < 0:
                    continue
                if to[i][0] != frm[j][0] or to[i][2] != frm[j][2]:
                    continue
                terms.append((i, rows[i], cols[j], j))
        for one in range(len(news)):
            for other in range(len(olds)):
                value = 0j
                for i, br, bc, j in terms:
                    value += braided[one, i] * braid[br, bc] * np.conj(straight[other, j])
                if value == 0:
                    continue
                factors.setdefault((news[one], olds[other]), complex(value))

    before: Dict[tuple, List[int]] = {}
    after: Dict[tuple, List[int]] = {}
    after_shapes = [None] * len(target.trees)
    for at, tree in enumerate(source.trees):
        before.setdefault(tuple(shape_of(tree, codomain)), []).append(at)
    for at, tree in enumerate(target.trees):
        shape = shape_of(tree, exchanged)
        back = [0] * count
        for axis in range(count):
            back[moved(axis)] = shape[axis]
        after_shapes[at] = shape
        after.setdefault(tuple(back), []).append(at)
    for shape in sorted(before):
        olds = before[shape]
        news = after.get(shape)
        if news is None:
            raise RuntimeError('the exchange has no image for a block of trees')
        width = int(np.prod(shape)) if shape else 1
        target_shape = after_shapes[news[0]]
        width_of = np.zeros(width, dtype=np.int64)
        for at in range(width):
            digits = [0] * len(shape)
            left = at
            for axis in range(len(shape) - 1, -1, -1):
                digits[axis] = left % shape[axis]
                left //= shape[axis]
            placed = [0] * count
            for axis in range(count):
                placed[moved(axis)] = digits[axis]
            index = 0
            for axis in range(count):
                index = index * target_shape[axis] + placed[axis]
            width_of[at] = index
        for one in news:
            for other in olds:
                factor = factors.get((one, other))
                if factor is None:
                    continue
                at = 0
                while at < width:
                    run = 1
                    while at + run < width and width_of[at + run] == width_of[at] + run:
                        run += 1
                    shift.add((source.offsets[other] + at) * stride,
                              (target.offsets[one] + int(width_of[at])) * stride, 1, 1, run * stride,
                              factor)
                    at += run
    return [shift]


def bend_regrouping(cat, codomain, domain, key, qtotal, from_codomain: bool) -> List[Shift]:
    '''`bend_regrouping` (`regroup.cpp:260-417`): the last leg of one side onto the end of the
    other, through F(prefix, charge, dual(root), q); the leg lands carrying the dual sector
    (`regroup.cpp:355-358`).'''
    giving = codomain if from_codomain else domain
    taking = domain if from_codomain else codomain
    moved_leg = giving[-1]
    kept_giving = giving[:-1]
    kept_taking = list(taking) + [moved_leg]
    out_codomain = kept_giving if from_codomain else kept_taking
    out_domain = kept_taking if from_codomain else kept_giving
    moved_side = layout_of(cat, codomain if from_codomain else domain,
                           key[0] if from_codomain else key[1])
    gains_side = layout_of(cat, domain if from_codomain else codomain,
                           key[1] if from_codomain else key[0])
    moved_root = key[0] if from_codomain else key[1]
    gains_root = key[1] if from_codomain else key[0]
    source_stride = (gains_side if from_codomain else moved_side).rows()

    shifts: List[Shift] = []
    at_of: Dict[BlockKey, int] = {}

    def index_of(to, rows, cols):
        found = at_of.get(to)
        if found is not None:
            return found
        at_of[to] = len(shifts)
        shifts.append(Shift(to, rows, cols))
        return at_of[to]

    out_layouts: Dict[Tuple[Sector, Sector], Tuple[SideLayout, SideLayout]] = {}
    for gains_at, gains_tree in enumerate(gains_side.trees):
        gains = gains_side.offsets[gains_at + 1] - gains_side.offsets[gains_at]
        for moved_at, moved_tree in enumerate(moved_side.trees):
            charge = moved_tree.sectors[-1]
            copies = moved_leg.mult_of(charge)
            span = moved_side.offsets[moved_at + 1] - moved_side.offsets[moved_at]
            inside = span // copies
            kept_tree = trimmed(cat, moved_tree)
            prefix = kept_tree.coupled
            copy = moved_tree.copies[-1] if moved_tree.size() >= 2 else 0
            first = prefix if from_codomain else gains_root
            middle = charge if from_codomain else cat.dual(charge)
            last = cat.dual(gains_root) if from_codomain else cat.dual(prefix)
            coefficients = cat.F_symbol(first, middle, last, qtotal)
            across = cat.f_rows(first, middle, last, qtotal)
            along = cat.f_columns(first, middle, last, qtotal)
            named = across if from_codomain else along
            free = along if from_codomain else across
            named_root = moved_root if from_codomain else cat.dual(moved_root)
            taken = [i for i, n in enumerate(named)
                     if n[0] == named_root and n[1] == copy and n[2] == key[2]]
            if not taken:
                continue
            source_row = moved_side.offsets[moved_at]
            source_column = gains_side.offsets[gains_at]
            for index in taken:
                for column, taken_side in enumerate(free):
                    value = coefficients[index, column]
                    if value == 0:
                        continue
                    extended_root = cat.dual(taken_side[0]) if from_codomain else taken_side[0]
                    ext_sectors = gains_tree.sectors + (cat.dual(charge),)
                    ext_inner = gains_tree.inner + ((gains_root,) if gains_tree.size() >= 2 else ())
                    ext_copies = gains_tree.copies + ((taken_side[1],) if gains_tree.size() >= 1 else ())
                    extended = FusionTree(ext_sectors, ext_inner, ext_copies, extended_root)
                    out_cod_root = prefix if from_codomain else extended_root
                    out_dom_root = extended_root if from_codomain else prefix
                    roots = (out_cod_root, out_dom_root)
                    found = out_layouts.get(roots)
                    if found is None:
                        found = (layout_of(cat, out_codomain, out_cod_root),
                                 layout_of(cat, out_domain, out_dom_root))
                        out_layouts[roots] = found
                    rows_l, cols_l = found
                    row_at = rows_l.index.get((kept_tree if from_codomain else extended).key())
                    col_at = cols_l.index.get((extended if from_codomain else kept_tree).key())
                    if row_at is None or col_at is None:
                        continue
                    row = rows_l.offsets[row_at]
                    column_offset = cols_l.offsets[col_at]
                    target_stride = cols_l.rows()
                    shift = shifts[index_of((out_cod_root, out_dom_root, taken_side[2]), rows_l.rows(),
                                            cols_l.rows())]
                    factor = complex(value)
                    if from_codomain:
                        for shifted in range(inside):
                            for gain in range(gains):
                                shift.add((source_row + shifted * copies) * source_stride + source_column + gain,
                                          (row + shifted) * target_stride + column_offset + gain * copies,
                                          source_stride, 1, copies, factor)
                    else:
                        for gain in range(gains):
                            for copy_at in range(copies):
                                shift.add((source_column + gain) * source_stride + source_row + copy_at,
                                          (row + gain * copies + copy_at) * target_stride + column_offset,
                                          copies, 1, inside, factor)
    return shifts


class _Applier:
    '''The pieces of one (plan, target) as flat index arrays (`cache.cpp:66-93` walk, vectorised;
    runs are expanded in plan order so repeated targets accumulate in that order).'''
    __slots__ = ('reads', 'writes', 'coef')

    def __init__(self, shifts: List[Shift], target):
        reads, writes, coef = [], [], []
        for shift in shifts:
            if shift.to != target or not shift.run:
                continue
            first = np.asarray(shift.first, dtype=np.int64)
            last = np.asarray(shift.last, dtype=np.int64)
            fs = np.asarray(shift.from_step, dtype=np.int64)
            ts = np.asarray(shift.to_step, dtype=np.int64)
            run = np.asarray(shift.run, dtype=np.int64)
            c = np.asarray(shift.coef, dtype=np.complex128)
            piece = np.repeat(np.arange(len(run)), run)
            starts = np.cumsum(run) - run
            i = np.arange(int(run.sum())) - np.repeat(starts, run)
            reads.append(first[piece] + i * fs[piece])
            writes.append(last[piece] + i * ts[piece])
            coef.append(c[piece])
        self.reads = np.concatenate(reads) if reads else np.zeros(0, dtype=np.int64)
        self.writes = np.concatenate(writes) if writes else np.zeros(0, dtype=np.int64)
        self.coef = np.concatenate(coef) if coef else np.zeros(0, dtype=np.complex128)

    def apply(self, source: np.ndarray, into: np.ndarray) -> None:
        if not self.reads.size:
            return
        src = source.reshape(-1)
        dst = into.reshape(-1)
        coef = self.coef if np.iscomplexobj(dst) else self.coef.real
        np.add.at(dst, self.writes, coef * src[self.reads])


# The plan cache of `TensorCaches` (`cache.cpp:197-218`), without the byte budget.
_plans: Dict[tuple, Tuple[List[Shift], Dict]] = {}


def _plan_of(kind, position, cat, codomain, domain, key, qtotal):
    q = (kind, position, id(cat), tuple(l.key() for l in codomain), tuple(l.key() for l in domain),
         key, qtotal)
    found = _plans.get(q)
    if found is None:
        if kind == 'exchange':
            shifts = exchange_regrouping(cat, codomain, domain, key, qtotal, position)
        else:
            shifts = bend_regrouping(cat, codomain, domain, key, qtotal, kind == 'to_domain')
        found = (shifts, {})
        _plans[q] = found
    return found


def clear_plans() -> None:
    _plans.clear()


class _Chain:
    '''`Chain` (`place.cpp:41-107`).'''

    def __init__(self, codomain, domain):
        self.codomain = list(codomain)
        self.domain = list(domain)
        self.steps = []   # (kind, position, codomain after, domain after)

    def _add(self, kind, position):
        self.steps.append((kind, position, list(self.codomain), list(self.domain)))

    def exchange(self, position):
        self.codomain[position], self.codomain[position + 1] = self.codomain[position + 1], self.codomain[position]
        self._add('exchange', position)

    def fold(self, from_codomain):
        if from_codomain:
            self.domain.append(self.codomain.pop())
        else:
            self.codomain.append(self.domain.pop())
        self._add('to_domain' if from_codomain else 'to_codomain', 0)

    def sort(self, order):
        which = list(range(len(self.codomain)))
        for at in range(len(order)):
            frm = at
            while which[frm] != order[at]:
                frm += 1
            while frm > at:
                self.exchange(frm - 1)
                which[frm - 1], which[frm] = which[frm], which[frm - 1]
                frm -= 1


def _carried(tensor: Tensor, steps) -> Tensor:
    '''`carried` (`place.cpp:127-180`).'''
    if not steps:
        return tensor
    out = Tensor(tensor.cat, steps[-1][2], steps[-1][3], tensor.qtotal, tensor.dtype)
    for start, block in tensor.blocks.items():
        live = {start: block}
        codomain, domain = tensor.codomain, tensor.domain
        for at, (kind, position, cod_after, dom_after) in enumerate(steps):
            last = at + 1 == len(steps)
            nxt: Dict[BlockKey, np.ndarray] = {}
            for key, buffer in live.items():
                shifts, appliers = _plan_of(kind, position, tensor.cat, codomain, domain, key, tensor.qtotal)
                targets = []
                for shift in shifts:
                    if all(seen[0] != shift.to for seen in targets):
                        targets.append((shift.to, (buffer.shape[0] if shift.rows == 0 else shift.rows,
                                                   buffer.shape[1] if shift.cols == 0 else shift.cols)))
                for target, shape in targets:
                    if last:
                        into = out.blocks[target]
                    else:
                        into = nxt.get(target)
                        if into is None:
                            into = np.zeros(shape, dtype=tensor.dtype)
                            nxt[target] = into
                    applier = appliers.get(target)
                    if applier is None:
                        applier = _Applier(shifts, target)
                        appliers[target] = applier
                    applier.apply(buffer, into)
            codomain, domain = cod_after, dom_after
            live = nxt
    return out


def place(tensor: Tensor, above: List[int], below: Optional[List[int]] = None) -> Tensor:
    '''`Tensor::place` (`place.cpp:183-247`).'''
    height, depth = len(tensor.codomain), len(tensor.domain)
    total = height + depth
    if below is None:
        named = [False] * total
        for leg in above:
            if 0 <= leg < total:
                named[leg] = True
        below = [leg for leg in range(total) if not named[leg]]
    named = [False] * total
    for side in (above, below):
        for leg in side:
            if leg < 0 or leg >= total:
                raise ValueError(f'the leg {leg} does not exist, the tensor has {total} legs')
            if named[leg]:
                raise ValueError(f'the leg {leg} is named twice')
            named[leg] = True
    if len(above) + len(below) != total:
        left_out = named.index(False)
        raise ValueError(f'the leg {left_out} is named on neither side')
    chain = _Chain(tensor.codomain, tensor.domain)
    descends = any(leg < height for leg in below)
    kept = 0
    while (not descends and kept < depth and kept < len(below) and below[kept] == height + kept):
        kept += 1
    for _ in range(kept, depth):
        chain.fold(False)

    def position_of(leg):
        return leg if leg < height else height + (depth - 1 - (leg - height))

    order = [position_of(leg) for leg in above]
    order += [position_of(leg) for leg in list(reversed(below))[:len(below) - kept]]
    chain.sort(order)
    for _ in range(kept, len(below)):
        chain.fold(True)
    return _carried(tensor, chain.steps)


def codomain_legs(t: Tensor) -> List[int]:
    return list(range(len(t.codomain)))


def domain_legs(t: Tensor) -> List[int]:
    return list(range(len(t.codomain), t.nlegs()))


# ----------------------------------------------------------------------------------------------
# Contraction


def _without(total, taken):
    named = set(taken)
    return [leg for leg in range(total) if leg not in named]


def _check_pairs(left: Tensor, right: Tensor, left_legs, right_legs):
    '''`check_pairs` (`contract.cpp:47-77`).'''
    if left.cat is not right.cat:
        raise ValueError('a contraction needs both tensors over one category')
    if left.dtype != right.dtype:
        raise ValueError('a contraction needs both tensors of one scalar type')
    if len(left_legs) != len(right_legs):
        raise ValueError(f'the two legs of a contraction must be given in pairs, got {len(left_legs)} '
                         f'and {len(right_legs)}')
    for t, legs, side in ((left, left_legs, 'first'), (right, right_legs, 'second')):
        seen = set()
        for leg in legs:
            if leg < 0 or leg >= t.nlegs():
                raise ValueError(f'the leg {leg} of the {side} tensor is out of range of its {t.nlegs()} legs')
            if leg in seen:
                raise ValueError(f'the leg {leg} of the {side} tensor is summed twice')
            seen.add(leg)
    la, ra = left.legs(), right.legs()
    for at, (l, r) in enumerate(zip(left_legs, right_legs)):
        if not la[l] == ra[r]:
            raise ValueError(f'the pair at position {at} joins legs that do not name one space')


def _require_single_closing(cat, above, below, charge, subject='the roots', what='a contraction'):
    if cat.n_symbol(above, cat.dual(below), charge) > 1:
        raise ValueError(f'{subject} {render(above)} and {render(below)} close to {render(charge)} in '
                         f'more than one way, which {what} does not recouple yet')


def align(left, right, left_legs, right_legs):
    '''`align` (`contract.cpp:108-115`): summed legs last on the left's domain, first on the right's
    codomain.'''
    _check_pairs(left, right, left_legs, right_legs)
    return (left.place(_without(left.nlegs(), left_legs), list(left_legs)),
            right.place(list(right_legs), _without(right.nlegs(), right_legs)))


def contract(left: Tensor, right: Tensor, left_legs, right_legs, charge: Optional[Sector] = None) -> Tensor:
    '''`contract` (`contract.cpp:117-175`): per-root block products, `addmm_` per pair.'''
    if charge is None:
        if right.qtotal != left.cat.trivial():
            raise ValueError('the second tensor of a contraction must carry the trivial charge')
        charge = left.qtotal
    _check_pairs(left, right, left_legs, right_legs)
    cat = left.cat
    if cat.n_symbol(left.qtotal, right.qtotal, charge) == 0:
        raise ValueError(f'the charge {render(charge)} is not an outcome of {render(left.qtotal)} and '
                         f'{render(right.qtotal)}')
    _require_single_closing(cat, left.qtotal, right.qtotal, charge, 'the charges')
    above, under = align(left, right, left_legs, right_legs)
    out = Tensor(cat, above.codomain, under.domain, charge, left.dtype)
    for k in above.blocks:
        _require_single_closing(cat, k[0], k[1], above.qtotal)
    for k in under.blocks:
        _require_single_closing(cat, k[0], k[1], under.qtotal)
    for k in out.blocks:
        _require_single_closing(cat, k[0], k[1], charge)
    rows: Dict[Sector, list] = {}
    for (r, s, c), b in under.blocks.items():
        rows.setdefault(r, []).append((s, b))
    for (r, s, c), lb in above.blocks.items():
        found = rows.get(s)
        if found is None:
            continue
        for domain_root, rb in found:
            if cat.n_symbol(r, cat.dual(domain_root), charge) == 0:
                continue
            if lb.shape[1] != rb.shape[0]:
                raise RuntimeError('the two blocks do not line up')
            out.block(r, domain_root)[...] += lb @ rb
    return out


# ----------------------------------------------------------------------------------------------
# einsum (`tentorch/linalg/graded/einsum.cpp`)

_MAX_CHARGES = 4096


def _side_of(text):
    labels, above = [], []
    for letter in text:
        if letter == ' ':
            continue
        labels.append(letter.lower())
        above.append('A' <= letter <= 'Z')
    return labels, above


def _position_of(labels, label):
    try:
        return labels.index(label)
    except ValueError:
        raise RuntimeError(f"the label '{label}' left the equation")


def _laid_out(tensor: Tensor, side, summed):
    '''`laid_out` (`einsum.cpp:65-81`).'''
    labels_in, above_in = side
    codomain = len(tensor.codomain)
    above, below = [], []
    for leg in range(tensor.nlegs()):
        stays = (leg < codomain) if labels_in[leg] in summed else above_in[leg]
        (above if stays else below).append(leg)
    labels = [labels_in[leg] for leg in above] + [labels_in[leg] for leg in below]
    return tensor.place(above, below), labels


def _partial_trace(tensor: Tensor):
    '''`partial_trace` (`einsum.cpp:95-159`).'''
    cat, qtotal = tensor.cat, tensor.qtotal
    out_codomain = tensor.codomain[:-1]
    out_domain = tensor.domain[:-1]
    out: Dict[Sector, Tensor] = {}
    for (r, s, c), matrix in tensor.blocks.items():
        _require_single_closing(cat, r, s, qtotal, 'the roots', 'a trace')
        rows = layout_of(cat, tensor.codomain, r)
        cols = layout_of(cat, tensor.domain, s)
        for row_at, above_tree in enumerate(rows.trees):
            row_span = rows.offsets[row_at + 1] - rows.offsets[row_at]
            kept_above = trimmed(cat, above_tree)
            for col_at, under_tree in enumerate(cols.trees):
                col_span = cols.offsets[col_at + 1] - cols.offsets[col_at]
                if above_tree.sectors[-1] != under_tree.sectors[-1]:
                    continue
                kept_under = trimmed(cat, under_tree)
                out_rows = layout_of(cat, out_codomain, kept_above.coupled)
                out_cols = layout_of(cat, out_domain, kept_under.coupled)
                new_row = out_rows.index.get(kept_above.key(), -1)
                new_col = out_cols.index.get(kept_under.key(), -1)
                if new_row < 0 or new_col < 0:
                    raise RuntimeError('a kept tree of a trace is not in the layout it was trimmed from')
                inside_rows = out_rows.offsets[new_row + 1] - out_rows.offsets[new_row]
                inside_cols = out_cols.offsets[new_col + 1] - out_cols.offsets[new_col]
                copies = row_span // inside_rows
                sub = matrix[rows.offsets[row_at]:rows.offsets[row_at] + row_span,
                             cols.offsets[col_at]:cols.offsets[col_at] + col_span]
                piece = np.einsum('ikjk->ij', sub.reshape(inside_rows, copies, inside_cols, copies))
                for closing, _ in cat.outcomes(kept_above.coupled, cat.dual(kept_under.coupled)):
                    _require_single_closing(cat, kept_above.coupled, kept_under.coupled, closing,
                                            'the kept roots', 'a trace')
                    if closing not in out:
                        out[closing] = Tensor(cat, out_codomain, out_domain, closing, tensor.dtype)
                    blk = out[closing].block(kept_above.coupled, kept_under.coupled)
                    r0, c0 = out_rows.offsets[new_row], out_cols.offsets[new_col]
                    blk[r0:r0 + inside_rows, c0:c0 + inside_cols] += piece
    return [out[k] for k in sorted(out)]


def _leg_size(cat, leg: Leg) -> float:
    return float(sum(m * cat.q_dim(s) for s, m in zip(leg.sectors, leg.multiplicities)))


def _tensor_size(labels, dims):
    out = 1.0
    for label in labels:
        out *= dims[label]
    return out


def _union(left, right):
    out = list(left)
    for label in right:
        if label not in out:
            out.append(label)
    return out


def _shared(left, right):
    return [label for label in left if label in right]


def _product_labels(left, right):
    shared = _shared(left, right)
    return [label for label in _union(left, right) if label not in shared]


def fold_order(inputs, dims):
    '''`fold_order` (`einsum.cpp:216-301`).'''
    count = len(inputs)
    order = []
    if count <= 12:
        masks = 1 << (count - 1)
        labels = [list(inputs[0]) for _ in range(masks)]
        for mask in range(1, masks):
            bit = (mask & -mask).bit_length() - 1
            labels[mask] = _union(labels[mask ^ (1 << bit)], inputs[bit + 1])
        for pass_ in range(2):
            cost = [float('inf')] * masks
            summed = [0] * masks
            last = [0] * masks
            cost[0] = 0.0
            for mask in range(1, masks):
                for bit in range(count - 1):
                    if mask & (1 << bit):
                        shared = _shared(labels[mask ^ (1 << bit)], inputs[bit + 1])
                        if pass_ == 0 and not shared:
                            continue
                        candidate = cost[mask ^ (1 << bit)] + _tensor_size(labels[mask], dims)
                        better = (candidate < cost[mask]
                                  or (candidate == cost[mask] and len(shared) > summed[mask])
                                  or (candidate == cost[mask] and len(shared) == summed[mask]
                                      and bit > last[mask]))
                        if better:
                            cost[mask] = candidate
                            summed[mask] = len(shared)
                            last[mask] = bit
            if cost[masks - 1] == float('inf'):
                continue
            mask = masks - 1
            while mask:
                order.append(last[mask] + 1)
                mask ^= 1 << last[mask]
            order.append(0)
            order.reverse()
            return order
    taken = [False] * count
    live = list(inputs[0])
    order.append(0)
    taken[0] = True
    for _ in range(1, count):
        best, best_size, found = -1, 0.0, False
        for pass_ in range(2):
            if found:
                break
            for at in range(count):
                if taken[at]:
                    continue
                if (not _shared(live, inputs[at])) != (pass_ == 1):
                    continue
                size = _tensor_size(_union(live, inputs[at]), dims)
                if not found or size < best_size:
                    found, best, best_size = True, at, size
        order.append(best)
        taken[best] = True
        live = _union(live, inputs[best])
    return order


def _combine(left_states, left_labels, right_states, right_labels):
    '''`combine` (`einsum.cpp:306-345`).'''
    def closings_of(cat, first, second):
        out = []
        for sector, _ in cat.outcomes(first, second):
            if sector not in out:
                out.append(sector)
        return out

    shared = _shared(left_labels, right_labels)
    left_legs = [_position_of(left_labels, label) for label in shared]
    right_legs = [_position_of(right_labels, label) for label in shared]
    folded = []
    for left in left_states:
        for right in right_states:
            above, under = align(left, right, left_legs, right_legs)
            for closed in closings_of(left.cat, left.qtotal, right.qtotal):
                folded.append(contract(above, under, domain_legs(above), codomain_legs(under), closed))
    return folded


def _contract_path(inputs, output, operands, order):
    '''`contract_path` (`einsum.cpp:351-376`).'''
    labels = list(inputs[0])
    states = [operands[0]]
    for step in range(1, len(order)):
        at = order[step]
        folded = _combine(states, labels, [operands[at]], inputs[at])
        if len(folded) > _MAX_CHARGES:
            raise ValueError(f'the products of the equation leave more than {_MAX_CHARGES} charges to sum')
        labels = _product_labels(labels, inputs[at])
        states = folded
    out_labels, out_above = output
    above, below = [], []
    for label, up in zip(out_labels, out_above):
        (above if up else below).append(_position_of(labels, label))
    return [state.place(above, below) for state in states]


def _self_contract(operands, inputs, spec):
    '''`self_contract` (`einsum.cpp:386-439`).'''
    prepared_ops, prepared_labels = [], []
    for at, (operand, labels) in enumerate(zip(operands, inputs)):
        alternatives = [operand]
        labels = list(labels)
        while True:
            first = second = -1
            for one in range(len(labels)):
                if first >= 0:
                    break
                for other in range(one + 1, len(labels)):
                    if labels[one] == labels[other]:
                        first, second = one, other
                        break
            if first < 0:
                break
            label = labels[first]
            codomain = len(alternatives[0].codomain)
            legs = alternatives[0].legs()
            if not legs[first] == legs[second]:
                raise ValueError(f"the label '{label}' of '{spec}' names two different spaces in the "
                                 f"operand {at}, and a trace sums a space with itself")
            above_leg = first if first < codomain else second
            below_leg = second if first < codomain else first
            above, below = [], []
            for leg in range(alternatives[0].nlegs()):
                if leg != above_leg and leg != below_leg:
                    (above if leg < codomain else below).append(leg)
            above.append(above_leg)
            below.append(below_leg)
            traced = []
            for alternative in alternatives:
                traced.extend(_partial_trace(alternative.place(above, below)))
            alternatives = traced
            del labels[second]
            del labels[first]
        prepared_ops.append(alternatives)
        prepared_labels.append(labels)
    return prepared_ops, prepared_labels


def einsum(spec: str, *operands: Tensor, outcharge: Optional[Sector] = None) -> Tensor:
    '''`sym::einsum` (`einsum.cpp:444-598`).'''
    if not operands:
        raise ValueError('an einsum needs at least one operand')
    if '->' not in spec:
        raise ValueError(f"the equation '{spec}' does not say what it leaves")
    left, right = spec.split('->', 1)
    groups = [_side_of(text) for text in left.split(',')]
    output = _side_of(right)
    if len(groups) != len(operands):
        raise ValueError(f"the equation '{spec}' names {len(groups)} operands but got {len(operands)}")
    cat = operands[0].cat
    charge = cat.trivial() if outcharge is None else tuple(outcharge)
    for at, operand in enumerate(operands):
        if operand.cat is not cat:
            raise ValueError('an einsum needs every operand over one category')
        if len(groups[at][0]) != operand.nlegs():
            raise ValueError(f"the group {at} of '{spec}' names {len(groups[at][0])} legs but the tensor "
                             f"has {operand.nlegs()}")
    seen = []
    for labels, _ in groups:
        for label in labels:
            if seen.count(label) >= 2:
                raise ValueError(f"the label '{label}' of '{spec}' names more than two legs")
            seen.append(label)
    for label in seen:
        count = seen.count(label)
        if count == 1 and label not in output[0]:
            raise ValueError(f"the label '{label}' of '{spec}' names one leg but is not kept in the output")
        if count == 2 and label in output[0]:
            raise ValueError(f"the label '{label}' of '{spec}' is summed over and still kept")
    for label in output[0]:
        if seen.count(label) != 1:
            raise ValueError(f"the output label '{label}' of '{spec}' does not name a leg of the operands")
    summed = []
    for label in seen:
        if seen.count(label) == 2 and label not in summed:
            summed.append(label)
    laid, inputs = [], []
    for operand, group in zip(operands, groups):
        t, labels = _laid_out(operand, group, summed)
        laid.append(t)
        inputs.append(labels)
    prepared_ops, prepared_labels = _self_contract(laid, inputs, spec)

    above_legs, below_legs = [], []
    for label, up in zip(*output):
        for at, labels in enumerate(prepared_labels):
            if label in labels:
                leg = prepared_ops[at][0].legs()[_position_of(labels, label)]
                (above_legs if up else below_legs).append(leg)
                break
    from .symmetry import reachable_charges
    closes = not output[0]
    for a in reachable_charges(cat, above_legs):
        for b in reachable_charges(cat, below_legs):
            if cat.n_symbol(a, cat.dual(b), charge) > 0:
                closes = True
    if not closes:
        raise ValueError(f"the output legs of '{spec}' cannot close to {render(charge)}")
    combinations = 1
    for alternatives in prepared_ops:
        if not alternatives:
            raise ValueError(f"the traces of '{spec}' leave nothing to sum")
        if len(alternatives) > _MAX_CHARGES // combinations:
            raise ValueError(f"the traces of '{spec}' leave more than {_MAX_CHARGES} combinations")
        combinations *= len(alternatives)
    dims = {}
    for alternatives, labels in zip(prepared_ops, prepared_labels):
        legs = alternatives[0].legs()
        for leg, label in zip(legs, labels):
            dims[label] = _leg_size(cat, leg)
    order = fold_order(prepared_labels, dims)
    total = None
    choice = [0] * len(prepared_ops)
    while True:
        ops = [prepared_ops[at][choice[at]] for at in range(len(choice))]
        for result in _contract_path(prepared_labels, output, ops, order):
            if result.qtotal != charge or not result.blocks:
                continue
            if total is None:
                total = result
            else:
                total += result
        digit = 0
        while digit < len(choice):
            choice[digit] += 1
            if choice[digit] == len(prepared_ops[digit]):
                choice[digit] = 0
                digit += 1
            else:
                break
        if digit == len(choice):
            break
    if total is None:
        raise ValueError(f"no way for the charges of '{spec}' to fuse to {render(charge)}")
    return total


# ----------------------------------------------------------------------------------------------
# Scalars (`tentorch/linalg/graded/compare.cpp`)


def inner(first: Tensor, second: Tensor) -> complex:
    '''`inner` (`compare.cpp:31-43`): Σ_blocks Σ conj(a)·b, no q_dim weight.'''
    if first.cat is not second.cat:
        raise ValueError('a scalar product needs both tensors over one category')
    if not (_same_legs(first.codomain, second.codomain) and _same_legs(first.domain, second.domain)):
        raise ValueError('a scalar product needs both tensors with the same legs')
    if first.qtotal != second.qtotal:
        return 0j
    total = 0j
    for key, b in first.blocks.items():
        total += complex(np.sum(b.conj() * second.blocks[key]))
    return total


def norm(t: Tensor) -> float:
    return float(np.sqrt(abs(inner(t, t))))


def is_close(first: Tensor, second: Tensor, rtol=1e-10, atol=1e-12) -> bool:
    if first.cat is not second.cat or first.qtotal != second.qtotal:
        return False
    if not (_same_legs(first.codomain, second.codomain) and _same_legs(first.domain, second.domain)):
        return False
    if first.blocks.keys() != second.blocks.keys():
        return False
    return all(b.shape == second.blocks[k].shape and np.allclose(b, second.blocks[k], rtol=rtol, atol=atol)
               for k, b in first.blocks.items())


# ----------------------------------------------------------------------------------------------
# Splittings (`tentorch/linalg/graded/svd.cpp`)


def svd(tensor: Tensor, left: Optional[List[int]] = None, max_bond: Optional[int] = None,
        tol: float = 0.0, absorb: int = 0):
    '''`split_over` (`svd.cpp:41-177`): per-block SVD, global ranking, tolerance, cap, and the
    sector/dual re-pairing.  Ties in the ranking are broken by (block, index) order; the reference's
    `std::sort` leaves them unspecified.'''
    if absorb not in (0, 1, 2, 3):
        raise ValueError(f'absorb has to be 0, 1, 2 or 3, got {absorb}')
    if max_bond is not None and max_bond < 1:
        raise ValueError(f'a bond needs at least one multiplet, got {max_bond}')
    arranged = tensor.place(left) if left is not None else tensor
    cat = arranged.cat
    factors = []
    for (r, s, c), m in arranged.blocks.items():
        if m.size == 0:
            continue
        u, v, vh = np.linalg.svd(m, full_matrices=False)
        factors.append({'key': (r, s, c), 'u': u, 'v': v, 'vh': vh, 'kept': 0, 'offset': 0})
    candidates = []
    for at, f in enumerate(factors):
        for index, value in enumerate(f['v']):
            candidates.append((float(value), at, index))
    candidates.sort(key=lambda cnd: -cnd[0])
    cut = len(candidates)
    for i, cnd in enumerate(candidates):
        if cnd[0] <= tol:
            cut = i
            break
    candidates = candidates[:cut]
    if max_bond is not None and len(candidates) > max_bond:
        candidates = candidates[:max_bond]
    for value, at, index in candidates:
        factors[at]['kept'] = max(factors[at]['kept'], index + 1)
    mults: Dict[Sector, int] = {}
    for f in factors:
        if f['kept']:
            s = f['key'][1]
            mults[s] = mults.get(s, 0) + f['kept']
    paired = {}
    for sector, count in mults.items():
        other = mults.get(cat.dual(sector))
        if other is not None:
            paired[sector] = min(count, other)
    if paired and paired != mults:
        for f in factors:
            f['kept'] = 0
        used: Dict[Sector, int] = {}
        for value, at, index in candidates:
            f = factors[at]
            s = f['key'][1]
            room = paired.get(s)
            if room is None or used.get(s, 0) >= room:
                continue
            used[s] = used.get(s, 0) + 1
            f['kept'] = max(f['kept'], index + 1)
        mults = paired
    nxt: Dict[Sector, int] = {}
    for f in factors:
        s = f['key'][1]
        f['offset'] = nxt.get(s, 0)
        nxt[s] = f['offset'] + f['kept']
    placed, run = {}, 0
    for sector in sorted(mults):
        placed[sector] = run
        run += mults[sector]
    if not mults:
        mults = {cat.trivial(): 1}
    bond = Leg(cat, sorted(mults), [mults[s] for s in sorted(mults)])
    first = Tensor(cat, arranged.codomain, [bond], arranged.qtotal, arranged.dtype)
    values = np.zeros(bond.multiplets(), dtype=arranged.dtype)
    second = Tensor(cat, [bond], arranged.domain, cat.trivial(), arranged.dtype)
    for f in factors:
        k = f['kept']
        if not k:
            continue
        r, s, c = f['key']
        u = f['u'][:, :k]
        vh = f['vh'][:k]
        v = f['v'][:k]
        if absorb == 3:
            u = u * np.sqrt(v)
            vh = np.sqrt(v)[:, None] * vh
        elif absorb == 1:
            u = u * v
        elif absorb == 2:
            vh = v[:, None] * vh
        o = f['offset']
        first.block(r, s, c)[:, o:o + k] = u
        second.block(s, s)[o:o + k, :] = vh
        values[placed[s] + o:placed[s] + o + k] = v
    return first, values, second


def svdvals(tensor: Tensor, left=None) -> np.ndarray:
    return svd(tensor, left)[1]


def _split_blocks(tensor: Tensor, pieces):
    '''`split_blocks` (`svd.cpp:196-223`).'''
    cat = tensor.cat
    mults: Dict[Sector, int] = {}
    nxt: Dict[Sector, int] = {}
    placed = []
    for key, lft, rgt, kept in pieces:
        if not kept:
            continue
        s = key[1]
        offset = nxt.get(s, 0)
        nxt[s] = offset + kept
        mults[s] = mults.get(s, 0) + kept
        placed.append((key, lft, rgt, kept, offset))
    if not mults:
        mults = {cat.trivial(): 1}
    bond = Leg(cat, sorted(mults), [mults[s] for s in sorted(mults)])
    first = Tensor(cat, tensor.codomain, [bond], tensor.qtotal, tensor.dtype)
    second = Tensor(cat, [bond], tensor.domain, cat.trivial(), tensor.dtype)
    for (r, s, c), lft, rgt, kept, offset in placed:
        first.block(r, s, c)[:, offset:offset + kept] = lft
        second.block(s, s)[offset:offset + kept, :] = rgt
    return first, second


def qr(tensor: Tensor, left=None):
    '''`qr_over` (`svd.cpp:225-236`).'''
    arranged = tensor.place(left) if left is not None else tensor
    pieces = []
    for key, m in arranged.blocks.items():
        if m.size == 0:
            continue
        q, r = np.linalg.qr(m, mode='reduced')
        pieces.append((key, q, r, q.shape[1]))
    return _split_blocks(arranged, pieces)


def rq(tensor: Tensor, left=None):
    '''`rq_over` (`svd.cpp:238-249`).'''
    arranged = tensor.place(left) if left is not None else tensor
    pieces = []
    for key, m in arranged.blocks.items():
        if m.size == 0:
            continue
        q, r = np.linalg.qr(m.T, mode='reduced')
        pieces.append((key, r.T, q.T, q.shape[1]))
    return _split_blocks(arranged, pieces)
