'''
Oracle check of the checker: the dense restatement of abelian block-sparse tensors.

TEST INFRASTRUCTURE ONLY.

For an abelian category every vertex is the number 1 (`tentorch/symmetry/cat/abelian/abelian.cpp:49-54`),
so the entry of a block at (codomain tree, copies; domain tree, copies) is a plain array entry.  Each
leg gets a dense axis of size `multiplets()`, sectors in the leg's sorted order and copies inside a
sector.  A domain leg is read in the codomain frame: its tree sector `c` names the dense index of
`dual(c)`, because a bend lands a leg carrying the dual sector (`regroup.cpp:355-358`).  In that frame
`place` is a transpose, and a contraction pairs the sector `c` of one leg with `dual(c)` of the other.
'''

from typing import List, Sequence

import numpy as np

from .symmetry import Leg, fusion_trees
from .tensor import Tensor


def _leg_offsets(leg: Leg):
    out, at = {}, 0
    for s, m in zip(leg.sectors, leg.multiplicities):
        out[s] = at
        at += m
    return out


def dense(t: Tensor) -> np.ndarray:
    '''The dense array of an abelian tensor, every leg in the codomain frame.  Each (codomain tree,
    domain tree) pair of a block is a box of the dense array (one index range per leg, copies
    mixed-radix with the last leg fastest), written in one slice assignment.'''
    legs = t.legs()
    ncod = len(t.codomain)
    out = np.zeros([leg.multiplets() for leg in legs], dtype=t.dtype)
    offs = [_leg_offsets(leg) for leg in legs]
    for (r, s, c), block in t.blocks.items():
        row_boxes = _boxes(t.cat, t.codomain, r, offs[:ncod], flip=False)
        col_boxes = _boxes(t.cat, t.domain, s, offs[ncod:], flip=True)
        row_at = 0
        for row_slices, row_shape in row_boxes:
            nr = int(np.prod(row_shape)) if row_shape else 1
            col_at = 0
            for col_slices, col_shape in col_boxes:
                nc = int(np.prod(col_shape)) if col_shape else 1
                part = block[row_at:row_at + nr, col_at:col_at + nc]
                out[tuple(row_slices + col_slices)] = part.reshape(tuple(row_shape) + tuple(col_shape))
                col_at += nc
            row_at += nr
    return out


def _boxes(cat, legs: Sequence[Leg], root, offs, flip: bool):
    '''The rows of one side as boxes, in block row order: one (slices, extents) per fusion tree.'''
    if not legs:
        return [([], [])]
    out = []
    for tree in fusion_trees(cat, legs, root):
        slices, extents = [], []
        for leg, sector, off in zip(legs, tree.sectors, offs):
            start = off[cat.dual(sector) if flip else sector]
            m = leg.mult_of(sector)
            slices.append(slice(start, start + m))
            extents.append(m)
        out.append((slices, extents))
    return out


def dense_by_entry(t: Tensor) -> np.ndarray:
    '''`dense` written entry by entry (the definition the box form is checked against).'''
    legs = t.legs()
    ncod = len(t.codomain)
    shape = [leg.multiplets() for leg in legs]
    out = np.zeros(shape, dtype=t.dtype)
    offs = [_leg_offsets(leg) for leg in legs]
    cat = t.cat
    for (r, s, c), block in t.blocks.items():
        row_entries = _entries(cat, t.codomain, r, offs[:ncod], flip=False)
        col_entries = _entries(cat, t.domain, s, offs[ncod:], flip=True)
        for ri, rix in enumerate(row_entries):
            for ci, cix in enumerate(col_entries):
                out[rix + cix] = block[ri, ci]
    return out


def _entries(cat, legs: Sequence[Leg], root, offs, flip: bool) -> List[tuple]:
    '''Dense index tuples of the rows of one side, in block row order (trees in `fusion_trees`
    order, then the copies mixed-radix with the last leg fastest, `layout.cpp:18-32`).'''
    import itertools
    rows = []
    for tree in fusion_trees(cat, legs, root):
        ranges = []
        for leg, sector, off in zip(legs, tree.sectors, offs):
            frame = cat.dual(sector) if flip else sector
            m = leg.mult_of(sector)
            ranges.append(range(off[frame], off[frame] + m))
        for idx in itertools.product(*ranges):
            rows.append(idx)
    if not legs:
        rows = [()]
    return rows


def flip(cat, array: np.ndarray, axis: int, leg: Leg) -> np.ndarray:
    '''Relabel one axis of a dense array by the dual of each sector: the index of `c` goes to the
    index of `dual(c)` (both carry the same multiplicity on a dual-symmetric leg).'''
    offs = _leg_offsets(leg)
    perm = np.zeros(leg.multiplets(), dtype=np.int64)
    for s, m in zip(leg.sectors, leg.multiplicities):
        d = cat.dual(s)
        perm[offs[d]:offs[d] + m] = np.arange(offs[s], offs[s] + m)
    return np.take(array, perm, axis=axis)
