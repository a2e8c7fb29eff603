'''
The splittings and the Krylov eigensolver of the MI355X backend.

`svd`/`svdvals`/`qr`/`rq` restate `tentorch/linalg/graded/svd.cpp:41-257`: the factorisation of
every block on the device (for the DMRG split, one batched eigendecomposition of the blocks' Gram
matrices by the device library; otherwise one torch factorisation per block), a global ranking of
the singular values on the host with the tolerance, the cap and the sector/dual re-pairing, then the
factors packed along a new bond leg that carries the domain root of each block.  Ties in the ranking are broken by (block,
index) order, a deterministic choice where the reference's `std::sort` leaves it open
(`svd.cpp:79-80`).

`lanczos` restates `tentorch/linalg/lanczos.py:8-64`; the Krylov basis and its scalars stay on the
device (`tt_dot`, `tt_multidot`, `tt_multiaxpy`, `tt_axpby`), with one readback per call.
'''

from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from ._native import engine
from .gram import DECOUPLING, GramSplit, _Stage, choose_tails
from .tensor import BlockKey, Tensor, einsum, einsum_program, inner, inner_device, krylov_steps, norm


def _arranged(tensor: Tensor, left: Optional[Sequence[int]]) -> Tensor:
    return tensor.place(list(left)) if left is not None else tensor


def _bond_leg(cat, mults: dict):
    if not mults:
        mults = {cat.trivial(): 1}
    keys = sorted(mults)
    return engine.Leg(cat, keys, [mults[k] for k in keys])


class SvdSettings:
    '''How `svd` factorises the blocks: 'gram' (the default on the GPU for absorb 1 and 2) takes the
    isometric factor from a batched symmetric eigendecomposition of each block's Gram matrix and the
    other factor as the exact projection of the block on it; 'exact' runs one SVD per block.
    `eigh` picks the Gram eigensolver: 'device' (this library's batched tridiagonalisation, divide
    and conquer and back-transformation for all blocks at once, orders up to 2048; larger Gram
    matrices go to torch) or 'torch' (`torch.linalg.eigh`).  `resolved`: a Gram pass's values are
    trusted down to this fraction of its largest; below, when the cut can reach there, the rows under
    `reliable` x the largest go through a further pass (see `tentorch_amd.gram`).'''
    method = 'gram'
    eigh = 'device'
    reliable = 1e-5
    resolved = 1e-7


def _eigenrows(packed: torch.Tensor, sizes, offsets):
    '''Eigenvectors of the symmetric matrices packed in `packed` (matrix b: sizes[b]^2 entries at
    offsets[b]) as the rows of one n x n matrix per problem, in ascending eigenvalue order: one call
    of the device library's batched solver for every matrix up to its order limit, torch beyond it
    (or for all of them under `SvdSettings.eigh = 'torch'`).  Nothing is read back.'''
    rows = [None] * len(sizes)
    device = SvdSettings.eigh == 'device'
    small = [b for b, g in enumerate(sizes) if device and g <= _native.EIGH_MAX_ORDER]
    if small:
        _, vectors = _native.eigh_batched(packed, [sizes[b] for b in small], [offsets[b] for b in small])
        for b in small:
            g, at = sizes[b], offsets[b]
            rows[b] = vectors[at:at + g * g].view(g, g)
    for b, g in enumerate(sizes):
        if rows[b] is None:
            at = offsets[b]
            _, columns = torch.linalg.eigh(packed[at:at + g * g].view(g, g))
            rows[b] = columns.transpose(0, 1)
    return rows


def _negated_grams(sides):
    '''-C C^T of every (g x n) matrix C, packed back to back in one buffer.'''
    sizes = [c.shape[0] for c in sides]
    offsets = np.concatenate([[0], np.cumsum([g * g for g in sizes])]).astype(np.int64).tolist()
    packed = torch.empty(max(offsets[-1], 1), dtype=torch.float64, device=sides[0].device)
    for c, g, at in zip(sides, sizes, offsets):
        gram = packed[at:at + g * g].view(g, g)
        torch.addmm(gram, c, c.transpose(0, 1), beta=0.0, alpha=-1.0, out=gram)
    return packed, sizes, offsets[:-1]


def _row_norms(matrices) -> torch.Tensor:
    '''The Euclidean norm of every row of every matrix, concatenated, on the device.'''
    return torch.cat([torch.linalg.vector_norm(m, dim=1) for m in matrices])


def _rotated(x: torch.Tensor, head: torch.Tensor, tail: torch.Tensor):
    '''(head, tail) rows under the rotation exp([[0, X^T], [-X, 0]]) through the third order, as
    `tentorch_amd.gram` applies it: with K = X^T X and U = X^T tail,
    head' = (1 - K/2) head + (1 - K/6) U and tail' = tail - X ((1 - K/6) head + U/2).'''
    xt = x.transpose(0, 1)
    k = torch.matmul(xt, x)
    eye = torch.eye(k.shape[0], dtype=k.dtype, device=k.device)
    a, b = eye - k / 2, eye - k / 6
    u = torch.matmul(xt, tail)
    new_head = torch.matmul(a, head) + torch.matmul(b, u)
    new_tail = tail - torch.matmul(x, torch.matmul(b, head) + u / 2)
    return new_head, new_tail


def _gram_split(blocks, left_isometric: bool, cut):
    '''The passes of `tentorch_amd.gram.GramSplit` (Gram eigenvectors, projected row norms,
    decoupled tail passes; its module text gives the numerics) in torch operations per block: the
    path for `SvdSettings.eigh = 'torch'` and for Gram orders above the device eigensolver's limit.
    With C = B (`left_isometric`) or B^T, returns per block (Q rows [r x g], P = Q C rows [r x n], the
    singular values on the host), r = min(g, n), in descending value order.'''
    sides = [b if left_isometric else b.transpose(0, 1) for b in blocks]
    packed, sizes, offsets = _negated_grams(sides)
    q = [rows.clone() for rows in _eigenrows(packed, sizes, offsets)]
    p = [torch.matmul(qb, cb) for qb, cb in zip(q, sides)]
    ranks = [min(c.shape) for c in sides]
    norms = _row_norms(p).cpu().numpy()
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    values = [norms[starts[b]:starts[b + 1]].copy() for b in range(len(sides))]    # all g rows
    first = [0] * len(sides)          # the first row of every block's current pass
    while True:
        tails = choose_tails(values, ranks, first, cut([v[:r] for v, r in zip(values, ranks)]),
                             SvdSettings.reliable, SvdSettings.resolved)
        if not tails:
            break
        for b, h in tails:
            gaps = values[b][first[b]:h][None, :] ** 2 - values[b][h:][:, None] ** 2
            scale = torch.as_tensor(1.0 / gaps, device=p[b].device)
            for _ in range(DECOUPLING):
                head, tail = p[b][first[b]:h], p[b][h:]
                x = torch.matmul(tail, head.transpose(0, 1)) * scale
                p[b] = torch.cat([p[b][:first[b]], *_rotated(x, head, tail)])
                q[b] = torch.cat([q[b][:first[b]], *_rotated(x, q[b][first[b]:h], q[b][h:])])
        tail_p = [p[b][h:] for b, h in tails]
        packed, t_sizes, t_offsets = _negated_grams(tail_p)
        for (b, h), w in zip(tails, _eigenrows(packed, t_sizes, t_offsets)):
            p[b] = torch.cat([p[b][:h], torch.matmul(w, p[b][h:])])
            q[b] = torch.cat([q[b][:h], torch.matmul(w, q[b][h:])])
        tail_norms = _row_norms([p[b][h:] for b, h in tails]).cpu().numpy()
        at = 0
        for b, h in tails:
            n = values[b].size - h
            values[b][h:] = tail_norms[at:at + n]
            at += n
            first[b] = h
    out = []
    for b, v in enumerate(values):
        v = v[:ranks[b]]
        qb, pb = q[b][:ranks[b]], p[b][:ranks[b]]
        order = np.argsort(-v, kind='stable')
        if np.any(order != np.arange(v.size)):
            index = torch.as_tensor(order, device=pb.device)
            qb, pb, v = qb[index], pb[index], v[order]
        out.append((qb, pb, v))
    return out


def _is_index(value) -> bool:
    return isinstance(value, (int, np.integer)) and not isinstance(value, bool)


def _bind_split_arguments(args, kwargs):
    '''The two call forms of the reference's `svd` binding (`bindings.cpp:367-382`), resolved as its
    overloads are, first form first:
      svd(tensor, max_bond=None, tol=0.0, absorb=0)
      svd(tensor, left, max_bond=None, tol=0.0, absorb=0)
    Returns (left, max_bond, tol, absorb); a call that fits neither form raises TypeError.'''

    def bind(names, defaults):
        if len(args) > len(names):
            return None
        bound = dict(zip(names, args))
        for name, value in kwargs.items():
            if name not in names or name in bound:
                return None
            bound[name] = value
        for name, value in defaults.items():
            bound.setdefault(name, value)
        if any(name not in bound for name in names):
            return None
        if bound['max_bond'] is not None and not _is_index(bound['max_bond']):
            return None
        if not isinstance(bound['tol'], (int, float, np.floating, np.integer)) or isinstance(bound['tol'], bool):
            return None
        if not _is_index(bound['absorb']):
            return None
        if 'left' in bound:
            left = bound['left']
            if left is None or isinstance(left, (str, bytes)) or not all(_is_index(i) for i in left):
                return None
        return bound

    first = bind(('max_bond', 'tol', 'absorb'), {'max_bond': None, 'tol': 0.0, 'absorb': 0})
    if first is not None:
        return None, first['max_bond'], float(first['tol']), int(first['absorb'])
    try:
        second = bind(('left', 'max_bond', 'tol', 'absorb'), {'max_bond': None, 'tol': 0.0, 'absorb': 0})
    except TypeError:
        second = None
    if second is not None:
        return [int(i) for i in second['left']], second['max_bond'], float(second['tol']), int(second['absorb'])
    raise TypeError('svd(): incompatible arguments; the supported forms are svd(tensor, max_bond=None, tol=0.0, '
                    'absorb=0) and svd(tensor, left, max_bond=None, tol=0.0, absorb=0)')


def svd(tensor: Tensor, *args, **kwargs) -> Tuple[Tensor, torch.Tensor, Tensor]:
    '''`split_over` (`svd.cpp:41-177`) under the reference binding's two call forms
    (`bindings.cpp:367-382`): `svd(tensor, max_bond=None, tol=0.0, absorb=0)` splits the tensor as its
    sides stand, `svd(tensor, left, max_bond=None, tol=0.0, absorb=0)` first brings the legs `left`
    to the codomain.  Returns (U, values, V) with U over (codomain | bond) and V over (bond | domain).'''
    left, max_bond, tol, absorb = _bind_split_arguments(args, kwargs)
    return _split(tensor, left, max_bond, tol, absorb)[:3]


def _split(tensor: Tensor, left: Optional[Sequence[int]], max_bond: Optional[int], tol: float, absorb: int,
           device_values: bool = True):
    '''`svd`, plus the kept singular values on the host (FP64, in the order of `values`), which the
    split reads back anyway to rank them: a caller that needs their sum needs no further
    synchronisation.  `device_values` False leaves `values` None for a caller that uses only the host
    copy (no upload).'''
    if absorb not in (0, 1, 2, 3):
        raise ValueError(f'absorb has to be 0, 1, 2 or 3, got {absorb}')
    if max_bond is not None and max_bond < 1:
        raise ValueError(f'a bond needs at least one multiplet, got {max_bond}')
    arranged = _arranged(tensor, left)
    cat = arranged.cat
    factors = []
    # (key, rows, cols, element offset) of every non-empty block, without building views
    blocks = arranged.structure.blocks()
    nonempty = [i for i, (_, rows, cols, _) in enumerate(blocks) if rows * cols]
    layout = [(BlockKey(*blocks[i][0]),) + tuple(blocks[i][1:]) for i in nonempty]
    gram = SvdSettings.method == 'gram' and absorb in (1, 2) and not arranged.dtype.is_complex
    device_split = None

    def views():
        return [(key, arranged.data[off:off + rows * cols].view(rows, cols)) for key, rows, cols, off in layout]

    if gram and layout:
        # absorb 2 keeps U isometric (values go right), absorb 1 keeps V isometric; the singular
        # values arrive on the host, each block's in descending order
        def floor(values):
            # the smallest value the ranking below can keep: the tolerance, or the first value the cap
            # drops (a block's tail that lies wholly below it is dropped whatever its exact values)
            ranked = np.sort(np.concatenate(values))[::-1]
            alive = ranked[ranked > tol]
            return max(tol, float(alive[max_bond])) if max_bond is not None and alive.size > max_bond else tol

        on_device = SvdSettings.eigh == 'device' and max(
            (rows if absorb == 2 else cols) for _, rows, cols, _ in layout) <= _native.EIGH_MAX_ORDER
        if on_device:
            device_split = GramSplit.of_buffer(arranged.data, [x[1] for x in layout], [x[2] for x in layout],
                                               [x[3] for x in layout], absorb == 2)
            for (key, _, _, _), v in zip(layout, device_split.factor(floor, SvdSettings.reliable, SvdSettings.resolved)):
                factors.append({'key': key, 'v': v, 'kept': 0, 'offset': 0})
        else:
            items = views()
            for (key, _), (q, projected, v) in zip(items, _gram_split([b for _, b in items], absorb == 2, floor)):
                factors.append({'key': key, 'q': q, 'p': projected, 'v': v, 'kept': 0, 'offset': 0})
        host = np.concatenate([f['v'] for f in factors])
        sizes = [f['v'].size for f in factors]
    else:
        for key, block in views():
            u, v, vh = torch.linalg.svd(block, full_matrices=False)
            factors.append({'key': key, 'u': u, 'v': v, 'vh': vh, 'kept': 0, 'offset': 0})
        host = torch.cat([f['v'] for f in factors]).double().cpu().numpy() if factors else np.zeros(0)
        sizes = [f['v'].numel() for f in factors]
    # The ranking keeps a leading run of every block, which needs each block's values descending
    # (svd.cpp:80).  torch.linalg.svd on ROCm can return values at the rounding level (~1e-14 of
    # the largest) slightly out of order; such a block is re-sorted with its singular vectors.
    if not gram:
        at = 0
        for f, size in zip(factors, sizes):
            segment = host[at:at + size]
            if size > 1 and np.any(segment[1:] > segment[:-1]):
                order = np.argsort(-segment, kind='stable')
                index = torch.as_tensor(order, device=f['v'].device)
                f['u'], f['v'], f['vh'] = f['u'][:, index], f['v'][index], f['vh'][index]
                host[at:at + size] = segment[order]
            at += size
    # the global ranking (svd.cpp:80-119): candidates largest first, ties in (block, index) order,
    # cut at the tolerance and the cap, dual sectors re-paired; the engine returns every block's kept
    # count and bond offset, the bond and the kept values in bond order
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64) if factors else np.zeros(1, dtype=np.int64)
    kept, offsets, bond_sectors, bond_mults, kept_values = engine.rank_split(
        arranged.structure, nonempty, host, starts.tolist(), -1 if max_bond is None else max_bond, tol)
    for fi, f in enumerate(factors):
        f['kept'], f['offset'] = kept[fi], offsets[fi]
    placed = dict(zip(bond_sectors, np.concatenate([[0], np.cumsum(bond_mults)[:-1]]).tolist()))
    bond = engine.Leg(cat, bond_sectors, bond_mults)
    first = Tensor(cat, arranged.codomain, [bond], arranged.qtotal, arranged.dtype, arranged.device)
    second = Tensor(cat, [bond], arranged.domain, cat.trivial(), arranged.dtype, arranged.device)
    values = None if gram else torch.zeros(bond.multiplets(), dtype=arranged.dtype, device=arranged.device)
    placements = []
    if gram and device_split is not None:
        # the blocks' places in the two new buffers, read once from their structures
        first_at = {key: (rows, cols, off) for key, rows, cols, off in first.structure.blocks()}
        second_at = {key[0]: (rows, cols, off) for key, rows, cols, off in second.structure.blocks()}
        first_base, second_base = first.data.data_ptr(), second.data.data_ptr()
        spans = []      # the byte range of every block a placement writes
    for fi, f in enumerate(factors):
        k = f['kept']
        if not k:
            continue
        r, s, c = f['key']
        o = f['offset']
        if gram and device_split is not None:
            # (block, kept, first block's address of column o, its row length, second block's address
            # of row o, its row length)
            frows, fcols, foff = first_at[(r, s, c)]
            srows, scols, soff = second_at[s]
            placements.append((fi, k, first_base + 8 * (foff + o), fcols, second_base + 8 * (soff + o * scols), scols))
            spans += [(first_base + 8 * foff, first_base + 8 * (foff + frows * fcols)),
                      (second_base + 8 * soff, second_base + 8 * (soff + srows * scols))]
            continue
        if gram:
            # the isometric factor is Q's leading rows, the other the projection P = Q C already
            # formed: two copies per block
            q, projected = f['q'][:k], f['p'][:k]
            if absorb == 2:
                first.block(r, s, c)[:, o:o + k].copy_(q.transpose(0, 1))
                second.block(s, s)[o:o + k, :].copy_(projected)
            else:
                first.block(r, s, c)[:, o:o + k].copy_(projected.transpose(0, 1))
                second.block(s, s)[o:o + k, :].copy_(q)
            continue
        v = f['v'][:k]
        u, vh = f['u'][:, :k], f['vh'][:k]
        if absorb == 3:
            u = u * v.sqrt()
            vh = v.sqrt().unsqueeze(1) * vh
        elif absorb == 1:
            u = u * v
        elif absorb == 2:
            vh = v.unsqueeze(1) * vh
        first.block(r, s, c)[:, o:o + k] = u
        second.block(s, s)[o:o + k, :] = vh
        values[placed[s] + o:placed[s] + o + k] = v.to(arranged.dtype)
    if gram and device_split is not None:
        device_split.assemble(placements, spans)
    if gram and device_values:
        # one copy from pinned memory, queued behind the products without waiting for them
        staged = torch.from_numpy(kept_values).to(arranged.dtype)
        if arranged.device.type == 'cuda':
            values = staged.pin_memory().to(arranged.device, non_blocking=True)
        else:
            values = staged
    return first, values, second, kept_values


def svdvals(tensor: Tensor, left: Optional[Sequence[int]] = None) -> torch.Tensor:
    '''`svdvals` (`svd.cpp:179-187`, binding `bindings.cpp:383-385`): every singular value, no cut.'''
    return _split(tensor, None if left is None else list(left), None, 0.0, 0)[1]


def _bond_factors(tensor: Tensor, pieces):
    """The two factors of `split_blocks` (`svd.cpp:196-223`), zero-filled: `pieces` lists (key, kept)
    per block in block order; returns the two tensors and, per piece, the column offset of its
    multiplets among those of its domain sector on the new bond."""
    cat = tensor.cat
    mults, nxt, offsets = {}, {}, []
    for key, kept in pieces:
        s = key.domain
        offsets.append(nxt.get(s, 0))
        if kept:
            nxt[s] = offsets[-1] + kept
            mults[s] = mults.get(s, 0) + kept
    bond = _bond_leg(cat, mults)
    first = Tensor(cat, tensor.codomain, [bond], tensor.qtotal, tensor.dtype, tensor.device)
    second = Tensor(cat, [bond], tensor.domain, cat.trivial(), tensor.dtype, tensor.device)
    return first, second, offsets


def _triangular_split(tensor: Tensor, left: Optional[Sequence[int]], mirrored: bool) -> Tuple[Tensor, Tensor]:
    """`qr_over` (`svd.cpp:225-236`) and, `mirrored`, `rq_over` (`svd.cpp:238-249`): each block B
    factored as Q R (qr: B = Q R, Q into the first factor; rq: B^T = Q R, so B = R^T Q^T with R^T into
    the first factor and Q^T into the second), k = min(rows, cols) multiplets per block on the bond.
    Real FP64 blocks go through one launch of the library's batched Householder QR, which writes
    straight into the two factors; complex blocks take `torch.linalg.qr` per block on the device."""
    arranged = _arranged(tensor, left)
    items = [(key, block) for key, block in arranged.blocks().items() if block.numel() != 0]
    kept = [min(block.shape) for _, block in items]
    first, second, offsets = _bond_factors(arranged, [(key, k) for (key, _), k in zip(items, kept)])
    if not items:
        return first, second
    if arranged.dtype.is_complex:
        for (key, block), o, k in zip(items, offsets, kept):
            r, s, c = key
            if mirrored:
                q, rr = torch.linalg.qr(block.transpose(0, 1), mode='reduced')
                lft, rgt = rr.transpose(0, 1), q.transpose(0, 1)
            else:
                lft, rgt = torch.linalg.qr(block, mode='reduced')
            first.block(r, s, c)[:, o:o + k] = lft
            second.block(s, s)[o:o + k, :] = rgt
        return first, second
    if arranged.dtype != torch.float64:
        raise ValueError(f'qr and rq run on float64 or complex128 tensors, got {arranged.dtype}')
    count = len(items)
    rows = np.array([b.shape[0] for _, b in items], dtype=np.int64)
    cols = np.array([b.shape[1] for _, b in items], dtype=np.int64)
    m, n = (cols, rows) if mirrored else (rows, cols)
    work_at = np.concatenate([[0], np.cumsum(m * n + np.minimum(m, n))]).astype(np.int64)
    work = torch.empty(max(int(work_at[-1]), 1), dtype=torch.float64, device=arranged.device)
    table = np.zeros(count, dtype=_native.QR_PROBLEM)
    lefts = [first.block(key.codomain, key.domain, key.copy) for key, _ in items]
    rights = [second.block(key.domain, key.domain) for key, _ in items]
    left_ptr = np.array([b.data_ptr() + 8 * o for b, o in zip(lefts, offsets)], dtype=np.int64)
    right_ptr = np.array([b.data_ptr() + 8 * o * b.shape[1] for b, o in zip(rights, offsets)], dtype=np.int64)
    left_ld = np.array([b.shape[1] for b in lefts], dtype=np.int64)
    right_ld = np.array([b.shape[1] for b in rights], dtype=np.int64)
    table['src'] = [b.data_ptr() for _, b in items]
    table['work'] = work.data_ptr() + 8 * work_at[:-1]
    table['m'], table['n'], table['src_ld'] = m, n, cols
    table['trans'] = 1 if mirrored else 0
    # qr: Q (rows x k) is the first factor's columns, R (k x cols) the second's rows; rq: R^T
    # (rows x k) is the first factor's columns, Q^T (k x cols) the second's rows
    table['q'], table['q_ld'] = (right_ptr, right_ld) if mirrored else (left_ptr, left_ld)
    table['r'], table['r_ld'] = (left_ptr, left_ld) if mirrored else (right_ptr, right_ld)
    table['q_trans'] = table['r_trans'] = 1 if mirrored else 0
    extents = _native.Extents()
    for t in (work, first.data, second.data, arranged.data):
        extents.add(t)
    _native.check_qr(table, extents)
    stage = _Stage(arranged.device, [table.nbytes])
    stage.put(0, table)
    stage.send()
    # the workspace and the staged table are stream-ordered: freed after this, they are reused only by
    # work queued behind the launch
    _native.qr_batched(table, stage.address(0))
    return first, second


def qr(tensor: Tensor, left: Optional[Sequence[int]] = None) -> Tuple[Tensor, Tensor]:
    """`qr_over` (`svd.cpp:225-236`)."""
    return _triangular_split(tensor, left, False)


def rq(tensor: Tensor, left: Optional[Sequence[int]] = None) -> Tuple[Tensor, Tensor]:
    """`rq_over` (`svd.cpp:238-249`)."""
    return _triangular_split(tensor, left, True)


class LanczosSettings:
    '''`device` runs the iteration with the Krylov basis and its scalars, the start norm included, on
    the GPU (one host synchronisation per call); `host` follows the reference's loop with one
    synchronisation per scalar product.'''
    mode = 'device'


def lanczos(action: Callable[[Tensor], Tensor], start: Tensor, n_iter: int = 5,
            tol: float = 1e-12) -> Tuple[float, Tensor]:
    '''
    `lanczos` (`tentorch/linalg/lanczos.py:8-64`): n_iter Krylov steps with two full
    reorthogonalisation passes, the tridiagonal solved in FP64 on the host, the Ritz vector normalised.
    '''
    if n_iter < 1:
        raise ValueError(f'Lanczos needs at least one step, not {n_iter}.')
    if LanczosSettings.mode == 'device' and start.dtype == torch.float64 and n_iter + 1 <= 32:
        return _lanczos_device(action, start, n_iter, tol)
    return _lanczos_host(action, start, n_iter, tol)


class EinsumAction:
    '''The action v -> einsum(spec, *operands with v in position `slot`) as a callable, which a
    Lanczos run recognises: its Krylov steps then go to the engine in one call (`krylov_steps`)
    instead of one einsum and one orthonormalisation per step from Python.'''

    def __init__(self, spec: str, operands: Sequence[Tensor], slot: int):
        self.spec, self.operands, self.slot = spec, list(operands), slot

    def with_vector(self, vector: Tensor) -> List[Tensor]:
        out = list(self.operands)
        out.insert(self.slot, vector)
        return out

    def __call__(self, vector: Tensor) -> Tensor:
        return einsum(self.spec, *self.with_vector(vector))


def _krylov_basis(action, start: Tensor, n_iter: int):
    '''The Krylov basis V (n_iter + 1 rows) and the device scalars of a Lanczos run.  `scalars` holds
    the alphas, the betas, each step's scratch (2 (k+1) + 1), the Ritz coefficients (n_iter + 1) and
    six work scalars; work[2] is the squared norm of the start.  Each Krylov step is the action, then
    one library call (`tt_lanczos_orthonormalize`).'''
    size_n = start.structure.size
    device = start.device
    V = torch.empty((n_iter + 1, size_n), dtype=torch.float64, device=device)
    scalars = torch.zeros(5 * n_iter + 10, dtype=torch.float64, device=device)
    alphas, betas = scalars[:n_iter], scalars[n_iter:2 * n_iter]
    scratch = scalars[2 * n_iter:4 * n_iter + 3]
    work = scalars[5 * n_iter + 4:]
    # v_0 = start / |start| with the norm kept on the device (the dot of `norm`, then a correctly
    # rounded square root and reciprocal, as on the host); a zero start is refused once the
    # scalars are read back
    _native.dot(torch.float64, start.data, start.data, work[2:3])
    torch.reciprocal(torch.sqrt(work[2:3]), out=work[3:4])
    _native.axpby(torch.float64, 0.0, start.data, 0.0, V[0], alpha_dev=work[3:4])
    if isinstance(action, EinsumAction) and n_iter + 1 <= 32:
        operands = action.with_vector(start)
        program = einsum_program(action.spec, operands)
        if program.out == start.structure:
            image = torch.empty(size_n, dtype=torch.float64, device=device)
            if krylov_steps(program, operands, action.slot, V, n_iter, image, scratch, alphas, betas):
                return V, scalars, work
    for k in range(n_iter):
        vector = Tensor._wrap(start.structure, V[k])
        image = action(vector)
        if image.structure != start.structure:
            raise ValueError('the action has to map the legs of the start vector to themselves')
        w = image.data if image.data.data_ptr() != V[k].data_ptr() else image.data.clone()
        _native.lanczos_orthonormalize(V, k, w, scratch, alphas[k:k + 1], betas[k:k + 1])
    return V, scalars, work


def _lanczos_device(action, start: Tensor, n_iter: int, tol: float) -> Tuple[float, Tensor]:
    '''The same recurrence with every scalar on the device.  Each Krylov step: the action, then one
    library call (`tt_lanczos_orthonormalize`): two classical Gram-Schmidt passes of w = Hv_k against
    the whole basis, alpha = its projections on v_k (<v_k, Hv_k> plus the second pass's correction),
    beta = |w|, v_{k+1} = w / beta.  The reference subtracts alpha v_k + beta_{k-1} v_{k-1} first and
    then runs two modified Gram-Schmidt passes (`lanczos.py:36-45`); the first full pass removes the
    same components, so the two agree up to rounding.  The alphas and betas are read back once; the
    break at beta < tol (`lanczos.py:46-48`) is applied to what was read.'''
    size_n = start.structure.size
    device = start.device
    V, scalars, work = _krylov_basis(action, start, n_iter)
    coef = scalars[4 * n_iter + 3:5 * n_iter + 4]
    host = scalars.cpu().numpy()
    if host[5 * n_iter + 6] == 0.0:
        raise ValueError('Lanczos needs a starting vector of non-zero norm.')
    a, b = host[:n_iter], host[n_iter:2 * n_iter]
    size = n_iter
    for k in range(n_iter):
        if b[k] < tol:
            size = k + 1
            break
    matrix = np.zeros((size, size))
    for step in range(size):
        matrix[step, step] = a[step]
    for step in range(size - 1):
        matrix[step, step + 1] = matrix[step + 1, step] = b[step]
    values, vectors = np.linalg.eigh(matrix)
    # from pinned memory, queued without waiting (a pageable copy would synchronise)
    coef[:size].copy_(torch.from_numpy(vectors[:, 0].copy()).pin_memory(), non_blocking=True)
    out = torch.zeros(size_n, dtype=torch.float64, device=device)
    _native.multiaxpy(V, size, coef, 1.0, out)
    _native.dot(torch.float64, out, out, work[0:1])
    torch.rsqrt(work[0:1], out=work[1:2])
    _native.axpby(torch.float64, 0.0, out, 0.0, out, alpha_dev=work[1:2])
    return float(values[0]), Tensor._wrap(start.structure, out)


class RitzValue:
    '''A Ritz value that may still be on the device.  `stage()` queues its copy (with the squared
    norm of the run's start vector) to pinned host memory behind an event; `get()` waits for that
    event, refuses a zero start as `lanczos` does, and returns the value.'''

    def __init__(self, device_scalars: Optional[torch.Tensor] = None, host_value: Optional[float] = None):
        self._device = device_scalars      # [ritz value, squared start norm]
        self._value = host_value
        self._staged = None
        self._ready = None

    def stage(self) -> 'RitzValue':
        if self._device is not None and self._staged is None:
            self._staged = torch.empty(2, dtype=torch.float64, pin_memory=True)
            self._staged.copy_(self._device, non_blocking=True)
            self._ready = torch.cuda.Event()
            self._ready.record()
        return self

    def get(self) -> float:
        if self._value is None:
            self.stage()
            self._ready.synchronize()
            host = self._staged.numpy()
            if host[1] == 0.0:
                raise ValueError('Lanczos needs a starting vector of non-zero norm.')
            self._value = float(host[0])
        return self._value


def lanczos_deferred(action: Callable[[Tensor], Tensor], start: Tensor, n_iter: int = 5,
                     tol: float = 1e-12) -> Tuple[RitzValue, Tensor]:
    '''`lanczos` with no host synchronisation: the tridiagonal is solved on the device
    (`tt_lanczos_ritz`), the Ritz vector formed from the basis rows before the stop
    (`tt_multiaxpy_counted`) and normalised there, and the Ritz value left on the device until the
    caller reads it (after a synchronisation it needs anyway).  Same recurrence as `lanczos`; the
    Ritz vector's sign makes its overlap with the start non-negative.  Runs that the device path does
    not take (complex vectors, more than 31 steps, host mode) run `lanczos` and return its value.'''
    if n_iter < 1:
        raise ValueError(f'Lanczos needs at least one step, not {n_iter}.')
    if not (LanczosSettings.mode == 'device' and start.dtype == torch.float64 and n_iter + 1 <= 32):
        value, vector = lanczos(action, start, n_iter, tol)
        return RitzValue(host_value=value), vector
    V, scalars, work = _krylov_basis(action, start, n_iter)
    alphas, betas = scalars[:n_iter], scalars[n_iter:2 * n_iter]
    coef = scalars[4 * n_iter + 3:5 * n_iter + 4]
    size = torch.empty(1, dtype=torch.int32, device=start.device)
    _native.lanczos_ritz(n_iter, alphas, betas, tol, coef, work[4:5], size)
    out = torch.zeros(start.structure.size, dtype=torch.float64, device=start.device)
    _native.multiaxpy(V, n_iter, coef, 1.0, out, count=size)
    _native.dot(torch.float64, out, out, work[0:1])
    torch.rsqrt(work[0:1], out=work[1:2])
    _native.axpby(torch.float64, 0.0, out, 0.0, out, alpha_dev=work[1:2])
    # [ritz value, squared start norm] side by side for one copy to the host
    work[5:6].copy_(work[2:3])
    return RitzValue(device_scalars=work[4:6]), Tensor._wrap(start.structure, out)


def _lanczos_host(action, start: Tensor, n_iter: int, tol: float) -> Tuple[float, Tensor]:
    '''The reference's loop with host scalars (`lanczos.py:27-64`).'''
    scale = norm(start)
    if scale == 0.0:
        raise ValueError('Lanczos needs a starting vector of non-zero norm.')
    basis = [start * (1.0 / scale)]
    alphas: List[float] = []
    betas: List[float] = []
    for step in range(n_iter):
        image = action(basis[-1])
        alphas.append(inner(basis[-1], image).real)
        image -= basis[-1] * alphas[-1]
        if step > 0:
            image -= basis[-2] * betas[-1]
        for _ in range(2):
            for vector in basis:
                image -= vector * inner(vector, image)
        beta = norm(image)
        if beta < tol:
            break
        betas.append(beta)
        basis.append(image * (1.0 / beta))
    size = len(alphas)
    matrix = np.zeros((size, size))
    for step in range(size):
        matrix[step, step] = alphas[step]
    for step, beta in enumerate(betas[:size - 1]):
        matrix[step, step + 1] = beta
        matrix[step + 1, step] = beta
    values, vectors = np.linalg.eigh(matrix)
    coefficients = vectors[:, 0]
    vector = basis[0] * float(coefficients[0])
    for step in range(1, size):
        vector += basis[step] * float(coefficients[step])
    return float(values[0]), vector * (1.0 / norm(vector))
