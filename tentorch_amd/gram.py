'''
The split's factorisation on the device: the singular triplets of every block of a graded split
(`tentorch/linalg/graded/svd.cpp:41-177`, which runs `at::linalg_svd` per block) from eigendecompositions
of Gram matrices, to the accuracy of an SVD, in a few batched launches whatever the block count.

Per block B (absorb 2: C = B, the left factor isometric; absorb 1: C = B^T), all blocks at once:

- pass 1: G = -C C^T (`tt_bgemm`), its eigenvectors Q as rows (`tt_eigh_batched`, descending
  singular value first), P = Q C, so C = Q^T P exactly; P and Q sit side by side in one row-major
  matrix [P | Q] per block.  The row norms of P (`tt_row_norms`) are the singular values: a Gram
  eigenvalue is off by about eps s_0^2, but a row norm of the projection on its eigenvector only to
  second order, (eps s_0^2 / s_i^2)^2 relative, sound down to about 1e-5 s_0.
- further passes, where the host's ranking could reach a block's values below what a pass resolves
  (`SvdSettings.resolved`, 1e-7 of the pass's largest value), over the block's tail (its rows below
  `SvdSettings.reliable`, 1e-5, of the pass's largest value, cut at the widest gap in the decade
  above): the tail T is decoupled from the pass's head rows H, whose eigenvectors leaked into it (a
  Gram pass of T alone would turn that leak into spurious values of about eps s_0^2 / s_h), by the
  rotation exp([[0, X^T], [-X, 0]]), X_ij = (T_P H_P^T)_ij / (|H_j|^2 - |T_i|^2) (|X| ~ 1e-6),
  applied to third order to [P | Q]; then T is diagonalised by the eigenvectors W of -T_P T_P^T:
  T <- W T.  T's norm is at most about 10 x reliable x the pass's largest, so the pass resolves its
  values to about eps x that.  The head rows keep their values, accurate to (eps s_0^2 / s^2)^2
  relative, which is below eps s_0 absolute for every s above the boundary.
- the factors: Q's rows and P's rows, in descending order of their norms, copied into the blocks of
  the two new tensors (`tt_copy_rows`).

Every launch reads a problem table built here, checked against the buffers it names, and uploaded
with the other tables of its phase in one copy.  The host reads the values back once per pass.
'''

from typing import Callable, List

import numpy as np
import torch

from . import _native


def _offsets(sizes) -> np.ndarray:
    '''Exclusive prefix sums (one more entry: the total).'''
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.int64))]).astype(np.int64)


def _bgemm(count: int, **fields):
    '''A bgemm problem table of `count` problems, fields given as scalars or arrays; numbered by
    64 x 64 tiles.  Returns (table, tiles).'''
    table = np.zeros(count, dtype=_native.BGEMM_PROBLEM)
    table['alpha'] = 1.0
    for name, value in fields.items():
        table[name] = value
    tiles = (-(-table['m'].astype(np.int64) // 64)) * (-(-table['n'].astype(np.int64) // 64))
    table['first'] = _offsets(tiles)[:-1]
    return table, int(tiles.sum())


def _rowwise(dtype, count: int, **fields):
    '''A row-problem or lincomb table, numbered by rows.  Returns (table, rows).'''
    table = np.zeros(count, dtype=dtype)
    for name, value in fields.items():
        table[name] = value
    rows = table['rows'].astype(np.int64)
    table['first'] = _offsets(rows)[:-1]
    return table, int(rows.sum())


class _Stage:
    '''The host arrays and launch tables of one phase in one device buffer: the device range is
    allocated first, so tables can name the addresses of what is staged with them; the host side
    is pinned memory, sent in one copy behind the phase's earlier work.  On a CPU device the buffer
    is host memory and the copy a plain one.'''

    def __init__(self, device, sizes):
        self.offsets = _offsets([(int(n) + 255) // 256 * 256 for n in sizes])
        total = max(int(self.offsets[-1]), 1)
        self.device_buffer = torch.empty(total, dtype=torch.uint8, device=device)
        self.host = torch.empty(total, dtype=torch.uint8, pin_memory=device.type == 'cuda')
        self.view = self.host.numpy()
        self.base = self.device_buffer.data_ptr()

    def address(self, part: int) -> int:
        return self.base + int(self.offsets[part])

    def put(self, part: int, array: np.ndarray) -> None:
        at = int(self.offsets[part])
        self.view[at:at + array.nbytes] = np.ascontiguousarray(array).view(np.uint8).reshape(-1)

    def send(self) -> None:
        self.device_buffer.copy_(self.host, non_blocking=True)


class _Launches:
    '''The launches of one phase: each table checked against `extents` (every buffer its operands may
    touch), staged with the phase's host arrays, and launched from its device address.'''

    def __init__(self, extents: '_native.Extents'):
        self.extents = extents
        self.steps = []

    def add(self, kind, table_items, source_rows=None):
        table, items = table_items
        if kind == 'gemm':
            checked = _native.check_bgemm(table, self.extents)
        elif kind == 'lincomb':
            checked = _native.check_lincomb(table, self.extents)
        else:
            checked = _native.check_rows(table, self.extents, source_rows, vector_dst=kind == 'norms')
        if checked != items:
            raise RuntimeError(f'{kind}: work-item count')
        self.steps.append((kind, table, items))

    def stage(self, stage: '_Stage', first_part: int) -> None:
        '''Puts the tables into `stage` from part `first_part` on, and sends it.'''
        self.addresses = []
        for i, (_, table, _) in enumerate(self.steps):
            stage.put(first_part + i, table)
            self.addresses.append(stage.address(first_part + i))
        stage.send()

    def run(self, first: int = 0, last: int = None):
        for (kind, table, items), address in list(zip(self.steps, self.addresses))[first:last]:
            if not table.size or not items:
                continue
            if kind == 'gemm':
                _native.bgemm(table, address, items)
            elif kind == 'norms':
                _native.row_norms(table, address, items)
            elif kind == 'lincomb':
                _native.lincomb_rows(table, address, items)
            else:
                _native.copy_rows(table, address, items)


def _extents_of(*tensors) -> '_native.Extents':
    extents = _native.Extents()
    for t in tensors:
        extents.add(t)
    return extents


def _doubles(device, count: int) -> torch.Tensor:
    return torch.empty(max(int(count), 1), dtype=torch.float64, device=device)


# rotations that decouple a tail from its head before the tail's Gram pass: one leaves a coupling
# of order |X|^2 relative, |X| being about eps x (largest / boundary value)^2 / (1 - gap^-2)
DECOUPLING = 1


# the tail of a pass starts at the widest gap among the values within these factors of the level
WINDOW = (1.0, 10.0)


def _tail_start(values: np.ndarray, first_below: int, level: float) -> int:
    '''Where a pass's tail starts: at the widest gap between consecutive values (rows in descending
    eigenvalue order) among the rows between WINDOW[0] and WINDOW[1] times `level`, so that no
    near-degenerate pair straddles the cut (the decoupling divides by the gap).'''
    window = np.nonzero((values > WINDOW[0] * level) & (values < WINDOW[1] * level))[0]
    if not window.size:
        return first_below
    lo, hi = max(1, int(window[0])), min(values.size - 1, int(window[-1]) + 1)
    if hi < lo:
        return first_below
    ratios = values[lo - 1:hi] / np.maximum(values[lo:hi + 1], np.finfo(np.float64).tiny)
    return lo + int(np.argmax(ratios))


def choose_tails(norms: List[np.ndarray], ranks: List[int], first: List[int], floor: float, reliable: float,
                 resolved: float):
    '''The blocks whose current pass needs another: (block, tail start).  `norms[b]` holds the row
    norms of block b's projection, all g rows (its singular values are the first ranks[b] of them
    once resolved); the current pass covers rows first[b]...  A pass's row norms are trusted down to
    `resolved` x its largest value (there the second-order error is about (eps / resolved^2)^2 = 1e-4
    relative); below that nothing is known of a row but its norm.  Another pass is needed when the
    smallest value the ranking can keep, `floor`, lies below twice that level and the unresolved rows
    could hold a value above half the floor: their singular values are at most their Frobenius norm
    (the coupling to the resolved rows only lowers them, by the Schur complement), so a set whose norm
    is below half the floor is dropped whatever its exact values.  The tail then holds the rows below
    about reliable x the largest, cut at a wide gap.'''
    count = len(norms)
    if not count:
        return []
    first = np.asarray(first, dtype=np.int64)
    pass_sizes = np.maximum(np.asarray(ranks, dtype=np.int64) - first, 0)
    # the largest value of every block's pass, in one reduction over the passes back to back
    tops = np.zeros(count)
    live = np.nonzero(pass_sizes > 0)[0]
    if live.size:
        values = np.concatenate([norms[b][first[b]:first[b] + pass_sizes[b]] for b in live.tolist()])
        starts = np.concatenate([[0], np.cumsum(pass_sizes[live])[:-1]]).astype(np.int64)
        tops[live] = np.maximum.reduceat(values, starts)
    candidate = (pass_sizes >= 2) & (tops > 0.0) & (floor < 2.0 * resolved * tops)
    tails = []
    for b in np.nonzero(candidate)[0].tolist():
        v = norms[b][first[b]:]
        top = tops[b]
        unresolved = v[v < 2.0 * resolved * top]
        if float(np.sqrt(np.sum(unresolved ** 2))) < 0.5 * floor:
            continue
        pass_values = v[:pass_sizes[b]]
        below = np.nonzero(pass_values < reliable * top)[0]
        if below.size:
            start = _tail_start(pass_values, int(below[0]), reliable * top)
            if start > 0:        # a pass whose first row is already in the tail has no head to leave
                tails.append((b, int(first[b]) + start))
    return tails


class GramSplit:
    '''The singular triplets of the blocks `blocks` (2-D row-major device views, FP64) for a split
    that keeps the left factor isometric (`left_isometric`, absorb 2) or the right one (absorb 1).
    `factor(cut, reliable, resolved)` runs the passes, `cut(values)` being the host ranking's floor,
    and returns each block's singular values (host, descending); `assemble` then writes the kept
    triplets into the new tensors' blocks.'''

    def __init__(self, blocks, left_isometric: bool):
        rows = np.array([b.shape[0] for b in blocks], dtype=np.int64)
        cols = np.array([b.shape[1] for b in blocks], dtype=np.int64)
        pointers = np.array([b.data_ptr() for b in blocks], dtype=np.int64)
        self._setup(rows, cols, pointers, blocks[0].device, pointers + 8 * rows * cols, left_isometric)

    @classmethod
    def of_buffer(cls, data: torch.Tensor, rows, cols, offsets, left_isometric: bool) -> 'GramSplit':
        '''The split of the blocks of one flat FP64 buffer: block b is rows[b] x cols[b], row-major, at
        element offset offsets[b].'''
        self = cls.__new__(cls)
        rows, cols = np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64)
        pointers = data.data_ptr() + 8 * np.asarray(offsets, dtype=np.int64)
        self._setup(rows, cols, pointers, data.device, pointers + 8 * rows * cols, left_isometric)
        return self

    def _setup(self, rows, cols, pointers, device, ends, left_isometric: bool):
        self.count = len(rows)
        self.left = left_isometric
        self.rows, self.cols = rows, cols
        self.g = rows.copy() if left_isometric else cols.copy()
        self.n = cols.copy() if left_isometric else rows.copy()
        self.r = np.minimum(self.g, self.n)
        self.ld = self.n + self.g
        self.device = device
        self.pq_at = _offsets(self.g * self.ld)
        self.pq = _doubles(self.device, self.pq_at[-1])
        self.pq_base = self.pq.data_ptr()
        self.block_ptr = pointers
        self.block_spans = list(zip(pointers.tolist(), ends.tolist()))

    def _row_ptr(self, b, row):
        '''Address of row `row` of block b's [P | Q] (vectorised over arrays).'''
        return self.pq_base + 8 * (self.pq_at[b] + row * self.ld[b])

    # -- pass 1
    def _first_pass(self):
        g, n, ld, count = self.g, self.n, self.ld, self.count
        sq_at, g_at = _offsets(g * g), _offsets(g)
        grams, vectors, norms = _doubles(self.device, sq_at[-1]), _doubles(self.device, sq_at[-1]), \
            _doubles(self.device, g_at[-1])
        extents = _extents_of(self.pq, grams, vectors, norms)
        for start, end in self.block_spans:
            extents.add_range(start, end)
        gram_at = grams.data_ptr() + 8 * sq_at[:-1]
        vec_at = vectors.data_ptr() + 8 * sq_at[:-1]
        p_at = self._row_ptr(np.arange(count), 0)
        launches = _Launches(extents)
        # -C C^T with C = B (B B^T) or C = B^T (B^T B); B is stored rows x cols
        launches.add('gemm', _bgemm(count, a=self.block_ptr, b=self.block_ptr, c=gram_at, m=g, n=g,
                                    k=self.cols if self.left else self.rows, lda=self.cols, ldb=self.cols, ldc=g,
                                    trans_a=0 if self.left else 1, trans_b=1 if self.left else 0, alpha=-1.0))
        # P = Q C into [P | Q][:, :n], Q beside it, the row norms of P
        launches.add('gemm', _bgemm(count, a=vec_at, b=self.block_ptr, c=p_at, m=g, n=n, k=g, lda=g, ldb=self.cols,
                                    ldc=ld, trans_b=0 if self.left else 1))
        launches.add('copy', _rowwise(_native.ROWS_PROBLEM, count, src=vec_at, dst=p_at + 8 * n, rows=g, cols=g,
                                      src_ld=g, dst_ld=ld))
        launches.add('norms', _rowwise(_native.ROWS_PROBLEM, count, src=p_at, dst=norms.data_ptr() + 8 * g_at[:-1],
                                       rows=g, cols=n, src_ld=ld, dst_ld=1))
        stage = _Stage(self.device, [t.nbytes for _, t, _ in launches.steps])
        launches.stage(stage, 0)
        launches.run(0, 1)
        _native.eigh_batched(grams, g.tolist(), sq_at[:-1].tolist(), vectors=vectors)
        launches.run(1)
        host = norms.cpu().numpy().copy()
        self.norms_flat, self.norms_at = host, g_at
        return [host[g_at[b]:g_at[b + 1]] for b in range(count)]     # views: updates reach the flat array

    # -- a further pass over the tails
    def _tail_pass(self, tails, values):
        '''One more pass over the tails: decoupled from their heads (DECOUPLING rotations), then
        diagonalised; the tails' new row norms are read back into `values`.  With one rotation the
        launch tables come from the engine (`TailPass`: laid out, filled and checked in one call);
        the launches, in the order of `_decouple` and `_diagonalise`, are the same.'''
        heads = [(b, self.first[b], h) for b, h in tails]
        if DECOUPLING == 1:
            self._tail_pass_tables(heads, values)
        else:
            tail_rows = None
            for iteration in range(DECOUPLING):
                tail_rows = self._decouple(heads, values, last=iteration + 1 == DECOUPLING)
            self._diagonalise(heads, values, tail_rows)
        for b, f, h in heads:
            self.first[b] = h

    def _tail_pass_tables(self, heads, values):
        '''`_decouple` (one rotation) and `_diagonalise` from the engine's tables: one staging
        buffer, ten table launches around the eigensolver, one readback of the tails' norms.'''
        b = [x[0] for x in heads]
        plan = _native.engine.TailPass(self.g.tolist(), self.n.tolist(), self.ld.tolist(), self.pq_at[:-1].tolist(),
                                       b, [x[1] for x in heads], [x[2] for x in heads])
        work = _doubles(self.device, plan.work_doubles)
        grams = _doubles(self.device, 2 * plan.gram_doubles + plan.norm_doubles)
        vectors, norms = grams[plan.gram_doubles:2 * plan.gram_doubles], grams[2 * plan.gram_doubles:]
        stage = _Stage(self.device, [plan.stage_bytes])
        flat = np.concatenate(values)
        at = _offsets([v.size for v in values])
        plan.fill(self.pq_base, self.pq.numel() * 8, work.data_ptr(), grams.data_ptr(), vectors.data_ptr(),
                  norms.data_ptr(), stage.base, stage.host.data_ptr(), flat, at)
        stage.send()
        tables = plan.tables()
        dtypes = (_native.BGEMM_PROBLEM, _native.LINCOMB_PROBLEM, _native.ROWS_PROBLEM)

        def launch(kind, count, items, offset):
            if not count or not items:
                return
            # the launch reads the staged table; the host copy stands in for its length
            table = np.empty(count, dtype=dtypes[kind])
            table.view(np.uint8)[:] = stage.view[offset:offset + table.nbytes]
            address = stage.base + offset
            if kind == 0:
                _native.bgemm(table, address, items)
            elif kind == 1:
                _native.lincomb_rows(table, address, items)
            else:
                _native.row_norms(table, address, items)

        for entry in tables[:8]:
            launch(*entry)
        orders = plan.orders
        _native.eigh_batched(grams[:max(plan.gram_doubles, 1)], orders, plan.gram_offsets, vectors=vectors)
        for entry in tables[8:]:
            launch(*entry)
        host = norms.cpu().numpy()
        self._keep = (work, grams, stage)
        start = 0
        for (bb, ff, hh), q in zip(heads, orders):
            values[bb][hh:] = host[start:start + q]
            start += q

    def _decouple(self, heads, values, last: bool):
        '''The rotation exp([[0, X^T], [-X, 0]]) of every (head, tail) pair of rows of [P | Q], with
        X_ij = (T_P H_P^T)_ij / (|H_j|^2 - |T_i|^2) (the rotation's generator to first order: with
        H_P H_P^T and T_P T_P^T diagonal, it solves X H_P H_P^T - T_P T_P^T X = T_P H_P^T, so the
        rotated tail is orthogonal to the rotated head), through the third order: with K = X^T X
        and U = X^T T, head <- (1 - K/2) H + (1 - K/6) U and tail <- T - X ((1 - K/6) H + U/2),
        orthogonal to |X|^4 / 24 (|X| is below about 1e-5 at a tail boundary chosen at a gap).  The
        heads go back in place; the tails too, unless `last`, when they stay in the returned buffer
        for the diagonalisation.'''
        b = np.array([x[0] for x in heads], dtype=np.int64)
        f = np.array([x[1] for x in heads], dtype=np.int64)
        h = np.array([x[2] for x in heads], dtype=np.int64)
        count = len(heads)
        e, q, w, n = h - f, self.g[b] - h, self.ld[b], self.n[b]
        # one work buffer: X (q x e), U, H', M (e x w), T' (q x w), K, A, B (e x e), per head
        sizes = np.stack([q * e, e * w, e * w, e * w, q * w, e * e, e * e, e * e])
        at = _offsets(sizes.T.reshape(-1)).reshape(-1)[:-1].reshape(count, 8)
        work = _doubles(self.device, sizes.sum())
        ptr = work.data_ptr() + 8 * at
        X, U, Hn, M, Tn, K, A, B = (ptr[:, j] for j in range(8))
        H, T = self._row_ptr(b, f), self._row_ptr(b, h)
        scale = np.concatenate([values[bb][ff:hh] ** 2 for bb, ff, hh in heads])
        shift = np.concatenate([values[bb][hh:] ** 2 for bb, ff, hh in heads])
        head_at, tail_at = _offsets(e)[:-1], _offsets(q)[:-1]
        copies = count if last else 2 * count
        counts = [count, 2 * count, 4 * count, 2 * count, count, count, copies]
        kinds = [(_native.BGEMM_PROBLEM, c) for c in counts[:2]] + [(_native.LINCOMB_PROBLEM, counts[2])] + \
            [(_native.BGEMM_PROBLEM, c) for c in counts[3:6]] + [(_native.LINCOMB_PROBLEM, counts[6])]
        stage = _Stage(self.device, [scale.nbytes, shift.nbytes] + [dt.itemsize * c for dt, c in kinds])
        stage.put(0, scale)
        stage.put(1, shift)
        scale_ptr, shift_ptr = stage.address(0) + 8 * head_at, stage.address(1) + 8 * tail_at
        extents = _extents_of(self.pq, work, stage.device_buffer)
        cat = np.concatenate
        launches = _Launches(extents)
        launches.add('gemm', _bgemm(count, a=T, b=H, c=X, m=q, n=e, k=n, lda=w, ldb=w, ldc=e, trans_b=1,
                                    col_scale=scale_ptr, row_shift=shift_ptr))
        launches.add('gemm', _bgemm(2 * count, a=cat([X, X]), b=cat([X, T]), c=cat([K, U]), m=cat([e, e]),
                                    n=cat([e, w]), k=cat([q, q]), lda=cat([e, e]), ldb=cat([e, w]), ldc=cat([e, w]),
                                    trans_a=1))
        launches.add('lincomb', _rowwise(_native.LINCOMB_PROBLEM, 4 * count, x=cat([K, K, U, T]), dst=cat([A, B, M, Tn]),
                                         rows=cat([e, e, e, q]), cols=cat([e, e, w, w]), x_ld=cat([e, e, w, w]),
                                         dst_ld=cat([e, e, w, w]),
                                         alpha=np.repeat([-0.5, -1.0 / 6.0, 0.5, 1.0], count),
                                         gamma=np.repeat([1.0, 1.0, 0.0, 0.0], count)))
        launches.add('gemm', _bgemm(2 * count, a=cat([A, B]), b=cat([H, H]), c=cat([Hn, M]), m=cat([e, e]),
                                    n=cat([w, w]), k=cat([e, e]), lda=cat([e, e]), ldb=cat([w, w]), ldc=cat([w, w]),
                                    beta=np.repeat([0.0, 1.0], count)))
        launches.add('gemm', _bgemm(count, a=B, b=U, c=Hn, m=e, n=w, k=e, lda=e, ldb=w, ldc=w, beta=1.0))
        launches.add('gemm', _bgemm(count, a=X, b=M, c=Tn, m=q, n=w, k=e, lda=e, ldb=w, ldc=w, alpha=-1.0, beta=1.0))
        back_src, back_dst = (Hn, H) if last else (cat([Hn, Tn]), cat([H, T]))
        back_rows = e if last else cat([e, q])
        launches.add('lincomb', _rowwise(_native.LINCOMB_PROBLEM, copies, x=back_src, dst=back_dst, rows=back_rows,
                                         cols=np.tile(w, copies // count), x_ld=np.tile(w, copies // count),
                                         dst_ld=np.tile(w, copies // count), alpha=1.0))
        launches.stage(stage, 2)
        launches.run()
        self._keep = (work, stage)
        return (Tn, q, work) if last else None

    def _diagonalise(self, heads, values, tail_rows):
        '''Each decoupled tail (q x w per block, at `Tn`) diagonalised: W, the eigenvectors of
        -T_P T_P^T as rows, gives the block's tail rows W T; their norms come back to the host.'''
        Tn, q, work = tail_rows
        b = np.array([x[0] for x in heads], dtype=np.int64)
        h = np.array([x[2] for x in heads], dtype=np.int64)
        count, w, n = len(heads), self.ld[b], self.n[b]
        sq_at, nr_at = _offsets(q * q), _offsets(q)
        grams, vectors, norms = _doubles(self.device, sq_at[-1]), _doubles(self.device, sq_at[-1]), \
            _doubles(self.device, nr_at[-1])
        T = self._row_ptr(b, h)
        gram_ptr, vec_ptr = grams.data_ptr() + 8 * sq_at[:-1], vectors.data_ptr() + 8 * sq_at[:-1]
        launches = _Launches(_extents_of(self.pq, work, grams, vectors, norms))
        launches.add('gemm', _bgemm(count, a=Tn, b=Tn, c=gram_ptr, m=q, n=q, k=n, lda=w, ldb=w, ldc=q, trans_b=1,
                                    alpha=-1.0))
        launches.add('gemm', _bgemm(count, a=vec_ptr, b=Tn, c=T, m=q, n=w, k=q, lda=q, ldb=w, ldc=w))
        launches.add('norms', _rowwise(_native.ROWS_PROBLEM, count, src=T, dst=norms.data_ptr() + 8 * nr_at[:-1],
                                       rows=q, cols=n, src_ld=w, dst_ld=1))
        stage = _Stage(self.device, [t.nbytes for _, t, _ in launches.steps])
        launches.stage(stage, 0)
        launches.run(0, 1)
        _native.eigh_batched(grams, q.tolist(), sq_at[:-1].tolist(), vectors=vectors)
        launches.run(1)
        host = norms.cpu().numpy()
        self._keep = None
        for i, (bb, ff, hh) in enumerate(heads):
            values[bb][hh:] = host[nr_at[i]:nr_at[i + 1]]

    def factor(self, cut: Callable, reliable: float, resolved: float) -> List[np.ndarray]:
        norms = self._first_pass()          # every row of every block; the first r are the values
        self.first = [0] * self.count
        ranks = self.r.tolist()
        rank_array = np.asarray(ranks, dtype=np.int64)
        while True:
            # the rule of choose_tails, run by the engine over the flat row norms
            floor = cut([v[:r] for v, r in zip(norms, ranks)])
            tails = [tuple(t) for t in _native.engine.choose_tails(
                self.norms_flat, self.norms_at, rank_array, np.asarray(self.first, dtype=np.int64), floor, reliable,
                resolved)]
            if not tails:
                break
            self._tail_pass(tails, norms)
        values = [v[:r] for v, r in zip(norms, ranks)]
        self.order = [np.argsort(-v, kind='stable').astype(np.int32) for v in values]
        return [v[o] for v, o in zip(values, self.order)]

    # -- the factors
    def assemble(self, placements, spans):
        '''`placements`: (block index, kept k, address of the first factor's block at its column offset,
        that block's row length, address of the second factor's block at its row offset, that block's
        row length) per kept block; `spans` the byte ranges of those blocks, which bound the copies.
        Writes Q's and P's leading k rows (in descending value order) into the two new tensors' blocks.'''
        if not placements:
            return
        count = len(placements)
        columns = np.array(placements, dtype=np.int64).T
        b, k, first_ptr, first_cols, second_ptr, second_cols = columns
        index = np.concatenate([self.order[bb][:kk] for bb, kk in zip(b, k)]).astype(np.int32)
        n, g, ld = self.n[b], self.g[b], self.ld[b]
        p_rows, q_rows = self._row_ptr(b, 0), self._row_ptr(b, 0) + 8 * n
        # absorb 2: first = Q^T (columns), second = P (rows); absorb 1: first = P^T, second = Q
        first_src, first_width = (q_rows, g) if self.left else (p_rows, n)
        second_src, second_width = (p_rows, n) if self.left else (q_rows, g)
        table_bytes = 2 * count * _native.ROWS_PROBLEM.itemsize
        stage = _Stage(self.device, [index.nbytes, table_bytes])
        stage.put(0, index)
        rows_ptr = stage.address(0) + 4 * _offsets(k)[:-1]
        extents = _extents_of(self.pq, stage.device_buffer)
        for start, end in spans:
            extents.add_range(start, end)
        cat = np.concatenate
        launches = _Launches(extents)
        launches.add('copy', _rowwise(_native.ROWS_PROBLEM, 2 * count, src=cat([first_src, second_src]),
                                      dst=cat([first_ptr, second_ptr]), index=cat([rows_ptr, rows_ptr]), rows=cat([k, k]),
                                      cols=cat([first_width, second_width]), src_ld=cat([ld, ld]),
                                      dst_ld=cat([first_cols, second_cols]),
                                      transpose=np.repeat([1, 0], count)),
                     source_rows=cat([self.r[b], self.r[b]]))
        launches.stage(stage, 1)
        launches.run()
        self._keep = stage
