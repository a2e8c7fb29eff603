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


def _table(rows, dtype, count_field):
    '''A problem table from tuples (the `first` field numbered here), and its item count.'''
    table = np.array(rows, dtype=dtype) if rows else np.zeros(0, dtype=dtype)
    if count_field == 'tiles':
        items = (-(-table['m'] // 64)) * (-(-table['n'] // 64))
    else:
        items = table['rows'].astype(np.int64)
    table['first'] = np.concatenate([[0], np.cumsum(items)[:-1]]) if table.size else table['first']
    return table, int(items.sum())


def _gemm(a, b, c, m, n, k, lda, ldb, ldc, trans_a=0, trans_b=0, alpha=1.0, beta=0.0, scale=0, shift=0):
    return (a, b, c, scale, shift, 0, m, n, k, lda, ldb, ldc, trans_a, trans_b, alpha, beta)


def _rows(src, dst, rows, cols, src_ld, dst_ld, transpose=0, index=0):
    return (src, dst, index, 0, rows, cols, src_ld, dst_ld, transpose, 0)


def _lincomb(x, y, dst, rows, cols, x_ld, y_ld, dst_ld, alpha, beta, gamma):
    return (x, y, dst, 0, rows, cols, x_ld, y_ld, dst_ld, 0, alpha, beta, gamma)


# rotations that decouple a tail from its head before the tail's Gram pass: one leaves a coupling
# of order |X|^2 relative, |X| being about eps x (largest / boundary value)^2 / (1 - gap^-2)
DECOUPLING = 1


class _Launches:
    '''The launches of one phase: tables collected, checked against `extents`, uploaded together.'''

    def __init__(self, extents: '_native.Extents'):
        self.extents = extents
        self.steps = []

    def gemm(self, problems):
        table, tiles = _table(problems, _native.BGEMM_PROBLEM, 'tiles')
        if _native.check_bgemm(table, self.extents) != tiles:
            raise RuntimeError('bgemm: tile count')
        self.steps.append(('gemm', table, tiles))

    def lincomb(self, problems):
        table, rows = _table(problems, _native.LINCOMB_PROBLEM, 'rows')
        if _native.check_lincomb(table, self.extents) != rows:
            raise RuntimeError('lincomb: row count')
        self.steps.append(('lincomb', table, rows))

    def rows(self, kind, problems, source_rows=None):
        table, rows = _table(problems, _native.ROWS_PROBLEM, 'rows')
        if _native.check_rows(table, self.extents, source_rows, vector_dst=kind == 'norms') != rows:
            raise RuntimeError('row problems: row count')
        self.steps.append((kind, table, rows))

    def upload(self):
        '''Uploads every table; returns the device buffer (keep it until the launches are queued).'''
        buffer, addresses = _native.upload(*[table for _, table, _ in self.steps])
        self.addresses = addresses
        return buffer

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


def _uploaded(array: np.ndarray, device) -> torch.Tensor:
    '''A host array on the device, copied from pinned memory without waiting.'''
    host = torch.from_numpy(np.ascontiguousarray(array))
    if device.type == 'cuda':
        return host.pin_memory().to(device, non_blocking=True)
    return host.clone()


def _ptr(t: torch.Tensor, offset: int = 0) -> int:
    return t.data_ptr() + 8 * int(offset)


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
    tails = []
    for b, v in enumerate(norms):
        pass_values = v[first[b]:ranks[b]]
        top = float(pass_values.max()) if pass_values.size else 0.0
        if pass_values.size < 2 or top == 0.0 or floor >= 2.0 * resolved * top:
            continue
        unresolved = v[first[b]:][v[first[b]:] < 2.0 * resolved * top]
        if float(np.sqrt(np.sum(unresolved ** 2))) < 0.5 * floor:
            continue
        below = np.nonzero(pass_values < reliable * top)[0]
        if below.size:
            start = _tail_start(pass_values, int(below[0]), reliable * top)
            if start > 0:        # a pass whose first row is already in the tail has no head to leave
                tails.append((b, first[b] + start))
    return tails


class GramSplit:
    '''The singular triplets of the blocks `blocks` (2-D row-major device views, FP64) for a split
    that keeps the left factor isometric (`left_isometric`, absorb 2) or the right one (absorb 1).
    `factor(cut, reliable)` runs the passes, `cut(values)` being the host ranking's floor, and
    returns each block's singular values (host, descending); `assemble` then writes the kept
    triplets into the new tensors' blocks.'''

    def __init__(self, blocks, left_isometric: bool):
        self.blocks = blocks
        self.left = left_isometric
        self.g = [b.shape[0] if left_isometric else b.shape[1] for b in blocks]
        self.n = [b.shape[1] if left_isometric else b.shape[0] for b in blocks]
        self.r = [min(g, n) for g, n in zip(self.g, self.n)]
        self.ld = [n + g for g, n in zip(self.g, self.n)]
        device = blocks[0].device
        self.pq_at = np.concatenate([[0], np.cumsum([g * ld for g, ld in zip(self.g, self.ld)])]).astype(np.int64)
        self.pq = torch.empty(max(int(self.pq_at[-1]), 1), dtype=torch.float64, device=device)
        self.device = device

    # -- pass 1
    def _first_pass(self):
        g, n, ld, r = self.g, self.n, self.ld, self.r
        sq_at = np.concatenate([[0], np.cumsum([x * x for x in g])]).astype(np.int64)
        grams = torch.empty(max(int(sq_at[-1]), 1), dtype=torch.float64, device=self.device)
        vectors = torch.empty_like(grams)
        g_at = np.concatenate([[0], np.cumsum(g)]).astype(np.int64)
        norms = torch.empty(max(int(g_at[-1]), 1), dtype=torch.float64, device=self.device)
        extents = _native.Extents().add(self.pq).add(grams).add(vectors).add(norms)
        gram_problems, p_problems, q_copies, norm_rows = [], [], [], []
        for b, block in enumerate(self.blocks):
            extents.add(block)
            rows_b, cols_b = block.shape
            base = block.data_ptr()
            gb, nb = g[b], n[b]
            # -C C^T with C = B (B B^T) or C = B^T (B^T B); B is stored rows_b x cols_b
            if self.left:
                gram_problems.append(_gemm(base, base, _ptr(grams, sq_at[b]), gb, gb, cols_b, cols_b, cols_b, gb,
                                           trans_a=0, trans_b=1, alpha=-1.0))
            else:
                gram_problems.append(_gemm(base, base, _ptr(grams, sq_at[b]), gb, gb, rows_b, cols_b, cols_b, gb,
                                           trans_a=1, trans_b=0, alpha=-1.0))
            # P = Q C into [P | Q][:, :n]
            p_problems.append(_gemm(_ptr(vectors, sq_at[b]), base, _ptr(self.pq, self.pq_at[b]), gb, nb, gb, gb,
                                    cols_b, ld[b], trans_a=0, trans_b=0 if self.left else 1))
            q_copies.append(_rows(_ptr(vectors, sq_at[b]), _ptr(self.pq, self.pq_at[b] + nb), gb, gb, gb, ld[b]))
            norm_rows.append(_rows(_ptr(self.pq, self.pq_at[b]), _ptr(norms, g_at[b]), gb, nb, ld[b], 1))
        launches = _Launches(extents)
        launches.gemm(gram_problems)
        launches.gemm(p_problems)
        launches.rows('copy', q_copies)
        launches.rows('norms', norm_rows)
        keep = launches.upload()
        launches.run(0, 1)
        _native.eigh_batched(grams, g, sq_at[:-1].tolist(), vectors=vectors)
        launches.run(1)
        host = norms.cpu().numpy()
        del keep
        return [host[g_at[b]:g_at[b + 1]].copy() for b in range(len(g))]

    # -- a further pass over the tails
    def _tail_pass(self, tails, values):
        '''One more pass over the tails: decoupled from their heads (DECOUPLING rotations), then
        diagonalised; the tails' new row norms are read back into `values`.'''
        heads = [(b, self.first[b], h) for b, h in tails]
        tail_rows = None
        for iteration in range(DECOUPLING):
            tail_rows = self._decouple(heads, values, last=iteration + 1 == DECOUPLING)
        self._diagonalise(heads, values, tail_rows)
        for b, f, h in heads:
            self.first[b] = h

    def _decouple(self, heads, values, last: bool):
        '''The rotation exp([[0, X^T], [-X, 0]]) of every (head, tail) pair of rows of [P | Q], with
        X_ij = (T_P H_P^T)_ij / (|H_j|^2 - |T_i|^2) (the rotation's generator to first order: with
        H_P H_P^T and T_P T_P^T diagonal, it solves X H_P H_P^T - T_P T_P^T X = T_P H_P^T, so the
        rotated tail is orthogonal to the rotated head), through the third order: with K = X^T X
        and U = X^T T, head <- (1 - K/2) H + (1 - K/6) U and tail <- T - X ((1 - K/6) H + U/2),
        orthogonal to |X|^4 / 24 (|X| is below about 1e-5 at a tail boundary chosen at a gap).  The
        heads go back in place; the tails too, unless `last`, when they stay in the returned buffer
        for the diagonalisation.'''
        n, ld, g = self.n, self.ld, self.g
        hh = [h - f for _, f, h in heads]
        t = [g[b] - h for b, _, h in heads]
        w = [ld[b] for b, _, _ in heads]

        def offsets(sizes):
            return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

        x_at, sq_at = offsets([a * c for a, c in zip(t, hh)]), offsets([a * a for a in hh])
        hw_at, tw_at = offsets([a * c for a, c in zip(hh, w)]), offsets([a * c for a, c in zip(t, w)])
        head_at, tail_at = offsets(hh), offsets(t)

        def buffer(size):
            return torch.empty(max(int(size), 1), dtype=torch.float64, device=self.device)

        x, u, hn, m, tn = buffer(x_at[-1]), buffer(hw_at[-1]), buffer(hw_at[-1]), buffer(hw_at[-1]), buffer(tw_at[-1])
        k_mat, a_mat, b_mat = (buffer(sq_at[-1]) for _ in range(3))
        scale = _uploaded(np.concatenate([values[b][f:h] ** 2 for b, f, h in heads]), self.device)
        shift = _uploaded(np.concatenate([values[b][h:] ** 2 for b, f, h in heads]), self.device)
        extents = _native.Extents().add(self.pq).add(scale).add(shift)
        for buf in (x, u, hn, m, tn, k_mat, a_mat, b_mat):
            extents.add(buf)
        steps = {k: [] for k in ('x', 'k', 'init', 'first', 'second', 'tail', 'back')}
        for i, (b, f, h) in enumerate(heads):
            base = self.pq_at[b]
            H, T = _ptr(self.pq, base + f * ld[b]), _ptr(self.pq, base + h * ld[b])
            X, U, Hn, M, Tn = (_ptr(x, x_at[i]), _ptr(u, hw_at[i]), _ptr(hn, hw_at[i]), _ptr(m, hw_at[i]),
                               _ptr(tn, tw_at[i]))
            K, A, B = (_ptr(buf, sq_at[i]) for buf in (k_mat, a_mat, b_mat))
            e, q = hh[i], t[i]
            steps['x'].append(_gemm(T, H, X, q, e, n[b], ld[b], ld[b], e, trans_b=1, scale=_ptr(scale, head_at[i]),
                                    shift=_ptr(shift, tail_at[i])))
            steps['k'].append(_gemm(X, X, K, e, e, q, e, e, e, trans_a=1))
            steps['k'].append(_gemm(X, T, U, e, w[i], q, e, ld[b], w[i], trans_a=1))
            steps['init'].append(_lincomb(K, 0, A, e, e, e, 0, e, -0.5, 0.0, 1.0))
            steps['init'].append(_lincomb(K, 0, B, e, e, e, 0, e, -1.0 / 6.0, 0.0, 1.0))
            steps['init'].append(_lincomb(U, 0, M, e, w[i], w[i], 0, w[i], 0.5, 0.0, 0.0))
            steps['init'].append(_lincomb(T, 0, Tn, q, w[i], ld[b], 0, w[i], 1.0, 0.0, 0.0))
            steps['first'].append(_gemm(A, H, Hn, e, w[i], e, e, ld[b], w[i]))
            steps['first'].append(_gemm(B, H, M, e, w[i], e, e, ld[b], w[i], beta=1.0))
            steps['second'].append(_gemm(B, U, Hn, e, w[i], e, e, w[i], w[i], beta=1.0))
            steps['tail'].append(_gemm(X, M, Tn, q, w[i], e, e, w[i], w[i], alpha=-1.0, beta=1.0))
            steps['back'].append(_lincomb(Hn, 0, H, e, w[i], w[i], 0, ld[b], 1.0, 0.0, 0.0))
            if not last:
                steps['back'].append(_lincomb(Tn, 0, T, q, w[i], w[i], 0, ld[b], 1.0, 0.0, 0.0))
        launches = _Launches(extents)
        for kind in ('x', 'k'):
            launches.gemm(steps[kind])
        launches.lincomb(steps['init'])
        for kind in ('first', 'second', 'tail'):
            launches.gemm(steps[kind])
        launches.lincomb(steps['back'])
        keep = launches.upload()
        launches.run()
        del keep
        self._keep = (scale, shift)
        return (tn, tw_at) if last else None

    def _diagonalise(self, heads, values, tail_rows):
        '''Each decoupled tail (in `tail_rows`, t x w per block) diagonalised: W, the eigenvectors of
        -T_P T_P^T as rows, gives the block's tail rows W T; their norms come back to the host.'''
        n, ld, g = self.n, self.ld, self.g
        tn, tw_at = tail_rows
        t = [g[b] - h for b, _, h in heads]
        sq_at = np.concatenate([[0], np.cumsum([a * a for a in t])]).astype(np.int64)
        nr_at = np.concatenate([[0], np.cumsum(t)]).astype(np.int64)
        grams = torch.empty(max(int(sq_at[-1]), 1), dtype=torch.float64, device=self.device)
        vectors, norms = torch.empty_like(grams), torch.empty(max(int(nr_at[-1]), 1), dtype=torch.float64,
                                                              device=self.device)
        extents = _native.Extents().add(self.pq).add(tn).add(grams).add(vectors).add(norms)
        gram_steps, rotate_steps, norm_steps = [], [], []
        for i, (b, f, h) in enumerate(heads):
            T, Tn, q = _ptr(self.pq, self.pq_at[b] + h * ld[b]), _ptr(tn, tw_at[i]), t[i]
            gram_steps.append(_gemm(Tn, Tn, _ptr(grams, sq_at[i]), q, q, n[b], ld[b], ld[b], q, trans_b=1,
                                    alpha=-1.0))
            rotate_steps.append(_gemm(_ptr(vectors, sq_at[i]), Tn, T, q, ld[b], q, q, ld[b], ld[b]))
            norm_steps.append(_rows(T, _ptr(norms, nr_at[i]), q, n[b], ld[b], 1))
        launches = _Launches(extents)
        launches.gemm(gram_steps)
        launches.gemm(rotate_steps)
        launches.rows('norms', norm_steps)
        keep = launches.upload()
        launches.run(0, 1)
        _native.eigh_batched(grams, t, sq_at[:-1].tolist(), vectors=vectors)
        launches.run(1)
        host = norms.cpu().numpy()
        del keep
        self._keep = None
        for i, (b, f, h) in enumerate(heads):
            values[b][h:] = host[nr_at[i]:nr_at[i + 1]]

    def factor(self, cut: Callable, reliable: float, resolved: float) -> List[np.ndarray]:
        norms = self._first_pass()          # every row of every block; the first r are the values
        self.first = [0] * len(self.blocks)
        while True:
            tails = choose_tails(norms, self.r, self.first, cut([v[:r] for v, r in zip(norms, self.r)]), reliable,
                                 resolved)
            if not tails:
                break
            self._tail_pass(tails, norms)
        values = [v[:r] for v, r in zip(norms, self.r)]
        self.order = [np.argsort(-v, kind='stable').astype(np.int32) for v in values]
        return [v[o] for v, o in zip(values, self.order)]

    # -- the factors
    def assemble(self, placements):
        '''`placements`: (block index, kept k, first-factor block view, its column offset, second-factor
        block view, its row offset) per kept block.  Writes Q's and P's leading k rows (in descending
        value order) into the two new tensors' blocks.'''
        if not placements:
            return
        index_host = np.concatenate([self.order[b][:k] for b, k, *_ in placements])
        index = _uploaded(index_host, self.device)
        at = np.concatenate([[0], np.cumsum([k for _, k, *_ in placements])]).astype(np.int64)
        extents = _native.Extents().add(self.pq).add(index)
        problems, sources = [], []
        for i, (b, k, first_block, column, second_block, row) in enumerate(placements):
            extents.add(first_block).add(second_block)
            base, n, g, ld = self.pq_at[b], self.n[b], self.g[b], self.ld[b]
            rows_pointer = index.data_ptr() + 4 * int(at[i])
            p_rows, q_rows = _ptr(self.pq, base), _ptr(self.pq, base + n)
            first_cols = first_block.shape[1]
            second_cols = second_block.shape[1]
            # absorb 2: first = Q^T (columns), second = P (rows); absorb 1: first = P^T, second = Q
            first_src, first_width = (q_rows, g) if self.left else (p_rows, n)
            second_src, second_width = (p_rows, n) if self.left else (q_rows, g)
            problems.append(_rows(first_src, _ptr(first_block, column), k, first_width, ld, first_cols, 1,
                                  rows_pointer))
            problems.append(_rows(second_src, _ptr(second_block, row * second_cols), k, second_width, ld,
                                  second_cols, 0, rows_pointer))
            sources += [self.r[b], self.r[b]]
        launches = _Launches(extents)
        launches.rows('copy', problems, sources)
        keep = launches.upload()
        launches.run()
        del keep
