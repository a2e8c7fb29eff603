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
  `SvdSettings.reliable`, 1e-5, of the pass's largest value, cut at a wide gap): the tail T is decoupled
  from the pass's head rows H, whose eigenvectors leaked into it (a Gram pass of T alone would turn
  that leak into spurious values of about eps s_0^2 / s_h), by the rotation
  exp([[0, X^T], [-X, 0]]), X = T_P H_P^T diag(1 / |H_P|^2) (|X| ~ 1e-6), applied to second order
  to [P | Q]; then T is diagonalised by the eigenvectors W of -T_P T_P^T: T <- W T.  T's norm is
  at most about reliable x the pass's largest, so the pass resolves its values to about eps x that.
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


def _gemm(a, b, c, m, n, k, lda, ldb, ldc, trans_a=0, trans_b=0, alpha=1.0, beta=0.0, scale=0):
    return (a, b, c, scale, 0, m, n, k, lda, ldb, ldc, trans_a, trans_b, alpha, beta)


def _rows(src, dst, rows, cols, src_ld, dst_ld, transpose=0, index=0):
    return (src, dst, index, 0, rows, cols, src_ld, dst_ld, transpose, 0)


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


def _tail_start(values: np.ndarray, first_below: int, level: float) -> int:
    '''Where a pass's tail starts: at the widest gap between consecutive values (rows in descending
    eigenvalue order) among the rows within a factor 10 of `level`, so that no near-degenerate pair
    straddles the cut (the decoupling divides by the gap).'''
    window = np.nonzero((values > 0.1 * level) & (values < 10.0 * level))[0]
    if not window.size:
        return first_below
    lo, hi = max(1, int(window[0])), min(values.size - 1, int(window[-1]) + 1)
    if hi < lo:
        return first_below
    ratios = values[lo - 1:hi] / np.maximum(values[lo:hi + 1], np.finfo(np.float64).tiny)
    return lo + int(np.argmax(ratios))


def choose_tails(values: List[np.ndarray], first: List[int], floor: float, reliable: float, resolved: float):
    '''The blocks whose current pass needs another: (block, tail start).  A pass's row norms are
    trusted down to `resolved` x its largest value (there the second-order error is about
    (eps / resolved^2)^2 = 1e-4 relative); another pass is needed when the smallest value the ranking
    can keep, `floor`, lies below twice that.  The tail then holds the rows below about reliable x
    the largest, cut at a wide gap.'''
    tails = []
    for b, v in enumerate(values):
        pass_values = v[first[b]:]
        top = float(pass_values.max()) if pass_values.size else 0.0
        if pass_values.size < 2 or top == 0.0 or floor >= 2.0 * resolved * top:
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
        r_at = np.concatenate([[0], np.cumsum(r)]).astype(np.int64)
        norms = torch.empty(max(int(r_at[-1]), 1), dtype=torch.float64, device=self.device)
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
            norm_rows.append(_rows(_ptr(self.pq, self.pq_at[b]), _ptr(norms, r_at[b]), r[b], nb, ld[b], 1))
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
        return [host[r_at[b]:r_at[b + 1]].copy() for b in range(len(g))]

    # -- a further pass over the tails
    def _tail_pass(self, tails, values):
        n, ld, r, g = self.n, self.ld, self.r, self.g
        heads = [(b, self.first[b], h) for b, h in tails]
        hh = [h - f for _, f, h in heads]
        t = [g[b] - h for b, _, h in heads]
        def offsets(sizes):
            return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        x_at, xtx_at, sq_at = offsets([a * c for a, c in zip(t, hh)]), offsets([a * a for a in hh]), offsets([a * a for a in t])
        hn_at = offsets([a * ld[b] for a, (b, _, _) in zip(hh, heads)])
        tn_at = offsets([a * ld[b] for a, (b, _, _) in zip(t, heads)])
        inv_at = offsets(hh)
        nr = [max(r[b] - h, 0) for b, _, h in heads]
        nr_at = offsets(nr)
        def buffer(size):
            return torch.empty(max(int(size), 1), dtype=torch.float64, device=self.device)
        x, xtx, xxt, hn, tn = buffer(x_at[-1]), buffer(xtx_at[-1]), buffer(sq_at[-1]), buffer(hn_at[-1]), buffer(tn_at[-1])
        grams, vectors, norms = buffer(sq_at[-1]), buffer(sq_at[-1]), buffer(nr_at[-1])
        inverse = np.concatenate([1.0 / values[b][f:h] ** 2 for b, f, h in heads])
        scale = _uploaded(inverse, self.device)
        extents = _native.Extents().add(self.pq).add(scale)
        for buf in (x, xtx, xxt, hn, tn, grams, vectors, norms):
            extents.add(buf)
        steps = {k: [] for k in ('x', 'xx', 'copy', 'acc1', 'acc2', 'gram', 'rotate', 'back', 'norms')}
        for i, (b, f, h) in enumerate(heads):
            base = self.pq_at[b]
            H, T = _ptr(self.pq, base + f * ld[b]), _ptr(self.pq, base + h * ld[b])
            X, Hn, Tn = _ptr(x, x_at[i]), _ptr(hn, hn_at[i]), _ptr(tn, tn_at[i])
            # X = T_P H_P^T diag(1 / |H_P|^2)
            steps['x'].append(_gemm(T, H, X, t[i], hh[i], n[b], ld[b], ld[b], hh[i], trans_b=1,
                                    scale=_ptr(scale, inv_at[i])))
            steps['xx'].append(_gemm(X, X, _ptr(xtx, xtx_at[i]), hh[i], hh[i], t[i], hh[i], hh[i], hh[i], trans_a=1))
            steps['xx'].append(_gemm(X, X, _ptr(xxt, sq_at[i]), t[i], t[i], hh[i], hh[i], hh[i], t[i], trans_b=1))
            steps['copy'].append(_rows(H, Hn, hh[i], ld[b], ld[b], ld[b]))
            steps['copy'].append(_rows(T, Tn, t[i], ld[b], ld[b], ld[b]))
            # Hn = H + X^T T - X^T X H / 2 and Tn = T - X H - X X^T T / 2, both rows of [P | Q]
            steps['acc1'].append(_gemm(X, T, Hn, hh[i], ld[b], t[i], hh[i], ld[b], ld[b], trans_a=1, beta=1.0))
            steps['acc1'].append(_gemm(X, H, Tn, t[i], ld[b], hh[i], hh[i], ld[b], ld[b], alpha=-1.0, beta=1.0))
            steps['acc2'].append(_gemm(_ptr(xtx, xtx_at[i]), H, Hn, hh[i], ld[b], hh[i], hh[i], ld[b], ld[b],
                                       alpha=-0.5, beta=1.0))
            steps['acc2'].append(_gemm(_ptr(xxt, sq_at[i]), T, Tn, t[i], ld[b], t[i], t[i], ld[b], ld[b],
                                       alpha=-0.5, beta=1.0))
            # -Tn_P Tn_P^T, diagonalised: T = W Tn; the head goes back in place
            steps['gram'].append(_gemm(Tn, Tn, _ptr(grams, sq_at[i]), t[i], t[i], n[b], ld[b], ld[b], t[i],
                                       trans_b=1, alpha=-1.0))
            steps['rotate'].append(_gemm(_ptr(vectors, sq_at[i]), Tn, T, t[i], ld[b], t[i], t[i], ld[b], ld[b]))
            steps['back'].append(_rows(Hn, H, hh[i], ld[b], ld[b], ld[b]))
            steps['norms'].append(_rows(T, _ptr(norms, nr_at[i]), nr[i], n[b], ld[b], 1))
        launches = _Launches(extents)
        launches.gemm(steps['x'])          # 0
        launches.gemm(steps['xx'])         # 1
        launches.rows('copy', steps['copy'])   # 2
        launches.gemm(steps['acc1'])       # 3
        launches.gemm(steps['acc2'])       # 4
        launches.gemm(steps['gram'])       # 5
        launches.gemm(steps['rotate'])     # 6
        launches.rows('copy', steps['back'])   # 7
        launches.rows('norms', steps['norms'])  # 8
        keep = launches.upload()
        launches.run(0, 6)
        _native.eigh_batched(grams, t, sq_at[:-1].tolist(), vectors=vectors)
        launches.run(6)
        host = norms.cpu().numpy()
        del keep, scale
        for i, (b, f, h) in enumerate(heads):
            values[b][h:] = host[nr_at[i]:nr_at[i + 1]]
            self.first[b] = h

    def factor(self, cut: Callable, reliable: float, resolved: float) -> List[np.ndarray]:
        values = self._first_pass()
        self.first = [0] * len(self.blocks)
        while True:
            tails = choose_tails(values, self.first, cut(values), reliable, resolved)
            if not tails:
                break
            self._tail_pass(tails, values)
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
