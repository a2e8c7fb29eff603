'''
A CPU interpreter of the device programs the host engine plans: it reads the descriptor tables
(gather-scale-scatter tiles, grouped-GEMM groups and terms) through the engine's inspection hooks
and applies them with numpy, exactly as the device library defines them (include/tentorch_amd.h).

Test infrastructure: it checks the host planner against the oracle without a GPU.  It is not a CPU
path of the product (the product has none).
'''

import numpy as np


def apply_move(tasks, sources, src: np.ndarray, dst: np.ndarray) -> None:
    '''One gather-scale-scatter table: every tile = Σ coef * source box, in source order.'''
    for dst_offset, ext, dstr, first, count in tasks:
        ext = list(ext)
        grids = np.indices(ext).reshape(len(ext), -1) if ext else np.zeros((0, 1), dtype=np.int64)
        doff = dst_offset + (np.asarray(dstr, dtype=np.int64)[:, None] * grids).sum(axis=0) if ext else np.array([dst_offset])
        acc = np.zeros(doff.shape, dtype=dst.dtype)
        for s in range(first, first + count):
            soffset, sstr, coef = sources[s]
            sstr = np.asarray(sstr[:len(ext)], dtype=np.int64)
            soff = soffset + (sstr[:, None] * grids).sum(axis=0) if ext else np.array([soffset])
            c = coef if np.iscomplexobj(dst) else coef.real
            acc = acc + c * src[soff]
        dst[doff] = acc


def apply_gemm(groups, terms, a: np.ndarray, b: np.ndarray, c: np.ndarray) -> None:
    '''One grouped GEMM: every block = Σ_terms A_t B_t, a block with no term = 0.'''
    for c_offset, m, n, first, count in groups:
        acc = np.zeros((m, n), dtype=c.dtype)
        for t in range(first, first + count):
            a_off, b_off, k = terms[t]
            acc = acc + a[a_off:a_off + m * k].reshape(m, k) @ b[b_off:b_off + k * n].reshape(k, n)
        c[c_offset:c_offset + m * n] = acc.reshape(-1)


def run_program(program, operands, dtype=np.float64) -> np.ndarray:
    '''Run a planned program on flat numpy operands; returns the flat result.'''
    slots = program.slots()
    buffers = []
    out = np.zeros(program.out.size, dtype=dtype)
    for kind, index, offset, size in slots:
        if kind == 'operand':
            buffers.append(operands[index])
        elif kind == 'output':
            buffers.append(out)
        else:
            buffers.append(np.zeros(size, dtype=dtype))
    moves = gemms = 0
    for kind, a, b, d, _ in program.ops():
        if kind == 'move':
            tasks, sources = program.move_tables(moves)
            moves += 1
            apply_move(tasks, sources, buffers[a], buffers[d])
        elif kind == 'gemm':
            groups, terms = program.gemm_tables(gemms)
            gemms += 1
            apply_gemm(groups, terms, buffers[a], buffers[b], buffers[d])
        elif kind == 'copy':
            buffers[d][:] = buffers[a]
        else:
            buffers[d][:] += buffers[a]
    return out
