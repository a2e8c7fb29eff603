'''
The device library through its C ABI (include/tentorch_amd.h), against numpy: the grouped block
GEMM (FP64 MFMA within 1e-12 relative, FP32 MFMA within 2e-5 relative to the operand scale), the
gather-scale-scatter tables the host engine plans (bit-exact against the oracle's leg moves), and the
fixed-order scalar product.
'''

import ctypes

import numpy as np
import pytest
import torch

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def lib():
    return cabi.load()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _random_gemm(rng, shapes, multi_term=False):
    '''Flat A, B buffers and the tables of one grouped GEMM over the given (m, n, k) blocks.'''
    groups, terms = [], []
    a_parts, b_parts = [], []
    a_at = b_at = c_at = 0
    for m, n, k in shapes:
        first = len(terms)
        ks = [k, max(1, k // 3)] if multi_term and k else ([k] if k else [])
        for kk in ks:
            a_parts.append(rng.standard_normal(m * kk))
            b_parts.append(rng.standard_normal(kk * n))
            terms.append((a_at, b_at, kk))
            a_at += m * kk
            b_at += kk * n
        groups.append((c_at, m, n, first, len(terms) - first))
        c_at += m * n
    a = np.concatenate(a_parts) if a_parts else np.zeros(1)
    b = np.concatenate(b_parts) if b_parts else np.zeros(1)
    return a, b, c_at, groups, terms


def _reference_gemm(a, b, c_size, groups, terms):
    c = np.zeros(c_size)
    for c_off, m, n, first, count in groups:
        acc = np.zeros((m, n))
        for t in range(first, first + count):
            ao, bo, k = terms[t]
            acc += a[ao:ao + m * k].reshape(m, k) @ b[bo:bo + k * n].reshape(k, n)
        c[c_off:c_off + m * n] = acc.reshape(-1)
    return c


SHAPES = [(1, 1, 1), (6, 1, 1), (64, 64, 16), (65, 63, 17), (130, 7, 200), (3, 129, 5), (200, 190, 70), (17, 33, 0),
          # deep products, which the library splits along the depth (partial sums added in order)
          (40, 30, 1500), (5, 70, 900), (300, 260, 700)]


@pytest.mark.parametrize('multi_term', [False, True])
@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_grouped_gemm_matches_numpy(lib, multi_term, precision):
    rng = np.random.default_rng(3)
    a, b, c_size, groups, terms = _random_gemm(rng, SHAPES, multi_term)
    ref = _reference_gemm(a, b, c_size, groups, terms)
    g, t = cabi.gemm_tables(groups, terms)
    plan = ctypes.c_void_p()
    assert lib.tt_ggemm_plan_create(g, len(groups), t, len(terms), ctypes.byref(plan)) == 0, cabi.error(lib)
    dtype = torch.float64 if precision == 'f64' else torch.float32
    code = cabi.TT_F64 if precision == 'f64' else cabi.TT_F32
    da, db = torch.tensor(a, dtype=dtype, device='cuda'), torch.tensor(b, dtype=dtype, device='cuda')
    dc = torch.full((c_size,), float('nan'), dtype=dtype, device='cuda')
    assert lib.tt_ggemm_execute(plan, code, code, code, code, da.data_ptr(), db.data_ptr(), dc.data_ptr(), _stream()) == 0, cabi.error(lib)
    got = dc.double().cpu().numpy()
    scale = np.abs(ref).max()
    tol = 1e-12 if precision == 'f64' else 2e-5
    assert np.all(np.isfinite(got))
    assert np.abs(got - ref).max() <= tol * scale
    # the per-block form gives the same blocks
    dc2 = torch.full((c_size,), float('nan'), dtype=dtype, device='cuda')
    assert lib.tt_ggemm_execute_plain(plan, code, da.data_ptr(), db.data_ptr(), dc2.data_ptr(), _stream()) == 0, cabi.error(lib)
    assert np.abs(dc2.double().cpu().numpy() - ref).max() <= tol * scale
    lib.tt_ggemm_plan_destroy(plan)


def test_mixed_precision_gemm_reads_fp64_writes_fp64(lib):
    rng = np.random.default_rng(4)
    a, b, c_size, groups, terms = _random_gemm(rng, [(100, 80, 60), (33, 5, 7)])
    ref = _reference_gemm(a, b, c_size, groups, terms)
    g, t = cabi.gemm_tables(groups, terms)
    plan = ctypes.c_void_p()
    assert lib.tt_ggemm_plan_create(g, len(groups), t, len(terms), ctypes.byref(plan)) == 0
    da = torch.tensor(a, dtype=torch.float64, device='cuda')
    db = torch.tensor(b, dtype=torch.float64, device='cuda')
    dc = torch.zeros(c_size, dtype=torch.float64, device='cuda')
    assert lib.tt_ggemm_execute(plan, cabi.TT_F64, cabi.TT_F64, cabi.TT_F64, cabi.TT_F32,
                                da.data_ptr(), db.data_ptr(), dc.data_ptr(), _stream()) == 0, cabi.error(lib)
    assert np.abs(dc.cpu().numpy() - ref).max() <= 2e-5 * np.abs(ref).max()
    lib.tt_ggemm_plan_destroy(plan)


def test_dot_is_fixed_order_and_accurate(lib):
    rng = np.random.default_rng(5)
    for n in (0, 1, 1000, 1 << 20):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        dx, dy = torch.tensor(x, device='cuda'), torch.tensor(y, device='cuda')
        out = torch.zeros(2, dtype=torch.float64, device='cuda')
        assert lib.tt_dot(cabi.TT_F64, n, dx.data_ptr(), dy.data_ptr(), out.data_ptr(), _stream()) == 0
        first = out[0].item()
        assert lib.tt_dot(cabi.TT_F64, n, dx.data_ptr(), dy.data_ptr(), out.data_ptr(), _stream()) == 0
        assert out[0].item() == first
        exact = float(np.dot(x, y))
        assert abs(first - exact) <= 1e-13 * max(1.0, np.abs(x).dot(np.abs(y)))


def test_leg_move_tables_are_bit_exact_against_the_oracle(lib):
    '''Every placement of a 4-leg U(1) tensor: the engine's table run on the GPU equals the oracle's
    per-entry chain walk bit for bit.'''
    import itertools
    from tentorch_amd._native import engine
    from tentorch_amd import u1, Leg
    import oracle.symmetry as osym
    import oracle.tensor as ot
    rng = np.random.default_rng(0)
    bond = ([(-2,), (0,), (2,)], [2, 3, 2])
    phys = ([(-1,), (1,)], [1, 1])
    o = ot.random_tensor(osym.u1, [osym.Leg(osym.u1, *bond), osym.Leg(osym.u1, *phys)],
                         [osym.Leg(osym.u1, *phys), osym.Leg(osym.u1, *bond)], (0,), rng)
    s = engine.Structure(u1, [Leg(u1, *bond), Leg(u1, *phys)], [Leg(u1, *phys), Leg(u1, *bond)], (0,))
    src = torch.tensor(o.flat(), device='cuda')
    for n_above in range(5):
        for perm in itertools.permutations(range(4)):
            above, below = list(perm[:n_above]), list(perm[n_above:])
            program = engine.plan_place(s, above, below)
            ref = o.place(above, below).flat()
            if 'move' not in program.op_kinds():
                continue
            tasks, sources = program.move_tables(0)
            tt, ss = cabi.gss_tables(tasks, sources)
            plan = ctypes.c_void_p()
            assert lib.tt_gss_plan_create(tt, len(tasks), ss, len(sources), ctypes.byref(plan)) == 0, cabi.error(lib)
            dst = torch.full((program.out.size,), float('nan'), dtype=torch.float64, device='cuda')
            assert lib.tt_gss_execute(plan, cabi.TT_F64, src.data_ptr(), dst.data_ptr(), _stream()) == 0
            assert np.array_equal(dst.cpu().numpy(), ref), (above, below)
            lib.tt_gss_plan_destroy(plan)


def test_krylov_kernels_match_numpy(lib):
    '''tt_multidot / tt_multiaxpy against numpy on a 7-row basis.'''
    rng = np.random.default_rng(8)
    n, k = 100003, 7
    V = torch.tensor(rng.standard_normal((k + 2, n)), device='cuda')
    w = torch.tensor(rng.standard_normal(n), device='cuda')
    out = torch.zeros(k, dtype=torch.float64, device='cuda')
    assert lib.tt_multidot(n, k, V.data_ptr(), V.stride(0), w.data_ptr(), out.data_ptr(), _stream()) == 0, cabi.error(lib)
    ref = V[:k].cpu().numpy() @ w.cpu().numpy()
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    before = w.clone()
    assert lib.tt_multiaxpy(n, k, out.data_ptr(), -0.5, V.data_ptr(), V.stride(0), w.data_ptr(), _stream()) == 0
    ref_w = before.cpu().numpy() - 0.5 * (out.cpu().numpy() @ V[:k].cpu().numpy())
    assert np.abs(w.cpu().numpy() - ref_w).max() <= 1e-12 * np.abs(ref_w).max()


@pytest.mark.parametrize('k,stop', [(1, None), (2, None), (5, None), (7, 3), (13, None), (31, 0), (32, 20)])
def test_lanczos_ritz_matches_numpy(lib, k, stop):
    '''tt_lanczos_ritz against numpy's eigh of the tridiagonal the reference builds: the size stops at
    the first beta below tol, the least eigenvalue to 1e-12, its eigenvector (first component made
    non-negative) to 1e-10, zero coefficients past the size.'''
    rng = np.random.default_rng(100 + k)
    a = rng.standard_normal(k)
    b = np.abs(rng.standard_normal(k)) + 0.1
    if stop is not None:
        b[stop] = 1e-15
    size = k
    for j in range(k):
        if b[j] < 1e-12:
            size = j + 1
            break
    matrix = np.diag(a[:size]) + np.diag(b[:size - 1], 1) + np.diag(b[:size - 1], -1)
    values, vectors = np.linalg.eigh(matrix)
    ref = vectors[:, 0] * (1.0 if vectors[0, 0] >= 0 else -1.0)
    da, db = torch.tensor(a, device='cuda'), torch.tensor(b, device='cuda')
    coef = torch.full((k + 1,), float('nan'), dtype=torch.float64, device='cuda')
    ritz = torch.zeros(1, dtype=torch.float64, device='cuda')
    dsize = torch.zeros(1, dtype=torch.int32, device='cuda')
    assert lib.tt_lanczos_ritz(k, da.data_ptr(), db.data_ptr(), 1e-12, coef.data_ptr(), ritz.data_ptr(),
                               dsize.data_ptr(), _stream()) == 0, cabi.error(lib)
    got = coef.cpu().numpy()
    assert int(dsize.item()) == size
    assert abs(ritz.item() - values[0]) <= 1e-12 * max(1.0, np.abs(values).max())
    assert np.abs(got[:size] - ref).max() <= 1e-10
    assert np.all(got[size:k] == 0.0)


def test_counted_multiaxpy_skips_rows_past_the_count(lib):
    '''tt_multiaxpy_counted reads only the first min(k, count) rows: rows past it hold NaN and the
    result stays finite and equal to numpy's.'''
    rng = np.random.default_rng(14)
    n, k, count = 10007, 6, 4
    V = torch.tensor(rng.standard_normal((k, n)), device='cuda')
    V[count:] = float('nan')
    coef = torch.tensor(rng.standard_normal(k), device='cuda')
    w = torch.tensor(rng.standard_normal(n), device='cuda')
    before = w.cpu().numpy()
    dcount = torch.tensor([count], dtype=torch.int32, device='cuda')
    assert lib.tt_multiaxpy_counted(n, k, dcount.data_ptr(), coef.data_ptr(), 0.5, V.data_ptr(), V.stride(0),
                                    w.data_ptr(), _stream()) == 0, cabi.error(lib)
    ref = before + 0.5 * (coef.cpu().numpy()[:count] @ V[:count].cpu().numpy())
    got = w.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_lanczos_orthonormalization_matches_numpy(lib):
    '''tt_lanczos_orthonormalize against numpy's two classical Gram-Schmidt passes on a 6-vector
    orthonormal basis: the projection on the last basis vector, the norm left, and the next basis
    vector (orthogonal to the basis to 1e-13).'''
    rng = np.random.default_rng(10)
    n, k = 50021, 5
    q, _ = np.linalg.qr(rng.standard_normal((n, k + 1)))
    w0 = rng.standard_normal(n) + 3.0 * q[:, k] - 2.0 * q[:, k - 1]
    V = torch.zeros((k + 2, n), dtype=torch.float64, device='cuda')
    V[:k + 1] = torch.tensor(q.T.copy(), device='cuda')
    w = torch.tensor(w0, device='cuda')
    coef = torch.zeros(2 * (k + 1) + 1, dtype=torch.float64, device='cuda')
    alpha = torch.zeros(1, dtype=torch.float64, device='cuda')
    beta = torch.zeros(1, dtype=torch.float64, device='cuda')
    assert lib.tt_lanczos_orthonormalize(n, k, V.data_ptr(), V.stride(0), w.data_ptr(), coef.data_ptr(),
                                         alpha.data_ptr(), beta.data_ptr(), _stream()) == 0, cabi.error(lib)
    first = q.T @ w0
    once = w0 - q @ first
    second = q.T @ once
    twice = once - q @ second
    left = np.linalg.norm(twice)
    assert abs(alpha.item() - (first[k] + second[k])) <= 1e-12 * abs(first[k])
    assert abs(beta.item() - left) <= 1e-12 * left
    nxt = V[k + 1].cpu().numpy()
    assert np.abs(nxt - twice / left).max() <= 1e-12 * np.abs(twice / left).max()
    assert np.abs(q.T @ nxt).max() <= 1e-13


def test_typed_leg_move_converts_between_fp64_and_fp32(lib):
    '''A composed move that reads FP64 and writes FP32 equals the FP64 move rounded to FP32.'''
    from tentorch_amd._native import engine
    from tentorch_amd import u1, Leg
    rng = np.random.default_rng(9)
    bond = ([(-2,), (0,), (2,)], [2, 3, 2])
    phys = ([(-1,), (1,)], [1, 1])
    s = engine.Structure(u1, [Leg(u1, *bond), Leg(u1, *phys)], [Leg(u1, *phys), Leg(u1, *bond)], (0,))
    program = engine.plan_place(s, [3, 0], [1, 2])
    tasks, sources = program.move_tables(0)
    tt, ss = cabi.gss_tables(tasks, sources)
    plan = ctypes.c_void_p()
    assert lib.tt_gss_plan_create(tt, len(tasks), ss, len(sources), ctypes.byref(plan)) == 0
    src = torch.tensor(rng.standard_normal(s.size), device='cuda')
    full = torch.zeros(program.out.size, dtype=torch.float64, device='cuda')
    low = torch.zeros(program.out.size, dtype=torch.float32, device='cuda')
    back = torch.zeros(s.size, dtype=torch.float64, device='cuda')
    assert lib.tt_gss_execute(plan, cabi.TT_F64, src.data_ptr(), full.data_ptr(), _stream()) == 0
    assert lib.tt_gss_execute_typed(plan, cabi.TT_F64, cabi.TT_F32, src.data_ptr(), low.data_ptr(), _stream()) == 0
    assert torch.equal(low, full.float())
    lib.tt_gss_plan_destroy(plan)


def test_plans_created_in_an_upload_batch_run_like_plans_created_alone(lib):
    '''Leg-move plans created inside nested tt_upload_batch_begin/end have their tables copied
    together when the outer batch ends. One run while both batches are open (its staged tables are
    issued first), one between the two ends and one after them give the bits of the same plans
    created outside any batch. An end with no batch open is refused.'''
    from tentorch_amd._native import engine
    from tentorch_amd import u1, Leg
    rng = np.random.default_rng(11)
    bond = ([(-2,), (0,), (2,)], [2, 3, 2])
    phys = ([(-1,), (1,)], [1, 1])
    s = engine.Structure(u1, [Leg(u1, *bond), Leg(u1, *phys)], [Leg(u1, *phys), Leg(u1, *bond)], (0,))
    src = torch.tensor(rng.standard_normal(s.size), device='cuda')
    moves = []
    for above, below in (([3, 0], [1, 2]), ([1], [3, 2, 0]), ([2, 3, 1], [0])):
        program = engine.plan_place(s, above, below)
        assert 'move' in program.op_kinds()
        tasks, sources = program.move_tables(0)
        moves.append((cabi.gss_tables(tasks, sources), len(tasks), len(sources), program.out.size))

    def create(move):
        (tasks, sources), num_tasks, num_sources, _ = move
        plan = ctypes.c_void_p()
        assert lib.tt_gss_plan_create(tasks, num_tasks, sources, num_sources, ctypes.byref(plan)) == 0, cabi.error(lib)
        return plan

    def run(plan, move):
        dst = torch.full((move[3],), float('nan'), dtype=torch.float64, device='cuda')
        assert lib.tt_gss_execute(plan, cabi.TT_F64, src.data_ptr(), dst.data_ptr(), _stream()) == 0, cabi.error(lib)
        return dst.cpu()

    alone = []
    for move in moves:
        plan = create(move)
        alone.append(run(plan, move))
        lib.tt_gss_plan_destroy(plan)
    assert lib.tt_upload_batch_end() != 0
    assert 'no batch is open' in cabi.error(lib)
    assert lib.tt_upload_batch_begin() == 0
    assert lib.tt_upload_batch_begin() == 0
    plans = [create(move) for move in moves]
    got = [run(plans[0], moves[0])]
    assert lib.tt_upload_batch_end() == 0
    got.append(run(plans[1], moves[1]))
    assert lib.tt_upload_batch_end() == 0
    got.append(run(plans[2], moves[2]))
    for a, b in zip(alone, got):
        assert torch.equal(a, b)
    for plan in plans:
        lib.tt_gss_plan_destroy(plan)


@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_every_launch_configuration_gives_identical_products(lib, precision):
    '''All tile shapes and both packings: the same values bit for bit (each output element sums its
    terms and k-slabs in one order whatever the tile; FP32 tiles permute the depths within each
    16-deep chunk, alike for every shape), within 1e-12 (FP64) or 2e-5 (FP32) of numpy.'''
    rng = np.random.default_rng(11)
    a, b, c_size, groups, terms = _random_gemm(rng, SHAPES, multi_term=True)
    ref = _reference_gemm(a, b, c_size, groups, terms)
    g, t = cabi.gemm_tables(groups, terms)
    dtype = torch.float64 if precision == 'f64' else torch.float32
    code = cabi.TT_F64 if precision == 'f64' else cabi.TT_F32
    tol = 1e-12 if precision == 'f64' else 2e-5
    da, db = torch.tensor(a, dtype=dtype, device='cuda'), torch.tensor(b, dtype=dtype, device='cuda')
    first = None
    for config in range(lib.tt_ggemm_num_configs()):
        plan = ctypes.c_void_p()
        assert lib.tt_ggemm_plan_create_config(g, len(groups), t, len(terms), config, ctypes.byref(plan)) == 0, cabi.error(lib)
        assert lib.tt_ggemm_plan_config(plan) == config
        dc = torch.full((c_size,), float('nan'), dtype=dtype, device='cuda')
        assert lib.tt_ggemm_execute(plan, code, code, code, code, da.data_ptr(), db.data_ptr(),
                                    dc.data_ptr(), _stream()) == 0, cabi.error(lib)
        lib.tt_ggemm_plan_destroy(plan)
        got = dc.double().cpu().numpy()
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), config
        if first is None:
            first = got
        assert np.array_equal(got, first), config


@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_deep_products_are_split_and_match_the_per_block_form(lib, precision):
    '''Products much deeper than they are wide run as depth parts plus one launch that adds the
    parts' partial sums in order: two launches for the whole plan, twice for the same values, within
    1e-12 (FP64) of the per-block form and of numpy, with the last rows and columns of ragged
    blocks written.'''
    rng = np.random.default_rng(13)
    shapes = [(40, 30, 1500), (5, 70, 900), (300, 260, 700), (33, 17, 40), (2, 3, 0)]
    a, b, c_size, groups, terms = _random_gemm(rng, shapes, multi_term=True)
    ref = _reference_gemm(a, b, c_size, groups, terms)
    g, t = cabi.gemm_tables(groups, terms)
    plan = ctypes.c_void_p()
    assert lib.tt_ggemm_plan_create(g, len(groups), t, len(terms), ctypes.byref(plan)) == 0, cabi.error(lib)
    dtype = torch.float64 if precision == 'f64' else torch.float32
    code = cabi.TT_F64 if precision == 'f64' else cabi.TT_F32
    da, db = torch.tensor(a, dtype=dtype, device='cuda'), torch.tensor(b, dtype=dtype, device='cuda')
    results = []
    for _ in range(2):
        dc = torch.full((c_size,), float('nan'), dtype=dtype, device='cuda')
        before = lib.tt_launch_count()
        assert lib.tt_ggemm_execute(plan, code, code, code, code, da.data_ptr(), db.data_ptr(), dc.data_ptr(),
                                    _stream()) == 0, cabi.error(lib)
        assert lib.tt_launch_count() - before == 2
        results.append(dc.double().cpu().numpy())
    assert np.array_equal(results[0], results[1])
    plain = torch.full((c_size,), float('nan'), dtype=dtype, device='cuda')
    assert lib.tt_ggemm_execute_plain(plan, code, da.data_ptr(), db.data_ptr(), plain.data_ptr(), _stream()) == 0, cabi.error(lib)
    lib.tt_ggemm_plan_destroy(plan)
    tol = 1e-12 if precision == 'f64' else 2e-5
    scale = np.abs(ref).max()
    assert np.all(np.isfinite(results[0]))
    assert np.abs(results[0] - ref).max() <= tol * scale
    assert np.abs(results[0] - plain.double().cpu().numpy()).max() <= tol * scale


@pytest.mark.parametrize('config', [0, 1, 5, 13])
def test_complex_grouped_gemm_matches_numpy(lib, config):
    '''complex128 block products (four FP64 MFMAs per step), grouped and per block.'''
    rng = np.random.default_rng(12)
    _, _, c_size, groups, terms = _random_gemm(rng, SHAPES, multi_term=True)
    a_size = max(t[0] + g[1] * t[2] for g in groups for t in terms[g[3]:g[3] + g[4]])
    b_size = max(t[1] + t[2] * g[2] for g in groups for t in terms[g[3]:g[3] + g[4]])
    a = rng.standard_normal(a_size) + 1j * rng.standard_normal(a_size)
    b = rng.standard_normal(b_size) + 1j * rng.standard_normal(b_size)
    ref = np.zeros(c_size, dtype=np.complex128)
    for c_off, m, n, first, count in groups:
        acc = np.zeros((m, n), dtype=np.complex128)
        for ao, bo, k in terms[first:first + count]:
            acc += a[ao:ao + m * k].reshape(m, k) @ b[bo:bo + k * n].reshape(k, n)
        ref[c_off:c_off + m * n] = acc.reshape(-1)
    g, t = cabi.gemm_tables(groups, terms)
    plan = ctypes.c_void_p()
    assert lib.tt_ggemm_plan_create_config(g, len(groups), t, len(terms), config, ctypes.byref(plan)) == 0, cabi.error(lib)
    da, db = torch.tensor(a, device='cuda'), torch.tensor(b, device='cuda')
    for plain in (False, True):
        dc = torch.full((c_size,), complex('nan'), dtype=torch.complex128, device='cuda')
        if plain:
            status = lib.tt_ggemm_execute_plain(plan, cabi.TT_C128, da.data_ptr(), db.data_ptr(), dc.data_ptr(), _stream())
        else:
            status = lib.tt_ggemm_execute(plan, cabi.TT_C128, cabi.TT_C128, cabi.TT_C128, cabi.TT_C128, da.data_ptr(),
                                          db.data_ptr(), dc.data_ptr(), _stream())
        assert status == 0, cabi.error(lib)
        assert np.abs(dc.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max(), plain
    lib.tt_ggemm_plan_destroy(plan)


def _eigh_cases(rng):
    '''Symmetric matrices of every kind the split meets: random, Gram matrices with geometric
    spectra and rank deficits, exact degeneracies, diagonal, zero and identity (small and above
    order 128), sizes 1 to 700.'''
    cases = []
    for n in (1, 2, 3, 5, 17, 64, 130):
        a = rng.standard_normal((n, n))
        cases.append(a + a.T)
    for n in (50, 384, 700):
        s = np.geomspace(1, 1e-9, n)
        s[(2 * n) // 3:] = 0
        u, _ = np.linalg.qr(rng.standard_normal((n, n)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        b = (u * s) @ v.T
        cases.append(b @ b.T)
    half = rng.standard_normal((100, 100))
    half = half + half.T
    p, _ = np.linalg.qr(rng.standard_normal((200, 200)))
    cases.append(p @ np.kron(np.eye(2), half) @ p.T)
    cases.append(np.diag(rng.standard_normal(40)))
    cases.append(np.zeros((9, 9)))
    cases.append(np.eye(30))
    # above the LDS tail's 128 rows: every step with nothing to reflect, every merge fully deflated
    cases.append(np.zeros((260, 260)))
    cases.append(2.5 * np.eye(300))
    return cases


@pytest.mark.parametrize('largest', [2048, 300, 192, 160, 128])
def test_batched_eigh_matches_numpy(largest):
    '''tt_eigh_batched on one batch of mixed sizes: eigenvalues within 1e-12 of numpy's (relative to
    the matrix scale), eigenvectors orthonormal to 1e-12 with residuals below 1e-12 of the scale.
    The tridiagonalisation runs one launch per step over many workgroups until the largest trailing
    matrix has 192 rows, then finishes every matrix in one workgroup with the trailing matrix in
    registers: a batch reaching order 2048 (the largest taken), one reaching 300 (matrices of order
    260, 200 and 193 hand over part-way, 193 after one stepped launch), one reaching exactly 192 (no
    stepped launch, the register workgroup's full size), one reaching 160 and one with nothing above
    128 (orders 16, 33 and 177 end a sixteen-column group exactly, just past one, and part-way).'''
    from tentorch_amd import _native
    assert _native.EIGH_MAX_ORDER == 2048
    rng = np.random.default_rng(21)
    cases = [c for c in _eigh_cases(rng) if c.shape[0] <= largest]
    for n in {2048: (1024, 2048), 300: (300, 193), 192: (192, 191, 177), 160: (160, 159, 33), 128: (16,)}[largest]:
        a = rng.standard_normal((n, n))
        cases.append(a + a.T)
    packed = torch.tensor(np.concatenate([c.reshape(-1) for c in cases]), device='cuda')
    # a workspace full of garbage: nothing may depend on its contents
    size = _native.eigh_workspace([c.shape[0] for c in cases])
    work = torch.full((size,), float('nan'), dtype=torch.float64, device='cuda')
    values, vectors = _native.eigh_batched(packed, [c.shape[0] for c in cases], work=work)
    values, vectors = values.cpu().numpy(), vectors.cpu().numpy()
    at = vat = 0
    for c in cases:
        n = c.shape[0]
        lam = values[vat:vat + n]
        rows = vectors[at:at + n * n].reshape(n, n)      # eigenvector j is row j
        ref = np.linalg.eigvalsh(c)
        scale = max(1.0, np.abs(ref).max())
        assert np.all(np.diff(lam) >= 0), n
        assert np.abs(lam - ref).max() <= 1e-12 * scale, n
        assert np.abs(rows @ rows.T - np.eye(n)).max() <= 1e-12, n
        assert np.abs(c @ rows.T - rows.T * lam).max() <= 1e-12 * scale, n
        at += n * n
        vat += n


def test_batched_gemm_with_transposes_matches_numpy():
    '''tt_bgemm over one table mixing every transpose pair, ragged sizes (1 to 150, not multiples of
    the 64 x 64 tile or the 16-deep slab), an empty depth, column scales, the division by column
    minus row shifts, alpha and beta.'''
    from tentorch_amd import _native
    from tentorch_amd.gram import _Launches, _Stage, _bgemm
    rng = np.random.default_rng(51)
    shapes = [(1, 1, 1), (17, 5, 33), (64, 64, 16), (65, 130, 150), (150, 3, 97), (40, 70, 0), (3, 129, 64)]
    cases, buffers = [], []
    extents = _native.Extents()
    for i, (m, n, k) in enumerate(shapes):
        ta, tb = i % 2, (i // 2) % 2
        a = rng.standard_normal((k, m) if ta else (m, k))
        b = rng.standard_normal((n, k) if tb else (k, n))
        c = rng.standard_normal((m, n))
        scale = rng.standard_normal(n) + 3.0 if i % 3 != 2 else None
        shift = rng.standard_normal(m) - 3.0 if i % 3 == 1 else None
        alpha, beta = 0.5 + i, (0.0 if i % 2 else -1.5)
        product = (a.T if ta else a) @ (b.T if tb else b)
        if shift is not None:
            product = product / (scale[None, :] - shift[:, None])
        elif scale is not None:
            product = product * scale
        ref = alpha * product + beta * c
        da, db, dc = (torch.tensor(x, device='cuda').contiguous() for x in (a, b, c))
        ds = torch.tensor(scale, device='cuda') if scale is not None else None
        dr = torch.tensor(shift, device='cuda') if shift is not None else None
        for t in (da, db, dc, ds, dr):
            if t is not None:
                extents.add(t)
        cases.append((dc, ref))
        buffers.append((da, db, dc, ds, dr))
        lda, ldb = (m if ta else k), (k if tb else n)
        buffers[-1] += (dict(a=da.data_ptr(), b=db.data_ptr(), c=dc.data_ptr(), m=m, n=n, k=k, lda=max(lda, 1),
                             ldb=max(ldb, 1), ldc=n, trans_a=ta, trans_b=tb, alpha=alpha, beta=beta,
                             col_scale=ds.data_ptr() if ds is not None else 0,
                             row_shift=dr.data_ptr() if dr is not None else 0),)
    problems = [entry[-1] for entry in buffers]
    launches = _Launches(extents)
    launches.add('gemm', _bgemm(len(problems), **{key: [p[key] for p in problems] for key in problems[0]}))
    stage = _Stage(torch.device('cuda'), [launches.steps[0][1].nbytes])
    launches.stage(stage, 0)
    launches.run()
    torch.cuda.synchronize()
    for dc, ref in cases:
        assert np.abs(dc.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_row_norms_and_row_copies_match_numpy():
    '''tt_row_norms and tt_copy_rows (with and without a row index, straight and transposed) on
    matrices with row strides longer than their rows.'''
    from tentorch_amd import _native
    from tentorch_amd.gram import _Launches, _Stage, _rowwise
    rng = np.random.default_rng(52)
    src = torch.tensor(rng.standard_normal((70, 90)), device='cuda')       # rows of 80, stride 90
    norms = torch.empty(70, dtype=torch.float64, device='cuda')
    straight = torch.full((70, 85), float('nan'), dtype=torch.float64, device='cuda')
    turned = torch.full((80, 75), float('nan'), dtype=torch.float64, device='cuda')
    order = rng.permutation(70)[:33].astype(np.int32)
    index = torch.tensor(order, device='cuda')
    extents = _native.Extents().add(src).add(norms).add(straight).add(turned).add(index)
    launches = _Launches(extents)
    launches.add('norms', _rowwise(_native.ROWS_PROBLEM, 1, src=src.data_ptr(), dst=norms.data_ptr(), rows=70, cols=80,
                                   src_ld=90, dst_ld=1))
    launches.add('copy', _rowwise(_native.ROWS_PROBLEM, 2, src=src.data_ptr(), dst=[straight.data_ptr(), turned.data_ptr()],
                                  rows=[70, 33], cols=80, src_ld=90, dst_ld=[85, 75], transpose=[0, 1],
                                  index=[0, index.data_ptr()]),
                 source_rows=[70, 70])
    stage = _Stage(torch.device('cuda'), [table.nbytes for _, table, _ in launches.steps])
    launches.stage(stage, 0)
    launches.run()
    host = src.cpu().numpy()[:, :80]
    assert np.abs(norms.cpu().numpy() - np.linalg.norm(host, axis=1)).max() <= 1e-13 * np.abs(host).max() * 10
    assert np.array_equal(straight.cpu().numpy()[:, :80], host)
    assert np.all(np.isnan(straight.cpu().numpy()[:, 80:]))
    assert np.array_equal(turned.cpu().numpy()[:, :33], host[order].T)
    assert np.all(np.isnan(turned.cpu().numpy()[:, 33:]))
