"""Dense complex LU with partial pivoting and the blocked solves (csrc/dev/bq_lu.hip), which
replace the reference's PARDISO / LAPACKE_zgesv calls (d3qd2_src/d3qd2_mkl_solver.c:18-148,
d3qd2_src/d3qd2_solve_bieq.c:753).  Checked against numpy's LAPACK on random systems of
sizes that cover partial panels (n not a multiple of the 32-column block), many and few
right-hand sides, and a matrix that needs pivoting (zero leading diagonal).

Tolerance: backward error |A x - b| / (|A| |x|) < 1e-13 and forward error vs numpy below
1e-9 (condition numbers of these random matrices are ~1e3)."""
import numpy as np
import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("n,nrhs", [(1, 1), (31, 3), (32, 1), (100, 7), (333, 130), (1000, 6), (2053, 64)])
def test_gpu_zgesv_random(gpu_lib, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    X = gpu_lib.zgesv(A, B)
    ref = np.linalg.solve(A, B)
    bw = np.abs(A @ X - B).max() / (np.abs(A).max() * np.abs(X).max() * n)
    fw = np.abs(X - ref).max() / np.abs(ref).max()
    print(n, nrhs, "backward", bw, "forward", fw)
    assert bw < 1e-13 and fw < 1e-9


@pytest.mark.gpu
def test_gpu_zgesv_needs_pivoting(gpu_lib):
    rng = np.random.default_rng(3)
    n = 97
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[0, 0] = 0.0
    A[40:45, 40:45] = 0.0
    b = rng.standard_normal(n) + 0j
    x = gpu_lib.zgesv(A, b)
    assert np.abs(A @ x - b).max() < 1e-11 * np.abs(b).max() * np.abs(A).max() * n


@pytest.mark.gpu
def test_gpu_zgesv_pivots_across_outer_panels(gpu_lib):
    """Zero diagonal blocks straddling the 256-column outer panel boundary and the 32-column
    steps: the interchanges chosen inside one outer panel must reach the columns on both
    sides of it."""
    rng = np.random.default_rng(11)
    n = 600
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[250:262, 250:262] = 0.0
    A[510:515, :] *= 1e-6
    B = rng.standard_normal((n, 40)) + 1j * rng.standard_normal((n, 40))
    X = gpu_lib.zgesv(A, B)
    ref = np.linalg.solve(A, B)
    bw = np.abs(A @ X - B).max() / (np.abs(A).max() * np.abs(X).max() * n)
    assert bw < 1e-13
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-8


# unit roundoff of the FP32 factorisation's updates: FP32 products, or bf16x3 products
# (x = hi + lo + r with |r| <= 2^-17 |x|; the dropped terms lo lo and r y leave ~2^-15)
UPDATE_U = {"fp32": 6e-8, "bf16x3": 2.0 ** -15}


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (31, 3), (100, 7), (333, 130), (1100, 300), (2053, 64)])
def test_gpu_cgesv_random(gpu_lib, n, nrhs, mode):
    """The complex FP32 LU (the preconditioner of the mixed-precision global solve) with
    either update arithmetic (the bf16x3 kernel takes the updates of order >= 128: the
    rank-512 trailing updates for n > 512 and the solve's block updates for >= 128
    right-hand sides): backward error within 16 u, forward error vs numpy's complex128 solve
    of the same (FP32-rounded) system within cond(A) x u x (n + 4), u the update's unit
    roundoff."""
    u = UPDATE_U[mode]
    rng = np.random.default_rng(7 * n + nrhs)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64).astype(complex)
    B = (rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))).astype(np.complex64).astype(complex)
    old = gpu_lib.lu_update_mode(mode)
    try:
        X = gpu_lib.cgesv(A, B).astype(complex)
    finally:
        gpu_lib.lu_update_mode(old)
    ref = np.linalg.solve(A, B)
    bw = np.abs(A @ X - B).max() / (np.abs(A).max() * np.abs(X).max() * n)
    fw = np.abs(X - ref).max() / np.abs(ref).max()
    cond = np.linalg.cond(A)
    print(mode, n, nrhs, "backward", bw, "forward", fw, "cond", cond)
    assert bw < 16 * u
    assert fw < cond * (n + 4) * u


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["z", "c"])
def test_gpu_lu_lookahead_on_off(gpu_lib, prec):
    """The factorisation's look-ahead (for order > 1024 the columns right of each 512-column
    panel are updated in two parts, the next panel's on the calling stream and the rest on a
    second stream while the next panel is factored) against the one-stream order, at n = 2600:
    five full panels and a partial one, so the second part runs 1576 .. 40 columns wide (the
    last below the 128 at which the FP32 path solves U12 by one product with inv(L11)).  A
    matrix with a small leading block forces interchanges across panels.  Both solves meet
    the backward-error bound of their precision and agree within cond(A) (n + 4) u."""
    n, nrhs = 2600, 5
    rng = np.random.default_rng(29)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[:700, :700] *= 1e-3
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    if prec == "c":
        A = A.astype(np.complex64).astype(complex)
        B = B.astype(np.complex64).astype(complex)
        u, bw_max = UPDATE_U["bf16x3"], 16 * UPDATE_U["bf16x3"]
        solve = lambda: gpu_lib.cgesv(A, B).astype(complex)
    else:
        u, bw_max = 1.1e-16, 1e-13
        solve = lambda: gpu_lib.zgesv(A, B)
    old = gpu_lib.lu_lookahead(True)
    try:
        X1 = solve()
        gpu_lib.lu_lookahead(False)
        X0 = solve()
    finally:
        gpu_lib.lu_lookahead(old)
    cond = np.linalg.cond(A)
    for X in (X1, X0):
        bw = np.abs(A @ X - B).max() / (np.abs(A).max() * np.abs(X).max() * n)
        assert bw < bw_max, bw
    d = np.abs(X1 - X0).max() / np.abs(X0).max()
    print(prec, "on/off", d, "cond", cond)
    assert d < cond * (n + 4) * u


def _bf16_split(x):
    """hi, lo (float64) of float32 x: hi = bf16(x), lo = bf16(x - hi), round to nearest even."""
    def rne(v):
        b = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
        b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
        return (b & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    x = np.asarray(x, np.float32)
    hi = rne(x)
    lo = rne((x - hi).astype(np.float32))
    return hi.astype(np.float64), lo.astype(np.float64)


def _bf16x3_product(A, B):
    """A B with every real product taken as hi hi + hi lo + lo hi, summed exactly (float64)."""
    arh, arl = _bf16_split(A.real)
    aih, ail = _bf16_split(A.imag)
    brh, brl = _bf16_split(B.real)
    bih, bil = _bf16_split(B.imag)
    p = lambda xh, xl, yh, yl: xh @ yh + xh @ yl + xl @ yh
    return (p(arh, arl, brh, brl) - p(aih, ail, bih, bil)) + 1j * (p(arh, arl, bih, bil) + p(aih, ail, brh, brl))


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (33, 47, 16), (130, 257, 129), (1000, 1300, 512), (1290, 1030, 300),
                                   (1500, 200, 77)])
def test_gpu_cgemm_split_matches_bf16x3(gpu_lib, m, n, k):
    """bq_cgemm_split_dev (C -= A B on the bf16 matrix cores) against the same bf16x3 products
    summed exactly on the host: the difference is the FP32 accumulation alone (<= 2e-6 of the
    largest |A B| entry; one dropped hi-lo term would leave ~2^-9).  The shapes cover single
    elements, k not a multiple of 16, partial 128 x 128 tiles, tile counts with and without
    the per-XCD remap (multiples of 8) and a partial last group of tile rows; the matrices are
    strided (leading dimensions m + 5, k + 3, m + 7) and the padding rows must stay untouched.
    Entries span 2^-6 .. 2^6 in magnitude."""
    import ctypes
    from hipmem import DeviceArray
    L = gpu_lib.lib()
    V = ctypes.c_void_p
    L.bq_cgemm_split_dev.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, ctypes.c_long,
                                     V, ctypes.c_long]
    rng = np.random.default_rng(m + 3 * n + 7 * k)

    def rnd(r, c):
        s = 2.0 ** rng.uniform(-6, 6, (r, c))
        return ((rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))) * s).astype(np.complex64)

    lda, ldb, ldc = m + 5, k + 3, m + 7
    Acm, Bcm, Ccm = rnd(k, lda), rnd(n, ldb), rnd(n, ldc)   # column-major storage: one row per column
    A, B, C0 = Acm[:, :m].T, Bcm[:, :k].T, Ccm[:, :m].T
    dA, dB, dC = DeviceArray(Acm), DeviceArray(Bcm), DeviceArray(Ccm)
    assert L.bq_cgemm_split_dev(None, m, n, k, dA.ptr, lda, dB.ptr, ldb, dC.ptr, ldc) == 0
    out = dC.get()
    got = out[:, :m].T.astype(complex)
    D = _bf16x3_product(A.astype(complex), B.astype(complex))
    exact = A.astype(complex) @ B.astype(complex)
    err = np.abs(got - (C0.astype(complex) - D)).max() / np.abs(D).max()
    split_err = np.abs(D - exact).max() / np.abs(exact).max()
    print(m, n, k, "vs bf16x3 sum", err, "bf16x3 vs exact", split_err)
    assert np.array_equal(out[:, m:], Ccm[:, m:])
    assert err < 2e-6
    assert split_err < 2.0 ** -13


@pytest.mark.gpu
def test_gpu_cgesv_needs_pivoting(gpu_lib):
    rng = np.random.default_rng(5)
    n = 600
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[250:262, 250:262] = 0.0
    A[0, 0] = 0.0
    b = rng.standard_normal(n) + 0j
    x = gpu_lib.cgesv(A, b).astype(complex)
    assert np.abs(A @ x - b).max() < 1e-5 * np.abs(A).max() * np.abs(x).max() * n


@pytest.mark.gpu
@pytest.mark.parametrize("n", [33, 300, 1100])
def test_gpu_small_rhs_solve_matches_blocked_solve(gpu_lib, n):
    """Up to 8 right-hand sides take the fused per-block solve (one workgroup per 256-row
    diagonal block and a skinny update of the rows outside it); 9 or more take the blocked
    solve with library products.  The same 8 columns solved both ways agree to rounding."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 9)) + 1j * rng.standard_normal((n, 9))
    X8 = gpu_lib.zgesv(A, B[:, :8])
    X9 = gpu_lib.zgesv(A, B)
    assert np.abs(X8 - X9[:, :8]).max() <= 1e-12 * np.abs(X9).max()
    bw = np.abs(A @ X8 - B[:, :8]).max() / (np.abs(A).max() * np.abs(X8).max() * n)
    assert bw < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("single", [0, 1])
def test_gpu_getrs_column_groups(gpu_lib, single):
    """Solves whose right-hand sides come in column groups that are zero in the permuted rows
    above gf (bq_zgetrs_groups_dev / bq_cgetrs_groups_dev, used for Z = M^-1 K in the block
    elimination): the forward solve of a group starts at the 512-row block holding gf.  With
    n = 2048, groups starting at gf = 0, 700 and 1500 (so the skip branch runs for two of
    them) and right-hand sides built zero in the permuted rows above their gf, the grouped
    solve must equal the plain one: 1e-12 (FP64), 1e-5 (FP32) of the largest entry."""
    import ctypes
    from hipmem import DeviceArray
    L = gpu_lib.lib()
    n, ng, per = 2048, 3, 40
    rng = np.random.default_rng(5 + single)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8.0 * np.sqrt(n) * np.eye(n)[::-1]
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    dA = DeviceArray(np.asfortranarray(A).T.astype(cdt))  # column-major storage
    L.bq_lu_inv_doubles.restype = ctypes.c_long
    ninv = L.bq_lu_inv_doubles(n)
    dinv = DeviceArray(np.zeros(ninv, rdt))
    dip = DeviceArray(np.zeros(n, np.int32))
    getrf = L.bq_cgetrf_dev if single else L.bq_zgetrf_dev
    getrs = L.bq_cgetrs_dev if single else L.bq_zgetrs_dev
    getrs_g = L.bq_cgetrs_groups_dev if single else L.bq_zgetrs_groups_dev
    V = ctypes.c_void_p
    getrf.argtypes = [V, ctypes.c_int, V, ctypes.c_long, V, V]
    getrs.argtypes = [V, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, V, V, ctypes.c_long]
    I = ctypes.POINTER(ctypes.c_int)
    getrs_g.argtypes = getrs.argtypes + [ctypes.c_int, I, I, I]
    L.bq_lu_perm_dev.argtypes = [V, ctypes.c_int, ctypes.c_int]
    L.bq_lu_perm_dev.restype = V
    assert getrf(None, n, dA.ptr, n, dip.ptr, dinv.ptr) == 0
    off = L.bq_lu_perm_dev(dinv.ptr, n, single) - dinv.ptr.value
    raw = dinv.get().view(np.uint8)[off:off + 4 * n].view(np.int32)
    assert sorted(raw.tolist()) == list(range(n))
    assert not np.array_equal(raw, np.arange(n))  # the anti-diagonal forces interchanges
    gfs = [0, 700, 1500]
    B = rng.standard_normal((n, ng * per)) + 1j * rng.standard_normal((n, ng * per))
    for g, gf in enumerate(gfs):
        B[raw[:gf], g * per:(g + 1) * per] = 0.0  # zero in permuted rows 0 .. gf-1
    dB1 = DeviceArray(np.asfortranarray(B).T.astype(cdt))
    dB2 = DeviceArray(np.asfortranarray(B).T.astype(cdt))
    assert getrs(None, n, ng * per, dA.ptr, n, dip.ptr, dinv.ptr, dB1.ptr, n) == 0
    g0 = (ctypes.c_int * ng)(*[g * per for g in range(ng)])
    gn = (ctypes.c_int * ng)(*([per] * ng))
    gf = (ctypes.c_int * ng)(*gfs)
    assert getrs_g(None, n, ng * per, dA.ptr, n, dip.ptr, dinv.ptr, dB2.ptr, n, ng, g0, gn, gf) == 0
    X1, X2 = dB1.get(), dB2.get()
    for d in (dA, dinv, dip, dB1, dB2):
        d.free()
    err = np.abs(X1 - X2).max() / np.abs(X1).max()
    ref = np.linalg.solve(A, B).T
    fw = np.abs(X1 - ref).max() / np.abs(ref).max()
    print("single" if single else "double", "grouped vs plain", err, "plain vs numpy", fw)
    assert err < (1e-5 if single else 1e-12)
    assert fw < (1e-4 if single else 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("single,n", [(0, 1024), (1, 1024), (1, 1536)])
def test_gpu_getrs_through_block_inverses(gpu_lib, single, n):
    """Few-column solves through the inverted 512-row diagonal blocks of both factors
    (bq_[zc]lu_blkinv_dev, bq_[zc]getrs_blkinv_dev; the preconditioner and dE/dH solves of the
    single-object reduction) against the plain solve with the same factors, for 1 to 8
    right-hand sides on a system whose anti-diagonal forces interchanges: 1e-12 (FP64) and 1e-5
    (FP32) of the largest entry, and the plain solve against numpy.  An order that is not a
    multiple of 512 has no block inverses and is refused."""
    import ctypes
    from hipmem import DeviceArray
    L = gpu_lib.lib()
    rng = np.random.default_rng(17 + n + single)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8.0 * np.sqrt(n) * np.eye(n)[::-1]
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    V = ctypes.c_void_p
    L.bq_lu_inv_doubles.restype = ctypes.c_long
    L.bq_lu_blkinv_elems.restype = ctypes.c_long
    L.bq_lu_blkinv_elems.argtypes = [ctypes.c_int]
    assert L.bq_lu_blkinv_elems(1000) == 0
    nw = L.bq_lu_blkinv_elems(n)
    assert nw == 2 * n * 512
    getrf = L.bq_cgetrf_dev if single else L.bq_zgetrf_dev
    getrs = L.bq_cgetrs_dev if single else L.bq_zgetrs_dev
    blkinv = L.bq_clu_blkinv_dev if single else L.bq_zlu_blkinv_dev
    getrs_w = L.bq_cgetrs_blkinv_dev if single else L.bq_zgetrs_blkinv_dev
    getrf.argtypes = [V, ctypes.c_int, V, ctypes.c_long, V, V]
    getrs.argtypes = [V, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, V, V, ctypes.c_long]
    blkinv.argtypes = [V, ctypes.c_int, V, ctypes.c_long, V, V]
    getrs_w.argtypes = [V, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, V, V, ctypes.c_long]
    dA = DeviceArray(np.asfortranarray(A).T.astype(cdt))
    dinv = DeviceArray(np.zeros(L.bq_lu_inv_doubles(n), rdt))
    dip = DeviceArray(np.zeros(n, np.int32))
    dW = DeviceArray(np.zeros(nw, cdt))
    assert getrf(None, n, dA.ptr, n, dip.ptr, dinv.ptr) == 0
    assert blkinv(None, n, dA.ptr, n, dinv.ptr, dW.ptr) == 0
    worst = 0.0
    for nrhs in (1, 4, 6, 8):
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        dB1 = DeviceArray(np.asfortranarray(B).T.astype(cdt))
        dB2 = DeviceArray(np.asfortranarray(B).T.astype(cdt))
        assert getrs(None, n, nrhs, dA.ptr, n, dip.ptr, dinv.ptr, dB1.ptr, n) == 0
        assert getrs_w(None, n, nrhs, dA.ptr, n, dinv.ptr, dW.ptr, dB2.ptr, n) == 0
        X1, X2 = dB1.get().T, dB2.get().T
        dB1.free()
        dB2.free()
        err = np.abs(X1 - X2).max() / np.abs(X1).max()
        ref = np.linalg.solve(A, B)
        fw = np.abs(X1 - ref).max() / np.abs(ref).max()
        print("single" if single else "double", n, nrhs, "block inverses vs plain", err, "plain vs numpy", fw)
        assert fw < (1e-4 if single else 1e-10)
        worst = max(worst, err)
    dB = DeviceArray(np.zeros((9, n), cdt))
    assert getrs_w(None, n, 9, dA.ptr, n, dinv.ptr, dW.ptr, dB.ptr, n) != 0  # more than 8 columns
    for d in (dA, dinv, dip, dW, dB):
        d.free()
    assert worst < (1e-5 if single else 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("single,n", [(0, 700), (0, 1536), (1, 1024), (1, 1300)])
def test_gpu_lu_inverse(gpu_lib, single, n):
    """bq_[zc]lu_inverse_dev (B^-1 of the single-object reduction): X = U^-1 L^-1 from the
    factors, whose forward sweep skips the zeros of L^-1, read through the factorization's
    permutation (A^-1[:, perm[i]] = X[:, i]) against the plain solve with the identity as
    right-hand sides (1e-12 FP64, 1e-5 FP32 of the largest entry) and against numpy's inverse;
    orders below, at and between multiples of 512 (full and partial blocks, both FP32 update
    paths), on systems whose anti-diagonal forces interchanges."""
    import ctypes
    from hipmem import DeviceArray
    L = gpu_lib.lib()
    rng = np.random.default_rng(31 + n + single)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8.0 * np.sqrt(n) * np.eye(n)[::-1]
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    V = ctypes.c_void_p
    L.bq_lu_inv_doubles.restype = ctypes.c_long
    getrf = L.bq_cgetrf_dev if single else L.bq_zgetrf_dev
    getrs = L.bq_cgetrs_dev if single else L.bq_zgetrs_dev
    inverse = L.bq_clu_inverse_dev if single else L.bq_zlu_inverse_dev
    getrf.argtypes = [V, ctypes.c_int, V, ctypes.c_long, V, V]
    getrs.argtypes = [V, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, V, V, ctypes.c_long]
    inverse.argtypes = [V, ctypes.c_int, V, ctypes.c_long, V, V, ctypes.c_long]
    L.bq_lu_perm_dev.argtypes = [V, ctypes.c_int, ctypes.c_int]
    L.bq_lu_perm_dev.restype = V
    dA = DeviceArray(np.asfortranarray(A).T.astype(cdt))
    dinv = DeviceArray(np.zeros(L.bq_lu_inv_doubles(n), rdt))
    dip = DeviceArray(np.zeros(n, np.int32))
    assert getrf(None, n, dA.ptr, n, dip.ptr, dinv.ptr) == 0
    off = L.bq_lu_perm_dev(dinv.ptr, n, single) - dinv.ptr.value
    perm = dinv.get().view(np.uint8)[off:off + 4 * n].view(np.int32).copy()
    assert sorted(perm.tolist()) == list(range(n))
    assert not np.array_equal(perm, np.arange(n))
    dX = DeviceArray(np.full((n, n), np.nan, cdt))  # every entry must be written
    assert inverse(None, n, dA.ptr, n, dinv.ptr, dX.ptr, n) == 0
    dI = DeviceArray(np.eye(n, dtype=cdt))
    assert getrs(None, n, n, dA.ptr, n, dip.ptr, dinv.ptr, dI.ptr, n) == 0
    X, Ainv = dX.get().T, dI.get().T
    for d in (dA, dinv, dip, dX, dI):
        d.free()
    Xn = np.empty_like(X)
    Xn[:, perm] = X  # A^-1[:, perm[i]] = X[:, i]
    assert np.isfinite(Xn).all()
    err = np.abs(Xn - Ainv).max() / np.abs(Ainv).max()
    ref = np.linalg.inv(A)
    fw = np.abs(Xn - ref).max() / np.abs(ref).max()
    print("single" if single else "double", n, "inverse vs plain solve", err, "vs numpy", fw)
    assert err < (1e-5 if single else 1e-12)
    assert fw < (1e-4 if single else 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("m,n", [(512, 3000), (200, 77)])
def test_gpu_cgemm_split_store_in_place(gpu_lib, m, n):
    """bq_cgemm_split_store_dev with B == C (C <- -A C in place, as the solves apply their
    inverted 512-row diagonal blocks): equal to -A C_old summed exactly with bf16x3 products,
    within the FP32 accumulation (2e-6 of the largest entry); the padding rows of the strided
    C (leading dimension m + 9) stay untouched."""
    import ctypes
    from hipmem import DeviceArray
    L = gpu_lib.lib()
    V = ctypes.c_void_p
    L.bq_cgemm_split_store_dev.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V,
                                           ctypes.c_long, V, ctypes.c_long]
    rng = np.random.default_rng(m + n)
    ldc = m + 9
    Acm = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))).astype(np.complex64)
    Ccm = (rng.standard_normal((n, ldc)) + 1j * rng.standard_normal((n, ldc))).astype(np.complex64)
    A, C0 = Acm.T, Ccm[:, :m].T
    dA, dC = DeviceArray(Acm), DeviceArray(Ccm)
    assert L.bq_cgemm_split_store_dev(None, m, n, m, dA.ptr, m, dC.ptr, ldc, dC.ptr, ldc) == 0
    out = dC.get()
    got = out[:, :m].T.astype(complex)
    D = _bf16x3_product(A.astype(complex), C0.astype(complex))
    err = np.abs(got + D).max() / np.abs(D).max()
    print(m, n, "in place vs -bf16x3 sum", err)
    assert np.array_equal(out[:, m:], Ccm[:, m:])
    assert err < 2e-6
