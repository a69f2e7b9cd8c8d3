"""Pins the partial-pivoting oracle (oracle_gtsv_*: LAPACK ?gtsv restated).

The reference's pivoting baseline calls MKL ?gtsv per system
(tridiagonalMKL.cpp:5-26). MKL is not in this image; scipy's LAPACK ?gtsv, an
independent build of the same published routine, is. The oracle performs
LAPACK's operations in LAPACK's order without contraction, so it is compared
with scipy bit for bit -- x and INFO -- on both non-dominant value kinds, on
small integer systems full of ties and zero pivots, and on planted mistakes
that this comparison must catch. Dense solves and the conditioning of the
"paired" kind pin what the comparison alone cannot (that the answer is the
solution, and that the kind really needs the interchanges).
"""
import numpy as np
import pytest
from scipy.linalg import lapack


def lapack_gtsv(a, b, c, d, n):
    """scipy's ?gtsv per system; returns (x, info) like oracle.gtsv."""
    f = lapack.sgtsv if a.dtype == np.float32 else lapack.dgtsv
    N = a.shape[0] // n
    x = np.empty_like(a)
    info = np.empty(N, dtype=np.int32)
    for s in range(N):
        sl = slice(s * n, (s + 1) * n)
        if n == 1:
            info[s] = 1 if b[sl][0] == 0 else 0
            x[sl] = d[sl] / b[sl] if info[s] == 0 else np.nan
            continue
        _, _, _, xs, inf = f(a[sl][1:].copy(), b[sl].copy(), c[sl][:-1].copy(),
                             d[sl].reshape(-1, 1).copy())
        info[s] = inf
        x[sl] = xs[:, 0] if inf == 0 else np.nan
    return x, info


def same_bits(x, y):
    """Equal element for element, NaN matching NaN."""
    both_nan = np.isnan(x) & np.isnan(y)
    return bool(np.all(both_nan | (x == y)))


@pytest.mark.parametrize("kind", ["paired", "general"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gtsv_bitwise_equals_lapack(oracle, kind, dtype):
    for n in (1, 2, 3, 4, 5, 16, 64, 333, 1024):
        a, b, c, d = oracle.generate(kind, 21, n, 24, dtype)
        x, info = oracle.gtsv(a, b, c, d, n)
        xl, infol = lapack_gtsv(a, b, c, d, n)
        assert np.array_equal(info, infol), n
        assert same_bits(x, xl), n


def integer_systems(seed, n, N, dtype):
    """Entries in {-1, 0, 1}: ties |D| == |l|, zero pivots and singular
    systems are common."""
    rng = np.random.default_rng(seed)
    a, b, c, d = (rng.integers(-1, 2, size=n * N).astype(dtype) for _ in range(4))
    d[d == 0] = 1
    return a, b, c, d


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gtsv_ties_and_singular_info_match_lapack(oracle, dtype):
    seen_singular = seen_solved = 0
    for n in (1, 2, 3, 4, 6):
        a, b, c, d = integer_systems(n, n, 600, dtype)
        x, info = oracle.gtsv(a, b, c, d, n)
        xl, infol = lapack_gtsv(a, b, c, d, n)
        assert np.array_equal(info, infol), n
        assert same_bits(x, xl), n
        assert np.all(np.isnan(x.reshape(-1, n)[info != 0]))
        seen_singular += int((info != 0).sum())
        seen_solved += int((info == 0).sum())
    assert seen_singular > 100 and seen_solved > 100


def test_gtsv_known_singular_cases(oracle):
    # zero pivot with nothing below it in row 0: info = 1
    a = np.array([0.0, 0.0, 1.0]); b = np.array([0.0, 1.0, 1.0])
    c = np.array([1.0, 1.0, 0.0]); d = np.ones(3)
    assert oracle.gtsv(a, b, c, d, 3)[1][0] == 1
    # [[1, 1], [1, 1]]: no interchange (tie), then D_1 = 0: info = n = 2
    x, info = oracle.gtsv(np.array([0.0, 1.0]), np.array([1.0, 1.0]),
                          np.array([1.0, 0.0]), np.array([1.0, 2.0]), 2)
    assert info[0] == 2 and np.all(np.isnan(x))


@pytest.mark.parametrize("n", [2, 5, 32, 64])
def test_gtsv_dense_solve(oracle, n):
    for kind in ("paired", "general"):
        a, b, c, d = oracle.generate(kind, 5, n, 16, np.float64)
        x, info = oracle.gtsv(a, b, c, d, n)
        assert np.all(info == 0)
        for s in range(16):
            sl = slice(s * n, (s + 1) * n)
            T = np.diag(b[sl]) + np.diag(a[sl][1:], -1) + np.diag(c[sl][:-1], 1)
            ref = np.linalg.solve(T, d[sl])
            err = np.max(np.abs(x[sl] - ref)) / np.max(np.abs(ref))
            tol = 1e-13 if kind == "paired" else 1e-15 * np.linalg.cond(T) * n
            assert err <= max(tol, 1e-13), (kind, s, err)


def test_paired_kind_needs_interchanges(oracle):
    n, N = 1024, 16
    a, b, c, d = oracle.generate("paired", 3, n, N, np.float64)
    for s in range(4):
        sl = slice(s * n, (s + 1) * n)
        T = np.diag(b[sl]) + np.diag(a[sl][1:], -1) + np.diag(c[sl][:-1], 1)
        assert np.linalg.cond(T) < 10
    ref = oracle.thomas(a, b, c, d, n=n)  # no pivoting, FP64
    x, info = oracle.gtsv(a, b, c, d, n)
    resid_g = oracle.residual(a, b, c, d, x, n=n).max()
    resid_t = oracle.residual(a, b, c, d, ref, n=n).max()
    assert resid_g < 1e-15 and resid_t > 1e3 * resid_g
    a32, b32, c32, d32 = oracle.generate("paired", 3, n, N, np.float32)
    x32, _ = oracle.gtsv(a32, b32, c32, d32, n)
    t32 = oracle.thomas_f32(a32, b32, c32, d32, n)
    x64, _ = oracle.gtsv(*(v.astype(np.float64) for v in (a32, b32, c32, d32)), n)
    scale = np.abs(x64).max()
    assert np.abs(x32 - x64).max() / scale < 1e-5
    assert np.abs(t32 - x64).max() / scale > 1e-3


# --------------------------------------------------------------------------
# planted mistakes: a numpy restatement with one slip each must fail the
# bitwise LAPACK comparison that pins the C oracle
# --------------------------------------------------------------------------

def _gtsv_variant(a, b, c, d, n, mistake):
    A, B, C, X = (np.asarray(v, dtype=np.float64).reshape(-1, n).copy() for v in (a, b, c, d))
    N = A.shape[0]
    D = B.copy()
    U = np.zeros((N, n)); U[:, :-1] = C[:, :-1]
    W = np.zeros((N, n)); W[:, :-1] = A[:, 1:]
    for i in range(n - 1):
        _variant_step(A, D, U, W, X, i, n, N, mistake)
    x = np.empty((N, n))
    with np.errstate(all="ignore"):
        x[:, -1] = X[:, -1] / D[:, -1]
        if n > 1:
            x[:, -2] = (X[:, -2] - U[:, -2] * x[:, -1]) / D[:, -2]
        for i in range(n - 3, -1, -1):
            x[:, i] = (X[:, i] - U[:, i] * x[:, i + 1] - W[:, i] * x[:, i + 2]) / D[:, i]
    return x.reshape(-1)


@np.errstate(all="ignore")
def _variant_step(A, D, U, W, X, i, n, N, mistake):
    """Elimination step i of the restatement, with the planted slip."""
    if mistake == "never_swap":
        swap = np.zeros(N, dtype=bool)
    elif mistake == "gt_not_ge":
        swap = ~(np.abs(D[:, i]) > np.abs(W[:, i]))
    else:
        swap = ~(np.abs(D[:, i]) >= np.abs(W[:, i]))
    ns = ~swap
    f = np.where(swap, D[:, i] / W[:, i], W[:, i] / D[:, i])
    Dn, Un, Xi, Xn = D[:, i + 1].copy(), U[:, i + 1].copy(), X[:, i].copy(), X[:, i + 1].copy()
    # no interchange
    D[ns, i + 1] = Dn[ns] - f[ns] * U[ns, i]
    X[ns, i + 1] = Xn[ns] - f[ns] * Xi[ns]
    W[ns, i] = 0
    # interchange
    D[swap, i] = W[swap, i]
    D[swap, i + 1] = U[swap, i] - f[swap] * Dn[swap]
    if i < n - 2:
        W[swap, i] = 0 if mistake == "drop_fill" else Un[swap]
        U[swap, i + 1] = -f[swap] * Un[swap]
    else:
        W[swap, i] = 0
    U[swap, i] = Dn[swap]
    X[swap, i] = Xn[swap]
    X[swap, i + 1] = Xi[swap] - f[swap] * Xn[swap]
    if mistake == "rhs_unswapped":
        X[swap, i] = Xi[swap]
        X[swap, i + 1] = Xn[swap] - f[swap] * Xi[swap]


@pytest.mark.parametrize("mistake", ["none", "never_swap", "gt_not_ge", "drop_fill",
                                     "rhs_unswapped"])
def test_gtsv_planted_mistakes_are_caught(mistake):
    caught = False
    cases = [integer_systems(11, 6, 400, np.float64)]
    from oracle import oracle as O
    cases.append(O.generate("paired", 4, 6, 400, np.float64))
    for a, b, c, d in cases:
        xl, info = lapack_gtsv(a, b, c, d, 6)
        ok = np.repeat(info == 0, 6)
        xv = _gtsv_variant(a, b, c, d, 6, mistake)
        if not same_bits(xv[ok], xl[ok]):
            caught = True
    assert caught == (mistake != "none")


@pytest.mark.parametrize("kind", ["paired", "general"])
def test_generator_ragged_pairs_follow_system_index(oracle, kind):
    offsets = np.array([0, 5, 11, 111, 114], dtype=np.int64)
    c_arrs = oracle.generate(kind, 3, 0, 0, np.float64, offsets=offsets)
    np_arrs = oracle.generate_np(kind, 3, 0, 0, np.float64, offsets=offsets)
    for u, v in zip(c_arrs, np_arrs):
        assert np.array_equal(u, v)
    if kind == "paired":
        a, b, c, d = c_arrs
        for s in range(len(offsets) - 1):
            o, e = offsets[s], offsets[s + 1]
            assert a[o] == 0 and c[e - 1] == 0
            assert np.all(a[o + 1:e:2] >= 1.0)   # odd rows: strong a
            assert np.all(c[o:e - 1:2] >= 1.0)   # even rows: strong c
            inner = b[o:e] if (e - o) % 2 == 0 else b[o:e - 1]
            assert np.all(np.abs(inner) <= 0.001)
