"""The CPU oracle of CG2D pinned by what the mathematics fixes (CPU only): the two sum orders against independent
formulations, the solution against a scipy sparse direct solve of the same assembled operator, symmetry, the
stopping rule and special cases of the reference's tests (test_cg2d.py:170-237), a textbook Jacobi-PCG, and
planted errors that must change the result."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from tests.conftest import maps

T, NY, NX, OL, N = 13, 98, 98, 4, 90


def _halving(v):
    """The tree sum as repeated halving: fold to 64 values, then v[:h] + v[h:] (lane 0 of a butterfly)."""
    v = np.asarray(v, np.float64)
    w = v[:64].copy()
    w[:len(v) - 64] = w[:len(v) - 64] + v[64:]
    while len(w) > 1:
        h = len(w) // 2
        w = w[:h] + w[h:]
    return w[0]


def test_tree_sum_is_the_halving_tree():
    rng = np.random.default_rng(0)
    for n in (64, 65, 90, 100, 128):
        for _ in range(5):
            v = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 8, size=n)
            assert oracle.tree_n(v) == _halving(v), n


def test_sum_orders():
    """'fortran' = strictly sequential per tile (np.cumsum is sequential) then 0 + tile1 + ... in order; 'tile' =
    the tree per row then per tile. Integer-valued data sums exactly in both; adversarial data separates them."""
    rng = np.random.default_rng(1)
    a = rng.normal(size=(T, N, N)) * 10.0 ** rng.uniform(-6, 6, size=(T, N, N))
    seq = 0.0
    for t in range(T):
        seq = seq + np.cumsum(np.concatenate([[0.0], a[t].reshape(-1)]))[-1]
    assert oracle.global_sum(a, "fortran") == seq
    tree = 0.0
    for t in range(T):
        tree = tree + _halving([_halving(row) for row in a[t]])
    assert oracle.global_sum(a, "tile") == tree
    ints = rng.integers(-1000, 1000, size=(T, N, N)).astype(np.float64)
    assert oracle.global_sum(ints, "fortran") == oracle.global_sum(ints, "tile") == ints.sum()
    assert oracle.global_sum(a, "fortran") != oracle.global_sum(a, "tile")


def _assemble(ops, norm):
    """(A / cg2dNorm) as a scipy matrix over the interior unknowns, halo couplings resolved through the T map."""
    src, comp, sign = maps().full("T")
    inner = np.zeros((T, NY, NX), bool)
    inner[:, OL:OL + N, OL:OL + N] = True
    idx = -np.ones(T * NY * NX, np.int64)
    idx[inner.reshape(-1)] = np.arange(inner.sum())
    # column of a stencil point: itself if interior, its exchange source if written, none if unwritten
    col = np.where(inner.reshape(-1), idx, np.where(comp > 0, idx[src], -1))
    aW, aS, aC = (o.reshape(-1) for o in ops[:3])
    p = np.nonzero(inner.reshape(-1))[0]
    rows, cols, vals = [], [], []
    for coef, nb in ((aW[p], p - 1), (aW[p + 1], p + 1), (aS[p], p - NX), (aS[p + NX], p + NX), (aC[p], p)):
        c = col[nb]
        k = (c >= 0) & (coef != 0)
        rows.append(idx[p][k])
        cols.append(c[k])
        vals.append(coef[k] / norm)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(p), len(p)))
    return A, inner


@pytest.fixture(scope="module")
def case():
    from mit_jax_amd import synthetic as S
    g, ops, norm = S.cg2d_case(0, "5point")
    return g, ops, norm, S.rhs(g, 1)


@pytest.mark.parametrize("order", ["fortran", "tile"])
def test_cg2d_matches_direct_solve(case, order):
    g, ops, norm, b = case
    x, d = oracle.cg2d(maps().full("T"), ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=1e-12, max_iters=2000,
                       sum_order=order)
    A, inner = _assemble(ops, norm)
    xs = spl.spsolve(A.tocsc(), b[inner])
    assert d["lastResidual"] < 1e-12
    assert np.abs(x[inner] - xs).max() <= 1e-10 * np.abs(xs).max()
    # the returned halos: the exchanged normalised first guess (0 here) where written, the first guess elsewhere
    assert np.all(x[~inner] == 0.0)


def test_operator_symmetric_and_matvec(case):
    """<A u, v> = <u, A v> over the exchange-coupled operator (reference test_cg2d.py:214-226), and the oracle's
    _matvec equals the assembled matrix product."""
    g, ops, norm, _ = case
    A, inner = _assemble(ops, norm)
    assert abs(A - A.T).max() == 0.0
    rng = np.random.default_rng(1)
    u = np.where(inner, rng.normal(size=inner.shape), 0.0)
    v = np.where(inner, rng.normal(size=inner.shape), 0.0)
    Au = oracle.cg2d_matvec(maps().full("T"), ops[:3], u, norm)
    Av = oracle.cg2d_matvec(maps().full("T"), ops[:3], v, norm)
    assert abs((Au * v).sum() - (u * Av).sum()) <= 1e-12 * abs((Au * v).sum())
    np.testing.assert_allclose(Au[inner], A @ u[inner], rtol=1e-13, atol=1e-13 * np.abs(Au).max())
    assert np.all(Au[~inner] == 0.0)


def test_zero_rhs_and_stopping_rule(case):
    """b = 0, x0 = 0: no iteration, x = 0 (cg2d.py:224-225, :241). A target just below the residual a solve
    stopped at needs exactly one more iteration (reference test_cg2d.py:187-189)."""
    g, ops, norm, b = case
    Tm = maps().full("T")
    x, d = oracle.cg2d(Tm, ops, np.zeros_like(b), np.zeros_like(b), cg2dNorm=norm, tolerance=1e-12, max_iters=100)
    assert d["numIters"] == 0 and np.all(x == 0.0) and d["rhsMax"] == 0.0
    x, d = oracle.cg2d(Tm, ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=1e-7, max_iters=1000)
    _, d2 = oracle.cg2d(Tm, ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=0.99 * d["lastResidual"],
                        max_iters=1000)
    assert d2["numIters"] == d["numIters"] + 1


def test_jacobi_pcg_textbook():
    """pW = pS = 0 makes the solver Jacobi-preconditioned CG: an independent numpy implementation of the textbook
    algorithm on the assembled matrix reaches the same solution and (within one) the same iteration count."""
    from mit_jax_amd import synthetic as S
    g, ops, norm = S.cg2d_case(0, "diagonal")
    b = S.rhs(g, 1)
    x, d = oracle.cg2d(maps().full("T"), ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=1e-12, max_iters=2000)
    A, inner = _assemble(ops, norm)
    An = A * norm                               # the normalised operator cg2d.F iterates with
    bb = b[inner] * norm
    rhsNorm = 1.0 / np.abs(bb).max()
    bb = bb * rhsNorm
    Dinv = 1.0 / An.diagonal()
    xx = np.zeros_like(bb)
    r = bb - An @ xx
    z = Dinv * r
    p = z.copy()
    rz = r @ z
    it = 0
    while not (r @ r < 1e-24) and it < 2000:
        q = An @ p
        a = rz / (p @ q)
        xx += a * p
        r -= a * q
        z = Dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    assert abs(it - d["numIters"]) <= 1
    assert np.abs(x[inner] - xx / rhsNorm).max() <= 1e-9 * np.abs(x[inner]).max()


def test_planted_errors_change_the_solution(case):
    """cg2dNorm x (1 + 1e-6) and one preconditioner value x (1 + 1e-6) each change x (reference
    test_cg2d.py:170-184), so an oracle with either mistake would fail the GPU parity gates."""
    g, ops, norm, b = case
    Tm = maps().full("T")
    ref, _ = oracle.cg2d(Tm, ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=1e-9, max_iters=1000)
    x, _ = oracle.cg2d(Tm, ops, b, np.zeros_like(b), cg2dNorm=norm * (1 + 1e-6), tolerance=1e-9, max_iters=1000)
    assert not np.array_equal(x, ref)
    bad = list(ops)
    bad[3] = ops[3].copy()
    k = np.flatnonzero(ops[3])[len(np.flatnonzero(ops[3])) // 2]  # a wet point with a nonzero pW
    bad[3].reshape(-1)[k] *= 1 + 1e-6
    x, _ = oracle.cg2d(Tm, bad, b, np.zeros_like(b), cg2dNorm=norm, tolerance=1e-9, max_iters=1000)
    assert not np.array_equal(x, ref)


def test_coeff_grad_is_the_transpose_of_the_coefficient_matvec(case):
    """The oracle's coefficient cotangent is the adjoint of dA -> -_matvec(dA, x) (core/cg2d.py:326-334): for random
    directions dA (each of aW, aS, aC separately) <g_k, dA_k> = -<lam, _matvec(dA, x)>, where the right-hand side
    is the oracle's C matvec (exchange + 5-point stencil), an independent path. _matvec is linear in dA, so the two
    agree to rounding. A shifted neighbour (x(i+1) for x(i-1)) or a dropped term fails by O(1)."""
    g, ops, norm, b = case
    Tm = maps().full("T")
    _, inner = _assemble(ops, norm)
    rng = np.random.default_rng(5)
    x = rng.normal(size=b.shape)  # halos arbitrary: the cotangent must only see exch(where(inner, x, 0))
    lam = np.where(inner, rng.normal(size=b.shape), 0.0)
    gW, gS, gC = oracle.cg2d_coeff_grad(Tm, x, lam, norm)
    zero = np.zeros_like(b)
    for k, gk in enumerate((gW, gS, gC)):
        dA = rng.normal(size=b.shape)
        A = [zero, zero, zero]
        A[k] = dA
        rhs = -(lam * oracle.cg2d_matvec(Tm, A, x, norm)).sum()
        lhs = (gk * dA).sum()
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs)), (k, lhs, rhs)
    # entries no interior stencil reads carry no cotangent
    ny, nx = inner.shape[1:]
    read = np.zeros((ny, nx), bool)
    read[4:94, 4:95] = True  # aW at i and i+1 of interior points
    assert np.all(gW[:, ~read] == 0.0)
