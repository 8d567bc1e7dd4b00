"""CPU check of the algebra the blocked BSE kernel implements (no GPU).

The kernel inverts M in place by column-pivoted Gauss-Jordan elimination with
implicit pivoting, one panel of b rows at a time:
  * panel factorisation (rows R, serial): for each row r of R pick the pivot
    column p = argmax |Re| + |Im| over unpivoted columns of the current row r,
    then   a[i][c] -= a[i][p] a[r][c] / piv,  a[i][p] /= piv   (i in R, i != r)
           a[r][c] = -a[r][c] / piv,  a[r][p] = 1 / piv;
  * block update of every other row S:  M[S, :] += M[S, P] W,  W = M'[R, :] - E_P,
    with E_P the unit rows e_{p_s} and M[S, P] taken before the update.
After all rows, a[r][c] = inv(M)[p_r][row_of(c)], row_of(c) the row whose pivot
column is c.  This restates that algorithm in numpy and checks it against
numpy.linalg.inv on random and on BSE-shaped matrices, for any row order.
"""
import numpy as np
import pytest


def blocked_cgj(M, b, order):
    a = M.astype(complex).copy()
    N = a.shape[0]
    pivoted = np.zeros(N, bool)
    pcol = np.empty(N, int)
    for k0 in range(0, N, b):
        R = order[k0:k0 + b]
        panel = a[R].copy()
        P = []
        for s, r in enumerate(R):
            row = panel[s]
            cand = np.where(pivoted, -1.0, np.abs(row.real) + np.abs(row.imag))
            p = int(np.argmax(cand))
            pivoted[p] = True
            pcol[r] = p
            inv = 1.0 / row[p]
            rowr, colp = row.copy(), panel[:, p].copy()
            for t in range(len(R)):
                if t != s:
                    panel[t] -= colp[t] * rowr * inv
                    panel[t, p] = colp[t] * inv
            panel[s] = -rowr * inv
            panel[s, p] = inv
            P.append(p)
        W = panel.copy()
        W[np.arange(len(P)), P] -= 1.0
        S = np.setdiff1d(np.arange(N), R)
        a[S] += a[np.ix_(S, P)] @ W
        a[R] = panel
    return a, pcol


def unpermute(a, pcol):
    """inv(M) from the kernel's storage: inv[p_r][row_of(c)] = a[r][c]."""
    N = a.shape[0]
    rowof = np.empty(N, int)
    rowof[pcol] = np.arange(N)
    X = np.empty_like(a)
    X[np.ix_(pcol, rowof)] = a
    return X


@pytest.mark.parametrize("N,b", [(12, 4), (16, 4), (13, 4), (20, 1), (24, 8)])
@pytest.mark.parametrize("perm", [False, True])
def test_blocked_cgj_inverts(N, b, perm):
    rng = np.random.default_rng(N * 10 + b)
    M = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    order = list(rng.permutation(N)) if perm else list(range(N))
    a, pcol = blocked_cgj(M, b, order)
    X = unpermute(a, pcol)
    assert np.allclose(X, np.linalg.inv(M), rtol=0, atol=1e-11 * np.abs(np.linalg.inv(M)).max())
    assert sorted(pcol) == list(range(N))


def test_blocked_cgj_bse_solution():
    """y = inv(A^T) 1 from the stored inverse of M = A^T (row sums of inv(M))."""
    rng = np.random.default_rng(3)
    N = 24
    chi0 = (rng.uniform(0.05, 1.0, N) + 1j * rng.uniform(-0.3, 0.3, N))
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    Gam = -1.5 + 0.15 * (X + X.T) / 2 / (1 + np.abs(np.subtract.outer(np.arange(N), np.arange(N))))
    A = Gam + np.diag(1 / chi0)
    a, pcol = blocked_cgj(A.T, 4, list(range(N)))
    y = unpermute(a, pcol).sum(axis=1)
    assert np.allclose(A.T @ y, np.ones(N), atol=1e-12)
    assert np.allclose(y, np.linalg.inv(A).sum(axis=0), atol=1e-12)
