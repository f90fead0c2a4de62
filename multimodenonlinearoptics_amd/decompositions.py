"""Bloch-Messiah decomposition of a real symplectic matrix, with the interface
of the reference's decompositions.py:250-353 (symplectic form
Omega = [[0, I], [-I, 0]]):

  bloch_messiah(S, tol=1e-10, rounding=9) -> (O1, D, O2),  S = O1 D O2

O1 and O2 are orthogonal and symplectic, D = diag(s_1..s_n, 1/s_1..1/s_n) with
s_1 >= ... >= s_n >= 1.  For the quadrature matrix Z of a Green's function
(multimode.calcQuadratureGreens) the s_k are the supermode squeezing factors
e^{r_k} and the columns of O1 the output supermodes.

Method.  The left polar factor P = sqrtm(S S^T) has eigenvalues in pairs
(s, 1/s), and Omega maps the s-eigenspace onto the 1/s-eigenspace.  The
eigenvectors x_k of the n largest eigenvalues (Gram matrix S S^T on the FP64
matrix cores, symmetric eigensolver on the host) together with y_k = -Omega x_k
form O1 = [x | y], which is orthogonal and symplectic by construction
(x_j^T Omega x_k = 0 across eigenspaces).  Eigenvalues within 10^-rounding of 1
(the passive modes) form one eigenspace E that Omega maps onto itself; there
the real Schur form of E^T Omega E pairs the basis into Lagrangian halves, so
degenerate and passive modes need no separate rotation step.  Finally
O2 = D^-1 O1^T S, so that O1 D O2 reproduces S to rounding.  The ordering and
sign conventions are those of the reference: squeezers in decreasing order,
O1^T Omega O1 = Omega; a matrix that is already orthogonal is returned as
(S, I, I).
"""
import numpy as np

from .multimode import _gemm_nt


def sympmat(n):
    """The symplectic form [[0, I_n], [-I_n, 0]] (decompositions.py:17-29 of the reference)."""
    out = np.zeros((2 * n, 2 * n))
    out[:n, n:] = np.eye(n)
    out[n:, :n] = -np.eye(n)
    return out


def _omega_times(X):
    """Omega @ X for the symplectic form of order X.shape[0] // 2."""
    n = X.shape[0] // 2
    return np.concatenate([X[n:], -X[:n]])


def bloch_messiah(S, tol=1e-10, rounding=9):
    S = np.asarray(S)
    if S.ndim != 2 or S.shape[0] != S.shape[1]:
        raise ValueError("The input matrix is not square")
    if S.shape[0] % 2 != 0:
        raise ValueError("The input matrix must have an even number of rows/columns")
    S = np.ascontiguousarray(S.real, dtype=np.float64)
    n = S.shape[0] // 2
    St = np.ascontiguousarray(S.T)

    # S^T Omega S = S^T (Omega S)
    error = np.linalg.norm(_gemm_nt(St, np.ascontiguousarray(_omega_times(S).T)) - sympmat(n)) / n
    if error >= tol:
        raise ValueError("The input matrix is not symplectic (error = %f)" % error)
    if np.linalg.norm(_gemm_nt(St, St) - np.eye(2 * n)) < tol:
        return S.copy(), np.eye(2 * n), np.eye(2 * n)

    lam, V = np.linalg.eigh(_gemm_nt(S, S))
    order = np.argsort(lam)[::-1]
    s = np.sqrt(np.clip(lam[order], 0.0, None))
    V = V[:, order]

    # passive modes: the eigenvalue-1 eigenspace, the 2d middle columns
    d = int(np.count_nonzero(np.abs(s[:n] - 1.0) < 10.0 ** (-rounding)))
    X = np.empty((2 * n, n))
    D = np.empty(2 * n)
    m = n - d
    X[:, :m] = V[:, :m]
    D[:m] = s[:m]
    D[n:n + m] = 1.0 / s[:m]
    if d:
        import scipy.linalg
        E = np.ascontiguousarray(V[:, m:n + d])
        Q = _gemm_nt(np.ascontiguousarray(E.T), np.ascontiguousarray(_omega_times(E).T))
        T, W = scipy.linalg.schur(0.5 * (Q - Q.T), output="real")
        # T is block diagonal with 2x2 blocks [[0, t], [-t, 0]], |t| = 1: the
        # first vector of each block spans a Lagrangian half with the others
        # principal values of the passive modes (Rayleigh quotients of P), decreasing
        pv = (W[:, 0::2] ** 2).T @ s[m:n + d]
        po = np.argsort(pv)[::-1]
        pv = pv[po]
        X[:, m:] = E @ W[:, 0::2][:, po]
        D[m:n] = pv
        D[n + m:] = 1.0 / pv
    # O1 = [[Xt, -Xb], [Xb, Xt]] is orthogonal exactly when U = Xt + i Xb is
    # unitary.  Eigenvectors of principal values 1 + delta with delta near the
    # eigensolver's resolution mix with their partners' (error ~ eps |P|^2 / delta);
    # the nearest unitary to U removes that mixing from O1 while moving each
    # vector only within its near-degenerate cluster.
    w, _, vh = np.linalg.svd(X[:n] + 1j * X[n:])
    U = w @ vh
    X = np.concatenate([U.real, U.imag])
    O1 = np.concatenate([X, -_omega_times(X)], axis=1)
    O2 = _gemm_nt(np.ascontiguousarray(O1.T), St) / D[:, None]
    return O1, np.diag(D), O2
