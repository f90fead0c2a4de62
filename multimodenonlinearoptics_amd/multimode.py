"""Gaussian-state quantities of Green's functions on the device (hbar = 2,
x = a^dagger + a, p = i (a^dagger - a)), with the names and conventions of the
reference's multimode helpers (multimode.py:15-36):

  calcQuadratureGreens(greenC, greenS) -> Z = [[Re(C+S), -Im(C-S)], [Im(C+S), Re(C-S)]]
  calcCovarianceMtx(Z, tol=1e-4)       -> Z Z^T, checking |det - 1| < tol

The products run on the FP64 matrix cores (nlm_gram); the determinant check is
done on the host (log-determinant) as in the reference.
"""
import numpy as np

from . import capi


def calcQuadratureGreens(greenC, greenS):
    C = np.ascontiguousarray(greenC, dtype=np.complex128)
    S = np.ascontiguousarray(greenS, dtype=np.complex128)
    if C.ndim != 2 or C.shape[0] != C.shape[1] or S.shape != C.shape:
        raise ValueError("greenC and greenS must be square matrices of the same shape")
    n = C.shape[0]
    Z = np.empty((2 * n, 2 * n), dtype=np.float64)
    capi.check(capi.lib().nlm_gaussian_covariance(n, C.ctypes.data, S.ctypes.data, Z.ctypes.data, None))
    return Z


def calcCovarianceMtx(Z, tol=1e-4):
    A = np.ascontiguousarray(Z, dtype=np.float64)
    if A.ndim != 2:
        raise ValueError("Z must be a matrix")
    m, k = A.shape
    cov = np.empty((m, m), dtype=np.float64)
    capi.check(capi.lib().nlm_gram(m, k, A.ctypes.data, cov.ctypes.data))
    sign, logdet = np.linalg.slogdet(cov)
    determinant = sign * np.exp(logdet)
    assert abs(determinant - 1) < tol, "det(C) = %f" % determinant
    return cov
