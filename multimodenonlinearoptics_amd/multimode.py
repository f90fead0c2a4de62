"""Gaussian-state quantities of Green's functions on the device (hbar = 2,
x = a^dagger + a, p = i (a^dagger - a)), with the names and conventions of the
reference's multimode helpers (multimode.py:15-36):

  calcQuadratureGreens(greenC, greenS) -> Z = [[Re(C+S), -Im(C-S)], [Im(C+S), Re(C-S)]]
  calcCovarianceMtx(Z, tol=1e-4)       -> Z Z^T, checking |det - 1| < tol
  blochMessiahEigs(Z)                  -> principal values of sqrtm(Z Z^T), ascending
  blochMessiahVecs(Z)                  -> (values, left vectors, right vectors)
  combineGreens(C1, S1, C2, S2)        -> Green's functions of two media in sequence
  ftGreens(C, S, toTime=True)          -> frequency <-> time domain Green's functions

The matrix products run on the FP64 matrix cores (nlm_gram, nlm_gemm_nt,
nlm_greens_compose) and the transforms on the device FFT; the determinant and
the symmetric eigensolver run on the host (LAPACK), as in the reference.
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


def _gemm_nt(A, B):
    """A @ B.T of real matrices on the FP64 matrix cores."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    if A.shape[1] != B.shape[1]:
        raise ValueError("inner dimensions differ")
    out = np.empty((A.shape[0], B.shape[0]), dtype=np.float64)
    capi.check(capi.lib().nlm_gemm_nt(A.shape[0], B.shape[0], A.shape[1], A.ctypes.data, B.ctypes.data,
                                      out.ctypes.data))
    return out


def _gram_eigh(Z):
    """Eigen-decomposition of Z Z^T (Gram matrix on the device, symmetric
    eigensolver on the host): returns the square roots of the eigenvalues
    (the principal values of the polar factor sqrtm(Z Z^T)) in ascending order
    and the orthonormal eigenvectors as columns."""
    lam, V = np.linalg.eigh(_gemm_nt(Z, Z))
    return np.sqrt(np.clip(lam, 0.0, None)), V


def blochMessiahEigs(Z):
    """Bloch-Messiah principal values (supermode uncertainties) of a quadrature
    matrix, ascending: the eigenvalues of sqrtm(Z Z^T) (multimode.py:140-148 of
    the reference)."""
    return _gram_eigh(np.asarray(Z, dtype=np.float64))[0]


def blochMessiahVecs(Z):
    """Bloch-Messiah decomposition from the eigenvectors of the polar factor,
    without canonical symplectic ordering (multimode.py:150-171 of the
    reference): returns (d, L, R) with d the principal values ordered
    s_1 >= ... >= s_N followed by the N smallest ascending, L the matching
    eigenvectors as columns and R the matching rows of L^T sqrtm(Z Z^T)^-1 Z, so
    that Z = L diag(d) R.  Degenerate principal values give an arbitrary basis
    of their eigenspace; bloch_messiah (decompositions module) resolves them."""
    Z = np.asarray(Z, dtype=np.float64)
    s, V = _gram_eigh(Z)
    N = Z.shape[0] // 2
    order = np.argsort(s)
    idx = np.concatenate([order[-1:-N - 1:-1], order[:N]])
    right = _gemm_nt(np.ascontiguousarray(V.T), np.ascontiguousarray(Z.T)) / s[:, None]
    return s[idx], V[:, idx], right[idx]


def combineGreens(Cfirst, Sfirst, Csecond, Ssecond):
    """Green's functions of two media in sequence (multimode.py:199-205 of the
    reference): C = C2 C1 + S2 conj(S1), S = C2 S1 + S2 conj(C1), on the FP64
    matrix cores (the composition of Cascade.computeGreensFunction)."""
    C1, S1, C2, S2 = [np.ascontiguousarray(a, dtype=np.complex128) for a in (Cfirst, Sfirst, Csecond, Ssecond)]
    if C1.shape != S1.shape or C2.shape != S2.shape or C2.shape[1] != C1.shape[0]:
        raise ValueError("incompatible Green's function shapes")
    M, K, Nc = C2.shape[0], C2.shape[1], C1.shape[1]
    C = np.empty((M, Nc), dtype=np.complex128)
    S = np.empty((M, Nc), dtype=np.complex128)
    capi.check(capi.lib().nlm_greens_compose(M, K, Nc, C2.ctypes.data, S2.ctypes.data, C1.ctypes.data,
                                             S1.ctypes.data, C.ctypes.data, S.ctypes.data))
    return C, S


def ftGreens(greenC, greenS, toTime=True):
    """Green's functions between the frequency and the time domain
    (multimode.py:173-186 of the reference), matrices in the centred (fftshift)
    order of computeGreensFunction.  With F the unscaled DFT matrix,
    a_t = F^-1 a_w gives C_t = F^-1 C_w F and S_t = F^-1 S_w conj(F) = N F^-1 S_w F^-1;
    the row transforms run on the device."""
    C = np.fft.ifftshift(np.asarray(greenC, dtype=np.complex128))
    S = np.fft.ifftshift(np.asarray(greenS, dtype=np.complex128))
    n = S.shape[0]

    def cols(x, inverse):  # transform along axis 0
        return capi.fft(np.ascontiguousarray(x.T), inverse=inverse).T

    if toTime:
        Cc = capi.fft(np.ascontiguousarray(cols(C, True)), inverse=False)
        Sc = capi.fft(np.ascontiguousarray(cols(S, True)), inverse=True) * n
    else:
        Cc = capi.fft(np.ascontiguousarray(cols(C, False)), inverse=True)
        Sc = capi.fft(np.ascontiguousarray(cols(S, False)), inverse=False) / n
    return np.fft.fftshift(Cc), np.fft.fftshift(Sc)
