"""Chi3GNLSE batch propagation restated with torch (CPU, complex128).

Used as the reverse-mode reference for the device adjoint: torch's autograd
differentiates this forward without any of the adjoint's hand-derived
formulas.  The forward follows Chi3GNLSE::DiffEq (solvers/Chi3GNLSE.cpp:39-92)
inside signalSimulationTemplate (src/_NonlinearMedium.hpp:160-239): half
dispersion step in, then per z step an RK4 step of
f(y) = IFFT((nl0 + nl1 omega) FFT(y IFFT(R FFT(r |y|^2)))) and a dispersion
step, half step out.  Its constants come from the medium (omega, the Raman
response) and the constructor arguments; test_gnlse_torch_reference checks it
against the oracle's batchSignalSimulation.
"""
import numpy as np
import torch


class GnlseTables:
    def __init__(self, kw, omega, raman):
        nz = kw["nZSteps"]
        dz = kw["relativeLength"] / (nz - 1)
        w = torch.as_tensor(np.asarray(omega), dtype=torch.float64)
        beta1, beta2, beta3 = kw.get("beta1", 0.0), kw["beta2"], kw.get("beta3", 0.0)
        D = w * (beta1 + w * (0.5 * beta2 + w * beta3 / 6))
        self.nz = nz
        self.disp = torch.exp(1j * dz * D)
        self.half_in = torch.exp(1j * (0.5 * dz) * D)
        self.half_out = torch.exp(-1j * (0.5 * dz) * D)
        self.filt = (1j * dz / kw["nlLength"]) + (1j * dz / kw["selfSteepLength"]) * w
        self.raman = torch.as_tensor(np.asarray(raman), dtype=torch.complex128)


def _f(y, t):
    s = (y.real ** 2 + y.imag ** 2).to(torch.complex128)
    v = torch.fft.ifft(torch.fft.fft(s) * t.raman)
    return torch.fft.ifft(torch.fft.fft(y * v) * t.filt)


def forward(X, t, in_time):
    """Output rows (domain of the input) for input rows X [B, N] (torch complex128)."""
    F = torch.fft.fft(X) if in_time else X
    a = torch.fft.ifft(F * t.half_in)
    for i in range(1, t.nz):
        k1 = _f(a, t)
        k2 = _f(a + 0.5 * k1, t)
        k3 = _f(a + 0.5 * k2, t)
        k4 = _f(a + k3, t)
        a = a + (((k1 + 2 * k2) + 2 * k3) + k4) / 6
        F = torch.fft.fft(a) * t.disp
        if i < t.nz - 1:
            a = torch.fft.ifft(F)
    F = F * t.half_out
    return torch.fft.ifft(F) if in_time else F


def loss_weights(rows, n, seed):
    r = np.random.default_rng(seed)
    w = r.uniform(0.0, 1.0, (rows, n))
    c = r.standard_normal((rows, n)) + 1j * r.standard_normal((rows, n))
    return w, c


def loss(Y, w, c):
    """sum w |Y|^2 + Re sum c Y, a real loss that weighs every output bin."""
    w = torch.as_tensor(w)
    c = torch.as_tensor(c)
    return (w * (Y.real ** 2 + Y.imag ** 2)).sum() + (c * Y).real.sum()


def grad_out(Y, w, c):
    """dL/dY in the convention dL/dRe + i dL/dIm: 2 w Y + conj(c)."""
    return 2 * w * Y + np.conj(c)


def reference_gradient(X, t, in_time, w, c):
    """(Y, dL/dX) with torch's autograd (dL/dRe + i dL/dIm)."""
    x = torch.as_tensor(np.asarray(X, dtype=np.complex128)).clone().requires_grad_(True)
    Y = forward(x, t, in_time)
    L = loss(Y, w, c)
    L.backward()
    return Y.detach().numpy(), x.grad.numpy()
