"""The torch restatement of the Chi3GNLSE batch (tests/gnlse_torch.py), which
supplies the exact reverse-mode reference of the device adjoint, against the
oracle: its forward must equal the oracle's batchSignalSimulation, and its
autograd gradient must match central finite differences of the oracle."""
import math

import numpy as np
import pytest

import gnlse_torch as GT
from conftest import rel_l2


def small_kwargs(N, nZ, length=0.02):
    return dict(relativeLength=length, nlLength=1 / 9, selfSteepLength=100.0, fr=0.18, fb=0.21, tau1=0.0122,
                tau2=0.032, tau3=0.096, beta2=-1.0, tMax=10.0, tPrecision=N, nZSteps=nZ)


def sech_inputs(N, rows, seed, tMax):
    r = np.random.default_rng(seed)
    tau = 2 * tMax / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    noise = (r.standard_normal((rows, N)) + 1j * r.standard_normal((rows, N))) / math.sqrt(2 * N)
    return (1 / np.cosh(tau))[None, :] + np.fft.ifft(noise * N, axis=1) / N


@pytest.mark.parametrize("in_time", [True, False])
def test_torch_forward_matches_oracle(oracle, in_time):
    N, nZ, rows = 128, 12, 2
    kw = small_kwargs(N, nZ)
    ref = oracle.Chi3GNLSE(**kw)
    X = sech_inputs(N, rows, 7, kw["tMax"])
    if not in_time:
        X = np.fft.fft(X, axis=1)
    t = GT.GnlseTables(kw, ref.omega, ref.ramanResponse)
    import torch
    got = GT.forward(torch.as_tensor(X), t, in_time).numpy()
    exp = ref.batchSignalSimulation(X, inTimeDomain=in_time)
    assert rel_l2(got, exp) < 1e-12


def test_torch_gradient_matches_oracle_fd(oracle):
    N, nZ, rows = 64, 8, 1
    kw = small_kwargs(N, nZ)
    ref = oracle.Chi3GNLSE(**kw)
    X = sech_inputs(N, rows, 11, kw["tMax"])
    w, c = GT.loss_weights(rows, N, 12)
    t = GT.GnlseTables(kw, ref.omega, ref.ramanResponse)
    _, g = GT.reference_gradient(X, t, True, w, c)

    def L(x):
        Y = ref.batchSignalSimulation(x, inTimeDomain=True)
        return float(np.sum(w * np.abs(Y) ** 2) + np.real(np.sum(c * Y)))

    r = np.random.default_rng(13)
    eps = 1e-6
    for _ in range(3):
        d = r.standard_normal((rows, N)) + 1j * r.standard_normal((rows, N))
        d /= np.linalg.norm(d)
        fd = (L(X + eps * d) - L(X - eps * d)) / (2 * eps)
        an = float(np.real(np.vdot(g, d)))
        assert abs(fd - an) <= 1e-6 * np.linalg.norm(g)
