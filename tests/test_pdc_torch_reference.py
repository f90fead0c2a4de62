"""The torch restatement of the Chi2PDC Green's function (tests/pdc_torch.py),
which supplies the exact directional derivatives for the device adjoint at the
differentiable-poling configuration, against the oracle: its C and S must equal
the oracle's computeGreensFunction, its forward-mode derivative must equal its
reverse-mode gradient, and both must match central differences of the oracle."""
import numpy as np
import pytest
import torch

import pdc_torch as PT
from cases import config4_parts
from conftest import rel_l2


def _oracle_S(O, kw, spec0, phi, pol):
    env = np.fft.ifft(spec0 * np.exp(1j * phi))
    m = O.Chi2PDC(customPump=env, polingProfile=list(pol), **kw)
    return m.computeGreensFunction()


@pytest.mark.parametrize("d0", [0.0, 1.7])
def test_torch_greens_matches_oracle(oracle, d0):
    kw, spec0, phi, pol = config4_parts(64, 40)
    kw = dict(kw, diffBeta0=d0, nlLength=0.8)
    T = PT.PdcGreens(kw, spec0)
    C, S = T.greens(torch.as_tensor(phi), torch.as_tensor(pol))
    Co, So = _oracle_S(oracle, kw, spec0, phi, pol)
    assert rel_l2(C.numpy(), Co) < 1e-12
    assert rel_l2(S.numpy(), So) < 1e-12


def test_forward_mode_equals_reverse_mode_and_oracle_fd(oracle):
    kw, spec0, phi, pol = config4_parts(32, 24)
    kw = dict(kw, nlLength=0.5, diffBeta0=0.9)
    T = PT.PdcGreens(kw, spec0)
    L, g_phi, g_pol = PT.reverse_gradient(T, phi, pol)
    N, nZ = phi.size, pol.size
    I = np.eye(N)

    def L_oracle(ph, pl):
        S = _oracle_S(oracle, kw, spec0, ph, pl)[1]
        return float(np.sum(np.abs(S @ S.T - 1.5 * I) ** 2))

    assert abs(L - L_oracle(phi, pol)) <= 1e-12 * abs(L)
    r = np.random.default_rng(3)
    for _ in range(3):
        d = r.standard_normal(N + nZ)
        d /= np.linalg.norm(d)
        dphi, dpol = d[:N], d[N:]
        L2, dL = PT.loss_and_directional_derivative(T, phi, pol, dphi, dpol)
        rev = float(np.dot(g_phi, dphi) + np.dot(g_pol, dpol))
        assert abs(L2 - L) <= 1e-13 * abs(L)
        assert abs(dL - rev) <= 1e-12 * max(abs(rev), 1e-300), (dL, rev)
        # the loss is far from its constant 1.5^2 N here (strong coupling), so a
        # step of 1e-4 leaves truncation and rounding below 1e-7 of the derivative
        eps = 1e-4
        fd = (L_oracle(phi + eps * dphi, pol + eps * dpol) - L_oracle(phi - eps * dphi, pol - eps * dpol)) / (2 * eps)
        assert abs(fd - dL) <= 1e-6 * abs(dL), (fd, dL)
