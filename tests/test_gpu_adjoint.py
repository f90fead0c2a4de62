"""Adjoint Green's function (single-mode Chi2PDC) against central finite
differences of the oracle's Green's function.

The loss functionals are real; gradients follow torch's convention
g = dL/dRe + i dL/dIm, so a directional derivative along d is Re(sum conj(g) d).
"""
import math

import numpy as np
import pytest

from cases import SEED0
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _kwargs(N, nZ, seed):
    r = np.random.default_rng(seed)
    kw = dict(relativeLength=1.0, nlLength=float(r.uniform(0.7, 1.3)), beta2=float(r.uniform(-1, 1)),
              beta2s=float(r.uniform(-1, 1)), beta1=float(r.uniform(-1, 1)), beta1s=float(r.uniform(-1, 1)),
              diffBeta0=float(r.uniform(-3, 3)), tMax=10.0, tPrecision=N, nZSteps=nZ)
    tau = 2 * kw["tMax"] / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    env = np.exp(-tau ** 2 / 2) * (1 + 0.2 * r.standard_normal(N)) * np.exp(0.3j * r.standard_normal(N))
    pol = r.uniform(0.5, 1.5, size=nZ) * r.choice([-1.0, 1.0], size=nZ)
    return kw, env.astype(np.complex128), pol


def _oracle_greens(O, kw, env, pol, inTimeDomain=False):
    m = O.Chi2PDC(customPump=env, polingProfile=list(pol), **kw)
    return m.computeGreensFunction(inTimeDomain=inTimeDomain)


def _linear_loss(C, S, Wc, Ws):
    return float(np.real(np.vdot(Wc, C) + np.vdot(Ws, S)))


@pytest.mark.parametrize("N,nZ,ckpt,intime", [(64, 48, 0, False), (64, 48, 1, False), (128, 61, 5, True),
                                              (512, 40, 0, False)])
def test_linear_functional_gradients_match_fd(nlm, oracle, N, nZ, ckpt, intime):
    kw, env, pol = _kwargs(N, nZ, SEED0 + N + nZ)
    r = np.random.default_rng(7)
    Wc = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    Ws = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    m = nlm.Chi2PDC(customPump=env, **kw)
    m.setPolingProfile(pol)
    C, S = m.greensForward(intime, True, ckpt)
    Co, So = _oracle_greens(oracle, kw, env, pol, intime)
    assert rel_l2(np.hstack([C, S]), np.hstack([Co, So])) <= 1e-10
    g_env, g_pol = m.greensAdjoint(Wc, Ws)
    assert g_env.shape == (N,) and g_pol.shape == (nZ,)

    def L(e, p):
        return _linear_loss(*_oracle_greens(oracle, kw, e, p, intime), Wc, Ws)

    d_env = r.standard_normal(N) + 1j * r.standard_normal(N)
    d_pol = r.standard_normal(nZ)
    for de, dp in ((d_env, 0 * d_pol), (0 * d_env, d_pol)):
        eps = 1e-5
        fd = (L(env + eps * de, pol + eps * dp) - L(env - eps * de, pol - eps * dp)) / (2 * eps)
        ad = float(np.real(np.vdot(g_env, de)) + np.dot(g_pol, dp))
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3), (ad, fd)
    # single poling entries, including the first and last step
    for i in (0, nZ // 2, nZ - 1):
        e_i = np.zeros(nZ)
        e_i[i] = 1.0
        eps = 1e-5
        fd = (L(env, pol + eps * e_i) - L(env, pol - eps * e_i)) / (2 * eps)
        assert abs(g_pol[i] - fd) <= 1e-6 * max(abs(fd), 1e-2 * np.abs(g_pol).max()), (i, g_pol[i], fd)


def test_checkpoint_interval_does_not_change_gradients(nlm):
    N, nZ = 128, 77
    kw, env, pol = _kwargs(N, nZ, SEED0 + 3)
    r = np.random.default_rng(11)
    Wc = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    Ws = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    m = nlm.Chi2PDC(customPump=env, **kw)
    m.setPolingProfile(pol)
    res = []
    for k in (0, 1, 3, 76, 200):
        m.greensForward(False, True, k)
        res.append(m.greensAdjoint(Wc, Ws))
    for ge, gp in res[1:]:
        assert rel_l2(ge, res[0][0]) <= 1e-12
        assert rel_l2(gp, res[0][1]) <= 1e-12


def test_autograd_squeezing_loss_matches_fd(nlm, oracle):
    import torch
    from multimodenonlinearoptics_amd.autograd import GreensFunction, squeezing_loss
    N, nZ = 64, 33
    kw, env, pol = _kwargs(N, nZ, SEED0 + 4)
    m = nlm.Chi2PDC(customPump=env, **kw)
    e = torch.tensor(env, requires_grad=True)
    p = torch.tensor(pol, requires_grad=True)
    C, S = GreensFunction.apply(m, e, p)
    loss = squeezing_loss(S)
    loss.backward()

    def L(ev, pv):
        S_ = _oracle_greens(oracle, kw, ev, pv)[1]
        return float(np.sum(np.abs(S_ @ S_.T - 1.5 * np.eye(N)) ** 2))

    assert abs(float(loss.detach()) - L(env, pol)) <= 1e-9 * abs(L(env, pol))
    r = np.random.default_rng(5)
    de = r.standard_normal(N) + 1j * r.standard_normal(N)
    dp = r.standard_normal(nZ)
    eps = 1e-5
    fd = (L(env + eps * de, pol + eps * dp) - L(env - eps * de, pol - eps * dp)) / (2 * eps)
    ad = float(np.real(np.vdot(e.grad.numpy(), de)) + np.dot(p.grad.numpy(), dp))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_adjoint_errors(nlm):
    N, nZ = 64, 20
    kw, env, pol = _kwargs(N, nZ, SEED0 + 5)
    m = nlm.Chi2PDC(customPump=env, **kw)
    with pytest.raises(Exception):
        m.greensAdjoint(np.zeros((N, N), complex), np.zeros((N, N), complex))  # no forward yet
    m.greensForward(False, True, 0)
    with pytest.raises(ValueError):
        m.greensAdjoint(np.zeros((N, N + 1), complex), np.zeros((N, N), complex))
    with pytest.raises(ValueError):
        m.greensAdjoint(np.zeros((N, N + 1), complex), np.zeros((N, N + 1), complex))
    with pytest.raises(ValueError):
        m.greensAdjoint(np.zeros(N * N, complex), np.zeros(N * N, complex))
    # non-contiguous and real inputs are converted (the buffers are otherwise used in place)
    r = np.random.default_rng(9)
    gC = r.standard_normal((N, 2 * N))[:, ::2] + 0j
    gS = r.standard_normal((N, N))
    a = m.greensAdjoint(gC, gS)
    b = m.greensAdjoint(np.ascontiguousarray(gC), gS.astype(complex))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    sfg = nlm.Chi2SFG(relativeLength=1, nlLength=1, nlLengthOrig=1, beta2=0.1, beta2s=0.1, beta2o=0.1, tMax=10,
                      tPrecision=64)
    with pytest.raises(ValueError):
        sfg.greensForward(False, True, 0)


@pytest.mark.parametrize("N,nZ,ckpt", [(1024, 33, 0), (2048, 17, 4), (2048, 50, 0), (256, 50, 7), (4096, 9, 3),
                                       (8192, 5, 2)])
def test_gradients_match_fd_of_device_forward(nlm, N, nZ, ckpt):
    """Larger sizes (one and two fields per workgroup in the reverse sweep, and
    512 / 1024 lanes per field from N = 4096): the finite differences come from
    the device forward, itself pinned to the oracle by the parity tests."""
    print("adjoint vs device FD, N=%d nZ=%d ckpt=%d" % (N, nZ, ckpt), flush=True)
    kw, env, pol = _kwargs(N, nZ, SEED0 + 7 * N + nZ)
    r = np.random.default_rng(13)
    Wc = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    Ws = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    m = nlm.Chi2PDC(customPump=env, **kw)
    m.setPolingProfile(pol)
    m.greensForward(False, True, ckpt)
    g_env, g_pol = m.greensAdjoint(Wc, Ws)

    def L(e, p):
        m.setPump(e)
        m.setPolingProfile(p)
        return _linear_loss(*m.computeGreensFunction(), Wc, Ws)

    for de, dp in ((r.standard_normal(N) + 1j * r.standard_normal(N), np.zeros(nZ)),
                   (np.zeros(N, complex), r.standard_normal(nZ))):
        eps = 1e-5
        fd = (L(env + eps * de, pol + eps * dp) - L(env - eps * de, pol - eps * dp)) / (2 * eps)
        ad = float(np.real(np.vdot(g_env, de)) + np.dot(g_pol, dp))
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3), (ad, fd)


def test_adjoint_refuses_a_changed_medium(nlm):
    """The adjoint reads the checkpoints and rows of the latest greensForward and the
    medium's current tables: a Green's run or a coefficient change in between must
    raise instead of mixing two forwards."""
    N, nZ = 64, 20
    kw, env, pol = _kwargs(N, nZ, SEED0 + 6)
    m = nlm.Chi2PDC(customPump=env, **kw)
    m.setPolingProfile(pol)
    g = (np.ones((N, N), complex), np.ones((N, N), complex))
    m.greensForward(False, True, 0)
    ref = m.greensAdjoint(*g)
    m.computeGreensFunction()
    with pytest.raises(RuntimeError):
        m.greensAdjoint(*g)
    m.greensForward(False, True, 0)
    m.setPolingProfile(pol[::-1].copy())
    with pytest.raises(RuntimeError):
        m.greensAdjoint(*g)
    m.setPolingProfile(pol)
    m.greensForward(False, True, 0)
    again = m.greensAdjoint(*g)
    assert np.array_equal(again[0], ref[0]) and np.array_equal(again[1], ref[1])
    # two autograd forwards on one medium: the first graph's backward must refuse
    import torch
    from multimodenonlinearoptics_amd.autograd import GreensFunction
    e = torch.tensor(env, requires_grad=True)
    p = torch.tensor(pol, requires_grad=True)
    _, S1 = GreensFunction.apply(m, e, p)
    _, S2 = GreensFunction.apply(m, e, p)
    with pytest.raises(RuntimeError):
        S1.abs().sum().backward()
    S2.abs().sum().backward()
    assert e.grad is not None and torch.isfinite(e.grad).all()
