"""Pins for the CPU oracle: closed forms, invariants and brute force.

Each test checks the oracle against something the mathematics fixes rather
than against its own formulas:
  * the DFT definition (naive O(N^2) sums) and numpy's FFT;
  * dispersion-only propagation, which is an exact phase exp(i z D(omega));
  * the nonlinear (SPM) pump with no dispersion, which is an exact phase;
  * constant-coefficient RK4, which equals the 4th-order Taylor polynomial of
    the per-point step matrix (built here from the coupled-mode equations of
    solvers/README.md, not from the oracle's stage code);
  * the Green's identity a_out = C a_in + S conj(a_in) with the fftshift
    conventions of tests/04 - Green's Function.py:98-111;
  * symplecticity of the PDC map and unitarity of the normalised SFG map;
  * hand-computed poling patterns and the Raman response formula.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2


def naive_dft(x, sign=-1):
    n = x.shape[-1]
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ W.T


@pytest.mark.parametrize("n", [2, 6, 8, 12, 14, 30, 32, 48, 64])
def test_fft_against_naive_dft(oracle, n):
    r = np.random.default_rng(n)
    x = r.standard_normal((3, n)) + 1j * r.standard_normal((3, n))
    assert rel_l2(oracle.fft(x), naive_dft(x)) < 1e-13
    assert rel_l2(oracle.fft(x, inverse=True), naive_dft(x, +1) / n) < 1e-13


@pytest.mark.parametrize("n", [128, 512, 1000, 4096])
def test_fft_against_numpy(oracle, n):
    r = np.random.default_rng(n)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    assert rel_l2(oracle.fft(x), np.fft.fft(x)) < 1e-14
    assert rel_l2(oracle.fft(x, inverse=True), np.fft.ifft(x)) < 1e-14


def test_axes_closed_form(oracle):
    N, tMax = 64, 7.0
    m = oracle.Chi2PDC(1, 1, 0.5, 0.5, tMax=tMax, tPrecision=N, nZSteps=10)
    k = np.fft.fftfreq(N) * N
    assert np.allclose(m.tau, 2 * tMax / N * k, rtol=0, atol=1e-15)
    # the reference's frequency axis is the negated numpy convention
    assert np.allclose(m.omega, -np.pi / tMax * k, rtol=0, atol=1e-15)


def _disp(w, b1, b2, b3):
    return w * (b1 + w * (0.5 * b2 + w * b3 / 6))


@pytest.mark.parametrize("in_time", [True, False])
def test_dispersion_only_is_exact_phase(oracle, in_time):
    N, nZ, L = 128, 40, 1.7
    b = dict(beta2=0.3, beta2s=-0.8, beta1=0.4, beta1s=1.0, beta3=0.05, beta3s=-0.1)
    m = oracle.Chi2PDC(relativeLength=L, nlLength=math.inf, tMax=15, tPrecision=N, nZSteps=nZ, **b)
    w = m.omega
    x = np.exp(-(m.tau - 1.0) ** 2) * np.exp(0.3j * m.tau)
    xf = np.fft.fft(x)
    m.runSignalSimulation(x if in_time else xf, inTimeDomain=in_time)
    D = _disp(w, b["beta1s"], b["beta2s"], b["beta3s"])
    dz = L / (nZ - 1)
    F = m.signalFreq
    assert rel_l2(F[-1], xf * np.exp(1j * L * D)) < 1e-13
    # interior rows carry half a dispersion step of offset
    for i in (0, 1, 17, nZ - 2):
        assert rel_l2(F[i], xf * np.exp(1j * (i + 0.5) * dz * D)) < 1e-13
    assert rel_l2(m.signalTime[-1], np.fft.ifft(xf * np.exp(1j * L * D))) < 1e-13


def test_linear_pump_grid_closed_form(oracle):
    N, nZ, L, zR = 64, 25, 1.3, 0.9
    m = oracle.Chi2PDC(relativeLength=L, nlLength=1, beta2=-0.7, beta2s=0.2, beta1=0.6, beta3=0.04,
                       rayleighLength=zR, intensityProfile=oracle.IntensityProfile.GaussianBeam,
                       pulseType=oracle.PulseType.Sech, tMax=8, tPrecision=N, nZSteps=nZ)
    m.runPumpSimulation()
    nZP = 2 * nZ - 1
    dzp = L / (nZP - 1)
    D = _disp(m.omega, 0.6, -0.7, 0.04)
    P0 = np.fft.fft(1 / np.cosh(m.tau))
    z = np.linspace(-L / 2, L / 2, nZP)
    prof = 1 / np.sqrt(1 + (z / zR) ** 2)
    for i in (0, 1, 10, nZP - 1):
        exp_f = P0 * np.exp(1j * i * dzp * D) * prof[i]
        assert rel_l2(m.pumpFreq[i], exp_f) < 1e-13
        assert rel_l2(m.pumpTime[i], np.fft.ifft(exp_f)) < 1e-13


def test_chi3_pump_pure_spm_is_exact(oracle):
    # beta2 = beta3 = 0: each pump row is a pure self-phase rotation
    N, nZ, L, Lnl = 64, 30, 0.8, 1.0
    m = oracle.Chi3(relativeLength=L, nlLength=Lnl, beta2=0, pulseType=oracle.PulseType.Sech, tMax=6,
                    tPrecision=N, nZSteps=nZ)
    m.runPumpSimulation()
    P0 = 1 / np.cosh(m.tau)
    nZP = 2 * nZ - 1
    z = np.arange(nZP) * L / (nZP - 1)
    exact = P0[None, :] * np.exp(1j * np.abs(P0[None, :]) ** 2 * z[:, None] / Lnl)
    assert rel_l2(m.pumpTime, exact) < 1e-13


def test_chi3_fundamental_soliton_is_stationary(oracle):
    # N = 1 soliton (beta2 = -1, L_NL = 1): |A(z)| stays sech(t) (Agrawal ch. 5)
    m = oracle.Chi3(relativeLength=math.pi / 2, nlLength=1, beta2=-1, pulseType=oracle.PulseType.Sech, tMax=10,
                    tPrecision=512, zPrecision=637)
    m.runPumpSimulation()
    err = np.abs(np.abs(m.pumpTime[-1]) - 1 / np.cosh(m.tau)).max()
    assert err < 1e-4


def taylor4(A):
    I = np.eye(A.shape[0])
    A2 = A @ A
    return I + A + A2 / 2 + A2 @ A / 6 + A2 @ A2 / 24


def realify(Mc, Mcc):
    """Real 2n x 2n matrix of the map a -> Mc a + Mcc conj(a) on (Re a, Im a)."""
    return np.block([[Mc.real + Mcc.real, -Mc.imag + Mcc.imag], [Mc.imag + Mcc.imag, Mc.real - Mcc.real]])


COUPLING = {
    # per-point coupled-mode equations of solvers/README.md with a constant real
    # pump p and unit poling, zero mismatch: da/dz = Mc a + Mcc conj(a)
    "Chi2PDC": lambda p, nl: (np.zeros((1, 1)), np.array([[1j * p / nl[0]]])),
    "Chi2SFG": lambda p, nl: (np.array([[0, 1j * p / nl[0]], [1j * p / nl[1], 0]]), np.zeros((2, 2))),
    "Chi2PDCII": lambda p, nl: (np.zeros((2, 2)), np.array([[0, 1j * p / nl[0]], [1j * p / nl[1], 0]])),
}


@pytest.mark.parametrize("name", list(COUPLING))
def test_rk4_equals_taylor_polynomial_constant_pump(oracle, name):
    N, nZ, L = 32, 21, 1.0
    p = 0.8
    nl = [0.9, 1.3]
    kw = dict(relativeLength=L, tMax=5, tPrecision=N, nZSteps=nZ, customPump=np.full(N, p, dtype=complex))
    if name == "Chi2PDC":
        m = oracle.Chi2PDC(nlLength=nl[0], beta2=0, beta2s=0, **kw)
        nS = 1
    elif name == "Chi2SFG":
        m = oracle.Chi2SFG(nlLength=nl[0], nlLengthOrig=nl[1], beta2=0, beta2s=0, beta2o=0, **kw)
        nS = 2
    else:
        m = oracle.Chi2PDCII(nlLengthI=nl[0], nlLengthII=nl[1], beta2=0, beta2s=0, beta2o=0, **kw)
        nS = 2
    m.runPumpSimulation()
    r = np.random.default_rng(1)
    x = (r.standard_normal((1, nS * N)) + 1j * r.standard_normal((1, nS * N))) * 0.5
    out = m.batchSignalSimulation(x, inTimeDomain=True, runPump=False)
    dz = L / (nZ - 1)
    Mc, Mcc = COUPLING[name](p, nl)
    T = np.linalg.matrix_power(taylor4(realify(Mc, Mcc) * dz), nZ - 1)
    for k in (0, 5, 31):
        a = x[0, k::N][:nS]
        v = T @ np.concatenate([a.real, a.imag])
        exp = v[:nS] + 1j * v[nS:]
        got = out[0, k::N][:nS]
        assert np.abs(got - exp).max() < 1e-13


def _green_identity(m, x_time):
    """a_out (frequency domain) from C, S with the reference's shift conventions."""
    C, S = m.computeGreensFunction()
    xf = np.fft.fft(x_time)
    out = C @ np.fft.fftshift(xf) + S @ np.fft.fftshift(xf.conj())
    return np.fft.ifftshift(out)


@pytest.mark.parametrize("name", ["Chi3", "Chi2PDC"])
def test_greens_identity(oracle, name):
    N, nZ = 64, 60
    if name == "Chi3":
        m = oracle.Chi3(relativeLength=1.0, nlLength=1, beta2=-1, pulseType=oracle.PulseType.Sech, tMax=10,
                        tPrecision=N, nZSteps=nZ)
    else:
        m = oracle.Chi2PDC(relativeLength=1.0, nlLength=0.7, beta2=0.4, beta2s=-0.3, beta1=0.5, beta1s=-0.2,
                           diffBeta0=1.1, tMax=10, tPrecision=N, nZSteps=nZ, poling=[0.3, 0.5, 0.2])
    r = np.random.default_rng(3)
    x = (r.standard_normal(N) + 1j * r.standard_normal(N)) * np.exp(-np.asarray(m.tau) ** 2 / 8)
    pred = _green_identity(m, x)
    m.runSignalSimulation(x, inTimeDomain=True)
    assert rel_l2(pred, m.signalFreq[-1]) < 1e-12


def test_pdc_map_is_symplectic(oracle):
    m = oracle.Chi2PDC(relativeLength=1.0, nlLength=1.0, beta2=0.2, beta2s=-0.4, beta1=0.3, beta1s=-0.3,
                       diffBeta0=0.7, tMax=12, tPrecision=128, nZSteps=400)
    C, S = m.computeGreensFunction()
    n = C.shape[0]
    assert np.abs(C @ C.conj().T - S @ S.conj().T - np.eye(n)).max() < 1e-9
    assert np.abs(C @ S.T - S @ C.T).max() < 1e-9


def test_sfg_normalised_map_is_unitary(oracle):
    m = oracle.Chi2SFG(relativeLength=1.0, nlLength=0.8, nlLengthOrig=1.4, beta2=0.2, beta2s=-0.4, beta2o=0.3,
                       beta1s=0.5, beta1o=-0.5, diffBeta0=0.6, tMax=12, tPrecision=64, nZSteps=400)
    C, S = m.computeGreensFunction(normalize=True)
    assert np.abs(S).max() < 1e-12
    assert np.abs(C @ C.conj().T - np.eye(C.shape[0])).max() < 1e-9


def test_poling_hand_cases(oracle):
    m = oracle.Chi2PDC(1, 1, 0.5, 0.5, tPrecision=16, nZSteps=10, poling=[1, 1])
    # cumulative [1, 2] * 10/2 = [5, 10]: steps 0-4 +1, step 5 = +1*(2*0-1), steps 6-9 -1
    assert np.array_equal(m.poling, [1, 1, 1, 1, 1, -1, -1, -1, -1, -1])
    m = oracle.Chi2PDC(1, 1, 0.5, 0.5, tPrecision=16, nZSteps=10, poling=[1.25, 2.75])
    # cumulative [1.25, 4] * 10/4 = [3.125, 10]: boundary step 3 = 2*0.125-1 = -0.75
    assert np.allclose(m.poling, [1, 1, 1, -0.75, -1, -1, -1, -1, -1, -1])
    m = oracle.Chi2PDC(1, 1, 0.5, 0.5, tPrecision=16, nZSteps=10)
    assert np.array_equal(m.poling, np.ones(10))
    with pytest.raises(ValueError, match="Poling contains invalid domain length"):
        oracle.Chi2PDC(1, 1, 0.5, 0.5, tPrecision=16, nZSteps=10, poling=[1, -1])
    with pytest.raises(ValueError, match="Poling period too small"):
        oracle.Chi2PDC(1, 1, 0.5, 0.5, tPrecision=16, nZSteps=4, poling=[1, 0.01, 0.01, 1])


def test_raman_response_formula(oracle):
    fr, fb, t1, t2, t3 = 0.18, 0.21, 0.3, 0.6, 1.5
    m = oracle.Chi3GNLSE(1.0, 1.0, 50.0, fr, fb, t1, t2, t3, -1.0, tMax=20, tPrecision=128, nZSteps=5)
    tau = np.asarray(m.tau)
    h = np.zeros(128)
    t = tau[:64]
    h[:64] = fr * (1 - fb) * (t1 / t2 ** 2 + 1 / t1) * np.exp(-t / t2) * np.sin(t / t1) \
        + fr * fb / t3 ** 2 * np.exp(-t / t3) * (2 * t3 - t)
    h[0] += 1 - fr
    assert rel_l2(m.ramanResponse, np.fft.fft(h)) < 1e-14


def test_gnlse_conserves_energy_without_self_steepening(oracle):
    m = oracle.Chi3GNLSE(0.5, 1.0, math.inf, 0.18, 0.21, 0.25, 0.6, 1.9, -1.0, tMax=20, tPrecision=256,
                         nZSteps=400)
    tau = np.asarray(m.tau)
    x = 2 / np.cosh(tau) + 0j
    m.runSignalSimulation(x)
    e = np.sum(np.abs(m.signalTime) ** 2, axis=1)
    assert np.abs(e / e[0] - 1).max() < 1e-8


@pytest.mark.parametrize("kw,msg", [
    (dict(tPrecision=511), "Invalid number of Frequencies"),
    (dict(tPrecision=0), "Invalid number of Frequencies"),
    (dict(tMax=-1), "Negative time span"),
    (dict(relativeLength=-1), "Non-positive length scale"),
    (dict(zPrecision=0), "Zero steps"),
])
def test_reference_errors(oracle, kw, msg):
    base = dict(relativeLength=1.0, nlLength=1.0, beta2=0.5, beta2s=0.5, tPrecision=64)
    base.update(kw)
    with pytest.raises(ValueError, match=msg):
        oracle.Chi2PDC(**base)


def test_unit_length_error(oracle):
    with pytest.raises(ValueError, match="No unit length scale provided"):
        oracle.Chi2PDC(relativeLength=2.0, nlLength=3.0, beta2=0.5, beta2s=0.5, tPrecision=64)


def test_fully_nonlinear_has_no_pump_methods(oracle):
    m = oracle.Chi2SHG(1.0, 1.0, 1.0, 0.2, 0.3, tPrecision=32, nZSteps=10)
    for call in (m.runPumpSimulation, m.computeGreensFunction, lambda: m.setPump(0)):
        with pytest.raises(RuntimeError, match="Object does not have this method."):
            call()


# ---------------------------------------------------------------------------
# Manley-Rowe invariants of the fully nonlinear chi(2) media.  With coupling
# constants kappa_p = 1/L_NL[p] the flows conserve weighted photon numbers
# pointwise in time for any unit-modulus phase mismatch, poling sign and
# beam-profile factor, and the dispersion step conserves each mode's energy,
# so the sums below drift only by RK4 truncation.  Each check fails if a
# coupling constant lands on the wrong mode or the factor 1/2 of the
# second-harmonic term moves.
# ---------------------------------------------------------------------------

def _energies(m, modes):
    return [float(np.sum(np.abs(np.asarray(m.signalTimes(q))) ** 2, axis=1)[[0, -1]].tolist()[0]) for q in modes], \
        [float(np.sum(np.abs(np.asarray(m.signalTimes(q))[-1]) ** 2)) for q in modes]


def _run_fnlm(m, nmodes, seed, amp=0.6):
    r = np.random.default_rng(seed)
    N = np.asarray(m.tau).size
    tau = np.asarray(m.tau)
    x = np.concatenate([amp * np.exp(-tau ** 2 / 2) * np.exp(1j * r.uniform(0, 6, N)) for _ in range(nmodes)])
    m.runSignalSimulation(x, inTimeDomain=True)
    first = [float(np.sum(np.abs(np.asarray(m.signalTimes(q))[0]) ** 2)) for q in range(nmodes)]
    last = [float(np.sum(np.abs(np.asarray(m.signalTimes(q))[-1]) ** 2)) for q in range(nmodes)]
    return np.array(first), np.array(last)


@pytest.mark.parametrize("name", ["Chi2SHG", "Chi2SHGXPM", "Chi2ASHG"])
def test_shg_manley_rowe(oracle, name):
    LH, LP = 0.7, 1.9
    kw = dict(relativeLength=1.0, nlLengthH=LH, nlLengthP=LP, beta2h=0.3, beta2p=-0.2, beta1h=0.4, beta1p=-0.1,
              tMax=10, tPrecision=64, nZSteps=800, rayleighLength=1.3)
    if name == "Chi2SHGXPM":
        kw["nlLengthChi3"] = 2.0
    if name == "Chi2ASHG":
        kw.update(diffBeta0Start=-2.0, diffBeta0End=3.0)
    else:
        kw["diffBeta0"] = 1.5
    m = getattr(oracle, name)(**kw)
    e0, e1 = _run_fnlm(m, 2, seed=4)
    # k[0] = i dz/L_P (...) conj(A) B,  k[1] = i dz/L_H (...) A^2 / 2   ->  |A|^2/L_H + 2|B|^2/L_P
    q0 = e0[0] / LH + 2 * e0[1] / LP
    q1 = e1[0] / LH + 2 * e1[1] / LP
    assert abs(q1 - q0) <= 1e-9 * q0
    assert abs(e1[1] - e0[1]) > 1e-3 * q0  # real conversion happened
    swapped = (e1[0] / LP + 2 * e1[1] / LH) - (e0[0] / LP + 2 * e0[1] / LH)
    assert abs(swapped) > 1e-4 * q0


def test_dsfg_manley_rowe(oracle):
    Lp, Ls, Ld = 0.8, 1.3, 2.1
    m = oracle.Chi2DSFG(relativeLength=1.0, nlLengthP=Lp, nlLengthS=Ls, nlLengthD=Ld, beta2p=0.2, beta2s=-0.3,
                        beta2d=0.1, beta1p=0.2, beta1s=0.0, beta1d=-0.3, diffBeta0=0.9, tMax=10, tPrecision=64,
                        nZSteps=800)
    e0, e1 = _run_fnlm(m, 3, seed=6)
    kp, ks, kd = 1 / Lp, 1 / Ls, 1 / Ld
    for a, b, ka, kb in ((0, 1, ks, kp), (1, 2, kd, ks)):
        q0 = ka * e0[a] + kb * e0[b]
        q1 = ka * e1[a] + kb * e1[b]
        assert abs(q1 - q0) <= 1e-9 * q0
    assert np.abs(e1 - e0).max() > 1e-3 * e0.sum()


# ---------------------------------------------------------------------------
# Cascade (src/Cascade.cpp:110-135): the composed Green's function applied to
# an input equals the signal propagated through the media one after another
# (src/Cascade.cpp:91-107).  Composing in the wrong order or conjugating the
# wrong factor fails this.
# ---------------------------------------------------------------------------

def test_cascade_composition_equals_sequential_propagation(oracle):
    N, nZ = 64, 50
    a = oracle.Chi2PDC(relativeLength=1.0, nlLength=0.8, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5,
                       diffBeta0=0.7, tMax=10, tPrecision=N, nZSteps=nZ)
    b = oracle.Chi2PDC(relativeLength=0.7, nlLength=1.0, beta2=-0.2, beta2s=0.5, beta1=-0.3, beta1s=0.4,
                       diffBeta0=-1.2, tMax=10, tPrecision=N, nZSteps=nZ + 7)
    cas = oracle.Cascade([a, b], [{0: 0}], False)
    C, S = cas.computeGreensFunction()
    r = np.random.default_rng(9)
    x = (r.standard_normal(N) + 1j * r.standard_normal(N)) * np.exp(-np.asarray(a.tau) ** 2 / 8)
    cas.runSignalSimulation(x, inTimeDomain=True)
    xf = np.fft.fft(x)
    pred = np.fft.ifftshift(C @ np.fft.fftshift(xf) + S @ np.fft.fftshift(xf.conj()))
    assert rel_l2(pred, b.signalFreq[-1]) < 1e-12
    # the reversed composition is a different operator
    Ca, Sa = a.computeGreensFunction(runPump=False)
    Cb, Sb = b.computeGreensFunction(runPump=False)
    assert rel_l2(Ca @ Cb + Sa @ Sb.conj(), C) > 1e-3
