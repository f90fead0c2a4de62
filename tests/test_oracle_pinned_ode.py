"""Pins for the couplings that the Taylor-polynomial and Manley-Rowe checks leave
open: each model's per-point coupled-mode equation, written from
solvers/README.md (or, where the code departs from it, the equation the code
implements, named at the test), is integrated by an independent high-order
method (scipy DOP853, rtol 1e-13) and compared with the oracle's split step.

With every dispersion coefficient zero the dispersion step is the identity, so
each time sample evolves on its own under its local pump value, and the pump
rows are the envelope times the beam-profile factor.  With dz = 2.5e-3 the RK4
truncation stays near 1e-11 of the output, while a wrong sign, factor,
conjugation, mismatch or mode, or a stage evaluated at the wrong z (a first- or
second-order global error), moves it by 1e-6 or more; each test also checks
that a plausible mistake of that kind is seen.

Also here: the self-steepening term of Chi3GNLSE (a conserved photon number),
setPump(other) (closed forms for a dispersion-only source), the chirp/delay
phase of setPump (a chirped, delayed Gaussian in closed form), the Sinc
envelope and the GaussianApodization profiles.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from conftest import rel_l2

N_T = 16          # time samples (each one an independent ODE here)
NZ = 401          # z steps: dz = 2.5e-3 over L = 1
L = 1.0
TOL = 1e-9


def tau_axis(N, tMax):
    return 2 * tMax / N * np.fft.fftshift(np.arange(-N // 2, N // 2))


def beam(z, zR):
    return 1.0 / np.sqrt(1.0 + ((z - L / 2) / zR) ** 2)


def integrate(rhs, a0, nmodes):
    """Solve da/dz = rhs(z, a) for a [nmodes, N_T] complex from z = 0 to L."""
    shape = (nmodes, N_T)

    def f(z, y):
        a = (y[:y.size // 2] + 1j * y[y.size // 2:]).reshape(shape)
        d = rhs(z, a).reshape(-1)
        return np.concatenate([d.real, d.imag])

    y0 = np.concatenate([a0.real.reshape(-1), a0.imag.reshape(-1)])
    sol = solve_ivp(f, (0.0, L), y0, method="DOP853", rtol=1e-13, atol=1e-15)
    assert sol.success
    y = sol.y[:, -1]
    return (y[:y.size // 2] + 1j * y[y.size // 2:]).reshape(shape)


def inputs(nmodes, seed, amp=0.7):
    r = np.random.default_rng(seed)
    return amp * (r.standard_normal((nmodes, N_T)) + 1j * r.standard_normal((nmodes, N_T)))


def pump_envelope(seed, tMax=5.0, amp=0.9):
    r = np.random.default_rng(seed)
    t = tau_axis(N_T, tMax)
    return amp * np.exp(-t ** 2 / 20) * np.exp(1j * r.uniform(0, 2 * np.pi, N_T))


def oracle_outputs(m, a0, run_pump=True):
    nmodes = a0.shape[0]
    if run_pump:
        m.runPumpSimulation()
    m.runSignalSimulation(a0.reshape(-1), inTimeDomain=True)
    return np.array([np.asarray(m.signalTimes(q))[-1] for q in range(nmodes)])


def zero_disp(names):
    return {n: 0.0 for n in names}


# ---------------------------------------------------------------------------
# undepleted media
# ---------------------------------------------------------------------------

def test_chi3_signal_coupling(oracle):
    """Chi3: A0' = i/L_NL (2 |Ap|^2 A0 + Ap^2 conj(A0)), the pump a pure
    self-phase rotation Ap = p exp(i |p|^2 z / L_NL) without dispersion."""
    Lnl = 0.8
    p = pump_envelope(1)
    m = oracle.Chi3(relativeLength=L, nlLength=Lnl, beta2=0.0, customPump=p, tMax=5.0, tPrecision=N_T, nZSteps=NZ,
                    intensityProfile=oracle.IntensityProfile.Constant)
    m.runPumpSimulation()
    a0 = inputs(1, 2)

    def rhs(factor):
        def f(z, a):
            P = p * np.exp(1j * np.abs(p) ** 2 * z / Lnl)
            return 1j / Lnl * (factor * np.abs(P) ** 2 * a + P ** 2 * np.conj(a))
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(2.0), a0, 1)) < TOL
    assert rel_l2(got, integrate(rhs(1.0), a0, 1)) > 1e-3  # the factor 2 is seen


def test_chi2sfgpdc_coupling(oracle):
    """Chi2SFGPDC as the code implements it (solvers/Chi2SFGPDC.cpp:50-57):
    A0' = i/L0 Ap A1 e^{i dk0 z}, A1' = i/L1 (Ap* A0 e^{-i dk0 z} + Ap conj(A1) e^{i dk1 z});
    L0 = nlLength, L1 = nlLengthOrig, dk0 = diffBeta0, dk1 = diffBeta0o (:22-29),
    Ap = p / sqrt(1 + ((z - L/2)/zR)^2).  solvers/README.md writes the amplification
    term with conj(A0); the code amplifies mode 1 on itself, and the oracle and the
    engine follow the code."""
    L0, L1, dk0, dk1, zR = 0.9, 1.3, 2.1, -1.4, 0.7
    p = pump_envelope(3)
    m = oracle.Chi2SFGPDC(relativeLength=L, nlLength=L0, nlLengthOrig=L1, customPump=p, diffBeta0=dk0, diffBeta0o=dk1,
                          rayleighLength=zR, tMax=5.0, tPrecision=N_T, nZSteps=NZ,
                          **zero_disp(["beta2", "beta2s", "beta2o"]))
    a0 = inputs(2, 4)

    def rhs(s1, amp_mode=1):
        def f(z, a):
            P = p * beam(z, zR)
            e0, e1 = np.exp(1j * dk0 * z), np.exp(1j * s1 * dk1 * z)
            return np.array([1j / L0 * P * a[1] * e0,
                             1j / L1 * (np.conj(P) * a[0] / e0 + P * np.conj(a[amp_mode]) * e1)])
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(1), a0, 2)) < TOL
    assert rel_l2(got, integrate(rhs(-1), a0, 2)) > 1e-3
    assert rel_l2(got, integrate(rhs(1, amp_mode=0), a0, 2)) > 1e-3  # the README's form


def _sfgii(oracle, a0, p, pol=None, nZ=NZ):
    L0, L1, dk0, dk1, dk2 = 0.9, 1.2, 1.3, -2.2, 0.8
    kw = dict(relativeLength=L, nlLengthZY=L0, nlLengthZZ=L1, customPump=p, diffBeta0z=dk0, diffBeta0y=dk1,
              diffBeta0=dk2, tMax=5.0, tPrecision=N_T, nZSteps=nZ,
              # the reference binds "intensityProfile" to the zPrecision slot (solvers/Chi2SFGII.cpp:97-106)
              intensityProfile=100, zPrecision=oracle.IntensityProfile.Constant)
    kw.update(zero_disp(["beta2", "beta2sz", "beta2sy", "beta2oz", "beta2oy"]))
    if pol is not None:
        kw["polingProfile"] = list(pol)
    return oracle.Chi2SFGII(**kw), (L0, L1, dk0, dk1, dk2)


def test_chi2sfgii_coupling(oracle):
    """Chi2SFGII (modes sz, sy, oz, oy): A0' = i Ap (A3/L0 e^{i dk0 z} + A2/L1 e^{i dk2 z}),
    A1' = i/L0 Ap A2 e^{i dk1 z}, A2' = i Ap* (A1/L0 e^{-i dk1 z} + A0/L1 e^{-i dk2 z}),
    A3' = i/L0 Ap* A0 e^{-i dk0 z}; L0, L1 = nlLengthZY, nlLengthZZ and dk0, dk1, dk2 =
    diffBeta0z, diffBeta0y, diffBeta0 (solvers/Chi2SFGII.cpp:20-31)."""
    p = pump_envelope(5)
    a0 = inputs(4, 6)
    m, (L0, L1, dk0, dk1, dk2) = _sfgii(oracle, a0, p)

    def rhs(swap):
        def f(z, a):
            e0, e1, e2 = np.exp(1j * dk0 * z), np.exp(1j * dk1 * z), np.exp(1j * dk2 * z)
            x0, x1 = (L1, L0) if swap else (L0, L1)
            return np.array([1j * p * (a[3] / x0 * e0 + a[2] / x1 * e2),
                             1j / x0 * p * a[2] * e1,
                             1j * np.conj(p) * (a[1] / x0 / e1 + a[0] / x1 / e2),
                             1j / x0 * np.conj(p) * a[0] / e0])
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(False), a0, 4)) < TOL
    assert rel_l2(got, integrate(rhs(True), a0, 4)) > 1e-3


def test_chi2sfgii_last_stage_uses_previous_poling(oracle):
    """The reference's fourth RK4 stage of Chi2SFGII takes the poling of the step's
    start, not its end (solvers/Chi2SFGII.cpp:86-89).  One step across a sign
    change of the poling, with a constant pump and no mismatch, is the classical
    RK4 step of the linear system with stage poling (p0, pm, pm, p0): it differs
    from the consistent choice (p0, pm, pm, p1) by O(dz)."""
    p = np.full(N_T, 0.8 + 0.3j)
    a0 = inputs(4, 7)
    pol = [1.0, -1.0]
    m, (L0, L1, _, _, _) = _sfgii(oracle, a0, p, pol=pol, nZ=2)
    # mismatch-free: rebuild without mismatch
    kw = dict(relativeLength=L, nlLengthZY=L0, nlLengthZZ=L1, customPump=p, tMax=5.0, tPrecision=N_T, nZSteps=2,
              intensityProfile=100, zPrecision=oracle.IntensityProfile.Constant, polingProfile=pol)
    kw.update(zero_disp(["beta2", "beta2sz", "beta2sy", "beta2oz", "beta2oy"]))
    m = oracle.Chi2SFGII(**kw)
    got = oracle_outputs(m, a0)

    def f(a, s):
        return s * np.array([1j * p * (a[3] / L0 + a[2] / L1), 1j / L0 * p * a[2],
                             1j * np.conj(p) * (a[1] / L0 + a[0] / L1), 1j / L0 * np.conj(p) * a[0]])

    def rk4(stage_pol):
        h = L  # one step of the whole length
        k1 = h * f(a0, stage_pol[0])
        k2 = h * f(a0 + k1 / 2, stage_pol[1])
        k3 = h * f(a0 + k2 / 2, stage_pol[2])
        k4 = h * f(a0 + k3, stage_pol[3])
        return a0 + (k1 + 2 * k2 + 2 * k3 + k4) / 6

    pm = 0.5 * (pol[0] + pol[1])
    assert rel_l2(got, rk4((pol[0], pm, pm, pol[0]))) < 1e-14
    assert rel_l2(got, rk4((pol[0], pm, pm, pol[1]))) > 1e-2


def test_chi2sfgopa_coupling(oracle):
    """Chi2SFGOPA (pumps F, H): A0' = i/L0 Ap0 A1 e^{i dk0 z} + i/L2 Ap1 conj(A1) e^{i dk1 z},
    A1' = i/L1 Ap0* A0 e^{-i dk0 z} + i/L3 Ap1 conj(A0) e^{i dk1 z};
    L0..L3 = nlLengthSFGh, SFGf, DOPAh, DOPAf, dk0, dk1 = diffBeta0SFG, diffBeta0DOPA
    (solvers/Chi2SFGOPA.cpp:26-34)."""
    Ls = (0.9, 1.2, 1.4, 0.8)
    dk0, dk1, zR = 1.7, -0.9, 0.9
    p0, p1 = pump_envelope(8), pump_envelope(9, amp=0.6)
    kw = dict(relativeLength=L, nlLengthSFGh=Ls[0], nlLengthSFGf=Ls[1], nlLengthDOPAh=Ls[2], nlLengthDOPAf=Ls[3],
              customPump=p0, diffBeta0SFG=dk0, diffBeta0DOPA=dk1, rayleighLength=zR, tMax=5.0, tPrecision=N_T,
              nZSteps=NZ)
    kw.update(zero_disp(["beta2F", "beta2H", "beta2h", "beta2f"]))
    m = oracle.Chi2SFGOPA(**kw)
    m.setPump(p1, 0, 0, 1)
    a0 = inputs(2, 10)

    def rhs(conj_second):
        def f(z, a):
            P0, P1 = p0 * beam(z, zR), p1 * beam(z, zR)
            e0, e1 = np.exp(1j * dk0 * z), np.exp(1j * dk1 * z)
            c = np.conj if conj_second else (lambda v: v)
            return np.array([1j / Ls[0] * P0 * a[1] * e0 + 1j / Ls[2] * P1 * c(a[1]) * e1,
                             1j / Ls[1] * np.conj(P0) * a[0] / e0 + 1j / Ls[3] * P1 * c(a[0]) * e1])
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(True), a0, 2)) < TOL
    assert rel_l2(got, integrate(rhs(False), a0, 2)) > 1e-3


def test_chi2sfgxpm_coupling(oracle):
    """Chi2SFGXPM as the code implements it (solvers/Chi2SFGXPM.cpp:50-57):
    A0' = (i/L0 Ap e^{i dk z} + 2i/Lc0 |Ap|^2) A1, A1' = (i/L1 Ap* e^{-i dk z} + 2i/Lc1 |Ap|^2) A0,
    with Lc0, Lc1 = nlLengthChi3, nlLengthChi3Orig (the constructor passes them halved,
    :21-24).  solvers/README.md puts each cross-phase term on the mode's own field;
    the code multiplies the other mode's field, and the oracle and the engine follow
    the code."""
    L0, L1, Lc0, Lc1, dk, zR = 0.9, 1.3, 1.7, 2.3, 1.6, 0.8
    p = pump_envelope(11)
    kw = dict(relativeLength=L, nlLength=L0, nlLengthOrig=L1, nlLengthChi3=Lc0, nlLengthChi3Orig=Lc1, customPump=p,
              diffBeta0=dk, rayleighLength=zR, tMax=5.0, tPrecision=N_T, nZSteps=NZ)
    kw.update(zero_disp(["beta2", "beta2s", "beta2o"]))
    m = oracle.Chi2SFGXPM(**kw)
    a0 = inputs(2, 12)

    def rhs(own_mode):
        def f(z, a):
            P = p * beam(z, zR)
            e = np.exp(1j * dk * z)
            I = np.abs(P) ** 2
            if own_mode:  # the README's form
                return np.array([1j / L0 * P * e * a[1] + 2j / Lc0 * I * a[0],
                                 1j / L1 * np.conj(P) / e * a[0] + 2j / Lc1 * I * a[1]])
            return np.array([(1j / L0 * P * e + 2j / Lc0 * I) * a[1], (1j / L1 * np.conj(P) / e + 2j / Lc1 * I) * a[0]])
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(False), a0, 2)) < TOL
    assert rel_l2(got, integrate(rhs(True), a0, 2)) > 1e-3


def test_chi2afc_chirped_mismatch(oracle):
    """Chi2AFC: the SFG equations (A0' = i/L0 Ap A1 e^{i phi}, A1' = i/L1 Ap* A0 e^{-i phi})
    with L0, L1 = (pi/2) nlLength, (pi/2) nlLengthOrig (solvers/Chi2AFC.cpp:17-23), no
    poling, and a phase whose rate varies linearly along the medium:
    phi(z) = (dk_s + z (dk_e - dk_s) / (2 L)) z / 2 (solvers/Chi2AFC.cpp:33-36)."""
    Ln, Lo, ks, ke, zR = 0.8, 1.1, -3.0, 4.0, 0.9
    p = pump_envelope(13)
    kw = dict(relativeLength=L, nlLength=Ln, nlLengthOrig=Lo, customPump=p, diffBeta0Start=ks, diffBeta0End=ke,
              rayleighLength=zR, tMax=5.0, tPrecision=N_T, nZSteps=NZ)
    kw.update(zero_disp(["beta2", "beta2s", "beta2o"]))
    m = oracle.Chi2AFC(**kw)
    a0 = inputs(2, 14)
    L0, L1 = 0.5 * math.pi * Ln, 0.5 * math.pi * Lo

    def rhs(chirp):
        def f(z, a):
            P = p * beam(z, zR)
            ph = 0.5 * (ks + z * chirp * (ke - ks) / (2 * L)) * z
            e = np.exp(1j * ph)
            return np.array([1j / L0 * P * a[1] * e, 1j / L1 * np.conj(P) * a[0] / e])
        return f

    got = oracle_outputs(m, a0)
    assert rel_l2(got, integrate(rhs(1.0), a0, 2)) < TOL
    assert rel_l2(got, integrate(rhs(2.0), a0, 2)) > 1e-3


# ---------------------------------------------------------------------------
# fully nonlinear media
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("profile", ["GaussianBeam", "GaussianApodization"])
def test_chi2shgopa_coupling(oracle, profile):
    """Chi2SHGOPA (modes: fundamental, second harmonic, two amplified fields):
    A0' = i r/L0 A1 conj(A0) e^{i dk0 z}, A1' = i r/L1 (A0^2/2 e^{-i dk0 z} + A2 A3 e^{i dk1 z}),
    A2' = i r/L2 A1 conj(A3) e^{-i dk1 z}, A3' = i r/L3 A1 conj(A2) e^{-i dk1 z};
    L0..L3 = nlLengthP, SH, PA1, PA2, dk0, dk1 = diffBeta0shg, diffBeta0opa
    (solvers/Chi2SHGOPA.cpp:19-29), r the beam's amplitude factor,
    1/sqrt(1 + ((z - L/2)/zR)^2) or exp(-((z - L/2)/zR)^2 / 2)
    (src/_FullyNonlinearMedium.hpp:40-51)."""
    Ls = (0.9, 1.3, 1.1, 0.7)
    dk0, dk1, zR = 1.4, -2.1, 0.6
    kw = dict(relativeLength=L, nlLengthP=Ls[0], nlLengthSH=Ls[1], nlLengthPA1=Ls[2], nlLengthPA2=Ls[3],
              diffBeta0shg=dk0, diffBeta0opa=dk1, rayleighLength=zR, tMax=5.0, tPrecision=N_T, nZSteps=NZ,
              intensityProfile=getattr(oracle.IntensityProfile, profile))
    kw.update(zero_disp(["beta2p", "beta2sh", "beta2pa1", "beta2pa2"]))
    m = oracle.Chi2SHGOPA(**kw)
    a0 = inputs(4, 15, amp=0.8)

    def r(z):
        x = (z - L / 2) / zR
        return 1 / np.sqrt(1 + x * x) if profile == "GaussianBeam" else np.exp(-0.5 * x * x)

    def rhs(half):
        def f(z, a):
            e0, e1 = np.exp(1j * dk0 * z), np.exp(1j * dk1 * z)
            return 1j * r(z) * np.array([a[1] * np.conj(a[0]) * e0 / Ls[0],
                                         (half * a[0] ** 2 / e0 + a[2] * a[3] * e1) / Ls[1],
                                         a[1] * np.conj(a[3]) / e1 / Ls[2],
                                         a[1] * np.conj(a[2]) / e1 / Ls[3]])
        return f

    got = oracle_outputs(m, a0, run_pump=False)
    assert rel_l2(got, integrate(rhs(0.5), a0, 4)) < TOL
    assert rel_l2(got, integrate(rhs(1.0), a0, 4)) > 1e-3


def test_gnlse_spm_coefficient(oracle):
    """Chi3GNLSE with fr = 0, no self-steepening and no dispersion is pure self-phase
    modulation per sample: A(z) = A(0) exp(i r |A(0)|^2 z / L_NL) with the intensity
    factor r of a Constant profile."""
    Lnl = 0.7
    m = oracle.Chi3GNLSE(L, Lnl, math.inf, 0.0, 0.21, 0.3, 0.6, 1.5, 0.0, tMax=5.0, tPrecision=N_T, nZSteps=NZ,
                         intensityProfile=oracle.IntensityProfile.Constant)
    a0 = inputs(1, 16, amp=0.4)
    got = oracle_outputs(m, a0, run_pump=False)[0]
    exact = a0[0] * np.exp(1j * np.abs(a0[0]) ** 2 * L / Lnl)
    assert rel_l2(got, exact) < TOL


def test_gnlse_self_steepening_conserves_photon_number(oracle):
    """Chi3GNLSE with fr = 0: A' = i/L_NL (1 + i (L_NL/L_ss) d/dt)(|A|^2 A) + dispersion
    (solvers/README.md; the constructor takes {nlLength, selfSteepLength},
    solvers/Chi3GNLSE.cpp:25-26).  In the reference's frequency convention
    (omega = -pi/tMax k, so d/dt is -i omega) the operator is (1 + s omega),
    s = L_NL/L_ss, and sum_k |A_k|^2 / (1 + s omega_k) is conserved by the flow
    (and by the dispersion step); the opposite sign is not."""
    N, Lnl, Lss = 256, 1.0, 40.0
    m = oracle.Chi3GNLSE(0.6, Lnl, Lss, 0.0, 0.21, 0.3, 0.6, 1.5, -1.0, beta3=0.05, tMax=20.0, tPrecision=N,
                         nZSteps=1500)
    tau = tau_axis(N, 20.0)
    x = 2.5 / np.cosh(tau) * np.exp(0.2j * tau)
    m.runSignalSimulation(x, inTimeDomain=True)
    w = np.asarray(m.omega)
    s = Lnl / Lss
    F = np.asarray(m.signalFreq)
    n_right = np.sum(np.abs(F) ** 2 / (1 + s * w), axis=1)
    n_wrong = np.sum(np.abs(F) ** 2 / (1 - s * w), axis=1)
    assert np.abs(n_right / n_right[0] - 1).max() < 1e-8
    assert np.abs(n_wrong / n_wrong[0] - 1).max() > 1e-4
    # and self-steepening does act: the energy sum |A_k|^2 is not conserved
    e = np.sum(np.abs(F) ** 2, axis=1)
    assert np.abs(e / e[0] - 1).max() > 1e-4


# ---------------------------------------------------------------------------
# pump set-up
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("nzs_factor", [1, 2])
def test_set_pump_from_dispersion_only_source(oracle, nzs_factor):
    """setPump(other, mode, delay) (src/_NonlinearMedium.cpp:506-534) from a source that
    only disperses its input X (nlLength = inf): the source's rows are
    X exp(i (j + 1/2) dz_s D_s), and with the frame phases removed the pump row r is
    X exp(i r dzp D_p + i beta1_p delay omega) wherever r falls on a source row.  With
    a source of 2 nZ - 1 steps every row does; with a source of nZ steps the rows in
    between are the mean of two neighbours, X cos(dz_s D_s / 2) times that phase
    (X (1 + exp(-i dz_s D_s / 2)) / 2 next to the source's last row, which carries
    no half-step offset)."""
    N, nZ, tMax = 64, 21, 8.0
    nZs = nZ if nzs_factor == 1 else 2 * nZ - 1
    src = oracle.Chi2PDC(relativeLength=1.3, nlLength=math.inf, beta2=1.0, beta2s=-0.7, beta1s=0.6, beta3s=0.05,
                         tMax=tMax, tPrecision=N, nZSteps=nZs)
    tau = tau_axis(N, tMax)
    x = np.exp(-(tau - 0.5) ** 2 / 2) * np.exp(0.4j * tau)
    src.runSignalSimulation(x, inTimeDomain=True)
    b1p, delay = 0.3, 0.9
    dst = oracle.Chi2PDC(relativeLength=1.0, nlLength=1.0, beta2=-0.5, beta2s=0.2, beta1=b1p, beta3=0.1, tMax=tMax,
                         tPrecision=N, nZSteps=nZ)
    dst.setPump(src, 0, delay)
    w = np.asarray(dst.omega)
    Dp = w * (b1p + w * (0.5 * -0.5 + w * 0.1 / 6))
    Ds = w * (0.6 + w * (0.5 * -0.7 + w * 0.05 / 6))
    dzs = 1.3 / (nZs - 1)
    dzp = 1.0 / (2 * nZ - 2)
    X = np.fft.fft(x)
    PF = np.asarray(dst.pumpFreq)
    PT = np.asarray(dst.pumpTime)
    for rr in range(2 * nZ - 1):
        exp = X * np.exp(1j * rr * dzp * Dp + 1j * b1p * delay * w)
        if nzs_factor == 1 and rr % 2 == 1:
            if rr < 2 * nZ - 3:
                exp = exp * np.cos(dzs * Ds / 2)
            else:  # the last interval ends on the source's final row, which has no half-step offset
                exp = exp * 0.5 * (1 + np.exp(-0.5j * dzs * Ds))
        assert rel_l2(PF[rr], exp) < 1e-12, rr
        if rr < 2 * nZ - 2:  # the reference transforms row nZ - 1 again instead of the last row (cpp:533)
            assert rel_l2(PT[rr], np.fft.ifft(exp)) < 1e-12, rr


def test_chirp_and_delay_give_a_chirped_delayed_gaussian(oracle):
    """setPump(Gaussian, chirp, delay): the spectrum of exp(-t^2/2) times
    exp(i (beta1 delay + beta2 chirp omega / 2) omega) (src/_NonlinearMedium.cpp:180-204).
    With the reference's transform (A_k = sum A(t) e^{+i omega t}, inverse with
    e^{-i omega t}) that is (1 - i c)^(-1/2) exp(-(t - t0)^2 / (2 (1 - i c))),
    c = beta2 chirp, t0 = beta1 delay, sampled on a window wide enough for the
    discrete transform to equal the continuous one."""
    N, tMax = 512, 25.0
    b1, b2, chirp, delay = 0.7, -0.6, 1.8, 2.0
    m = oracle.Chi2PDC(relativeLength=1.0, nlLength=1.0, beta2=b2, beta2s=0.3, beta1=b1, tMax=tMax, tPrecision=N,
                       nZSteps=10, intensityProfile=oracle.IntensityProfile.Constant)
    m.setPump(oracle.PulseType.Gaussian, chirp, delay)
    m.runPumpSimulation()
    t = tau_axis(N, tMax)
    c, t0 = b2 * chirp, b1 * delay
    exact = (1 - 1j * c) ** -0.5 * np.exp(-(t - t0) ** 2 / (2 * (1 - 1j * c)))
    assert rel_l2(np.asarray(m.pumpTime)[0], exact) < 1e-12
    # a sign error in either phase term moves the pulse or flips its chirp
    wrong = (1 + 1j * c) ** -0.5 * np.exp(-(t - t0) ** 2 / (2 * (1 + 1j * c)))
    assert rel_l2(np.asarray(m.pumpTime)[0], wrong) > 1e-2


def test_sinc_envelope(oracle):
    """PulseType.Sinc: sin(t)/t with the value 1 at t = 0 (src/_NonlinearMedium.cpp:194-196)."""
    N, tMax = 128, 12.0
    m = oracle.Chi2PDC(relativeLength=1.0, nlLength=1.0, beta2=0.3, beta2s=0.3, pulseType=oracle.PulseType.Sinc,
                       tMax=tMax, tPrecision=N, nZSteps=5, intensityProfile=oracle.IntensityProfile.Constant)
    m.runPumpSimulation()
    t = tau_axis(N, tMax)
    exact = np.sinc(t / np.pi)
    assert np.abs(np.asarray(m.pumpTime)[0] - exact).max() < 1e-15


def test_gaussian_apodization_pump_profile(oracle):
    """IntensityProfile.GaussianApodization scales pump row r by
    exp(-((z_r - L/2)/zR)^2 / 2), z_r = r L / (2 nZ - 2) (src/_NonlinearMedium.cpp:242-247)."""
    N, nZ, Lm, zR = 64, 25, 1.3, 0.5
    m = oracle.Chi2PDC(relativeLength=Lm, nlLength=1, beta2=-0.7, beta2s=0.2, beta1=0.6, rayleighLength=zR,
                       intensityProfile=oracle.IntensityProfile.GaussianApodization, pulseType=oracle.PulseType.Sech,
                       tMax=8, tPrecision=N, nZSteps=nZ)
    m.runPumpSimulation()
    nZP = 2 * nZ - 1
    dzp = Lm / (nZP - 1)
    w = np.asarray(m.omega)
    D = w * (0.6 + w * (0.5 * -0.7))
    P0 = np.fft.fft(1 / np.cosh(tau_axis(N, 8)))
    for i in (0, 1, 12, 24, nZP - 1):
        prof = math.exp(-0.5 * ((i * dzp - Lm / 2) / zR) ** 2)
        assert rel_l2(np.asarray(m.pumpFreq)[i], P0 * np.exp(1j * i * dzp * D) * prof) < 1e-13
