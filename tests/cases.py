"""Seeded solver configurations shared by the oracle and the engine tests.

Each builder returns (class name, kwargs) given the module whose enums to use
(the product's `nonlinearmedium` or the oracle), so the two sides are built
from identical arguments.
"""
import math

import numpy as np

UNDEPLETED = ["Chi3", "Chi2PDC", "Chi2SFG", "Chi2AFC", "Chi2SFGPDC", "Chi2SFGII", "Chi2PDCII", "Chi2SFGOPA",
              "Chi2SFGXPM"]
DEPLETED = ["Chi2SHG", "Chi2SHGOPA", "Chi2DSFG", "Chi2SHGXPM", "Chi2ASHG", "Chi3GNLSE"]
ALL = UNDEPLETED + DEPLETED
NMODES = {"Chi3": 1, "Chi2PDC": 1, "Chi2SFG": 2, "Chi2AFC": 2, "Chi2SFGPDC": 2, "Chi2SFGII": 4, "Chi2PDCII": 2,
          "Chi2SFGOPA": 2, "Chi2SFGXPM": 2, "Chi2SHG": 2, "Chi2SHGOPA": 4, "Chi2DSFG": 3, "Chi2SHGXPM": 2,
          "Chi2ASHG": 2, "Chi3GNLSE": 1}


def small_case(name, mod, N=64, nZ=48, seed=0, profile="GaussianBeam", poled=True):
    """Random but well-conditioned parameters for solver `name`."""
    r = np.random.default_rng(seed)
    u = lambda a, b: float(r.uniform(a, b))
    IP = mod.IntensityProfile
    prof = getattr(IP, profile)
    zR = u(0.8, 2.0) if profile != "Constant" else math.inf
    tMax = 10.0
    pol = list(r.uniform(0.5, 1.5, size=7)) if poled else []
    common = dict(tMax=tMax, tPrecision=N, nZSteps=nZ)
    if name == "Chi3":
        return name, dict(relativeLength=u(0.5, 1.5), nlLength=1.0, beta2=-1.0, beta3=u(-0.1, 0.1),
                          pulseType=mod.PulseType.Sech, rayleighLength=zR, intensityProfile=prof, **common)
    if name == "Chi2PDC":
        return name, dict(relativeLength=1.0, nlLength=u(0.6, 1.4), beta2=u(-1, 1), beta2s=u(-1, 1),
                          beta1=u(-1, 1), beta1s=u(-1, 1), beta3=u(-.1, .1), beta3s=u(-.1, .1),
                          diffBeta0=u(-3, 3), rayleighLength=zR, intensityProfile=prof, poling=pol, **common)
    if name in ("Chi2SFG", "Chi2PDCII", "Chi2SFGPDC", "Chi2AFC"):
        names = {"Chi2SFG": ("nlLength", "nlLengthOrig"), "Chi2PDCII": ("nlLengthI", "nlLengthII"),
                 "Chi2SFGPDC": ("nlLength", "nlLengthOrig"), "Chi2AFC": ("nlLength", "nlLengthOrig")}[name]
        kw = dict(relativeLength=1.0, beta2=u(-1, 1), beta2s=u(-1, 1), beta2o=u(-1, 1), beta1=u(-1, 1),
                  beta1s=u(-1, 1), beta1o=u(-1, 1), beta3=u(-.1, .1), beta3s=u(-.1, .1), beta3o=u(-.1, .1),
                  rayleighLength=zR, intensityProfile=prof, **common)
        kw[names[0]] = u(0.6, 1.4)
        kw[names[1]] = u(0.6, 1.4)
        if name == "Chi2AFC":
            kw.update(diffBeta0Start=u(-3, 0), diffBeta0End=u(0, 3))
        else:
            kw.update(diffBeta0=u(-3, 3), poling=pol)
            if name == "Chi2SFGPDC":
                kw["diffBeta0o"] = u(-3, 3)
        return name, kw
    if name == "Chi2SFGII":
        # the reference binds "intensityProfile" to the zPrecision slot and vice versa
        kw = dict(relativeLength=1.0, nlLengthZY=u(0.6, 1.4), nlLengthZZ=u(0.6, 1.4), beta2=u(-1, 1),
                  beta2sz=u(-1, 1), beta2sy=u(-1, 1), beta2oz=u(-1, 1), beta2oy=u(-1, 1), beta1=u(-1, 1),
                  beta1sz=u(-1, 1), beta1sy=u(-1, 1), beta1oz=u(-1, 1), beta1oy=u(-1, 1), diffBeta0z=u(-3, 3),
                  diffBeta0y=u(-3, 3), diffBeta0=u(-3, 3), rayleighLength=zR, poling=pol,
                  intensityProfile=100, zPrecision=prof, **common)
        return name, kw
    if name == "Chi2SFGOPA":
        return name, dict(relativeLength=1.0, nlLengthSFGh=u(.6, 1.4), nlLengthSFGf=u(.6, 1.4),
                          nlLengthDOPAh=u(.6, 1.4), nlLengthDOPAf=u(.6, 1.4), beta2F=u(-1, 1), beta2H=u(-1, 1),
                          beta2h=u(-1, 1), beta2f=u(-1, 1), beta1F=u(-1, 1), beta1H=u(-1, 1), beta1h=u(-1, 1),
                          beta1f=u(-1, 1), diffBeta0SFG=u(-3, 3), diffBeta0DOPA=u(-3, 3), rayleighLength=zR,
                          intensityProfile=prof, poling=pol, **common)
    if name == "Chi2SFGXPM":
        return name, dict(relativeLength=1.0, nlLength=u(.6, 1.4), nlLengthOrig=u(.6, 1.4), nlLengthChi3=u(1, 3),
                          nlLengthChi3Orig=u(1, 3), beta2=u(-1, 1), beta2s=u(-1, 1), beta2o=u(-1, 1),
                          beta1=u(-1, 1), beta1s=u(-1, 1), beta1o=u(-1, 1), diffBeta0=u(-3, 3), rayleighLength=zR,
                          intensityProfile=prof, poling=pol, **common)
    if name == "Chi2SHG":
        return name, dict(relativeLength=1.0, nlLengthH=u(.6, 1.4), nlLengthP=u(.6, 1.4), beta2h=u(-1, 1),
                          beta2p=u(-1, 1), beta1h=u(-1, 1), beta1p=u(-1, 1), beta3h=u(-.1, .1), beta3p=u(-.1, .1),
                          diffBeta0=u(-3, 3), rayleighLength=zR, intensityProfile=prof, poling=pol, **common)
    if name == "Chi2SHGOPA":
        return name, dict(relativeLength=1.0, nlLengthP=u(.6, 1.4), nlLengthSH=u(.6, 1.4), nlLengthPA1=u(.6, 1.4),
                          nlLengthPA2=u(.6, 1.4), beta2p=u(-1, 1), beta2sh=u(-1, 1), beta2pa1=u(-1, 1),
                          beta2pa2=u(-1, 1), beta1p=u(-1, 1), beta1sh=u(-1, 1), beta1pa1=u(-1, 1), beta1pa2=u(-1, 1),
                          diffBeta0shg=u(-3, 3), diffBeta0opa=u(-3, 3), rayleighLength=zR, intensityProfile=prof,
                          poling=pol, **common)
    if name == "Chi2DSFG":
        return name, dict(relativeLength=1.0, nlLengthP=u(.6, 1.4), nlLengthS=u(.6, 1.4), nlLengthD=u(.6, 1.4),
                          beta2p=u(-1, 1), beta2s=u(-1, 1), beta2d=u(-1, 1), beta1p=u(-1, 1), beta1s=u(-1, 1),
                          beta1d=u(-1, 1), diffBeta0=u(-3, 3), rayleighLength=zR, intensityProfile=prof, poling=pol,
                          **common)
    if name == "Chi2SHGXPM":
        return name, dict(relativeLength=1.0, nlLengthH=u(.6, 1.4), nlLengthP=u(.6, 1.4), nlLengthChi3=u(1, 3),
                          beta2h=u(-1, 1), beta2p=u(-1, 1), beta1h=u(-1, 1), beta1p=u(-1, 1), diffBeta0=u(-3, 3),
                          rayleighLength=zR, intensityProfile=prof, poling=pol, **common)
    if name == "Chi2ASHG":
        return name, dict(relativeLength=1.0, nlLengthH=u(.6, 1.4), nlLengthP=u(.6, 1.4), beta2h=u(-1, 1),
                          beta2p=u(-1, 1), beta1h=u(-1, 1), beta1p=u(-1, 1), diffBeta0Start=u(-3, 0),
                          diffBeta0End=u(0, 3), rayleighLength=zR, intensityProfile=prof, **common)
    if name == "Chi3GNLSE":
        return name, dict(relativeLength=u(0.3, 0.8), nlLength=1.0, selfSteepLength=u(20, 100), fr=0.18, fb=0.21,
                          tau1=0.0122 * 20, tau2=0.032 * 20, tau3=0.096 * 20, beta2=-1.0, beta3=u(-.05, .05),
                          rayleighLength=zR, intensityProfile=prof, **common)
    raise KeyError(name)


def make(mod, name, kw):
    obj = getattr(mod, name)(**kw)
    if name == "Chi2SFGOPA":
        # second pump: a delayed Gaussian
        tau = np.asarray(obj.tau)
        obj.setPump(np.exp(-0.5 * (tau - 1.0) ** 2) * 0.7, 0, 0, 1)
    return obj


def random_inputs(name, N, rows, seed, nch=None):
    r = np.random.default_rng(seed)
    nch = nch or NMODES[name]
    shape = (rows, nch * N)
    x = (r.standard_normal(shape) + 1j * r.standard_normal(shape)) * 0.3
    if name in ("Chi3GNLSE",):
        x *= 0.5
    return x


# ---------------------------------------------------------------------------
# the configurations quoted by the benchmark definition
# ---------------------------------------------------------------------------

SEED0 = 20261202


def config1(mod):
    """Chi3 soliton pump N=512, tMax=10, z=pi/2 (zPrecision 637 -> nZ=1000), Gaussian signal."""
    fiber = mod.Chi3(relativeLength=math.pi / 2, nlLength=1, beta2=-1, beta3=0, pulseType=mod.PulseType.Sech,
                     tMax=10, tPrecision=512, zPrecision=637)
    r = np.random.default_rng(SEED0 + 1)
    phi = r.uniform(0, 2 * math.pi)
    tau = np.asarray(fiber.tau)
    sig = np.exp(-tau ** 2 / 2) * np.exp(1j * phi)
    return fiber, sig


def config2_kwargs():
    r = np.random.default_rng(SEED0 + 2)
    w = r.uniform(0.5, 2)
    b1, b1s = r.uniform(-2, 2, size=2)
    b2, b2s = r.uniform(-1, 1, size=2)
    d0 = r.uniform(-5, 5)
    return dict(relativeLength=1.0, nlLength=1.0, beta2=float(b2), beta2s=float(b2s), beta1=float(b1),
                beta1s=float(b1s), diffBeta0=float(d0), tMax=20.0, tPrecision=512, nZSteps=1000), float(w)


def config2(mod, N=512, nZ=1000):
    """Chi2PDC Green's function: N=512, tMax=20, nZ=1000, seeded Gaussian pump and dispersion."""
    kw, w = config2_kwargs()
    kw = dict(kw, tPrecision=N, nZSteps=nZ)
    tau = 2 * kw["tMax"] / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    kw["customPump"] = np.exp(-tau ** 2 / (2 * w * w)).astype(np.complex128)
    return mod.Chi2PDC(**kw)


def config4_parts(N=1024, nZ=2000):
    """Differentiable Chi2PDC: N=1024, nZ=2000, tMax=20, Gaussian pump (w=1) with a
    seeded spectral phase phi ~ N(0, 0.5^2), per-step poling p ~ U(-1, 1).

    Returns (kwargs, spectrum0, phi, poling): the pump envelope is
    IFFT(spectrum0 * exp(i phi)) with spectrum0 = FFT(exp(-tau^2/2))."""
    r = np.random.default_rng(SEED0 + 4)
    b1s = float(r.uniform(-2, 2))
    b2, b2s = (float(v) for v in r.uniform(-1, 1, size=2))
    phi = r.normal(0.0, 0.5, size=N)
    pol = r.uniform(-1.0, 1.0, size=nZ)
    tMax = 20.0
    tau = 2 * tMax / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    spec0 = np.fft.fft(np.exp(-tau ** 2 / 2).astype(np.complex128))
    kw = dict(relativeLength=1.0, nlLength=1.0, beta2=b2, beta2s=b2s, beta1=0.0, beta1s=b1s, diffBeta0=0.0,
              tMax=tMax, tPrecision=N, nZSteps=nZ)
    return kw, spec0, phi, pol


def config4(mod, N=1024, nZ=2000):
    """The differentiable Chi2PDC medium of config4_parts with its pump and poling set.

    Returns (medium, spectrum0, phi, poling)."""
    kw, spec0, phi, pol = config4_parts(N, nZ)
    env = np.fft.ifft(spec0 * np.exp(1j * phi))
    m = mod.Chi2PDC(customPump=env, polingProfile=list(pol), **kw)
    return m, spec0, phi, pol


def config3_instances(count=2000, seed=SEED0 + 3, sizes=(256, 512, 1024, 2048, 4096), nz_range=(200, 4000)):
    """Heterogeneous batch: `count` Chi2SHG / Chi2PDC media (p = 1/2) with nFreqs drawn
    from `sizes`, nZSteps ~ U{nz_range}, tMax=20, L=1, L_NL=1, beta1 ~ U(-2, 2),
    beta2 ~ U(-1, 1) per mode; the pulse (SHG: the fundamental input; PDC: the pump)
    is a Gaussian or sech (p = 1/2) of width U(0.3, 3) with spectral phase
    a omega + b omega^2, a, b ~ U(-1, 1); the PDC signal input is exp(-tau^2/2).

    Returns a list of (class name, kwargs, time-domain input row)."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        name = "Chi2SHG" if r.uniform() < 0.5 else "Chi2PDC"
        N = int(r.choice(sizes))
        nZ = int(r.integers(nz_range[0], nz_range[1] + 1))
        tMax = 20.0
        tau = 2 * tMax / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
        omega = -np.pi / tMax * np.fft.fftshift(np.arange(-N // 2, N // 2))
        w = r.uniform(0.3, 3.0)
        shape = np.exp(-0.5 * (tau / w) ** 2) if r.uniform() < 0.5 else 1 / np.cosh(tau / w)
        a, b = r.uniform(-1, 1, size=2)
        pulse = np.fft.ifft(np.fft.fft(shape.astype(np.complex128)) * np.exp(1j * (a * omega + b * omega ** 2)))
        b1 = r.uniform(-2, 2, size=2)
        b2 = r.uniform(-1, 1, size=2)
        if name == "Chi2SHG":
            kw = dict(relativeLength=1.0, nlLengthH=1.0, nlLengthP=1.0, beta2h=float(b2[0]), beta2p=float(b2[1]),
                      beta1h=float(b1[0]), beta1p=float(b1[1]), tMax=tMax, tPrecision=N, nZSteps=nZ)
            x = pulse
        else:
            kw = dict(relativeLength=1.0, nlLength=1.0, beta2=float(b2[0]), beta2s=float(b2[1]), beta1=float(b1[0]),
                      beta1s=float(b1[1]), tMax=tMax, tPrecision=N, nZSteps=nZ, customPump=pulse)
            x = np.exp(-tau ** 2 / 2).astype(np.complex128)
        out.append((name, kw, x))
    return out


def config4_adjoint(m, spec0, phi, pol, r=1.5, ckpt_every=0):
    """Loss ||S S^T - r I||_F^2 of a config4 medium (pump phase phi, poling pol
    already set) and its gradient from the device adjoint: (L, dL/dphi, dL/dpol).
    The envelope gradient is chained to the phase through env = IFFT(X),
    X = spec0 exp(i phi)."""
    N = spec0.size
    C, S = m.greensForward(False, True, ckpt_every)
    M = S @ S.T - r * np.eye(N)
    gS = 4.0 * (M @ np.conj(S))
    g_env, g_pol = m.greensAdjoint(np.zeros_like(C), gS)
    X = spec0 * np.exp(1j * phi)
    gX = np.fft.fft(g_env) / N
    g_phi = np.real(np.conj(gX) * 1j * X)
    return float(np.sum(np.abs(M) ** 2)), g_phi, g_pol


def config4_directions(count, N=1024, nZ=2000, seed=SEED0 + 5):
    """Seeded unit directions in (phase, poling) space: a list of (dphi, dpol)."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = r.standard_normal(N + nZ)
        d /= np.linalg.norm(d)
        out.append((d[:N], d[N:]))
    return out


def config4_loss_difference(m, spec0, plus, minus, r=1.5):
    """L(plus) - L(minus) for (phi, pol) pairs on the device forward, evaluated as
    Re sum conj(M+ + M-) (P+ - P-) with P = S S^T, M = P - r I, so the constant
    r I cancels exactly instead of in floating point."""
    def s_at(ph, pl):
        m.setPump(np.fft.ifft(spec0 * np.exp(1j * ph)))
        m.setPolingProfile(pl)
        return m.computeGreensFunction()[1]

    Sp, Sm = s_at(*plus), s_at(*minus)
    Pp, Pm = Sp @ Sp.T, Sm @ Sm.T
    I = np.eye(spec0.size)
    return float(np.real(np.sum(np.conj(Pp - r * I + Pm - r * I) * (Pp - Pm))))


def config4_fd(m, spec0, phi, pol, dphi, dpol, eps):
    """Central difference of the config4 loss along (dphi, dpol) with step eps."""
    return config4_loss_difference(m, spec0, (phi + eps * dphi, pol + eps * dpol),
                                   (phi - eps * dphi, pol - eps * dpol)) / (2 * eps)


def config3_check_sample(insts, random_count=16, seed=SEED0 + 30):
    """Indices of the config-3 instances checked against the oracle: every Chi2PDC
    instance of nFreqs 4096 with more than 3000 z steps (the longest in-kernel pump
    generation and the batch's critical path), the longest instance of every
    (model, nFreqs) class, and `random_count` others drawn uniformly."""
    pick = set()
    longest = {}
    for i, (name, kw, _) in enumerate(insts):
        key = (name, kw["tPrecision"])
        if name == "Chi2PDC" and kw["tPrecision"] == 4096 and kw["nZSteps"] > 3000:
            pick.add(i)
        if key not in longest or kw["nZSteps"] > insts[longest[key]][1]["nZSteps"]:
            longest[key] = i
    pick.update(longest.values())
    r = np.random.default_rng(seed)
    rest = [i for i in range(len(insts)) if i not in pick]
    pick.update(int(i) for i in r.choice(rest, size=min(random_count, len(rest)), replace=False))
    return sorted(pick)


def oracle_outputs(O, insts, indices, workers=None):
    """The oracle's batchSignalSimulation of the given config-3 instances (time-domain
    input), computed on a pool of host threads (the oracle releases the GIL):
    {index: output row}."""
    import concurrent.futures as cf
    import os

    def one(i):
        name, kw, x = insts[i]
        return i, getattr(O, name)(**kw).batchSignalSimulation(x[None, :], inTimeDomain=True)[0]

    workers = workers or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return dict(ex.map(one, indices))
