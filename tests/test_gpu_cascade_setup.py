"""Branches of the set-up and of Cascade that the other parity tests do not reach,
against the oracle (rel. L2 <= 1e-10):

* Cascade with a shared pump: each medium's envelope is the previous medium's
  last pump row (src/Cascade.cpp:75-89);
* Cascade.runSignalSimulation: the last rows of the previous medium's signal
  modes feed the next medium through the mode map (src/Cascade.cpp:91-107);
* Cascade.batchSignalSimulation: each medium's batch output feeds the next one
  (src/Cascade.cpp:138-167; the reference builds the mapped inputs but passes the
  previous outputs on, and so does the engine);
* PulseType.Sinc, the chirp/delay phase of setPump and the GaussianApodization
  profile of the pump (src/_NonlinearMedium.cpp:180-204, 242-247) and of a fully
  nonlinear medium (src/_FullyNonlinearMedium.hpp:40-63).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _pdc(mod, i, **extra):
    kws = [dict(relativeLength=1.0, nlLength=0.8, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5, diffBeta0=0.7,
                tMax=10, tPrecision=128, nZSteps=50, rayleighLength=0.9),
           dict(relativeLength=0.7, nlLength=1.0, beta2=-0.2, beta2s=0.5, beta1=-0.3, beta1s=0.4, diffBeta0=-1.2,
                tMax=10, tPrecision=128, nZSteps=57, rayleighLength=1.4),
           dict(relativeLength=0.5, nlLength=1.0, beta2=0.1, beta2s=0.2, beta3=0.05, diffBeta0=0.4, tMax=10,
                tPrecision=128, nZSteps=41)]
    kw = dict(kws[i], **extra)
    return mod.Chi2PDC(**kw)


def _sfg(mod, i):
    kws = [dict(relativeLength=1.0, nlLength=0.9, nlLengthOrig=1.2, beta2=0.2, beta2s=-0.3, beta2o=0.4, beta1s=0.3,
                beta1o=-0.2, diffBeta0=0.6, tMax=10, tPrecision=128, nZSteps=45),
           dict(relativeLength=0.8, nlLength=1.0, nlLengthOrig=0.7, beta2=-0.1, beta2s=0.3, beta2o=-0.2, beta1s=-0.4,
                beta1o=0.1, diffBeta0=-0.9, tMax=10, tPrecision=128, nZSteps=39)]
    return mod.Chi2SFG(**kws[i])


@pytest.mark.parametrize("share", [True, False])
def test_cascade_shared_pump_greens(nlm, oracle, share):
    dm = [_pdc(nlm, i) for i in range(3)]
    om = [_pdc(oracle, i) for i in range(3)]
    dev = nlm.Cascade(dm, [{0: 0}, {0: 0}], share)
    ref = oracle.Cascade(om, [{0: 0}, {0: 0}], share)
    tau = np.asarray(om[0].tau)
    pump = np.exp(-tau ** 2 / 2) * np.exp(0.3j * tau)
    dev.setPump(pump)
    ref.setPump(pump)
    Cd, Sd = dev.computeGreensFunction()
    Co, So = ref.computeGreensFunction()
    assert rel_l2(Cd, Co) < TOL and rel_l2(Sd, So) < TOL
    # with a shared pump each later medium starts from its predecessor's last pump row
    for d, o in zip(dm, om):
        assert rel_l2(np.asarray(d.pumpTime), np.asarray(o.pumpTime)) < TOL
    if share:  # row 0 of medium 1 carries its own beam factor at z = 0: L = 0.7, zR = 1.4
        prof = 1 / np.sqrt(1 + (0.35 / 1.4) ** 2)
        assert rel_l2(np.asarray(dm[1].pumpTime)[0], prof * np.asarray(dm[0].pumpTime)[-1]) < 1e-14


def test_cascade_run_signal_simulation(nlm, oracle):
    # swap the two modes between the media
    dm = [_sfg(nlm, 0), _sfg(nlm, 1)]
    om = [_sfg(oracle, 0), _sfg(oracle, 1)]
    dev = nlm.Cascade(dm, [{0: 1, 1: 0}], False)
    ref = oracle.Cascade(om, [{0: 1, 1: 0}], False)
    dev.runPumpSimulation()
    ref.runPumpSimulation()
    r = np.random.default_rng(3)
    tau = np.asarray(om[0].tau)
    x = (r.standard_normal(256) + 1j * r.standard_normal(256)) * np.tile(np.exp(-tau ** 2 / 6), 2)
    dev.runSignalSimulation(x, inTimeDomain=True)
    ref.runSignalSimulation(x, inTimeDomain=True)
    for d, o in zip(dm, om):
        for q in range(2):
            assert rel_l2(np.asarray(d.signalFreqs(q)), np.asarray(o.signalFreqs(q))) < TOL
            assert rel_l2(np.asarray(d.signalTimes(q)), np.asarray(o.signalTimes(q))) < TOL


@pytest.mark.parametrize("in_time", [True, False])
def test_cascade_batch_signal_simulation(nlm, oracle, in_time):
    dm = [_sfg(nlm, 0), _sfg(nlm, 1)]
    om = [_sfg(oracle, 0), _sfg(oracle, 1)]
    dev = nlm.Cascade(dm, [{0: 1, 1: 0}], False)
    ref = oracle.Cascade(om, [{0: 1, 1: 0}], False)
    r = np.random.default_rng(4)
    X = (r.standard_normal((5, 256)) + 1j * r.standard_normal((5, 256))) * 0.5
    got = dev.batchSignalSimulation(X, inTimeDomain=in_time)
    exp = ref.batchSignalSimulation(X, inTimeDomain=in_time)
    assert got.shape == exp.shape
    assert rel_l2(got, exp) < TOL
    got1 = dev.batchSignalSimulation(X, inTimeDomain=in_time, useOutput=[0, 1])
    exp1 = ref.batchSignalSimulation(X, inTimeDomain=in_time, useOutput=[0, 1])
    assert rel_l2(got1, exp1) < TOL


@pytest.mark.parametrize("pulse", ["Sinc", "Gaussian", "Sech"])
def test_pump_shapes_with_chirp_and_delay(nlm, oracle, pulse):
    kw = dict(relativeLength=1.0, nlLength=0.9, beta2=0.4, beta2s=-0.3, beta1=0.6, beta1s=0.2, diffBeta0=0.5,
              tMax=12, tPrecision=256, nZSteps=60)
    dev = nlm.Chi2PDC(pulseType=getattr(nlm.PulseType, pulse), chirp=0.7, delay=-0.4, **kw)
    ref = oracle.Chi2PDC(pulseType=getattr(oracle.PulseType, pulse), chirp=0.7, delay=-0.4, **kw)
    Cd, Sd = dev.computeGreensFunction()
    Co, So = ref.computeGreensFunction()
    assert rel_l2(Cd, Co) < TOL and rel_l2(Sd, So) < TOL
    assert rel_l2(np.asarray(dev.pumpTime), np.asarray(ref.pumpTime)) < TOL
    # the setPump(pulseType, chirp, delay) method on an existing medium
    dev.setPump(getattr(nlm.PulseType, pulse), -1.1, 0.8)
    ref.setPump(getattr(oracle.PulseType, pulse), -1.1, 0.8)
    Cd, Sd = dev.computeGreensFunction()
    Co, So = ref.computeGreensFunction()
    assert rel_l2(Cd, Co) < TOL and rel_l2(Sd, So) < TOL


def test_gaussian_apodization(nlm, oracle):
    kw = dict(relativeLength=1.0, nlLength=0.9, beta2=0.4, beta2s=-0.3, beta1=0.6, beta1s=0.2, diffBeta0=0.5,
              tMax=12, tPrecision=128, nZSteps=70, rayleighLength=0.4)
    dev = nlm.Chi2PDC(intensityProfile=nlm.IntensityProfile.GaussianApodization, **kw)
    ref = oracle.Chi2PDC(intensityProfile=oracle.IntensityProfile.GaussianApodization, **kw)
    Cd, Sd = dev.computeGreensFunction()
    Co, So = ref.computeGreensFunction()
    assert rel_l2(Cd, Co) < TOL and rel_l2(Sd, So) < TOL
    assert rel_l2(np.asarray(dev.pumpFreq), np.asarray(ref.pumpFreq)) < TOL
    # fully nonlinear: the coupling scales by exp(-((z - L/2)/zR)^2 / 2)
    kws = dict(relativeLength=1.0, nlLengthH=0.8, nlLengthP=1.3, beta2h=0.2, beta2p=-0.3, beta1h=0.2, diffBeta0=0.8,
               tMax=10, tPrecision=128, nZSteps=90, rayleighLength=0.5)
    dev = nlm.Chi2SHG(intensityProfile=nlm.IntensityProfile.GaussianApodization, **kws)
    ref = oracle.Chi2SHG(intensityProfile=oracle.IntensityProfile.GaussianApodization, **kws)
    r = np.random.default_rng(6)
    X = (r.standard_normal((3, 256)) + 1j * r.standard_normal((3, 256))) * 0.6
    assert rel_l2(dev.batchSignalSimulation(X, inTimeDomain=True), ref.batchSignalSimulation(X, inTimeDomain=True)) < TOL
    # and GNLSE's intensity factor
    kwg = dict(relativeLength=0.3, nlLength=1.0, selfSteepLength=50.0, fr=0.18, fb=0.21, tau1=0.3, tau2=0.6, tau3=1.5,
               beta2=-1.0, tMax=10, tPrecision=128, nZSteps=60, rayleighLength=0.2)
    dev = nlm.Chi3GNLSE(intensityProfile=nlm.IntensityProfile.GaussianApodization, **kwg)
    ref = oracle.Chi3GNLSE(intensityProfile=oracle.IntensityProfile.GaussianApodization, **kwg)
    tau = np.asarray(ref.tau)
    X = np.array([2 / np.cosh(tau) + 0j])
    assert rel_l2(dev.batchSignalSimulation(X, inTimeDomain=True), ref.batchSignalSimulation(X, inTimeDomain=True)) < TOL
