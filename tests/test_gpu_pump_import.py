"""setPump(other, modeIndex, delay) (src/_NonlinearMedium.cpp:506-534): a medium's
pump grid interpolated from another medium's signal grid with the frame phases,
then used for propagation.  Pump grids and Green's functions against the oracle
(rel L2 <= 1e-10)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pair(mod, N, nzA, nzB):
    a = mod.Chi2PDC(relativeLength=1.3, nlLength=1.0, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5, diffBeta0=0.7,
                    tMax=10, tPrecision=N, nZSteps=nzA)
    b = mod.Chi2PDC(relativeLength=0.8, nlLength=1.0, beta2=-0.2, beta2s=0.5, beta1=-0.3, beta1s=0.4,
                    diffBeta0=-1.2, tMax=10, tPrecision=N, nZSteps=nzB)
    return a, b


@pytest.mark.parametrize("N,nzA,nzB,delay", [(128, 60, 41, 0.0), (256, 97, 97, 0.35), (512, 200, 33, -0.2)])
def test_pump_from_signal_grid(nlm, oracle, N, nzA, nzB, delay):
    print("pump import N=%d nzA=%d nzB=%d delay=%g" % (N, nzA, nzB, delay), flush=True)
    r = np.random.default_rng(N + nzA)
    tau = 2 * 10.0 / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    x = (np.exp(-tau ** 2 / 2) * (1 + 0.1 * r.standard_normal(N))).astype(np.complex128)
    res = []
    for mod in (nlm, oracle):
        a, b = _pair(mod, N, nzA, nzB)
        a.runPumpSimulation()
        a.runSignalSimulation(x, True)
        b.setPump(a, 0, delay)
        C, S = b.computeGreensFunction(runPump=False)
        res.append((np.array(b.pumpFreq), np.array(b.pumpTime), C, S))
    for got, exp in zip(res[0], res[1]):
        assert rel_l2(got, exp) < 1e-10
