"""Chi3GNLSE batches with nFreqs 16384 and 32768 (transforms staged through HBM
as N = P x 4096 four-step FFTs, z loop replayed as a hipGraph) against the
oracle's batchSignalSimulation (rel. L2 <= 1e-10)."""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-10


def config5_kwargs(N=16384, nZ=20000):
    return dict(relativeLength=math.pi / 2, nlLength=1 / 9, selfSteepLength=100.0, fr=0.18, fb=0.21, tau1=0.0122,
                tau2=0.032, tau3=0.096, beta2=-1.0, tMax=50.0, tPrecision=N, nZSteps=nZ)


def inputs(N, rows, seed, tMax=50.0):
    r = np.random.default_rng(seed)
    tau = 2 * tMax / N * np.fft.fftshift(np.arange(-N // 2, N // 2))
    sech = 1 / np.cosh(tau)
    noise = (r.standard_normal((rows, N)) + 1j * r.standard_normal((rows, N))) / math.sqrt(2) / math.sqrt(N)
    return (sech[None, :] + np.fft.ifft(noise * N, axis=1) / N).astype(np.complex128)


@pytest.mark.parametrize("N,nZ,rows,in_time", [(16384, 40, 2, True), (16384, 33, 1, False), (32768, 9, 1, True)])
def test_staged_gnlse_matches_oracle(nlm, oracle, N, nZ, rows, in_time):
    kw = config5_kwargs(N, nZ)
    # a short medium keeps the step size near that of the full configuration
    # (pi/2 over 20000 steps), where RK4 is stable, at a few dozen steps
    kw["relativeLength"] = 0.01
    X = inputs(N, rows, seed=N + nZ)
    if not in_time:
        X = np.fft.fft(X, axis=1)
    dev = nlm.Chi3GNLSE(**kw)
    ref = oracle.Chi3GNLSE(**kw)
    got = dev.batchSignalSimulation(X, inTimeDomain=in_time)
    exp = ref.batchSignalSimulation(X, inTimeDomain=in_time, nThreads=rows)
    assert got.shape == exp.shape
    assert rel_l2(got, exp) < TOL
    # a second run reuses the captured graph
    again = dev.batchSignalSimulation(X, inTimeDomain=in_time)
    assert np.array_equal(got, again)


def test_staged_gnlse_grid_run_is_refused(nlm):
    m = nlm.Chi3GNLSE(**config5_kwargs(16384, 10))
    with pytest.raises(ValueError, match="above 8192"):
        m.runSignalSimulation(np.ones(16384, complex))
