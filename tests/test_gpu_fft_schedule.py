"""FFT schedule variants of the step-table propagation and their autotuner: each
variant (LDS padding, position of the radix-2/4 stage), forced through
NLM_FFT_SCHEDULE, agrees with the oracle, and a Green's launch of at least 128
fields records the autotuner's choice (nlm_fft_schedule)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _kw(N, nZ):
    return dict(relativeLength=1.0, nlLength=0.8, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5, diffBeta0=0.7,
                tMax=20, tPrecision=N, nZSteps=nZ)


@pytest.mark.parametrize("N,nZ", [(512, 30), (1024, 20), (2048, 12), (4096, 9), (8192, 5)])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_schedule_variants_match_oracle(nlm, oracle, monkeypatch, N, nZ, variant):
    print("schedule variant %d, N=%d nZ=%d" % (variant, N, nZ), flush=True)
    monkeypatch.setenv("NLM_FFT_SCHEDULE", str(variant))
    kw = _kw(N, nZ)
    dev = nlm.Chi2PDC(**kw)
    ref = oracle.Chi2PDC(**kw)
    if N <= 1024:
        Cd, Sd = dev.computeGreensFunction()
        Co, So = ref.computeGreensFunction()
        assert rel_l2(Cd, Co) < 1e-10 and rel_l2(Sd, So) < 1e-10
    r = np.random.default_rng(N + variant)
    X = r.standard_normal((3, N)) + 1j * r.standard_normal((3, N))
    got = dev.batchSignalSimulation(X, inTimeDomain=True)
    exp = ref.batchSignalSimulation(X, inTimeDomain=True)
    assert rel_l2(got, exp) < 1e-10


def test_autotuner_records_a_choice(nlm):
    from multimodenonlinearoptics_amd import capi
    L = capi.lib()
    m = nlm.Chi2PDC(**_kw(1024, 60))
    C0, S0 = m.computeGreensFunction()
    v = L.nlm_fft_schedule(1024, 0, 2048)
    assert 0 <= v <= 3
    # later solves reuse the choice and reproduce the first one's outputs exactly
    C1, S1 = m.computeGreensFunction()
    assert np.array_equal(C0, C1) and np.array_equal(S0, S1)
    assert L.nlm_fft_schedule(1024, 0, 2048) == v
    # N = 512 is not tuned: it keeps schedule 0
    m512 = nlm.Chi2PDC(**_kw(512, 60))
    m512.computeGreensFunction()
    assert L.nlm_fft_schedule(512, 0, 1024) == -1



@pytest.mark.parametrize("N,nZ", [(1024, 40), (2048, 50)])
def test_adjoint_after_each_forward_schedule(nlm, monkeypatch, N, nZ):
    """The Green's adjoint behind a forward run with each schedule variant of the
    table kernel (the sweep keeps its default schedule): the gradients agree."""
    from cases import SEED0
    from test_gpu_adjoint import _kwargs
    kw, env, pol = _kwargs(N, nZ, SEED0 + 11 * N)
    r = np.random.default_rng(2)
    Wc = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    Ws = r.standard_normal((N, N)) + 1j * r.standard_normal((N, N))
    res = []
    for v in range(4):
        print("forward schedule variant %d, N=%d" % (v, N), flush=True)
        monkeypatch.setenv("NLM_FFT_SCHEDULE", str(v))
        m = nlm.Chi2PDC(customPump=env, **kw)
        m.setPolingProfile(pol)
        m.greensForward(False, True, 0)
        res.append(m.greensAdjoint(Wc, Ws))
    for ge, gp in res[1:]:
        assert rel_l2(ge, res[0][0]) < 1e-11 and rel_l2(gp, res[0][1]) < 1e-11
