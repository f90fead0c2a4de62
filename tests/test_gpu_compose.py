"""Green's function composition of two media (Cascade::computeGreensFunction,
src/Cascade.cpp:128-129: C = C2 C1 + S2 conj(S1), S = C2 S1 + S2 conj(C1)) on the
FP64 matrix cores, through the C-ABI on device buffers and through
Cascade.computeGreensFunction against the oracle.  Tolerance: relative L2 <= 1e-13
(products of FP64 values summed in a different order)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _compose_dev(C2, S2, C1, S1):
    from multimodenonlinearoptics_amd import capi
    L = capi.lib()
    M, K = C2.shape
    Nc = C1.shape[1]
    dev = [capi.DeviceArray(a) for a in (C2, S2, C1, S1)]
    out = [capi.DeviceArray(shape=(M, Nc)) for _ in range(2)]
    rc = L.nlm_greens_compose_dev(M, K, Nc, *[d.ptr for d in dev + out], None)
    assert rc == 0, L.nlm_last_error()
    assert capi.hip().hipStreamSynchronize(None) == 0
    return out[0].numpy(), out[1].numpy()


@pytest.mark.parametrize("M,K,Nc", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (100, 48, 300), (31, 130, 33),
                                    (512, 512, 512)])
def test_compose_matches_numpy(M, K, Nc):
    print("compose M=%d K=%d Nc=%d" % (M, K, Nc), flush=True)
    r = np.random.default_rng(M * 1000 + K + Nc)

    def rnd(a, b):
        return r.standard_normal((a, b)) + 1j * r.standard_normal((a, b))

    C2, S2, C1, S1 = rnd(M, K), rnd(M, K), rnd(K, Nc), rnd(K, Nc)
    Cg, Sg = _compose_dev(C2, S2, C1, S1)
    assert rel_l2(Cg, C2 @ C1 + S2 @ S1.conj()) < 1e-13
    assert rel_l2(Sg, C2 @ S1 + S2 @ C1.conj()) < 1e-13


def test_compose_host_entry_and_empty_k(nlm):
    from multimodenonlinearoptics_amd import capi
    L = capi.lib()
    r = np.random.default_rng(3)
    M, K, Nc = 40, 24, 56
    C2, S2 = [np.ascontiguousarray(r.standard_normal((M, K)) + 1j * r.standard_normal((M, K))) for _ in range(2)]
    C1, S1 = [np.ascontiguousarray(r.standard_normal((K, Nc)) + 1j * r.standard_normal((K, Nc))) for _ in range(2)]
    Co, So = np.empty((M, Nc), np.complex128), np.empty((M, Nc), np.complex128)
    ptr = [C.c_void_p(a.ctypes.data) for a in (C2, S2, C1, S1, Co, So)]
    assert L.nlm_greens_compose(M, K, Nc, *ptr) == 0
    assert rel_l2(Co, C2 @ C1 + S2 @ S1.conj()) < 1e-13
    assert rel_l2(So, C2 @ S1 + S2 @ C1.conj()) < 1e-13
    Cz, Sz = _compose_dev(np.zeros((M, 0), np.complex128), np.zeros((M, 0), np.complex128),
                          np.zeros((0, Nc), np.complex128), np.zeros((0, Nc), np.complex128))
    assert not Cz.any() and not Sz.any()


def test_cascade_greens_matches_oracle(nlm, oracle):
    N, nZ = 256, 60
    kw_a = dict(relativeLength=1.0, nlLength=0.8, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5, diffBeta0=0.7,
                tMax=10, tPrecision=N, nZSteps=nZ)
    kw_b = dict(relativeLength=0.7, nlLength=1.0, beta2=-0.2, beta2s=0.5, beta1=-0.3, beta1s=0.4, diffBeta0=-1.2,
                tMax=10, tPrecision=N, nZSteps=nZ + 7)
    dev = nlm.Cascade([nlm.Chi2PDC(**kw_a), nlm.Chi2PDC(**kw_b)], [{0: 0}], False)
    ref = oracle.Cascade([oracle.Chi2PDC(**kw_a), oracle.Chi2PDC(**kw_b)], [{0: 0}], False)
    Cd, Sd = dev.computeGreensFunction()
    Co, So = ref.computeGreensFunction()
    assert rel_l2(Cd, Co) < 1e-10
    assert rel_l2(Sd, So) < 1e-10


def test_three_media_cascade_matches_oracle(nlm, oracle):
    N = 128
    kws = [dict(relativeLength=1.0, nlLength=0.8, beta2=0.3, beta2s=-0.4, beta1=0.2, beta1s=-0.5, diffBeta0=0.7,
                tMax=10, tPrecision=N, nZSteps=40),
           dict(relativeLength=0.5, nlLength=1.0, beta2=-0.3, beta2s=0.2, diffBeta0=0.3, tMax=10, tPrecision=N,
                nZSteps=33),
           dict(relativeLength=0.7, nlLength=1.0, beta2=-0.2, beta2s=0.5, beta1=-0.3, beta1s=0.4, diffBeta0=-1.2,
                tMax=10, tPrecision=N, nZSteps=47)]
    dev = nlm.Cascade([nlm.Chi2PDC(**k) for k in kws], [{0: 0}, {0: 0}], False)
    ref = oracle.Cascade([oracle.Chi2PDC(**k) for k in kws], [{0: 0}, {0: 0}], False)
    for in_time in (False, True):
        Cd, Sd = dev.computeGreensFunction(inTimeDomain=in_time)
        Co, So = ref.computeGreensFunction(inTimeDomain=in_time)
        assert rel_l2(Cd, Co) < 1e-10
        assert rel_l2(Sd, So) < 1e-10
