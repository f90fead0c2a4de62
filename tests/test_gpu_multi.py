"""Heterogeneous batches (MultiBatch, nlm_multi_*): every instance against the
oracle's batchSignalSimulation of the same medium and input (rel. L2 <= 1e-10),
and against the engine's own per-medium path; every solver has a batch kernel."""
import numpy as np
import pytest

from cases import config3_instances
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _build(mod, insts):
    return [getattr(mod, name)(**kw) for name, kw, _ in insts]


def test_mixed_batch_matches_oracle(nlm, oracle):
    insts = config3_instances(count=14, seed=77, sizes=(32, 64, 128, 256, 512), nz_range=(20, 90))
    media = _build(nlm, insts)
    ref = _build(oracle, insts)
    xs = [x for _, _, x in insts]
    b = nlm.MultiBatch(media, xs, inTimeDomain=True)
    assert len(b) == len(insts)
    ms = b.run()
    assert ms > 0
    for i, (m, r, x) in enumerate(zip(media, ref, xs)):
        got = b.output(i)
        exp = r.batchSignalSimulation(x[None, :], inTimeDomain=True)[0]
        assert got.shape == exp.shape
        assert rel_l2(got, exp) < TOL, (i, insts[i][0], insts[i][1]["tPrecision"], rel_l2(got, exp))
        own = m.batchSignalSimulation(x[None, :], inTimeDomain=True)[0]
        assert rel_l2(got, own) < 1e-12
    # a second launch of the captured graph gives the same outputs
    first = [b.output(i) for i in range(len(b))]
    b.run()
    for i in range(len(b)):
        assert np.array_equal(first[i], b.output(i))
    cls = b.classes()
    assert sum(c[2] for c in cls) == len(insts)
    assert all(1 <= c[3] <= c[2] for c in cls)


def test_mixed_batch_frequency_domain_and_sizes(nlm, oracle):
    insts = config3_instances(count=6, seed=5, sizes=(1024, 2048, 4096), nz_range=(6, 12))
    media = _build(nlm, insts)
    ref = _build(oracle, insts)
    xs = [np.fft.fft(x) for _, _, x in insts]
    b = nlm.MultiBatch(media, xs, inTimeDomain=False, autotune=False)
    b.run()
    for i, (r, x) in enumerate(zip(ref, xs)):
        exp = r.batchSignalSimulation(x[None, :], inTimeDomain=False)[0]
        assert rel_l2(b.output(i), exp) < TOL, i


@pytest.mark.parametrize("stage", [True, False])
def test_large_pdc_class_matches_oracle(nlm, oracle, stage):
    """Chi2PDC at N = 4096 (512 lanes per field) with instances of 3, 70 and 150
    steps and a Chi2SHG medium of the same size in the same batch: the pump rows
    staged in a device buffer at the start of the launch (chunks of 32 rows, the
    last ones partial), or generated in the kernel (the phase re-evaluated
    exactly every 64 rows)."""
    r = np.random.default_rng(41)
    kws = [dict(relativeLength=1.0, nlLength=0.5 + 0.3 * j, beta2=0.2 * j - 0.1, beta2s=-0.3, beta1=0.1 * j,
                diffBeta0=0.5 - j, tMax=20, tPrecision=4096, nZSteps=nz) for j, nz in enumerate((3, 70, 150))]
    insts = [("Chi2PDC", kw) for kw in kws]
    insts.append(("Chi2SHG", dict(relativeLength=1.0, nlLengthH=2.0, nlLengthP=1.0, beta2h=-0.2, beta2p=0.1,
                                  tMax=20, tPrecision=4096, nZSteps=40)))
    media = [getattr(nlm, n)(**kw) for n, kw in insts]
    ref = [getattr(oracle, n)(**kw) for n, kw in insts]
    cols = [kw["tPrecision"] for _, kw in insts]
    xs = [r.standard_normal(c) + 1j * r.standard_normal(c) for c in cols]
    b = nlm.MultiBatch(media, xs, inTimeDomain=True, autotune=False, stagePumpRows=stage)
    b.run()
    for i, (rm, x) in enumerate(zip(ref, xs)):
        print("large-N instance", i, insts[i][0], insts[i][1]["nZSteps"], "staged", stage, flush=True)
        exp = rm.batchSignalSimulation(x[None, :], inTimeDomain=True)[0]
        assert rel_l2(b.output(i), exp) < TOL, (i, rel_l2(b.output(i), exp))
    first = [b.output(i) for i in range(len(b))]
    b.run()
    for i in range(len(b)):
        assert np.array_equal(first[i], b.output(i))


def test_mixed_batch_errors(nlm):
    big = nlm.Chi3(relativeLength=1, nlLength=1, beta2=-1, tMax=10, tPrecision=8192, nZSteps=10)
    with pytest.raises(ValueError, match="power of two, 32..4096"):
        nlm.MultiBatch([big], [np.zeros(8192, complex)])
    pdc = nlm.Chi2PDC(relativeLength=1, nlLength=1, beta2=0.1, beta2s=0.1, tMax=10, tPrecision=64, nZSteps=10)
    with pytest.raises(ValueError, match="Signals not of correct length"):
        nlm.MultiBatch([pdc], [np.zeros(60, complex)])
    with pytest.raises(ValueError):
        nlm.MultiBatch([pdc, pdc], [np.zeros(64, complex)])


def test_config3_workload_matches_oracle(nlm, oracle):
    """The heterogeneous-batch configuration at its workload: 2000 Chi2SHG /
    Chi2PDC media (nFreqs 256..4096, nZSteps 200..4000) in one graph launch, as the
    benchmark runs it.  Checked against the oracle on every Chi2PDC medium of
    nFreqs 4096 with more than 3000 steps (in-kernel pump rows over up to 8000 rows),
    the longest medium of every class and 16 others."""
    from cases import config3_check_sample, oracle_outputs
    insts = config3_instances(2000)
    idx = config3_check_sample(insts)
    print("config 3: %d media, %d checked (longest nZ %d)" % (len(insts), len(idx),
                                                              max(insts[i][1]["nZSteps"] for i in idx)), flush=True)
    media = _build(nlm, insts)
    b = nlm.MultiBatch(media, [x for _, _, x in insts], inTimeDomain=True, autotune=True)
    b.run()
    exp = oracle_outputs(oracle, insts, idx)
    worst = 0.0
    for i in idx:
        err = rel_l2(b.output(i), exp[i])
        worst = max(worst, err)
        assert err < TOL, (i, insts[i][0], insts[i][1]["tPrecision"], insts[i][1]["nZSteps"], err)
    print("config 3: worst rel. L2 %.2e" % worst, flush=True)


def test_every_solver_in_one_batch(nlm, oracle):
    """All 15 solvers (two sizes each) in one heterogeneous batch: single linear
    pumps generated in the kernel, Chi3's nonlinear pump and Chi2SFGOPA's two pumps
    from the media's device grids, Chi3GNLSE with its in-stage transforms."""
    from cases import ALL, make, random_inputs, small_case
    media, ref, xs, names = [], [], [], []
    for j, name in enumerate(ALL):
        for N, nZ in ((32, 17 + j), (128, 40 - j)):
            _, kw = small_case(name, nlm, N=N, nZ=nZ, seed=100 + j)
            _, kwo = small_case(name, oracle, N=N, nZ=nZ, seed=100 + j)
            media.append(make(nlm, name, kw))
            ref.append(make(oracle, name, kwo))
            xs.append(random_inputs(name, N, 1, seed=200 + j)[0])
            names.append((name, N))
    b = nlm.MultiBatch(media, xs, inTimeDomain=True, autotune=False)
    b.run()
    for i, (r, x) in enumerate(zip(ref, xs)):
        print("instance", i, names[i], flush=True)
        exp = r.batchSignalSimulation(x[None, :], inTimeDomain=True)[0]
        got = b.output(i)
        assert got.shape == exp.shape
        assert rel_l2(got, exp) < TOL, (names[i], rel_l2(got, exp))
    # classes: one per (kernel, size, generated pump); Chi2AFC shares Chi2SFG's, Chi2ASHG Chi2SHG's
    assert sum(c[2] for c in b.classes()) == len(media)


def test_batch_needs_power_of_two_lengths(nlm):
    from cases import make, random_inputs, small_case
    _, kw = small_case("Chi2PDC", nlm, N=48, nZ=10, seed=1)
    m = make(nlm, "Chi2PDC", kw)
    with pytest.raises(ValueError, match="power of two"):
        nlm.MultiBatch([m], [random_inputs("Chi2PDC", 48, 1, seed=1)[0]], inTimeDomain=True)
