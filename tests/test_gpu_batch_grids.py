"""z-resolved batches (batchSignalGrids, an extension): runSignalSimulation
(src/_NonlinearMedium.cpp:253-260, src/_NonlinearMedium.hpp:219-238) of every input
row in one launch.  Each row's frequency and time grids against the oracle's
runSignalSimulation of that row alone (rel L2 <= 1e-10)."""
import numpy as np
import pytest

from cases import NMODES, make, random_inputs, small_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,N,rows,in_time,mode", [
    ("Chi2PDC", 64, 3, True, 0),
    ("Chi3", 128, 2, False, 0),
    ("Chi2SHG", 64, 3, True, 0),
    ("Chi2SFG", 128, 2, True, 1),
    ("Chi3GNLSE", 256, 2, True, 0),
])
def test_batch_grids_match_oracle_rows(nlm, oracle, name, N, rows, in_time, mode):
    print("batch grids %s N=%d rows=%d" % (name, N, rows), flush=True)
    nZ = 30
    _, kw = small_case(name, nlm, N=N, nZ=nZ, seed=11)
    _, kwo = small_case(name, oracle, N=N, nZ=nZ, seed=11)
    dev = make(nlm, name, kw)
    ref = make(oracle, name, kwo)
    if name not in ("Chi2SHG", "Chi3GNLSE"):
        dev.runPumpSimulation()
        ref.runPumpSimulation()
    X = random_inputs(name, N, rows, seed=12, nch=1)
    if not in_time:
        X = np.fft.fft(X, axis=1)
    GF, GT = dev.batchSignalGrids(X, in_time, mode)
    assert GF.shape == (rows, NMODES[name], dev.nZSteps, N)
    for r in range(rows):
        ref.runSignalSimulation(X[r], in_time, mode)
        for m in range(NMODES[name]):
            assert rel_l2(GF[r, m], ref.signalFreqs(m)) < 1e-10
            assert rel_l2(GT[r, m], ref.signalTimes(m)) < 1e-10
    # the medium's own grid views hold row 0's propagation
    for m in range(NMODES[name]):
        assert rel_l2(dev.signalFreqs(m), GF[0, m]) == 0.0
