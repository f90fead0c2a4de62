"""The reference's own call sequence through its import name: the construction
and calls of tests/04 - Green's Function.py:26-46, 76-116 (plots left out), run
verbatim in a fresh interpreter that only has the repository on sys.path, with
the Green's function then compared with the oracle and the notebook's own check
(propagated signal = C x + S x*) asserted."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, rel_l2

pytestmark = pytest.mark.gpu

NOTEBOOK = r'''
import numpy as np

from numpy.fft import fft, ifft, fftshift

from nonlinearmedium import Chi3

nFreqs = 512
fiberS = Chi3(relativeLength=np.pi / 2,
              nlLength=1,
              beta2=-1,
              pulseType=1,
              tPrecision=nFreqs, zPrecision=100)

C, S = fiberS.computeGreensFunction(nThreads=4)

pulse = np.zeros(nFreqs)
pulse[nFreqs//4 : 3*nFreqs//4] = 1 - np.abs(np.arange(-nFreqs // 4, nFreqs // 4)) / (nFreqs // 4)
pulse = fftshift(pulse)
pulseFreq = fft(pulse)

fiberS.runSignalSimulation(1j * pulse)

output = C @ fftshift(1j * pulseFreq) + S @ fftshift((1j * pulseFreq).conj())
outputTime = fftshift(ifft(fftshift(output)))
'''


def test_reference_notebook_sequence(oracle, tmp_path):
    out = tmp_path / "nb.npz"
    script = ("import sys\nsys.path.insert(0, %r)\n" % REPO + NOTEBOOK +
              "\nnp.savez(%r, C=C, S=S, sim=fftshift(fiberS.signalTime[-1]), gf=outputTime,"
              " isinst=isinstance(fiberS, sys.modules['nonlinearmedium']._NonlinearMedium))\n" % str(out))
    r = subprocess.run([sys.executable, "-c", script], cwd=str(tmp_path), timeout=120, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    d = np.load(out)
    assert bool(d["isinst"])
    # the notebook's own check: the propagated signal equals C x + S x*
    assert rel_l2(d["gf"], d["sim"]) < 1e-12
    ref = oracle.Chi3(relativeLength=np.pi / 2, nlLength=1, beta2=-1, pulseType=1, tPrecision=512, zPrecision=100)
    Co, So = ref.computeGreensFunction(nThreads=4)
    assert rel_l2(d["C"], Co) < 1e-10 and rel_l2(d["S"], So) < 1e-10
