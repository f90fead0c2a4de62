"""The differentiable-poling configuration at full size: Chi2PDC, N=1024,
nZ=2000, loss ||S S^T - 1.5 I||_F^2, gradient w.r.t. the pump spectral phase
(1024 reals) and the per-step poling (2000 reals) from the device adjoint.

Central differences with a step of 1e-6 are limited here by the rounding of S
(the loss is 2304 and nearly constant; a step of 1e-6 moves it by ~1e-12 of
itself), about 1e-4 of the directional derivative.  The references are
therefore (a) torch forward-mode AD of the torch restatement (tests/pdc_torch.py,
pinned to the oracle), exact up to rounding, and (b) Richardson-extrapolated
central differences of the device forward at steps 4e-3 and 2e-3, whose own
error (truncation O(eps^4) plus rounding / eps) is below 1e-7 here.  Both must
agree with the adjoint to 1e-6 of each directional derivative."""
import numpy as np
import pytest

import cases
from cases import config4, config4_adjoint, config4_directions, config4_fd, config4_parts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def adjoint(nlm):
    m, spec0, phi, pol = config4(nlm)
    L, g_phi, g_pol = config4_adjoint(m, spec0, phi, pol)
    return m, spec0, phi, pol, L, g_phi, g_pol


def test_adjoint_matches_forward_mode_derivative(adjoint, tmp_path):
    import json
    import os
    import subprocess
    import sys
    m, spec0, phi, pol, L, g_phi, g_pol = adjoint
    kw = config4_parts()[0]
    dirs = config4_directions(3)
    src, dst = tmp_path / "in.npz", tmp_path / "out.npz"
    np.savez(src, kw=json.dumps(kw), spec0=spec0, phi=phi, pol=pol, dphi=np.array([d[0] for d in dirs]),
             dpol=np.array([d[1] for d in dirs]))
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, os.path.join(here, "pdc_torch.py"), str(src), str(dst)], timeout=240,
                       capture_output=True, text=True)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    res = np.load(dst)
    for k, ((dphi, dpol), Lt, dL) in enumerate(zip(dirs, res["L"], res["dL"])):
        ad = float(np.dot(g_phi, dphi) + np.dot(g_pol, dpol))
        print("direction %d: adjoint %.12e forward-mode %.12e rel %.2e" % (k, ad, dL, abs(ad - dL) / abs(dL)),
              flush=True)
        assert abs(Lt - L) <= 1e-10 * abs(L)
        assert abs(ad - dL) <= 1e-6 * abs(dL), (k, ad, dL)


def test_adjoint_matches_extrapolated_differences(adjoint):
    m, spec0, phi, pol, L, g_phi, g_pol = adjoint
    for k, (dphi, dpol) in enumerate(config4_directions(8)):
        f2 = config4_fd(m, spec0, phi, pol, dphi, dpol, 2e-3)
        f4 = config4_fd(m, spec0, phi, pol, dphi, dpol, 4e-3)
        rich = (4 * f2 - f4) / 3
        ad = float(np.dot(g_phi, dphi) + np.dot(g_pol, dpol))
        print("direction %d: adjoint %.12e extrapolated FD %.12e rel %.2e" % (k, ad, rich, abs(ad - rich) / abs(rich)),
              flush=True)
        assert abs(ad - rich) <= 1e-6 * abs(rich), (k, ad, rich)
