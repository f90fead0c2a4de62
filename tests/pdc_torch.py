"""Chi2PDC Green's function restated with torch (complex128), as a function of the
pump's spectral phase and the per-step poling.

Used as the exact-derivative reference of the device Green's adjoint: torch's
forward-mode AD (torch.func.jvp) differentiates this forward along a direction
without any of the adjoint's hand-derived formulas and without the rounding
floor of finite differences.  test_pdc_torch_reference pins it to the oracle.

The forward follows, for one signal mode, frequency-domain impulse inputs and
no normalisation:

* runPumpSimulation (src/_NonlinearMedium.cpp:225-250): pumpFreq row r =
  FFT(envelope) exp(i (r dzp) D_p), pumpTime row r = its inverse transform, row 0
  the envelope itself; the beam profile is 1 (Constant, or GaussianBeam with an
  infinite Rayleigh length);
* Chi2PDC::DiffEq (solvers/Chi2PDC.cpp:27-47): k1 = (p[i-1] nl m(i-1)) P[2i-2]
  conj(a), k2 = (p_mid nl m(i-1/2)) P[2i-1] conj(a + k1/2), k3 likewise with k2,
  k4 = (p[i] nl m(i)) P[2i] conj(a + k3), p_mid = (p[i-1] + p[i]) / 2,
  m(x) = exp(i diffBeta0 x dz), nl = i dz / nlLength;
* the optimized branch of signalSimulationTemplate (src/_NonlinearMedium.hpp:
  199-218, 235-238): half dispersion step in, RK4 step and full dispersion step
  per z step, half step out;
* computeGreensFunction (src/_NonlinearMedium.cpp:263-375): C row j =
  (o(e_j) - i o(i e_j)) / 2, S row j = (o(e_j) + i o(i e_j)) / 2, transposed,
  quadrants swapped (fftshift2, :447-467).
"""
import math

import numpy as np
import torch


class PdcGreens:
    """Tables of one Chi2PDC medium (constructor kwargs of the solver) on `device`.

    `spec0` is the pump spectrum before the phase: envelope = IFFT(spec0 exp(i phi))."""

    def __init__(self, kw, spec0, device="cpu"):
        N, nZ = kw["tPrecision"], kw["nZSteps"]
        if kw.get("rayleighLength", math.inf) != math.inf:
            raise ValueError("the restatement covers a unit beam profile only")
        z = kw["relativeLength"]
        self.N, self.nZ = N, nZ
        self.dz = z / (nZ - 1)
        self.dzp = z / (2 * nZ - 2)
        dev = torch.device(device)
        self.device = dev
        tMax = kw.get("tMax", 10.0)
        k = np.fft.fftshift(np.arange(-N // 2, N // 2))
        omega = (-math.pi / tMax) * k
        w = torch.as_tensor(omega, dtype=torch.float64, device=dev)

        def disp(b1, b2, b3):
            return w * (b1 + w * (0.5 * b2 + w * b3 / 6))

        Dp = disp(kw.get("beta1", 0.0), kw["beta2"], kw.get("beta3", 0.0))
        Ds = disp(kw.get("beta1s", 0.0), kw["beta2s"], kw.get("beta3s", 0.0))
        r = torch.arange(2 * nZ - 1, dtype=torch.float64, device=dev)
        self.pump_phase = torch.exp(1j * (r * self.dzp)[:, None] * Dp[None, :])
        self.disp_step = torch.exp(1j * self.dz * Ds)
        self.half_in = torch.exp((0.5j * self.dz) * Ds)
        self.half_out = torch.exp((-0.5j * self.dz) * Ds)
        nl = 1j / kw["nlLength"] * self.dz
        d0 = kw.get("diffBeta0", 0.0)
        i = torch.arange(1, nZ, dtype=torch.float64, device=dev)
        self.c_prev = nl * torch.exp(1j * d0 * ((i - 1) * self.dz))
        self.c_intm = nl * torch.exp(1j * d0 * ((i - 0.5) * self.dz))
        self.c_curr = nl * torch.exp(1j * d0 * (i * self.dz))
        self.spec0 = torch.as_tensor(np.asarray(spec0, dtype=np.complex128), device=dev)

    def envelope(self, phi):
        return torch.fft.ifft(self.spec0 * torch.exp(1j * phi))

    def pump_time(self, env):
        rows = torch.fft.ifft(torch.fft.fft(env)[None, :] * self.pump_phase[1:])
        return torch.cat([env[None, :], rows])

    def greens(self, phi, pol):
        """(C, S) for spectral phase phi [N] and per-step poling pol [nZ] (float64 tensors)."""
        N = self.N
        P = self.pump_time(self.envelope(phi))
        pol = pol.to(torch.complex128)
        pm = 0.5 * (pol[:-1] + pol[1:])
        m1 = (pol[:-1] * self.c_prev)[:, None] * P[0:-1:2]
        m2 = (pm * self.c_intm)[:, None] * P[1::2]
        m3 = (pol[1:] * self.c_curr)[:, None] * P[2::2]
        eye = torch.eye(N, dtype=torch.complex128, device=self.device)
        F = torch.cat([eye, 1j * eye]) * self.half_in[None, :]
        a = torch.fft.ifft(F)
        for i in range(self.nZ - 1):
            k1 = m1[i] * a.conj()
            k2 = m2[i] * (a + 0.5 * k1).conj()
            k3 = m2[i] * (a + 0.5 * k2).conj()
            k4 = m3[i] * (a + k3).conj()
            a = a + (k1 + 2 * k2 + 2 * k3 + k4) * (1.0 / 6.0)
            F = torch.fft.fft(a) * self.disp_step
            a = torch.fft.ifft(F)
        out = F * self.half_out
        o1, o2 = out[:N], out[N:]
        C = (0.5 * o1 - 0.5j * o2).transpose(0, 1)
        S = (0.5 * o1 + 0.5j * o2).transpose(0, 1)
        return torch.fft.fftshift(C, dim=(0, 1)), torch.fft.fftshift(S, dim=(0, 1))


def squeezing_loss(S, r=1.5):
    """||S S^T - r I||_F^2."""
    n = S.shape[0]
    M = S @ S.transpose(0, 1).contiguous() - r * torch.eye(n, dtype=S.dtype, device=S.device)
    return (M.real ** 2 + M.imag ** 2).sum()


def loss_and_directional_derivative(T, phi, pol, dphi, dpol, r=1.5):
    """(L, dL along (dphi, dpol)) of the squeezing loss by forward-mode AD."""
    def f(ph, pl):
        return squeezing_loss(T.greens(ph, pl)[1], r)

    dev = T.device
    args = (torch.as_tensor(phi, dtype=torch.float64, device=dev), torch.as_tensor(pol, dtype=torch.float64, device=dev))
    tans = (torch.as_tensor(dphi, dtype=torch.float64, device=dev), torch.as_tensor(dpol, dtype=torch.float64, device=dev))
    with torch.no_grad():
        L, dL = torch.func.jvp(f, args, tans)
    return float(L), float(dL)


def reverse_gradient(T, phi, pol, r=1.5):
    """(L, dL/dphi, dL/dpol) of the squeezing loss by torch's reverse mode."""
    dev = T.device
    ph = torch.as_tensor(phi, dtype=torch.float64, device=dev).clone().requires_grad_(True)
    pl = torch.as_tensor(pol, dtype=torch.float64, device=dev).clone().requires_grad_(True)
    L = squeezing_loss(T.greens(ph, pl)[1], r)
    L.backward()
    return float(L.detach()), ph.grad.cpu().numpy(), pl.grad.cpu().numpy()


def run_directions(kw, spec0, phi, pol, dirs, device):
    """[(L, dL)] for each (dphi, dpol) in dirs."""
    T = PdcGreens(kw, spec0, device=device)
    out = []
    for k, (dphi, dpol) in enumerate(dirs):
        out.append(loss_and_directional_derivative(T, phi, pol, dphi, dpol))
        print("forward-mode direction %d: L=%.15e dL=%.15e" % (k, out[-1][0], out[-1][1]), flush=True)
    return out


def main(argv):
    """python pdc_torch.py IN.npz OUT.npz: forward-mode directional derivatives on
    the GPU in a process of its own.  (The engine links the system HIP runtime and
    torch its bundled copy; the two do not share a device within one process.)
    IN holds kw (JSON text), spec0, phi, pol, dphi [K, N], dpol [K, nZ]."""
    import json
    d = np.load(argv[1])
    kw = json.loads(str(d["kw"]))
    dirs = list(zip(d["dphi"], d["dpol"]))
    res = run_directions(kw, d["spec0"], d["phi"], d["pol"], dirs, "cuda")
    np.savez(argv[2], L=np.array([r[0] for r in res]), dL=np.array([r[1] for r in res]))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main(sys.argv))
