"""PyTorch autograd over the MI355X engine's adjoints.

``GreensFunction.apply(medium, envelope, poling)`` returns the Green's
matrices (C, S) of a single-mode Chi2PDC medium as complex128 tensors, for a
time-domain pump envelope (complex, length nFreqs) and a per-z-step poling
profile (real, length nZSteps).  Its backward runs the device adjoint (forward
checkpoints of the field, every step when they fit in a quarter of the device
memory, else every few steps with the states between them rebuilt) and returns
dL/d envelope and dL/d poling in torch's complex-gradient convention.

The forward replaces computeGreensFunction (src/_NonlinearMedium.cpp:263-375)
for a Chi2PDC medium (solvers/Chi2PDC.cpp:27-47) after setPump(customPump)
(src/_NonlinearMedium.cpp:207-222) and a per-step poling.

``BatchSignal.apply(medium, inputs, inTimeDomain)`` is batchSignalSimulation of a
Chi3GNLSE medium (src/_FullyNonlinearMedium.hpp:8-11, solvers/Chi3GNLSE.cpp:39-92)
for a [rows, nFreqs] complex128 tensor; its backward runs the device adjoint
(checkpointed forward, segment recomputation) and returns dL/d inputs.
"""
import numpy as np
import torch


def _check_serial(ctx):
    # the medium keeps the checkpoints of its latest forward only
    if ctx.medium.forwardSerial != ctx.serial:
        raise RuntimeError("the medium ran another differentiable forward after this one; its checkpoints are "
                           "gone: run backward before the next forward on the same medium, or use one medium "
                           "per graph")


class GreensFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, medium, envelope, poling, ckpt_every=0):
        env = envelope.detach().cpu().numpy().astype(np.complex128)
        pol = poling.detach().cpu().numpy().astype(np.float64)
        medium.setPump(env)
        medium.setPolingProfile(pol)
        C, S = medium.greensForward(False, True, int(ckpt_every))
        ctx.medium = medium
        ctx.serial = medium.forwardSerial
        ctx.device = envelope.device
        return torch.from_numpy(C).to(envelope.device), torch.from_numpy(S).to(envelope.device)

    @staticmethod
    def backward(ctx, gC, gS):
        _check_serial(ctx)
        gC = np.zeros(0) if gC is None else gC.detach().cpu().numpy()
        gS = np.zeros(0) if gS is None else gS.detach().cpu().numpy()
        n = ctx.medium.signalFreq.shape[1]
        if gC.size == 0:
            gC = np.zeros((n, n), dtype=np.complex128)
        if gS.size == 0:
            gS = np.zeros((n, n), dtype=np.complex128)
        g_env, g_pol = ctx.medium.greensAdjoint(gC.astype(np.complex128), gS.astype(np.complex128))
        return (None, torch.from_numpy(g_env).to(ctx.device), torch.from_numpy(g_pol).to(ctx.device), None)


class BatchSignal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, medium, inputs, inTimeDomain=True, ckpt_every=0):
        x = inputs.detach().cpu().numpy().astype(np.complex128)
        out = medium.batchForward(x, bool(inTimeDomain), int(ckpt_every))
        ctx.medium = medium
        ctx.serial = medium.forwardSerial
        ctx.device = inputs.device
        return torch.from_numpy(out).to(inputs.device)

    @staticmethod
    def backward(ctx, g):
        _check_serial(ctx)
        gin = ctx.medium.batchAdjoint(g.detach().cpu().numpy().astype(np.complex128))
        return None, torch.from_numpy(gin).to(ctx.device), None, None


def band_energy_loss(out_time, omega, lo=5.0, hi=15.0):
    """Spectral energy of the output rows (time domain) in lo <= omega <= hi."""
    F = torch.fft.fft(out_time, dim=-1)
    band = torch.as_tensor((np.asarray(omega) >= lo) & (np.asarray(omega) <= hi), device=out_time.device)
    return (F.abs() ** 2 * band).sum()


def squeezing_loss(S, r=1.5):
    """||S S^T - r I||_F^2 (the differentiable-poling benchmark loss)."""
    n = S.shape[0]
    M = S @ S.transpose(0, 1) - r * torch.eye(n, dtype=S.dtype, device=S.device)
    return (M.abs() ** 2).sum()
