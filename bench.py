"""Benchmark: batched Green's-function solve of a Chi2PDC medium on one MI355X.

Workload (the benchmark definition's Green's configuration): Chi2PDC, nFreqs=512,
tMax=20, nZSteps=1000, seeded Gaussian pump (width ~ U(0.5, 2)) and dispersion
(beta1, beta1s ~ U(-2, 2); beta2, beta2s ~ U(-1, 1); diffBeta0 ~ U(-5, 5)),
uniform poling.  One step = one computeGreensFunction(): the pump grid
(2nZ-1 rows), the 2N impulse propagations of nZ-1 split steps each in one
launch, and the C/S assembly, with C and S left in HBM (the host copy is timed
separately as end_to_end).  Metric: field-steps/s, a field-step being one
propagation advanced by one RK4 step plus one dispersion step (2*512*999 per
solve).

Reported beside it:
  plain         the same kernel run as one propagation per launch, serially
                (sampled on --plain-sample impulses, scaled to the full solve);
  end_to_end    computeGreensFunction() including the copy of C and S to the host;
  roofline      the propagation kernel's achieved FP64 rate (algorithmic flops
                10 N log2 N + 80 N per field-step) against the 78.6 TFLOP/s peak,
                from HIP events on the kernel's own stream;
  cpu_baseline  the C oracle (a restatement of the reference C++ path with a
                kissfft-class FFT and the reference's std::thread fan-out), compiled
                on this host with the reference's release flags, timed on every core
                this process may use (at most 16) and on one thread, median of 5 after
                a warm-up, with the host's CPU model and core counts;
  rel_err       relative L2 difference of (C, S) against that CPU result;
  config1       the plain Chi3 solve (pump and signal z-grids, N=512, nZ=1000);
  hetero        2000 heterogeneous Chi2SHG/Chi2PDC media in one graph launch;
  gnlse         the Chi3GNLSE batch at nFreqs=16384 through HBM-staged transforms;
  gnlse_grad    its input gradient by the adjoint;
  compose       the Cascade Green's composition on the FP64 matrix cores;
  grad          the differentiable Chi2PDC (N=1024, nZ=2000): adjoint cost and its
                directional derivatives against finite differences and torch
                forward-mode AD.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 (vector and matrix) peak
# HBM bytes per Green's launch of the propagation kernel: rocprofv3 FETCH_SIZE and
# WRITE_SIZE, each scaled by the counter/actual ratio of a 256 MiB copy kernel
# (x2 and x1); profiles/k_prop_tab_chi2pdc_greens_n512_nz1000.md
TRAFFIC_BYTES_PER_LAUNCH = 1.40e8
HBM_PEAK_GBS = 8000.0
N_FREQS, NZ = 512, 1000


def flops_per_field_step(n):
    """Algorithmic FP64 flops of one Chi2PDC field-step: two FFTs (5 N log2 N each),
    the RK4 stages (72 N), the dispersion phase (6 N) and the inverse scale (2 N)."""
    return 10 * n * math.log2(n) + 80 * n


def grad_benchmark(nlm, reps=3, n_dirs=32, fd_eps=1e-6, fwd_mode_dirs=4, ckpt_every=0, n=1024, nz=2000):
    """Differentiable Chi2PDC (N=1024, nZ=2000 by default): loss ||S S^T - 1.5 I||_F^2 of the
    Green's function S, gradient w.r.t. the pump spectral phase (1024 reals) and the
    per-step poling (2000 reals) by the adjoint kernel with ~sqrt(nZ) checkpoints.

    Reports the backward time over one forward Green's solve (no checkpoints), both
    as wall time of the calls and as kernel time, and the directional derivative
    along n_dirs seeded unit directions against three references:
      fd_eps       central differences of the device forward with step 1e-6 (their
                   rounding floor here is ~1e-4 of the derivative: the loss is 2304
                   and a step of 1e-6 moves it by ~1e-12 of itself);
      fd_rich      Richardson extrapolation of central differences at steps 4e-3
                   and 2e-3 (truncation O(eps^4), rounding ~1e-13 / eps);
      forward mode torch forward-mode AD of the torch restatement (tests/pdc_torch.py,
                   pinned to the oracle) on fwd_mode_dirs directions, run in a
                   process of its own.
    n_dirs = 0 times the forward and the backward only (other sizes)."""
    import numpy as np
    from cases import config4, config4_adjoint, config4_directions, config4_fd, config4_parts
    m, spec0, phi, pol = config4(nlm, n, nz)
    N, nZ = spec0.size, pol.size
    m.enableTiming(True)
    m.greensOnDevice()
    m.sync()
    t_fwd, k_fwd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        m.greensOnDevice()
        m.sync()
        t_fwd.append(time.perf_counter() - t0)
        k_fwd.append(m.lastTiming()[1])
    t_ck, t_bwd, k_bwd = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        C, S = m.greensForward(False, True, ckpt_every)
        t_ck.append(time.perf_counter() - t0)
        M = S @ S.T - 1.5 * np.eye(N)
        gS = 4.0 * (M @ np.conj(S))
        t0 = time.perf_counter()
        m.greensAdjoint(np.zeros_like(C), gS)
        t_bwd.append(time.perf_counter() - t0)
        tm = m.lastTiming()
        k_bwd.append(tm[1] + tm[2])
    timing = {
        "workload": "Chi2PDC N=%d nZ=%d, loss ||S S^T - 1.5 I||_F^2, grad w.r.t. pump spectral phase "
                    "(%d) and per-step poling (%d)" % (N, nZ, N, nZ),
        "forward_s": min(t_fwd), "forward_ckpt_s": min(t_ck), "backward_s": min(t_bwd),
        "forward_kernel_ms": min(k_fwd), "backward_kernel_ms": min(k_bwd),
        # device work of the backward (sweep, reductions, chain rule) over that of the
        # forward solve; the wall-clock ratio also holds the upload of dL/dC, dL/dS
        "grad_vs_forward": min(k_bwd) / min(k_fwd),
        "grad_vs_forward_wall": min(t_bwd) / min(t_fwd),
    }
    if n_dirs <= 0:
        return timing
    L0, g_phi, g_pol = config4_adjoint(m, spec0, phi, pol, ckpt_every=ckpt_every)
    dirs = config4_directions(n_dirs, N, nZ)
    ads = [float(np.dot(g_phi, dp) + np.dot(g_pol, dq)) for dp, dq in dirs]
    raw, rich = [], []
    for dp, dq in dirs:
        raw.append(config4_fd(m, spec0, phi, pol, dp, dq, fd_eps))
        f2 = config4_fd(m, spec0, phi, pol, dp, dq, 2e-3)
        f4 = config4_fd(m, spec0, phi, pol, dp, dq, 4e-3)
        rich.append((4 * f2 - f4) / 3)
    m.setPump(np.fft.ifft(spec0 * np.exp(1j * phi)))
    m.setPolingProfile(pol)

    def rel(a, b):
        return abs(a - b) / max(abs(b), 1e-300)

    fwd = forward_mode_reference(config4_parts(N, nZ)[0], spec0, phi, pol, dirs[:fwd_mode_dirs]) \
        if fwd_mode_dirs > 0 else None
    out = dict(timing)
    out.update({
        "loss": L0,
        "directions": n_dirs,
        "fd_rich_max_rel_err": max(rel(a, r) for a, r in zip(ads, rich)),
        "fd_rich_steps": [4e-3, 2e-3],
        "fd_eps": fd_eps,
        "fd_eps_max_rel_err": max(rel(a, f) for a, f in zip(ads, raw)),
        "fd_eps_noise_floor": max(rel(f, r) for f, r in zip(raw, rich)),
        "pairs": [[a, r, f] for a, r, f in zip(ads[:6], rich[:6], raw[:6])],
        "grad_norm": float(np.sqrt(np.sum(g_phi ** 2) + np.sum(g_pol ** 2))),
    })
    if fwd is not None:
        out["forward_mode_directions"] = len(fwd)
        out["forward_mode_max_rel_err"] = max(rel(a, d) for a, (_, d) in zip(ads, fwd))
        out["forward_mode_loss_rel_err"] = max(rel(Lt, L0) for Lt, _ in fwd)
    return out


def forward_mode_reference(kw, spec0, phi, pol, dirs, timeout=300):
    """[(L, dL)] along dirs from torch forward-mode AD of tests/pdc_torch.py, on the
    GPU in a child process (torch's bundled HIP runtime and the engine's do not
    share a device within one process); None if that fails."""
    import subprocess
    import tempfile
    import numpy as np
    with tempfile.TemporaryDirectory() as d:
        src, dst = os.path.join(d, "in.npz"), os.path.join(d, "out.npz")
        np.savez(src, kw=json.dumps(kw), spec0=spec0, phi=phi, pol=pol, dphi=np.array([x[0] for x in dirs]),
                 dpol=np.array([x[1] for x in dirs]))
        r = subprocess.run([sys.executable, os.path.join(HERE, "tests", "pdc_torch.py"), src, dst], timeout=timeout,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            print("forward-mode reference failed:\n" + r.stdout[-2000:], file=sys.stderr)
            return None
        res = np.load(dst)
        return list(zip(res["L"].tolist(), res["dL"].tolist()))


def flops_shg(n):
    """Algorithmic FP64 flops of one Chi2SHG field-step (two modes)."""
    return 2 * (10 * n * math.log2(n) + 8 * n) + 154 * n


def hetero_benchmark(nlm, count=2000, reps=3, plain_sample=40, cpu_sample=8, cpu_cores=1):
    """Heterogeneous batch: `count` Chi2SHG / Chi2PDC media of mixed nFreqs (256..4096)
    and nZSteps (200..4000), one input row each, run as one MultiBatch graph launch.

    Reports field-steps/s of the whole batch (inputs resident in HBM, device time of
    the graph launch), the same instances run one by one through
    batchSignalSimulation (sampled, scaled by algorithmic cost), and the C oracle on
    the cpu_sample costliest instances (N = 4096, nZ near 4000) plus cpu_sample random
    ones, on cpu_cores host threads: its throughput and the worst rel. L2 error of
    those outputs against it."""
    import numpy as np
    from cases import config3_instances
    from oracle import oracle as O
    insts = config3_instances(count)
    t0 = time.perf_counter()
    media = [getattr(nlm, name)(**kw) for name, kw, _ in insts]
    xs = [x for _, _, x in insts]
    b = nlm.MultiBatch(media, xs, inTimeDomain=True, autotune=True)
    t_create = time.perf_counter() - t0
    b.run()
    dev_ms, wall = [], []
    for _ in range(reps):
        w0 = time.perf_counter()
        dev_ms.append(b.run())
        wall.append(time.perf_counter() - w0)
    steps = [kw["nZSteps"] - 1 for _, kw, _ in insts]
    cost = [(flops_shg(kw["tPrecision"]) if name == "Chi2SHG" else flops_per_field_step(kw["tPrecision"])) * s
            for (name, kw, _), s in zip(insts, steps)]
    total_steps, total_flops = sum(steps), sum(cost)
    ms = min(dev_ms)
    # plain: one instance per call, a uniform sample scaled by algorithmic cost
    r = np.random.default_rng(1)
    idx = sorted(r.choice(count, size=min(plain_sample, count), replace=False).tolist())
    media[idx[0]].batchSignalSimulation(xs[idx[0]][None, :], inTimeDomain=True)
    p0 = time.perf_counter()
    for i in idx:
        media[i].batchSignalSimulation(xs[i][None, :], inTimeDomain=True)
    plain_s = (time.perf_counter() - p0) * total_flops / sum(cost[i] for i in idx)
    # oracle on the costliest instances and a random sample: parity and CPU throughput
    from cases import oracle_outputs
    order = sorted(range(count), key=lambda i: -cost[i])
    pick = sorted(set(order[:cpu_sample]) | set(r.choice(count, size=cpu_sample, replace=False).tolist()))
    c0 = time.perf_counter()
    exp = oracle_outputs(O, insts, pick, workers=cpu_cores)
    cpu_t = time.perf_counter() - c0
    cpu_steps = sum(steps[i] for i in pick)
    worst = max(float(np.linalg.norm(b.output(i) - exp[i]) / np.linalg.norm(exp[i])) for i in pick)
    classes = [list(c) for c in b.classes()]
    return {
        "workload": "%d Chi2SHG/Chi2PDC instances, nFreqs in {256..4096}, nZSteps U[200,4000], one graph launch"
                    % count,
        "instances": count, "field_steps": total_steps, "algorithmic_flops": total_flops,
        "ms": ms, "wall_ms": min(wall) * 1e3, "create_s": t_create,
        "value": total_steps / (ms * 1e-3), "unit": "field-steps/s",
        "tflops": total_flops / (ms * 1e-3) / 1e12,
        "plain_s": plain_s, "plain_sample": len(idx), "batched_vs_plain": plain_s / (ms * 1e-3),
        "cpu": {"field_steps_per_s": cpu_steps / cpu_t, "cores": cpu_cores, "kind": "port",
                "sample": "%d instances (the %d costliest and %d random) on %d threads, %.1f s"
                          % (len(pick), cpu_sample, cpu_sample, cpu_cores, cpu_t)},
        "rel_err_sample": len(pick),
        "rel_err_max": worst, "classes": classes,
    }


def flops_gnlse(n):
    """Algorithmic FP64 flops of one Chi3GNLSE field-step (16 RK transforms + 2)."""
    return 90 * n * math.log2(n) + 154 * n


def gnlse_benchmark(nlm, rows=64, nz=20000, n=16384, cpu_steps=6):
    """Chi3GNLSE, nFreqs=16384, tMax=50, nZSteps=20000, batch of 64 sech pulses
    (soliton order 3) with seeded one-photon noise spectra: one batchSignalSimulation,
    transforms staged through HBM.  The oracle times a few steps of one field.
    A smaller nz shortens the medium in proportion (same step size dz), which
    keeps the run stable for profiling."""
    import numpy as np
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from test_gpu_big import config5_kwargs, inputs
    kw = config5_kwargs(n, nz)
    kw["relativeLength"] *= (nz - 1) / 19999.0
    X = inputs(n, rows, seed=20261202 + 5)
    m = nlm.Chi3GNLSE(**kw)
    m.enableTiming(True)
    t0 = time.perf_counter()
    m.batchSignalSimulation(X, inTimeDomain=True)
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    Y = m.batchSignalSimulation(X, inTimeDomain=True)
    wall = time.perf_counter() - t0
    dev_ms = m.lastTiming()[1]
    steps = rows * (nz - 1)
    kwc = dict(kw, nZSteps=cpu_steps + 1, relativeLength=kw["relativeLength"] * cpu_steps / (nz - 1))
    ref = O.Chi3GNLSE(**kwc)
    c0 = time.perf_counter()
    ref.batchSignalSimulation(X[:1], inTimeDomain=True)
    cpu_s = time.perf_counter() - c0
    return {
        "workload": "Chi3GNLSE nFreqs=%d tMax=50 nZSteps=%d, batch of %d sech (N=3) + noise" % (n, nz, rows),
        "field_steps": steps, "device_ms": dev_ms, "wall_s": wall, "first_call_s": t_first,
        "value": steps / (dev_ms * 1e-3), "unit": "field-steps/s",
        "tflops": flops_gnlse(n) * steps / (dev_ms * 1e-3) / 1e12,
        "finite": bool(np.isfinite(Y).all()),
        "cpu": {"field_steps_per_s": cpu_steps / cpu_s, "cores": 1, "kind": "port",
                "sample": "%d steps of one field" % cpu_steps},
    }


def gnlse_grad_benchmark(nlm, rows=64, nz=20000, n=16384, fwd_ms=None, fd_dirs=2, fd_eps=1e-6, fd_nz=2001,
                         fd_eps_full=1e-8):
    """Gradient of the spectral energy in omega in [5, 15] of the Chi3GNLSE batch
    above w.r.t. its 64 input fields: checkpointed forward (batchForward), then
    the adjoint (batchAdjoint), timed at the full length.  Central differences
    check it twice: at the full length with a step of 1e-8 (after the soliton
    fission the loss is far from linear over 1e-6: that step gives even the wrong
    sign), and on the first fd_nz steps (same step size) with 1e-6."""
    import numpy as np
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from test_gpu_big import config5_kwargs, inputs
    X = inputs(n, rows, seed=20261202 + 5)

    def medium(steps):
        kw = config5_kwargs(n, steps)
        kw["relativeLength"] *= (steps - 1) / 19999.0
        m = nlm.Chi3GNLSE(**kw)
        m.enableTiming(True)
        return m

    def loss_of(m, Y):
        band = (np.asarray(m.omega) >= 5) & (np.asarray(m.omega) <= 15)
        F = np.fft.fft(Y, axis=1)
        return float(np.sum(np.abs(F[:, band]) ** 2)), n * np.fft.ifft(2 * F * band[None, :], axis=1)

    m = medium(nz)
    t0 = time.perf_counter()
    Y = m.batchForward(X, inTimeDomain=True)
    t_fwd = time.perf_counter() - t0
    fwd_ckpt_ms = m.lastTiming()[1]
    L0, gY = loss_of(m, Y)
    t0 = time.perf_counter()
    g = m.batchAdjoint(gY)
    t_bwd = time.perf_counter() - t0
    bwd_ms = m.lastTiming()[1]
    gn = float(np.linalg.norm(g))
    r = np.random.default_rng(20261202 + 6)
    full_pairs = []
    for _ in range(fd_dirs):
        d = r.standard_normal(X.shape) + 1j * r.standard_normal(X.shape)
        d /= np.linalg.norm(d)
        lp, _ = loss_of(m, m.batchSignalSimulation(X + fd_eps_full * d, inTimeDomain=True))
        lm, _ = loss_of(m, m.batchSignalSimulation(X - fd_eps_full * d, inTimeDomain=True))
        full_pairs.append([(lp - lm) / (2 * fd_eps_full), float(np.real(np.vdot(g, d)))])
    del m

    mf = medium(fd_nz)
    _, gYf = loss_of(mf, mf.batchForward(X, inTimeDomain=True))
    gf = mf.batchAdjoint(gYf)
    gfn = float(np.linalg.norm(gf))
    pairs = []
    for _ in range(fd_dirs):
        d = r.standard_normal(X.shape) + 1j * r.standard_normal(X.shape)
        d /= np.linalg.norm(d)
        lp, _ = loss_of(mf, mf.batchSignalSimulation(X + fd_eps * d, inTimeDomain=True))
        lm, _ = loss_of(mf, mf.batchSignalSimulation(X - fd_eps * d, inTimeDomain=True))
        pairs.append([(lp - lm) / (2 * fd_eps), float(np.real(np.vdot(gf, d)))])
    base = fwd_ms if fwd_ms else fwd_ckpt_ms
    return {
        "workload": "d(spectral energy in omega in [5,15])/d(inputs), Chi3GNLSE nFreqs=%d nZSteps=%d, batch %d"
                    % (n, nz, rows),
        "loss": L0, "forward_ckpt_ms": fwd_ckpt_ms, "backward_ms": bwd_ms, "forward_ckpt_s": t_fwd,
        "backward_s": t_bwd, "grad_vs_forward": bwd_ms / base, "grad_norm": gn,
        "fd_full_eps": fd_eps_full, "fd_full_pairs": full_pairs,
        "fd_full_max_rel_err": max(abs(a - b) / max(abs(a), 1e-300) for a, b in full_pairs),
        "fd_nz": fd_nz, "fd_eps": fd_eps, "fd_pairs": pairs, "fd_grad_norm": gfn,
        "fd_max_err_over_grad_norm": max(abs(a - b) for a, b in pairs) / gfn if gfn > 0 else None,
        "finite": bool(np.isfinite(g).all()),
    }


def compose_benchmark(n=512, reps=20):
    """Green's composition of two media (Cascade::computeGreensFunction,
    src/Cascade.cpp:128-129) at n x n (two 512-point Chi2PDC stages): the FP64
    matrix-core kernel on device buffers, timed with HIP events on the stream it
    runs on."""
    import ctypes as C
    import numpy as np
    from multimodenonlinearoptics_amd import capi
    L, H = capi.lib(), capi.hip()
    r = np.random.default_rng(20261202 + 7)
    host = [r.standard_normal((n, n)) + 1j * r.standard_normal((n, n)) for _ in range(4)]
    dev = [capi.DeviceArray(a) for a in host]
    out = [capi.DeviceArray(shape=(n, n)) for _ in range(2)]
    stream, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p()
    H.hipStreamCreate(C.byref(stream))
    H.hipEventCreate(C.byref(e0))
    H.hipEventCreate(C.byref(e1))
    args = [d.ptr for d in dev + out] + [stream]

    def run():
        if L.nlm_greens_compose_dev(n, n, n, *args) != 0:
            raise RuntimeError(L.nlm_last_error().decode())

    run()
    H.hipStreamSynchronize(stream)
    H.hipEventRecord(e0, stream)
    for _ in range(reps):
        run()
    H.hipEventRecord(e1, stream)
    H.hipEventSynchronize(e1)
    ms_total = C.c_float()
    H.hipEventElapsedTime(C.byref(ms_total), e0, e1)
    ms = ms_total.value / reps
    H.hipEventDestroy(e0)
    H.hipEventDestroy(e1)
    H.hipStreamDestroy(stream)
    C2, S2, C1, S1 = host
    ref = C2 @ C1 + S2 @ S1.conj()
    err = float(np.linalg.norm(out[0].numpy() - ref) / np.linalg.norm(ref))
    flops = 32.0 * n ** 3  # two complex products per output matrix: 4 x 8 n^3
    return {"workload": "C = C2 C1 + S2 conj(S1), S = C2 S1 + S2 conj(C1), %d x %d complex" % (n, n),
            "kernel_ms": ms, "tflops": flops / (ms * 1e-3) / 1e12,
            "frac_fp64_peak": flops / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, "rel_err": err}


def host_cpu():
    """The host's CPU model and count, and the cores this process may use: its
    affinity mask, capped by a cgroup CPU quota and by 16 cores (the host share of
    one GPU on a shared multi-GPU node)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    cores = min(aff, max(1, int(quota)) if quota else aff, 16)
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "cores": cores}


def median_time(fn, reps=5):
    """Median wall time of fn() over reps runs after one warm-up run."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def config1_benchmark(nlm, O, reps=5):
    """Plain Chi3 (NLSE) solve: N=1 soliton pump, nFreqs=512, tMax=10, zPrecision 637
    (nZSteps 1000), runPumpSimulation (1998 serial nonlinear pump steps) then
    runSignalSimulation of a seeded Gaussian signal (999 steps), the full pump
    (1999 x 512) and signal (1000 x 512) grids compared with the oracle.  Times are
    wall clock of the two calls (grids left on the device) and with the four grids
    read back; the oracle runs the same calls on one thread (the path is serial)."""
    import numpy as np
    from cases import config1
    fiber, sig = config1(nlm)
    steps = (2 * fiber.nZSteps - 2) + (fiber.nZSteps - 1)

    def run():
        fiber.runPumpSimulation()
        fiber.runSignalSimulation(sig)
        fiber.sync()

    def run_read():
        run()
        return [np.asarray(g) for g in (fiber.pumpTime, fiber.pumpFreq, fiber.signalTime, fiber.signalFreq)]

    t_dev = median_time(run, reps)
    t_e2e = median_time(run_read, reps)
    got = run_read()
    ref, rsig = config1(O)

    def run_ref():
        ref.runPumpSimulation()
        ref.runSignalSimulation(rsig)

    t_cpu = median_time(run_ref, reps)
    exp = [np.asarray(g) for g in (ref.pumpTime, ref.pumpFreq, ref.signalTime, ref.signalFreq)]
    errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(got, exp)]
    return {"workload": "Chi3 soliton pump + Gaussian signal, nFreqs=512, nZSteps=1000, full z-resolved grids",
            "field_steps": steps, "s": t_dev, "value": steps / t_dev, "unit": "field-steps/s",
            "s_with_grid_download": t_e2e, "cpu": {"s": t_cpu, "field_steps_per_s": steps / t_cpu, "cores": 1,
                                                   "kind": "port", "sample": "the whole run, median of 5"},
            "vs_cpu": t_cpu / t_dev,
            "rel_err": {"pumpTime": errs[0], "pumpFreq": errs[1], "signalTime": errs[2], "signalFreq": errs[3]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the GPU's clocks rise over the first ~20 ms of work (2.36 -> 2.03 ms per
    # headline launch over its first ten launches): 20 untimed solves, ~45 ms
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity check")
    ap.add_argument("--no-grad", action="store_true", help="skip the adjoint (gradient) measurement")
    ap.add_argument("--no-hetero", action="store_true", help="skip the heterogeneous-batch measurement")
    ap.add_argument("--hetero-count", type=int, default=2000)
    ap.add_argument("--hetero-plain", type=int, default=40,
                    help="instances of the batch run one by one for the plain comparison (--only hetero)")
    ap.add_argument("--hetero-cpu", type=int, default=8,
                    help="costliest and random instances checked against the oracle (--only hetero)")
    ap.add_argument("--no-gnlse", action="store_true", help="skip the nFreqs=16384 Chi3GNLSE measurement")
    ap.add_argument("--no-gnlse-grad", action="store_true", help="skip the Chi3GNLSE input-gradient measurement")
    ap.add_argument("--no-compose", action="store_true", help="skip the Green's composition measurement")
    ap.add_argument("--no-config1", action="store_true", help="skip the plain Chi3 solve measurement")
    ap.add_argument("--only", choices=["hetero", "gnlse", "grad", "gnlse_grad", "compose", "config1"], default=None,
                    help="run only this section and print its record (for profiling)")
    ap.add_argument("--gnlse-nz", type=int, default=20000,
                    help="nZSteps of the Chi3GNLSE section (the medium is shortened in proportion)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads for the CPU baseline (default: every core this process may use, at most 16)")
    ap.add_argument("--grad-dirs", type=int, default=32,
                    help="seeded directions of the config-4 finite-difference check")
    ap.add_argument("--grad-n", type=int, default=1024, help="nFreqs of the config-4 medium (--only grad)")
    ap.add_argument("--grad-nz", type=int, default=2000, help="nZSteps of the config-4 medium (--only grad)")
    ap.add_argument("--ckpt-every", type=int, default=0,
                    help="checkpoint interval of the config-4 forward (0: the engine's choice)")
    ap.add_argument("--fwd-mode-dirs", type=int, default=4,
                    help="directions of the config-4 forward-mode reference (0: skip)")
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="only run a device-to-device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe_mb > 0:
        import torch
        n = int(args.copy_probe_mb * (1 << 20)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the split-step engine runs on a single MI355X "
              "(no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import multimodenonlinearoptics_amd as pkg
    from cases import config2
    from oracle import oracle as O

    host = host_cpu()
    cores = args.cpu_threads or host["cores"]
    # the CPU leg runs the oracle compiled here with the reference's release flags
    oracle_build = {"flags": "in-tree (-O3 -march=x86-64-v3, no contraction)"}
    if not args.no_cpu:
        import tempfile
        path = os.path.join(tempfile.mkdtemp(prefix="nlm_oracle_"), "libnlm_oracle_native.so")
        try:
            oracle_build = {"flags": " ".join(O.build_native(path)[1:7])}
            O.use_library(path)
        except Exception as e:  # keep the in-tree build
            oracle_build["native_build_error"] = str(e)[:200]

    nlm = pkg.nonlinearmedium
    if args.only:
        if args.only == "hetero":
            out = hetero_benchmark(nlm, count=args.hetero_count, cpu_cores=cores, plain_sample=args.hetero_plain,
                                   cpu_sample=args.hetero_cpu)
        elif args.only == "config1":
            out = config1_benchmark(nlm, O)
        elif args.only == "gnlse":
            out = gnlse_benchmark(nlm, nz=args.gnlse_nz)
        elif args.only == "gnlse_grad":
            out = gnlse_grad_benchmark(nlm, nz=args.gnlse_nz)
        elif args.only == "compose":
            out = compose_benchmark()
        else:
            out = grad_benchmark(nlm, n_dirs=args.grad_dirs, fwd_mode_dirs=args.fwd_mode_dirs,
                                 ckpt_every=args.ckpt_every, n=args.grad_n, nz=args.grad_nz)
        print(json.dumps(out))
        return
    medium = config2(nlm, N_FREQS, NZ)
    medium.enableTiming(True)
    field_steps = 2 * N_FREQS * (NZ - 1)

    for _ in range(args.warmup):
        medium.greensOnDevice()
    medium.sync()

    prop_ms, pump_ms, asm_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        medium.greensOnDevice()
        tp = medium.lastTiming()
        pump_ms.append(tp[0])
        prop_ms.append(tp[1])
        asm_ms.append(tp[2])
    medium.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = field_steps / (ms_per_step * 1e-3)

    prop_avg = sum(prop_ms) / len(prop_ms)
    # the FFT schedule the autotuner chose for this launch shape (kernel k_prop_tab<9 + 32 v>;
    # -1: not tuned, the default schedule 0)
    from multimodenonlinearoptics_amd import capi
    sched = int(capi.lib().nlm_fft_schedule(N_FREQS, 0, 2 * N_FREQS))
    # end to end: computeGreensFunction with C and S copied to host numpy arrays
    e2e_s = median_time(lambda: medium.computeGreensFunction(), 5)
    fl = flops_per_field_step(N_FREQS)
    achieved_tflops = fl * field_steps / (prop_avg * 1e-3) / 1e12

    # plain: one propagation per launch, serially (same kernel)
    medium.runPumpSimulation()
    tau = np.asarray(medium.tau)
    sample = max(0, args.plain_sample)
    X = np.zeros((1, N_FREQS), dtype=np.complex128)
    medium.batchSignalSimulation(X, runPump=False)  # warm
    tp0 = time.perf_counter()
    for j in range(sample):
        X[:] = 0
        X[0, j % N_FREQS] = 1.0 if j % 2 == 0 else 1j
        medium.batchSignalSimulation(X, runPump=False)
    tp1 = time.perf_counter()
    plain_per_prop = (tp1 - tp0) / max(sample, 1)
    plain_solve_s = plain_per_prop * 2 * N_FREQS
    plain_value = field_steps / plain_solve_s if sample else None

    rec = {
        "metric": "field-steps/s FP64 (rel.err<=1e-10), batched Green's solve vs plain GPU",
        "value": value,
        "unit": "field-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian pump and dispersion, impulse inputs)",
        "config": {"workload": "Chi2PDC computeGreensFunction, nFreqs=512, tMax=20, nZSteps=1000, "
                               "2x512 impulse propagations per solve",
                   "nFreqs": N_FREQS, "nZSteps": NZ, "field_steps_per_step": field_steps},
        "kernel_ms": {"pump": sum(pump_ms) / len(pump_ms), "propagation": prop_avg,
                      "assembly": sum(asm_ms) / len(asm_ms)},
        "plain": {"value": plain_value, "unit": "field-steps/s", "sample_propagations": sample,
                  "s_per_solve": plain_solve_s},
        "batched_vs_plain": value / plain_value if plain_value else None,
        "end_to_end": {"s_per_solve": e2e_s, "value": field_steps / e2e_s, "unit": "field-steps/s",
                       "note": "computeGreensFunction() with C and S (2 x 4 MB) copied to host, median of 5"},
        "grad_vs_forward": None,
        "grad": None,
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_BYTES_PER_LAUNCH,
                     "kernel": "k_prop_tab<%d> (Chi2PDC Green's launch, FFT schedule %d)"
                               % (9 + 32 * max(sched, 0), max(sched, 0)),
                     "flops_per_field_step": fl,
                     "bound_note": "FP64 vector ALU and LDS exchanges; the kernel issues no MFMA. 78.6 TFLOP/s is "
                                   "MI355X's FP64 peak for both the vector units and the matrix cores"},
    }

    if not args.no_config1:
        rec["config1"] = config1_benchmark(nlm, O)

    if not args.no_hetero:
        rec["hetero"] = hetero_benchmark(nlm, count=args.hetero_count, cpu_cores=cores)

    if not args.no_gnlse:
        rec["gnlse"] = gnlse_benchmark(nlm, nz=args.gnlse_nz)

    if not args.no_gnlse_grad:
        rec["gnlse_grad"] = gnlse_grad_benchmark(nlm, nz=args.gnlse_nz,
                                                 fwd_ms=rec["gnlse"]["device_ms"] if "gnlse" in rec else None)

    if not args.no_compose:
        rec["compose"] = compose_benchmark()

    if not args.no_grad:
        g = grad_benchmark(nlm, n_dirs=args.grad_dirs, fwd_mode_dirs=args.fwd_mode_dirs,
                           ckpt_every=args.ckpt_every)
        rec["grad"] = g
        rec["grad_vs_forward"] = g["grad_vs_forward"]

    if not args.no_cpu:
        ref = config2(O, N_FREQS, NZ)
        t_all = median_time(lambda: ref.computeGreensFunction(nThreads=cores), 5)
        Co, So = ref.computeGreensFunction(nThreads=cores)
        # one thread: 32 of the 1024 impulse propagations of the same solve
        ref.runPumpSimulation()
        X = np.zeros((32, N_FREQS), dtype=np.complex128)
        for j in range(32):
            X[j, (j * 37) % N_FREQS] = 1.0 if j % 2 == 0 else 1j
        t_one = median_time(lambda: ref.batchSignalSimulation(X, inTimeDomain=False, runPump=False, nThreads=1), 5)
        one_value = 32 * (NZ - 1) / t_one
        C, S = medium.computeGreensFunction()
        num = np.linalg.norm(np.hstack([C - Co, S - So]))
        den = np.linalg.norm(np.hstack([Co, So]))
        cpu_value = field_steps / t_all
        rec["cpu_baseline"] = {"value": cpu_value, "unit": "field-steps/s", "cores": cores, "kind": "port",
                               "sample": "the full Green's solve (2x512 propagations x 999 steps) on %d threads, "
                                         "median of 5 after a warm-up: %.3f s" % (cores, t_all),
                               "value_1thread": one_value,
                               "sample_1thread": "32 of its propagations on one thread, median of 5: %.3f s" % t_one,
                               "host": host, "build": oracle_build}
        rec["vs_cpu"] = value / cpu_value
        rec["vs_cpu_1thread"] = value / one_value
        rec["rel_err"] = float(num / den)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
