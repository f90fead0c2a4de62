"""Build the MI355X engine in-tree.

Produces, next to this file:
  libnlm_hip.so                     HIP kernels + C-ABI (include/nlm.h), gfx950
  nonlinearmedium<EXT_SUFFIX>       pybind11 module with the reference's classes

Every translation unit is compiled by its own hipcc/g++ process so the build
runs in parallel.  Usage: python -m multimodenonlinearoptics_amd.build [--force]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("NLM_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-Wall", "-Wno-unused-function"]
# model ids with their own propagation kernels (Chi2AFC and Chi2ASHG reuse 2 and 9)
PROP_MODELS = [0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14]

LIB = os.path.join(HERE, "libnlm_hip.so")
EXT = os.path.join(HERE, "nonlinearmedium" + sysconfig.get_config_var("EXT_SUFFIX"))


def _deps(*names):
    return [os.path.join(CSRC, n) for n in names] + [os.path.join(INCLUDE, "nlm.h")]


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(force=False, jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    dev_hdrs = ["nlm_fft.hpp", "nlm_models.hpp", "nlm_prop.hpp"]
    steps = []
    objs = []
    o = os.path.join(OBJ, "nlm_device.o")
    objs.append(o)
    deps = _deps("nlm_device.hip", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_device.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_adjoint.o")
    objs.append(o)
    deps = _deps("nlm_adjoint.hip", "nlm_adjoint.hpp", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_adjoint.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_big.o")
    objs.append(o)
    deps = _deps("nlm_big.hip", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_big.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_multi.o")
    objs.append(o)
    deps = _deps("nlm_multi.hip", "nlm_multi.hpp", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_multi.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_compose.o")
    objs.append(o)
    deps = _deps("nlm_compose.hip", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_compose.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_any.o")
    objs.append(o)
    deps = _deps("nlm_any.hip", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_any.hip"), "-o", o])
    o = os.path.join(OBJ, "nlm_noalloc.o")
    objs.append(o)
    deps = _deps("nlm_noalloc.hip", "nlm_internal.hpp", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "nlm_noalloc.hip"), "-o", o])
    for v in range(4):  # FFT schedule variants of the step-table kernel (chosen by the autotuner)
        o = os.path.join(OBJ, "nlm_prop_tab_v%d.o" % v)
        objs.append(o)
        deps = _deps("nlm_prop_tab.hip", *dev_hdrs)
        if force or _stale(o, deps):
            steps.append([HIPCC] + HIP_FLAGS + ["-DNLM_TAB_VARIANT=%d" % v, "-c",
                                                os.path.join(CSRC, "nlm_prop_tab.hip"), "-o", o])
    # the single-wave sizes of the step-table kernel, with the memory-clause scheduler
    o = os.path.join(OBJ, "nlm_prop_tab_small.o")
    objs.append(o)
    deps = _deps("nlm_prop_tab.hip", *dev_hdrs)
    if force or _stale(o, deps):
        steps.append([HIPCC] + HIP_FLAGS + ["-mllvm", "-amdgpu-sched-strategy=max-memory-clause", "-DNLM_TAB_SMALL",
                                            "-c", os.path.join(CSRC, "nlm_prop_tab.hip"), "-o", o])
    for m in PROP_MODELS:
        o = os.path.join(OBJ, "nlm_prop_%d.o" % m)
        objs.append(o)
        deps = _deps("nlm_prop_inst.hip", *dev_hdrs)
        if force or _stale(o, deps):
            steps.append([HIPCC] + HIP_FLAGS + ["-DNLM_INST_MODEL=%d" % m, "-c",
                                                os.path.join(CSRC, "nlm_prop_inst.hip"), "-o", o])
    for m in PROP_MODELS:  # heterogeneous-batch kernels of each model
        o = os.path.join(OBJ, "nlm_multi_%d.o" % m)
        objs.append(o)
        deps = _deps("nlm_multi_inst.hip", "nlm_multi.hpp", *dev_hdrs)
        if force or _stale(o, deps):
            steps.append([HIPCC] + HIP_FLAGS + ["-DNLM_INST_MODEL=%d" % m, "-c",
                                                os.path.join(CSRC, "nlm_multi_inst.hip"), "-o", o])
    host_srcs = ["nlm_host.cpp", "nlm_pybind.cpp"]
    import pybind11
    host_flags = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE, "-I" + CSRC,
                  "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
                  "-fvisibility=hidden"]
    host_objs = []
    for src in host_srcs:
        o = os.path.join(OBJ, src.replace(".cpp", ".o"))
        host_objs.append(o)
        deps = _deps(src, "nlm_host.hpp")
        if force or _stale(o, deps):
            steps.append(["g++"] + host_flags + ["-c", os.path.join(CSRC, src), "-o", o])

    jobs = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, steps))
    if verbose:
        for s, out in zip(steps, outs):
            print(" ".join(s[:1] + s[-3:]))
            if out.strip():
                print(out)

    if force or steps or _stale(LIB, objs):
        tmp = LIB + ".tmp%d" % os.getpid()
        _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
        os.replace(tmp, LIB)
    if force or steps or _stale(EXT, host_objs + [LIB]):
        tmp = EXT + ".tmp%d" % os.getpid()
        _run(["g++", "-shared", "-fPIC", "-o", tmp] + host_objs +
             ["-L" + HERE, "-lnlm_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, EXT)
    return LIB, EXT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.jobs, verbose=a.verbose))
    sys.exit(0)
