"""CPU oracle: Python classes over the C restatement in nlm_oracle.c.

TEST INFRASTRUCTURE ONLY.  Used by tests/, by __graft_entry__.smoke() as the
parity reference, and by bench.py's cpu_baseline leg.  The product never imports it.

Every class mirrors the reference's pybind11 binding of the same name (argument
names, order and defaults) and maps its arguments onto the base-class
constructor exactly as the reference's C++ constructor does:

  Chi3        solvers/Chi3.cpp:18-24, 75-80      Chi2SHG     solvers/Chi2SHG.cpp:16-22, 66-71
  Chi2PDC     solvers/Chi2PDC.cpp:18-24, 53-60   Chi2SHGOPA  solvers/Chi2SHGOPA.cpp:19-29, 90-98
  Chi2SFG     solvers/Chi2SFG.cpp:18-25, 62-70   Chi2DSFG    solvers/Chi2DSFG.cpp:16-22, 71-77
  Chi2AFC     solvers/Chi2AFC.cpp:17-23, 57-65   Chi2SHGXPM  solvers/Chi2SHGXPM.cpp:18-24, 72-78
  Chi2SFGPDC  solvers/Chi2SFGPDC.cpp:22-29,...   Chi2ASHG    solvers/Chi2ASHG.cpp:15-20, 61-66
  Chi2SFGII   solvers/Chi2SFGII.cpp:20-31, 95-106 (kwarg/slot swap kept)
  Chi2PDCII   solvers/Chi2PDCII.cpp:18-26, 60-68 Chi3GNLSE   solvers/Chi3GNLSE.cpp:13-36, 98-105
  Chi2SFGOPA  solvers/Chi2SFGOPA.cpp:26-34, 53-61
  Chi2SFGXPM  solvers/Chi2SFGXPM.cpp:16-23, 60-69
  Cascade     src/Cascade.cpp:3-167 (composition restated with numpy products)
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnlm_oracle.so")
_lib = None

INF = float("inf")
Gaussian, Sech, Sinc = 0, 1, 2
GaussianBeam, Constant, GaussianApodization = 0, 1, 2

MODEL_IDS = {
    "Chi3": 0, "Chi2PDC": 1, "Chi2SFG": 2, "Chi2AFC": 3, "Chi2SFGPDC": 4, "Chi2SFGII": 5,
    "Chi2PDCII": 6, "Chi2SFGOPA": 7, "Chi2SFGXPM": 8, "Chi2SHG": 9, "Chi2SHGOPA": 10,
    "Chi2DSFG": 11, "Chi2SHGXPM": 12, "Chi2ASHG": 13, "Chi3GNLSE": 14,
}


def build(force=False):
    """Compile nlm_oracle.c into libnlm_oracle.so (gcc, -O3)."""
    src = os.path.join(_HERE, "nlm_oracle.c")
    if not force and os.path.exists(_LIB_PATH) and os.path.getmtime(_LIB_PATH) >= os.path.getmtime(src):
        return _LIB_PATH
    import subprocess
    tmp = _LIB_PATH + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-O3", "-fno-tree-vectorize", "-march=x86-64-v3", "-ffp-contract=off", "-fno-math-errno", "-fcx-limited-range", "-fPIC", "-shared", "-pthread",
           "-o", tmp, src, "-lm"]
    subprocess.check_call(cmd)
    os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


# the reference's release flags (CMakeLists.txt:30; -fcx-limited-range stands for
# Eigen's packet complex products, which skip the C99 NaN recovery)
NATIVE_FLAGS = ["-O3", "-DNDEBUG", "-march=native", "-flto", "-fno-math-errno", "-fcx-limited-range"]


def build_native(path):
    """Compile nlm_oracle.c with the reference's release flags for the host it runs
    on (used by the CPU-baseline timing; the in-tree build above, portable and
    without contraction, is the parity reference).  Returns the gcc command."""
    import subprocess
    src = os.path.join(_HERE, "nlm_oracle.c")
    cmd = ["gcc"] + NATIVE_FLAGS + ["-fPIC", "-shared", "-pthread", "-o", path, src, "-lm"]
    subprocess.check_call(cmd)
    return cmd


def use_library(path):
    """Load the oracle from `path` instead of the in-tree build (before first use)."""
    global _LIB_PATH, _lib
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


class _Params(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("n_signal_modes", C.c_int), ("n_pump_modes", C.c_int),
        ("can_be_poled", C.c_int), ("fully_nonlinear", C.c_int),
        ("relative_length", C.c_double),
        ("n_nl", C.c_int), ("nl_length", C.POINTER(C.c_double)),
        ("n_beta2", C.c_int), ("beta2", C.POINTER(C.c_double)),
        ("n_beta2s", C.c_int), ("beta2s", C.POINTER(C.c_double)),
        ("n_beta1", C.c_int), ("beta1", C.POINTER(C.c_double)),
        ("n_beta1s", C.c_int), ("beta1s", C.POINTER(C.c_double)),
        ("n_beta3", C.c_int), ("beta3", C.POINTER(C.c_double)),
        ("n_beta3s", C.c_int), ("beta3s", C.POINTER(C.c_double)),
        ("n_diff", C.c_int), ("diff_beta0", C.POINTER(C.c_double)),
        ("rayleigh_length", C.c_double), ("t_max", C.c_double),
        ("t_precision", C.c_int), ("z_precision", C.c_int), ("nz_override", C.c_int),
        ("intensity_profile", C.c_int), ("chirp", C.c_double), ("delay", C.c_double),
        ("n_custom_pump", C.c_int), ("custom_pump", C.POINTER(C.c_double)),
        ("pulse_type", C.c_int),
        ("n_poling", C.c_int), ("poling", C.POINTER(C.c_double)),
        ("n_poling_profile", C.c_int), ("poling_profile", C.POINTER(C.c_double)),
        ("gnlse", C.c_double * 5),
    ]


def lib():
    global _lib
    if _lib is None:
        if _LIB_PATH == os.path.join(_HERE, "libnlm_oracle.so"):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_void_p
        u8 = C.POINTER(C.c_uint8)
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.POINTER(_Params), C.POINTER(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_dims.argtypes = [vp, ip]
        L.orc_scalars.argtypes = [vp, dp]
        for f in ("orc_tau", "orc_omega", "orc_poling", "orc_raman"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = dp
        for f in ("orc_pump_freq", "orc_pump_time", "orc_signal_freq", "orc_signal_time", "orc_envelope"):
            getattr(L, f).argtypes = [vp, C.c_int]
            getattr(L, f).restype = dp
        L.orc_set_pump_type.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_set_pump_custom.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_set_pump_other.argtypes = [vp, vp, C.c_int, C.c_double, C.c_int]
        L.orc_run_pump.argtypes = [vp]
        L.orc_run_signal.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_greens_shape.argtypes = [vp, u8, C.c_int, u8, C.c_int, ip]
        L.orc_greens.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, u8, C.c_int, u8, C.c_int, dp, dp]
        L.orc_batch_shape.argtypes = [vp, C.c_int, C.c_int, u8, C.c_int, ip]
        L.orc_batch.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8, C.c_int, dp]
        L.orc_fft.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc == 0:
        return
    msg = lib().orc_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(lst):
    a = np.ascontiguousarray(np.asarray(lst, dtype=np.uint8).reshape(-1))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8)), a.size


def fft(x, inverse=False):
    """The reference's transform conventions on the last axis."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    n = x.shape[-1]
    out = np.empty_like(x)
    _check(lib().orc_fft(n, x.size // n, _dptr(x.view(np.float64)), _dptr(out.view(np.float64)), int(inverse)))
    return out


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(-1))


class _NonlinearMedium:
    """Oracle twin of the reference's _NonlinearMedium (src/_NonlinearMedium.hpp:20-142)."""

    _fnlm = False

    def _build(self, model, nS, nP, poled, relativeLength, nl, beta2, beta2s, customPump, pulseType,
               beta1, beta1s, beta3, beta3s, diffBeta0, rayleighLength, tMax, tPrecision, zPrecision,
               intensityProfile, chirp, delay, poling, nZSteps=None, polingProfile=None, gnlse=None):
        p = _Params()
        keep = []

        def arr(vals):
            a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).reshape(-1))
            keep.append(a)
            return a.size, _dptr(a)

        p.model = MODEL_IDS[model]
        p.n_signal_modes, p.n_pump_modes, p.can_be_poled = nS, nP, int(poled)
        p.fully_nonlinear = int(self._fnlm)
        p.relative_length = relativeLength
        p.n_nl, p.nl_length = arr(nl)
        p.n_beta2, p.beta2 = arr(beta2)
        p.n_beta2s, p.beta2s = arr(beta2s)
        p.n_beta1, p.beta1 = arr(beta1)
        p.n_beta1s, p.beta1s = arr(beta1s)
        p.n_beta3, p.beta3 = arr(beta3)
        p.n_beta3s, p.beta3s = arr(beta3s)
        p.n_diff, p.diff_beta0 = arr(diffBeta0)
        p.rayleigh_length, p.t_max = rayleighLength, tMax
        p.t_precision, p.z_precision = int(tPrecision), int(zPrecision)
        p.nz_override = int(nZSteps) if nZSteps else 0
        p.intensity_profile = int(intensityProfile)
        p.chirp, p.delay = chirp, delay
        cp = _c128(customPump if customPump is not None else [])
        keep.append(cp)
        p.n_custom_pump, p.custom_pump = cp.size, _dptr(cp.view(np.float64))
        p.pulse_type = int(pulseType)
        p.n_poling, p.poling = arr(poling if poling is not None else [])
        p.n_poling_profile, p.poling_profile = arr(polingProfile if polingProfile is not None else [])
        if gnlse is not None:
            for i, v in enumerate(gnlse):
                p.gnlse[i] = v
        h = C.c_void_p()
        _check(lib().orc_create(C.byref(p), C.byref(h)))
        self._h = h
        d = (C.c_int32 * 8)()
        lib().orc_dims(h, d)
        self._N, self._nZ, self._nZP, self._nS, self._nP, _, self._nPol, _ = list(d)
        s = (C.c_double * 5)()
        lib().orc_scalars(h, s)
        self._z, self._dz, self._dzp, self._tMax, self._zR = list(s)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.orc_destroy(h)
            self._h = None

    # -- views (copies of the C arrays) --------------------------------------
    def _vec(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)

    def _grid(self, ptr, rows):
        if rows <= 0:
            return np.zeros((0, self._N), dtype=np.complex128)
        return np.ctypeslib.as_array(ptr, shape=(rows * self._N * 2,)).view(np.complex128).reshape(rows, self._N).copy()

    @property
    def omega(self):
        return self._vec(lib().orc_omega(self._h), self._N)

    @property
    def tau(self):
        return self._vec(lib().orc_tau(self._h), self._N)

    @property
    def poling(self):
        return self._vec(lib().orc_poling(self._h), self._nPol)

    def signalFreqs(self, i=0):
        return self._grid(lib().orc_signal_freq(self._h, i), self._nZ)

    def signalTimes(self, i=0):
        return self._grid(lib().orc_signal_time(self._h, i), self._nZ)

    def pumpFreqs(self, i=0):
        return self._grid(lib().orc_pump_freq(self._h, i), self._nZP)

    def pumpTimes(self, i=0):
        return self._grid(lib().orc_pump_time(self._h, i), self._nZP)

    signalFreq = property(lambda self: self.signalFreqs(0))
    signalTime = property(lambda self: self.signalTimes(0))
    pumpFreq = property(lambda self: self.pumpFreqs(0))
    pumpTime = property(lambda self: self.pumpTimes(0))

    def envelope(self, i=0):
        return self._grid(lib().orc_envelope(self._h, i), 1)[0]

    @property
    def nZSteps(self):
        return self._nZ

    # -- methods ------------------------------------------------------------------
    def setPump(self, *args, **kw):
        first = args[0] if args else kw.pop("pulseType", kw.pop("customPump", kw.pop("other", None)))
        rest = list(args[1:])
        if isinstance(first, _NonlinearMedium):
            names = ["modeIndex", "delay", "pumpIndex"]
            vals = [0, 0.0, 0]
            for i, v in enumerate(rest):
                vals[i] = v
            for i, n in enumerate(names):
                if n in kw:
                    vals[i] = kw[n]
            for alt, i in (("signalIndex", 0), ("delayLength", 1)):
                if alt in kw:
                    vals[i] = kw[alt]
            _check(lib().orc_set_pump_other(self._h, first._h, int(vals[0]), float(vals[1]), int(vals[2])))
            return
        names = ["chirp", "delay", "pumpIndex"]
        vals = [0.0, 0.0, 0]
        for i, v in enumerate(rest):
            vals[i] = v
        for i, n in enumerate(names):
            if n in kw:
                vals[i] = kw[n]
        if np.ndim(first) == 0:
            _check(lib().orc_set_pump_type(self._h, int(first), float(vals[0]), float(vals[1]), int(vals[2])))
        else:
            a = _c128(first)
            _check(lib().orc_set_pump_custom(self._h, _dptr(a.view(np.float64)), a.size, float(vals[0]),
                                             float(vals[1]), int(vals[2])))

    def runPumpSimulation(self):
        _check(lib().orc_run_pump(self._h))

    def runSignalSimulation(self, inputProf, inTimeDomain=True, inputMode=0):
        a = _c128(inputProf)
        _check(lib().orc_run_signal(self._h, _dptr(a.view(np.float64)), a.size, int(inTimeDomain), int(inputMode)))

    def computeGreensFunction(self, inTimeDomain=False, runPump=True, nThreads=1, normalize=False,
                              useInput=(), useOutput=()):
        ui, uip, nui = _u8(list(useInput))
        uo, uop, nuo = _u8(list(useOutput))
        shp = (C.c_int32 * 2)()
        _check(lib().orc_greens_shape(self._h, uip, nui, uop, nuo, shp))
        Cm = np.zeros((shp[0], shp[1]), dtype=np.complex128)
        Sm = np.zeros((shp[0], shp[1]), dtype=np.complex128)
        _check(lib().orc_greens(self._h, int(inTimeDomain), int(runPump), int(nThreads), int(normalize),
                                uip, nui, uop, nuo, _dptr(Cm.view(np.float64)), _dptr(Sm.view(np.float64))))
        return Cm, Sm

    def batchSignalSimulation(self, inputProfs, inTimeDomain=False, runPump=True, nThreads=1, inputMode=0,
                              useOutput=()):
        return self._batch(inputProfs, inTimeDomain, runPump, nThreads, inputMode, useOutput)

    def _batch(self, inputProfs, inTimeDomain, runPump, nThreads, inputMode, useOutput):
        X = np.ascontiguousarray(np.asarray(inputProfs, dtype=np.complex128))
        if X.ndim == 1:
            X = X.reshape(1, -1)
        uo, uop, nuo = _u8(list(useOutput))
        shp = (C.c_int32 * 2)()
        _check(lib().orc_batch_shape(self._h, X.shape[0], X.shape[1], uop, nuo, shp))
        out = np.zeros((shp[0], shp[1]), dtype=np.complex128)
        _check(lib().orc_batch(self._h, _dptr(X.view(np.float64)), X.shape[0], X.shape[1], int(inTimeDomain),
                               int(runPump), int(nThreads), int(inputMode), uop, nuo, _dptr(out.view(np.float64))))
        return out


class _FullyNonlinearMedium(_NonlinearMedium):
    """Oracle twin of src/_FullyNonlinearMedium.hpp:6-69."""

    _fnlm = True

    def batchSignalSimulation(self, inputProfs, inTimeDomain=False, nThreads=1, inputMode=0, useOutput=()):
        return self._batch(inputProfs, inTimeDomain, False, nThreads, inputMode, useOutput)


def _ext(kw):
    return dict(nZSteps=kw.pop("nZSteps", None), polingProfile=kw.pop("polingProfile", None))


class Chi3(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, beta2, customPump=(), pulseType=Gaussian, beta3=0,
                 rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam,
                 chirp=0, **kw):
        self._build("Chi3", 1, 1, False, relativeLength, [nlLength], [beta2], [beta2], customPump, pulseType,
                    [0], [0], [beta3], [beta3], [], rayleighLength, tMax, tPrecision, zPrecision,
                    intensityProfile, chirp, 0, None, **_ext(kw))


class Chi2PDC(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, beta2, beta2s, customPump=(), pulseType=Gaussian, beta1=0,
                 beta1s=0, beta3=0, beta3s=0, diffBeta0=0, rayleighLength=INF, tMax=10, tPrecision=512,
                 zPrecision=100, intensityProfile=GaussianBeam, chirp=0, delay=0, poling=(), **kw):
        self._build("Chi2PDC", 1, 1, True, relativeLength, [nlLength], [beta2], [beta2s], customPump, pulseType,
                    [beta1], [beta1s], [beta3], [beta3s], [diffBeta0], rayleighLength, tMax, tPrecision,
                    zPrecision, intensityProfile, chirp, delay, poling, **_ext(kw))


class Chi2SFG(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, nlLengthOrig, beta2, beta2s, beta2o, customPump=(),
                 pulseType=Gaussian, beta1=0, beta1s=0, beta1o=0, beta3=0, beta3s=0, beta3o=0, diffBeta0=0,
                 rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam,
                 chirp=0, delay=0, poling=(), **kw):
        self._build("Chi2SFG", 2, 1, True, relativeLength, [nlLength, nlLengthOrig], [beta2], [beta2s, beta2o],
                    customPump, pulseType, [beta1], [beta1s, beta1o], [beta3], [beta3s, beta3o], [diffBeta0],
                    rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, chirp, delay, poling,
                    **_ext(kw))


class Chi2AFC(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, nlLengthOrig, beta2, beta2s, beta2o, customPump=(),
                 pulseType=Gaussian, beta1=0, beta1s=0, beta1o=0, beta3=0, beta3s=0, beta3o=0, diffBeta0Start=0,
                 diffBeta0End=0, rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100,
                 intensityProfile=GaussianBeam, chirp=0, delay=0, **kw):
        self._build("Chi2AFC", 2, 1, False, relativeLength, [0.5 * math.pi * nlLength, 0.5 * math.pi * nlLengthOrig],
                    [beta2], [beta2s, beta2o], customPump, pulseType, [beta1], [beta1s, beta1o], [beta3],
                    [beta3s, beta3o], [diffBeta0Start, diffBeta0End], rayleighLength, tMax, tPrecision, zPrecision,
                    intensityProfile, chirp, delay, None, **_ext(kw))


class Chi2SFGPDC(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, nlLengthOrig, beta2, beta2s, beta2o, customPump=(),
                 pulseType=Gaussian, beta1=0, beta1s=0, beta1o=0, beta3=0, beta3s=0, beta3o=0, diffBeta0=0,
                 diffBeta0o=0, rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100,
                 intensityProfile=GaussianBeam, chirp=0, delay=0, poling=(), **kw):
        self._build("Chi2SFGPDC", 2, 1, True, relativeLength, [nlLength, nlLengthOrig], [beta2], [beta2s, beta2o],
                    customPump, pulseType, [beta1], [beta1s, beta1o], [beta3], [beta3s, beta3o],
                    [diffBeta0, diffBeta0o], rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, chirp,
                    delay, poling, **_ext(kw))


class Chi2SFGII(_NonlinearMedium):
    # The reference binds the Python names "intensityProfile" and "zPrecision" to
    # the C++ slots (zPrecision, intensityProfile) respectively
    # (solvers/Chi2SFGII.cpp:97-106); the same slot order is kept here.
    def __init__(self, relativeLength, nlLengthZY, nlLengthZZ, beta2, beta2sz, beta2sy, beta2oz, beta2oy,
                 customPump=(), pulseType=Gaussian, beta1=0, beta1sz=0, beta1sy=0, beta1oz=0, beta1oy=0, beta3=0,
                 beta3sz=0, beta3sy=0, beta3oz=0, beta3oy=0, diffBeta0z=0, diffBeta0y=0, diffBeta0=0,
                 rayleighLength=INF, tMax=10, tPrecision=512, intensityProfile=GaussianBeam, zPrecision=100,
                 chirp=0, delay=0, poling=(), **kw):
        zslot, pslot = intensityProfile, zPrecision
        if int(pslot) not in (0, 1, 2):
            raise TypeError("__init__(): incompatible constructor arguments")
        self._build("Chi2SFGII", 4, 1, True, relativeLength, [nlLengthZY, nlLengthZZ], [beta2],
                    [beta2sz, beta2sy, beta2oz, beta2oy], customPump, pulseType, [beta1],
                    [beta1sz, beta1sy, beta1oz, beta1oy], [beta3], [beta3sz, beta3sy, beta3oz, beta3oy],
                    [diffBeta0z, diffBeta0y, diffBeta0], rayleighLength, tMax, tPrecision, zslot, pslot, chirp,
                    delay, poling, **_ext(kw))


class Chi2PDCII(_NonlinearMedium):
    def __init__(self, relativeLength, nlLengthI, nlLengthII, beta2, beta2s, beta2o, customPump=(),
                 pulseType=Gaussian, beta1=0, beta1s=0, beta1o=0, beta3=0, beta3s=0, beta3o=0, diffBeta0=0,
                 rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam,
                 chirp=0, delay=0, poling=(), **kw):
        self._build("Chi2PDCII", 2, 1, True, relativeLength, [nlLengthI, nlLengthII], [beta2], [beta2s, beta2o],
                    customPump, pulseType, [beta1], [beta1s, beta1o], [beta3], [beta3s, beta3o], [diffBeta0],
                    rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, chirp, delay, poling,
                    **_ext(kw))


class Chi2SFGOPA(_NonlinearMedium):
    def __init__(self, relativeLength, nlLengthSFGh, nlLengthSFGf, nlLengthDOPAh, nlLengthDOPAf, beta2F, beta2H,
                 beta2h, beta2f, customPump=(), pulseType=Gaussian, beta1F=0, beta1H=0, beta1h=0, beta1f=0,
                 beta3F=0, beta3H=0, beta3h=0, beta3f=0, diffBeta0SFG=0, diffBeta0DOPA=0, rayleighLength=INF,
                 tMax=10, tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam, chirp=0, delay=0,
                 poling=(), **kw):
        self._build("Chi2SFGOPA", 2, 2, True, relativeLength,
                    [nlLengthSFGh, nlLengthSFGf, nlLengthDOPAh, nlLengthDOPAf], [beta2F, beta2H], [beta2h, beta2f],
                    customPump, pulseType, [beta1F, beta1H], [beta1h, beta1f], [beta3F, beta3H], [beta3h, beta3f],
                    [diffBeta0SFG, diffBeta0DOPA], rayleighLength, tMax, tPrecision, zPrecision, intensityProfile,
                    chirp, delay, poling, **_ext(kw))


class Chi2SFGXPM(_NonlinearMedium):
    def __init__(self, relativeLength, nlLength, nlLengthOrig, nlLengthChi3, nlLengthChi3Orig, beta2, beta2s,
                 beta2o, customPump=(), pulseType=Gaussian, beta1=0, beta1s=0, beta1o=0, beta3=0, beta3s=0,
                 beta3o=0, diffBeta0=0, rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100,
                 intensityProfile=GaussianBeam, chirp=0, delay=0, poling=(), **kw):
        self._build("Chi2SFGXPM", 2, 1, True, relativeLength,
                    [nlLength, nlLengthOrig, nlLengthChi3 / 2, nlLengthChi3Orig / 2], [beta2], [beta2s, beta2o],
                    customPump, pulseType, [beta1], [beta1s, beta1o], [beta3], [beta3s, beta3o], [diffBeta0],
                    rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, chirp, delay, poling,
                    **_ext(kw))


def _fnlm_build(self, model, nS, poled, relativeLength, nl, beta2s, beta1s, beta3s, diff, rayleighLength, tMax,
                tPrecision, zPrecision, intensityProfile, poling, kw, gnlse=None):
    self._build(model, nS, 0, poled, relativeLength, nl, [], beta2s, None, 0, [], beta1s, [], beta3s, diff,
                rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, 0, 0, poling, gnlse=gnlse, **_ext(kw))


class Chi2SHG(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLengthH, nlLengthP, beta2h, beta2p, beta1h=0, beta1p=0, beta3h=0,
                 beta3p=0, diffBeta0=0, rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100,
                 intensityProfile=GaussianBeam, poling=(), **kw):
        _fnlm_build(self, "Chi2SHG", 2, True, relativeLength, [nlLengthP, nlLengthH], [beta2p, beta2h],
                    [beta1p, beta1h], [beta3p, beta3h], [diffBeta0], rayleighLength, tMax, tPrecision, zPrecision,
                    intensityProfile, poling, kw)


class Chi2SHGOPA(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLengthP, nlLengthSH, nlLengthPA1, nlLengthPA2, beta2p, beta2sh, beta2pa1,
                 beta2pa2, beta1p=0, beta1sh=0, beta1pa1=0, beta1pa2=0, beta3p=0, beta3sh=0, beta3pa1=0,
                 beta3pa2=0, diffBeta0shg=0, diffBeta0opa=0, rayleighLength=INF, tMax=10, tPrecision=512,
                 zPrecision=100, intensityProfile=GaussianBeam, poling=(), **kw):
        _fnlm_build(self, "Chi2SHGOPA", 4, True, relativeLength, [nlLengthP, nlLengthSH, nlLengthPA1, nlLengthPA2],
                    [beta2p, beta2sh, beta2pa1, beta2pa2], [beta1p, beta1sh, beta1pa1, beta1pa2],
                    [beta3p, beta3sh, beta3pa1, beta3pa2], [diffBeta0shg, diffBeta0opa], rayleighLength, tMax,
                    tPrecision, zPrecision, intensityProfile, poling, kw)


class Chi2DSFG(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLengthP, nlLengthS, nlLengthD, beta2p, beta2s, beta2d, beta1p=0,
                 beta1s=0, beta1d=0, beta3p=0, beta3s=0, beta3d=0, diffBeta0=0, rayleighLength=INF, tMax=10,
                 tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam, poling=(), **kw):
        _fnlm_build(self, "Chi2DSFG", 3, True, relativeLength, [nlLengthP, nlLengthS, nlLengthD],
                    [beta2p, beta2s, beta2d], [beta1p, beta1s, beta1d], [beta3p, beta3s, beta3d], [diffBeta0],
                    rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, poling, kw)


class Chi2SHGXPM(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLengthH, nlLengthP, nlLengthChi3, beta2h, beta2p, beta1h=0, beta1p=0,
                 beta3h=0, beta3p=0, diffBeta0=0, rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100,
                 intensityProfile=GaussianBeam, poling=(), **kw):
        _fnlm_build(self, "Chi2SHGXPM", 2, True, relativeLength, [nlLengthP, nlLengthH, nlLengthChi3],
                    [beta2p, beta2h], [beta1p, beta1h], [beta3p, beta3h], [diffBeta0], rayleighLength, tMax,
                    tPrecision, zPrecision, intensityProfile, poling, kw)


class Chi2ASHG(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLengthH, nlLengthP, beta2h, beta2p, beta1h=0, beta1p=0, beta3h=0,
                 beta3p=0, diffBeta0Start=0, diffBeta0End=0, rayleighLength=INF, tMax=10, tPrecision=512,
                 zPrecision=100, intensityProfile=GaussianBeam, **kw):
        _fnlm_build(self, "Chi2ASHG", 2, False, relativeLength, [0.5 * math.pi * nlLengthP, 0.5 * math.pi * nlLengthH],
                    [beta2p, beta2h], [beta1p, beta1h], [beta3p, beta3h], [diffBeta0Start, diffBeta0End],
                    rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, None, kw)


class Chi3GNLSE(_FullyNonlinearMedium):
    def __init__(self, relativeLength, nlLength, selfSteepLength, fr, fb, tau1, tau2, tau3, beta2, beta3=0,
                 rayleighLength=INF, tMax=10, tPrecision=512, zPrecision=100, intensityProfile=GaussianBeam, **kw):
        _fnlm_build(self, "Chi3GNLSE", 1, False, relativeLength, [nlLength, selfSteepLength], [beta2], [0], [beta3],
                    [], rayleighLength, tMax, tPrecision, zPrecision, intensityProfile, None, kw,
                    gnlse=(fr, fb, tau1, tau2, tau3))

    @property
    def ramanResponse(self):
        return self._grid(lib().orc_raman(self._h), 1)[0]


SOLVERS = {c.__name__: c for c in (Chi3, Chi2PDC, Chi2SFG, Chi2AFC, Chi2SFGPDC, Chi2SFGII, Chi2PDCII, Chi2SFGOPA,
                                   Chi2SFGXPM, Chi2SHG, Chi2SHGOPA, Chi2DSFG, Chi2SHGXPM, Chi2ASHG, Chi3GNLSE)}


class Cascade:
    """Oracle twin of src/Cascade.cpp:3-167 (Green's composition with numpy)."""

    def __init__(self, inputMedia, modeConnections, sharePump):
        if len(inputMedia) == 0:
            raise ValueError("Cascade must contain at least one medium")
        self.media = list(inputMedia)
        N, tMax = self.media[0]._N, self.media[0]._tMax
        for m in self.media:
            if m._N != N or m._tMax != tMax:
                raise ValueError("Medium does not have same time and frequency axes as the first")
        if len(modeConnections) != len(self.media) - 1:
            raise ValueError("Must have one connection per pair of adjacent media")
        for i, con in enumerate(modeConnections):
            if len(con) == 0:
                raise ValueError("Connections missing between media!")
            for k, v in con.items():
                if v >= self.media[i]._nS or k >= self.media[i + 1]._nS:
                    raise ValueError("Invalid connections, out of range")
        self.connections = [dict(c) for c in modeConnections]
        self.sharedPump = bool(sharePump)
        self._N = N

    def setPump(self, *args, **kw):
        """src/Cascade.cpp:57-72: the first medium's pump when it is shared, else every medium's."""
        for m in (self.media[:1] if self.sharedPump else self.media):
            m.setPump(*args, **kw)

    def runPumpSimulation(self):
        if not self.sharedPump:
            for m in self.media:
                m.runPumpSimulation()
        else:
            self.media[0].runPumpSimulation()
            for i in range(1, len(self.media)):
                last = self.media[i - 1].pumpTimes(0)[-1]
                self.media[i].setPump(last)
                self.media[i].runPumpSimulation()

    def runSignalSimulation(self, inputProf, inTimeDomain=True, inputMode=0):
        self.media[0].runSignalSimulation(inputProf, inTimeDomain, inputMode)
        for i in range(1, len(self.media)):
            con = self.connections[i - 1]
            mx = max(con.keys())
            nxt = np.zeros(self._N * (mx + 1), dtype=np.complex128)
            for k, v in con.items():
                nxt[k * self._N:(k + 1) * self._N] = self.media[i - 1].signalFreqs(v)[-1]
            self.media[i].runSignalSimulation(nxt, False)

    def computeGreensFunction(self, inTimeDomain=False, runPump=True, nThreads=1, normalize=False,
                              useInput=(), useOutput=()):
        if runPump:
            self.runPumpSimulation()
        n = len(self.media)
        Cg, Sg = self.media[0].computeGreensFunction(inTimeDomain, False, nThreads, normalize, useInput,
                                                     useOutput if n == 1 else ())
        for i in range(1, n):
            C2, S2 = self.media[i].computeGreensFunction(inTimeDomain, False, nThreads, normalize, (),
                                                         useOutput if i == n - 1 else ())
            Cg, Sg = C2 @ Cg + S2 @ Sg.conj(), C2 @ Sg + S2 @ Cg.conj()
        return Cg, Sg

    def batchSignalSimulation(self, inputProfs, inTimeDomain=False, runPump=True, nThreads=1, inputMode=0,
                              useOutput=()):
        if runPump:
            self.runPumpSimulation()
        out = self.media[0]._batch(inputProfs, inTimeDomain, False, nThreads, inputMode, ())
        for i in range(1, len(self.media)):
            if i != len(self.media) - 1:
                out = self.media[i]._batch(out, inTimeDomain, False, nThreads, 0, ())
            else:
                out = self.media[i]._batch(out, inTimeDomain, False, nThreads, 0, useOutput)
        return out


class PulseType:
    Gaussian, Sech, Sinc = 0, 1, 2


class IntensityProfile:
    GaussianBeam, Constant, GaussianApodization = 0, 1, 2
