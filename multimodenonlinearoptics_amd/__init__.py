"""MI355X-native split-step propagation engine for `nonlinearmedium`.

The pybind11 module with the reference's classes is built in-tree by
``python -m multimodenonlinearoptics_amd.build`` and reached as
``multimodenonlinearoptics_amd.nonlinearmedium``, or under its reference name
through the top-level ``nonlinearmedium`` module (``from nonlinearmedium import
Chi3``).  It is loaded on first use, so the build module and the pure-Python
helpers import on an unbuilt tree.  All propagation runs on the GPU through
libnlm_hip.so; there is no CPU path, and using the module on an unbuilt tree
raises ImportError naming the build command.
"""
import importlib
import sys

_EXT = __name__ + ".nonlinearmedium"


def load():
    """The built extension module; registered as ``nonlinearmedium`` too."""
    mod = sys.modules.get(_EXT)
    if mod is None:
        try:
            mod = importlib.import_module(_EXT)
        except ImportError as e:
            raise ImportError(
                "the MI355X extension of multimodenonlinearoptics_amd is not built; run "
                "`python -m multimodenonlinearoptics_amd.build` (needs hipcc for gfx950): %s" % e
            ) from e
    if sys.modules.get("nonlinearmedium") is not mod:
        sys.modules["nonlinearmedium"] = mod
    return mod


def __getattr__(name):
    if name == "nonlinearmedium":
        return load()
    raise AttributeError("module %r has no attribute %r" % (__name__, name))


__all__ = ["nonlinearmedium", "load"]
