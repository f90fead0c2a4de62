"""MI355X-native split-step propagation engine for `nonlinearmedium`.

Importing this package loads the in-tree pybind11 module `nonlinearmedium`
(built by ``python -m multimodenonlinearoptics_amd.build``) and registers it
under its reference name, so ``import nonlinearmedium`` works afterwards.  All
propagation runs on the GPU through libnlm_hip.so; there is no CPU path, and
the import fails loudly if the extension has not been built.
"""
import importlib
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    nonlinearmedium = importlib.import_module(__name__ + ".nonlinearmedium")
except ImportError as e:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError(
        "the MI355X extension of multimodenonlinearoptics_amd is not built; run "
        "`python -m multimodenonlinearoptics_amd.build` (needs hipcc for gfx950): %s" % e
    ) from e

sys.modules.setdefault("nonlinearmedium", nonlinearmedium)

__all__ = ["nonlinearmedium"]
