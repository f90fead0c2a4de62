"""The reference's import name, ``nonlinearmedium`` (src/nlmModulePy.cpp:12).

With the repository on ``sys.path``, ``import nonlinearmedium`` or ``from
nonlinearmedium import Chi3`` loads the MI355X engine's pybind11 module
(multimodenonlinearoptics_amd/nonlinearmedium*.so) and puts it in this module's
place, so existing callers run unchanged.
"""
import sys

from multimodenonlinearoptics_amd import load as _load

sys.modules[__name__] = _load()
