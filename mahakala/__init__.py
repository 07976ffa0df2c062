"""Import name of the reference package, bound to ``mahakala_amd``.

``import mahakala as ma`` and ``from mahakala.grmhd.athenak import
AthenakFluidModel`` work unchanged: each submodule name below is the same
module object as its ``mahakala_amd`` counterpart.
"""
import importlib
import sys

import mahakala_amd as _impl
from mahakala_amd import *  # noqa: F401,F403
from mahakala_amd import __all__  # noqa: F401

for _name in ("geodesics", "transfer", "images", "electrons", "constants", "grmhd", "grmhd.grmhd",
              "grmhd.athenak", "grmhd.synthetic"):
    _mod = importlib.import_module("mahakala_amd." + _name)
    sys.modules["mahakala." + _name] = _mod
    if "." not in _name:
        setattr(sys.modules[__name__], _name, _mod)
del _name, _mod, importlib, sys, _impl
