"""MI355X-native 3D adaptive FMM (HIP, gfx950). See afmm.py for the binding.

The directory name is not a Python identifier; load it by path, e.g.
importlib.util.spec_from_file_location("on_galaxy_amd", ".../on-galaxy_amd/__init__.py").
"""
import importlib.util as _ilu
import os as _os

_spec = _ilu.spec_from_file_location("on_galaxy_amd_afmm", _os.path.join(_os.path.dirname(__file__), "afmm.py"))
afmm = _ilu.module_from_spec(_spec)
_spec.loader.exec_module(afmm)

AdaptiveFMM3D = afmm.AdaptiveFMM3D
AfmmError = afmm.AfmmError
