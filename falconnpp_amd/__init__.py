"""Falconn++ on AMD MI355X (gfx950).

* ``falconnpp_amd.FalconnPP`` -- the pybind11 module with the reference's
  Python API (``FalconnPP.FalconnPP(n_points, n_features)``, ``setIndexParam``,
  ``build``, ``set_qProbes``, ``query``, ...).  Putting this directory on
  ``sys.path`` also makes ``import FalconnPP`` work as with the reference.
* ``falconnpp_amd.engine`` -- ctypes access to the C ABI (device pointers,
  inspection getters, FP64 brute force).
"""
import importlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def __getattr__(name):
    if name == "FalconnPP":
        import sys
        from .engine import share_hip_runtime
        share_hip_runtime()
        if HERE not in sys.path:
            sys.path.insert(0, HERE)
        mod = importlib.import_module("FalconnPP")
        globals()["FalconnPP"] = mod
        return mod
    raise AttributeError(name)
