"""MI355X-native chain-propagator and concentration path with the Python surface
of the `langevinfts` module of Langevin_FTS.

    from langevin_fts_amd import langevinfts     # PlatformSelector, factories, solvers (pybind11)
    from langevin_fts_amd import capi            # ctypes binding of the C ABI (libfts_hip.so)

Both native libraries are built in-tree by `python -m langevin_fts_amd.build`.
There is no CPU fallback: using the package without the built HIP library raises.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIBRARY = os.path.join(PKG_DIR, "libfts_hip.so")


def __getattr__(name):
    if name == "langevinfts":
        import importlib

        try:
            mod = importlib.import_module(".langevinfts", __name__)
        except ImportError as e:
            raise ImportError("langevin_fts_amd.langevinfts is not built; run `python -m langevin_fts_amd.build`") from e
        globals()["langevinfts"] = mod
        return mod
    raise AttributeError(name)
