"""Loads the gpu-amr_amd package (its directory name is not an identifier)."""
import importlib.util
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def load_package():
    if "gpu_amr_amd" in sys.modules:
        return sys.modules["gpu_amr_amd"]
    pkg = ROOT / "gpu-amr_amd"
    spec = importlib.util.spec_from_file_location("gpu_amr_amd", pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpu_amr_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
