"""Load the product package (directory hoomd-rs_amd/) and its workloads."""
import importlib.util
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "hoomd-rs_amd"


def pkg():
    if "hoomd_rs_amd" in sys.modules:
        return sys.modules["hoomd_rs_amd"]
    spec = importlib.util.spec_from_file_location("hoomd_rs_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hoomd_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workloads():
    if "hoomd_rs_amd_workloads" in sys.modules:
        return sys.modules["hoomd_rs_amd_workloads"]
    spec = importlib.util.spec_from_file_location("hoomd_rs_amd_workloads", PKG_DIR / "workloads.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hoomd_rs_amd_workloads"] = mod
    spec.loader.exec_module(mod)
    return mod
