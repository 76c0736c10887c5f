"""Loads the product package (its directory name is not an identifier) and the
CPU oracle for the tests."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "transient-analysis-grounding-systems_amd")


def load():
    if "tags_amd" in sys.modules:
        return sys.modules["tags_amd"]
    spec = importlib.util.spec_from_file_location(
        "tags_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tags_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def oracle():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O
    return O
