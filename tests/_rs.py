"""Shared helpers for tests: import the product package (directory name starts with a digit)."""
import importlib
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
rs = importlib.import_module("3d_ravesim_amd")
lib = rs._lib


def header_functions():
    text = open(os.path.join(ROOT, "include", "ravesim.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(rsim_[a-z0-9_]+)\s*\(", text, re.M)))
