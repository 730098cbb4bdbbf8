"""Locations of the in-tree native libraries."""
from __future__ import annotations

import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(PKG_DIR)                 # genesis-2.4_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_DIR = os.path.join(PKG_ROOT, "lib")


def lib_path(name: str) -> str:
    p = os.path.join(LIB_DIR, name)
    if not os.path.exists(p):
        raise OSError(
            f"{name} is not built ({p} missing); run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C genesis-2.4_amd/csrc`")
    return p
