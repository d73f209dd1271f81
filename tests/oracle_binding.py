"""Test helper: the CPU oracle behind the same Python API as ``peclet.dem.Simulation``.

Only tests, smoke() and the bench's CPU-baseline leg use this; the product never loads the oracle.
"""
from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "peclet-dem_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from peclet.dem import Simulation  # noqa: E402
from peclet.dem._binding import Binding  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "libpdref.so")
_oracle = None


def oracle_binding() -> Binding:
    global _oracle
    if _oracle is None:
        _oracle = Binding(ORACLE_LIB, "pdref_", product=False)
    return _oracle


def oracle_sim(capacity: int) -> Simulation:
    return Simulation(capacity, _binding=oracle_binding())
