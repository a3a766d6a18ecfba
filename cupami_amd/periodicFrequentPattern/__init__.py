"""Periodic-frequent itemset miners on the gfx950 support path."""

from ._temporal import TemporalInput
from .cuPFEclat import cuPFEclat

__all__ = ["TemporalInput", "cuPFEclat"]
