"""Frequent-itemset miners on the gfx950 support path."""

from .cuApriori import cuApriori, Apriori
from .cuEclat import cuEclat, Eclat
from .cuFPGrowth import cuFPGrowth, FPGrowth

__all__ = ["cuApriori", "Apriori", "cuEclat", "Eclat", "cuFPGrowth", "FPGrowth"]
