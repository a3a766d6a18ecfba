"""MI355X (gfx950) support-counting path for cuPAMI's frequent-itemset miners.

Public names follow the reference package (``cupami/__init__.py:14-22``) for
the miners this build covers.  Importing needs no GPU; mining needs
``libcupami_hip.so`` (``python -m cupami_amd.build``) and a gfx950 device.
"""

from .frequentPattern import cuApriori, Apriori, cuEclat, Eclat, cuFPGrowth, FPGrowth
from .periodicFrequentPattern import TemporalInput, cuPFEclat

__version__ = "0.1.0"

__all__ = ["cuApriori", "Apriori", "cuEclat", "Eclat", "cuFPGrowth", "FPGrowth",
           "TemporalInput", "cuPFEclat"]
