"""MI355X-native Bethe-Salpeter kernel builder and exciton solver (Xatu ExcitonTB hot path)."""
from __future__ import annotations

import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "lib" / "libxatu_bse.so"


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python -m xatu_amd.build)")
    from . import _xatu  # noqa: F401  (links libxatu_bse.so)
    return _xatu


_xatu = _load()
SystemConfiguration = _xatu.SystemConfiguration
CRYSTALConfiguration = _xatu.CRYSTALConfiguration
ExcitonConfiguration = _xatu.ExcitonConfiguration
SystemTB = _xatu.SystemTB
ExcitonTB = _xatu.ExcitonTB
ResultTB = _xatu.ResultTB
detectDegeneracies = _xatu.detectDegeneracies

__all__ = ["SystemConfiguration", "CRYSTALConfiguration", "ExcitonConfiguration", "SystemTB", "ExcitonTB", "ResultTB",
           "detectDegeneracies", "LIB_PATH"]
