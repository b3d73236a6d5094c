"""MI355X-native engine for Bayadera's ensemble MCMC (stretch move) path.

``engine`` mirrors the reference's OpenCL engine types behind its protocols,
``models`` its device model library; both run through the C-ABI of
``libbayadera_hip.so`` (include/bayadera_hip.h), which loads on first use.
"""
from . import engine, models  # noqa: F401

__all__ = ["models", "engine"]
