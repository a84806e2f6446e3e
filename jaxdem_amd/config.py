"""Precision of the arrays this package creates.

The reference computes in float32 unless JAX's 64-bit mode is on
(``jax.config.update("jax_enable_x64", True)``; its test suite turns it on
in tests/conftest.py:4).  The same switch here:

    import jaxdem_amd as jd
    jd.config.update("jax_enable_x64", True)

makes ``State.create``, ``System.create`` and the domain, material, law and
grid parameters float64; a float64 state then steps through the float64
instances of the step kernels (colliders naive and CellList, laws without
history).  The flag is read when arrays are created: states and systems made
before a change keep their precision.
"""

from __future__ import annotations

import numpy as np
import torch


class Config:
    """The ``jax_enable_x64`` switch (jax/_src/config.py)."""

    def __init__(self) -> None:
        self.jax_enable_x64 = False

    def update(self, name: str, value) -> None:
        if name != "jax_enable_x64":
            raise KeyError(f"unknown option {name!r}; jaxdem_amd reads only 'jax_enable_x64'")
        self.jax_enable_x64 = bool(value)


config = Config()


def real_dtype() -> type:
    """numpy floating type of new arrays: float64 with x64, else float32."""
    return np.float64 if config.jax_enable_x64 else np.float32


def torch_real() -> torch.dtype:
    return torch.float64 if config.jax_enable_x64 else torch.float32


__all__ = ["Config", "config", "real_dtype", "torch_real"]
