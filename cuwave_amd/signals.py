"""Source wavelets (cuwave/signals.py:5-28), on the host in float64."""

from __future__ import annotations

import numpy as np


def sineburst(t: np.ndarray, amplitude: float, frequency: float, cycles: int) -> np.ndarray:
    """Hann-windowed sine burst of `cycles` periods at `frequency`, zero outside its support."""
    mask = (t > 0) & (t <= cycles / frequency)
    return (amplitude * mask * np.sin(2 * np.pi * frequency * t)
            * np.sin(np.pi * frequency * t / cycles) ** 2)


def ricker(t: np.ndarray, amplitude: float, frequency: float, delay: float | None = None) -> np.ndarray:
    """Ricker wavelet at `frequency`, centred at `delay` (one period by default)."""
    if delay is None:
        delay = 1.0 / frequency
    arg = (np.pi * frequency * (t - delay)) ** 2
    return amplitude * (1.0 - 2.0 * arg) * np.exp(-arg)
