"""``MetalEngine``: the engine config object users pass to ``collect``
(python/polars_metal/_engine.py:8-32).  Same fields; ``device`` names the HIP
device ordinal instead of an MTLDevice."""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class MetalEngine:
    device: int | None = None
    debug: bool = False
    force_fusion: bool = False
