"""polars_metal on MI355X: the relational hot path (filter, group-by, fused
aggregation) of polars-metal, executed by hand-written HIP kernels for
gfx950 behind the C ABI in include/pmh.h.

``_native`` mirrors the reference's PyO3 module function for function
(crates/polars-metal-core/src/lib.rs:57-83).  When Polars is importable, the
engine hook (``MetalEngine`` + patched ``LazyFrame.collect``) is installed by
``install()``; Polars is optional for the kernel-level API.
"""
from __future__ import annotations

from ._engine import MetalEngine

__all__ = ["MetalEngine", "install", "native"]


def native():
    """The ``_native`` module (raises ImportError if libpmh.so is not built)."""
    from . import _lib, _native
    _lib.load()
    return _native


def install() -> bool:
    """Install the Polars engine hook; returns False when Polars is absent."""
    try:
        import polars  # noqa: F401
    except Exception:
        return False
    from . import _hook
    _hook.install()
    return True
