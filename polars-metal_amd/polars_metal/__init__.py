"""polars_metal on MI355X: the relational hot path (filter, group-by, fused
aggregation) of polars-metal, executed by hand-written HIP kernels for
gfx950 behind the C ABI in include/pmh.h.

``_native`` mirrors the reference's PyO3 module function for function
(crates/polars-metal-core/src/lib.rs:57-83).  ``udf.dispatch_groupby`` is
the group-by dispatch of the reference's engine hook
(python/polars_metal/_udf.py:955-1127) over Arrow tables.  ``install``
points a loaded copy of the reference's ``_udf`` module at both.
"""
from __future__ import annotations

from ._engine import MetalEngine

__all__ = ["MetalEngine", "install", "native"]


def native():
    """The ``_native`` module (raises ImportError if libpmh.so is not built)."""
    from . import _lib, _native
    _lib.load()
    return _native


def install(reference_udf) -> None:
    """Route the reference's engine hook through this backend.

    ``reference_udf`` is the reference's ``polars_metal._udf`` module object,
    imported by the deployment under whatever package name it gives the
    reference (this package already owns ``polars_metal``).  Its ``_native``
    binding (python/polars_metal/_udf.py:49) becomes this package's
    ``_native``, so every compare, Kleene, compaction and group-by call the
    hook makes lands on the HIP kernels, and its ``_dispatch_groupby``
    (_udf.py:955) goes through ``udf.dispatch_groupby_pydf``, which runs a
    GroupBy over Filter or Scan as one fused device call instead of Polars'
    CPU compaction followed by ``execute_groupby``; a shape it does not
    lower (NotImplementedError) is handed to the reference's own dispatch.
    The previous bindings are kept as ``_native_reference`` and
    ``_dispatch_groupby_reference``.
    """
    from . import _native, udf
    for attr in ("_native", "_dispatch_groupby"):
        if not hasattr(reference_udf, attr):
            raise TypeError(f"polars_metal.install: {reference_udf!r} has no {attr}; pass the reference's _udf module")
    if getattr(reference_udf._dispatch_groupby, "__polars_metal_amd__", False):
        return
    original = reference_udf._dispatch_groupby

    def dispatch_groupby(df_pydf, wire_plan):
        try:
            return udf.dispatch_groupby_pydf(df_pydf, wire_plan)
        except NotImplementedError:
            return original(df_pydf, wire_plan)

    dispatch_groupby.__polars_metal_amd__ = True
    reference_udf._native_reference = reference_udf._native
    reference_udf._dispatch_groupby_reference = original
    reference_udf._native = _native
    reference_udf._dispatch_groupby = dispatch_groupby
