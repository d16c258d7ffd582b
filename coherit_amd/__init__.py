"""coherit_amd — MI355X-native matrix-free REML engine (GPU_REML's hot path).

Python host code on PyTorch-ROCm over hand-written gfx950 HIP kernels
reached through a C-ABI (``include/coherit_amd.h``).  Modules mirror the
reference interfaces: ``geno_stream``/``block_backend`` (resident panel),
``kv_impl`` (operators), ``pcg``, ``precond``, ``slq``, ``reml`` and
``reml_model``.
"""
from __future__ import annotations

__version__ = "0.1.0"

_LAZY = {
    "GenoBlockStreamer": ("geno_stream", "GenoBlockStreamer"),
    "BedBlockStreamer": ("geno_stream", "BedBlockStreamer"),
    "BedGenoSource": ("sources", "BedGenoSource"),
    "ArraySource": ("sources", "ArraySource"),
    "SyntheticSource": ("synth", "SyntheticSource"),
    "DensePackedBlockDescriptor": ("block_backend", "DensePackedBlockDescriptor"),
    "pcg_solve": ("pcg", "pcg_solve"),
    "fit_reml": ("reml", "fit_reml"),
    "standardize_response": ("reml", "standardize_response"),
    "FitConfig": ("reml_model", "FitConfig"),
    "InfinitesimalREMLFitter": ("reml_model", "InfinitesimalREMLFitter"),
}


def __getattr__(name):
    if name in _LAZY:
        import importlib

        mod, attr = _LAZY[name]
        return getattr(importlib.import_module(f"{__name__}.{mod}"), attr)
    raise AttributeError(name)


__all__ = sorted(_LAZY)
