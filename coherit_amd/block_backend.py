"""Per-call block descriptors of the operator plugin API (block_backend.py:10-17).

Kept field for field so callers that inspect ``block_backend_plan()`` work
unchanged; ``storage`` names the resident HBM panel instead of a tmpfile.
The sparse rare-variant descriptor is out of scope (dense panels only).
"""
from __future__ import annotations

import dataclasses
from typing import Literal

BlockBackendKind = Literal["dense_packed"]


@dataclasses.dataclass(frozen=True)
class DensePackedBlockDescriptor:
    call_idx: int
    snp_start: int
    true_width: int
    packed_width: int
    storage: str = "hbm_resident_packed"
    kind: BlockBackendKind = "dense_packed"


__all__ = ["BlockBackendKind", "DensePackedBlockDescriptor"]
