"""apg-sgm stereo engine on one MI355X (gfx950): the reference's Python API
(StereoSession, StereoResult, PipelineStats, StageTiming and the enums) over
the in-tree libapg_sgm_amd.so.  The extension module _core is built in-tree
by `make -C apg-sgm_amd` (or __graft_entry__.build()); there is no fallback.
"""
from __future__ import annotations

try:
    from ._core import (
        HierarchicalFallback,
        MemoryPolicy,
        PipelineStats,
        QualityMode,
        StageTiming,
        StatusCode,
        StereoResult,
        StereoSession,
        VolumeBackend,
    )
except ImportError as err:
    raise ImportError(
        "apg_sgm._core is not built: run `make -C apg-sgm_amd` (it links the in-tree libapg_sgm_amd.so)"
    ) from err

__version__ = "1.0.0"
__all__ = [
    "HierarchicalFallback",
    "MemoryPolicy",
    "PipelineStats",
    "QualityMode",
    "StageTiming",
    "StatusCode",
    "StereoResult",
    "StereoSession",
    "VolumeBackend",
    "__version__",
]
