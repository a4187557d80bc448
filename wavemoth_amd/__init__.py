"""wavemoth_amd: MI355X (gfx950) a_lm -> HEALPix synthesis behind wavemoth's
plan/execute API.  ``ShtPlan`` mirrors wavemoth/lib.pyx; ``torch_sht``
holds the device-tensor entry points."""
from ._native import (WAVEMOTH_BATCH_NO_GRAPH, WAVEMOTH_ESTIMATE, WAVEMOTH_MEASURE,
                      WAVEMOTH_MIXED_BF16X3, WAVEMOTH_MIXED_BF16X6, WAVEMOTH_MIXED_F16X3,
                      WAVEMOTH_MIXED_F32,
                      WAVEMOTH_MMAJOR, WAVEMOTH_NO_RESOURCE_COPY, WAVEMOTH_SEEDS_ON_THE_FLY,
                      last_error)
from .batch import BatchPlan, random_requests
from .alm import lm_to_idx_mmajor, nalm, power_law_alm
from .lib import ShtPlan, legendre_transform, legendre_transform_auxdata, query_resourcefile

__all__ = ["ShtPlan", "legendre_transform", "legendre_transform_auxdata", "BatchPlan", "random_requests", "query_resourcefile", "nalm", "lm_to_idx_mmajor", "power_law_alm",
           "last_error", "WAVEMOTH_MMAJOR", "WAVEMOTH_ESTIMATE", "WAVEMOTH_MEASURE",
           "WAVEMOTH_NO_RESOURCE_COPY", "WAVEMOTH_SEEDS_ON_THE_FLY", "WAVEMOTH_BATCH_NO_GRAPH",
           "WAVEMOTH_MIXED_F32", "WAVEMOTH_MIXED_BF16X3", "WAVEMOTH_MIXED_BF16X6",
           "WAVEMOTH_MIXED_F16X3"]
