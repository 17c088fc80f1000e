"""ferrousalign_amd: FerrousAlign's banded seed extension (ksw_extend2
semantics) on MI355X (gfx950) behind a C ABI (include/ferrousalign_amd.h)."""
from ._native import FaError  # noqa: F401
from .banded_swa import (  # noqa: F401
    BandedPairWiseSW,
    BatchExtensionResult,
    BatchedExtensionJob,
    Context,
    DeviceBatch,
    DeviceReference,
    ExtensionDirection,
    ExtensionJobBatch,
    OutScore,
    SoAInputs,
    bwa_fill_scmat,
    execute_batch_simd_scoring,
    extend_arrays,
    pack_pac,
    scalar_banded_swa,
    scalar_banded_swa_batch,
    simd_banded_swa_batch8_int16_soa,
    simd_banded_swa_batch16_int16_soa,
    simd_banded_swa_batch16_soa,
    simd_banded_swa_batch32_int16_soa,
    simd_banded_swa_batch32_soa,
    simd_banded_swa_batch64_soa,
    soa_transform,
)
