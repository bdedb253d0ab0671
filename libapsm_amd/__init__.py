"""libapsm_amd: MI355X (gfx950) APSM / LLS multi-user detector.

Python mirror of the reference library's detector API (fraunhoferhhi/libapsm,
cpp/lib/noma/noma_detector.cuh:65-75, noma_detector_factory.cuh:70-74,
cpp/lib/apsm/apsm_noma_detector.cuh:83-115, cpp/lib/noma/lls_detector.cuh:75-105)
over the C ABI of libapsm_amd.so (include/apsm_capi.h). Every computation runs in
hand-written HIP kernels; this module only converts arrays and forwards calls. If the
shared library is missing, importing a detector raises instead of falling back.

Array conventions (cpp/lib/noma/types.cuh:69-100): received signals are complex
[antenna][sample]; pilots are complex [sample] (or [user][sample] for a multi-user
handle); estimates come back as complex64 [sample] (or [user][sample]).
"""
from ._lib import ApsmError, DeviceArray, PinnedArray, lib, library_path  # noqa: F401
from .batch import ApsmBatch  # noqa: F401
from .serve import ApsmServer, Frame  # noqa: F401
from .detector import (  # noqa: F401
    APSM_VERSIONS,
    ApsmNomaDetector,
    ApsmNomaDetectorFactory,
    LlsNomaDetector,
    LlsNomaDetectorFactory,
    NomaDetector,
    NomaDetectorFactory,
    NomaTrainHistory,
    apsm_get_version_string,
    apsm_parameters,
)
