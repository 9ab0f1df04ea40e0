"""ctypes mirror of include/pfSolve.h and include/pfsolve_amd.h.

Field order and types follow the C header one to one; tests/test_abi.py
compares every offset with the library's own offsetof table.
"""
from __future__ import annotations

import ctypes as C

u64 = C.c_uint64
i64 = C.c_int64
i32 = C.c_int
dbl = C.c_double
vp = C.c_void_p
vpp = C.POINTER(C.c_void_p)
u64p = C.POINTER(C.c_uint64)

RUNTIME_MSIZE = 1
RUNTIME_OFFSETM = 2
RUNTIME_OFFSETV = 4
RUNTIME_OFFSETSOLUTION = 8
RUNTIME_INPUTZEROPAD = 16
RUNTIME_OUTPUTZEROPAD = 32
RUNTIME_SCALEC = 64
RUNTIME_INPUTBUFFERSTRIDE = 128
RUNTIME_OUTPUTBUFFERSTRIDE = 256
RUNTIME_OFFSETT = 512

BLOCK_READ_REAL_WRITE_REAL = 1
BLOCK_READ_COMPLEX_R_WRITE_REAL = 2
BLOCK_READ_COMPLEX_I_WRITE_REAL = 3
BLOCK_READ_COMPLEX_STRIDED_WRITE_COMPLEX_PACKED = 4
BLOCK_READ_REAL_WRITE_COMPLEX_R = 5
BLOCK_READ_REAL_WRITE_COMPLEX_I = 6
BLOCK_READ_COMPLEX_PACKED_WRITE_COMPLEX_STRIDED = 7
BLOCK_ADD_OUTPUTBUFFER_VALUES = 20
BLOCK_SCALEC = 100
BLOCK_SCALED = 200
BLOCK_SCALESPHLAPLA = 300
BLOCK_SCALESPHLAPLB = 400

JWT_DISABLE_TRIDIAGONAL_SOLVE = 1
JWT_DISABLE_MATVECMUL_SOLVE = 2
JWT_DIAGONAL_MATVECMUL = 3
JWT_FUSED_BAND_TRI_BIDIAG = 4
JWT_USE_PARALLEL_THOMAS = 10

def _result_codes() -> dict:
    """PfSolveResult enumerators, read from include/pfSolve.h."""
    import os
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pfSolve.h")
    text = open(hdr).read()
    return {k: int(v) for k, v in re.findall(r"\b(PFSOLVE_(?:SUCCESS|ERROR_\w+))\s*=\s*(\d+)", text)}


globals().update(_result_codes())


class PfSolveConfiguration(C.Structure):
    _fields_ = [
        ("FFTdim", u64),
        ("size", u64 * 3),
        ("device", C.POINTER(C.c_int)),
        ("stream", vpp),
        ("num_streams", u64),
        ("userTempBuffer", u64),
        ("bufferNum", u64),
        ("tempBufferNum", u64),
        ("inputBufferNum", u64),
        ("outputBufferNum", u64),
        ("kernelNum", u64),
        ("bufferSize", u64p),
        ("tempBufferSize", u64p),
        ("inputBufferSize", u64p),
        ("outputBufferSize", u64p),
        ("kernelSize", u64p),
        ("buffer", vpp),
        ("tempBuffer", vpp),
        ("inputBuffer", vpp),
        ("outputBuffer", vpp),
        ("kernel", vpp),
        ("bufferOffset", u64),
        ("tempBufferOffset", u64),
        ("inputBufferOffset", u64),
        ("outputBufferOffset", u64),
        ("kernelOffset", u64),
        ("specifyOffsetsAtLaunch", u64),
        ("s_dx", dbl),
        ("s_dy", dbl),
        ("s_dz", dbl),
        ("s_dt_D", dbl),
        ("compute_Pf", i64),
        ("finiteDifferences", i64),
        ("logicBlock", u64 * 3),
        ("upperBanded", i32),
        ("jw_type", i32),
        ("block", i32),
        ("lshift", i32),
        ("LDA", i32),
        ("KU", i32),
        ("KL", i32),
        ("ALT_dgbmv", i32),
        ("Msplit", u64 * 3),
        ("jw_control_bitmask", i64),
        ("M_size", i64),
        ("offsetM", i64),
        ("offsetV", i64),
        ("offsetSolution", i64),
        ("inputZeropad", i64 * 2),
        ("outputZeropad", i64 * 2),
        ("scaleC", dbl),
        ("coalescedMemory", u64),
        ("aimThreads", u64),
        ("numSharedBanks", u64),
        ("inverseReturnToInputBuffer", u64),
        ("numberBatches", u64),
        ("useUint64", u64),
        ("omitDimension", u64 * 3),
        ("performBandwidthBoost", u64),
        ("groupedBatch", u64 * 3),
        ("doublePrecision", u64),
        ("quadDoubleDoublePrecision", u64),
        ("quadDoubleDoublePrecisionDoubleMemory", u64),
        ("halfPrecision", u64),
        ("halfPrecisionMemoryOnly", u64),
        ("doublePrecisionFloatMemory", u64),
        ("performR2C", u64),
        ("performDCT", u64),
        ("disableMergeSequencesR2C", u64),
        ("normalize", u64),
        ("disableReorderFourStep", u64),
        ("useLUT", i64),
        ("useLUT_4step", i64),
        ("makeForwardPlanOnly", u64),
        ("makeInversePlanOnly", u64),
        ("bufferStride", u64 * 3),
        ("isInputFormatted", u64),
        ("isOutputFormatted", u64),
        ("inputBufferStride", u64 * 3),
        ("outputBufferStride", u64 * 3),
        ("considerAllAxesStrided", u64),
        ("keepShaderCode", u64),
        ("printMemoryLayout", u64),
        ("disableCaching", i32),
        ("saveApplicationToString", u64),
        ("loadApplicationFromString", u64),
        ("loadApplicationString", vp),
        ("disableSetLocale", u64),
        ("fixMaxRadixBluestein", u64),
        ("forceBluesteinSequenceSize", u64),
        ("useCustomBluesteinPaddingPattern", u64),
        ("primeSizes", u64p),
        ("paddedSizes", u64p),
        ("fixMinRaderPrimeMult", u64),
        ("fixMaxRaderPrimeMult", u64),
        ("fixMinRaderPrimeFFT", u64),
        ("fixMaxRaderPrimeFFT", u64),
        ("numConsecutiveJWIterations", i32),
        ("useMultipleInputBuffers", i32),
        ("performZeropadding", u64 * 3),
        ("fft_zeropad_left", u64 * 3),
        ("fft_zeropad_right", u64 * 3),
        ("frequencyZeroPadding", u64),
        ("performConvolution", u64),
        ("conjugateConvolution", u64),
        ("crossPowerSpectrumNormalization", u64),
        ("coordinateFeatures", u64),
        ("matrixConvolution", u64),
        ("symmetricKernel", u64),
        ("numberKernels", u64),
        ("kernelConvolution", u64),
        ("registerBoost", u64),
        ("registerBoostNonPow2", u64),
        ("registerBoost4Step", u64),
        ("swapTo3Stage4Step", u64),
        ("devicePageSize", u64),
        ("localPageSize", u64),
        ("computeCapabilityMajor", u64),
        ("computeCapabilityMinor", u64),
        ("maxComputeWorkGroupCount", u64 * 3),
        ("maxComputeWorkGroupSize", u64 * 3),
        ("maxThreadsNum", u64),
        ("sharedMemorySizeStatic", u64),
        ("sharedMemorySize", u64),
        ("sharedMemorySizePow2", u64),
        ("warpSize", u64),
        ("halfThreads", u64),
        ("allocateTempBuffer", u64),
        ("reorderFourStep", u64),
        ("maxCodeLength", i64),
        ("maxTempLength", i64),
        ("autoCustomBluesteinPaddingPattern", u64),
        ("useRaderUintLUT", u64),
        ("vendorID", u64),
        ("stream_event", vpp),
        ("streamCounter", u64),
        ("streamID", u64),
        ("useStrict32BitAddress", i64),
        ("offsetT", i64),
        ("autotune", i64),
        ("kernelVariant", i64),
        ("inputBlock", i64),
        ("outputBlock", i64),
        ("inputBlockLshift", i64),
        ("outputBlockLshift", i64),
        ("inputBlockScale", C.c_double),
        ("outputBlockScale", C.c_double),
        ("inputBlockOffset", i64),
    ]


class PfSolveLaunchParams(C.Structure):
    _fields_ = [
        ("buffer", vpp),
        ("tempBuffer", vpp),
        ("inputBuffer", vpp),
        ("outputBuffer", vpp),
        ("kernel", vpp),
        ("bufferOffset", u64),
        ("tempBufferOffset", u64),
        ("inputBufferOffset", u64),
        ("outputBufferOffset", u64),
        ("kernelOffset", u64),
        ("M_size", u64),
        ("offsetM", u64),
        ("offsetV", u64),
        ("offsetSolution", u64),
        ("inputZeropad", u64 * 2),
        ("outputZeropad", u64 * 2),
        ("scaleC", C.c_longdouble),
        ("inputBufferStride", u64),
        ("outputBufferStride", u64),
        ("batchSize", u64),
        ("offsetT", u64),
    ]


class PfSolveApplication(C.Structure):
    _fields_ = [
        ("configuration", PfSolveConfiguration),
        ("localFFTPlan", vp),
        ("localFFTPlan_inverse", vp),
        ("kernelName", C.c_char * 200),
        ("actualNumBatches", u64),
        ("firstAxis", u64),
        ("lastAxis", u64),
        ("useBluesteinFFT", u64 * 3),
        ("bufferRaderUintLUT", (vp * 4) * 3),
        ("bufferBluestein", vp * 3),
        ("bufferBluesteinFFT", vp * 3),
        ("bufferBluesteinIFFT", vp * 3),
        ("bufferRaderUintLUTSize", (u64 * 4) * 3),
        ("bufferBluesteinSize", u64 * 3),
        ("applicationBluesteinString", vp * 3),
        ("applicationBluesteinStringSize", u64 * 3),
        ("numRaderFFTPrimes", u64),
        ("rader_primes", u64 * 30),
        ("rader_buffer_size", u64 * 30),
        ("raderFFTkernel", vp * 30),
        ("applicationStringOffsetRader", u64),
        ("currentApplicationStringPos", u64),
        ("applicationStringSize", u64),
        ("saveApplicationString", vp),
    ]


class PfSolveGeometry(C.Structure):
    _fields_ = [
        ("lanes", C.c_int32),
        ("rows_per_lane", C.c_int32),
        ("waves", C.c_int32),
        ("precision", C.c_int32),
        ("table_bytes", i64),
        ("lds_bytes", i64),
        ("fast_rows_per_lane", C.c_int32),
        ("fast_waves", C.c_int32),
        ("fast_store_waves", C.c_int32),
        ("pad", C.c_int32),
    ]


BATCH_CLASSES = 5


class PfSolveBatchSet(C.Structure):
    """pfsolve_batch_set (include/pfsolve_amd.h)."""
    _fields_ = [
        ("N", i64),
        ("nsys", i64),
        ("matrices", vp),
        ("offsetV", i64),
        ("offsetT", i64),
        ("offsetM", i64),
        ("rhs", vp),
        ("stride", i64),
        ("offsetSolution", i64),
        ("naturalOrder", C.c_int32),
        ("KL", C.c_int32),
        ("KU", C.c_int32),
        ("pad", C.c_int32),
        ("scaleC", C.c_double),
    ]


class PfSolveBatchTuning(C.Structure):
    """pfsolve_batch_tuning: per row class c (R = 2 << c)."""
    _fields_ = [
        ("waves", C.c_int32 * BATCH_CLASSES),
        ("unit_pairs", C.c_int32 * BATCH_CLASSES),
        ("grid_per_cu", C.c_int32 * BATCH_CLASSES),
        ("rows_per_lane_shift", C.c_int32 * BATCH_CLASSES),
        ("kernel", C.c_int32 * BATCH_CLASSES),
        ("order", C.c_int32),
        ("systems_per_wave", C.c_int32),
    ]


class PfSolveBatchInfo(C.Structure):
    _fields_ = [
        ("sets", i64),
        ("systems", i64),
        ("units", i64),
        ("table_bytes", i64),
        ("algorithmic_bytes", i64),
        ("classes_used", C.c_int32),
        ("launches", C.c_int32),
        ("tuning", PfSolveBatchTuning),
    ]


class PfSolveServeConfig(C.Structure):
    _fields_ = [("hbm_budget", i64), ("max_inflight", C.c_int32), ("streams", C.c_int32),
                ("max_batch", C.c_int32), ("realtime", C.c_int32), ("fail_launch", C.c_int32),
                ("dry_run", C.c_int32)]


class PfSolveServeRequest(C.Structure):
    _fields_ = [
        ("N", i64), ("nsys", i64), ("matrix_key", u64), ("matrices", vp),
        ("offsetV", i64), ("offsetT", i64), ("offsetM", i64),
        ("naturalOrder", C.c_int32), ("KL", C.c_int32), ("KU", C.c_int32), ("pad", C.c_int32),
        ("scaleC", C.c_double), ("rhs_in", vp), ("rhs_out", vp),
        ("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
        ("status", C.c_int32), ("pad2", C.c_int32),
    ]


class PfSolveServeStats(C.Structure):
    _fields_ = [
        ("requests", i64), ("rejected", i64), ("systems", i64), ("launches", i64),
        ("wall_s", C.c_double), ("systems_per_s", C.c_double),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_batch", C.c_double),
        ("peak_bytes", i64), ("resident_keys", i64), ("resident_bytes", i64), ("evictions", i64),
        ("failed", i64), ("integrity_failures", i64), ("peak_keys", i64),
    ]


def declare(lib: C.CDLL) -> None:
    """Argument and result types of every exported entry point."""
    P = C.POINTER
    lib.initializePfSolve.argtypes = [P(PfSolveApplication), PfSolveConfiguration]
    lib.initializePfSolve.restype = C.c_int
    lib.PfSolveAppend.argtypes = [P(PfSolveApplication), C.c_int, P(PfSolveLaunchParams)]
    lib.PfSolveAppend.restype = C.c_int
    lib.deletePfSolve.argtypes = [P(PfSolveApplication)]
    lib.deletePfSolve.restype = None
    lib.PfSolveSync.argtypes = [P(PfSolveApplication)]
    lib.PfSolveSync.restype = C.c_int
    lib.getPfSolveErrorString.argtypes = [C.c_int]
    lib.getPfSolveErrorString.restype = C.c_char_p
    lib.PfSolveGetVersion.argtypes = []
    lib.PfSolveGetVersion.restype = C.c_int
    lib.pfsolve_sizeof_configuration.restype = C.c_size_t
    lib.pfsolve_sizeof_launch_params.restype = C.c_size_t
    lib.pfsolve_sizeof_application.restype = C.c_size_t
    lib.pfsolve_offsetof_configuration.argtypes = [C.c_char_p]
    lib.pfsolve_offsetof_configuration.restype = i64
    lib.pfsolve_offsetof_launch_params.argtypes = [C.c_char_p]
    lib.pfsolve_offsetof_launch_params.restype = i64
    lib.pfsolve_validate_configuration.argtypes = [P(PfSolveConfiguration), P(PfSolveGeometry)]
    lib.pfsolve_validate_configuration.restype = C.c_int
    lib.pfsolve_plan_geometry.argtypes = [P(PfSolveApplication), P(PfSolveGeometry)]
    lib.pfsolve_plan_geometry.restype = C.c_int
    lib.pfsolve_reorder_i0.argtypes = [i64, i64, i64, i64]
    lib.pfsolve_reorder_i0.restype = i64
    lib.pfsolve_kernel_supported.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.pfsolve_kernel_supported.restype = C.c_int
    lib.pfsolve_last_error.restype = C.c_char_p


def declare_ext(lib: C.CDLL) -> None:
    lib.pfsolve_append_noalloc.argtypes = [C.POINTER(PfSolveConfiguration), C.POINTER(PfSolveLaunchParams)]
    lib.pfsolve_append_noalloc.restype = C.c_int
    lib.pfsolve_noalloc_scratch_bytes.argtypes = []
    lib.pfsolve_noalloc_scratch_bytes.restype = C.c_size_t
    lib.pfsolve_set_kernel_events.argtypes = [C.POINTER(PfSolveApplication), C.c_void_p, C.c_void_p]
    lib.pfsolve_set_kernel_events.restype = C.c_int
    lib.pfsolve_set_wait_polls.argtypes = [C.POINTER(PfSolveApplication), C.c_int32]
    lib.pfsolve_set_wait_polls.restype = C.c_int
    P = C.POINTER
    lib.pfsolve_batch_create.argtypes = [P(vp), P(PfSolveBatchSet), i64, P(PfSolveBatchTuning)]
    lib.pfsolve_batch_create.restype = C.c_int
    for name in ("pfsolve_batch_run", "pfsolve_batch_run_plain", "pfsolve_batch_capture_graph",
                 "pfsolve_batch_launch_graph", "pfsolve_batch_prepare", "pfsolve_batch_solve",
                 "pfsolve_batch_sync"):
        f = getattr(lib, name)
        f.argtypes = [vp, vp]
        f.restype = C.c_int
    lib.pfsolve_batch_capture_graph_mode.argtypes = [vp, vp, C.c_int]
    lib.pfsolve_batch_capture_graph_mode.restype = C.c_int
    lib.pfsolve_batch_capture_graph_steps.argtypes = [vp, vp, C.c_int, C.c_int]
    lib.pfsolve_batch_capture_graph_steps.restype = C.c_int
    lib.pfsolve_batch_tune.argtypes = [vp, vp, C.c_int, P(PfSolveBatchTuning)]
    lib.pfsolve_batch_tune.restype = C.c_int
    lib.pfsolve_batch_tune_mode.argtypes = [vp, vp, C.c_int, C.c_int, P(PfSolveBatchTuning)]
    lib.pfsolve_batch_tune_mode.restype = C.c_int
    lib.pfsolve_batch_set_tuning.argtypes = [vp, P(PfSolveBatchTuning)]
    lib.pfsolve_batch_set_tuning.restype = C.c_int
    lib.pfsolve_batch_set_wait_polls.argtypes = [vp, C.c_int32]
    lib.pfsolve_batch_set_wait_polls.restype = C.c_int
    lib.pfsolve_batch_get_info.argtypes = [vp, P(PfSolveBatchInfo)]
    lib.pfsolve_batch_get_info.restype = C.c_int
    lib.pfsolve_batch_destroy.argtypes = [vp]
    lib.pfsolve_batch_destroy.restype = None
    lib.pfsolve_tuning_save.argtypes = [C.c_char_p]
    lib.pfsolve_tuning_save.restype = C.c_int
    lib.pfsolve_tuning_load.argtypes = [C.c_char_p]
    lib.pfsolve_tuning_load.restype = C.c_int
    lib.pfsolve_tuning_clear.argtypes = []
    lib.pfsolve_tuning_clear.restype = None
    lib.pfsolve_sizeof_ext.argtypes = [C.c_char_p]
    lib.pfsolve_sizeof_ext.restype = C.c_size_t
    lib.pfsolve_serve.argtypes = [P(PfSolveServeConfig), P(PfSolveServeRequest), i64, P(PfSolveServeStats)]
    lib.pfsolve_serve.restype = C.c_int
