/*
 * pfSolve.h -- drop-in C interface of the MI355X (gfx950) Jones-Worland step.
 *
 * Source-compatible with the plan surface of the reference PfSolve headers for
 * the JW path: a caller that builds against the reference's pfSolve.h with
 * VKFFT_BACKEND=2 (HIP) switches its include path to this directory and links
 * libpfsolve_gfx950.so.  The reference implements these as static inline
 * functions that generate and JIT-compile kernels; here they are exported
 * extern "C" symbols that launch precompiled gfx950 kernels.
 *
 * Interfaces replaced (reference file:line):
 *   initializePfSolve      PfSolve/pfSolve_AppManagement/pfSolve_InitializeApp.h:1480
 *   PfSolveAppend          PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:80
 *   deletePfSolve          PfSolve/pfSolve_AppManagement/pfSolve_DeleteApp.h:28
 *   PfSolveSync            PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:28
 *   getPfSolveErrorString  PfSolve/pfSolve_Structs/pfSolve_Structs.h:545
 *   PfSolveGetVersion      PfSolve/pfSolve.h:101-103
 *   PfSolveConfiguration   PfSolve/pfSolve_Structs/pfSolve_Structs.h:119-381
 *   PfSolveLaunchParams    PfSolve/pfSolve_Structs/pfSolve_Structs.h:383-449
 *   PfSolveResult          PfSolve/pfSolve_Structs/pfSolve_Structs.h:450-543
 *   RUNTIME_/BLOCK_/JWT_   PfSolve/pfSolve_Structs/pfSolve_Structs.h:73-99
 *
 * PfSolveConfiguration holds every field of the reference's HIP build
 * (VKFFT_BACKEND=2), in its order and at its offsets, so that a caller's
 * assignments compile unchanged; the fields of the reference's FFT machinery
 * (Bluestein/Rader tuning, convolution, R2C/DCT, register boost, paging) are
 * accepted and ignored.  Extension fields are appended after the last
 * reference field.
 */
#ifndef PFSOLVE_H
#define PFSOLVE_H

#include <stdint.h>
#include <stddef.h>

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#define PFSOLVE_ZERO_INIT {}
extern "C" {
#else
#define PFSOLVE_ZERO_INIT {0}
#endif

#define pfLD long double
#define pfUINT uint64_t
#define pfINT int64_t

/* runtime-parameter bits of jw_control_bitmask (pfSolve_Structs.h:73-81) */
#define RUNTIME_MSIZE 1
#define RUNTIME_OFFSETM 2
#define RUNTIME_OFFSETV 4
#define RUNTIME_OFFSETSOLUTION 8
#define RUNTIME_INPUTZEROPAD 16
#define RUNTIME_OUTPUTZEROPAD 32
#define RUNTIME_SCALEC 64
#define RUNTIME_INPUTBUFFERSTRIDE 128
#define RUNTIME_OUTPUTBUFFERSTRIDE 256

/* block plan type digits (pfSolve_Structs.h:83-94) */
#define BLOCK_READ_REAL_WRITE_REAL 1
#define BLOCK_READ_COMPLEX_R_WRITE_REAL 2
#define BLOCK_READ_COMPLEX_I_WRITE_REAL 3
#define BLOCK_READ_COMPLEX_STRIDED_WRITE_COMPLEX_PACKED 4
#define BLOCK_READ_REAL_WRITE_COMPLEX_R 5
#define BLOCK_READ_REAL_WRITE_COMPLEX_I 6
#define BLOCK_READ_COMPLEX_PACKED_WRITE_COMPLEX_STRIDED 7
#define BLOCK_ADD_OUTPUTBUFFER_VALUES 20
#define BLOCK_SCALEC 100
#define BLOCK_SCALED 200
#define BLOCK_SCALESPHLAPLA 300
#define BLOCK_SCALESPHLAPLB 400

/* jw_type digits (pfSolve_Structs.h:96-99) */
#define JWT_DISABLE_TRIDIAGONAL_SOLVE 1
#define JWT_DISABLE_MATVECMUL_SOLVE 2
#define JWT_DIAGONAL_MATVECMUL 3
#define JWT_USE_PARALLEL_THOMAS 10

typedef struct {
    /* required parameters */
    uint64_t FFTdim;
    uint64_t size[3];            /* N (rows), B (systems), batches of B */

    hipDevice_t* device;
    hipStream_t* stream;
    uint64_t num_streams;

    /* data parameters */
    uint64_t userTempBuffer;
    uint64_t bufferNum;
    uint64_t tempBufferNum;
    uint64_t inputBufferNum;
    uint64_t outputBufferNum;
    uint64_t kernelNum;
    uint64_t* bufferSize;
    uint64_t* tempBufferSize;
    uint64_t* inputBufferSize;
    uint64_t* outputBufferSize;
    uint64_t* kernelSize;
    void** buffer;               /* matrices ("inputs") */
    void** tempBuffer;
    void** inputBuffer;
    void** outputBuffer;         /* right-hand sides, solved in place */
    void** kernel;               /* scratch of the long-N path */
    uint64_t bufferOffset;
    uint64_t tempBufferOffset;
    uint64_t inputBufferOffset;
    uint64_t outputBufferOffset;
    uint64_t kernelOffset;
    uint64_t specifyOffsetsAtLaunch;

    double s_dx;
    double s_dy;
    double s_dz;
    double s_dt_D;
    int64_t compute_Pf;
    int64_t finiteDifferences;
    uint64_t logicBlock[3];

    int upperBanded;
    int jw_type;
    int block;
    int lshift;
    int LDA;
    int KU;
    int KL;
    int ALT_dgbmv;
    uint64_t Msplit[3];

    int64_t jw_control_bitmask;

    int64_t M_size;
    int64_t offsetM;
    int64_t offsetV;
    int64_t offsetSolution;
    int64_t inputZeropad[2];
    int64_t outputZeropad[2];
    double scaleC;

    uint64_t coalescedMemory;
    uint64_t aimThreads;
    uint64_t numSharedBanks;
    uint64_t inverseReturnToInputBuffer;
    uint64_t numberBatches;
    uint64_t useUint64;
    uint64_t omitDimension[3];
    uint64_t performBandwidthBoost;
    uint64_t groupedBatch[3];

    uint64_t doublePrecision;
    pfUINT quadDoubleDoublePrecision;
    pfUINT quadDoubleDoublePrecisionDoubleMemory;
    uint64_t halfPrecision;
    uint64_t halfPrecisionMemoryOnly;
    uint64_t doublePrecisionFloatMemory;

    uint64_t performR2C;
    uint64_t performDCT;
    uint64_t disableMergeSequencesR2C;
    uint64_t normalize;
    uint64_t disableReorderFourStep;
    int64_t useLUT;
    int64_t useLUT_4step;
    uint64_t makeForwardPlanOnly;
    uint64_t makeInversePlanOnly;

    uint64_t bufferStride[3];
    uint64_t isInputFormatted;
    uint64_t isOutputFormatted;
    uint64_t inputBufferStride[3];
    uint64_t outputBufferStride[3];

    uint64_t considerAllAxesStrided;
    uint64_t keepShaderCode;
    uint64_t printMemoryLayout;

    int disableCaching;
    uint64_t saveApplicationToString;
    uint64_t loadApplicationFromString;
    void* loadApplicationString;
    uint64_t disableSetLocale;

    /* Bluestein / Rader tuning of the reference's FFT machinery: ignored */
    uint64_t fixMaxRadixBluestein;
    uint64_t forceBluesteinSequenceSize;
    uint64_t useCustomBluesteinPaddingPattern;
    uint64_t* primeSizes;
    uint64_t* paddedSizes;
    uint64_t fixMinRaderPrimeMult;
    uint64_t fixMaxRaderPrimeMult;
    uint64_t fixMinRaderPrimeFFT;
    uint64_t fixMaxRaderPrimeFFT;

    int numConsecutiveJWIterations;
    int useMultipleInputBuffers;

    uint64_t performZeropadding[3];
    uint64_t fft_zeropad_left[3];
    uint64_t fft_zeropad_right[3];
    uint64_t frequencyZeroPadding;

    /* convolution controls of the FFT machinery: ignored */
    uint64_t performConvolution;
    uint64_t conjugateConvolution;
    uint64_t crossPowerSpectrumNormalization;
    uint64_t coordinateFeatures;
    uint64_t matrixConvolution;
    uint64_t symmetricKernel;
    uint64_t numberKernels;
    uint64_t kernelConvolution;

    /* register over-utilisation and paging of the FFT machinery: ignored */
    uint64_t registerBoost;
    uint64_t registerBoostNonPow2;
    uint64_t registerBoost4Step;
    uint64_t swapTo3Stage4Step;
    uint64_t devicePageSize;
    uint64_t localPageSize;

    /* filled from the device at plan creation (may be overridden) */
    uint64_t computeCapabilityMajor;
    uint64_t computeCapabilityMinor;
    uint64_t maxComputeWorkGroupCount[3];
    uint64_t maxComputeWorkGroupSize[3];
    uint64_t maxThreadsNum;
    uint64_t sharedMemorySizeStatic;
    uint64_t sharedMemorySize;
    uint64_t sharedMemorySizePow2;
    uint64_t warpSize;
    uint64_t halfThreads;
    uint64_t allocateTempBuffer;
    uint64_t reorderFourStep;
    int64_t maxCodeLength;
    int64_t maxTempLength;
    uint64_t autoCustomBluesteinPaddingPattern;
    uint64_t useRaderUintLUT;
    uint64_t vendorID;
    hipEvent_t* stream_event;
    uint64_t streamCounter;
    uint64_t streamID;
    int64_t useStrict32BitAddress;

    /* ---- extension fields (no reference counterpart), appended ---- */
    int64_t offsetT;             /* fused step (jw_type 14): tridiagonal block T */
    int64_t autotune;            /* 0: built-in table, 1: run the autotuner at plan time */
    int64_t kernelVariant;       /* 0: automatic, 1: general kernel (long systems: the chunk kernel), 2: wave-per-system FP64 kernel */
    /* Fused block operators around a JW step (block codes as in `block`: layout +
     * 20 for accumulate + 100/200/300/400 for ScaleC/ScaleD/SphLaplA/SphLaplB).
     * inputBlock (layout 4) reads complex (interleaved re, im) systems from
     * inputBuffer and feeds their re/im halves to the step as a pair of real
     * systems; outputBlock (layout 7, optional accumulate) writes the result as
     * complex systems to outputBuffer, or, when 0, the step writes its real
     * systems there as usual.  0 / 0 = no fusion; an output block needs an input
     * block.  Input strides count complex elements (inputBufferStride), input
     * system 0 starts at inputBlockOffset; outputBufferStride and offsetSolution
     * count complex elements with an output block, doubles without.  FP64,
     * N <= 512; inputZeropad selects the rows read, outputZeropad those written.
     * No counterpart in the reference, where these are separate block plans
     * (PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_block.h). */
    int64_t inputBlock;
    int64_t outputBlock;
    int64_t inputBlockLshift;
    int64_t outputBlockLshift;
    double inputBlockScale;      /* the input operator's s */
    double outputBlockScale;     /* the output operator's s */
    int64_t inputBlockOffset;
} PfSolveConfiguration;

typedef struct {
    void** buffer;
    void** tempBuffer;
    void** inputBuffer;
    void** outputBuffer;
    void** kernel;
    uint64_t bufferOffset;
    uint64_t tempBufferOffset;
    uint64_t inputBufferOffset;
    uint64_t outputBufferOffset;
    uint64_t kernelOffset;

    uint64_t M_size;             /* RUNTIME_MSIZE - not implemented (as in the reference) */
    uint64_t offsetM;            /* RUNTIME_OFFSETM */
    uint64_t offsetV;            /* RUNTIME_OFFSETV */
    uint64_t offsetSolution;     /* RUNTIME_OFFSETSOLUTION */
    uint64_t inputZeropad[2];    /* RUNTIME_INPUTZEROPAD */
    uint64_t outputZeropad[2];   /* RUNTIME_OUTPUTZEROPAD */
    long double scaleC;          /* RUNTIME_SCALEC */
    uint64_t inputBufferStride;  /* RUNTIME_INPUTBUFFERSTRIDE */
    uint64_t outputBufferStride; /* RUNTIME_OUTPUTBUFFERSTRIDE */

    uint64_t batchSize;

    /* ---- extension fields, appended ---- */
    uint64_t offsetT;            /* RUNTIME_OFFSETT (fused step) */
} PfSolveLaunchParams;

typedef enum PfSolveResult {
    PFSOLVE_SUCCESS = 0,
    PFSOLVE_ERROR_MALLOC_FAILED = 1,
    PFSOLVE_ERROR_INSUFFICIENT_CODE_BUFFER = 2,
    PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER = 3,
    PFSOLVE_ERROR_PLAN_NOT_INITIALIZED = 4,
    PFSOLVE_ERROR_NULL_TEMP_PASSED = 5,
    PFSOLVE_ERROR_MATH_FAILED = 6,
    PFSOLVE_ERROR_INVALID_PHYSICAL_DEVICE = 1001,
    PFSOLVE_ERROR_INVALID_DEVICE = 1002,
    PFSOLVE_ERROR_INVALID_QUEUE = 1003,
    PFSOLVE_ERROR_INVALID_COMMAND_POOL = 1004,
    PFSOLVE_ERROR_INVALID_FENCE = 1005,
    PFSOLVE_ERROR_ONLY_FORWARD_FFT_INITIALIZED = 1006,
    PFSOLVE_ERROR_ONLY_INVERSE_FFT_INITIALIZED = 1007,
    PFSOLVE_ERROR_INVALID_CONTEXT = 1008,
    PFSOLVE_ERROR_INVALID_PLATFORM = 1009,
    PFSOLVE_ERROR_ENABLED_saveApplicationToString = 1010,
    PFSOLVE_ERROR_EMPTY_FILE = 1011,
    PFSOLVE_ERROR_EMPTY_FFTdim = 2001,
    PFSOLVE_ERROR_EMPTY_size = 2002,
    PFSOLVE_ERROR_EMPTY_bufferSize = 2003,
    PFSOLVE_ERROR_EMPTY_buffer = 2004,
    PFSOLVE_ERROR_EMPTY_tempBufferSize = 2005,
    PFSOLVE_ERROR_EMPTY_tempBuffer = 2006,
    PFSOLVE_ERROR_EMPTY_inputBufferSize = 2007,
    PFSOLVE_ERROR_EMPTY_inputBuffer = 2008,
    PFSOLVE_ERROR_EMPTY_outputBufferSize = 2009,
    PFSOLVE_ERROR_EMPTY_outputBuffer = 2010,
    PFSOLVE_ERROR_EMPTY_kernelSize = 2011,
    PFSOLVE_ERROR_EMPTY_kernel = 2012,
    PFSOLVE_ERROR_EMPTY_applicationString = 2013,
    PFSOLVE_ERROR_EMPTY_useCustomBluesteinPaddingPattern_arrays = 2014,
    PFSOLVE_ERROR_UNSUPPORTED_RADIX = 3001,
    PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH = 3002,
    PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH_R2C = 3003,
    PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH_DCT = 3004,
    PFSOLVE_ERROR_UNSUPPORTED_FFT_OMIT = 3005,
    PFSOLVE_ERROR_FAILED_TO_ALLOCATE = 4001,
    PFSOLVE_ERROR_FAILED_TO_MAP_MEMORY = 4002,
    PFSOLVE_ERROR_FAILED_TO_ALLOCATE_COMMAND_BUFFERS = 4003,
    PFSOLVE_ERROR_FAILED_TO_BEGIN_COMMAND_BUFFER = 4004,
    PFSOLVE_ERROR_FAILED_TO_END_COMMAND_BUFFER = 4005,
    PFSOLVE_ERROR_FAILED_TO_SUBMIT_QUEUE = 4006,
    PFSOLVE_ERROR_FAILED_TO_WAIT_FOR_FENCES = 4007,
    PFSOLVE_ERROR_FAILED_TO_RESET_FENCES = 4008,
    PFSOLVE_ERROR_FAILED_TO_CREATE_DESCRIPTOR_POOL = 4009,
    PFSOLVE_ERROR_FAILED_TO_CREATE_DESCRIPTOR_SET_LAYOUT = 4010,
    PFSOLVE_ERROR_FAILED_TO_ALLOCATE_DESCRIPTOR_SETS = 4011,
    PFSOLVE_ERROR_FAILED_TO_CREATE_PIPELINE_LAYOUT = 4012,
    PFSOLVE_ERROR_FAILED_SHADER_PREPROCESS = 4013,
    PFSOLVE_ERROR_FAILED_SHADER_PARSE = 4014,
    PFSOLVE_ERROR_FAILED_SHADER_LINK = 4015,
    PFSOLVE_ERROR_FAILED_SPIRV_GENERATE = 4016,
    PFSOLVE_ERROR_FAILED_TO_CREATE_SHADER_MODULE = 4017,
    PFSOLVE_ERROR_FAILED_TO_CREATE_INSTANCE = 4018,
    PFSOLVE_ERROR_FAILED_TO_SETUP_DEBUG_MESSENGER = 4019,
    PFSOLVE_ERROR_FAILED_TO_FIND_PHYSICAL_DEVICE = 4020,
    PFSOLVE_ERROR_FAILED_TO_CREATE_DEVICE = 4021,
    PFSOLVE_ERROR_FAILED_TO_CREATE_FENCE = 4022,
    PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_POOL = 4023,
    PFSOLVE_ERROR_FAILED_TO_CREATE_BUFFER = 4024,
    PFSOLVE_ERROR_FAILED_TO_ALLOCATE_MEMORY = 4025,
    PFSOLVE_ERROR_FAILED_TO_BIND_BUFFER_MEMORY = 4026,
    PFSOLVE_ERROR_FAILED_TO_FIND_MEMORY = 4027,
    PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE = 4028,
    PFSOLVE_ERROR_FAILED_TO_COPY = 4029,
    PFSOLVE_ERROR_FAILED_TO_CREATE_PROGRAM = 4030,
    PFSOLVE_ERROR_FAILED_TO_COMPILE_PROGRAM = 4031,
    PFSOLVE_ERROR_FAILED_TO_GET_CODE_SIZE = 4032,
    PFSOLVE_ERROR_FAILED_TO_GET_CODE = 4033,
    PFSOLVE_ERROR_FAILED_TO_DESTROY_PROGRAM = 4034,
    PFSOLVE_ERROR_FAILED_TO_LOAD_MODULE = 4035,
    PFSOLVE_ERROR_FAILED_TO_GET_FUNCTION = 4036,
    PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY = 4037,
    PFSOLVE_ERROR_FAILED_TO_MODULE_GET_GLOBAL = 4038,
    PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL = 4039,
    PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD = 4040,
    PFSOLVE_ERROR_FAILED_TO_ADD_NAME_EXPRESSION = 4041,
    PFSOLVE_ERROR_FAILED_TO_INITIALIZE = 4042,
    PFSOLVE_ERROR_FAILED_TO_SET_DEVICE_ID = 4043,
    PFSOLVE_ERROR_FAILED_TO_GET_DEVICE = 4044,
    PFSOLVE_ERROR_FAILED_TO_CREATE_CONTEXT = 4045,
    PFSOLVE_ERROR_FAILED_TO_CREATE_PIPELINE = 4046,
    PFSOLVE_ERROR_FAILED_TO_SET_KERNEL_ARG = 4047,
    PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_QUEUE = 4048,
    PFSOLVE_ERROR_FAILED_TO_RELEASE_COMMAND_QUEUE = 4049,
    PFSOLVE_ERROR_FAILED_TO_ENUMERATE_DEVICES = 4050,
    PFSOLVE_ERROR_FAILED_TO_GET_ATTRIBUTE = 4051,
    PFSOLVE_ERROR_FAILED_TO_CREATE_EVENT = 4052,
    PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_LIST = 4053,
    PFSOLVE_ERROR_FAILED_TO_DESTROY_COMMAND_LIST = 4054,
    PFSOLVE_ERROR_FAILED_TO_SUBMIT_BARRIER = 4055
} PfSolveResult;

/* Opaque plan state of one application. */
struct pfsolve_plan;

/* The reference's application (pfSolve_Structs.h:1281-1341, HIP build) field for
 * field; the FFT machinery's Bluestein / Rader members stay zero. */
typedef struct {
    PfSolveConfiguration configuration;
    struct pfsolve_plan* localFFTPlan;   /* opaque; NULL before init and after delete */
    struct pfsolve_plan* localFFTPlan_inverse;   /* always NULL: a JW plan has no inverse */
    char kernelName[200];
    uint64_t actualNumBatches;           /* size[2] of the plan */
    uint64_t firstAxis;
    uint64_t lastAxis;
    uint64_t useBluesteinFFT[3];
    void* bufferRaderUintLUT[3][4];
    void* bufferBluestein[3];
    void* bufferBluesteinFFT[3];
    void* bufferBluesteinIFFT[3];
    uint64_t bufferRaderUintLUTSize[3][4];
    uint64_t bufferBluesteinSize[3];
    void* applicationBluesteinString[3];
    uint64_t applicationBluesteinStringSize[3];
    uint64_t numRaderFFTPrimes;
    uint64_t rader_primes[30];
    uint64_t rader_buffer_size[30];
    void* raderFFTkernel[30];
    uint64_t applicationStringOffsetRader;
    uint64_t currentApplicationStringPos;
    uint64_t applicationStringSize;
    void* saveApplicationString;
} PfSolveApplication;

/* Creates a plan.  On failure the application is already torn down. */
PfSolveResult initializePfSolve(PfSolveApplication* app, PfSolveConfiguration inputLaunchConfiguration);
/* Enqueues one JW step (or block / dgbmv pass) on configuration.stream[0];
 * asynchronous: the caller synchronises.  `inverse` is ignored, as in the
 * reference's JW path. */
PfSolveResult PfSolveAppend(PfSolveApplication* app, int inverse, PfSolveLaunchParams* launchParams);
void deletePfSolve(PfSolveApplication* app);
PfSolveResult PfSolveSync(PfSolveApplication* app);
const char* getPfSolveErrorString(PfSolveResult result);
int PfSolveGetVersion(void);

#ifdef __cplusplus
}
#endif
#endif
