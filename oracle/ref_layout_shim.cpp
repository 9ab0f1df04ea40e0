// ref_layout_shim.cpp -- exports a few static inline host functions of the
// reference headers, compiled from the reference sources where they lie
// (oracle/Makefile target `ref`, output under oracle/_ref/).  TEST
// INFRASTRUCTURE ONLY: used to pin the oracle's restatement of the layout
// maths, the error-string table and the struct layouts against the
// reference's own code.
//   reorder_i0             PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142
//   get_ALT_stride         PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:30-64
//   reorder_i_alt          PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:144-219
//   getPfSolveErrorString  PfSolve/pfSolve_Structs/pfSolve_Structs.h:545-733
//   PfSolveGetVersion      PfSolve/pfSolve.h:101-103
//   PfSolveConfiguration, PfSolveLaunchParams, PfSolveApplication (HIP build)
//                          PfSolve/pfSolve_Structs/pfSolve_Structs.h:119-381, 383-449, 1281-1341
#include <cstddef>
#include <cstring>

#include "pfSolve.h"

// every member of the HIP build of each struct, in declaration order
#define CFG_FIELDS(X) \
    X(FFTdim) X(size) X(device) X(stream) X(num_streams) X(userTempBuffer) X(bufferNum) X(tempBufferNum) \
    X(inputBufferNum) X(outputBufferNum) X(kernelNum) X(bufferSize) X(tempBufferSize) X(inputBufferSize) \
    X(outputBufferSize) X(kernelSize) X(buffer) X(tempBuffer) X(inputBuffer) X(outputBuffer) X(kernel) \
    X(bufferOffset) X(tempBufferOffset) X(inputBufferOffset) X(outputBufferOffset) X(kernelOffset) \
    X(specifyOffsetsAtLaunch) X(s_dx) X(s_dy) X(s_dz) X(s_dt_D) X(compute_Pf) X(finiteDifferences) \
    X(logicBlock) X(upperBanded) X(jw_type) X(block) X(lshift) X(LDA) X(KU) X(KL) X(ALT_dgbmv) X(Msplit) \
    X(jw_control_bitmask) X(M_size) X(offsetM) X(offsetV) X(offsetSolution) X(inputZeropad) \
    X(outputZeropad) X(scaleC) X(coalescedMemory) X(aimThreads) X(numSharedBanks) \
    X(inverseReturnToInputBuffer) X(numberBatches) X(useUint64) X(omitDimension) X(performBandwidthBoost) \
    X(groupedBatch) X(doublePrecision) X(quadDoubleDoublePrecision) X(quadDoubleDoublePrecisionDoubleMemory) \
    X(halfPrecision) X(halfPrecisionMemoryOnly) X(doublePrecisionFloatMemory) X(performR2C) X(performDCT) \
    X(disableMergeSequencesR2C) X(normalize) X(disableReorderFourStep) X(useLUT) X(useLUT_4step) \
    X(makeForwardPlanOnly) X(makeInversePlanOnly) X(bufferStride) X(isInputFormatted) X(isOutputFormatted) \
    X(inputBufferStride) X(outputBufferStride) X(considerAllAxesStrided) X(keepShaderCode) \
    X(printMemoryLayout) X(disableCaching) X(saveApplicationToString) X(loadApplicationFromString) \
    X(loadApplicationString) X(disableSetLocale) X(fixMaxRadixBluestein) X(forceBluesteinSequenceSize) \
    X(useCustomBluesteinPaddingPattern) X(primeSizes) X(paddedSizes) X(fixMinRaderPrimeMult) \
    X(fixMaxRaderPrimeMult) X(fixMinRaderPrimeFFT) X(fixMaxRaderPrimeFFT) X(numConsecutiveJWIterations) \
    X(useMultipleInputBuffers) X(performZeropadding) X(fft_zeropad_left) X(fft_zeropad_right) \
    X(frequencyZeroPadding) X(performConvolution) X(conjugateConvolution) X(crossPowerSpectrumNormalization) \
    X(coordinateFeatures) X(matrixConvolution) X(symmetricKernel) X(numberKernels) X(kernelConvolution) \
    X(registerBoost) X(registerBoostNonPow2) X(registerBoost4Step) X(swapTo3Stage4Step) X(devicePageSize) \
    X(localPageSize) X(computeCapabilityMajor) X(computeCapabilityMinor) X(maxComputeWorkGroupCount) \
    X(maxComputeWorkGroupSize) X(maxThreadsNum) X(sharedMemorySizeStatic) X(sharedMemorySize) \
    X(sharedMemorySizePow2) X(warpSize) X(halfThreads) X(allocateTempBuffer) X(reorderFourStep) \
    X(maxCodeLength) X(maxTempLength) X(autoCustomBluesteinPaddingPattern) X(useRaderUintLUT) X(vendorID) \
    X(stream_event) X(streamCounter) X(streamID) X(useStrict32BitAddress)

#define LP_FIELDS(X) \
    X(buffer) X(tempBuffer) X(inputBuffer) X(outputBuffer) X(kernel) X(bufferOffset) X(tempBufferOffset) \
    X(inputBufferOffset) X(outputBufferOffset) X(kernelOffset) X(M_size) X(offsetM) X(offsetV) \
    X(offsetSolution) X(inputZeropad) X(outputZeropad) X(scaleC) X(inputBufferStride) X(outputBufferStride) \
    X(batchSize)

#define APP_FIELDS(X) \
    X(configuration) X(localFFTPlan) X(localFFTPlan_inverse) X(kernelName) X(actualNumBatches) X(firstAxis) \
    X(lastAxis) X(useBluesteinFFT) X(bufferRaderUintLUT) X(bufferBluestein) X(bufferBluesteinFFT) \
    X(bufferBluesteinIFFT) X(bufferRaderUintLUTSize) X(bufferBluesteinSize) X(applicationBluesteinString) \
    X(applicationBluesteinStringSize) X(numRaderFFTPrimes) X(rader_primes) X(rader_buffer_size) \
    X(raderFFTkernel) X(applicationStringOffsetRader) X(currentApplicationStringPos) X(applicationStringSize) \
    X(saveApplicationString)

namespace {
struct field { const char* name; long off; long size; };
#define CFG_ENTRY(f) {#f, (long)offsetof(PfSolveConfiguration, f), (long)sizeof(((PfSolveConfiguration*)0)->f)},
#define LP_ENTRY(f) {#f, (long)offsetof(PfSolveLaunchParams, f), (long)sizeof(((PfSolveLaunchParams*)0)->f)},
#define APP_ENTRY(f) {#f, (long)offsetof(PfSolveApplication, f), (long)sizeof(((PfSolveApplication*)0)->f)},
const field k_cfg[] = {CFG_FIELDS(CFG_ENTRY)};
const field k_lp[] = {LP_FIELDS(LP_ENTRY)};
const field k_app[] = {APP_FIELDS(APP_ENTRY)};

const field* table(int which, long* n)
{
    switch (which) {
    case 0: *n = sizeof(k_cfg) / sizeof(k_cfg[0]); return k_cfg;
    case 1: *n = sizeof(k_lp) / sizeof(k_lp[0]); return k_lp;
    default: *n = sizeof(k_app) / sizeof(k_app[0]); return k_app;
    }
}
}  // namespace

extern "C" {
int ref_reorder_i0(int i, int M_size, int M_size_align, int Msplit0)
{
    return reorder_i0(i, M_size, M_size_align, Msplit0);
}
int ref_get_ALT_stride(int M_size_align, int Msplit0) { return get_ALT_stride(M_size_align, Msplit0); }
int ref_reorder_i_alt(int i, int M_size, int M_size_align, int Msplit0)
{
    return reorder_i_alt(i, M_size, M_size_align, Msplit0);
}
const char* ref_error_string(int code) { return getPfSolveErrorString((PfSolveResult)code); }
int ref_version(void) { return PfSolveGetVersion(); }

// which: 0 = PfSolveConfiguration, 1 = PfSolveLaunchParams, 2 = PfSolveApplication
long ref_struct_size(int which)
{
    return which == 0 ? (long)sizeof(PfSolveConfiguration)
                      : which == 1 ? (long)sizeof(PfSolveLaunchParams) : (long)sizeof(PfSolveApplication);
}
long ref_field_count(int which)
{
    long n = 0;
    table(which, &n);
    return n;
}
// name, offset and size of member i (declaration order); NULL past the end
const char* ref_field(int which, long i, long* off, long* size)
{
    long n = 0;
    const field* t = table(which, &n);
    if (i < 0 || i >= n) return nullptr;
    *off = t[i].off;
    *size = t[i].size;
    return t[i].name;
}
}
