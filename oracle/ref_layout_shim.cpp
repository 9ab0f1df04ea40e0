// ref_layout_shim.cpp -- exports a few static inline host functions of the
// reference headers, compiled from the reference sources where they lie
// (oracle/Makefile target `ref`, output under oracle/_ref/).  TEST
// INFRASTRUCTURE ONLY: used to pin the oracle's restatement of the layout
// maths and the error-string table against the reference's own code.
//   reorder_i0             PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142
//   get_ALT_stride         PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:30-64
//   getPfSolveErrorString  PfSolve/pfSolve_Structs/pfSolve_Structs.h:545-733
//   PfSolveGetVersion      PfSolve/pfSolve.h:101-103
#include "pfSolve.h"

extern "C" {
int ref_reorder_i0(int i, int M_size, int M_size_align, int Msplit0)
{
    return reorder_i0(i, M_size, M_size_align, Msplit0);
}
int ref_get_ALT_stride(int M_size_align, int Msplit0) { return get_ALT_stride(M_size_align, Msplit0); }
const char* ref_error_string(int code) { return getPfSolveErrorString((PfSolveResult)code); }
int ref_version(void) { return PfSolveGetVersion(); }
}
