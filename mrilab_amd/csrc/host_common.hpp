// Host-side helpers shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mrilab_amd.h"

namespace mrs {

// Records msg as mrs_last_error() (per calling thread) and returns st.
mrs_status fail(mrs_status st, const std::string &msg);

// Validates the seven scan structs (NULL pointers, dims, counts, line sizes).
mrs_status check_args(const mrs_vobj *vo, const mrs_vmag *vm, const mrs_vcoi *vc, const mrs_vctl *vt,
                      const mrs_vseq *vs, const mrs_vvar *vv, const mrs_vsig *vg);

// Writes the decoder's final VVar into the caller's struct, keeping the
// caller's rfAmp/rfPhase/rfFreq arrays and filling them with rf[0..2].
void copy_vvar_out(const mrs_vvar &src, const std::vector<double> *rf, int tx, mrs_vvar *dst);

// The k-space a plan's last run left in device memory ([Ty][Rx][n] fp64,
// as VSig.Sx/Sy), after the plan's stream has drained.
struct PlanSignal {
    const double *sx = nullptr, *sy = nullptr;
    int64_t n = 0;
    int types = 0, rx = 0, device = 0;
    hipStream_t stream = nullptr;
};
mrs_status plan_signal(mrs_plan *plan, PlanSignal *out);

}  // namespace mrs

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::mrs::fail(e_ == hipErrorOutOfMemory ? MRS_ERR_NOMEM : MRS_ERR_DEVICE,      \
                               std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)
