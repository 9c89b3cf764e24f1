// DoScanAtCPU: MRiLab's MEX entry point (Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:70), on the engine's C ABI
// (mrs_DoScanAtCPU). Exports only mexFunction; MATLAB calls it with no arguments.
#include "mrs_mex_gateway.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    (void)nlhs;
    (void)plhs;
    (void)nrhs;
    (void)prhs;
    mrs_mex::run(mrs_DoScanAtCPU, false, 0);
}
