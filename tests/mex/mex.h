/*
 * mex.h of the mock MEX harness (tests/mex): the part of MATLAB's MEX and
 * matrix API that the engine's own gateway (mrilab_amd/mex/) calls, with
 * MATLAB's names, signatures and class numbers, so that the gateway compiles
 * unchanged against it. TEST INFRASTRUCTURE ONLY: it tests this project's
 * gateway; nothing of the reference is built with it.
 */
#ifndef MRS_TEST_MEX_H
#define MRS_TEST_MEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxSTRUCT_CLASS = 2,
    mxCHAR_CLASS = 4,
    mxDOUBLE_CLASS = 6,
    mxSINGLE_CLASS = 7,
    mxINT32_CLASS = 12
} mxClassID;

const mxArray *mexGetVariablePtr(const char *workspace, const char *name);
mxArray *mxGetField(const mxArray *pm, mwIndex index, const char *fieldname);
void *mxGetData(const mxArray *pm);
size_t mxGetNumberOfElements(const mxArray *pm);
mwSize mxGetNumberOfDimensions(const mxArray *pm);
const mwSize *mxGetDimensions(const mxArray *pm);
mxClassID mxGetClassID(const mxArray *pm);
int mxGetString(const mxArray *pm, char *str, mwSize strlen);
int mexEvalString(const char *command);
void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...);
int mxUnshareArray(mxArray *pa, int noDeepCopy);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif
#endif
