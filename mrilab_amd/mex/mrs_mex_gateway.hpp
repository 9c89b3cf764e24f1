// MEX gateway of MRiLab's scan entry points on the C ABI of mrilab_amd.h.
//
// MRiLab calls its scan MEX functions with no arguments
// (Src/Main/SimuPanel.m:800-897, BatchSim.m:160); they read and write the
// seven global structs VObj, VMag, VCoi, VCtl, VSeq, VVar and VSig in place
// (Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:70-210). The gateway does the
// same on the engine's C structs:
//   * binds every field the engine reads to the MATLAB array's data,
//     refusing a field of the wrong class with mexErrMsgIdAndTxt (the
//     reference reinterprets it silently), with the classes of
//     Src/Main/DoDataTypeConv.m;
//   * runs the engine, whose outputs land in the MATLAB arrays (VObj.M*,
//     VSig.*) and in the C copy of VVar's scalars, which it writes back;
//   * on an Ext code the engine hands over, syncs VVar to MATLAB, evaluates
//     DoGPUFetch (GPU entry points) and DoExtPlugin as the reference does
//     (DoScanAtGPU.cu:566,592; DoScanAtCPU.cpp:360), then re-fetches every
//     array pointer, since MATLAB may have reallocated the globals it edited
//     (DoScanAtCPU.cpp:366-405), and reads VVar back.
// Each DoXScanAtY.cpp beside this header defines mexFunction with one call.
// Build (MATLAB): mex DoScanAtGPU.cpp -I<repo>/include -L<repo>/mrilab_amd -lmrilab_amd
#pragma once

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "mex.h"
#include "mrilab_amd.h"

namespace mrs_mex {

// Inside an Ext callback the engine's frames are on the stack, so an error
// there must not leave through mexErrMsgIdAndTxt: it is thrown as CbError,
// caught by the callback, which returns nonzero; the engine then fails and
// the gateway reports the message after the engine has returned.
struct CbError {
    std::string msg;
};
inline bool &in_callback() {
    static thread_local bool on = false;
    return on;
}
inline std::string &callback_error() {
    static thread_local std::string msg;
    return msg;
}

[[noreturn]] inline void error(const char *id, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (in_callback()) throw CbError{buf};
    mexErrMsgIdAndTxt(id, "%s", buf);
    throw CbError{buf};  // mexErrMsgIdAndTxt does not return
}

inline mxArray *global(const char *name) {
    mxArray *a = const_cast<mxArray *>(mexGetVariablePtr("global", name));
    if (!a) error("MRiLab:global", "global %s is not defined", name);
    return a;
}

inline const char *class_name(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS: return "double";
        case mxSINGLE_CLASS: return "single";
        case mxINT32_CLASS: return "int32";
        case mxCHAR_CLASS: return "char";
        default: return "another class";
    }
}

// The field's array, of class cls with at least n elements (n > 0), else a
// MATLAB error naming the field.
inline mxArray *field(const mxArray *s, const char *sname, const char *name, mxClassID cls, size_t n = 0) {
    mxArray *f = mxGetField(s, 0, name);
    if (!f) error("MRiLab:field", "%s.%s is missing", sname, name);
    if (mxGetClassID(f) != cls)
        error("MRiLab:dtype", "%s.%s must be %s, not %s", sname, name, class_name(cls), class_name(mxGetClassID(f)));
    if (n && mxGetNumberOfElements(f) < n)
        error("MRiLab:size", "%s.%s has %d elements, at least %d expected", sname, name, (int)mxGetNumberOfElements(f),
              (int)n);
    return f;
}

template <class T>
inline T *data(const mxArray *s, const char *sname, const char *name, mxClassID cls, size_t n = 0) {
    return (T *)mxGetData(field(s, sname, name, cls, n));
}

template <class T>
inline T scalar(const mxArray *s, const char *sname, const char *name, mxClassID cls) {
    return *data<T>(s, sname, name, cls, 1);
}

inline bool is_on(const mxArray *s, const char *name) {  // VCtl.MultiTransmit = 'on' / 'off'
    mxArray *f = mxGetField(s, 0, name);
    if (!f || mxGetClassID(f) != mxCHAR_CLASS) return false;
    char buf[8] = {0};
    mxGetString(f, buf, sizeof buf);
    return std::strcmp(buf, "on") == 0;
}

// The engine's view of the globals, and the sizes it needs.
struct Scan {
    mrs_vobj vobj;
    mrs_vmag vmag;
    mrs_vcoi vcoi;
    mrs_vctl vctl;
    mrs_vseq vseq;
    mrs_vvar vvar;
    mrs_vsig vsig;
    size_t nvox = 0, ntot = 0;
    int k_planes = 0;  // multi-pool: planes of VObj.K in the entry point's layout, 0 = single pool
};

// VObj, VMag and VCoi arrays (bound again after every Ext callback).
inline void bind_maps(Scan &S) {
    const mxArray *O = global("VObj"), *M = global("VMag"), *C = global("VCoi");
    const size_t nv = S.nvox, Ty = (size_t)S.vobj.TypeNum, Tx = (size_t)S.vcoi.TxCoilNum;
    const size_t Rx = (size_t)S.vcoi.RxCoilNum;
    S.vobj.Mx = data<float>(O, "VObj", "Mx", mxSINGLE_CLASS, S.ntot);
    S.vobj.My = data<float>(O, "VObj", "My", mxSINGLE_CLASS, S.ntot);
    S.vobj.Mz = data<float>(O, "VObj", "Mz", mxSINGLE_CLASS, S.ntot);
    S.vobj.Rho = data<float>(O, "VObj", "Rho", mxSINGLE_CLASS, nv * Ty);
    S.vobj.T1 = data<float>(O, "VObj", "T1", mxSINGLE_CLASS, nv * Ty);
    S.vobj.T2 = data<float>(O, "VObj", "T2", mxSINGLE_CLASS, nv * Ty);
    if (S.k_planes) {
        S.vobj.K = data<float>(O, "VObj", "K", mxSINGLE_CLASS, nv * (size_t)S.k_planes);
        S.vobj.TypeFlag = mxGetField(O, 0, "TypeFlag") ? data<double>(O, "VObj", "TypeFlag", mxDOUBLE_CLASS, Ty)
                                                       : nullptr;
    }
    S.vmag.dB0 = data<float>(M, "VMag", "dB0", mxSINGLE_CLASS, nv);
    S.vmag.dWRnd = data<float>(M, "VMag", "dWRnd", mxSINGLE_CLASS, S.ntot);
    S.vmag.Gzgrid = data<float>(M, "VMag", "Gzgrid", mxSINGLE_CLASS, nv);
    S.vmag.Gygrid = data<float>(M, "VMag", "Gygrid", mxSINGLE_CLASS, nv);
    S.vmag.Gxgrid = data<float>(M, "VMag", "Gxgrid", mxSINGLE_CLASS, nv);
    S.vcoi.TxCoilmg = data<float>(C, "VCoi", "TxCoilmg", mxSINGLE_CLASS, nv * Tx);
    S.vcoi.TxCoilpe = data<float>(C, "VCoi", "TxCoilpe", mxSINGLE_CLASS, nv * Tx);
    S.vcoi.RxCoilx = data<float>(C, "VCoi", "RxCoilx", mxSINGLE_CLASS, S.vcoi.RxCoilDefault != 0 ? nv : nv * Rx);
    S.vcoi.RxCoily = data<float>(C, "VCoi", "RxCoily", mxSINGLE_CLASS, S.vcoi.RxCoilDefault != 0 ? nv : nv * Rx);
}

// VSig arrays (bound again after every Ext callback).
inline void bind_sig(Scan &S) {
    const mxArray *G = global("VSig");
    const size_t N = (size_t)S.vsig.SignalNum;
    const size_t n = N * (size_t)S.vobj.TypeNum * (size_t)S.vcoi.RxCoilNum;
    S.vsig.Sx = data<double>(G, "VSig", "Sx", mxDOUBLE_CLASS, n);
    S.vsig.Sy = data<double>(G, "VSig", "Sy", mxDOUBLE_CLASS, n);
    S.vsig.Kz = data<double>(G, "VSig", "Kz", mxDOUBLE_CLASS, N);
    S.vsig.Ky = data<double>(G, "VSig", "Ky", mxDOUBLE_CLASS, N);
    S.vsig.Kx = data<double>(G, "VSig", "Kx", mxDOUBLE_CLASS, N);
    if (S.vctl.RunMode == 1) {
        const size_t nuts = (size_t)S.vseq.n_uts;
        S.vsig.Mx = data<float>(G, "VSig", "Mx", mxSINGLE_CLASS, S.ntot * nuts);
        S.vsig.My = data<float>(G, "VSig", "My", mxSINGLE_CLASS, S.ntot * nuts);
        S.vsig.Mz = data<float>(G, "VSig", "Mz", mxSINGLE_CLASS, S.ntot * nuts);
        S.vsig.Muts = data<double>(G, "VSig", "Muts", mxDOUBLE_CLASS, nuts);
    }
}

// VVar's scalars into the C struct; its rf arrays are copied into the
// struct's arrays (the engine keeps its own copies during a scan).
inline void vvar_from_matlab(const mxArray *V, mrs_vvar *v, int tx) {
    struct {
        const char *name;
        double *dst;
    } D[] = {{"t", &v->t},           {"dt", &v->dt},         {"rfCoil", &v->rfCoil}, {"rfRef", &v->rfRef},
             {"GzAmp", &v->GzAmp},   {"GyAmp", &v->GyAmp},   {"GxAmp", &v->GxAmp},   {"ADC", &v->ADC},
             {"Ext", &v->Ext},       {"Kz", &v->Kz},         {"Ky", &v->Ky},         {"Kx", &v->Kx},
             {"gpuFetch", &v->gpuFetch}};
    for (auto &d : D) *d.dst = scalar<double>(V, "VVar", d.name, mxDOUBLE_CLASS);
    // Plugin_LockK.m creates the keepers on first use
    struct {
        const char *name;
        double *dst;
    } K[] = {{"KzKeeper", &v->KzKeeper}, {"KyKeeper", &v->KyKeeper}, {"KxKeeper", &v->KxKeeper}};
    for (auto &k : K) *k.dst = mxGetField(V, 0, k.name) ? scalar<double>(V, "VVar", k.name, mxDOUBLE_CLASS) : 0.0;
    struct {
        const char *name;
        int32_t *dst;
    } I[] = {{"utsi", &v->utsi}, {"rfi", &v->rfi},   {"Gzi", &v->Gzi},   {"Gyi", &v->Gyi},
             {"Gxi", &v->Gxi},   {"ADCi", &v->ADCi}, {"Exti", &v->Exti}, {"TRCount", &v->TRCount}};
    for (auto &i : I) *i.dst = scalar<int32_t>(V, "VVar", i.name, mxINT32_CLASS);
    const double *a = data<double>(V, "VVar", "rfAmp", mxDOUBLE_CLASS, (size_t)tx);
    const double *p = data<double>(V, "VVar", "rfPhase", mxDOUBLE_CLASS, (size_t)tx);
    const double *f = data<double>(V, "VVar", "rfFreq", mxDOUBLE_CLASS, (size_t)tx);
    if (v->rfAmp != a) std::memcpy(v->rfAmp, a, sizeof(double) * tx);
    if (v->rfPhase != p) std::memcpy(v->rfPhase, p, sizeof(double) * tx);
    if (v->rfFreq != f) std::memcpy(v->rfFreq, f, sizeof(double) * tx);
}

// The C struct's VVar state into the MATLAB fields, in place.
inline void vvar_to_matlab(const mrs_vvar *v, mxArray *V, int tx) {
    struct {
        const char *name;
        double val;
    } D[] = {{"t", v->t},         {"dt", v->dt},         {"rfCoil", v->rfCoil}, {"rfRef", v->rfRef},
             {"GzAmp", v->GzAmp}, {"GyAmp", v->GyAmp},   {"GxAmp", v->GxAmp},   {"ADC", v->ADC},
             {"Ext", v->Ext},     {"Kz", v->Kz},         {"Ky", v->Ky},         {"Kx", v->Kx},
             {"gpuFetch", v->gpuFetch}};
    for (auto &d : D) *data<double>(V, "VVar", d.name, mxDOUBLE_CLASS, 1) = d.val;
    struct {
        const char *name;
        double val;
    } K[] = {{"KzKeeper", v->KzKeeper}, {"KyKeeper", v->KyKeeper}, {"KxKeeper", v->KxKeeper}};
    for (auto &k : K)
        if (mxGetField(V, 0, k.name)) *data<double>(V, "VVar", k.name, mxDOUBLE_CLASS, 1) = k.val;
    struct {
        const char *name;
        int32_t val;
    } I[] = {{"utsi", v->utsi}, {"rfi", v->rfi},   {"Gzi", v->Gzi},   {"Gyi", v->Gyi},
             {"Gxi", v->Gxi},   {"ADCi", v->ADCi}, {"Exti", v->Exti}, {"TRCount", v->TRCount}};
    for (auto &i : I) *data<int32_t>(V, "VVar", i.name, mxINT32_CLASS, 1) = i.val;
    double *a = data<double>(V, "VVar", "rfAmp", mxDOUBLE_CLASS, (size_t)tx);
    double *p = data<double>(V, "VVar", "rfPhase", mxDOUBLE_CLASS, (size_t)tx);
    double *f = data<double>(V, "VVar", "rfFreq", mxDOUBLE_CLASS, (size_t)tx);
    if (v->rfAmp != a) std::memcpy(a, v->rfAmp, sizeof(double) * tx);
    if (v->rfPhase != p) std::memcpy(p, v->rfPhase, sizeof(double) * tx);
    if (v->rfFreq != f) std::memcpy(f, v->rfFreq, sizeof(double) * tx);
}

// Everything the engine reads, bound to the globals. k_planes: planes of
// VObj.K in the multi-pool entry point's layout (0 for a single pool).
inline void bind(Scan &S, bool gpu, int model) {
    // break copy-on-write so that the in-place writes reach the globals
    // (DoScanAtCPU.cpp:118-124)
    const char *names[7] = {"VObj", "VMag", "VCoi", "VCtl", "VVar", "VSeq", "VSig"};
    for (const char *n : names) mxUnshareArray(global(n), true);
    const mxArray *O = global("VObj"), *C = global("VCoi"), *T = global("VCtl"), *Q = global("VSeq");
    const mxArray *G = global("VSig");
    std::memset(&S.vobj, 0, sizeof S.vobj);
    std::memset(&S.vmag, 0, sizeof S.vmag);
    std::memset(&S.vcoi, 0, sizeof S.vcoi);
    std::memset(&S.vctl, 0, sizeof S.vctl);
    std::memset(&S.vseq, 0, sizeof S.vseq);
    std::memset(&S.vsig, 0, sizeof S.vsig);
    S.vobj.Gyro = scalar<double>(O, "VObj", "Gyro", mxDOUBLE_CLASS);
    S.vobj.SpinNum = scalar<int32_t>(O, "VObj", "SpinNum", mxINT32_CLASS);
    S.vobj.TypeNum = scalar<int32_t>(O, "VObj", "TypeNum", mxINT32_CLASS);
    const mxArray *mz = field(O, "VObj", "Mz", mxSINGLE_CLASS);
    const mwSize nd = mxGetNumberOfDimensions(mz);
    const mwSize *dims = mxGetDimensions(mz);
    S.vobj.dims[0] = (int32_t)dims[0];
    S.vobj.dims[1] = nd > 1 ? (int32_t)dims[1] : 1;
    S.vobj.dims[2] = nd > 2 ? (int32_t)dims[2] : 1;  // a 2-D object has one slice (DoScanAtCPU.cpp:216-222)
    S.nvox = (size_t)S.vobj.dims[0] * S.vobj.dims[1] * S.vobj.dims[2];
    S.ntot = S.nvox * (size_t)S.vobj.SpinNum * (size_t)S.vobj.TypeNum;
    if (model) {
        const int Ty = S.vobj.TypeNum;
        S.k_planes = model == 3 ? 2 * (Ty - 1) : Ty * Ty;  // CEST: 2*(TypeNum-1) planes
    }
    S.vcoi.TxCoilNum = scalar<int32_t>(C, "VCoi", "TxCoilNum", mxINT32_CLASS);
    S.vcoi.RxCoilNum = scalar<int32_t>(C, "VCoi", "RxCoilNum", mxINT32_CLASS);
    S.vcoi.TxCoilDefault = scalar<double>(C, "VCoi", "TxCoilDefault", mxDOUBLE_CLASS);
    S.vcoi.RxCoilDefault = scalar<double>(C, "VCoi", "RxCoilDefault", mxDOUBLE_CLASS);
    S.vctl.CS = data<double>(T, "VCtl", "CS", mxDOUBLE_CLASS, (size_t)S.vobj.TypeNum);
    S.vctl.TRNum = scalar<int32_t>(T, "VCtl", "TRNum", mxINT32_CLASS);
    S.vctl.RunMode = scalar<int32_t>(T, "VCtl", "RunMode", mxINT32_CLASS);
    S.vctl.MaxThreadNum = scalar<int32_t>(T, "VCtl", "MaxThreadNum", mxINT32_CLASS);
    S.vctl.GPUIndex = gpu ? scalar<int32_t>(T, "VCtl", "GPUIndex", mxINT32_CLASS) : 0;
    S.vctl.MultiTransmit = is_on(T, "MultiTransmit") ? 1 : 0;
    S.vctl.MasterTxCoil = mxGetField(T, 0, "MasterTxCoil")
                              ? (int32_t)std::lround(scalar<double>(T, "VCtl", "MasterTxCoil", mxDOUBLE_CLASS))
                              : 1;
    const char *lines[12] = {"utsLine", "tsLine", "rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine",
                             "GzAmpLine", "GyAmpLine", "GxAmpLine", "ADCLine", "ExtLine", "flagsLine"};
    const double **dst[12] = {&S.vseq.utsLine,   &S.vseq.tsLine,    &S.vseq.rfAmpLine, &S.vseq.rfPhaseLine,
                              &S.vseq.rfFreqLine, &S.vseq.rfCoilLine, &S.vseq.GzAmpLine, &S.vseq.GyAmpLine,
                              &S.vseq.GxAmpLine,  &S.vseq.ADCLine,    &S.vseq.ExtLine,   &S.vseq.flagsLine};
    for (int i = 0; i < 12; i++) *dst[i] = data<double>(Q, "VSeq", lines[i], mxDOUBLE_CLASS);
    auto numel = [&](const char *n) { return (int64_t)mxGetNumberOfElements(mxGetField(Q, 0, n)); };
    S.vseq.n_ts = numel("tsLine");
    S.vseq.n_uts = numel("utsLine");
    S.vseq.n_rf = numel("rfAmpLine");
    S.vseq.n_Gz = numel("GzAmpLine");
    S.vseq.n_Gy = numel("GyAmpLine");
    S.vseq.n_Gx = numel("GxAmpLine");
    S.vseq.n_ADC = numel("ADCLine");
    S.vseq.n_Ext = numel("ExtLine");
    if (numel("flagsLine") != 6 * S.vseq.n_ts) error("MRiLab:size", "VSeq.flagsLine must be 6 x numel(VSeq.tsLine)");
    if (numel("rfCoilLine") != S.vseq.n_rf)
        error("MRiLab:size", "VSeq.rfCoilLine and VSeq.rfAmpLine differ in length");
    S.vsig.SignalNum = scalar<int32_t>(G, "VSig", "SignalNum", mxINT32_CLASS);
    bind_maps(S);
    bind_sig(S);
    std::memset(&S.vvar, 0, sizeof S.vvar);
    const mxArray *V = global("VVar");
    const int tx = S.vcoi.TxCoilNum;
    // the rf arrays are MATLAB's (written in place at the end)
    S.vvar.rfAmp = data<double>(V, "VVar", "rfAmp", mxDOUBLE_CLASS, (size_t)tx);
    S.vvar.rfPhase = data<double>(V, "VVar", "rfPhase", mxDOUBLE_CLASS, (size_t)tx);
    S.vvar.rfFreq = data<double>(V, "VVar", "rfFreq", mxDOUBLE_CLASS, (size_t)tx);
    vvar_from_matlab(V, &S.vvar, tx);
}

struct CallbackCtx {
    Scan *scan;
    bool gpu;
};

struct CallbackScope {
    CallbackScope() { in_callback() = true; }
    ~CallbackScope() { in_callback() = false; }
};

// DoGPUFetch.m sets VVar.gpuFetch from VVar.Ext; the engine fetches the
// state when it is nonzero.
inline int gpu_fetch_cb(void *user, double ext, int *fetch) {
    CallbackScope scope;
    try {
        mxArray *V = global("VVar");
        *data<double>(V, "VVar", "Ext", mxDOUBLE_CLASS, 1) = ext;
        if (mexEvalString("DoGPUFetch") != 0) {
            callback_error() = "DoGPUFetch failed";
            return 1;
        }
        *fetch = scalar<double>(global("VVar"), "VVar", "gpuFetch", mxDOUBLE_CLASS) != 0;
        return 0;
    } catch (const CbError &e) {
        callback_error() = e.msg;
        return 1;
    }
    (void)user;
}

// DoExtPlugin.m with VVar synced both ways and every pointer re-fetched.
inline int ext_cb(void *user, double ext, mrs_vvar *vvar, mrs_ext_state *st) {
    CallbackScope scope;
    try {
        CallbackCtx *c = (CallbackCtx *)user;
        Scan &S = *c->scan;
        mxArray *V = global("VVar");
        const int tx = S.vcoi.TxCoilNum;
        vvar_to_matlab(vvar, V, tx);
        *data<double>(V, "VVar", "Ext", mxDOUBLE_CLASS, 1) = ext;
        if (mexEvalString("DoExtPlugin") != 0) {
            callback_error() = "Extended process encounters ERROR!";  // DoScanAtCPU.cpp:361-363
            return 1;
        }
        // DoScanAtCPU.cpp:366-405: MATLAB may have moved any global array
        V = global("VVar");
        bind_maps(S);
        bind_sig(S);
        // the entry point's VObj and VSig are S's: bound in place above;
        // VMag and VCoi are the engine's copies
        *st->vmag = S.vmag;
        st->vcoi->TxCoilmg = S.vcoi.TxCoilmg;
        st->vcoi->TxCoilpe = S.vcoi.TxCoilpe;
        st->vcoi->RxCoilx = S.vcoi.RxCoilx;
        st->vcoi->RxCoily = S.vcoi.RxCoily;
        // the engine writes the final rf state into the entry point's VVar
        // arrays: point them at the current MATLAB arrays
        S.vvar.rfAmp = data<double>(V, "VVar", "rfAmp", mxDOUBLE_CLASS, (size_t)tx);
        S.vvar.rfPhase = data<double>(V, "VVar", "rfPhase", mxDOUBLE_CLASS, (size_t)tx);
        S.vvar.rfFreq = data<double>(V, "VVar", "rfFreq", mxDOUBLE_CLASS, (size_t)tx);
        vvar_from_matlab(V, vvar, tx);
        return 0;
    } catch (const CbError &e) {
        callback_error() = e.msg;
        return 1;
    }
}

typedef mrs_status (*entry_fn)(mrs_vobj *, const mrs_vmag *, const mrs_vcoi *, const mrs_vctl *, const mrs_vseq *,
                               mrs_vvar *, mrs_vsig *, const mrs_callbacks *, const mrs_options *);

// The whole MEX call: bind, run with callbacks, write VVar back.
// gpu: the GPU entry points evaluate DoGPUFetch before DoExtPlugin;
// model: 0 single pool, 1 GM, 2 MT, 3 CEST, 4 ME (VObj.K layout).
inline void run(entry_fn fn, bool gpu, int model) {
    Scan S;
    bind(S, gpu, model);
    CallbackCtx ctx{&S, gpu};
    mrs_callbacks cb;
    cb.gpu_fetch = gpu ? gpu_fetch_cb : nullptr;
    cb.ext = ext_cb;
    cb.user = &ctx;
    mrs_options opt;
    std::memset(&opt, 0, sizeof opt);
    opt.model = model;
    callback_error().clear();
    const mrs_status rc = fn(&S.vobj, &S.vmag, &S.vcoi, &S.vctl, &S.vseq, &S.vvar, &S.vsig, &cb, &opt);
    if (rc == MRS_ERR_CALLBACK && !callback_error().empty())
        error("MRiLab:callback", "%s", callback_error().c_str());
    if (rc != MRS_OK) error("MRiLab:engine", "%s", mrs_last_error());
    vvar_to_matlab(&S.vvar, global("VVar"), S.vcoi.TxCoilNum);
}

}  // namespace mrs_mex
