// Local SAR on gfx950 (Lib/src/engine/CPUEngine/DoCalSARAtCPU.cpp, SARKernel.h).
//
// Host: sar_decode runs the reference's event loop (DoCalSARAtCPU.cpp:191-355:
// coincident-event flags, rf hold of the smaller magnitude, VVar reset,
// dt == 0 skip, dt < 0 TR boundary, Muts, sampling before and after every
// step). Everything but the per-voxel power sums depends only on the
// sequence, so the loop updates VVar/Muts/tRealSample/SARi on the host and
// emits a record stream: STEP (fp32 dt and each coil's fp32 amplitude, only
// when some amplitude is nonzero: SARKernel.h:17-22 skips the rest) and
// SAMPLE (slot of VSig.SAR).
// Device: sar_kernel gives each thread one voxel with its E fields (3 x Tx),
// ECon and running SAR in registers and walks the records wave-uniformly,
// in the reference's fp32 operation order with contraction off
// (SARKernel.h:34-57, DoCalSARAtCPU.cpp:225-233), so the results equal the
// reference arithmetic bit for bit. Per voxel-step: 3 x (Tx multiply-adds)
// + 3 squares, 3 scalings, 3 adds; per sample 3 multiply-adds on HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "device_common.hpp"
#include "host_common.hpp"
#include "mrilab_amd.h"

namespace mrs {

constexpr uint32_t kSarStep = 0, kSarSample = 1;
constexpr int kSarMaxTx = 16;

struct SarDecoded {
    int stride = 2;
    std::vector<uint32_t> rec;  // [kind | (word0)] see sar_decode
    int64_t n_steps = 0, n_samples = 0;
};

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// DoCalSARAtCPU.cpp:191-355 without the per-voxel arithmetic.
static mrs_status sar_decode(const mrs_sar_obj *o, const mrs_vseq *vs, mrs_vvar *vv, mrs_sar_sig *g, SarDecoded &d) {
    const int Tx = o->TxCoilNum;
    d.stride = (2 + Tx + 1) & ~1;
    const int64_t MaxStep = vs->n_ts, MaxutsStep = vs->n_uts, MaxrfStep = vs->n_rf, MaxSARStep = g->n_samples;
    auto sample = [&]() {  // :223-238 / :344-354
        const int64_t at = (int64_t)d.rec.size();
        d.rec.resize(at + d.stride, 0u);
        d.rec[at] = (uint32_t)g->SARi;
        d.rec[at + 1] = kSarSample;
        g->tRealSample[g->SARi] = g->Muts[vv->utsi];
        g->SARi++;
        d.n_samples++;
    };
    int64_t i = 0;
    double flag[6];
    while (i < MaxStep - 1) {
        for (double &f : flag) f = 0;
        if (vs->tsLine[i] != vs->tsLine[i + 1]) {  // :200-236
            for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[i * 6 + f];
            i++;
        } else {
            for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[i * 6 + f];
            while (vs->tsLine[i] == vs->tsLine[i + 1]) {
                for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[(i + 1) * 6 + f];
                i++;
                if (i == MaxStep - 1) break;
            }
            i++;
        }
        if (g->SARi < MaxSARStep) {  // :239-255
            if (g->Muts[vv->utsi] >= g->tSample[g->SARi]) sample();
            if (g->SARi == MaxSARStep) return MRS_OK;
        }
        vv->t = vs->utsLine[vv->utsi];  // :258-260
        {
            const int64_t nxt = std::min<int64_t>(vv->utsi + 1, MaxutsStep - 1);
            vv->dt = vs->utsLine[nxt] - vs->utsLine[vv->utsi];
            vv->utsi = (int32_t)nxt;
        }
        if (flag[0] >= 1) {  // :262-291
            for (int64_t j = 0; j < flag[0]; j++) {
                if (vv->rfi >= MaxrfStep) return fail(MRS_ERR_ARG, "rf flags exceed rfAmpLine");
                vv->rfCoil = vs->rfCoilLine[vv->rfi];
                const int c = (int)vv->rfCoil;
                if (c < 1 || c > Tx) return fail(MRS_ERR_ARG, "rfCoilLine names a coil outside 1..TxCoilNum");
                for (int64_t s = vv->rfi + 1; s < MaxrfStep; s++) {
                    if (vv->rfCoil == vs->rfCoilLine[s]) {
                        const int64_t r = vv->rfi;
                        vv->rfAmp[c - 1] = std::fabs(vs->rfAmpLine[r]) <= std::fabs(vs->rfAmpLine[s])
                                               ? vs->rfAmpLine[r] : vs->rfAmpLine[s];
                        vv->rfPhase[c - 1] = std::fabs(vs->rfPhaseLine[r]) <= std::fabs(vs->rfPhaseLine[s])
                                                 ? vs->rfPhaseLine[r] : vs->rfPhaseLine[s];
                        vv->rfFreq[c - 1] = std::fabs(vs->rfFreqLine[r]) <= std::fabs(vs->rfFreqLine[s])
                                                ? vs->rfFreqLine[r] : vs->rfFreqLine[s];
                        break;
                    }
                }
                vv->rfi++;
            }
        }
        if (flag[0] + flag[1] + flag[2] + flag[3] + flag[4] + flag[5] == 0) {  // :293-298
            for (int c = 0; c < Tx; c++) vv->rfAmp[c] = vv->rfPhase[c] = vv->rfFreq[c] = 0;
            vv->Ext = 0;
        }
        if (vv->dt == 0) continue;  // :301-303
        if (vv->dt < 0) {           // :304-313
            vv->TRCount++;
            if (vv->utsi > 0) g->Muts[vv->utsi] = g->Muts[vv->utsi - 1];
            continue;
        }
        float amp_sum = 0.f;  // SARKernel.h:13-19
        for (int c = 0; c < Tx; c++) amp_sum += std::fabs((float)vv->rfAmp[c]);
        if (amp_sum != 0.f) {
            const int64_t at = (int64_t)d.rec.size();
            d.rec.resize(at + d.stride, 0u);
            d.rec[at] = fbits((float)vv->dt);
            d.rec[at + 1] = kSarStep;
            for (int c = 0; c < Tx; c++) d.rec[at + 2 + c] = fbits((float)vv->rfAmp[c]);
            d.n_steps++;
        }
        if (vv->utsi > 0) g->Muts[vv->utsi] = g->Muts[vv->utsi - 1] + vv->dt;  // :342
        if (g->SARi < MaxSARStep) {                                           // :344-354
            if (g->Muts[vv->utsi] >= g->tSample[g->SARi]) sample();
            if (g->SARi == MaxSARStep) return MRS_OK;
        }
    }
    return MRS_OK;
}

struct SarParams {
    const float *ex, *ey, *ez;  // [Tx][nvox]
    const float *econ;          // [3][nvox]
    float *base;                // [3][nvox] VVar.SAR
    float *sar;                 // [n_samples][nvox] VSig.SAR
    int64_t nvox;
    int64_t n_rec;
    int stride, tx;
};

template <int TX>
__global__ __launch_bounds__(256) void sar_kernel(SarParams p, const uint32_t *__restrict__ rec_g) {
#pragma clang fp contract(off)
    const crec_t rec = (crec_t)rec_g;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= p.nvox) return;  // no barriers or cross-lane work below
    float ex[TX], ey[TX], ez[TX];
#pragma unroll
    for (int c = 0; c < TX; c++) {
        const bool on = c < p.tx;
        ex[c] = on ? p.ex[(int64_t)c * p.nvox + v] : 0.f;
        ey[c] = on ? p.ey[(int64_t)c * p.nvox + v] : 0.f;
        ez[c] = on ? p.ez[(int64_t)c * p.nvox + v] : 0.f;
    }
    float sx = p.base[v], sy = p.base[p.nvox + v], sz = p.base[2 * p.nvox + v];
    const float cx = p.econ[v], cy = p.econ[p.nvox + v], cz = p.econ[2 * p.nvox + v];
    for (int64_t r = 0; r < p.n_rec; r++) {
        const crec_t q = rec + r * p.stride;
        if (q[1] == kSarStep) {
            // SARKernel.h:34-57: E = sum_c E_c * a_c over nonzero a_c, then
            // E^2, times dt, added to the running SAR
            float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
            for (int c = 0; c < TX; c++) {
                if (c >= p.tx) break;
                const float a = __uint_as_float(q[2 + c]);
                if (a != 0.f) {
                    bx = bx + ex[c] * a;
                    by = by + ey[c] * a;
                    bz = bz + ez[c] * a;
                }
            }
            const float dt = __uint_as_float(q[0]);
            bx = bx * bx;
            by = by * by;
            bz = bz * bz;
            bx = bx * dt;
            by = by * dt;
            bz = bz * dt;
            sx = sx + bx;
            sy = sy + by;
            sz = sz + bz;
        } else {
            // DoCalSARAtCPU.cpp:225-233: VSig.SAR(:, s) += SAR_k * ECon_k, k = x, y, z in turn
            float *dst = p.sar + (int64_t)q[0] * p.nvox + v;
            float a = *dst;
            a = a + sx * cx;
            a = a + sy * cy;
            a = a + sz * cz;
            *dst = a;
        }
    }
    p.base[v] = sx;
    p.base[p.nvox + v] = sy;
    p.base[2 * p.nvox + v] = sz;
}

struct SDev {
    void *p = nullptr;
    ~SDev() {
        if (p) (void)hipFree(p);
    }
};

static mrs_status supload(SDev &b, const void *src, size_t bytes, hipStream_t st) {
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 8));
    if (bytes && src) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return MRS_OK;
}

}  // namespace mrs

using namespace mrs;

extern "C" {

mrs_status mrs_DoCalSARAtGPU(const mrs_sar_obj *o, const mrs_vseq *vs, mrs_vvar *vv, mrs_sar_sig *g) {
    if (!o || !vs || !vv || !g) return fail(MRS_ERR_ARG, "NULL struct pointer");
    if (!o->ECon || !o->TxE1x || !o->TxE1y || !o->TxE1z) return fail(MRS_ERR_ARG, "ECon/TxE1 pointer is NULL");
    if (!g->SARBase || !g->SAR || !g->Muts || !g->tSample || !g->tRealSample)
        return fail(MRS_ERR_ARG, "SAR output pointer is NULL");
    if (!vv->rfAmp || !vv->rfPhase || !vv->rfFreq) return fail(MRS_ERR_ARG, "VVar rf arrays are NULL");
    if (!vs->tsLine || !vs->utsLine || !vs->flagsLine || (vs->n_rf > 0 && (!vs->rfAmpLine || !vs->rfCoilLine)))
        return fail(MRS_ERR_ARG, "VSeq line pointer is NULL");
    if (o->dims[0] <= 0 || o->dims[1] <= 0 || o->dims[2] <= 0) return fail(MRS_ERR_ARG, "MassDen dims must be positive");
    if (o->TxCoilNum < 1 || o->TxCoilNum > kSarMaxTx) return fail(MRS_ERR_UNSUPPORTED, "TxCoilNum must be 1..16");
    if (vs->n_uts < 1 || vs->n_ts < 1) return fail(MRS_ERR_ARG, "empty sequence");
    if (vv->utsi < 0 || vv->utsi >= vs->n_uts || vv->rfi < 0 || g->SARi < 0)
        return fail(MRS_ERR_ARG, "VVar.utsi/rfi or SARi out of range");
    const int64_t nvox = (int64_t)o->dims[0] * o->dims[1] * o->dims[2];
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail(MRS_ERR_DEVICE, "no HIP device visible");
    HIP_TRY(hipSetDevice(0));
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct S {
        hipStream_t s;
        ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    // the event loop (host only: it writes VVar, tRealSample and SARi) runs
    // on a thread of its own while the per-voxel arrays go to the device
    SarDecoded d;
    g->n_rf_steps = 0;
    mrs_status drc = MRS_OK;
    std::string dmsg;
    std::thread decoder([&]() {
        drc = sar_decode(o, vs, vv, g, d);
        if (drc) dmsg = mrs_last_error();  // the message is thread-local
    });
    struct J {
        std::thread &t;
        ~J() {
            if (t.joinable()) t.join();
        }
    } joiner{decoder};
    const int Tx = o->TxCoilNum;
    const size_t eb = (size_t)nvox * Tx * sizeof(float), vb = (size_t)nvox * sizeof(float);
    SDev ex, ey, ez, econ, base, sar, rec;
    mrs_status rc;
    if ((rc = supload(ex, o->TxE1x, eb, st)) || (rc = supload(ey, o->TxE1y, eb, st)) ||
        (rc = supload(ez, o->TxE1z, eb, st)) || (rc = supload(econ, o->ECon, 3 * vb, st)) ||
        (rc = supload(base, g->SARBase, 3 * vb, st)) ||
        (rc = supload(sar, g->SAR, (size_t)g->n_samples * vb, st)))
        return rc;
    decoder.join();
    if (drc) return fail(drc, dmsg);
    g->n_rf_steps = d.n_steps;
    if (d.rec.empty()) {  // nothing reaches the voxels
        HIP_TRY(hipStreamSynchronize(st));
        return MRS_OK;
    }
    if ((rc = supload(rec, d.rec.data(), d.rec.size() * 4, st))) return rc;
    SarParams p;
    p.ex = (const float *)ex.p;
    p.ey = (const float *)ey.p;
    p.ez = (const float *)ez.p;
    p.econ = (const float *)econ.p;
    p.base = (float *)base.p;
    p.sar = (float *)sar.p;
    p.nvox = nvox;
    p.n_rec = (int64_t)d.rec.size() / d.stride;
    p.stride = d.stride;
    p.tx = Tx;
    const dim3 grid((unsigned)((nvox + 255) / 256)), block(256);
    const uint32_t *rp = (const uint32_t *)rec.p;
    if (Tx <= 1)
        sar_kernel<1><<<grid, block, 0, st>>>(p, rp);
    else if (Tx <= 2)
        sar_kernel<2><<<grid, block, 0, st>>>(p, rp);
    else if (Tx <= 4)
        sar_kernel<4><<<grid, block, 0, st>>>(p, rp);
    else if (Tx <= 8)
        sar_kernel<8><<<grid, block, 0, st>>>(p, rp);
    else
        sar_kernel<16><<<grid, block, 0, st>>>(p, rp);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g->SARBase, base.p, 3 * vb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(g->SAR, sar.p, (size_t)g->n_samples * vb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return MRS_OK;
}

}  // extern "C"
