// MI355X Bloch scan engine: forward kernel, signal reduction, plans and
// the C ABI of include/mrilab_amd.h.
//
// Design (gfx950, 64-wide wavefronts):
//  * One thread owns one voxel of one type and SPT of its isochromats; the
//    magnetization stays in VGPRs for every step of a launch. Maps are read
//    once per launch from the MRiLab column-major layout, which is already
//    structure-of-arrays, so every load and store is coalesced.
//  * Step records are wave-uniform. They are read through the scalar data
//    cache (address space 4), so record fields land in SGPRs and the record
//    flags drive scalar branches; no LDS or VGPRs are spent on them.
//  * Free precession with the same dt and fields as the previous step reuses
//    the per-spin rotation (cos, sin); relaxation factors are recomputed only
//    when dt changes.
//  * At an ADC step each thread forms sum_spins(Mx + iMy); the signal per
//    coil is that times conj(C_c) summed over the voxels. With 4 or 8 coils
//    each wave stages 8 samples of its 64 voxels' sums in LDS and one MFMA
//    chain (16x16x4 f32, K over voxel and component) projects them onto the
//    coil maps held beside them; with other coil counts each thread forms
//    its coil products and a transposed butterfly inside the wavefront
//    leaves one coil value per lane group. The waves of a block add theirs
//    in LDS, and each block writes one partial per sample. A second kernel
//    sums the block partials of each sample in a fixed order (compensated
//    fp32 per lane, fp64 across lanes), demodulates by rfRef and adds into
//    Sx/Sy. No atomics: results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "decode.hpp"
#include "device_common.hpp"
#include "mrilab_amd.h"
#include "host_common.hpp"
#include "multipool.hpp"
#include "spin_step.hpp"

namespace mrs {

constexpr int kMaxTypes = 16;
constexpr double kPi = 3.14159265359;  // DoScanAtCPU.cpp:63

struct FwdParams {
    float *Mx, *My, *Mz;
    const float *dWRnd, *Rho, *T1, *T2, *dB0, *gx, *gy, *gz;
    const float *txmg, *txpe, *rxx, *rxy;
    float *partial;                 // [adc_local][nblk][NV]
    float *capx, *capy, *capz;      // RunMode 1 capture [slot][ntot]
    const int32_t *cap_slot;        // per record
    int64_t nvox;
    int S, Ty, Tx, Rx, ngroups;
    int rx_default, precision;
    int stride;
    float gyro;
    float cs[kMaxTypes];            // 2*pi*CS per type, rounded as BlochKernel.h:42
};

// GEN = false: single Tx coil and no RunMode-1 capture (the common case);
// GEN = true: any Tx coil count and capture. PREC: relaxation form
// (0 reference rounding, 1 accurate). Up to 4 spins per thread the kernel
// asks for 4 waves per SIMD (at most 128 registers): measured on the 3-D
// bSSFP scan with 8 coils, 3 waves per SIMD ran 14 % slower, and the few
// values the limit moves to scratch are spilled before the step loop and
// reloaded after it.
// MFC: the MFMA coil projection (4 or 8 coils), else the butterfly.
// DEFER: free steps between ADC samples, rf steps and spoilers gather their
// rotation angles and apply them once (bloch_step), else every step rotates.
template <int SPT, int NRX, bool GEN, int PREC, bool MFC, bool DEFER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPT <= 4 ? 4 : 1)))
void bloch_fwd_kernel(FwdParams p, const uint32_t *__restrict__ rec_g, int r0, int r1) {
    constexpr int NV = 2 * NRX;
    constexpr bool MF = MFC && (NRX == 4 || NRX == 8);
    extern __shared__ float red[];  // [kRedBatch][nwaves][NV], then (MF) kMfmaWaveFloats floats per wave
    const crec_t rec = (crec_t)rec_g;
    const int stride = GEN ? p.stride : 8;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int type = blockIdx.y / p.ngroups;
    const int grp = blockIdx.y - type * p.ngroups;
    const int64_t nvox = p.nvox;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = v < nvox;
    const int64_t vv = valid ? v : nvox - 1;
    const int64_t tv = vv + (int64_t)type * nvox;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;

    VoxelConst k;
    k.x = p.gx[vv];
    k.y = p.gy[vv];
    k.z = p.gz[vv];
    k.t1 = p.T1[tv];
    k.t2 = p.T2[tv];
    k.m0s = p.Rho[tv] / (float)p.S;  // BlochKernel.h:365
    const float w0 = p.dB0[vv] * p.gyro;  // BlochKernel.h:38
    const float cst = p.cs[type];

    SpinState<SPT> st;
    int64_t sidx[SPT];
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        sidx[s] = vv + nvox * ((int64_t)(grp * SPT + s) + (int64_t)p.S * type);
        st.mx[s] = p.Mx[sidx[s]];
        st.my[s] = p.My[sidx[s]];
        st.mz[s] = p.Mz[sidx[s]];
        st.wb[s] = (w0 + p.dWRnd[sidx[s]]) + cst;  // BlochKernel.h:38-42
        st.cc[s] = 1.f;
        st.ss[s] = 0.f;
    }
    float crx[NRX], cry[NRX];
#pragma unroll
    for (int c = 0; c < NRX; c++) {
        if (p.rx_default) {
            crx[c] = 1.f;
            cry[c] = 0.f;
        } else if (c < p.Rx) {
            crx[c] = p.rxx[vv + (int64_t)c * nvox];
            cry[c] = p.rxy[vv + (int64_t)c * nvox];
        } else {
            crx[c] = 0.f;
            cry[c] = 0.f;
        }
    }
    if (!valid) {  // their state mirrors the last voxel's; keep it out of the sums
#pragma unroll
        for (int c = 0; c < NRX; c++) crx[c] = cry[c] = 0.f;
    }
    float *const mstage = MF ? red + kRedBatch * nwaves * NV + wave * kMfmaWaveFloats : nullptr;
    float *const mcoil = MF ? mstage + 2 * kMfmaRows * kMfmaStride : nullptr;
    if constexpr (MF) {
        // B operand: MFMA step m of lane (g, j) = 16 g + j reads the map of
        // voxel 16 g + m for column j (Cx of coil j/2 for even j, Cy for odd
        // j), at float ((m/4) * 64 + lane) * 4 + m % 4; this lane is voxel
        // 16 g + m with g = lane / 16, m = lane % 16
        const int g = lane >> 4, m = lane & 15;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int c = j >> 1;
            mcoil[((m >> 2) * 64 + 16 * g + j) * 4 + (m & 3)] = c < NRX ? ((j & 1) ? cry[c] : crx[c]) : 0.f;
        }
    }
    k.wtx_r = k.wtx_i = 0.f;
    if (!GEN || p.Tx == 1) {
        float sp, cp;
        sincos_cr(p.txpe[vv], &sp, &cp);
        const float mg = p.txmg[vv];
        k.wtx_r = mg * cp;
        k.wtx_i = mg * sp;
    }
    const TxMaps txm{p.txmg, p.txpe, vv, nvox, GEN ? p.Tx : 1};

    int slot = 0;
    int adc_batch0 = 0;
    // MF: project the staged samples [row0, row0 + nrows) of the red batch
    // onto the coils and leave this wave's sums in red
    auto mfma_project = [&](int row0, int nrows) __attribute__((always_inline)) {
        if constexpr (MF) {
            // this wave's LDS writes precede its reads (LDS is in order per wave)
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int g = lane >> 4, j = lane & 15;
            // A: row j of the staged samples, voxels 16 g .. 16 g + 15
            const float4 *const arow = (const float4 *)(mstage + j * kMfmaStride + 16 * g);
            const float4 *const brow = (const float4 *)(mcoil + lane * 4);
            mfma_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int qq = 0; qq < 4; qq++) {
                const float4 a = arow[qq], b = brow[qq * 64];
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, acc, 0, 0, 0);
            }
            // lane holds P[4g + i][j], i = 0..3: rows 2s, 2s+1 of samples
            // s = 2g (i = 0, 1) and 2g + 1 (i = 2, 3); the neighbouring
            // lane (j ^ 1) holds the other column of the same coil
            const float sgn = (j & 1) ? -1.f : 1.f;
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const float o = dpp_partner<1>(acc[2 * t + 1]) + sgn * acc[2 * t];
                if (j < NV && 2 * g + t < nrows) red[((row0 + 2 * g + t) * nwaves + wave) * NV + j] = o;
            }
        }
    };
    for (int r = r0; r < r1; ++r) {
        const crec_t q = rec + r * stride;
        const uint32_t flags = q[5];
        if (flags & REC_SPOIL) {
            // a real (scalar) branch: the spoiler is rare, and a select here
            // would cost 2*SPT VALU on every record
            __asm__ volatile("");
            st.spoil();
        }
        // DoScanAtCPU.cpp:445-491, before the step
        auto adc_sample = [&]() __attribute__((always_inline)) {
            if constexpr (DEFER) apply_pending(st);
            // lanes past the last voxel carry zero coil weights
            float sx = st.mx[0], sy = st.my[0];
#pragma unroll
            for (int s = 1; s < SPT; s++) {
                sx += st.mx[s];
                sy += st.my[s];
            }
            if constexpr (MF) {
                const int row = 2 * (slot & (kMfmaRows - 1));
                mstage[row * kMfmaStride + lane] = sx;
                mstage[(row + 1) * kMfmaStride + lane] = sy;
                if ((++slot & (kMfmaRows - 1)) == 0) mfma_project(slot - kMfmaRows, kMfmaRows);
                if (slot == kRedBatch) {
                    __syncthreads();
                    flush_partials(red, kRedBatch, nwaves, NV, p.partial, adc_batch0, nblk, blk, NV);
                    __syncthreads();
                    adc_batch0 += kRedBatch;
                    slot = 0;
                }
                return;
            }
            float val[NV];
#pragma unroll
            for (int c = 0; c < NRX; c++) {
                val[2 * c] = sx * crx[c] + sy * cry[c];
                val[2 * c + 1] = sy * crx[c] - sx * cry[c];
            }
            WaveSum<NV, 32>::run(val, lane);
            constexpr int G = 64 / NV;  // lanes per value after the butterfly
            if ((lane & (G - 1)) == 0) red[(slot * nwaves + wave) * NV + lane / G] = val[0];
            if (++slot == kRedBatch) {
                __syncthreads();
                flush_partials(red, kRedBatch, nwaves, NV, p.partial, adc_batch0, nblk, blk, NV);
                __syncthreads();
                adc_batch0 += kRedBatch;
                slot = 0;
            }
        };
        // an ADC run is this step and then (sample, step) pairs with the
        // cached rotation; every other record is one pass of this loop
        const bool run = (flags & REC_ADCRUN) != 0;
        if (flags & REC_ADC) adc_sample();
        if (!(flags & REC_NOSTEP))
            bloch_step<SPT, GEN, PREC, DEFER>(st, k, q, run ? flags & ((1u << REC_REP_SHIFT) - 1) : flags, txm);
        if (run) {
            int left = (int)(flags >> REC_REP_SHIFT);  // (sample, step) pairs after this step
            if constexpr (MF) {
                // whole batches of staged samples: the sample counter, the
                // projection and the flush test once per batch
                while (left > 0 && (slot & (kMfmaRows - 1)) != 0) {
                    adc_sample();
                    free_step<SPT, PREC>(st, k.m0s);
                    left--;
                }
                for (; left >= kMfmaRows; left -= kMfmaRows) {
#pragma unroll
                    for (int j = 0; j < kMfmaRows; j++) {
                        float sx = st.mx[0], sy = st.my[0];
#pragma unroll
                        for (int s = 1; s < SPT; s++) {
                            sx += st.mx[s];
                            sy += st.my[s];
                        }
                        mstage[2 * j * kMfmaStride + lane] = sx;
                        mstage[(2 * j + 1) * kMfmaStride + lane] = sy;
                        free_step<SPT, PREC>(st, k.m0s);
                    }
                    slot += kMfmaRows;
                    mfma_project(slot - kMfmaRows, kMfmaRows);
                    if (slot == kRedBatch) {
                        __syncthreads();
                        flush_partials(red, kRedBatch, nwaves, NV, p.partial, adc_batch0, nblk, blk, NV);
                        __syncthreads();
                        adc_batch0 += kRedBatch;
                        slot = 0;
                    }
                }
            }
            for (; left > 0; left--) {
                adc_sample();
                free_step<SPT, PREC>(st, k.m0s);
            }
        }
        if constexpr (DEFER && GEN) {
            if (flags & REC_CAPTURE) apply_pending(st);
        }
        if (GEN && (flags & REC_CAPTURE) && valid) {  // DoScanAtCPU.cpp:544-552
            const int64_t base = (int64_t)p.cap_slot[r] * nvox * p.S * p.Ty;
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                p.capx[base + sidx[s]] = st.mx[s];
                p.capy[base + sidx[s]] = st.my[s];
                p.capz[base + sidx[s]] = st.mz[s];
            }
        }
    }
    if (slot > 0) {
        if (MF && (slot & (kMfmaRows - 1)) != 0)
            mfma_project(slot & ~(kMfmaRows - 1), slot & (kMfmaRows - 1));
        __syncthreads();
        flush_partials(red, slot, nwaves, NV, p.partial, adc_batch0, nblk, blk, NV);
    }
    if constexpr (DEFER) apply_pending(st);
    if (valid) {
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            p.Mx[sidx[s]] = st.mx[s];
            p.My[sidx[s]] = st.my[s];
            p.Mz[sidx[s]] = st.mz[s];
        }
    }
}


// Neumaier-compensated fp32 accumulation of block partials per lane, then
// an fp64 butterfly across the wave; one wave per (sample, type). The
// partials of type t are blocks [t*bstep, t*bstep + bcount) at value offset
// t*vstep within each block's vstride values.
template <int NV>
__global__ __launch_bounds__(64) void signal_reduce_kernel(const float *__restrict__ partial, int nblk, int bstep,
                                                           int bcount, int vstride, int vstep, int Rx,
                                                           int rx_default, const double *__restrict__ adc_ref,
                                                           int64_t adc0, int64_t N, double *Sx, double *Sy,
                                                           int c0, int rx_total) {
    __shared__ double tot[NV];
    const int a = blockIdx.x;
    const int t = blockIdx.y;
    const int lane = threadIdx.x;
    float sum[NV], comp[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        sum[j] = 0.f;
        comp[j] = 0.f;
    }
    const int b0 = t * bstep;
    const int voff = t * vstep;
    for (int b = lane; b < bcount; b += 64) {
        const float *src = partial + ((int64_t)a * nblk + b0 + b) * vstride + voff;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const float xv = src[j];
            const float s = sum[j] + xv;
            if (fabsf(sum[j]) >= fabsf(xv))
                comp[j] += (sum[j] - s) + xv;
            else
                comp[j] += (xv - s) + sum[j];
            sum[j] = s;
        }
    }
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double d = (double)sum[j] + (double)comp[j];
        for (int o = 32; o >= 1; o >>= 1) d += __shfl_xor(d, o);
        if (lane == j) tot[j] = d;
    }
    __syncthreads();
    const double ref = adc_ref[adc0 + a];
    const double cr = cos(-ref), sr = sin(-ref);
    const int64_t n = adc0 + a;
    for (int c = lane; c < Rx; c += 64) {
        const int cc = rx_default ? 0 : c;
        double re = tot[2 * cc], im = tot[2 * cc + 1];
        if (ref != 0) {  // DoScanAtCPU.cpp:472-479: multiply by exp(-i*rfRef)
            const double r2 = re * cr - im * sr;
            const double i2 = re * sr + im * cr;
            re = r2;
            im = i2;
        }
        Sx[(int64_t)t * rx_total * N + (int64_t)(c0 + c) * N + n] += re;
        Sy[(int64_t)t * rx_total * N + (int64_t)(c0 + c) * N + n] += im;
    }
}

}  // namespace mrs

// ===========================================================================
// Host side
// ===========================================================================
using namespace mrs;

static thread_local std::string g_err;

namespace mrs {
mrs_status fail(mrs_status st, const std::string &msg) {
    g_err = msg;
    return st;
}
}  // namespace mrs

typedef void (*fwd_fn)(FwdParams, const uint32_t *, int64_t, int64_t);

template <int SPT, int NRX, bool GEN, int PREC, bool MFC, bool DEFER>
static void launch_fwd(dim3 grid, dim3 block, size_t lds, hipStream_t st, const FwdParams &p,
                       const uint32_t *rec, int r0, int r1) {
    bloch_fwd_kernel<SPT, NRX, GEN, PREC, MFC, DEFER><<<grid, block, lds, st>>>(p, rec, r0, r1);
}

typedef void (*launch_fn)(dim3, dim3, size_t, hipStream_t, const FwdParams &, const uint32_t *, int, int);

template <int SPT, bool GEN, int PREC, bool DEFER>
static launch_fn pick_nrx(int nrx, bool mfma) {
    switch (nrx) {
        case 1: return launch_fwd<SPT, 1, GEN, PREC, false, DEFER>;
        case 2: return launch_fwd<SPT, 2, GEN, PREC, false, DEFER>;
        case 4: return mfma ? launch_fwd<SPT, 4, GEN, PREC, true, DEFER> : launch_fwd<SPT, 4, GEN, PREC, false, DEFER>;
        case 8: return mfma ? launch_fwd<SPT, 8, GEN, PREC, true, DEFER> : launch_fwd<SPT, 8, GEN, PREC, false, DEFER>;
        case 16: return launch_fwd<SPT, 16, GEN, PREC, false, DEFER>;
        case 32: return launch_fwd<SPT, 32, GEN, PREC, false, DEFER>;
    }
    return nullptr;
}

template <bool GEN, int PREC, bool DEFER>
static launch_fn pick_spt(int spt, int nrx, bool mfma) {
    switch (spt) {
        case 1: return pick_nrx<1, GEN, PREC, DEFER>(nrx, mfma);
        case 2: return pick_nrx<2, GEN, PREC, DEFER>(nrx, mfma);
        case 4: return pick_nrx<4, GEN, PREC, DEFER>(nrx, mfma);
        case 8: return pick_nrx<8, GEN, PREC, DEFER>(nrx, mfma);
    }
    return nullptr;
}

// Whether the forward kernel for nrx (padded) receive coils sums the coil
// signal with the MFMA projection (coil_sum option 0) or the butterfly (1).
static bool use_mfma(int nrx, int coil_sum) { return coil_sum == 0 && (nrx == 4 || nrx == 8); }

template <bool GEN, int PREC>
static launch_fn pick_defer(int spt, int nrx, bool mfma, bool defer) {
    return defer ? pick_spt<GEN, PREC, true>(spt, nrx, mfma) : pick_spt<GEN, PREC, false>(spt, nrx, mfma);
}

// Forward kernel for spt spins per thread, nrx (padded) receive coils,
// the general variant (multi-Tx, capture) or not, relaxation form prec, the
// coil sum of use_mfma, and deferred or stepwise free steps.
static launch_fn pick_fwd(int spt, int nrx, bool gen, int prec, bool mfma, bool defer) {
    if (prec)
        return gen ? pick_defer<true, 1>(spt, nrx, mfma, defer) : pick_defer<false, 1>(spt, nrx, mfma, defer);
    return gen ? pick_defer<true, 0>(spt, nrx, mfma, defer) : pick_defer<false, 0>(spt, nrx, mfma, defer);
}

void launch_signal_reduce(int nv, dim3 grid, hipStream_t st, const float *partial, int nblk, int bstep,
                          int bcount, int vstride, int vstep, int Rx, int rxd, const double *ref, int64_t adc0,
                          int64_t N, double *Sx, double *Sy, int c0, int rx_total) {
#define MRS_RED(NV_)                                                                                         \
    signal_reduce_kernel<NV_><<<grid, 64, 0, st>>>(partial, nblk, bstep, bcount, vstride, vstep, Rx, rxd, ref, \
                                                   adc0, N, Sx, Sy, c0, rx_total)
    switch (nv) {
        case 2: MRS_RED(2); break;
        case 4: MRS_RED(4); break;
        case 8: MRS_RED(8); break;
        case 16: MRS_RED(16); break;
        case 32: MRS_RED(32); break;
        case 64: MRS_RED(64); break;
    }
#undef MRS_RED
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct mrs_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int R = 0, C = 0, Z = 0, S = 0, Ty = 0, Tx = 0, Rx = 0;
    int64_t nvox = 0, ntot = 0;
    int rx_default = 1, precision = 0, run_mode = 0;
    int spt = 1, nrx = 1, block = 256, ngroups = 1, gx = 1, nblk = 1, nv = 2;
    // receive coils per kernel pass and the number of passes: a scan with
    // more coils than one kernel sums re-runs its program once per group
    // of coils, from the same starting state (Msave)
    int rx_grp = 1, n_rxg = 1;
    int64_t max_adc = 0;
    int64_t signal_num = 0;
    int64_t sig_from = 0;     // samples below it were already added into VSig (Ext callbacks)
    float gyro = 0.f;
    float cs[kMaxTypes] = {0};
    // device memory
    DevBuf M0[3], M[3], Msave[3], dWRnd, Rho, T1, T2, dB0, gxg, gyg, gzg, txmg, txpe, rxx, rxy;
    DevBuf rec, cap_slot, adc_ref, partial, Sx, Sy, cap[3];
    size_t rec_cap_words = 0, adc_cap = 0;
    int64_t n_uts = 0;
    // decoded program
    Decoded dec;
    std::vector<int64_t> launch_r0, launch_r1, launch_adc0, launch_nadc;
    mrs_vvar vvar_final;
    std::vector<double> rf_final[3];
    bool ran = false;
    float last_ms = 0.f;
    int timing = 0;
    int model = 0, scheme = 0;        // 0 single pool; 1 GM, 2 MT, 3 CEST, 4 ME multi-pool
    int coil_sum = 0;                 // mrs_options.coil_sum
    int free_steps = 0;               // mrs_options.free_steps
    bool defer = false;               // the program's free steps are deferred (DEFER kernels)
    int variant = 0;                  // model - 1 for the multi-pool engines
    DevBuf K;
    int flag[kGmMaxTypes] = {0};
    std::vector<hipEvent_t> tev;  // per launch: fwd begin, fwd end, reduce end
    double fwd_ms = 0.0, red_ms = 0.0;
    // batched sequences (mrs_multipool_batch_*): one phantom, nseq record
    // streams, each from the initial state; 0 = one scan
    int nseq = 0;
    DevBuf seq_tab, vox_perm, Mout[3];
    // exact multi-pool scheme: loop records (fold_periods) and the
    // per-thread slots their kernels keep a block's map in
    int64_t n_loops = 0;
    DevBuf loop_cols;
    size_t loop_cols_n = 0;
    std::vector<int32_t> seq_r0, seq_r1, seq_slot;    // records; first partial slot in its group
    std::vector<int64_t> seq_nadc, seq_adc_off;       // samples; offset in the adc_ref table
    std::vector<int64_t> seq_signum, seq_sig_off;     // VSig.SignalNum; offset in Sx/Sy
    std::vector<int> grp_begin, grp_end;              // launch groups of whole sequences
    std::vector<std::vector<double>> seq_k[3];        // k-space per sequence (x, y, z)
    std::vector<mrs_vvar> seq_vvar;
    std::vector<std::vector<double>> seq_rf[3];
    ~mrs_plan() {
        for (auto e : tev) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <class T>
static hipError_t dalloc(DevBuf &b, size_t n) {
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
    }
    if (n == 0) n = 1;
    return hipMalloc(&b.p, n * sizeof(T));
}

static mrs_status upload(DevBuf &b, const void *src, size_t bytes, hipStream_t st) {
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 4));
    if (bytes) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return MRS_OK;
}

// Rate matrix of the multi-pool kernels, [Ty][Ty][nvox] with plane q*Ty + t
// the rate from pool t into pool q, from the K layout of each reference
// engine: GM plane q*Ty + t (BlochKernelGM.cuh:417,423), ME plane t*Ty + q
// (BlochKernelME.cuh:327,333, diagonal kept), MT planes 1 (water -> bound)
// and 2 (bound -> water) (BlochKernelMT.cuh:151-154), CEST planes 0 (water
// -> MT), Ty-1 (MT -> water), t-1 (water -> CEST pool t) and t+Ty-2 (CEST
// pool t -> water) (BlochKernelCEST.cuh:157-160, :379-384).
static void canonical_rates(const float *K, int variant, int Ty, int64_t nvox, std::vector<float> &R) {
    R.assign((size_t)Ty * Ty * nvox, 0.f);
    auto set = [&](int a, int b, int plane) {
        std::copy(K + (size_t)plane * nvox, K + (size_t)(plane + 1) * nvox, R.begin() + (size_t)(b * Ty + a) * nvox);
    };
    switch (variant) {
        case kVarME:
            for (int a = 0; a < Ty; a++)
                for (int b = 0; b < Ty; b++) set(a, b, a * Ty + b);
            break;
        case kVarMT:
            set(0, 1, 1);
            set(1, 0, 2);
            break;
        case kVarCEST:
            set(0, 1, 0);
            set(1, 0, Ty - 1);
            for (int t = 2; t < Ty; t++) {
                set(0, t, t - 1);
                set(t, 0, t + Ty - 2);
            }
            break;
        default:
            for (int a = 0; a < Ty; a++)
                for (int b = 0; b < Ty; b++)
                    if (a != b) set(a, b, b * Ty + a);
    }
}

mrs_status mrs::check_args(const mrs_vobj *vo, const mrs_vmag *vm, const mrs_vcoi *vc, const mrs_vctl *vt,
                             const mrs_vseq *vs, const mrs_vvar *vv, const mrs_vsig *vg) {
    if (!vo || !vm || !vc || !vt || !vs || !vv || !vg) return fail(MRS_ERR_ARG, "NULL struct pointer");
    if (!vo->Mx || !vo->My || !vo->Mz || !vo->Rho || !vo->T1 || !vo->T2)
        return fail(MRS_ERR_ARG, "VObj array pointer is NULL");
    if (!vm->dB0 || !vm->dWRnd || !vm->Gxgrid || !vm->Gygrid || !vm->Gzgrid)
        return fail(MRS_ERR_ARG, "VMag array pointer is NULL");
    if (!vc->TxCoilmg || !vc->TxCoilpe || !vc->RxCoilx || !vc->RxCoily)
        return fail(MRS_ERR_ARG, "VCoi array pointer is NULL");
    if (vo->dims[0] <= 0 || vo->dims[1] <= 0 || vo->dims[2] <= 0 || vo->SpinNum <= 0 || vo->TypeNum <= 0)
        return fail(MRS_ERR_ARG, "VObj dims, SpinNum and TypeNum must be positive");
    if (vo->TypeNum > kMaxTypes) return fail(MRS_ERR_UNSUPPORTED, "TypeNum above 16");
    if (vc->TxCoilNum <= 0 || vc->RxCoilNum <= 0) return fail(MRS_ERR_ARG, "coil counts must be positive");
    if (!vt->CS) return fail(MRS_ERR_ARG, "VCtl.CS is NULL");
    if (!vs->utsLine || !vs->tsLine || !vs->flagsLine || !vs->rfAmpLine || !vs->rfPhaseLine || !vs->rfFreqLine ||
        !vs->rfCoilLine || !vs->GzAmpLine || !vs->GyAmpLine || !vs->GxAmpLine || !vs->ADCLine || !vs->ExtLine)
        return fail(MRS_ERR_ARG, "VSeq line pointer is NULL");
    if (vs->n_ts < 2 || vs->n_uts < 1) return fail(MRS_ERR_ARG, "VSeq.tsLine needs at least two entries");
    if (!vv->rfAmp || !vv->rfPhase || !vv->rfFreq) return fail(MRS_ERR_ARG, "VVar rf arrays are NULL");
    if (vt->RunMode == 1 && (!vg->Mx || !vg->My || !vg->Mz || !vg->Muts))
        return fail(MRS_ERR_ARG, "RunMode 1 needs VSig.Mx/My/Mz/Muts");
    if (vt->MultiTransmit && (vt->MasterTxCoil < 1 || vt->MasterTxCoil > vc->TxCoilNum))
        return fail(MRS_ERR_ARG, "VCtl.MasterTxCoil outside 1..TxCoilNum");
    if (const char *bad = check_lines(*vs, *vv, vc->TxCoilNum)) return fail(MRS_ERR_ARG, bad);
    return MRS_OK;
}

// Splits [0, n_rec) into launches of at most max_adc ADC samples.
static void partition(mrs_plan *P) {
    P->launch_r0.clear();
    P->launch_r1.clear();
    P->launch_adc0.clear();
    P->launch_nadc.clear();
    const int64_t n = P->dec.n_rec();
    const int stride = P->dec.stride;
    int64_t r0 = 0, nadc = 0, adc0 = 0;
    for (int64_t r = 0; r < n; r++) {
        const int64_t na = rec_adc_samples(P->dec.rec[r * stride + 5]);
        if (na && nadc + na > P->max_adc) {
            P->launch_r0.push_back(r0);
            P->launch_r1.push_back(r);
            P->launch_adc0.push_back(adc0);
            P->launch_nadc.push_back(nadc);
            adc0 += nadc;
            r0 = r;
            nadc = 0;
        }
        nadc += na;
    }
    if (n > r0) {
        P->launch_r0.push_back(r0);
        P->launch_r1.push_back(n);
        P->launch_adc0.push_back(adc0);
        P->launch_nadc.push_back(nadc);
    }
}

static mrs_status plan_setup(mrs_plan *P, const mrs_vobj *vo, const mrs_vmag *vm, const mrs_vcoi *vc,
                             const mrs_vctl *vt, const mrs_vseq *vs, const mrs_vsig *vg, const mrs_options *opt) {
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (vt->GPUIndex < 0 || vt->GPUIndex >= ndev)
        return fail(MRS_ERR_DEVICE, "VCtl.GPUIndex " + std::to_string(vt->GPUIndex) + " is not a visible device");
    P->device = vt->GPUIndex;
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&P->ev0));
    HIP_TRY(hipEventCreate(&P->ev1));
    P->R = vo->dims[0];
    P->C = vo->dims[1];
    P->Z = vo->dims[2];
    P->S = vo->SpinNum;
    P->Ty = vo->TypeNum;
    P->Tx = vc->TxCoilNum;
    P->Rx = vc->RxCoilNum;
    P->nvox = (int64_t)P->R * P->C * P->Z;
    P->ntot = P->nvox * P->S * P->Ty;
    P->rx_default = vc->RxCoilDefault != 0;
    P->precision = opt ? opt->precision_mode : 0;
    P->timing = opt ? opt->timing : 0;
    P->model = opt ? opt->model : 0;
    P->scheme = opt ? opt->multipool_scheme : 0;
    P->coil_sum = opt ? opt->coil_sum : 0;
    P->free_steps = opt ? opt->free_steps : 0;
    if (P->model < 0 || P->model > 4) return fail(MRS_ERR_ARG, "unknown engine model");
    if (P->scheme < 0 || P->scheme > 2) return fail(MRS_ERR_ARG, "multipool_scheme must be 0, 1 or 2");
    if (P->free_steps < 0 || P->free_steps > 2) return fail(MRS_ERR_ARG, "free_steps must be 0, 1 or 2");
    if (P->model >= 1) {
        P->variant = P->model - 1;
        if (!vo->K) return fail(MRS_ERR_ARG, "multi-pool scan needs VObj.K");
        if (vo->TypeNum < 2 || vo->TypeNum > kGmMaxTypes)
            return fail(MRS_ERR_UNSUPPORTED, "multi-pool engine supports TypeNum 2.." + std::to_string(kGmMaxTypes));
        if (P->variant == kVarGM && !vo->TypeFlag) return fail(MRS_ERR_ARG, "the GM engine needs VObj.TypeFlag");
        if (P->variant == kVarMT && vo->TypeNum != 2)
            return fail(MRS_ERR_ARG, "the MT engine takes two pools (free water, bound)");
        if (P->variant == kVarCEST && vo->TypeNum < 3)
            return fail(MRS_ERR_ARG, "the CEST engine takes water, a bound pool and CEST pools (TypeNum >= 3)");
        for (int t = 0; t < vo->TypeNum; t++) {
            switch (P->variant) {
                case kVarGM: P->flag[t] = (int)vo->TypeFlag[t]; break;
                case kVarMT:
                case kVarCEST: P->flag[t] = t == 1; break;  // pool 1 is the bound pool
                default: P->flag[t] = 0;                    // ME: every pool free
            }
        }
    }
    P->run_mode = vt->RunMode;
    P->signal_num = vg->SignalNum;
    P->n_uts = vs->n_uts;
    P->gyro = (float)vo->Gyro;
    for (int t = 0; t < P->Ty; t++)
        P->cs[t] = (float)((double)((float)vt->CS[t] * 2.0f) * kPi);  // BlochKernel.h:42
    // spins per thread: largest of 8, 4, 2, 1 dividing SpinNum; when not
    // given, also small enough that the scan has a thread for every wave
    // slot of the device (256 CUs x 4 SIMDs x 4 waves x 64 lanes) if it can
    int spt = opt && opt->spins_per_thread > 0 ? opt->spins_per_thread : 8;
    while (spt > 1 && (P->S % spt != 0 || spt > P->S)) spt >>= 1;
    if (!(opt && opt->spins_per_thread > 0)) {
        const int64_t iso = P->nvox * (int64_t)P->S * (P->model >= 1 ? 1 : P->Ty);
        while (spt > 1 && iso / spt < (int64_t)256 * 4 * 4 * 64) spt >>= 1;
    }
    if (spt != 1 && spt != 2 && spt != 4 && spt != 8) return fail(MRS_ERR_ARG, "spins_per_thread must be 1, 2, 4 or 8");
    if (P->S % spt) return fail(MRS_ERR_ARG, "spins_per_thread must divide SpinNum");
    if (P->model >= 1) spt = 1;
    P->spt = spt;
    P->ngroups = P->S / spt;
    int block = opt && opt->block_size > 0 ? opt->block_size : (P->nvox >= (int64_t)256 * 512 ? 256 : 64);
    if (block % 64 || block > 256) return fail(MRS_ERR_ARG, "block_size must be 64, 128, 192 or 256");
    P->block = block;
    P->gx = (int)((P->nvox + block - 1) / block);
    const int gy = P->model >= 1 ? P->ngroups : P->Ty * P->ngroups;
    if (gy > 65535) return fail(MRS_ERR_UNSUPPORTED, "TypeNum*SpinNum/spins_per_thread too large");
    P->nblk = P->gx * gy;
    // coils per kernel pass: 32 for the single-pool kernel, 8 for the
    // multi-pool kernel (4 with 5 or 6 pools, whose signal values per
    // isochromat would exceed a wavefront)
    const int rxn = P->rx_default ? 1 : P->Rx;
    P->rx_grp = std::min(rxn, P->model >= 1 ? (P->Ty <= 4 ? 8 : 4) : 32);
    P->n_rxg = (rxn + P->rx_grp - 1) / P->rx_grp;
    int nrx = 1;
    while (nrx < P->rx_grp) nrx <<= 1;
    P->nrx = nrx;
    P->nv = P->model >= 1 ? gm_values_per_isochromat(P->Ty, nrx) : 2 * nrx;
    if (P->model >= 1 && !pick_gm(P->Ty, nrx, P->scheme, P->variant))
        return fail(MRS_ERR_UNSUPPORTED, "no multi-pool kernel for this pool and coil count");
    const size_t budget = (size_t)512 << 20;  // partial-buffer budget
    int64_t max_adc = (int64_t)(budget / ((size_t)P->nblk * P->nv * sizeof(float)));
    max_adc = std::max<int64_t>(kRedBatch, std::min<int64_t>(max_adc, 65536));
    if (opt && opt->max_adc_per_launch > 0) max_adc = opt->max_adc_per_launch;
    P->max_adc = max_adc;
    if (!pick_fwd(P->spt, P->nrx, true, 0, false, false)) return fail(MRS_ERR_UNSUPPORTED, "no kernel for this coil count");

    hipStream_t st = P->stream;
    const size_t vb = (size_t)P->nvox * sizeof(float);
    mrs_status rc;
    const float *src[3] = {vo->Mx, vo->My, vo->Mz};
    for (int k = 0; k < 3; k++) {
        if ((rc = upload(P->M0[k], src[k], P->ntot * sizeof(float), st))) return rc;
        HIP_TRY(hipMalloc(&P->M[k].p, P->ntot * sizeof(float)));
    }
    if ((rc = upload(P->dWRnd, vm->dWRnd, P->ntot * sizeof(float), st))) return rc;
    if (P->model >= 1) {
        std::vector<float> rates;
        canonical_rates(vo->K, P->variant, P->Ty, P->nvox, rates);
        if ((rc = upload(P->K, rates.data(), rates.size() * sizeof(float), st))) return rc;
        HIP_TRY(hipStreamSynchronize(st));  // rates is a host temporary
    }
    if ((rc = upload(P->Rho, vo->Rho, vb * P->Ty, st))) return rc;
    if ((rc = upload(P->T1, vo->T1, vb * P->Ty, st))) return rc;
    if ((rc = upload(P->T2, vo->T2, vb * P->Ty, st))) return rc;
    if ((rc = upload(P->dB0, vm->dB0, vb, st))) return rc;
    if ((rc = upload(P->gxg, vm->Gxgrid, vb, st))) return rc;
    if ((rc = upload(P->gyg, vm->Gygrid, vb, st))) return rc;
    if ((rc = upload(P->gzg, vm->Gzgrid, vb, st))) return rc;
    if ((rc = upload(P->txmg, vc->TxCoilmg, vb * P->Tx, st))) return rc;
    if ((rc = upload(P->txpe, vc->TxCoilpe, vb * P->Tx, st))) return rc;
    if (!P->rx_default) {
        if ((rc = upload(P->rxx, vc->RxCoilx, vb * P->Rx, st))) return rc;
        if ((rc = upload(P->rxy, vc->RxCoily, vb * P->Rx, st))) return rc;
    } else {
        HIP_TRY(hipMalloc(&P->rxx.p, 4));
        HIP_TRY(hipMalloc(&P->rxy.p, 4));
    }
    const size_t nsig = (size_t)std::max<int64_t>(1, P->signal_num) * P->Ty * P->Rx;
    HIP_TRY(hipMalloc(&P->Sx.p, nsig * sizeof(double)));
    HIP_TRY(hipMalloc(&P->Sy.p, nsig * sizeof(double)));
    if (P->run_mode == 1) {
        const size_t ncap = (size_t)P->ntot * P->n_uts;
        if (ncap * 3 * sizeof(float) > ((size_t)64 << 30))
            return fail(MRS_ERR_UNSUPPORTED, "RunMode 1 capture above 64 GiB of device memory");
        for (int k = 0; k < 3; k++) HIP_TRY(hipMalloc(&P->cap[k].p, ncap * sizeof(float)));
    }
    return MRS_OK;
}

// Uploads the decoded records of P->dec (replacing earlier ones).
// Exact multi-pool scheme (multipool_scheme 1): periodic blocks of the
// record stream become loop records when the block recurs often enough to
// repay running the block once per basis state of the 3*Ty-dimensional
// state (plus the zero state); scheme 2 steps every record. The slots a
// launch's threads keep a block's map in ((3 Ty + 2) * 3 Ty doubles each)
// are capped at kLoopColsMax bytes; a plan that would need more steps
// every record.
constexpr double kLoopColsMax = 8.0 * (1 << 30);
static void fold_exact(const mrs_plan *P, Decoded &d, int64_t threads, int64_t *loops) {
    if (P->model < 1 || P->scheme != 1) return;
    const double D = 3.0 * P->Ty;
    if ((double)threads * (D + 2) * D * sizeof(double) > kLoopColsMax) return;
    *loops += fold_periods(d, 3 * (int64_t)P->Ty + 2, (int64_t)1 << 20);
}

// Slots for loop records: (3 Ty + 2) * 3 Ty doubles per thread of a launch.
static mrs_status ensure_loop_cols(mrs_plan *P, int64_t threads) {
    if (!P->n_loops) return MRS_OK;
    const int64_t D = 3 * (int64_t)P->Ty;
    const size_t need = (size_t)(threads * (D + 2) * D);
    if (need <= P->loop_cols_n && P->loop_cols.p) return MRS_OK;
    HIP_TRY(hipStreamSynchronize(P->stream));
    HIP_TRY(dalloc<double>(P->loop_cols, need));
    P->loop_cols_n = need;
    return MRS_OK;
}

static mrs_status upload_program(mrs_plan *P) {
    hipStream_t st = P->stream;
    const size_t words = P->dec.rec.size();
    if (words >= ((size_t)1 << 31))
        return fail(MRS_ERR_UNSUPPORTED, "more than 2^31 record words in one program");
    if (words > P->rec_cap_words || !P->rec.p) {
        HIP_TRY(hipStreamSynchronize(st));
        if (P->rec.p) (void)hipFree(P->rec.p);
        P->rec.p = nullptr;
        HIP_TRY(hipMalloc(&P->rec.p, std::max<size_t>(words, 8) * 4));
        P->rec_cap_words = std::max<size_t>(words, 8);
        if (P->run_mode == 1) {
            if (P->cap_slot.p) (void)hipFree(P->cap_slot.p);
            P->cap_slot.p = nullptr;
            HIP_TRY(hipMalloc(&P->cap_slot.p, std::max<size_t>(P->dec.cap_slot.size(), 1) * 4));
        }
    }
    if (words) HIP_TRY(hipMemcpyAsync(P->rec.p, P->dec.rec.data(), words * 4, hipMemcpyHostToDevice, st));
    if (P->run_mode == 1 && !P->dec.cap_slot.empty())
        HIP_TRY(hipMemcpyAsync(P->cap_slot.p, P->dec.cap_slot.data(), P->dec.cap_slot.size() * 4,
                               hipMemcpyHostToDevice, st));
    const size_t nadc = P->dec.adc_ref.size();
    if (nadc > P->adc_cap || !P->adc_ref.p) {
        HIP_TRY(hipStreamSynchronize(st));
        if (P->adc_ref.p) (void)hipFree(P->adc_ref.p);
        P->adc_ref.p = nullptr;
        HIP_TRY(hipMalloc(&P->adc_ref.p, std::max<size_t>(nadc, 1) * sizeof(double)));
        P->adc_cap = std::max<size_t>(nadc, 1);
    }
    if (nadc)
        HIP_TRY(hipMemcpyAsync(P->adc_ref.p, P->dec.adc_ref.data(), nadc * sizeof(double), hipMemcpyHostToDevice,
                               st));
    partition(P);
    int64_t maxn = 0;
    for (auto n : P->launch_nadc) maxn = std::max(maxn, n);
    const size_t need = (size_t)std::max<int64_t>(maxn, 1) * P->nblk * P->nv;
    if (!P->partial.p || need > (size_t)P->max_adc * P->nblk * P->nv) {
        if (P->partial.p) {
            HIP_TRY(hipStreamSynchronize(st));
            (void)hipFree(P->partial.p);
            P->partial.p = nullptr;
        }
        HIP_TRY(hipMalloc(&P->partial.p, (size_t)std::max<int64_t>(P->max_adc, 1) * P->nblk * P->nv * sizeof(float)));
    }
    return MRS_OK;
}

// Kernel parameters of the plan's multi-pool launches.
static GmParams gm_params(mrs_plan *P) {
    GmParams gp;
    std::memset(&gp, 0, sizeof gp);
    gp.Mx = (float *)P->M[0].p;
    gp.My = (float *)P->M[1].p;
    gp.Mz = (float *)P->M[2].p;
    gp.dWRnd = (const float *)P->dWRnd.p;
    gp.Rho = (const float *)P->Rho.p;
    gp.T1 = (const float *)P->T1.p;
    gp.T2 = (const float *)P->T2.p;
    gp.K = (const float *)P->K.p;
    gp.dB0 = (const float *)P->dB0.p;
    gp.gx = (const float *)P->gxg.p;
    gp.gy = (const float *)P->gyg.p;
    gp.gz = (const float *)P->gzg.p;
    gp.txmg = (const float *)P->txmg.p;
    gp.txpe = (const float *)P->txpe.p;
    gp.rxx = (const float *)P->rxx.p;
    gp.rxy = (const float *)P->rxy.p;
    gp.partial = (float *)P->partial.p;
    gp.nvox = P->nvox;
    gp.S = P->S;
    gp.Ty = P->Ty;
    gp.Tx = P->Tx;
    gp.Rx = P->Rx;
    gp.rx_default = P->rx_default;
    gp.stride = P->dec.stride;
    gp.gyro = P->gyro;
    for (int t = 0; t < P->Ty; t++) {
        gp.cs[t] = P->cs[t];
        gp.flag[t] = P->flag[t];
    }
    gp.variant = P->variant;
    gp.self_x = P->variant == kVarME;
    gp.loop_cols = P->n_loops ? (double *)P->loop_cols.p : nullptr;
    return gp;
}

// Receive coils of pass g: the first coil and the count.
static int group_c0(const mrs_plan *P, int g) { return g * P->rx_grp; }
static int group_rx(const mrs_plan *P, int g) {
    return P->rx_default ? P->Rx : std::min(P->rx_grp, P->Rx - g * P->rx_grp);
}

// Multi-pool launches: one thread per (voxel, isochromat), all pools;
// receive coils of pass g.
static mrs_status execute_gm(mrs_plan *P, int g) {
    GmParams gp = gm_params(P);
    gp.rxx += (size_t)group_c0(P, g) * P->nvox;
    gp.rxy += (size_t)group_c0(P, g) * P->nvox;
    gp.Rx = group_rx(P, g);
    gm_launch_fn fn = pick_gm(P->Ty, P->nrx, P->scheme, P->variant);
    const dim3 grid(P->gx, P->ngroups), block(P->block);
    const size_t lds = (size_t)kRedBatch * (P->block / 64) * P->nv * sizeof(float);
    const uint32_t *rec = (const uint32_t *)P->rec.p;
    const size_t nl = P->launch_r0.size();
    const size_t e0 = 3 * nl * (size_t)g;  // timing events of this pass
    if (P->timing) {
        while (P->tev.size() < e0 + 3 * nl) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            P->tev.push_back(e);
        }
    }
    for (size_t l = 0; l < nl; l++) {
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l], P->stream));
        fn(grid, block, lds, P->stream, gp, rec, (int)P->launch_r0[l], (int)P->launch_r1[l]);
        HIP_TRY(hipGetLastError());
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l + 1], P->stream));
        const int64_t na = P->launch_nadc[l];
        if (na > 0) {
            launch_signal_reduce(2 * P->nrx, dim3((unsigned)na, P->Ty), P->stream, (const float *)P->partial.p,
                                 P->nblk, 0, P->nblk, P->nv, 2 * P->nrx, gp.Rx, P->rx_default,
                                 (const double *)P->adc_ref.p, P->launch_adc0[l], P->signal_num, (double *)P->Sx.p,
                                 (double *)P->Sy.p, group_c0(P, g), P->Rx);
            HIP_TRY(hipGetLastError());
        }
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l + 2], P->stream));
    }
    return MRS_OK;
}

// Runs every launch of the uploaded program on the plan stream for the
// receive coils of pass g. adc_base: index of the program's first ADC
// sample in the scan.
static mrs_status execute_group(mrs_plan *P, int64_t adc_base, int g) {
    FwdParams fp;
    std::memset(&fp, 0, sizeof fp);
    fp.Mx = (float *)P->M[0].p;
    fp.My = (float *)P->M[1].p;
    fp.Mz = (float *)P->M[2].p;
    fp.dWRnd = (const float *)P->dWRnd.p;
    fp.Rho = (const float *)P->Rho.p;
    fp.T1 = (const float *)P->T1.p;
    fp.T2 = (const float *)P->T2.p;
    fp.dB0 = (const float *)P->dB0.p;
    fp.gx = (const float *)P->gxg.p;
    fp.gy = (const float *)P->gyg.p;
    fp.gz = (const float *)P->gzg.p;
    fp.txmg = (const float *)P->txmg.p;
    fp.txpe = (const float *)P->txpe.p;
    fp.rxx = (const float *)P->rxx.p + (size_t)group_c0(P, g) * P->nvox;
    fp.rxy = (const float *)P->rxy.p + (size_t)group_c0(P, g) * P->nvox;
    fp.partial = (float *)P->partial.p;
    fp.capx = (float *)P->cap[0].p;
    fp.capy = (float *)P->cap[1].p;
    fp.capz = (float *)P->cap[2].p;
    fp.cap_slot = (const int32_t *)P->cap_slot.p;
    fp.nvox = P->nvox;
    fp.S = P->S;
    fp.Ty = P->Ty;
    fp.Tx = P->Tx;
    fp.Rx = group_rx(P, g);
    fp.ngroups = P->ngroups;
    fp.rx_default = P->rx_default;
    fp.precision = P->precision;
    fp.stride = P->dec.stride;
    fp.gyro = P->gyro;
    for (int t = 0; t < P->Ty; t++) fp.cs[t] = P->cs[t];
    if (P->model >= 1) return execute_gm(P, g);
    const bool mfma = use_mfma(P->nrx, P->coil_sum);
    launch_fn fn = pick_fwd(P->spt, P->nrx, P->Tx != 1 || P->run_mode == 1, P->precision != 0, mfma, P->defer);
    const dim3 grid(P->gx, P->Ty * P->ngroups), block(P->block);
    const size_t lds = (size_t)kRedBatch * (P->block / 64) * P->nv * sizeof(float) +
                       (mfma ? (size_t)(P->block / 64) * kMfmaWaveFloats * sizeof(float) : 0);
    const uint32_t *rec = (const uint32_t *)P->rec.p;
    const size_t nl = P->launch_r0.size();
    const size_t e0 = 3 * nl * (size_t)g;  // timing events of this pass
    if (P->timing) {
        while (P->tev.size() < e0 + 3 * nl) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            P->tev.push_back(e);
        }
    }
    for (size_t l = 0; l < nl; l++) {
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l], P->stream));
        fn(grid, block, lds, P->stream, fp, rec, (int)P->launch_r0[l], (int)P->launch_r1[l]);
        HIP_TRY(hipGetLastError());
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l + 1], P->stream));
        const int64_t na = P->launch_nadc[l];
        if (na > 0) {
            launch_signal_reduce(P->nv, dim3((unsigned)na, P->Ty), P->stream, (const float *)P->partial.p, P->nblk,
                                 P->gx * P->ngroups, P->gx * P->ngroups, P->nv, 0, fp.Rx, P->rx_default,
                                 (const double *)P->adc_ref.p, P->launch_adc0[l], P->signal_num, (double *)P->Sx.p,
                                 (double *)P->Sy.p, group_c0(P, g), P->Rx);
            HIP_TRY(hipGetLastError());
        }
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[e0 + 3 * l + 2], P->stream));
    }
    (void)adc_base;
    return MRS_OK;
}

// Runs the uploaded program: once, or once per group of receive coils, each
// pass from the state the program starts from.
static mrs_status execute(mrs_plan *P, int64_t adc_base) {
    if (P->n_rxg == 1) return execute_group(P, adc_base, 0);
    const size_t bytes = (size_t)P->ntot * sizeof(float);
    for (int k = 0; k < 3; k++) {
        if (!P->Msave[k].p) HIP_TRY(hipMalloc(&P->Msave[k].p, bytes));
        HIP_TRY(hipMemcpyAsync(P->Msave[k].p, P->M[k].p, bytes, hipMemcpyDeviceToDevice, P->stream));
    }
    mrs_status rc;
    for (int g = 0; g < P->n_rxg; g++) {
        if (g > 0)
            for (int k = 0; k < 3; k++)
                HIP_TRY(hipMemcpyAsync(P->M[k].p, P->Msave[k].p, bytes, hipMemcpyDeviceToDevice, P->stream));
        if ((rc = execute_group(P, adc_base, g))) return rc;
    }
    return MRS_OK;
}

static mrs_status reset_state(mrs_plan *P) {
    hipStream_t st = P->stream;
    for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(P->M[k].p, P->M0[k].p, P->ntot * sizeof(float), hipMemcpyDeviceToDevice, st));
    const size_t nsig = (size_t)std::max<int64_t>(1, P->signal_num) * P->Ty * P->Rx;
    HIP_TRY(hipMemsetAsync(P->Sx.p, 0, nsig * sizeof(double), st));
    HIP_TRY(hipMemsetAsync(P->Sy.p, 0, nsig * sizeof(double), st));
    return MRS_OK;
}

void mrs::copy_vvar_out(const mrs_vvar &src, const std::vector<double> *rf, int tx, mrs_vvar *dst) {
    double *a = dst->rfAmp, *p = dst->rfPhase, *f = dst->rfFreq;
    *dst = src;
    dst->rfAmp = a;
    dst->rfPhase = p;
    dst->rfFreq = f;
    for (int c = 0; c < tx; c++) {
        a[c] = rf[0][c];
        p[c] = rf[1][c];
        f[c] = rf[2][c];
    }
}

// One run of a batched multi-pool plan: per launch group one launch over
// (voxel tile, sequence) blocks, then the signal reduction of each sequence.
static mrs_status run_batch(mrs_plan *P) {
    hipStream_t st = P->stream;
    const size_t nsig = (size_t)std::max<int64_t>(1, P->seq_sig_off.back());
    HIP_TRY(hipMemsetAsync(P->Sx.p, 0, nsig * sizeof(double), st));
    HIP_TRY(hipMemsetAsync(P->Sy.p, 0, nsig * sizeof(double), st));
    HIP_TRY(hipEventRecord(P->ev0, st));
    GmParams gp = gm_params(P);
    for (int k = 0; k < 3; k++) {
        (&gp.Mx)[k] = (float *)P->M0[k].p;  // every sequence starts from the initial state
        gp.Mout[k] = (float *)P->Mout[k].p;
    }
    gp.seq_tab = (const int32_t *)P->seq_tab.p;
    gp.vox_perm = (const int32_t *)P->vox_perm.p;
    gp.tiles = P->gx;
    gm_launch_fn fn = pick_gm(P->Ty, P->nrx, P->scheme, P->variant);
    const size_t lds = (size_t)kRedBatch * (P->block / 64) * P->nv * sizeof(float);
    const size_t ng = P->grp_begin.size();
    if (P->timing) {
        while (P->tev.size() < 3 * ng) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            P->tev.push_back(e);
        }
    }
    const int nblk = P->gx * P->ngroups;
    for (size_t g = 0; g < ng; g++) {
        gp.seq_base = P->grp_begin[g];
        gp.nseq_l = P->grp_end[g] - P->grp_begin[g];
        const dim3 grid((unsigned)(P->gx * gp.nseq_l), P->ngroups), block(P->block);
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[3 * g], st));
        fn(grid, block, lds, st, gp, (const uint32_t *)P->rec.p, 0, 0);
        HIP_TRY(hipGetLastError());
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[3 * g + 1], st));
        for (int q = P->grp_begin[g]; q < P->grp_end[g]; q++) {
            const int64_t na = P->seq_nadc[q];
            if (na <= 0) continue;
            launch_signal_reduce(2 * P->nrx, dim3((unsigned)na, P->Ty), st,
                                 (const float *)P->partial.p + (int64_t)P->seq_slot[q] * nblk * P->nv, nblk, 0, nblk,
                                 P->nv, 2 * P->nrx, P->Rx, P->rx_default,
                                 (const double *)P->adc_ref.p + P->seq_adc_off[q], 0, P->seq_signum[q],
                                 (double *)P->Sx.p + P->seq_sig_off[q], (double *)P->Sy.p + P->seq_sig_off[q], 0,
                                 P->Rx);
            HIP_TRY(hipGetLastError());
        }
        if (P->timing) HIP_TRY(hipEventRecord(P->tev[3 * g + 2], st));
    }
    HIP_TRY(hipEventRecord(P->ev1, st));
    P->ran = true;
    return MRS_OK;
}

// Adds the device signal of samples [from, to) of every type and coil into
// VSig.Sx/Sy (index t*Rx*N + c*N + n).
static mrs_status add_signal(mrs_plan *P, mrs_vsig *vsig, int64_t from, int64_t to) {
    if (to <= from) return MRS_OK;
    const int64_t N = P->signal_num, n = to - from;
    std::vector<double> hx((size_t)n), hy((size_t)n);
    for (int t = 0; t < P->Ty; t++)
        for (int c = 0; c < P->Rx; c++) {
            const int64_t off = (int64_t)t * P->Rx * N + (int64_t)c * N + from;
            HIP_TRY(hipMemcpy(hx.data(), (const double *)P->Sx.p + off, n * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hy.data(), (const double *)P->Sy.p + off, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; i++) {
                vsig->Sx[off + i] += hx[(size_t)i];
                vsig->Sy[off + i] += hy[(size_t)i];
            }
        }
    return MRS_OK;
}

// Uploads the current magnetization and every map the kernels read from the
// host structs into the plan's device buffers, after an Ext callback that
// had the state fetched (DoScanAtGPU.cu:633-659: VObj, VMag and VCoi
// arrays, and VObj.K for the multi-pool engines, DoGMScanAtGPU.cu:714).
static mrs_status refresh_maps(mrs_plan *P, const mrs_vobj *vo, const mrs_vmag *vm, const mrs_vcoi *vc) {
    const size_t vb = (size_t)P->nvox * sizeof(float), sb = (size_t)P->ntot * sizeof(float);
    struct {
        DevBuf *d;
        const void *h;
        size_t n;
        const char *name;
    } L[] = {{&P->M[0], vo->Mx, sb, "VObj.Mx"},          {&P->M[1], vo->My, sb, "VObj.My"},
             {&P->M[2], vo->Mz, sb, "VObj.Mz"},          {&P->dWRnd, vm->dWRnd, sb, "VMag.dWRnd"},
             {&P->Rho, vo->Rho, vb * P->Ty, "VObj.Rho"}, {&P->T1, vo->T1, vb * P->Ty, "VObj.T1"},
             {&P->T2, vo->T2, vb * P->Ty, "VObj.T2"},    {&P->dB0, vm->dB0, vb, "VMag.dB0"},
             {&P->gxg, vm->Gxgrid, vb, "VMag.Gxgrid"},   {&P->gyg, vm->Gygrid, vb, "VMag.Gygrid"},
             {&P->gzg, vm->Gzgrid, vb, "VMag.Gzgrid"},   {&P->txmg, vc->TxCoilmg, vb * P->Tx, "VCoi.TxCoilmg"},
             {&P->txpe, vc->TxCoilpe, vb * P->Tx, "VCoi.TxCoilpe"}};
    for (auto &l : L) {
        if (!l.h) return fail(MRS_ERR_CALLBACK, std::string("the Ext callback left ") + l.name + " NULL");
        HIP_TRY(hipMemcpy(l.d->p, l.h, l.n, hipMemcpyHostToDevice));
    }
    if (!P->rx_default) {
        if (!vc->RxCoilx || !vc->RxCoily) return fail(MRS_ERR_CALLBACK, "the Ext callback left a VCoi.RxCoil map NULL");
        HIP_TRY(hipMemcpy(P->rxx.p, vc->RxCoilx, vb * P->Rx, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->rxy.p, vc->RxCoily, vb * P->Rx, hipMemcpyHostToDevice));
    }
    if (P->model >= 1) {
        if (!vo->K) return fail(MRS_ERR_CALLBACK, "the Ext callback left VObj.K NULL");
        std::vector<float> rates;
        canonical_rates(vo->K, P->variant, P->Ty, P->nvox, rates);
        HIP_TRY(hipMemcpy(P->K.p, rates.data(), rates.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return MRS_OK;
}

extern "C" {

const char *mrs_last_error(void) { return g_err.c_str(); }
// free_steps 0 defers free precession when the decoded program has enough
// deferrable steps per pending rotation (deferral_pays), 2 always, 1 never.
static void choose_deferral(mrs_plan *P) {
    int64_t steps = 0, rotations = 0;
    deferral_counts(P->dec, &steps, &rotations);
    P->defer = P->free_steps == 2 || (P->free_steps == 0 && deferral_pays(steps, rotations));
}

int32_t mrs_abi_version(void) { return MRS_ABI_VERSION; }
int32_t mrs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -1;
    return n;
}

mrs_status mrs_plan_create(const mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, const mrs_vvar *vvar, const mrs_vsig *vsig,
                           const mrs_options *opt, mrs_plan **out) {
    if (!out) return fail(MRS_ERR_ARG, "out is NULL");
    *out = nullptr;
    mrs_status rc = check_args(vobj, vmag, vcoi, vctl, vseq, vvar, vsig);
    if (rc) return rc;
    mrs_plan *P = new (std::nothrow) mrs_plan();
    if (!P) return fail(MRS_ERR_NOMEM, "host allocation failed");
    if ((rc = plan_setup(P, vobj, vmag, vcoi, vctl, vseq, vsig, opt))) {
        delete P;
        return rc;
    }
    P->dec.tx = P->Tx;
    P->dec.coil_words = P->model >= 1 ? 4 : 2;
    P->dec.stride = rec_stride(P->Tx, P->dec.coil_words);
    Decoder dec(vobj, vcoi, vctl, vseq, vvar, false);
    dec.run(P->dec);
    dec.flush_tail(P->dec);
    compress_runs(P->dec, P->model == 0 ? P->max_adc : 0);
    choose_deferral(P);
    fold_exact(P, P->dec, (int64_t)P->gx * P->ngroups * P->block, &P->n_loops);
    if ((int64_t)P->dec.adc_ref.size() > P->signal_num) {
        delete P;
        return fail(MRS_ERR_ARG, "VSig.SignalNum is smaller than the number of ADC samples");
    }
    P->vvar_final = dec.vvar();
    P->rf_final[0].assign(dec.vvar().rfAmp, dec.vvar().rfAmp + P->Tx);
    P->rf_final[1].assign(dec.vvar().rfPhase, dec.vvar().rfPhase + P->Tx);
    P->rf_final[2].assign(dec.vvar().rfFreq, dec.vvar().rfFreq + P->Tx);
    if ((rc = upload_program(P)) || (rc = ensure_loop_cols(P, (int64_t)P->gx * P->ngroups * P->block))) {
        delete P;
        return rc;
    }
    if (hipStreamSynchronize(P->stream) != hipSuccess) {
        delete P;
        return fail(MRS_ERR_DEVICE, "upload failed");
    }
    *out = P;
    return MRS_OK;
}

mrs_status mrs_plan_run(mrs_plan *P) {
    if (!P) return fail(MRS_ERR_ARG, "plan is NULL");
    HIP_TRY(hipSetDevice(P->device));
    if (P->nseq) return run_batch(P);
    mrs_status rc = reset_state(P);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(P->ev0, P->stream));
    if ((rc = execute(P, 0))) return rc;
    HIP_TRY(hipEventRecord(P->ev1, P->stream));
    P->ran = true;
    return MRS_OK;
}

mrs_status mrs_plan_sync(mrs_plan *P) {
    if (!P) return fail(MRS_ERR_ARG, "plan is NULL");
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipStreamSynchronize(P->stream));
    if (P->ran) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, P->ev0, P->ev1) == hipSuccess) P->last_ms = ms;
        if (P->timing) {
            double f = 0.0, r = 0.0;
            const size_t nl = P->nseq ? P->grp_begin.size() : P->launch_r0.size() * (size_t)P->n_rxg;
            for (size_t l = 0; l < nl && 3 * l + 2 < P->tev.size(); l++) {
                float a = 0.f, b = 0.f;
                if (hipEventElapsedTime(&a, P->tev[3 * l], P->tev[3 * l + 1]) == hipSuccess) f += a;
                if (hipEventElapsedTime(&b, P->tev[3 * l + 1], P->tev[3 * l + 2]) == hipSuccess) r += b;
            }
            P->fwd_ms = f;
            P->red_ms = r;
        }
        P->ran = false;
    }
    return MRS_OK;
}


mrs_status mrs_plan_fetch(mrs_plan *P, mrs_vobj *vobj, mrs_vvar *vvar, mrs_vsig *vsig) {
    if (!P) return fail(MRS_ERR_ARG, "plan is NULL");
    if (P->nseq) return fail(MRS_ERR_ARG, "a batched multi-pool plan is fetched with mrs_multipool_batch_fetch");
    mrs_status rc = mrs_plan_sync(P);
    if (rc) return rc;
    if (vobj) {
        float *dst[3] = {vobj->Mx, vobj->My, vobj->Mz};
        for (int k = 0; k < 3; k++)
            HIP_TRY(hipMemcpy(dst[k], P->M[k].p, P->ntot * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (vsig) {
        mrs_status rc2 = add_signal(P, vsig, P->sig_from, P->signal_num);
        if (rc2) return rc2;
        // the decoded k-space holds the samples from sig_from on
        const size_t k0 = (size_t)P->sig_from;
        for (size_t n = 0; n < P->dec.kx.size(); n++) {
            vsig->Kz[k0 + n] += P->dec.kz[n];
            vsig->Ky[k0 + n] += P->dec.ky[n];
            vsig->Kx[k0 + n] += P->dec.kx[n];
        }
        if (P->run_mode == 1) {
            const size_t ncap = (size_t)P->ntot * P->n_uts;
            float *dst[3] = {vsig->Mx, vsig->My, vsig->Mz};
            // slots never written keep the caller's values: copy only captured slots
            std::vector<char> hit(P->n_uts, 0);
            for (auto s : P->dec.cap_slot)
                if (s >= 0) hit[s] = 1;
            for (int k = 0; k < 3; k++) {
                std::vector<float> tmp(ncap);
                HIP_TRY(hipMemcpy(tmp.data(), P->cap[k].p, ncap * sizeof(float), hipMemcpyDeviceToHost));
                for (int64_t u = 0; u < P->n_uts; u++)
                    if (hit[u]) std::memcpy(dst[k] + u * P->ntot, tmp.data() + u * P->ntot, P->ntot * sizeof(float));
            }
            for (int64_t u = 0; u < P->n_uts; u++)
                if (hit[u]) vsig->Muts[u] = P->dec.muts[u];
        }
    }
    if (vvar) copy_vvar_out(P->vvar_final, P->rf_final, P->Tx, vvar);
    return MRS_OK;
}

}  // extern "C"

mrs_status mrs::plan_signal(mrs_plan *P, PlanSignal *out) {
    mrs_status rc = mrs_plan_sync(P);
    if (rc) return rc;
    out->sx = (const double *)P->Sx.p;
    out->sy = (const double *)P->Sy.p;
    out->n = P->signal_num;
    out->types = P->Ty;
    out->rx = P->Rx;
    out->device = P->device;
    out->stream = P->stream;
    return MRS_OK;
}

extern "C" {

mrs_status mrs_plan_get_info(const mrs_plan *P, mrs_plan_info *info) {
    if (!P || !info) return fail(MRS_ERR_ARG, "NULL argument");
    info->n_steps = P->dec.n_steps;
    info->n_adc = (int64_t)P->dec.adc_ref.size();
    info->n_iso = P->ntot;
    info->n_launches = (int64_t)(P->nseq ? P->grp_begin.size() : P->launch_r0.size());
    info->n_rf_steps = P->dec.n_rf;
    info->n_cached_steps = P->dec.n_same;
    info->block_size = P->block;
    info->spins_per_thread = P->spt;
    info->grid_blocks = P->nblk;
    info->last_run_ms = P->last_ms;
    info->fwd_kernel_ms = P->fwd_ms;
    info->reduce_kernel_ms = P->red_ms;
    info->deferred_free_steps = P->defer ? 1 : 0;
    return MRS_OK;
}

void mrs_plan_destroy(mrs_plan *P) {
    if (!P) return;
    (void)hipSetDevice(P->device);
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    delete P;
}

mrs_status mrs_multipool_batch_create(int32_t n_seq, const mrs_vobj *vobj, const mrs_vmag *vmag,
                                      const mrs_vcoi *vcoi, const mrs_vctl *vctl, const mrs_vseq *const *vseq,
                                      const mrs_vvar *const *vvar, const mrs_vsig *const *vsig,
                                      const mrs_options *opt, mrs_plan **out) {
    if (!out) return fail(MRS_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_seq < 1 || !vseq || !vvar || !vsig) return fail(MRS_ERR_ARG, "n_seq must be positive, with VSeq/VVar/VSig arrays");
    if (!opt || opt->model < 1 || opt->model > 4)
        return fail(MRS_ERR_ARG, "a batched multi-pool plan needs options.model 1..4 (GM, MT, CEST, ME)");
    mrs_status rc;
    for (int q = 0; q < n_seq; q++) {
        if (!vseq[q] || !vvar[q] || !vsig[q]) return fail(MRS_ERR_ARG, "NULL VSeq/VVar/VSig of a sequence");
        if ((rc = check_args(vobj, vmag, vcoi, vctl, vseq[q], vvar[q], vsig[q]))) return rc;
    }
    if (vctl->RunMode != 0) return fail(MRS_ERR_UNSUPPORTED, "batched sequences run with RunMode 0");
    mrs_plan *P = new (std::nothrow) mrs_plan();
    if (!P) return fail(MRS_ERR_NOMEM, "host allocation failed");
    struct Guard {
        mrs_plan *p;
        ~Guard() {
            if (p) mrs_plan_destroy(p);
        }
    } guard{P};
    if ((rc = plan_setup(P, vobj, vmag, vcoi, vctl, vseq[0], vsig[0], opt))) return rc;
    if (P->n_rxg > 1)
        return fail(MRS_ERR_UNSUPPORTED, "batched multi-pool sequences take up to " + std::to_string(P->rx_grp) +
                                             " receive coils");
    P->nseq = n_seq;
    P->dec = Decoded();
    P->dec.tx = P->Tx;
    P->dec.coil_words = 4;
    P->dec.stride = rec_stride(P->Tx, 4);
    int64_t sig_off = 0;
    for (int q = 0; q < n_seq; q++) {
        Decoded d;
        d.tx = P->Tx;
        d.coil_words = 4;
        d.stride = P->dec.stride;
        Decoder dec(vobj, vcoi, vctl, vseq[q], vvar[q], false);
        dec.run(d);
        dec.flush_tail(d);
        compress_runs(d, 0);
        fold_exact(P, d, (int64_t)P->gx * n_seq * P->ngroups * P->block, &P->n_loops);
        if ((int64_t)d.adc_ref.size() > vsig[q]->SignalNum)
            return fail(MRS_ERR_ARG, "VSig.SignalNum of sequence " + std::to_string(q) +
                                         " is smaller than its number of ADC samples");
        P->seq_r0.push_back((int32_t)P->dec.n_rec());
        P->dec.rec.insert(P->dec.rec.end(), d.rec.begin(), d.rec.end());
        if (P->dec.rec.size() >= ((size_t)1 << 31)) return fail(MRS_ERR_UNSUPPORTED, "more than 2^31 record words");
        P->seq_r1.push_back((int32_t)P->dec.n_rec());
        P->seq_adc_off.push_back((int64_t)P->dec.adc_ref.size());
        P->seq_nadc.push_back((int64_t)d.adc_ref.size());
        P->dec.adc_ref.insert(P->dec.adc_ref.end(), d.adc_ref.begin(), d.adc_ref.end());
        P->dec.n_steps += d.n_steps;
        P->dec.n_rf += d.n_rf;
        P->dec.n_same += d.n_same;
        P->seq_k[0].push_back(std::move(d.kx));
        P->seq_k[1].push_back(std::move(d.ky));
        P->seq_k[2].push_back(std::move(d.kz));
        P->seq_vvar.push_back(dec.vvar());
        P->seq_rf[0].emplace_back(dec.vvar().rfAmp, dec.vvar().rfAmp + P->Tx);
        P->seq_rf[1].emplace_back(dec.vvar().rfPhase, dec.vvar().rfPhase + P->Tx);
        P->seq_rf[2].emplace_back(dec.vvar().rfFreq, dec.vvar().rfFreq + P->Tx);
        P->seq_signum.push_back(vsig[q]->SignalNum);
        P->seq_sig_off.push_back(sig_off);
        sig_off += (int64_t)vsig[q]->SignalNum * P->Ty * P->Rx;
    }
    P->seq_sig_off.push_back(sig_off);
    // launch groups: whole sequences whose block partials fit the budget
    const int nblk = P->gx * P->ngroups;
    const int64_t budget = opt->max_adc_per_launch > 0
                               ? (int64_t)opt->max_adc_per_launch
                               : ((int64_t)512 << 20) / ((int64_t)nblk * P->nv * (int64_t)sizeof(float));
    int64_t used = 0, most = 1;
    P->seq_slot.assign(n_seq, 0);
    P->grp_begin.push_back(0);
    for (int q = 0; q < n_seq; q++) {
        const int64_t na = P->seq_nadc[q];
        if (na > budget)
            return fail(MRS_ERR_UNSUPPORTED, "sequence " + std::to_string(q) +
                                                 " has more ADC samples than one batched launch holds; scan it alone");
        if (used + na > budget) {
            P->grp_end.push_back(q);
            P->grp_begin.push_back(q);
            used = 0;
        }
        P->seq_slot[q] = (int32_t)used;
        used += na;
        most = std::max(most, used);
    }
    P->grp_end.push_back(n_seq);
    {
        int most_seq = 1;
        for (size_t g = 0; g < P->grp_begin.size(); g++) most_seq = std::max(most_seq, P->grp_end[g] - P->grp_begin[g]);
        if ((rc = ensure_loop_cols(P, (int64_t)P->gx * most_seq * P->ngroups * P->block))) return rc;
    }
    hipStream_t st = P->stream;
    for (int k = 0; k < 2; k++) {
        DevBuf &b = k ? P->Sy : P->Sx;
        (void)hipFree(b.p);
        b.p = nullptr;
        HIP_TRY(hipMalloc(&b.p, (size_t)std::max<int64_t>(1, sig_off) * sizeof(double)));
    }
    const size_t words = P->dec.rec.size();
    HIP_TRY(hipMalloc(&P->rec.p, std::max<size_t>(words, 8) * 4));
    if (words) HIP_TRY(hipMemcpyAsync(P->rec.p, P->dec.rec.data(), words * 4, hipMemcpyHostToDevice, st));
    const size_t nadc = P->dec.adc_ref.size();
    HIP_TRY(hipMalloc(&P->adc_ref.p, std::max<size_t>(nadc, 1) * sizeof(double)));
    if (nadc)
        HIP_TRY(hipMemcpyAsync(P->adc_ref.p, P->dec.adc_ref.data(), nadc * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMalloc(&P->partial.p, (size_t)most * nblk * P->nv * sizeof(float)));
    std::vector<int32_t> tab((size_t)4 * n_seq, 0);
    for (int q = 0; q < n_seq; q++) {
        tab[4 * q] = P->seq_r0[q];
        tab[4 * q + 1] = P->seq_r1[q];
        tab[4 * q + 2] = P->seq_slot[q];
    }
    HIP_TRY(hipMalloc(&P->seq_tab.p, tab.size() * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(P->seq_tab.p, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    for (int k = 0; k < 3; k++) HIP_TRY(hipMalloc(&P->Mout[k].p, (size_t)n_seq * P->ntot * sizeof(float)));
    if (P->scheme != 0) {
        // the exact scheme's cost grows with a voxel's fastest rate (its
        // Taylor substeps): order the voxels costliest first, so the waves
        // that hold them share tiles and are dispatched first
        std::vector<float> rates;
        canonical_rates(vobj->K, P->variant, P->Ty, P->nvox, rates);
        std::vector<float> cost((size_t)P->nvox, 0.f);
        for (int a = 0; a < P->Ty * P->Ty; a++)
            for (int64_t v = 0; v < P->nvox; v++) cost[v] += std::fabs(rates[(size_t)a * P->nvox + v]);
        std::vector<int32_t> perm((size_t)P->nvox);
        for (int64_t v = 0; v < P->nvox; v++) perm[v] = (int32_t)v;
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
        HIP_TRY(hipMalloc(&P->vox_perm.p, perm.size() * sizeof(int32_t)));
        HIP_TRY(hipMemcpy(P->vox_perm.p, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipStreamSynchronize(st));
    *out = P;
    guard.p = nullptr;
    return MRS_OK;
}

mrs_status mrs_multipool_batch_fetch(mrs_plan *P, mrs_vobj *const *vobj, mrs_vvar *const *vvar,
                                     mrs_vsig *const *vsig) {
    if (!P) return fail(MRS_ERR_ARG, "plan is NULL");
    if (!P->nseq) return fail(MRS_ERR_ARG, "not a batched multi-pool plan (use mrs_plan_fetch)");
    mrs_status rc = mrs_plan_sync(P);
    if (rc) return rc;
    for (int q = 0; q < P->nseq; q++) {
        if (vobj && vobj[q]) {
            float *dst[3] = {vobj[q]->Mx, vobj[q]->My, vobj[q]->Mz};
            for (int k = 0; k < 3; k++) {
                if (!dst[k]) return fail(MRS_ERR_ARG, "NULL VObj.Mx/My/Mz of a sequence");
                HIP_TRY(hipMemcpy(dst[k], (const float *)P->Mout[k].p + (size_t)q * P->ntot,
                                  P->ntot * sizeof(float), hipMemcpyDeviceToHost));
            }
        }
        if (vsig && vsig[q]) {
            const int64_t n = P->seq_sig_off[q + 1] - P->seq_sig_off[q];
            std::vector<double> hx((size_t)n), hy((size_t)n);
            if (n) {
                HIP_TRY(hipMemcpy(hx.data(), (const double *)P->Sx.p + P->seq_sig_off[q], n * sizeof(double),
                                  hipMemcpyDeviceToHost));
                HIP_TRY(hipMemcpy(hy.data(), (const double *)P->Sy.p + P->seq_sig_off[q], n * sizeof(double),
                                  hipMemcpyDeviceToHost));
            }
            for (int64_t i = 0; i < n; i++) {
                vsig[q]->Sx[i] += hx[(size_t)i];
                vsig[q]->Sy[i] += hy[(size_t)i];
            }
            for (size_t i = 0; i < P->seq_k[0][q].size(); i++) {
                vsig[q]->Kx[i] += P->seq_k[0][q][i];
                vsig[q]->Ky[i] += P->seq_k[1][q][i];
                vsig[q]->Kz[i] += P->seq_k[2][q][i];
            }
        }
        if (vvar && vvar[q]) {
            const std::vector<double> rf[3] = {P->seq_rf[0][q], P->seq_rf[1][q], P->seq_rf[2][q]};
            copy_vvar_out(P->seq_vvar[q], rf, P->Tx, vvar[q]);
        }
    }
    return MRS_OK;
}

mrs_status mrs_DoMultiPoolBatchAtGPU(int32_t n_seq, const mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                                     const mrs_vctl *vctl, const mrs_vseq *const *vseq, mrs_vvar *const *vvar,
                                     mrs_vsig *const *vsig, mrs_vobj *const *vobj_out, const mrs_options *opt) {
    mrs_plan *P = nullptr;
    mrs_status rc = mrs_multipool_batch_create(n_seq, vobj, vmag, vcoi, vctl, vseq, (const mrs_vvar *const *)vvar,
                                               (const mrs_vsig *const *)vsig, opt, &P);
    if (rc) return rc;
    rc = mrs_plan_run(P);
    if (!rc) rc = mrs_multipool_batch_fetch(P, vobj_out, vvar, vsig);
    mrs_plan_destroy(P);
    return rc;
}

// One call with callbacks: decode until an Ext code that needs the caller,
// run the steps decoded so far, hand over to the callback, continue.
// cpu_engine: the state is always fetched, as the CPU engine's plugins share
// its arrays (DoScanAtCPU.cpp:360-405); otherwise gpu_fetch() decides, or
// DoGPUFetch.m's table when it is not set.
static mrs_status scan_with_callbacks(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                                      const mrs_vctl *vctl, const mrs_vseq *vseq, mrs_vvar *vvar,
                                      mrs_vsig *vsig, const mrs_callbacks *cb, const mrs_options *opt,
                                      bool cpu_engine) {
    mrs_plan *P = new (std::nothrow) mrs_plan();
    if (!P) return fail(MRS_ERR_NOMEM, "host allocation failed");
    struct Guard {
        mrs_plan *p;
        ~Guard() { mrs_plan_destroy(p); }
    } guard{P};
    mrs_status rc = plan_setup(P, vobj, vmag, vcoi, vctl, vseq, vsig, opt);
    if (rc) return rc;
    if ((rc = reset_state(P))) return rc;
    // the callback may re-point VMag and VCoi arrays (MATLAB reallocates the
    // globals it edits); later re-uploads read these copies
    mrs_vmag vm_cur = *vmag;
    mrs_vcoi vc_cur = *vcoi;
    Decoder dec(vobj, vcoi, vctl, vseq, vvar, true);
    Decoded all;  // accumulated k-space, muts and capture slots across chunks
    int64_t adc_base = 0, flushed = 0;
    for (;;) {
        P->dec = Decoded();
        P->dec.tx = P->Tx;
        P->dec.coil_words = P->model >= 1 ? 4 : 2;
        P->dec.stride = rec_stride(P->Tx, P->dec.coil_words);
        if (!all.muts.empty()) P->dec.muts = all.muts;
        const DecodeStop stop = dec.run(P->dec);
        if (stop == DECODE_END) dec.flush_tail(P->dec);
            compress_runs(P->dec, P->model == 0 ? P->max_adc : 0);
            choose_deferral(P);
        fold_exact(P, P->dec, (int64_t)P->gx * P->ngroups * P->block, &P->n_loops);
        if (adc_base + (int64_t)P->dec.adc_ref.size() > P->signal_num)
            return fail(MRS_ERR_ARG, "VSig.SignalNum is smaller than the number of ADC samples");
        if ((rc = upload_program(P)) || (rc = ensure_loop_cols(P, (int64_t)P->gx * P->ngroups * P->block)))
            return rc;
        // the launches of this chunk write samples adc_base + [0, n)
        for (auto &a : P->launch_adc0) a += adc_base;
        {
            // adc_ref indices are chunk-local; shift the table by uploading it at its offset
            std::vector<double> full(adc_base + P->dec.adc_ref.size());
            std::copy(P->dec.adc_ref.begin(), P->dec.adc_ref.end(), full.begin() + adc_base);
            if (P->adc_ref.p) (void)hipFree(P->adc_ref.p);
            P->adc_ref.p = nullptr;
            HIP_TRY(hipMalloc(&P->adc_ref.p, std::max<size_t>(full.size(), 1) * sizeof(double)));
            if (!full.empty())
                HIP_TRY(hipMemcpy(P->adc_ref.p, full.data(), full.size() * sizeof(double), hipMemcpyHostToDevice));
            P->adc_cap = 0;
        }
        if ((rc = execute(P, adc_base))) return rc;
        HIP_TRY(hipStreamSynchronize(P->stream));
        adc_base += (int64_t)P->dec.adc_ref.size();
        all.kx.insert(all.kx.end(), P->dec.kx.begin(), P->dec.kx.end());
        all.ky.insert(all.ky.end(), P->dec.ky.begin(), P->dec.ky.end());
        all.kz.insert(all.kz.end(), P->dec.kz.begin(), P->dec.kz.end());
        all.cap_slot.insert(all.cap_slot.end(), P->dec.cap_slot.begin(), P->dec.cap_slot.end());
        all.muts = P->dec.muts;
        if (stop == DECODE_END) break;
        // the samples so far go into VSig before the callback (DoScanAtGPU.cu:527-557)
        if ((rc = add_signal(P, vsig, flushed, adc_base))) return rc;
        for (int64_t n = flushed; n < adc_base; n++) {
            vsig->Kz[n] += all.kz[(size_t)n];
            vsig->Ky[n] += all.ky[(size_t)n];
            vsig->Kx[n] += all.kx[(size_t)n];
        }
        flushed = adc_base;
        const double ext = dec.ext_code();
        int fetch = 0;
        if (cpu_engine) {
            fetch = 1;
        } else if (cb->gpu_fetch) {  // DoGPUFetch.m
            if (cb->gpu_fetch(cb->user, ext, &fetch) != 0) return fail(MRS_ERR_CALLBACK, "gpu_fetch callback failed");
        } else {
            fetch = ext == 6 || ext == 8;  // DoGPUFetch.m:12-15
        }
        if (fetch) {
            float *dst[3] = {vobj->Mx, vobj->My, vobj->Mz};
            for (int k = 0; k < 3; k++)
                HIP_TRY(hipMemcpy(dst[k], P->M[k].p, P->ntot * sizeof(float), hipMemcpyDeviceToHost));
        }
        mrs_ext_state es;
        std::memset(&es, 0, sizeof es);
        es.vobj = vobj;
        es.vmag = &vm_cur;
        es.vcoi = &vc_cur;
        es.vsig = vsig;
        es.fetched = fetch != 0;
        dec.vvar().gpuFetch = fetch != 0;
        if (cb->ext && cb->ext(cb->user, ext, &dec.vvar(), &es) != 0)
            return fail(MRS_ERR_CALLBACK, "Extended process encounters ERROR!");
        dec.vvar().gpuFetch = 0;  // DoScanAtGPU.cu:686-687
        if (fetch && (rc = refresh_maps(P, vobj, &vm_cur, &vc_cur))) return rc;
    }
    // outputs: the samples from `flushed` on
    P->sig_from = flushed;
    P->dec.kx.assign(all.kx.begin() + flushed, all.kx.end());
    P->dec.ky.assign(all.ky.begin() + flushed, all.ky.end());
    P->dec.kz.assign(all.kz.begin() + flushed, all.kz.end());
    P->dec.cap_slot.swap(all.cap_slot);
    P->dec.muts.swap(all.muts);
    P->vvar_final = dec.vvar();
    P->rf_final[0].assign(dec.vvar().rfAmp, dec.vvar().rfAmp + P->Tx);
    P->rf_final[1].assign(dec.vvar().rfPhase, dec.vvar().rfPhase + P->Tx);
    P->rf_final[2].assign(dec.vvar().rfFreq, dec.vvar().rfFreq + P->Tx);
    P->ran = false;
    return mrs_plan_fetch(P, vobj, vvar, vsig);
}

static mrs_status scan_entry(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt, bool cpu_engine) {
    mrs_status rc = check_args(vobj, vmag, vcoi, vctl, vseq, vvar, vsig);
    if (rc) return rc;
    if (cb && (cb->ext || cb->gpu_fetch))
        return scan_with_callbacks(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt, cpu_engine);
    mrs_plan *P = nullptr;
    if ((rc = mrs_plan_create(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, opt, &P))) return rc;
    rc = mrs_plan_run(P);
    if (!rc) rc = mrs_plan_fetch(P, vobj, vvar, vsig);
    mrs_plan_destroy(P);
    return rc;
}

mrs_status mrs_DoScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                           const mrs_options *opt) {
    return scan_entry(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt, false);
}

mrs_status mrs_DoGMScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt) {
    mrs_options o;
    std::memset(&o, 0, sizeof o);
    if (opt) o = *opt;
    o.model = 1;
    return mrs_DoScanAtGPU(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, &o);
}

static mrs_status multipool_entry(int model, mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                                  const mrs_vctl *vctl, const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig,
                                  const mrs_callbacks *cb, const mrs_options *opt) {
    mrs_options o;
    std::memset(&o, 0, sizeof o);
    if (opt) o = *opt;
    o.model = model;
    return mrs_DoScanAtGPU(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, &o);
}

mrs_status mrs_DoMTScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt) {
    return multipool_entry(2, vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt);
}

mrs_status mrs_DoCESTScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                               const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                               const mrs_options *opt) {
    return multipool_entry(3, vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt);
}

mrs_status mrs_DoMEScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt) {
    return multipool_entry(4, vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt);
}

mrs_status mrs_DoScanAtCPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, mrs_vvar *vvar, mrs_vsig *vsig, const mrs_callbacks *cb,
                           const mrs_options *opt) {
    return scan_entry(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, opt, true);
}

}  // extern "C"
