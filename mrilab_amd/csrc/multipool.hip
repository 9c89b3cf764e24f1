// Multi-pool (generalized model) Bloch kernel for gfx950.
//
// MRiLab's GM engine (Lib/src/engine/GPUEngine/BlochKernelGM.cuh:6-487)
// evolves TypeNum pools per isochromat: free pools (TypeFlag 0) rotate,
// relax and exchange; bound pools (TypeFlag 1) carry only Mz and are
// saturated by rf through a super-Lorentzian line. Two schemes:
//   SCHEME 0: the reference's first-order operator split in fp32, in its
//             statement order (rotate, relax, then exchange with the
//             start-of-step magnetization of the other pools);
//   SCHEME 1: the exact solution of the same linear Bloch-McConnell system
//             over each held step, in fp64: the Taylor series of the action
//             of exp(A*dt) on the state, with substeps of h*|A| <= 4 and the
//             structured generator applied directly (3x3 rotation-relaxation
//             blocks, scalar bound-pool rows, diagonal exchange couplings).
// One thread owns one isochromat of one voxel with all its pools; all pool
// state stays in registers (fp64 for SCHEME 1) across the launch.
//
// The MT, CEST and ME engines (BlochKernelMT.cuh, BlochKernelCEST.cuh,
// BlochKernelME.cuh) are variants of the same model, with K converted to one
// rate matrix on the host:
//   ME:   GM with every pool free and a nonzero diagonal of K applied;
//   MT:   water (pool 0) and a bound pool (1) exchanging longitudinally only,
//         water relaxing through the combined exponent exp(-dt/T1 - dt*k);
//   CEST: MT plus free CEST pools t >= 2 exchanging with water only.
// MT and CEST skip the voxel on the water pool's maps and sample water only
// (WC: the water-centred split of those two kernels).
#include "multipool.hpp"

#include "decode.hpp"
#include "device_common.hpp"

namespace mrs {

__device__ __forceinline__ float expf_cr(float x) { return (float)exp((double)x); }

// Bound-pool lineshape G(dw) as BlochKernelGM.cuh:339-344 evaluates it: a
// 1000-point Riemann sum over u = 0, 0.001, ... (u accumulated by repeated
// addition); T is float for the split scheme and double for the exact one.
template <typename T>
__device__ T superlorentzian(double dw, double t2b) {
    T u = 0, G = 0;
    const T c = (T)sqrt(2.0 / 3.14159265359);
    const T w = (T)dw, t2 = (T)t2b;
    for (int k = 0; k < 1000; k++) {
        const T den = (T)3 * u * u - (T)1;
        const T b = (w * t2) / den;
        T e;
        if constexpr (sizeof(T) == 4)
            e = expf_cr(-2 * b * b);
        else
            e = exp(-2 * b * b);
        G += (c * (t2 / fabs(den)) * e) * (T)0.001;
        u += (T)0.001;
    }
    return G;
}

// longest Taylor series of one exact substep (h*|A| <= 4 needs about 35 terms)
constexpr int kTaylorMax = 48;
// 1/k for the Taylor coefficients h/k (a uniform index: read through the
// scalar cache instead of a division per term)
__constant__ const double kInvK[kTaylorMax + 1] = {
    0.0, 1.0 / 1, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7, 1.0 / 8, 1.0 / 9, 1.0 / 10,
    1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17, 1.0 / 18, 1.0 / 19,
    1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26, 1.0 / 27, 1.0 / 28,
    1.0 / 29, 1.0 / 30, 1.0 / 31, 1.0 / 32, 1.0 / 33, 1.0 / 34, 1.0 / 35, 1.0 / 36, 1.0 / 37,
    1.0 / 38, 1.0 / 39, 1.0 / 40, 1.0 / 41, 1.0 / 42, 1.0 / 43, 1.0 / 44, 1.0 / 45, 1.0 / 46,
    1.0 / 47, 1.0 / 48};

// signal values per isochromat and ADC sample: 2*nrx per free pool, padded
// to a power of two
#define NVT_OF(TY, NRX) (gm_nvt((TY), (NRX)))
constexpr int gm_nvt(int ty, int nrx) {
    return (2 * nrx * ty <= 2) ? 2 : (2 * nrx * ty <= 4) ? 4 : (2 * nrx * ty <= 8) ? 8 : (2 * nrx * ty <= 16) ? 16
           : (2 * nrx * ty <= 32) ? 32 : 64;
}

template <int TY, typename T>
struct GmState {
    T mx[TY], my[TY], mz[TY];
};

// Waves per SIMD the kernel is compiled for (the register budget): measured
// on the 256x256 3-pool phantom with 40 batched offsets, the split scheme
// runs 20 % faster at 3 waves (168 VGPRs) than at 2, and the exact scheme
// 23 % faster at 2 (256 VGPRs, a few values spilled outside the Taylor
// loop) than at 1; 3 waves for the exact scheme spill inside the series and
// ran 44 % slower. Above 3 pools the budget of 2 waves (split) or 1-2
// (exact) avoids heavy spills.
constexpr int gm_waves_min(int ty, int scheme) { return scheme == 0 ? (ty <= 3 ? 3 : 1) : (ty <= 3 ? 2 : 1); }
constexpr int gm_waves_max(int ty, int scheme) { return scheme == 0 ? (ty <= 3 ? 3 : 2) : 2; }

// Loop records of the exact scheme (kept out of line: they run once per
// basis state and loop, and inlined they would hold registers across the
// whole step loop). Entry e of a thread's slots is at e * gs.
// Column b of the block's map: the basis image less the constant c (b = 0
// stores c itself).
template <int D>
__device__ __attribute__((noinline)) void loop_store_column(double *colA, double *colC, int64_t gs, int b,
                                                            const double *sv) {
    if (b == 0) {
        for (int i = 0; i < D; i++) colC[(int64_t)i * gs] = sv[i];
    } else {
        for (int i = 0; i < D; i++) colA[((int64_t)(b - 1) * D + i) * gs] = sv[i] - colC[(int64_t)i * gs];
    }
}

// Zero columns of A for the components the block is not run for (bit i
// of mask clear): their state entries are zero, so the columns only need
// to be finite.
template <int D>
__device__ __attribute__((noinline)) void loop_zero_columns(double *colA, int64_t gs, uint32_t mask) {
    for (int b = 0; b < D; b++)
        if (!((mask >> b) & 1u))
            for (int i = 0; i < D; i++) colA[((int64_t)b * D + i) * gs] = 0.0;
}

// The waiting state mapped n times by x -> A x + c, into xs.
template <int D>
__device__ __attribute__((noinline)) void loop_apply(const double *colA, const double *colC, const double *keep,
                                                     int64_t gs, int n, double *xs) {
    double x[D];
#pragma unroll
    for (int i = 0; i < D; i++) x[i] = keep[(int64_t)i * gs];
    for (int k = 0; k < n; k++) {
        double y[D];
#pragma unroll
        for (int i = 0; i < D; i++) y[i] = colC[(int64_t)i * gs];
#pragma unroll 1
        for (int b = 0; b < D; b++) {
            double xb = x[0];
#pragma unroll
            for (int i = 1; i < D; i++) xb = b == i ? x[i] : xb;
#pragma unroll
            for (int i = 0; i < D; i++) y[i] += xb * colA[((int64_t)b * D + i) * gs];
        }
#pragma unroll
        for (int i = 0; i < D; i++) x[i] = y[i];
    }
#pragma unroll
    for (int i = 0; i < D; i++) xs[i] = x[i];
}

template <int TY, int NRX, int SCHEME, bool WC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gm_waves_min(TY, SCHEME),
                                                                     gm_waves_max(TY, SCHEME)))) void bloch_gm_kernel(
    GmParams p, const uint32_t *__restrict__ rec_g, int r0, int r1) {
    using T = typename std::conditional<SCHEME == 0, float, double>::type;
    constexpr int NV = 2 * NRX;
    constexpr int NVT = NVT_OF(TY, NRX);
    static_assert(NV * TY <= 64, "too many signal values per isochromat");
    extern __shared__ float red[];  // [kRedBatch][nwaves][NVT]
    const crec_t rec = (crec_t)rec_g;
    const int cw = 4;
    const int stride = p.stride;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int64_t nvox = p.nvox;
    const int s = blockIdx.y;
    // batched sequences: block x is (tile, sequence), sequence-minor so the
    // costliest tiles of every sequence are dispatched first
    const bool batched = p.seq_tab != nullptr;
    const int tile = batched ? (int)blockIdx.x / p.nseq_l : (int)blockIdx.x;
    const int sq = batched ? p.seq_base + (int)blockIdx.x % p.nseq_l : 0;
    if (batched) {
        r0 = p.seq_tab[4 * sq];
        r1 = p.seq_tab[4 * sq + 1];
    }
    float *const partial = batched ? p.partial + (int64_t)p.seq_tab[4 * sq + 2] * p.tiles * gridDim.y * NVT_OF(TY, NRX)
                                   : p.partial;
    const int64_t vt = (int64_t)tile * blockDim.x + threadIdx.x;
    const bool valid = vt < nvox;
    const int64_t vv = !valid ? (p.vox_perm ? p.vox_perm[nvox - 1] : nvox - 1) : (p.vox_perm ? p.vox_perm[vt] : vt);
    const int nblk = (batched ? p.tiles : (int)gridDim.x) * gridDim.y;
    const int blk = blockIdx.y * (batched ? p.tiles : (int)gridDim.x) + tile;

    const float x = p.gx[vv], y = p.gy[vv], z = p.gz[vv];
    const float dB0 = p.dB0[vv];
    const float w0 = dB0 * p.gyro;
    GmState<TY, T> st;
    int64_t sidx[TY];
    float rho[TY], t1[TY], t2[TY], wb[TY];
    bool skip[TY];
    float kmat[TY][TY];  // kmat[t][q]: rate from pool t into pool q
#pragma unroll
    for (int t = 0; t < TY; t++) {
        sidx[t] = vv + nvox * ((int64_t)s + (int64_t)p.S * t);
        st.mx[t] = p.Mx[sidx[t]];
        st.my[t] = p.My[sidx[t]];
        st.mz[t] = p.Mz[sidx[t]];
        const int64_t tv = vv + (int64_t)t * nvox;
        rho[t] = p.Rho[tv];
        t1[t] = p.T1[tv];
        t2[t] = p.T2[tv];
        skip[t] = t1[t] == 0.f || t2[t] == 0.f || rho[t] == 0.f;  // BlochKernelGM.cuh:176
        wb[t] = (w0 + p.dWRnd[sidx[t]]) + p.cs[t];
#pragma unroll
        for (int q = 0; q < TY; q++)
            kmat[t][q] = (q == t && !p.self_x) ? 0.f : p.K[vv + nvox * ((int64_t)q * TY + t)];
    }
    // MT / CEST: the water pool's maps decide for the whole voxel
    const bool wcv = p.variant == kVarMT || p.variant == kVarCEST;
    if (WC || wcv) {
#pragma unroll
        for (int t = 1; t < TY; t++) skip[t] = skip[0];
    }
    // exact scheme: the constant parts of each pool's rows of the generator
    // (BlochKernelGM.cuh's rates, exchange K, T1/T2), arranged so that a
    // Taylor term is the same straight-line code for every pool: a skipped
    // pool has all-zero rows (its state stays as it is), a bound pool zero
    // transverse rows, and MT/CEST water exchanges no transverse
    // magnetization with the bound pool (xm)
    using ET = typename std::conditional<SCHEME == 1, double, float>::type;
    ET kz[TY][TY];   // kz[t][q]: inflow rate q -> t into Mz of t
    ET kt[TY][TY];   // the same into Mx, My of t (zero unless t is a free pool)
    ET dxc[TY], dzc[TY], affc[TY], fm[TY], bm[TY], xm[TY], rsum[TY];
    if constexpr (SCHEME == 1) {
#pragma unroll
        for (int t = 0; t < TY; t++) {
            const bool fr = p.flag[t] == 0;
            const bool on = !skip[t];
            double ko = 0.0, kox = 0.0, kin = 0.0;
#pragma unroll
            for (int q = 0; q < TY; q++) {
                if (q == t) continue;  // self-exchange cancels exactly
                ko += kmat[t][q];
                if (!(wcv && p.flag[q] != 0)) kox += kmat[t][q];  // MT/CEST: no transverse loss to bound
                kin += kmat[q][t];
            }
            const double r1 = 1.0 / (double)t1[t];
            const double r2 = fr ? 1.0 / (double)t2[t] : 0.0;
#pragma unroll
            for (int q = 0; q < TY; q++) {
                kz[t][q] = (on && q != t) ? (double)kmat[q][t] : 0.0;
                kt[t][q] = fr ? kz[t][q] : 0.0;
            }
            dxc[t] = (on && fr) ? r2 + kox : 0.0;
            dzc[t] = on ? r1 + ko : 0.0;
            affc[t] = on ? r1 * (double)rho[t] / p.S : 0.0;
            fm[t] = (on && fr) ? 1.0 : 0.0;
            bm[t] = (on && !fr) ? 1.0 : 0.0;
            xm[t] = (wcv && !fr) ? 0.0 : 1.0;
            rsum[t] = on ? r1 + r2 + ko + kin : 0.0;
        }
    }
    (void)kz;
    (void)kt;
    (void)dxc;
    (void)dzc;
    (void)affc;
    (void)fm;
    (void)bm;
    (void)xm;
    (void)rsum;
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
    // per-dt factors of the split scheme
    float e1[TY], e2[TY], m0[TY], kout[TY], fin[TY][TY];
#pragma unroll
    for (int t = 0; t < TY; t++) {
        e1[t] = e2[t] = kout[t] = 1.f;
        m0[t] = 0.f;
#pragma unroll
        for (int q = 0; q < TY; q++) fin[t][q] = 0.f;
    }
    // lineshape cache per pool (key: rf frequency, T2 of the bound pool)
    float gkey[TY];
    double gval[TY];
    bool gok[TY];
#pragma unroll
    for (int t = 0; t < TY; t++) gok[t] = false;
    bool have_e = false;
    int slot = 0;
    int adc_batch0 = 0;
    // first Tx coil of this voxel (magnitude, sin and cos of its phase):
    // constant over the launch, so rf steps read no maps
    const float tx_mg0 = p.txmg[vv];
    float tx_sp0, tx_cp0;
    sincos_cr(p.txpe[vv], &tx_sp0, &tx_cp0);
    // free-precession rotation of each free pool, kept while the steps
    // repeat with the same dt and fields (REC_SAME) and for folded repeats
    float fsn[TY], fcs[TY];
    bool have_frot = false;

    // the step of record q: dt, gradient phase rate, rf frequency, effective
    // B1 (gamma-scaled) and the bound pools' saturation rates
    // (gv: the summed gradient term for the fp64 schemes; gt: the three
    // terms the split scheme adds in the reference's order)
    auto rec_fields = [&](const crec_t q, uint32_t flags, float &dt, bool &rf, float &gv, GradTerms &gt, float &wrf,
                          float &bx, float &by, float (&wsat)[TY]) __attribute__((always_inline)) {
        dt = __uint_as_float(q[0]);
        rf = (flags & REC_RF) != 0;
        gv = z * __uint_as_float(q[1]) + y * __uint_as_float(q[2]) + x * __uint_as_float(q[3]);
        gt = grad_terms_xyz(z, y, x, __uint_as_float(q[1]), __uint_as_float(q[2]), __uint_as_float(q[3]));
        wrf = __uint_as_float(q[4]);
        // effective B1 (gamma-scaled) and bound-pool saturation rates
        bx = 0.f;
        by = 0.f;
#pragma unroll
        for (int t = 0; t < TY; t++) wsat[t] = 0.f;
        if (rf) {
            double wsum = 0.0;
            int nact = 0;
            for (int c = 0; c < p.Tx; c++) {
                const float br = __uint_as_float(q[REC_HEAD + cw * c]);
                const float bi = __uint_as_float(q[REC_HEAD + cw * c + 1]);
                const float A = __uint_as_float(q[REC_HEAD + cw * c + 2]);
                const float F = __uint_as_float(q[REC_HEAD + cw * c + 3]);
                if (A == 0.f) continue;
                float sp = tx_sp0, cp = tx_cp0, mg = tx_mg0;
                if (c > 0) {
                    sincos_cr(p.txpe[vv + (int64_t)c * nvox], &sp, &cp);
                    mg = p.txmg[vv + (int64_t)c * nvox];
                }
                bx += mg * (cp * br - sp * bi);
                by += mg * (cp * bi + sp * br);
                // BlochKernelGM.cuh:323-380: W = pi*(gamma*B1)^2*G, averaged over active coils
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    if (p.flag[t] != 1 || skip[t]) continue;
                    if (!gok[t] || gkey[t] != F) {
                        gkey[t] = F;
                        gok[t] = true;
                        gval[t] = (double)superlorentzian<T>(2 * 3.14159265359 * (double)F + (double)dB0 * p.gyro,
                                                             (double)t2[t]);
                    }
                    const float bg = (mg * A) * p.gyro;
                    if (SCHEME == 0)
                        wsat[t] += (float)(3.14159265359f * bg * bg * (float)gval[t]);
                    else
                        wsat[t] += (float)(3.14159265359 * (double)bg * bg * gval[t]);
                }
                wsum += 1.0;
                nact++;
            }
            if (nact > 1) {
#pragma unroll
                for (int t = 0; t < TY; t++) wsat[t] = wsat[t] / (float)nact;
            }
            (void)wsum;
        }
    };
    // exact step of record q (not a no-step record) on the state sv (pool t
    // at sv[3t .. 3t+2]), its folded repeats included
    auto exact_record = [&](const crec_t q, uint32_t flags, double *sv) __attribute__((always_inline)) {
        if constexpr (SCHEME == 1) {
            float dt, gv, wrf, bx, by;
            GradTerms gt;
            bool rf;
            float wsat[TY];
            rec_fields(q, flags, dt, rf, gv, gt, wrf, bx, by, wsat);
            const int nrep = 1 + (int)(flags >> REC_REP_SHIFT);
            // exact step: x <- exp(A*dt) x with the affine M0 term; per step
            // the rotation of each free pool (zero for others) and the bound
            // pools' saturation
            double omt[TY], bxt[TY], byt[TY], dzs[TY];
            double nrm = 0.0;
#pragma unroll
            for (int t = 0; t < TY; t++) {
                omt[t] = fm[t] * ((double)wb[t] + (double)gv + (rf ? (double)wrf : 0.0));
                bxt[t] = fm[t] * (double)bx;
                byt[t] = fm[t] * (double)by;
                dzs[t] = dzc[t] + bm[t] * (double)wsat[t];
                // the row sum of pool t's rows bounds |A| (max row sum)
                nrm = fmax(nrm, rsum[t] + fabs(omt[t]) + fabs(bxt[t]) + fabs(byt[t]) + bm[t] * (double)wsat[t]);
            }
            // substeps of h*|A| <= 4; the series runs until every active lane's
            // term is below 1e-13 of its state (at most kTaylorMax terms): the
            // truncation adds at most 1e-13 relative per step, 2e-8 over the
            // 200,000 steps of a 2 s saturation train, under the fp32
            // rounding of the state kept between launches and of the output
            int nsub = (int)ceil(nrm * (double)dt * 0.25);
            if (nsub < 1) nsub = 1;
            const double h = (double)dt / nsub;
            const double hn = h * nrm;
            double amax = 0.0;
#pragma unroll
            for (int t = 0; t < TY; t++) amax = fmax(amax, fabs(affc[t]));
            auto apply = [&](const double *in, double *out) __attribute__((always_inline)) {
                double ixm[TY], iym[TY];
#pragma unroll
                for (int q = 0; q < TY; q++) {
                    ixm[q] = in[3 * q] * xm[q];
                    iym[q] = in[3 * q + 1] * xm[q];
                }
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    const double ix = in[3 * t], iy = in[3 * t + 1], iz = in[3 * t + 2];
                    double ox = omt[t] * iy - byt[t] * iz - dxc[t] * ix;
                    double oy = bxt[t] * iz - omt[t] * ix - dxc[t] * iy;
                    double oz = byt[t] * ix - bxt[t] * iy - dzs[t] * iz;
#pragma unroll
                    for (int q = 0; q < TY; q++) {
                        if (q == t) continue;
                        ox += kt[t][q] * ixm[q];
                        oy += kt[t][q] * iym[q];
                        oz += kz[t][q] * in[3 * q + 2];
                    }
                    out[3 * t] = ox;
                    out[3 * t + 1] = oy;
                    out[3 * t + 2] = oz;
                }
            };
            // x <- exp(A*dt) x + (exp(A*dt) - I) A^-1 b, substep by substep:
            // x += sum_k h^k/k! A^(k-1) (A x + b)
            auto step_exact = [&](double *xv) __attribute__((always_inline)) {
                for (int sub = 0; sub < nsub; sub++) {
                    double term[3 * TY], nt[3 * TY];
                    double xmax = 0.0;
#pragma unroll
                    for (int i = 0; i < 3 * TY; i++) xmax = fmax(xmax, fabs(xv[i]));
                    const double tol = 1e-13 * (xmax + h * amax);
                    apply(xv, nt);
#pragma unroll
                    for (int t = 0; t < TY; t++) nt[3 * t + 2] += affc[t];
#pragma unroll
                    for (int i = 0; i < 3 * TY; i++) {
                        term[i] = h * nt[i];
                        xv[i] += term[i];
                    }
                    for (int kk = 2; kk <= kTaylorMax; kk++) {
                        apply(term, nt);
                        const double f = h * kInvK[kk];
                        double tm = 0.0;
#pragma unroll
                        for (int i = 0; i < 3 * TY; i++) {
                            term[i] = f * nt[i];
                            xv[i] += term[i];
                            tm = fmax(tm, fabs(term[i]));
                        }
                        // terms shrink monotonically once kk exceeds h*|A|
                        if (__all((double)kk > hn && tm <= tol)) break;
                    }
                }
            };
            if (!rf && nrep > 1) {
                // held free step: the map of one step is affine and the same for
                // every repeat; take its columns from basis states (transverse
                // columns are complex: without rf, Mx + iMy of the free pools
                // evolve complex-linearly) and apply it nrep times
                double cz[TY];   // the affine constant (longitudinal only)
                double tr[TY][TY], ti[TY][TY], lz[TY][TY];
                // basis b = 0: zero state; 1 + t: unit Mx of pool t; 1 + TY + t:
                // unit Mz of pool t. Every basis state is stepped (skipped and
                // bound pools included), so the columns are those of the direct
                // step; one copy of the step code, entries stored by
                // compile-time index.
#pragma unroll 1
                for (int b = 0; b <= 2 * TY; b++) {
                    const bool bz = b > TY;
                    const int tb = bz ? b - 1 - TY : b - 1;
                    const int ib = bz ? 3 * tb + 2 : 3 * tb;
                    double e[3 * TY];
#pragma unroll
                    for (int i = 0; i < 3 * TY; i++) e[i] = (b > 0 && i == ib) ? 1.0 : 0.0;
                    step_exact(e);
#pragma unroll
                    for (int q = 0; q < TY; q++) {
                        if (b == 0) cz[q] = e[3 * q + 2];
#pragma unroll
                        for (int t = 0; t < TY; t++) {
                            if (b > 0 && t == tb) {
                                if (bz) {
                                    lz[q][t] = e[3 * q + 2] - cz[q];
                                } else {
                                    tr[q][t] = e[3 * q];
                                    ti[q][t] = e[3 * q + 1];
                                }
                            }
                        }
                    }
                }
                for (int k = 0; k < nrep; k++) {
                    double ox[TY], oy[TY], oz[TY];
#pragma unroll
                    for (int t = 0; t < TY; t++) {
                        ox[t] = sv[3 * t];
                        oy[t] = sv[3 * t + 1];
                        oz[t] = sv[3 * t + 2];
                    }
#pragma unroll
                    for (int q = 0; q < TY; q++) {
                        double ax = 0.0, ay = 0.0, az = cz[q];
#pragma unroll
                        for (int t = 0; t < TY; t++) {
                            ax += tr[q][t] * ox[t] - ti[q][t] * oy[t];
                            ay += tr[q][t] * oy[t] + ti[q][t] * ox[t];
                            az += lz[q][t] * oz[t];
                        }
                        sv[3 * q] = ax;
                        sv[3 * q + 1] = ay;
                        sv[3 * q + 2] = az;
                    }
                }
            } else {
                for (int k = 0; k < nrep; k++) step_exact(sv);
            }
        }
    };
    // loop records (exact scheme): the block [loop_start, loop_end] runs for
    // basis state loop_b (0 the zero state, b the unit state of component
    // b - 1; -1: no loop open); this thread's slots of
    // loop_cols hold A ((3 TY)^2 entries), c (3 TY) and the waiting state
    // (3 TY), entry e at e * (threads of the grid)
    int loop_start = 0, loop_end = -1, loop_n = 0, loop_b = -1;
    uint32_t loop_mask = 0;  // components whose basis states run
    (void)loop_mask;
    const int64_t lgs = (int64_t)gridDim.x * gridDim.y * blockDim.x;
    double *const lcolA = p.loop_cols ? p.loop_cols + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x +
                                             threadIdx.x
                                      : nullptr;
    double *const lcolC = lcolA ? lcolA + (int64_t)9 * TY * TY * lgs : nullptr;
    double *const lkeep = lcolC ? lcolC + (int64_t)3 * TY * lgs : nullptr;
    (void)loop_start;
    (void)loop_end;
    (void)loop_n;
    (void)loop_b;
    (void)lkeep;
    for (int r = r0; r < r1; ++r) {
        const crec_t q = rec + r * stride;
        const uint32_t flags = q[5];
        if (flags & REC_SPOIL) {  // Plugin_IdealSpoiler.m:10-11 zeroes every pool
#pragma unroll
            for (int t = 0; t < TY; t++) {
                st.mx[t] = 0;
                st.my[t] = 0;
            }
        }
        if (flags & REC_ADC) {  // free pools only (BlochKernelGM.cuh:192)
            float val[NVT];
#pragma unroll
            for (int j = 0; j < NVT; j++) val[j] = 0.f;
#pragma unroll
            for (int t = 0; t < TY; t++) {
                if (WC || wcv) {
                    if (t != 0) continue;  // water only (BlochKernelMT.cuh:230)
                } else if (p.flag[t] != 0) {
                    continue;
                }
                const float sx = (valid && !skip[t]) ? (float)st.mx[t] : 0.f;
                const float sy = (valid && !skip[t]) ? (float)st.my[t] : 0.f;
#pragma unroll
                for (int c = 0; c < NRX; c++) {
                    val[t * NV + 2 * c] = sx * crx[c] + sy * cry[c];
                    val[t * NV + 2 * c + 1] = sy * crx[c] - sx * cry[c];
                }
            }
            WaveSum<NVT, 32>::run(val, lane);
            constexpr int G = 64 / NVT;
            if ((lane & (G - 1)) == 0) red[(slot * nwaves + wave) * NVT + lane / G] = val[0];
            if (++slot == kRedBatch) {
                __syncthreads();
                flush_partials(red, kRedBatch, nwaves, NVT, partial, adc_batch0, nblk, blk, NVT);
                __syncthreads();
                adc_batch0 += kRedBatch;
                slot = 0;
            }
        }
        if constexpr (SCHEME == 1) {
            constexpr int D = 3 * TY;
            if (flags & REC_LOOP) {
                // the block's affine map x -> A x + c: c is the image of the
                // zero state, column b of A the image of unit vector b less
                // c; the block runs once per basis state (the state itself
                // waits in loop_cols), then the map is applied N times
                loop_start = r + 1;
                loop_end = r + (int)q[6];
                loop_n = 1 + (int)(flags >> REC_REP_SHIFT);
                loop_b = 0;
                // a bound pool's transverse entries have zero rows; where the
                // whole wave holds zeros there, they stay zero and their
                // basis states need no run
                loop_mask = (1u << D) - 1;
#pragma unroll
                for (int t = 0; t < TY; t++)
                    if (p.flag[t] != 0 && __all(st.mx[t] == 0.0 && st.my[t] == 0.0))
                        loop_mask &= ~(3u << (3 * t));
                if (loop_mask != (1u << D) - 1) loop_zero_columns<D>(lcolA, lgs, loop_mask);
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    lkeep[(int64_t)(3 * t) * lgs] = st.mx[t];
                    lkeep[(int64_t)(3 * t + 1) * lgs] = st.my[t];
                    lkeep[(int64_t)(3 * t + 2) * lgs] = st.mz[t];
                    st.mx[t] = st.my[t] = st.mz[t] = 0.0;
                }
                continue;
            }
            if (!(flags & REC_NOSTEP)) {
                double sv[D];
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    sv[3 * t] = st.mx[t];
                    sv[3 * t + 1] = st.my[t];
                    sv[3 * t + 2] = st.mz[t];
                }
                exact_record(q, flags, sv);
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    st.mx[t] = sv[3 * t];
                    st.my[t] = sv[3 * t + 1];
                    st.mz[t] = sv[3 * t + 2];
                }
            }
            if (loop_b >= 0 && r == loop_end) {
                double sv[D];
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    sv[3 * t] = st.mx[t];
                    sv[3 * t + 1] = st.my[t];
                    sv[3 * t + 2] = st.mz[t];
                }
                loop_store_column<D>(lcolA, lcolC, lgs, loop_b, sv);
                const uint32_t rest = loop_mask >> loop_b;  // components from loop_b on
                if (rest != 0) {
                    loop_b += __builtin_ctz(rest) + 1;
                    // the next basis state through the block
#pragma unroll
                    for (int t = 0; t < TY; t++) {
                        st.mx[t] = (3 * t == loop_b - 1) ? 1.0 : 0.0;
                        st.my[t] = (3 * t + 1 == loop_b - 1) ? 1.0 : 0.0;
                        st.mz[t] = (3 * t + 2 == loop_b - 1) ? 1.0 : 0.0;
                    }
                    r = loop_start - 1;
                    continue;
                }
                double xs[D];
                loop_apply<D>(lcolA, lcolC, lkeep, lgs, loop_n, xs);
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    st.mx[t] = xs[3 * t];
                    st.my[t] = xs[3 * t + 1];
                    st.mz[t] = xs[3 * t + 2];
                }
                loop_b = -1;
            }
            continue;
        }
        if (flags & REC_NOSTEP) continue;
        float dt, gv, wrf, bx, by;
        GradTerms gt;
        bool rf;
        float wsat[TY];
        rec_fields(q, flags, dt, rf, gv, gt, wrf, bx, by, wsat);
        const int nrep = 1 + (int)(flags >> REC_REP_SHIFT);
        if constexpr (SCHEME == 0) {
            if (rf) {
                have_frot = false;
            } else if (!((flags & REC_SAME) && have_frot)) {
#pragma unroll
                for (int t = 0; t < TY; t++) sincos_cr(ref_dw(wb[t], gt) * dt, &fsn[t], &fcs[t]);
                have_frot = true;
            }
        }
        // rotation of free pool t over this step (BlochKernel.h:146-344 form)
        // (pool values passed in, so the arrays stay in registers)
        auto rotate_free = [&](float wbt, float fst, float fct, float &ax, float &ay,
                               float &az) __attribute__((always_inline)) {
            const float w = rf ? ref_dw(wbt, gt) + wrf : ref_dw(wbt, gt);
            if (rf) {
                // applied once: the hardware reciprocal square root and the
                // 1.5 ulp fp32 sincos (as the single-pool rf step); below
                // |Omega| = 1e-15 rad/s the rotation is skipped
                const float om2 = (bx * bx + by * by) + w * w;
                if (om2 > 1e-30f) {
                    const float inv = __builtin_amdgcn_rsqf(om2);
                    const float om = om2 * inv;
                    const float nx = bx * inv, ny = by * inv, nz = w * inv;
                    float sn, cs;
                    sincos_fast(om * dt, &sn, &cs);
                    const float kk = (1.f - cs) * (nx * ax + ny * ay + nz * az);
                    const float rx = ax * cs - (ny * az - nz * ay) * sn + nx * kk;
                    const float ry = ay * cs - (nz * ax - nx * az) * sn + ny * kk;
                    const float rz = az * cs - (nx * ay - ny * ax) * sn + nz * kk;
                    ax = rx;
                    ay = ry;
                    az = rz;
                }
            } else {
                // the cached sincos of the step's angle
                const float rx = ax * fct + ay * fst;
                const float ry = ay * fct - ax * fst;
                ax = rx;
                ay = ry;
            }
        };
        if constexpr (SCHEME == 0 && !WC) {
#pragma clang fp contract(off)
            if ((flags & REC_NEWDT) || !have_e) {
                have_e = true;
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    e1[t] = expf_cr(-dt / t1[t]);
                    e2[t] = expf_cr(-dt / t2[t]);
                    m0[t] = (rho[t] * (1.f - e1[t])) / (float)p.S;
                    float ko = 1.f;
#pragma unroll
                    for (int qq = 0; qq < TY; qq++) {  // the diagonal is zero except for ME
                        if (kmat[t][qq] != 0.f) ko = ko * expf_cr(-dt * kmat[t][qq]);
                        fin[t][qq] = kmat[qq][t] != 0.f ? 1.f - expf_cr(-dt * kmat[qq][t]) : 0.f;
                    }
                    kout[t] = ko;
                }
            }
            for (int k = 0; k < nrep; k++) {
                float Bx[TY], By[TY], Bz[TY];
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    Bx[t] = st.mx[t];
                    By[t] = st.my[t];
                    Bz[t] = st.mz[t];
                }
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    if (skip[t] || (p.flag[t] != 0 && p.flag[t] != 1)) continue;
                    float ax = Bx[t], ay = By[t], az = Bz[t];
                    if (p.flag[t] == 0) {
                        rotate_free(wb[t], fsn[t], fcs[t], ax, ay, az);
                        ax = ax * e2[t];
                        ay = ay * e2[t];
                    } else if (rf) {
                        az = az * expf_cr(-wsat[t] * dt);
                    }
                    az = az * e1[t];
                    // inflow from every pool: fin is 0 where nothing flows
                    // (adding 0 * B leaves the sum as it is; no per-lane branch)
                    float kx = 0.f, ky = 0.f, kz = 0.f;
#pragma unroll
                    for (int qq = 0; qq < TY; qq++) {
                        if (p.flag[t] == 0) {
                            kx = kx + fin[t][qq] * Bx[qq];
                            ky = ky + fin[t][qq] * By[qq];
                        }
                        kz = kz + fin[t][qq] * Bz[qq];
                    }
                    if (p.flag[t] == 0) {
                        st.mx[t] = ax * kout[t] + kx;
                        st.my[t] = ay * kout[t] + ky;
                    }
                    st.mz[t] = az * kout[t] + kz + m0[t];
                }
            }
        } else if constexpr (SCHEME == 0 && WC) {
#pragma clang fp contract(off)
            // MT (TY == 2) and CEST split: BlochKernelMT.cuh:322-414,
            // BlochKernelCEST.cuh:361-610
            if ((flags & REC_NEWDT) || !have_e) {
                have_e = true;
                const float k12 = kmat[0][1], k21 = kmat[1][0];
                e2[0] = expf_cr(-dt / t2[0]);
                e1[0] = expf_cr(-dt / t1[0] - dt * k12);  // water: one combined exponent
                m0[0] = (rho[0] * (1.f - expf_cr(-dt / t1[0]))) / (float)p.S;
                kout[0] = 1.f - expf_cr(-dt * k21);       // water gain from the bound pool
                e1[1] = expf_cr(-dt / t1[1] - dt * k21);
                m0[1] = (rho[1] * (1.f - expf_cr(-dt / t1[1]))) / (float)p.S;
                kout[1] = 1.f - expf_cr(-dt * k12);       // bound gain from water
#pragma unroll
                for (int t = 2; t < TY; t++) {
                    e2[t] = expf_cr(-dt / t2[t]);
                    e1[t] = expf_cr(-dt / t1[t]);
                    m0[t] = (rho[t] * (1.f - e1[t])) / (float)p.S;
                    fin[0][t] = kmat[0][t] != 0.f ? expf_cr(-dt * kmat[0][t]) : 1.f;  // water -> t survival
                    fin[t][0] = kmat[t][0] != 0.f ? expf_cr(-dt * kmat[t][0]) : 1.f;  // t -> water survival
                }
            }
            for (int k = 0; k < nrep && !skip[0]; k++) {
                float Bx[TY], By[TY], Bz[TY], ox[TY], oy[TY], oz[TY];
#pragma unroll
                for (int t = 0; t < TY; t++) {
                    Bx[t] = ox[t] = st.mx[t];
                    By[t] = oy[t] = st.my[t];
                    Bz[t] = oz[t] = st.mz[t];
                }
#pragma unroll
                for (int t = 0; t < TY; t++)
                    if (t != 1) rotate_free(wb[t], fsn[t], fcs[t], ox[t], oy[t], oz[t]);
                // water
                float ko = 1.f, kx = 0.f, ky = 0.f, kz = 0.f;
#pragma unroll
                for (int t = 2; t < TY; t++) {
                    // survival factors are 1 where no rate is set (ko * 1 and
                    // 0 * B change nothing; no per-lane branches)
                    ko = ko * fin[0][t];
                    const float f = 1.f - fin[t][0];
                    kx = kx + f * Bx[t];
                    ky = ky + f * By[t];
                    kz = kz + f * Bz[t];
                }
                if constexpr (TY == 2) {
                    st.mx[0] = ox[0] * e2[0];
                    st.my[0] = oy[0] * e2[0];
                    st.mz[0] = oz[0] * e1[0] + m0[0] + Bz[1] * kout[0];
                } else {
                    st.mx[0] = ox[0] * e2[0] * ko + kx;
                    st.my[0] = oy[0] * e2[0] * ko + ky;
                    st.mz[0] = oz[0] * e1[0] * ko + kz + Bz[1] * kout[0] + m0[0];
                }
                // CEST pools
#pragma unroll
                for (int t = 2; t < TY; t++) {
                    const float bxc = ox[t] * e2[t], byc = oy[t] * e2[t], bzc = oz[t] * e1[t];
                    const float kt = fin[t][0];
                    const float f = 1.f - fin[0][t];
                    const float ix = f * Bx[0], iy = f * By[0], iz = f * Bz[0];
                    st.mx[t] = bxc * kt + ix;
                    st.my[t] = byc * kt + iy;
                    st.mz[t] = bzc * kt + iz + m0[t];
                }
                // bound pool
                float bzb = Bz[1];
                if (rf) bzb = Bz[1] * expf_cr(-wsat[1] * dt);
                st.mz[1] = bzb * e1[1] + m0[1] + Bz[0] * kout[1];
            }
        }
    }
    if (slot > 0) {
        __syncthreads();
        flush_partials(red, slot, nwaves, NVT, partial, adc_batch0, nblk, blk, NVT);
    }
    if (valid) {
        const int64_t ob = batched ? (int64_t)sq * nvox * p.S * TY : 0;
        float *ox = batched ? p.Mout[0] + ob : p.Mx;
        float *oy = batched ? p.Mout[1] + ob : p.My;
        float *oz = batched ? p.Mout[2] + ob : p.Mz;
#pragma unroll
        for (int t = 0; t < TY; t++) {
            ox[sidx[t]] = (float)st.mx[t];
            oy[sidx[t]] = (float)st.my[t];
            oz[sidx[t]] = (float)st.mz[t];
        }
    }
}

template <int TY, int NRX, int SCHEME, bool WC>
static void launch_t(dim3 g, dim3 b, size_t lds, hipStream_t st, const GmParams &p, const uint32_t *rec, int r0,
                     int r1) {
    bloch_gm_kernel<TY, NRX, SCHEME, WC><<<g, b, lds, st>>>(p, rec, r0, r1);
}

template <int TY, int NRX>
static gm_launch_fn pick_scheme(int scheme, int variant) {
    if (scheme) return launch_t<TY, NRX, 1, false>;
    if (variant == kVarMT || variant == kVarCEST) return launch_t<TY, NRX, 0, true>;
    return launch_t<TY, NRX, 0, false>;
}

template <int TY>
static gm_launch_fn pick_nrx(int nrx, int scheme, int variant) {
    switch (nrx) {
        case 1: return pick_scheme<TY, 1>(scheme, variant);
        case 2: return pick_scheme<TY, 2>(scheme, variant);
        case 4: return pick_scheme<TY, 4>(scheme, variant);
        case 8:
            if constexpr (2 * 8 * TY <= 64) return pick_scheme<TY, 8>(scheme, variant);
            return nullptr;
    }
    return nullptr;
}

gm_launch_fn pick_gm(int ty, int nrx, int scheme, int variant) {
    if (variant == kVarMT && ty != 2) return nullptr;
    if (variant == kVarCEST && ty < 3) return nullptr;
    switch (ty) {
        case 2: return pick_nrx<2>(nrx, scheme, variant);
        case 3: return pick_nrx<3>(nrx, scheme, variant);
        case 4: return pick_nrx<4>(nrx, scheme, variant);
        case 5: return pick_nrx<5>(nrx, scheme, variant);
        case 6: return pick_nrx<6>(nrx, scheme, variant);
    }
    return nullptr;
}

int gm_values_per_isochromat(int ty, int nrx) {
    int n = 2;
    while (n < 2 * nrx * ty) n <<= 1;
    return n;
}

}  // namespace mrs
