// Signal to image on gfx950: Cartesian reconstruction of the scan's k-space
// and k-space noise.
//
// Cartesian recon (Src/Recon/Default/DoCartRecon.m:3-91), per echo:
//  * recon_gather_kernel: one thread per element of the image volume of one
//    coil and echo. It sums the types of VSig.Sx/Sy (:12-16), drops the last
//    readout sample (:25-27), zero-fills (:29-40), applies the permutes
//    (:41-73) and the input fftshift in one pass, writing the complex fp64
//    FFT input and the permuted k-space (VImg.Sx/Sy, :86-87);
//  * hipFFT: one batched 3-D inverse Z2Z transform over the coils, fp64 as
//    MATLAB's ifftn;
//  * recon_out_kernel: the output fftshift and 1/N scale, then Real, Imag,
//    Mag and Phase (:79-84).
// Every kernel reads and writes each element once: HBM-bound elementwise
// passes around the library FFT.
//
// k-space noise (Src/Main/DoAddNoise.m:1-21): scale * N(0,1) added to Sx and
// Sy. The normals come from Philox-4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by the seed
// and counted by the element index, and the Box-Muller transform in fp64:
// reproducible for a seed and independent of the launch shape. MATLAB's
// randn stream itself is not reproduced.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "host_common.hpp"
#include "mrilab_amd.h"

namespace mrs {

struct ReconDims {
    int64_t a0, a1, a2;  // k-space after the readout drop (and zero filling): [a0][a1][a2]
    int64_t f0, f1, f2;  // image volume after the permutes
    int64_t o0, o1, o2;  // zero-filling offsets of the acquired block
    int64_t rf, rp, sl;  // acquired (kept) readout, phase, slices
};

static mrs_status recon_layout(const mrs_recon_desc *d, ReconDims &g) {
    if (!d) return fail(MRS_ERR_ARG, "recon descriptor is NULL");
    if (d->res_freq < 2 || d->res_phase < 1 || d->slice_num < 1 || d->te_per_tr < 1 || d->rx < 1 || d->types < 1)
        return fail(MRS_ERR_ARG, "recon sizes must be positive (ResFreq >= 2)");
    g.rf = d->res_freq - 1;  // DoCartRecon.m:25-28
    g.rp = d->res_phase;
    g.sl = d->slice_num;
    const bool zf = d->zf_kx > 0 || d->zf_ky > 0 || d->zf_kz > 0;
    if (zf) {
        if (d->zf_kx <= 0 || d->zf_ky <= 0 || d->zf_kz <= 0)
            return fail(MRS_ERR_ARG, "zero filling needs ZF_Kx, ZF_Ky and the ZF_Kz factor");
        // DoCartRecon.m:31-37 (1-based there): x at ZF_Kx/2 - ResFreq/2 + 1,
        // y at ZF_Ky/2 - ResPhase/2 + 1, z at f*floor(S/2) - floor(S/2) + 1;
        // MATLAB's indexing needs integers
        const int64_t tx = (int64_t)d->zf_kx - g.rf, ty = (int64_t)d->zf_ky - g.rp;
        if (tx < 0 || ty < 0 || (tx & 1) || (ty & 1))
            return fail(MRS_ERR_ARG, "zero-filling offsets (ZF/2 - size/2) must be non-negative integers");
        g.a0 = d->zf_kx;
        g.a1 = d->zf_ky;
        g.a2 = (int64_t)d->zf_kz * g.sl;
        g.o0 = tx / 2;
        g.o1 = ty / 2;
        g.o2 = (int64_t)d->zf_kz * (g.sl / 2) - g.sl / 2;
    } else {
        g.a0 = g.rf;
        g.a1 = g.rp;
        g.a2 = g.sl;
        g.o0 = g.o1 = g.o2 = 0;
    }
    // permute [2 1 3 4] (:41-42), undone by the second permute when swap_xy
    if (d->swap_xy) {
        g.f0 = g.a0;
        g.f1 = g.a1;
    } else {
        g.f0 = g.a1;
        g.f1 = g.a0;
    }
    g.f2 = g.a2;
    return MRS_OK;
}

struct GatherParams {
    const double *sx, *sy;  // VSig [Ty][Rx][N]
    double2 *X;             // FFT input [Rx][f2][f1][f0] (column-major [f0,f1,f2,Rx])
    double *ksx, *ksy;      // VImg.Sx/Sy of this echo [Rx][f2][f1][f0]
    int64_t N, vol;         // samples per coil and type; f0*f1*f2
    ReconDims g;
    int rf_full, te, e, rx, types, swap;
};

__global__ __launch_bounds__(256) void recon_gather_kernel(GatherParams p) {
    const int64_t total = p.vol * p.rx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / p.vol;
        int64_t r = i - c * p.vol;
        const int64_t j0 = r % p.g.f0;
        r /= p.g.f0;
        const int64_t j1 = r % p.g.f1;
        const int64_t j2 = r / p.g.f1;
        // element (j0, j1, j2) of the permuted volume -> (b0, b1, b2) before the permutes
        const int64_t b0 = p.swap ? j0 : j1, b1 = p.swap ? j1 : j0, b2 = j2;
        // acquired sample (ro, pe, sl) of the zero-filled block
        const int64_t ro = b0 - p.g.o0, pe = b1 - p.g.o1, sl = b2 - p.g.o2;
        double vx = 0.0, vy = 0.0;
        if (ro >= 0 && ro < p.g.rf && pe >= 0 && pe < p.g.rp && sl >= 0 && sl < p.g.sl) {
            // reshape [ResFreq*TEPerTR, ResPhase, SliceNum, Rx, Ty] (:12-13)
            const int64_t n = (int64_t)p.e * p.rf_full + ro + (int64_t)p.rf_full * p.te * (pe + p.g.rp * sl);
            for (int t = 0; t < p.types; t++) {  // sum over types (:15-16)
                const int64_t k = ((int64_t)t * p.rx + c) * p.N + n;
                vx += p.sx[k];
                vy += p.sy[k];
            }
        }
        p.ksx[i] = vx;
        p.ksy[i] = vy;
        // fftshift: element j goes to (j + floor(f/2)) mod f
        const int64_t s0 = (j0 + p.g.f0 / 2) % p.g.f0, s1 = (j1 + p.g.f1 / 2) % p.g.f1,
                      s2 = (j2 + p.g.f2 / 2) % p.g.f2;
        p.X[c * p.vol + s0 + p.g.f0 * (s1 + p.g.f1 * s2)] = make_double2(vx, vy);
    }
}

// The same gather when the volume is the transposed k-space (no second
// permute): the source runs along the readout, the output along the phase
// direction, so a 32x32 tile goes through LDS and both sides stay
// coalesced. Block 32x8; blockIdx.x walks the (phase, readout) tiles and
// blockIdx.y the (slice, coil) planes.
__global__ __launch_bounds__(256) void recon_gather_t_kernel(GatherParams p, int tiles0, int64_t planes) {
    __shared__ double tx_[32][33], ty_[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t tile = blockIdx.x;
    const int64_t x0 = (tile % tiles0) * 32;  // j0: phase (output fastest)
    const int64_t y0 = (tile / tiles0) * 32;  // j1: readout (source fastest)
    for (int64_t pl = blockIdx.y; pl < planes; pl += gridDim.y) {
        const int64_t c = pl / p.g.f2, j2 = pl - c * p.g.f2;
        const int64_t sl = j2 - p.g.o2;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t j0 = x0 + ty + 8 * k, j1 = y0 + tx;
            const int64_t ro = j1 - p.g.o0, pe = j0 - p.g.o1;
            double vx = 0.0, vy = 0.0;
            if (j0 < p.g.f0 && j1 < p.g.f1 && ro >= 0 && ro < p.g.rf && pe >= 0 && pe < p.g.rp && sl >= 0 &&
                sl < p.g.sl) {
                const int64_t n = (int64_t)p.e * p.rf_full + ro + (int64_t)p.rf_full * p.te * (pe + p.g.rp * sl);
                for (int t = 0; t < p.types; t++) {
                    const int64_t kk = ((int64_t)t * p.rx + c) * p.N + n;
                    vx += p.sx[kk];
                    vy += p.sy[kk];
                }
            }
            tx_[ty + 8 * k][tx] = vx;
            ty_[ty + 8 * k][tx] = vy;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t j0 = x0 + tx, j1 = y0 + ty + 8 * k;
            if (j0 < p.g.f0 && j1 < p.g.f1) {
                const double vx = tx_[tx][ty + 8 * k], vy = ty_[tx][ty + 8 * k];
                const int64_t i = c * p.vol + j0 + p.g.f0 * (j1 + p.g.f1 * j2);
                p.ksx[i] = vx;
                p.ksy[i] = vy;
                const int64_t s0 = (j0 + p.g.f0 / 2) % p.g.f0, s1 = (j1 + p.g.f1 / 2) % p.g.f1,
                              s2 = (j2 + p.g.f2 / 2) % p.g.f2;
                p.X[c * p.vol + s0 + p.g.f0 * (s1 + p.g.f1 * s2)] = make_double2(vx, vy);
            }
        }
        __syncthreads();
    }
}

struct OutParams {
    const double2 *X;  // ifftn result (unscaled) [Rx][f2][f1][f0]
    double *re, *im, *mag, *ph;
    int64_t vol, f0, f1, f2;
    int rx;
    double inv_n;
};

__global__ __launch_bounds__(256) void recon_out_kernel(OutParams p) {
    const int64_t total = p.vol * p.rx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / p.vol;
        int64_t r = i - c * p.vol;
        const int64_t k0 = r % p.f0;
        r /= p.f0;
        const int64_t k1 = r % p.f1;
        const int64_t k2 = r / p.f1;
        // output fftshift: element k of the ifftn result lands at (k + floor(f/2)) mod f
        const int64_t d0 = (k0 + p.f0 / 2) % p.f0, d1 = (k1 + p.f1 / 2) % p.f1, d2 = (k2 + p.f2 / 2) % p.f2;
        const int64_t o = c * p.vol + d0 + p.f0 * (d1 + p.f1 * d2);
        const double2 v = p.X[i];
        const double x = v.x * p.inv_n, y = v.y * p.inv_n;
        p.re[o] = x;
        p.im[o] = y;
        p.mag[o] = hypot(x, y);
        p.ph[o] = atan2(y, x);
    }
}

// Philox-4x32-10 (SC'11): 10 rounds of two 32x32->64 multiplies, the key
// bumped by the Weyl constants between rounds.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

// 53-bit uniform in (0, 1) from two words.
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(256) void noise_kernel(double *sx, double *sy, int64_t n, double scale, uint32_t k0,
                                                    uint32_t k1) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u};
        philox4x32_10(c, k0, k1);
        const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
        const double rr = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        sx[i] += scale * (rr * cs);
        sy[i] += scale * (rr * sn);
    }
}

// Echo-train k-space to Cartesian order (DoEPIRecon.m:8-26,
// DoFSERecon.m:8-22): the signal of each shot is ETL echoes of R samples;
// echo e (0-based) of shot p fills phase line e*ShotNum + p, and with
// flip_even (EPI) the odd-indexed echoes, read out backwards, are
// reversed along the readout. One thread per input value; the same
// mapping for every coil and type block of N values.
__global__ __launch_bounds__(256) void echo_train_reorder_kernel(const double *__restrict__ sx,
                                                                 const double *__restrict__ sy, double *ox, double *oy,
                                                                 int64_t total, int64_t N, int R, int etl, int shots,
                                                                 int flip_even) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t blk = i / N, n = i - blk * N;
    const int64_t per = (int64_t)R * etl;  // samples of one shot
    const int64_t ps = n / per;
    const int a = (int)(n - ps * per);
    const int p = (int)(ps % shots);
    const int64_t sl = ps / shots;
    const int e = a / R, m = a - e * R;
    const int mo = (flip_even && (e & 1)) ? R - 1 - m : m;
    const int64_t no = mo + (int64_t)R * ((int64_t)e * shots + p + (int64_t)etl * shots * sl);
    ox[blk * N + no] = sx[i];
    oy[blk * N + no] = sy[i];
}

struct RDev {
    void *p = nullptr;
    ~RDev() {
        if (p) (void)hipFree(p);
    }
};

static mrs_status rmalloc(RDev &b, size_t bytes) {
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 8));
    return MRS_OK;
}

static unsigned grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 256 * 64));
}

#define FFT_TRY(expr)                                                                            \
    do {                                                                                         \
        const hipfftResult r_ = (expr);                                                          \
        if (r_ != HIPFFT_SUCCESS) return fail(MRS_ERR_DEVICE, std::string("hipFFT error ") +     \
                                                                    std::to_string((int)r_) + \
                                                                    " in " #expr);               \
    } while (0)

// The last shape's hipFFT plan and FFT buffer, kept for the next call
// (creating a plan costs tens of ms; a transform of config-1 size well
// under one). Calls hold the lock for their whole run, so concurrent
// recons serialise instead of sharing the buffer. Kept until the process
// ends or the shape changes.
struct FftCache {
    std::mutex mu;
    bool made = false;
    int device = -1, rx = 0;
    int64_t f0 = 0, f1 = 0, f2 = 0;
    hipfftHandle h = 0;
    void *X = nullptr;
};
static FftCache g_fft;

static mrs_status fft_cache_get(int device, const ReconDims &g, int rx) {
    if (g_fft.made && g_fft.device == device && g_fft.rx == rx && g_fft.f0 == g.f0 && g_fft.f1 == g.f1 &&
        g_fft.f2 == g.f2)
        return MRS_OK;
    if (g_fft.made) {
        (void)hipfftDestroy(g_fft.h);
        if (g_fft.X) (void)hipFree(g_fft.X);
        g_fft.made = false;
        g_fft.X = nullptr;
    }
    const int64_t vol = g.f0 * g.f1 * g.f2;
    HIP_TRY(hipMalloc(&g_fft.X, (size_t)(vol * rx) * sizeof(double2)));
    int n[3] = {(int)g.f2, (int)g.f1, (int)g.f0};  // row-major dims: f0 varies fastest
    const hipfftResult r = hipfftPlanMany(&g_fft.h, 3, n, nullptr, 1, (int)vol, nullptr, 1, (int)vol, HIPFFT_Z2Z, rx);
    if (r != HIPFFT_SUCCESS) {
        (void)hipFree(g_fft.X);
        g_fft.X = nullptr;
        return fail(MRS_ERR_DEVICE, "hipfftPlanMany failed with " + std::to_string((int)r));
    }
    g_fft.made = true;
    g_fft.device = device;
    g_fft.rx = rx;
    g_fft.f0 = g.f0;
    g_fft.f1 = g.f1;
    g_fft.f2 = g.f2;
    return MRS_OK;
}

// Reconstruction from device-resident k-space sx/sy ([Ty][Rx][N]) into the
// VImg arrays (device pointers when dev_out, else host).
static mrs_status recon_device(const double *sx, const double *sy, int64_t N, const mrs_recon_desc *d,
                               mrs_vimg *img, bool dev_out, hipStream_t st) {
    ReconDims g;
    mrs_status rc = recon_layout(d, g);
    if (rc) return rc;
    if (!img || !img->Real || !img->Imag || !img->Mag || !img->Phase || !img->Sx || !img->Sy)
        return fail(MRS_ERR_ARG, "VImg array pointer is NULL");
    if (N != (int64_t)d->res_freq * d->te_per_tr * d->res_phase * d->slice_num)
        return fail(MRS_ERR_ARG, "SignalNum must equal ResFreq*TEPerTR*ResPhase*SliceNum");
    const int64_t vol = g.f0 * g.f1 * g.f2, per_echo = vol * d->rx;
    if (g.f0 > INT32_MAX || g.f1 > INT32_MAX || g.f2 > INT32_MAX || per_echo > ((int64_t)1 << 40))
        return fail(MRS_ERR_UNSUPPORTED, "image volume too large");
    RDev o[6];
    double *out[6];
    double *const dst[6] = {img->Real, img->Imag, img->Mag, img->Phase, img->Sx, img->Sy};
    for (int k = 0; k < 6; k++) {
        if (dev_out) {
            out[k] = dst[k];
        } else {
            if ((rc = rmalloc(o[k], (size_t)per_echo * sizeof(double)))) return rc;
            out[k] = (double *)o[k].p;
        }
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_fft.mu);
    if ((rc = fft_cache_get(dev, g, d->rx))) return rc;
    FFT_TRY(hipfftSetStream(g_fft.h, st));
    struct { void *p; } X{g_fft.X};
    for (int e = 0; e < d->te_per_tr; e++) {
        const int64_t off = (dev_out ? (int64_t)e * per_echo : 0);
        GatherParams gp;
        gp.sx = sx;
        gp.sy = sy;
        gp.X = (double2 *)X.p;
        gp.ksx = out[4] + off;
        gp.ksy = out[5] + off;
        gp.N = N;
        gp.vol = vol;
        gp.g = g;
        gp.rf_full = d->res_freq;
        gp.te = d->te_per_tr;
        gp.e = e;
        gp.rx = d->rx;
        gp.types = d->types;
        gp.swap = d->swap_xy ? 1 : 0;
        if (d->swap_xy) {
            recon_gather_kernel<<<grid_for(per_echo), 256, 0, st>>>(gp);
        } else {
            const int64_t t0 = (g.f0 + 31) / 32, t1 = (g.f1 + 31) / 32, planes = g.f2 * d->rx;
            if (t0 * t1 > INT32_MAX) return fail(MRS_ERR_UNSUPPORTED, "image plane too large");
            const dim3 grid((unsigned)(t0 * t1), (unsigned)std::min<int64_t>(planes, 65535));
            recon_gather_t_kernel<<<grid, 256, 0, st>>>(gp, (int)t0, planes);
        }
        HIP_TRY(hipGetLastError());
        FFT_TRY(hipfftExecZ2Z(g_fft.h, (hipfftDoubleComplex *)X.p, (hipfftDoubleComplex *)X.p, HIPFFT_BACKWARD));
        OutParams op;
        op.X = (const double2 *)X.p;
        op.re = out[0] + off;
        op.im = out[1] + off;
        op.mag = out[2] + off;
        op.ph = out[3] + off;
        op.vol = vol;
        op.f0 = g.f0;
        op.f1 = g.f1;
        op.f2 = g.f2;
        op.rx = d->rx;
        op.inv_n = 1.0 / (double)vol;
        recon_out_kernel<<<grid_for(per_echo), 256, 0, st>>>(op);
        HIP_TRY(hipGetLastError());
        if (!dev_out) {
            for (int k = 0; k < 6; k++)
                HIP_TRY(hipMemcpyAsync(dst[k] + (int64_t)e * per_echo, out[k], (size_t)per_echo * sizeof(double),
                                       hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));  // the buffers are reused by the next echo
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    return MRS_OK;
}

}  // namespace mrs

using namespace mrs;

extern "C" {

mrs_status mrs_recon_dims(const mrs_recon_desc *desc, int64_t dims[5]) {
    if (!dims) return fail(MRS_ERR_ARG, "dims is NULL");
    ReconDims g;
    mrs_status rc = recon_layout(desc, g);
    if (rc) return rc;
    dims[0] = g.f0;
    dims[1] = g.f1;
    dims[2] = g.f2;
    dims[3] = desc->rx;
    dims[4] = desc->te_per_tr;
    return MRS_OK;
}

mrs_status mrs_DoEchoTrainRecon(const mrs_vsig *vsig, const mrs_recon_desc *desc, int32_t etl, int32_t shots,
                                int32_t flip_even, mrs_vimg *img, int32_t device) {
    if (!vsig || !vsig->Sx || !vsig->Sy) return fail(MRS_ERR_ARG, "VSig.Sx/Sy is NULL");
    if (!desc) return fail(MRS_ERR_ARG, "recon descriptor is NULL");
    if (etl < 1 || shots < 1) return fail(MRS_ERR_ARG, "echo train length and shot count must be positive");
    if (desc->te_per_tr != 1) return fail(MRS_ERR_ARG, "echo-train recon takes TEPerTR = 1");
    if ((int64_t)desc->res_phase != (int64_t)etl * shots)
        return fail(MRS_ERR_ARG, "res_phase must be the echo train length times the shot count");
    if (desc->res_freq < 1 || desc->slice_num < 1 || desc->rx < 1 || desc->types < 1)
        return fail(MRS_ERR_ARG, "recon sizes must be positive");
    const int64_t N = vsig->SignalNum;
    if (N != (int64_t)desc->res_freq * etl * shots * desc->slice_num)
        return fail(MRS_ERR_ARG, "SignalNum must be ResFreq * ETL * ShotNum * SliceNum");
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
    const int64_t total = N * desc->rx * desc->types;
    const size_t bytes = (size_t)total * sizeof(double);
    RDev ix, iy, ox, oy;
    mrs_status rc;
    const double *sx = vsig->Sx, *sy = vsig->Sy;
    if (!device) {
        if ((rc = rmalloc(ix, bytes)) || (rc = rmalloc(iy, bytes))) return rc;
        HIP_TRY(hipMemcpyAsync(ix.p, vsig->Sx, bytes, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(iy.p, vsig->Sy, bytes, hipMemcpyHostToDevice, st));
        sx = (const double *)ix.p;
        sy = (const double *)iy.p;
    }
    if ((rc = rmalloc(ox, bytes)) || (rc = rmalloc(oy, bytes))) return rc;
    const int64_t nblk = (total + 255) / 256;
    if (nblk > INT32_MAX) return fail(MRS_ERR_ARG, "signal too large");
    echo_train_reorder_kernel<<<dim3((unsigned)nblk), dim3(256), 0, st>>>(
        sx, sy, (double *)ox.p, (double *)oy.p, total, N, desc->res_freq, etl, shots, flip_even ? 1 : 0);
    HIP_TRY(hipGetLastError());
    rc = recon_device((const double *)ox.p, (const double *)oy.p, N, desc, img, device != 0, st);
    HIP_TRY(hipStreamSynchronize(st));  // the temporaries are freed on return
    return rc;
}

mrs_status mrs_DoCartRecon(const mrs_vsig *vsig, const mrs_recon_desc *desc, mrs_vimg *img, int32_t device) {
    if (!vsig || !vsig->Sx || !vsig->Sy) return fail(MRS_ERR_ARG, "VSig.Sx/Sy is NULL");
    if (!desc) return fail(MRS_ERR_ARG, "recon descriptor is NULL");
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
    const int64_t N = vsig->SignalNum;
    const size_t in_bytes = (size_t)N * desc->rx * desc->types * sizeof(double);
    if (device) return recon_device(vsig->Sx, vsig->Sy, N, desc, img, true, st);
    RDev ix, iy;
    mrs_status rc;
    if ((rc = rmalloc(ix, in_bytes)) || (rc = rmalloc(iy, in_bytes))) return rc;
    HIP_TRY(hipMemcpyAsync(ix.p, vsig->Sx, in_bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(iy.p, vsig->Sy, in_bytes, hipMemcpyHostToDevice, st));
    return recon_device((const double *)ix.p, (const double *)iy.p, N, desc, img, false, st);
}

mrs_status mrs_plan_recon(mrs_plan *plan, const mrs_recon_desc *desc, mrs_vimg *img, int32_t device) {
    if (!plan) return fail(MRS_ERR_ARG, "plan is NULL");
    if (!desc) return fail(MRS_ERR_ARG, "recon descriptor is NULL");
    PlanSignal ps;
    mrs_status rc = plan_signal(plan, &ps);
    if (rc) return rc;
    if (ps.types != desc->types || ps.rx != desc->rx)
        return fail(MRS_ERR_ARG, "recon TypeNum/RxCoilNum differ from the plan's");
    HIP_TRY(hipSetDevice(ps.device));
    return recon_device(ps.sx, ps.sy, ps.n, desc, img, device != 0, ps.stream);
}

mrs_status mrs_DoAddNoise(double *Sx, double *Sy, int64_t n, double scale, uint64_t seed, int32_t device) {
    if (n < 0) return fail(MRS_ERR_ARG, "n must be non-negative");
    if (n == 0) return MRS_OK;
    if (!Sx || !Sy) return fail(MRS_ERR_ARG, "Sx/Sy is NULL");
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
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (device) {
        noise_kernel<<<grid_for(n), 256, 0, st>>>(Sx, Sy, n, scale, k0, k1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
        return MRS_OK;
    }
    RDev dx, dy;
    mrs_status rc;
    if ((rc = rmalloc(dx, (size_t)n * 8)) || (rc = rmalloc(dy, (size_t)n * 8))) return rc;
    HIP_TRY(hipMemcpyAsync(dx.p, Sx, (size_t)n * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dy.p, Sy, (size_t)n * 8, hipMemcpyHostToDevice, st));
    noise_kernel<<<grid_for(n), 256, 0, st>>>((double *)dx.p, (double *)dy.p, n, scale, k0, k1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(Sx, dx.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(Sy, dy.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return MRS_OK;
}

double mrs_noise_scale(double bandwidth, double noise_level, double nex, double b0, double rfreq, double rphase,
                       double rslice, double res_freq, double res_phase, double slice_num) {
    // DoAddNoise.m:8-16 reference values
    const double bw_ref = 1e3, b0_ref = 1.5, nex_ref = 1, adc_ref = 1e4, vol_ref = 1e-9, noise_ref = 1;
    return (std::sqrt(bandwidth / bw_ref) * (noise_level / noise_ref)) /
           (std::sqrt(nex / nex_ref) * (b0 / b0_ref) * (rfreq * rphase * rslice / vol_ref) *
            std::sqrt(res_freq * res_phase * slice_num / adc_ref));
}

}  // extern "C"
