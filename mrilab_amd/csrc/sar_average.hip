// N-gram local SAR averaging on gfx950
// (Lib/src/utilities/DoSARAverageAtCPU/DoSARAverageAtCPU.cpp).
//
// The reference grows, per voxel with MassDen != 0, a voxel-index "sphere"
// S_r (r = 1, 2, ...; rado = floor(sqrt(r^2 - o^2)) columns per slice
// offset o, radm = floor(sqrt(rado^2 - m^2)) rows per column offset m)
// until the mass of S_r exceeds grams*1e-3 kg, summing every S_r afresh,
// then interpolates the SAR and Power sums of S_(r-1) and S_r to exactly
// that mass (:69-127). Only the mass decides the radius, so the work splits:
//  - sar_radius_kernel, one thread per voxel: the radius search on MassDen
//    alone, each S_r summed in the reference's order; it leaves the final
//    radius R (0 = voxel left at zero), the interpolation weight and the
//    voxel-count denominator.
//  - sar_average_kernel, one thread per voxel and time point: one walk over
//    S_R in (o, m, n) order, summing SAR and Power into S_R accumulators
//    and, for the elements inside S_(R-1), into S_(R-1) ones. S_(R-1) is a
//    subset of S_R and both walks are lexicographic in (o, m, n), so each
//    accumulator receives the reference's sequence of additions: results
//    equal the reference's bit for bit (fp64, contraction off), from 2|S_R|
//    reads per voxel and time point where the reference reads the sum over
//    r of 2|S_r|.
// Both kernels hand each XCD a contiguous slab of the volume (workgroups
// are dispatched to the 8 XCDs round robin), so the overlapping spheres of
// neighbouring voxels hit the same L2.
//
// The reference loops forever when the volume's whole mass is at most
// grams*1e-3 (S_r stops growing once it covers the volume). The mass of that
// covering sphere, summed in its (o, m, n) order, is the sum over the volume
// in memory order for every voxel; the host entry point checks it before the
// search and returns MRS_ERR_ARG instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "host_common.hpp"
#include "mrilab_amd.h"

namespace mrs {

namespace {

constexpr int kAvgBlock = 256;
constexpr int kXcds = 8;

// floor(sqrt(v)) for 0 <= v < 2^31, exact.
__device__ __forceinline__ int isqrt_i(int v) {
    int k = (int)sqrtf((float)v);
    while (k * k > v) k--;
    while ((k + 1) * (k + 1) <= v) k++;
    return k;
}

// Tile of this workgroup: the grid (a multiple of 8) is cut into 8 runs of
// consecutive tiles, run x going to the workgroups that land on XCD x.
__device__ __forceinline__ int64_t xcd_tile() {
    const int64_t per = gridDim.x / kXcds;
    return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

struct RadiusParams {
    const double *mass, *sar;  // MassDen; SAR's first time point
    int *R;                    // out: final radius, 0 = voxel stays 0
    double *coefw, *den;       // out: DoSARAverageAtCPU.cpp:113-114 per voxel
    int *status;               // out: set to 1 if a search hits rcap
    int64_t nvox;
    int d0, d1, d2, rcap;
    double vol, gkg;  // (dx*dy)*dz, grams*1e-3
};

// Rows are walked in chunks of kChunk elements whose loads are all issued
// before the in-order additions, so a wave keeps several loads in flight.
constexpr int kChunk = 2;

__global__ __launch_bounds__(kAvgBlock) void sar_radius_kernel(RadiusParams p) {
#pragma clang fp contract(off)
    const int64_t v = xcd_tile() * kAvgBlock + threadIdx.x;
    if (v >= p.nvox) return;
    const int j = (int)(v % p.d0), i = (int)((v / p.d0) % p.d1), s = (int)(v / ((int64_t)p.d0 * p.d1));
    if (p.mass[v] == 0) {  // :74
        p.R[v] = 0;
        return;
    }
    const bool hot = p.sar[v] > 0;  // :99, the centre's first time point
    const int64_t plane = (int64_t)p.d0 * p.d1;
    double totwei = 0, totwei2 = 0;
    int totpx = 0, totpx2 = 0, r = 0;
    while (totwei <= p.gkg) {  // :81-111
        if (r >= p.rcap) {     // unreachable once the host's mass check passed
            p.status[0] = 1;
            p.R[v] = 0;
            return;
        }
        totwei2 = totwei;
        totpx2 = totpx;
        r++;
        double w = 0;
        int cnt = 0;
        const int o0 = max(-r, -s), o1 = min(r, p.d2 - 1 - s);
        for (int o = o0; o <= o1; o++) {
            const int rado = isqrt_i(r * r - o * o);
            const int m0 = max(-rado, -i), m1 = min(rado, p.d1 - 1 - i);
            for (int m = m0; m <= m1; m++) {
                const int radm = isqrt_i(rado * rado - m * m);
                const int n0 = max(-radm, -j), n1 = min(radm, p.d0 - 1 - j);
                const double *row = p.mass + (j + (int64_t)(i + m) * p.d0 + (int64_t)(s + o) * plane);
                int n = n0;
                for (; n + kChunk - 1 <= n1; n += kChunk) {
                    double x[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; u++) x[u] = row[n + u];
#pragma unroll
                    for (int u = 0; u < kChunk; u++) w = w + p.vol * x[u];
                }
                for (; n <= n1; n++) w = w + p.vol * row[n];
                cnt += max(0, n1 - n0 + 1);
            }
        }
        totwei = w;
        totpx = hot ? cnt : 0;
    }
    const double coefw = (p.gkg - totwei2) / (totwei - totwei2);  // :113
    const double totpxs = coefw * (double)(totpx - totpx2);        // :114
    p.R[v] = totpx > 0 ? r : 0;                                    // :119
    p.coefw[v] = coefw;
    p.den[v] = (double)totpx2 + totpxs;
}

struct AverageParams {
    const double *sar, *pow;  // [d0, d1, d2, T]
    double *ave_sar, *ave_pow;
    const int *R;
    const double *coefw, *den;
    int64_t nvox;
    int d0, d1, d2;
    int t0;  // first window of this launch; block y covers windows t0 + y*TT .. + TT-1
};

// Adds elements [nb, ne] of TT rows (window stride nvox) to the S_R sums
// and, with BOTH, to the S_(R-1) sums, each in increasing n.
template <int TT, bool BOTH>
__device__ __forceinline__ void add_row(const double *rs, const double *rp, int64_t tstride, int nb, int ne,
                                        double *s2, double *q2, double *s1, double *q1) {
#pragma clang fp contract(off)
    int n = nb;
    for (; n + kChunk - 1 <= ne; n += kChunk) {
        double a[TT][kChunk], b[TT][kChunk];
#pragma unroll
        for (int t = 0; t < TT; t++)
#pragma unroll
            for (int u = 0; u < kChunk; u++) {
                a[t][u] = rs[t * tstride + n + u];
                b[t][u] = rp[t * tstride + n + u];
            }
#pragma unroll
        for (int t = 0; t < TT; t++)
#pragma unroll
            for (int u = 0; u < kChunk; u++) {
                s2[t] = s2[t] + a[t][u];
                q2[t] = q2[t] + b[t][u];
                if (BOTH) {
                    s1[t] = s1[t] + a[t][u];
                    q1[t] = q1[t] + b[t][u];
                }
            }
    }
    for (; n <= ne; n++) {
#pragma unroll
        for (int t = 0; t < TT; t++) {
            const double x = rs[t * tstride + n], y = rp[t * tstride + n];
            s2[t] = s2[t] + x;
            q2[t] = q2[t] + y;
            if (BOTH) {
                s1[t] = s1[t] + x;
                q1[t] = q1[t] + y;
            }
        }
    }
}

template <int TT>
__global__ __launch_bounds__(kAvgBlock) void sar_average_kernel(AverageParams p) {
#pragma clang fp contract(off)
    const int64_t v = xcd_tile() * kAvgBlock + threadIdx.x;
    if (v >= p.nvox) return;
    const int r = p.R[v];
    const int j = (int)(v % p.d0), i = (int)((v / p.d0) % p.d1), s = (int)(v / ((int64_t)p.d0 * p.d1));
    const int64_t plane = (int64_t)p.d0 * p.d1;
    const int64_t off = (int64_t)(p.t0 + blockIdx.y * TT) * p.nvox;
    if (r == 0) {
#pragma unroll
        for (int t = 0; t < TT; t++) {
            p.ave_sar[off + t * p.nvox + v] = 0.0;
            p.ave_pow[off + t * p.nvox + v] = 0.0;
        }
        return;
    }
    const double *sar = p.sar + off, *pw = p.pow + off;
    // S_R (2) and S_(R-1) (1) per window; S_0 is empty: the first search
    // step interpolates from zero sums (:75-78)
    double s2[TT], s1[TT], q2[TT], q1[TT];
#pragma unroll
    for (int t = 0; t < TT; t++) s2[t] = s1[t] = q2[t] = q1[t] = 0;
    const int r1 = r - 1;
    const int o0 = max(-r, -s), o1 = min(r, p.d2 - 1 - s);
    for (int o = o0; o <= o1; o++) {
        const int rado = isqrt_i(r * r - o * o);
        const int rado1 = (r1 > 0 && abs(o) <= r1) ? isqrt_i(r1 * r1 - o * o) : -1;
        const int m0 = max(-rado, -i), m1 = min(rado, p.d1 - 1 - i);
        for (int m = m0; m <= m1; m++) {
            const int radm = isqrt_i(rado * rado - m * m);
            const int radm1 = abs(m) <= rado1 ? isqrt_i(rado1 * rado1 - m * m) : -1;
            const int n0 = max(-radm, -j), n1 = min(radm, p.d0 - 1 - j);
            const int64_t base = j + (int64_t)(i + m) * p.d0 + (int64_t)(s + o) * plane;
            // [n0, n1] in order, split at S_(R-1)'s row [lo, hi] (empty
            // when radm1 < 0): the three pieces partition the row
            const int lo = radm1 >= 0 ? -radm1 : 1, hi = radm1 >= 0 ? radm1 : 0;
            add_row<TT, false>(sar + base, pw + base, p.nvox, n0, min(n1, lo - 1), s2, q2, s1, q1);
            add_row<TT, true>(sar + base, pw + base, p.nvox, max(n0, lo), min(n1, hi), s2, q2, s1, q1);
            add_row<TT, false>(sar + base, pw + base, p.nvox, max(n0, hi + 1), n1, s2, q2, s1, q1);
        }
    }
    const double cw = p.coefw[v], den = p.den[v];
#pragma unroll
    for (int t = 0; t < TT; t++) {
        const double totSARs = cw * (s2[t] - s1[t]), totPowers = cw * (q2[t] - q1[t]);  // :117-118
        p.ave_sar[off + t * p.nvox + v] = (s1[t] + totSARs) / den;                      // :120-124
        p.ave_pow[off + t * p.nvox + v] = (q1[t] + totPowers) / den;
    }
}

// Per workgroup, the sum and absolute sum of vol*MassDen over a grid-stride
// share of the volume (any order; the host adds the kMassBlocks pairs).
constexpr int kMassBlocks = 128;

__global__ __launch_bounds__(1024) void mass_total_kernel(const double *mass, int64_t n, double vol, double *out) {
#pragma clang fp contract(off)
    __shared__ double ls[1024], la[1024];
    double a = 0, b = 0;
    for (int64_t k = (int64_t)blockIdx.x * 1024 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 1024) {
        const double w = vol * mass[k];
        a += w;
        b += fabs(w);
    }
    ls[threadIdx.x] = a;
    la[threadIdx.x] = b;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            ls[threadIdx.x] += ls[threadIdx.x + h];
            la[threadIdx.x] += la[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = ls[0];
        out[2 * blockIdx.x + 1] = la[0];
    }
}

// The same sum in memory order, the order of the covering sphere's walk.
__global__ __launch_bounds__(64) void mass_total_ordered_kernel(const double *mass, int64_t n, double vol,
                                                                double *out) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    double w = 0;
    for (int64_t k = 0; k < n; k++) w = w + vol * mass[k];
    out[0] = w;
}

struct ADev {
    void *p = nullptr;
    ~ADev() {
        if (p) (void)hipFree(p);
    }
};

mrs_status amalloc(ADev &b, size_t bytes) {
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 8));
    return MRS_OK;
}

// Sets the search-status word to 0 (a kernel rather than a memset, so that
// code running this on a captured stream records no memset node).
__global__ void clear_status_kernel(int *status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *status = 0;
}

// Device buffers of one average beyond the inputs and outputs: per voxel the
// radius, weight and denominator; the status word and the mass partials.
struct AvgScratch {
    int *R;
    double *coefw, *den;
    int *status;
    double *tot;  // 2 * kMassBlocks + 2 doubles
};

mrs_status check_avg_args(const double *SAR, const double *Power, const double *MassDen, const int64_t dims[4],
                          double dx, double dy, double dz, double grams, double *aveSAR, double *avePower) {
    if (!SAR || !Power || !MassDen || !aveSAR || !avePower) return fail(MRS_ERR_ARG, "NULL array pointer");
    if (!dims) return fail(MRS_ERR_ARG, "dims is NULL");
    const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2], T = dims[3];
    if (d0 <= 0 || d1 <= 0 || d2 <= 0 || T <= 0) return fail(MRS_ERR_ARG, "dims must be positive");
    if (d0 + d1 + d2 > 40000) return fail(MRS_ERR_UNSUPPORTED, "rows + columns + slices must be at most 40000");
    if (d0 * d1 * d2 > ((int64_t)1 << 40) / T) return fail(MRS_ERR_UNSUPPORTED, "volume too large");
    if (!std::isfinite(dx) || !std::isfinite(dy) || !std::isfinite(dz) || !std::isfinite(grams))
        return fail(MRS_ERR_ARG, "dx, dy, dz and grams must be finite");
    return MRS_OK;
}

// The average on device arrays, enqueued on st (synchronized before return).
mrs_status sar_average_device(const double *sar, const double *pw, const double *mass, const int64_t dims[4],
                              double dx, double dy, double dz, double grams, double *out_s, double *out_p,
                              const AvgScratch &sc, hipStream_t st) {
    const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2], T = dims[3];
    const int64_t nvox = d0 * d1 * d2;
    const double vol = dx * dy * dz, gkg = grams * 1e-3;

    // the search ends for every voxel iff the covering sphere's mass,
    // summed in memory order, exceeds gkg; decide from an any-order sum and
    // its rounding bound, and sum in order only when the bound straddles gkg
    double part[2 * kMassBlocks], tot[2] = {0, 0};
    double *dtot = sc.tot + 2;
    mass_total_kernel<<<kMassBlocks, 1024, 0, st>>>(mass, nvox, vol, dtot);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(part, dtot, sizeof(part), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < kMassBlocks; b++) {
        tot[0] += part[2 * b];
        tot[1] += part[2 * b + 1];
    }
    const double bound = 2.0 * (double)nvox * std::ldexp(1.0, -53) * tot[1];
    bool ends = tot[0] - bound > gkg;
    if (!ends && !(tot[0] + bound <= gkg)) {
        mass_total_ordered_kernel<<<1, 64, 0, st>>>(mass, nvox, vol, dtot);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(tot, dtot, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        ends = !(tot[0] <= gkg);  // a NaN mass ends the search at r = 1 as in the reference
    }
    if (!ends)
        return fail(MRS_ERR_ARG, "the volume's mass does not exceed grams*1e-3 kg: the N-gram sphere never closes");

    const int64_t tiles = (nvox + kAvgBlock - 1) / kAvgBlock;
    const int64_t grid = (tiles + kXcds - 1) / kXcds * kXcds;
    if (grid > INT32_MAX) return fail(MRS_ERR_UNSUPPORTED, "volume too large");
    clear_status_kernel<<<1, 64, 0, st>>>(sc.status);
    HIP_TRY(hipGetLastError());
    RadiusParams rp;
    rp.mass = mass;
    rp.sar = sar;
    rp.R = sc.R;
    rp.coefw = sc.coefw;
    rp.den = sc.den;
    rp.status = sc.status;
    rp.nvox = nvox;
    rp.d0 = (int)d0;
    rp.d1 = (int)d1;
    rp.d2 = (int)d2;
    rp.rcap = (int)(d0 + d1 + d2);
    rp.vol = vol;
    rp.gkg = gkg;
    sar_radius_kernel<<<(unsigned)grid, kAvgBlock, 0, st>>>(rp);
    HIP_TRY(hipGetLastError());
    AverageParams ap;
    ap.sar = sar;
    ap.pow = pw;
    ap.ave_sar = out_s;
    ap.ave_pow = out_p;
    ap.R = rp.R;
    ap.coefw = rp.coefw;
    ap.den = rp.den;
    ap.nvox = nvox;
    ap.d0 = rp.d0;
    ap.d1 = rp.d1;
    ap.d2 = rp.d2;
    // windows in pairs, then one, per sphere walk (grid.y <= 65535); chunks
    // of 2 elements and pairs of windows measured fastest on the 10-gram
    // average of 128x128x64 voxels (4, 8 and window groups of 4 were 5-30 %
    // slower)
    for (int64_t t = 0; t < T;) {
        const int tt = T - t >= 2 ? 2 : 1;
        const int64_t groups = std::min<int64_t>((T - t) / tt, 65535);
        ap.t0 = (int)t;
        const dim3 g2((unsigned)grid, (unsigned)groups);
        if (tt == 2)
            sar_average_kernel<2><<<g2, kAvgBlock, 0, st>>>(ap);
        else
            sar_average_kernel<1><<<g2, kAvgBlock, 0, st>>>(ap);
        HIP_TRY(hipGetLastError());
        t += groups * tt;
    }
    int status = 0;
    HIP_TRY(hipMemcpyAsync(&status, sc.status, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (status) return fail(MRS_ERR_DEVICE, "N-gram radius search passed the covering sphere");
    return MRS_OK;
}

// Scratch of mrs_DoSARAverageAtGPU_noalloc, for volumes of at most
// kNoAllocMaxVoxels voxels: the status word and mass partials (4 KiB), then
// per voxel the radius (int) and the weight and denominator (double).
constexpr int64_t kNoAllocMaxVoxels = (int64_t)1 << 20;
constexpr size_t kScratchHead = 4096;
constexpr size_t kScratchBytes = kScratchHead + (size_t)kNoAllocMaxVoxels * (sizeof(int) + 2 * sizeof(double));
static_assert((2 * kMassBlocks + 2) * sizeof(double) <= kScratchHead, "mass partials exceed the scratch head");

// One copy on every GPU the library runs on, held until the process ends.
// In host code the name stands for the symbol, not a device address: the
// address of this GPU's copy comes from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

// This device's scratch address (looked up once per device, with it current).
mrs_status scratch_address(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return fail(MRS_ERR_DEVICE, "device index out of range");
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(MRS_ERR_DEVICE, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch_ptr[device] = ptr;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return MRS_OK;
}

}  // namespace

}  // namespace mrs

using namespace mrs;

extern "C" {

mrs_status mrs_DoSARAverageAtGPU(const double *SAR, const double *Power, const double *MassDen, const int64_t dims[4],
                                 double dx, double dy, double dz, double grams, double *aveSAR, double *avePower,
                                 int32_t device) {
    mrs_status rc = check_avg_args(SAR, Power, MassDen, dims, dx, dy, dz, grams, aveSAR, avePower);
    if (rc) return rc;
    const int64_t nvox = dims[0] * dims[1] * dims[2], T = dims[3];
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
    const size_t vb = (size_t)nvox * sizeof(double), tb = vb * (size_t)T;
    ADev dsar, dpow, dmass, dout_s, dout_p, dR, dcw, dden, dstat;
    const double *sar = SAR, *pw = Power, *mass = MassDen;
    double *out_s = aveSAR, *out_p = avePower;
    if (!device) {
        if ((rc = amalloc(dsar, tb)) || (rc = amalloc(dpow, tb)) || (rc = amalloc(dmass, vb)) ||
            (rc = amalloc(dout_s, tb)) || (rc = amalloc(dout_p, tb)))
            return rc;
        HIP_TRY(hipMemcpyAsync(dsar.p, SAR, tb, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dpow.p, Power, tb, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dmass.p, MassDen, vb, hipMemcpyHostToDevice, st));
        sar = (const double *)dsar.p;
        pw = (const double *)dpow.p;
        mass = (const double *)dmass.p;
        out_s = (double *)dout_s.p;
        out_p = (double *)dout_p.p;
    }
    if ((rc = amalloc(dR, (size_t)nvox * sizeof(int))) || (rc = amalloc(dcw, vb)) || (rc = amalloc(dden, vb)) ||
        (rc = amalloc(dstat, (2 * kMassBlocks + 2) * sizeof(double))))
        return rc;
    AvgScratch sc{(int *)dR.p, (double *)dcw.p, (double *)dden.p, (int *)dstat.p, (double *)dstat.p};
    if ((rc = sar_average_device(sar, pw, mass, dims, dx, dy, dz, grams, out_s, out_p, sc, st))) return rc;
    if (!device) {
        HIP_TRY(hipMemcpyAsync(aveSAR, out_s, tb, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(avePower, out_p, tb, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return MRS_OK;
}

mrs_status mrs_DoSARAverageAtGPU_noalloc(const double *SAR, const double *Power, const double *MassDen,
                                         const int64_t dims[4], double dx, double dy, double dz, double grams,
                                         double *aveSAR, double *avePower, void *stream) {
    mrs_status rc = check_avg_args(SAR, Power, MassDen, dims, dx, dy, dz, grams, aveSAR, avePower);
    if (rc) return rc;
    const int64_t nvox = dims[0] * dims[1] * dims[2];
    if (nvox > kNoAllocMaxVoxels)
        return fail(MRS_ERR_UNSUPPORTED, "the allocation-free SAR average takes at most " +
                                             std::to_string(kNoAllocMaxVoxels) + " voxels, got " +
                                             std::to_string(nvox));
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    unsigned char *base = nullptr;
    if ((rc = scratch_address(device, &base))) return rc;
    AvgScratch sc;
    sc.status = (int *)base;
    sc.tot = (double *)base;  // the status word shares the first 16 bytes; partials from tot + 2
    sc.R = (int *)(base + kScratchHead);
    sc.coefw = (double *)(base + kScratchHead + (size_t)kNoAllocMaxVoxels * sizeof(int));
    sc.den = sc.coefw + kNoAllocMaxVoxels;
    return sar_average_device(SAR, Power, MassDen, dims, dx, dy, dz, grams, aveSAR, avePower, sc,
                              (hipStream_t)stream);
}

}  // extern "C"
