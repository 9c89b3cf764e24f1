// Batched Bloch scans on gfx950: many independent single-pool scans (own
// phantom, coils and sequence each, uneven lengths) in one persistent launch.
//
//  * Work item = one wavefront tile: 64 lanes x SPT isochromats of one scan.
//    Tiles are ordered longest-scan-first and dealt round-robin to the
//    resident waves of a grid sized to the device (static LPT schedule, no
//    atomics, every wave exits after its last tile).
//  * A tile runs its scan's record stream (wave-uniform scalar loads) with
//    the same per-step code as the single-scan kernel (bloch_step), starts
//    from the initial state (M0) and writes the final state (M), so runs can
//    be replayed without resetting anything.
//  * At an ADC record the wave sums its isochromats per receive coil
//    (transposed butterfly) and writes one fp32 partial per tile; a second
//    kernel adds the tiles of each (scan, sample) in fp64 in a fixed order,
//    demodulates by rfRef and writes the scan's Sx/Sy.
//  * Both kernels are captured once into a hipGraph and replayed per run.
//  * The autotuner times every (spins per thread, block size) shape that
//    the batch admits and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "decode.hpp"
#include "device_common.hpp"
#include "host_common.hpp"
#include "mrilab_amd.h"
#include "spin_step.hpp"

namespace mrs {

constexpr double kBatchPi = 3.14159265359;  // DoScanAtCPU.cpp:63

struct BatchScan {
    uint64_t rec_off;   // first record word of the scan
    int64_t nvox;       // voxels
    int64_t iso_off;    // first isochromat in the concatenated state arrays
    int64_t vox_off;    // first voxel in the concatenated per-voxel maps
    int64_t rx_off;     // first element in the concatenated receive maps
    int64_t part_off;   // first partial vector of the scan
    int64_t adc_off;    // first sample of the scan in the concatenated sample arrays
    int64_t sig_off;    // first element of the scan's Sx/Sy
    int64_t N;          // VSig.SignalNum of the scan
    int32_t n_rec;
    int32_t S;
    int32_t Rx, rx_default;
    int32_t n_tiles;
    float cs, gyro;
    int32_t pad_;
    int64_t relax_off;  // single-voxel scan: first record's relaxation factors, else -1
};

struct BatchParams {
    const float *M0x, *M0y, *M0z;  // initial state, concatenated [scan][nvox*S]
    float *Mx, *My, *Mz;           // final state, same layout
    const float *dWRnd;            // same layout
    const float *Rho, *T1, *T2, *dB0, *gx, *gy, *gz, *txmg, *txpe;  // [scan][nvox]
    const float *rxx, *rxy;        // [scan][Rx][nvox]
    float *partial;                // [part][NV]
    const BatchScan *scans;
    const int32_t *tile_scan, *tile_idx;  // per work item, longest scan first
    const float *relax;                   // [record][2] of single-voxel scans (relax_off)
    int32_t n_tiles;
};

template <int SPT, int NRX, int PREC, bool DEFER>
__global__ __launch_bounds__(256) void bloch_batch_kernel(BatchParams p, const uint32_t *__restrict__ rec_g) {
    constexpr int NV = 2 * NRX;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwb = blockDim.x >> 6;
    const int nw = gridDim.x * nwb;
    for (int t = blockIdx.x * nwb + wave; t < p.n_tiles; t += nw) {
        const int sc = __builtin_amdgcn_readfirstlane(p.tile_scan[t]);
        const int ti = __builtin_amdgcn_readfirstlane(p.tile_idx[t]);
        const BatchScan &d = p.scans[sc];
        const int64_t nvox = d.nvox;
        const int S = d.S;
        const int64_t npair = nvox * (S / SPT);  // (voxel, spin group) pairs
        const int64_t pr = (int64_t)ti * 64 + lane;
        const bool valid = pr < npair;
        const int64_t prr = valid ? pr : npair - 1;
        const int64_t v = prr % nvox;
        const int grp = (int)(prr / nvox);
        const int64_t vg = d.vox_off + v;

        VoxelConst k;
        k.x = p.gx[vg];
        k.y = p.gy[vg];
        k.z = p.gz[vg];
        k.t1 = p.T1[vg];
        k.t2 = p.T2[vg];
        k.m0s = p.Rho[vg] / (float)S;  // BlochKernel.h:365
        {
            float sp, cp;
            sincos_cr(p.txpe[vg], &sp, &cp);
            const float mg = p.txmg[vg];
            k.wtx_r = mg * cp;
            k.wtx_i = mg * sp;
        }
        const float w0 = p.dB0[vg] * d.gyro;  // BlochKernel.h:38
        SpinState<SPT> st;
        int64_t sidx[SPT];
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            sidx[s] = d.iso_off + v + nvox * (int64_t)(grp * SPT + s);
            st.mx[s] = p.M0x[sidx[s]];
            st.my[s] = p.M0y[sidx[s]];
            st.mz[s] = p.M0z[sidx[s]];
            st.wb[s] = (w0 + p.dWRnd[sidx[s]]) + d.cs;  // BlochKernel.h:38-42
            st.cc[s] = 1.f;
            st.ss[s] = 0.f;
        }
        float crx[NRX], cry[NRX];
#pragma unroll
        for (int c = 0; c < NRX; c++) {
            if (d.rx_default) {
                crx[c] = c == 0 ? 1.f : 0.f;
                cry[c] = 0.f;
            } else if (c < d.Rx) {
                crx[c] = p.rxx[d.rx_off + v + (int64_t)c * nvox];
                cry[c] = p.rxy[d.rx_off + v + (int64_t)c * nvox];
            } else {
                crx[c] = 0.f;
                cry[c] = 0.f;
            }
        }
        if (!valid) {  // their state mirrors the last voxel's; keep it out of the sums
#pragma unroll
            for (int c = 0; c < NRX; c++) crx[c] = cry[c] = 0.f;
        }
        const TxMaps txm{p.txmg, p.txpe, vg, nvox, 1};
        const crec_t rec = (crec_t)rec_g + d.rec_off;
        const int n_rec = d.n_rec;
        // a single-voxel scan (one T1, T2 for the wave) reads its relaxation
        // factors per record instead of evaluating exp per lane
        const bool have_rlx = d.relax_off >= 0;
        const crec_t rlx = (crec_t)p.relax + 2 * (have_rlx ? d.relax_off : 0);
        const int64_t part0 = d.part_off + ti;
        const int ntile = d.n_tiles;
        int64_t adc = 0;
        // Each scan has a record stream of its own, so its records are not
        // shared through the caches with other waves: every record (8 words,
        // one scalar load) and its relaxation pair are fetched one step
        // ahead, and the step waits only for data requested a step before.
        uint32_t nx[8];
        uint32_t nr[2] = {0u, 0u};
        if (n_rec > 0) {
#pragma unroll
            for (int j = 0; j < 8; j++) nx[j] = rec[j];
            if (have_rlx) {
                nr[0] = rlx[0];
                nr[1] = rlx[1];
            }
        }
        for (int r = 0; r < n_rec; ++r) {
            uint32_t q[8];
#pragma unroll
            for (int j = 0; j < 8; j++) q[j] = nx[j];
            const RelaxIn rin{have_rlx, __uint_as_float(nr[0]), __uint_as_float(nr[1])};
            if (r + 1 < n_rec) {
                const crec_t qn = rec + (r + 1) * 8;
#pragma unroll
                for (int j = 0; j < 8; j++) nx[j] = qn[j];
                if (have_rlx) {
                    nr[0] = rlx[2 * (r + 1)];
                    nr[1] = rlx[2 * (r + 1) + 1];
                }
            }
            const uint32_t flags = q[5];
            if (flags & REC_SPOIL) {
                __asm__ volatile("");
                st.spoil();
            }
            if (flags & REC_ADC) {  // DoScanAtCPU.cpp:445-491, before the step
                if constexpr (DEFER) apply_pending(st);
                // lanes past the last voxel carry zero coil weights
                float sx = st.mx[0], sy = st.my[0];
#pragma unroll
                for (int s = 1; s < SPT; s++) {
                    sx += st.mx[s];
                    sy += st.my[s];
                }
                float val[NV];
#pragma unroll
                for (int c = 0; c < NRX; c++) {
                    val[2 * c] = sx * crx[c] + sy * cry[c];
                    val[2 * c + 1] = sy * crx[c] - sx * cry[c];
                }
                WaveSum<NV, 32>::run(val, lane);
                constexpr int G = 64 / NV;
                if ((lane & (G - 1)) == 0) p.partial[(part0 + adc * ntile) * NV + lane / G] = val[0];
                adc++;
            }
            if (!(flags & REC_NOSTEP))
                bloch_step<SPT, false, PREC, DEFER>(st, k, (const uint32_t *)q, flags, txm, rin);
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
}

// One thread per (scan, sample): adds the scan's tile partials in fp64 in
// tile order, demodulates by rfRef (DoScanAtCPU.cpp:472-479) and writes
// Sx/Sy[c*N + n] of the scan.
__global__ __launch_bounds__(256) void batch_reduce_kernel(const float *__restrict__ partial, int nv,
                                                           const BatchScan *__restrict__ scans,
                                                           const int32_t *__restrict__ samp_scan,
                                                           const double *__restrict__ adc_ref, int64_t n_total,
                                                           double *Sx, double *Sy) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const BatchScan &d = scans[samp_scan[g]];
    const int64_t a = g - d.adc_off;
    const double ref = adc_ref[g];
    const double cr = cos(-ref), sr = sin(-ref);
    const float *src = partial + (d.part_off + a * d.n_tiles) * nv;
    for (int c = 0; c < d.Rx; c++) {
        const int cc = d.rx_default ? 0 : c;
        double re = 0.0, im = 0.0;
        for (int t = 0; t < d.n_tiles; t++) {
            re += (double)src[(int64_t)t * nv + 2 * cc];
            im += (double)src[(int64_t)t * nv + 2 * cc + 1];
        }
        if (ref != 0) {
            const double r2 = re * cr - im * sr;
            const double i2 = re * sr + im * cr;
            re = r2;
            im = i2;
        }
        Sx[d.sig_off + (int64_t)c * d.N + a] = re;
        Sy[d.sig_off + (int64_t)c * d.N + a] = im;
    }
}

typedef void (*batch_fn)(dim3, dim3, hipStream_t, const BatchParams &, const uint32_t *);
typedef int (*occ_fn)(int block);

template <int SPT, int NRX, int PREC, bool DEFER>
static void launch_batch(dim3 g, dim3 b, hipStream_t st, const BatchParams &p, const uint32_t *rec) {
    bloch_batch_kernel<SPT, NRX, PREC, DEFER><<<g, b, 0, st>>>(p, rec);
}

template <int SPT, int NRX, int PREC, bool DEFER>
static int occupancy_batch(int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bloch_batch_kernel<SPT, NRX, PREC, DEFER>, block, 0) !=
        hipSuccess)
        return 1;
    return std::max(n, 1);
}

template <int SPT, int PREC, bool D>
static void pick_batch_nrx(int nrx, batch_fn *f, occ_fn *o) {
    switch (nrx) {
        case 1: *f = launch_batch<SPT, 1, PREC, D>; *o = occupancy_batch<SPT, 1, PREC, D>; return;
        case 2: *f = launch_batch<SPT, 2, PREC, D>; *o = occupancy_batch<SPT, 2, PREC, D>; return;
        case 4: *f = launch_batch<SPT, 4, PREC, D>; *o = occupancy_batch<SPT, 4, PREC, D>; return;
        case 8: *f = launch_batch<SPT, 8, PREC, D>; *o = occupancy_batch<SPT, 8, PREC, D>; return;
        case 16: *f = launch_batch<SPT, 16, PREC, D>; *o = occupancy_batch<SPT, 16, PREC, D>; return;
        case 32: *f = launch_batch<SPT, 32, PREC, D>; *o = occupancy_batch<SPT, 32, PREC, D>; return;
    }
    *f = nullptr;
    *o = nullptr;
}

template <int PREC, bool D>
static void pick_batch_spt(int spt, int nrx, batch_fn *f, occ_fn *o) {
    switch (spt) {
        case 1: pick_batch_nrx<1, PREC, D>(nrx, f, o); return;
        case 2: pick_batch_nrx<2, PREC, D>(nrx, f, o); return;
        case 4: pick_batch_nrx<4, PREC, D>(nrx, f, o); return;
        case 8: pick_batch_nrx<8, PREC, D>(nrx, f, o); return;
    }
    *f = nullptr;
    *o = nullptr;
}

// Batch kernel for spt spins per thread, nrx (padded) coils, relaxation
// form prec, and deferred or stepwise free steps.
static void pick_batch(int spt, int nrx, int prec, bool defer, batch_fn *f, occ_fn *o) {
    if (prec)
        defer ? pick_batch_spt<1, true>(spt, nrx, f, o) : pick_batch_spt<1, false>(spt, nrx, f, o);
    else
        defer ? pick_batch_spt<0, true>(spt, nrx, f, o) : pick_batch_spt<0, false>(spt, nrx, f, o);
}

struct BDev {
    void *p = nullptr;
    ~BDev() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
static mrs_status bupload(BDev &b, const std::vector<T> &v, hipStream_t st) {
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
    }
    const size_t bytes = v.size() * sizeof(T);
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 8));
    if (bytes) HIP_TRY(hipMemcpyAsync(b.p, v.data(), bytes, hipMemcpyHostToDevice, st));
    return MRS_OK;
}

}  // namespace mrs

using namespace mrs;

struct mrs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n = 0, prec = 0, timing = 0, free_steps = 0;
    bool defer = false;  // deferred free precession (DEFER kernel)
    int spt = 1, block = 64, grid = 1, nrx = 1, nv = 2, cus = 1;
    bool tuned = false;
    int n_candidates = 0;
    float tune_best = 0.f, tune_worst = 0.f;
    double decode_ms = 0, tune_ms = 0, create_ms = 0;
    float last_ms = 0.f;
    bool ran = false;
    std::vector<Decoded> dec;
    std::vector<mrs_vvar> vvar_final;
    std::vector<std::array<double, 3>> rf_final;
    std::vector<int64_t> ntot, nvox, nsig;
    std::vector<BatchScan> hscan;
    std::vector<int64_t> steps;
    int64_t total_iso = 0, total_adc = 0, total_iso_steps = 0, n_tiles = 0;
    BDev M0[3], M[3], dWRnd, Rho, T1, T2, dB0, gx, gy, gz, txmg, txpe, rxx, rxy;
    BDev rec, scans, tile_scan, tile_idx, partial, adc_ref, samp_scan, Sx, Sy, relax;
    size_t partial_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~mrs_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Lays out the tiles for spins-per-thread spt: per-scan tile counts and
// partial offsets, the longest-first work list, and the partial buffer.
static mrs_status batch_layout(mrs_batch *B, int spt) {
    int64_t part = 0;
    for (int i = 0; i < B->n; i++) {
        BatchScan &d = B->hscan[i];
        const int64_t npair = d.nvox * (d.S / spt);
        d.n_tiles = (int32_t)((npair + 63) / 64);
        d.part_off = part;
        part += (int64_t)B->dec[i].adc_ref.size() * d.n_tiles;
    }
    std::vector<int32_t> order(B->n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->steps[a] > B->steps[b]; });
    std::vector<int32_t> ts, ti;
    for (int i : order)
        for (int t = 0; t < B->hscan[i].n_tiles; t++) {
            ts.push_back(i);
            ti.push_back(t);
        }
    B->n_tiles = (int64_t)ts.size();
    if (B->n_tiles >= ((int64_t)1 << 31)) return fail(MRS_ERR_UNSUPPORTED, "more than 2^31 tiles");
    hipStream_t st = B->stream;
    mrs_status rc;
    if ((rc = bupload(B->scans, B->hscan, st))) return rc;
    if ((rc = bupload(B->tile_scan, ts, st))) return rc;
    if ((rc = bupload(B->tile_idx, ti, st))) return rc;
    const size_t need = (size_t)std::max<int64_t>(part, 1) * B->nv;
    if (need > B->partial_cap) {
        HIP_TRY(hipStreamSynchronize(st));
        if (B->partial.p) (void)hipFree(B->partial.p);
        B->partial.p = nullptr;
        HIP_TRY(hipMalloc(&B->partial.p, need * sizeof(float)));
        B->partial_cap = need;
    }
    return MRS_OK;
}

static BatchParams batch_params(mrs_batch *B) {
    BatchParams p;
    std::memset(&p, 0, sizeof p);
    p.M0x = (const float *)B->M0[0].p;
    p.M0y = (const float *)B->M0[1].p;
    p.M0z = (const float *)B->M0[2].p;
    p.Mx = (float *)B->M[0].p;
    p.My = (float *)B->M[1].p;
    p.Mz = (float *)B->M[2].p;
    p.dWRnd = (const float *)B->dWRnd.p;
    p.Rho = (const float *)B->Rho.p;
    p.T1 = (const float *)B->T1.p;
    p.T2 = (const float *)B->T2.p;
    p.dB0 = (const float *)B->dB0.p;
    p.gx = (const float *)B->gx.p;
    p.gy = (const float *)B->gy.p;
    p.gz = (const float *)B->gz.p;
    p.txmg = (const float *)B->txmg.p;
    p.txpe = (const float *)B->txpe.p;
    p.rxx = (const float *)B->rxx.p;
    p.rxy = (const float *)B->rxy.p;
    p.partial = (float *)B->partial.p;
    p.scans = (const BatchScan *)B->scans.p;
    p.tile_scan = (const int32_t *)B->tile_scan.p;
    p.tile_idx = (const int32_t *)B->tile_idx.p;
    p.relax = (const float *)B->relax.p;
    p.n_tiles = (int32_t)B->n_tiles;
    return p;
}

// Enqueues the Bloch kernel and the reduction with the current layout.
static mrs_status batch_enqueue(mrs_batch *B, int block) {
    batch_fn f;
    occ_fn o;
    pick_batch(B->spt, B->nrx, B->prec, B->defer, &f, &o);
    if (!f) return fail(MRS_ERR_UNSUPPORTED, "no batched kernel for this shape");
    const int wpb = block / 64;
    const int64_t need = (B->n_tiles + wpb - 1) / wpb;
    const int64_t resident = (int64_t)o(block) * B->cus;  // persistent grid: what fits at once
    B->grid = (int)std::max<int64_t>(1, std::min(need, resident));
    B->block = block;
    const BatchParams p = batch_params(B);
    if (B->n_tiles > 0) {
        f(dim3(B->grid), dim3(block), B->stream, p, (const uint32_t *)B->rec.p);
        HIP_TRY(hipGetLastError());
    }
    if (B->total_adc > 0) {
        const int64_t nb = (B->total_adc + 255) / 256;
        batch_reduce_kernel<<<dim3((unsigned)nb), dim3(256), 0, B->stream>>>(
            (const float *)B->partial.p, B->nv, (const BatchScan *)B->scans.p, (const int32_t *)B->samp_scan.p,
            (const double *)B->adc_ref.p, B->total_adc, (double *)B->Sx.p, (double *)B->Sy.p);
        HIP_TRY(hipGetLastError());
    }
    return MRS_OK;
}

// Device time of one batch run with this block size: an untimed run when
// the kernel has not run yet (warm), a timed screening run, then the mean of
// reps runs unless the screening run already took longer than cutoff ms (a
// shape that cannot win).
static mrs_status batch_time(mrs_batch *B, int block, int reps, float cutoff, bool warm, float *ms) {
    mrs_status rc;
    if (warm && (rc = batch_enqueue(B, block))) return rc;
    HIP_TRY(hipEventRecord(B->ev0, B->stream));
    if ((rc = batch_enqueue(B, block))) return rc;
    HIP_TRY(hipEventRecord(B->ev1, B->stream));
    HIP_TRY(hipEventSynchronize(B->ev1));
    float w = 0.f;
    HIP_TRY(hipEventElapsedTime(&w, B->ev0, B->ev1));
    if (w > cutoff) {
        *ms = w;
        return MRS_OK;
    }
    HIP_TRY(hipEventRecord(B->ev0, B->stream));
    for (int r = 0; r < reps; r++)
        if ((rc = batch_enqueue(B, block))) return rc;
    HIP_TRY(hipEventRecord(B->ev1, B->stream));
    HIP_TRY(hipEventSynchronize(B->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, B->ev0, B->ev1));
    *ms = t / reps;
    return MRS_OK;
}

static mrs_status batch_capture(mrs_batch *B) {
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
    HIP_TRY(hipStreamSynchronize(B->stream));
    HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    mrs_status rc = batch_enqueue(B, B->block);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(MRS_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    B->graph = g;
    HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    return MRS_OK;
}

extern "C" {

mrs_status mrs_batch_create(int32_t n, mrs_vobj *const *vobj, const mrs_vmag *const *vmag,
                            const mrs_vcoi *const *vcoi, const mrs_vctl *const *vctl, const mrs_vseq *const *vseq,
                            mrs_vvar *const *vvar, mrs_vsig *const *vsig, const mrs_options *opt, mrs_batch **out) {
    if (!out) return fail(MRS_ERR_ARG, "out is NULL");
    *out = nullptr;
    const auto t_create = std::chrono::steady_clock::now();
    if (n <= 0) return fail(MRS_ERR_ARG, "n_scans must be positive");
    if (!vobj || !vmag || !vcoi || !vctl || !vseq || !vvar || !vsig) return fail(MRS_ERR_ARG, "NULL array argument");
    mrs_status rc;
    {
        // argument checks (they read every time point of every sequence)
        // spread over host threads; the first failing scan is reported
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<mrs_status> st_(n, MRS_OK);
        std::vector<std::string> msg(n);
        std::vector<std::thread> pool;
        for (int w = 0; w < nth; w++)
            pool.emplace_back([&, w]() {
                for (int i = w; i < n; i += nth) {
                    st_[i] = check_args(vobj[i], vmag[i], vcoi[i], vctl[i], vseq[i], vvar[i], vsig[i]);
                    if (st_[i]) msg[i] = mrs_last_error();  // the message is thread-local
                }
            });
        for (auto &th : pool) th.join();
        for (int i = 0; i < n; i++)
            if (st_[i]) return fail(st_[i], "scan " + std::to_string(i) + ": " + msg[i]);
    }
    for (int i = 0; i < n; i++) {
        const std::string who = "scan " + std::to_string(i) + ": ";
        if (vobj[i]->TypeNum != 1) return fail(MRS_ERR_UNSUPPORTED, who + "batched scans are single-pool (TypeNum 1)");
        if (vcoi[i]->TxCoilNum != 1) return fail(MRS_ERR_UNSUPPORTED, who + "batched scans have one Tx coil");
        if (vctl[i]->RunMode != 0) return fail(MRS_ERR_UNSUPPORTED, who + "batched scans need RunMode 0");
        if (vctl[i]->GPUIndex != vctl[0]->GPUIndex) return fail(MRS_ERR_ARG, who + "all scans need the same GPUIndex");
    }
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (vctl[0]->GPUIndex < 0 || vctl[0]->GPUIndex >= ndev)
        return fail(MRS_ERR_DEVICE, "VCtl.GPUIndex " + std::to_string(vctl[0]->GPUIndex) + " is not a visible device");
    mrs_batch *B = new (std::nothrow) mrs_batch();
    if (!B) return fail(MRS_ERR_NOMEM, "host allocation failed");
    struct Guard {
        mrs_batch *b;
        ~Guard() { delete b; }
    } guard{B};
    B->n = n;
    B->device = vctl[0]->GPUIndex;
    B->prec = opt ? opt->precision_mode : 0;
    B->timing = opt ? opt->timing : 0;
    B->free_steps = opt ? opt->free_steps : 0;
    if (B->free_steps < 0 || B->free_steps > 2) return fail(MRS_ERR_ARG, "free_steps must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(B->device));
    HIP_TRY(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&B->ev0));
    HIP_TRY(hipEventCreate(&B->ev1));
    HIP_TRY(hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, B->device));

    // decode every sequence (independent, spread over host threads)
    const auto t0 = std::chrono::steady_clock::now();
    B->dec.resize(n);
    B->vvar_final.resize(n);
    B->rf_final.resize(n);
    {
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int w = 0; w < nth; w++)
            pool.emplace_back([&, w]() {
                for (int i = w; i < n; i += nth) {
                    Decoded &d = B->dec[i];
                    d.tx = 1;
                    d.coil_words = 2;
                    d.stride = rec_stride(1, 2);
                    Decoder dc(vobj[i], vcoi[i], vctl[i], vseq[i], vvar[i], false);
                    dc.run(d);
                    dc.flush_tail(d);
                    compress_runs(d);
                    B->vvar_final[i] = dc.vvar();
                    B->rf_final[i] = {dc.vvar().rfAmp[0], dc.vvar().rfPhase[0], dc.vvar().rfFreq[0]};
                }
            });
        for (auto &th : pool) th.join();
    }
    B->decode_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    {
        // one kernel runs every scan: deferral as the plan chooses it, over
        // the steps and pending rotations of all scans
        int64_t steps = 0, rotations = 0;
        for (int i = 0; i < n; i++) {
            int64_t a = 0, b = 0;
            deferral_counts(B->dec[i], &a, &b);
            steps += a;
            rotations += b;
        }
        B->defer = B->free_steps == 2 || (B->free_steps == 0 && deferral_pays(steps, rotations));
    }

    // concatenate inputs
    B->hscan.resize(n);
    B->ntot.resize(n);
    B->nvox.resize(n);
    B->nsig.resize(n);
    B->steps.resize(n);
    int maxrx = 1;
    int64_t iso = 0, vox = 0, rxo = 0, recw = 0, adc = 0, sig = 0;
    for (int i = 0; i < n; i++) {
        const mrs_vobj *o = vobj[i];
        const int64_t nv = (int64_t)o->dims[0] * o->dims[1] * o->dims[2];
        const int rxd = vcoi[i]->RxCoilDefault != 0;
        const int Rx = vcoi[i]->RxCoilNum;
        if ((int64_t)B->dec[i].adc_ref.size() > vsig[i]->SignalNum)
            return fail(MRS_ERR_ARG, "scan " + std::to_string(i) + ": VSig.SignalNum is smaller than the number of ADC samples");
        if ((int64_t)B->dec[i].n_rec() >= ((int64_t)1 << 31) / 8)
            return fail(MRS_ERR_UNSUPPORTED, "scan " + std::to_string(i) + ": too many records");
        BatchScan &d = B->hscan[i];
        std::memset(&d, 0, sizeof d);
        d.rec_off = (uint64_t)recw;
        d.nvox = nv;
        d.iso_off = iso;
        d.vox_off = vox;
        d.rx_off = rxo;
        d.adc_off = adc;
        d.sig_off = sig;
        d.N = vsig[i]->SignalNum;
        d.n_rec = (int32_t)B->dec[i].n_rec();
        d.S = o->SpinNum;
        d.Rx = Rx;
        d.rx_default = rxd;
        d.cs = (float)((double)((float)vctl[i]->CS[0] * 2.0f) * kBatchPi);  // BlochKernel.h:42
        d.gyro = (float)o->Gyro;
        B->ntot[i] = nv * o->SpinNum;
        B->nvox[i] = nv;
        B->nsig[i] = (int64_t)vsig[i]->SignalNum * Rx;
        B->steps[i] = B->dec[i].n_steps;
        B->total_iso_steps += B->ntot[i] * B->dec[i].n_steps;
        iso += B->ntot[i];
        vox += nv;
        if (!rxd) {
            rxo += nv * Rx;
            maxrx = std::max(maxrx, Rx);
        }
        recw += (int64_t)B->dec[i].rec.size();
        adc += (int64_t)B->dec[i].adc_ref.size();
        sig += B->nsig[i];
    }
    if (maxrx > 32) return fail(MRS_ERR_UNSUPPORTED, "more than 32 receive coils");
    B->total_iso = iso;
    B->total_adc = adc;
    while (B->nrx < maxrx) B->nrx <<= 1;
    B->nv = 2 * B->nrx;
    hipStream_t st = B->stream;
    {
        std::vector<float> a(iso), b(iso), c(iso), w(iso);
        std::vector<float> rho(vox), t1(vox), t2(vox), db0(vox), ggx(vox), ggy(vox), ggz(vox), mg(vox), pe(vox);
        std::vector<float> rx(std::max<int64_t>(rxo, 1)), ry(std::max<int64_t>(rxo, 1));
        for (int i = 0; i < n; i++) {
            const BatchScan &d = B->hscan[i];
            const int64_t nt = B->ntot[i], nv = B->nvox[i];
            std::memcpy(a.data() + d.iso_off, vobj[i]->Mx, nt * 4);
            std::memcpy(b.data() + d.iso_off, vobj[i]->My, nt * 4);
            std::memcpy(c.data() + d.iso_off, vobj[i]->Mz, nt * 4);
            std::memcpy(w.data() + d.iso_off, vmag[i]->dWRnd, nt * 4);
            std::memcpy(rho.data() + d.vox_off, vobj[i]->Rho, nv * 4);
            std::memcpy(t1.data() + d.vox_off, vobj[i]->T1, nv * 4);
            std::memcpy(t2.data() + d.vox_off, vobj[i]->T2, nv * 4);
            std::memcpy(db0.data() + d.vox_off, vmag[i]->dB0, nv * 4);
            std::memcpy(ggx.data() + d.vox_off, vmag[i]->Gxgrid, nv * 4);
            std::memcpy(ggy.data() + d.vox_off, vmag[i]->Gygrid, nv * 4);
            std::memcpy(ggz.data() + d.vox_off, vmag[i]->Gzgrid, nv * 4);
            std::memcpy(mg.data() + d.vox_off, vcoi[i]->TxCoilmg, nv * 4);
            std::memcpy(pe.data() + d.vox_off, vcoi[i]->TxCoilpe, nv * 4);
            if (!d.rx_default) {
                std::memcpy(rx.data() + d.rx_off, vcoi[i]->RxCoilx, nv * d.Rx * 4);
                std::memcpy(ry.data() + d.rx_off, vcoi[i]->RxCoily, nv * d.Rx * 4);
            }
        }
        if ((rc = bupload(B->M0[0], a, st)) || (rc = bupload(B->M0[1], b, st)) || (rc = bupload(B->M0[2], c, st)) ||
            (rc = bupload(B->dWRnd, w, st)) || (rc = bupload(B->Rho, rho, st)) || (rc = bupload(B->T1, t1, st)) ||
            (rc = bupload(B->T2, t2, st)) || (rc = bupload(B->dB0, db0, st)) || (rc = bupload(B->gx, ggx, st)) ||
            (rc = bupload(B->gy, ggy, st)) || (rc = bupload(B->gz, ggz, st)) || (rc = bupload(B->txmg, mg, st)) ||
            (rc = bupload(B->txpe, pe, st)) || (rc = bupload(B->rxx, rx, st)) || (rc = bupload(B->rxy, ry, st)))
            return rc;
        for (int k = 0; k < 3; k++) HIP_TRY(hipMalloc(&B->M[k].p, std::max<int64_t>(iso, 1) * sizeof(float)));
        HIP_TRY(hipStreamSynchronize(st));
    }
    {
        // the record streams, concatenated without zero-filling (every word
        // is copied from a scan's stream), scans copied in parallel
        std::unique_ptr<uint32_t[]> words(new (std::nothrow) uint32_t[(size_t)std::max<int64_t>(recw, 2)]);
        if (!words) return fail(MRS_ERR_NOMEM, "host allocation of the record streams failed");
        std::vector<double> ref((size_t)std::max<int64_t>(adc, 1));
        std::vector<int32_t> ss((size_t)std::max<int64_t>(adc, 1));
        // relaxation factors of single-voxel scans per record (the factors
        // depend only on dt and the voxel's T1, T2), as bloch_step rounds
        // them; scans in parallel, each with a small cache keyed on dt (a
        // train repeats a few step durations)
        int64_t nrlx = 0;
        for (int i = 0; i < n; i++) {
            BatchScan &d = B->hscan[i];
            d.relax_off = -1;
            if (d.nvox != 1) continue;
            d.relax_off = nrlx / 2;
            nrlx += 2 * B->dec[i].n_rec();
        }
        std::vector<float> rlx((size_t)nrlx);
        auto relax_scan = [&](int i) {
            const BatchScan &d = B->hscan[i];
            if (d.relax_off < 0) return;
            const float t1 = vobj[i]->T1[0], t2 = vobj[i]->T2[0];
            const std::vector<uint32_t> &w = B->dec[i].rec;
            float *out = rlx.data() + 2 * d.relax_off;
            constexpr int kCache = 8;
            uint32_t key[kCache];
            float v2[kCache], v1[kCache];
            int nk = 0, next = 0;
            for (int64_t r = 0; r < B->dec[i].n_rec(); r++) {
                const uint32_t kdt = w[(size_t)r * 8];
                int hit = -1;
                for (int c = 0; c < nk; c++)
                    if (key[c] == kdt) hit = c;
                if (hit < 0) {
                    float dt;
                    std::memcpy(&dt, &kdt, 4);
                    float f2, f1;
                    if (B->prec == 0) {  // BlochKernel.h:348-368 in fp32, exp rounded from fp64
                        float a2 = (1.0f / t2) * dt;
                        a2 = -a2;
                        if (a2 < -200.f) a2 = -200.f;
                        float a1 = (1.0f / t1) * dt;
                        a1 = -a1;
                        if (a1 < -200.f) a1 = -200.f;
                        f2 = (float)std::exp((double)a2);
                        f1 = (float)std::exp((double)a1);
                    } else {
                        f2 = (float)(-std::expm1(-(double)dt / (double)t2));
                        f1 = (float)(-std::expm1(-(double)dt / (double)t1));
                    }
                    hit = next;
                    key[hit] = kdt;
                    v2[hit] = f2;
                    v1[hit] = f1;
                    next = (next + 1) % kCache;
                    if (nk < kCache) nk++;
                }
                out[2 * r] = v2[hit];
                out[2 * r + 1] = v1[hit];
            }
        };
        {
            const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (int t = 0; t < nth; t++)
                pool.emplace_back([&, t] {
                    for (int i = t; i < n; i += nth) relax_scan(i);
                });
            for (auto &th : pool) th.join();
        }
        if ((rc = bupload(B->relax, rlx, st))) return rc;
        {
            const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (int t = 0; t < nth; t++)
                pool.emplace_back([&, t] {
                    for (int i = t; i < n; i += nth) {
                        const BatchScan &d = B->hscan[i];
                        std::copy(B->dec[i].rec.begin(), B->dec[i].rec.end(), words.get() + d.rec_off);
                        std::copy(B->dec[i].adc_ref.begin(), B->dec[i].adc_ref.end(), ref.begin() + d.adc_off);
                        std::fill(ss.begin() + d.adc_off, ss.begin() + d.adc_off + B->dec[i].adc_ref.size(), i);
                    }
                });
            for (auto &th : pool) th.join();
        }
        for (int i = 0; i < n; i++) std::vector<uint32_t>().swap(B->dec[i].rec);  // keep only the host-side outputs
        {
            const size_t bytes = (size_t)std::max<int64_t>(recw, 2) * sizeof(uint32_t);
            HIP_TRY(hipMalloc(&B->rec.p, bytes));
            HIP_TRY(hipMemcpyAsync(B->rec.p, words.get(), bytes, hipMemcpyHostToDevice, st));
        }
        if ((rc = bupload(B->adc_ref, ref, st)) || (rc = bupload(B->samp_scan, ss, st))) return rc;
        HIP_TRY(hipMalloc(&B->Sx.p, std::max<int64_t>(sig, 1) * sizeof(double)));
        HIP_TRY(hipMalloc(&B->Sy.p, std::max<int64_t>(sig, 1) * sizeof(double)));
        HIP_TRY(hipMemsetAsync(B->Sx.p, 0, std::max<int64_t>(sig, 1) * sizeof(double), st));
        HIP_TRY(hipMemsetAsync(B->Sy.p, 0, std::max<int64_t>(sig, 1) * sizeof(double), st));
        HIP_TRY(hipStreamSynchronize(st));
    }

    // shape: fixed by the options or tuned on the device
    int g = 0;
    for (int i = 0; i < n; i++) g = std::gcd(g, B->hscan[i].S);
    const int want_spt = opt ? opt->spins_per_thread : 0;
    const int want_block = opt ? opt->block_size : 0;
    std::vector<int> spts, blocks;
    for (int s : {1, 2, 4, 8})
        if (g % s == 0 && (want_spt <= 0 || want_spt == s)) spts.push_back(s);
    for (int b : {64, 128, 256})
        if (want_block <= 0 || want_block == b) blocks.push_back(b);
    if (spts.empty()) return fail(MRS_ERR_ARG, "spins_per_thread must be 1, 2, 4 or 8 and divide every SpinNum");
    if (blocks.empty()) return fail(MRS_ERR_ARG, "block_size must be 64, 128 or 256");
    int best_spt = spts[0], best_block = blocks[0];
    const auto t_tune = std::chrono::steady_clock::now();
    if (spts.size() * blocks.size() > 1) {
        float best = 1e30f, worst = 0.f;
        for (int s : spts) {
            B->spt = s;
            if ((rc = batch_layout(B, s))) return rc;
            // a shape with fewer active threads than one wave per SIMD
            // leaves SIMDs idle that a smaller spins-per-thread would use:
            // not timed
            if (s != spts[0] && B->total_iso / s < (int64_t)B->cus * 4 * 64) continue;
            for (int b : blocks) {
                float ms = 0.f;
                if ((rc = batch_time(B, b, 2, 1.3f * best, b == blocks[0], &ms))) return rc;
                B->n_candidates++;
                worst = std::max(worst, ms);
                if (ms < best) {
                    best = ms;
                    best_spt = s;
                    best_block = b;
                }
            }
        }
        B->tuned = true;
        B->tune_best = best;
        B->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_tune).count();
        B->tune_worst = worst;
    }
    B->spt = best_spt;
    B->block = best_block;
    if ((rc = batch_layout(B, best_spt))) return rc;
    if ((rc = batch_capture(B))) return rc;
    B->create_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_create).count();
    guard.b = nullptr;
    *out = B;
    return MRS_OK;
}

mrs_status mrs_batch_run(mrs_batch *B) {
    if (!B) return fail(MRS_ERR_ARG, "batch is NULL");
    HIP_TRY(hipSetDevice(B->device));
    if (B->timing) HIP_TRY(hipEventRecord(B->ev0, B->stream));
    HIP_TRY(hipGraphLaunch(B->exec, B->stream));
    if (B->timing) HIP_TRY(hipEventRecord(B->ev1, B->stream));
    B->ran = true;
    return MRS_OK;
}

mrs_status mrs_batch_sync(mrs_batch *B) {
    if (!B) return fail(MRS_ERR_ARG, "batch is NULL");
    HIP_TRY(hipSetDevice(B->device));
    HIP_TRY(hipStreamSynchronize(B->stream));
    if (B->timing && B->ran) HIP_TRY(hipEventElapsedTime(&B->last_ms, B->ev0, B->ev1));
    return MRS_OK;
}

mrs_status mrs_batch_fetch(mrs_batch *B, mrs_vobj *const *vobj, mrs_vvar *const *vvar, mrs_vsig *const *vsig) {
    if (!B) return fail(MRS_ERR_ARG, "batch is NULL");
    mrs_status rc = mrs_batch_sync(B);
    if (rc) return rc;
    if (vobj) {
        std::vector<float> h(std::max<int64_t>(B->total_iso, 1));
        for (int k = 0; k < 3; k++) {
            const void *src = B->ran ? B->M[k].p : B->M0[k].p;
            HIP_TRY(hipMemcpy(h.data(), src, B->total_iso * sizeof(float), hipMemcpyDeviceToHost));
            for (int i = 0; i < B->n; i++) {
                if (!vobj[i]) continue;
                float *dst = k == 0 ? vobj[i]->Mx : k == 1 ? vobj[i]->My : vobj[i]->Mz;
                std::memcpy(dst, h.data() + B->hscan[i].iso_off, B->ntot[i] * sizeof(float));
            }
        }
    }
    if (vsig) {
        int64_t total = 0;
        for (int i = 0; i < B->n; i++) total += B->nsig[i];
        std::vector<double> hx(std::max<int64_t>(total, 1)), hy(std::max<int64_t>(total, 1));
        if (total) {
            HIP_TRY(hipMemcpy(hx.data(), B->Sx.p, total * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hy.data(), B->Sy.p, total * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < B->n; i++) {
            mrs_vsig *g = vsig[i];
            if (!g) continue;
            const int64_t o = B->hscan[i].sig_off;
            if (B->ran)
                for (int64_t j = 0; j < B->nsig[i]; j++) {
                    g->Sx[j] += hx[o + j];
                    g->Sy[j] += hy[o + j];
                }
            const Decoded &d = B->dec[i];
            for (size_t s = 0; s < d.kx.size(); s++) {
                g->Kz[s] += d.kz[s];
                g->Ky[s] += d.ky[s];
                g->Kx[s] += d.kx[s];
            }
        }
    }
    if (vvar)
        for (int i = 0; i < B->n; i++) {
            if (!vvar[i]) continue;
            std::vector<double> rf[3] = {{B->rf_final[i][0]}, {B->rf_final[i][1]}, {B->rf_final[i][2]}};
            copy_vvar_out(B->vvar_final[i], rf, 1, vvar[i]);
        }
    return MRS_OK;
}

mrs_status mrs_batch_get_info(const mrs_batch *B, mrs_batch_info *info) {
    if (!B || !info) return fail(MRS_ERR_ARG, "NULL argument");
    info->n_scans = B->n;
    info->n_tiles = B->n_tiles;
    info->n_iso = B->total_iso;
    info->n_iso_steps = B->total_iso_steps;
    info->n_adc = B->total_adc;
    info->spins_per_thread = B->spt;
    info->block_size = B->block;
    info->grid_blocks = B->grid;
    info->tuned = B->tuned ? 1 : 0;
    info->n_candidates = B->n_candidates;
    info->last_run_ms = B->last_ms;
    info->tune_best_ms = B->tune_best;
    info->tune_worst_ms = B->tune_worst;
    info->deferred_free_steps = B->defer ? 1 : 0;
    info->decode_ms = B->decode_ms;
    info->tune_ms = B->tune_ms;
    info->create_ms = B->create_ms;
    return MRS_OK;
}

void mrs_batch_destroy(mrs_batch *B) {
    if (!B) return;
    (void)hipSetDevice(B->device);
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    delete B;
}

mrs_status mrs_DoScanBatchAtGPU(int32_t n, mrs_vobj *const *vobj, const mrs_vmag *const *vmag,
                                const mrs_vcoi *const *vcoi, const mrs_vctl *const *vctl,
                                const mrs_vseq *const *vseq, mrs_vvar *const *vvar, mrs_vsig *const *vsig,
                                const mrs_options *opt) {
    mrs_batch *B = nullptr;
    mrs_status rc = mrs_batch_create(n, vobj, vmag, vcoi, vctl, vseq, vvar, vsig, opt, &B);
    if (rc) return rc;
    rc = mrs_batch_run(B);
    if (!rc) rc = mrs_batch_fetch(B, vobj, vvar, vsig);
    mrs_batch_destroy(B);
    return rc;
}

}  // extern "C"
