// Adjoint (backward-in-time) Bloch kernels for rf pulse design on gfx950.
//
// No reference counterpart: MRiLab only simulates. The forward model is the
// engine's rf step (BlochKernel.h:146-306 as a Rodrigues rotation about
// -Omega, then BlochKernel.h:348-368 relaxation) over a fixed-dt excitation
// with complex rf u_k (T) and gradients G_k; the loss is
// L = sum_v |sum_s Mxy_s,N - target_v|^2. Gradients w.r.t. every rf sample
// and the per-voxel T1, T2 and dB0 come from reverse mode:
//   lambda_N = dL/dM_N,  lambda' = E lambda_{k+1},  lambda_k = R^T lambda',
//   dL/dOmega_k = dt (u - a psi x u + b psi x (psi x u)), u = lambda' x M',
//   psi = -Omega dt, a = (1 - cos th)/th^2, b = (th - sin th)/th^3,
//   dL/dT2 += (lx Mx' + ly My') e2 dt/T2^2, dL/dT1 += lz (Mz' - m0) e1 dt/T1^2.
//
//  * The longitudinal state is carried as the deviation d = Mz - m0 (m0 =
//    Rho/SpinNum): relaxation is then d*e1 exactly and a rotation adds
//    m0*(R z - z), so d keeps full fp32 relative precision near equilibrium,
//    where dL/dT1 ~ lz*(Mz' - m0) would otherwise cancel.
//  * Forward kernel: one thread per voxel x SPT isochromats; state in VGPRs;
//    every K steps the state goes to a checkpoint buffer in HBM; at the end
//    each thread writes its share of the voxel's Mxy.
//  * Backward kernel: the same threads walk the checkpoint segments from the
//    last to the first. Two ways to get the states of a segment back:
//    - rebuild (REBUILD = true): start from the checkpoint at the segment's
//      end and undo each step, M' = E^-1 M_{j+1}, M_j = R^T M'. No memory
//      traffic besides the checkpoint; the inverse relaxation amplifies
//      rounding by E^-K at most, so the host picks this only where
//      K*dt/T1 and K*dt/T2 stay below ln 8 in every voxel (the forward
//      kernel flags the blocks that break this);
//    - store: recompute the forward from the checkpoint at the segment's
//      start, storing M' = R M and (sin, cos) of each step in a scratch
//      buffer (coalesced [step][component][isochromat]).
//    Then the adjoint runs back through the segment. The rf gradient of each step is summed over
//    the wave (butterfly), the block (LDS) and the grid (block partials plus
//    a fixed-order fp64 reduction), like the engine's ADC sums. Per-voxel
//    gradients accumulate in fp32 within a segment and in fp64 across them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "host_common.hpp"
#include "mrilab_amd.h"

namespace mrs {

struct ExcParams {
    const float *x, *y, *z, *mg, *pe, *rho;  // [nvox]
    const float *T1, *T2, *dB0;              // [nvox] parameters
    const float *target;                     // [nvox][2]
    const float *dWRnd, *M0x, *M0y, *M0z;    // [S][nvox]
    const float *q;                          // [n][3] gamma*Gz, gamma*Gy, gamma*Gx
    const float *rf;                         // [n][2] rf (T)
    float *ck;                               // [C+1][3][n_iso], slot c holds M_{cK} (c >= 1), slot C M_n
    float *mxy_part;                         // [ngroups][nvox][2]
    float *scratch;                          // [K][5][n_iso]
    float *rf_part;                          // [n][nblk][2], step order reversed
    double *vox_part;                        // [ngroups][nvox][3]
    int *amp;                                // [nblk] 1 if a voxel of the block cannot rebuild segments
    const float *resid;                      // [nvox][2] m_v - target_v
    int64_t nvox, n_iso;
    int S, n, K, C;
    float gyro, dt;
};

struct ExcVoxel {
    float x, y, z, wr, wi, m0s, e1, e2, t1, t2;
    float a1, a2;  // -dt/T1, -dt/T2 (clamped at -200)
};

// Rebuilding a segment backwards multiplies rounding by at most
// exp(K dt / T) <= 8.
constexpr float kLnMaxAmp = 2.0794415f;

__device__ __forceinline__ void exc_setup(const ExcParams &p, int64_t vv, ExcVoxel &k) {
    k.x = p.x[vv];
    k.y = p.y[vv];
    k.z = p.z[vv];
    float sp, cp;
    sincos_cr(p.pe[vv], &sp, &cp);
    k.wr = p.mg[vv] * cp;
    k.wi = p.mg[vv] * sp;
    k.m0s = p.rho[vv] / (float)p.S;
    k.t1 = p.T1[vv];
    k.t2 = p.T2[vv];
    // relaxation factors rounded as the engine's reference mode (the
    // longitudinal relaxation itself is applied to the deviation Mz - m0)
    float a2 = (1.0f / k.t2) * p.dt;
    a2 = -a2;
    if (a2 < -200.f) a2 = -200.f;
    float a1 = (1.0f / k.t1) * p.dt;
    a1 = -a1;
    if (a1 < -200.f) a1 = -200.f;
    k.e2 = (float)exp((double)a2);
    k.e1 = (float)exp((double)a1);
    k.a1 = a1;
    k.a2 = a2;
}

// B1 (gamma-scaled, rotated by the Tx map) and the gradient off-resonance of step j
__device__ __forceinline__ void exc_fields(const ExcParams &p, const ExcVoxel &k, int j, float &bx, float &by,
                                           float &gv) {
    const crec_t q = (crec_t)p.q + 3 * j;
    const crec_t u = (crec_t)p.rf + 2 * j;
    gv = k.z * __uint_as_float(q[0]) + k.y * __uint_as_float(q[1]) + k.x * __uint_as_float(q[2]);
    const float br = p.gyro * __uint_as_float(u[0]), bi = p.gyro * __uint_as_float(u[1]);
    bx = k.wr * br - k.wi * bi;
    by = k.wr * bi + k.wi * br;
}

// (Mx, My, d) <- R(-Omega dt) applied to M = (Mx, My, d + m0), with d the
// deviation of Mz from m0; returns sin, cos of |Omega| dt (0, 1 without
// rotation). R z - z = (-ny sn + nx nz k, nx sn + ny nz k, -k (nx^2 + ny^2)),
// k = 1 - cos.
// (Mx, My, d) <- R^T M (rotation by +|Omega| dt about n = Omega/|Omega|) in
// the deviation form of exc_rotate: R^T z - z = (ny sn + nx nz k,
// -nx sn + ny nz k, -k (nx^2 + ny^2)).
__device__ __forceinline__ void exc_unrotate(float nx, float ny, float nz, float sn, float cs, float m0, float &mx,
                                             float &my, float &md) {
    const float ax = mx, ay = my, az = md;
    const float k1 = 1.f - cs;
    const float kk = k1 * (nx * ax + ny * ay + nz * az);
    mx = ax * cs + (ny * az - nz * ay) * sn + nx * kk + m0 * (nx * nz * k1 + ny * sn);
    my = ay * cs + (nz * ax - nx * az) * sn + ny * kk + m0 * (ny * nz * k1 - nx * sn);
    md = az * cs + (nx * ay - ny * ax) * sn + nz * kk - m0 * (k1 * (nx * nx + ny * ny));
}

__device__ __forceinline__ void exc_rotate(float bx, float by, float w, float dt, float m0, float &mx, float &my,
                                           float &md, float &sn, float &cs) {
    const float om2 = (bx * bx + by * by) + w * w;
    sn = 0.f;
    cs = 1.f;
    // the hardware reciprocal square root (1 ulp) for the axis and |Omega|,
    // as the engine's rf step; the backward pass forms |Omega| the same way,
    // so its (sin, cos) are the forward's. Below |Omega| = 1e-15 rad/s the
    // rotation is skipped (rsq stays off subnormals).
    if (om2 > 1e-30f) {
        const float inv = __builtin_amdgcn_rsqf(om2);
        const float om = om2 * inv;
        const float nx = bx * inv, ny = by * inv, nz = w * inv;
        sincos_fast(om * dt, &sn, &cs);
        const float ax = mx, ay = my, az = md;
        const float k1 = 1.f - cs;
        const float kk = k1 * (nx * ax + ny * ay + nz * az);
        mx = ax * cs - (ny * az - nz * ay) * sn + nx * kk + m0 * (nx * nz * k1 - ny * sn);
        my = ay * cs - (nz * ax - nx * az) * sn + ny * kk + m0 * (ny * nz * k1 + nx * sn);
        md = az * cs - (nx * ay - ny * ax) * sn + nz * kk - m0 * (k1 * (nx * nx + ny * ny));
    }
}

template <int SPT>
__global__ __launch_bounds__(256) void exc_fwd_kernel(ExcParams p) {
    const int64_t nvox = p.nvox;
    const int grp = blockIdx.y;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = v < nvox;
    const int64_t vv = valid ? v : nvox - 1;
    ExcVoxel k;
    exc_setup(p, vv, k);
    const float w0 = p.dB0[vv] * p.gyro;
    float mx[SPT], my[SPT], mz[SPT], wb[SPT];  // mz: deviation Mz - m0
    int64_t sidx[SPT];
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        sidx[s] = vv + nvox * (int64_t)(grp * SPT + s);
        mx[s] = p.M0x[sidx[s]];
        my[s] = p.M0y[sidx[s]];
        mz[s] = p.M0z[sidx[s]] - k.m0s;
        wb[s] = w0 + p.dWRnd[sidx[s]];
    }
    int seg = 0, left = p.K;
    for (int j = 0; j < p.n; j++) {
        if (left == 0) {  // checkpoint M_j, j = seg*K
            seg++;
            left = p.K;
            if (valid) {
                float *c = p.ck + (int64_t)seg * 3 * p.n_iso;
#pragma unroll
                for (int s = 0; s < SPT; s++) {
                    c[sidx[s]] = mx[s];
                    c[p.n_iso + sidx[s]] = my[s];
                    c[2 * p.n_iso + sidx[s]] = mz[s];
                }
            }
        }
        left--;
        float bx, by, gv;
        exc_fields(p, k, j, bx, by, gv);
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            float sn, cs;
            exc_rotate(bx, by, wb[s] + gv, p.dt, k.m0s, mx[s], my[s], mz[s], sn, cs);
            mx[s] = mx[s] * k.e2;
            my[s] = my[s] * k.e2;
            mz[s] = mz[s] * k.e1;
        }
    }
    if (valid) {
        float sx = 0.f, sy = 0.f;
        float *c = p.ck + (int64_t)p.C * 3 * p.n_iso;  // M_n, where the last segment is rebuilt from
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            sx += mx[s];
            sy += my[s];
            c[sidx[s]] = mx[s];
            c[p.n_iso + sidx[s]] = my[s];
            c[2 * p.n_iso + sidx[s]] = mz[s];
        }
        p.mxy_part[(grp * nvox + v) * 2] = sx;
        p.mxy_part[(grp * nvox + v) * 2 + 1] = sy;
    }
    const float kdt = (float)p.K;
    const int bad = valid && (kdt * k.a1 < -kLnMaxAmp || kdt * k.a2 < -kLnMaxAmp);
    const int any_bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) p.amp[blockIdx.y * gridDim.x + blockIdx.x] = any_bad;
}

// Per voxel: m_v = sum over spin groups (fixed order, fp64); resid = m - target.
__global__ __launch_bounds__(256) void exc_resid_kernel(const float *__restrict__ part, int ngroups, int64_t nvox,
                                                        const float *__restrict__ target, float *resid,
                                                        double *mxy) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double sx = 0, sy = 0;
    for (int g = 0; g < ngroups; g++) {
        sx += (double)part[(g * nvox + v) * 2];
        sy += (double)part[(g * nvox + v) * 2 + 1];
    }
    mxy[2 * v] = sx;
    mxy[2 * v + 1] = sy;
    resid[2 * v] = (float)(sx - (double)target[2 * v]);
    resid[2 * v + 1] = (float)(sy - (double)target[2 * v + 1]);
}

// L = sum_v |m_v - target_v|^2 in fp64, one block, fixed order.
__global__ __launch_bounds__(256) void exc_loss_kernel(const double *__restrict__ mxy,
                                                       const float *__restrict__ target, int64_t nvox,
                                                       double *loss) {
    __shared__ double red[256];
    double acc = 0;
    for (int64_t v = threadIdx.x; v < nvox; v += 256) {
        const double rx = mxy[2 * v] - (double)target[2 * v], ry = mxy[2 * v + 1] - (double)target[2 * v + 1];
        acc += rx * rx + ry * ry;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0];
}

template <int SPT, bool REBUILD>
__global__ __launch_bounds__(256) void exc_bwd_kernel(ExcParams p) {
    extern __shared__ float red[];  // [kRedBatch][nwaves][2]
    const int64_t nvox = p.nvox;
    const int64_t n_iso = p.n_iso;
    const int grp = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = v < nvox;
    const int64_t vv = valid ? v : nvox - 1;
    ExcVoxel k;
    exc_setup(p, vv, k);
    const float w0 = p.dB0[vv] * p.gyro;
    const float dt = p.dt;
    const float gdt = p.gyro * dt;
    const float wr = valid ? k.wr * gdt : 0.f, wi = valid ? k.wi * gdt : 0.f;  // rf chain rule factors
    float wb[SPT], lx[SPT], ly[SPT], lz[SPT];
    int64_t sidx[SPT];
    const float r2x = 2.f * p.resid[2 * vv], r2y = 2.f * p.resid[2 * vv + 1];
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        sidx[s] = vv + nvox * (int64_t)(grp * SPT + s);
        wb[s] = w0 + p.dWRnd[sidx[s]];
        lx[s] = r2x;  // lambda_N = dL/dM_N
        ly[s] = r2y;
        lz[s] = 0.f;
    }
    double gT1 = 0, gT2 = 0, gB0 = 0;
    int slot = 0, t0 = 0;  // t = n - 1 - j counts processed steps
    // inverse relaxation factors of the rebuild
    const float ie2 = (float)exp(-(double)k.a2), ie1 = (float)exp(-(double)k.a1);
    for (int c = p.C - 1; c >= 0; c--) {
        const int j0 = c * p.K;
        const int j1 = min(p.n, j0 + p.K);
        float mx[SPT], my[SPT], mz[SPT];
        if constexpr (REBUILD) {
            // M_{j1} from the checkpoint at the segment's end
            const float *ck = p.ck + (int64_t)(c + 1) * 3 * n_iso;
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                mx[s] = ck[sidx[s]];
                my[s] = ck[n_iso + sidx[s]];
                mz[s] = ck[2 * n_iso + sidx[s]];
            }
        } else {
            // recompute the segment from its checkpoint; keep M' and (sin, cos)
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                if (c == 0) {
                    mx[s] = p.M0x[sidx[s]];
                    my[s] = p.M0y[sidx[s]];
                    mz[s] = p.M0z[sidx[s]] - k.m0s;
                } else {
                    const float *ck = p.ck + (int64_t)c * 3 * n_iso;
                    mx[s] = ck[sidx[s]];
                    my[s] = ck[n_iso + sidx[s]];
                    mz[s] = ck[2 * n_iso + sidx[s]];
                }
            }
            for (int j = j0; j < j1; j++) {
                float bx, by, gv;
                exc_fields(p, k, j, bx, by, gv);
                float *sc = p.scratch + (int64_t)(j - j0) * 5 * n_iso;
#pragma unroll
                for (int s = 0; s < SPT; s++) {
                    float sn, cs;
                    exc_rotate(bx, by, wb[s] + gv, dt, k.m0s, mx[s], my[s], mz[s], sn, cs);
                    if (valid) {
                        sc[sidx[s]] = mx[s];
                        sc[n_iso + sidx[s]] = my[s];
                        sc[2 * n_iso + sidx[s]] = mz[s];
                        sc[3 * n_iso + sidx[s]] = sn;
                        sc[4 * n_iso + sidx[s]] = cs;
                    }
                    mx[s] = mx[s] * k.e2;
                    my[s] = my[s] * k.e2;
                    mz[s] = mz[s] * k.e1;
                }
            }
        }
        // adjoint through the segment
        float sT1 = 0.f, sT2 = 0.f, sB0 = 0.f;
        for (int j = j1 - 1; j >= j0; j--) {
            float bx, by, gv;
            exc_fields(p, k, j, bx, by, gv);
            const float b2 = bx * bx + by * by;
            const float *sc = REBUILD ? nullptr : p.scratch + (int64_t)(j - j0) * 5 * n_iso;
            float gx = 0.f, gy = 0.f;
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                const float w = wb[s] + gv;
                const float om2 = b2 + w * w;
                // below |Omega| = 1e-15 rad/s: a zero axis and angle, for which the
                // steps below reduce to the identity exactly (no branches)
                const float inv = om2 > 1e-30f ? __builtin_amdgcn_rsqf(om2) : 0.f;
                const float nx = bx * inv, ny = by * inv, nz = w * inv, th = (om2 * inv) * dt;
                float Mx, My, Md, sn, cs;
                if constexpr (REBUILD) {
                    // M' = E^-1 M_{j+1}; (sin, cos) as the forward computed them
                    Mx = mx[s] * ie2;
                    My = my[s] * ie2;
                    Md = mz[s] * ie1;
                    sincos_fast(th, &sn, &cs);
                } else {
                    Mx = sc[sidx[s]];
                    My = sc[n_iso + sidx[s]];
                    Md = sc[2 * n_iso + sidx[s]];
                    sn = sc[3 * n_iso + sidx[s]];
                    cs = sc[4 * n_iso + sidx[s]];
                }
                const float Mz = Md + k.m0s;
                const float px = lx[s] * k.e2, py = ly[s] * k.e2, pz = lz[s] * k.e1;  // lambda'
                sT2 += lx[s] * Mx + ly[s] * My;
                sT1 += lz[s] * Md;
                // u = lambda' x M'
                const float ux = py * Mz - pz * My, uy = pz * Mx - px * Mz, uz = px * My - py * Mx;
                // a = (1 - cos)/th^2, b = (th - sin)/th^3 (series for small th)
                // (selects, both forms computed; the hardware reciprocal, 1 ulp,
                // suffices for gradient factors)
                const float t2 = th * th;
                const float i2 = __builtin_amdgcn_rcpf(t2);
                const bool small = th < 0.05f;
                const float a = small ? 0.5f - t2 * (1.0f / 24) + t2 * t2 * (1.0f / 720) : (1.f - cs) * i2;
                const float b = small ? (1.0f / 6) - t2 * (1.0f / 120) + t2 * t2 * (1.0f / 5040)
                                      : (th - sn) * i2 * __builtin_amdgcn_rcpf(th);
                // psi = -th * n; c1 = psi x u; c2 = psi x c1
                const float qx = -th * nx, qy = -th * ny, qz = -th * nz;
                const float c1x = qy * uz - qz * uy, c1y = qz * ux - qx * uz, c1z = qx * uy - qy * ux;
                const float c2x = qy * c1z - qz * c1y, c2y = qz * c1x - qx * c1z, c2z = qx * c1y - qy * c1x;
                const float gox = ux - a * c1x + b * c2x;  // dL/dOmega / dt
                const float goy = uy - a * c1y + b * c2y;
                const float goz = uz - a * c1z + b * c2z;
                sB0 += goz;
                gx += gox;
                gy += goy;
                // lambda_j = R^T lambda' (rotation by +th about n)
                const float kk = (1.f - cs) * (nx * px + ny * py + nz * pz);
                lx[s] = px * cs + (ny * pz - nz * py) * sn + nx * kk;
                ly[s] = py * cs + (nz * px - nx * pz) * sn + ny * kk;
                lz[s] = pz * cs + (nx * py - ny * px) * sn + nz * kk;
                if constexpr (REBUILD) {
                    // M_j = R^T M'
                    mx[s] = Mx;
                    my[s] = My;
                    mz[s] = Md;
                    exc_unrotate(nx, ny, nz, sn, cs, k.m0s, mx[s], my[s], mz[s]);
                }
            }
            // rf gradient of step j: (gx, gy) through B1 = gamma * w_tx * u
            // lanes past the last voxel read scratch they never wrote: select, never multiply
            float val[2] = {valid ? wr * gx + wi * gy : 0.f, valid ? wr * gy - wi * gx : 0.f};
            WaveSum<2, 32>::run(val, lane);
            if ((lane & 31) == 0) red[(slot * nwaves + wave) * 2 + lane / 32] = val[0];
            if (++slot == kRedBatch) {
                __syncthreads();
                flush_partials(red, kRedBatch, nwaves, 2, p.rf_part, t0, nblk, blk, 2);
                __syncthreads();
                t0 += kRedBatch;
                slot = 0;
            }
        }
        gT1 += (double)sT1;
        gT2 += (double)sT2;
        gB0 += (double)sB0;
    }
    if (slot > 0) {
        __syncthreads();
        flush_partials(red, slot, nwaves, 2, p.rf_part, t0, nblk, blk, 2);
    }
    if (valid) {
        double *o = p.vox_part + (grp * nvox + v) * 3;
        o[0] = gT1 * (double)k.e1 * (double)dt / ((double)k.t1 * (double)k.t1);
        o[1] = gT2 * (double)k.e2 * (double)dt / ((double)k.t2 * (double)k.t2);
        o[2] = gB0 * (double)p.gyro * (double)dt;
    }
}

// g_rf[j] = sum over blocks of the partials of step j (fixed order, fp64).
__global__ __launch_bounds__(256) void exc_rf_reduce_kernel(const float *__restrict__ part, int nblk, int n,
                                                            float *g_rf) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int t = n - 1 - j;
    double gx = 0, gy = 0;
    for (int b = 0; b < nblk; b++) {
        gx += (double)part[((int64_t)t * nblk + b) * 2];
        gy += (double)part[((int64_t)t * nblk + b) * 2 + 1];
    }
    g_rf[2 * j] = (float)gx;
    g_rf[2 * j + 1] = (float)gy;
}

__global__ __launch_bounds__(256) void exc_vox_reduce_kernel(const double *__restrict__ part, int ngroups,
                                                             int64_t nvox, float *g_T1, float *g_T2,
                                                             float *g_dB0) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double a = 0, b = 0, c = 0;
    for (int g = 0; g < ngroups; g++) {
        a += part[(g * nvox + v) * 3];
        b += part[(g * nvox + v) * 3 + 1];
        c += part[(g * nvox + v) * 3 + 2];
    }
    g_T1[v] = (float)a;
    g_T2[v] = (float)b;
    g_dB0[v] = (float)c;
}

struct ADev {
    void *p = nullptr;
    ~ADev() {
        if (p) (void)hipFree(p);
    }
};

static mrs_status amalloc(ADev &b, size_t bytes) {
    HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 8));
    return MRS_OK;
}

static mrs_status aupload(ADev &b, const void *src, size_t bytes, hipStream_t st) {
    mrs_status rc = amalloc(b, bytes);
    if (rc) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return MRS_OK;
}

}  // namespace mrs

using namespace mrs;

struct mrs_adjoint {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t nvox = 0, n_iso = 0;
    int S = 1, n = 0, K = 64, C = 1, spt = 4, block = 256, gx = 1, ngroups = 1;
    float gyro = 0.f, dt = 0.f;
    float fwd_ms = 0.f, bwd_ms = 0.f;
    bool have_fwd = false;
    int segment_mode = 0;
    bool rebuild = false;       // the last forward pass allows rebuilding segments
    bool last_rebuilt = false;  // what the last backward pass did
    std::vector<int> amp_host;
    ADev x, y, z, mg, pe, rho, T1, T2, dB0, target, dWRnd, M0[3], q, rf, ck, mxy_part, scratch, rf_part, vox_part,
        resid, mxy, loss, g_rf, g_T1, g_T2, g_dB0, amp;
    ~mrs_adjoint() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static ExcParams exc_params(mrs_adjoint *A) {
    ExcParams p;
    std::memset(&p, 0, sizeof p);
    p.x = (const float *)A->x.p;
    p.y = (const float *)A->y.p;
    p.z = (const float *)A->z.p;
    p.mg = (const float *)A->mg.p;
    p.pe = (const float *)A->pe.p;
    p.rho = (const float *)A->rho.p;
    p.T1 = (const float *)A->T1.p;
    p.T2 = (const float *)A->T2.p;
    p.dB0 = (const float *)A->dB0.p;
    p.target = (const float *)A->target.p;
    p.dWRnd = (const float *)A->dWRnd.p;
    p.M0x = (const float *)A->M0[0].p;
    p.M0y = (const float *)A->M0[1].p;
    p.M0z = (const float *)A->M0[2].p;
    p.q = (const float *)A->q.p;
    p.rf = (const float *)A->rf.p;
    p.ck = (float *)A->ck.p;
    p.mxy_part = (float *)A->mxy_part.p;
    p.scratch = (float *)A->scratch.p;
    p.rf_part = (float *)A->rf_part.p;
    p.vox_part = (double *)A->vox_part.p;
    p.amp = (int *)A->amp.p;
    p.resid = (const float *)A->resid.p;
    p.nvox = A->nvox;
    p.n_iso = A->n_iso;
    p.S = A->S;
    p.n = A->n;
    p.K = A->K;
    p.C = A->C;
    p.gyro = A->gyro;
    p.dt = A->dt;
    return p;
}

static void launch_exc_fwd(int spt, dim3 g, dim3 b, hipStream_t st, const ExcParams &p) {
    switch (spt) {
        case 1: exc_fwd_kernel<1><<<g, b, 0, st>>>(p); break;
        case 2: exc_fwd_kernel<2><<<g, b, 0, st>>>(p); break;
        case 4: exc_fwd_kernel<4><<<g, b, 0, st>>>(p); break;
        case 8: exc_fwd_kernel<8><<<g, b, 0, st>>>(p); break;
    }
}

template <bool REBUILD>
static void launch_exc_bwd_t(int spt, dim3 g, dim3 b, size_t lds, hipStream_t st, const ExcParams &p) {
    switch (spt) {
        case 1: exc_bwd_kernel<1, REBUILD><<<g, b, lds, st>>>(p); break;
        case 2: exc_bwd_kernel<2, REBUILD><<<g, b, lds, st>>>(p); break;
        case 4: exc_bwd_kernel<4, REBUILD><<<g, b, lds, st>>>(p); break;
        case 8: exc_bwd_kernel<8, REBUILD><<<g, b, lds, st>>>(p); break;
    }
}

static void launch_exc_bwd(bool rebuild, int spt, dim3 g, dim3 b, size_t lds, hipStream_t st, const ExcParams &p) {
    if (rebuild)
        launch_exc_bwd_t<true>(spt, g, b, lds, st, p);
    else
        launch_exc_bwd_t<false>(spt, g, b, lds, st, p);
}

// Copies n floats into dst (device buffer of the handle) from host or device memory.
static mrs_status put(void *dst, const float *src, size_t n, int device, hipStream_t st) {
    if (!src) return fail(MRS_ERR_ARG, "NULL parameter pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(float), device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           st));
    return MRS_OK;
}

extern "C" {

mrs_status mrs_adjoint_create(const mrs_vobj *vo, const mrs_vmag *vm, const mrs_vcoi *vc, const mrs_excitation *ex,
                              const mrs_options *opt, mrs_adjoint **out) {
    if (!out) return fail(MRS_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!vo || !vm || !vc || !ex) return fail(MRS_ERR_ARG, "NULL struct pointer");
    if (!vo->Mx || !vo->My || !vo->Mz || !vo->Rho) return fail(MRS_ERR_ARG, "VObj array pointer is NULL");
    if (!vm->dWRnd || !vm->Gxgrid || !vm->Gygrid || !vm->Gzgrid) return fail(MRS_ERR_ARG, "VMag array pointer is NULL");
    if (!vc->TxCoilmg || !vc->TxCoilpe) return fail(MRS_ERR_ARG, "VCoi Tx map pointer is NULL");
    if (!ex->grad || !ex->target) return fail(MRS_ERR_ARG, "excitation grad/target is NULL");
    if (vo->TypeNum != 1) return fail(MRS_ERR_UNSUPPORTED, "the adjoint is single-pool (TypeNum 1)");
    if (vc->TxCoilNum != 1) return fail(MRS_ERR_UNSUPPORTED, "the adjoint takes one Tx coil");
    if (ex->n_steps <= 0 || !(ex->dt > 0)) return fail(MRS_ERR_ARG, "n_steps and dt must be positive");
    if (vo->dims[0] <= 0 || vo->dims[1] <= 0 || vo->dims[2] <= 0 || vo->SpinNum <= 0)
        return fail(MRS_ERR_ARG, "VObj dims and SpinNum must be positive");
    mrs_adjoint *A = new (std::nothrow) mrs_adjoint();
    if (!A) return fail(MRS_ERR_NOMEM, "host allocation failed");
    struct Guard {
        mrs_adjoint *a;
        ~Guard() { delete a; }
    } guard{A};
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    const int dev = 0;
    if (ndev <= 0) return fail(MRS_ERR_DEVICE, "no HIP device visible");
    A->device = dev;
    HIP_TRY(hipSetDevice(dev));
    HIP_TRY(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
    for (auto &e : A->ev) HIP_TRY(hipEventCreate(&e));
    A->nvox = (int64_t)vo->dims[0] * vo->dims[1] * vo->dims[2];
    A->S = vo->SpinNum;
    A->n_iso = A->nvox * A->S;
    A->n = ex->n_steps;
    A->K = ex->checkpoint > 0 ? ex->checkpoint : 64;
    A->C = (A->n + A->K - 1) / A->K;
    if (ex->segment_mode != 0 && ex->segment_mode != 1) return fail(MRS_ERR_ARG, "segment_mode must be 0 or 1");
    A->segment_mode = ex->segment_mode;
    A->gyro = (float)vo->Gyro;
    A->dt = (float)ex->dt;
    // 2 spins per thread keeps the backward kernel at occupancy 4 (121 VGPRs)
    int spt = opt && opt->spins_per_thread > 0 ? opt->spins_per_thread : 2;
    while (spt > 1 && A->S % spt) spt >>= 1;
    if (spt != 1 && spt != 2 && spt != 4 && spt != 8) return fail(MRS_ERR_ARG, "spins_per_thread must be 1, 2, 4 or 8");
    A->spt = spt;
    A->ngroups = A->S / spt;
    if (A->ngroups > 65535) return fail(MRS_ERR_UNSUPPORTED, "SpinNum / spins_per_thread above 65535");
    A->block = opt && opt->block_size > 0 ? opt->block_size : 256;
    if (A->block % 64 || A->block > 256) return fail(MRS_ERR_ARG, "block_size must be 64, 128, 192 or 256");
    A->gx = (int)((A->nvox + A->block - 1) / A->block);
    const int nblk = A->gx * A->ngroups;
    hipStream_t st = A->stream;
    const size_t vb = (size_t)A->nvox * sizeof(float), ib = (size_t)A->n_iso * sizeof(float);
    // per-step gamma*G products, rounded like the engine's step records
    std::vector<float> q((size_t)A->n * 3);
    for (int j = 0; j < A->n; j++) {
        q[3 * j] = A->gyro * ex->grad[3 * j + 2];
        q[3 * j + 1] = A->gyro * ex->grad[3 * j + 1];
        q[3 * j + 2] = A->gyro * ex->grad[3 * j];
    }
    mrs_status rc;
    if ((rc = aupload(A->x, vm->Gxgrid, vb, st)) || (rc = aupload(A->y, vm->Gygrid, vb, st)) ||
        (rc = aupload(A->z, vm->Gzgrid, vb, st)) || (rc = aupload(A->mg, vc->TxCoilmg, vb, st)) ||
        (rc = aupload(A->pe, vc->TxCoilpe, vb, st)) || (rc = aupload(A->rho, vo->Rho, vb, st)) ||
        (rc = aupload(A->target, ex->target, 2 * vb, st)) || (rc = aupload(A->dWRnd, vm->dWRnd, ib, st)) ||
        (rc = aupload(A->M0[0], vo->Mx, ib, st)) || (rc = aupload(A->M0[1], vo->My, ib, st)) ||
        (rc = aupload(A->M0[2], vo->Mz, ib, st)) || (rc = aupload(A->q, q.data(), q.size() * 4, st)) ||
        (rc = amalloc(A->rf, (size_t)A->n * 2 * 4)) || (rc = amalloc(A->T1, vb)) || (rc = amalloc(A->T2, vb)) ||
        (rc = amalloc(A->dB0, vb)) || (rc = amalloc(A->ck, (size_t)(A->C + 1) * 3 * ib)) ||
        (rc = amalloc(A->mxy_part, (size_t)A->ngroups * 2 * vb)) || (rc = amalloc(A->amp, (size_t)nblk * 4)) ||
        (rc = amalloc(A->rf_part, (size_t)A->n * nblk * 2 * 4)) ||
        (rc = amalloc(A->vox_part, (size_t)A->ngroups * A->nvox * 3 * 8)) || (rc = amalloc(A->resid, 2 * vb)) ||
        (rc = amalloc(A->mxy, (size_t)A->nvox * 2 * 8)) || (rc = amalloc(A->loss, 8)) ||
        (rc = amalloc(A->g_rf, (size_t)A->n * 2 * 4)) || (rc = amalloc(A->g_T1, vb)) ||
        (rc = amalloc(A->g_T2, vb)) || (rc = amalloc(A->g_dB0, vb)))
        return rc;
    A->amp_host.assign(nblk, 0);
    HIP_TRY(hipStreamSynchronize(st));
    guard.a = nullptr;
    *out = A;
    return MRS_OK;
}

mrs_status mrs_adjoint_forward(mrs_adjoint *A, const float *rf, const float *T1, const float *T2, const float *dB0,
                               int32_t device, double *loss, double *mxy) {
    if (!A) return fail(MRS_ERR_ARG, "adjoint handle is NULL");
    HIP_TRY(hipSetDevice(A->device));
    hipStream_t st = A->stream;
    mrs_status rc;
    if ((rc = put(A->rf.p, rf, (size_t)A->n * 2, device, st)) || (rc = put(A->T1.p, T1, A->nvox, device, st)) ||
        (rc = put(A->T2.p, T2, A->nvox, device, st)) || (rc = put(A->dB0.p, dB0, A->nvox, device, st)))
        return rc;
    const ExcParams p = exc_params(A);
    HIP_TRY(hipEventRecord(A->ev[0], st));
    launch_exc_fwd(A->spt, dim3(A->gx, A->ngroups), dim3(A->block), st, p);
    HIP_TRY(hipGetLastError());
    const unsigned nb = (unsigned)((A->nvox + 255) / 256);
    exc_resid_kernel<<<nb, 256, 0, st>>>((const float *)A->mxy_part.p, A->ngroups, A->nvox,
                                         (const float *)A->target.p, (float *)A->resid.p, (double *)A->mxy.p);
    exc_loss_kernel<<<1, 256, 0, st>>>((const double *)A->mxy.p, (const float *)A->target.p, A->nvox,
                                       (double *)A->loss.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(A->ev[1], st));
    if (loss) HIP_TRY(hipMemcpyAsync(loss, A->loss.p, 8, hipMemcpyDeviceToHost, st));
    if (mxy) HIP_TRY(hipMemcpyAsync(mxy, A->mxy.p, (size_t)A->nvox * 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(A->amp_host.data(), A->amp.p, A->amp_host.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventElapsedTime(&A->fwd_ms, A->ev[0], A->ev[1]));
    A->rebuild = A->segment_mode == 0;
    for (int f : A->amp_host) A->rebuild = A->rebuild && f == 0;
    A->have_fwd = true;
    return MRS_OK;
}

mrs_status mrs_adjoint_backward(mrs_adjoint *A, float *g_rf, float *g_T1, float *g_T2, float *g_dB0,
                                int32_t device) {
    if (!A) return fail(MRS_ERR_ARG, "adjoint handle is NULL");
    if (!A->have_fwd) return fail(MRS_ERR_ARG, "mrs_adjoint_backward needs a forward pass first");
    if (!g_rf || !g_T1 || !g_T2 || !g_dB0) return fail(MRS_ERR_ARG, "NULL gradient pointer");
    HIP_TRY(hipSetDevice(A->device));
    hipStream_t st = A->stream;
    if (!A->rebuild && !A->scratch.p) {
        mrs_status rc = amalloc(A->scratch, (size_t)A->K * 5 * (size_t)A->n_iso * sizeof(float));
        if (rc) return rc;
    }
    const ExcParams p = exc_params(A);
    const size_t lds = (size_t)kRedBatch * (A->block / 64) * 2 * sizeof(float);
    HIP_TRY(hipEventRecord(A->ev[2], st));
    launch_exc_bwd(A->rebuild, A->spt, dim3(A->gx, A->ngroups), dim3(A->block), lds, st, p);
    HIP_TRY(hipGetLastError());
    const int nblk = A->gx * A->ngroups;
    exc_rf_reduce_kernel<<<(unsigned)((A->n + 255) / 256), 256, 0, st>>>((const float *)A->rf_part.p, nblk, A->n,
                                                                         (float *)A->g_rf.p);
    const unsigned nb = (unsigned)((A->nvox + 255) / 256);
    exc_vox_reduce_kernel<<<nb, 256, 0, st>>>((const double *)A->vox_part.p, A->ngroups, A->nvox,
                                              (float *)A->g_T1.p, (float *)A->g_T2.p, (float *)A->g_dB0.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(A->ev[3], st));
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpyAsync(g_rf, A->g_rf.p, (size_t)A->n * 2 * 4, kind, st));
    HIP_TRY(hipMemcpyAsync(g_T1, A->g_T1.p, (size_t)A->nvox * 4, kind, st));
    HIP_TRY(hipMemcpyAsync(g_T2, A->g_T2.p, (size_t)A->nvox * 4, kind, st));
    HIP_TRY(hipMemcpyAsync(g_dB0, A->g_dB0.p, (size_t)A->nvox * 4, kind, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventElapsedTime(&A->bwd_ms, A->ev[2], A->ev[3]));
    A->last_rebuilt = A->rebuild;
    return MRS_OK;
}

mrs_status mrs_adjoint_get_info(const mrs_adjoint *A, mrs_adjoint_info *info) {
    if (!A || !info) return fail(MRS_ERR_ARG, "NULL argument");
    info->n_iso = A->n_iso;
    info->n_steps = A->n;
    info->checkpoint = A->K;
    info->n_checkpoints = A->C;
    info->spins_per_thread = A->spt;
    info->block_size = A->block;
    info->segments_rebuilt = A->last_rebuilt ? 1 : 0;
    info->reserved = 0;
    info->forward_ms = A->fwd_ms;
    info->backward_ms = A->bwd_ms;
    return MRS_OK;
}

void mrs_adjoint_destroy(mrs_adjoint *A) {
    if (!A) return;
    (void)hipSetDevice(A->device);
    if (A->stream) (void)hipStreamSynchronize(A->stream);
    delete A;
}

}  // extern "C"
