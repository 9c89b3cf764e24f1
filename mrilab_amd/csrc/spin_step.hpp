// Per-thread spin state and the Bloch step of one decoded record, shared by
// the single-scan forward kernel and the batched kernel.
//
// A thread owns SPT isochromats of one voxel and type. bloch_step() applies
// one step record (BlochKernel.h:7-397): rf rotation or free precession,
// then T2/T1 relaxation, with the cached free-precession rotation reused for
// REC_SAME steps and repeated for the folded run length of the record, or
// (DEFER) with free precession gathered as a pending rotation angle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "decode.hpp"
#include "device_common.hpp"

namespace mrs {

// Per-voxel inputs of a thread that stay constant over a launch.
struct VoxelConst {
    float x, y, z;        // voxel position (m)
    float t1, t2;         // relaxation times (s)
    float m0s;            // Rho / SpinNum (BlochKernel.h:365)
    float wtx_r, wtx_i;   // single-Tx B1+ weight mg * exp(i * pe)
};

// Multi-Tx rf inputs (GEN kernels): per-coil B1+ maps of this voxel.
struct TxMaps {
    const float *mg, *pe;
    int64_t v, nvox;
    int tx;
};

template <int SPT>
struct SpinState {
    float mx[SPT], my[SPT], mz[SPT];
    float wb[SPT];            // off-resonance without gradients (rad/s)
    float cc[SPT], ss[SPT];   // cached free-precession rotation (times e2 in reference mode)
    // rotation by (wb[s] - wb[0]) * dt for s >= 1 (times e2 in reference mode),
    // refreshed with dt: single-use rotations of spin s are composed from
    // spin 0's by angle addition
    float cd[SPT], sd[SPT];
    float e1 = 1.f, e2 = 1.f, m0e = 0.f;  // reference relaxation factors
    float d1 = 0.f, d2 = 0.f;             // accurate relaxation factors
    bool have_rot = false, have_e = false;
    // Deferred free precession (DEFER kernels): the summed rotation angle of
    // the free steps not yet applied to Mx, My, per spin, and the product of
    // their transverse relaxation factors; apply_pending() applies both.
    double ph[SPT] = {};
    float gp = 1.f;
    bool pend = false;

    __device__ __forceinline__ void spoil() {  // Plugin_IdealSpoiler.m:10-11
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            mx[s] = 0.f;
            my[s] = 0.f;
            ph[s] = 0.0;
        }
        gp = 1.f;
        pend = false;
    }
};

// The gradient terms of record q for this voxel.
template <typename Q>
__device__ __forceinline__ GradTerms grad_terms(const VoxelConst &k, Q q) {
    return grad_terms_xyz(k.z, k.y, k.x, __uint_as_float(q[1]), __uint_as_float(q[2]), __uint_as_float(q[3]));
}
// Applies the deferred free precession: rotation of each spin about -z by
// its summed angle (sin and cos of the fp64 sum within 1.5 ulp: the rotation
// is applied once, so its rounding does not build up as a cached rotation's
// would), then the summed transverse relaxation factor.
template <int SPT>
__device__ __forceinline__ void apply_pending(SpinState<SPT> &st) {
    if (!st.pend) return;
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        float sn, cs;
        sincos_fast_d(st.ph[s], &sn, &cs);
        const float ax = st.mx[s], ay = st.my[s];
        const float rx = ax * cs + ay * sn;
        const float ry = ay * cs - ax * sn;
        st.mx[s] = rx * st.gp;
        st.my[s] = ry * st.gp;
        st.ph[s] = 0.0;
    }
    st.gp = 1.f;
    st.pend = false;
}

// The relaxation of one deferred free step: Mz as the reference rounds it
// (PREC 0) or in the accurate form (PREC 1); the transverse factor goes
// into the pending product.
template <int SPT, int PREC>
__device__ __forceinline__ void relax_step(SpinState<SPT> &st, float m0s) {
    if constexpr (PREC == 0) {
        st.gp = st.gp * st.e2;
#pragma unroll
        for (int s = 0; s < SPT; s++) st.mz[s] = relax_mz_ref(st.mz[s], st.e1, st.m0e);
    } else {
        st.gp = st.gp - st.gp * st.d2;
#pragma unroll
        for (int s = 0; s < SPT; s++) st.mz[s] = st.mz[s] + (m0s - st.mz[s]) * st.d1;
    }
}

// One step of free precession + relaxation with a cached rotation
// (BlochKernel.h:314-368); PREC selects the relaxation form. In reference
// mode the cached (c, s) already carry the T2 factor e2 (6 instead of 8
// operations per isochromat-step; the transverse rounding differs from the
// reference's rotate-then-scale by about an ulp per step, which damps out
// with T2, whereas Mz keeps the reference's rounding exactly).
template <int SPT, int PREC>
__device__ __forceinline__ void free_step(SpinState<SPT> &st, float m0s) {
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        const float ax = st.mx[s], ay = st.my[s];
        const float rx = ax * st.cc[s] + ay * st.ss[s];
        const float ry = ay * st.cc[s] - ax * st.ss[s];
        if constexpr (PREC == 0) {
            st.mx[s] = rx;
            st.my[s] = ry;
            st.mz[s] = relax_mz_ref(st.mz[s], st.e1, st.m0e);
        } else {
            st.mx[s] = rx - rx * st.d2;
            st.my[s] = ry - ry * st.d2;
            st.mz[s] = st.mz[s] + (m0s - st.mz[s]) * st.d1;
        }
    }
}

// Relaxation factors of one record computed on the host for a single-voxel
// scan (one T1, T2 for the wave): {e2, e1} (PREC 0) or {d2, d1} (PREC 1),
// rounded exactly as the device branch rounds them; on = false evaluates
// them per lane.
struct RelaxIn {
    bool on;
    float a, b;
};

// Applies the step of record q (flags already read; REC_NOSTEP excluded by
// the caller). q is any word accessor: a scalar-path pointer into the record
// stream, or a record already held in registers. GEN: any Tx coil count
// (TxMaps), else one Tx coil (VoxelConst weight). PREC: 0 = the reference's
// fp32 relaxation rounding, 1 = accurate. DEFER: a free step without an ADC
// sample only adds its rotation angle (the reference's fp32 angle, times the
// record's step count) to the spin's pending fp64 sum and relaxes Mz step
// by step; the caller applies the pending rotation (apply_pending) before
// anything reads or turns Mx, My (an ADC sample, a capture, the end of the
// launch); rf steps apply it here. Free precession commutes with itself, so
// this changes the rounding only: one rotation by the exact sum of the
// reference's angles instead of one fp32 rotation per step.
template <int SPT, bool GEN, int PREC, bool DEFER = false, typename Q>
__device__ __forceinline__ void bloch_step(SpinState<SPT> &st, const VoxelConst &k, Q q, uint32_t flags,
                                           const TxMaps &txm, RelaxIn relax = RelaxIn{false, 0.f, 0.f}) {
    const float dt = __uint_as_float(q[0]);
    if ((flags & REC_NEWDT) || !st.have_e) {
        st.have_e = true;
        if (relax.on) {
            if constexpr (PREC == 0) {
                st.e2 = relax.a;
                st.e1 = relax.b;
                st.m0e = k.m0s * (1.0f - st.e1);
            } else {
                st.d2 = relax.a;
                st.d1 = relax.b;
            }
        } else if constexpr (PREC == 0) {
            // BlochKernel.h:348-368 in fp32: E = exp(max(-dt/T, -200)), with
            // exp rounded correctly from fp64
            float a2 = (1.0f / k.t2) * dt;
            a2 = -a2;
            if (a2 < -200.f) a2 = -200.f;
            float a1 = (1.0f / k.t1) * dt;
            a1 = -a1;
            if (a1 < -200.f) a1 = -200.f;
            st.e2 = (float)exp((double)a2);
            st.e1 = (float)exp((double)a1);
            st.m0e = k.m0s * (1.0f - st.e1);
        } else {
            st.d2 = (float)(-expm1(-(double)dt / (double)k.t2));
            st.d1 = (float)(-expm1(-(double)dt / (double)k.t1));
        }
        // (deferred kernels rotate per step only at ADC samples, where the
        // cached path serves single-use rotations too)
        if constexpr (!DEFER) {
#pragma unroll
            for (int s = 1; s < SPT; s++) {
                sincos_cr((st.wb[s] - st.wb[0]) * dt, &st.sd[s], &st.cd[s]);
                if constexpr (PREC == 0) {
                    st.sd[s] *= st.e2;
                    st.cd[s] *= st.e2;
                }
            }
        }
    }
    const GradTerms gt = grad_terms(k, q);
    if (flags & REC_RF) {
        if constexpr (DEFER) apply_pending(st);
        const float wrf = __uint_as_float(q[4]);
        float bx, by;
        if (!GEN || txm.tx == 1) {
            const float br = __uint_as_float(q[REC_HEAD]), bi = __uint_as_float(q[REC_HEAD + 1]);
            bx = k.wtx_r * br - k.wtx_i * bi;
            by = k.wtx_r * bi + k.wtx_i * br;
        } else {
            bx = 0.f;
            by = 0.f;
            for (int c = 0; c < txm.tx; c++) {
                const float br = __uint_as_float(q[REC_HEAD + 2 * c]);
                const float bi = __uint_as_float(q[REC_HEAD + 2 * c + 1]);
                float sp, cp;
                sincos_cr(txm.pe[txm.v + (int64_t)c * txm.nvox], &sp, &cp);
                const float mg = txm.mg[txm.v + (int64_t)c * txm.nvox];
                bx += mg * (cp * br - sp * bi);
                by += mg * (cp * bi + sp * br);
            }
        }
        const float b2 = bx * bx + by * by;
#pragma unroll
        for (int s = 0; s < SPT; s++) {
            // rotation by |Omega|*dt about -Omega, Omega = (bx, by, w): the
            // matrix of BlochKernel.h:146-306 in Rodrigues form
            // (the rotation is applied once, so the hardware reciprocal
            // square root, 1 ulp, and the 1.5 ulp fp32 sincos keep the
            // step within a few ulp of the correctly rounded one; below
            // |Omega| = 1e-15 rad/s the rotation, under 1e-15 rad per
            // second of dt, is skipped, which keeps rsq off subnormals)
            const float w = ref_dw(st.wb[s], gt) + wrf;
            const float om2 = b2 + w * w;
            if (om2 > 1e-30f) {
                const float inv = __builtin_amdgcn_rsqf(om2);
                const float om = om2 * inv;
                const float nx = bx * inv, ny = by * inv, nz = w * inv;
                float sn, cs;
                sincos_fast(om * dt, &sn, &cs);
                const float ax = st.mx[s], ay = st.my[s], az = st.mz[s];
                const float kk = (1.f - cs) * (nx * ax + ny * ay + nz * az);
                st.mx[s] = ax * cs - (ny * az - nz * ay) * sn + nx * kk;
                st.my[s] = ay * cs - (nz * ax - nx * az) * sn + ny * kk;
                st.mz[s] = az * cs - (nx * ay - ny * ax) * sn + nz * kk;
            }
        }
        if constexpr (PREC == 0) {
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                st.mx[s] = st.mx[s] * st.e2;
                st.my[s] = st.my[s] * st.e2;
                st.mz[s] = relax_mz_ref(st.mz[s], st.e1, st.m0e);
            }
        } else {
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                st.mx[s] = st.mx[s] - st.mx[s] * st.d2;
                st.my[s] = st.my[s] - st.my[s] * st.d2;
                st.mz[s] = st.mz[s] + (k.m0s - st.mz[s]) * st.d1;
            }
        }
    } else {
        if constexpr (DEFER) {
            if (!(flags & REC_ADC)) {
                const int n = 1 + (int)(flags >> REC_REP_SHIFT);
#pragma unroll
                for (int s = 0; s < SPT; s++) st.ph[s] += (double)(ref_dw(st.wb[s], gt) * dt) * (double)n;
                st.pend = true;
                st.have_rot = false;  // the next REC_SAME step computes its rotation
                if (n & 1) relax_step<SPT, PREC>(st, k.m0s);
                for (int r = n & 1; r < n; r += 2) {
                    relax_step<SPT, PREC>(st, k.m0s);
                    relax_step<SPT, PREC>(st, k.m0s);
                }
                return;
            }
        }
        if (!DEFER && (flags & REC_ONCE)) {
            // rotation used by this step only: spin 0's angle, then angle
            // addition with the cached per-spin offsets (about 2 ulp, not
            // accumulated since no later step reuses it)
            float s0, c0;
            sincos_fast(ref_dw(st.wb[0], gt) * dt, &s0, &c0);
            if constexpr (PREC == 0) {
                st.ss[0] = s0 * st.e2;
                st.cc[0] = c0 * st.e2;
            } else {
                st.ss[0] = s0;
                st.cc[0] = c0;
            }
#pragma unroll
            for (int s = 1; s < SPT; s++) {
                st.cc[s] = c0 * st.cd[s] - s0 * st.sd[s];
                st.ss[s] = s0 * st.cd[s] + c0 * st.sd[s];
            }
            st.have_rot = false;
        } else if (!((flags & REC_SAME) && st.have_rot)) {
            // a REC_SAME step has this step's dt, so e2 stays valid with the cache
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                sincos_cr(ref_dw(st.wb[s], gt) * dt, &st.ss[s], &st.cc[s]);
                if constexpr (PREC == 0) {
                    st.ss[s] *= st.e2;
                    st.cc[s] *= st.e2;
                }
            }
            st.have_rot = true;
        }
        // this step, then the folded repeats of it (in pairs: the compiler
        // then alternates registers instead of copying the state each step)
        const int n = 1 + (int)(flags >> REC_REP_SHIFT);
        if (n & 1) free_step<SPT, PREC>(st, k.m0s);
        for (int r = n & 1; r < n; r += 2) {
            free_step<SPT, PREC>(st, k.m0s);
            free_step<SPT, PREC>(st, k.m0s);
        }
    }
}

}  // namespace mrs
