// Device helpers shared by the Bloch kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrs {

constexpr int kRedBatch = 16;          // ADC samples buffered in LDS before a flush

// Step records are wave-uniform: reading them through address space 4 puts
// them on the scalar path (s_load into SGPRs).
typedef const __attribute__((address_space(4))) uint32_t *crec_t;


// sin and cos of an fp64 angle (sincos_cr: of an fp32 angle), correctly
// rounded to fp32 (up to rare near-midpoint cases): the reduction by pi/2 (two-part constant, FMA, exact
// enough for |x| up to 1e15 rad) and the Taylor polynomials of sin and cos
// on [-pi/4, pi/4] (truncation error below 3e-14) run in fp64, which on
// gfx950 issues at the fp32 FMA rate when packed fp32 is off. Beyond 1e15
// rad per step (never reached by a physical scan) the angle is first
// reduced modulo the fp64 value of 2*pi, which keeps the rotation unitary
// but not its phase. MRiLab's CPU engine takes sin/cos from a
// 24-bit-accurate vector library; a 1.5 ulp fp32 sincos drifts visibly
// from it over 1e5 cached rotations, this one does not.
__device__ __forceinline__ void sincos_cr_d(double x, float *sp, float *cp) {
    if (fabs(x) > 1.0e15) x = fmod(x, 6.283185307179586477);
    const double k = rint(x * 0.63661977236758134308);   // 2/pi
    double r = fma(-k, 1.5707963267948966192, x);        // pi/2, high part
    r = fma(-k, 6.1232339957367660e-17, r);              // pi/2, low part
    const double r2 = r * r;
    const double sp_ = r + r * r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 +
                       r2 * (-1.0 / 39916800 + r2 * (1.0 / 6227020800.0))))));
    const double cp_ = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 +
                       r2 * (-1.0 / 3628800 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0)))))));
    // quadrant: rounding commutes with negation, so the swap and the signs
    // are applied to the fp32 values (selects, not branches on fp64 pairs)
    const int q = (int)(((long long)k) & 3);
    const float sf = (float)sp_, cf = (float)cp_;
    const float so = (q & 1) ? cf : sf;
    const float co = (q & 1) ? sf : cf;
    *sp = (q & 2) ? -so : so;
    *cp = ((q + 1) & 2) ? -co : co;
}
__device__ __forceinline__ void sincos_cr(float xf, float *sp, float *cp) { sincos_cr_d((double)xf, sp, cp); }

// The Taylor polynomials of sincos_fast on the reduced angle r in
// [-pi/4, pi/4] and the quadrant q of the angle (its multiple of pi/2, mod 4).
__device__ __forceinline__ void sincos_poly(float r, int q, float *sp, float *cp) {
    const float r2 = r * r;
    const float s = __builtin_fmaf(r * r2, __builtin_fmaf(r2, __builtin_fmaf(r2, __builtin_fmaf(r2, 2.75573192e-06f,
                                   -1.98412698e-04f), 8.33333333e-03f), -1.66666667e-01f), r);
    const float c = __builtin_fmaf(r2, __builtin_fmaf(r2, __builtin_fmaf(r2, __builtin_fmaf(r2, __builtin_fmaf(r2,
                                   -2.75573192e-07f, 2.48015873e-05f), -1.38888889e-03f), 4.16666667e-02f), -0.5f), 1.0f);
    const float so = (q & 1) ? c : s;
    const float co = (q & 1) ? s : c;
    *sp = (q & 2) ? -so : so;
    *cp = ((q + 1) & 2) ? -co : co;
}

// sin and cos of an fp32 angle in fp32, within 1.5 ulp: Cody-Waite
// reduction by pi/2 (three-part constant, FMA) and the Taylor polynomials of
// sin (to r^9) and cos (to r^10) on [-pi/4, pi/4]. For rotations applied
// once; a rotation reused over a run of steps takes sincos_cr, whose
// rounding does not bias |R| over many applications.
__device__ __forceinline__ void sincos_fast(float x, float *sp, float *cp) {
    const float k = __builtin_rintf(x * 0.636619772367581343f);
    if (__builtin_fabsf(k) > 65536.f) {
        sincos_cr(x, sp, cp);
        return;
    }
    float r = __builtin_fmaf(-k, 1.57079637050628662109375f, x);
    r = __builtin_fmaf(-k, -4.37113882867379e-08f, r);
    r = __builtin_fmaf(-k, -1.77635684e-15f, r);
    sincos_poly(r, ((int)k) & 3, sp, cp);
}

// sin and cos of an fp64 angle in fp32, within 1.5 ulp: the reduction by
// pi/2 in fp64 (as sincos_cr_d), then sincos_fast's fp32 polynomials. For
// rotations applied once, such as the summed angle of deferred free steps.
__device__ __forceinline__ void sincos_fast_d(double x, float *sp, float *cp) {
    if (fabs(x) > 1.0e15) x = fmod(x, 6.283185307179586477);
    const double k = rint(x * 0.63661977236758134308);
    double r = fma(-k, 1.5707963267948966192, x);
    r = fma(-k, 6.1232339957367660e-17, r);
    sincos_poly((float)r, (int)(((long long)k) & 3), sp, cp);
}

// The gradient terms of the off-resonance, position times gamma*G, each
// rounded as BlochKernel.h:44-51 rounds them (gamma*G comes rounded in the
// record).
struct GradTerms {
    float z, y, x;
};
__device__ __forceinline__ GradTerms grad_terms_xyz(float z, float y, float x, float gz, float gy, float gx) {
#pragma clang fp contract(off)
    return {z * gz, y * gy, x * gx};
}
// Off-resonance of a spin (rad/s) in BlochKernel.h:38-51's order: the spin's
// field-free term wb, then the z, y and x gradient terms one at a time, each
// sum rounded. bSSFP makes the signal sensitive to the phase each spin
// gathers per TR, so the angle of a step is kept bit-equal to the reference's.
__device__ __forceinline__ float ref_dw(float wb, GradTerms g) {
#pragma clang fp contract(off)
    float w = wb + g.z;
    w = w + g.y;
    return w + g.x;
}

// Lane exchange for one butterfly stage across lane bit O. For O = 32 and
// 16, gfx950's v_permlane{32,16}_swap exchange whole rows; for O <= 8, DPP
// row_mirror / row_half_mirror / quad_perm pair each lane with one whose bit
// O differs (not always lane ^ O, but the same pair in both directions).
template <int O>
__device__ __forceinline__ float dpp_partner(float x) {
    constexpr int ctrl = O == 8 ? 0x140 : O == 4 ? 0x141 : O == 2 ? 0x4E : 0xB1;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xF, 0xF, true));
}

// Transposed butterfly: N values per lane, 64 lanes. Each stage halves the
// number of values a lane keeps while pairing lanes across lane bit O (the
// lane with bit O set keeps the upper half); when one value is left the
// remaining stages sum plainly. Afterwards lane l holds the wave total of
// value l >> (6 - log2 N). With the row swaps a pair costs one swap and
// one add: after swap(X = v[i], Y = v[i+H]) the two outputs hold, in every
// lane, exactly the two addends of the value that lane keeps.
template <int N, int O>
struct WaveSum {
    static __device__ __forceinline__ void run(float *v, int lane) {
        if constexpr (O >= 1) {
            if constexpr (N > 1) {
                constexpr int H = N / 2;
                if constexpr (O >= 16) {
#pragma unroll
                    for (int i = 0; i < H; i++) {
                        const auto r = O == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]), __float_as_uint(v[i + H]), false, false)
                                               : __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]), __float_as_uint(v[i + H]), false, false);
                        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
                    }
                } else {
                    const bool up = (lane & O) != 0;
#pragma unroll
                    for (int i = 0; i < H; i++) {
                        const float lo = v[i] + dpp_partner<O>(v[i]);
                        const float hi = v[i + H] + dpp_partner<O>(v[i + H]);
                        v[i] = up ? hi : lo;
                    }
                }
                WaveSum<H, O / 2>::run(v, lane);
            } else {
                if constexpr (O >= 16) {
                    const auto r = O == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(v[0]), __float_as_uint(v[0]), false, false)
                                           : __builtin_amdgcn_permlane16_swap(__float_as_uint(v[0]), __float_as_uint(v[0]), false, false);
                    // r[0] holds the partner's value in the upper rows, r[1] in the lower rows
                    v[0] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
                } else {
                    v[0] += dpp_partner<O>(v[0]);
                }
                WaveSum<1, O / 2>::run(v, lane);
            }
        }
    }
};

// Mz relaxation exactly as BlochKernel.h:363-368 rounds it: the product and
// the sum are rounded separately. Near equilibrium the per-step change is
// below half an ulp of Mz and fp32 recovery stalls; where it stalls depends
// on this rounding, so it must not be fused into an FMA.
__device__ __forceinline__ float relax_mz_ref(float mz, float e1, float m0e) {
#pragma clang fp contract(off)
    const float t = mz * e1;
    return t + m0e;
}

// Block-level flush of buffered per-wave ADC values: sums the waves of the
// block in LDS order and writes one partial per (sample, value).
// MFMA coil projection (forward kernel, 4 or 8 receive coils). Each wave
// stages its lanes' per-sample spin sums in LDS, kMfmaRows samples at a
// time: row 2s holds sum Mx and row 2s+1 sum My of staged sample s, one
// float per voxel, rows kMfmaStride floats apart (the 4-float pad keeps the
// 16-byte row reads of 8 lanes on distinct banks). Every kMfmaRows samples
// one chain of 16 16x16x4 f32 MFMAs, K running over the wave's 64 voxels,
// forms the real products
//   P[2s+p][2c+q] = sum_v m_p(s, v) C_q(v, c),  m_0 = sum Mx, m_1 = sum My,
//                                              C_0 = Cx, C_1 = Cy
// with the B operand (the coil maps) stored in LDS in the order the MFMA
// steps read it. Every one of the 256 products is used: lane (g, j) holds
// rows 4g..4g+3 of column j, and one exchange with the neighbouring lane
// gives
//   Re_c(s) = P[2s][2c] + P[2s+1][2c+1] = sum_v (sx Cx + sy Cy)
//   Im_c(s) = P[2s+1][2c] - P[2s][2c+1] = sum_v (sy Cx - sx Cy)
// in column 2c (Re) and 2c+1 (Im), samples 2g and 2g+1.
constexpr int kMfmaRows = 8;
constexpr int kMfmaStride = 64 + 4;
constexpr int kMfmaCoils = 8;
// per wave: 2 kMfmaRows staged rows, then 16 x 64 B-operand floats
constexpr int kMfmaWaveFloats = 2 * kMfmaRows * kMfmaStride + 16 * 64;

typedef float mfma_v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void flush_partials(const float *red, int nslot, int nwaves, int nv, float *partial,
                                               int64_t adc0, int nblk, int blk, int vstride) {
    for (int i = threadIdx.x; i < nslot * nv; i += blockDim.x) {
        const int sl = i / nv, j = i - sl * nv;
        float acc = 0.f;
#pragma unroll 4
        for (int w = 0; w < nwaves; w++) acc += red[(sl * nwaves + w) * nv + j];
        partial[((adc0 + sl) * nblk + blk) * vstride + j] = acc;
    }
}

}  // namespace mrs
