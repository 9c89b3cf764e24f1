/*
 * bloch_oracle.c -- CPU oracle of the single-pool Bloch scan.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline, never as part of the product path.
 *
 * It restates MRiLab's CPU engine:
 *   event loop   Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:246-553
 *   Bloch step   Lib/src/engine/CPUEngine/BlochKernel.h:7-397
 *   Ext plugins  Src/Main/DoExtPlugin.m:11-33, Src/Plugin/Plugin_*.m
 * The same source is built twice:
 *   ORC_REAL=float  -> liborc_f32.so: every vector pass of BlochKernel.h is
 *                      one rounded fp32 operation in the reference's order
 *                      (libm in place of IPP's A24 functions; exp rounded
 *                      correctly from fp64).
 *   ORC_REAL=double -> liborc_f64.so: the same discrete scheme in fp64.
 * The reference itself cannot be built here (it needs mex.h and Intel IPP,
 * neither of which exists in this environment).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/mrilab_amd.h"

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

#define ORC_PI 3.14159265359 /* DoScanAtCPU.cpp:63 */

/* Planted mistakes, for the oracle's own tests only: tests/test_oracle_pins.py
 * compiles a copy of this file with -DORC_MUTATION=n and checks that one of
 * the closed-form pins fails on it. Every library that checks the engine is
 * built with the default 0, where each ORC_MUT(n) below is constant false.
 *   1 multi-Tx phase from atan2(re, im) instead of atan2(im, re)
 *   2 multi-Tx rf frequency added for the first active coil only
 *   3 multi-Tx rf frequency of a silent coil (amplitude 0) added
 *   4 bound-pool saturation W summed over active Tx coils, not averaged
 *   5 split exchange with the end-of-step state of pools updated earlier
 *   6 MT bound pool exchanging with water's end-of-step Mz
 *   7 ReleaseK (Ext 4) leaving Kx unrestored */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif
#define ORC_MUT(n) (ORC_MUTATION == (n))

static int g_threads = 0; /* 0: min(slices, MaxThreadNum) like the reference */
static int g_relax = 0;   /* 0: reference order; 1: expm1 form (error study only) */

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return g_threads; }
void orc_set_relax_mode(int m) { g_relax = m; }

int orc_real_size(void) { return (int)sizeof(real); }

static real r_sin(real x) { return sizeof(real) == 4 ? (real)sinf((float)x) : (real)sin((double)x); }
static real r_cos(real x) { return sizeof(real) == 4 ? (real)cosf((float)x) : (real)cos((double)x); }
static real r_atan(real x) { return sizeof(real) == 4 ? (real)atanf((float)x) : (real)atan((double)x); }
static real r_atan2(real y, real x) { return sizeof(real) == 4 ? (real)atan2f((float)y, (float)x) : (real)atan2((double)y, (double)x); }
static real r_sqrt(real x) { return sizeof(real) == 4 ? (real)sqrtf((float)x) : (real)sqrt((double)x); }
/* exp of the relaxation factors: fp32 build rounds exp correctly (from
 * fp64), the canonical fp32 exp; IPP's ippsExp_32f_I is not available and
 * any accurate fp32 exp agrees with it except near rounding midpoints. */
static real r_exp(real x) { return (real)exp((double)x); }

/* The rf rotation of BlochKernel.h:124-306 for one isochromat: b1 is the
 * gamma-scaled B1 magnitude, b2 the negated phase -(pe + phase), dW the
 * off-resonance including the rf frequency; each statement is one IPP pass. */
static void orc_rf_rotate(real dW, real b1, real b2, real dt, real mx, real my, real mz,
                          real *omx, real *omy, real *omz)
{
    real nmx, nmy, nmz;
    /* beta = atan(dW / b1), alpha = sqrt(dW^2 + b1^2) * dt, :124-133 */
    real q = dW / b1;
    real beta = r_atan(q);
    real s2 = dW * dW; b1 = b1 * b1; s2 = b1 + s2; s2 = r_sqrt(s2);
    real alpha = s2 * dt;
    real sA = r_sin(alpha), sB = r_sin(beta), cA = r_cos(alpha), cB = r_cos(beta);
    real sP = r_sin(b2), msP = sP * -1, cP = r_cos(b2), mcP = cP * -1;
    real t1, t2, t3, t4, acc;
    /* Mx, :146-196 */
    t1 = sB * cA; t1 = mcP * t1; t3 = sA * sP; t1 = t3 + t1; t1 = sB * t1; t1 = t1 * -1;
    t3 = cB * cB; t3 = cP * t3; t1 = t3 + t1; t3 = t1;
    t1 = cP * t1; t3 = sP * t3;
    t2 = sA * sB; t2 = cP * t2; t4 = cA * sP; t2 = t4 + t2; t4 = t2;
    t2 = sP * t2; t4 = mcP * t4;
    t2 = t1 + t2; t2 = mx * t2;
    t4 = t3 + t4; t4 = my * t4; t4 = t4 * -1;
    acc = t2 + t4;
    t1 = sB * cB; t1 = cP * t1;
    t2 = sB * cA; t2 = mcP * t2;
    t3 = sA * sP;
    t2 = t3 + t2; t2 = cB * t2;
    t1 = t2 + t1; t1 = mz * t1;
    nmx = acc + t1;
    /* My, :205-254 */
    t1 = cA * sB; t1 = sP * t1; t3 = sA * cP; t1 = t3 + t1; t1 = sB * t1;
    t3 = cB * cB; t3 = sP * t3; t1 = t3 + t1; t3 = t1;
    t1 = sP * t1; t3 = cP * t3;
    t2 = sA * sB; t2 = msP * t2; t4 = cA * cP; t2 = t4 + t2; t4 = t2;
    t2 = cP * t2; t4 = msP * t4;
    t2 = t1 + t2; t2 = my * t2;
    t4 = t3 + t4; t4 = mx * t4; t4 = t4 * -1;
    acc = t2 + t4;
    t1 = cB * sB; t1 = msP * t1;
    t2 = cA * sB; t2 = sP * t2;
    t3 = sA * cP;
    t2 = t3 + t2; t2 = cB * t2;
    t1 = t2 + t1; t1 = mz * t1;
    nmy = acc + t1;
    /* Mz, :263-299 */
    t1 = cA * cB; t1 = sB * t1; t1 = t1 * -1;
    t3 = cB * sB; t1 = t3 + t1; t3 = t1;
    t1 = cP * t1; t3 = sP * t3;
    t2 = cB * sA; t2 = msP * t2;
    t4 = cB * sA; t4 = cP * t4;
    t2 = t1 + t2; t2 = mx * t2;
    t4 = t3 + t4; t4 = my * t4; t4 = t4 * -1;
    acc = t2 + t4;
    t1 = sB * sB;
    t2 = cB * cB; t2 = cA * t2;
    t1 = t2 + t1; t1 = mz * t1;
    nmz = acc + t1;
    *omx = nmx; *omy = nmy; *omz = nmz;
}

/*
 * One Bloch step for the voxels of one slice of one (type, spin).
 * BlochKernel.h:7-397; every statement below is one elementwise IPP pass
 * of the reference, applied per element.
 */
static void orc_bloch_slice(real Gyro, const double *CS, int SpinNum,
                            const float *Rho, const float *T1, const float *T2,
                            float *Mz, float *My, float *Mx,
                            const float *dB0, const float *dWRnd,
                            const float *Gzgrid, const float *Gygrid, const float *Gxgrid,
                            const float *TxCoilmg, const float *TxCoilpe,
                            real dt, const double *rfAmp, const double *rfPhase, const double *rfFreq,
                            real GzAmp, real GyAmp, real GxAmp,
                            int Typei, int64_t SpinMxNum, int64_t SpinMxSliceNum, int TxCoilNum,
                            real *stMx, real *stMy, real *stMz)
{
    real rfAmpSum = 0;
    for (int i = 0; i < TxCoilNum; i++)
        rfAmpSum += (real)fabs((double)(real)rfAmp[i]);          /* :54-56 */
    const real csterm = (real)((double)((real)CS[Typei] * 2) * ORC_PI); /* :42 */
    const real gzg = GzAmp * Gyro, gyg = GyAmp * Gyro, gxg = GxAmp * Gyro;

    for (int64_t v = 0; v < SpinMxNum; v++) {
        real mx = stMx ? stMx[v] : (real)Mx[v];
        real my = stMy ? stMy[v] : (real)My[v];
        real mz = stMz ? stMz[v] : (real)Mz[v];
        real nmx, nmy, nmz = mz;
        /* E1 off-resonance, :38-51 */
        real dW = (real)dB0[v] * Gyro;
        dW = dW + (real)dWRnd[v];
        dW = dW + csterm;
        real buf = (real)Gzgrid[v] * gzg; dW = dW + buf;
        buf = (real)Gygrid[v] * gyg; dW = dW + buf;
        buf = (real)Gxgrid[v] * gxg; dW = dW + buf;

        if (rfAmpSum != 0) {
            real b1, b2;
            if (TxCoilNum == 1) {                                   /* :58-72 */
                real A = (real)rfAmp[0], P = (real)rfPhase[0], F = (real)rfFreq[0];
                if (A < 0) { A = (real)fabs((double)A); P = (real)((double)P + ORC_PI); }
                b1 = (real)TxCoilmg[v] * A;
                b2 = (real)TxCoilpe[v] + P;
                dW = dW + (real)((double)(F * 2) * ORC_PI);
                b1 = b1 * Gyro;
                b2 = b2 * -1;
            } else {                                                /* :73-121 */
                real b5 = 0, b6 = 0;
                int nact = 0;
                for (int i = 0; i < TxCoilNum; i++) {
                    real A = (real)rfAmp[i], P = (real)rfPhase[i], F = (real)rfFreq[i];
                    if (ORC_MUT(3) && A == 0) dW = dW + (real)((double)(F * 2) * ORC_PI);
                    if (A != 0) {
                        if (A < 0) { A = (real)fabs((double)A); P = (real)((double)P + ORC_PI); }
                        real m = (real)TxCoilmg[v + i * SpinMxSliceNum * SpinMxNum] * A;
                        real ph = (real)TxCoilpe[v + i * SpinMxSliceNum * SpinMxNum] + P;
                        real re = m * r_cos(ph), im = m * r_sin(ph);   /* PolarToCart */
                        b5 = b5 + re; b6 = b6 + im;
                        if (!(ORC_MUT(2) && nact > 0)) dW = dW + (real)((double)(F * 2) * ORC_PI);
                        nact++;
                    }
                }
                b1 = r_sqrt(b5 * b5 + b6 * b6);                     /* CartToPolar */
                b2 = ORC_MUT(1) ? r_atan2(b5, b6) : r_atan2(b6, b5);
                b1 = b1 * Gyro;
                b2 = b2 * -1;
            }
            orc_rf_rotate(dW, b1, b2, dt, mx, my, mz, &nmx, &nmy, &nmz);
        } else {
            /* free precession, :314-343 */
            real alpha = dW * dt;
            real sA = r_sin(alpha), cA = r_cos(alpha);
            real t1 = mx * cA, t2 = my * sA;
            nmx = t1 + t2;
            real t3 = my * cA, t4 = mx * sA; t4 = t4 * -1;
            nmy = t3 + t4;
        }
        /* relaxation, :346-368 */
        real e2 = (real)1 / (real)T2[v]; e2 = e2 * dt; e2 = e2 * -1;
        if (e2 < -200) e2 = -200;
        e2 = r_exp(e2);
        real e1 = (real)1 / (real)T1[v]; e1 = e1 * dt; e1 = e1 * -1;
        if (e1 < -200) e1 = -200;
        e1 = r_exp(e1);
        if (g_relax == 0) {
            nmx = e2 * nmx;
            nmy = e2 * nmy;
            real r3 = e1 * -1; r3 = r3 + 1;
            real r4 = (real)Rho[v] / (real)SpinNum;
            r3 = r4 * r3;
            r4 = nmz * e1;
            nmz = r3 + r4;
        } else {
            /* M' = M - M*d with d = -expm1(-dt/T) computed without the
             * cancellation of 1 - exp(-dt/T); not the reference's order. */
            real d2 = (real)-expm1(-(double)dt / (double)T2[v]);
            real d1 = (real)-expm1(-(double)dt / (double)T1[v]);
            real m0 = (real)Rho[v] / (real)SpinNum;
            nmx = nmx - nmx * d2;
            nmy = nmy - nmy * d2;
            nmz = nmz + (m0 - nmz) * d1;
        }
        if (stMx) { stMx[v] = nmx; stMy[v] = nmy; stMz[v] = nmz; }
        else { Mx[v] = (float)nmx; My[v] = (float)nmy; Mz[v] = (float)nmz; }
    }
}

/* fp32 sum in 16 interleaved partial sums, like a vectorised ippsSum_32f
 * with ippAlgHintFast; fp64 build: plain fp64 sum. */
static double orc_sum(const real *x, int64_t n)
{
    if (sizeof(real) == 4) {
        float p[16] = {0};
        for (int64_t i = 0; i < n; i++) p[i & 15] += (float)x[i];
        for (int w = 8; w >= 1; w >>= 1)
            for (int i = 0; i < w; i++) p[i] += p[i + w];
        return (double)p[0];
    }
    double s = 0;
    for (int64_t i = 0; i < n; i++) s += (double)x[i];
    return s;
}

typedef struct orc_ctx {
    mrs_vobj *vobj; mrs_vmag *vmag; mrs_vcoi *vcoi; const mrs_vctl *vctl;
    const mrs_vseq *vseq; mrs_vvar *vvar; mrs_vsig *vsig; const mrs_callbacks *cb;
    int64_t nvox;   /* R*C*Z */
    int64_t nxy;    /* R*C */
    int64_t nz;
    real *wMx, *wMy, *wMz; /* fp64 build keeps the state in double */
    int variant;    /* multi-pool model: ORC_GM, ORC_MT, ORC_CEST, ORC_ME */
} orc_ctx;

enum { ORC_GM = 0, ORC_MT = 1, ORC_CEST = 2, ORC_ME = 3 };

/* ======================================================================
 * Generalized multi-pool model (MRiLab's GM engine,
 * Lib/src/engine/GPUEngine/BlochKernelGM.cuh:110-487). Pools t with
 * TypeFlag 0 are free (3 components), TypeFlag 1 bound (Mz only).
 * K is [R,C,Z,Ty,Ty]: plane p*Ty + t is the rate from pool t into pool p
 * (BlochKernelGM.cuh:417 "go away", :423 "come in").
 * ==================================================================== */

/* Super-Lorentzian lineshape of the bound pool as BlochKernelGM.cuh:339-344
 * evaluates it: a 1000-point Riemann sum over u = 0, 0.001, ..., with u
 * accumulated by repeated addition. dw = 2*pi*rfFreq + gamma*dB0 (rad/s). */
static double orc_sl_lineshape(double dw, double T2b)
{
    real u = 0, G = 0;
    const real c = (real)sqrt(2.0 / ORC_PI);
    for (int p = 0; p < 1000; p++) {
        real den = (real)3 * u * u - 1;
        real b = ((real)dw * (real)T2b) / den;
        G += (c * ((real)T2b / (real)fabs((double)den)) * r_exp(-2 * b * b)) * (real)0.001;
        u += (real)0.001;
    }
    return (double)G;
}

double orc_superlorentzian(double dw, double T2b) { return orc_sl_lineshape(dw, T2b); }

typedef struct gm_cache {   /* per (voxel, spin, pool): lineshape cache key and value */
    double key_f, key_t2, G;
    int valid;
} gm_cache;

/* Saturation rate W of a bound pool (BlochKernelGM.cuh:323-380). */
static double orc_gm_W(const orc_ctx *C, int64_t v, int t, double T2b, const double *rfAmp,
                       const double *rfFreq, gm_cache *gc)
{
    const int Tx = C->vcoi->TxCoilNum;
    const double gyro = (real)C->vobj->Gyro;
    const double dB0 = C->vmag->dB0[v];
    if (Tx == 1) {
        real A = (real)fabs((double)(real)rfAmp[0]);
        real F = (real)rfFreq[0];
        real b1 = (real)C->vcoi->TxCoilmg[v] * A;
        if (!gc->valid || gc->key_f != F || gc->key_t2 != T2b) {
            gc->key_f = F; gc->key_t2 = T2b; gc->valid = 1;
            gc->G = orc_sl_lineshape(2 * ORC_PI * F + dB0 * gyro, T2b);
        }
        real bg = b1 * (real)gyro;
        return (double)((real)ORC_PI * bg * bg * (real)gc->G);
    }
    double W = 0; int n = 0;
    for (int c = 0; c < Tx; c++) {
        real A = (real)rfAmp[c], F = (real)rfFreq[c];
        if (A == 0) continue;
        A = (real)fabs((double)A);
        real b1 = (real)C->vcoi->TxCoilmg[v + (int64_t)c * C->nvox] * A;
        if (!gc->valid || gc->key_f != F || gc->key_t2 != T2b) {
            gc->key_f = F; gc->key_t2 = T2b; gc->valid = 1;
            gc->G = orc_sl_lineshape(2 * ORC_PI * F + dB0 * gyro, T2b);
        }
        real bg = b1 * (real)gyro;
        W += (double)((real)ORC_PI * bg * bg * (real)gc->G);
        n++;
    }
    if (ORC_MUT(4)) return W;
    return n ? W / n : 0.0;
}

/* Effective field of a free pool on an rf step: gamma*B1 magnitude and the
 * negated phase -(pe + phase) as BlochKernel.h:58-121 forms them, plus the
 * rf frequency term added to dW. Returns 0 if rf is off. */
static int orc_rf_field(const orc_ctx *C, int64_t v, const double *rfAmp, const double *rfPhase,
                        const double *rfFreq, real *b1o, real *b2o, real *dW)
{
    const int Tx = C->vcoi->TxCoilNum;
    const real Gyro = (real)C->vobj->Gyro;
    real rfAmpSum = 0;
    for (int i = 0; i < Tx; i++) rfAmpSum += (real)fabs((double)(real)rfAmp[i]);
    if (rfAmpSum == 0) return 0;
    real b1, b2;
    if (Tx == 1) {
        real A = (real)rfAmp[0], P = (real)rfPhase[0], F = (real)rfFreq[0];
        if (A < 0) { A = (real)fabs((double)A); P = (real)((double)P + ORC_PI); }
        b1 = (real)C->vcoi->TxCoilmg[v] * A;
        b2 = (real)C->vcoi->TxCoilpe[v] + P;
        *dW = *dW + (real)((double)(F * 2) * ORC_PI);
    } else {
        real b5 = 0, b6 = 0;
        int nact = 0;
        for (int i = 0; i < Tx; i++) {
            real A = (real)rfAmp[i], P = (real)rfPhase[i], F = (real)rfFreq[i];
            if (ORC_MUT(3) && A == 0) *dW = *dW + (real)((double)(F * 2) * ORC_PI);
            if (A == 0) continue;
            if (A < 0) { A = (real)fabs((double)A); P = (real)((double)P + ORC_PI); }
            real m = (real)C->vcoi->TxCoilmg[v + (int64_t)i * C->nvox] * A;
            real ph = (real)C->vcoi->TxCoilpe[v + (int64_t)i * C->nvox] + P;
            b5 = b5 + m * r_cos(ph); b6 = b6 + m * r_sin(ph);
            if (!(ORC_MUT(2) && nact > 0)) *dW = *dW + (real)((double)(F * 2) * ORC_PI);
            nact++;
        }
        b1 = r_sqrt(b5 * b5 + b6 * b6);
        b2 = ORC_MUT(1) ? r_atan2(b5, b6) : r_atan2(b6, b5);
    }
    *b1o = b1 * Gyro;
    *b2o = b2 * -1;
    return 1;
}

/* Off-resonance of a free pool (BlochKernelGM.cuh:242-245). */
static real orc_gm_dW(const orc_ctx *C, int64_t v, int64_t si, int t, real gz, real gy, real gx)
{
    const real Gyro = (real)C->vobj->Gyro;
    real dW = (real)C->vmag->dB0[v] * Gyro;
    dW = dW + (real)C->vmag->dWRnd[si];
    dW = dW + (real)((double)((real)C->vctl->CS[t] * 2) * ORC_PI);
    dW = dW + (real)C->vmag->Gzgrid[v] * (gz * Gyro);
    dW = dW + (real)C->vmag->Gygrid[v] * (gy * Gyro);
    dW = dW + (real)C->vmag->Gxgrid[v] * (gx * Gyro);
    return dW;
}

#define GM_MAXT 8

static int orc_gm_skip(const orc_ctx *C, int64_t v, int t)
{
    const int64_t tv = v + (int64_t)t * C->nvox;
    return C->vobj->T2[tv] == 0 || C->vobj->T1[tv] == 0 || C->vobj->Rho[tv] == 0;  /* :176 */
}

/* Pools frozen at a step: GM and ME skip each pool with a zero T1, T2 or
 * Rho (BlochKernelGM.cuh:176, BlochKernelME.cuh:173); MT and CEST skip the
 * whole voxel when the water pool's are zero (BlochKernelMT.cuh:121,
 * BlochKernelCEST.cuh:133). */
static int orc_vskip(const orc_ctx *C, int64_t v, int t)
{
    if (C->variant == ORC_MT || C->variant == ORC_CEST) return orc_gm_skip(C, v, 0);
    return orc_gm_skip(C, v, t);
}

/* 0 free pool, 1 bound pool: GM reads TypeFlag; MT and CEST have the bound
 * pool at index 1; ME pools are all free. */
static int orc_pool_flag(const orc_ctx *C, int t)
{
    switch (C->variant) {
    case ORC_MT: case ORC_CEST: return t == 1;
    case ORC_ME: return 0;
    default: return (int)C->vobj->TypeFlag[t];
    }
}

/* Exchange rate from pool a into pool b, read from K as each reference
 * kernel reads it:
 *   GM   plane b*Ty + a (BlochKernelGM.cuh:417 go away, :423 come in)
 *   ME   plane a*Ty + b (BlochKernelME.cuh:327, :333), diagonal included
 *   MT   plane 1 = water -> bound, plane 2 = bound -> water
 *        (BlochKernelMT.cuh:151-154)
 *   CEST plane 0 = water -> MT, plane Ty-1 = MT -> water, plane t-1 =
 *        water -> CEST pool t, plane t+Ty-2 = CEST pool t -> water
 *        (BlochKernelCEST.cuh:157-160, :379-384)  */
static real orc_rate(const orc_ctx *C, int64_t v, int a, int b)
{
    const int Ty = C->vobj->TypeNum;
    const int64_t nvox = C->nvox;
    const float *K = C->vobj->K;
    switch (C->variant) {
    case ORC_ME: return K[v + nvox * ((int64_t)a * Ty + b)];
    case ORC_MT:
        if (a == 0 && b == 1) return K[v + nvox * 1];
        if (a == 1 && b == 0) return K[v + nvox * 2];
        return 0;
    case ORC_CEST:
        if (a == 0 && b == 1) return K[v];
        if (a == 1 && b == 0) return K[v + nvox * (int64_t)(Ty - 1)];
        if (a == 0 && b >= 2) return K[v + nvox * (int64_t)(b - 1)];
        if (a >= 2 && b == 0) return K[v + nvox * (int64_t)(a + Ty - 2)];
        return 0;
    default:
        if (a == b) return 0;
        return K[v + nvox * ((int64_t)b * Ty + a)];
    }
}

/* One operator-split step of all pools of one (voxel, spin):
 * BlochKernelGM.cuh:159-476. */
static void orc_gm_split_voxel(orc_ctx *C, int64_t v, int s, real dt, real gz, real gy, real gx,
                               const double *rfAmp, const double *rfPhase, const double *rfFreq,
                               gm_cache *gcs)
{
    const int Ty = C->vobj->TypeNum, S = C->vobj->SpinNum;
    const int64_t nvox = C->nvox;
    real Bx[GM_MAXT], By[GM_MAXT], Bz[GM_MAXT];
    int64_t si[GM_MAXT];
    for (int t = 0; t < Ty; t++) {
        si[t] = v + nvox * ((int64_t)s + (int64_t)S * t);
        Bx[t] = C->wMx ? C->wMx[si[t]] : (real)C->vobj->Mx[si[t]];
        By[t] = C->wMy ? C->wMy[si[t]] : (real)C->vobj->My[si[t]];
        Bz[t] = C->wMz ? C->wMz[si[t]] : (real)C->vobj->Mz[si[t]];
    }
    for (int t = 0; t < Ty; t++) {
        if (orc_vskip(C, v, t)) continue;
        const int64_t tv = v + (int64_t)t * nvox;
        const real Rho = C->vobj->Rho[tv], T1 = C->vobj->T1[tv], T2 = C->vobj->T2[tv];
        real mx = Bx[t], my = By[t], mz = Bz[t];
        real bx = mx, by = my, bz = mz;
        const int flag = orc_pool_flag(C, t);
        if (flag == 0) {
            real dW = orc_gm_dW(C, v, si[t], t, gz, gy, gx);
            real b1, b2;
            if (orc_rf_field(C, v, rfAmp, rfPhase, rfFreq, &b1, &b2, &dW)) {
                orc_rf_rotate(dW, b1, b2, dt, mx, my, mz, &bx, &by, &bz);
            } else {
                real alpha = dW * dt, sA = r_sin(alpha), cA = r_cos(alpha);
                bx = mx * cA + my * sA;
                by = my * cA - mx * sA;
                bz = mz;
            }
        } else if (flag == 1) {
            real rfAmpSum = 0;
            for (int i = 0; i < C->vcoi->TxCoilNum; i++) rfAmpSum += (real)fabs((double)(real)rfAmp[i]);
            if (rfAmpSum != 0) {
                real W = (real)orc_gm_W(C, v, t, T2, rfAmp, rfFreq, &gcs[s * Ty + t]);
                bz = mz * r_exp(-W * dt);
            } else {
                bz = mz;
            }
        } else {
            continue;
        }
        /* relaxation (:397-405, :443-447) */
        real e1 = r_exp(-dt / T1);
        real m0 = (Rho * (1 - e1)) / (real)S;
        if (flag == 0) {
            real e2 = r_exp(-dt / T2);
            bx = bx * e2;
            by = by * e2;
        }
        bz = bz * e1;
        /* exchange with start-of-step values of the other pools (:408-433);
         * ME also applies a nonzero diagonal of K (BlochKernelME.cuh:323-338) */
        real kout = 1, kx = 0, ky = 0, kz = 0;
        for (int p = 0; p < Ty; p++) {
            if (p == t && C->variant != ORC_ME) continue;
            real ko = orc_rate(C, v, t, p);
            if (ko != 0) kout = kout * r_exp(-dt * ko);
            real ki = orc_rate(C, v, p, t);
            if (ki != 0) {
                real f = 1 - r_exp(-dt * ki);
                if (flag == 0) { kx = kx + f * Bx[p]; ky = ky + f * By[p]; }
                kz = kz + f * Bz[p];
            }
        }
        if (flag == 0) {
            mx = bx * kout + kx;
            my = by * kout + ky;
        }
        mz = bz * kout + kz + m0;
        if (ORC_MUT(5)) { if (flag == 0) { Bx[t] = mx; By[t] = my; } Bz[t] = mz; }
        if (C->wMx) {
            if (flag == 0) { C->wMx[si[t]] = mx; C->wMy[si[t]] = my; }
            C->wMz[si[t]] = mz;
        } else {
            if (flag == 0) { C->vobj->Mx[si[t]] = (float)mx; C->vobj->My[si[t]] = (float)my; }
            C->vobj->Mz[si[t]] = (float)mz;
        }
    }
}

/* One split step of the MT and CEST kernels for one (voxel, spin): water
 * (pool 0) and the bound pool exchange longitudinally through a combined
 * exponent; CEST pools (t >= 2) exchange all components with water.
 * BlochKernelMT.cuh:238-414, BlochKernelCEST.cuh:271-610. */
static void orc_wc_split_voxel(orc_ctx *C, int64_t v, int s, real dt, real gz, real gy, real gx,
                               const double *rfAmp, const double *rfPhase, const double *rfFreq,
                               gm_cache *gcs)
{
    const int Ty = C->vobj->TypeNum, S = C->vobj->SpinNum;
    const int64_t nvox = C->nvox;
    if (orc_gm_skip(C, v, 0)) return;
    real Bx[GM_MAXT], By[GM_MAXT], Bz[GM_MAXT];
    int64_t si[GM_MAXT];
    for (int t = 0; t < Ty; t++) {
        si[t] = v + nvox * ((int64_t)s + (int64_t)S * t);
        Bx[t] = C->wMx ? C->wMx[si[t]] : (real)C->vobj->Mx[si[t]];
        By[t] = C->wMy ? C->wMy[si[t]] : (real)C->vobj->My[si[t]];
        Bz[t] = C->wMz ? C->wMz[si[t]] : (real)C->vobj->Mz[si[t]];
    }
    real ox[GM_MAXT], oy[GM_MAXT], oz[GM_MAXT];
    const real k12 = orc_rate(C, v, 0, 1), k21 = orc_rate(C, v, 1, 0);
    /* free pools: rotation */
    for (int t = 0; t < Ty; t++) {
        if (t == 1) continue;
        real mx = Bx[t], my = By[t], mz = Bz[t], bx = mx, by = my, bz = mz;
        real dW = orc_gm_dW(C, v, si[t], t, gz, gy, gx);
        real b1, b2;
        if (orc_rf_field(C, v, rfAmp, rfPhase, rfFreq, &b1, &b2, &dW)) {
            orc_rf_rotate(dW, b1, b2, dt, mx, my, mz, &bx, &by, &bz);
        } else {
            real alpha = dW * dt, sA = r_sin(alpha), cA = r_cos(alpha);
            bx = mx * cA + my * sA;
            by = my * cA - mx * sA;
            bz = mz;
        }
        ox[t] = bx; oy[t] = by; oz[t] = bz;
    }
    /* water */
    {
        const int64_t tv = v;
        const real Rho = C->vobj->Rho[tv], T1 = C->vobj->T1[tv], T2 = C->vobj->T2[tv];
        const real e2 = r_exp(-dt / T2);
        const real e1k = r_exp(-dt / T1 - dt * k12);
        const real m0 = (Rho * (1 - r_exp(-dt / T1))) / (real)S;
        const real ek = 1 - r_exp(-dt * k21);
        real kout = 1, kx = 0, ky = 0, kz = 0;
        for (int t = 2; t < Ty; t++) {
            const real ko = orc_rate(C, v, 0, t);
            if (ko != 0) kout = kout * r_exp(-dt * ko);
            const real ki = orc_rate(C, v, t, 0);
            if (ki != 0) {
                const real f = 1 - r_exp(-dt * ki);
                kx = kx + f * Bx[t]; ky = ky + f * By[t]; kz = kz + f * Bz[t];
            }
        }
        real mx, my, mz;
        if (C->variant == ORC_MT) {
            mx = ox[0] * e2;
            my = oy[0] * e2;
            mz = oz[0] * e1k + m0 + Bz[1] * ek;
        } else {
            mx = ox[0] * e2 * kout + kx;
            my = oy[0] * e2 * kout + ky;
            mz = oz[0] * e1k * kout + kz + Bz[1] * ek + m0;
        }
        ox[0] = mx; oy[0] = my; oz[0] = mz;
    }
    /* CEST pools */
    for (int t = 2; t < Ty; t++) {
        const int64_t tv = v + (int64_t)t * nvox;
        const real Rho = C->vobj->Rho[tv], T1 = C->vobj->T1[tv], T2 = C->vobj->T2[tv];
        const real e2 = r_exp(-dt / T2), e1 = r_exp(-dt / T1);
        const real m0 = (Rho * (1 - e1)) / (real)S;
        real bx = ox[t] * e2, by = oy[t] * e2, bz = oz[t] * e1;
        real kout = 1, kx = 0, ky = 0, kz = 0;
        const real ko = orc_rate(C, v, t, 0);
        if (ko != 0) kout = r_exp(-dt * ko);
        const real ki = orc_rate(C, v, 0, t);
        if (ki != 0) {
            const real f = 1 - r_exp(-dt * ki);
            kx = f * Bx[0]; ky = f * By[0]; kz = f * Bz[0];
        }
        ox[t] = bx * kout + kx;
        oy[t] = by * kout + ky;
        oz[t] = bz * kout + kz + m0;
    }
    /* bound pool */
    {
        const int64_t tv = v + nvox;
        const real Rho = C->vobj->Rho[tv], T1 = C->vobj->T1[tv], T2 = C->vobj->T2[tv];
        real rfAmpSum = 0;
        for (int i = 0; i < C->vcoi->TxCoilNum; i++) rfAmpSum += (real)fabs((double)(real)rfAmp[i]);
        real bz = Bz[1];
        if (rfAmpSum != 0) {
            real W = (real)orc_gm_W(C, v, 1, T2, rfAmp, rfFreq, &gcs[s * Ty + 1]);
            bz = Bz[1] * r_exp(-W * dt);
        }
        const real e1k = r_exp(-dt / T1 - dt * k21);
        const real m0 = (Rho * (1 - r_exp(-dt / T1))) / (real)S;
        const real ek = 1 - r_exp(-dt * k12);
        oz[1] = bz * e1k + m0 + (ORC_MUT(6) ? oz[0] : Bz[0]) * ek;
    }
    for (int t = 0; t < Ty; t++) {
        if (C->wMx) {
            if (t != 1) { C->wMx[si[t]] = ox[t]; C->wMy[si[t]] = oy[t]; }
            C->wMz[si[t]] = oz[t];
        } else {
            if (t != 1) { C->vobj->Mx[si[t]] = (float)ox[t]; C->vobj->My[si[t]] = (float)oy[t]; }
            C->vobj->Mz[si[t]] = (float)oz[t];
        }
    }
}

/* x <- expm(A*dt) x for an n x n generator (row-major), in fp64: Taylor
 * series of the action on x with substeps of h*|A|_1 <= 0.5. */
static void orc_expm_action(int n, const double *A, double dt, double *x)
{
    double nrm = 0;
    for (int j = 0; j < n; j++) {
        double cs = 0;
        for (int i = 0; i < n; i++) cs += fabs(A[i * n + j]);
        if (cs > nrm) nrm = cs;
    }
    int sub = (int)ceil(nrm * dt / 0.5);
    if (sub < 1) sub = 1;
    const double h = dt / sub;
    double term[4 * GM_MAXT], nt[4 * GM_MAXT], y[4 * GM_MAXT];
    for (int st = 0; st < sub; st++) {
        for (int i = 0; i < n; i++) { term[i] = x[i]; y[i] = x[i]; }
        for (int k = 1; k <= 40; k++) {
            double tn = 0, yn = 0;
            for (int i = 0; i < n; i++) {
                double acc = 0;
                for (int j = 0; j < n; j++) acc += A[i * n + j] * term[j];
                nt[i] = acc * h / k;
            }
            for (int i = 0; i < n; i++) { term[i] = nt[i]; y[i] += nt[i]; tn += fabs(nt[i]); yn += fabs(y[i]); }
            if (tn <= 1e-18 * (yn + 1e-300)) break;
        }
        for (int i = 0; i < n; i++) x[i] = y[i];
    }
}

/* One step of all pools of one (voxel, spin) as the exact solution of the
 * linear Bloch-McConnell system the split scheme approximates: state
 * (Mx,My,Mz per pool, 1), generator = rotation M x Omega_t, relaxation,
 * saturation W of bound pools, exchange k_{t->p} = K[p*Ty+t]. */
static void orc_gm_expm_voxel(orc_ctx *C, int64_t v, int s, double dt, double gz, double gy, double gx,
                              const double *rfAmp, const double *rfPhase, const double *rfFreq,
                              gm_cache *gcs)
{
    const int Ty = C->vobj->TypeNum, S = C->vobj->SpinNum;
    const int64_t nvox = C->nvox;
    const int n = 3 * Ty + 1;
    double A[(3 * GM_MAXT + 1) * (3 * GM_MAXT + 1)];
    double x[3 * GM_MAXT + 1];
    int64_t si[GM_MAXT];
    memset(A, 0, sizeof(double) * n * n);
    const double gyro = (double)(real)C->vobj->Gyro;
    for (int t = 0; t < Ty; t++) {
        si[t] = v + nvox * ((int64_t)s + (int64_t)S * t);
        x[3 * t] = C->wMx ? (double)C->wMx[si[t]] : C->vobj->Mx[si[t]];
        x[3 * t + 1] = C->wMy ? (double)C->wMy[si[t]] : C->vobj->My[si[t]];
        x[3 * t + 2] = C->wMz ? (double)C->wMz[si[t]] : C->vobj->Mz[si[t]];
    }
    x[n - 1] = 1.0;
    /* rf field (same for all free pools) */
    const int Tx = C->vcoi->TxCoilNum;
    double rfsum = 0;
    for (int i = 0; i < Tx; i++) rfsum += fabs(rfAmp[i]);
    double bxr = 0, byr = 0, wrf = 0;
    if (rfsum != 0) {
        for (int i = 0; i < Tx; i++) {
            double Aa = rfAmp[i], P = rfPhase[i];
            if (Aa == 0) continue;
            if (Aa < 0) { Aa = -Aa; P += ORC_PI; }
            double m = (double)C->vcoi->TxCoilmg[v + (int64_t)i * nvox] * Aa * gyro;
            double ph = (double)C->vcoi->TxCoilpe[v + (int64_t)i * nvox] + P;
            bxr += m * cos(ph); byr += m * sin(ph);
            wrf += 2 * ORC_PI * rfFreq[i];
        }
    }
    /* MT and CEST: water and the bound pool exchange longitudinally only */
    const int wc = C->variant == ORC_MT || C->variant == ORC_CEST;
    for (int t = 0; t < Ty; t++) {
        if (orc_vskip(C, v, t)) continue;
        const int64_t tv = v + (int64_t)t * nvox;
        const double Rho = C->vobj->Rho[tv], R1 = 1.0 / C->vobj->T1[tv], R2 = 1.0 / C->vobj->T2[tv];
        const int flag = orc_pool_flag(C, t);
        double kout = 0, koutx = 0;
        for (int p = 0; p < Ty; p++) {
            if (p == t) continue;
            const double r = orc_rate(C, v, t, p);
            kout += r;
            if (!(wc && orc_pool_flag(C, p) == 1)) koutx += r;
        }
        const int z = 3 * t + 2;
        if (flag == 0) {
            const int xi = 3 * t, yi = 3 * t + 1;
            double w = (double)C->vmag->dB0[v] * gyro + (double)C->vmag->dWRnd[si[t]] +
                       2 * ORC_PI * C->vctl->CS[t] + gyro * ((double)C->vmag->Gzgrid[v] * gz +
                       (double)C->vmag->Gygrid[v] * gy + (double)C->vmag->Gxgrid[v] * gx);
            if (rfsum != 0) w += wrf;
            /* dM/dt = M x Omega, Omega = (bx, by, w) */
            A[xi * n + yi] += w;   A[xi * n + z] -= byr;
            A[yi * n + z] += bxr;  A[yi * n + xi] -= w;
            A[z * n + xi] += byr;  A[z * n + yi] -= bxr;
            A[xi * n + xi] -= R2 + koutx;
            A[yi * n + yi] -= R2 + koutx;
            for (int p = 0; p < Ty; p++) {
                if (p == t || (wc && orc_pool_flag(C, p) == 1)) continue;
                double ki = orc_rate(C, v, p, t);
                A[xi * n + 3 * p] += ki;
                A[yi * n + 3 * p + 1] += ki;
            }
        } else if (flag == 1) {
            double W = 0;
            if (rfsum != 0) W = orc_gm_W(C, v, t, C->vobj->T2[tv], rfAmp, rfFreq, &gcs[s * Ty + t]);
            A[z * n + z] -= W;
        } else {
            continue;
        }
        A[z * n + z] -= R1 + kout;
        A[z * n + (n - 1)] += R1 * Rho / S;
        for (int p = 0; p < Ty; p++) {
            if (p == t) continue;
            A[z * n + 3 * p + 2] += orc_rate(C, v, p, t);
        }
    }
    orc_expm_action(n, A, dt, x);
    for (int t = 0; t < Ty; t++) {
        if (C->wMx) { C->wMx[si[t]] = x[3 * t]; C->wMy[si[t]] = x[3 * t + 1]; C->wMz[si[t]] = x[3 * t + 2]; }
        else { C->vobj->Mx[si[t]] = (float)x[3 * t]; C->vobj->My[si[t]] = (float)x[3 * t + 1]; C->vobj->Mz[si[t]] = (float)x[3 * t + 2]; }
    }
}

/* Ext plugins executed natively: Src/Main/DoExtPlugin.m:11-33 */
static int orc_ext(orc_ctx *c, double ext)
{
    mrs_vvar *V = c->vvar;
    int64_t ntot = c->nvox * c->vobj->SpinNum * c->vobj->TypeNum;
    switch ((int)ext) {
    case 1: V->Kz = 0; V->Ky = 0; V->Kx = 0; return 0;              /* Plugin_ResetK.m:7-9 */
    case 2: V->Kz = -V->Kz; V->Ky = -V->Ky; V->Kx = -V->Kx; return 0; /* Plugin_ReverseK.m */
    case 3: V->KzKeeper = V->Kz; V->KyKeeper = V->Ky; V->KxKeeper = V->Kx; return 0; /* LockK */
    case 4:                                                       /* Plugin_ReleaseK.m */
        V->Kz = V->KzKeeper; V->Ky = V->KyKeeper;
        if (!ORC_MUT(7)) V->Kx = V->KxKeeper;
        return 0;
    case 6:                                                       /* Plugin_IdealSpoiler.m:10-11 */
        for (int64_t i = 0; i < ntot; i++) {
            if (c->wMx) { c->wMx[i] = 0; c->wMy[i] = 0; }
            c->vobj->Mx[i] = 0; c->vobj->My[i] = 0;
        }
        return 0;
    case 7:                                                       /* Plugin_rfRef.m:7-12 */
        if (c->vctl->MultiTransmit == 0) V->rfRef = V->rfPhase[0];
        else V->rfRef = V->rfPhase[c->vctl->MasterTxCoil - 1];
        return 0;
    default:
        if (c->cb && c->cb->ext) {
            if (c->wMx)
                for (int64_t i = 0; i < ntot; i++) {
                    c->vobj->Mx[i] = (float)c->wMx[i]; c->vobj->My[i] = (float)c->wMy[i];
                    c->vobj->Mz[i] = (float)c->wMz[i];
                }
            /* the CPU engine's plugins share its arrays: VObj holds the
             * current state, VSig the samples so far, and the arrays are
             * read again through the (possibly re-pointed) structs after
             * the call (DoScanAtCPU.cpp:360-405) */
            mrs_ext_state st;
            memset(&st, 0, sizeof st);
            st.vobj = c->vobj; st.vmag = c->vmag; st.vcoi = c->vcoi; st.vsig = c->vsig;
            st.fetched = 1;
            int rc = c->cb->ext(c->cb->user, ext, V, &st);
            if (c->wMx)
                for (int64_t i = 0; i < ntot; i++) {
                    c->wMx[i] = c->vobj->Mx[i]; c->wMy[i] = c->vobj->My[i]; c->wMz[i] = c->vobj->Mz[i];
                }
            return rc;
        }
        return 0;
    }
}

static void orc_capture(orc_ctx *c)
{
    /* RunMode 1 copy of the full state into VSig.M*[utsi], DoScanAtCPU.cpp:433-435,546-548 */
    int64_t ntot = c->nvox * c->vobj->SpinNum * c->vobj->TypeNum;
    int64_t off = (int64_t)c->vvar->utsi * ntot;
    for (int64_t i = 0; i < ntot; i++) {
        c->vsig->Mx[off + i] = c->wMx ? (float)c->wMx[i] : c->vobj->Mx[i];
        c->vsig->My[off + i] = c->wMy ? (float)c->wMy[i] : c->vobj->My[i];
        c->vsig->Mz[off + i] = c->wMz ? (float)c->wMz[i] : c->vobj->Mz[i];
    }
}

/*
 * The scan: DoScanAtCPU.cpp:246-553 with the MEX globals replaced by the
 * ABI structs. Returns 0 on success, MRS_ERR_CALLBACK if a callback failed.
 */
static int orc_scan(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                    const mrs_vctl *vctl, const mrs_vseq *vseq, mrs_vvar *vvar,
                    mrs_vsig *vsig, const mrs_callbacks *cb, int model, int variant)
{
    orc_ctx C;
    memset(&C, 0, sizeof C);
    C.variant = variant;
    /* Ext callbacks may re-point VMag and VCoi arrays: the scan reads them
     * through these copies */
    mrs_vmag vmag_cur = *vmag;
    mrs_vcoi vcoi_cur = *vcoi;
    C.vobj = vobj; C.vmag = &vmag_cur; C.vcoi = &vcoi_cur; C.vctl = vctl; C.vseq = vseq;
    C.vvar = vvar; C.vsig = vsig; C.cb = cb;
    const int64_t SpinMxNum = (int64_t)vobj->dims[0] * vobj->dims[1];
    const int64_t SpinMxSliceNum = vobj->dims[2] > 0 ? vobj->dims[2] : 1;
    C.nxy = SpinMxNum; C.nz = SpinMxSliceNum; C.nvox = SpinMxNum * SpinMxSliceNum;
    const int S = vobj->SpinNum, Ty = vobj->TypeNum, Tx = vcoi->TxCoilNum, Rx = vcoi->RxCoilNum;
    const int64_t ntot = C.nvox * S * Ty;
    const int64_t SignalNum = vsig->SignalNum;
    const real Gyro = (real)vobj->Gyro;

    if (sizeof(real) == 8) {       /* fp64 state */
        C.wMx = (real *)malloc(sizeof(real) * ntot);
        C.wMy = (real *)malloc(sizeof(real) * ntot);
        C.wMz = (real *)malloc(sizeof(real) * ntot);
        for (int64_t i = 0; i < ntot; i++) { C.wMx[i] = vobj->Mx[i]; C.wMy[i] = vobj->My[i]; C.wMz[i] = vobj->Mz[i]; }
    }
    real *buf1 = (real *)malloc(sizeof(real) * C.nvox);
    real *buf3 = (real *)malloc(sizeof(real) * C.nvox);
    gm_cache *gcs = model ? (gm_cache *)calloc((size_t)C.nvox * S * Ty, sizeof(gm_cache)) : NULL;

    int nthreads = g_threads > 0 ? g_threads
                 : (SpinMxSliceNum < vctl->MaxThreadNum ? (int)SpinMxSliceNum : vctl->MaxThreadNum); /* :225-228 */
    if (nthreads < 1) nthreads = 1;

    const double *tsLine = vseq->tsLine, *flagsLine = vseq->flagsLine, *utsLine = vseq->utsLine;
    const int64_t MaxStep = vseq->n_ts, MaxutsStep = vseq->n_uts, MaxrfStep = vseq->n_rf;
    const int64_t MaxGzStep = vseq->n_Gz, MaxGyStep = vseq->n_Gy, MaxGxStep = vseq->n_Gx;
    int64_t i = 0, Signali = 0;
    double flag[6];
    int status = 0;

    while (i < MaxStep - 1) {                                        /* :246 */
        for (int k = 0; k < 6; k++) flag[k] = 0;
        if (tsLine[i] != tsLine[i + 1]) {                           /* :254-284 */
            for (int k = 0; k < 6; k++) flag[k] += flagsLine[i * 6 + k];
            i++;
        } else {
            for (int k = 0; k < 6; k++) flag[k] += flagsLine[i * 6 + k];
            while (tsLine[i] == tsLine[i + 1]) {
                for (int k = 0; k < 6; k++) flag[k] += flagsLine[(i + 1) * 6 + k];
                i++;
                if (i == MaxStep - 1) break;
            }
            i++;
        }
        /* :287-289 */
        vvar->t = utsLine[vvar->utsi];
        int64_t un = vvar->utsi + 1 < MaxutsStep - 1 ? vvar->utsi + 1 : MaxutsStep - 1;
        vvar->dt = utsLine[un] - utsLine[vvar->utsi];
        vvar->utsi = (int32_t)un;

        if (flag[0] >= 1) {                                          /* :291-320 */
            for (int64_t j = 0; j < flag[0]; j++) {
                vvar->rfCoil = vseq->rfCoilLine[vvar->rfi];
                int TxCoili = (int)vvar->rfCoil;
                int64_t s = vvar->rfi + 1;
                while (s < MaxrfStep) {
                    if (vvar->rfCoil == vseq->rfCoilLine[s]) {
                        const double *A = vseq->rfAmpLine, *P = vseq->rfPhaseLine, *F = vseq->rfFreqLine;
                        vvar->rfAmp[TxCoili - 1] = fabs(A[vvar->rfi]) <= fabs(A[s]) ? A[vvar->rfi] : A[s];
                        vvar->rfPhase[TxCoili - 1] = fabs(P[vvar->rfi]) <= fabs(P[s]) ? P[vvar->rfi] : P[s];
                        vvar->rfFreq[TxCoili - 1] = fabs(F[vvar->rfi]) <= fabs(F[s]) ? F[vvar->rfi] : F[s];
                        break;
                    }
                    s++;
                }
                vvar->rfi++;
            }
        }
        if (flag[1] == 1) {                                          /* :322-329 */
            const double *L = vseq->GzAmpLine; int64_t n1 = vvar->Gzi + 1 < MaxGzStep - 1 ? vvar->Gzi + 1 : MaxGzStep - 1;
            vvar->GzAmp = fabs(L[vvar->Gzi]) <= fabs(L[n1]) ? L[vvar->Gzi] : L[vvar->Gzi + 1];
            vvar->Gzi++;
        }
        if (flag[2] == 1) {                                          /* :331-338 */
            const double *L = vseq->GyAmpLine; int64_t n1 = vvar->Gyi + 1 < MaxGyStep - 1 ? vvar->Gyi + 1 : MaxGyStep - 1;
            vvar->GyAmp = fabs(L[vvar->Gyi]) <= fabs(L[n1]) ? L[vvar->Gyi] : L[vvar->Gyi + 1];
            vvar->Gyi++;
        }
        if (flag[3] == 1) {                                          /* :340-347 */
            const double *L = vseq->GxAmpLine; int64_t n1 = vvar->Gxi + 1 < MaxGxStep - 1 ? vvar->Gxi + 1 : MaxGxStep - 1;
            vvar->GxAmp = fabs(L[vvar->Gxi]) <= fabs(L[n1]) ? L[vvar->Gxi] : L[vvar->Gxi + 1];
            vvar->Gxi++;
        }
        vvar->ADC = 0;                                               /* :349-354 */
        if (flag[4] == 1) { vvar->ADC = vseq->ADCLine[vvar->ADCi]; vvar->ADCi++; }
        if (flag[5] == 1) {                                          /* :356-409 */
            vvar->Ext = vseq->ExtLine[vvar->Exti];
            if (vvar->Ext != 0) {
                if (orc_ext(&C, vvar->Ext)) { status = MRS_ERR_CALLBACK; goto done; }
            }
            vvar->Exti++;
        }
        if (flag[0] + flag[1] + flag[2] + flag[3] + flag[4] + flag[5] == 0) { /* :411-420 */
            for (int k = 0; k < Tx; k++) { vvar->rfAmp[k] = 0; vvar->rfPhase[k] = 0; vvar->rfFreq[k] = 0; }
            vvar->GzAmp = 0; vvar->GyAmp = 0; vvar->GxAmp = 0; vvar->ADC = 0; vvar->Ext = 0;
        }
        if (vvar->dt == 0) continue;                                 /* :423-425 */
        if (vvar->dt < 0) {                                          /* :426-440 */
            vvar->TRCount++;
            if (vctl->RunMode == 1) {
                orc_capture(&C);
                vsig->Muts[vvar->utsi] = vsig->Muts[vvar->utsi - 1];
            }
            continue;
        }
        for (int Typei = 0; Typei < Ty; Typei++) {
            for (int Spini = 0; Spini < S; Spini++) {
                const int64_t base = (int64_t)Typei * S * C.nvox + (int64_t)Spini * C.nvox;
                /* GM samples free pools only (BlochKernelGM.cuh:192), ME every pool
                 * (BlochKernelME.cuh:186-220), MT and CEST the water pool
                 * (BlochKernelMT.cuh:230, BlochKernelCEST.cuh:255) */
                const int sample = vvar->ADC == 1 &&
                    (model == 0 || (variant == ORC_MT || variant == ORC_CEST ? Typei == 0
                                                                            : orc_pool_flag(&C, Typei) == 0));
                if (sample) {                                        /* :445-491 */
                    for (int c = 0; c < Rx; c++) {
                        const float *Cx = C.vcoi->RxCoilx + (int64_t)c * C.nvox, *Cy = C.vcoi->RxCoily + (int64_t)c * C.nvox;
                        const real cr = (real)cos(-vvar->rfRef), msr = (real)-sin(-vvar->rfRef);
                        const real sr = (real)sin(-vvar->rfRef);
                        for (int64_t v = 0; v < C.nvox; v++) {
                            real mx = C.wMx ? C.wMx[base + v] : (real)vobj->Mx[base + v];
                            real my = C.wMy ? C.wMy[base + v] : (real)vobj->My[base + v];
                            real b1, b2, b3, b4;
                            if (vcoi->RxCoilDefault == 0) {
                                b1 = mx * (real)Cx[v]; b2 = my * (real)Cy[v]; b3 = b1 + b2;
                                b1 = mx * (real)Cy[v]; b1 = b1 * -1; b2 = my * (real)Cx[v]; b4 = b1 + b2;
                                b1 = b3; b2 = b4;
                            } else {
                                b1 = mx; b2 = my; b3 = mx; b4 = my;
                            }
                            if (vvar->rfRef != 0) {
                                b1 = b1 * cr; b2 = b2 * msr; b1 = b2 + b1;
                                b3 = b3 * sr; b4 = b4 * cr; b3 = b4 + b3;
                            } else {
                                b3 = b4;
                            }
                            if (model && orc_vskip(&C, v, Typei)) { b1 = 0; b3 = 0; }
                            buf1[v] = b1; buf3[v] = b3;
                        }
                        vsig->Sx[(int64_t)Typei * Rx * SignalNum + (int64_t)c * SignalNum + Signali] += orc_sum(buf1, C.nvox);
                        vsig->Sy[(int64_t)Typei * Rx * SignalNum + (int64_t)c * SignalNum + Signali] += orc_sum(buf3, C.nvox);
                    }
                }
                if (model) continue;  /* multi-pool steps run per voxel below */
                /* Bloch step over slices, OpenMP as :494-527 */
                const real rdt = (real)vvar->dt;
                const real gz = (real)vvar->GzAmp, gy = (real)vvar->GyAmp, gx = (real)vvar->GxAmp;
                int64_t Slicei;
#pragma omp parallel for num_threads(nthreads) schedule(static)
                for (Slicei = 0; Slicei < SpinMxSliceNum; Slicei++) {
                    const int64_t so = Slicei * SpinMxNum;
                    const int64_t po = (int64_t)Typei * C.nvox + so;
                    orc_bloch_slice(Gyro, vctl->CS, S,
                                    vobj->Rho + po, vobj->T1 + po, vobj->T2 + po,
                                    vobj->Mz + base + so, vobj->My + base + so, vobj->Mx + base + so,
                                    C.vmag->dB0 + so, C.vmag->dWRnd + base + so,
                                    C.vmag->Gzgrid + so, C.vmag->Gygrid + so, C.vmag->Gxgrid + so,
                                    C.vcoi->TxCoilmg + so, C.vcoi->TxCoilpe + so,
                                    rdt, vvar->rfAmp, vvar->rfPhase, vvar->rfFreq, gz, gy, gx,
                                    Typei, SpinMxNum, SpinMxSliceNum, Tx,
                                    C.wMx ? C.wMx + base + so : NULL,
                                    C.wMy ? C.wMy + base + so : NULL,
                                    C.wMz ? C.wMz + base + so : NULL);
                }
            }
        }
        if (model) {
            const real rdt = (real)vvar->dt;
            const real gz = (real)vvar->GzAmp, gy = (real)vvar->GyAmp, gx = (real)vvar->GxAmp;
            int64_t v;
#pragma omp parallel for num_threads(g_threads > 0 ? g_threads : 1) schedule(static)
            for (v = 0; v < C.nvox; v++) {
                for (int sp = 0; sp < S; sp++) {
                    gm_cache *gc = gcs + (size_t)v * S * Ty;
                    if (model == 1 && (variant == ORC_MT || variant == ORC_CEST))
                        orc_wc_split_voxel(&C, v, sp, rdt, gz, gy, gx, vvar->rfAmp, vvar->rfPhase, vvar->rfFreq, gc);
                    else if (model == 1)
                        orc_gm_split_voxel(&C, v, sp, rdt, gz, gy, gx, vvar->rfAmp, vvar->rfPhase, vvar->rfFreq, gc);
                    else
                        orc_gm_expm_voxel(&C, v, sp, vvar->dt, vvar->GzAmp, vvar->GyAmp, vvar->GxAmp,
                                          vvar->rfAmp, vvar->rfPhase, vvar->rfFreq, gc);
                }
            }
        }
        if (vvar->ADC == 1) {                                        /* :531-537 */
            vsig->Kz[Signali] += vvar->Kz;
            vsig->Ky[Signali] += vvar->Ky;
            vsig->Kx[Signali] += vvar->Kx;
            Signali++;
        }
        vvar->Kz += vvar->GzAmp * vvar->dt * (vobj->Gyro / (2 * ORC_PI)); /* :540-542 */
        vvar->Ky += vvar->GyAmp * vvar->dt * (vobj->Gyro / (2 * ORC_PI));
        vvar->Kx += vvar->GxAmp * vvar->dt * (vobj->Gyro / (2 * ORC_PI));
        if (vctl->RunMode == 1) {                                    /* :544-552 */
            orc_capture(&C);
            vsig->Muts[vvar->utsi] = vsig->Muts[vvar->utsi - 1] + vvar->dt;
        }
    }
done:
    if (C.wMx) {
        for (int64_t q = 0; q < ntot; q++) {
            vobj->Mx[q] = (float)C.wMx[q]; vobj->My[q] = (float)C.wMy[q]; vobj->Mz[q] = (float)C.wMz[q];
        }
        free(C.wMx); free(C.wMy); free(C.wMz);
    }
    free(buf1); free(buf3); free(gcs);
    return status;
}

/* The single-pool scan of DoScanAtCPU.cpp:70. */
int orc_DoScanAtCPU(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                    const mrs_vctl *vctl, const mrs_vseq *vseq, mrs_vvar *vvar,
                    mrs_vsig *vsig, const mrs_callbacks *cb)
{
    return orc_scan(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, 0, 0);
}

/* The multi-pool scans (same decoder): variant ORC_GM (DoGMScanAtGPU.cu),
 * ORC_MT (DoMTScanAtGPU.cu), ORC_CEST (DoCESTScanAtGPU.cu) or ORC_ME
 * (DoMEScanAtGPU.cu); scheme 0 is the reference kernel's operator split,
 * scheme 1 the fp64 matrix exponential of the same generator. */
int orc_DoGMScan(mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                 const mrs_vctl *vctl, const mrs_vseq *vseq, mrs_vvar *vvar,
                 mrs_vsig *vsig, const mrs_callbacks *cb, int scheme, int variant)
{
    if (!vobj->K || vobj->TypeNum > GM_MAXT || vobj->TypeNum < 2) return MRS_ERR_ARG;
    if (variant == ORC_GM && !vobj->TypeFlag) return MRS_ERR_ARG;
    if (variant == ORC_MT && vobj->TypeNum != 2) return MRS_ERR_ARG;
    if (variant == ORC_CEST && vobj->TypeNum < 3) return MRS_ERR_ARG;
    return orc_scan(vobj, vmag, vcoi, vctl, vseq, vvar, vsig, cb, scheme ? 2 : 1, variant);
}
