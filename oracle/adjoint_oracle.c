/*
 * adjoint_oracle.c -- CPU oracle of the excitation problem and its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/ may load this code, as the checker
 * of the HIP adjoint; the product path never does.
 *
 * MRiLab has no gradient code; this is the discrete forward model the HIP
 * adjoint differentiates, stated plainly, and its reverse-mode derivative.
 * The forward step is the engine's rf step (the rotation of
 * BlochKernel.h:146-306 in Rodrigues form, then the relaxation of
 * BlochKernel.h:348-368): per isochromat and step k, with Omega =
 * (bx, by, w), bx + i*by = gamma * mg * exp(i*pe) * rf_k,
 * w = gamma*dB0 + dWRnd + gamma*(G_k . r):
 *     M' = R(-Omega*dt) M,  M_{k+1} = E M' + (0, 0, m0 (1 - e1)),
 *     E = diag(e2, e2, e1), e = exp(-dt/T), m0 = Rho / SpinNum.
 * Loss: L = sum_v |sum_s (Mx + i My)_s,N - target_v|^2 over voxels v.
 * Reverse mode: lambda_N = dL/dM_N, lambda' = E lambda_{k+1},
 *     dL/dT2 += (lx' M'x + ly' M'y)/e2 * e2 * dt/T2^2 (per unit of lambda'),
 *     dL/dT1 += lz (M'z - m0) e1 dt / T1^2,
 *     dL/dOmega = dt (u - a psi x u + b psi x (psi x u)), u = lambda' x M',
 *         psi = -Omega dt, theta = |psi|, a = (1-cos)/theta^2,
 *         b = (theta - sin)/theta^3 (the transposed left Jacobian of SO(3)),
 *     lambda_k = R^T lambda'.
 * Parity of this file is pinned by central finite differences of its own
 * fp64 forward (tests/test_oracle_adjoint.py), not by a reference.
 * Built in the same two precisions as bloch_oracle.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

typedef struct orc_exc {
    int32_t nvox, S, n;
    double dt, gyro;
    const float *x, *y, *z;       /* [nvox] positions (m) */
    const float *mg, *pe;         /* [nvox] single-Tx B1+ map */
    const float *dWRnd;           /* [S][nvox] */
    const float *rho;             /* [nvox] */
    const float *M0x, *M0y, *M0z; /* [S][nvox] initial state */
    const float *grad;            /* [n][3] Gx, Gy, Gz (T/m) */
    const float *target;          /* [nvox][2] */
} orc_exc;

static real a_exp(real x) { return (real)exp((double)x); }
static real a_sqrt(real x) { return sizeof(real) == 4 ? (real)sqrtf((float)x) : (real)sqrt((double)x); }
static void a_sincos(real x, real *s, real *c)
{
    /* correctly rounded from fp64 in the fp32 build, as the engine's sincos */
    *s = (real)sin((double)x);
    *c = (real)cos((double)x);
}

typedef struct {
    real gam, wr, wi, x, y, z, wb, e1, e2, m0, t1, t2;
} iso_const;

static void iso_setup(const orc_exc *P, int64_t v, int s, const double *T1, const double *T2,
                      const double *dB0, iso_const *k)
{
    const real gam = (real)(float)P->gyro;
    real sp, cp;
    a_sincos((real)P->pe[v], &sp, &cp);
    k->gam = gam;
    k->wr = (real)P->mg[v] * cp;
    k->wi = (real)P->mg[v] * sp;
    k->x = P->x[v]; k->y = P->y[v]; k->z = P->z[v];
    k->wb = (real)dB0[v] * gam + (real)P->dWRnd[v + (int64_t)P->nvox * s];
    k->t1 = (real)T1[v];
    k->t2 = (real)T2[v];
    real a2 = ((real)1 / k->t2) * (real)P->dt; a2 = -a2; if (a2 < (real)-200) a2 = (real)-200;
    real a1 = ((real)1 / k->t1) * (real)P->dt; a1 = -a1; if (a1 < (real)-200) a1 = (real)-200;
    k->e2 = a_exp(a2);
    k->e1 = a_exp(a1);
    k->m0 = (real)P->rho[v] / (real)P->S;
}

/* Omega of step j for one isochromat */
static void omega(const orc_exc *P, const iso_const *k, const double *rf, int j, real *bx, real *by, real *w)
{
    const float *g = P->grad + 3 * (size_t)j;
    const real qz = sizeof(real) == 4 ? (real)((float)k->gam * g[2]) : k->gam * (real)g[2];
    const real qy = sizeof(real) == 4 ? (real)((float)k->gam * g[1]) : k->gam * (real)g[1];
    const real qx = sizeof(real) == 4 ? (real)((float)k->gam * g[0]) : k->gam * (real)g[0];
    const real gv = k->z * qz + k->y * qy + k->x * qx;
    const real br = k->gam * (real)rf[2 * j], bi = k->gam * (real)rf[2 * j + 1];
    *bx = k->wr * br - k->wi * bi;
    *by = k->wr * bi + k->wi * br;
    *w = k->wb + gv;
}

/* M <- R(-Omega dt) M (sgn = +1) or R(+Omega dt) M (sgn = -1) */
static void rotate(real bx, real by, real w, real dt, int sgn, real *m)
{
    const real om2 = bx * bx + by * by + w * w;
    if (!(om2 > 0)) return;
    const real om = a_sqrt(om2), inv = (real)1 / om;
    const real nx = bx * inv, ny = by * inv, nz = w * inv;
    real sn, cs;
    a_sincos(om * dt, &sn, &cs);
    if (sgn < 0) sn = -sn;
    const real ax = m[0], ay = m[1], az = m[2];
    const real kk = ((real)1 - cs) * (nx * ax + ny * ay + nz * az);
    m[0] = ax * cs - (ny * az - nz * ay) * sn + nx * kk;
    m[1] = ay * cs - (nz * ax - nx * az) * sn + ny * kk;
    m[2] = az * cs - (nx * ay - ny * ax) * sn + nz * kk;
}

static void step_fwd(const iso_const *k, real bx, real by, real w, real dt, real *m)
{
    rotate(bx, by, w, dt, +1, m);
    m[0] = m[0] * k->e2;
    m[1] = m[1] * k->e2;
    const real t = m[2] * k->e1;
    m[2] = t + k->m0 * ((real)1 - k->e1);
}

/* forward of one isochromat; states[j] = M_j (j = 0..n) when states != NULL */
static void iso_forward(const orc_exc *P, const iso_const *k, int64_t v, int s, const double *rf, real *m,
                        real *states)
{
    const int64_t i = v + (int64_t)P->nvox * s;
    m[0] = P->M0x[i]; m[1] = P->M0y[i]; m[2] = P->M0z[i];
    if (states) memcpy(states, m, 3 * sizeof(real));
    for (int j = 0; j < P->n; j++) {
        real bx, by, w;
        omega(P, k, rf, j, &bx, &by, &w);
        step_fwd(k, bx, by, w, (real)P->dt, m);
        if (states) memcpy(states + 3 * (size_t)(j + 1), m, 3 * sizeof(real));
    }
}

double orc_exc_loss(const orc_exc *P, const double *rf, const double *T1, const double *T2, const double *dB0,
                    double *mxy)
{
    double L = 0;
    for (int64_t v = 0; v < P->nvox; v++) {
        double sx = 0, sy = 0;
        for (int s = 0; s < P->S; s++) {
            iso_const k;
            real m[3];
            iso_setup(P, v, s, T1, T2, dB0, &k);
            iso_forward(P, &k, v, s, rf, m, NULL);
            sx += (double)m[0];
            sy += (double)m[1];
        }
        if (mxy) { mxy[2 * v] = sx; mxy[2 * v + 1] = sy; }
        const double rx = sx - P->target[2 * v], ry = sy - P->target[2 * v + 1];
        L += rx * rx + ry * ry;
    }
    return L;
}

static void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double orc_exc_grad(const orc_exc *P, const double *rf, const double *T1, const double *T2, const double *dB0,
                    double *g_rf, double *g_T1, double *g_T2, double *g_dB0)
{
    const int n = P->n;
    const double dt = P->dt;
    real *states = (real *)malloc(sizeof(real) * 3 * (size_t)(n + 1));
    memset(g_rf, 0, sizeof(double) * 2 * (size_t)n);
    double L = 0;
    for (int64_t v = 0; v < P->nvox; v++) {
        /* voxel residual from the forward pass */
        double sx = 0, sy = 0;
        for (int s = 0; s < P->S; s++) {
            iso_const k;
            real m[3];
            iso_setup(P, v, s, T1, T2, dB0, &k);
            iso_forward(P, &k, v, s, rf, m, NULL);
            sx += (double)m[0];
            sy += (double)m[1];
        }
        const double rx = sx - P->target[2 * v], ry = sy - P->target[2 * v + 1];
        L += rx * rx + ry * ry;
        double gt1 = 0, gt2 = 0, gb0 = 0;
        for (int s = 0; s < P->S; s++) {
            iso_const k;
            real m[3];
            iso_setup(P, v, s, T1, T2, dB0, &k);
            iso_forward(P, &k, v, s, rf, m, states);
            double lam[3] = {2 * rx, 2 * ry, 0};
            for (int j = n - 1; j >= 0; j--) {
                real bx, by, w;
                omega(P, &k, rf, j, &bx, &by, &w);
                real mp[3] = {states[3 * j], states[3 * j + 1], states[3 * j + 2]};
                rotate(bx, by, w, (real)dt, +1, mp); /* M' = R M_j */
                const double Mp[3] = {mp[0], mp[1], mp[2]};
                const double e1 = k.e1, e2 = k.e2, t1 = k.t1, t2 = k.t2;
                gt2 += (lam[0] * Mp[0] + lam[1] * Mp[1]) * e2 * dt / (t2 * t2);
                gt1 += lam[2] * (Mp[2] - (double)k.m0) * e1 * dt / (t1 * t1);
                double lp[3] = {lam[0] * e2, lam[1] * e2, lam[2] * e1};
                double u[3], c1[3], c2[3];
                cross(lp, Mp, u);
                const double psi[3] = {-(double)bx * dt, -(double)by * dt, -(double)w * dt};
                const double th = sqrt(psi[0] * psi[0] + psi[1] * psi[1] + psi[2] * psi[2]);
                double a, b;
                if (th < 1e-3) {
                    const double t2_ = th * th;
                    a = 0.5 - t2_ / 24 + t2_ * t2_ / 720;
                    b = 1.0 / 6 - t2_ / 120 + t2_ * t2_ / 5040;
                } else {
                    a = (1 - cos(th)) / (th * th);
                    b = (th - sin(th)) / (th * th * th);
                }
                cross(psi, u, c1);
                cross(psi, c1, c2);
                double gO[3];
                for (int q = 0; q < 3; q++) gO[q] = dt * (u[q] - a * c1[q] + b * c2[q]);
                const double gam = k.gam;
                gb0 += gam * gO[2];
                g_rf[2 * j] += gam * ((double)k.wr * gO[0] + (double)k.wi * gO[1]);
                g_rf[2 * j + 1] += gam * (-(double)k.wi * gO[0] + (double)k.wr * gO[1]);
                /* lambda_j = R^T lambda' */
                real lr[3] = {(real)lp[0], (real)lp[1], (real)lp[2]};
                if (sizeof(real) == 8) {
                    rotate(bx, by, w, (real)dt, -1, lr);
                    lam[0] = lr[0]; lam[1] = lr[1]; lam[2] = lr[2];
                } else {
                    /* keep the adjoint in fp64 even in the fp32 build */
                    const double om2 = (double)bx * bx + (double)by * by + (double)w * w;
                    if (om2 > 0) {
                        const double om = sqrt(om2), nx = bx / om, ny = by / om, nz = w / om;
                        const double sn = sin(om * dt), cs = cos(om * dt);
                        const double kk = (1 - cs) * (nx * lp[0] + ny * lp[1] + nz * lp[2]);
                        lam[0] = lp[0] * cs + (ny * lp[2] - nz * lp[1]) * sn + nx * kk;
                        lam[1] = lp[1] * cs + (nz * lp[0] - nx * lp[2]) * sn + ny * kk;
                        lam[2] = lp[2] * cs + (nx * lp[1] - ny * lp[0]) * sn + nz * kk;
                    } else {
                        lam[0] = lp[0]; lam[1] = lp[1]; lam[2] = lp[2];
                    }
                }
            }
        }
        g_T1[v] = gt1;
        g_T2[v] = gt2;
        g_dB0[v] = gb0;
    }
    free(states);
    return L;
}
