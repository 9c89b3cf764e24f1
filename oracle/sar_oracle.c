/* Local SAR, CPU restatement of MRiLab's SAR engine. Test infrastructure
 * only: the tests and the bench's CPU leg check mrs_DoCalSARAtGPU against
 * it; the product never calls it.
 *
 * orc_DoCalSAR follows Lib/src/engine/CPUEngine/DoCalSARAtCPU.cpp:191-355
 * (event loop, rf hold, SAR sampling before and after each step, TR
 * boundaries) and SARKernel.h:8-63 (per-voxel power accumulation) statement
 * by statement, with IPP's elementwise calls written as loops in the same
 * operation order and FP contraction off (the Makefile's -ffp-contract=off):
 *   E_k = sum_c (TxE1_k,c * (float)rfAmp_c)   over coils with rfAmp_c != 0
 *   SAR_k += (E_k * E_k) * dt                    k = x, y, z
 *   VSig.SAR(:, s) = ((VSig.SAR + SAR_x*ECon_x) + SAR_y*ECon_y) + SAR_z*ECon_z
 * The ORC_REAL = double build evaluates the same statements with double
 * temporaries (the accumulators stay the float arrays of VVar/VSig).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/mrilab_amd.h"

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

int orc_get_threads(void); /* bloch_oracle.c: threads set by orc_set_threads, 0 = one */

/* SARKernel.h:8-63 over all voxels (the reference splits them by slice
 * across OpenMP threads; each voxel is independent). */
static int sar_kernel(const mrs_sar_obj *o, float *base, int64_t nvox, real dt, const double *rfAmp) {
    const int Tx = o->TxCoilNum;
    float amp_sum = 0.f; /* SARKernel.h:13-19 */
    for (int i = 0; i < Tx; i++) amp_sum += fabsf((float)rfAmp[i]);
    if (amp_sum == 0.f) return 0;
    int64_t v;
    const int nth = orc_get_threads() > 0 ? orc_get_threads() : 1;
#pragma omp parallel for num_threads(nth) schedule(static)
    for (v = 0; v < nvox; v++) {
        real bx = 0, by = 0, bz = 0; /* ippsZero_32f :30-32 */
        for (int i = 0; i < Tx; i++) { /* :34-45 */
            const real a = (real)(float)rfAmp[i];
            if (a != 0) {
                bx = bx + (real)o->TxE1x[(int64_t)i * nvox + v] * a;
                by = by + (real)o->TxE1y[(int64_t)i * nvox + v] * a;
                bz = bz + (real)o->TxE1z[(int64_t)i * nvox + v] * a;
            }
        }
        bx = bx * bx; /* ippsSqr_32f_I :47-49 */
        by = by * by;
        bz = bz * bz;
        bx = bx * dt; /* ippsMulC_32f_I :51-53 */
        by = by * dt;
        bz = bz * dt;
        base[v] = (float)((real)base[v] + bx); /* ippsAdd_32f_I :55-57 */
        base[nvox + v] = (float)((real)base[nvox + v] + by);
        base[2 * nvox + v] = (float)((real)base[2 * nvox + v] + bz);
    }
    return 1;
}

/* DoCalSARAtCPU.cpp:223-238 / :329-353: store one SAR sample. */
static void sar_sample(const mrs_sar_obj *o, mrs_sar_sig *g, int64_t nvox, double mut) {
    float *dst = g->SAR + (int64_t)g->SARi * nvox;
    for (int k = 0; k < 3; k++) {
        const float *b = g->SARBase + (int64_t)k * nvox;
        const float *e = o->ECon + (int64_t)k * nvox;
        for (int64_t v = 0; v < nvox; v++) {
            const float t = b[v] * e[v]; /* ippsMul_32f into buffer1 */
            dst[v] = dst[v] + t;         /* ippsAdd_32f_I */
        }
    }
    g->tRealSample[g->SARi] = mut;
    g->SARi++;
}

/* Returns 0, or -1 on bad arguments. */
int orc_DoCalSAR(const mrs_sar_obj *o, const mrs_vseq *vs, mrs_vvar *vv, mrs_sar_sig *g) {
    if (!o || !vs || !vv || !g) return -1;
    const int64_t nvox = (int64_t)o->dims[0] * o->dims[1] * o->dims[2];
    const int64_t MaxStep = vs->n_ts, MaxutsStep = vs->n_uts, MaxrfStep = vs->n_rf;
    const int64_t MaxSARStep = g->n_samples;
    const int Tx = o->TxCoilNum;
    int64_t i = 0;
    double flag[6];
    g->n_rf_steps = 0;
    while (i < MaxStep - 1) { /* :191 */
        for (int f = 0; f < 6; f++) flag[f] = 0;
        if (vs->tsLine[i] != vs->tsLine[i + 1]) { /* :200-236: sum the flags of coincident events */
            for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[i * 6 + f];
            i++;
        } else {
            for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[i * 6 + f];
            while (vs->tsLine[i] == vs->tsLine[i + 1]) {
                for (int f = 0; f < 6; f++) flag[f] += vs->flagsLine[(i + 1) * 6 + f];
                i++;
                if (i == MaxStep - 1) break;
            }
            i++;
        }
        /* :239-255 sample before the step, no miss for the start point */
        if (g->SARi < MaxSARStep) {
            if (g->Muts[vv->utsi] >= g->tSample[g->SARi]) sar_sample(o, g, nvox, g->Muts[vv->utsi]);
            if (g->SARi == MaxSARStep) return 0;
        }
        /* :258-260 */
        vv->t = vs->utsLine[vv->utsi];
        {
            const int64_t nxt = vv->utsi + 1 < MaxutsStep - 1 ? vv->utsi + 1 : MaxutsStep - 1;
            vv->dt = vs->utsLine[nxt] - vs->utsLine[vv->utsi];
            vv->utsi = (int32_t)nxt;
        }
        if (flag[0] >= 1) { /* :262-291 rf lines: hold the smaller magnitude of this and the next same-coil sample */
            for (int64_t j = 0; j < flag[0]; j++) {
                vv->rfCoil = vs->rfCoilLine[vv->rfi];
                const int c = (int)vv->rfCoil;
                int64_t s = vv->rfi + 1;
                while (s < MaxrfStep) {
                    if (vv->rfCoil == vs->rfCoilLine[s]) {
                        vv->rfAmp[c - 1] = fabs(vs->rfAmpLine[vv->rfi]) <= fabs(vs->rfAmpLine[s])
                                               ? vs->rfAmpLine[vv->rfi] : vs->rfAmpLine[s];
                        vv->rfPhase[c - 1] = fabs(vs->rfPhaseLine[vv->rfi]) <= fabs(vs->rfPhaseLine[s])
                                                 ? vs->rfPhaseLine[vv->rfi] : vs->rfPhaseLine[s];
                        vv->rfFreq[c - 1] = fabs(vs->rfFreqLine[vv->rfi]) <= fabs(vs->rfFreqLine[s])
                                                ? vs->rfFreqLine[vv->rfi] : vs->rfFreqLine[s];
                        break;
                    }
                    s++;
                }
                vv->rfi++;
            }
        }
        if (flag[0] + flag[1] + flag[2] + flag[3] + flag[4] + flag[5] == 0) { /* :293-298 */
            memset(vv->rfAmp, 0, sizeof(double) * Tx);
            memset(vv->rfPhase, 0, sizeof(double) * Tx);
            memset(vv->rfFreq, 0, sizeof(double) * Tx);
            vv->Ext = 0;
        }
        if (vv->dt == 0) continue; /* :301-303 */
        if (vv->dt < 0) {          /* :304-313 TR boundary */
            vv->TRCount++;
            if (vv->utsi > 0) g->Muts[vv->utsi] = g->Muts[vv->utsi - 1];
            continue;
        }
        g->n_rf_steps += sar_kernel(o, g->SARBase, nvox, (real)(float)vv->dt, vv->rfAmp); /* :315-340 */
        if (vv->utsi > 0) g->Muts[vv->utsi] = g->Muts[vv->utsi - 1] + vv->dt; /* :342 */
        /* :344-354 sample after the step, no miss for the end point */
        if (g->SARi < MaxSARStep) {
            if (g->Muts[vv->utsi] >= g->tSample[g->SARi]) sar_sample(o, g, nvox, g->Muts[vv->utsi]);
            if (g->SARi == MaxSARStep) return 0;
        }
    }
    return 0;
}
