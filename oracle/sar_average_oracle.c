/* N-gram local SAR averaging, CPU restatement of MRiLab's
 * DoSARAverageAtCPU. Test infrastructure only: the tests and the bench's
 * CPU leg check mrs_DoSARAverageAtGPU against it; the product never calls
 * it.
 *
 * orc_DoSARAverage follows Lib/src/utilities/DoSARAverageAtCPU/
 * DoSARAverageAtCPU.cpp:69-127 statement by statement, in fp64 with FP
 * contraction off (the Makefile's -ffp-contract=off), so each sum is formed
 * in the reference's order:
 *   - per voxel with MassDen != 0, the radius r = 1, 2, ... grows until the
 *     mass of the voxel-index "sphere" S_r exceeds grams*1e-3 (kg), each
 *     S_r summed afresh in (o over slices, m over columns, n over rows)
 *     order, with rado = floor(sqrt(r^2 - o^2)) and
 *     radm = floor(sqrt(rado^2 - m^2)) (:81-111);
 *   - the voxel count and the SAR and Power sums of S_r only count when the
 *     centre voxel's SAR (first time point) is > 0 (:99-106);
 *   - the result interpolates between S_(r-1) and S_r to exactly
 *     grams*1e-3 of mass (:113-126); voxels it skips stay 0.
 * One departure: the reference loops forever when the whole volume weighs
 * no more than grams*1e-3; this restatement stops at the radius whose
 * sphere covers the volume and returns -1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_get_threads(void); /* bloch_oracle.c: threads set by orc_set_threads, 0 = one */

/* Voxels (j, i, s) of slices [s0, s1); dims = {rows, cols, slices, times}.
 * Returns 0, -1 (radius search without end) or -2 (bad arguments). */
int orc_DoSARAverage(const double *SAR, const double *Power, const double *MassDen, const int64_t dims[4],
                     double dx, double dy, double dz, double grams, double *aveSAR, double *avePower,
                     int64_t s0, int64_t s1) {
    const int64_t D0 = dims[0], D1 = dims[1], D2 = dims[2], T = dims[3];
    if (D0 <= 0 || D1 <= 0 || D2 <= 0 || T <= 0 || s0 < 0 || s1 > D2 || s0 > s1) return -2;
    const int64_t shift = D0 * D1 * D2;
    const int64_t rcap = D0 + D1 + D2; /* S_rcap covers the volume */
    const int nth = orc_get_threads() > 0 ? orc_get_threads() : 1;
    int err = 0;
    int64_t row;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 1) reduction(| : err)
    for (row = s0 * D1; row < s1 * D1; row++) {
        const int64_t s = row / D1, i = row % D1;
        double *totSAR = (double *)calloc((size_t)T * 4, sizeof(double));
        double *totSAR2 = totSAR + T, *totPower = totSAR + 2 * T, *totPower2 = totSAR + 3 * T;
        for (int64_t j = 0; j < D0; j++) {
            const int64_t c = j + i * D0 + s * D0 * D1;
            for (int64_t t = 0; t < T; t++) {
                aveSAR[t * shift + c] = 0.0;
                avePower[t * shift + c] = 0.0;
            }
            if (MassDen[c] == 0) continue;
            double totwei = 0, totwei2 = 0;
            int64_t totpx = 0, totpx2 = 0, radius = 0;
            for (int64_t t = 0; t < T; t++) totSAR[t] = totPower[t] = 0;
            while (totwei <= grams * 1e-3) {
                if (radius >= rcap) {
                    err = 1;
                    break;
                }
                totwei2 = totwei;
                memcpy(totSAR2, totSAR, (size_t)T * sizeof(double));
                memcpy(totPower2, totPower, (size_t)T * sizeof(double));
                totpx2 = totpx;
                totwei = 0;
                for (int64_t t = 0; t < T; t++) totSAR[t] = totPower[t] = 0;
                totpx = 0;
                radius++;
                for (int64_t o = -radius; o <= radius; o++) {
                    const int64_t rado = (int64_t)floor(sqrt((double)(radius * radius - o * o)));
                    for (int64_t m = -rado; m <= rado; m++) {
                        const int64_t radm = (int64_t)floor(sqrt((double)(rado * rado - m * m)));
                        for (int64_t n = -radm; n <= radm; n++) {
                            if (i + m >= 0 && i + m < D1 && j + n >= 0 && j + n < D0 && s + o >= 0 && s + o < D2) {
                                const int64_t q = j + n + (i + m) * D0 + (s + o) * D0 * D1;
                                totwei = totwei + dx * dy * dz * MassDen[q];
                                if (SAR[c] > 0) {
                                    totpx++;
                                    for (int64_t t = 0; t < T; t++) {
                                        totSAR[t] = totSAR[t] + SAR[t * shift + q];
                                        totPower[t] = totPower[t] + Power[t * shift + q];
                                    }
                                }
                            }
                        }
                    }
                }
            }
            if (err) continue;
            const double coefw = (grams * 1e-3 - totwei2) / (totwei - totwei2);
            const double totpxs = coefw * (double)(totpx - totpx2);
            for (int64_t t = 0; t < T; t++) {
                const double totSARs = coefw * (totSAR[t] - totSAR2[t]);
                const double totPowers = coefw * (totPower[t] - totPower2[t]);
                if (totpx > 0) {
                    aveSAR[t * shift + c] = (totSAR2[t] + totSARs) / ((double)totpx2 + totpxs);
                    avePower[t * shift + c] = (totPower2[t] + totPowers) / ((double)totpx2 + totpxs);
                }
            }
        }
        free(totSAR);
    }
    return err ? -1 : 0;
}
