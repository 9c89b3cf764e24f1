// Multi-pool (generalized model) kernel launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrs {

constexpr int kGmMaxTypes = 6;

struct GmParams {
    float *Mx, *My, *Mz;                 // [nvox, S, Ty]
    const float *dWRnd, *Rho, *T1, *T2;  // [nvox, S, Ty], [nvox, Ty]
    const float *K;                      // [nvox, Ty, Ty]: plane q*Ty + t = rate t -> q
    const float *dB0, *gx, *gy, *gz;     // [nvox]
    const float *txmg, *txpe, *rxx, *rxy;
    float *partial;                      // [adc_local][nblk][nvt]
    int64_t nvox;
    int S, Ty, Tx, Rx, rx_default, stride;
    float gyro;
    float cs[kGmMaxTypes];               // 2*pi*CS per type, rounded as the reference
    int flag[kGmMaxTypes];               // 0 free, 1 bound, per type
    int variant;                         // kVarGM, kVarMT, kVarCEST, kVarME
    int self_x;                          // ME: a nonzero diagonal of K exchanges too
    // Batched sequences (one phantom, many independent record streams):
    // seq_tab[4*q .. 4*q+3] = {first record, end record, first ADC slot of
    // the launch's partial buffer, 0} of sequence q; nullptr for one scan.
    // In batched launches block x maps to (tile x / nseq_l, sequence
    // seq_base + x % nseq_l), each sequence starts from Mx/My/Mz and its
    // final state goes to Mout[c] + q * nvox * S * Ty; vox_perm[i] is the
    // voxel of thread i of a tile row (costliest voxels first).
    const int32_t *seq_tab;
    const int32_t *vox_perm;
    float *Mout[3];
    int seq_base, nseq_l, tiles;  // tiles: voxel tiles of one sequence (grid x / nseq_l)
    // Exact scheme with loop records: (3 Ty + 2) * 3 Ty doubles per thread of
    // the launch grid (entry e of thread i at e * threads + i); nullptr when
    // the record stream has no loops.
    double *loop_cols;
};

// Multi-pool layouts of the reference engines (the K layout each entry point
// takes; the plan converts every layout to the rate matrix above).
enum { kVarGM = 0, kVarMT = 1, kVarCEST = 2, kVarME = 3 };

typedef void (*gm_launch_fn)(dim3, dim3, size_t, hipStream_t, const GmParams &, const uint32_t *, int, int);

// Kernel for TypeNum ty (2..6; 5 and 6 with at most 4 padded receive
// coils), padded coil count nrx (1, 2, 4, 8), scheme
// (0 split fp32, 1 exact fp64) and variant; nullptr if unsupported.
gm_launch_fn pick_gm(int ty, int nrx, int scheme, int variant);
// Signal values each isochromat contributes per ADC sample (power of two).
int gm_values_per_isochromat(int ty, int nrx);

}  // namespace mrs
