// Multi-pool (generalized model) kernel launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrs {

constexpr int kGmMaxTypes = 4;

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
};

// Multi-pool layouts of the reference engines (the K layout each entry point
// takes; the plan converts every layout to the rate matrix above).
enum { kVarGM = 0, kVarMT = 1, kVarCEST = 2, kVarME = 3 };

typedef void (*gm_launch_fn)(dim3, dim3, size_t, hipStream_t, const GmParams &, const uint32_t *, int, int);

// Kernel for TypeNum ty (2..4), padded coil count nrx (1, 2, 4, 8), scheme
// (0 split fp32, 1 exact fp64) and variant; nullptr if unsupported.
gm_launch_fn pick_gm(int ty, int nrx, int scheme, int variant);
// Signal values each isochromat contributes per ADC sample (power of two).
int gm_values_per_isochromat(int ty, int nrx);

}  // namespace mrs
