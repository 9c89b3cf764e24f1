/*
 * mrilab_amd.h -- C ABI of the MI355X Bloch simulation engine.
 *
 * The engine replaces MRiLab's scan MEX functions. Those take no MEX
 * arguments and exchange everything through seven MATLAB global structs
 * (VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig). This ABI passes the same
 * fields, under the same names, as plain C structs of pointers and sizes.
 * A MEX gateway fills them with mxGetData() pointers and calls one
 * entry point. Arrays keep MATLAB's column-major layout:
 *   per-voxel maps   [R, C, Z]              flat index r + R*c + R*C*z
 *   per-spin arrays  [R, C, Z, SpinNum, TypeNum]
 *   per-type maps    [R, C, Z, TypeNum]
 *   coil maps        [R, C, Z, CoilNum]
 *
 * Ownership: the library never allocates caller-visible memory. Outputs
 * are written in place into caller buffers, as the reference mutates the
 * MATLAB-owned arrays after mxUnshareArray (DoScanAtCPU.cpp:118-124).
 * Device memory lives inside the library (per call, or in an mrs_plan).
 *
 * Errors: every entry point returns an mrs_status; mrs_last_error()
 * gives a message for the calling thread. Nothing throws across the ABI.
 * Threading: calls on different plans may run on different threads; one
 * plan is not re-entrant. Callbacks run on the calling thread.
 */
#ifndef MRILAB_AMD_H
#define MRILAB_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRS_ABI_VERSION 3

typedef enum mrs_status {
    MRS_OK = 0,
    MRS_ERR_ARG = 1,          /* NULL pointer, bad size or inconsistent dims */
    MRS_ERR_DEVICE = 2,       /* HIP runtime error or invalid GPUIndex */
    MRS_ERR_CALLBACK = 3,     /* an Ext callback returned nonzero */
    MRS_ERR_UNSUPPORTED = 4,  /* outside the engine's limits (coil count ...) */
    MRS_ERR_NOMEM = 5         /* device or host allocation failed */
} mrs_status;

/* ---- VObj: phantom (DoScanAtCPU.cpp:127-135, DoGMScanAtGPU.cu K/TypeFlag) */
typedef struct mrs_vobj {
    double Gyro;              /* rad/s/T */
    int32_t SpinNum;          /* isochromats per voxel */
    int32_t TypeNum;          /* pools / species */
    int32_t dims[3];          /* R, C, Z of VObj.Mz (Z = 1 for a 2-D object) */
    float *Mx, *My, *Mz;      /* [R,C,Z,SpinNum,TypeNum], in: start, out: end */
    const float *Rho;         /* [R,C,Z,TypeNum] proton density */
    const float *T1, *T2;     /* [R,C,Z,TypeNum] seconds */
    const float *K;           /* multi-pool only, layout per entry point */
    const double *TypeFlag;   /* multi-pool only: [TypeNum], 0 free, 1 bound */
} mrs_vobj;

/* ---- VMag: field maps (DoScanAtCPU.cpp:137-141) */
typedef struct mrs_vmag {
    const float *dB0;         /* [R,C,Z] Tesla */
    const float *dWRnd;       /* [R,C,Z,SpinNum,TypeNum] rad/s */
    const float *Gzgrid, *Gygrid, *Gxgrid;  /* [R,C,Z] metres */
} mrs_vmag;

/* ---- VCoi: coils (DoScanAtCPU.cpp:143-150) */
typedef struct mrs_vcoi {
    const float *TxCoilmg, *TxCoilpe;  /* [R,C,Z,TxCoilNum] B1+ magnitude, phase */
    const float *RxCoilx, *RxCoily;    /* [R,C,Z,RxCoilNum] B1- x, y */
    int32_t TxCoilNum, RxCoilNum;
    double TxCoilDefault, RxCoilDefault; /* nonzero: uniform coil */
} mrs_vcoi;

/* ---- VCtl: controls (DoScanAtCPU.cpp:152-155, Plugin_rfRef.m:7-11) */
typedef struct mrs_vctl {
    const double *CS;         /* [TypeNum] chemical shift, Hz */
    int32_t TRNum;
    int32_t RunMode;          /* 0 imaging, 1 per-step magnetization capture */
    int32_t MaxThreadNum;     /* accepted for compatibility; unused on GPU */
    int32_t GPUIndex;
    int32_t MultiTransmit;    /* 0 = 'off', 1 = 'on' */
    int32_t MasterTxCoil;     /* 1-based, used by Ext 7 when MultiTransmit on */
} mrs_vctl;

/* ---- VSeq: sequence lines (DoScanAtCPU.cpp:157-175) */
typedef struct mrs_vseq {
    const double *utsLine, *tsLine;
    const double *rfAmpLine, *rfPhaseLine, *rfFreqLine, *rfCoilLine;
    const double *GzAmpLine, *GyAmpLine, *GxAmpLine;
    const double *ADCLine, *ExtLine;
    const double *flagsLine;  /* [6, n_ts] column-major: rf,Gz,Gy,Gx,ADC,Ext */
    int64_t n_ts, n_uts, n_rf, n_Gz, n_Gy, n_Gx, n_ADC, n_Ext;
} mrs_vseq;

/* ---- VVar: decoder state, read at start and written back at the end
 * (DoScanAtCPU.cpp:177-199, DoPreScan.m:517-547, Plugin_LockK.m) */
typedef struct mrs_vvar {
    double t, dt;
    double *rfAmp, *rfPhase, *rfFreq;   /* [TxCoilNum] held rf values */
    double rfCoil, rfRef;
    double GzAmp, GyAmp, GxAmp, ADC, Ext;
    double Kz, Ky, Kx;                  /* running k-space position, 1/m */
    double KzKeeper, KyKeeper, KxKeeper;
    double gpuFetch;
    int32_t utsi, rfi, Gzi, Gyi, Gxi, ADCi, Exti, TRCount;
} mrs_vvar;

/* ---- VSig: outputs (DoScanAtCPU.cpp:201-210) */
typedef struct mrs_vsig {
    double *Sx, *Sy;          /* [SignalNum*TypeNum*RxCoilNum], index
                                 t*Rx*N + c*N + n; accumulated with += */
    double *Kz, *Ky, *Kx;     /* [SignalNum], accumulated with += */
    int32_t SignalNum;
    float *Mx, *My, *Mz;      /* RunMode 1 only: [R,C,Z,S,Ty, n_uts] */
    double *Muts;             /* RunMode 1 only: [n_uts] */
} mrs_vsig;

/* ---- Ext callbacks (DoScanAtCPU.cpp:356-409, DoScanAtGPU.cu:479-662)
 * Ext 1-4 (k reset/reverse/lock/release), 6 (ideal spoiler) and 7 (rfRef)
 * are executed by the engine itself. Any other nonzero Ext code (5 timer,
 * 8 motion, 9 real-time recon, negative user codes) calls ext() if it is
 * set, and is a no-op otherwise; both return 0 on success.
 *
 * At the call (DoScanAtGPU.cu:483-560): every step before the Ext event has
 * run, and VSig.Sx/Sy and VSig.Kx/Ky/Kz hold all samples acquired so far
 * (added with +=, as the scan adds the later ones at its end); vvar is the
 * decoder state at the event, and changes to it steer the rest of the scan.
 * When the state is fetched (DoGPUFetch.m), VObj.Mx/My/Mz hold the current
 * magnetization before ext() and, after it, everything the scan reads from
 * the host is read again from the (possibly re-pointed) arrays of
 * mrs_ext_state: VObj.Mx/My/Mz/Rho/T1/T2/K, VMag.dB0/dWRnd/G*grid and the
 * Tx and Rx coil maps (DoScanAtGPU.cu:596-659, pointers re-fetched as
 * :617-659 re-fetches them). Whether it is fetched: mrs_DoScanAtCPU always
 * (the CPU engine's plugins share its arrays, DoScanAtCPU.cpp:366-384); the
 * GPU entry points ask gpu_fetch() if it is set, else follow DoGPUFetch.m
 * (codes 6 and 8 fetch). */
typedef struct mrs_ext_state {
    mrs_vobj *vobj;           /* the caller's VObj */
    mrs_vmag *vmag;           /* the scan's VMag (a copy; its pointers may be re-pointed) */
    mrs_vcoi *vcoi;           /* the scan's VCoi (likewise) */
    mrs_vsig *vsig;           /* the caller's VSig */
    int32_t fetched;          /* 1 when VObj.Mx/My/Mz hold the current state */
    int32_t reserved;
} mrs_ext_state;

typedef struct mrs_callbacks {
    int (*gpu_fetch)(void *user, double ext, int *fetch);
    int (*ext)(void *user, double ext, mrs_vvar *vvar, mrs_ext_state *state);
    void *user;
} mrs_callbacks;

/* ---- Engine options; NULL means defaults */
typedef struct mrs_options {
    int32_t block_size;       /* threads per block, 0 = auto */
    int32_t spins_per_thread; /* 0 = auto */
    int32_t max_adc_per_launch; /* ADC samples per kernel launch, 0 = auto */
    int32_t multipool_scheme; /* 0 = operator split as DoGMScanAtGPU,
                                 1 = exact FP64 matrix exponential, blocks
                                 of steps that repeat (saturation trains)
                                 applied as the block's exact affine map,
                                 2 = exact FP64, every step stepped */
    int32_t precision_mode;   /* 0 = reference: relaxation factors rounded
                                 as BlochKernel.h:348-368 rounds them;
                                 1 = accurate: expm1-based relaxation */
    int32_t timing;           /* 1 = time every kernel launch with HIP events
                                 on the plan stream (mrs_plan_info) */
    int32_t model;            /* 0 = single-pool engine; multi-pool engines
                                 1 = GM, 2 = MT, 3 = CEST, 4 = ME (VObj.K in
                                 that engine's layout); set by the entry
                                 point, or by the caller of mrs_plan_create */
    int32_t coil_sum;         /* single-pool forward kernel with 4 or 8 receive
                                 coils: 0 = MFMA projection of LDS-staged
                                 samples onto the coil maps, 1 = per-lane
                                 coil products and a wavefront butterfly
                                 (other coil counts always take the latter) */
    int32_t free_steps;       /* single-pool engines, free precession:
                                 2 = deferred: a free step without an ADC
                                 sample adds its rotation angle (the
                                 reference's fp32 angle) to a per-isochromat
                                 fp64 sum and relaxes Mz; the summed rotation
                                 is applied once before the next ADC sample,
                                 rf step, capture or launch end;
                                 1 = every free step rotates, as
                                 BlochKernel.h does;
                                 0 = deferred when the decoded sequence has
                                 at least 8 such steps per summed rotation,
                                 else stepwise (the info structs report the
                                 choice). Both keep each step's angle as the
                                 reference rounds it and agree to fp32
                                 rounding. */
} mrs_options;

/* Thread-local message of the last failed call on this thread. */
const char *mrs_last_error(void);
int32_t mrs_abi_version(void);
/* Number of visible HIP devices, or a negative value on error. */
int32_t mrs_device_count(void);

/* Single-pool scan with DoScanAtCPU semantics, run on the GPU. Any number
 * of receive coils: a kernel pass sums up to 32; with more, the scan's
 * program runs once per group of 32 coils from the same state.
 * Replaces: mexFunction, Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:70 */
mrs_status mrs_DoScanAtCPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                           const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, mrs_vvar *vvar,
                           mrs_vsig *vsig, const mrs_callbacks *cb,
                           const mrs_options *opt);

/* Single-pool scan, GPU entry point.
 * Replaces: mexFunction, Lib/src/engine/GPUEngine/DoScanAtGPU.cu:76 */
mrs_status mrs_DoScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                           const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, mrs_vvar *vvar,
                           mrs_vsig *vsig, const mrs_callbacks *cb,
                           const mrs_options *opt);

/* Generalized multi-pool scan (free and bound pools with exchange, MT
 * saturation), TypeNum 2..6, any number of receive coils (a kernel pass
 * sums 8, or 4 with 5 or 6 pools; more coils re-run the scan per group of
 * coils from the same state). VObj.K is [R,C,Z,TypeNum,TypeNum] with
 * K(:,:,:,a,b) the rate from pool a into pool b; VObj.TypeFlag is 0 for a
 * free pool and 1 for a bound pool. mrs_options.multipool_scheme picks the
 * reference's operator split (0) or the fp64 matrix exponential (1).
 * Replaces: mexFunction, Lib/src/engine/GPUEngine/DoGMScanAtGPU.cu:76 */
mrs_status mrs_DoGMScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                             const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar,
                             mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt);

/* Two-pool magnetization transfer: pool 0 free water, pool 1 bound
 * (super-Lorentzian saturation). VObj.K is [R,C,Z,4]: K(:,:,:,2) the rate
 * water -> bound, K(:,:,:,3) bound -> water; water and bound exchange
 * longitudinally only; signal is water only. TypeFlag is not read.
 * Replaces: mexFunction, Lib/src/engine/GPUEngine/DoMTScanAtGPU.cu:76 */
mrs_status mrs_DoMTScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                             const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar,
                             mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt);

/* CEST: pool 0 water, 1 MT bound pool, 2..TypeNum-1 CEST pools (TypeNum
 * 3..6). VObj.K is [R,C,Z,2*(TypeNum-1)]: plane 1 water -> MT, plane TypeNum
 * MT -> water, plane t water -> CEST pool t, plane t+TypeNum-1 CEST pool t
 * -> water (1-based planes); signal is water only.
 * Replaces: mexFunction, Lib/src/engine/GPUEngine/DoCESTScanAtGPU.cu:76 */
mrs_status mrs_DoCESTScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                               const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                               const mrs_vseq *vseq, mrs_vvar *vvar,
                               mrs_vsig *vsig, const mrs_callbacks *cb,
                               const mrs_options *opt);

/* Multi-exchange: every pool free, VObj.K [R,C,Z,TypeNum^2] with plane
 * a*TypeNum + b (0-based) the rate from pool a into pool b (the transpose
 * of the GM layout, diagonal applied); signal per pool.
 * Replaces: mexFunction, Lib/src/engine/GPUEngine/DoMEScanAtGPU.cu:76 */
mrs_status mrs_DoMEScanAtGPU(mrs_vobj *vobj, const mrs_vmag *vmag,
                             const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                             const mrs_vseq *vseq, mrs_vvar *vvar,
                             mrs_vsig *vsig, const mrs_callbacks *cb,
                             const mrs_options *opt);

/* ---- Plan API: decode once, keep maps and step records resident in HBM,
 * run many times. mrs_DoScanAtGPU is plan_create + plan_run + plan_fetch +
 * plan_destroy. No reference counterpart (the reference re-uploads per
 * call); used by benchmarks and by callers that rerun one scan. */
typedef struct mrs_plan mrs_plan;

mrs_status mrs_plan_create(const mrs_vobj *vobj, const mrs_vmag *vmag,
                           const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                           const mrs_vseq *vseq, const mrs_vvar *vvar,
                           const mrs_vsig *vsig, const mrs_options *opt,
                           mrs_plan **out);
/* Resets device magnetization to the plan's initial state, zeroes the
 * device signal, and runs every step on the plan's stream; asynchronous. */
mrs_status mrs_plan_run(mrs_plan *plan);
mrs_status mrs_plan_sync(mrs_plan *plan);
/* Copies final magnetization, signal and k-space into caller buffers
 * (VObj.Mx/My/Mz, VSig.Sx/Sy/Kx/Ky/Kz with +=) and the final decoder
 * state into vvar. Synchronizes first. */
mrs_status mrs_plan_fetch(mrs_plan *plan, mrs_vobj *vobj, mrs_vvar *vvar,
                          mrs_vsig *vsig);
/* Statistics of the decoded plan. */
typedef struct mrs_plan_info {
    int64_t n_steps;          /* decoded steps with dt > 0 */
    int64_t n_adc;            /* ADC samples */
    int64_t n_iso;            /* isochromats = R*C*Z*SpinNum*TypeNum */
    int64_t n_launches;       /* kernel launches per run */
    int64_t n_rf_steps;       /* steps with rf on */
    int64_t n_cached_steps;   /* free steps reusing the previous rotation */
    int32_t block_size, spins_per_thread, grid_blocks;
    float   last_run_ms;      /* device time of the last run, 0 if unknown */
    double  fwd_kernel_ms;    /* timing = 1: summed Bloch-kernel time of the last run */
    double  reduce_kernel_ms; /* timing = 1: summed signal-reduction time */
    int32_t deferred_free_steps; /* 1 if free precession is deferred (mrs_options.free_steps) */
} mrs_plan_info;
mrs_status mrs_plan_get_info(const mrs_plan *plan, mrs_plan_info *info);
void mrs_plan_destroy(mrs_plan *plan);

/* ---------------------------------------------------------------------------
 * Batched multi-pool sequences: one phantom under n_seq independent
 * sequences, e.g. the saturation offsets of a Z-spectrum, each simulated
 * from VObj's magnetization. No reference counterpart: MRiLab runs such
 * offsets as successive TRs of one scan (DoGMScanAtGPU.cu:76), the state of
 * one carrying into the next. Every (voxel, sequence) pair is a thread of
 * one launch grid (launch groups of whole sequences when their ADC samples
 * exceed one launch's partial buffer), so the device is filled even when the
 * phantom alone is one wave per SIMD. options.model picks the engine (1 GM,
 * 2 MT, 3 CEST, 4 ME, VObj.K in that engine's layout) and
 * options.multipool_scheme the scheme; RunMode 0, no Ext callbacks, at
 * most one kernel pass of receive coils (8, or 4 with 5 or 6 pools). Each
 * sequence's outputs are what the single-scan entry point writes for that
 * sequence alone. Run, sync, info and destroy with mrs_plan_*.
 * ------------------------------------------------------------------------- */
mrs_status mrs_multipool_batch_create(int32_t n_seq, const mrs_vobj *vobj, const mrs_vmag *vmag,
                                      const mrs_vcoi *vcoi, const mrs_vctl *vctl, const mrs_vseq *const *vseq,
                                      const mrs_vvar *const *vvar, const mrs_vsig *const *vsig,
                                      const mrs_options *options, mrs_plan **out);
/* Per sequence q: final magnetization into vobj[q]->Mx/My/Mz, signal and
 * k-space added into vsig[q], decoder state into vvar[q] (any array, or
 * entry, may be NULL). */
mrs_status mrs_multipool_batch_fetch(mrs_plan *plan, mrs_vobj *const *vobj, mrs_vvar *const *vvar,
                                     mrs_vsig *const *vsig);
/* create + run + fetch (final states into vobj_out) + destroy. */
mrs_status mrs_DoMultiPoolBatchAtGPU(int32_t n_seq, const mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                                     const mrs_vctl *vctl, const mrs_vseq *const *vseq, mrs_vvar *const *vvar,
                                     mrs_vsig *const *vsig, mrs_vobj *const *vobj_out, const mrs_options *options);

/* ---------------------------------------------------------------------------
 * Batched scans. No reference counterpart: MRiLab's batch runner
 * (Src/Main/BatchSim.m:129-166) loads saved scans and calls the engine
 * (DoScanAtGPU.cu:76) on each in turn. Here n independent single-pool scans,
 * each with its own phantom, coils and sequence of any length, run in one
 * persistent kernel launch replayed through a captured hipGraph. Each scan's
 * outputs are what mrs_DoScanAtGPU writes for it (up to fp32 summation
 * order in the coil sums).
 * Limits: TypeNum 1, TxCoilNum 1, RunMode 0, one GPUIndex for all scans, no
 * Ext callbacks (Ext 1-4, 6 and 7 run natively, other codes are ignored).
 * options: spins_per_thread / block_size 0 = chosen by the autotuner at
 * create (every valid shape is timed on the device, the fastest is kept);
 * precision_mode as for single scans; timing = 1 times each run.
 * ------------------------------------------------------------------------- */
typedef struct mrs_batch mrs_batch;
mrs_status mrs_batch_create(int32_t n_scans, mrs_vobj *const *vobj, const mrs_vmag *const *vmag,
                            const mrs_vcoi *const *vcoi, const mrs_vctl *const *vctl,
                            const mrs_vseq *const *vseq, mrs_vvar *const *vvar, mrs_vsig *const *vsig,
                            const mrs_options *options, mrs_batch **out);
/* Enqueues one run of every scan (graph replay); asynchronous. */
mrs_status mrs_batch_run(mrs_batch *batch);
mrs_status mrs_batch_sync(mrs_batch *batch);
/* Adds the last run's signals and k-space into each VSig, writes the final
 * magnetization into each VObj and the final decoder state into each VVar
 * (any of the three array arguments may be NULL). */
mrs_status mrs_batch_fetch(mrs_batch *batch, mrs_vobj *const *vobj, mrs_vvar *const *vvar,
                           mrs_vsig *const *vsig);
typedef struct mrs_batch_info {
    int64_t n_scans;
    int64_t n_tiles;          /* wavefront work items (scan, 64*SPT isochromats) */
    int64_t n_iso;            /* isochromats over all scans */
    int64_t n_iso_steps;      /* sum over scans of isochromats x steps */
    int64_t n_adc;            /* ADC samples over all scans */
    int32_t spins_per_thread, block_size, grid_blocks;
    int32_t tuned;            /* 1 if the autotuner chose spins_per_thread and block_size */
    int32_t n_candidates;     /* shapes timed by the autotuner */
    float   last_run_ms;      /* timing = 1: device time of the last run */
    float   tune_best_ms, tune_worst_ms;
    double  decode_ms;        /* host time spent decoding the sequences at create */
    double  tune_ms;          /* host time spent in the autotuner at create */
    double  create_ms;        /* host time of mrs_batch_create as a whole */
    int32_t deferred_free_steps; /* 1 if free precession is deferred (mrs_options.free_steps) */
} mrs_batch_info;
mrs_status mrs_batch_get_info(const mrs_batch *batch, mrs_batch_info *info);
void mrs_batch_destroy(mrs_batch *batch);
/* create + run + fetch + destroy. */
mrs_status mrs_DoScanBatchAtGPU(int32_t n_scans, mrs_vobj *const *vobj, const mrs_vmag *const *vmag,
                                const mrs_vcoi *const *vcoi, const mrs_vctl *const *vctl,
                                const mrs_vseq *const *vseq, mrs_vvar *const *vvar, mrs_vsig *const *vsig,
                                const mrs_options *options);

/* ---------------------------------------------------------------------------
 * Adjoint rf design. No reference counterpart: MRiLab simulates only. The
 * forward model is the engine's rf step (BlochKernel.h:146-306 rotation,
 * BlochKernel.h:348-368 relaxation, reference fp32 rounding) over a fixed-dt
 * excitation with complex rf rf[k] (T, [n_steps][2]) and gradients
 * grad[k] (Gx, Gy, Gz in T/m) on the phantom of VObj/VMag/VCoi (single pool,
 * one Tx coil; VObj.Mx/My/Mz is the initial state, VObj.Rho sets M0;
 * VObj.T1/T2 and VMag.dB0 are replaced by the parameters of each call).
 * Loss L = sum over voxels of |sum over isochromats (Mx + iMy) - target|^2.
 * mrs_adjoint_backward returns dL/drf, dL/dT1, dL/dT2 and dL/ddB0 by a
 * backward-in-time pass that keeps the forward state every `checkpoint`
 * steps. Between checkpoints it either rebuilds the states backwards from
 * the checkpoint at the segment's end (inverse relaxation and rotation, no
 * scratch memory) or recomputes the segment forwards into a scratch buffer
 * of checkpoint * 5 * n_iso floats; rebuilding is used only where undoing a
 * segment's relaxation amplifies rounding by at most 8 (K*dt <= 2.08*T2 and
 * K*dt <= 2.08*T1 in every voxel).
 * ------------------------------------------------------------------------- */
typedef struct mrs_excitation {
    int32_t n_steps;
    int32_t checkpoint;       /* K; 0 = 64 */
    int32_t segment_mode;     /* 0 = rebuild segments backwards where that is
                                 safe, else recompute them into scratch;
                                 1 = always recompute into scratch */
    int32_t reserved;
    double  dt;               /* s */
    const float *grad;        /* [n_steps][3] Gx, Gy, Gz (T/m), host */
    const float *target;      /* [nvox][2] target transverse magnetization, host */
} mrs_excitation;
typedef struct mrs_adjoint mrs_adjoint;
mrs_status mrs_adjoint_create(const mrs_vobj *vobj, const mrs_vmag *vmag, const mrs_vcoi *vcoi,
                              const mrs_excitation *exc, const mrs_options *options, mrs_adjoint **out);
/* rf [n_steps][2] (T), T1/T2 [nvox] (s), dB0 [nvox] (T); device = 1 when the
 * pointers are device memory of GPU 0, 0 for host memory. Writes the loss
 * and, if mxy is not NULL, the voxel sums [nvox][2] (host, fp64). */
mrs_status mrs_adjoint_forward(mrs_adjoint *adj, const float *rf, const float *T1, const float *T2,
                               const float *dB0, int32_t device, double *loss, double *mxy);
/* Gradients of the loss of the last forward pass; same memory convention. */
mrs_status mrs_adjoint_backward(mrs_adjoint *adj, float *g_rf, float *g_T1, float *g_T2, float *g_dB0,
                                int32_t device);
typedef struct mrs_adjoint_info {
    int64_t n_iso, n_steps;
    int32_t checkpoint, n_checkpoints, spins_per_thread, block_size;
    int32_t segments_rebuilt; /* 1 if the last backward pass rebuilt the
                                 segments backwards, 0 if it used scratch */
    int32_t reserved;
    float   forward_ms;       /* device time of the last forward pass */
    float   backward_ms;      /* device time of the last backward pass */
} mrs_adjoint_info;
mrs_status mrs_adjoint_get_info(const mrs_adjoint *adj, mrs_adjoint_info *info);
void mrs_adjoint_destroy(mrs_adjoint *adj);

/* ---------------------------------------------------------------------------
 * Signal to image: Cartesian reconstruction and k-space noise.
 * ------------------------------------------------------------------------- */
typedef struct mrs_recon_desc {
    int32_t res_freq;         /* VCtl.ResFreq: readout samples per echo */
    int32_t res_phase;        /* VCtl.ResPhase */
    int32_t slice_num;        /* VCtl.SliceNum */
    int32_t te_per_tr;        /* VCtl.TEPerTR */
    int32_t rx;               /* VCoi.RxCoilNum */
    int32_t types;            /* VObj.TypeNum */
    int32_t zf_kx, zf_ky;     /* VCtl.ZF_Kx, ZF_Ky; 0 = no zero filling */
    int32_t zf_kz;            /* the factor of VCtl.ZF_Kz ('x2' -> 2) */
    int32_t swap_xy;          /* 1 when VCtl.ScanPlane/FreqDir select the
                                 second permute (DoCartRecon.m:45-73) */
} mrs_recon_desc;
/* VImg: images and the permuted k-space, each [d0][d1][d2][Rx][TEPerTR]
 * column-major (dims from mrs_recon_dims), fp64. */
typedef struct mrs_vimg {
    double *Real, *Imag, *Mag, *Phase;
    double *Sx, *Sy;
} mrs_vimg;
/* d0..d2 of the image volume, then Rx and TEPerTR. */
mrs_status mrs_recon_dims(const mrs_recon_desc *desc, int64_t dims[5]);
/* Reconstructs VSig.Sx/Sy (SignalNum = ResFreq*TEPerTR*ResPhase*SliceNum;
 * index t*Rx*N + c*N + n) into VImg: per echo the types are summed, the
 * last readout sample dropped, k-space zero-filled and permuted, then
 * fftshift(ifftn(fftshift(.))) per coil, in fp64 on hipFFT. device = 1 when
 * vsig->Sx/Sy and the VImg arrays are device memory of GPU 0.
 * Replaces: DoCartRecon, Src/Recon/Default/DoCartRecon.m:3-91 */
mrs_status mrs_DoCartRecon(const mrs_vsig *vsig, const mrs_recon_desc *desc, mrs_vimg *img, int32_t device);
/* EPI and FSE recon: the signal of an echo-train scan (per shot EPI_ETL or
 * FSE_ETL echoes of ResFreq samples, SignalNum = ResFreq*ETL*ShotNum*
 * SliceNum, TEPerTR 1) put in Cartesian order (echo e of shot p on phase
 * line e*ShotNum + p; flip_even = 1 for EPI reverses the readout of every
 * second echo), then the mrs_DoCartRecon pipeline with res_phase =
 * ETL*ShotNum.
 * Replaces: DoEPIRecon, Src/Recon/Default/DoEPIRecon.m:1-95 (flip_even 1)
 *           DoFSERecon, Src/Recon/Default/DoFSERecon.m:1-90 (flip_even 0) */
mrs_status mrs_DoEchoTrainRecon(const mrs_vsig *vsig, const mrs_recon_desc *desc, int32_t etl, int32_t shots,
                                int32_t flip_even, mrs_vimg *img, int32_t device);
/* The same from the k-space a plan's last run left in device memory (no
 * download of the signal). */
mrs_status mrs_plan_recon(mrs_plan *plan, const mrs_recon_desc *desc, mrs_vimg *img, int32_t device);
/* Adds scale * N(0, 1) noise to VSig.Sx and VSig.Sy (n values each), with
 * a counter-based generator (Philox-4x32-10, Box-Muller in fp64) seeded by
 * seed: the same seed gives the same noise on any device. scale is the
 * DoAddNoise.m:8-16 NoiseScale; mrs_noise_scale computes it.
 * Replaces: DoAddNoise, Src/Main/DoAddNoise.m:1-21 */
mrs_status mrs_DoAddNoise(double *Sx, double *Sy, int64_t n, double scale, uint64_t seed, int32_t device);
/* NoiseScale of DoAddNoise.m:8-16 from VCtl.BandWidth, NoiseLevel, NEX, B0,
 * RFreq, RPhase, RSlice, ResFreq, ResPhase, SliceNum. */
double mrs_noise_scale(double bandwidth, double noise_level, double nex, double b0, double rfreq, double rphase,
                       double rslice, double res_freq, double res_phase, double slice_num);

/* ---------------------------------------------------------------------------
 * Local SAR (DoCalSARAtCPU.cpp, SARKernel.h). Per voxel and step with rf on,
 * E = sum over Tx coils of rfAmp_c * TxE1_c (x, y, z components; the rf
 * phase does not enter) and VVar.SAR += dt * E^2 per component; at each
 * VSig.tSample time VSig.SAR(:, s) += sum_k VVar.SAR_k * VObj.ECon_k.
 * ------------------------------------------------------------------------- */
typedef struct mrs_sar_obj {
    int32_t dims[3];          /* VObj.MassDen dims R, C, Z (Z = 1 for 2-D) */
    const float *ECon;        /* VObj.ECon [R,C,Z,3] */
    const float *TxE1x, *TxE1y, *TxE1z;  /* VCoi [R,C,Z,TxCoilNum] */
    int32_t TxCoilNum;
    int32_t TRNum;
} mrs_sar_obj;
typedef struct mrs_sar_sig {
    float *SARBase;           /* VVar.SAR [R,C,Z,3], running, in/out */
    int32_t SARi;             /* VVar.SARi, in/out */
    int32_t reserved;
    float *SAR;               /* VSig.SAR [R,C,Z,n_samples], accumulated with += */
    double *Muts;             /* VSig.Muts [n_uts] */
    const double *tSample;    /* VSig.tSample [n_samples] */
    double *tRealSample;      /* VSig.tRealSample [n_samples] */
    int64_t n_samples;
    int64_t n_rf_steps;       /* out: steps with rf on that the call summed */
} mrs_sar_sig;
/* Runs the SAR event loop over VSeq from VVar's state, writes VVar and the
 * sampled SAR back.
 * Replaces: mexFunction, Lib/src/engine/CPUEngine/DoCalSARAtCPU.cpp:63 */
mrs_status mrs_DoCalSARAtGPU(const mrs_sar_obj *obj, const mrs_vseq *vseq, mrs_vvar *vvar, mrs_sar_sig *sig);

/* N-gram local SAR average ([aveSAR, avePower] = DoSARAverageAtCPU(SAR,
 * Power, MassDen, dx, dy, dz, grams)). SAR, Power, aveSAR and avePower are
 * fp64 [rows, cols, slices, times] (dims[0..3]), MassDen fp64 [rows, cols,
 * slices] in kg/m^3, dx/dy/dz the voxel size in m. Per voxel with
 * MassDen != 0 and SAR > 0 at the first time point, a voxel-index sphere
 * grows one voxel of radius at a time until it holds more than grams*1e-3
 * kg, and the sums of the last two spheres are interpolated to exactly that
 * mass; other voxels get 0. Results equal the reference's arithmetic bit
 * for bit. Returns MRS_ERR_ARG when the volume's mass does not exceed
 * grams*1e-3 kg (the reference loops forever there). device != 0: all
 * arrays are device memory of GPU 0.
 * Replaces: mexFunction, Lib/src/utilities/DoSARAverageAtCPU/DoSARAverageAtCPU.cpp:19 */
mrs_status mrs_DoSARAverageAtGPU(const double *SAR, const double *Power, const double *MassDen, const int64_t dims[4],
                                 double dx, double dy, double dz, double grams, double *aveSAR, double *avePower,
                                 int32_t device);
/* mrs_DoSARAverageAtGPU on device arrays of the calling thread's current GPU
 * that allocates no memory: no hipMalloc, hipFree or allocating library
 * call. No counterpart in the reference; for callers that cannot allocate
 * (code under stream capture, serving loops, harnesses that forbid device
 * allocation). The per-voxel radius, weight and denominator and the mass
 * partials live in a static device array of 20 MiB + 4 KiB that every GPU
 * the library runs on holds until the process ends, so the volume may have
 * at most 2^20 voxels (rows*cols*slices; MRS_ERR_UNSUPPORTED above). All
 * calls on one GPU share that array: calls that overlap on one GPU race.
 * stream: a hipStream_t, NULL for the default stream; the call synchronizes
 * it (the mass check reads two values back). */
mrs_status mrs_DoSARAverageAtGPU_noalloc(const double *SAR, const double *Power, const double *MassDen,
                                         const int64_t dims[4], double dx, double dy, double dz, double grams,
                                         double *aveSAR, double *avePower, void *stream);

/* ---------------------------------------------------------------------------
 * Sequence compiler: a PSD XML file (MRiLab's MRiLabSeq format: CVs, then
 * Pulses sections of rf / GzSS / GyPE / GxR / ADC / Ext elements whose
 * attributes are MATLAB expressions of CVn, VCtl.x, VVar.x) and the VCtl
 * scan parameters to the VSeq lines, as DoPulseGen runs the PSD function
 * that DoWriteXML2m writes. Elements: rfSinc, rfRect, rfGaussian, rfFermi,
 * the adiabatic rfHyperbolicSecant, rfTanhTan, rfBIR and rfBIREF,
 * GzCartesian, GzSelective, GzSelective2, {Gx,Gy,Gz}Trapezoid,
 * {Gx,Gy,Gz}AreaTrapezoid(2), GyCartesian, GxCartesian, ADCCartesian, the EPI
 * (GyEPI, GxEPI, ADCEPI), FSE (GyFSE, GxFSE, ADCFSE), radial (GyRadial,
 * GxRadial, ADCRadial) and spiral (GySpiral, GxSpiral, ADCSpiral) elements,
 * ExtBit; any other element is an MRS_ERR_ARG naming it.
 * ------------------------------------------------------------------------- */
typedef struct mrs_seq_param {
    const char *name;         /* VCtl field, e.g. "TR", "FlipAng", "ME_TEs" */
    const double *value;      /* n numbers (VCtl.x(i) indexes them from 1) */
    int32_t n;
    int32_t reserved;
    const char *text;         /* text fields (TEAnchor, MultiTransmit, DP_Flag);
                                 NULL for numbers */
} mrs_seq_param;
typedef struct mrs_seq mrs_seq;  /* a compiled sequence, owned by the library */
typedef struct mrs_seq_lines {   /* caller buffers of the sizes mrs_seq_counts gives */
    double *utsLine, *tsLine;
    double *rfAmpLine, *rfPhaseLine, *rfFreqLine, *rfCoilLine;
    double *GzAmpLine, *GyAmpLine, *GxAmpLine, *ADCLine, *ExtLine;
    double *flagsLine;        /* [6, n_ts] column-major */
} mrs_seq_lines;
/* Compiles psd_xml (the file's text) for the VCtl parameters (TR, TE,
 * FlipAng, FOVFreq/FOVPhase/FOVSlice, ResFreq/ResPhase, SliceNum,
 * FirstPhNum/SecondPhNum, BandWidth, RFreq, MinUpdRate, MasterTxCoil
 * (default 1), TEAnchor (default 'Middle'), TEAnchorTime (default 0),
 * MultiTransmit (default 'off'), DP_Flag/DP_TR/DP_FlipAng/DP_Num, the
 * technique fields the elements read (MaxGrad, MaxSlewRate, RPhase;
 * EPI_ETL/EPI_ESP/EPI_ShotNum/EPI_EchoShifting; FSE_ETL/FSE_ESP/
 * FSE_ShotNum; R_AngPattern/R_AngRange (text)/R_SpokeNum/R_SampPerSpoke;
 * S_Lamda/S_SlewRate/S_SlewRate0/S_Gradient/S_ShotNum) and whatever the
 * PSD's expressions name) and gyro = VObj.Gyro (rad/s/T).
 * Replaces: DoPulseGen(Simuh), Src/Main/DoPulseGen.m:2-117 with
 * Src/Main/DoWriteXML2m.m:11-337 and the Macro/SeqElem elements above
 * (the gradient and update-rate checks of DoPulseGen.m:119-134 are
 * mrs_seq_check). */
mrs_status mrs_seq_compile(const char *psd_xml, const mrs_seq_param *vctl, int32_t n_params, double gyro,
                           mrs_seq **out);
/* counts = {n_ts, n_uts, n_rf, n_Gz, n_Gy, n_Gx, n_ADC, n_Ext, TRs}. */
/* The checks DoPulseGen makes after compiling (Src/Main/DoPulseGen.m:
 * 119-134): DoMultiTransmitChk.m when VCtl.MultiTransmit is 'on' and
 * tx_coil_num > 0 (no CoilID above tx_coil_num = VCoi.TxCoilNum, the
 * master coil used), DoGradChk.m (slew rate dG/dt of each
 * gradient line within MaxSlewRate + 0.5 %, amplitude within MaxGrad +
 * 0.5 %, GzSS, GyP, GxR in turn) then DoUpdRateChk.m (successive time
 * points of each Tx coil's rf and of each gradient line at least
 * 0.99 MinUpdRate apart, steps below eps/2 ignored). MRS_ERR_ARG names the
 * first failing check; mrs_seq_compile does not call it. */
mrs_status mrs_seq_check(const mrs_seq *seq, double max_grad, double max_slew_rate, double min_upd_rate,
                         int32_t tx_coil_num);
mrs_status mrs_seq_counts(const mrs_seq *seq, int64_t counts[9]);
/* Copies the lines into the caller's buffers. */
mrs_status mrs_seq_copy(const mrs_seq *seq, const mrs_seq_lines *dst);
void mrs_seq_destroy(mrs_seq *seq);

#ifdef __cplusplus
}
#endif
#endif /* MRILAB_AMD_H */
