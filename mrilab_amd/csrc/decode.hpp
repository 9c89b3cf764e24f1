// Host-side event decoder: turns MRiLab's VSeq lines into a stream of
// fixed-size step records that the GPU kernels read wave-uniformly.
#pragma once

#include <cstdint>
#include <vector>

#include "mrilab_amd.h"

namespace mrs {

// Record flag bits (word 5 of a record).
enum : uint32_t {
    REC_ADC = 1u << 0,      // take an ADC sample before the step
    REC_RF = 1u << 1,       // rf is on during the step
    REC_SPOIL = 1u << 2,    // zero Mx, My before anything else (Ext 6)
    REC_SAME = 1u << 3,     // free step with the same dt and fields as the previous step
    REC_CAPTURE = 1u << 4,  // RunMode 1: copy the state after the step
    REC_NOSTEP = 1u << 5,   // no precession or relaxation (dt <= 0 event carrier)
    REC_NEWDT = 1u << 6,    // dt differs from the previous step's
    REC_ONCE = 1u << 7,     // free step whose rotation no later step reuses
    REC_ADCRUN = 1u << 8,   // ADC run: the repeats are (ADC sample, free step) pairs
    REC_LOOP = 1u << 9,     // loop head (with REC_NOSTEP): the next word-6 records repeat
};
// Bits 10..31 of the flags word: the repeat count. For a plain record, the
// number of further plain free steps with this record's dt and fields (no
// ADC, rf, spoiler or capture) that follow it; for an REC_ADCRUN record,
// the number of further steps that each take an ADC sample first and then
// step with the same dt and fields. compress_runs() folds them. For a loop
// head, the number of further passes over its block (fold_periods()).
constexpr int REC_REP_SHIFT = 10;
constexpr uint32_t REC_REP_MAX = (1u << 22) - 1;

// Record layout, 32-bit words:
//   0 dt (s)            1 gamma*Gz (rad/s/m)  2 gamma*Gy   3 gamma*Gx
//   4 2*pi*rf frequency summed over active Tx coils (rad/s)
//   5 flags
//   then per Tx coil c (coil_words = 2, or 4 for the multi-pool engine):
//     gamma*A_c*cos(phi_c), gamma*A_c*sin(phi_c) [, |A_c| (T), rf frequency_c (Hz)]
// The stride is rounded up to an even number of words.
constexpr int REC_HEAD = 6;
inline int rec_stride(int tx, int coil_words = 2) { return (REC_HEAD + coil_words * tx + 1) & ~1; }

struct Decoded {
    int tx = 1;
    int coil_words = 2;
    int stride = 8;
    std::vector<uint32_t> rec;       // n_rec * stride words
    std::vector<double> adc_ref;     // rfRef at each ADC sample
    std::vector<double> kx, ky, kz;  // k-space at each ADC sample
    std::vector<int32_t> cap_slot;   // per record (RunMode 1), else empty
    std::vector<double> muts;        // RunMode 1: VSig.Muts values by utsi
    int64_t n_steps = 0, n_rf = 0, n_same = 0;
    int64_t n_rec() const { return stride ? (int64_t)rec.size() / stride : 0; }
};

// Folds runs of plain cached free steps into the record before them
// (repeat count in the flags word; not when records carry RunMode-1
// capture slots), then marks free-step records whose rotation is used by
// that step only (REC_ONCE: no repeats and the next step is not REC_SAME).
// With max_adc_run > 0, a free step with an ADC sample also absorbs the
// ADC steps with the same dt and fields that follow it (REC_ADCRUN, at
// most max_adc_run samples per record) instead of plain steps; only
// kernels that execute REC_ADCRUN may be given such a stream.
void compress_runs(Decoded &d, int64_t max_adc_run = 0);
// Periodic blocks for the exact multi-pool kernel: a block of records that
// recurs at least min_repeats times in a row, holds no ADC sample, spoiler,
// capture or loop, and is at most max_block records long, is kept once
// behind a loop head (REC_LOOP | REC_NOSTEP, further passes in the repeat
// bits, the block length in word 6). Records compare equal when every word
// matches except the flag bits that only the split scheme reads (REC_SAME,
// REC_NEWDT, REC_ONCE). Returns the number of loops made.
int64_t fold_periods(Decoded &d, int64_t min_repeats, int64_t max_block);
// Deferred free precession (the single-pool kernels' DEFER form): the free
// steps that would be deferred (free steps without an ADC sample) and the
// pending rotations they cause (one at each later ADC sample, rf step or
// capture, and one at the end of the stream). A pending rotation costs about
// what kDeferStepsPerRotation deferred steps save, so deferral pays when the
// steps reach that many per rotation.
constexpr int64_t kDeferStepsPerRotation = 8;
void deferral_counts(const Decoded &d, int64_t *steps, int64_t *rotations);
inline bool deferral_pays(int64_t steps, int64_t rotations) {
    return steps > 0 && steps >= kDeferStepsPerRotation * rotations;
}
// ADC samples a record takes.
inline int64_t rec_adc_samples(uint32_t flags) {
    if (!(flags & REC_ADC)) return 0;
    return (flags & REC_ADCRUN) ? 1 + (int64_t)(flags >> REC_REP_SHIFT) : 1;
}

// Checks that decoding VSeq from the state in vvar stays inside the lines:
// every flagsLine entry a finite non-negative integer; per line, the start
// index in VVar plus the flags that advance it within the line's length
// (rf: rfi + the sum of the rf row <= n_rf; Gz/Gy/Gx/ADC/Ext likewise);
// every rfCoilLine value an integer in 1..tx; 0 <= utsi < n_uts. The
// reference reads past the arrays on such input (DoScanAtCPU.cpp:291-409).
// Returns nullptr when the lines are consistent, else what is wrong.
const char *check_lines(const mrs_vseq &seq, const mrs_vvar &vvar, int tx);

// Stop reasons of Decoder::run.
enum DecodeStop { DECODE_END = 0, DECODE_CALLBACK = 1 };

// The event loop of DoScanAtCPU.cpp:246-553 without the Bloch step: it
// advances the VVar state through the VSeq lines exactly as the reference
// does and emits one record per time point with dt > 0.
class Decoder {
public:
    Decoder(const mrs_vobj *vobj, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
            const mrs_vseq *vseq, const mrs_vvar *vvar_in, bool callbacks);
    // Decodes until the end of the sequence or, when callbacks are set, an
    // Ext code the engine does not execute itself. On DECODE_CALLBACK the
    // pending code is ext_code(); the caller runs the callback on vvar()
    // and calls run() again.
    DecodeStop run(Decoded &out);
    double ext_code() const { return ext_pending_; }
    mrs_vvar &vvar() { return v_; }
    // Pending Ext 6 not yet attached to a record; flush_tail() emits a
    // no-step record carrying it (end of sequence).
    void flush_tail(Decoded &out);

private:
    void emit(Decoded &out, bool step, bool adc, bool capture);
    const mrs_vobj *vobj_;
    const mrs_vcoi *vcoi_;
    const mrs_vctl *vctl_;
    const mrs_vseq *vseq_;
    bool callbacks_;
    mrs_vvar v_;
    std::vector<double> rfAmp_, rfPhase_, rfFreq_;
    int64_t i_ = 0;
    bool spoil_pending_ = false;
    double ext_pending_ = 0;
    bool resume_after_ext_ = false;
    double flag_[6] = {0, 0, 0, 0, 0, 0};
    // previous step record fields for REC_SAME
    bool prev_free_ = false;
    bool have_prev_dt_ = false;
    uint32_t prev_dt_ = 0;
    uint32_t prev_key_[4] = {0, 0, 0, 0};
};

}  // namespace mrs
