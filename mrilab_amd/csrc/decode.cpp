// Event decoder: the time-point loop of MRiLab's CPU engine
// (Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:246-553) restated as a
// producer of step records. The Bloch arithmetic itself runs on the GPU.
#include "decode.hpp"

#include <algorithm>

#include <cmath>
#include <cstring>
#include <unordered_map>

namespace mrs {

static constexpr double kPi = 3.14159265359;  // DoScanAtCPU.cpp:63

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

Decoder::Decoder(const mrs_vobj *vobj, const mrs_vcoi *vcoi, const mrs_vctl *vctl,
                 const mrs_vseq *vseq, const mrs_vvar *vvar_in, bool callbacks)
    : vobj_(vobj), vcoi_(vcoi), vctl_(vctl), vseq_(vseq), callbacks_(callbacks) {
    const int tx = vcoi->TxCoilNum;
    v_ = *vvar_in;
    rfAmp_.assign(vvar_in->rfAmp, vvar_in->rfAmp + tx);
    rfPhase_.assign(vvar_in->rfPhase, vvar_in->rfPhase + tx);
    rfFreq_.assign(vvar_in->rfFreq, vvar_in->rfFreq + tx);
    v_.rfAmp = rfAmp_.data();
    v_.rfPhase = rfPhase_.data();
    v_.rfFreq = rfFreq_.data();
}

void Decoder::emit(Decoded &out, bool step, bool adc, bool capture) {
    const int tx = vcoi_->TxCoilNum;
    const int stride = out.stride;
    const size_t base = out.rec.size();
    out.rec.resize(base + stride, 0u);
    uint32_t *w = out.rec.data() + base;
    uint32_t flags = 0;
    if (spoil_pending_) {
        flags |= REC_SPOIL;
        spoil_pending_ = false;
    }
    if (!step) {
        flags |= REC_NOSTEP;
    } else {
        const float gyro = (float)vobj_->Gyro;
        const float dt = (float)v_.dt;
        // BlochKernel.h:44-51 multiplies the float amplitude by the float gamma
        w[0] = fbits(dt);
        w[1] = fbits((float)v_.GzAmp * gyro);
        w[2] = fbits((float)v_.GyAmp * gyro);
        w[3] = fbits((float)v_.GxAmp * gyro);
        float amp_sum = 0.f;  // BlochKernel.h:54-56, in float
        for (int c = 0; c < tx; c++) amp_sum += std::fabs((float)rfAmp_[c]);
        float wrf = 0.f;
        if (amp_sum != 0.f) {
            flags |= REC_RF;
            out.n_rf++;
            for (int c = 0; c < tx; c++) {
                float A = (float)rfAmp_[c], P = (float)rfPhase_[c], Fq = (float)rfFreq_[c];
                if (tx > 1 && A == 0.f) continue;  // BlochKernel.h:80, inactive Tx coil
                if (A < 0.f) {                     // BlochKernel.h:62-65
                    A = std::fabs(A);
                    P = (float)((double)P + kPi);
                }
                wrf += (float)((double)(Fq * 2.0f) * kPi);  // BlochKernel.h:69
                const double b = (double)gyro * (double)A;
                const int cw = out.coil_words;
                w[REC_HEAD + cw * c] = fbits((float)(b * std::cos((double)P)));
                w[REC_HEAD + cw * c + 1] = fbits((float)(b * std::sin((double)P)));
                if (cw == 4) {  // bound-pool saturation inputs (BlochKernelGM.cuh:325-327)
                    w[REC_HEAD + cw * c + 2] = fbits(A);
                    w[REC_HEAD + cw * c + 3] = fbits(Fq);
                }
            }
            prev_free_ = false;
        } else {
            const bool same = prev_free_ && prev_key_[0] == w[0] && prev_key_[1] == w[1] &&
                              prev_key_[2] == w[2] && prev_key_[3] == w[3];
            if (same) {
                flags |= REC_SAME;
                out.n_same++;
            }
            prev_free_ = true;
            std::memcpy(prev_key_, w, sizeof prev_key_);
        }
        w[4] = fbits(wrf);
        if (!have_prev_dt_ || prev_dt_ != w[0]) flags |= REC_NEWDT;
        have_prev_dt_ = true;
        prev_dt_ = w[0];
        out.n_steps++;
    }
    if (adc) {
        flags |= REC_ADC;
        out.adc_ref.push_back(v_.rfRef);
        out.kz.push_back(v_.Kz);
        out.ky.push_back(v_.Ky);
        out.kx.push_back(v_.Kx);
    }
    if (vctl_->RunMode == 1) {
        if (capture) flags |= REC_CAPTURE;
        out.cap_slot.push_back(capture ? v_.utsi : -1);
    }
    w[5] = flags;
}

void compress_runs(Decoded &d, int64_t max_adc_run) {
    const int st = d.stride;
    if (d.cap_slot.empty()) {
        const int64_t n = d.n_rec();
        const uint32_t plain = REC_ADC | REC_RF | REC_SPOIL | REC_CAPTURE | REC_NOSTEP | REC_NEWDT;
        const uint32_t low = (1u << REC_REP_SHIFT) - 1;
        const int64_t run_cap = std::min<int64_t>(max_adc_run, (int64_t)REC_REP_MAX + 1);
        int64_t out = -1;      // index of the last kept record
        bool open = false;     // the last kept record accepts plain repeats
        bool open_adc = false; // the last kept record accepts ADC-run repeats
        for (int64_t r = 0; r < n; r++) {
            const uint32_t *w = d.rec.data() + r * st;
            const uint32_t f = w[5];
            if (open_adc && (f & low) == (REC_ADC | REC_SAME)) {
                uint32_t *k = d.rec.data() + out * st;
                const uint32_t rep = ((k[5] & REC_ADCRUN) ? (k[5] >> REC_REP_SHIFT) : 0) + 1;
                k[5] = (k[5] & low) | REC_ADCRUN | (rep << REC_REP_SHIFT);
                if ((int64_t)rep + 1 >= run_cap) open_adc = false;
                continue;
            }
            if (open && (f & REC_SAME) && !(f & plain)) {
                uint32_t *k = d.rec.data() + out * st;
                const uint32_t rep = (k[5] >> REC_REP_SHIFT) + 1;
                k[5] = (k[5] & low) | (rep << REC_REP_SHIFT);
                if (rep == REC_REP_MAX) open = false;
                continue;
            }
            out++;
            if (out != r) std::memmove(d.rec.data() + out * st, w, st * sizeof(uint32_t));
            const bool free_step = !(f & (REC_RF | REC_NOSTEP | REC_CAPTURE));
            open_adc = run_cap > 1 && free_step && (f & REC_ADC);
            open = free_step && !open_adc;
        }
        d.rec.resize((size_t)(out + 1) * st);
    }
    // single-use rotations: scan backwards remembering whether the next
    // step record reuses the rotation of the one before it
    bool next_same = false;
    for (int64_t r = d.n_rec() - 1; r >= 0; r--) {
        uint32_t &f = d.rec[(size_t)r * st + 5];
        if (f & REC_NOSTEP) continue;
        const bool free_step = !(f & REC_RF);
        if (free_step && (f >> REC_REP_SHIFT) == 0 && !next_same) f |= REC_ONCE;
        next_same = (f & REC_SAME) != 0;
    }
}

void deferral_counts(const Decoded &d, int64_t *steps, int64_t *rotations) {
    int64_t n = 0, k = 0;
    bool pend = false;
    for (int64_t r = 0; r < d.n_rec(); r++) {
        const uint32_t f = d.rec[(size_t)r * d.stride + 5];
        if (f & REC_SPOIL) pend = false;  // the spoiler discards the pending rotation
        if (f & (REC_ADC | REC_RF | REC_CAPTURE)) {
            if (pend) k++;
            pend = false;
        }
        if (!(f & (REC_NOSTEP | REC_ADC | REC_RF))) {
            n += 1 + (int64_t)(f >> REC_REP_SHIFT);
            pend = !(f & REC_CAPTURE);
        }
    }
    *steps = n;
    *rotations = k + (pend ? 1 : 0);
}

int64_t fold_periods(Decoded &d, int64_t min_repeats, int64_t max_block) {
    const int st = d.stride;
    const int64_t n = d.n_rec();
    if (n < 2 * min_repeats || !d.cap_slot.empty()) return 0;
    const uint32_t ignore = REC_SAME | REC_NEWDT | REC_ONCE;
    const uint32_t barrier = REC_ADC | REC_SPOIL | REC_CAPTURE | REC_LOOP;
    const uint32_t *w = d.rec.data();
    auto word = [&](int64_t r, int k) { return k == 5 ? w[r * st + 5] & ~ignore : w[r * st + k]; };
    auto equal = [&](int64_t a, int64_t b) {
        for (int k = 0; k < st; k++)
            if (word(a, k) != word(b, k)) return false;
        return true;
    };
    // next[r]: the next record with the same content as r (hash, then compare)
    std::vector<int64_t> next((size_t)n, -1);
    {
        std::unordered_map<uint64_t, int64_t> last;
        last.reserve((size_t)n);
        for (int64_t r = n - 1; r >= 0; r--) {
            uint64_t h = 1469598103934665603ull;
            for (int k = 0; k < st; k++) h = (h ^ word(r, k)) * 1099511628211ull;
            auto it = last.find(h);
            if (it != last.end() && equal(r, it->second)) next[r] = it->second;
            last[h] = r;
        }
    }
    // barrier records before r: a block [a, b) is loop-free of them when
    // bar[b] == bar[a]
    std::vector<int64_t> bar((size_t)n + 1, 0);
    for (int64_t r = 0; r < n; r++) bar[r + 1] = bar[r] + ((w[r * st + 5] & barrier) != 0);
    std::vector<uint32_t> out;
    out.reserve(d.rec.size());
    int64_t loops = 0;
    int64_t r = 0;
    while (r < n) {
        int64_t best_len = 0, best_rep = 0;
        int tries = 0;
        for (int64_t p = next[r]; p > r && p - r <= max_block && tries < 8; p = next[p], tries++) {
            const int64_t L = p - r;
            if (bar[p] != bar[r]) break;  // longer blocks hold the same barrier
            int64_t N = 1;
            while (r + (N + 1) * L <= n && bar[r + (N + 1) * L] == bar[r]) {
                bool same = true;
                for (int64_t i = 0; i < L && same; i++) same = equal(r + i, r + N * L + i);
                if (!same) break;
                N++;
            }
            if (N >= min_repeats && N * L > best_rep * best_len) {
                best_len = L;
                best_rep = N;
            }
            if (N >= min_repeats) break;
        }
        if (best_rep >= min_repeats && best_rep - 1 <= (int64_t)REC_REP_MAX) {
            std::vector<uint32_t> head((size_t)st, 0);
            head[5] = REC_LOOP | REC_NOSTEP | ((uint32_t)(best_rep - 1) << REC_REP_SHIFT);
            head[6] = (uint32_t)best_len;
            out.insert(out.end(), head.begin(), head.end());
            out.insert(out.end(), w + r * st, w + (r + best_len) * st);
            r += best_rep * best_len;
            loops++;
        } else {
            out.insert(out.end(), w + r * st, w + (r + 1) * st);
            r++;
        }
    }
    if (loops) d.rec.swap(out);
    return loops;
}

const char *check_lines(const mrs_vseq &S, const mrs_vvar &V, int tx) {
    double sum[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < S.n_ts; i++)
        for (int k = 0; k < 6; k++) {
            const double f = S.flagsLine[i * 6 + k];
            if (!(f >= 0) || f != std::floor(f) || f > 1e9) return "VSeq.flagsLine entries must be non-negative integers";
            sum[k] += f;
        }
    const int64_t start[6] = {V.rfi, V.Gzi, V.Gyi, V.Gxi, V.ADCi, V.Exti};
    const int64_t len[6] = {S.n_rf, S.n_Gz, S.n_Gy, S.n_Gx, S.n_ADC, S.n_Ext};
    static const char *msg[6] = {
        "VVar.rfi plus the rf flags of VSeq.flagsLine exceeds the length of VSeq.rfAmpLine",
        "VVar.Gzi plus the Gz flags of VSeq.flagsLine exceeds the length of VSeq.GzAmpLine",
        "VVar.Gyi plus the Gy flags of VSeq.flagsLine exceeds the length of VSeq.GyAmpLine",
        "VVar.Gxi plus the Gx flags of VSeq.flagsLine exceeds the length of VSeq.GxAmpLine",
        "VVar.ADCi plus the ADC flags of VSeq.flagsLine exceeds the length of VSeq.ADCLine",
        "VVar.Exti plus the Ext flags of VSeq.flagsLine exceeds the length of VSeq.ExtLine"};
    for (int k = 0; k < 6; k++)
        if (start[k] < 0 || (double)start[k] + sum[k] > (double)len[k]) return msg[k];
    for (int64_t j = 0; j < S.n_rf; j++) {
        const double c = S.rfCoilLine[j];
        if (!(c >= 1 && c <= tx) || c != std::floor(c)) return "VSeq.rfCoilLine names a coil outside 1..TxCoilNum";
    }
    if (V.utsi < 0 || V.utsi >= S.n_uts) return "VVar.utsi is outside VSeq.utsLine";
    return nullptr;
}

void Decoder::flush_tail(Decoded &out) {
    // a trailing Ext 6 changes the final state only
    if (spoil_pending_) emit(out, false, false, false);
}

DecodeStop Decoder::run(Decoded &out) {
    const mrs_vseq &S = *vseq_;
    const double *ts = S.tsLine, *fl = S.flagsLine, *uts = S.utsLine;
    const int64_t MaxStep = S.n_ts, MaxutsStep = S.n_uts, MaxrfStep = S.n_rf;
    const int tx = vcoi_->TxCoilNum;
    const double kfac = vobj_->Gyro / (2 * kPi);
    if (out.muts.empty() && vctl_->RunMode == 1) out.muts.assign(MaxutsStep, 0.0);

    for (;;) {
        if (!resume_after_ext_) {
            if (i_ >= MaxStep - 1) return DECODE_END;  // DoScanAtCPU.cpp:246
            double *flag = flag_;
            for (int k = 0; k < 6; k++) flag[k] = 0;
            int64_t i = i_;
            if (ts[i] != ts[i + 1]) {                   // :254-284
                for (int k = 0; k < 6; k++) flag[k] += fl[i * 6 + k];
                i++;
            } else {
                for (int k = 0; k < 6; k++) flag[k] += fl[i * 6 + k];
                while (ts[i] == ts[i + 1]) {
                    for (int k = 0; k < 6; k++) flag[k] += fl[(i + 1) * 6 + k];
                    i++;
                    if (i == MaxStep - 1) break;
                }
                i++;
            }
            i_ = i;
            // :287-289
            v_.t = uts[v_.utsi];
            const int64_t un = v_.utsi + 1 < MaxutsStep - 1 ? v_.utsi + 1 : MaxutsStep - 1;
            v_.dt = uts[un] - uts[v_.utsi];
            v_.utsi = (int32_t)un;
            if (flag[0] >= 1) {                         // :291-320
                for (int64_t j = 0; j < flag[0]; j++) {
                    v_.rfCoil = S.rfCoilLine[v_.rfi];
                    const int c = (int)v_.rfCoil;
                    for (int64_t s = v_.rfi + 1; s < MaxrfStep; s++) {
                        if (v_.rfCoil == S.rfCoilLine[s]) {
                            const double *A = S.rfAmpLine, *P = S.rfPhaseLine, *Fq = S.rfFreqLine;
                            rfAmp_[c - 1] = std::fabs(A[v_.rfi]) <= std::fabs(A[s]) ? A[v_.rfi] : A[s];
                            rfPhase_[c - 1] = std::fabs(P[v_.rfi]) <= std::fabs(P[s]) ? P[v_.rfi] : P[s];
                            rfFreq_[c - 1] = std::fabs(Fq[v_.rfi]) <= std::fabs(Fq[s]) ? Fq[v_.rfi] : Fq[s];
                            break;
                        }
                    }
                    v_.rfi++;
                }
            }
            struct { double *amp; const double *line; int64_t n; int32_t *idx; } G[3] = {
                {&v_.GzAmp, S.GzAmpLine, S.n_Gz, &v_.Gzi},
                {&v_.GyAmp, S.GyAmpLine, S.n_Gy, &v_.Gyi},
                {&v_.GxAmp, S.GxAmpLine, S.n_Gx, &v_.Gxi}};
            for (int g = 0; g < 3; g++) {                // :322-347
                if (flag[1 + g] == 1) {
                    const int64_t k = *G[g].idx;
                    const int64_t k1 = k + 1 < G[g].n - 1 ? k + 1 : G[g].n - 1;
                    *G[g].amp = std::fabs(G[g].line[k]) <= std::fabs(G[g].line[k1]) ? G[g].line[k] : G[g].line[k + 1];
                    (*G[g].idx)++;
                }
            }
            v_.ADC = 0;                                  // :349-354
            if (flag[4] == 1) {
                v_.ADC = S.ADCLine[v_.ADCi];
                v_.ADCi++;
            }
            if (flag[5] == 1) {                          // :356-409
                v_.Ext = S.ExtLine[v_.Exti];
                if (v_.Ext != 0) {
                    const int code = (int)v_.Ext;
                    if (v_.Ext == code && code == 1) {          // Plugin_ResetK.m
                        v_.Kz = 0; v_.Ky = 0; v_.Kx = 0;
                    } else if (v_.Ext == code && code == 2) {   // Plugin_ReverseK.m
                        v_.Kz = -v_.Kz; v_.Ky = -v_.Ky; v_.Kx = -v_.Kx;
                    } else if (v_.Ext == code && code == 3) {   // Plugin_LockK.m
                        v_.KzKeeper = v_.Kz; v_.KyKeeper = v_.Ky; v_.KxKeeper = v_.Kx;
                    } else if (v_.Ext == code && code == 4) {   // Plugin_ReleaseK.m
                        v_.Kz = v_.KzKeeper; v_.Ky = v_.KyKeeper; v_.Kx = v_.KxKeeper;
                    } else if (v_.Ext == code && code == 6) {   // Plugin_IdealSpoiler.m
                        spoil_pending_ = true;
                    } else if (v_.Ext == code && code == 7) {   // Plugin_rfRef.m
                        v_.rfRef = vctl_->MultiTransmit == 0 ? rfPhase_[0]
                                                             : rfPhase_[vctl_->MasterTxCoil - 1];
                    } else if (callbacks_) {
                        ext_pending_ = v_.Ext;
                        resume_after_ext_ = true;
                        return DECODE_CALLBACK;
                    }
                }
                v_.Exti++;
            }
        } else {
            resume_after_ext_ = false;
            v_.Exti++;
        }
        const double *flag = flag_;
        if (flag[0] + flag[1] + flag[2] + flag[3] + flag[4] + flag[5] == 0) {  // :411-420
            for (int k = 0; k < tx; k++) { rfAmp_[k] = 0; rfPhase_[k] = 0; rfFreq_[k] = 0; }
            v_.GzAmp = 0; v_.GyAmp = 0; v_.GxAmp = 0; v_.ADC = 0; v_.Ext = 0;
        }
        if (v_.dt == 0) continue;                        // :423-425
        if (v_.dt < 0) {                                 // :426-440
            v_.TRCount++;
            if (vctl_->RunMode == 1) {
                emit(out, false, false, true);
                out.muts[v_.utsi] = out.muts[v_.utsi - 1];
            }
            continue;
        }
        emit(out, true, v_.ADC == 1, true);
        // :540-542
        v_.Kz += v_.GzAmp * v_.dt * kfac;
        v_.Ky += v_.GyAmp * v_.dt * kfac;
        v_.Kx += v_.GxAmp * v_.dt * kfac;
        if (vctl_->RunMode == 1) out.muts[v_.utsi] = out.muts[v_.utsi - 1] + v_.dt;  // :549
    }
}

}  // namespace mrs
