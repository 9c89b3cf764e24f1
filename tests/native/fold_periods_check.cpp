// Host check of fold_periods (mrilab_amd/csrc/decode.cpp): a record stream
// with loops unfolds to the stream it came from, and only blocks that may
// fold (no ADC, spoiler, capture; enough repeats) do; and of deferral_counts
// on a mixed stream. Prints "ok" or the first failure and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "decode.hpp"

using namespace mrs;

static const int kStride = 10;  // multi-pool records with one Tx coil
static const uint32_t kIgnore = REC_SAME | REC_NEWDT | REC_ONCE;

static std::vector<uint32_t> rec_words(uint32_t id, uint32_t flags) {
    std::vector<uint32_t> w(kStride, 0);
    w[0] = 0x3727c5acu;  // 1e-5 s
    w[1] = id;
    w[4] = id * 2654435761u;
    w[5] = flags;
    w[6] = id ^ 0x5bd1e995u;
    return w;
}

static void push(Decoded &d, uint32_t id, uint32_t flags = 0) {
    const std::vector<uint32_t> w = rec_words(id, flags);
    d.rec.insert(d.rec.end(), w.begin(), w.end());
}

// The stream a kernel steps through: loop heads replaced by their passes.
static std::vector<uint32_t> unfold(const std::vector<uint32_t> &rec) {
    std::vector<uint32_t> out;
    const size_t n = rec.size() / kStride;
    for (size_t r = 0; r < n; r++) {
        const uint32_t *w = rec.data() + r * kStride;
        if (w[5] & REC_LOOP) {
            const uint32_t L = w[6], N = 1 + (w[5] >> REC_REP_SHIFT);
            for (uint32_t k = 0; k < N; k++) out.insert(out.end(), w + kStride, w + (1 + L) * kStride);
            r += L;
        } else {
            out.insert(out.end(), w, w + kStride);
        }
    }
    return out;
}

static bool same_stream(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++) {
        const uint32_t m = (i % kStride == 5) ? ~kIgnore : ~0u;
        if ((a[i] & m) != (b[i] & m)) return false;
    }
    return true;
}

static int fails = 0;
static void check(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        fails++;
    }
}

static int64_t loops_in(const Decoded &d) {
    int64_t k = 0;
    for (size_t r = 0; r < d.rec.size() / kStride; r++)
        if (d.rec[r * kStride + 5] & REC_LOOP) k++;
    return k;
}

int main() {
    // a saturation train: prefix, 50 periods of a symmetric pulse (its
    // first record recurs inside the pulse) and a held gap, a readout
    {
        Decoded d;
        d.stride = kStride;
        push(d, 900, REC_NEWDT);
        for (int k = 0; k < 50; k++) {
            for (int j = 0; j < 21; j++) push(d, 1000 + (uint32_t)std::abs(j - 10), REC_RF | (k == 0 && j == 0 ? REC_NEWDT : 0));
            push(d, 2000, 1999u << REC_REP_SHIFT);
        }
        push(d, 3000, REC_SPOIL);
        push(d, 3001, REC_ADC);
        const std::vector<uint32_t> orig = d.rec;
        const int64_t n = fold_periods(d, 11, 1 << 20);
        check(n == 1 && loops_in(d) == 1, "one loop over the 50 periods");
        check(d.rec.size() / kStride == 1 + 1 + 22 + 2, "prefix, head, one period, tail");
        const uint32_t *h = d.rec.data() + kStride;
        check((h[5] & REC_LOOP) && (h[5] & REC_NOSTEP) && (h[5] >> REC_REP_SHIFT) == 49 && h[6] == 22,
              "head: 49 further passes of 22 records");
        check(same_stream(unfold(d.rec), orig), "train unfolds to the input");
    }
    // too few repeats, or an ADC sample, spoiler or capture in the block: no loop
    for (uint32_t bar : {0u, (uint32_t)REC_ADC, (uint32_t)REC_SPOIL, (uint32_t)REC_CAPTURE}) {
        Decoded d;
        d.stride = kStride;
        const int reps = bar ? 40 : 10;
        for (int k = 0; k < reps; k++) {
            push(d, 1, REC_RF);
            push(d, 2, bar);
            push(d, 3, REC_RF);
        }
        const std::vector<uint32_t> orig = d.rec;
        check(fold_periods(d, 11, 1 << 20) == 0 && d.rec == orig, bar ? "barrier records never fold" : "10 < 11 repeats");
    }
    // a single record repeated (a rectangular rf pulse) folds with L = 1
    {
        Decoded d;
        d.stride = kStride;
        for (int k = 0; k < 300; k++) push(d, 7, REC_RF);
        const std::vector<uint32_t> orig = d.rec;
        check(fold_periods(d, 11, 1 << 20) == 1 && d.rec.size() == 2 * kStride && d.rec[6] == 1 &&
                  (d.rec[5] >> REC_REP_SHIFT) == 299, "300 equal rf records: one loop of one record");
        check(same_stream(unfold(d.rec), orig), "rf pulse unfolds to the input");
    }
    // random streams with periodic stretches: unfold(fold(x)) == x
    std::mt19937 rng(12345);
    for (int trial = 0; trial < 400; trial++) {
        Decoded d;
        d.stride = kStride;
        const int parts = 1 + (int)(rng() % 6);
        for (int p = 0; p < parts; p++) {
            if (rng() % 2) {
                const int L = 1 + (int)(rng() % 9), N = 1 + (int)(rng() % 30);
                std::vector<std::pair<uint32_t, uint32_t>> blk;
                for (int i = 0; i < L; i++) {
                    const uint32_t fl = (rng() % 7 == 0) ? REC_ADC : (rng() % 2 ? REC_RF : 0);
                    blk.push_back({(uint32_t)(rng() % 5), fl});
                }
                for (int k = 0; k < N; k++)
                    for (auto &b : blk) push(d, b.first, b.second | ((rng() % 3 == 0) ? REC_NEWDT : 0));
            } else {
                const int n = (int)(rng() % 20);
                for (int i = 0; i < n; i++) push(d, (uint32_t)(rng() % 4), rng() % 2 ? REC_RF : 0);
            }
        }
        const std::vector<uint32_t> orig = d.rec;
        fold_periods(d, 3 + (int64_t)(rng() % 10), 1 << 20);
        if (!same_stream(unfold(d.rec), orig)) {
            std::printf("FAIL: random stream %d does not unfold to its input\n", trial);
            fails++;
            break;
        }
        for (size_t r = 0; r < d.rec.size() / kStride; r++) {
            const uint32_t *w = d.rec.data() + r * kStride;
            if (!(w[5] & REC_LOOP)) continue;
            for (uint32_t i = 1; i <= w[6]; i++)
                if (w[i * kStride + 5] & (REC_ADC | REC_SPOIL | REC_CAPTURE | REC_LOOP)) {
                    std::printf("FAIL: random stream %d: a barrier record inside a loop\n", trial);
                    fails++;
                }
            r += w[6];
        }
    }
    // deferral_counts: deferred free steps (their step counts) and the pending
    // rotations they cause (at an ADC sample, rf step or capture, and at the
    // end); ADC and rf steps are not deferred, a spoiler drops what is pending
    {
        Decoded d;
        d.stride = kStride;
        push(d, 1, 4u << REC_REP_SHIFT);  // 5 free steps
        push(d, 2, REC_ADC);              // rotation 1; a sample, not deferred
        push(d, 3, 0);                    // 1 free step
        push(d, 4, REC_RF);               // rotation 2
        push(d, 5, REC_RF);               // nothing pending
        push(d, 6, 2u << REC_REP_SHIFT);  // 3 free steps
        push(d, 7, REC_NOSTEP | REC_SPOIL); // spoiler event: the pending rotation is dropped
        push(d, 8, REC_NOSTEP | REC_ADC);   // a sample with nothing pending
        push(d, 9, 9u << REC_REP_SHIFT);    // 10 free steps, rotation 3 at the end
        int64_t steps = 0, rot = 0;
        deferral_counts(d, &steps, &rot);
        check(steps == 5 + 1 + 3 + 10 && rot == 3, "deferral counts of a mixed stream");
        check(deferral_pays(20, 2) && !deferral_pays(15, 2) && !deferral_pays(0, 0),
              "deferral pays from 8 steps per rotation");
    }
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
