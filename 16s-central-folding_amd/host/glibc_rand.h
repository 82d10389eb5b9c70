// glibc rand()/srand() as a self-contained generator object.
//
// The reference draws every random number from glibc's rand() (TYPE_3 state,
// the default): ion placement and velocities at set-up (fast.cpp:509-566,
// 5354-5403), HB valence resolution and thermal noise every step
// (fast.cpp:2255-2300, 2598-2604).  The host reproduces that stream without
// touching the process-global generator.
//
// Algorithm (glibc random_r.c, TYPE_3, degree 31, separation 3):
//   r[0] = seed (0 -> 1); r[i] = 16807 * r[i-1] mod (2^31 - 1) for i = 1..30,
//   computed with Schrage's decomposition in signed arithmetic;
//   r[31..33] = r[0..2]; r[i] = r[i-31] + r[i-3] (mod 2^32) for i >= 34;
//   the k-th output (k >= 0) is r[k + 344] >> 1 (the first 310 values are
//   discarded by srand()).
#pragma once
#include <cstdint>

namespace tis_host {

class GlibcRand {
  public:
    static constexpr int32_t kRandMax = 2147483647;

    explicit GlibcRand(uint32_t seed = 1) { reseed(seed); }

    void reseed(uint32_t seed) {
        int32_t s = (int32_t)seed;
        if (s == 0) s = 1;
        uint32_t r[344];
        r[0] = (uint32_t)s;
        int64_t word = s;
        for (int i = 1; i < 31; ++i) {
            // 16807 * word mod (2^31 - 1) without overflow (Schrage).
            int64_t hi = word / 127773;
            int64_t lo = word % 127773;
            word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            r[i] = (uint32_t)(int32_t)word;
        }
        for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
        for (int i = 34; i < 344; ++i) r[i] = r[i - 31] + r[i - 3];
        // ring_[q] = r[313 + q]: the 31 values preceding output index 344.
        for (int q = 0; q < 31; ++q) ring_[q] = r[313 + q];
        pos_ = 0;
        draws_ = 0;
    }

    // One rand() value in [0, RAND_MAX].
    int32_t next() {
        ++draws_;
        return (int32_t)(next_raw() >> 1);
    }

    long long draws() const { return draws_; }

  private:
    // Produces r[i] for the next i and advances.  ring_[pos_] is r[i-31],
    // ring_[(pos_ + 28) % 31] is r[i-3].
    uint32_t next_raw() {
        uint32_t v = ring_[pos_] + ring_[(pos_ + 28) % 31];
        ring_[pos_] = v;
        pos_ = (pos_ + 1) % 31;
        return v;
    }

    uint32_t ring_[31];
    int pos_ = 0;
    long long draws_ = 0;
};

}  // namespace tis_host
