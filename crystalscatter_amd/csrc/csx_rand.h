// The C library's rand() sequence (glibc's default generator, the additive feedback generator
// of random() with a 31-word table), restated so that its state can be read, advanced by n
// draws without drawing them, and continued on the device.  The reference draws every random
// value with rand() after srand(seed) (zuf(), Source/sc_calc_generic_cpu.cpp:1085 and
// Source/sc_mainCons.cpp:850-853).
//
// The sequence, as published (P. Selinger, "The GLIBC pseudo-random number generator", 2007):
//   r[0] = seed (1 when seed is 0), r[i] = 16807 r[i-1] mod (2^31 - 1) for 0 < i < 31,
//   r[i] = r[i-31] for 31 <= i < 34, r[i] = r[i-3] + r[i-31] mod 2^32 for i >= 34,
//   the k-th rand() value is r[k + 344] >> 1.
// From r[34] on every term obeys the recurrence, so the last 31 terms are a full state, and the
// term n places ahead is a fixed linear combination of them: with f(x) = x^31 - x^28 - 1 and
// x^n = sum_j a_j x^j mod f, r[i + n] = sum_j a_j r[i + j] (mod 2^32).  The coefficients come
// from square-and-multiply on polynomials mod f, so a jump of n draws costs O(31^2 log n).
#pragma once

#include <cstdint>

namespace csr {

constexpr int kDeg = 31;   // table length
constexpr int kLag = 28;   // r[i] = r[i - 31] + r[i - 31 + kLag]
constexpr int kRows = kDeg * kDeg;

// The last 31 terms: w[(pos + j) % 31] is the j-th oldest.
struct State {
    uint32_t w[kDeg];
    int pos;
};

// The next term (its rand() value is the term >> 1).
inline uint32_t step(State& s) {
    const uint32_t v = s.w[s.pos] + s.w[(s.pos + kLag) % kDeg];
    s.w[s.pos] = v;
    s.pos = s.pos + 1 == kDeg ? 0 : s.pos + 1;
    return v;
}

// rand() value of a term.
inline int32_t value(uint32_t term) { return static_cast<int32_t>(term >> 1); }

inline void seed(State& s, uint32_t seed) {
    int64_t r[kDeg];
    r[0] = static_cast<int32_t>(seed == 0 ? 1u : seed);
    for (int i = 1; i < kDeg; ++i) {
        // 16807 r mod (2^31 - 1) without overflow (Schrage), the table holding signed words
        const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int64_t v = 16807 * lo - 2836 * hi;
        if (v < 0) v += 2147483647;
        r[i] = v;
    }
    // terms r[3] .. r[33], oldest first (r[31..33] repeat r[0..2])
    for (int j = 0; j < kDeg; ++j) s.w[j] = static_cast<uint32_t>(r[(j + 3) % kDeg]);
    s.pos = 0;
    for (int k = 0; k < 10 * kDeg; ++k) step(s);  // the terms before the first rand() value
}

// The state's terms oldest first, and back.
inline void window(const State& s, uint32_t out[kDeg]) {
    for (int j = 0; j < kDeg; ++j) out[j] = s.w[(s.pos + j) % kDeg];
}
inline void set_window(State& s, const uint32_t in[kDeg]) {
    for (int j = 0; j < kDeg; ++j) s.w[j] = in[j];
    s.pos = 0;
}

// c = a b mod f (coefficients mod 2^32, lowest degree first).
inline void polymulmod(const uint32_t a[kDeg], const uint32_t b[kDeg], uint32_t c[kDeg]) {
    uint32_t t[2 * kDeg - 1] = {};
    for (int i = 0; i < kDeg; ++i)
        for (int j = 0; j < kDeg; ++j) t[i + j] += a[i] * b[j];
    // x^d = x^(d-31) (x^28 + 1), from the top down
    for (int d = 2 * kDeg - 2; d >= kDeg; --d) {
        t[d - kDeg + kLag] += t[d];
        t[d - kDeg] += t[d];
    }
    for (int i = 0; i < kDeg; ++i) c[i] = t[i];
}

// x^n mod f.
inline void xpow(uint64_t n, uint32_t out[kDeg]) {
    uint32_t r[kDeg] = {}, b[kDeg] = {}, t[kDeg];
    r[0] = 1;
    b[1] = 1;
    while (n) {
        if (n & 1) {
            polymulmod(r, b, t);
            for (int i = 0; i < kDeg; ++i) r[i] = t[i];
        }
        n >>= 1;
        if (n) {
            polymulmod(b, b, t);
            for (int i = 0; i < kDeg; ++i) b[i] = t[i];
        }
    }
    for (int i = 0; i < kDeg; ++i) out[i] = r[i];
}

// The jump whose first row is x^n mod f: rows[m][j] with r[i + n + m] = sum_j rows[m][j] r[i + j],
// m, j < 31 (row m is x^(n+m) mod f).
inline void poly_rows(const uint32_t xn[kDeg], uint32_t rows[kRows]) {
    uint32_t a[kDeg];
    for (int j = 0; j < kDeg; ++j) a[j] = xn[j];
    for (int m = 0; m < kDeg; ++m) {
        for (int j = 0; j < kDeg; ++j) rows[m * kDeg + j] = a[j];
        // times x: x^31 = x^28 + 1
        const uint32_t top = a[kDeg - 1];
        for (int j = kDeg - 1; j > 0; --j) a[j] = a[j - 1];
        a[0] = top;
        a[kLag] += top;
    }
}

// The jump by n draws.
inline void jump_rows(uint64_t n, uint32_t rows[kRows]) {
    uint32_t a[kDeg];
    xpow(n, a);
    poly_rows(a, rows);
}

inline void apply(const uint32_t rows[kRows], const uint32_t in[kDeg], uint32_t out[kDeg]) {
    for (int m = 0; m < kDeg; ++m) {
        uint32_t v = 0;
        for (int j = 0; j < kDeg; ++j) v += rows[m * kDeg + j] * in[j];
        out[m] = v;
    }
}

// Advance the state by the jump whose rows are given.
inline void advance(State& s, const uint32_t rows[kRows]) {
    uint32_t in[kDeg], out[kDeg];
    window(s, in);
    apply(rows, in, out);
    set_window(s, out);
}

}  // namespace csr

// The generator handle of the C ABI.
struct csx_rand {
    csr::State s;
};
