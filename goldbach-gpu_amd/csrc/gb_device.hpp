// gb_device.hpp -- device-side building blocks for gfx950 (wave64, 160 KB LDS):
// exact small-modulus reduction, 64-bit Montgomery Miller-Rabin, the word
// funnel shift used by the partition marker, and the geometry shared by the
// sieve and marking kernels.
//
// Everything here is integer-exact.  The fp64 paths only ever hold integers
// below 2^53, where fp64 arithmetic is exact, and every quotient estimate is
// corrected by one compare.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbk {

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
// Every window (fused tile or plain segment) starting at even n0 holds the
// odd q = n0 - 2*HALO - 1 + 2i at bit i, so q = n - p for n = n0 + 2j sits at
// bit HALO + j - (p-1)/2.  Primes p <= 2*HALO+1 are answered from the window.
constexpr uint32_t HALO = 4096;              // odds below the first n
constexpr uint32_t HALO_P = 2 * HALO + 1;    // largest p served by the window: 8193
constexpr uint32_t WIN_PAD = 64;             // extra window bits read by the last word
constexpr uint32_t SP_MAX = 1u << 22;        // sieving-prime table bound (odd primes <= 2^22)
constexpr int NPAT = 17;                     // odd primes 3..61 applied as word patterns
constexpr uint32_t CHUNK = 32;               // marks per medium-prime work item ...
constexpr uint32_t WAVE_MARKS = 8;           // marks per lane of a wheel wave item (64 lanes: 512 marks)
constexpr uint32_t CHUNK_SPAN = 48;          // ... spanning 48p bits: odd multiples p*m with 3 | m are
                                             // skipped (the pattern of 3 already cleared them)
constexpr uint32_t MR_PRESIEVE = 1u << 22;   // pre-sieve bound of the tiles decided by Baillie-PSW (the whole table)
// ... of tiles with sparse proofs: only one window word in several is tested, so fewer sieving
// primes (5 % more survivors) save more set-up than the extra tests cost (64 high windows
// 183.1 -> 180.1 ms; 2^20 and 2^19: 181.9, 182.1)
constexpr uint32_t MR_PRESIEVE_SPARSE = 1u << 21;

struct Tables {
    const uint32_t* sp;     // odd primes <= SP_MAX ascending (sp[0] = 3, sp[17] = 67)
    const double* inv;      // 1.0 / p (rounded)
    const uint32_t* c32;    // 2^32 mod p
    const uint32_t* d5;     // (-4 / 6) mod p: from a multiple 6k+1 of p to the next 6k'+5 one, in k
    const uint32_t* items;  // medium-prime work items: (prime index << 12) | chunk
    const uint32_t* items_p;     // the prime of each item (saves a dependent load)
    const uint32_t* item_start;  // item_start[k] = first item of medium prime k
    const uint64_t* p2bits; // odd-only bitset of primes <= p2_limit (bit (q-3)/2)
    uint64_t p2_limit;      // odd upper end of p2bits
    uint64_t cover;         // sp holds every odd prime <= cover
    uint32_t nsp;           // entries in sp
    uint32_t n_med;         // medium primes: sp[17 .. n_med) are split into items
    uint32_t k_wave;        // wheel: sp[17 .. k_wave) use wave items (lanes interleave one run)
};

// Window geometry for a window starting at even n0 and covering `want` bits.
struct WinGeom {
    uint64_t qv0;   // q at bit `lead` (odd, >= 3)
    uint32_t lead;  // bits [0, lead) stand for q < 3 and are zero
    uint32_t nbits; // materialised bits (clipped so that q <= 2^64-1)
};

__host__ __device__ inline WinGeom window_geom(uint64_t n0, uint64_t want) {
    WinGeom g;
    if (n0 >= 2ull * HALO + 4) {
        g.qv0 = n0 - 2ull * HALO - 1;
        g.lead = 0;
    } else {
        g.lead = (uint32_t)(HALO + 2 - n0 / 2);
        g.qv0 = 3;
    }
    // last q = qv0 + 2(nbits - 1 - lead) must stay <= 2^64 - 1
    uint64_t room = (~0ull - g.qv0) / 2 + 1 + g.lead;
    g.nbits = (uint32_t)(want < room ? want : room);
    return g;
}

__host__ __device__ inline uint64_t geom_q_top(const WinGeom& g) {
    return g.qv0 + 2ull * (g.nbits - 1 - g.lead);
}

// ---------------------------------------------------------------------------
// Exact arithmetic helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t isqrt64(uint64_t x) {
    uint64_t r = (uint64_t)sqrt((double)x);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > x) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= x) ++r;
    return r;
}

// x mod p for odd p < 2^22 (inv = 1/p, c32 = 2^32 mod p).  Two fp64 steps,
// each on an integer below 2^53, each corrected by one compare.
__device__ __forceinline__ uint32_t mod_small(uint64_t x, uint32_t p, double inv, uint32_t c32) {
    const double pd = (double)p;
    double d = (double)(uint32_t)(x >> 32);
    double q = floor(d * inv);
    double r = fma(-q, pd, d);
    r = r < 0.0 ? r + pd : (r >= pd ? r - pd : r);
    double y = fma(r, (double)c32, (double)(uint32_t)x);
    q = floor(y * inv);
    r = fma(-q, pd, y);
    r = r < 0.0 ? r + pd : (r >= pd ? r - pd : r);
    return (uint32_t)r;
}

// (rbase + d) mod p from r0 = rbase mod p, for d < 2^52: y = d + r0 < 2^53 is
// exact in fp64 and one corrected quotient estimate finishes the reduction.
__device__ __forceinline__ uint32_t mod_from_base(uint64_t d, uint32_t r0, uint32_t p, double inv) {
    const double pd = (double)p;
    const double y = (double)d + (double)r0;
    const double q = floor(y * inv);
    double r = fma(-q, pd, y);
    r = r < 0.0 ? r + pd : (r >= pd ? r - pd : r);
    return (uint32_t)r;
}

// (rbase + d) mod p as above with the fp64 value of d precomputed (uniform per
// tile); the fix-ups run on integers: the estimate leaves r in [-p, 2p).
__device__ __forceinline__ uint32_t mod_base_u32(double df, uint32_t r0, uint32_t p, double inv) {
    const double y = df + (double)r0;
    const double q = floor(y * inv);
    uint32_t u = (uint32_t)(int32_t)fma(-q, (double)p, y);
    u = min(u, u + p);   // r < 0: u wrapped high, u + p is the residue
    u = min(u, u - p);   // r >= p: u - p is the residue; else u - p wraps high
    return u;
}

// First bit index >= lb holding an odd multiple of p that is >= p^2, given
// r = q_lo mod p for the odd q_lo at bit lb.
__device__ __forceinline__ uint64_t first_mark_r(uint64_t q_lo, uint32_t r, uint32_t lb, uint32_t p) {
    uint64_t d = r ? (uint64_t)(p - r) : 0;
    if (d & 1) d += p;                 // q_lo + d odd
    const uint64_t pp = (uint64_t)p * p;
    if (q_lo < pp && q_lo + d < pp) return lb + (pp - q_lo) / 2;
    return lb + d / 2;
}

// First bit index >= lb holding an odd multiple m of p with m >= p^2, where
// bit lb holds the odd q_lo.  Returns an index; the caller bounds it.
__device__ __forceinline__ uint64_t first_mark(uint64_t q_lo, uint32_t lb, uint32_t p, double inv, uint32_t c32) {
    uint32_t r = mod_small(q_lo, p, inv, c32);
    uint64_t d = r ? (uint64_t)(p - r) : 0;
    if (d & 1) d += p;                 // q_lo + d odd
    uint64_t pp = (uint64_t)p * p;
    if (q_lo < pp && q_lo + d < pp) return lb + (pp - q_lo) / 2;
    return lb + d / 2;
}

// 64-bit funnel: bits [s, s+64) of the 128-bit value hi:lo, s uniform in [0, 64).
__device__ __forceinline__ uint64_t funnel64(uint64_t lo, uint64_t hi, uint32_t s) {
    uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
    uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
    uint32_t r0, r1;
    if (s < 32) {
        r0 = __builtin_amdgcn_alignbit(l1, l0, s);
        r1 = __builtin_amdgcn_alignbit(h0, l1, s);
    } else {
        r0 = __builtin_amdgcn_alignbit(h0, l1, s - 32);
        r1 = __builtin_amdgcn_alignbit(h1, h0, s - 32);
    }
    return ((uint64_t)r1 << 32) | r0;
}

// ---------------------------------------------------------------------------
// Montgomery arithmetic modulo odd n < 2^64, R = 2^64 (Montgomery 1985).
// REDC in subtractive form: with m = lo * n^{-1} mod R, (ab - mn)/R = hi - mulhi(m, n)
// lies in (-n, n), so one conditional add of n finishes the reduction and no
// 65-bit intermediate ever appears.
// ---------------------------------------------------------------------------
struct Mont {
    uint64_t n;
    uint64_t ninv;  // n^{-1} mod 2^64
    uint64_t one;   // R mod n
};

// The products are written on 32-bit limbs so that each is one v_mad_u64_u32 and the low
// half of ab is formed once: m = lo(ab) * n^{-1} takes three word products and the high half
// of mn four, since its low half equals lo(ab) (left to itself the compiler rebuilds lo(ab)
// inside m = a * (b * n^{-1})).
//
// LAZY (only for n < 2^62): operands and results lie in [0, 2n) instead of [0, n).  Then
// ab < 4n^2 < Rn, so (ab - mn)/R still lies in (-n, n), and adding n unconditionally leaves
// it in (0, 2n): the compare and the two selects of the exact form drop out of every product.
//
// With an offset c in [n, 3n] and operands in [0, 4n) (n < 2^60, so ab < 16n^2 < Rn),
// (ab - mn)/R + c lies in (0, 4n): c = 3n - x subtracts x < n inside the product.
template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_redc(uint64_t hi, uint32_t t0, uint32_t t1, const Mont& M, uint64_t off) {
    const uint32_t n0 = (uint32_t)M.n, n1 = (uint32_t)(M.n >> 32);
    const uint32_t i0 = (uint32_t)M.ninv, i1 = (uint32_t)(M.ninv >> 32);
    const uint64_t q = (uint64_t)t0 * i0;
    const uint32_t m0 = (uint32_t)q, m1 = (uint32_t)(q >> 32) + t0 * i1 + t1 * i0;
    const uint64_t u = (uint64_t)m1 * n0 + __umulhi(m0, n0);     // the low word of m0 n0 is t0
    const uint64_t v = (uint64_t)m0 * n1 + (uint32_t)u;
    const uint64_t mh = (uint64_t)m1 * n1 + (u >> 32) + (v >> 32);
    if constexpr (LAZY) return hi - mh + off;
    const uint64_t t = hi - mh;
    return hi < mh ? t + M.n : t;
}

template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_redc(uint64_t hi, uint32_t t0, uint32_t t1, const Mont& M) {
    return mont_redc<LAZY>(hi, t0, t1, M, M.n);
}

template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const Mont& M, uint64_t off) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t x = (uint64_t)a1 * b0 + (p00 >> 32);
    const uint64_t y = (uint64_t)a0 * b1 + (uint32_t)x;
    const uint64_t hi = (uint64_t)a1 * b1 + (x >> 32) + (y >> 32);
    return mont_redc<LAZY>(hi, (uint32_t)p00, (uint32_t)y, M, off);
}

template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_sqr(uint64_t a, const Mont& M, uint64_t off) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint64_t p00 = (uint64_t)a0 * a0;
    const uint64_t c = (uint64_t)a0 * a1;
    const uint64_t y = (uint64_t)(uint32_t)(c << 1) + (p00 >> 32);
    const uint64_t hi = (uint64_t)a1 * a1 + (c >> 31) + (y >> 32);
    return mont_redc<LAZY>(hi, (uint32_t)p00, (uint32_t)y, M, off);
}

template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const Mont& M) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t x = (uint64_t)a1 * b0 + (p00 >> 32);
    const uint64_t y = (uint64_t)a0 * b1 + (uint32_t)x;
    const uint64_t hi = (uint64_t)a1 * b1 + (x >> 32) + (y >> 32);
    return mont_redc<LAZY>(hi, (uint32_t)p00, (uint32_t)y, M);
}

// a^2: the cross product a0 a1 once, doubled
template <bool LAZY = false>
__device__ __forceinline__ uint64_t mont_sqr(uint64_t a, const Mont& M) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint64_t p00 = (uint64_t)a0 * a0;
    const uint64_t c = (uint64_t)a0 * a1;
    const uint64_t y = (uint64_t)(uint32_t)(c << 1) + (p00 >> 32);
    const uint64_t hi = (uint64_t)a1 * a1 + (c >> 31) + (y >> 32);
    return mont_redc<LAZY>(hi, (uint32_t)p00, (uint32_t)y, M);
}

constexpr uint64_t MONT_LAZY_MAX = 1ull << 62;    // [0, 2n) needs 4n <= R
constexpr uint64_t MONT_LAZY4_MAX = 1ull << 60;   // [0, 4n) needs 16n <= R

// x in [0, 2n) to [0, n)
__device__ __forceinline__ uint64_t mont_canon(uint64_t x, uint64_t n) { return x >= n ? x - n : x; }

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t n) {
    // a, b < n: (a + b) mod n without 65-bit overflow
    uint64_t s = a + b;
    return (s < a || s >= n) ? s - n : s;
}

// 2^64 mod n for odd n > 1.  From 2^40 up the quotient Q = floor(2^64 / n) is below 2^24,
// so the rounded fp64 quotient q0 is Q - 1, Q or Q + 1 (relative error under 2^-52); with
// q = max(q0, 2) - 1 <= Q the remainder 2^64 - q n is exact in 64 bits and below 3n.
__device__ __forceinline__ uint64_t pow64_mod(uint64_t n) {
    if (n < (1ull << 40)) return (0ull - n) % n;
    if (n > (1ull << 63)) return 0ull - n;
    const uint64_t q0 = (uint64_t)(18446744073709551616.0 / (double)n);
    uint64_t r = 0ull - (max(q0, (uint64_t)2) - 1) * n;
    r = r >= n ? r - n : r;
    return r >= n ? r - n : r;
}

__device__ __forceinline__ Mont mont_setup(uint64_t n) {
    Mont M;
    M.n = n;
    uint64_t x = n;                    // n*n == 1 mod 8 for odd n: 3 correct bits
    for (int i = 0; i < 5; ++i) x *= 2 - n * x;   // 6, 12, 24, 48, 96 bits
    M.ninv = x;
    M.one = pow64_mod(n);              // 2^64 mod n
    return M;
}

// Strong probable-prime test to base a (0 < a < n, a < 2^32), n odd > 2, n - 1 = d * 2^r.
__device__ __forceinline__ bool sprp(uint64_t a, uint64_t d, uint32_t r, const Mont& M) {
    // a in Montgomery form: a * R mod n by double-and-add over a's bits
    uint64_t am = 0;
    for (int b = 31 - __clz((uint32_t)a); b >= 0; --b) {
        am = add_mod(am, am, M.n);
        if ((a >> b) & 1) am = add_mod(am, M.one, M.n);
    }
    uint64_t x = am;
    int top = 63 - __clzll((long long)d);
    for (int b = top - 1; b >= 0; --b) {
        x = mont_sqr(x, M);
        if ((d >> b) & 1) x = mont_mul(x, am, M);
    }
    const uint64_t minus_one = M.n - M.one;
    if (x == M.one || x == minus_one) return true;
    for (uint32_t i = 1; i < r; ++i) {
        x = mont_sqr(x, M);
        if (x == minus_one) return true;
        if (x == M.one) return false;
    }
    return false;
}

// 2^d in Montgomery form by a left-to-right ladder whose multiply step is a doubling, the
// result in [0, n).  MODE 0: exact products.  MODE 1 (n < 2^62): the ladder runs in
// [0, 2n).  MODE 2 (n < 2^60): it runs in [0, 4n), where a product lands in (0, 2n) and
// its doubling needs no reduction.
template <int MODE>
__device__ __forceinline__ uint64_t pow2_ladder(uint64_t d, const Mont& M) {
    constexpr bool LAZY = MODE != 0;
    const uint64_t md = LAZY ? 2 * M.n : M.n;   // MODE 0, 1: the ladder's values lie in [0, md)
    uint64_t x = add_mod(M.one, M.one, M.n);    // 2 in Montgomery form
    const int top = 63 - __clzll((long long)d);
    for (int b = top - 1; b >= 0; --b) {
        x = mont_sqr<LAZY>(x, M);
        if constexpr (MODE == 2) {
            if ((d >> b) & 1) x += x;
        } else {
            if ((d >> b) & 1) x = add_mod(x, x, md);
        }
    }
    if constexpr (MODE == 2) x = mont_canon(x, 2 * M.n);
    return LAZY ? mont_canon(x, M.n) : x;
}

// Strong test to base 2.
__device__ __forceinline__ bool sprp2(uint64_t d, uint32_t r, const Mont& M) {
    uint64_t x = M.n < MONT_LAZY4_MAX  ? pow2_ladder<2>(d, M)
                 : M.n < MONT_LAZY_MAX ? pow2_ladder<1>(d, M)
                                       : pow2_ladder<0>(d, M);
    const uint64_t minus_one = M.n - M.one;
    if (x == M.one || x == minus_one) return true;
    for (uint32_t i = 1; i < r; ++i) {
        x = mont_sqr(x, M);
        if (x == minus_one) return true;
        if (x == M.one) return false;
    }
    return false;
}

// Miller-Rabin bases.  Below psi_7 = 341550071728321: the prime bases 2, 3, 5, ... in
// order, the first k of which are exact below psi_k, the smallest odd composite
// passing all k (OEIS A014233).  From psi_7 up: Sinclair's seven bases 2, 325, 9375,
// 28178, 450775, 9780504, 1795265022, exact for every n < 2^64 (J. Sinclair, 2011; every
// base is below n there, so none reduces to 0).  Both give the answer of the twelve prime
// bases 2..37 (Sorenson & Webster 2015).
constexpr uint64_t MR_PSI7 = 341550071728321ull;

__device__ __forceinline__ uint32_t mr_base(uint32_t k, bool wide) {
    if (wide) {
        constexpr uint32_t S[7] = {2u, 325u, 9375u, 28178u, 450775u, 9780504u, 1795265022u};
        return S[k];
    }
    return k < 6 ? (0x0D0B07050302ull >> (8 * k)) & 0xFF : (0x251F1D171311ull >> (8 * (k - 6))) & 0xFF;
}

__host__ __device__ inline bool mr_wide(uint64_t n) { return n >= MR_PSI7; }

__host__ __device__ inline uint32_t mr_bases(uint64_t n) {
    if (n < 2047ull) return 1;
    if (n < 1373653ull) return 2;
    if (n < 25326001ull) return 3;
    if (n < 3215031751ull) return 4;
    if (n < 2152302898747ull) return 5;
    if (n < 3474749660383ull) return 6;
    return 7;
}

// Strong test of odd n > 37 to base k of the set `wide` selects (k = 0: base 2).
__device__ inline bool mr_test(uint64_t n, uint32_t k, bool wide) {
    const Mont M = mont_setup(n);
    uint64_t d = n - 1;
    const uint32_t r = __builtin_ctzll(d);
    d >>= r;
    return k == 0 ? sprp2(d, r, M) : sprp(mr_base(k, wide), d, r, M);
}

// ---------------------------------------------------------------------------
// Strong Lucas test (Baillie & Wagstaff 1980, Selfridge's parameters: the first D
// of 5, -7, 9, -11, ... with Jacobi (D/n) = -1, P = 1, Q = (1 - D)/4).  With the
// base-2 strong test it is the Baillie-PSW test, which no composite below 2^64
// passes (Gilchrist 2013, over Feitsma's list of base-2 strong pseudoprimes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t n) {
    return a >= b ? a - b : a + (n - b);
}

__device__ __forceinline__ uint64_t half_mod(uint64_t x, uint64_t n) {   // x / 2 mod n, x < n, n odd
    return (x & 1) ? (x >> 1) + (n >> 1) + 1 : x >> 1;
}

// Jacobi symbol (a/m), m odd > 0
__device__ inline int jacobi32(uint32_t a, uint32_t m) {
    int t = 1;
    a %= m;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            if ((m & 7) == 3 || (m & 7) == 5) t = -t;
        }
        const uint32_t x = a;
        a = m;
        m = x;
        if ((a & 3) == 3 && (m & 3) == 3) t = -t;
        a %= m;
    }
    return m == 1 ? t : 0;
}

// x mod c for c < 2^16, with c32 = 2^32 mod c
__device__ __forceinline__ uint32_t mod_u64_small(uint64_t x, uint32_t c, uint32_t c32) {
    return (((uint32_t)(x >> 32) % c) * c32 + (uint32_t)x % c) % c;
}

// a^{-1} mod m for odd m > 1 by the extended Euclidean algorithm; 0 when gcd(a, m) > 1
__device__ inline uint32_t inv_mod32(uint32_t a, uint32_t m) {
    int32_t r0 = (int32_t)m, r1 = (int32_t)a, x0 = 0, x1 = 1;
    while (r1) {
        const int32_t t = r0 / r1, r2 = r0 - t * r1, x2 = x0 - t * x1;
        r0 = r1; r1 = r2; x0 = x1; x1 = x2;
    }
    if (r0 != 1) return 0;
    return (uint32_t)(x0 < 0 ? x0 + (int32_t)m : x0);
}

// Jacobi symbols (a/m) of a table of small odd m as bit masks over a < m, built by the
// compiler from jacobi_c (the same algorithm as jacobi32).
__host__ __device__ constexpr int jacobi_c(uint32_t a, uint32_t m) {
    int t = 1;
    a %= m;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            if ((m & 7) == 3 || (m & 7) == 5) t = -t;
        }
        const uint32_t x = a;
        a = m;
        m = x;
        if ((a & 3) == 3 && (m & 3) == 3) t = -t;
        a %= m;
    }
    return m == 1 ? t : 0;
}

__host__ __device__ constexpr uint64_t jacobi_mask(uint32_t m, int v) {
    uint64_t s = 0;
    for (uint32_t a = 0; a < m; ++a)
        if (jacobi_c(a, m) == v) s |= 1ull << a;
    return s;
}

// x^{-1} mod c for x = 1..c-1 (c <= 15) as 4-bit fields at 4x, 0 where no inverse exists
__host__ __device__ constexpr uint64_t inv_nibbles(uint32_t c) {
    uint64_t s = 0;
    for (uint32_t x = 1; x < c; ++x)
        for (uint32_t y = 1; y < c; ++y)
            if (x * y % c == 1) s |= (uint64_t)y << (4 * x);
    return s;
}

// c^{-1} mod 2^64 for odd c (Newton: each step doubles the correct low bits)
__host__ __device__ constexpr uint64_t inv_pow2_64(uint64_t c) {
    uint64_t x = c;
    for (int i = 0; i < 6; ++i) x *= 2 - c * x;
    return x;
}
static_assert(inv_nibbles(7) == 0x6325410ull && inv_pow2_64(3) * 3 == 1 && inv_pow2_64(11) * 11 == 1, "inverse tables");

// Selfridge's candidates D = 5, -7, 9, ..., -47 (|D| <= 47).  Each is 1 mod 4, so by
// reciprocity (D/n) = (n mod |D| / |D|) for every odd n; n mod |D| comes from n modulo one
// of 218243025 = 3^3 5^2 7 11 13 17 19, 31367009 = 23 29 31 37 41 and 2021 = 43 47.
struct SelfridgeD {
    int32_t D;
    uint32_t m, group;
    uint64_t minus, zero;   // a with (a/m) = -1, with (a/m) = 0
};
__host__ __device__ constexpr SelfridgeD selfridge_d(int32_t D, uint32_t group) {
    const uint32_t m = (uint32_t)(D < 0 ? -D : D);
    return SelfridgeD{D, m, group, jacobi_mask(m, -1), jacobi_mask(m, 0)};
}
constexpr int SELFRIDGE_N = 22;
constexpr SelfridgeD SELFRIDGE[SELFRIDGE_N] = {
    selfridge_d(5, 0),   selfridge_d(-7, 0),  selfridge_d(9, 0),   selfridge_d(-11, 0), selfridge_d(13, 0),
    selfridge_d(-15, 0), selfridge_d(17, 0),  selfridge_d(-19, 0), selfridge_d(21, 0),  selfridge_d(-23, 1),
    selfridge_d(25, 0),  selfridge_d(-27, 0), selfridge_d(29, 1),  selfridge_d(-31, 1), selfridge_d(33, 0),
    selfridge_d(-35, 0), selfridge_d(37, 1),  selfridge_d(-39, 0), selfridge_d(41, 1),  selfridge_d(-43, 2),
    selfridge_d(45, 0),  selfridge_d(-47, 2)};
static_assert(SELFRIDGE[0].minus == 0xC && SELFRIDGE[0].zero == 0x1, "(a/5) = -1 for a = 2, 3");
static_assert(SELFRIDGE[1].minus == 0x68 && SELFRIDGE[2].minus == 0 && SELFRIDGE[2].zero == 0x49, "(a/7), (a/9)");
static_assert(SELFRIDGE[5].minus == 0x6880 && SELFRIDGE[5].zero == 0x1669, "(a/15)");
static_assert(218243025u == 27u * 25u * 7u * 11u * 13u * 17u * 19u && 31367009u == 23u * 29u * 31u * 37u * 41u &&
              2021u == 43u * 47u, "residue groups");

// Strong Lucas probable-prime test of odd n > 2^32 (a perfect square, for which no D
// exists, is composite).  Returns 1 (passes), 0 (composite) or -1 when no D below 2^16
// turned up (the first D is small for every non-square; the caller then decides by
// Miller-Rabin).
//
// D: the table above settles nearly every n (a lane needs 2.2 candidates on average, a wave
// of 64 primes about 9 and rarely more than 20): a candidate costs a remainder by a constant
// and two mask tests, and a wave leaves the table once every lane is decided.  Past |D| = 47
// the search goes on with jacobi32; a square never meets (D/n) = -1, so it is only looked
// for there.
//
// The ladder runs on W_k = V_{2k} / Q^k, the Lucas V sequence with parameters (P', 1),
// P' = P^2/Q - 2 = 1/Q - 2 mod n (the roots of x^2 - x + Q squared and divided by Q have
// product 1 and sum P').  With n + 1 = d 2^s, d odd, m = (d - 1)/2, and D, Q units mod n:
//   U_d = 0  <=>  W_{m+1} = W_m        (D U_d = 2 V_{d+1} - V_d,  V_d = V_{d+1} + Q V_{d-1})
//   V_d = 0  <=>  W_{m+1} + W_m = 0
//   V_{d 2^r} = 0  <=>  W_{d 2^{r-1}} = 0   for 1 <= r < s.
// A step of a Q = 1 sequence, (W_k, W_{k+1}) -> (W_{2k}, W_{2k+1}) or (W_{2k+1}, W_{2k+2}),
// is one product and one square whichever the bit (W_{2k} = W_k^2 - 2, W_{2k+1} = W_k W_{k+1}
// - P'), so lanes whose bits differ do not diverge.
// (W_k, W_{k+1}) -> (W_{2k}, W_{2k+1}) or (W_{2k+1}, W_{2k+2}) over bits b..0 of m, the
// values in [0, n) on entry and return.  MODE 0: exact products.  MODE 1 (n < 2^62): the
// ladder runs in [0, 2n).  MODE 2 (n < 2^60): it runs in [0, 4n) and the subtractions of
// P' and 2 ride in the products' offsets.
template <int MODE>
__device__ __forceinline__ void w_ladder(uint64_t m, int b, uint64_t& w0, uint64_t& w1, uint64_t pp, uint64_t two,
                                         const Mont& M) {
    if constexpr (MODE == 2) {
        const uint64_t c1 = 3 * M.n - pp, c2 = 3 * M.n - two;
        for (; b >= 0; --b) {
            const bool bit = (m >> b) & 1;
            const uint64_t x = mont_mul<true>(w0, w1, M, c1);                 // W_{2k+1}
            const uint64_t h = bit ? w1 : w0;
            const uint64_t y = mont_sqr<true>(h, M, c2);                      // W_{2k+2} or W_{2k}
            w0 = bit ? x : y;
            w1 = bit ? y : x;
        }
        w0 = mont_canon(mont_canon(w0, 2 * M.n), M.n);
        w1 = mont_canon(mont_canon(w1, 2 * M.n), M.n);
    } else {
        constexpr bool LAZY = MODE == 1;
        const uint64_t md = LAZY ? 2 * M.n : M.n;   // the ladder's values lie in [0, md)
        for (; b >= 0; --b) {
            const bool bit = (m >> b) & 1;
            const uint64_t x = sub_mod(mont_mul<LAZY>(w0, w1, M), pp, md);    // W_{2k+1}
            const uint64_t h = bit ? w1 : w0;
            const uint64_t y = sub_mod(mont_sqr<LAZY>(h, M), two, md);       // W_{2k+2} or W_{2k}
            w0 = bit ? x : y;
            w1 = bit ? y : x;
        }
        if (LAZY) {
            w0 = mont_canon(w0, M.n);
            w1 = mont_canon(w1, M.n);
        }
    }
}

__device__ inline int strong_lucas(uint64_t n, const Mont& M) {
    const uint32_t res[3] = {(uint32_t)(n % 218243025ull), (uint32_t)(n % 31367009ull), (uint32_t)(n % 2021ull)};
    int64_t D = 0;                                                              // 0: undecided
#pragma unroll
    for (int k = 0; k < SELFRIDGE_N; ++k) {
        if (D == 0) {
            const uint32_t a = res[SELFRIDGE[k].group] % SELFRIDGE[k].m;
            if ((SELFRIDGE[k].zero >> a) & 1) return 0;                          // |D| < n shares a factor
            if ((SELFRIDGE[k].minus >> a) & 1) D = SELFRIDGE[k].D;
        }
    }
    const bool D_from_table = D != 0;
    if (D == 0) {
        const uint64_t r = isqrt64(n);
        if (r * r == n) return 0;
        D = 49;
        for (;;) {
            const uint32_t ad = (uint32_t)(D < 0 ? -D : D);
            if (ad >= 65536) return -1;
            const uint32_t c32 = (0xFFFFFFFFu % ad + 1) % ad;                   // 2^32 mod |D|
            const int j = jacobi32(mod_u64_small(n, ad, c32), ad);             // = (D/n): D = 1 mod 4
            if (j == -1) break;
            if (j == 0) return 0;
            D = D > 0 ? -D - 2 : -D + 2;
        }
    }
    // 1/Q in Montgomery form: halve R mod n for the power of two in |Q|, then divide by the
    // odd part c exactly, (x + t n)/c with x + t n = 0 mod c, as a product with c^{-1} mod 2^64
    const int64_t Q = (1 - D) / 4;
    const uint32_t aq = (uint32_t)(Q < 0 ? -Q : Q);
    const uint32_t e = __builtin_ctz(aq), c = aq >> e;
    uint64_t qi = M.one;
    for (uint32_t i = 0; i < e; ++i) qi = half_mod(qi, M.n);
    if (c > 1 && D_from_table) {
        // |D| <= 47: c is 3, 5, 7, 9 or 11, all dividing 218243025; the inverses mod c and
        // mod 2^64 come from compile-time constants, selected without a branch
        const uint64_t tab = c == 3 ? inv_nibbles(3) : c == 5 ? inv_nibbles(5) : c == 7 ? inv_nibbles(7)
                           : c == 9 ? inv_nibbles(9) : inv_nibbles(11);
        const uint64_t ci = c == 3 ? inv_pow2_64(3) : c == 5 ? inv_pow2_64(5) : c == 7 ? inv_pow2_64(7)
                          : c == 9 ? inv_pow2_64(9) : inv_pow2_64(11);
        const uint32_t inv = (uint32_t)(tab >> (4 * (res[0] % c))) & 15u;
        if (!inv) return 0;                                                     // Q shares a factor with n
        const uint32_t t = (c - (uint32_t)(qi % 218243025ull) % c * inv % c) % c;
        qi = (qi + (uint64_t)t * M.n) * ci;
    } else if (c > 1) {
        const uint32_t c32 = (0xFFFFFFFFu % c + 1) % c;
        const uint32_t inv = inv_mod32(mod_u64_small(n, c, c32), c);
        if (!inv) return 0;                                                     // Q shares a factor with n
        const uint32_t t = (c - mod_u64_small(qi, c, c32) * inv % c) % c;
        uint64_t ci = c;                                                        // c^{-1} mod 2^64
        for (int i = 0; i < 5; ++i) ci *= 2 - (uint64_t)c * ci;
        qi = (qi + (uint64_t)t * M.n) * ci;
    }
    if (Q < 0 && qi) qi = M.n - qi;
    const uint64_t two = add_mod(M.one, M.one, M.n);
    const uint64_t pp = sub_mod(qi, two, M.n);                                  // P'
    uint64_t d = n + 1;
    const uint32_t s = __builtin_ctzll(d);
    const uint64_t m = d >> (s + 1);
    uint64_t w0 = two, w1 = pp;                                                 // W_0, W_1
    int b = -1;
    if (m) {                                                                    // top bit: W_1, W_2
        w0 = pp;
        w1 = sub_mod(mont_sqr(pp, M), two, M.n);
        b = 62 - __clzll((long long)m);
    }
    if (M.n < MONT_LAZY4_MAX)
        w_ladder<2>(m, b, w0, w1, pp, two, M);
    else if (M.n < MONT_LAZY_MAX)
        w_ladder<1>(m, b, w0, w1, pp, two, M);
    else
        w_ladder<0>(m, b, w0, w1, pp, two, M);
    if (w0 == w1 || add_mod(w0, w1, M.n) == 0) return 1;                       // U_d = 0 or V_d = 0
    uint64_t w = sub_mod(mont_mul(w0, w1, M), pp, M.n);                         // W_d
    for (uint32_t i = 1; i < s; ++i) {
        if (w == 0) return 1;                                                   // V_{d 2^i} = 0
        w = sub_mod(mont_sqr(w, M), two, M.n);
    }
    return 0;
}

// The Lucas half of Baillie-PSW for odd n > 2^32 that passed the base-2 strong test.
__device__ inline bool bpsw_lucas(uint64_t n) {
    const Mont M = mont_setup(n);
    const int r = strong_lucas(n, M);
    if (r >= 0) return r == 1;
    uint64_t d = n - 1;
    const uint32_t s = __builtin_ctzll(d);
    d >>= s;
    for (uint32_t k = 1; k < 7; ++k)
        if (!sprp(mr_base(k, true), d, s, M)) return false;
    return true;
}

// Deterministic for every 64-bit n.  Small n (< 41) are answered directly;
// n even is composite.
__device__ inline bool is_prime_u64(uint64_t n) {
    if (n < 41) return (0x20A08A28ACull >> n) & 1;   // primes 2..37 as a mask
    if ((n & 1) == 0) return false;
    if (n % 3 == 0 || n % 5 == 0 || n % 7 == 0) return false;
    Mont M = mont_setup(n);
    uint64_t d = n - 1;
    uint32_t r = __builtin_ctzll(d);
    d >>= r;
    if (!sprp2(d, r, M)) return false;
    const uint32_t nb = mr_bases(n);
    const bool wide = mr_wide(n);
#pragma unroll 1
    for (uint32_t i = 1; i < nb; ++i)
        if (!sprp(mr_base(i, wide), d, r, M)) return false;
    return true;
}

// Exact primality using the resident p2bits table when q is small enough.
__device__ __forceinline__ bool is_prime_q(uint64_t q, const Tables& T) {
    if (q <= T.p2_limit) {
        if (q < 3) return q == 2;
        if ((q & 1) == 0) return false;
        uint64_t b = (q - 3) >> 1;
        return (T.p2bits[b >> 6] >> (b & 63)) & 1;
    }
    return is_prime_u64(q);
}

// Per-block statistics record, merged by the reduce kernel.
struct StatRec {
    uint64_t phase2;
    uint64_t failures;
    uint64_t first_fail;   // UINT64_MAX when none
    uint64_t sum_pmin;
    uint64_t max_pmin;
    uint64_t argmax;       // UINT64_MAX when none
    uint64_t evens;
    uint64_t pad;
};

}  // namespace gbk
