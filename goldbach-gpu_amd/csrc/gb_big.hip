// gb_big.hip -- strong probable-prime tests of q = n - p for even n of any size
// (up to BIG_LMAX 32-bit words), one workgroup per (p, base) task.
//
// The reference's big_check (big_check.cpp:83-118) asks GMP whether q = n - p
// is prime for p = 2, 3, 5, ... and stops at the first yes.  Here a
// workgroup (1024 threads; 256 below 384 words) holds q and its Montgomery constants in LDS and
// runs one strong test (Miller-Rabin with a fixed base) on it:
//   * products as column sums: thread (wave w, lane l) owns columns
//     64 g + l for the groups g = w, w + waves, ...; a column adds a_i * b_(k-i)
//     over i in 96 bits (one v_mad_u64_u32 and a 64-bit add per term, four
//     independent accumulators, a_i read four words at a time);
//     a_i is the same word for the whole wave, b_(k-i) 64 consecutive words,
//     and 64 zero words either side of every multiplicand keep all lanes of
//     a wave on one loop range;
//   * carries resolved by blocks of consecutive words per thread, the
//     carry into each block taken from a workgroup scan over the blocks'
//     carry functions c -> a + [c >= theta] (closed under composition);
//   * Montgomery reduction by separate products: M = (T mod R) * (-q^-1 mod R)
//     mod R, U = (T + M q) / R, R = 2^(32 L), with -q^-1 mod R built by
//     Newton doubling from the word inverse;
//   * R mod q without a division: q's top bit t - 1 gives 2^(t-1) < q, and
//     32 L - t + 1 modular doublings lift it to 2^(32 L);
//   * base 2 multiplies by doubling, other bases b < 128 by one word-by-b
//     product with its quotient by q read off the top words (no Montgomery
//     product); b R mod q, the start value, comes from R mod q by doublings
//     and additions.
#include <hip/hip_runtime.h>

#include "gb_big.hpp"

namespace gbk {
namespace {

constexpr uint32_t NW_MAX = BIG_BS / 64;  // waves of the largest workgroup (scan slots)

// threads and waves of this launch (256 or 1024: the host picks by the size of n)
__device__ __forceinline__ uint32_t bw() { return blockDim.x; }
__device__ __forceinline__ uint32_t nw() { return blockDim.x >> 6; }
constexpr uint32_t GUARD = 64;            // zero words either side of a multiplicand
constexpr uint32_t INF = 0xFFFFFFFFu;     // carry-function threshold never reached

// LDS word offsets for a q of L words.
struct Layout {
    uint32_t x, y, q, qi, t, m, s0, s1, s2, d, e, rq, qmr, scan, total;
};

__host__ __device__ inline uint32_t round4(uint32_t w) { return (w + 3) & ~3u; }

// Every buffer starts on a 16-byte boundary (multiplicands are read four words at a time).
__host__ __device__ inline Layout layout(uint32_t L) {
    Layout o{};
    uint32_t w = 0;
    const uint32_t gl = round4(L) + 2 * GUARD;
    o.x = w + GUARD; w += gl;
    o.y = w + GUARD; w += gl;
    o.q = w + GUARD; w += gl;
    o.qi = w + GUARD; w += gl;
    o.t = w; w += round4(2 * L);
    o.m = w; w += round4(L);
    o.s0 = w; w += round4(2 * L);
    o.s1 = w; w += round4(2 * L);
    o.s2 = w; w += round4(2 * L);
    o.d = w; w += round4(L + 1);
    o.e = w; w += round4(L + 1);
    o.rq = w; w += round4(L);
    o.qmr = w; w += round4(L);
    o.scan = w; w += 4 * NW_MAX + 8;        // wave totals, top carry, t, s, C, qhat
    o.total = w;
    return o;
}

// carry function c -> a + [c >= th]; compose(f, g) = g after f
struct CF {
    uint32_t a, th;
};

__device__ inline CF compose(CF f, CF g) {
    CF r;
    r.a = g.a + (f.a >= g.th ? 1u : 0u);
    r.th = (f.a + 1 == g.th) ? f.th : INF;
    return r;
}

__device__ inline uint32_t apply(CF f, uint32_t c) { return f.a + (c >= f.th ? 1u : 0u); }

// The carry function of the lane that CTRL (a DPP control) names, or a marked one (a = MARK)
// where that lane does not exist or ROW_MASK leaves the row out.
constexpr uint32_t MARK = 0xFFFFFFFFu;   // no carry function has a this large

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ CF dpp_cf(CF f) {
    CF g;
    g.a = (uint32_t)__builtin_amdgcn_update_dpp((int)MARK, (int)f.a, CTRL, ROW_MASK, 0xF, false);
    g.th = (uint32_t)__builtin_amdgcn_update_dpp((int)INF, (int)f.th, CTRL, ROW_MASK, 0xF, false);
    return g;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(CF& inc) {
    const CF g = dpp_cf<CTRL, ROW_MASK>(inc);
    const CF c = compose(g, inc);
    if (g.a != MARK) inc = c;
}

// Workgroup scan of the threads' carry functions over blocks 0..nb-1 (thread t = block t).
// Returns the carry into this thread's block; *top gets the carry out of block nb-1.
// Within a wave: DPP row shifts 1, 2, 4, 8, then row broadcasts 15 and 31 (no LDS).
__device__ uint32_t scan_carry(CF f, uint32_t nb, uint32_t* sc, uint32_t* top) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    CF inc = f;
    scan_step<0x111, 0xF>(inc);    // row_shr:1
    scan_step<0x112, 0xF>(inc);    // row_shr:2
    scan_step<0x114, 0xF>(inc);    // row_shr:4
    scan_step<0x118, 0xF>(inc);    // row_shr:8
    scan_step<0x142, 0xA>(inc);    // row_bcast:15 into rows 1 and 3
    scan_step<0x143, 0xC>(inc);    // row_bcast:31 into rows 2 and 3
    const CF exc = dpp_cf<0x138, 0xF>(inc);   // wave_shr:1: the lane before's inclusive function
    if (lane == 63) {
        sc[2 * wv] = inc.a;
        sc[2 * wv + 1] = inc.th;
    }
    __syncthreads();
    uint32_t c = 0;                                  // carry into this wave
    for (uint32_t v = 0; v < wv; ++v) c = apply(CF{sc[2 * v], sc[2 * v + 1]}, c);
    const uint32_t cin = lane == 0 ? c : apply(exc, c);
    if (tid == nb - 1) sc[2 * NW_MAX] = apply(f, cin);
    __syncthreads();
    *top = sc[2 * NW_MAX];
    return cin;
}

struct Big {
    uint32_t L;
    uint32_t *x, *y, *q, *qi, *t, *m, *s0, *s1, *s2, *d, *e, *rq, *qmr, *sc;
};

// (acc, hi) += a b: one v_mad_u64_u32 with its carry-out counted in hi
__device__ __forceinline__ void mac(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(hi)
        : "v"(a), "v"(b)
        : "vcc");
}

// The column value acc_0 + ... + acc_3 + 2^64 (hi_0 + ... + hi_3) as three words, optionally doubled.
__device__ __forceinline__ void put_col(const Big& b, int k, uint64_t a0, uint64_t a1, uint64_t a2, uint64_t a3,
                                        uint32_t h0, uint32_t h1, uint32_t h2, uint32_t h3, bool dbl) {
    uint64_t lo = a0, hi = (uint64_t)h0 + h1 + h2 + h3;
    lo += a1;
    hi += lo < a1;
    lo += a2;
    hi += lo < a2;
    lo += a3;
    hi += lo < a3;
    if (dbl) {
        hi = (hi << 1) | (lo >> 63);
        lo <<= 1;
    }
    b.s0[k] = (uint32_t)lo;
    b.s1[k] = (uint32_t)(lo >> 32);
    b.s2[k] = (uint32_t)hi;
}

// Column sums of A (la words, 16-byte aligned) times B (lb words, guarded) for the
// columns k in [64 floor(kbeg / 64), kend), each started from add[k] when add is given;
// written as three words per column (s0 + 2^32 s1 + 2^64 s2).
__device__ void cols(const Big& b, const uint32_t* A, uint32_t la, const uint32_t* B, uint32_t lb, uint32_t kbeg,
                     uint32_t kend, const uint32_t* add) {
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t g = kbeg / 64 + wv; 64 * g < kend; g += nw()) {
        const int kmin = 64 * (int)g, kmax = kmin + 63;
        const int k = kmin + (int)lane;
        const int ilo = kmin - (int)lb + 1 > 0 ? kmin - (int)lb + 1 : 0;
        const int ihi = kmax < (int)la - 1 ? kmax : (int)la - 1;
        uint64_t a0 = (add && k < (int)kend) ? add[k] : 0u, a1 = 0, a2 = 0, a3 = 0;
        uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0;
        const uint32_t* bk = B + k;
        int i = ilo;
        for (; i <= ihi && (i & 3); ++i) mac(a0, h0, A[i], bk[-i]);
        for (; i + 7 <= ihi; i += 8) {
            const uint4 u = *reinterpret_cast<const uint4*>(A + i);
            const uint4 v = *reinterpret_cast<const uint4*>(A + i + 4);
            mac(a0, h0, u.x, bk[-i]);
            mac(a1, h1, u.y, bk[-i - 1]);
            mac(a2, h2, u.z, bk[-i - 2]);
            mac(a3, h3, u.w, bk[-i - 3]);
            mac(a0, h0, v.x, bk[-i - 4]);
            mac(a1, h1, v.y, bk[-i - 5]);
            mac(a2, h2, v.z, bk[-i - 6]);
            mac(a3, h3, v.w, bk[-i - 7]);
        }
        for (; i <= ihi; ++i) mac(a0, h0, A[i], bk[-i]);
        if (k < (int)kend) put_col(b, k, a0, a1, a2, a3, h0, h1, h2, h3, false);
    }
    __syncthreads();
}

// Column sums of X^2 (X: L words, guarded and 16-byte aligned) for all 2L columns:
// twice the terms x_i x_(k-i) with i < k - i, plus x_(k/2)^2 for even k.
__device__ void cols_sqr(const Big& b, const uint32_t* X) {
    const int L = (int)b.L;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t g = wv; 64 * g < 2 * b.L; g += nw()) {
        const int kmin = 64 * (int)g;
        const int k = kmin + (int)lane;
        const int ilo = kmin - L + 1 > 0 ? kmin - L + 1 : 0;        // lanes below their own start read zeros
        const int iend = k > 0 ? (k - 1) / 2 : -1;                    // i < k - i
        const int iend_w = kmin > 0 ? (kmin - 1) / 2 : -1;           // the wave's smallest end
        uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0;
        const uint32_t* bk = X + k;
        int i = ilo;
        for (; i <= iend_w && (i & 3); ++i) mac(a0, h0, X[i], bk[-i]);
        for (; i + 7 <= iend_w; i += 8) {
            const uint4 u = *reinterpret_cast<const uint4*>(X + i);
            const uint4 v = *reinterpret_cast<const uint4*>(X + i + 4);
            mac(a0, h0, u.x, bk[-i]);
            mac(a1, h1, u.y, bk[-i - 1]);
            mac(a2, h2, u.z, bk[-i - 2]);
            mac(a3, h3, u.w, bk[-i - 3]);
            mac(a0, h0, v.x, bk[-i - 4]);
            mac(a1, h1, v.y, bk[-i - 5]);
            mac(a2, h2, v.z, bk[-i - 6]);
            mac(a3, h3, v.w, bk[-i - 7]);
        }
        for (; i <= iend; ++i) mac(a0, h0, X[i], bk[-i]);            // the lane's own tail (< 40 terms)
        uint64_t dg = 0;
        if (!(k & 1) && k / 2 < L) dg = (uint64_t)X[k / 2] * X[k / 2];
        if (k < 2 * L) {
            put_col(b, k, a0, a1, a2, a3, h0, h1, h2, h3, true);
            // add the square term (the doubled value stays below 2^77)
            uint64_t lo = (uint64_t)b.s0[k] | ((uint64_t)b.s1[k] << 32);
            const uint64_t s = lo + dg;
            b.s0[k] = (uint32_t)s;
            b.s1[k] = (uint32_t)(s >> 32);
            b.s2[k] += s < dg;
        }
    }
    __syncthreads();
}

// Resolves the column sums of words [jlo, jhi) into out[0, jhi - jlo), adding the 64-bit
// `extra` at word jlo; spills of columns below jlo and beyond jhi - 1 are dropped.
// Returns the carry out of word jhi - 1.
__device__ uint32_t carry_range(const Big& b, uint32_t jlo, uint32_t jhi, uint32_t* out, uint64_t extra) {
    const uint32_t tid = threadIdx.x, n = jhi - jlo;
    const uint32_t B = (n + bw() - 1) / bw();
    const uint32_t nb = (n + B - 1) / B;
    const uint32_t j0 = tid * B, j1 = j0 + B < n ? j0 + B : n;
    uint64_t c = 0;
    uint32_t r0 = 0;
    bool pr = true;
    for (uint32_t jj = j0; jj < j1; ++jj) {
        const uint32_t j = jlo + jj;
        uint64_t v = (uint64_t)b.s0[j] + (jj >= 1 ? b.s1[j - 1] : 0u) + (jj >= 2 ? b.s2[j - 2] : 0u) + c;
        if (jj == 0) v += (uint32_t)extra;
        if (jj == 1) v += (uint32_t)(extra >> 32);
        const uint32_t r = (uint32_t)v;
        c = v >> 32;
        out[jj] = r;
        if (jj == j0) r0 = r;
        else pr = pr && r == 0xFFFFFFFFu;
    }
    CF f{(uint32_t)c, INF};
    if (j0 < j1 && pr && r0 >= 0xFFFFFFF0u) f.th = 0u - r0;   // 2^32 - r0 <= 16
    uint32_t top;
    uint32_t ci = scan_carry(f, nb, b.sc, &top);
    for (uint32_t jj = j0; jj < j1 && ci; ++jj) {
        const uint64_t v = (uint64_t)out[jj] + ci;
        out[jj] = (uint32_t)v;
        ci = (uint32_t)(v >> 32);
    }
    __syncthreads();
    return top;
}

__device__ inline uint32_t carry(const Big& b, uint32_t n, uint32_t* out) { return carry_range(b, 0, n, out, 0); }

// dst = src - q when top or src >= q, else src (src: L words plus the top bit; src may be dst).
__device__ void sub_if_ge(const Big& b, const uint32_t* src, uint32_t top, uint32_t* dst) {
    const uint32_t L = b.L, tid = threadIdx.x;
    const uint32_t B = (L + bw() - 1) / bw();
    const uint32_t nb = (L + B - 1) / B;
    const uint32_t j0 = tid * B, j1 = j0 + B < L ? j0 + B : L;
    uint32_t bo = 0;
    bool allz = true;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t v = (uint64_t)src[j] - b.q[j] - bo;
        b.e[j] = (uint32_t)v;
        bo = (uint32_t)(v >> 63);
        allz = allz && (uint32_t)v == 0;
    }
    CF f{bo, (j0 < j1 && allz) ? 1u : INF};
    uint32_t fin;
    uint32_t bi = scan_carry(f, nb, b.sc, &fin);
    const bool take = top || fin == 0;
    for (uint32_t j = j0; j < j1; ++j) {
        uint32_t v = b.e[j];
        if (bi) {
            bi = v == 0 ? 1u : 0u;
            v -= 1;
        }
        dst[j] = take ? v : src[j];
    }
    __syncthreads();
}

// x = 2 x mod q (x < q)
__device__ void dbl_mod(const Big& b, uint32_t* x) {
    const uint32_t L = b.L;
    for (uint32_t j = threadIdx.x; j < L; j += bw()) b.d[j] = (x[j] << 1) | (j ? x[j - 1] >> 31 : 0u);
    const uint32_t top = x[L - 1] >> 31;
    __syncthreads();
    sub_if_ge(b, b.d, top, x);
}

// x = x + a mod q (x, a < q)
__device__ void add_mod(const Big& b, uint32_t* x, const uint32_t* a) {
    const uint32_t L = b.L, tid = threadIdx.x;
    const uint32_t B = (L + bw() - 1) / bw();
    const uint32_t nb = (L + B - 1) / B;
    const uint32_t j0 = tid * B, j1 = j0 + B < L ? j0 + B : L;
    uint32_t c = 0;
    bool allf = true;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t v = (uint64_t)x[j] + a[j] + c;
        b.d[j] = (uint32_t)v;
        c = (uint32_t)(v >> 32);
        allf = allf && (uint32_t)v == 0xFFFFFFFFu;
    }
    CF f{c, (j0 < j1 && allf) ? 1u : INF};
    uint32_t top;
    uint32_t ci = scan_carry(f, nb, b.sc, &top);
    for (uint32_t j = j0; j < j1 && ci; ++j) {
        const uint64_t v = (uint64_t)b.d[j] + ci;
        b.d[j] = (uint32_t)v;
        ci = (uint32_t)(v >> 32);
    }
    __syncthreads();
    sub_if_ge(b, b.d, top, x);
}

// x = k x mod q for 2 < k < 128 (x < q; x in Montgomery form stays so): t = k x in L + 1
// words, qhat = floor(t / q) from the top words (off by at most one), y = t - qhat q, and
// one correction (add q when y < 0, subtract q when y >= q).  Three scans, O(L) work: the
// multiply step of a strong test to a small base, in place of a Montgomery product.
__device__ void mul_small_mod(const Big& b, uint32_t* x, uint32_t k) {
    const uint32_t L = b.L, tid = threadIdx.x;
    const uint32_t B = (L + bw() - 1) / bw();
    const uint32_t nb = (L + B - 1) / B;
    const uint32_t j0 = tid * B, j1 = j0 + B < L ? j0 + B : L;
    // t = k x into d, the word above into t_top (carries below k < 2^7)
    uint32_t c = 0, r0 = 0;
    bool pr = true;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t v = (uint64_t)x[j] * k + c;
        b.d[j] = (uint32_t)v;
        c = (uint32_t)(v >> 32);
        if (j == j0) r0 = (uint32_t)v;
        else pr = pr && (uint32_t)v == 0xFFFFFFFFu;
    }
    CF f{c, INF};
    if (j0 < j1 && pr && r0 >= 0xFFFFFF00u) f.th = 0u - r0;
    uint32_t t_top;
    uint32_t ci = scan_carry(f, nb, b.sc, &t_top);
    for (uint32_t j = j0; j < j1 && ci; ++j) {
        const uint64_t v = (uint64_t)b.d[j] + ci;
        b.d[j] = (uint32_t)v;
        ci = (uint32_t)(v >> 32);
    }
    __syncthreads();
    if (tid == 0) {   // q >= 2^(32(L-1)) - 2^24, so its top two words carry the ratio to 2^-31
        const double tt = (double)t_top * 18446744073709551616.0 + (double)b.d[L - 1] * 4294967296.0 + b.d[L - 2];
        const double qq = (double)b.q[L - 1] * 4294967296.0 + b.q[L - 2] + b.q[L - 3] * (1.0 / 4294967296.0);
        double r = floor(tt / qq);
        r = r < 0.0 ? 0.0 : (r > (double)k ? (double)k : r);
        b.sc[2 * NW_MAX + 5] = (uint32_t)r;
    }
    __syncthreads();
    const uint32_t qh = b.sc[2 * NW_MAX + 5];
    // y = t - qh q in place in d; a borrow-in bi below 2^8 adds a borrow out iff the block is below bi
    uint32_t bo = 0;
    bool allz = true;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t m = (uint64_t)qh * b.q[j] + bo;
        const uint32_t lo = (uint32_t)m, dj = b.d[j], r = dj - lo;
        b.d[j] = r;
        bo = (uint32_t)(m >> 32) + (dj < lo ? 1u : 0u);
        if (j == j0) r0 = r;
        else allz = allz && r == 0;
    }
    CF g{bo, INF};
    if (j0 < j1 && allz && r0 < 256u) g.th = r0 + 1;
    uint32_t bfin;
    uint32_t bi = scan_carry(g, nb, b.sc, &bfin);
    for (uint32_t j = j0; j < j1 && bi; ++j) {
        const uint32_t v = b.d[j];
        b.d[j] = v - bi;
        bi = v < bi ? 1u : 0u;
    }
    __syncthreads();
    if (t_top < bfin) {                              // qh one too large: y + q, its carry out dropped
        uint32_t cc = 0;
        bool allf = true;
        for (uint32_t j = j0; j < j1; ++j) {
            const uint64_t v = (uint64_t)b.d[j] + b.q[j] + cc;
            b.d[j] = (uint32_t)v;
            cc = (uint32_t)(v >> 32);
            allf = allf && (uint32_t)v == 0xFFFFFFFFu;
        }
        uint32_t unused;
        uint32_t cj = scan_carry(CF{cc, (j0 < j1 && allf) ? 1u : INF}, nb, b.sc, &unused);
        for (uint32_t j = j0; j < j1; ++j) {
            const uint64_t v = (uint64_t)b.d[j] + cj;
            x[j] = (uint32_t)v;
            cj = (uint32_t)(v >> 32);
        }
        __syncthreads();
    } else {
        sub_if_ge(b, b.d, t_top - bfin, x);          // qh exact or one too small
    }
}

// Montgomery reduction of the product T = b.t (2L words, T < q R): x = T / R mod q.
// U = (T + M q) / R with M = (T mod R)(-1/q) mod R.  Of M q only the columns from L - 3 up
// are summed: the low L words of T + M q are zero, so the carry C out of them is an integer
// that the top three low columns fix: with W = S_(L-1) 2^64 + S_(L-2) 2^32 + S_(L-3) (T
// folded in), C = ceil(W / 2^96), the columns below L - 3 adding less than 2^-19 R.
__device__ void mont_reduce(const Big& b, uint32_t* x) {
    const uint32_t L = b.L;
    cols(b, b.t, L, b.qi, L, 0, L, nullptr);
    carry(b, L, b.m);                                        // M
    cols(b, b.m, L, b.q, L, L - 3, 2 * L, b.t);              // columns L-3 .. 2L-1 of T + M q
    if (threadIdx.x == 0) {
        const uint32_t k = L - 3;
        // W in 32-bit positions 0..4, then C = ceil(W / 2^96)
        uint64_t w0 = b.s0[k];
        uint64_t w1 = (uint64_t)b.s1[k] + b.s0[k + 1];
        uint64_t w2 = (uint64_t)b.s2[k] + b.s1[k + 1] + b.s0[k + 2];
        uint64_t w3 = (uint64_t)b.s2[k + 1] + b.s1[k + 2];
        uint64_t w4 = b.s2[k + 2];
        w2 += w1 >> 32;
        w1 &= 0xFFFFFFFFu;
        w3 += w2 >> 32;
        w2 &= 0xFFFFFFFFu;
        w4 += w3 >> 32;
        w3 &= 0xFFFFFFFFu;
        const uint64_t cc = w3 + (w4 << 32) + ((w0 | w1 | w2) ? 1u : 0u);
        b.sc[2 * NW_MAX + 3] = (uint32_t)cc;
        b.sc[2 * NW_MAX + 4] = (uint32_t)(cc >> 32);
    }
    __syncthreads();
    const uint64_t cc = (uint64_t)b.sc[2 * NW_MAX + 3] | ((uint64_t)b.sc[2 * NW_MAX + 4] << 32);
    const uint32_t top = carry_range(b, L, 2 * L, b.t + L, cc);   // U = T_hi + (M q)_hi + C
    sub_if_ge(b, b.t + L, top, x);
}

// x = x y / R mod q (x, y < q)
__device__ void mont_mul(const Big& b, uint32_t* x, const uint32_t* y) {
    cols(b, x, b.L, y, b.L, 0, 2 * b.L, nullptr);
    carry(b, 2 * b.L, b.t);
    mont_reduce(b, x);
}

// x = x^2 / R mod q (x < q)
__device__ void mont_sqr(const Big& b, uint32_t* x) {
    cols_sqr(b, x);
    carry(b, 2 * b.L, b.t);
    mont_reduce(b, x);
}

__device__ bool equal(const Big& b, const uint32_t* u, const uint32_t* v) {
    int diff = 0;
    for (uint32_t j = threadIdx.x; j < b.L; j += bw()) diff |= u[j] != v[j];
    return __syncthreads_or(diff) == 0;
}

__device__ inline uint32_t ebit(const Big& b, uint32_t j) {   // bit j of q - 1 (q odd)
    const uint32_t w = b.q[j >> 5] - (j < 32 ? 1u : 0u);
    return (w >> (j & 31)) & 1u;
}

}  // namespace

__global__ __launch_bounds__(BIG_BS) void big_mr_kernel(const uint32_t* __restrict__ n, uint32_t L,
                                                        const BigTask* __restrict__ tasks,
                                                        uint32_t* __restrict__ pass) {
    extern __shared__ uint4 lds4[];                  // 16-byte aligned base
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds4);
    const Layout o = layout(L);
    Big b;
    b.L = L;
    b.x = lds + o.x; b.y = lds + o.y; b.q = lds + o.q; b.qi = lds + o.qi;
    b.t = lds + o.t; b.m = lds + o.m; b.s0 = lds + o.s0; b.s1 = lds + o.s1; b.s2 = lds + o.s2;
    b.d = lds + o.d; b.e = lds + o.e; b.rq = lds + o.rq; b.qmr = lds + o.qmr; b.sc = lds + o.scan;
    const BigTask task = tasks[blockIdx.x];
    const uint32_t tid = threadIdx.x;

    // guarded buffers start as zero; q = n
    for (uint32_t j = tid; j < o.t; j += bw()) lds[j] = 0;
    __syncthreads();
    for (uint32_t j = tid; j < L; j += bw()) b.q[j] = n[j];
    __syncthreads();
    if (tid == 0) {                                  // q = n - p (n > p: the borrow stops)
        uint64_t v = (uint64_t)b.q[0] - task.p;
        b.q[0] = (uint32_t)v;
        for (uint32_t j = 1; (v >> 63) && j < L; ++j) {
            v = (uint64_t)b.q[j] - 1;
            b.q[j] = (uint32_t)v;
        }
        // -1/q mod 2^32 by Newton from the 3-bit inverse q0 (q0 q0 = 1 mod 8)
        const uint32_t q0 = b.q[0];
        uint32_t inv = q0;
        for (int k = 0; k < 4; ++k) inv *= 2u - q0 * inv;
        b.qi[0] = 0u - inv;
        // bit length of q
        uint32_t top = L - 1;
        while (top && b.q[top] == 0) --top;
        b.sc[2 * NW_MAX + 1] = 32 * top + (32 - __builtin_clz(b.q[top]));
        uint32_t s = 1;                              // q - 1 = 2^s d
        while (!ebit(b, s)) ++s;
        b.sc[2 * NW_MAX + 2] = s;
    }
    __syncthreads();
    const uint32_t tbits = b.sc[2 * NW_MAX + 1], s = b.sc[2 * NW_MAX + 2];

    // -1/q mod R: y <- y (2 + q y) doubles the words that are right
    for (uint32_t m = 1; m < L;) {
        const uint32_t m2 = 2 * m < L ? 2 * m : L;
        cols(b, b.q, m2, b.qi, m, 0, m2, nullptr);
        carry(b, m2, b.d);                           // d = q y mod 2^(32 m2)
        if (tid == 0) {                              // d += 2
            uint64_t v = (uint64_t)b.d[0] + 2;
            b.d[0] = (uint32_t)v;
            for (uint32_t j = 1; (v >> 32) && j < m2; ++j) {
                v = (uint64_t)b.d[j] + 1;
                b.d[j] = (uint32_t)v;
            }
        }
        __syncthreads();
        cols(b, b.d, m2, b.qi, m, 0, m2, nullptr);
        carry(b, m2, b.qi);                          // y = y d mod 2^(32 m2)
        m = m2;
    }

    // R mod q: 2^(t-1) < q, then 32 L - t + 1 doublings
    for (uint32_t j = tid; j < L; j += bw()) b.rq[j] = 0;
    __syncthreads();
    if (tid == 0) b.rq[(tbits - 1) >> 5] = 1u << ((tbits - 1) & 31);
    __syncthreads();
    for (uint32_t k = tbits - 1; k < 32 * L; ++k) dbl_mod(b, b.rq);
    // q - (R mod q): the Montgomery form of -1
    for (uint32_t j = tid; j < L; j += bw()) b.d[j] = b.rq[j];
    __syncthreads();
    {
        const uint32_t B = (L + bw() - 1) / bw(), nb = (L + B - 1) / B;
        const uint32_t j0 = tid * B, j1 = j0 + B < L ? j0 + B : L;
        uint32_t bo = 0;
        bool allz = true;
        for (uint32_t j = j0; j < j1; ++j) {
            const uint64_t v = (uint64_t)b.q[j] - b.d[j] - bo;
            b.qmr[j] = (uint32_t)v;
            bo = (uint32_t)(v >> 63);
            allz = allz && (uint32_t)v == 0;
        }
        uint32_t fin;
        uint32_t bi = scan_carry(CF{bo, (j0 < j1 && allz) ? 1u : INF}, nb, b.sc, &fin);
        for (uint32_t j = j0; j < j1 && bi; ++j) {
            bi = b.qmr[j] == 0 ? 1u : 0u;
            b.qmr[j] -= 1;
        }
        __syncthreads();
    }

    // y = base R mod q
    const uint32_t base = task.base;
    for (uint32_t j = tid; j < L; j += bw()) b.y[j] = b.rq[j];
    __syncthreads();
    for (int k = 30 - __builtin_clz(base); k >= 0; --k) {
        dbl_mod(b, b.y);
        if ((base >> k) & 1) add_mod(b, b.y, b.rq);
    }

    // x = base^d, d = (q - 1) / 2^s, from the top bit of q - 1 down to bit s
    for (uint32_t j = tid; j < L; j += bw()) b.x[j] = b.y[j];
    __syncthreads();
    for (int j = (int)tbits - 2; j >= (int)s; --j) {
        mont_sqr(b, b.x);
        if (ebit(b, (uint32_t)j)) {
            if (base == 2) dbl_mod(b, b.x);
            else mul_small_mod(b, b.x, base);
        }
    }
    bool ok = equal(b, b.x, b.rq) || equal(b, b.x, b.qmr);
    for (uint32_t r = 1; r < s && !ok; ++r) {
        mont_sqr(b, b.x);
        if (equal(b, b.x, b.qmr)) ok = true;
        else if (equal(b, b.x, b.rq)) break;
    }
    if (tid == 0) pass[blockIdx.x] = ok ? 1u : 0u;
}

// r[i] = n mod ell[i]
__global__ void big_residue_kernel(const uint32_t* __restrict__ n, uint32_t L, const uint32_t* __restrict__ ell,
                                   uint32_t nell, uint32_t* __restrict__ r) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nell) return;
    const uint64_t l = ell[i];
    uint64_t v = 0;
    for (int j = (int)L - 1; j >= 0; --j) v = ((v << 32) | n[j]) % l;
    r[i] = (uint32_t)v;
}

// bad bit (p - 1) / 2 for every odd p <= p_max with p = n mod ell, i.e. ell | n - p.
// Block (x, y): prime ell[i0 + x], the p of [y span, (y + 1) span).
__global__ void big_filter_kernel(const uint32_t* __restrict__ ell, const uint32_t* __restrict__ r, uint32_t i0,
                                  uint64_t p_max, uint64_t span, uint32_t* __restrict__ bad) {
    const uint32_t i = i0 + blockIdx.x;
    const uint64_t l = ell[i], l2 = 2 * l;
    const uint64_t ro = (r[i] & 1) ? r[i] : r[i] + l;   // the odd p = n mod ell: p = ro mod 2 ell
    const uint64_t lo = blockIdx.y * span;
    if (lo > p_max) return;
    const uint64_t hi = lo + span - 1 < p_max ? lo + span - 1 : p_max;
    uint64_t p = lo + (ro + l2 - lo % l2) % l2;
    for (p += l2 * threadIdx.x; p <= hi; p += l2 * blockDim.x) atomicOr(&bad[p >> 6], 1u << ((p >> 1) & 31));
}

size_t big_lds_bytes(uint32_t L) { return (size_t)layout(L).total * 4; }

uint32_t big_threads(uint32_t L) { return L >= 384 ? BIG_BS : 256; }

hipError_t launch_big_mr(hipStream_t s, const uint32_t* n, uint32_t L, const BigTask* tasks, uint32_t ntasks,
                         uint32_t* pass) {
    if (L < 2 || L > BIG_LMAX || ntasks == 0) return hipErrorInvalidValue;
    const size_t lds = big_lds_bytes(L);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(big_mr_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(big_mr_kernel, dim3(ntasks), dim3(big_threads(L)), lds, s, n, L, tasks, pass);
    return hipGetLastError();
}

hipError_t launch_big_filter(hipStream_t s, const uint32_t* n, uint32_t L, const uint32_t* ell, uint32_t nell,
                             uint32_t nsmall, uint32_t* r, uint64_t p_max, uint32_t* bad) {
    if (nell == 0) return hipSuccess;
    hipLaunchKernelGGL(big_residue_kernel, dim3((nell + 255) / 256), dim3(256), 0, s, n, L, ell, nell, r);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the first nsmall primes hit many p: their ranges are cut in 64 pieces
    const uint64_t span = (p_max + 64) / 64;
    if (nsmall) {
        hipLaunchKernelGGL(big_filter_kernel, dim3(nsmall, 64), dim3(64), 0, s, ell, r, 0u, p_max, span, bad);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nell > nsmall)
        hipLaunchKernelGGL(big_filter_kernel, dim3(nell - nsmall, 1), dim3(64), 0, s, ell, r, nsmall, p_max,
                           p_max + 1, bad);
    return hipGetLastError();
}

}  // namespace gbk
