// gb_big.hip -- strong probable-prime tests of q = n - p for even n of any size
// (up to BIG_LMAX 32-bit words), one workgroup per (p, base) task.
//
// The reference's big_check (big_check.cpp:83-118) asks GMP whether q = n - p
// is prime for p = 2, 3, 5, ... and stops at the first yes.  Here a
// workgroup of 256 threads holds q and its Montgomery constants in LDS and
// runs one strong test (Miller-Rabin with a fixed base) on it:
//   * products as column sums: thread (wave w, lane l) owns columns
//     64 g + l for the groups g = w, w + 4, ...; a column adds a_i * b_(k-i)
//     over i in 96 bits (one v_mad_u64_u32 and a 64-bit add per term);
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
//   * base 2 multiplies by doubling; other bases by a Montgomery product with
//     b R mod q (built from R mod q by doublings and additions).
#include <hip/hip_runtime.h>

#include "gb_big.hpp"

namespace gbk {
namespace {

constexpr uint32_t BW = BIG_BS;           // threads per workgroup
constexpr uint32_t NW = BW / 64;          // waves per workgroup
constexpr uint32_t GUARD = 64;            // zero words either side of a multiplicand
constexpr uint32_t INF = 0xFFFFFFFFu;     // carry-function threshold never reached

// LDS word offsets for a q of L words.
struct Layout {
    uint32_t x, y, q, qi, t, m, s0, s1, s2, d, e, rq, qmr, scan, total;
};

__host__ __device__ inline Layout layout(uint32_t L) {
    Layout o{};
    uint32_t w = 0;
    const uint32_t gl = L + 2 * GUARD;
    o.x = w + GUARD; w += gl;
    o.y = w + GUARD; w += gl;
    o.q = w + GUARD; w += gl;
    o.qi = w + GUARD; w += gl;
    o.t = w; w += 2 * L;
    o.m = w; w += L;
    o.s0 = w; w += 2 * L;
    o.s1 = w; w += 2 * L;
    o.s2 = w; w += 2 * L;
    o.d = w; w += L + 1;
    o.e = w; w += L + 1;
    o.rq = w; w += L;
    o.qmr = w; w += L;
    o.scan = w; w += 4 * NW + 8;
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

// Workgroup scan of the threads' carry functions over blocks 0..nb-1 (thread t = block t).
// Returns the carry into this thread's block; *top gets the carry out of block nb-1.
__device__ uint32_t scan_carry(CF f, uint32_t nb, uint32_t* sc, uint32_t* top) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    CF inc = f;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        CF g;
        g.a = __shfl_up(inc.a, off, 64);
        g.th = __shfl_up(inc.th, off, 64);
        if (lane >= off) inc = compose(g, inc);
    }
    CF exc;
    exc.a = __shfl_up(inc.a, 1, 64);
    exc.th = __shfl_up(inc.th, 1, 64);
    if (lane == 63) {
        sc[2 * wv] = inc.a;
        sc[2 * wv + 1] = inc.th;
    }
    __syncthreads();
    uint32_t c = 0;                                  // carry into this wave
    for (uint32_t v = 0; v < wv; ++v) c = apply(CF{sc[2 * v], sc[2 * v + 1]}, c);
    const uint32_t cin = lane == 0 ? c : apply(exc, c);
    if (tid == nb - 1) sc[2 * NW] = apply(f, cin);
    __syncthreads();
    *top = sc[2 * NW];
    return cin;
}

struct Big {
    uint32_t L;
    uint32_t *x, *y, *q, *qi, *t, *m, *s0, *s1, *s2, *d, *e, *rq, *qmr, *sc;
};

// Column sums of A (la words) times B (lb words, guarded) for columns k < kend,
// each started from add[k] when add is given; written as three words per column.
__device__ void cols(const Big& b, const uint32_t* A, uint32_t la, const uint32_t* B, uint32_t lb, uint32_t kend,
                     const uint32_t* add) {
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t g = wv; 64 * g < kend; g += NW) {
        const int kmin = 64 * (int)g, kmax = kmin + 63;
        const int k = kmin + (int)lane;
        const int ilo = kmin - (int)lb + 1 > 0 ? kmin - (int)lb + 1 : 0;
        const int ihi = kmax < (int)la - 1 ? kmax : (int)la - 1;
        uint32_t lo = (add && k < (int)kend) ? add[k] : 0u;
        uint64_t hi = 0;
        const uint32_t* bk = B + k;
        int i = ilo;
        for (; i + 3 <= ihi; i += 4) {
            uint64_t p0 = (uint64_t)A[i] * bk[-i] + lo;
            lo = (uint32_t)p0;
            hi += p0 >> 32;
            uint64_t p1 = (uint64_t)A[i + 1] * bk[-i - 1] + lo;
            lo = (uint32_t)p1;
            hi += p1 >> 32;
            uint64_t p2 = (uint64_t)A[i + 2] * bk[-i - 2] + lo;
            lo = (uint32_t)p2;
            hi += p2 >> 32;
            uint64_t p3 = (uint64_t)A[i + 3] * bk[-i - 3] + lo;
            lo = (uint32_t)p3;
            hi += p3 >> 32;
        }
        for (; i <= ihi; ++i) {
            const uint64_t p = (uint64_t)A[i] * bk[-i] + lo;
            lo = (uint32_t)p;
            hi += p >> 32;
        }
        if (k < (int)kend) {
            b.s0[k] = lo;
            b.s1[k] = (uint32_t)hi;
            b.s2[k] = (uint32_t)(hi >> 32);
        }
    }
    __syncthreads();
}

// Resolves the column sums of words [0, n) into out; returns the carry out of word n-1
// (contributions of the column sums beyond word n-1 are dropped).
__device__ uint32_t carry(const Big& b, uint32_t n, uint32_t* out) {
    const uint32_t tid = threadIdx.x;
    const uint32_t B = (n + BW - 1) / BW;
    const uint32_t nb = (n + B - 1) / B;
    const uint32_t j0 = tid * B, j1 = j0 + B < n ? j0 + B : n;
    uint64_t c = 0;
    uint32_t r0 = 0;
    bool pr = true;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t v = (uint64_t)b.s0[j] + (j >= 1 ? b.s1[j - 1] : 0u) + (j >= 2 ? b.s2[j - 2] : 0u) + c;
        const uint32_t r = (uint32_t)v;
        c = v >> 32;
        out[j] = r;
        if (j == j0) r0 = r;
        else pr = pr && r == 0xFFFFFFFFu;
    }
    CF f{(uint32_t)c, INF};
    if (j0 < j1 && pr && r0 >= 0xFFFFFFF0u) f.th = 0u - r0;   // 2^32 - r0 <= 16
    uint32_t top;
    uint32_t ci = scan_carry(f, nb, b.sc, &top);
    for (uint32_t j = j0; j < j1 && ci; ++j) {
        const uint64_t v = (uint64_t)out[j] + ci;
        out[j] = (uint32_t)v;
        ci = (uint32_t)(v >> 32);
    }
    __syncthreads();
    return top;
}

// dst = src - q when top or src >= q, else src (src: L words plus the top bit; src may be dst).
__device__ void sub_if_ge(const Big& b, const uint32_t* src, uint32_t top, uint32_t* dst) {
    const uint32_t L = b.L, tid = threadIdx.x;
    const uint32_t B = (L + BW - 1) / BW;
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
    for (uint32_t j = threadIdx.x; j < L; j += BW) b.d[j] = (x[j] << 1) | (j ? x[j - 1] >> 31 : 0u);
    const uint32_t top = x[L - 1] >> 31;
    __syncthreads();
    sub_if_ge(b, b.d, top, x);
}

// x = x + a mod q (x, a < q)
__device__ void add_mod(const Big& b, uint32_t* x, const uint32_t* a) {
    const uint32_t L = b.L, tid = threadIdx.x;
    const uint32_t B = (L + BW - 1) / BW;
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

// x = x * y / R mod q (x, y < q; y may be x)
__device__ void mont_mul(const Big& b, uint32_t* x, const uint32_t* y) {
    const uint32_t L = b.L;
    cols(b, x, L, y, L, 2 * L, nullptr);
    carry(b, 2 * L, b.t);                            // T = x y
    cols(b, b.t, L, b.qi, L, L, nullptr);
    carry(b, L, b.m);                                // M = (T mod R)(-1/q) mod R
    cols(b, b.m, L, b.q, L, 2 * L, b.t);
    const uint32_t top = carry(b, 2 * L, b.t);       // T + M q, low half zero
    sub_if_ge(b, b.t + L, top, x);
}

__device__ bool equal(const Big& b, const uint32_t* u, const uint32_t* v) {
    int diff = 0;
    for (uint32_t j = threadIdx.x; j < b.L; j += BW) diff |= u[j] != v[j];
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
    extern __shared__ uint32_t lds[];
    const Layout o = layout(L);
    Big b;
    b.L = L;
    b.x = lds + o.x; b.y = lds + o.y; b.q = lds + o.q; b.qi = lds + o.qi;
    b.t = lds + o.t; b.m = lds + o.m; b.s0 = lds + o.s0; b.s1 = lds + o.s1; b.s2 = lds + o.s2;
    b.d = lds + o.d; b.e = lds + o.e; b.rq = lds + o.rq; b.qmr = lds + o.qmr; b.sc = lds + o.scan;
    const BigTask task = tasks[blockIdx.x];
    const uint32_t tid = threadIdx.x;

    // guarded buffers start as zero; q = n
    for (uint32_t j = tid; j < o.t; j += BW) lds[j] = 0;
    __syncthreads();
    for (uint32_t j = tid; j < L; j += BW) b.q[j] = n[j];
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
        b.sc[2 * NW + 1] = 32 * top + (32 - __builtin_clz(b.q[top]));
        uint32_t s = 1;                              // q - 1 = 2^s d
        while (!ebit(b, s)) ++s;
        b.sc[2 * NW + 2] = s;
    }
    __syncthreads();
    const uint32_t tbits = b.sc[2 * NW + 1], s = b.sc[2 * NW + 2];

    // -1/q mod R: y <- y (2 + q y) doubles the words that are right
    for (uint32_t m = 1; m < L;) {
        const uint32_t m2 = 2 * m < L ? 2 * m : L;
        cols(b, b.q, m2, b.qi, m, m2, nullptr);
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
        cols(b, b.d, m2, b.qi, m, m2, nullptr);
        carry(b, m2, b.qi);                          // y = y d mod 2^(32 m2)
        m = m2;
    }

    // R mod q: 2^(t-1) < q, then 32 L - t + 1 doublings
    for (uint32_t j = tid; j < L; j += BW) b.rq[j] = 0;
    __syncthreads();
    if (tid == 0) b.rq[(tbits - 1) >> 5] = 1u << ((tbits - 1) & 31);
    __syncthreads();
    for (uint32_t k = tbits - 1; k < 32 * L; ++k) dbl_mod(b, b.rq);
    // q - (R mod q): the Montgomery form of -1
    for (uint32_t j = tid; j < L; j += BW) b.d[j] = b.rq[j];
    __syncthreads();
    {
        const uint32_t B = (L + BW - 1) / BW, nb = (L + B - 1) / B;
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
    for (uint32_t j = tid; j < L; j += BW) b.y[j] = b.rq[j];
    __syncthreads();
    for (int k = 30 - __builtin_clz(base); k >= 0; --k) {
        dbl_mod(b, b.y);
        if ((base >> k) & 1) add_mod(b, b.y, b.rq);
    }

    // x = base^d, d = (q - 1) / 2^s, from the top bit of q - 1 down to bit s
    for (uint32_t j = tid; j < L; j += BW) b.x[j] = b.y[j];
    __syncthreads();
    for (int j = (int)tbits - 2; j >= (int)s; --j) {
        mont_mul(b, b.x, b.x);
        if (ebit(b, (uint32_t)j)) {
            if (base == 2) dbl_mod(b, b.x);
            else mont_mul(b, b.x, b.y);
        }
    }
    bool ok = equal(b, b.x, b.rq) || equal(b, b.x, b.qmr);
    for (uint32_t r = 1; r < s && !ok; ++r) {
        mont_mul(b, b.x, b.x);
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

// bad bit (p - 1) / 2 for every odd p <= p_max with p = n mod ell, i.e. ell | n - p
__global__ void big_filter_kernel(const uint32_t* __restrict__ ell, const uint32_t* __restrict__ r, uint32_t nell,
                                  uint64_t p_max, uint32_t* __restrict__ bad) {
    const uint32_t i = blockIdx.x;
    if (i >= nell) return;
    const uint64_t l = ell[i];
    uint64_t p = r[i];
    if (!(p & 1)) p += l;                            // odd p only (l odd)
    for (p += 2 * l * threadIdx.x; p <= p_max; p += 2 * l * blockDim.x)
        atomicOr(&bad[p >> 6], 1u << ((p >> 1) & 31));
}

size_t big_lds_bytes(uint32_t L) { return (size_t)layout(L).total * 4; }

hipError_t launch_big_mr(hipStream_t s, const uint32_t* n, uint32_t L, const BigTask* tasks, uint32_t ntasks,
                         uint32_t* pass) {
    if (L < 2 || L > BIG_LMAX || ntasks == 0) return hipErrorInvalidValue;
    const size_t lds = big_lds_bytes(L);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(big_mr_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(big_mr_kernel, dim3(ntasks), dim3(BIG_BS), lds, s, n, L, tasks, pass);
    return hipGetLastError();
}

hipError_t launch_big_filter(hipStream_t s, const uint32_t* n, uint32_t L, const uint32_t* ell, uint32_t nell,
                             uint32_t* r, uint64_t p_max, uint32_t* bad) {
    if (nell == 0) return hipSuccess;
    hipLaunchKernelGGL(big_residue_kernel, dim3((nell + 255) / 256), dim3(256), 0, s, n, L, ell, nell, r);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(big_filter_kernel, dim3(nell), dim3(64), 0, s, ell, r, nell, p_max, bad);
    return hipGetLastError();
}

}  // namespace gbk
