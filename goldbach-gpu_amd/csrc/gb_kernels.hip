// gb_kernels.hip -- hand-written gfx950 kernels for the segmented double-sieve
// Goldbach verifier.
//
//   sieve_tile      odd-only prime bitset of one window tile in LDS:
//                   word patterns for p <= 61, LDS atomics for 67 <= p,
//                   medium primes split into 32-mark work items so a wave's
//                   lanes stride p 32-bit words apart (distinct LDS banks),
//                   Baillie-PSW on pre-sieve survivors above 2^44.
//   mark_words      Phase 1 + in-window Phase 2: a lane owns 64 consecutive
//                   even n; for each odd prime p (wave-uniform) the 64 q = n-p
//                   are 64 consecutive window bits, so one funnel shift + OR
//                   marks 64 n.  Newly set bits give p_min exactly.
//   fused_kernel    persistent workgroups pull tiles from a device counter and
//                   run sieve_tile -> mark_words entirely in LDS (no HBM round
//                   trip of the window).
//   slow_kernel     the rare n whose p_min lies above the window's reach:
//                   primes from the table, q by bitset or Miller-Rabin.
//
// Reference semantics restated: goldbach.cu:160-179 (q resolution), :198-259
// (segment sieve), :288-311 (Phase 1), :362-373 (Phase 2), :184-195 (count).
#include "gb_kernels.hpp"

#include <utility>

namespace gbk {

// ---------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------
__host__ __device__ constexpr uint32_t pat_p(int i) {
    constexpr uint32_t P[NPAT] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    return P[i];
}

// bits 0, p, 2p, ... below 64
__host__ __device__ constexpr uint64_t pat_mask(uint32_t p) {
    uint64_t m = 0;
    for (uint32_t b = 0; b < 64; b += p) m |= 1ull << b;
    return m;
}

__device__ __forceinline__ uint32_t count_le(const uint32_t* sp, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)sp[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Clearing bits of the LDS sieve window.  One LDS instruction per mark: a CU
// retires about one wave64 LDS instruction per 4.5 clocks (ds_and_b32 and
// plain stores alike; ds_mskor_b32 is slower).  Along a run i, i+p, i+2p, ...
// the clear-mask ~(1 << (i & 31)) rotates left by p mod 32 per step, so a
// mark costs one v_alignbit_b32 for the mask, two ops for the word address
// and the index add (the window sits at a fixed LDS address).

// ~(1u << (i & 31)) as a rotation of ~1.
__device__ __forceinline__ uint32_t clear_mask(uint32_t i) { return __builtin_amdgcn_alignbit(~1u, ~1u, 0u - i); }

__device__ __forceinline__ void lds_clear_bit(uint32_t* win32, uint32_t i) { atomicAnd(&win32[i >> 5], clear_mask(i)); }

// bits i, i + p, ... below end
// rot = 0 - p serves as the mask's rotation (left by p mod 32) and, negated, as the step; it
// is formed in asm so that the compiler neither folds i - rot back into i + p nor recomputes
// rot inside the loop when registers run short (which would cost a VALU op per mark).
__device__ __forceinline__ void clear_run(uint32_t* win32, uint32_t i, uint32_t p, uint32_t end) {
    uint32_t mask = clear_mask(i);
    uint32_t rot;
    asm("v_sub_u32 %0, 0, %1" : "=v"(rot) : "v"(p));
    for (; i < end; i -= rot) {
        atomicAnd(&win32[i >> 5], mask);
        mask = __builtin_amdgcn_alignbit(mask, mask, rot);
    }
}

// exactly N bits i, i + p, ..., i + (N-1)p
template <int N>
__device__ __forceinline__ void clear_run_n(uint32_t* win32, uint32_t i, uint32_t p) {
    uint32_t mask = clear_mask(i);
    const uint32_t rot = 0u - p;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        atomicAnd(&win32[i >> 5], mask);
        mask = __builtin_amdgcn_alignbit(mask, mask, rot);
        i += p;
    }
}

// ---------------------------------------------------------------------------
// Block-wide exclusive prefix sum of one count per thread (wave scans through
// lane shuffles, wave totals through wsum[BS/64] in LDS).  *total = block sum.
// ---------------------------------------------------------------------------
template <int BS>
__device__ uint32_t block_scan(uint32_t c, uint32_t tid, uint32_t* __restrict__ wsum, uint32_t* total) {
    const uint32_t lane = tid & 63, wv = tid >> 6;
    uint32_t x = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) {
        const uint32_t v = wsum[w];
        before += (uint32_t)w < wv ? v : 0;
        all += v;
    }
    __syncthreads();   // wsum may be rewritten by the next scan
    *total = all;
    return before + x - c;
}

// Primality of the pre-sieve survivors of a tile (bit i <-> odd q = qb + 2i, every
// q > 2^44) by Baillie-PSW.  Survivors are packed into an LDS queue so all lanes
// run tests of equal length instead of each lane walking its own word's survivors:
//   pass 0: every survivor gets the base-2 strong test,
//   pass 1: every survivor left gets the strong Lucas test;
// a failed test clears the bit.  No composite below 2^64 passes both, so the answer is
// the one of the Miller-Rabin bases 2..37; nearly every base-2 survivor is prime, and
// one Lucas test (a product and a square per bit) costs about two base-2 tests.
// With stride > 1 only the survivors in every stride-th word are tested; the other words are
// cleared, so a q left unproven reads as composite and the marking resolves an n only through a
// proven prime (each n still meets about 65 / stride proven q among its 1,027 window primes).
// copy (nullable): the pre-sieved window is saved there first, for resolve_open.
template <int BS>
__device__ void mr_tile(uint32_t* __restrict__ win32, uint32_t nbits, uint64_t qb, uint32_t tid,
                        uint32_t* __restrict__ queue, uint32_t qcap, uint32_t* __restrict__ wsum, uint32_t stride,
                        uint64_t* __restrict__ copy) {
    const uint64_t* win64 = reinterpret_cast<const uint64_t*>(win32);
    const uint32_t nw64 = (nbits + 63) >> 6;
    if (stride > 1) {
        for (uint32_t k = tid; k < nw64; k += BS) {
            if (copy) copy[k] = win64[k];
            if (k % stride) reinterpret_cast<uint64_t*>(win32)[k] = 0;
        }
        __syncthreads();
    }
    for (uint32_t pass = 0; pass < 2; ++pass) {
        for (uint32_t wb = 0; wb < nw64; wb += BS) {
            const uint32_t k = wb + tid;
            uint64_t w = k < nw64 ? win64[k] : 0;
            if (k == nw64 - 1 && (nbits & 63)) w &= (1ull << (nbits & 63)) - 1;
            uint32_t total;
            const uint32_t off = block_scan<BS>((uint32_t)__popcll(w), tid, wsum, &total);
            for (uint32_t base = 0; base < total; base += qcap) {
                uint32_t pos = off;
                for (uint64_t x = w; x; x &= x - 1, ++pos)
                    if (pos - base < qcap) queue[pos - base] = 64 * k + __builtin_ctzll(x);
                __syncthreads();
                const uint32_t n = total - base < qcap ? total - base : qcap;
                for (uint32_t u = tid; u < n; u += BS) {
                    const uint32_t i = queue[u];
                    const uint64_t q = qb + 2ull * i;
                    if (!(pass == 0 ? mr_test(q, 0, false) : bpsw_lucas(q))) lds_clear_bit(win32, i);
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// sieve_tile: window bits [g0, g0 + nbits) of a window whose bit i (>= lead)
// holds q = qv0 + 2(i - lead).  Result in win32[0 .. ceil(nbits/32)).
// pr[k] describes sieving prime sp[k] relative to the odd rbase <= q_lo
// (q_lo - rbase < 2^52): its first multiple sits e bits above rbase, so in
// this tile at bit lb + ((e - (q_lo - rbase)/2) mod p), one fp64 reduction.
// ---------------------------------------------------------------------------

// (e - db) mod p for y = e - db exact in fp64 (|y| < 2^52): the quotient
// estimate is off by at most one, so the remainder lies in (-p, 2p) and two
// unsigned min fix-ups finish it.
__device__ __forceinline__ uint32_t first_off(double y, uint32_t p, double inv) {
    const double q = floor(y * inv);
    uint32_t u = (uint32_t)(int32_t)fma(-q, (double)p, y);
    u = min(u, u + p);   // negative: u wrapped high, u + p is the residue
    u = min(u, u - p);   // >= p: u - p is the residue; else u - p wraps high
    return u;
}

// a[i] through a 32-bit byte offset from the table's (uniform) base, so the load can take
// the scalar-base + 32-bit vector-offset form instead of a 64-bit address per lane (every
// table read this way holds well under 2^32 bytes)
template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ a, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + (i * (uint32_t)sizeof(T)));
}

// MR = false leaves the Baillie-PSW step out of the instantiation, for windows whose q all lie
// below (cover + 1)^2 (the callers check): the strong tests would otherwise set the register
// budget, and so the occupancy, of every plain sieve launch.
// Returns whether the tile's survivors went through the strong tests (q above the table's reach).
template <int BS, bool MR = true>
__device__ bool sieve_tile(uint32_t* __restrict__ win32, uint32_t* __restrict__ scratch, uint32_t scratch_cap,
                           uint32_t* __restrict__ wsum, uint64_t qv0, uint32_t lead, uint64_t g0, uint32_t nbits,
                           const Tables& T, uint32_t tid, const PrimeRes* __restrict__ pr, uint64_t rbase,
                           uint32_t proof_stride = 1, uint64_t* __restrict__ copy = nullptr) {
    uint32_t* med_i0 = scratch;   // medium-prime first marks; the Miller-Rabin queue afterwards
    uint64_t* win64 = reinterpret_cast<uint64_t*>(win32);
    const uint32_t nw64 = (nbits + 63) >> 6;
    const uint64_t b_lo = g0 > lead ? g0 : (uint64_t)lead;
    const bool any_valid = b_lo < g0 + nbits;
    const uint32_t lb = any_valid ? (uint32_t)(b_lo - g0) : nbits;
    const uint64_t q_lo = any_valid ? qv0 + 2ull * (b_lo - lead) : 3;
    const uint64_t q_top = any_valid ? q_lo + 2ull * (nbits - 1 - lb) : 3;
    const uint64_t s = isqrt64(q_top);
    const bool mr = s > T.cover;
    const uint32_t np = count_le(T.sp, T.nsp, mr ? (uint64_t)(proof_stride > 1 ? MR_PRESIEVE_SPARSE : MR_PRESIEVE) : s);
    const double dbf = (double)((q_lo - rbase) >> 1);   // bits from rbase to q_lo, < 2^51

    // 1) primes 3..61 as per-word masks: bit j of word k is q(64k+j); the
    //    multiples of p in a word sit at j0, j0+p, ... with j0 stepping down by
    //    64*BS mod p between this thread's words.
    uint32_t j0[NPAT];
#pragma unroll
    for (int i = 0; i < NPAT; ++i) {
        const uint32_t p = pat_p(i);
        const uint32_t r = (uint32_t)((q_lo % p + p - (2ull * lb) % p) % p);   // q at local bit 0, mod p
        const uint32_t res = (r + (128u * tid) % p) % p;
        const uint32_t neg = res ? p - res : 0;
        j0[i] = (neg * ((p + 1) / 2)) % p;
    }
    const bool low = g0 < lead;   // only the window's first tile holds q < 3
    for (uint32_t k = tid; k < nw64; k += BS) {
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < NPAT; ++i) {
            const uint32_t p = pat_p(i);
            acc |= pat_mask(p) << j0[i];
            const uint32_t t = j0[i] - (64u * BS) % p;   // wraps high when negative
            j0[i] = min(t, t + p);
        }
        uint64_t w = ~acc;
        if (low) {
            const uint64_t gb = g0 + 64ull * k;
            if (gb < lead) {
                const uint64_t zb = lead - gb;
                w = zb >= 64 ? 0 : (w & (~0ull << zb));
            }
        }
        win64[k] = w;
    }

    // 2) first marks of the medium primes (split into CHUNK-mark items)
    const uint32_t kmed_end = np < T.n_med ? np : T.n_med;
    //    An entry holds the first mark p*m0 with 3 not dividing m0 and, in bit
    //    31, which of the strides p, 2p follows it: the odd m prime to 3 step
    //    +2, +4 from m = 5 mod 6 and +4, +2 from m = 1 mod 6.
    const uint32_t ql3 = (uint32_t)(q_lo % 3);
    for (uint32_t k = NPAT + tid; k < kmed_end; k += BS) {
        uint64_t i0 = nbits;
        uint32_t two_p_first = 0;
        if (any_valid) {
            const PrimeRes e = pr[k];
            i0 = lb + first_off((double)e.e - dbf, e.p, e.inv);
            const uint64_t pp = (uint64_t)e.p * e.p;   // the first mark is p^2 or above
            if (q_lo < pp) i0 = max(i0, lb + ((pp - q_lo) >> 1));
            if (i0 < nbits) {
                const uint32_t q3 = (ql3 + 2 * ((uint32_t)(i0 - lb) % 3)) % 3;   // p*m0 mod 3
                if (q3 == 0) {
                    i0 += e.p;                                   // m0 + 2 = 5 mod 6
                } else {
                    two_p_first = (q3 * (e.p % 3)) % 3 == 1;     // m0 = 1 mod 6
                }
            }
        }
        med_i0[k - NPAT] = (i0 < nbits ? (uint32_t)i0 : nbits) | (two_p_first << 31);
    }
    __syncthreads();

    // the pattern primes are themselves prime (marks below never touch q < 67^2)
    if (any_valid && qv0 <= 61 && tid < NPAT) {
        const uint32_t p = pat_p(tid);
        if (p >= qv0) {
            const uint64_t bi = lead + (p - qv0) / 2;
            if (bi >= g0 && bi < g0 + nbits) atomicOr(&win32[(bi - g0) >> 5], 1u << ((bi - g0) & 31));
        }
    }

    // marks at raised issue priority, as in the wheel tiles
    __builtin_amdgcn_s_setprio(1);
    // 3a) medium items: CHUNK marks each (unrolled when the item is whole);
    //     the next item's words are loaded before this item's marks issue
    const uint32_t n_items = kmed_end > NPAT ? T.item_start[kmed_end] : 0;
    {
        uint32_t u = tid, it_n = 0, p_n = 0;
        if (u < n_items) {
            it_n = T.items[u];
            p_n = T.items_p[u];
        }
        while (u < n_items) {
            const uint32_t it = it_n, p = p_n, un = u + BS;
            if (un < n_items) {
                it_n = T.items[un];
                p_n = T.items_p[un];
            }
            const uint32_t f = med_i0[(it >> 12) - NPAT];
            const uint32_t i = (f & 0x7FFFFFFFu) + (it & 4095u) * (CHUNK_SPAN * p);
            const uint32_t s1 = (f >> 31) ? 2 * p : p;   // the two runs: i + 3pt and i + s1 + 3pt
            if (i + (CHUNK_SPAN - 1) * p < nbits) {
                clear_run_n<CHUNK / 2>(win32, i, 3 * p);
                clear_run_n<CHUNK / 2>(win32, i + s1, 3 * p);
            } else {
                clear_run(win32, i, 3 * p, nbits);
                clear_run(win32, i + s1, 3 * p, nbits);
            }
            u = un;
        }
    }
    // 3b) large primes, one per thread; the next prime's entry is loaded
    //     before this prime's marks issue.  Primes p <= sqrt(q_lo) need no
    //     p^2 adjustment (all but the last few of a tile).
    __builtin_amdgcn_s_setprio(2);
    if (any_valid) {
        const uint64_t root_lo = isqrt64(q_lo);
        uint32_t k = T.n_med + tid;
        PrimeRes nx{0, 0, 0.0};
        if (k < np) nx = pr[k];
        while (k < np) {
            const PrimeRes e = nx;
            const uint32_t kn = k + BS;
            if (kn < np) nx = pr[kn];
            uint32_t i0 = lb + first_off((double)e.e - dbf, e.p, e.inv);
            if (e.p > root_lo) {
                const uint64_t pp = (uint64_t)e.p * e.p;
                if (q_lo < pp) {
                    const uint64_t f = lb + ((pp - q_lo) >> 1);
                    i0 = f < nbits ? max(i0, (uint32_t)f) : nbits;
                }
            }
            clear_run(win32, i0, e.p, nbits);
            k = kn;
        }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();

    // 4) above 2^44 the table cannot reach sqrt(q): survivors of the pre-sieve
    //    are decided by Baillie-PSW (exact below 2^64).
    if constexpr (MR) {
        if (mr && any_valid) {
            mr_tile<BS>(win32, nbits, qv0 + 2 * g0 - 2ull * lead, tid, scratch, scratch_cap, wsum, proof_stride,
                        copy);
            return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------
// statistics
// ---------------------------------------------------------------------------
struct LaneStats {
    uint64_t ph2, fail, first, sum, mx, arg;
};

__device__ __forceinline__ LaneStats lane_stats_zero() {
    LaneStats s;
    s.ph2 = 0;
    s.fail = 0;
    s.first = ~0ull;
    s.sum = 0;
    s.mx = 0;
    s.arg = ~0ull;
    return s;
}

__device__ __forceinline__ void merge_max(uint64_t& mx, uint64_t& arg, uint64_t m2, uint64_t a2) {
    if (m2 > mx || (m2 == mx && a2 < arg)) { mx = m2; arg = a2; }
}

__device__ __forceinline__ void wave_reduce(LaneStats& s) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s.ph2 += __shfl_xor(s.ph2, o);
        s.fail += __shfl_xor(s.fail, o);
        const uint64_t f = __shfl_xor(s.first, o);
        s.first = f < s.first ? f : s.first;
        s.sum += __shfl_xor(s.sum, o);
        const uint64_t m2 = __shfl_xor(s.mx, o), a2 = __shfl_xor(s.arg, o);
        merge_max(s.mx, s.arg, m2, a2);
    }
}

// Block-wide merge of lane statistics into one record (all threads must call).
template <int BS>
__device__ void block_store(LaneStats s, uint64_t evens, StatRec* out, uint32_t tid, LaneStats* red) {
    wave_reduce(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        LaneStats t = red[0];
        for (int w = 1; w < BS / 64; ++w) {
            const LaneStats& u = red[w];
            t.ph2 += u.ph2;
            t.fail += u.fail;
            t.first = u.first < t.first ? u.first : t.first;
            t.sum += u.sum;
            merge_max(t.mx, t.arg, u.mx, u.arg);
        }
        StatRec r;
        r.phase2 = t.ph2;
        r.failures = t.fail;
        r.first_fail = t.first;
        r.sum_pmin = t.sum;
        r.max_pmin = t.mx;
        r.argmax = t.arg;
        r.evens = evens;
        r.pad = 0;
        *out = r;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Unrolled marking over the odd primes p <= HALO_P.  Every such p reads the
// window at the fixed offset HALO - (p-1)/2, so word index and shift are
// compile-time constants: LDS reads carry immediate offsets (repeated words
// are loaded once) and each prime costs two alignbit + two OR per lane.
// ---------------------------------------------------------------------------
struct MarkTable {
    uint16_t p[NMARK];
};

constexpr MarkTable make_mark_table() {
    MarkTable t{};
    int n = 0;
    for (uint32_t q = 3; q <= HALO_P && n < NMARK; q += 2) {
        bool pr = true;
        for (uint32_t d = 3; d * d <= q; d += 2)
            if (q % d == 0) { pr = false; break; }
        if (pr) t.p[n++] = (uint16_t)q;
    }
    return t;
}

constexpr MarkTable kMark = make_mark_table();
static_assert(kMark.p[0] == 3 && kMark.p[NMARK - 1] == 8191, "odd primes <= 8193");

// Bits [S, S+64) of hi:lo with S a compile-time constant: two v_alignbit_b32
// (a 64-bit shift pair would cost three instructions).
template <uint32_t S>
__device__ __forceinline__ uint64_t funnel_c(uint64_t lo, uint64_t hi) {
    const uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
    const uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
    if constexpr (S == 0) {
        return lo;
    } else if constexpr (S < 32) {
        return ((uint64_t)__builtin_amdgcn_alignbit(h0, l1, S) << 32) | __builtin_amdgcn_alignbit(l1, l0, S);
    } else if constexpr (S == 32) {
        return ((uint64_t)h0 << 32) | l1;
    } else {
        return ((uint64_t)__builtin_amdgcn_alignbit(h1, h0, S - 32) << 32) | __builtin_amdgcn_alignbit(h0, l1, S - 32);
    }
}

template <int K>
__device__ __forceinline__ void mark_one(const uint64_t* __restrict__ w, uint64_t& v) {
    if constexpr (K < NMARK) {
        constexpr uint32_t p = kMark.p[K];
        constexpr uint32_t off = HALO - (p - 1) / 2;
        v |= funnel_c<off & 63>(w[off >> 6], w[(off >> 6) + 1]);
    }
}

template <int K, int... I>
__device__ __forceinline__ void mark_group(const uint64_t* __restrict__ w, uint64_t& v,
                                           std::integer_sequence<int, I...>) {
    (mark_one<K + I>(w, v), ...);
}

// primes [K, KEND) in groups of 8; with EXIT the wave stops once all 64 words are set.
// The empty asm between groups keeps each group's window reads inside it (no
// hoisting of all ~65 words into registers up front).
template <int K, int KEND, bool EXIT>
__device__ __forceinline__ void mark_span(const uint64_t* __restrict__ w, uint64_t& v) {
    if constexpr (K < KEND && K < NMARK) {
        mark_group<K>(w, v, std::make_integer_sequence<int, 8>{});
        if constexpr (K + 8 < KEND && K + 8 < NMARK) {
            asm volatile("" ::: "memory");
            if (!EXIT || !__all(v == ~0ull)) mark_span<K + 8, KEND, EXIT>(w, v);
        }
    }
}

// Compaction stages: every word runs primes [0, MK1); words still incomplete
// are packed 64 at a time and run [MK1, MK2); the rest are packed again and
// run to completion.  A wave then pays for the slowest of 64 similar words
// instead of the slowest of 64 arbitrary ones.
constexpr int MK1 = 48;
constexpr int MK2 = 96;

// Up to 128 pending (word, state) entries, two register slots per lane.
struct Pending {
    uint32_t w0, w1;
    uint64_t v0, v1;
    uint32_t n;   // wave-uniform count
};

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Append the lanes with `take` set to P.  One forward permute moves every
// lane: takers to P.n + rank (mod 64), the others to the lanes after them, so
// each lane receives exactly one value.
__device__ __forceinline__ void pend_push(Pending& P, bool take, uint32_t w, uint64_t v, uint32_t lane) {
    const uint64_t T = __ballot(take);
    const uint32_t U = __popcll(T);
    if (U == 0) return;
    const uint32_t rank = take ? lane_rank(T) : U + lane_rank(~T);
    const int dst = (int)(((P.n + rank) & 63u) << 2);
    const uint32_t rw = (uint32_t)__builtin_amdgcn_ds_permute(dst, (int)w);
    const uint32_t rlo = (uint32_t)__builtin_amdgcn_ds_permute(dst, (int)(uint32_t)v);
    const uint32_t rhi = (uint32_t)__builtin_amdgcn_ds_permute(dst, (int)(uint32_t)(v >> 32));
    const uint32_t off = (lane - P.n) & 63u;
    if (off < U) {
        const uint64_t rv = ((uint64_t)rhi << 32) | rlo;
        if (P.n + off < 64) { P.w0 = rw; P.v0 = rv; }
        else { P.w1 = rw; P.v1 = rv; }
    }
    P.n += U;
}

__device__ __forceinline__ void pend_pop(Pending& P) {
    P.w0 = P.w1;
    P.v0 = P.v1;
    P.n -= 64;
}

// ---------------------------------------------------------------------------
// mark_words: evens n = n_first + 2j, j in [0, cnt).  win64 bit (HALO + j - (p-1)/2)
// holds q = n - p.  Lanes own 64-even words; the prime loop is wave-uniform.
// The compacted unrolled path serves p_small >= HALO_P (every marking prime
// is a Phase-1 prime) without p_min statistics; the table-driven loop serves
// the rest (small p_small, the Phase-2 test hook, p_min statistics and the
// per-n p_min output).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t word_valid(uint32_t w, uint32_t nw, uint32_t cnt) {
    return (w == nw - 1 && (cnt & 63)) ? ((1ull << (cnt & 63)) - 1) : ~0ull;
}

// The n a tile with sparse proofs leaves open, kept in LDS for its second pass (resolve_open):
// their offsets j (n = n_first + 2j); past cap they go to the slow list.
struct OpenList {
    uint32_t* j;
    uint32_t* count;   // entries appended (may exceed cap)
    uint32_t cap;
};

// Results of one word whose marking is complete: verified bits, Phase-2
// count, and the n no window prime resolved (failures, or the open or slow list).
__device__ __forceinline__ void finish_word(uint32_t w, uint64_t v1, uint64_t v, uint64_t valid, uint64_t n_first,
                                            const MarkParams& mp, uint64_t* __restrict__ verified,
                                            uint32_t* __restrict__ pmin_out, const SlowList& sl, uint32_t win_idx,
                                            LaneStats& st, uint32_t step = 2, const OpenList* ol = nullptr) {
    st.ph2 += __popcll(~v1 & valid);
    uint64_t u = ~v & valid;
    if (u) {
        if (mp.fast_final) {
            st.fail += __popcll(u);
            const uint64_t n = n_first + (uint64_t)step * (64ull * w + __builtin_ctzll(u));
            st.first = n < st.first ? n : st.first;
            if (pmin_out)
                while (u) {
                    const uint32_t b = __builtin_ctzll(u);
                    u &= u - 1;
                    pmin_out[64u * w + b] = 0;
                }
        } else {
            while (u) {
                const uint32_t b = __builtin_ctzll(u);
                u &= u - 1;
                if (ol) {
                    const uint32_t o = atomicAdd(ol->count, 1u);
                    if (o < ol->cap) {
                        ol->j[o] = 64u * w + b;
                        continue;
                    }
                }
                const uint32_t slot = atomicAdd(sl.count, 1u);
                if (slot < sl.cap) {
                    sl.n[slot] = n_first + (uint64_t)step * (64ull * w + b);
                    sl.win[slot] = win_idx;
                } else {
                    atomicOr(sl.overflow, 1u);
                }
            }
        }
    }
    if (verified) verified[w] = v1 & valid;
}

template <int BS>
__device__ void mark_words_fast(const uint64_t* __restrict__ win64, uint64_t n_first, uint32_t cnt,
                                const MarkParams& mp, uint64_t* __restrict__ verified, const SlowList& sl,
                                uint32_t win_idx, LaneStats& st, uint32_t tid, const OpenList* ol) {
    const uint32_t lane = tid & 63, wave = tid >> 6;
    const uint32_t nw = (cnt + 63) >> 6;
    Pending P1, P2;
    P1.n = 0; P2.n = 0;
    P1.w0 = P1.w1 = P2.w0 = P2.w1 = 0;
    P1.v0 = P1.v1 = P2.v0 = P2.v1 = ~0ull;
    auto done = [&](uint32_t w, uint64_t v) {
        finish_word(w, v, v, word_valid(w, nw, cnt), n_first, mp, verified, nullptr, sl, win_idx, st, 2, ol);
    };
    // stage B on the 64 entries in P1's first slot; incomplete ones move to P2
    auto run_b = [&](uint32_t live) {
        const bool ok = lane < live;
        const uint32_t w = ok ? P1.w0 : 0;
        uint64_t v = ok ? P1.v0 : ~0ull;
        mark_span<MK1, MK2, true>(win64 + w, v);
        const bool unf = v != ~0ull;
        if (ok && !unf) done(w, v);
        pend_push(P2, unf, w, v, lane);
    };
    auto run_c = [&](uint32_t live) {
        const bool ok = lane < live;
        const uint32_t w = ok ? P2.w0 : 0;
        uint64_t v = ok ? P2.v0 : ~0ull;
        mark_span<MK2, NMARK, true>(win64 + w, v);
        if (ok) done(w, v);
    };
    for (uint32_t base = wave * 64; base < nw; base += BS) {
        const uint32_t w = base + lane;
        const bool act = w < nw;
        const uint32_t wr = act ? w : 0;
        uint64_t v = act ? ~word_valid(w, nw, cnt) : ~0ull;
        if (act && w == 0 && n_first == 4) v |= 1;   // 4 = 2 + 2, the only partition with p = 2
        mark_span<0, MK1, false>(win64 + wr, v);
        const bool unf = v != ~0ull;
        if (act && !unf) done(w, v);
        pend_push(P1, unf, wr, v, lane);
        if (P1.n >= 64) {
            run_b(64);
            pend_pop(P1);
        }
        if (P2.n >= 64) {
            run_c(64);
            pend_pop(P2);
        }
    }
    if (P1.n) {
        run_b(P1.n);
        P1.n = 0;
    }
    while (P2.n) {
        const uint32_t live = P2.n < 64 ? P2.n : 64;
        run_c(live);
        if (P2.n >= 64) pend_pop(P2);
        else P2.n = 0;
    }
}

// ol (nullable): the open list of a tile with sparse proofs (the unrolled path serves those).
template <int BS>
__device__ void mark_words(const uint64_t* __restrict__ win64, uint64_t n_first, uint32_t cnt,
                           const MarkParams& mp, const uint32_t* __restrict__ sp, uint64_t* __restrict__ verified,
                           uint32_t* __restrict__ pmin_out, const SlowList& sl, uint32_t win_idx, LaneStats& st,
                           uint32_t tid, const OpenList* ol = nullptr) {
    const bool want = mp.flags & 1u;
    if (mp.K == NMARK && mp.p_small >= HALO_P && pmin_out == nullptr && !want) {
        mark_words_fast<BS>(win64, n_first, cnt, mp, verified, sl, win_idx, st, tid, ol);
        return;
    }
    const uint32_t lane = tid & 63, wave = tid >> 6;
    const uint32_t nw = (cnt + 63) >> 6;
    for (uint32_t base = wave * 64; base < nw; base += BS) {
        const uint32_t w = base + lane;
        const bool act = w < nw;
        const uint32_t wr = act ? w : 0;
        const uint64_t valid = act ? word_valid(w, nw, cnt) : 0;
        uint64_t v = ~valid;
        const bool has4 = act && w == 0 && n_first == 4;
        if (has4 && mp.bound >= 2) {
            v |= 1;   // 4 = 2 + 2, the only partition with p = 2
            if (want) {
                st.sum += 2;
                merge_max(st.mx, st.arg, 2, 4);
                if (pmin_out) pmin_out[0] = 2;
            }
        }
        uint64_t v1 = ~valid | ((has4 && mp.p_small >= 2) ? 1ull : 0ull);
        bool crossed = mp.p_small < 3;
        for (uint32_t k = 0; k < mp.K; ++k) {
            const uint32_t p = sp[k];
            if (!crossed && p > mp.p_small) {
                v1 = v;
                crossed = true;
            }
            const uint32_t off = HALO - ((p - 1) >> 1);
            const uint32_t g = off >> 6, s = off & 63;
            const uint64_t R = funnel64(win64[wr + g], win64[wr + g + 1], s);
            if (want) {
                uint64_t nb = R & ~v;
                if (nb) {
                    st.sum += (uint64_t)__popcll(nb) * p;
                    merge_max(st.mx, st.arg, p, n_first + 2ull * (64ull * w + __builtin_ctzll(nb)));
                    if (pmin_out)
                        while (nb) {
                            const uint32_t b = __builtin_ctzll(nb);
                            nb &= nb - 1;
                            pmin_out[64u * w + b] = p;
                        }
                }
            }
            v |= R;
            if ((k & 7) == 7 && __all(v == ~0ull)) break;
        }
        if (!crossed) v1 = v;
        if (act) finish_word(w, v1, v, valid, n_first, mp, verified, pmin_out, sl, win_idx, st);
    }
}

// ---------------------------------------------------------------------------
// Second pass of a tile with sparse proofs.  An open n met no proven prime q = n - p with p
// prime, so its other candidates are the pre-sieve survivors of the unproven words: q = n - p
// with p <= 8191 prime and q's window bit set in a word left unproven.  A lane walks them in
// p order by reading the saved window downward from q = n - 3 (one word per step, the set bits
// from the top) and checking p against the bitset of odd primes; a group of L lanes
// (L = BS / m, at most 8, inside one wave) takes an n and tests its candidates L at a time by
// Baillie-PSW until one is prime (a survivor is prime about 70 % of the time), so a tile's few
// open n keep its waves busy for a round or two.  win64 holds the saved pre-sieve window; an n
// with no prime candidate goes to the slow list.
// ---------------------------------------------------------------------------
struct OpenWalk {
    uint32_t kw;     // window word being read
    uint32_t kend;   // the lowest word that may hold a candidate (p <= 8191)
    uint64_t bits;   // its set bits not yet taken
};

// The walk of the open n = n_first + 2j from q = n - 3 (window bit top - 1, top = HALO + j) down to
// q = n - 8191 (bit top - 4095).
__device__ __forceinline__ OpenWalk open_walk(const uint64_t* __restrict__ win64, uint32_t top, uint32_t stride) {
    OpenWalk wk;
    wk.kw = (top - 1) >> 6;
    wk.kend = (top - (HALO_P - 3) / 2) >> 6;
    wk.bits = wk.kw % stride ? win64[wk.kw] & (~0ull >> (63 - ((top - 1) & 63))) : 0;
    return wk;
}

// An open n no candidate resolved: to the slow list, which decides it exactly.
__device__ __forceinline__ void open_to_slow(const SlowList& sl, uint64_t n, uint32_t win_idx) {
    const uint32_t slot = atomicAdd(sl.count, 1u);
    if (slot < sl.cap) {
        sl.n[slot] = n;
        sl.win[slot] = win_idx;
    } else {
        atomicOr(sl.overflow, 1u);
    }
}

// Next candidate bit of the walk (p ascending), or ~0u when none is left.
__device__ __forceinline__ uint32_t open_next(OpenWalk& wk, const uint64_t* __restrict__ win64, uint32_t top,
                                              uint32_t stride, const uint32_t* __restrict__ podd) {
    for (;;) {
        while (wk.bits) {
            const uint32_t t = 63 - __builtin_clzll(wk.bits);
            wk.bits &= ~(1ull << t);
            const uint32_t b = 64 * wk.kw + t;
            const uint32_t h = top - b;   // (p - 1) / 2, p = 2h + 1
            if (h > (HALO_P - 3) / 2) return ~0u;
            if ((podd[h >> 5] >> (h & 31)) & 1) return b;
        }
        if (wk.kw == wk.kend) return ~0u;
        --wk.kw;
        wk.bits = wk.kw % stride ? win64[wk.kw] : 0;
    }
}

template <int BS>
__device__ void resolve_open(const uint64_t* __restrict__ win64, uint64_t n_first, const uint32_t* __restrict__ open_j,
                             uint32_t m, uint32_t* __restrict__ next, uint32_t stride,
                             const uint32_t* __restrict__ podd, const SlowList& sl, uint32_t win_idx, LaneStats& st,
                             uint32_t tid) {
    if (m >= BS) {
        // more open n than lanes: a lane takes the next n from the list whenever its own is decided
        bool have = false;
        uint32_t top = 0;
        uint64_t n = 0;
        OpenWalk wk{};
        auto take = [&]() {
            const uint32_t i = atomicAdd(next, 1u);
            have = i < m;
            if (have) {
                top = HALO + open_j[i];
                n = n_first + 2ull * (top - HALO);
                wk = open_walk(win64, top, stride);
            }
        };
        take();
        while (__any(have)) {
            uint64_t q = 0;
            bool has = false;
            if (have) {
                const uint32_t b = open_next(wk, win64, top, stride, podd);
                if (b != ~0u) {
                    q = n - (2ull * (top - b) + 1);
                    has = true;
                } else {
                    open_to_slow(sl, n, win_idx);
                    take();
                }
            }
            if (has && mr_test(q, 0, false) && bpsw_lucas(q)) {
                st.ph2 -= 1;   // counted unverified by the marking
                take();
            }
        }
        return;
    }
    const uint32_t L = m >= BS / 8 ? BS / m : 8u;   // lanes per open n
    const uint32_t G = 64 / L;                      // groups per wave
    const uint32_t lane = tid & 63, r = lane % L, lead = lane - r;
    const uint64_t gmask = ((1ull << L) - 1) << lead;
    for (uint32_t base = 0; base < m; base += (BS / 64) * G) {
        const uint32_t i = base + (tid >> 6) * G + lane / L;
        const bool mine = lane < G * L && i < m;
        const uint32_t top = HALO + (mine ? open_j[i] : 0);
        const uint64_t n = n_first + 2ull * (top - HALO);
        OpenWalk wk = open_walk(win64, top, stride);
        bool open = mine, found = false;
        uint32_t skip = r;   // candidates this lane passes over before its next test
        while (__any(open)) {
            uint64_t q = 0;
            bool has = false;
            if (open) {
                uint32_t b = open_next(wk, win64, top, stride, podd);
                while (b != ~0u && skip) {
                    --skip;
                    b = open_next(wk, win64, top, stride, podd);
                }
                if (b != ~0u) {
                    q = n - (2ull * (top - b) + 1);
                    has = true;
                    skip = L - 1;
                } else {
                    open = false;   // this lane's share of the candidates is used up
                }
            }
            const bool prime = has && mr_test(q, 0, false) && bpsw_lucas(q);
            if (__ballot(prime) & gmask) {
                found = true;
                open = false;
            }
        }
        if (mine && r == 0) {
            if (found) st.ph2 -= 1;   // counted unverified by the marking
            else open_to_slow(sl, n, win_idx);
        }
    }
}

// ---------------------------------------------------------------------------
// plain sieve kernel: one tile of tile_bits window bits per workgroup
// ---------------------------------------------------------------------------
template <int BS, bool MR>
__global__ __launch_bounds__(BS) void sieve_window_kernel(Tables T, uint64_t qv0, uint32_t lead, uint64_t nbits,
                                                          uint32_t tile_bits, uint64_t* __restrict__ out,
                                                          int tail_ones, const PrimeRes* __restrict__ pr) {
    __shared__ __align__(16) uint32_t win32[SIEVE_WIN_WORDS];   // fixed LDS address: folds into ds offsets
    extern __shared__ uint32_t lds[];   // [scan words | medium first marks, then the Miller-Rabin queue]
    uint32_t* wsum = lds;
    uint32_t* scratch = lds + BS / 64;
    const uint32_t tid = threadIdx.x;
    const uint64_t g0 = (uint64_t)blockIdx.x * tile_bits;
    if (g0 >= nbits) return;
    const uint32_t tb = (uint32_t)(nbits - g0 < tile_bits ? nbits - g0 : tile_bits);
    sieve_tile<BS, MR>(win32, scratch, SIEVE_MRQ, wsum, qv0, lead, g0, tb, T, tid, pr, qv0);
    const uint64_t* win64 = reinterpret_cast<const uint64_t*>(win32);
    const uint32_t nw = (tb + 63) >> 6;
    const uint64_t wo = g0 >> 6;
    for (uint32_t k = tid; k < nw; k += BS) {
        uint64_t w = win64[k];
        if (k == nw - 1 && (tb & 63)) {
            const uint64_t tail = ~0ull << (tb & 63);
            w = tail_ones ? (w | tail) : (w & ~tail);
        }
        out[wo + k] = w;
    }
}

// ---------------------------------------------------------------------------
// plain marking kernel: words_per_block verified words per workgroup; the
// window slice (plus the 64-word halo) is staged HBM -> LDS once.
// ---------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void mark_segment_kernel(Tables T, MarkParams mp, uint64_t n0, uint64_t cnt,
                                                          const uint64_t* __restrict__ window, uint64_t window_words,
                                                          uint32_t wpb, uint64_t* __restrict__ verified,
                                                          uint32_t* __restrict__ pmin_out, SlowList sl,
                                                          StatRec* __restrict__ rec) {
    extern __shared__ uint32_t lds[];
    uint64_t* win64 = reinterpret_cast<uint64_t*>(lds);
    LaneStats* red = reinterpret_cast<LaneStats*>(lds + 2 * (wpb + HALO / 64 + 1));
    const uint32_t tid = threadIdx.x;
    const uint64_t w0 = (uint64_t)blockIdx.x * wpb;          // first verified word
    const uint64_t j0 = w0 * 64;
    const uint32_t lcnt = (uint32_t)(cnt - j0 < (uint64_t)wpb * 64 ? cnt - j0 : (uint64_t)wpb * 64);
    const uint32_t nwl = (lcnt + 63) >> 6;
    const uint32_t stage = nwl + HALO / 64 + 1;
    for (uint32_t k = tid; k < stage; k += BS) {
        const uint64_t gw = w0 + k;
        win64[k] = gw < window_words ? window[gw] : 0;
    }
    __syncthreads();
    LaneStats st = lane_stats_zero();
    mark_words<BS>(win64, n0 + 2 * j0, lcnt, mp, T.sp, verified ? verified + w0 : nullptr,
                   pmin_out ? pmin_out + j0 : nullptr, sl, 0, st, tid);
    block_store<BS>(st, lcnt, &rec[blockIdx.x], tid, red);
}

// ---------------------------------------------------------------------------
// fused persistent kernel: tiles of tile_evens evens from a device counter;
// sieve and marking share the LDS window, nothing but results leave the CU.
// ---------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void fused_kernel(Tables T, MarkParams mp, const WinDesc* __restrict__ wins,
                                                   uint32_t nwin, uint64_t ntiles, uint32_t tile_evens,
                                                   uint64_t* __restrict__ tile_counter, SlowList sl,
                                                   StatRec* __restrict__ rec, const PrimeRes* __restrict__ pr,
                                                   uint64_t* __restrict__ wcopy) {
    // the window at a fixed LDS address (sized for the largest tile); dynamic
    // LDS after it: [stats | tile id | scan words | open count | scratch]
    __shared__ __align__(16) uint32_t win32[FUSED_WIN_WORDS];
    extern __shared__ uint32_t lds[];
    LaneStats* red = reinterpret_cast<LaneStats*>(lds);
    uint64_t& s_tile = *reinterpret_cast<uint64_t*>(red + BS / 64);
    uint32_t* wsum = reinterpret_cast<uint32_t*>(&s_tile + 1);
    uint32_t* open_cnt = wsum + BS / 64;          // [open n appended | next open n to take]
    uint32_t* podd = open_cnt + 2;                // bit h: 2h + 1 is an odd prime <= 8191
    uint32_t* scratch = podd + (HALO_P + 63) / 64;   // medium first marks, the Miller-Rabin queue, then the open list
    const uint32_t tid = threadIdx.x;
    // sparse proofs: this workgroup's slot for the saved pre-sieve window
    uint64_t* const copy = (mp.flags & MP_SPARSE) ? wcopy + (uint64_t)blockIdx.x * FUSED_COPY_WORDS : nullptr;
    if (copy) {
        for (uint32_t k = tid; k < (HALO_P + 63) / 64; k += BS) podd[k] = 0;
        __syncthreads();
        for (uint32_t k = tid; k < NMARK; k += BS) {
            const uint32_t h = T.sp[k] >> 1;
            atomicOr(&podd[h >> 5], 1u << (h & 31));
        }
    }
    for (;;) {
        if (tid == 0) {
            s_tile = atomicAdd((unsigned long long*)tile_counter, 1ull);
            open_cnt[0] = 0;
            open_cnt[1] = 0;
        }
        __syncthreads();
        const uint64_t t = s_tile;
        __syncthreads();
        if (t >= ntiles) break;
        // window holding tile t
        uint32_t lo = 0, hi = nwin;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (wins[mid].tile0 <= t) lo = mid; else hi = mid;
        }
        const WinDesc wd = wins[lo];
        const uint64_t lt = t - wd.tile0;
        const uint64_t n0 = wd.start + 2ull * tile_evens * lt;
        const uint64_t rem = wd.evens - (uint64_t)tile_evens * lt;
        const uint32_t cnt = (uint32_t)(rem < tile_evens ? rem : tile_evens);
        const WinGeom g = window_geom(n0, (uint64_t)cnt + HALO + WIN_PAD);
        const bool sparse = sieve_tile<BS>(win32, scratch, FUSED_MRQ, wsum, g.qv0, g.lead, 0, g.nbits, T, tid,
                                           pr + wd.r0off, wd.rbase, copy ? mp.proof_stride : 1u, copy) && copy;
        LaneStats st = lane_stats_zero();
        const OpenList ol{scratch, open_cnt, FUSED_MRQ};
        mark_words<BS>(reinterpret_cast<const uint64_t*>(win32), n0, cnt, mp, T.sp, nullptr, nullptr, sl, lo, st,
                       tid, sparse ? &ol : nullptr);
        if (sparse) {
            __syncthreads();   // the open list is complete
            const uint32_t m = *open_cnt < FUSED_MRQ ? *open_cnt : FUSED_MRQ;
            if (m) {
                // the saved window back into LDS (written by this workgroup: the barriers since order it)
                uint64_t* win64 = reinterpret_cast<uint64_t*>(win32);
                for (uint32_t k = tid; k < (g.nbits + 63) >> 6; k += BS) win64[k] = copy[k];
                __syncthreads();
                resolve_open<BS>(win64, n0, scratch, m, open_cnt + 1, mp.proof_stride, podd, sl, lo, st, tid);
            }
        }
        block_store<BS>(st, cnt, &rec[t], tid, red);
    }
}

// ===========================================================================
// Wheel (mod 6) form of the fused fast path.
//
// The window keeps only q prime to 6, as two streams in LDS: bit i of stream
// S1 / S5 holds q = 6(K0 + i) + 1 / + 5.  The even n of a tile are split by
// n mod 6 into three classes, class c holding n = n0 + c + 6j (j < cls_len,
// n0 = 6(K0 + WH_HALO)).  For a prime p the q = n - p of class c lie in stream
// r = (c - p) mod 6 at the fixed bit offset WH_HALO + (c - p - r)/6 from j;
// when r = 3 they are all multiples of 3 and class c skips p.  Against the
// odd-only window this drops the multiples of 3 from every sieve pass, packs
// 1.5x more n into the same LDS, and lets each class OR only the primes that
// can give it a q prime to 6 (class 0 all of them, classes 2 and 4 half).
// ===========================================================================

__host__ __device__ constexpr uint32_t inv6_mod(uint32_t p) { return p % 6 == 1 ? (5 * p + 1) / 6 : (p + 1) / 6; }

// first bit b >= 0 with p | 6(k + b) + r, from km = k mod p (p prime to 6)
__device__ __forceinline__ uint32_t wheel_first(uint32_t km, uint32_t r, uint32_t p, uint32_t inv6) {
    const uint32_t x = (6 * km + r) % p;
    return (uint32_t)(((uint64_t)(x ? p - x : 0) * inv6) % p);
}

// First mark of the run off, off + p, ... in stream r (q = 6(K0 + i) + r) at or
// above p^2: q = p must stay, and the multiples below p^2 are cleared by their
// smaller factors anyway.
__device__ __forceinline__ uint32_t wheel_p2(uint32_t off, uint32_t p, uint32_t r, int64_t K0, uint32_t end) {
    const uint64_t pp = (uint64_t)p * p;
    const int64_t kmin = (int64_t)((pp - r + 5) / 6);   // smallest k with 6k + r >= p^2
    if (K0 + (int64_t)off >= kmin) return off;
    const uint64_t gap = (uint64_t)(kmin - K0 - (int64_t)off);
    const uint64_t i = off + (gap + p - 1) / p * p;
    return i >= end ? end : (uint32_t)i;
}

// (e - t lp) mod p for a tile t stream lengths above the window's base (WheelRes): 32-bit and
// fp32 arithmetic only, exact for t < 2^17.
__device__ __forceinline__ uint32_t wheel_off(const WheelRes& E, uint32_t t, float tf) {
    const uint32_t q = (uint32_t)(tf * E.ratio);                 // t lp / p, less one at most, or plus one
    uint32_t r = __umul24(t, E.lp) - __umul24(q, E.p);           // t lp mod p + {-p, 0, p}, wrapped
    r = min(r, r + E.p);                                         // negative: add p
    r = min(r, r - E.p);                                         // >= p: subtract p
    const uint32_t o = E.e - r;
    return min(o, o + E.p);                                      // e < r: add p
}

template <int BS, uint32_t SW>
__device__ void sieve_wheel_tile(uint32_t* __restrict__ wh, uint32_t* __restrict__ med0, int64_t K0, uint32_t nbits,
                                 const Tables& T, uint32_t tid, const WheelRes* __restrict__ pr, uint32_t tl,
                                 uint32_t np) {
    const uint32_t nw64 = (nbits + 63) >> 6;
    // 1) primes 5..61 as per-word masks in each stream; bits standing for q <= 1 are zero
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const uint32_t r = s ? 5u : 1u;
        uint64_t* sw = reinterpret_cast<uint64_t*>(wh + s * SW);
        uint32_t j0[NPAT - 1];
#pragma unroll
        for (int i = 1; i < NPAT; ++i) {
            const uint32_t p = pat_p(i);
            const uint32_t km = (uint32_t)(((uint64_t)(K0 + (int64_t)WH_HALO) % p + p - (WH_HALO % p)) % p);
            const uint32_t kt = (km + (64u * tid) % p) % p;      // k of this thread's first bit, mod p
            j0[i - 1] = wheel_first(kt, r, p, inv6_mod(p));
        }
        for (uint32_t k = tid; k < nw64; k += BS) {
            uint64_t acc = 0;
#pragma unroll
            for (int i = 1; i < NPAT; ++i) {
                const uint32_t p = pat_p(i);
                acc |= pat_mask(p) << j0[i - 1];
                const uint32_t t = j0[i - 1] - (64u * BS) % p;   // wraps high when negative
                j0[i - 1] = min(t, t + p);
            }
            uint64_t w = ~acc;
            const int64_t kw = K0 + 64 * (int64_t)k;             // k of the word's bit 0
            if (kw <= 0) {                                        // q = 6k + r <= 1 below k = 0 (+ q = 1)
                const int64_t zb = -kw + (r == 1 ? 1 : 0);        // low bits to clear
                w = zb >= 64 ? 0 : (w & (~0ull << zb));
            }
            sw[k] = w;
        }
    }
    // 2) first marks of the medium primes in both streams (np: the primes up to the square root
    //    of the tile's last q, from wheel_tiles_kernel)
    const uint32_t kmed_end = np < T.n_med ? np : T.n_med;
    const float tf = (float)tl;
    for (uint32_t k = NPAT + tid; k < kmed_end; k += BS) {
        const WheelRes e = pr[k];
        const uint32_t o1 = wheel_off(e, tl, tf);
        uint32_t o5 = o1 + T.d5[k];
        o5 = min(o5, o5 - e.p);
        med0[2 * (k - NPAT)] = wheel_p2(o1, e.p, 1, K0, nbits);
        med0[2 * (k - NPAT) + 1] = wheel_p2(o5, e.p, 5, K0, nbits);
    }
    __syncthreads();
    // pattern primes are themselves prime
    if (tid >= 1 && tid < NPAT) {
        const uint32_t p = pat_p(tid);
        const int64_t i = (int64_t)(p / 6) - K0;
        if (i >= 0 && i < (int64_t)nbits) atomicOr(&wh[(p % 6 == 5 ? SW : 0) + (uint32_t)(i >> 5)], 1u << (i & 31));
    }
    // The marks below are LDS-bound while the other workgroup on the CU is often in its
    // VALU-bound class marking: raising these waves' issue priority keeps the LDS fed
    // (measured: [4, 1e12] 159.7 -> 154.2 ms; the level itself hardly matters).
    __builtin_amdgcn_s_setprio(1);
    // 3a) wave items: a wavefront takes 64 * WAVE_MARKS marks of one prime in
    //     one stream, lane l the marks l, l + 64, ...  Those sit 64p bits apart:
    //     the same bit of words 2p apart, so a lane's clear-mask never changes
    //     and a mark costs one address add; the lanes' words lie ~p/32 apart.
    const uint32_t kw_end = kmed_end < T.k_wave ? kmed_end : T.k_wave;
    const uint32_t n_wave = kw_end > NPAT ? T.item_start[kw_end] : 0;
    {
        const uint32_t lane = tid & 63;
        for (uint32_t u = __builtin_amdgcn_readfirstlane(tid >> 6); u < n_wave; u += BS / 64) {
            const uint32_t it = T.items[u], p = T.items_p[u];
            const uint32_t sidx = (it >> 12) & 1u;
            uint32_t* S = wh + sidx * SW;
            const uint32_t i0 = med0[2 * ((it >> 13) - NPAT) + sidx];
            const uint32_t first = (it & 4095u) * (64 * WAVE_MARKS);   // the item's first mark number
            uint32_t b = i0 + (first + lane) * p;
            const uint32_t mask = clear_mask(b);
            if (i0 + (first + 64 * WAVE_MARKS - 1) * p < nbits) {
                uint32_t* wp = S + (b >> 5);
#pragma unroll
                for (uint32_t t = 0; t < WAVE_MARKS; ++t) atomicAnd(wp + 2 * p * t, mask);
            } else {
                for (uint32_t t = 0; t < WAVE_MARKS && b < nbits; ++t, b += 64 * p) atomicAnd(&S[b >> 5], mask);
            }
        }
    }
    // 3b) lane items: CHUNK consecutive marks of one prime in one stream per lane
    const uint32_t n_items = kmed_end > NPAT ? T.item_start[kmed_end] : 0;
    {
        uint32_t u = n_wave + tid, it_n = 0, p_n = 0;
        if (u < n_items) {
            it_n = ld32(T.items, u);
            p_n = ld32(T.items_p, u);
        }
        while (u < n_items) {
            const uint32_t it = it_n, p = p_n, un = u + BS;
            if (un < n_items) {
                it_n = ld32(T.items, un);
                p_n = ld32(T.items_p, un);
            }
            const uint32_t sidx = (it >> 12) & 1u;
            const uint32_t sb = sidx * (32 * SW);   // the stream's first bit in wh (a multiple of 32)
            const uint32_t i = med0[2 * ((it >> 13) - NPAT) + sidx] + (it & 4095u) * (CHUNK * p);
            // marks indexed from wh itself: no per-mark add of the stream's base address
            if (i + (CHUNK - 1) * p < nbits)
                clear_run_n<CHUNK>(wh, sb + i, p);
            else
                clear_run(wh, sb + i, p, sb + nbits);
            u = un;
        }
    }
    // 3c) large primes, one per thread, both streams
    __builtin_amdgcn_s_setprio(2);
    {
        const int64_t kroot = (int64_t)(isqrt64(6ull * (uint64_t)(K0 > 0 ? K0 : 0) + 1) );
        uint32_t k = T.n_med + tid;
        WheelRes nx{0, 0, 0, 0.0f};
        uint32_t d_n = 0;
        if (k < np) {
            nx = ld32(pr, k);
            d_n = ld32(T.d5, k);
        }
        while (k < np) {
            const WheelRes e = nx;
            const uint32_t d5 = d_n, kn = k + BS;
            if (kn < np) {
                nx = ld32(pr, kn);
                d_n = ld32(T.d5, kn);
            }
            uint32_t o1 = wheel_off(e, tl, tf);
            uint32_t o5 = o1 + d5;
            o5 = min(o5, o5 - e.p);
            if ((int64_t)e.p > kroot) {   // p^2 may lie above the tile's first q
                o1 = wheel_p2(o1, e.p, 1, K0, nbits);
                o5 = wheel_p2(o5, e.p, 5, K0, nbits);
            }
            clear_run(wh, o1, e.p, nbits);
            clear_run(wh + SW, o5, e.p, nbits);
            k = kn;
        }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
}

// marking of one class C (n = n_first + 6j): stream and offset per prime are compile-time
template <int C, int K>
__device__ __forceinline__ void wmark_one(const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s5,
                                          uint64_t& v) {
    if constexpr (K < NMARK) {
        constexpr int p = kMark.p[K];
        constexpr int r = ((C - p) % 6 + 6) % 6;
        if constexpr (r != 3) {
            constexpr int off = (int)WH_HALO + (C - p - r) / 6;
            static_assert(off >= 0, "halo covers every marking prime");
            const uint64_t* S = r == 1 ? s1 : s5;
            v |= funnel_c<(uint32_t)off & 63>(S[off >> 6], S[(off >> 6) + 1]);
        }
    }
}

template <int C, int K, int... I>
__device__ __forceinline__ void wmark_group(const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s5,
                                            uint64_t& v, std::integer_sequence<int, I...>) {
    (wmark_one<C, K + I>(s1, s5, v), ...);
}

template <int C, int K, int KEND, bool EXIT>
__device__ __forceinline__ void wmark_span(const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s5,
                                           uint64_t& v) {
    if constexpr (K < KEND && K < NMARK) {
        wmark_group<C, K>(s1, s5, v, std::make_integer_sequence<int, 8>{});
        if constexpr (K + 8 < KEND && K + 8 < NMARK) {
            asm volatile("" ::: "memory");
            if (!EXIT || !__all(v == ~0ull)) wmark_span<C, K + 8, KEND, EXIT>(s1, s5, v);
        }
    }
}

// Compaction stage bounds in prime indices (class 0 can use every prime, class
// 2 the primes 1 mod 6, class 4 those 5 mod 6), set by a cost model of the
// stages on the minimal partitions of 3 x 96,000 n near 1e12.
template <int C> constexpr int WK1 = C == 0 ? 40 : (C == 2 ? 72 : 56);
template <int C> constexpr int WK2 = C == 0 ? 56 : (C == 2 ? 112 : 88);

// Word-0 bit of n = 6 in class 0 (n = n_first + 6j): the tile base n0 is 0 or 6 there.
__device__ __forceinline__ uint64_t six_bit(uint64_t n_first) {
    return n_first == 0 ? 2ull : (n_first == 6 ? 1ull : 0ull);
}

// bits j of word w inside [jlo, jhi)
__device__ __forceinline__ uint64_t range_mask(uint32_t w, uint32_t jlo, uint32_t jhi) {
    const int64_t a = (int64_t)jlo - 64 * (int64_t)w, b = (int64_t)jhi - 64 * (int64_t)w;
    const uint64_t hi = b >= 64 ? ~0ull : (b <= 0 ? 0 : ((1ull << b) - 1));
    const uint64_t lo = a <= 0 ? ~0ull : (a >= 64 ? 0 : (~0ull << a));
    return hi & lo;
}

template <int BS, int C>
__device__ void wheel_mark_class(const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s5, uint64_t n_first,
                                 uint32_t nw, uint32_t jlo, uint32_t jhi, const MarkParams& mp, const SlowList& sl,
                                 uint32_t win_idx, LaneStats& st, uint32_t tid) {
    const uint32_t lane = tid & 63, wave = tid >> 6;
    Pending P1, P2;
    P1.n = 0; P2.n = 0;
    P1.w0 = P1.w1 = P2.w0 = P2.w1 = 0;
    P1.v0 = P1.v1 = P2.v0 = P2.v1 = ~0ull;
    auto done = [&](uint32_t w, uint64_t v) {
        finish_word(w, v, v, range_mask(w, jlo, jhi), n_first, mp, nullptr, nullptr, sl, win_idx, st, 6);
    };
    auto run_b = [&](uint32_t live) {
        const bool ok = lane < live;
        const uint32_t w = ok ? P1.w0 : 0;
        uint64_t v = ok ? P1.v0 : ~0ull;
        wmark_span<C, WK1<C>, WK2<C>, true>(s1 + w, s5 + w, v);
        const bool unf = v != ~0ull;
        if (ok && !unf) done(w, v);
        pend_push(P2, unf, w, v, lane);
    };
    auto run_c = [&](uint32_t live) {
        const bool ok = lane < live;
        const uint32_t w = ok ? P2.w0 : 0;
        uint64_t v = ok ? P2.v0 : ~0ull;
        wmark_span<C, WK2<C>, NMARK, true>(s1 + w, s5 + w, v);
        if (ok) done(w, v);
    };
    for (uint32_t base = wave * 64; base < nw; base += BS) {
        const uint32_t w = base + lane;
        const bool act = w < nw;
        const uint32_t wr = act ? w : 0;
        const uint64_t valid = act ? range_mask(w, jlo, jhi) : 0;
        uint64_t v = ~valid;
        if (act && w == 0) {
            if (C == 4 && n_first == 4) v |= 1;   // 4 = 2 + 2
            if (C == 0) v |= six_bit(n_first);    // 6 = 3 + 3 (q = 3 is not in the wheel)
        }
        wmark_span<C, 0, WK1<C>, false>(s1 + wr, s5 + wr, v);
        const bool unf = v != ~0ull;
        if (act && !unf) done(w, v);
        pend_push(P1, unf, wr, v, lane);
        if (P1.n >= 64) {
            run_b(64);
            pend_pop(P1);
        }
        if (P2.n >= 64) {
            run_c(64);
            pend_pop(P2);
        }
    }
    if (P1.n) {
        run_b(P1.n);
        P1.n = 0;
    }
    while (P2.n) {
        const uint32_t live = P2.n < 64 ? P2.n : 64;
        run_c(live);
        if (P2.n >= 64) pend_pop(P2);
        else P2.n = 0;
    }
}

// Table-driven marking of class c: p_min statistics, P_SMALL below the window
// primes and the Phase-2 test hook (the unrolled path serves the rest).
// Prime 3 gives class 0 only q = 3, i.e. n = 6, which the wheel does not hold.
template <int BS>
__device__ void wheel_mark_generic(const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s5, uint64_t n0,
                                   uint32_t c, uint32_t nw, uint32_t jlo, uint32_t jhi, const MarkParams& mp,
                                   const uint32_t* __restrict__ sp, const SlowList& sl, uint32_t win_idx,
                                   LaneStats& st, uint32_t tid) {
    const bool want = mp.flags & 1u;
    const uint32_t lane = tid & 63, wave = tid >> 6;
    const uint64_t n_first = n0 + c;
    for (uint32_t base = wave * 64; base < nw; base += BS) {
        const uint32_t w = base + lane;
        const bool act = w < nw;
        const uint32_t wr = act ? w : 0;
        const uint64_t valid = act ? range_mask(w, jlo, jhi) : 0;
        uint64_t v = ~valid;
        const bool has4 = act && c == 4 && w == 0 && n_first == 4 && (valid & 1);
        if (has4 && mp.bound >= 2) {
            v |= 1;   // 4 = 2 + 2, the only partition with p = 2
            if (want) {
                st.sum += 2;
                merge_max(st.mx, st.arg, 2, 4);
            }
        }
        uint64_t v1 = ~valid | ((has4 && mp.p_small >= 2) ? 1ull : 0ull);
        bool crossed = mp.p_small < 3;
        const uint64_t bit6 = (act && c == 0 && w == 0) ? six_bit(n_first) : 0;
        for (uint32_t k = 0; k < mp.K; ++k) {
            const uint32_t p = sp[k];
            if (!crossed && p > mp.p_small) {
                v1 = v;
                crossed = true;
            }
            const uint32_t r = (c + 6 * 1366 - p) % 6;   // (c - p) mod 6, p <= 8191
            uint64_t R;
            if (r == 3) {
                R = p == 3 ? bit6 : 0ull;               // 6 = 3 + 3
            } else {
                const uint32_t off = WH_HALO - (p + r - c) / 6;
                const uint64_t* S = (r == 1 ? s1 : s5) + wr + (off >> 6);
                R = funnel64(S[0], S[1], off & 63);
            }
            if (want) {
                const uint64_t nb = R & ~v;
                if (nb) {
                    st.sum += (uint64_t)__popcll(nb) * p;
                    merge_max(st.mx, st.arg, p, n_first + 6ull * (64ull * w + __builtin_ctzll(nb)));
                }
            }
            v |= R;
            if ((k & 7) == 7 && __all(v == ~0ull)) break;
        }
        if (!crossed) v1 = v;
        if (act) finish_word(w, v1, v, valid, n_first, mp, nullptr, nullptr, sl, win_idx, st, 6);
    }
}

// j range of class c inside [n_lo, n_hi] for n = n0 + c + 6j, j < len
__device__ __forceinline__ void class_range(uint64_t n0, uint32_t c, uint32_t len, uint64_t n_lo, uint64_t n_hi,
                                            uint32_t& jlo, uint32_t& jhi) {
    const uint64_t a = n0 + c;
    jlo = a >= n_lo ? 0u : (uint32_t)min((n_lo - a + 5) / 6, (uint64_t)len);
    jhi = a > n_hi ? 0u : (uint32_t)min((n_hi - a) / 6 + 1, (uint64_t)len);
    if (jhi < jlo) jhi = jlo;
}

// Tile t of a wheel launch (n0 = the window's first multiple of 6 below its start + 6 cls_len
// (t - tile0)) sieves stream bits [0, cls_len + WH_HALO + 64) from k = n0/6 - WH_HALO; its
// last q is 6(K0 + nbits - 1) + 5.
__device__ __forceinline__ uint64_t wheel_tile_n0(const WinDesc& wd, uint64_t t, uint32_t cls_len) {
    return (wd.start - wd.start % 6) + 6ull * cls_len * (t - wd.tile0);
}

// info[t] = (window index << 32) | number of sieving primes up to the square root of the tile's
// last q: the two binary searches of a tile, done once per launch for all tiles in parallel.
__global__ void wheel_tiles_kernel(Tables T, const WinDesc* __restrict__ wins, uint32_t nwin, uint64_t ntiles,
                                   uint32_t cls_len, uint64_t* __restrict__ info) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t lo = 0, hi = nwin;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (wins[mid].tile0 <= t) lo = mid; else hi = mid;
    }
    const WinDesc wd = wins[lo];
    const uint32_t nbits = cls_len + WH_HALO + 64;
    const int64_t K0 = (int64_t)(wheel_tile_n0(wd, t, cls_len) / 6) - (int64_t)WH_HALO;
    const uint64_t q_top = 6ull * (uint64_t)(K0 + (int64_t)nbits - 1) + 5;
    info[t] = ((uint64_t)lo << 32) | count_le(T.sp, T.nsp, isqrt64(q_top));
}

template <int BS, uint32_t CMAX>
__global__ __launch_bounds__(BS, 4) void wheel_kernel(Tables T, MarkParams mp, const WinDesc* __restrict__ wins,
                                                   uint64_t ntiles, uint32_t cls_len,
                                                   uint64_t* __restrict__ tile_counter, SlowList sl,
                                                   StatRec* __restrict__ rec, const WheelRes* __restrict__ pr,
                                                   const uint64_t* __restrict__ tinfo) {
    constexpr uint32_t SW = wh_stream_words(CMAX);
    __shared__ __align__(16) uint32_t wh[2 * SW];   // [S1 | S5] at fixed LDS addresses
    extern __shared__ uint32_t lds[];                               // [stats | tile id | medium first marks]
    LaneStats* red = reinterpret_cast<LaneStats*>(lds);
    uint64_t& s_tile = *reinterpret_cast<uint64_t*>(red + BS / 64);
    uint32_t* med0 = reinterpret_cast<uint32_t*>(&s_tile + 1);
    const uint32_t tid = threadIdx.x;
    const uint32_t nbits = cls_len + WH_HALO + 64;
    const uint32_t nw = cls_len / 64;
    const uint64_t* s1 = reinterpret_cast<const uint64_t*>(wh);
    const uint64_t* s5 = reinterpret_cast<const uint64_t*>(wh + SW);
    for (;;) {
        if (tid == 0) s_tile = atomicAdd((unsigned long long*)tile_counter, 1ull);
        __syncthreads();
        const uint64_t t = s_tile;
        __syncthreads();
        if (t >= ntiles) break;
        const uint64_t ti = tinfo[t];
        const uint32_t lo = (uint32_t)(ti >> 32);
        const WinDesc wd = wins[lo];
        const uint64_t n_lo = wd.start, n_hi = wd.start + 2 * (wd.evens - 1);
        const uint64_t n0 = wheel_tile_n0(wd, t, cls_len);
        const int64_t K0 = (int64_t)(n0 / 6) - (int64_t)WH_HALO;
        sieve_wheel_tile<BS, SW>(wh, med0, K0, nbits, T, tid, pr + wd.r0off, (uint32_t)(t - wd.tile0), (uint32_t)ti);
        LaneStats st = lane_stats_zero();
        uint32_t jlo0, jhi0, jlo2, jhi2, jlo4, jhi4;
        class_range(n0, 0, cls_len, n_lo, n_hi, jlo0, jhi0);
        class_range(n0, 2, cls_len, n_lo, n_hi, jlo2, jhi2);
        class_range(n0, 4, cls_len, n_lo, n_hi, jlo4, jhi4);
        if (mp.K == NMARK && mp.p_small >= HALO_P && !(mp.flags & 1u)) {
            if (jhi0 > jlo0) wheel_mark_class<BS, 0>(s1, s5, n0, nw, jlo0, jhi0, mp, sl, lo, st, tid);
            if (jhi2 > jlo2) wheel_mark_class<BS, 2>(s1, s5, n0 + 2, nw, jlo2, jhi2, mp, sl, lo, st, tid);
            if (jhi4 > jlo4) wheel_mark_class<BS, 4>(s1, s5, n0 + 4, nw, jlo4, jhi4, mp, sl, lo, st, tid);
        } else {
            if (jhi0 > jlo0) wheel_mark_generic<BS>(s1, s5, n0, 0, nw, jlo0, jhi0, mp, T.sp, sl, lo, st, tid);
            if (jhi2 > jlo2) wheel_mark_generic<BS>(s1, s5, n0, 2, nw, jlo2, jhi2, mp, T.sp, sl, lo, st, tid);
            if (jhi4 > jlo4) wheel_mark_generic<BS>(s1, s5, n0, 4, nw, jlo4, jhi4, mp, T.sp, sl, lo, st, tid);
        }
        block_store<BS>(st, (uint64_t)(jhi0 - jlo0) + (jhi2 - jlo2) + (jhi4 - jlo4), &rec[t], tid, red);
    }
}

// pr[wd.r0off + k] for the wheel: e = first bit of S1 (q = 6(kbase + e) + 1) holding a multiple of
// p, and the constants of the tile offset (cls_len mod p and its ratio to p)
__global__ void residue_wheel_kernel(Tables T, const WinDesc* __restrict__ wins, uint32_t cls_len,
                                     WheelRes* __restrict__ pr) {
    const WinDesc wd = wins[blockIdx.y];
    const uint64_t kb = (uint64_t)(wd.kbase + (int64_t)WH_HALO);   // >= 0
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < wd.np; k += gridDim.x * blockDim.x) {
        const uint32_t p = T.sp[k];
        WheelRes e;
        e.p = p;
        e.lp = cls_len % p;
        e.ratio = (float)((double)e.lp / (double)p);
        if (p % 3 == 0) {
            e.e = 0;
        } else {
            const uint32_t km = (mod_small(kb, p, T.inv[k], T.c32[k]) + p - (WH_HALO % p)) % p;   // kbase mod p
            e.e = wheel_first(km, 1, p, inv6_mod(p));
        }
        pr[wd.r0off + k] = e;
    }
}

// ---------------------------------------------------------------------------
// slow path: entries whose p_min exceeds the window's reach (p > sp[K-1])
// ---------------------------------------------------------------------------
__device__ int partition_chunk(const Tables& T, uint64_t n, uint64_t half, uint64_t bound, uint64_t base,
                               uint32_t lane, uint64_t& res);

__global__ void slow_kernel(Tables T, MarkParams mp, SlowList sl, uint64_t* __restrict__ res_p,
                            const WinDesc* __restrict__ wins, uint64_t* __restrict__ verified,
                            uint32_t* __restrict__ pmin_out, uint64_t pmin_n0) {
    const uint32_t cnt = *sl.count < sl.cap ? *sl.count : sl.cap;
    if (mp.flags & MP_SPARSE) {
        // sparse proofs: an n here may have a partner among the window primes whose q was left
        // unproven, so a wave per n scans the candidates from p = 3, 64 at a time, as a single
        // query does (base-2 filter, the other bases on the lowest survivor)
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
        for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < cnt; i += nwaves) {
            const uint64_t n = sl.n[i];
            uint64_t found = 0;
            for (uint64_t base = 0;; base += 64)
                if (partition_chunk(T, n, n / 2, mp.bound, base, lane, found) != 0) break;
            if (lane == 0) {
                res_p[i] = found;
                const WinDesc wd = wins[sl.win[i]];
                if (found && found <= mp.p_small && verified) {
                    const uint64_t bit = wd.word0 * 64 + (n - wd.start) / 2;
                    atomicOr((unsigned long long*)&verified[bit >> 6], 1ull << (bit & 63));
                }
                if (pmin_out) pmin_out[(n - pmin_n0) / 2] = (uint32_t)found;
            }
        }
        return;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint64_t n = sl.n[i];
        const uint64_t half = n / 2;
        uint64_t found = 0;
        uint32_t k = mp.K;
        for (; k < T.nsp; ++k) {
            const uint64_t p = T.sp[k];
            if (p > half || p > mp.bound) break;
            if (is_prime_q(n - p, T)) { found = p; break; }
        }
        if (!found && k == T.nsp)
            for (uint64_t p = (uint64_t)T.sp[T.nsp - 1] + 2; p <= half && p <= mp.bound; p += 2)
                if (is_prime_u64(p) && is_prime_q(n - p, T)) { found = p; break; }
        res_p[i] = found;
        const WinDesc wd = wins[sl.win[i]];
        const uint64_t j = (n - wd.start) / 2;
        if (found && found <= mp.p_small && verified) {
            const uint64_t bit = wd.word0 * 64 + j;
            atomicOr((unsigned long long*)&verified[bit >> 6], 1ull << (bit & 63));
        }
        if (pmin_out) pmin_out[(n - pmin_n0) / 2] = (uint32_t)found;
    }
}

// ---------------------------------------------------------------------------
// reduce: one workgroup per window
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_kernel(MarkParams mp, const StatRec* __restrict__ rec,
                                                      const uint64_t* __restrict__ rec_lo,
                                                      const uint64_t* __restrict__ rec_hi, SlowList sl,
                                                      const uint64_t* __restrict__ res_p, StatRec* __restrict__ out,
                                                      SegClock* __restrict__ clk, const uint64_t* __restrict__ t0,
                                                      uint64_t seq) {
    const uint32_t w = blockIdx.x, tid = threadIdx.x;
    LaneStats s = lane_stats_zero();
    uint64_t evens = 0;
    for (uint64_t r = rec_lo[w] + tid; r < rec_hi[w]; r += blockDim.x) {
        const StatRec& x = rec[r];
        s.ph2 += x.phase2;
        s.fail += x.failures;
        s.first = x.first_fail < s.first ? x.first_fail : s.first;
        s.sum += x.sum_pmin;
        merge_max(s.mx, s.arg, x.max_pmin, x.argmax);
        evens += x.evens;
    }
    const uint32_t cnt = *sl.count < sl.cap ? *sl.count : sl.cap;
    for (uint32_t i = tid; i < cnt; i += blockDim.x) {
        if (sl.win[i] != w) continue;
        const uint64_t n = sl.n[i], p = res_p[i];
        if (p) {
            if (p <= mp.p_small) s.ph2 -= 1;   // counted unverified by the window pass
            if (mp.flags & 1u) {
                s.sum += p;
                merge_max(s.mx, s.arg, p, n);
            }
        } else {
            s.fail += 1;
            s.first = n < s.first ? n : s.first;
        }
    }
    __shared__ uint64_t s_ev[256 / 64];
    wave_reduce(s);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) evens += __shfl_xor(evens, o);
    if ((tid & 63) == 0) s_ev[tid >> 6] = evens;
    __shared__ LaneStats red[256 / 64];
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        LaneStats t = red[0];
        uint64_t ev = s_ev[0];
        for (int i = 1; i < 256 / 64; ++i) {
            t.ph2 += red[i].ph2;
            t.fail += red[i].fail;
            t.first = red[i].first < t.first ? red[i].first : t.first;
            t.sum += red[i].sum;
            merge_max(t.mx, t.arg, red[i].mx, red[i].arg);
            ev += s_ev[i];
        }
        StatRec r;
        r.phase2 = t.ph2;
        r.failures = t.fail;
        r.first_fail = t.first;
        r.sum_pmin = t.sum;
        r.max_pmin = t.mx;
        r.argmax = t.arg;
        r.evens = ev;
        r.pad = *sl.overflow;   // the slow list overflowed: these statistics are incomplete
        out[w] = r;
        if (clk) {
            clk->t0 = *t0;
            clk->t1 = (uint64_t)wall_clock64();
            __threadfence_system();   // the record and the times before seq
            __hip_atomic_store(&clk->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// pr[wins[w].r0off + k] for k < wins[w].np (grid.y = window)
// e = smallest b >= 0 with p | rbase + 2b: b = (-rbase mod p) / 2 when that is
// even, else (-rbase mod p + p) / 2 (rbase odd, so rbase + 2b is the odd multiple).
__global__ void residue_kernel(Tables T, const WinDesc* __restrict__ wins, PrimeRes* __restrict__ pr) {
    const WinDesc wd = wins[blockIdx.y];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < wd.np; k += gridDim.x * blockDim.x) {
        const uint32_t p = T.sp[k];
        const uint32_t r = mod_small(wd.rbase, p, T.inv[k], T.c32[k]);
        const uint32_t t = r ? p - r : 0;
        PrimeRes e;
        e.p = p;
        e.e = (t & 1) ? (t + p) >> 1 : t >> 1;
        e.inv = T.inv[k];
        pr[wd.r0off + k] = e;
    }
}

// Set-up of one plain segment in a single launch: its descriptor, record bounds and cleared
// slow-list control (block 0), and the sieving-prime entries of its window (all blocks), so no
// copy precedes the segment's kernels.
__global__ void segment_prep_kernel(Tables T, WinDesc wd, uint64_t nrec, WinDesc* __restrict__ wins,
                                    uint64_t* __restrict__ rec_lo, uint64_t* __restrict__ rec_hi,
                                    uint64_t* __restrict__ slow_ctl, PrimeRes* __restrict__ pr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        wins[0] = wd;
        rec_lo[0] = 0;
        rec_hi[0] = nrec;
        slow_ctl[0] = 0;
        slow_ctl[1] = 0;
        slow_ctl[2] = (uint64_t)wall_clock64();   // the segment's start, for its kernel time
    }
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < wd.np; k += gridDim.x * blockDim.x) {
        const uint32_t p = T.sp[k];
        const uint32_t r = mod_small(wd.rbase, p, T.inv[k], T.c32[k]);
        const uint32_t t = r ? p - r : 0;
        PrimeRes e;
        e.p = p;
        e.e = (t & 1) ? (t + p) >> 1 : t >> 1;
        e.inv = T.inv[k];
        pr[wd.r0off + k] = e;
    }
}

__global__ void fill_kernel(uint64_t* p, uint64_t words, uint64_t value) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// table builders
// ---------------------------------------------------------------------------
// bits_out[w] bit b <-> odd q = 2(64w+b)+3 <= limit, by trial division (tiny limits only)
__global__ void seed_primes_kernel(uint32_t limit, uint64_t* bits_out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nbits = (limit - 3) / 2 + 1;
    if (w >= (nbits + 63) / 64) return;
    uint64_t word = 0;
    for (uint32_t b = 0; b < 64; ++b) {
        const uint32_t i = 64 * w + b;
        if (i >= nbits) break;
        const uint32_t q = 2 * i + 3;
        bool pr = true;
        for (uint32_t d = 3; d * d <= q; d += 2)
            if (q % d == 0) { pr = false; break; }
        if (pr) word |= 1ull << b;
    }
    bits_out[w] = word;
}

// single workgroup: list of odd q <= qlimit whose bit is set (bit i <-> 2i+3)
__global__ __launch_bounds__(1024) void compact_primes_kernel(const uint64_t* __restrict__ bits, uint32_t nwords,
                                                              uint32_t qlimit, uint32_t* __restrict__ out,
                                                              uint32_t cap, uint32_t* __restrict__ count_out) {
    __shared__ uint32_t cnt[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (nwords + 1023) / 1024;
    const uint32_t a = tid * per, b = a + per < nwords ? a + per : nwords;
    const uint32_t nbits = (qlimit - 3) / 2 + 1;
    uint32_t c = 0;
    for (uint32_t w = a; w < b; ++w) {
        uint64_t x = bits[w];
        if (64ull * (w + 1) > nbits) x &= (64ull * w >= nbits) ? 0 : ((1ull << (nbits - 64 * w)) - 1);
        c += __popcll(x);
    }
    cnt[tid] = c;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t v = tid >= o ? cnt[tid - o] : 0;
        __syncthreads();
        cnt[tid] += v;
        __syncthreads();
    }
    uint32_t pos = cnt[tid] - c;
    for (uint32_t w = a; w < b; ++w) {
        uint64_t x = bits[w];
        if (64ull * (w + 1) > nbits) x &= (64ull * w >= nbits) ? 0 : ((1ull << (nbits - 64 * w)) - 1);
        while (x) {
            const uint32_t bb = __builtin_ctzll(x);
            x &= x - 1;
            if (pos < cap) out[pos] = 2 * (64 * w + bb) + 3;
            ++pos;
        }
    }
    if (tid == 1023) *count_out = cnt[1023];
}

// d5 = (-4 * 6^-1) mod p, closed form: 6^-1 = (5p+1)/6 for p = 1 mod 6 and (p+1)/6
// for p = 5 mod 6, giving (2p-2)/3 and (p-2)/3 (0 for p = 3, which the wheel never sieves).
__global__ void prime_aux_kernel(const uint32_t* __restrict__ sp, uint32_t n, double* __restrict__ inv,
                                 uint32_t* __restrict__ c32, uint32_t* __restrict__ d5) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = sp[k];
    inv[k] = 1.0 / (double)p;
    c32[k] = (uint32_t)((1ull << 32) % p);
    if (d5) d5[k] = p % 3 == 0 ? 0u : (p % 6 == 1 ? (2 * p - 2) / 3 : (p - 2) / 3);
}

// ---------------------------------------------------------------------------
// serving: one wave per query; lanes try 64 consecutive primes, base-2 SPRP
// filters, the lowest surviving lane is confirmed with the remaining bases
// spread over lanes 1 .. mr_bases(q)-1.  The answer is the minimal p (scan order of cpu_goldbach.cpp:19-22).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool small_factor(uint64_t q) {
    return q % 3 == 0 || q % 5 == 0 || q % 7 == 0 || q % 11 == 0 || q % 13 == 0 || q % 17 == 0 ||
           q % 19 == 0 || q % 23 == 0;
}

// One 64-candidate chunk of the minimal-partition scan for even n > 4: lane l
// tries the (base + l)-th odd prime (odd numbers tested by Miller-Rabin past
// the table).  Returns 1 with the minimal p of the chunk in res, -1 when no
// candidate is in range any more (no partition up to bound: res = 0), 0 when
// the chunk holds none and the scan goes on.  Wave-uniform.
__device__ int partition_chunk(const Tables& T, uint64_t n, uint64_t half, uint64_t bound, uint64_t base,
                               uint32_t lane, uint64_t& res) {
    const uint64_t k = base + lane;
    uint64_t p;
    bool isp;
    if (k < T.nsp) {
        p = T.sp[k];
        isp = true;
    } else {
        p = (uint64_t)T.sp[T.nsp - 1] + 2 * (k - T.nsp + 1);
        isp = p <= half && p <= bound && is_prime_u64(p);
    }
    const bool in = p <= half && p <= bound;
    if (__ballot(in) == 0) {
        res = 0;
        return -1;
    }
    const uint64_t q = n - p;
    bool pass = false;
    if (in && isp) {
        if (q <= T.p2_limit) {
            pass = is_prime_q(q, T);
        } else if (q < 41) {
            // a Phase-2 bound below 41 leaves small q here: small_factor would reject the
            // primes 3..23 themselves and the strong tests need q > 37
            pass = (0x20A08A28ACull >> q) & 1;
        } else if (!small_factor(q)) {
            pass = mr_test(q, 0, false);
        }
    }
    uint64_t ball = __ballot(pass);
    while (ball) {
        const uint32_t L = __ffsll((unsigned long long)ball) - 1;
        const uint64_t qq = __shfl(q, L);
        bool ok = true;
        if (qq > T.p2_limit && qq >= 41 && lane >= 1 && lane < mr_bases(qq)) ok = mr_test(qq, lane, mr_wide(qq));
        if (__all(ok)) {
            res = __shfl(p, L);
            return 1;
        }
        ball &= ball - 1;
    }
    return 0;
}

// partition_chunk with the strong-test bases spread over the workgroup's SERVE_BASES waves:
// wave w tests every candidate of the chunk to base w (those q whose base count is below w+1
// pass it), and a candidate is prime when it passes in every wave, so the chunk's answer takes
// the time of one strong test instead of base 2 followed by the others.  pass[SERVE_BASES] is
// the workgroup's LDS.  Returns as partition_chunk; uniform over the workgroup.
__device__ int partition_chunk_wide(const Tables& T, uint64_t n, uint64_t half, uint64_t bound, uint64_t base,
                                    uint32_t lane, uint32_t w, uint64_t* pass, uint64_t& res) {
    const uint64_t k = base + lane;
    uint64_t p;
    bool isp;
    if (k < T.nsp) {
        p = T.sp[k];
        isp = true;
    } else {
        p = (uint64_t)T.sp[T.nsp - 1] + 2 * (k - T.nsp + 1);
        isp = p <= half && p <= bound && is_prime_u64(p);
    }
    const bool in = p <= half && p <= bound;
    if (__ballot(in) == 0) {   // the same p in every wave: uniform over the workgroup
        res = 0;
        return -1;
    }
    const uint64_t q = n - p;
    bool ok = false;
    if (in && isp) {
        if (q <= T.p2_limit) {
            ok = is_prime_q(q, T);
        } else if (q < 41) {
            ok = (0x20A08A28ACull >> q) & 1;   // as in partition_chunk
        } else if (!small_factor(q)) {
            ok = w < mr_bases(q) ? mr_test(q, w, mr_wide(q)) : true;
        }
    }
    const uint64_t m = __ballot(ok);
    if (lane == 0) pass[w] = m;
    __syncthreads();
    uint64_t all = ~0ull;
#pragma unroll
    for (uint32_t v = 0; v < SERVE_BASES; ++v) all &= pass[v];
    __syncthreads();   // pass[] is written again by the next chunk
    if (all) {
        res = __shfl(p, __ffsll((unsigned long long)all) - 1);
        return 1;
    }
    return 0;
}

__global__ __launch_bounds__(64 * SERVE_WAVES) void min_partition_kernel(Tables T, const uint64_t* __restrict__ ns,
                                                                         uint32_t count, uint64_t bound,
                                                                         uint64_t* __restrict__ p_out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t qi = blockIdx.x * SERVE_WAVES + (threadIdx.x >> 6);
    if (qi >= count) return;   // uniform per wave
    const uint64_t n = ns[qi];
    uint64_t res = 0;
    if (n >= 4 && !(n & 1)) {
        if (n == 4) {
            res = bound >= 2 ? 2 : 0;
        } else {
            for (uint64_t base = 0;; base += 64)
                if (partition_chunk(T, n, n / 2, bound, base, lane, res) != 0) break;
        }
    }
    if (lane == 0) p_out[qi] = res;
}

// Continuous-batching step over slots [0, count): a slot whose adm entry is nonzero takes that
// n as a new query (scan from the first candidate); otherwise it resumes the query its
// device-resident state holds (n = 0: the slot is free).  One wave per slot advances its query
// by up to `chunks` chunks; out[slot] (pinned host memory) gets the answer and whether it is
// final, and the state keeps the position of a query that goes on (or frees the slot).
// A serving step's completion: each workgroup counts itself once its answers are out (its
// writer fenced them at system scope first); the last one resets the count for the lane's next
// step and publishes seq to the host, which waits on that pinned word instead of an event.  For
// the wide step (at most SERVE_WIDE_MAX workgroups, one fence each).
__device__ __forceinline__ void step_signal(const StepSignal& sg) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = atomicAdd(sg.blocks_done, 1u);
        if (prev == gridDim.x - 1) {
            atomicExch(sg.blocks_done, 0u);
            __threadfence_system();
            __hip_atomic_store(sg.seq_out, sg.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// The narrow step's completion, launched behind it on the lane's stream: a system-scope fence per
// wave of a 16,384-slot step costs more than the step, so the step's answers are ordered before
// seq by the kernel boundary instead.
__global__ void serve_signal_kernel(StepSignal sg) {
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(sg.seq_out, sg.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(64 * SERVE_WAVES) void serve_step_kernel(Tables T, const uint64_t* __restrict__ adm,
                                                                      ServeSlot* __restrict__ slots, uint32_t count,
                                                                      uint64_t bound, uint32_t chunks,
                                                                      ServeDone* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t qi = blockIdx.x * SERVE_WAVES + (threadIdx.x >> 6);
    if (qi < count) {   // uniform per wave
        const uint64_t a = adm[qi];
        uint64_t n, base;
        if (a) {
            n = a;
            base = 0;
        } else {
            const ServeSlot st = slots[qi];
            n = st.n;
            base = st.k;
        }
        uint64_t res = 0;
        uint32_t state = 1;
        if (n == 0) {
            state = 2;   // a free slot
        } else if (n < 4 || (n & 1)) {
            res = 0;
        } else if (n == 4) {
            res = bound >= 2 ? 2 : 0;
        } else {
            state = 0;
            for (uint32_t c = 0; c < chunks; ++c, base += 64) {
                if (partition_chunk(T, n, n / 2, bound, base, lane, res) != 0) {
                    state = 1;
                    break;
                }
            }
        }
        if (lane == 0) {
            out[qi] = ServeDone{res, state, 0};
            if (state != 2) slots[qi] = state ? ServeSlot{0, 0} : ServeSlot{n, base};
        }
    }
}

// serve_step_kernel for steps with few occupied slots: one workgroup of SERVE_BASES waves per
// slot (partition_chunk_wide), so a query's step takes one strong test's time, not two.
__global__ __launch_bounds__(64 * SERVE_BASES) void serve_step_wide_kernel(Tables T, const uint64_t* __restrict__ adm,
                                                                           ServeSlot* __restrict__ slots,
                                                                           uint64_t bound, uint32_t chunks,
                                                                           ServeDone* __restrict__ out,
                                                                           StepSignal sg) {
    __shared__ uint64_t pass[SERVE_BASES];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t qi = blockIdx.x;
    const uint64_t a = adm[qi];
    uint64_t n, base;
    if (a) {
        n = a;
        base = 0;
    } else {
        const ServeSlot st = slots[qi];
        n = st.n;
        base = st.k;
    }
    uint64_t res = 0;
    uint32_t state = 1;
    if (n == 0) {   // a free slot (uniform over the workgroup)
        state = 2;
    } else if (n < 4 || (n & 1)) {
        res = 0;
    } else if (n == 4) {
        res = bound >= 2 ? 2 : 0;
    } else {
        state = 0;
        for (uint32_t c = 0; c < chunks; ++c, base += 64) {
            if (partition_chunk_wide(T, n, n / 2, bound, base, lane, w, pass, res) != 0) {
                state = 1;
                break;
            }
        }
    }
    if (threadIdx.x == 0) {
        out[qi] = ServeDone{res, state, 0};
        if (state != 2) slots[qi] = state ? ServeSlot{0, 0} : ServeSlot{n, base};
        __threadfence_system();
    }
    step_signal(sg);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename K>
static hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

hipError_t launch_sieve_window(hipStream_t s, const Tables& T, uint64_t qv0, uint32_t lead, uint64_t nbits,
                               uint32_t tile_bits, uint64_t* out, int tail_ones, const PrimeRes* pr) {
    if (nbits == 0) return hipSuccess;
    const uint64_t ntiles = (nbits + tile_bits - 1) / tile_bits;
    if (ntiles > 0x7FFFFFFFull || tile_bits > SIEVE_TILE_MAX || (tile_bits & 63)) return hipErrorInvalidValue;
    const uint32_t med = T.n_med > NPAT ? T.n_med - NPAT : 1;
    const size_t lds = (size_t)(SIEVE_BS / 64 + (med > SIEVE_MRQ ? med : SIEVE_MRQ)) * 4;
    // the largest q of the window: bit nbits - 1 (bits below `lead` stand for no q)
    const uint64_t span = nbits - 1 >= lead ? nbits - 1 - lead : 0;
    const uint64_t c1 = (uint64_t)T.cover + 1;
    const bool mr = span > (~0ull - qv0) / 2 || qv0 + 2 * span >= c1 * c1;
    hipError_t e = mr ? allow_lds(sieve_window_kernel<SIEVE_BS, true>, lds) : allow_lds(sieve_window_kernel<SIEVE_BS, false>, lds);
    if (e != hipSuccess) return e;
    if (mr)
        hipLaunchKernelGGL((sieve_window_kernel<SIEVE_BS, true>), dim3((uint32_t)ntiles), dim3(SIEVE_BS), lds, s, T, qv0,
                           lead, nbits, tile_bits, out, tail_ones, pr);
    else
        hipLaunchKernelGGL((sieve_window_kernel<SIEVE_BS, false>), dim3((uint32_t)ntiles), dim3(SIEVE_BS), lds, s, T, qv0,
                           lead, nbits, tile_bits, out, tail_ones, pr);
    return hipGetLastError();
}

hipError_t launch_mark_segment(hipStream_t s, const Tables& T, const MarkParams& mp, uint64_t n0, uint64_t cnt,
                               const uint64_t* window, uint64_t window_words, uint32_t wpb, uint64_t* verified,
                               uint32_t* pmin_out, SlowList sl, StatRec* rec, uint32_t* nrec_out) {
    const uint64_t nw = (cnt + 63) / 64;
    const uint64_t blocks = (nw + wpb - 1) / wpb;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const size_t lds = (size_t)(wpb + HALO / 64 + 1) * 8 + sizeof(LaneStats) * (MARK_BS / 64);
    hipError_t e = allow_lds(mark_segment_kernel<MARK_BS>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mark_segment_kernel<MARK_BS>, dim3((uint32_t)blocks), dim3(MARK_BS), lds, s, T, mp, n0, cnt,
                       window, window_words, wpb, verified, pmin_out, sl, rec);
    if (nrec_out) *nrec_out = (uint32_t)blocks;
    return hipGetLastError();
}

hipError_t launch_residues(hipStream_t s, const Tables& T, const WinDesc* wins, uint32_t nwin, uint32_t max_np,
                           PrimeRes* pr) {
    if (!nwin || !max_np) return hipSuccess;
    uint32_t bx = (max_np + 255) / 256;
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(residue_kernel, dim3(bx, nwin), dim3(256), 0, s, T, wins, pr);
    return hipGetLastError();
}

static size_t fused_dyn_lds(uint32_t n_med) {
    const uint32_t med = n_med > NPAT ? n_med - NPAT : 1;
    return sizeof(LaneStats) * (FUSED_BS / 64) + 8 +
           (size_t)(FUSED_BS / 64 + 2 + (HALO_P + 63) / 64 + (med > FUSED_MRQ ? med : FUSED_MRQ)) * 4;
}

size_t fused_lds_bytes(uint32_t tile_evens, uint32_t n_med) {
    (void)tile_evens;   // the window is sized for the largest tile
    return (size_t)FUSED_WIN_WORDS * 4 + fused_dyn_lds(n_med);
}

hipError_t launch_fused(hipStream_t s, const Tables& T, const MarkParams& mp, const WinDesc* wins, uint32_t nwin,
                        uint64_t ntiles, uint32_t tile_evens, uint32_t grid, uint64_t* tile_counter,
                        SlowList sl, StatRec* rec, const PrimeRes* pr, uint64_t* wcopy) {
    if (tile_evens > FUSED_TILE_MAX || (tile_evens & 63)) return hipErrorInvalidValue;
    if ((mp.flags & MP_SPARSE) && (!wcopy || mp.proof_stride < 2)) return hipErrorInvalidValue;
    const size_t lds = fused_dyn_lds(T.n_med);
    hipError_t e = allow_lds(fused_kernel<FUSED_BS>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fused_kernel<FUSED_BS>, dim3(grid), dim3(FUSED_BS), lds, s, T, mp, wins, nwin, ntiles,
                       tile_evens, tile_counter, sl, rec, pr, wcopy);
    return hipGetLastError();
}

size_t wheel_lds_bytes(uint32_t n_med, uint32_t wg_per_cu) {
    const uint32_t med = n_med > NPAT ? n_med - NPAT : 1;
    if (wg_per_cu >= 2)
        return (size_t)2 * wh_stream_words(WH_CLASS_HALF) * 4 + sizeof(LaneStats) * (FUSED_BS / 128) + 8 +
               (size_t)2 * med * 4;
    return (size_t)2 * wh_stream_words(WH_CLASS_MAX) * 4 + sizeof(LaneStats) * (FUSED_BS / 64) + 8 +
           (size_t)2 * med * 4;
}

hipError_t launch_residues_wheel(hipStream_t s, const Tables& T, const WinDesc* wins, uint32_t nwin,
                                 uint32_t max_np, uint32_t cls_len, PrimeRes* pr) {
    if (!nwin || !max_np) return hipSuccess;
    uint32_t bx = (max_np + 255) / 256;
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(residue_wheel_kernel, dim3(bx, nwin), dim3(256), 0, s, T, wins, cls_len,
                       reinterpret_cast<WheelRes*>(pr));
    return hipGetLastError();
}

hipError_t launch_wheel(hipStream_t s, const Tables& T, const MarkParams& mp, const WinDesc* wins, uint32_t nwin,
                        uint64_t ntiles, uint32_t cls_len, uint32_t grid, uint32_t wg_per_cu, uint64_t* tile_counter,
                        SlowList sl, StatRec* rec, const PrimeRes* pr, uint64_t* tinfo) {
    if (cls_len > WH_CLASS_MAX || cls_len < 64 || (cls_len & 63)) return hipErrorInvalidValue;
    if (!ntiles) return hipSuccess;
    hipLaunchKernelGGL(wheel_tiles_kernel, dim3((uint32_t)((ntiles + 255) / 256)), dim3(256), 0, s, T, wins, nwin,
                       ntiles, cls_len, tinfo);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const uint32_t med = T.n_med > NPAT ? T.n_med - NPAT : 1;
    if (wg_per_cu >= 2) {   // two half-size workgroups per CU: one's LDS phase can overlap the other's VALU phase
        if (cls_len > WH_CLASS_HALF) return hipErrorInvalidValue;
        constexpr int BS = FUSED_BS / 2;
        const size_t lds = sizeof(LaneStats) * (BS / 64) + 8 + (size_t)2 * med * 4;
        hipError_t e = allow_lds(wheel_kernel<BS, WH_CLASS_HALF>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((wheel_kernel<BS, WH_CLASS_HALF>), dim3(grid), dim3(BS), lds, s, T, mp, wins, ntiles,
                           cls_len, tile_counter, sl, rec, reinterpret_cast<const WheelRes*>(pr), tinfo);
        return hipGetLastError();
    }
    const size_t lds = sizeof(LaneStats) * (FUSED_BS / 64) + 8 + (size_t)2 * med * 4;
    hipError_t e = allow_lds(wheel_kernel<FUSED_BS, WH_CLASS_MAX>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((wheel_kernel<FUSED_BS, WH_CLASS_MAX>), dim3(grid), dim3(FUSED_BS), lds, s, T, mp, wins,
                       ntiles, cls_len, tile_counter, sl, rec, reinterpret_cast<const WheelRes*>(pr), tinfo);
    return hipGetLastError();
}

hipError_t launch_slow(hipStream_t s, const Tables& T, const MarkParams& mp, SlowList sl, uint64_t* res_p,
                       const WinDesc* wins, uint64_t* verified, uint32_t* pmin_out, uint64_t pmin_n0) {
    // a wave per n with sparse proofs: more waves, so the strong tests' latency is hidden
    const uint32_t blocks = (mp.flags & MP_SPARSE) ? 2048u : 256u;
    hipLaunchKernelGGL(slow_kernel, dim3(blocks), dim3(256), 0, s, T, mp, sl, res_p, wins, verified, pmin_out, pmin_n0);
    return hipGetLastError();
}

hipError_t launch_reduce(hipStream_t s, const MarkParams& mp, const StatRec* rec, const uint64_t* rec_lo,
                         const uint64_t* rec_hi, uint32_t nwin, SlowList sl, const uint64_t* res_p, StatRec* out,
                         SegClock* clk, const uint64_t* t0, uint64_t seq) {
    if (clk && (nwin != 1 || !t0)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reduce_kernel, dim3(nwin), dim3(256), 0, s, mp, rec, rec_lo, rec_hi, sl, res_p, out, clk, t0,
                       seq);
    return hipGetLastError();
}

hipError_t launch_segment_prep(hipStream_t s, const Tables& T, const WinDesc& wd, uint64_t nrec, WinDesc* wins,
                               uint64_t* rec_lo, uint64_t* rec_hi, uint64_t* slow_ctl, PrimeRes* pr) {
    uint32_t bx = (wd.np + 255) / 256;
    if (bx < 1) bx = 1;
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(segment_prep_kernel, dim3(bx), dim3(256), 0, s, T, wd, nrec, wins, rec_lo, rec_hi, slow_ctl, pr);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, uint64_t* p, uint64_t words, uint64_t value) {
    if (!words) return hipSuccess;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, words, value);
    return hipGetLastError();
}

hipError_t launch_seed_primes(hipStream_t s, uint32_t limit, uint64_t* bits_out) {
    const uint32_t nbits = (limit - 3) / 2 + 1;
    const uint32_t nw = (nbits + 63) / 64;
    hipLaunchKernelGGL(seed_primes_kernel, dim3((nw + 63) / 64), dim3(64), 0, s, limit, bits_out);
    return hipGetLastError();
}

hipError_t launch_compact_primes(hipStream_t s, const uint64_t* bits, uint32_t nwords, uint32_t qlimit,
                                 uint32_t* primes_out, uint32_t cap, uint32_t* count_out) {
    hipLaunchKernelGGL(compact_primes_kernel, dim3(1), dim3(1024), 0, s, bits, nwords, qlimit, primes_out, cap,
                       count_out);
    return hipGetLastError();
}

hipError_t launch_prime_aux(hipStream_t s, const uint32_t* sp, uint32_t n, double* inv, uint32_t* c32, uint32_t* d5) {
    hipLaunchKernelGGL(prime_aux_kernel, dim3((n + 255) / 256), dim3(256), 0, s, sp, n, inv, c32, d5);
    return hipGetLastError();
}

hipError_t launch_serve_step(hipStream_t s, const Tables& T, const uint64_t* adm, ServeSlot* slots, uint32_t count,
                             uint64_t bound, uint32_t chunks, ServeDone* out, StepSignal sg) {
    if (!count || !sg.blocks_done || !sg.seq_out) return hipErrorInvalidValue;   // a step must signal
    if (count <= SERVE_WIDE_MAX) {
        hipLaunchKernelGGL(serve_step_wide_kernel, dim3(count), dim3(64 * SERVE_BASES), 0, s, T, adm, slots, bound,
                           chunks, out, sg);
        return hipGetLastError();
    }
    const uint32_t blocks = (count + SERVE_WAVES - 1) / SERVE_WAVES;
    hipLaunchKernelGGL(serve_step_kernel, dim3(blocks), dim3(64 * SERVE_WAVES), 0, s, T, adm, slots, count, bound,
                       chunks, out);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    hipLaunchKernelGGL(serve_signal_kernel, dim3(1), dim3(64), 0, s, sg);
    return hipGetLastError();
}

hipError_t launch_min_partition(hipStream_t s, const Tables& T, const uint64_t* ns, uint32_t count, uint64_t bound,
                                uint64_t* p_out) {
    if (!count) return hipSuccess;
    const uint32_t blocks = (count + SERVE_WAVES - 1) / SERVE_WAVES;
    hipLaunchKernelGGL(min_partition_kernel, dim3(blocks), dim3(64 * SERVE_WAVES), 0, s, T, ns, count, bound, p_out);
    return hipGetLastError();
}

// ===========================================================================
// Goldbach partition counts c(n) = #{p <= n/2 : p and n - p prime} for a
// range of even n (cpu_goldbach.cpp:29-36, goldbach_count, restated for a
// range).  With P the odd prime bitset (bit a <-> 2a + 3) and K = (n - 6)/2,
// c(n) = [n = 4] + #{a <= K/2 : P[a] and P[K - a]}: an anti-diagonal of P
// against itself.  R[j] = P[M - j] (M the last index) turns it into
// popcount(P[64w..] & R[64w + M - K ..]) over words w <= K/128.
//
// count_kernel: workgroup (a, r) takes the 1024 n of index i = 65536 a + r + 64 t
// (thread t, r < 64), so K = (n - 6)/2 is the same mod 64 for all of them and R is
// shifted by one amount for the whole workgroup: per chunk of CP_CHUNK words of P,
// the matching R words are staged in LDS already shifted (two v_alignbit_b32 per
// word, once per workgroup).  Per thread and word: two LDS reads (P broadcast, R
// at the thread's own offset, consecutive across lanes), an AND and a popcount.
// ===========================================================================
namespace {
constexpr uint32_t CP_CHUNK = 2048;
}

// R[w] = bit-reverse of P over indices [0, M]: bit j of R is P[M - j]; words past M are zero.
__global__ void reverse_bits_kernel(const uint64_t* __restrict__ P, uint64_t M, uint64_t nwords,
                                    uint64_t* __restrict__ R) {
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t out = 0;
        const uint64_t j0 = 64 * w;
        if (j0 <= M) {
            // P bits [M - j0 - 63, M - j0], read as a 64-bit window ending at M - j0
            const int64_t hi_bit = (int64_t)(M - j0);          // P index of R bit j0
            const int64_t lo_bit = hi_bit - 63;                // may be negative
            uint64_t win;
            if (lo_bit >= 0) {
                const uint64_t q = (uint64_t)lo_bit >> 6, sh = (uint64_t)lo_bit & 63;
                win = funnel64(P[q], sh ? P[q + 1] : 0, (uint32_t)sh);
            } else {
                win = P[0] << (uint32_t)(-lo_bit);             // bits below index 0 are absent
            }
            // win bit t <-> P[lo_bit + t]; R bit (j0 + u) <-> P[hi_bit - u] = win bit (63 - u)
            const uint32_t lo = (uint32_t)win, hi = (uint32_t)(win >> 32);
            out = ((uint64_t)__builtin_bitreverse32(lo) << 32) | __builtin_bitreverse32(hi);
            if (M - j0 < 63) out &= (2ull << (M - j0)) - 1;    // bits j > M are zero
        }
        R[w] = out;
    }
}

// Narrow ranges (fewer than 65536 n): a workgroup takes 1024 consecutive even n in a
// block of 8192 (wave v, lane l: n = base + 2(16g + v) + 128l); the n of a wave are 64
// apart in K, so the wave shares one shift of R, applied per word in the loop.
__global__ __launch_bounds__(1024) void count_kernel_narrow(const uint64_t* __restrict__ P, const uint64_t* __restrict__ R,
                                                     uint64_t M, uint64_t pwords, uint64_t start, uint64_t n_end,
                                                     uint64_t nchunks, uint64_t* __restrict__ counts) {
    __shared__ uint64_t sP[CP_CHUNK];
    __shared__ uint64_t sR[CP_CHUNK + 66];
    const uint64_t b = blockIdx.x >> 2;
    const uint32_t g = blockIdx.x & 3;
    const uint32_t tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const uint64_t nb = start + 8192 * b;
    const uint64_t n = nb + 2 * (16 * g + wv) + 128 * (uint64_t)lane;
    const bool valid = n <= n_end && n >= 6;
    const uint64_t K = valid ? (n - 6) / 2 : 0;
    const uint64_t amax = K / 2;                      // a <= K/2: p <= n/2
    const uint64_t wl = amax >> 6;                    // this lane's last word
    const uint64_t last_mask = (amax & 63) == 63 ? ~0ull : ((2ull << (amax & 63)) - 1);
    // the workgroup's n span, n_lo..n_hi (K up to K_hi)
    const uint64_t n_lo = nb + 2 * (16 * g) > 6 ? nb + 2 * (16 * g) : 6;
    const uint64_t n_hi_raw = nb + 2 * (16 * g + 15) + 128 * 63;
    const uint64_t n_hi = n_hi_raw < n_end ? n_hi_raw : n_end;
    if (n_hi < n_lo) return;                          // uniform: nothing to count here
    const uint64_t K_hi = (n_hi - 6) / 2;
    const uint64_t s_min = M - K_hi, w_hi = (K_hi / 2) >> 6;
    const uint64_t s = M - K;
    // lanes of a wave are 64 apart in K, so s mod 64 is the wave's (taken from lane 0's
    // nominal n, which may be 4: K = -1 there, still right mod 64)
    const int64_t n_w = (int64_t)(nb + 2 * (16 * g + wv));
    const uint32_t sh = __builtin_amdgcn_readfirstlane((uint32_t)(((int64_t)M - (n_w - 6) / 2) & 63));
    const uint32_t ro = (uint32_t)((s >> 6) - (s_min >> 6));
    uint32_t acc = 0;                                 // c(n) < 2^32 for n <= 2^36
    for (uint64_t chunk = blockIdx.y; chunk < nchunks; chunk += gridDim.y) {
        const uint64_t w0 = chunk * CP_CHUNK;
        if (w0 > w_hi) break;                         // uniform
        const uint64_t rb = w0 + (s_min >> 6);
        for (uint32_t i = tid; i < CP_CHUNK; i += 1024) sP[i] = w0 + i < pwords ? P[w0 + i] : 0;
        for (uint32_t i = tid; i < CP_CHUNK + 66; i += 1024) sR[i] = rb + i < pwords ? R[rb + i] : 0;
        __syncthreads();
        if (valid && wl >= w0) {
            // full words first, then the lane's last word (a <= amax) under its mask
            // R read as 32-bit words so the wave-uniform shift stays below 32: the lane's
            // R bits for word i start at bit 64(ro + i) + sh of the staged R
            const bool ends_here = wl - w0 < CP_CHUNK;
            const uint32_t full = ends_here ? (uint32_t)(wl - w0) : CP_CHUNK;
            const uint32_t* sR32 = reinterpret_cast<const uint32_t*>(sR);
            const uint32_t bi = 2 * ro + (sh >> 5), s5 = sh & 31;
            uint32_t x0 = sR32[bi];
            uint32_t i = 0;
            for (; i < full; ++i) {
                const uint32_t x1 = sR32[bi + 2 * i + 1], x2 = sR32[bi + 2 * i + 2];
                const uint64_t pw = sP[i];
                acc += __builtin_popcount(__builtin_amdgcn_alignbit(x1, x0, s5) & (uint32_t)pw) +
                       __builtin_popcount(__builtin_amdgcn_alignbit(x2, x1, s5) & (uint32_t)(pw >> 32));
                x0 = x2;
            }
            if (ends_here) {
                const uint32_t x1 = sR32[bi + 2 * i + 1], x2 = sR32[bi + 2 * i + 2];
                const uint64_t pw = sP[i] & last_mask;
                acc += __builtin_popcount(__builtin_amdgcn_alignbit(x1, x0, s5) & (uint32_t)pw) +
                       __builtin_popcount(__builtin_amdgcn_alignbit(x2, x1, s5) & (uint32_t)(pw >> 32));
            }
        }
        __syncthreads();
    }
    if (valid && acc) atomicAdd((unsigned long long*)&counts[(n - start) / 2], (unsigned long long)acc);
}

// acc += popcount(x) in one v_bcnt_u32_b32 (the compiler pairs two counts into a v_add3_u32 instead)
__device__ __forceinline__ void bcnt_acc(uint32_t& acc, uint32_t x) {
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x));
}

__global__ __launch_bounds__(1024) void count_kernel(const uint64_t* __restrict__ P, const uint64_t* __restrict__ R,
                                                     uint64_t M, uint64_t pwords, uint64_t K0, uint64_t count,
                                                     uint64_t nchunks, uint64_t* __restrict__ counts) {
    __shared__ uint64_t sP[CP_CHUNK];
    __shared__ uint64_t sR[CP_CHUNK + 1026];
    const uint32_t tid = threadIdx.x;
    const uint64_t r = blockIdx.x & 63, a = blockIdx.x >> 6;
    const uint64_t i_first = 65536 * a + r;                // this workgroup's n: i = i_first + 64 t
    if (i_first >= count) return;                          // uniform
    const uint64_t i = i_first + 64ull * tid;
    const uint64_t i_last = i_first + 64ull * ((count - 1 - i_first) / 64 < 1023 ? (count - 1 - i_first) / 64 : 1023);
    if ((int64_t)K0 + (int64_t)i_last < 0) return;        // uniform: only n = 4 here (the host adds it)
    const bool valid = i < count;
    // K = K0 + i with K0 = (start - 6)/2 (start = 4: K0 = -1, and i = 0 is n = 4, counted on the host)
    const int64_t Ks = (int64_t)K0 + (int64_t)i;
    const bool live = valid && Ks >= 0;
    const uint64_t K = live ? (uint64_t)Ks : 0;
    const uint64_t K_hi = (uint64_t)((int64_t)K0 + (int64_t)i_last);
    const uint64_t s_min = M - K_hi, w_hi = (K_hi / 2) >> 6;
    const uint32_t sh = (uint32_t)((M - K_hi) & 63);        // = (M - K) mod 64 for every thread here
    const uint64_t s = M - K;
    const uint32_t ro = live ? (uint32_t)((s >> 6) - (s_min >> 6)) : 0;   // <= 1024
    // the largest ro here: the workgroup's first live n (K >= 0) has the smallest K
    const int64_t K_first = (int64_t)K0 + (int64_t)i_first >= 0 ? (int64_t)K0 + (int64_t)i_first : (int64_t)K0 + (int64_t)i_first + 64;
    const uint32_t ro_max = (uint32_t)(((M - (uint64_t)K_first) >> 6) - (s_min >> 6));
    const uint64_t amax = K / 2;
    const uint64_t wl = amax >> 6;
    const uint64_t last_mask = (amax & 63) == 63 ? ~0ull : ((2ull << (amax & 63)) - 1);
    uint32_t acc = 0;
    for (uint64_t chunk = blockIdx.y; chunk < nchunks; chunk += gridDim.y) {
        const uint64_t w0 = chunk * CP_CHUNK;
        if (w0 > w_hi) break;                              // uniform
        const uint64_t rb = w0 + (s_min >> 6);
        for (uint32_t j = tid; j < CP_CHUNK; j += 1024) sP[j] = w0 + j < pwords ? P[w0 + j] : 0;
        for (uint32_t j = tid; j < CP_CHUNK + ro_max + 2; j += 1024) {
            const uint64_t x0 = rb + j < pwords ? R[rb + j] : 0, x1 = rb + j + 1 < pwords ? R[rb + j + 1] : 0;
            sR[j] = funnel64(x0, x1, sh);
        }
        __syncthreads();
        if (live && wl >= w0) {
            const bool ends_here = wl - w0 < CP_CHUNK;
            const uint32_t full = ends_here ? (uint32_t)(wl - w0) : CP_CHUNK;
            const uint64_t* rr = sR + ro;
            uint32_t j = 0;
            for (; j + 4 <= full; j += 4) {                 // unrolled by hand: asm blocks unrolling
                const uint64_t v0 = rr[j] & sP[j], v1 = rr[j + 1] & sP[j + 1];
                const uint64_t v2 = rr[j + 2] & sP[j + 2], v3 = rr[j + 3] & sP[j + 3];
                bcnt_acc(acc, (uint32_t)v0);
                bcnt_acc(acc, (uint32_t)(v0 >> 32));
                bcnt_acc(acc, (uint32_t)v1);
                bcnt_acc(acc, (uint32_t)(v1 >> 32));
                bcnt_acc(acc, (uint32_t)v2);
                bcnt_acc(acc, (uint32_t)(v2 >> 32));
                bcnt_acc(acc, (uint32_t)v3);
                bcnt_acc(acc, (uint32_t)(v3 >> 32));
            }
            for (; j < full; ++j) {
                const uint64_t v = rr[j] & sP[j];
                bcnt_acc(acc, (uint32_t)v);
                bcnt_acc(acc, (uint32_t)(v >> 32));
            }
            if (ends_here) {
                const uint64_t v = rr[j] & sP[j] & last_mask;
                acc += __builtin_popcount((uint32_t)v) + __builtin_popcount((uint32_t)(v >> 32));
            }
        }
        __syncthreads();
    }
    if (live && acc) atomicAdd((unsigned long long*)&counts[i], (unsigned long long)acc);
}

hipError_t launch_count_partitions(hipStream_t s, const uint64_t* P, uint64_t* R, uint64_t M, uint64_t pwords,
                                   uint64_t start, uint64_t n_end, uint64_t* counts) {
    uint64_t blocks = (pwords + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(reverse_bits_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, P, M, pwords, R);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t count = (n_end - start) / 2 + 1;
    if (count < 65536) {
        const uint64_t nb = (count + 4095) / 4096;              // blocks of 8192 in n
        const uint64_t w_top = (((n_end >= 6 ? n_end - 6 : 0) / 2) / 2 >> 6) + 1;
        const uint64_t nchunks = (w_top + CP_CHUNK - 1) / CP_CHUNK;
        const uint32_t gy = (uint32_t)(nchunks < 65535 ? nchunks : 65535);
        hipLaunchKernelGGL(count_kernel_narrow, dim3((uint32_t)(nb * 4), gy), dim3(1024), 0, s, P, R, M, pwords, start,
                           n_end, nchunks, counts);
        return hipGetLastError();
    }
    const uint64_t K0 = (start - 6) / 2;                   // wraps to 2^64 - 1 for start = 4 (K = -1)
    const uint64_t nb = (count + 65535) / 65536 * 64;       // 64 residues of i mod 64 per 65536 n
    const uint64_t w_top = ((n_end - 6) / 2 / 2 >> 6) + 1;
    const uint64_t nchunks = (w_top + CP_CHUNK - 1) / CP_CHUNK;
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const uint32_t gy = (uint32_t)(nchunks < 65535 ? nchunks : 65535);
    hipLaunchKernelGGL(count_kernel, dim3((uint32_t)nb, gy), dim3(1024), 0, s, P, R, M, pwords,
                       start == 4 ? ~0ull : K0, count, nchunks, counts);
    return hipGetLastError();
}

}  // namespace gbk
