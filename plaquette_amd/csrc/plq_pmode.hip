// plq_pmode.hip -- the reference CPU path's own random stream on the device.
//
// The CPU LocalUpdate draws from one sequential stream, RandRng = rand_pcg's
// Lcg128Xsl64 seeded by rand_core's seed_from_u64 (rng.rs:30-50), in site
// order: colour 0's sites in lattice order, then colour 1's
// (updater.rs:316-324).  A Metropolis step draws one f64 only when the move
// is uphill (updater.rs:77-96); a heat-bath step always draws one
// (updater.rs:109-167).  Within a colour pass every site's energy change is
// fixed (its neighbours have the other colour), so which sites draw is known
// before any draw is made, and the k-th drawing site of the pass uses the
// stream's (base + k + 1)-th state.  Per colour pass:
//   pm_count   draws per block of 4,096 sites (16 per lane, in array order,
//              which is lattice order restricted to the colour);
//   pm_scan    exclusive prefix over the blocks, and the pass total;
//   pm_update  each lane jumps the LCG to its first draw (a = multiplier,
//              c = increment: s_{n+k} = A_k s_n + C_k, from a host table of
//              (A, C) for k = 2^b, at most 64 multiply-adds) and steps
//              sequentially through its 16 sites;
//   pm_advance the base state moves past the pass's draws.
// A draw is x = XSL-RR(state), u = (x >> 11) 2^-53; every decision is an
// integer compare of m = x >> 11 with a 53-bit threshold the host derives from
// the reference's f64 arithmetic (glibc exp), so no device floating point is
// involved.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

typedef unsigned __int128 u128;

constexpr int PM_RUN = 16;    // sites per lane
constexpr int PM_BLOCK = 256; // lanes per block
constexpr uint64_t NO_DRAW = ~0ull;

__device__ __forceinline__ u128 make128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

__device__ __forceinline__ uint64_t xsl_rr(u128 s) {
    const uint32_t rot = (uint32_t)(s >> 122);
    const uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// A lane's walk over consecutive entries of a colour array: row r, position
// i in the row and the row's coordinates, advanced without divisions.
struct PmCursor {
    int32_t r, i;
    int32_t x[MAX_D];
};

__device__ __forceinline__ PmCursor pm_cursor(const Geom &g, int64_t w) {
    PmCursor k;
    k.r = (int32_t)(w / g.H);
    k.i = (int32_t)(w - (int64_t)k.r * g.H);
    row_coords(g, k.r, k.x);
    return k;
}

template <int D>
__device__ __forceinline__ void pm_next(const Geom &g, PmCursor &k) {
    if (++k.i < g.H) return;
    k.i = 0;
    ++k.r;
#pragma unroll
    for (int mu = 1; mu < D; ++mu) {
        if (++k.x[mu] < g.L[mu]) return;
        k.x[mu] = 0;
    }
}

// The class of the cursor's site of colour c: (current state, down neighbours).
template <int D>
__device__ __forceinline__ uint32_t site_class(const uint8_t *__restrict__ own, const uint8_t *__restrict__ oth,
                                               const Geom &g, int c, const PmCursor &k) {
    int32_t prs = 0;
#pragma unroll
    for (int mu = 1; mu < D; ++mu) prs += k.x[mu];
    const int32_t x0 = 2 * k.i + ((prs + c) & 1);
    const int32_t L0 = g.L[0];
    const int64_t rowbase = (int64_t)k.r * g.H;
    const int32_t xp = (x0 + 1 == L0) ? 0 : x0 + 1;
    const int32_t xm = (x0 == 0) ? L0 - 1 : x0 - 1;
    uint32_t d = oth[rowbase + (xp >> 1)] + oth[rowbase + (xm >> 1)];
#pragma unroll
    for (int mu = 1; mu < D; ++mu) {
        d += oth[(int64_t)row_step(g, k.r, k.x, mu, +1) * g.H + k.i];
        d += oth[(int64_t)row_step(g, k.r, k.x, mu, -1) * g.H + k.i];
    }
    return (uint32_t)own[rowbase + k.i] * (2 * D + 1) + d;
}

template <int D>
__global__ __launch_bounds__(PM_BLOCK) void pm_count(const uint8_t *__restrict__ own, const uint8_t *__restrict__ oth,
                                                    Geom g, int c, const uint64_t *__restrict__ cls,
                                                    uint32_t *__restrict__ blk, int64_t work) {
    __shared__ uint32_t wsum[PM_BLOCK / 64];
    const int64_t w0 = ((int64_t)blockIdx.x * PM_BLOCK + threadIdx.x) * PM_RUN;
    uint32_t n = 0;
    if (w0 < work) {
        PmCursor k = pm_cursor(g, w0);
        for (int q = 0; q < PM_RUN; ++q) {
            if (w0 + q >= work) break;
            n += cls[site_class<D>(own, oth, g, c, k)] != NO_DRAW ? 1u : 0u;
            pm_next<D>(g, k);
        }
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < PM_BLOCK / 64; ++q) t += wsum[q];
        blk[blockIdx.x] = t;
    }
}

// Exclusive prefix of blk[0..nb) into off[0..nb), off[nb] = total; one block.
__global__ __launch_bounds__(1024) void pm_scan(const uint32_t *__restrict__ blk, uint32_t *__restrict__ off, int nb) {
    __shared__ uint64_t part[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const uint64_t v = i < nb ? blk[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
            const uint64_t t = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) off[i] = (uint32_t)(carry + part[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) off[nb] = (uint32_t)carry;
}

__device__ __forceinline__ u128 jump(u128 s, uint64_t k, const uint64_t *__restrict__ jt) {
    for (int b = 0; k != 0; ++b, k >>= 1)
        if (k & 1u) s = make128(jt[4 * b], jt[4 * b + 1]) * s + make128(jt[4 * b + 2], jt[4 * b + 3]);
    return s;
}

template <int D, int MODE>
__global__ __launch_bounds__(PM_BLOCK) void pm_update(uint8_t *__restrict__ own, const uint8_t *__restrict__ oth,
                                                     Geom g, int c, const uint64_t *__restrict__ cls,
                                                     const uint32_t *__restrict__ off,
                                                     const uint64_t *__restrict__ state,
                                                     const uint64_t *__restrict__ jt, int64_t work) {
    __shared__ uint32_t wave_n[PM_BLOCK / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t w0 = ((int64_t)blockIdx.x * PM_BLOCK + threadIdx.x) * PM_RUN;
    uint32_t cl[PM_RUN];
    uint32_t n = 0;
    PmCursor k{};
    if (w0 < work) k = pm_cursor(g, w0);
#pragma unroll
    for (int q = 0; q < PM_RUN; ++q) {
        cl[q] = 0xffffffffu;
        if (w0 + q < work) {
            cl[q] = site_class<D>(own, oth, g, c, k);
            n += cls[cl[q]] != NO_DRAW ? 1u : 0u;
            pm_next<D>(g, k);
        }
    }
    // draws before this lane within the block: a wave scan, then the waves before
    uint32_t incl = n;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wave_n[wv] = incl;
    __syncthreads();
    uint64_t before = off[blockIdx.x] + (incl - n);
    for (int q = 0; q < wv; ++q) before += wave_n[q];
    const u128 mul = make128(jt[0], jt[1]), inc = make128(jt[2], jt[3]);
    u128 s = jump(make128(state[0], state[1]), before, jt);
#pragma unroll
    for (int q = 0; q < PM_RUN; ++q) {
        if (cl[q] == 0xffffffffu) continue;
        const uint64_t t = cls[cl[q]];
        const int64_t w = w0 + q;
        const uint32_t cur = own[w];
        uint32_t nxt;
        if (t == NO_DRAW) {
            nxt = 1u - cur;  // Metropolis downhill: accepted without a draw
        } else {
            s = mul * s + inc;
            const uint64_t m = xsl_rr(s) >> 11;
            const bool below = m < t;
            nxt = MODE == 0 ? (below ? 1u - cur : cur) : (below ? 0u : 1u);
        }
        own[w] = (uint8_t)nxt;
    }
}

__global__ void pm_advance(uint64_t *__restrict__ state, const uint32_t *__restrict__ off, int nb,
                           const uint64_t *__restrict__ jt) {
    const u128 s = jump(make128(state[0], state[1]), off[nb], jt);
    state[0] = (uint64_t)s;
    state[1] = (uint64_t)(s >> 64);
}

}  // namespace

// --------------------------------------------------------------- host side ---

static const u128 PCG_MUL = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;

void pm_seed(uint64_t seed, uint64_t out[4]) {
    // rand_core seed_from_u64: a PCG32 expansion of the u64 into the 32-byte
    // seed; rand_pcg Lcg128Xsl64::from_seed: state = bytes 0..16,
    // increment = bytes 16..32 | 1, then one step from (state + increment).
    uint8_t bytes[32];
    uint64_t st = seed;
    for (int chunk = 0; chunk < 8; ++chunk) {
        st = st * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xorshifted = (uint32_t)(((st >> 18) ^ st) >> 27);
        const uint32_t rot = (uint32_t)(st >> 59);
        const uint32_t x = (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
        std::memcpy(bytes + 4 * chunk, &x, 4);
    }
    uint64_t w[4];
    std::memcpy(w, bytes, 32);
    const u128 inc = (((u128)w[3] << 64) | w[2]) | 1u;
    u128 s = ((u128)w[1] << 64) | w[0];
    s = (s + inc) * PCG_MUL + inc;
    out[0] = (uint64_t)s;
    out[1] = (uint64_t)(s >> 64);
    out[2] = (uint64_t)inc;
    out[3] = (uint64_t)(inc >> 64);
}

// One Lcg128Xsl64 output (rand_pcg): step, then XSL-RR of the new state.
static uint64_t pm_host_next(u128 &s, u128 inc) {
    s = s * PCG_MUL + inc;
    const uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    const unsigned rot = (unsigned)(s >> 122);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

void pm_hot_start(uint64_t seed, uint32_t q, uint8_t *out, size_t n, uint64_t state_out[4]) {
    uint64_t st[4];
    pm_seed(seed, st);
    u128 s = ((u128)st[1] << 64) | st[0];
    const u128 inc = ((u128)st[3] << 64) | st[2];
    // rand's random_range(0..q) for a usize bound below 2^32 samples a u32
    // (the low half of next_u64) by Canon's method: hi(x * q), and when the
    // low half exceeds 2^32 - q, one more draw decides a carry into it.
    for (size_t i = 0; i < n; ++i) {
        const uint64_t w = (uint64_t)(uint32_t)pm_host_next(s, inc) * q;
        uint32_t result = (uint32_t)(w >> 32);
        const uint32_t lo = (uint32_t)w;
        if (lo > (uint32_t)(0u - q)) {
            const uint32_t hi2 = (uint32_t)(((uint64_t)(uint32_t)pm_host_next(s, inc) * q) >> 32);
            if ((uint64_t)lo + hi2 > 0xffffffffull) result += 1;
        }
        out[i] = (uint8_t)result;
    }
    state_out[0] = (uint64_t)s;
    state_out[1] = (uint64_t)(s >> 64);
    state_out[2] = (uint64_t)inc;
    state_out[3] = (uint64_t)(inc >> 64);
}

void pm_jump_table(const uint64_t state[4], std::vector<uint64_t> &jt) {
    jt.assign(4 * 64, 0);
    u128 A = PCG_MUL, C = ((u128)state[3] << 64) | state[2];
    for (int b = 0; b < 64; ++b) {
        jt[4 * b] = (uint64_t)A;
        jt[4 * b + 1] = (uint64_t)(A >> 64);
        jt[4 * b + 2] = (uint64_t)C;
        jt[4 * b + 3] = (uint64_t)(C >> 64);
        C = (A + 1) * C;  // two steps of (A, C): s -> A (A s + C) + C
        A = A * A;
    }
}

bool pm_class_table(int D, int kernel, double beta, double j, double h, std::vector<uint64_t> &cls, std::string &err) {
    const int NN = 2 * D;
    if (!std::isfinite(beta) || !std::isfinite(j) || !std::isfinite(h)) {
        err = "beta, J and h must be finite";
        return false;
    }
    cls.assign(2 * (NN + 1), NO_DRAW);
    const double two53 = 9007199254740992.0;
    for (int cur = 0; cur < 2; ++cur)
        for (int d = 0; d <= NN; ++d) {
            // updater.rs:77-167 with ising/mod.rs:119-143: decode(0) = +1
            const double nsum = (double)(NN - 2 * d);
            const double ds = cur == 0 ? -2.0 : 2.0;  // decode(1 - cur) - decode(cur)
            const double delta = -ds * (j * nsum + h);
            uint64_t T;
            if (kernel == 0) {
                if (delta <= 0.0) continue;  // downhill: no draw
                const double w = std::exp(-beta * delta);
                const double scaled = std::ceil(w * two53);  // m 2^-53 < w  <=>  m < ceil(w 2^53)
                T = scaled >= two53 ? (uint64_t)1 << 53 : (uint64_t)scaled;
            } else {
                const double d0 = cur == 0 ? 0.0 : delta, d1 = cur == 1 ? 0.0 : delta;
                const double fl = std::fmin(std::fmin(INFINITY, d0), d1);
                const double w0 = std::exp(-beta * (d0 - fl)), w1 = std::exp(-beta * (d1 - fl));
                const double total = 0.0 + w0 + w1;
                // state 0 iff fl(m 2^-53 total) < w0, monotone in m: T = #{m : true}
                uint64_t lo = 0, hi = (uint64_t)1 << 53;
                while (lo < hi) {
                    const uint64_t mid = lo + (hi - lo) / 2;
                    const double u = (double)mid * (1.0 / two53);
                    if (u * total < w0) lo = mid + 1;
                    else hi = mid;
                }
                T = lo;
            }
            cls[cur * (NN + 1) + d] = T;
        }
    return true;
}

int64_t pm_blocks(const Geom &g) {
    const int64_t work = (int64_t)g.R * g.H;
    const int64_t per = (int64_t)PM_BLOCK * PM_RUN;
    return (work + per - 1) / per;
}

void launch_pm_pass(const PmodeArgs &a) {
    const int64_t work = (int64_t)a.g.R * a.g.H;
    const int64_t nb = pm_blocks(a.g);
    dim3 grid((uint32_t)nb), block(PM_BLOCK);
#define PLQ_PM(DD)                                                                                                   \
    pm_count<DD><<<grid, block, 0, a.stream>>>(a.own, a.oth, a.g, a.color, a.cls, a.blk, work);                     \
    pm_scan<<<dim3(1), dim3(1024), 0, a.stream>>>(a.blk, a.off, (int)nb);                                           \
    if (a.mode == 0)                                                                                                 \
        pm_update<DD, 0><<<grid, block, 0, a.stream>>>(a.own, a.oth, a.g, a.color, a.cls, a.off, a.state, a.jt, work); \
    else                                                                                                             \
        pm_update<DD, 1><<<grid, block, 0, a.stream>>>(a.own, a.oth, a.g, a.color, a.cls, a.off, a.state, a.jt, work); \
    pm_advance<<<dim3(1), dim3(1), 0, a.stream>>>(a.state, a.off, (int)nb, a.jt)
    switch (a.g.D) {
        case 1: PLQ_PM(1); break;
        case 2: PLQ_PM(2); break;
        case 3: PLQ_PM(3); break;
        default: PLQ_PM(4); break;
    }
#undef PLQ_PM
}

}  // namespace plq
