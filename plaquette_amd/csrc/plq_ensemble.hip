// plq_ensemble.hip -- batched Z2 gauge replicas of uneven sizes on gfx950.
//
// The reference runs one GpuChain per lattice (device.rs:590-701) with the
// link checkerboard of link_schedule.wgsl:23-40 (2D colour passes per sweep,
// dir-major, parity 0 then 1) and the staple pricing of
// gauge/model.wgsl:136-147.  Here one launch runs a colour pass of every
// replica of an ensemble at once; a replica-exchange step (no reference
// counterpart) swaps inverse temperatures between replicas of equal shape.
//
// Layout: the D links of a site are one word (byte mu = U_mu(x), D <= 4),
// and sites are colour-interleaved as in plq_common.hpp (Geom): colour c of
// replica r is the word array words + rep.off[c], row-major with H = L0 / 2
// words per row.  A pass (mu, c) rewrites byte mu of the colour-c words only;
// every other byte it reads (the other directions, and direction mu of the
// other colour) is not written during the pass, which is the link-colouring
// argument of link_schedule.wgsl.  The new link is written as a single byte,
// so no lane ever rewrites a byte another lane reads.
//
// Each block belongs to one replica (blocks never straddle replicas), so the
// replica's geometry and its threshold table are uniform per block.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_ensemble.hpp"
#include "plq_internal.hpp"

namespace plq {

// n / d for n < 2^24 through a 64-bit reciprocal: q = (n * M) >> 40 with
// M = ceil(2^40 / d) is exact while n * d < 2^40.
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint64_t M) { return (uint32_t)(((uint64_t)n * M) >> 40); }

struct LinkSite {
    uint32_t x0, r;
    uint32_t x[4];
};

// One step of a site along axis a (sign +1/-1) with periodic wrap.
__device__ __forceinline__ LinkSite lstep(const EnsReplica &R, const LinkSite &s, int a, int sign) {
    LinkSite t = s;
    uint32_t l = (uint32_t)R.L[a];
    if (a == 0) {
        t.x0 = sign > 0 ? (s.x0 + 1u == l ? 0u : s.x0 + 1u) : (s.x0 == 0u ? l - 1u : s.x0 - 1u);
    } else {
        uint32_t xa = s.x[a];
        uint32_t st = (uint32_t)R.rstride[a];
        if (sign > 0) {
            t.x[a] = xa + 1u == l ? 0u : xa + 1u;
            t.r = xa + 1u == l ? s.r - (l - 1u) * st : s.r + st;
        } else {
            t.x[a] = xa == 0u ? l - 1u : xa - 1u;
            t.r = xa == 0u ? s.r + (l - 1u) * st : s.r - st;
        }
    }
    return t;
}

// The word of a site: colour (x0 + row parity) & 1, index r * H + x0 / 2.
__device__ __forceinline__ uint32_t lword(const uint32_t *__restrict__ words, const EnsReplica &R, const LinkSite &s,
                                          uint32_t rowpar) {
    uint32_t col = (s.x0 + rowpar) & 1u;
    return ld32(words, R.off[col] + s.r * (uint32_t)R.H + (s.x0 >> 1));
}

__device__ __forceinline__ uint32_t row_parity(const LinkSite &s) { return (s.x[1] + s.x[2] + s.x[3]) & 1u; }

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int d) { return (w >> (8 * d)) & 1u; }

// Block -> replica: the replicas' block ranges are consecutive; a uniform
// scan over the (at most a few hundred) replicas.
__device__ __forceinline__ int block_replica(const EnsReplica *__restrict__ reps, int nrep, uint32_t b) {
    int r = 0;
    while (r + 1 < nrep && reps[r + 1].block0 <= b) ++r;
    return r;
}

// One (mu, c) pass of every replica.  sweep = *sweep_base + sweep_off.
template <int D, int MODE>
__global__ __launch_bounds__(256) void ens_link_pass(uint32_t *__restrict__ words, const EnsReplica *__restrict__ reps,
                                                     int nrep, const uint32_t *__restrict__ tables,
                                                     const uint32_t *__restrict__ beta_slot,
                                                     const uint32_t *__restrict__ sweep_base, uint32_t sweep_off,
                                                     int mu, int c, int per_lane) {
    constexpr int K = 2 * (D - 1);
    __shared__ uint32_t T[2 * (K + 1)];
    const int ri = block_replica(reps, nrep, blockIdx.x);
    const EnsReplica R = reps[ri];
    if (threadIdx.x < 2 * (K + 1)) T[threadIdx.x] = tables[beta_slot[ri] * (2 * (K + 1)) + threadIdx.x];
    __syncthreads();
    const uint32_t skey = (*sweep_base + sweep_off) * GOLDEN;
    const uint32_t half = R.nsites >> 1;
    const uint32_t base = (blockIdx.x - R.block0) * blockDim.x * (uint32_t)per_lane + threadIdx.x;
    for (int k = 0; k < per_lane; ++k) {
        const uint32_t item = base + (uint32_t)k * blockDim.x;
        if (item >= half) break;
        LinkSite s;
        s.r = fdiv(item, R.mH);
        const uint32_t i = item - s.r * (uint32_t)R.H;
        uint32_t rest = s.r;
        s.x[0] = 0;
        s.x[1] = 0;
        s.x[2] = 0;
        s.x[3] = 0;
        if (D >= 2) {
            uint32_t q = fdiv(rest, R.mL[1]);
            s.x[1] = rest - q * (uint32_t)R.L[1];
            rest = q;
        }
        if (D >= 3) {
            uint32_t q = fdiv(rest, R.mL[2]);
            s.x[2] = rest - q * (uint32_t)R.L[2];
            rest = q;
        }
        if (D >= 4) s.x[3] = rest;
        const uint32_t rp = row_parity(s);
        s.x0 = 2u * i + ((rp + (uint32_t)c) & 1u);
        const uint32_t wx = lword(words, R, s, rp);
        const LinkSite xm = lstep(R, s, mu, +1);
        const uint32_t wxm = lword(words, R, xm, row_parity(xm));
        int neg = 0;
#pragma unroll
        for (int nu = 0; nu < D; ++nu) {
            if (nu == mu) continue;
            const LinkSite xpn = lstep(R, s, nu, +1);
            const LinkSite xmn = lstep(R, s, nu, -1);
            const LinkSite xmnp = lstep(R, xmn, mu, +1);
            const uint32_t wpn = lword(words, R, xpn, row_parity(xpn));
            const uint32_t wmn = lword(words, R, xmn, row_parity(xmn));
            const uint32_t wmnp = lword(words, R, xmnp, row_parity(xmnp));
            // forward: U_nu(x+mu) U_mu(x+nu) U_nu(x); backward: U_mu(x-nu) U_nu(x-nu+mu) U_nu(x-nu)
            neg += (int)(byte_of(wxm, nu) ^ byte_of(wpn, mu) ^ byte_of(wx, nu));
            neg += (int)(byte_of(wmn, mu) ^ byte_of(wmnp, nu) ^ byte_of(wmn, nu));
        }
        const uint32_t cur = byte_of(wx, mu);
        const uint32_t t = T[cur * (K + 1) + (uint32_t)neg];
        const uint32_t site = s.r * (uint32_t)R.L[0] + s.x0;
        const uint32_t mm = keyed_bits_k(R.seed32, site * (uint32_t)D + (uint32_t)mu, skey);
        const bool hit = mm < t;
        const uint32_t nxt = MODE == 0 ? (cur ^ (hit ? 1u : 0u)) : (hit ? 0u : 1u);
        if (nxt != cur) {
            uint8_t *bytes = reinterpret_cast<uint8_t *>(words);
            const uint32_t widx = R.off[c] + s.r * (uint32_t)R.H + i;
            bytes[(size_t)widx * 4u + (uint32_t)mu] = (uint8_t)nxt;
        }
    }
}

// Signed plaquette sum of every replica into sums[r] (zeroed by the caller).
template <int D>
__global__ __launch_bounds__(256) void ens_plaquette(const uint32_t *__restrict__ words,
                                                     const EnsReplica *__restrict__ reps, int nrep,
                                                     unsigned long long *__restrict__ sums, int per_lane) {
    __shared__ long long lds[8];
    const int ri = block_replica(reps, nrep, blockIdx.x);
    const EnsReplica R = reps[ri];
    long long acc = 0;
    const uint32_t base = (blockIdx.x - R.block0) * blockDim.x * (uint32_t)per_lane + threadIdx.x;
    for (int k = 0; k < 2 * per_lane; ++k) {
        // every site once: items [0, nsites) in pairs of half-lattices
        const uint32_t item = base + (uint32_t)(k >> 1) * blockDim.x;
        const uint32_t half = R.nsites >> 1;
        if (item >= half) break;
        LinkSite s;
        s.r = fdiv(item, R.mH);
        const uint32_t i = item - s.r * (uint32_t)R.H;
        uint32_t rest = s.r;
        s.x[0] = s.x[1] = s.x[2] = s.x[3] = 0;
        if (D >= 2) {
            uint32_t q = fdiv(rest, R.mL[1]);
            s.x[1] = rest - q * (uint32_t)R.L[1];
            rest = q;
        }
        if (D >= 3) {
            uint32_t q = fdiv(rest, R.mL[2]);
            s.x[2] = rest - q * (uint32_t)R.L[2];
            rest = q;
        }
        if (D >= 4) s.x[3] = rest;
        const uint32_t rp = row_parity(s);
        s.x0 = 2u * i + ((rp + (uint32_t)(k & 1)) & 1u);
        const uint32_t wx = lword(words, R, s, rp);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const LinkSite sa = lstep(R, s, a, +1);
            const uint32_t wa = lword(words, R, sa, row_parity(sa));
#pragma unroll
            for (int b = a + 1; b < D; ++b) {
                const LinkSite sb = lstep(R, s, b, +1);
                const uint32_t wb = lword(words, R, sb, row_parity(sb));
                // U_a(x) U_b(x+a) U_a(x+b) U_b(x)  (lattice.rs:420-430)
                uint32_t x = byte_of(wx, a) ^ byte_of(wa, b) ^ byte_of(wb, a) ^ byte_of(wx, b);
                acc += 1 - 2 * (int)x;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[w];
        atomicAdd(&sums[ri], (unsigned long long)t);
    }
}

__global__ void ens_zero(unsigned long long *__restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Replica exchange between neighbouring inverse temperatures of each group
// (replicas of one shape).  slot_of[r] = index into betas of replica r;
// rep_at[g][k] = replica holding the k-th smallest beta of group g.  Round n
// pairs (k, k+1) with k = n mod 2, 2 + n mod 2, ...  The pair (k, k+1) of
// group g swaps iff  x = P_a - P_b  (a at beta_k, b at beta_k+1) gives
// delta = (beta_k+1 - beta_k) J x >= 0, or the keyed draw is below the
// host-built threshold for x (f64 exp, as every other acceptance here).
// One thread per pair; pairs of one round are disjoint.
__global__ void ens_exchange(const EnsGroupPair *__restrict__ pairs, int npairs, uint32_t *__restrict__ slot_of,
                             uint32_t *__restrict__ rep_at, const unsigned long long *__restrict__ sums,
                             const uint32_t *__restrict__ ex_tables, const uint32_t *__restrict__ round_base,
                             uint32_t round_off, uint32_t ex_seed, unsigned long long *__restrict__ accepted) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const EnsGroupPair P = pairs[p];
    const uint32_t round = *round_base + round_off;
    if ((uint32_t)P.k % 2u != round % 2u) return;
    const uint32_t a = rep_at[P.rep_at0 + P.k];
    const uint32_t b = rep_at[P.rep_at0 + P.k + 1];
    const long long x = (long long)sums[a] - (long long)sums[b];
    bool swap;
    if (P.positive == 2u || (P.positive ? x >= 0 : x <= 0)) {  // delta >= 0 (2: equal betas)
        swap = true;
    } else {
        long long ax = x < 0 ? -x : x;
        uint32_t t = ax < (long long)P.table_len ? ex_tables[P.table0 + (uint32_t)ax] : 0u;
        uint32_t m = keyed_bits(ex_seed, (uint32_t)p, round);
        swap = m < t;
    }
    if (swap) {
        rep_at[P.rep_at0 + P.k] = b;
        rep_at[P.rep_at0 + P.k + 1] = a;
        uint32_t sa = slot_of[a];
        slot_of[a] = slot_of[b];
        slot_of[b] = sa;
        atomicAdd(&accepted[p], 1ull);
    }
}

// counters[0] = sweep base, counters[1] = exchange round, counters[2] = records
__global__ void ens_tick(uint32_t *__restrict__ counters, uint32_t sweeps, uint32_t rounds) {
    counters[0] += sweeps;
    counters[1] += rounds;
}

// Copy the current sums and beta slots into the measurement ring.
__global__ void ens_record(const unsigned long long *__restrict__ sums, const uint32_t *__restrict__ slot_of, int nrep,
                           long long *__restrict__ ring_sums, uint32_t *__restrict__ ring_slots, uint32_t cap,
                           uint32_t *__restrict__ counters) {
    const uint32_t slot = counters[2] % cap;
    for (int i = threadIdx.x; i < nrep; i += blockDim.x) {
        ring_sums[(size_t)slot * nrep + i] = (long long)sums[i];
        ring_slots[(size_t)slot * nrep + i] = slot_of[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) counters[2] += 1;
}

// lattice order (link = site * D + dir) <-> words
__global__ void ens_pack(const uint8_t *__restrict__ lat, uint32_t *__restrict__ words, EnsReplica R, int D) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= R.nsites) return;
    uint32_t r = s / (uint32_t)R.L[0];
    uint32_t x0 = s - r * (uint32_t)R.L[0];
    uint32_t rest = r, par = x0;
    for (int a = 1; a < D; ++a) {
        par += rest % (uint32_t)R.L[a];
        rest /= (uint32_t)R.L[a];
    }
    uint32_t w = 0;
    for (int d = 0; d < D; ++d) w |= (uint32_t)lat[(size_t)s * D + d] << (8 * d);
    words[R.off[par & 1u] + r * (uint32_t)R.H + (x0 >> 1)] = w;
}

__global__ void ens_unpack(uint8_t *__restrict__ lat, const uint32_t *__restrict__ words, EnsReplica R, int D) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= R.nsites) return;
    uint32_t r = s / (uint32_t)R.L[0];
    uint32_t x0 = s - r * (uint32_t)R.L[0];
    uint32_t rest = r, par = x0;
    for (int a = 1; a < D; ++a) {
        par += rest % (uint32_t)R.L[a];
        rest /= (uint32_t)R.L[a];
    }
    uint32_t w = words[R.off[par & 1u] + r * (uint32_t)R.H + (x0 >> 1)];
    for (int d = 0; d < D; ++d) lat[(size_t)s * D + d] = (uint8_t)((w >> (8 * d)) & 0xffu);
}

// ------------------------------------------------------------- launchers ---

void launch_ens_pack(const uint8_t *lat, uint32_t *words, const EnsReplica &R, int D, hipStream_t s) {
    ens_pack<<<dim3((R.nsites + 255) / 256), dim3(256), 0, s>>>(lat, words, R, D);
}

void launch_ens_unpack(uint8_t *lat, const uint32_t *words, const EnsReplica &R, int D, hipStream_t s) {
    ens_unpack<<<dim3((R.nsites + 255) / 256), dim3(256), 0, s>>>(lat, words, R, D);
}

void launch_ens_pass(const EnsPassArgs &a) {
    dim3 grid(a.blocks), block(a.block_size);
#define PLQ_ENS(DD)                                                                                          \
    if (a.mode == 0)                                                                                         \
        ens_link_pass<DD, 0><<<grid, block, 0, a.stream>>>(a.words, a.reps, a.nrep, a.tables, a.beta_slot,   \
                                                           a.sweep_base, a.sweep_off, a.mu, a.c, a.per_lane); \
    else                                                                                                     \
        ens_link_pass<DD, 1><<<grid, block, 0, a.stream>>>(a.words, a.reps, a.nrep, a.tables, a.beta_slot,   \
                                                           a.sweep_base, a.sweep_off, a.mu, a.c, a.per_lane)
    switch (a.D) {
        case 2: PLQ_ENS(2); break;
        case 3: PLQ_ENS(3); break;
        default: PLQ_ENS(4); break;
    }
#undef PLQ_ENS
}

void launch_ens_plaquette(const uint32_t *words, const EnsReplica *reps, int nrep, int D, uint32_t blocks,
                          int block_size, int per_lane, unsigned long long *sums, hipStream_t s) {
    ens_zero<<<dim3((nrep + 255) / 256), dim3(256), 0, s>>>(sums, nrep);
    dim3 grid(blocks), block(block_size);
    switch (D) {
        case 2: ens_plaquette<2><<<grid, block, 0, s>>>(words, reps, nrep, sums, per_lane); break;
        case 3: ens_plaquette<3><<<grid, block, 0, s>>>(words, reps, nrep, sums, per_lane); break;
        default: ens_plaquette<4><<<grid, block, 0, s>>>(words, reps, nrep, sums, per_lane); break;
    }
}

void launch_ens_exchange(const EnsGroupPair *pairs, int npairs, uint32_t *slot_of, uint32_t *rep_at,
                         const unsigned long long *sums, const uint32_t *ex_tables, const uint32_t *round_base,
                         uint32_t round_off, uint32_t ex_seed, unsigned long long *accepted, hipStream_t s) {
    if (npairs <= 0) return;
    ens_exchange<<<dim3((npairs + 63) / 64), dim3(64), 0, s>>>(pairs, npairs, slot_of, rep_at, sums, ex_tables,
                                                             round_base, round_off, ex_seed, accepted);
}

void launch_ens_tick(uint32_t *counters, uint32_t sweeps, uint32_t rounds, hipStream_t s) {
    ens_tick<<<dim3(1), dim3(1), 0, s>>>(counters, sweeps, rounds);
}

void launch_ens_record(const unsigned long long *sums, const uint32_t *slot_of, int nrep, long long *ring_sums,
                       uint32_t *ring_slots, uint32_t cap, uint32_t *counters, hipStream_t s) {
    ens_record<<<dim3(1), dim3(256), 0, s>>>(sums, slot_of, nrep, ring_sums, ring_slots, cap, counters);
}

}  // namespace plq
