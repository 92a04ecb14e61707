// plq_ensemble.hip -- batched Z2 gauge replicas of uneven sizes on gfx950.
//
// The reference runs one GpuChain per lattice (device.rs:590-701) with the
// link checkerboard of link_schedule.wgsl:23-40 (2D colour passes per sweep,
// dir-major, parity 0 then 1) and the staple pricing of
// gauge/model.wgsl:136-147.  Here one launch runs a colour pass of every
// replica of an ensemble at once; a replica-exchange step (no reference
// counterpart) swaps inverse temperatures between replicas of equal shape.
//
// Layout: replicas of one shape are packed four to a word (EnsQuad): byte j
// of a word is replica rep[j]'s link, and a quad has one word array per
// (direction mu, colour c), each colour-interleaved as in plq_common.hpp
// (Geom: site (x0, row r) of colour c at r * H + x0 / 2, H = L0 / 2).  A lane
// updates the link (x, mu) of all four replicas at once: the coordinates,
// the 18 neighbour words and the staple products (XORs of 0/1 bytes, summed
// per byte) are shared; only the draw, keyed by each replica's own seed, and
// the threshold of each replica's current beta are per byte.  A pass (mu, c)
// rewrites words of array (mu, c) only and reads arrays it does not write,
// which is the link-colouring argument of link_schedule.wgsl; each word is
// owned by one lane.  Every neighbour's colour is fixed by its distance: one
// step away is the other colour, two steps the same (extents are even).
//
// Each block belongs to one quad (blocks never straddle quads), so the
// geometry and the four threshold tables are uniform per block.
#include <hip/hip_cooperative_groups.h>
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

// One step of a site along axis A (SIGN +1/-1) with periodic wrap.  The
// axis is a template parameter so every coordinate lives in a register (a
// runtime index into the coordinate array would put it in scratch memory).
template <int A, int SIGN>
__device__ __forceinline__ LinkSite lstep(const EnsQuad &Q, const LinkSite &s) {
    LinkSite t = s;
    const uint32_t l = (uint32_t)Q.L[A];
    if (A == 0) {
        t.x0 = SIGN > 0 ? (s.x0 + 1u == l ? 0u : s.x0 + 1u) : (s.x0 == 0u ? l - 1u : s.x0 - 1u);
    } else {
        const uint32_t xa = s.x[A];
        const uint32_t st = (uint32_t)Q.rstride[A];
        if (SIGN > 0) {
            t.x[A] = xa + 1u == l ? 0u : xa + 1u;
            t.r = xa + 1u == l ? s.r - (l - 1u) * st : s.r + st;
        } else {
            t.x[A] = xa == 0u ? l - 1u : xa - 1u;
            t.r = xa == 0u ? s.r + (l - 1u) * st : s.r - st;
        }
    }
    return t;
}

// Index of a site within its colour array.
__device__ __forceinline__ uint32_t cidx(const EnsQuad &Q, const LinkSite &s) { return s.r * (uint32_t)Q.H + (s.x0 >> 1); }

// Word array of (direction mu, colour c).
__device__ __forceinline__ uint32_t arr(const EnsQuad &Q, uint32_t half, int mu, uint32_t c) {
    return Q.woff + ((uint32_t)mu * 2u + c) * half;
}

// Site of colour c at index item of its colour array.
template <int D>
__device__ __forceinline__ LinkSite site_of(const EnsQuad &Q, uint32_t item, uint32_t c) {
    LinkSite s;
    s.r = fdiv(item, Q.mH);
    const uint32_t i = item - s.r * (uint32_t)Q.H;
    uint32_t rest = s.r;
    s.x[0] = s.x[1] = s.x[2] = s.x[3] = 0;
    if (D >= 2) {
        uint32_t q = fdiv(rest, Q.mL[1]);
        s.x[1] = rest - q * (uint32_t)Q.L[1];
        rest = q;
    }
    if (D >= 3) {
        uint32_t q = fdiv(rest, Q.mL[2]);
        s.x[2] = rest - q * (uint32_t)Q.L[2];
        rest = q;
    }
    if (D >= 4) s.x[3] = rest;
    s.x0 = 2u * i + ((s.x[1] + s.x[2] + s.x[3] + c) & 1u);
    return s;
}

// Negative staple products of link (x, MU) in the (MU, NU) plane, one count
// per byte (replica):
//   forward  U_NU(x+MU) U_MU(x+NU) U_NU(x)
//   backward U_MU(x-NU) U_NU(x-NU+MU) U_NU(x-NU)
// x has colour c; x+MU, x+NU, x-NU have colour 1-c; x-NU+MU has colour c.
template <int MU, int NU>
__device__ __forceinline__ uint32_t staple_negs4(const uint32_t *__restrict__ words, const EnsQuad &Q, uint32_t half,
                                                 const LinkSite &s, uint32_t c, uint32_t item, uint32_t ixp) {
    const uint32_t o = 1u - c;
    const LinkSite xpn = lstep<NU, +1>(Q, s);
    const LinkSite xmn = lstep<NU, -1>(Q, s);
    const LinkSite xmnp = lstep<MU, +1>(Q, xmn);
    const uint32_t imn = cidx(Q, xmn);
    const uint32_t fw = ld32(words, arr(Q, half, NU, o) + ixp) ^ ld32(words, arr(Q, half, MU, o) + cidx(Q, xpn)) ^
                        ld32(words, arr(Q, half, NU, c) + item);
    const uint32_t bw = ld32(words, arr(Q, half, MU, o) + imn) ^ ld32(words, arr(Q, half, NU, c) + cidx(Q, xmnp)) ^
                        ld32(words, arr(Q, half, NU, o) + imn);
    return fw + bw;
}

// Logical block of this workgroup.  Workgroups are dispatched round-robin
// over the 8 XCDs; with the grid a multiple of 8 (padded blocks carry
// NO_QUAD), XCD x runs logical blocks [x G/8, (x+1) G/8), i.e. a contiguous
// run of quads, so each XCD's L2 holds only its replicas' links.
constexpr uint16_t NO_QUAD = 0xffff;
__device__ __forceinline__ uint32_t ens_logical_block() {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    if ((G & 7u) != 0) return b;
    return (b & 7u) * (G >> 3) + (b >> 3);
}

// One (MU, c) pass of every replica.  sweep = *sweep_base + sweep_off.
template <int D, int MODE, int MU>
__global__ __launch_bounds__(256) void ens_link_pass(uint32_t *__restrict__ words, const EnsQuad *__restrict__ quads,
                                                     const uint16_t *__restrict__ block_quad,
                                                     const uint32_t *__restrict__ tables,
                                                     const uint32_t *__restrict__ beta_slot,
                                                     const uint32_t *__restrict__ sweep_base, uint32_t sweep_off,
                                                     int c, int per_lane) {
    constexpr int K = 2 * (D - 1);
    constexpr int TL = 2 * (K + 1);
    __shared__ uint32_t T[4 * TL];
    const uint32_t lb = ens_logical_block();
    const int qi = block_quad[lb];
    if (qi == NO_QUAD) return;
    const EnsQuad Q = quads[qi];
    if (threadIdx.x < 4 * TL) {
        const int j = threadIdx.x / TL, e = threadIdx.x - j * TL;
        const int rj = quads[qi].rep[j];  // runtime index: read from memory, not the register copy
        T[threadIdx.x] = rj >= 0 ? tables[beta_slot[rj] * TL + e] : 0u;
    }
    __syncthreads();
    const uint32_t uc = (uint32_t)c;
    const uint32_t skey = (*sweep_base + sweep_off) * GOLDEN;
    const uint32_t half = Q.nsites >> 1;
    const uint32_t base = (lb - Q.block0) * blockDim.x * (uint32_t)per_lane + threadIdx.x;
    for (int k = 0; k < per_lane; ++k) {
        const uint32_t item = base + (uint32_t)k * blockDim.x;
        if (item >= half) break;
        const LinkSite s = site_of<D>(Q, item, uc);
        const uint32_t ixp = cidx(Q, lstep<MU, +1>(Q, s));
        uint32_t neg = 0;
        if (MU != 0) neg += staple_negs4<MU, 0>(words, Q, half, s, uc, item, ixp);
        if (D > 1 && MU != 1) neg += staple_negs4<MU, 1>(words, Q, half, s, uc, item, ixp);
        if (D > 2 && MU != 2) neg += staple_negs4<MU, (D > 2 ? 2 : 0)>(words, Q, half, s, uc, item, ixp);
        if (D > 3 && MU != 3) neg += staple_negs4<MU, (D > 3 ? 3 : 0)>(words, Q, half, s, uc, item, ixp);
        const uint32_t widx = arr(Q, half, MU, uc) + item;
        const uint32_t wcur = ld32(words, widx);
        const uint32_t link = (s.r * (uint32_t)Q.L[0] + s.x0) * (uint32_t)D + (uint32_t)MU;
        uint32_t wnew = wcur;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (Q.rep[j] < 0) continue;
            const uint32_t cur = (wcur >> (8 * j)) & 1u;
            const uint32_t n = (neg >> (8 * j)) & 0xffu;
            const uint32_t t = T[j * TL + cur * (K + 1) + n];
            const bool hit = keyed_bits_k(Q.seed32[j], link, skey) < t;
            const uint32_t nxt = MODE == 0 ? (cur ^ (hit ? 1u : 0u)) : (hit ? 0u : 1u);
            wnew = (wnew & ~(0xffu << (8 * j))) | (nxt << (8 * j));
        }
        if (wnew != wcur) words[widx] = wnew;
    }
}

// Index deltas of one-step moves in a colour array: moving a site along axis
// a changes its index r * H + x0 / 2 by a delta that depends only on that
// axis's coordinate (and on x0's parity for axis 0), so moves along several
// axes add their deltas.
struct EnsDeltas {
    uint32_t p[4], m[4];  // successor / predecessor along each axis (mod 2^32)
};

template <int D>
__device__ __forceinline__ EnsDeltas ens_deltas(const EnsQuad &Q, const LinkSite &s) {
    EnsDeltas d;
    const uint32_t H = (uint32_t)Q.H, L0 = (uint32_t)Q.L[0];
    // axis 0: x0 odd moves to the next index going up, x0 even going down
    const bool odd = (s.x0 & 1u) != 0;
    d.p[0] = odd ? (s.x0 + 1u == L0 ? 0u - (H - 1u) : 1u) : 0u;
    d.m[0] = odd ? 0u : (s.x0 == 0u ? H - 1u : 0u - 1u);
#pragma unroll
    for (int a = 1; a < 4; ++a) {
        if (a >= D) break;
        const uint32_t l = (uint32_t)Q.L[a];
        const uint32_t st = (uint32_t)Q.rstride[a] * H;  // uniform
        d.p[a] = s.x[a] + 1u == l ? 0u - (l - 1u) * st : st;
        d.m[a] = s.x[a] == 0u ? (l - 1u) * st : 0u - st;
    }
    return d;
}

// One (MU, c) pass of every replica, as ens_link_pass, with the staple words
// addressed by index deltas from the lane's own index (no per-neighbour
// coordinate arithmetic) and the four replicas' decisions taken as in the
// uniform-parity site pass: the class bytes cur * (K + 1) + n of all four
// replicas in one dword, and per replica one compare of the unshifted hash
// y against threshold * 256 - 1 into VCC and one add-with-carry.  Every
// threshold of every beta slot must be at least 1 (the host checks).
// The body of one logical block lb of a (MU, c) pass, for ens_link_pass2
// (one launch per pass) and ens_sweeps_coop (every pass of a block of
// sweeps in one launch).  T is the block's LDS threshold table.
template <int D, int MODE, int MU>
__device__ __forceinline__ void ens_pass2_block(uint32_t *__restrict__ words, const EnsQuad *__restrict__ quads,
                                                const uint16_t *__restrict__ block_quad,
                                                const uint32_t *__restrict__ qtab, uint32_t kmix, int c,
                                                int per_lane, uint32_t lb, uint32_t *T) {
    constexpr int K = 2 * (D - 1);
    constexpr int TL = 2 * (K + 1);
    const int qi = block_quad[lb];
    if (qi == NO_QUAD) return;  // the same for the whole block
    const EnsQuad Q = quads[qi];
    if (threadIdx.x < 4 * TL) T[threadIdx.x] = qtab[(uint32_t)qi * (4 * TL) + threadIdx.x];
    __syncthreads();
    const uint32_t uc = (uint32_t)c, uo = 1u - uc;
    const uint32_t half = Q.nsites >> 1;
    // bytes of the word that hold a replica (unused bytes keep their value)
    uint32_t used = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) used |= Q.rep[j] >= 0 ? 0xffu << (8 * j) : 0u;
    // the word arrays this pass reads, as uniform bases
    const uint32_t *Wc[4], *Wo[4];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        Wc[a] = words + arr(Q, half, a, uc);
        Wo[a] = words + arr(Q, half, a, uo);
    }
    uint32_t *own = words + arr(Q, half, MU, uc);
    const char *Tb = reinterpret_cast<const char *>(T);
    const uint32_t base = (lb - Q.block0) * blockDim.x * (uint32_t)per_lane + threadIdx.x;
    for (int k = 0; k < per_lane; ++k) {
        const uint32_t item = base + (uint32_t)k * blockDim.x;
        if (item >= half) break;
        const LinkSite s = site_of<D>(Q, item, uc);
        const EnsDeltas dl = ens_deltas<D>(Q, s);
        const uint32_t ixp = item + dl.p[MU];
        uint32_t neg = 0;
#pragma unroll
        for (int nu = 0; nu < D; ++nu) {
            if (nu == MU) continue;
            const uint32_t ixpn = item + dl.p[nu];
            const uint32_t imn = item + dl.m[nu];
            const uint32_t imnp = imn + dl.p[MU];
            // forward U_NU(x+MU) U_MU(x+NU) U_NU(x), backward U_MU(x-NU) U_NU(x-NU+MU) U_NU(x-NU)
            const uint32_t fw = ld32(Wo[nu], ixp) ^ ld32(Wo[MU], ixpn) ^ ld32(Wc[nu], item);
            const uint32_t bw = ld32(Wo[MU], imn) ^ ld32(Wc[nu], imnp) ^ ld32(Wo[nu], imn);
            neg += fw + bw;
        }
        const uint32_t wcur = ld32(own, item);
        const uint32_t link = (s.r * (uint32_t)Q.L[0] + s.x0) * (uint32_t)D + (uint32_t)MU;
        const uint32_t cls4 = wcur * (uint32_t)(K + 1) + neg;  // per byte: cur * (K + 1) + n <= 2K + 1
        uint32_t bits = 0;  // rejections, replica 3 first, so that replica j lands in bit j
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            const uint32_t x = Q.seed32[j] ^ link;
            const uint32_t y = lowbias32_tail(lowbias32_mid(x ^ (x >> 16)) ^ kmix);
            const uint32_t t2 =
                *reinterpret_cast<const uint32_t *>(Tb + 4 * (j * TL) + (((cls4 >> (8 * j)) & 0xffu) << 2));
            asm("v_cmp_gt_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
                : "+v"(bits)
                : "v"(y), "v"(t2)
                : "vcc");
        }
        const uint32_t raw = bits * 0x00204081u;  // bit j -> bit 8 j
        uint32_t wnew = MODE == 0 ? (wcur ^ (~raw & 0x01010101u)) : (raw & 0x01010101u);
        wnew = (wnew & used) | (wcur & ~used);
        if (wnew != wcur) own[item] = wnew;
    }
}

// One (MU, c) pass of every replica, as ens_link_pass, with the staple words
// addressed by index deltas from the lane's own index (no per-neighbour
// coordinate arithmetic) and the four replicas' decisions taken as in the
// uniform-parity site pass: the class bytes cur * (K + 1) + n of all four
// replicas in one dword, and per replica one compare of the unshifted hash
// y against threshold * 256 - 1 into VCC and one add-with-carry.  Every
// threshold of every beta slot must be at least 1 (the host checks).
template <int D, int MODE, int MU>
__global__ __launch_bounds__(256) void ens_link_pass2(uint32_t *__restrict__ words, const EnsQuad *__restrict__ quads,
                                                      const uint16_t *__restrict__ block_quad,
                                                      const uint32_t *__restrict__ qtab,
                                                      const uint32_t *__restrict__ sweep_base, uint32_t sweep_off,
                                                      int c, int per_lane) {
    constexpr int TL = 2 * (2 * (D - 1) + 1);
    __shared__ uint32_t T[4 * TL];  // T[j * TL + cur * (K + 1) + n] = threshold * 256 - 1
    const uint32_t kmix = sweep_mix((*sweep_base + sweep_off) * GOLDEN);
    ens_pass2_block<D, MODE, MU>(words, quads, block_quad, qtab, kmix, c, per_lane, ens_logical_block(), T);
}

// Every colour pass of nsweeps sweeps of every replica in one cooperative
// launch: the grid's blocks walk the logical blocks of a pass, and a grid
// barrier separates consecutive passes (a link's staples hold links of the
// other passes).  Bit for bit the per-pass launches; measured slower, since a
// grid barrier costs more than a launch boundary on this GPU.
template <int D, int MODE>
__global__ __launch_bounds__(256) void ens_sweeps_coop(uint32_t *__restrict__ words, const EnsQuad *__restrict__ quads,
                                                       const uint16_t *__restrict__ block_quad,
                                                       const uint32_t *__restrict__ qtab,
                                                       const uint32_t *__restrict__ sweep_base, uint32_t nsweeps,
                                                       int per_lane, uint32_t nblocks) {
    constexpr int TL = 2 * (2 * (D - 1) + 1);
    __shared__ uint32_t T[4 * TL];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const uint32_t sweep0 = *sweep_base;  // the counter ticks after the launch
    for (uint32_t k = 0; k < nsweeps; ++k) {
        const uint32_t kmix = sweep_mix((sweep0 + k) * GOLDEN);
#pragma unroll
        for (int mu = 0; mu < D; ++mu)
            for (int c = 0; c < 2; ++c) {
                for (uint32_t lb = blockIdx.x; lb < nblocks; lb += gridDim.x) {
                    __syncthreads();  // the previous logical block is done with T
                    switch (mu) {
                        case 0: ens_pass2_block<D, MODE, 0>(words, quads, block_quad, qtab, kmix, c, per_lane, lb, T); break;
                        case 1: ens_pass2_block<D, MODE, 1>(words, quads, block_quad, qtab, kmix, c, per_lane, lb, T); break;
                        case 2: ens_pass2_block<D, MODE, (2 < D ? 2 : 0)>(words, quads, block_quad, qtab, kmix, c, per_lane, lb, T); break;
                        default: ens_pass2_block<D, MODE, (3 < D ? 3 : 0)>(words, quads, block_quad, qtab, kmix, c, per_lane, lb, T); break;
                    }
                }
                grid.sync();
            }
    }
}

// Signed plaquette sum of every replica into sums[r] (zeroed by the caller):
// per site and plane (A < B), U_A(x) U_B(x+A) U_A(x+B) U_B(x)
// (lattice.rs:420-430), negative when the XOR of the four bits is 1.
template <int A, int B>
__device__ __forceinline__ uint32_t plaq_bits4(const uint32_t *__restrict__ words, const EnsQuad &Q, uint32_t half,
                                               const LinkSite &s, uint32_t c, uint32_t item) {
    const uint32_t o = 1u - c;
    return ld32(words, arr(Q, half, A, c) + item) ^ ld32(words, arr(Q, half, B, o) + cidx(Q, lstep<A, +1>(Q, s))) ^
           ld32(words, arr(Q, half, A, o) + cidx(Q, lstep<B, +1>(Q, s))) ^ ld32(words, arr(Q, half, B, c) + item);
}

template <int D>
__global__ __launch_bounds__(256) void ens_plaquette(const uint32_t *__restrict__ words,
                                                     const EnsQuad *__restrict__ quads,
                                                     const uint16_t *__restrict__ block_quad,
                                                     unsigned long long *__restrict__ sums, int per_lane) {
    constexpr int PLANES = D * (D - 1) / 2;
    __shared__ long long lds[4][8];
    const uint32_t lb = ens_logical_block();
    const int qi = block_quad[lb];
    if (qi == NO_QUAD) return;
    const EnsQuad Q = quads[qi];
    const uint32_t half = Q.nsites >> 1;
    uint32_t negs = 0;   // per byte: negative plaquettes, at most 6 per site and 8 sites
    uint32_t sites = 0;
    const uint32_t base = (lb - Q.block0) * blockDim.x * (uint32_t)per_lane + threadIdx.x;
    for (int k = 0; k < 2 * per_lane; ++k) {
        // every site once: both colours of each of the lane's items
        const uint32_t item = base + (uint32_t)(k >> 1) * blockDim.x;
        if (item >= half) break;
        const uint32_t c = (uint32_t)(k & 1);
        const LinkSite s = site_of<D>(Q, item, c);
        uint32_t n = plaq_bits4<0, 1>(words, Q, half, s, c, item);
        if (D >= 3) n += plaq_bits4<0, (D >= 3 ? 2 : 1)>(words, Q, half, s, c, item) +
                         plaq_bits4<1, (D >= 3 ? 2 : 1)>(words, Q, half, s, c, item);
        if (D >= 4) n += plaq_bits4<0, (D >= 4 ? 3 : 1)>(words, Q, half, s, c, item) +
                         plaq_bits4<1, (D >= 4 ? 3 : 1)>(words, Q, half, s, c, item) +
                         plaq_bits4<2, (D >= 4 ? 3 : 1)>(words, Q, half, s, c, item);
        negs += n;
        sites += 1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        long long acc = (long long)sites * PLANES - 2 * (long long)((negs >> (8 * j)) & 0xffu);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) lds[j][threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int rj = quads[qi].rep[threadIdx.x];  // runtime index: read from memory
        if (rj >= 0) {
            long long t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[threadIdx.x][w];
            atomicAdd(&sums[rj], (unsigned long long)t);
        }
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
__device__ __forceinline__ void ens_exchange_pair(int p, uint32_t round, const EnsGroupPair *__restrict__ pairs,
                                                  uint32_t *__restrict__ slot_of, uint32_t *__restrict__ rep_at,
                                                  const unsigned long long *__restrict__ sums,
                                                  const uint32_t *__restrict__ ex_tables, uint32_t ex_seed,
                                                  unsigned long long *__restrict__ accepted) {
    const EnsGroupPair P = pairs[p];
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


// EnsQtab's entries from the beta slots, by the threads of one workgroup.
__device__ __forceinline__ void ens_qtab_fill(const EnsQtab &qt, const uint32_t *__restrict__ slot_of) {
    const int per_quad = 4 * qt.TL, n = qt.nquads * per_quad;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int q = i / per_quad, r = i - q * per_quad, j = r / qt.TL, e = r - j * qt.TL;
        const int rep = qt.quads[q].rep[j];
        qt.qtab[i] = rep >= 0 ? (qt.tables[slot_of[rep] * (uint32_t)qt.TL + (uint32_t)e] << 8) - 1u : 0u;
    }
}

__global__ __launch_bounds__(256) void ens_qtab(EnsQtab qt, const uint32_t *__restrict__ slot_of) {
    ens_qtab_fill(qt, slot_of);
}

// counters[0] = sweep base, counters[1] = exchange round, counters[2] = records
__global__ void ens_tick(uint32_t *__restrict__ counters, uint32_t sweeps, uint32_t rounds) {
    counters[0] += sweeps;
    counters[1] += rounds;
}


// The measurement tail of a block of sweeps in one workgroup (one launch
// where four small ones ran): the sums and beta slots into the ring, then
// (npairs > 0) one exchange round, then the sums cleared for the next
// block's ens_plaquette, the fast pass's thresholds rewritten for the new
// slots, and the counters advanced as ens_tick does.  Barriers order the
// phases: the record reads the slots before the exchange swaps them, the
// thresholds read them after, and every thread reads the counters before
// thread 0 moves them.
__global__ __launch_bounds__(256) void ens_post(unsigned long long *__restrict__ sums, uint32_t *__restrict__ slot_of,
                                                int nrep, long long *__restrict__ ring_sums,
                                                uint32_t *__restrict__ ring_slots, uint32_t cap,
                                                uint32_t *__restrict__ counters, const EnsGroupPair *__restrict__ pairs,
                                                int npairs, uint32_t *__restrict__ rep_at,
                                                const uint32_t *__restrict__ ex_tables, uint32_t ex_seed,
                                                unsigned long long *__restrict__ accepted, uint32_t sweeps,
                                                EnsQtab qt) {
    const uint32_t slot = counters[2] % cap;
    const uint32_t round = counters[1];
    for (int i = threadIdx.x; i < nrep; i += blockDim.x) {
        ring_sums[(size_t)slot * nrep + i] = (long long)sums[i];
        ring_slots[(size_t)slot * nrep + i] = slot_of[i];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < npairs; p += blockDim.x)
        ens_exchange_pair(p, round, pairs, slot_of, rep_at, sums, ex_tables, ex_seed, accepted);
    __syncthreads();
    if (npairs > 0) ens_qtab_fill(qt, slot_of);
    for (int i = threadIdx.x; i < nrep; i += blockDim.x) sums[i] = 0ull;
    if (threadIdx.x == 0) {
        counters[0] += sweeps;
        counters[1] += npairs > 0 ? 1u : 0u;
        counters[2] += 1u;
    }
}

// lattice order (link = site * D + dir) <-> byte j of the quad's words
__global__ void ens_pack(const uint8_t *__restrict__ lat, uint32_t *__restrict__ words, EnsQuad Q, int j, int D) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Q.nsites) return;
    uint32_t r = s / (uint32_t)Q.L[0];
    uint32_t x0 = s - r * (uint32_t)Q.L[0];
    uint32_t rest = r, par = x0;
    for (int a = 1; a < D; ++a) {
        par += rest % (uint32_t)Q.L[a];
        rest /= (uint32_t)Q.L[a];
    }
    const uint32_t half = Q.nsites >> 1, c = par & 1u, idx = r * (uint32_t)Q.H + (x0 >> 1);
    uint8_t *bytes = reinterpret_cast<uint8_t *>(words);
    for (int d = 0; d < D; ++d)
        bytes[((size_t)Q.woff + ((uint32_t)d * 2u + c) * half + idx) * 4u + (uint32_t)j] = lat[(size_t)s * D + d];
}

__global__ void ens_unpack(uint8_t *__restrict__ lat, const uint32_t *__restrict__ words, EnsQuad Q, int j, int D) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Q.nsites) return;
    uint32_t r = s / (uint32_t)Q.L[0];
    uint32_t x0 = s - r * (uint32_t)Q.L[0];
    uint32_t rest = r, par = x0;
    for (int a = 1; a < D; ++a) {
        par += rest % (uint32_t)Q.L[a];
        rest /= (uint32_t)Q.L[a];
    }
    const uint32_t half = Q.nsites >> 1, c = par & 1u, idx = r * (uint32_t)Q.H + (x0 >> 1);
    const uint8_t *bytes = reinterpret_cast<const uint8_t *>(words);
    for (int d = 0; d < D; ++d)
        lat[(size_t)s * D + d] = bytes[((size_t)Q.woff + ((uint32_t)d * 2u + c) * half + idx) * 4u + (uint32_t)j];
}

void launch_ens_pack(const uint8_t *lat, uint32_t *words, const EnsQuad &Q, int j, int D, hipStream_t s) {
    ens_pack<<<dim3((Q.nsites + 255) / 256), dim3(256), 0, s>>>(lat, words, Q, j, D);
}

void launch_ens_unpack(uint8_t *lat, const uint32_t *words, const EnsQuad &Q, int j, int D, hipStream_t s) {
    ens_unpack<<<dim3((Q.nsites + 255) / 256), dim3(256), 0, s>>>(lat, words, Q, j, D);
}

template <int D, int MODE>
static void ens_pass_mu(const EnsPassArgs &a) {
    dim3 grid(a.blocks), block(a.block_size);
#define PLQ_ENS_MU(M)                                                                                             \
    if (a.fast)                                                                                                   \
        ens_link_pass2<D, MODE, M><<<grid, block, 0, a.stream>>>(a.words, a.quads, a.block_quad, a.qtab,         \
                                                                 a.sweep_base, a.sweep_off, a.c, a.per_lane);     \
    else                                                                                                          \
        ens_link_pass<D, MODE, M><<<grid, block, 0, a.stream>>>(a.words, a.quads, a.block_quad, a.tables,        \
                                                                a.beta_slot, a.sweep_base, a.sweep_off, a.c,      \
                                                                a.per_lane)
    switch (a.mu) {
        case 0: PLQ_ENS_MU(0); break;
        case 1: PLQ_ENS_MU(1); break;
        case 2: PLQ_ENS_MU(2 < D ? 2 : 0); break;
        default: PLQ_ENS_MU(3 < D ? 3 : 0); break;
    }
#undef PLQ_ENS_MU
}

template <int D, int MODE>
static hipError_t ens_coop(const EnsPassArgs &a, uint32_t nsweeps) {
    int bpc = 0, cus = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, ens_sweeps_coop<D, MODE>, a.block_size, 0);
    if (e != hipSuccess) return e;
    if (bpc < 1) return hipErrorInvalidConfiguration;
    // one block per CU: a grid barrier's cost grows with the blocks it joins
    // (measured 1.2 ms per sweep with every resident block, against 0.3 ms)
    uint32_t grid = (uint32_t)cus;
    if (grid > a.blocks) grid = a.blocks;
    uint32_t *words = a.words;
    const EnsQuad *quads = a.quads;
    const uint16_t *block_quad = a.block_quad;
    const uint32_t *qtab = a.qtab, *sweep_base = a.sweep_base;
    int per_lane = a.per_lane;
    uint32_t nblocks = a.blocks;
    void *args[] = {&words, &quads, &block_quad, &qtab, &sweep_base, &nsweeps, &per_lane, &nblocks};
    return hipLaunchCooperativeKernel((const void *)ens_sweeps_coop<D, MODE>, dim3(grid), dim3(a.block_size), args, 0,
                                      a.stream);
}

hipError_t launch_ens_sweeps_coop(const EnsPassArgs &a, uint32_t nsweeps) {
    switch (a.D * 2 + a.mode) {
        case 4: return ens_coop<2, 0>(a, nsweeps);
        case 5: return ens_coop<2, 1>(a, nsweeps);
        case 6: return ens_coop<3, 0>(a, nsweeps);
        case 7: return ens_coop<3, 1>(a, nsweeps);
        case 8: return ens_coop<4, 0>(a, nsweeps);
        default: return ens_coop<4, 1>(a, nsweeps);
    }
}

void launch_ens_pass(const EnsPassArgs &a) {
    switch (a.D * 2 + a.mode) {
        case 4: ens_pass_mu<2, 0>(a); break;
        case 5: ens_pass_mu<2, 1>(a); break;
        case 6: ens_pass_mu<3, 0>(a); break;
        case 7: ens_pass_mu<3, 1>(a); break;
        case 8: ens_pass_mu<4, 0>(a); break;
        default: ens_pass_mu<4, 1>(a); break;
    }
}

void launch_ens_zero(unsigned long long *sums, int nrep, hipStream_t s) {
    ens_zero<<<dim3((nrep + 255) / 256), dim3(256), 0, s>>>(sums, nrep);
}

void launch_ens_post(unsigned long long *sums, uint32_t *slot_of, int nrep, long long *ring_sums, uint32_t *ring_slots,
                     uint32_t cap, uint32_t *counters, const EnsGroupPair *pairs, int npairs, uint32_t *rep_at,
                     const uint32_t *ex_tables, uint32_t ex_seed, unsigned long long *accepted, uint32_t sweeps,
                     const EnsQtab &qt, hipStream_t s) {
    ens_post<<<dim3(1), dim3(256), 0, s>>>(sums, slot_of, nrep, ring_sums, ring_slots, cap, counters, pairs, npairs,
                                           rep_at, ex_tables, ex_seed, accepted, sweeps, qt);
}

void launch_ens_qtab(const EnsQtab &qt, const uint32_t *slot_of, hipStream_t s) {
    ens_qtab<<<dim3(1), dim3(256), 0, s>>>(qt, slot_of);
}

void launch_ens_plaquette(const uint32_t *words, const EnsQuad *quads, const uint16_t *block_quad, int nrep, int D,
                          uint32_t blocks, int block_size, int per_lane, unsigned long long *sums, hipStream_t s,
                          bool zero_first) {
    if (zero_first) launch_ens_zero(sums, nrep, s);
    dim3 grid(blocks), block(block_size);
    switch (D) {
        case 2: ens_plaquette<2><<<grid, block, 0, s>>>(words, quads, block_quad, sums, per_lane); break;
        case 3: ens_plaquette<3><<<grid, block, 0, s>>>(words, quads, block_quad, sums, per_lane); break;
        default: ens_plaquette<4><<<grid, block, 0, s>>>(words, quads, block_quad, sums, per_lane); break;
    }
}


void launch_ens_tick(uint32_t *counters, uint32_t sweeps, uint32_t rounds, hipStream_t s) {
    ens_tick<<<dim3(1), dim3(1), 0, s>>>(counters, sweeps, rounds);
}


}  // namespace plq
