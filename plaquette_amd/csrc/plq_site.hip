// plq_site.hip -- checkerboard kernels for site fields (Ising, Potts) on gfx950.
//
// One launch is one colour pass of the reference's checkerboard
// (site_schedule.wgsl:15-31, updater.rs:316-324): every site of colour c is
// updated from its 2D neighbours, which all carry the other colour and are
// not written during the pass.  Unlike the reference, a launch covers only the
// sites of its colour: the two colours live in separate int8 arrays
// (plq_common.hpp, Geom), so a lane never loads a site it will not update and
// every neighbour index is computed from coordinates in registers instead of
// read from the 2D-entry u32 neighbour table (device.rs:741-747).
//
// Decisions: the draw m = keyed_bits(seed32, site, sweep) (rng.wgsl:47-52) is
// compared with an integer threshold of the site's class, built on the host
// from the reference arithmetic (plq_thresholds.cpp).  The draw is computed
// only when the threshold leaves the outcome open (0 < T < 2^24), which
// matches the reference's "downhill takes no draw" rule without changing any
// result, since keyed draws have no stream position.
//
// Observables are fused: on a recording sweep the colour-0 pass counts the
// new colour-0 states, and the colour-1 pass counts the colour-1 states and
// sums each colour-1 site's bond term against its (final) colour-0
// neighbours.  Every bond has exactly one colour-1 end on an even lattice, so
// that sum is the full bond sum of ising/mod.rs:97-117 and potts/mod.rs:237-271.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

// ---------------------------------------------------------------- layout ---

__global__ void pack_sites_kernel(const uint8_t *__restrict__ lat, uint8_t *__restrict__ c0,
                                  uint8_t *__restrict__ c1, Geom g) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    int32_t r = (int32_t)(s / g.L[0]);
    int32_t x0 = (int32_t)(s - (int64_t)r * g.L[0]);
    int32_t x[MAX_D];
    row_coords(g, r, x);
    int32_t par = x0;
    for (int mu = 1; mu < MAX_D; ++mu) par += x[mu];
    int64_t idx = (int64_t)r * g.H + (x0 >> 1);
    if (par & 1) c1[idx] = lat[s];
    else c0[idx] = lat[s];
}

__global__ void unpack_sites_kernel(uint8_t *__restrict__ lat, const uint8_t *__restrict__ c0,
                                    const uint8_t *__restrict__ c1, Geom g) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    int32_t r = (int32_t)(s / g.L[0]);
    int32_t x0 = (int32_t)(s - (int64_t)r * g.L[0]);
    int32_t x[MAX_D];
    row_coords(g, r, x);
    int32_t par = x0;
    for (int mu = 1; mu < MAX_D; ++mu) par += x[mu];
    int64_t idx = (int64_t)r * g.H + (x0 >> 1);
    lat[s] = (par & 1) ? c1[idx] : c0[idx];
}

// -------------------------------------------------------------- reduction ---

// Sum `v` over the block and let thread 0 return it.  Block of up to 1024.
__device__ __forceinline__ long long block_sum(long long v, long long *lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    long long t = 0;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) t += lds[w];
    }
    return t;
}

// Logical block index: hardware dispatches consecutive blocks round-robin over
// the 8 XCDs; remapping gives each XCD a contiguous 1/8 of the rows so the
// neighbour rows a block reads sit in the same XCD's L2.
__device__ __forceinline__ int64_t logical_block() {
    uint32_t b = blockIdx.x, G = gridDim.x;
    if ((G & 7u) != 0) return b;
    return (int64_t)(b & 7u) * (G >> 3) + (b >> 3);
}

// Write this block's partial sums of a record, column-major: column k of
// block b at part[k * gridDim.x + b], columns [bond, 0, counts[0..q)].
// finalize_columns_kernel folds both colour passes' columns into the record.
// Every lane of the block calls it, at the kernel's end: a DPP sum per wave
// and value (each lane's values, and a wave's sums, fit in an int), one
// barrier, then thread k adds value k over the waves.
__device__ __forceinline__ void emit_partials(const MeasArgs &m, long long bond, const int *counts, int q) {
    __shared__ int red[16][1 + MAX_Q_DEV];
    const int64_t G = gridDim.x;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const int s0 = wave_sum_i32((int)bond);
    if (lane == 0) red[wave][0] = s0;
    for (int a = 0; a < q; ++a) {
        const int s = wave_sum_i32(counts[a]);
        if (lane == 0) red[wave][1 + a] = s;
    }
    __syncthreads();
    const int k = (int)threadIdx.x;
    if (k <= q) {
        long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w][k];
        m.part[(k == 0 ? 0 : 1 + k) * G + blockIdx.x] = t;
    } else if (k == q + 1) {
        m.part[G + blockIdx.x] = 0;
    }
}

// Fold the column-major partials of the two colour passes (na and nb blocks)
// into a record: one workgroup per record column, each column a contiguous
// run read coalesced; columns >= nk are cleared.
// With ctr non-null the record goes to slot ctr[1] of a ring of cap records
// at rec (a replayed graph's record).
__global__ __launch_bounds__(256) void finalize_columns_kernel(const long long *__restrict__ pa, int64_t na,
                                                               const long long *__restrict__ pb, int64_t nb, int nk,
                                                               long long *__restrict__ rec,
                                                               const uint32_t *__restrict__ ctr) {
    __shared__ long long lds[16];
    if (ctr) rec += (size_t)ctr[1] * REC_K;
    const int k = blockIdx.x;
    if (k >= nk) {
        if (threadIdx.x == 0) rec[k] = 0;
        return;
    }
    // eight loads in flight per lane: a loop of dependent steps would wait
    // out each load's latency in turn
    auto col = [](const long long *__restrict__ src, int64_t n) -> long long {
        long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int64_t b = threadIdx.x;
        for (; b + 7 * 256 < n; b += 8 * 256)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += src[b + u * 256];
        for (; b < n; b += 256) acc[0] += src[b];
        return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    };
    const long long v = col(pa + k * na, na) + col(pb + k * nb, nb);
    long long t = block_sum(v, lds);
    if (threadIdx.x == 0) rec[k] = t;
}

// Sum two runs of block partials (one per colour pass) into one record: each
// lane accumulates whole rows, then one wave-and-block reduction of all REC_K
// columns at once.
__global__ __launch_bounds__(1024) void finalize_record_kernel(const long long *__restrict__ pa, int64_t na,
                                                               const long long *__restrict__ pb, int64_t nb,
                                                               long long *__restrict__ rec) {
    __shared__ long long acc[16][REC_K];
    long long v[REC_K];
#pragma unroll
    for (int k = 0; k < REC_K; ++k) v[k] = 0;
    for (int64_t b = threadIdx.x; b < na + nb; b += blockDim.x) {
        const long long *row = b < na ? pa + b * REC_K : pb + (b - na) * REC_K;
#pragma unroll
        for (int k = 0; k < REC_K; ++k) v[k] += row[k];
    }
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < REC_K; ++k) {
        long long t = v[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane == 0) acc[wave][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < REC_K) {
        long long t = 0;
        int nw = (int)(blockDim.x >> 6);
        for (int w = 0; w < nw; ++w) t += acc[w][threadIdx.x];
        rec[threadIdx.x] = t;
    }
}

// ------------------------------------------------------- generic site pass ---

// One site per lane, any D <= 4, any even extents, Ising or Potts.
// MODE: 0 Metropolis, 1 heat bath, 2 measure only (no update).
template <int D, int MODEL, int MODE>
__global__ __launch_bounds__(256) void site_pass_generic(const uint8_t *__restrict__ oth, uint8_t *__restrict__ own,
                                                         Geom g, int c, uint32_t seed, uint32_t skey, int q,
                                                         const uint32_t *__restrict__ table,
                                                         MeasArgs m,
                                                         int64_t work, int64_t span) {
    constexpr int NN = 2 * D;
    const int64_t lb = logical_block();
    const int64_t begin = lb * span;
    int64_t end = begin + span;
    if (end > work) end = work;
    long long bond = 0;
    int counts[MAX_Q_DEV];
    for (int a = 0; a < MAX_Q_DEV; ++a) counts[a] = 0;
    const int L0 = g.L[0];
    for (int64_t w = begin + threadIdx.x; w < end; w += blockDim.x) {
        int32_t r = (int32_t)(w / g.H);
        int32_t i = (int32_t)(w - (int64_t)r * g.H);
        int32_t x[MAX_D];
        row_coords(g, r, x);
        int32_t prs = 0;
        for (int mu = 1; mu < D; ++mu) prs += x[mu];
        int32_t p = (prs + c) & 1;
        int32_t x0 = 2 * i + p;
        uint32_t site = (uint32_t)((int64_t)r * L0 + x0);
        int64_t rowbase = (int64_t)r * g.H;
        uint8_t nb[NN];
        {
            int32_t xp = (x0 + 1 == L0) ? 0 : x0 + 1;
            int32_t xm = (x0 == 0) ? L0 - 1 : x0 - 1;
            nb[0] = oth[rowbase + (xp >> 1)];
            nb[1] = oth[rowbase + (xm >> 1)];
#pragma unroll
            for (int mu = 1; mu < D; ++mu) {
                nb[2 * mu] = oth[(int64_t)row_step(g, r, x, mu, +1) * g.H + i];
                nb[2 * mu + 1] = oth[(int64_t)row_step(g, r, x, mu, -1) * g.H + i];
            }
        }
        int cur = own[rowbase + i];
        int nxt = cur;
        if (MODEL == 0) {  // Ising: class (cur, d)
            int d = 0;
#pragma unroll
            for (int k = 0; k < NN; ++k) d += nb[k];
            if (MODE != 2) {
                uint32_t t = table[cur * (NN + 1) + d];
                bool hit;
                if (t == 0u) hit = false;
                else if (t >= DRAW_ONE) hit = true;
                else hit = keyed_bits_k(seed, site, skey) < t;
                if (MODE == 0) nxt = hit ? 1 - cur : cur;
                else nxt = hit ? 0 : 1;
                own[rowbase + i] = (uint8_t)nxt;
            }
            if (m.on) {
                counts[nxt] += 1;
                if (c == 1) bond += (long long)((1 - 2 * nxt) * (NN - 2 * d));
            }
        } else {  // Potts
            int n[MAX_Q_DEV];
            for (int a = 0; a < q; ++a) n[a] = 0;
#pragma unroll
            for (int k = 0; k < NN; ++k) n[nb[k]] += 1;
            if (MODE == 0) {
                int prop;
                if (q == 2) {
                    prop = 1 - cur;
                } else {
                    uint32_t alternatives = (uint32_t)(q - 1);
                    float u = keyed_uniform_f32(keyed_bits_k(seed ^ PROPOSAL_KEY, site, skey));
                    uint32_t dr = (uint32_t)(u * (float)alternatives);
                    if (dr >= alternatives) dr = alternatives - 1;
                    prop = (int)dr + ((int)dr >= cur ? 1 : 0);
                }
                int change = n[prop] - n[cur];
                uint32_t t = table[((uint32_t)cur * q + prop) * (2 * NN + 1) + (change + NN)];
                bool hit;
                if (t == 0u) hit = false;
                else if (t >= DRAW_ONE) hit = true;
                else hit = keyed_bits_k(seed, site, skey) < t;
                nxt = hit ? prop : cur;
                own[rowbase + i] = (uint8_t)nxt;
            } else if (MODE == 1) {
                uint32_t enc = 0, base = 1;
                for (int a = 0; a < q - 1; ++a) {
                    enc += (uint32_t)n[a] * base;
                    base *= (uint32_t)(NN + 1);
                }
                uint32_t cls = (uint32_t)cur * base + enc;
                const uint32_t *T = table + (int64_t)cls * (q - 1);
                uint32_t m = keyed_bits_k(seed, site, skey);
                nxt = q - 1;
                for (int s = 0; s < q - 1; ++s)
                    if (m < T[s]) {
                        nxt = s;
                        break;
                    }
                own[rowbase + i] = (uint8_t)nxt;
            }
            if (m.on) {
                counts[nxt] += 1;
                if (c == 1) bond += (long long)n[nxt];
            }
        }
    }
    if (m.on) emit_partials(m, bond, counts, q);
}

// ------------------------------------------------------ SWAR Ising pass ---

// Four same-colour sites per lane (one dword of a colour row).  Requires
// power-of-two extents with L0 >= 8.  State bytes are 0/1, so the per-byte
// count of "down" neighbours is the plain dword sum of the 2D neighbour words
// (at most 8 per byte, no carries between bytes).
template <int D, int MODE>
__global__ __launch_bounds__(256) void ising_pass_swar(const uint32_t *__restrict__ oth, uint32_t *__restrict__ own,
                                                       SwarGeom sg, int c, uint32_t seed, uint32_t skey,
                                                       const uint32_t *__restrict__ table,
                                                       MeasArgs m, int64_t work,
                                                       int64_t span) {
    constexpr int NN = 2 * D;
    constexpr int K = NN + 1;
    __shared__ uint32_t T[2 * K];
    if (threadIdx.x < 2 * K) T[threadIdx.x] = table[threadIdx.x];
    __syncthreads();
    // 32-bit offsets throughout (N / 8 dwords < 2^32): every load is an SGPR
    // base plus a 32-bit VGPR offset, with no 64-bit address arithmetic.
    const uint32_t lb = (uint32_t)logical_block();
    const uint32_t begin = lb * (uint32_t)span;
    uint32_t end = begin + (uint32_t)span;
    if (end > (uint32_t)work) end = (uint32_t)work;
    int bond = 0;
    int down = 0, total = 0;
    const uint32_t h4mask = (1u << sg.lh4) - 1u;
    for (uint32_t w = begin + threadIdx.x; w < end; w += blockDim.x) {
        uint32_t r = w >> sg.lh4;
        uint32_t i4 = w & h4mask;
        uint32_t x1 = r & (sg.L1 - 1u);
        uint32_t x2 = 0, x3 = 0;
        if (D >= 3) x2 = (r >> sg.ll1) & (sg.L2 - 1u);
        if (D >= 4) x3 = r >> (sg.ll1 + sg.ll2);
        uint32_t prs = x1 + x2 + x3;
        uint32_t p = (prs + (uint32_t)c) & 1u;
        const uint32_t rb = r << sg.lh4;
        // The along-axis-0 neighbours: the other colour's dword at the same
        // position, and the same bytes shifted one site forward (p = 1) or
        // back (p = 0), borrowing a byte from the adjacent dword of the row
        // (wrapping at the row ends).  Branch-free: p alternates row by row.
        uint32_t a0 = ld32(oth, rb + i4);
        uint32_t adj = ld32(oth, rb + ((i4 + (p ? 1u : h4mask)) & h4mask));
        uint32_t sh = p ? ((a0 >> 8) | (adj << 24)) : ((a0 << 8) | (adj >> 24));
        uint32_t dsum = a0 + sh;
        if (D >= 2) {
            uint32_t rp = (x1 + 1u == sg.L1) ? r - (sg.L1 - 1u) : r + 1u;
            uint32_t rm = (x1 == 0u) ? r + (sg.L1 - 1u) : r - 1u;
            dsum += ld32(oth, (rp << sg.lh4) + i4) + ld32(oth, (rm << sg.lh4) + i4);
        }
        if (D >= 3) {
            uint32_t st = sg.L1;
            uint32_t span2 = (sg.L2 - 1u) * st;
            uint32_t rp = (x2 + 1u == sg.L2) ? r - span2 : r + st;
            uint32_t rm = (x2 == 0u) ? r + span2 : r - st;
            dsum += ld32(oth, (rp << sg.lh4) + i4) + ld32(oth, (rm << sg.lh4) + i4);
        }
        if (D >= 4) {
            uint32_t st = sg.L1 * sg.L2;
            uint32_t span3 = (sg.L3 - 1u) * st;
            uint32_t rp = (x3 + 1u == sg.L3) ? r - span3 : r + st;
            uint32_t rm = (x3 == 0u) ? r + span3 : r - st;
            dsum += ld32(oth, (rp << sg.lh4) + i4) + ld32(oth, (rm << sg.lh4) + i4);
        }
        uint32_t s4 = ld32((const uint32_t *)own, rb + i4);
        uint32_t out = s4;
        if (MODE != 2) {
            out = 0;
            const uint32_t site0 = (r << (sg.lh4 + 3)) + (i4 << 3) + p;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t s = (s4 >> (8 * k)) & 1u;
                uint32_t d = (dsum >> (8 * k)) & 0xffu;
                uint32_t t = T[s * K + d];
                bool hit;
                if (t == 0u) hit = false;
                else if (t >= DRAW_ONE) hit = true;
                else hit = keyed_bits_k(seed, site0 + 2u * k, skey) < t;
                uint32_t nxt = (MODE == 0) ? (s ^ (hit ? 1u : 0u)) : (hit ? 0u : 1u);
                out |= nxt << (8 * k);
            }
            st32(own, rb + i4, out);
        }
        if (m.on) {
            down += __popc(out);
            total += 4;
            if (c == 1) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    int s = (int)((out >> (8 * k)) & 1u);
                    int d = (int)((dsum >> (8 * k)) & 0xffu);
                    bond += (1 - 2 * s) * (NN - 2 * d);
                }
            }
        }
    }
    if (m.on) {
        int counts[2] = {total - down, down};
        emit_partials(m, (long long)bond, counts, 2);
    }
}

// ----------------------------------------------------- 16-site Ising pass ---


// Spread the low four bits of x to bit 0 of each byte: bit k of x lands at
// bit 8k of x * (1 + 2^7 + 2^14 + 2^21), and no other partial product falls
// on a multiple of 8 or collides with another.
__device__ __forceinline__ uint32_t spread4(uint32_t x) { return (x * 0x00204081u) & 0x01010101u; }

// Sixteen same-colour sites per lane: one 16-byte chunk of a colour row.
// Requires power-of-two extents with L0 >= 32.  Per site the work is one
// LDS threshold read, the keyed draw and one compare; everything else is
// done per dword (4 sites) or per chunk (16 sites):
//   - the "down" neighbour counts of 4 sites are one dword sum;
//   - the class byte of each site, scaled to an LDS byte address, is built
//     for 4 sites at once ((dsum << 3) | (state << 2) indexes a copy of the
//     table reordered to [d][state]);
//   - the site indices of a chunk are site0 | 2k (site0 has bits 1..4 clear),
//     so the first xor-shift of the first hash round, (x ^ x >> 16) with
//     x = seed ^ site, is computed once per chunk and per site only the
//     constant 2k is xored in;
//   - the 16 decisions are collected as bits and spread back to bytes with
//     one multiply per dword.
// Everything a 16-site chunk reads, loaded ahead of its use so that a lane
// computes one chunk while the next chunk's loads are in flight.
template <int D>
struct X16Chunk {
    uint4 a;                        // the other colour's 16 bytes at this position
    uint4 s;                        // this colour's 16 states
    uint4 rows[D > 1 ? 2 * (D - 1) : 1];  // the other colour's bytes in the neighbour rows
    uint32_t adj;                   // the other colour's dword the axis-0 shift pulls a byte from
    uint32_t site0;                 // lattice index of the chunk's first site
    uint32_t p;                     // x0 parity of this colour in the row
};

template <int D>
__device__ __forceinline__ void x16_load(const uint4 *__restrict__ oth, const uint4 *__restrict__ own,
                                         const SwarGeom &sg, uint32_t c, uint32_t lh16, uint32_t cmask, uint32_t w,
                                         X16Chunk<D> &k) {
    const uint32_t r = w >> lh16;
    const uint32_t ch = w & cmask;
    const uint32_t x1 = r & (sg.L1 - 1u);
    uint32_t x2 = 0, x3 = 0;
    if (D >= 3) x2 = (r >> sg.ll1) & (sg.L2 - 1u);
    if (D >= 4) x3 = r >> (sg.ll1 + sg.ll2);
    const uint32_t p = (x1 + x2 + x3 + c) & 1u;
    const uint32_t rc = r << lh16;  // chunk index of the row start
    k.p = p;
    k.site0 = (r << (sg.lh4 + 3u)) + (ch << 5) + p;
    // p = 1 needs the first byte of the next chunk, p = 0 the last byte of the
    // previous one (wrapping at the row ends).  Computed without a branch: a
    // branch here let the compiler wait for the first load inside it, before
    // the other loads were issued.
    const uint32_t nch = (ch + (p ? 1u : cmask)) & cmask;
    const uint32_t adj_dw = ((rc + nch) << 2) + (p ^ 1u) * 3u;
    k.a = ld32(oth, w);
    k.adj = ld32(reinterpret_cast<const uint32_t *>(oth), adj_dw);
    if (D >= 2) {
        k.rows[0] = ld32(oth, ((((x1 + 1u == sg.L1) ? r - (sg.L1 - 1u) : r + 1u)) << lh16) + ch);
        k.rows[1] = ld32(oth, ((((x1 == 0u) ? r + (sg.L1 - 1u) : r - 1u)) << lh16) + ch);
    }
    if (D >= 3) {
        const uint32_t st = sg.L1, span2 = (sg.L2 - 1u) * sg.L1;
        k.rows[2] = ld32(oth, (((x2 + 1u == sg.L2) ? r - span2 : r + st) << lh16) + ch);
        k.rows[3] = ld32(oth, (((x2 == 0u) ? r + span2 : r - st) << lh16) + ch);
    }
    if (D >= 4) {
        const uint32_t st = sg.L1 * sg.L2, span3 = (sg.L3 - 1u) * st;
        k.rows[4] = ld32(oth, (((x3 + 1u == sg.L3) ? r - span3 : r + st) << lh16) + ch);
        k.rows[5] = ld32(oth, (((x3 == 0u) ? r + span3 : r - st) << lh16) + ch);
    }
    k.s = ld32(own, w);
}

// Sixteen same-colour sites per lane: one 16-byte chunk of a colour row.
// Requires power-of-two extents with L0 >= 32.  Per site the work is one
// LDS threshold read, the keyed draw and one compare; everything else is
// done per dword (4 sites) or per chunk (16 sites):
//   - the "down" neighbour counts of 4 sites are one dword sum;
//   - the class byte of each site, scaled to an LDS byte address, is built
//     for 4 sites at once ((dsum << 3) | (state << 2) indexes a copy of the
//     table reordered to [d][state]);
//   - the site indices of a chunk are site0 | 2k (site0 has bits 1..4 clear),
//     so the first xor-shift of the first hash round, (x ^ x >> 16) with
//     x = seed ^ site, is computed once per chunk and per site only the
//     constant 2k is xored in;
//   - the 16 decisions are collected as bits and spread back to bytes with
//     one multiply per dword.
// A lane walks its block's chunks with a stride of blockDim.x.  (Loading the
// next chunk ahead of the current one's compute was measured slower: it
// raises the register count from ~50 to 83 and occupancy from 8 to 5 waves,
// and the pass is bound by VALU issue, not by load latency.)
template <int D, int MODE>
__global__ __launch_bounds__(256) void ising_pass_x16(const uint4 *__restrict__ oth, uint4 *__restrict__ own,
                                                      SwarGeom sg, int c, uint32_t seed, uint32_t skey,
                                                      const uint32_t *__restrict__ table, MeasArgs m, int64_t work,
                                                      int64_t span) {
    constexpr int NN = 2 * D;
    constexpr int K = NN + 1;
    __shared__ uint32_t T[2 * 16];  // T[d * 2 + s], byte address (d << 3) | (s << 2)
    if (threadIdx.x < 2 * K) {
        int s = threadIdx.x / K, d = threadIdx.x % K;
        T[d * 2 + s] = table[threadIdx.x];
    }
    __syncthreads();
    const uint32_t lb = (uint32_t)logical_block();
    const uint32_t begin = lb * (uint32_t)span;
    uint32_t end = begin + (uint32_t)span;
    if (end > (uint32_t)work) end = (uint32_t)work;
    int bond = 0;
    int down = 0, total = 0;
    const uint32_t lh16 = sg.lh4 - 2u;  // log2(chunks per colour row)
    const uint32_t cmask = (1u << lh16) - 1u;
    const char *Tb = reinterpret_cast<const char *>(T);
    const uint32_t uc = (uint32_t)c;
    for (uint32_t w = begin + threadIdx.x; w < end; w += blockDim.x) {
        X16Chunk<D> cur;
        x16_load<D>(oth, own, sg, uc, lh16, cmask, w, cur);
        // The draws depend on the site and the sweep only, not on the
        // lattice: all sixteen are computed here, between issuing the loads
        // and the first use of their data, so the load latency is covered by
        // about 60% of the chunk's VALU work.  The scheduling barriers keep
        // the compiler from sinking the loads below the hashes or hoisting
        // the waits above them.
        uint32_t mm[16];
        const uint32_t kmix = sweep_mix(skey);  // keyed_bits_k's merged inter-round xor-shift
        __builtin_amdgcn_sched_barrier(0);
        if (MODE != 2) {
            const uint32_t bx = seed ^ cur.site0;
            const uint32_t v0 = bx ^ (bx >> 16);
#pragma unroll
            for (int k = 0; k < 16; ++k) mm[k] = lowbias32_tail(lowbias32_mid(v0 ^ (2u * (uint32_t)k)) ^ kmix) >> 8;
        }
        __builtin_amdgcn_sched_barrier(0);
        const uint4 a = cur.a;
        const uint32_t p = cur.p;
        // axis-0 neighbours: the same bytes shifted one site forward (p = 1)
        // or back (p = 0); selected without branching since p alternates
        // from row to row inside a wave
        const uint32_t sel = p ? 1u : 3u;
        uint4 ds;
        ds.x = a.x + __builtin_amdgcn_alignbyte(p ? a.y : a.x, p ? a.x : cur.adj, sel);
        ds.y = a.y + __builtin_amdgcn_alignbyte(p ? a.z : a.y, p ? a.y : a.x, sel);
        ds.z = a.z + __builtin_amdgcn_alignbyte(p ? a.w : a.z, p ? a.z : a.y, sel);
        ds.w = a.w + __builtin_amdgcn_alignbyte(p ? cur.adj : a.w, p ? a.w : a.z, sel);
#pragma unroll
        for (int j = 0; j < (D > 1 ? 2 * (D - 1) : 0); ++j) {
            ds.x += cur.rows[j].x;
            ds.y += cur.rows[j].y;
            ds.z += cur.rows[j].z;
            ds.w += cur.rows[j].w;
        }
        const uint4 s = cur.s;
        uint4 out = s;
        if (MODE != 2) {
            // m < t <=> m - t is negative (both below 2^31); the four sign
            // bytes are gathered into one dword with two byte permutes, and
            // their sign bits become the per-site hit bytes.
            auto decide4 = [&](uint32_t sw, uint32_t dw, uint32_t kbase) -> uint32_t {
                const uint32_t addr4 = (dw << 3) | (sw << 2);
                uint32_t diff[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t addr = (addr4 >> (8 * k)) & 0xffu;
                    uint32_t t = *reinterpret_cast<const uint32_t *>(Tb + addr);
                    diff[k] = mm[kbase + k] - t;
                }
                // byte 3 of diff[0..3] -> bytes 0..3
                uint32_t lo = __builtin_amdgcn_perm(diff[1], diff[0], 0x0c0c0703u);
                uint32_t hi = __builtin_amdgcn_perm(diff[3], diff[2], 0x07030c0cu);
                uint32_t bits = ((lo | hi) >> 7) & 0x01010101u;
                return MODE == 0 ? (sw ^ bits) : (bits ^ 0x01010101u);
            };
            out.x = decide4(s.x, ds.x, 0);
            out.y = decide4(s.y, ds.y, 4);
            out.z = decide4(s.z, ds.z, 8);
            out.w = decide4(s.w, ds.w, 12);
            st32(own, w, out);
        }
        if (m.on) {
            down += __popc(out.x) + __popc(out.y) + __popc(out.z) + __popc(out.w);
            total += 16;
            if (c == 1) {
                uint32_t o4[4] = {out.x, out.y, out.z, out.w};
                uint32_t d4[4] = {ds.x, ds.y, ds.z, ds.w};
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        int sn = (int)((o4[j] >> (8 * k)) & 1u);
                        int dn = (int)((d4[j] >> (8 * k)) & 0xffu);
                        bond += (1 - 2 * sn) * (NN - 2 * dn);
                    }
            }
        }
    }
    if (m.on) {
        int counts[2] = {total - down, down};
        emit_partials(m, (long long)bond, counts, 2);
    }
}

// Byte-offset access: a uniform base plus a 32-bit byte offset.
template <class T>
__device__ __forceinline__ T ldb(const T *base, uint32_t off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <class T>
__device__ __forceinline__ void stb(T *base, uint32_t off, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// (x & m) | (y & ~m): one v_bfi_b32.
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t x, uint32_t y) { return (x & m) | (y & ~m); }

// Offsets of one wave's chunk rows that do not depend on the lane: the
// wave's rows share x2 and x3, so their axis-2 and axis-3 neighbours are a
// uniform delta from the row's own offset.
struct X16UWave {
    uint32_t base;      // byte offset of the wave's first row
    uint32_t dz[2];     // axis-2 successor / predecessor deltas (mod 2^32)
    uint32_t dw[2];     // axis-3 successor / predecessor deltas
};

// One 16-site chunk of ising_pass_x16u whose rows have axis-0 colour offset
// P (wave-uniform).  A is the chunk's byte offset in a colour array.
template <int D, int MODE, int P>
__device__ __forceinline__ void x16u_chunk(const uint4 *__restrict__ oth, uint4 *__restrict__ own, const X16UGeom &ug,
                                           const X16UWave &wv, uint32_t A, uint32_t seed, uint32_t kmix,
                                           const char *Tb, int c, int on, int &bond, int &down, int &total) {
    constexpr int NN = 2 * D;
    constexpr int NR = 2 * (D - 1);
    // Axis-1 neighbours: the x1 field of the byte offset replaced by its
    // periodic successor or predecessor (one add and one bit-field insert);
    // axes 2 and 3: the wave's uniform deltas.
    const uint32_t S1 = 1u << ug.sh, M1 = ug.m1 << ug.sh;
    const uint32_t cm = (1u << (ug.lh16 + 4)) - 16u;  // chunk field of the offset
    uint32_t rowoff[NR];
    rowoff[0] = bfi(M1, A + S1, A);
    rowoff[1] = bfi(M1, A - S1, A);
    if (D > 2) {
        rowoff[2] = A + wv.dz[0];
        rowoff[3] = A + wv.dz[1];
    }
    if (D > 3) {
        rowoff[4] = A + wv.dw[0];
        rowoff[5] = A + wv.dw[1];
    }
    // The byte the axis-0 shift pulls in: the first byte of the next chunk
    // (P = 1) or the last byte of the previous one (P = 0), along the row.
    const uint32_t adjoff = P ? bfi(cm, A + 16u, A) : bfi(cm, A - 16u, A) + 12u;
    // Even waves of a block issue their loads first and compute the sixteen
    // draws while they are in flight; odd waves compute the draws first and
    // then load, which spreads the pass's read burst over time (measured
    // about 1% faster than every wave loading first).
    uint4 a, s, rows[NR];
    uint32_t adj;
    auto load = [&]() {
        a = ldb(oth, A);
        adj = ldb(reinterpret_cast<const uint32_t *>(oth), adjoff);
#pragma unroll
        for (int k = 0; k < NR; ++k) rows[k] = ldb(oth, rowoff[k]);
        s = ldb(own, A);
    };
    const bool late = (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) & 1u) != 0;
    if (!late) load();
    // The draws (site0 = 2A + P has bits 1..4 clear); y = the final hash
    // before its 8-bit shift, so that m < T  <=>  y <= T * 256 - 1.
    //
    // Site k hashes v0 ^ 2k, v0 = x ^ (x >> 16) with x = seed ^ (2A + P).  Bits
    // 1..4 of v0, b, are bits 1..4 of seed ^ (seed >> 16) xor bits 16..19 of
    // A, and every lane of the wave has the same A >> 16 (its rows differ
    // below bit 16 - sh), so b is wave-uniform.  Then
    // (v0 ^ 2k) * C1 = (v0 & ~30) * C1 + (b ^ k) * 2 C1 (mod 2^32): the first
    // multiply of every site is one add of a wave-uniform offset.
    uint32_t y[16];
    __builtin_amdgcn_sched_barrier(0);
    {
        const uint32_t bx = seed ^ (2u * A + (uint32_t)P);
        const uint32_t v0 = bx ^ (bx >> 16);
        const uint32_t b = __builtin_amdgcn_readfirstlane(((seed >> 1) ^ (seed >> 17) ^ (wv.base >> 16)) & 15u);
        const uint32_t q0 = (v0 & ~30u) * 0x7feb352du;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t t = q0 + (b ^ (uint32_t)k) * 0xffd66a5au;  // 2 * 0x7feb352d (mod 2^32)
            y[k] = lowbias32_tail(((t ^ (t >> 15)) * 0x846ca68bu) ^ kmix);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (late) load();
    // Down-neighbour counts, four sites per dword.
    uint32_t ds[4];
    const uint32_t av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t sh0 = P ? __builtin_amdgcn_alignbyte(j < 3 ? av[j + 1] : adj, av[j], 1)
                               : __builtin_amdgcn_alignbyte(av[j], j > 0 ? av[j - 1] : adj, 3);
        ds[j] = av[j] + sh0;
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        ds[0] += rows[k].x;
        ds[1] += rows[k].y;
        ds[2] += rows[k].z;
        ds[3] += rows[k].w;
    }
    const uint32_t sv[4] = {s.x, s.y, s.z, s.w};
    // Rejections as bits, site 15 first, so that site i lands in bit i: a
    // compare into VCC and an add-with-carry per site.  Class byte d + CS s
    // (d <= 2D < CS), its threshold at LDS byte 4 (d + CS s).
    constexpr int CS_LOG = D == 4 ? 4 : 3;
    uint32_t bits = 0;
#pragma unroll
    for (int j = 3; j >= 0; --j) {
        const uint32_t cls4 = (sv[j] << CS_LOG) + ds[j];
#pragma unroll
        for (int k = 3; k >= 0; --k) {
            // the class byte scaled to its LDS address; byte 0 as one SDWA op too
            // (the compiler folds (x & 0xff) << 2 into a shift and a mask)
            uint32_t addr;
            if (k == 0)
                asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
                    : "=v"(addr)
                    : "v"(2u), "v"(cls4));
            else
                addr = ((cls4 >> (8 * k)) & 0xffu) << 2;
            const uint32_t t2 = *reinterpret_cast<const uint32_t *>(Tb + addr);
            asm("v_cmp_gt_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
                : "+v"(bits)
                : "v"(y[4 * j + k]), "v"(t2)
                : "vcc");
        }
    }
    // Spread each dword's four bits to its four bytes.
    uint32_t ov[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t raw = ((bits >> (4 * j)) & 0xfu) * 0x00204081u;
        ov[j] = MODE == 0 ? (sv[j] ^ (~raw & 0x01010101u)) : (raw & 0x01010101u);
    }
    stb(own, A, make_uint4(ov[0], ov[1], ov[2], ov[3]));
    if (on) {
        const int dn = (int)(__popc(ov[0]) + __popc(ov[1]) + __popc(ov[2]) + __popc(ov[3]));
        down += dn;
        total += 16;
        if (c == 1) {
            // sum over the 16 sites of (1 - 2 s)(NN - 2 d) = 16 NN - 2 sum d - 2 NN sum s + 4 sum s d,
            // with the byte sums of d and of d where s = 1 taken by v_sad_u8
            uint32_t dsum = 0, sd = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dsum = __builtin_amdgcn_sad_u8(ds[j], 0u, dsum);
                sd = __builtin_amdgcn_sad_u8(ds[j] & (ov[j] * 0xffu), 0u, sd);
            }
            bond += 16 * NN - 2 * (int)dsum - 2 * NN * dn + 4 * (int)sd;
        }
    }
}

// Sixteen same-colour sites per lane, as ising_pass_x16, with the lanes of a
// wave on rows of one parity.  Logical chunk u (64 per wave) goes to row
// (wave >> 1) * 2K + 2 * (lane >> lh16) + (wave & 1), chunk lane & (2^lh16 - 1),
// K = 64 >> lh16 rows per wave: a pair of waves covers 2K consecutive rows,
// which lie on one axis-1 line (2K <= L1), so the axis-0 colour offset
// p = (x1 + x2 + x3 + c) & 1 is the same for the whole wave and the axis-0
// neighbour bytes are selected by a scalar branch.  Every threshold T >= 1
// (the host checks), so m < T is y <= T * 256 - 1 with y the hash before
// its final 8-bit shift, decided by one compare and one add-with-carry per
// site.  Requires D >= 2 and span a multiple of 64 (whole waves per range).
template <int D, int MODE>
__global__ __launch_bounds__(256) void ising_pass_x16u(const uint4 *__restrict__ oth, uint4 *__restrict__ own,
                                                       X16UGeom ug, int c, uint32_t seed, uint32_t skey,
                                                       const uint32_t *__restrict__ ctr, uint32_t sweep_off,
                                                       const uint32_t *__restrict__ table, MeasArgs m, uint32_t work,
                                                       uint32_t span) {
    constexpr int NN = 2 * D;
    constexpr int K = NN + 1;
    // T[d + CS s] = threshold * 256 - 1: the class stride CS exceeds the
    // largest down-neighbour count 2D (8 in 4D takes a stride of 16)
    constexpr int CS = D == 4 ? 16 : 8;
    __shared__ uint32_t T[2 * CS];
    __shared__ long long lds[16];
    if (threadIdx.x < 2 * K) {
        int s = threadIdx.x / K, d = threadIdx.x % K;
        T[d + CS * s] = (table[threadIdx.x] << 8) - 1u;
    }
    __syncthreads();
    // An earlier record, folded here instead of in a launch of its own.
    // After the table's loads: a branch on the argument before them delayed
    // every wave's start (each wave runs one chunk per lane), 6.5 -> 6.8 us.
    if (m.fold.rec) fold_record(m.fold, lds);
    const uint32_t begin = (uint32_t)logical_block() * span;
    uint32_t end = begin + span;
    if (end > work) end = work;
    // a replayed graph reads its sweep from the device counter
    const uint32_t kmix = sweep_mix(ctr ? (ctr[0] + sweep_off) * GOLDEN : skey);
    const char *Tb = reinterpret_cast<const char *>(T);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave in block
    // the lane's part of the byte offset: row 2 (lane >> lh16), chunk lane & mask
    const uint32_t lanepart = ((lane >> ug.lh16) << (ug.sh + 1)) | ((lane & ((1u << ug.lh16) - 1u)) << 4);
    const uint32_t S2 = 1u << (ug.sh + ug.ll1), S3 = 1u << (ug.sh + ug.ll1 + ug.ll2);
    int bond = 0, down = 0, total = 0;
    for (uint32_t u0 = begin; u0 < end; u0 += blockDim.x) {
        if (u0 + (wib << 6) >= end) break;  // whole waves: the wave's range is empty
        const uint32_t wave = (u0 >> 6) + wib;
        const uint32_t row0 = ((wave >> 1) << ug.lk2) + (wave & 1u);
        X16UWave wv;
        wv.base = row0 << ug.sh;
        const uint32_t x2 = (row0 >> ug.ll1) & ug.m2;
        const uint32_t x3 = row0 >> (ug.ll1 + ug.ll2);
        wv.dz[0] = x2 == ug.m2 ? 0u - ug.m2 * S2 : S2;
        wv.dz[1] = x2 == 0u ? ug.m2 * S2 : 0u - S2;
        wv.dw[0] = x3 == ug.m3 ? 0u - ug.m3 * S3 : S3;
        wv.dw[1] = x3 == 0u ? ug.m3 * S3 : 0u - S3;
        const uint32_t p = ((row0 & 1u) + (D > 2 ? x2 : 0u) + (D > 3 ? x3 : 0u) + (uint32_t)c) & 1u;
        const uint32_t A = wv.base + lanepart;
        if (p) x16u_chunk<D, MODE, 1>(oth, own, ug, wv, A, seed, kmix, Tb, c, m.on, bond, down, total);
        else x16u_chunk<D, MODE, 0>(oth, own, ug, wv, A, seed, kmix, Tb, c, m.on, bond, down, total);
    }
    if (m.on) {
        int counts[2] = {total - down, down};
        emit_partials(m, (long long)bond, counts, 2);
    }
}

// ---------------------------------------------------- fused 3D sweep ---

// Both colour passes of one 3D Ising sweep in one launch, for L0 = 256
// (eight 16-byte chunks per colour row) and L1, L2 multiples of 16.
// Workgroup b owns a tile of 16 x 16 rows (x1, x2):
//   1. it updates colour 0 on the tile and on the tile's four face halos
//      (the rows one step beyond it along x1 and x2), keeping every new
//      colour-0 row in LDS and storing the tile's own rows to n0.  A halo
//      row's update is the one its own tile makes: a site's update reads
//      only the old configuration around it and its keyed draw;
//   2. after a barrier it updates colour 1 on the tile from the new
//      colour-0 rows in LDS and its old colour-1 rows, storing to n1.
// The old arrays o0, o1 are only read, so no workgroup sees another's new
// values; the caller swaps old and new arrays after the sweep.  The bits
// are the two colour passes' (same draws, same decisions, same order of
// dependence).  Rows of a group of eight lie two apart along x1 (tile rows
// and the x2 halos: one x2, as ising_pass_x16u's waves) or along x2 (the
// x1 halos), so the axis-0 colour offset is uniform over the group.
constexpr int X16F_T = 16;               // tile rows along x1 and x2
constexpr int X16F_S = X16F_T + 2;       // LDS row slots along each axis, halos included
constexpr int X16F_GROUPS0 = 40;         // colour-0 row groups: 32 tile, 4 x2-halo, 4 x1-halo
constexpr int X16F_GROUPS1 = 32;         // colour-1 row groups: the tile's

// LDS row slot of tile-local rows (yl, zl), each in -1 .. 16.
__device__ __forceinline__ uint32_t x16f_slot(int yl, int zl) { return (uint32_t)((zl + 1) * X16F_S + (yl + 1)); }

// The sixteen draws of a chunk at colour-array byte offset A with axis-0
// colour offset P (y = the hash before its final 8-bit shift).  UNI: the
// chunk's wave has one value of A >> 16 (see x16u_chunk), so the first
// multiply of every draw is an add of a wave-uniform offset.
template <int P, bool UNI>
__device__ __forceinline__ void x16f_draws(uint32_t seed, uint32_t A, uint32_t kmix, uint32_t (&y)[16]) {
    const uint32_t bx = seed ^ (2u * A + (uint32_t)P);
    const uint32_t v0 = bx ^ (bx >> 16);
    if (UNI) {
        const uint32_t b = __builtin_amdgcn_readfirstlane(((seed >> 1) ^ (seed >> 17) ^ (A >> 16)) & 15u);
        const uint32_t q0 = (v0 & ~30u) * 0x7feb352du;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t t = q0 + (b ^ (uint32_t)k) * 0xffd66a5au;  // 2 * 0x7feb352d (mod 2^32)
            y[k] = lowbias32_tail(((t ^ (t >> 15)) * 0x846ca68bu) ^ kmix);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) y[k] = lowbias32_tail(lowbias32_mid(v0 ^ (2u * (uint32_t)k)) ^ kmix);
    }
}

// The new bytes of a chunk: a = the other colour's chunk of the same row,
// adj = its dword holding the byte the axis-0 shift pulls in, rows = the
// other colour's chunks of the four rows one step along x1 and x2, s = the
// chunk's own old bytes; as x16u_chunk's decision for D = 3.
template <int MODE, int P>
__device__ __forceinline__ uint4 x16f_decide(const uint4 &a, uint32_t adj, const uint4 (&rows)[4], const uint4 &s,
                                             const uint32_t (&y)[16], const char *Tb, uint32_t (&ds)[4]) {
    const uint32_t av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t sh0 = P ? __builtin_amdgcn_alignbyte(j < 3 ? av[j + 1] : adj, av[j], 1)
                               : __builtin_amdgcn_alignbyte(av[j], j > 0 ? av[j - 1] : adj, 3);
        ds[j] = av[j] + sh0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ds[0] += rows[k].x;
        ds[1] += rows[k].y;
        ds[2] += rows[k].z;
        ds[3] += rows[k].w;
    }
    const uint32_t sv[4] = {s.x, s.y, s.z, s.w};
    uint32_t bits = 0;
#pragma unroll
    for (int j = 3; j >= 0; --j) {
        const uint32_t cls4 = (sv[j] << 3) + ds[j];
#pragma unroll
        for (int k = 3; k >= 0; --k) {
            uint32_t addr;
            if (k == 0)
                asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
                    : "=v"(addr)
                    : "v"(2u), "v"(cls4));
            else
                addr = ((cls4 >> (8 * k)) & 0xffu) << 2;
            const uint32_t t2 = *reinterpret_cast<const uint32_t *>(Tb + addr);
            asm("v_cmp_gt_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
                : "+v"(bits)
                : "v"(y[4 * j + k]), "v"(t2)
                : "vcc");
        }
    }
    uint32_t ov[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t raw = ((bits >> (4 * j)) & 0xfu) * 0x00204081u;
        ov[j] = MODE == 0 ? (sv[j] ^ (~raw & 0x01010101u)) : (raw & 0x01010101u);
    }
    return make_uint4(ov[0], ov[1], ov[2], ov[3]);
}

// One colour-0 chunk: loads from the old arrays, the new bytes to LDS slot
// `slot` and (tile rows) to n0.
template <int MODE, int P, bool UNI>
__device__ __forceinline__ void x16f_chunk0(const uint4 *__restrict__ o0, const uint4 *__restrict__ o1,
                                            uint4 *__restrict__ n0, uint4 *R0, const X16UGeom &ug, uint32_t A,
                                            uint32_t slot, uint32_t chunk, bool store, uint32_t seed, uint32_t kmix,
                                            const char *Tb, int &down) {
    const uint32_t S1 = 1u << ug.sh, M1 = ug.m1 << ug.sh;
    const uint32_t S2 = 1u << (ug.sh + ug.ll1), M2 = ug.m2 << (ug.sh + ug.ll1);
    const uint32_t cm = (1u << (ug.lh16 + 4)) - 16u;
    const uint32_t adjoff = P ? bfi(cm, A + 16u, A) : bfi(cm, A - 16u, A) + 12u;
    uint4 rows[4];
    const uint4 a = ldb(o1, A);
    const uint32_t adj = ldb(reinterpret_cast<const uint32_t *>(o1), adjoff);
    rows[0] = ldb(o1, bfi(M1, A + S1, A));
    rows[1] = ldb(o1, bfi(M1, A - S1, A));
    rows[2] = ldb(o1, bfi(M2, A + S2, A));
    rows[3] = ldb(o1, bfi(M2, A - S2, A));
    const uint4 s = ldb(o0, A);
    uint32_t y[16];
    __builtin_amdgcn_sched_barrier(0);
    x16f_draws<P, UNI>(seed, A, kmix, y);
    __builtin_amdgcn_sched_barrier(0);
    uint32_t ds[4];
    const uint4 ov = x16f_decide<MODE, P>(a, adj, rows, s, y, Tb, ds);
    R0[slot * 8u + chunk] = ov;
    if (store) {
        stb(n0, A, ov);
        down += (int)(__popc(ov.x) + __popc(ov.y) + __popc(ov.z) + __popc(ov.w));
    }
}

// One colour-1 chunk of the tile: the new colour-0 neighbours from LDS.
template <int MODE, int P, bool MEAS>
__device__ __forceinline__ void x16f_chunk1(const uint4 &s, uint4 *__restrict__ n1, const uint4 *R0, uint32_t A,
                                            int yl, int zl, uint32_t chunk, uint32_t seed, uint32_t kmix,
                                            const char *Tb, int &down, int &bond) {
    const uint32_t sl = x16f_slot(yl, zl) * 8u;
    const uint4 a = R0[sl + chunk];
    const uint32_t *R0w = reinterpret_cast<const uint32_t *>(R0);
    const uint32_t adj = P ? R0w[(sl + ((chunk + 1u) & 7u)) * 4u] : R0w[(sl + ((chunk - 1u) & 7u)) * 4u + 3u];
    uint4 rows[4];
    rows[0] = R0[x16f_slot(yl + 1, zl) * 8u + chunk];
    rows[1] = R0[x16f_slot(yl - 1, zl) * 8u + chunk];
    rows[2] = R0[x16f_slot(yl, zl + 1) * 8u + chunk];
    rows[3] = R0[x16f_slot(yl, zl - 1) * 8u + chunk];
    uint32_t y[16];
    __builtin_amdgcn_sched_barrier(0);
    x16f_draws<P, true>(seed, A, kmix, y);
    __builtin_amdgcn_sched_barrier(0);
    uint32_t ds[4];
    const uint4 ov = x16f_decide<MODE, P>(a, adj, rows, s, y, Tb, ds);
    stb(n1, A, ov);
    if (MEAS) {
        // as x16u_chunk's colour-1 bond sum: 16 NN - 2 sum d - 2 NN sum s + 4 sum s d with NN = 6
        const uint32_t ovv[4] = {ov.x, ov.y, ov.z, ov.w};
        const int dn = (int)(__popc(ov.x) + __popc(ov.y) + __popc(ov.z) + __popc(ov.w));
        uint32_t dsum = 0, sd = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            dsum = __builtin_amdgcn_sad_u8(ds[q], 0u, dsum);
            sd = __builtin_amdgcn_sad_u8(ds[q] & (ovv[q] * 0xffu), 0u, sd);
        }
        down += dn;
        bond += 16 * 6 - 2 * (int)dsum - 2 * 6 * dn + 4 * (int)sd;
    }
}

template <int MODE, bool MEAS>
__global__ __launch_bounds__(1024) void ising_sweep_x16f(const uint4 *__restrict__ o0, const uint4 *__restrict__ o1,
                                                         uint4 *__restrict__ n0, uint4 *__restrict__ n1, X16UGeom ug,
                                                         uint32_t seed, uint32_t skey, SweepThresholds th,
                                                         RecordFold fold, uint32_t tiles1, MeasArgs m) {
    __shared__ uint32_t T[16];
    __shared__ uint4 R0[X16F_S * X16F_S * 8];
    __shared__ long long lds[16];
    if (threadIdx.x == 0) {
        // T[d + 8 s] = threshold * 256 - 1 for own spin s and d down neighbours (d <= 6)
#pragma unroll
        for (int i = 0; i < SWEEP_TH_LEN; ++i) T[i % 7 + 8 * (i / 7)] = (th.t[i] << 8) - 1u;
    }
    __syncthreads();
    if (fold.rec) fold_record(fold, lds);
    const uint32_t lb = (uint32_t)logical_block();
    const uint32_t y0 = (lb % tiles1) * X16F_T, z0 = (lb / tiles1) * X16F_T;
    const uint32_t kmix = sweep_mix(skey);
    const char *Tb = reinterpret_cast<const char *>(T);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j = lane >> 3, chunk = lane & 7u;
    int down = 0, bond = 0;  // with MEAS: the record's sums over the tile's sites
    // colour 0: groups wave, wave + 16, wave + 32
    for (uint32_t g = wave; g < (uint32_t)X16F_GROUPS0; g += 16u) {
        int yl, zl;
        bool uni = true, store = false;
        if (g < 32u) {
            zl = (int)(g >> 1);
            yl = (int)(g & 1u) + 2 * (int)j;
            store = true;
        } else if (g < 36u) {
            zl = ((g - 32u) >> 1) ? X16F_T : -1;
            yl = (int)(g & 1u) + 2 * (int)j;
        } else {
            yl = ((g - 36u) >> 1) ? X16F_T : -1;
            zl = (int)(g & 1u) + 2 * (int)j;
            uni = false;
        }
        const uint32_t y = (y0 + (uint32_t)yl) & ug.m1, z = (z0 + (uint32_t)zl) & ug.m2;
        const uint32_t A = ((y | (z << ug.ll1)) << ug.sh) | (chunk << 4);
        const uint32_t p = __builtin_amdgcn_readfirstlane((y + z) & 1u);
        const uint32_t slot = x16f_slot(yl, zl);
        if (uni) {
            if (p) x16f_chunk0<MODE, 1, true>(o0, o1, n0, R0, ug, A, slot, chunk, store, seed, kmix, Tb, down);
            else x16f_chunk0<MODE, 0, true>(o0, o1, n0, R0, ug, A, slot, chunk, store, seed, kmix, Tb, down);
        } else {
            if (p) x16f_chunk0<MODE, 1, false>(o0, o1, n0, R0, ug, A, slot, chunk, false, seed, kmix, Tb, down);
            else x16f_chunk0<MODE, 0, false>(o0, o1, n0, R0, ug, A, slot, chunk, false, seed, kmix, Tb, down);
        }
    }
    __syncthreads();
    // colour 1: groups wave, wave + 16
    for (uint32_t g = wave; g < (uint32_t)X16F_GROUPS1; g += 16u) {
        const int zl = (int)(g >> 1), yl = (int)(g & 1u) + 2 * (int)j;
        const uint32_t y = y0 + (uint32_t)yl, z = z0 + (uint32_t)zl;
        const uint32_t A = ((y | (z << ug.ll1)) << ug.sh) | (chunk << 4);
        const uint32_t p = __builtin_amdgcn_readfirstlane((y + z + 1u) & 1u);
        const uint4 s = ldb(o1, A);
        if (p) x16f_chunk1<MODE, 1, MEAS>(s, n1, R0, A, yl, zl, chunk, seed, kmix, Tb, down, bond);
        else x16f_chunk1<MODE, 0, MEAS>(s, n1, R0, A, yl, zl, chunk, seed, kmix, Tb, down, bond);
    }
    if (MEAS) {
        // every lane updated 16 sites of each colour per tile-row group: 32 groups of 64 lanes per colour
        const int total = 16 * 2 * (X16F_GROUPS1 / 16);
        int counts[2] = {total - down, down};
        emit_partials(m, (long long)bond, counts, 2);
    }
}

bool x16f_applies(const Geom &g, int uniform) {
    return uniform && g.D == 3 && g.L[0] == 256 && g.L[1] % X16F_T == 0 && g.L[2] % X16F_T == 0 &&
           g.L[1] >= 2 * X16F_T && g.L[2] >= 2 * X16F_T;
}

int64_t x16f_tiles(const Geom &g) { return (int64_t)(g.L[1] / X16F_T) * (g.L[2] / X16F_T); }

void launch_site_sweep_x16f(const uint8_t *o0, const uint8_t *o1, uint8_t *n0, uint8_t *n1, const Geom &g,
                            const X16UGeom &ug, int mode, uint32_t seed, uint32_t skey, const SweepThresholds &th,
                            const RecordFold &fold, const MeasArgs &m, hipStream_t s) {
    const uint32_t tiles1 = (uint32_t)(g.L[1] / X16F_T);
    dim3 grid((uint32_t)x16f_tiles(g)), block(1024);
#define PLQ_X16F(MODE, MEAS)                                                                                      \
    ising_sweep_x16f<MODE, MEAS><<<grid, block, 0, s>>>((const uint4 *)o0, (const uint4 *)o1, (uint4 *)n0,       \
                                                        (uint4 *)n1, ug, seed, skey, th, fold, tiles1, m)
    if (mode == 0) {
        if (m.on) PLQ_X16F(0, true);
        else PLQ_X16F(0, false);
    } else {
        if (m.on) PLQ_X16F(1, true);
        else PLQ_X16F(1, false);
    }
#undef PLQ_X16F
}

// --------------------------------------------------------------- launchers ---

static void grid_for_work(int64_t work, int64_t max_blocks, int *blocks, int64_t *span) {
    int64_t need = (work + 255) / 256;
    if (need < 1) need = 1;
    int64_t G = need < max_blocks ? need : max_blocks;
    if (G >= 8) G = (G / 8) * 8;
    *blocks = (int)G;
    *span = (work + G - 1) / G;
}

static uint32_t log2u(int64_t v) {
    uint32_t k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
}

X16UGeom x16u_geom(const Geom &g) {
    X16UGeom u{};
    u.lh16 = log2u(g.L[0] / 32);
    u.sh = u.lh16 + 4;
    u.lk2 = 7 - u.lh16;
    u.ll1 = log2u(g.L[1]);
    u.ll2 = log2u(g.L[2]);
    u.m1 = (uint32_t)g.L[1] - 1u;
    u.m2 = (uint32_t)g.L[2] - 1u;
    u.m3 = (uint32_t)g.L[3] - 1u;
    return u;
}

bool x16u_applies(const Geom &g, const uint32_t *host_table, size_t table_len, int64_t max_blocks) {
    if (g.D < 2 || g.L[0] < 32 || g.L[0] > 2048) return false;
    for (int mu = 0; mu < g.D; ++mu)
        if (g.L[mu] & (g.L[mu] - 1)) return false;
    const int64_t chunks_per_row = g.L[0] / 32;
    const int64_t rows_per_pair = 2 * (64 / chunks_per_row);  // 2K
    if (g.L[1] < rows_per_pair) return false;
    // whole waves per block span: work and the grid are powers of two
    const int64_t work = (int64_t)g.R * chunks_per_row;
    if (work < 256 || work > (int64_t(1) << 31)) return false;
    int64_t G = work / 256 < max_blocks ? work / 256 : max_blocks;
    if (G & (G - 1)) return false;
    if (g.N > (int64_t(1) << 32)) return false;
    for (size_t k = 0; k < table_len; ++k)
        if (host_table[k] == 0 || host_table[k] > DRAW_ONE) return false;
    return true;
}

void launch_pack_sites(const uint8_t *lat, uint8_t *c0, uint8_t *c1, const Geom &g, hipStream_t s) {
    int64_t blocks = (g.N + 255) / 256;
    pack_sites_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(lat, c0, c1, g);
}

void launch_unpack_sites(uint8_t *lat, const uint8_t *c0, const uint8_t *c1, const Geom &g, hipStream_t s) {
    int64_t blocks = (g.N + 255) / 256;
    unpack_sites_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(lat, c0, c1, g);
}

void launch_finalize_record(const long long *pa, int64_t na, const long long *pb, int64_t nb, long long *rec,
                            hipStream_t s) {
    finalize_record_kernel<<<dim3(1), dim3(1024), 0, s>>>(pa, na, pb, nb, rec);
}

void launch_finalize_columns(const long long *pa, int64_t na, const long long *pb, int64_t nb, int nk, long long *rec,
                             hipStream_t s) {
    finalize_columns_kernel<<<dim3(REC_K), dim3(256), 0, s>>>(pa, na, pb, nb, nk, rec, nullptr);
}

void launch_finalize_columns_ring(const long long *pa, int64_t na, const long long *pb, int64_t nb, int nk,
                                  long long *rec_base, const uint32_t *ctr, hipStream_t s) {
    finalize_columns_kernel<<<dim3(REC_K), dim3(256), 0, s>>>(pa, na, pb, nb, nk, rec_base, ctr);
}

// Device counters of a site chain's graph replays: ctr[0] the sweep, ctr[1]
// the record slot.
__global__ void site_ctr_set(uint32_t *ctr, uint32_t sweep, uint32_t slot) {
    ctr[0] = sweep;
    ctr[1] = slot;
}

__global__ void site_ctr_tick(uint32_t *ctr, uint32_t sweeps, uint32_t records, uint32_t cap) {
    ctr[0] += sweeps;
    if (records) ctr[1] = (uint32_t)(((uint64_t)ctr[1] + records) % cap);
}

void launch_site_ctr_set(uint32_t *ctr, uint32_t sweep, uint32_t slot, hipStream_t s) {
    site_ctr_set<<<dim3(1), dim3(1), 0, s>>>(ctr, sweep, slot);
}

void launch_site_ctr_tick(uint32_t *ctr, uint32_t sweeps, uint32_t records, uint32_t cap, hipStream_t s) {
    site_ctr_tick<<<dim3(1), dim3(1), 0, s>>>(ctr, sweeps, records, cap);
}

static int64_t pass_work(const SitePassArgs &a) { return (int64_t)a.g.R * a.g.H / a.vec; }

int site_pass_blocks(const SitePassArgs &a) {
    int blocks;
    int64_t span;
    grid_for_work(pass_work(a), a.max_blocks, &blocks, &span);
    return blocks;
}

template <int D>
static void dispatch_generic(const SitePassArgs &a, int blocks, int64_t work, int64_t span) {
    dim3 grid((uint32_t)blocks), block(256);
    const uint8_t *oth = a.oth;
    uint8_t *own = a.own;
#define PLQ_GEN(MODEL, MODE)                                                                                    \
    site_pass_generic<D, MODEL, MODE><<<grid, block, 0, a.stream>>>(oth, own, a.g, a.color, a.seed, a.skey, a.q, \
                                                                    a.table, a.m, work, span)
    if (a.model == 0) {
        if (a.mode == 0) PLQ_GEN(0, 0);
        else if (a.mode == 1) PLQ_GEN(0, 1);
        else PLQ_GEN(0, 2);
    } else {
        if (a.mode == 0) PLQ_GEN(1, 0);
        else if (a.mode == 1) PLQ_GEN(1, 1);
        else PLQ_GEN(1, 2);
    }
#undef PLQ_GEN
}

template <int D>
static void dispatch_x4(const SitePassArgs &a, int blocks, int64_t work, int64_t span) {
    dim3 grid((uint32_t)blocks), block(256);
    const uint32_t *oth = (const uint32_t *)a.oth;
    uint32_t *own = (uint32_t *)a.own;
#define PLQ_X4(MODE)                                                                                         \
    ising_pass_swar<D, MODE><<<grid, block, 0, a.stream>>>(oth, own, a.sg, a.color, a.seed, a.skey, a.table, \
                                                           a.m, work, span)
    if (a.mode == 0) PLQ_X4(0);
    else if (a.mode == 1) PLQ_X4(1);
    else PLQ_X4(2);
#undef PLQ_X4
}

template <int D>
static void dispatch_x16u(const SitePassArgs &a, int blocks, int64_t work, int64_t span) {
    dim3 grid((uint32_t)blocks), block(256);
    const uint4 *oth = (const uint4 *)a.oth;
    uint4 *own = (uint4 *)a.own;
    if (a.mode == 0)
        ising_pass_x16u<D, 0><<<grid, block, 0, a.stream>>>(oth, own, a.ug, a.color, a.seed, a.skey, a.ctr,
                                                            a.sweep_off, a.table, a.m, (uint32_t)work,
                                                            (uint32_t)span);
    else
        ising_pass_x16u<D, 1><<<grid, block, 0, a.stream>>>(oth, own, a.ug, a.color, a.seed, a.skey, a.ctr,
                                                            a.sweep_off, a.table, a.m, (uint32_t)work,
                                                            (uint32_t)span);
}

template <int D>
static void dispatch_x16(const SitePassArgs &a, int blocks, int64_t work, int64_t span) {
    dim3 grid((uint32_t)blocks), block(256);
    const uint4 *oth = (const uint4 *)a.oth;
    uint4 *own = (uint4 *)a.own;
#define PLQ_X16(MODE)                                                                                       \
    ising_pass_x16<D, MODE><<<grid, block, 0, a.stream>>>(oth, own, a.sg, a.color, a.seed, a.skey, a.table, \
                                                          a.m, work, span)
    if (a.mode == 0) PLQ_X16(0);
    else if (a.mode == 1) PLQ_X16(1);
    else PLQ_X16(2);
#undef PLQ_X16
}

void launch_site_pass(const SitePassArgs &a_in) {
    int blocks;
    int64_t span;
    int64_t work = pass_work(a_in);
    grid_for_work(work, a_in.max_blocks, &blocks, &span);
    // An earlier record's fold rides on the uniform pass when that pass
    // writes no partials of its own and has a workgroup per column; else it
    // is launched first.
    SitePassArgs a = a_in;
    const bool x16u = a.vec == 16 && a.uniform && a.mode != 2 && span % 64 == 0 && a.g.D >= 2;
    if (a.m.fold.rec && (!x16u || a.m.on || blocks < a.m.fold.nk)) {
        launch_finalize_columns(a.m.fold.pa, a.m.fold.na, a.m.fold.pb, a.m.fold.nb, a.m.fold.nk, a.m.fold.rec,
                                a.stream);
        a.m.fold = RecordFold{};
    }
#define PLQ_BY_D(FN)                                   \
    switch (a.g.D) {                                   \
        case 1: FN<1>(a, blocks, work, span); break;   \
        case 2: FN<2>(a, blocks, work, span); break;   \
        case 3: FN<3>(a, blocks, work, span); break;   \
        default: FN<4>(a, blocks, work, span); break;  \
    }
    if (x16u) {
        switch (a.g.D) {
            case 2: dispatch_x16u<2>(a, blocks, work, span); break;
            case 3: dispatch_x16u<3>(a, blocks, work, span); break;
            default: dispatch_x16u<4>(a, blocks, work, span); break;
        }
    } else if (a.vec == 16) {
        PLQ_BY_D(dispatch_x16)
    } else if (a.vec == 4) {
        PLQ_BY_D(dispatch_x4)
    } else {
        PLQ_BY_D(dispatch_generic)
    }
#undef PLQ_BY_D
}

}  // namespace plq
