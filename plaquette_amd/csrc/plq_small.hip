// plq_small.hip -- the whole chain of a small Ising lattice in one workgroup.
//
// For lattices of a few thousand sites (2D 64^2: 4,096) a colour pass is a
// few microseconds of launch overhead around well under a microsecond of
// work.  This kernel keeps both colour arrays in LDS and runs many sweeps of
// the reference checkerboard (site_schedule.wgsl:15-31, updater.rs:316-324)
// in one launch: colour 0, barrier, colour 1, barrier, per sweep.  The draws
// are the same keyed draws (rng.wgsl:47-52) keyed by the global site index
// and the chain's sweep counter, and the decisions the same integer
// thresholds as the multi-launch kernels, so the chain is bit-identical to
// them whatever path runs it.
//
// Each lane owns K sites of each colour for the whole launch and computes
// their LDS offsets once: per site one word 2*offset + parity (which is also
// the site's lattice index, since site = r*L0 + 2*i + parity and
// offset = r*L0/2 + i) and, per axis, the two neighbour offsets packed in one
// word.  A recorded sweep also sums the bond terms (colour-1 sites against
// their updated colour-0 neighbours, so every bond once) and the down count,
// reduced across the block into one record of the chain's ring
// (ising/mod.rs:97-117, 177-186).
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

constexpr int SMALL_THREADS = 1024;
constexpr uint32_t NO_SITE = 0xffffffffu;
constexpr uint32_t SMALL_RING = 64;

struct SmallGeom {
    uint32_t lg[MAX_D];     // log2 of each extent (0 for unused axes)
    uint32_t rshift[MAX_D]; // bit position of axis mu >= 1 within the row index
    uint32_t half;          // sites of one colour (N / 2)
};

typedef int SmallRing[SMALL_RING][SMALL_THREADS / 64][2];

// The record stream of a launch.  A recorded sweep's per-wave (bond, down)
// sums go to ring[pend] before the colour-1 barrier; when SMALL_RING are
// pending, and at the end of the launch, wave 0 turns them into records, one
// lane per record.  When more are pending than the record ring holds, only
// the newest rec_cap are written, as sequential writes would leave them.
// The next writes to ring[] come after the next sweep's colour-0 barrier,
// which wave 0 reaches only after its reads.
struct SmallRecorder {
    long long *rec;
    uint64_t rec_cap;
    uint64_t slot;   // record-ring slot of the first pending record
    uint64_t togo;   // sweeps before the next record
    uint64_t every;
    uint32_t pend;   // recorded sweeps held in ring[]
    uint32_t half;

    __device__ SmallRecorder(long long *r, uint64_t cap, uint64_t first, uint64_t ev, uint64_t phase, uint32_t h)
        : rec(r), rec_cap(cap), slot(ev ? first % cap : 0), togo(ev ? ev - 1 - phase % ev : 0), every(ev), pend(0),
          half(h) {}

    __device__ bool recording() const { return every != 0 && togo == 0; }

    // Publish this wave's sums (every lane of every wave calls this, before
    // the colour-1 barrier of a recorded sweep).  A wave that owns no site
    // publishes zeros without reducing.  The two sums travel in one word,
    // down * 2^16 + bond: a wave's |bond| and down count are below 2^15.
    __device__ void publish(SmallRing &ring, int bond, int ndown, bool wave_owns_sites) const {
        int packed = 0;
        if (wave_owns_sites) packed = wave_sum_i32(ndown * 65536 + bond);
        if ((threadIdx.x & 63) == 0) {
            const int b = ((packed + 32768) & 0xffff) - 32768;
            ring[pend][threadIdx.x >> 6][0] = b;
            ring[pend][threadIdx.x >> 6][1] = (packed - b) >> 16;
        }
    }

    __device__ void flush(const SmallRing &ring) {
        const uint32_t tid = threadIdx.x;
        if (tid < pend && tid + rec_cap >= pend) {
            int b = 0, nd = 0;
#pragma unroll
            for (int wv = 0; wv < SMALL_THREADS / 64; ++wv) {
                b += ring[tid][wv][0];
                nd += ring[tid][wv][1];
            }
            long long *out = rec + ((slot + tid) % rec_cap) * REC_K;
            out[0] = b;
            out[1] = 0;
            out[2] = 2 * (long long)half - nd;
            out[3] = nd;
            for (int q = 4; q < REC_K; ++q) out[q] = 0;
        }
        slot = (slot + pend) % rec_cap;
        pend = 0;
    }

    // After the sweep's last barrier.
    __device__ void end_sweep(const SmallRing &ring) {
        if (recording()) {
            pend += 1;
            if (pend == SMALL_RING) flush(ring);
            togo = every - 1;
        } else if (every != 0) {
            togo -= 1;
        }
    }

    __device__ void finish(const SmallRing &ring) {
        if (pend) flush(ring);
    }
};

template <int D, int K, int MODE>
__global__ __launch_bounds__(SMALL_THREADS) void small_chain_kernel(uint8_t *__restrict__ c0g,
                                                                    uint8_t *__restrict__ c1g, SmallGeom sg,
                                                                    uint32_t seed, uint32_t sweep0, uint32_t nsweeps,
                                                                    const uint32_t *__restrict__ table,
                                                                    long long *__restrict__ rec, uint64_t rec_cap,
                                                                    uint64_t rec_first, uint64_t every,
                                                                    uint64_t phase) {
    constexpr int NN = 2 * D;
    __shared__ uint8_t lat[2][SMALL_MAX_SITES / 2];
    __shared__ uint32_t tab[2 * (NN + 1)];
    __shared__ SmallRing ring;
    const uint32_t tid = threadIdx.x;
    const uint32_t half = sg.half;

    for (uint32_t i = tid; i < half; i += SMALL_THREADS) {
        lat[0][i] = c0g[i];
        lat[1][i] = c1g[i];
    }
    if (tid < 2 * (NN + 1)) tab[tid] = table[tid];

    // Per-lane site words: own[c][k] = 2*offset + parity (= lattice index),
    // nb[c][k][mu] = (offset of the - neighbour << 16) | offset of the + one.
    uint32_t own[2][K];
    uint32_t h1[2][K];  // the first hash round of the site's draw, the same every sweep
    uint32_t nb[2][K][D];
    const uint32_t lh = sg.lg[0] - 1;  // log2(L0 / 2)
    const uint32_t L0m = (1u << sg.lg[0]) - 1u;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            uint32_t w = tid + (uint32_t)k * SMALL_THREADS;
            if (w >= half) {
                own[c][k] = NO_SITE;
                h1[c][k] = 0;
#pragma unroll
                for (int mu = 0; mu < D; ++mu) nb[c][k][mu] = 0;
                continue;
            }
            uint32_t r = w >> lh;
            uint32_t i = w & ((1u << lh) - 1u);
            uint32_t prs = 0;
#pragma unroll
            for (int mu = 1; mu < D; ++mu) prs += (r >> sg.rshift[mu]) & ((1u << sg.lg[mu]) - 1u);
            uint32_t p = (prs + (uint32_t)c) & 1u;
            uint32_t x0 = 2 * i + p;
            uint32_t rowbase = r << lh;
            own[c][k] = 2 * w + p;
            {
                const uint32_t x = seed ^ own[c][k];
                h1[c][k] = lowbias32_mid(x ^ (x >> 16));
            }
            uint32_t xp = (x0 + 1) & L0m, xm = (x0 - 1) & L0m;
            nb[c][k][0] = (rowbase + (xp >> 1)) | ((rowbase + (xm >> 1)) << 16);
#pragma unroll
            for (int mu = 1; mu < D; ++mu) {
                uint32_t Lm = (1u << sg.lg[mu]) - 1u;
                uint32_t x = (r >> sg.rshift[mu]) & Lm;
                uint32_t rp = r + ((((x + 1) & Lm) - x) << sg.rshift[mu]);
                uint32_t rm = r + ((((x - 1) & Lm) - x) << sg.rshift[mu]);
                nb[c][k][mu] = ((rp << lh) + i) | (((rm << lh) + i) << 16);
            }
        }
    __syncthreads();

    SmallRecorder records(rec, rec_cap, rec_first, every, phase, half);
    for (uint32_t t = 0; t < nsweeps; ++t) {
        const uint32_t kmix = sweep_mix((sweep0 + t) * GOLDEN);  // keyed_bits_k's merged xor-shift
        const bool record = records.recording();
        int bond = 0, ndown = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            uint8_t *mine = lat[c];
            const uint8_t *oth = lat[1 - c];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t op = own[c][k];
                if (op == NO_SITE) continue;
                int d = 0;
#pragma unroll
                for (int mu = 0; mu < D; ++mu) {
                    uint32_t pr = nb[c][k][mu];
                    d += oth[pr & 0xffffu] + oth[pr >> 16];
                }
                const uint32_t off = op >> 1;
                const int cur = mine[off];
                const uint32_t th = tab[cur * (NN + 1) + d];
                bool hit;
                if (th == 0u) hit = false;
                else if (th >= DRAW_ONE) hit = true;
                else hit = (lowbias32_tail(h1[c][k] ^ kmix) >> 8) < th;  // keyed_bits_k(seed, op, sweep key)
                const int nxt = MODE == 0 ? (hit ? 1 - cur : cur) : (hit ? 0 : 1);
                mine[off] = (uint8_t)nxt;
                if (record) {
                    ndown += nxt;
                    if (c == 1) bond += (1 - 2 * nxt) * (NN - 2 * d);
                }
            }
            if (c == 1 && record) records.publish(ring, bond, ndown, (tid & ~63u) < half);
            __syncthreads();
        }
        records.end_sweep(ring);
    }
    records.finish(ring);  // the last recorded sweep's barrier has published its entries

    for (uint32_t i = tid; i < half; i += SMALL_THREADS) {
        c0g[i] = lat[0][i];
        c1g[i] = lat[1][i];
    }
}

// Four same-colour sites per lane, one dword of a colour row (L0 >= 8).  The
// per-site work shrinks to the keyed draw, one LDS threshold read and a
// compare; the neighbour counts of the four sites are one dword sum (the
// axis-0 neighbours are the other colour's dword shifted by a byte, the other
// axes' the other colour's dwords in the neighbour rows), and the decisions
// are gathered back into the dword with byte permutes.  The draws and
// thresholds are those of the per-site kernel above, so the two agree bit for
// bit.
template <int D, int MODE>
__global__ __launch_bounds__(SMALL_THREADS) void small_chain_x4(uint8_t *__restrict__ c0g, uint8_t *__restrict__ c1g,
                                                               SmallGeom sg, uint32_t seed, uint32_t sweep0,
                                                               uint32_t nsweeps, const uint32_t *__restrict__ table,
                                                               long long *__restrict__ rec, uint64_t rec_cap,
                                                               uint64_t rec_first, uint64_t every, uint64_t phase) {
    constexpr int NN = 2 * D;
    constexpr int K = NN + 1;
    __shared__ uint32_t lat[2][SMALL_MAX_SITES / 8];
    __shared__ uint32_t T[2 * 16];  // T[d * 2 + s], byte address (d << 3) | (s << 2)
    __shared__ SmallRing ring;
    const uint32_t tid = threadIdx.x;
    const uint32_t half = sg.half;
    const uint32_t nw = half >> 2;  // dwords of one colour

    const uint32_t *c0w = reinterpret_cast<const uint32_t *>(c0g);
    const uint32_t *c1w = reinterpret_cast<const uint32_t *>(c1g);
    for (uint32_t i = tid; i < nw; i += SMALL_THREADS) {
        lat[0][i] = c0w[i];
        lat[1][i] = c1w[i];
    }
    if (tid < 2 * K) {
        const int s = (int)tid / K, d = (int)tid % K;
        T[d * 2 + s] = table[tid];
    }

    // This lane's dword w = tid: row r, dword j of the row (H4 = L0 / 8 per row).
    const bool active = tid < nw;
    const uint32_t lh4 = sg.lg[0] - 3;
    const uint32_t r = tid >> lh4, j = tid & ((1u << lh4) - 1u);
    uint32_t prs = 0, rowp[D > 1 ? D - 1 : 1], rowm[D > 1 ? D - 1 : 1];
#pragma unroll
    for (int mu = 1; mu < D; ++mu) {
        const uint32_t Lm = (1u << sg.lg[mu]) - 1u;
        const uint32_t x = (r >> sg.rshift[mu]) & Lm;
        prs += x;
        const uint32_t rp = r + ((((x + 1) & Lm) - x) << sg.rshift[mu]);
        const uint32_t rm = r + ((((x - 1) & Lm) - x) << sg.rshift[mu]);
        rowp[mu - 1] = (rp << lh4) + j;
        rowm[mu - 1] = (rm << lh4) + j;
    }
    uint32_t p[2], adj[2], u[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        p[c] = (prs + (uint32_t)c) & 1u;
        // p = 1 needs the next dword's first byte, p = 0 the previous dword's last
        adj[c] = (r << lh4) + ((j + (p[c] ? 1u : (1u << lh4) - 1u)) & ((1u << lh4) - 1u));
        // site k of the dword is site0 + 2k, and site0 has bits 1..4 clear, so
        // the first xor-shift of the draw, (x ^ x >> 16) with x = seed ^ site,
        // is v0 ^ 2k.  The first hash round does not depend on the sweep: it
        // is computed once, and a sweep's draw is the second round of it.
        const uint32_t bx = seed ^ ((r << sg.lg[0]) + (j << 3) + p[c]);
        const uint32_t v0 = bx ^ (bx >> 16);
#pragma unroll
        for (int k = 0; k < 4; ++k) u[c][k] = lowbias32_mid(v0 ^ (2u * (uint32_t)k));
    }
    const char *Tb = reinterpret_cast<const char *>(T);
    __syncthreads();

    SmallRecorder records(rec, rec_cap, rec_first, every, phase, half);
    for (uint32_t t = 0; t < nsweeps; ++t) {
        const uint32_t kmix = sweep_mix((sweep0 + t) * GOLDEN);  // keyed_bits_k's merged xor-shift
        const bool record = records.recording();
        int bond = 0, ndown = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (active) {
                uint32_t *mine = lat[c];
                const uint32_t *oth = lat[1 - c];
                const uint32_t a = oth[tid], aj = oth[adj[c]];
                uint32_t ds = a + (p[c] ? __builtin_amdgcn_alignbyte(aj, a, 1) : __builtin_amdgcn_alignbyte(a, aj, 3));
#pragma unroll
                for (int mu = 0; mu < D - 1; ++mu) ds += oth[rowp[mu]] + oth[rowm[mu]];
                const uint32_t s4 = mine[tid];
                const uint32_t addr4 = (ds << 3) | (s4 << 2);
                uint32_t diff[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t tk = *reinterpret_cast<const uint32_t *>(Tb + ((addr4 >> (8 * k)) & 0xffu));
                    const uint32_t m = lowbias32_tail(u[c][k] ^ kmix) >> 8;
                    diff[k] = m - tk;  // negative iff m < t (both below 2^31)
                }
                const uint32_t lo = __builtin_amdgcn_perm(diff[1], diff[0], 0x0c0c0703u);
                const uint32_t hi = __builtin_amdgcn_perm(diff[3], diff[2], 0x07030c0cu);
                const uint32_t hits = ((lo | hi) >> 7) & 0x01010101u;
                const uint32_t out = MODE == 0 ? (s4 ^ hits) : (hits ^ 0x01010101u);
                mine[tid] = out;
                if (record) {
                    const int downs = __popc(out);
                    ndown += downs;
                    if (c == 1) {
                        // sum over the four sites of (1 - 2 s)(NN - 2 d), with
                        // byte sums taken by one multiply (each below 256)
                        const int dsum = (int)((ds * 0x01010101u) >> 24);
                        const int sd = (int)(((ds & (out * 0xffu)) * 0x01010101u) >> 24);
                        bond += 4 * NN - 2 * dsum - 2 * NN * downs + 4 * sd;
                    }
                }
            }
            if (c == 1 && record) records.publish(ring, bond, ndown, (tid & ~63u) < nw);
            __syncthreads();
        }
        records.end_sweep(ring);
    }
    records.finish(ring);

    uint32_t *c0o = reinterpret_cast<uint32_t *>(c0g);
    uint32_t *c1o = reinterpret_cast<uint32_t *>(c1g);
    for (uint32_t i = tid; i < nw; i += SMALL_THREADS) {
        c0o[i] = lat[0][i];
        c1o[i] = lat[1][i];
    }
}

template <int D, int MODE>
void launch_small_k(const SmallChainArgs &a, const SmallGeom &sg, int k) {
    dim3 grid(1), block(SMALL_THREADS);
#define PLQ_SMALL(KK)                                                                                           \
    small_chain_kernel<D, KK, MODE><<<grid, block, 0, a.stream>>>(a.c0, a.c1, sg, a.seed, a.sweep0, a.nsweeps, \
                                                                  a.table, a.rec, a.rec_cap, a.rec_first,      \
                                                                  a.every, a.phase)
    if (k <= 1) PLQ_SMALL(1);
    else if (k <= 2) PLQ_SMALL(2);
    else PLQ_SMALL(4);
#undef PLQ_SMALL
}

template <int D>
void launch_small_d(const SmallChainArgs &a, const SmallGeom &sg, int k) {
    if (sg.lg[0] >= 3) {  // L0 >= 8: whole dwords of a colour row
        dim3 grid(1), block(SMALL_THREADS);
#define PLQ_SMALL_X4(MD)                                                                                  \
    small_chain_x4<D, MD><<<grid, block, 0, a.stream>>>(a.c0, a.c1, sg, a.seed, a.sweep0, a.nsweeps, a.table, \
                                                        a.rec, a.rec_cap, a.rec_first, a.every, a.phase)
        if (a.mode == 0) PLQ_SMALL_X4(0);
        else PLQ_SMALL_X4(1);
#undef PLQ_SMALL_X4
        return;
    }
    if (a.mode == 0) launch_small_k<D, 0>(a, sg, k);
    else launch_small_k<D, 1>(a, sg, k);
}

}  // namespace

// Up to four sites of each colour per lane (more spills registers in 4D).
int64_t small_chain_max_sites(int D) { return D <= 3 ? SMALL_MAX_SITES : SMALL_MAX_SITES / 2; }

void launch_small_chain(const SmallChainArgs &a) {
    SmallGeom sg{};
    uint32_t shift = 0;
    for (int mu = 0; mu < MAX_D; ++mu) {
        sg.lg[mu] = mu < a.D ? a.lg[mu] : 0;
        if (mu >= 1) {
            sg.rshift[mu] = shift;
            shift += sg.lg[mu];
        }
    }
    sg.half = (uint32_t)(a.N / 2);
    int k = (int)((sg.half + SMALL_THREADS - 1) / SMALL_THREADS);
    switch (a.D) {
        case 1: launch_small_d<1>(a, sg, k); break;
        case 2: launch_small_d<2>(a, sg, k); break;
        case 3: launch_small_d<3>(a, sg, k); break;
        default: launch_small_d<4>(a, sg, k); break;
    }
}

}  // namespace plq
