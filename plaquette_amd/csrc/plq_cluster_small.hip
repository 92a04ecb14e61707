// plq_cluster_small.hip -- whole Swendsen-Wang / Wolff chains of a small
// lattice in one workgroup.
//
// The general cluster sweep is four kernels (plq_cluster.hip), so on a
// lattice of a few thousand sites a sweep is a few microseconds of launch
// latency around a fraction of a microsecond of work, and a Wolff step (one
// cluster) costs the same as a full Swendsen-Wang sweep.  Here labels, the
// union-find forest and the new cluster labels stay in LDS for up to 4,096
// sweeps per launch, one barrier-separated phase after another:
//
//   1. every site its own root;
//   2. bond draws and unions (cluster_prelude.wgsl:61-80: bond = site*D + axis
//      to the forward neighbour, open iff the labels agree and m < T_p);
//   3. every site's root, the minimum site of its cluster (links always put
//      the larger root under the smaller), and for Wolff the seed site's root
//      (cluster_relabel.wgsl:55-66);
//   4. each root draws its cluster's label keyed on itself
//      (cluster_relabel.wgsl:39-90: Redraw, or ForcedChange for Wolff);
//   5. every site takes its root's label; a recorded sweep then sums the
//      agreeing bonds and the label counts.
//
// The bonds, roots and draws are those of the four-kernel sweep, so the two
// give the same labels bit for bit; the forest of the last sweep is written
// back for the cluster statistics.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

constexpr int CCS_THREADS = 1024;
constexpr uint8_t CCS_NO_LABEL = 0xff;

__device__ __forceinline__ int32_t ccs_find(int32_t *P, int32_t x) {
    int32_t p = __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        const int32_t gp = __hip_atomic_load(&P[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp != p) __hip_atomic_store(&P[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

// Root of x without compression.  Used once the unions are done, while every
// lane re-points its own sites at their roots: the only writes are then
// P[l] = root(l) by l's owner, so every pointer a walk reads stays an ancestor.
__device__ __forceinline__ int32_t ccs_root(const int32_t *P, int32_t x) {
    int32_t p = __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return x;
}

__device__ __forceinline__ void ccs_union(int32_t *P, int32_t a, int32_t b) {
    for (;;) {
        a = ccs_find(P, a);
        b = ccs_find(P, b);
        if (a == b) return;
        const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
        const int32_t old = atomicCAS(&P[hi], hi, lo);
        if (old == hi) return;
        a = old;  // hi gained a parent meanwhile; continue from it
        b = lo;
    }
}

// CCS_K sites per lane: 8 (N <= 8,192) below four dimensions, 4 (N <= 4,096)
// in four, where eight would spill registers.
template <int D, int CCS_K>
__global__ __launch_bounds__(CCS_THREADS) void cluster_small_kernel(SmallClusterArgs a) {
    __shared__ uint8_t lab[SMALL_CC_MAX_SITES];
    __shared__ uint8_t newl[SMALL_CC_MAX_SITES];
    __shared__ int32_t P[SMALL_CC_MAX_SITES];
    __shared__ int wsum[CCS_THREADS / 64][2 + MAX_Q_DEV];
    __shared__ int32_t seed_root;
    const uint32_t tid = threadIdx.x;
    const uint32_t N = a.N;
    const int q = a.q;

    for (uint32_t i = tid; i < N; i += CCS_THREADS) lab[i] = a.labels[i];

    // forward neighbour of each of this lane's sites along each axis, two
    // 16-bit site indices per word
    uint32_t fwdp[CCS_K][(D + 1) / 2];
#pragma unroll
    for (int k = 0; k < CCS_K; ++k) {
        const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
        uint32_t rest = l, stride = 1;
#pragma unroll
        for (int h = 0; h < (D + 1) / 2; ++h) fwdp[k][h] = 0;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const uint32_t L = (uint32_t)a.L[mu];
            const uint32_t x = rest % L;
            rest /= L;
            const uint32_t f = (x + 1 == L ? l - x * stride : l + stride) & 0xffffu;
            fwdp[k][mu >> 1] |= f << (16 * (mu & 1));
            stride *= L;
        }
    }
    auto fwd = [&](int k, int mu) -> uint32_t { return (fwdp[k][mu >> 1] >> (16 * (mu & 1))) & 0xffffu; };
    __syncthreads();

    const uint32_t lane = tid & 63u;
    const uint64_t below = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
    uint64_t slot = a.every ? a.rec_first % a.rec_cap : 0;
    uint64_t togo = a.every ? a.every - 1 - a.phase % a.every : 0;
    int moved = 0;
    for (uint32_t t = 0; t < a.nsweeps; ++t) {
        const uint32_t sweep = a.sweep0 + t;
        const uint32_t skey = sweep * GOLDEN;
        // 1. Axis 0 by a wave scan: the 64 lanes of a wave hold consecutive
        //    sites, so a run of open axis-0 bonds inside a row is a run of
        //    lanes, and every site points at the first site of its run (the
        //    smaller index, as every union keeps) with no atomics.  The bond
        //    from a wave's last lane to the next wave and the row's wrap bond
        //    are joined in phase 2 with the other axes' bonds.
        uint32_t edge = 0;  // bit k: lane 63's axis-0 bond to the next wave is open
        int32_t rstart[CCS_K];  // the first site of each site's run: the unions start there, one hop up
#pragma unroll
        for (int k = 0; k < CCS_K; ++k) {
            const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
            const bool inrow = l < N && fwd(k, 0) == l + 1;
            const bool open = inrow && lab[l + 1] == lab[l] && keyed_bits_k(a.bond_seed, l * (uint32_t)D, skey) < a.bond_t;
            const uint64_t m = __ballot(open);
            const uint64_t starts = (~m << 1) | 1ull;  // bit j: lane j begins a run
            const uint32_t first = 63u - (uint32_t)__clzll(starts & below);
            rstart[k] = (int32_t)(l - lane + first);
            if (l < N) P[l] = rstart[k];
            if (open && lane == 63u) edge |= 1u << k;
        }
        __syncthreads();
        // 2. The remaining bonds, by union-find in LDS.
#pragma unroll
        for (int k = 0; k < CCS_K; ++k) {
            const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
            if (l >= N) continue;
            if (edge & (1u << k)) ccs_union(P, rstart[k], (int32_t)(l + 1));
            const uint8_t s = lab[l];
#pragma unroll
            for (int mu = 0; mu < D; ++mu) {
                const uint32_t lp = fwd(k, mu);
                if (mu == 0 && (lp == l + 1 || lp == l)) continue;  // inside the row (phase 1), or an extent of 1
                if (lab[lp] == s && keyed_bits_k(a.bond_seed, l * (uint32_t)D + (uint32_t)mu, skey) < a.bond_t)
                    ccs_union(P, rstart[k], (int32_t)lp);
            }
        }
        __syncthreads();
        // 3.
        if (a.seeded && tid == 0) {
            const float u = keyed_uniform_f32(keyed_bits(a.seed_seed, 0u, sweep));
            uint32_t ss = (uint32_t)(u * (float)N);
            if (ss >= N) ss = N - 1u;
            seed_root = ccs_root(P, (int32_t)ss);
        }
        // every site re-pointed at its root
#pragma unroll
        for (int k = 0; k < CCS_K; ++k) {
            const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
            if (l < N) __hip_atomic_store(&P[l], ccs_root(P, (int32_t)l), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        // 4.
#pragma unroll
        for (int k = 0; k < CCS_K; ++k) {
            const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
            if (l >= N || P[l] != (int32_t)l) continue;
            uint8_t nl = CCS_NO_LABEL;
            if (!a.seeded || (int32_t)l == seed_root) {
                const float u = keyed_uniform_f32(keyed_bits_k(a.root_seed, l, skey));
                if (a.forced) {
                    uint32_t step = (uint32_t)(u * (float)(uint32_t)(q - 1));
                    if (step >= (uint32_t)(q - 1)) step = (uint32_t)(q - 2);
                    nl = (uint8_t)((lab[l] + 1u + step) % (uint32_t)q);
                } else {
                    uint32_t draw = (uint32_t)(u * (float)(uint32_t)q);
                    if (draw >= (uint32_t)q) draw = (uint32_t)(q - 1);
                    nl = (uint8_t)draw;
                }
            }
            newl[l] = nl;
        }
        __syncthreads();
        // 5.
        moved = 0;
#pragma unroll
        for (int k = 0; k < CCS_K; ++k) {
            const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
            if (l >= N) continue;
            const uint8_t nl = newl[P[l]];
            if (nl != CCS_NO_LABEL) {
                lab[l] = nl;
                moved += 1;
            }
        }
        __syncthreads();
        const bool record = a.every != 0 && togo == 0;
        if (record) {
            // potts/mod.rs:124-139, 237-271: agreeing forward bonds and the
            // label counts of the new configuration.  A lane's label counts
            // are packed four bits each (at most 8 sites per lane).
            int agree = 0;
            uint64_t packed = 0;
#pragma unroll
            for (int k = 0; k < CCS_K; ++k) {
                const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
                if (l >= N) continue;
                const uint8_t sl = lab[l];
                packed += 1ull << (4u * sl);
#pragma unroll
                for (int mu = 0; mu < D; ++mu) agree += lab[fwd(k, mu)] == sl;
            }
            const int aw = wave_sum_i32(agree);
            if (lane == 0) {
                wsum[tid >> 6][0] = aw;
                wsum[tid >> 6][1] = 0;
            }
            for (int v = 0; v < q; ++v) {
                const int cw = wave_sum_i32((int)((packed >> (4 * v)) & 0xfu));
                if (lane == 0) wsum[tid >> 6][2 + v] = cw;
            }
            __syncthreads();
            if (tid < (uint32_t)REC_K) {
                long long sum = 0;
                if ((int)tid < 2 + q)
                    for (int wv = 0; wv < CCS_THREADS / 64; ++wv) sum += wsum[wv][tid];
                a.rec[slot * REC_K + tid] = sum;
            }
            slot = slot + 1 == a.rec_cap ? 0 : slot + 1;
            togo = a.every - 1;
            // the next writes to wsum[] come after the next sweep's first
            // barrier, which the lanes reading it reach only after their reads
        } else if (a.every != 0) {
            togo -= 1;
        }
    }

    for (uint32_t i = tid; i < N; i += CCS_THREADS) a.labels[i] = lab[i];
    // the last sweep's forest: every site points at its root (parent <= site)
#pragma unroll
    for (int k = 0; k < CCS_K; ++k) {
        const uint32_t l = tid + (uint32_t)k * CCS_THREADS;
        if (l < N) a.parent[l] = P[l];
    }
    // the last sweep's relabelled sites, as the four-kernel sweep's partials
    const int mw = wave_sum_i32(a.nsweeps ? moved : 0);
    __syncthreads();
    if ((tid & 63) == 0) wsum[tid >> 6][0] = mw;
    __syncthreads();
    if (tid < (uint32_t)REC_K) {
        long long v = 0;
        if (tid == 1)
            for (int wv = 0; wv < CCS_THREADS / 64; ++wv) v += wsum[wv][0];
        a.partials[tid] = v;
    }
}

}  // namespace

void launch_small_cluster(const SmallClusterArgs &a) {
    dim3 grid(1), block(CCS_THREADS);
    switch (a.D) {
        case 1: cluster_small_kernel<1, 8><<<grid, block, 0, a.stream>>>(a); break;
        case 2: cluster_small_kernel<2, 8><<<grid, block, 0, a.stream>>>(a); break;
        case 3: cluster_small_kernel<3, 8><<<grid, block, 0, a.stream>>>(a); break;
        default: cluster_small_kernel<4, 4><<<grid, block, 0, a.stream>>>(a); break;
    }
}

int64_t small_cluster_max_sites(int D) { return D <= 3 ? SMALL_CC_MAX_SITES : SMALL_CC_MAX_SITES / 2; }

}  // namespace plq
