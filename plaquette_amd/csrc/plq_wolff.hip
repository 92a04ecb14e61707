// plq_wolff.hip -- Wolff moves grown from the seed site on gfx950.
//
// The reference's device Wolff move runs the full Swendsen-Wang bond
// decomposition and relabels only the cluster of a keyed seed site
// (cluster_relabel.wgsl:55-66); its CPU move grows that one cluster from the
// seed, testing each bond at most once (cluster.rs:88-111).  With the bonds
// keyed by (seed, bond, sweep), both reach the same cluster: the seed's
// connected component over the sweep's open bonds.  Its root, the minimum
// site, keys the new label.  So a sweep may grow the cluster instead of
// decomposing the lattice, and the labels come out bit for bit the same.
//
// wolff_grow is one workgroup of 1024 lanes, and the growth lives in its
// LDS: the cluster's sites (a queue of up to 6,144) and an open-addressing
// set of the sites reached (8,192 slots).  Level by level, every lane takes
// sites of the frontier (queue[lo, hi)), loads the labels of their 2D
// neighbours and draws their bonds (bond = site * D + axis, open iff the
// labels agree and m < bond_t, the draw of cc_local and cc_boundary), and
// inserts each site an open bond reaches into the set by compare-and-swap,
// so a site joins once; inserted sites are appended to the queue and the
// root is their minimum.  A level thus costs one round trip to the labels
// and two barriers.  When the frontier empties within the level budget and
// the queue, the kernel relabels the cluster (the draw of cc_relabel*),
// copies the queue to global memory (for the partition's rebuild), writes
// the relabel partials the full sweep would (block 0: [0, size, 0 ...]; the
// others zero) and sets gate = 0, so the gated kernels of the full
// decomposition return at once.  Otherwise it leaves the labels alone and
// sets gate = 1: the full decomposition runs.  Growth pays for clusters a
// few levels deep on large lattices (high temperature, or the small clusters
// of the critical size distribution), and the level budget bounds a failed
// attempt near half the decomposition it precedes.  In adaptive mode the
// kernel keeps running means of an attempt's duration (wall clock) and of
// its success, and attempts only where the mean duration is below the mean
// success times the time a grown sweep saves, and otherwise every 32nd
// sweep (every 256th where the expected cost is twice the saving or more)
// to keep the means current; the chain's states do not depend on the path
// taken.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

constexpr int GROW_THREADS = 1024;

template <int D>
__global__ __launch_bounds__(GROW_THREADS) void wolff_grow(uint8_t *__restrict__ labels, CCGeom cg,
                                                           uint32_t bond_seed, uint32_t root_seed,
                                                           uint32_t seed_seed, uint32_t skey, uint32_t sweep,
                                                           uint32_t bond_t, int q, int forced,
                                                           int32_t *__restrict__ queue_out, uint32_t qcap,
                                                           uint32_t level_cap, int adaptive, float gain_ticks,
                                                           WolffState *__restrict__ st,
                                                           long long *__restrict__ partials, int64_t nparts) {
    __shared__ uint32_t set[WOLFF_SET_SLOTS];  // site + 1, 0 = empty
    __shared__ uint32_t lq[WOLFF_QUEUE];
    __shared__ uint32_t s_tail, s_root, s_lab;
    __shared__ int s_go;
    const uint32_t tid = threadIdx.x;
    const uint64_t t0 = wall_clock64();
    if (tid == 0) {
        int go = 1;
        if (adaptive) {
            // expected cost over expected saving; explore every 32nd sweep
            // near the break-even, every 256th where growth is far from it
            const float r = st->cost / fmaxf(st->success * gain_ticks, 1.0f);
            go = r < 1.0f || st->since >= (r < 2.0f ? 31 : 255);
            st->since = go ? 0 : st->since + 1;
        }
        s_go = go;
        if (!go) {
            st->gate = 1;
            st->full += 1ull;
        }
    }
    for (uint32_t i = tid; i < WOLFF_SET_SLOTS; i += GROW_THREADS) set[i] = 0u;
    __syncthreads();
    if (!s_go) return;
    // true when y was not in the set and is now; a full set (more sites
    // reached than the queue holds, so the attempt fails anyway) refuses
    auto insert = [&](uint32_t y) -> bool {
        uint32_t h = (y * 0x9e3779b1u) >> (32 - WOLFF_SET_BITS);
        for (uint32_t probe = 0; probe < WOLFF_SET_SLOTS; ++probe) {
            const uint32_t old = atomicCAS(&set[h], 0u, y + 1u);
            if (old == 0u) return true;
            if (old == y + 1u) return false;
            h = (h + 1u) & (WOLFF_SET_SLOTS - 1u);
        }
        return false;
    };
    if (tid == 0) {
        // the seed site (cluster_relabel.wgsl:59-62), as seed_cluster_root draws it
        const float u = keyed_uniform_f32(keyed_bits(seed_seed, 0u, sweep));
        uint32_t ss = (uint32_t)(u * (float)(uint32_t)cg.N);
        if (ss >= (uint32_t)cg.N) ss = (uint32_t)cg.N - 1u;
        s_lab = labels[ss];
        (void)insert(ss);
        lq[0] = ss;
        s_tail = 1u;
        s_root = ss;
    }
    __syncthreads();
    const uint32_t lab = s_lab;
    uint32_t lo = 0, hi = 1, levels = 0;
    bool closed = true;
    while (lo < hi) {
        if (levels == level_cap) {
            closed = false;
            break;
        }
        for (uint32_t i = lo + tid; i < hi; i += GROW_THREADS) {
            // once the queue has overflowed the attempt fails: stop early
            if (__hip_atomic_load(&s_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > qcap) break;
            const uint32_t x = lq[i];
            uint32_t nb[2 * D], bond[2 * D];
            uint32_t rest = x;
#pragma unroll
            for (int mu = 0; mu < D; ++mu) {
                const uint32_t L = (uint32_t)cg.L[mu];
                const uint32_t qt = mu + 1 < D ? (cg.mul_div ? cdiv(rest, cg.mL[mu]) : rest / L) : 0u;
                const uint32_t c = mu + 1 < D ? rest - qt * L : rest;
                rest = qt;
                const uint32_t sd = (uint32_t)cg.stride[mu];
                // L = 1: both neighbours are x itself, which is in the set
                nb[2 * mu] = c + 1u >= L ? x - c * sd : x + sd;
                nb[2 * mu + 1] = c == 0u ? x + (L - 1u) * sd : x - sd;
                bond[2 * mu] = x * (uint32_t)D + (uint32_t)mu;
                bond[2 * mu + 1] = nb[2 * mu + 1] * (uint32_t)D + (uint32_t)mu;
            }
            uint32_t agree[2 * D];
#pragma unroll
            for (int k = 0; k < 2 * D; ++k) agree[k] = labels[nb[k]];
#pragma unroll
            for (int k = 0; k < 2 * D; ++k) {
                const bool open = keyed_bits_k(bond_seed, bond[k], skey) < bond_t;
                if (!open || agree[k] != lab) continue;
                if (!insert(nb[k])) continue;
                const uint32_t pos = atomicAdd(&s_tail, 1u);
                if (pos < qcap) lq[pos] = nb[k];
                atomicMin(&s_root, nb[k]);
            }
        }
        __syncthreads();
        const uint32_t t = s_tail;
        __syncthreads();  // every lane has read the tail before the next level appends
        if (t > qcap) {
            closed = false;
            break;
        }
        lo = hi;
        hi = t;
        ++levels;
    }
    // running means over about the last 8 attempts
    auto account = [&](float ok) {
        const float dt = (float)(wall_clock64() - t0);
        st->cost += 0.125f * (dt - st->cost);
        st->success += 0.125f * (ok - st->success);
    };
    if (!closed) {
        if (tid == 0) {
            st->gate = 1;
            st->full += 1ull;
            if (adaptive) account(0.0f);
        }
        return;
    }
    const uint32_t size = hi;
    const uint32_t nl = relabel_draw(keyed_bits_k(root_seed, s_root, skey), lab, q, forced);
    for (uint32_t i = tid; i < size; i += GROW_THREADS) {
        const uint32_t y = lq[i];
        labels[y] = (uint8_t)nl;
        queue_out[i] = (int32_t)y;
    }
    for (int64_t i = tid; i < nparts * REC_K; i += GROW_THREADS) partials[i] = i == 1 ? (long long)size : 0ll;
    if (tid == 0) {
        st->gate = 0;
        st->size = size;
        st->old_label = (int)lab;
        st->grown += 1ull;
        if (adaptive) account(1.0f);
    }
}

__global__ __launch_bounds__(256) void wolff_restore(uint8_t *__restrict__ dst, const int32_t *__restrict__ queue,
                                                     const WolffState *__restrict__ st) {
    if (st->gate != 0) return;
    const uint32_t size = st->size;
    const uint8_t lab = (uint8_t)st->old_label;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x)
        dst[queue[i]] = lab;
}

}  // namespace

void launch_wolff_grow(const WolffGrowArgs &a) {
    const uint32_t skey = a.sweep * GOLDEN;
#define PLQ_WG(DD)                                                                                                \
    wolff_grow<DD><<<dim3(1), dim3(GROW_THREADS), 0, a.stream>>>(                                               \
        a.labels, a.cg, a.seed32 ^ BOND_KEY, a.seed32 ^ ROOT_SALT, a.seed32 ^ SEED_SALT, skey, a.sweep, a.bond_t, \
        a.q, a.forced, a.queue, a.qcap < WOLFF_QUEUE ? a.qcap : WOLFF_QUEUE, a.level_cap, a.adaptive, a.gain_ticks, \
        a.state, a.partials, a.nparts)
    switch (a.cg.D) {
        case 1: PLQ_WG(1); break;
        case 2: PLQ_WG(2); break;
        case 3: PLQ_WG(3); break;
        default: PLQ_WG(4); break;
    }
#undef PLQ_WG
}

void launch_wolff_restore(uint8_t *dst, const int32_t *queue, const WolffState *state, hipStream_t s) {
    wolff_restore<<<dim3(256), dim3(256), 0, s>>>(dst, queue, state);
}

}  // namespace plq
