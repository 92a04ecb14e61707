// plq_cluster.hip -- Swendsen-Wang / Wolff cluster sweep on gfx950.
//
// The reference builds clusters on the device by repeated min-label
// propagation with a host read-back every four passes
// (cluster_prelude.wgsl:61-147, gpu_cluster.rs:492-526).  Its fixed point
// gives every site the minimum site index of its cluster, and the relabel
// draw is keyed on that root (cluster_relabel.wgsl:39-90).  Any algorithm
// reaching the same partition and the same root reproduces the reference's
// labels bit for bit, so this file uses union-find instead:
//
//   1. cc_local_x16 (power-of-two boxes) or cc_local: one workgroup per box
//                 of the lattice: draw the box's internal bonds
//                 (cluster_prelude.wgsl:61-80; bond = site*D + axis, open
//                 iff the labels agree and m < T_p) and union their ends in
//                 LDS, then store each site's box-local root as a global
//                 site index.
//   2. cc_boundary one lane per box-face site: draw the bonds that leave a
//                 box and union their ends in the global forest with
//                 compare-and-swap.
//   3. cc_compress one lane per box-face site again: point every local root
//                 a face bond joined straight at its cluster root.
//   4. cc_relabel_box (the boxes of cc_local_x16) or cc_relabel: the
//                 cluster root is two loads away; draw the cluster's new
//                 label keyed on it (Redraw or ForcedChange; for Wolff only
//                 the seed site's cluster) and store it.
//
// Every union links the larger root under the smaller one, so parent[x] <= x
// always holds and the root of a tree is the minimum site index in it: the
// same root the reference's propagation converges to (and the first-site
// order the CPU Forest compacts by, cluster.rs:196-226).  Every loop here
// walks strictly decreasing indices or retries only after another lane made
// progress, so each terminates; an explicit hop budget (the pass cap of
// gpu_cluster.rs:620-625) additionally lets the host force the
// non-convergence error.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

__device__ __forceinline__ int32_t load_relaxed(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A Wolff growth (plq_wolff.hip) gates the kernels of its sweep: each returns
// at once when *gate == skip (a null gate never skips).
__device__ __forceinline__ bool gated(const int *gate, int skip) { return gate != nullptr && *gate == skip; }

// ---------------------------------------------------------------- local ---

__device__ __forceinline__ int32_t lds_load(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x with path splitting: every node on the walk is re-pointed at its
// grandparent (an ancestor, so parent <= child still holds under concurrent
// unions); the grandparent just read is the next node's parent.
// Word k of a lane's 16 LDS words (16-byte aligned), through the 128-bit load that holds it
// (k is a constant once the caller's loop is unrolled).
__device__ __forceinline__ int32_t lane_word(const int32_t *P, uint32_t l0, int k, int4 &q) {
    if ((k & 3) == 0) q = reinterpret_cast<const int4 *>(P + l0)[k >> 2];
    return (k & 3) == 0 ? q.x : (k & 3) == 1 ? q.y : (k & 3) == 2 ? q.z : q.w;
}

// p: x's parent as read earlier (any ancestor of x will do).
__device__ __forceinline__ int32_t lds_find_from(int32_t *P, int32_t x, int32_t p) {
    while (p != x) {
        const int32_t gp = lds_load(&P[p]);
        if (gp != p) lds_store(&P[x], gp);
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ int32_t lds_find(int32_t *P, int32_t x) { return lds_find_from(P, x, lds_load(&P[x])); }

__device__ __forceinline__ void lds_union(int32_t *P, int32_t a, int32_t b) {
    for (;;) {
        a = lds_find(P, a);
        b = lds_find(P, b);
        if (a == b) return;
        int32_t hi = a > b ? a : b, lo = a > b ? b : a;
        int32_t old = atomicCAS(&P[hi], hi, lo);
        if (old == hi) return;
        a = old;  // hi gained a parent meanwhile; continue from it
        b = lo;
    }
}

// n / d for n < 2^24 through a 64-bit reciprocal: q = (n * M) >> 40 with
// M = ceil(2^40 / d) is exact while n * d < 2^40.

// POW2: every box has the same power-of-two extents, so a local index splits
// into box coordinates by shifts and masks instead of reciprocal divisions.
template <int D, bool POW2>
__global__ __launch_bounds__(1024) void cc_local(const uint8_t *__restrict__ labels, int32_t *__restrict__ parent,
                                                 CCGeom cg, uint32_t bond_seed, uint32_t skey, uint32_t bond_t,
                                                 const int *__restrict__ gate, int gate_skip) {
    if (gated(gate, gate_skip)) return;
    extern __shared__ int32_t P[];
    uint32_t o[MAX_D], e[MAX_D], ls[MAX_D];
    uint64_t me[MAX_D];
    uint32_t rest = blockIdx.x;
    uint32_t nloc = 1;
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        const uint32_t nb = (uint32_t)cg.nbox[mu];
        const uint32_t q = rest / nb;  // once per block
        const uint32_t b = rest - q * nb;
        rest = q;
        o[mu] = b * (uint32_t)cg.B[mu];
        const uint32_t left = (uint32_t)cg.L[mu] - o[mu];
        e[mu] = left < (uint32_t)cg.B[mu] ? left : (uint32_t)cg.B[mu];
        me[mu] = ((uint64_t(1) << 40) + e[mu] - 1) / e[mu];
        ls[mu] = nloc;
        nloc *= e[mu];
    }
    // local index -> (box coordinates, global site)
    uint32_t lsh[MAX_D];
    if (POW2) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) lsh[mu] = (uint32_t)(31 - __clz((int)ls[mu]));
    }
    auto locate = [&](uint32_t l, uint32_t *y) -> uint32_t {
        uint32_t r = l, g = 0;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            if (POW2) {
                y[mu] = (l >> lsh[mu]) & (e[mu] - 1u);
            } else {
                uint32_t q = mu + 1 < D ? cdiv(r, me[mu]) : 0u;
                y[mu] = mu + 1 < D ? r - q * e[mu] : r;
                r = q;
            }
            g += (o[mu] + y[mu]) * (uint32_t)cg.stride[mu];
        }
        return g;
    };
    // Axis 0 by a wave scan: lanes hold consecutive local indices, so a run of
    // open axis-0 bonds is a run of lanes; every site points at the first site
    // of its run (parent <= child, as every union keeps), with no atomics.
    // Bonds across a row end (x = e0 - 1) are cut here; the wrap bond of a box
    // that spans axis 0 and the link between two waves of one run are joined
    // below with the other axes' bonds.
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
    for (uint32_t base = threadIdx.x - lane; base < nloc; base += blockDim.x) {
        const uint32_t l = base + lane;
        bool open = false;
        if (l < nloc) {
            uint32_t y[MAX_D];
            const uint32_t gsite = locate(l, y);
            if (y[0] + 1u < e[0] && l + 1u < nloc) {
                const uint8_t lab = labels[gsite];
                open = labels[gsite + 1u] == lab &&
                       keyed_bits_k(bond_seed, gsite * (uint32_t)D, skey) < bond_t;
            }
        }
        const uint64_t m = __ballot(open);
        const uint64_t starts = (~m << 1) | 1ull;  // bit j: lane j begins a run
        const uint32_t first = 63u - (uint32_t)__clzll(starts & below);
        if (l < nloc) P[l] = (int32_t)(base + first);
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < nloc; l += blockDim.x) {
        uint32_t y[MAX_D];
        const uint32_t gsite = locate(l, y);
        const uint8_t lab = labels[gsite];
        // the axis-0 link from the last lane of a wave to the next wave
        if ((l & 63u) == 63u && y[0] + 1u < e[0] && l + 1u < nloc && labels[gsite + 1u] == lab &&
            keyed_bits_k(bond_seed, gsite * (uint32_t)D, skey) < bond_t)
            lds_union(P, (int32_t)l, (int32_t)(l + 1u));
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            uint32_t lp, gp;
            if (mu > 0 && y[mu] + 1u < e[mu]) {
                lp = l + ls[mu];
                gp = gsite + (uint32_t)cg.stride[mu];
            } else if (y[mu] + 1u == e[mu] && e[mu] == (uint32_t)cg.L[mu]) {  // the box spans the axis: the wrap bond is internal
                lp = l - (e[mu] - 1u) * ls[mu];
                gp = gsite - (uint32_t)(cg.L[mu] - 1) * (uint32_t)cg.stride[mu];
            } else {
                continue;  // axis 0 inside the box (the scan), or a bond leaving the box (cc_boundary)
            }
            if (labels[gp] != lab) continue;
            const uint32_t bond = gsite * (uint32_t)D + (uint32_t)mu;
            if (keyed_bits_k(bond_seed, bond, skey) < bond_t) lds_union(P, (int32_t)l, (int32_t)lp);
        }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < nloc; l += blockDim.x) {
        const uint32_t root = (uint32_t)lds_find(P, (int32_t)l);
        uint32_t y[MAX_D], yr[MAX_D];
        const uint32_t gsite = locate(l, y);
        const uint32_t groot = root == l ? gsite : locate(root, yr);
        parent[gsite] = (int32_t)groot;
    }
}

// Bit k (k < 4) set where byte k of a equals byte k of b.  Zero bytes of
// z = a ^ b have bit 7 clear in ((z & 0x7f..) + 0x7f..) | z; the flags at
// bits 0, 8, 16, 24 of f are gathered to bits 24..27 by one multiply with
// 2^24 + 2^17 + 2^10 + 2^3 (every other partial product lands below bit 24
// or above bit 31, on a position no other one takes, so nothing carries).
__device__ __forceinline__ uint32_t equal_bytes4(uint32_t a, uint32_t b) {
    const uint32_t z = a ^ b;
    const uint32_t f = (~(((z & 0x7f7f7f7fu) + 0x7f7f7f7fu) | z) >> 7) & 0x01010101u;
    return (f * 0x01020408u) >> 24;
}

__device__ __forceinline__ uint32_t equal_bytes16(const uint4 &a, const uint4 &b) {
    return equal_bytes4(a.x, b.x) | (equal_bytes4(a.y, b.y) << 4) | (equal_bytes4(a.z, b.z) << 8) |
           (equal_bytes4(a.w, b.w) << 12);
}

// Geometry of the boxes cc_local_x16 labels (every extent a power of two,
// box extent along axis 0 a multiple of 16): box origin, extents, and the
// bit offsets of the local coordinates (lsh) and of the lattice ones (Lsh).
template <int D>
struct X16Box {
    uint32_t o[MAX_D], e[MAX_D], lsh[MAX_D], Lsh[MAX_D];
    uint32_t nloc;
    __device__ __forceinline__ X16Box(const CCGeom &cg, uint32_t box) {
        uint32_t lg = 0, Lg = 0, rest = box;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const uint32_t nb = (uint32_t)cg.nbox[mu];
            const uint32_t q = rest / nb;
            e[mu] = (uint32_t)cg.B[mu];
            o[mu] = (rest - q * nb) * e[mu];
            rest = q;
            lsh[mu] = lg;
            lg += (uint32_t)(31 - __clz((int)e[mu]));
            Lsh[mu] = Lg;
            Lg += (uint32_t)(31 - __clz(cg.L[mu]));
        }
        nloc = 1u << lg;
    }
    // global site of box-local index l
    __device__ __forceinline__ uint32_t locate(const CCGeom &cg, uint32_t l) const {
        uint32_t g = 0;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) g += (o[mu] + ((l >> lsh[mu]) & (e[mu] - 1u))) * (uint32_t)cg.stride[mu];
        return g;
    }
};

// Membership of a global site in a box of power-of-two extents, with few
// scalar operands: the coordinate bits above the box extents (mask out)
// are the box origin's (g0) exactly for the box's sites, and the box-local
// index gathers each coordinate's low bits (shifted right by sh[mu], masked
// by msk[mu]).
template <int D>
struct BoxMember {
    uint32_t g0, out, sh[MAX_D], msk[MAX_D];
    __device__ __forceinline__ BoxMember(const CCGeom &cg, const X16Box<D> &bx) {
        g0 = 0;
        out = 0;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            g0 |= bx.o[mu] << bx.Lsh[mu];
            out |= (((uint32_t)cg.L[mu] - 1u) & ~(bx.e[mu] - 1u)) << bx.Lsh[mu];
            sh[mu] = bx.Lsh[mu] - bx.lsh[mu];
            msk[mu] = (bx.e[mu] - 1u) << bx.lsh[mu];
        }
    }
    // whether site p lies in the box, and its box-local index if so
    __device__ __forceinline__ bool local(uint32_t p, uint32_t &l) const {
        l = 0;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) l |= (p >> sh[mu]) & msk[mu];
        return ((p ^ g0) & out) == 0u;
    }
};

// The box labelling of cc_local, sixteen sites per lane: lane c owns the
// box-local indices 16c .. 16c+15, one segment of an axis-0 row.  `me` holds
// the lane's 16 labels; up(l, g) returns the 16 labels at box-local index l
// (global site g) of the row one step up, next(l, g) the label after the
// segment.  All threads of the workgroup call it.
//   1. Bond bits: equal bytes and the keyed draws give a 16-bit mask of open
//      bonds per axis.  P[16c + k] is the first site of k's run of open
//      axis-0 bonds inside the segment, and the segment's axis-0 mask goes
//      to H[c].
//   2. Unions, after a barrier, between run starts: the axis-0 bond out of
//      the segment's last site, and each open bond of the other axes unless
//      the bond one site back joins the same two runs (the runs of the row
//      one step up are read from its segment's H).
//   3. After a barrier, the roots of the run starts; every site of a run
//      stores the run's root as a global index (four 16-byte stores).
// The draws are the reference's (bond = site * D + axis); for D = 1, 2, 4
// the site index of a segment's first site times D is a multiple of 16 D,
// so the bond index is that base or-ed with k D + axis, and the first
// xor-shift of the hash is done once per segment.
template <int D, class Up, class Next>
__device__ __forceinline__ void x16_label(int32_t *P, uint16_t *H, int32_t *__restrict__ parent, const CCGeom &cg,
                                          const X16Box<D> &bx, uint32_t bond_seed, uint32_t skey, uint32_t bond_t,
                                          const uint4 &me, Up up, Next next) {
    const uint32_t c = threadIdx.x;
    const bool act = c < (bx.nloc >> 4);
    const uint32_t l0 = c << 4;
    uint32_t y[MAX_D];
    uint32_t gsite0 = 0;
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        y[mu] = (l0 >> bx.lsh[mu]) & (bx.e[mu] - 1u);
        gsite0 += (bx.o[mu] + y[mu]) * (uint32_t)cg.stride[mu];
    }
    const uint32_t kmix = sweep_mix(skey);
    constexpr bool XOR_BONDS = (D & (D - 1)) == 0;
    const uint32_t sb = bond_seed ^ (gsite0 * (uint32_t)D);
    const uint32_t v0 = sb ^ (sb >> 16);
    // the 16 bonds of axis mu: bit k set iff open (labels agree and m < t)
    auto draw16 = [&](uint32_t eq, int mu) -> uint32_t {
        uint32_t open = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t m = XOR_BONDS
                                   ? lowbias32_tail(lowbias32_mid(v0 ^ (uint32_t)(k * D + mu)) ^ kmix) >> 8
                                   : keyed_bits_k(bond_seed, (gsite0 + (uint32_t)k) * (uint32_t)D + (uint32_t)mu, skey);
            open |= (m < bond_t ? 1u : 0u) << k;
        }
        return open & eq;
    };
    uint32_t h = 0, starts = 1;
    uint32_t vmask[MAX_D];
    uint32_t nbr_l0[MAX_D];
    bool x_out = false;
    uint32_t x_to = 0;
    if (act) {
        // axis 0: site k to k + 1; site 15 to the next segment (or the row's
        // first segment when the box spans axis 0), when that stays in the box
        const bool x_in = y[0] + 16u < bx.e[0];
        const bool x_wrap = !x_in && bx.e[0] == (uint32_t)cg.L[0];
        x_out = x_in || x_wrap;
        x_to = x_in ? l0 + 16u : l0 + 16u - bx.e[0];
        const uint32_t nb = x_out ? next(x_to, x_in ? gsite0 + 16u : gsite0 + 16u - (uint32_t)cg.L[0]) : 0u;
        uint4 nx;
        nx.x = __builtin_amdgcn_alignbyte(me.y, me.x, 1);
        nx.y = __builtin_amdgcn_alignbyte(me.z, me.y, 1);
        nx.z = __builtin_amdgcn_alignbyte(me.w, me.z, 1);
        nx.w = __builtin_amdgcn_alignbyte(nb, me.w, 1);
        uint32_t eq0 = equal_bytes16(me, nx);
        if (!x_out) eq0 &= 0x7fffu;
        h = draw16(eq0, 0);
        starts = ((~h << 1) | 1u) & 0xffffu;  // bit k: site k begins a run inside the segment
#pragma unroll
        for (int mu = 1; mu < D; ++mu) {
            vmask[mu] = 0;
            uint32_t gn;
            if (y[mu] + 1u < bx.e[mu]) {
                nbr_l0[mu] = l0 + (1u << bx.lsh[mu]);
                gn = gsite0 + (uint32_t)cg.stride[mu];
            } else if (bx.e[mu] == (uint32_t)cg.L[mu]) {  // the box spans the axis: the wrap bond is internal
                nbr_l0[mu] = l0 - ((bx.e[mu] - 1u) << bx.lsh[mu]);
                gn = gsite0 - (uint32_t)(cg.L[mu] - 1) * (uint32_t)cg.stride[mu];
            } else {
                continue;  // bonds leaving the box (cc_boundary)
            }
            vmask[mu] = draw16(equal_bytes16(me, up(nbr_l0[mu], gn)), mu);
        }
    }
    if (act) {
        H[c] = (uint16_t)(h & 0x7fffu);
        // the lane's 16 words as four 128-bit stores (P is 16-byte aligned);
        // 32-bit stores at the 16-word lane stride meet on a few LDS banks
        int32_t run[16];
        int32_t r = (int32_t)l0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if ((starts >> k) & 1u) r = (int32_t)(l0 + (uint32_t)k);
            run[k] = r;
        }
        int4 *own = reinterpret_cast<int4 *>(P + l0);
#pragma unroll
        for (int d = 0; d < 4; ++d) own[d] = make_int4(run[4 * d], run[4 * d + 1], run[4 * d + 2], run[4 * d + 3]);
    }
    __syncthreads();
    if (act) {
        if (x_out && ((h >> 15) & 1u)) lds_union(P, (int32_t)(l0 + 31u - (uint32_t)__clz((int)starts)), (int32_t)x_to);
#pragma unroll
        for (int mu = 1; mu < D; ++mu) {
            uint32_t v = vmask[mu];
            if (!v) continue;
            const uint32_t sn = ((~(uint32_t)H[nbr_l0[mu] >> 4] << 1) | 1u) & 0xffffu;  // the neighbour's run starts
            // drop a bond that continues both runs of the bond one site back:
            // that bond joins the same two runs
            uint32_t need = v & ~((v << 1) & ~starts & ~sn);
            while (need) {
                const uint32_t k = (uint32_t)__builtin_ctz(need);
                need &= need - 1u;
                const uint32_t le = (2u << k) - 1u;
                const uint32_t a = 31u - (uint32_t)__clz((int)(starts & le));
                const uint32_t b = 31u - (uint32_t)__clz((int)(sn & le));
                lds_union(P, (int32_t)(l0 + a), (int32_t)(nbr_l0[mu] + b));
            }
        }
    }
    __syncthreads();
    if (act) {
        // the run starts' parents in four 128-bit loads; no union runs in
        // this phase, so a parent read here is an ancestor with the same root
        int4 q4;
        int32_t out[16];
        int32_t g = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int32_t p0 = lane_word(P, l0, k, q4);
            if ((starts >> k) & 1u) {
                const uint32_t r = (uint32_t)lds_find_from(P, (int32_t)(l0 + (uint32_t)k), p0);
                g = (int32_t)(r == l0 + (uint32_t)k ? gsite0 + (uint32_t)k : bx.locate(cg, r));
            }
            out[k] = g;
        }
        int4 *dst = reinterpret_cast<int4 *>(parent + gsite0);
        dst[0] = make_int4(out[0], out[1], out[2], out[3]);
        dst[1] = make_int4(out[4], out[5], out[6], out[7]);
        dst[2] = make_int4(out[8], out[9], out[10], out[11]);
        dst[3] = make_int4(out[12], out[13], out[14], out[15]);
    }
}

// cc_local for power-of-two boxes whose extent along axis 0 is a multiple of
// 16: x16_label with the labels read from global memory.
template <int D>
__global__ __launch_bounds__(1024) void cc_local_x16(const uint8_t *__restrict__ labels, int32_t *__restrict__ parent,
                                                     CCGeom cg, uint32_t bond_seed, uint32_t skey, uint32_t bond_t,
                                                     const int *__restrict__ gate, int gate_skip, RecordFold fold) {
    if (gated(gate, gate_skip)) return;
    extern __shared__ int4 box_lds[];  // its own name: extern LDS arrays of one name share one alignment
    int32_t *P = reinterpret_cast<int32_t *>(box_lds);
    __shared__ uint16_t H[1024];
    if (fold.rec) {
        // the previous recorded sweep's record, folded here instead of in a launch of its own
        __shared__ long long red[16];
        fold_record(fold, red);
    }
    const X16Box<D> bx(cg, blockIdx.x);
    uint4 me = make_uint4(0u, 0u, 0u, 0u);
    if (threadIdx.x < (bx.nloc >> 4)) me = *reinterpret_cast<const uint4 *>(labels + bx.locate(cg, threadIdx.x << 4));
    x16_label<D>(
        P, H, parent, cg, bx, bond_seed, skey, bond_t, me,
        [&](uint32_t, uint32_t g) { return *reinterpret_cast<const uint4 *>(labels + g); },
        [&](uint32_t, uint32_t g) { return (uint32_t)labels[g]; });
}

// ------------------------------------------------------------- boundary ---

__device__ __forceinline__ int32_t g_find(int32_t *P, int32_t x, uint32_t cap, uint32_t &hops) {
    int32_t p = load_relaxed(&P[x]);
    while (p != x) {
        if (++hops > cap) return x;
        int32_t gp = load_relaxed(&P[p]);
        if (gp != p) store_relaxed(&P[x], gp);
        x = p;
        p = load_relaxed(&P[x]);
    }
    return x;
}

// Root of x in a forest no lane modifies during the kernel: plain (cached)
// loads, no path compression.
__device__ __forceinline__ int32_t find_ro(const int32_t *__restrict__ P, int32_t x, uint32_t cap, uint32_t &hops) {
    int32_t p = P[x];
    while (p != x) {
        if (++hops > cap) return x;
        x = p;
        p = P[x];
    }
    return x;
}

// Face site t of the box faces: along axis mu, the last coordinate of each
// box (only where the box does not span the axis), each layer holding
// N / L[mu] sites.  Sets the site, its forward neighbour across the face and
// whether their bond is open; false past the last face site.
template <int D>
__device__ __forceinline__ bool face_bond(const uint8_t *__restrict__ labels, const CCGeom &cg, int64_t t,
                                          uint32_t bond_seed, uint32_t skey, uint32_t bond_t, int64_t &gsite,
                                          int64_t &gp) {
    int mu = 0;
    int64_t layer_sites = 0, count = 0;
    for (; mu < D; ++mu) {
        if (cg.B[mu] >= cg.L[mu]) continue;
        layer_sites = cg.N / cg.L[mu];
        count = (int64_t)cg.nbox[mu] * layer_sites;
        if (t < count) break;
        t -= count;
    }
    if (mu >= D) return false;
    int32_t k = (int32_t)(t / layer_sites);
    int64_t rem = t - (int64_t)k * layer_sites;  // index over the other axes
    int32_t xm = (k + 1) * cg.B[mu];
    if (xm > cg.L[mu]) xm = cg.L[mu];
    xm -= 1;
    // rebuild the site: rem enumerates the other axes in lattice order
    gsite = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        if (a == mu) {
            gsite += (int64_t)xm * cg.stride[a];
        } else {
            int32_t ya = (int32_t)(rem % cg.L[a]);
            rem /= cg.L[a];
            gsite += (int64_t)ya * cg.stride[a];
        }
    }
    gp = (xm + 1 == cg.L[mu]) ? gsite - (int64_t)xm * cg.stride[mu] : gsite + cg.stride[mu];
    return labels[gsite] == labels[gp] && keyed_bits_k(bond_seed, (uint32_t)(gsite * D + mu), skey) < bond_t;
}

constexpr int BND_SLOTS = 512;  // LDS pair set of a 256-lane boundary workgroup

// One lane per face site: the bonds that leave a box, joined in the global
// forest.  The walks start from the two box-local roots (the sites' parents
// as cc_local stored them), so a face site's own parent is never re-pointed
// and stays its local root, which cc_compress relies on.
template <int D>
__global__ __launch_bounds__(256) void cc_boundary(const uint8_t *__restrict__ labels, int32_t *__restrict__ parent,
                                                   CCGeom cg, uint32_t bond_seed, uint32_t skey, uint32_t bond_t,
                                                   uint32_t cap, int *__restrict__ flag, const int *__restrict__ gate,
                                                   int gate_skip) {
    if (gated(gate, gate_skip)) return;
    // The workgroup's distinct (parent, parent) pairs: face sites along a
    // face mostly pair the same two box-local roots, and a bond whose ends
    // have the same parents as another open bond joins the same two sets.
    // Only the lane that first enters a pair in this LDS set walks and links.
    __shared__ unsigned long long pairs[BND_SLOTS];
    for (int i = threadIdx.x; i < BND_SLOTS; i += blockDim.x) pairs[i] = ~0ull;
    __syncthreads();
    int64_t gsite = 0, gp = 0;
    const bool open = face_bond<D>(labels, cg, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, bond_seed, skey,
                                   bond_t, gsite, gp);
    if (!open) return;
    const int32_t pa = load_relaxed(&parent[gsite]);
    const int32_t pb = load_relaxed(&parent[gp]);
    const unsigned long long key = ((unsigned long long)(uint32_t)pa << 32) | (uint32_t)pb;
    uint32_t h = (uint32_t)((key * 0x9e3779b97f4a7c15ull) >> 55) & (BND_SLOTS - 1);
    for (int probe = 0; probe < 8; ++probe, h = (h + 1u) & (BND_SLOTS - 1)) {
        const unsigned long long old = atomicCAS(&pairs[h], ~0ull, key);
        if (old == ~0ull) break;  // first: this lane links the pair
        if (old == key) return;   // another lane links it
    }                             // (a crowded probe links it again: harmless)
    uint32_t hops = 1;  // the hop from each site to its local root
    int32_t a = pa, b = pb;
    for (;;) {
        a = g_find(parent, a, cap, hops);
        b = g_find(parent, b, cap, hops);
        if (hops > cap) {
            atomicOr(flag, 1);
            return;
        }
        if (a == b) return;
        int32_t hi = a > b ? a : b, lo = a > b ? b : a;
        int32_t old = atomicCAS(&parent[hi], hi, lo);
        if (old == hi) return;
        if (++hops > cap) {
            atomicOr(flag, 1);
            return;
        }
        a = old;
        b = lo;
    }
}

// After cc_boundary: every box-local root that a face bond joined is pointed
// straight at its cluster root, so that every site is two loads from its
// cluster root (parent[parent[s]]).  Such a local root x first met the
// forest through an open face bond with an end s whose parent was x (s = x,
// or a site of x's component).  Each open face bond's lane reads the parents
// of its two ends, walks to the root (read only: the forest no longer
// changes, and a path split could store a non-root ancestor over another
// lane's store) and stores the root into both ends and both parents.  Every
// store writes the final root of that cluster.  Of the lanes storing into an
// end s, the first to read reads s's original parent, so x is reached: for
// s = x directly, otherwise through s's parent.  With cross set, cross[root]
// = tag marks the cluster as one that crosses a box face (cc_sizes_box).
template <int D>
__global__ __launch_bounds__(256) void cc_compress(const uint8_t *__restrict__ labels, int32_t *__restrict__ parent,
                                                   CCGeom cg, uint32_t bond_seed, uint32_t skey, uint32_t bond_t,
                                                   uint32_t cap, int *__restrict__ flag, uint8_t *__restrict__ cross,
                                                   uint8_t tag, const int *__restrict__ gate, int gate_skip) {
    if (gated(gate, gate_skip)) return;
    int64_t gsite = 0, gp = 0;
    if (!face_bond<D>(labels, cg, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, bond_seed, skey, bond_t, gsite,
                      gp))
        return;
    const int32_t pa = load_relaxed(&parent[gsite]);
    const int32_t pb = load_relaxed(&parent[gp]);
    uint32_t hops = 1;
    int32_t root = pa;
    for (int32_t up = load_relaxed(&parent[root]); up != root; up = load_relaxed(&parent[root])) {
        if (++hops > cap) {
            atomicOr(flag, 1);
            return;
        }
        root = up;
    }
    if (cross) cross[root] = tag;  // the cluster crosses a face (cc_sizes_box)
    if (pa != root) {
        store_relaxed(&parent[gsite], root);
        store_relaxed(&parent[pa], root);
    }
    if (pb != root) {
        store_relaxed(&parent[gp], root);
        store_relaxed(&parent[pb], root);
    }
}

// -------------------------------------------------------------- relabel ---

// Block totals of the first nv (uniform) of NV per-lane values, for blocks
// of at most 4 waves with every lane active: DPP wave sums into LDS, one
// barrier, then thread k adds value k over the waves and returns it (threads
// k >= nv return 0).  Each lane's value and each wave's sum fit in an int.
constexpr int BLK_WAVES = 4;
template <int NV>
__device__ __forceinline__ long long blk_sums(const int (&v)[NV], int nv, int (*wsum)[NV]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (k < nv) {
            const int s = wave_sum_i32(v[k]);
            if (lane == 0) wsum[wave][k] = s;
        }
    }
    __syncthreads();
    long long t = 0;
    if ((int)threadIdx.x < nv)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += wsum[w][threadIdx.x];
    return t;
}

// The new label of every site, four sites per lane (one dword of labels, one
// int4 of parents).  A site's parent is its box-local root (cc_local stores
// it; cc_boundary and cc_compress re-point only local roots), and after
// cc_compress the local root's parent is the cluster root, the minimum site
// of the cluster, on which the label draw is keyed (cluster_relabel.wgsl:
// 39-90): Redraw takes floor(u q), ForcedChange adds 1 + floor(u (q - 1)) to
// the cluster's label (every site of a cluster holds that label, since only
// agreeing sites are bonded).  Seeded (Wolff): only the cluster of the seed
// site, drawn from the sweep's seed key (cluster_relabel.wgsl:59-62),
// changes.  The hop budget counts the hop from a site to its local root and
// the one above it.  The last block also takes the N mod 4 tail.  Partials
// per block: [0, relabelled sites, 0 ...].
constexpr uint32_t NO_LABEL = 0xffu;

// Number of zero bytes of x (exact: no borrow between bytes).
__device__ __forceinline__ int zero_bytes(uint32_t x) {
    uint32_t t = (x & 0x7f7f7f7fu) + 0x7f7f7f7fu;
    return __popc(~(t | x | 0x7f7f7f7fu));
}

// The seed site's cluster root (Wolff), for the relabel kernels: the seed
// site from the keyed draw, scaled in f32 and clamped
// (cluster_relabel.wgsl:59-62).
__device__ __forceinline__ int32_t seed_cluster_root(const int32_t *__restrict__ parent, int64_t N, uint32_t seed_seed,
                                                     uint32_t sweep, uint32_t cap, int *__restrict__ flag) {
    const float u = keyed_uniform_f32(keyed_bits(seed_seed, 0u, sweep));
    uint32_t ss = (uint32_t)(u * (float)(uint32_t)N);
    if (ss >= (uint32_t)N) ss = (uint32_t)N - 1u;
    uint32_t hops = 0;
    const int32_t r = find_ro(parent, (int32_t)ss, cap, hops);
    if (hops > cap) atomicOr(flag, 1);
    return r;
}

__global__ __launch_bounds__(256) void cc_relabel(uint8_t *__restrict__ labels, const int32_t *__restrict__ parent,
                                                  int64_t N, uint32_t root_seed, uint32_t seed_seed, uint32_t skey,
                                                  uint32_t sweep, int q, int seeded, int forced, uint32_t cap,
                                                  int *__restrict__ flag, long long *__restrict__ partials,
                                                  int64_t span, const int *__restrict__ gate, int gate_skip) {
    if (gated(gate, gate_skip)) return;
    __shared__ int32_t seed_root;
    if (seeded) {
        if (threadIdx.x == 0) seed_root = seed_cluster_root(parent, N, seed_seed, sweep, cap, flag);
        __syncthreads();
    }
    const int32_t sroot = seeded ? seed_root : -1;
    bool over = false;
    // new label of a site of label lab in the cluster of root r (NO_LABEL: unchanged)
    auto relabel = [&](int32_t r, uint32_t lab) -> uint32_t {
        if (seeded && r != sroot) return NO_LABEL;
        return relabel_draw(keyed_bits_k(root_seed, (uint32_t)r, skey), lab, q, forced);
    };
    long long moved = 0;
    const int64_t nd = N >> 2;
    int64_t begin = (int64_t)blockIdx.x * span;
    int64_t end = begin + span;
    if (end > nd) end = nd;
    uint32_t *w32 = reinterpret_cast<uint32_t *>(labels);
    const int4 *p4 = reinterpret_cast<const int4 *>(parent);
    // After cc_compress a site's cluster root is parent[parent[s]]: one
    // gather per distinct parent in the dword.
    for (int64_t v = begin + threadIdx.x; v < end; v += blockDim.x) {
        const int4 p = p4[v];
        const uint32_t w = w32[v];
        int32_t r[4];
        r[0] = parent[p.x];
        r[1] = p.y != p.x ? parent[p.y] : -1;
        r[2] = p.z != p.y ? parent[p.z] : -1;
        r[3] = p.w != p.z ? parent[p.w] : -1;
        if (p.y == p.x) r[1] = r[0];
        if (p.z == p.y) r[2] = r[1];
        if (p.w == p.z) r[3] = r[2];
        // the hop budget: one hop to the local root, one more above it
        const bool deeper = r[0] != p.x || r[1] != p.y || r[2] != p.z || r[3] != p.w;
        over = over || (deeper ? 2u : 1u) > cap;
        uint32_t out = w;
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t nl = relabel(r[k], (w >> (8 * k)) & 0xffu);
            if (nl != NO_LABEL) {
                out = (out & ~(0xffu << (8 * k))) | (nl << (8 * k));
                ++cnt;
            }
        }
        if (cnt) w32[v] = out;
        moved += cnt;
    }
    if (blockIdx.x == gridDim.x - 1) {
        for (int64_t s = (nd << 2) + threadIdx.x; s < N; s += blockDim.x) {
            const int32_t lr = parent[s];
            const int32_t r = parent[lr];
            over = over || (r != lr ? 2u : 1u) > cap;
            const uint32_t nl = relabel(r, labels[s]);
            if (nl != NO_LABEL) {
                labels[s] = (uint8_t)nl;
                moved += 1;
            }
        }
    }
    if (over) atomicOr(flag, 1);
    __shared__ int wsum[BLK_WAVES][1];
    const int mv[1] = {(int)moved};
    const long long t = blk_sums<1>(mv, 1, wsum);
    if (threadIdx.x == 0) {
        long long *dst = partials + (int64_t)blockIdx.x * REC_K;
        dst[0] = 0;
        dst[1] = t;
        for (int k = 2; k < REC_K; ++k) dst[k] = 0;
    }
}

// Relabel of one box, for the boxes cc_local_x16 labels: the box's parents
// are loaded into LDS (P) with coalesced 16-byte loads, and a site's parent
// p is then either a site of the box, whose own parent (from LDS) is the
// cluster root after cc_compress, or a cluster root outside the box that
// cc_compress stored into a face site.  No global gathers.  Lane c takes the
// 16 sites of one axis-0 row segment (its labels in w, updated in place);
// consecutive sites with the same parent share the draw.  Returns the
// relabelled sites; the caller synchronises before P is reused.
//
// A lane's own 16 words of P move as four 128-bit LDS accesses (P is
// declared 16-byte aligned and l0 is a multiple of 16): 32-bit accesses at
// the 16-word lane stride would meet on a few LDS banks.
template <int D>
__device__ __forceinline__ void box_load_parents(int32_t *P, const int32_t *__restrict__ parent, uint32_t l0,
                                                 uint32_t gsite0) {
    const int4 *src = reinterpret_cast<const int4 *>(parent + gsite0);
    int4 *dst = reinterpret_cast<int4 *>(P + l0);
#pragma unroll
    for (int d = 0; d < 4; ++d) dst[d] = src[d];
}

// The new labels of a lane's 16 sites (w, in place), the roots found as
// the sites are visited; returns the relabelled sites.  hops collects, as
// bits, the lengths of the walks taken (1: the parent is the root, 2: its
// parent is); the caller checks the longest against the hop budget once
// (an integer, not a lane mask held through the loop).
template <int D>
__device__ __forceinline__ int box_relabel16(const int32_t *P, const BoxMember<D> &bm, uint32_t l0,
                                             uint32_t (&w)[4], uint32_t root_seed, uint32_t skey, int q, int seeded,
                                             int forced, int32_t seed_root, uint32_t &hops) {
    int32_t prev = -1;
    uint32_t nl = NO_LABEL;
    int cnt = 0;
    int4 q4;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int32_t p = lane_word(P, l0, k, q4);
        if (p != prev) {
            prev = p;
            // p inside the box: its parent, from LDS; else p is a root
            uint32_t local;
            const int32_t root = bm.local((uint32_t)p, local) ? P[local] : p;
            hops |= root != p ? 2u : 1u;
            const uint32_t lab = (w[k >> 2] >> (8 * (k & 3))) & 0xffu;
            nl = (seeded && root != seed_root) ? NO_LABEL
                                               : relabel_draw(keyed_bits_k(root_seed, (uint32_t)root, skey), lab, q, forced);
        }
        if (nl != NO_LABEL) {
            const int sh = 8 * (k & 3);
            w[k >> 2] = (w[k >> 2] & ~(0xffu << sh)) | (nl << sh);
            ++cnt;
        }
    }
    return cnt;
}

// Whether the walks box_relabel16 / box_roots16 recorded in hops exceed
// the hop budget.
__device__ __forceinline__ bool hops_over(uint32_t hops, uint32_t cap) { return (hops >= 2u ? 2u : hops) > cap; }

// The cluster roots of a lane's 16 sites, kept for the counting that
// follows (the relabel-and-count and sizes kernels).
template <int D>
__device__ __forceinline__ void box_roots16(const int32_t *P, const BoxMember<D> &bm, uint32_t l0,
                                            int32_t (&root)[16], uint32_t &hops) {
    int32_t prev = -1, r = -1;
    int4 q4;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int32_t p = lane_word(P, l0, k, q4);
        if (p != prev) {
            prev = p;
            // p inside the box: its parent, from LDS; else p is a root
            uint32_t local;
            r = bm.local((uint32_t)p, local) ? P[local] : p;
            hops |= r != p ? 2u : 1u;
        }
        root[k] = r;
    }
}

// The new labels of a lane's 16 sites (w, in place) from their roots
// (box_roots16); consecutive sites with the same root share the draw.
// Returns the relabelled sites.
__device__ __forceinline__ int box_relabel_roots16(const int32_t (&root)[16], uint32_t (&w)[4], uint32_t root_seed,
                                                   uint32_t skey, int q, int seeded, int forced, int32_t seed_root) {
    int32_t prev = -1;
    uint32_t nl = NO_LABEL;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (root[k] != prev) {
            prev = root[k];
            const uint32_t lab = (w[k >> 2] >> (8 * (k & 3))) & 0xffu;
            nl = (seeded && prev != seed_root)
                     ? NO_LABEL
                     : relabel_draw(keyed_bits_k(root_seed, (uint32_t)prev, skey), lab, q, forced);
        }
        if (nl != NO_LABEL) {
            const int sh = 8 * (k & 3);
            w[k >> 2] = (w[k >> 2] & ~(0xffu << sh)) | (nl << sh);
            ++cnt;
        }
    }
    return cnt;
}


// The moved count of the workgroup into its partials row [0, moved, 0 ...].
__device__ __forceinline__ void box_partials(long long moved, int *wsum, long long *__restrict__ partials) {
    const int wv = wave_sum_i32((int)moved);
    if ((threadIdx.x & 63u) == 0u) wsum[threadIdx.x >> 6] = wv;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += wsum[i];
        long long *dst = partials + (int64_t)blockIdx.x * REC_K;
        dst[0] = 0;
        dst[1] = t;
        for (int k = 2; k < REC_K; ++k) dst[k] = 0;
    }
}

// The relabel one box per workgroup (box_relabel16); a workgroup walks boxes
// blockIdx.x, blockIdx.x + gridDim.x, ...  Partials per block: [0,
// relabelled sites, 0 ...].
template <int D>
__global__ __launch_bounds__(1024, 8) void cc_relabel_box(uint8_t *__restrict__ labels,
                                                       const int32_t *__restrict__ parent, CCGeom cg,
                                                       uint32_t root_seed, uint32_t seed_seed, uint32_t skey,
                                                       uint32_t sweep, int q, int seeded, int forced, uint32_t cap,
                                                       int *__restrict__ flag, long long *__restrict__ partials,
                                                       const int *__restrict__ gate, int gate_skip) {
    if (gated(gate, gate_skip)) return;
    extern __shared__ int4 box_lds[];  // its own name: extern LDS arrays of one name share one alignment
    int32_t *P = reinterpret_cast<int32_t *>(box_lds);
    __shared__ int32_t seed_root;
    __shared__ int wsum[16];
    if (seeded && threadIdx.x == 0) seed_root = seed_cluster_root(parent, cg.N, seed_seed, sweep, cap, flag);
    const uint32_t c = threadIdx.x;
    const uint32_t l0 = c << 4;
    long long moved = 0;
    uint32_t hops = 0;
    for (int64_t box = blockIdx.x; box < cg.nboxes; box += gridDim.x) {
        const X16Box<D> bx(cg, (uint32_t)box);
        const bool act = c < (bx.nloc >> 4);
        const uint32_t gsite0 = bx.locate(cg, l0);
        __syncthreads();  // the previous box's readers are done with P
        if (act) box_load_parents<D>(P, parent, l0, gsite0);
        __syncthreads();
        if (act) {
            uint4 *dst = reinterpret_cast<uint4 *>(labels + gsite0);
            const uint4 l4 = *dst;
            uint32_t w[4] = {l4.x, l4.y, l4.z, l4.w};
            const int cnt =
                box_relabel16<D>(P, BoxMember<D>(cg, bx), l0, w, root_seed, skey, q, seeded, forced, seed_root, hops);
            if (cnt) *dst = make_uint4(w[0], w[1], w[2], w[3]);
            moved += cnt;
        }
    }
    if (hops_over(hops, cap)) atomicOr(flag, 1);
    box_partials(moved, wsum, partials);
}

// Bytes of the halo of forward neighbours outside a box that
// cc_relabel_box_meas stages in LDS (at most one per lane, and at most
// 3 nloc, so labels and halo fit the nloc parents' LDS).
constexpr int RELABEL_HALO_MAX = 1024;

// cc_relabel_box for a Swendsen-Wang sweep (every cluster redrawn, no seed
// cluster, no forced change, no gate), with the Potts energy sums of the new
// labels (potts_measure_vec's columns: agreeing forward bonds, label counts)
// counted in the same pass, so a recorded sweep reads its labels once.
// After the relabel the box's new labels (bytes 0 .. nloc) and the halo
// (from byte nloc) replace its parents in LDS.  A forward neighbour outside
// the box belongs to another workgroup's box, which may be relabelling it
// now, but its new label is fixed: its cluster root is parent[parent[s]]
// (the forest after cc_compress, which no kernel of the relabel writes) and
// the redraw is keyed on the root alone.  Those neighbours form the halo: per
// axis mu >= 1 with the box narrower than the lattice, the face beyond the
// box's last layer (nloc / e[mu] bytes, axis 0 contiguous inside it), then
// axis 0's (nloc / e[0] bytes, one per box row).  Partials per block: the
// relabel's row [0, moved, 0 ...] in partials, and the measure's columns
// [agree, 0, counts[0..q)] column-major in mpart (column k of block b at
// mpart[k * gridDim.x + b]).
//
// Halo index h: the global site of its neighbour, and its byte in LDS.
template <int D>
__device__ __forceinline__ uint32_t halo_site(const CCGeom &cg, const X16Box<D> &bx, uint32_t h) {
    uint32_t off = 0;
#pragma unroll
    for (int k = 1; k <= D; ++k) {
        const int mu = k < D ? k : 0;  // axes 1 .. D-1, then axis 0
        if (bx.e[mu] == (uint32_t)cg.L[mu]) continue;
        const uint32_t lg = (uint32_t)(31 - __clz((int)bx.e[mu]));
        const uint32_t n = bx.nloc >> lg;
        if (h < off + n) {
            const uint32_t f = h - off;
            const uint32_t lo = f & ((1u << bx.lsh[mu]) - 1u);
            const uint32_t l = lo | ((bx.e[mu] - 1u) << bx.lsh[mu]) | ((f >> bx.lsh[mu]) << (bx.lsh[mu] + lg));
            const uint32_t mu_out = (bx.o[mu] + bx.e[mu]) & ((uint32_t)cg.L[mu] - 1u);
            return bx.locate(cg, l) + (mu_out - (bx.o[mu] + bx.e[mu] - 1u)) * (uint32_t)cg.stride[mu];
        }
        off += n;
    }
    return 0;
}

// LDS byte offsets of the forward neighbours lane c's segment (box-local
// l0 .. l0 + 15) compares with: nx the byte after the segment along axis 0
// (the next segment's first, the row's first when the box is as wide as the
// lattice, or the row's halo byte), up[mu] the 16 bytes one step along axis
// mu >= 1 (16-byte aligned).
template <int D>
__device__ __forceinline__ void meas_offsets(const CCGeom &cg, const X16Box<D> &bx, uint32_t l0, uint32_t &nx,
                                             uint32_t (&up)[MAX_D]) {
    uint32_t hoff = bx.nloc;
#pragma unroll
    for (int mu = 1; mu < D; ++mu) {
        const uint32_t lg = (uint32_t)(31 - __clz((int)bx.e[mu]));
        const uint32_t y = (l0 >> bx.lsh[mu]) & (bx.e[mu] - 1u);
        if (y + 1u < bx.e[mu])
            up[mu] = l0 + (1u << bx.lsh[mu]);
        else if (bx.e[mu] == (uint32_t)cg.L[mu])
            up[mu] = l0 - (y << bx.lsh[mu]);
        else
            up[mu] = hoff + ((l0 & ((1u << bx.lsh[mu]) - 1u)) | ((l0 >> (bx.lsh[mu] + lg)) << bx.lsh[mu]));
        if (bx.e[mu] != (uint32_t)cg.L[mu]) hoff += bx.nloc >> lg;
    }
    const uint32_t lg0 = (uint32_t)(31 - __clz((int)bx.e[0]));
    const uint32_t x0 = l0 & (bx.e[0] - 1u);
    nx = x0 + 16u < bx.e[0] ? l0 + 16u : bx.e[0] == (uint32_t)cg.L[0] ? l0 + 16u - bx.e[0] : hoff + (l0 >> lg0);
}

template <int D>
__global__ __launch_bounds__(1024, 8) void cc_relabel_box_meas(uint8_t *__restrict__ labels,
                                                            const int32_t *__restrict__ parent, CCGeom cg,
                                                            uint32_t root_seed, uint32_t skey, int q, uint32_t cap,
                                                            int *__restrict__ flag, long long *__restrict__ partials,
                                                            long long *__restrict__ mpart, uint32_t nhalo) {
    extern __shared__ int4 box_lds[];
    int32_t *P = reinterpret_cast<int32_t *>(box_lds);
    uint8_t *R = reinterpret_cast<uint8_t *>(box_lds);  // after the relabel: new labels, then the halo
    __shared__ int wsum[16][3 + MAX_Q_DEV];
    const uint32_t c = threadIdx.x;
    const uint32_t l0 = c << 4;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    int moved = 0, agree = 0;
    if (lane == 0)
        for (int a = 0; a < q; ++a) wsum[wave][3 + a] = 0;  // each wave's label counts, summed over its boxes
    uint32_t hops = 0;
    for (int64_t box = blockIdx.x; box < cg.nboxes; box += gridDim.x) {
        // the box geometry is dead after the relabel: the count uses the
        // LDS offsets computed here
        const X16Box<D> bx(cg, (uint32_t)box);
        const bool act = c < (bx.nloc >> 4);
        const uint32_t gsite0 = bx.locate(cg, l0);
        // lanes past the halo load their own first site's parents instead:
        // no mask held over the relabel
        const uint32_t hs = c < nhalo ? halo_site<D>(cg, bx, c) : gsite0;
        const uint32_t hb = bx.nloc + c;
        uint32_t nx = 0, up[MAX_D] = {};
        meas_offsets<D>(cg, bx, l0, nx, up);
        __syncthreads();  // the previous box's readers are done with the LDS
        // the halo site's parent beside the box's parents, its root during
        // the relabel (a barrier waits for every load in flight)
        const int32_t hp = parent[hs];
        if (act) box_load_parents<D>(P, parent, l0, gsite0);
        __syncthreads();
        const int32_t hr = parent[hp];
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        if (act) {
            uint4 *dst = reinterpret_cast<uint4 *>(labels + gsite0);
            const uint4 l4 = *dst;
            w[0] = l4.x;
            w[1] = l4.y;
            w[2] = l4.z;
            w[3] = l4.w;
            moved += box_relabel16<D>(P, BoxMember<D>(cg, bx), l0, w, root_seed, skey, q, 0, 0, -1, hops);
            *dst = make_uint4(w[0], w[1], w[2], w[3]);
        }
        const uint32_t hl = relabel_draw(keyed_bits_k(root_seed, (uint32_t)hr, skey), 0u, q, 0);
        __syncthreads();  // every lane is done with P
        if (act) reinterpret_cast<uint4 *>(box_lds)[c] = make_uint4(w[0], w[1], w[2], w[3]);
        if (c < nhalo) R[hb] = (uint8_t)hl;
        __syncthreads();
        if (act) {
            const uint32_t n0 = R[nx];
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const uint32_t nxt = d < 3 ? w[d < 3 ? d + 1 : 0] : n0;
                agree += zero_bytes(w[d] ^ ((w[d] >> 8) | (nxt << 24)));
            }
#pragma unroll
            for (int mu = 1; mu < D; ++mu) {
                const uint4 u = *reinterpret_cast<const uint4 *>(R + up[mu]);
                agree += zero_bytes(w[0] ^ u.x) + zero_bytes(w[1] ^ u.y) + zero_bytes(w[2] ^ u.z) +
                         zero_bytes(w[3] ^ u.w);
            }
        }
        // label counts of the box: a wave sum per label, kept by lane 0 (a
        // loop that is not unrolled: unrolled, the per-label conditions
        // took scalar registers from the relabel)
#pragma unroll 1
        for (int a = 0; a < q; ++a) {
            int n = 0;
            if (act)
#pragma unroll
                for (int d = 0; d < 4; ++d) n += zero_bytes(w[d] ^ (0x01010101u * (uint32_t)a));
            const int s = wave_sum_i32(n);
            if (lane == 0) wsum[wave][3 + a] += s;
        }
    }
    if (hops_over(hops, cap)) atomicOr(flag, 1);
    // block sums: value 0 the moved sites, 1 + k measure column k
    const int nv = 3 + q;
    {
        const int s0 = wave_sum_i32(moved), s1 = wave_sum_i32(agree);
        if (lane == 0) {
            wsum[wave][0] = s0;
            wsum[wave][1] = s1;
            wsum[wave][2] = 0;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += wsum[i][threadIdx.x];
        if (threadIdx.x == 0) {
            long long *dst = partials + (int64_t)blockIdx.x * REC_K;
            dst[0] = 0;
            dst[1] = t;
            for (int k = 2; k < REC_K; ++k) dst[k] = 0;
        } else {
            mpart[(int64_t)(threadIdx.x - 1) * gridDim.x + blockIdx.x] = t;
        }
    }
}


// --------------------------------------------------------- measurement ---

// Potts energy sums of a lattice-ordered label field: agreeing forward bonds
// (potts/mod.rs:237-271) and label counts (potts/mod.rs:124-139).
template <int D>
__global__ __launch_bounds__(256) void potts_measure_kernel(const uint8_t *__restrict__ labels, CCGeom cg, int q,
                                                            long long *__restrict__ partials, int64_t span) {
    long long agree = 0;
    int counts[MAX_Q_DEV];
    for (int a = 0; a < MAX_Q_DEV; ++a) counts[a] = 0;
    int64_t begin = (int64_t)blockIdx.x * span;
    int64_t end = begin + span;
    if (end > cg.N) end = cg.N;
    for (int64_t s64 = begin + threadIdx.x; s64 < end; s64 += blockDim.x) {
        const uint32_t s = (uint32_t)s64;
        const uint8_t lab = labels[s];
        counts[lab] += 1;
        uint32_t rest = s;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const uint32_t q = mu + 1 < D ? (cg.mul_div ? cdiv(rest, cg.mL[mu]) : rest / (uint32_t)cg.L[mu]) : 0u;
            const uint32_t x = mu + 1 < D ? rest - q * (uint32_t)cg.L[mu] : rest;
            rest = q;
            const uint32_t st = (uint32_t)cg.stride[mu];
            const uint32_t gp = (x + 1u == (uint32_t)cg.L[mu]) ? s - x * st : s + st;
            agree += labels[gp] == lab;
        }
    }
    __shared__ int wsum[BLK_WAVES][2 + MAX_Q_DEV];
    int vals[2 + MAX_Q_DEV];
    vals[0] = (int)agree;
    vals[1] = 0;
#pragma unroll
    for (int a = 0; a < MAX_Q_DEV; ++a) vals[2 + a] = counts[a];
    const long long t = blk_sums<2 + MAX_Q_DEV>(vals, 2 + q, wsum);
    if (threadIdx.x < (unsigned)REC_K) partials[(int64_t)blockIdx.x * REC_K + threadIdx.x] = t;
}

// The same sums 4 W sites per lane and step, for L0 a multiple of 4 W (and,
// for W = 4, 16-byte aligned labels): a step's labels never straddle a row;
// agreement along axis 0 compares each byte with the next (the row's first
// byte after its last), along the other axes with the labels one stride on.
// Column-major partials: column k of block b at part[k * gridDim.x + b],
// columns [agree, 0, counts[0..q)].
template <int D, int W>
__global__ __launch_bounds__(256) void potts_measure_vec(const uint8_t *__restrict__ labels, CCGeom cg, int q,
                                                         long long *__restrict__ part, int64_t span) {
    constexpr uint32_t S = 4u * W;
    const int64_t nu = cg.N / S;
    long long agree = 0;
    int counts[MAX_Q_DEV];
#pragma unroll
    for (int a = 0; a < MAX_Q_DEV; ++a) counts[a] = 0;
    int64_t begin = (int64_t)blockIdx.x * span;
    int64_t end = begin + span;
    if (end > nu) end = nu;
    const uint32_t L0 = (uint32_t)cg.L[0];
    auto load = [&](uint32_t site, uint32_t (&w)[W]) {
        if constexpr (W == 4) {
            const uint4 l4 = *reinterpret_cast<const uint4 *>(labels + site);
            w[0] = l4.x;
            w[1] = l4.y;
            w[2] = l4.z;
            w[3] = l4.w;
        } else {
#pragma unroll
            for (int d = 0; d < W; ++d) w[d] = reinterpret_cast<const uint32_t *>(labels)[(site >> 2) + d];
        }
    };
    for (int64_t v = begin + threadIdx.x; v < end; v += blockDim.x) {
        const uint32_t s0 = (uint32_t)v * S;
        uint32_t w[W];
        load(s0, w);
        uint32_t rest = s0;
        uint32_t x[MAX_D];
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const uint32_t qt = mu + 1 < D ? (cg.mul_div ? cdiv(rest, cg.mL[mu]) : rest / (uint32_t)cg.L[mu]) : 0u;
            x[mu] = mu + 1 < D ? rest - qt * (uint32_t)cg.L[mu] : rest;
            rest = qt;
        }
        const uint32_t nb = x[0] + S == L0 ? labels[s0 - x[0]] : labels[s0 + S];
        int a0 = 0;
#pragma unroll
        for (int d = 0; d < W; ++d) {
            const uint32_t nxt = d + 1 < W ? w[d + 1 < W ? d + 1 : 0] : nb;
            a0 += zero_bytes(w[d] ^ ((w[d] >> 8) | (nxt << 24)));
        }
#pragma unroll
        for (int mu = 1; mu < D; ++mu) {
            const uint32_t st = (uint32_t)cg.stride[mu];
            const uint32_t gp = (x[mu] + 1u == (uint32_t)cg.L[mu]) ? s0 - x[mu] * st : s0 + st;
            uint32_t u[W];
            load(gp, u);
#pragma unroll
            for (int d = 0; d < W; ++d) a0 += zero_bytes(w[d] ^ u[d]);
        }
        agree += a0;
#pragma unroll
        for (int a = 0; a < MAX_Q_DEV; ++a)
            if (a < q)
#pragma unroll
                for (int d = 0; d < W; ++d) counts[a] += zero_bytes(w[d] ^ (0x01010101u * (uint32_t)a));
    }
    const int64_t G = gridDim.x;
    __shared__ int wsum[BLK_WAVES][2 + MAX_Q_DEV];
    int vals[2 + MAX_Q_DEV];
    vals[0] = (int)agree;
    vals[1] = 0;
#pragma unroll
    for (int a = 0; a < MAX_Q_DEV; ++a) vals[2 + a] = counts[a];
    const long long t = blk_sums<2 + MAX_Q_DEV>(vals, 2 + q, wsum);
    if ((int)threadIdx.x < 2 + q) part[(int64_t)threadIdx.x * G + blockIdx.x] = t;
}

// ---------------------------------------------------- roots and sizes ---

__global__ void cc_resolve(int32_t *__restrict__ parent, int32_t *__restrict__ roots, int64_t N, uint32_t cap,
                           int *__restrict__ flag) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    uint32_t hops = 0;
    int32_t r = find_ro(parent, (int32_t)s, cap, hops);
    if (hops > cap) atomicOr(flag, 1);
    roots[s] = r;
}

// Cluster sizes.  A block takes CC_SPAN consecutive sites, a wave 64 at a
// time (coalesced); each run of equal roots among a wave's 64 sites is one
// add into an LDS table keyed by root (open addressing, linear probing), and
// the table goes to the global counts with one atomic per root it holds.  A
// cluster spanning many blocks so costs one global atomic per block instead
// of one per run; a run whose probe finds the table crowded goes straight to
// the global count.
constexpr int CC_SLOTS = 4096;
constexpr int64_t CC_SPAN = 16384;
constexpr int CC_PROBES = 8;

__global__ __launch_bounds__(256) void cc_count(const int32_t *__restrict__ roots, int32_t *__restrict__ sizes,
                                                int64_t N) {
    __shared__ int32_t key[CC_SLOTS];
    __shared__ int32_t cnt[CC_SLOTS];
    for (int i = threadIdx.x; i < CC_SLOTS; i += blockDim.x) {
        key[i] = -1;
        cnt[i] = 0;
    }
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * CC_SPAN;
    const int64_t b1 = b0 + CC_SPAN < N ? b0 + CC_SPAN : N;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t above = lane == 63u ? 0ull : (~0ull << (lane + 1));
    for (int64_t base = b0 + (threadIdx.x - lane); base < b1; base += blockDim.x) {
        const int64_t s = base + lane;
        const bool in = s < b1;
        const int32_t r = in ? roots[s] : -1;
        const int32_t prev = __shfl_up(r, 1, 64);
        const bool head = in && (lane == 0u || prev != r);
        // run length: up to the next head, or the end of the wave's sites
        const uint64_t heads = __ballot(head) | (__ballot(!in) & ~0ull);
        if (!head) continue;
        const uint64_t later = heads & above;
        const uint32_t stop = later ? (uint32_t)__ffsll((unsigned long long)later) - 1u : 64u;
        const int run = (int)(stop - lane);
        uint32_t h = ((uint32_t)r * 0x9e3779b1u) >> (32 - 12);
        bool done = false;
        for (int p = 0; p < CC_PROBES && !done; ++p, h = (h + 1u) & (CC_SLOTS - 1)) {
            int32_t k = key[h];
            if (k == -1) {
                const int32_t old = atomicCAS(&key[h], -1, r);
                k = old == -1 ? r : old;
            }
            if (k == r) {
                atomicAdd(&cnt[h], run);
                done = true;
            }
        }
        if (!done) atomicAdd(&sizes[r], run);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CC_SLOTS; i += blockDim.x)
        if (key[i] != -1) atomicAdd(&sizes[key[i]], cnt[i]);
}

// hist[0..39] clusters by floor(log2 size); hist[40] clusters; hist[41]
// largest.  A grid-stride loop over the roots; per block, LDS bins, wave
// sums for the count and a wave maximum for the largest size, then one
// global atomic per bin.
__global__ __launch_bounds__(256) void cc_hist(const int32_t *__restrict__ roots, const int32_t *__restrict__ sizes,
                                               int64_t N, unsigned long long *__restrict__ hist) {
    __shared__ unsigned int h[40];
    __shared__ unsigned int nclu, largest;
    if (threadIdx.x < 40) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        nclu = 0;
        largest = 0;
    }
    __syncthreads();
    unsigned int mine = 0, big = 0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += (int64_t)gridDim.x * blockDim.x) {
        if (roots[s] == (int32_t)s) {
            const uint32_t sz = (uint32_t)sizes[s];
            atomicAdd(&h[31 - __clz(sz)], 1u);
            ++mine;
            big = sz > big ? sz : big;
        }
    }
    const int tot = wave_sum_i32((int)mine);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned int other = (unsigned int)__shfl_xor((int)big, o, 64);
        big = other > big ? other : big;
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&nclu, (unsigned int)tot);
        atomicMax(&largest, big);
    }
    __syncthreads();
    if (threadIdx.x < 40 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    if (threadIdx.x == 0) {
        atomicAdd(&hist[40], (unsigned long long)nclu);
        atomicMax(&hist[41], (unsigned long long)largest);
    }
}

// A workgroup's histogram as one row of HIST_K partials (bins, count,
// largest), summed by cc_hist_reduce: with thousands of workgroups, atomics
// on the few hot bins would serialise at their cache line.
constexpr int HIST_K = 42;
// The workgroups' histogram rows are stored column-major, entry c of row r
// at part[c * HIST_ROWS + r], so that cc_hist_reduce reads each column
// contiguously; the statistics kernels of one call write at most
// 2 * MAX_PASS_BLOCKS rows.
constexpr int64_t HIST_ROWS = 2 * (int64_t)MAX_PASS_BLOCKS;

__device__ __forceinline__ void hist_row(const unsigned int *h, unsigned int nclu, unsigned int largest,
                                         unsigned long long *part, int64_t row) {
    if (threadIdx.x < 40) part[threadIdx.x * HIST_ROWS + row] = h[threadIdx.x];
    if (threadIdx.x == 0) {
        part[40 * HIST_ROWS + row] = nclu;
        part[41 * HIST_ROWS + row] = largest;
    }
}

// Column c of `rows` partial rows into out[c]: the sum, the maximum for the
// largest size (column 41).  One workgroup per column; workgroup HIST_K sums
// the relabelled-site column of the relabel's partials (REC_K per row, `rrows`
// rows) into out[HIST_K + 1] and clears out[HIST_K .. HIST_K + REC_K).
__global__ __launch_bounds__(256) void cc_hist_reduce(const unsigned long long *__restrict__ part, int64_t rows,
                                                      const long long *__restrict__ rpart, int64_t rrows,
                                                      unsigned long long *__restrict__ out) {
    __shared__ unsigned long long acc[4];
    const int c = blockIdx.x;
    unsigned long long v = 0;
    if (c == HIST_K) {
        if (threadIdx.x < REC_K && threadIdx.x != 1) out[HIST_K + threadIdx.x] = 0;
        for (int64_t r = threadIdx.x; r < rrows; r += blockDim.x) v += (unsigned long long)rpart[r * REC_K + 1];
    } else {
        for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) {
            const unsigned long long x = part[c * HIST_ROWS + r];
            v = c == 41 ? (x > v ? x : v) : v + x;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long y = __shfl_xor(v, o, 64);
        v = c == 41 ? (y > v ? y : v) : v + y;
    }
    if ((threadIdx.x & 63u) == 0u) acc[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = acc[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = c == 41 ? (acc[w] > t ? acc[w] : t) : t + acc[w];
        out[c == HIST_K ? HIST_K + 1 : c] = t;
    }
}

// Per-lane share of a size histogram: the eight smallest bins, the cluster
// count and the largest size in registers, larger bins in the workgroup's
// LDS bins h; flush() adds the lane shares to h, nclu and largest by wave
// sums (all lanes of the workgroup call it), so hot counters see one atomic
// per wave instead of one per cluster.
struct LaneHist {
    static constexpr int NB = 8;
    int small[NB];
    unsigned int count, big;
    __device__ __forceinline__ LaneHist() : count(0), big(0) {
#pragma unroll
        for (int k = 0; k < NB; ++k) small[k] = 0;
    }
    __device__ __forceinline__ void add(uint32_t sz, unsigned int *h) {
        const int bin = 31 - __clz(sz);
        if (bin < NB) {
#pragma unroll
            for (int k = 0; k < NB; ++k) small[k] += bin == k ? 1 : 0;
        } else {
            atomicAdd(&h[bin], 1u);
        }
        ++count;
        big = sz > big ? sz : big;
    }
    __device__ __forceinline__ void flush(unsigned int *h, unsigned int *nclu, unsigned int *largest) {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int t = wave_sum_i32(small[k]);
            if ((threadIdx.x & 63u) == 0u && t) atomicAdd(&h[k], (unsigned int)t);
        }
        const int tot = wave_sum_i32((int)count);
        unsigned int m = big;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned int other = (unsigned int)__shfl_xor((int)m, o, 64);
            m = other > m ? other : m;
        }
        if ((threadIdx.x & 63u) == 0u) {
            atomicAdd(nclu, (unsigned int)tot);
            atomicMax(largest, m);
        }
    }
};

// Cluster sizes for the boxes cc_local_x16 labels, from the forest of a
// finished sweep (every site two loads from its cluster root: the box's
// parents in LDS as in cc_relabel_box).  Per box, the members of each run of
// equal roots in a lane's 16 sites are added at once: to an LDS count at the
// root's slot when the root lies in the box, else to an LDS table of outside
// roots that goes to ext[root] in global memory at the end of the box.  Every root's own site lies in its box, so each root's box stores
// its count in sizes[root]; a cluster's size is sizes[root] + ext[root].
// A workgroup walks boxes blockIdx.x, blockIdx.x + gridDim.x, ...
// LDS table of the roots outside a box: 8 sub-tables of 128 slots, a lane
// using sub-table lane & 7, so a root met by many lanes (the percolating
// cluster's) spreads its atomics over 8 slots
constexpr int SZ_SUB = 128, SZ_SLOTS = 8 * SZ_SUB;

// The counting of one box (all threads of the workgroup call it, after the
// roots are known): P becomes the box's per-slot counts, outside roots go
// through the LDS table okey / ocnt to ext, and each counted root either
// stores its count in sizes (when marked as crossing) or is binned in lh.
template <int D>
__device__ __forceinline__ void box_count16(int32_t *P, int32_t *okey, int32_t *ocnt, const CCGeom &cg,
                                            const X16Box<D> &bx, bool act, uint32_t l0, uint32_t gsite0,
                                            const int32_t (&root)[16], int32_t *__restrict__ sizes,
                                            int32_t *__restrict__ ext, const uint8_t *__restrict__ cross,
                                            uint8_t tag, LaneHist &lh, unsigned int *hb) {
    // box-local slot of global site g, or ~0u outside the box
    auto slot = [&](int32_t g) -> uint32_t {
        uint32_t local = 0;
        bool inside = true;
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const uint32_t y = (((uint32_t)g >> bx.Lsh[mu]) & ((uint32_t)cg.L[mu] - 1u)) - bx.o[mu];
            inside = inside && y < bx.e[mu];
            local |= y << bx.lsh[mu];
        }
        return inside ? local : ~0u;
    };
    __syncthreads();  // every read of the parents is done: P becomes the counts
    if (act) {
#pragma unroll
        for (int d = 0; d < 4; ++d) reinterpret_cast<int4 *>(P + l0)[d] = make_int4(0, 0, 0, 0);
    }
    __syncthreads();
    // n sites of the outside root r into the lane's sub-table
    auto outside_add = [&](int32_t r, int n) {
        const uint32_t sub = (threadIdx.x & 7u) * SZ_SUB;
        uint32_t h = ((uint32_t)r * 0x9e3779b1u) >> (32 - 7);
        bool done = false;
        for (int probe = 0; probe < 8 && !done; ++probe, h = (h + 1u) & (SZ_SUB - 1)) {
            int32_t key = okey[sub + h];
            if (key == -1) {
                const int32_t old = atomicCAS(&okey[sub + h], -1, r);
                key = old == -1 ? r : old;
            }
            if (key == r) {
                atomicAdd(&ocnt[sub + h], n);
                done = true;
            }
        }
        if (!done) atomicAdd(&ext[r], n);  // table crowded: straight to the global count
    };
    // Runs of one outside root (a percolating cluster's, in most boxes) are
    // summed in registers while they follow each other in the segment
    int32_t orr = -1;
    int olen = 0;
    if (act) {
        int len = 1;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k == 15 || root[k + 1] != root[k]) {
                const int32_t r = root[k];
                const uint32_t sl = slot(r);
                if (sl != ~0u) {
                    atomicAdd(&P[sl], len);
                } else if (r == orr) {
                    olen += len;
                } else {
                    if (olen) outside_add(orr, olen);
                    orr = r;
                    olen = len;
                }
                len = 1;
            } else {
                ++len;
            }
        }
    }
    {
        // then the wave's pending runs of its first pending root go in as one
        const uint64_t pend = __ballot(olen > 0);
        if (pend) {
            const int leader = __ffsll((unsigned long long)pend) - 1;
            const int32_t r0 = __shfl(orr, leader, 64);
            const bool same = olen > 0 && orr == r0;
            int v = same ? olen : 0;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (same) olen = 0;
            if ((int)(threadIdx.x & 63u) == leader) olen = v;
        }
        if (olen) outside_add(orr, olen);
    }
    __syncthreads();
    if (act) {
        // the segment's 16 crossing marks in one load
        const uint4 cm = *reinterpret_cast<const uint4 *>(cross + gsite0);
        const uint32_t t4 = 0x01010101u * tag;
        const uint32_t crossing = equal_bytes4(cm.x, t4) | (equal_bytes4(cm.y, t4) << 4) |
                                  (equal_bytes4(cm.z, t4) << 8) | (equal_bytes4(cm.w, t4) << 12);
        int4 q4;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int32_t n = lane_word(P, l0, k, q4);
            if (!n) continue;
            if ((crossing >> k) & 1u)
                sizes[gsite0 + (uint32_t)k] = n;  // cc_hist_cross adds the other boxes' share and bins it
            else
                lh.add((uint32_t)n, hb);
        }
    }
    // sub-tables 1..7 fold into sub-table 0, so a root goes to global
    // memory once per box (the percolating root's ext counter takes one
    // atomic per box, not eight)
    __syncthreads();
    for (int i = SZ_SUB + threadIdx.x; i < SZ_SLOTS; i += blockDim.x) {
        const int32_t r = okey[i];
        if (r == -1) continue;
        uint32_t h = ((uint32_t)r * 0x9e3779b1u) >> (32 - 7);
        bool done = false;
        for (int probe = 0; probe < 8 && !done; ++probe, h = (h + 1u) & (SZ_SUB - 1)) {
            int32_t key = okey[h];
            if (key == -1) {
                const int32_t old = atomicCAS(&okey[h], -1, r);
                key = old == -1 ? r : old;
            }
            if (key == r) {
                atomicAdd(&ocnt[h], ocnt[i]);
                done = true;
            }
        }
        if (!done) atomicAdd(&ext[r], ocnt[i]);
        okey[i] = -1;
        ocnt[i] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SZ_SUB; i += blockDim.x) {
        if (okey[i] != -1) {
            atomicAdd(&ext[okey[i]], ocnt[i]);
            okey[i] = -1;
            ocnt[i] = 0;
        }
    }
}

template <int D>
__global__ __launch_bounds__(1024, 8) void cc_sizes_box(const int32_t *__restrict__ parent, CCGeom cg,
                                                     int32_t *__restrict__ sizes, int32_t *__restrict__ ext,
                                                     const uint8_t *__restrict__ cross, uint8_t tag,
                                                     unsigned long long *__restrict__ hist) {  // hist: this block's row
    extern __shared__ int4 box_lds[];  // its own name: extern LDS arrays of one name share one alignment
    int32_t *P = reinterpret_cast<int32_t *>(box_lds);
    __shared__ int32_t okey[SZ_SLOTS], ocnt[SZ_SLOTS];
    __shared__ unsigned int hb[40];
    __shared__ unsigned int nclu, largest;
    if (threadIdx.x < 40) hb[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        nclu = 0;
        largest = 0;
    }
    for (int i = threadIdx.x; i < SZ_SLOTS; i += blockDim.x) {
        okey[i] = -1;
        ocnt[i] = 0;
    }
    const uint32_t c = threadIdx.x;
    const uint32_t l0 = c << 4;
    LaneHist lh;
    uint32_t hops = 0;  // not checked: the partition is final
    for (int64_t box = blockIdx.x; box < cg.nboxes; box += gridDim.x) {
        const X16Box<D> bx(cg, (uint32_t)box);
        const bool act = c < (bx.nloc >> 4);
        const uint32_t gsite0 = bx.locate(cg, l0);
        __syncthreads();  // the previous box's counts are out
        if (act) box_load_parents<D>(P, parent, l0, gsite0);
        __syncthreads();
        int32_t root[16];
        if (act) box_roots16<D>(P, BoxMember<D>(cg, bx), l0, root, hops);
        box_count16<D>(P, okey, ocnt, cg, bx, act, l0, gsite0, root, sizes, ext, cross, tag, lh, hb);
    }
    lh.flush(hb, &nclu, &largest);
    __syncthreads();
    hist_row(hb, nclu, largest, hist, blockIdx.x);
}

// One sweep's relabel (as cc_relabel_box) and the counting of its clusters
// (as cc_sizes_box) in one pass over the box's parents, for a recorded sweep
// whose cluster statistics are kept.
template <int D>
__global__ __launch_bounds__(1024, 8) void cc_relabel_count(uint8_t *__restrict__ labels,
                                                         const int32_t *__restrict__ parent, CCGeom cg,
                                                         uint32_t root_seed, uint32_t seed_seed, uint32_t skey,
                                                         uint32_t sweep, int q, int seeded, int forced,
                                                         uint32_t cap, int *__restrict__ flag,
                                                         long long *__restrict__ partials,
                                                         int32_t *__restrict__ sizes, int32_t *__restrict__ ext,
                                                         const uint8_t *__restrict__ cross, uint8_t tag,
                                                         unsigned long long *__restrict__ hist) {
    extern __shared__ int4 box_lds[];  // its own name: extern LDS arrays of one name share one alignment
    int32_t *P = reinterpret_cast<int32_t *>(box_lds);
    __shared__ int32_t okey[SZ_SLOTS], ocnt[SZ_SLOTS];
    __shared__ unsigned int hb[40];
    __shared__ unsigned int nclu, largest;
    __shared__ int32_t seed_root;
    __shared__ int wsum[16];
    if (threadIdx.x < 40) hb[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        nclu = 0;
        largest = 0;
        if (seeded) seed_root = seed_cluster_root(parent, cg.N, seed_seed, sweep, cap, flag);
    }
    for (int i = threadIdx.x; i < SZ_SLOTS; i += blockDim.x) {
        okey[i] = -1;
        ocnt[i] = 0;
    }
    const uint32_t c = threadIdx.x;
    const uint32_t l0 = c << 4;
    LaneHist lh;
    long long moved = 0;
    uint32_t hops = 0;
    for (int64_t box = blockIdx.x; box < cg.nboxes; box += gridDim.x) {
        const X16Box<D> bx(cg, (uint32_t)box);
        const bool act = c < (bx.nloc >> 4);
        const uint32_t gsite0 = bx.locate(cg, l0);
        __syncthreads();  // the previous box's counts are out
        if (act) box_load_parents<D>(P, parent, l0, gsite0);
        __syncthreads();
        int32_t root[16];
        if (act) {
            box_roots16<D>(P, BoxMember<D>(cg, bx), l0, root, hops);
            uint4 *dst = reinterpret_cast<uint4 *>(labels + gsite0);
            const uint4 l4 = *dst;
            uint32_t w[4] = {l4.x, l4.y, l4.z, l4.w};
            const int cnt = box_relabel_roots16(root, w, root_seed, skey, q, seeded, forced, seed_root);
            if (cnt) *dst = make_uint4(w[0], w[1], w[2], w[3]);
            moved += cnt;
        }
        box_count16<D>(P, okey, ocnt, cg, bx, act, l0, gsite0, root, sizes, ext, cross, tag, lh, hb);
    }
    if (hops_over(hops, cap)) atomicOr(flag, 1);
    lh.flush(hb, &nclu, &largest);
    __syncthreads();
    hist_row(hb, nclu, largest, hist, blockIdx.x);
    box_partials(moved, wsum, partials);
}

// The crossing clusters' share of the histogram: every site marked with
// this sweep's tag that is a root (parent[s] == s) has size sizes[s] +
// ext[s]; ext[s] is cleared for the next call (only crossing roots receive
// ext counts).  A mark left by an older sweep with the same tag only sends
// a cluster through this path, where its size is still right.  Sixteen
// marks per lane and step, visited in a loop over the matching ones.
__global__ __launch_bounds__(256) void cc_hist_cross(const int32_t *__restrict__ parent,
                                                     const int32_t *__restrict__ sizes, int32_t *__restrict__ ext,
                                                     const uint8_t *__restrict__ cross, uint8_t tag, int64_t N,
                                                     unsigned long long *__restrict__ hist, int64_t row0) {
    __shared__ unsigned int h[40];
    __shared__ unsigned int nclu, largest;
    if (threadIdx.x < 40) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        nclu = 0;
        largest = 0;
    }
    __syncthreads();
    LaneHist lh;
    // the three loads of a visit are independent, so they overlap
    auto visit = [&](int64_t s) {
        const int32_t p = parent[s], x = ext[s], z = sizes[s];
        if (p != (int32_t)s) return;
        if (x) ext[s] = 0;
        lh.add((uint32_t)(z + x), h);
    };
    const uint32_t t4 = 0x01010101u * tag;
    const int64_t n16 = N >> 4;
    const uint4 *c16 = reinterpret_cast<const uint4 *>(cross);
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n16; v += (int64_t)gridDim.x * blockDim.x) {
        const uint4 m = c16[v];
        // the marked sites of the 16 as a mask, then one visit per mark: a
        // wave loops as often as its busiest lane has marks
        uint32_t hit = equal_bytes4(m.x, t4) | (equal_bytes4(m.y, t4) << 4) | (equal_bytes4(m.z, t4) << 8) |
                       (equal_bytes4(m.w, t4) << 12);
        while (hit) {
            const uint32_t k = (uint32_t)__builtin_ctz(hit);
            hit &= hit - 1u;
            visit((v << 4) + k);
        }
    }
    for (int64_t s = (n16 << 4) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N;
         s += (int64_t)gridDim.x * blockDim.x)
        if (cross[s] == tag) visit(s);
    lh.flush(h, &nclu, &largest);
    __syncthreads();
    hist_row(h, nclu, largest, hist, row0 + blockIdx.x);
}

// ------------------------------------------------------------ launchers ---

// Boxes: start from the whole lattice and halve the longest box edge until a
// box fits the LDS budget; a 2D 4096^2 lattice gets 128 x 128 boxes, so only
// 2 x 128 of each box's 2 x 16384 forward bonds leave it.
CCGeom make_ccgeom(int D, const int64_t *shape, int32_t max_box_sites) {
    CCGeom cg{};
    cg.D = D;
    cg.N = 1;
    for (int mu = 0; mu < MAX_D; ++mu) {
        cg.L[mu] = mu < D ? (int32_t)shape[mu] : 1;
        cg.stride[mu] = cg.N;
        if (mu < D) cg.N *= shape[mu];
        cg.B[mu] = cg.L[mu];
    }
    for (;;) {
        int64_t vol = 1;
        int big = 0;
        for (int mu = 0; mu < D; ++mu) {
            vol *= cg.B[mu];
            if (cg.B[mu] > cg.B[big]) big = mu;
        }
        if (vol <= max_box_sites) break;
        cg.B[big] = (cg.B[big] + 1) / 2;
    }
    int64_t maxL = 1;
    for (int mu = 0; mu < D; ++mu) maxL = cg.L[mu] > maxL ? cg.L[mu] : maxL;
    cg.mul_div = (double)cg.N * (double)maxL < 1099511627776.0 ? 1 : 0;  // 2^40
    cg.nboxes = 1;
    cg.box_sites = 1;
    for (int mu = 0; mu < MAX_D; ++mu) {
        cg.mL[mu] = ((uint64_t(1) << 40) + (uint64_t)cg.L[mu] - 1) / (uint64_t)cg.L[mu];
        cg.nbox[mu] = (cg.L[mu] + cg.B[mu] - 1) / cg.B[mu];
        cg.nboxes *= cg.nbox[mu];
        cg.box_sites *= cg.B[mu];
    }
    return cg;
}

static int64_t face_sites(const CCGeom &cg) {
    int64_t total = 0;
    for (int mu = 0; mu < cg.D; ++mu)
        if (cg.B[mu] < cg.L[mu]) total += (int64_t)cg.nbox[mu] * (cg.N / cg.L[mu]);
    return total;
}

// Halo bytes of cc_relabel_box_meas: the faces beyond each box along the
// axes where the box is narrower than the lattice.
static uint32_t relabel_halo_sites(const CCGeom &cg) {
    uint32_t n = 0;
    for (int mu = 0; mu < cg.D; ++mu)
        if (cg.B[mu] < cg.L[mu]) n += (uint32_t)(cg.box_sites / cg.B[mu]);
    return n;
}

int64_t launch_cluster_sweep(const ClusterSweepArgs &a) {
    const CCGeom &cg = a.cg;
    uint32_t skey = a.sweep * GOLDEN;
    uint32_t bond_seed = a.seed32 ^ BOND_KEY;
    size_t lds = (size_t)cg.box_sites * sizeof(int32_t);
    dim3 gl((uint32_t)cg.nboxes), bl(1024);
    int64_t faces = face_sites(cg);
    bool pow2 = true;
    for (int mu = 0; mu < cg.D; ++mu)
        pow2 = pow2 && (cg.L[mu] & (cg.L[mu] - 1)) == 0 && (cg.B[mu] & (cg.B[mu] - 1)) == 0;
    dim3 gb((uint32_t)((faces + 255) / 256)), bb(256);
    // sixteen sites per lane: whole 16-site segments of axis-0 rows, one per lane
    const bool x16 = cluster_boxes_x16(cg, a.labels);
    // an earlier record's fold rides on cc_local_x16 when it has a workgroup
    // per column and no gate; else it is launched first
    RecordFold fold = a.fold;
    if (fold.rec && !(x16 && a.gate == nullptr && cg.nboxes >= fold.nk)) {
        launch_finalize_columns(fold.pa, fold.na, fold.pb, fold.nb, fold.nk, fold.rec, a.stream);
        fold = RecordFold{};
    }
#define PLQ_CC(DD)                                                                                          \
    if (x16)                                                                                                \
        cc_local_x16<DD><<<gl, bl, lds, a.stream>>>(a.labels, a.parent, cg, bond_seed, skey, a.bond_t,      \
                                                    a.gate, a.gate_skip, fold);                              \
    else if (pow2)                                                                                          \
        cc_local<DD, true><<<gl, bl, lds, a.stream>>>(a.labels, a.parent, cg, bond_seed, skey, a.bond_t,    \
                                                      a.gate, a.gate_skip);                                  \
    else                                                                                                    \
        cc_local<DD, false><<<gl, bl, lds, a.stream>>>(a.labels, a.parent, cg, bond_seed, skey, a.bond_t,   \
                                                       a.gate, a.gate_skip);                                 \
    if (faces > 0) {                                                                                        \
        cc_boundary<DD><<<gb, bb, 0, a.stream>>>(a.labels, a.parent, cg, bond_seed, skey, a.bond_t, a.cap, \
                                                 a.flag, a.gate, a.gate_skip);                              \
        cc_compress<DD><<<gb, bb, 0, a.stream>>>(a.labels, a.parent, cg, bond_seed, skey, a.bond_t, a.cap, \
                                                 a.flag, a.cross, cross_tag(a.sweep), a.gate, a.gate_skip); \
    }
    switch (cg.D) {
        case 1: PLQ_CC(1); break;
        case 2: PLQ_CC(2); break;
        case 3: PLQ_CC(3); break;
        default: PLQ_CC(4); break;
    }
#undef PLQ_CC
    if (a.label_only) return 0;
    int64_t need = (cg.N + 255) / 256;
    int64_t G = need < a.max_blocks ? need : a.max_blocks;
    if (x16 && a.count_sizes) {
        const int64_t Gb = cg.nboxes < a.max_blocks ? cg.nboxes : a.max_blocks;
        const uint8_t tag = cross_tag(a.sweep);
        switch (cg.D) {
#define PLQ_RC(DD)                                                                                           \
    cc_relabel_count<DD><<<dim3((uint32_t)Gb), bl, lds, a.stream>>>(                                       \
        a.labels, a.parent, cg, a.seed32 ^ ROOT_SALT, a.seed32 ^ SEED_SALT, skey, a.sweep, a.q, a.seeded,   \
        a.forced, a.cap, a.flag, a.partials, a.sizes, a.ext, a.cross, tag, a.hist_part)
            case 1: PLQ_RC(1); break;
            case 2: PLQ_RC(2); break;
            case 3: PLQ_RC(3); break;
            default: PLQ_RC(4); break;
#undef PLQ_RC
        }
        return Gb;
    }
    if (x16 && a.meas_rec && cluster_relabel_measures(a)) {
        const int64_t Gb = cg.nboxes < a.max_blocks ? cg.nboxes : a.max_blocks;
        const uint32_t nhalo = relabel_halo_sites(cg);
        switch (cg.D) {
#define PLQ_RM(DD)                                                                                              \
    cc_relabel_box_meas<DD><<<dim3((uint32_t)Gb), bl, lds, a.stream>>>(a.labels, a.parent, cg, a.seed32 ^ ROOT_SALT, \
                                                                     skey, a.q, a.cap, a.flag, a.partials,        \
                                                                     a.meas_part, nhalo)
            case 1: PLQ_RM(1); break;
            case 2: PLQ_RM(2); break;
            case 3: PLQ_RM(3); break;
            default: PLQ_RM(4); break;
#undef PLQ_RM
        }
        if (!a.defer_fold) launch_finalize_columns(a.meas_part, Gb, nullptr, 0, 2 + a.q, a.meas_rec, a.stream);
        return Gb;
    }
    if (x16) {
        const int64_t Gb = cg.nboxes < a.max_blocks ? cg.nboxes : a.max_blocks;
        switch (cg.D) {
#define PLQ_RB(DD)                                                                                            \
    cc_relabel_box<DD><<<dim3((uint32_t)Gb), bl, lds, a.stream>>>(a.labels, a.parent, cg, a.seed32 ^ ROOT_SALT, \
                                                                a.seed32 ^ SEED_SALT, skey, a.sweep, a.q,     \
                                                                a.seeded, a.forced, a.cap, a.flag, a.partials, \
                                                                a.gate, a.gate_skip)
            case 1: PLQ_RB(1); break;
            case 2: PLQ_RB(2); break;
            case 3: PLQ_RB(3); break;
            default: PLQ_RB(4); break;
#undef PLQ_RB
        }
        return Gb;
    }
    const int64_t nd = cg.N >> 2;
    const int64_t span4 = nd > 0 ? (nd + G - 1) / G : 1;
    cc_relabel<<<dim3((uint32_t)G), dim3(256), 0, a.stream>>>(a.labels, a.parent, cg.N, a.seed32 ^ ROOT_SALT,
                                                              a.seed32 ^ SEED_SALT, skey, a.sweep, a.q, a.seeded,
                                                              a.forced, a.cap, a.flag, a.partials, span4, a.gate,
                                                              a.gate_skip);
    return G;
}

bool cluster_relabel_measures(const ClusterSweepArgs &a) {
    const uint32_t nhalo = relabel_halo_sites(a.cg);
    return !a.label_only && !a.count_sizes && !a.seeded && !a.forced && a.gate == nullptr && a.q >= 1 &&
           a.q <= MAX_Q_DEV && cluster_boxes_x16(a.cg, a.labels) && nhalo <= (uint32_t)RELABEL_HALO_MAX &&
           nhalo <= 3u * (uint32_t)a.cg.box_sites;
}

RecordFold cluster_record_fold(const ClusterSweepArgs &a, int64_t blocks) {
    RecordFold f{};
    f.pa = a.meas_part;
    f.na = blocks;
    f.nk = 2 + a.q;
    f.rec = a.meas_rec;
    return f;
}

int64_t cluster_relabel_blocks(const ClusterSweepArgs &a) {
    const CCGeom &cg = a.cg;
    if (a.label_only) return 0;
    if (cluster_boxes_x16(cg, a.labels)) return cg.nboxes < a.max_blocks ? cg.nboxes : a.max_blocks;
    const int64_t need = (cg.N + 255) / 256;
    return need < a.max_blocks ? need : a.max_blocks;
}

void launch_potts_measure(const uint8_t *labels, const CCGeom &cg, int q, long long *partials, int64_t max_blocks,
                          long long *rec, hipStream_t s, RecordFold *defer) {
    if (defer) *defer = RecordFold{};
    if (cg.L[0] % 4 == 0) {
        // sixteen sites per lane and step where rows hold whole 16-byte vectors
        const bool wide = cg.L[0] % 16 == 0 && (reinterpret_cast<uintptr_t>(labels) & 15u) == 0;
        const int64_t nu = cg.N / (wide ? 16 : 4);
        int64_t need = (nu + 255) / 256;
        int64_t G = need < max_blocks ? need : max_blocks;
        int64_t span = (nu + G - 1) / G;
        dim3 g((uint32_t)G), b(256);
#define PLQ_PM(DD)                                                                           \
    if (wide)                                                                                \
        potts_measure_vec<DD, 4><<<g, b, 0, s>>>(labels, cg, q, partials, span);           \
    else                                                                                     \
        potts_measure_vec<DD, 1><<<g, b, 0, s>>>(labels, cg, q, partials, span)
        switch (cg.D) {
            case 1: PLQ_PM(1); break;
            case 2: PLQ_PM(2); break;
            case 3: PLQ_PM(3); break;
            default: PLQ_PM(4); break;
        }
#undef PLQ_PM
        if (defer) {
            defer->pa = partials;
            defer->na = G;
            defer->nk = 2 + q;
            defer->rec = rec;
        } else {
            launch_finalize_columns(partials, G, nullptr, 0, 2 + q, rec, s);
        }
        return;
    }
    int64_t need = (cg.N + 255) / 256;
    int64_t G = need < max_blocks ? need : max_blocks;
    int64_t span = (cg.N + G - 1) / G;
    dim3 g((uint32_t)G), b(256);
    switch (cg.D) {
        case 1: potts_measure_kernel<1><<<g, b, 0, s>>>(labels, cg, q, partials, span); break;
        case 2: potts_measure_kernel<2><<<g, b, 0, s>>>(labels, cg, q, partials, span); break;
        case 3: potts_measure_kernel<3><<<g, b, 0, s>>>(labels, cg, q, partials, span); break;
        default: potts_measure_kernel<4><<<g, b, 0, s>>>(labels, cg, q, partials, span); break;
    }
    launch_finalize_record(partials, G, nullptr, 0, rec, s);
}

bool cluster_boxes_x16(const CCGeom &cg, const void *labels) {
    bool pow2 = true;
    for (int mu = 0; mu < cg.D; ++mu)
        pow2 = pow2 && (cg.L[mu] & (cg.L[mu] - 1)) == 0 && (cg.B[mu] & (cg.B[mu] - 1)) == 0;
    return pow2 && cg.B[0] % 16 == 0 && cg.box_sites <= 16 * 1024 && (reinterpret_cast<uintptr_t>(labels) & 15u) == 0;
}

uint8_t cross_tag(uint32_t sweep) { return (uint8_t)(1u + sweep % 255u); }

void launch_cluster_sizes_box(const int32_t *parent, const CCGeom &cg, int32_t *sizes, int32_t *ext,
                              const uint8_t *cross, uint32_t sweep, unsigned long long *part,
                              const long long *relabel_part, int64_t relabel_rows, unsigned long long *hist,
                              int64_t max_blocks, bool counted, hipStream_t s) {
    if (max_blocks > MAX_PASS_BLOCKS) max_blocks = MAX_PASS_BLOCKS;  // the rows fit HIST_ROWS
    const int64_t Gb = cg.nboxes < max_blocks ? cg.nboxes : max_blocks;
    const size_t lds = (size_t)cg.box_sites * sizeof(int32_t);
    const uint8_t tag = cross_tag(sweep);
    const dim3 g((uint32_t)Gb), b(1024);
    if (!counted) switch (cg.D) {
        case 1: cc_sizes_box<1><<<g, b, lds, s>>>(parent, cg, sizes, ext, cross, tag, part); break;
        case 2: cc_sizes_box<2><<<g, b, lds, s>>>(parent, cg, sizes, ext, cross, tag, part); break;
        case 3: cc_sizes_box<3><<<g, b, lds, s>>>(parent, cg, sizes, ext, cross, tag, part); break;
        default: cc_sizes_box<4><<<g, b, lds, s>>>(parent, cg, sizes, ext, cross, tag, part); break;
    }
    int64_t hb = (cg.N + 4095) / 4096 < max_blocks ? (cg.N + 4095) / 4096 : max_blocks;
    if (hb < 1) hb = 1;
    cc_hist_cross<<<dim3((uint32_t)hb), dim3(256), 0, s>>>(parent, sizes, ext, cross, tag, cg.N, part, Gb);
    cc_hist_reduce<<<dim3(HIST_K + 1), dim3(256), 0, s>>>(part, Gb + hb, relabel_part, relabel_rows, hist);
}

void launch_cluster_resolve(int32_t *parent, int32_t *roots, const CCGeom &cg, uint32_t cap, int *flag,
                            int32_t *sizes, unsigned long long *hist, hipStream_t s) {
    dim3 g((uint32_t)((cg.N + 255) / 256)), b(256);
    cc_resolve<<<g, b, 0, s>>>(parent, roots, cg.N, cap, flag);
    if (sizes && hist) {
        cc_count<<<dim3((uint32_t)((cg.N + CC_SPAN - 1) / CC_SPAN)), b, 0, s>>>(roots, sizes, cg.N);
        const int64_t hb = (cg.N + 255) / 256 < 2048 ? (cg.N + 255) / 256 : 2048;
        cc_hist<<<dim3((uint32_t)hb), b, 0, s>>>(roots, sizes, cg.N, hist);
    }
}

}  // namespace plq
