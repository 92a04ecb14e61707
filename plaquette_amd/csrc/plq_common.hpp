// plq_common.hpp -- pieces shared by the host library and the gfx950 kernels:
// the keyed counter-based stream, the colour-interleaved site layout, and the
// launch-independent constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plq {

// Stream salts: metropolis.wgsl:30, cluster_prelude.wgsl:29,
// cluster_relabel.wgsl:33 and :37.
constexpr uint32_t PROPOSAL_KEY = 0x85ebca6bu;
constexpr uint32_t BOND_KEY = 0x27d4eb2fu;
constexpr uint32_t ROOT_SALT = 0x85ebca6bu;
constexpr uint32_t SEED_SALT = 0xc2b2ae35u;
constexpr uint32_t GOLDEN = 0x9e3779b9u;
constexpr uint32_t DRAW_ONE = 1u << 24;  // 2^24: every 24-bit draw is below it

// lowbias32 (rng.wgsl:34-42): two xor-shift-multiply rounds on u32.
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

// The 24-bit integer behind keyed_uniform (rng.wgsl:47-52):
// keyed_uniform = m * 2^-24 exactly, so every comparison of it against a
// threshold is an integer comparison of m.
__host__ __device__ __forceinline__ uint32_t keyed_bits(uint32_t seed, uint32_t var, uint32_t sweep) {
    uint32_t k = lowbias32(seed ^ var);
    k = lowbias32(k ^ (sweep * GOLDEN));
    return k >> 8;
}

// lowbias32 after its first xor-shift (the caller supplies v ^ (v >> 16)).
__host__ __device__ __forceinline__ uint32_t lowbias32_tail(uint32_t v) {
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

// lowbias32 between its first and its last xor-shift.
__host__ __device__ __forceinline__ uint32_t lowbias32_mid(uint32_t v) {
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    return v;
}

// The sweep's key as the merged draw below consumes it.
__host__ __device__ __forceinline__ uint32_t sweep_mix(uint32_t sweep_key) { return sweep_key ^ (sweep_key >> 16); }

// The same draw with the sweep term hoisted: sweep_key = sweep * GOLDEN.
// With u the first round before its last xor-shift (h = u ^ u >> 16), the
// second round starts from (h ^ k) ^ ((h ^ k) >> 16) = u ^ (k ^ k >> 16),
// since h >> 16 = u >> 16: the two xor-shifts between the rounds collapse
// into one xor with sweep_mix(k), two operations fewer per draw.
__host__ __device__ __forceinline__ uint32_t keyed_bits_k(uint32_t seed, uint32_t var, uint32_t sweep_key) {
    const uint32_t x = seed ^ var;
    return lowbias32_tail(lowbias32_mid(x ^ (x >> 16)) ^ sweep_mix(sweep_key)) >> 8;
}

// f32 uniform for the draws the reference scales in f32 (proposal, relabel,
// Wolff seed): m * 2^-24, exact.
__host__ __device__ __forceinline__ float keyed_uniform_f32(uint32_t m) {
    return (float)m * (1.0f / 16777216.0f);
}

// A cluster's new label from its keyed draw (cluster_relabel.wgsl:39-90):
// Redraw takes floor(u q), ForcedChange adds 1 + floor(u (q - 1)) to the
// cluster's label lab; both scaled in f32 and clamped.
__host__ __device__ __forceinline__ uint32_t relabel_draw(uint32_t u_bits, uint32_t lab, int q, int forced) {
    const float u = keyed_uniform_f32(u_bits);
    if (forced) {
        uint32_t step = (uint32_t)(u * (float)(uint32_t)(q - 1));
        if (step >= (uint32_t)(q - 1)) step = (uint32_t)(q - 2);
        return (lab + 1u + step) % (uint32_t)q;
    }
    uint32_t draw = (uint32_t)(u * (float)(uint32_t)q);
    return draw >= (uint32_t)q ? (uint32_t)(q - 1) : draw;
}

// n / L by a multiplication, M = ceil(2^40 / L): exact while n * L < 2^40.
__host__ __device__ __forceinline__ uint32_t cdiv(uint32_t n, uint64_t M) { return (uint32_t)(((uint64_t)n * M) >> 40); }

// fold_seed (device.rs:761-763).
__host__ __device__ __forceinline__ uint32_t fold_seed(uint64_t seed) {
    return (uint32_t)(seed ^ (seed >> 32));
}

// Element access through a 32-bit byte offset: the address is a uniform base
// plus a zero-extended 32-bit offset, which the compiler issues as a global
// load/store with an SGPR base and a VGPR offset (no 64-bit VALU address
// arithmetic per access).  Callers guarantee idx * sizeof(T) < 2^32.
template <class T>
__device__ __forceinline__ T ld32(const T *base, uint32_t idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T *base, uint32_t idx, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}

constexpr int MAX_D = 4;

// Colour-interleaved site layout.  A site (x0, x1, ..., x_{D-1}) lives in
// row r = x1 + L1*(x2 + L2*x3) at position x0.  Its colour is
// c = (x0 + x1 + ... ) & 1, and it is stored in colour array c at
// r * H + (x0 >> 1), H = L0 / 2.  Every row of a colour array holds the H
// sites of that colour in increasing x0, so a lane reading four consecutive
// bytes reads four same-colour sites two lattice steps apart.
struct Geom {
    int32_t D;
    int32_t L[MAX_D];      // extents; unused axes are 1
    int32_t H;             // L[0] / 2
    int32_t R;             // rows = N / L[0]
    int32_t rstride[MAX_D];// row stride of axis mu >= 1 (rstride[1] = 1)
    int64_t N;             // sites
};

// Decompose a row index into the coordinates of axes 1..D-1.
__host__ __device__ __forceinline__ void row_coords(const Geom &g, int32_t r, int32_t *x) {
    int32_t rest = r;
    for (int mu = 1; mu < MAX_D; ++mu) {
        if (mu < g.D) {
            x[mu] = rest % g.L[mu];
            rest /= g.L[mu];
        } else {
            x[mu] = 0;
        }
    }
}

// The row reached from row r (with coordinates x) by one step along axis
// mu >= 1, sign +1 or -1, under periodic wrap.
__host__ __device__ __forceinline__ int32_t row_step(const Geom &g, int32_t r, const int32_t *x, int mu, int sign) {
    int32_t l = g.L[mu];
    int32_t xn = x[mu] + sign;
    int32_t delta = sign;
    if (xn == l) delta = 1 - l;
    if (xn < 0) delta = l - 1;
    return r + delta * g.rstride[mu];
}

// Sum of v over the 64 lanes of a wave, returned in every lane; every lane
// must be active.  DPP butterflies give each lane its quad's sum
// (quad_perm [1,0,3,2] then [2,3,0,1]) and then its row's sum (row_ror 4 and
// 8: a lane adds the quads four and eight lanes round its row of 16), and the
// four row sums are read into a scalar.  A few cycles, where a shuffle
// reduction waits on six LDS-crossbar round trips.
__device__ __forceinline__ int wave_sum_i32(int v) {
    v += __builtin_amdgcn_mov_dpp(v, 0xb1, 0xf, 0xf, false);
    v += __builtin_amdgcn_mov_dpp(v, 0x4e, 0xf, 0xf, false);
    v += __builtin_amdgcn_mov_dpp(v, 0x124, 0xf, 0xf, false);
    v += __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, false);
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
           __builtin_amdgcn_readlane(v, 48);
}

// Deterministic f32 exponential: Cody-Waite range reduction (Cody & Waite,
// "Software Manual for the Elementary Functions", 1980) with a split ln 2 and
// a degree-7 Taylor polynomial on |r| <= ln2/2, scaled by 2^n through the
// exponent bits.  Only IEEE single adds and multiplies (no fused operations
// under -ffp-contract=off), so host and device compute the same bits; within
// 1.1 ulp of exp on [-87, 88], 0 below, +inf above.
__host__ __device__ __forceinline__ float plq_expf(float x) {
    if (x > 88.0f) return __builtin_inff();
    if (x < -87.0f) return 0.0f;
    const float n = __builtin_rintf(x * 1.44269502f);
    float r = x - n * 0.693145751953125f;
    r = r - n * 1.42860677e-06f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const uint32_t bits = (uint32_t)((int)n + 127) << 23;
    return p * __builtin_bit_cast(float, bits);
}

}  // namespace plq
