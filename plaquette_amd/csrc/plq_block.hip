// plq_block.hip -- learned block proposal for the 2D Ising checkerboard.
//
// The lattice is cut into 8x8 blocks coloured (bx + by) & 1; the blocks of
// one colour share no bond, so a pass updates all of them at once.  For each
// block a 3-layer CNN (3x3 kernels, 32 channels, tanh) maps the 32 spins of
// the block's boundary ring to 64 logits, one per interior site.  The
// middle layer's taps are two sites apart (dilation 2), so that the three
// layers reach all 64 sites from the ring (depths 0, 0-2 and 0-3).  The
// block proposal draws the interior as 64 independent Bernoulli variables
// (state 0, spin +1, with probability sigmoid(logit)).  Because the logits
// depend on the ring only, which the move leaves unchanged, the proposal
// density of the reverse move is the same product, and the
// Metropolis-Hastings ratio is
//     exp(-beta dE) q(x) / q(x'),  log q(x) - log q(x') = sum_p ([x_p = 0] - [x'_p = 0]) l_p
// with dE the exact Ising energy change of the block (ising/mod.rs:97-143).
//
// The network runs on the matrix cores: the 32 -> 32 layer is a 64 x 32 x 288
// product per block, done as 72 v_mfma_f32_16x16x32_bf16 with BF16 weights
// and activations and FP32 accumulation.  In the pass kernel the 1 -> 32
// layer is 8 more (K = the 9 taps, zero-padded) and the 32 -> 1 layer 4 more
// (per position the channel sums of each tap, then 9 shifted adds); training
// runs those two layers as 288 FP32 FMAs per lane.  The acceptance test uses only IEEE single adds,
// multiplies and divides with a fixed summation order and the deterministic
// exponential plq_expf, so given the logits it is reproducible bit for bit
// on any IEEE machine.
//
// Training minimises the mean binary cross-entropy of each block's interior
// given its ring over configurations of a chain: one workgroup of one wave
// runs forward and backward for a contiguous run of blocks, accumulating its
// gradient in LDS in a fixed order, and a second kernel sums the workgroups'
// gradients in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

constexpr int B = BN_BLOCK;      // 8
constexpr int CH = BN_CHANNELS;  // 32
constexpr int HS = 40;           // LDS row stride (bf16) of an activation: 80 B, conflict-free b128 reads
constexpr int TS = 72;           // LDS row stride (bf16) of the transposed gradient
constexpr int DIL2 = 2;          // dilation of the middle layer: its taps are 2 sites apart

__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint32_t bits) { return __uint_as_float(bits << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ bf16x8 lds_frag(const uint16_t *p) {
    return as_bf16x8(*reinterpret_cast<const u32x4 *>(p));
}
__device__ __forceinline__ bf16x8 zero_frag() { return as_bf16x8(u32x4{0u, 0u, 0u, 0u}); }

// tanh(x) = 1 - 2 / (1 + e^{2x}) with the hardware exponential and
// reciprocal: absolute error ~1e-7, far below the BF16 step of an activation.
__device__ __forceinline__ float tanh_fast(float x) {
    return 1.0f - 2.0f * __frcp_rn(1.0f + __expf(2.0f * x));
}

__device__ __forceinline__ bool inside(int y, int x) { return (unsigned)y < (unsigned)B && (unsigned)x < (unsigned)B; }

// State of lattice site (X, Y) in the colour-interleaved layout, periodic.
__device__ __forceinline__ int site_state(const uint8_t *c0, const uint8_t *c1, int L0, int L1, int X, int Y) {
    X = X < 0 ? X + L0 : (X >= L0 ? X - L0 : X);
    Y = Y < 0 ? Y + L1 : (Y >= L1 ? Y - L1 : Y);
    int64_t idx = (int64_t)Y * (L0 >> 1) + (X >> 1);
    return ((X + Y) & 1) ? c1[idx] : c0[idx];
}

struct alignas(16) Wave {
    float patch[100];          // 10x10 input: ring spins +-1, interior and corners 0
    uint16_t h1[B * B * HS];   // layer-1 activations, [position][channel]
    uint16_t h2[B * B * HS];   // layer-2 activations
};

// Input patch of block (bx, by); lanes cover entries lane and lane + 64.
__device__ __forceinline__ void gather_patch(const uint8_t *c0, const uint8_t *c1, int L0, int L1, int bx, int by,
                                             int lane, float *patch) {
    for (int k = lane; k < 100; k += 64) {
        int yy = k / 10, xx = k - 10 * (k / 10);
        bool ring_row = (yy == 0 || yy == 9) && xx >= 1 && xx <= 8;
        bool ring_col = (xx == 0 || xx == 9) && yy >= 1 && yy <= 8;
        float v = 0.0f;
        if (ring_row || ring_col) v = 1.0f - 2.0f * (float)site_state(c0, c1, L0, L1, bx * B + xx - 1, by * B + yy - 1);
        patch[k] = v;
    }
}

// Layer 1 (1 -> 32, 3x3, tanh) for the lane's position.
__device__ __forceinline__ void layer1(const BlockNetView &n, const float *patch, int lane, uint16_t *h1) {
    const int y = lane >> 3, x = lane & 7;
    float in[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) in[t] = patch[(y + t / 3) * 10 + (x + t % 3)];
#pragma unroll 4
    for (int c = 0; c < CH; c += 2) {
        float a0 = n.b1[c], a1 = n.b1[c + 1];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            a0 += n.w1[t * CH + c] * in[t];
            a1 += n.w1[t * CH + c + 1] * in[t];
        }
        uint32_t pk = (uint32_t)bf16_bits(tanh_fast(a0)) | ((uint32_t)bf16_bits(tanh_fast(a1)) << 16);
        *reinterpret_cast<uint32_t *>(&h1[lane * HS + c]) = pk;
    }
}

// Layer 2 (32 -> 32, 3x3, tanh) on the matrix cores.  Position tile mt holds
// positions 16 mt .. 16 mt + 15; A rows are positions, K the 32 input
// channels of one tap, B columns 16 output channels.
// With PRELOADED the B fragments come from wB[t * 2 + nt] (registers), else
// from the packed weights in global memory.
template <bool PRELOADED>
__device__ __forceinline__ void layer2(const BlockNetView &n, const bf16x8 *wB, const uint16_t *h1, int lane,
                                       uint16_t *h2) {
    const int row = lane & 15, kb = (lane >> 4) * 8;
#pragma unroll 1
    for (int mt = 0; mt < 4; ++mt) {
        const int p = 16 * mt + row, y = p >> 3, x = p & 7;
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y + DIL2 * (t / 3 - 1), xx = x + DIL2 * (t % 3 - 1);
            bf16x8 a = inside(yy, xx) ? lds_frag(&h1[(yy * B + xx) * HS + kb]) : zero_frag();
            bf16x8 b0, b1;
            if (PRELOADED) {
                b0 = wB[t * 2 + 0];
                b1 = wB[t * 2 + 1];
            } else {
                b0 = as_bf16x8(reinterpret_cast<const u32x4 *>(n.w2f)[(t * 2 + 0) * 64 + lane]);
                b1 = as_bf16x8(reinterpret_cast<const u32x4 *>(n.w2f)[(t * 2 + 1) * 64 + lane]);
            }
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
        }
        // D: lane holds column (lane & 15), rows 4 * (lane >> 4) + i
        const int col = lane & 15;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = 16 * mt + 4 * (lane >> 4) + i;
            h2[q * HS + col] = bf16_bits(tanh_fast(acc0[i] + n.b2[col]));
            h2[q * HS + 16 + col] = bf16_bits(tanh_fast(acc1[i] + n.b2[16 + col]));
        }
    }
}

// The matrix-core forms of layers 1 and 3 in the pass kernel.  Their B
// fragments (n.w13f) are packed by block_pack_kernel:
//   layer 1, tile nt (channels 16 nt ..): lane l, element j is W1 at tap
//     3 (l >> 4) + j, channel 16 nt + (l & 15), for j < 3 and l < 48 (else 0);
//   layer 3: lane l, element j is W3 at tap l & 15 (below 9, else 0),
//     channel 8 (l >> 4) + j.
constexpr int W13F_L3 = 2 * 64 * 8;  // offset of the layer-3 fragments (uint16)
constexpr int W13F = 3 * 64 * 8;     // uint16 entries of n.w13f
constexpr int CBS = 17;              // row stride (floats) of layer 3's per-tap sums

// Layer 1 (1 -> 32, 3x3, tanh) on the matrix cores.  A rows are the 16
// positions of tile mt, K the 9 taps (three per group of 16 lanes, the rest
// zero), B columns 16 output channels.  The patch values (+-1, 0) are exact
// in BF16 and the weights are the BF16-rounded W1 of the vector form, so the
// products are the same; only the order of the FP32 sums differs.  The
// accumulator starts at the channel's bias.
__device__ __forceinline__ void layer1_mfma(const BlockNetView &n, const float *patch, int lane, uint16_t *h1) {
    const int row = lane & 15, kq = lane >> 4, col = lane & 15;
    const bf16x8 b0 = as_bf16x8(reinterpret_cast<const u32x4 *>(n.w13f)[0 * 64 + lane]);
    const bf16x8 b1 = as_bf16x8(reinterpret_cast<const u32x4 *>(n.w13f)[1 * 64 + lane]);
    const float bias0 = n.b1[col], bias1 = n.b1[16 + col];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int p = 16 * mt + row, y = p >> 3, x = p & 7;
        uint32_t lo = 0u, hi = 0u;
        if (kq < 3) {
            // taps 3 kq .. 3 kq + 2: patch rows y + kq, columns x .. x + 2
            const float *pr = &patch[(y + kq) * 10 + x];
            lo = (uint32_t)bf16_bits(pr[0]) | ((uint32_t)bf16_bits(pr[1]) << 16);
            hi = (uint32_t)bf16_bits(pr[2]);
        }
        const bf16x8 a = as_bf16x8(u32x4{lo, hi, 0u, 0u});
        f32x4 acc0 = {bias0, bias0, bias0, bias0}, acc1 = {bias1, bias1, bias1, bias1};
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
        // D: lane holds column (lane & 15), rows 4 * (lane >> 4) + i
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = 16 * mt + 4 * kq + i;
            h1[q * HS + col] = bf16_bits(tanh_fast(acc0[i]));
            h1[q * HS + 16 + col] = bf16_bits(tanh_fast(acc1[i]));
        }
    }
}

// Layer 3 (32 -> 1, 3x3) on the matrix cores, first half: per position p and
// tap t the channel sum C[p][t] = sum_c W3[t][c] h2[p][c], one 16 x 16 x 32
// product per tile of 16 positions (B columns are the taps, 9 of 16 used),
// into cbuf[p * CBS + t] (a row stride of 17 floats keeps the position-wise
// reads of the second half on distinct LDS banks).
__device__ __forceinline__ void layer3_mfma_taps(const BlockNetView &n, const uint16_t *h2, int lane, float *cbuf) {
    const int row = lane & 15, kb = (lane >> 4) * 8, col = lane & 15;
    const bf16x8 b = as_bf16x8(reinterpret_cast<const u32x4 *>(n.w13f + W13F_L3)[lane]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const bf16x8 a = lds_frag(&h2[(16 * mt + row) * HS + kb]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) cbuf[(16 * mt + 4 * (lane >> 4) + i) * CBS + col] = acc[i];
    }
}

// Second half: the logit of the lane's position, the bias plus C[p + tap
// offset][t] over the taps whose position lies in the block, in tap order.
__device__ __forceinline__ float layer3_mfma_sum(const BlockNetView &n, const float *cbuf, int lane) {
    const int y = lane >> 3, x = lane & 7;
    float s = n.b3[0];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
        if (inside(yy, xx)) s += cbuf[(yy * B + xx) * CBS + t];
    }
    return s;
}

// Layer 3 (32 -> 1, 3x3): the logit of the lane's position.
__device__ __forceinline__ float layer3(const BlockNetView &n, const uint16_t *h2, int lane) {
    const int y = lane >> 3, x = lane & 7;
    float s = n.b3[0];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
        if (!inside(yy, xx)) continue;
        const uint16_t *row = &h2[(yy * B + xx) * HS];
#pragma unroll
        for (int c8 = 0; c8 < CH; c8 += 8) {
            u32x4 v = *reinterpret_cast<const u32x4 *>(row + c8);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s += n.w3[t * CH + c8 + 2 * k] * bf16_float(v[k] & 0xffffu);
                s += n.w3[t * CH + c8 + 2 * k + 1] * bf16_float(v[k] >> 16);
            }
        }
    }
    return s;
}

__device__ __forceinline__ int wave_sum_int(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------ MH pass ---

// mode 0: one Metropolis-Hastings pass over the blocks of `colour`;
// mode 1: logits of every block, no update.  Four waves per workgroup, one
// block per wave.
__global__ __launch_bounds__(256) void block_pass_kernel(uint8_t *__restrict__ c0, uint8_t *__restrict__ c1,
                                                         int L0, int L1, int mode, int colour, uint32_t seed32,
                                                         uint32_t bstep, double beta, double jc, double hc,
                                                         BlockNetView net, float *__restrict__ logits_out,
                                                         unsigned long long *__restrict__ acc_slots) {
    __shared__ Wave waves[4];
    __shared__ int wave_acc[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    Wave &W = waves[w];
    const int nb0 = L0 / B, nb1 = L1 / B;
    const int count = mode == 0 ? nb0 * nb1 / 2 : nb0 * nb1;
    const int k = blockIdx.x * 4 + w;
    const bool active = k < count;
    int bx = 0, by = 0;
    if (mode == 0) {
        by = k / (nb0 >> 1);
        bx = 2 * (k - by * (nb0 >> 1)) + ((by + colour) & 1);
    } else {
        by = k / nb0;
        bx = k - by * nb0;
    }
    const int px = lane & 7, py = lane >> 3;
    const int X = bx * B + px, Y = by * B + py;

    bf16x8 wB[18];
#pragma unroll
    for (int f = 0; f < 18; ++f) wB[f] = as_bf16x8(reinterpret_cast<const u32x4 *>(net.w2f)[f * 64 + lane]);

    int cur = 0;
    if (active) {
        gather_patch(c0, c1, L0, L1, bx, by, lane, W.patch);
        cur = site_state(c0, c1, L0, L1, X, Y);
    }
    __syncthreads();
    if (active) layer1_mfma(net, W.patch, lane, W.h1);
    __syncthreads();
    if (active) layer2<true>(net, wB, W.h1, lane, W.h2);
    __syncthreads();
    // layer 3's per-tap channel sums go where h1 was (h1 is dead after layer 2)
    static_assert(B * B * CBS * sizeof(float) <= sizeof(W.h1), "layer-3 sums fit in h1");
    float *cbuf = reinterpret_cast<float *>(W.h1);
    if (active) layer3_mfma_taps(net, W.h2, lane, cbuf);
    __syncthreads();
    if (active) {
        const float l = layer3_mfma_sum(net, cbuf, lane);
        const int64_t b = (int64_t)by * nb0 + bx;
        if (logits_out) logits_out[b * 64 + lane] = l;
        if (mode == 0) {
            const uint32_t site = (uint32_t)Y * (uint32_t)L0 + (uint32_t)X;
            const uint32_t m = keyed_bits(seed32 ^ BLOCK_PROPOSAL_KEY, site, bstep);
            const float pu = 1.0f / (1.0f + plq_expf(-l));
            const int prop = ((float)m < pu * 16777216.0f) ? 0 : 1;
            float v = (float)((cur == 0) - (prop == 0)) * l;
            for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
            const int s = 1 - 2 * cur, t = 1 - 2 * prop;
            const int s_r = __shfl(s, (lane + 1) & 63, 64), t_r = __shfl(t, (lane + 1) & 63, 64);
            const int s_d = __shfl(s, (lane + 8) & 63, 64), t_d = __shfl(t, (lane + 8) & 63, 64);
            int ring = 0;  // the patch is zero inside the block, so this sums ring neighbours only
            ring += (int)W.patch[(py + 1) * 10 + px] + (int)W.patch[(py + 1) * 10 + px + 2];
            ring += (int)W.patch[py * 10 + px + 1] + (int)W.patch[(py + 2) * 10 + px + 1];
            int db = (t - s) * ring;
            if (px < B - 1) db += t * t_r - s * s_r;
            if (py < B - 1) db += t * t_d - s * s_d;
            const int dbond = wave_sum_int(db), dmag = wave_sum_int(t - s);
            const double de = jc * (double)dbond + hc * (double)dmag;  // -Delta E
            const float A = (float)(beta * de) + v;
            const uint32_t ma = keyed_bits(seed32 ^ BLOCK_ACCEPT_KEY, (uint32_t)b, bstep);
            const bool acc = A >= 0.0f || (float)ma < plq_expf(A) * 16777216.0f;
            if (acc) {
                int64_t idx = (int64_t)Y * (L0 >> 1) + (X >> 1);
                (((X + Y) & 1) ? c1 : c0)[idx] = (uint8_t)prop;
            }
            if (lane == 0) wave_acc[w] = acc ? 1 : 0;
        }
    }
    // Accepted moves per workgroup, accumulated in the workgroup's own slot:
    // passes run in stream order, so the read-modify-write needs no atomic.
    if (mode == 0) {
        if (!active && lane == 0) wave_acc[w] = 0;
        __syncthreads();
        if (threadIdx.x == 0)
            acc_slots[blockIdx.x] += (unsigned long long)(wave_acc[0] + wave_acc[1] + wave_acc[2] + wave_acc[3]);
    }
}

// ------------------------------------------------------------- training ---

struct alignas(16) TrainLds {
    float patch[100];
    uint16_t h1[B * B * HS];
    float g3[B * B];
    uint16_t g2[B * B * HS];              // dL/d(pre-activation of layer 2), BF16, [position][channel]
    union {
        struct {
            uint16_t h2[B * B * HS];
            uint16_t g2t[CH * TS];        // the same, transposed [channel][position]
        } a;
        float g1[B * B * CH];             // dL/d(pre-activation of layer 1), [position][channel]
    } u;
    float dw2[9 * CH * CH];               // [tap][out][in]
};

// One wave per workgroup; workgroup g trains on blocks [g*per, (g+1)*per)
// and writes its gradient sums (and loss sum) to part[g * stride ...].
__global__ __launch_bounds__(64) void block_train_kernel(const uint8_t *__restrict__ c0,
                                                         const uint8_t *__restrict__ c1, int L0, int L1, int per,
                                                         BlockNetView net, float *__restrict__ part, int stride) {
    __shared__ TrainLds S;
    const int lane = threadIdx.x;
    const int nb0 = L0 / B, nblocks = nb0 * (L1 / B);
    const int k0 = blockIdx.x * per;
    const int k1 = min(nblocks, k0 + per);
    const int py = lane >> 3, px = lane & 7;
    const int row = lane & 15, kb = (lane >> 4) * 8, col = lane & 15;

    for (int e = lane; e < 9 * CH * CH; e += 64) S.dw2[e] = 0.0f;
    float a_w1[5] = {0, 0, 0, 0, 0}, a_w3[5] = {0, 0, 0, 0, 0};
    float a_b2[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) a_b2[c] = 0.0f;
    float a_b1[2] = {0.f, 0.f}, a_b3 = 0.f, a_loss = 0.f;
    // lane's dW1 / dW3 entries: channel lane & 31, taps (lane >> 5) + 2 i
    const int ec = lane & 31, et0 = lane >> 5;

    for (int kk = k0; kk < k1; ++kk) {
        const int by = kk / nb0, bx = kk - by * nb0;
        __syncthreads();
        gather_patch(c0, c1, L0, L1, bx, by, lane, S.patch);
        const int cur = site_state(c0, c1, L0, L1, bx * B + px, by * B + py);
        __syncthreads();
        layer1(net, S.patch, lane, S.h1);
        __syncthreads();
        layer2<false>(net, nullptr, S.h1, lane, S.u.a.h2);
        __syncthreads();
        // loss and dL/dlogit
        const float l = layer3(net, S.u.a.h2, lane);
        const float yv = cur == 0 ? 1.0f : 0.0f;
        a_loss += (l > 0.0f ? l + log1pf(expf(-l)) : log1pf(expf(l))) - yv * l;
        const float g3 = 1.0f / (1.0f + expf(-l)) - yv;
        a_b3 += g3;
        S.g3[lane] = g3;
        __syncthreads();
        // dW3[t][c] += sum_p g3[p] h2[p + t][c]
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int t = et0 + 2 * i;
            if (t >= 9) break;
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            float s = 0.f;
#pragma unroll 8
            for (int p = 0; p < 64; ++p) {
                const int yy = (p >> 3) + dy, xx = (p & 7) + dx;
                if (inside(yy, xx)) s += S.g3[p] * bf16_float(S.u.a.h2[(yy * B + xx) * HS + ec]);
            }
            a_w3[i] += s;
        }
        // g2[q][c] = (sum_t W3[t][c] g3[q - t]) (1 - h2[q][c]^2)
        {
            float gq[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = py - (t / 3 - 1), xx = px - (t % 3 - 1);
                gq[t] = inside(yy, xx) ? S.g3[yy * B + xx] : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                float s = 0.f;
#pragma unroll
                for (int t = 0; t < 9; ++t) s += net.w3[t * CH + c] * gq[t];
                const float h = bf16_float(S.u.a.h2[lane * HS + c]);
                const float g = s * (1.0f - h * h);
                a_b2[c] += g;
                const uint16_t gb = bf16_bits(g);
                S.g2[lane * HS + c] = gb;
                S.u.a.g2t[c * TS + lane] = gb;
            }
        }
        __syncthreads();
        // dW2[t][co][ci] += sum_p g2[p][co] h1[p + t][ci]: M = co, N = ci, K = positions
#pragma unroll 1
        for (int t = 0; t < 9; ++t) {
            const int dy = DIL2 * (t / 3 - 1), dx = DIL2 * (t % 3 - 1);
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const int p0 = 32 * ks + kb;  // a row of 8 positions: y = p0 / 8, x = 0..7
                    bf16x8 a = lds_frag(&S.u.a.g2t[(16 * mt + row) * TS + p0]);
                    const int yy = (p0 >> 3) + dy;
                    bf16x8 bv[2];
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        uint32_t w[4];
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            uint32_t lo = 0, hi = 0;
                            const int x0 = 2 * jj + dx, x1 = 2 * jj + 1 + dx;
                            if (inside(yy, x0)) lo = S.h1[(yy * B + x0) * HS + 16 * nt + col];
                            if (inside(yy, x1)) hi = S.h1[(yy * B + x1) * HS + 16 * nt + col];
                            w[jj] = lo | (hi << 16);
                        }
                        bv[nt] = as_bf16x8(u32x4{w[0], w[1], w[2], w[3]});
                    }
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv[0], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv[1], acc1, 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = 16 * mt + 4 * (lane >> 4) + i;
                    S.dw2[(t * CH + co) * CH + col] += acc0[i];
                    S.dw2[(t * CH + co) * CH + 16 + col] += acc1[i];
                }
            }
        }
        __syncthreads();
        // g1[q][ci] = (sum_t sum_co W2[t][co][ci] g2[q - t][co]) (1 - h1[q][ci]^2):
        // M = positions, N = ci, K = co per tap
#pragma unroll 1
        for (int mt = 0; mt < 4; ++mt) {
            const int p = 16 * mt + row, y = p >> 3, x = p & 7;
            f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = y - DIL2 * (t / 3 - 1), xx = x - DIL2 * (t % 3 - 1);
                bf16x8 a = inside(yy, xx) ? lds_frag(&S.g2[(yy * B + xx) * HS + kb]) : zero_frag();
                bf16x8 b0 = as_bf16x8(reinterpret_cast<const u32x4 *>(net.w2t)[(t * 2 + 0) * 64 + lane]);
                bf16x8 b1 = as_bf16x8(reinterpret_cast<const u32x4 *>(net.w2t)[(t * 2 + 1) * 64 + lane]);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int q = 16 * mt + 4 * (lane >> 4) + i;
                const float h0 = bf16_float(S.h1[q * HS + col]), h1v = bf16_float(S.h1[q * HS + 16 + col]);
                const float g0 = acc0[i] * (1.0f - h0 * h0), g1 = acc1[i] * (1.0f - h1v * h1v);
                S.u.g1[q * CH + col] = g0;
                S.u.g1[q * CH + 16 + col] = g1;
                a_b1[0] += g0;
                a_b1[1] += g1;
            }
        }
        __syncthreads();
        // dW1[t][c] += sum_q g1[q][c] patch[q + t]
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int t = et0 + 2 * i;
            if (t >= 9) break;
            float s = 0.f;
#pragma unroll 8
            for (int q = 0; q < 64; ++q)
                s += S.u.g1[q * CH + ec] * S.patch[((q >> 3) + t / 3) * 10 + (q & 7) + t % 3];
            a_w1[i] += s;
        }
    }
    __syncthreads();

    // Write this workgroup's sums in parameter order.
    float *out = part + (int64_t)blockIdx.x * stride;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const int t = et0 + 2 * i;
        if (t < 9) {
            out[BN_OFF_W1 + t * CH + ec] = a_w1[i];
            out[BN_OFF_W3 + t * CH + ec] = a_w3[i];
        }
    }
    for (int e = lane; e < 9 * CH * CH; e += 64) out[BN_OFF_W2 + e] = S.dw2[e];
    // b1: lanes l, l+16, l+32, l+48 share channels (l & 15) and 16 + (l & 15)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float v = a_b1[nt];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (lane < 16) out[BN_OFF_B1 + 16 * nt + lane] = v;
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        float v = a_b2[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) out[BN_OFF_B2 + c] = v;
    }
    float v3 = a_b3, vl = a_loss;
    for (int off = 32; off > 0; off >>= 1) {
        v3 += __shfl_xor(v3, off, 64);
        vl += __shfl_xor(vl, off, 64);
    }
    if (lane == 0) {
        out[BN_OFF_B3] = v3;
        out[BN_PARAMS] = vl;
    }
}

// grad[e] = (sum over workgroups of part[g][e]) / nsites, e <= BN_PARAMS (the
// last entry is the mean loss), in two fixed-order stages: stage 1 sums runs
// of REDUCE_RUN workgroups into part2[run][e], stage 2 sums the runs.
constexpr int REDUCE_RUN = 16;

__global__ void block_reduce1_kernel(const float *__restrict__ part, int nwg, int stride,
                                     float *__restrict__ part2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int run = blockIdx.y;
    if (e > BN_PARAMS) return;
    const int g0 = run * REDUCE_RUN, g1 = min(nwg, g0 + REDUCE_RUN);
    float s = 0.f;
    for (int g = g0; g < g1; ++g) s += part[(int64_t)g * stride + e];
    part2[(int64_t)run * stride + e] = s;
}

__global__ void block_reduce2_kernel(const float *__restrict__ part2, int nruns, int stride, float inv_sites,
                                     float *__restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > BN_PARAMS) return;
    float s = 0.f;
    for (int r = 0; r < nruns; ++r) s += part2[(int64_t)r * stride + e];
    grad[e] = s * inv_sites;
}

__global__ void block_adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                  const float *__restrict__ grad, float lr, float beta1, float beta2, float c1,
                                  float c2, float eps) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= BN_PARAMS) return;
    const float g = grad[e];
    const float mm = beta1 * m[e] + (1.0f - beta1) * g;
    const float vv = beta2 * v[e] + (1.0f - beta2) * g * g;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// BF16 forms of the weights: W1 and W3 rounded (kept as float for the VALU
// layers), and W2 as matrix-core B fragments for the forward pass
// (w2f[t][nt][lane][j] = W2[t][16 nt + (lane & 15)][8 (lane >> 4) + j]) and
// for the input gradient (w2t[t][nt][lane][j] = W2[t][8 (lane >> 4) + j][16 nt + (lane & 15)]).
__global__ void block_pack_kernel(const float *__restrict__ w, float *__restrict__ w1r, float *__restrict__ w3r,
                                  uint16_t *__restrict__ w2f, uint16_t *__restrict__ w2t, uint16_t *__restrict__ w13f) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 9 * CH) {
        w1r[e] = bf16_float(bf16_bits(w[BN_OFF_W1 + e]));
        w3r[e] = bf16_float(bf16_bits(w[BN_OFF_W3 + e]));
    }
    if (e < W13F) {
        const int j = e & 7, lane = (e >> 3) & 63;
        uint16_t v = 0;
        if (e < W13F_L3) {
            // layer 1, channel tile nt: tap 3 (lane >> 4) + j, channel 16 nt + (lane & 15)
            const int nt = e >> 9, kq = lane >> 4, t = 3 * kq + j;
            if (kq < 3 && j < 3) v = bf16_bits(w[BN_OFF_W1 + t * CH + 16 * nt + (lane & 15)]);
        } else {
            // layer 3: tap lane & 15, channel 8 (lane >> 4) + j
            const int t = lane & 15, ch = 8 * (lane >> 4) + j;
            if (t < 9) v = bf16_bits(w[BN_OFF_W3 + t * CH + ch]);
        }
        w13f[e] = v;
    }
    if (e < 9 * 2 * 64 * 8) {
        const int j = e & 7, lane = (e >> 3) & 63, nt = (e >> 9) & 1, t = e >> 10;
        const int a = 16 * nt + (lane & 15), b = 8 * (lane >> 4) + j;
        w2f[e] = bf16_bits(w[BN_OFF_W2 + (t * CH + a) * CH + b]);
        w2t[e] = bf16_bits(w[BN_OFF_W2 + (t * CH + b) * CH + a]);
    }
}

}  // namespace

void launch_block_pack(const float *master, float *w1r, float *w3r, uint16_t *w2f, uint16_t *w2t, uint16_t *w13f,
                       hipStream_t s) {
    block_pack_kernel<<<dim3((9 * 2 * 64 * 8 + 255) / 256), dim3(256), 0, s>>>(master, w1r, w3r, w2f, w2t, w13f);
}

void launch_block_pass(const BlockPassArgs &a) {
    const int nb = (a.L0 / BN_BLOCK) * (a.L1 / BN_BLOCK);
    const int count = a.mode == 0 ? nb / 2 : nb;
    block_pass_kernel<<<dim3((count + 3) / 4), dim3(256), 0, a.stream>>>(
        a.c0, a.c1, a.L0, a.L1, a.mode, a.colour, a.seed32, a.bstep, a.beta, a.j, a.h, a.net, a.logits, a.acc_slots);
}

int block_pass_groups(int L0, int L1) { return ((L0 / BN_BLOCK) * (L1 / BN_BLOCK) / 2 + 3) / 4; }

int block_train_groups(int L0, int L1) {
    const int nb = (L0 / BN_BLOCK) * (L1 / BN_BLOCK);
    return nb < BLOCK_TRAIN_GROUPS ? nb : BLOCK_TRAIN_GROUPS;
}

void launch_block_grad(const BlockGradArgs &a) {
    const int nb = (a.L0 / BN_BLOCK) * (a.L1 / BN_BLOCK);
    const int nwg = block_train_groups(a.L0, a.L1);
    const int per = (nb + nwg - 1) / nwg;
    block_train_kernel<<<dim3(nwg), dim3(64), 0, a.stream>>>(a.c0, a.c1, a.L0, a.L1, per, a.net, a.partials,
                                                             BLOCK_PART_STRIDE);
    const int nruns = (nwg + REDUCE_RUN - 1) / REDUCE_RUN;
    float *part2 = a.partials + (int64_t)BLOCK_TRAIN_GROUPS * BLOCK_PART_STRIDE;
    block_reduce1_kernel<<<dim3((BN_PARAMS + 1 + 255) / 256, nruns), dim3(256), 0, a.stream>>>(
        a.partials, nwg, BLOCK_PART_STRIDE, part2);
    block_reduce2_kernel<<<dim3((BN_PARAMS + 1 + 255) / 256), dim3(256), 0, a.stream>>>(
        part2, nruns, BLOCK_PART_STRIDE, 1.0f / (float)((int64_t)a.L0 * a.L1), a.grad);
}

void launch_block_adam(float *w, float *m, float *v, const float *grad, float lr, float beta1, float beta2,
                       float c1, float c2, float eps, hipStream_t s) {
    block_adam_kernel<<<dim3((BN_PARAMS + 255) / 256), dim3(256), 0, s>>>(w, m, v, grad, lr, beta1, beta2, c1, c2,
                                                                         eps);
}

}  // namespace plq
