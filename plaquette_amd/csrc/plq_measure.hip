// plq_measure.hip -- per-configuration measurements on gfx950: the
// two-point correlator's pair sums (models/mod.rs:44-60), Polyakov loops
// (gauge/mod.rs:178-201) and the Wilson rectangle table (gauge/mod.rs:224-268).
//
// The reference walks every site for every displacement, one state at a
// time.  Here a configuration is first packed into bit planes: plane k holds
// one bit of every variable, row R (the line of axis-0 sites with the same
// x1..x_{D-1}) in W = ceil(L0 / 32) words, bit x0 % 32 of word x0 / 32, bits
// past L0 zero.  A translation along an axis a >= 1 is then a change of row,
// and along axis 0 a 32-bit window read from the row with wrap at L0, so every
// pair, loop or line sum becomes XOR / AND over words and a popcount:
//
//   correlator: agreements of s(x) and s(x + r e_mu) = popc(AND_p ~(a_p ^ b_p))
//               over the state's bit planes; the Ising sign-product sum is
//               2 * agreements - N;
//   Polyakov:   the parity of the L_dir links of each line (XOR down the line);
//   Wilson:     with line products  Lr_mu(x) = XOR_{k<r} U_mu(x + k e_mu),
//               the r-by-t loop from x (r along mu, then t along nu, walked as
//               lattice.rs:467-486 walks Loop::rectangle, lattice.rs:532-555)
//               has parity Lr_mu(x) ^ Lt_nu(x + r e_mu) ^ Lr_mu(x + t e_nu) ^ Lt_nu(x).
//
// Every result is an integer count; the host turns counts into the
// reference's f64 values with the reference's own final division.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

namespace {

constexpr int MEAS_BLOCK = 256;

__host__ __device__ __forceinline__ uint32_t valid_bits(int L0, int w) {
    const int rem = L0 - 32 * w;
    return rem >= 32 ? 0xffffffffu : ((1u << rem) - 1u);
}

__host__ __device__ __forceinline__ uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

// Bits x0 = (p + b) mod L0 of a row, b = 0..31, for 0 <= p < L0.  A window
// that crosses the end of the row continues from x0 = 0 (several times over
// when L0 < 32).
__host__ __device__ __forceinline__ uint32_t row_window(const uint32_t *row, int W, int L0, int p) {
    uint32_t out = 0;
    int got = 0;
    while (got < 32) {
        int take = L0 - p;
        if (take > 32 - got) take = 32 - got;
        const int wi = p >> 5, sh = p & 31;
        const uint32_t lo = row[wi];
        const uint32_t hi = (sh != 0 && wi + 1 < W) ? row[wi + 1] : 0u;
        uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
        if (take < 32) v &= (1u << take) - 1u;
        out |= v << got;
        got += take;
        p = 0;
    }
    return out;
}

// The row d steps along axis a >= 1 from row R (0 <= d < L[a]).
__host__ __device__ __forceinline__ int32_t shift_row(const MeasGeom &m, int32_t R, int a, int d) {
    const int32_t x = (R / m.rstride[a]) % m.L[a];
    int32_t nx = x + d;
    if (nx >= m.L[a]) nx -= m.L[a];
    return R + (nx - x) * m.rstride[a];
}

// Word w of row R of `plane` translated by d steps along axis a: bit b holds
// the variable at (x + d e_a) for the site x of bit b (0 <= d < L[a]).
__host__ __device__ __forceinline__ uint32_t shifted_word(const uint32_t *plane, const MeasGeom &m, int32_t R, int w,
                                                          int a, int d) {
    if (a == 0) {
        int p = 32 * w + d;
        if (p >= m.L[0]) p -= m.L[0];
        return row_window(plane + (int64_t)R * m.W, m.W, m.L[0], p);
    }
    return plane[(int64_t)shift_row(m, R, a, d) * m.W + w];
}

}  // namespace

// The work of one lane on one plane word idx = R * W + w, shared by the
// kernels below and by host-side checks of their indexing.
namespace meas {

// Plane v * B + p gets bit p of variable v of every site (V variables per
// site in lattice order: sites V = 1, links V = D with variable = direction).
template <int V, int B>
__host__ __device__ __forceinline__ void pack_word(const uint8_t *lat, const MeasGeom &m, uint32_t *bits,
                                                   int64_t idx) {
    const int64_t total = (int64_t)m.R * m.W;
    const int32_t R = (int32_t)(idx / m.W);
    const int w = (int)(idx - (int64_t)R * m.W);
    int n = m.L[0] - 32 * w;
    if (n > 32) n = 32;
    const uint8_t *src = lat + ((int64_t)R * m.L[0] + 32 * w) * V;
    uint32_t acc[V * B];
#pragma unroll
    for (int k = 0; k < V * B; ++k) acc[k] = 0;
    for (int b = 0; b < n; ++b) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const uint32_t s = src[b * V + v];
#pragma unroll
            for (int p = 0; p < B; ++p) acc[v * B + p] |= ((s >> p) & 1u) << b;
        }
    }
#pragma unroll
    for (int k = 0; k < V * B; ++k) bits[(int64_t)k * total + idx] = acc[k];
}

// Pair k, numbered axis after axis with r = 0..L_mu/2 (models/mod.rs:50-59).
__host__ __device__ __forceinline__ void pair_of(const MeasGeom &m, int k, int &mu, int &r) {
    mu = 0;
    r = k;
    while (r > m.L[mu] / 2) {
        r -= m.L[mu] / 2 + 1;
        ++mu;
    }
}

// Sites of word idx where s(x) == s(x + r e_mu), over P state planes.
template <int P>
__host__ __device__ __forceinline__ uint32_t pair_word(const uint32_t *bits, const MeasGeom &m, int mu, int r,
                                                       int64_t idx) {
    const int64_t total = (int64_t)m.R * m.W;
    const int32_t R = (int32_t)(idx / m.W);
    const int w = (int)(idx - (int64_t)R * m.W);
    uint32_t e = valid_bits(m.L[0], w);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const uint32_t *plane = bits + (int64_t)p * total;
        e &= ~(plane[idx] ^ shifted_word(plane, m, R, w, mu, r));
    }
    return popc32(e);
}

// Lines along dir with link product -1 in item idx: a row for dir = 0 (idx < R),
// a (line row, word) for dir >= 1 (idx < R / L_dir * W).
__host__ __device__ __forceinline__ uint32_t polyakov_item(const uint32_t *plane, const MeasGeom &m, int dir,
                                                           int64_t idx) {
    if (dir == 0) {
        uint32_t x = 0;
        for (int w = 0; w < m.W; ++w) x ^= plane[idx * m.W + w];
        return popc32(x) & 1u;
    }
    const int32_t rs = m.rstride[dir], L = m.L[dir];
    const int32_t j = (int32_t)(idx / m.W);
    const int w = (int)(idx - (int64_t)j * m.W);
    const int32_t R0 = (j / rs) * rs * L + (j % rs);  // the line's row with x_dir = 0
    uint32_t x = 0;
    for (int k = 0; k < L; ++k) x ^= plane[(int64_t)(R0 + k * rs) * m.W + w];
    return popc32(x & valid_bits(m.L[0], w));
}

// Line products Lr_mu(x) = XOR_{k<r} U_mu(x + k e_mu), r = 1..max, of word idx:
// plane mu * max + r - 1.
__host__ __device__ __forceinline__ void lines_word(const uint32_t *links, const MeasGeom &m, int max,
                                                    uint32_t *lines, int mu, int64_t idx) {
    const int64_t total = (int64_t)m.R * m.W;
    const int32_t R = (int32_t)(idx / m.W);
    const int w = (int)(idx - (int64_t)R * m.W);
    const uint32_t *u = links + (int64_t)mu * total;
    uint32_t acc = 0;
    for (int r = 1; r <= max; ++r) {
        acc ^= shifted_word(u, m, R, w, mu, r - 1);
        lines[((int64_t)mu * max + r - 1) * total + idx] = acc;
    }
}

// Combination k = ((r - 1) * max + (t - 1)) * D (D - 1) + o over the ordered
// direction pairs o = (mu, nu), mu != nu, in the reference's order
// (gauge/mod.rs:244-246).
__host__ __device__ __forceinline__ void combo_of(const MeasGeom &m, int max, int k, int &r, int &t, int &mu,
                                                  int &nu) {
    const int O = m.D * (m.D - 1);
    const int o = k % O, rt = k / O;
    r = rt / max + 1;
    t = rt % max + 1;
    mu = o / (m.D - 1);
    const int jn = o % (m.D - 1);
    nu = jn + (jn >= mu ? 1 : 0);
}

// Loops of word idx with product -1: Lr_mu(x) ^ Lt_nu(x + r e_mu) ^ Lr_mu(x + t e_nu) ^ Lt_nu(x).
__host__ __device__ __forceinline__ uint32_t wilson_word(const uint32_t *lines, const MeasGeom &m, int max, int r,
                                                         int t, int mu, int nu, int64_t idx) {
    const int64_t total = (int64_t)m.R * m.W;
    const uint32_t *lr = lines + ((int64_t)mu * max + r - 1) * total;
    const uint32_t *lt = lines + ((int64_t)nu * max + t - 1) * total;
    const int32_t R = (int32_t)(idx / m.W);
    const int w = (int)(idx - (int64_t)R * m.W);
    const uint32_t x = lr[idx] ^ lt[idx] ^ shifted_word(lt, m, R, w, mu, r) ^ shifted_word(lr, m, R, w, nu, t);
    return popc32(x & valid_bits(m.L[0], w));
}

}  // namespace meas

namespace {

// Sum of per-lane counts of one workgroup into *out: wave reduction, one
// atomic per wave.
__device__ __forceinline__ void add_count(uint32_t cnt, unsigned long long *out) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

template <int V, int B>
__global__ __launch_bounds__(MEAS_BLOCK) void meas_pack_kernel(const uint8_t *__restrict__ lat, MeasGeom m,
                                                               uint32_t *__restrict__ bits) {
    const int64_t idx = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x;
    if (idx < (int64_t)m.R * m.W) meas::pack_word<V, B>(lat, m, bits, idx);
}

// Agreements of pair k = k0 + blockIdx.y.
template <int P>
__global__ __launch_bounds__(MEAS_BLOCK) void meas_pairs_kernel(const uint32_t *__restrict__ bits, MeasGeom m,
                                                                int k0, unsigned long long *__restrict__ agree) {
    const int k = k0 + (int)blockIdx.y;
    int mu, r;
    meas::pair_of(m, k, mu, r);
    const int64_t total = (int64_t)m.R * m.W;
    uint32_t cnt = 0;
    for (int64_t idx = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * MEAS_BLOCK)
        cnt += meas::pair_word<P>(bits, m, mu, r, idx);
    add_count(cnt, agree + k);
}

__global__ __launch_bounds__(MEAS_BLOCK) void meas_polyakov_kernel(const uint32_t *__restrict__ plane, MeasGeom m,
                                                                   int dir, int64_t items,
                                                                   unsigned long long *__restrict__ odd) {
    uint32_t cnt = 0;
    for (int64_t idx = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x; idx < items;
         idx += (int64_t)gridDim.x * MEAS_BLOCK)
        cnt += meas::polyakov_item(plane, m, dir, idx);
    add_count(cnt, odd);
}

__global__ __launch_bounds__(MEAS_BLOCK) void meas_lines_kernel(const uint32_t *__restrict__ links, MeasGeom m,
                                                                int max, uint32_t *__restrict__ lines) {
    const int64_t idx = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x;
    if (idx < (int64_t)m.R * m.W) meas::lines_word(links, m, max, lines, (int)blockIdx.y, idx);
}

// Loops with product -1 for combination k = k0 + blockIdx.y.
__global__ __launch_bounds__(MEAS_BLOCK) void meas_wilson_kernel(const uint32_t *__restrict__ lines, MeasGeom m,
                                                                 int max, int k0, unsigned long long *__restrict__ odd) {
    const int k = k0 + (int)blockIdx.y;
    int r, t, mu, nu;
    meas::combo_of(m, max, k, r, t, mu, nu);
    const int64_t total = (int64_t)m.R * m.W;
    uint32_t cnt = 0;
    for (int64_t idx = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * MEAS_BLOCK)
        cnt += meas::wilson_word(lines, m, max, r, t, mu, nu, idx);
    add_count(cnt, odd + k);
}

// Zero n counters (a fill kernel rather than hipMemsetAsync, so the clear is
// an ordinary kernel node should a caller capture the stream).
__global__ __launch_bounds__(MEAS_BLOCK) void meas_clear_kernel(unsigned long long *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * MEAS_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * MEAS_BLOCK)
        p[i] = 0ull;
}

constexpr int MAX_GRID_Y = 65535;

// Workgroups along x for `ny` independent sums over `total` words: enough to
// fill the chip with a few thousand workgroups in all.
uint32_t x_blocks(int64_t total, int64_t ny) {
    const int64_t need = (total + MEAS_BLOCK - 1) / MEAS_BLOCK;
    int64_t want = (4096 + ny - 1) / ny;
    if (want < 1) want = 1;
    return (uint32_t)(need < want ? need : want);
}

}  // namespace

MeasGeom meas_geom(int D, const int64_t *shape) {
    MeasGeom m{};
    m.D = D;
    for (int a = 0; a < MAX_D; ++a) m.L[a] = a < D ? (int32_t)shape[a] : 1;
    m.N = 1;
    for (int a = 0; a < D; ++a) m.N *= shape[a];
    m.R = (int32_t)(m.N / m.L[0]);
    m.W = (m.L[0] + 31) / 32;
    m.rstride[0] = 0;
    int32_t s = 1;
    for (int a = 1; a < MAX_D; ++a) {
        m.rstride[a] = s;
        s *= m.L[a];
    }
    return m;
}

int meas_state_planes(int q) {
    int p = 1;
    while ((1 << p) < q) ++p;
    return p;
}

void launch_meas_pack(const uint8_t *lat, const MeasGeom &m, int vars_per_site, int bits_per_var, uint32_t *bits,
                      hipStream_t s) {
    const int64_t total = (int64_t)m.R * m.W;
    const dim3 grid((uint32_t)((total + MEAS_BLOCK - 1) / MEAS_BLOCK)), block(MEAS_BLOCK);
    switch (vars_per_site * 8 + bits_per_var) {
        case 9: meas_pack_kernel<1, 1><<<grid, block, 0, s>>>(lat, m, bits); break;
        case 10: meas_pack_kernel<1, 2><<<grid, block, 0, s>>>(lat, m, bits); break;
        case 11: meas_pack_kernel<1, 3><<<grid, block, 0, s>>>(lat, m, bits); break;
        case 12: meas_pack_kernel<1, 4><<<grid, block, 0, s>>>(lat, m, bits); break;
        case 17: meas_pack_kernel<2, 1><<<grid, block, 0, s>>>(lat, m, bits); break;
        case 25: meas_pack_kernel<3, 1><<<grid, block, 0, s>>>(lat, m, bits); break;
        default: meas_pack_kernel<4, 1><<<grid, block, 0, s>>>(lat, m, bits); break;
    }
}

void launch_meas_clear(unsigned long long *p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t need = (n + MEAS_BLOCK - 1) / MEAS_BLOCK;
    meas_clear_kernel<<<dim3((uint32_t)(need < 64 ? need : 64)), dim3(MEAS_BLOCK), 0, s>>>(p, n);
}

void launch_meas_pairs(const uint32_t *bits, const MeasGeom &m, int planes, int npairs, unsigned long long *agree,
                       hipStream_t s) {
    const int64_t total = (int64_t)m.R * m.W;
    for (int k0 = 0; k0 < npairs; k0 += MAX_GRID_Y) {
        const int ny = npairs - k0 < MAX_GRID_Y ? npairs - k0 : MAX_GRID_Y;
        const dim3 grid(x_blocks(total, ny), (uint32_t)ny), block(MEAS_BLOCK);
        switch (planes) {
            case 1: meas_pairs_kernel<1><<<grid, block, 0, s>>>(bits, m, k0, agree); break;
            case 2: meas_pairs_kernel<2><<<grid, block, 0, s>>>(bits, m, k0, agree); break;
            case 3: meas_pairs_kernel<3><<<grid, block, 0, s>>>(bits, m, k0, agree); break;
            default: meas_pairs_kernel<4><<<grid, block, 0, s>>>(bits, m, k0, agree); break;
        }
    }
}

int64_t meas_polyakov_items(const MeasGeom &m, int dir) {
    return dir == 0 ? (int64_t)m.R : (int64_t)(m.R / m.L[dir]) * m.W;
}

void launch_meas_polyakov(const uint32_t *plane, const MeasGeom &m, int dir, unsigned long long *odd, hipStream_t s) {
    const int64_t items = meas_polyakov_items(m, dir);
    meas_polyakov_kernel<<<dim3(x_blocks(items, 1)), dim3(MEAS_BLOCK), 0, s>>>(plane, m, dir, items, odd);
}

void launch_meas_lines(const uint32_t *links, const MeasGeom &m, int max, uint32_t *lines, hipStream_t s) {
    const int64_t total = (int64_t)m.R * m.W;
    const dim3 grid((uint32_t)((total + MEAS_BLOCK - 1) / MEAS_BLOCK), (uint32_t)m.D), block(MEAS_BLOCK);
    meas_lines_kernel<<<grid, block, 0, s>>>(links, m, max, lines);
}

void launch_meas_wilson(const uint32_t *lines, const MeasGeom &m, int max, unsigned long long *odd, hipStream_t s) {
    const int64_t total = (int64_t)m.R * m.W;
    const int ncomb = max * max * m.D * (m.D - 1);
    for (int k0 = 0; k0 < ncomb; k0 += MAX_GRID_Y) {
        const int ny = ncomb - k0 < MAX_GRID_Y ? ncomb - k0 : MAX_GRID_Y;
        const dim3 grid(x_blocks(total, ny), (uint32_t)ny), block(MEAS_BLOCK);
        meas_wilson_kernel<<<grid, block, 0, s>>>(lines, m, max, k0, odd);
    }
}

}  // namespace plq
