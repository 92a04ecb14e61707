// plq_link.hip -- checkerboard kernels for link fields (Z2 gauge) on gfx950.
//
// The reference colours a link by its direction and its base site's parity
// and runs 2D passes per sweep, dir-major (link_schedule.wgsl:23-40,
// updater.rs:325-343); each pass prices a link against the 2(D-1) staples of
// the plaquettes through it (gauge/model.wgsl:136-147), read through a
// 6(D-1)-entry u32 table per link (gauge/gpu.rs:54-110; 24 B/link... 72 B in
// 4D).  Here links are stored in 2D int8 arrays, one per (direction, base
// parity), each in the colour-interleaved site layout of plq_common.hpp, and
// the staple links are reached by coordinate steps computed in registers.
//
// For link (x, mu) and every nu != mu the two plaquettes in the (mu, nu)
// plane contribute
//     forward : U_nu(x+mu) U_mu(x+nu) U_nu(x)
//     backward: U_mu(x-nu) U_nu(x-nu+mu) U_nu(x-nu)
// and the energy change depends only on how many of these 2(D-1) products
// are negative (plus the link's own state), which indexes the host-built
// threshold table.  The sum of products is independent of the order the
// reference lists the staple groups in (lattice.rs:136-163), so the class is
// exactly the reference's staple_sum.
#include <hip/hip_runtime.h>

#include "plq_common.hpp"
#include "plq_internal.hpp"

namespace plq {

// A site in (x0, row) coordinates with its row coordinates cached.
struct SitePos {
    int32_t x0;
    int32_t r;
    int32_t x[MAX_D];
};

__device__ __forceinline__ SitePos step_site(const Geom &g, const SitePos &s, int mu, int sign) {
    SitePos t = s;
    if (mu == 0) {
        int32_t l = g.L[0];
        int32_t v = s.x0 + sign;
        t.x0 = v == l ? 0 : (v < 0 ? l - 1 : v);
    } else {
        t.r = row_step(g, s.r, s.x, mu, sign);
        int32_t l = g.L[mu];
        int32_t v = s.x[mu] + sign;
        t.x[mu] = v == l ? 0 : (v < 0 ? l - 1 : v);
    }
    return t;
}

__device__ __forceinline__ int site_colour(const Geom &g, const SitePos &s) {
    int32_t par = s.x0;
    for (int mu = 1; mu < MAX_D; ++mu) par += s.x[mu];
    return par & 1;
}

__device__ __forceinline__ uint8_t link_at(uint8_t *const *arr, const Geom &g, const SitePos &s, int dir) {
    int c = site_colour(g, s);
    return arr[dir * 2 + c][(int64_t)s.r * g.H + (s.x0 >> 1)];
}

__global__ void pack_links_kernel(const uint8_t *__restrict__ lat, uint8_t *const *__restrict__ arr, Geom g) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    int32_t r = (int32_t)(s / g.L[0]);
    int32_t x0 = (int32_t)(s - (int64_t)r * g.L[0]);
    int32_t x[MAX_D];
    row_coords(g, r, x);
    int32_t par = x0;
    for (int mu = 1; mu < MAX_D; ++mu) par += x[mu];
    int64_t idx = (int64_t)r * g.H + (x0 >> 1);
    for (int mu = 0; mu < g.D; ++mu) arr[mu * 2 + (par & 1)][idx] = lat[s * g.D + mu];
}

__global__ void unpack_links_kernel(uint8_t *__restrict__ lat, uint8_t *const *__restrict__ arr, Geom g) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    int32_t r = (int32_t)(s / g.L[0]);
    int32_t x0 = (int32_t)(s - (int64_t)r * g.L[0]);
    int32_t x[MAX_D];
    row_coords(g, r, x);
    int32_t par = x0;
    for (int mu = 1; mu < MAX_D; ++mu) par += x[mu];
    int64_t idx = (int64_t)r * g.H + (x0 >> 1);
    for (int mu = 0; mu < g.D; ++mu) lat[s * g.D + mu] = arr[mu * 2 + (par & 1)][idx];
}

// One (dir, colour) pass: a lane per base site of that colour.
template <int D, int MODE, int MU>
__global__ __launch_bounds__(256) void link_pass_kernel(uint8_t *const *__restrict__ arr, Geom g, int c,
                                                        uint32_t seed, uint32_t skey,
                                                        const uint32_t *__restrict__ table) {
    constexpr int mu = MU;  // a template parameter, so every coordinate index is static (no scratch)
    constexpr int K = 2 * (D - 1);
    __shared__ uint32_t T[2 * (K + 1)];
    if (threadIdx.x < 2 * (K + 1)) T[threadIdx.x] = table[threadIdx.x];
    __syncthreads();
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (int64_t)g.R * g.H) return;
    SitePos s;
    s.r = (int32_t)(w / g.H);
    int32_t i = (int32_t)(w - (int64_t)s.r * g.H);
    row_coords(g, s.r, s.x);
    int32_t prs = 0;
    for (int a = 1; a < D; ++a) prs += s.x[a];
    s.x0 = 2 * i + ((prs + c) & 1);
    int neg = 0;
    SitePos xp = step_site(g, s, mu, +1);
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        SitePos xpn = step_site(g, s, nu, +1);
        SitePos xmn = step_site(g, s, nu, -1);
        SitePos xmnp = step_site(g, xmn, mu, +1);
        int f = link_at(arr, g, xp, nu) ^ link_at(arr, g, xpn, mu) ^ link_at(arr, g, s, nu);
        int b = link_at(arr, g, xmn, mu) ^ link_at(arr, g, xmnp, nu) ^ link_at(arr, g, xmn, nu);
        neg += f + b;  // states are 0/1 and a product is negative iff the xor is 1
    }
    uint8_t *own = arr[mu * 2 + c];
    int64_t idx = (int64_t)s.r * g.H + i;
    uint32_t cur = own[idx];
    uint32_t t = T[cur * (K + 1) + neg];
    bool hit;
    if (t == 0u) hit = false;
    else if (t >= DRAW_ONE) hit = true;
    else {
        uint32_t link = (uint32_t)(((int64_t)s.r * g.L[0] + s.x0) * D + mu);
        hit = keyed_bits_k(seed, link, skey) < t;
    }
    uint32_t nxt = (MODE == 0) ? (cur ^ (hit ? 1u : 0u)) : (hit ? 0u : 1u);
    own[idx] = (uint8_t)nxt;
}

// Signed plaquette sum (gauge/mod.rs:99-122): one lane per site, all pairs.
template <int D>
__global__ __launch_bounds__(256) void plaquette_sum_kernel(uint8_t *const *__restrict__ arr, Geom g,
                                                            long long *__restrict__ partials, int64_t span) {
    __shared__ long long lds[16];
    long long acc = 0;
    int64_t begin = (int64_t)blockIdx.x * span;
    int64_t end = begin + span;
    if (end > g.N) end = g.N;
    for (int64_t sidx = begin + threadIdx.x; sidx < end; sidx += blockDim.x) {
        SitePos s;
        s.r = (int32_t)(sidx / g.L[0]);
        s.x0 = (int32_t)(sidx - (int64_t)s.r * g.L[0]);
        row_coords(g, s.r, s.x);
        for (int mu = 0; mu < D; ++mu)
            for (int nu = mu + 1; nu < D; ++nu) {
                SitePos sm = step_site(g, s, mu, +1), sn = step_site(g, s, nu, +1);
                int x = link_at(arr, g, s, mu) ^ link_at(arr, g, sm, nu) ^ link_at(arr, g, sn, mu) ^
                        link_at(arr, g, s, nu);
                acc += 1 - 2 * x;
            }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) lds[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) t += lds[wv];
        partials[(int64_t)blockIdx.x * REC_K] = t;
        for (int k = 1; k < REC_K; ++k) partials[(int64_t)blockIdx.x * REC_K + k] = 0;
    }
}

void launch_pack_links(const uint8_t *lat, uint8_t *const *arr, const Geom &g, hipStream_t s) {
    int64_t blocks = (g.N + 255) / 256;
    pack_links_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(lat, arr, g);
}

void launch_unpack_links(uint8_t *lat, uint8_t *const *arr, const Geom &g, hipStream_t s) {
    int64_t blocks = (g.N + 255) / 256;
    unpack_links_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(lat, arr, g);
}

void launch_link_pass(const LinkPassArgs &a) {
    int64_t work = (int64_t)a.g.R * a.g.H;
    dim3 grid((uint32_t)((work + 255) / 256)), block(256);
#define PLQ_LINK_MU(DD, MD, M) \
    link_pass_kernel<DD, MD, M><<<grid, block, 0, a.stream>>>(a.arr, a.g, a.color, a.seed, a.skey, a.table)
#define PLQ_LINK(DD, MD)                                 \
    switch (a.dir) {                                     \
        case 0: PLQ_LINK_MU(DD, MD, 0); break;           \
        case 1: PLQ_LINK_MU(DD, MD, 1); break;           \
        case 2: PLQ_LINK_MU(DD, MD, (2 < DD ? 2 : 0)); break; \
        default: PLQ_LINK_MU(DD, MD, (3 < DD ? 3 : 0)); break; \
    }
    switch (a.g.D * 2 + a.mode) {
        case 4: PLQ_LINK(2, 0); break;
        case 5: PLQ_LINK(2, 1); break;
        case 6: PLQ_LINK(3, 0); break;
        case 7: PLQ_LINK(3, 1); break;
        case 8: PLQ_LINK(4, 0); break;
        default: PLQ_LINK(4, 1); break;
    }
#undef PLQ_LINK
#undef PLQ_LINK_MU
}

void launch_plaquette_sum(uint8_t *const *arr, const Geom &g, long long *partials, int64_t max_blocks,
                          long long *rec, hipStream_t s) {
    int64_t need = (g.N + 255) / 256;
    int64_t G = need < max_blocks ? need : max_blocks;
    int64_t span = (g.N + G - 1) / G;
    switch (g.D) {
        case 2: plaquette_sum_kernel<2><<<dim3((uint32_t)G), dim3(256), 0, s>>>(arr, g, partials, span); break;
        case 3: plaquette_sum_kernel<3><<<dim3((uint32_t)G), dim3(256), 0, s>>>(arr, g, partials, span); break;
        default: plaquette_sum_kernel<4><<<dim3((uint32_t)G), dim3(256), 0, s>>>(arr, g, partials, span); break;
    }
    launch_finalize_record(partials, G, nullptr, 0, rec, s);
}

}  // namespace plq
