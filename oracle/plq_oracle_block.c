/*
 * plq_oracle_block.c -- CPU oracle of the learned block proposal.
 *
 * TEST INFRASTRUCTURE ONLY (see plq_oracle.h).  The learned block proposal
 * has no counterpart in the reference; what this file restates is the
 * sampler it plugs into -- the 2D Ising action of ising/mod.rs:97-143 with
 * the keyed draws of rng.wgsl:34-52 -- and the network the device kernels
 * evaluate, so that the device can be checked against it:
 *
 *   orc_expf              the deterministic f32 exponential both sides use in
 *                         the Metropolis-Hastings test: Cody-Waite range
 *                         reduction (Cody & Waite, "Software Manual for the
 *                         Elementary Functions", 1980) with a split ln 2 and a
 *                         degree-7 Taylor polynomial on |r| <= ln2/2, scaled
 *                         by 2^n through the exponent bits.  Only IEEE single
 *                         adds and multiplies, so any compiler that keeps
 *                         them unfused computes the same bits.
 *   orc_blocknet_logits   forward pass of the 3-layer CNN (boundary ring ->
 *                         64 logits; middle layer dilated by 2) in FP64,
 *                         optionally reproducing the
 *                         device's BF16 rounding of weights and activations.
 *   orc_blocknet_loss_grad  mean binary cross-entropy of the interior given
 *                         the ring, and its gradient (straight through the
 *                         BF16 roundings when they are reproduced).
 *   orc_block_mh_pass     one block-colour pass of the block
 *                         Metropolis-Hastings update with given logits, in
 *                         the exact f32 arithmetic of the device kernel.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "plq_oracle.h"

#define BN_B ORC_BN_BLOCK
/* the middle layer's taps are two sites apart (dilation 2) */
#define DIL2 2
#define BN_C ORC_BN_CHANNELS
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + 9 * BN_C)
#define OFF_W2 (OFF_B1 + BN_C)
#define OFF_B2 (OFF_W2 + 9 * BN_C * BN_C)
#define OFF_W3 (OFF_B2 + BN_C)
#define OFF_B3 (OFF_W3 + 9 * BN_C)

float orc_expf(float x) {
    if (x > 88.0f) return INFINITY;
    if (x < -87.0f) return 0.0f;
    const float log2e = 1.44269502f;        /* 0x3fb8aa3b */
    const float ln2_hi = 0.693145751953125f; /* 0x3f317200: 15 significant bits, n*ln2_hi is exact */
    const float ln2_lo = 1.42860677e-06f;    /* 0x35bfbe8e: ln 2 - ln2_hi */
    float n = rintf(x * log2e);
    float r = x - n * ln2_hi;
    r = r - n * ln2_lo;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int e = (int)n;
    uint32_t bits = (uint32_t)(e + 127) << 23;
    float scale;
    memcpy(&scale, &bits, 4);
    return p * scale;
}

float orc_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

static double rnd(double v, int round) { return round ? (double)orc_bf16((float)v) : v; }

/* Site (X, Y) of the lattice, periodic. */
static int64_t site_at(int64_t L0, int64_t L1, int64_t X, int64_t Y) {
    X = ((X % L0) + L0) % L0;
    Y = ((Y % L1) + L1) % L1;
    return Y * L0 + X;
}

/* The 10x10 input patch of block (bx, by): ring spins +-1 (state 0 -> +1),
 * interior and corners 0. */
static void gather_patch(int64_t L0, int64_t L1, const uint8_t *state, int64_t bx, int64_t by, double in[10][10]) {
    memset(in, 0, sizeof(double) * 100);
    for (int k = 0; k < BN_B; ++k) {
        int64_t X = bx * BN_B + k, Y = by * BN_B + k;
        in[0][k + 1] = 1.0 - 2.0 * state[site_at(L0, L1, X, by * BN_B - 1)];
        in[9][k + 1] = 1.0 - 2.0 * state[site_at(L0, L1, X, by * BN_B + BN_B)];
        in[k + 1][0] = 1.0 - 2.0 * state[site_at(L0, L1, bx * BN_B - 1, Y)];
        in[k + 1][9] = 1.0 - 2.0 * state[site_at(L0, L1, bx * BN_B + BN_B, Y)];
    }
}

typedef struct {
    double w1[9][BN_C], b1[BN_C], w2[9][BN_C][BN_C], b2[BN_C], w3[9][BN_C], b3;
} net_t;

static void load_net(const float *p, int round, net_t *n) {
    for (int t = 0; t < 9; ++t)
        for (int c = 0; c < BN_C; ++c) {
            n->w1[t][c] = rnd(p[OFF_W1 + t * BN_C + c], round);
            n->w3[t][c] = rnd(p[OFF_W3 + t * BN_C + c], round);
            for (int ci = 0; ci < BN_C; ++ci)
                n->w2[t][c][ci] = rnd(p[OFF_W2 + (t * BN_C + c) * BN_C + ci], round);
        }
    for (int c = 0; c < BN_C; ++c) {
        n->b1[c] = p[OFF_B1 + c];
        n->b2[c] = p[OFF_B2 + c];
    }
    n->b3 = p[OFF_B3];
}

/* Activations of one block.  h1, h2 indexed [y][x][c] over the 8x8 grid. */
typedef struct {
    double in[10][10];
    double h1[BN_B][BN_B][BN_C], h2[BN_B][BN_B][BN_C], l[BN_B][BN_B];
} act_t;

static int inside(int y, int x) { return y >= 0 && y < BN_B && x >= 0 && x < BN_B; }

static void forward(const net_t *n, int round, act_t *a) {
    for (int y = 0; y < BN_B; ++y)
        for (int x = 0; x < BN_B; ++x)
            for (int c = 0; c < BN_C; ++c) {
                double s = n->b1[c];
                for (int t = 0; t < 9; ++t) s += n->w1[t][c] * a->in[y + 1 + t / 3 - 1][x + 1 + t % 3 - 1];
                a->h1[y][x][c] = rnd(tanh(s), round);
            }
    for (int y = 0; y < BN_B; ++y)
        for (int x = 0; x < BN_B; ++x)
            for (int co = 0; co < BN_C; ++co) {
                double s = n->b2[co];
                for (int t = 0; t < 9; ++t) {
                    int yy = y + DIL2 * (t / 3 - 1), xx = x + DIL2 * (t % 3 - 1);
                    if (!inside(yy, xx)) continue;
                    for (int ci = 0; ci < BN_C; ++ci) s += n->w2[t][co][ci] * a->h1[yy][xx][ci];
                }
                a->h2[y][x][co] = rnd(tanh(s), round);
            }
    for (int y = 0; y < BN_B; ++y)
        for (int x = 0; x < BN_B; ++x) {
            double s = n->b3;
            for (int t = 0; t < 9; ++t) {
                int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
                if (!inside(yy, xx)) continue;
                for (int c = 0; c < BN_C; ++c) s += n->w3[t][c] * a->h2[yy][xx][c];
            }
            a->l[y][x] = s;
        }
}

void orc_blocknet_logits(const float *params, int round, int64_t L0, int64_t L1, const uint8_t *state,
                         double *out) {
    net_t *n = malloc(sizeof *n);
    act_t *a = malloc(sizeof *a);
    load_net(params, round, n);
    int64_t nb0 = L0 / BN_B, nb1 = L1 / BN_B;
    for (int64_t by = 0; by < nb1; ++by)
        for (int64_t bx = 0; bx < nb0; ++bx) {
            gather_patch(L0, L1, state, bx, by, a->in);
            forward(n, round, a);
            double *o = out + (by * nb0 + bx) * BN_B * BN_B;
            for (int p = 0; p < BN_B * BN_B; ++p) o[p] = a->l[p / BN_B][p % BN_B];
        }
    free(a);
    free(n);
}

double orc_blocknet_loss_grad(const float *params, int round, int64_t L0, int64_t L1, const uint8_t *state,
                              double *grad) {
    net_t *n = malloc(sizeof *n);
    act_t *a = malloc(sizeof *a);
    double(*g2)[BN_B][BN_C] = malloc(sizeof(double) * BN_B * BN_B * BN_C);
    double(*g1)[BN_B][BN_C] = malloc(sizeof(double) * BN_B * BN_B * BN_C);
    load_net(params, round, n);
    memset(grad, 0, sizeof(double) * ORC_BN_PARAMS);
    int64_t nb0 = L0 / BN_B, nb1 = L1 / BN_B;
    double loss = 0.0;
    for (int64_t by = 0; by < nb1; ++by)
        for (int64_t bx = 0; bx < nb0; ++bx) {
            gather_patch(L0, L1, state, bx, by, a->in);
            forward(n, round, a);
            double g3[BN_B][BN_B];
            for (int y = 0; y < BN_B; ++y)
                for (int x = 0; x < BN_B; ++x) {
                    double l = a->l[y][x];
                    double yv = state[site_at(L0, L1, bx * BN_B + x, by * BN_B + y)] == 0 ? 1.0 : 0.0;
                    /* BCE with logits: softplus(l) - y l, written stably */
                    loss += (l > 0 ? l + log1p(exp(-l)) : log1p(exp(l))) - yv * l;
                    g3[y][x] = 1.0 / (1.0 + exp(-l)) - yv;
                    grad[OFF_B3] += g3[y][x];
                }
            /* layer 3 */
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < BN_C; ++c) {
                    double s = 0.0;
                    for (int y = 0; y < BN_B; ++y)
                        for (int x = 0; x < BN_B; ++x) {
                            int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
                            if (inside(yy, xx)) s += g3[y][x] * a->h2[yy][xx][c];
                        }
                    grad[OFF_W3 + t * BN_C + c] += s;
                }
            for (int y = 0; y < BN_B; ++y)
                for (int x = 0; x < BN_B; ++x)
                    for (int c = 0; c < BN_C; ++c) {
                        double s = 0.0;
                        for (int t = 0; t < 9; ++t) {
                            int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
                            if (inside(yy, xx)) s += n->w3[t][c] * g3[yy][xx];
                        }
                        double h = a->h2[y][x][c];
                        g2[y][x][c] = s * (1.0 - h * h);
                        grad[OFF_B2 + c] += g2[y][x][c];
                    }
            /* the device feeds g2 to the matrix cores in BF16 */
            if (round)
                for (int y = 0; y < BN_B; ++y)
                    for (int x = 0; x < BN_B; ++x)
                        for (int c = 0; c < BN_C; ++c) g2[y][x][c] = rnd(g2[y][x][c], 1);
            /* layer 2 */
            for (int t = 0; t < 9; ++t)
                for (int co = 0; co < BN_C; ++co)
                    for (int ci = 0; ci < BN_C; ++ci) {
                        double s = 0.0;
                        for (int y = 0; y < BN_B; ++y)
                            for (int x = 0; x < BN_B; ++x) {
                                int yy = y + DIL2 * (t / 3 - 1), xx = x + DIL2 * (t % 3 - 1);
                                if (inside(yy, xx)) s += g2[y][x][co] * a->h1[yy][xx][ci];
                            }
                        grad[OFF_W2 + (t * BN_C + co) * BN_C + ci] += s;
                    }
            for (int y = 0; y < BN_B; ++y)
                for (int x = 0; x < BN_B; ++x)
                    for (int ci = 0; ci < BN_C; ++ci) {
                        double s = 0.0;
                        for (int t = 0; t < 9; ++t) {
                            int yy = y - DIL2 * (t / 3 - 1), xx = x - DIL2 * (t % 3 - 1);
                            if (!inside(yy, xx)) continue;
                            for (int co = 0; co < BN_C; ++co) s += n->w2[t][co][ci] * g2[yy][xx][co];
                        }
                        double h = a->h1[y][x][ci];
                        g1[y][x][ci] = s * (1.0 - h * h);
                        grad[OFF_B1 + ci] += g1[y][x][ci];
                    }
            /* layer 1 */
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < BN_C; ++c) {
                    double s = 0.0;
                    for (int y = 0; y < BN_B; ++y)
                        for (int x = 0; x < BN_B; ++x) s += g1[y][x][c] * a->in[y + 1 + t / 3 - 1][x + 1 + t % 3 - 1];
                    grad[OFF_W1 + t * BN_C + c] += s;
                }
        }
    double nsites = (double)(L0 * L1);
    for (int e = 0; e < ORC_BN_PARAMS; ++e) grad[e] /= nsites;
    free(g1);
    free(g2);
    free(a);
    free(n);
    return loss / nsites;
}

int64_t orc_block_mh_pass(int64_t L0, int64_t L1, uint8_t *state, const float *logits, int colour, uint32_t seed32,
                          uint32_t bstep, double beta, double j, double h) {
    int64_t nb0 = L0 / BN_B, nb1 = L1 / BN_B, accepted = 0;
    const uint32_t kp = seed32 ^ ORC_BLOCK_PROPOSAL_KEY, ka = seed32 ^ ORC_BLOCK_ACCEPT_KEY;
    for (int64_t by = 0; by < nb1; ++by)
        for (int64_t bx = 0; bx < nb0; ++bx) {
            if (((bx + by) & 1) != colour) continue;
            int64_t b = by * nb0 + bx;
            int cur[64], prop[64];
            float v[64];
            for (int p = 0; p < 64; ++p) {
                int64_t site = site_at(L0, L1, bx * BN_B + p % BN_B, by * BN_B + p / BN_B);
                float l = logits[b * 64 + p];
                uint32_t m = orc_keyed_bits(kp, (uint32_t)site, bstep);
                float pu = 1.0f / (1.0f + orc_expf(-l));
                cur[p] = state[site];
                prop[p] = ((float)m < pu * 16777216.0f) ? 0 : 1;
                v[p] = (float)((cur[p] == 0) - (prop[p] == 0)) * l;
            }
            /* log q(x) - log q(x') summed in the device's butterfly order */
            for (int off = 32; off > 0; off >>= 1) {
                float w[64];
                for (int i = 0; i < 64; ++i) w[i] = v[i] + v[i ^ off];
                memcpy(v, w, sizeof w);
            }
            float S = v[0];
            long long dbond = 0, dmag = 0;
            for (int p = 0; p < 64; ++p) {
                int px = p % BN_B, py = p / BN_B;
                int s = 1 - 2 * cur[p], t = 1 - 2 * prop[p];
                dmag += t - s;
                if (px < BN_B - 1) dbond += t * (1 - 2 * prop[p + 1]) - s * (1 - 2 * cur[p + 1]);
                if (py < BN_B - 1) dbond += t * (1 - 2 * prop[p + BN_B]) - s * (1 - 2 * cur[p + BN_B]);
                int ring = 0;
                int64_t X = bx * BN_B + px, Y = by * BN_B + py;
                if (px == 0) ring += 1 - 2 * state[site_at(L0, L1, X - 1, Y)];
                if (px == BN_B - 1) ring += 1 - 2 * state[site_at(L0, L1, X + 1, Y)];
                if (py == 0) ring += 1 - 2 * state[site_at(L0, L1, X, Y - 1)];
                if (py == BN_B - 1) ring += 1 - 2 * state[site_at(L0, L1, X, Y + 1)];
                dbond += (long long)(t - s) * ring;
            }
            double de = j * (double)dbond + h * (double)dmag; /* -Delta E */
            float A = (float)(beta * de) + S;
            uint32_t ma = orc_keyed_bits(ka, (uint32_t)b, bstep);
            int acc = A >= 0.0f || (float)ma < orc_expf(A) * 16777216.0f;
            if (acc) {
                accepted += 1;
                for (int p = 0; p < 64; ++p)
                    state[site_at(L0, L1, bx * BN_B + p % BN_B, by * BN_B + p / BN_B)] = (uint8_t)prop[p];
            }
        }
    return accepted;
}
