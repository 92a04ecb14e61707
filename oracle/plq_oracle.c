/*
 * plq_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see plq_oracle.h).
 *
 * A sequential restatement of the reference's CPU checkerboard and cluster
 * semantics driven by the reference's keyed device stream.  Every function
 * cites the reference lines it follows.  Compiled with -ffp-contract=off and
 * SSE2 doubles so that each f32/f64 operation rounds exactly once, as in the
 * Rust and WGSL sources.
 */
#include <omp.h>

#include "plq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define PROPOSAL_KEY 0x85ebca6bu /* metropolis.wgsl:30 */
#define BOND_KEY 0x27d4eb2fu     /* cluster_prelude.wgsl:29 */
#define ROOT_SALT 0x85ebca6bu    /* cluster_relabel.wgsl:33 */
#define SEED_SALT 0xc2b2ae35u    /* cluster_relabel.wgsl:37 */

/* ---------------------------------------------------------------- RNG --- */

/* rng.wgsl:34-42 */
uint32_t orc_lowbias32(uint32_t x) {
    uint32_t v = x;
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

/* rng.wgsl:47-52, the integer part: k >> 8 */
uint32_t orc_keyed_bits(uint32_t seed, uint32_t var, uint32_t sweep) {
    uint32_t k = seed;
    k = orc_lowbias32(k ^ var);
    k = orc_lowbias32(k ^ (sweep * 0x9e3779b9u));
    return k >> 8;
}

/* rng.wgsl:51: f32(k >> 8) * (1.0 / 16777216.0), exact */
float orc_keyed_uniform(uint32_t seed, uint32_t var, uint32_t sweep) {
    return (float)orc_keyed_bits(seed, var, sweep) * (1.0f / 16777216.0f);
}

/* device.rs:761-763 */
uint32_t orc_fold_seed(uint64_t seed) { return (uint32_t)(seed ^ (seed >> 32)); }

static double u64_of(uint32_t m) { return (double)m * (1.0 / 16777216.0); }

/* ----------------------------------------------------------- geometry --- */

int64_t orc_n_sites(int32_t dim, const int64_t *shape) {
    int64_t n = 1;
    for (int mu = 0; mu < dim; ++mu) n *= shape[mu];
    return n;
}

/* lattice.rs:55-100: mixed radix, axis 0 fastest, columns +0,-0,+1,-1,... */
void orc_neighbor_table(int32_t dim, const int64_t *shape, int64_t *out) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t stride[ORC_MAX_DIM];
    stride[0] = 1;
    for (int mu = 1; mu < dim; ++mu) stride[mu] = stride[mu - 1] * shape[mu - 1];
    for (int64_t site = 0; site < n; ++site) {
        for (int mu = 0; mu < dim; ++mu) {
            int64_t l = shape[mu], s = stride[mu];
            int64_t coord = (site / s) % l;
            int64_t base = site - coord * s;
            out[site * 2 * dim + 2 * mu] = base + ((coord + 1) % l) * s;
            out[site * 2 * dim + 2 * mu + 1] = base + ((coord + l - 1) % l) * s;
        }
    }
}

/* lattice.rs:98: coordinate-sum parity */
void orc_parity_table(int32_t dim, const int64_t *shape, uint8_t *out) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t stride[ORC_MAX_DIM];
    stride[0] = 1;
    for (int mu = 1; mu < dim; ++mu) stride[mu] = stride[mu - 1] * shape[mu - 1];
    for (int64_t site = 0; site < n; ++site) {
        int64_t sum = 0;
        for (int mu = 0; mu < dim; ++mu) sum += (site / stride[mu]) % shape[mu];
        out[site] = (uint8_t)(sum % 2);
    }
}

static int64_t fwd_neighbor(int32_t dim, const int64_t *shape, int64_t site, int mu) {
    (void)dim;
    int64_t s = 1;
    for (int a = 0; a < mu; ++a) s *= shape[a];
    int64_t l = shape[mu];
    int64_t coord = (site / s) % l;
    return site - coord * s + ((coord + 1) % l) * s;
}

/* lattice.rs:273-313 (pair ordinal, lexicographic) and 420-430 (link order) */
void orc_plaquette_links(int32_t dim, const int64_t *shape, int64_t plaquette, int64_t out[4]) {
    int64_t n_pairs = (int64_t)dim * (dim - 1) / 2;
    int64_t site = plaquette / n_pairs;
    int64_t rest = plaquette % n_pairs;
    int mu = 0;
    for (;;) {
        int64_t in_group = dim - 1 - mu;
        if (rest < in_group) break;
        rest -= in_group;
        ++mu;
    }
    int nu = mu + 1 + (int)rest;
    int64_t site_mu = fwd_neighbor(dim, shape, site, mu);
    int64_t site_nu = fwd_neighbor(dim, shape, site, nu);
    out[0] = site * dim + mu;
    out[1] = site_mu * dim + nu;
    out[2] = site_nu * dim + mu;
    out[3] = site * dim + nu;
}

/* lattice.rs:136-163: for each plaquette in index order, append to each of
 * its four links the other three in plaquette order */
void orc_staple_table(int32_t dim, const int64_t *shape, int64_t *out) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t n_links = n * dim;
    int64_t stride = 6 * (int64_t)(dim - 1);
    int64_t n_plaq = n * (int64_t)dim * (dim - 1) / 2;
    int32_t *filled = (int32_t *)calloc((size_t)n_links, sizeof(int32_t));
    for (int64_t p = 0; p < n_plaq; ++p) {
        int64_t links[4];
        orc_plaquette_links(dim, shape, p, links);
        for (int i = 0; i < 4; ++i) {
            int64_t link = links[i];
            int64_t base = link * stride + (int64_t)filled[link] * 3;
            int slot = 0;
            for (int jx = 0; jx < 4; ++jx)
                if (jx != i) out[base + slot++] = links[jx];
            filled[link] += 1;
        }
    }
    free(filled);
}

/* ------------------------------------------------------ energy deltas --- */

static int decode(uint8_t s) { return 1 - 2 * (int)s; } /* models/mod.rs:30-32 */

typedef struct {
    const orc_local_desc *d;
    uint8_t *state;
    const int64_t *nbr;     /* site fields */
    const int64_t *staples; /* link fields */
    int nn;                 /* 2*dim */
} ctx_t;

/* f64 profile: ising/mod.rs:119-143, potts/mod.rs:273-309, gauge/mod.rs:287-328 */
static double delta_f64(const ctx_t *c, int64_t var, int cur, int prop) {
    const orc_local_desc *d = c->d;
    if (prop == cur) return 0.0;
    if (d->model == ORC_MODEL_ISING) {
        int64_t ds = (int64_t)(decode((uint8_t)prop) - decode((uint8_t)cur));
        int64_t nsum = 0;
        for (int k = 0; k < c->nn; ++k) nsum += decode(c->state[c->nbr[var * c->nn + k]]);
        return -(double)ds * (d->j * (double)nsum + d->h);
    } else if (d->model == ORC_MODEL_POTTS) {
        int64_t change = 0;
        for (int k = 0; k < c->nn; ++k) {
            int s = c->state[c->nbr[var * c->nn + k]];
            change += (int64_t)(s == prop) - (int64_t)(s == cur);
        }
        double offset = d->offsets[prop] - d->offsets[cur];
        return -d->j * (double)change - offset;
    } else {
        int64_t ds = (int64_t)(decode((uint8_t)prop) - decode((uint8_t)cur));
        int64_t stride = 6 * (int64_t)(d->dim - 1);
        int64_t sum = 0;
        for (int g = 0; g < 2 * (d->dim - 1); ++g) {
            int32_t prod = 1;
            for (int k = 0; k < 3; ++k) prod *= decode(c->state[c->staples[var * stride + g * 3 + k]]);
            sum += prod;
        }
        return -d->j * (double)(ds * sum);
    }
}

/* f32 profile: ising/model.wgsl:32-38, potts/model.wgsl:92-105, gauge/model.wgsl:136-147 */
static float delta_f32(const ctx_t *c, int64_t var, int cur, int prop) {
    const orc_local_desc *d = c->d;
    float j32 = (float)d->j;
    if (d->model == ORC_MODEL_ISING) {
        int32_t nsum = 0;
        for (int k = 0; k < c->nn; ++k) nsum += decode(c->state[c->nbr[var * c->nn + k]]);
        float two_s = 2.0f * (float)decode((uint8_t)cur);
        float field = j32 * (float)nsum;
        field = field + (float)d->h;
        return two_s * field;
    } else if (d->model == ORC_MODEL_POTTS) {
        int32_t change = 0;
        for (int k = 0; k < c->nn; ++k) {
            int s = c->state[c->nbr[var * c->nn + k]];
            if (s == prop) change += 1;
            if (s == cur) change -= 1;
        }
        float offset = (float)d->offsets[prop] - (float)d->offsets[cur];
        float a = -j32 * (float)change;
        return a - offset;
    } else {
        int64_t stride = 6 * (int64_t)(d->dim - 1);
        int32_t sum = 0;
        for (int g = 0; g < 2 * (d->dim - 1); ++g) {
            int32_t prod = 1;
            for (int k = 0; k < 3; ++k) prod *= decode(c->state[c->staples[var * stride + g * 3 + k]]);
            sum += prod;
        }
        float a = 2.0f * j32;
        a = a * (float)decode((uint8_t)cur);
        return a * (float)sum;
    }
}

/* the canonical f32 exponential: one rounding of the f64 libm value */
static float exp_f32(float x) { return (float)exp((double)x); }

/* ------------------------------------------------------ local kernels --- */

/* metropolis.wgsl:36-55 / updater.rs:659-668 in K-mode: q = 2 flips with no
 * draw; q > 2 draws keyed on (seed ^ PROPOSAL_KEY, var, sweep) through an f32
 * multiply, clamps, and skips past the current state */
static int propose(const orc_local_desc *d, int64_t var, uint32_t sweep, int cur) {
    if (d->q == 2) return 1 - cur;
    uint32_t alternatives = (uint32_t)(d->q - 1);
    float u = orc_keyed_uniform(d->seed32 ^ PROPOSAL_KEY, (uint32_t)var, sweep);
    float scaled = u * (float)alternatives;
    uint32_t draw = (uint32_t)scaled;
    if (draw >= alternatives) draw = alternatives - 1;
    int prop = (int)draw;
    if (prop >= cur) prop += 1;
    return prop;
}

/* updater.rs:77-96 (f64) / metropolis.wgsl:57-62 (f32) */
static void metropolis(const ctx_t *c, int64_t var, uint32_t sweep) {
    const orc_local_desc *d = c->d;
    int cur = c->state[var];
    int prop = propose(d, var, sweep, cur);
    uint32_t m = orc_keyed_bits(d->seed32, (uint32_t)var, sweep);
    int accept;
    if (d->profile == ORC_PROFILE_F64) {
        double delta = delta_f64(c, var, cur, prop);
        accept = (delta <= 0.0) || (u64_of(m) < exp(-d->beta * delta));
    } else {
        float delta = delta_f32(c, var, cur, prop);
        float u = (float)m * (1.0f / 16777216.0f);
        float arg = -(float)d->beta * delta;
        accept = (delta <= 0.0f) || (u < exp_f32(arg));
    }
    if (accept) c->state[var] = (uint8_t)prop;
}

/* updater.rs:109-167 (f64) / heat_bath.wgsl:28-62 (f32) */
static void heat_bath(const ctx_t *c, int64_t var, uint32_t sweep) {
    const orc_local_desc *d = c->d;
    int q = d->q;
    int cur = c->state[var];
    uint32_t m = orc_keyed_bits(d->seed32, (uint32_t)var, sweep);
    int chosen = q - 1;
    if (d->profile == ORC_PROFILE_F64) {
        double deltas[ORC_MAX_Q], weights[ORC_MAX_Q];
        double floor_ = INFINITY;
        for (int s = 0; s < q; ++s) {
            deltas[s] = (s == cur) ? 0.0 : delta_f64(c, var, cur, s);
            floor_ = fmin(floor_, deltas[s]);
        }
        double total = 0.0;
        for (int s = 0; s < q; ++s) {
            weights[s] = exp(-d->beta * (deltas[s] - floor_));
            total += weights[s];
        }
        double target = u64_of(m) * total;
        double cumulative = 0.0;
        for (int s = 0; s < q - 1; ++s) {
            cumulative += weights[s];
            if (target < cumulative) { chosen = s; break; }
        }
    } else {
        float deltas[ORC_MAX_Q], weights[ORC_MAX_Q];
        float lowest = 0.0f;
        for (int s = 0; s < q; ++s) {
            float de = 0.0f;
            if (s != cur) de = delta_f32(c, var, cur, s);
            deltas[s] = de;
            lowest = fminf(lowest, de);
        }
        float total = 0.0f;
        float beta32 = (float)d->beta;
        for (int s = 0; s < q; ++s) {
            float diff = deltas[s] - lowest;
            weights[s] = exp_f32(-beta32 * diff);
            total = total + weights[s];
        }
        float u = (float)m * (1.0f / 16777216.0f);
        float threshold = u * total;
        float cumulative = 0.0f;
        for (int s = 0; s + 1 < q; ++s) {
            cumulative = cumulative + weights[s];
            if (threshold < cumulative) { chosen = s; break; }
        }
    }
    c->state[var] = (uint8_t)chosen;
}

/* updater.rs:291-347, Schedule::Checkerboard */
int orc_local_sweeps(const orc_local_desc *d, uint8_t *state, uint32_t sweep0, uint32_t n_sweeps) {
    int dim = d->dim;
    if (dim < 1 || dim > ORC_MAX_DIM || d->q < 2 || d->q > ORC_MAX_Q) return -1;
    if (d->model == ORC_MODEL_Z2 && dim < 2) return -1;
    int64_t n = orc_n_sites(dim, d->shape);
    uint8_t *parity = (uint8_t *)malloc((size_t)n);
    orc_parity_table(dim, d->shape, parity);
    ctx_t c;
    c.d = d;
    c.state = state;
    c.nn = 2 * dim;
    int64_t *nbr = NULL, *staples = NULL;
    if (d->model == ORC_MODEL_Z2) {
        staples = (int64_t *)malloc((size_t)(n * dim * 6 * (dim - 1)) * sizeof(int64_t));
        orc_staple_table(dim, d->shape, staples);
    } else {
        nbr = (int64_t *)malloc((size_t)(n * 2 * dim) * sizeof(int64_t));
        orc_neighbor_table(dim, d->shape, nbr);
    }
    c.nbr = nbr;
    c.staples = staples;
    for (uint32_t t = 0; t < n_sweeps; ++t) {
        uint32_t sweep = sweep0 + t;
        if (d->model != ORC_MODEL_Z2) {
            for (int color = 0; color < 2; ++color)
                for (int64_t site = 0; site < n; ++site)
                    if (parity[site] == color) {
                        if (d->kernel == ORC_KERNEL_METROPOLIS) metropolis(&c, site, sweep);
                        else heat_bath(&c, site, sweep);
                    }
        } else {
            for (int dir = 0; dir < dim; ++dir)
                for (int color = 0; color < 2; ++color)
                    for (int64_t site = 0; site < n; ++site)
                        if (parity[site] == color) {
                            int64_t link = site * dim + dir;
                            if (d->kernel == ORC_KERNEL_METROPOLIS) metropolis(&c, link, sweep);
                            else heat_bath(&c, link, sweep);
                        }
        }
    }
    free(parity);
    free(nbr);
    free(staples);
    return 0;
}

/* OpenMP's thread count for a parallel region (OMP_NUM_THREADS, else the CPUs). */
int32_t orc_max_threads(void) { return omp_get_max_threads(); }

/* The same sweeps with every colour pass spread over OpenMP threads.  In
 * the keyed stream a variable's decision reads only variables of the other
 * colour (or, for links, other directions) and its own draw, so the result is
 * the sequential one whatever the thread count.  Returns the wall time of
 * the sweeps in seconds (tables excluded), or -1 on a bad descriptor. */
double orc_local_sweeps_mt(const orc_local_desc *d, uint8_t *state, uint32_t sweep0, uint32_t n_sweeps,
                           int32_t threads) {
    int dim = d->dim;
    if (dim < 1 || dim > ORC_MAX_DIM || d->q < 2 || d->q > ORC_MAX_Q) return -1.0;
    if (d->model == ORC_MODEL_Z2 && dim < 2) return -1.0;
    if (threads < 1) threads = omp_get_max_threads();
    int64_t n = orc_n_sites(dim, d->shape);
    uint8_t *parity = (uint8_t *)malloc((size_t)n);
    orc_parity_table(dim, d->shape, parity);
    ctx_t c;
    c.d = d;
    c.state = state;
    c.nn = 2 * dim;
    int64_t *nbr = NULL, *staples = NULL;
    if (d->model == ORC_MODEL_Z2) {
        staples = (int64_t *)malloc((size_t)(n * dim * 6 * (dim - 1)) * sizeof(int64_t));
        orc_staple_table(dim, d->shape, staples);
    } else {
        nbr = (int64_t *)malloc((size_t)(n * 2 * dim) * sizeof(int64_t));
        orc_neighbor_table(dim, d->shape, nbr);
    }
    c.nbr = nbr;
    c.staples = staples;
    const double t0 = omp_get_wtime();
    for (uint32_t t = 0; t < n_sweeps; ++t) {
        uint32_t sweep = sweep0 + t;
        const int ndir = d->model == ORC_MODEL_Z2 ? dim : 1;
        for (int dir = 0; dir < ndir; ++dir)
            for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static) num_threads(threads)
                for (int64_t site = 0; site < n; ++site)
                    if (parity[site] == color) {
                        int64_t var = d->model == ORC_MODEL_Z2 ? site * dim + dir : site;
                        if (d->kernel == ORC_KERNEL_METROPOLIS) metropolis(&c, var, sweep);
                        else heat_bath(&c, var, sweep);
                    }
            }
    }
    const double secs = omp_get_wtime() - t0;
    free(parity);
    free(nbr);
    free(staples);
    return secs;
}

/* Single-variable decisions for an Ising site with a given neighbour sum,
 * built from the same per-site arithmetic as the sweeps above, with a fake
 * one-site context whose neighbours carry the requested sum. */
static int ising_decide(const orc_local_desc *d, int32_t cur, int64_t nsum, uint32_t m, int hb) {
    int nn = 2 * d->dim;
    if (nsum < -nn || nsum > nn || ((nsum + nn) & 1)) return -1;
    uint8_t st[1 + 2 * ORC_MAX_DIM];
    int64_t nbr[2 * ORC_MAX_DIM];
    int64_t ups = (nsum + nn) / 2;
    st[0] = (uint8_t)cur;
    for (int k = 0; k < nn; ++k) {
        st[1 + k] = (uint8_t)(k < ups ? 0 : 1);
        nbr[k] = 1 + k;
    }
    ctx_t c;
    c.d = d;
    c.state = st;
    c.nbr = nbr;
    c.staples = NULL;
    c.nn = nn;
    /* evaluate with the draw m by temporarily keying nothing: replicate the
     * per-site arithmetic inline */
    int prop = 1 - cur;
    if (!hb) {
        if (d->profile == ORC_PROFILE_F64) {
            double delta = delta_f64(&c, 0, cur, prop);
            return (delta <= 0.0) || (u64_of(m) < exp(-d->beta * delta));
        } else {
            float delta = delta_f32(&c, 0, cur, prop);
            float u = (float)m * (1.0f / 16777216.0f);
            float arg = -(float)d->beta * delta;
            return (delta <= 0.0f) || (u < exp_f32(arg));
        }
    }
    if (d->profile == ORC_PROFILE_F64) {
        double deltas[2];
        deltas[cur] = 0.0;
        deltas[prop] = delta_f64(&c, 0, cur, prop);
        double floor_ = fmin(fmin(INFINITY, deltas[0]), deltas[1]);
        double w0 = exp(-d->beta * (deltas[0] - floor_));
        double w1 = exp(-d->beta * (deltas[1] - floor_));
        double total = 0.0;
        total += w0;
        total += w1;
        double target = u64_of(m) * total;
        return (target < 0.0 + w0) ? 0 : 1;
    } else {
        float deltas[2];
        deltas[cur] = 0.0f;
        deltas[prop] = delta_f32(&c, 0, cur, prop);
        float lowest = fminf(fminf(0.0f, deltas[0]), deltas[1]);
        float beta32 = (float)d->beta;
        float w0 = exp_f32(-beta32 * (deltas[0] - lowest));
        float w1 = exp_f32(-beta32 * (deltas[1] - lowest));
        float total = 0.0f;
        total = total + w0;
        total = total + w1;
        float u = (float)m * (1.0f / 16777216.0f);
        float threshold = u * total;
        return (threshold < 0.0f + w0) ? 0 : 1;
    }
}

int orc_ising_metropolis_accept(const orc_local_desc *d, int32_t cur, int64_t nsum, uint32_t m) {
    return ising_decide(d, cur, nsum, m, 0);
}

int orc_ising_heat_bath_choice(const orc_local_desc *d, int32_t cur, int64_t nsum, uint32_t m) {
    return ising_decide(d, cur, nsum, m, 1);
}

/* ------------------------------------------------------------ cluster --- */

static int64_t uf_find(int64_t *parent, int64_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

/* cluster.rs:151-227 up to canonical relabelling: union by min index so the
 * root of every component is its minimum site (the fixed point of
 * cluster_prelude.wgsl:113-147, and the first-site order of cluster.rs:196-226) */
int orc_components(int32_t dim, const int64_t *shape, const uint8_t *bonds, int64_t *roots) {
    int64_t n = orc_n_sites(dim, shape);
    for (int64_t s = 0; s < n; ++s) roots[s] = s;
    for (int64_t site = 0; site < n; ++site)
        for (int mu = 0; mu < dim; ++mu)
            if (bonds[site * dim + mu]) {
                int64_t a = uf_find(roots, site);
                int64_t b = uf_find(roots, fwd_neighbor(dim, shape, site, mu));
                if (a < b) roots[b] = a;
                else if (b < a) roots[a] = b;
            }
    for (int64_t s = 0; s < n; ++s) roots[s] = uf_find(roots, s);
    return 0;
}

/* updater.rs:499-501 (p in f64); gpu_cluster.rs:289,317 (p as f32);
 * cluster_prelude.wgsl:61-80 (bond = site*D + axis, keyed with BOND_KEY) */
int orc_cluster_bonds(const orc_cluster_desc *d, const uint8_t *state, uint32_t sweep, uint8_t *bonds) {
    int dim = d->dim;
    int64_t n = orc_n_sites(dim, d->shape);
    double p = 1.0 - exp(-d->beta * d->gap);
    float p32 = (float)p;
    for (int64_t site = 0; site < n; ++site)
        for (int mu = 0; mu < dim; ++mu) {
            int64_t bond = site * dim + mu;
            int64_t partner = fwd_neighbor(dim, d->shape, site, mu);
            int agree = state[site] == state[partner];
            uint32_t m = orc_keyed_bits(d->seed32 ^ BOND_KEY, (uint32_t)bond, sweep);
            int drawn;
            if (d->profile == ORC_PROFILE_F64) drawn = u64_of(m) < p;
            else drawn = ((float)m * (1.0f / 16777216.0f)) < p32;
            bonds[bond] = (uint8_t)(agree && drawn);
        }
    return 0;
}

/* cluster_relabel.wgsl:39-90 */
int orc_cluster_sweeps(const orc_cluster_desc *d, uint8_t *state, uint32_t sweep0,
                       uint32_t n_sweeps, int64_t *roots_out) {
    int dim = d->dim;
    if (dim < 1 || dim > ORC_MAX_DIM || d->q < 2) return -1;
    int64_t n = orc_n_sites(dim, d->shape);
    uint8_t *bonds = (uint8_t *)malloc((size_t)(n * dim));
    int64_t *roots = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    uint32_t q = (uint32_t)d->q;
    for (uint32_t t = 0; t < n_sweeps; ++t) {
        uint32_t sweep = sweep0 + t;
        orc_cluster_bonds(d, state, sweep, bonds);
        orc_components(dim, d->shape, bonds, roots);
        int64_t seed_root = -1;
        if (d->seeded) {
            float u = orc_keyed_uniform(d->seed32 ^ SEED_SALT, 0u, sweep);
            uint32_t seed_site = (uint32_t)(u * (float)(uint32_t)n);
            if (seed_site >= (uint32_t)n) seed_site = (uint32_t)n - 1;
            seed_root = roots[seed_site];
        }
        for (int64_t site = 0; site < n; ++site) {
            int64_t root = roots[site];
            if (d->seeded && root != seed_root) continue;
            float u = orc_keyed_uniform(d->seed32 ^ ROOT_SALT, (uint32_t)root, sweep);
            if (d->forced) {
                uint32_t step = (uint32_t)(u * (float)(q - 1));
                if (step >= q - 1) step = q - 2;
                state[site] = (uint8_t)((state[site] + 1 + step) % q);
            } else {
                uint32_t draw = (uint32_t)(u * (float)q);
                if (draw >= q) draw = q - 1;
                state[site] = (uint8_t)draw;
            }
        }
    }
    if (roots_out) memcpy(roots_out, roots, (size_t)n * sizeof(int64_t));
    free(bonds);
    free(roots);
    return 0;
}

/* -------------------------------------------------------- observables --- */

/* ising/mod.rs:97-117 (forward bonds once) and 66-70 */
void orc_ising_sums(int32_t dim, const int64_t *shape, const uint8_t *state,
                    int64_t *bond_sum, int64_t *spin_sum) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t b = 0, s = 0;
    for (int64_t site = 0; site < n; ++site) {
        int si = decode(state[site]);
        s += si;
        for (int mu = 0; mu < dim; ++mu) b += si * decode(state[fwd_neighbor(dim, shape, site, mu)]);
    }
    *bond_sum = b;
    *spin_sum = s;
}

/* potts/mod.rs:237-271 (agree over forward bonds) and 124-139 (label counts) */
void orc_potts_sums(int32_t dim, const int64_t *shape, int32_t q, const uint8_t *state,
                    int64_t *agree, int64_t *counts) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t a = 0;
    for (int k = 0; k < q; ++k) counts[k] = 0;
    for (int64_t site = 0; site < n; ++site) {
        counts[state[site]] += 1;
        for (int mu = 0; mu < dim; ++mu) a += state[site] == state[fwd_neighbor(dim, shape, site, mu)];
    }
    *agree = a;
}

/* gauge/mod.rs:99-122 */
int64_t orc_plaquette_sum(int32_t dim, const int64_t *shape, const uint8_t *links) {
    int64_t n = orc_n_sites(dim, shape);
    int64_t n_plaq = n * (int64_t)dim * (dim - 1) / 2;
    int64_t sum = 0;
    for (int64_t p = 0; p < n_plaq; ++p) {
        int64_t l[4];
        orc_plaquette_links(dim, shape, p, l);
        sum += decode(links[l[0]]) * decode(links[l[1]]) * decode(links[l[2]]) * decode(links[l[3]]);
    }
    return sum;
}

/* --------------------------------------------------------- statistics --- */

static double mean_of(const double *x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    return s / (double)n;
}

/* statistics.rs:83-137 */
orc_estimate orc_reduce(const double *series, int64_t n) {
    orc_estimate e;
    if (n < 2) {
        e.mean = n == 1 ? series[0] : 0.0;
        e.stderr_ = 0.0;
        e.tau_int = 0.5;
        e.n_eff = (double)n;
        return e;
    }
    double nf = (double)n;
    double mean = mean_of(series, n);
    double c0 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double dx = series[i] - mean;
        c0 += dx * dx;
    }
    c0 /= nf;
    if (c0 == 0.0) {
        e.mean = mean;
        e.stderr_ = 0.0;
        e.tau_int = 0.5;
        e.n_eff = nf;
        return e;
    }
    double c = 5.0, tau = 0.5, tau_int = 0.5;
    for (int64_t t = 1; t < n; ++t) {
        double ct = 0.0;
        for (int64_t i = 0; i < n - t; ++i) ct += (series[i] - mean) * (series[i + t] - mean);
        ct /= nf;
        tau += ct / c0;
        tau_int = tau;
        if ((double)t >= c * tau) break;
    }
    if (tau_int < 0.5) tau_int = 0.5;
    e.mean = mean;
    e.tau_int = tau_int;
    e.n_eff = nf / (2.0 * tau_int);
    e.stderr_ = sqrt(2.0 * tau_int * c0 / nf);
    return e;
}

typedef double (*estimator_fn)(const double *means, const double *params);

/* statistics.rs:152-211 */
static orc_derived jackknife(const double **features, int nf_, int64_t n, int64_t block_len,
                             estimator_fn est, const double *params) {
    orc_derived r;
    double means[8];
    for (int j = 0; j < nf_; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += features[j][i];
        means[j] = s / (double)n;
    }
    r.value = est(means, params);
    if (block_len < 1) block_len = 1;
    int64_t nb = n / block_len;
    r.n_blocks = nb;
    if (nb < 2) {
        r.stderr_ = NAN;
        return r;
    }
    int64_t used = block_len * nb;
    double *partials = (double *)malloc((size_t)(nb * nf_) * sizeof(double));
    double total_used[8] = {0};
    for (int j = 0; j < nf_; ++j)
        for (int64_t b = 0; b < nb; ++b) {
            double s = 0.0;
            for (int64_t i = b * block_len; i < b * block_len + block_len; ++i) s += features[j][i];
            partials[b * nf_ + j] = s;
            total_used[j] += s;
        }
    double leave_n = (double)(used - block_len);
    double *thetas = (double *)malloc((size_t)nb * sizeof(double));
    for (int64_t b = 0; b < nb; ++b) {
        for (int j = 0; j < nf_; ++j) means[j] = (total_used[j] - partials[b * nf_ + j]) / leave_n;
        thetas[b] = est(means, params);
    }
    double theta_bar = mean_of(thetas, nb);
    double factor = ((double)nb - 1.0) / (double)nb;
    double spread = 0.0;
    for (int64_t b = 0; b < nb; ++b) spread += (thetas[b] - theta_bar) * (thetas[b] - theta_bar);
    r.stderr_ = sqrt(factor * spread);
    free(partials);
    free(thetas);
    return r;
}

/* statistics.rs:216-219 */
static int64_t block_len_for(const double **series, int ns, int64_t n) {
    double tau = 0.5;
    for (int k = 0; k < ns; ++k) tau = fmax(tau, orc_reduce(series[k], n).tau_int);
    int64_t b = (int64_t)ceil(2.0 * tau);
    return b < 1 ? 1 : b;
}

static double est_specific_heat(const double *m, const double *p) {
    return p[0] * p[0] * fmax(m[1] - m[0] * m[0], 0.0) / p[1];
}
static double est_susceptibility(const double *m, const double *p) {
    return p[0] * p[1] * fmax(m[1] - m[0] * m[0], 0.0);
}
static double est_binder(const double *m, const double *p) {
    (void)p;
    return 1.0 - m[1] / (3.0 * m[0] * m[0]);
}

/* statistics.rs:232-244 */
orc_derived orc_specific_heat(const double *e, int64_t n, double beta, double n_sites) {
    double mean_e = mean_of(e, n);
    double *f0 = (double *)malloc((size_t)n * sizeof(double));
    double *f1 = (double *)malloc((size_t)n * sizeof(double));
    double *e2 = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        f0[i] = e[i] - mean_e;
        f1[i] = f0[i] * f0[i];
        e2[i] = e[i] * e[i];
    }
    const double *raw[2] = {e, e2};
    int64_t bl = block_len_for(raw, 2, n);
    const double *feat[2] = {f0, f1};
    double params[2] = {beta, n_sites};
    orc_derived r = jackknife(feat, 2, n, bl, est_specific_heat, params);
    free(f0);
    free(f1);
    free(e2);
    return r;
}

/* statistics.rs:248-264 */
orc_derived orc_susceptibility(const double *mg, int64_t n, double beta, double n_sites) {
    double *a = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i) a[i] = fabs(mg[i] / n_sites);
    double mean_abs = mean_of(a, n);
    double *f0 = (double *)malloc((size_t)n * sizeof(double));
    double *f1 = (double *)malloc((size_t)n * sizeof(double));
    double *a2 = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        f0[i] = a[i] - mean_abs;
        f1[i] = f0[i] * f0[i];
        a2[i] = a[i] * a[i];
    }
    const double *raw[2] = {a, a2};
    int64_t bl = block_len_for(raw, 2, n);
    const double *feat[2] = {f0, f1};
    double params[2] = {beta, n_sites};
    orc_derived r = jackknife(feat, 2, n, bl, est_susceptibility, params);
    free(a);
    free(f0);
    free(f1);
    free(a2);
    return r;
}

/* statistics.rs:277-284 */
orc_derived orc_binder(const double *mg, int64_t n) {
    double *m2 = (double *)malloc((size_t)n * sizeof(double));
    double *m4 = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        m2[i] = mg[i] * mg[i];
        m4[i] = m2[i] * m2[i];
    }
    const double *raw[2] = {m2, m4};
    int64_t bl = block_len_for(raw, 2, n);
    orc_derived r = jackknife(raw, 2, n, bl, est_binder, NULL);
    free(m2);
    free(m4);
    return r;
}

/* ---------------------------------------------------- P-mode (PCG64) --- */

typedef unsigned __int128 u128;
static const u128 PCG_MUL = (((u128)0x2360ED051FC65DA4ull) << 64) | (u128)0x4385DF649FCCF645ull;

static u128 pcg_state(const orc_pcg64 *g) { return ((u128)g->state_hi << 64) | g->state_lo; }
static u128 pcg_inc(const orc_pcg64 *g) { return ((u128)g->inc_hi << 64) | g->inc_lo; }
static void pcg_set(orc_pcg64 *g, u128 s) {
    g->state_lo = (uint64_t)s;
    g->state_hi = (uint64_t)(s >> 64);
}

/* rand_core seed_from_u64: PCG32 expansion of the u64 into the 32-byte seed;
 * rand_pcg Lcg128Xsl64::from_seed / from_state_incr */
void orc_pcg64_seed_from_u64(orc_pcg64 *g, uint64_t state) {
    uint8_t seed[32];
    for (int chunk = 0; chunk < 8; ++chunk) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        uint64_t s = state;
        uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
        uint32_t rot = (uint32_t)(s >> 59);
        uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
        memcpy(seed + 4 * chunk, &x, 4);
    }
    uint64_t w[4];
    memcpy(w, seed, 32);
    u128 st = ((u128)w[1] << 64) | w[0];
    u128 inc = (((u128)w[3] << 64) | w[2]) | 1;
    g->inc_lo = (uint64_t)inc;
    g->inc_hi = (uint64_t)(inc >> 64);
    st = st + inc;
    st = st * PCG_MUL + inc;
    pcg_set(g, st);
}

uint64_t orc_pcg64_next_u64(orc_pcg64 *g) {
    u128 s = pcg_state(g) * PCG_MUL + pcg_inc(g);
    pcg_set(g, s);
    uint32_t rot = (uint32_t)(s >> 122);
    uint64_t xsl = (uint64_t)(s >> 64) ^ (uint64_t)s;
    return (xsl >> rot) | (xsl << ((64 - rot) & 63));
}

double orc_pcg64_next_f64(orc_pcg64 *g) {
    return (double)(orc_pcg64_next_u64(g) >> 11) * (1.0 / 9007199254740992.0);
}

/* LocalUpdate{kernel, Checkerboard} over an Ising site field, as written in
 * updater.rs:77-167 and 316-324 with RandRng (rng.rs:30-50): usize neighbour
 * table, u8 parity table, a full scan per colour, f64 exp per uphill site */
double orc_pmode_ising_sweeps(int32_t dim, const int64_t *shape, int32_t kernel,
                              double beta, double j, double h, uint64_t seed,
                              uint8_t *state, uint32_t n_sweeps) {
    orc_pcg64 g;
    orc_pcg64_seed_from_u64(&g, seed);
    return orc_pmode_ising_sweeps_from(dim, shape, kernel, beta, j, h, &g, state, n_sweeps);
}

/* The same sweeps continuing a given stream (a run's generator after its hot
 * start draws, ising/run_config.rs:237-251); *gen is advanced. */
double orc_pmode_ising_sweeps_from(int32_t dim, const int64_t *shape, int32_t kernel,
                                   double beta, double j, double h, orc_pcg64 *gen,
                                   uint8_t *state, uint32_t n_sweeps) {
    int64_t n = orc_n_sites(dim, shape);
    int nn = 2 * dim;
    int64_t *nbr = (int64_t *)malloc((size_t)(n * nn) * sizeof(int64_t));
    uint8_t *parity = (uint8_t *)malloc((size_t)n);
    orc_neighbor_table(dim, shape, nbr);
    orc_parity_table(dim, shape, parity);
    orc_pcg64 g = *gen;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint32_t t = 0; t < n_sweeps; ++t)
        for (int color = 0; color < 2; ++color)
            for (int64_t site = 0; site < n; ++site) {
                if (parity[site] != color) continue;
                int cur = state[site];
                int64_t nsum = 0;
                const int64_t *row = nbr + site * nn;
                for (int k = 0; k < nn; ++k) nsum += decode(state[row[k]]);
                int64_t ds = (int64_t)(decode((uint8_t)(1 - cur)) - decode((uint8_t)cur));
                double delta = -(double)ds * (j * (double)nsum + h);
                if (kernel == ORC_KERNEL_METROPOLIS) {
                    if (delta <= 0.0 || orc_pcg64_next_f64(&g) < exp(-beta * delta))
                        state[site] = (uint8_t)(1 - cur);
                } else {
                    double d0 = cur == 0 ? 0.0 : delta, d1 = cur == 1 ? 0.0 : delta;
                    double fl = fmin(fmin(INFINITY, d0), d1);
                    double w0 = exp(-beta * (d0 - fl)), w1 = exp(-beta * (d1 - fl));
                    double total = 0.0 + w0 + w1;
                    double target = orc_pcg64_next_f64(&g) * total;
                    state[site] = (uint8_t)((target < w0) ? 0 : 1);
                }
            }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *gen = g;
    free(nbr);
    free(parity);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------ ensemble --- */

typedef struct {
    int n_groups;
    int group_of[4096];
    int g_first[4096];   /* offset of group g in rep_at */
    int g_size[4096];
} ens_groups;

static int same_shape(const orc_ensemble_desc *d, int a, int b) {
    for (int k = 0; k < d->dim; ++k)
        if (d->shapes[a * d->dim + k] != d->shapes[b * d->dim + k]) return 0;
    return 1;
}

/* groups by first appearance of a shape; members ordered by ascending beta
 * (stable), giving the pair order every exchange round walks */
static void ens_build(const orc_ensemble_desc *d, ens_groups *G, uint32_t *rep_at) {
    int R = d->n_replicas;
    G->n_groups = 0;
    int placed = 0;
    for (int r = 0; r < R; ++r) {
        int g = -1;
        for (int s = 0; s < r; ++s)
            if (same_shape(d, r, s)) { g = G->group_of[s]; break; }
        if (g < 0) { g = G->n_groups++; G->g_size[g] = 0; }
        G->group_of[r] = g;
        G->g_size[g] += 1;
    }
    for (int g = 0; g < G->n_groups; ++g) {
        G->g_first[g] = placed;
        int k = 0;
        for (int r = 0; r < R; ++r)
            if (G->group_of[r] == g) rep_at[placed + k++] = (uint32_t)r;
        /* stable insertion sort by beta */
        for (int i = 1; i < k; ++i) {
            uint32_t v = rep_at[placed + i];
            int jx = i - 1;
            while (jx >= 0 && d->betas[rep_at[placed + jx]] > d->betas[v]) {
                rep_at[placed + jx + 1] = rep_at[placed + jx];
                --jx;
            }
            rep_at[placed + jx + 1] = v;
        }
        placed += k;
    }
}

int orc_ensemble_init(const orc_ensemble_desc *d, uint32_t *slot_of, uint32_t *rep_at) {
    if (d->n_replicas < 1 || d->n_replicas > 4096) return -1;
    ens_groups G;
    ens_build(d, &G, rep_at);
    for (int r = 0; r < d->n_replicas; ++r) slot_of[r] = (uint32_t)r;
    return 0;
}

int orc_ensemble_sweeps(const orc_ensemble_desc *d, uint8_t *state, uint32_t *slot_of, uint32_t *rep_at,
                        uint64_t *sweeps_done, uint64_t *rounds_done, uint64_t n_sweeps, int64_t *rec_sums,
                        uint32_t *rec_slots, uint64_t rec_cap, uint64_t *n_rec, uint64_t *accepted) {
    int R = d->n_replicas;
    if (R < 1 || R > 4096) return -1;
    ens_groups G;
    uint32_t order0[4096];
    ens_build(d, &G, order0); /* group layout only; rep_at carries the current order */
    size_t *off = (size_t *)malloc(sizeof(size_t) * (size_t)R);
    size_t o = 0;
    for (int r = 0; r < R; ++r) {
        off[r] = o;
        o += (size_t)orc_n_sites(d->dim, d->shapes + r * d->dim) * (size_t)d->dim;
    }
    int64_t *sums = (int64_t *)malloc(sizeof(int64_t) * (size_t)R);
    for (uint64_t t = 0; t < n_sweeps; ++t) {
        uint32_t sweep = (uint32_t)*sweeps_done;
        for (int r = 0; r < R; ++r) {
            orc_local_desc ld;
            memset(&ld, 0, sizeof ld);
            ld.dim = d->dim;
            for (int k = 0; k < d->dim; ++k) ld.shape[k] = d->shapes[r * d->dim + k];
            ld.model = ORC_MODEL_Z2;
            ld.q = 2;
            ld.kernel = d->kernel;
            ld.profile = d->profile;
            ld.beta = d->betas[slot_of[r]];
            ld.j = d->j;
            ld.seed32 = d->seed32s[r];
            orc_local_sweeps(&ld, state + off[r], sweep, 1);
        }
        *sweeps_done += 1;
        if (!d->measure_every || *sweeps_done % d->measure_every != 0) continue;
        for (int r = 0; r < R; ++r) sums[r] = orc_plaquette_sum(d->dim, d->shapes + r * d->dim, state + off[r]);
        if (rec_sums && n_rec && *n_rec < rec_cap) {
            for (int r = 0; r < R; ++r) {
                rec_sums[*n_rec * R + r] = sums[r];
                if (rec_slots) rec_slots[*n_rec * R + r] = slot_of[r];
            }
            *n_rec += 1;
        }
        if (!d->exchange) continue;
        uint32_t round = (uint32_t)*rounds_done;
        uint32_t p = 0;
        for (int g = 0; g < G.n_groups; ++g) {
            int first = G.g_first[g];
            for (int k = 0; k + 1 < G.g_size[g]; ++k, ++p) {
                if ((uint32_t)(k % 2) != round % 2) continue;
                uint32_t a = rep_at[first + k], b = rep_at[first + k + 1];
                /* betas of the k-th and (k+1)-th smallest of the group */
                double bk = d->betas[order0[first + k]], bk1 = d->betas[order0[first + k + 1]];
                double s = (bk1 - bk) * d->j;
                int64_t x = sums[a] - sums[b];
                double delta = s * (double)x;
                int swap = delta >= 0.0;
                if (!swap) swap = u64_of(orc_keyed_bits(d->ex_seed32, p, round)) < exp(delta);
                if (swap) {
                    rep_at[first + k] = b;
                    rep_at[first + k + 1] = a;
                    uint32_t sa = slot_of[a];
                    slot_of[a] = slot_of[b];
                    slot_of[b] = sa;
                    if (accepted) accepted[p] += 1;
                }
            }
        }
        *rounds_done += 1;
    }
    free(off);
    free(sums);
    return 0;
}
