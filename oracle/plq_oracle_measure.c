/*
 * plq_oracle_measure.c -- CPU oracle of the per-configuration measurements
 * (test infrastructure only, like the rest of oracle/): the two-point
 * correlator's raw pair sums, the Polyakov loop and the Wilson rectangle
 * table.  Each function is a sequential restatement of the reference's loop,
 * site by site and link by link, with none of the device's bit-plane or
 * line-product algebra, so the two are independent computations.
 *
 * Lattice conventions (lattice.rs:55-100, 353-361, 467-486): site index is
 * mixed radix with axis 0 fastest, link index = site * D + dir, state 0 is
 * the +1 spin / link (models/mod.rs:30-32).
 */
#include <stdint.h>
#include <stdlib.h>

#include "plq_oracle.h"

static int32_t sign_of(uint8_t s) { return 1 - 2 * (int32_t)s; }

static void strides_of(int32_t dim, const int64_t *shape, int64_t *stride) {
    stride[0] = 1;
    for (int mu = 1; mu < dim; ++mu) stride[mu] = stride[mu - 1] * shape[mu - 1];
}

/* lattice.rs:353-361: strip the `dir` coordinate and put back (coord + delta) mod L. */
static int64_t site_shift(const int64_t *shape, const int64_t *stride, int64_t site, int dir, int64_t delta) {
    int64_t s = stride[dir], l = shape[dir];
    int64_t coord = (site / s) % l;
    int64_t base = site - coord * s;
    return base + ((coord + delta % l) % l) * s;
}

/* models/mod.rs:44-60: for every site, axis and r = 0..=L_mu/2, add pair(s_i, s_{i + r e_mu}).
 * pair_kind 0: decode(a) * decode(b) (ising/mod.rs:89-91); 1: a == b (potts/mod.rs:230). */
void orc_axis_pair_sums(int32_t dim, const int64_t *shape, const uint8_t *state, int32_t pair_kind, int64_t *out) {
    int64_t stride[4];
    strides_of(dim, shape, stride);
    int64_t off[4];
    int64_t total = 0;
    for (int mu = 0; mu < dim; ++mu) {
        off[mu] = total;
        total += shape[mu] / 2 + 1;
    }
    for (int64_t k = 0; k < total; ++k) out[k] = 0;
    const int64_t n = orc_n_sites(dim, shape);
    for (int64_t site = 0; site < n; ++site) {
        const uint8_t si = state[site];
        for (int mu = 0; mu < dim; ++mu) {
            for (int64_t r = 0; r <= shape[mu] / 2; ++r) {
                const uint8_t sj = state[site_shift(shape, stride, site, mu, r)];
                out[off[mu] + r] += pair_kind == 0 ? (int64_t)(sign_of(si) * sign_of(sj)) : (int64_t)(si == sj);
            }
        }
    }
}

/* gauge/mod.rs:178-201: the product of the L_dir links along `dir` from every
 * site with coordinate 0 on that axis (lattice.rs:467-486 for a run of Plus
 * steps: take the link at the current site, then move). */
int64_t orc_polyakov_sum(int32_t dim, const int64_t *shape, const uint8_t *links, int32_t dir, int64_t *lines) {
    int64_t stride[4];
    strides_of(dim, shape, stride);
    const int64_t n = orc_n_sites(dim, shape);
    int64_t total = 0, count = 0;
    for (int64_t site = 0; site < n; ++site) {
        if ((site / stride[dir]) % shape[dir] != 0) continue;
        int32_t prod = 1;
        int64_t cur = site;
        for (int64_t k = 0; k < shape[dir]; ++k) {
            prod *= sign_of(links[cur * dim + dir]);
            cur = site_shift(shape, stride, cur, dir, 1);
        }
        total += prod;
        count += 1;
    }
    if (lines) *lines = count;
    return total;
}

/* Loop::rectangle (lattice.rs:532-555) walked by loop_links (lattice.rs:467-486):
 * r Plus steps along mu, t Plus along nu, r Minus along mu, t Minus along nu.
 * A Plus step takes the link at the current site and moves; a Minus step moves
 * back first and takes the link at the new site. */
static int32_t rectangle_product(int32_t dim, const int64_t *shape, const int64_t *stride, const uint8_t *links,
                                 int64_t base, int mu, int64_t r, int nu, int64_t t) {
    int32_t prod = 1;
    int64_t cur = base;
    const int dirs[4] = {mu, nu, mu, nu};
    const int64_t lens[4] = {r, t, r, t};
    for (int leg = 0; leg < 4; ++leg) {
        const int d = dirs[leg];
        for (int64_t k = 0; k < lens[leg]; ++k) {
            if (leg < 2) {
                prod *= sign_of(links[cur * dim + d]);
                cur = site_shift(shape, stride, cur, d, 1);
            } else {
                cur = site_shift(shape, stride, cur, d, shape[d] - 1);
                prod *= sign_of(links[cur * dim + d]);
            }
        }
    }
    return prod;
}

/* gauge/mod.rs:224-268: max = min(max_side, shortest / 2); totals[r * (max + 1) + t]
 * is the sum of the r-by-t loop products over every ordered direction pair
 * (mu != nu) and every base site, for r, t >= 1 (entries with r or t = 0 are
 * left 0: the table's 1.0 anchor).  *placements = D (D - 1) N.  Returns max. */
int32_t orc_wilson_sums(int32_t dim, const int64_t *shape, const uint8_t *links, int32_t max_side, int64_t *totals,
                        int64_t *placements) {
    int64_t stride[4];
    strides_of(dim, shape, stride);
    int64_t shortest = shape[0];
    for (int mu = 1; mu < dim; ++mu)
        if (shape[mu] < shortest) shortest = shape[mu];
    int32_t max = (int64_t)max_side < shortest / 2 ? max_side : (int32_t)(shortest / 2);
    const int64_t n = orc_n_sites(dim, shape);
    for (int64_t k = 0; k < (int64_t)(max + 1) * (max + 1); ++k) totals[k] = 0;
    for (int32_t r = 1; r <= max; ++r)
        for (int32_t t = 1; t <= max; ++t) {
            int64_t total = 0;
            for (int mu = 0; mu < dim; ++mu)
                for (int nu = 0; nu < dim; ++nu) {
                    if (mu == nu) continue;
                    for (int64_t site = 0; site < n; ++site)
                        total += rectangle_product(dim, shape, stride, links, site, mu, r, nu, t);
                }
            totals[r * (max + 1) + t] = total;
        }
    if (placements) *placements = (int64_t)dim * (dim - 1) * n;
    return max;
}
