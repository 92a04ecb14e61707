// plq_thresholds.cpp -- host-side acceptance tables for the local kernels.
//
// Every accept / choose decision of the reference kernels compares the keyed
// uniform u = m * 2^-24 (m a 24-bit integer, rng.wgsl:47-52) against a value
// that depends only on the variable's class: its current state and the
// multiset of states it is priced against.  Because the arithmetic producing
// that value is a fixed sequence of IEEE operations, and because u*x is
// monotone in m, each decision is "m < T[class]" for an integer T that the
// host finds by evaluating the reference's own predicate on m (binary search
// over [0, 2^24]).  The device then does no floating point at all, and the
// answer cannot depend on the device's exp, FMA contraction or rounding mode.
//
// Two arithmetic profiles:
//   f64 -- the CPU path: updater.rs:77-96 (step), 109-167 (heat_bath_step),
//          with ising/mod.rs:119-143, potts/mod.rs:273-309,
//          gauge/mod.rs:287-328 for the deltas; glibc exp (what Rust's
//          f64::exp calls on Linux).
//   f32 -- the WGSL kernels: metropolis.wgsl:57-62, heat_bath.wgsl:28-62 with
//          the model.wgsl deltas, exp_f32(x) = (float)exp((double)x).
// This file must be compiled with -ffp-contract=off (no FMA contraction).
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "plq_tables.hpp"

namespace plq {
namespace {

constexpr uint32_t ONE = 1u << 24;

double u64_of(uint32_t m) { return (double)m * (1.0 / 16777216.0); }
float u32_of(uint32_t m) { return (float)m * (1.0f / 16777216.0f); }
float exp_f32(float x) { return (float)std::exp((double)x); }

// Number of m in [0, 2^24) for which pred(m) holds, pred monotone (true then
// false).
template <class P>
uint32_t count_prefix(P pred) {
    uint32_t lo = 0, hi = ONE;  // invariant: pred true below lo, false at/after hi
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (pred(mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int decode(int s) { return 1 - 2 * s; }

// Metropolis on a two-state variable whose flip costs delta: accept iff
// delta <= 0 || u < exp(-beta delta).
uint32_t metropolis_f64(double beta, double delta) {
    if (delta <= 0.0) return ONE;
    double x = std::exp(-beta * delta);
    return count_prefix([&](uint32_t m) { return u64_of(m) < x; });
}

uint32_t metropolis_f32(float beta, float delta) {
    if (delta <= 0.0f) return ONE;
    float x = exp_f32(-beta * delta);
    return count_prefix([&](uint32_t m) { return u32_of(m) < x; });
}

// Heat bath over q states with deltas[cur] = 0 (updater.rs:127-163):
// thresholds T[0..q-2], chosen = first s with m < T[s], else q-1.
void heat_bath_f64(double beta, const double *deltas, int q, uint32_t *out) {
    double floor_ = INFINITY;
    for (int s = 0; s < q; ++s) floor_ = std::fmin(floor_, deltas[s]);
    double weights[32];
    double total = 0.0;
    for (int s = 0; s < q; ++s) {
        weights[s] = std::exp(-beta * (deltas[s] - floor_));
        total += weights[s];
    }
    double cumulative = 0.0;
    for (int s = 0; s < q - 1; ++s) {
        cumulative += weights[s];
        double cum = cumulative;
        out[s] = count_prefix([&](uint32_t m) { return u64_of(m) * total < cum; });
    }
}

void heat_bath_f32(float beta, const float *deltas, int q, uint32_t *out) {
    float lowest = 0.0f;
    for (int s = 0; s < q; ++s) lowest = std::fmin(lowest, deltas[s]);
    float weights[32];
    float total = 0.0f;
    for (int s = 0; s < q; ++s) {
        float diff = deltas[s] - lowest;
        weights[s] = exp_f32(-beta * diff);
        total = total + weights[s];
    }
    float cumulative = 0.0f;
    for (int s = 0; s + 1 < q; ++s) {
        cumulative = cumulative + weights[s];
        float cum = cumulative;
        out[s] = count_prefix([&](uint32_t m) { return u32_of(m) * total < cum; });
    }
}

// Ising delta for the flip of state `cur` with neighbour sum nsum.
double ising_delta_f64(const TableParams &p, int cur, int nsum) {
    int64_t ds = (int64_t)(decode(1 - cur) - decode(cur));
    return -(double)ds * (p.j * (double)nsum + p.h);
}
float ising_delta_f32(const TableParams &p, int cur, int nsum) {
    float two_s = 2.0f * (float)decode(cur);
    float field = (float)p.j * (float)nsum;
    field = field + (float)p.h;
    return two_s * field;
}
// Z2 delta for the flip of link state `cur` with staple sum S.
double z2_delta_f64(const TableParams &p, int cur, int S) {
    int64_t ds = (int64_t)(decode(1 - cur) - decode(cur));
    return -p.j * (double)(ds * (int64_t)S);
}
float z2_delta_f32(const TableParams &p, int cur, int S) {
    float a = 2.0f * (float)p.j;
    a = a * (float)decode(cur);
    return a * (float)S;
}
// Potts delta for cur -> prop with change = n_prop - n_cur.
double potts_delta_f64(const TableParams &p, int cur, int prop, int change) {
    if (prop == cur) return 0.0;
    double offset = p.offsets[prop] - p.offsets[cur];
    return -p.j * (double)change - offset;
}
float potts_delta_f32(const TableParams &p, int cur, int prop, int change) {
    if (prop == cur) return 0.0f;
    float offset = (float)p.offsets[prop] - (float)p.offsets[cur];
    float a = -(float)p.j * (float)change;
    return a - offset;
}

// Two-state table over classes (cur, k), k in [0, K]: value(cur, k) is the
// flip's delta.  Metropolis: T = accept threshold.  Heat bath: T = threshold
// below which the new state is 0.
template <class F64, class F32>
void two_state_table(const TableParams &p, int K, F64 d64, F32 d32, std::vector<uint32_t> &out) {
    out.assign(2 * (K + 1), 0);
    for (int cur = 0; cur < 2; ++cur)
        for (int k = 0; k <= K; ++k) {
            uint32_t t;
            if (p.kernel == 0) {
                t = p.profile == 0 ? metropolis_f64(p.beta, d64(cur, k)) : metropolis_f32((float)p.beta, d32(cur, k));
            } else if (p.profile == 0) {
                double deltas[2];
                deltas[cur] = 0.0;
                deltas[1 - cur] = d64(cur, k);
                heat_bath_f64(p.beta, deltas, 2, &t);
            } else {
                float deltas[2];
                deltas[cur] = 0.0f;
                deltas[1 - cur] = d32(cur, k);
                heat_bath_f32((float)p.beta, deltas, 2, &t);
            }
            out[cur * (K + 1) + k] = t;
        }
}

}  // namespace

int64_t potts_hb_classes(int D, int q) {
    int64_t b = 2 * D + 1, n = 1;
    for (int a = 0; a < q - 1; ++a) n *= b;
    return n * q;
}

bool build_table(const TableParams &p, std::vector<uint32_t> &out, std::string &err) {
    if (!std::isfinite(p.beta) || !std::isfinite(p.j) || !std::isfinite(p.h)) {
        err = "beta, j and h must be finite";
        return false;
    }
    const int D = p.D;
    if (p.model == 0) {  // Ising: class (cur, d), d = neighbours in state 1
        const int K = 2 * D;
        two_state_table(
            p, K, [&](int cur, int d) { return ising_delta_f64(p, cur, K - 2 * d); },
            [&](int cur, int d) { return ising_delta_f32(p, cur, K - 2 * d); }, out);
        return true;
    }
    if (p.model == 2) {  // Z2: class (cur, k), k = negative staple products
        const int K = 2 * (D - 1);
        two_state_table(
            p, K, [&](int cur, int k) { return z2_delta_f64(p, cur, K - 2 * k); },
            [&](int cur, int k) { return z2_delta_f32(p, cur, K - 2 * k); }, out);
        return true;
    }
    // Potts
    const int q = p.q;
    const int nn = 2 * D;
    if (p.kernel == 0) {
        // class (cur, prop, change + nn)
        const int W = 2 * nn + 1;
        out.assign((size_t)q * q * W, 0);
        for (int cur = 0; cur < q; ++cur)
            for (int prop = 0; prop < q; ++prop)
                for (int c = -nn; c <= nn; ++c) {
                    uint32_t t = 0;
                    if (prop != cur)
                        t = p.profile == 0 ? metropolis_f64(p.beta, potts_delta_f64(p, cur, prop, c))
                                           : metropolis_f32((float)p.beta, potts_delta_f32(p, cur, prop, c));
                    out[((size_t)cur * q + prop) * W + (c + nn)] = t;
                }
        return true;
    }
    // heat bath: class (cur, n_0..n_{q-2}) in base B = nn + 1, (q-1) thresholds each
    int64_t classes = potts_hb_classes(D, q);
    if (classes * (q - 1) > (int64_t)(1 << 22)) {
        err = "the Potts heat-bath table for this q and dimension exceeds 4M entries";
        return false;
    }
    const int B = nn + 1;
    int64_t per_cur = classes / q;
    out.assign((size_t)(classes * (q - 1)), 0);
    std::vector<int> n(q);
    for (int cur = 0; cur < q; ++cur)
        for (int64_t enc = 0; enc < per_cur; ++enc) {
            int64_t rest = enc;
            int sum = 0;
            for (int a = 0; a < q - 1; ++a) {
                n[a] = (int)(rest % B);
                rest /= B;
                sum += n[a];
            }
            if (sum > nn) continue;  // unreachable class
            n[q - 1] = nn - sum;
            uint32_t *dst = &out[(size_t)((cur * per_cur + enc) * (q - 1))];
            if (p.profile == 0) {
                double deltas[32];
                for (int s = 0; s < q; ++s) deltas[s] = potts_delta_f64(p, cur, s, n[s] - n[cur]);
                heat_bath_f64(p.beta, deltas, q, dst);
            } else {
                float deltas[32];
                for (int s = 0; s < q; ++s) deltas[s] = potts_delta_f32(p, cur, s, n[s] - n[cur]);
                heat_bath_f32((float)p.beta, deltas, q, dst);
            }
        }
    return true;
}

// Bond-open threshold: open iff agree && u < p, p = 1 - exp(-beta * gap)
// (updater.rs:499-501) in f64, or cast to f32 (gpu_cluster.rs:289, 317).
uint32_t bond_threshold(double beta, double gap, int profile) {
    double p = 1.0 - std::exp(-beta * gap);
    if (profile == 0) return count_prefix([&](uint32_t m) { return u64_of(m) < p; });
    float p32 = (float)p;
    return count_prefix([&](uint32_t m) { return u32_of(m) < p32; });
}

}  // namespace plq
