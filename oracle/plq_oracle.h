/*
 * plq_oracle.h -- CPU oracle for the plaquette data-parallel update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (plaquette_amd, libplaquette_hip.so) never links, loads or calls
 * anything here, and must fail loudly when its HIP extension is missing.
 *
 * It restates, sequentially and with the reference's own data structures
 * (neighbour table, parity table, staple table built in plaquette order), the
 * semantics of swvasquez/plaquette's CPU checkerboard path:
 *
 *   - keyed RNG            src/wgsl/rng.wgsl:34-52, src/device.rs:761-763
 *   - geometry             src/lattice.rs:55-163, 259-313, 420-441
 *   - Metropolis step      src/updater.rs:77-96, 659-668; src/wgsl/metropolis.wgsl:32-63
 *   - heat bath step       src/updater.rs:109-167; src/wgsl/heat_bath.wgsl:28-62
 *   - checkerboard order   src/updater.rs:291-347 (site: colour 0 then 1;
 *                          link: for dir, colour 0 then 1, link = site*D+dir)
 *   - energy deltas        src/models/ising/mod.rs:119-143,
 *                          src/models/potts/mod.rs:273-309,
 *                          src/models/gauge/mod.rs:287-328  (f64 profile);
 *                          src/models/{ising,potts,gauge}/model.wgsl (f32 profile)
 *   - cluster sweep        src/wgsl/cluster_prelude.wgsl:61-147,
 *                          src/wgsl/cluster_relabel.wgsl:39-90,
 *                          src/updater.rs:447-613, src/cluster.rs:44-227
 *   - observables          src/models/ising/mod.rs:97-186,
 *                          src/models/mod.rs:44-60 (correlator pair sums),
 *                          src/models/gauge/mod.rs:178-268 (Polyakov, Wilson),
 *                          src/models/potts/mod.rs:142-181, 237-365,
 *                          src/models/gauge/mod.rs:99-122
 *   - statistics (FP64)    src/statistics.rs:83-317
 *   - PCG64 CPU stream     src/rng.rs:30-50 (rand_pcg 0.10.2 Lcg128Xsl64,
 *                          rand_core 0.10.1 seed_from_u64) -- parity unpinned:
 *                          no reference test fixes its output values.
 *
 * Exactness contract ("K-mode"): the reference checkerboard order, driven by
 * the reference's keyed per-variable stream m = hash(seed32, var, sweep) >> 8,
 * with the acceptance arithmetic of one of two profiles:
 *   PLQ_PROFILE_F64: the CPU path's f64 arithmetic (u = m * 2^-24 as f64,
 *                    glibc exp), i.e. updater.rs:88-91 and 141-163;
 *   PLQ_PROFILE_F32: the WGSL kernels' f32 arithmetic with the canonical
 *                    exp_f32(x) = (float)exp((double)x).
 */
#ifndef PLQ_ORACLE_H
#define PLQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MODEL_ISING = 0, ORC_MODEL_POTTS = 1, ORC_MODEL_Z2 = 2 };
enum { ORC_KERNEL_METROPOLIS = 0, ORC_KERNEL_HEAT_BATH = 1 };
enum { ORC_PROFILE_F64 = 0, ORC_PROFILE_F32 = 1 };
enum { ORC_MAX_DIM = 8, ORC_MAX_Q = 16 };

typedef struct {
    int32_t dim;
    int64_t shape[ORC_MAX_DIM];
    int32_t model;      /* ORC_MODEL_* ; Z2 lives on links, the others on sites */
    int32_t q;          /* 2 for Ising / Z2 */
    int32_t kernel;     /* ORC_KERNEL_* */
    int32_t profile;    /* ORC_PROFILE_* */
    double beta;
    double j;
    double h;                 /* Ising field */
    double offsets[ORC_MAX_Q];/* Potts per-label offsets h[a] */
    uint32_t seed32;          /* fold_seed(seed) */
} orc_local_desc;

typedef struct {
    int32_t dim;
    int64_t shape[ORC_MAX_DIM];
    int32_t q;
    int32_t seeded;     /* 1 = Wolff extent (relabel only the seed's cluster) */
    int32_t forced;     /* 1 = forced change, 0 = redraw */
    int32_t profile;    /* bond test in f64 (CPU) or f32 (WGSL) */
    double beta;
    double gap;         /* bond energy gap: J (Potts), 2J (Ising) */
    uint32_t seed32;
} orc_cluster_desc;

/* keyed RNG (rng.wgsl:34-52) */
uint32_t orc_lowbias32(uint32_t x);
uint32_t orc_keyed_bits(uint32_t seed, uint32_t var, uint32_t sweep); /* 24-bit m */
float orc_keyed_uniform(uint32_t seed, uint32_t var, uint32_t sweep);
uint32_t orc_fold_seed(uint64_t seed);

/* geometry (lattice.rs) */
int64_t orc_n_sites(int32_t dim, const int64_t *shape);
/* 2*dim neighbour entries per site, order +0,-0,+1,-1,... */
void orc_neighbor_table(int32_t dim, const int64_t *shape, int64_t *out);
void orc_parity_table(int32_t dim, const int64_t *shape, uint8_t *out);
/* number of plaquettes = n_sites * dim*(dim-1)/2; 4 links each */
void orc_plaquette_links(int32_t dim, const int64_t *shape, int64_t plaquette, int64_t out[4]);
/* staples: 6*(dim-1) link indices per link, groups in increasing plaquette order */
void orc_staple_table(int32_t dim, const int64_t *shape, int64_t *out);

/* K-mode checkerboard sweeps; state is lattice-ordered (site, or link=site*D+dir) */
int orc_local_sweeps(const orc_local_desc *d, uint8_t *state, uint32_t sweep0, uint32_t n_sweeps);
/* orc_local_sweeps with each colour pass over `threads` OpenMP threads (0: all);
 * bit-identical to it.  Returns the seconds spent sweeping. */
int32_t orc_max_threads(void);
double orc_local_sweeps_mt(const orc_local_desc *d, uint8_t *state, uint32_t sweep0, uint32_t n_sweeps,
                           int32_t threads);
/* The acceptance table entries the device uses are checked against the literal
 * per-site arithmetic through these: the Metropolis / heat bath decision of a
 * single variable with a given 24-bit draw. */
int orc_ising_metropolis_accept(const orc_local_desc *d, int32_t cur, int64_t nsum, uint32_t m);
int orc_ising_heat_bath_choice(const orc_local_desc *d, int32_t cur, int64_t nsum, uint32_t m);

/* K-mode cluster sweeps; roots_out (n_sites, may be NULL) receives the
 * min-site-index root of each site's cluster in the LAST sweep, sizes of the
 * seeded cluster etc. are derivable from it. */
int orc_cluster_sweeps(const orc_cluster_desc *d, uint8_t *state, uint32_t sweep0,
                       uint32_t n_sweeps, int64_t *roots_out);
/* Only the bond draw of one sweep (bond = site*D+axis), 1 open / 0 closed. */
int orc_cluster_bonds(const orc_cluster_desc *d, const uint8_t *state, uint32_t sweep, uint8_t *bonds);
/* Connected components of the open bonds, canonical: root = min site index. */
int orc_components(int32_t dim, const int64_t *shape, const uint8_t *bonds, int64_t *roots);

/* Batched Z2 replicas with replica exchange (no reference counterpart; the
 * semantics are those documented for plq_ensemble_* in plaquette_hip.h).
 * state concatenates the replicas' lattice-order links; slot_of / rep_at /
 * counters carry the exchange state between calls (rep_at in group order:
 * groups by first appearance of a shape, members by ascending beta). */
typedef struct {
    int32_t dim, n_replicas, kernel, profile;
    const int64_t *shapes;
    const double *betas;
    const uint32_t *seed32s;
    double j;
    uint64_t measure_every;
    int32_t exchange;
    uint32_t ex_seed32;   /* fold_seed(exchange_seed) ^ 0x68e31da4 */
} orc_ensemble_desc;
int orc_ensemble_init(const orc_ensemble_desc *d, uint32_t *slot_of, uint32_t *rep_at);
int orc_ensemble_sweeps(const orc_ensemble_desc *d, uint8_t *state, uint32_t *slot_of, uint32_t *rep_at,
                        uint64_t *sweeps_done, uint64_t *rounds_done, uint64_t n_sweeps, int64_t *rec_sums,
                        uint32_t *rec_slots, uint64_t rec_cap, uint64_t *n_rec, uint64_t *accepted);

/* observables (integer sums) */
void orc_ising_sums(int32_t dim, const int64_t *shape, const uint8_t *state,
                    int64_t *bond_sum, int64_t *spin_sum);
void orc_potts_sums(int32_t dim, const int64_t *shape, int32_t q, const uint8_t *state,
                    int64_t *agree, int64_t *counts);
int64_t orc_plaquette_sum(int32_t dim, const int64_t *shape, const uint8_t *links);

/* statistics (statistics.rs:83-317), FP64 */
typedef struct { double mean, stderr_, tau_int, n_eff; } orc_estimate;
typedef struct { double value, stderr_; int64_t n_blocks; } orc_derived;
orc_estimate orc_reduce(const double *series, int64_t n);
orc_derived orc_specific_heat(const double *e, int64_t n, double beta, double n_sites);
orc_derived orc_susceptibility(const double *m, int64_t n, double beta, double n_sites);
orc_derived orc_binder(const double *m, int64_t n);

/* P-mode: the reference CPU path as written (PCG64 stream, f64, table
 * geometry), used as the timed CPU baseline.  Parity unpinned. */
typedef struct { uint64_t state_lo, state_hi, inc_lo, inc_hi; } orc_pcg64;
void orc_pcg64_seed_from_u64(orc_pcg64 *g, uint64_t seed);
uint64_t orc_pcg64_next_u64(orc_pcg64 *g);
double orc_pcg64_next_f64(orc_pcg64 *g);
/* Runs n_sweeps of LocalUpdate{Metropolis|HeatBath, Checkerboard} on a D-dim
 * Ising site field with the PCG64 stream; returns wall seconds spent inside the
 * sweeps (table construction excluded). */
double orc_pmode_ising_sweeps(int32_t dim, const int64_t *shape, int32_t kernel,
                              double beta, double j, double h, uint64_t seed,
                              uint8_t *state, uint32_t n_sweeps);
double orc_pmode_ising_sweeps_from(int32_t dim, const int64_t *shape, int32_t kernel,
                                   double beta, double j, double h, orc_pcg64 *gen,
                                   uint8_t *state, uint32_t n_sweeps);

/* ---- learned block proposal (plq_oracle_block.c) ----------------------------
 * 2D Ising, 8x8 blocks, block b = by * (L0/8) + bx, block colour (bx + by) & 1,
 * block site p = py * 8 + px.  Network parameters (FP32, ORC_BN_PARAMS):
 *   W1[9][32], b1[32], W2[9][32 out][32 in], b2[32], W3[9][32], b3
 * tap t = (dy + 1) * 3 + (dx + 1); layer outputs tanh except the logits.
 * The logit of a site is that of state 0 (spin +1).  `round` = 1 reproduces
 * the device's BF16 rounding of weights and activations. */
#define ORC_BN_BLOCK 8
#define ORC_BN_CHANNELS 32
#define ORC_BN_PARAMS 9857
#define ORC_BLOCK_PROPOSAL_KEY 0x3c6ef372u
#define ORC_BLOCK_ACCEPT_KEY 0xa54ff53au
float orc_expf(float x);
float orc_bf16(float f);
void orc_blocknet_logits(const float *params, int round, int64_t L0, int64_t L1, const uint8_t *state,
                         double *out);
double orc_blocknet_loss_grad(const float *params, int round, int64_t L0, int64_t L1, const uint8_t *state,
                              double *grad);
/* One pass over the blocks of `colour`; returns the number accepted. */
int64_t orc_block_mh_pass(int64_t L0, int64_t L1, uint8_t *state, const float *logits, int colour, uint32_t seed32,
                          uint32_t bstep, double beta, double j, double h);

/* ---- per-configuration measurements (plq_oracle_measure.c) -----------------
 * Correlator pair sums (models/mod.rs:44-60): out holds, axis after axis,
 * shape[mu] / 2 + 1 sums each; pair_kind 0 = Ising sign product, 1 = Potts
 * agreement.  Polyakov loop (gauge/mod.rs:178-201): returns the sum of the line
 * products, *lines the line count.  Wilson rectangles (gauge/mod.rs:224-268):
 * returns max = min(max_side, shortest / 2); totals is (max + 1)^2, row r,
 * column t, the sum over orientations and sites for r, t >= 1. */
void orc_axis_pair_sums(int32_t dim, const int64_t *shape, const uint8_t *state, int32_t pair_kind, int64_t *out);
int64_t orc_polyakov_sum(int32_t dim, const int64_t *shape, const uint8_t *links, int32_t dir, int64_t *lines);
int32_t orc_wilson_sums(int32_t dim, const int64_t *shape, const uint8_t *links, int32_t max_side, int64_t *totals,
                        int64_t *placements);

#ifdef __cplusplus
}
#endif
#endif
