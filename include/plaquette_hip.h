/*
 * plaquette_hip.h -- C ABI of the MI355X (gfx950) device path of plaquette.
 *
 * This is the drop-in boundary: the extern "C" surface a Rust `extern "C"`
 * block in the reference's src/device.rs and src/gpu_cluster.rs binds in place
 * of the wgpu device.  Plain pointers and sizes only; no HIP or torch types.
 *
 * Conventions shared by every entry point:
 *   - Return value: PLQ_OK (0) or a negative PLQ_E* code; plq_last_error()
 *     then holds a message (thread-local).  These codes replace the
 *     reference's panics (device.rs:618-646, 114-127; gpu_cluster.rs:267-283,
 *     512-518) and its `Option` for a missing adapter (device.rs:42-77).
 *   - States cross the ABI as uint8 state indices in lattice-index order:
 *     sites mixed-radix with axis 0 fastest (lattice.rs:75-78), links
 *     link = site * D + dir (lattice.rs:259-263).  The reference uses one u32
 *     word per variable (device.rs:721-732); the device layout (colour-
 *     interleaved int8) stays private to the library.
 *   - One handle is single-owner; calls on it serialise (the reference's
 *     `&mut self`).  Different handles may be driven from different threads.
 *   - `advance` only enqueues work on the handle's stream; `next`, `read`,
 *     `measure`, `series` and `sync` wait for it (device.rs:383-386, 462-470).
 *   - Randomness is the reference's keyed device stream (rng.wgsl:34-52) keyed
 *     by fold_seed(seed) (device.rs:761-763), the variable index and the
 *     chain's global sweep counter, so results are independent of the launch
 *     shape.  Acceptance tests are exact integer comparisons against
 *     thresholds the host derives from the reference's f64 CPU arithmetic
 *     (PLQ_PROFILE_F64, updater.rs:77-167) or the WGSL f32 arithmetic
 *     (PLQ_PROFILE_F32, metropolis.wgsl / heat_bath.wgsl, with
 *     exp_f32(x) = (float)exp((double)x)).
 */
#ifndef PLAQUETTE_HIP_H
#define PLAQUETTE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLQ_OK 0
#define PLQ_EINVAL (-22)  /* bad descriptor: odd extent, q out of range, size mismatch */
#define PLQ_ENODEV (-19)  /* no HIP device (reference: Gpu::new() -> None) */
#define PLQ_ENOMEM (-12)  /* device allocation failed / lattice too large */
#define PLQ_ENOCONV (-62) /* cluster labelling exceeded its pass cap */
#define PLQ_EDEVICE (-5)  /* a HIP runtime call or kernel launch failed */

#define PLQ_MAX_DIM 8
#define PLQ_MAX_Q 16

enum { PLQ_MODEL_ISING = 0, PLQ_MODEL_POTTS = 1, PLQ_MODEL_Z2 = 2 };
enum { PLQ_KERNEL_METROPOLIS = 0, PLQ_KERNEL_HEAT_BATH = 1 };
/* PLQ_PROFILE_CPU_STREAM (Ising site chains): the CPU path itself -- its f64
 * arithmetic and its own sequential stream, rand_pcg Lcg128Xsl64 seeded by
 * rand_core seed_from_u64(seed) (rng.rs:30-50), consumed in the CPU's site
 * order (colour 0's sites in lattice order, then colour 1's) with one draw
 * per uphill Metropolis site or per heat-bath site (updater.rs:77-167,
 * 316-324).  The draws are not keyed by the sweep, so a chain continues from
 * its stream state (plq_chain_pcg_state / plq_chain_set_pcg_state). */
enum { PLQ_PROFILE_F64 = 0, PLQ_PROFILE_F32 = 1, PLQ_PROFILE_CPU_STREAM = 2 };
enum { PLQ_EXTENT_ALL = 0, PLQ_EXTENT_SEEDED = 1 };
enum { PLQ_RELABEL_REDRAW = 0, PLQ_RELABEL_FORCED_CHANGE = 1 };

/* A local-update chain: GpuChain::build + the model's gpu_chain constructor
 * (device.rs:609-681; models/ising/gpu.rs:57-95, potts/gpu.rs, gauge/gpu.rs:54-110).
 * Ising and Potts live on sites, Z2 on links (cell follows from the model). */
typedef struct {
    int32_t dim;                   /* 1..4 (Z2: 2..4) */
    int64_t shape[PLQ_MAX_DIM];    /* every extent even (device.rs:714-719) */
    int32_t model;                 /* PLQ_MODEL_* */
    int32_t q;                     /* 2 for Ising and Z2; 2..8 for Potts */
    int32_t kernel;                /* PLQ_KERNEL_* */
    int32_t profile;               /* PLQ_PROFILE_* */
    double beta;
    double j;                      /* coupling J */
    double h;                      /* Ising field */
    double offsets[PLQ_MAX_Q];     /* Potts per-label offsets h[a] */
    uint64_t seed;                 /* folded with fold_seed on the host */
    uint64_t sweeps_between;       /* Chain::next advances this many sweeps */
} plq_local_desc;

/* A cluster chain: GpuClusterChain::new (gpu_cluster.rs:256-472) with the
 * model's bond gap (BondAction::bond_energy_gap: J for Potts, 2J for Ising). */
typedef struct {
    int32_t dim;                   /* 1..4 */
    int64_t shape[PLQ_MAX_DIM];    /* any positive extents */
    int32_t q;                     /* 2..16 */
    int32_t extent;                /* PLQ_EXTENT_ALL (Swendsen-Wang) / _SEEDED (Wolff) */
    int32_t relabel;               /* PLQ_RELABEL_* */
    int32_t profile;               /* bond test: f64 p (CPU) or f32 p (WGSL) */
    double beta;
    double gap;                    /* >= 0 and finite (updater.rs:447-478) */
    uint64_t seed;
    uint64_t sweeps_between;
} plq_cluster_desc;

/* One measurement: exact integer sums, reduced on the device, and the model
 * energy formed from them in f64 (ising/mod.rs:97-117,
 * potts/mod.rs:237-271, gauge/mod.rs:275-286). */
typedef struct {
    uint64_t sweep;                /* sweeps done when measured */
    int64_t n_vars;
    int64_t bond_sum;              /* Ising: sum over bonds of s_i s_j; Potts: agreeing bonds;
                                      Z2: signed plaquette sum */
    int64_t spin_sum;              /* Ising: magnetization M = sum s_i; otherwise 0 */
    int64_t counts[PLQ_MAX_Q];     /* label populations */
    double energy;                 /* H of the configuration */
} plq_obs;

/* Statistics of the last cluster sweep. */
typedef struct {
    uint64_t sweep;
    int64_t n_clusters;
    int64_t largest;               /* size of the largest cluster */
    int64_t moved;                 /* sites relabelled (Wolff: the seed's cluster size) */
    int64_t size_hist[40];         /* clusters with size in [2^k, 2^(k+1)) */
} plq_cluster_stats;

typedef struct plq_chain plq_chain;

/* Gpu::new (device.rs:42-77): number of usable HIP devices (0 -> PLQ_ENODEV). */
int plq_device_count(int32_t *count);
const char *plq_last_error(void);

/* GpuChain::build (device.rs:609-681).  `start` holds n state indices < q. */
int plq_chain_create_local(const plq_local_desc *desc, const uint8_t *start, size_t n,
                           plq_chain **out);
/* GpuClusterChain::new (gpu_cluster.rs:256-472). */
int plq_chain_create_cluster(const plq_cluster_desc *desc, const uint8_t *start, size_t n,
                             plq_chain **out);
/* GpuChain::advance / GpuClusterChain::advance (device.rs:387-399,
 * gpu_cluster.rs:477-526): enqueue `sweeps` sweeps, return without waiting,
 * for cluster chains too.  A cluster labelling that exceeds its pass cap sets
 * a device flag; the next call that waits for the chain (next, next_batch,
 * read, measure, series, cluster_roots, cluster_stats, cluster_series, sync)
 * returns PLQ_ENOCONV.  The flag is sticky: like the reference, which panics
 * (gpu_cluster.rs:512-518), such a chain holds no defined configuration and
 * every later waiting call returns PLQ_ENOCONV; destroy it. */
int plq_chain_advance(plq_chain *chain, uint64_t sweeps);
/* Iterator::next (device.rs:692-701, gpu_cluster.rs:630-637): advance
 * sweeps_between sweeps and copy the configuration out. */
int plq_chain_next(plq_chain *chain, uint8_t *out, size_t n);
int plq_chain_read(plq_chain *chain, uint8_t *out, size_t n);
/* `count` samples, each after sweeps_between sweeps, copied on the device into a
 * staging buffer and read back in one transfer: the sample batch of
 * GpuChain::run_batch (device.rs:431-488; GPU_BATCH = 64 at device.rs:497-501).
 * out holds count * n_vars states, sample after sample, in lattice order. */
int plq_chain_next_batch(plq_chain *chain, uint64_t count, uint8_t *out, size_t n);
int plq_chain_write(plq_chain *chain, const uint8_t *in, size_t n);
/* On-device measurement of the current configuration (replaces the host
 * readback + measure of device.rs:435-494 and ising/mod.rs:177-186). */
int plq_chain_measure(plq_chain *chain, plq_obs *out);
/* Record a measurement every `every` sweeps during advance, fused into the
 * sweep kernels where the layout allows; up to `capacity` records are kept
 * (older ones are dropped first).  every = 0 turns recording off. */
int plq_chain_record(plq_chain *chain, uint64_t every, uint64_t capacity);
int plq_chain_series(plq_chain *chain, plq_obs *out, uint64_t capacity, uint64_t *n_out);
int plq_chain_sweeps_done(plq_chain *chain, uint64_t *out);
/* Resume: with keyed draws a checkpoint is (states, sweep counter, seed). */
int plq_chain_set_sweep(plq_chain *chain, uint64_t sweep);
int plq_chain_sync(plq_chain *chain);
/* The hipStream_t the chain's kernels run on (for event timing). */
int plq_chain_stream(plq_chain *chain, void **stream);
/* Cluster chains only: the canonical partition of the last sweep (each site's
 * cluster root, the cluster's minimum site index) and its statistics. */
int plq_chain_cluster_roots(plq_chain *chain, int64_t *roots, size_t n);
int plq_chain_cluster_stats(plq_chain *chain, plq_cluster_stats *out);
/* Cluster chains only: with recording on (plq_chain_record), also keep the
 * statistics of each recorded sweep's partition (as plq_chain_cluster_stats
 * gives them) in a device ring beside the records, read back by
 * plq_chain_cluster_series without a synchronisation per sweep.  No
 * counterpart in the reference, which computes cluster statistics on the host
 * from read-back configurations.  While on, the chain stays on the
 * four-kernel sweep.  Calling it discards the kept statistics. */
int plq_chain_record_clusters(plq_chain *chain, int on);
int plq_chain_cluster_series(plq_chain *chain, plq_cluster_stats *out, uint64_t capacity, uint64_t *n_out);
/* Cluster chains only: lower the labelling pass cap (gpu_cluster.rs:620-625);
 * used to force PLQ_ENOCONV from the host. */
int plq_chain_set_pass_cap(plq_chain *chain, uint64_t cap);
/* Wolff chains only: how many sweeps grew the seed's cluster on the device
 * (the CPU's grow_cluster, cluster.rs:88-111, over the sweep's keyed bonds)
 * and how many ran the full decomposition of the reference's device move
 * (cluster_relabel.wgsl:55-66) after a growth gave up or was skipped; both
 * zero while growth is off.  No counterpart in the reference: the states do
 * not depend on the path, so this only reports the cost split.  Waits for the
 * chain's stream. */
int plq_chain_wolff_counts(plq_chain *chain, uint64_t *grown, uint64_t *full);
int plq_chain_destroy(plq_chain *chain);
/* PLQ_PROFILE_CPU_STREAM chains only: the stream state as four u64
 * (state lo, state hi, increment lo, increment hi) -- the two u128 fields of
 * rand_pcg's Lcg128Xsl64 -- read, or replaced to continue a CPU run's stream. */
int plq_chain_pcg_state(plq_chain *chain, uint64_t *out4);
int plq_chain_set_pcg_state(plq_chain *chain, const uint64_t *in4);
/* Configuration::hot (configuration.rs:80-87) as a run config builds it
 * (ising/run_config.rs:237-251): RandRng::seed_from_u64(seed), then one
 * random_range(0..q) (rng.rs:46-49) per variable in index order.  state4 (may
 * be NULL) receives the stream state after the draws, which a
 * PLQ_PROFILE_CPU_STREAM chain continues from (plq_chain_set_pcg_state).
 * Host-only; no device is needed. */
int plq_hot_start(uint64_t seed, int32_t q, uint8_t *out, size_t n, uint64_t *state4);

/* ---------------------------------------------------------------------------
 * Per-configuration measurements, computed where the configuration lives (the
 * reference reads the configuration back and walks it on the CPU).  The sums
 * are exact integers and the f64 results use the reference's own final
 * division, so every value equals the reference's bit for bit.
 * ------------------------------------------------------------------------- */
/* Two-point correlator of a site chain along every axis: Ising::correlator
 * (ising/mod.rs:83-94) for model PLQ_MODEL_ISING (q = 2), Potts::correlator
 * (potts/mod.rs:218-233) for PLQ_MODEL_POTTS, both over axis_pair_sums
 * (models/mod.rs:44-60).  Entries run axis after axis, r = 0..shape[mu]/2,
 * so len = sum over mu of (shape[mu]/2 + 1).  `sums` (may be NULL) receives
 * the raw pair sums (Ising sign products, Potts agreement counts), `values`
 * (may be NULL) sum / N (Ising) or sum / N - 1/q (Potts).  A local site chain
 * takes its own model; a cluster chain either (Ising needs q = 2). */
int plq_chain_correlator(plq_chain *chain, int32_t model, int64_t *sums, double *values, size_t len);
/* The same correlator without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * between HIP graph captures, tight serving loops, harnesses that forbid
 * hipMalloc).  The bit planes and counters live in a fixed device array of
 * the library (2^21 plane words + 8,192 counters, 8.06 MiB per device, held
 * until the process ends): a configuration needing more than 2^21 words
 * (ceil(log2 q) * N / L0 * ceil(L0 / 32)) or more than 8,192 entries gets
 * PLQ_EINVAL; plq_chain_correlator has no such limit.  The counters are read
 * back into the caller's `sums` (or, if NULL, `values`) storage, so at least
 * one of them must be given.  Calls on one device share the array; the
 * library serialises them. */
int plq_chain_correlator_noalloc(plq_chain *chain, int32_t model, int64_t *sums, double *values, size_t len);
/* Z2 gauge chains: Z2Gauge::polyakov_loop (gauge/mod.rs:178-201) along dir. */
int plq_chain_polyakov_loop(plq_chain *chain, int32_t dir, double *value);
/* Z2 gauge chains: Z2Gauge::wilson_rectangles (gauge/mod.rs:224-268).
 * side = min(max_side, shortest extent / 2) goes to *side; `table` (NULL to
 * ask for the side only) receives the (side + 1)^2 entries row-major,
 * table[r * (side + 1) + t], with len >= (side + 1)^2. */
int plq_chain_wilson_rectangles(plq_chain *chain, int32_t max_side, double *table, size_t len, int32_t *side);

/* ---------------------------------------------------------------------------
 * Batched Z2 gauge ensembles (no reference counterpart: the reference builds
 * one GpuChain per lattice, gauge/gpu.rs:54-110).  Many replicas of uneven
 * shapes advance together, one launch per colour pass for all of them; every
 * replica keeps the reference's link checkerboard and keyed draws
 * (link_schedule.wgsl:23-40, gauge/model.wgsl:136-147), keyed by its own seed.
 * Every `measure_every` sweeps the plaquette sums of all replicas are recorded
 * and, with `exchange` set, one replica-exchange round swaps inverse
 * temperatures between replicas of equal shape whose betas are neighbours:
 * round n pairs the k-th and (k+1)-th smallest beta for k = n mod 2,
 * n mod 2 + 2, ...; the pair swaps iff delta = (beta_k+1 - beta_k) * J * (P_a -
 * P_b) >= 0 or m < T, m the keyed draw (fold_seed(exchange_seed) ^ 0x68e31da4,
 * global pair index, round) and T the f64 threshold of exp(delta).  A block of
 * measure_every sweeps (with its measurement and exchange) can be replayed as
 * one captured hipGraph.
 * ------------------------------------------------------------------------- */
typedef struct {
    int32_t dim;                 /* 2..4 */
    int32_t n_replicas;
    int32_t kernel;              /* PLQ_KERNEL_* */
    int32_t profile;             /* PLQ_PROFILE_* */
    const int64_t *shapes;       /* n_replicas x dim, every extent even */
    const double *betas;         /* replica r starts at betas[r] */
    const uint64_t *seeds;       /* per replica */
    double j;
    uint64_t measure_every;      /* sweeps between recorded measurements; 0 = none */
    int32_t exchange;            /* 1 = an exchange round after every measurement */
    uint64_t exchange_seed;
    uint64_t record_capacity;    /* measurements kept (ring) */
} plq_ensemble_desc;

typedef struct plq_ensemble plq_ensemble;

/* `start` concatenates the replicas' lattice-order link states
 * (replica r: n_sites_r * dim bytes, link = site * dim + dir). */
int plq_ensemble_create(const plq_ensemble_desc *desc, const uint8_t *start, size_t n, plq_ensemble **out);
int plq_ensemble_advance(plq_ensemble *e, uint64_t sweeps);
int plq_ensemble_read(plq_ensemble *e, uint8_t *out, size_t n);
/* Current signed plaquette sum of every replica (synchronises). */
int plq_ensemble_plaquette_sums(plq_ensemble *e, int64_t *sums, size_t n);
/* Current inverse temperature of every replica. */
int plq_ensemble_betas(plq_ensemble *e, double *betas, size_t n);
/* Recorded measurements, oldest first: sums[k * n_replicas + r] and the beta
 * slot (index into the descriptor's betas) replica r held when measured. */
int plq_ensemble_series(plq_ensemble *e, int64_t *sums, uint32_t *slots, uint64_t capacity, uint64_t *n_out);
/* Accepted swaps and attempts per exchange pair (pairs in group order). */
int plq_ensemble_exchange_counts(plq_ensemble *e, uint64_t *accepted, uint64_t *attempted, size_t npairs,
                                 size_t *npairs_out);
/* Launch shape: lanes per block, links per lane, and how a block of sweeps
 * is launched: use_graph 0 one launch per colour pass, 1 those launches
 * replayed from a captured hipGraph, 2 every pass of the block in one
 * cooperative launch with grid barriers between the passes (bit for bit the
 * same; measured slower on MI355X, so the autotuner does not pick it there). */
int plq_ensemble_set_launch(plq_ensemble *e, int32_t block_size, int32_t per_lane, int32_t use_graph);
/* Time the candidate launch shapes on a scratch copy of the state and keep
 * the fastest; the chain's state and counters are left as they were. */
int plq_ensemble_tune(plq_ensemble *e, int32_t *block_size, int32_t *per_lane, int32_t *use_graph,
                      double *best_us_per_sweep);
int plq_ensemble_sweeps_done(plq_ensemble *e, uint64_t *out);
int plq_ensemble_stream(plq_ensemble *e, void **stream);
int plq_ensemble_sync(plq_ensemble *e);
int plq_ensemble_destroy(plq_ensemble *e);

/* ---------------------------------------------------------------------------
 * Learned block proposal (no reference counterpart: the reference's local
 * updates change one variable at a time, updater.rs:291-347).  A 2D Ising
 * chain whose extents are multiples of 16 is cut into 8x8 blocks coloured
 * (bx + by) & 1.  A 3-layer CNN (3x3 kernels, 32 channels, tanh; BF16 weights
 * and activations, FP32 accumulation) maps the 32 spins of a block's boundary
 * ring to 64 logits, and the proposal draws the block's interior as 64
 * independent Bernoulli variables, state 0 with probability sigmoid(logit).
 * The move is accepted by Metropolis-Hastings with the true Ising action
 * (ising/mod.rs:97-143): A = beta (J dBond + h dMag) + sum_p ([x_p = 0] -
 * [x'_p = 0]) logit_p, accepted iff A >= 0 or m < exp(A) 2^24.  Draws are
 * keyed like the chain's own (rng.wgsl:47-52): proposal bits by
 * (seed ^ 0x3c6ef372, site, block sweep), acceptance by (seed ^ 0xa54ff53a,
 * block index by * L0/8 + bx, block sweep); the acceptance arithmetic is IEEE
 * single with a deterministic exp, reproducible bit for bit given the logits.
 * Parameters (PLQ_BN_PARAMS floats): W1[9][32], b1[32], W2[9][32 out][32 in],
 * b2[32], W3[9][32], b3, tap t = (dy + 1) * 3 + (dx + 1).
 * Calls that change the network wait for all work on the device first.
 * ------------------------------------------------------------------------- */
#define PLQ_BN_PARAMS 9857
typedef struct plq_blocknet plq_blocknet;

/* Glorot-uniform weights drawn from a splitmix64 stream of `seed`, zero biases. */
int plq_blocknet_create(uint64_t seed, plq_blocknet **out);
int plq_blocknet_params(plq_blocknet *net, float *out, size_t n);
/* Replace the parameters and reset the Adam state. */
int plq_blocknet_set_params(plq_blocknet *net, const float *in, size_t n);
/* Mean binary cross-entropy of every block's interior given its ring over the
 * chain's current configuration, and its gradient. */
int plq_blocknet_gradient(plq_blocknet *net, plq_chain *chain, float *grad, size_t n, double *loss);
/* `steps` times: advance the chain `sweeps_between` sweeps, then one Adam step
 * (0.9, 0.999, 1e-8) on the gradient over its blocks.  losses (nullable) gets
 * each step's mean loss. */
int plq_blocknet_train(plq_blocknet *net, plq_chain *chain, uint64_t steps, uint64_t sweeps_between, double lr,
                       float *losses);
/* Logits of every block of the chain's current configuration, indexed
 * (by * L0/8 + bx) * 64 + py * 8 + px. */
int plq_blocknet_logits(plq_blocknet *net, plq_chain *chain, float *out, size_t n);
int plq_blocknet_destroy(plq_blocknet *net);

/* Enqueue `sweeps` block sweeps (blocks of colour 0, then colour 1). */
int plq_chain_block_sweeps(plq_chain *chain, plq_blocknet *net, uint64_t sweeps);
/* Enqueue `steps` times: one block sweep, then `plain_sweeps` sweeps of the
 * chain's own kernel (recording as plq_chain_record set it up). */
int plq_chain_advance_with_blocks(plq_chain *chain, plq_blocknet *net, uint64_t steps, uint64_t plain_sweeps);
/* Accepted and proposed block moves so far, and block sweeps done. */
int plq_chain_block_stats(plq_chain *chain, uint64_t *accepted, uint64_t *proposed, uint64_t *block_sweeps);
/* The logits each block used in its latest block pass (layout as above). */
int plq_chain_block_logits(plq_chain *chain, float *out, size_t n);

/* Host-side helpers shared with the device path, exposed for tests. */
uint32_t plq_fold_seed(uint64_t seed);
/* The integer acceptance table the local kernels compare the 24-bit draw
 * against: Ising/Z2 classes are (current state, number of "down" (state 1)
 * neighbours or negative staple products).  Returns the number of entries. */
int plq_threshold_table(const plq_local_desc *desc, uint32_t *out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
