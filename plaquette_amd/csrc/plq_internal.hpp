// plq_internal.hpp -- declarations shared between the host library and the
// kernel translation units.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "plq_common.hpp"

namespace plq {

// Record `msg` as the calling thread's last error and return `code`.
int host_fail(int code, const std::string &msg);

constexpr int MAX_Q_DEV = 16;
// One measurement record on the device: [bond_sum, unused, counts[16]].
constexpr int REC_K = 18;
constexpr int MAX_PASS_BLOCKS = 2048;

// Log2 form of a power-of-two geometry for the SWAR Ising pass.
struct SwarGeom {
    uint32_t lh4;   // log2(L0 / 8): dwords per colour row
    uint32_t L1, L2, L3;
    uint32_t ll1, ll2;
};

// A record fold (launch_finalize_columns' work) carried into the first
// kernel of the next sweep instead of a launch of its own: workgroups
// k < nk of that kernel sum column k of the column-major partials of up to
// two passes (pa: na blocks, pb: nb blocks) into rec[k], and workgroup 0
// clears rec[nk .. REC_K).  rec == nullptr: nothing to fold.
struct RecordFold {
    const long long *pa;
    int64_t na;
    const long long *pb;
    int64_t nb;
    int nk;
    long long *rec;
};

// The fold, called by every thread of a workgroup (it synchronises the
// workgroup); lds holds at least 16 entries.  Column sums keep eight loads
// in flight per lane.
__device__ __forceinline__ void fold_record(const RecordFold &f, long long *lds) {
    const int k = (int)blockIdx.x;
    if (k == 0)
        for (int j = f.nk + (int)threadIdx.x; j < REC_K; j += (int)blockDim.x) f.rec[j] = 0;
    if (k >= f.nk) return;
    auto col = [](const long long *__restrict__ src, int64_t n) -> long long {
        long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int64_t step = blockDim.x;
        int64_t b = threadIdx.x;
        for (; b + 7 * step < n; b += 8 * step)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += src[b + u * step];
        for (; b < n; b += step) acc[0] += src[b];
        return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    };
    long long v = col(f.pa + k * f.na, f.na);
    if (f.pb) v += col(f.pb + k * f.nb, f.nb);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63u) == 0u) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[w];
        f.rec[k] = t;
    }
    __syncthreads();  // lds is free again
}

// Fused measurement of a colour pass: partial sums per block, column-major
// (column k of block b at part[k * blocks + b]), folded into a record by
// launch_finalize_columns after the colour-1 pass.
struct MeasArgs {
    long long *part;       // this pass's partials
    int on;                // this pass measures
    RecordFold fold;       // ising_pass_x16u only: an earlier record's fold
};

// Geometry of the sixteen-sites-per-lane pass whose waves hold rows of one
// parity (ising_pass_x16u): lanes are mapped so that the 64 lanes of a wave
// cover 2^(6 - lh16) rows two apart, so every row of a wave has the same
// colour offset along axis 0.  Byte offsets in a colour array are
// (row << sh) | (chunk << 4).
struct X16UGeom {
    uint32_t lh16;       // log2(16-byte chunks per colour row), 0..6
    uint32_t sh;         // lh16 + 4: log2(bytes per colour row)
    uint32_t lk2;        // 7 - lh16: log2 of the rows a pair of waves covers
    uint32_t ll1, ll2;   // log2 L1, log2 L2
    uint32_t m1, m2, m3; // L1 - 1, L2 - 1, L3 - 1
};

// The 3D Ising threshold table (own spin s, d of 6 neighbours down: entry
// d + 7 s), passed by value to the fused sweep so that a launch fills its LDS
// copy from kernel arguments instead of a dependent global load.
constexpr int SWEEP_TH_LEN = 14;
struct SweepThresholds {
    uint32_t t[SWEEP_TH_LEN];
};

struct SitePassArgs {
    const uint8_t *oth;  // the other colour's array (read)
    uint8_t *own;        // this colour's array (updated)
    Geom g;
    SwarGeom sg;
    X16UGeom ug;
    int uniform;         // 1: ising_pass_x16u applies (see x16u_applies)
    int vec;             // sites per lane: 1 (generic), 4 or 16 (power-of-two Ising)
    int model;           // 0 Ising, 1 Potts
    int mode;            // 0 Metropolis, 1 heat bath, 2 measure only
    int color;
    uint32_t seed;
    uint32_t skey;       // sweep * 0x9e3779b9
    const uint32_t *ctr; // uniform pass only, non-null in a graph: sweep = ctr[0] + sweep_off
    uint32_t sweep_off;
    int q;
    const uint32_t *table;
    MeasArgs m;
    int64_t max_blocks;
    hipStream_t stream;
};

void launch_pack_sites(const uint8_t *lat, uint8_t *c0, uint8_t *c1, const Geom &g, hipStream_t s);
void launch_unpack_sites(uint8_t *lat, const uint8_t *c0, const uint8_t *c1, const Geom &g, hipStream_t s);
void launch_site_pass(const SitePassArgs &a);
// Whether the uniform-parity sixteen-site pass serves this geometry and
// threshold table (every threshold at least 1; shapes as X16UGeom needs).
bool x16u_applies(const Geom &g, const uint32_t *host_table, size_t table_len, int64_t max_blocks);
// The fused two-colour 3D sweep (ising_sweep_x16f): both colour passes of a
// sweep in one launch, reading o0/o1 and writing n0/n1 (the caller swaps
// them after), for chains the uniform pass serves with L0 = 256 and L1, L2
// multiples of 16 (at least 32); one workgroup per 16 x 16 tile of rows.
// With m.on the launch also writes the record's partial sums (one column
// entry per tile, as the colour passes' partials) into m.part.
bool x16f_applies(const Geom &g, int uniform);
int64_t x16f_tiles(const Geom &g);
void launch_site_sweep_x16f(const uint8_t *o0, const uint8_t *o1, uint8_t *n0, uint8_t *n1, const Geom &g,
                            const X16UGeom &ug, int mode, uint32_t seed, uint32_t skey, const SweepThresholds &th,
                            const RecordFold &fold, const MeasArgs &m, hipStream_t s);
X16UGeom x16u_geom(const Geom &g);
int site_pass_blocks(const SitePassArgs &a);
void launch_finalize_record(const long long *pa, int64_t na, const long long *pb, int64_t nb, long long *rec,
                            hipStream_t s);
void launch_finalize_columns(const long long *pa, int64_t na, const long long *pb, int64_t nb, int nk, long long *rec,
                             hipStream_t s);
// The same into slot ctr[1] of the record ring at rec_base (graph replays).
void launch_finalize_columns_ring(const long long *pa, int64_t na, const long long *pb, int64_t nb, int nk,
                                  long long *rec_base, const uint32_t *ctr, hipStream_t s);
// A site chain's graph counters: ctr[0] = sweep, ctr[1] = record slot; the
// tick advances them by one replay (sweeps, and records into a ring of cap).
void launch_site_ctr_set(uint32_t *ctr, uint32_t sweep, uint32_t slot, hipStream_t s);
void launch_site_ctr_tick(uint32_t *ctr, uint32_t sweeps, uint32_t records, uint32_t cap, hipStream_t s);

// One-workgroup chain of a small power-of-two Ising lattice (plq_small.hip):
// nsweeps sweeps in one launch with the lattice held in LDS.
constexpr int64_t SMALL_MAX_SITES = 8192;
struct SmallChainArgs {
    uint8_t *c0, *c1;     // colour arrays, read at the start and written back at the end
    int D;
    uint32_t lg[MAX_D];   // log2 of each extent; L0 >= 2
    int64_t N;
    int mode;             // 0 Metropolis, 1 heat bath
    uint32_t seed;
    uint32_t sweep0;      // chain sweep counter of the first sweep
    uint32_t nsweeps;
    const uint32_t *table;
    long long *rec;       // record ring (REC_K per slot), rec_cap slots
    uint64_t rec_cap;
    uint64_t rec_first;   // ring position of this launch's first record
    uint64_t every;       // record after sweep t iff (phase + t + 1) % every == 0; 0 = never
    uint64_t phase;
    hipStream_t stream;
};
int64_t small_chain_max_sites(int D);
void launch_small_chain(const SmallChainArgs &a);

// --------------------------------------------------- learned block proposal ---

constexpr int BN_BLOCK = 8;
constexpr int BN_CHANNELS = 32;
// FP32 parameters: W1[9][32], b1[32], W2[9][32 out][32 in], b2[32], W3[9][32], b3
constexpr int BN_OFF_W1 = 0;
constexpr int BN_OFF_B1 = BN_OFF_W1 + 9 * BN_CHANNELS;
constexpr int BN_OFF_W2 = BN_OFF_B1 + BN_CHANNELS;
constexpr int BN_OFF_B2 = BN_OFF_W2 + 9 * BN_CHANNELS * BN_CHANNELS;
constexpr int BN_OFF_W3 = BN_OFF_B2 + BN_CHANNELS;
constexpr int BN_OFF_B3 = BN_OFF_W3 + 9 * BN_CHANNELS;
constexpr int BN_PARAMS = BN_OFF_B3 + 1;  // 9857
constexpr uint32_t BLOCK_PROPOSAL_KEY = 0x3c6ef372u;  // proposal draws: keyed (seed ^ this, site, block sweep)
constexpr uint32_t BLOCK_ACCEPT_KEY = 0xa54ff53au;    // acceptance draw: keyed (seed ^ this, block, block sweep)
constexpr int BLOCK_TRAIN_GROUPS = 512;               // workgroups of the gradient kernel
constexpr int BLOCK_PART_STRIDE = 9984;               // floats per workgroup partial (>= BN_PARAMS + 1)
constexpr int BLOCK_W2_FRAG = 9 * 2 * 64 * 8;         // bf16 entries of one packed W2 form

// Device forms of the network: biases straight from the FP32 parameters,
// W1 / W3 rounded to BF16 (held as float), W2 as matrix-core fragments.
struct BlockNetView {
    const float *w1, *b1, *b2, *w3, *b3;
    const uint16_t *w2f, *w2t;
    const uint16_t *w13f;  // MFMA B fragments of layers 1 and 3 (3 * 64 * 8 BF16)
};

struct BlockPassArgs {
    uint8_t *c0, *c1;     // the chain's colour arrays
    int L0, L1;           // multiples of 16
    int mode;             // 0: Metropolis-Hastings pass over `colour`; 1: logits of every block only
    int colour;
    uint32_t seed32, bstep;
    double beta, j, h;
    BlockNetView net;
    float *logits;        // nullable: n_blocks * 64 logits of the blocks evaluated
    unsigned long long *acc_slots;  // mode 0: accepted moves, one slot per workgroup (block_pass_groups)
    hipStream_t stream;
};
struct BlockGradArgs {
    const uint8_t *c0, *c1;
    int L0, L1;
    BlockNetView net;
    float *partials;      // (BLOCK_TRAIN_GROUPS + BLOCK_TRAIN_GROUPS / 16) * BLOCK_PART_STRIDE
    float *grad;          // BN_PARAMS + 1: mean gradient, then the mean loss
    hipStream_t stream;
};
void launch_block_pack(const float *master, float *w1r, float *w3r, uint16_t *w2f, uint16_t *w2t, uint16_t *w13f,
                       hipStream_t s);
constexpr int BLOCK_W13_FRAG = 3 * 64 * 8;  // uint16 entries of BlockNetView::w13f
void launch_block_pass(const BlockPassArgs &a);
int block_pass_groups(int L0, int L1);  // workgroups of a mode-0 pass
int block_train_groups(int L0, int L1);
void launch_block_grad(const BlockGradArgs &a);
void launch_block_adam(float *w, float *m, float *v, const float *grad, float lr, float beta1, float beta2,
                       float c1, float c2, float eps, hipStream_t s);

// ------------------------------------------- the CPU stream (plq_pmode.hip) ---

struct PmodeArgs {
    uint8_t *own;            // the colour being updated
    const uint8_t *oth;
    Geom g;
    int color;
    int mode;                // 0 Metropolis, 1 heat bath
    const uint64_t *cls;     // per class (state, down neighbours): 53-bit threshold, or ~0 (no draw)
    uint32_t *blk, *off;     // pm_blocks entries, pm_blocks + 1 entries
    uint64_t *state;         // [state lo, state hi, increment lo, increment hi]
    const uint64_t *jt;      // 64 x (A lo, A hi, C lo, C hi): k = 2^b steps
    hipStream_t stream;
};
// rand_core seed_from_u64 + rand_pcg Lcg128Xsl64 (state lo, hi, increment lo, hi).
void pm_seed(uint64_t seed, uint64_t out[4]);
// Configuration::hot (configuration.rs:80-87) from RandRng::seed_from_u64(seed):
// n states by random_range(0..q) (rng.rs:46-49), and the stream state after them.
void pm_hot_start(uint64_t seed, uint32_t q, uint8_t *out, size_t n, uint64_t state_out[4]);
void pm_jump_table(const uint64_t state[4], std::vector<uint64_t> &jt);
bool pm_class_table(int D, int kernel, double beta, double j, double h, std::vector<uint64_t> &cls,
                    std::string &err);
int64_t pm_blocks(const Geom &g);
void launch_pm_pass(const PmodeArgs &a);

// ------------------------------------------------------------------ links ---

struct LinkPassArgs {
    uint8_t *const *arr;  // 2*D arrays, index dir*2 + colour (device pointers, in a device array)
    Geom g;
    int dir;
    int color;
    int mode;             // 0 Metropolis, 1 heat bath
    uint32_t seed;
    uint32_t skey;
    const uint32_t *table;
    hipStream_t stream;
};
void launch_pack_links(const uint8_t *lat, uint8_t *const *arr, const Geom &g, hipStream_t s);
void launch_unpack_links(uint8_t *lat, uint8_t *const *arr, const Geom &g, hipStream_t s);
void launch_link_pass(const LinkPassArgs &a);
// Plaquette sum into rec[0] (partials reduced on the device).
void launch_plaquette_sum(uint8_t *const *arr, const Geom &g, long long *partials, int64_t max_blocks,
                          long long *rec, hipStream_t s);

// ---------------------------------------------------------------- cluster ---

// Plain lattice-order geometry cut into boxes for the LDS union-find.
struct CCGeom {
    int32_t D;
    int32_t L[MAX_D];
    int64_t stride[MAX_D];
    uint64_t mL[MAX_D];   // ceil(2^40 / L[mu]) for division by multiplication
    int32_t mul_div;      // 1 when N * max(L) < 2^40, so the reciprocal division is exact
    int32_t B[MAX_D];     // box extents
    int32_t nbox[MAX_D];  // boxes per axis
    int64_t N;
    int64_t nboxes;
    int32_t box_sites;    // max sites in one box (LDS entries)
};
CCGeom make_ccgeom(int D, const int64_t *shape, int32_t max_box_sites);

struct ClusterSweepArgs {
    uint8_t *labels;      // lattice order
    int32_t *parent;      // union-find forest, lattice order
    uint8_t *cross;       // nullable: N bytes, cc_compress marks the roots of face-crossing clusters
    bool count_sizes;     // the relabel also counts cluster sizes (box kernels only) into:
    int32_t *sizes, *ext;
    unsigned long long *hist_part;  // per-workgroup histogram rows (as launch_cluster_sizes_box's part)
    CCGeom cg;
    uint32_t seed32;
    uint32_t sweep;
    uint32_t bond_t;      // open iff agree && m < bond_t
    int q;
    int seeded;
    int forced;
    uint32_t cap;         // hop budget of every find / union
    int *flag;            // set to 1 when a budget is exceeded
    long long *partials;  // REC_K per relabel block: [0, moved, counts...]
    int64_t max_blocks;
    // Gating by a Wolff growth (plq_wolff.hip): with gate non-null every
    // kernel of the sweep returns at once when *gate == gate_skip.
    const int *gate;
    int gate_skip;
    bool label_only;      // build the partition only: no relabel, no partials
    // With meas_rec non-null and cluster_relabel_measures(a), the relabel
    // also counts the Potts record of the new labels (potts_measure_vec's
    // columns) into meas_part (REC_K * max_blocks entries), folded into
    // meas_rec after it, so no separate measurement is needed.
    long long *meas_part;
    long long *meas_rec;
    // An earlier record's fold for this sweep's first kernel (cc_local_x16;
    // other paths run launch_finalize_columns first), and with defer_fold
    // the fused record above is left for the caller to fold later: the
    // caller takes cluster_record_fold(a, blocks) as its pending fold.
    RecordFold fold;
    bool defer_fold;
    hipStream_t stream;
};
// Enqueue one cluster sweep; returns the number of relabel blocks whose
// partials were written.
int64_t launch_cluster_sweep(const ClusterSweepArgs &a);
// The number of relabel blocks launch_cluster_sweep gives the same arguments.
int64_t cluster_relabel_blocks(const ClusterSweepArgs &a);
// Whether the relabel of a sweep with these arguments can count the sweep's
// Potts record itself (the per-box relabel of a Swendsen-Wang sweep: no seed
// cluster, no forced change, no gate, no size counting; q <= MAX_Q_DEV).
bool cluster_relabel_measures(const ClusterSweepArgs &a);
// The fold of a fused record left by a sweep with defer_fold (blocks: what
// launch_cluster_sweep returned).
RecordFold cluster_record_fold(const ClusterSweepArgs &a, int64_t blocks);

// Wolff moves grown from the seed site (plq_wolff.hip).  One workgroup
// grows the seed's cluster level by level over the sweep's keyed bonds and,
// when it closes within the level and queue budgets, relabels it and sets
// gate = 0, so the full decomposition of the same sweep (gated) is skipped;
// otherwise gate = 1 and the full decomposition runs.
struct WolffState {
    int gate;                // 0: the last sweep grew its cluster; 1: the full decomposition ran
    int since;               // sweeps since the last attempt (adaptive mode)
    float cost;              // running mean of an attempt's duration, wall-clock ticks (adaptive mode)
    float success;           // running mean of the attempts' success (adaptive mode)
    int old_label;           // the grown cluster's label before the move
    uint32_t size;           // the grown cluster's size (queue[0 .. size))
    unsigned long long grown, full;  // sweeps by growth, by the full decomposition
};
struct WolffGrowArgs {
    uint8_t *labels;
    CCGeom cg;
    uint32_t seed32;
    uint32_t sweep;
    uint32_t bond_t;
    int q;
    int forced;
    int32_t *queue;          // min(qcap, WOLFF_QUEUE) entries: the grown cluster's sites
    uint32_t qcap;
    uint32_t level_cap;
    int adaptive;            // attempt only where it is expected to pay, and every 32nd or 256th sweep
    float gain_ticks;        // wall-clock ticks a grown sweep saves against the decomposition
    WolffState *state;
    long long *partials;     // relabel partials: nparts rows of REC_K
    int64_t nparts;
    hipStream_t stream;
};
// LDS of the growth: the cluster's sites and the set of sites reached.
constexpr uint32_t WOLFF_SET_BITS = 13;
constexpr uint32_t WOLFF_SET_SLOTS = 1u << WOLFF_SET_BITS;
constexpr uint32_t WOLFF_QUEUE = 6144;
void launch_wolff_grow(const WolffGrowArgs &a);
// dst[queue[i]] = the grown cluster's old label for i < size, when gate == 0:
// the labels the last grown sweep started from, given its result in dst.
void launch_wolff_restore(uint8_t *dst, const int32_t *queue, const WolffState *state, hipStream_t s);
// With defer non-null, a measurement whose partials are column-major leaves
// their fold in *defer instead of launching it (defer->rec stays null when
// the fold was launched).
void launch_potts_measure(const uint8_t *labels, const CCGeom &cg, int q, long long *partials,
                          int64_t max_blocks, long long *rec, hipStream_t s, RecordFold *defer = nullptr);
// Canonical roots of the current forest, and (when sizes/hist are non-null)
// cluster sizes and the size histogram (hist[0..39] by floor(log2 size),
// hist[40] cluster count, hist[41] largest size).
// Whether the sweep uses the 16-sites-per-lane box kernels (power-of-two
// boxes, box extent along axis 0 a multiple of 16, 16-byte aligned labels).
bool cluster_boxes_x16(const CCGeom &cg, const void *labels);
// Size histogram of the last sweep's clusters on those boxes (hist zeroed by
// the caller; ext all zero on entry and on return).
// cross: the marks cc_compress leaves on the roots of face-crossing clusters
// (N bytes, tag cross_tag(sweep) of the sweep whose partition is counted).
// part: scratch for 2 * MAX_PASS_BLOCKS rows of 42 partials (stored
// column-major, 42 * 2 * MAX_PASS_BLOCKS entries); hist[42 ..] also
// takes the relabel's partials summed (relabel_rows rows of REC_K; only the
// relabelled-site column is kept).
// counted: the sweep's relabel already counted the sizes (ClusterSweepArgs::count_sizes).
void launch_cluster_sizes_box(const int32_t *parent, const CCGeom &cg, int32_t *sizes, int32_t *ext,
                              const uint8_t *cross, uint32_t sweep, unsigned long long *part,
                              const long long *relabel_part, int64_t relabel_rows, unsigned long long *hist,
                              int64_t max_blocks, bool counted, hipStream_t s);
uint8_t cross_tag(uint32_t sweep);
void launch_cluster_resolve(int32_t *parent, int32_t *roots, const CCGeom &cg, uint32_t cap, int *flag,
                            int32_t *sizes, unsigned long long *hist, hipStream_t s);

// --- small-lattice cluster chains (plq_cluster_small.hip) ------------------
// Swendsen-Wang / Wolff sweeps of a lattice of at most
// small_cluster_max_sites(D) sites in one workgroup, up to 4,096 sweeps per
// launch.
constexpr int SMALL_CC_MAX_SITES = 8192;
struct SmallClusterArgs {
    uint8_t *labels;        // lattice order, read at the start and written back at the end
    int32_t *parent;        // receives the last sweep's forest (every site -> its root)
    long long *partials;    // REC_K entries: [0, sites relabelled by the last sweep, 0, ...]
    int D;
    int32_t L[MAX_D];
    uint32_t N;
    int q;
    int seeded, forced;     // Wolff: Extent::Seeded, Relabel::ForcedChange
    uint32_t bond_seed, root_seed, seed_seed;
    uint32_t bond_t;        // open iff agree && m < bond_t
    uint32_t sweep0, nsweeps;
    long long *rec;         // record ring: [agree, 0, counts[0..q)] per recorded sweep
    uint64_t rec_cap, rec_first, every, phase;  // as SmallChainArgs
    hipStream_t stream;
};
void launch_small_cluster(const SmallClusterArgs &a);
int64_t small_cluster_max_sites(int D);

// --- per-configuration measurements (plq_measure.hip) ----------------------
// Bit planes of a lattice-order configuration: row R = the axis-0 line with
// coordinates x1..x_{D-1} (R = x1 + L1 * (x2 + L2 * x3)), W words per row,
// bit x0 % 32 of word x0 / 32, bits past L0 zero; plane k at k * R * W.
struct MeasGeom {
    int32_t D;
    int32_t L[MAX_D];        // extents; unused axes 1
    int32_t rstride[MAX_D];  // row stride of axis a >= 1
    int32_t R;               // rows, N / L0
    int32_t W;               // words per row, ceil(L0 / 32)
    int64_t N;
};
MeasGeom meas_geom(int D, const int64_t *shape);
// Bit planes that hold a state 0..q-1.
int meas_state_planes(int q);
// Plane v * bits_per_var + p <- bit p of variable v (vars_per_site variables
// per site, lat[site * vars_per_site + v]); vars_per_site 1 with
// bits_per_var 1..4, or 2..4 with bits_per_var 1.
void launch_meas_pack(const uint8_t *lat, const MeasGeom &m, int vars_per_site, int bits_per_var, uint32_t *bits,
                      hipStream_t s);
// p[0..n) <- 0 by a fill kernel.
void launch_meas_clear(unsigned long long *p, int64_t n, hipStream_t s);
// agree[k] += sites where s(x) == s(x + r e_mu), pair k numbered axis after
// axis over r = 0..L_mu/2; agree must start at zero.
void launch_meas_pairs(const uint32_t *bits, const MeasGeom &m, int planes, int npairs, unsigned long long *agree,
                       hipStream_t s);
// *odd += lines along dir whose link product is -1; plane = links along dir.
int64_t meas_polyakov_items(const MeasGeom &m, int dir);
void launch_meas_polyakov(const uint32_t *plane, const MeasGeom &m, int dir, unsigned long long *odd, hipStream_t s);
// D * max line-product planes from the D link planes (max < every extent).
void launch_meas_lines(const uint32_t *links, const MeasGeom &m, int max, uint32_t *lines, hipStream_t s);
// odd[((r-1) * max + t-1) * D(D-1) + o] += r-by-t loops with product -1 in
// ordered direction pair o; 2 * max <= every extent.
void launch_meas_wilson(const uint32_t *lines, const MeasGeom &m, int max, unsigned long long *odd, hipStream_t s);

}  // namespace plq
