// plq_abi.hip -- the extern "C" boundary (include/plaquette_hip.h): chain
// objects, validation, the sweep loop and the host side of measurement.
//
// A chain owns one HIP stream and its device buffers.  Sweeps are enqueued
// asynchronously; the entry points that hand data back synchronise.  The
// sweep counter is the reference's global per-chain counter (device.rs:217,
// 393-395; gpu_cluster.rs:173, 525): every colour pass of a sweep keys its
// draws with the same value.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/plaquette_hip.h"
#include "plq_common.hpp"
#include "plq_chain.hpp"
#include "plq_internal.hpp"
#include "plq_tables.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

#define PLQ_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PLQ_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace plq {
int host_fail(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace plq

namespace {

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
uint32_t ilog2(int64_t v) {
    uint32_t r = 0;
    while ((int64_t(1) << (r + 1)) <= v) ++r;
    return r;
}

}  // namespace


namespace {

int check_device(int *dev) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(PLQ_ENODEV, "no HIP device is available (the reference's Gpu::new() would return None)");
    PLQ_HIP(hipGetDevice(dev));
    return PLQ_OK;
}

int validate_states(const uint8_t *start, size_t n, int q) {
    for (size_t i = 0; i < n; ++i)
        if (start[i] >= q) {
            char buf[128];
            snprintf(buf, sizeof buf, "state %u at index %zu is not below q = %d", start[i], i, q);
            return fail(PLQ_EINVAL, buf);
        }
    return PLQ_OK;
}

int alloc(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PLQ_ENOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " +
                                    hipGetErrorString(e));
    }
    return PLQ_OK;
}

#define PLQ_TRY(expr)           \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != PLQ_OK) {    \
            return rc_;         \
        }                       \
    } while (0)

// Build the geometry of the colour-interleaved layout.
void make_geom(plq_chain *c) {
    plq::Geom &g = c->g;
    g.D = c->D;
    g.N = c->n_sites;
    for (int mu = 0; mu < plq::MAX_D; ++mu) g.L[mu] = mu < c->D ? (int32_t)c->shape[mu] : 1;
    g.H = g.L[0] / 2;
    g.R = (int32_t)(c->n_sites / g.L[0]);
    int32_t st = 1;
    g.rstride[0] = 0;
    for (int mu = 1; mu < plq::MAX_D; ++mu) {
        g.rstride[mu] = st;
        st *= g.L[mu];
    }
}

int upload_table(plq_chain *c) {
    plq::TableParams tp{};
    tp.D = c->D;
    tp.model = c->model;
    tp.q = c->q;
    tp.kernel = c->kernel;
    tp.profile = c->profile == PLQ_PROFILE_CPU_STREAM ? PLQ_PROFILE_F64 : c->profile;
    tp.beta = c->beta;
    tp.j = c->j;
    tp.h = c->h;
    for (int a = 0; a < PLQ_MAX_Q; ++a) tp.offsets[a] = c->offsets[a];
    std::vector<uint32_t> table;
    std::string err;
    if (!plq::build_table(tp, table, err)) return fail(PLQ_EINVAL, err);
    // The uniform-parity 16-site pass (power-of-two Ising shapes, every
    // threshold >= 1); PLQ_X16U=0 keeps the chain on ising_pass_x16, which
    // tests use to check that both give identical bits.
    c->x16u = 0;
    if (c->kind == KIND_SITE && c->model == PLQ_MODEL_ISING && c->vec == 16 && !c->pmode &&
        plq::x16u_applies(c->g, table.data(), table.size(), c->pass_blocks)) {
        c->x16u = 1;
        c->ug = plq::x16u_geom(c->g);
        if (const char *e = getenv("PLQ_X16U"))
            if (atoi(e) == 0) c->x16u = 0;
    }
    if (c->d_table) (void)hipFree(c->d_table);
    c->d_table = nullptr;
    PLQ_TRY(alloc((void **)&c->d_table, table.size() * sizeof(uint32_t)));
    PLQ_HIP(hipMemcpy(c->d_table, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->table_len = table.size();
    c->h_table = table;
    return PLQ_OK;
}

int alloc_scratch(plq_chain *c) {
    PLQ_TRY(alloc((void **)&c->d_lat, (size_t)c->n_vars));
    PLQ_TRY(alloc((void **)&c->d_partials, sizeof(long long) * 2 * plq::MAX_PASS_BLOCKS * plq::REC_K));
    PLQ_TRY(alloc((void **)&c->d_meas, sizeof(long long) * plq::REC_K));
    return PLQ_OK;
}

// The CPU stream: 53-bit class thresholds, the seeded stream state and its
// jump table.
int pmode_upload_state(plq_chain *c, const uint64_t st[4]) {
    std::vector<uint64_t> jt;
    plq::pm_jump_table(st, jt);
    PLQ_HIP(hipMemcpy(c->d_pm_state, st, 4 * sizeof(uint64_t), hipMemcpyHostToDevice));
    PLQ_HIP(hipMemcpy(c->d_pm_jt, jt.data(), jt.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    return PLQ_OK;
}

int pmode_setup(plq_chain *c, uint64_t seed) {
    std::vector<uint64_t> cls;
    std::string err;
    if (!plq::pm_class_table(c->D, c->kernel, c->beta, c->j, c->h, cls, err)) return fail(PLQ_EINVAL, err);
    const int64_t nb = plq::pm_blocks(c->g);
    PLQ_TRY(alloc((void **)&c->d_pm_cls, cls.size() * sizeof(uint64_t)));
    PLQ_TRY(alloc((void **)&c->d_pm_state, 4 * sizeof(uint64_t)));
    PLQ_TRY(alloc((void **)&c->d_pm_jt, 4 * 64 * sizeof(uint64_t)));
    PLQ_TRY(alloc((void **)&c->d_pm_blk, (size_t)nb * sizeof(uint32_t)));
    PLQ_TRY(alloc((void **)&c->d_pm_off, (size_t)(nb + 1) * sizeof(uint32_t)));
    PLQ_HIP(hipMemcpy(c->d_pm_cls, cls.data(), cls.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    uint64_t st[4];
    plq::pm_seed(seed, st);
    return pmode_upload_state(c, st);
}

// ---------------------------------------------------------- local sweeps ---

// The arguments of one colour pass.  When `meas` is set the pass leaves its
// column-major partial sums in its half of d_partials (colour 0 first).
plq::SitePassArgs site_args(plq_chain *c, int color, uint32_t sweep, int mode, int meas) {
    plq::SitePassArgs a{};
    a.oth = c->col[1 - color];
    a.own = c->col[color];
    a.g = c->g;
    a.sg = c->sg;
    a.ug = c->ug;
    a.uniform = c->x16u;
    a.vec = c->vec;
    a.model = c->model == PLQ_MODEL_POTTS ? 1 : 0;
    a.mode = mode;
    a.color = color;
    a.seed = c->seed32;
    a.skey = sweep * plq::GOLDEN;
    a.q = c->q;
    a.table = c->d_table;
    a.max_blocks = c->pass_blocks;
    a.stream = c->stream;
    a.m.part = c->d_partials + (color ? plq::MAX_PASS_BLOCKS * plq::REC_K : 0);
    a.m.on = meas;
    return a;
}

// Enqueue the fold of a record left pending by the last sweep, if any.
void flush_fold(plq_chain *c) {
    const plq::RecordFold f = c->fold;
    if (!f.rec) return;
    plq::launch_finalize_columns(f.pa, f.na, f.pb, f.nb, f.nk, f.rec, c->stream);
    c->fold = plq::RecordFold{};
}

// Enqueue the two colour passes of one sweep (mode 2: measure only, no
// update); with rec non-null, fold their partial sums into rec.  A pending
// fold goes to the colour-0 pass (launch_site_pass runs it first where that
// pass cannot take it); between direct sweeps (fold_defer) this sweep's own
// fold waits for the next sweep.
void enqueue_site_passes(plq_chain *c, uint32_t sweep, int mode, long long *rec) {
    plq::SitePassArgs a0 = site_args(c, 0, sweep, mode, rec != nullptr);
    plq::SitePassArgs a1 = site_args(c, 1, sweep, mode, rec != nullptr);
    a0.m.fold = c->fold;
    c->fold = plq::RecordFold{};
    plq::launch_site_pass(a0);
    plq::launch_site_pass(a1);
    if (!rec) return;
    plq::RecordFold f{};
    f.pa = a0.m.part;
    f.na = plq::site_pass_blocks(a0);
    f.pb = a1.m.part;
    f.nb = plq::site_pass_blocks(a1);
    f.nk = 2 + c->q;
    f.rec = rec;
    c->fold = f;
    if (!c->fold_defer) flush_fold(c);
}

// Enqueue a measurement of the current site field into rec (REC_K entries).
void enqueue_site_measure(plq_chain *c, long long *rec) { enqueue_site_passes(c, 0, 2, rec); }

void enqueue_measure(plq_chain *c, long long *rec) {
    if (c->kind == KIND_SITE) {
        enqueue_site_measure(c, rec);
    } else if (c->kind == KIND_LINK) {
        plq::launch_plaquette_sum(c->d_links, c->g, c->d_partials, plq::MAX_PASS_BLOCKS, rec, c->stream);
    } else {
        plq::launch_potts_measure(c->labels, c->cg, c->q, c->d_partials, plq::MAX_PASS_BLOCKS, rec, c->stream);
    }
}

// Statistics of the last sweep's partition into dst (CLU_SLOT entries):
// dst[0..39] clusters by floor(log2 size), dst[40] clusters, dst[41] the
// largest size, dst[42 ..] the last relabel's partials summed ([0, moved,
// 0 ...]; on the box path only the moved column).  Enqueue only; allocates
// the work buffers on first use.
int ensure_stats_buffers(plq_chain *c) {
    if (!c->sizes) PLQ_TRY(alloc((void **)&c->sizes, sizeof(int32_t) * (size_t)c->n_sites));
    if (plq::cluster_boxes_x16(c->cg, c->labels)) {
        // ext stays zero between calls
        if (!c->ext) {
            PLQ_TRY(alloc((void **)&c->ext, sizeof(int32_t) * (size_t)c->n_sites));
            PLQ_HIP(hipMemsetAsync(c->ext, 0, sizeof(int32_t) * (size_t)c->n_sites, c->stream));
        }
        if (!c->d_hist_part)
            PLQ_TRY(alloc((void **)&c->d_hist_part, sizeof(unsigned long long) * 42 * 2 * plq::MAX_PASS_BLOCKS));
    }
    return PLQ_OK;
}

int enqueue_cluster_stats(plq_chain *c, unsigned long long *dst) {
    PLQ_TRY(ensure_stats_buffers(c));
    if (plq::cluster_boxes_x16(c->cg, c->labels)) {
        // sizes counted per box from the finished forest (by the sweep's
        // relabel when it was recorded); cc_hist_reduce writes every entry
        // of dst
        plq::launch_cluster_sizes_box(c->parent, c->cg, c->sizes, c->ext, c->d_cross, c->cross_sweep,
                                      c->d_hist_part, c->d_partials + plq::MAX_PASS_BLOCKS * plq::REC_K,
                                      c->last_relabel_blocks, dst, plq::MAX_PASS_BLOCKS, c->sizes_counted,
                                      c->stream);
        c->sizes_counted = false;  // the counts are consumed (ext cleared)
        return PLQ_OK;
    } else {
        PLQ_HIP(hipMemsetAsync(dst, 0, sizeof(unsigned long long) * 42, c->stream));
        if (!c->roots) PLQ_TRY(alloc((void **)&c->roots, sizeof(int32_t) * (size_t)c->n_sites));
        PLQ_HIP(hipMemsetAsync(c->sizes, 0, sizeof(int32_t) * (size_t)c->n_sites, c->stream));
        uint32_t cap = (uint32_t)(c->pass_cap > 0xffffffffull ? 0xffffffffull : c->pass_cap);
        plq::launch_cluster_resolve(c->parent, c->roots, c->cg, cap, c->d_flag, c->sizes, dst, c->stream);
    }
    plq::launch_finalize_record(c->d_partials + plq::MAX_PASS_BLOCKS * plq::REC_K, c->last_relabel_blocks, nullptr,
                                0, reinterpret_cast<long long *>(dst + 42), c->stream);
    return PLQ_OK;
}

// A cluster-statistics slot as plq_cluster_stats.
void fill_cluster_stats(const unsigned long long *h, uint64_t sweep, plq_cluster_stats *out) {
    std::memset(out, 0, sizeof *out);
    out->sweep = sweep;
    for (int k = 0; k < 40; ++k) out->size_hist[k] = (int64_t)h[k];
    out->n_clusters = (int64_t)h[40];
    out->largest = (int64_t)h[41];
    out->moved = (int64_t)h[42 + 1];
}

long long *rec_slot(plq_chain *c) {
    uint64_t slot = c->rec_count % c->rec_cap;
    c->rec_sweeps[slot] = c->sweeps_done;
    c->rec_count += 1;
    return c->d_rec + slot * plq::REC_K;
}

int enqueue_local_sweep(plq_chain *c) {
    uint32_t sweep = (uint32_t)c->sweeps_done;
    bool record = c->rec_every && ((c->sweeps_done + 1) % c->rec_every == 0);
    int mode = c->kernel == PLQ_KERNEL_METROPOLIS ? 0 : 1;
    if (c->kind == KIND_SITE && c->pmode) {
        for (int color = 0; color < 2; ++color) {
            plq::PmodeArgs a{};
            a.own = c->col[color];
            a.oth = c->col[1 - color];
            a.g = c->g;
            a.color = color;
            a.mode = mode;
            a.cls = c->d_pm_cls;
            a.blk = c->d_pm_blk;
            a.off = c->d_pm_off;
            a.state = c->d_pm_state;
            a.jt = c->d_pm_jt;
            a.stream = c->stream;
            plq::launch_pm_pass(a);
        }
        c->sweeps_done += 1;
        if (record) enqueue_site_measure(c, rec_slot(c));
    } else if (c->kind == KIND_SITE) {
        c->sweeps_done += 1;
        if (c->fused && c->x16u && !c->site_graph && c->h_table.size() == (size_t)plq::SWEEP_TH_LEN &&
            (!record || plq::x16f_tiles(c->g) <= plq::MAX_PASS_BLOCKS)) {
            // both colour passes in one launch; a pending record fold rides on
            // it, and its own record (one partial row per tile) is folded by
            // the next sweep's first kernel or at once
            plq::MeasArgs m{};
            m.part = c->d_partials;
            m.on = record ? 1 : 0;
            if (record && c->fold.rec) {
                // the pending fold reads its partials during this launch:
                // write this record's into a half it does not read
                long long *h1 = c->d_partials + plq::MAX_PASS_BLOCKS * plq::REC_K;
                auto reads = [&](const long long *h) { return c->fold.pa == h || c->fold.pb == h; };
                if (reads(m.part)) {
                    if (!reads(h1)) m.part = h1;
                    else flush_fold(c);
                }
            }
            long long *rec = record ? rec_slot(c) : nullptr;
            plq::SweepThresholds th;
            for (int i = 0; i < plq::SWEEP_TH_LEN; ++i) th.t[i] = c->h_table[(size_t)i];
            plq::launch_site_sweep_x16f(c->col[0], c->col[1], c->col_alt[0], c->col_alt[1], c->g, c->ug, mode,
                                        c->seed32, sweep * plq::GOLDEN, th, c->fold, m, c->stream);
            c->fold = plq::RecordFold{};
            std::swap(c->col[0], c->col_alt[0]);
            std::swap(c->col[1], c->col_alt[1]);
            if (rec) {
                plq::RecordFold f{};
                f.pa = m.part;
                f.na = plq::x16f_tiles(c->g);
                f.nk = 2 + c->q;
                f.rec = rec;
                c->fold = f;
                if (!c->fold_defer) flush_fold(c);
            }
        } else {
            enqueue_site_passes(c, sweep, mode, record ? rec_slot(c) : nullptr);
        }
    } else {
        for (int dir = 0; dir < c->D; ++dir)
            for (int color = 0; color < 2; ++color) {
                plq::LinkPassArgs a{};
                a.arr = c->d_links;
                a.g = c->g;
                a.dir = dir;
                a.color = color;
                a.mode = mode;
                a.seed = c->seed32;
                a.skey = sweep * plq::GOLDEN;
                a.table = c->d_table;
                a.stream = c->stream;
                plq::launch_link_pass(a);
            }
        c->sweeps_done += 1;
        if (record) enqueue_measure(c, rec_slot(c));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return PLQ_OK;
}

// Enqueue `sweeps` sweeps of a small chain, at most 4096 per launch, with
// the records that fall inside them.
int enqueue_small_sweeps(plq_chain *c, uint64_t sweeps) {
    while (sweeps > 0) {
        uint32_t n = sweeps > 4096 ? 4096u : (uint32_t)sweeps;
        plq::SmallChainArgs a{};
        a.c0 = c->col[0];
        a.c1 = c->col[1];
        a.D = c->D;
        for (int mu = 0; mu < c->D; ++mu) a.lg[mu] = (uint32_t)ilog2(c->shape[mu]);
        a.N = c->n_sites;
        a.mode = c->kernel == PLQ_KERNEL_METROPOLIS ? 0 : 1;
        a.seed = c->seed32;
        a.sweep0 = (uint32_t)c->sweeps_done;
        a.nsweeps = n;
        a.table = c->d_table;
        a.rec = c->d_rec;
        a.rec_cap = c->rec_cap;
        a.rec_first = c->rec_cap ? c->rec_count % c->rec_cap : 0;
        a.every = c->rec_every;
        a.phase = c->rec_every ? c->sweeps_done % c->rec_every : 0;
        a.stream = c->stream;
        if (c->rec_every)
            for (uint64_t s = c->sweeps_done; s < c->sweeps_done + n; ++s)
                if ((s + 1) % c->rec_every == 0) {
                    c->rec_sweeps[c->rec_count % c->rec_cap] = s + 1;
                    c->rec_count += 1;
                }
        plq::launch_small_chain(a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        c->sweeps_done += n;
        sweeps -= n;
    }
    return PLQ_OK;
}

int check_flag(plq_chain *c) {
    if (c->kind != KIND_CLUSTER) return PLQ_OK;
    int flag = 0;
    PLQ_HIP(hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    if (flag) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "cluster labelling did not converge within its budget of %llu hops on %lld sites; "
                 "the configuration is undefined",
                 (unsigned long long)c->pass_cap, (long long)c->n_sites);
        return fail(PLQ_ENOCONV, buf);
    }
    return PLQ_OK;
}

int enqueue_cluster_sweep(plq_chain *c) {
    plq::ClusterSweepArgs a{};
    a.labels = c->labels;
    a.parent = c->parent;
    a.cg = c->cg;
    a.seed32 = c->seed32;
    a.sweep = (uint32_t)c->sweeps_done;
    a.bond_t = c->bond_t;
    a.q = c->q;
    a.seeded = c->seeded;
    a.forced = c->forced;
    a.cap = (uint32_t)(c->pass_cap > 0xffffffffull ? 0xffffffffull : c->pass_cap);
    a.flag = c->d_flag;
    a.partials = c->d_partials + plq::MAX_PASS_BLOCKS * plq::REC_K;  // the measure kernels use the first half
    a.cross = c->d_cross;
    c->cross_sweep = a.sweep;
    // a recorded sweep whose cluster statistics are kept counts them in its relabel
    const bool record_next = c->rec_every && ((c->sweeps_done + 1) % c->rec_every == 0);
    a.count_sizes = false;
    c->sizes_counted = false;
    if (record_next && c->rec_clusters && plq::cluster_boxes_x16(c->cg, c->labels)) {
        PLQ_TRY(ensure_stats_buffers(c));
        a.count_sizes = true;
        a.sizes = c->sizes;
        a.ext = c->ext;
        a.hist_part = c->d_hist_part;
        c->sizes_counted = true;
    }
    a.max_blocks = plq::MAX_PASS_BLOCKS;
    a.stream = c->stream;
    // A Wolff move may grow the seed's cluster first; the decomposition then
    // runs only where the growth gave up (gate = 1).  Runs with per-sweep
    // cluster statistics or a lowered hop budget always decompose.
    const bool grow = c->wolff_mode && !a.count_sizes && !c->rec_clusters && c->pass_cap >= (uint64_t)c->n_sites + 64;
    if (grow) {
        plq::WolffGrowArgs g{};
        g.labels = c->labels;
        g.cg = c->cg;
        g.seed32 = c->seed32;
        g.sweep = a.sweep;
        g.bond_t = c->bond_t;
        g.q = c->q;
        g.forced = c->forced;
        g.queue = c->d_queue;
        g.qcap = c->wolff_qcap;
        g.level_cap = c->wolff_levels;
        g.adaptive = c->wolff_mode == 1;
        g.gain_ticks = c->wolff_gain_ticks;
        g.state = c->d_wolff;
        g.partials = a.partials;
        g.nparts = plq::cluster_relabel_blocks(a);
        g.stream = c->stream;
        plq::launch_wolff_grow(g);
        a.gate = &c->d_wolff->gate;
        a.gate_skip = 0;
    }
    // A recorded Swendsen-Wang sweep counts its record in the relabel, into
    // the slot rec_slot gives below (rec_count does not change before then).
    const bool fused = record_next && c->relabel_measures && plq::cluster_relabel_measures(a);
    if (fused) {
        a.meas_part = c->d_partials;
        a.meas_rec = c->d_rec + (c->rec_count % c->rec_cap) * plq::REC_K;
        a.defer_fold = c->fold_defer;
    }
    // a pending fold goes to this sweep's first kernel
    a.fold = c->fold;
    c->fold = plq::RecordFold{};
    c->last_relabel_blocks = plq::launch_cluster_sweep(a);
    if (fused && a.defer_fold) c->fold = plq::cluster_record_fold(a, c->last_relabel_blocks);
    c->have_partition = true;
    c->grow_pending = grow;
    c->grow_sweep = a.sweep;
    c->sweeps_done += 1;
    bool record = c->rec_every && (c->sweeps_done % c->rec_every == 0);
    if (record) {
        long long *rec = rec_slot(c);
        // between direct sweeps the measurement's fold waits for the next
        // sweep's first kernel
        if (!fused && c->fold_defer)
            plq::launch_potts_measure(c->labels, c->cg, c->q, c->d_partials, plq::MAX_PASS_BLOCKS, rec, c->stream,
                                      &c->fold);
        else if (!fused)
            enqueue_measure(c, rec);
        if (c->rec_clusters) {
            const uint64_t slot = (c->rec_count - 1) % c->rec_cap;
            PLQ_TRY(enqueue_cluster_stats(c, c->d_clu + slot * CLU_SLOT));
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return PLQ_OK;
}

// The partition of the last sweep when that sweep may have grown its
// cluster instead of decomposing the lattice: the labels it started from
// (the result with the grown cluster's old label restored) decomposed again
// with the sweep's keys into parent, all gated to run only where the growth
// succeeded (gate = 0).  Enqueue only.
int rebuild_grown_partition(plq_chain *c) {
    if (!c->grow_pending) return PLQ_OK;
    PLQ_HIP(hipMemcpyAsync(c->d_lat, c->labels, (size_t)c->n_sites, hipMemcpyDeviceToDevice, c->stream));
    plq::launch_wolff_restore(c->d_lat, c->d_queue, c->d_wolff, c->stream);
    plq::ClusterSweepArgs a{};
    a.labels = c->d_lat;
    a.parent = c->parent;
    a.cg = c->cg;
    a.seed32 = c->seed32;
    a.sweep = c->grow_sweep;
    a.bond_t = c->bond_t;
    a.q = c->q;
    a.seeded = c->seeded;
    a.forced = c->forced;
    a.cap = (uint32_t)(c->pass_cap > 0xffffffffull ? 0xffffffffull : c->pass_cap);
    a.flag = c->d_flag;
    a.cross = c->d_cross;
    a.max_blocks = plq::MAX_PASS_BLOCKS;
    a.gate = &c->d_wolff->gate;
    a.gate_skip = 1;
    a.label_only = true;
    a.stream = c->stream;
    (void)plq::launch_cluster_sweep(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    c->grow_pending = false;
    return PLQ_OK;
}

int read_lattice(plq_chain *c, uint8_t *out, size_t n) {
    if (n != (size_t)c->n_vars) return fail(PLQ_EINVAL, "output length does not match the variable count");
    if (c->kind == KIND_SITE) plq::launch_unpack_sites(c->d_lat, c->col[0], c->col[1], c->g, c->stream);
    else if (c->kind == KIND_LINK) plq::launch_unpack_links(c->d_lat, c->d_links, c->g, c->stream);
    const uint8_t *src = c->kind == KIND_CLUSTER ? c->labels : c->d_lat;
    PLQ_HIP(hipMemcpyAsync(out, src, n, hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return check_flag(c);
}

int write_lattice(plq_chain *c, const uint8_t *in, size_t n) {
    if (n != (size_t)c->n_vars) return fail(PLQ_EINVAL, "input length does not match the variable count");
    PLQ_TRY(validate_states(in, n, c->q));
    uint8_t *dst = c->kind == KIND_CLUSTER ? c->labels : c->d_lat;
    PLQ_HIP(hipMemcpyAsync(dst, in, n, hipMemcpyHostToDevice, c->stream));
    if (c->kind == KIND_SITE) plq::launch_pack_sites(c->d_lat, c->col[0], c->col[1], c->g, c->stream);
    else if (c->kind == KIND_LINK) plq::launch_pack_links(c->d_lat, c->d_links, c->g, c->stream);
    PLQ_HIP(hipStreamSynchronize(c->stream));
    c->have_partition = false;
    c->grow_pending = false;
    return PLQ_OK;
}

void fill_obs(const plq_chain *c, const long long *rec, uint64_t sweep, plq_obs *o) {
    std::memset(o, 0, sizeof *o);
    o->sweep = sweep;
    o->n_vars = c->n_vars;
    o->bond_sum = rec[0];
    for (int a = 0; a < PLQ_MAX_Q; ++a) o->counts[a] = rec[2 + a];
    if (c->kind == KIND_LINK) {
        o->energy = -c->j * (double)rec[0];  // gauge/mod.rs:275-286
    } else if (c->model == PLQ_MODEL_ISING) {
        o->spin_sum = rec[2] - rec[3];
        o->energy = -c->j * (double)o->bond_sum - c->h * (double)o->spin_sum;  // ising/mod.rs:97-117
    } else {
        double offset = 0.0;  // potts/mod.rs:253-271
        bool symmetric = true;
        for (int a = 0; a < c->q; ++a) symmetric = symmetric && c->offsets[a] == 0.0;
        if (!symmetric)
            for (int a = 0; a < c->q; ++a) offset += c->offsets[a] * (double)rec[2 + a];
        o->energy = -c->j * (double)o->bond_sum - offset;
    }
}

int max_box_sites(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 16384;
    // One int32 parent per box site; stay within 64 KiB of LDS per block so
    // the launch needs no opt-in attribute and two boxes fit one CU.
    size_t lds = prop.sharedMemPerBlock < 65536 ? prop.sharedMemPerBlock : 65536;
    int sites = 1024;
    while ((size_t)sites * 2 * sizeof(int32_t) <= lds) sites *= 2;
    return sites;
}

}  // namespace

extern "C" {

const char *plq_last_error(void) { return g_error.c_str(); }

uint32_t plq_fold_seed(uint64_t seed) { return plq::fold_seed(seed); }

int plq_device_count(int32_t *count) {
    if (!count) return fail(PLQ_EINVAL, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    if (*count <= 0) return fail(PLQ_ENODEV, "no HIP device is available");
    return PLQ_OK;
}

int plq_threshold_table(const plq_local_desc *d, uint32_t *out, size_t cap) {
    if (!d) return fail(PLQ_EINVAL, "descriptor is null");
    if (d->profile == PLQ_PROFILE_CPU_STREAM)
        return fail(PLQ_EINVAL, "the CPU-stream profile compares 53-bit draws, not against a 24-bit table");
    plq::TableParams tp{};
    tp.D = d->dim;
    tp.model = d->model;
    tp.q = d->q;
    tp.kernel = d->kernel;
    tp.profile = d->profile;
    tp.beta = d->beta;
    tp.j = d->j;
    tp.h = d->h;
    for (int a = 0; a < PLQ_MAX_Q; ++a) tp.offsets[a] = d->offsets[a];
    std::vector<uint32_t> t;
    std::string err;
    if (!plq::build_table(tp, t, err)) return fail(PLQ_EINVAL, err);
    if (out) {
        if (cap < t.size()) return fail(PLQ_EINVAL, "table buffer too small");
        std::memcpy(out, t.data(), t.size() * sizeof(uint32_t));
    }
    return (int)t.size();
}

int plq_chain_create_local(const plq_local_desc *d, const uint8_t *start, size_t n, plq_chain **out) {
    if (!d || !out) return fail(PLQ_EINVAL, "descriptor or output handle is null");
    *out = nullptr;
    bool link = d->model == PLQ_MODEL_Z2;
    if (d->model < 0 || d->model > 2) return fail(PLQ_EINVAL, "unknown model");
    if (d->kernel < 0 || d->kernel > 1) return fail(PLQ_EINVAL, "unknown kernel");
    if (d->profile < 0 || d->profile > 2) return fail(PLQ_EINVAL, "unknown threshold profile");
    if (d->profile == PLQ_PROFILE_CPU_STREAM && d->model != PLQ_MODEL_ISING)
        return fail(PLQ_EINVAL, "the CPU-stream profile is implemented for Ising site chains");
    if (d->dim < (link ? 2 : 1) || d->dim > plq::MAX_D)
        return fail(PLQ_EINVAL, link ? "the Z2 gauge device path supports dimensions 2..4"
                                     : "the site device path supports dimensions 1..4");
    if ((d->model != PLQ_MODEL_POTTS && d->q != 2) || d->q < 2 || d->q > 8)
        return fail(PLQ_EINVAL, "q must be 2 for Ising and Z2, and 2..8 for Potts");
    int64_t ns = 1;
    for (int mu = 0; mu < d->dim; ++mu) {
        if (d->shape[mu] < 2 || (d->shape[mu] & 1)) {
            return fail(PLQ_EINVAL, "the parallel checkerboard needs even lattice extents (device.rs:714-719)");
        }
        ns *= d->shape[mu];
        if (ns > (int64_t(1) << 31)) return fail(PLQ_ENOMEM, "lattice exceeds 2^31 sites");
    }
    int64_t nv = link ? ns * d->dim : ns;
    if (nv > (int64_t(1) << 32) - 1) return fail(PLQ_ENOMEM, "variable index exceeds 32 bits");
    if (!start || n != (size_t)nv) return fail(PLQ_EINVAL, "start configuration and lattice disagree on variable count");
    PLQ_TRY(validate_states(start, n, d->q));
    int dev = 0;
    PLQ_TRY(check_device(&dev));

    plq_chain *c = new plq_chain();
    c->kind = link ? KIND_LINK : KIND_SITE;
    c->device = dev;
    c->D = d->dim;
    for (int mu = 0; mu < d->dim; ++mu) c->shape[mu] = d->shape[mu];
    c->n_sites = ns;
    c->n_vars = nv;
    c->model = d->model;
    c->q = d->q;
    c->kernel = d->kernel;
    c->profile = d->profile;
    c->beta = d->beta;
    c->j = d->j;
    c->h = d->h;
    for (int a = 0; a < PLQ_MAX_Q; ++a) c->offsets[a] = d->offsets[a];
    c->seed32 = plq::fold_seed(d->seed);
    c->sweeps_between = d->sweeps_between;
    make_geom(c);
    bool all_pow2 = true;
    for (int mu = 0; mu < d->dim; ++mu) all_pow2 = all_pow2 && is_pow2(d->shape[mu]);
    // Sites per lane of the Ising kernel: 16 when L0 >= 32, 4 when L0 >= 8
    // (power-of-two extents), else the one-site-per-lane kernel.
    // PLQ_SITES_PER_LANE (1, 4 or 16) caps the choice; tests use it to check
    // that every launch shape produces identical bits.
    c->vec = 1;
    if (c->kind == KIND_SITE && d->model == PLQ_MODEL_ISING && all_pow2)
        c->vec = d->shape[0] >= 32 ? 16 : (d->shape[0] >= 8 ? 4 : 1);
    if (const char *cap = getenv("PLQ_SITES_PER_LANE")) {
        int v = atoi(cap);
        if (v == 1 || v == 4 || v == 16)
            while (c->vec > v) c->vec /= 4;
    }
    // Small power-of-two Ising lattices run many sweeps per launch in one
    // workgroup; PLQ_SMALL_CHAIN=0 keeps them on the per-pass kernels (tests
    // use it to check that both paths produce identical bits).
    c->pmode = d->profile == PLQ_PROFILE_CPU_STREAM;
    c->small = c->kind == KIND_SITE && d->model == PLQ_MODEL_ISING && all_pow2 && !c->pmode &&
               ns <= plq::small_chain_max_sites(d->dim);
    if (const char *sm = getenv("PLQ_SMALL_CHAIN"))
        if (atoi(sm) == 0) c->small = false;
    // Grid cap of a colour pass (8..2048 blocks of 256 lanes); lanes walk
    // their block's span, so a smaller grid gives each lane several chunks
    // to pipeline.  PLQ_PASS_BLOCKS overrides the default.
    if (const char *pb = getenv("PLQ_PASS_BLOCKS")) {
        long v = atol(pb);
        if (v >= 8 && v <= plq::MAX_PASS_BLOCKS) c->pass_blocks = v;
    }
    // PLQ_SITE_GRAPH=1 replays a uniform-pass chain's sweeps from captured
    // graphs.  Off by default: measured on MI355X, a pass inside a replay
    // takes 8.0 us against 6.8 us launched directly (256^3), so replays lose
    // about 6% of the whole job.  Tests compare the two paths.
    if (const char *sg = getenv("PLQ_SITE_GRAPH")) c->site_graph = atoi(sg) != 0 ? 1 : 0;
    if (c->vec > 1) {
        c->sg.lh4 = ilog2(d->shape[0] / 8);
        c->sg.L1 = d->dim > 1 ? (uint32_t)d->shape[1] : 1;
        c->sg.L2 = d->dim > 2 ? (uint32_t)d->shape[2] : 1;
        c->sg.L3 = d->dim > 3 ? (uint32_t)d->shape[3] : 1;
        c->sg.ll1 = ilog2(c->sg.L1);
        c->sg.ll2 = ilog2(c->sg.L2);
    }
    int rc = PLQ_OK;
    auto bail = [&](int code) {
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PLQ_EDEVICE, "hipStreamCreate failed"));
    if ((rc = upload_table(c)) != PLQ_OK) return bail(rc);
    if ((rc = alloc_scratch(c)) != PLQ_OK) return bail(rc);
    if (c->pmode && (rc = pmode_setup(c, d->seed)) != PLQ_OK) return bail(rc);
    size_t half = (size_t)(ns / 2) + 16;
    if (c->kind == KIND_SITE) {
        for (int k = 0; k < 2; ++k)
            if ((rc = alloc((void **)&c->col[k], half)) != PLQ_OK) return bail(rc);
        // The fused two-colour sweep (3D Ising, L0 = 256): on where its
        // 16 x 16-row tiles give a workgroup to every CU (at least 256
        // tiles); PLQ_FUSED_SWEEP=1 wherever its shape applies, 0 off.
        if (c->model == PLQ_MODEL_ISING && !c->pmode && c->vec == 16 && plq::x16f_applies(c->g, 1)) {
            int on = plq::x16f_tiles(c->g) >= 256 ? 1 : 0;
            if (const char *e = getenv("PLQ_FUSED_SWEEP")) on = atoi(e) != 0 ? 1 : 0;
            if (on) {
                for (int k = 0; k < 2; ++k)
                    if ((rc = alloc((void **)&c->col_alt[k], half)) != PLQ_OK) return bail(rc);
                c->fused = true;
            }
        }
    } else {
        for (int k = 0; k < 2 * d->dim; ++k)
            if ((rc = alloc((void **)&c->links[k], half)) != PLQ_OK) return bail(rc);
        if ((rc = alloc((void **)&c->d_links, sizeof(uint8_t *) * 2 * plq::MAX_D)) != PLQ_OK) return bail(rc);
        if (hipMemcpy(c->d_links, c->links, sizeof(uint8_t *) * 2 * plq::MAX_D, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(PLQ_EDEVICE, "hipMemcpy of link pointers failed"));
    }
    if ((rc = write_lattice(c, start, n)) != PLQ_OK) return bail(rc);
    *out = c;
    return PLQ_OK;
}

int plq_chain_create_cluster(const plq_cluster_desc *d, const uint8_t *start, size_t n, plq_chain **out) {
    if (!d || !out) return fail(PLQ_EINVAL, "descriptor or output handle is null");
    *out = nullptr;
    if (d->dim < 1 || d->dim > plq::MAX_D) return fail(PLQ_EINVAL, "the cluster device path supports dimensions 1..4");
    if (d->q < 2 || d->q > PLQ_MAX_Q) return fail(PLQ_EINVAL, "a cluster redraw needs 2 <= q <= 16 states");
    if (d->extent < 0 || d->extent > 1 || d->relabel < 0 || d->relabel > 1)
        return fail(PLQ_EINVAL, "unknown extent or relabel rule");
    if (d->q == 2 && d->extent == PLQ_EXTENT_ALL && d->relabel == PLQ_RELABEL_FORCED_CHANGE)
        return fail(PLQ_EINVAL,
                    "at q = 2 a forced change of every cluster is the deterministic global flip "
                    "(updater.rs:447-478)");
    if (!std::isfinite(d->gap) || d->gap < 0.0)
        return fail(PLQ_EINVAL, "the bond probability needs a finite, non-negative gap");
    if (!std::isfinite(d->beta)) return fail(PLQ_EINVAL, "beta must be finite");
    if (d->profile < 0 || d->profile > 1) return fail(PLQ_EINVAL, "unknown threshold profile");
    int64_t ns = 1;
    for (int mu = 0; mu < d->dim; ++mu) {
        if (d->shape[mu] < 1) return fail(PLQ_EINVAL, "every extent must be positive");
        ns *= d->shape[mu];
        if (ns > (int64_t(1) << 31) - 1) return fail(PLQ_ENOMEM, "lattice exceeds 2^31 - 1 sites");
    }
    if (ns * d->dim > (int64_t(1) << 32) - 1) return fail(PLQ_ENOMEM, "bond index exceeds 32 bits");
    if (!start || n != (size_t)ns) return fail(PLQ_EINVAL, "start configuration and lattice disagree on site count");
    PLQ_TRY(validate_states(start, n, d->q));
    int dev = 0;
    PLQ_TRY(check_device(&dev));

    plq_chain *c = new plq_chain();
    c->kind = KIND_CLUSTER;
    c->device = dev;
    c->D = d->dim;
    for (int mu = 0; mu < d->dim; ++mu) c->shape[mu] = d->shape[mu];
    c->n_sites = ns;
    c->n_vars = ns;
    c->model = PLQ_MODEL_POTTS;
    c->q = d->q;
    c->profile = d->profile;
    c->beta = d->beta;
    c->j = 1.0;
    c->seed32 = plq::fold_seed(d->seed);
    c->sweeps_between = d->sweeps_between;
    c->seeded = d->extent == PLQ_EXTENT_SEEDED;
    c->forced = d->relabel == PLQ_RELABEL_FORCED_CHANGE;
    c->bond_t = plq::bond_threshold(d->beta, d->gap, d->profile);
    c->cg = plq::make_ccgeom(d->dim, d->shape, max_box_sites(dev));
    // The reference's cap counts propagation passes (gpu_cluster.rs:620-625);
    // here every find and union walks strictly decreasing indices, so the
    // site count bounds any walk.
    c->pass_cap = (uint64_t)ns + 64;
    // Small lattices run whole cluster sweeps in one workgroup;
    // PLQ_SMALL_CLUSTER=0 keeps them on the four-kernel sweep (tests compare
    // the two).
    c->small = ns <= plq::small_cluster_max_sites(d->dim);
    if (const char *sm = getenv("PLQ_SMALL_CLUSTER"))
        if (atoi(sm) == 0) c->small = false;
    // PLQ_RELABEL_MEASURE=0 records SW sweeps with a separate measurement
    // pass (tests compare the two).
    if (const char *rm = getenv("PLQ_RELABEL_MEASURE"))
        if (atoi(rm) == 0) c->relabel_measures = false;
    // Wolff moves grown from the seed (plq_wolff.hip): adaptive by default
    // on lattices of 2^22 sites or more, where a decomposition costs several
    // times the fixed cost of a grown move; PLQ_WOLFF_GROW=0/1/2 (off,
    // adaptive, every sweep), PLQ_WOLFF_LEVELS and PLQ_WOLFF_QUEUE set the
    // mode and the budgets (tests drive both paths).
    if (c->seeded && !c->small) {
        c->wolff_mode = ns >= (int64_t(1) << 22) ? 1 : 0;
        if (const char *g = getenv("PLQ_WOLFF_GROW")) c->wolff_mode = std::max(0, std::min(2, atoi(g)));
        // Measured on MI355X (2D Ising, 1024^2 to 4096^2): a Wolff sweep by
        // the decomposition costs about 25 us + 7.5 ps per site, a level of
        // growth about 1.2 us, and a grown sweep's gated launches about
        // 10 us.  The level budget keeps a failed attempt near half a
        // decomposition.
        const double full_us = 25.0 + 7.5e-6 * (double)ns;
        c->wolff_levels = (uint32_t)std::max(8.0, std::min(128.0, 0.5 * full_us / 1.2));
        int rate_khz = 0;
        if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || rate_khz <= 0)
            rate_khz = 100000;
        c->wolff_gain_ticks = (float)((full_us - 10.0) * 1e-3 * (double)rate_khz);
        c->wolff_qcap = (uint32_t)std::min<int64_t>(ns, plq::WOLFF_QUEUE);
        if (const char *v = getenv("PLQ_WOLFF_LEVELS")) c->wolff_levels = (uint32_t)std::max(1, atoi(v));
        if (const char *v = getenv("PLQ_WOLFF_QUEUE"))
            c->wolff_qcap = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(c->wolff_qcap, atoll(v)));
    }
    int rc = PLQ_OK;
    auto bail = [&](int code) {
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PLQ_EDEVICE, "hipStreamCreate failed"));
    if ((rc = alloc_scratch(c)) != PLQ_OK) return bail(rc);
    if (c->wolff_mode) {
        if ((rc = alloc((void **)&c->d_wolff, sizeof(plq::WolffState))) != PLQ_OK ||
            (rc = alloc((void **)&c->d_queue, sizeof(int32_t) * (size_t)c->wolff_qcap)) != PLQ_OK)
            return bail(rc);
        plq::WolffState st0{};
        st0.gate = 1;
        st0.success = 0.5f;
        if (hipMemcpy(c->d_wolff, &st0, sizeof st0, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(PLQ_EDEVICE, "Wolff state initialisation failed"));
    }
    if ((rc = alloc((void **)&c->labels, (size_t)ns)) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&c->parent, (size_t)ns * sizeof(int32_t))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&c->d_flag, sizeof(int))) != PLQ_OK) return bail(rc);
    if (plq::cluster_boxes_x16(c->cg, c->labels)) {
        // crossing marks of cc_compress, zero (no sweep's tag) to start
        if ((rc = alloc((void **)&c->d_cross, (size_t)ns)) != PLQ_OK) return bail(rc);
        if (hipMemset(c->d_cross, 0, (size_t)ns) != hipSuccess) return bail(fail(PLQ_EDEVICE, "hipMemset failed"));
    }
    if (hipMemset(c->d_flag, 0, sizeof(int)) != hipSuccess)
        return bail(fail(PLQ_EDEVICE, "hipMemset failed"));
    if ((rc = write_lattice(c, start, n)) != PLQ_OK) return bail(rc);
    *out = c;
    return PLQ_OK;
}

int plq_chain_pcg_state(plq_chain *c, uint64_t *out4) {
    if (!c || !out4) return fail(PLQ_EINVAL, "chain or buffer is null");
    if (!c->pmode) return fail(PLQ_EINVAL, "only PLQ_PROFILE_CPU_STREAM chains have a stream state");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipMemcpyAsync(out4, c->d_pm_state, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return PLQ_OK;
}

int plq_chain_set_pcg_state(plq_chain *c, const uint64_t *in4) {
    if (!c || !in4) return fail(PLQ_EINVAL, "chain or buffer is null");
    if (!c->pmode) return fail(PLQ_EINVAL, "only PLQ_PROFILE_CPU_STREAM chains have a stream state");
    if ((in4[2] & 1u) == 0) return fail(PLQ_EINVAL, "the increment of an Lcg128Xsl64 stream is odd");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return pmode_upload_state(c, in4);
}

int plq_chain_destroy(plq_chain *c) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    delete c;
    return PLQ_OK;
}

namespace {

// Cluster sweeps of a small lattice, up to 4,096 per launch of one workgroup
// (plq_cluster_small.hip), with the record-ring bookkeeping of
// enqueue_small_sweeps.
int enqueue_small_cluster_sweeps(plq_chain *c, uint64_t sweeps) {
    while (sweeps > 0) {
        const uint32_t n = sweeps > 4096 ? 4096u : (uint32_t)sweeps;
        plq::SmallClusterArgs a{};
        a.labels = c->labels;
        a.parent = c->parent;
        a.partials = c->d_partials + plq::MAX_PASS_BLOCKS * plq::REC_K;
        a.D = c->D;
        for (int mu = 0; mu < plq::MAX_D; ++mu) a.L[mu] = mu < c->D ? (int32_t)c->shape[mu] : 1;
        a.N = (uint32_t)c->n_sites;
        a.q = c->q;
        a.seeded = c->seeded;
        a.forced = c->forced;
        a.bond_seed = c->seed32 ^ plq::BOND_KEY;
        a.root_seed = c->seed32 ^ plq::ROOT_SALT;
        a.seed_seed = c->seed32 ^ plq::SEED_SALT;
        a.bond_t = c->bond_t;
        a.sweep0 = (uint32_t)c->sweeps_done;
        a.nsweeps = n;
        a.rec = c->d_rec;
        a.rec_cap = c->rec_cap;
        a.rec_first = c->rec_cap ? c->rec_count % c->rec_cap : 0;
        a.every = c->rec_every;
        a.phase = c->rec_every ? c->sweeps_done % c->rec_every : 0;
        a.stream = c->stream;
        if (c->rec_every)
            for (uint64_t s = c->sweeps_done; s < c->sweeps_done + n; ++s)
                if ((s + 1) % c->rec_every == 0) {
                    c->rec_sweeps[c->rec_count % c->rec_cap] = s + 1;
                    c->rec_count += 1;
                }
        plq::launch_small_cluster(a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        c->sweeps_done += n;
        sweeps -= n;
    }
    c->last_relabel_blocks = 1;
    c->have_partition = true;
    return PLQ_OK;
}

// Whether the chain's sweeps can be replayed from a captured graph: a site
// chain on the uniform pass (the sweep key of that pass can come from a
// device counter), recording every sweep up to every 64th or not at all.
bool site_graph_ok(const plq_chain *c) {
    return c->kind == KIND_SITE && c->x16u && !c->small && !c->pmode && c->site_graph && c->rec_every <= 64;
}

void drop_site_graph(plq_chain *c) {
    if (c->site_exec) (void)hipGraphExecDestroy(c->site_exec);
    c->site_exec = nullptr;
    c->exec_sweeps = 0;
}

// Capture `B` sweeps of the two colour passes, keyed by the device counter,
// the record of the last one when recording, and the counters' tick.
int build_site_graph(plq_chain *c, uint32_t B, bool records) {
    drop_site_graph(c);
    if (!c->d_ctr) PLQ_TRY(alloc((void **)&c->d_ctr, 2 * sizeof(uint32_t)));
    const int mode = c->kernel == PLQ_KERNEL_METROPOLIS ? 0 : 1;
    hipGraph_t g = nullptr;
    PLQ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t k = 0; k < B; ++k) {
        const bool rec = records && k + 1 == B;
        plq::SitePassArgs a0 = site_args(c, 0, 0, mode, rec);
        plq::SitePassArgs a1 = site_args(c, 1, 0, mode, rec);
        a0.ctr = a1.ctr = c->d_ctr;
        a0.sweep_off = a1.sweep_off = k;
        plq::launch_site_pass(a0);
        plq::launch_site_pass(a1);
        if (rec)
            plq::launch_finalize_columns_ring(a0.m.part, plq::site_pass_blocks(a0), a1.m.part,
                                              plq::site_pass_blocks(a1), 2 + c->q, c->d_rec, c->d_ctr, c->stream);
    }
    plq::launch_site_ctr_tick(c->d_ctr, B, records ? 1u : 0u, records ? (uint32_t)c->rec_cap : 1u, c->stream);
    hipError_t err = hipStreamEndCapture(c->stream, &g);
    if (err == hipSuccess) err = hipGraphInstantiate(&c->site_exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (err != hipSuccess) {
        c->site_exec = nullptr;
        (void)hipGetLastError();
        return fail(PLQ_EDEVICE, std::string("site sweep graph capture failed: ") + hipGetErrorString(err));
    }
    c->exec_sweeps = B;
    c->exec_records = records;
    return PLQ_OK;
}

// Site sweeps in graph replays of `B` sweeps (B = the record interval, or 16
// without records) where a replay fits the record phase, and one by one
// around them.  The counters on the device are set from the host before the
// first replay after sweeps enqueued directly.
int enqueue_site_graph_sweeps(plq_chain *c, uint64_t sweeps) {
    const bool records = c->rec_every != 0;
    const uint32_t B = records ? (uint32_t)c->rec_every : 16u;
    while (sweeps > 0) {
        const uint64_t into = records ? c->sweeps_done % c->rec_every : 0;
        if (into == 0 && sweeps >= B && c->site_graph) {
            if (!c->site_exec || c->exec_sweeps != B || c->exec_records != records) {
                if (build_site_graph(c, B, records) != PLQ_OK) {
                    c->site_graph = 0;  // replays off for this chain; sweeps go on directly
                    continue;
                }
            }
            if (!c->ctr_valid) {
                const uint32_t slot = records ? (uint32_t)(c->rec_count % c->rec_cap) : 0u;
                plq::launch_site_ctr_set(c->d_ctr, (uint32_t)c->sweeps_done, slot, c->stream);
                c->ctr_valid = true;
            }
            PLQ_HIP(hipGraphLaunch(c->site_exec, c->stream));
            c->sweeps_done += B;
            if (records) {
                c->rec_sweeps[c->rec_count % c->rec_cap] = c->sweeps_done;
                c->rec_count += 1;
            }
            sweeps -= B;
            continue;
        }
        PLQ_TRY(enqueue_local_sweep(c));
        c->ctr_valid = false;
        sweeps -= 1;
    }
    return PLQ_OK;
}

// Enqueue `sweeps` sweeps of any chain kind (no synchronisation).
int enqueue_sweeps(plq_chain *c, uint64_t sweeps) {
    if (c->kind == KIND_SITE && c->small) return enqueue_small_sweeps(c, sweeps);
    if (site_graph_ok(c)) return enqueue_site_graph_sweeps(c, sweeps);
    // The one-workgroup cluster sweep cannot overrun the default hop budget
    // (every walk is shorter than the site count); a lowered budget takes the
    // four-kernel path, which counts hops.
    if (c->kind == KIND_CLUSTER && c->small && !c->rec_clusters && c->pass_cap >= (uint64_t)c->n_sites + 64 &&
        sweeps > 0)
        return enqueue_small_cluster_sweeps(c, sweeps);
    // Direct sweeps: a sweep's record is folded by the next sweep's first
    // kernel, and the last one's after the loop.
    c->fold_defer = true;
    int rc = PLQ_OK;
    for (uint64_t t = 0; t < sweeps && rc == PLQ_OK; ++t)
        rc = c->kind == KIND_CLUSTER ? enqueue_cluster_sweep(c) : enqueue_local_sweep(c);
    c->fold_defer = false;
    flush_fold(c);
    return rc;
}

}  // namespace

int plq_chain_advance(plq_chain *c, uint64_t sweeps) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    // Enqueue only, for cluster chains too: a labelling that overran its hop
    // budget sets a sticky device flag, which the next call that waits for
    // the chain's stream (next, read, measure, series, sync, ...) reports.
    return enqueue_sweeps(c, sweeps);
}

int plq_chain_next_batch(plq_chain *c, uint64_t count, uint8_t *out, size_t n) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    if (count == 0) return PLQ_OK;
    if (!out) return fail(PLQ_EINVAL, "output buffer is null");
    const size_t nv = (size_t)c->n_vars;
    if (count > SIZE_MAX / nv || n != (size_t)count * nv)
        return fail(PLQ_EINVAL, "the output holds count * n_vars states");
    PLQ_HIP(hipSetDevice(c->device));
    const size_t bytes = (size_t)count * nv;
    if (bytes > c->batch_cap) {
        PLQ_HIP(hipStreamSynchronize(c->stream));
        if (c->d_batch) (void)hipFree(c->d_batch);
        c->d_batch = nullptr;
        c->batch_cap = 0;
        PLQ_TRY(alloc((void **)&c->d_batch, bytes));
        c->batch_cap = bytes;
    }
    // device.rs:431-488: every sample after sweeps_between sweeps, copied into
    // its slot of the staging buffer, and the batch read back in one transfer.
    for (uint64_t k = 0; k < count; ++k) {
        PLQ_TRY(enqueue_sweeps(c, c->sweeps_between));
        uint8_t *slot = c->d_batch + k * nv;
        if (c->kind == KIND_SITE) plq::launch_unpack_sites(slot, c->col[0], c->col[1], c->g, c->stream);
        else if (c->kind == KIND_LINK) plq::launch_unpack_links(slot, c->d_links, c->g, c->stream);
        else PLQ_HIP(hipMemcpyAsync(slot, c->labels, nv, hipMemcpyDeviceToDevice, c->stream));
    }
    PLQ_HIP(hipMemcpyAsync(out, c->d_batch, bytes, hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return check_flag(c);
}

int plq_chain_next(plq_chain *c, uint8_t *out, size_t n) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    PLQ_TRY(plq_chain_advance(c, c->sweeps_between));
    std::lock_guard<std::mutex> lk(c->mu);
    return read_lattice(c, out, n);
}

int plq_hot_start(uint64_t seed, int32_t q, uint8_t *out, size_t n, uint64_t *state4) {
    if (!out && n) return fail(PLQ_EINVAL, "output buffer is null");
    if (q < 2 || q > PLQ_MAX_Q) return fail(PLQ_EINVAL, "a hot start needs 2 <= q <= 16 states");
    uint64_t st[4];
    plq::pm_hot_start(seed, (uint32_t)q, out, n, st);
    if (state4) std::memcpy(state4, st, sizeof st);
    return PLQ_OK;
}

int plq_chain_read(plq_chain *c, uint8_t *out, size_t n) {
    if (!c || !out) return fail(PLQ_EINVAL, "chain or buffer is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    return read_lattice(c, out, n);
}

int plq_chain_write(plq_chain *c, const uint8_t *in, size_t n) {
    if (!c || !in) return fail(PLQ_EINVAL, "chain or buffer is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    return write_lattice(c, in, n);
}

int plq_chain_measure(plq_chain *c, plq_obs *out) {
    if (!c || !out) return fail(PLQ_EINVAL, "chain or output is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    enqueue_measure(c, c->d_meas);
    long long rec[plq::REC_K];
    PLQ_HIP(hipMemcpyAsync(rec, c->d_meas, sizeof rec, hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    fill_obs(c, rec, c->sweeps_done, out);
    return check_flag(c);
}

int plq_chain_record(plq_chain *c, uint64_t every, uint64_t capacity) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    drop_site_graph(c);  // a replay holds the record ring's address and capacity
    c->ctr_valid = false;
    if (c->d_rec) (void)hipFree(c->d_rec);
    if (c->d_clu) (void)hipFree(c->d_clu);
    c->d_rec = nullptr;
    c->d_clu = nullptr;
    c->rec_every = 0;
    c->rec_cap = 0;
    c->rec_count = 0;
    c->rec_sweeps.clear();
    if (every == 0) return PLQ_OK;
    if (capacity == 0) return fail(PLQ_EINVAL, "recording needs a positive capacity");
    PLQ_TRY(alloc((void **)&c->d_rec, sizeof(long long) * plq::REC_K * capacity));
    if (c->rec_clusters) PLQ_TRY(alloc((void **)&c->d_clu, sizeof(unsigned long long) * CLU_SLOT * capacity));
    c->clu_from = 0;
    c->rec_every = every;
    c->rec_cap = capacity;
    c->rec_sweeps.assign(capacity, 0);
    return PLQ_OK;
}

int plq_chain_series(plq_chain *c, plq_obs *out, uint64_t capacity, uint64_t *n_out) {
    if (!c || !n_out) return fail(PLQ_EINVAL, "chain or count is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    uint64_t have = c->rec_count < c->rec_cap ? c->rec_count : c->rec_cap;
    *n_out = have;
    if (!out || have == 0) return PLQ_OK;
    if (capacity < have) return fail(PLQ_EINVAL, "series buffer too small");
    std::vector<long long> host(have * plq::REC_K > 0 ? c->rec_cap * plq::REC_K : 1);
    PLQ_HIP(hipMemcpyAsync(host.data(), c->d_rec, sizeof(long long) * plq::REC_K * c->rec_cap,
                           hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    uint64_t first = c->rec_count - have;
    for (uint64_t k = 0; k < have; ++k) {
        uint64_t slot = (first + k) % c->rec_cap;
        fill_obs(c, &host[slot * plq::REC_K], c->rec_sweeps[slot], &out[k]);
    }
    return check_flag(c);
}

int plq_chain_record_clusters(plq_chain *c, int on) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    if (c->kind != KIND_CLUSTER) return fail(PLQ_EINVAL, "not a cluster chain");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    c->rec_clusters = on != 0;
    c->clu_from = c->rec_count;
    if (c->d_clu) (void)hipFree(c->d_clu);
    c->d_clu = nullptr;
    if (c->rec_clusters && c->rec_cap)
        PLQ_TRY(alloc((void **)&c->d_clu, sizeof(unsigned long long) * CLU_SLOT * c->rec_cap));
    return PLQ_OK;
}

int plq_chain_cluster_series(plq_chain *c, plq_cluster_stats *out, uint64_t capacity, uint64_t *n_out) {
    if (!c || !n_out) return fail(PLQ_EINVAL, "chain or count is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    const uint64_t kept = c->rec_count - c->clu_from;
    uint64_t have = c->rec_clusters && c->d_clu ? (kept < c->rec_cap ? kept : c->rec_cap) : 0;
    *n_out = have;
    if (!out || have == 0) return PLQ_OK;
    if (capacity < have) return fail(PLQ_EINVAL, "series buffer too small");
    std::vector<unsigned long long> host(c->rec_cap * CLU_SLOT);
    PLQ_HIP(hipMemcpyAsync(host.data(), c->d_clu, sizeof(unsigned long long) * host.size(), hipMemcpyDeviceToHost,
                           c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    uint64_t first = c->rec_count - have;
    for (uint64_t k = 0; k < have; ++k) {
        uint64_t slot = (first + k) % c->rec_cap;
        fill_cluster_stats(&host[slot * CLU_SLOT], c->rec_sweeps[slot], &out[k]);
    }
    return check_flag(c);
}

int plq_chain_sweeps_done(plq_chain *c, uint64_t *out) {
    if (!c || !out) return fail(PLQ_EINVAL, "chain or output is null");
    *out = c->sweeps_done;
    return PLQ_OK;
}

int plq_chain_set_sweep(plq_chain *c, uint64_t sweep) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    c->sweeps_done = sweep;
    c->ctr_valid = false;
    return PLQ_OK;
}

int plq_chain_sync(plq_chain *c) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return check_flag(c);
}

int plq_chain_stream(plq_chain *c, void **stream) {
    if (!c || !stream) return fail(PLQ_EINVAL, "chain or output is null");
    *stream = (void *)c->stream;
    return PLQ_OK;
}

int plq_chain_set_pass_cap(plq_chain *c, uint64_t cap) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    if (c->kind != KIND_CLUSTER) return fail(PLQ_EINVAL, "only cluster chains have a labelling budget");
    std::lock_guard<std::mutex> lk(c->mu);
    c->pass_cap = cap;
    return PLQ_OK;
}

int plq_chain_wolff_counts(plq_chain *c, uint64_t *grown, uint64_t *full) {
    if (!c || !grown || !full) return fail(PLQ_EINVAL, "chain or output is null");
    if (c->kind != KIND_CLUSTER || !c->seeded) return fail(PLQ_EINVAL, "not a Wolff chain");
    std::lock_guard<std::mutex> lk(c->mu);
    *grown = 0;
    *full = 0;
    if (!c->d_wolff) return PLQ_OK;
    PLQ_HIP(hipSetDevice(c->device));
    plq::WolffState st{};
    PLQ_HIP(hipMemcpyAsync(&st, c->d_wolff, sizeof st, hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    *grown = st.grown;
    *full = st.full;
    return check_flag(c);
}

int plq_chain_cluster_roots(plq_chain *c, int64_t *roots, size_t n) {
    if (!c || !roots) return fail(PLQ_EINVAL, "chain or output is null");
    if (c->kind != KIND_CLUSTER) return fail(PLQ_EINVAL, "not a cluster chain");
    if (n != (size_t)c->n_sites) return fail(PLQ_EINVAL, "output length does not match the site count");
    if (!c->have_partition) return fail(PLQ_EINVAL, "no cluster sweep has run since the last write");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    if (!c->roots) PLQ_TRY(alloc((void **)&c->roots, sizeof(int32_t) * (size_t)c->n_sites));
    PLQ_TRY(rebuild_grown_partition(c));
    uint32_t cap = (uint32_t)(c->pass_cap > 0xffffffffull ? 0xffffffffull : c->pass_cap);
    plq::launch_cluster_resolve(c->parent, c->roots, c->cg, cap, c->d_flag, nullptr, nullptr, c->stream);
    std::vector<int32_t> host((size_t)c->n_sites);
    PLQ_HIP(hipMemcpyAsync(host.data(), c->roots, sizeof(int32_t) * host.size(), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < host.size(); ++i) roots[i] = host[i];
    return check_flag(c);
}

int plq_chain_cluster_stats(plq_chain *c, plq_cluster_stats *out) {
    if (!c || !out) return fail(PLQ_EINVAL, "chain or output is null");
    if (c->kind != KIND_CLUSTER) return fail(PLQ_EINVAL, "not a cluster chain");
    if (!c->have_partition) return fail(PLQ_EINVAL, "no cluster sweep has run since the last write");
    std::lock_guard<std::mutex> lk(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    if (!c->sizes) PLQ_TRY(alloc((void **)&c->sizes, sizeof(int32_t) * (size_t)c->n_sites));
    if (!c->hist) PLQ_TRY(alloc((void **)&c->hist, sizeof(unsigned long long) * CLU_SLOT));
    PLQ_TRY(rebuild_grown_partition(c));
    PLQ_TRY(enqueue_cluster_stats(c, c->hist));
    unsigned long long h[CLU_SLOT];
    PLQ_HIP(hipMemcpyAsync(h, c->hist, sizeof h, hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    fill_cluster_stats(h, c->sweeps_done, out);
    return check_flag(c);
}

}  // extern "C"
