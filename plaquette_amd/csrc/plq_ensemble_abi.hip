// plq_ensemble_abi.hip -- host side of batched Z2 ensembles: validation,
// layout, threshold tables, replica-exchange tables, the sweep-block
// schedule (direct launches or a captured hipGraph) and the launch-shape
// autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plaquette_hip.h"
#include "plq_common.hpp"
#include "plq_ensemble.hpp"
#include "plq_internal.hpp"
#include "plq_tables.hpp"

namespace {

constexpr uint32_t EXCHANGE_SALT = 0x68e31da4u;

#define ENS_HIP(expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return plq::host_fail(PLQ_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_));     \
    } while (0)

#define ENS_TRY(expr)        \
    do {                     \
        int rc_ = (expr);    \
        if (rc_ != PLQ_OK) { \
            return rc_;      \
        }                    \
    } while (0)

uint64_t magic40(uint32_t d) { return ((uint64_t(1) << 40) + d - 1) / d; }

int dalloc(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        return plq::host_fail(PLQ_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return PLQ_OK;
}

}  // namespace

struct plq_ensemble {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 0, nrep = 0, kernel = 0, profile = 0;
    double j = 1.0;
    std::vector<plq::EnsReplica> reps;  // per replica geometry and seed (host side)
    std::vector<double> betas;
    std::vector<size_t> lat_off;  // per replica offset in the lattice-order concatenation
    size_t lat_bytes = 0;
    std::vector<plq::EnsQuad> quads;  // replicas of one shape, four to a word
    std::vector<int> rep_quad, rep_byte;
    plq::EnsQuad *d_quads = nullptr;
    uint16_t *d_block_rep = nullptr;  // quad of each pass block
    size_t block_rep_cap = 0;
    uint32_t *d_words = nullptr;
    size_t nwords = 0;
    uint8_t *d_lat = nullptr;
    uint32_t *d_tables = nullptr;
    uint32_t *d_slot_of = nullptr;
    uint32_t *d_qtab = nullptr;  // the fast pass's thresholds per quad (plq::EnsQtab)
    uint32_t *d_rep_at = nullptr;
    std::vector<plq::EnsGroupPair> pairs;
    plq::EnsGroupPair *d_pairs = nullptr;
    uint32_t *d_ex_tables = nullptr;
    unsigned long long *d_sums = nullptr;
    unsigned long long *d_accepted = nullptr;
    uint32_t *d_counters = nullptr;  // sweep base, round, records
    long long *d_ring_sums = nullptr;
    uint32_t *d_ring_slots = nullptr;
    uint64_t rec_cap = 1;
    uint64_t sweeps_done = 0, rounds_done = 0, records_done = 0;
    uint64_t measure_every = 0;
    int exchange = 0;
    uint32_t ex_seed32 = 0;
    int block_size = 256, per_lane = 1, use_graph = 1;
    int fast = 0;  // ens_link_pass2 (every threshold >= 1)
    uint32_t total_blocks = 0;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_sweeps = 0;
    bool graph_measures = false;

    ~plq_ensemble() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (void *p : {(void *)d_quads, (void *)d_block_rep, (void *)d_words, (void *)d_lat, (void *)d_tables, (void *)d_slot_of,
                        (void *)d_rep_at, (void *)d_pairs, (void *)d_ex_tables, (void *)d_sums, (void *)d_accepted,
                        (void *)d_counters, (void *)d_ring_sums, (void *)d_ring_slots, (void *)d_qtab})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Blocks of each replica for the current launch shape; uploads the replica table.
plq::EnsQtab qtab_of(const plq_ensemble *e) {
    plq::EnsQtab qt{};
    qt.quads = e->d_quads;
    qt.nquads = (int)e->quads.size();
    qt.tables = e->d_tables;
    qt.qtab = e->d_qtab;
    qt.TL = 2 * (2 * e->D - 1);
    return qt;
}

int layout_blocks(plq_ensemble *e) {
    uint32_t b = 0;
    uint32_t per_block = (uint32_t)e->block_size * (uint32_t)e->per_lane;
    for (auto &Q : e->quads) {
        Q.block0 = b;
        uint32_t half = Q.nsites / 2;
        b += (half + per_block - 1) / per_block;
    }
    // Pad the grid to a multiple of 8 so the kernels can give each XCD a
    // contiguous run of quads; padded blocks map to no quad (0xffff).
    const uint32_t used = b;
    b = (b + 7u) & ~7u;
    e->total_blocks = b;
    const int nq = (int)e->quads.size();
    std::vector<uint16_t> block_rep(b, (uint16_t)0xffff);
    for (int q = 0; q < nq; ++q) {
        uint32_t end = q + 1 < nq ? e->quads[q + 1].block0 : used;
        for (uint32_t k = e->quads[q].block0; k < end; ++k) block_rep[k] = (uint16_t)q;
    }
    if (e->block_rep_cap < b) {
        if (e->d_block_rep) (void)hipFree(e->d_block_rep);
        e->d_block_rep = nullptr;
        e->block_rep_cap = 0;
        ENS_TRY(dalloc((void **)&e->d_block_rep, sizeof(uint16_t) * b));
        e->block_rep_cap = b;
    }
    ENS_HIP(hipMemcpy(e->d_block_rep, block_rep.data(), sizeof(uint16_t) * b, hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_quads, e->quads.data(), sizeof(plq::EnsQuad) * e->quads.size(), hipMemcpyHostToDevice));
    plq::launch_ens_qtab(qtab_of(e), e->d_slot_of, e->stream);
    ENS_HIP(hipGetLastError());
    if (e->graph_exec) {
        (void)hipGraphExecDestroy(e->graph_exec);
        e->graph_exec = nullptr;
    }
    return PLQ_OK;
}

void enqueue_block(plq_ensemble *e, uint32_t nsweeps, bool measure) {
    plq::EnsPassArgs a{};
    a.words = e->d_words;
    a.quads = e->d_quads;
    a.block_quad = e->d_block_rep;
    a.nrep = e->nrep;
    a.D = e->D;
    a.tables = e->d_tables;
    a.beta_slot = e->d_slot_of;
    a.qtab = e->d_qtab;
    a.sweep_base = e->d_counters;
    a.mode = e->kernel == PLQ_KERNEL_METROPOLIS ? 0 : 1;
    a.fast = e->fast;
    a.blocks = e->total_blocks;
    a.block_size = e->block_size;
    a.per_lane = e->per_lane;
    a.stream = e->stream;
    // use_graph = 2: the block's sweeps in one cooperative launch (grid
    // barriers between the passes); a launch the device refuses falls back
    // to per-pass launches for good
    bool coop_done = false;
    if (e->use_graph == 2 && e->fast && nsweeps > 0) {
        if (plq::launch_ens_sweeps_coop(a, nsweeps) == hipSuccess) coop_done = true;
        else {
            (void)hipGetLastError();
            e->use_graph = 0;
        }
    }
    for (uint32_t k = 0; k < nsweeps && !coop_done; ++k) {
        a.sweep_off = k;
        for (int mu = 0; mu < e->D; ++mu)
            for (int c = 0; c < 2; ++c) {
                a.mu = mu;
                a.c = c;
                plq::launch_ens_pass(a);
            }
    }
    if (measure) {
        // the sums are zero here: ens_post clears them after each record
        plq::launch_ens_plaquette(e->d_words, e->d_quads, e->d_block_rep, e->nrep, e->D, e->total_blocks, e->block_size,
                                  e->per_lane, e->d_sums, e->stream, false);
        plq::launch_ens_post(e->d_sums, e->d_slot_of, e->nrep, e->d_ring_sums, e->d_ring_slots, (uint32_t)e->rec_cap,
                             e->d_counters, e->d_pairs, e->exchange ? (int)e->pairs.size() : 0, e->d_rep_at,
                             e->d_ex_tables, e->ex_seed32, e->d_accepted, nsweeps, qtab_of(e), e->stream);
        return;
    }
    plq::launch_ens_tick(e->d_counters, nsweeps, 0, e->stream);
}

void account(plq_ensemble *e, uint32_t nsweeps, bool measure) {
    e->sweeps_done += nsweeps;
    if (measure) {
        e->records_done += 1;
        if (e->exchange) e->rounds_done += 1;
    }
}

int build_graph(plq_ensemble *e, uint32_t nsweeps, bool measure) {
    if (e->graph_exec) {
        (void)hipGraphExecDestroy(e->graph_exec);
        e->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    ENS_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    enqueue_block(e, nsweeps, measure);
    ENS_HIP(hipStreamEndCapture(e->stream, &g));
    hipError_t err = hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) {
        e->graph_exec = nullptr;
        return plq::host_fail(PLQ_EDEVICE, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(err));
    }
    e->graph_sweeps = nsweeps;
    e->graph_measures = measure;
    return PLQ_OK;
}

int advance(plq_ensemble *e, uint64_t sweeps) {
    const uint32_t block = e->measure_every ? (uint32_t)e->measure_every : 10u;
    const bool measure = e->measure_every != 0;
    while (sweeps > 0) {
        uint64_t into = measure ? e->sweeps_done % e->measure_every : 0;
        if (e->use_graph == 1 && into == 0 && sweeps >= block) {
            if (!e->graph_exec || e->graph_sweeps != block || e->graph_measures != measure)
                ENS_TRY(build_graph(e, block, measure));
            ENS_HIP(hipGraphLaunch(e->graph_exec, e->stream));
            account(e, block, measure);
            sweeps -= block;
            continue;
        }
        uint64_t n = measure ? std::min<uint64_t>(sweeps, e->measure_every - into) : std::min<uint64_t>(sweeps, block);
        bool reaches = measure && into + n == e->measure_every;
        enqueue_block(e, (uint32_t)n, reaches);
        account(e, (uint32_t)n, reaches);
        sweeps -= n;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return plq::host_fail(PLQ_EDEVICE, std::string("launch failed: ") + hipGetErrorString(err));
    return PLQ_OK;
}

// Exchange threshold table of one pair: T[ax] = number of 24-bit draws m
// with m * 2^-24 < exp(s * x) for the x of magnitude ax on the rejecting side
// (s = (beta_k+1 - beta_k) * J); entries stop once T reaches 0.
std::vector<uint32_t> exchange_table(double s) {
    std::vector<uint32_t> t;
    if (s == 0.0) return std::vector<uint32_t>(1, 1u << 24);  // equal betas always swap
    for (int64_t ax = 0;; ++ax) {
        double x = s >= 0.0 ? -(double)ax : (double)ax;
        double delta = s * x;
        double w = std::exp(delta);
        double scaled = w * 16777216.0;  // exact scaling by 2^24
        uint32_t T;
        if (delta >= 0.0) T = 1u << 24;
        else {
            // count m in [0, 2^24) with m * 2^-24 < w: m < w * 2^24
            double c = std::ceil(scaled);
            T = c >= 16777216.0 ? (1u << 24) : (uint32_t)c;
        }
        t.push_back(T);
        if (T == 0 || t.size() >= (size_t(1) << 22)) break;
    }
    return t;
}

}  // namespace

extern "C" {

int plq_ensemble_create(const plq_ensemble_desc *d, const uint8_t *start, size_t n, plq_ensemble **out) {
    if (!d || !out) return plq::host_fail(PLQ_EINVAL, "descriptor or output handle is null");
    *out = nullptr;
    if (d->dim < 2 || d->dim > 4) return plq::host_fail(PLQ_EINVAL, "ensembles support dimensions 2..4");
    if (d->n_replicas < 1 || !d->shapes || !d->betas || !d->seeds)
        return plq::host_fail(PLQ_EINVAL, "an ensemble needs at least one replica with shape, beta and seed");
    if (d->n_replicas > 65535) return plq::host_fail(PLQ_EINVAL, "an ensemble holds at most 65535 replicas");
    if (d->kernel < 0 || d->kernel > 1 || d->profile < 0 || d->profile > 1)
        return plq::host_fail(PLQ_EINVAL, "unknown kernel or threshold profile");
    if (!std::isfinite(d->j)) return plq::host_fail(PLQ_EINVAL, "j must be finite");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return plq::host_fail(PLQ_ENODEV, "no HIP device is available");
    plq_ensemble *e = new plq_ensemble();
    auto bail = [&](int code) {
        delete e;
        return code;
    };
    (void)hipGetDevice(&e->device);
    e->D = d->dim;
    e->nrep = d->n_replicas;
    e->kernel = d->kernel;
    e->profile = d->profile;
    e->j = d->j;
    e->measure_every = d->measure_every;
    e->exchange = d->exchange ? 1 : 0;
    e->ex_seed32 = plq::fold_seed(d->exchange_seed) ^ EXCHANGE_SALT;
    e->rec_cap = d->record_capacity ? d->record_capacity : 1;
    if (e->exchange && !e->measure_every)
        return bail(plq::host_fail(PLQ_EINVAL, "replica exchange needs measure_every > 0"));
    // replicas
    uint32_t woff = 0;
    size_t loff = 0;
    for (int r = 0; r < e->nrep; ++r) {
        plq::EnsReplica R{};
        int64_t ns = 1;
        for (int a = 0; a < 4; ++a) R.L[a] = 1;
        for (int a = 0; a < e->D; ++a) {
            int64_t l = d->shapes[r * e->D + a];
            if (l < 2 || (l & 1)) return bail(plq::host_fail(PLQ_EINVAL, "every extent must be even and >= 2"));
            R.L[a] = (int32_t)l;
            ns *= l;
        }
        if (ns * e->D > (int64_t(1) << 31)) return bail(plq::host_fail(PLQ_ENOMEM, "replica too large"));
        if (!std::isfinite(d->betas[r])) return bail(plq::host_fail(PLQ_EINVAL, "beta must be finite"));
        R.nsites = (uint32_t)ns;
        R.H = R.L[0] / 2;
        int32_t st = 1;
        R.rstride[0] = 0;
        for (int a = 1; a < 4; ++a) {
            R.rstride[a] = st;
            st *= R.L[a];
        }
        R.mH = magic40((uint32_t)R.H);
        for (int a = 0; a < 4; ++a) R.mL[a] = magic40((uint32_t)R.L[a]);
        R.off[0] = woff;
        R.off[1] = woff + R.nsites / 2;
        woff += R.nsites;
        R.seed32 = plq::fold_seed(d->seeds[r]);
        e->reps.push_back(R);
        e->betas.push_back(d->betas[r]);
        e->lat_off.push_back(loff);
        loff += (size_t)ns * e->D;
    }
    e->nwords = woff;
    e->lat_bytes = loff;
    if (!start || n != loff)
        return bail(plq::host_fail(PLQ_EINVAL, "start does not hold every replica's links"));
    for (size_t i = 0; i < n; ++i)
        if (start[i] > 1) return bail(plq::host_fail(PLQ_EINVAL, "link states must be 0 or 1"));
    // thresholds per beta slot
    std::vector<uint32_t> tables;
    for (int r = 0; r < e->nrep; ++r) {
        plq::TableParams tp{};
        tp.D = e->D;
        tp.model = 2;
        tp.q = 2;
        tp.kernel = e->kernel;
        tp.profile = e->profile;
        tp.beta = e->betas[r];
        tp.j = e->j;
        std::vector<uint32_t> t;
        std::string err;
        if (!plq::build_table(tp, t, err)) return bail(plq::host_fail(PLQ_EINVAL, err));
        tables.insert(tables.end(), t.begin(), t.end());
    }
    // The index-delta pass compares against threshold * 256 - 1, which needs
    // every threshold in [1, 2^24]; PLQ_ENS_FAST=0 keeps the first pass (tests
    // compare the two).
    e->fast = 1;
    for (uint32_t t : tables)
        if (t == 0 || t > plq::DRAW_ONE) e->fast = 0;
    if (const char *f = getenv("PLQ_ENS_FAST"))
        if (atoi(f) == 0) e->fast = 0;
    // exchange groups: replicas of one shape, ordered by beta
    std::vector<uint32_t> rep_at;
    std::vector<uint32_t> ex_tables;
    std::map<std::vector<int32_t>, std::vector<int>> groups;
    std::vector<std::vector<int32_t>> order;
    for (int r = 0; r < e->nrep; ++r) {
        std::vector<int32_t> key(e->reps[r].L, e->reps[r].L + 4);
        if (!groups.count(key)) order.push_back(key);
        groups[key].push_back(r);
    }
    for (auto &key : order) {
        std::vector<int> members = groups[key];
        std::stable_sort(members.begin(), members.end(), [&](int a, int b) { return e->betas[a] < e->betas[b]; });
        uint32_t base = (uint32_t)rep_at.size();
        for (int m : members) rep_at.push_back((uint32_t)m);
        for (size_t k = 0; k + 1 < members.size(); ++k) {
            plq::EnsGroupPair P{};
            P.k = (uint32_t)k;
            P.rep_at0 = base;
            double s = (e->betas[members[k + 1]] - e->betas[members[k]]) * e->j;
            P.positive = s == 0.0 ? 2u : (s > 0.0 ? 1u : 0u);
            std::vector<uint32_t> t = exchange_table(s);
            P.table0 = (uint32_t)ex_tables.size();
            P.table_len = (uint32_t)t.size();
            ex_tables.insert(ex_tables.end(), t.begin(), t.end());
            e->pairs.push_back(P);
        }
    }
    // quads: the replicas of each shape in descriptor order, four to a word
    e->rep_quad.assign(e->nrep, -1);
    e->rep_byte.assign(e->nrep, -1);
    uint32_t qwoff = 0;
    for (auto &key : order) {
        const std::vector<int> &members = groups[key];
        for (size_t m0 = 0; m0 < members.size(); m0 += 4) {
            const plq::EnsReplica &R = e->reps[members[m0]];
            plq::EnsQuad Q{};
            for (int a = 0; a < 4; ++a) {
                Q.L[a] = R.L[a];
                Q.rstride[a] = R.rstride[a];
                Q.mL[a] = R.mL[a];
                Q.rep[a] = -1;
            }
            Q.H = R.H;
            Q.nsites = R.nsites;
            Q.mH = R.mH;
            Q.woff = qwoff;
            qwoff += (uint32_t)e->D * R.nsites;
            for (size_t jj = 0; jj < 4 && m0 + jj < members.size(); ++jj) {
                const int r = members[m0 + jj];
                Q.rep[jj] = r;
                Q.seed32[jj] = e->reps[r].seed32;
                e->rep_quad[r] = (int)e->quads.size();
                e->rep_byte[r] = (int)jj;
            }
            e->quads.push_back(Q);
        }
    }
    e->nwords = qwoff;
    if (e->quads.size() > 65535) return bail(plq::host_fail(PLQ_EINVAL, "too many replica quads"));
    std::vector<uint32_t> slot_of(e->nrep);
    for (int r = 0; r < e->nrep; ++r) slot_of[r] = (uint32_t)r;
    int rc;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(plq::host_fail(PLQ_EDEVICE, "hipStreamCreate failed"));
    if ((rc = dalloc((void **)&e->d_quads, sizeof(plq::EnsQuad) * e->quads.size())) ||
        (rc = dalloc((void **)&e->d_words, sizeof(uint32_t) * e->nwords)) ||
        (rc = dalloc((void **)&e->d_lat, e->lat_bytes)) ||
        (rc = dalloc((void **)&e->d_tables, sizeof(uint32_t) * tables.size())) ||
        (rc = dalloc((void **)&e->d_slot_of, sizeof(uint32_t) * e->nrep)) ||
        (rc = dalloc((void **)&e->d_qtab, sizeof(uint32_t) * e->quads.size() * 4 * (2 * (2 * e->D - 1)))) ||
        (rc = dalloc((void **)&e->d_rep_at, sizeof(uint32_t) * rep_at.size())) ||
        (rc = dalloc((void **)&e->d_pairs, sizeof(plq::EnsGroupPair) * std::max<size_t>(1, e->pairs.size()))) ||
        (rc = dalloc((void **)&e->d_ex_tables, sizeof(uint32_t) * std::max<size_t>(1, ex_tables.size()))) ||
        (rc = dalloc((void **)&e->d_sums, sizeof(unsigned long long) * e->nrep)) ||
        (rc = dalloc((void **)&e->d_accepted, sizeof(unsigned long long) * std::max<size_t>(1, e->pairs.size()))) ||
        (rc = dalloc((void **)&e->d_counters, sizeof(uint32_t) * 4)) ||
        (rc = dalloc((void **)&e->d_ring_sums, sizeof(long long) * e->rec_cap * e->nrep)) ||
        (rc = dalloc((void **)&e->d_ring_slots, sizeof(uint32_t) * e->rec_cap * e->nrep)))
        return bail(rc);
    if (hipMemcpy(e->d_tables, tables.data(), sizeof(uint32_t) * tables.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(e->d_slot_of, slot_of.data(), sizeof(uint32_t) * e->nrep, hipMemcpyHostToDevice) ||
        hipMemcpy(e->d_rep_at, rep_at.data(), sizeof(uint32_t) * rep_at.size(), hipMemcpyHostToDevice) ||
        (e->pairs.size() &&
         hipMemcpy(e->d_pairs, e->pairs.data(), sizeof(plq::EnsGroupPair) * e->pairs.size(), hipMemcpyHostToDevice)) ||
        (ex_tables.size() &&
         hipMemcpy(e->d_ex_tables, ex_tables.data(), sizeof(uint32_t) * ex_tables.size(), hipMemcpyHostToDevice)) ||
        hipMemset(e->d_accepted, 0, sizeof(unsigned long long) * std::max<size_t>(1, e->pairs.size())) ||
        hipMemset(e->d_sums, 0, sizeof(unsigned long long) * e->nrep) ||
        hipMemset(e->d_counters, 0, sizeof(uint32_t) * 4) ||
        hipMemset(e->d_words, 0, sizeof(uint32_t) * e->nwords) ||  // unused quad bytes stay 0
        hipMemcpy(e->d_lat, start, n, hipMemcpyHostToDevice))
        return bail(plq::host_fail(PLQ_EDEVICE, "uploading the ensemble failed"));
    for (int r = 0; r < e->nrep; ++r)
        plq::launch_ens_pack(e->d_lat + e->lat_off[r], e->d_words, e->quads[e->rep_quad[r]], e->rep_byte[r], e->D,
                             e->stream);
    if ((rc = layout_blocks(e)) != PLQ_OK) return bail(rc);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return bail(plq::host_fail(PLQ_EDEVICE, "pack failed"));
    *out = e;
    return PLQ_OK;
}

int plq_ensemble_destroy(plq_ensemble *e) {
    if (!e) return plq::host_fail(PLQ_EINVAL, "ensemble is null");
    delete e;
    return PLQ_OK;
}

int plq_ensemble_advance(plq_ensemble *e, uint64_t sweeps) {
    if (!e) return plq::host_fail(PLQ_EINVAL, "ensemble is null");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    return advance(e, sweeps);
}

int plq_ensemble_read(plq_ensemble *e, uint8_t *out, size_t n) {
    if (!e || !out) return plq::host_fail(PLQ_EINVAL, "ensemble or buffer is null");
    if (n != e->lat_bytes) return plq::host_fail(PLQ_EINVAL, "buffer length does not match the ensemble");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    for (int r = 0; r < e->nrep; ++r)
        plq::launch_ens_unpack(e->d_lat + e->lat_off[r], e->d_words, e->quads[e->rep_quad[r]], e->rep_byte[r], e->D,
                               e->stream);
    ENS_HIP(hipMemcpyAsync(out, e->d_lat, n, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    return PLQ_OK;
}

int plq_ensemble_plaquette_sums(plq_ensemble *e, int64_t *sums, size_t n) {
    if (!e || !sums) return plq::host_fail(PLQ_EINVAL, "ensemble or buffer is null");
    if (n != (size_t)e->nrep) return plq::host_fail(PLQ_EINVAL, "one sum per replica");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    plq::launch_ens_plaquette(e->d_words, e->d_quads, e->d_block_rep, e->nrep, e->D, e->total_blocks, e->block_size, e->per_lane,
                              e->d_sums, e->stream, true);
    std::vector<unsigned long long> h(e->nrep);
    ENS_HIP(hipMemcpyAsync(h.data(), e->d_sums, sizeof(unsigned long long) * e->nrep, hipMemcpyDeviceToHost,
                           e->stream));
    plq::launch_ens_zero(e->d_sums, e->nrep, e->stream);  // a measured block's plaquette adds into zeroed sums
    ENS_HIP(hipStreamSynchronize(e->stream));
    for (int r = 0; r < e->nrep; ++r) sums[r] = (int64_t)h[r];
    return PLQ_OK;
}

int plq_ensemble_betas(plq_ensemble *e, double *betas, size_t n) {
    if (!e || !betas) return plq::host_fail(PLQ_EINVAL, "ensemble or buffer is null");
    if (n != (size_t)e->nrep) return plq::host_fail(PLQ_EINVAL, "one beta per replica");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    std::vector<uint32_t> slot(e->nrep);
    ENS_HIP(hipMemcpyAsync(slot.data(), e->d_slot_of, sizeof(uint32_t) * e->nrep, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    for (int r = 0; r < e->nrep; ++r) betas[r] = e->betas[slot[r]];
    return PLQ_OK;
}

int plq_ensemble_series(plq_ensemble *e, int64_t *sums, uint32_t *slots, uint64_t capacity, uint64_t *n_out) {
    if (!e || !n_out) return plq::host_fail(PLQ_EINVAL, "ensemble or count is null");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    uint64_t have = std::min<uint64_t>(e->records_done, e->rec_cap);
    *n_out = have;
    if (!sums || have == 0) return PLQ_OK;
    if (capacity < have) return plq::host_fail(PLQ_EINVAL, "series buffer too small");
    std::vector<long long> hs(e->rec_cap * e->nrep);
    std::vector<uint32_t> hl(e->rec_cap * e->nrep);
    ENS_HIP(hipMemcpyAsync(hs.data(), e->d_ring_sums, sizeof(long long) * hs.size(), hipMemcpyDeviceToHost,
                           e->stream));
    ENS_HIP(hipMemcpyAsync(hl.data(), e->d_ring_slots, sizeof(uint32_t) * hl.size(), hipMemcpyDeviceToHost,
                           e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    uint64_t first = e->records_done - have;
    for (uint64_t k = 0; k < have; ++k) {
        uint64_t slot = (first + k) % e->rec_cap;
        for (int r = 0; r < e->nrep; ++r) {
            sums[k * e->nrep + r] = hs[slot * e->nrep + r];
            if (slots) slots[k * e->nrep + r] = hl[slot * e->nrep + r];
        }
    }
    return PLQ_OK;
}

int plq_ensemble_exchange_counts(plq_ensemble *e, uint64_t *accepted, uint64_t *attempted, size_t npairs,
                                 size_t *npairs_out) {
    if (!e || !npairs_out) return plq::host_fail(PLQ_EINVAL, "ensemble or count is null");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    *npairs_out = e->pairs.size();
    if (!accepted || e->pairs.empty()) return PLQ_OK;
    if (npairs < e->pairs.size()) return plq::host_fail(PLQ_EINVAL, "buffer too small");
    std::vector<unsigned long long> h(e->pairs.size());
    ENS_HIP(hipMemcpyAsync(h.data(), e->d_accepted, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                           e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    for (size_t p = 0; p < h.size(); ++p) {
        accepted[p] = h[p];
        if (attempted) {
            // rounds with the pair's parity among rounds 0 .. rounds_done - 1
            uint64_t k = e->pairs[p].k % 2;
            attempted[p] = e->rounds_done > k ? (e->rounds_done - k + 1) / 2 : 0;
        }
    }
    return PLQ_OK;
}

int plq_ensemble_set_launch(plq_ensemble *e, int32_t block_size, int32_t per_lane, int32_t use_graph) {
    if (!e) return plq::host_fail(PLQ_EINVAL, "ensemble is null");
    if (block_size != 64 && block_size != 128 && block_size != 256)
        return plq::host_fail(PLQ_EINVAL, "block size must be 64, 128 or 256");
    if (per_lane < 1 || per_lane > 8) return plq::host_fail(PLQ_EINVAL, "links per lane must be 1..8");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    ENS_HIP(hipStreamSynchronize(e->stream));
    if (use_graph < 0 || use_graph > 2) return plq::host_fail(PLQ_EINVAL, "use_graph must be 0, 1 or 2");
    e->block_size = block_size;
    e->per_lane = per_lane;
    e->use_graph = use_graph;
    return layout_blocks(e);
}

int plq_ensemble_tune(plq_ensemble *e, int32_t *block_size, int32_t *per_lane, int32_t *use_graph,
                      double *best_us) {
    if (!e) return plq::host_fail(PLQ_EINVAL, "ensemble is null");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    ENS_HIP(hipStreamSynchronize(e->stream));
    // snapshot everything a sweep block mutates
    std::vector<uint32_t> words(e->nwords), slot(e->nrep), counters(4);
    std::vector<uint32_t> rep_at;
    size_t n_rep_at = 0;
    for (auto &P : e->pairs) n_rep_at = std::max<size_t>(n_rep_at, P.rep_at0 + P.k + 2);
    n_rep_at = std::max<size_t>(n_rep_at, (size_t)e->nrep);
    rep_at.resize(n_rep_at);
    std::vector<unsigned long long> acc(std::max<size_t>(1, e->pairs.size()));
    std::vector<long long> ring(e->rec_cap * e->nrep);
    std::vector<uint32_t> ring_slots(e->rec_cap * e->nrep);
    ENS_HIP(hipMemcpy(words.data(), e->d_words, sizeof(uint32_t) * e->nwords, hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(slot.data(), e->d_slot_of, sizeof(uint32_t) * e->nrep, hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(rep_at.data(), e->d_rep_at, sizeof(uint32_t) * e->nrep, hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(counters.data(), e->d_counters, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(acc.data(), e->d_accepted, sizeof(unsigned long long) * acc.size(), hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(ring.data(), e->d_ring_sums, sizeof(long long) * ring.size(), hipMemcpyDeviceToHost));
    ENS_HIP(hipMemcpy(ring_slots.data(), e->d_ring_slots, sizeof(uint32_t) * ring_slots.size(),
                      hipMemcpyDeviceToHost));
    const uint64_t saved_sweeps = e->sweeps_done, saved_rounds = e->rounds_done, saved_records = e->records_done;
    const int saved_bs = e->block_size, saved_pl = e->per_lane, saved_g = e->use_graph;
    hipEvent_t t0, t1;
    ENS_HIP(hipEventCreate(&t0));
    ENS_HIP(hipEventCreate(&t1));
    double best = 1e300;
    int bbs = saved_bs, bpl = saved_pl, bg = saved_g;
    const uint64_t block = e->measure_every ? e->measure_every : 10;
    const uint64_t trial = block * 3;
    for (int g : {1, 0, 2})
        for (int bs : {64, 128, 256})
            for (int pl : {1, 2, 4}) {
                e->block_size = bs;
                e->per_lane = pl;
                e->use_graph = g;
                ENS_TRY(layout_blocks(e));
                ENS_TRY(advance(e, block));  // warm (and graph build)
                ENS_HIP(hipEventRecord(t0, e->stream));
                ENS_TRY(advance(e, trial));
                ENS_HIP(hipEventRecord(t1, e->stream));
                ENS_HIP(hipEventSynchronize(t1));
                float ms = 0;
                ENS_HIP(hipEventElapsedTime(&ms, t0, t1));
                double us = 1e3 * ms / (double)trial;
                if (us < best) {
                    best = us;
                    bbs = bs;
                    bpl = pl;
                    bg = g;
                }
            }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    // restore
    ENS_HIP(hipMemcpy(e->d_words, words.data(), sizeof(uint32_t) * e->nwords, hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_slot_of, slot.data(), sizeof(uint32_t) * e->nrep, hipMemcpyHostToDevice));
    plq::launch_ens_qtab(qtab_of(e), e->d_slot_of, e->stream);
    ENS_HIP(hipGetLastError());
    ENS_HIP(hipMemcpy(e->d_rep_at, rep_at.data(), sizeof(uint32_t) * e->nrep, hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_counters, counters.data(), sizeof(uint32_t) * 4, hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_accepted, acc.data(), sizeof(unsigned long long) * acc.size(), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_ring_sums, ring.data(), sizeof(long long) * ring.size(), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(e->d_ring_slots, ring_slots.data(), sizeof(uint32_t) * ring_slots.size(),
                      hipMemcpyHostToDevice));
    e->sweeps_done = saved_sweeps;
    e->rounds_done = saved_rounds;
    e->records_done = saved_records;
    e->block_size = bbs;
    e->per_lane = bpl;
    e->use_graph = bg;
    ENS_TRY(layout_blocks(e));
    if (block_size) *block_size = bbs;
    if (per_lane) *per_lane = bpl;
    if (use_graph) *use_graph = bg;
    if (best_us) *best_us = best;
    return PLQ_OK;
}

int plq_ensemble_sweeps_done(plq_ensemble *e, uint64_t *out) {
    if (!e || !out) return plq::host_fail(PLQ_EINVAL, "ensemble or output is null");
    *out = e->sweeps_done;
    return PLQ_OK;
}

int plq_ensemble_stream(plq_ensemble *e, void **stream) {
    if (!e || !stream) return plq::host_fail(PLQ_EINVAL, "ensemble or output is null");
    *stream = (void *)e->stream;
    return PLQ_OK;
}

int plq_ensemble_sync(plq_ensemble *e) {
    if (!e) return plq::host_fail(PLQ_EINVAL, "ensemble is null");
    std::lock_guard<std::mutex> lk(e->mu);
    ENS_HIP(hipSetDevice(e->device));
    ENS_HIP(hipStreamSynchronize(e->stream));
    return PLQ_OK;
}

}  // extern "C"
