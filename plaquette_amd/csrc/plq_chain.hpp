// plq_chain.hpp -- the chain object behind the opaque plq_chain handle of
// the C ABI, shared by the translation units that drive chains.  Private to
// the library.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/plaquette_hip.h"
#include "plq_internal.hpp"

enum plq_chain_kind { KIND_SITE = 0, KIND_LINK = 1, KIND_CLUSTER = 2 };

// Entries of one cluster-statistics slot: 40 size bins, the cluster count,
// the largest size, then the last relabel's REC_K partial sums.
constexpr size_t CLU_SLOT = 42 + plq::REC_K;

struct plq_chain {
    std::mutex mu;
    plq_chain_kind kind;
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 0;
    int64_t shape[PLQ_MAX_DIM] = {0};
    int64_t n_sites = 0, n_vars = 0;
    int model = 0, q = 2, kernel = 0, profile = 0;
    double beta = 0, j = 1, h = 0;
    double offsets[PLQ_MAX_Q] = {0};
    uint32_t seed32 = 0;
    uint64_t sweeps_done = 0, sweeps_between = 1;
    // local fields
    plq::Geom g{};
    plq::SwarGeom sg{};
    int vec = 1;                           // sites per lane of the site kernel
    int x16u = 0;                          // the uniform-parity 16-site pass serves this chain
    plq::X16UGeom ug{};
    bool small = false;                    // whole sweeps in one workgroup (plq_small.hip, plq_cluster_small.hip)
    int64_t pass_blocks = plq::MAX_PASS_BLOCKS;  // grid cap of a colour pass
    uint8_t *col[2] = {nullptr, nullptr};  // site colours
    // The fused two-colour sweep reads col and writes col_alt, then the two
    // pairs swap (ising_sweep_x16f).
    uint8_t *col_alt[2] = {nullptr, nullptr};
    bool fused = false;
    uint8_t *links[2 * plq::MAX_D] = {nullptr};
    uint8_t **d_links = nullptr;           // device copy of the link array pointers
    uint32_t *d_table = nullptr;
    size_t table_len = 0;
    std::vector<uint32_t> h_table;  // host copy of d_table
    // cluster
    plq::CCGeom cg{};
    uint8_t *labels = nullptr;
    int32_t *parent = nullptr;
    int32_t *roots = nullptr;
    int32_t *sizes = nullptr;
    int32_t *ext = nullptr;                // cluster counts from other boxes (box-counted sizes), zero between calls
    uint8_t *d_cross = nullptr;            // roots of face-crossing clusters, marked by cc_compress
    uint32_t cross_sweep = 0;              // the sweep whose marks are the latest
    unsigned long long *d_hist_part = nullptr;  // per-workgroup histogram rows of the statistics kernels
    bool sizes_counted = false;            // the last sweep's relabel counted its sizes into sizes / ext / rows
    unsigned long long *hist = nullptr;
    int *d_flag = nullptr;
    uint32_t bond_t = 0;
    int seeded = 0, forced = 0;
    uint64_t pass_cap = 0;
    int64_t last_relabel_blocks = 0;
    bool have_partition = false;
    bool relabel_measures = true;          // a recorded SW sweep's relabel also counts its record
    // A record whose fold waits for the next sweep's first kernel (while
    // fold_defer: inside one enqueue of direct sweeps, flushed at its end).
    plq::RecordFold fold{};
    bool fold_defer = false;
    // Wolff moves grown from the seed (plq_wolff.hip)
    int wolff_mode = 0;                    // 0 off, 1 adaptive, 2 an attempt every sweep
    uint32_t wolff_levels = 0, wolff_qcap = 0;
    float wolff_gain_ticks = 0;            // wall-clock ticks a grown sweep saves (adaptive mode)
    plq::WolffState *d_wolff = nullptr;
    int32_t *d_queue = nullptr;            // wolff_qcap sites of the grown cluster
    bool grow_pending = false;             // the last sweep may have grown: its partition is built on demand
    uint32_t grow_sweep = 0;
    // hipGraph replays of a uniform-pass site chain's sweeps
    int site_graph = 0;                    // replays on (PLQ_SITE_GRAPH=1; a failed capture turns them off)
    uint32_t *d_ctr = nullptr;             // replay counters: sweep, record slot
    bool ctr_valid = false;                // d_ctr holds sweeps_done and the next record slot
    hipGraphExec_t site_exec = nullptr;
    uint32_t exec_sweeps = 0;              // sweeps per replay
    bool exec_records = false;             // a replay ends with a record
    // shared scratch
    uint8_t *d_lat = nullptr;              // lattice-order staging
    long long *d_partials = nullptr;       // 2 * MAX_PASS_BLOCKS * REC_K
    long long *d_meas = nullptr;           // REC_K (+ cluster stats)
    long long *d_rec = nullptr;            // rec_cap * REC_K
    bool rec_clusters = false;             // recorded cluster sweeps also keep their partition's statistics
    unsigned long long *d_clu = nullptr;   // rec_cap * CLU_SLOT: histogram, count, largest, relabel sums
    uint64_t clu_from = 0;                 // rec_count when the cluster statistics began
    uint64_t rec_every = 0, rec_cap = 0, rec_count = 0;
    std::vector<uint64_t> rec_sweeps;
    std::vector<hipEvent_t> pending;
    // the CPU stream (PLQ_PROFILE_CPU_STREAM, plq_pmode.hip)
    bool pmode = false;
    uint64_t *d_pm_cls = nullptr;    // per class threshold
    uint64_t *d_pm_state = nullptr;  // state lo, hi, increment lo, hi
    uint64_t *d_pm_jt = nullptr;     // jump table
    uint32_t *d_pm_blk = nullptr, *d_pm_off = nullptr;
    // learned block proposal (plq_block_abi.hip)
    uint64_t block_sweeps = 0;                       // key of the block passes' draws
    unsigned long long *d_block_counters = nullptr;  // accepted moves, one slot per pass workgroup
    float *d_block_logits = nullptr;                 // logits of each block's last pass
    // per-configuration measurements (plq_measure_abi.hip)
    uint32_t *d_meas_bits = nullptr;           // bit planes (and Wilson line planes)
    size_t meas_bits_cap = 0;                  // words
    unsigned long long *d_meas_counts = nullptr;
    size_t meas_counts_cap = 0;
    // sample batches (plq_chain_next_batch)
    uint8_t *d_batch = nullptr;
    size_t batch_cap = 0;

    ~plq_chain() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (site_exec) (void)hipGraphExecDestroy(site_exec);
        for (auto p : {(void *)col[0], (void *)col[1], (void *)col_alt[0], (void *)col_alt[1], (void *)d_links, (void *)d_table, (void *)labels,
                       (void *)parent, (void *)roots, (void *)sizes, (void *)ext, (void *)hist, (void *)d_flag,
                       (void *)d_lat,
                       (void *)d_partials, (void *)d_meas, (void *)d_rec, (void *)d_block_counters, (void *)d_clu, (void *)d_cross,
                       (void *)d_hist_part,
                       (void *)d_block_logits, (void *)d_pm_cls, (void *)d_pm_state, (void *)d_pm_jt, (void *)d_pm_blk,
                       (void *)d_pm_off, (void *)d_meas_bits, (void *)d_meas_counts, (void *)d_batch,
                       (void *)d_wolff, (void *)d_queue, (void *)d_ctr})
            if (p) (void)hipFree(p);
        for (int k = 0; k < 2 * plq::MAX_D; ++k)
            if (links[k]) (void)hipFree(links[k]);
        if (stream) (void)hipStreamDestroy(stream);
    }
};
