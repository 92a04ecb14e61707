// plq_ensemble.hpp -- device-side descriptors of a batched replica ensemble.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plq {

struct EnsReplica {
    int32_t L[4];         // extents (unused axes 1)
    int32_t H;            // L0 / 2 words per colour row
    int32_t rstride[4];   // row stride of axis a >= 1
    uint32_t nsites;
    uint32_t off[2];      // word offsets of the two colour arrays
    uint64_t mH;          // ceil(2^40 / H)
    uint64_t mL[4];       // ceil(2^40 / L[a])
    uint32_t block0;      // first block of this replica in a pass launch
    uint32_t seed32;      // fold_seed of the replica's seed
};

// Four replicas of one shape share their link words: byte j of every word
// is replica rep[j]'s link (rep[j] < 0: an unused byte).  Word arrays per
// (direction mu, colour c): woff + (mu * 2 + c) * nsites / 2, each in the
// colour-interleaved site order r * H + x0 / 2.
struct EnsQuad {
    int32_t L[4];         // extents (unused axes 1)
    int32_t H;            // L0 / 2 sites per colour row
    int32_t rstride[4];   // row stride of axis a >= 1
    uint32_t nsites;
    uint64_t mH;          // ceil(2^40 / H)
    uint64_t mL[4];       // ceil(2^40 / L[a])
    uint32_t woff;        // first word of the quad
    uint32_t block0;      // first block of this quad in a pass launch
    int32_t rep[4];       // replica in byte j, or -1
    uint32_t seed32[4];   // fold_seed of each replica's seed
};

struct EnsGroupPair {
    uint32_t k;           // the pair (k, k+1) of its group's sorted betas
    uint32_t rep_at0;     // offset of the group's rep_at list
    uint32_t positive;    // 1 if (beta_k+1 - beta_k) J >= 0
    uint32_t table0;      // offset of its threshold table (indexed by |P_a - P_b|)
    uint32_t table_len;
};

// The fast pass's thresholds per quad, kept in step with the replicas' beta
// slots: qtab[(q * 4 + j) * TL + e] = tables[slot_of[r] * TL + e] * 256 - 1
// for the replica r in byte j of quad q (0 for an empty byte), TL = 2 (2D - 1).
// A pass's workgroup then fills its LDS table with one load per entry
// instead of the chain quads -> slot_of -> tables.
struct EnsQtab {
    const EnsQuad *quads;
    int nquads;
    const uint32_t *tables;
    uint32_t *qtab;
    int TL;
};

struct EnsPassArgs {
    uint32_t *words;
    const EnsQuad *quads;
    const uint16_t *block_quad;  // quad of each (logical) block
    int nrep;
    int D;
    const uint32_t *tables;      // per beta slot: 2 * (2D - 1) thresholds
    const uint32_t *beta_slot;   // per replica
    const uint32_t *qtab;        // per quad: the fast pass's 4 x TL thresholds * 256 - 1 (EnsQtab)
    const uint32_t *sweep_base;  // device sweep counter
    uint32_t sweep_off;
    int mu, c;
    int mode;                    // 0 Metropolis, 1 heat bath
    int fast;                    // every threshold >= 1: ens_link_pass2
    uint32_t blocks;
    int block_size;
    int per_lane;
    hipStream_t stream;
};

// Replica in byte j of quad Q: lattice order (link = site * D + dir) <-> words.
void launch_ens_pack(const uint8_t *lat, uint32_t *words, const EnsQuad &Q, int j, int D, hipStream_t s);
void launch_ens_unpack(uint8_t *lat, const uint32_t *words, const EnsQuad &Q, int j, int D, hipStream_t s);
void launch_ens_pass(const EnsPassArgs &a);
// Every colour pass of nsweeps sweeps in one cooperative launch with grid
// barriers between the passes (ens_link_pass2's body; a.fast must hold).
hipError_t launch_ens_sweeps_coop(const EnsPassArgs &a, uint32_t nsweeps);
// Per-replica signed plaquette sums added into sums (cleared first with
// zero_first).
void launch_ens_plaquette(const uint32_t *words, const EnsQuad *quads, const uint16_t *block_quad, int nrep, int D,
                          uint32_t blocks, int block_size, int per_lane, unsigned long long *sums, hipStream_t s,
                          bool zero_first);
void launch_ens_zero(unsigned long long *sums, int nrep, hipStream_t s);
// A measured block's tail in one launch (ens_post): record, exchange round
// (npairs > 0), sums cleared, counters advanced by sweeps and the round.
// The exchange round rewrites the fast pass's thresholds of every quad (qt).
void launch_ens_post(unsigned long long *sums, uint32_t *slot_of, int nrep, long long *ring_sums, uint32_t *ring_slots,
                     uint32_t cap, uint32_t *counters, const EnsGroupPair *pairs, int npairs, uint32_t *rep_at,
                     const uint32_t *ex_tables, uint32_t ex_seed, unsigned long long *accepted, uint32_t sweeps,
                     const EnsQtab &qt, hipStream_t s);
// EnsQtab's thresholds from the current beta slots (after any host write of them).
void launch_ens_qtab(const EnsQtab &qt, const uint32_t *slot_of, hipStream_t s);
void launch_ens_tick(uint32_t *counters, uint32_t sweeps, uint32_t rounds, hipStream_t s);

}  // namespace plq
