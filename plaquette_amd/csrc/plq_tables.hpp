// plq_tables.hpp -- host-side integer acceptance tables (see plq_thresholds.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace plq {

struct TableParams {
    int D;
    int model;    // 0 Ising, 1 Potts, 2 Z2
    int q;
    int kernel;   // 0 Metropolis, 1 heat bath
    int profile;  // 0 f64 (CPU arithmetic), 1 f32 (WGSL arithmetic)
    double beta, j, h;
    double offsets[16];
};

// Table layouts:
//   Ising   2*(2D+1) entries, index cur*(2D+1) + d, d = neighbours in state 1.
//   Z2      2*(2D-1) entries, index cur*(2D-1) + k, k = negative staple products.
//   Ising/Z2 Metropolis: flip iff m < T.  Heat bath: new state 0 iff m < T, else 1.
//   Potts Metropolis: q*q*(4D+1), index (cur*q + prop)*(4D+1) + change + 2D,
//       accept iff m < T.
//   Potts heat bath: q * (2D+1)^(q-1) classes of (q-1) thresholds, class
//       (cur * (2D+1)^(q-1) + sum_a n_a (2D+1)^a) over a < q-1; chosen = first s
//       with m < T[s], else q-1.
bool build_table(const TableParams &p, std::vector<uint32_t> &out, std::string &err);
int64_t potts_hb_classes(int D, int q);
uint32_t bond_threshold(double beta, double gap, int profile);

}  // namespace plq
