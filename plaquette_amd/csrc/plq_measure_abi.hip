// plq_measure_abi.hip -- C ABI of the per-configuration measurements: the
// correlator of a site chain and the Polyakov loop and Wilson rectangle table
// of a Z2 gauge chain, computed where the configuration lives.  Counts come
// back as integers; the f64 results are formed here with the reference's own
// final operations, so they are the reference's values bit for bit.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plaquette_hip.h"
#include "plq_chain.hpp"
#include "plq_internal.hpp"

// Scratch of plq_chain_correlator_noalloc: bit planes, then counters.  A
// file-scope device array, so that entry point never allocates; every device
// the library runs on holds its own copy (8.06 MiB) until the process ends.
constexpr int64_t NOALLOC_PLANE_WORDS = int64_t(1) << 21;  // 8 MiB of bit-plane words
constexpr int64_t NOALLOC_COUNTS = 8192;                    // 64 KiB of u64 counters
static __device__ __align__(256) uint32_t scratch[NOALLOC_PLANE_WORDS + 2 * NOALLOC_COUNTS];

namespace {

int fail(int code, const std::string &msg) { return plq::host_fail(code, msg); }

#define PLQ_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PLQ_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define PLQ_TRY(expr)        \
    do {                     \
        int rc_ = (expr);    \
        if (rc_ != PLQ_OK) { \
            return rc_;      \
        }                    \
    } while (0)

int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return PLQ_OK;
}

// The device address of `scratch` on `device` (current): in host code the
// array's name is a host-side stand-in, so kernels and copies take this
// address, found once per device with hipGetSymbolAddress.
constexpr int NOALLOC_MAX_DEVICES = 64;

int scratch_address(int device, uint32_t **out) {
    static std::mutex mu;
    static uint32_t *addr[NOALLOC_MAX_DEVICES] = {};
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(PLQ_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PLQ_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        addr[device] = static_cast<uint32_t *>(p);
    }
    *out = addr[device];
    return PLQ_OK;
}

// Calls on one device share its scratch: they are serialised by this lock.
std::mutex &scratch_mutex(int device) {
    static std::mutex mu[NOALLOC_MAX_DEVICES];
    return mu[device];
}

// Grow the chain's measurement buffers to `words` plane words and `counts`
// counters; the counters are cleared on the chain's stream.
int ensure_meas(plq_chain *c, size_t words, size_t counts) {
    if (words > c->meas_bits_cap) {
        PLQ_HIP(hipStreamSynchronize(c->stream));
        if (c->d_meas_bits) (void)hipFree(c->d_meas_bits);
        c->d_meas_bits = nullptr;
        c->meas_bits_cap = 0;
        hipError_t e = hipMalloc((void **)&c->d_meas_bits, words * sizeof(uint32_t));
        if (e != hipSuccess)
            return fail(PLQ_ENOMEM, std::string("hipMalloc of measurement planes failed: ") + hipGetErrorString(e));
        c->meas_bits_cap = words;
    }
    if (counts > c->meas_counts_cap) {
        PLQ_HIP(hipStreamSynchronize(c->stream));
        if (c->d_meas_counts) (void)hipFree(c->d_meas_counts);
        c->d_meas_counts = nullptr;
        c->meas_counts_cap = 0;
        hipError_t e = hipMalloc((void **)&c->d_meas_counts, counts * sizeof(unsigned long long));
        if (e != hipSuccess)
            return fail(PLQ_ENOMEM, std::string("hipMalloc of measurement counters failed: ") + hipGetErrorString(e));
        c->meas_counts_cap = counts;
    }
    PLQ_HIP(hipMemsetAsync(c->d_meas_counts, 0, counts * sizeof(unsigned long long), c->stream));
    return PLQ_OK;
}

// The configuration in lattice order (site, or site * D + dir for links).
const uint8_t *lattice_bytes(plq_chain *c) {
    if (c->kind == KIND_SITE) {
        plq::launch_unpack_sites(c->d_lat, c->col[0], c->col[1], c->g, c->stream);
        return c->d_lat;
    }
    if (c->kind == KIND_LINK) {
        plq::launch_unpack_links(c->d_lat, c->d_links, c->g, c->stream);
        return c->d_lat;
    }
    return c->labels;
}

// A cluster chain whose labelling overran its budget holds no configuration.
int check_cluster(plq_chain *c) {
    if (c->kind != KIND_CLUSTER) return PLQ_OK;
    int flag = 0;
    PLQ_HIP(hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    if (flag) return fail(PLQ_ENOCONV, "cluster labelling did not converge; the configuration is undefined");
    return PLQ_OK;
}

int64_t pair_count(const plq_chain *c) {
    int64_t n = 0;
    for (int mu = 0; mu < c->D; ++mu) n += c->shape[mu] / 2 + 1;
    return n;
}

int wilson_side(const plq_chain *c, int32_t max_side) {
    int64_t shortest = c->shape[0];
    for (int mu = 1; mu < c->D; ++mu)
        if (c->shape[mu] < shortest) shortest = c->shape[mu];
    return (int64_t)max_side < shortest / 2 ? max_side : (int)(shortest / 2);
}

int read_counts(plq_chain *c, std::vector<unsigned long long> &out) {
    PLQ_HIP(hipMemcpyAsync(out.data(), c->d_meas_counts, out.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return PLQ_OK;
}

// Validation shared by both correlator entry points; *npairs receives the
// number of (axis, displacement) entries.
int correlator_args(const plq_chain *c, int32_t model, size_t len, int64_t *npairs) {
    if (c->kind == KIND_LINK) return fail(PLQ_EINVAL, "the correlator is measured on site chains (Ising, Potts)");
    if (model != PLQ_MODEL_ISING && model != PLQ_MODEL_POTTS)
        return fail(PLQ_EINVAL, "the correlator convention is PLQ_MODEL_ISING or PLQ_MODEL_POTTS");
    if (c->kind == KIND_SITE && model != c->model)
        return fail(PLQ_EINVAL, "the correlator convention must be the chain's model");
    if (model == PLQ_MODEL_ISING && c->q != 2) return fail(PLQ_EINVAL, "the Ising correlator needs q = 2");
    *npairs = pair_count(c);
    if (len != (size_t)*npairs)
        return fail(PLQ_EINVAL, "the output holds one entry per axis and displacement r = 0..L/2 (sum of L/2 + 1)");
    if (*npairs > (int64_t(1) << 30)) return fail(PLQ_EINVAL, "too many displacements");
    return PLQ_OK;
}

// models/mod.rs:44-60 returns the raw sums; ising/mod.rs:89-93 and
// potts/mod.rs:230-232 scale them.  agree may alias sums or values: entry k
// is read before entry k of either output is written.
void correlator_values(const plq_chain *c, int32_t model, int64_t npairs, const unsigned long long *agree,
                       int64_t *sums, double *values) {
    const double n = (double)c->n_sites;
    const double floor = 1.0 / (double)c->q;
    for (int64_t k = 0; k < npairs; ++k) {
        unsigned long long a;
        std::memcpy(&a, agree + k, sizeof a);  // agree may be the caller's int64 or double storage
        const int64_t s = model == PLQ_MODEL_ISING ? 2 * (int64_t)a - c->n_sites : (int64_t)a;
        if (sums) sums[k] = s;
        if (values) values[k] = model == PLQ_MODEL_ISING ? (double)s / n : (double)s / n - floor;
    }
}

}  // namespace

int plq_chain_correlator(plq_chain *c, int32_t model, int64_t *sums, double *values, size_t len) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lk(c->mu);
    int64_t npairs = 0;
    PLQ_TRY(correlator_args(c, model, len, &npairs));
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_TRY(check_cluster(c));
    const plq::MeasGeom m = plq::meas_geom(c->D, c->shape);
    const int planes = plq::meas_state_planes(c->q);
    PLQ_TRY(ensure_meas(c, (size_t)planes * m.R * m.W, (size_t)npairs));
    plq::launch_meas_pack(lattice_bytes(c), m, 1, planes, c->d_meas_bits, c->stream);
    plq::launch_meas_pairs(c->d_meas_bits, m, planes, (int)npairs, c->d_meas_counts, c->stream);
    PLQ_TRY(launch_status());
    std::vector<unsigned long long> agree((size_t)npairs);
    PLQ_TRY(read_counts(c, agree));
    correlator_values(c, model, npairs, agree.data(), sums, values);
    return PLQ_OK;
}

int plq_chain_correlator_noalloc(plq_chain *c, int32_t model, int64_t *sums, double *values, size_t len) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    if (!sums && !values) return fail(PLQ_EINVAL, "sums and values are both null: nothing to return");
    std::lock_guard<std::mutex> lk(c->mu);
    int64_t npairs = 0;
    PLQ_TRY(correlator_args(c, model, len, &npairs));
    const plq::MeasGeom m = plq::meas_geom(c->D, c->shape);
    const int planes = plq::meas_state_planes(c->q);
    const int64_t words = (int64_t)planes * m.R * m.W;
    if (words > NOALLOC_PLANE_WORDS || npairs > NOALLOC_COUNTS) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "the configuration needs %lld bit-plane words and %lld counters; the no-allocation "
                 "scratch holds %lld and %lld (use plq_chain_correlator)",
                 (long long)words, (long long)npairs, (long long)NOALLOC_PLANE_WORDS, (long long)NOALLOC_COUNTS);
        return fail(PLQ_EINVAL, buf);
    }
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_TRY(check_cluster(c));
    uint32_t *base = nullptr;
    PLQ_TRY(scratch_address(c->device, &base));
    // One device array serves every chain on the device: hold it from the
    // first launch to the end of the readback.
    std::lock_guard<std::mutex> sk(scratch_mutex(c->device));
    uint32_t *bits = base;
    unsigned long long *counts = reinterpret_cast<unsigned long long *>(base + NOALLOC_PLANE_WORDS);
    plq::launch_meas_clear(counts, npairs, c->stream);
    plq::launch_meas_pack(lattice_bytes(c), m, 1, planes, bits, c->stream);
    plq::launch_meas_pairs(bits, m, planes, (int)npairs, counts, c->stream);
    PLQ_TRY(launch_status());
    // The counters come back into the caller's own 8-byte entries and are
    // turned into the results in place.
    void *host = sums ? (void *)sums : (void *)values;
    PLQ_HIP(hipMemcpyAsync(host, counts, (size_t)npairs * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    correlator_values(c, model, npairs, reinterpret_cast<const unsigned long long *>(host), sums, values);
    return PLQ_OK;
}

int plq_chain_polyakov_loop(plq_chain *c, int32_t dir, double *value) {
    if (!c || !value) return fail(PLQ_EINVAL, "chain or output is null");
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->kind != KIND_LINK) return fail(PLQ_EINVAL, "the Polyakov loop is measured on Z2 gauge chains");
    if (dir < 0 || dir >= c->D) return fail(PLQ_EINVAL, "direction out of range");
    PLQ_HIP(hipSetDevice(c->device));
    const plq::MeasGeom m = plq::meas_geom(c->D, c->shape);
    const size_t plane = (size_t)m.R * m.W;
    PLQ_TRY(ensure_meas(c, (size_t)c->D * plane, 1));
    plq::launch_meas_pack(lattice_bytes(c), m, c->D, 1, c->d_meas_bits, c->stream);
    plq::launch_meas_polyakov(c->d_meas_bits + (size_t)dir * plane, m, dir, c->d_meas_counts, c->stream);
    PLQ_TRY(launch_status());
    std::vector<unsigned long long> odd(1);
    PLQ_TRY(read_counts(c, odd));
    // gauge/mod.rs:193-200: the sum of the +-1 line products over the line count.
    const int64_t lines = c->n_sites / c->shape[dir];
    *value = (double)(lines - 2 * (int64_t)odd[0]) / (double)lines;
    return PLQ_OK;
}

int plq_chain_wilson_rectangles(plq_chain *c, int32_t max_side, double *table, size_t len, int32_t *side) {
    if (!c || !side) return fail(PLQ_EINVAL, "chain or side output is null");
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->kind != KIND_LINK) return fail(PLQ_EINVAL, "Wilson loops are measured on Z2 gauge chains");
    if (max_side < 0) return fail(PLQ_EINVAL, "max_side must be non-negative");
    const int max = wilson_side(c, max_side);
    *side = max;
    if (!table) return PLQ_OK;
    const size_t cells = (size_t)(max + 1) * (max + 1);
    if (len < cells) return fail(PLQ_EINVAL, "the table needs (side + 1)^2 entries");
    for (size_t k = 0; k < cells; ++k) table[k] = 1.0;  // gauge/mod.rs:237-238: the 1.0 anchor
    if (max == 0) return PLQ_OK;
    const int64_t ncomb = (int64_t)max * max * c->D * (c->D - 1);
    if (ncomb > (int64_t(1) << 30)) return fail(PLQ_EINVAL, "too many rectangle sizes");
    PLQ_HIP(hipSetDevice(c->device));
    const plq::MeasGeom m = plq::meas_geom(c->D, c->shape);
    const size_t plane = (size_t)m.R * m.W;
    PLQ_TRY(ensure_meas(c, (size_t)c->D * plane * (size_t)(1 + max), (size_t)ncomb));
    uint32_t *links = c->d_meas_bits;
    uint32_t *lines = c->d_meas_bits + (size_t)c->D * plane;
    plq::launch_meas_pack(lattice_bytes(c), m, c->D, 1, links, c->stream);
    plq::launch_meas_lines(links, m, max, lines, c->stream);
    plq::launch_meas_wilson(lines, m, max, c->d_meas_counts, c->stream);
    PLQ_TRY(launch_status());
    std::vector<unsigned long long> odd((size_t)ncomb);
    PLQ_TRY(read_counts(c, odd));
    // gauge/mod.rs:239-264: the sum of the +-1 loop products over every
    // orientation and site, divided by the placement count D (D - 1) N.
    const int O = c->D * (c->D - 1);
    const int64_t placements = (int64_t)O * c->n_sites;
    for (int r = 1; r <= max; ++r)
        for (int t = 1; t <= max; ++t) {
            int64_t total = placements;
            for (int o = 0; o < O; ++o) total -= 2 * (int64_t)odd[((size_t)(r - 1) * max + (t - 1)) * O + o];
            table[(size_t)r * (max + 1) + t] = (double)total / (double)placements;
        }
    return PLQ_OK;
}
