// plq_block_abi.hip -- C ABI of the learned block proposal: the network
// object (FP32 parameters, Adam state and their BF16 device forms) and the
// chain entry points that train it on a chain's configurations and run the
// block Metropolis-Hastings update on a chain.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plaquette_hip.h"
#include "plq_chain.hpp"
#include "plq_internal.hpp"

struct plq_blocknet {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr;  // FP32 parameters and Adam moments
    float *w1r = nullptr, *w3r = nullptr;            // BF16-rounded W1, W3
    uint16_t *w2f = nullptr, *w2t = nullptr;         // W2 fragments (forward, input gradient)
    uint16_t *w13f = nullptr;                        // W1 and W3 fragments of the pass kernel
    float *grad = nullptr;                           // BN_PARAMS + 1
    float *partials = nullptr;                       // per-workgroup gradient sums, then per-run sums
    float *losses = nullptr;                         // per training step
    size_t losses_cap = 0;
    uint64_t adam_steps = 0;

    plq::BlockNetView view() const {
        plq::BlockNetView n{};
        n.w1 = w1r;
        n.b1 = w + plq::BN_OFF_B1;
        n.b2 = w + plq::BN_OFF_B2;
        n.w3 = w3r;
        n.b3 = w + plq::BN_OFF_B3;
        n.w2f = w2f;
        n.w2t = w2t;
        n.w13f = w13f;
        return n;
    }

    ~plq_blocknet() {
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        for (void *p : {(void *)w, (void *)m, (void *)v, (void *)w1r, (void *)w3r, (void *)w2f, (void *)w2t, (void *)w13f,
                        (void *)grad, (void *)partials, (void *)losses})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

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

int alloc(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PLQ_ENOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " +
                                    hipGetErrorString(e));
    }
    return PLQ_OK;
}

int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PLQ_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return PLQ_OK;
}

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// The chain must be a 2D Ising site chain whose extents are multiples of 16,
// so the 8x8 blocks tile it and the block colouring is consistent under wrap.
int check_chain(const plq_chain *c) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    if (c->kind != KIND_SITE || c->model != PLQ_MODEL_ISING || c->D != 2)
        return fail(PLQ_EINVAL, "the block proposal needs a 2D Ising chain");
    if (c->shape[0] % 16 != 0 || c->shape[1] % 16 != 0)
        return fail(PLQ_EINVAL, "the block proposal needs extents that are multiples of 16 (8x8 blocks, two colours)");
    return PLQ_OK;
}

int ensure_chain_buffers(plq_chain *c) {
    if (!c->d_block_counters) {
        const size_t slots = (size_t)plq::block_pass_groups((int)c->shape[0], (int)c->shape[1]);
        PLQ_TRY(alloc((void **)&c->d_block_counters, slots * sizeof(unsigned long long)));
        PLQ_HIP(hipMemsetAsync(c->d_block_counters, 0, slots * sizeof(unsigned long long), c->stream));
    }
    if (!c->d_block_logits) PLQ_TRY(alloc((void **)&c->d_block_logits, sizeof(float) * (size_t)c->n_sites));
    return PLQ_OK;
}

int pack(plq_blocknet *n, hipStream_t s) {
    plq::launch_block_pack(n->w, n->w1r, n->w3r, n->w2f, n->w2t, n->w13f, s);
    return launch_status();
}

int enqueue_grad(plq_blocknet *n, plq_chain *c) {
    plq::BlockGradArgs a{};
    a.c0 = c->col[0];
    a.c1 = c->col[1];
    a.L0 = (int)c->shape[0];
    a.L1 = (int)c->shape[1];
    a.net = n->view();
    a.partials = n->partials;
    a.grad = n->grad;
    a.stream = c->stream;
    plq::launch_block_grad(a);
    return launch_status();
}

int enqueue_block_sweep(plq_chain *c, plq_blocknet *n) {
    for (int colour = 0; colour < 2; ++colour) {
        plq::BlockPassArgs a{};
        a.c0 = c->col[0];
        a.c1 = c->col[1];
        a.L0 = (int)c->shape[0];
        a.L1 = (int)c->shape[1];
        a.mode = 0;
        a.colour = colour;
        a.seed32 = c->seed32;
        a.bstep = (uint32_t)c->block_sweeps;
        a.beta = c->beta;
        a.j = c->j;
        a.h = c->h;
        a.net = n->view();
        a.logits = c->d_block_logits;
        a.acc_slots = c->d_block_counters;
        a.stream = c->stream;
        plq::launch_block_pass(a);
        PLQ_TRY(launch_status());
    }
    c->block_sweeps += 1;
    return PLQ_OK;
}

}  // namespace

extern "C" {

int plq_blocknet_create(uint64_t seed, plq_blocknet **out) {
    if (!out) return fail(PLQ_EINVAL, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(PLQ_ENODEV, "no HIP device is available");
    plq_blocknet *n = new plq_blocknet();
    auto bail = [&](int code) {
        delete n;
        return code;
    };
    if (hipGetDevice(&n->device) != hipSuccess) return bail(fail(PLQ_EDEVICE, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PLQ_EDEVICE, "hipStreamCreate failed"));
    const size_t P = plq::BN_PARAMS;
    int rc;
    for (float **p : {&n->w, &n->m, &n->v})
        if ((rc = alloc((void **)p, P * sizeof(float))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->w1r, 9 * plq::BN_CHANNELS * sizeof(float))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->w3r, 9 * plq::BN_CHANNELS * sizeof(float))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->w2f, plq::BLOCK_W2_FRAG * sizeof(uint16_t))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->w2t, plq::BLOCK_W2_FRAG * sizeof(uint16_t))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->w13f, plq::BLOCK_W13_FRAG * sizeof(uint16_t))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->grad, (P + 1) * sizeof(float))) != PLQ_OK) return bail(rc);
    if ((rc = alloc((void **)&n->partials, (size_t)(plq::BLOCK_TRAIN_GROUPS + plq::BLOCK_TRAIN_GROUPS / 16) *
                                               plq::BLOCK_PART_STRIDE * sizeof(float))) != PLQ_OK)
        return bail(rc);
    // Glorot-uniform weights (fan_in + fan_out of the 3x3 kernels) from a
    // splitmix64 stream of the seed; zero biases.
    std::vector<float> w(P, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int cnt, double fan) {
        const double lim = std::sqrt(6.0 / fan);
        for (int e = 0; e < cnt; ++e) {
            double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
            w[off + e] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(plq::BN_OFF_W1, 9 * plq::BN_CHANNELS, 9.0 * (1 + plq::BN_CHANNELS));
    fill(plq::BN_OFF_W2, 9 * plq::BN_CHANNELS * plq::BN_CHANNELS, 9.0 * 2 * plq::BN_CHANNELS);
    fill(plq::BN_OFF_W3, 9 * plq::BN_CHANNELS, 9.0 * (plq::BN_CHANNELS + 1));
    if (hipMemcpy(n->w, w.data(), P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, P * sizeof(float)) != hipSuccess || hipMemset(n->v, 0, P * sizeof(float)) != hipSuccess)
        return bail(fail(PLQ_EDEVICE, "parameter upload failed"));
    if ((rc = pack(n, n->stream)) != PLQ_OK) return bail(rc);
    if (hipStreamSynchronize(n->stream) != hipSuccess) return bail(fail(PLQ_EDEVICE, "hipStreamSynchronize failed"));
    *out = n;
    return PLQ_OK;
}

int plq_blocknet_params(plq_blocknet *n, float *out, size_t cnt) {
    if (!n || !out) return fail(PLQ_EINVAL, "net or buffer is null");
    if (cnt != (size_t)plq::BN_PARAMS) return fail(PLQ_EINVAL, "the network has PLQ_BN_PARAMS parameters");
    std::lock_guard<std::mutex> lk(n->mu);
    PLQ_HIP(hipSetDevice(n->device));
    PLQ_HIP(hipDeviceSynchronize());
    PLQ_HIP(hipMemcpy(out, n->w, cnt * sizeof(float), hipMemcpyDeviceToHost));
    return PLQ_OK;
}

int plq_blocknet_set_params(plq_blocknet *n, const float *in, size_t cnt) {
    if (!n || !in) return fail(PLQ_EINVAL, "net or buffer is null");
    if (cnt != (size_t)plq::BN_PARAMS) return fail(PLQ_EINVAL, "the network has PLQ_BN_PARAMS parameters");
    for (size_t e = 0; e < cnt; ++e)
        if (!std::isfinite(in[e])) return fail(PLQ_EINVAL, "parameters must be finite");
    std::lock_guard<std::mutex> lk(n->mu);
    PLQ_HIP(hipSetDevice(n->device));
    PLQ_HIP(hipDeviceSynchronize());  // no pass may still read the old weights
    PLQ_HIP(hipMemcpy(n->w, in, cnt * sizeof(float), hipMemcpyHostToDevice));
    PLQ_HIP(hipMemset(n->m, 0, cnt * sizeof(float)));
    PLQ_HIP(hipMemset(n->v, 0, cnt * sizeof(float)));
    n->adam_steps = 0;
    PLQ_TRY(pack(n, n->stream));
    PLQ_HIP(hipStreamSynchronize(n->stream));
    return PLQ_OK;
}

int plq_blocknet_gradient(plq_blocknet *n, plq_chain *c, float *grad, size_t cnt, double *loss) {
    if (!n || !grad) return fail(PLQ_EINVAL, "net or buffer is null");
    PLQ_TRY(check_chain(c));
    if (cnt != (size_t)plq::BN_PARAMS) return fail(PLQ_EINVAL, "the network has PLQ_BN_PARAMS parameters");
    std::lock_guard<std::mutex> ln(n->mu);
    std::lock_guard<std::mutex> lc(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_TRY(enqueue_grad(n, c));
    std::vector<float> host(cnt + 1);
    PLQ_HIP(hipMemcpyAsync(host.data(), n->grad, (cnt + 1) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(grad, host.data(), cnt * sizeof(float));
    if (loss) *loss = host[cnt];
    return PLQ_OK;
}

int plq_blocknet_train(plq_blocknet *n, plq_chain *c, uint64_t steps, uint64_t sweeps_between, double lr,
                       float *losses) {
    if (!n) return fail(PLQ_EINVAL, "net is null");
    PLQ_TRY(check_chain(c));
    if (!(lr > 0.0) || !std::isfinite(lr)) return fail(PLQ_EINVAL, "the learning rate must be positive and finite");
    std::lock_guard<std::mutex> ln(n->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_HIP(hipDeviceSynchronize());  // no pass may still read the weights this changes
    if (losses && n->losses_cap < steps) {
        if (n->losses) (void)hipFree(n->losses);
        n->losses = nullptr;
        n->losses_cap = 0;
        PLQ_TRY(alloc((void **)&n->losses, steps * sizeof(float)));
        n->losses_cap = steps;
    }
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (uint64_t t = 0; t < steps; ++t) {
        PLQ_TRY(plq_chain_advance(c, sweeps_between));
        std::lock_guard<std::mutex> lc(c->mu);
        PLQ_TRY(enqueue_grad(n, c));
        n->adam_steps += 1;
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)n->adam_steps)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)n->adam_steps)));
        plq::launch_block_adam(n->w, n->m, n->v, n->grad, (float)lr, beta1, beta2, c1, c2, eps, c->stream);
        PLQ_TRY(launch_status());
        PLQ_TRY(pack(n, c->stream));
        if (losses)
            PLQ_HIP(hipMemcpyAsync(n->losses + t, n->grad + plq::BN_PARAMS, sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
    }
    if (losses && steps)
        PLQ_HIP(hipMemcpyAsync(losses, n->losses, steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return PLQ_OK;
}

int plq_blocknet_logits(plq_blocknet *n, plq_chain *c, float *out, size_t cnt) {
    if (!n || !out) return fail(PLQ_EINVAL, "net or buffer is null");
    PLQ_TRY(check_chain(c));
    if (cnt != (size_t)c->n_sites) return fail(PLQ_EINVAL, "the logits buffer holds one float per site");
    std::lock_guard<std::mutex> ln(n->mu);
    std::lock_guard<std::mutex> lc(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_TRY(ensure_chain_buffers(c));
    float *tmp = nullptr;
    PLQ_TRY(alloc((void **)&tmp, cnt * sizeof(float)));
    plq::BlockPassArgs a{};
    a.c0 = c->col[0];
    a.c1 = c->col[1];
    a.L0 = (int)c->shape[0];
    a.L1 = (int)c->shape[1];
    a.mode = 1;
    a.net = n->view();
    a.logits = tmp;
    a.acc_slots = nullptr;
    a.stream = c->stream;
    plq::launch_block_pass(a);
    int rc = launch_status();
    if (rc == PLQ_OK && hipMemcpyAsync(out, tmp, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(PLQ_EDEVICE, "logit readback failed");
    if (rc == PLQ_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PLQ_EDEVICE, "sync failed");
    (void)hipFree(tmp);
    return rc;
}

int plq_blocknet_destroy(plq_blocknet *n) {
    delete n;
    return PLQ_OK;
}

int plq_chain_block_sweeps(plq_chain *c, plq_blocknet *n, uint64_t sweeps) {
    if (!n) return fail(PLQ_EINVAL, "net is null");
    PLQ_TRY(check_chain(c));
    std::lock_guard<std::mutex> lc(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    PLQ_TRY(ensure_chain_buffers(c));
    for (uint64_t t = 0; t < sweeps; ++t) PLQ_TRY(enqueue_block_sweep(c, n));
    return PLQ_OK;
}

int plq_chain_advance_with_blocks(plq_chain *c, plq_blocknet *n, uint64_t steps, uint64_t plain_sweeps) {
    if (!n) return fail(PLQ_EINVAL, "net is null");
    PLQ_TRY(check_chain(c));
    for (uint64_t t = 0; t < steps; ++t) {
        {
            std::lock_guard<std::mutex> lc(c->mu);
            PLQ_HIP(hipSetDevice(c->device));
            PLQ_TRY(ensure_chain_buffers(c));
            PLQ_TRY(enqueue_block_sweep(c, n));
        }
        PLQ_TRY(plq_chain_advance(c, plain_sweeps));
    }
    return PLQ_OK;
}

int plq_chain_block_stats(plq_chain *c, uint64_t *accepted, uint64_t *proposed, uint64_t *block_sweeps) {
    if (!c) return fail(PLQ_EINVAL, "chain is null");
    std::lock_guard<std::mutex> lc(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    unsigned long long acc = 0;
    if (c->d_block_counters) {
        std::vector<unsigned long long> slots((size_t)plq::block_pass_groups((int)c->shape[0], (int)c->shape[1]));
        PLQ_HIP(hipMemcpyAsync(slots.data(), c->d_block_counters, slots.size() * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
        PLQ_HIP(hipStreamSynchronize(c->stream));
        for (unsigned long long v : slots) acc += v;
    }
    if (accepted) *accepted = acc;
    if (proposed) *proposed = c->block_sweeps * (uint64_t)(c->n_sites / 64);
    if (block_sweeps) *block_sweeps = c->block_sweeps;
    return PLQ_OK;
}

int plq_chain_block_logits(plq_chain *c, float *out, size_t cnt) {
    PLQ_TRY(check_chain(c));
    if (!out) return fail(PLQ_EINVAL, "buffer is null");
    if (cnt != (size_t)c->n_sites) return fail(PLQ_EINVAL, "the logits buffer holds one float per site");
    std::lock_guard<std::mutex> lc(c->mu);
    PLQ_HIP(hipSetDevice(c->device));
    if (!c->d_block_logits) return fail(PLQ_EINVAL, "no block sweep has run on this chain");
    PLQ_HIP(hipMemcpyAsync(out, c->d_block_logits, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PLQ_HIP(hipStreamSynchronize(c->stream));
    return PLQ_OK;
}

}  // extern "C"
