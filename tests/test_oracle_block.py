"""Pins of the CPU oracle of the learned block proposal (no GPU).

The block proposal has no reference counterpart, so its oracle is pinned by
what the mathematics fixes: the exponential against libm, the network
gradient against finite differences of the loss, the logits' dependence on
the boundary ring only, tanh's oddness, the all-zero network in closed form,
and the Metropolis-Hastings pass against exactly solvable targets (J = 0,
where the Boltzmann weight is a product over sites).
"""
import math

import numpy as np
import pytest

import oracle as O


def test_expf_is_within_one_and_a_half_ulp_of_libm():
    xs = np.concatenate([np.linspace(-87.0, 88.0, 4001), np.random.default_rng(0).uniform(-30, 30, 4000)])
    worst = 0.0
    for x in xs.astype(np.float32):
        want = math.exp(float(x))
        worst = max(worst, abs(O.expf(float(x)) - want) / float(np.spacing(np.float32(want))))
    assert worst < 1.5
    assert O.expf(0.0) == 1.0
    assert O.expf(-88.0) == 0.0 and O.expf(89.0) == math.inf
    assert O.expf(1.0) == float(np.float32(math.e))


def test_bf16_rounds_to_nearest_even():
    assert O.bf16(1.0) == 1.0
    assert O.bf16(1.0 + 2.0 ** -8) == 1.0            # tie, to even
    assert O.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, to even (up)
    assert O.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert O.bf16(-2.5) == -2.5


def _params(seed, scale=0.25, zero_bias=False):
    p = (np.random.default_rng(seed).standard_normal(O.BN_PARAMS) * scale).astype(np.float32)
    if zero_bias:
        p[288:320] = 0
        p[9536:9568] = 0
        p[9856] = 0
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = _params(4)
    shape = (16, 16)
    st = rng.integers(0, 2, 256).astype(np.uint8)
    _, g = O.blocknet_loss_grad(params, shape, st, round_bf16=False)
    # every layer's weights and biases
    idx = list(rng.choice(O.BN_PARAMS, 30, replace=False)) + [0, 287, 288, 319, 320, 9535, 9536, 9567, 9568,
                                                                9855, 9856]
    for e in idx:
        p1, p2 = params.copy(), params.copy()
        p1[e] += np.float32(1e-3)
        p2[e] -= np.float32(1e-3)
        l1, _ = O.blocknet_loss_grad(p1, shape, st, round_bf16=False)
        l2, _ = O.blocknet_loss_grad(p2, shape, st, round_bf16=False)
        fd = (l1 - l2) / (float(p1[e]) - float(p2[e]))
        assert abs(g[e] - fd) <= 1e-4 * abs(fd) + 1e-9, (e, g[e], fd)


def test_rounded_gradient_stays_close_to_the_exact_one():
    rng = np.random.default_rng(5)
    params = _params(6)
    st = rng.integers(0, 2, 32 * 16).astype(np.uint8)
    l0, g0 = O.blocknet_loss_grad(params, (32, 16), st, round_bf16=False)
    l1, g1 = O.blocknet_loss_grad(params, (32, 16), st, round_bf16=True)
    assert abs(l0 - l1) < 1e-2 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)


def test_logits_depend_on_the_boundary_ring_only():
    rng = np.random.default_rng(7)
    params = _params(8)
    L = 32
    st = rng.integers(0, 2, L * L).astype(np.uint8)
    base = O.blocknet_logits(params, (L, L), st)
    b = 1 * (L // 8) + 2  # block (bx, by) = (2, 1)
    # interior sites and the diagonal (same-colour) neighbours' corners
    st2 = st.copy()
    for y in range(8, 16):
        for x in range(16, 24):
            st2[y * L + x] ^= 1
    st2[7 * L + 15] ^= 1
    st2[16 * L + 24] ^= 1
    assert np.array_equal(O.blocknet_logits(params, (L, L), st2)[b], base[b])
    # a ring site changes them
    st3 = st.copy()
    st3[7 * L + 20] ^= 1
    assert not np.array_equal(O.blocknet_logits(params, (L, L), st3)[b], base[b])


def test_every_interior_site_sees_the_ring():
    # Receptive field: layer 1 reaches depth 0 from the ring, the dilated
    # middle layer depths 0-2, layer 3 depths 0-3, so every one of the 64
    # logits -- the centre 2x2 included -- moves when some ring spin flips.
    # With an undilated middle layer the centre's logits would stay at the
    # bias for every ring.
    params = _params(18)
    L = 16
    st = np.random.default_rng(19).integers(0, 2, L * L).astype(np.uint8)
    base = O.blocknet_logits(params, (L, L), st, round_bf16=False)[0]
    moved = np.zeros(64, bool)
    ring = [(-1, x) for x in range(8)] + [(8, x) for x in range(8)] + [(y, -1) for y in range(8)] + \
           [(y, 8) for y in range(8)]
    for y, x in ring:
        st2 = st.copy()
        st2[(y % L) * L + (x % L)] ^= 1
        moved |= np.abs(O.blocknet_logits(params, (L, L), st2, round_bf16=False)[0] - base) > 1e-12
    assert moved.all(), np.flatnonzero(~moved)


def test_unbiased_network_is_odd_under_global_flip():
    rng = np.random.default_rng(9)
    params = _params(10, zero_bias=True)
    st = rng.integers(0, 2, 256).astype(np.uint8)
    a = O.blocknet_logits(params, (16, 16), st, round_bf16=False)
    b = O.blocknet_logits(params, (16, 16), 1 - st, round_bf16=False)
    assert np.allclose(a, -b, atol=1e-12)


def test_zero_network_is_the_bias_in_closed_form():
    params = np.zeros(O.BN_PARAMS, np.float32)
    params[9856] = 0.75
    st = np.random.default_rng(11).integers(0, 2, 256).astype(np.uint8)
    lg = O.blocknet_logits(params, (16, 16), st)
    assert np.all(lg == 0.75)
    loss, g = O.blocknet_loss_grad(params, (16, 16), st)
    up = float(np.mean(st == 0))
    assert abs(loss - (math.log1p(math.exp(0.75)) - up * 0.75)) < 1e-12
    assert abs(g[9856] - (1 / (1 + math.exp(-0.75)) - up)) < 1e-12


def _run_mh(shape, logits, beta, j, h, passes, seed32=0x1234567, start=None):
    L0, L1 = shape
    st = np.zeros(L0 * L1, np.uint8) if start is None else start.copy()
    acc = mags = 0.0
    for t in range(passes):
        for colour in (0, 1):
            st, a = O.block_mh_pass(shape, st, logits, colour, seed32, t, beta, j, h)
            acc += a
        mags += float(np.mean(1.0 - 2.0 * st))
    return st, acc / (passes * (L0 // 8) * (L1 // 8)), mags / passes


@pytest.mark.parametrize("offset", [0.0, 0.3, -0.3])
def test_block_mh_samples_the_product_target_at_zero_coupling(offset):
    # J = 0: pi(s) is a product with <s> = tanh(beta h), whatever the
    # proposal; logits away from the exact conditional 2 beta h exercise the
    # q-ratio: without it the chain samples pi * q, whose <s> is 0.64 at
    # offset 0.3 against tanh(0.3) = 0.29.  The start is drawn from pi: an
    # independence sampler leaves an atypical start (all +1) slowly.
    beta, h = 1.0, 0.3
    noise = np.random.default_rng(12).normal(0, 0.3, 4 * 64).astype(np.float32)
    logits = (np.float32(2 * beta * h + offset) + noise).astype(np.float32)
    p_up = math.exp(beta * h) / (2 * math.cosh(beta * h))
    start = (np.random.default_rng(13).random(256) >= p_up).astype(np.uint8)
    _, acc, m = _run_mh((16, 16), logits, beta, 0.0, h, 3000, start=start)
    assert 0.005 < acc <= 1.0
    assert abs(m - math.tanh(beta * h)) < 0.02, (m, acc)


def test_exact_conditional_proposal_is_always_accepted():
    beta, h = 0.7, -0.2
    logits = np.full(4 * 64, np.float32(2 * beta * h), np.float32)
    _, acc, _ = _run_mh((16, 16), logits, beta, 0.0, h, 300)
    assert acc > 0.999


# Stationarity at J != 0.  pi must be invariant under one block sweep: from
# (nearly) independent equilibrium samples of the plain heat-bath chain, one
# block sweep (both colours) must leave <E> unchanged.  The paired shift is a
# sharp test: a double-counted ring bond, a dropped q-ratio or a dropped
# in-block bond move it by 10 to 130 standard errors.  Logits: the ring-only
# mean-field logits of tests/block_logits.py.  (A chain of block sweeps alone
# is not used: with nearly deterministic proposals it has a slow mode of small
# amplitude that the tau_int window misses, so its error bars are not
# trustworthy.)  The moved samples are also compared with Kaufman's exact
# E/N of the torus where J = 1, h = 0.
@pytest.mark.parametrize("shape,beta,j,h,n", [((16, 16), 0.6, 1.0, 0.0, 3000),
                                              ((32, 32), 0.4406868, 1.0, 0.0, 1500),
                                              ((16, 32), 0.5, 0.8, 0.15, 2000)])
def test_block_sweep_leaves_the_boltzmann_law_invariant(shape, beta, j, h, n):
    import block_logits as B
    import exact_ising2d as X

    N = shape[0] * shape[1]
    d = O.local_desc(list(shape), kernel=O.KERNEL_HEAT_BATH, beta=beta, j=j, h=h, seed=n)
    st = O.local_sweeps(d, np.random.default_rng(n).integers(0, 2, N).astype(np.uint8), 0, 300)
    shifts, after, acc, t = [], [], 0, 300
    for k in range(n):
        st = O.local_sweeps(d, st, t, 10)
        t += 10
        b0, m0 = O.ising_sums(list(shape), st)
        x = st
        for colour in (0, 1):
            x, a = O.block_mh_pass(shape, x, B.mean_field_logits(shape, x, beta, j, h), colour, 0xB10C + n, k,
                                   beta, j, h)
            acc += a
        b1, m1 = O.ising_sums(list(shape), x)
        shifts.append((-j * (b1 - b0) - h * (m1 - m0)) / N)
        after.append((-j * b1 - h * m1) / N)
    shifts = np.array(shifts)
    rate = acc / (n * (shape[0] // 8) * (shape[1] // 8))
    assert rate > 0.1, rate
    err = shifts.std() / math.sqrt(n)
    assert abs(shifts.mean()) < 4.5 * err, (shifts.mean(), err, rate)
    if j == 1.0 and h == 0.0:
        est = O.reduce(np.array(after))
        want = X.energy_per_site(shape[0], shape[1], beta)
        assert abs(est.mean - want) < 5 * est.stderr, (est.mean, want, est.stderr)
