"""Learned block proposal on the GPU against the CPU oracle.

The network (BF16 matrix-core products with FP32 accumulation) is compared
with the oracle's FP64 evaluation of the same BF16-rounded weights and
activations, within a tolerance stated per test.  The Metropolis-Hastings
pass is compared bit for bit: given the logits the device used, the
oracle's IEEE-single restatement must reproduce every accepted block.
"""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BETA_C = 0.4406868


def _params(seed, scale=0.25):
    return (np.random.default_rng(seed).standard_normal(O.BN_PARAMS) * scale).astype(np.float32)


def _chain(gpu, shape, beta, seed, start, j=1.0, h=0.0):
    return gpu.ising_gpu_chain(list(shape), j, h, beta, seed, start)


def _hot(n, seed):
    return np.random.default_rng(seed).integers(0, 2, n).astype(np.uint8)


@pytest.mark.parametrize("shape", [(32, 32), (64, 48)])
def test_logits_match_the_oracle_network(gpu, shape):
    # Tolerance: BF16 activations can round the other way when the sums are
    # formed in FP32 (device) rather than FP64 (oracle); one such flip (2^-8
    # relative in one activation) moves a logit by well under 0.03, and flips
    # are rare, so the mean error stays below 1e-4 (measured: max 2.5e-6).
    net = gpu.BlockNet(1)
    params = _params(2)
    net.params = params
    n = shape[0] * shape[1]
    st = _hot(n, 3)
    c = _chain(gpu, shape, BETA_C, 4, st)
    got = net.logits(c)
    want = O.blocknet_logits(params, shape, st, round_bf16=True)
    err = np.abs(got - want)
    print("logit error max", err.max(), "mean", err.mean(), "scale", np.abs(want).mean())
    assert err.max() < 0.03 and err.mean() < 1e-4
    c.close()
    net.close()


def test_logits_depend_on_the_ring_only(gpu):
    net = gpu.BlockNet(5)
    L = 32
    st = _hot(L * L, 6)
    c = _chain(gpu, (L, L), BETA_C, 1, st)
    base = net.logits(c)
    st2 = st.copy()
    for y in range(8, 16):
        for x in range(16, 24):
            st2[y * L + x] ^= 1
    c.write(st2)
    b = 1 * (L // 8) + 2
    assert np.array_equal(net.logits(c)[b], base[b])
    c.close()
    net.close()


@pytest.mark.parametrize("shape,beta,j,h,scale", [((32, 32), BETA_C, 1.0, 0.0, 0.25), ((64, 48), 0.3, 0.8, 0.1, 0.5),
                                                  ((16, 32), 0.6, 1.0, -0.2, 0.05)])
def test_block_pass_is_bit_exact_given_the_logits(gpu, shape, beta, j, h, scale):
    net = gpu.BlockNet(7)
    net.params = _params(8, scale)
    n = shape[0] * shape[1]
    c = _chain(gpu, shape, beta, 99, _hot(n, 9), j=j, h=h)
    seed32 = O.fold_seed(99)
    total = 0
    for t in range(12):
        before = c.read()
        c.block_sweeps(net, 1)
        lg = c.block_logits()
        after = c.read()
        st, a0 = O.block_mh_pass(shape, before, lg, 0, seed32, t, beta, j, h)
        st, a1 = O.block_mh_pass(shape, st, lg, 1, seed32, t, beta, j, h)
        assert np.array_equal(st, after), f"sweep {t}: {(st != after).sum()} sites differ"
        total += a0 + a1
    acc, prop, done = c.block_stats()
    assert done == 12 and prop == 12 * n // 64 and acc == total
    c.close()
    net.close()


def test_block_pass_colours_see_the_updated_other_colour(gpu):
    # The colour-1 logits are computed after the colour-0 pass: they match
    # the oracle network on the intermediate configuration.
    shape = (32, 32)
    net = gpu.BlockNet(11)
    params = np.zeros(O.BN_PARAMS, np.float32)
    params[9856] = 0.0  # uniform proposals: many colour-0 blocks change
    rng = np.random.default_rng(12)
    params[:] = rng.standard_normal(O.BN_PARAMS).astype(np.float32) * 0.05
    net.params = params
    c = _chain(gpu, shape, 0.05, 3, _hot(1024, 13))
    before = c.read()
    c.block_sweeps(net, 1)
    lg = c.block_logits()
    mid, _ = O.block_mh_pass(shape, before, lg, 0, O.fold_seed(3), 0, 0.05, 1.0, 0.0)
    want = O.blocknet_logits(params, shape, mid)
    colour1 = [(by * 4 + bx) for by in range(4) for bx in range(4) if (bx + by) & 1]
    assert np.abs(lg[colour1] - want[colour1]).max() < 0.03
    c.close()
    net.close()


def test_gradient_matches_the_oracle(gpu):
    # Tolerance: the device sums in FP32 where the oracle sums in FP64, and a
    # BF16 rounding of an activation or of the layer-2 gradient can flip;
    # relative L2 error below 1e-3 (measured: 6e-8).
    shape = (32, 32)
    net = gpu.BlockNet(21)
    params = _params(22)
    net.params = params
    st = _hot(1024, 23)
    c = _chain(gpu, shape, BETA_C, 5, st)
    loss, g = net.gradient(c)
    wl, wg = O.blocknet_loss_grad(params, shape, st, round_bf16=True)
    rel = np.linalg.norm(g - wg) / np.linalg.norm(wg)
    print("loss", loss, wl, "grad rel err", rel)
    assert abs(loss - wl) < 1e-5 * abs(wl)
    assert rel < 1e-3
    for lo, hi in ((0, 288), (288, 320), (320, 9536), (9536, 9568), (9568, 9856), (9856, 9857)):
        assert np.linalg.norm(g[lo:hi] - wg[lo:hi]) <= 1e-3 * np.linalg.norm(wg[lo:hi]) + 1e-7, (lo, hi)
    c.close()
    net.close()


def test_training_lowers_the_loss_and_is_deterministic(gpu):
    outs = []
    for _ in range(2):
        net = gpu.BlockNet(31)
        c = _chain(gpu, (64, 64), BETA_C, 8, _hot(4096, 32))
        c.advance(200)
        losses = net.train(c, 150, 2, 3e-3)
        outs.append((losses, net.params))
        c.close()
        net.close()
    (l1, p1), (l2, p2) = outs
    assert np.array_equal(p1, p2) and np.array_equal(l1, l2)
    assert l1[-30:].mean() < l1[:10].mean() - 0.02, (l1[:10].mean(), l1[-30:].mean())


def test_block_update_samples_the_product_target_at_zero_coupling(gpu):
    # J = 0: pi is a product with <s> = tanh(beta h) whatever the proposal;
    # the constant logit 2 beta h + 0.3 exercises the q-ratio (without it
    # the chain would sample pi q, with <s> = 0.64).
    beta, h = 1.0, 0.3
    params = np.zeros(O.BN_PARAMS, np.float32)
    params[9856] = 2 * beta * h + 0.3
    net = gpu.BlockNet(0)
    net.params = params
    p_up = math.exp(beta * h) / (2 * math.cosh(beta * h))
    start = (np.random.default_rng(41).random(64 * 64) >= p_up).astype(np.uint8)
    c = _chain(gpu, (64, 64), beta, 42, start, j=0.0, h=h)
    ms = []
    for _ in range(400):
        c.block_sweeps(net, 5)
        ms.append(c.measure().spin_sum / 4096)
    acc, prop, _ = c.block_stats()
    m = float(np.mean(ms))
    print("acceptance", acc / prop, "<s>", m)
    assert 0.05 < acc / prop < 1.0
    assert abs(m - math.tanh(beta * h)) < 0.01
    c.close()
    net.close()


# -- the mixed chain (config 5's sampler): block sweeps between plain sweeps -----

MIXED_CASES = [
    # shape, beta, J, h, plain sweeps per step, kernel; 32^2 takes the
    # one-workgroup small chain for its plain sweeps, 64x48 the colour passes
    ((32, 32), BETA_C, 1.0, 0.0, 1, O.KERNEL_METROPOLIS),
    ((64, 48), 0.35, 0.8, 0.1, 2, O.KERNEL_HEAT_BATH),
    ((32, 64), 0.6, 1.0, -0.15, 3, O.KERNEL_METROPOLIS),
    # config 5's lattice: 4,096 blocks, the uniform-parity colour passes
    ((512, 512), BETA_C, 1.0, 0.0, 1, O.KERNEL_METROPOLIS),
]


@pytest.mark.parametrize("shape,beta,j,h,plain,kernel", MIXED_CASES)
def test_mixed_chain_is_bit_exact_given_the_logits(gpu, shape, beta, j, h, plain, kernel):
    # One advance_with_blocks step at a time: the oracle runs the block pass
    # of both colours with the logits the device used, then the plain
    # sweeps with the chain's own sweep counter.
    net = gpu.BlockNet(51)
    net.params = _params(52, 0.3)
    n = shape[0] * shape[1]
    start = _hot(n, 53)
    c = gpu.ising_gpu_chain(list(shape), j, h, beta, 77, start, kernel=kernel)
    d = O.local_desc(list(shape), kernel=kernel, beta=beta, j=j, h=h, seed=77)
    seed32 = O.fold_seed(77)
    st, sweep, total = start.copy(), 0, 0
    for t in range(12):
        c.advance_with_blocks(net, 1, plain)
        lg = c.block_logits()
        st, a0 = O.block_mh_pass(shape, st, lg, 0, seed32, t, beta, j, h)
        st, a1 = O.block_mh_pass(shape, st, lg, 1, seed32, t, beta, j, h)
        st = O.local_sweeps(d, st, sweep, plain)
        sweep += plain
        total += a0 + a1
        got = c.read()
        assert np.array_equal(got, st), f"step {t}: {(got != st).sum()} sites differ"
    acc, prop, done = c.block_stats()
    assert done == 12 and acc == total and c.sweeps_done == sweep
    c.close()
    net.close()


@pytest.mark.parametrize("shape,beta,j,h,plain,kernel", MIXED_CASES)
def test_mixed_chain_of_many_steps_in_one_call_is_bit_exact(gpu, shape, beta, j, h, plain, kernel):
    # Many steps per call, with recording on.  A bias-only network makes
    # every logit exactly b3 on both sides, so the oracle needs no device
    # logits; the block sweep keys (block sweep counter) and the plain sweep
    # keys (chain sweep counter) must interleave as on the device.
    b3 = np.float32(0.375 if h >= 0 else -0.25)
    params = np.zeros(O.BN_PARAMS, np.float32)
    params[O.BN_PARAMS - 1] = b3
    net = gpu.BlockNet(0)
    net.params = params
    n = shape[0] * shape[1]
    start = _hot(n, 61)
    small_beta = 0.12
    c = gpu.ising_gpu_chain(list(shape), j, h, small_beta, 78, start, kernel=kernel)
    c.record(plain, 1024)
    d = O.local_desc(list(shape), kernel=kernel, beta=small_beta, j=j, h=h, seed=78)
    seed32 = O.fold_seed(78)
    lg = np.full((n // 64, 64), b3, np.float32)
    st, sweep, total, recs = start.copy(), 0, 0, []
    for steps in (5, 7):
        c.advance_with_blocks(net, steps, plain)
        for _ in range(steps):
            t = len(recs)
            st, a0 = O.block_mh_pass(shape, st, lg, 0, seed32, t, small_beta, j, h)
            st, a1 = O.block_mh_pass(shape, st, lg, 1, seed32, t, small_beta, j, h)
            st = O.local_sweeps(d, st, sweep, plain)
            sweep += plain
            total += a0 + a1
            recs.append(O.ising_sums(list(shape), st))
        got = c.read()
        assert np.array_equal(got, st), f"after {len(recs)} steps: {(got != st).sum()} sites differ"
    acc, prop, done = c.block_stats()
    assert done == 12 and acc == total and total > 0
    series = c.series()
    assert [(o.bond_sum, o.spin_sum) for o in series] == recs
    c.close()
    net.close()


def test_block_sweep_leaves_the_boltzmann_law_invariant_on_the_device(gpu):
    # As the oracle's check at J != 0: a block sweep applied to equilibrium
    # configurations of the plain chain must not move <E>.  The network is
    # trained briefly on the chain, so that blocks are accepted often enough
    # for the paired shift to be sharp.  Tolerance: 4.5 standard errors of
    # the mean paired shift.
    shape, beta, n = (32, 32), BETA_C, 1024
    net = gpu.BlockNet(71)
    c = _chain(gpu, shape, beta, 72, _hot(n, 73))
    c.advance(500)
    net.train(c, 300, 2, 3e-3)
    shifts = []
    for _ in range(3000):
        c.advance(4)
        e0 = c.measure().energy
        c.block_sweeps(net, 1)
        shifts.append((c.measure().energy - e0) / n)
    acc, prop, _ = c.block_stats()
    shifts = np.array(shifts)
    err = shifts.std() / math.sqrt(shifts.size)
    print("acceptance", acc / prop, "shift", shifts.mean(), "+-", err)
    assert acc / prop > 0.01
    assert abs(shifts.mean()) < 4.5 * err
    c.close()
    net.close()


def test_mixed_chain_matches_kaufman(gpu):
    # The mixed chain (a block sweep, then one plain sweep) at 32^2 and
    # beta_c against the exact finite-torus E/N and specific heat; 4.5
    # standard errors on the tau_int-inflated and jackknifed errors.
    import exact_ising2d as X
    from plaquette_amd import statistics as S

    L, beta = 32, BETA_C
    n = L * L
    net = gpu.BlockNet(81)
    c = _chain(gpu, (L, L), beta, 82, _hot(n, 83))
    c.advance(1000)
    net.train(c, 300, 2, 3e-3)
    c.advance_with_blocks(net, 2000, 1)
    steps = 100000
    c.record(1, steps)
    c.advance_with_blocks(net, steps, 1)
    e = np.array([o.energy for o in c.series()])
    acc, prop, _ = c.block_stats()
    c.close()
    net.close()
    assert e.size == steps and acc / prop > 0.01
    est = S.reduce(e / n)
    want = X.energy_per_site(L, L, beta)
    assert abs(est.mean - want) < 4.5 * est.stderr, (est.mean, want, est.stderr)
    cv = S.specific_heat(e, beta, n)
    want_c = X.specific_heat(L, L, beta)
    assert abs(cv.value - want_c) < 4.5 * cv.stderr, (cv.value, want_c, cv.stderr)


def test_trained_network_beats_the_untrained_one(gpu):
    # 64^2 at beta_c.  The same initial weights before and after 600 Adam
    # steps on the plain chain's configurations (logits then scaled by 1.5,
    # the calibrated sharpness): the trained network's blocks are accepted
    # about 16% of the time and block sweeps alone move E between almost
    # every pair of measurements, while the untrained network's proposals
    # are almost never accepted (7e-5), so its block sweeps barely move the
    # chain.  Block sweeps alone decorrelate E slowly (tau_int about 1,000
    # block sweeps, printed), and interleaved with plain sweeps the trained
    # network does not measurably lower tau_int of E or |M| per step at this
    # size; the bench reports those numbers.
    from plaquette_amd import statistics as S

    L, n = 64, 64 * 64
    c = _chain(gpu, (L, L), BETA_C, 91, _hot(n, 92))
    c.advance(3000)
    trained = gpu.BlockNet(93)
    untrained = gpu.BlockNet(93)
    trained.train(c, 600, 5, 3e-3)
    p = trained.params
    p[9568:9857] *= 1.5
    trained.params = p
    st0 = c.read()
    c.close()
    out = {}
    for name, net in (("trained", trained), ("untrained", untrained)):
        # block sweeps alone, E measured every 10 of them
        ch = _chain(gpu, (L, L), BETA_C, 94, st0)
        es = []
        for _ in range(2000):
            ch.block_sweeps(net, 10)
            es.append(ch.measure().energy / n)
        acc, prop, _ = ch.block_stats()
        out[name] = (acc / prop, np.array(es))
        ch.close()
    trained.close()
    untrained.close()
    acc_t, e_t = out["trained"]
    acc_u, e_u = out["untrained"]
    est = S.reduce(e_t)
    print("acceptance", acc_t, acc_u, "tau_int(E) of block sweeps, in units of 10:", est.tau_int)
    moved_t, moved_u = (np.diff(e_t) != 0).mean(), (np.diff(e_u) != 0).mean()
    msd_t, msd_u = np.mean(np.diff(e_t) ** 2), np.mean(np.diff(e_u) ** 2)
    print("E changed between measurements:", moved_t, moved_u, "mean squared change:", msd_t, msd_u)
    assert acc_t > 0.05 and acc_u < 1e-3     # measured 0.16 and 7e-5
    assert moved_t > 0.9 and moved_u < 0.1   # the untrained network almost never moves the chain
    assert msd_t > 10 * msd_u                # and the trained one moves E far more per block sweep
