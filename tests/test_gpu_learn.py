"""The learned proposal on the GPU: descriptor, network forward and gradients
against the CPU oracle (oracle/net_oracle.py); the sampler with learned moves
keeps the Boltzmann distribution (two-particle closed form, equilibrium
against the reference program's runs); training lowers the loss."""
import math
import os

import numpy as np
import pytest

import net_oracle as no
from conftest import golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _rotations(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    q[np.linalg.det(q) < 0, :, 0] *= -1
    return q


def test_descriptor_matches_oracle(pp, oracle):
    """BF16 descriptors of the GPU equal the oracle's (a value may differ by
    one BF16 step where the double-precision inputs round differently)."""
    for seed, (n, rho) in enumerate([(60, 0.5), (200, 0.8), (30, 0.05)]):
        r, O, box = oracle.generate(n, min(rho, 0.7), seed=seed + 1)
        idx = np.arange(n, dtype=np.int32)
        got = pp.net_descriptor(r, O, box, idx).astype(np.float64)
        want = np.array([no.descriptor(r, O, box, i) for i in range(n)])
        diff = got != want
        assert diff.mean() < 1e-3, (seed, diff.sum())
        rel = np.abs(got - want)[diff] / np.maximum(np.abs(want[diff]), 1e-30)
        assert (rel <= 2 ** -7).all()
    # an isolated particle: all zeros
    x = pp.net_descriptor(np.array([[1., 1., 1.], [5., 5., 5.]]), np.array([np.eye(3)] * 2), np.full(3, 10.), [0])
    assert not x.any()


def _rand_inputs(rng, m):
    x = np.zeros((m, no.IN))
    for t in range(no.NB):
        x[:, t * 7:t * 7 + 3] = rng.uniform(-1.5, 1.5, (m, 3))
        q = rng.normal(size=(m, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        q *= np.sign(q[:, 3:4])
        x[:, t * 7 + 3:t * 7 + 7] = q
    # fewer neighbours in some rows
    for i in range(m):
        k = rng.integers(0, no.NB + 1)
        x[i, k * 7:] = 0
    return x


def test_forward_matches_oracle(pp):
    rng = np.random.default_rng(1)
    with pp.ProposalNet(seed=5) as net:
        w, b = net.get()
        for m in (1, 17, 300):
            x = _rand_inputs(rng, m)
            got = net.forward(x).astype(np.float64)
            want = no.forward(w, b, x)
            err = np.abs(got - want)
            assert err.max() < 2e-2 * (1 + np.abs(want).max()), (m, err.max())
            assert np.linalg.norm(got - want) < 5e-3 * np.linalg.norm(want)
        # weights set from the host, read back unchanged; hidden layers
        # scaled up so that every layer matters
        w2 = w * 3.0
        b2 = b + rng.normal(scale=0.1, size=b.shape).astype(np.float32)
        net.set(w2, b2)
        w3, b3 = net.get()
        assert np.array_equal(w3, w2) and np.array_equal(b3, b2)
        x = _rand_inputs(rng, 64)
        got = net.forward(x).astype(np.float64)
        want = no.forward(w3, b3, x)
        assert np.linalg.norm(got - want) < 1e-2 * np.linalg.norm(want)


def test_loss_and_gradients_match_oracle(pp):
    rng = np.random.default_rng(2)
    with pp.ProposalNet(seed=9) as net:
        w, b = net.get()
        # a last layer of larger weights: gradients through every output
        w = w.copy()
        w[no.W_OFF[3]:] *= 10
        net.set(w, b)
        for m in (32, 45, 1000):
            x = _rand_inputs(rng, m)
            t = np.hstack([rng.normal(scale=0.05, size=(m, 3)), rng.normal(size=(m, 3))])
            t[:, 3:] /= np.linalg.norm(t[:, 3:], axis=1, keepdims=True)
            loss, gw, gb = net.loss_grad(x, t)
            lo, ow, ob = no.loss_and_grads(w, b, x, t.astype(np.float32).astype(np.float64))
            assert abs(loss - lo) < 1e-3 * (1 + abs(lo)), (m, loss, lo)
            for g, o in ((gw, ow), (gb, ob)):
                g = g.astype(np.float64)
                cos = g @ o / (np.linalg.norm(g) * np.linalg.norm(o))
                assert cos > 0.999, (m, cos)
                assert abs(np.linalg.norm(g) / np.linalg.norm(o) - 1) < 0.02
            # padded rows / columns stay without gradient
            assert not gw[no.W_OFF[0] + no.IN * no.N[0]:no.W_OFF[1]].any()
            assert not gw[no.W_OFF[3]:].reshape(no.H, no.OUT_PAD)[:, no.OUT:].any()


def _skewed_net(pp, seed, drift=0.08, spread=(0.4, 0.05, 0.1), kappa=20.0, hidden_scale=1.0):
    """A network whose proposal is far from symmetric: a drift along the body
    x axis, anisotropic spread, an axis concentrated around body z, and (with
    hidden_scale > 0) a dependence on the neighbourhood."""
    net = pp.ProposalNet(seed=seed)
    w, b = net.get()
    w = w.copy()
    b = b.copy()
    w[no.W_OFF[3]:] *= hidden_scale
    bl = b[no.B_OFF[3]:]
    bl[0:3] = [drift, 0.0, 0.0]
    bl[3:6] = np.log(spread)
    for c in range(no.MIX):
        bl[6 + 5 * c:6 + 5 * c + 5] = [0.0, 0.0, 1.0, math.log(kappa), 0.0]
    net.set(w, b)
    return net


@pytest.mark.parametrize("hidden_scale", [0.0, 3.0])
def test_two_particle_bonding_probability_with_learned_moves(pp, hidden_scale):
    """The closed form of the two-particle test holds when half the colour
    passes use a skewed learned proposal: the Hastings ratio is exact."""
    delta, cosmax, T, L = 0.12, 0.92, 0.125, 5.0
    Vb = 16 * math.pi * (delta ** 3 + 3 * delta ** 2 + 3 * delta) * (1 - cosmax) ** 2 / 3
    wgt = Vb * math.exp(1 / T)
    p_exact = wgt / (wgt + L ** 3 - 4 * math.pi / 3 - Vb)
    R = 256
    rng = np.random.default_rng(0)
    rs, Os, bs = [], [], []
    for _ in range(R):
        rs.append(np.array([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]]) + rng.uniform(0, 0.3, (2, 3)))
        Os.append(np.array([np.eye(3), np.eye(3)]))
        bs.append(np.array([L, L, L]))
    prm = [dict(T=T, disp_max=0.8, theta_max=1.5)] * R
    # kappa 2.5: reverse axis -a is e^{-5 cos} as likely as a, so some of the
    # skewed moves are still accepted
    net = _skewed_net(pp, 3, kappa=2.5, hidden_scale=hidden_scale)
    with pp.System.batch(rs, Os, bs, prm, seed=23) as s:
        s.attach(net, learned_fraction=0.5, theta_max=1.5)
        s.do_ensemble(5000)
        samples = []
        for _ in range(100):
            s.do_ensemble(200)
            samples.append(s.counters_all()["bonds"])
        tried, acc = s.learn_counters(0)
    net.close()
    assert tried > 0 and 0 < acc < tried
    per_rep = np.array(samples).mean(axis=0)
    m, se = per_rep.mean(), per_rep.std(ddof=1) / np.sqrt(R)
    assert abs(m - p_exact) < 4 * se + 2e-3, (m, se, p_exact)


def _reference_start(pp, tmp_path, n, rho):
    import subprocess
    gen = os.path.join(ROOT, "patchyparticles_amd", "bin", "generator_hip")
    subprocess.run([gen, str(n), str(rho)], cwd=tmp_path, check=True, capture_output=True)
    _, box, r, O = pp.read_rrr(str(tmp_path / "generated.rrr"))
    return r, O, box


def _equilibrium(pp, s, n, sweeps, every, discard):
    e = []
    for _ in range(sweeps // every):
        s.do_ensemble(every)
        e.append(-s.counters_all()["bonds"] / n)
    for k in range(s.R):
        assert not s.energy(k)[3]
    e = np.array(e)
    return e[int(discard * len(e)):].mean(axis=0)


@pytest.mark.parametrize("kind", ["random", "trained"])
def test_equilibrium_with_learned_moves_matches_reference_runs(pp, tmp_path, kind):
    """nvt rho = 0.5, T = 0.2, N = 300: the mean energy per particle with learned
    moves (a random network or one trained on the run itself) against the 8
    reference runs, within 3 combined standard errors."""
    ref = golden("stats_ref.json")["nvt_rho05_T020"]
    r, O, box = _reference_start(pp, tmp_path, ref["n"], ref["rho0"])
    R = 32
    with pp.System.batch([r] * R, [O] * R, [box] * R, [dict(T=0.2)] * R, seed=99) as s:
        if kind == "random":
            net = _skewed_net(pp, 11, drift=0.02, spread=(0.06, 0.03, 0.03), kappa=3.0, hidden_scale=3.0)
            frac = 0.5
        else:
            net = pp.ProposalNet(seed=12)
            s.do_ensemble(2000)
            l0, l1 = s.train(net, steps=300, batch=2048, lr=1e-3, pos_noise=0.05, theta_max=0.1)
            assert l1 < l0
            frac = 0.8
        s.attach(net, learned_fraction=frac, theta_max=0.1)
        e_rep = _equilibrium(pp, s, ref["n"], ref["sweeps"], ref["print_every"], ref["discard"])
        tried, acc = s.learn_counters(0)
        assert tried > 0 and acc > 0
    net.close()
    m, se = float(e_rep.mean()), float(e_rep.std(ddof=1) / math.sqrt(R))
    tol = 3 * math.hypot(se, ref["E_per_N_se"])
    assert abs(m - ref["E_per_N"]) < tol, (kind, m, se, ref["E_per_N"], ref["E_per_N_se"])


def test_training_lowers_the_loss(pp, oracle):
    """Denoising training on a bonded fluid: the loss of the move back falls
    well below that of the untrained network, and the trained network's
    gradients at the end still match the oracle's."""
    r, O, box = oracle.generate(2000, 0.5, seed=4)
    with pp.System(r, O, box, T=0.15, seed=3) as s, pp.ProposalNet(seed=2) as net:
        s.do_ensemble(500)
        l0, _ = s.train(net, steps=1, batch=4096, lr=1e-9)
        a, b = s.train(net, steps=400, batch=4096, lr=1e-3, pos_noise=0.05, theta_max=0.1, sweeps_between=1)
        assert b < l0 - 1.0, (l0, a, b)
        # the trained network on real descriptors: forward against the oracle
        rr, OO, bx = s.state(0)
        idx = np.arange(0, 2000, 7, dtype=np.int32)
        x = pp.net_descriptor(rr, OO, bx, idx)
        w, bb = net.get()
        got = net.forward(x).astype(np.float64)
        want = no.forward(w, bb, x.astype(np.float64))
        assert np.linalg.norm(got - want) < 1e-2 * np.linalg.norm(want)


def _jumps(s, sweeps):
    """Mean squared displacement and mean 1 - cos(rotation angle) per
    attempted move, from states one sweep apart (the expected squared jump of
    a move; a particle rarely moves twice in one sweep of one trial per cell)."""
    tr = rot = 0.0
    att = 0
    for _ in range(sweeps):
        r0, O0, box = s.state(0)
        c0 = s.counters()["tries"][0]
        s.do_ensemble(1)
        r1, O1, _ = s.state(0)
        d = r1 - r0
        d -= box * np.rint(d / box)
        tr += float((d * d).sum())
        rot += float((1 - (np.einsum("nij,nij->n", O1, O0) - 1) / 2).sum())
        att += s.counters()["tries"][0] - c0
    return tr / att, rot / att


def test_trained_network_gains_over_plain_moves(pp, oracle):
    """What training buys, per attempted move at rho = 0.5, T = 0.12 with one
    trial per non-empty cell and pass: the trained proposal's squared jump
    beats the plain uniform move's (Disp_max 0.05) and the untrained
    network's.  Measured: about 1.7x plain and 3x untrained in translation;
    rotations stay below the plain move's (the learned axis density
    concentrates), and the energy's autocorrelation time per attempt is
    unchanged (bench.py reports both)."""
    r, O, box = oracle.generate(4000, 0.5, seed=7)
    with pp.System(r, O, box, T=0.12, seed=5, trials_per_cell=8) as s:
        s.do_ensemble(1500)
        st = s.state(0)
    with pp.System(*st, T=0.12, seed=9, trials_per_cell=1) as s:
        plain = _jumps(s, 20)
    res = {}
    for name, steps in (("untrained", 0), ("trained", 1500)):
        with pp.ProposalNet(seed=3) as net:
            if steps:
                with pp.System(*st, T=0.12, seed=13, trials_per_cell=8) as s:
                    s.train(net, steps=steps, batch=4096, lr=1e-3, pos_noise=0.03, theta_max=0.1)
            with pp.System(*st, T=0.12, seed=9, trials_per_cell=1) as s:
                s.attach(net, 1.0, 0.1)
                res[name] = _jumps(s, 20)
                tried, acc = s.learn_counters(0)
                assert tried > 0
    assert res["trained"][0] > 1.3 * plain[0], (res, plain)
    assert res["trained"][0] > 2.0 * res["untrained"][0], (res, plain)
    assert res["trained"][1] > 1.5 * res["untrained"][1], (res, plain)
