"""The CPU oracle of the learned proposal (oracle/net_oracle.py) pinned by what
the mathematics fixes: rigid-motion and permutation invariance of the
descriptor, the rotation its quaternions encode, the normalisation of the
proposal density, the vMF mean cosine, and finite differences of every
gradient."""
import math

import numpy as np
import pytest

import net_oracle as no


def _rotations(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    q[np.linalg.det(q) < 0, :, 0] *= -1
    return q


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40, 0.0], np.float32)
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    want = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5]
    assert np.array_equal(no.bf16(x)[:5], want)
    assert no.bf16(x)[7] == 0.0
    # any float: a BF16 value (low 16 bits clear) no farther than either neighbour
    v = np.random.default_rng(2).normal(scale=100.0, size=20000).astype(np.float32)
    r = no.bf16(v).astype(np.float32)
    assert not (r.view(np.uint32) & 0xFFFF).any()
    up = ((v.view(np.uint32) >> 16) << 16).view(np.float32)
    hi = (((v.view(np.uint32) >> 16) + 1) << 16).view(np.float32)
    err = np.abs(r.astype(np.float64) - v)
    assert (err <= np.abs(up.astype(np.float64) - v)).all() and (err <= np.abs(hi.astype(np.float64) - v)).all()


def test_quaternion_encodes_the_rotation():
    rng = np.random.default_rng(3)
    Ms = list(_rotations(rng, 200))
    # each branch of the method: rotations by nearly pi about each axis
    for ax in np.eye(3):
        t = math.pi - 1e-3
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        Ms.append(np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K)
    for M in Ms:
        q = no.quat_of(M)
        assert abs(np.linalg.norm(q) - 1) < 1e-12 and q[3] >= 0
        assert np.allclose(no.quat_matrix(q), M, atol=1e-12)


def _cluster(rng, n, box):
    """n non-overlapping particles around the box centre."""
    r = []
    while len(r) < n:
        p = box / 2 + rng.uniform(-1.6, 1.6, 3)
        if all(np.linalg.norm(p - q) > 1.0 for q in r):
            r.append(p)
    return np.array(r), _rotations(rng, n)


def test_descriptor_invariances():
    rng = np.random.default_rng(5)
    box = np.full(3, 30.0)
    r, O = _cluster(rng, 24, box)
    i = 0
    d0 = no.descriptor(r, O, box, i, round_bf16=False)
    assert np.count_nonzero(d0) > 7  # a neighbour or more within 1.5
    # a global rotation about the box centre: r -> c + Q (r - c), O -> O Q^T
    Q = _rotations(rng, 1)[0]
    r2 = box / 2 + (r - box / 2) @ Q.T
    d1 = no.descriptor(r2, O @ Q.T, box, i, round_bf16=False)
    assert np.allclose(d0, d1, atol=1e-12)
    # a translation across the periodic boundary
    r3 = np.mod(r + np.array([14.9, -7.3, 29.0]), box)
    assert np.allclose(no.descriptor(r3, O, box, i, round_bf16=False), d0, atol=1e-12)
    # relabelling the other particles
    perm = np.concatenate([[0], 1 + rng.permutation(len(r) - 1)])
    assert np.allclose(no.descriptor(r[perm], O[perm], box, 0, round_bf16=False), d0, atol=1e-12)
    # each slot: body-frame separation and relative rotation of a real neighbour
    d = no.min_image(r - r[i], box)
    for t in range(no.NB):
        f = d0[t * 7:(t + 1) * 7]
        if not f.any():
            break
        j = int(np.argmin(np.linalg.norm(d @ O[i].T - f[:3], axis=1)))
        assert np.allclose(O[i].T @ f[:3], d[j], atol=1e-12)
        assert np.allclose(no.quat_matrix(f[3:]) @ O[i], O[j], atol=1e-12)
        assert np.linalg.norm(d[j]) < no.RCUT


def test_descriptor_keeps_the_nearest_twelve_in_order():
    rng = np.random.default_rng(6)
    box = np.full(3, 20.0)
    # 20 points on shells of growing radius around particle 0
    rad = np.linspace(1.0, 1.45, 20)
    v = rng.normal(size=(20, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = np.vstack([[10, 10, 10], 10 + v * rad[:, None]])
    O = np.broadcast_to(np.eye(3), (21, 3, 3)).copy()
    d = no.descriptor(r, O, box, 0, round_bf16=False)
    got = np.linalg.norm(d.reshape(12, 7)[:, :3], axis=1)
    assert np.allclose(got, rad[:12])


def _random_y(rng, m):
    y = rng.normal(size=(m, no.OUT))
    y[:, 3:6] = rng.uniform(-3.0, -1.0, (m, 3))
    for c in range(no.MIX):
        y[:, 6 + 5 * c + 3] = rng.uniform(-1.0, 3.5, m)
    return y


def test_proposal_density_is_normalised():
    """q(d, a) integrates to 1 over R^3 x S^2: the Gaussian part on a grid,
    the vMF mixture with a Fibonacci lattice of the sphere."""
    rng = np.random.default_rng(7)
    y = _random_y(rng, 3)
    n = 200000
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = math.pi * (1 + 5 ** 0.5) * k
    A = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    for row in y:
        Y = np.broadcast_to(row, (n, no.OUT))
        # fix d at mu: log q(mu, a) - log N(mu) = log mixture(a)
        p = no.decode(row[None])
        d = np.broadcast_to(p["mu"][0], (n, 3))
        lgm = no.logq(Y, d, A) - (-p["ls"][0] - 0.5 * no.LOG_2PI).sum()
        assert abs(np.exp(lgm).mean() * 4 * math.pi - 1) < 2e-3
        # the Gaussian along one axis, the others at their means
        s = math.exp(p["ls"][0][0])
        xs = np.linspace(-8 * s, 8 * s, 4001)
        D = np.tile(p["mu"][0], (len(xs), 1))
        D[:, 0] += xs
        a = np.tile([0., 0., 1.], (len(xs), 1))
        lq = no.logq(np.broadcast_to(row, (len(xs), no.OUT)), D, a)
        lq -= no.logq(row[None], p["mu"], np.array([[0., 0., 1.]]))[0]
        lq += -p["ls"][0][0] - 0.5 * no.LOG_2PI
        assert abs(np.exp(lq).sum() * (xs[1] - xs[0]) - 1) < 1e-6


@pytest.mark.parametrize("kappa", [0.05, 1.0, 7.5, 150.0])
def test_vmf_mean_cosine(kappa):
    c = np.linspace(-1, 1, 400001)
    f = 2 * math.pi * np.exp(no.log_vmf(np.full_like(c, math.log(kappa)), c))
    h = c[1] - c[0]
    assert abs(f.sum() * h - 1) < 1e-4 * max(1, kappa / 10)
    mean = (c * f).sum() * h
    want = 1 / math.tanh(kappa) - 1 / kappa
    assert abs(mean - want) < 1e-4 * max(1, kappa / 10)


def test_loss_gradient_wrt_outputs_by_finite_differences():
    rng = np.random.default_rng(8)
    m = 6
    y = _random_y(rng, m)
    d = rng.normal(scale=0.1, size=(m, 3))
    a = rng.normal(size=(m, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    loss, g = no.loss_grad_y(y, d, a)
    assert np.allclose(loss, -no.logq(y, d, a), atol=1e-12)
    h = 1e-6
    for e in range(no.OUT):
        yp, ym = y.copy(), y.copy()
        yp[:, e] += h
        ym[:, e] -= h
        fd = (-no.logq(yp, d, a) + no.logq(ym, d, a)) / (2 * h)
        assert np.allclose(g[:, e], fd, rtol=1e-5, atol=1e-6), e
    # clamped log sigma / log kappa: no gradient
    y2 = y.copy()
    y2[:, 3] = -9.0
    y2[:, 6 + 3] = 9.5
    _, g2 = no.loss_grad_y(y2, d, a)
    assert not g2[:, 3].any() and not g2[:, 9].any()


def _small_net(rng, scale=0.15):
    w = np.zeros(no.N_W)
    b = np.zeros(no.N_B)
    for l in range(4):
        kin = no.IN if l == 0 else no.H
        nout = no.OUT if l == 3 else no.H
        W = np.zeros((no.K[l], no.N[l]))
        W[:kin, :nout] = rng.normal(scale=scale, size=(kin, nout))
        w[no.W_OFF[l]:no.W_OFF[l + 1]] = W.reshape(-1)
        b[no.B_OFF[l]:no.B_OFF[l] + nout] = rng.normal(scale=0.05, size=nout)
    return w, b


def test_network_gradients_by_finite_differences():
    rng = np.random.default_rng(9)
    w, b = _small_net(rng)
    m = 8
    x = rng.normal(scale=0.5, size=(m, no.IN))
    t = np.hstack([rng.normal(scale=0.1, size=(m, 3)), rng.normal(size=(m, 3))])
    t[:, 3:] /= np.linalg.norm(t[:, 3:], axis=1, keepdims=True)
    loss, gw, gb = no.loss_and_grads(w, b, x, t, exact=True)
    h = 1e-6

    def f(w2, b2):
        return no.loss_and_grads(w2, b2, x, t, exact=True)[0]

    for l in range(4):
        kin = no.IN if l == 0 else no.H
        nout = no.OUT if l == 3 else no.H
        for _ in range(12):
            k, n = rng.integers(kin), rng.integers(nout)
            e = no.W_OFF[l] + k * no.N[l] + n
            wp, wm = w.copy(), w.copy()
            wp[e] += h
            wm[e] -= h
            fd = (f(wp, b) - f(wm, b)) / (2 * h)
            assert abs(gw[e] - fd) < 1e-5 * (1 + abs(fd)), (l, k, n, gw[e], fd)
        for _ in range(6):
            e = no.B_OFF[l] + rng.integers(nout)
            bp, bm = b.copy(), b.copy()
            bp[e] += h
            bm[e] -= h
            fd = (f(w, bp) - f(w, bm)) / (2 * h)
            assert abs(gb[e] - fd) < 1e-5 * (1 + abs(fd))
    # padded rows and columns carry no gradient
    assert not gw[no.W_OFF[0] + no.IN * no.N[0]:no.W_OFF[1]].any()
    W3 = gw[no.W_OFF[3]:].reshape(no.H, no.OUT_PAD)
    assert not W3[:, no.OUT:].any()


def test_forward_matches_torch_linear_layers():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    w, b = _small_net(rng)
    x = rng.normal(size=(5, no.IN))
    y = no.forward(w, b, x, exact=True)
    a = torch.zeros(5, no.IN_PAD, dtype=torch.float64)
    a[:, :no.IN] = torch.from_numpy(x)
    for l in range(4):
        W = torch.from_numpy(w[no.W_OFF[l]:no.W_OFF[l + 1]].reshape(no.K[l], no.N[l]))
        a = torch.nn.functional.linear(a, W.T.contiguous(), torch.from_numpy(b[no.B_OFF[l]:no.B_OFF[l + 1]]))
        if l < 3:
            a = torch.relu(a)
    assert np.allclose(y, a[:, :no.OUT].numpy(), atol=1e-12)
