"""CPU restatement (numpy, float64) of the learned proposal of the
rototranslation move: neighbourhood descriptor, the MLP, the proposal density
q and the gradient of -log q through the network.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the GPU
kernels.  The product package never imports this module.

The learned move has no counterpart in the reference (it replaces the uniform
draws of MC_move_rototranslate, MC.c:138-161, inside the same Metropolis
test); what this module restates is this project's own definition:

- descriptor: the 12 nearest other particles within 1.5 (minimum image,
  utils.c's periodic distance), nearest first, each as its separation in the
  body frame of the moving particle (O d, O's rows the body axes) and the
  quaternion (x, y, z, w >= 0) of Oq O^T; zero-padded; values rounded to BF16;
- network: 96 -> 128 -> 128 -> 128 -> 32, ReLU, weights and activations rounded
  to BF16 between layers, biases and accumulation wide;
- proposal: Gaussian displacement (mu, log sigma clamped to [-6.2, -0.5]) in
  the body frame, and a 4-component von Mises-Fisher mixture over the rotation
  axis (direction normalised, log kappa clamped to [-3, 8], softmax weights).

Pins (tests/test_oracle_net.py): the descriptor by rigid-motion and
permutation invariance and by the rotation its quaternion encodes; log q by
numerical normalisation over R^3 x S^2 and by the vMF mean-cosine closed form;
the gradients by central finite differences; the forward pass by torch's own
float64 linear layers.
"""
from __future__ import annotations

import math

import numpy as np

NB, FEAT = 12, 7
IN, IN_PAD, H, OUT, OUT_PAD, MIX = NB * FEAT, 96, 128, 26, 32, 4
RCUT = 1.5
LS_MIN, LS_MAX = -6.2, -0.5
LK_MIN, LK_MAX = -3.0, 8.0
LOG_2PI = math.log(2 * math.pi)
K = (IN_PAD, H, H, H)     # inputs of layer l
N = (H, H, H, OUT_PAD)    # outputs of layer l
W_OFF = np.concatenate([[0], np.cumsum([k * n for k, n in zip(K, N)])]).astype(int)
B_OFF = np.concatenate([[0], np.cumsum(N)]).astype(int)
N_W, N_B = int(W_OFF[-1]), int(B_OFF[-1])


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(w, b, round_weights=True):
    """Weight matrices (K_l x N_l) and biases of the flat layout."""
    Ws, bs = [], []
    for l in range(4):
        W = np.asarray(w[W_OFF[l]:W_OFF[l + 1]], np.float64).reshape(K[l], N[l])
        Ws.append(bf16(W) if round_weights else W)
        bs.append(np.asarray(b[B_OFF[l]:B_OFF[l + 1]], np.float64))
    return Ws, bs


# ---------------------------------------------------------------------------
# descriptor

def quat_of(M):
    """Quaternion (x, y, z, w), w >= 0, of the rotation matrix M (rows): the
    largest-diagonal branch choice of Shepperd's method."""
    tr = M[0, 0] + M[1, 1] + M[2, 2]
    if tr > 0:
        S = math.sqrt(tr + 1.) * 2.
        w, x, y, z = 0.25 * S, (M[2, 1] - M[1, 2]) / S, (M[0, 2] - M[2, 0]) / S, (M[1, 0] - M[0, 1]) / S
    elif M[0, 0] > M[1, 1] and M[0, 0] > M[2, 2]:
        S = math.sqrt(1. + M[0, 0] - M[1, 1] - M[2, 2]) * 2.
        w, x, y, z = (M[2, 1] - M[1, 2]) / S, 0.25 * S, (M[0, 1] + M[1, 0]) / S, (M[0, 2] + M[2, 0]) / S
    elif M[1, 1] > M[2, 2]:
        S = math.sqrt(1. + M[1, 1] - M[0, 0] - M[2, 2]) * 2.
        w, x, y, z = (M[0, 2] - M[2, 0]) / S, (M[0, 1] + M[1, 0]) / S, 0.25 * S, (M[1, 2] + M[2, 1]) / S
    else:
        S = math.sqrt(1. + M[2, 2] - M[0, 0] - M[1, 1]) * 2.
        w, x, y, z = (M[1, 0] - M[0, 1]) / S, (M[0, 2] + M[2, 0]) / S, (M[1, 2] + M[2, 1]) / S, 0.25 * S
    sg = -1. if w < 0 else 1.
    return np.array([sg * x, sg * y, sg * z, sg * w])


def quat_matrix(q):
    """Rotation matrix of the unit quaternion (x, y, z, w)."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def min_image(d, box):
    return d - box * np.rint(d / box)


def descriptor(r, O, box, i, round_bf16=True):
    """Descriptor (84 values) of particle i, every other particle a candidate."""
    r = np.asarray(r, np.float64).reshape(-1, 3)
    O = np.asarray(O, np.float64).reshape(-1, 3, 3)
    box = np.asarray(box, np.float64)
    d = min_image(r - r[i], box)
    r2 = (d * d).sum(1)
    others = [j for j in range(len(r)) if j != i and r2[j] < RCUT * RCUT]
    # nearest first; the distance compared in single precision, ties by index
    others.sort(key=lambda j: (np.float32(r2[j]), j))
    out = np.zeros(IN)
    for t, j in enumerate(others[:NB]):
        out[t * FEAT:t * FEAT + 3] = O[i] @ d[j]
        out[t * FEAT + 3:t * FEAT + 7] = quat_of(O[j] @ O[i].T)
    if round_bf16:
        out = bf16(out.astype(np.float32))
    return out


# ---------------------------------------------------------------------------
# network

def forward(w, b, x, round_act=True, keep=False, exact=False):
    """y (m, 26) of descriptors x (m, 84); with keep, also the layer inputs
    and the ReLU masks for the backward pass.  exact: no BF16 rounding at all
    (a smooth function of the weights, for finite differences)."""
    Ws, bs = layers(w, b, round_weights=not exact)
    round_act = round_act and not exact
    m = x.shape[0]
    a = np.zeros((m, IN_PAD))
    a[:, :IN] = x if exact else bf16(x)
    acts, masks = [a], []
    for l in range(4):
        z = a @ Ws[l] + bs[l]
        if l < 3:
            masks.append(z > 0)
            a = np.maximum(z, 0.)
            if round_act:
                a = bf16(a)
            acts.append(a)
        else:
            y = z[:, :OUT]
    return (y, acts, masks, Ws) if keep else y


def decode(y):
    y = np.asarray(y, np.float64)
    p = {"mu": y[:, 0:3], "ls_raw": y[:, 3:6]}
    p["ls"] = np.clip(p["ls_raw"], LS_MIN, LS_MAX)
    c = y[:, 6:26].reshape(-1, MIX, 5)
    u = c[..., :3]
    p["nrm"] = np.sqrt((u * u).sum(-1) + 1e-12)
    p["dir"] = u / p["nrm"][..., None]
    p["lk_raw"] = c[..., 3]
    p["lk"] = np.clip(p["lk_raw"], LK_MIN, LK_MAX)
    lg = c[..., 4]
    mx = lg.max(1, keepdims=True)
    p["lw"] = lg - (mx + np.log(np.exp(lg - mx).sum(1, keepdims=True)))
    return p


def log_vmf(lk, c):
    """log of kappa / (4 pi sinh kappa) exp(kappa c) on S^2."""
    k = np.exp(lk)
    return lk - LOG_2PI - k - np.log(-np.expm1(-2 * k)) + k * c


def logq(y, d, a):
    """log density of body-frame displacement d (m, 3) and unit axis a (m, 3)."""
    p = decode(y)
    z = (d - p["mu"]) * np.exp(-p["ls"])
    lg = (-0.5 * z * z - p["ls"] - 0.5 * LOG_2PI).sum(1)
    cs = (p["dir"] * a[:, None, :]).sum(-1)
    l = p["lw"] + log_vmf(p["lk"], cs)
    mx = l.max(1, keepdims=True)
    return lg + (mx + np.log(np.exp(l - mx).sum(1, keepdims=True)))[:, 0]


def loss_grad_y(y, d, a):
    """-log q per row and its gradient with respect to y (m, 26); zero
    gradient for a clamped log sigma / log kappa."""
    p = decode(y)
    m = y.shape[0]
    g = np.zeros((m, OUT))
    s2inv = np.exp(-2 * p["ls"])
    diff = d - p["mu"]
    z2 = diff * diff * s2inv
    loss = (0.5 * z2 + p["ls"] + 0.5 * LOG_2PI).sum(1)
    g[:, 0:3] = -diff * s2inv
    inside = (p["ls_raw"] > LS_MIN) & (p["ls_raw"] < LS_MAX)
    g[:, 3:6] = np.where(inside, 1 - z2, 0.)
    cs = (p["dir"] * a[:, None, :]).sum(-1)
    l = p["lw"] + log_vmf(p["lk"], cs)
    mx = l.max(1, keepdims=True)
    lse = mx + np.log(np.exp(l - mx).sum(1, keepdims=True))
    loss -= lse[:, 0]
    resp = np.exp(l - lse)
    wgt = np.exp(p["lw"])
    k = np.exp(p["lk"])
    kcoth = k * (1 + np.exp(-2 * k)) / (-np.expm1(-2 * k))
    gc = g[:, 6:26].reshape(m, MIX, 5)
    gc[..., :3] = -(resp * k)[..., None] * (a[:, None, :] - cs[..., None] * p["dir"]) / p["nrm"][..., None]
    lk_in = (p["lk_raw"] > LK_MIN) & (p["lk_raw"] < LK_MAX)
    gc[..., 3] = np.where(lk_in, -resp * (1 + k * cs - kcoth), 0.)
    gc[..., 4] = wgt - resp
    g[:, 6:26] = gc.reshape(m, 20)
    return loss, g


def loss_and_grads(w, b, x, target, round_act=True, exact=False):
    """Mean -log q over the rows and its gradient with respect to the flat
    weights and biases (float64 backward through the rounded forward)."""
    y, acts, masks, Ws = forward(w, b, x, round_act=round_act, keep=True, exact=exact)
    m = x.shape[0]
    loss, gy = loss_grad_y(y, target[:, :3], target[:, 3:6])
    dz = np.zeros((m, OUT_PAD))
    dz[:, :OUT] = gy / m
    gw = np.zeros(N_W)
    gb = np.zeros(N_B)
    for l in range(3, -1, -1):
        gw[W_OFF[l]:W_OFF[l + 1]] = (acts[l].T @ dz).reshape(-1)
        gb[B_OFF[l]:B_OFF[l + 1]] = dz.sum(0)
        if l > 0:
            dz = (dz @ Ws[l].T) * masks[l - 1]
    return float(loss.mean()), gw, gb
