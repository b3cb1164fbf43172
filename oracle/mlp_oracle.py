"""CPU oracle of the learned step-size proposal (test infrastructure only).

The reference has no learned step size; its controller is the PI rule of
sources/integrations.F90:177,186 (RK5) and :466-490 (DOP853).  This module
restates, in float64 NumPy, what the device library adds on top of it:

  * ``features`` -- the 16 inputs the network sees after an accepted step
    (the device's step_features), from the reference's own corner indexing
    and weights (sources/interpolations.F90:140-149, :207-210);
  * ``forward`` / ``gradient`` -- the 16 -> 64 -> 64 -> 1 GELU network and
    the gradient of its mean-squared loss, optionally emulating the device's
    BF16 rounding points (inputs, weights, activations, the backward
    products' operands) so the device's matrix-core results can be compared
    to float64 arithmetic.

Pinning: the float64 backward is checked against central finite differences
(tests/test_mlp.py), GELU against its closed form at known points, and the
features against closed forms in uniform and linearly varying fields.

Only tests/ and ``__graft_entry__.smoke`` import this module.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 16, 64
W1, B1 = 0, NIN * NH
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
MU = B3 + 1
SIG = MU + NIN
NPARAMS = SIG + NIN
NTRAIN = MU
GELU_C = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even) via float32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * z * z)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(w1=p[W1:B1].reshape(NIN, NH), b1=p[B1:W2], w2=p[W2:B2].reshape(NH, NH),
                b2=p[B2:W3], w3=p[W3:B3], b3=p[B3], mu=p[MU:SIG], sig=p[SIG:NPARAMS])


def targets(records, alpha=0.0, lo=-4.0, hi=4.0):
    r = np.asarray(records, dtype=np.float64)
    return np.clip(r[:, 16] - alpha * r[:, 17] / r[:, 19], lo, hi)


def _net(p, feat, emulate):
    q = unpack(p)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    x = rnd((np.asarray(feat, dtype=np.float64) - q["mu"]) / q["sig"])
    w1, w2 = rnd(q["w1"]), rnd(q["w2"])
    z1 = x @ w1 + q["b1"]
    a1 = rnd(gelu(z1))
    z2 = a1 @ w2 + q["b2"]
    a2 = rnd(gelu(z2))
    y = a2 @ q["w3"] + q["b3"]
    return q, rnd, x, w1, w2, z1, a1, z2, a2, y


def forward(p, feat, emulate=True):
    return _net(p, feat, emulate)[-1]


def gradient(p, records, alpha=0.0, emulate=True):
    """(loss, grad) of mean((y - target)^2); grad in the packed order of the
    first NTRAIN parameters.  With emulate, the backward products take BF16
    operands where the device's matrix cores do."""
    r = np.asarray(records, dtype=np.float64)
    q, rnd, x, w1, w2, z1, a1, z2, a2, y = _net(p, r[:, :NIN], emulate)
    t = targets(r, alpha)
    n = r.shape[0]
    res = y - t
    loss = float(np.mean(res * res))
    dy = 2.0 * res / n
    g = np.zeros(NTRAIN)
    g[W3:B3] = a2.T @ dy
    g[B3] = dy.sum()
    dz2 = dy[:, None] * q["w3"][None, :] * gelu_grad(z2)
    g[B2:W3] = dz2.sum(0)
    dz2r = rnd(dz2)
    g[W2:B2] = (a1.T @ dz2r).ravel()
    dz1 = (dz2r @ w2.T) * gelu_grad(z1)
    g[B1:W2] = dz1.sum(0)
    g[W1:B1] = (x.T @ rnd(dz1)).ravel()
    return loss, g


def init_params(seed=0, scale=1.0):
    """A random parameter vector for tests (He-uniform hidden layers)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    p[W1:B1] = rng.uniform(-1, 1, NIN * NH) * np.sqrt(6.0 / NIN) * scale
    p[B1:W2] = rng.normal(0, 0.1, NH) * scale
    p[W2:B2] = rng.uniform(-1, 1, NH * NH) * np.sqrt(6.0 / NH) * scale
    p[B2:W3] = rng.normal(0, 0.1, NH) * scale
    p[W3:B3] = rng.uniform(-1, 1, NH) * np.sqrt(6.0 / NH) * scale
    p[B3] = 0.1
    p[MU:SIG] = rng.normal(0, 0.5, NIN)
    p[SIG:NPARAMS] = rng.uniform(0.5, 2.0, NIN)
    return p


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
def _cell(pos, dm, xl, xs):
    """Left/right node indices and in-cell fractions, the reference's two
    floors (interpolations.F90:140-149), 0-based, wrapped into the grid."""
    l, r, xd = np.zeros(3, int), np.zeros(3, int), np.zeros(3)
    for i in range(3):
        x = (pos[i] - xl[i]) / xs[i]
        xq = x + 1.0 / dm[i]
        r[i] = int(np.floor(dm[i] * (xq - np.floor(xq)))) % dm[i]
        xp = dm[i] * (x - np.floor(x))
        pl = np.floor(xp)
        xd[i] = xp - pl
        l[i] = int(pl) % dm[i]
    return l, r, xd


def features(f1, bb, vec, h, err, pi_ratio, rejected, params, mode="vb"):
    """The 16 network inputs at state vec after an accepted step of size h
    with error ratio err; f1 = U (mode "vb") or E ("eb"), fields (dm1, dm2,
    dm3, 3); params is the reference's params vector."""
    dm = f1.shape[:3]
    qom, vun = params[0], params[1]
    xl, xs = params[11:14], params[14:17]
    l, r, xd = _cell(vec[:3], dm, xl, xs)
    xu = 1.0 - xd
    U, B, bc = np.zeros(3), np.zeros(3), np.zeros(8)
    for j in range(8):
        bits = (j >> 2, (j >> 1) & 1, j & 1)
        ix = [r[a] if bits[a] else l[a] for a in range(3)]
        w = np.prod([xd[a] if bits[a] else xu[a] for a in range(3)])
        u = f1[ix[0], ix[1], ix[2], :]
        b = bb[ix[0], ix[1], ix[2], :]
        U += w * u
        B += w * b
        bc[j] = np.linalg.norm(b)
    E = U if mode == "eb" else -np.cross(U, B)
    bm, em = np.linalg.norm(B), np.linalg.norm(E)
    u = np.asarray(vec[3:6], dtype=np.float64)
    um = np.linalg.norm(u)
    g = np.sqrt(um * um + 1.0)
    cp = u @ B / (um * bm)
    qb = abs(qom) * bm
    tf = np.inf
    for i in range(3):
        vel = u[i] / (g * vun) * (dm[i] / xs[i])
        dist = 1.0 - xd[i] if vel > 0 else xd[i]
        with np.errstate(divide="ignore", invalid="ignore"):
            ti = dist / abs(vel)
        tf = ti if ti < tf else tf
    faces = [bc[[0, 1, 2, 3]], bc[[4, 5, 6, 7]], bc[[0, 1, 4, 5]], bc[[2, 3, 6, 7]],
             bc[[0, 2, 4, 6]], bc[[1, 3, 5, 7]]]
    with np.errstate(divide="ignore", invalid="ignore"):
        f = [np.log(bm), np.log(em / bm)] + [np.log(fc.mean() / bm) for fc in faces] + [
            np.log(g), cp, np.log(h * qb / g), np.log(err), np.log(tf / h), pi_ratio,
            1.0 if rejected else 0.0,
            np.log(um * np.sqrt(max(1.0 - cp * cp, 0.0)) / qb * dm[0] / xs[0])]
    f = np.clip(np.asarray(f, dtype=np.float64), -30.0, 30.0)
    return np.where(np.isnan(f), 0.0, f)
