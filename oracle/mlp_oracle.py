"""CPU oracle of the learned warm start (test infrastructure only).

Restates, in numpy, what libvbd_ws.so computes (newton_amd/csrc/vbd_warmstart.hip):
the 48 per-particle features, the 48-128-128-128-3 SiLU MLP forward pass, its
backward pass for the weighted mean-squared error, Adam, and the 1-ring low-pass
of predicted corrections (``smooth``) -- with the same
BF16 rounding points (features; H_l = bf16(silu(Z_l)); dZ_l = bf16(dH_l *
silu'(Z_l)); dZ_4 = bf16(dY)), the products accumulated in float64 here and in
FP32 on the MFMA units there.  ``bf16=False`` drops every rounding, giving the
exact float64 gradient that the tests pin with finite differences.

The warm start has no counterpart in the reference (Newton's SolverVBD starts
from the inertial guess, particle_vbd_kernels.py:1350-1375), so this oracle
restates this repository's own algorithm; the pins are the mathematics
(finite differences, Adam's closed form on one step) and torch's BF16
conversion for the rounding.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this.
"""

from __future__ import annotations

import numpy as np

N_IN, N_HIDDEN, N_OUT = 48, 128, 3
SHAPES = [("W1", (N_IN, N_HIDDEN)), ("b1", (N_HIDDEN,)), ("W2", (N_HIDDEN, N_HIDDEN)), ("b2", (N_HIDDEN,)),
          ("W3", (N_HIDDEN, N_HIDDEN)), ("b3", (N_HIDDEN,)), ("W4", (N_HIDDEN, N_OUT)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_bits(x) -> np.ndarray:
    """Round-to-nearest-even float32 -> BF16 bit patterns (uint16); NaN stays NaN."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return r.astype(np.uint16)


def bf16_to_f32(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x) -> np.ndarray:
    return bf16_to_f32(bf16_bits(x))


def unpack(params) -> dict:
    p = np.asarray(params)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(named: dict) -> np.ndarray:
    return np.concatenate([np.asarray(named[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(x, params, bf16: bool = True):
    """x: [n,48] (BF16 values as float) -> (out [n,3], H list, Z list), float64."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    # FP32 biases as the device holds them (exact float64 when rounding is off)
    rb = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if bf16 else (lambda a: np.asarray(a))
    P = unpack(params)
    h = np.asarray(x, dtype=np.float64)
    hs, zs = [], []
    for l in (1, 2, 3):
        z = h @ r(P[f"W{l}"]).astype(np.float64) + rb(P[f"b{l}"])
        h = r(_silu(z)).astype(np.float64)
        zs.append(z)
        hs.append(h)
    out = h @ r(P["W4"]).astype(np.float64) + rb(P["b4"])
    return out, hs, zs


def loss_and_grad(x, params, target, row_weight=None, bf16: bool = True):
    """Weighted MSE (mean over rows and outputs) and its gradient w.r.t. the parameter vector (float64)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    w = np.ones(n) if row_weight is None else np.asarray(row_weight, dtype=np.float64)
    P = unpack(params)
    out, hs, zs = forward(x, params, bf16)
    diff = out - np.asarray(target, dtype=np.float64)
    loss = float(np.sum(w[:, None] * diff * diff) / (3.0 * n))
    dz = r(2.0 * w[:, None] * diff / (3.0 * n)).astype(np.float64)
    grads = {"W4": hs[2].T @ dz, "b4": dz.sum(axis=0)}
    for l in (3, 2, 1):
        dh = dz @ r(P[f"W{l + 1}"]).astype(np.float64).T
        dz = r(dh * _dsilu(zs[l - 1])).astype(np.float64)
        prev = hs[l - 2] if l > 1 else x
        grads[f"W{l}"] = prev.T @ dz
        grads[f"b{l}"] = dz.sum(axis=0)
    return loss, pack(grads)


def adam(params, grad, m, v, lr, beta1, beta2, eps, step):
    """One Adam step in float32 (returns new params, m, v)."""
    f = np.float32
    g = np.asarray(grad, dtype=f)
    m = f(beta1) * np.asarray(m, dtype=f) + (f(1.0) - f(beta1)) * g
    v = f(beta2) * np.asarray(v, dtype=f) + (f(1.0) - f(beta2)) * g * g
    c1 = f(1.0) - f(beta1) ** f(step)
    c2 = f(1.0) - f(beta2) ** f(step)
    p = np.asarray(params, dtype=f) - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p, m, v


def features(nbr_off, nbr, x0, y, corr, acc, mass, area, dt, scales) -> np.ndarray:
    """The 48 feature columns per particle (float32 arithmetic, neighbour order), before BF16 rounding."""
    f32 = np.float32
    x0 = np.asarray(x0, f32)
    y = np.asarray(y, f32)
    corr = np.zeros_like(x0) if corr is None else np.asarray(corr, f32)
    acc_ = None if acc is None else np.asarray(acc, f32)
    sd, sc, sl, mm, ma = (f32(s) for s in scales)
    dt = f32(dt)
    n = x0.shape[0]
    out = np.zeros((n, N_IN), dtype=f32)
    for i in range(n):
        a = (y[i] - x0[i]) / sd
        c = corr[i] / sc
        j0, j1 = int(nbr_off[i]), int(nbr_off[i + 1])
        deg = j1 - j0
        inv = f32(1.0) / f32(deg) if deg > 0 else f32(0.0)
        sa = np.zeros(3, f32)
        scs = np.zeros(3, f32)
        sdx = np.zeros(3, f32)
        sdy = np.zeros(3, f32)
        spin = np.zeros(3, f32)
        mxa, mna, mxc, mnc = a.copy(), a.copy(), c.copy(), c.copy()
        slen = f32(0.0)
        sabsc = f32(0.0)
        mxl = mnl = mxabsc = f32(0.0)
        npin = 0
        for jj in range(j0, j1):
            j = int(nbr[jj])
            aj = (y[j] - x0[j]) / sd
            cj = corr[j] / sc
            dx = x0[j] - x0[i]
            sdy = sdy + (y[j] - y[i])
            ln = f32(np.sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2])) / sl
            ac = f32(np.sqrt(cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2]))
            if jj == j0:
                mxa, mna, mxc, mnc = aj.copy(), aj.copy(), cj.copy(), cj.copy()
                mxl = mnl = ln
                mxabsc = ac
            else:
                mxa, mna = np.maximum(mxa, aj), np.minimum(mna, aj)
                mxc, mnc = np.maximum(mxc, cj), np.minimum(mnc, cj)
                mxl, mnl = max(mxl, ln), min(mnl, ln)
                mxabsc = max(mxabsc, ac)
            sa = sa + aj
            scs = scs + cj
            sdx = sdx + dx
            slen = f32(slen + ln)
            sabsc = f32(sabsc + ac)
            if mass[j] == 0.0:
                npin += 1
                spin = spin + dx
        absa = f32(np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]))
        absc = f32(np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]))
        row = out[i]
        row[0:3] = a
        row[3:6] = c
        row[6:9] = sa * inv - a if deg > 0 else 0.0
        row[9:12] = scs * inv - c if deg > 0 else 0.0
        row[12:15] = mxa - a
        row[15:18] = mna - a
        row[18:21] = mxc - c
        row[21:24] = mnc - c
        row[24:27] = sdx * inv / sl
        row[27] = slen * inv
        row[28] = mxl
        row[29] = mnl
        row[30:33] = acc_[i] * dt * dt / sd if acc_ is not None else 0.0
        row[33] = absa
        row[34] = absc
        row[35] = sabsc * inv if deg > 0 else absc
        row[36] = mxabsc if deg > 0 else absc
        row[37] = f32(mass[i]) / mm
        row[38] = f32(area[i]) / ma
        row[39] = f32(npin) * inv
        row[40] = f32(deg) / f32(6.0)
        row[41:44] = sdy * inv / sl
        row[44:47] = spin * inv / sl
        row[47] = 1.0
    return out


__all__ = ["N_PARAMS", "adam", "bf16_bits", "bf16_round", "bf16_to_f32", "features", "forward", "loss_and_grad",
           "pack", "unpack"]


def smooth(nbr_off, nbr, active, v, sweeps: int) -> np.ndarray:
    """``vbd_ws_smooth``: ``sweeps`` sweeps of v_i <- (v_i + sum_j active_j v_j) / (1 + degree_i) for
    active particles (inactive ones become 0), in float32 with the kernel's order of operations
    (the particle's own value, then its neighbours in CSR order, then one multiply by the float32
    reciprocal of 1 + degree)."""
    off = np.asarray(nbr_off, dtype=np.int64)
    nb = np.asarray(nbr, dtype=np.int64)
    act = np.asarray(active, dtype=np.float32).reshape(-1)
    v = np.array(v, dtype=np.float32, copy=True)
    n = v.shape[0]
    deg = np.diff(off)
    inv = (np.float32(1.0) / (1 + deg).astype(np.float32)).astype(np.float32)
    for _ in range(sweeps):
        acc = v.copy()
        for slot in range(int(deg.max()) if n else 0):
            has = np.nonzero(deg > slot)[0]
            j = nb[off[has] + slot]
            acc[has] = (acc[has] + act[j, None] * v[j]).astype(np.float32)
        out = (acc * inv[:, None]).astype(np.float32)
        out[act == 0] = 0.0
        v = out
    return v
