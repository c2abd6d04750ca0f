"""CPU oracle of the learned precision selector (TEST INFRASTRUCTURE ONLY).

The reference has no learned selector; this module states, in plain numpy,
what the selector's GPU kernels compute, so that tests can check them:
  * tile_features -- the 12 per-tile features of a conserved FP64 state;
  * mlp_forward / mlp_grads -- the 12 -> 64 -> 64 -> 1 network (ReLU,
    sigmoid), BF16 operands with FP32 accumulation as the MFMA kernels do,
    and the gradients of the weighted binary cross-entropy with respect to
    the FP32 master weights (straight-through the BF16 rounding);
  * tile_errors -- the per-tile label source: the largest scaled
    difference between an FP32-flux and an FP64-flux stage update (MHD
    variables; psi excluded).
Only tests/ use it.
"""
from __future__ import annotations

import numpy as np

NFEAT = 12
NIN = 16          # features padded to the MFMA K dimension
H = 64
GAMMA = 5.0 / 3.0


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def _prims(u):
    rho = np.maximum(u[0], 1e-12)
    vx, vy, vz = u[1] / rho, u[2] / rho, u[3] / rho
    bx, by, bz = u[5], u[6], u[7]
    b2 = bx * bx + by * by + bz * bz
    p = (GAMMA - 1.0) * (u[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * b2)
    return rho, vx, vy, vz, p, bx, by, bz, b2


def tile_features(u, tx, ty, ch):
    """Features of every tile (row-major tile order), shape (ntiles, 12).

    u: conserved state (9, ny, nx) float64; tiles of tx x ty cells (partial
    tiles at the high edges); ch: the step's c_h.  Neighbour differences use
    the right/top neighbours, div B the central difference (periodic)."""
    ny, nx = u.shape[1:]
    rho, vx, vy, vz, p, bx, by, bz, b2 = _prims(u)
    pp = np.maximum(p, 1e-12)
    bmag = np.sqrt(b2)
    beta = np.clip(2.0 * pp / np.maximum(b2, 1e-30), 1e-6, 1e6)
    v2 = vx * vx + vy * vy + vz * vz
    cs2 = GAMMA * pp / rho
    mach = np.minimum(np.sqrt(v2 / cs2), 100.0)
    right = lambda a: np.roll(a, -1, axis=1)
    top = lambda a: np.roll(a, -1, axis=0)
    left = lambda a: np.roll(a, 1, axis=1)
    down = lambda a: np.roll(a, 1, axis=0)
    drho = np.maximum(np.abs(right(rho) - rho), np.abs(top(rho) - rho)) / rho
    dp = np.maximum(np.abs(right(pp) - pp), np.abs(top(pp) - pp)) / pp
    dbv = np.maximum(np.sqrt((right(bx) - bx) ** 2 + (right(by) - by) ** 2 + (right(bz) - bz) ** 2),
                     np.sqrt((top(bx) - bx) ** 2 + (top(by) - by) ** 2 + (top(bz) - bz) ** 2))
    db = dbv / np.maximum(bmag, 1e-12)
    h = 1.0 / max(nx, ny)
    divb = 0.5 * (right(bx) - left(bx)) * nx + 0.5 * (top(by) - down(by)) * ny
    divb_n = np.abs(divb) * h / (bmag + 1e-12)
    bn_min = np.minimum(np.abs(bx), np.abs(by)) / (bmag + 1e-12)
    vch = np.sqrt(v2) / ch
    cf = np.sqrt(cs2 + b2 / rho) / ch
    ke = 0.5 * rho * v2
    me = 0.5 * b2
    psi = np.abs(u[8])
    tiles_x, tiles_y = -(-nx // tx), -(-ny // ty)
    out = np.empty((tiles_x * tiles_y, NFEAT))
    for t in range(tiles_x * tiles_y):
        j0, i0 = (t // tiles_x) * ty, (t % tiles_x) * tx
        sl = (slice(j0, min(j0 + ty, ny)), slice(i0, min(i0 + tx, nx)))
        n = (sl[0].stop - sl[0].start) * (sl[1].stop - sl[1].start)
        out[t] = [
            np.log10(beta[sl].min()), np.log10(beta[sl].max()),
            np.log10(drho[sl].max() + 1e-12), np.log10(dp[sl].max() + 1e-12),
            np.log10(db[sl].max() + 1e-12), mach[sl].max(),
            np.log10(psi[sl].max() / (bmag[sl].max() + 1e-12) + 1e-12),
            np.log10(divb_n[sl].max() + 1e-12), np.log10(bn_min[sl].min() + 1e-12),
            vch[sl].max(), cf[sl].max(),
            np.log10(ke[sl].sum() / n / (me[sl].sum() / n + 1e-12) + 1e-12),
        ]
    return out


def tile_errors(u_fp32flux, u_fp64flux, tx, ty):
    """Largest scaled stage-update difference per tile: max over the tile's
    cells and the 8 MHD variables (rho, momentum, E, B; not the GLM
    potential psi, which has no scale of its own) of |a - b| / (|b| + s_k),
    s_k the grid RMS (of b) of rho, |momentum|, E or |B| for variable k."""
    a = np.asarray(u_fp32flux, dtype=np.float64)[:8]
    b = np.asarray(u_fp64flux, dtype=np.float64)[:8]
    ny, nx = b.shape[1:]
    ms = np.mean(b * b, axis=(1, 2))
    group = [ms[0], ms[1:4].sum(), ms[1:4].sum(), ms[1:4].sum(), ms[4],
             ms[5:8].sum(), ms[5:8].sum(), ms[5:8].sum()]
    scale = (np.sqrt(np.array(group)) + 1e-30)[:, None, None]
    e = (np.abs(a - b) / (np.abs(b) + scale)).max(axis=0)
    tiles_x, tiles_y = -(-nx // tx), -(-ny // ty)
    out = np.empty(tiles_x * tiles_y)
    for t in range(tiles_x * tiles_y):
        j0, i0 = (t // tiles_x) * ty, (t % tiles_x) * tx
        out[t] = e[j0:j0 + ty, i0:i0 + tx].max()
    return out


def mlp_forward(x, p, bf16_ops=True):
    """x: (n, 12) standardised features; p: dict W1 (16,64), b1, W2 (64,64),
    b2, w3 (64,), b3.  Returns the intermediate arrays and the probability.
    With bf16_ops the matmul operands are rounded to bfloat16 and the
    products accumulated in float32 (the MFMA contract)."""
    if bf16_ops:
        r, acc = bf16, np.float32
    else:  # plain float64 network (for finite-difference checks)
        r, acc = (lambda a: np.asarray(a, dtype=np.float64)), np.float64
    xp = np.zeros((x.shape[0], NIN), dtype=acc)
    xp[:, :NFEAT] = x
    z1 = (r(xp).astype(np.float64) @ r(p["W1"]).astype(np.float64)).astype(acc) + p["b1"]
    a1 = np.maximum(z1, 0.0)
    z2 = (r(a1).astype(np.float64) @ r(p["W2"]).astype(np.float64)).astype(acc) + p["b2"]
    a2 = np.maximum(z2, 0.0)
    z3 = a2.astype(np.float64) @ p["w3"].astype(np.float64) + p["b3"]
    prob = 1.0 / (1.0 + np.exp(-z3))
    return {"xp": xp, "z1": z1, "a1": a1, "z2": z2, "a2": a2, "z3": z3, "p": prob}


def mlp_grads(x, y, w, p, bf16_ops=True):
    """Gradients of L = sum_s w_s * BCE(p_s, y_s) / sum_s w_s with respect to
    W1, b1, W2, b2, w3, b3 (float64), straight through the BF16 rounding."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    f = mlp_forward(x, p, bf16_ops)
    wsum = float(np.sum(w))
    dz3 = (f["p"] - y) * w / wsum
    g = {"w3": f["a2"].astype(np.float64).T @ dz3, "b3": np.sum(dz3)}
    da2 = np.outer(dz3, p["w3"])
    dz2 = da2 * (f["z2"] > 0)
    g["W2"] = r(f["a1"]).astype(np.float64).T @ dz2
    g["b2"] = dz2.sum(axis=0)
    da1 = dz2 @ r(p["W2"]).astype(np.float64).T
    dz1 = da1 * (f["z1"] > 0)
    g["W1"] = r(f["xp"]).astype(np.float64).T @ dz1
    g["b1"] = dz1.sum(axis=0)
    loss = -np.sum(w * (y * np.log(np.maximum(f["p"], 1e-12)) +
                        (1 - y) * np.log(np.maximum(1 - f["p"], 1e-12)))) / wsum
    return g, loss


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.standard_normal((NIN, H)) * np.sqrt(2.0 / NFEAT),
         "b1": np.zeros(H), "W2": rng.standard_normal((H, H)) * np.sqrt(2.0 / H),
         "b2": np.zeros(H), "w3": rng.standard_normal(H) * np.sqrt(1.0 / H), "b3": np.zeros(1)}
    p["W1"][NFEAT:] = 0.0
    return {k: v.astype(np.float32) for k, v in p.items()}
