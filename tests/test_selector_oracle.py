"""CPU tests pinning the selector oracle (oracle/selector_oracle.py).

  * bf16 rounding: known bit patterns, round-to-nearest-even ties;
  * MLP gradients: central finite differences of the loss (float64 path);
  * tile features: closed forms on uniform and single-mode states.
"""
import numpy as np
import pytest

import selector_oracle as so


def test_bf16_rounding_bits():
    vals = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-9, 1.0 + 2.0**-9, -2.5, 0.0], dtype=np.float32)
    got = so.bf16(vals)
    # 1+2^-8 is exactly representable in bf16 (8 mantissa bits -> spacing 2^-7? no: 7 bits)
    # bf16 has 7 stored mantissa bits: spacing at 1.0 is 2^-7.
    exp = np.array([1.0, 1.0, 1.0 + 2.0**-7, 1.0, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(got, exp)
    # ties round to even: 1 + 2^-8 is halfway between 1 and 1+2^-7 -> 1 (even)
    assert so.bf16(np.float32(1.0 + 2.0**-8)) == np.float32(1.0)
    assert so.bf16(np.float32(1.0 + 2.0**-7 + 2.0**-8)) == np.float32(1.0 + 2.0**-6)
    assert np.isnan(so.bf16(np.float32(np.nan)))


def test_mlp_gradients_finite_differences():
    rng = np.random.default_rng(0)
    p = {k: v.astype(np.float64) for k, v in so.init_params(3).items()}
    x = rng.standard_normal((40, so.NFEAT))
    y = (rng.random(40) < 0.3).astype(np.float64)
    w = np.where(y > 0, 3.0, 1.0)

    def loss(q):
        return so.mlp_grads(x, y, w, q, bf16_ops=False)[1]

    g, _ = so.mlp_grads(x, y, w, p, bf16_ops=False)
    eps = 1e-6
    for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
        flat = p[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(12, flat.size), replace=False)
        for k in idx:
            if name == "W1" and k // so.H >= so.NFEAT:
                continue  # padding rows carry zero input
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[name].reshape(-1)[k] += eps
            lp = loss(q)
            q[name].reshape(-1)[k] -= 2 * eps
            lm = loss(q)
            num = (lp - lm) / (2 * eps)
            ana = np.asarray(g[name]).reshape(-1)[k]
            assert abs(num - ana) <= 1e-6 + 1e-4 * abs(num), (name, k, num, ana)


def test_features_uniform_state():
    """A uniform state: zero neighbour jumps, beta from the closed form."""
    n = 32
    rho, p, vx, bx, by = 1.3, 0.7, 0.4, 0.3, 0.4
    u = np.zeros((9, n, n))
    u[0] = rho
    u[1] = rho * vx
    u[5], u[6] = bx, by
    u[4] = p / (so.GAMMA - 1) + 0.5 * rho * vx * vx + 0.5 * (bx * bx + by * by)
    f = so.tile_features(u, 16, 15, ch=2.0)
    assert f.shape == (2 * 3, so.NFEAT)
    beta = 2 * p / (bx * bx + by * by)
    assert np.allclose(f[:, 0], np.log10(beta)) and np.allclose(f[:, 1], np.log10(beta))
    assert np.allclose(f[:, 2:5], -12.0)
    assert np.allclose(f[:, 5], vx / np.sqrt(so.GAMMA * p / rho))
    assert np.allclose(f[:, 8], np.log10(min(bx, by) / 0.5 + 1e-12))
    assert np.allclose(f[:, 9], vx / 2.0)
    assert np.allclose(f[:, 11], np.log10(0.5 * rho * vx * vx / (0.5 * 0.25) + 1e-12))


def test_tile_errors_closed_form():
    b = np.ones((9, 8, 8))
    a = b.copy()
    a[4, 2, 3] += 0.5          # one cell of tile 0 (tiles 4x4)
    e = so.tile_errors(a, b, 4, 4)
    assert e.shape == (4,)
    assert e[0] == pytest.approx(0.25)   # 0.5 / (|1| + rms 1)
    assert np.all(e[1:] == 0.0)
    # vector components are scaled by the vector's RMS: |m| = sqrt(3) here
    a2 = b.copy()
    a2[2, 6, 6] += 0.5                   # momentum y, tile 3
    e2 = so.tile_errors(a2, b, 4, 4)
    assert e2[3] == pytest.approx(0.5 / (1.0 + np.sqrt(3.0)))
    assert np.all(e2[:3] == 0.0)
    # the GLM potential does not enter
    a[8, 5, 5] += 3.0
    assert np.array_equal(so.tile_errors(a, b, 4, 4), e)


def test_features_psi_divb_fast_speed_closed_forms(oracle):
    """Features 6 (|psi|/|B|), 7 (|div B| h/|B|) and 10 (c_f/c_h) on a
    single-mode field Bx = b0 + a sin(2 pi x), By = c, with a constant psi:
    the central difference of a sine is cos(2 pi x) sin(2 pi dx)/dx, and c_f
    of a field perpendicular to x is the reference's fast speed
    (mhdFastSpeedX, through the pinned physics oracle)."""
    nx, ny = 48, 30
    rho, p, psi0, b0, amp, c = 1.1, 0.6, 0.02, 0.3, 0.2, 0.25
    x = (np.arange(nx) + 0.5) / nx
    X = np.broadcast_to(x, (ny, nx))
    bx = b0 + amp * np.sin(2 * np.pi * X)
    u = np.zeros((9, ny, nx))
    u[0] = rho
    u[5], u[6] = bx, c
    u[4] = p / (so.GAMMA - 1) + 0.5 * (bx * bx + c * c)
    u[8] = psi0
    ch = 1.7
    tx, ty = 16, 15
    f = so.tile_features(u, tx, ty, ch=ch)
    bmag = np.sqrt(bx * bx + c * c)
    divb = amp * np.cos(2 * np.pi * X) * np.sin(2 * np.pi / nx) * nx   # central difference
    divb_n = np.abs(divb) * (1.0 / nx) / bmag
    tiles_x = nx // tx
    for t in range(f.shape[0]):
        j0, i0 = (t // tiles_x) * ty, (t % tiles_x) * tx
        sl = (slice(j0, j0 + ty), slice(i0, i0 + tx))
        assert f[t, 6] == pytest.approx(np.log10(psi0 / (bmag[sl].max() + 1e-12) + 1e-12), abs=1e-12)
        assert f[t, 7] == pytest.approx(np.log10(divb_n[sl].max() + 1e-12), abs=1e-9)
    # c_f/c_h with Bx = 0: the x fast speed of the reference's formula
    u2 = u.copy()
    u2[5] = 0.0
    u2[6] = bmag
    u2[4] = p / (so.GAMMA - 1) + 0.5 * bmag * bmag
    f2 = so.tile_features(u2, tx, ty, ch=ch)
    W = np.zeros((nx * ny, 9))
    W[:, 0], W[:, 4], W[:, 6] = rho, p, bmag.reshape(-1)
    cf = oracle.fast_speed(W, direction=0).reshape(ny, nx)
    for t in range(f2.shape[0]):
        j0, i0 = (t // tiles_x) * ty, (t % tiles_x) * tx
        sl = (slice(j0, j0 + ty), slice(i0, i0 + tx))
        assert f2[t, 10] == pytest.approx(cf[sl].max() / ch, rel=1e-12)
