"""Seeded input states for function-level parity tests (AoS, shape (n, 9)).

Covers the regimes the reference path meets: smooth Orszag-Tang-like
states, strong jumps, the HLLD degenerate branches (B_n = 0, UL = UR,
near-zero Alfven denominators), supersonic faces (sL >= 0, sR <= 0),
near-vacuum and non-physical states (negative pressure, NaN, inf).
"""
import numpy as np

GAMMA = 5.0 / 3.0


def prim_to_cons(W):
    W = np.asarray(W, dtype=np.float64)
    U = np.empty_like(W)
    rho, vx, vy, vz, p, bx, by, bz, psi = W.T
    U[:, 0] = rho
    U[:, 1] = rho * vx
    U[:, 2] = rho * vy
    U[:, 3] = rho * vz
    U[:, 4] = p / (GAMMA - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    U[:, 5:9] = W[:, 5:9]
    return U


def random_prims(rng, n, vmax=2.0, bmax=1.0):
    W = np.empty((n, 9))
    W[:, 0] = rng.uniform(0.05, 3.0, n)
    W[:, 1:4] = rng.uniform(-vmax, vmax, (n, 3))
    W[:, 4] = rng.uniform(0.01, 3.0, n)
    W[:, 5:8] = rng.uniform(-bmax, bmax, (n, 3))
    W[:, 8] = rng.uniform(-0.1, 0.1, n)
    return W


def face_pairs(seed=1234, n=4096):
    """Pairs (UL, UR) of conserved states with labelled regimes."""
    rng = np.random.default_rng(seed)
    WL = random_prims(rng, n)
    # Right states: small perturbations for half the pairs, independent for the rest.
    WR = np.where(rng.random((n, 1)) < 0.5, WL * (1.0 + 0.05 * rng.standard_normal((n, 9))),
                  random_prims(rng, n))
    WR[:, 0] = np.abs(WR[:, 0]) + 1e-3
    WR[:, 4] = np.abs(WR[:, 4]) + 1e-3
    UL, UR = prim_to_cons(WL), prim_to_cons(WR)
    k = n // 16
    # Degenerate normal field: B_n = 0 and equal psi on both sides.
    UL[0:k, 5] = 0.0
    UR[0:k, 5] = 0.0
    UL[0:k, 6] = 0.0
    UR[0:k, 6] = 0.0
    UL[0:k, 8] = UR[0:k, 8] = 0.0
    # Identical states (HLLD(U,U)).
    UR[k:2 * k] = UL[k:2 * k]
    # Supersonic to the right and to the left.
    for sl, sign in ((slice(2 * k, 3 * k), 1.0), (slice(3 * k, 4 * k), -1.0)):
        W = random_prims(rng, k, vmax=0.1, bmax=0.1)
        W[:, 1] = sign * 50.0
        W2 = W.copy()
        W2[:, 0] *= 1.1
        UL[sl], UR[sl] = prim_to_cons(W), prim_to_cons(W2)
    # Near-vacuum on one side.
    UR[4 * k:5 * k, 0] = 1e-13
    UR[4 * k:5 * k, 1:4] = 0.0
    # Negative-pressure (non-physical) right states: tiny E.
    UR[5 * k:6 * k, 4] = 1e-6
    # Strong shear with tiny B (near-degenerate Alfven denominator).
    W = random_prims(rng, k, bmax=1e-6)
    W2 = W.copy()
    W2[:, 2] = -W[:, 2]
    UL[6 * k:7 * k], UR[6 * k:7 * k] = prim_to_cons(W), prim_to_cons(W2)
    # Non-finite entries.
    UL[7 * k, 0] = np.nan
    UR[7 * k + 1, 4] = np.inf
    UL[7 * k + 2, 1] = -np.inf
    UL[7 * k + 3, :] = 0.0
    # Signed zeros in the transverse slots.
    UL[7 * k + 4:8 * k, 3] = -0.0
    UL[7 * k + 4:8 * k, 7] = -0.0
    return UL, UR


def cell_states(seed=99, n=4096):
    """Conserved states for recovery / wave-speed tests."""
    rng = np.random.default_rng(seed)
    U = prim_to_cons(random_prims(rng, n))
    k = n // 8
    U[0:k, 0] = rng.uniform(-1e-12, 2e-12, k)          # around the density floor
    U[k:2 * k, 4] = rng.uniform(0.0, 1e-3, k)          # negative recovered pressure
    U[2 * k, :] = np.nan
    U[2 * k + 1, 4] = np.inf
    U[2 * k + 2, :] = 0.0
    U[2 * k + 3:3 * k, 3] = -0.0
    return U
