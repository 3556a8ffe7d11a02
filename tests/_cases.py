"""Seeded synthetic inputs shared by the CPU and GPU tests."""
import math

import numpy as np

from nuclearcgm_amd import config as C


def grid(n, box):
    a = 0.5 * box
    h = 2 * a / (n - 1)
    x = -a + h * np.arange(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")  # flattened index = i + n (j + n k)
    return a, h, X.reshape(-1), Y.reshape(-1), Z.reshape(-1)


def random_fields(n, seed=1):
    rng = np.random.default_rng(seed)
    N3 = n ** 3
    f = 20.0 + 5.0 * rng.random(N3)
    gf = rng.standard_normal(3 * N3)
    U = -50.0 + 10.0 * rng.standard_normal(N3)
    W = rng.standard_normal(3 * N3)
    return f, gf, U, W


def random_states(n, m, seed=2):
    rng = np.random.default_rng(seed)
    rows = 2 * n ** 3
    return np.asfortranarray(rng.standard_normal((rows, m)) + 1j * rng.standard_normal((rows, m)))


def smooth_fields(n, box, seed=3):
    """Woods-Saxon-like smooth fields with a position-dependent mass and a
    radial spin-orbit form factor (physically shaped, seeded wobble)."""
    a, h, X, Y, Z = grid(n, box)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rng = np.random.default_rng(seed)
    R0, d = 3.0, 0.6
    fws = 1.0 / (1.0 + np.exp((r - R0) / d))
    dfws = -fws * (1 - fws) / d  # d/dr
    hb2m = C.HBAR ** 2 / (2 * C.MNUC)
    c = 0.2 * hb2m
    f = hb2m - c * fws
    gr = -c * dfws
    rr = np.where(r > 0, r, 1.0)
    gf = np.concatenate([gr * X / rr, gr * Y / rr, gr * Z / rr])
    U = -50.0 * fws * (1 + 0.05 * rng.standard_normal(r.shape))
    wso = 30.0 * dfws
    W = np.concatenate([wso * X / rr, wso * Y / rr, wso * Z / rr])
    return f, gf, U, W


def o16_config(n=16, box=18.0, pairing=False, M=None, mix=0.25, max_iter=400, tol=1e-6):
    d = {"functional": "SLy4", "maxIterations": max_iter, "energyTol": tol, "densityMix": mix,
         "box": {"boxSize": box, "axisMeshPoints": n}, "nucleus": {"A": 16, "Z": 8}}
    if pairing:
        d["pairing"] = {"type": "BCS", "window": 5.0, "V0": 250.0, "HFbasisSize": M or 12}
    cfg = C.config_from_input(d)
    if not pairing:
        cfg.M_n = cfg.M_p = M or 8
    return cfg


def guess(cfg, seed=7, amp=1e-3, ho=None):
    """HO |n l j m_j> start (hbar omega = 41 A^-1/3), trapezoid-normalised,
    plus a seeded uniform +-amp perturbation of Re and Im, renormalised.
    `ho` is the HO generator (oracle or library) returning Euclidean-normalised
    columns."""
    M = max(cfg.M_n, cfg.M_p)
    b = C.ho_length(cfg.A)
    g = ho(cfg.n, cfg.box_size, M, b)
    g = trap_normalize(g, cfg.n, cfg.box_size)
    rng = np.random.default_rng(seed)
    g = g + amp * (rng.uniform(-1, 1, g.shape) + 1j * rng.uniform(-1, 1, g.shape))
    g = trap_normalize(np.asfortranarray(g), cfg.n, cfg.box_size)
    e0 = C.ho_energies(M, cfg.A)
    return (np.asfortranarray(g[:, : cfg.M_n]), np.asfortranarray(g[:, : cfg.M_p]), e0[: cfg.M_n], e0[: cfg.M_p])


def trap_weights(n, box):
    a = 0.5 * box
    h = 2 * a / (n - 1)
    w1 = np.ones(n)
    w1[0] = w1[-1] = 0.5
    return (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).reshape(-1) * h ** 3


def trap_normalize(psi, n, box):
    w = trap_weights(n, box)
    psi = np.array(psi, dtype=np.complex128, order="F")
    for c in range(psi.shape[1]):
        d = np.abs(psi[0::2, c]) ** 2 + np.abs(psi[1::2, c]) ** 2
        psi[:, c] /= math.sqrt(float(np.dot(w, d)))
    return psi
