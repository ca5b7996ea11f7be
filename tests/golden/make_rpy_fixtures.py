"""Writes rpy_n256.npz: a polydisperse sphere configuration, 64 force
right-hand sides and U = M F computed with numpy's FP64 dense matrix
product (the library routine the mobility is checked against).

  python tests/golden/make_rpy_fixtures.py

Configuration (as the 4,096-sphere workload, at 256 spheres): centres
uniform in a cube of side 2 (256/4096)^(1/3) m, rejection-sampled to a gap
of at least 0.1 max(a_i, a_j); radii uniform in [0.01, 0.03] m; mu = 1;
F ~ N(0, 1) rounded to FP32."""
from pathlib import Path

import numpy as np


def spheres(n, side, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    rad = np.zeros(n)
    k = 0
    while k < n:
        x = rng.uniform(0, side, 3)
        a = rng.uniform(0.01, 0.03)
        if k:
            r = np.linalg.norm(pos[:k] - x, axis=1)
            gap = r - rad[:k] - a
            if np.any(gap < 0.1 * np.maximum(rad[:k], a)):
                continue
        pos[k], rad[k] = x, a
        k += 1
    return pos.astype(np.float32), rad.astype(np.float32)


def dense_rpy(pos, rad, mu):
    n = len(pos)
    p = pos.astype(np.float64)
    a = rad.astype(np.float64)
    d = p[None, :, :] - p[:, None, :]                 # x_q - x_p
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, 1.0)
    r = np.sqrt(r2)
    s2 = a[:, None] ** 2 + a[None, :] ** 2
    c0 = 1.0 / (8 * np.pi * mu * r)
    ci = c0 * (1 + s2 / (3 * r2))
    cr = c0 * (1 - s2 / r2) / r2
    M = ci[:, :, None, None] * np.eye(3) + cr[:, :, None, None] * d[:, :, :, None] * d[:, :, None, :]
    for i in range(n):
        M[i, i] = np.eye(3) / (6 * np.pi * mu * a[i])
    return M.transpose(0, 2, 1, 3).reshape(3 * n, 3 * n)


def main():
    n, K, mu = 256, 64, 1.0
    pos, rad = spheres(n, 2.0 * (n / 4096) ** (1 / 3), 20261206)
    rng = np.random.default_rng(7)
    F = rng.standard_normal((3 * n, K)).astype(np.float32)
    M = dense_rpy(pos, rad, mu)
    U = M @ F.astype(np.float64)
    out = Path(__file__).resolve().parent / "rpy_n256.npz"
    np.savez_compressed(out, pos=pos, rad=rad, mu=mu, F=F, U=U)
    print(f"wrote {out}: n={n} K={K} |U|max={np.abs(U).max():.3e}")


if __name__ == "__main__":
    main()
