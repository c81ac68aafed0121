"""CPU restatement of the device lognormal random field (test infrastructure only).

kappa = exp(s g) per cell, g = white noise (counter-based hash of (seed, index),
Box-Muller) on a grid extended by R cells per side, convolved along x, y, z
with a Gaussian of standard deviation corr_len / sqrt(2) (in cells: divided by
the cell size), truncated at 4 standard deviations and normalised to sum 1,
then divided by the exact standard deviation sqrt(prod_d sum_k w_k^2).
Only tests/ may import this module.
"""
import numpy as np

M64 = (1 << 64) - 1


def _mix64(z):
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def unif(seed, i):
    i = np.asarray(i, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64((seed * 0x9E3779B97F4A7C15) & M64) + (i + np.uint64(1)) * np.uint64(0xD1B54A32D192ED03)
    z = _mix64(z)
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def noise(seed, n):
    i = np.arange(n, dtype=np.uint64)
    u1, u2 = unif(seed, 2 * i), unif(seed, 2 * i + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def gauss_weights(sd_cells):
    R = max(1, int(np.ceil(4.0 * sd_cells)))
    k = np.arange(-R, R + 1)
    w = np.exp(-0.5 * (k / sd_cells) ** 2)
    return w / w.sum(), R


def lognormal_cells(cells, h, seed, corr_len, log_scale):
    """kappa per cell, array of shape cells[::-1] (z, y, x) flattened x fastest."""
    dim = len(cells)
    ws, Rs = [], []
    for d in range(dim):
        w, R = gauss_weights(corr_len / np.sqrt(2.0) / h[d])
        ws.append(w)
        Rs.append(R)
    ext = [cells[d] + 2 * Rs[d] for d in range(dim)]
    g = noise(seed, int(np.prod(ext))).reshape(ext[::-1])  # (z, y, x) with x fastest
    for d in range(dim):
        ax = dim - 1 - d
        w, R = ws[d], Rs[d]
        n_out = g.shape[ax] - 2 * R
        acc = np.zeros(g.shape[:ax] + (n_out,) + g.shape[ax + 1:])
        for k in range(2 * R + 1):
            sl = [slice(None)] * g.ndim
            sl[ax] = slice(k, k + n_out)
            acc = acc + w[k] * g[tuple(sl)]
        g = acc
    var = np.prod([np.sum(w * w) for w in ws])
    return np.exp(log_scale / np.sqrt(var) * g).ravel()
