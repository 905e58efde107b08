"""The mass anchor's density sum (lbmCpuReference.ts:433-440) does not depend on its order.

Every step form of the library (per-cell or split, 1, 2 or 4 cells per thread, one tunnel or a batch, per-kind
launches or the one persistent launch) reduces the post-boundary densities in a different order. They still find
the same anchor factor because every density summed is a float in [0.5, 1.8]: the collision clamps rho
(lbmCpuReference.ts:233) and the inlet and outlet cells take 0.85 rho + 0.15. Such a float is a multiple of 2^-24,
and FP64 holds every multiple of 2^-24 below 2^29, so each FP64 addition of partial sums is exact. The one-launch
form adds the same values as 64-bit integers in units of 2^-32. These checks pin that arithmetic on the CPU;
the GPU tests compare the step forms themselves with the anchor firing.
"""
import numpy as np


def _densities(n, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.8, n).astype(np.float32)
    rho[: n // 64] = np.float32(0.5)
    rho[n // 64: n // 32] = np.float32(1.8)
    return rng.permutation(rho)


def _exact_units(rho, frac_bits):
    """Sum of rho in units of 2^-frac_bits, as a Python integer (exact)."""
    q = rho.astype(np.float64) * float(2 ** frac_bits)
    assert np.all(q == np.round(q)), "a density is not a multiple of the unit"
    return int(q.astype(np.int64).sum(dtype=np.int64))


def _orders(rho, seed):
    """FP64 sums of rho in several orders and groupings."""
    d = rho.astype(np.float64)
    rng = np.random.default_rng(seed)
    out = [
        float(np.sum(d)),                                    # pairwise
        float(np.cumsum(d)[-1]),                             # sequential
        float(np.cumsum(d[::-1])[-1]),                       # sequential, reversed
        float(np.cumsum(d[rng.permutation(d.size)])[-1]),    # sequential, shuffled
    ]
    for block in (64, 256, 512, 1000):                       # block partials, then their sum (as the kernels)
        pad = (-d.size) % block
        parts = np.concatenate([d, np.zeros(pad)]).reshape(-1, block).sum(axis=1)
        out.append(float(np.cumsum(parts[rng.permutation(parts.size)])[-1]))
    return out


def test_fp64_density_sums_are_order_independent_and_exact():
    rho = _densities(1 << 21, 7)
    sums = _orders(rho, 11)
    exact = _exact_units(rho, 24)
    assert all(s == sums[0] for s in sums), sums
    assert sums[0] * 2.0 ** 24 == exact


def test_fixed_point_total_equals_the_fp64_sum():
    """The one-launch form's total: each thread's per-tile sum of a few densities in FP64 (exact), converted to
    2^-32 units with no rounding, added as integers, and read back as a double."""
    rho = _densities(1 << 20, 3)
    groups = rho.astype(np.float64).reshape(-1, 4).sum(axis=1)       # 4 cells per thread and tile
    q = groups * 2.0 ** 32
    assert np.all(q == np.round(q))
    total = int(q.astype(np.int64).sum(dtype=np.int64))
    assert float(total) / 2.0 ** 32 == float(np.sum(rho.astype(np.float64)))
    assert total == _exact_units(rho, 32)


def test_bound_at_the_largest_grid():
    """1024 x 512 x 512 cells at the largest density stay below 2^29, where FP64 stops holding every multiple of
    2^-24: the largest total has 29 integer and 24 fraction bits, 53 in all."""
    n = 1024 * 512 * 512
    assert n * 1.8 < 2.0 ** 29
    big = 1.8 * n
    assert (big + 2.0 ** -24) - big == 2.0 ** -24


def test_order_dependence_returns_outside_the_clamp():
    """Negative control: densities far below the clamp are not multiples of 2^-24, and the orders disagree."""
    rng = np.random.default_rng(5)
    rho = np.concatenate([_densities(1 << 16, 1), rng.uniform(1e-6, 1e-3, 1 << 16).astype(np.float32)])
    sums = _orders(rng.permutation(rho), 2)
    assert len(set(sums)) > 1
