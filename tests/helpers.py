"""Shared helpers for the tests (inputs and comparisons; no reference code)."""
import numpy as np

from rtnurbs import scenes


def pcg_hash_np(x):
    """PCG hash of uint32 arrays (the published Jarzynski-Olano PCG variant the
    reference uses as pcg_hash, math_utils.h:12-17)."""
    x = np.asarray(x, np.uint32)
    with np.errstate(over="ignore"):
        x = (x * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        w = (((x >> ((x >> np.uint32(28)) + np.uint32(4))) ^ x) * np.uint32(277803737)).astype(np.uint32)
        return ((w >> np.uint32(22)) ^ w).astype(np.uint32)


def jitter_np(n, seed, k):
    """Raster jitter of rays 0..n-1 on axis k, as the ray kernel defines it:
    24 bits of pcg(pcg(2r+k) ^ seed*golden) scaled by 2^-24."""
    r = np.arange(n, dtype=np.uint32)
    with np.errstate(over="ignore"):
        h = pcg_hash_np(pcg_hash_np(r * np.uint32(2) + np.uint32(k)) ^ np.uint32((seed * 0x9E3779B9) & 0xffffffff))
    return ((h >> np.uint32(8)).astype(np.float32) * np.float32(5.9604645e-08)).astype(np.float32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what=""):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.nonzero(bits(a) != bits(b))
    assert len(bad[0]) == 0, f"{what}: {len(bad[0])} differing values, first at {tuple(x[0] for x in bad)}: " \
                             f"{a[tuple(x[0] for x in bad)]} vs {b[tuple(x[0] for x in bad)]}"


def c2_small(n_surf=6, seed=3):
    """A few config-2-style random surfaces (mixed degree 2..5), unrefined."""
    rng = np.random.default_rng(500 + seed)
    return [scenes.random_surface(rng) for _ in range(n_surf)]


def nodes_equal(a, b):
    """Node arrays equal: links exactly, boxes by value (fminf/fmaxf may pick
    either signed zero when both operands compare equal)."""
    assert len(a) == len(b), (len(a), len(b))
    assert np.array_equal(a["left_first"], b["left_first"])
    assert np.array_equal(a["prim_count"], b["prim_count"])
    assert np.array_equal(a["b"], b["b"])
