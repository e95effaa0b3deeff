"""The label corners of a march sample are the cell record's labels.

The reference point-samples the segmentation at floor(p + d) for the 8 corner
offsets d (deepdrr/projector/project_kernel.cu:421-425); the kernels read the
labels stored for cell clamp(floor(p), -1, n-1), i.e. L[clamp(floor(p) + d)].
The two agree because every sample position is p = fl(q - 1) for an FP32 q
(q = fma(alpha, r, s)), and then fl(p + 1) = p + 1 exactly while |p| < 2^22;
beyond that the sample lies so far outside any volume (dimensions <= 16384)
that floor(p) + 1, floor(p + 1) and floor(p) clamp to the same edge voxel.
These tests check both statements on FP32 arithmetic directly.
"""
import numpy as np

F = np.float32


def _positions(q):
    q = np.asarray(q, F)
    return (q - F(1)).astype(F)


def _corner_ok(px):
    b = np.floor(px)
    return np.floor((px + F(1)).astype(F)) == (b + F(1)).astype(F)


def test_floor_p_plus_one_is_floor_p_plus_one_below_2_22():
    rng = np.random.default_rng(7)
    n = 2_000_000
    ints = rng.integers(-(1 << 22), 1 << 22, n).astype(F)
    qs = [
        rng.uniform(-(2.0 ** 22), 2.0 ** 22, n),
        rng.uniform(-2.0, 2.0, n),
        rng.uniform(-1e-3, 1e-3, n),
        ints, np.nextafter(ints, F(-np.inf)), np.nextafter(ints, F(np.inf)),
        ints + F(0.5), np.nextafter(ints + F(0.5), F(-np.inf)),
        # a strided sweep over every FP32 bit pattern
        np.arange(0, 1 << 32, 211, dtype=np.uint64).astype(np.uint32).view(F),
    ]
    for q in qs:
        q = np.asarray(q, F)
        q = q[np.isfinite(q) & (np.abs(q) < F(2.0 ** 22))]
        px = _positions(q)
        assert _corner_ok(px).all()


def test_the_identity_needs_positions_of_that_form():
    # A position that is not fl(q - 1) can break it: p = -2^-30 gives
    # fl(p + 1) = 1 = floor(p) + 2.  Such a p is unreachable from a float q.
    px = np.array([-(2.0 ** -30)], F)
    assert not _corner_ok(px).all()
    q = np.nextafter(F(1), F(0))                 # the float just below 1
    assert _positions([q])[0] == F(-(2.0 ** -24))  # the closest reachable p below 0


def test_far_positions_clamp_to_the_same_voxel():
    n = 16384
    for px in (F(2.0 ** 22), F(3.0e7), F(-(2.0 ** 22)), F(-3.0e7)):
        b = np.floor(px)
        regular = np.clip(np.array([b, b + 1]), 0, n - 1)
        literal = np.clip(np.array([np.floor(px), np.floor((px + F(1)).astype(F))]), 0, n - 1)
        assert (regular == literal).all()
