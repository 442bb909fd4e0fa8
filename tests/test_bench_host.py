"""Host-side pieces of bench.py that feed the reported figures (CPU):

* ``stratified_bonds``: the CPU baseline's sample of the GPU sweep's bond
  shapes -- the groups partition every bond once, follow the cost order,
  each is represented by one of its own bonds, and carries the mean Lanczos
  matvec count of its bonds;
* ``measured_traffic``: the committed H_eff traffic profile the bench line
  scales is found and is a measurement of this workload (bytes above the
  algorithmic minimum, with the command that produced it).
"""

import numpy as np

import bench


def synthetic_sweep(rng, n=64, d=4, cap=1024):
    """(shapes, iters) of a two-site sweep: bond b between sites b and b+1,
    dimensions growing from the edges to the cap."""
    shapes, iters = [], []
    for half in (range(1, n), range(n - 1, 0, -1)):
        for b in half:
            dl = min(cap, d ** min(b - 1, n - b + 1), int(rng.integers(cap // 2, cap + 1)))
            dr = min(cap, d ** min(b + 1, n - b - 1), int(rng.integers(cap // 2, cap + 1)))
            shapes.append((b, max(1, dl), d, d, max(1, dr)))
            iters.append(int(rng.integers(1, 5)))
    return shapes, iters


def cost(s):
    _, Dl, d1, d2, Dr = s
    return (min(Dl * d1, d2 * Dr), Dl * Dr)


def test_stratified_bonds_partition_and_order():
    rng = np.random.default_rng(4)
    for strata in (1, 2, 4, 7):
        shapes, iters = synthetic_sweep(rng)
        sample = bench.stratified_bonds(shapes, iters, strata)
        assert len(sample) == strata
        assert sum(cnt for _, _, _, cnt in sample) == len(shapes)
        rows = sorted(zip(shapes, iters), key=lambda r: cost(r[0]))
        start = 0
        prev = None
        for Dl, Dr, it, cnt in sample:
            grp = rows[start:start + cnt]
            start += cnt
            # the representative is a bond of its own group
            assert any(s[1] == Dl and s[4] == Dr for s, _ in grp)
            # mean matvec count of the group, at least one
            assert it == max(1, int(round(sum(i for _, i in grp) / len(grp))))
            c = cost(next(s for s, _ in grp if s[1] == Dl and s[4] == Dr))
            assert prev is None or c >= prev
            prev = c
        assert start == len(shapes)


def test_stratified_bonds_more_strata_than_bonds():
    shapes = [(1, 4, 4, 4, 16), (2, 16, 4, 4, 4)]
    sample = bench.stratified_bonds(shapes, [2, 3], strata=4)
    assert sorted((Dl, Dr, it, cnt) for Dl, Dr, it, cnt in sample) == [(4, 16, 2, 1), (16, 4, 3, 1)]


def test_measured_traffic_profile():
    prof = bench.measured_traffic("hubbard", 1024, "fp32")
    assert prof is not None
    ratio = prof["traffic_bytes_per_apply"] / prof["algorithmic_bytes_per_apply"]
    assert 1.0 < ratio < 10.0
    assert "bench.py" in prof["profiled_command"] and prof["applies"] > 0
    assert bench.measured_traffic("hubbard", 999, "fp32") is None
