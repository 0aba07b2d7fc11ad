"""The move step's fp64 arithmetic on the device, bit for bit.

scd_eval_improvement runs dev.hpp's improvement() -- the device function
every move kernel inlines -- on the rows of tests/golden/improve_table.npz,
whose results the reference's own calculate_improvement (src/scd.cpp:173-219,
built from its sources) wrote, including 1,000 rows whose int32 products wrap
and the r <= 2 / empty-community edge cases.  scd_eval_candidate adds the
candidate set-up of src/scd.cpp:280-307 (own community without the vertex,
halved internal edges, p_in from the size) and is compared with the oracle's
candidate step, which the reference's per-iteration labels pin."""
import numpy as np
import pytest

import oracle_bind as orc
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x, np.float64).view(np.uint64)


def test_improvement_table_bit_exact_on_device(engine):
    t = np.load(f"{GOLDEN}/improve_table.npz")
    print(f"[gpu] eval_improvement {len(t['r'])} rows", flush=True)
    got = engine.eval_improvement(t["r"], t["d_in"], t["d_out"], t["c_out"], t["p_in"], t["p_ext"])
    bad = np.flatnonzero(bits(got) != t["bits"])
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    # the table holds wrapping rows and both signs of the result
    d_out = t["d_out"].astype(np.int64)
    assert (d_out * (d_out - 1) >= 2 ** 31).sum() >= 500
    assert (got > 0).any() and (got < 0).any() and (got == 0).any()


def candidate_rows(seed, count):
    rng = np.random.default_rng(seed)
    size = rng.integers(0, 300, count).astype(np.int32)
    deg = rng.integers(1, 3000, count).astype(np.uint32)
    freq = np.minimum(rng.integers(1, 200, count), deg).astype(np.uint32)
    ext = rng.integers(0, 20000, count).astype(np.int32)
    inte = (2 * rng.integers(0, 5000, count)).astype(np.int32)
    label = rng.integers(0, 50, count).astype(np.uint32)
    my = np.where(rng.random(count) < 0.4, label, rng.integers(0, 50, count)).astype(np.uint32)
    # hub-sized rows: wrapping products and boundary counts near 2^31
    k = count // 5
    deg[:k] = rng.integers(40000, 400000, k)
    freq[:k] = np.minimum(rng.integers(1, 70000, k), deg[:k])
    ext[:k] = rng.integers(2 ** 30, 2 ** 31 - 1, k)
    size[:k] = rng.integers(0, 70000, k)
    # tiny communities (size 0, 1, 2 after removing the vertex)
    size[k:2 * k] = rng.integers(0, 4, k)
    return size, ext, inte, label, freq, deg, my


@pytest.mark.parametrize("seed", [1, 2])
def test_candidate_setup_bit_exact_on_device(engine, seed):
    rows = candidate_rows(seed, 4000)
    for cc in (0.0, 0.3183098861837907, 0.6):
        got = engine.eval_candidate(*rows, cc)
        want = np.array([orc.candidate_improvement(int(s), int(e), int(i), int(l), int(f), int(d), int(m), cc)
                         for s, e, i, l, f, d, m in zip(*rows)])
        bad = np.flatnonzero(bits(got) != bits(want))
        assert bad.size == 0, f"cc={cc}: {bad.size} rows differ, first {bad[:5]}"
        assert (got != 0).sum() > 1000
