"""The chunked moving averages (buff.add_chunk over lbm_mv_avg_chunk, host
code of the D2Q9 library) against the per-iteration restatement of
lbm/src/utils/buff.py:51-77 (buff.add + buff.mv_avg, np.sum window sums): the
averages, the returned observable and growth, and the convergence counter and
flag agree bit for bit for histories of several lengths split into chunks of
several sizes.  Runs on the CPU."""
import numpy as np
import pytest

from lbm_amd.buff import buff


def _series(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    return 3.2 + 0.3 * np.sin(t * 0.013) + 1e-3 * rng.standard_normal(n) - 2.0 * np.exp(-t / 50.0)


def _reference(vals, dt, ct, nb):
    b = buff("drag", dt, ct, nb, "/tmp/")
    out = []
    for v in vals:
        b.add(v)
        out.append(b.mv_avg())
    return b, np.array(out)


@pytest.mark.parametrize("n,chunks", [(9, [9]), (40, [1, 3, 36]), (2500, [1000, 1, 7, 1492]),
                                      (30001, [1000] * 30 + [1])])
def test_add_chunk_equals_mv_avg(n, chunks):
    assert sum(chunks) == n
    vals = _series(n, n)
    dt, ct, nb = 0.0123, 5e-3, 40
    ref, ref_out = _reference(vals, dt, ct, nb)
    b = buff("drag", dt, ct, nb, "/tmp/")
    got, i = [], 0
    for k in chunks:
        obs, growth = b.add_chunk(vals[i:i + k])
        got.append(np.stack([obs, growth], 1))
        i += k
    got = np.concatenate(got)
    assert np.array_equal(got, ref_out)
    for name in ("buff", "avg1_buff", "avg2_buff", "avg3_buff"):
        assert np.array_equal(getattr(b, name), getattr(ref, name)), name
    assert (b.it, b.obs, b.obs_cv_cnt, b.obs_cv) == (ref.it, ref.obs, ref.obs_cv_cnt, ref.obs_cv)


def test_convergence_flag_sets_on_the_same_iteration():
    """A history that settles: obs_cv must switch on at the same iteration
    (obs_cv_it) on both paths, inside a chunk."""
    n = 6000
    vals = 1.0 + np.exp(-np.arange(n) / 300.0)
    dt, ct, nb = 1.0, 1e-6, 200
    ref, _ = _reference(vals, dt, ct, nb)
    b = buff("lift", dt, ct, nb, "/tmp/")
    for i0 in range(0, n, 512):
        b.add_chunk(vals[i0:i0 + 512])
    assert ref.obs_cv and ref.obs_cv_it is not None and ref.obs_cv_it % 512 != 0
    assert (b.obs_cv, b.obs_cv_it) == (ref.obs_cv, ref.obs_cv_it)
