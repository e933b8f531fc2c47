"""Waveform segment storage and the segment records (CPU).

``add_segments`` is ``add`` for many segments at once, ``set_segments`` and
``copy_segments`` replace a generator's segments; all of them must leave
the same segments, gradient tags, amplitude bounds and device records as
the one-at-a-time form.  The records list each generator's segments in
ascending (t_start, t_end, amp, omega, phase) order with their tags.
"""

import math

import numpy as np
import pytest

from superspinsim_amd.sweep import (BatchSimulator, Waveform,
                                    waveform_records)

RNG = np.random.default_rng(11)


def _segments(count, shuffle=True):
    t = np.cumsum(RNG.uniform(1e-9, 5e-9, count + 1))
    rows = [(t[k], t[k + 1], RNG.normal(), RNG.uniform(0, 1e10),
             RNG.uniform(-3, 3)) for k in range(count)]
    tags = [(int(k % 7) - 1, float(RNG.uniform(0.5, 2))) for k in
            range(count)]
    order = RNG.permutation(count) if shuffle else np.arange(count)
    return [rows[k] for k in order], [tags[k] for k in order]


def _one_at_a_time(base, per_gen):
    w = Waveform(len(base), base)
    for g, (rows, tags) in enumerate(per_gen):
        for (ts, te, a, om, ph), (pa, dp) in zip(rows, tags):
            w.add(g, ts, te, a, om, ph, param=pa, dparam=dp)
    return w


def _arrays(base, per_gen):
    w = Waveform(len(base), base)
    for g, (rows, tags) in enumerate(per_gen):
        if not rows:
            continue
        r = np.array(rows)
        t = np.array(tags)
        w.add_segments(g, r[:, 0], r[:, 1], r[:, 2], r[:, 3], r[:, 4],
                       param=t[:, 0].astype(np.int64), dparam=t[:, 1])
    return w


def _same(a, b):
    assert a.segments == b.segments
    assert a.params == b.params
    assert np.array_equal(a.base, b.base)
    assert np.array_equal(a.abs_max(), b.abs_max())
    for x, y in zip(a.bounds(), b.bounds()):
        assert np.array_equal(x, y)
    ra = waveform_records([a], a.n_gen, with_tags=True)
    rb = waveform_records([b], b.n_gen, with_tags=True)
    assert ra[0].tobytes() == rb[0].tobytes()
    for x, y in zip(ra[1:], rb[1:]):
        assert np.array_equal(x, y)


def test_add_segments_equals_add():
    base = [1.0, -0.5, 0.0, 2.0]
    per_gen = [_segments(5), ([], []), _segments(30), _segments(1)]
    _same(_one_at_a_time(base, per_gen), _arrays(base, per_gen))


def test_set_and_copy_segments_equal_add():
    base = [0.0, 1.0, 3.0]
    per_gen = [_segments(4), _segments(9), _segments(2)]
    want = _one_at_a_time(base, per_gen)
    got = Waveform(3, base)
    for g, (rows, tags) in enumerate(per_gen):
        got.set_segments(g, rows, tags)
    _same(want, got)
    copy = Waveform(3, base)
    for g in range(3):
        copy.copy_segments(g, want, g)
    _same(want, copy)
    # untagged by default
    bare = Waveform(1).set_segments(0, per_gen[0][0])
    assert bare.params[0] == ((-1, 1.0),) * 4
    with pytest.raises(ValueError):
        Waveform(1).set_segments(0, per_gen[0][0], per_gen[0][1][:2])


def test_mixed_add_forms_keep_insertion_order():
    rows, tags = _segments(6, shuffle=False)
    w = Waveform(1)
    w.add(0, *rows[0], param=tags[0][0], dparam=tags[0][1])
    r = np.array(rows[1:4])
    w.add_segments(0, r[:, 0], r[:, 1], r[:, 2], r[:, 3], r[:, 4],
                   param=[t[0] for t in tags[1:4]],
                   dparam=[t[1] for t in tags[1:4]])
    for row, tag in zip(rows[4:], tags[4:]):
        w.add(0, *row, param=tag[0], dparam=tag[1])
    assert w.segments[0] == tuple(tuple(x) for x in rows)
    assert w.params[0] == tuple(tags)
    # read-only views: mutation fails loudly instead of being ignored
    with pytest.raises(AttributeError):
        w.segments[0].append(rows[0])
    assert w.segment_count(0) == 6
    assert math.isclose(w.abs_max()[0], max(abs(x[2]) for x in rows))


def test_records_sorted_with_their_tags():
    rows, tags = _segments(12)
    w = Waveform(2, [0.5, 0.0])
    w.set_segments(1, rows, tags)
    recs, segs, index, scale = waveform_records([w], 2, with_tags=True)
    assert recs["seg_begin"][0, 1] == 0 and recs["seg_count"][0, 1] == 12
    order = sorted(range(12), key=lambda k: rows[k])
    assert np.array_equal(segs, np.array([rows[k] for k in order]))
    assert index.tolist() == [tags[k][0] for k in order]
    assert scale.tolist() == [tags[k][1] for k in order]


def test_scalar_broadcast_and_empty_batch():
    w = Waveform(2)
    w.add_segments(1, [0.0, 1e-9], [1e-9, 2e-9], 0.25, omega=3.0)
    assert w.segments[1] == ((0.0, 1e-9, 0.25, 3.0, 0.0),
                             (1e-9, 2e-9, 0.25, 3.0, 0.0))
    assert w.params[1] == ((-1, 1.0), (-1, 1.0))
    w.add_segments(0, [], [], [])
    assert w.segment_count(0) == 0 and w.abs_max()[0] == 0.0


def test_kernel_waves_keep_segments_and_tags():
    rows, tags = _segments(8)
    w = Waveform(3, [1.0, 0.0, 0.7]).set_segments(1, rows, tags)
    kw = BatchSimulator.kernel_waves([w], [None, 1, 2])[0]
    assert kw.base.tolist() == [1.0, 0.0, 0.7]
    assert kw.segments[1] == w.segments[1] and kw.params[1] == w.params[1]
    assert kw.segment_count(0) == 0 and kw.segment_count(2) == 0
    assert kw.abs_max()[1] == w.abs_max()[1]
