"""Host-side checks of the batch ABI that need no GPU: admission sizes and force order."""
import numpy as np
import pytest

import d3fgpu
from bench import ba_graph, link_defaults, phyllotaxis


def job(n, nd, forces):
    return d3fgpu.Job(phyllotaxis(n, nd)[:, :nd], forces, ndim=nd)


def test_job_bytes_grow_with_forces():
    n = 1000
    src, dst = ba_graph(n, 2, 1)
    bias, s, d = link_defaults(n, src, dst)
    plain = job(n, 3, [d3fgpu.center()]).bytes()
    tree = job(n, 3, [d3fgpu.many_body(np.full(n, -30.0)), d3fgpu.center()]).bytes()
    full = job(n, 3, [d3fgpu.link(src, dst, s, d, bias), d3fgpu.many_body(np.full(n, -30.0)),
                      d3fgpu.collide(np.full(n, 3.0)), d3fgpu.center()]).bytes()
    assert 0 < plain < tree < full
    assert tree - plain >= 240 * n  # the tree's scratch share


def test_force_order_is_enforced():
    n = 100
    ok = job(n, 2, [d3fgpu.many_body(np.full(n, -30.0)), d3fgpu.center()])
    assert ok.bytes() > 0
    bad = job(n, 2, [d3fgpu.center(), d3fgpu.many_body(np.full(n, -30.0))])
    with pytest.raises(d3fgpu.FgError) as e:
        bad.bytes()
    assert e.value.status == -4
    twice = job(n, 2, [d3fgpu.center(), d3fgpu.center()])
    with pytest.raises(d3fgpu.FgError):
        twice.bytes()


def test_ba_graph_shape():
    src, dst = ba_graph(10000, 2, 7)
    assert len(src) == 2 * 9999
    assert np.all(dst < src)
    deg = np.bincount(np.concatenate([src, dst]), minlength=10000)
    assert deg.min() >= 2 and deg.max() > 50  # heavy tail of preferential attachment
