"""rtx_trace_closest_noalloc (the allocation-free closest-hit query) against
rtx_trace_closest on the same rays: both run the same traversal kernel on the
same inputs, so hits must agree bit for bit.  Two calls in a row, the second
with fewer rays on another scene, so that stale contents of the library's
fixed device array would show; then the size limit."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from parity import log

pytestmark = pytest.mark.gpu


def _rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    org = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rtxamd.rays_array(org, d, 1e-3, 1e6)


def _same(a, b):
    t, p, u, v = a
    t2, p2, u2, v2 = b
    return (np.array_equal(t.view(np.uint32), t2.view(np.uint32)) and np.array_equal(p, p2)
            and np.array_equal(u.view(np.uint32), u2.view(np.uint32))
            and np.array_equal(v.view(np.uint32), v2.view(np.uint32)))


def test_noalloc_query_equals_allocating_query():
    with rtxamd.Renderer(0) as r:
        room = r.upload(rtxamd.HostScene.generate(abi.SCENE_ROOM, 1))
        cornell = r.upload(rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1))
        n1 = abi.RTX_NOALLOC_MAX_RAYS
        log(f"no-alloc query: room, {n1} rays")
        rays1 = _rays(n1, [-10, 0, -10], [10, 10, 10], seed=11)
        a1 = r.trace_closest_noalloc(room, rays1)
        b1 = r.trace_closest(room, rays1)
        assert _same(a1, b1)
        assert (a1[1] >= 0).mean() > 0.5  # most rays hit something in the closed room
        log("no-alloc query: cornell, 1000 rays")
        rays2 = _rays(1000, [-1, -1, 0], [1, 1, 2], seed=12)
        a2 = r.trace_closest_noalloc(cornell, rays2)
        b2 = r.trace_closest(cornell, rays2)
        assert _same(a2, b2)
        assert len(a2[0]) == 1000
        # the same rays again on the same scene: the result does not depend on the call before
        a3 = r.trace_closest_noalloc(room, rays1)
        assert _same(a3, b1)
        # empty call, then the limit
        e = r.trace_closest_noalloc(room, rays1[:0])
        assert len(e[0]) == 0
        big = np.concatenate([rays1, rays1[:1]])
        with pytest.raises(rtxamd.RtxError) as ex:
            r.trace_closest_noalloc(room, big)
        assert ex.value.code == abi.RTX_E_INVALID
        room.close()
        cornell.close()
