"""Reconstructions of one size on many streams with parameters that change
from call to call (GPU).

A size keeps at most four pooled workspaces, so a fifth and sixth stream
take over busy ones and queue behind their previous users' work.  When the
HIO parameters change, the workspace's captured cycle graph is rebuilt; the
old graph may still have replays queued on the previous user's stream and
must outlive them.  Every output must equal the same reconstruction run
alone, bit for bit, in graph and in eager mode.
"""
import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _alone(M, beta, cut):
    x = M.copy()
    im.shrink_wrap(x, n_cycles=3, n_hio_cycles=6, target_error=FLT_MIN, hio_beta=beta, cutoff=cut)
    return x


@pytest.mark.parametrize("graphs", [1, 0])
def test_alternating_parameters_on_six_streams(graphs):
    obj, M = syn.workload(192, 192, n_shapes=8, central=96, index=31)
    combos = [(0.9, 0.2), (0.7, 0.2), (0.9, 0.25), (0.7, 0.25), (0.8, 0.2), (0.6, 0.3)]
    refs = {c: _alone(M, *c) for c in combos}
    n = M.size
    src = im.DeviceBuffer.from_array(M)
    bufs = [im.DeviceBuffer(4 * n) for _ in range(6)]
    streams = [im.stream_create() for _ in range(6)]
    im.set_option(im.OPT_CYCLE_GRAPHS, graphs)
    try:
        for rnd in range(4):
            order = combos[rnd % 6:] + combos[:rnd % 6]
            for i, (s, b) in enumerate(zip(streams, bufs)):
                beta, cut = order[i]
                im.check(im.lib().imresh_memcpy_d2d(b.ptr, src.ptr, 4 * n, s))
                p = im.default_params(n_cycles=3, n_hio_cycles=6, target_error=FLT_MIN, hio_beta=beta,
                                      cutoff=cut)
                im.shrink_wrap_device(b.ptr, 192, 192, p, stream=s, want_result=False)
            for s in streams:
                im.sync(s)
            for i, b in enumerate(bufs):
                out = b.download(np.float32, M.shape)
                assert np.array_equal(out, refs[order[i]]), (rnd, i, order[i])
    finally:
        im.set_option(im.OPT_CYCLE_GRAPHS, 1)
        for s in streams:
            im.stream_destroy(s)
