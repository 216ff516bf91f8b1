"""Reentrancy of the cone backprojector (GPU test).

Calls that bring no workspace share one scaled-sinogram scratch per
geometry; the library serialises them on the device (a lock around the
enqueue and an event between callers).  Two backprojections of different
sinograms on one geometry, enqueued back to back on two streams, and from
two host threads with a stream each, must equal the serial results bit for
bit (the BP is deterministic: one wavefront per voxel column, records in a
fixed order)."""

import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from geometries import CB, centred

pytestmark = pytest.mark.gpu


def _geom():
    from ccpi.reconstruction.projectors import Geometry
    n = 96
    g = CB((n, n, n), (1, 1, 1), np.linspace(0, 2 * np.pi, 180, endpoint=False),
           centred(160, 1.2), centred(160, 1.2), 200.0, 400.0)
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape, cone=True,
                    source=g.source, detector_x=g.det_x)


def _serial(G, sinos):
    from ccpi.reconstruction.projectors import backward_project
    out = [backward_project(G, s) for s in sinos]
    torch.cuda.synchronize()
    return [o.cpu() for o in out]


def test_two_streams_match_serial():
    from ccpi.reconstruction.projectors import backward_project
    G = _geom()
    gen = torch.Generator(device="cuda").manual_seed(11)
    sinos = [torch.rand(G.sino_shape, device="cuda", generator=gen) for _ in range(2)]
    ref = _serial(G, sinos)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for st in streams:
        st.wait_stream(torch.cuda.current_stream())
    outs = [[], []]
    for rep in range(4):
        for q in range(2):
            with torch.cuda.stream(streams[q]):
                outs[q].append(backward_project(G, sinos[q]))
    torch.cuda.synchronize()
    for q in range(2):
        for o in outs[q]:
            assert torch.equal(o.cpu(), ref[q])


def test_two_host_threads_match_serial():
    from ccpi.reconstruction.projectors import backward_project
    G = _geom()
    gen = torch.Generator(device="cuda").manual_seed(12)
    sinos = [torch.rand(G.sino_shape, device="cuda", generator=gen) for _ in range(2)]
    ref = _serial(G, sinos)
    outs = [[], []]
    errors = []
    barrier = threading.Barrier(2)

    def work(q):
        try:
            st = torch.cuda.Stream()
            st.wait_stream(torch.cuda.default_stream())
            barrier.wait(timeout=60)
            with torch.cuda.stream(st):
                for rep in range(4):
                    outs[q].append(backward_project(G, sinos[q]))
            st.synchronize()
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(q,)) for q in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    torch.cuda.synchronize()
    for q in range(2):
        assert len(outs[q]) == 4
        for o in outs[q]:
            assert torch.equal(o.cpu(), ref[q])
