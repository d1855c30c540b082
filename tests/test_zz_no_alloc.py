"""sdk_solve_nodes_noalloc (the subset solver with its buffers in the library's
fixed device array, for callers that cannot allocate) against sdk_solve_nodes
on the same inputs: the same kernels, so the same bits. Two calls in a row,
the second on a smaller image with fewer nodes, so that it runs over what the
first left in the array."""
import ctypes

import numpy as np
import pytest

from helpers import P, corpus_images, corpus_nodes, engine_module

pytestmark = pytest.mark.gpu

sa = engine_module()
NOALLOC_MAX_PIXELS = 1 << 20
ERR_CAPACITY = -7


@pytest.fixture(scope="module")
def dev():
    lib = sa.load_dev()
    assert lib.sdk_device_count() >= 1, "no GPU visible"
    return lib


def solve_host(dev, ref, dfm, dim, nodes, guesses, modes, lm=True, tol=1e-3):
    h, w = ref.shape
    n = len(nodes)
    io = np.zeros((n, 5), np.int32)
    fo = np.zeros((n, 7), np.float32)
    f32 = 1 if dfm.dtype == np.float32 else 0
    rc = dev.sdk_solve_nodes(P(np.ascontiguousarray(ref, np.float32)), P(np.ascontiguousarray(dfm)), f32, w, h, dim,
                             0, int(lm), ctypes.c_float(tol), n, P(np.array(nodes, np.int32).reshape(-1)),
                             P(np.ascontiguousarray(guesses, np.float32)), P(np.array(modes, np.int32)), P(io), P(fo))
    assert rc == 0
    return io, fo


class DevArrays:
    """Device copies of host arrays (the caller's inputs and outputs)."""

    def __init__(self, dev):
        self.dev, self.ptrs = dev, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.dev.sdk_malloc(max(a.nbytes, 1))
        assert p
        self.ptrs.append(p)
        assert self.dev.sdk_memcpy(p, P(a), a.nbytes, 0) == 0
        return p

    def get(self, p, like):
        out = np.zeros_like(like)
        assert self.dev.sdk_memcpy(P(out), p, out.nbytes, 1) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.dev.sdk_free(p)


def solve_noalloc(dev, ref, dfm, dim, nodes, guesses, modes, lm=True, tol=1e-3):
    h, w = ref.shape
    n = len(nodes)
    m = DevArrays(dev)
    try:
        f32 = 1 if dfm.dtype == np.float32 else 0
        io = np.full((n, 5), -1, np.int32)
        fo = np.full((n, 7), np.nan, np.float32)
        dio, dfo = m.put(io), m.put(fo)
        rc = dev.sdk_solve_nodes_noalloc(m.put(np.ascontiguousarray(ref, np.float32)), m.put(dfm), f32, w, h, dim, 0,
                                         int(lm), ctypes.c_float(tol), n, m.put(np.array(nodes, np.int32)),
                                         m.put(np.ascontiguousarray(guesses, np.float32)),
                                         m.put(np.array(modes, np.int32)), dio, dfo, None)
        assert rc == 0
        # the blocking copies on the null stream wait for the enqueued solve
        return m.get(dio, io), m.get(dfo, fo)
    finally:
        m.free()


def cases():
    ref, dfm = corpus_images(1)
    nodes = corpus_nodes()
    n = len(nodes)
    rng = np.random.default_rng(7)
    g = np.zeros((n, 6), np.float32)
    g[:, 0] = 2.1 + rng.uniform(-1.5, 1.5, n)
    g[:, 1] = 1.4 + rng.uniform(-1.5, 1.5, n)
    modes = np.array([1, 2, 0] * (n // 3 + 1))[:n]
    g[modes == 0, :] = 0.0
    big = (ref, dfm.astype(np.float32), nodes, g, modes)
    # smaller: a 320 x 288 crop, u8 deformed frame, the nodes inside it
    H, W = 288, 320
    keep = [i for i, (x, y) in enumerate(nodes) if 20 <= x < W - 20 and 20 <= y < H - 20]
    small = (np.ascontiguousarray(ref[:H, :W]), np.ascontiguousarray(np.clip(np.rint(dfm[:H, :W]), 0, 255)
                                                                      .astype(np.uint8)),
             [nodes[i] for i in keep], g[keep], modes[keep])
    return big, small


def test_noalloc_matches_solve_nodes_twice(dev):
    big, small = cases()
    assert len(small[2]) < len(big[2]) and small[0].size < big[0].size
    for ref, dfm, nodes, g, modes in (big, small):
        print(f"noalloc: {ref.shape[1]}x{ref.shape[0]} image, {len(nodes)} nodes, def {dfm.dtype}", flush=True)
        io_h, fo_h = solve_host(dev, ref, dfm, 27, nodes, g, modes)
        io_n, fo_n = solve_noalloc(dev, ref, dfm, 27, nodes, g, modes)
        assert io_h[:, 0].sum() > len(nodes) // 2
        assert np.array_equal(io_n, io_h)
        assert np.array_equal(fo_n.view(np.uint32), fo_h.view(np.uint32))


def test_noalloc_rejects_sizes_beyond_its_array(dev):
    """Beyond the array's capacity the call returns its own code and enqueues
    nothing (the check is on the host)."""
    w = h = 1025   # w * h > SDK_NOALLOC_MAX_PIXELS
    assert w * h > NOALLOC_MAX_PIXELS
    z = ctypes.c_void_p(1)   # never dereferenced: the size check comes first
    rc = dev.sdk_solve_nodes_noalloc(z, z, 1, w, h, 21, 0, 1, ctypes.c_float(1e-3), 1, z, z, z, z, z, None)
    assert rc == ERR_CAPACITY
    rc = dev.sdk_solve_nodes_noalloc(z, z, 1, 64, 64, 21, 0, 1, ctypes.c_float(1e-3), 16385, z, z, z, z, z, None)
    assert rc == ERR_CAPACITY
