"""uft_mma_design_change_noalloc (the allocation-free design change) against
uft_mma_design_change through the C-ABI: two calls in a row, the second on a
smaller input, each checked bit for bit (a max and a copy are exact), plus
the argument checks.  The test allocates the device arrays for the
allocation-free call itself, through the same HIP runtime that libuft.so
links (the entry point may not allocate; the test may)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs, from libuft.so's own runtime."""

    def __init__(self):
        import uft
        uft.lib()  # loads /opt/rocm/lib/libamdhip64.so.7 as libuft.so's dependency
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        return p

    def put(self, p, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, p, n):
        out = np.zeros(n)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out


def _ref_change(nx, ny, nz, x, xold):
    """ch and the updated xold from the context entry point on the same values."""
    from uft.capi import cantilever
    ctx = cantilever(8, nx, ny, nz, Emax=1.0)
    ctx.set("X", x)
    ctx.set("Xold", xold)
    ch = ctx.design_change()
    out = ctx.get("Xold")
    ctx.close()
    return ch, out


def test_noalloc_matches_design_change():
    import uft
    L = uft.lib()
    H = Hip()
    rng = np.random.default_rng(17)
    side = C.c_void_p()
    assert H.L.hipStreamCreate(C.byref(side)) == 0  # default priority
    ch_dev = H.malloc(8)
    H.put(ch_dev, [-1.0])
    # first call: 16,000 values (63 block partials) on a created stream, largest change 5 in the
    # last block; second call: 250 values (1 block) on the null stream, largest change below 1,
    # so partials left over from the first call would show
    bufs = []
    for (nx, ny, nz), big, s in (((40, 20, 20), 5.0, side), ((10, 5, 5), None, None)):
        n = nx * ny * nz
        x = rng.uniform(0.01, 1.0, n)
        xold = np.clip(x + rng.uniform(-0.2, 0.2, n), 0.0, 1.0)
        if big is not None:
            xold[n - 3] = x[n - 3] + big
        print(f"[case] noalloc n={n} stream={'created' if s else 'null'}", flush=True)
        ch_ref, xold_ref = _ref_change(nx, ny, nz, x, xold)
        xd, xod = H.malloc(8 * n), H.malloc(8 * n)
        bufs += [xd, xod]
        H.put(xd, x)
        H.put(xod, xold)
        rc = L.uft_mma_design_change_noalloc(0, xd, xod, n, ch_dev, s)
        assert rc == 0, L.uft_last_error(None)
        assert H.L.hipStreamSynchronize(s) == 0
        ch = H.get(ch_dev, 1)[0]
        assert ch == ch_ref, (n, ch, ch_ref)
        assert ch == np.max(np.abs(x - xold))
        xo = H.get(xod, n)
        assert np.array_equal(xo, xold_ref)
        assert np.array_equal(xo, x)
    assert ch < 1.0
    for p in bufs + [ch_dev]:
        H.L.hipFree(p)
    H.L.hipStreamDestroy(side)


def test_noalloc_arguments():
    import uft
    L = uft.lib()
    H = Hip()
    p = H.malloc(8)
    H.put(p, [-1.0])
    # n = 0: ch = 0, no element touched
    assert L.uft_mma_design_change_noalloc(0, None, None, 0, p, None) == 0
    assert H.L.hipStreamSynchronize(None) == 0
    assert H.get(p, 1)[0] == 0.0
    # beyond the scratch array's capacity: rejected before any launch
    assert L.uft_mma_design_change_noalloc(0, p, p, (1 << 24) + 1, p, None) == 1
    assert b"outside [0, 16777216]" in L.uft_last_error(None)
    assert L.uft_mma_design_change_noalloc(0, p, p, -1, p, None) == 1
    assert L.uft_mma_design_change_noalloc(0, None, p, 10, p, None) == 1
    assert L.uft_mma_design_change_noalloc(0, p, p, 10, None, None) == 1
    assert L.uft_mma_design_change_noalloc(-1, p, p, 10, p, None) == 1
    assert L.uft_mma_design_change_noalloc(64, p, p, 10, p, None) == 1
    assert b"out of range" in L.uft_last_error(None)
    H.L.hipFree(p)
