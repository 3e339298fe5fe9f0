"""xatu_bse_matvec_noalloc: the FP64 product for callers that cannot allocate, checked
against xatu_bse_matvec (same kernels) and the oracle's matrix on two calls in a row, the
second smaller, sharing the GPU's scratch; sizes beyond the scratch are refused."""
import numpy as np
import pytest

from cases import MODELS, gpu_context, oracle_case, rel_err

pytestmark = pytest.mark.gpu


class DeviceArray:
    """Device memory from the HIP runtime the library itself links (libamdhip64 of the ROCm
    install): torch bundles a second HIP runtime that cannot share the process's GPU with it."""

    def __init__(self, nbytes):
        import ctypes
        from xatu_amd import capi
        capi.load()  # the library's HIP runtime is loaded by now; dlopen returns that one
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        import ctypes
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, a):
        import ctypes
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(a.nbytes), 2) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


def device_product(ctx, x):
    x = np.asfortranarray(x, dtype=np.complex128)  # column-major N x nvec
    y = np.full_like(x, np.nan, order="F")
    dx, dy = DeviceArray(x.nbytes), DeviceArray(y.nbytes)
    dx.upload(x)
    dy.upload(y)
    ctx.matvec_noalloc(dx.ptr.value, dy.ptr.value, x.shape[1])
    dy.download(y)
    dx.free()
    dy.free()
    return y


def test_noalloc_product_two_calls():
    print("[no-alloc] hBN 72x72 (N=5184), 37 vectors; then hBN 20x20 (N=400), 5 vectors", flush=True)
    rng = np.random.default_rng(41)
    big = oracle_case(MODELS / "hBN.model", 72, 1, 0, (1, 1, 10), exchange=True)
    small = oracle_case(MODELS / "hBN.model", 20, 1, 0, (1, 1, 10), exchange=True, Q=(0.02, 0, 0))
    ctx = gpu_context(big)
    ctx.build(0)
    x = rng.normal(size=(big.N, 37)) + 1j * rng.normal(size=(big.N, 37))
    y = device_product(ctx, x)
    y_ref = ctx.matvec(x, precision=0)
    # fewer column splits fit in the scratch at this N: same kernels, another summation order
    assert rel_err(y, y_ref) < 1e-13
    ctx.close()
    ctx = gpu_context(small)
    ctx.build(0)
    x = rng.normal(size=(small.N, 5)) + 1j * rng.normal(size=(small.N, 5))
    y = device_product(ctx, x)
    # the same splits fit: identical to the allocating product, bit for bit
    assert np.array_equal(y, ctx.matvec(x, precision=0))
    assert rel_err(y, small.bse() @ x) < 1e-12
    ctx.close()


def test_noalloc_refusals():
    from xatu_amd import capi
    ex = oracle_case(MODELS / "hBN.model", 12, 1, 0, (1, 1, 10))
    ctx = gpu_context(ex)
    with pytest.raises(capi.XatuError) as e:  # null vectors
        ctx.matvec_noalloc(0, 0, 1)
    assert e.value.code == capi.XATU_ERR_INVALID_ARGUMENT
    xd = DeviceArray(16 * ex.N)
    with pytest.raises(capi.XatuError) as e:  # no FP64 matrix resident
        ctx.matvec_noalloc(xd.ptr.value, xd.ptr.value, 1)
    assert e.value.code == capi.XATU_ERR_LOGIC
    # N = 257^2 = 66049 > 65536: beyond the scratch, refused before anything runs
    n = 257
    nk = n * n
    ctx.set_lattice(np.array([[2.5, 0, 0]]), np.zeros((2, 3)), [1, 1], nk)
    ctx.set_bands(1, 1, np.zeros((2, nk), order="F"), np.zeros((2, 2, nk), dtype=np.complex128, order="F"),
                  np.arange(nk, dtype=np.int32)[:, None])
    with pytest.raises(capi.XatuError) as e:
        ctx.matvec_noalloc(xd.ptr.value, xd.ptr.value, 1)
    assert e.value.code == capi.XATU_ERR_INVALID_ARGUMENT
    assert "65536" in str(e.value)
    xd.free()
    ctx.close()
