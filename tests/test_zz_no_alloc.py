"""The allocation-free environment-table entry point (pt_env_tables_noalloc) against the
allocating one (pt_unit_env_tables): same FP64 kernels, tables equal bit for bit, on two calls
in a row on one device (the second with a smaller map, so stale scratch contents from the
first call would show), and maps beyond the fixed scratch are refused.

Device buffers come from the HIP runtime that libpathed itself uses (hipMalloc through
ctypes, by the test, never by the entry point under test)."""
import ctypes as C

import numpy as np
import pytest

import pathed_capi as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The libamdhip64 instance libpathed is linked against (found in this process's maps)."""
    P.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert len(paths) >= 1, "libamdhip64 is not loaded"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def fill_nan(self):
        self.upload(np.full(self.nbytes // 4, np.nan, np.float32))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


@pytest.fixture(scope="module")
def hip():
    if P.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return _hip()


def _synthetic_env(w, h, seed):
    rng = np.random.default_rng(seed)
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = rng.lognormal(0.0, 1.0, (h, w, 3))
    rgba[h // 3, w // 5, :3] = 1e4  # a sun texel
    rgba[h - 1, :, :3] = 0.0        # a black row
    return rgba


def _noalloc_tables(hip, rgba, sin_weight):
    h, w = rgba.shape[:2]
    src = DeviceBuffer(hip, rgba.nbytes)
    rows = DeviceBuffer(hip, h * w * 8)
    marg = DeviceBuffer(hip, h * 8)
    pdf = DeviceBuffer(hip, h * w * 4)
    try:
        src.upload(rgba)
        for b in (rows, marg, pdf):
            b.fill_nan()
        P.check(P.lib().pt_env_tables_noalloc(src.ptr, w, h, int(sin_weight), rows.ptr, marg.ptr, pdf.ptr, None))
        assert hip.hipDeviceSynchronize() == 0
        return (rows.download(np.float32, (h, w, 2)), marg.download(np.float32, (h, 2)), pdf.download(np.float32, (h, w)))
    finally:
        for b in (src, rows, marg, pdf):
            b.free()


@pytest.mark.parametrize("sin_weight", [0, 1])
def test_noalloc_tables_equal_allocating_entry_point(hip, sin_weight):
    print(f"case: pt_env_tables_noalloc vs pt_unit_env_tables, sin weight {sin_weight}", flush=True)
    for w, h, seed in [(1024, 512, 1), (96, 40, 2)]:
        rgba = _synthetic_env(w, h, seed + 10 * sin_weight)
        rows, marg, pdf = _noalloc_tables(hip, rgba, sin_weight)
        rp, ra, mp, ma, tp = P.env_tables(rgba, sin_theta_weight=bool(sin_weight))
        assert np.array_equal(rows[..., 0], rp.astype(np.float32)), (w, h)
        assert np.array_equal(rows[..., 1].view(np.int32), ra), (w, h)
        assert np.array_equal(marg[:, 0], mp.astype(np.float32)), (w, h)
        assert np.array_equal(marg[:, 1].view(np.int32), ma), (w, h)
        assert np.array_equal(pdf, tp), (w, h)


def test_noalloc_refuses_maps_beyond_its_scratch(hip):
    print("case: pt_env_tables_noalloc size limits", flush=True)
    L = P.lib()
    dummy = DeviceBuffer(hip, 64)
    try:
        ptr = dummy.ptr
        # rejected on the host before anything is enqueued: no kernel sees these sizes
        assert L.pt_env_tables_noalloc(ptr, 4096, 1024, 0, ptr, ptr, ptr, None) == P.ERR_INVALID
        assert b"exceeds the fixed scratch" in L.pt_last_error()
        assert L.pt_env_tables_noalloc(ptr, 1, 8192, 0, ptr, ptr, ptr, None) == P.ERR_INVALID
        assert L.pt_env_tables_noalloc(None, 4, 4, 0, ptr, ptr, ptr, None) == P.ERR_INVALID
    finally:
        dummy.free()
