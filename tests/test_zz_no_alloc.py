"""The allocation-free tendency (ssg_tendency_noalloc) against the existing
entry point ssg_tendency on the same inputs, through the C ABI.

Inputs are non-band-limited random fields, so the Nyquist rows and columns
of every symbol take part.  Device memory for the inputs and outputs comes
from torch (the caller's side); the entry point itself allocates nothing,
which is checked with the device's free-memory count.  Calls run in a row on
decreasing grids, so each one starts from the previous call's scratch
contents.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12   # FP64; the two paths differ only in how the twiddle table is formed


def _hip_free_bytes():
    # by SONAME, so the runtime already in the process is the one asked
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx))


def _reference(gpu, b, psi, quirk):
    ny, nx = b.shape
    d = gpu.make_domain(nx, ny, 1)
    f = gpu.allocate_fields(d)
    f.phis = psi
    f.bs = b
    t = gpu.compute_tendency(f, d, ab2_reference_aliasing=bool(quirk))
    d.close()
    return t


def test_noalloc_tendency_matches_ssg_tendency(gpu):
    import torch
    cases = [(1024, 1024, 0), (256, 256, 1), (128, 64, 0), (64, 128, 1)]   # largest first
    dev = {}
    for nx, ny, _ in cases:
        b, psi = _fields(nx, ny, nx + 7 * ny)
        dev[(nx, ny)] = (b, psi, torch.from_numpy(b).cuda(), torch.from_numpy(psi).cuda(),
                         torch.full((ny, nx), np.nan, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    first = True
    for rep in range(2):
        for nx, ny, quirk in cases:
            b, psi, tb, tp, to = dev[(nx, ny)]
            print(f"noalloc tendency {nx}x{ny} quirk={quirk} rep={rep}", flush=True)
            free_before = _hip_free_bytes()
            gpu.compute_tendency_noalloc(tb.data_ptr(), tp.data_ptr(), to.data_ptr(), nx, ny,
                                         ab2_reference_aliasing=bool(quirk), stream=stream)
            torch.cuda.synchronize()
            if not first:   # the first call may load the library's code object (and its array)
                assert _hip_free_bytes() == free_before, "the allocation-free call allocated device memory"
            first = False
            got = to.cpu().numpy()
            ref = _reference(gpu, b, psi, quirk)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < RTOL, (nx, ny, quirk, err)
            to.fill_(np.nan)
            torch.cuda.synchronize()


def test_noalloc_tendency_rejects(gpu):
    lib = gpu.load_library()
    desc = gpu._DomainDesc(2048, 1024, 1, 2 * np.pi, 2 * np.pi, 1.0, gpu.SSG_FP64, 0)
    one = ctypes.c_void_p(16)   # never dereferenced: the size check comes first
    rc = lib.ssg_tendency_noalloc(ctypes.byref(desc), one, one, one, 0, None)
    assert rc == gpu.SSG_ERR_UNSUPPORTED
    assert b"allocation-free path holds" in lib.ssg_last_error(None)
    desc = gpu._DomainDesc(64, 64, 1, 2 * np.pi, 2 * np.pi, 1.0, gpu.SSG_FP32_MIXED, 0)
    assert lib.ssg_tendency_noalloc(ctypes.byref(desc), one, one, one, 0, None) == gpu.SSG_ERR_UNSUPPORTED
    desc = gpu._DomainDesc(64, 64, 1, 2 * np.pi, 2 * np.pi, 1.0, gpu.SSG_FP64, 0)
    assert lib.ssg_tendency_noalloc(ctypes.byref(desc), None, one, one, 0, None) == gpu.SSG_ERR_ARG
    assert b"null" in lib.ssg_last_error(None)
