"""phase_infer_execute / phase_infer_execute_device (fft_cffi.cpp:2058-2174) against the
float restatement (oracle/cpu_ref.cpp ref_phase_linear, ref_phase_minimum).

Mode 0 computes phases in double: the float outputs agree to within 2 ulps of the
magnitude. Mode 1 chains log, two transforms and exp: relative L2 within 4 *
fp64.tolerance(N), plus the one-pole closed form (tests/test_oracle.py) on the GPU."""
import ctypes

import numpy as np
import pytest

from fftfree_amd import CffiFft
from oracle import fp64, ref

pytestmark = pytest.mark.gpu
fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.fixture(scope="module")
def lib():
    f = CffiFft()
    yield f.lib
    f.close()


def run(lib, N, hop, half, mode, mag):
    h = lib.phase_init(N, hop, int(half), mode, 0)
    assert h
    try:
        re, im = np.full_like(mag, np.nan), np.full_like(mag, np.nan)
        got = lib.phase_infer_execute(h, fp(mag), mag.shape[0], fp(re), fp(im))
        return got, re, im
    finally:
        lib.phase_free(h)


@pytest.mark.parametrize("case", [(1024, 256, True, 3000), (512, 100, True, 777), (256, 256, False, 50),
                                  (1000, 250, True, 20), (2048, 512, True, 100), (1024, 333, True, 64)])
def test_phase_linear_vs_oracle(lib, case):
    """Odd and even hops, block counts that leave the last 64-frame block partial, 257 to
    1025 bins, against the reference's double running sum."""
    N, H, half, F = case
    bins = N // 2 + 1 if half else N
    mag = np.random.default_rng(N).random((F, bins)).astype(np.float32) * 10
    got, re, im = run(lib, N, H, half, 0, mag)
    assert got == F
    rr, ri = ref.phase_infer(mag, N, H, half=half, mode=0)
    tol = 2 * np.finfo(np.float32).eps * np.maximum(mag, 1e-30)
    assert np.all(np.abs(re - rr) <= tol) and np.all(np.abs(im - ri) <= tol)


@pytest.mark.parametrize("case", [(1024, True, 300), (256, True, 64), (512, False, 40), (8192, True, 12),
                                  (1, True, 5), (2, True, 7), (1, False, 3), (1024, False, 33), (512, True, 61)])
def test_phase_minimum_vs_oracle(lib, case):
    """N = 512 and 1024 run the fused kernel (one wave per frame, both transforms and the
    fold in registers and LDS; frame counts that leave the last workgroup partly empty),
    the other N the log / C2C / fold / C2C / exp launches: both against the float
    restatement of the reference."""
    N, half, F = case
    bins = N // 2 + 1 if half else N
    mag = (np.random.default_rng(N + 1).random((F, bins)).astype(np.float32) + 0.05) * 3
    got, re, im = run(lib, N, 0, half, 1, mag)
    assert got == F
    rr, ri = ref.phase_infer(mag, N, 0, half=half, mode=1)
    assert fp64.rel_l2(re + 1j * im, rr + 1j * ri) <= 4 * fp64.tolerance(N)
    if N == 1:
        # one bin: log m, a 1-point transform each way, exp -> (m, 0); the reference always
        # sets the DC bin (fft_cffi.cpp:2113-2114)
        assert np.allclose(re, mag, rtol=1e-6) and not im.any()


def test_phase_minimum_one_pole_on_device(lib):
    import torch
    N, a = 1024, 0.6
    k = np.arange(N // 2 + 1)
    h = 1.0 / (1.0 - a * np.exp(-2j * np.pi * k / N))
    mag = np.abs(h)[None, :].astype(np.float32).repeat(8, axis=0)
    d = torch.from_numpy(mag).cuda()
    re, im = torch.empty_like(d), torch.empty_like(d)
    hd = lib.phase_init(N, 0, 1, 1, 0)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.phase_infer_execute_device(hd, vp(d), 8, vp(re), vp(im), None) == 8
    torch.cuda.synchronize()
    lib.phase_free(hd)
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    assert np.max(np.abs(got - h[None, :])) <= 1e-5 * np.max(np.abs(h))


def test_phase_refusals(lib):
    mag = np.ones((4, 513), np.float32)
    out = np.zeros_like(mag)
    assert not lib.phase_init(0, 0, 1, 0, 0)
    h = lib.phase_init(1024, 256, 1, 7, 0)  # unknown mode: init succeeds, execute returns 0
    assert lib.phase_infer_execute(h, fp(mag), 4, fp(out), fp(out)) == 0
    assert lib.phase_infer_execute(h, fp(mag), 0, fp(out), fp(out)) == 0
    lib.phase_free(h)
    h = lib.phase_init(1000, 250, 1, 1, 0)  # minimum phase needs a power-of-two plan
    m2 = np.ones((2, 501), np.float32)
    o2 = np.zeros_like(m2)
    assert lib.phase_infer_execute(h, fp(m2), 2, fp(o2), fp(o2)) == 0
    lib.phase_free(h)
    lib.phase_free(None)
