"""dpca_orthonormalize_noalloc (no reference counterpart: CholQR2 for callers
that cannot allocate, its intermediates in a static device array) against the
allocating dpca_orthonormalize, through the C ABI.

Device buffers come from the HIP runtime libdpca itself links (hipMalloc /
hipMemcpy through ctypes), so the test process holds one runtime.  Bar: the
same kernels in the same order, so the factor must match the allocating entry
point's to 1e-13 (FP64; bitwise in practice, which is printed)."""
import ctypes

import numpy as np
import pytest

import distpca_amd as D

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    D.lib()  # libdpca first: the runtime it links is the one we use
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc(hip, A, sync=True):
    m, k = A.shape
    kp = -(-k // 16) * 16
    Ap = np.zeros((m, kp))
    Ap[:, :k] = A
    nbytes = Ap.nbytes
    dA, dQ = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dA), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dQ), nbytes) == 0
    try:
        assert hip.hipMemcpy(dA, Ap.ctypes.data, nbytes, H2D) == 0
        shifted = ctypes.c_int(-1)
        rc = D.lib().dpca_orthonormalize_noalloc(dA, m, k, dQ,
                                                 ctypes.byref(shifted) if sync else None, 0, None)
        assert rc == 0, D.lib().dpca_last_error().decode()
        if not sync:
            assert hip.hipDeviceSynchronize() == 0
        Qp = np.empty_like(Ap)
        assert hip.hipMemcpy(Qp.ctypes.data, dQ, nbytes, D2H) == 0
    finally:
        hip.hipFree(dA)
        hip.hipFree(dQ)
    assert not Qp[:, k:].any()  # padding columns stay zero
    return Qp[:, :k], shifted.value


def _check(A, Q, shifted, sync=True):
    ref, ref_shifted = D.orthonormalize(A)
    diff = np.abs(Q - ref).max()
    print(f"[gpu] noalloc {A.shape}: max |diff| {diff:.2e}, bitwise {np.array_equal(Q, ref)}, "
          f"shifted {shifted} (allocating: {ref_shifted})", flush=True)
    assert diff <= 1e-13
    if sync:
        assert shifted == int(ref_shifted)
    assert np.abs(Q.T @ Q - np.eye(A.shape[1])).max() < 1e-12


def test_noalloc_matches_allocating_two_calls():
    hip = _hip()
    rng = np.random.default_rng(21)
    # two calls in a row, the second smaller (the array keeps the first's
    # partials; the second must not see them)
    A1 = rng.standard_normal((5000, 40)) @ np.diag(np.logspace(0, 3, 40))
    Q1, s1 = _noalloc(hip, A1)
    _check(A1, Q1, s1)
    A2 = rng.standard_normal((1200, 20))
    Q2, s2 = _noalloc(hip, A2)
    _check(A2, Q2, s2)


def test_noalloc_unsynchronised_and_shifted():
    hip = _hip()
    rng = np.random.default_rng(22)
    A = rng.standard_normal((3000, 64))
    Q, s = _noalloc(hip, A, sync=False)
    _check(A, Q, s, sync=False)
    # condition ~1e12: the unshifted Cholesky fails, the shifted fallback runs
    U = np.linalg.qr(rng.standard_normal((3000, 16)))[0]
    V = np.linalg.qr(rng.standard_normal((16, 16)))[0]
    B = (U * np.logspace(0, -12, 16)) @ V
    Q, s = _noalloc(hip, B)
    _check(B, Q, s)
    assert s == 1


def test_noalloc_rejects_beyond_limits():
    D.lib()
    dummy = ctypes.c_void_p(1)
    other = ctypes.c_void_p(2)
    assert D.lib().dpca_orthonormalize_noalloc(dummy, 1000, 65, other, None, 0, None) == -1
    assert "exceeds its limits" in D.lib().dpca_last_error().decode()
    assert D.lib().dpca_orthonormalize_noalloc(dummy, 1 << 30, 40, other, None, 0, None) == -1
    assert D.lib().dpca_orthonormalize_noalloc(dummy, 10, 4, dummy, None, 0, None) == -1
    assert "bad arguments" in D.lib().dpca_last_error().decode()
