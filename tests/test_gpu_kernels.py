"""Unit checks of single HIP kernels of the path through their check entry points."""
import ctypes

import numpy as np
import pytest

import sfb_amd as S
from sfb_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


# K not a multiple of 4 or of the 16-deep k-step, several row and column tiles
@pytest.mark.parametrize("rows,K,N", [(16, 4, 16), (64, 16, 64), (70, 37, 23), (130, 256, 71), (300, 80, 200),
                                      (33, 81, 90), (65, 3, 17), (64, 1, 1)])
def test_gemm_f64_matches_numpy(rows, K, N):
    rng = np.random.default_rng(rows + K + N)
    A = rng.normal(size=(K, rows))
    B = rng.normal(size=(K, N))
    C = np.empty((N, rows))
    L.check(L.lib().sfb_gemm_f64(0, rows, K, N, L.ptr(A), L.ptr(B), L.ptr(C)))
    ref = B.T @ A
    assert np.max(np.abs(C - ref)) <= 1e-13 * np.max(np.abs(ref)) * K


@pytest.mark.parametrize("rows,K,N", [(16, 4, 16), (70, 37, 23), (130, 256, 71)])
def test_gemm_f32_matches_numpy(rows, K, N):
    rng = np.random.default_rng(rows * 3 + K + N)
    A = rng.normal(size=(K, rows))
    B = rng.normal(size=(K, N))
    C = np.empty((N, rows))
    L.check(L.lib().sfb_gemm_f32(0, rows, K, N, L.ptr(A), L.ptr(B), L.ptr(C)))
    ref = B.astype(np.float32).astype(np.float64).T @ A.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(C - ref)) <= 1e-6 * np.max(np.abs(ref))


# the mixed path's kernels: FP32 storage, 8-deep k-tiles (K and K1 off multiples of 8, column
# counts whose B tile leaves the last load round partial), the split-K merged form and Cin
@pytest.mark.parametrize("rows,K,K1,N,cin", [(16, 4, 0, 16, False), (70, 37, 0, 23, False), (130, 256, 0, 71, False),
                                             (65, 3, 0, 17, False), (64, 1, 0, 1, False), (33, 9, 0, 80, True),
                                             (300, 80, 0, 200, False), (70, 37, 20, 23, True),
                                             (130, 141, 71, 71, True), (64, 12, 5, 48, False)])
def test_gemm_f32_stored_matches_numpy(rows, K, K1, N, cin):
    rng = np.random.default_rng(rows * 7 + K + 3 * K1 + N)
    A = rng.normal(size=(K, rows))
    B = rng.normal(size=(K, N))
    Cin = rng.normal(size=(N, rows)) if cin else None
    C = np.empty((N, rows))
    L.check(L.lib().sfb_gemm_f32_stored(0, rows, K, K1, N, L.ptr(A), L.ptr(B), L.ptr(Cin) if cin else None,
                                        L.ptr(C)))
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    ref = f32(B).T @ f32(A) + (f32(Cin) if cin else 0.0)
    scale = np.abs(f32(B)).T @ np.abs(f32(A)) + (np.abs(f32(Cin)) if cin else 0.0)
    err = np.max(np.abs(C - ref) / scale)
    print(f"rows {rows} K {K} K1 {K1} N {N}: max |C - ref| / (|B|^T |A| + |Cin|) = {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("rows,K,N", [(16, 4, 16), (70, 37, 23), (130, 256, 71), (300, 80, 200), (64, 1, 1)])
@pytest.mark.parametrize("terms", [2, 3])
def test_gemm_bf16_split_matches_numpy(rows, K, N, terms):
    """Split-BF16 operands: the FP32-rounded operands' product to ~2^-24 (hi + mid + lo) or
    ~2^-16 (hi + lo) of sum |a||b|; also checks the bf16 MFMA's operand and result layout."""
    rng = np.random.default_rng(rows * 5 + K + N + terms)
    A = rng.normal(size=(K, rows))
    B = rng.normal(size=(K, N))
    C = np.empty((N, rows))
    L.check(L.lib().sfb_gemm_bf16(0, rows, K, N, L.ptr(A), L.ptr(B), L.ptr(C), terms))
    A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    ref = B32.T @ A32
    scale = np.abs(B32).T @ np.abs(A32)
    err = np.max(np.abs(C - ref) / scale)
    print(f"terms {terms}: max |C - ref| / (|B|^T |A|) = {err:.2e}", flush=True)
    assert err <= (4e-7 if terms == 3 else 4e-5)


def test_phi_transform_matches_direct_dft():
    nside, lmax = 8, 16
    am = S.AnlmModes.from_nl(2, lmax, 500.0, 1000.0, nside=nside)
    plan = S.Plan(am)
    npix = 12 * nside * nside
    maps = np.random.default_rng(3).normal(size=(2, npix))
    K = 2 * nside
    out = np.empty((2 * (lmax + 1), K, 4))
    L.check(L.lib().sfb_phi_transform(plan.handle, 2, L.ptr(maps), L.ptr(out)))
    for k in range(K):
        st, npr, z, phi0 = O.ring_info(nside, k + 1)
        sts = st if k == K - 1 else npix - npr - st
        j = np.arange(npr)
        for m in range(lmax + 1):
            e = np.exp(-1j * m * (phi0 + 2 * np.pi * j / npr))
            for q in range(2):
                FN = np.sum(maps[q, st:st + npr] * e)
                FS = 0 if k == K - 1 else np.sum(maps[q, sts:sts + npr] * e)
                E, Od = (FN, FN) if k == K - 1 else (FN + FS, FN - FS)
                got_e = out[2 * m, k, 2 * q] + 1j * out[2 * m, k, 2 * q + 1]
                got_o = out[2 * m + 1, k, 2 * q] + 1j * out[2 * m + 1, k, 2 * q + 1]
                assert abs(got_e - E) < 1e-12 * (1 + abs(E)), (k, m, q, got_e, E)
                assert abs(got_o - Od) < 1e-12 * (1 + abs(Od)), (k, m, q, got_o, Od)


@pytest.mark.parametrize("nside,lmax,nmap", [(64, 200, 37), (256, 400, 19)])
def test_phi_transform_large_rings(nside, lmax, nmap):
    """Every ring pair against numpy's FFT (F_m = e^{-i m phi_0} FFT(ring)[m mod npr]) at
    ring sizes where a polar block holds one to sixty-four sequence batches (frequency
    counts from 1 to 128) and a map count that leaves the last tile of each ring part-full."""
    am = S.AnlmModes.from_nl(2, lmax, 500.0, 1000.0, nside=nside)
    plan = S.Plan(am)
    npix = 12 * nside * nside
    maps = np.random.default_rng(nside).normal(size=(nmap, npix))
    K = 2 * nside
    out = np.empty((2 * (lmax + 1), K, 2 * nmap))
    L.check(L.lib().sfb_phi_transform(plan.handle, nmap, L.ptr(maps), L.ptr(out)))
    got = out[:, :, 0::2] + 1j * out[:, :, 1::2]  # [(m, parity), k, q]
    m = np.arange(lmax + 1)
    worst = 0.0
    for k in range(K):
        st, npr, z, phi0 = O.ring_info(nside, k + 1)
        eq = k == K - 1
        sts = st if eq else npix - npr - st
        # e^{-i m phi_0} with m phi_0 reduced mod 2 pi in integers (phi_0 = shift pi / npr;
        # m phi_0 in floating point errs by up to 1e-13 at m = 400 on small rings)
        shift = int(round(phi0 * npr / np.pi))
        ph = np.exp(-1j * np.pi * ((m * shift) % (2 * npr)) / npr)[:, None]
        FN = ph * np.fft.fft(maps[:, st:st + npr], axis=1)[:, m % npr].T
        FS = np.zeros_like(FN) if eq else ph * np.fft.fft(maps[:, sts:sts + npr], axis=1)[:, m % npr].T
        E, Od = (FN, FN) if eq else (FN + FS, FN - FS)
        scale = np.sqrt(npr)
        worst = max(worst, np.abs(got[0::2, k] - E).max() / scale, np.abs(got[1::2, k] - Od).max() / scale)
    print(f"nside {nside}: max |F - FFT| / sqrt(npr) = {worst:.2e}", flush=True)
    assert worst < 1e-13
