"""Pins for the oracle's major-cycle functions (current/tMajorACC/tMajorACC.cc):
findPeak, subtractPsf, deconvolve (Hogbom), fftExec's shifted transform,
fftFix and the visibility subtraction.  tMajorACC needs FFTW and an OpenACC
compiler, so no compiled reference backs these: the checks are closed forms,
invariants and brute force."""
from __future__ import annotations

import cmath

import numpy as np
import pytest

import oracle as O


def _img(rng, g):
    return (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))).astype(np.complex64)


# ---- findPeak (tMajorACC.cc:689-701) -----------------------------------------

def test_find_peak_matches_first_argmax_of_abs_real():
    rng = np.random.default_rng(1)
    for g in (1, 7, 64):
        im = _img(rng, g)
        v, p = O.find_peak(im)
        re = im.real.ravel()
        assert p == int(np.argmax(np.abs(re)))
        assert v == re[p]


def test_find_peak_ties_and_zeros():
    im = np.zeros((8, 8), np.complex64)
    assert O.find_peak(im) == (0.0, 0)
    im.flat[3] = 2.0
    im.flat[5] = -2.0
    assert O.find_peak(im) == (2.0, 3)        # first of equal magnitudes
    im.flat[3] = -2.0
    im.flat[5] = 2.0
    assert O.find_peak(im) == (-2.0, 3)
    im[:] = 0
    im.flat[9] = 1j * 5                       # imaginary parts are ignored
    im.flat[10] = 0.5
    assert O.find_peak(im) == (0.5, 10)


# ---- subtractPsf (tMajorACC.cc:723-750) --------------------------------------

def test_subtract_delta_psf_changes_one_pixel():
    g = 16
    psf = np.zeros((g, g), np.complex64)
    psf[g // 2, g // 2] = 1.0
    rng = np.random.default_rng(2)
    res = _img(rng, g)
    peak = 3 * g + 11
    out = O.subtract_psf(psf, res, peak, (g // 2) * g + g // 2, 2.5, 0.1)
    diff = (out - res).ravel()
    assert np.count_nonzero(diff) == 1
    amp = np.float32(np.float32(0.1) * np.float32(2.5))
    assert out.ravel()[peak].real == np.float32(res.ravel()[peak].real - amp)
    assert out.ravel()[peak].imag == res.ravel()[peak].imag


@pytest.mark.parametrize("rx,ry", [(0, 0), (5, 9), (15, 15), (8, 8)])
def test_subtract_ones_psf_covers_the_overlap_rectangle(rx, ry):
    g = 16
    px = py = 8
    psf = np.ones((g, g), np.complex64)
    res = np.zeros((g, g), np.complex64)
    out = O.subtract_psf(psf, res, ry * g + rx, py * g + px, 1.0, 0.5)
    changed = out.real != 0
    # the PSF (g x g, peak at (px, py)) placed with its peak on (rx, ry)
    ys = slice(max(0, ry - py), min(g, ry + g - py))
    xs = slice(max(0, rx - px), min(g, rx + g - px))
    want = np.zeros((g, g), bool)
    want[ys, xs] = True
    assert np.array_equal(changed, want)
    assert np.all(out.real[want] == -0.5)
    assert np.all(out.imag == 0)


def test_subtract_psf_aligned_when_peak_on_diagonal():
    g = 12
    rng = np.random.default_rng(3)
    psf = _img(rng, g)
    px = py = 5
    rx, ry = 7, 2
    out = O.subtract_psf(psf, np.zeros((g, g), np.complex64), ry * g + rx, py * g + px, 1.0, 1.0)
    for y in range(g):
        for x in range(g):
            sy, sx = y - (ry - py), x - (rx - px)
            want = -psf[sy, sx].real if (0 <= sy < g and 0 <= sx < g) else 0.0
            assert out[y, x].real == np.float32(want)


def test_subtract_psf_row_offset_off_diagonal():
    """With the PSF peak off the diagonal the reference offsets PSF rows by
    ry - px (not ry - py): row y of the residual takes PSF row y - ry + px."""
    g = 10
    psf = np.zeros((g, g), np.complex64)
    for sy in range(g):
        psf[sy, :] = sy + 1                      # row-coded PSF
    px, py = 2, 6
    rx, ry = 4, 5
    out = O.subtract_psf(psf, np.zeros((g, g), np.complex64), ry * g + rx, py * g + px, 1.0, 1.0)
    for y in range(max(0, ry - py), min(g, ry + g - py)):
        sy = y - ry + px
        row = out[y].real
        if 0 <= sy < g:
            assert np.all(row[max(0, rx - px):] == -(sy + 1))
        else:
            assert np.all(row == 0)


# ---- deconvolve (tMajorACC.cc:800-840) ---------------------------------------

def test_hogbom_delta_psf_closed_form():
    g = 32
    psf = np.zeros((g, g), np.complex64)
    psf[g // 2, g // 2] = 1.0
    dirty = np.zeros((g, g), np.complex64)
    A = 3.0
    dirty[7, 21] = A
    n = 40
    res, model, (pv, pp), (dv, dp) = O.deconvolve(dirty, psf, niters=n, gain=0.1)
    assert (pv, pp) == (1.0, (g // 2) * g + g // 2)
    assert (dv, dp) == (A, 7 * g + 21)
    assert res[7, 21].real == pytest.approx(A * 0.9 ** n, rel=1e-5)
    assert model[7, 21].real == pytest.approx(A * (1 - 0.9 ** n), rel=1e-5)
    assert np.count_nonzero(model) == 1


def test_hogbom_delta_psf_conserves_flux_per_pixel():
    g = 24
    rng = np.random.default_rng(4)
    psf = np.zeros((g, g), np.complex64)
    psf[g // 2, g // 2] = 1.0
    dirty = np.zeros((g, g), np.complex64)
    for _ in range(6):
        dirty[rng.integers(g), rng.integers(g)] = rng.uniform(-5, 5)
    res, model, _, _ = O.deconvolve(dirty, psf, niters=200, gain=0.1)
    assert np.allclose(res.real + model.real, dirty.real, rtol=0, atol=1e-5)
    # every iteration removes part of the current peak, so the largest residual shrinks
    assert np.abs(res.real).max() < np.abs(dirty.real).max()


# ---- fftExec (tMajorACC.cc:488-546), fftFix (:620-627) ---------------------

def test_shift_copy_is_half_roll_on_both_axes():
    a = np.arange(36).reshape(6, 6).astype(np.complex128)
    assert np.array_equal(O.fft_shift_copy(a), np.roll(np.roll(a, 3, 0), 3, 1))
    with pytest.raises(ValueError):
        O.fft_shift_copy(np.zeros((5, 5)))


def _dft2_bruteforce(x, sign):
    g = x.shape[0]
    out = np.zeros((g, g), np.complex128)
    for k1 in range(g):
        for k2 in range(g):
            acc = 0j
            for n1 in range(g):
                for n2 in range(g):
                    acc += x[n1, n2] * cmath.exp(sign * 2j * cmath.pi * (k1 * n1 + k2 * n2) / g)
            out[k1, k2] = acc
    return out


@pytest.mark.parametrize("forward", [True, False])
def test_fft_exec_vs_bruteforce_dft(forward):
    rng = np.random.default_rng(5)
    x = _img(rng, 8).astype(np.complex128)
    sh = np.roll(np.roll(x, 4, 0), 4, 1)
    want = _dft2_bruteforce(sh, -1 if forward else +1)
    want = np.roll(np.roll(want, 4, 0), 4, 1)
    assert np.allclose(O.fft_exec(x, forward), want, rtol=0, atol=1e-9)


def test_fft_exec_centre_delta_and_round_trip():
    g = 16
    x = np.zeros((g, g), np.complex64)
    x[g // 2, g // 2] = 1.0
    assert np.allclose(O.fft_exec(x, True), 1.0)
    rng = np.random.default_rng(6)
    y = _img(rng, g)
    back = O.fft_exec(O.fft_exec(y, True), False) / (g * g)
    assert np.allclose(back, y, atol=1e-6)


def test_fft_fix_and_vis_subtract():
    rng = np.random.default_rng(7)
    a = _img(rng, 9)
    f = O.fft_fix(a, 0.25)
    assert np.array_equal(f.real, (a.real * np.float32(0.25)).astype(np.float32))
    assert np.all(f.imag == 0)
    d = _img(rng, 5).ravel()
    m = _img(rng, 5).ravel()
    assert np.array_equal(O.vis_subtract(d, m), d - m)
