"""Major-cycle pieces on the device (current/tMajorACC/tMajorACC.cc) against
the oracle: the shifted FFT and fftFix (FP32 transform vs numpy's FP64 one,
max|d|/max|ref| <= 1e-5), findPeak and Hogbom CLEAN (bit-exact: same float
operations in the same order per pixel), the visibility subtraction
(bit-exact), PSF gridding (<= 1e-5) and one small end-to-end cycle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import rel_err, small_problem
from tcv import major as M

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _img(rng, g, real=False):
    a = rng.standard_normal((g, g))
    if not real:
        a = a + 1j * rng.standard_normal((g, g))
    return a.astype(np.complex64)


@pytest.mark.parametrize("g", [8, 256, 4096])
@pytest.mark.parametrize("forward", [True, False])
def test_fft_exec_vs_oracle(ctx, g, forward):
    rng = np.random.default_rng(g + forward)
    x = _img(rng, g)
    d = ctx.to_device(x)
    M.fftExec(ctx, d, g, forward)
    assert rel_err(d.download(), O.fft_exec(x, forward)) <= TOL


def test_fft_exec_fix_and_fix_alone(ctx):
    g = 512
    rng = np.random.default_rng(9)
    x = _img(rng, g)
    scale = 1.0 / 12345.0
    want = O.fft_fix(O.fft_exec(x, False).astype(np.complex64), scale)
    d = ctx.to_device(x)
    M.fftExecFix(ctx, d, g, False, scale)
    got = d.download()
    assert np.all(got.imag == 0)
    assert rel_err(got, want) <= TOL
    d2 = ctx.to_device(x)
    M.fftFix(ctx, d2, scale)
    assert np.array_equal(d2.download(), O.fft_fix(x, scale))


def test_fft_rejects_odd_size(ctx):
    d = ctx.zeros((7, 7), np.complex64)
    with pytest.raises(tcv.TcvError):
        M.fftExec(ctx, d, 7, True)


def test_find_peak_bitexact(ctx):
    rng = np.random.default_rng(10)
    for g in (1, 33, 1024):
        x = _img(rng, g)
        assert M.findPeak(ctx, ctx.to_device(x)) == O.find_peak(x)
    z = np.zeros((64, 64), np.complex64)
    assert M.findPeak(ctx, ctx.to_device(z)) == (0.0, 0)
    z.flat[100] = -3.0
    z.flat[2000] = 3.0
    z.flat[4000] = 2.0 + 9j
    assert M.findPeak(ctx, ctx.to_device(z)) == O.find_peak(z) == (-3.0, 100)


def _dirty_and_psf(seed, g, nSrc=30):
    """A dirty image and PSF with realistic structure: point sources convolved
    with a PSF that has sidelobes, peak at the centre."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:g, 0:g]
    c = g // 2
    r2 = ((xx - c) ** 2 + (yy - c) ** 2) / (g / 16) ** 2
    psf = (np.exp(-r2) + 0.2 * np.cos(0.9 * (xx - c)) * np.exp(-r2 / 9)).astype(np.float32)
    psf /= psf[c, c]
    sky = np.zeros((g, g), np.float32)
    for _ in range(nSrc):
        sky[rng.integers(g), rng.integers(g)] += rng.uniform(-1, 2)
    dirty = np.real(np.fft.ifft2(np.fft.fft2(sky) * np.fft.fft2(np.fft.ifftshift(psf))))
    return dirty.astype(np.complex64), psf.astype(np.complex64)


@pytest.mark.parametrize("g,niters", [(64, 50), (512, 200)])
def test_hogbom_bitexact(ctx, g, niters):
    dirty, psf = _dirty_and_psf(g, g)
    model0 = np.zeros_like(dirty)
    model0[3, 4] = 0.5 + 0.25j                       # the model accumulates
    r_ref, m_ref, (pv, pp), (dv, dp) = O.deconvolve(dirty, psf, model0, niters, 0.1)
    d_res, d_psf, d_model = ctx.to_device(dirty), ctx.to_device(psf), ctx.to_device(model0)
    (gpv, gpp), (gdv, gdp) = M.deconvolve(ctx, d_res, g, d_psf, g, d_model, niters, 0.1)
    assert (gpv, gpp) == (pv, pp)
    assert (gdv, gdp) == (dv, dp)
    assert np.array_equal(d_res.download(), r_ref)
    assert np.array_equal(d_model.download(), m_ref)


@pytest.mark.parametrize("g,pw", [(96, 40), (101, 37)])
def test_hogbom_small_psf_and_off_diagonal_peak(ctx, g, pw):
    """PSF narrower than the image, with its peak off the diagonal: the
    reference's row offset (ry - px) is reproduced, rows it would read outside
    the PSF are skipped, exactly as the oracle does.  Widths that are a
    multiple of four take the four-pixel path, others the one-pixel path."""
    rng = np.random.default_rng(12 + g)
    dirty = _img(rng, g, real=True)
    psf = (0.1 * _img(rng, pw, real=True)).astype(np.complex64)
    psf[9, 23] = 5.0                                   # peak at px=23, py=9
    r_ref, m_ref, pk_ref, dk_ref = O.deconvolve(dirty, psf, None, 60, 0.1)
    d_res, d_psf, d_model = ctx.to_device(dirty), ctx.to_device(psf), ctx.zeros((g, g), np.complex64)
    pk, dk = M.deconvolve(ctx, d_res, g, d_psf, pw, d_model, 60, 0.1)
    assert (pk, dk) == (pk_ref, dk_ref)
    assert np.array_equal(d_res.download(), r_ref)
    assert np.array_equal(d_model.download(), m_ref)


def test_hogbom_twice_reuses_workspace(ctx):
    dirty, psf = _dirty_and_psf(5, 128)
    for niters in (30, 7):
        r_ref, m_ref, _, _ = O.deconvolve(dirty, psf, None, niters, 0.1)
        d_res = ctx.to_device(dirty)
        d_model = ctx.zeros((128, 128), np.complex64)
        M.deconvolve(ctx, d_res, 128, ctx.to_device(psf), 128, d_model, niters, 0.1, peaks=False)
        assert np.array_equal(d_res.download(), r_ref)
        assert np.array_equal(d_model.download(), m_ref)


def test_vis_subtract_bitexact(ctx):
    rng = np.random.default_rng(13)
    for n in (1, 1000, 1 << 20):
        a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        d = ctx.to_device(a)
        M.subtractVis(ctx, d, ctx.to_device(b))
        assert np.array_equal(d.download(), O.vis_subtract(a, b))


def test_psf_gridding_vs_oracle(ctx):
    Sz, gSize = 65, 384
    iu, iv, co, _, C = small_problem(14, 3000, gSize, Sz)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    plan = tcv.Plan(ctx, vis)
    g = ctx.zeros((gSize, gSize), np.complex64)
    M.gridKernel(ctx, plan, None, ctx.to_device(C), g, isPSF=True)
    ones = np.ones(iu.size, np.complex64)
    assert rel_err(g.download(), O.grid(ones, iu, iv, co, C, gSize, Sz)) <= TOL


def test_major_cycle_small_vs_oracle(ctx):
    """Two cycles of tMajorACC's loop at a small size, every stage compared
    with the oracle chain started from the same visibilities."""
    mc = M.MajorCycle(ctx, nSamples=3000, wSize=9, nChan=1, cellSize=5.0, gSize=512,
                      baseline=200, nMajor=2, nMinor=20, nSources=25)
    s = 2 * mc.support + 1
    iu, iv, co = mc.vis.iu.download(), mc.vis.iv.download(), mc.vis.cOffset.download()
    C = mc.C.download()
    data = mc.data.download()
    # the visibilities are the degridded FFT of the true sky
    sky_grid = O.fft_exec(mc.sky, True).astype(np.complex64)
    assert rel_err(data, O.degrid(sky_grid, iu, iv, co, C, s)) <= 1e-4
    scale = np.float32(1.0 / float(mc.nVis))
    psf_img = O.fft_fix(O.fft_exec(O.grid(np.ones_like(data), iu, iv, co, C, 512, s), False)
                        .astype(np.complex64), scale)
    d = data.copy()
    for it in range(2):
        mc.cycle(it)
        if it == 0:
            assert rel_err(mc.psf.download(), psf_img) <= TOL
        img = O.fft_fix(O.fft_exec(O.grid(d, iu, iv, co, C, 512, s), False).astype(np.complex64),
                        scale)
        _, model, _, _ = O.deconvolve(img, psf_img, None, 20, 0.1)
        mg = O.fft_exec(model, True).astype(np.complex64)
        mvis = O.degrid(mg, iu, iv, co, C, s)
        d = O.vis_subtract(d, mvis)
        # same clean components, values within the FFT's rounding
        got_model_vis = mc.model_vis.download()
        assert rel_err(got_model_vis, mvis) <= 1e-4, it
        assert rel_err(mc.data.download(), d) <= 1e-4, it
