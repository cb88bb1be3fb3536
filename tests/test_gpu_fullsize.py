"""The spectral workload at its full size (1 048 576 samples x 8 channels,
4096^2 grid, width 129, wSize 129: 1.4e11 grid points a pass) on the default
plan, checked through properties that hold at any size, so the oracle need
not run there:

* degridding a grid of ones gives every visibility the sum of its C slice;
* gridding conserves the total: sum(grid) = sum_v data_v * sum(slice_v);
* the two kernels are adjoint: <G, grid(d)> = <degrid(G), d> (bilinear);
* and, for index mapping at that size, 2000 sampled visibilities of the
  degridded random grid equal the oracle's values for them (the oracle is
  cheap per visibility).

Sums on the host are FP64.  Tolerances are relative to the matching sum of
magnitudes (sum_v |d_v| sum |slice_v| and the like): FP32 accumulation of
the 16 641 taps of a slice stays within 1e-5 of that scale."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from tcv import setup as S

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def spectral(ctx):
    prob = S.spectral_problem()
    job = tcv.GridJob(ctx, prob)
    s2 = int(prob.sSize[0]) ** 2
    C = prob.C.reshape(-1, s2)
    slice_sum = C.sum(axis=1, dtype=np.complex128)
    slice_abs = np.abs(C).sum(axis=1, dtype=np.float64)
    _, _, _, co = job.offsets_host()
    assert job.nBad == 0 and (co % s2 == 0).all()
    idx = co // s2
    yield prob, job, slice_sum[idx], slice_abs[idx]
    job.plan.close()


def test_fullsize_default_plan(spectral):
    prob, job, _, _ = spectral
    info = job.plan.info()
    assert (info["tile_w"], info["tile_h"], info["kernel"], info["degrid_kernel"]) == (128, 64, 3, 3)
    assert info["out_of_range"] == 0 and info["entries"] > prob.nVis


def test_fullsize_degrid_of_ones_is_slice_sum(ctx, spectral):
    prob, job, ssum, sabs = spectral
    ones = ctx.to_device(np.ones((prob.gSize, prob.gSize), np.complex64))
    job.plan.degrid(ones, job.C, job.out)
    out = job.out.download().astype(np.complex128)
    err = np.abs(out - ssum) / np.maximum(sabs, 1e-30)
    assert err.max() <= TOL, float(err.max())


def test_fullsize_grid_conserves_total(ctx, spectral):
    prob, job, ssum, sabs = spectral
    job.zero_grid()
    job.grid("tiled")
    total = job.grid_d.download().sum(dtype=np.complex128)
    d = prob.data.astype(np.complex128)
    want = (d * ssum).sum()
    scale = (np.abs(d) * sabs).sum()
    assert abs(total - want) <= TOL * scale, (total, want, scale)


def test_fullsize_grid_degrid_adjoint(ctx, spectral):
    prob, job, _, sabs = spectral
    rng = np.random.default_rng(41)
    g = prob.gSize
    G = (rng.standard_normal((g, g), dtype=np.float32) +
         1j * rng.standard_normal((g, g), dtype=np.float32)).astype(np.complex64)
    job.plan.degrid(ctx.to_device(G), job.C, job.out)
    out32 = job.out.download()
    out = out32.astype(np.complex128)
    # sampled visibilities against the oracle
    iu, iv, _, co = job.offsets_host()
    pick = rng.choice(prob.nVis, 2000, replace=False)
    s = int(prob.sSize[0])
    want = O.degrid(G, iu[pick], iv[pick], co[pick], prob.C, s)
    assert np.abs(out32[pick] - want).max() <= TOL * np.abs(want).max()
    job.zero_grid()
    job.grid("tiled")
    grid = job.grid_d.download()
    d = prob.data.astype(np.complex128)
    lhs = (G.astype(np.complex128) * grid).sum()
    rhs = (out * d).sum()
    scale = (np.abs(d) * sabs).sum() * float(np.abs(G).max())
    assert abs(lhs - rhs) <= TOL * scale, (lhs, rhs, scale)
