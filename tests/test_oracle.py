"""Pins for the oracle (CPU, no GPU).

Three kinds of evidence, chosen so a plausible mistake (dropped term, wrong
sign or index, transposed operand) fails at least one:
  * golden fixtures written from the compiled reference (tests/golden/),
  * closed forms and invariants computed independently here,
  * brute force on tiny inputs in pure Python.
"""
from __future__ import annotations

import hashlib
import math
import os

import numpy as np
import pytest

import oracle as O
from helpers import golden, rel_err, small_problem


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---- a1 LCG --------------------------------------------------------------------

def test_lcg_closed_form_and_reference_draws():
    meta, arrs = golden("attic")
    lcg = O.LCG()
    mine = [lcg.next() for _ in range(64)]
    # closed form: 64-bit state, (uint32)(state / 65536) % (2^31 - 1)
    state, expect = 1, []
    for _ in range(64):
        state = (state * 1103515245 + 12345) % (1 << 64)
        expect.append(((state >> 16) & 0xFFFFFFFF) % 2147483647)
    assert mine == expect
    assert mine == arrs["lcg_draws"].tolist()
    assert mine[:6] == [16838, 1507104382, 373696386, 1701364844, 1461254476, 1173460475]


# ---- a2/a4 attic set-up ------------------------------------------------------------

@pytest.fixture(scope="module")
def attic():
    meta, arrs = golden("attic")
    u, v, w = O.attic_uvw(160000, 2000)
    freq = O.attic_freq(1)
    sup, wcs, C = O.attic_initC(freq[0], 5.0, 33, 2000)
    iu, iv, co = O.attic_initCOffset(u, v, w, freq, 5.0, wcs, 33, 4096, sup)
    return dict(meta=meta, arrs=arrs, u=u, v=v, w=w, freq=freq, sup=sup, wcs=wcs, C=C,
                iu=iu, iv=iv, co=co)


def test_attic_scalars(attic):
    f0 = 1.4e9 / 2.998e8
    assert attic["freq"][0] == f0
    # printed by the reference: Support = 64, W cellsize = 566.034, 268 MB
    assert attic["sup"] == int(1.5 * math.sqrt(2000 * 5.0 * f0) / 5.0) == 64
    assert attic["wcs"] == pytest.approx(566.034, abs=5e-4)
    assert attic["wcs"] == attic["meta"]["wCellSize"]
    assert attic["C"].nbytes // (1024 * 1024) == 268
    assert attic["C"].size == 129 * 129 * 8 * 8 * 33


def test_attic_uvw_vs_reference(attic):
    m, a = attic["meta"], attic["arrs"]
    for k in ("u", "v", "w"):
        assert sha(attic[k]) == m["sha256"][k]
    assert attic["u"].min() >= -1000 and attic["u"].max() <= 1000


def test_attic_C_closed_form_and_reference(attic):
    C = attic["C"]
    assert sha(C) == attic["meta"]["sha256"]["C"]
    assert not C.imag.any()
    # closed form before normalisation, recomputed here; the normalisation
    # is one common float factor
    f0, wcs, s, c = attic["freq"][0], attic["wcs"], 129, 64
    raw = []
    idx = []
    rng = np.random.default_rng(0)
    for _ in range(200):
        k, osj, osi, j, i = (int(rng.integers(0, n)) for n in (33, 8, 8, s, s))
        w = float(k - 33 // 2)
        r2 = (j - c + osj / 8) ** 2 + (i - c + osi / 8) ** 2
        val = math.exp(-r2) if w == 0 else math.cos(r2 / (w * math.sqrt(abs(w) * wcs * f0) / 5.0))
        raw.append(np.float32(val))
        idx.append(i + s * (j + s * (osi + 8 * (osj + 8 * k))))
    raw = np.array(raw, np.float32)
    # the w = 0 plane's centre tap is exp(0) = 1, so it holds the common
    # normalisation factor exactly; every tap is float(raw) * that factor
    centre = 64 + s * (64 + s * (0 + 8 * (0 + 8 * 16)))
    scale = C.real[centre]
    assert np.array_equal(C.real[idx], raw * scale)
    # the factor is 2112 / (float running sum of |C|): that sum lies between
    # 2^24 (where adding |C| <= 1 stalls) and the exact sum
    sumC = np.float32(33 * 64) / scale
    assert 2.0 ** 23 < sumC < 2.0 ** 25


def test_attic_offsets_closed_form_and_reference(attic):
    m = attic["meta"]
    for k in ("iu", "iv"):
        assert sha(attic[k]) == m["sha256"][k]
    assert sha(attic["co"]) == m["sha256"]["cOffset"]
    # expected first samples (derived by hand from the closed form)
    assert attic["iu"][:3].tolist() == [1114, 2593, 1998]
    assert attic["iv"][:3].tolist() == [2424, 2385, 2638]
    assert attic["co"][:3].tolist() == [12647160, 17156871, 24595398]
    # closed form in Python floats for a sample of visibilities
    f0, wcs = attic["freq"][0], attic["wcs"]
    for i in range(0, 160000, 997):
        us = f0 * attic["u"][i] / 5.0
        vs = f0 * attic["v"][i] / 5.0
        iu, iv = math.floor(us), math.floor(vs)
        fu, fv = int(8 * (us - iu)), int(8 * (vs - iv))
        woff = 16 + int(f0 * attic["w"][i] / wcs)
        assert attic["iu"][i] == iu + 2048 and attic["iv"][i] == iv + 2048
        assert attic["co"][i] == 129 * 129 * (fu + 8 * (fv + 8 * woff))
        assert 0 <= fu < 8 and 0 <= fv < 8 and 8 <= woff <= 24


@pytest.mark.slow
def test_attic_grid_degrid_vs_reference(attic):
    """Full config-1 grid and degrid, bit-identical to the compiled reference."""
    data = np.ones(160000, np.complex64)
    g = O.grid(data, attic["iu"], attic["iv"], attic["co"], attic["C"], 4096, 129)
    assert sha(g) == attic["meta"]["sha256"]["grid"]
    out = O.degrid(g, attic["iu"], attic["iv"], attic["co"], attic["C"], 129)
    assert sha(out) == attic["meta"]["sha256"]["outdata"]
    # conservation: sum of the grid = sum over visibilities of the slice sums
    slice_sums = np.array([attic["C"][c:c + 129 * 129].astype(np.complex128).sum()
                           for c in attic["co"][:2000]])
    g2 = O.grid(data[:2000], attic["iu"][:2000], attic["iv"][:2000], attic["co"][:2000],
                attic["C"], 4096, 129)
    assert abs(g2.astype(np.complex128).sum() - slice_sums.sum()) <= 1e-4 * np.abs(slice_sums).sum()


# ---- a5/a6 current generation -----------------------------------------------------

@pytest.mark.parametrize("rt", [0, 1, 2, 3, 4])
def test_current_setup_vs_reference(rt):
    meta, arrs = golden(f"current_rt{rt}")
    p = O.current_params(rt)
    for k in ("gSize", "uvCellSize", "wSize", "support", "overSample", "wCellSize"):
        assert p[k] == meta[k], k
    u, v, w = O.current_uvw(p["baseline"])
    assert u.size == meta["nSamples"]
    for k, a in (("u", u), ("v", v), ("w", w)):
        assert sha(a) == meta["sha256"][k]
    sS, c0, C = O.current_initC(p["uvCellSize"], p["wSize"], p["support"], p["overSample"],
                                p["wCellSize"])
    assert sha(C) == meta["sha256"]["C"]
    assert np.array_equal(sS, arrs["sSize"]) and np.array_equal(c0, arrs["cOffset0"])
    # per-plane normalisation: sum |C| = overSample^2 on every plane (double sum of floats)
    os2 = p["overSample"] ** 2
    for k in range(0, p["wSize"], max(1, p["wSize"] // 7)):
        n = int(sS[k]) ** 2 * os2
        tot = np.abs(C[c0[k]:c0[k] + n].astype(np.complex128)).sum()
        assert tot == pytest.approx(os2, rel=1e-5)
    # widths odd, >= 7 except the single-plane cases, centre plane narrowest
    assert (sS % 2 == 1).all()
    wn = np.array([1.0e9 / 2.998e8])
    iu, iv, wp, co, npp = O.current_initCOffset(u, v, w, wn, p["uvCellSize"], p["wCellSize"],
                                                p["wSize"], p["gSize"], p["overSample"], sS, c0)
    for k, a in (("iu", iu), ("iv", iv), ("wPlane", wp), ("cOffset", co)):
        assert sha(a) == meta["sha256"][k]
    assert np.array_equal(npp, arrs["numPerPlane"])
    assert int((npp * sS.astype(np.int64) ** 2).sum()) == meta["nPixelsGridded"]


@pytest.mark.parametrize("rt", [1, 3])
def test_current_grid_vs_reference(rt):
    meta, arrs = golden(f"current_rt{rt}")
    p = O.current_params(rt)
    u, v, w = O.current_uvw(p["baseline"])
    sS, c0, C = O.current_initC(p["uvCellSize"], p["wSize"], p["support"], p["overSample"],
                                p["wCellSize"])
    iu, iv, wp, co, npp = O.current_initCOffset(u, v, w, np.array([1.0e9 / 2.998e8]),
                                                p["uvCellSize"], p["wCellSize"], p["wSize"],
                                                p["gSize"], p["overSample"], sS, c0)
    g = O.grid(np.ones(iu.size, np.complex64), iu, iv, co, C, p["gSize"], sS, wPlane=wp)
    assert sha(g) == meta["sha256"]["grid"]
    out = O.degrid(g, iu, iv, co, C, sS, wPlane=wp)
    assert sha(out) == meta["sha256"]["outdata"]


# ---- a7/a8 kernels: invariants and brute force ------------------------------------

def test_single_visibility_patch_equals_slice():
    iu, iv, co, data, C = small_problem(1, 1, 300, 129)
    data[:] = 1
    g = O.grid(data, iu, iv, co, C, 300, 129)
    u0, v0 = iu[0] - 64, iv[0]
    assert np.array_equal(g[v0:v0 + 129, u0:u0 + 129], C[co[0]:co[0] + 129 * 129].reshape(129, 129))
    g[v0:v0 + 129, u0:u0 + 129] = 0
    assert not g.any()


def test_brute_force_tiny_complex():
    """Pure-Python loops with complex arithmetic on a tiny problem."""
    S, gSize = 5, 16
    iu, iv, co, data, C = small_problem(2, 7, gSize, S, os_=2, nPlanes=2)
    ref = [[0j] * gSize for _ in range(gSize)]
    for d in range(7):
        for y in range(S):
            for x in range(S):
                ref[iv[d] + y][iu[d] - 2 + x] += complex(data[d]) * complex(C[co[d] + y * S + x])
    ref = np.array(ref)
    g = O.grid(data, iu, iv, co, C, gSize, S)
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()
    # degrid brute force (no conjugation)
    G = (np.arange(gSize * gSize).reshape(gSize, gSize) * (1 - 0.5j) / 100).astype(np.complex64)
    out = O.degrid(G, iu, iv, co, C, S)
    for d in range(7):
        s = sum(complex(G[iv[d] + y, iu[d] - 2 + x]) * complex(C[co[d] + y * S + x])
                for y in range(S) for x in range(S))
        assert abs(out[d] - s) <= 1e-5 * max(abs(s), 1)


def test_linearity_and_adjointness():
    S, gSize = 33, 128
    iu, iv, co, data, C = small_problem(3, 300, gSize, S)
    a, b = np.complex64(1.5 - 0.25j), np.complex64(-0.75 + 2j)
    d2 = np.roll(data, 7).astype(np.complex128)
    g_lin = O.grid_f64(a * data.astype(np.complex128) + b * d2, iu, iv, co, C, gSize, S)
    g_sep = a * O.grid_f64(data, iu, iv, co, C, gSize, S) + b * O.grid_f64(d2, iu, iv, co, C, gSize, S)
    assert np.abs(g_lin - g_sep).max() <= 1e-9 * np.abs(g_sep).max()
    rng = np.random.default_rng(9)
    G = rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))
    lhs = np.sum(G * O.grid_f64(data, iu, iv, co, C, gSize, S))
    rhs = np.sum(O.degrid_f64(G, iu, iv, co, C, S) * data)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_degrid_of_ones_is_slice_sum():
    """The attic CUDA check (attic/tConvolveCuda/tConvolveCuda.cc:395-420),
    extended to the imaginary part with complex C."""
    S, gSize = 65, 200
    iu, iv, co, data, C = small_problem(4, 100, gSize, S)
    ones = np.ones((gSize, gSize), np.complex64)
    out = O.degrid(ones, iu, iv, co, C, S)
    ref = np.array([C[c:c + S * S].astype(np.complex128).sum() for c in co])
    assert rel_err(out, ref) <= 1e-5


def test_fp32_oracle_close_to_fp64():
    S, gSize = 129, 400
    iu, iv, co, data, C = small_problem(5, 2000, gSize, S, cluster=40)
    g32 = O.grid(data, iu, iv, co, C, gSize, S)
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, S)
    assert rel_err(g32, g64) <= 1e-5
    o32 = O.degrid(g32, iu, iv, co, C, S)
    o64 = O.degrid_f64(g64, iu, iv, co, C, S)
    assert rel_err(o32, o64) <= 1e-5


def test_planted_mistakes_are_caught():
    """Mistakes planted in the oracle's inputs on the CPU (transposed slice,
    conjugated C, dropped imaginary term) must fail the brute-force pins."""
    S, gSize = 5, 16
    iu, iv, co, data, C = small_problem(6, 3, gSize, S, os_=1, nPlanes=3)
    good = O.grid(data, iu, iv, co, C, gSize, S)
    Ct = C.copy()
    for c in set(co.tolist()):
        Ct[c:c + S * S] = C[c:c + S * S].reshape(S, S).T.reshape(-1)
    assert rel_err(O.grid(data, iu, iv, co, Ct, gSize, S), good) > 1e-3
    assert rel_err(O.grid(data, iu, iv, co, np.conj(C), gSize, S), good) > 1e-3
    assert rel_err(O.grid(data.real.astype(np.complex64), iu, iv, co, C, gSize, S), good) > 1e-3


def test_cpu_baseline_runner():
    S, gSize = 33, 256
    iu, iv, co, data, C = small_problem(7, 400, gSize, S)
    tg, td = O.cpu_baseline(2, 300, data, iu, iv, co, C, gSize, S)
    assert tg > 0 and td > 0


REF_TIMER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                         "ref_time_attic")


@pytest.mark.skipif(not os.access(REF_TIMER, os.X_OK), reason="reference timing driver not built")
def test_reference_timing_driver_computes_the_oracle(tmp_path):
    """The CPU-baseline driver runs the reference's own kernels: per thread, the
    degridded value of its first visibility from its own grid equals the
    oracle's, summed over the threads (printed as `check`)."""
    import json
    import subprocess
    S, gSize, n, threads, per = 17, 256, 600, 3, 150
    iu, iv, co, data, C = small_problem(41, n, gSize, S)
    for name, a in (("iu.i32", iu.astype(np.int32)), ("iv.i32", iv.astype(np.int32)),
                    ("cOffset.i32", co.astype(np.int32)), ("data.c64", data.astype(np.complex64)),
                    ("C.c64", C.astype(np.complex64))):
        a.tofile(tmp_path / name)
    (tmp_path / "params.txt").write_text(f"{gSize} {S // 2}\n")
    out = subprocess.run([REF_TIMER, str(tmp_path), str(threads), str(per)], capture_output=True,
                         text=True, check=True, timeout=120)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    want = 0.0
    for t in range(threads):
        sl = slice(t * per, (t + 1) * per)
        g = O.grid(data[sl], iu[sl], iv[sl], co[sl], C, gSize, S)
        want += abs(O.degrid(g, iu[sl], iv[sl], co[sl], C, S)[0])
    assert r["check"] == pytest.approx(want, rel=1e-5)
    assert r["grid_s"] > 0 and r["degrid_s"] > 0


@pytest.mark.parametrize("Sz", [129, 33])
def test_oracle_refuses_footprints_past_the_grid(Sz):
    """The oracle's C loops, like the reference's, do no bounds checks, so its
    wrappers refuse footprints that leave the grid (rows [iv, iv + s) past the
    last row, columns [iu - s/2, iu - s/2 + s) past either edge) or whose
    slice leaves C, instead of writing past the output array (a heap overrun
    that glibc reports later, at an unrelated allocation, as an abort).  The
    seeded test problems draw rows in [0, gSize - s] for the same reason."""
    gSize = 384
    iu, iv, co, data, C = small_problem(11 + Sz, 3000, gSize, Sz, complex_=True)
    assert iv.min() >= 0 and iv.max() <= gSize - Sz
    assert (iu - Sz // 2).min() >= 0 and (iu - Sz // 2 + Sz).max() <= gSize
    O.grid(data, iu, iv, co, C, gSize, Sz)        # in range: accepted
    for bad in ("row", "col", "slice"):
        iu2, iv2, co2 = iu.copy(), iv.copy(), co.copy()
        if bad == "row":
            iv2[7] = gSize - Sz + 1                  # last footprint row == gSize
        elif bad == "col":
            iu2[7] = gSize - Sz // 2                 # last column == gSize
        else:
            co2[7] = C.size - Sz * Sz + 1
        with pytest.raises(ValueError):
            O.grid(data, iu2, iv2, co2, C, gSize, Sz)
        with pytest.raises(ValueError):
            O.degrid(np.zeros((gSize, gSize), np.complex64), iu2, iv2, co2, C, Sz)


def test_grid_window_equals_full_grid_cells():
    """The windowed oracle gives exactly the full oracle grid's cells (FP32 and
    FP64), for windows on the grid's corners and edges, inside, and across
    the footprints of visibilities that only partly overlap them."""
    Sz, gSize = 65, 300
    iu, iv, co, data, C = small_problem(61, 1500, gSize, Sz)
    full = O.grid(data, iu, iv, co, C, gSize, Sz)
    full64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    for x0, y0, w, h in ((0, 0, 64, 32), (gSize - 50, gSize - 20, 50, 20), (0, 137, 300, 5),
                         (101, 0, 7, 300), (140, 151, 128, 64), (299, 299, 1, 1)):
        win = O.grid_window(data, iu, iv, co, C, gSize, Sz, x0, y0, w, h)
        assert np.array_equal(win, full[y0:y0 + h, x0:x0 + w]), (x0, y0)
        win64 = O.grid_window(data, iu, iv, co, C, gSize, Sz, x0, y0, w, h, f64=True)
        assert np.array_equal(win64, full64[y0:y0 + h, x0:x0 + w]), (x0, y0)
    # per-plane widths
    widths = np.array([33, 65, 17], np.int32)
    rng = np.random.default_rng(62)
    wp = rng.integers(0, 3, iu.size).astype(np.int32)
    sizes = widths.astype(np.int64) ** 2 * 64
    c0 = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    co2 = (c0[wp] + rng.integers(0, 64, iu.size) * widths[wp].astype(np.int64) ** 2).astype(np.int32)
    C2 = (rng.standard_normal(int(sizes.sum())) + 1j * rng.standard_normal(int(sizes.sum()))).astype(np.complex64)
    full2 = O.grid(data, iu, iv, co2, C2, gSize, widths, wPlane=wp)
    win2 = O.grid_window(data, iu, iv, co2, C2, gSize, widths, 90, 200, 100, 60, wPlane=wp)
    assert np.array_equal(win2, full2[200:260, 90:190])
    with pytest.raises(ValueError):
        O.grid_window(data, iu, iv, co, C, gSize, Sz, 250, 0, 64, 64)


@pytest.mark.parametrize("rt", [1])
def test_sort_by_plane_matches_reference_counting_sort(rt):
    """setSort(1): the product's stable sort by w-plane gives the same arrays as
    the reference's counting sort (current/tConvolveMPI/Benchmark.cc:275-298)
    on the full current-generation offsets; the oracle's sort is itself pinned
    by its defining properties (a permutation, planes contiguous and in order,
    input order kept inside each plane)."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "askap-benchmarks_amd"))
    from tcv.benchmark import sort_by_plane
    p = O.current_params(rt)
    u, v, w = O.current_uvw(p["baseline"])
    sS, c0, _ = O.current_initC(p["uvCellSize"], p["wSize"], p["support"], p["overSample"],
                                p["wCellSize"], table=False)
    wn = np.array([1.0e9 / 2.998e8])
    iu, iv, wp, co, npp = O.current_initCOffset(u, v, w, wn, p["uvCellSize"], p["wCellSize"],
                                                p["wSize"], p["gSize"], p["overSample"], sS, c0)
    ref = O.sort_by_plane(iu, iv, wp, co, npp)
    # the oracle's own pins
    assert np.all(np.diff(ref[2]) >= 0)
    assert np.array_equal(np.bincount(ref[2], minlength=p["wSize"]), npp)
    def key(a, b, c, d):   # iu, iv < 2^13, plane < 2^8, cOffset < 2^29
        return np.sort((d.astype(np.int64) << 34) | (c.astype(np.int64) << 26) |
                       (b.astype(np.int64) << 13) | a.astype(np.int64))
    assert max(iu.max(), iv.max()) < 1 << 13 and wp.max() < 1 << 8 and co.max() < 1 << 29
    assert np.array_equal(key(*ref), key(iu, iv, wp, co))
    for k in np.unique(wp)[:: max(1, p["wSize"] // 9)]:
        sel = wp == k
        start = int(npp[:k].sum())
        assert np.array_equal(ref[3][start:start + int(npp[k])], co[sel])
        assert np.array_equal(ref[0][start:start + int(npp[k])], iu[sel])
    got = sort_by_plane(iu, iv, wp, co)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
