"""Generate the golden fixtures from the compiled reference.

Builds oracle/_ref/ (the reference's own Benchmark.cc compiled from
/root/reference with a checker driver, see oracle/Makefile), runs it, and
writes small fixtures into this directory:

  attic.json / attic.npz          config-1 workload of attic/tConvolveMPI
  current_rt{0..4}.json / .npz    current/tConvolveMPI runType 0..4

Each JSON holds the scalars and a SHA-256 of every full array the reference
produced (u, v, w, iu, iv, cOffset, wPlane, C, grid, outdata); each NPZ holds
the first samples and a seeded selection of grid pixels, so tests can pin
values without the reference tree.  Run from the repository root:

    python tests/golden/make_golden.py [scratch_dir]
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(REPO, "oracle")
NHEAD = 256


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pixel_sample(grid: np.ndarray, n: int, seed: int):
    """Seeded selection of nonzero pixels plus the largest-magnitude ones."""
    g = grid.reshape(-1)
    nz = np.flatnonzero(g)
    rng = np.random.default_rng(seed)
    pick = rng.choice(nz, size=min(n, nz.size), replace=False)
    top = np.argsort(-np.abs(g))[:16]
    idx = np.unique(np.concatenate([pick, top])).astype(np.int64)
    return idx, g[idx]


def run_attic(scratch: str):
    out = os.path.join(scratch, "attic")
    os.makedirs(out, exist_ok=True)
    subprocess.run([os.path.join(ORACLE, "_ref", "ref_attic"), out], check=True)
    ld = lambda n, t: np.fromfile(os.path.join(out, n), dtype=t)
    arrays = {
        "u": ld("u.f64", np.float64), "v": ld("v.f64", np.float64),
        "w": ld("w.f64", np.float64), "iu": ld("iu.i32", np.int32),
        "iv": ld("iv.i32", np.int32), "cOffset": ld("cOffset.i32", np.int32),
        "C": ld("C.c64", np.complex64), "grid": ld("grid.c64", np.complex64),
        "outdata": ld("outdata.c64", np.complex64),
    }
    sc = ld("scalars.f64", np.float64)
    meta = {
        "source": "attic/tConvolveMPI/Benchmark.cc (compiled, config-1 constants)",
        "nSamples": 160000, "nChan": 1, "gSize": 4096, "cellSize": 5.0,
        "baseline": 2000, "wSize": 33, "overSample": 8,
        "support": int(sc[0]), "wCellSize": float(sc[1]),
        "sha256": {k: sha(v) for k, v in arrays.items()},
        "grid_sum": [float(arrays["grid"].real.astype(np.float64).sum()),
                     float(arrays["grid"].imag.astype(np.float64).sum())],
        "grid_maxabs": float(np.abs(arrays["grid"]).max()),
        "outdata_maxabs": float(np.abs(arrays["outdata"]).max()),
    }
    pidx, pval = pixel_sample(arrays["grid"], 2048, 1)
    npz = {
        "lcg_draws": ld("lcg_draws.i32", np.int32),
        "u": arrays["u"][:NHEAD], "v": arrays["v"][:NHEAD], "w": arrays["w"][:NHEAD],
        "iu": arrays["iu"][:NHEAD], "iv": arrays["iv"][:NHEAD],
        "cOffset": arrays["cOffset"][:NHEAD],
        "C_head": arrays["C"][:NHEAD],
        "C_plane16_os00_row64": arrays["C"][(129 * 129 * 64 * 16) + 129 * 64:(129 * 129 * 64 * 16) + 129 * 65],
        "grid_pix_idx": pidx, "grid_pix_val": pval,
        "outdata_head": arrays["outdata"][:NHEAD],
    }
    with open(os.path.join(HERE, "attic.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "attic.npz"), **npz)
    shutil.rmtree(out)


def run_current(scratch: str, rt: int, grid: bool):
    out = os.path.join(scratch, f"cur{rt}")
    os.makedirs(out, exist_ok=True)
    cmd = [os.path.join(ORACLE, "_ref", "ref_current"), out, str(rt)]
    if grid:
        cmd.append("grid")
    subprocess.run(cmd, check=True)
    ld = lambda n, t: np.fromfile(os.path.join(out, n), dtype=t)
    arrays = {
        "u": ld("u.f64", np.float64), "v": ld("v.f64", np.float64),
        "w": ld("w.f64", np.float64), "iu": ld("iu.i32", np.int32),
        "iv": ld("iv.i32", np.int32), "wPlane": ld("wPlane.i32", np.int32),
        "cOffset": ld("cOffset.i32", np.int32), "sSize": ld("sSize.i32", np.int32),
        "cOffset0": ld("cOffset0.i32", np.int32), "C": ld("C.c64", np.complex64),
    }
    if grid:
        arrays["grid"] = ld("grid.c64", np.complex64)
        arrays["outdata"] = ld("outdata.c64", np.complex64)
    sc = ld("scalars.f64", np.float64)
    meta = {
        "source": f"current/tConvolveMPI/Benchmark.cc (compiled, runType {rt})",
        "runType": rt, "gSize": int(sc[0]), "uvCellSize": float(sc[1]),
        "wSize": int(sc[2]), "support": int(sc[3]), "overSample": int(sc[4]),
        "wCellSize": float(sc[5]), "baseline": float(sc[6]),
        "nPixelsGridded": int(sc[7]), "nSamples": int(arrays["iu"].size),
        "sha256": {k: sha(v) for k, v in arrays.items()},
    }
    npz = {
        "u": arrays["u"][:NHEAD], "v": arrays["v"][:NHEAD], "w": arrays["w"][:NHEAD],
        "iu": arrays["iu"][:NHEAD], "iv": arrays["iv"][:NHEAD],
        "wPlane": arrays["wPlane"][:NHEAD], "cOffset": arrays["cOffset"][:NHEAD],
        "sSize": arrays["sSize"], "cOffset0": arrays["cOffset0"],
        "numPerPlane": np.bincount(arrays["wPlane"], minlength=int(sc[2])).astype(np.int64),
        "C_head": arrays["C"][:NHEAD],
    }
    if grid:
        meta["grid_maxabs"] = float(np.abs(arrays["grid"]).max())
        meta["outdata_maxabs"] = float(np.abs(arrays["outdata"]).max())
        pidx, pval = pixel_sample(arrays["grid"], 2048, 2 + rt)
        npz["grid_pix_idx"] = pidx
        npz["grid_pix_val"] = pval
        npz["outdata_head"] = arrays["outdata"][:NHEAD]
    with open(os.path.join(HERE, f"current_rt{rt}.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, f"current_rt{rt}.npz"), **npz)
    shutil.rmtree(out)


def main():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    scratch = sys.argv[1] if len(sys.argv) > 1 else tempfile.mkdtemp(prefix="tcv_golden_")
    run_attic(scratch)
    for rt in range(5):
        run_current(scratch, rt, grid=(rt != 4))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
