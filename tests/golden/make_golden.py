"""Write the golden vectors under tests/golden/ from the reference itself.

Runs in the build container only: it needs oracle/_ref/libref.so, built by
``make -C oracle ref`` from the reference's own Src/FADDEEVA.c, Src/SVD.c,
Src/ALLOCATION.c and Src/LOG_ERROR.c, and the reference data file
niris_example/niriswavelength.fits.  Outputs are data only (inputs and the
reference's outputs):

  faddeeva_ref.npz   (nu, y) grids per nfigures 4..8 -> H, L of Faddeeva()
  svd_ref.npz        9x9 matrices -> U (in place), W, V of svd_dbl(), and the
                     chain U_k = left factor of svd(U_{k-1}) that trial k of an
                     LM iteration solves with
  niris_wavelength.npy  the 34-sample NIRIS grid (FITS BITPIX -64, big-endian)
  niriswavelength.fits  that file itself (a real CFITSIO-written FITS image,
                     read by the host program's FITS reader in the tests)
  control/*_input    the reference's example control files (data: keyword =
                     value text), parsed by the host program's control-file
                     parser and by meinv.control in the tests

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO / "oracle"))
sys.path.insert(0, str(REPO / "milne-eddington_amd"))

import oracle as O  # noqa: E402
from meinv.workloads import get  # noqa: E402

REF_FITS = Path("/root/reference/niris_example/niriswavelength.fits")


def faddeeva_grid(rng):
    """Points in every region of the dispatcher: asymptotic rings, Hui, 916,
    both special lines, and the region boundaries for nfigures 6."""
    pts = []
    pts.append(np.stack([rng.uniform(-45, 45, 600), rng.uniform(0, 1.5, 600)], 1))
    pts.append(np.stack([rng.uniform(-12, 12, 600), rng.uniform(0, 3.0, 600)], 1))
    pts.append(np.stack([rng.uniform(-1500, 1500, 100), rng.uniform(0, 2, 100)], 1))
    r = np.sqrt(np.array([110.0, 180.0, 1.6e3, 1.451e6]))
    ring = []
    for rr in r:
        for d in (-1e-9, 0.0, 1e-9):
            ring.append([rr + d, 0.3])
            ring.append([-(rr + d), 0.7])
    pts.append(np.array(ring))
    special = [[0.0, 0.0], [0.0, 0.5], [0.0, 1.5], [0.0, 30.0], [1e-3, 0.0], [0.15, 0.0], [0.3, 0.0],
               [-2.5, 0.0], [7.0, 0.0], [1.0, 1.0], [1.0, 0.999999], [-3.0, 1e-12], [2.0, 1e-6]]
    pts.append(np.array(special))
    return np.concatenate(pts, 0)


def read_fits_1d(path: Path) -> np.ndarray:
    raw = path.read_bytes()
    cards = {}
    off = 0
    while True:
        block = raw[off:off + 2880].decode("ascii")
        off += 2880
        done = False
        for i in range(0, 2880, 80):
            card = block[i:i + 80]
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                cards[key] = card[10:].split("/")[0].strip()
        if done:
            break
    assert int(cards["BITPIX"]) == -64 and int(cards["NAXIS"]) == 1
    n = int(cards["NAXIS1"])
    return np.frombuffer(raw[off:off + 8 * n], dtype=">f8").astype(np.float64)


def normal_matrices(n, rng):
    """Realistic LM normal matrices J^T W J + gradients from the oracle's
    Jacobians on the 60-sample NIRIS-like grid."""
    w = get("niris64")
    cfg = w.config()
    orc = O.Oracle(cfg, w.wavelengths())
    out = []
    for _ in range(n):
        lo, hi = w.ranges()[:, 0], w.ranges()[:, 1]
        p = lo + (hi - lo) * rng.uniform(size=9)
        if rng.uniform() < 0.3:
            p[0] = rng.uniform(0, 20)  # weak field: nearly singular theta/phi rows
        syn, J = orc.synth(p, jac=True)
        J = J[0].reshape(4, 9, -1).transpose(0, 2, 1).reshape(-1, 9)
        iv = 1.0 / (0.01 * p[7]) ** 2
        out.append(iv * (J.T @ J))
    return np.array(out)


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref.so missing: run make -C oracle ref")
    rng = np.random.default_rng(20261200)

    grids, Hs, Ls, figs = [], [], [], []
    for nf in (4, 5, 6, 7, 8):
        g = faddeeva_grid(rng)
        H, L = O.ref_faddeeva(g[:, 0], g[:, 1], nf)
        grids.append(g)
        Hs.append(H)
        Ls.append(L)
        figs.append(np.full(len(g), nf, dtype=np.int8))
    np.savez_compressed(HERE / "faddeeva_ref.npz", nu=np.concatenate([g[:, 0] for g in grids]),
                        y=np.concatenate([g[:, 1] for g in grids]), nfig=np.concatenate(figs),
                        H=np.concatenate(Hs), L=np.concatenate(Ls))

    mats = [rng.normal(size=(9, 9)) for _ in range(8)]
    nm = normal_matrices(12, rng)
    mats += list(nm)
    A = rng.normal(size=(9, 9))
    A = A @ A.T
    A[:, 4] = 0.0
    A[4, :] = 0.0
    mats.append(A)
    mats.append(np.diag(np.arange(1.0, 10.0)))
    mats = np.array(mats)
    U = np.empty_like(mats)
    W = np.empty((len(mats), 9))
    V = np.empty_like(mats)
    for i, m in enumerate(mats):
        U[i], W[i], V[i] = O.ref_svd(m)
    # H1 chains from four normal matrices: chain[c, k] = U_k (U_0 = left factor of H)
    chain = np.empty((4, 6, 9, 9))
    for c in range(4):
        m = nm[c]
        for k in range(6):
            m, _, _ = O.ref_svd(m)
            chain[c, k] = m
    g = rng.normal(size=(len(mats), 9))
    x = np.array([O.ref_svbksb(U[i], V[i], W[i], g[i]) for i in range(len(mats))])
    np.savez_compressed(HERE / "svd_ref.npz", A=mats, U=U, W=W, V=V, chain_src=nm[:4], chain=chain,
                        b=g, x=x)

    if REF_FITS.exists():
        np.save(HERE / "niris_wavelength.npy", read_fits_1d(REF_FITS))
        shutil.copyfile(REF_FITS, HERE / "niriswavelength.fits")
    (HERE / "control").mkdir(exist_ok=True)
    ref = REF_FITS.parents[1]
    for src, dst in ((ref / "niris_example" / "input", "niris_example_input"),
                     (ref / "examples" / "niris_example" / "input", "niris_input"),
                     (ref / "examples" / "hinode_example" / "input", "hinode_input"),
                     (ref / "examples" / "hmi_example" / "input", "hmi_input")):
        if src.exists():
            shutil.copyfile(src, HERE / "control" / dst)
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".npy")))


if __name__ == "__main__":
    main()
