"""The oracle against the reference's own outputs (tests/golden, written by
tests/golden/make_golden.py from the reference's Faddeeva and SVD built from
its sources): bit-for-bit equality is the bar."""
import numpy as np

import oracle as O


def test_faddeeva_bit_exact_all_nfigures(golden):
    d = np.load(golden / "faddeeva_ref.npz")
    for nf in (4, 5, 6, 7, 8):
        m = d["nfig"] == nf
        H, L, _ = O.faddeeva(d["nu"][m], d["y"][m], nf)
        assert np.array_equal(H, d["H"][m]), nf
        assert np.array_equal(L, d["L"][m]), nf


def test_svd_bit_exact(golden):
    d = np.load(golden / "svd_ref.npz")
    for i in range(len(d["A"])):
        U, W, V, st = O.svd(d["A"][i])
        assert st == 0
        assert np.array_equal(U, d["U"][i]), i
        assert np.array_equal(W, d["W"][i]), i
        assert np.array_equal(V, d["V"][i]), i
        x = O.svbksb(d["U"][i], d["V"][i], d["W"][i], d["b"][i])
        assert np.array_equal(x, d["x"][i]), i


def test_h1_chain_bit_exact(golden):
    """Trial k of an LM iteration solves with U_{k-1}, the left factor the
    previous trial's SVD wrote over the matrix (H1): the chain must match."""
    d = np.load(golden / "svd_ref.npz")
    for c in range(len(d["chain_src"])):
        m = d["chain_src"][c]
        for k in range(d["chain"].shape[1]):
            m, _, _, _ = O.svd(m)
            assert np.array_equal(m, d["chain"][c, k]), (c, k)


def test_svd_reconstruction_and_numpy_singular_values(golden):
    d = np.load(golden / "svd_ref.npz")
    for A in d["A"]:
        U, W, V, _ = O.svd(A)
        np.testing.assert_allclose(U @ np.diag(W) @ V.T, A, atol=1e-12 * np.abs(A).max())
        np.testing.assert_allclose(np.sort(W), np.sort(np.linalg.svd(A, compute_uv=False)),
                                   atol=1e-12 * W.max())
        assert np.all(W >= 0)
