"""Init_Guess, INV_INIT's derived constants and INVERSION's run policy,
against values derived by hand from the reference's code
(Src/ME_SOLVER.c:445-676, Src/INV_INIT.c:144-177, Src/LM_FIT.c:701-724)."""
import numpy as np

import oracle as O
from meinv.workloads import get

C_LIGHT = 299792458.0


def test_line_centre_index_on_the_niris_grid(golden):
    lam = np.load(golden / "niris_wavelength.npy")
    assert lam.size == 34 and lam[13] == 15648.5
    cfg = get("niris64").config()
    info = O.Oracle(cfg, lam).info()
    # H4: lambda[13] equals lambda0, so the last index strictly below it is 12
    assert info["ic"] == 12
    ncut = 34 // 5
    assert info["ncut"] == ncut
    # window: half-width max(2, ic-ncut, nl-1-ncut-ic) = 15, clipped at 0
    assert (info["i0"], info["i1"]) == (0, 24)
    lo = max(-20.0, (lam[ncut] - 15648.5) * C_LIGHT / 1e3)
    hi = min(20.0, (lam[34 - 1 - ncut] - 15648.5) * C_LIGHT / 1e3)
    assert info["vlos"] == (lo, hi)


def test_derived_constants_centred_grids():
    for name, (ic, i0, i1) in {"niris64": (29, 11, 47), "hmi4k": (2, 0, 4), "hinode": (55, 21, 89)}.items():
        w = get(name)
        info = O.Oracle(w.config(), w.wavelengths()).info()
        assert (info["ic"], info["i0"], info["i1"]) == (ic, i0, i1), name


def _grid():
    w = get("niris64")
    return w, O.Oracle(w.config(), w.wavelengths())


def test_init_guess_unpolarised_absorption():
    w, orc = _grid()
    lam = w.wavelengths()
    lc = w.lambda0 + 0.1
    I = 1.0 - 0.5 * np.exp(-((lam - lc) / 0.2) ** 2)
    prof = np.stack([I, 0 * I, 0 * I, 0 * I])[None]
    r = orc.init_guess(prof)
    g = r["guess"]
    iL = int(np.argmin(I))
    assert r["fail"] == 0
    assert r["normp"] == 1.0
    np.testing.assert_allclose(r["norm0"], (0.01 * I.mean()) ** 2, rtol=1e-15)
    # no V signal: B_los = 0, B_pos = 100, theta = pi/2; H8: phi = pi/2
    assert g[0] == 100.0
    np.testing.assert_allclose(g[1], np.pi / 2, rtol=1e-15)
    np.testing.assert_allclose(g[2], np.pi / 2, rtol=1e-15)
    # v from the I minimum (five-point check holds on a smooth dip)
    np.testing.assert_allclose(g[3], (lam[iL] - w.lambda0) / w.lambda0 * C_LIGHT / 1e3, rtol=1e-12)
    # 15640 < lambda0 < 15660 table, S = Imax
    assert (g[4], g[5], g[6], g[8]) == (60.0, 0.5, 15.0, 0.7)
    assert g[7] == I.max()
    np.testing.assert_allclose(r["limit7"], [0.6 * I.max(), 1.4 * I.max()], rtol=1e-15)
    np.testing.assert_allclose(r["step7"], 0.1 * I.max(), rtol=1e-15)


def test_init_guess_longitudinal_field():
    w, orc = _grid()
    lam = w.wavelengths()
    info = orc.info()
    I = 1.0 - 0.4 * np.exp(-((lam - w.lambda0) / 0.25) ** 2)
    V = 0.08 * (lam - w.lambda0) / 0.25 * np.exp(-((lam - w.lambda0) / 0.25) ** 2)
    prof = np.stack([I, 0 * I, 0 * I, V])[None]
    g = orc.init_guess(prof)["guess"]
    win = slice(info["i0"], info["i1"] + 1)
    iVmin = info["i0"] + int(np.argmin(V[win]))
    iVmax = info["i0"] + int(np.argmax(V[win]))
    Imean = I.mean()
    LV = 0.5 * (100 * V[win].max() / Imean - 100 * V[win].min() / Imean)
    Blos = 20 + 100 * LV * 1.0
    if iVmin < iVmax:
        Blos = -Blos
    assert LV > 1.5  # above the NIRIS V threshold: v from the V extrema
    Bpos = 100.0
    B = np.hypot(Blos, Bpos)
    np.testing.assert_allclose(g[0], min(B, 3000), rtol=1e-14)
    np.testing.assert_allclose(g[1], np.arccos(Blos / B), rtol=1e-14)
    # LV > 1.5 (NIRIS threshold): v from the V extrema midpoint
    dl = 0.5 * (lam[iVmin] + lam[iVmax]) - w.lambda0
    np.testing.assert_allclose(g[3], dl / w.lambda0 * C_LIGHT / 1e3, rtol=1e-12, atol=1e-12)


def test_init_guess_reads_before_the_i_block_at_nl6():
    """H7: with nl = 6 the five-point check reads I[iL-2] = I[-1] for iL = 1,
    which in the reference's contiguous buffer is the previous pixel's V[5];
    before the first pixel it is taken as NaN (check fails)."""
    w = get("hmi4k")
    orc = O.Oracle(w.config(), w.wavelengths())
    lam = w.wavelengths()
    I = np.array([0.9, 0.5, 0.7, 0.8, 0.95, 1.0])
    pix = np.stack([I, np.zeros(6), np.zeros(6), np.zeros(6)])
    prev_hi = pix.copy()
    prev_hi[3, 5] = 2.0   # I[-1] = 2.0 > I[0]: the check passes
    prev_lo = pix.copy()
    prev_lo[3, 5] = 0.1   # I[-1] = 0.1 < I[0]: the check fails
    v_min = (lam[1] - w.lambda0) / w.lambda0 * C_LIGHT / 1e3
    assert np.isclose(orc.init_guess(np.stack([prev_hi, pix]), pix=1)["guess"][3], v_min, rtol=1e-12)
    g_lo = orc.init_guess(np.stack([prev_lo, pix]), pix=1)["guess"][3]
    g_first = orc.init_guess(pix[None], pix=0)["guess"][3]
    # centroid over ic +- 7 (the whole grid here)
    d = I.max() - I
    cen = (np.sum(lam * d) / np.sum(d) - w.lambda0) / w.lambda0 * C_LIGHT / 1e3
    np.testing.assert_allclose([g_lo, g_first], [cen, cen], rtol=1e-12)


def test_run_policy_counts_and_saves():
    crit = 1.0
    # num_run = 1, nothing converges: all 4 runs, best is the minimum
    n, best = O.run_policy(1, crit, [5.0, 4.0, 6.0, 3.0])
    assert (n, best) == (4, 3)
    # every run under the criterion: no break before irun 4 ... except the
    # irun >= 2*num_run test (best < 2*crit) at irun 2 -> 3 runs; H3: each
    # under-criterion run overwrites the saved one
    n, best = O.run_policy(1, crit, [0.5, 0.6, 0.7, 0.1])
    assert (n, best) == (3, 2)
    # first run just above the criterion but under twice it: stop at irun 2
    n, best = O.run_policy(1, crit, [1.5, 3.0, 4.0, 0.1])
    assert (n, best) == (3, 0)
    # num_run = 2: minimum of 5 runs, break at irun 4 under the criterion
    n, best = O.run_policy(2, crit, [0.1, 0.9, 0.5, 0.8, 0.7, 0.2])
    assert (n, best) == (5, 4)
    # num_run = 2 without convergence: stops at irun 4 once best < 2*crit
    n, best = O.run_policy(2, crit, [3.0, 1.7, 5.0, 5.0, 5.0, 5.0, 5.0, 5.0])
    assert (n, best) == (5, 1)
    n, best = O.run_policy(2, crit, [9.0] * 8)
    assert (n, best) == (8, 0)


def test_invalid_pixels_give_zero_records():
    w, orc = _grid()
    prof, _ = orc.gen_map(w.seed, w.ranges(), 4)
    prof = prof.astype(np.float64)
    prof[1, 0, 3] = np.nan
    prof[2, 0] *= 0.01      # mean(I) below Icriteria = 0.1
    out = orc.invert(prof)
    assert np.all(out["res"][1] == 0) and np.all(out["res"][2] == 0)
    assert np.all(out["res"][[0, 3], 9] > 0)
    assert list(out["status"][[1, 2]]) == [1, 1]


def test_init_guess_transverse_field_and_the_b_table():
    """B_pos = min(100 + LCoeffi * Lmax * factor, 2200) with Lmax the largest
    sqrt(Q^2 + U^2) over i0..i1 in percent of mean(I) (Src/ME_SOLVER.c:525-
    545); B = hypot(B_los, B_pos), theta = acos(B_los / B); off the NIRIS line
    the starting Doppler width, damping, eta0 and beta come from the B > 1100
    or B <= 1100 row of the table (:631-660).  H8: sqrt(Q^2+U^2) is compared
    with the threshold 5 in absolute units, so phi stays pi/2."""
    w = get("hmi4k")
    orc = O.Oracle(w.config(), w.wavelengths())
    lam = w.wavelengths()
    info = orc.info()
    win = slice(info["i0"], info["i1"] + 1)
    I = 1.0 - 0.5 * np.exp(-((lam - w.lambda0 - 0.01) / 0.08) ** 2)
    shape = np.exp(-((lam - w.lambda0) / 0.08) ** 2)
    for amp, row in ((0.06, (20.0, 0.5, 20.0, 0.3)), (0.01, (30.0, 0.5, 5.0, 0.15))):
        Q, U = 0.6 * amp * shape, -0.8 * amp * shape
        prof = np.stack([I, Q, U, 0 * I])[None]
        g = orc.init_guess(prof)["guess"]
        Imean = I.mean()
        Lmax = 100 * np.max(np.hypot(Q[win], U[win])) / Imean
        Bpos = min(100 + 250 * Lmax * 1.0, 2200)
        np.testing.assert_allclose(g[0], Bpos, rtol=1e-14)           # B_los = 0
        np.testing.assert_allclose(g[1], np.pi / 2, rtol=1e-15)
        np.testing.assert_allclose(g[2], np.pi / 2, rtol=1e-15)      # H8
        assert (g[4], g[5], g[6], g[8]) == row, (amp, g)
        assert (Bpos > 1100) == (row[0] == 20.0)
    # the cap: a very strong Q/U signal gives B_pos = 2200
    prof = np.stack([I, 0.5 * shape, 0 * I, 0 * I])[None]
    assert orc.init_guess(prof)["guess"][0] == 2200.0


def test_init_guess_weak_line_doubles_the_field_coefficients():
    """A line shallower than 10% (Imin > 0.9 Imax) doubles the V and L
    coefficients (factor 2, Src/ME_SOLVER.c:523-531), and v_los cannot come
    from the I minimum (the five-point check needs Imin < 0.9 Imax): below the
    V threshold it is the I-weighted centroid over ic +- 7."""
    w, orc = _grid()
    lam = w.wavelengths()
    info = orc.info()
    win = slice(info["i0"], info["i1"] + 1)
    I = 1.0 - 0.05 * np.exp(-((lam - w.lambda0 - 0.05) / 0.25) ** 2)
    V = 0.004 * (lam - w.lambda0) / 0.25 * np.exp(-((lam - w.lambda0) / 0.25) ** 2)
    Q = 0.002 * np.exp(-((lam - w.lambda0) / 0.25) ** 2)
    prof = np.stack([I, Q, 0 * I, V])[None]
    g = orc.init_guess(prof)["guess"]
    Imean = I.mean()
    LV = 0.5 * (100 * V[win].max() / Imean - 100 * V[win].min() / Imean)
    assert 0.1 < LV < 1.5
    iVmin = info["i0"] + int(np.argmin(V[win]))
    iVmax = info["i0"] + int(np.argmax(V[win]))
    Blos = 20 + 100 * LV * 2.0
    Blos = -Blos if iVmin < iVmax else Blos
    Bpos = min(100 + 250 * (100 * np.max(np.abs(Q[win])) / Imean) * 2.0, 2200)
    B = np.hypot(Blos, Bpos)
    np.testing.assert_allclose(g[0], B, rtol=1e-13)
    np.testing.assert_allclose(g[1], np.arccos(Blos / B), rtol=1e-13)
    a, b = max(info["ic"] - 7, 0), min(info["ic"] + 7, w.nl - 1)
    d = I.max() - I[a:b + 1]
    cen = np.sum(lam[a:b + 1] * d) / np.sum(d) - w.lambda0
    np.testing.assert_allclose(g[3], cen / w.lambda0 * C_LIGHT / 1e3, rtol=1e-11)
