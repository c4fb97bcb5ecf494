"""The oracle's LM loop control, Get_Regul, Random_Jump and learned policy,
pinned by hand-derived values on constructed sequences, by the Gauss-Newton
equations of an augmented residual vector, by recovering the Gaussian draws
from the jumped parameters and by statistical tests of the draws
(Src/LM_FIT.c:124-201, 354-396, 531-661, 665-742)."""
import numpy as np
import pytest
from scipy import stats

import oracle as O
from meinv.workloads import get


def _cfg(**over):
    return get("niris64").config(**over)


# ------------------------------------------------------------ LM loop control
def test_every_trial_rejected_stops_at_the_damping_bound():
    """Damp 1e-3 grows x5 per rejected trial; the run stops once it exceeds
    lambda_limits[1] = 1e3: 1e-3 * 5^9 = 1953 > 1e3 after the 9th trial
    (1e-3 * 5^8 = 391 is still under it).  No iteration completes."""
    r = O.lm_replay(_cfg(), chi0=10.0, trial_chi=[11.0])
    assert (r["trials"], r["iters"], r["chisq"]) == (9, 0, 10.0)
    # lambda_limits = (1e-5, 6e-3): 5e-3 after one rejection, 2.5e-2 after two
    r = O.lm_replay(_cfg(lambda_limits=(1e-5, 6e-3)), chi0=10.0, trial_chi=[11.0])
    assert (r["trials"], r["iters"]) == (2, 0)
    # lambda_reject = 10: 1e-3 * 10^7 = 1e4 > 1e3 after 7 trials (10^6 = 1e3 is not above it)
    r = O.lm_replay(_cfg(lambda_reject=10.0), chi0=10.0, trial_chi=[11.0])
    assert r["trials"] == 7
    # an equal chi2 is a rejection (the accept test is strict, Src/LM_FIT.c:591)
    r = O.lm_replay(_cfg(), chi0=10.0, trial_chi=[10.0])
    assert r["trials"] == 9


def test_damping_is_reset_after_every_iteration():
    """Each iteration may reject 8 trials and accept the 9th: with the damping
    reset to `damp` after every iteration (Src/LM_FIT.c:633) three such
    iterations run and a fourth with 9 rejections ends the run: 36 trials.
    Without the reset the second iteration would stop after one rejection."""
    seq = []
    chi = 100.0
    for _ in range(3):
        seq += [chi + 1.0] * 8
        chi *= 0.5
        seq += [chi]
    seq += [chi + 1.0] * 9
    r = O.lm_replay(_cfg(), chi0=100.0, trial_chi=seq)
    assert (r["trials"], r["iters"], r["chisq"]) == (36, 3, 12.5)


def test_convergence_is_relative_to_the_previous_chi2():
    """Converged when (chi2 - chi2_new) / chi2 < convg_criteria = 1e-3
    (Src/LM_FIT.c:57-58): the loop breaks without counting the iteration."""
    # gain 9.995e-4 of the old chi2 (but 1.0005e-3 of the new one): converged
    r = O.lm_replay(_cfg(), chi0=1.0, trial_chi=[1.0 - 9.995e-4])
    assert (r["trials"], r["iters"]) == (1, 0)
    assert r["chisq"] == 1.0 - 9.995e-4
    # gain exactly 1e-3 is not below it: one more iteration, then converged
    r = O.lm_replay(_cfg(), chi0=1000.0, trial_chi=[999.0, 998.5])
    assert (r["trials"], r["iters"], r["chisq"]) == (2, 1, 998.5)


def test_iteration_cap():
    """Every trial accepted with a large gain: num_iter iterations, one trial
    each; num_iter = 0 runs no iteration at all."""
    seq = [2.0 ** -k for k in range(1, 80)]
    r = O.lm_replay(_cfg(), chi0=1.0, trial_chi=seq)
    assert (r["trials"], r["iters"]) == (60, 60)
    assert r["chisq"] == 2.0 ** -60
    r = O.lm_replay(_cfg(num_iter=7), chi0=1.0, trial_chi=seq)
    assert (r["trials"], r["iters"], r["chisq"]) == (7, 7, 2.0 ** -7)
    r = O.lm_replay(_cfg(num_iter=0), chi0=1.0, trial_chi=seq)
    assert (r["trials"], r["iters"], r["chisq"]) == (0, 0, 1.0)


def test_h2_state_is_the_last_trial_chi2():
    """LM->chisq after a run is its last trial's chi2 (Src/LM_FIT.c:520),
    accepted or not."""
    r = O.lm_replay(_cfg(), chi0=10.0, trial_chi=[5.0, 5.0 - 1e-6])
    assert r["lm_chisq"] == 5.0 - 1e-6
    r = O.lm_replay(_cfg(), chi0=10.0, trial_chi=[5.0, 6.0])
    assert r["lm_chisq"] == 6.0 and r["trials"] == 10


def test_regularisation_ratio_updates():
    """ratio = 1 at the start of a run; if penalty(P) > 0.1 * (previous run's
    last trial chi2) it becomes 0.1 * chi2 / penalty (Src/LM_FIT.c:553-560).
    On an accepted trial, if penalty(P_trial) > 0.1 * chi2_trial it becomes
    0.1 * chi2_old / penalty(P_trial) (:596-598).  Each normal-matrix build
    uses the ratio current at that point."""
    # start: penalty 0.5 > 0.1 * 2 -> ratio 0.1 * 1 / 0.5 = 0.2; trial 1
    # accepted (0.5 < 1) with penalty 0.3 > 0.05 -> 0.1 * 1 / 0.3; trial 2
    # accepted with penalty 1e-3 < 0.1 * 0.3: unchanged; trial 3 converges
    r = O.lm_replay(_cfg(), chi0=1.0, pen0=0.5, lm_chisq=2.0, trial_chi=[0.5, 0.3, 0.2999999],
                    trial_pen=[0.3, 1e-3, 1e-3], regul=True)
    np.testing.assert_allclose(r["ratios"], [0.2, 0.1 / 0.3, 0.1 / 0.3], rtol=1e-15)
    assert r["iters"] == 2
    # previous run's chi2 large: the start test is false, ratio stays 1
    r = O.lm_replay(_cfg(), chi0=1.0, pen0=0.5, lm_chisq=10.0, trial_chi=[0.5, 0.4999999],
                    trial_pen=[1e-3, 1e-3], regul=True)
    np.testing.assert_allclose(r["ratios"], [1.0, 1.0])
    # a rejected trial never changes the ratio, whatever its penalty
    r = O.lm_replay(_cfg(), chi0=1.0, pen0=0.0, lm_chisq=np.inf, trial_chi=[2.0, 0.5, 0.4999999],
                    trial_pen=[5.0, 1e-3, 1e-3], regul=True)
    np.testing.assert_allclose(r["ratios"], [1.0, 1.0])
    # regularisation off: the ratio is never computed
    r = O.lm_replay(_cfg(), chi0=1.0, pen0=0.5, lm_chisq=2.0, trial_chi=[0.5, 0.4999999], trial_pen=[0.3, 0.3])
    np.testing.assert_allclose(r["ratios"], [1.0, 1.0])


# ------------------------------------------------------------------ Get_Regul
@pytest.mark.parametrize("fixed", [(), (0, 4), (1,), (2, 3, 5)])
def test_regularised_normal_equations_are_the_augmented_least_squares(fixed):
    """Get_Regul adds ratio * w_k to H's diagonal and subtracts ratio * w_k *
    (p_k - c_k) from the gradient of each regularised free parameter, at its
    packed index; that is the Gauss-Newton system of the residual vector
    extended by sqrt(ratio w_k) (p_k - c_k) with unit sensitivity.  The
    penalty is sum w_k (p_k - c_k)^2 over the regularised free parameters."""
    regul = (1, 1, 0, 1, 0, 1, 1, 0, 1)
    value = (100.0, 0.5 * np.pi, 0.3, 0.2, 30.0, 0.5, 10.0, 1.0, 0.4)
    weight = (1e-5, 1e-4, 3.0, 0.2, 0.05, 80.0, 2e-3, 1.0, 0.7)
    fx = tuple(1 if k in fixed else 0 for k in range(9))
    w = get("niris64")
    orc = O.Oracle(w.config(fixed=fx, regul=regul, regul_value=value, regul_weight=weight), w.wavelengths())
    free = [k for k in range(9) if not fx[k]]
    n = len(free)
    rng = np.random.default_rng(len(fixed))
    A = rng.normal(size=(n, n))
    H = A @ A.T
    g = rng.normal(size=n)
    par = np.array([1500.0, 1.1, 0.4, -0.7, 25.0, 0.9, 14.0, 1.05, 0.25])
    ratio = 0.37
    pen, H2, g2 = orc.regul(par, ratio, H, g)
    # augmented residuals: one row per regularised free parameter
    rows = [(ip, k) for ip, k in enumerate(free) if regul[k]]
    Jr = np.zeros((len(rows), n))
    r = np.zeros(len(rows))
    for j, (ip, k) in enumerate(rows):
        s = np.sqrt(ratio * weight[k])
        Jr[j, ip] = s
        r[j] = s * (par[k] - value[k])
    np.testing.assert_allclose(H2, H + Jr.T @ Jr, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(g2, g - Jr.T @ r, rtol=1e-14, atol=1e-14)
    want = sum(weight[k] * (par[k] - value[k]) ** 2 for k in range(9) if regul[k] and not fx[k])
    np.testing.assert_allclose(pen, want, rtol=1e-14)


# ---------------------------------------------------------------- Random_Jump
LIM = np.array([[5, 6000], [0, np.pi], [0, np.pi], [-20, 20], [1, 1000], [0, 5], [0, 200], [0, 10], [0, 1]],
               dtype=float)
BEST = np.array([800.0, 1.0, 0.3, 0.5, 40.0, 0.6, 12.0, 1.1, 0.4])
GUESS = np.array([1500.0, 2.0, 1.2, -0.8, 25.0, 0.5, 20.0, 0.9, 0.3])


@pytest.mark.parametrize("irun", [1, 2, 3])
def test_jump_runs_1_to_3_rotate_the_guess_azimuth(irun):
    """Runs 1-3 restart from the best run with phi = phi_guess + pi/4 irun,
    folded into [0, pi) by bounds_check."""
    p = O.jump(7, 123, irun, BEST, GUESS, LIM)
    want = BEST.copy()
    want[2] = np.fmod(GUESS[2] + np.pi / 4 * irun, np.pi)
    np.testing.assert_array_equal(np.delete(p, 2), np.delete(want, 2))
    np.testing.assert_allclose(p[2], want[2], rtol=0, atol=1e-15)


@pytest.mark.parametrize("irun", [4, 5, 9, 39])
def test_jump_runs_from_4_draw_in_the_reference_order(irun):
    """Runs >= 4 (Src/LM_FIT.c:185-197): the eight draws, in the reference's
    order, must be recoverable from the jumped parameters as the first eight
    N(0,1) values of the pixel's two jump blocks; B, theta, phi, v, a, eta0
    and beta move around the guess, the Doppler width around the best, S is
    the best's.  Wide limits keep bounds_check from clipping (the folds of
    theta and phi are undone modulo their period)."""
    seed, key = 20261203, 98765
    p = O.jump(seed, key, irun, BEST, GUESS, LIM)
    n = np.concatenate([O.gauss4(seed, key, irun, O.STREAM_JUMP), O.gauss4(seed, key, irun, O.STREAM_JUMP + 1)])
    rec = np.empty(8)
    rec[0] = (p[0] / GUESS[0] - 1) / 0.6
    # theta: fmod 2 pi, negative + 2 pi, > pi -> 2 pi - theta; try both branches
    th = GUESS[1] + np.pi / 8 * n[1]
    assert np.isclose(p[1], min(th % (2 * np.pi), 2 * np.pi - th % (2 * np.pi)), atol=1e-13)
    rec[1] = n[1]
    dphi = (p[2] - (GUESS[2] + np.pi / 4 * irun)) % np.pi
    rec[2] = min(dphi, dphi - np.pi, key=abs) / (np.pi / 4 * 0.3)
    rec[3] = (p[3] - GUESS[3]) / 5.0
    rec[4] = (p[4] / BEST[4] - 1) / 0.2
    rec[5] = (p[5] - GUESS[5]) / 0.1
    rec[6] = (p[6] - GUESS[6]) / 5.0
    rec[7] = (p[8] - GUESS[8]) / 0.1
    assert p[7] == BEST[7]
    # draws whose parameter stayed inside the limits are recovered exactly
    inside = np.array([LIM[0, 0] < p[0] < LIM[0, 1], True, abs(rec[2]) < 1 / 0.6, LIM[3, 0] < p[3] < LIM[3, 1],
                       LIM[4, 0] < p[4] < LIM[4, 1], LIM[5, 0] < p[5] < LIM[5, 1], LIM[6, 0] < p[6] < LIM[6, 1],
                       LIM[8, 0] < p[8] < LIM[8, 1]])
    assert inside.sum() >= 6
    np.testing.assert_allclose(rec[inside], n[inside], rtol=1e-9, atol=1e-9)


def test_jump_draws_are_standard_normal_and_independent():
    """The Box-Muller draws behind the jumps: 400k values pass a
    Kolmogorov-Smirnov test against N(0,1); the four values of a block and
    consecutive runs are uncorrelated; different pixels get different draws."""
    g = np.array([O.gauss4(11, pix, run, O.STREAM_JUMP + s) for pix in range(2500) for run in range(4, 24)
                  for s in (0, 1)])
    flat = g.ravel()
    assert stats.kstest(flat, "norm").pvalue > 1e-3
    assert abs(flat.mean()) < 5e-3 and abs(flat.std() - 1) < 5e-3
    c = np.corrcoef(g.T)
    assert np.max(np.abs(c - np.eye(4))) < 0.01
    assert not np.array_equal(O.gauss4(11, 1, 4, O.STREAM_JUMP), O.gauss4(11, 2, 4, O.STREAM_JUMP))


def test_restart_key_is_the_global_pixel_index():
    """orc_invert's pix_base: inverting the second half of a batch with base
    256 gives the records of the whole batch (num_run = 5 jumps at irun >= 4)."""
    w = get("niris64")
    orc = O.Oracle(w.config(num_run=5), w.wavelengths())
    prof, _ = orc.gen_map(w.seed, w.ranges(), 96)
    prof = prof.astype(np.float64)
    full = orc.invert(prof, nthreads=8)
    part = orc.invert(prof[48:], pix_base=48, nthreads=8)
    np.testing.assert_array_equal(part["res"], full["res"][48:])
    assert np.all(full["runs"][full["res"][:, 9] > 0] >= 5)
    other = orc.invert(prof[48:], pix_base=0, nthreads=8)
    assert np.any(other["res"] != full["res"][48:])


# ------------------------------------------------------------ learned policy
def test_learned_policy_one_run_from_a_good_prediction():
    """A prediction at the truth converges in one LM run (no heuristic runs,
    no fallback status); a prediction far off falls back to the heuristic
    runs, keeps the better record, and sets ME_PIX_FALLBACK."""
    w = get("hmi4k")
    orc = O.Oracle(w.config(num_run=1), w.wavelengths())
    prof, truth = orc.gen_map(w.seed, w.ranges(), 200, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
    prof = prof.astype(np.float64)
    good = orc.invert(prof, init=truth, learned=True, nthreads=8)
    conv = good["res"][:, 9] <= w.chisq_criteria
    assert conv.mean() > 0.9
    assert np.all(good["runs"][conv] == 1) and np.all((good["status"][conv] & 4) == 0)
    # heuristic result for comparison
    plain = orc.invert(prof, nthreads=8)
    bad = truth.copy()
    bad[:, 6] = 1.0         # eta0 at its lower bound
    bad[:, 4] = 80.0        # Doppler width at its upper bound
    bad[:, 0] = 3000.0
    lrn = orc.invert(prof, init=bad, learned=True, nthreads=8)
    fb = (lrn["status"] & 4) != 0
    assert fb.mean() > 0.3
    # a fallback pixel runs the learned run plus the heuristic runs (>= 3 at
    # num_run = 1) and keeps the better of the two records
    assert np.all(lrn["runs"][fb] >= 1 + 3)
    better = lrn["res"][fb, 9] <= plain["res"][fb, 9]
    assert better.all()
    took_plain = lrn["res"][fb, 9] == plain["res"][fb, 9]
    np.testing.assert_array_equal(lrn["res"][fb][took_plain], plain["res"][fb][took_plain])
    # without fallback the record is the learned run's
    assert np.all(lrn["runs"][~fb] == 1)


def test_learned_run_damping_bound_applies_to_the_learned_run_only():
    """learned_lambda_limit = 6e-3 ends an iteration of the learned run at its
    second rejected trial (1e-3 * 5^2 > 6e-3 >= 1e-3 * 5), so a pixel without
    fallback runs at most 2 trials per iteration it starts; the heuristic runs
    of a fallback keep lambda_limits[1] (9 rejected trials end an iteration)."""
    w = get("hmi4k")
    prof, truth = O.Oracle(w.config(), w.wavelengths()).gen_map(w.seed, w.ranges(), 200, ff_frac=w.ff_frac,
                                                                 ff_bmax=w.ff_bmax, sigma=w.sigma)
    prof = prof.astype(np.float64)
    start = truth.copy()
    start[:, 0] *= 1.3          # a start near, not at, the truth
    start[:, 3] += 0.3
    ref = O.Oracle(w.config(num_run=1, learned_lambda_limit=0.0), w.wavelengths()).invert(
        prof, init=start, learned=True, nthreads=8)
    cap = O.Oracle(w.config(num_run=1, learned_lambda_limit=6e-3), w.wavelengths()).invert(
        prof, init=start, learned=True, nthreads=8)
    direct = ((ref["status"] & 4) == 0) & ((cap["status"] & 4) == 0) & (ref["runs"] > 0)
    assert direct.mean() > 0.5
    assert np.all(cap["trials"][direct] <= 2 * (cap["iters"][direct] + 1))
    # under the reference bound a learned run ends with 9 rejected trials
    # unless it converges on the relative-gain test, so more trials run
    assert np.any(ref["trials"][direct] > 2 * (ref["iters"][direct] + 1))
    assert cap["trials"].sum() < ref["trials"].sum()
    # from a poor start most pixels fall back: their heuristic runs are the
    # plain inversion's (same Philox keys), with the reference bound, so
    # subtracting the plain inversion's counts leaves the learned run's own
    bad = truth.copy()
    bad[:, 6], bad[:, 4], bad[:, 0] = 1.0, 80.0, 3000.0
    plain = O.Oracle(w.config(num_run=1), w.wavelengths()).invert(prof, nthreads=8)
    for lim, bounded in ((0.0, False), (6e-3, True)):
        r = O.Oracle(w.config(num_run=1, learned_lambda_limit=lim), w.wavelengths()).invert(
            prof, init=bad, learned=True, nthreads=8)
        fb = (r["status"] & 4) != 0
        assert fb.mean() > 0.3
        assert np.all(r["runs"][fb] == plain["runs"][fb] + 1)
        lt = r["trials"][fb] - plain["trials"][fb]
        li = r["iters"][fb] - plain["iters"][fb]
        assert np.all(lt >= 1) and np.all(li >= 0)
        if bounded:
            assert np.all(lt <= 2 * (li + 1))
        else:
            assert np.any(lt > 2 * (li + 1))


def test_learned_run_default_bound_is_one_trial_per_iteration():
    """The default learned_lambda_limit (4e-3: 1e-3 * 5 > 4e-3) ends an
    iteration of the learned run at its first rejected trial: a pixel without
    fallback runs exactly one trial per iteration it starts (iters accepted
    ones plus the last), fewer trials than under 6e-3, and the records of
    pixels that converge within the bound agree with the two-trial run's
    where both accept only first trials."""
    w = get("hmi4k")
    cfg = w.config(num_run=1)
    assert abs(cfg.learned_lambda_limit - 4e-3) < 1e-15
    prof, truth = O.Oracle(cfg, w.wavelengths()).gen_map(w.seed, w.ranges(), 200, ff_frac=w.ff_frac,
                                                           ff_bmax=w.ff_bmax, sigma=w.sigma)
    prof = prof.astype(np.float64)
    start = truth.copy()
    start[:, 0] *= 1.3
    start[:, 3] += 0.3
    one = O.Oracle(cfg, w.wavelengths()).invert(prof, init=start, learned=True, nthreads=8)
    two = O.Oracle(w.config(num_run=1, learned_lambda_limit=6e-3), w.wavelengths()).invert(
        prof, init=start, learned=True, nthreads=8)
    direct = ((one["status"] & 4) == 0) & (one["runs"] > 0)
    assert direct.mean() > 0.5
    assert np.all(one["trials"][direct] == one["iters"][direct] + 1)
    assert one["trials"].sum() < two["trials"].sum()
    # a two-trial run that never rejected a trial took the same steps
    same = direct & ((two["status"] & 4) == 0) & (two["trials"] == two["iters"] + 1)
    assert same.any()
    np.testing.assert_array_equal(one["res"][same], two["res"][same])


def test_learned_run_s_bound_reaches_lines_without_continuum():
    """Init_Guess bounds S to [0.6, 1.4] Imax (Src/ME_SOLVER.c:662-666): a line
    that never reaches the continuum in the window (truth S above 1.4 Imax)
    cannot be fitted by any heuristic run.  learned_s_limit = 2 lets the
    learned run go up to 2 Imax: started at the truth it converges there, with
    S above 1.4 Imax and no fallback; with 0 (Init_Guess's bound) the same
    starts are clamped and fall back.  Heuristic inversions never see it."""
    w = get("hmi4k")
    orc0 = O.Oracle(w.config(num_run=1, learned_s_limit=0.0), w.wavelengths())
    orc2 = O.Oracle(w.config(num_run=1, learned_s_limit=2.0), w.wavelengths())
    prof, truth = orc0.gen_map(w.seed, w.ranges(), 3000, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
    prof = prof.astype(np.float64)
    imax = prof[:, 0, :].max(axis=1)
    hi = truth[:, 7] > 1.45 * imax
    assert 40 <= hi.sum() <= 300
    p, t, im = prof[hi], truth[hi], imax[hi]
    a = orc0.invert(p, init=t, learned=True, nthreads=8)
    b = orc2.invert(p, init=t, learned=True, nthreads=8)
    crit = w.chisq_criteria
    # Init_Guess's bound holds for every run of a: every record has S <= 1.4
    # Imax; a third of these pixels miss the criterion from the clamped
    # truth and fall back (at nl = 6 the others fit within the bound)
    assert np.all(a["res"][:, 7] <= 1.4 * im * (1 + 1e-12))
    assert np.mean((a["status"] & 4) != 0) > 0.2
    conv_b = (b["res"][:, 9] > 0) & (b["res"][:, 9] <= crit)
    direct = (b["status"] & 4) == 0
    assert np.mean(conv_b & direct) > 0.97
    assert b["trials"].sum() < a["trials"].sum()
    assert np.all(b["res"][direct, 7] <= 2.0 * im[direct] * (1 + 1e-12))
    assert np.mean(b["res"][direct, 7] > 1.4 * im[direct]) > 0.9
    # heuristic inversions: identical records whatever the learned bound
    plain0, plain2 = orc0.invert(prof[:200], nthreads=8), orc2.invert(prof[:200], nthreads=8)
    np.testing.assert_array_equal(plain0["res"], plain2["res"])
