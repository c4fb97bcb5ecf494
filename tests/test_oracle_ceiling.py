"""The convergence ceiling of the hmi4k distribution under the reference's
bounds: Init_Guess bounds S to [0.6, 1.4] x Imax, the largest observed I of
the pixel (Src/ME_SOLVER.c:662-666), and on the 6-sample HMI window a wide,
strong line never reaches the continuum, so some truths have S above 1.4 Imax.
Starting each pixel's LM run at its truth (the oracle's learned policy fed
the truths) with num_run = 1, the converged fraction stays below 99 %, and
every pixel that fails to converge is one of those: without Gaussian restarts
no run leaves the bounded-S basin.  (Restarts, num_run >= 5, find other fits
under the criterion for some of them.)  The learned run's own S bound of
2 Imax (learned_s_limit) lifts the ceiling: from the truth every pixel
converges."""
import numpy as np

import oracle as O
from meinv.workloads import get


def test_starting_at_the_truth_converges_below_99_percent():
    w = get("hmi4k")
    orc = O.Oracle(w.config(learned_s_limit=0.0), w.wavelengths())
    n = 8192
    prof, truth = orc.gen_map(w.seed, w.ranges(), n, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
    p64 = prof.astype(np.float64)
    res = orc.invert(p64, init=truth, learned=True, nthreads=8)["res"]
    conv = (res[:, 9] > 0) & (res[:, 9] <= w.chisq_criteria)
    imax = p64[:, 0, :].max(axis=1)
    s_bound = truth[:, 7] > 1.4 * imax
    print(f"from the truth: converged {conv.mean():.4f}; S above 1.4 Imax {s_bound.mean():.4f}; "
          f"not converged with S in bounds {np.mean(~conv & ~s_bound):.4f}")
    assert 0.98 <= conv.mean() < 0.99
    assert not np.any(~conv & ~s_bound)
    assert s_bound.mean() > 0.03
    # the learned run's default bound, 2 Imax
    orc2 = O.Oracle(w.config(learned_s_limit=2.0), w.wavelengths())
    res2 = orc2.invert(p64, init=truth, learned=True, nthreads=8)["res"]
    conv2 = (res2[:, 9] > 0) & (res2[:, 9] <= w.chisq_criteria)
    print(f"from the truth, S up to 2 Imax: converged {conv2.mean():.4f}")
    assert conv2.mean() >= 0.999
