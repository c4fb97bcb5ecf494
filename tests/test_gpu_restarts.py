"""The restart regime on the device: num_run >= 2, where runs 1-3 rotate the
azimuth of the best run and runs >= 4 are Gaussian jumps (Random_Jump,
Src/LM_FIT.c:157-201) drawn from Philox4x32-10, and the run-count policy
(Src/LM_FIT.c:701-724).  The oracle draws the same Philox streams, so the
device must agree with it as often as the oracle agrees with itself."""
import numpy as np
import pytest

import oracle as O
from gpu_util import chaos_floor, gpu_inversion, oracle_map, within_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("num_run", [2, 5])
@pytest.mark.parametrize("name", ["niris64", "hmi4k"])
def test_inversion_fp64_parity_restarts(name, num_run):
    n = 512
    w, _, prof, truth = oracle_map(name, n)
    orc = O.Oracle(w.config(num_run=num_run), w.wavelengths())
    prof64 = prof.astype(np.float64)
    floor, ref = chaos_floor(orc, prof64)
    with gpu_inversion(name, num_run=num_run) as inv:
        out = inv.invert(prof64, per_pixel=True)
    ok = within_parity(out["res"], ref["res"])
    runs = out["stats"].per_pixel["runs"]
    same_runs = np.mean(runs == ref["runs"])
    jumped = np.mean(runs >= 5)
    st = out["stats"]
    print(f"{name} num_run={num_run}: {ok.mean():.4f} within tolerance (oracle self {floor:.4f}); same run count "
          f"{same_runs:.4f}; pixels with Gaussian jumps (>= 5 runs) {jumped:.4f}; runs/px {runs.mean():.2f}; "
          f"converged {st.n_converged}/{st.n_valid}")
    # num_run >= 2 runs at least 5 LM runs per pixel (H9): every pixel
    # reaches irun = 4, the first Gaussian jump
    assert np.all(runs[ref["res"][:, 9] > 0] >= 5)
    assert ok.mean() >= floor - 0.03


def test_restart_draws_follow_the_pixel_base():
    """me_ctx_set_pixel_base: the second half of a batch inverted on its own
    with base 256 equals the whole batch's second half bit for bit; with base
    0 the pixels that jumped differ (their draws belong to pixels 0..255)."""
    n = 512
    w, orc, prof, truth = oracle_map("niris64", n)
    prof64 = prof.astype(np.float64)
    with gpu_inversion("niris64", num_run=5) as inv:
        full = inv.invert(prof64, per_pixel=True)
        inv.set_pixel_base(256)
        part = inv.invert(prof64[256:])
        inv.set_pixel_base(0)
        wrong = inv.invert(prof64[256:])
    np.testing.assert_array_equal(part["res"], full["res"][256:])
    jumped = full["stats"].per_pixel["runs"][256:] >= 5
    assert jumped.sum() > 20
    differ = np.any(wrong["res"] != full["res"][256:], axis=1)
    assert differ[jumped].mean() > 0.5
    assert not differ[~jumped].any()
    # the oracle keys its draws the same way: the device's second half agrees
    # with the oracle's, inverted with base 256, at the restart chaos floor
    orc5 = O.Oracle(w.config(num_run=5), w.wavelengths())
    ref = orc5.invert(prof64[256:], pix_base=256, nthreads=8)
    ok = within_parity(part["res"], ref["res"]).mean()
    print(f"second half with base 256 vs the oracle: {ok:.4f} within tolerance")
    assert ok >= 0.75
