"""Learned mode judged by the FP64 oracle: the network's predictions are
copied off the device and handed to the oracle's learned policy (one LM run
from the prediction; heuristic runs only above chisq_criteria; the better
record kept), on the same hmi4k pixels the engine inverts."""
import numpy as np
import pytest

import oracle as O
from gpu_util import DEG, device_tensor, gpu_inversion, within_parity, within_physical
from meinv import ME_FP32, ME_INIT_LEARNED
from meinv.workloads import get

pytestmark = pytest.mark.gpu

N = 4096


@pytest.fixture(scope="module")
def learned_case():
    import torch
    w = get("hmi4k")
    ranges = w.ranges()
    with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED) as inv:
        inv.mlp_train(w.seed + 100, ranges, 262144, batch=2048, epochs=4, lr=2e-3, ff_frac=w.ff_frac,
                      ff_bmax=w.ff_bmax, sigma=w.sigma)
        params = inv.mlp_params()
        d_prof = device_tensor((N, 4, w.nl), torch.float32)
        inv.gen_map_device(w.seed, ranges, N, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                           sigma=w.sigma)
        d_init = device_tensor((N, 9), torch.float64)
        inv.mlp_predict_device(d_prof.data_ptr(), N, d_init.data_ptr())
        torch.cuda.synchronize()
        prof = d_prof.cpu().numpy()
        pred = d_init.cpu().numpy()
        d_res = device_tensor((N, 10), torch.float64)
        inv.invert_device(d_prof.data_ptr(), N, d_res.data_ptr())
        res32_net = d_res.cpu().numpy()
    return w, params, prof, pred, res32_net


@pytest.mark.parametrize("s_limit", [0.0, 2.0])
def test_learned_fp64_engine_matches_the_oracle(learned_case, s_limit):
    """FP64 engine in learned mode seeded by the predictions vs the oracle's
    learned policy seeded by the same values: agreement at the chaos floor,
    with the learned run's S bound at Init_Guess's 1.4 Imax (0) and at the
    default 2 Imax."""
    w, params, prof, pred, _ = learned_case
    prof64 = prof.astype(np.float64)
    orc = O.Oracle(w.config(init_mode=ME_INIT_LEARNED, learned_s_limit=s_limit), w.wavelengths())
    ref = orc.invert(prof64, init=pred, learned=True, nthreads=8)
    rng = np.random.default_rng(0)
    self_ = orc.invert(prof64 * (1 + 1e-15 * rng.standard_normal(prof64.shape)), init=pred, learned=True, nthreads=8)
    floor = within_parity(self_["res"], ref["res"]).mean()
    with gpu_inversion("hmi4k", init_mode=ME_INIT_LEARNED, learned_s_limit=s_limit) as inv:
        out = inv.invert(prof64, init=pred, per_pixel=True)
    ok = within_parity(out["res"], ref["res"])
    fb_dev = (out["stats"].per_pixel["status"] & 4) != 0
    fb_orc = (ref["status"] & 4) != 0
    conv = np.mean(out["res"][:, 9] <= w.chisq_criteria)
    conv_o = np.mean(ref["res"][:, 9] <= w.chisq_criteria)
    print(f"learned FP64 engine vs oracle: {ok.mean():.4f} within tolerance (oracle self {floor:.4f}); fallback "
          f"{fb_dev.mean():.4f} / {fb_orc.mean():.4f}; converged {conv:.4f} / {conv_o:.4f}")
    assert np.mean(fb_dev == fb_orc) >= 0.99
    assert ok.mean() >= floor - 0.03


def test_learned_fp32_error_against_the_oracle(learned_case):
    """The headline form (FP32 synthesis, network prediction) against the FP64
    oracle's learned policy from the same predictions: per-parameter error and
    the fraction of pixels within 0.1 G / 0.01 deg / 1 m/s."""
    w, params, prof, pred, res32 = learned_case
    orc = O.Oracle(w.config(init_mode=ME_INIT_LEARNED), w.wavelengths())
    ref = orc.invert(prof.astype(np.float64), init=pred, learned=True, nthreads=8)["res"]
    conv_a = np.mean(res32[:, 9] <= w.chisq_criteria)
    conv_b = np.mean(ref[:, 9] <= w.chisq_criteria)
    both = (res32[:, 9] <= w.chisq_criteria) & (ref[:, 9] <= w.chisq_criteria)
    phys_all = within_physical(res32, ref).mean()
    phys = within_physical(res32[both], ref[both]).mean()
    d = np.abs(res32[both] - ref[both])
    print(f"learned FP32 engine vs FP64 oracle: converged {conv_a:.4f} / {conv_b:.4f}; within 0.1 G / 0.01 deg / "
          f"1 m/s: {phys_all:.4f} of all pixels, {phys:.4f} of pixels converged in both; median |dB| "
          f"{np.median(d[:, 0]):.3g} G, |dtheta| {np.median(d[:, 1]) * DEG:.3g} deg, |dv| "
          f"{np.median(d[:, 3]) * 1e3:.3g} m/s")
    assert abs(conv_a - conv_b) <= 0.02
    assert phys >= 0.6


# Learned run bounded to two trials per LM iteration (learned_lambda_limit:
# 1e-3 * 5^2 > 6e-3 ends an iteration at its second rejected trial); the
# heuristic runs of the fallback keep lambda_limits[1]
LEARNED_LIMIT = 6e-3


def test_learned_two_trial_bound_matches_the_oracle(learned_case):
    """FP64 engine and the oracle's learned policy under the same learned-run
    bound, seeded by the same predictions: the learned run never solves with a
    rounding-decided matrix, so its records agree on (nearly) every pixel that
    does not fall back; fallback pixels agree at the chaos floor."""
    w, params, prof, pred, _ = learned_case
    prof64 = prof.astype(np.float64)
    cfg = w.config(init_mode=ME_INIT_LEARNED, learned_lambda_limit=LEARNED_LIMIT)
    orc = O.Oracle(cfg, w.wavelengths())
    ref = orc.invert(prof64, init=pred, learned=True, nthreads=8)
    with gpu_inversion("hmi4k", init_mode=ME_INIT_LEARNED, learned_lambda_limit=LEARNED_LIMIT) as inv:
        out = inv.invert(prof64, init=pred, per_pixel=True)
    pp = out["stats"].per_pixel
    ok = within_parity(out["res"], ref["res"])
    fb_dev = (pp["status"] & 4) != 0
    fb_orc = (ref["status"] & 4) != 0
    direct = ~fb_dev & ~fb_orc
    print(f"learned FP64 engine vs oracle, two-trial learned run: {ok.mean():.4f} within tolerance, "
          f"{ok[direct].mean():.4f} of pixels without fallback; fallback {fb_dev.mean():.4f} / {fb_orc.mean():.4f}")
    assert np.mean(fb_dev == fb_orc) >= 0.995
    assert ok[direct].mean() >= 0.99
    # a pixel without fallback ran one LM run of at most two trials per
    # iteration: its iterations that accepted a trial, plus the last one
    assert np.all(pp["trials"][direct] <= 2 * (pp["iters"][direct] + 1))
    assert np.all(ref["trials"][direct] <= 2 * (ref["iters"][direct] + 1))


def test_learned_two_trial_bound_keeps_the_records(learned_case):
    """FP32 engine, headline form: the bound changes the learned run's trial
    count, not its outcome; converged fraction and chi2 distribution against
    the reference damping bound, on the device's own predictions."""
    import torch
    w, params, prof, pred, _ = learned_case
    d_prof = torch.from_numpy(prof).to("cuda:0")
    d_res = device_tensor((N, 10), torch.float64)
    recs, trials = [], []
    for lim in (0.0, LEARNED_LIMIT):
        with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED, learned_lambda_limit=lim) as inv:
            inv.mlp_import(params, w.ranges())
            st = inv.invert_device(d_prof.data_ptr(), N, d_res.data_ptr())
            recs.append(d_res.cpu().numpy())
            trials.append(st.n_trial)
    a, b = recs
    ca = (a[:, 9] > 0) & (a[:, 9] <= w.chisq_criteria)
    cb = (b[:, 9] > 0) & (b[:, 9] <= w.chisq_criteria)
    same = np.mean(np.all(a == b, axis=1))
    # the chi2 distribution below its tail (the top percent is the pixels
    # that converge nowhere, whose last record depends on which run ends last)
    qa = np.quantile(a[a[:, 9] > 0, 9], [0.1, 0.5, 0.9])
    qb = np.quantile(b[b[:, 9] > 0, 9], [0.1, 0.5, 0.9])
    print(f"two-trial learned run: trials {trials[1]} vs {trials[0]}, converged {cb.mean():.4f} vs {ca.mean():.4f}, "
          f"identical records {same:.4f}, chi2 quantiles {qb} vs {qa}")
    assert trials[1] < trials[0]
    assert cb.mean() >= ca.mean() - 0.002
    assert same >= 0.99
    assert np.allclose(qa, qb, rtol=0.02)


def test_learned_records_do_not_depend_on_the_launch_shape(learned_case):
    """Learned records are the same bit for bit under every launch shape of
    one kernel build: the FP64-SVD build (the learned default on the FP32
    path) with 256- and 128-lane blocks, and the FP32-SVD build
    (svd_fp32 = 1) at one and two waves per SIMD."""
    import torch
    w, params, prof, _, _ = learned_case
    d_prof = torch.from_numpy(prof).to("cuda:0")
    d_res = device_tensor((N, 10), torch.float64)
    for svd32, shapes in ((0, ((256, 4), (128, 4))), (1, ((256, 4), (256, 8)))):
        recs = []
        for block, waves in shapes:
            with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED, block_size=block,
                               waves_per_cu=waves, svd_fp32=svd32) as inv:
                inv.mlp_import(params, w.ranges())
                inv.invert_device(d_prof.data_ptr(), N, d_res.data_ptr())
                recs.append(d_res.cpu().numpy())
        assert np.mean(recs[0][:, 9] > 0) > 0.9
        np.testing.assert_array_equal(recs[1], recs[0])
