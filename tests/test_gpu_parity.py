"""Device path against the CPU FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from gpu_util import DEG, chaos_floor, device_tensor, gpu_inversion, oracle_map, within_parity, within_physical
from meinv import ME_FP32, Inversion
from meinv.workloads import get

pytestmark = pytest.mark.gpu


def _random_params(w, n, seed=1):
    rng = np.random.default_rng(seed)
    r = w.ranges()
    p = r[:, 0] + (r[:, 1] - r[:, 0]) * rng.uniform(size=(n, 9))
    p[:, 5] = np.maximum(p[:, 5], 1e-3)
    return p


@pytest.mark.parametrize("name,nfig", [("niris64", 6), ("hinode", 6), ("hmi4k", 6), ("niris64", 5), ("hmi4k", 5),
                                       ("niris64", 4), ("hmi4k", 4)])
def test_synthesis_fp64_matches_oracle(name, nfig):
    """FP64 synthesis and Jacobian, every Faddeeva region split that
    Nfigures selects (6: Algorithm 916 / Hui p6; 5 and 4: Humlicek W4 and Hui
    p6, Src/FADDEEVA.c:510-549, 581-583)."""
    w = get(name)
    orc = O.Oracle(w.config(nfigures=nfig), w.wavelengths())
    par = _random_params(w, 300)
    with gpu_inversion(name, nfigures=nfig) as inv:
        syn, J = inv.synth(par, jac=True)
    ref, Jr = orc.synth(par, jac=True)
    scale = par[:, 7][:, None, None]
    assert np.abs(syn - ref).max() / 1.2 <= 1e-12 * 10 or np.max(np.abs(syn - ref) / scale) <= 1e-12
    jscale = np.abs(Jr).max(axis=(1, 3), keepdims=True) + 1e-300
    assert np.max(np.abs(J - Jr) / jscale) <= 1e-9


@pytest.mark.parametrize("name", ["niris64", "hinode", "hmi4k"])
def test_synthesis_fp32_error_against_fp64_oracle(name):
    """FP32 synthesis + Jacobian, error stated against the FP64 oracle."""
    w = get(name)
    orc = O.Oracle(w.config(), w.wavelengths())
    par = _random_params(w, 300, seed=2)
    with gpu_inversion(name, precision=ME_FP32) as inv:
        syn, J = inv.synth(par, jac=True)
    ref, Jr = orc.synth(par, jac=True)
    err = np.max(np.abs(syn - ref) / par[:, 7][:, None, None])
    jscale = np.abs(Jr).max(axis=(1, 3), keepdims=True) + 1e-300
    jerr = np.max(np.abs(J - Jr) / jscale)
    print(f"{name}: FP32 Stokes max error {err:.2e} of S; Jacobian max error {jerr:.2e} of row scale")
    # the quotient-rule rows (D*N' - D'*N)/D^2 cancel in FP32 near the
    # continuum; the LM only needs a descent direction, chi2 decides in FP64
    assert err <= 2e-5
    assert jerr <= 2e-2


def test_map_generator_matches_oracle():
    import torch
    for name in ("niris64", "hmi4k"):
        w, orc, prof, truth = oracle_map(name, 2000)
        with gpu_inversion(name) as inv:
            d_prof = device_tensor((2000, 4, w.nl), torch.float32)
            d_truth = device_tensor((2000, 9), torch.float64)
            inv.gen_map_device(w.seed, w.ranges(), 2000, d_prof.data_ptr(), d_truth.data_ptr(),
                               ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
            torch.cuda.synchronize()
            gp, gt = d_prof.cpu().numpy(), d_truth.cpu().numpy()
        np.testing.assert_array_equal(gt, truth)
        # FP64 synthesis + Box-Muller on device vs glibc: equal after the
        # float32 rounding except at rare rounding boundaries (1 ulp)
        ulp = np.spacing(np.abs(prof).astype(np.float32))
        assert np.all(np.abs(gp - prof) <= ulp)
        assert np.mean(gp == prof) > 0.999


def _parity(name, n, init_mode, **over):
    w, orc0, prof, truth = oracle_map(name, n)
    orc = O.Oracle(w.config(**over), w.wavelengths())
    prof64 = prof.astype(np.float64)
    init = None
    if init_mode == "explicit":
        init = np.array([orc.init_guess(prof64, pix=i)["guess"] for i in range(n)])
        init[:, 3] += 0.1  # a start different from the heuristic, fed to both
    floor, ref = chaos_floor(orc, prof64, init=init)
    with gpu_inversion(name, **over) as inv:
        out = inv.invert(prof64, init=init, per_pixel=True)
    ok = within_parity(out["res"], ref["res"])
    return ok, floor, out, ref


@pytest.mark.parametrize("name", ["niris64", "hinode", "hmi4k"])
def test_inversion_fp64_parity_two_trial_cap(name):
    """With at most two trials per LM iteration (lambda_limits = 1e-5, 6e-3)
    no trial solves with a rounding-decided matrix, and the FP64 device path
    must agree with the oracle pixel by pixel at 0.1 G / 0.01 deg / 1 m/s
    (and 1e-6 relative on the other parameters and chi2)."""
    ok, floor, out, ref = _parity(name, 1024 if name == "niris64" else 512, "heuristic",
                                  lambda_limits=(1e-5, 6e-3))
    print(f"{name} two-trial cap: {ok.mean():.4f} of pixels within tolerance (oracle self {floor:.4f})")
    assert ok.mean() >= 0.98
    assert ok.mean() >= floor - 0.01


@pytest.mark.parametrize("wexp", [-120, 120])
@pytest.mark.parametrize("name", ["niris64", "hmi4k"])
def test_inversion_fp64_parity_extreme_weights(name, wexp):
    """Stokes weights of 2^-120 or 2^120 scale the normal matrix by 2^-240 or
    2^240 (and chi2 with it): the SVD, its Givens pairs and the solve must
    handle magnitudes far from 1.  Two-trial cap, as in the parity test above:
    agreement with the oracle pixel by pixel.  (Much further out, near
    2^+-300, the reference's own QR shift overflows or underflows and its
    parameters turn to NaN.)"""
    wt = (2.0 ** wexp,) * 4
    ok, floor, out, ref = _parity(name, 512, "heuristic", weights=wt, lambda_limits=(1e-5, 6e-3))
    print(f"{name} weights 2^{wexp}: {ok.mean():.4f} of pixels within tolerance (oracle self {floor:.4f})")
    assert out["stats"].n_svd_fail == 0
    assert ok.mean() >= 0.98
    assert ok.mean() >= floor - 0.01


@pytest.mark.parametrize("init_mode", ["heuristic", "explicit"])
@pytest.mark.parametrize("name", ["niris64", "hinode", "hmi4k"])
def test_inversion_fp64_parity_default_policy(name, init_mode):
    """Default LM policy (up to 9 trials per iteration), num_run = 1,
    identical starting points: agreement must reach the oracle's own
    agreement under a 1e-15 input perturbation, less a 3% sampling margin."""
    ok, floor, out, ref = _parity(name, 1024 if name == "niris64" else 512, init_mode)
    same_runs = np.mean(out["stats"].per_pixel["runs"] == ref["runs"])
    st = out["stats"]
    print(f"{name} {init_mode}: {ok.mean():.4f} within tolerance, oracle self-agreement {floor:.4f}; "
          f"same run count {same_runs:.4f}; converged {st.n_converged}/{st.n_valid}")
    assert ok.mean() >= floor - 0.03
    assert st.n_valid == len(ok)


def test_inversion_fp32_error_against_fp64_oracle():
    w, orc, prof, truth = oracle_map("niris64", 1024)
    ref = orc.invert(prof.astype(np.float64), nthreads=8)
    with gpu_inversion("niris64", precision=ME_FP32) as inv:
        out = inv.invert(prof)
    a, b = out["res"], ref["res"]
    conv_a = np.mean(a[:, 9] <= w.chisq_criteria)
    conv_b = np.mean(b[:, 9] <= w.chisq_criteria)
    d = np.abs(a - b)
    med = [np.median(d[:, 0]), np.median(d[:, 1]) * DEG, np.median(d[:, 3]) * 1e3]
    phys = within_physical(a, b).mean()
    print(f"FP32 vs FP64 oracle: {phys:.4f} of pixels within 0.1 G / 0.01 deg / 1 m/s; converged "
          f"{conv_a:.4f} (oracle {conv_b:.4f}); median |dB| {med[0]:.3g} G, |dtheta| {med[1]:.3g} deg, "
          f"|dv| {med[2]:.3g} m/s")
    assert abs(conv_a - conv_b) <= 0.05
    assert phys >= 0.6


def test_invalid_and_empty_batches():
    w, orc, prof, truth = oracle_map("niris64", 8)
    p = prof.astype(np.float64)
    p[2, 0, 5] = np.nan
    p[5, 0] *= 0.01
    with gpu_inversion("niris64") as inv:
        out = inv.invert(p, per_pixel=True, fit=True)
        assert np.all(out["res"][[2, 5]] == 0) and np.all(out["fit"][[2, 5]] == 0)
        assert list(out["stats"].per_pixel["status"][[2, 5]] & 1) == [1, 1]
        assert out["stats"].n_valid == 6
        empty = inv.invert(np.zeros((0, 4, w.nl)))
        assert empty["res"].shape == (0, 10)
        one = inv.invert(p[:1])
        ref = orc.invert(p[:1])
        assert within_parity(one["res"], ref["res"]).all()


def test_fit_profiles_are_the_best_model():
    w, orc, prof, truth = oracle_map("niris64", 64)
    with gpu_inversion("niris64") as inv:
        out = inv.invert(prof, fit=True)
        syn = inv.synth(out["res"][:, :9])
    np.testing.assert_allclose(out["fit"], syn, atol=1e-14)


@pytest.mark.parametrize("fixed", [(5, 8), (1, 2), (0, 4, 6, 7)])
@pytest.mark.parametrize("name", ["niris64", "hmi4k"])
def test_inversion_fp64_parity_fixed_parameters(name, fixed):
    """Fixed parameters (npar < 9): the device packs the free rows/columns
    into a zero-padded 9x9 system; agreement with the oracle's packed
    npar x npar path at the two-trial cap and under the default policy."""
    fx = tuple(1 if k in fixed else 0 for k in range(9))
    n = 512
    ok2, floor2, out2, _ = _parity(name, n, "heuristic", fixed=fx, lambda_limits=(1e-5, 6e-3))
    ok, floor, out, ref = _parity(name, n, "heuristic", fixed=fx)
    print(f"{name} fixed {fixed}: two-trial cap {ok2.mean():.4f} (oracle self {floor2:.4f}); default "
          f"{ok.mean():.4f} (oracle self {floor:.4f})")
    assert ok2.mean() >= floor2 - 0.02
    assert ok.mean() >= floor - 0.03


@pytest.mark.parametrize("name", ["niris64", "hmi4k"])
def test_inversion_fp64_parity_regularisation(name):
    """Get_Regul on the device (Src/LM_FIT.c:354-396): the reference's
    default theta regularisation plus an eta term, against the oracle with the
    per-pixel reading of H2 (no LM state before a pixel's first run, which is
    what a pixel-parallel engine can reproduce); the oracle's agreement with
    its own sequential reading (state carried from pixel to pixel) is printed
    as the size of the H2 effect."""
    over = dict(regul_theta=True, regul=(0, 1, 0, 0, 0, 0, 1, 0, 0),
                regul_value=(0, 0.5 * np.pi, 0, 0, 0, 0, 10.0, 0, 0), regul_weight=(0, 1e-4, 0, 0, 0, 0, 2e-3, 0, 0))
    ok, floor, out, ref = _parity(name, 512, "heuristic", **over)
    w, orc0, prof, truth = oracle_map(name, 512)
    seq = O.Oracle(w.config(**over), w.wavelengths()).invert(prof.astype(np.float64), h2=O.ORC_H2_SEQUENTIAL)
    h2 = within_parity(seq["res"], ref["res"]).mean()
    print(f"{name} regularised: {ok.mean():.4f} within tolerance (oracle self {floor:.4f}); oracle sequential vs "
          f"per-pixel H2 {h2:.4f}")
    assert ok.mean() >= floor - 0.03


def test_inversion_hmi_ref_rotates_phi_only():
    """HMI_REF (Src/LM_FIT.c:730-733): the same fit with phi moved by pi/2
    mod pi, and the oracle's result with the same setting."""
    w, orc0, prof, truth = oracle_map("hmi4k", 256)
    with gpu_inversion("hmi4k") as inv:
        base = inv.invert(prof)["res"]
    with gpu_inversion("hmi4k", hmi_ref=1) as inv:
        rot = inv.invert(prof)["res"]
    keep = [k for k in range(10) if k != 2]
    np.testing.assert_array_equal(rot[:, keep], base[:, keep])
    phi = base[:, 2] + 0.5 * np.pi
    phi = np.where(phi > np.pi, phi - np.pi, phi)
    valid = base[:, 9] > 0
    np.testing.assert_array_equal(rot[valid, 2], phi[valid])
    ok, floor, out, ref = _parity("hmi4k", 256, "heuristic", hmi_ref=1)
    assert ok.mean() >= floor - 0.03


def test_region_counts_switch_changes_counters_only():
    """me_ctx_set_region_counts: the FP32 path's per-region Faddeeva counts are
    gathered only when on; the inversion result is the same either way, and
    with counting on the FP32 path counts what the FP64 path counts."""
    w, orc, prof, truth = oracle_map("hmi4k", 2048)
    with gpu_inversion("hmi4k", precision=ME_FP32) as inv:
        inv.set_region_counts(False)
        off = inv.invert(prof)
        inv.set_region_counts(True)
        on = inv.invert(prof)
    np.testing.assert_array_equal(on["res"], off["res"])
    assert sum(off["stats"].n_fadd) == 0
    n_on = np.array(on["stats"].n_fadd)
    assert n_on.sum() > 0
    # every synthesis evaluates 3 Faddeeva calls per wavelength
    st = on["stats"]
    assert n_on.sum() == 3 * w.nl * (st.n_jac + st.n_trial)


def test_autotuner_picks_a_valid_shape_and_keeps_results():
    """me_ctx_autotune: the chosen block size and resident waves are one of
    the candidates (every SIMD of a CU holds a wave: four or eight per CU),
    stay in the context, and the inversion is the same under every launch
    shape, for both kernel builds (FP64 SVD at one wave per SIMD, the default,
    and the FP32 SVD's build at one and two)."""
    import torch
    w = get("hmi4k")
    n = 20000
    with gpu_inversion("hmi4k", precision=ME_FP32) as inv:
        d_prof = device_tensor((n, 4, w.nl), torch.float32)
        inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                           sigma=w.sigma)
        d_res = device_tensor((n, 10), torch.float64)
        inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        before = d_res.cpu().numpy()
        tune = inv.autotune(d_prof.data_ptr(), n)
        print(tune)
        assert tune["block_size"] in (64, 128, 256) and tune["waves_per_cu"] >= 1 and tune["kernel_ms"] > 0
        assert tune["waves_per_cu"] % (tune["block_size"] // 64) == 0
        assert tune["waves_per_cu"] in (4, 8)
        d_res.zero_()
        inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        np.testing.assert_array_equal(d_res.cpu().numpy(), before)
    for block, waves in ((64, 0), (128, 0), (256, 4), (128, 4)):
        with gpu_inversion("hmi4k", precision=ME_FP32, block_size=block, waves_per_cu=waves) as inv:
            d_res.zero_()
            inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            np.testing.assert_array_equal(d_res.cpu().numpy(), before)
    # the FP32 SVD's build (svd_fp32 = 1): its own records, the same at one
    # and two waves per SIMD
    outs = []
    for block, waves in ((256, 0), (256, 8), (256, 4), (128, 4)):
        with gpu_inversion("hmi4k", precision=ME_FP32, svd_fp32=1, block_size=block, waves_per_cu=waves) as inv:
            d_res.zero_()
            inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            outs.append(d_res.cpu().numpy())
    assert np.mean(outs[0][:, 9] > 0) > 0.9
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.gpu
def test_split_jacobian_is_independent_of_the_launch_shape():
    """At nl >= 12 the FP32 path shares each synthesis pass across the wave
    (fixed wavelength parts per pixel, whichever lanes compute them): the
    results must not depend on which pixels share a wave, i.e. on the block
    size and the number of resident waves, in either kernel build (FP64 SVD,
    the default here, and FP32 SVD at one and two waves per SIMD)."""
    import torch
    w = get("niris64")
    n = 6000
    d_prof = device_tensor((n, 4, w.nl), torch.float32)
    d_res = device_tensor((n, 10), torch.float64)
    first = True
    for svd32, shapes in ((0, ((256, 0), (64, 0), (128, 2))), (1, ((256, 0), (256, 4), (128, 4)))):
        outs = []
        for block, waves in shapes:
            with gpu_inversion("niris64", precision=ME_FP32, block_size=block, waves_per_cu=waves,
                               svd_fp32=svd32) as inv:
                if first:
                    inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), None, ff_frac=w.ff_frac,
                                       ff_bmax=w.ff_bmax, sigma=w.sigma)
                    first = False
                d_res.zero_()
                inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
                outs.append(d_res.cpu().numpy())
        assert np.mean(outs[0][:, 9] > 0) > 0.9
        np.testing.assert_array_equal(outs[1], outs[0])
        np.testing.assert_array_equal(outs[2], outs[0])


@pytest.mark.parametrize("name,n", [("hinode", 512), ("hmi4k", 2048)])
def test_inversion_fp32_plain_error_against_fp64_oracle(name, n):
    """FP32 plain inversion (hinode: nl = 112, the Jacobian pass split over
    two wavelength halves; hmi4k: the headline precision, nl = 6) against the
    FP64 oracle from the same heuristic start: converged fraction within 2
    points; within-tolerance fraction and median differences reported."""
    w, orc, prof, truth = oracle_map(name, n)
    ref = orc.invert(prof.astype(np.float64), nthreads=8)
    with gpu_inversion(name, precision=ME_FP32) as inv:
        out = inv.invert(prof)
    a, b = out["res"], ref["res"]
    conv_a = np.mean(a[:, 9] <= w.chisq_criteria)
    conv_b = np.mean(b[:, 9] <= w.chisq_criteria)
    d = np.abs(a - b)
    phys = within_physical(a, b).mean()
    print(f"{name} FP32 plain vs FP64 oracle: {phys:.4f} of pixels within 0.1 G / 0.01 deg / 1 m/s; converged "
          f"{conv_a:.4f} (oracle {conv_b:.4f}); median |dB| {np.median(d[:, 0]):.3g} G, |dtheta| "
          f"{np.median(d[:, 1]) * DEG:.3g} deg, |dv| {np.median(d[:, 3]) * 1e3:.3g} m/s")
    assert abs(conv_a - conv_b) <= 0.02
    assert phys >= 0.6


@pytest.mark.parametrize("name,n", [("hmi4k", 2048), ("niris64", 1024)])
def test_fp32_svd_against_fp64_svd_and_the_oracle(name, n):
    """The FP32 path with its SVD in FP32 (svd_fp32 = 1: the two-wave build)
    and in FP64 (svd_fp32 = 0), each against the FP64 oracle, with at most
    two trials per LM iteration (lambda_limits = 1e-5, 6e-3: no trial solves
    with a rounding-decided matrix).  The FP32 synthesis already sets how many
    pixels agree with the oracle at 0.1 G / 0.01 deg / 1 m/s; the FP32 SVD
    lowers that fraction (hmi4k, nl = 6: 0.83 against 0.95; the normal
    matrix's conditioning is why the default keeps the SVD in FP64), keeps the
    converged fraction within a point of the FP64 SVD's, and never fails to
    converge."""
    over = dict(lambda_limits=(1e-5, 6e-3))
    w, _, prof, _ = oracle_map(name, n)
    orc = O.Oracle(w.config(**over), w.wavelengths())
    ref = orc.invert(prof.astype(np.float64), nthreads=8)
    phys, conv, fails = {}, {}, {}
    for svd32 in (0, 1):
        with gpu_inversion(name, precision=ME_FP32, svd_fp32=svd32, **over) as inv:
            out = inv.invert(prof)
        phys[svd32] = within_physical(out["res"], ref["res"]).mean()
        conv[svd32] = np.mean(out["res"][:, 9] <= w.chisq_criteria)
        fails[svd32] = out["stats"].n_svd_fail
    conv_ref = np.mean(ref["res"][:, 9] <= w.chisq_criteria)
    print(f"{name} FP32 path vs FP64 oracle, two-trial cap: within tolerance {phys[1]:.4f} (FP32 SVD) / "
          f"{phys[0]:.4f} (FP64 SVD); converged {conv[1]:.4f} / {conv[0]:.4f} (oracle {conv_ref:.4f})")
    assert fails[1] == 0
    assert phys[1] >= phys[0] - 0.2
    assert abs(conv[1] - conv[0]) <= 0.01


@pytest.mark.parametrize("name", ["hmi4k", "niris64"])
def test_fp32_svd_build_with_fixed_parameters_regularisation_and_restarts(name):
    """The FP32 SVD's two-wave build on the paths the learned and plain
    defaults leave aside there: fixed parameters (the packed system), the
    theta regularisation and num_run = 5 (restarts from the start parameters
    and the best run, Gaussian jumps from run 4 on, which that build keeps in
    the scratch slot and LDS).  Against the FP64 SVD's build and the FP64
    oracle on the same pixels: converged fractions within 3 points."""
    n = 1024
    over = dict(num_run=5, fixed=(0, 0, 0, 0, 0, 1, 0, 0, 1), regul_theta=True)
    w, _, prof, _ = oracle_map(name, n)
    orc = O.Oracle(w.config(**over), w.wavelengths())
    ref = orc.invert(prof.astype(np.float64), nthreads=8)
    conv_ref = np.mean((ref["res"][:, 9] > 0) & (ref["res"][:, 9] <= w.chisq_criteria))
    conv, runs = {}, {}
    for svd32 in (0, 1):
        with gpu_inversion(name, precision=ME_FP32, svd_fp32=svd32, **over) as inv:
            out = inv.invert(prof, per_pixel=True)
        r = out["res"]
        conv[svd32] = np.mean((r[:, 9] > 0) & (r[:, 9] <= w.chisq_criteria))
        runs[svd32] = out["stats"].per_pixel["runs"].max()
        assert out["stats"].n_svd_fail == 0
    print(f"{name} fixed (5, 8) + theta regularisation + num_run 5: converged {conv[1]:.4f} (FP32 SVD) / "
          f"{conv[0]:.4f} (FP64 SVD) / {conv_ref:.4f} (oracle); most runs per pixel {runs[1]} / {runs[0]}")
    assert runs[1] >= 5
    assert abs(conv[1] - conv[0]) <= 0.03
    assert abs(conv[1] - conv_ref) <= 0.03


def test_synthesis_fp64_special_lines(golden):
    """Faddeeva's special lines on the device (Src/FADDEEVA.c:461-468): a = 0
    puts every call on y = 0 (exp(-nu^2), Dawson); B = 0 and v_los = 0 on the
    NIRIS grid, whose sample 13 is lambda0 exactly, put nu = 0 there (erfcx)."""
    lam = np.load(golden / "niris_wavelength.npy")
    w = get("niris64")
    orc = O.Oracle(w.config(), lam)
    par = _random_params(w, 64, seed=5)
    par[:32, 5] = 0.0                       # y = 0 everywhere
    par[32:, 0] = 0.0                       # B = 0 ...
    par[32:, 3] = 0.0                       # ... and v_los = 0: nu = 0 at lambda[13]
    _, c0 = orc.synth_counted(par[0])
    _, c1 = orc.synth_counted(par[40])
    assert c0[6] + c0[7] == 3 * lam.size and c0[6] > 0 and c1[7] == 3
    with Inversion(w.config(), lam) as inv:
        syn, J = inv.synth(par, jac=True)
    ref, Jr = orc.synth(par, jac=True)
    assert np.max(np.abs(syn - ref) / par[:, 7][:, None, None]) <= 1e-12
    jscale = np.abs(Jr).max(axis=(1, 3), keepdims=True) + 1e-300
    assert np.max(np.abs(J - Jr) / jscale) <= 1e-9


def test_context_constants_match_the_oracle(golden):
    """me_ctx_info (INV_INIT's derived constants, Src/INV_INIT.c:121-177)
    against the oracle: the NIRIS grid (H4: ic = 12) and the workload grids."""
    grids = [("niris64", np.load(golden / "niris_wavelength.npy"))]
    grids += [(name, get(name).wavelengths()) for name in ("niris64", "hinode", "hmi4k")]
    for name, lam in grids:
        cfg = get(name).config()
        ref = O.Oracle(cfg, lam).info()
        with Inversion(cfg, lam) as inv:
            got = inv.info()
        assert (got["ic"], got["i0"], got["i1"], got["ncut"]) == (ref["ic"], ref["i0"], ref["i1"], ref["ncut"]), name
        assert tuple(got["vlos"]) == tuple(ref["vlos"]), name
    assert O.Oracle(get("niris64").config(), grids[0][1]).info()["ic"] == 12
