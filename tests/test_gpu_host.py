"""The MEINV-style host program end to end on the GPU: control file + FITS
cube in, res.fits / fit.fits out, equal to the library called directly on the
same solution box; chunked runs, Cache_prof, and an interrupted Cache_inv run
resumed to the same result."""
import subprocess

import numpy as np
import pytest

from conftest import REPO
from gpu_util import oracle_map
from meinv import Inversion
from meinv.control import parse_text, to_config
from test_host import _read_fits, _write_fits

pytestmark = pytest.mark.gpu

HOST = REPO / "milne-eddington_amd" / "host" / "meinv"
NX, NY = 37, 23
BOX = (3, 33, 2, 20)        # x0, x1, y0, y1 (inclusive), as sol_box


def _setup(tmp_path, **keys):
    w, orc, prof, truth = oracle_map("niris64", NX * NY)
    cube = prof.reshape(NY, NX, 4, w.nl).astype(np.float32)
    _write_fits(tmp_path / "cube.fits", cube, -32)
    _write_fits(tmp_path / "wav.fits", w.wavelengths(), -64)
    kv = {
        "lines": f"{w.lambda0}, {w.geff}",
        "data_path": tmp_path / "cube.fits",
        "wavelength_path": tmp_path / "wav.fits",
        "result_path": tmp_path / "res.fits",
        "fit_path": tmp_path / "fit.fits",
        "cache_path": tmp_path / "cache",
        "num_run": 1, "chisq_criteria": 1.0, "Icriteria": 0.1,
        "Bounds_Dopp": "100, 400", "Bounds_Damp": "0, 1.5", "Bounds_Eta": "1, 50", "Bounds_Beta": "0.1, 0.9",
        "Btheta_Regularization": "NO",
        "sol_box": ", ".join(map(str, BOX)),
    }
    kv.update(keys)
    text = "".join(f"{k} = {v}\n" for k, v in kv.items())
    (tmp_path / "input").write_text(text)
    x0, x1, y0, y1 = BOX
    box = cube[y0:y1 + 1, x0:x1 + 1].reshape(-1, 4, w.nl).astype(np.float64)
    with Inversion(to_config(parse_text(text)), w.wavelengths()) as inv:
        ref = inv.invert(box, fit=True)
    return w, ref


def _run(tmp_path, *args):
    r = subprocess.run([str(HOST), str(tmp_path / "input"), *map(str, args)], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_host_run_matches_the_library(tmp_path):
    w, ref = _setup(tmp_path, nprofiles=50, output_fit="YES")
    out = _run(tmp_path, "--batch", 100)
    nxb, nyb = BOX[1] - BOX[0] + 1, BOX[3] - BOX[2] + 1
    assert f"box {nxb}x{nyb}" in out
    cards, res = _read_fits(tmp_path / "res.fits")
    assert res.shape == (nyb, nxb, 10)
    np.testing.assert_array_equal(res.reshape(-1, 10), ref["res"])
    _, fit = _read_fits(tmp_path / "fit.fits")
    assert fit.shape == (nyb, nxb, 4, w.nl)
    np.testing.assert_array_equal(fit.reshape(-1, 4, w.nl), ref["fit"])
    assert np.mean(ref["res"][:, 9] <= 1.0) > 0.5


def test_host_cache_prof_reads_the_box_at_once(tmp_path):
    w, ref = _setup(tmp_path, Cache_prof="YES", output_fit="YES")
    out = _run(tmp_path)
    assert "in 1 chunks" in out
    _, res = _read_fits(tmp_path / "res.fits")
    np.testing.assert_array_equal(res.reshape(-1, 10), ref["res"])
    assert not (tmp_path / "fit.fits").exists()   # Cache_prof turns output_fit off, as MEINV does


def test_host_interrupted_run_resumes_from_the_cache(tmp_path):
    w, ref = _setup(tmp_path, nprofiles=50, Cache_inv="YES")
    out = _run(tmp_path, "--max-chunks", 5, "--batch", 100)
    assert "in 5 chunks" in out
    raw = (tmp_path / "cache").read_bytes()
    header = np.frombuffer(raw[4:16], dtype="<i4")
    flags = np.frombuffer(raw[16:], dtype="<i4")
    nxb, nyb = BOX[1] - BOX[0] + 1, BOX[3] - BOX[2] + 1
    assert raw[:4] == b"cach" and list(header) == [nxb, nyb, nxb * nyb // 50 + 1]
    assert list(flags[:5]) == [1] * 5 and not flags[5:].any()
    _, part = _read_fits(tmp_path / "res.fits")
    part = part.reshape(-1, 10)
    np.testing.assert_array_equal(part[:250], ref["res"][:250])
    assert not part[250:].any()
    out = _run(tmp_path, "--batch", 100)
    assert "(resumed)" in out
    _, res = _read_fits(tmp_path / "res.fits")
    np.testing.assert_array_equal(res.reshape(-1, 10), ref["res"])
    flags = np.frombuffer((tmp_path / "cache").read_bytes()[16:], dtype="<i4")
    assert flags[:-1].all()


def test_host_resume_with_restarts_equals_an_uninterrupted_run(tmp_path):
    """num_run = 2 (Gaussian jumps at irun >= 4): the host keys each pixel's
    restart draws by its index in the box (me_ctx_set_pixel_base), so a run
    stopped after 5 chunks and resumed with other batch boundaries ends with
    the records of one uninterrupted batch."""
    w, ref = _setup(tmp_path, nprofiles=50, Cache_inv="YES", num_run=2)
    _run(tmp_path, "--max-chunks", 5, "--batch", 100)
    out = _run(tmp_path, "--batch", 150)
    assert "(resumed)" in out
    _, res = _read_fits(tmp_path / "res.fits")
    np.testing.assert_array_equal(res.reshape(-1, 10), ref["res"])


def test_batch_driver_inverts_every_cube_with_one_context(tmp_path, monkeypatch):
    """meinv.batch (Tools/RunMEINV.py's command line): two cubes in a
    directory, one host process, res/<stem>_res.fits each equal to the library
    on that cube."""
    from meinv import batch
    w, orc, prof, truth = oracle_map("niris64", 2 * 300)
    data = tmp_path / "data"
    data.mkdir()
    cubes = []
    for k in range(2):
        cube = prof[300 * k:300 * (k + 1)].reshape(15, 20, 4, w.nl).astype(np.float32)
        _write_fits(data / f"scan{k}.fits", cube, -32)
        cubes.append(cube)
    _write_fits(tmp_path / "wav.fits", w.wavelengths(), -64)
    settings = dict(batch.SETTINGS, wavelength_path=str(tmp_path / "wav.fits"), lines=f"{w.lambda0}, {w.geff}",
                    num_run=1, Icriteria=0.1, chisq_criteria=1.0, nprofiles=64)
    monkeypatch.chdir(tmp_path)
    outs = batch.run(data, settings=settings)
    text = (tmp_path / "input").read_text()
    cfg = to_config(parse_text(text))
    for k, out in enumerate(outs):
        assert out == data / "res" / f"scan{k}_res.fits"
        _, res = _read_fits(out)
        with Inversion(cfg, w.wavelengths()) as inv:
            ref = inv.invert(cubes[k].reshape(-1, 4, w.nl).astype(np.float64))
        np.testing.assert_array_equal(res.reshape(-1, 10), ref["res"])


def _read_net(path):
    raw = path.read_bytes()
    assert raw[:8] == b"MEINVMLP"
    npar = int(np.frombuffer(raw[8:16], dtype="<i8")[0])
    ranges = np.frombuffer(raw[16:16 + 144], dtype="<f8").reshape(9, 2)
    params = np.frombuffer(raw[160:160 + 4 * npar], dtype="<f4")
    assert len(raw) == 160 + 4 * npar
    return params, ranges


def test_host_learned_mode_trains_saves_and_reuses_the_network(tmp_path):
    """--learned: train on the control file's bounds and save the network,
    then a second run reads it back; both runs and the library (same network
    imported, learned mode) give the same records."""
    from meinv import ME_INIT_LEARNED
    w, _ = _setup(tmp_path, nprofiles=100)
    out = _run(tmp_path, "--learned", "--train-profiles", 65536, "--train-epochs", 2, "--save-mlp", tmp_path / "net.bin")
    assert "network trained on 65536 profiles" in out
    _, res1 = _read_fits(tmp_path / "res.fits")
    params, ranges = _read_net(tmp_path / "net.bin")
    assert ranges[1, 1] == pytest.approx(np.pi) and ranges[7].tolist() == [0.5, 1.2]
    (tmp_path / "res.fits").unlink()
    out = _run(tmp_path, "--learned", "--mlp", tmp_path / "net.bin")
    assert "network read from" in out
    _, res2 = _read_fits(tmp_path / "res.fits")
    np.testing.assert_array_equal(res2, res1)
    text = (tmp_path / "input").read_text()
    cube = _read_fits_cube(tmp_path / "cube.fits")
    x0, x1, y0, y1 = BOX
    box = cube[y0:y1 + 1, x0:x1 + 1].reshape(-1, 4, w.nl).astype(np.float64)
    cfg = to_config(parse_text(text))
    cfg.init_mode = ME_INIT_LEARNED
    with Inversion(cfg, w.wavelengths()) as inv:
        inv.mlp_import(params, ranges)
        ref = inv.invert(box)
    np.testing.assert_array_equal(res2.reshape(-1, 10), ref["res"])
    assert np.mean(ref["res"][:, 9] <= 1.0) > 0.5


def _read_fits_cube(path):
    """float32 cube written by _write_fits (BITPIX -32)."""
    raw = path.read_bytes()
    hdr = raw[:2880].decode("ascii")
    cards = {hdr[80 * i:80 * i + 8].strip(): hdr[80 * i + 10:80 * i + 80].split("/")[0].strip() for i in range(36)}
    shape = [int(cards[f"NAXIS{k}"]) for k in range(int(cards["NAXIS"]), 0, -1)]
    n = int(np.prod(shape))
    return np.frombuffer(raw[2880:2880 + 4 * n], dtype=">f4").reshape(shape)
