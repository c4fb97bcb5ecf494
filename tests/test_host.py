"""The host program's FITS reader/writer and control-file parser (no GPU):
against the reference's own CFITSIO-written wavelength file, against
independently written FITS files of every BITPIX, and against meinv.control on
the reference's example control files."""
import json
import math
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO
from meinv.control import ControlFileError, parse_text, to_config

HOST = REPO / "milne-eddington_amd" / "host" / "meinv"


def _run(*args, check=True):
    if not HOST.exists():
        pytest.fail(f"{HOST} not built (run __graft_entry__.build())")
    r = subprocess.run([str(HOST), *map(str, args)], capture_output=True, text=True)
    if check and r.returncode != 0:
        raise AssertionError(f"meinv {' '.join(map(str, args))} failed: {r.stderr}")
    return r


def _write_fits(path, data, bitpix, bscale=None, bzero=None, extra_cards=0):
    """Independent FITS writer: numpy big-endian data, fixed-format cards."""
    cards = ["SIMPLE  =                    T", f"BITPIX  = {bitpix:20d}", f"NAXIS   = {data.ndim:20d}"]
    for k, n in enumerate(data.shape[::-1]):
        cards.append(f"NAXIS{k + 1:<3d}= {n:20d}")
    if bscale is not None:
        cards.append(f"BSCALE  = {bscale:20.10E}")
    if bzero is not None:
        cards.append(f"BZERO   = {bzero:20.10E}")
    cards += [f"COMMENT filler card {i}" for i in range(extra_cards)]
    cards.append("END")
    hdr = "".join(c.ljust(80) for c in cards)
    hdr += " " * ((2880 - len(hdr) % 2880) % 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    raw = np.ascontiguousarray(data).astype(dt).tobytes()
    raw += b"\0" * ((2880 - len(raw) % 2880) % 2880)
    path.write_bytes(hdr.encode("ascii") + raw)


def _read_fits(path):
    raw = path.read_bytes()
    cards, off = {}, 0
    while True:
        block = raw[off:off + 2880].decode("ascii")
        off += 2880
        done = False
        for i in range(36):
            c = block[80 * i:80 * i + 80]
            if c.startswith("END"):
                done = True
                break
            if c[8:10] == "= ":
                cards[c[:8].strip()] = c[10:].split("/")[0].strip()
        if done:
            break
    shape = [int(cards[f"NAXIS{k}"]) for k in range(int(cards["NAXIS"]), 0, -1)]
    n = int(np.prod(shape))
    assert cards["BITPIX"] == "-64"
    return cards, np.frombuffer(raw[off:off + 8 * n], dtype=">f8").reshape(shape)


def test_reads_the_reference_wavelength_file():
    info = json.loads(_run("--fits-info", GOLDEN / "niriswavelength.fits").stdout)
    assert info["bitpix"] == -64 and info["naxes"] == [34]
    vals = np.array([float(x) for x in _run("--fits-read", GOLDEN / "niriswavelength.fits", 0, 34).stdout.split()])
    np.testing.assert_array_equal(vals, np.load(GOLDEN / "niris_wavelength.npy"))


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_reads_every_bitpix_with_scaling(tmp_path, bitpix):
    rng = np.random.default_rng(abs(bitpix))
    shape = (3, 5, 4, 7)          # [ny][nx][4][nl] in C order: NAXIS1 = 7
    if bitpix == 8:
        raw = rng.integers(0, 255, size=shape)
    elif bitpix > 0:
        raw = rng.integers(-30000, 30000, size=shape)
    else:
        raw = rng.normal(size=shape)
    scaled = bitpix in (16, 32)
    bscale, bzero = (2.5e-4, 1.25) if scaled else (None, None)
    p = tmp_path / f"cube{bitpix}.fits"
    _write_fits(p, raw, bitpix, bscale, bzero, extra_cards=40)   # header spans two blocks
    info = json.loads(_run("--fits-info", p).stdout)
    assert info["bitpix"] == bitpix and info["naxes"] == [7, 4, 5, 3]
    assert info["data_offset"] == 2 * 2880
    first, count = 11, 123
    got = np.array([float(x) for x in _run("--fits-read", p, first, count).stdout.split()])
    ref = raw.reshape(-1)[first:first + count].astype(np.float32 if bitpix == -32 else np.float64)
    ref = ref.astype(np.float64)
    if scaled:
        ref = bzero + bscale * ref
    np.testing.assert_array_equal(got, ref)


def test_writes_a_standard_double_image(tmp_path):
    p = tmp_path / "demo.fits"
    _run("--fits-demo", p)
    raw = p.read_bytes()
    assert len(raw) % 2880 == 0
    cards, data = _read_fits(p)
    assert cards["SIMPLE"] == "T" and cards["NAXIS"] == "3"
    assert (cards["NAXIS1"], cards["NAXIS2"], cards["NAXIS3"]) == ("3", "4", "2")
    np.testing.assert_array_equal(data.reshape(-1), np.arange(24.0))


def test_rejects_files_that_are_not_fits(tmp_path):
    p = tmp_path / "bad.fits"
    p.write_bytes(b"NOT A FITS FILE".ljust(2880))
    r = _run("--fits-info", p, check=False)
    assert r.returncode != 0 and "not a FITS file" in r.stderr
    p.write_bytes(("SIMPLE  =                    T".ljust(80) + "BITPIX  =                  -64".ljust(80)).encode())
    r = _run("--fits-info", p, check=False)
    assert r.returncode != 0


def _python_config(text):
    return to_config(parse_text(text))


def _compare(js, cfg):
    assert js["lambda0"] == cfg.lambda0 and js["geff"] == cfg.geff
    for k in ("num_iter", "num_run", "nfigures", "hmi_ref", "output_fit"):
        assert js[k] == getattr(cfg, k), k
    for k in ("convg_criteria", "chisq_criteria", "damp", "lambda_accept", "lambda_reject", "vcoeffi", "lcoeffi",
              "icriteria", "svd_threshold"):
        assert js[k] == getattr(cfg, k), k
    assert js["weights"] == list(cfg.weights)
    assert js["lambda_limits"] == list(cfg.lambda_limits)
    assert js["bounds"] == [cfg.bounds[p][i] for p in range(9) for i in range(2)]
    assert js["fixed"] == list(cfg.fixed)
    assert js["regul"] == list(cfg.regul)
    assert js["regul_value"] == list(cfg.regul_value)
    assert js["regul_weight"] == list(cfg.regul_weight)
    assert js["step"] == list(cfg.step)


@pytest.mark.parametrize("name", ["niris_example_input", "niris_input", "hinode_input", "hmi_input"])
def test_control_parser_matches_the_python_mirror_on_reference_examples(name):
    path = GOLDEN / "control" / name
    js = json.loads(_run("--dump-config", path).stdout)
    _compare(js, _python_config(path.read_text()))


CUSTOM = """
lines = 6302.5, 2.5
lines = 6301.5, 1.67   ! only the first line is used
data_path = /data/cube.fits   # trailing comment
wavelength_path=/data/wav.fits
num_iter = 40
num_run = 2
weights = 1, 4, 4, 3
sol_box = 10, 20, 5, 9
output_fit = yes
Bounds_Btheta = 0.1, 0.9
Bounds_Dopp = 100, 400
Dopp_fixed = YES
Beta_fixed = Yes
Btheta_Regularization = NO, 0.5, 1e-4
Eta_Regularization = YES, 10.0, 2e-3
Bphi_Regularization = YES, 0.25, 1
HMI_REF = YES
Nfigures = 5
Step = 500, 1, 1, 1, 10, 0.1, 10, 1e4, 0.1
nprofiles = 1000
Cache_inv = YES
unknown_key = 3
num_iter = 45
"""


def test_control_parser_custom_keywords(tmp_path):
    p = tmp_path / "input"
    p.write_text(CUSTOM)
    js = json.loads(_run("--dump-config", p).stdout)
    cfg = _python_config(CUSTOM)
    _compare(js, cfg)
    assert js["num_iter"] == 45 and js["lambda0"] == 6302.5
    assert js["sol_box"] == [10, 20, 5, 9] and js["nprofiles"] == 1000 and js["cache_inv"] == 1
    assert js["data_path"] == "/data/cube.fits" and js["wavelength_path"] == "/data/wav.fits"
    assert js["fixed"] == [0, 0, 0, 0, 1, 0, 0, 0, 1]
    assert js["regul"] == [0, 0, 1, 0, 0, 0, 1, 0, 0]
    assert math.isclose(js["regul_value"][2], 0.25 * math.pi)
    assert math.isclose(js["bounds"][2], 0.1 * math.pi)
    assert js["warnings"] == 2          # the unknown keyword and the redefinition


@pytest.mark.parametrize("bad,msg", [
    ("data_path = a\nwavelength_path = b\n", "required keyword missing"),
    ("lines = 6302.5, 2.5\ndata_path = a\nwavelength_path = b\nlambda_accept = 1.2\n", "larger than 1.5"),
    ("lines = 6302.5, 2.5\ndata_path = a\nwavelength_path = b\nweights = 1, 2, 3\n", "4 floats"),
])
def test_control_parser_errors(tmp_path, bad, msg):
    p = tmp_path / "input"
    p.write_text(bad)
    r = _run("--dump-config", p, check=False)
    assert r.returncode != 0 and msg in r.stderr
    with pytest.raises(ControlFileError):
        _python_config(bad)


def test_learned_mode_options_are_checked(tmp_path):
    """--mlp / --save-mlp only with --learned; training sizes validated (all
    before any device call)."""
    _write_fits(tmp_path / "wav.fits", np.linspace(6172.0, 6174.0, 6), -64)
    p = tmp_path / "input"
    p.write_text(f"lines = 6173.34, 2.5\ndata_path = {tmp_path / 'cube.fits'}\n"
                 f"wavelength_path = {tmp_path / 'wav.fits'}\n")
    r = _run(p, "--mlp", tmp_path / "net.bin", check=False)
    assert r.returncode != 0 and "need --learned" in r.stderr
    r = _run(p, "--learned", "--train-epochs", 0, check=False)
    assert r.returncode != 0 and "--train-epochs E needs E >= 1" in r.stderr
    r = _run(p, "--learned", "--train-profiles", 100, check=False)
    assert r.returncode != 0 and "--train-profiles N needs N >= 4096" in r.stderr
