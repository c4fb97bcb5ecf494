"""Batch inversion of every FITS cube in a directory.

Same command line, control settings and output names as the reference's
Tools/RunMEINV.py (lines 20-82): ``python -m meinv.batch <data_path> np=<n>``
writes ``input`` in the current directory and leaves ``<data_path>/res/<stem>_res.fits``
per cube.  The reference starts one MPI job of MEINV per file; here one host
process inverts all cubes with one GPU context (the wavelength grid, its
derived constants and the device buffers are set up once).  ``np`` is
accepted for compatibility and has no effect: the cubes go to the one GPU.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HOST = Path(__file__).resolve().parents[1] / "host" / "meinv"

# the control settings of Tools/RunMEINV.py (NIRIS Fe I 15648.5)
SETTINGS = {
    "verbose": 0,
    "num_run": 5,
    "lines": "15648.5, 3.0",
    "wavelength_path": "../niris_example/niriswavelength.fits",
    "weights": "1, 8, 8, 12",
    "Bounds_Bmod": "5, 5000",
    "Bounds_Vlos": "-20, 20",
    "Bounds_Dopp": "50, 140",
    "Bounds_Damp": "0.4, 1.5",
    "Bounds_Eta": "2, 50",
    "Bounds_Beta": "0.1, 0.9",
    "chisq_criteria": 5,
    "LCoeffi": 300,
    "VCoeffi": 300,
}


def write_control(path: Path, settings: dict) -> None:
    with open(path, "w") as fp:
        for k, v in settings.items():
            fp.write(f"{k} = {v}\n")


def run(data_dir, nproc: int = 6, settings: dict | None = None, host: Path = HOST,
        extra_args: tuple = ()) -> list[Path]:
    """Invert every ``*.fits`` in ``data_dir``; returns the result files."""
    path = Path(data_dir)
    if not path.exists():
        raise ValueError("path does not exist")
    files = sorted(path.glob("*.fits"))
    if not files:
        raise ValueError("No data file found")
    outdir = path / "res"
    outdir.mkdir(exist_ok=True)
    cfg = dict(SETTINGS if settings is None else settings)
    cfg["data_path"] = str(files[0])
    cfg["result_path"] = str(outdir / f"{files[0].stem}_res.fits")
    write_control(Path("input"), cfg)
    print(f"nproc = {nproc} (one GPU)")
    for i, f in enumerate(files):
        print(f"Processing: {i + 1}/{len(files)}  {f.name}")
    cmd = [str(host), "input", "--files", *map(str, files), "--result-dir", str(outdir), *extra_args]
    subprocess.run(cmd, check=True)
    return [outdir / f"{f.stem}_res.fits" for f in files]


def main(argv: list[str]) -> int:
    if len(argv) < 3:
        raise ValueError("Usage: python -m meinv.batch <data_path> np=<core_number>")
    path = Path(argv[1].strip())
    print("path = ", path)
    nproc = 6
    if "np=" in argv[2].lower():
        try:
            nproc = int(argv[2].split("=")[1])
        except ValueError as e:
            raise ValueError("Invalid np format") from e
    run(path, nproc)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
