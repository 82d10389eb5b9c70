"""The drop-in driver tis_run (replaces main() of fast.cpp and the _DD second
deck) on short runs: the reference's command line and input files, its three
output files with their names and formats, the rolling copies, restart from the
last whole frame of the Coordinates file, and the device observables files.

Formats (fast.cpp:6288-6314): Contacts one line of n_st3 "%+.3f " values per
output step; Coordinates a raw frame of x[0..N], y[0..N], z[0..N] doubles per
output step; Energies "E2_HB E3_HB E2_STACK E3_STACK Rg" per output step.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "tis_run")
DATA = os.path.join(PKG, "data", "16SCD")
ARGS = ["-D0.850", "-E3.975", "-d6.5", "-T0.615", "-c0.005", "-f0.050", "-s1"]
TAG = "D0.850_E3.975_d6.500_T0.615_cM0.0050_f0.050_s1"


def run(cwd, *extra):
    r = subprocess.run([EXE, *ARGS, f"-I{DATA}", "-P32", "-B250", *extra], cwd=cwd, capture_output=True, text=True,
                       timeout=150)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def read_frames(path, n):
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size % (3 * (n + 1)) == 0
    return raw.reshape(-1, 3, n + 1)


def test_outputs_restart_and_observables(cfg1, tmp_path):
    d = cfg1[0]
    n = d["n_particles"]
    run(tmp_path, "-N2000", "-W500", "-O100")
    en = np.loadtxt(tmp_path / f"Energies_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    ct = np.loadtxt(tmp_path / f"Contacts_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    fr = read_frames(tmp_path / f"Coordinates_16SCD_EPS_Mg_Cl_K_{TAG}.dat", n)
    # outputs at klok = 500, 1000, 1500, 2000
    assert en.shape == (4, 5) and ct.shape == (4, d["n_st3"]) and fr.shape[0] == 4
    assert np.all(fr[:, :, 0] == 0.0) and np.isfinite(fr).all()
    # column 4 of Energies is the sum of the Contacts line (E3_STACK), column 5 Rg of the beads
    assert np.allclose(en[:, 3], ct.sum(axis=1), atol=5e-3 * d["n_st3"])
    beads = fr[:, :, 1:d["n_beads"] + 1]
    rg = np.sqrt(((beads - beads.mean(axis=2, keepdims=True)) ** 2).sum(axis=1).mean(axis=1))
    assert np.allclose(en[:, 4], rg, rtol=1e-5)   # "%le": 7 significant digits
    # the run ended on an output step: rolling copies with index 2000 / 2500 = 0
    for stem in ("Contacts", "Coordinates", "Energies"):
        assert (tmp_path / f"{stem}_16SCD_EPS_Mg_Cl_K_{TAG}_0.dat").exists(), stem
    # device observables, one line per output step
    ob = np.loadtxt(tmp_path / f"Observables_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    mg = np.loadtxt(tmp_path / f"MgBound_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    n_p = int(np.sum(np.asarray(d["cls"])[:d["n_beads"]] == 0))
    assert ob.shape == (4, 13) and mg.shape == (4, n_p)
    assert np.all(ob[:, 1] >= 4) and np.all((mg >= 0) & (mg <= 1))
    assert np.allclose(ob[:, 2], en[:, 4], rtol=0.02)   # window-mean Rg near the frame Rg
    # restart: continue from the last frame to klok 3000 (two more frames)
    run(tmp_path, "-N3000", "-W500", "-O100")
    fr2 = read_frames(tmp_path / f"Coordinates_16SCD_EPS_Mg_Cl_K_{TAG}.dat", n)
    en2 = np.loadtxt(tmp_path / f"Energies_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    assert fr2.shape[0] == 6 and en2.shape == (6, 5)
    assert np.array_equal(fr2[:4], fr)
    # the chain continues from the last frame (ions may wrap across the box)
    nb = d["n_beads"]
    assert np.abs(fr2[4][:, 1:nb + 1] - fr2[3][:, 1:nb + 1]).max() < 20.0


def test_partial_frame_is_refused(cfg1, tmp_path):
    n = cfg1[0]["n_particles"]
    with open(tmp_path / f"Coordinates_16SCD_EPS_Mg_Cl_K_{TAG}.dat", "wb") as f:
        f.write(b"\0" * (8 * 3 * (n + 1) + 16))
    r = subprocess.run([EXE, *ARGS, f"-I{DATA}", "-N100"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "partial frame" in r.stderr
