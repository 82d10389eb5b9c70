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


TAG2 = "D0.850_E3.975_d6.500_T0.615_cM0.0050_f0.050_s11"


def test_second_deck(cfg1, tmp_path):
    """-G, the second deck of the _DD driver (DD.cpp:6405-6544): files named
    ..._s1<seed>.dat (DD.cpp:6136-6140), klok restarted at 1, the run continuing
    from deck 1's final state, the same writers.  Output every step (-W1), so a
    deck's frame at klok k holds the state before step k."""
    d = cfg1[0]
    n, nb = d["n_particles"], d["n_beads"]
    run(tmp_path, "-N40", "-W1", "-G")
    f1 = read_frames(tmp_path / f"Coordinates_16SCD_EPS_Mg_Cl_K_{TAG}.dat", n)
    f2 = read_frames(tmp_path / f"Coordinates_16SCD_EPS_Mg_Cl_K_{TAG2}.dat", n)
    e2 = np.loadtxt(tmp_path / f"Energies_16SCD_EPS_Mg_Cl_K_{TAG2}.dat")
    c2 = np.loadtxt(tmp_path / f"Contacts_16SCD_EPS_Mg_Cl_K_{TAG2}.dat")
    # klok restarts at 1: deck 2 writes one frame at each of its klok 1..40
    # (continuing deck 1's klok past -N would write none)
    assert f1.shape[0] == 40 and f2.shape[0] == 40
    assert e2.shape == (40, 5) and c2.shape == (40, d["n_st3"])
    # deck 2's first frame is deck 1's final state: one step after deck 1's last
    # frame (a bead moves well under 0.5 A per step), far from deck 1's start
    one = np.abs(f2[0][:, 1:nb + 1] - f1[-1][:, 1:nb + 1]).max()
    per_step = np.abs(np.diff(f1[:, :, 1:nb + 1], axis=0)).max()
    print("deck 2 start vs deck 1 end", one, "largest bead step of deck 1", per_step)
    assert 0.0 < one <= 2.0 * per_step < 1.0
    assert np.abs(f2[0][:, 1:nb + 1] - f1[0][:, 1:nb + 1]).max() > 2.0 * one
    # the writers of deck 1: E3_STACK is the sum of the Contacts line, Rg of the beads
    assert np.all(f2[:, :, 0] == 0.0) and np.isfinite(f2).all()
    assert np.allclose(e2[:, 3], c2.sum(axis=1), atol=5e-3 * d["n_st3"])
    beads = f2[:, :, 1:nb + 1]
    rg = np.sqrt(((beads - beads.mean(axis=2, keepdims=True)) ** 2).sum(axis=1).mean(axis=1))
    assert np.allclose(e2[:, 4], rg, rtol=1e-5)
    # rolling copies of both decks at the last step (40 / (5 * 1) = 8)
    for t in (TAG, TAG2):
        for stem in ("Contacts", "Coordinates", "Energies"):
            assert (tmp_path / f"{stem}_16SCD_EPS_Mg_Cl_K_{t}_8.dat").exists(), (stem, t)


def test_observables_leave_the_output_files_alone(tmp_path):
    """A run with device observables (-O) writes the Energies and Contacts files of
    a run without them: a sample evaluates the terms at the current state, and the
    files keep the energies of the last step's force call (fast.cpp:6276-6316).
    The two runs differ only by the summation order of the FP64 force atomics."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run(a, "-N1000", "-W500", "-O100")
    run(b, "-N1000", "-W500")
    ea = np.loadtxt(a / f"Energies_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    eb = np.loadtxt(b / f"Energies_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    ca = np.loadtxt(a / f"Contacts_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    cb = np.loadtxt(b / f"Contacts_16SCD_EPS_Mg_Cl_K_{TAG}.dat")
    print("largest Energies difference (relative)", np.abs(ea - eb).max() / np.abs(eb).max())
    assert ea.shape == eb.shape == (2, 5)
    assert np.allclose(ea, eb, rtol=2e-6, atol=0)          # "%le": 7 significant digits
    assert np.allclose(ca, cb, rtol=0, atol=1.5e-3)         # "%+.3f"
