"""Regenerates the golden fixtures in this folder (run in the build container only).

Sources, all read where they lie:
  * the reference's own committed outputs of its set-up code
    (analysis/make_ninfo/16SCD_ninfo.cg.pdb, hb-dist.edit.ninfo, tbs-dist.edit.ninfo,
    analysis/Analysis_storedvalues_n*.txt) -> ninfo_16scd.npz, ion_counts.npz;
  * the reference translation unit compiled by oracle/Makefile ("make -C oracle ref")
    and driven by oracle/ref_harness.cpp -> ref_cfg1.npz (5 mM MgCl2, 50 mM KCl,
    350 A, T = 0.596, seed 1: topology checksums, initial state, per-term forces,
    pair-list checksums, HB statuses, energies, a 200-step trajectory with the glibc
    stream srand(777)) and ref_ninfo_state.npz (the c = 0 set-up of the ninfo run).

Usage:  python tests/golden/make_fixtures.py
"""
import ast
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
REF = "/root/reference"
sys.path.insert(0, HERE)
from recfile import read_records  # noqa: E402


def sha(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def sorted_pairs(flat):
    p = flat.reshape(-1, 2).astype(np.int64) - 1
    p.sort(axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))].astype(np.int32)


def run_harness(args, out):
    exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True,
                   stdout=subprocess.DEVNULL)
    subprocess.run([exe] + args + ["-O" + out], check=True, cwd=os.path.join(REF, "src"))
    return read_records(out)


def xyz(r, p):
    return np.stack([r[p + "_x"], r[p + "_y"], r[p + "_z"]], 1)


def ninfo_fixtures():
    mk = os.path.join(REF, "analysis", "make_ninfo")
    hb = []
    with open(os.path.join(mk, "hb-dist.edit.ninfo")) as fh:
        for line in fh:
            t = line.split()
            if not t or t[0] != "hb-dist":
                continue
            # hb-dist id 1 1 k11 k21 k11 k21 E r 5.0 T|S nHB atoms...
            hb.append((int(t[4]), int(t[5]), float(t[8]), float(t[9]), 1 if t[11] == "T" else 0, int(t[12])))
    hb = np.array(hb, dtype=np.float64)
    tbs = []
    with open(os.path.join(mk, "tbs-dist.edit.ninfo")) as fh:
        for line in fh:
            t = line.split()
            if not t or t[0] != "tbs-dist":
                continue
            tbs.append((int(t[4]), int(t[5]), float(t[8]), float(t[9]), int(t[11]), int(t[12])))
    tbs = np.array(tbs, dtype=np.float64)
    pdb = []
    with open(os.path.join(mk, "16SCD_ninfo.cg.pdb")) as fh:
        for line in fh:
            if line.startswith("ATOM"):
                pdb.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
    pdb = np.array(pdb)
    np.savez_compressed(os.path.join(HERE, "ninfo_16scd.npz"),
                        hb_k11=hb[:, 0].astype(np.int32), hb_k21=hb[:, 1].astype(np.int32),
                        hb_E=hb[:, 2], hb_r=hb[:, 3], hb_tertiary=hb[:, 4].astype(np.int8),
                        hb_n_atom=hb[:, 5].astype(np.int16),
                        tbs_k11=tbs[:, 0].astype(np.int32), tbs_k21=tbs[:, 1].astype(np.int32),
                        tbs_E=tbs[:, 2], tbs_r=tbs[:, 3], tbs_ii=tbs[:, 4].astype(np.int8),
                        tbs_jj=tbs[:, 5].astype(np.int8), cg_xyz=pdb)
    an = os.path.join(REF, "analysis")
    counts = {}
    for key in ("nmps", "nMgs", "nKs", "nCls"):
        with open(os.path.join(an, f"Analysis_storedvalues_{key}.txt")) as fh:
            counts[key] = ast.literal_eval(fh.read())
    conc = np.array(sorted(counts["nmps"].keys()))
    np.savez_compressed(os.path.join(HERE, "ion_counts.npz"), c_mg=conc,
                        n_total=np.array([counts["nmps"][c] for c in conc]),
                        n_mg=np.array([counts["nMgs"][c] for c in conc]),
                        n_k=np.array([counts["nKs"][c] for c in conc]),
                        n_cl=np.array([counts["nCls"][c] for c in conc]))


def ref_fixtures(tmp):
    r = run_harness(["-D0.850", "-E3.975", "-d6.5", "-T0.596", "-c0.005", "-f0.050", "-s1",
                     "-N200", "-K50", "-Q777", "-R12345"], os.path.join(tmp, "ref_cfg1.bin"))
    sr, lr = sorted_pairs(r["list_SR"]), sorted_pairs(r["list_LR"])
    out = dict(
        n_particles=r["NT_atom"], rand_after_init=r["rand_after_init"],
        pos=xyz(r, "pos"), pos_old=xyz(r, "old"), vel=xyz(r, "vel"),
        maxi_key=r["maxi_key"].astype(np.int8), part_key=r["part_key"].astype(np.int8),
        topo_sha=np.array([sha(r[k]) for k in (
            "hb_k10", "hb_k11", "hb_k12", "hb_k20", "hb_k21", "hb_k22", "hb_E", "hb_r", "hb_theta1",
            "hb_theta2", "hb_psi", "hb_psi1", "hb_psi2", "hb_dode", "VALENCE", "ATOM_HB_n", "ATOM_HB_v",
            "HB_ATOM_n", "HB_ATOM_v", "st_i", "st_j", "st_r", "st_theta1", "st_theta2", "st_psi",
            "st_psi1", "st_psi2", "st_E", "ATOM_ST0", "ATOM_ST1", "ST_EXCESS", "atom_key0",
            "atom_key1", "amino_key")]),
        sr_pairs=sr, lr_count=np.array([len(lr)]), lr_sha=np.array([sha(lr)]),
        hb_status0=r["hb_status0"].astype(np.int8), hb_excess0=r["HB_EXCESS0"].astype(np.int16),
        F_stack=xyz(r, "F_stack"), F_hb=xyz(r, "F_hb"), F_poly=xyz(r, "F_poly"),
        F_SR=xyz(r, "F_SR"), F_LR=xyz(r, "F_LR"),
        st_energy=r["st_energy"], st_status=r["st_status"].astype(np.int8),
        hb_energy_nz_idx=np.nonzero(r["hb_energy"])[0].astype(np.int32),
        hb_energy_nz=r["hb_energy"][np.nonzero(r["hb_energy"])[0]],
        hb_rng_seed=r["hb_rng_seed"], hb_rand_calls=r["hb_rand_calls"],
        energies=np.array([r[k][0] for k in ("E2_HB", "E3_HB", "E2_STACK", "E3_STACK", "E_WCA_SR",
                                             "E_COUL_SR", "E_COUL_LR", "Rg0")]),
        traj_rng_seed=r["traj_rng_seed"], traj_draws=r["traj_draws"], traj_rebuilt=r["traj_rebuilt"],
        traj_progress=r["traj_progress"],
        traj_E=np.stack([r["traj_E2_HB"], r["traj_E3_HB"], r["traj_E2_STACK"], r["traj_E3_STACK"]], 1),
    )
    for s in (50, 100, 200):
        out[f"t{s}_pos"] = xyz(r, f"t{s}_pos")
        out[f"t{s}_vel"] = xyz(r, f"t{s}_vel")
    np.savez_compressed(os.path.join(HERE, "ref_cfg1.npz"), **out)
    # The set-up of the ninfo run (c = 0, T = 0.615), for the .cg.pdb comparison.
    r0 = run_harness(["-D0.850", "-E3.975", "-d6.500", "-T0.615", "-c0.0000", "-f0.050", "-s1"],
                     os.path.join(tmp, "ref_ninfo.bin"))
    np.savez_compressed(os.path.join(HERE, "ref_ninfo_state.npz"), pos=xyz(r0, "pos"),
                        sr_count=np.array([len(r0["list_SR"]) // 2]),
                        lr_count=np.array([len(r0["list_LR"]) // 2]),
                        rand_after_init=r0["rand_after_init"])


if __name__ == "__main__":
    import tempfile

    ninfo_fixtures()
    with tempfile.TemporaryDirectory(prefix="tis_fixtures_") as tmp:   # the harness's record files
        ref_fixtures(tmp)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
