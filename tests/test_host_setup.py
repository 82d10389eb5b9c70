"""Host set-up (tis_build_16scd) against the reference's own outputs.

Pins: the reference's committed ninfo export of its tables and initial state
(analysis/make_ninfo: 16SCD_ninfo.cg.pdb coordinates to 3 dp, hb-dist and
tbs-dist lists), its stored ion/particle counts for 17 concentrations, glibc
rand() itself, and the dump of the compiled reference (ref_cfg1.npz) for
bit-exact positions, velocities, rand() consumption and table checksums.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import tis_ctypes as T
from conftest import GOLDEN


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ninfo():
    return dict(np.load(os.path.join(GOLDEN, "ninfo_16scd.npz")))


@pytest.fixture(scope="module")
def sys_ninfo(lib):
    # the ninfo run: -D0.850 -E3.975 -d6.500 -T0.615 -c0.0000 -f0.050 -s1
    d, draws = T.build_16scd(lib, T=0.615, c_mg=0.0, c_k=0.050, seed=1)
    return d, draws


def test_sizes_match_reference_trace(sys_ninfo):
    d, _ = sys_ninfo
    # analysis/make_ninfo/trace and 16SCD_ninfo.cg.pdb header
    assert d["n_particles"] == 3272
    assert d["n_beads"] == 1037
    assert d["n_units"] == 691
    assert d["n_hb"] == 22637
    assert (d["n_mg"], d["n_cl"], d["n_k"]) == (0, 945, 1290)
    assert d["n_st3"] == 49 and d["n_st2"] == 343
    assert d["n_hb_sites"] == 2758


def test_cg_pdb_coordinates(sys_ninfo, ninfo):
    """Beads (centroids, principal-axis rotation, recentring) match the exported
    .cg.pdb to its 3 decimals.  Its ions come from the exporter's own copy of the
    set-up code (analysis/make_ninfo/16SCD_ninfo.cpp), whose rand() consumption
    differs from the simulator's; the ions are pinned instead bit-exactly against
    the compiled simulator (ref_ninfo_state.npz)."""
    d, draws = sys_ninfo
    xyz = d["pos"].reshape(-1, 3)
    ref = ninfo["cg_xyz"]
    assert ref.shape == xyz.shape
    nb = d["n_beads"]
    assert np.abs(np.round(xyz[:nb], 3) - ref[:nb]).max() <= 1.0001e-3
    sim = dict(np.load(os.path.join(GOLDEN, "ref_ninfo_state.npz")))
    assert np.array_equal(xyz, sim["pos"])
    assert draws == int(sim["rand_after_init"][0])


def test_hb_list_matches_ninfo(sys_ninfo, ninfo):
    d, _ = sys_ninfo
    hb = d["hb_beads"].reshape(-1, 6)
    assert np.array_equal(hb[:, 1] + 1, ninfo["hb_k11"])
    assert np.array_equal(hb[:, 4] + 1, ninfo["hb_k21"])
    assert np.abs(np.round(d["hb_E"], 3) - ninfo["hb_E"]).max() <= 1.0001e-3
    # hb_r is printed with %7.2f from the same geometry (canonical codes use fixed values)
    assert np.abs(d["hb_geo"].reshape(-1, 6)[:, 0] - ninfo["hb_r"]).max() <= 5.0001e-3
    assert np.array_equal(d["hb_dode"], ninfo["hb_tertiary"].astype(np.int32))
    # merged atom-level HBs: the energy is a whole multiple of -hs_D * 0.73753
    units = d["hb_E"] / (-3.975 * 0.73753)
    assert np.allclose(units, np.round(units), atol=1e-9) and units.min() >= 1


def test_tertiary_stacks_match_ninfo(sys_ninfo, ninfo):
    d, _ = sys_ninfo
    b = d["st3_beads"].reshape(-1, 6)
    assert np.array_equal(b[:, 1] + 1, ninfo["tbs_k11"])
    assert np.array_equal(b[:, 4] + 1, ninfo["tbs_k21"])
    assert np.allclose(d["st3_E"], ninfo["tbs_E"], atol=1e-9)
    assert np.abs(d["st3_geo"].reshape(-1, 6)[:, 0] - ninfo["tbs_r"]).max() <= 5.0001e-3
    # |ii|, |jj| = 2 when ST_EXCESS of the face is 1, else 1 (ninfo.cpp:5859-5865); the
    # sign was added by hand from the stacks file: '+' = 3' face (odd unit), '-' = 5' face
    faces = d["st3_face"].reshape(-1, 2)
    flag = np.where(d["st_excess"][faces - 1] == 1, 2, 1) * np.where(faces % 2 == 1, 1, -1)
    assert np.array_equal(flag[:, 0], ninfo["tbs_ii"]) and np.array_equal(flag[:, 1], ninfo["tbs_jj"])


def test_ion_counts_formula(lib):
    counts = dict(np.load(os.path.join(GOLDEN, "ion_counts.npz")))
    for c, n, nmg, nk, ncl in zip(counts["c_mg"], counts["n_total"], counts["n_mg"], counts["n_k"],
                                  counts["n_cl"]):
        d, _ = T.build_16scd(lib, T=0.615, c_mg=float(c), c_k=0.050, seed=1)
        assert (d["n_particles"], d["n_mg"], d["n_k"], d["n_cl"]) == (n, nmg, nk, ncl), c


def test_glibc_stream_matches_libc(lib):
    libc = C.CDLL("libc.so.6")
    for seed in (0, 1, 615, 596, 777, 12345, 2 ** 31 + 7, 2 ** 32 - 1):
        libc.srand(C.c_uint(seed))
        want = np.array([libc.rand() for _ in range(3000)], dtype=np.int32)
        got = T.glibc_stream(lib, seed, 3000)
        assert np.array_equal(got, want), seed
        assert np.array_equal(T.glibc_stream(lib, seed, 1000, skip=2000), want[2000:])


def test_cfg1_bit_exact(cfg1, golden_cfg1):
    d, draws = cfg1
    g = golden_cfg1
    assert d["n_particles"] == int(g["n_particles"][0])
    assert draws == int(g["rand_after_init"][0])
    assert np.array_equal(d["pos"].reshape(-1, 3), g["pos"])
    assert np.array_equal(d["pos_old"].reshape(-1, 3), g["pos_old"])
    assert np.array_equal(d["vel"].reshape(-1, 3), g["vel"])
    assert np.array_equal(d["type"] + 1, g["maxi_key"])
    assert np.array_equal(d["cls"], g["part_key"])


def test_cfg1_table_checksums(cfg1, golden_cfg1):
    """The reference's tables, re-derived from our 0-based system, hash the same."""
    d, _ = cfg1
    hb = d["hb_beads"].reshape(-1, 6) + 1
    geo = d["hb_geo"].reshape(-1, 6)
    so, si = d["hb_site_off"], d["hb_site_idx"]
    ho, hi = d["site_hb_off"], d["site_hb_idx"]
    n3 = d["n_st3"]
    st_i = np.concatenate([np.zeros(n3, np.int32), np.zeros(d["n_st2"], np.int32)])
    # st_i / st_j are unit indices; recover them from the bead tables
    u0 = d["unit_bead0"]
    unit_of_bead0 = {int(b): u + 1 for u, b in enumerate(u0)}
    b3 = d["st3_beads"].reshape(-1, 6)
    st_i[:n3] = [unit_of_bead0[int(b)] for b in b3[:, 0]]
    st_j = np.array([unit_of_bead0[int(b)] for b in b3[:, 3]], np.int32)
    b2 = d["st2_beads"].reshape(-1, 7)
    st_i[n3:] = [unit_of_bead0[int(b)] for b in b2[:, 1]]
    st_r = np.concatenate([d["st3_geo"].reshape(-1, 6)[:, 0], d["st2_r0"]])
    st_E = np.concatenate([d["st3_E"], d["st2_E"]])
    faces = d["st3_face"].reshape(-1, 2)
    f0 = np.concatenate([faces[:, 0], st_i[n3:]]).astype(np.int32)
    f1 = np.concatenate([faces[:, 1], st_i[n3:] + 1]).astype(np.int32)
    a1 = np.where(d["unit_bead1"] >= 0, d["unit_bead1"] + 1, 0).astype(np.int32)
    ours = [hb[:, 0], hb[:, 1], hb[:, 2], hb[:, 3], hb[:, 4], hb[:, 5], d["hb_E"], geo[:, 0], geo[:, 1],
            geo[:, 2], geo[:, 3], geo[:, 4], geo[:, 5], d["hb_dode"], d["site_valence"],
            np.diff(so).astype(np.int32), si + 1, np.diff(ho).astype(np.int32), hi + 1, st_i, st_j, st_r,
            d["st3_geo"].reshape(-1, 6)[:, 1], d["st3_geo"].reshape(-1, 6)[:, 2],
            d["st3_geo"].reshape(-1, 6)[:, 3], d["st3_geo"].reshape(-1, 6)[:, 4],
            d["st3_geo"].reshape(-1, 6)[:, 5], st_E, f0, f1, d["st_excess"], u0 + 1, a1,
            d["unit_kind"].astype(np.int32)]
    want = list(golden_cfg1["topo_sha"])
    names = ["hb_k10", "hb_k11", "hb_k12", "hb_k20", "hb_k21", "hb_k22", "hb_E", "hb_r", "hb_theta1",
             "hb_theta2", "hb_psi", "hb_psi1", "hb_psi2", "hb_dode", "VALENCE", "ATOM_HB_n", "ATOM_HB_v",
             "HB_ATOM_n", "HB_ATOM_v", "st_i", "st_j", "st_r", "st_theta1", "st_theta2", "st_psi",
             "st_psi1", "st_psi2", "st_E", "ATOM_ST0", "ATOM_ST1", "ST_EXCESS", "atom_key0",
             "atom_key1", "amino_key"]
    for name, arr, w in zip(names, ours, want):
        dt = np.float64 if arr.dtype.kind == "f" else np.int32
        assert sha(np.asarray(arr, dtype=dt)) == w, name


def test_bad_inputs_report_errors(lib, tmp_path):
    with pytest.raises(T.TisError) as e:
        T.build_16scd(lib, data_dir=str(tmp_path))
    assert e.value.code == -3          # TIS_EIO
    with pytest.raises(T.TisError) as e:
        T.build_16scd(lib, T=-1.0)
    assert e.value.code == -1          # TIS_EINVAL
