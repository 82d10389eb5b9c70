"""Host set-up of synthetic chains (tis_build_synthetic), on the CPU.

The reference builds only 16SCD; a synthetic chain reuses the pinned 16SCD
pieces (backbone term table, secondary-stack table, HB site table, ion
placement) on a self-avoiding walk.  The checks here hold by construction of
the walk or by physics, and tie the synthetic tables to the 16SCD builder's
(pinned against the reference's own outputs elsewhere):
  - every bond length and bond angle sits at its term's equilibrium value;
  - no two beads other than bonded / 1-3 partners are closer than 3.5 A, and
    no ion overlaps anything (minimum image);
  - the system is neutral;
  - HB candidates have the geometry measured on the built chain, so the FP64
    oracle evaluates each at its full strength at t = 0;
  - secondary-stack parameters equal the 16SCD builder's for the same bases.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle_ctypes as O
import tis_ctypes as T

KIND_TYPE = {"A": 2, "G": 3, "C": 4, "U": 5}     # maxi_key - 1


@pytest.fixture(scope="module")
def big(lib):
    """The large synthetic workload: 1,500 nt, 600 A box, 10 mM MgCl2 + 100 mM KCl."""
    return T.build_synthetic(lib)


@pytest.fixture(scope="module")
def small(lib):
    return T.build_synthetic(lib, n_nt=200, n_hb=30, radius=40.0, box=250.0, c_mg=0.005, c_k=0.05, seed=3)


def excluded_pairs(d):
    ex = set()
    for i, j in d["bond_ij"].reshape(-1, 2):
        ex.add((min(i, j), max(i, j)))
    for i, _, k in d["angle_ijk"].reshape(-1, 3):
        ex.add((min(i, k), max(i, k)))
    return ex


def test_sizes_and_neutrality(big):
    d = big
    assert d["n_beads"] == 3 * 1500 - 1 and d["n_units"] == 2 * 1500 - 1
    assert d["n_particles"] == d["n_beads"] + d["n_mg"] + d["n_cl"] + d["n_k"] == 32914
    assert (d["n_mg"], d["n_k"]) == (1300, 13007)    # int(c * 6.022e-4 * 600^3)
    q = d["type_q"][d["type"]].sum()
    assert q == 0
    assert d["n_hb"] == 400
    assert np.all(d["box"] == 600.0)
    assert d["recentre"] == (d["n_beads"] + 1) // 2 - 1
    kinds = d["unit_kind"].tobytes().decode()
    assert kinds[1::2] == "P" * 1499 and set(kinds[0::2]) <= set("ACGU")


def test_beads_follow_the_16scd_layout(big):
    d = big
    t, c = d["type"], d["cls"]
    kinds = d["unit_kind"].tobytes().decode()
    nt = kinds[0::2]
    n = len(nt)
    s = 3 * np.arange(n)                       # S(n) = 3n - 3, B = S + 1, P(n+1) = S + 2
    assert np.all(t[s] == 1) and np.all(c[s] == 1)
    assert np.all(t[s + 1] == [KIND_TYPE[k] for k in nt]) and np.all(c[s + 1] == 2)
    assert np.all(t[s[:-1] + 2] == 0) and np.all(c[s[:-1] + 2] == 0)
    assert np.all(d["unit_bead0"][0::2] == s) and np.all(d["unit_bead1"][0::2] == s + 1)


def test_bonds_and_angles_at_equilibrium(big):
    d = big
    p = d["pos"].reshape(-1, 3)
    b, kr = d["bond_ij"].reshape(-1, 2), d["bond_kr"].reshape(-1, 2)
    r = np.linalg.norm(p[b[:, 0]] - p[b[:, 1]], axis=1)
    assert np.abs(r - kr[:, 1]).max() < 1e-9
    a, kt = d["angle_ijk"].reshape(-1, 3), d["angle_kt"].reshape(-1, 2)
    u, v = p[a[:, 0]] - p[a[:, 1]], p[a[:, 2]] - p[a[:, 1]]
    th = np.arccos(np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1))
    assert np.abs(th - kt[:, 1]).max() < 1e-7
    # base bonds + S-P + P-S bonds; B-S-P x2, P-S-P, S-P-S angles
    assert d["n_bonds"] == 1500 + 2 * 1499 and d["n_angles"] == 2 * 1499 + 1498 + 1499


def test_walk_is_self_avoiding(big):
    d = big
    nb = d["n_beads"]
    p = d["pos"].reshape(-1, 3)[:nb]
    close = cKDTree(p).query_pairs(3.5 - 1e-9, output_type="ndarray")
    ex = excluded_pairs(d)
    bad = [tuple(x) for x in close if (min(x), max(x)) not in ex]
    assert not bad


def test_ions_do_not_overlap(big):
    d = big
    p = np.mod(d["pos"].reshape(-1, 3), 600.0)
    R = d["type_R"][d["type"]]
    tree = cKDTree(p, boxsize=600.0)
    pairs = tree.query_pairs(2 * R.max(), output_type="ndarray")
    nb = d["n_beads"]
    pairs = pairs[(pairs[:, 0] >= nb) | (pairs[:, 1] >= nb)]
    dv = p[pairs[:, 0]] - p[pairs[:, 1]]
    dv -= 600.0 * np.round(dv / 600.0)
    r = np.linalg.norm(dv, axis=1)
    dmin = R[pairs[:, 0]] + R[pairs[:, 1]] - 0.2 * 1.5852    # add_crowder's overlap test
    assert np.all(r >= dmin)


def geometry(p, b):
    def ang(a, c, e):
        u, v = a - c, e - c
        return np.arccos(np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1))

    def dih(p1, p2, p3, p4):
        b1, b2, b3 = p2 - p1, p3 - p2, p4 - p3
        n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
        return np.arctan2(np.linalg.norm(b2, axis=1) * np.sum(b1 * n2, 1), np.sum(n1 * n2, 1))

    q = [p[b[:, k]] for k in range(6)]
    return np.stack([np.linalg.norm(q[1] - q[4], axis=1), ang(q[0], q[1], q[4]), ang(q[3], q[4], q[1]),
                     dih(q[0], q[1], q[4], q[3]), dih(q[4], q[1], q[0], q[2]), dih(q[1], q[4], q[3], q[5])], 1)


def test_hb_candidates(big):
    d = big
    p = d["pos"].reshape(-1, 3)
    b = d["hb_beads"].reshape(-1, 6)
    g = d["hb_geo"].reshape(-1, 6)
    # base-atom bead triples (S, B, P(n+1)) of two nucleotides >= 4 apart
    assert np.all(b[:, 1] == b[:, 0] + 1) and np.all(b[:, 2] == b[:, 0] + 2)
    assert np.all(b[:, 4] == b[:, 3] + 1) and np.all(b[:, 5] == b[:, 3] + 2)
    i, j = b[:, 0] // 3, b[:, 3] // 3
    assert np.all(j >= i + 4)
    assert np.bincount(np.concatenate([i, j])).max() <= 2
    assert np.all((g[:, 0] >= 5.0) & (g[:, 0] <= 7.0))
    assert np.abs(geometry(p, b) - g).max() < 1e-9
    assert np.array_equal(d["hb_dode"], (j - i > 50).astype(np.int32))
    # sites: m atom-level bonds between distinct base sites of the two nucleotides
    off, idx = d["hb_site_off"], d["hb_site_idx"]
    so = d["site_hb_off"]
    # sites are numbered unit by unit; recover each site's unit and slot
    # from the per-unit site counts of initialize_HB_atoms (P 2, A 6, C 5, G 7, U 5)
    kinds = d["unit_kind"].tobytes().decode()
    counts = np.array([{"P": 2, "A": 6, "C": 5, "G": 7, "U": 5}[k] for k in kinds])
    first = np.concatenate([[0], np.cumsum(counts)])
    assert first[-1] == d["n_hb_sites"]
    site_unit = np.repeat(np.arange(len(kinds)), counts)
    slot = np.arange(d["n_hb_sites"]) - first[site_unit]
    for k in range(d["n_hb"]):
        s = idx[off[k]:off[k + 1]]
        m = len(s) // 2
        gc = {kinds[2 * i[k]], kinds[2 * j[k]]} == {"G", "C"}
        assert m == (3 if gc else 2)
        assert np.all(site_unit[s[0::2]] == 2 * i[k]) and np.all(site_unit[s[1::2]] == 2 * j[k])
        assert np.all(slot[s] >= 2)
        assert len(set(s[0::2])) == m and len(set(s[1::2])) == m
        assert d["hb_E"][k] == pytest.approx(-m * 3.975 * 0.73753, rel=1e-12)
    # the site -> HB lists are the transpose of the HB -> site lists
    pairs_a = sorted((int(s), k) for k in range(d["n_hb"]) for s in idx[off[k]:off[k + 1]])
    pairs_b = sorted((q, int(h)) for q in range(d["n_hb_sites"]) for h in d["site_hb_idx"][so[q]:so[q + 1]])
    assert pairs_a == pairs_b


def test_secondary_stacks_match_16scd_parameters(lib, big):
    d = big
    kinds = d["unit_kind"].tobytes().decode()
    sb = d["st2_beads"].reshape(-1, 7)
    assert np.all(np.diff(sb, axis=1) == 1)            # P S B P S B P, consecutive
    n = (sb[:, 1] + 3) // 3                            # 1-based nucleotide of S
    assert np.all(n >= 2) and np.all(n <= 1498)
    frac = d["n_st2"] / 1497
    assert 0.55 < frac < 0.65
    ref, _ = T.build_16scd(lib, T=0.596, c_mg=0.0, c_k=0.05, seed=1)
    rk = ref["unit_kind"].tobytes().decode()
    table = {}
    for q, row in enumerate(ref["st2_beads"].reshape(-1, 7)):
        u = row[1]
        ui = list(ref["unit_bead0"]).index(u)
        table[(rk[ui], rk[ui + 2])] = (ref["st2_r0"][q], ref["st2_E"][q])
    for q in range(d["n_st2"]):
        key = (kinds[2 * n[q] - 2], kinds[2 * n[q]])
        r0, e = table[key]
        assert d["st2_r0"][q] == r0 and d["st2_E"][q] == e
    ex = d["st_excess"]
    faces = np.zeros(d["n_units"], dtype=np.int64)
    for q in range(d["n_st2"]):
        u = 2 * n[q] - 2                                # 0-based unit of the 5' nucleoside
        faces[u] += 1
        faces[u + 1] += 1
    assert np.array_equal(ex, faces - 1)
    assert d["n_st3"] == 0


def test_velocities_are_maxwellian(big):
    d = big
    v = d["vel"].reshape(-1, 3)
    m = d["type_mass"][d["type"]]
    ke = 0.5 * np.sum(m[:, None] * v * v) / v.size
    assert ke == pytest.approx(0.5 * 0.596, rel=0.02)
    old = d["pos_old"].reshape(-1, 3)
    assert np.abs(d["pos"].reshape(-1, 3) - old - 0.05 * v).max() < 1e-9


def test_seeded_and_reproducible(lib, small):
    again = T.build_synthetic(lib, n_nt=200, n_hb=30, radius=40.0, box=250.0, c_mg=0.005, c_k=0.05, seed=3)
    for k in ("pos", "vel", "hb_beads", "st2_beads", "unit_kind"):
        assert np.array_equal(small[k], again[k]), k
    other = T.build_synthetic(lib, n_nt=200, n_hb=30, radius=40.0, box=250.0, c_mg=0.005, c_k=0.05, seed=4)
    assert not np.array_equal(small["pos"], other["pos"])


def test_oracle_gives_every_hb_its_full_strength_at_t0(lib, small):
    """At the built geometry every candidate's deviation Q is 0, so once no
    site can be over its valence (valences raised to 8 here) the HB energy is
    the sum of the candidate energies."""
    d = dict(small)
    d["site_valence"] = np.full_like(d["site_valence"], 8)
    assert np.bincount(d["hb_site_idx"]).max() <= 8
    h = T.SystemHandle(d)
    o = O.Oracle(h)
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    o.set_stream(O.glibc_stream(5, 4000))
    _, _, e = o.eval(T.TERM_HB)
    e2 = d["hb_E"][d["hb_dode"] == 0].sum()
    e3 = d["hb_E"][d["hb_dode"] == 1].sum()
    assert e["e2_hb"] == pytest.approx(e2, rel=1e-12)
    assert e["e3_hb"] == pytest.approx(e3, rel=1e-12)


def test_bad_arguments(lib):
    with pytest.raises(T.TisError) as e:
        T.build_synthetic(lib, n_nt=4)
    assert e.value.code == -1
    with pytest.raises(T.TisError) as e:
        T.build_synthetic(lib, radius=200.0, box=600.0)
    assert e.value.code == -1
    with pytest.raises(T.TisError) as e:
        T.build_synthetic(lib, n_nt=1500, radius=30.0, box=600.0)
    assert e.value.code == -3 and "radius too small" in str(e.value)
