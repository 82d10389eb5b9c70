"""tis_build_chain: the set-up of any chain from files in the reference's formats.

The reference maps 16SCD residue numbers to nucleotides by hard-coded offsets
(res - 561, or res - 567 from 842 on: fast.cpp:3783-3795, 4295-4305) and
hard-codes 15 stack-face exceptions (fast.cpp:4458-4491).  The generic builder
takes the numbering from the PDB and the exceptions as an argument.  Pins:
the 16SCD inputs renumbered (shifted, or made gap-free with the files mapped
through the reference's own rule) build a system identical, array for array,
to tis_build_16scd's, which is itself pinned bit-exactly against the compiled
reference elsewhere.
"""
import os

import numpy as np
import pytest

import tis_ctypes as T

SRC = T.DATA_16SCD
FACES = [583, 587, 658, 687, 700, 701, 721, 727, 760, 782, 802, 814, 881, 892, 898]


def ref_index(res):
    """The reference's 16SCD numbering (fast.cpp:3783-3795)."""
    return res - 561 if res < 842 else res - 567


def rewrite(tmp, renum, keep=None):
    """Copies the 16SCD inputs into tmp with every residue number r replaced by
    renum(r); keep(r1, r2) may drop HB / stack lines."""
    pdb_out = []
    with open(os.path.join(SRC, "16SCD.pdb")) as f:
        for line in f:
            tok = line.split()
            tok[5] = str(renum(int(tok[5])))
            pdb_out.append(" ".join(tok[:9]) + "\n")
    with open(os.path.join(tmp, "chain.pdb"), "w") as f:
        f.writelines(pdb_out)
    with open(os.path.join(SRC, "16SCD_16SCD_unprocessed_bonds.dat")) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    with open(os.path.join(tmp, "bonds.dat"), "w") as f:
        for tok in lines:
            r1, r2 = int(tok[1]), int(tok[2])
            if keep and not keep(r1, r2):
                continue
            f.write(" ".join([tok[0], str(renum(r1)), str(renum(r2))] + tok[3:]) + "\n")
    with open(os.path.join(SRC, "16SCD_16SCD_unprocessed_stacks.dat")) as f:
        pairs = [ln.split() for ln in f if ln.strip()]
    with open(os.path.join(tmp, "stacks.dat"), "w") as f:
        for a, b in pairs:
            r1, r2 = int(a[1:]), int(b[1:])
            if keep and not keep(r1, r2):
                continue
            f.write(f"{a[0]}{renum(r1)} {b[0]}{renum(r2)}\n")
    return {k: os.path.join(tmp, v) for k, v in
            (("pdb", "chain.pdb"), ("bonds", "bonds.dat"), ("stacks", "stacks.dat"))}


def build(lib, files, faces, **kw):
    return T.build_chain(lib, files["pdb"], files["bonds"], files["stacks"],
                         os.path.join(SRC, "MAXIKEY_Mg_Cl_K.dat"), faces, **kw)


def same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        va, vb = a[k], b[k]
        if isinstance(va, np.ndarray):
            assert np.array_equal(va, vb), k
        else:
            assert va == vb, k


@pytest.fixture(scope="module")
def ref16(lib):
    return T.build_16scd(lib, T=0.596, c_mg=0.005, c_k=0.050, seed=1)


def test_shifted_numbering_builds_the_same_system(lib, ref16, tmp_path):
    files = rewrite(str(tmp_path), lambda r: r + 1000)
    got = build(lib, files, [r + 1000 for r in FACES], T=0.596, c_mg=0.005, c_k=0.050, seed=1)
    same(got[0], ref16[0])
    assert got[1] == ref16[1]


def test_gap_free_numbering_builds_the_same_system(lib, ref16, tmp_path):
    """Residues renumbered 1..346 in chain order; the files' numbers (including
    the 842-847 gap numbers the HB list uses) mapped through the reference's rule."""
    files = rewrite(str(tmp_path), ref_index)
    got = build(lib, files, [ref_index(r) for r in FACES], T=0.596, c_mg=0.005, c_k=0.050, seed=1)
    same(got[0], ref16[0])


def test_face_exceptions_are_an_argument(lib, ref16, tmp_path):
    files = rewrite(str(tmp_path), lambda r: r)
    got, _ = build(lib, files, [], T=0.596, c_mg=0.005, c_k=0.050, seed=1)
    ex0, ex1 = ref16[0]["st_excess"], got["st_excess"]
    faces = [2 * ref_index(r) - 2 for r in FACES]          # 1-based units
    diff = np.nonzero(ex1 != ex0)[0] + 1
    assert sorted(diff.tolist()) == sorted(faces)
    assert np.all(ex1[np.array(faces) - 1] == ex0[np.array(faces) - 1] + 1)


def test_input_errors(lib, tmp_path):
    files = rewrite(str(tmp_path), lambda r: r)
    # a residue number used twice in the PDB
    with open(files["pdb"]) as f:
        lines = f.readlines()
    first = lines[0].split()[5]
    dup = [ln.replace(" 563 ", f" {first} ") for ln in lines]
    bad_pdb = os.path.join(str(tmp_path), "dup.pdb")
    with open(bad_pdb, "w") as f:
        f.writelines(dup)
    with pytest.raises(T.TisError) as e:
        build(lib, dict(files, pdb=bad_pdb), FACES)
    assert e.value.code == -3 and "twice" in str(e.value)
    # an HB naming a residue past the chain's end
    bad_bonds = os.path.join(str(tmp_path), "far.dat")
    with open(bad_bonds, "w") as f:
        f.write("CAN11 563 5000 N6 O4 TER\n")
    with pytest.raises(T.TisError) as e:
        build(lib, dict(files, bonds=bad_bonds), FACES)
    assert e.value.code == -3 and "outside the chain" in str(e.value)
    # a base HB on the last nucleotide needs the 3' phosphate it does not have
    bad_bonds2 = os.path.join(str(tmp_path), "end.dat")
    with open(bad_bonds2, "w") as f:
        f.write("CAN11 563 913 N6 O4 TER\n")
    with pytest.raises(T.TisError) as e:
        build(lib, dict(files, bonds=bad_bonds2), FACES)
    assert e.value.code == -3 and "outside the chain" in str(e.value)
    with pytest.raises(T.TisError) as e:
        build(lib, files, FACES, recentre_bead=5000)
    assert e.value.code == -3 and "recentring bead" in str(e.value)
