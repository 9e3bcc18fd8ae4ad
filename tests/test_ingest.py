"""Structure and curve ingest (src/structure.rs, src/io.rs) on the CPU: the
reference's own tests on its fixture files (tests/golden/structures/, copied
from the reference's examples/data) and the edge cases of its selection
rules.  PDB/mmCIF reading in the reference goes through pdbtbx 0.12; the
expected values here are those the reference's tests and code fix."""
import math
from pathlib import Path

import numpy as np
import pytest

import crabsaxs_amd as csx

DATA = Path(__file__).parent / "golden" / "structures"
PDB = DATA / "test_structure.pdb"


def names(s):
    return [(a.chain_id, a.residue_id, a.atom_name) for a in s.atoms]


def test_parses_known_atom_count_and_altloc():  # structure.rs:354-368
    s = csx.Structure.from_pdb_file(PDB)
    assert len(s) == 2
    ca = next(a for a in s.atoms if a.atom_name == "CA")
    assert abs(ca.occupancy - 0.8) < 1e-12 and ca.pos[0] == 0.1


def test_selected_conformer_can_be_treated_as_full_occupancy():  # structure.rs:370-388
    s = csx.Structure.from_pdb_file_with_options(
        PDB, csx.ParseOptions(occupancy_policy=csx.OccupancyPolicy.SelectedConformerFull))
    assert next(a for a in s.atoms if a.atom_name == "CA").occupancy == 1.0
    assert next(a for a in s.atoms if a.atom_name == "N").occupancy == 1.0


def test_applies_hetatm_water_and_model_policies():  # structure.rs:390-426
    assert len(csx.Structure.from_pdb_file(PDB, csx.ParseOptions(include_hetatm=True))) == 2
    with_water = csx.Structure.from_pdb_file(
        PDB, csx.ParseOptions(include_hetatm=True, include_waters=True))
    assert len(with_water) == 3
    assert [a.atom_name for a in with_water.atoms] == ["CA", "N", "O"]
    second = csx.Structure.from_pdb_file(PDB, csx.ParseOptions(model_index=1))
    assert len(second) == 1 and second.atoms[0].pos[0] == 9.0
    with pytest.raises(csx.ParseError, match="model 2 is not present"):
        csx.Structure.from_pdb_file(PDB, csx.ParseOptions(model_index=2))
    models = csx.Structure.from_pdb_file_models_with_options(PDB)  # structure.rs:330-346
    assert [len(m) for m in models] == [2, 1]


def test_parses_first_model_from_mmcif():  # structure.rs:428-433
    s = csx.Structure.from_pdb_file(DATA / "test_structure.cif")
    assert len(s) == 2 and s.atoms[0].chain_id == "A"
    assert names(s) == [("A", 1, "CA"), ("A", 1, "O")]
    assert [a.element for a in s.atoms] == [csx.Element.C, csx.Element.O]
    second = csx.Structure.from_pdb_file(DATA / "test_structure.cif",
                                         csx.ParseOptions(model_index=1))
    assert len(second) == 1 and second.atoms[0].pos[0] == -1.2


def test_malformed_structure_returns_an_error():  # structure.rs:435-438
    with pytest.raises(csx.ParseError):
        csx.Structure.from_pdb_file(DATA / "malformed.pdb")
    with pytest.raises(csx.IoError):
        csx.Structure.from_pdb_file(DATA / "does_not_exist.pdb")


def test_small_fixture_is_two_carbons():  # tests/end_to_end.rs:4-10 input
    s = csx.Structure.from_pdb_file(DATA / "small.pdb")
    assert len(s) == 2 and list(s.element) == [csx.Element.C, csx.Element.C]
    assert np.array_equal(s.xyz, [[0, 0, 0], [2, 0, 0]]) and list(s.occupancy) == [1.0, 1.0]


def line(rec, serial, name, alt, res, chain, seq, x, y, z, occ, el, icode=" "):
    return (f"{rec:<6}{serial:>5} {name:<4}{alt}{res:>3} {chain}{seq:>4}{icode}   "
            f"{x:>8.3f}{y:>8.3f}{z:>8.3f}{occ:>6.2f}{10.0:>6.2f}          {el:>2}")


def pdb(*lines):
    return "\n".join(lines) + "\nEND\n"


def test_altloc_policies_and_same_position_occupancy():
    text = pdb(line("ATOM", 1, "CB", "A", "SER", "A", 5, 1, 0, 0, 0.40, "C"),
               line("ATOM", 2, "CB", "B", "SER", "A", 5, 2, 0, 0, 0.60, "C"),
               line("ATOM", 3, "OG", "A", "SER", "A", 5, 3, 0, 0, 0.50, "O"),
               line("ATOM", 4, "OG", "B", "SER", "A", 5, 3, 0, 0, 0.50, "O"),
               line("ATOM", 5, "CA", " ", "SER", "A", 5, 0, 0, 0, 1.00, "C"))
    hi = csx.Structure.parse(text, 1)
    assert names(hi) == [("A", 5, "CA"), ("A", 5, "CB"), ("A", 5, "OG")]
    by = {a.atom_name: a for a in hi.atoms}
    assert by["CB"].pos[0] == 2 and by["CB"].occupancy == 0.6
    # equal occupancies: the last alternative (Iterator::max_by); identical
    # positions restore min(occ * (count + 1), 1) (structure.rs:303-309)
    assert by["OG"].occupancy == 1.0 and by["CA"].occupancy == 1.0
    first = csx.Structure.parse(text, 1, csx.ParseOptions(altloc_policy=csx.AltlocPolicy.First))
    assert {a.atom_name: a.pos[0] for a in first.atoms}["CB"] == 1
    full = csx.Structure.parse(text, 1, csx.ParseOptions(
        altloc_policy=csx.AltlocPolicy.First,
        occupancy_policy=csx.OccupancyPolicy.SelectedConformerFull))
    assert {a.atom_name: a.occupancy for a in full.atoms} == {"CA": 1.0, "CB": 1.0, "OG": 1.0}
    every = csx.Structure.parse(text, 1, csx.ParseOptions(altloc_policy=csx.AltlocPolicy.All))
    assert len(every) == 5
    assert [a.occupancy for a in every.atoms if a.atom_name == "CB"] == [0.4, 0.6]


def test_residue_filters_chains_elements_and_order():
    text = pdb(line("ATOM", 1, "N", " ", "GLY", "B", 2, 0, 0, 0, 1.0, "N"),
               line("ATOM", 2, "CA", " ", "GLY", "A", 10, 1, 0, 0, 1.0, "C"),
               line("ATOM", 3, "N", " ", "GLY", "A", 3, 2, 0, 0, 1.0, "N"),
               line("HETATM", 4, "C1", " ", "NAG", "A", 400, 3, 0, 0, 1.0, "C"),
               line("HETATM", 5, "ZN", " ", "ZN", "A", 500, 4, 0, 0, 1.0, "ZN"),
               line("HETATM", 6, "O", " ", "HOH", "A", 600, 5, 0, 0, 1.0, "O"),
               line("ATOM", 7, "SE", " ", "MSE", "A", 11, 6, 0, 0, 1.0, "SE"),
               line("ATOM", 8, "CB", " ", "GLY", "A", 10, 7, 0, 0, -0.5, "C"))
    with pytest.raises(csx.UnsupportedElement, match="SE"):
        csx.Structure.parse(text, 1)
    s = csx.Structure.parse(text, 1, csx.ParseOptions(allow_unknown_elements=True))
    # glycans kept by default, other HETATM and water dropped; sorted by
    # (chain, residue number, atom name)
    assert names(s) == [("A", 3, "N"), ("A", 10, "CA"), ("A", 10, "CB"), ("A", 11, "SE"),
                        ("A", 400, "C1"), ("B", 2, "N")]
    assert s.atoms[3].element == csx.Element.Other and s.atoms[2].occupancy == 0.0
    het = csx.Structure.parse(text, 1, csx.ParseOptions(allow_unknown_elements=True,
                                                        include_hetatm=True))
    assert ("A", 500, "ZN") in names(het) and ("A", 600, "O") not in names(het)
    no_gly = csx.Structure.parse(text, 1, csx.ParseOptions(allow_unknown_elements=True,
                                                           include_glycans=False))
    assert ("A", 400, "C1") not in names(no_gly)
    only_b = csx.Structure.parse(text, 1, csx.ParseOptions(chains=["B"]))
    assert names(only_b) == [("B", 2, "N")]
    assert csx.residue_kind("hoh") == csx.ResidueKind.Water
    assert csx.residue_kind(" NAG ") == csx.ResidueKind.Glycan
    assert csx.residue_kind("DA") == csx.ResidueKind.Polymer
    assert csx.residue_kind("LIG") == csx.ResidueKind.Other


def test_element_from_name_and_united_atom_typing():
    # no element column: Element::from_symbol(atom name) (structure.rs:255-258)
    text = pdb(line("ATOM", 1, "CA", " ", "ALA", "A", 1, 0, 0, 0, 1.0, ""),
               line("ATOM", 2, "N", " ", "ALA", "A", 1, 1, 0, 0, 1.0, ""))
    s = csx.Structure.parse(text, 1)
    assert [a.element for a in s.atoms] == [csx.Element.Ca, csx.Element.N]
    typed = csx.Structure.parse(pdb(line("ATOM", 1, "CA", " ", "ALA", "A", 1, 0, 0, 0, 1.0, "C"),
                                    line("ATOM", 2, "N", " ", "ALA", "A", 1, 1, 0, 0, 1.0, "N")), 1)
    ref = csx.Structure([csx.Atom(pos=(0, 0, 0), element=csx.Element.C, atom_name="CA",
                                  residue_name="ALA"),
                         csx.Atom(pos=(1, 0, 0), element=csx.Element.N, atom_name="N",
                                  residue_name="ALA")])
    assert np.array_equal(typed.group, ref.group) and typed.group[0] >= 0


def test_mmcif_tokens():
    text = """data_x
loop_
_atom_site.group_PDB
_atom_site.type_symbol
_atom_site.label_atom_id
_atom_site.label_alt_id
_atom_site.label_comp_id
_atom_site.label_asym_id
_atom_site.label_seq_id
_atom_site.Cartn_x
_atom_site.Cartn_y
_atom_site.Cartn_z
_atom_site.occupancy
_atom_site.auth_seq_id
_atom_site.auth_asym_id
_atom_site.pdbx_PDB_model_num
# a comment
ATOM C "C5'" . DA X 1 1.5 2.5(3) 3.5 1.00 7 B 1
HETATM O O . HOH X . 4.0 0.0 0.0 ? 8 B 1
ATOM N 'N1' . DA X 1 -1.0 0.0 0.0 0.5 7 B 1
"""
    s = csx.Structure.parse(text)
    assert names(s) == [("B", 7, "C5'"), ("B", 7, "N1")]
    assert s.atoms[0].pos == (1.5, 2.5, 3.5) and s.atoms[1].occupancy == 0.5
    w = csx.Structure.parse(text, 2, csx.ParseOptions(include_hetatm=True, include_waters=True))
    assert ("B", 8, "O") in names(w) and w.atoms[-1].occupancy == 1.0
    with pytest.raises(csx.ParseError):
        csx.Structure.parse("data_y\nloop_\n_atom_site.id\n_atom_site.Cartn_x\n1\n", 2)


# ------------------------------------------------------------------ .dat --

def test_parses_validated_three_column_curve():  # io.rs:178-184
    c = csx.ExperimentalCurve.parse_dat("header\n# comment\n0.01 2.0 0.1\n0.02,1.5,0.2\n0.03 0.0 0.0\n")
    assert list(c.q) == [0.01, 0.02] and list(c.sigma) == [0.1, 0.2] and c.has_errors


def test_accepts_two_column_data_and_rejects_bad_rows():  # io.rs:186-201
    c = csx.ExperimentalCurve.parse_dat("0.01 2.0\n0.02 1.0\n")
    assert not c.has_errors and list(c.sigma) == [1.0, 1.0]
    for bad in ("0.01 2.0 0\n", "0.02 2.0 0.1\n0.01 1.0 0.1\n", "-0.01 2.0 0.1\n"):
        with pytest.raises(csx.ParseError):
            csx.ExperimentalCurve.parse_dat(bad)


def test_normalizes_nm_inverse_angstrom_header():  # io.rs:203-209
    c = csx.ExperimentalCurve.parse_dat("# q (1/nm) I sigma\n1.0 2.0 0.1\n2.0 1.0 0.1\n")
    assert list(c.q) == [0.1, 0.2] and c.q_unit() == "angstrom^-1"


def test_dat_rules_and_fixture(tmp_path):
    # q > 1 on the first row without a header: nm^-1 (io.rs:61-65)
    assert list(csx.ExperimentalCurve.parse_dat("1.5 2 0.1\n2.5 1 0.1\n").q) == [1.5 * 0.1, 2.5 * 0.1]
    with pytest.raises(csx.ParseError, match="expected q and intensity on line 2"):
        csx.ExperimentalCurve.parse_dat("0.1 1\n0.2\n")
    with pytest.raises(csx.ParseError, match="invalid numeric value on line 1"):
        csx.ExperimentalCurve.parse_dat("0.1 abc\n")
    with pytest.raises(csx.ParseError, match="no data rows found"):
        csx.ExperimentalCurve.parse_dat("# nothing\n\n")
    with pytest.raises(csx.ParseError, match="finite"):
        csx.ExperimentalCurve.parse_dat("0.1 inf 0.1\n")
    c = csx.ExperimentalCurve.parse_dat("0.1 1 0.1 extra\n0.1 2 0.3\n")  # equal q allowed
    assert list(c.intensity) == [1.0, 2.0]
    f = csx.ExperimentalCurve.from_dat_file(DATA / "test_experimental.dat")
    assert np.allclose(f.q, [0.01, 0.02, 0.03, 0.04, 0.05]) and f.has_errors and len(f) == 5
    out = tmp_path / "c.dat"
    csx.write_curve_dat(out, [0.01, 0.25], [2.0, 0.0])
    assert out.read_text() == "# q (1/A) intensity\n1.00000000e-2 2.00000000e0\n2.50000000e-1 0.00000000e0\n"
    csx.write_curve_csv(tmp_path / "c.csv", [0.01], [123.0])
    assert (tmp_path / "c.csv").read_text() == "q,intensity\n1.00000000e-2,1.23000000e2\n"
    back = csx.ExperimentalCurve.from_dat_file(out)
    assert list(back.q) == [0.01, 0.25] and list(back.intensity) == [2.0, 0.0]
