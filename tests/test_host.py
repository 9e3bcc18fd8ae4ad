"""CPU tests of the library's host side: ABI exports, typing, tables, grids,
generator, and error behaviour without a device (this container has none)."""
import ctypes
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest

import crabsaxs_amd as csx
from crabsaxs_amd import _lib
import oracle_py as O

GOLDEN = Path(__file__).resolve().parent / "golden"

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "crabsaxs.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


@pytest.mark.skipif(csx.device_count() > 0, reason="checks the no-device path")
def test_no_device_is_a_clean_error():
    with pytest.raises(csx.NoDevice):
        csx.Context(0)


def test_element_symbols():
    assert csx.Element.from_symbol(" c ") == csx.Element.C
    assert csx.Element.from_symbol("D") == csx.Element.H
    assert csx.Element.from_symbol("Zn") == csx.Element.Zn
    assert csx.Element.from_symbol("Xx") == csx.Element.Other


# formfactor.rs:445-524: electron counts at q = 0.
@pytest.mark.parametrize("residue,name,element,electrons", [
    ("ALA", "CB", csx.Element.C, 9.0), ("GLY", "CA", csx.Element.C, 8.0),
    ("LYS", "NZ", csx.Element.N, 9.0), ("ASP", "OD1", csx.Element.O, 8.5),
    ("DA", "C2'", csx.Element.C, 8.0), ("A", "C2'", csx.Element.C, 7.0),
    ("A", "N6", csx.Element.N, 9.0), ("DG", "N1", csx.Element.N, 8.0),
    ("DT", "C7", csx.Element.C, 9.0),
])
def test_united_atom_electron_counts(residue, name, element, electrons):
    atom = csx.Atom(pos=(0, 0, 0), element=element, residue_name=residue, atom_name=name)
    assert abs(csx.form_factor_for_atom(atom, True, 0.0) - electrons) < 0.02


def test_element_electron_counts_and_decay():
    for el, z in [(csx.Element.C, 6), (csx.Element.Na, 11), (csx.Element.Mg, 12),
                  (csx.Element.Cl, 17), (csx.Element.K, 19), (csx.Element.Ca, 20),
                  (csx.Element.Fe, 26), (csx.Element.Zn, 30), (csx.Element.I, 53)]:
        atom = csx.Atom(pos=(0, 0, 0), element=el, residue_name="HOH", atom_name="X")
        f0 = csx.form_factor_for_atom(atom, False, 0.0)
        assert abs(f0 - z) < 0.02
        assert csx.form_factor_for_atom(atom, False, 0.5) < f0


def test_explicit_hydrogens_not_double_counted():
    h = csx.Atom(pos=(0, 0, 0), element=csx.Element.H, residue_name="ALA", atom_name="HB1")
    assert csx.form_factor_for_atom(h, True, 0.1) == 0.0
    assert csx.form_factor_for_atom(h, False, 0.1) > 0.0


NAMES = [("ALA", "CB"), ("ARG", "NE"), ("ARG", "NH2"), ("ASH", "OD2"), ("GLU", "OE1"),
         ("GLH", "OE2"), ("GLH", "OE1"), ("CYS", "SG"), ("CYX", "SG"), ("HIE", "NE2"),
         ("HIP", "ND1"), ("HID", "CE1"), ("ILE", "CD1"), ("LEU", "CB"), ("LYS", "NZ"),
         ("MSE", "CE"), ("PHE", "CZ"), ("TYR", "OH"), ("TRP", "CH2"), ("TRP", "NE1"),
         ("PRO", "N"), ("PCA", "CG"), ("GLY", "CA"), ("VAL", "CG2"), ("THR", "OG1"),
         (" ser ", " og "), ("DA", "C8"), ("DC", "C8"), ("U", "O2'"), ("DU", "O2'"),
         ("T", "C5M"), ("RC", "N4"), ("G", "N2"), ("A", "OP1"), ("C", "O1P"), ("DT", "C1*"),
         ("UNK", "CA"), ("HOH", "O"), ("ALA", "N"), ("RU", "C6"), ("THY", "N3")]


def test_typing_matches_oracle_restatement():
    lib = _lib.load()
    for res, name in NAMES:
        a = lib.csx_united_atom_group(res.encode(), name.encode())
        b = O.lib().or_united_atom_group(res.encode(), name.encode())
        assert a == b, (res, name, a, b)


def test_form_factor_volume_water_scale_match_oracle_bitwise():
    lib = _lib.load()
    for el in range(15):
        for grp in (-1, 0, 3, 6, 10, 13):
            for imp in (0, 1):
                assert lib.csx_excluded_volume(el, grp, imp) == O.lib().or_excluded_volume(el, grp, imp)
                for q in (0.0, 0.013, 0.37, 1.0):
                    assert lib.csx_form_factor(el, grp, 0.7, imp, q) == \
                        O.lib().or_form_factor(el, grp, 0.7, imp, q)
    for q in (0.0, 0.1, 0.5, 1.3):
        assert csx.water_form_factor(q) == O.lib().or_water_form_factor(q)
        for c1 in (0.95, 1.0, 1.05):
            assert csx.excluded_volume_scale(q, c1) == O.lib().or_excluded_volume_scale(q, c1)
    assert abs(csx.water_form_factor(0.0) - 10.0) < 1e-3  # solvent.rs:774-777


def two_atoms():
    return csx.Structure([csx.Atom(pos=(-1, 0, 0), element=csx.Element.C, residue_name="TST"),
                          csx.Atom(pos=(1, 0, 0), element=csx.Element.O, residue_name="TST",
                                   atom_name="O")])


def test_generator_deterministic_and_shaped():
    a = csx.generate_globule(1200, 16.0, 1.0, seed=7)
    b = csx.generate_globule(1200, 16.0, 1.0, seed=7)
    c = csx.generate_globule(1200, 16.0, 1.0, seed=8)
    assert np.array_equal(a.xyz, b.xyz) and np.array_equal(a.element, b.element)
    assert not np.array_equal(a.xyz, c.xyz)
    assert np.abs(a.xyz.mean(axis=0)).max() < 1e-9
    frac = np.bincount(a.element, minlength=15)[[1, 2, 3, 4]] / 1200
    assert abs(frac[0] - 0.63) < 0.05 and abs(frac[2] - 0.19) < 0.05
    ca = a.xyz[::8]
    step = np.linalg.norm(np.diff(ca, axis=0), axis=1)
    assert np.allclose(step, 3.8, atol=1e-9)
    d = np.linalg.norm(a.xyz[1:8] - a.xyz[0], axis=1)
    assert np.all((d >= 1.2 - 1e-12) & (d <= 1.8 + 1e-12))
    e = csx.generate_globule(2000, 20.0, 2.0, seed=1)
    span = e.xyz.max(axis=0) - e.xyz.min(axis=0)
    assert span[2] > 1.4 * span[0]
    ch = csx.generate_domain_chain(6000, 3, seed=5)
    assert ch.xyz.shape == (6000, 3) and np.isfinite(ch.xyz).all()


def test_shell_sites_are_between_radii():
    s = csx.generate_globule(800, 14.0, 1.0, seed=2)
    sites = csx.generate_shell_sites(s, 3.0, 3.0, 6.0)
    assert len(sites) > 0
    d = np.min(np.linalg.norm(sites[:, None, :] - s.xyz[None, :, :], axis=2), axis=1)
    assert np.all(d > 3.0 - 1e-6) and np.all(d <= 6.0 + 1e-6)


def test_combine_solvent_basis_matches_oracle():
    rng = np.random.default_rng(0)
    q = np.linspace(0, 0.6, 31)
    basis = rng.normal(size=(6, 31)) * 100 + 300
    for c1, c2 in [(1.0, 0.0), (1.04, 0.35), (0.95, -1.0)]:
        a = csx.combine_solvent_basis(basis, q, c1, c2)
        b = O.combine_solvent_basis(basis, q, c1, c2)
        assert np.array_equal(a, b)


def _header_table(name):
    """A numeric array of include/crabsaxs_tables.h (rows for 2-D arrays)."""
    import re
    text = (Path(__file__).resolve().parents[1] / "include" / "crabsaxs_tables.h").read_text()
    m = re.search(name + r"\[[^\]]*\](?:\[[^\]]*\])?\s*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    rows = re.findall(r"\{([^{}]*)\}", body)
    if rows:
        return [[float(v) for v in r.split(",") if v.strip()] for r in rows]
    return [float(v) for v in body.split(",") if v.strip()]


def test_tables_equal_the_reference_sources():
    """include/crabsaxs_tables.h against the numbers parsed from the
    reference's own sources (tests/golden/make_tables_fixture.py): every
    coefficient exactly, and the library's form factors, displaced volumes
    and hydrogen counts evaluated from them."""
    ref = json.loads((GOLDEN / "reference_tables.json").read_text())
    assert ref["elements"] == [e.name for e in csx.Element]
    cm = ref["cromer_mann"]
    assert _header_table("CSX_CM_A") == [r["a"] for r in cm]
    assert _header_table("CSX_CM_B") == [r["b"] for r in cm]
    assert _header_table("CSX_CM_C") == [r["c"] for r in cm]
    ua = ref["united_atom"]
    assert _header_table("CSX_UA_A") == [r["a"] for r in ua]
    assert _header_table("CSX_UA_B") == [r["b"] for r in ua]
    assert _header_table("CSX_UA_C") == [r["c"] for r in ua]
    assert _header_table("CSX_UA_HYDROGENS") == [r["hydrogens"] for r in ua]
    assert _header_table("CSX_EXCLUDED_VOLUME") == ref["excluded_volume"]
    assert _header_table("CSX_VDW_RADIUS") == ref["van_der_waals_radius"]
    lib = _lib.load()
    qs = [0.0, 0.05, 0.3, 0.7, 1.2]

    def gauss(a, b, c, q):
        s2 = (q / (4 * math.pi)) ** 2
        return sum(ak * math.exp(-bk * s2) for ak, bk in zip(a, b)) + c
    for e, r in enumerate(cm):
        for q in qs:
            want = gauss(r["a"], r["b"], r["c"], q)
            assert abs(lib.csx_form_factor(e, -1, 1.0, 0, q) - want) <= 1e-14 * max(abs(want), 1)
        assert lib.csx_excluded_volume(e, -1, 0) == ref["excluded_volume"][e]
    for g, r in enumerate(ua):
        for q in qs:
            want = gauss(r["a"], r["b"], r["c"], q)
            assert abs(lib.csx_form_factor(1, g, 1.0, 1, q) - want) <= 1e-14 * max(abs(want), 1)
        assert lib.csx_excluded_volume(1, g, 1) == pytest.approx(
            ref["excluded_volume"][1] + r["hydrogens"] * ref["excluded_volume"][0], abs=1e-12)
