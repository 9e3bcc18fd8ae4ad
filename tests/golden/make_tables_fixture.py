"""Write tests/golden/reference_tables.json: the coefficient tables of the
reference, read as text from its Rust sources.

The tables are the published constants the reference transcribes
(Cromer-Mann and united-atom Gaussians, implicit-hydrogen counts,
displaced volumes, van der Waals radii).  This script parses them from
  src/structure.rs:90-106    Element (the element order)
  src/formfactor.rs:52-145   Cromer-Mann constants and coefficients_for
  src/formfactor.rs:167-278  UnitedAtomGroup, its hydrogen counts and
                             coefficients
  src/solvent.rs:235-253     van_der_waals_radius_for
  src/solvent.rs:406-424     excluded_volume_for
and writes the numbers only, so that include/crabsaxs_tables.h (which the
library, the oracle and the golden-vector script all read) is checked
against a source that does not go through it.

Run in the build container, where the reference checkout lives:
    python tests/golden/make_tables_fixture.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")


def enum_variants(text: str, name: str) -> list[str]:
    body = re.search(r"enum\s+" + name + r"\s*\{(.*?)\}", text, re.S).group(1)
    return [v.strip() for v in body.split(",") if v.strip()]


def numbers(s: str) -> list[float]:
    return [float(v) for v in s.split(",") if v.strip()]


def match_arms(text: str, fn: str) -> dict[str, float]:
    """`Variant => value` arms of the match in function `fn`."""
    body = re.search(r"fn\s+" + fn + r"\b.*?\{(.*?)\n\}", text, re.S).group(1)
    return {m.group(1): float(m.group(2))
            for m in re.finditer(r"Element::(\w+)\s*=>\s*(-?[\d.]+)", body)}


def main() -> None:
    structure = (REF / "src" / "structure.rs").read_text()
    formfactor = (REF / "src" / "formfactor.rs").read_text()
    solvent = (REF / "src" / "solvent.rs").read_text()
    elements = enum_variants(structure, "Element")

    consts = {}
    for m in re.finditer(r"const\s+(\w+):\s*CromerMann\s*=\s*CromerMann\s*\{\s*a:\s*\[([^\]]*)\],"
                         r"\s*b:\s*\[([^\]]*)\],\s*c:\s*(-?[\d.]+),", formfactor):
        consts[m.group(1)] = {"a": numbers(m.group(2)), "b": numbers(m.group(3)),
                              "c": float(m.group(4))}
    body = re.search(r"fn coefficients_for\(element: Element\).*?\{(.*?)\n\}", formfactor,
                     re.S).group(1)
    cromer_mann = {}
    for m in re.finditer(r"Element::(\w+)\s*=>\s*(\w+),", body):
        cromer_mann[m.group(1)] = consts[m.group(2)]
    cromer_mann["Other"] = {"a": [0.0] * 4, "b": [0.0] * 4, "c": 0.0}

    groups = enum_variants(formfactor, "UnitedAtomGroup")
    united = {}
    for m in re.finditer(r"Self::(\w+)\s*=>\s*FiveGaussian\s*\{\s*a:\s*\[([^\]]*)\],"
                         r"\s*b:\s*\[([^\]]*)\],\s*c:\s*(-?[\d.]+),", formfactor):
        united[m.group(1)] = {"a": numbers(m.group(2)), "b": numbers(m.group(3)),
                              "c": float(m.group(4))}
    hbody = re.search(r"fn implicit_hydrogen_count\(self\).*?\{(.*?)\n    \}", formfactor,
                      re.S).group(1)
    for m in re.finditer(r"((?:\|?\s*Self::\w+\s*)+)=>\s*([\d.]+)", hbody):
        for g in re.findall(r"Self::(\w+)", m.group(1)):
            united[g]["hydrogens"] = float(m.group(2))

    out = {
        "source": "crabSAXS src/structure.rs:90-106, src/formfactor.rs:52-278, "
                  "src/solvent.rs:235-253,406-424 (numbers parsed from the text)",
        "elements": elements,
        "cromer_mann": [cromer_mann[e] for e in elements],
        "united_atom_groups": groups,
        "united_atom": [united[g] for g in groups],
        "van_der_waals_radius": [match_arms(solvent, "van_der_waals_radius_for")[e]
                                 for e in elements],
        "excluded_volume": [match_arms(solvent, "excluded_volume_for")[e] for e in elements],
    }
    assert all(len(r["a"]) == 4 and len(r["b"]) == 4 for r in out["cromer_mann"])
    assert all(len(r["a"]) == 5 and len(r["b"]) == 5 and "hydrogens" in r
               for r in out["united_atom"])
    (HERE / "reference_tables.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(elements)} elements, {len(groups)} united-atom groups")


if __name__ == "__main__":
    main()
