"""Writes tests/golden/spc27_refdata.json from the reference tree.

Run once in a container that has the reference checkout:
    python tests/golden/make_golden.py [/root/reference]

It extracts data only:
  - the 243-atom water box of src/gromacs/nbnxm/tests/spc27_coords.h
    (coordinates and box),
  - the double-precision reference forces and energies of the 18 files
    src/gromacs/nbnxm/tests/refdata/NbnxmKernelTest_Coulomb{Ewald,EwaldTwin,ReactionField}_
    Vdw{CutCombGeom,CutCombLB,CutCombNone,ForceSwitch,PotSwitch,EwaldCombGeom}.xml,
and records the test options of src/gromacs/nbnxm/tests/kernel_test.cpp that
produced them (cut-off 0.9 nm, Ewald rtol 1e-6, PotShift Coulomb, three atom
types; kernel_test.cpp:402-447 for the per-case options: twin cut-off
rvdw = 0.7 nm, rvdw_switch = rvdw - 0.2 nm, LJ-PME rtol 1e-4 with the
geometric grid rule, reaction field with epsilon_rf = 0).
"""
import json
import os
import re
import sys
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))
COULOMB = ["Ewald", "EwaldTwin", "ReactionField"]
VDW = ["CutCombGeom", "CutCombLB", "CutCombNone", "ForceSwitch", "PotSwitch", "EwaldCombGeom"]
CASES = [f"Coulomb{c}_Vdw{v}" for c in COULOMB for v in VDW]


def parse_coords(path):
    text = open(path).read()
    box = re.search(r"spc27Box\s*=\s*\{\s*\{\s*([-\d.eE+]+)", text)
    body = text[text.index("spc27Coordinates = {"):]
    triples = re.findall(r"\{\s*([-\d.eE+]+)\s*,\s*([-\d.eE+]+)\s*,\s*([-\d.eE+]+)\s*\}", body)
    coords = [[float(a), float(b), float(c)] for a, b, c in triples]
    return float(box.group(1)), coords


def parse_xml(path):
    root = ET.parse(path).getroot()
    out = {}
    for seq in root.findall("Sequence"):
        name = seq.get("Name")
        if name == "Forces":
            forces = []
            for vec in seq.findall("Vector"):
                vals = {r.get("Name"): float(r.text) for r in vec.findall("Real")}
                forces.append([vals["X"], vals["Y"], vals["Z"]])
            out["forces"] = forces
        else:
            out[name] = [float(r.text) for r in seq.findall("Real")]
    for real in root.findall("Real"):
        out[real.get("Name")] = float(real.text)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    testdir = os.path.join(ref, "src", "gromacs", "nbnxm", "tests")
    box, coords = parse_coords(os.path.join(testdir, "spc27_coords.h"))
    assert len(coords) == 243, len(coords)
    data = {
        "source": "src/gromacs/nbnxm/tests/spc27_coords.h and refdata/NbnxmKernelTest_*.xml",
        "options": {
            "rc": 0.9, "rlist": 0.9, "ewald_rtol": 1e-6, "epsilon_r": 1.0,
            "charge_O": -0.8476, "charge_H": 0.4238,
            "sigma_O": 0.316557, "epsilon_O": 0.650194,
            "sigma_H": 0.04, "epsilon_H": 0.192464,
            "ewald_rtol_lj": 1e-4, "epsilon_rf": 0.0, "rvdw_twin": 0.7, "switch_width": 0.2,
            "note": "kernel_test.cpp:104-129,150-271: first third of atoms uncharged, "
                    "hydrogens of the last third carry LJ, water exclusions; "
                    "kernel_test.cpp:402-447: per-case interaction options; "
                    "kernel_test.cpp:577-611: LB-combined parameters for every case but CutCombGeom",
        },
        "box": box,
        "coordinates": coords,
        "cases": {},
    }
    for case in CASES:
        xml = parse_xml(os.path.join(testdir, "refdata", f"NbnxmKernelTest_{case}.xml"))
        assert len(xml["forces"]) == 243
        data["cases"][case] = {
            "forces": xml["forces"],
            "vdw_energy": xml["VdW energy"],
            "coulomb_energy": xml["Coulomb energy"],
        }
    with open(os.path.join(HERE, "spc27_refdata.json"), "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", os.path.join(HERE, "spc27_refdata.json"))


if __name__ == "__main__":
    main()
