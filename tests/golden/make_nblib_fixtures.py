"""Writes the nblib known-answer fixtures as JSON data.

Inputs are the reference's nblib test systems (api/nblib/tests/testsystems.cpp:
ParticleLibrary :53-95, charges :97-101, water/methanol molecules :103-145,
Argon state :244-268, SPC-methanol state :313-335); the expected forces are
the reference's own refdata files, read here from the reference tree:
  api/nblib/tests/refdata/NBlibTest_ArgonForcesAreCorrect.xml
  api/nblib/tests/refdata/NBlibTest_SpcMethanolForcesAreCorrect.xml
Both tests run plain-C kernels with CoulombType::Cutoff at rc = 1.0 nm
(api/nblib/tests/nbkernelsystem.cpp:69-85,187-203).

Run from the repository root:  python tests/golden/make_nblib_fixtures.py
"""
import json
import os
import xml.etree.ElementTree as ET

REF = "/root/reference/api/nblib/tests/refdata"
HERE = os.path.dirname(os.path.abspath(__file__))

C6 = {"Ow": 0.0026173456, "H": 0.0, "OMet": 0.0022619536, "CMet": 0.0088755241, "Ar": 0.0062647225}
C12 = {"Ow": 2.634129e-06, "H": 0.0, "OMet": 1.505529e-06, "CMet": 2.0852922e-05, "Ar": 9.847044e-06}


def read_forces(fname):
    root = ET.parse(os.path.join(REF, fname)).getroot()
    out = []
    for vec in root.iter("Vector"):
        out.append([float(r.text) for r in vec.findall("Real")])
    return out


def main():
    argon = {
        "source": "api/nblib/tests/testsystems.cpp:244-268; refdata NBlibTest_ArgonForcesAreCorrect.xml",
        "box": 6.05449,
        "coulomb": "cutoff",
        "rc": 1.0,
        "type_names": ["Ar"],
        "c6": [C6["Ar"]],
        "c12": [C12["Ar"]],
        "x": [[0.794, 1.439, 0.610], [1.397, 0.673, 1.916], [0.659, 1.080, 0.573],
              [1.105, 0.090, 3.431], [1.741, 1.291, 3.432], [1.936, 1.441, 5.873],
              [0.960, 2.246, 1.659], [0.382, 3.023, 2.793], [0.053, 4.857, 4.242],
              [2.655, 5.057, 2.211], [4.114, 0.737, 0.614], [5.977, 5.104, 5.217]],
        "types": [0] * 12,
        "charges": [0.0] * 12,
        "exclusions": [[i] for i in range(12)],
        "forces": read_forces("NBlibTest_ArgonForcesAreCorrect.xml"),
    }
    # Methanol (Me1 CMet +0.176, O2 OMet -0.574, H3 H +0.398) then SPC water
    # (Ow -0.82, H +0.41, H +0.41); type ids follow the order of type names
    # below, which the physics does not depend on.
    names = ["CMet", "OMet", "H", "Ow"]
    spc = {
        "source": "api/nblib/tests/testsystems.cpp:53-145,313-335; refdata NBlibTest_SpcMethanolForcesAreCorrect.xml",
        "box": 3.01,
        "coulomb": "cutoff",
        "rc": 1.0,
        "type_names": names,
        "c6": [C6[n] for n in names],
        "c12": [C12[n] for n in names],
        "x": [[1.970, 1.460, 1.209], [1.978, 1.415, 1.082], [1.905, 1.460, 1.030],
              [1.555, 1.511, 0.703], [1.498, 1.495, 0.784], [1.496, 1.521, 0.623]],
        "types": [0, 1, 2, 3, 2, 2],
        "charges": [0.176, -0.574, 0.398, -0.82, 0.41, 0.41],
        "exclusions": [[0, 1, 2], [1, 0, 2], [2, 0, 1], [3, 4, 5], [4, 3, 5], [5, 3, 4]],
        "forces": read_forces("NBlibTest_SpcMethanolForcesAreCorrect.xml"),
    }
    for name, data in (("nblib_argon.json", argon), ("nblib_spc_methanol.json", spc)):
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(data, fh, indent=1)
            fh.write("\n")


if __name__ == "__main__":
    main()
