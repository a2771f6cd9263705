"""Copies the CHEMKIN-II inputs of gpuFoam's H2 and GRI tutorials
(tutorials/{h2,gri}/chemkin/{chem.inp,therm.dat}, mechanism data) into
tests/golden/chemkin/<name>/ and extracts the SENKIN constant-pressure
ignition results of tutorials/{h2,gri}/chemkin/senk.out -- the initial
conditions, the T(t) trace and the ignition time -- into
tests/golden/chemkin/<name>/senkin.json (the validation data that
tutorials/h2/validation/Allrun plots against).  Run once where the gpuFoam
sources are present:

    python tests/golden/make_chemkin_fixtures.py [/path/to/gpuFoam]
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def senkin(path):
    text = open(path).read()
    t, T = [], []
    for m in re.finditer(r"^ Time \(sec\) =\s*(\S+)\s+T \(K\) =\s*(\S+)", text, re.M):
        t.append(float(m.group(1)))
        T.append(float(m.group(2)))
    p_atm = float(re.search(r"Pressure \(atm\)\s*=\s*(\S+)", text).group(1))
    T0 = float(re.search(r"Temperature \(K\)\s*=\s*(\S+)", text).group(1))
    X = {}
    block = text.split("Mole Fractions:")[1].split("Time Integration")[0]
    for m in re.finditer(r"^\s*(\S+)\s*=\s*(\S+)", block, re.M):
        X[m.group(1)] = float(m.group(2))
    ign = float(re.search(r"Ignition Time \(sec\)\s*=\s*(\S+)", text).group(1))
    crit = float(re.search(r"Temp criteria \(K\)\s*=\s*(\S+)", text).group(1))
    return {"source": "senk.out (CHEMKIN-II SENKIN, constant pressure)",
            "pressure_atm": p_atm, "T0": T0, "mole_fractions": X,
            "ignition_time": ign, "ignition_T": crit, "time": t, "T": T}


def main(ref="/root/reference"):
    for name in ("h2", "gri"):
        src = os.path.join(ref, "tutorials", name, "chemkin")
        dst = os.path.join(HERE, "chemkin", name)
        os.makedirs(dst, exist_ok=True)
        for f in ("chem.inp", "therm.dat"):
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        with open(os.path.join(dst, "senkin.json"), "w") as f:
            json.dump(senkin(os.path.join(src, "senk.out")), f, indent=0)
        print("wrote", dst)


if __name__ == "__main__":
    main(*sys.argv[1:])
