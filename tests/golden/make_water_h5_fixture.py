"""Build the water HDF5 fixtures from the reference's documentation data.

* tests/golden/water.h5: the file richmol wrote with h5py in
  docs/source/notebooks/rotations_quickstart.ipynb (dip, pol and h0 of water,
  J <= 5, D2 symmetry) -- a data file, copied as it is;
* tests/golden/water_h5_notebook.json: what that notebook printed for the
  same tensors before storing them (its saved cell outputs): the rotational
  energies per (J, symmetry, index), the entries of dip.tomat(form='full',
  cart='x'), pol.tomat(form='full', cart='xz') and h0.tomat(form='full',
  cart='0') it lists, the assignment (J, symmetry, m, k) of the first 20
  full-form states, and the matrix dimensions of the m-filtered
  Hamiltonians.

    python tests/golden/make_water_h5_fixture.py /root/reference
"""
import ast
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _outputs(nb):
    cells = [c for c in json.load(open(nb))["cells"] if c["cell_type"] == "code"]
    return ["".join("".join(o.get("text", [])) for o in c.get("outputs", [])) for c in cells]


_ENTRY = re.compile(r"^\s*\((\d+), (\d+)\)\t\((.*)\)\s*$")


def _entries(text):
    out = []
    for line in text.splitlines():
        m = _ENTRY.match(line)
        if m:
            z = complex(m.group(3))
            out.append([int(m.group(1)), int(m.group(2)), z.real, z.imag])
    return out


def main(ref_root):
    nbdir = os.path.join(ref_root, "docs", "source", "notebooks")
    outs = _outputs(os.path.join(nbdir, "rotations_quickstart.ipynb"))
    fix = {"source": "docs/source/notebooks/rotations_quickstart.ipynb (saved outputs)"}
    energy_txt = next(t for t in outs if t.startswith("J  sym #"))
    fix["energies"] = []
    for line in energy_txt.splitlines()[1:]:
        w = line.split()
        if len(w) >= 4:
            fix["energies"].append([int(w[0]), w[1], int(w[2]), float(w[3])])
    mat_txt = next(t for t in outs if "dipole X:" in t)
    dip_txt, pol_txt = mat_txt.split("polarizability XZ:")[0], mat_txt.split("polarizability XZ:")[1]
    pol_txt, h0_txt = pol_txt.split("field-free Hamiltonian:")
    fix["shape"] = [int(x) for x in re.search(r"matrix dimensions: \((\d+), (\d+)\)", mat_txt).groups()]
    fix["dip_x"] = _entries(dip_txt.replace("dipole X:", "\n"))
    fix["pol_xz"] = _entries(pol_txt)
    fix["h0"] = _entries(h0_txt)
    asg_txt = next(t for t in outs if t.startswith("0 J = "))
    fix["assign"] = []
    for line in asg_txt.splitlines():
        m = re.match(r"^(\d+) J = ([\d.]+) , sym = (\w+) , m = (-?\d+) , k = (\(.*\))$", line)
        if m:
            kq, ke = ast.literal_eval(m.group(5))
            fix["assign"].append([int(m.group(1)), float(m.group(2)), m.group(3), int(m.group(4)), kq, ke])
    dims_txt = next(t for t in outs if "matrix dimensions for m =" in t)
    fix["m_filter_dims"] = [[int(a), int(b)] for a, b in
                            re.findall(r"matrix dimensions for m = (-?\d+): \((\d+), \d+\)", dims_txt)]
    with open(os.path.join(HERE, "water_h5_notebook.json"), "w") as f:
        json.dump(fix, f, indent=0)
    shutil.copyfile(os.path.join(nbdir, "water.h5"), os.path.join(HERE, "water.h5"))
    print({k: len(v) if isinstance(v, list) else v for k, v in fix.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
