"""Write tests/golden/1azp.npz: the 1AZP complex of the reference's own unit
tests (tests/1azp/1azp_receptor.pdb, 1azp_ligand.pdb; used by pydock.rs:540-594
and dna.rs:1112-1180) as per-atom arrays under the pydock and dna parameter
sets.

    python tests/golden/make_1azp.py /root/reference

Inputs read as text, nothing of the reference is run: the two PDB files (ATOM
records, fixed columns) and the atom-typing tables (`hashmap![...]` literals)
of amber.rs (pydock) and dna.rs (dna).  The typing rule is restated from
pydock.rs:120-157 / dna.rs:344-401:

  atom_id = "RES-NAME"; type = AMBER_TYPES[atom_id], else for H1/H2/H3 the
  residue's "RES-H" entry, else the element fallback "*-<first letter>", else
  "C"; charge = ELE_CHARGES[atom_id], else NT_ELE_CHARGES[atom_id], else 0;
  epsilon = VDW_CHARGES[type] (default 0.086); radius = VDW_RADII[type]
  (default 1.908); heavy = the atom name does not start with 'H' (dna.rs:397).

Atoms are kept in file order (one chain per file, no alternate locations or
insertion codes in these two files, so pdbtbx's chain/residue/atom walk is the
file order).  The output holds data only: coordinates, per-atom parameters,
and the residue/atom names they were derived from.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_tables(path):
    """name -> dict of the `static ref NAME: HashMap<..> = hashmap![k => v, ...];`
    literals of one source file (a repeated key keeps its last value, as the
    macro's successive inserts do)."""
    text = open(path).read()
    out = {}
    for m in re.finditer(r"static ref (\w+)\s*:\s*HashMap<[^>]*>\s*=\s*hashmap!\[(.*?)\];", text, re.S):
        d = {}
        for k, v in re.findall(r'"([^"]+)"\s*=>\s*("[^"]*"|[-+0-9.eE]+)', m.group(2)):
            d[k] = v.strip('"') if v.startswith('"') else float(v)
        out[m.group(1)] = d
    return out


def read_pdb(path):
    res, names, xyz = [], [], []
    for line in open(path):
        if not line.startswith(("ATOM  ", "HETATM")):
            continue
        if line[16] not in " A":
            raise ValueError("alternate locations are not handled: " + line)
        names.append(line[12:16].strip())
        res.append(line[17:20].strip())
        xyz.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
    return res, names, np.asarray(xyz, np.float64)


def type_atoms(tables, res, names):
    T, E, NT = tables["AMBER_TYPES"], tables["ELE_CHARGES"], tables["NT_ELE_CHARGES"]
    VC, VR = tables["VDW_CHARGES"], tables["VDW_RADII"]
    ele, eps, rad, fallback = [], [], [], 0
    for r, a in zip(res, names):
        atom_id = "%s-%s" % (r, a)
        if atom_id in T:
            amber = T[atom_id]
        elif a in ("H1", "H2", "H3") and ("%s-H" % r) in T:
            atom_id = "%s-H" % r
            amber = T[atom_id]
        else:
            fallback += 1
            atom_id = "*-%s" % (a[:1] or "C").upper()
            amber = T.get(atom_id, "C")
        ele.append(E.get(atom_id, NT.get(atom_id, 0.0)))
        eps.append(VC.get(amber, 0.086))
        rad.append(VR.get(amber, 1.908))
    return np.asarray(ele), np.asarray(eps), np.asarray(rad), fallback


def main(ref):
    out = {}
    for side in ("receptor", "ligand"):
        res, names, xyz = read_pdb(os.path.join(ref, "tests", "1azp", "1azp_%s.pdb" % side))
        k = side[:3]
        out[k + "_xyz"] = xyz
        out[k + "_res"] = np.asarray(res)
        out[k + "_name"] = np.asarray(names)
        out[k + "_heavy"] = np.asarray([not a.startswith("H") for a in names], np.uint8)
        for fam, src in (("pydock", "amber.rs"), ("dna", "dna.rs")):
            ele, eps, rad, nfb = type_atoms(read_tables(os.path.join(ref, "src", src)), res, names)
            out["%s_%s_ele" % (k, fam)] = ele
            out["%s_%s_eps" % (k, fam)] = eps
            out["%s_%s_rad" % (k, fam)] = rad
            print("%s %s: %d atoms, %d typed by the element fallback, net charge %.4f"
                  % (side, fam, len(names), nfb, ele.sum()))
    np.savez_compressed(os.path.join(HERE, "1azp.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
