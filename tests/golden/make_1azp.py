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

cpydock (cpydock.rs:172-332) types with amber.rs as pydock does and adds, per
atom, the hydrogen flag from the PDB element column (name fallback), the
desolvation coefficient and radius from the ASP tables of cpydock.rs (the
ASP_CHARGES / ASP_RADII arrays, the asp_type, radius_per_asp and des_energy
match arms, read as text) after the residue renames of cpydock.rs:110-118.
The areas themselves (Lee & Richards over each monomer's heavy atoms with the
desolvation radii) are not stored: they are what the SASA under test computes.

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


def read_asp(path):
    """The desolvation tables of cpydock.rs as data: ASP charges / radii by
    ASP type, {residue: {atom: type}}, and the two amber-type fallbacks."""
    text = open(path).read()

    def arr(name):
        m = re.search(r"const %s: \[f64; \d+\] = \[(.*?)\];" % name, text, re.S)
        return [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]

    def arms(body):
        out = {}
        for keys, val in re.findall(r'((?:"[^"]+"\s*\|?\s*)+)=>\s*([-0-9.eE]+)', body):
            for k in re.findall(r'"([^"]+)"', keys):
                out[k] = val
        return out

    def fn_body(name, start):
        i = text.index("fn %s(" % name)
        j = text.index(start, i)
        depth, k = 0, j
        while True:  # the balanced block that starts at `start`'s brace
            c = text[k]
            depth += c == "{"
            depth -= c == "}"
            k += 1
            if depth == 0 and c == "}":
                return text[j:k]

    asp = {}
    for res, body in re.findall(r'"(\w+)" => match atom_name \{(.*?)_ => return None', fn_body("asp_type", "match res_name")):
        asp[res] = {a: int(v) for a, v in arms(body).items()}
    rad_fb = {k: float(v) for k, v in arms(fn_body("radius_per_asp", "match amber_type")).items()}
    des_fb = {k: float(v) for k, v in arms(fn_body("des_energy", "match amber_type")).items()}
    return arr("ASP_CHARGES"), arr("ASP_RADII"), asp, rad_fb, des_fb


def desolvation(asp, res, names, amber_types):
    """cpydock.rs:104-137: the coefficient and radius of each atom."""
    charges, radii, types, rad_fb, des_fb = asp
    des, rad = [], []
    for r, a, t in zip(res, names, amber_types):
        if a == "OXT":
            r, a = "ASP", "OD1"
        r = {"CYX": "CYS", "HSD": "HID", "HSP": "HIP", "HSE": "HIE"}.get(r, r)
        if r == "HID":
            r = "HIP"
        k = types.get(r, {}).get(a)
        if k is not None:
            des.append(charges[k])
            rad.append(radii[k])
        else:
            des.append(des_fb.get(t, 0.0))
            rad.append(rad_fb.get(t, 0.0))
    return np.asarray(des), np.asarray(rad)


def read_pdb(path):
    res, names, xyz, elem = [], [], [], []
    for line in open(path):
        if not line.startswith(("ATOM  ", "HETATM")):
            continue
        if line[16] not in " A":
            raise ValueError("alternate locations are not handled: " + line)
        names.append(line[12:16].strip())
        elem.append("".join(c for c in line[76:78] if c.isalpha()))
        res.append(line[17:20].strip())
        xyz.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
    return res, names, np.asarray(xyz, np.float64), elem


def type_atoms(tables, res, names):
    T, E, NT = tables["AMBER_TYPES"], tables["ELE_CHARGES"], tables["NT_ELE_CHARGES"]
    VC, VR = tables["VDW_CHARGES"], tables["VDW_RADII"]
    ele, eps, rad, types, fallback = [], [], [], [], 0
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
        types.append(amber)
        ele.append(E.get(atom_id, NT.get(atom_id, 0.0)))
        eps.append(VC.get(amber, 0.086))
        rad.append(VR.get(amber, 1.908))
    return np.asarray(ele), np.asarray(eps), np.asarray(rad), types, fallback


def main(ref):
    out = {}
    for side in ("receptor", "ligand"):
        res, names, xyz, elem = read_pdb(os.path.join(ref, "tests", "1azp", "1azp_%s.pdb" % side))
        k = side[:3]
        out[k + "_xyz"] = xyz
        out[k + "_res"] = np.asarray(res)
        out[k + "_name"] = np.asarray(names)
        out[k + "_heavy"] = np.asarray([not a.startswith("H") for a in names], np.uint8)
        for fam, src in (("pydock", "amber.rs"), ("dna", "dna.rs")):
            ele, eps, rad, types, nfb = type_atoms(read_tables(os.path.join(ref, "src", src)), res, names)
            if fam == "pydock":  # cpydock types with the same tables
                des, desrad = desolvation(read_asp(os.path.join(ref, "src", "cpydock.rs")), res, names,
                                          types)
                hyd = np.asarray([0 if (e == "H" if e else a.startswith("H")) else 1
                                  for e, a in zip(elem, names)], np.int32)
                out[k + "_cpy_des"], out[k + "_cpy_desrad"], out[k + "_cpy_hyd"] = des, desrad, hyd
                print("%s cpydock: %d heavy, %d with a desolvation radius, sum des %.5f"
                      % (side, hyd.sum(), int((desrad > 0).sum()), des.sum()))
            out["%s_%s_ele" % (k, fam)] = ele
            out["%s_%s_eps" % (k, fam)] = eps
            out["%s_%s_rad" % (k, fam)] = rad
            print("%s %s: %d atoms, %d typed by the element fallback, net charge %.4f"
                  % (side, fam, len(names), nfb, ele.sum()))
    np.savez_compressed(os.path.join(HERE, "1azp.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
