"""Writes tests/golden/flex_vectors.npz with oracle/_ref/ref_flex_driver, the
generator compiled with -DEnableFlexibleLinkers from the reference's own
src/Geometry.cpp, src/Potential.cpp, src/Residue.cpp,
src/TorsionalLookupMatrix.cpp, src/AminoAcids.cpp, src/AminoAcid.cpp,
src/vector3f.cpp and src/Quaternion.cpp by oracle/Makefile.

Run in the container that holds the reference tree:
    make -C oracle ref_flex && python tests/golden/make_flex_golden.py

Cases (inputs and expected outputs only):
  ala8, angiotensin   all-flexible peptides of tests/testMolecule.cpp:26-31
                      (box 15 and 1000, centred at the origin)
  ubq_seg             ubiquitin with the flexible tail 72-75 (box 1000),
                      tests/testMolecule.cpp:11-20
  conf1_flex          data/conf1 with 1b's tail 72-75 flexible (box 118.4),
                      tests/testReplica.cpp:30-58
  diubq               the two-chain topology of tests/testGeometry.cpp:6-28
  linked, linked_mixed  seeded linked-domain systems in the reference formats

Per case: {case}_{input} arrays, the topology of every molecule
({case}_m{m}_{key}) and the energies ({case}_inter, {case}_m{m}_internal_true,
_internal_false, _gpu_intra: raw LJ, raw DH, raw bond, angle product,
torsion, total).
"""
from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from cgppd_amd import formats, synth  # noqa: E402

DRIVER = ROOT / "oracle" / "_ref" / "ref_flex_driver"
DATA = ROOT / "tests" / "golden" / "data"
OUT = ROOT / "tests" / "golden" / "flex_vectors.npz"

LIST_KEYS = ["bonds", "angles", "torsions", "local", "crankshaft", "flex", "segment_bonds", "indirect"]


def system_text(acids, case):
    lines = [f"box {case['box']:.9g}", f"nmol {len(case['mol_off']) - 1}"]
    for m in range(len(case["mol_off"]) - 1):
        a, b = case["mol_off"][m], case["mol_off"][m + 1]
        segs = case["segments"][m]
        lines.append(f"mol {b - a} {int(case['all_flexible'][m])} {len(segs)}")
        for sg in segs:
            lines.append("seg %d %s" % (len(sg), " ".join(str(v) for v in sg)))
        for i in range(a, b):
            p = case["pos"][i]
            lines.append("%s %s %d %d %.9g %.9g %.9g" % (acids[int(case["types"][i])].sname.upper(),
                                                        chr(int(case["chain_code"][i])) if case["chain_code"][i] != 32
                                                        else "_", case["chain_index"][i], case["res_seq"][i],
                                                        p[0], p[1], p[2]))
    return "\n".join(lines) + "\n"


def run_case(name, case, aa_file, pair_file, tors_file, acids, fx):
    text = system_text(acids, case)
    res = subprocess.run([str(DRIVER), "system", str(aa_file), str(pair_file), str(tors_file)], input=text,
                         capture_output=True, text=True, check=True)
    M = len(case["mol_off"]) - 1
    m = -1
    branches = {}
    for line in res.stdout.strip().split("\n"):
        f = line.split()
        key = f[0]
        if key == "mol":
            m = int(f[1])
            fx[f"{name}_m{m}_flexible"] = np.array([int(f[3])], dtype=np.int32)
            branches[m] = ([], [0], [])
            domains = [[0], []]
        elif key in LIST_KEYS:
            fx[f"{name}_m{m}_{key}"] = np.array([int(v) for v in f[2:]], dtype=np.int32)
        elif key == "domains":
            pass
        elif key == "domain":
            domains[1].extend(int(v) for v in f[2:])
            domains[0].append(len(domains[1]))
            fx[f"{name}_m{m}_domain_off"] = np.array(domains[0], dtype=np.int32)
            fx[f"{name}_m{m}_domain_res"] = np.array(domains[1], dtype=np.int32)
        elif key == "branch":
            ri, nb, n = int(f[1]), int(f[2]), int(f[3])
            keys, off, res_ = branches[m]
            keys.append((ri, nb))
            res_.extend(int(v) for v in f[4:4 + n])
            off.append(len(res_))
        elif key == "uids":
            fx[f"{name}_m{m}_uids"] = np.array([int(v) for v in f[1:]], dtype=np.int32).reshape(-1, 3)
        elif key == "inter":
            fx[f"{name}_inter"] = np.array([float(v) for v in f[1:]])
        elif key in ("internal_true", "internal_false", "gpu_intra"):
            fx[f"{name}_m{int(f[1])}_{key}"] = np.array([float(v) for v in f[2:]])
    for mm, (keys, off, res_) in branches.items():
        fx[f"{name}_m{mm}_branch_key"] = np.array(keys, dtype=np.int32).reshape(-1, 2)
        fx[f"{name}_m{mm}_branch_off"] = np.array(off, dtype=np.int32)
        fx[f"{name}_m{mm}_branch_res"] = np.array(res_, dtype=np.int32)
    for k in ("mol_off", "types", "chain_index", "chain_code", "res_seq", "all_flexible"):
        fx[f"{name}_{k}"] = np.asarray(case[k], dtype=np.int32)
    fx[f"{name}_pos"] = np.asarray(case["pos"], dtype=np.float32)
    fx[f"{name}_box"] = np.array([case["box"]], dtype=np.float32)
    seg_mol, seg_off, seg_res = [], [0], []
    for mm in range(M):
        for sg in case["segments"][mm]:
            seg_mol.append(mm)
            seg_res.extend(sg)
            seg_off.append(len(seg_res))
    fx[f"{name}_seg_mol"] = np.array(seg_mol, dtype=np.int32)
    fx[f"{name}_seg_off"] = np.array(seg_off, dtype=np.int32)
    fx[f"{name}_seg_res"] = np.array(seg_res, dtype=np.int32)
    ff = formats.load_force_field(aa_file, pair_file)
    fx[f"{name}_eps"] = ff.eps.astype(np.float32)
    fx[f"{name}_sigma"] = ff.sigma
    fx[f"{name}_charge"] = ff.charge
    V, D = formats.read_torsions(tors_file, ff.acids)
    fx[f"{name}_torsion_v"] = V
    fx[f"{name}_torsion_delta"] = D


def pdb_case(pdbs, acids, box, all_flexible, segments, centre):
    types, pos, ci, cc, rs, off = [], [], [], [], [], [0]
    for path in pdbs:
        mol = formats.read_pdb(path, acids)
        xyz = mol.centred()[1] if centre else mol.xyz
        types.append(mol.types)
        pos.append(xyz)
        ci += mol.chain_index
        cc += [ord(c) for c in mol.chain_ids]
        rs += mol.res_seq
        off.append(off[-1] + len(mol.types))
    return dict(mol_off=np.array(off), types=np.concatenate(types), pos=np.concatenate(pos), box=box,
                chain_index=np.array(ci), chain_code=np.array(cc), res_seq=np.array(rs),
                all_flexible=np.array(all_flexible), segments=segments)


def spec_case(spec):
    return dict(mol_off=spec.mol_off, types=spec.types, pos=spec.positions(), box=spec.box,
                chain_index=spec.chain_index, chain_code=spec.chain_code, res_seq=spec.res_seq,
                all_flexible=np.array(spec.all_flexible, dtype=np.int32), segments=spec.segments)


def main():
    if not DRIVER.exists():
        raise SystemExit(f"{DRIVER} missing: run `make -C oracle ref_flex`")
    fx = {}
    aa_file, pair_file, tors_file = DATA / "AminoAcids", DATA / "pair_potentials_MJ1996", \
        DATA / "torsional_pair_potentials"
    acids = formats.read_amino_acids(aa_file)
    cases = {
        "ala8": pdb_case([DATA / "ala8.pdb"], acids, 15.0, [1], [[]], True),
        "angiotensin": pdb_case([DATA / "angiotensin.pdb"], acids, 1000.0, [1], [[]], True),
        "ubq_seg": pdb_case([DATA / "1UBQ.pdb"], acids, 1000.0, [0], [[[72, 73, 74, 75]]], True),
        "conf1_flex": pdb_case([DATA / "conf1" / "1a.pdb", DATA / "conf1" / "1b.pdb"], acids, 118.4, [0, 0],
                               [[], [[72, 73, 74, 75]]], False),
    }
    # tests/testGeometry.cpp:6-28: chains A and B of 76 residues, resSeq 1+c+r
    n = 152
    cases["diubq"] = dict(mol_off=np.array([0, n]), types=np.zeros(n, dtype=np.int32),
                          pos=np.column_stack([np.arange(n) * 3.8, np.zeros(n), np.zeros(n)]).astype(np.float32),
                          box=1000.0, chain_index=np.repeat([0, 1], 76), chain_code=np.repeat([65, 66], 76),
                          res_seq=np.array([1 + c + r for c in range(2) for r in range(76)]),
                          all_flexible=np.array([0]), segments=[[[72, 73, 74, 75, 123], [148, 149, 150, 151]]])
    for name, case in cases.items():
        run_case(name, case, aa_file, pair_file, tors_file, acids, fx)

    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        specs = {
            "linked": synth.config(2),
            "linked_mixed": synth.make_linked_system("linked_mixed", 2, 40, 6, 0.2, 120.0, 4, 288.0, 500.0, 31,
                                                     extra_rigid=(30, 55)),
        }
        for name, spec in specs.items():
            aa, pt, tt = td / f"{name}_aa", td / f"{name}_pairs", td / f"{name}_tors"
            formats.write_amino_acids(aa, spec.ff.acids)
            formats.write_pair_table(pt, spec.ff.acids, spec.ff.eps)
            formats.write_torsions(tt, spec.ff.acids, spec.torsion_v, spec.torsion_delta)
            run_case(name, spec_case(spec), aa, pt, tt, formats.read_amino_acids(aa), fx)
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(fx)} arrays)")


if __name__ == "__main__":
    main()
