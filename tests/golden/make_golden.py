"""Writes tests/golden/ref_vectors.npz with oracle/_ref/ref_driver, the
generator compiled from the reference's own energy sources (src/Potential.cpp,
src/Residue.cpp, src/AminoAcids.cpp, src/AminoAcid.cpp, src/vector3f.cpp,
src/Quaternion.cpp) by oracle/Makefile.

Run in the container that holds the reference tree:
    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (inputs and expected outputs only):
  consts                      RT_to_kcalmol, DH_CONVERSION_FACTOR, r0_constant
  conf{i}_*                   the ten reference conformations (box 1000)
  sys_{name}_*                seeded synthetic systems in the reference formats,
                              incl. molecules straddling the periodic boundary
  rot_in / rot_out            Quaternion(angle, axis).normalize().rotateVector(v)
  nrm_in / nrm_out            Vector3f::normalize
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

DRIVER = ROOT / "oracle" / "_ref" / "ref_driver"
DATA = ROOT / "tests" / "golden" / "data"
OUT = ROOT / "tests" / "golden" / "ref_vectors.npz"


def run(args, stdin):
    res = subprocess.run([str(DRIVER), *args], input=stdin, capture_output=True, text=True, check=True)
    return res.stdout


def energy_case(aa_file, pair_file, acids, mol_off, types, pos, box):
    lines = [f"box {box:.9g}", f"nmol {len(mol_off) - 1}"]
    lines += [str(mol_off[m + 1] - mol_off[m]) for m in range(len(mol_off) - 1)]
    for t, p in zip(types, pos):
        lines.append(f"{acids[int(t)].sname} {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}")
    out = run(["energy", str(aa_file), str(pair_file)], "\n".join(lines) + "\n").split("\n")
    lj, dh, tot = (float(x) for x in out[0].split())
    M = len(mol_off) - 1
    mat = np.array([[float(x) for x in out[1 + I].split()] for I in range(M)])
    return np.array([lj, dh, tot]), mat


def main():
    if not DRIVER.exists():
        raise SystemExit(f"{DRIVER} missing: run `make -C oracle ref`")
    fx = {}
    fx["consts"] = np.array([float(x) for x in run(["consts"], "").split()], dtype=np.float32)

    aa_file, pair_file = DATA / "AminoAcids", DATA / "pair_potentials_MJ1996"
    acids = formats.read_amino_acids(aa_file)
    for i in range(1, 11):
        a = formats.read_pdb(DATA / f"conf{i}" / "1a.pdb", acids)
        b = formats.read_pdb(DATA / f"conf{i}" / "1b.pdb", acids)
        types = np.concatenate([a.types, b.types])
        pos = np.concatenate([a.xyz, b.xyz])
        off = np.array([0, len(a.types), len(types)], dtype=np.int32)
        e, mat = energy_case(aa_file, pair_file, acids, off, types, pos, 1000.0)
        fx[f"conf{i}_energy"] = e

    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        systems = []
        s0 = synth.config(0)
        systems.append(("cfg0", s0, s0.positions()))
        small = synth.make_system("mixed8", [20, 57, 300, 33, 120, 76, 201, 45], 160.7, 4, 288.0, 450.0, 77)
        systems.append(("mixed8", small, small.positions()))
        # the same chains pushed across the periodic boundary and to large
        # coordinates: exercises the minimum image and the rounding of x/L
        wrapped = small.positions().copy()
        for m in range(small.n_mol):
            a, b = small.mol_off[m], small.mol_off[m + 1]
            shift = np.float32(small.box) * np.float32((m % 3) - 1) + np.float32(0.37 * m)
            wrapped[a:b] += shift
        systems.append(("mixed8wrap", small, wrapped.astype(np.float32)))
        tight = synth.make_system("tight4", [76, 76, 40, 24], 95.0, 2, 288.0, 450.0, 91)
        systems.append(("tight4", tight, tight.positions()))
        for name, spec, pos in systems:
            aa, pt = td / f"{name}_aa", td / f"{name}_pairs"
            formats.write_amino_acids(aa, spec.ff.acids)
            formats.write_pair_table(pt, spec.ff.acids, spec.ff.eps)
            # read back through the reference loaders: what the driver uses
            e, mat = energy_case(aa, pt, spec.ff.acids, spec.mol_off, spec.types, pos, spec.box)
            fx[f"sys_{name}_mol_off"] = np.asarray(spec.mol_off, dtype=np.int32)
            fx[f"sys_{name}_types"] = np.asarray(spec.types, dtype=np.int32)
            fx[f"sys_{name}_pos"] = np.asarray(pos, dtype=np.float32)
            fx[f"sys_{name}_box"] = np.array([spec.box], dtype=np.float32)
            fx[f"sys_{name}_eps"] = spec.ff.eps.astype(np.float32)
            fx[f"sys_{name}_sigma"] = spec.ff.sigma
            fx[f"sys_{name}_charge"] = spec.ff.charge
            fx[f"sys_{name}_energy"] = e
            fx[f"sys_{name}_matrix"] = mat

    rng = np.random.default_rng(5)
    n = 64
    rot_in = np.column_stack([
        rng.uniform(0.05, 0.5, n), rng.uniform(-0.5, 0.5, (n, 3)), rng.uniform(-30, 30, (n, 3)).astype(np.float32),
    ])
    text = "\n".join(f"{r[0]:.17g} {r[1]:.17g} {r[2]:.17g} {r[3]:.17g} {np.float32(r[4]):.9g} "
                     f"{np.float32(r[5]):.9g} {np.float32(r[6]):.9g}" for r in rot_in)
    fx["rot_in"] = rot_in
    fx["rot_out"] = np.array([[float(x) for x in l.split()] for l in run(["rotate"], text + "\n").strip().split("\n")],
                             dtype=np.float32)
    nrm_in = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    text = "\n".join(f"{v[0]:.9g} {v[1]:.9g} {v[2]:.9g}" for v in nrm_in)
    fx["nrm_in"] = nrm_in
    fx["nrm_out"] = np.array([[float(x) for x in l.split()] for l in run(["normalize"], text + "\n").strip().split("\n")],
                             dtype=np.float32)
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(fx)} arrays)")


if __name__ == "__main__":
    main()
