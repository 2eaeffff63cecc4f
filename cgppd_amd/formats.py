"""Readers and writers for cgppd's data formats.

* AminoAcids file     -- src/AminoAcids.cpp:29-73 (five non-comment lines per
                         type: name, 3-letter name, 1-letter symbol, van der
                         Waals radius, charge; index = file order)
* pair-potential file -- src/AminoAcids.cpp:75-138 (label row, then upper-
                         triangle rows, mirrored)
* PDB (C-alpha beads) -- src/Molecule.cpp:506-612 reader,
                         src/Replica.cpp:1124-1168 writer

The C++ host library (cgppd_amd/csrc/host) reads the same formats for the
drop-in Replica/Simulation API; these Python versions serve the test and
benchmark harness.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

AA_COUNT = 20
_FLOAT_PREFIX = re.compile(r"\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)")
_INT_PREFIX = re.compile(r"\s*([-+]?\d+)")


@dataclass
class AminoAcid:
    name: str
    sname: str
    symbol: str
    radius: float
    charge: float


@dataclass
class ForceField:
    """Per-type radius/charge and the raw 20x20 contact energies e_ij [RT]."""

    acids: list
    eps: np.ndarray  # (20, 20) float32, AminoAcids::LJpotentials

    @property
    def sigma(self) -> np.ndarray:
        return np.array([a.radius for a in self.acids], dtype=np.float32)

    @property
    def charge(self) -> np.ndarray:
        return np.array([a.charge for a in self.acids], dtype=np.float32)

    def index(self, name: str) -> int:
        return amino_acid_index(self.acids, name)


def _data_lines(path):
    with open(path, "r", errors="replace") as fh:
        for raw in fh:
            yield raw.rstrip("\n").rstrip("\r")


def read_amino_acids(path) -> list:
    """src/AminoAcids.cpp:29-73."""
    lines = list(_data_lines(path))
    acids = []
    k = 0
    while k < len(lines):
        line = lines[k]
        if line != "" and not line.startswith("#"):
            if k + 4 >= len(lines):
                raise ValueError(f"{path}: truncated amino-acid entry at line {k + 1}")
            name = line[:19]
            sname = lines[k + 1][:3]
            symbol = lines[k + 2][:1]
            m = _FLOAT_PREFIX.match(lines[k + 3])
            radius = float(m.group(1)) if m else 0.0
            m = _FLOAT_PREFIX.match(lines[k + 4])
            charge = float(m.group(1)) if m else 0.0
            acids.append(AminoAcid(name, sname, symbol, np.float32(radius).item(), np.float32(charge).item()))
            k += 5
        else:
            k += 1
    if len(acids) != AA_COUNT:
        raise ValueError(f"{path}: expected {AA_COUNT} amino acids, found {len(acids)}")
    return acids


def amino_acid_index(acids, name: str) -> int:
    """src/AminoAcids.cpp:145-181: 3-letter, 1-letter or full name, searched
    from the last entry down; -1 when absent."""
    if len(name) < 1:
        return -1
    if len(name) == 3:
        key = name[0].upper() + name[1:].lower()
        for i in range(AA_COUNT - 1, -1, -1):
            if acids[i].sname == key:
                return i
        return -1
    if len(name) == 1:
        key = name.upper()
        for i in range(AA_COUNT - 1, -1, -1):
            if acids[i].symbol == key:
                return i
        return -1
    key = name[0].upper() + name[1:]
    for i in range(AA_COUNT - 1, -1, -1):
        if acids[i].name == key:
            return i
    return -1


def read_pair_table(path, acids) -> np.ndarray:
    """src/AminoAcids.cpp:75-138."""
    eps = np.zeros((AA_COUNT, AA_COUNT), dtype=np.float32)
    labels = None
    start = 0
    for line in _data_lines(path):
        if line == "" or line.startswith("#"):
            continue
        toks = line.split()
        if labels is None:
            labels = [amino_acid_index(acids, t) for t in toks]
            if len(labels) != AA_COUNT or min(labels) < 0:
                raise ValueError(f"{path}: bad label row")
            continue
        if start >= AA_COUNT:
            break
        vals = toks[1:]
        if len(vals) < AA_COUNT - start:
            raise ValueError(f"{path}: row {start} has {len(vals)} values, needs {AA_COUNT - start}")
        for j in range(start, AA_COUNT):
            v = np.float32(float(vals[j - start]))
            eps[labels[start], labels[j]] = v
            eps[labels[j], labels[start]] = v
        start += 1
    if labels is None or start != AA_COUNT:
        raise ValueError(f"{path}: incomplete pair table")
    return eps


def load_force_field(aa_path, pair_path) -> ForceField:
    acids = read_amino_acids(aa_path)
    return ForceField(acids, read_pair_table(pair_path, acids))


def write_amino_acids(path, acids):
    with open(path, "w") as fh:
        fh.write("#Amino acid class initialisation data\n#name\n#abbriviation\n#short symbol\n"
                 "#van der Waal radius\n#residue charge\n")
        for a in acids:
            fh.write(f"{a.name}\n{a.sname}\n{a.symbol}\n{a.radius:.9g}\n{a.charge:.9g}\n\n")


def write_pair_table(path, acids, eps):
    labels = [a.sname for a in acids]
    with open(path, "w") as fh:
        fh.write("# synthetic contact energies e_ij [RT]\n")
        fh.write("     " + " ".join(f"{s:>5}" for s in labels) + "\n")
        for i in range(AA_COUNT):
            row = " ".join(f"{float(eps[i, j]):.9g}" for j in range(i, AA_COUNT))
            fh.write(f"{labels[i]} {row}\n")


@dataclass
class PDBMolecule:
    types: np.ndarray       # (n,) int32
    xyz: np.ndarray         # (n, 3) float32
    chain_ids: list = field(default_factory=list)
    res_seq: list = field(default_factory=list)
    chain_count: int = 1
    chain_index: list = field(default_factory=list)  # TER records seen before each residue

    def centred(self):
        """Centre (float sum / count) and relative positions as
        src/Molecule.cpp:590-611 computes them."""
        c = np.zeros(3, dtype=np.float32)
        for p in self.xyz:
            c = (c + p).astype(np.float32)
        c = (c / np.float32(len(self.xyz))).astype(np.float32)
        return c, (self.xyz - c).astype(np.float32)


def _field_float(line, col):
    m = _FLOAT_PREFIX.match(line, col) if col < len(line) else None
    return np.float32(float(m.group(1))).item() if m else 0.0


def read_pdb(path, acids) -> PDBMolecule:
    """src/Molecule.cpp:506-612: CA atoms only, TER counts chains, END stops."""
    types, xyz, chains, seqs, chain_index = [], [], [], [], []
    chain_count = 0
    with open(path, "r", errors="replace") as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith("ATOM"):
                name = line[12:].split()[0] if len(line) > 12 and line[12:].split() else ""
                if name != "CA":
                    continue
                res_name = line[17:].split()[0] if len(line) > 17 else ""
                chain = line[21] if len(line) > 21 else " "
                m = _INT_PREFIX.match(line, 22) if len(line) > 22 else None
                seq = int(m.group(1)) if m else 0
                types.append(amino_acid_index(acids, res_name))
                xyz.append((_field_float(line, 30), _field_float(line, 38), _field_float(line, 46)))
                chains.append(chain)
                seqs.append(seq)
                chain_index.append(chain_count)
            elif line.startswith("TER"):
                chain_count += 1
            elif line.startswith("END"):
                break
    if types and chain_count == 0:
        chain_count = 1
    return PDBMolecule(np.array(types, dtype=np.int32), np.array(xyz, dtype=np.float32).reshape(-1, 3),
                       chains, seqs, chain_count, chain_index)


def read_torsions(path, acids):
    """Torsion table (src/TorsionalLookupMatrix.cpp:27-52): lines
    'AAA BBB V n sigma_degrees'; returns V and the phase in radians, each
    (20, 20, 4) float64, indexed [type_a, type_b, n - 1]."""
    V = np.zeros((AA_COUNT, AA_COUNT, 4))
    D = np.zeros((AA_COUNT, AA_COUNT, 4))
    for line in _data_lines(path):
        if line == "" or line.startswith("#"):
            continue
        f = line.split()
        a, b = amino_acid_index(acids, f[0][:3]), amino_acid_index(acids, f[1][:3])
        n = int(f[3])
        V[a, b, n - 1] = float(f[2])
        D[a, b, n - 1] = float(f[4]) / 57.29577951308232
    return V, D


def write_torsions(path, acids, V, delta):
    """Writes a torsion table in the layout read_torsions reads (phase in
    degrees, as data/torsional_pair_potentials stores it)."""
    with open(path, "w") as fh:
        fh.write("# NAME NAME V N SIGMA\n")
        for a in range(AA_COUNT):
            for b in range(AA_COUNT):
                for n in range(4):
                    fh.write("%s %s %12.6f %d %12.6f\n" % (acids[a].sname.upper(), acids[b].sname.upper(), V[a, b, n],
                                                           n + 1, delta[a, b, n] * 57.29577951308232))


def write_pdb(path, acids, types, xyz, chain="A", res_seq=None):
    """C-alpha PDB in the column layout src/Molecule.cpp:526-575 reads."""
    with open(path, "w") as fh:
        for k, (t, p) in enumerate(zip(types, xyz)):
            seq = res_seq[k] if res_seq is not None else k + 1
            fh.write("ATOM  %5d  CA  %3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                     % (k + 1, acids[int(t)].sname.upper(), chain, seq, p[0], p[1], p[2], 0.0, 1.0))
        fh.write("TER\nEND\n")
