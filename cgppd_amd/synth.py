"""Seeded synthetic systems for the benchmark configurations.

Every system is expressed in the reference's own terms (AminoAcids-style
per-type radius/charge, a 20x20 contact-energy table, C-alpha chains with
centre + relative coordinates) and can be written out in the reference file
formats (cgppd_amd.formats).

Configurations (rigid chains, one C-alpha bead per residue):
  cfg0  2 x 76 residues, box 150 A, 20 replicas 288-450 K, steps 2 A / 0.2 rad
  cfg1  40 chains, lengths ~ U[20,300], 400 uM total molecule concentration
        (box 549.6 A), 32 replicas 288-450 K
  cfg4  256 chains, lengths ~ log-U[20,2000], 800 uM (box 810 A), 64 replicas

Chains are self-avoiding random walks with 3.8 A bonds grown inside a sphere
so that the radius of gyration is about 12 A at 76 residues (scaled as
n^0.38 for other lengths).  Residue types are uniform over the 20 types;
charges K/R +1, D/E -1, H +0.5; contact energies e_ij ~ U[-7,-1] RT
(symmetric); radii sigma_i ~ U[4.6,7.4] A.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .formats import AminoAcid, ForceField

# Reference residue order (data/AminoAcids, src/AminoAcids.cpp:29-73).
ACID_NAMES = [
    ("Alanine", "Ala", "A"), ("Argnine", "Arg", "R"), ("Asparagine", "Asn", "N"),
    ("Aspartic_acid", "Asp", "D"), ("Cysteine", "Cys", "C"), ("Glutamic_acid", "Glu", "E"),
    ("Glutamine", "Gln", "Q"), ("Glycine", "Gly", "G"), ("Leucine", "Leu", "L"),
    ("Lysine", "Lys", "K"), ("Methionine", "Met", "M"), ("Phenylalanine", "Phe", "F"),
    ("Proline", "Pro", "P"), ("Serine", "Ser", "S"), ("Threonine", "Thr", "T"),
    ("Tryptophan", "Trp", "W"), ("Tyrosine", "Tyr", "Y"), ("Valine", "Val", "V"),
    ("Histine", "His", "H"), ("Isoleucine", "Ile", "I"),
]
CHARGE_BY_SYMBOL = {"K": 1.0, "R": 1.0, "D": -1.0, "E": -1.0, "H": 0.5}

AVOGADRO = 6.02214129e23


def synthetic_force_field(seed: int) -> ForceField:
    rng = np.random.default_rng(seed)
    sig = rng.uniform(4.6, 7.4, size=20).astype(np.float32)
    acids = [AminoAcid(n, s, c, float(sig[k]), CHARGE_BY_SYMBOL.get(c, 0.0)) for k, (n, s, c) in enumerate(ACID_NAMES)]
    e = rng.uniform(-7.0, -1.0, size=(20, 20))
    e = np.triu(e) + np.triu(e, 1).T
    return ForceField(acids, e.astype(np.float32))


def box_for_concentration(n_mol: int, micromolar: float) -> float:
    """Cubic box edge [A] holding n_mol molecules at a total concentration."""
    vol_m3 = n_mol / (micromolar * 1e-6 * 1e3 * AVOGADRO)
    return float(vol_m3 ** (1.0 / 3.0) * 1e10)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def compact_chain(n: int, rng, bond: float = 3.8, min_dist: float = 4.0, rg76: float = 12.0) -> np.ndarray:
    """Self-avoiding walk with fixed bond length grown inside a sphere sized
    for the target radius of gyration; returns (n, 3) float64 centred at 0."""
    rg = rg76 * (n / 76.0) ** 0.38
    radius = max(rg * np.sqrt(5.0 / 3.0), bond * 1.5)
    for _attempt in range(200):
        pts = np.zeros((n, 3))
        k = 1
        stuck = 0
        while k < n:
            dirs = rng.normal(size=(48, 3))
            dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
            cand = pts[k - 1] + bond * dirs
            ok = np.linalg.norm(cand, axis=1) < radius
            if k > 1:
                d = np.linalg.norm(cand[:, None, :] - pts[None, : k - 1, :], axis=2)
                ok &= (d >= min_dist).all(axis=1)
            idx = np.flatnonzero(ok)
            if idx.size:
                pts[k] = cand[idx[0]]
                k += 1
                stuck = 0
            else:
                stuck += 1
                back = min(k - 1, 2 + stuck)
                k -= back
                if stuck > 50:
                    break
        if k == n:
            return pts - pts.mean(axis=0)
    raise RuntimeError(f"could not grow a compact chain of {n} residues")


@dataclass
class SystemSpec:
    name: str
    ff: ForceField
    box: float
    lengths: np.ndarray      # (M,)
    mol_off: np.ndarray      # (M+1,) int32
    types: np.ndarray        # (N,) int32
    rel: np.ndarray          # (N, 3) float32, relative to each molecule centre
    centers: np.ndarray      # (M, 3) float32
    n_rep: int
    t_min: float
    t_max: float
    translate_step: float | None  # None: the reference ladder
    rotate_step: float | None
    mc_steps: int
    exchange_every: int

    @property
    def n_mol(self):
        return len(self.lengths)

    @property
    def n_res(self):
        return int(self.mol_off[-1])

    def positions(self) -> np.ndarray:
        pos = np.empty_like(self.rel)
        for m in range(self.n_mol):
            a, b = self.mol_off[m], self.mol_off[m + 1]
            pos[a:b] = self.rel[a:b] + self.centers[m]
        return pos

    def pairs_full(self) -> int:
        n = self.lengths.astype(np.int64)
        return int((n.sum() ** 2 - (n * n).sum()) // 2)

    def temperatures(self) -> np.ndarray:
        """src/Simulation.cpp:590 + src/Replica.cpp:86: float ratio, double
        power, stored as float."""
        R = self.n_rep
        if R == 1:
            return np.array([np.float32(self.t_min)], dtype=np.float64)
        ratio = float(np.float32(self.t_max) / np.float32(self.t_min))
        g = ratio ** (1.0 / (R - 1))
        return np.array([np.float32(np.float32(self.t_min) * g ** i) for i in range(R)], dtype=np.float64)

    def step_sizes(self):
        """Fixed steps, or the reference ladder (src/Simulation.cpp:591-592,
        src/Replica.cpp:87-88, src/definitions.h:27-32)."""
        R = self.n_rep
        if self.translate_step is not None:
            t = np.full(R, self.translate_step, dtype=np.float32)
        else:
            g = (2.0 / 0.5) ** (1.0 / max(R - 1, 1))
            t = np.array([np.float32(0.5 * g ** i) for i in range(R)], dtype=np.float32)
        if self.rotate_step is not None:
            rs = np.full(R, self.rotate_step, dtype=np.float32)
        else:
            g = (0.4 / 0.2) ** (1.0 / max(R - 1, 1))
            rs = np.array([np.float32(0.2 * g ** i) for i in range(R)], dtype=np.float32)
        return t, rs


def _place(lengths, radii, box, rng):
    centers = np.zeros((len(lengths), 3))
    for m in range(len(lengths)):
        for _ in range(100000):
            c = rng.uniform(0.0, box, size=3)
            if m == 0:
                break
            d = centers[:m] - c
            d -= box * np.round(d / box)
            if (np.linalg.norm(d, axis=1) > radii[:m] + radii[m] + 4.0).all():
                break
        else:
            raise RuntimeError("could not place molecules without overlap")
        centers[m] = c
    return centers


def make_system(name: str, lengths, box: float, n_rep: int, t_min: float, t_max: float, seed: int,
                translate_step=2.0, rotate_step=0.2, mc_steps=10000, exchange_every=1000) -> SystemSpec:
    rng = np.random.default_rng(seed)
    ff = synthetic_force_field(seed + 1)
    lengths = np.asarray(lengths, dtype=np.int64)
    mol_off = np.zeros(len(lengths) + 1, dtype=np.int32)
    mol_off[1:] = np.cumsum(lengths)
    N = int(mol_off[-1])
    types = rng.integers(0, 20, size=N).astype(np.int32)
    rel = np.zeros((N, 3), dtype=np.float32)
    radii = np.zeros(len(lengths))
    for m, n in enumerate(lengths):
        pts = compact_chain(int(n), rng) @ _random_rotation(rng).T
        rel[mol_off[m]:mol_off[m + 1]] = pts.astype(np.float32)
        radii[m] = np.linalg.norm(pts, axis=1).max()
    centers = _place(lengths, radii, box, rng).astype(np.float32)
    return SystemSpec(name, ff, float(np.float32(box)), lengths, mol_off, types, rel, centers, n_rep, t_min, t_max,
                      translate_step, rotate_step, mc_steps, exchange_every)


def config(k: int, seed: int | None = None) -> SystemSpec:
    """The synthetic system of benchmark configuration k (0, 1 or 4)."""
    if k == 0:
        s = 1000 if seed is None else seed
        return make_system("pair_2x76", [76, 76], 150.0, 20, 288.0, 450.0, s, 2.0, 0.2, 10000, 1000)
    if k == 1:
        s = 1001 if seed is None else seed
        rng = np.random.default_rng(s + 7)
        lengths = rng.integers(20, 301, size=40)
        return make_system("crowded40", lengths, box_for_concentration(40, 400.0), 32, 288.0, 450.0, s,
                           2.0, 0.2, 50000, 1000)
    if k == 4:
        s = 1004 if seed is None else seed
        rng = np.random.default_rng(s + 7)
        lengths = np.exp(rng.uniform(np.log(20.0), np.log(2000.0), size=256)).astype(np.int64)
        return make_system("crowded256", lengths, box_for_concentration(256, 800.0), 64, 288.0, 450.0, s,
                           2.0, 0.2, 50000, 1000)
    raise ValueError(f"no synthetic rigid system for configuration {k}")
