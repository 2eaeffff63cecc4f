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
  cfg2  flexible linkers: one molecule of four 76-residue domains joined by
        6-residue linkers (322 residues); 20 % of each domain's residues
        (its tails next to the linkers) and the linkers are flexible
        segments; seeded 20x20x4 torsion table; 64 replicas 288-500 K

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
    # Placed residues are binned in cubic cells wider than min_dist.  Every
    # candidate lies within `bond` of the previous residue, so only residues
    # within bond + min_dist of it can be closer than min_dist to one: the
    # cells covering that cube hold all of them, and the test below on that
    # subset takes the same decisions as the test against every residue.
    cell = min_dist * 1.125
    reach = int(np.ceil((bond + min_dist) / cell))
    for _attempt in range(200):
        pts = np.zeros((n, 3))
        grid: dict = {}
        cell_of = [None] * n

        def place(i, p):
            if cell_of[i] is not None:
                grid[cell_of[i]].discard(i)
            c = tuple(np.floor(p / cell).astype(np.int64))
            grid.setdefault(c, set()).add(i)
            cell_of[i] = c

        place(0, pts[0])
        k = 1
        stuck = 0
        while k < n:
            dirs = rng.normal(size=(48, 3))
            dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
            cand = pts[k - 1] + bond * dirs
            ok = np.linalg.norm(cand, axis=1) < radius
            if k > 1:
                cx, cy, cz = np.floor(pts[k - 1] / cell).astype(np.int64)
                near = [i for a in range(cx - reach, cx + reach + 1) for b in range(cy - reach, cy + reach + 1)
                        for c in range(cz - reach, cz + reach + 1) for i in grid.get((a, b, c), ()) if i < k - 1]
                if near:
                    d = np.linalg.norm(cand[:, None, :] - pts[None, near, :], axis=2)
                    ok &= (d >= min_dist).all(axis=1)
            idx = np.flatnonzero(ok)
            if idx.size:
                pts[k] = cand[idx[0]]
                place(k, pts[k])
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
    # flexible topology (src/structures.h:49-75 moldata/segdata): per residue
    # chain index (TERs before it), chain code and resSeq; per molecule the
    # all-flexible flag and segments (0-based residue indices in the molecule)
    chain_index: np.ndarray | None = None
    chain_code: np.ndarray | None = None
    res_seq: np.ndarray | None = None
    all_flexible: list | None = None
    segments: list | None = None
    torsion_v: np.ndarray | None = None      # (20, 20, 4) kcal/mol
    torsion_delta: np.ndarray | None = None  # (20, 20, 4) radians

    @property
    def flexible(self) -> bool:
        return bool(self.all_flexible and any(self.all_flexible)) or bool(self.segments and any(self.segments))

    def molecule_is_flexible(self, m) -> bool:
        return bool((self.all_flexible and self.all_flexible[m]) or (self.segments and self.segments[m]))

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
        """Residue pairs of one full evaluation: every inter-molecular pair
        plus every intra-molecular pair of flexible molecules."""
        n = self.lengths.astype(np.int64)
        pairs = int((n.sum() ** 2 - (n * n).sum()) // 2)
        for m in range(self.n_mol):
            if self.molecule_is_flexible(m):
                pairs += int(n[m] * (n[m] - 1) // 2)
        return pairs

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


def synthetic_torsions(seed: int):
    """Seeded 20x20x4 torsion table (V ~ U[0, 1.6] kcal/mol, phase ~
    U[0, 360) degrees), rounded as the text format stores it and converted
    to radians as src/TorsionalLookupMatrix.cpp:48 does."""
    rng = np.random.default_rng(seed)
    V = np.round(rng.uniform(0.0, 1.6, size=(20, 20, 4)), 6)
    deg = np.round(rng.uniform(0.0, 360.0, size=(20, 20, 4)), 6)
    return V, deg / 57.29577951308232


def _min_dist_ok(pts, others, lim=4.0):
    if len(others) == 0 or len(pts) == 0:
        return True
    d = np.linalg.norm(pts[:, None, :] - others[None, :, :], axis=2)
    return bool((d >= lim).all())


def _linker(last, u, n, rng, bond=3.8):
    """n linker residues after `last` along direction u: a zig-zag whose
    every other residue is offset perpendicular to u (bond length `bond`,
    bond angles about 120-140 degrees, never collinear)."""
    d = bond * np.sqrt(0.75)          # axial advance per bond
    h = bond * 0.5                    # perpendicular offset
    a = np.cross(u, rng.normal(size=3))
    a /= np.linalg.norm(a)
    b = np.cross(u, a)
    pts = []
    for t in range(n + 1):            # the last point is the next domain's first residue
        off = np.zeros(3)
        if t % 2 == 0:
            th = 0.9 * t
            off = h * (np.cos(th) * a + np.sin(th) * b)
        pts.append(last + d * (t + 1) * u + off)
    return np.array(pts)


def _angles_ok(pts, lim=0.995):
    v1 = pts[:-2] - pts[1:-1]
    v2 = pts[2:] - pts[1:-1]
    c = (v1 * v2).sum(1) / (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1))
    return bool((np.abs(c) < lim).all())


def linked_domains(n_domains, domain_len, linker_len, rng, bond=3.8):
    """One chain of compact domains joined by zig-zag linkers with bond
    length `bond`; returns (n, 3) float64 positions centred at 0."""
    pts = compact_chain(domain_len, rng) @ _random_rotation(rng).T
    chain = [pts]
    for _k in range(1, n_domains):
        done = False
        for _attempt in range(4000):
            prev = np.concatenate(chain)
            last = prev[-1]
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            lk = _linker(last, u, linker_len, rng, bond)
            linker, first = lk[:-1], lk[-1]
            dom = compact_chain(domain_len, rng) @ _random_rotation(rng).T
            dom = dom - dom[0] + first
            cand = np.concatenate([prev, linker, dom])
            if (_min_dist_ok(linker, prev[:-1]) and _min_dist_ok(dom, np.concatenate([prev, linker[:-1]]))
                    and _angles_ok(cand[len(prev) - 3:len(prev) + linker_len + 3])):
                chain += [linker, dom]
                done = True
                break
        if not done:
            raise RuntimeError("could not place a linked domain")
    out = np.concatenate(chain)
    if not _angles_ok(out, 0.99999):
        raise RuntimeError("collinear residues in a linked chain")
    return out - out.mean(axis=0)


def make_linked_system(name, n_domains, domain_len, linker_len, flex_frac, box, n_rep, t_min, t_max, seed,
                       translate_step=2.0, rotate_step=0.2, mc_steps=10000, exchange_every=1000,
                       extra_rigid=()) -> SystemSpec:
    """One flexible molecule of linked domains (plus optional rigid chains
    of the given lengths).  Each domain's flexible residues are its tails
    next to the linkers; every linker with the flanking flexible tails forms
    one segment."""
    rng = np.random.default_rng(seed)
    ff = synthetic_force_field(seed + 1)
    tv, td = synthetic_torsions(seed + 2)
    n_flex = int(round(flex_frac * domain_len))
    pts = linked_domains(n_domains, domain_len, linker_len, rng)
    segments = []
    stride = domain_len + linker_len
    for k in range(n_domains - 1):
        tail = n_flex if k == 0 else n_flex - n_flex // 2
        head = n_flex if k == n_domains - 2 else n_flex // 2
        a = k * stride + domain_len - tail
        b = (k + 1) * stride + head - 1
        segments.append(list(range(a, b + 1)))
    lengths = [len(pts)] + [int(n) for n in extra_rigid]
    mol_off = np.zeros(len(lengths) + 1, dtype=np.int32)
    mol_off[1:] = np.cumsum(lengths)
    N = int(mol_off[-1])
    types = rng.integers(0, 20, size=N).astype(np.int32)
    rel = np.zeros((N, 3), dtype=np.float32)
    radii = np.zeros(len(lengths))
    rel[: len(pts)] = pts.astype(np.float32)
    radii[0] = np.linalg.norm(pts, axis=1).max()
    for m, n in enumerate(extra_rigid, start=1):
        q = compact_chain(int(n), rng) @ _random_rotation(rng).T
        rel[mol_off[m]:mol_off[m + 1]] = q.astype(np.float32)
        radii[m] = np.linalg.norm(q, axis=1).max()
    centers = _place(lengths, radii, box, rng).astype(np.float32)
    res_seq = np.concatenate([np.arange(1, n + 1) for n in lengths]).astype(np.int32)
    spec = SystemSpec(name, ff, float(np.float32(box)), np.asarray(lengths, dtype=np.int64), mol_off, types, rel,
                      centers, n_rep, t_min, t_max, translate_step, rotate_step, mc_steps, exchange_every,
                      chain_index=np.zeros(N, dtype=np.int32), chain_code=np.full(N, ord("A"), dtype=np.int32),
                      res_seq=res_seq, all_flexible=[False] * len(lengths),
                      segments=[segments] + [[] for _ in extra_rigid], torsion_v=tv, torsion_delta=td)
    return spec


def config(k: int, seed: int | None = None) -> SystemSpec:
    """The synthetic system of benchmark configuration k (0, 1, 2 or 4)."""
    if k == 0:
        s = 1000 if seed is None else seed
        return make_system("pair_2x76", [76, 76], 150.0, 20, 288.0, 450.0, s, 2.0, 0.2, 10000, 1000)
    if k == 1:
        s = 1001 if seed is None else seed
        rng = np.random.default_rng(s + 7)
        lengths = rng.integers(20, 301, size=40)
        return make_system("crowded40", lengths, box_for_concentration(40, 400.0), 32, 288.0, 450.0, s,
                           2.0, 0.2, 50000, 1000)
    if k == 2:
        s = 1002 if seed is None else seed
        return make_linked_system("linked4x76", 4, 76, 6, 0.2, 300.0, 64, 288.0, 500.0, s, 2.0, 0.2, 10000, 1000)
    if k == 4:
        s = 1004 if seed is None else seed
        rng = np.random.default_rng(s + 7)
        lengths = np.exp(rng.uniform(np.log(20.0), np.log(2000.0), size=256)).astype(np.int64)
        return make_system("crowded256", lengths, box_for_concentration(256, 800.0), 64, 288.0, 450.0, s,
                           2.0, 0.2, 50000, 1000)
    raise ValueError(f"no synthetic system for configuration {k}")
