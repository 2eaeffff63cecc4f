import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
DATA = GOLDEN / "data"

# Expected totals of tests/testReferenceConformations.cpp:83-94 (kcal/mol LJ, DH)
REFERENCE_CONFORMATIONS = [
    (-0.0796731, -0.212202), (-1.30422, 0.264889), (-8.97201, -1.17831), (-5.81976, -1.67169),
    (-2.08631e-05, -5.76825e-05), (-4.74252, -0.749355), (-4.11734, -1.26302), (-9.08724, -1.44068),
    (-7.83475, -1.94273), (-7.34204, -1.06544),
]
# Kim 2008 values, tests/testReferenceConformations.cpp:119-130
KIM_RESULTS = [
    (-0.08092694, -0.21315828), (-1.32250603, 0.26608889), (-9.09456258, -1.18374210), (-5.90453581, -1.67963588),
    (-0.00002116, -0.00005795), (-4.81153585, -0.75302819), (-4.18365757, -1.26891119), (-9.22300158, -1.44730149),
    (-7.95237776, -1.95173336), (-7.44768961, -1.07043531),
]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcgppd_hip.so)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN / "ref_vectors.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def mj_ff():
    from cgppd_amd import formats
    return formats.load_force_field(DATA / "AminoAcids", DATA / "pair_potentials_MJ1996")


@pytest.fixture(scope="session")
def conformations(mj_ff):
    """The ten reference conformations: (mol_off, types, pos) with box 1000."""
    from cgppd_amd import formats
    out = []
    for i in range(1, 11):
        a = formats.read_pdb(DATA / f"conf{i}" / "1a.pdb", mj_ff.acids)
        b = formats.read_pdb(DATA / f"conf{i}" / "1b.pdb", mj_ff.acids)
        types = np.concatenate([a.types, b.types]).astype(np.int32)
        pos = np.concatenate([a.xyz, b.xyz]).astype(np.float32)
        out.append((np.array([0, len(a.types), len(types)], dtype=np.int32), types, pos))
    return out


def golden_system(golden, name):
    g = lambda k: golden[f"sys_{name}_{k}"]  # noqa: E731
    return dict(mol_off=g("mol_off"), types=g("types"), pos=g("pos"), box=float(g("box")[0]), eps=g("eps"),
                sigma=g("sigma"), charge=g("charge"), energy=g("energy"), matrix=g("matrix"))


GOLDEN_SYSTEMS = ["cfg0", "mixed8", "mixed8wrap", "tight4"]

FLEX_CASES = ["ala8", "angiotensin", "ubq_seg", "conf1_flex", "diubq", "linked", "linked_mixed"]
FLEX_ENERGY_CASES = ["ala8", "angiotensin", "ubq_seg", "conf1_flex", "linked", "linked_mixed"]


@pytest.fixture(scope="session")
def flex_golden():
    with np.load(GOLDEN / "flex_vectors.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def flex_case(fg, name):
    """A flexible-linker golden case as a system dict (inputs only)."""
    g = lambda k: fg[f"{name}_{k}"]  # noqa: E731
    mol_off = g("mol_off")
    M = len(mol_off) - 1
    segs = [[] for _ in range(M)]
    so, sr = g("seg_off"), g("seg_res")
    for k, m in enumerate(g("seg_mol")):
        segs[int(m)].append([int(v) for v in sr[so[k]:so[k + 1]]])
    return dict(name=name, mol_off=mol_off, types=g("types"), pos=g("pos"), box=float(g("box")[0]),
                chain_code=g("chain_code"), chain_index=g("chain_index"), res_seq=g("res_seq"),
                all_flexible=[int(v) for v in g("all_flexible")], segments=segs, eps=g("eps"), sigma=g("sigma"),
                charge=g("charge"), torsion_v=g("torsion_v"), torsion_delta=g("torsion_delta"))


@pytest.fixture(scope="session")
def gpu_engine_factory():
    """Creates engines on device 0; skips when this machine has no GPU.  A
    missing library raises: the product has no CPU fallback."""
    from cgppd_amd import engine
    if engine.device_count() < 1:
        pytest.skip("no GPU visible")

    def make():
        return engine.Engine(0)

    return make
