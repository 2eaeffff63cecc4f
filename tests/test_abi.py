"""The C-ABI library loads and exports every symbol include/*.h declares;
argument checking works without a GPU; the data-format readers and the
synthetic systems behave as documented."""
import ctypes as C
import re

import numpy as np
import pytest

from conftest import DATA, ROOT


def declared_symbols(header):
    text = (ROOT / "include" / header).read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(cgppd_\w+)\s*\(", text, re.M)))


def test_hip_library_exports_every_declared_symbol():
    from cgppd_amd import _lib
    lib = C.CDLL(str(_lib.HIP_LIB_PATH))
    names = declared_symbols("cgppd_hip.h")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == names


def test_null_and_range_arguments_rejected():
    from cgppd_amd import _lib
    lib = _lib.load()
    assert lib.cgppd_create(None, 0) == -1
    assert lib.cgppd_destroy(None) == -1
    assert lib.cgppd_device_count(None) == -1
    assert lib.cgppd_set_box(None, 10.0) == -1
    assert lib.cgppd_get_sizes(None, None, None, None) == -1
    assert lib.cgppd_last_error(None) == b"null context"


def test_create_without_device_fails_cleanly():
    from cgppd_amd import engine
    if engine.device_count() > 0:
        pytest.skip("a GPU is visible here")
    with pytest.raises(engine.CgppdError):
        engine.Engine(0)


def test_model_constants_match_reference_build(golden):
    from cgppd_amd import engine
    rt, dhc, r0 = engine.model_constants()
    assert np.float32(rt) == golden["consts"][0]
    assert np.float32(dhc) == golden["consts"][1]
    assert np.float32(r0) == golden["consts"][2]


def test_amino_acid_file(mj_ff):
    names = [a.sname for a in mj_ff.acids]
    assert names[:5] == ["Ala", "Arg", "Asn", "Asp", "Cys"]
    assert names[18:] == ["His", "Ile"]
    assert mj_ff.sigma[0] == np.float32(5.0) and mj_ff.charge[1] == 1.0 and mj_ff.charge[18] == np.float32(0.5)
    assert mj_ff.index("LYS") == 9 and mj_ff.index("K") == 9 and mj_ff.index("Xyz") == -1


def test_pair_table_symmetric_and_placed(mj_ff):
    eps = mj_ff.eps
    np.testing.assert_array_equal(eps, eps.T)
    cys, met, pro = mj_ff.index("Cys"), mj_ff.index("Met"), mj_ff.index("Pro")
    assert eps[cys, cys] == np.float32(-5.44)
    assert eps[cys, met] == np.float32(-4.99)
    assert eps[pro, pro] == np.float32(-1.75)


def test_pdb_reader(mj_ff):
    from cgppd_amd import formats
    m = formats.read_pdb(DATA / "conf1" / "1a.pdb", mj_ff.acids)
    assert len(m.types) == 24 and m.chain_count == 1
    assert m.types[0] == mj_ff.index("TYR")
    np.testing.assert_array_equal(m.xyz[0], np.float32([-8.846, -3.579, 20.533]))


def test_format_round_trip(tmp_path):
    from cgppd_amd import formats, synth
    spec = synth.make_system("rt", [15, 22], 80.0, 1, 300.0, 300.0, 2)
    formats.write_amino_acids(tmp_path / "aa", spec.ff.acids)
    formats.write_pair_table(tmp_path / "pairs", spec.ff.acids, spec.ff.eps)
    ff = formats.load_force_field(tmp_path / "aa", tmp_path / "pairs")
    np.testing.assert_array_equal(ff.eps, spec.ff.eps)
    np.testing.assert_array_equal(ff.sigma, spec.ff.sigma)
    pos = spec.positions()[:15]
    formats.write_pdb(tmp_path / "m.pdb", ff.acids, spec.types[:15], pos)
    back = formats.read_pdb(tmp_path / "m.pdb", ff.acids)
    np.testing.assert_array_equal(back.types, spec.types[:15])
    np.testing.assert_allclose(back.xyz, pos, atol=5e-4)


def test_synthetic_chain_geometry():
    from cgppd_amd import synth
    rng = np.random.default_rng(3)
    for n in (20, 76, 300):
        pts = synth.compact_chain(n, rng)
        bonds = np.linalg.norm(np.diff(pts, axis=0), axis=1)
        np.testing.assert_allclose(bonds, 3.8, rtol=1e-9)
        d = np.linalg.norm(pts[:, None] - pts[None], axis=2)
        iu = np.triu_indices(n, 2)
        assert d[iu].min() >= 4.0 - 1e-9
        rg = np.sqrt(((pts - pts.mean(0)) ** 2).sum(1).mean())
        assert 0.55 * 12.0 * (n / 76) ** 0.38 < rg < 1.1 * 12.0 * (n / 76) ** 0.38


def test_synthetic_configs_shapes():
    from cgppd_amd import synth
    s0 = synth.config(0)
    assert list(s0.lengths) == [76, 76] and s0.box == 150.0 and s0.n_rep == 20
    assert s0.pairs_full() == 5776
    T = s0.temperatures()
    assert T[0] == np.float32(288.0) and abs(T[-1] - 450.0) < 1e-3
    assert np.all(np.diff(np.log(T)) > 0)
    assert synth.box_for_concentration(40, 400.0) == pytest.approx(549.6, abs=0.1)
    assert synth.box_for_concentration(256, 800.0) == pytest.approx(810.0, abs=0.1)
    charged = {i for i, a in enumerate(s0.ff.acids) if a.charge != 0}
    assert {s0.ff.acids[i].symbol for i in charged} == {"K", "R", "D", "E", "H"}
