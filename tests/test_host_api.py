"""The C++ host API (libcgppd_host.so): the reference's loaders, argfile/CLI
grammar, Replica::E on the CPU, and -- on a GPU -- EonDevice, EonDeviceNC,
EonDeviceAsync + SumGridResults and Simulation::run through the C ABI."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import oracle
from conftest import DATA, REFERENCE_CONFORMATIONS, ROOT

HOST = ROOT / "cgppd_amd" / "libcgppd_host.so"
CLI = ROOT / "cgppd_amd" / "cgppd"


@pytest.fixture(scope="module")
def host():
    if not HOST.exists():
        raise RuntimeError("libcgppd_host.so is not built: run python -m cgppd_amd.build")
    lib = C.CDLL(str(HOST))
    lib.cgppd_host_last_error.restype = C.c_char_p
    lib.cgppd_host_simulation_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_void_p)]
    for n in ("cgppd_host_simulation_destroy", "cgppd_host_simulation_sizes", "cgppd_host_replica_state",
              "cgppd_host_replica_E", "cgppd_host_setup_device", "cgppd_host_replica_EonDevice",
              "cgppd_host_simulation_run", "cgppd_host_replica_internal", "cgppd_host_replica_set_rigid_only"):
        getattr(lib, n).argtypes = None
    return lib


def write_config(tmp_path, spec, n_rep=4, mcsteps=400, resteps=4, seed=5, extra=""):
    from cgppd_amd import formats
    aa, pairs = tmp_path / "AminoAcids", tmp_path / "pairs"
    formats.write_amino_acids(aa, spec.ff.acids)
    formats.write_pair_table(pairs, spec.ff.acids, spec.ff.eps)
    lines = [f"mcsteps {mcsteps}", f"resteps {resteps}", f"replicas {n_rep}", f"boundary {spec.box:.6f}",
             "temperaturemax 450", "temperaturemin 288", "samplefrequency 100", "sampleafter 0",
             f"prefix {tmp_path.name}", f"seed {seed}", f"aminoacids {aa}", f"pairpotentials {pairs}", extra, "",
             "files", ""]
    pos = spec.positions()
    for m in range(spec.n_mol):
        a, b = spec.mol_off[m], spec.mol_off[m + 1]
        pdb = tmp_path / f"mol{m}.pdb"
        formats.write_pdb(pdb, spec.ff.acids, spec.types[a:b], pos[a:b])
        lines.append(f"t(0,0,0) r(0,0,0,0) {pdb} mol{m}")
    cfg = tmp_path / "config"
    cfg.write_text("\n".join(lines) + "\n")
    return cfg


def create(host, argv):
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    h = C.c_void_p()
    rc = host.cgppd_host_simulation_create(len(argv), arr, C.byref(h))
    if rc != 0:
        raise RuntimeError(host.cgppd_host_last_error().decode())
    return h


def test_loader_matches_python_reader(host, mj_ff):
    eps = np.zeros(400, np.float32)
    sig = np.zeros(20, np.float32)
    q = np.zeros(20, np.float32)
    host.cgppd_host_load_force_field(str(DATA / "AminoAcids").encode(), str(DATA / "pair_potentials_MJ1996").encode(),
                                     eps.ctypes.data_as(C.c_void_p), sig.ctypes.data_as(C.c_void_p),
                                     q.ctypes.data_as(C.c_void_p))
    np.testing.assert_array_equal(eps.reshape(20, 20), mj_ff.eps)
    np.testing.assert_array_equal(sig, mj_ff.sigma)
    np.testing.assert_array_equal(q, mj_ff.charge)


def test_missing_files_raise(host):
    rc = host.cgppd_host_load_force_field(b"/nonexistent/aa", b"/nonexistent/pairs", None, None, None)
    assert rc == -1 and b"Failed to open" in host.cgppd_host_last_error()
    with pytest.raises(RuntimeError, match="No configuration file"):
        create(host, ["cgppd"])
    with pytest.raises(RuntimeError, match="Failed to open config file"):
        create(host, ["cgppd", "-f", "/nonexistent/config"])


def test_argfile_and_cli_override(host, tmp_path):
    from cgppd_amd import synth
    spec = synth.make_system("host", [20, 35, 28], 90.0, 4, 288.0, 450.0, 41)
    cfg = write_config(tmp_path, spec, n_rep=4)
    h = create(host, ["cgppd", "-f", str(cfg), "-r", "6", "-x", "400"])
    n_rep, n_mol, n_res = C.c_int(), C.c_int(), C.c_int()
    host.cgppd_host_simulation_sizes(h, C.byref(n_rep), C.byref(n_mol), C.byref(n_res))
    assert (n_rep.value, n_mol.value, n_res.value) == (6, 3, 83)
    temps = []
    for r in range(6):
        off = np.zeros(4, np.int32)
        ty = np.zeros(83, np.int32)
        pos = np.zeros(83 * 3, np.float32)
        T, ts, rs = C.c_double(), C.c_float(), C.c_double()
        host.cgppd_host_replica_state(h, r, off.ctypes.data_as(C.c_void_p), ty.ctypes.data_as(C.c_void_p),
                                      pos.ctypes.data_as(C.c_void_p), C.byref(T), C.byref(ts), C.byref(rs))
        temps.append(T.value)
        np.testing.assert_array_equal(off, spec.mol_off)
        np.testing.assert_array_equal(ty, spec.types)
        np.testing.assert_allclose(pos.reshape(-1, 3), spec.positions(), atol=6e-4)  # PDB holds 3 decimals
        # src/Simulation.cpp:590-592 ladders
        assert ts.value == pytest.approx(0.5 * 4 ** (r / 5), rel=1e-6)
        assert rs.value == pytest.approx(0.2 * 2 ** (r / 5), rel=1e-6)
    assert temps[0] == pytest.approx(288.0) and temps[-1] == pytest.approx(400.0, rel=1e-6)
    host.cgppd_host_simulation_destroy(h)


def test_replica_E_matches_oracle_on_reference_conformation(host, tmp_path, mj_ff):
    cfg = tmp_path / "config"
    cfg.write_text("replicas 1\nboundary 1000\n"
                   f"aminoacids {DATA / 'AminoAcids'}\npairpotentials {DATA / 'pair_potentials_MJ1996'}\n\nfiles\n"
                   f"t(0,0,0) r(0,0,0,0) {DATA / 'conf3' / '1a.pdb'} a\n"
                   f"t(0,0,0) r(0,0,0,0) {DATA / 'conf3' / '1b.pdb'} b\n")
    h = create(host, ["cgppd", "-f", str(cfg)])
    lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
    host.cgppd_host_replica_E(h, 0, C.byref(lj), C.byref(dh), C.byref(tot))
    elj, edh = REFERENCE_CONFORMATIONS[2]
    rt, dhc, _ = oracle.constants()
    assert abs(lj.value - elj / np.float32(rt)) < 1e-5
    assert abs(dh.value - edh / np.float32(dhc)) < 1e-5
    host.cgppd_host_simulation_destroy(h)


def flexible_conf1_config(tmp_path):
    """tests/testReplica.cpp:30-58: data/conf1 with 1b's residues 73-76
    (1-based in the argfile) a flexible segment, box 118.4."""
    cfg = tmp_path / "config"
    cfg.write_text("replicas 1\nboundary 118.4\n"
                   f"aminoacids {DATA / 'AminoAcids'}\npairpotentials {DATA / 'pair_potentials_MJ1996'}\n"
                   f"torsions {DATA / 'torsional_pair_potentials'}\n\nfiles\n"
                   f"t(0,0,0) r(0,0,0,0) {DATA / 'conf1' / '1a.pdb'} a\n"
                   f"t(0,0,0) r(0,0,0,0) {DATA / 'conf1' / '1b.pdb'} b\n\nsegments\nb 73 74 75 76\n")
    return cfg


def raw_potential(lj, dh, bond, angle, torsion):
    """Potential(LJ, DH, bond, angle, torsion) in kcal/mol -> raw (src/Potential.cpp:116-126)."""
    rt, dhc, _ = oracle.constants()
    return np.array([lj / np.float32(rt), dh / np.float32(dhc), bond / (0.5 * 378), np.exp(angle / -10.0), torsion])


def test_flexible_replica_E_matches_reference_test(host, tmp_path):
    """(replica.E() + replica.internal_molecule_E(true)).almost_equal(expected)
    (tests/testReplica.cpp:81-82) on the C++ host API's CPU path."""
    h = create(host, ["cgppd", "-f", str(flexible_conf1_config(tmp_path))])
    lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
    host.cgppd_host_replica_E(h, 0, C.byref(lj), C.byref(dh), C.byref(tot))
    out = np.zeros(6)
    assert host.cgppd_host_replica_internal(h, 0, 1, out.ctypes.data_as(C.c_void_p)) == 0, \
        host.cgppd_host_last_error()
    got = np.array([lj.value + out[0], dh.value + out[1], out[2], out[3], out[4]])
    want = raw_potential(-0.068205, -0.240383, 0.180426, 1.555213, 5.420859)
    assert (np.abs(got - want) < 1e-5).all(), (got, want)
    rigid = raw_potential(-0.079673, -0.212202, 0.0, 0.0, 0.0)
    assert abs(lj.value - rigid[0]) < 1e-5 and abs(dh.value - rigid[1]) < 1e-5
    host.cgppd_host_simulation_destroy(h)


@pytest.mark.gpu
def test_flexible_device_entry_points(host, tmp_path, gpu_engine_factory):
    """tests/testReplica.cpp:84-91: EonDevice() + internal_molecule_E(false).total()
    == Approx(flexible total), EonDeviceNC() == Approx(rigid total), and the
    asynchronous path."""
    h = create(host, ["cgppd", "-f", str(flexible_conf1_config(tmp_path))])
    assert host.cgppd_host_setup_device(h, 0) == 0, host.cgppd_host_last_error()
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    assert host.cgppd_host_replica_EonDevice(h, 0, C.byref(a), C.byref(b), C.byref(c)) == 0, \
        host.cgppd_host_last_error()
    out = np.zeros(6)
    assert host.cgppd_host_replica_internal(h, 0, 0, out.ctypes.data_as(C.c_void_p)) == 0
    flexible = -0.068205 - 0.240383 + 0.180426 + 1.555213 + 5.420859
    approx = lambda x, y: abs(x - y) < 100 * np.finfo(np.float32).eps * (1 + max(abs(x), abs(y))) + 5e-6  # noqa: E731
    assert approx(a.value + out[5], flexible), (a.value + out[5], flexible)
    assert approx(c.value + out[5], flexible)
    assert approx(b.value, -0.079673 - 0.212202)
    host.cgppd_host_simulation_destroy(h)


@pytest.mark.gpu
def test_rigid_potential_only_override(host, tmp_path, gpu_engine_factory):
    """tests/testReplica.cpp:96-111: with calculate_rigid_potential_only set
    and the data sent again, E(), EonDevice(), EonDeviceNC() and the
    asynchronous path all give the rigid totals (LJ -0.079673, DH -0.212202);
    cleared again, EonDevice counts the flexible residues' pairs once more."""
    h = create(host, ["cgppd", "-f", str(flexible_conf1_config(tmp_path))])
    assert host.cgppd_host_setup_device(h, 0) == 0, host.cgppd_host_last_error()
    rigid = -0.079673 - 0.212202
    approx = lambda x, y: abs(x - y) < 100 * np.finfo(np.float32).eps * (1 + max(abs(x), abs(y))) + 5e-6  # noqa: E731
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    assert host.cgppd_host_replica_EonDevice(h, 0, C.byref(a), C.byref(b), C.byref(c)) == 0
    flex_pairs = a.value
    assert not approx(flex_pairs, rigid)  # the flexible tail's own pairs count
    assert host.cgppd_host_replica_set_rigid_only(h, 0, 1) == 0
    assert host.cgppd_host_replica_EonDevice(h, 0, C.byref(a), C.byref(b), C.byref(c)) == 0, \
        host.cgppd_host_last_error()
    lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
    host.cgppd_host_replica_E(h, 0, C.byref(lj), C.byref(dh), C.byref(tot))
    assert approx(tot.value, rigid)
    for v in (a.value, b.value, c.value):
        assert approx(v, rigid), (a.value, b.value, c.value)
    assert host.cgppd_host_replica_set_rigid_only(h, 0, 0) == 0
    assert host.cgppd_host_replica_EonDevice(h, 0, C.byref(a), C.byref(b), C.byref(c)) == 0
    assert abs(a.value - flex_pairs) <= 1e-9 * abs(flex_pairs) + 1e-12
    host.cgppd_host_simulation_destroy(h)


@pytest.mark.gpu
def test_device_entry_points_match_E(host, tmp_path, gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_system("hostgpu", [20, 35, 28, 60], 90.0, 4, 288.0, 450.0, 43)
    cfg = write_config(tmp_path, spec, n_rep=4)
    h = create(host, ["cgppd", "-f", str(cfg)])
    assert host.cgppd_host_setup_device(h, 0) == 0, host.cgppd_host_last_error()
    for r in range(4):
        lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
        host.cgppd_host_replica_E(h, r, C.byref(lj), C.byref(dh), C.byref(tot))
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        assert host.cgppd_host_replica_EonDevice(h, r, C.byref(a), C.byref(b), C.byref(c)) == 0
        for v in (a.value, b.value, c.value):
            assert abs(v - tot.value) <= 1e-5 * abs(tot.value) + 1e-9
    host.cgppd_host_simulation_destroy(h)


@pytest.mark.gpu
def test_simulation_run_on_device(host, tmp_path, gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_system("hostrun", [20, 35, 28], 90.0, 4, 288.0, 450.0, 47)
    cfg = write_config(tmp_path, spec, n_rep=4, mcsteps=400, resteps=4)
    h = create(host, ["cgppd", "-f", str(cfg)])
    e = np.zeros(4)
    t = np.zeros(4)
    rc = host.cgppd_host_simulation_run(h, e.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p))
    assert rc == 0, host.cgppd_host_last_error()
    for r in range(4):
        lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
        host.cgppd_host_replica_E(h, r, C.byref(lj), C.byref(dh), C.byref(tot))
        assert abs(e[r] - tot.value) <= 1e-5 * abs(tot.value) + 1e-6
    assert sorted(t) == pytest.approx(sorted([288.0 * (450.0 / 288.0) ** (k / 3) for k in range(4)]), rel=1e-6)
    out = ROOT / "output"
    host.cgppd_host_simulation_destroy(h)
    assert any(out.glob("*/acceptance_ratios"))


@pytest.mark.gpu
def test_cli_runs(tmp_path, gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_system("cli", [20, 35], 90.0, 2, 288.0, 450.0, 53)
    cfg = write_config(tmp_path, spec, n_rep=2, mcsteps=200, resteps=2)
    res = subprocess.run([str(CLI), "-f", str(cfg)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "2 replicas" in res.stdout
    assert (tmp_path / "output").exists()
    # src/Simulation.cpp:1181-1194: one line per molecule file after the header
    idx = list((tmp_path / "output").glob("*/fileindex"))
    assert len(idx) == 1
    lines = idx[0].read_text().splitlines()
    assert lines[0] == "index molecule_file_path crowder(Y/N)"
    assert len(lines) == 1 + spec.n_mol and all(l.split()[-1] == "N" for l in lines[1:])


def crowded_config(tmp_path):
    """Two chains and eight small crowders (a `crowders` section) with the
    argfile key `repulsivecrowding 1`."""
    from cgppd_amd import synth
    spec = synth.make_system("crowdmc", [40, 25, 1, 1, 3, 1, 2, 1, 1, 1], 60.0, 2, 288.0, 450.0, 14, 2.0, 0.2)
    n_nc = 2
    cfg = write_config(tmp_path, spec, n_rep=2, extra="repulsivecrowding 1")
    text = cfg.read_text().splitlines()
    at = next(i for i, line in enumerate(text) if f"mol{n_nc}.pdb" in line)
    cfg.write_text("\n".join(text[:at] + ["", "crowders"] + text[at:]) + "\n")
    return spec, n_nc, cfg


def host_replica_system(host, h, spec, r=0):
    n_res = int(spec.mol_off[-1])
    off = np.zeros(spec.n_mol + 1, np.int32)
    ty = np.zeros(n_res, np.int32)
    pos = np.zeros(n_res * 3, np.float32)
    T, ts, rs = C.c_double(), C.c_float(), C.c_double()
    host.cgppd_host_replica_state(h, r, off.ctypes.data_as(C.c_void_p), ty.ctypes.data_as(C.c_void_p),
                                  pos.ctypes.data_as(C.c_void_p), C.byref(T), C.byref(ts), C.byref(rs))
    return off, ty, pos.reshape(-1, 3)


def test_repulsive_crowding_host_E(host, tmp_path):
    """The CPU Replica::E under `repulsivecrowding 1` (src/Replica.cpp:432-483):
    crowder pairs as (6/r)^12 below 15 A without DH, as the oracle's crowding
    model gives on the same coordinates."""
    spec, n_nc, cfg = crowded_config(tmp_path)
    h = create(host, ["cgppd", "-f", str(cfg)])
    off, ty, pos = host_replica_system(host, h, spec)
    args = (off, ty, spec.ff.charge, spec.ff.sigma, spec.ff.eps, pos, spec.box)
    plain = oracle.total(*oracle.energy_ref(*args))
    try:
        oracle.set_repulsive_crowding(int(off[n_nc]))
        lj_w, dh_w = oracle.energy_ref(*args)
    finally:
        oracle.set_repulsive_crowding(None)
    lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
    host.cgppd_host_replica_E(h, 0, C.byref(lj), C.byref(dh), C.byref(tot))
    # the same reference-semantics arithmetic on both sides, summed in a
    # different order (molecule pairs vs residue runs): 1e-9 relative
    assert abs(lj.value - lj_w) <= 1e-9 * abs(lj_w) + 1e-12, (lj.value, lj_w)
    assert abs(dh.value - dh_w) <= 1e-9 * abs(dh_w) + 1e-12, (dh.value, dh_w)
    assert abs(tot.value - plain) > 1e-2
    host.cgppd_host_simulation_destroy(h)


@pytest.mark.gpu
def test_repulsive_crowding_argfile_key(host, tmp_path, gpu_engine_factory):
    """The argfile key `repulsivecrowding 1` (the reference's compile-time
    REPULSIVE_CROWDING, src/definitions.h:18) reaches the device: EonDevice
    of a system with a crowders section equals the oracle's energy with
    (6/r)^12 crowder pairs, and differs from the plain model."""
    spec, n_nc, cfg = crowded_config(tmp_path)
    h = create(host, ["cgppd", "-f", str(cfg)])
    assert host.cgppd_host_setup_device(h, 0) == 0, host.cgppd_host_last_error()
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    assert host.cgppd_host_replica_EonDevice(h, 0, C.byref(a), C.byref(b), C.byref(c)) == 0, \
        host.cgppd_host_last_error()
    off, ty, pos = host_replica_system(host, h, spec)
    args = (off, ty, spec.ff.charge, spec.ff.sigma, spec.ff.eps, pos, spec.box)
    plain = oracle.total(*oracle.energy_fp64(*args))
    try:
        oracle.set_repulsive_crowding(int(off[n_nc]))
        want = oracle.total(*oracle.energy_fp64(*args))
    finally:
        oracle.set_repulsive_crowding(None)
    # (force field through its text files: 1e-4)
    assert abs(a.value - want) <= 1e-4 * abs(want) + 1e-5, (a.value, want)
    assert abs(a.value - plain) > 1e-2, (a.value, plain)
    # the CPU Replica::E follows the same switch (src/Replica.cpp:432-483)
    lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
    host.cgppd_host_replica_E(h, 0, C.byref(lj), C.byref(dh), C.byref(tot))
    assert abs(tot.value - want) <= 1e-4 * abs(want) + 1e-5, (tot.value, want)
    host.cgppd_host_simulation_destroy(h)
