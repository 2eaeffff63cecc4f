"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py -- as the checker and as the timed CPU
baseline -- and by nothing in the product (cgppd_amd/).

Parity pins (see tests/test_oracle.py):
  orc_energy_ref       ten reference conformations (tests/testReferenceConformations.cpp:83-94),
                       Kim 2008 comparison (:119-189), bit-exact vs oracle/_ref
  orc_energy_fp64      closed-form two-residue cases, agreement with orc_energy_ref
  orc_philox4x32_10    Random123 known-answer vectors
  moves                bit-exact vs oracle/_ref (reference Quaternion/Vector3f),
                       rigid-body invariants
  orc_metropolis       acceptance frequencies on a grid of uniforms against
                       min(1, exp(-dE kcal / (R T))) with R and kcal from
                       scipy.constants; exact stationary distributions of
                       small state spaces equal to their Boltzmann weights
  orc_remc_exchange    acceptance frequency against min(1, exp(delta)),
                       swaps temperatures only
  orc_remc_mc_steps_fast  energies against orc_energy_fp64; the decisions of
                       orc_remc_mc_steps on the same streams
  flexible linkers     topology (oracle/topology.py), bonded and intra-molecular
                       energies against oracle/_ref/ref_flex_driver and
                       tests/testMolecule.cpp / tests/testReplica.cpp values;
                       flexible moves against tests/testMolecule.cpp:146-283
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_DRIVER = HERE / "_ref" / "ref_driver"

_i32p = C.POINTER(C.c_int)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class FlexState(C.Structure):
    _fields_ = [(name, _i32p) for name in (
        "mol_flexible", "tag", "bond_off", "bond", "angle_off", "angle", "tors_off", "tors", "ind_off", "ind",
        "local_off", "local", "crank_off", "crank", "flexr_off", "flexr", "nb_off", "nb", "br_off", "br")] + [
        ("tor_v", _f64p), ("tor_d", _f64p)]


class RemcState(C.Structure):
    _fields_ = [
        ("n_mol", C.c_int), ("n_res", C.c_int), ("n_rep", C.c_int),
        ("mol_off", _i32p), ("type", _i32p),
        ("charge", _f32p), ("sigma", _f32p), ("eps", _f32p),
        ("box", C.c_float),
        ("seed", C.c_uint64),
        ("center", _f32p), ("rel", _f32p), ("pos", _f32p),
        ("temperature", _f64p), ("translate_step", _f32p), ("rotate_step", _f32p),
        ("energy", _f64p), ("emat", _f64p), ("counters", _i64p), ("step", _i64p),
        ("label", _i32p), ("rank_of", _i32p), ("rep_at_rank", _i32p),
        ("exchange_round", _i64p), ("exchange_stats", _i64p),
        ("flex", C.POINTER(FlexState)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        sysargs = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        for name in ("orc_energy_ref", "orc_energy_fp64"):
            getattr(L, name).argtypes = sysargs + [_f64p, _f64p]
            getattr(L, name).restype = None
        L.orc_energy_nc_ref.argtypes = sysargs + [_f64p, _f64p]
        L.orc_energy_pair_ref.argtypes = [_i32p, C.c_int, C.c_int, _i32p, _f32p, _f32p, _f32p, _f32p, C.c_float,
                                          _f64p, _f64p]
        L.orc_energy_matrix_ref.argtypes = sysargs + [_f64p]
        L.orc_set_repulsive_crowding.argtypes = [C.c_int]
        L.orc_set_repulsive_crowding.restype = None
        for name in ("orc_rt_to_kcal", "orc_dh_factor", "orc_r0_factor"):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = []
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_mc_proposal.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_apply_move.argtypes = [C.c_int, _f64p, C.c_float, C.c_double, C.c_double, C.c_float, _f32p, _f32p,
                                     C.c_int, _f32p, _f32p, _f32p]
        L.orc_remc_init.argtypes = [C.POINTER(RemcState)]
        L.orc_remc_mc_steps.argtypes = [C.POINTER(RemcState), C.c_int]
        L.orc_remc_exchange.argtypes = [C.POINTER(RemcState)]
        L.orc_remc_reference_run.argtypes = [C.POINTER(RemcState), C.c_int, C.c_int]
        L.orc_remc_reference_run.restype = C.c_double
        L.orc_remc_refresh_fast.argtypes = [C.POINTER(RemcState)]
        L.orc_remc_mc_steps_fast.argtypes = [C.POINTER(RemcState), C.c_int, C.c_int]
        L.orc_energy_fast.argtypes = [C.POINTER(RemcState), C.c_int]
        L.orc_energy_fast.restype = C.c_double
        L.orc_metropolis.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_metropolis.restype = C.c_int
        L.orc_metropolis_ref.argtypes = [C.c_float, C.c_float, C.c_double]
        L.orc_metropolis_ref.restype = C.c_int
        for name in ("orc_internal_ref", "orc_internal_fp64"):
            getattr(L, name).argtypes = [C.POINTER(FlexState), C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p, _f32p,
                                         C.c_float, C.c_int, _f64p]
            getattr(L, name).restype = None
        L.orc_internal_total.argtypes = [_f64p]
        L.orc_internal_total.restype = C.c_double
        L.orc_bond_term.argtypes = [_f32p, _f32p, C.c_float]
        L.orc_bond_term.restype = C.c_double
        L.orc_angle_term.argtypes = [_f32p, _f32p, _f32p]
        L.orc_angle_term.restype = C.c_double
        L.orc_torsion_term.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, _f64p, _f64p]
        L.orc_torsion_term.restype = C.c_double
        L.orc_local_translate.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, _f32p, C.c_float]
        L.orc_crankshaft.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_float]
        L.orc_flex_move.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, _f64p, C.c_double, C.c_double, C.c_int,
                                    _i32p, C.c_float]
        L.orc_mc_proposal_flex.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_int, _i32p, _i32p, _i32p, _f64p,
                                           _f64p, C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _sys(mol_off, types, charge, sigma, eps, pos, box):
    mo = np.ascontiguousarray(mol_off, dtype=np.int32)
    ty = np.ascontiguousarray(types, dtype=np.int32)
    q = np.ascontiguousarray(charge, dtype=np.float32)
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    e = np.ascontiguousarray(eps, dtype=np.float32).reshape(400)
    x = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    return mo, ty, q, s, e, x, float(box)


def constants():
    L = lib()
    return L.orc_rt_to_kcal(), L.orc_dh_factor(), L.orc_r0_factor()


def total(lj, dh):
    rt, dhc, _ = constants()
    return dh * float(np.float32(dhc)) + lj * float(np.float32(rt))


def energy_ref(mol_off, types, charge, sigma, eps, pos, box):
    """Reference semantics, raw (LJ, DH)."""
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    lj, dh = C.c_double(), C.c_double()
    lib().orc_energy_ref(len(mo) - 1, _p(mo, C.c_int), _p(ty, C.c_int), _p(q, C.c_float), _p(s, C.c_float),
                         _p(e, C.c_float), _p(x, C.c_float), L, C.byref(lj), C.byref(dh))
    return lj.value, dh.value


def energy_nc_ref(n_nc, mol_off, types, charge, sigma, eps, pos, box):
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    lj, dh = C.c_double(), C.c_double()
    lib().orc_energy_nc_ref(int(n_nc), _p(mo, C.c_int), _p(ty, C.c_int), _p(q, C.c_float), _p(s, C.c_float),
                            _p(e, C.c_float), _p(x, C.c_float), L, C.byref(lj), C.byref(dh))
    return lj.value, dh.value


def energy_pair_ref(mol_off, a, b, types, charge, sigma, eps, pos, box):
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    lj, dh = C.c_double(), C.c_double()
    lib().orc_energy_pair_ref(_p(mo, C.c_int), int(a), int(b), _p(ty, C.c_int), _p(q, C.c_float),
                              _p(s, C.c_float), _p(e, C.c_float), _p(x, C.c_float), L, C.byref(lj), C.byref(dh))
    return lj.value, dh.value


def set_repulsive_crowding(first_crowder_residue):
    """Repulsive crowding for every energy function and the MC oracle: pairs
    with a residue at or past first_crowder_residue (the crowder molecules)
    take (6/r)^12 of LJ below 15 A and no DH (src/Replica.cpp:416-483).
    None switches it off, the reference's default (src/definitions.h:18)."""
    lib().orc_set_repulsive_crowding(-1 if first_crowder_residue is None else int(first_crowder_residue))


def energy_fp64(mol_off, types, charge, sigma, eps, pos, box):
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    lj, dh = C.c_double(), C.c_double()
    lib().orc_energy_fp64(len(mo) - 1, _p(mo, C.c_int), _p(ty, C.c_int), _p(q, C.c_float), _p(s, C.c_float),
                          _p(e, C.c_float), _p(x, C.c_float), L, C.byref(lj), C.byref(dh))
    return lj.value, dh.value


def energy_matrix_ref(mol_off, types, charge, sigma, eps, pos, box):
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    M = len(mo) - 1
    mat = np.zeros((M, M))
    lib().orc_energy_matrix_ref(M, _p(mo, C.c_int), _p(ty, C.c_int), _p(q, C.c_float), _p(s, C.c_float),
                                _p(e, C.c_float), _p(x, C.c_float), L, _p(mat, C.c_double))
    return mat


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, out)
    return list(out)


def metropolis(delta, temperature, u, reference=False):
    """The oracle's Metropolis decision (0 downhill, 1 accepted by the
    Boltzmann test, 2 rejected); reference=True: float delta and exponent."""
    f = lib().orc_metropolis_ref if reference else lib().orc_metropolis
    return int(f(float(delta), float(temperature), float(u)))


def proposal(seed, r, step, n_mol):
    m, rot = C.c_int(), C.c_int()
    axis = (C.c_double * 3)()
    u = C.c_double()
    lib().orc_mc_proposal(C.c_uint64(seed), int(r), int(step), int(n_mol), C.byref(m), C.byref(rot), axis,
                          C.byref(u))
    return m.value, rot.value, list(axis), u.value


def apply_move(is_rotate, axis_raw, tstep, rcos, rsin, box, c_in, rel_in):
    """One rigid move of n residues; returns (c_out, rel_out, pos_out)."""
    ax = np.ascontiguousarray(axis_raw, dtype=np.float64)
    ci = np.ascontiguousarray(c_in, dtype=np.float32)
    ri = np.ascontiguousarray(rel_in, dtype=np.float32).reshape(-1, 3)
    co = np.zeros(3, dtype=np.float32)
    ro = np.zeros_like(ri)
    po = np.zeros_like(ri)
    lib().orc_apply_move(int(is_rotate), _p(ax, C.c_double), float(tstep), float(rcos), float(rsin), float(box),
                         _p(ci, C.c_float), _p(co, C.c_float), len(ri), _p(ri, C.c_float), _p(ro, C.c_float),
                         _p(po, C.c_float))
    return co, ro, po


class Flex:
    """The flexible topology of a system as the C oracle takes it (global
    residue indices, caller order), built by oracle/topology.py.  `system`
    is a SystemSpec or a dict with mol_off, types, chain_code, chain_index,
    res_seq, all_flexible, segments; tor_v/tor_d the (20, 20, 4) table."""

    def __init__(self, system, tor_v, tor_d):
        from . import topology

        d = topology._get(system)
        mol_off = np.asarray(d["mol_off"], dtype=np.int64)
        M = len(mol_off) - 1
        self.graphs = topology.molecule_graphs(d)
        uids = topology.assign_uids(d, self.graphs)
        N = int(mol_off[-1])
        tag = np.zeros((N, 4), dtype=np.int32)
        tag[:, :3] = uids
        tag[:, 3] = np.asarray(d["res_seq"], dtype=np.int32)
        self.uids = uids
        self.mol_flexible = np.array([1 if (d["all_flexible"][m] or d["segments"][m]) else 0 for m in range(M)],
                                     dtype=np.int32)
        lists = {k: ([0], []) for k in ("bond", "angle", "tors", "ind", "local", "crank", "flexr")}
        nb_off, nb, br_off, br = [0], [], [0], []
        for m, g in enumerate(self.graphs):
            a = int(mol_off[m])
            flexible = bool(self.mol_flexible[m])
            n_m = int(mol_off[m + 1]) - a
            # crankshaft needs the residues before and after in the caller's
            # order (src/Molecule.cpp:306,318): residues at the molecule's ends
            # are left out (the reference would read outside the molecule)
            crank = [v for v in g.crankshaft if 0 < v < n_m - 1]
            src = {"bond": g.bonds, "angle": g.angles, "tors": g.torsions, "ind": g.indirect, "local": g.local,
                   "crank": crank, "flexr": g.flex}
            for k, items in src.items():
                off, flat = lists[k]
                if flexible:
                    for it in items:
                        flat.extend([a + v for v in it] if isinstance(it, tuple) else [a + it])
                off.append(len(flat) // {"bond": 2, "angle": 3, "tors": 4, "ind": 2}.get(k, 1))
            if flexible:
                for ri in g.flex:
                    for n_ in g.neighbours(ri):
                        nb.append(a + n_)
                        br.extend(a + v for v in g.flex_branch(ri, n_))
                        br_off.append(len(br))
                    nb_off.append(len(nb))
        arr = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))  # noqa: E731
        self.arrays = {"mol_flexible": arr(self.mol_flexible), "tag": arr(tag)}
        for k, (off, flat) in lists.items():
            self.arrays[f"{k}_off"] = arr(off)
            self.arrays[k] = arr(flat if flat else [0])
        self.arrays.update(nb_off=arr(nb_off), nb=arr(nb if nb else [0]), br_off=arr(br_off),
                           br=arr(br if br else [0]))
        self.tor_v = np.ascontiguousarray(tor_v, dtype=np.float64).reshape(-1)
        self.tor_d = np.ascontiguousarray(tor_d, dtype=np.float64).reshape(-1)
        a_ = self.arrays
        self.st = FlexState(*[_p(a_[n], C.c_int) for n, _t in FlexState._fields_[:-2]],
                            _p(self.tor_v, C.c_double), _p(self.tor_d, C.c_double))


def internal(flex, m, mol_off, types, charge, sigma, eps, pos, box, mode, fp64=False):
    """Internal energy of molecule m: (raw LJ, raw DH, raw bond, angle
    product, torsion).  mode 0 Molecule::E(true), 1 Molecule::E(false),
    2 the GPU kernel's intra-molecular pairs."""
    mo, ty, q, s, e, x, L = _sys(mol_off, types, charge, sigma, eps, pos, box)
    out = np.zeros(5)
    fn = lib().orc_internal_fp64 if fp64 else lib().orc_internal_ref
    fn(C.byref(flex.st), int(m), _p(mo, C.c_int), _p(ty, C.c_int), _p(q, C.c_float), _p(s, C.c_float),
       _p(e, C.c_float), _p(x, C.c_float), L, int(mode), _p(out, C.c_double))
    return out


def internal_total(v):
    a = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_internal_total(_p(a, C.c_double))


def local_translate(pos, rel, c, ri, v, box):
    pos, rel, c = (np.ascontiguousarray(a, dtype=np.float32).copy() for a in (pos, rel, c))
    vv = np.ascontiguousarray(v, dtype=np.float32)
    lib().orc_local_translate(_p(pos, C.c_float), _p(rel, C.c_float), _p(c, C.c_float), len(pos), int(ri),
                              _p(vv, C.c_float), float(box))
    return pos, rel, c


def crankshaft(pos, rel, c, ri, angle, flip, box):
    pos, rel, c = (np.ascontiguousarray(a, dtype=np.float32).copy() for a in (pos, rel, c))
    lib().orc_crankshaft(_p(pos, C.c_float), _p(rel, C.c_float), _p(c, C.c_float), len(pos), int(ri),
                         float(np.cos(angle / 2.0)), float(np.sin(angle / 2.0)), int(flip), float(box))
    return pos, rel, c


def flex_move(pos, rel, c, ri, axis, angle, branch, box):
    pos, rel, c = (np.ascontiguousarray(a, dtype=np.float32).copy() for a in (pos, rel, c))
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    b = np.ascontiguousarray(branch, dtype=np.int32)
    lib().orc_flex_move(_p(pos, C.c_float), _p(rel, C.c_float), _p(c, C.c_float), len(pos), int(ri),
                        _p(ax, C.c_double), float(np.cos(angle / 2.0)), float(np.sin(angle / 2.0)), len(b),
                        _p(b, C.c_int), float(box))
    return pos, rel, c


def proposal_flex(seed, r, step, mol_flexible):
    mf = np.ascontiguousarray(mol_flexible, dtype=np.int32)
    m, kind = C.c_int(), C.c_int()
    axis = (C.c_double * 3)()
    u = C.c_double()
    extra = (C.c_uint32 * 4)()
    lib().orc_mc_proposal_flex(C.c_uint64(seed), int(r), int(step), len(mf), _p(mf, C.c_int), C.byref(m),
                               C.byref(kind), axis, C.byref(u), extra)
    return m.value, kind.value, list(axis), u.value, list(extra)


class Remc:
    """Owns numpy arrays for an orc_remc state (all replicas).  Flexible
    systems (spec.flexible) carry their topology (Flex)."""

    def __init__(self, spec, n_rep=None, seed=0, temperatures=None, steps=None):
        R = spec.n_rep if n_rep is None else n_rep
        M, N = spec.n_mol, spec.n_res
        self.mol_off = np.ascontiguousarray(spec.mol_off, dtype=np.int32)
        self.types = np.ascontiguousarray(spec.types, dtype=np.int32)
        self.charge = np.ascontiguousarray(spec.ff.charge, dtype=np.float32)
        self.sigma = np.ascontiguousarray(spec.ff.sigma, dtype=np.float32)
        self.eps = np.ascontiguousarray(spec.ff.eps, dtype=np.float32).reshape(400)
        self.center = np.ascontiguousarray(np.broadcast_to(spec.centers, (R, M, 3)), dtype=np.float32).copy()
        self.rel = np.ascontiguousarray(np.broadcast_to(spec.rel, (R, N, 3)), dtype=np.float32).copy()
        self.pos = np.zeros((R, N, 3), dtype=np.float32)
        if temperatures is None:
            temperatures = spec.temperatures() if R == spec.n_rep else np.linspace(spec.t_min, spec.t_max, R)
        self.temperature = np.ascontiguousarray(temperatures, dtype=np.float64).copy()
        if steps is None:
            ts, rs = spec.step_sizes()
            if R != spec.n_rep:
                ts, rs = np.resize(ts, R), np.resize(rs, R)
        else:
            ts, rs = steps
        self.translate_step = np.ascontiguousarray(ts, dtype=np.float32).copy()
        self.rotate_step = np.ascontiguousarray(rs, dtype=np.float32).copy()
        self.energy = np.zeros(R)
        self.emat = np.zeros((R, M, M))
        self.counters = np.zeros((R, 3), dtype=np.int64)
        self.step = np.zeros(R, dtype=np.int64)
        self.label = np.arange(R, dtype=np.int32)
        order = np.argsort(self.temperature, kind="stable").astype(np.int32)
        self.rep_at_rank = np.ascontiguousarray(order)
        self.rank_of = np.zeros(R, dtype=np.int32)
        self.rank_of[order] = np.arange(R, dtype=np.int32)
        self.exchange_round = np.zeros(1, dtype=np.int64)
        self.exchange_stats = np.zeros(2, dtype=np.int64)
        self.st = RemcState(M, N, R, _p(self.mol_off, C.c_int), _p(self.types, C.c_int),
                            _p(self.charge, C.c_float), _p(self.sigma, C.c_float), _p(self.eps, C.c_float),
                            float(spec.box), C.c_uint64(int(seed)),
                            _p(self.center, C.c_float), _p(self.rel, C.c_float), _p(self.pos, C.c_float),
                            _p(self.temperature, C.c_double), _p(self.translate_step, C.c_float),
                            _p(self.rotate_step, C.c_float), _p(self.energy, C.c_double), _p(self.emat, C.c_double),
                            _p(self.counters, C.c_int64), _p(self.step, C.c_int64), _p(self.label, C.c_int),
                            _p(self.rank_of, C.c_int), _p(self.rep_at_rank, C.c_int),
                            _p(self.exchange_round, C.c_int64), _p(self.exchange_stats, C.c_int64), None)
        self.flex = None
        if getattr(spec, "flexible", False):
            self.flex = Flex(spec, spec.torsion_v, spec.torsion_delta)
            self.st.flex = C.pointer(self.flex.st)
        lib().orc_remc_init(C.byref(self.st))

    def mc_steps(self, n):
        lib().orc_remc_mc_steps(C.byref(self.st), int(n))

    def exchange(self):
        lib().orc_remc_exchange(C.byref(self.st))

    def refresh_fast(self):
        lib().orc_remc_refresh_fast(C.byref(self.st))

    def mc_steps_fast(self, n, threads=1):
        lib().orc_remc_mc_steps_fast(C.byref(self.st), int(n), int(threads))

    def energy_fast(self, r):
        return lib().orc_energy_fast(C.byref(self.st), int(r))

    def reference_run(self, n, threads):
        return lib().orc_remc_reference_run(C.byref(self.st), int(n), int(threads))
