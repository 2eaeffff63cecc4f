"""Python handle over the C ABI (include/cgppd_hip.h), for the test and
benchmark harness.  All compute happens in libcgppd_hip.so on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import f32, f64, i32, ptr


class CgppdError(RuntimeError):
    pass


class Engine:
    """One device context: force field, box, system, R resident replicas."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.cgppd_create(C.byref(h), int(device))
        if rc != 0:
            raise CgppdError(f"cgppd_create(device={device}) failed with status {rc}")
        self.h = h
        self.n_mol = self.n_res = self.n_rep = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.cgppd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.cgppd_last_error(self.h)
            raise CgppdError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # set-up -------------------------------------------------------------
    def set_force_field(self, eps, sigma, charge):
        e, s, q = f32(eps).reshape(400), f32(sigma), f32(charge)
        self._check(self.lib.cgppd_set_force_field(self.h, ptr(e, C.c_float), ptr(s, C.c_float), ptr(q, C.c_float)),
                    "set_force_field")

    def set_box(self, box):
        self._check(self.lib.cgppd_set_box(self.h, float(box)), "set_box")

    def set_system(self, mol_off, types, n_noncrowder=None):
        mo, ty = i32(mol_off), i32(types)
        n_mol = len(mo) - 1
        nc = n_mol if n_noncrowder is None else int(n_noncrowder)
        self._check(self.lib.cgppd_set_system(self.h, n_mol, ptr(mo, C.c_int), ptr(ty, C.c_int), nc), "set_system")
        self.n_mol, self.n_res = n_mol, int(mo[-1])

    def set_torsion_table(self, v, delta):
        vv, dd = f64(v).reshape(-1), f64(delta).reshape(-1)
        self._check(self.lib.cgppd_set_torsion_table(self.h, ptr(vv, C.c_double), ptr(dd, C.c_double)),
                    "set_torsion_table")

    def set_topology(self, molecule, chain_code, chain_index, all_flexible=False, segments=()):
        code, idx = i32(chain_code), i32(chain_index)
        off, res = _lib._segments_csr(list(segments))
        self._check(self.lib.cgppd_set_topology(self.h, int(molecule), ptr(code, C.c_int), ptr(idx, C.c_int),
                                                int(bool(all_flexible)), len(segments), ptr(off, C.c_int),
                                                ptr(res, C.c_int)), "set_topology")

    def set_flexible_spec(self, spec):
        """Torsion table and the flexible molecules' topology of a spec
        (SystemSpec or dict with mol_off, chain_code, chain_index,
        all_flexible, segments, torsion_v, torsion_delta)."""
        g = spec if isinstance(spec, dict) else spec.__dict__
        self.set_torsion_table(g["torsion_v"], g["torsion_delta"])
        mol_off = np.asarray(g["mol_off"])
        for m in range(len(mol_off) - 1):
            if g["all_flexible"][m] or g["segments"][m]:
                a, b = int(mol_off[m]), int(mol_off[m + 1])
                self.set_topology(m, g["chain_code"][a:b], g["chain_index"][a:b], g["all_flexible"][m],
                                  g["segments"][m])

    def energy_internal(self, r=0):
        """internal_molecule_E(false) of replica r: kcal/mol LJ, DH (indirect
        neighbours, negated), bond, angle, torsion."""
        out = np.zeros(5)
        self._check(self.lib.cgppd_energy_internal(self.h, int(r), ptr(out, C.c_double)), "energy_internal")
        return out

    def set_replicas(self, centers, rel, temperature=None, translate_step=None, rotate_step=None, label=None,
                     seed=0):
        c, r = f32(centers), f32(rel)
        R = c.shape[0] if c.ndim == 3 else 1
        t = None if temperature is None else f64(temperature)
        ts = None if translate_step is None else f32(translate_step)
        rs = None if rotate_step is None else f32(rotate_step)
        lb = None if label is None else i32(label)
        self._check(self.lib.cgppd_set_replicas(self.h, R, ptr(c, C.c_float), ptr(r, C.c_float), ptr(t, C.c_double),
                                                ptr(ts, C.c_float), ptr(rs, C.c_float), ptr(lb, C.c_int),
                                                C.c_uint64(int(seed))), "set_replicas")
        self.n_rep = R

    def load_spec(self, spec, n_rep=None, seed=0, n_noncrowder=None):
        """Force field, box, topology and R identical replicas of a
        cgppd_amd.synth.SystemSpec (temperatures and steps from the spec);
        molecules from n_noncrowder on are crowders."""
        R = spec.n_rep if n_rep is None else n_rep
        self.set_force_field(spec.ff.eps, spec.ff.sigma, spec.ff.charge)
        self.set_box(spec.box)
        self.set_system(spec.mol_off, spec.types, n_noncrowder)
        if getattr(spec, "flexible", False):
            self.set_flexible_spec(spec)
        temps = spec.temperatures() if R == spec.n_rep else np.linspace(spec.t_min, spec.t_max, R)
        ts, rs = spec.step_sizes()
        if R != spec.n_rep:
            ts, rs = np.resize(ts, R), np.resize(rs, R)
        self.set_replicas(np.broadcast_to(spec.centers, (R,) + spec.centers.shape),
                          np.broadcast_to(spec.rel, (R,) + spec.rel.shape), temps, ts, rs, np.arange(R), seed)

    def upload_positions(self, r, xyz):
        x = f32(xyz)
        self._check(self.lib.cgppd_upload_positions(self.h, int(r), ptr(x, C.c_float)), "upload_positions")

    def update_molecule(self, r, m, xyz):
        x = f32(xyz)
        self._check(self.lib.cgppd_update_molecule(self.h, int(r), int(m), ptr(x, C.c_float)), "update_molecule")

    # energies -----------------------------------------------------------
    def _energy3(self, fn, r, what):
        lj, dh, tot = C.c_double(), C.c_double(), C.c_double()
        self._check(fn(self.h, int(r), C.byref(lj), C.byref(dh), C.byref(tot)), what)
        return lj.value, dh.value, tot.value

    def energy(self, r=0):
        return self._energy3(self.lib.cgppd_energy, r, "energy")

    def energy_nc(self, r=0):
        return self._energy3(self.lib.cgppd_energy_nc, r, "energy_nc")

    def energy_nc_all(self):
        out = np.zeros(self.n_rep)
        self._check(self.lib.cgppd_energy_nc_all(self.h, ptr(out, C.c_double)), "energy_nc_all")
        return out

    def energy_async(self, r=0):
        self._check(self.lib.cgppd_energy_async(self.h, int(r)), "energy_async")

    def energy_wait(self, r=0):
        return self._energy3(self.lib.cgppd_energy_wait, r, "energy_wait")

    def energy_all(self, matrix=False):
        R, M = self.n_rep, self.n_mol
        lj, dh, tot = np.zeros(R), np.zeros(R), np.zeros(R)
        mat = np.zeros((R, M, M)) if matrix else None
        self._check(self.lib.cgppd_energy_all(self.h, ptr(lj, C.c_double), ptr(dh, C.c_double),
                                              ptr(tot, C.c_double), ptr(mat, C.c_double)), "energy_all")
        return (lj, dh, tot, mat) if matrix else (lj, dh, tot)

    # MC -----------------------------------------------------------------
    def mc_prepare(self):
        self._check(self.lib.cgppd_mc_prepare(self.h), "mc_prepare")

    def mc_steps(self, n):
        self._check(self.lib.cgppd_mc_steps(self.h, int(n)), "mc_steps")

    def exchange(self):
        self._check(self.lib.cgppd_exchange(self.h), "exchange")

    def run(self, n_intervals, steps_per_interval, use_graph=True):
        self._check(self.lib.cgppd_run(self.h, int(n_intervals), int(steps_per_interval), int(bool(use_graph))), "run")

    def synchronize(self):
        self._check(self.lib.cgppd_synchronize(self.h), "synchronize")

    def stats(self):
        R = self.n_rep
        e, t = np.zeros(R), np.zeros(R)
        lab = np.zeros(R, dtype=np.int32)
        cnt = np.zeros((R, 3), dtype=np.int64)
        st = np.zeros(R, dtype=np.int64)
        ex = np.zeros(2, dtype=np.int64)
        self._check(self.lib.cgppd_get_replica_stats(self.h, ptr(e, C.c_double), ptr(t, C.c_double),
                                                     ptr(lab, C.c_int), ptr(cnt, C.c_int64), ptr(st, C.c_int64),
                                                     ptr(ex, C.c_int64)), "get_replica_stats")
        return {"energy": e, "temperature": t, "label": lab, "counters": cnt, "steps": st, "exchange": ex}

    def pair_evaluations(self):
        n = C.c_int64()
        self._check(self.lib.cgppd_get_pair_evaluations(self.h, C.byref(n)), "get_pair_evaluations")
        return n.value

    def coordinates(self, r):
        N, M = self.n_res, self.n_mol
        pos, rel, cen = np.zeros((N, 3), np.float32), np.zeros((N, 3), np.float32), np.zeros((M, 3), np.float32)
        self._check(self.lib.cgppd_get_coordinates(self.h, int(r), ptr(pos, C.c_float), ptr(rel, C.c_float),
                                                   ptr(cen, C.c_float)), "get_coordinates")
        return pos, rel, cen

    def energy_matrix(self, r):
        M = self.n_mol
        mat = np.zeros((M, M))
        self._check(self.lib.cgppd_get_energy_matrix(self.h, int(r), ptr(mat, C.c_double)), "get_energy_matrix")
        return mat

    # surrogate ------------------------------------------------------------
    PARAM_SHAPES = (("W1", (80, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("w3", (128,)),
                    ("b3", (1,)))

    @classmethod
    def flatten_params(cls, p):
        return np.concatenate([np.asarray(p[k], dtype=np.float32).reshape(-1) for k, _ in cls.PARAM_SHAPES])

    @classmethod
    def unflatten_params(cls, flat):
        out, o = {}, 0
        for k, shape in cls.PARAM_SHAPES:
            n = int(np.prod(shape))
            out[k] = np.asarray(flat[o:o + n], dtype=np.float32).reshape(shape)
            o += n
        return out

    def surrogate_init(self, seed=0):
        self._check(self.lib.cgppd_surrogate_init(self.h, C.c_uint64(int(seed))), "surrogate_init")

    def surrogate_set_params(self, params):
        flat = f32(self.flatten_params(params) if isinstance(params, dict) else params)
        self._check(self.lib.cgppd_surrogate_set_params(self.h, ptr(flat, C.c_float)), "surrogate_set_params")

    def surrogate_params(self):
        n = C.c_int()
        self.lib.cgppd_surrogate_param_count(C.byref(n))
        flat = np.zeros(n.value, dtype=np.float32)
        self._check(self.lib.cgppd_surrogate_get_params(self.h, ptr(flat, C.c_float)), "surrogate_get_params")
        return self.unflatten_params(flat)

    def surrogate_predict(self, x):
        xx = f32(x).reshape(-1, 80)
        y = np.zeros(len(xx), dtype=np.float32)
        self._check(self.lib.cgppd_surrogate_predict(self.h, ptr(xx, C.c_float), len(xx), ptr(y, C.c_float)),
                    "surrogate_predict")
        return y

    def surrogate_gradient(self, x, t):
        xx, tt = f32(x).reshape(-1, 80), f32(t)
        n = C.c_int()
        self.lib.cgppd_surrogate_param_count(C.byref(n))
        g = np.zeros(n.value, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.cgppd_surrogate_gradient(self.h, ptr(xx, C.c_float), ptr(tt, C.c_float), len(xx),
                                                      ptr(g, C.c_float), C.byref(loss)), "surrogate_gradient")
        return loss.value, self.unflatten_params(g)

    def surrogate_add_samples(self, x, t):
        xx, tt = f32(x).reshape(-1, 80), f32(t)
        self._check(self.lib.cgppd_surrogate_add_samples(self.h, ptr(xx, C.c_float), ptr(tt, C.c_float), len(xx)),
                    "surrogate_add_samples")

    def surrogate_collect(self):
        self._check(self.lib.cgppd_surrogate_collect(self.h), "surrogate_collect")

    def surrogate_sample_count(self):
        n = C.c_int64()
        self._check(self.lib.cgppd_surrogate_sample_count(self.h, C.byref(n)), "surrogate_sample_count")
        return n.value

    def surrogate_clear_samples(self):
        self._check(self.lib.cgppd_surrogate_clear_samples(self.h), "surrogate_clear_samples")

    def surrogate_train(self, steps, batch=1024, lr=1e-3, seed=0):
        loss = C.c_double()
        self._check(self.lib.cgppd_surrogate_train(self.h, int(steps), int(batch), float(lr), C.c_uint64(int(seed)),
                                                   C.byref(loss)), "surrogate_train")
        return loss.value

    def surrogate_features(self, r):
        M = self.n_mol
        x = np.zeros((M * (M - 1) // 2, 80), dtype=np.float32)
        self._check(self.lib.cgppd_surrogate_features(self.h, int(r), ptr(x, C.c_float)), "surrogate_features")
        return x

    def set_delayed_acceptance(self, on):
        self._check(self.lib.cgppd_set_delayed_acceptance(self.h, int(bool(on))), "set_delayed_acceptance")

    def stage1_rejections(self):
        out = np.zeros(self.n_rep, dtype=np.int64)
        self._check(self.lib.cgppd_get_stage1_rejections(self.h, ptr(out, C.c_int64)), "get_stage1_rejections")
        return out

    def set_launch(self, wg_per_cu):
        self._check(self.lib.cgppd_set_launch(self.h, int(wg_per_cu)), "set_launch")

    def set_replica_groups(self, groups):
        """MC steps of `groups` contiguous replica groups on streams of their own."""
        self._check(self.lib.cgppd_set_replica_groups(self.h, int(groups)), "set_replica_groups")

    def replica_groups(self):
        g = C.c_int()
        self._check(self.lib.cgppd_get_replica_groups(self.h, C.byref(g)), "get_replica_groups")
        return g.value

    LAYOUTS = {"auto": 0, "ilanes": 1, "jlanes": 2, "jlanes4": 3, "resident": 4}

    def set_pair_layout(self, layout):
        """Pair-kernel layout: "auto", "ilanes" or "jlanes" (or 0 / 1 / 2)."""
        code = self.LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        self._check(self.lib.cgppd_set_pair_layout(self.h, code), "set_pair_layout")

    def set_window(self, slots):
        """Speculative windows of MC steps: 0 automatic, 1 off, 2, 4 or 8 slots."""
        self._check(self.lib.cgppd_set_window(self.h, int(slots)), "set_window")

    def window(self):
        """Slots per window the next MC steps use (1: one step per launch pair)."""
        w = C.c_int()
        self._check(self.lib.cgppd_get_window(self.h, C.byref(w)), "get_window")
        return w.value

    def _pair_launch(self):
        lay, wg = C.c_int(), C.c_int()
        self._check(self.lib.cgppd_get_pair_launch(self.h, C.byref(lay), C.byref(wg)), "get_pair_launch")
        return lay.value, wg.value

    def pair_layout(self):
        """The layout the next pair-kernel launch uses ("ilanes" or "jlanes")."""
        names = {v: k for k, v in self.LAYOUTS.items()}
        return names[self._pair_launch()[0]]

    def pair_wg_per_cu(self):
        return self._pair_launch()[1]

    def autotune(self, launches=20):
        """Times the pair-kernel candidates on the pending proposals and keeps
        the fastest; returns {"layout", "wg_per_cu", "window", "ms"} (ms per
        MC step of all replicas) and the per-candidate table."""
        lay, wg, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.cgppd_autotune(self.h, int(launches), C.byref(lay), C.byref(wg), C.byref(ms)),
                    "autotune")
        names = {v: k for k, v in self.LAYOUTS.items()}
        table = np.zeros(21)
        self._check(self.lib.cgppd_autotune_table(self.h, ptr(table, C.c_double)), "autotune_table")
        timings = {f"{lay}/{w}": round(float(table[i * 7 + w - 2]) * 1e3, 2)
                   for i, lay in enumerate(("jlanes", "jlanes4", "ilanes")) for w in range(2, 9)
                   if table[i * 7 + w - 2] > 0}
        return {"layout": names[lay.value], "wg_per_cu": wg.value, "window": self.window(), "ms": ms.value,
                "us_by_candidate": timings}

    def set_repulsive_crowding(self, on):
        """(6/r)^12 below 15 A and no DH for pairs with a crowder molecule
        (the reference's REPULSIVE_CROWDING, src/Replica.cpp:416-483)."""
        self._check(self.lib.cgppd_set_repulsive_crowding(self.h, int(bool(on))), "set_repulsive_crowding")

    def set_rigid_potential_only(self, on):
        """Replica::calculate_rigid_potential_only (src/Replica.cpp:753-757):
        flexible molecules carry no internal energy."""
        self._check(self.lib.cgppd_set_rigid_potential_only(self.h, int(bool(on))), "set_rigid_potential_only")

    def set_image_shift(self, on):
        self._check(self.lib.cgppd_set_image_shift(self.h, int(bool(on))), "set_image_shift")

    def set_timing(self, on):
        self._check(self.lib.cgppd_set_timing(self.h, int(bool(on))), "set_timing")

    def last_kernel_ms(self):
        ms, n = C.c_double(), C.c_int64()
        self._check(self.lib.cgppd_last_kernel_ms(self.h, C.byref(ms), C.byref(n)), "last_kernel_ms")
        return ms.value, n.value


def model_constants():
    lib = _lib.load()
    a, b, c = C.c_float(), C.c_float(), C.c_float()
    lib.cgppd_model_constants(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def device_count() -> int:
    lib = _lib.load()
    n = C.c_int()
    lib.cgppd_device_count(C.byref(n))
    return n.value
