"""Statistical equivalence of the device REMC with a CPU REMC, and of
delayed acceptance with plain Metropolis, on the two-chain system of the
first benchmark configuration (2 x 76 residues, box 150 A, 20 replicas at
288-450 K, 2 A / 0.2 rad moves, an exchange every 1000 steps).

The reference's own REMC test (tests/testSimulation.cpp:97-150) diffs a
seeded run against stored files, which needs GSL's exact random streams;
this replaces it.  Each chain is sampled every 100 MC steps over 10^6
steps after a 20 000-step warm-up; at four temperature ranks chosen by a fixed rule (the
rank nearest 300 K and every fifth rank above it) the interaction energy
and the centre-to-centre distance of the replica at that temperature are
compared by two-sample Kolmogorov-Smirnov tests whose sample sizes are the
effective ones (the delayed-acceptance chain, whose stage-one rejections
can lengthen its autocorrelation at the hottest rank to about 100 samples,
is sampled over twice as many steps), n / tau_int (Sokal's automatic window, c = 5).  Each
comparison needs at least 100 effective samples on both sides and a
p-value above 1e-3 (eight comparisons per pair of chains: a family-wise
false-alarm rate below 1 %).  A comparison across temperatures shows that
the tests can tell different distributions apart.

The CPU chain is the oracle's REMC (oracle/oracle.c, orc_remc_mc_steps_fast:
the same moves and Metropolis law with FP64 pair terms) with another seed.
"""
import numpy as np
import pytest

import oracle
from cgppd_amd.surrogate_study import ks_2samp, tau_int

pytestmark = pytest.mark.gpu

CHUNK, WARMUP, CHUNKS, EXCHANGE = 100, 20000, 10000, 1000
MIN_EFFECTIVE, MIN_P = 100.0, 1e-3


def ranks_of(temps):
    k300 = int(np.argmin(np.abs(np.sort(temps) - 300.0)))
    return [k300 + 5 * i for i in range(4) if k300 + 5 * i < len(temps)]


def centre_distance(c, box):
    v = c[1] - c[0]
    v = v - box * np.rint(v / box)
    return float(np.linalg.norm(v))


class DeviceChain:
    def __init__(self, e, spec):
        self.e, self.spec, self.step = e, spec, 0

    def advance(self, n):
        done = 0
        while done < n:
            k = min(n - done, EXCHANGE - self.step % EXCHANGE)
            self.e.mc_steps(k)
            done += k
            self.step += k
            if self.step % EXCHANGE == 0:
                self.e.exchange()

    def record(self, ranks):
        s = self.e.stats()
        order = np.argsort(s["temperature"], kind="stable")
        en = [s["energy"][order[k]] for k in ranks]
        di = [centre_distance(self.e.coordinates(int(order[k]))[2], self.spec.box) for k in ranks]
        return en, di


class OracleChain:
    def __init__(self, st, spec, threads):
        self.st, self.spec, self.threads, self.step = st, spec, threads, 0
        st.refresh_fast()

    def advance(self, n):
        done = 0
        while done < n:
            k = min(n - done, EXCHANGE - self.step % EXCHANGE)
            self.st.mc_steps_fast(k, self.threads)
            done += k
            self.step += k
            if self.step % EXCHANGE == 0:
                self.st.exchange()

    def record(self, ranks):
        order = np.argsort(self.st.temperature, kind="stable")
        en = [self.st.energy[order[k]] for k in ranks]
        di = [centre_distance(self.st.center[order[k]], self.spec.box) for k in ranks]
        return en, di


def sample(chain, ranks, chunks=CHUNKS):
    chain.advance(WARMUP)
    import time
    t0 = time.perf_counter()
    E, D = [], []
    for c in range(chunks):
        if c % 2000 == 0:
            print(f"{type(chain).__name__}: chunk {c} of {chunks}, {time.perf_counter() - t0:.1f} s", flush=True)
        chain.advance(CHUNK)
        en, di = chain.record(ranks)
        E.append(en)
        D.append(di)
    return np.array(E), np.array(D)


def compare(a, b, label):
    """tau-corrected KS test of two recorded series; returns the result."""
    ta, tb = tau_int(a), tau_int(b)
    res = ks_2samp(a, b, ta, tb)
    print(f"{label}: tau {ta:.1f} / {tb:.1f}, n_eff {res['n_eff'][0]:.0f} / {res['n_eff'][1]:.0f}, "
          f"D {res['D']:.3f}, p {res['p_value']:.3g}", flush=True)
    return res


def assert_same(A, B, ranks, temps, what):
    results = []
    for i, k in enumerate(ranks):
        for name, x, y in (("energy", A[0][:, i], B[0][:, i]), ("distance", A[1][:, i], B[1][:, i])):
            results.append((k, name, compare(x, y, f"{what}, {temps[k]:.1f} K, {name}")))
    for k, name, res in results:
        assert min(res["n_eff"]) >= MIN_EFFECTIVE, (k, name, res)
        assert res["p_value"] > MIN_P, (k, name, res)


@pytest.fixture(scope="module")
def system():
    from cgppd_amd import synth
    spec = synth.config(0)
    return spec, np.sort(spec.temperatures()), ranks_of(spec.temperatures())


@pytest.fixture(scope="module")
def device_plain(system, gpu_engine_factory):
    spec, temps, ranks = system
    e = gpu_engine_factory()
    e.load_spec(spec, seed=11)
    e.mc_prepare()
    out = sample(DeviceChain(e, spec), ranks)
    e.close()
    return out


def test_device_remc_matches_cpu_remc(system, device_plain):
    import os
    spec, temps, ranks = system
    st = oracle.Remc(spec, seed=12)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    cpu = sample(OracleChain(st, spec, threads), ranks)
    assert_same(device_plain, cpu, ranks, temps, "device vs CPU")
    # the same test tells two temperatures apart
    res = compare(device_plain[0][:, 0], cpu[0][:, -1], f"device {temps[ranks[0]]:.1f} K vs CPU {temps[ranks[-1]]:.1f} K")
    assert res["p_value"] < 1e-6


def test_delayed_acceptance_matches_plain(system, device_plain, gpu_engine_factory):
    spec, temps, ranks = system
    e = gpu_engine_factory()
    e.load_spec(spec, seed=13)
    e.mc_prepare()
    e.surrogate_init(seed=4)
    for _ in range(20):
        e.run(1, 1000, use_graph=True)
        e.surrogate_collect()
    e.surrogate_train(400, batch=1024, lr=2e-3, seed=9)
    e.set_delayed_acceptance(True)
    da = sample(DeviceChain(e, spec), ranks, chunks=2 * CHUNKS)
    rej = e.stage1_rejections().sum()
    print(f"stage-one rejections {rej}", flush=True)
    assert rej > 0
    e.close()
    assert_same(device_plain, da, ranks, temps, "delayed acceptance vs plain")
