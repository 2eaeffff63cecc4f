"""Pins the oracle's Metropolis decision (src/Replica.cpp:385-410), which
judges every accept/reject decision the device takes.

The law is checked without restating the oracle's code: acceptance
frequencies over a fine grid of uniforms are compared with the Boltzmann
factor built from scipy's physical constants (the molar gas constant and
the thermochemical calorie), at several energies and temperatures; the
exact stationary distribution of a small state space under the decision
is compared with the Boltzmann weights of its states (detailed balance).
A unit slip (J vs kcal), a sign slip or an inverted comparison fails them.
"""
import math

import numpy as np
import pytest
from scipy import constants as sc

import oracle

GRID = 20000
U = (np.arange(GRID) + 0.5) / GRID          # uniforms on a fine grid of [0, 1)
KCAL_J = 1000.0 * sc.calorie                # 4184 J


def boltzmann(delta_kcal, temperature):
    return math.exp(-delta_kcal * KCAL_J / (sc.R * temperature))


def acceptance(delta, temperature, reference=False):
    """Fraction of the uniform grid the oracle accepts, and the decision codes."""
    codes = np.array([oracle.metropolis(delta, temperature, u, reference) for u in U])
    return float((codes < 2).mean()), codes


@pytest.mark.parametrize("reference", [False, True])
@pytest.mark.parametrize("delta,temperature", [(1.0, 300.0), (0.25, 288.0), (2.5, 450.0), (0.05, 500.0),
                                               (4.0, 300.0)])
def test_acceptance_frequency_is_boltzmann(delta, temperature, reference):
    p, codes = acceptance(delta, temperature, reference)
    assert abs(p - min(1.0, boltzmann(delta, temperature))) < 2.0 / GRID
    # uphill moves are never counted as downhill
    assert not (codes == 0).any()


def test_one_kcal_at_300k():
    """dE = 1 kcal/mol at 300 K is accepted with probability 0.1869."""
    p, _ = acceptance(1.0, 300.0)
    assert abs(p - 0.1869) < 1e-4


@pytest.mark.parametrize("reference", [False, True])
def test_downhill_and_flat_moves(reference):
    for u in (0.0, 0.5, 1.0 - 2.0 ** -53):
        assert oracle.metropolis(-1e-6, 300.0, u, reference) == 0
        assert oracle.metropolis(-50.0, 300.0, u, reference) == 0
        # no change: exp(0) = 1 > u, accepted by the Boltzmann test
        assert oracle.metropolis(0.0, 300.0, u, reference) == 1
    assert oracle.metropolis(1e4, 300.0, 0.0, reference) == 2


def test_acceptance_grows_with_temperature():
    ps = [acceptance(1.0, t)[0] for t in (250.0, 300.0, 400.0, 600.0)]
    assert all(a < b for a, b in zip(ps, ps[1:]))


def test_two_state_detailed_balance():
    """P(a->b) / P(b->a) = exp(-(E_b - E_a) / kT) for a symmetric proposal."""
    for de, t in ((0.7, 300.0), (1.8, 420.0), (0.2, 288.0)):
        up, _ = acceptance(de, t)
        down, _ = acceptance(-de, t)
        assert down == 1.0
        assert abs(up / down - boltzmann(de, t)) < 2.0 / GRID


def test_stationary_distribution_is_boltzmann():
    """Five states, uniform proposals to the other four: the exact stationary
    vector of the transition matrix the decision defines equals the
    normalised Boltzmann weights."""
    E = np.array([0.0, 0.4, -0.3, 1.2, 0.9])
    t = 310.0
    n = len(E)
    P = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                P[i, j] = acceptance(E[j] - E[i], t)[0] / (n - 1)
        P[i, i] = 1.0 - P[i].sum()
    w, v = np.linalg.eig(P.T)
    pi = np.real(v[:, np.argmin(abs(w - 1.0))])
    pi /= pi.sum()
    b = np.array([boltzmann(e, t) for e in E])
    b /= b.sum()
    np.testing.assert_allclose(pi, b, rtol=2e-3)


def test_simulated_two_state_chain_occupancy():
    """A two-state chain driven by the decision with seeded uniforms spends
    the Boltzmann fraction of its steps in the upper state."""
    rng = np.random.default_rng(11)
    de, t = 0.5, 300.0
    n, state, upper = 40000, 0, 0
    for u in rng.random(n):
        d = de if state == 0 else -de
        if oracle.metropolis(d, t, u) < 2:
            state = 1 - state
        upper += state
    q = boltzmann(de, t)
    want = q / (1.0 + q)
    # the chain's steps are correlated (lag-1 autocorrelation below 0.6 here):
    # a 6-sigma band of the inflated binomial error
    assert abs(upper / n - want) < 6.0 * math.sqrt(want * (1 - want) / n * 4.0)
