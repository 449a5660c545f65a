"""Synthetic constant-pressure kinetics networks and their initial conditions.

The reference ships no chemistry mechanism: ``mechanism/`` only holds SCons
files for pyJac output (mechanism/SConscript:1-22), and the pyJac example
(examples/pyJac/h2.cti, 13 species / 27 reactions) needs Cantera and pyJac.
This module generates seeded stand-in networks of the sizes the workloads
name (13, 53 and 111 species) whose right-hand side and analytic Jacobian are
evaluated by the library's built-in kinetics callbacks.

Model (state ``y = (T, Y_1..Y_ns)``, parameter = pressure in Pa):

* species are formulas over four elements (H, C, O, N; at most 4, 3, 3, 2
  atoms, 8 in total), so
  molecular weights follow from the formula and every reaction conserves
  elements, hence mass (sum Y_k = 1 is invariant);
* thermo per species: ``cp/R = a0 + a1 T + a2 T^2``, with enthalpy and entropy
  offsets ``bh`` [K] and ``bs``;
* reactions are reversible and mass-action with 1-3 slots per side of equal
  element content; reverse rates follow from the equilibrium constant; each
  reaction is written in its favourable direction at 1300 K and kept only if
  that direction has 2 or 3 reactant slots (no unimolecular forward steps,
  which bounds the fastest first-order rates near 1e12 /s);
* ``kf = A (T / 1000 K)^beta exp(-Ea / (R T))`` with A log-uniform in
  [1e8, 1e16] (mol, cm, s units), beta uniform in [-1, 2], Ea uniform in
  [0, 3e5] J/mol.
"""
from __future__ import annotations

import dataclasses
import itertools
import math

import numpy as np

R_GAS = 8.31446261815324
P_ATM = 101325.0
ELEMENT_W = np.array([1.008, 12.011, 15.999, 14.007]) * 1e-3  # kg/mol


@dataclasses.dataclass
class KineticsMechanism:
    ns: int
    nr: int
    W: np.ndarray        # (ns,) kg/mol
    thermo: np.ndarray   # (ns, 5): a0, a1, a2, bh, bs
    slots: np.ndarray    # (nr, 6) int32: reactants 0..2, products 3..5, -1 empty
    lnA: np.ndarray      # (nr,)
    beta: np.ndarray     # (nr,)
    EaR: np.ndarray      # (nr,) K
    formulas: np.ndarray  # (ns, 4) element counts
    seed: int = 0

    @property
    def n(self) -> int:
        return self.ns + 1

    def save(self, path: str) -> None:
        np.savez(path, **{f.name: getattr(self, f.name) for f in dataclasses.fields(self)})

    def write_binary(self, path: str) -> None:
        """The network file ``accelerInt_load_kinetics`` reads (and the command-line
        driver's ``--mech``): int32 ns, int32 nr, W, thermo, slots (int32), lnA,
        beta, EaR."""
        with open(path, "wb") as f:
            np.array([self.ns, self.nr], dtype=np.int32).tofile(f)
            for a, t in ((self.W, np.float64), (self.thermo, np.float64), (self.slots, np.int32),
                         (self.lnA, np.float64), (self.beta, np.float64), (self.EaR, np.float64)):
                np.ascontiguousarray(a, dtype=t).tofile(f)

    @classmethod
    def load(cls, path: str) -> "KineticsMechanism":
        with np.load(path, allow_pickle=False) as z:
            kw = {f.name: z[f.name] for f in dataclasses.fields(cls)}
        kw["ns"], kw["nr"], kw["seed"] = int(kw["ns"]), int(kw["nr"]), int(kw["seed"])
        return cls(**kw)

    # thermodynamics (same formulas as the device code)
    def g_RT(self, T: float) -> np.ndarray:
        return _g_RT(self.thermo, T)

    def ln_kc(self, T: float) -> np.ndarray:
        g = self.g_RT(T)
        lnC0 = math.log(1e-6 * P_ATM / (R_GAS * T))
        out = np.zeros(self.nr)
        for j in range(self.nr):
            sl = self.slots[j]
            nu_g = sum(-g[s] for s in sl[:3] if s >= 0) + sum(g[s] for s in sl[3:] if s >= 0)
            dnu = int((sl[3:] >= 0).sum() - (sl[:3] >= 0).sum())
            out[j] = -nu_g + dnu * lnC0
        return out


def _g_RT(thermo, T):
    a0, a1, a2, bh, bs = thermo.T
    h = a0 * T + a1 * T * T / 2 + a2 * T ** 3 / 3 + bh
    s = a0 * math.log(T) + a1 * T + a2 * T * T / 2 + bs
    return h / T - s


def _formulas(rng, ns):
    space = [f for f in itertools.product(range(5), range(4), range(4), range(3)) if 0 < sum(f) <= 8]
    # favour small molecules: weight by 1/atoms^1.5
    w = np.array([1.0 / sum(f) ** 1.5 for f in space])
    idx = rng.choice(len(space), size=ns, replace=False, p=w / w.sum())
    return np.array([space[i] for i in idx], dtype=np.int64)


def generate(ns: int, nr: int, seed: int, max_tries: int = 200) -> KineticsMechanism:
    """Seeded synthetic network with ``ns`` species and ``nr`` reactions."""
    for attempt in range(max_tries):
        rng = np.random.default_rng([seed, attempt])
        F = _formulas(rng, ns)
        mech = _try_generate(rng, F, ns, nr, seed)
        if mech is not None:
            return mech
    raise RuntimeError("could not generate a %d-species / %d-reaction network" % (ns, nr))


def _try_generate(rng, F, ns, nr, seed):
    atoms = F.sum(1)
    W = F @ ELEMENT_W
    a0 = 2.5 + 0.9 * (atoms - 1) + rng.uniform(0.0, 1.0, ns)
    a1 = rng.uniform(0.0, 1.5e-3, ns) * atoms
    a2 = -rng.uniform(0.0, 2.0e-7, ns) * atoms
    bh = rng.uniform(-6.0e3, 6.0e3, ns)
    bs = rng.uniform(5.0, 25.0, ns)
    thermo = np.stack([a0, a1, a2, bh, bs], 1)

    # index multisets of 1..3 species by element vector
    by_comp = {}
    for k in (1, 2, 3):
        for ms in itertools.combinations_with_replacement(range(ns), k):
            comp = tuple(F[list(ms)].sum(0))
            by_comp.setdefault(comp, []).append(ms)

    g1300 = _g_RT(thermo, 1300.0)
    lnC0 = math.log(1e-6 * P_ATM / (R_GAS * 1300.0))

    def favour(reac, prod):
        # ln Kc at 1300 K plus the concentration factor at a typical species
        # concentration of 1e-5 mol/cm^3: > 0 when forward is favourable
        dnu = len(prod) - len(reac)
        lnkc = -(sum(g1300[s] for s in prod) - sum(g1300[s] for s in reac)) + dnu * lnC0
        return lnkc + dnu * math.log(1e-5)

    slots, seen = [], set()
    tries = 0
    while len(slots) < nr and tries < 200 * nr:
        tries += 1
        k = 2 if rng.uniform() < 0.8 else 3
        reac = tuple(sorted(rng.choice(ns, size=k, replace=True)))
        comp = tuple(F[list(reac)].sum(0))
        cands = [p for p in by_comp.get(comp, []) if not set(p) & set(reac)]
        if not cands:
            continue
        prod = cands[rng.integers(len(cands))]
        if favour(reac, prod) < 0:
            reac, prod = prod, reac
        if len(reac) < 2:  # keep every forward step bi- or termolecular
            continue
        key = (reac, prod)
        if key in seen or (prod, reac) in seen:
            continue
        seen.add(key)
        row = [-1] * 6
        row[:len(reac)] = reac
        row[3:3 + len(prod)] = prod
        slots.append(row)
    if len(slots) < nr:
        return None
    slots = np.array(slots, dtype=np.int32)
    lnA = np.log(10.0) * rng.uniform(8.0, 16.0, nr)
    beta = rng.uniform(-1.0, 2.0, nr)
    EaR = rng.uniform(0.0, 3.0e5, nr) / R_GAS
    mech = KineticsMechanism(ns=ns, nr=nr, W=W, thermo=thermo, slots=slots, lnA=lnA, beta=beta,
                             EaR=EaR, formulas=F, seed=seed)
    return mech


# ---- workloads -----------------------------------------------------------------

#: name -> (species, reactions, generator seed)
NETWORKS = {14: (13, 27, 20261243), 54: (53, 325, 20261204), 112: (111, 784, 20261206)}


def network(n_states: int) -> KineticsMechanism:
    ns, nr, seed = NETWORKS[n_states]
    return generate(ns, nr, seed)


def initial_conditions(mech: KineticsMechanism, num: int, seed: int, T_range=(1000.0, 1600.0),
                       P_atm_range=(1.0, 25.0), alpha: float = 0.5):
    """Column-major y (N, num) = (T, Y...) and pressure [Pa] per condition:
    T uniform, P uniform in atm, mass fractions Dirichlet(alpha)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(*T_range, num)
    P = rng.uniform(*P_atm_range, num) * P_ATM
    Y = rng.dirichlet(np.full(mech.ns, alpha), size=num)
    y = np.empty((mech.ns + 1, num))
    y[0] = T
    y[1:] = Y.T
    return np.ascontiguousarray(y), P


def van_der_pol_conditions(num: int, seed: int):
    """mu = 10^U(1,3), y0 = (U[-2,2], U[-1,1])."""
    rng = np.random.default_rng(seed)
    mu = 10.0 ** rng.uniform(1.0, 3.0, num)
    y = np.stack([rng.uniform(-2.0, 2.0, num), rng.uniform(-1.0, 1.0, num)])
    return np.ascontiguousarray(y), mu
