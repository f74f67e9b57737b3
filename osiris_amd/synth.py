"""Synthetic fluids and cell inputs for the five benchmark/parity workloads.

The reference holds no fixtures for this path (its only input is one example
deck, compsim/test.dat), so every workload is generated here from a fixed seed
with numpy's PCG64.  Component constants of the three-component workload are
the ones the workload names; molecular weight, critical volume, critical Z and
parachor come from the reference's component library (compsim/Data_Input.cpp:
657-763, entries C1, nC4, nC10).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PR, SRK = 0, 1


@dataclass
class Fluid:
    nc: int
    eos: int                 # PR or SRK
    Tc: np.ndarray           # K
    Pc: np.ndarray           # Pa
    omega: np.ndarray
    kij: np.ndarray          # [nc, nc] binary interaction k_ij (symmetric, zero diagonal)
    MW: np.ndarray = field(default=None)       # g/mol
    Vc: np.ndarray = field(default=None)       # cm^3/mol (library units)
    Zc: np.ndarray = field(default=None)
    parachor: np.ndarray = field(default=None)

    def bic(self):
        """1 - k_ij, the form the reference stores (Data_Input.cpp:977-990)."""
        return 1.0 - self.kij


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def fluid_c1_nc4_nc10():
    """Three-component PR fluid of the 10x10x1 dual-porosity workload."""
    Tc = np.array([190.6, 425.1, 617.7])
    Pc = np.array([4.599e6, 3.796e6, 2.110e6])
    om = np.array([0.011, 0.200, 0.490])
    kij = np.zeros((3, 3))
    kij[0, 2] = kij[2, 0] = 0.04
    return Fluid(3, PR, Tc, Pc, om, kij,
                 MW=np.array([16.043, 58.123, 142.285]),
                 Vc=np.array([98.6, 255.0, 600.0]),
                 Zc=np.array([0.2862, 0.2739, 0.2465]),
                 parachor=np.array([74.05, 193.90, 440.69]))


def fluid_pseudo(nc, eos, seed):
    """nc pseudo-components: Tc~U(190,800) K, Pc~U(1.5,4.6) MPa,
    omega~U(0.01,0.9), k_0j~U(0,0.15) for pairs with the lightest component,
    else 0.  Components are ordered by Tc so component 0 is the methane-like
    one.  MW, Vc, Zc and parachor follow smooth correlations in Tc (they only
    enter viscosity and interfacial tension, not the flash)."""
    g = rng(seed)
    Tc = np.sort(g.uniform(190.0, 800.0, nc))
    Pc = g.uniform(1.5e6, 4.6e6, nc)
    om = g.uniform(0.01, 0.9, nc)
    kij = np.zeros((nc, nc))
    k0 = g.uniform(0.0, 0.15, nc)
    for j in range(1, nc):
        kij[0, j] = kij[j, 0] = k0[j]
    MW = 16.0 + 0.45 * (Tc - 190.0) + 1.2e-4 * (Tc - 190.0) ** 2
    Zc = 0.2901 - 0.0879 * om
    Vc = Zc * 8.3144621 * Tc / Pc * 1e6
    parachor = 74.0 + 2.4 * (MW - 16.0)
    return Fluid(nc, eos, Tc, Pc, om, kij, MW=MW, Vc=Vc, Zc=Zc, parachor=parachor)


@dataclass
class FlashCells:
    T: np.ndarray        # [n] K
    P: np.ndarray        # [n] Pa
    z: np.ndarray        # [nc, n] feed mole fractions (structure of arrays)
    psat: np.ndarray     # [n] saturation pressure used to place P (-1 if none)


def flash_cells(fluid, n, seed, psat_fn, alpha=1.5, T_range=(330.0, 420.0),
                band=(0.85, 1.15), p_fallback=(5e6, 25e6)):
    """Independent flash cells with pressure placed around each cell's
    saturation pressure: T~U(T_range), z~Dirichlet(alpha), P = Psat*U(band).
    psat_fn(T, P0, kind, x, y) -> (psat, xm) is the saturation routine
    (bubble point of z first, dew point where that fails); cells with no
    saturation pressure get P~U(p_fallback)."""
    g = rng(seed)
    nc = fluid.nc
    T = g.uniform(T_range[0], T_range[1], n)
    z = g.dirichlet(np.full(nc, alpha), n).T.copy()
    u = g.uniform(band[0], band[1], n)
    pf = g.uniform(p_fallback[0], p_fallback[1], n)
    P0 = np.full(n, 20e6)
    ps, _ = psat_fn(T, P0, np.ones(n, np.int8), z, z)
    bad = ~(ps > 0)
    if bad.any():
        pd, _ = psat_fn(T, P0, -np.ones(n, np.int8), z, z)
        ps = np.where(bad, pd, ps)
    ok = ps > 0
    P = np.where(ok, ps * u, pf)
    return FlashCells(T=T, P=P, z=z, psat=np.where(ok, ps, -1.0))
