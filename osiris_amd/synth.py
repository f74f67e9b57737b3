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


# --------------------------------------------------------------- grids
# Saturation tables of the reference's example deck (compsim/test.dat:26-55),
# columns SAT, KR, KRO, PC (Pa).
SGT_TEST_DAT = np.array([
    [0, 0, 1, 0], [0.031, 0.0131, 0.6702, 40667.85714], [0.063, 0.0173, 0.55, 51765.35714],
    [0.094, 0.0226, 0.4395, 62035.71429], [0.125, 0.0297, 0.3395, 72030.35714],
    [0.188, 0.0515, 0.1721, 94087.5], [0.298, 0.1399, 0.028, 256483.2143],
    [0.386, 0.3939, 0.002, 419085.7143], [0.438, 0.5456, 0, 1671517.857],
    [0.565, 0.6215, 0, 2566900], [0.684, 0.6784, 0, 3995100], [0.8, 1, 0, 26824242.86]])
SWT_TEST_DAT = np.array([
    [0.2, 0, 1, 1349897.143], [0.229, 0.0001, 0.7407, 656200], [0.255, 0.0003, 0.6829, 413571.4286],
    [0.308, 0.0012, 0.5722, 151642.8571], [0.334, 0.0023, 0.5194, 129585.7143],
    [0.412, 0.0102, 0.3715, 86850], [0.464, 0.0219, 0.1526, 58106.78571],
    [0.557, 0.0416, 0.0822, 30328.57143], [0.606, 0.0721, 0, 9098.571429],
    [0.647, 0.1448, 0, 0], [0.7, 0.178, 0, 0], [0.8, 0.2604, 0, 0], [1, 1, 0, 0]])


@dataclass
class Wells:
    """Wells as the reference keeps them (deck WELLS, Data_Input.cpp:1185-1330):
    welli[w] = (type 0..8, Ix, Iy, first layer, last layer), 0-based blocks;
    wellf[w] = the type's nc + 6 values (Pwf or rate, limit, skin, rw, for
    types 3 and 8 the injection mole fractions, then start and end time in
    s); condition = WellCondition; sim_time = simTime, s; re_f = Re_f, m."""
    welli: np.ndarray          # [nw, 5] int32
    wellf: np.ndarray          # [nw, nc + 6]
    condition: np.ndarray      # [nw] int8
    sim_time: float = 0.0
    re_f: float = 0.0

    @property
    def nw(self):
        return int(self.welli.shape[0])


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    phi_m: np.ndarray
    phi_f: np.ndarray
    perm_m: np.ndarray      # [3, n] mD
    perm_f: np.ndarray
    tor: np.ndarray
    swt: np.ndarray
    sgt: np.ndarray
    refP: float = 101325.0
    cpor: float = 0.0
    dcpor: float = 0.0
    sigma: float = 1.0 / 3.0
    lf: float = 0.01
    const_diff: bool = False
    cdiff: float = 0.0
    T: float = 360.0
    dt: float = 86400.0
    wat_ro: float = 1000.0
    wat_mu: float = 1.0
    wells: "Wells" = None

    @property
    def n(self):
        return self.nx * self.ny * self.nz


@dataclass
class State:
    label_m: np.ndarray      # [n] int8
    label_f: np.ndarray
    Po_m: np.ndarray         # [n]
    Po_f: np.ndarray
    x_m: np.ndarray          # [nc, n]
    y_m: np.ndarray
    x_f: np.ndarray
    y_f: np.ndarray
    S_m: np.ndarray          # [3, n]
    S_f: np.ndarray
    rho_prev_m: np.ndarray = None   # [2, n] or None
    pcoeff: float = 1.0
    pre_m: np.ndarray = None        # [nc+1, n] or None
    pre_f: np.ndarray = None
    z_m: np.ndarray = None          # [nc, n] feed the state was flashed from (or None)
    z_f: np.ndarray = None


def dual_porosity_case(fluid, nx, ny, nz, seed, flash_fn, dx=10.0, dy=10.0, dz=5.0,
                       p_range=(5e6, 25e6), alpha=2.0, T=360.0, sw_m=0.2, sw_f=0.0,
                       phi_m=0.2, phi_f=0.01, k_m=0.1, k_f=100.0, k_sigma=0.5, **grid_kw):
    """A dual-porosity grid and its initial state.  Per cell and continuum
    p ~ U(p_range), z ~ Dirichlet(alpha); the state is the flash of (p, z)
    (labels, x, y and the So/Sg split of 1 - Sw), as the reference's
    initialisation produces it (AllFlash, Init_Calcs.cpp:949-962).
    flash_fn(T, P, z, sw) -> dict(label, x, y, so, sg)."""
    g = rng(seed)
    n = nx * ny * nz
    grid = Grid(nx, ny, nz, np.full(nx, dx), np.full(ny, dy), np.full(nz, dz),
                np.full(n, phi_m), np.full(n, phi_f),
                np.repeat(np.exp(np.log(k_m) + k_sigma * g.normal(size=n))[None], 3, 0).copy(),
                np.repeat(np.exp(np.log(k_f) + k_sigma * g.normal(size=n))[None], 3, 0).copy(),
                np.ones(n), SWT_TEST_DAT.copy(), SGT_TEST_DAT.copy(), T=T, **grid_kw)
    st, feeds = {}, {}
    for cont, sw in (("m", sw_m), ("f", sw_f)):
        P = g.uniform(p_range[0], p_range[1], n)
        z = g.dirichlet(np.full(fluid.nc, alpha), n).T.copy()
        r = flash_fn(T, P, z, sw)
        st[cont] = (P, r, sw)
        feeds[cont] = z
    def S(r, sw):
        return np.stack([np.full(n, sw), r["so"], r["sg"]])
    (Pm, rm, swm), (Pf, rf, swf) = st["m"], st["f"]
    state = State(rm["label"].astype(np.int8), rf["label"].astype(np.int8), Pm, Pf,
                  rm["x"], rm["y"], rf["x"], rf["y"], S(rm, swm), S(rf, swf),
                  pcoeff=float(max(Pm.max(), Pf.max())), z_m=feeds["m"], z_f=feeds["f"])
    return grid, state
