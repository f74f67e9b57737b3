"""Input deck of the reference (Data_Input, compsim/Data_Input.cpp:13-1349) --
the subset the per-cell Newton path consumes -- into the fluid, grid and
initial-state structures of this library.

Token rules are the reference's (MIfstream.cpp): a token is a run of
characters 33..126; a keyword is found by searching the whole deck from the
start for the first token equal to it (FileSearch), and its data are the
tokens after it (ReadWord); numbers are read with C atof/atoi semantics (the
longest numeric prefix, 0 if none).  Keywords read, with their units:

  GRID nx ny nz; DI/DJ/DK VAR v.. | CON v (m); NC pnc unc; MODEL PR|SRK;
  COMPNAME names (the reference's component library, Data_Input.cpp:657-951);
  BIC k_ij lower triangle by rows (bic = 1 - k_ij, :977-990);
  POR, FPOR (VAR zeros become 1e-11, :175, 285), PERMI/J/K,
  FPERMI/J/K (mD), TOR (default 1) -- each VAR | CON | IVAR | KVAR;
  SWT n, SGT n (rows SAT KR KRO PC(Pa)); RESTEMP (deg C, +273.15);
  WATERPROPS rho mu; REFP, CPOR, DCPOR (default 0); SIGMA (default 1);
  CDIF (constant diffusion coefficient, switches the correlation off);
  Lf, Re_f (m); DT, TOTALT (hours, x 3600); DTMAX (default 864000);
  ABSTOL, RELTOL, DIVTOL, MAXITERS, SOLVERTYPE (defaults 1e-50, 1e-18, 1e3,
  5000, bcgs, :232-266); DBS option pairs; INITCOND 0 (per-cell arrays with
  n*v repeats, ECLStar :1941-1975), 1 (per layer) or 3 (uniform) with
  IPRESS, FIPRESS (Pa), IWS, FIWS, IGC, FIGC; WELLS (wells_table turns them
  into the table the assembler completes).

Refused with DeckError where the reference's own reading is undefined:
JVAR distributions (its loops test the wrong index, :150, 190, 300, 347,
383, 419, 456, 492, 528), INITCOND 2 (its grading module does not exist,
:1133), user components
(NC with unc > 0) and the library's O2, SO2, H2, H2O (no parachor is ever
set for them: fluidProp is malloc'd, :1443-1446, and neither the library nor
CPlus_Props writes it).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

from . import synth

TIMECHFACT = 3600.0     # hours -> s (Functions.h:243)

# The reference's component library (Data_Input.cpp:657-951):
# name: (MW g/mol, Tc K, Pc Pa, Vc cm^3/mol, acentric factor, parachor, Zc);
# None where the library sets no value.
LIBRARY = {
    "C1": (16.043, 190.56, 4599000, 98.6, 0.008, 74.05, 0.2862),
    "C2": (30.070, 305.32, 4872000, 145.5, 0.0995, 112.91, 0.2793),
    "C3": (44.096, 369.83, 4248000, 200, 0.1523, 154.03, 0.2763),
    "iC4": (58.123, 408.14, 3648000, 262.7, 0.1770, 185.32, 0.2824),
    "nC4": (58.123, 425.12, 3796000, 255, 0.2002, 193.90, 0.2739),
    "iC5": (72.150, 460.43, 3381000, 305.8, 0.2275, 229.37, 0.2701),
    "nC5": (72.150, 469.7, 3370000, 313, 0.2515, 236.00, 0.2701),
    "nC6": (86.177, 507.6, 3025000, 371, 0.3013, 276.71, 0.2659),
    "nC7": (100.204, 540.2, 2740000, 428, 0.3495, 318.44, 0.2611),
    "nC8": (114.231, 568.7, 2490000, 486, 0.3996, 359.33, 0.2559),
    "nC9": (128.258, 594.6, 2290000, 544, 0.4435, 399.57, 0.2520),
    "nC10": (142.285, 617.7, 2110000, 600, 0.4923, 440.69, 0.2465),
    "nC11": (156.312, 639, 1949000, 659, 0.5303, 482.00, 0.2419),
    "nC12": (170.338, 658, 1820000, 716, 0.5764, 522.26, 0.2382),
    "nC13": (184.365, 675, 1680000, 775, 0.6174, 536.77, 0.2320),
    "nC14": (198.392, 693, 1570000, 830, 0.6430, 606.05, 0.2262),
    "nC15": (212.419, 708, 1480000, 889, 0.6863, 647.43, 0.2235),
    "nC16": (226.446, 723, 1400000, 944, 0.7174, 688.50, 0.2199),
    "nC17": (240.473, 736, 1340000, 1000, 0.7697, 730.05, 0.2190),
    "nC18": (254.5, 747, 1270000, 1060, 0.8114, 771.95, 0.2168),
    "nC19": (268.527, 758, 1210000, 1120, 0.8522, 813.85, 0.2150),
    "nC20": (282.553, 768, 1160000, 1170, 0.9069, 853.67, 0.2126),
    "nC21": (296.580, 781.7, 1147000, 1198, 0.9220, 897.64, 0.2114),
    "nC22": (310.610, 791.8, 1101000, 1253, 0.9550, 939.55, 0.2095),
    "nC23": (324.630, 801.3, 1059000, 1307, 0.9890, 981.43, 0.2078),
    "nC24": (338.680, 810.4, 1019000, 1362, 1.0190, 1023.40, 0.2061),
    "CO2": (44.010, 304.19, 7382000, 94, 0.2276, 82.00, 0.2744),
    "O2": (31.999, 154.58, 5043000, 73.4, 0.0218, None, 0.2880),
    "N2": (28.014, 126.1, 3394000, 90.1, 0.0403, 61.12, 0.2917),
    "H2S": (34.082, 373.53, 8963000, 98.5, 0.0827, 85.50, 0.2843),
    "SO2": (64.065, 430.75, 7884000, 122, 0.2451, None, 0.2686),
    "H2": (2.016, 33.18, 1313000, 64.2, 0.2150, None, 0.3053),
    "H2O": (18.015, 647.13, 22055000, 56, 0.3449, None, 0.2294),
}

# well data counts per well type after the five integers (Data_Input.cpp:1190-1330)
_WELL_REALS = {0: 6, 1: 6, 2: 6, 4: 6, 5: 6, 6: 6, 7: 6}


class DeckError(ValueError):
    """The deck lacks a keyword the reference requires, has one in a form
    the reference rejects, or asks for something the reference leaves
    undefined."""


_NUM = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.IGNORECASE)
_INT = re.compile(r"[+-]?\d+")


def atof(tok):
    """C atof: the longest numeric prefix, 0.0 if there is none."""
    m = _NUM.match(tok)
    return float(m.group(0)) if m else 0.0


def atoi(tok):
    """C atoi: the longest integer prefix, 0 if there is none."""
    m = _INT.match(tok)
    return int(m.group(0)) if m else 0


def tokens(text):
    """MIfstream::ReadWord's tokens: runs of characters 33..126."""
    data = text.encode("latin-1", "replace") if isinstance(text, str) else bytes(text)
    return [m.group(0).decode("latin-1") for m in re.finditer(rb"[\x21-\x7e]+", data)]


class _Reader:
    def __init__(self, toks):
        self.t = toks
        self.pos = len(toks)
        self.kw = ""

    def search(self, kw):
        """FileSearch: position after the first token equal to kw."""
        for i, tok in enumerate(self.t):
            if tok == kw:
                self.pos, self.kw = i + 1, kw
                return True
        return False

    def need(self, kw):
        if not self.search(kw):
            raise DeckError(f"no {kw} keyword in the deck")

    def word(self):
        if self.pos >= len(self.t):
            raise DeckError(f"incorrect {self.kw} keyword format in the deck")
        w = self.t[self.pos]
        self.pos += 1
        return w

    def num(self):
        return atof(self.word())

    def int(self):
        return atoi(self.word())


def _distribution(r, kw, nx, ny, nz, zero_to=None, required=True, default=None):
    """A per-cell array in VAR | CON | IVAR | KVAR form, cell index
    iz*nx*ny + iy*nx + ix (the deck's k-outer, i-inner order)."""
    if not r.search(kw):
        if required:
            raise DeckError(f"no {kw} keyword in the deck")
        return np.full(nx * ny * nz, default, dtype=np.float64)
    mode = r.word()
    n = nx * ny * nz
    if mode == "VAR":
        a = np.array([r.num() for _ in range(n)])
        if zero_to is not None:
            a[a == 0] = zero_to
        return a
    if mode == "CON":
        return np.full(n, r.num())
    if mode == "IVAR":
        v = np.array([r.num() for _ in range(nx)])
        return np.tile(v, ny * nz)
    if mode == "KVAR":
        v = np.array([r.num() for _ in range(nz)])
        return np.repeat(v, nx * ny)
    if mode == "JVAR":
        raise DeckError(f"{kw} JVAR: the reference's JVAR loops test the wrong index and read no defined "
                        "values (Data_Input.cpp:150, 190, 300, 347, 383, 419, 456, 492, 528)")
    raise DeckError(f"incorrect {kw} keyword format in the deck")


def _dims(r, kw, count):
    r.need(kw)
    mode = r.word()
    if mode == "VAR":
        return np.array([r.num() for _ in range(count)])
    if mode == "CON":
        return np.full(count, r.num())
    raise DeckError(f"incorrect {kw} keyword format in the deck")


def ecl_star(tok):
    """ECLStar (Data_Input.cpp:1941-1975): "n*v" -> (n, v), "v" -> (1, v);
    the count is the text before the first '*', the value the non-'*'
    characters after it."""
    if "*" not in tok:
        return 1, atof(tok)
    a, b = tok.split("*", 1)
    return atoi(a), atof(b.replace("*", ""))


def _ecl_values(r, count):
    """count values in the n*v repeat form, as the INITCOND 0 loops read
    them (Data_Input.cpp:996-1054): a repeat count below 1 leaves the
    reference's counter negative, so it never reads again and that value
    fills the rest of the keyword; repeats beyond the last cell are dropped."""
    out = np.empty(count)
    i = 0
    while i < count:
        times, v = ecl_star(r.word())
        if times < 1:
            out[i:] = v
            break
        m = min(times, count - i)
        out[i:i + m] = v
        i += m
    return out


@dataclass
class Deck:
    fluid: synth.Fluid
    names: list
    grid: synth.Grid
    init: dict
    run: dict
    wells: list = field(default_factory=list)


def parse(text):
    """Parse a deck (its text or bytes) into a Deck."""
    r = _Reader(tokens(text))
    r.need("GRID")
    nx, ny, nz = r.int(), r.int(), r.int()
    if min(nx, ny, nz) < 1:
        raise DeckError("incorrect GRID keyword format in the deck")
    n = nx * ny * nz
    r.need("NC")
    pnc, unc = r.int(), r.int()
    if unc:
        raise DeckError("user components (NC with a second count > 0) have no parachor in the reference "
                        "(CPlus_Props sets none; fluidProp is malloc'd, Data_Input.cpp:1443-1446)")
    nc = pnc
    if nc < 1:
        raise DeckError("incorrect NC keyword format in the deck")
    r.need("SWT")
    nswt = r.int()
    r.need("SGT")
    nsgt = r.int()
    r.need("WELLS")
    nwell = r.int()
    ndbs = 0
    if r.search("DBS"):
        ndbs = r.int()
    r.need("RESTEMP")
    T = r.num() + 273.15
    r.need("WATERPROPS")
    wat_ro, wat_mu = r.num(), r.num()
    dx, dy, dz = _dims(r, "DI", nx), _dims(r, "DJ", ny), _dims(r, "DK", nz)
    tor = _distribution(r, "TOR", nx, ny, nz, required=False, default=1.0)
    phi_m = _distribution(r, "POR", nx, ny, nz, zero_to=1e-11)
    const_diff, cdiff = False, 0.0
    if r.search("CDIF"):
        const_diff, cdiff = True, r.num()
    sigma = r.num() if r.search("SIGMA") else 1.0
    dt_max = r.num() if r.search("DTMAX") else 864000.0
    atol = r.num() if r.search("ABSTOL") else 1e-50
    rtol = r.num() if r.search("RELTOL") else 1e-18
    dtol = r.num() if r.search("DIVTOL") else 1e3
    maxit = r.int() if r.search("MAXITERS") else 5000
    solver = r.word() if r.search("SOLVERTYPE") else "bcgs"
    dbs = []
    if r.search("DBS"):
        r.word()
        dbs = [(r.word(), r.word()) for _ in range(ndbs)]
    phi_f = _distribution(r, "FPOR", nx, ny, nz, zero_to=1e-11)
    r.need("Lf")
    lf = r.num()
    r.need("Re_f")
    re_f = r.num()
    perm_m = np.stack([_distribution(r, k, nx, ny, nz) for k in ("PERMI", "PERMJ", "PERMK")])
    perm_f = np.stack([_distribution(r, k, nx, ny, nz) for k in ("FPERMI", "FPERMJ", "FPERMK")])
    refP = r.num() if r.search("REFP") else 0.0
    cpor = r.num() if r.search("CPOR") else 0.0
    dcpor = r.num() if r.search("DCPOR") else 0.0
    r.need("MODEL")
    model = r.word()
    if model not in ("PR", "SRK"):
        raise DeckError("incorrect EOS MODEL keyword format in the deck")
    r.need("COMPNAME")
    names = [r.word() for _ in range(nc)]
    props = []
    for nm in names:
        if nm not in LIBRARY:
            raise DeckError(f"unknown component {nm}")
        if LIBRARY[nm][5] is None:
            raise DeckError(f"component {nm}: the reference's library sets no parachor for it (its value "
                            "is uninitialised memory there, Data_Input.cpp:1443-1446)")
        props.append(LIBRARY[nm])
    props = np.array(props, dtype=np.float64)
    # saturation tables (zero entries stay zero: RELPERM0 = 0, Functions.h:266)
    r.search("SWT")
    r.word()
    swt = np.array([[r.num() for _ in range(4)] for _ in range(nswt)])
    r.search("SGT")
    r.word()
    sgt = np.array([[r.num() for _ in range(4)] for _ in range(nsgt)])
    kij = np.zeros((nc, nc))
    if r.search("BIC"):
        for i in range(1, nc):
            for j in range(i):
                kij[i, j] = kij[j, i] = r.num()
    init = _initial(r, nx, ny, nz, nc)
    wells = _wells(r, nwell, nc)
    r.need("DT")
    dt = r.num() * TIMECHFACT
    r.need("TOTALT")
    total = r.num() * TIMECHFACT

    fluid = synth.Fluid(nc, synth.PR if model == "PR" else synth.SRK, props[:, 1].copy(), props[:, 2].copy(),
                        props[:, 4].copy(), kij, MW=props[:, 0].copy(), Vc=props[:, 3].copy(),
                        Zc=props[:, 6].copy(), parachor=props[:, 5].copy())
    grid = synth.Grid(nx, ny, nz, dx, dy, dz, phi_m, phi_f, perm_m, perm_f, tor, swt, sgt, refP=refP,
                      cpor=cpor, dcpor=dcpor, sigma=sigma, lf=lf, const_diff=const_diff, cdiff=cdiff, T=T,
                      dt=dt, wat_ro=wat_ro, wat_mu=wat_mu)
    run = {"dt": dt, "total_time": total, "dt_max": dt_max, "re_f": re_f, "solver_type": solver,
           "ksp": {"rtol": rtol, "atol": atol, "dtol": dtol, "maxit": maxit}, "dbs": dbs}
    return Deck(fluid, names, grid, init, run, wells)


def load(path):
    with open(path, "rb") as f:
        return parse(f.read())


def _initial(r, nx, ny, nz, nc):
    """INITCOND (Data_Input.cpp:991-1179): oil pressure, water saturation and
    feed of both continua, before the initial flash."""
    n = nx * ny * nz
    r.need("INITCOND")
    kind = r.int()
    out = {"kind": kind}
    if kind == 0:
        for key, kw in (("Po_m", "IPRESS"), ("Po_f", "FIPRESS"), ("Sw_m", "IWS"), ("Sw_f", "FIWS")):
            r.need(kw)
            out[key] = _ecl_values(r, n)
        for key, kw in (("z_m", "IGC"), ("z_f", "FIGC")):
            r.need(kw)
            out[key] = _ecl_values(r, nc * n).reshape(nc, n)
    elif kind == 1:
        for key, kw in (("Po_m", "IPRESS"), ("Po_f", "FIPRESS"), ("Sw_m", "IWS"), ("Sw_f", "FIWS")):
            r.need(kw)
            out[key] = np.repeat(np.array([r.num() for _ in range(nz)]), nx * ny)
        for key, kw in (("z_m", "IGC"), ("z_f", "FIGC")):
            r.need(kw)
            layers = np.array([[r.num() for _ in range(nc)] for _ in range(nz)])   # [nz][nc]
            out[key] = np.repeat(layers.T, nx * ny, axis=1)
    elif kind == 3:
        for key, kw in (("Po_m", "IPRESS"), ("Po_f", "FIPRESS"), ("Sw_m", "IWS"), ("Sw_f", "FIWS")):
            r.need(kw)
            out[key] = np.full(n, r.num())
        for key, kw in (("z_m", "IGC"), ("z_f", "FIGC")):
            r.need(kw)
            out[key] = np.repeat(np.array([r.num() for _ in range(nc)])[:, None], n, axis=1)
    elif kind == 2:
        raise DeckError("INITCOND 2 (compositional grading): the reference's grading module does not exist "
                        "(Data_Input.cpp:1133)")
    else:
        raise DeckError("incorrect INITCOND keyword format in the deck")
    return out


def _wells(r, nwell, nc):
    """WELLS (Data_Input.cpp:1180-1334): type, 0-based block indices (i, j,
    k first, k last), and the type's values with the two times in seconds."""
    if not r.search("WELLS"):
        return []
    r.word()
    wells = []
    for _ in range(nwell):
        ints = [r.int() - 1 for _ in range(5)]
        wtype = ints[0] + 1
        nreal = _WELL_REALS.get(wtype, nc + 6 if wtype in (3, 8) else None)
        if nreal is None:
            raise DeckError("incorrect WELL keyword format in the deck")
        vals = [r.num() for _ in range(nreal)]
        for j in (nreal - 2, nreal - 1):
            vals[j] *= TIMECHFACT
        wells.append({"type": wtype, "i": ints[1], "j": ints[2], "k_first": ints[3], "k_last": ints[4],
                      "values": vals})
    return wells


def wells_table(deck, sim_time, condition=None):
    """The deck's wells as a synth.Wells (welli rows type, Ix, Iy, first and
    last layer; wellf rows of nc + 6 values, the type's values first) at
    simTime = sim_time s, with WellCondition `condition` (default all false,
    as Init_Calcs.cpp:19 starts them) and the deck's Re_f."""
    nc = deck.fluid.nc
    nw = len(deck.wells)
    wi = np.zeros((nw, 5), np.int32)
    wf = np.zeros((nw, nc + 6))
    for i, w in enumerate(deck.wells):
        wi[i] = (w["type"], w["i"], w["j"], w["k_first"], w["k_last"])
        wf[i, :len(w["values"])] = w["values"]
    cond = np.zeros(nw, np.int8) if condition is None else np.asarray(condition, np.int8)
    return synth.Wells(wi, wf, cond, sim_time=float(sim_time), re_f=float(deck.run["re_f"]))


def initial_state(deck, flash_fn):
    """The initial state AllFlash leaves (Init_Calcs.cpp:571-738, 949-962):
    every cell of both continua flashed at its oil pressure and feed; labels,
    x, y; So, Sg the Z-weighted split of 1 - Sw; PCoeff the largest oil
    pressure (CalcPCoeff).  flash_fn(T, P, z) -> dict of numpy arrays label,
    x, y, L, Zl, Zg (e.g. the device flash)."""
    ini = deck.init
    parts = {}
    for k in ("m", "f"):
        P, z, sw = ini["Po_" + k], ini["z_" + k], ini["Sw_" + k]
        res = flash_fn(deck.grid.T, P, z)
        L, Zl, Zg = res["L"], res["Zl"], res["Zg"]
        nd = Zl * L / (Zl * L + Zg * (1 - L))
        S = np.stack([sw, nd * (1 - sw), (1 - nd) * (1 - sw)])
        parts[k] = (res["label"].astype(np.int8), P.astype(np.float64), res["x"], res["y"], S, z)
    (lm, Pm, xm, ym, Sm, zm), (lf, Pf, xf, yf, Sf, zf) = parts["m"], parts["f"]
    pcoeff = float(max(0.0, Pm.max(), Pf.max()))
    return synth.State(lm, lf, Pm, Pf, xm, ym, xf, yf, Sm, Sf, pcoeff=pcoeff, z_m=zm, z_f=zf)
