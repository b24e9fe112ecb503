"""Reader for the reference's input decks and angle files, as far as the
solver step needs them: grid, crystals, crystal-plasticity materials,
element-to-material map, box sizes, boundary conditions, loading and the
nonlinear analysis parameters.  It produces what cpfft_create and the
load-step driver take (per-voxel Kocks angles, cpfft_material, dt, BC_all,
isNBC, tolerances), so a deck that drives CPFFT drives this solver.

Commands follow the reference's input driver (FFT_finite_3d.f:14-160):
  project [name] <name>                    structure name, 8 characters (:64-78)
  number of grid <N>                       (FFT_finite_3d.f:82-93)
  crystal <k> / properties ...             incrystal.f:13-1011
  material <name> / properties cp ...      inmat.f:154-313 (mm10 only)
  sizes of x_direction a y_direction b z_direction c   (box edges, output)
  elements / <list> material <name> ...    inelem.f
  blocking automatic                       (ignored: one voxel per element)
  strains / F_ij v | P_ij v ...            inlodcase.f:11-148
  loading <name> / step <list> constraints c ...       inlod.f:11-105
  nonlinear analysis parameters / maximum iterations n,
      convergence tolerance NR a CG b, time step dt    indypm.f:14-68
  output results steps <list>              load steps with result files (oudriv.f:85-160)
  output model / data ... (ignored), compute, stop

Scanner rules kept from the reference (scan.f, incrystal.f:987-994): a line
whose first token is `c` is a comment; a trailing comma continues a
statement on the next line; inside a property list a comma anywhere makes
the reader take the next line, so the rest of the line after a comma is not
read (the example deck's `harden_n 5,48` sets harden_n = 5).

Material defaults are the reference's (mod_crystals.f:245-410): e 69000,
nu 0.33, harden_n 20, theta_0 100, tau_y 0, tau_v 0, voce_m 1, gamma_bar
1e10, alter_mode off, local tolerances atol = atol1 = 1e-5, rtol = 5e-5,
rtol1 = 1e-5, miter 30.  `elastic_type isotropic` is the Lame stiffness:
its compliance has shear terms 2(1+nu)/e (mod_crystals.f:1795-1832), i.e.
mu = e / (2 (1 + nu)).  The material's `tolerance` (dmatprp(100)) is stored
by the reference but never read by mm10 (drive_eps_sig.f:732), so it is
ignored here.
"""
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .configs import Material

# slip_type names of incrystal.f:57-84 that this solver implements (all but
# single, hcp6, hcp18 and halite)
SLIP_TYPES = {"fcc": 1, "bcc": 2, "roters": 6, "bcc12": 7, "bcc48": 8}
BC_NAMES = ["xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"]


class DeckError(ValueError):
    pass


@dataclass
class Crystal:
    slip_type: int = 1
    elastic_type: int = 1        # 1 isotropic, 2 cubic
    e: float = 69000.0
    nu: float = 0.33
    mu: float = None             # cubic shear modulus (elastic_type cubic)
    harden_n: float = 20.0
    theta_0: float = 100.0
    tau_y: float = 0.0
    tau_v: float = 0.0
    voche_m: float = 1.0
    gamma_bar: float = 1.0e10
    alter_mode: bool = False
    hardening: str = "voce"
    atol: float = 1.0e-5
    atol1: float = 1.0e-5
    rtol: float = 5.0e-5
    rtol1: float = 1.0e-5
    miter: int = 30


@dataclass
class CPMaterial:
    model: str = "cp"
    crystal_type: int = 1
    angles: tuple = None          # `angles a b c` (orientation_input single)
    orientation_file: str = None  # `filename '...'` (orientation_input file)
    angle_type: str = "degrees"
    angle_convention: str = "kocks"
    n_crystals: int = 1


@dataclass
class Deck:
    N: int = 0
    crystals: dict = field(default_factory=dict)
    materials: dict = field(default_factory=dict)
    elements: list = field(default_factory=list)   # material name per element (1-based order)
    sizes: tuple = (1.0, 1.0, 1.0)
    FP_max: np.ndarray = field(default_factory=lambda: np.zeros(9))
    isNBC: np.ndarray = field(default_factory=lambda: np.zeros(9, np.int32))
    mults: dict = field(default_factory=dict)      # step -> constraints multiplier
    maxIter: int = None
    tolNR: float = None
    tolPCG: float = None
    dt: float = None
    computed: bool = False
    directory: str = "."
    stname: str = ""                               # structure name (project)
    out_steps: list = field(default_factory=list)  # steps of `output results`

    # ----- what the solver takes
    @property
    def nstep(self):
        return max(self.mults) if self.mults else 0

    def bc(self):
        """(BC_all [nstep][9], isNBC [9]) exactly as inlod fills them
        (inlod.f:57-63): the running sum of FP_max * mults(step), plus one
        on the diagonal components whose F is prescribed."""
        n = self.nstep
        m = np.array([self.mults.get(s, 0.0) for s in range(1, n + 1)])
        BC = np.zeros((n, 9))
        for s in range(n):
            BC[s] = (BC[s - 1] if s else 0.0) + self.FP_max * m[s]
        for d in (0, 4, 8):
            if not self.isNBC[d]:
                BC[:, d] += 1.0
        return BC, self.isNBC.copy()

    def material(self):
        """cpfft Material of the deck's single mm10 crystal type"""
        names = sorted(set(self.elements))
        kinds = {self.materials[n].crystal_type for n in names}
        if len(kinds) != 1:
            raise DeckError("the solver takes one crystal type for all voxels")
        cr = self.crystals[kinds.pop()]
        if cr.hardening != "voce":
            raise DeckError(f"hardening {cr.hardening} is not implemented (Voce only)")
        if cr.slip_type not in SLIP_TYPES.values():
            raise DeckError(f"slip_type {cr.slip_type} is not implemented ({', '.join(SLIP_TYPES)})")
        if cr.elastic_type == 1:
            mu = cr.e / (2.0 * (1.0 + cr.nu))
        elif cr.elastic_type == 2:
            if cr.mu is None:
                raise DeckError("elastic_type cubic needs mu")
            mu = cr.mu
        else:
            raise DeckError("elastic_type orthotropic is not implemented")
        return Material(slip_set=cr.slip_type, e=cr.e, nu=cr.nu, mu=mu, rate_n=cr.harden_n,
                        tau_y=cr.tau_y, tau_v=cr.tau_v, theta_0=cr.theta_0,
                        voche_m=cr.voche_m, gamma_bar=cr.gamma_bar,
                        alter_mode=int(cr.alter_mode), atol=cr.atol, atol1=cr.atol1,
                        rtol=cr.rtol, rtol1=cr.rtol1, miter=cr.miter, jac_exact=0)

    def angles(self):
        """per-voxel Kocks angles in degrees, [3][N^3], voxel = element - 1"""
        V = self.N ** 3
        if len(self.elements) != V:
            raise DeckError(f"{len(self.elements)} elements for a {self.N}^3 grid")
        out = np.zeros((3, V))
        by_file = {}
        for e, name in enumerate(self.elements):
            m = self.materials[name]
            if m.orientation_file:
                path = os.path.join(self.directory, m.orientation_file)
                if path not in by_file:
                    by_file[path] = read_angle_file(path, V)
                a = by_file[path][:, e]
            elif m.angles is not None:
                a = np.asarray(m.angles, float)
            else:
                raise DeckError(f"material {name} has no orientation")
            if m.angle_type == "radians":
                a = np.degrees(a)
            out[:, e] = a
        return out


# ------------------------------------------------------------------ scanner
_TOKEN = re.compile(r"'[^']*'|\"[^\"]*\"|,|[^\s,]+")


def _tokens(line):
    return _TOKEN.findall(line)


def _is_comment(toks):
    return not toks or toks[0].lower() == "c"


def _num(tok):
    return float(tok.lower().replace("d", "e"))


def _int_list(toks):
    """integer list of the scanner's list syntax (scan.f trscan_list):
    n, n-m, n to m, with `by k`, separated by blanks or commas"""
    out = []
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == ",":
            i += 1
            continue
        m = re.fullmatch(r"(\d+)-(\d+)", t)
        if m or (i + 2 < len(toks) and toks[i + 1].lower() == "to"):
            if m:
                a, b = int(m.group(1)), int(m.group(2))
                i += 1
            else:
                a, b = int(t), int(toks[i + 2])
                i += 3
            step = 1
            if i + 1 < len(toks) and toks[i].lower() == "by":
                step = int(toks[i + 1])
                i += 2
            out.extend(range(a, b + 1, step))
        elif t.isdigit():
            out.append(int(t))
            i += 1
        else:
            break
    return out, toks[i:]


def read_angle_file(path, nelem):
    """Angle file of inmat_cp (`orientation_input file`): records
    `element, psi, theta, phi`, read in order as read_defs does
    (mod_crystals.f:2233-2300): for each element the file is read forward to
    the record of that element.  Returns [3][nelem]."""
    recs = []
    with open(path) as f:
        for ln in f:
            vals = [v for v in re.split(r"[\s,]+", ln.strip()) if v]
            if len(vals) >= 4:
                recs.append((int(float(vals[0])), [_num(v) for v in vals[1:4]]))
    out = np.zeros((3, nelem))
    pos = 0
    for el in range(1, nelem + 1):
        while pos < len(recs) and recs[pos][0] != el:
            pos += 1
        if pos == len(recs):
            raise DeckError(f"{path}: no record for element {el}")
        out[:, el - 1] = recs[pos][1]
        pos += 1
    return out


def _property_list(lines, i, first):
    """Tokens of a property list that starts with `first` (the rest of the
    line after `properties`).  A comma ends the line's part of the list and
    the list goes on with the next line; a line without a comma ends it."""
    toks = []
    cur = first
    while True:
        if "," in cur:
            toks += cur[:cur.index(",")]
            if i >= len(lines):
                break
            cur = _tokens(lines[i])
            i += 1
            continue
        toks += cur
        break
    return toks, i


def _crystal_props(cr, toks):
    i = 0
    kv = {"e": "e", "nu": "nu", "mu": "mu", "harden_n": "harden_n", "theta_0": "theta_0",
          "tau_y": "tau_y", "tau_v": "tau_v", "voche_m": "voche_m", "voce_m": "voche_m",
          "gamma_bar": "gamma_bar", "eps_dot_0_y": "gamma_bar", "atol": "atol",
          "atol1": "atol1", "rtol": "rtol", "rtol1": "rtol1"}
    while i < len(toks):
        k = toks[i].lower()
        if k == "slip_type":
            name = toks[i + 1].lower()
            if name not in SLIP_TYPES:
                raise DeckError(f"slip_type {name} is not implemented ({', '.join(SLIP_TYPES)})")
            cr.slip_type = SLIP_TYPES[name]
            i += 2
        elif k == "elastic_type":
            cr.elastic_type = {"isotropic": 1, "cubic": 2, "orthotropic": 3}[toks[i + 1].lower()]
            i += 2
        elif k == "hardening":
            cr.hardening = toks[i + 1].lower()
            i += 2
        elif k == "alter_mode":
            cr.alter_mode = toks[i + 1].lower() == "on"
            i += 2
        elif k == "miter":
            cr.miter = int(toks[i + 1])
            i += 2
        elif k in kv:
            setattr(cr, kv[k], _num(toks[i + 1]))
            i += 2
        else:
            raise DeckError(f"crystal property {toks[i]} is not supported")


def _material_props(m, toks):
    i = 0
    while i < len(toks):
        k = toks[i].lower()
        if k == "cp":
            i += 1
        elif k in ("rho", "alpha", "tolerance"):
            i += 2
        elif k == "debug":
            i += 2
        elif k == "angle_convention":
            if toks[i + 1].lower() != "kocks":
                raise DeckError("only angle_convention kocks exists (inmat.f:176-189)")
            i += 2
        elif k == "angle_type":
            m.angle_type = toks[i + 1].lower()
            i += 2
        elif k == "n_crystals":
            m.n_crystals = int(toks[i + 1])
            if m.n_crystals != 1:
                raise DeckError("one crystal per material point")
            i += 2
        elif k in ("crystal_input", "orientation_input"):
            i += 2
        elif k == "crystal_type":
            m.crystal_type = int(toks[i + 1])
            i += 2
        elif k == "filename":
            m.orientation_file = toks[i + 1].strip("'\"")
            i += 2
        elif k == "angles":
            m.angles = tuple(_num(t) for t in toks[i + 1:i + 4])
            i += 4
        else:
            raise DeckError(f"material property {toks[i]} is not supported (cp materials only)")


def read_deck(path):
    """Parse a deck file; returns a Deck (paths of angle files are taken
    relative to the deck's directory)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    d = Deck(directory=os.path.dirname(os.path.abspath(path)))
    i = 0

    def next_card():
        nonlocal i
        while i < len(lines):
            toks = _tokens(lines[i])
            i += 1
            if not _is_comment(toks):
                return toks
        return None

    toks = next_card()
    while toks is not None:
        cmd = toks[0].lower()
        if cmd == "project":
            rest = toks[1:]
            if rest and rest[0].lower() == "name":
                rest = rest[1:]
            if not rest:
                raise DeckError("project: a name must follow")
            d.stname = rest[0].strip("'\"")[:8]
            toks = next_card()
        elif cmd == "output":
            rest = [t.lower() for t in toks[1:]]
            if rest and rest[0].startswith("result"):
                if len(rest) < 2 or not rest[1].startswith("step"):
                    raise DeckError("output results: steps <list> must follow")
                steps, _ = _int_list(toks[3:])
                if not steps or min(steps) <= 0:
                    raise DeckError("output results: a list of positive steps")
                d.out_steps = sorted(set(steps))  # the last `output results` wins (oudriv.f:129)
            toks = next_card()
        elif cmd == "blocking":
            toks = next_card()
        elif cmd == "number":
            rest = [t for t in toks[1:] if t.lower() not in ("of", "grid")]
            d.N = int(rest[0])
            toks = next_card()
        elif cmd == "crystal":
            k = int(toks[1])
            props = next_card()
            if props is None or props[0].lower() != "properties":
                raise DeckError("crystal: a properties line must follow")
            plist, i = _property_list(lines, i, props[1:])
            cr = Crystal()
            _crystal_props(cr, plist)
            d.crystals[k] = cr
            toks = next_card()
        elif cmd == "material":
            name = toks[1]
            props = next_card()
            if props is None or props[0].lower() != "properties":
                raise DeckError("material: a properties line must follow")
            plist, i = _property_list(lines, i, props[1:])
            if not plist or plist[0].lower() != "cp":
                d.materials[name] = CPMaterial(model=plist[0].lower() if plist else "?")
            else:
                m = CPMaterial()
                _material_props(m, plist)
                d.materials[name] = m
            toks = next_card()
        elif cmd == "sizes":
            vals = {}
            rest = toks[1:]
            for j, t in enumerate(rest):
                for ax in ("x", "y", "z"):
                    if t.lower().startswith(ax + "_dir"):
                        vals[ax] = _num(rest[j + 1])
            d.sizes = (vals.get("x", 1.0), vals.get("y", 1.0), vals.get("z", 1.0))
            toks = next_card()
        elif cmd == "elements":
            emap = {}
            toks = next_card()
            while toks is not None:
                ids, rest = _int_list(toks)
                if not ids or not rest or rest[0].lower() != "material":
                    break
                for e in ids:
                    emap[e] = rest[1]
                toks = next_card()
            n = max(emap) if emap else 0
            if sorted(emap) != list(range(1, n + 1)):
                raise DeckError("elements must cover 1..n")
            d.elements = [emap[e] for e in range(1, n + 1)]
        elif cmd == "strains":
            d.FP_max = np.zeros(9)
            d.isNBC = np.zeros(9, np.int32)
            toks = next_card()
            while toks is not None and re.fullmatch(r"[FPfp]_[xyz]{2}", toks[0]):
                comp = BC_NAMES.index(toks[0][2:].lower())
                d.FP_max[comp] = _num(toks[1])
                d.isNBC[comp] = 1 if toks[0][0] in "Pp" else 0
                toks = next_card()
        elif cmd == "loading":
            toks = next_card()
            while toks is not None and toks[0].lower() == "step":
                ids, rest = _int_list(toks[1:])
                rest = [t for t in rest if t != ","]
                if not rest or not rest[0].lower().startswith("constr"):
                    raise DeckError("loading: step <list> constraints <value>")
                for s in ids:
                    d.mults[s] = _num(rest[1])
                toks = next_card()
        elif cmd == "nonlinear":
            toks = next_card()
            while toks is not None:
                w = toks[0].lower()
                if w == "maximum":
                    d.maxIter = int([t for t in toks[1:] if t.isdigit()][0])
                elif w == "convergence":
                    rest = toks[1:]
                    for j, t in enumerate(rest):
                        if t == "NR":
                            d.tolNR = _num(rest[j + 1])
                        elif t == "CG":
                            d.tolPCG = _num(rest[j + 1])
                elif w == "time":
                    d.dt = _num([t for t in toks[1:] if t.lower() != "step"][0])
                else:
                    break
                toks = next_card()
        elif cmd == "compute":
            d.computed = True
            toks = next_card()
        elif cmd == "stop":
            break
        else:
            raise DeckError(f"unknown command {toks[0]}")
    for k, v in (("number of grid", d.N), ("time step", d.dt), ("maximum iterations", d.maxIter),
                 ("NR tolerance", d.tolNR), ("CG tolerance", d.tolPCG)):
        if not v:
            raise DeckError(f"the deck sets no {k} (the reference has no default, mod_fft.f:40-42)")
    return d
