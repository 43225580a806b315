"""Physics tables: PUMAS dE/dx loader, MDF recipes and the MaterialTable layout.

Mirrors the reference's table interface (src/Physics.jl:64-116, :534-599;
src/DEDXLoader.jl:51-191).  A PhysicsTables object holds the 19 arrays of
every material on one knot grid as a numpy array [material][array][knot]
(the layout dp_upload_tables copies to HBM) and uploads itself lazily to
the device of a Context.
"""
import ctypes as C
import math
import os
import re
import warnings
import xml.etree.ElementTree as ET

import numpy as np

from . import _lib
from ._lib import (T_CSDA_RANGE, T_CSDA_STOP, T_CSDA_TIME, T_ENERGY, T_LOG_ENERGY,
                   T_MAGNETIC_ROTATION, T_NARRAYS, T_SPIN_FACTOR, T_IONIZATION, T_BREMS, T_PAIR,
                   T_PHOTO)

MUON_MASS = 0.10565839        # src/Constants.jl:47
MUON_C_TAU = 658.654          # src/Constants.jl:26
LARMOR_FACTOR = 0.299792458   # src/Constants.jl:29


class DEDXData:
    """Parsed PUMAS dE/dx table (src/DEDXLoader.jl:24-43), GeV and kg/m2 units."""

    def __init__(self, name, mass, zoa, cols):
        self.material_name = name
        self.particle_mass = mass
        self.ZoA = zoa
        (self.kinetic_energy, self.momentum, self.ionization, self.bremsstrahlung,
         self.pair_production, self.photonuclear, self.radiative_total, self.stopping_power,
         self.csda_range, self.density_effect, self.beta) = [np.array(c, dtype=np.float64)
                                                              for c in cols]


def _parse_header(lines):
    """parse_header, src/DEDXLoader.jl:51-92."""
    name, mass, zoa, start = "", 0.0, 0.0, None
    for i, line in enumerate(lines):
        s = line.strip()
        if "Incident particle is a Muon with M =" in s:
            m = re.search(r"M = ([\d.]+) MeV", s)
            if m:
                mass = float(m.group(1)) * 1e-3
        if "Index =" in s and ":" in s:
            m = re.search(r"Index = \d+: (\w+)", s)
            if m:
                name = m.group(1)
        if "<Z/A>" in s:
            m = re.search(r"<Z/A> = ([\d.]+)", s)
            if m:
                zoa = float(m.group(1))
        if "[MeV]" in s and "[MeV/c]" in s:
            start = i + 1
            break
    return name, mass, zoa, start


def load_dedx_file(path):
    """load_dedx_file, src/DEDXLoader.jl:100-191.

    The data section starts after the units line; when a file has no units
    line but a column header (as in test/test_dedx_first_record.jl), the data
    starts at the first row that parses as 11 numbers, so the first physical
    record is kept.
    """
    with open(path) as f:
        lines = f.read().splitlines()
    name, mass, zoa, start = _parse_header(lines)
    if start is None:
        start = 0
    cols = [[] for _ in range(11)]
    conv_dedx, conv_range = 1e-4, 10.0
    for line in lines[start:]:
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        parts = s.split()
        if len(parts) < 11:
            continue
        try:
            v = [float(p) for p in parts[:11]]
        except ValueError:
            continue
        cols[0].append(v[0] * 1e-3)
        cols[1].append(v[1] * 1e-3)
        for j in range(2, 8):
            cols[j].append(v[j] * conv_dedx)
        cols[8].append(v[8] * conv_range)
        cols[9].append(v[9])
        cols[10].append(v[10])
    if not cols[0]:
        raise ValueError(f"Could not find data section in DEDX file {path}")
    return DEDXData(name, mass, zoa, cols)


# --------------------------------------------------------------- MDF recipes
class Element:
    def __init__(self, name, Z, A, I):
        self.name, self.Z, self.A, self.I = name, float(Z), float(A), float(I)


class BaseMaterial:
    """Material recipe: elements with mass fractions (src/Materials.jl:29-133)."""

    def __init__(self, name, density, elements, fractions, I=None):
        self.name = name
        self.density = float(density)
        self.elements = list(elements)
        tot = sum(fractions)
        self.fractions = [f / tot for f in fractions]
        self.ZoA = sum(f * e.Z / e.A for e, f in zip(self.elements, self.fractions))
        if I is None:
            # mean excitation energy from the fraction-weighted log I of the
            # elements (Bragg additivity on Z/A-weighted log I)
            num = sum(f * e.Z / e.A * math.log(e.I) for e, f in zip(self.elements, self.fractions))
            I = math.exp(num / self.ZoA)
        self.I = float(I)


def builtin_materials():
    """The reference's built-in recipes (ELEMENTS, STANDARD_ROCK, AIR, WATER;
    src/Materials.jl:35-127): elements with Z, A and I (GeV), material
    density (kg/m3), mean excitation energy and mass fractions.  These are the
    materials create_physics tabulates when it has no PUMAS dE/dx file."""
    E = {n: Element(n, Z, A, I * 1e-9) for n, Z, A, I in (
        ("H", 1, 1.00794, 19.2), ("C", 6, 12.0107, 78.0), ("N", 7, 14.0067, 82.0),
        ("O", 8, 15.9994, 95.0), ("Na", 11, 22.9898, 149.0), ("Mg", 12, 24.305, 156.0),
        ("Al", 13, 26.9815, 166.0), ("Si", 14, 28.0855, 173.0), ("K", 19, 39.0983, 190.0),
        ("Ca", 20, 40.078, 191.0), ("Fe", 26, 55.845, 286.0), ("Ar", 18, 39.948, 188.0))}
    mats = {
        "StandardRock": (2650.0, 136.4e-9, ("O", "Si", "Al", "Fe", "Ca", "Na", "Mg", "K"),
                         (0.466, 0.277, 0.081, 0.050, 0.036, 0.028, 0.021, 0.026)),
        "Air": (1.205, 85.7e-9, ("N", "O", "Ar"), (0.755, 0.232, 0.013)),
        "Water": (1000.0, 75.0e-9, ("H", "O"), (0.1119, 0.8881)),
    }
    return {n: BaseMaterial(n, rho, [E[e] for e in els], fr, I)
            for n, (rho, I, els, fr) in mats.items()}


def create_energy_grid(n_energies, K_min, K_max):
    """create_energy_grid (src/Physics.jl:125-127): log-spaced kinetic energies."""
    return np.array([K_min * (K_max / K_min) ** (i / (n_energies - 1))
                     for i in range(n_energies)])


def electronic_stopping_power(material, energy, mass=MUON_MASS):
    """electronic_stopping_power (src/Materials.jl:486-520), GeV m2/kg."""
    r, keep = _recipe(material)
    return float(_lib.load().dp_electronic_stopping_power(C.byref(r), mass, float(energy)))


def electronic_density_effect(material, gamma):
    """The Fano density effect delta(gamma) (src/Materials.jl:442-480)."""
    r, keep = _recipe(material)
    return float(_lib.load().dp_density_effect(C.byref(r), float(gamma)))


def load_mdf(path):
    """Elements, materials and composites of a PUMAS MDF file (src/Loader.jl:115-287)."""
    root = ET.parse(path).getroot()
    elements = {}
    for e in root.findall("element"):
        elements[e.get("name")] = Element(e.get("name"), e.get("Z"), e.get("A"),
                                          float(e.get("I")) * 1e-9)
    materials = {}
    for m in root.findall("material"):
        comps = m.findall("component")
        els = [elements[c.get("name")] for c in comps]
        fr = [float(c.get("fraction")) for c in comps]
        I = float(m.get("I")) * 1e-9 if m.get("I") is not None else None
        materials[m.get("name")] = BaseMaterial(m.get("name"), float(m.get("density")) * 1e3,
                                                els, fr, I)
    composites = {}
    for m in root.findall("composite"):
        composites[m.get("name")] = [(c.get("name"), float(c.get("fraction")))
                                     for c in m.findall("component")]
    return elements, materials, composites


# ------------------------------------------------------------ table building
def dedx_file_name(material):
    """File naming of find_dedx_file (src/DEDXLoader.jl:199-245)."""
    snake = re.sub(r"(?<!^)(?=[A-Z])", "_", material).lower()
    return f"{snake}.txt"


def csda_columns_from_dedx(dedx, mass):
    """The columns create_material_table_from_dedx copies from the file and
    the integrated proper time (src/Physics.jl:146-181), plus the
    kinematic columns (:256-263).  Returns {array index: values}."""
    n = len(dedx.kinetic_energy)
    e = dedx.kinetic_energy.copy()
    s = dedx.stopping_power.copy()
    t = np.zeros(n)
    for i in range(1, n):
        K, Kp = e[i], e[i - 1]
        g = 1.0 + K / mass
        gp = 1.0 + Kp / mass
        b = math.sqrt(1.0 - 1.0 / (g * g))
        bp = math.sqrt(1.0 - 1.0 / (gp * gp))
        gm = (g + gp) / 2
        bm = (b + bp) / 2
        avg = (s[i] + s[i - 1]) / 2
        dK = K - Kp
        t[i] = t[i - 1] + dK / (avg * bm * gm)
    mag = np.zeros(n)
    spin = np.zeros(n)
    for i in range(n):
        K = e[i]
        g = 1.0 + K / mass
        b = math.sqrt(1.0 - 1.0 / (g * g))
        p = math.sqrt(K * (K + 2 * mass))
        en = K + mass
        spin[i] = K * (en + mass) / (en * en)
        mag[i] = LARMOR_FACTOR / (b * p)
    return {T_ENERGY: e, T_LOG_ENERGY: np.log(e), T_CSDA_STOP: s,
            T_CSDA_RANGE: dedx.csda_range.copy(), T_CSDA_TIME: t,
            T_IONIZATION: dedx.ionization.copy(), T_BREMS: dedx.bremsstrahlung.copy(),
            T_PAIR: dedx.pair_production.copy(), T_PHOTO: dedx.photonuclear.copy(),
            T_MAGNETIC_ROTATION: mag, T_SPIN_FACTOR: spin}


class PhysicsTables:
    """PhysicsTables{Float64} (src/Physics.jl:109-116) in the device layout."""

    def __init__(self, names, data, scalars, mass=MUON_MASS, ctau=MUON_C_TAU, recipes=None,
                 complete=False):
        self.names = list(names)
        self.materials = {n: i for i, n in enumerate(self.names)}
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self.scalars = np.ascontiguousarray(scalars, dtype=np.float64)
        self.mass = float(mass)
        self.ctau = float(ctau)
        self.recipes = recipes or {}
        self.complete = complete  # False: only the CSDA columns are filled
        self.cutoff = 0.05
        self.elastic_ratio = 0.05
        self._dev = {}

    @property
    def n_materials(self):
        return self.data.shape[0]

    @property
    def n_knots(self):
        return self.data.shape[2]

    def table(self, m, a):
        return self.data[m, a]

    def get_material_index(self, name):
        """get_material_index, src/Physics.jl:606-608 (0-based, -1 if absent)."""
        return self.materials.get(name, -1)

    def supported_max(self):
        return float(self.data[:, T_ENERGY, -1].min())

    def device_handle(self, ctx):
        """Upload once per context; returns the dp_tables handle."""
        key = id(ctx)
        h = self._dev.get(key)
        if h is None:
            n_el, Z, A, I, f = self.recipe_arrays()
            self._keep = (n_el, Z, A, I, f)
            desc = _lib.TableDesc(self.n_materials, self.n_knots, _lib.ptr(self.data),
                                  _lib.ptr(self.scalars), self.mass, self.ctau,
                                  _lib.ptr(n_el, C.c_int32), _lib.ptr(Z), _lib.ptr(A),
                                  _lib.ptr(I), _lib.ptr(f), self.cutoff, self.elastic_ratio)
            out = C.c_void_p()
            _lib.check(ctx.lib.dp_upload_tables(ctx.h, C.byref(desc), C.byref(out)))
            h = out
            self._dev[key] = (ctx, h)
            return h
        return h[1]

    def recipe_arrays(self):
        """Element recipes in the flat layout of dp_table_desc (materials
        without a recipe get zero elements)."""
        n_el, Z, A, I, f = [], [], [], [], []
        for name in self.names:
            r = self.recipes.get(name)
            els = list(zip(r.elements, r.fractions)) if r is not None else []
            n_el.append(len(els))
            for e, fr in els:
                Z.append(e.Z); A.append(e.A); I.append(e.I); f.append(fr)
        return (np.array(n_el, dtype=np.int32), np.array(Z + [0.0]), np.array(A + [0.0]),
                np.array(I + [0.0]), np.array(f + [0.0]))

    def release(self):
        for ctx, h in self._dev.values():
            if ctx.h:
                ctx.lib.dp_free_tables(ctx.h, h)
        self._dev.clear()


def default_data_dir():
    """The package's material data: PUMAS dE/dx tables (pumas-materials
    dedx/muon) and the element recipes materials.xml."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def create_physics_csda(materials=("StandardRock", "Air"), data_dir=None, mdf=None):
    """create_physics (src/Physics.jl:534-599) for tabulated materials, CSDA
    columns only (energies, log energies, stopping power, range, proper time,
    per-process stopping powers, magnetic rotation, spin factor)."""
    data_dir = data_dir or default_data_dir()
    mdf = mdf or os.path.join(data_dir, "materials.xml")
    _, recipes, _ = load_mdf(mdf)
    names = list(materials)
    cols, scal = [], []
    K = None
    for name in names:
        path = os.path.join(data_dir, "pumas-materials", "dedx", "muon", dedx_file_name(name))
        d = load_dedx_file(path)
        c = csda_columns_from_dedx(d, MUON_MASS)
        n = len(d.kinetic_energy)
        if K is None:
            K = n
        elif n != K:
            raise ValueError("all materials must share one energy grid")
        arr = np.zeros((T_NARRAYS, n))
        for k, v in c.items():
            arr[k] = v
        cols.append(arr)
        rec = recipes.get(name)
        scal.append([rec.density if rec else 0.0, rec.I if rec else 0.0, d.ZoA, 0.0])
    return PhysicsTables(names, np.stack(cols), np.array(scal), recipes=recipes, complete=False)


# ----------------------------------------------------- full table builder
def _recipe(mat):
    Z = np.array([e.Z for e in mat.elements])
    A = np.array([e.A for e in mat.elements])
    I = np.array([e.I for e in mat.elements])
    f = np.array(mat.fractions, dtype=np.float64)
    keep = (Z, A, I, f)
    r = _lib.MaterialRecipe(len(Z), _lib.ptr(Z), _lib.ptr(A), _lib.ptr(I), _lib.ptr(f), mat.I,
                            mat.density, mat.ZoA)
    return r, keep


def flatten_composite(name, components, fractions):
    """flatten_to_base_material (src/Materials.jl:1081-1118): merged element
    fractions, harmonic-mean density (composite_density, :1066-1074)."""
    tot = sum(fractions)
    fr = [f / tot for f in fractions]
    merged = {}
    order = []
    for cf, m in zip(fr, components):
        for e, ef in zip(m.elements, m.fractions):
            if e.name not in merged:
                merged[e.name] = [e, 0.0]
                order.append(e.name)
            merged[e.name][1] += cf * ef
    els = [merged[n][0] for n in order]
    efs = [merged[n][1] for n in order]
    rho_inv = sum(f / m.density for f, m in zip(fr, components))
    density = sum(fr) / rho_inv
    return BaseMaterial(name, density, els, efs)


def create_physics(materials=("StandardRock", "Water", "Air"), composites=(), data_dir=None,
                   mdf=None, cutoff=0.05, elastic_ratio=0.05, n_energies=200, K_min=1e-3,
                   K_max=1e9, dedx_dir=None):
    """create_physics (src/Physics.jl:534-599): every MaterialTable array,
    built by the library's host table builder.  A material with a PUMAS dE/dx
    file in dedx_dir (default: the package's pumas-materials/dedx/muon) is
    tabulated on the file's knots with its MDF recipe
    (create_material_table_from_dedx); one without it, as in the reference,
    on create_energy_grid(n_energies, K_min, K_max) from the empirical model
    (create_material_table_empirical) with the reference's built-in recipe
    of that name (or the MDF one), with a warning.  All materials must end
    up on one knot grid."""
    data_dir = data_dir or default_data_dir()
    mdf = mdf or os.path.join(data_dir, "materials.xml")
    dedx_dir = dedx_dir or os.path.join(data_dir, "pumas-materials", "dedx", "muon")
    _, recipes, comp_defs = load_mdf(mdf)
    builtin = builtin_materials()
    lib = _lib.load()
    settings = _lib.PhysicsSettings(cutoff, elastic_ratio, MUON_MASS)
    names, tabs, scal = [], [], []
    K = None
    for name in materials:
        path = os.path.join(dedx_dir, dedx_file_name(name))
        if os.path.isfile(path):
            d = load_dedx_file(path)
            n = len(d.kinetic_energy)
            rec = recipes[name]
            r, keep = _recipe(rec)
            cols = [d.kinetic_energy, d.ionization, d.bremsstrahlung, d.pair_production,
                    d.photonuclear, d.stopping_power, d.csda_range]
            cols = [np.ascontiguousarray(c) for c in cols]
            dc = _lib.DedxColumns(n, *[_lib.ptr(c) for c in cols], d.ZoA)
            out = np.zeros((T_NARRAYS, n))
            _lib.check(lib.dp_build_material_table(C.byref(r), C.byref(dc), C.byref(settings),
                                                   _lib.ptr(out)))
            zoa = d.ZoA
        else:
            warnings.warn(f"PUMAS dE/dx file not found for {name} at {path}. "
                          "Using empirical formulas.")
            rec = builtin.get(name) or recipes[name]
            recipes[name] = rec
            energies = np.ascontiguousarray(create_energy_grid(n_energies, K_min, K_max))
            n = len(energies)
            r, keep = _recipe(rec)
            out = np.zeros((T_NARRAYS, n))
            _lib.check(lib.dp_build_material_table_empirical(C.byref(r), _lib.ptr(energies), n,
                                                             C.byref(settings), _lib.ptr(out)))
            zoa = rec.ZoA
        K = n if K is None else K
        if n != K:
            raise ValueError("all materials must share one energy grid")
        names.append(name)
        tabs.append(out)
        scal.append([rec.density, rec.I, zoa, 0.0])
    for cname in composites:
        parts = comp_defs[cname]
        comps = [recipes[p] for p, _ in parts]
        fr = np.array([f for _, f in parts])
        fr = fr / fr.sum()
        flat = flatten_composite(cname, comps, list(fr))
        r, keep = _recipe(flat)
        ptrs = (_lib.dbl_p * len(parts))(*[_lib.ptr(tabs[names.index(p)]) for p, _ in parts])
        out = np.zeros((T_NARRAYS, K))
        _lib.check(lib.dp_build_composite_table(len(parts), ptrs, _lib.ptr(fr), K, C.byref(r),
                                                C.byref(settings), _lib.ptr(out)))
        names.append(cname)
        tabs.append(out)
        scal.append([flat.density, flat.I, flat.ZoA, 0.0])
        recipes[cname] = flat
    ph = PhysicsTables(names, np.stack(tabs), np.array(scal), recipes=recipes, complete=True)
    ph.cutoff, ph.elastic_ratio = cutoff, elastic_ratio
    return ph
