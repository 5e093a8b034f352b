"""Run descriptions shared by the Python bindings, the tests and bench.py.

`ScenarioParams` mirrors the C struct `vx_scenario_params`
(veritas_amd/csrc/vx_params.h): the reference's Input/Particles settings
(Settings.hpp:4-31) plus the scenario values that the reference leaves to user
hooks (veritas.cpp:7-115).  Values are normalised (c/omega_p, m_e c, m_e c^2).
"""
import ctypes
import math


class ScenarioParams(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("np", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("r", ctypes.c_int32),
        ("quadrature_depth", ctypes.c_int32),
        ("modes", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("cfl", ctypes.c_double),
        ("min_efficiency", ctypes.c_double),
        ("refinement_criteria", ctypes.c_double),
        ("size_weight", ctypes.c_double),
        ("pre_length", ctypes.c_double),
        ("post_length", ctypes.c_double),
        ("n_ref", ctypes.c_double),
        ("length", ctypes.c_double),
        ("pmin", ctypes.c_double),
        ("pmax", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("k", ctypes.c_double),
        ("drift", ctypes.c_double),
        ("a0", ctypes.c_double),
        ("n_over_nc", ctypes.c_double),
        ("ramp", ctypes.c_double),
        ("fwhm_cycles", ctypes.c_double),
        ("lambda0", ctypes.c_double),
        ("gap", ctypes.c_double),
    ]

    def replace(self, **kw):
        out = ScenarioParams()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(self))
        for k, v in kw.items():
            setattr(out, k, v)
        return out


LANDAU, TWO_STREAM, LASER = 0, 1, 2


def _base(**kw):
    p = ScenarioParams(kind=LANDAU, nx=64, np=256, levels=1, r=2, quadrature_depth=1,
                       modes=8, seed=42, cfl=0.4, min_efficiency=0.6,
                       refinement_criteria=1e-8, size_weight=0.0, pre_length=0.0,
                       post_length=0.0, n_ref=1.0e24, length=math.pi / 2, pmin=-1.0,
                       pmax=1.0, theta=0.01, amplitude=0.01, k=4.0, drift=0.2, a0=0.5,
                       n_over_nc=0.01, ramp=50.0, fwhm_cycles=10.0, lambda0=1.0e-6, gap=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def landau(**kw):
    """Relativistic 1D1P Landau damping: x in [0, pi/2) with Nx=64, p in [-1, 1]
    with Np=256, f = (1 + 0.01 cos 4x) MJ(theta=0.01), one electron species
    neutralised by immobile ions (SolverManager.cpp:25)."""
    return _base(**kw)


def two_stream(**kw):
    """Relativistic two-stream: two MJ beams (theta=1e-3) drifting at p=+-0.2,
    density 0.5 each, perturbed by 1e-3 * sum_{m=1..8} cos(m x/16 + phi_m) with
    seeded phases; x length 32 pi with Nx=4096, p in [-0.6, 0.6] with Np=4096."""
    d = dict(kind=TWO_STREAM, nx=4096, np=4096, length=32 * math.pi, pmin=-0.6, pmax=0.6,
             theta=1e-3, amplitude=1e-3, drift=0.2, modes=8)
    d.update(kw)
    return _base(**d)


def laser(**kw):
    """Linearly polarised Gaussian pulse (a0 = 0.5, FWHM 10 cycles) entering
    from the left into underdense plasma: plateau density 0.01 n_c after a
    10-wavelength vacuum gap and a 50-wavelength linear ramp, in a
    400-wavelength domain; p in [-2, 2] m_e c.  Base grid Nx = 8192, Np = 512,
    two refinement levels at ratio 2, cells flagged where f > 1e-6 max f (plus
    the reference's gradient terms).  `wavelengths` sets the domain length in
    laser wavelengths (lengths are otherwise in c/omega_p)."""
    n_over_nc = kw.get("n_over_nc", 0.01)
    waves = kw.pop("wavelengths", 400.0)
    lam = 2 * math.pi * math.sqrt(n_over_nc)  # one wavelength in c/omega_p
    # "flag where f > 1e-6 max f" as the reference's criterion: the error
    # estimate adds sizeWeight / fMax * r^(depth - maxDepth) * |f| to its
    # gradient terms (Rectangle.hpp:128-130; Rectangle.cpp:68-74) and flags
    # above refinementCriteria, so sizeWeight = refinementCriteria / 1e-6 flags
    # every cell with f > 1e-6 fMax on the coarsest level (2e-6 fMax one level
    # finer), and the gradient terms add cells where f varies
    d = dict(kind=LASER, nx=8192, np=512, levels=3, r=2, length=waves * lam, pmin=-2.0,
             pmax=2.0, theta=1e-3, a0=0.5, n_over_nc=n_over_nc, ramp=50.0, fwhm_cycles=10.0,
             lambda0=1.0e-6, refinement_criteria=1e-3, min_efficiency=0.6, size_weight=1e3)
    d.update(kw)
    return _base(**d)


# ---------------------------------------------------------------- reference inputs and hooks

MAX_SPECIES = 8

# physical constants of the reference (veritas.hpp:22-29; veritas.cpp:15-16)
C = 299792458.0
EPS0 = 8.854187817e-12
M_E = 9.10938291e-31
Q_E = 1.60217657e-19


class SettingsDesc(ctypes.Structure):
    """The C struct `vx_settings_desc` (veritas_amd/csrc/vx_params.h): the
    reference's Input and Particles (Settings.hpp:4-31) in SI units, plus
    Settings::quadratureDepth.  Per-species arrays hold `n_species` entries."""
    _fields_ = [
        ("minEfficiency", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("refinementCriteria", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("sizeWeight", ctypes.c_double),
        ("preLength", ctypes.c_double),
        ("postLength", ctypes.c_double),
        ("nx", ctypes.c_uint32),
        ("r", ctypes.c_uint32),
        ("Lfinest", ctypes.c_uint32),
        ("quadratureDepth", ctypes.c_int32),
        ("n_species", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("mass", ctypes.c_double * MAX_SPECIES),
        ("charge", ctypes.c_double * MAX_SPECIES),
        ("dp", ctypes.c_double * MAX_SPECIES),
        ("pmin", ctypes.c_double * MAX_SPECIES),
        ("np", ctypes.c_uint32 * MAX_SPECIES),
    ]

    @classmethod
    def make(cls, nx, dx, species, Lfinest=1, r=2, cfl=0.4, minEfficiency=0.6,
             refinementCriteria=1e-8, sizeWeight=0.0, preLength=0.0, postLength=0.0,
             quadratureDepth=1):
        """`species`: list of dicts with mass, charge, np, dp, pmin (SI)."""
        d = cls(minEfficiency=minEfficiency, dx=dx, refinementCriteria=refinementCriteria, cfl=cfl,
                sizeWeight=sizeWeight, preLength=preLength, postLength=postLength, nx=nx, r=r,
                Lfinest=Lfinest, quadratureDepth=quadratureDepth, n_species=len(species))
        if not 1 <= len(species) <= MAX_SPECIES:
            raise ValueError("1 to 8 species")
        for i, s in enumerate(species):
            d.mass[i], d.charge[i] = s["mass"], s["charge"]
            d.np[i], d.dp[i], d.pmin[i] = s["np"], s["dp"], s["pmin"]
        return d


_INIT = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                         ctypes.c_int32)
_FIELD = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_double)
_REFINE = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                           ctypes.c_int32, ctypes.c_int32)
_OVERRIDE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(SettingsDesc))


class HooksC(ctypes.Structure):
    """The C struct `vx_hooks`."""
    _fields_ = [("user", ctypes.c_void_p), ("initial_distribution", _INIT), ("get_by", _FIELD),
                ("get_bz", _FIELD), ("refinement_override", _REFINE),
                ("settings_override", _OVERRIDE)]


class Hooks:
    """The reference's user hooks (Settings.hpp:45-64) as Python callables:
    initial_distribution(x, p, species) -> f (SI), get_by(x, t) / get_bz(x, t)
    -> B at the left edge, refinement_override(x, p, depth, species) -> bool,
    settings_override(desc) edits a SettingsDesc in place.  Only the first is
    required.  The object keeps the C callbacks alive; keep it for as long as
    a run made with it lives."""

    def __init__(self, initial_distribution, get_by=None, get_bz=None, refinement_override=None,
                 settings_override=None):
        self._init = _INIT(lambda u, x, p, s: float(initial_distribution(x, p, s)))
        self._by = _FIELD(lambda u, x, t: float(get_by(x, t))) if get_by else _FIELD()
        self._bz = _FIELD(lambda u, x, t: float(get_bz(x, t))) if get_bz else _FIELD()
        self._ref = (_REFINE(lambda u, x, p, d, s: int(bool(refinement_override(x, p, d, s))))
                     if refinement_override else _REFINE())
        self._ovr = (_OVERRIDE(lambda u, d: settings_override(d.contents))
                     if settings_override else _OVERRIDE())
        self.c = HooksC(None, self._init, self._by, self._bz, self._ref, self._ovr)
