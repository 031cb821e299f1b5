"""The reference's XML configuration files, read for the hip platform.

Restates what alsvinn's setup code reads (alsfvm/src/config/SimulatorSetup.cpp:
102-232, 247-470; alsuq/src/config/Setup.cpp:69-366): the <fvm> section
(grid, equation and its parameters, flux, reconstruction, cfl and integrator
with the "auto" rules, boundary, end time, Python initial data, writer) and
the <uq> section (samples, sampleStart, generator, parameters, statistics).
Options the reference rejects are rejected the same way; options that exist
in the reference but not on this platform (other fluxes, reconstructions,
writers, distributions, diffusion, functionals) raise ConfigError naming
them.

Initial data follow PythonInitialData (alsfvm/src/init/PythonInitialData.cpp:
84-283): the configuration's Python file runs with `from math import *` and
`from numpy import *`, the equation and sample parameters as globals (floats
for length 1, numpy arrays otherwise). A file that defines init_global gets
one array per primitive variable (shape (nx, ny, nz), x fastest in memory)
and nx, ny, nz, ax, ay, az, bx, by, bz; any other file is a per-cell snippet
that sets the primitive variables from x, y, z (cell midpoints) and i, j, k.
Primitive variables become conserved ones as Euler::computeConserved does
(Euler.hpp:236-245).

Sample parameters are drawn as the reference's Runner draws them
(Runner.cpp:42-50, SampleGenerator.cpp:29-56): per sample, per parameter in
name order, per component, from one process-wide std::mt19937_64 with the
draw counter of STLMersenne.cpp:36-85, scaled by Uniform (x*(b-a)+a).
"""
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .sampling import MT19937_64, canonical


class ConfigError(RuntimeError):
    pass


FVM_NODES = {"name", "platform", "boundary", "flux", "endTime", "equation", "equationParameters",
             "reconstruction", "cfl", "integrator", "initialData", "writer", "grid", "diffusion",
             "functionals"}
EQUATIONS = {"burgers": _lib.EQ_BURGERS, "euler1": _lib.EQ_EULER1, "euler2": _lib.EQ_EULER2,
             "euler3": _lib.EQ_EULER3}
FLUXES = {"hll": _lib.FLUX_HLL, "hll3": _lib.FLUX_HLL3, "godunov": _lib.FLUX_GODUNOV,
          "central": _lib.FLUX_CENTRAL, "rusanov": _lib.FLUX_RUSANOV, "roe": _lib.FLUX_ROE}
RECONSTRUCTIONS = {"none": _lib.RECON_NONE, "weno2": _lib.RECON_WENO2, "eno2": _lib.RECON_ENO2,
                   "eno3": _lib.RECON_ENO3, "eno4": _lib.RECON_ENO4, "mc": _lib.RECON_MC,
                   "mm": _lib.RECON_MM, "wenof2": _lib.RECON_WENOF2, "weno3": _lib.RECON_WENO3}
INTEGRATORS = {"forwardeuler": _lib.INT_FE, "rungekutta2": _lib.INT_RK2, "rungekutta3": _lib.INT_RK3,
               "rungekutta4": _lib.INT_RK4}
BOUNDARIES = {"periodic": _lib.BC_PERIODIC, "neumann": _lib.BC_NEUMANN}
PRIMITIVES = {"burgers": ["u"], "euler1": ["rho", "ux", "p"], "euler2": ["rho", "ux", "uy", "p"],
              "euler3": ["rho", "ux", "uy", "uz", "p"]}
STATISTICS = {"meanvar", "structure_cube", "structure_basic", "onepointmoment", "bv"}


def _text(node, path, default=None):
    n = node.find(path)
    if n is None or n.text is None:
        if default is not None:
            return default
        raise ConfigError("missing configuration option %s" % path)
    return n.text.strip()


def _vec(s, typ):
    parts = s.split()
    if len(parts) < 3:
        raise ConfigError("expected three values, got %r" % s)
    return tuple(typ(x) for x in parts[:3])


@dataclass
class Writer:
    type: str
    basename: str
    number_of_saves: int
    write_initial: bool = False


@dataclass
class Stat:
    name: str
    number_of_saves: int
    writer: Writer
    params: dict = field(default_factory=dict)
    write_initial: bool = False


@dataclass
class UQParameter:
    name: str
    length: int
    type: str
    lower: float = 0.0
    upper: float = 1.0


@dataclass
class Config:
    name: str
    platform: str
    lower: tuple
    upper: tuple
    dims: tuple
    boundary: str
    flux: str
    end_time: float
    equation: str
    equation_parameters: dict
    reconstruction: str
    cfl: float
    integrator: str
    python_file: str
    python_program: str
    initial_parameters: dict
    writer: Writer = None
    samples: int = 1
    sample_start: int = 0
    generator: str = "auto"
    uq_parameters: list = field(default_factory=list)
    stats: list = field(default_factory=list)
    has_uq: bool = False
    base_path: str = "."

    @property
    def gamma(self):
        return float(self.equation_parameters.get("gamma", 1.4))

    def scheme(self, precision=_lib.FP64):
        from .ensemble import make_scheme
        return make_scheme(EQUATIONS[self.equation], FLUXES[self.flux],
                           RECONSTRUCTIONS[self.reconstruction], INTEGRATORS[self.integrator],
                           BOUNDARIES[self.boundary], gamma=self.gamma, cfl=self.cfl,
                           precision=precision)


def _writer(node, default_saves=None):
    typ = _text(node, "type")
    base = _text(node, "basename")
    saves = node.find("numberOfSaves")
    n = int(saves.text.strip()) if saves is not None else default_saves
    if n is None:
        raise ConfigError("writer needs numberOfSaves")
    wi = node.find("writeInitialTimestep")
    initial = wi is not None and wi.text.strip().lower() in ("1", "true")
    if typ != "netcdf":
        raise ConfigError("writer type %s is not available on the hip platform (netcdf is)" % typ)
    return Writer(typ, base, n, initial)


def read_config(path):
    """Parses a reference configuration file (<config><fvm>..</fvm><uq>..</uq></config>)."""
    if not os.path.exists(path):
        raise ConfigError("Input file does not exist\n%s" % path)
    base_path = os.path.dirname(os.path.abspath(path))
    root = ET.parse(path).getroot()
    if root.tag != "config":
        root = root.find("config")
        if root is None:
            raise ConfigError("no <config> element")
    fvm = root.find("fvm")
    if fvm is None:
        raise ConfigError("no <fvm> section")
    for child in fvm:
        if child.tag not in FVM_NODES:
            raise ConfigError("Unsupported configuration option %s" % child.tag)

    lower = _vec(_text(fvm, "grid/lowerCorner"), float)
    upper = _vec(_text(fvm, "grid/upperCorner"), float)
    dims = _vec(_text(fvm, "grid/dimension"), int)
    boundary = _text(fvm, "boundary")
    if boundary not in BOUNDARIES:
        raise ConfigError("Unknown boundary conditions %s" % boundary)
    equation = _text(fvm, "equation")
    if equation == "euler":  # the reference's older name for 2D Euler
        equation = "euler2"
    if equation not in EQUATIONS:
        raise ConfigError("equation %s is not available on the hip platform" % equation)
    flux = _text(fvm, "flux")
    if flux not in FLUXES:
        raise ConfigError("flux %s is not available on the hip platform" % flux)
    # the equation/flux pairs of numerical_flux_list.hpp:41-113 that exist here
    if (equation == "burgers") != (flux in ("godunov", "rusanov", "roe")) and flux != "central":
        raise ConfigError("flux %s is not available for %s" % (flux, equation))
    recon = _text(fvm, "reconstruction")
    if recon not in RECONSTRUCTIONS:
        raise ConfigError("reconstruction %s is not available on the hip platform" % recon)
    if recon == "wenof2" and equation == "burgers":
        # ReconstructionFactory.cpp:102-117
        raise ConfigError("We do not support WENO2F for Burgers.")
    cfl_s = _text(fvm, "cfl")
    cfl = (0.9 if recon == "none" else 0.475) if cfl_s == "auto" else float(cfl_s)
    integ = _text(fvm, "integrator")
    if integ == "auto":
        integ = "forwardeuler" if recon == "none" else "rungekutta2"
    if integ not in INTEGRATORS:
        raise ConfigError("integrator %s is not available on the hip platform" % integ)
    eq_params = {}
    ep = fvm.find("equationParameters")
    if ep is not None:
        for c in ep:
            eq_params[c.tag] = float(c.text.strip())
    diff = fvm.find("diffusion")
    if diff is not None and _text(diff, "name", "none") != "none":
        raise ConfigError("diffusion %s is not available on the hip platform" % _text(diff, "name"))
    fun = fvm.find("functionals")
    if fun is not None and len(list(fun)) > 0:
        raise ConfigError("functionals are not available on the hip platform")

    init = fvm.find("initialData")
    if init is None or init.find("python") is None:
        raise ConfigError("Unknown initial data.")
    pyfile = os.path.join(base_path, _text(init, "python"))
    with open(pyfile) as f:
        program = f.read()
    init_params = {k: np.array([v]) for k, v in eq_params.items()}
    pn = init.find("parameters")
    if pn is not None:
        for prm in pn:
            name = _text(prm, "name")
            length = int(_text(prm, "length"))
            if length == 1 and prm.find("value") is not None:
                vals = [float(_text(prm, "value"))]
            else:
                vals = [float(v.text) for v in prm.findall("values/value")]
                if len(vals) == 1 and length > 1:
                    vals = vals * length
            init_params[name] = np.array(vals, dtype=np.float64)

    writer = _writer(fvm.find("writer")) if fvm.find("writer") is not None else None
    cfg = Config(name=_text(fvm, "name", "alsvinn"), platform=_text(fvm, "platform", "hip"),
                 lower=lower, upper=upper, dims=dims, boundary=boundary, flux=flux,
                 end_time=float(_text(fvm, "endTime")), equation=equation,
                 equation_parameters=eq_params, reconstruction=recon, cfl=cfl, integrator=integ,
                 python_file=pyfile, python_program=program, initial_parameters=init_params,
                 writer=writer, base_path=base_path)

    uq = root.find("uq")
    if uq is not None:
        cfg.has_uq = True
        cfg.samples = int(_text(uq, "samples"))
        cfg.sample_start = int(_text(uq, "sampleStart", "0"))
        cfg.generator = _text(uq, "generator", "auto")
        if cfg.generator not in ("auto", "stlmersenne"):
            raise ConfigError("generator %s is not available on the hip platform" % cfg.generator)
        pn = uq.find("parameters")
        for prm in (pn if pn is not None else []):
            typ = _text(prm, "type")
            if typ != "uniform":
                raise ConfigError("distribution %s is not available on the hip platform" % typ)
            cfg.uq_parameters.append(UQParameter(_text(prm, "name"), int(_text(prm, "length")), typ,
                                                 float(_text(prm, "lower", "0")),
                                                 float(_text(prm, "upper", "1"))))
        sn = uq.find("stats")
        for st in (sn if sn is not None else []):
            name = _text(st, "name")
            if name not in STATISTICS:
                raise ConfigError("statistics %s is not available on the hip platform" % name)
            saves = st.find("numberOfSaves")
            n = int(saves.text.strip()) if saves is not None else None
            w = _writer(st.find("writer"), n)
            params = {}
            for key in ("p", "numberOfH", "direction"):
                if st.find(key) is not None:
                    params[key] = float(_text(st, key))
            if name in ("onepointmoment", "bv"):
                # OnePointMoment.cpp:21-23, BoundedVariation.cpp:23-26: integer p
                p = params.get("p")
                if p is None or p != int(p):
                    raise ConfigError("statistics %s needs an integer <p>" % name)
                if name == "bv" and cfg.dims[2] > 1:
                    raise ConfigError("bv is not available for 3D grids (the reference's "
                                      "getTotalVariation rejects them)")
            wi = st.find("writeInitialTimestep")
            initial = wi is not None and wi.text.strip().lower() in ("1", "true")
            cfg.stats.append(Stat(name, n if n is not None else w.number_of_saves, w, params, initial))
    return cfg


class ReferenceSampler:
    """The process-wide std::mt19937_64 and draw counter shared by every
    STLMersenne generator (STLMersenne.cpp:25-85)."""

    def __init__(self):
        self.eng = MT19937_64()
        self.count = 0

    def generate(self, component, sample, dimension):
        if self.count // dimension < sample:
            n = sample * dimension - self.count
            self.eng.discard(n)
            self.count += n
        if self.count % dimension < component:
            n = component - self.count % dimension
            self.eng.discard(n)
            self.count += n
        self.count += 1
        return float(canonical(self.eng.raw(1))[0])


def sample_parameters(cfg):
    """Per sample (sampleStart, sampleStart+1, ...): {name: array} as the
    reference Runner draws them."""
    sampler = ReferenceSampler()
    params = sorted(cfg.uq_parameters, key=lambda p: p.name)
    out = []
    for i in range(cfg.samples):
        s = cfg.sample_start + i
        d = {}
        for prm in params:
            x = np.array([sampler.generate(c, s, prm.length) for c in range(prm.length)])
            d[prm.name] = x * (prm.upper - prm.lower) + prm.lower
        out.append(d)
    return out


def _globals(cfg, params):
    ns = {"__name__": "__alsvinn_initial_data__"}
    exec("from math import *\ntry:\n    from numpy import *\nexcept ImportError:\n    pass\n", ns)
    merged = dict(cfg.initial_parameters)
    merged.update(params or {})
    for k, v in merged.items():
        v = np.asarray(v, dtype=np.float64)
        ns[k] = float(v[0]) if v.size == 1 else v
    return ns


def primitive_volume(cfg, params=None):
    """Primitive variables on the interior cells, shape (nvar, nz, ny, nx)."""
    from .initial_data import midpoints
    nx, ny, nz = cfg.dims
    names = PRIMITIVES[cfg.equation]
    ns = _globals(cfg, params)
    prim = np.zeros((len(names), nz, ny, nx))
    if "init_global" in cfg.python_program:
        exec(compile(cfg.python_program, cfg.python_file, "exec"), ns)
        arrays = [prim[k].transpose(2, 1, 0) for k in range(len(names))]  # (nx, ny, nz) views
        (ax, ay, az), (bx, by, bz) = cfg.lower, cfg.upper
        ns["init_global"](*arrays, nx, ny, nz, ax, ay, az, bx, by, bz)
    else:
        body = "".join("    %s = 0.0\n" % n for n in names)
        body += "".join("    " + line + "\n" for line in cfg.python_program.splitlines())
        body += "".join("    output['%s'] = %s\n" % (n, n) for n in names)
        exec(compile("def initial_data(x, y, z, i, j, k, output):\n" + body, cfg.python_file, "exec"), ns)
        f = ns["initial_data"]
        xs = midpoints(nx, cfg.lower[0], cfg.upper[0])
        ys = midpoints(ny, cfg.lower[1], cfg.upper[1]) if ny > 1 else [cfg.lower[1] + 0.5 * 1.0]
        zs = midpoints(nz, cfg.lower[2], cfg.upper[2]) if nz > 1 else [cfg.lower[2] + 0.5 * 1.0]
        out = {}
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    f(float(xs[i]), float(ys[j]), float(zs[k]), i, j, k, out)
                    for v, n in enumerate(names):
                        prim[v, k, j, i] = float(out[n])
    return prim


def conserved_volume(cfg, ghosts, params=None):
    """The sample's initial conserved volume with ghost layers (zero; the
    device fills them), (nvar, tz, ty, tx) in the reference's layout."""
    from .initial_data import conserved_euler
    prim = primitive_volume(cfg, params)
    nx, ny, nz = cfg.dims
    gy = ghosts if ny > 1 else 0
    gz = ghosts if nz > 1 else 0
    vol = np.zeros((prim.shape[0], nz + 2 * gz, ny + 2 * gy, nx + 2 * ghosts))
    inner = vol[:, gz:gz + nz, gy:gy + ny, ghosts:ghosts + nx]
    if cfg.equation == "burgers":
        inner[0] = prim[0]
    else:
        nsd = prim.shape[0] - 2
        cons = conserved_euler(prim[0], [prim[1 + d] for d in range(nsd)], prim[-1], cfg.gamma)
        for v in range(len(cons)):
            inner[v] = cons[v]
    return vol
