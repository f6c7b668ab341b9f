"""Specification-file front end: a ``.exahype`` project file read the way the
reference's runtime parser reads it, mapped onto this library's solvers and
grids the way the toolkit's solver factory maps it onto kernels.

Reading (E/Parser.cpp:72-183).  Every line outside ``/* ... */`` comments is
one of
  ``key = values``          a runtime parameter,
  ``key const = values``    a compile-time parameter (``const`` is dropped),
  ``word word ...``         a group begin (``solver ADER-DG Name``); group
                            ends (``end ...``) are dropped,
and contributes its tokens, in order, to one token stream: the key, then the
values split at commas and blanks (kept whole if they contain ``{`` or
``}``).  Queries take "the token after the n-th occurrence of A, then after
the m-th occurrence of B" (E/Parser.cpp:199-272); numbers are read with C's
``atoi`` / ``atof`` semantics (leading numeric prefix, 0 otherwise).

Mapping.  ``kernel`` selects the family (Toolkit/src/eu/exahype/solvers/
SolverFactory.java:39-88): ``generic::fluxes::nonlinear`` and
``generic::fluxes::linear`` an ADER-DG solver (``linear`` the isLinear
branch), ``generic::godunov`` and ``generic::musclhancock`` a finite-volumes
solver.  The regular mesh is the runner's: a cubical bounding box of side
max(width), the coarsest level L with side / 3^(L-1) <= maximum-mesh-size (at
least 3; E/solvers/Solver.cpp:147-156, E/runners/Runner.cpp:321-345), and
ceil(width_d / h) cells along axis d (the scaled domain).
"""
import math
import re

_COMMENT_BEGIN = re.compile(r"(/\*)")
_COMMENT_END = re.compile(r"(\*/)")
_GROUP = re.compile(r"^(\s|\t)*([a-zA-Z][^=]+)+")
_CONST_PARAMETER = re.compile(
    r"^(\s|\t)*([A-Za-z](\w|[^a-zA-Z\d\s\t])*)(\s|\t)+const(\s|\t)*=(\s|\t)*(([^\s\t]|,\s*)+)(\s|\t)*$")
_PARAMETER = re.compile(r"^(\s|\t)*([A-Za-z](\w|[^a-zA-Z\d\s\t])*)(\s|\t)*=(\s|\t)*(([^\s\t]|,\s*)+)(\s|\t)*$")
_NO_SPLITTING = re.compile(r"\}|\{")
_COMMA_SEPARATED = re.compile(r"((\w|[^a-zA-Z,\s\t])+)")
_WHITESPACE_SEPARATED = re.compile(r"([^\s\t]+)")
_NAME_MULT = re.compile(r"([A-Za-z]\w*):([0-9]+)")

NO_TOKEN = "notoken"

SOLVER_TYPES = ("ADER-DG", "Finite-Volumes", "Limiting-ADER-DG")
TIME_STEPPINGS = ("global", "globalfixed")


class SpecError(ValueError):
    """An interpretation error of the specification (the reference logs it and
    marks the parser invalid)."""


def atoi(s):
    """C atoi: the leading integer of s, 0 without one."""
    m = re.match(r"\s*([+-]?\d+)", s)
    return int(m.group(1)) if m else 0


def atof(s):
    """C atof for decimal input: the leading floating-point number of s, 0.0 without one."""
    m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", s)
    return float(m.group(1)) if m else 0.0


def tokenize(text):
    """The token stream of a specification (E/Parser.cpp:72-170)."""
    tokens = []
    in_comment = 0
    for line_number, line in enumerate(text.split("\n")):
        if _COMMENT_BEGIN.search(line):
            in_comment += 1
        if _COMMENT_END.search(line):
            in_comment -= 1
        if in_comment < 0:
            raise SpecError("additional multi-line comment end in line %d" % line_number)
        if in_comment > 0:
            continue
        m = _PARAMETER.search(line)
        rhs = None
        if m:
            tokens.append(m.group(2))
            rhs = m.group(6)
        else:
            m = _CONST_PARAMETER.search(line)
            if m:
                tokens.append(m.group(2))
                rhs = m.group(7)
        if rhs is not None:
            if _NO_SPLITTING.search(rhs):
                tokens.append(rhs)
            else:
                tokens.extend(t.group(0) for t in _COMMA_SEPARATED.finditer(rhs))
            continue
        if _GROUP.search(line):
            words = _WHITESPACE_SEPARATED.findall(line)
            if words and words[0] != "end":
                tokens.extend(words)
    if in_comment > 0:
        raise SpecError("a multi-line comment was not closed")
    return tokens


def compute_mesh_level(mesh_size, domain_size):
    """exahype::solvers::Solver::computeMeshLevel (E/solvers/Solver.cpp:147-156)."""
    result = 1
    current = float("inf")
    while current > mesh_size:
        current = domain_size / 3 ** result
        result += 1
    return max(3, result)


class Spec:
    """A parsed specification: the reference parser's queries over the token
    stream.  Solver numbers count from 0 in order of appearance."""

    def __init__(self, text):
        self.tokens = tokenize(text)

    @classmethod
    def from_file(cls, path):
        with open(path) as f:
            return cls(f.read())

    # --- token lookups (E/Parser.cpp:199-272) -----------------------------
    def token_after(self, token0, occurrence0=1, token1=None, occurrence1=1, skip=0):
        t = self.tokens
        i = 0
        while i < len(t) and (t[i] != token0 or occurrence0 > 1):
            if t[i] == token0:
                occurrence0 -= 1
            i += 1
        if token1 is not None:
            while i < len(t) and (t[i] != token1 or occurrence1 > 1):
                if t[i] == token1:
                    occurrence1 -= 1
                i += 1
        i += skip + 1
        return t[i] if i < len(t) else NO_TOKEN

    def _solver_key(self, k, key, skip=0):
        return self.token_after("solver", k + 1, key, 1, skip)

    # --- project and domain ---------------------------------------------
    @property
    def project_name(self):
        return self.token_after("exahype-project")

    @property
    def dimension(self):
        d = atoi(self.token_after("computational-domain", 1, "dimension"))
        if d not in (2, 3):
            raise SpecError("dimension: %r" % d)
        return d

    def _triple(self, key, sentinel):
        out = []
        for i in range(self.dimension):
            tok = self.token_after("computational-domain", 1, key, 1, i)
            if i == 2 and tok == sentinel:
                raise SpecError("%s: not enough values for 3 dimensions" % key)
            out.append(atof(tok))
        return tuple(out)

    @property
    def domain_size(self):
        return self._triple("width", "offset")

    @property
    def offset(self):
        return self._triple("offset", "end-time")

    @property
    def end_time(self):
        t = atof(self.token_after("computational-domain", 1, "end-time"))
        if t <= 0:
            raise SpecError("invalid end-time")
        return t

    # --- solvers ------------------------------------------------------------
    @property
    def num_solvers(self):
        return sum(1 for t in self.tokens if t == "solver")

    def solver_type(self, k):
        tok = self.token_after("solver", k + 1)
        if tok not in SOLVER_TYPES:
            raise SpecError("solver %d: type %r is invalid" % (k, tok))
        return tok

    def identifier(self, k):
        return self.token_after("solver", k + 1, skip=1)

    def _counted(self, k, key):
        """variables / parameters: a number or name:multiplicity entries
        (E/Parser.cpp:628-698)."""
        tok = self._solver_key(k, key)
        n = atoi(tok)
        if n >= 1:
            return n
        n, i = 0, 1
        m = _NAME_MULT.search(tok)
        while m:
            n += int(m.group(2))
            tok = self._solver_key(k, key, i)
            i += 1
            m = _NAME_MULT.search(tok)
        return n

    def variables(self, k):
        n = self._counted(k, "variables")
        if n < 1:
            raise SpecError("solver %d: variables must be greater than zero" % k)
        return n

    def parameters(self, k):
        return self._counted(k, "parameters")

    def order(self, k):
        return atoi(self._solver_key(k, "order"))

    def patch_size(self, k):
        return atoi(self._solver_key(k, "patch-size"))

    def maximum_mesh_size(self, k):
        h = atof(self._solver_key(k, "maximum-mesh-size"))
        if h <= 0.0:
            raise SpecError("solver %d: maximum-mesh-size must be greater than zero" % k)
        return h

    def maximum_mesh_depth(self, k):
        tok = self._solver_key(k, "maximum-mesh-depth")
        return 0 if tok == NO_TOKEN else atoi(tok)

    def time_stepping(self, k):
        tok = self._solver_key(k, "time-stepping")
        if tok not in TIME_STEPPINGS:
            raise SpecError("solver %d: time-stepping %r is invalid" % (k, tok))
        return tok

    def kernel(self, k):
        return self._solver_key(k, "kernel")

    def limiter_kernel(self, k):
        return self._solver_key(k, "limiter-kernel")

    def dmp_relaxation_parameter(self, k):
        return atof(self._solver_key(k, "dmp-relaxation-parameter"))

    def dmp_difference_scaling(self, k):
        return atof(self._solver_key(k, "dmp-difference-scaling"))

    def steps_till_cured(self, k):
        return atoi(self._solver_key(k, "steps-till-cured"))

    # --- optimisation section -------------------------------------------------
    def fuse_algorithmic_steps(self):
        """`fuse-algorithmic-steps` (E/Parser.cpp:452-469): on / off, off when
        absent."""
        tok = self.token_after("optimisation", 1, "fuse-algorithmic-steps")
        if tok == NO_TOKEN:
            return False
        if tok not in ("on", "off"):
            raise SpecError("fuse-algorithmic-steps has to be either on or off: %r" % tok)
        return tok == "on"

    def fuse_algorithmic_steps_factor(self):
        """`fuse-algorithmic-steps-factor` (E/Parser.cpp:486-501): strtod of
        the token, an error unless a number in [0, 1]."""
        tok = self.token_after("optimisation", 1, "fuse-algorithmic-steps-factor")
        m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", tok)
        v = float(m.group(1)) if m else 0.0
        if not m or v < 0.0 or v > 1.0:
            raise SpecError("fuse-algorithmic-steps-factor must lie in [0, 1]: %r" % tok)
        return v

    # --- the regular mesh the runner builds ----------------------------------
    def mesh(self, k):
        """(cells per axis, mesh width h, scaled domain width) of solver k's
        regular mesh."""
        box = max(self.domain_size)
        level = compute_mesh_level(self.maximum_mesh_size(k), box)
        h = box / 3 ** (level - 1)
        # Runner::determineScaledDomainSize (E/runners/Runner.cpp:389-398):
        # std::ceil(domain / meshSize), no tolerance
        cells = tuple(int(math.ceil(w / h)) for w in self.domain_size)
        return cells, h, tuple(n * h for n in cells)

    def solver_config(self, k):
        """The kernel family and sizes of solver k, as the toolkit's factory
        selects them."""
        stype = self.solver_type(k)
        kernel = self.kernel(k)
        cells, h, width = self.mesh(k)
        cfg = dict(type=stype, identifier=self.identifier(k), nvar=self.variables(k), npar=self.parameters(k),
                   kernel=kernel, cells=cells, h=h, width=width, offset=self.offset, dimension=self.dimension,
                   end_time=self.end_time, time_stepping=self.time_stepping(k))
        if stype == "Finite-Volumes":
            schemes = {"generic::godunov": "godunov", "generic::musclhancock": "musclhancock"}
            if kernel not in schemes:
                raise SpecError("solver %d: finite-volumes kernel %r is not built" % (k, kernel))
            cfg.update(scheme=schemes[kernel], patch_size=self.patch_size(k))
        else:
            fam, _, kind = kernel.rpartition("::")
            if fam != "generic::fluxes" or kind.lower() not in ("nonlinear", "linear"):
                raise SpecError("solver %d: ADER-DG kernel %r is not built" % (k, kernel))
            cfg.update(order=self.order(k), linear=kind.lower() == "linear")
            if stype == "Limiting-ADER-DG":
                cfg.update(limiter_kernel=self.limiter_kernel(k),
                           dmp_relaxation_parameter=self.dmp_relaxation_parameter(k),
                           dmp_difference_scaling=self.dmp_difference_scaling(k),
                           steps_till_cured=self.steps_till_cured(k))
        return cfg


def create(spec, k=0, pde=None, velocity=(0.0, 0.0, 0.0), gamma=1.4, device=0, refine=None, patch_capacity=0):
    """The solver and resident grid of solver k of a Spec, with the user PDE
    given by `pde` (a PDE_* id of the compiled PDEs; the spec names none).
    Returns (solver, grid).  ADER-DG runs in 3D.  A Limiting-ADER-DG solver
    gives the adaptive limiting grid (exahype_amd.AMRGrid) on the unit cube:
    the regular mesh as level 0, at most maximum-mesh-depth finer levels
    (two at most), the spec's DMP parameters and steps-till-cured; the
    refinement criterion is user code in the reference, here `refine` =
    (centre, radius, half widths per refined level) of the grid's refinement
    shell (None: no refinement)."""
    import exahype_amd as E
    cfg = spec.solver_config(k)
    n = cfg["cells"]
    off, width = cfg["offset"], cfg["width"]
    if cfg["type"] == "Finite-Volumes":
        scheme = E.FV_GODUNOV if cfg["scheme"] == "godunov" else E.FV_MUSCLHANCOCK
        s = E.FiniteVolumesSolver(cfg["dimension"], cfg["patch_size"], scheme=scheme, gamma=gamma, pde=pde,
                                  device=device)
        if cfg["dimension"] == 2:
            g = E.FVGrid(s, n[0], n[1], 1, offset=off + (0.0,), width=width + (1.0,))
        else:
            g = E.FVGrid(s, n[0], n[1], n[2], offset=off, width=width)
        return s, g
    if cfg["dimension"] != 3:
        raise SpecError("ADER-DG kernels are built for 3 dimensions")
    if cfg["type"] == "Limiting-ADER-DG":
        if any(o != 0.0 for o in off) or any(w != 1.0 for w in width) or len(set(n)) != 1:
            raise SpecError("the adaptive limiting grid covers the unit cube with n^3 cells (offset 0, width 1)")
        if cfg["linear"]:
            raise SpecError("Limiting-ADER-DG: the limiting grid runs the nonlinear kernels")
        depth = min(spec.maximum_mesh_depth(k), 2)
        centre, radius, widths = refine if refine is not None else ((0.5, 0.5, 0.5), 0.0, ())
        levels = 1 + (min(depth, len(widths)) if refine is not None else 0)
        s = E.ADERDGSolver(cfg["order"], nvar=cfg["nvar"], npar=cfg["npar"], gamma=gamma, pde=pde,
                           velocity=velocity, device=device)
        acfg = E.amr_config(n[0], levels, centre, radius, list(widths)[:levels - 1],
                            relax=cfg["dmp_relaxation_parameter"], scaling=cfg["dmp_difference_scaling"],
                            steps_till_cured=cfg["steps_till_cured"], patch_capacity=patch_capacity)
        return s, E.AMRGrid(s, acfg)
    s = E.ADERDGSolver(cfg["order"], nvar=cfg["nvar"], npar=cfg["npar"], gamma=gamma, pde=pde, velocity=velocity,
                       linear=cfg["linear"], device=device)
    g = E.Grid(s, n[0], n[1], n[2], offset=off, width=width)
    if spec.fuse_algorithmic_steps():  # Runner.cpp:506-509
        g.set_fused_time_stepping(True, spec.fuse_algorithmic_steps_factor())
    return s, g
