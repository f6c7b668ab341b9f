"""ctypes binding of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Cfg(C.Structure):
    _fields_ = [("order", C.c_int), ("nvar", C.c_int), ("npar", C.c_int),
                ("gamma", C.c_double), ("cfl", C.c_double), ("pde", C.c_int),
                ("adv", C.c_double * 3), ("grav", C.c_double), ("linear", C.c_int),
                ("point_source", C.c_int), ("src_x0", C.c_double * 3), ("src_amp", C.c_double * 5),
                ("src_f0", C.c_double), ("src_t0", C.c_double)]


class IC(C.Structure):
    _fields_ = [("kind", C.c_int), ("xc", C.c_double * 3), ("L", C.c_double), ("eps", C.c_double),
                ("uinf", C.c_double * 3), ("rho_inf", C.c_double), ("p_inf", C.c_double),
                ("amp", C.c_double), ("nmodes", C.c_int), ("k", (C.c_double * 3) * 8),
                ("a", C.c_double * 8), ("phi", C.c_double * 8), ("r0", C.c_double),
                ("rho_in", C.c_double), ("p_in", C.c_double), ("rho_out", C.c_double),
                ("p_out", C.c_double), ("q0", C.c_double * 8)]


class BC(C.Structure):
    _fields_ = [("kind", C.c_int), ("ic", IC)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("offset", C.c_double * 3), ("width", C.c_double * 3),
                ("luh", C.POINTER(C.c_double)), ("lduh", C.POINTER(C.c_double)),
                ("lQhbnd", C.POINTER(C.c_double)), ("lFhbnd", C.POINTER(C.c_double)),
                ("iters", C.POINTER(C.c_int)), ("t", C.c_double), ("periodic", C.c_int * 3), ("bc", BC)]


class FVGridS(C.Structure):
    _fields_ = [("dim", C.c_int), ("scheme", C.c_int), ("patch", C.c_int), ("ghost", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("offset", C.c_double * 3), ("width", C.c_double * 3),
                ("luh", C.POINTER(C.c_double)), ("luh_new", C.POINTER(C.c_double)),
                ("periodic", C.c_int * 3), ("bc", BC), ("t", C.c_double)]


_lib = None
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_flux.argtypes = [P(Cfg), _D, _D]
        L.orc_eigenvalues.argtypes = [P(Cfg), _D, C.c_int, _D]
        L.orc_source.argtypes = [P(Cfg), _D, _D]
        L.orc_ncp.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_ncp_linear.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_sample_cartesian_vertices.argtypes = [P(Cfg), _D, _D]
        L.orc_coefficient_matrix.argtypes = [P(Cfg), _D, C.c_int, _D]
        L.orc_space_time_predictor_linear.argtypes = [P(Cfg), _D, _D, _D, _D, _D, _D, _D, _D, C.c_double, _D]
        L.orc_lagrange.argtypes = [C.c_int, C.c_int, C.c_double]
        L.orc_lagrange.restype = C.c_double
        L.orc_user_point_source.argtypes = [P(Cfg), _D, C.c_double, C.c_double, _D, _D]
        L.orc_point_source.argtypes = [P(Cfg), C.c_double, C.c_double, _D, _D, _D]
        L.orc_volume_integral_linear.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_surface_integral_linear.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_riemann_solver_linear.argtypes = [P(Cfg), _D, _D, _D, _D, C.c_double, C.c_int]
        L.orc_fused_source.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_space_time_predictor.argtypes = [P(Cfg), _D, _D, _D, _D, _D, _D, _D, _D, C.c_double]
        L.orc_space_time_predictor.restype = C.c_int
        L.orc_volume_integral.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_surface_integral.argtypes = [P(Cfg), _D, _D, _D]
        L.orc_riemann_solver.argtypes = [P(Cfg), _D, _D, _D, _D, C.c_double, C.c_int]
        L.orc_solution_update.argtypes = [P(Cfg), _D, _D, C.c_double]
        L.orc_stable_time_step_size.argtypes = [P(Cfg), _D, _D]
        L.orc_stable_time_step_size.restype = C.c_double
        L.orc_ic_point.argtypes = [P(Cfg), P(IC), _D, C.c_double, _D]
        L.orc_solution_adjustment.argtypes = [P(Cfg), P(IC), _D, _D, _D, C.c_double]
        L.orc_grid_alloc.argtypes = [P(Cfg), P(Grid), C.c_int, C.c_int, C.c_int]
        L.orc_grid_alloc.restype = C.c_int
        L.orc_grid_free.argtypes = [P(Grid)]
        L.orc_grid_init.argtypes = [P(Cfg), P(Grid), P(IC)]
        L.orc_grid_stable_dt.argtypes = [P(Cfg), P(Grid), C.c_int]
        L.orc_grid_stable_dt.restype = C.c_double
        L.orc_grid_step.argtypes = [P(Cfg), P(Grid), C.c_double, C.c_int]
        L.orc_grid_step.restype = C.c_double
        L.orc_grid_predict_cells.argtypes = [P(Cfg), P(Grid), C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_fv_rusanov.argtypes = [P(Cfg), C.c_int, _D, _D, _D, _D, C.c_int]
        L.orc_fv_rusanov.restype = C.c_double
        L.orc_fv_stable_dt.argtypes = [P(Cfg), C.c_int, C.c_int, C.c_int, _D, _D, C.c_double]
        L.orc_fv_stable_dt.restype = C.c_double
        L.orc_godunov_update.argtypes = [P(Cfg), C.c_int, C.c_int, _D, _D, _D, C.c_double, C.c_double]
        L.orc_godunov_update.restype = C.c_double
        L.orc_musclhancock_update.argtypes = [P(Cfg), C.c_int, _D, _D, _D, C.c_double, C.c_double]
        L.orc_fv_patch_ic.argtypes = [P(Cfg), P(IC), C.c_int, C.c_int, _D, _D, _D]
        L.orc_boundary_conditions.argtypes = [P(Cfg), P(BC), _D, _D, _D, _D, _D, _D, C.c_double, C.c_double, C.c_int,
                                              C.c_int]
        L.orc_musclhancock_update.restype = C.c_double
        for fn in ("orc_fv_boundary_layer_extraction", "orc_fv_ghost_layer_filling_at_boundary",
                   "orc_fv_ghost_layer_filling"):
            getattr(L, fn).argtypes = [P(Cfg), C.c_int, C.c_int, C.c_int, _D, _D, P(C.c_int)]
        L.orc_fv_boundary_conditions.argtypes = [P(Cfg), P(BC), C.c_int, C.c_int, C.c_int, _D, _D, _D, _D,
                                                 C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_fv_grid_alloc.argtypes = [P(Cfg), P(FVGridS), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_fv_grid_alloc.restype = C.c_int
        L.orc_fv_grid_free.argtypes = [P(FVGridS)]
        L.orc_fv_grid_init.argtypes = [P(Cfg), P(FVGridS), P(IC)]
        L.orc_fv_grid_stable_dt.argtypes = [P(Cfg), P(FVGridS), C.c_int]
        L.orc_fv_grid_stable_dt.restype = C.c_double
        L.orc_fv_grid_step.argtypes = [P(Cfg), P(FVGridS), C.c_double, _D, C.c_int]
        L.orc_fv_grid_step.restype = C.c_double
        L.orc_limiter_matrices.argtypes = [C.c_int, _D, _D, _D]
        L.orc_limiter_matrices.restype = C.c_int
        L.orc_project_on_fv.argtypes = [C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_project_on_dg.argtypes = [C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_cell_min_max.argtypes = [C.c_int, C.c_int, C.c_int, _D, _D, _D]
        L.orc_dmp.argtypes = [C.c_int, C.c_int, C.c_int, _D, C.c_double, C.c_double, _D, _D]
        L.orc_dmp.restype = C.c_int
        _I = C.POINTER(C.c_int)
        for nm in ("orc_face_unknowns_prolongation", "orc_face_unknowns_restriction"):
            getattr(L, nm).argtypes = [C.c_int, C.c_int, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int, _I]
        for nm in ("orc_volume_unknowns_prolongation", "orc_volume_unknowns_restriction"):
            getattr(L, nm).argtypes = [C.c_int, C.c_int, _D, _D, C.c_int, C.c_int, _I]
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def cfg(order, nvar=5, npar=0, gamma=1.4, cfl=0.9, pde=0, adv=(0.0, 0.0, 0.0), grav=0.0, linear=0,
        point_source=None):
    """pde 0 Euler, 1 linear advection (velocity adv), 2 EulerFV 2D, 3 Euler
    with gravity grav along -z (algebraic source), 4 linear advection in
    non-conservative form (velocity adv, no flux), 5 Euler with the test
    non-conservative product of the reference's Euler kernel test.
    point_source: None, or dict(x0=(3,), amp=(5,), f0=, t0=) -- the Ricker
    force of the linear predictor's usePointSource arm."""
    c = Cfg(order, nvar, npar, gamma, cfl, pde)
    for i in range(3):
        c.adv[i] = adv[i]
    c.grav = grav
    c.linear = int(linear)
    if point_source is not None:
        c.point_source = 1
        for i in range(3):
            c.src_x0[i] = point_source["x0"][i]
        for i in range(5):
            c.src_amp[i] = point_source["amp"][i]
        c.src_f0, c.src_t0 = point_source["f0"], point_source["t0"]
    return c


def lagrange(order, i, xi):
    """Lagrange basis polynomial of GL node i of the order's nodes at xi."""
    return lib().orc_lagrange(order, i, xi)


def user_point_source(c, x, t, dt):
    """(force[nvar], x0[3]) of the compiled user pointSource."""
    f, x0 = np.zeros(c.nvar), np.zeros(3)
    lib().orc_user_point_source(C.byref(c), ptr(dvec(x)), t, dt, ptr(f), ptr(x0))
    return f, x0


def point_source(c, t, dt, centre, dx):
    """pointSource.cpph: PSi [B+1][z][y][x][nVar]."""
    B = c.order + 1
    psi = np.zeros((B + 1) * B ** 3 * c.nvar)
    lib().orc_point_source(C.byref(c), t, dt, ptr(dvec(centre)), ptr(dvec(dx)), ptr(psi))
    return psi


def source(c, Q):
    S = np.zeros(c.nvar)
    lib().orc_source(C.byref(c), ptr(dvec(Q)), ptr(S))
    return S


def ncp(c, Q, gradQ):
    out = np.zeros(c.nvar)
    lib().orc_ncp(C.byref(c), ptr(dvec(Q)), ptr(dvec(gradQ)), ptr(out))
    return out


def dvec(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def flux(c, Q):
    F = np.zeros(3 * c.nvar)
    lib().orc_flux(C.byref(c), ptr(dvec(Q)), ptr(F))
    return F.reshape(3, c.nvar)


def eigenvalues(c, Q, d):
    lam = np.zeros(16)
    lib().orc_eigenvalues(C.byref(c), ptr(dvec(Q)), d, ptr(lam))
    return lam[:c.nvar]


def space_time_predictor(c, luh, dx, dt):
    B = c.order + 1
    nd, nv = c.nvar + c.npar, c.nvar
    luh = dvec(luh)
    out = dict(lQhbnd=np.zeros(6 * B * B * nd), lFhbnd=np.zeros(6 * B * B * nv),
               lQi=np.zeros(B ** 4 * nd), lFi=np.zeros(B ** 4 * 4 * nv),
               lQhi=np.zeros(B ** 3 * nd), lFhi=np.zeros(4 * B ** 3 * nv))
    it = lib().orc_space_time_predictor(C.byref(c), ptr(out["lQhbnd"]), ptr(out["lFhbnd"]), ptr(out["lQi"]),
                                        ptr(out["lFi"]), ptr(out["lQhi"]), ptr(out["lFhi"]), ptr(luh),
                                        ptr(dvec(dx)), dt)
    out["iters"] = it
    return out


def space_time_predictor_linear(c, luh, dx, dt, psi=None):
    """Cauchy-Kowalevski predictor: lQi [B+1][z][y][x][nData] (time
    derivatives), lFi [B][z][y][x][3][nVar], lQhi, lFhi [4][z][y][x][nVar]
    (block 3 zero), lQhbnd, lFhbnd.  psi: point forces (point_source) or
    None."""
    B = c.order + 1
    nd, nv = c.nvar + c.npar, c.nvar
    out = dict(lQhbnd=np.zeros(6 * B * B * nd), lFhbnd=np.zeros(6 * B * B * nv),
               lQi=np.zeros((B + 1) * B ** 3 * nd), lFi=np.zeros(B ** 4 * 3 * nv),
               lQhi=np.zeros(B ** 3 * nd), lFhi=np.zeros(4 * B ** 3 * nv))
    lib().orc_space_time_predictor_linear(C.byref(c), ptr(out["lQhbnd"]), ptr(out["lFhbnd"]), ptr(out["lQi"]),
                                          ptr(out["lFi"]), ptr(out["lQhi"]), ptr(out["lFhi"]), ptr(dvec(luh)),
                                          ptr(dvec(dx)), dt, None if psi is None else ptr(dvec(psi)))
    return out


def volume_integral_linear(c, lFhi, dx):
    lduh = np.zeros((c.order + 1) ** 3 * c.nvar)
    lib().orc_volume_integral_linear(C.byref(c), ptr(lduh), ptr(dvec(lFhi)), ptr(dvec(dx)))
    return lduh


def surface_integral_linear(c, lduh, lFhbnd, dx):
    lduh = dvec(lduh).copy()
    lib().orc_surface_integral_linear(C.byref(c), ptr(lduh), ptr(dvec(lFhbnd)), ptr(dvec(dx)))
    return lduh


def riemann_solver_linear(c, QL, QR, dt, direction):
    B = c.order + 1
    FL, FR = np.zeros(B * B * c.nvar), np.zeros(B * B * c.nvar)
    lib().orc_riemann_solver_linear(C.byref(c), ptr(FL), ptr(FR), ptr(dvec(QL)), ptr(dvec(QR)), dt, direction)
    return FL, FR


def sample_cartesian_vertices(c, luh):
    out = np.zeros((c.order + 1) ** 3 * c.nvar)
    lib().orc_sample_cartesian_vertices(C.byref(c), ptr(dvec(luh)), ptr(out))
    return out


def coefficient_matrix(c, Q, d):
    Bn = np.zeros(c.nvar * c.nvar)
    lib().orc_coefficient_matrix(C.byref(c), ptr(dvec(Q)), d, ptr(Bn))
    return Bn.reshape(c.nvar, c.nvar)


def volume_integral(c, lFhi, dx):
    B = c.order + 1
    lduh = np.zeros(B ** 3 * c.nvar)
    lib().orc_volume_integral(C.byref(c), ptr(lduh), ptr(dvec(lFhi)), ptr(dvec(dx)))
    return lduh


def surface_integral(c, lduh, lFhbnd, dx):
    lduh = dvec(lduh).copy()
    lib().orc_surface_integral(C.byref(c), ptr(lduh), ptr(dvec(lFhbnd)), ptr(dvec(dx)))
    return lduh


def riemann_solver(c, FL, FR, QL, QR, dt, direction):
    FL, FR = dvec(FL).copy(), dvec(FR).copy()
    lib().orc_riemann_solver(C.byref(c), ptr(FL), ptr(FR), ptr(dvec(QL)), ptr(dvec(QR)), dt, direction)
    return FL, FR


def solution_update(c, luh, lduh, dt):
    luh = dvec(luh).copy()
    lib().orc_solution_update(C.byref(c), ptr(luh), ptr(dvec(lduh)), dt)
    return luh


def stable_time_step_size(c, luh, dx):
    return lib().orc_stable_time_step_size(C.byref(c), ptr(dvec(luh)), ptr(dvec(dx)))


BC_OUTFLOW, BC_REFLECTING, BC_EXACT = 1, 2, 3


def make_bc(kind, ic=None):
    b = BC()
    b.kind = kind
    if ic is not None:
        b.ic = ic
    return b


def boundary_conditions(c, bc, FIn, QIn, centre, size, t, dt, face, normal):
    """(FOut, QOut) of one face: the user boundaryValues at its GL nodes."""
    B = c.order + 1
    FOut, QOut = np.zeros(B * B * c.nvar), np.zeros(B * B * (c.nvar + c.npar))
    lib().orc_boundary_conditions(C.byref(c), C.byref(bc), ptr(FOut), ptr(QOut), ptr(dvec(FIn)), ptr(dvec(QIn)),
                                  ptr(dvec(centre)), ptr(dvec(size)), t, dt, face, normal)
    return FOut, QOut


def make_ic(kind=0, **kw):
    ic = IC()
    ic.kind = kind
    for k, v in kw.items():
        cur = getattr(ic, k)
        if isinstance(cur, (int, float)):
            setattr(ic, k, v)
        elif k == "k":
            for m, row in enumerate(v):
                for j in range(3):
                    ic.k[m][j] = row[j]
        else:
            for i, x in enumerate(v):
                cur[i] = x
    return ic


def solution_adjustment(c, ic, centre, dx, t=0.0):
    B = c.order + 1
    luh = np.zeros(B ** 3 * (c.nvar + c.npar))
    lib().orc_solution_adjustment(C.byref(c), C.byref(ic), ptr(luh), ptr(dvec(centre)), ptr(dvec(dx)), t)
    return luh


class OracleGrid:
    """Periodic regular grid of ADER-DG cells driven by the oracle."""

    def __init__(self, c, nx, ny, nz):
        self.c = c
        self.g = Grid()
        if lib().orc_grid_alloc(C.byref(c), C.byref(self.g), nx, ny, nz) != 0:
            raise MemoryError("oracle grid allocation failed")
        self.ncells = nx * ny * nz
        B = c.order + 1
        self.cell_size = B ** 3 * (c.nvar + c.npar)

    def __del__(self):
        try:
            lib().orc_grid_free(C.byref(self.g))
        except Exception:
            pass

    def init(self, ic):
        lib().orc_grid_init(C.byref(self.c), C.byref(self.g), C.byref(ic))

    def set_boundary(self, periodic, bc):
        """periodic[d] = 0: the domain faces of axis d take the boundary conditions bc."""
        for d in range(3):
            self.g.periodic[d] = int(periodic[d])
        self.g.bc = bc

    def luh(self):
        n = self.ncells * self.cell_size
        return np.ctypeslib.as_array(self.g.luh, shape=(n,))

    def iters(self):
        return np.ctypeslib.as_array(self.g.iters, shape=(self.ncells,))

    def set_luh(self, a):
        self.luh()[:] = np.asarray(a, dtype=np.float64).ravel()

    def stable_dt(self, nthreads=0):
        return lib().orc_grid_stable_dt(C.byref(self.c), C.byref(self.g), nthreads or os.cpu_count())

    def step(self, dt, nthreads=0):
        return lib().orc_grid_step(C.byref(self.c), C.byref(self.g), dt, nthreads or os.cpu_count())

    def predict_cells(self, dt, first, count, nthreads=0):
        lib().orc_grid_predict_cells(C.byref(self.c), C.byref(self.g), dt, first, count,
                                     nthreads or os.cpu_count())


class FusedTimeStepping:
    """Time step data of the reference's fused algorithmic steps (global time
    stepping), restated: dt_c is the dt of the next step (its predictor and
    corrector), dt_p the optimistic predictor dt, reruns the steps whose
    optimistic dt was unstable.

    start(stable): ADERDGSolver::initSolver zeroes the predictor dt
    (E/solvers/ADERDGSolver.cpp:754-756); the first time-step-size
    computation then takes used = stable (TimeStepSizeComputation.cpp:166-168),
    is not unstable and sets next = min(stable, 0.5 (used + w stable))
    (:182-183, updateMinNextPredictorTimeStepSize :657-670), which
    startNewTimeStep makes the predictor dt (ADERDGSolver.cpp:556-570) and
    the fused initialisation the corrector dt
    (FusedTimeSteppingInitialisation.cpp:96-99).

    after_step(stable): the end of a fused iteration, stable the minimum over
    the updated cells (TimeStepSizeComputation.cpp:158-187 then
    startNewTimeStep): used > stable sets the predictor dt, and with it the
    next corrector dt, to w stable and the next optimistic dt to
    min(stable, w stable), and the predictor is recomputed with it
    (Runner.cpp:1114-1120); otherwise the corrector keeps used and the next
    optimistic dt is min(stable, 0.5 (used + w stable))."""

    def __init__(self, weight):
        self.w = weight
        self.dt_c = 0.0
        self.dt_p = 0.0
        self.reruns = 0

    def start(self, stable):
        used = stable
        nxt = min(stable, 0.5 * (used + self.w * stable))
        self.dt_p = nxt
        self.dt_c = nxt

    def after_step(self, stable):
        used = self.dt_p if self.dt_p != 0.0 else stable
        if used > stable:
            self.dt_c = self.w * stable
            self.dt_p = min(stable, self.w * stable)
            self.reruns += 1
        else:
            self.dt_c = used
            self.dt_p = min(stable, 0.5 * (used + self.w * stable))
        return self.dt_c


def fv_rusanov(c, dim, qL, qR, normal):
    fL, fR = np.zeros(16), np.zeros(16)
    s = lib().orc_fv_rusanov(C.byref(c), dim, ptr(fL), ptr(fR), ptr(dvec(qL)), ptr(dvec(qR)), normal)
    return fL[:c.nvar], fR[:c.nvar], s


def fv_stable_dt(c, dim, patch, ghost, luh, dx, cfl):
    return lib().orc_fv_stable_dt(C.byref(c), dim, patch, ghost, ptr(dvec(luh)), ptr(dvec(dx)), cfl)


def godunov_update(c, dim, patch, luh_new, luh, dx, dt, cfl):
    luh_new = dvec(luh_new).copy()
    a = lib().orc_godunov_update(C.byref(c), dim, patch, ptr(luh_new), ptr(dvec(luh)), ptr(dvec(dx)), dt, cfl)
    return luh_new, a


def fv_patch_ic(c, ic, patch, ghost, centre, dx):
    """Initial conditions of one 3D FV patch at its subcell centres (ghosts zero)."""
    S = patch + 2 * ghost
    out = np.zeros(S ** 3 * (c.nvar + c.npar))
    lib().orc_fv_patch_ic(C.byref(c), C.byref(ic), patch, ghost, ptr(dvec(centre)), ptr(dvec(dx)), ptr(out))
    return out


def musclhancock_update(c, patch, luh_new, luh, dx, dt, cfl=0.99):
    luh_new = dvec(luh_new).copy()
    a = lib().orc_musclhancock_update(C.byref(c), patch, ptr(luh_new), ptr(dvec(luh)), ptr(dvec(dx)), dt, cfl)
    return luh_new, a


def _pos(p):
    return (C.c_int * 3)(*[int(x) for x in p])


def fv_boundary_layer_extraction(c, dim, P, G, luh, pos):
    """boundaryLayerExtraction: the G layers of luh next to face pos, as a face block."""
    out = np.zeros(P ** (dim - 1) * G * (c.nvar + c.npar))
    lib().orc_fv_boundary_layer_extraction(C.byref(c), dim, P, G, ptr(out), ptr(dvec(luh)), _pos(pos))
    return out


def fv_ghost_layer_filling_at_boundary(c, dim, P, G, luh, luhbnd, pos):
    out = dvec(luh).copy()
    lib().orc_fv_ghost_layer_filling_at_boundary(C.byref(c), dim, P, G, ptr(out), ptr(dvec(luhbnd)), _pos(pos))
    return out


def fv_ghost_layer_filling(c, dim, P, G, luh, luh_nb, pos):
    out = dvec(luh).copy()
    lib().orc_fv_ghost_layer_filling(C.byref(c), dim, P, G, ptr(out), ptr(dvec(luh_nb)), _pos(pos))
    return out


def fv_boundary_conditions(c, bc, dim, P, G, state_in, centre, size, t, dt, face, normal):
    out = np.zeros(P ** (dim - 1) * G * (c.nvar + c.npar))
    lib().orc_fv_boundary_conditions(C.byref(c), C.byref(bc), dim, P, G, ptr(out), ptr(dvec(state_in)),
                                     ptr(dvec(centre)), ptr(dvec(size)), t, dt, face, normal)
    return out


class OracleFVGrid:
    """Periodic grid of FV patches driven by the oracle (scheme 0 Godunov, 1 MUSCL-Hancock)."""

    def __init__(self, c, dim, scheme, patch, nx, ny, nz=1):
        self.c = c
        self.g = FVGridS()
        if lib().orc_fv_grid_alloc(C.byref(c), C.byref(self.g), dim, scheme, patch, nx, ny, nz) != 0:
            raise MemoryError("oracle FV grid allocation failed")
        S = patch + 2 * self.g.ghost
        self.patch_size = S ** dim * (c.nvar + c.npar)
        self.npatch = nx * ny * (nz if dim == 3 else 1)

    def __del__(self):
        try:
            lib().orc_fv_grid_free(C.byref(self.g))
        except Exception:
            pass

    def init(self, ic):
        lib().orc_fv_grid_init(C.byref(self.c), C.byref(self.g), C.byref(ic))

    def set_boundary(self, periodic, bc):
        """periodic[d] = 0: the domain faces of axis d take the boundary values bc."""
        for d in range(3):
            self.g.periodic[d] = int(periodic[d])
        self.g.bc = bc

    def luh(self):
        return np.ctypeslib.as_array(self.g.luh, shape=(self.npatch * self.patch_size,))

    def stable_dt(self, nthreads=0):
        return lib().orc_fv_grid_stable_dt(C.byref(self.c), C.byref(self.g), nthreads or os.cpu_count())

    def step(self, dt, nthreads=0):
        adm = C.c_double(0.0)
        nxt = lib().orc_fv_grid_step(C.byref(self.c), C.byref(self.g), dt, C.byref(adm), nthreads or os.cpu_count())
        return nxt, adm.value


def limiter_matrices(N):
    B, L = N + 1, 2 * N + 1
    a, b, c = np.zeros(B * L), np.zeros(B * B), np.zeros(L * B)
    assert lib().orc_limiter_matrices(N, ptr(a), ptr(b), ptr(c)) == 0
    return a.reshape(B, L), b.reshape(B, B), c.reshape(L, B)


def project_on_fv(N, luh, ghost=0, nvar=5):
    S = 2 * N + 1 + 2 * ghost
    lim = np.zeros(S ** 3 * nvar)
    lib().orc_project_on_fv(N, nvar, ghost, ptr(dvec(luh)), ptr(lim))
    return lim


def project_on_dg(N, lim, ghost=0, nvar=5):
    luh = np.zeros((N + 1) ** 3 * nvar)
    lib().orc_project_on_dg(N, nvar, ghost, ptr(dvec(lim)), ptr(luh))
    return luh


def cell_min_max(N, luh, nobs=5, nvar=5):
    mn, mx = np.zeros(nobs), np.zeros(nobs)
    lib().orc_cell_min_max(N, nvar, nobs, ptr(dvec(luh)), ptr(mn), ptr(mx))
    return mn, mx


def dmp(N, luh, bmin, bmax, relax, scaling, nobs=5, nvar=5):
    bmin, bmax = dvec(bmin).copy(), dvec(bmax).copy()
    ok = lib().orc_dmp(N, nvar, nobs, ptr(dvec(luh)), relax, scaling, ptr(bmin), ptr(bmax))
    return bool(ok), bmin, bmax


def _ivec(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def face_unknowns_prolongation(N, qc, fc, coarse_level, fine_level, subface, nvar=5, npar=0):
    """Fine face (Q, F) from the coarse face; amrRoutines.cpph:45-144."""
    B2 = (N + 1) ** 2
    qf, ff = np.zeros(B2 * (nvar + npar)), np.zeros(B2 * nvar)
    idx, ip = _ivec(subface)
    lib().orc_face_unknowns_prolongation(N, nvar, npar, ptr(qf), ptr(ff), ptr(dvec(qc)), ptr(dvec(fc)),
                                         coarse_level, fine_level, ip)
    return qf, ff


def face_unknowns_restriction(N, qc, fc, qf, ff, coarse_level, fine_level, subface, nvar=5, npar=0):
    """Adds the restriction of the fine face into copies of (qc, fc); amrRoutines.cpph:182-258."""
    qc, fc = dvec(qc).copy(), dvec(fc).copy()
    idx, ip = _ivec(subface)
    lib().orc_face_unknowns_restriction(N, nvar, npar, ptr(qc), ptr(fc), ptr(dvec(qf)), ptr(dvec(ff)),
                                        coarse_level, fine_level, ip)
    return qc, fc


def volume_unknowns_prolongation(N, luh_coarse, coarse_level, fine_level, subcell, ndata=5):
    """amrRoutines.cpph:289-374"""
    out = np.zeros((N + 1) ** 3 * ndata)
    idx, ip = _ivec(subcell)
    lib().orc_volume_unknowns_prolongation(N, ndata, ptr(out), ptr(dvec(luh_coarse)), coarse_level, fine_level, ip)
    return out


def volume_unknowns_restriction(N, luh_coarse, luh_fine, coarse_level, fine_level, subcell, ndata=5):
    """Adds into a copy of luh_coarse; amrRoutines.cpph:410-493"""
    out = dvec(luh_coarse).copy()
    idx, ip = _ivec(subcell)
    lib().orc_volume_unknowns_restriction(N, ndata, ptr(out), ptr(dvec(luh_fine)), coarse_level, fine_level, ip)
    return out
