"""exahype_amd: MI355X (gfx950) ADER-DG / finite-volume kernels.

Python host side over the C ABI in include/exahype_amd.h.  The classes mirror
the kernel surface of the reference's solver plugins:

* ``ADERDGSolver``: spaceTimePredictor, volumeIntegral, surfaceIntegral,
  riemannSolver, solutionUpdate, stableTimeStepSize, adjustSolution -- the
  virtuals of exahype::solvers::ADERDGSolver (E/solvers/ADERDGSolver.h:994-1135),
  same argument meaning and array layouts.
* ``Grid``: a device-resident periodic mesh advanced by whole time steps.

The HIP library is required: importing works everywhere, but creating a solver
raises ``RuntimeError`` when the library is missing or no gfx950 GPU is usable.
There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libexahype_amd.so")

EXA_OK = 0
PREC_FP64 = 0
PREC_FP32 = 1
PREC_FP32_MFMA = 2
PREC_BF16_MFMA = 3
CONTRACT_DEFAULT, CONTRACT_PENCIL, CONTRACT_NODE, CONTRACT_PENCIL_PAIR, CONTRACT_STAGED = 0, 1, 2, 3, 4
PDE_EULER = 0
PDE_EULER_GRAVITY = 2
PDE_ADVECTION = 3          # linear advection of 5 variables, flux form
PDE_ADVECTION_NCP = 4      # the same in non-conservative form (no flux)
PDE_EULER_NCP_TEST = 5     # Euler with the reference kernel test's non-conservative product
PDE_EULERFV2D = 1
FV_GODUNOV = 0
FV_MUSCLHANCOCK = 1

_D = C.POINTER(C.c_double)
_lib = None


class AderdgConfig(C.Structure):
    _fields_ = [("order", C.c_int), ("nvar", C.c_int), ("npar", C.c_int), ("cfl", C.c_double),
                ("gamma", C.c_double), ("pde", C.c_int), ("precision", C.c_int), ("fp64_sweeps", C.c_int),
                ("waves_per_simd", C.c_int), ("contraction", C.c_int), ("gravity", C.c_double),
                ("velocity", C.c_double * 3), ("linear", C.c_int), ("point_source", C.c_int),
                ("source_x0", C.c_double * 3), ("source_amp", C.c_double * 5), ("source_f0", C.c_double),
                ("source_t0", C.c_double)]


class EnsembleMember(C.Structure):
    _fields_ = [("order", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("offset", C.c_double * 3), ("width", C.c_double * 3), ("gamma", C.c_double)]


class FVConfig(C.Structure):
    _fields_ = [("dim", C.c_int), ("patch", C.c_int), ("scheme", C.c_int), ("nvar", C.c_int), ("npar", C.c_int),
                ("cfl", C.c_double), ("gamma", C.c_double), ("pde", C.c_int)]


class InitialCondition(C.Structure):
    """exa_ic: pointwise initial condition evaluated at the Gauss-Legendre nodes."""
    _fields_ = [("kind", C.c_int), ("xc", C.c_double * 3), ("L", C.c_double), ("eps", C.c_double),
                ("uinf", C.c_double * 3), ("rho_inf", C.c_double), ("p_inf", C.c_double),
                ("amp", C.c_double), ("nmodes", C.c_int), ("k", (C.c_double * 3) * 8),
                ("a", C.c_double * 8), ("phi", C.c_double * 8), ("r0", C.c_double),
                ("rho_in", C.c_double), ("p_in", C.c_double), ("rho_out", C.c_double),
                ("p_out", C.c_double), ("q0", C.c_double * 8)]


class BoundaryCondition(C.Structure):
    """exa_bc: the boundary values of a domain face (EXA_BC_*)."""
    _fields_ = [("kind", C.c_int), ("exact", InitialCondition)]


BC_OUTFLOW, BC_REFLECTING, BC_EXACT, BC_USER = 1, 2, 3, 4
PDE_USER_FIRST = 16


def pde_lookup(name):
    """The id of a PDE registered by a project (exa_pde_register)."""
    pid = C.c_int(0)
    _check(load_library().exa_pde_lookup(name.encode(), C.byref(pid)))
    return pid.value


def make_bc(kind, exact=None):
    """BoundaryCondition of a kind BC_*; exact: the InitialCondition whose
    state BC_EXACT imposes."""
    b = BoundaryCondition()
    b.kind = kind
    if exact is not None:
        b.exact = exact
    return b


# every symbol declared in include/exahype_amd.h
EXPORTS = [
    "exa_last_error", "exa_version", "exa_init", "exa_device_synchronize", "exa_aderdg_create", "exa_aderdg_destroy",
    "exa_aderdg_space_time_predictor", "exa_aderdg_space_time_predictor_ps", "exa_aderdg_point_source", "exa_aderdg_volume_integral", "exa_aderdg_surface_integral",
    "exa_aderdg_riemann_solver", "exa_aderdg_solution_update", "exa_aderdg_stable_time_step_size",
    "exa_aderdg_adjust_solution", "exa_aderdg_boundary_conditions", "exa_grid_set_boundary",
    "exa_pde_register", "exa_pde_lookup",
    "exa_aderdg_batch_prediction", "exa_aderdg_batch_riemann",
    "exa_aderdg_batch_update", "exa_aderdg_batch_pending", "exa_aderdg_batch_flush", "exa_grid_create", "exa_grid_destroy", "exa_grid_num_cells",
    "exa_grid_upload", "exa_grid_download", "exa_grid_set_initial_conditions", "exa_grid_sample_cartesian",
    "exa_grid_compute_dt",
    "exa_grid_set_dt", "exa_grid_set_fused_time_stepping", "exa_grid_fused_reruns", "exa_grid_step", "exa_grid_sync",
    "exa_grid_picard_iterations",
    "exa_grid_use_graph", "exa_grid_enable_timing", "exa_grid_kernel_times",
    "exa_grid_total_picard_iterations",
    "exa_fv_create", "exa_fv_destroy", "exa_fv_solution_update", "exa_fv_stable_time_step_size",
    "exa_fv_riemann_solver", "exa_fv_ghost_layer_filling", "exa_fv_ghost_layer_filling_at_boundary",
    "exa_fv_boundary_layer_extraction", "exa_fv_boundary_conditions", "exa_fv_grid_set_boundary",
    "exa_fv_grid_create", "exa_fv_grid_destroy", "exa_fv_grid_upload",
    "exa_fv_grid_download", "exa_fv_grid_set_initial_conditions", "exa_fv_grid_compute_dt", "exa_fv_grid_step",
    "exa_fv_grid_sync", "exa_fv_grid_enable_timing", "exa_fv_grid_kernel_times", "exa_limiter_project_on_fv", "exa_limiter_project_on_dg", "exa_limiter_dmp",
    "exa_amr_face_unknowns_prolongation", "exa_amr_face_unknowns_restriction",
    "exa_amr_volume_unknowns_prolongation", "exa_amr_volume_unknowns_restriction",
    "exa_amr_prolongation_batched", "exa_amr_restriction_batched", "exa_amr_last_kernel_ms",
    "exa_amr_restriction_batched_noalloc",
    "exa_amr_mesh_cells", "exa_amr_grid_create", "exa_amr_grid_destroy", "exa_amr_grid_num_cells", "exa_amr_grid_cells",
    "exa_amr_grid_upload", "exa_amr_grid_download", "exa_amr_grid_set_initial_conditions", "exa_amr_grid_compute_dt",
    "exa_amr_grid_use_graph", "exa_amr_grid_enable_timing", "exa_amr_grid_step", "exa_amr_grid_sync",
    "exa_amr_grid_limiter_status", "exa_amr_grid_patches", "exa_amr_grid_counters", "exa_amr_grid_phase_times",
    "exa_amr_grid_total_picard_iterations",
    "exa_device_malloc", "exa_device_free", "exa_memcpy_htod", "exa_memcpy_dtoh", "exa_device_copy",
    "exa_selftest_mfma",
    "exa_ensemble_create", "exa_ensemble_destroy", "exa_ensemble_set_initial_conditions", "exa_ensemble_compute_dt",
    "exa_ensemble_use_graph", "exa_ensemble_step", "exa_ensemble_sync", "exa_ensemble_download",
    "exa_ensemble_total_picard_iterations", "exa_autotune",
    "exa_grid_limiter_snapshot", "exa_grid_limiter_check", "exa_grid_limiter_project",
]


def load_library(path=LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("exahype_amd HIP library not built: %s (run python -m exahype_amd.build)" % path)
    L = C.CDLL(path)
    P, I, LL = C.POINTER, C.c_int, C.c_longlong
    L.exa_last_error.restype = C.c_char_p
    L.exa_version.restype = C.c_char_p
    L.exa_init.argtypes = [I]
    L.exa_device_synchronize.argtypes = []
    L.exa_aderdg_create.argtypes = [P(AderdgConfig), P(C.c_void_p)]
    L.exa_aderdg_destroy.argtypes = [C.c_void_p]
    L.exa_aderdg_space_time_predictor.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, _D, _D, _D, C.c_double, P(I)]
    L.exa_aderdg_space_time_predictor_ps.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, _D, _D, _D, C.c_double, _D,
                                                     P(I)]
    L.exa_aderdg_point_source.argtypes = [C.c_void_p, C.c_double, C.c_double, _D, _D, _D]
    L.exa_aderdg_volume_integral.argtypes = [C.c_void_p, _D, _D, _D]
    L.exa_aderdg_surface_integral.argtypes = [C.c_void_p, _D, _D, _D]
    L.exa_aderdg_riemann_solver.argtypes = [C.c_void_p, _D, _D, _D, _D, C.c_double, I]
    L.exa_aderdg_batch_prediction.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, C.c_double]
    L.exa_aderdg_batch_riemann.argtypes = [C.c_void_p, _D, _D, _D, _D, C.c_double, I]
    L.exa_aderdg_batch_update.argtypes = [C.c_void_p, _D, _D, _D, _D, C.c_double]
    L.exa_aderdg_batch_pending.argtypes = [C.c_void_p, P(LL)]
    L.exa_aderdg_batch_flush.argtypes = [C.c_void_p]
    L.exa_aderdg_solution_update.argtypes = [C.c_void_p, _D, _D, C.c_double]
    L.exa_aderdg_stable_time_step_size.argtypes = [C.c_void_p, _D, _D, _D]
    L.exa_aderdg_adjust_solution.argtypes = [C.c_void_p, _D, _D, _D, C.c_double, P(InitialCondition)]
    L.exa_aderdg_boundary_conditions.argtypes = [C.c_void_p, P(BoundaryCondition), _D, _D, _D, _D, _D, _D,
                                                 C.c_double, C.c_double, I, I]
    L.exa_grid_set_boundary.argtypes = [C.c_void_p, P(I), P(BoundaryCondition)]
    L.exa_pde_register.argtypes = [C.c_char_p, C.c_void_p, P(I)]
    L.exa_pde_lookup.argtypes = [C.c_char_p, P(I)]
    L.exa_grid_create.argtypes = [C.c_void_p, I, I, I, _D, _D, P(C.c_void_p)]
    L.exa_grid_destroy.argtypes = [C.c_void_p]
    L.exa_grid_num_cells.argtypes = [C.c_void_p, P(LL)]
    L.exa_grid_upload.argtypes = [C.c_void_p, _D]
    L.exa_grid_download.argtypes = [C.c_void_p, _D]
    L.exa_grid_sample_cartesian.argtypes = [C.c_void_p, _D]
    L.exa_grid_set_initial_conditions.argtypes = [C.c_void_p, P(InitialCondition)]
    L.exa_grid_compute_dt.argtypes = [C.c_void_p, _D]
    L.exa_grid_set_dt.argtypes = [C.c_void_p, C.c_double]
    L.exa_grid_set_fused_time_stepping.argtypes = [C.c_void_p, I, C.c_double]
    L.exa_grid_fused_reruns.argtypes = [C.c_void_p, P(LL)]
    L.exa_grid_step.argtypes = [C.c_void_p, I]
    L.exa_grid_sync.argtypes = [C.c_void_p, _D, _D]
    L.exa_grid_picard_iterations.argtypes = [C.c_void_p, P(I)]
    L.exa_grid_use_graph.argtypes = [C.c_void_p, I]
    L.exa_grid_enable_timing.argtypes = [C.c_void_p, I]
    L.exa_grid_kernel_times.argtypes = [C.c_void_p, _D, P(LL)]
    L.exa_grid_total_picard_iterations.argtypes = [C.c_void_p, P(LL)]
    L.exa_fv_create.argtypes = [P(FVConfig), P(C.c_void_p)]
    L.exa_fv_destroy.argtypes = [C.c_void_p]
    L.exa_fv_solution_update.argtypes = [C.c_void_p, _D, _D, _D, C.c_double, _D]
    L.exa_fv_stable_time_step_size.argtypes = [C.c_void_p, _D, _D, _D]
    L.exa_fv_riemann_solver.argtypes = [C.c_void_p, _D, _D, _D, _D, I, _D]
    for fn in ("exa_fv_ghost_layer_filling", "exa_fv_ghost_layer_filling_at_boundary",
               "exa_fv_boundary_layer_extraction"):
        getattr(L, fn).argtypes = [C.c_void_p, _D, _D, P(I)]
    L.exa_fv_boundary_conditions.argtypes = [C.c_void_p, P(BoundaryCondition), _D, _D, _D, _D, C.c_double,
                                             C.c_double, I, I]
    L.exa_fv_grid_set_boundary.argtypes = [C.c_void_p, P(I), P(BoundaryCondition)]
    L.exa_fv_grid_create.argtypes = [C.c_void_p, I, I, I, _D, _D, P(C.c_void_p)]
    L.exa_fv_grid_destroy.argtypes = [C.c_void_p]
    L.exa_fv_grid_upload.argtypes = [C.c_void_p, _D]
    L.exa_fv_grid_download.argtypes = [C.c_void_p, _D]
    L.exa_fv_grid_set_initial_conditions.argtypes = [C.c_void_p, P(InitialCondition)]
    L.exa_fv_grid_compute_dt.argtypes = [C.c_void_p, _D]
    L.exa_fv_grid_step.argtypes = [C.c_void_p, I]
    L.exa_fv_grid_sync.argtypes = [C.c_void_p, _D, _D, _D]
    L.exa_fv_grid_enable_timing.argtypes = [C.c_void_p, I]
    L.exa_fv_grid_kernel_times.argtypes = [C.c_void_p, _D, P(LL)]
    L.exa_amr_last_kernel_ms.argtypes = [_D]
    L.exa_limiter_project_on_fv.argtypes = [I, _D, I, _D]
    L.exa_limiter_project_on_dg.argtypes = [I, _D, I, _D]
    L.exa_limiter_dmp.argtypes = [I, _D, I, C.c_double, C.c_double, _D, _D, P(I)]
    L.exa_device_malloc.argtypes = [P(C.c_void_p), LL]
    L.exa_device_free.argtypes = [C.c_void_p]
    L.exa_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, LL]
    L.exa_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, LL]
    L.exa_device_copy.argtypes = [C.c_void_p, C.c_void_p, LL, P(C.c_float)]
    L.exa_selftest_mfma.argtypes = [I, C.c_void_p, C.c_void_p, C.c_void_p]
    L.exa_autotune.argtypes = [I, I, I, P(I), P(I), _D]
    L.exa_grid_limiter_snapshot.argtypes = [C.c_void_p, I]
    L.exa_grid_limiter_check.argtypes = [C.c_void_p, C.c_double, C.c_double, P(LL), P(I)]
    L.exa_grid_limiter_project.argtypes = [C.c_void_p, I, C.c_void_p, LL, P(I)]
    L.exa_ensemble_create.argtypes = [P(EnsembleMember), I, P(AderdgConfig), P(C.c_void_p)]
    L.exa_ensemble_destroy.argtypes = [C.c_void_p]
    L.exa_ensemble_set_initial_conditions.argtypes = [C.c_void_p, P(InitialCondition)]
    L.exa_ensemble_compute_dt.argtypes = [C.c_void_p]
    L.exa_ensemble_use_graph.argtypes = [C.c_void_p, I]
    L.exa_ensemble_step.argtypes = [C.c_void_p, I]
    L.exa_ensemble_sync.argtypes = [C.c_void_p, _D, _D]
    L.exa_ensemble_download.argtypes = [C.c_void_p, I, _D]
    L.exa_ensemble_total_picard_iterations.argtypes = [C.c_void_p, P(LL)]
    for nm in ("exa_amr_face_unknowns_prolongation", "exa_amr_face_unknowns_restriction"):
        getattr(L, nm).argtypes = [I, I, I, _D, _D, _D, _D, I, I, P(I)]
    for nm in ("exa_amr_volume_unknowns_prolongation", "exa_amr_volume_unknowns_restriction"):
        getattr(L, nm).argtypes = [I, I, I, _D, _D, I, I, P(I)]
    for nm in ("exa_amr_prolongation_batched", "exa_amr_restriction_batched", "exa_amr_restriction_batched_noalloc"):
        getattr(L, nm).argtypes = [I, I, I, C.c_void_p, LL, C.c_void_p, LL, P(AmrTransfer), I]
    L.exa_amr_mesh_cells.argtypes = [C.POINTER(AMRConfig), P(I), LL, P(LL)]
    L.exa_amr_grid_create.argtypes = [C.c_void_p, C.POINTER(AMRConfig), P(C.c_void_p)]
    L.exa_amr_grid_destroy.argtypes = [C.c_void_p]
    L.exa_amr_grid_num_cells.argtypes = [C.c_void_p, P(LL)]
    L.exa_amr_grid_cells.argtypes = [C.c_void_p, P(I)]
    L.exa_amr_grid_upload.argtypes = [C.c_void_p, _D]
    L.exa_amr_grid_download.argtypes = [C.c_void_p, _D]
    L.exa_amr_grid_set_initial_conditions.argtypes = [C.c_void_p, P(InitialCondition)]
    L.exa_amr_grid_compute_dt.argtypes = [C.c_void_p, _D]
    L.exa_amr_grid_use_graph.argtypes = [C.c_void_p, I]
    L.exa_amr_grid_enable_timing.argtypes = [C.c_void_p, I]
    L.exa_amr_grid_step.argtypes = [C.c_void_p, I]
    L.exa_amr_grid_sync.argtypes = [C.c_void_p, _D, _D]
    L.exa_amr_grid_limiter_status.argtypes = [C.c_void_p, P(I)]
    L.exa_amr_grid_patches.argtypes = [C.c_void_p, _D, P(I), LL, P(LL)]
    L.exa_amr_grid_counters.argtypes = [C.c_void_p, P(LL)]
    L.exa_amr_grid_phase_times.argtypes = [C.c_void_p, _D, P(LL)]
    L.exa_amr_grid_total_picard_iterations.argtypes = [C.c_void_p, P(LL)]
    _lib = L
    return L


EXA_ERR_NUMERICS = -5


class NumericsError(RuntimeError):
    """The state holds non-finite values or a negative pressure: its stable
    time step is not a positive number (EXA_ERR_NUMERICS)."""


def _check(rc):
    if rc != EXA_OK:
        msg = "exahype_amd error %d: %s" % (rc, _lib.exa_last_error().decode())
        raise (NumericsError if rc == EXA_ERR_NUMERICS else RuntimeError)(msg)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "float64 C-contiguous arrays required"
    return a.ctypes.data_as(_D)


def _vec3(x):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (3,)))


def device_synchronize():
    """Wait for all queued GPU work of this process."""
    _check(load_library().exa_device_synchronize())


def make_ic(kind=0, **kw):
    """Build an InitialCondition: kind 0 vortex+modes, 1 Sod sphere, 2 constant, 3 sine modes
    (linear-advection PDEs), 4 entropy wave, 5 Gaussian density pulse."""
    ic = InitialCondition()
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


class ADERDGSolver:
    """Per-cell kernel surface of exahype::solvers::ADERDGSolver on the GPU."""

    def __init__(self, order, nvar=5, npar=0, cfl=0.9, gamma=1.4, precision=PREC_FP64, fp64_sweeps=1,
                 waves_per_simd=0, contraction=0, device=0, gravity=None, pde=None, velocity=(0.0, 0.0, 0.0),
                 linear=False, point_source=None):
        """gravity: None for the Euler equations, g for Euler with constant
        gravity along -z (EXA_PDE_EULER_GRAVITY, an algebraic source).
        pde: one of the PDE_* ids to select another compiled PDE (the
        advection PDEs take their velocity from `velocity`).
        linear: the linear kernel family (Cauchy-Kowalevski predictor,
        coefficient-matrix Riemann solver), for PDE_ADVECTION_NCP and
        PDE_EULER_NCP_TEST.
        point_source (linear solvers): None, or dict(x0=(3,), amp=(5,), f0=,
        t0=) -- a Ricker-wavelet point force at x0 (usePointSource)."""
        L = load_library()
        _check(L.exa_init(device))
        if pde is None:
            pde = PDE_EULER if gravity is None else PDE_EULER_GRAVITY
        self.cfg = AderdgConfig(order, nvar, npar, cfl, gamma, pde, precision, fp64_sweeps, waves_per_simd,
                                contraction, 0.0 if gravity is None else float(gravity),
                                (C.c_double * 3)(*[float(a) for a in velocity]), 1 if linear else 0)
        if point_source is not None:
            self.cfg.point_source = 1
            for d in range(3):
                self.cfg.source_x0[d] = float(point_source["x0"][d])
            for v in range(5):
                self.cfg.source_amp[v] = float(point_source["amp"][v])
            self.cfg.source_f0, self.cfg.source_t0 = float(point_source["f0"]), float(point_source["t0"])
        h = C.c_void_p()
        _check(L.exa_aderdg_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.order, self.nvar, self.npar = order, nvar, npar
        self.linear = bool(linear)
        self.B = order + 1
        self._pending = []
        self.ndata = nvar + npar

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_aderdg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- virtual kernel entry points (reference names) ---
    def pointSource(self, t, dt, centre, dx):
        """ADERDGSolver::pointSource: the cell's point forces [B+1][z][y][x][nVar]."""
        psi = np.zeros((self.B + 1) * self.B ** 3 * self.nvar)
        _check(_lib.exa_aderdg_point_source(self._h, float(t), float(dt), _p(_vec3(centre)), _p(_vec3(dx)),
                                            _p(psi)))
        return psi

    def spaceTimePredictor(self, luh, dx, dt, pointForceSources=None):
        B, nd, nv = self.B, self.ndata, self.nvar
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        # a linear solver returns the B+1 time derivatives and the 3 products per time level
        out = dict(lQhbnd=np.zeros(6 * B * B * nd), lFhbnd=np.zeros(6 * B * B * nv),
                   lQi=np.zeros((B + 1) * B ** 3 * nd if self.linear else B ** 4 * nd),
                   lFi=np.zeros(B ** 4 * (3 if self.linear else 4) * nv), lQhi=np.zeros(B ** 3 * nd),
                   lFhi=np.zeros(4 * B ** 3 * nv))
        it = C.c_int(0)
        psi = None if pointForceSources is None else np.ascontiguousarray(pointForceSources, dtype=np.float64)
        _check(_lib.exa_aderdg_space_time_predictor_ps(self._h, _p(out["lQhbnd"]), _p(out["lFhbnd"]),
                                                       _p(out["lQi"]), _p(out["lFi"]), _p(out["lQhi"]),
                                                       _p(out["lFhi"]), _p(luh), _p(_vec3(dx)), float(dt),
                                                       _p(psi), C.byref(it)))
        out["iters"] = it.value
        return out

    def volumeIntegral(self, lFhi, dx):
        lduh = np.zeros(self.B ** 3 * self.nvar)
        _check(_lib.exa_aderdg_volume_integral(self._h, _p(lduh), _p(np.ascontiguousarray(lFhi, np.float64)),
                                               _p(_vec3(dx))))
        return lduh

    def surfaceIntegral(self, lduh, lFhbnd, dx):
        lduh = np.array(lduh, dtype=np.float64, copy=True)
        _check(_lib.exa_aderdg_surface_integral(self._h, _p(lduh), _p(np.ascontiguousarray(lFhbnd, np.float64)),
                                                _p(_vec3(dx))))
        return lduh

    def riemannSolver(self, FL, FR, QL, QR, dt, normalNonZero):
        FL = np.array(FL, dtype=np.float64, copy=True)
        FR = np.array(FR, dtype=np.float64, copy=True)
        _check(_lib.exa_aderdg_riemann_solver(self._h, _p(FL), _p(FR), _p(np.ascontiguousarray(QL, np.float64)),
                                              _p(np.ascontiguousarray(QR, np.float64)), float(dt),
                                              int(normalNonZero)))
        return FL, FR

    # deferred forms: outputs are written into the given (float64, contiguous,
    # writable) arrays by batch_flush(), which the caller invokes at the end of
    # a sweep; the arrays are kept alive here until then
    def batch_prediction(self, lQhbnd, lFhbnd, lduh, luh, dx, dt):
        self._pending.extend((lQhbnd, lFhbnd, lduh))
        _check(_lib.exa_aderdg_batch_prediction(self._h, _p(lQhbnd), _p(lFhbnd), _p(lduh),
                                                _p(np.ascontiguousarray(luh, np.float64)), _p(_vec3(dx)), float(dt)))

    def batch_riemann(self, FL, FR, QL, QR, dt, normalNonZero):
        self._pending.extend((FL, FR))
        _check(_lib.exa_aderdg_batch_riemann(self._h, _p(FL), _p(FR), _p(np.ascontiguousarray(QL, np.float64)),
                                             _p(np.ascontiguousarray(QR, np.float64)), float(dt), int(normalNonZero)))

    def batch_update(self, luh, lduh, lFhbnd, dx, dt):
        self._pending.extend((luh, lduh))
        _check(_lib.exa_aderdg_batch_update(self._h, _p(luh), _p(lduh), _p(np.ascontiguousarray(lFhbnd, np.float64)),
                                            _p(_vec3(dx)), float(dt)))

    def batch_pending(self):
        c = (C.c_longlong * 3)()
        _check(_lib.exa_aderdg_batch_pending(self._h, c))
        return tuple(c)

    def batch_flush(self):
        try:
            _check(_lib.exa_aderdg_batch_flush(self._h))
        finally:
            self._pending = []

    def solutionUpdate(self, luh, lduh, dt):
        luh = np.array(luh, dtype=np.float64, copy=True)
        _check(_lib.exa_aderdg_solution_update(self._h, _p(luh), _p(np.ascontiguousarray(lduh, np.float64)),
                                               float(dt)))
        return luh

    def stableTimeStepSize(self, luh, dx):
        dt = C.c_double(0.0)
        _check(_lib.exa_aderdg_stable_time_step_size(self._h, _p(np.ascontiguousarray(luh, np.float64)),
                                                     _p(_vec3(dx)), C.byref(dt)))
        return dt.value

    def adjustSolution(self, centre, dx, t, ic):
        luh = np.zeros(self.B ** 3 * self.ndata)
        _check(_lib.exa_aderdg_adjust_solution(self._h, _p(luh), _p(_vec3(centre)), _p(_vec3(dx)), float(t),
                                               C.byref(ic)))
        return luh

    def boundaryConditions(self, bc, FIn, QIn, centre, size, t, dt, faceIndex, normalNonZero):
        """ADERDGSolver::boundaryConditions of one domain face: (FOut, QOut)
        in the reference layout [B^2][nVar] / [B^2][nData]."""
        B2 = self.B * self.B
        FOut, QOut = np.zeros(B2 * self.nvar), np.zeros(B2 * self.ndata)
        _check(_lib.exa_aderdg_boundary_conditions(self._h, C.byref(bc), _p(FOut), _p(QOut),
                                                   _p(np.ascontiguousarray(FIn, np.float64)),
                                                   _p(np.ascontiguousarray(QIn, np.float64)), _p(_vec3(centre)),
                                                   _p(_vec3(size)), float(t), float(dt), int(faceIndex),
                                                   int(normalNonZero)))
        return FOut, QOut


class Grid:
    """Regular mesh of ADER-DG cells resident in HBM (periodic unless
    set_boundary closes an axis)."""

    def __init__(self, solver, nx, ny, nz, offset=(0.0, 0.0, 0.0), width=(1.0, 1.0, 1.0)):
        self.solver = solver
        h = C.c_void_p()
        _check(_lib.exa_grid_create(solver._h, nx, ny, nz, _p(_vec3(offset)), _p(_vec3(width)), C.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.offset, self.width = tuple(offset), tuple(width)
        self.ncells = nx * ny * nz
        self.cell_size = solver.B ** 3 * solver.nvar

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64).ravel()
        assert luh.size == self.ncells * self.cell_size
        _check(_lib.exa_grid_upload(self._h, _p(luh)))

    def download(self):
        out = np.zeros(self.ncells * self.cell_size)
        _check(_lib.exa_grid_download(self._h, _p(out)))
        return out

    def sample_cartesian(self):
        """The solution at every cell's (N+1)^3 equidistant vertices (the
        values of a Cartesian-vertices plot): [cell][vertex, x fastest][nVar]."""
        out = np.zeros(self.ncells * self.solver.B ** 3 * self.solver.nvar)
        _check(_lib.exa_grid_sample_cartesian(self._h, _p(out)))
        return out.reshape(self.ncells, self.solver.B ** 3, self.solver.nvar)

    def set_initial_conditions(self, ic):
        """ic: an InitialCondition, or None for a registered PDE's adjustPointSolution."""
        _check(_lib.exa_grid_set_initial_conditions(self._h, None if ic is None else C.byref(ic)))

    def set_boundary(self, periodic, bc=None):
        """periodic: three flags; the faces of a non-periodic axis take bc
        (a BoundaryCondition)."""
        per = (C.c_int * 3)(*[1 if p else 0 for p in periodic])
        _check(_lib.exa_grid_set_boundary(self._h, per, None if bc is None else C.byref(bc)))

    def compute_dt(self):
        dt = C.c_double(0.0)
        _check(_lib.exa_grid_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def set_dt(self, dt):
        _check(_lib.exa_grid_set_dt(self._h, float(dt)))

    def set_fused_time_stepping(self, enable=True, weight=1.0):
        """The reference's fused stepping (`fuse-algorithmic-steps = on`,
        `fuse-algorithmic-steps-factor = weight`): optimistic time steps,
        weight * stable after an unstable one."""
        _check(_lib.exa_grid_set_fused_time_stepping(self._h, 1 if enable else 0, float(weight)))

    def fused_reruns(self):
        """Steps whose optimistic dt exceeded the stable one (the reference
        reran the predictor there)."""
        n = C.c_longlong(0)
        _check(_lib.exa_grid_fused_reruns(self._h, C.byref(n)))
        return n.value

    def step(self, nsteps=1):
        _check(_lib.exa_grid_step(self._h, int(nsteps)))

    # limiter (troubled-cell detection and batched FV projection)
    def limiter_snapshot(self, nobs=5):
        _check(_lib.exa_grid_limiter_snapshot(self._h, nobs))

    def limiter_check(self, relax=1e-4, scaling=1e-3):
        n = C.c_longlong(0)
        flags = np.zeros(self.ncells, np.int32)
        _check(_lib.exa_grid_limiter_check(self._h, float(relax), float(scaling), C.byref(n),
                                           flags.ctypes.data_as(C.POINTER(C.c_int))))
        return n.value, flags

    def limiter_project(self, ghost, ntroubled, out=None):
        """Troubled cells' subcell patches.  With out (a DeviceArray of at
        least ntroubled * V * S^3 doubles) the patches stay on the device and
        the cell list is returned; without, returns (cells, patches[n][V*S^3])."""
        S = 2 * self.solver.order + 1 + 2 * ghost
        per = 5 * S ** 3
        d = out if out is not None else DeviceArray(max(1, ntroubled) * per)
        cells = np.zeros(max(1, ntroubled), np.int32)
        _check(_lib.exa_grid_limiter_project(self._h, ghost, d.ptr, d.n, cells.ctypes.data_as(C.POINTER(C.c_int))))
        if out is not None:
            return cells[:ntroubled]
        res = d.numpy()[:ntroubled * per].reshape(ntroubled, per)
        d.close()
        return cells[:ntroubled], res

    def sync(self):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        _check(_lib.exa_grid_sync(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def picard_iterations(self):
        it = np.zeros(self.ncells, dtype=np.int32)
        _check(_lib.exa_grid_picard_iterations(self._h, it.ctypes.data_as(C.POINTER(C.c_int))))
        return it

    def total_picard_iterations(self):
        t = C.c_longlong(0)
        _check(_lib.exa_grid_total_picard_iterations(self._h, C.byref(t)))
        return t.value

    def use_graph(self, enable=True):
        _check(_lib.exa_grid_use_graph(self._h, 1 if enable else 0))

    def enable_timing(self, enable=True):
        _check(_lib.exa_grid_enable_timing(self._h, 1 if enable else 0))

    def kernel_times(self):
        t = np.zeros(3)
        c = np.zeros(3, dtype=np.int64)
        _check(_lib.exa_grid_kernel_times(self._h, _p(t), c.ctypes.data_as(C.POINTER(C.c_longlong))))
        return t, c


class FiniteVolumesSolver:
    """Per-patch kernel surface of exahype::solvers::FiniteVolumesSolver on the GPU
    (E/solvers/FiniteVolumesSolver.h:235-400).  Patches include their ghost layers
    in the reference layout [(z)][y][x][nData]."""

    def __init__(self, dim, patch, scheme=FV_GODUNOV, cfl=0.9, gamma=1.4, pde=None, device=0):
        L = load_library()
        _check(L.exa_init(device))
        if pde is None:
            pde = PDE_EULER if dim == 3 else PDE_EULERFV2D
        self.cfg = FVConfig(dim, patch, scheme, 5, 0, cfl, gamma, pde)
        h = C.c_void_p()
        _check(L.exa_fv_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.dim, self.patch = dim, patch
        self.ghost = 2 if scheme == FV_MUSCLHANCOCK else 1
        self.side = patch + 2 * self.ghost
        self.patch_size = self.side ** dim * 5

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_fv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solutionUpdate(self, luh_new, luh, dx, dt):
        luh_new = np.array(luh_new, dtype=np.float64, copy=True)
        adm = C.c_double(0.0)
        _check(_lib.exa_fv_solution_update(self._h, _p(luh_new), _p(np.ascontiguousarray(luh, np.float64)),
                                           _p(_vec3(dx)), float(dt), C.byref(adm)))
        return luh_new, adm.value

    def stableTimeStepSize(self, luh, dx):
        dt = C.c_double(0.0)
        _check(_lib.exa_fv_stable_time_step_size(self._h, _p(np.ascontiguousarray(luh, np.float64)), _p(_vec3(dx)),
                                                 C.byref(dt)))
        return dt.value

    def riemannSolver(self, qL, qR, normalNonZero):
        fL, fR, smax = np.zeros(5), np.zeros(5), C.c_double(0.0)
        _check(_lib.exa_fv_riemann_solver(self._h, _p(fL), _p(fR), _p(np.ascontiguousarray(qL, np.float64)),
                                          _p(np.ascontiguousarray(qR, np.float64)), int(normalNonZero),
                                          C.byref(smax)))
        return fL, fR, smax.value

    # boundary hooks (positions: a unit vector towards the face)
    def _block_size(self):
        return self.patch ** (self.dim - 1) * self.ghost * 5

    def ghostLayerFilling(self, luh, luhNeighbour, neighbourPosition):
        luh = np.array(luh, dtype=np.float64, copy=True)
        _check(_lib.exa_fv_ghost_layer_filling(self._h, _p(luh), _p(np.ascontiguousarray(luhNeighbour, np.float64)),
                                               (C.c_int * 3)(*neighbourPosition)))
        return luh

    def ghostLayerFillingAtBoundary(self, luh, luhbnd, boundaryPosition):
        luh = np.array(luh, dtype=np.float64, copy=True)
        _check(_lib.exa_fv_ghost_layer_filling_at_boundary(self._h, _p(luh),
                                                           _p(np.ascontiguousarray(luhbnd, np.float64)),
                                                           (C.c_int * 3)(*boundaryPosition)))
        return luh

    def boundaryLayerExtraction(self, luh, boundaryPosition):
        out = np.zeros(self._block_size())
        _check(_lib.exa_fv_boundary_layer_extraction(self._h, _p(out), _p(np.ascontiguousarray(luh, np.float64)),
                                                     (C.c_int * 3)(*boundaryPosition)))
        return out

    def boundaryConditions(self, bc, stateIn, centre, size, t, dt, faceIndex, normalNonZero):
        out = np.zeros(self._block_size())
        _check(_lib.exa_fv_boundary_conditions(self._h, C.byref(bc), _p(out),
                                               _p(np.ascontiguousarray(stateIn, np.float64)), _p(_vec3(centre)),
                                               _p(_vec3(size)), float(t), float(dt), int(faceIndex),
                                               int(normalNonZero)))
        return out


class FVGrid:
    """Grid of finite-volume patches resident in HBM (periodic unless
    set_boundary closes an axis)."""

    def __init__(self, solver, nx, ny, nz=1, offset=(0.0, 0.0, 0.0), width=(1.0, 1.0, 1.0)):
        self.solver = solver
        h = C.c_void_p()
        _check(_lib.exa_fv_grid_create(solver._h, nx, ny, nz, _p(_vec3(offset)), _p(_vec3(width)), C.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.offset, self.width = tuple(offset), tuple(width)
        self.npatch = nx * ny * nz

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_fv_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64).ravel()
        assert luh.size == self.npatch * self.solver.patch_size
        _check(_lib.exa_fv_grid_upload(self._h, _p(luh)))

    def download(self):
        out = np.zeros(self.npatch * self.solver.patch_size)
        _check(_lib.exa_fv_grid_download(self._h, _p(out)))
        return out

    def set_initial_conditions(self, ic):
        _check(_lib.exa_fv_grid_set_initial_conditions(self._h, C.byref(ic)))

    def set_boundary(self, periodic, bc=None):
        per = (C.c_int * 3)(*[1 if p else 0 for p in periodic])
        _check(_lib.exa_fv_grid_set_boundary(self._h, per, None if bc is None else C.byref(bc)))

    def compute_dt(self):
        dt = C.c_double(0.0)
        _check(_lib.exa_fv_grid_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def step(self, nsteps=1):
        _check(_lib.exa_fv_grid_step(self._h, int(nsteps)))

    def sync(self):
        t, dt, adm = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
        _check(_lib.exa_fv_grid_sync(self._h, C.byref(t), C.byref(dt), C.byref(adm)))
        return t.value, dt.value, adm.value

    def enable_timing(self, on=True):
        _check(_lib.exa_fv_grid_enable_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        """(ms summed, launches) per phase since the last call: ghost layers,
        MUSCL-Hancock states, update, time step."""
        t = np.zeros(4)
        c = (C.c_longlong * 4)()
        _check(_lib.exa_fv_grid_kernel_times(self._h, _p(t), c))
        return t, np.array(list(c))


def projectOnFVLimiterSpace(luh, order, ghost=0, lim=None):
    """Subcell averages of the DG polynomial (2N+1 subcells per direction)."""
    L = load_library()
    _check(L.exa_init(0))
    S = 2 * order + 1 + 2 * ghost
    lim = np.zeros(S ** 3 * 5) if lim is None else np.array(lim, dtype=np.float64, copy=True)
    _check(L.exa_limiter_project_on_fv(order, _p(np.ascontiguousarray(luh, np.float64)), ghost, _p(lim)))
    return lim


def projectOnDGSpace(lim, order, ghost=0):
    L = load_library()
    _check(L.exa_init(0))
    luh = np.zeros((order + 1) ** 3 * 5)
    _check(L.exa_limiter_project_on_dg(order, _p(np.ascontiguousarray(lim, np.float64)), ghost, _p(luh)))
    return luh


def discreteMaximumPrincipleAndMinAndMaxSearch(luh, order, bmin, bmax, relax, scaling, nobs=5):
    L = load_library()
    _check(L.exa_init(0))
    bmin = np.array(bmin, dtype=np.float64, copy=True)
    bmax = np.array(bmax, dtype=np.float64, copy=True)
    ok = C.c_int(0)
    _check(L.exa_limiter_dmp(order, _p(np.ascontiguousarray(luh, np.float64)), nobs, float(relax), float(scaling),
                             _p(bmin), _p(bmax), C.byref(ok)))
    return bool(ok.value), bmin, bmax


# ---------------------------------------------------------------------------
# AMR level transfers (kernels/aderdg/generic/c/3d/amrRoutines.cpph)

class AmrTransfer(C.Structure):
    _fields_ = [("fine", C.c_longlong), ("coarse", C.c_longlong), ("levels", C.c_int), ("index", C.c_int * 3)]


def _iarr(v, n):
    a = (C.c_int * n)()
    for i in range(n):
        a[i] = int(v[i])
    return a


def faceUnknownsProlongation(lQhbndCoarse, lFhbndCoarse, coarseGridLevel, fineGridLevel, subfaceIndex, order,
                             nvar=5, npar=0):
    """Fine face (lQhbnd, lFhbnd) of one sub-face; amrRoutines.cpph:45-144."""
    L = load_library()
    _check(L.exa_init(0))
    B2 = (order + 1) ** 2
    qf, ff = np.zeros(B2 * (nvar + npar)), np.zeros(B2 * nvar)
    _check(L.exa_amr_face_unknowns_prolongation(order, nvar, npar, _p(qf), _p(ff),
                                                _p(np.ascontiguousarray(lQhbndCoarse, np.float64)),
                                                _p(np.ascontiguousarray(lFhbndCoarse, np.float64)),
                                                coarseGridLevel, fineGridLevel, _iarr(subfaceIndex, 2)))
    return qf, ff


def faceUnknownsRestriction(lQhbndCoarse, lFhbndCoarse, lQhbndFine, lFhbndFine, coarseGridLevel, fineGridLevel,
                            subfaceIndex, order, nvar=5, npar=0):
    """Returns copies of the coarse face data with the fine face's restriction added."""
    L = load_library()
    _check(L.exa_init(0))
    qc = np.array(lQhbndCoarse, dtype=np.float64, copy=True)
    fc = np.array(lFhbndCoarse, dtype=np.float64, copy=True)
    _check(L.exa_amr_face_unknowns_restriction(order, nvar, npar, _p(qc), _p(fc),
                                               _p(np.ascontiguousarray(lQhbndFine, np.float64)),
                                               _p(np.ascontiguousarray(lFhbndFine, np.float64)),
                                               coarseGridLevel, fineGridLevel, _iarr(subfaceIndex, 2)))
    return qc, fc


def volumeUnknownsProlongation(luhCoarse, coarseGridLevel, fineGridLevel, subcellIndex, order, nvar=5, npar=0):
    L = load_library()
    _check(L.exa_init(0))
    out = np.zeros((order + 1) ** 3 * (nvar + npar))
    _check(L.exa_amr_volume_unknowns_prolongation(order, nvar, npar, _p(out),
                                                  _p(np.ascontiguousarray(luhCoarse, np.float64)),
                                                  coarseGridLevel, fineGridLevel, _iarr(subcellIndex, 3)))
    return out


def volumeUnknownsRestriction(luhCoarse, luhFine, coarseGridLevel, fineGridLevel, subcellIndex, order, nvar=5,
                              npar=0):
    L = load_library()
    _check(L.exa_init(0))
    out = np.array(luhCoarse, dtype=np.float64, copy=True)
    _check(L.exa_amr_volume_unknowns_restriction(order, nvar, npar, _p(out),
                                                 _p(np.ascontiguousarray(luhFine, np.float64)),
                                                 coarseGridLevel, fineGridLevel, _iarr(subcellIndex, 3)))
    return out


AMR_TRANSFER_DTYPE = np.dtype([("fine", np.int64), ("coarse", np.int64), ("levels", np.int32),
                               ("index", np.int32, (3,))], align=True)


def amr_transfer_array(fine, coarse, levels, index):
    """Transfer list as a numpy structured array (columns: fine and coarse
    offsets in doubles, level difference, sub-cell index per axis)."""
    fine = np.asarray(fine)
    a = np.zeros(fine.size, AMR_TRANSFER_DTYPE)
    a["fine"], a["coarse"], a["levels"] = fine, coarse, levels
    idx = np.asarray(index).reshape(fine.size, -1)
    a["index"][:, :idx.shape[1]] = idx
    return a


def _transfers(lst):
    if isinstance(lst, np.ndarray):
        if lst.dtype != AMR_TRANSFER_DTYPE or not lst.flags.c_contiguous:
            raise ValueError("transfer arrays come from amr_transfer_array()")
        return lst.ctypes.data_as(C.POINTER(AmrTransfer))
    arr = (AmrTransfer * max(1, len(lst)))()
    for i, (fine, coarse, levels, idx) in enumerate(lst):
        arr[i].fine, arr[i].coarse, arr[i].levels = int(fine), int(coarse), int(levels)
        for a in range(len(idx)):
            arr[i].index[a] = int(idx[a])
    return arr


class DeviceArray:
    """float64 array in device memory of the current device (exa_device_malloc)."""

    def __init__(self, n_or_host):
        L = load_library()
        _check(L.exa_init(0))
        host = None if np.isscalar(n_or_host) else np.ascontiguousarray(n_or_host, np.float64).ravel()
        self.n = int(n_or_host) if host is None else host.size
        p = C.c_void_p()
        _check(L.exa_device_malloc(C.byref(p), 8 * max(self.n, 1)))
        self.ptr = p
        if host is not None and self.n:
            _check(L.exa_memcpy_htod(self.ptr, host.ctypes.data_as(C.c_void_p), 8 * self.n))

    def upload(self, host):
        host = np.ascontiguousarray(host, np.float64).ravel()
        if host.size != self.n:
            raise ValueError("size mismatch")
        _check(_lib.exa_memcpy_htod(self.ptr, host.ctypes.data_as(C.c_void_p), 8 * self.n))

    def numpy(self):
        out = np.empty(self.n)
        if self.n:
            _check(_lib.exa_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n))
        return out

    def close(self):
        if getattr(self, "ptr", None):
            _lib.exa_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_copy(dst, src, nbytes):
    """Streaming device copy (library kernel); returns its duration in ms."""
    ms = C.c_float(0.0)
    _check(load_library().exa_device_copy(dst.ptr, src.ptr, int(nbytes), C.byref(ms)))
    return ms.value


def selftest_mfma(kind, A, Bm):
    """D = A @ B through the matrix cores (kind 0 FP32, 1 BF16) with the
    operand layouts of the mixed-precision predictor."""
    L = load_library()
    _check(L.exa_init(0))
    A = np.ascontiguousarray(A, np.float32)
    Bm = np.ascontiguousarray(Bm, np.float32)
    D = np.zeros((16, 16), np.float32)
    _check(L.exa_selftest_mfma(kind, A.ctypes.data_as(C.c_void_p), Bm.ctypes.data_as(C.c_void_p),
                               D.ctypes.data_as(C.c_void_p)))
    return D


def _dev(a):
    if not isinstance(a, DeviceArray):
        raise ValueError("expected a DeviceArray")
    return a.ptr, a.n


def amrProlongationBatched(dim, order, ndata, fine, coarse, transfers):
    """Batched prolongations between device tensors; transfers = [(fine_off, coarse_off, levels, index)]."""
    L = load_library()
    _check(L.exa_init(0))
    pf, nf = _dev(fine)
    pc, nc = _dev(coarse)
    _check(L.exa_amr_prolongation_batched(dim, order, ndata, pf, nf, pc, nc, _transfers(transfers), len(transfers)))


def amr_last_kernel_ms():
    """Kernel duration (HIP events) of the last batched AMR call on this device."""
    ms = C.c_double(0.0)
    _check(load_library().exa_amr_last_kernel_ms(C.byref(ms)))
    return ms.value


def amrRestrictionBatched(dim, order, ndata, coarse, fine, transfers):
    """Batched restrictions (added into coarse); runs of one coarse offset must be consecutive."""
    L = load_library()
    _check(L.exa_init(0))
    pf, nf = _dev(fine)
    pc, nc = _dev(coarse)
    _check(L.exa_amr_restriction_batched(dim, order, ndata, pc, nc, pf, nf, _transfers(transfers), len(transfers)))


def amrRestrictionBatchedNoAlloc(dim, order, ndata, coarse, fine, transfers):
    """amrRestrictionBatched through the allocation-free entry point (at most
    EXA_AMR_NOALLOC_MAX_TRANSFERS transfers; exa_init must have run)."""
    L = load_library()
    _check(L.exa_init(0))
    pf, nf = _dev(fine)
    pc, nc = _dev(coarse)
    _check(L.exa_amr_restriction_batched_noalloc(dim, order, ndata, pc, nc, pf, nf, _transfers(transfers),
                                                 len(transfers)))


AMR_NOALLOC_MAX_TRANSFERS = 16384


# ---------------------------------------------------------------------------
# Ensembles of independent grids

class AMRConfig(C.Structure):
    _fields_ = [("base_cells", C.c_int), ("levels", C.c_int), ("centre", C.c_double * 3), ("radius", C.c_double),
                ("width", C.c_double * 2), ("dmp_relaxation", C.c_double), ("dmp_scaling", C.c_double),
                ("steps_till_cured", C.c_int), ("patch_capacity", C.c_longlong)]


def amr_config(base_cells, levels, centre, radius, widths, relax=1e-4, scaling=1e-3, steps_till_cured=0,
               patch_capacity=0):
    c = AMRConfig()
    c.base_cells, c.levels, c.radius = base_cells, levels, radius
    for a in range(3):
        c.centre[a] = centre[a]
    for l in range(2):
        c.width[l] = widths[l] if l < len(widths) else 0.0
    c.dmp_relaxation, c.dmp_scaling, c.steps_till_cured = relax, scaling, steps_till_cured
    c.patch_capacity = patch_capacity
    return c


def amr_mesh_cells(cfg):
    """Leaves [n][4] = (level, ix, iy, iz) of the mesh a configuration builds (host only)."""
    L = load_library()
    n = C.c_longlong(0)
    _check(L.exa_amr_mesh_cells(C.byref(cfg), None, 0, C.byref(n)))
    out = np.zeros((max(n.value, 1), 4), np.int32)
    _check(L.exa_amr_mesh_cells(C.byref(cfg), out.ctypes.data_as(C.POINTER(C.c_int)), n.value, C.byref(n)))
    return out[:n.value]


AMR_STATUS_NAMES = ("Ok", "NeighbourOfTroubled4", "NeighbourOfTroubled3", "NeighbourOfTroubled2",
                    "NeighbourOfTroubled1", "Troubled")
AMR_COUNTER_NAMES = ("irregular_steps", "refinement_requests", "coarse_fine_limiter", "mirrored_ghost_faces",
                     "cannot_communicate", "patch_overflow", "troubled", "fv_updates")


class AMRGrid:
    """Limiting ADER-DG on a static adaptive mesh resident in HBM (leaves of
    up to three levels, MUSCL-Hancock subcell limiter on the finest level)."""

    def __init__(self, solver, cfg):
        self.solver = solver
        self.cfg = cfg
        h = C.c_void_p()
        _check(_lib.exa_amr_grid_create(solver._h, C.byref(cfg), C.byref(h)))
        self._h = h
        n = C.c_longlong(0)
        _check(_lib.exa_amr_grid_num_cells(self._h, C.byref(n)))
        self.ncells = n.value
        self.cell_size = solver.B ** 3 * solver.nvar

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_amr_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self):
        out = np.zeros((self.ncells, 4), np.int32)
        _check(_lib.exa_amr_grid_cells(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def upload(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64).ravel()
        assert luh.size == self.ncells * self.cell_size
        _check(_lib.exa_amr_grid_upload(self._h, _p(luh)))

    def download(self):
        out = np.zeros(self.ncells * self.cell_size)
        _check(_lib.exa_amr_grid_download(self._h, _p(out)))
        return out

    def set_initial_conditions(self, ic):
        _check(_lib.exa_amr_grid_set_initial_conditions(self._h, C.byref(ic)))

    def compute_dt(self):
        dt = C.c_double(0.0)
        _check(_lib.exa_amr_grid_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def use_graph(self, enable):
        _check(_lib.exa_amr_grid_use_graph(self._h, int(bool(enable))))

    def enable_timing(self, enable):
        _check(_lib.exa_amr_grid_enable_timing(self._h, int(bool(enable))))

    def step(self, nsteps=1):
        _check(_lib.exa_amr_grid_step(self._h, int(nsteps)))

    def sync(self):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        _check(_lib.exa_amr_grid_sync(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def limiter_status(self):
        out = np.zeros(self.ncells, np.int32)
        _check(_lib.exa_amr_grid_limiter_status(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def patches(self):
        """(cells, interiors [k][z][y][x][5]) of the limiter patches."""
        n = C.c_longlong(0)
        _check(_lib.exa_amr_grid_patches(self._h, None, None, 0, C.byref(n)))
        P = 2 * self.solver.order + 1
        out = np.zeros(max(n.value, 1) * 5 * P ** 3)
        cells = np.zeros(max(n.value, 1), np.int32)
        _check(_lib.exa_amr_grid_patches(self._h, _p(out), cells.ctypes.data_as(C.POINTER(C.c_int)), n.value,
                                         C.byref(n)))
        return cells[:n.value], out[:n.value * 5 * P ** 3].reshape(n.value, 5 * P ** 3)

    def counters(self):
        c = (C.c_longlong * 8)()
        _check(_lib.exa_amr_grid_counters(self._h, c))
        return dict(zip(AMR_COUNTER_NAMES, list(c)))

    def total_picard_iterations(self):
        n = C.c_longlong(0)
        _check(_lib.exa_amr_grid_total_picard_iterations(self._h, C.byref(n)))
        return n.value

    def phase_times(self):
        ms = np.zeros(6)
        n = C.c_longlong(0)
        _check(_lib.exa_amr_grid_phase_times(self._h, _p(ms), C.byref(n)))
        return ms, n.value


class Ensemble:
    """Independent periodic grids (order, cells per direction, gamma per
    member) advanced by shared launches, one hipGraph per step."""

    def __init__(self, members, precision=PREC_FP64, fp64_sweeps=1, cfl=0.9, waves_per_simd=0, contraction=0):
        L = load_library()
        _check(L.exa_init(0))
        arr = (EnsembleMember * len(members))()
        self.members = []
        for i, m in enumerate(members):
            order, (nx, ny, nz), gamma = m["order"], m["cells"], m.get("gamma", 1.4)
            arr[i].order, arr[i].nx, arr[i].ny, arr[i].nz, arr[i].gamma = order, nx, ny, nz, gamma
            for d in range(3):
                arr[i].offset[d] = m.get("offset", (0.0, 0.0, 0.0))[d]
                arr[i].width[d] = m.get("width", (1.0, 1.0, 1.0))[d]
            self.members.append(dict(order=order, cells=(nx, ny, nz), gamma=gamma))
        cfg = AderdgConfig(0, 5, 0, cfl, 1.4, PDE_EULER, precision, fp64_sweeps, waves_per_simd, contraction)
        h = C.c_void_p()
        _check(L.exa_ensemble_create(arr, len(members), C.byref(cfg), C.byref(h)))
        self._h = h
        self.n = len(members)

    def close(self):
        if getattr(self, "_h", None):
            _lib.exa_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_initial_conditions(self, ics):
        arr = (InitialCondition * self.n)(*ics)
        _check(_lib.exa_ensemble_set_initial_conditions(self._h, arr))

    def compute_dt(self):
        _check(_lib.exa_ensemble_compute_dt(self._h))
        return self.sync()[1]

    def use_graph(self, enable=True):
        _check(_lib.exa_ensemble_use_graph(self._h, 1 if enable else 0))

    def step(self, nsteps=1):
        _check(_lib.exa_ensemble_step(self._h, int(nsteps)))

    def sync(self):
        t, dt = np.zeros(self.n), np.zeros(self.n)
        _check(_lib.exa_ensemble_sync(self._h, _p(t), _p(dt)))
        return t, dt

    def download(self, i):
        m = self.members[i]
        nx, ny, nz = m["cells"]
        out = np.zeros(nx * ny * nz * (m["order"] + 1) ** 3 * 5)
        _check(_lib.exa_ensemble_download(self._h, i, _p(out)))
        return out

    def total_picard_iterations(self):
        n = C.c_longlong(0)
        _check(_lib.exa_ensemble_total_picard_iterations(self._h, C.byref(n)))
        return n.value


def autotune(order, precision=PREC_FP64, ncells=32):
    """Times the predictor variants of one order/precision on this device and
    records the fastest; returns (waves_per_simd, contraction, times_ms)."""
    L = load_library()
    _check(L.exa_init(0))
    w, c = C.c_int(0), C.c_int(0)
    ms = np.zeros(6)
    _check(L.exa_autotune(order, precision, ncells, C.byref(w), C.byref(c), _p(ms)))
    return w.value, c.value, ms
