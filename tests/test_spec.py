"""The specification-file front end (exahype_amd/spec.py): the runtime
parser's token stream and queries (E/Parser.cpp), the runner's regular mesh
(E/solvers/Solver.cpp:147-156, E/runners/Runner.cpp:321-345) and the
toolkit's kernel selection (SolverFactory.java:39-88)."""
import os

import pytest

from exahype_amd import spec as S

FV_SPEC = """/**
 * A finite-volumes project: Euler in 2D on patches of 10 x 10 volumes.
 */
exahype-project EulerPatches
  peano-kernel-path const = ./Peano
  output-directory const  = ./out

  computational-domain
    dimension const = 2
    width           = 1.0, 1.0
    offset          = 0.0, 0.0
    end-time        = 0.4
  end computational-domain

  /* solver block */
  solver Finite-Volumes PatchSolver
    variables const   = rho:1,j:3,E:1
    patch-size const  = 10
    maximum-mesh-size = 2e-2
    time-stepping     = global
    kernel const      = generic::godunov
    language const    = C
    plot vtu::Cartesian::cells::ascii Writer
      variables const = 5
      time            = 0.0
      repeat          = 0.01
      output          = ./variables
    end plot
  end solver
end exahype-project
"""

DG_SPEC = """exahype-project Vortex
  computational-domain
    dimension const = 3
    width           = 1.0, 1.0, 1.0
    offset          = 0.0, 0.0, 0.0
    end-time        = 0.1
  end computational-domain
  solver ADER-DG EulerSolver
    variables const    = 5
    order const        = 5
    maximum-mesh-size  = 0.05
    time-stepping      = globalfixed
    kernel const       = generic::fluxes::nonlinear
    language const     = C
  end solver
  solver ADER-DG LinearSolver
    variables const    = q:5
    parameters const   = 0
    order const        = 3
    maximum-mesh-size  = 0.3
    time-stepping      = global
    kernel const       = generic::fluxes::linear
  end solver
end exahype-project
"""


def test_token_stream():
    toks = S.tokenize(FV_SPEC)
    assert toks == ["exahype-project", "EulerPatches", "peano-kernel-path", "./Peano", "output-directory", "./out",
                    "computational-domain", "dimension", "2", "width", "1.0", "1.0", "offset", "0.0", "0.0",
                    "end-time", "0.4", "solver", "Finite-Volumes", "PatchSolver", "variables", "rho:1", "j:3", "E:1",
                    "patch-size", "10", "maximum-mesh-size", "2e-2", "time-stepping", "global", "kernel",
                    "generic::godunov", "language", "C", "plot", "vtu::Cartesian::cells::ascii", "Writer",
                    "variables", "5", "time", "0.0", "repeat", "0.01", "output", "./variables"]


def test_fv_queries_and_mesh():
    sp = S.Spec(FV_SPEC)
    assert sp.project_name == "EulerPatches"
    assert sp.dimension == 2 and sp.domain_size == (1.0, 1.0) and sp.offset == (0.0, 0.0)
    assert sp.end_time == 0.4 and sp.num_solvers == 1
    assert sp.solver_type(0) == "Finite-Volumes" and sp.identifier(0) == "PatchSolver"
    assert sp.variables(0) == 5 and sp.parameters(0) == 0 and sp.patch_size(0) == 10
    assert sp.maximum_mesh_size(0) == 0.02 and sp.time_stepping(0) == "global"
    cells, h, width = sp.mesh(0)
    assert cells == (81, 81) and h == pytest.approx(1 / 81) and width == pytest.approx((1.0, 1.0))
    cfg = sp.solver_config(0)
    assert cfg["scheme"] == "godunov" and cfg["patch_size"] == 10


def test_aderdg_queries_and_kernels():
    sp = S.Spec(DG_SPEC)
    assert sp.num_solvers == 2 and sp.dimension == 3
    c0, c1 = sp.solver_config(0), sp.solver_config(1)
    assert (c0["order"], c0["nvar"], c0["linear"], c0["time_stepping"]) == (5, 5, False, "globalfixed")
    assert c0["cells"] == (27, 27, 27)  # 1/27 <= 0.05 < 1/9
    assert (c1["order"], c1["nvar"], c1["npar"], c1["linear"]) == (3, 5, 0, True)
    assert c1["cells"] == (9, 9, 9)  # the level is at least 3


def test_mesh_level_rule():
    assert S.compute_mesh_level(2e-2, 1.0) == 5
    assert S.compute_mesh_level(1 / 81, 1.0) == 5
    assert S.compute_mesh_level(0.5, 1.0) == 3
    assert S.compute_mesh_level(0.05, 2.0) == 5  # 2/81 <= 0.05 < 2/27
    sp = S.Spec(DG_SPEC.replace("width           = 1.0, 1.0, 1.0", "width           = 2.0, 1.0, 0.5"))
    cells, h, width = sp.mesh(0)
    assert h == pytest.approx(2 / 81) and cells == (81, 41, 21)  # scaled domain: ceil(w/h) cells


def test_c_number_semantics():
    assert S.atoi("rho:1") == 0 and S.atoi("10") == 10 and S.atoi("2e-2") == 2
    assert S.atof("2e-2") == 0.02 and S.atof("notoken") == 0.0 and S.atof(".5x") == 0.5


def test_interpretation_errors():
    with pytest.raises(S.SpecError):
        S.Spec(DG_SPEC.replace("ADER-DG EulerSolver", "DG EulerSolver")).solver_type(0)
    with pytest.raises(S.SpecError):
        S.Spec("/* never closed\nexahype-project X\n")
    with pytest.raises(S.SpecError):
        S.Spec(FV_SPEC.replace("generic::godunov", "user::defined")).solver_config(0)


REF = "/root/reference/Demonstrators"


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's demonstrators are not present")
def test_reference_demonstrator_specs():
    sp = S.Spec.from_file(os.path.join(REF, "EulerFV", "EulerFV.exahype"))
    cfg = sp.solver_config(0)
    assert (cfg["type"], cfg["scheme"], cfg["patch_size"], cfg["nvar"]) == ("Finite-Volumes", "godunov", 10, 5)
    assert cfg["cells"] == (81, 81) and sp.end_time == 0.4
    tp = S.Spec.from_file(os.path.join(REF, "TrivialProject", "TrivialProject.exahype"))
    assert tp.num_solvers == 0 and tp.dimension == 2 and tp.end_time == 10.0


def _with_optimisation(fuse, factor):
    block = ("  optimisation\n    fuse-algorithmic-steps        = %s\n"
             "    fuse-algorithmic-steps-factor = %s\n  end optimisation\n" % (fuse, factor))
    return DG_SPEC.replace("  solver ADER-DG EulerSolver", block + "  solver ADER-DG EulerSolver", 1)


def test_fused_time_stepping_options():
    sp = S.Spec(_with_optimisation("on", "0.99"))
    assert sp.fuse_algorithmic_steps() is True and sp.fuse_algorithmic_steps_factor() == 0.99
    assert S.Spec(_with_optimisation("off", "0.5")).fuse_algorithmic_steps() is False
    assert S.Spec(DG_SPEC).fuse_algorithmic_steps() is False  # absent: off (E/Parser.cpp:459-460)
    assert sp.solver_config(0)["order"] == 5  # the section does not disturb the solver queries
    for bad in ("1.5", "-0.1", "x"):
        with pytest.raises(S.SpecError):
            S.Spec(_with_optimisation("on", bad)).fuse_algorithmic_steps_factor()
    with pytest.raises(S.SpecError):
        S.Spec(_with_optimisation("maybe", "0.5")).fuse_algorithmic_steps()


LIMITING_SPEC = """exahype-project SodSphere
  computational-domain
    dimension const = 3
    width           = 1.0, 1.0, 1.0
    offset          = 0.0, 0.0, 0.0
    end-time        = 0.1
  end computational-domain
  optimisation
    fuse-algorithmic-steps        = on
    fuse-algorithmic-steps-factor = 0.99
  end optimisation
  solver Limiting-ADER-DG EulerSolver
    variables const          = 5
    order const              = 3
    maximum-mesh-size        = 0.12
    maximum-mesh-depth       = 2
    time-stepping            = global
    kernel const             = generic::fluxes::nonlinear
    language const           = C
    limiter-kernel const     = generic::musclhancock
    limiter-language const   = C
    dmp-relaxation-parameter = 1e-3
    dmp-difference-scaling   = 1e-2
    steps-till-cured         = 0
  end solver
end exahype-project
"""


def test_limiting_solver_queries():
    sp = S.Spec(LIMITING_SPEC)
    c = sp.solver_config(0)
    assert c["type"] == "Limiting-ADER-DG" and c["order"] == 3 and c["cells"] == (9, 9, 9)
    assert (c["dmp_relaxation_parameter"], c["dmp_difference_scaling"], c["steps_till_cured"]) == (1e-3, 1e-2, 0)
    assert sp.maximum_mesh_depth(0) == 2 and c["limiter_kernel"] == "generic::musclhancock"
