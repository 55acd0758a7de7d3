"""The closed-form pins of tests/test_oracle_fem_pins.py detect wrong FEM
constants (CPU only).

Each case copies the oracle sources, changes one constant of the hex FEM
restatement (a scale, a mass, a source amplitude or time, a time-stepping
weight), builds the copy, and runs the pin that covers the constant against
it in a child process (EMS_ORACLE_LIB selects the build): the pin must fail.
A control build of the unchanged sources must pass the same pins.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, file, original text, mutated text, pin call)
MUTANTS = [
    ("symplectic operator scale 1/(hp)^2 -> 1/h^2", "fem_oracle.c",
     "sc = 4.0 / (h * p) / (h * p) / cfg->mu / cfg->eps;", "sc = 4.0 / h / h / cfg->mu / cfg->eps;",
     "t.test_cavity_eigenvalue_converges(3, (9, 12))"),
    ("leapfrog operator mass 8/p^3 -> 8", "fem_oracle.c",
     "/ cfg->mu / cfg->eps / (8.0 / p / p / p);", "/ cfg->mu / cfg->eps / 8.0;",
     "t.test_leapfrog_operator_is_dt2_over_mass_times_symplectic(2)"),
    ("symplectic mass 8/p^3 -> 8/p^2", "fem_oracle.c",
     "const double m = 8.0 / p / p / p;", "const double m = 8.0 / p / p;",
     "t.test_point_source_gauss_law('symplectic_2', 2, 4.0)"),
    ("symplectic source amplitude 8/(eps (hp)^3) -> 8/(eps h^3)", "fem_oracle.c",
     "* 8.0 / s->c.eps / hp / hp / hp;", "* 8.0 / s->c.eps / s->c.h / s->c.h / s->c.h;",
     "t.test_point_source_gauss_law('symplectic_2', 3, 4.0)"),
    ("symplectic source time t+off+D/2 -> t+off", "fem_oracle.c",
     "ems_oracle_fem_source(s->t_now + offset + 0.5 * deltat)", "ems_oracle_fem_source(s->t_now + offset)",
     "t.test_point_source_gauss_law('symplectic_2', 1, 4.0)"),
    ("symplectic_4 middle-stage offset x1 dt -> 0", "fem_oracle.c",
     "symplectic_stage(s, x2 * s->c.dt, x1 * s->c.dt);", "symplectic_stage(s, x2 * s->c.dt, 0.0);",
     "t.test_point_source_gauss_law('symplectic_4', 1, 16.0)"),
    ("leapfrog source amplitude dt^2 -> dt", "fem_oracle.c",
     "ems_oracle_fem_source(s->t_now) * s->c.dt * s->c.dt / s->c.eps",
     "ems_oracle_fem_source(s->t_now) * s->c.dt / s->c.eps",
     "t.test_point_source_gauss_law('leapfrog', 2, 4.0)"),
    ("leapfrog source time t_n -> t_n+1", "fem_oracle.c",
     "ems_oracle_fem_source(s->t_now) * s->c.dt * s->c.dt / s->c.eps",
     "ems_oracle_fem_source(s->t_now + s->c.dt) * s->c.dt * s->c.dt / s->c.eps",
     "t.test_point_source_gauss_law('leapfrog', 1, 4.0)"),
    ("fem_abm AM3 field weights (5, 8, -1) -> (8, 5, -1)", "fem1_oracle.c",
     "E3[q][i] = (5.0 * H2[q][i] + 8.0 * H1[q][i] - H0[q][i])",
     "E3[q][i] = (8.0 * H2[q][i] + 5.0 * H1[q][i] - H0[q][i])",
     "t.test_abm_weights_are_adams_weights()"),
    ("fem_abm AB2 predictor weights (3, -1)/2 -> (2, -1)/2", "fem1_oracle.c",
     "H3[q][i] = -(3.0 * E1[q][i] - E0[q][i]) * dt / 8.0 / mu / h;",
     "H3[q][i] = -(2.0 * E1[q][i] - E0[q][i]) * dt / 8.0 / mu / h;",
     "t.test_abm_weights_are_adams_weights()"),
    ("fem_abm AM3 source weights (5, 8, -1) -> (8, 5, -1)", "fem1_oracle.c",
     "add_points(s, E2, (5.0 * sf_abm(t + dt, h) + 8.0 * sf_abm(t, h) - sf_abm(t - dt, h))",
     "add_points(s, E2, (8.0 * sf_abm(t + dt, h) + 5.0 * sf_abm(t, h) - sf_abm(t - dt, h))",
     "t.test_fem1_point_source_gauss_law('fem_abm', 8.0)"),
    ("fem point source time t_n -> t_n+1", "fem1_oracle.c",
     "add_points(s, E2, sf_fem(t, c->mu, c->h) * c->dt * c->dt",
     "add_points(s, E2, sf_fem(t + c->dt, c->mu, c->h) * c->dt * c->dt",
     "t.test_fem1_point_source_gauss_law('fem', 4.0)"),
]

CONTROL = ("t.test_cavity_eigenvalue_converges(3, (9, 12)); t.test_cavity_eigenvalue_order_2(); "
           "t.test_leapfrog_operator_is_dt2_over_mass_times_symplectic(2); "
           "[t.test_point_source_gauss_law(*a) for a in [('symplectic_2', 1, 4.0), ('symplectic_2', 2, 4.0), "
           "('symplectic_2', 3, 4.0), ('symplectic_4', 1, 16.0), ('leapfrog', 1, 4.0), ('leapfrog', 2, 4.0)]]; "
           "t.test_abm_weights_are_adams_weights(); t.test_fem1_point_source_gauss_law('fem_abm', 8.0); "
           "t.test_fem1_point_source_gauss_law('fem', 4.0)")


def build_copy(tmp, fname=None, old=None, new=None):
    src = tmp / "src"
    src.mkdir()
    for f in os.listdir(ORACLE):
        if f.endswith("_oracle.c"):
            shutil.copy(os.path.join(ORACLE, f), src / f)
    if fname:
        text = (src / fname).read_text()
        assert text.count(old) == 1, "mutation site not found exactly once: %r" % old
        (src / fname).write_text(text.replace(old, new))
    lib = tmp / "libems_oracle.so"
    srcs = sorted(str(src / f) for f in os.listdir(src))
    subprocess.run(["gcc", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-shared", "-o", str(lib)]
                   + srcs + ["-lm"], check=True)
    return lib


def run_pin(lib, call):
    env = dict(os.environ, EMS_ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    code = "import sys; sys.path.insert(0, %r); import tests.test_oracle_fem_pins as t; %s" % (ROOT, call)
    return subprocess.run([sys.executable, "-c", code], env=env, cwd=ROOT, capture_output=True, text=True)


def test_control_build_passes(tmp_path):
    r = run_pin(build_copy(tmp_path), CONTROL)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("name,fname,old,new,call", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_its_pin(tmp_path, name, fname, old, new, call):
    r = run_pin(build_copy(tmp_path, fname, old, new), call)
    assert r.returncode != 0, "pin passed on a mutated oracle: " + name
    assert "AssertionError" in r.stderr, r.stderr[-3000:]
