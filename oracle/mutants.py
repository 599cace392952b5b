"""Mutation check of the oracle's pin tests.

TEST INFRASTRUCTURE ONLY.  Each mutant below plants one plausible one-line mistake (a
dropped term, a wrong sign or index, a swapped operand, a wrong constant) into a copy of
cs3d_oracle.cpp, builds it into a temporary directory and runs tests/test_oracle_pins.py
against it through CS3D_ORACLE_LIB.  A pin suite that lets a mutant pass does not pin the
function the mutant touches.  The real oracle/liboracle.so is never rebuilt or replaced.

    python oracle/mutants.py            # all mutants; exit 1 if any survives or the unchanged
                                        # source (the control) fails
    python oracle/mutants.py pressure   # selected mutants by name
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cs3d_oracle.cpp")
CXXFLAGS = ["-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
            "-std=c++17", "-shared"]

ACOS_H = os.path.join(ROOT, "cellsim3d_amd", "csrc", "cs3d_acos.h")
ACOS_INCLUDE = '#include "../cellsim3d_amd/csrc/cs3d_acos.h"'

# name, original text (must occur exactly once), replacement[, file: the oracle source by
# default, or ACOS_H for the acos the oracle shares with the kernels]
MUTANTS = [
    ("angle_map",
     "t1 = th0[3 * ni + 0];  // aij\n        t2 = th0[3 * ni + 1];  // ajk",
     "t1 = th0[3 * ni + 1];  // aij\n        t2 = th0[3 * ni + 0];  // ajk"),
    ("spring_rest_length", "T r0 = R0[i * S + atom];", "T r0 = R0[((i + 1) % 3) * S + atom];"),
    ("friction_projection", "V3<T> vTau = vij - dot(vij, nrm) * nrm;", "V3<T> vTau = vij;"),
    ("pentagon_factor", "if (f < 12) fac = (T)0.2f;", "if (f < 12) fac = (T)0.166666667f;"),
    ("com_divisor", "*dst[a] = s[0] / T(180);", "*dst[a] = s[0] / T(181);"),
    ("contact_sign", "cf.x += -as * E * (ar - R) / R * dX;", "cf.x += as * E * (ar - R) / R * dX;"),
    ("grid_dims", "int xd = (int)((e[1] - e[0]) / DL + T(1));", "int xd = (int)((e[1] - e[0]) / DL + T(2));"),
    ("pressure_term", "(T(3) * Pr[rank])", "(T(2) * Pr[rank])"),
    ("division_side", "if (pd < T(0) || pd == 0)", "if (pd > T(0))"),
    ("bending_neighbour",
     "nodeF = nodeF + angle_force(pos - nP, P(b) - nP, t2, k, st);",
     "nodeF = nodeF + angle_force(pos - nP, P(c) - nP, t2, k, st);"),
    ("integrate_half_kick",
     "0.5 * (double)(dt * FC[a] + dt * FD[a] + root_dt * fr)",
     "1.0 * (double)(dt * FC[a] + dt * FD[a] + root_dt * fr)"),
    ("bbox_max", "if (mx < s[a][t]) mx = s[a][t];", "if (mx < s[a][t] && t > 1) mx = s[a][t];"),
    ("internal_damping_sign", "force = force - gint * (vi - vn);", "force = force - gint * (vn - vi);"),
    ("philox_multiplier", "const uint32_t m0 = 0xD2511F53u", "const uint32_t m0 = 0xD2511F55u"),
    ("philox_weyl", "k1 += 0xBB67AE85u;", "k1 += 0xBB67AE86u;"),
    ("ln_series_term", "q = q * z2 + 0.2f;", "q = q * z2 + 0.25f;"),
    ("polar_factor", "std::sqrt(-2.0f * ln_unit(s) / s)", "std::sqrt(-1.0f * ln_unit(s) / s)"),
    ("polar_reject", "s > 0.0f && s < 1.0f", "s > 0.0f && s < 1.5f"),
    ("sys_centre_count",
     "for (int a = 0; a < 3; ++a) c3[a] = c3[a] / (T)n;",
     "for (int a = 0; a < 3; ++a) c3[a] = c3[a] / (T)(n + 1);"),
    ("sys_centre_sign", "X[g] = X[g] - c3[0];", "X[g] = X[g] + c3[0];"),
    ("acos_coefficient", "f = fma(f, z, 0x1.6db6db6e31f13p-5);", "f = fma(f, z, 0x1.6db6db6e31f13p-6);", ACOS_H),
    ("acos_negative_branch", "const int ki = big ? (x > 0.0 ? -2 : 2) : 1;",
     "const int ki = big ? (x > 0.0 ? -2 : 1) : 1;", ACOS_H),
    ("acos_half_angle", "fma(-0.5, ax, 0.5) :", "fma(-0.5, ax, 0.5000001) :", ACOS_H),
    ("acos_pi_low", "-m * 0x1.1a62633145c07p-54", "m * 0x1.1a62633145c07p-54", ACOS_H),
    ("acos_sign", "const double s = big ? sqrt(z) : x;", "const double s = big ? sqrt(z) : ax;", ACOS_H),
]


def run(name, old, new, tmp, path=SRC):
    """Builds the oracle with `old` replaced by `new` in `path` (unchanged when old is None) and
    runs the pin tests on it.  The result starts with "killed" (a test failed or the process
    crashed), "passed" or "error"."""
    text = open(SRC).read()
    if old is not None:
        target = open(path).read()
        if target.count(old) != 1:
            return "error: anchor not found once"
        target = target.replace(old, new, 1)
        if path == SRC:
            text = target
        else:  # a mutated copy of the shared acos header, included by absolute path
            hdr = os.path.join(tmp, name + "_acos.h")
            with open(hdr, "w") as f:
                f.write(target)
            if text.count(ACOS_INCLUDE) != 1:
                return "error: acos include not found once"
            text = text.replace(ACOS_INCLUDE, '#include "%s"' % hdr)
    cpp = os.path.join(tmp, name + ".cpp")
    so = os.path.join(tmp, name + ".so")
    with open(cpp, "w") as f:
        f.write(text)
    cc = subprocess.run(["g++", *CXXFLAGS, "-I", HERE, "-o", so, cpp], capture_output=True, text=True)
    if cc.returncode:
        return "error: does not compile: " + cc.stderr.strip().splitlines()[-1]
    env = dict(os.environ, CS3D_ORACLE_LIB=so, OMP_NUM_THREADS="2")
    t = subprocess.run([sys.executable, "-m", "pytest", "tests/test_oracle_pins.py", "-x", "-q",
                        "-p", "no:cacheprovider", "-m", "not gpu"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    last = t.stdout.strip().splitlines()[-1] if t.stdout.strip() else "(no output)"
    if t.returncode == 0:
        return "passed: " + last
    if t.returncode == 1 or t.returncode < 0:  # a test failed, or the process died on a signal
        return "killed: " + last
    return f"error: pytest exit {t.returncode}: " + last


def check(names=(), workers=4):
    """Runs the unchanged source (as "control", which must pass) and the chosen mutants (all by
    default), `workers` at a time; returns {name: result}."""
    chosen = [("control", None, None)] + [m for m in MUTANTS if not names or m[0] in names]
    with tempfile.TemporaryDirectory() as tmp, ThreadPoolExecutor(workers) as ex:
        futs = [(m[0], ex.submit(run, m[0], m[1], m[2], tmp, *m[3:])) for m in chosen]
        return {name: f.result() for name, f in futs}


def main(argv):
    res = check(argv)
    for name, r in res.items():
        print(f"{name:24s} {r}")
    ok = res.pop("control").startswith("passed")
    survivors = sum(not r.startswith("killed") for r in res.values())
    print(f"{len(res) - survivors} of {len(res)} mutants killed" + ("" if ok else "; the control FAILED"))
    return 1 if survivors or not ok else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
