"""Mutation check of the oracle's pins.

Each case plants one plausible mistake in a copy of oracle/oracle.cpp (a
dropped branch, a dx/dz swap, a lost cell of a 1-D candidate, ...), builds it,
and runs tests/test_oracle.py against that build: at least one pin must fail.
A pin set that passes a broken oracle pins nothing, so this is what shows the
pins bind each part of the local solver (src/modeling/modeling.cu:308-357),
the initialisation, the padding, the picks and the qP slowness.  The mistakes
are planted in the oracle, on the CPU; nothing here touches the GPU path.
The migration restatement (weights, indices, convolutions, Ricker wavelet) is
checked the same way against tests/test_migration_oracle.py.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")

# (name, text in oracle.cpp (must occur exactly once), replacement)
_A_HEAD = ("if ((tv <= te + c.dx * Sref) && (te <= tv + c.dz * Sref) && (te - tev >= R(0)) &&\n"
           "      (tv - tev >= R(0))) {\n    const R ta = tev + te - tv;\n    const R tb = tev - te + tv;\n    t1 =")
_B_HEAD = ("} else if ((te - tev <= Sref * c.dz * c.dz / rsqrt_(c.dx * c.dx + c.dz * c.dz)) &&\n"
           "             (te - tev > R(0))) {\n    t2")
_C_HEAD = ("} else if ((tv - tev <= Sref * c.dx * c.dx / rsqrt_(c.dx * c.dx + c.dz * c.dz)) &&\n"
           "             (tv - tev > R(0))) {\n    t3")
_B_BODY = "t2 = te + c.dx * rsqrt_(Sref * Sref - ((te - tev) / c.dz) * ((te - tev) / c.dz));"
_C_BODY = "t3 = tv + c.dz * rsqrt_(Sref * Sref - ((tv - tev) / c.dx) * ((tv - tev) / c.dx));"
_X1D = ("const R t1d2 = te + c.dx * rmin(S[std::max(i - 1, 1) + (size_t)j1 * nzz],\n"
        "                                  S[std::min(i, nzz - 1) + (size_t)j1 * nzz]);")
_Z1D = ("const R t1d1 = tv + c.dz * rmin(S[i1 + (size_t)std::max(j - 1, 1) * nzz],\n"
        "                                  S[i1 + (size_t)std::min(j, nxx - 1) * nzz]);")

MUTANTS = [
    ("branch_B_removed", _B_HEAD, _B_HEAD.replace("} else if ((te", "} else if (false && (te")),
    ("branch_B_value_dropped", _B_BODY, ";"),
    ("branch_C_removed", _C_HEAD, _C_HEAD.replace("} else if ((tv", "} else if (false && (tv")),
    ("branch_C_value_dropped", _C_BODY, ";"),
    ("B_factor_dx_to_dz", _B_BODY, _B_BODY.replace("te + c.dx *", "te + c.dz *")),
    ("B_quotient_dz_to_dx", _B_BODY, _B_BODY.replace("/ c.dz)", "/ c.dx)")),
    ("B_threshold_dz2_to_dx2", _B_HEAD, _B_HEAD.replace("Sref * c.dz * c.dz /", "Sref * c.dx * c.dx /")),
    ("C_factor_dz_to_dx", _C_BODY, _C_BODY.replace("tv + c.dz *", "tv + c.dx *")),
    ("C_quotient_dx_to_dz", _C_BODY, _C_BODY.replace("/ c.dx)", "/ c.dz)")),
    ("C_threshold_dx2_to_dz2", _C_HEAD, _C_HEAD.replace("Sref * c.dx * c.dx /", "Sref * c.dz * c.dz /")),
    ("A_guard_first_two_conjuncts_dropped", _A_HEAD,
     _A_HEAD.replace("(tv <= te + c.dx * Sref) && (te <= tv + c.dz * Sref) && ", "")),
    ("A_guard_dx_dz_swapped", _A_HEAD,
     _A_HEAD.replace("(tv <= te + c.dx * Sref) && (te <= tv + c.dz * Sref)",
                     "(tv <= te + c.dz * Sref) && (te <= tv + c.dx * Sref)")),
    ("A_ta_tb_swapped", _A_HEAD,
     _A_HEAD.replace("const R ta = tev + te - tv;\n    const R tb = tev - te + tv;",
                     "const R tb = tev + te - tv;\n    const R ta = tev - te + tv;")),
    ("A_numerator_weights_swapped", "t1 = ((tb * c.dz2i + ta * c.dx2i) +\n",
     "t1 = ((tb * c.dx2i + ta * c.dz2i) +\n"),
    ("x_edge_1d_one_cell_i", _X1D, "const R t1d2 = te + c.dx * S[std::min(i, nzz - 1) + (size_t)j1 * nzz];"),
    ("x_edge_1d_one_cell_i_minus_1", _X1D, "const R t1d2 = te + c.dx * S[std::max(i - 1, 1) + (size_t)j1 * nzz];"),
    ("x_edge_1d_clamp_dropped", _X1D, _X1D.replace("std::max(i - 1, 1)", "(i - 1)")),
    ("z_edge_1d_one_cell_j", _Z1D, "const R t1d1 = tv + c.dz * S[i1 + (size_t)std::min(j, nxx - 1) * nzz];"),
    ("z_edge_1d_one_cell_j_minus_1", _Z1D, "const R t1d1 = tv + c.dz * S[i1 + (size_t)std::max(j - 1, 1) * nzz];"),
    ("z_edge_1d_spacing_dz_to_dx", _Z1D, _Z1D.replace("tv + c.dz *", "tv + c.dx *")),
    ("final_min_is_max", "  return rmin(t1D, t2D);\n}", "  return std::fmax(t1D, t2D);\n}"),
    ("init_uses_source_node_slowness", "T[index] = S[index] * rsqrt_(ddx * ddx + ddz * ddz);",
     "T[index] = S[sIdz + (size_t)sIdx * nzz] * rsqrt_(ddx * ddx + ddz * ddz);"),
    ("expand_rows_from_wrong_edge", "out[(nzz - i - 1) + j * nzz] = out[(nzz - NB - 1) + j * nzz];",
     "out[(nzz - i - 1) + j * nzz] = out[NB + j * nzz];"),
    ("catmull_rom_coefficient", "d * (R(2) * P[0] - R(5) * P[1] + R(4) * P[2] - P[3] +",
     "d * (R(2) * P[0] - R(4) * P[1] + R(4) * P[2] - P[3] +"),
    ("qp_christoffel_cross_term_sign", "(Cm[3] + Cm[8]) * p0 * p1;", "(Cm[3] - Cm[8]) * p0 * p1;"),
    ("compress_no_offset", "static_cast<uint16_t>(1.0f + (COMPRESS - 1)", "static_cast<uint16_t>(0.0f + (COMPRESS - 1)"),
    ("compress_constant_code_0", "(mx == mn) ? uint16_t(1)", "(mx == mn) ? uint16_t(0)"),
]

# Kirchhoff migration (src/migration/migration.cu:54-73,197-261,398-756), checked
# by tests/test_migration_oracle.py
_TAPER = "const float taper = std::exp(-0.5f * ((x - cmp) / (sigma + KDM_EPS)) * ((x - cmp) / (sigma + KDM_EPS)));"
MIG_MUTANTS = [
    ("kdm_dTs_dx_over_dz", "const float dTs_dx = 0.5f * (at(Ts, i, j + 1) - at(Ts, i, j - 1)) / dx;",
     "const float dTs_dx = 0.5f * (at(Ts, i, j + 1) - at(Ts, i, j - 1)) / dz;"),
    ("kdm_d2Tr_dz2_over_dxdz", "const float d2Tr_dz2 = (at(Tr, i + 1, j) - 2.0f * at(Tr, i, j) + at(Tr, i - 1, j)) / (dz * dz);",
     "const float d2Tr_dz2 = (at(Tr, i + 1, j) - 2.0f * at(Tr, i, j) + at(Tr, i - 1, j)) / (dx * dz);"),
    ("kdm_cross_derivative_half_stencil", "at(Ts, i - 1, j - 1)) / (4.0f * dx * dz);", "at(Ts, i - 1, j - 1)) / (2.0f * dx * dz);"),
    ("kdm_jterm_norms_multiplied", "std::sqrt(detHs * detHr) * S[c] / (norm_Ts * norm_Tr);",
     "std::sqrt(detHs * detHr) * S[c] * (norm_Ts * norm_Tr);"),
    ("kdm_normal_sign_in_angle", "std::fmax(-1.0f, ux_s * nx_n + uz_s * nz_n)", "std::fmax(-1.0f, ux_s * nx_n - uz_s * nz_n)"),
    ("kdm_obliquity_without_sqrt", "const float obliquity = std::sqrt(std::fmax(0.0f, cos_s) * std::fmax(0.0f, cos_r));",
     "const float obliquity = std::fmax(0.0f, cos_s) * std::fmax(0.0f, cos_r);"),
    ("kdm_spreading_source_only", "const float G = 1.0f / std::sqrt(R_s * R_r);", "const float G = 1.0f / std::sqrt(R_s * R_s);"),
    ("kdm_reflection_factor_0_5", "const float Rf = 1.0f + 0.2f * std::cos(theta);", "const float Rf = 1.0f + 0.5f * std::cos(theta);"),
    ("kdm_taper_in_z", _TAPER, _TAPER.replace("(x - cmp)", "(z - cmp)")),
    ("kdm_aperture_not_in_radians", "const float tan_ap = std::tan((float)((double)aperture * M_PI / 180.0));",
     "const float tan_ap = std::tan(aperture);"),
    ("kdm_reflection_angle_not_halved", "const float ra = 0.5f * std::acos(", "const float ra = 1.0f * std::acos("),
    ("kdm_time_index_rounded", "const int tId = f2i_rd(T / dt);", "const int tId = f2i_rd(T / dt + 0.5f);"),
    ("kdm_angle_index_transposed", "mId = (size_t)(i - NB) + (size_t)aId * nz + (size_t)cmpId * na * nz;",
     "mId = (size_t)aId + (size_t)(i - NB) * na + (size_t)cmpId * na * nz;"),
    ("conv_shift_without_tenth", "const int shift = nw / 2 + nw / 10;\n  for (int t = 0;",
     "const int shift = nw / 2;\n  for (int t = 0;"),
    ("conv_adjoint_is_a_convolution", "acc += (double)w[m] * (double)d[k0 + m];",
     "acc += (double)w[nw - 1 - m] * (double)d[k0 + m];"),
    ("conv_no_nfft_scale", "o[k] = (float)(acc * nfft);", "o[k] = (float)acc;"),
    ("ricker_fc_without_sqrt_pi", "const float fc = fmax / (3.0f * std::sqrt((float)M_PI));\n  const int nw",
     "const float fc = fmax / 3.0f;\n  const int nw"),
    ("ricker_without_factor_2", "w[k] = (1.0f - 2.0f * arg) * std::exp(-arg);", "w[k] = (1.0f - arg) * std::exp(-arg);"),
]

_CXX = ["g++", "-O1", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]


def _build(text, path):
    src = path[:-3] + ".cpp"
    with open(src, "w") as f:
        f.write(text)
    subprocess.run(_CXX + ["-o", path, src], check=True, capture_output=True)


def _pins_pass(lib_path, tests="test_oracle.py"):
    env = dict(os.environ, SEISFAT_ORACLE_LIB=lib_path)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", tests), "-x", "-q",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode == 0, r.stdout[-600:]


def _check(mutants, tests):
    base = open(SRC).read()
    for name, old, new in mutants:
        assert base.count(old) == 1, f"{name}: anchor text occurs {base.count(old)} times"
        assert old != new, name
    with tempfile.TemporaryDirectory() as tmp:
        def run(case):
            name, old, new = case
            path = os.path.join(tmp, f"{name}.so")
            _build(base.replace(old, new), path)
            return name, _pins_pass(path, tests)

        # the unmutated source built the same way must pass (else the harness proves nothing)
        path = os.path.join(tmp, "control.so")
        _build(base, path)
        ok, out = _pins_pass(path, tests)
        assert ok, "unmutated oracle fails its pins:\n" + out
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            results = list(ex.map(run, mutants))
    survivors = [name for name, (passed, _) in results if passed]
    assert not survivors, f"planted mistakes that no pin catches: {survivors}"


def test_every_planted_mistake_fails_a_pin():
    _check(MUTANTS, "test_oracle.py")


def test_every_planted_migration_mistake_fails_a_pin():
    _check(MIG_MUTANTS, "test_migration_oracle.py")
