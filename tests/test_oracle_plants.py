"""The oracle's pins catch the mistakes they are meant to catch.

Each case copies the oracle's sources, plants one plausible slip in the copy
(a wrong stage index in the Runge-Kutta flux combination, a flipped 1/48 term,
a wrong speed or force factor, a wrong tableau weight or stage time, a
transposed or mis-layered coarse-to-fine fill), builds
it, and runs the pins named for it against the planted build: at least one of
them must fail.  The oracle in oracle/ is not touched."""
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

import _oracle
import test_clustering as K
import test_oracle_amr_fill as F
import test_oracle_vlasov as V

ORACLE = os.path.join(_oracle.ROOT, "oracle")
SOURCES = ("oracle.cpp", "oracle_amr.cpp", "oracle_capi.cpp", "oracle.hpp", "oracle_consts.h")

# name -> (file, [(text, planted text)], pins of which at least one must fail)
PLANTS = {
    # Rectangle.cpp:1396-1517: sum_k a_k FxH_k with the previous stage's flux
    "rk_stage_index": ("oracle.cpp",
                       [("sx = sx + a[k] * FxH[k][idx];", "sx = sx + a[k] * FxH[k - 1][idx];")],
                       [V.test_free_streaming_pulse_converges_at_third_order]),
    # the same slip in the p-flux combination
    "rk_stage_index_p": ("oracle.cpp",
                         [("sp_ = sp_ + a[k] * FpH[k][idx];", "sp_ = sp_ + a[k] * FpH[k - 1][idx];")],
                         [V.test_time_dependent_force_converges_at_the_order_of_the_tableau]),
    # Rectangle.cpp:1330-1345 and 1366-1381: the product term's sign, both axes
    "w48_sign": ("oracle.cpp",
                 [("FxH[step][NS(i, j)] = dx_inv * (fm * am + w3",
                   "FxH[step][NS(i, j)] = dx_inv * (fm * am - w3"),
                  ("FpH[step][NS(i, j)] = dp_inv * (fm * am + w3",
                   "FpH[step][NS(i, j)] = dp_inv * (fm * am - w3")],
                 [V.test_linear_profile_is_exact_in_time_and_fourth_order_in_dp,
                  V.test_free_streaming_pulse_converges_at_third_order]),
    # Rectangle.cpp:1278-1286: the cell-averaged x speed 1 % off
    "ex_factor": ("oracle.cpp",
                  [("ex[NS(i, j)] = dp_inv * cc *", "ex[NS(i, j)] = 1.01 * dp_inv * cc *")],
                  [V.test_free_streaming_pulse_converges_at_third_order,
                   V.test_linear_profile_is_exact_in_time_and_fourth_order_in_dp]),
    # Rectangle.cpp:1294-1305: the electric force with the wrong sign
    "force_sign": ("oracle.cpp",
                   [("double Em = q * GetEfield(i);", "double Em = -q * GetEfield(i);")],
                   [V.test_uniform_force_shifts_p_at_the_exact_rate,
                    V.test_time_dependent_force_converges_at_the_order_of_the_tableau]),
    # one weight of the tableau 0.1 % off (the weights then sum to 1.0007)
    "rk_weight": ("oracle_consts.h",
                  [("0.6805652953093346, -0.27524053099500667",
                    "0.6805652953093346 * 1.001, -0.27524053099500667")],
                  [V.test_time_dependent_force_converges_at_the_order_of_the_tableau]),
    # Mesh.cpp:480-490, 541-551: ties between equal inflections resolved toward
    # the later one (both axes)
    "inflection_tie": ("oracle_amr.cpp",
                       [("if (range > std::abs(inflectIdx.at(i) - center)) {",
                         "if (range >= std::abs(inflectIdx.at(i) - center)) {")],
                       [K.check_oracle_cases]),
    # Mesh.cpp:722-741: the bisection rounded up instead of down
    "bisect_round": ("oracle_amr.cpp",
                     [("cutPosition = std::floor((lengthx / 2) + rectangle.first.first);",
                       "cutPosition = std::ceil((lengthx / 2) + rectangle.first.first);")],
                     [K.check_oracle_cases]),
    # Rectangle.cpp:343-415: the r x r interpolants stored transposed
    "fill_transpose": ("oracle_amr.cpp",
                       [("interpolants[k * r + l] = part[l];", "interpolants[l * r + k] = part[l];")],
                       [F.test_ghost_fill_is_exact_for_bilinear_data]),
    # Rectangle.cpp:389-395: the ghost layer next to a lower p edge given the
    # sub-cell row of the layer behind it
    "fill_layer": ("oracle_amr.cpp",
                   [("red[2 * k] = interpolants[r * k + r - 1];", "red[2 * k] = interpolants[r * k + r - 2];")],
                   [F.test_ghost_fill_is_exact_for_bilinear_data]),
    # Rectangle.cpp:385-387: the mean correction taken from the next coarse cell
    "fill_correction_cell": ("oracle_amr.cpp",
                             [("double correction = g[NS(i_coarse, j_coarse)] - 1.0",
                               "double correction = g[NS(i_coarse, j_coarse + 1)] - 1.0")],
                             [F.test_interpolated_blocks_average_to_their_coarse_cell]),
    # Rectangle.cpp:76-89: one constant of the interpolation table's linear row
    # 1 % off (its quadratic row has no effect at r = 2: the sub-cell weights
    # of that term are zero on both halves)
    "fill_table": ("oracle.hpp",
                   [("0.104166666666667, -0.708333333333334,", "0.105208333333334, -0.708333333333334,")],
                   [F.test_ghost_fill_is_exact_for_bilinear_data]),
    # Rectangle.cpp:904-912: the transfer's sub-cells transposed
    "transfer_transpose": ("oracle_amr.cpp",
                           [("rect->f[0][rect->NS(fi, fj)] = in[l + r * k];",
                             "rect->f[0][rect->NS(fi, fj)] = in[k + r * l];")],
                           [F.test_regrid_transfer_is_exact_for_bilinear_data]),
    # Rectangle.cpp:476-517: a same-level corner block's outer ghost column
    # copied from the inner one
    "corner_same_level": ("oracle_amr.cpp",
                          [("g[P.NS(-2, i * r + j)] = same(s, P.x_pos - 2, P.p_pos + i * r + j);",
                            "g[P.NS(-2, i * r + j)] = same(s, P.x_pos - 1, P.p_pos + i * r + j);")],
                          [F.test_ghost_fill_is_exact_for_bilinear_data]),
    # Settings.cpp:166-179: the third stage time 0.333 instead of 0.332
    "stage_clock": ("oracle.cpp",
                    [("time += (0.332 - 0.5) * dt;", "time += (0.333 - 0.5) * dt;"),
                     ("time += (0.62 - 0.332) * dt;", "time += (0.62 - 0.333) * dt;")],
                    [V.test_time_dependent_force_converges_at_the_order_of_the_tableau]),
}


def _build(name, workdir):
    fname, subs, _ = PLANTS[name]
    d = os.path.join(workdir, name)
    os.makedirs(d)
    for src in SOURCES:
        with open(os.path.join(ORACLE, src)) as fh:
            text = fh.read()
        if src == fname:
            for old, new in subs:
                assert text.count(old) >= 1, (name, old)
                text = text.replace(old, new)
        with open(os.path.join(d, src), "w") as fh:
            fh.write(text)
    out = os.path.join(d, "libplanted.so")
    # -I oracle/ resolves the oracle's "../veritas_amd/..." includes
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                           "-shared", "-I", ORACLE, *[os.path.join(d, s) for s in SOURCES[:3]],
                           "-o", out])
    return out


@pytest.fixture(scope="module")
def planted_builds():
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
            paths = dict(zip(PLANTS, ex.map(lambda n: _build(n, tmp), PLANTS)))
        yield paths


def _fails(pin):
    try:
        pin()
    except (AssertionError, FloatingPointError, ValueError):
        return True
    return False


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(name, planted_builds):
    pins = PLANTS[name][2]
    with _oracle.library(planted_builds[name]):
        failed = [p.__name__ for p in pins if _fails(p)]
    assert failed, f"no pin caught the planted {name}"


def test_unplanted_oracle_passes_the_same_pins():
    for pin in {p for _, _, pins in PLANTS.values() for p in pins}:
        pin()
