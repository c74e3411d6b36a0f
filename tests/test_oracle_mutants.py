"""The pins of the CPU oracle have teeth: each test plants one plausible
mistake in a copy of oracle/pfm_oracle.c (a dropped term, a flipped sign, a
wrong index or factor), builds it with the oracle's own flags, points the
helpers at the copy and asserts that the pin guarding that code now fails.

A mutant "fails" a pin when the pin raises AssertionError, or when the
mutated oracle itself stops with an error (PfmError) on the pin's input."""
import os
import subprocess

import pytest

from pfmds_amd import capi
from tests import harness
from tests import test_oracle_growth as growth
from tests import test_oracle_guards as guards
from tests import test_oracle_pins as pins

ORACLE_DIR = os.path.join(harness.ROOT, "oracle")
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-Wall",
          "-Wno-unknown-pragmas"]

# name: (text in pfm_oracle.c, replacement, pin, pin arguments[, which
# occurrence when the text appears more than once])
MUTANTS = {
    "min_image_sign": (
        "d = d - hb[k] * (copysign(1.0, d - hb[k]) + copysign(1.0, d + hb[k]));",
        "d = d + hb[k] * (copysign(1.0, d - hb[k]) + copysign(1.0, d + hb[k]));",
        pins.test_lists_brute_force, ("cu",)),
    "list_cut_unsquared": (
        "double rc2 = nl->rcut * nl->rcut;", "double rc2 = nl->rcut;",
        pins.test_lists_brute_force, ("lj_gas",)),
    "lessnnum_wrong_order": (
        "if (less == -1 && i < j) less = nn;", "if (less == -1 && i > j) less = nn;",
        pins.test_lists_brute_force, ("lj_gas",)),
    "lj_cutoff_derivative_dropped": (
        "- V * (V - 1.0) * dfcut(r, R1, R2));", ");",
        pins.test_forces_lj_two_groups, ()),
    "lj1g_quirk_corrected": (
        "*dfr = t * x2 * (-30.0 + 60.0 * x - 30.0 * x2);",
        "*dfr = t * x2 * (-30.0 + 60.0 * x - 30.0 * x2) / (R2 - R1);",
        pins.test_forces_lj1g_cutoff_quirk, ()),
    "lj1g_energy_cut_power": (
        "return 1.0 + (-10.0 * a * a * a * w * w",
        "return 1.0 + (-10.0 * a * a * w * w * w",
        pins.test_lj_dimer_energy, ()),
    "rjl_density_of_one_atom": (
        "(1.0 / Eb[i] + 1.0 / Eb[nl->nlist[e]])", "(2.0 / Eb[i])",
        pins.test_forces_rjl, ()),
    "rjl_energy_exponent": (
        "Eb2 = Eb2 + exp(-2.0 * q * (r / r0 - 1.0))", "Eb2 = Eb2 + exp(-q * (r / r0 - 1.0))",
        pins.test_rjl_perfect_fcc_closed_form, ()),
    "tb_bond_order_counts_itself": (
        "if (p != q && nl->moddr[g] < R2) {", "if (nl->moddr[g] < R2) {",
        pins.test_tb_perfect_graphene_closed_form, ()),
    "tb_neighbour_angle_sign": (
        "c1 * (-drq[k] * rr + cosi * drl[k] / (rl * rl));", "c1 * (drq[k] * rr + cosi * drl[k] / (rl * rl));",
        pins.test_forces_tb, ()),
    "tb_cutoff_derivative_dropped": (
        "double c2 = dfcut(rq, R1, R2) * a0 * (1.0 + c02 / d02 - c02 / t);", "double c2 = 0.0;",
        pins.test_forces_tb_switching_zone, ()),
    "morsec_normal_term_dropped_graphene_side": (
        "b = (2.0 * c.delt * V2 * V3 / nd * fc);", "b = 0.0;",
        pins.test_forces_graphene_metal, ("morsec", True), (0, 2)),
    "morsec_normal_term_dropped_metal_side": (
        "b = (2.0 * c.delt * V2 * V3 / nd * fc);", "b = 0.0;",
        pins.test_forces_graphene_metal, ("morsec", True), (1, 2)),
    "ljc_radial_exponent_graphene_side": (
        "(6.0 + c.delt) * V2 * V3", "(6.0 - c.delt) * V2 * V3",
        pins.test_forces_graphene_metal, ("ljc", True), (0, 2)),
    "ljc_normal_term_metal_side": (
        "b = (c.delt * V2 * V3 / nd * fc);", "b = (V2 * V3 / nd * fc);",
        pins.test_forces_graphene_metal, ("ljc", True), (1, 2)),
    "normal_derivative_sign": (
        "f[k] = f[k] - coef * nj[k] * w;", "f[k] = f[k] + coef * nj[k] * w;",
        pins.test_forces_graphene_metal, ("morsec", False)),
    "normal_not_oriented": (
        "if (g[2] < 0.0) { g[0] = -g[0];", "if (0) { g[0] = -g[0];",
        pins.test_forces_graphene_metal, ("ljc", False)),
    "kick_full_step": (
        "h->force[3 * i + k] / h->mass[i] / MASS_COEF * h->ts[1];",
        "h->force[3 * i + k] / h->mass[i] / MASS_COEF * h->ts[0];",
        pins.test_nve_time_reversible_and_conserving, ()),
    "nhc_scale_sign": (
        "c->s = exp(-v[0] * h->ts[1]);", "c->s = exp(v[0] * h->ts[1]);",
        pins.test_nhc_conserved_quantity, (3,)),
    "nhc_energy_degrees_of_freedom": (
        "+ 3.0 * c->L * kt * c->x[0];", "+ c->L * kt * c->x[0];",
        pins.test_nhc_conserved_quantity, (1,)),
    "nvms_keeps_uphill_velocity": (
        "if (fv > 0.0 && ff > 1e-12) {", "if (ff > 1e-12) {",
        pins.test_nvms_velocity_rule, ()),
    "nvms_z_atoms_keep_uphill_velocity": (
        "if (!(fv > 0.0)) v[0] = v[1] = v[2] = 0.0;", "if (!(fv > -1.0)) v[0] = v[1] = v[2] = 0.0;",
        pins.test_nvms_velocity_rule, ()),
    "nvms_projection_unnormalised": (
        "v[k] = fv / ff * f[k];", "v[k] = fv * f[k];",
        pins.test_nvms_velocity_rule, ()),
    "rebo_dihedral_sign": (
        "bdh[e] = T * bdh[e];", "bdh[e] = -T * bdh[e];",
        pins.test_forces_rebo_buckled, ()),
    "rebo_bsp_exponent": (
        "bsp[e] = pow(1. + bsp[e], -0.5);", "bsp[e] = pow(1. + bsp[e], -1.0);",
        pins.test_rebo_perfect_graphene_closed_form, ()),
    "rebo_truncated_two_hops": (
        "truncated_nl(&t, nl, inl, 3);", "truncated_nl(&t, nl, inl, 2);",
        pins.test_forces_rebo_buckled, ()),
    "rebo_partner_bond_order_dropped": (
        "double bq = q < nl->nmax ? bsp[(size_t)j * nl->nmax + q] : 0.0;", "double bq = bsp[e];",
        pins.test_forces_rebo_buckled, ()),
    "nvms_signed_energy_change": (
        "fabs(h->pe - h->pe_prev) <= ms_de", "(h->pe - h->pe_prev) <= ms_de",
        pins.test_nvms_lj_dimer_minimum, ()),
    "group_change_period_from_zero": (
        "(step - c->ts1) % c->frec == 0", "step % c->frec == 0",
        growth.test_growth_counts_follow_the_rule, ()),
    "check_positions_tolerance_too_wide": (
        "const double tol = 0.0000001;", "const double tol = 0.000001;",
        guards.test_check_positions_tolerance, (0, "L+1.1e-7", True)),
    "drift_wraps_until_inside": (
        "if (x > h->box[k]) x = x - h->box[k];", "if (x > h->box[k]) x = fmod(x, h->box[k]);",
        guards.test_check_positions_after_overshooting_drift, ()),
    "invert_z_lower_half_direction": (
        "z < mid && vz > 0.0", "z < mid && vz < 0.0",
        guards.test_invert_z_is_a_reflecting_wall, ()),
    "invert_z_upper_border": (
        "hi = 0.9 * h->box[2]", "hi = 0.95 * h->box[2]",
        guards.test_invert_z_is_a_reflecting_wall, ()),
    "zero_momentum_unweighted": (
        "mcv[k] += h->mass[i] * h->vel[3 * i + k];", "mcv[k] += h->vel[3 * i + k];",
        guards.test_zero_momentum, ()),
    "zero_momentum_every_atom": (
        "for (int k = 0; k < 3; ++k) h->vel[3 * g->idx[ind] + k] -= mcv[k];",
        "for (int k = 0; k < 3; ++k) for (int a = 0; a < h->N; ++a) if (a == g->idx[ind] || a % 5 == 0) "
        "h->vel[3 * a + k] -= (a == g->idx[ind]) ? mcv[k] : 0.0 * mcv[k] + 1e-3;",
        guards.test_zero_momentum, ()),
    "nhc_first_kick_eighth_step": (
        "v[0] = v[0] + kedif / q[0] * h->ts[2];", "v[0] = v[0] + kedif / q[0] * h->ts[3];",
        guards.test_nhc_step_sequence, (1,), (0, 2)),
    "nhc_second_half_old_kinetic_energy": (
        "kedif = 2.0 * ke * (c->s * c->s) - 3.0 * c->L * kt;", "kedif = 2.0 * ke - 3.0 * c->L * kt;",
        guards.test_nhc_step_sequence, (2,)),
    "nhc_chain_force_unscaled": (
        "/ q[i] * h->ts[2] * b;", "/ q[i] * h->ts[2];",
        guards.test_nhc_step_sequence, (3,), (0, 2)),
    "z_kick_moves_every_component": (
        "for (int k = only_z ? 2 : 0; k < 3; ++k)\n", "for (int k = 0; k < 3; ++k)\n",
        guards.test_z_only_integration, ()),
    "group_change_rows_over_all_members": (
        "for (int ind = 0; ind < g1->n; ++ind) {", "for (int ind = 0; ind < g1->size; ++ind) {",
        growth.test_growth_moves_and_lists_only_joined_atoms, ()),
}


def build_mutant(tmp_path, old, new, site=(0, 1)):
    """Replace occurrence site[0] of `old` (which must occur site[1] times)."""
    src = open(os.path.join(ORACLE_DIR, "pfm_oracle.c")).read()
    nth, count = site
    assert src.count(old) == count, f"{old!r} occurs {src.count(old)} times, expected {count}"
    at = -1
    for _ in range(nth + 1):
        at = src.index(old, at + 1)
    path = tmp_path / "pfm_oracle_mutant.c"
    path.write_text(src[:at] + new + src[at + len(old):])
    so = tmp_path / "liborc_mutant.so"
    # -I oracle/: "pfm_oracle.h" and "../include/pfm.h" resolve as for the oracle
    subprocess.run(["gcc", *CFLAGS, "-I", ORACLE_DIR, "-shared", "-o", str(so), str(path), "-lm"], check=True,
                   capture_output=True, timeout=120)
    return str(so)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pin_fails_on_mutant(name, tmp_path, monkeypatch):
    old, new, pin, args, *site = MUTANTS[name]
    so = build_mutant(tmp_path, old, new, *site)
    monkeypatch.setattr(harness, "ORACLE_LIB", so)
    with pytest.raises((AssertionError, capi.PfmError)):
        pin(*args)


def test_unmutated_copy_passes(tmp_path, monkeypatch):
    """The same build route with no mutation passes a pin, so a failure above
    comes from the planted mistake and not from the copy or the flags."""
    so = build_mutant(tmp_path, "static inline double pow6", "static inline double pow6")
    monkeypatch.setattr(harness, "ORACLE_LIB", so)
    pins.test_forces_rjl()
