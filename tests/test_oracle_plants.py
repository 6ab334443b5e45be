"""Planted mistakes in the oracle fail the closed-form tests.

Each case copies oracle/hc_oracle.c, replaces one expression with a
plausible mistake (a dropped factor, a wrong constant, swapped image axes, a
wrong emission law), builds the copy with the oracle's own flags, and runs
the closed-form tests of tests/test_oracle_closed_forms.py that must catch
it against that build: at least one of them has to fail."""
import os
import subprocess

import pytest

import test_atmosphere as at
import test_filters as fi
import test_oracle_diffusion as dif
import test_oracle_closed_forms as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "hc_oracle.c")
CFLAGS = ["-O1", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
          "-w"]

# (name, original text, planted text, tests that must catch it)
PLANTS = [
    ("tauint drops (1 - albedo)",
     "(1.0 - P->op.albedo[id]) * g->densarr[k] * dt", "g->densarr[k] * dt",
     ["radial"]),
    ("jmean without the stellar radius",
     "t->jmean[cell] += dt * c->rmincgs;", "t->jmean[cell] += dt;",
     ["radial", "oblique"]),
    ("interaction point overshoots",
     "double ds = (tau - tsum) / opac;", "double ds = 1.01 * (tau - tsum) / opac;",
     ["interact"]),
    ("peel weight exp(-tau/2)", "factor = HCO_EXP(-tsum);", "factor = HCO_EXP(-0.5 * tsum);",
     ["peel"]),
    ("walk ignores the sublimation mask",
     "if (c->g->tdust[k] < c->tsub_ev) opac", "if (1) opac", ["mask"]),
    ("image axes swapped",
     "ximage = yp1 * cospe - xp1 * sinpe;", "ximage = xp1 * cospe - yp1 * sinpe;", ["image"]),
    ("direct peel weight 2 cos", ": 4.0 * P->sip * cosnorm;", ": 2.0 * P->sip * cosnorm;",
     ["direct"]),
    ("cos(theta) bin of 1 - cos^2",
     "it = (int)((double)s->nmu * (1.0 - P->cost) / 2.0) + 1;",
     "it = (int)((double)s->nmu * (1.0 - P->cost * fabs(P->cost)) / 2.0) + 1;", ["bins"]),
    ("emission mu = xi instead of sqrt(xi)", ": sqrt(xran);", ": xran;",
     ["dilution", "aperture"]),
    ("tfinal with 2 sigma T^4", "(na * ltot) / (4.0 * SIGT", "(na * ltot) / (2.0 * SIGT",
     ["dilution"]),
    ("star re-emission at half weight",
     "if (lp < 0.0) lp = lp + R2P;\n  P->sip = 1.0;\n  P->sqp = P->sup = P->svp = 0.0;\n"
     "  initp_sample(c, P);\n  initp_transform(P, cosb, sinb, lp);\n  P->nub = star_freq(c);",
     "if (lp < 0.0) lp = lp + R2P;\n  P->sip = 0.5;\n  P->sqp = P->sup = P->svp = 0.0;\n"
     "  initp_sample(c, P);\n  initp_transform(P, cosb, sinb, lp);\n  P->nub = star_freq(c);",
     ["reemit"]),
    ("atmosfreq interpolates from the wrong node",
     "(HCO_LOG10(xi) - s->loghnu[i - 1])", "(HCO_LOG10(xi) - s->loghnu[i])", ["atmos"]),
    ("darkening steps two bins", "(s->prob[n] - s->prob[n - 1]) * s->delmu;",
     "(s->prob[n] - s->prob[n - 1]) * 2.0 * s->delmu;", ["limb"]),
    ("limb peel weight 2 (mu + mu^2)", "? 12.0 / 5.0 * P->sip", "? 2.0 * P->sip",
     ["limbpeel"]),
    ("filter weight from the node above",
     "iw = iw - 1;\n        const double wt = s->filtphi[(size_t)s->nwfiltmax * ifilt + (size_t)(iw - 1)];",
     "const double wt = s->filtphi[(size_t)s->nwfiltmax * ifilt + (size_t)(iw - 1)];",
     ["filter"]),
    ("filter range closed at the top", "if (freq > fw[0] && freq < fw[nf - 1]) {",
     "if (freq > fw[0] && freq <= fw[nf - 1]) {", ["filteredge"]),
    ("isotrp without the square", "while (r2 * r2 > fabs(mur)) {",
     "while (r2 > fabs(mur)) {", ["isotrp"]),
    ("radial surface ring area without 2",
     "dA = 2.0 * PI * (ra[ir - 1] * ra[ir - 1])", "dA = PI * (ra[ir - 1] * ra[ir - 1])",
     ["difradial"]),
    ("theta surface move onto the cone itself", "cosbnew = 1.00001 * cost_wall(g, it - 1);",
     "cosbnew = cost_wall(g, it - 1);", ["diftheta"]),
    ("column density of the next cell",
     "const size_t km = cell_index(g, i - 1, it, ip) + c->ncell * id;",
     "const size_t km = cell_index(g, i, it, ip) + c->ncell * id;", ["difcolumn"]),
    ("filter range closed at the bottom", "if (freq > fw[0] && freq < fw[nf - 1]) {",
     "if (freq >= fw[0] && freq < fw[nf - 1]) {", ["filteredge"]),
    ("filter Q unweighted", "t->image_b_q[q] += sqp * wt;", "t->image_b_q[q] += sqp;",
     ["filter"]),
]


def _checks():
    def transparent(fn):
        return lambda: fn(cf.run_transparent_pass())
    return {
        "radial": [cf.test_radial_ray_deposits_kappa_abs_rho_times_shell_thickness],
        "oblique": [lambda: cf.test_oblique_ray_shell_chords_and_cell_split(0.35, 0.7, 0.4)],
        "interact": [cf.test_interaction_point_at_tau_over_opacity],
        "peel": [lambda: cf.test_peel_weight_is_exp_minus_kappa_rho_exit_distance(2)],
        "mask": [cf.test_peel_skips_sublimated_cells_star_and_tau_above_25],
        "image": [cf.test_image_pixel_of_a_known_point_is_its_sky_projection],
        "direct": [transparent(cf.test_direct_stellar_peel_weight_is_four_cos)],
        "bins": [transparent(cf.test_direct_packets_fill_cos_theta_and_phi_bins_uniformly)],
        "aperture": [transparent(cf.test_aperture_holds_the_disk_flux_inside_its_radius)],
        "dilution": [cf.test_optically_thin_grey_shell_reaches_the_dilution_temperature],
        "reemit": [cf.test_scattering_shell_returns_every_packet_and_the_stellar_spectrum],
        "atmos": [lambda: at.test_atmosfreq_maps_nodes_to_nodes_and_is_monotone(at.ORACLE),
                  at.test_atmosfreq_fills_brackets_with_their_probability_and_log_law],
        "limb": [at.test_limb_darkening_cosine_law],
        "limbpeel": [at.test_limb_darkened_disk_flux_and_peel_weight],
        "isotrp": [dif.test_isotrp_is_lambertian_about_the_normal],
        "difradial": [lambda: dif.test_radial_surface_move_direction_and_temperatures(
            dif.radial_model())],
        "diftheta": [lambda: dif.test_theta_surfaces_up_and_down(dif.thick_model())],
        "difcolumn": [lambda: dif.test_column_sums_with_absorbed_counts(dif.radial_model())],
        "filter": [fi.test_filter_image_pixel_weight_and_stokes],
        "filteredge": [fi.test_filter_image_outside_every_node_range_is_empty],
    }


@pytest.mark.parametrize("name,orig,planted,checks", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_mistake_is_caught(tmp_path, name, orig, planted, checks):
    src = open(SRC).read()
    assert src.count(orig) >= 1, f"{name}: expression not found in the oracle"
    c = tmp_path / "hc_oracle.c"
    c.write_text(src.replace(orig, planted))
    so = tmp_path / "libplant.so"
    subprocess.run(["gcc"] + CFLAGS + ["-I", os.path.join(ROOT, "oracle"), "-o", str(so), str(c),
                                       "-lm"], check=True)
    table = _checks()
    old = cf.ORACLE
    cf.ORACLE = at.ORACLE = fi.ORACLE = dif.ORACLE = str(so)
    caught = []
    try:
        for key in checks:
            for fn in table[key]:
                try:
                    fn()
                except AssertionError:
                    caught.append(key)
    finally:
        cf.ORACLE = at.ORACLE = fi.ORACLE = dif.ORACLE = old
    assert caught, f"planted mistake '{name}' passed every check {checks}"
