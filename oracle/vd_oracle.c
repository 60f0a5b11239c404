/*
 * vd_oracle.c -- CPU FP64 restatement of Vidyut's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  The product (vidyut3d_amd/) never links or loads it.
 *
 * Every function cites the reference file:line it restates.  Where the
 * reference delegates to AMReX (absent from the reference tree, version
 * unpinned), the convention adopted here is stated in the comment; the
 * HIP library implements the same conventions.
 *
 * Single level, single box per level (box == domain), Array4 layout.
 */
#include "vd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Source/UnivConstants.H:4-19 */
#define N_A 6.02214076e23
#define K_B 1.380649e-23
#define PI 3.14159265358979323846264338327950288
#define ECHARGE 1.60217662e-19
#define EPS0 8.854187817e-12
#define EV (ECHARGE / K_B)
#define ME 9.10938188e-31
#define M_AMU 1.66054e-27
#define FIVEBYTHREE (5.0 / 3.0)
#define TWOTHIRD (2.0 / 3.0)

static inline double* a4p(const vd_array4* a, int i, int j, int k, int n) {
    return a->p + (i - a->begin[0]) + (int64_t)(j - a->begin[1]) * a->jstride +
           (int64_t)(k - a->begin[2]) * a->kstride + (int64_t)n * a->nstride;
}
#define AT(a, i, j, k, n) (*a4p((a), (i), (j), (k), (n)))
#define ATV(a, c, n) (*a4p((a), (c)[0], (c)[1], (c)[2], (n)))

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int box_is_domain(const vd_box* b, const vd_geom* g) {
    for (int d = 0; d < 3; d++) {
        int lo = d < g->dim ? g->dom_lo[d] : 0, hi = d < g->dim ? g->dom_hi[d] : 0;
        if (b->lo[d] != lo || b->hi[d] != hi) return 0;
    }
    return 1;
}

/* Source/Vidyut.H:149-272 defaults */
void or_params_default(vd_params* p) {
    memset(p, 0, sizeof(*p));
    p->gas_temperature = 300.0;
    p->gas_pressure = 101325.0;
    p->hyp_order = 1;
    p->weno_scheme = 3;
    p->elecenergy_solve = 0;
    p->do_reactions = 1;
    p->do_transport = 1;
    p->do_spacechrg = 1;
    p->do_bg_reactions = 0;
    p->bound_specden = 0;
    p->min_species_density = 1e-20;
    p->min_electron_density = 1e12;
    p->min_electron_temp = 300.0;
    p->floor_jh = 0;
    p->const_ele_trans = 0;
    p->ele_mob = -1.0e24;
    p->ele_diff = 1.0e24;
    p->voltage_profile = 0;
    p->voltage_amp_1 = 0.0;
    p->voltage_amp_2 = 0.0;
    p->voltage_freq = 1.0;
    p->voltage_dur = 0.0;
    p->voltage_center = 0.0;
    p->linsolve_reltol = 1e-10;
    p->linsolve_abstol = 1e-10;
    p->linsolve_maxiter = 100;
    p->linsolve_max_coarsening_level = 10;
    p->mg_pre_smooth = 2;
    p->mg_post_smooth = 2;
    p->mg_precision = 0;
    p->adaptive_dt = 0;
    p->advective_cfl = 0.7;
    p->diffusive_cfl = 1.0;
    p->dielectric_cfl = 1.0;
    p->dt_min = 1e-14;
    p->dt_max = 1e-9;
    p->adaptive_dt_delay = 0.0;
    p->dt_stretch = DBL_MAX;
    for (int d = 0; d < 3; d++) {
        p->pot_bc_lo[d] = p->pot_bc_hi[d] = VD_ROBINBC;
        p->eden_bc_lo[d] = p->eden_bc_hi[d] = VD_ROBINBC;
        p->eenrg_bc_lo[d] = p->eenrg_bc_hi[d] = VD_ROBINBC;
        p->ion_bc_lo[d] = p->ion_bc_hi[d] = VD_ROBINBC;
        p->neutral_bc_lo[d] = p->neutral_bc_hi[d] = VD_ROBINBC;
    }
}

/* ------------------------------------------------------------ ghost fill
 * Vidyut::FillPatch level 0 (Source/GridWork.cpp:112-124):
 * FillPatchSingleLevel with bcspec int_dir (periodic) or foextrap
 * (Source/Vidyut.cpp:93-104).  foextrap copies the nearest valid cell;
 * corner ghosts take the cell found by clamping/wrapping every index. */
int or_fill_patch(const vd_array4* src, vd_array4* dst, const vd_box* box, const vd_geom* g,
                  int scomp, int ncomp) {
    if (!box_is_domain(box, g)) return VD_EINVAL;
    int n[3];
    for (int d = 0; d < 3; d++) n[d] = box->hi[d] - box->lo[d] + 1;
#pragma omp parallel for collapse(2)
    for (int k = dst->begin[2]; k < dst->end[2]; k++)
        for (int j = dst->begin[1]; j < dst->end[1]; j++)
            for (int i = dst->begin[0]; i < dst->end[0]; i++) {
                int c[3] = {i, j, k};
                for (int d = 0; d < g->dim; d++) {
                    int lo = box->lo[d], hi = box->hi[d];
                    if (c[d] < lo || c[d] > hi) {
                        if (g->is_periodic[d]) {
                            int r = (c[d] - lo) % n[d];
                            if (r < 0) r += n[d];
                            c[d] = lo + r;
                        } else {
                            c[d] = c[d] < lo ? lo : hi;
                        }
                    }
                }
                for (int q = scomp; q < scomp + ncomp; q++) AT(dst, i, j, k, q) = ATV(src, c, q);
            }
    return VD_OK;
}

/* ------------------------------------------------------------ transport
 * specMob / specDiff of the case plug-in (template: test/models/
 * Axisymmetric_streamer_Ar/Chemistry.H:180-240 with constant mobilities):
 * index nspec is the electron energy with 5/3 of the electron values. */
static double spec_mob(const vd_mech_desc* m, int s) {
    if (s == m->nspec) return FIVEBYTHREE * m->mob[m->e_idx];
    return m->mob[s];
}
static double spec_diff(const vd_mech_desc* m, int s, double Te, double Tg) {
    if (s == m->nspec) {
        double Te_in_ev = Te / EV;
        return FIVEBYTHREE * Te_in_ev * fabs(m->mob[m->e_idx]);
    }
    switch (m->diff_kind[s]) {
        case VD_DIFF_EINSTEIN_TE: return (Te / EV) * fabs(m->mob[s]);
        case VD_DIFF_EINSTEIN_TG: return (Tg / EV) * fabs(m->mob[s]);
        default: return m->diff[s];
    }
}

/* a vd_transport_fit coefficient (the state-dependent specMob / specDiff
 * of a case Chemistry.H, e.g. test/verification/Streamer/Chemistry.H:
 * 165-200, test/verification/He_RF_1d/Chemistry.H specMob) */
static double tr_eval(const vd_transport_fit* f, double Te, double Tg, double nd, double emag) {
    const double* c = f->c;
    double l = c[0];
    if (c[1] != 0.0) l += c[1] * log(emag);
    if (c[2] != 0.0) l += c[2] * log(Te);
    if (c[3] != 0.0) l += c[3] * log(Tg);
    if (c[4] != 0.0) l += c[4] * log(nd);
    if (c[5] != 0.0 || c[6] != 0.0 || c[7] != 0.0 || c[8] != 0.0) {
        double e = 1.5 * Te / EV;
        if (c[5] != 0.0) l += c[5] * log(e);
        if (c[6] != 0.0) l += c[6] / e;
        if (c[7] != 0.0) l += c[7] / (e * e);
        if (c[8] != 0.0) l += c[8] / (e * e * e);
    }
    if (c[9] != 0.0) {
        double x = emag / nd;
        l += c[9] * log(1.0 + c[10] * (x * x) + c[11] * (x * x * x * x));
    }
    return f->sign * exp(l);
}

/* specMob(specid, Te, ndens, |E|, Tg) / specDiff of the case plug-in
 * (index nspec = the electron energy) */
static double spec_mob_st(const vd_mech_desc* m, int s, double Te, double Tg, double nd, double emag) {
    const vd_transport_fit* f = &m->mob_fit[s];
    if (f->kind == VD_TR_FIT) return tr_eval(f, Te, Tg, nd, emag);
    if (s == m->nspec) {
        const vd_transport_fit* fe = &m->mob_fit[m->e_idx];
        if (fe->kind == VD_TR_FIT) return FIVEBYTHREE * tr_eval(fe, Te, Tg, nd, emag);
    }
    return spec_mob(m, s);
}
static double spec_diff_st(const vd_mech_desc* m, int s, double Te, double Tg, double nd, double emag) {
    const vd_transport_fit* f = &m->diff_fit[s];
    if (f->kind == VD_TR_FIT) return tr_eval(f, Te, Tg, nd, emag);
    const int e = m->e_idx;
    if (s == m->nspec) {
        if (m->diff_fit[e].kind == VD_TR_FIT) return FIVEBYTHREE * tr_eval(&m->diff_fit[e], Te, Tg, nd, emag);
        if (m->mob_fit[e].kind == VD_TR_FIT)
            return FIVEBYTHREE * (Te / EV) * fabs(tr_eval(&m->mob_fit[e], Te, Tg, nd, emag));
        return spec_diff(m, s, Te, Tg);
    }
    if (m->mob_fit[s].kind == VD_TR_FIT && m->diff_kind[s] != VD_DIFF_CONST) {
        double mu = fabs(tr_eval(&m->mob_fit[s], Te, Tg, nd, emag));
        return (m->diff_kind[s] == VD_DIFF_EINSTEIN_TE ? Te / EV : Tg / EV) * mu;
    }
    return spec_diff(m, s, Te, Tg);
}

/* ndens = sum of the species densities and |E| = sqrt(sum_d E_d^2) at one
 * cell (the arguments every specMob / specDiff call site forms) */
static void cell_nd_emag(const vd_array4* phi, const int c[3], int ns, int dim, double* nd, double* emag) {
    double n = 0.0, es = 0.0;
    for (int sp = 0; sp < ns; sp++) n += ATV(phi, c, sp);
    for (int d = 0; d < dim; d++) es += pow(ATV(phi, c, VD_EFX(ns) + d), 2.0);
    *nd = n;
    *emag = sqrt(es);
}

/* plasmachem_transport::compute_vel (Source/PlasmaChem.H:186-206) */
static double compute_vel(const vd_mech_desc* m, const vd_array4* phi, const int c[3], int dir,
                          int specid, double Tg) {
    int ns = m->nspec;
    /* user_transport::compute_vel: a prescribed velocity */
    if (m->vel_set[specid]) return m->vel[specid][dir];
    double efieldmag = sqrt(pow(ATV(phi, c, VD_EFX(ns)), 2.0) + pow(ATV(phi, c, VD_EFY(ns)), 2.0) +
                            pow(ATV(phi, c, VD_EFZ(ns)), 2.0));
    double ndens = 0.0;
    for (int sp = 0; sp < ns; sp++) ndens += ATV(phi, c, sp);
    double Te = ATV(phi, c, VD_ETEMP(ns));
    double mu = spec_mob_st(m, specid, Te, Tg, ndens, efieldmag);
    return mu * ATV(phi, c, VD_EFX(ns) + dir);
}

/* gk(), weno() (Source/compute_explicit_flux.H:16-112) */
static double gk(double wk, double g) {
    return wk * (g + g * g - 3.0 * g * wk + wk * wk) / (g * g + wk * (1 - 2.0 * g));
}

double or_weno(double phi_im2, double phi_im1, double phi_i, double phi_ip1, double phi_ip2,
               int scheme) {
    double eps = 1e-20;
    double p1 = (1.0 / 3.0) * phi_im2 - (7.0 / 6.0) * phi_im1 + (11.0 / 6.0) * phi_i;
    double p2 = -(1.0 / 6.0) * phi_im1 + (5.0 / 6.0) * phi_i + (1.0 / 3.0) * phi_ip1;
    double p3 = (1.0 / 3.0) * phi_i + (5.0 / 6.0) * phi_ip1 - (1.0 / 6.0) * phi_ip2;
    double b1 = (13.0 / 12.0) * pow((phi_im2 - 2.0 * phi_im1 + phi_i), 2.0) +
                (1.0 / 4.0) * pow((phi_im2 - 4.0 * phi_im1 + 3.0 * phi_i), 2.0);
    double b2 = (13.0 / 12.0) * pow((phi_im1 - 2.0 * phi_i + phi_ip1), 2.0) +
                (1.0 / 4.0) * pow((phi_im1 - phi_ip1), 2.0);
    double b3 = (13.0 / 12.0) * pow((phi_i - 2.0 * phi_ip1 + phi_ip2), 2.0) +
                (1.0 / 4.0) * pow((3.0 * phi_i - 4.0 * phi_ip1 + phi_ip2), 2.0);
    double g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0;
    double a1, a2, a3, w1, w2, w3;
    if (scheme == 1) {
        a1 = g1 / pow((b1 + eps), 2.0);
        a2 = g2 / pow((b2 + eps), 2.0);
        a3 = g3 / pow((b3 + eps), 2.0);
    } else if (scheme == 2) {
        a1 = g1 / pow((b1 + eps), 2.0);
        a2 = g2 / pow((b2 + eps), 2.0);
        a3 = g3 / pow((b3 + eps), 2.0);
        double s = a1 + a2 + a3;
        w1 = a1 / s;
        w2 = a2 / s;
        w3 = a3 / s;
        a1 = gk(w1, g1);
        a2 = gk(w2, g2);
        a3 = gk(w3, g3);
    } else {
        double tau = fabs(b1 - b3);
        a1 = g1 * (1.0 + pow(tau / (b1 + eps), 2.0));
        a2 = g2 * (1.0 + pow(tau / (b2 + eps), 2.0));
        a3 = g3 * (1.0 + pow(tau / (b3 + eps), 2.0));
    }
    double s = a1 + a2 + a3;
    w1 = a1 / s;
    w2 = a2 / s;
    w3 = a3 / s;
    return w1 * p1 + w2 * p2 + w3 * p3;
}

#define HEPS 1.0e-30
/* get_slope_lim / WAF / upwind (Source/compute_explicit_flux.H:130-198) */
static double slope_lim(double c, double dl, double dm, double dr) {
    double r = (c >= 0.0) ? dl / (dm + HEPS) : dr / (dm + HEPS);
    double lim;
    if (r <= 0.0)
        lim = 1.0;
    else if (r > 0.0 && r <= 1.0)
        lim = 1.0 - (1.0 - fabs(c)) * r;
    else
        lim = fabs(c);
    return lim;
}
static double waf_flux(double vL, double vR, double pLm1, double pL, double pR, double pRp1,
                       double dx, double dt) {
    double dl = pL - pLm1, dm = pR - pL, dr = pRp1 - pR;
    double vm = 0.5 * (vL + vR);
    double sgn = copysign(1.0, vm);
    double c = vm * dt / dx;
    double fL = pL * vL, fR = pR * vR;
    double lim = slope_lim(c, dl, dm, dr);
    return 0.5 * (fL + fR) - 0.5 * sgn * lim * (fR - fL);
}
static double upwind_flux(double vL, double vR, double pL, double pR) {
    double fL = pL * vL, fR = pR * vR;
    double vm = 0.5 * (vL + vR);
    double sgn = copysign(1.0, vm);
    return fL * 0.5 * (1.0 + sgn) + fR * 0.5 * (1.0 - sgn);
}

/* compute_flux (Source/compute_explicit_flux.H:200-326) on one face */
static double face_flux(const vd_mech_desc* m, const vd_array4* phi, int i, int j, int k, int d,
                        int specid, const int bclo[3], const int bchi[3], const vd_geom* g,
                        const vd_params* prm, double dt) {
    double Tg = prm->gas_temperature;
    int face[3] = {i, j, k}, left[3] = {i, j, k}, right[3] = {i, j, k};
    left[d] -= 1;
    double pL = ATV(phi, left, specid), pR = ATV(phi, right, specid);
    double vL = compute_vel(m, phi, left, d, specid, Tg);
    double vR = compute_vel(m, phi, right, d, specid, Tg);
    double f = 0.0;
    if (prm->hyp_order == 1) {
        f = upwind_flux(vL, vR, pL, pR);
    } else if (prm->hyp_order == 2) {
        int l1[3] = {i, j, k}, r1[3] = {i, j, k};
        l1[d] -= 2;
        r1[d] += 1;
        double dxmin = g->dx[0];
        for (int q = 1; q < g->dim; q++) dxmin = fmin(dxmin, g->dx[q]);
        f = waf_flux(vL, vR, ATV(phi, l1, specid), pL, pR, ATV(phi, r1, specid), dxmin, dt);
    } else {
        int l1[3] = {i, j, k}, r1[3] = {i, j, k}, l2[3] = {i, j, k}, r2[3] = {i, j, k};
        l1[d] -= 2;
        r1[d] += 1;
        l2[d] -= 3;
        r2[d] += 2;
        double vLm1 = compute_vel(m, phi, l1, d, specid, Tg);
        double vLm2 = compute_vel(m, phi, l2, d, specid, Tg);
        double vRp1 = compute_vel(m, phi, r1, d, specid, Tg);
        double vRp2 = compute_vel(m, phi, r2, d, specid, Tg);
        double ulm2 = ATV(phi, l2, specid), ulm1 = ATV(phi, l1, specid), ul = pL, ur = pR;
        double urp1 = ATV(phi, r1, specid), urp2 = ATV(phi, r2, specid);
        double flm2 = ulm2 * vLm2, flm1 = ulm1 * vLm1, fl = ul * vL, fr = ur * vR;
        double frp1 = urp1 * vRp1, frp2 = urp2 * vRp2;
        double lam = fabs(0.5 * (vL + vR));
        int sch = prm->weno_scheme;
        /* weno_reconstruct (compute_explicit_flux.H:114-128) */
        double fphalf = or_weno(flm2, flm1, fl, fr, frp1, sch);
        double fmhalf = or_weno(frp2, frp1, fr, fl, flm1, sch);
        double uphalf = or_weno(ulm2, ulm1, ul, ur, urp1, sch);
        double umhalf = or_weno(urp2, urp1, ur, ul, ulm1, sch);
        f = 0.5 * (fphalf + lam * uphalf) + 0.5 * (fmhalf - lam * umhalf);
    }
    if (face[d] == g->dom_lo[d] && bclo[d] == VD_IHNEUBC) f = 0.0;
    if (face[d] == g->dom_hi[d] + 1 && bchi[d] == VD_IHNEUBC) f = 0.0;
    return f;
}

static int check_flux_args(const vd_params* prm) {
    if (prm->hyp_order != 1 && prm->hyp_order != 2 && prm->hyp_order != 5) return VD_EINVAL;
    if (prm->hyp_order == 5 && (prm->weno_scheme < 1 || prm->weno_scheme > 3)) return VD_EINVAL;
    return VD_OK;
}

/* Vidyut::compute_scalar_transport_flux (Source/ScalarSolve.cpp:170-255) */
int or_compute_scalar_transport_flux(const vd_mech_desc* m, const vd_array4* sb, vd_array4 flux[3],
                                     const vd_box* box, const vd_geom* g, const int bclo[3],
                                     const int bchi[3], const vd_params* prm, double time,
                                     double dt, int specid) {
    (void)time;
    int st = check_flux_args(prm);
    if (st) return st;
    for (int d = 0; d < g->dim; d++) {
        int hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
        hi[d] += 1;
#pragma omp parallel for collapse(2)
        for (int k = box->lo[2]; k <= hi[2]; k++)
            for (int j = box->lo[1]; j <= hi[1]; j++)
                for (int i = box->lo[0]; i <= hi[0]; i++)
                    AT(&flux[d], i, j, k, 0) =
                        face_flux(m, sb, i, j, k, d, specid, bclo, bchi, g, prm, dt);
    }
    return VD_OK;
}

/* Vidyut::compute_dsdt (Source/ScalarSolve.cpp:17-55) */
int or_compute_dsdt(const vd_array4 flux[3], const vd_array4* rxn, vd_array4* dsdt,
                    const vd_box* box, const vd_geom* g, int specid) {
#pragma omp parallel for collapse(2)
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                double v = (AT(&flux[0], i, j, k, 0) - AT(&flux[0], i + 1, j, k, 0)) / g->dx[0] +
                           AT(rxn, i, j, k, specid);
                if (g->dim > 1)
                    v += (AT(&flux[1], i, j, k, 0) - AT(&flux[1], i, j + 1, k, 0)) / g->dx[1];
                if (g->dim == 3)
                    v += (AT(&flux[2], i, j, k, 0) - AT(&flux[2], i, j, k + 1, 0)) / g->dx[2];
                AT(dsdt, i, j, k, 0) = v;
            }
    return VD_OK;
}

/* ------------------------------------------------------------ chemistry */

/* modified-Arrhenius rate k = A*T^beta*exp(-Ea/T) in the form emitted by
 * the reference's mechanism generator (py_utils/ceptr_noThermo/ceptr/
 * production.py): A*exp(beta*log(T) - Ea/T) */
static double arrhenius(const vd_reaction* r, double T) {
    return r->A * exp(r->beta * log(T) - r->Ea / T);
}

/* Source/reaction_source.H:10-64 (table form; NUM_ALL_SPECIES = NS) */
static void table_source(const vd_mech_desc* m, const double* n, double Te, double Tg,
                         double* src) {
    int ns = m->nspec;
    for (int sp = 0; sp < ns + 1; sp++) src[sp] = 0.0;
    for (int r = 0; r < m->nreac; r++) {
        const vd_reaction* R = &m->reac[r];
        double k = arrhenius(R, R->temp_is_te ? Te : Tg);
        double concmult = 1.0;
        for (int q = 0; q < R->nreact; q++) concmult *= pow(n[R->react_sp[q]], R->react_nu[q]);
        double rate = k * concmult;
        for (int q = 0; q < R->nreact; q++) src[R->react_sp[q]] -= R->react_nu[q] * rate;
        for (int q = 0; q < R->nprod; q++) src[R->prod_sp[q]] += R->prod_nu[q] * rate;
        src[ns] += R->energy * rate;
    }
}

/* CKWC -> productionRate -> comp_ener_exch of a generated Chemistry.H
 * (template test/models/ionelec_DC_1d/Chemistry.H:240-357).  C is mol/cm^3
 * on entry, wdot mol/cm^3/s on exit. */
static void ckwc_source(const vd_mech_desc* m, double T, double* C, double* wdot, double Te,
                        double* ener) {
    int ns = m->nspec, E = m->e_idx;
    for (int id = 0; id < ns; ++id) C[id] *= 1.0e6;
    /* productionRate */
    for (int i = 0; i < ns; ++i) wdot[i] = 0.0;
    double ne_pr = C[E] * 6.02214085774e23;
    double Ue = 1.5 * Te * ne_pr * 1.380649e-23;
    for (int r = 0; r < m->nreac; r++) {
        const vd_reaction* R = &m->reac[r];
        double k_f = arrhenius(R, R->temp_is_te ? Te : T);
        int order = 0;
        for (int q = 0; q < R->nreact; q++) order += (int)R->react_nu[q];
        for (int o = 1; o < order; o++) k_f *= 6.02214076e23;
        double cprod = 1.0;
        int first = 1;
        for (int q = 0; q < R->nreact; q++)
            for (int c = 0; c < (int)R->react_nu[q]; c++) {
                if (first) {
                    cprod = C[R->react_sp[q]];
                    first = 0;
                } else {
                    cprod = cprod * C[R->react_sp[q]];
                }
            }
        double qf = k_f * cprod;
        double qr = 0.0;
        double qdot = qf - qr;
        /* net stoichiometry, ascending species */
        for (int sp = 0; sp < ns; sp++) {
            double net = 0.0;
            int hit = 0;
            for (int q = 0; q < R->nreact; q++)
                if (R->react_sp[q] == sp) {
                    net -= R->react_nu[q];
                    hit = 1;
                }
            for (int q = 0; q < R->nprod; q++)
                if (R->prod_sp[q] == sp) {
                    net += R->prod_nu[q];
                    hit = 1;
                }
            if (!hit || net == 0.0) continue;
            if (net == 1.0)
                wdot[sp] += qdot;
            else if (net == -1.0)
                wdot[sp] -= qdot;
            else
                wdot[sp] += net * qdot;
        }
        /* comp_ener_exch (ionelec_DC_1d/Chemistry.H:240-287) */
        double eV2J = 1.602176620800000e-19;
        double NA = 6.02214085774e23;
        double kB = 1.380649e-23;
        double ne = C[E] * NA;
        double mee = Ue / ne;
        double sefact = (tanh((Te - T) / 0.5) + 1.0) / 2.0;
        if (R->rxntype == 1 || R->rxntype == 2) *ener -= (R->eexci * eV2J) * NA * (qf - qr) * sefact;
        if (R->rxntype == 3) *ener -= mee * NA * (qf - qr);
        if (R->rxntype == 5) {
            double qq = C[E] * C[R->elidx] * k_f;
            *ener -= 3.0 * kB * (m->molwt[E] / m->molwt[R->elidx]) * (Te - T) * qq * NA;
        }
    }
    if (m->elastic_loss) {
        /* ionelec_DC_1d/Chemistry.H:334-341 */
        double ndens = 0.0;
        for (int sp = 0; sp < ns; sp++) ndens += C[sp] * N_A;
        double mu = spec_mob_st(m, E, Te, T, ndens, 0.0);
        double nu = -ECHARGE / ME / mu;
        double term = 3.0 / 2.0 * K_B * C[E] * N_A * (Te - T) * nu *
                      (2.0 * ME / (m->elastic_partner_amu * M_AMU));
        *ener -= term;
    }
    for (int id = 0; id < ns; ++id) {
        C[id] *= 1.0e-6;
        wdot[id] *= 1.0e-6;
    }
}

/* g/RT of species k at gas temperature T: gibbs() of a generated
 * Chemistry.H (test/Reactor/CO2_H2_Ar_mechanism_reduced/Chemistry.H:
 * 415-834), c0/T + c1 + c2 ln T + c3 T + c4 T^2 + c5 T^3 + c6 T^4 summed
 * left to right, the low range below gibbs_tmid */
static double gen_gibbs(const vd_mech_desc* m, int k, double T) {
    const double T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;
    const double invT = 1.0 / T, logT = log(T);
    const double* c = (m->gibbs_tmid[k] > 0.0 && !(T < m->gibbs_tmid[k])) ? m->gibbs_hi[k] : m->gibbs_lo[k];
    return c[0] * invT + c[1] + c[2] * logT + c[3] * T + c[4] * T2 + c[5] * T3 + c[6] * T4;
}

/* g/RT of species k (VD_CHEM_GEN thermo), for the tests */
double or_gen_gibbs(const vd_mech_desc* m, int k, double T) { return gen_gibbs(m, k, T); }

/* CKWC -> productionRate -> comp_ener_exch of a mechanism read from its
 * generated Chemistry.H (VD_CHEM_GEN; test/Reactor/
 * CO2_H2_Ar_mechanism_reduced/Chemistry.H:841-6399): the prelude of
 * :913-950, then per reaction block k_f by its rate kind, qf = [mixture *]
 * k_f * prod C, qr from the g/RT sum, the wdot updates as written and the
 * energy exchange of :841-885.  C is mol/cm^3 on entry, wdot mol/cm^3/s on
 * exit. */
static void gen_source(const vd_mech_desc* m, double T, double* C, double* wdot, double Te, double* ener) {
    const int ns = m->nspec, E = m->e_idx;
    for (int id = 0; id < ns; ++id) C[id] *= 1.0e6;
    const double invT = 1.0 / T;
    const double logT = log(T / 300.0);
    const double refC = 101325 / 8.31446 * invT;
    const double refCinv = 1 / refC;
    for (int i = 0; i < ns; ++i) wdot[i] = 0.0;
    double mixture = 0.0;
    for (int i = 0; i < ns; ++i) mixture += C[i];
    double g_RT[VD_MAX_SPECIES];
    for (int k = 0; k < ns; k++) g_RT[k] = gen_gibbs(m, k, T);
    const double ne_pr = C[E] * 6.02214085774e23;
    const double Ue = 1.5 * Te * ne_pr * 1.380649e-23;
    const double TeeV = Te / 11595.0;
    const double logTe = log(TeeV);
    const double invTeeV = (Te == 0) ? 1.0 : 1.0 / (TeeV);
    double Te_pow[9], invTe_pow[4];
    for (int j = 0; j < 9; j++) Te_pow[j] = pow(logTe, j);
    for (int j = 0; j < 4; j++) invTe_pow[j] = pow(invTeeV, j + 1);
    /* comp_ener_exch constants */
    const double eV2J = 1.602176620800000e-19, NA = 6.02214085774e23, kB = 1.380649e-23;
    const double mee = Ue / (C[E] * NA);
    const double sefact = (tanh((Te - T) / 0.5) + 1.0) / 2.0;
    for (int r = 0; r < m->nreac; r++) {
        const vd_reaction* R = &m->reac[r];
        double k_f, sum;
        switch (R->rate_kind) {
        case VD_RATE_ARRH:
            k_f = R->A * exp(R->beta * logT - R->Ea * invT);
            break;
        case VD_RATE_TEPOW:
            k_f = R->A * pow(Te / R->tref, R->beta);
            if (R->tgas_div) k_f = k_f / T;
            break;
        case VD_RATE_TEARRH:
            k_f = R->A * pow(Te, R->beta) * exp(-R->Ea / Te);
            break;
        case VD_RATE_JANEV:
            sum = 0.0;
            for (int j = 0; j < 9; j++) sum += R->fit[j] * Te_pow[j];
            k_f = R->A * exp(sum) * 6.02214085774e23;
            break;
        case VD_RATE_FFIT:
            sum = 0.0;
            for (int j = 0; j < 4; j++) sum += R->fit[j] * invTe_pow[j];
            k_f = R->A * exp(sum) * 6.02214085774e23;
            break;
        default: /* VD_RATE_JANEV_FFIT */
            sum = 0.0;
            if (TeeV < R->fit_switch_ev) {
                for (int j = 0; j < 9; j++) sum += R->fit[j] * Te_pow[j];
                k_f = R->A * exp(sum) * 6.02214085774e23;
            } else {
                for (int j = 0; j < 4; j++) sum += R->fit2[j] * invTe_pow[j];
                k_f = R->A2 * exp(sum) * 6.02214085774e23;
            }
            break;
        }
        if (R->kmax > 0.0 && R->kmax < k_f) k_f = R->kmax; /* std::min(k_f, kmax) */
        const double pre = R->third_body ? mixture * k_f : k_f;
        double cprod = 1.0;
        int first = 1;
        for (int q = 0; q < R->nreact; q++)
            for (int c = 0; c < (int)R->react_nu[q]; c++) {
                cprod = first ? C[R->react_sp[q]] : cprod * C[R->react_sp[q]];
                first = 0;
            }
        const double qf = pre * cprod;
        double qr = 0.0;
        if (R->reversible) {
            double s = 0.0;
            for (int q = 0; q < R->ng; q++) {
                const double t = R->g_nu[q] == 1.0 ? g_RT[R->g_sp[q]]
                                 : R->g_nu[q] == -1.0 ? -g_RT[R->g_sp[q]]
                                                      : R->g_nu[q] * g_RT[R->g_sp[q]];
                s = q == 0 ? t : s + t;
            }
            double v = pre * exp(-(s));
            for (int p = 0; p < R->refc_pow; p++) v = v * refC;
            for (int p = 0; p < -R->refc_pow; p++) v = v * refCinv;
            double rp = 1.0;
            first = 1;
            for (int q = 0; q < R->nrev; q++)
                for (int c = 0; c < (int)R->rev_nu[q]; c++) {
                    rp = first ? C[R->rev_sp[q]] : rp * C[R->rev_sp[q]];
                    first = 0;
                }
            qr = v * rp;
        }
        const double qdot = qf - qr;
        for (int q = 0; q < R->nwd; q++) {
            const double nu = R->wd_nu[q];
            double* w = &wdot[R->wd_sp[q]];
            if (nu == 1.0)
                *w += qdot;
            else if (nu == -1.0)
                *w -= qdot;
            else if (nu > 0.0)
                *w += nu * qdot;
            else
                *w -= (-nu) * qdot;
        }
        const int rt = R->rxntype;
        if (rt == 1 || rt == 2) *ener -= (R->eexci * eV2J) * NA * (qf - qr) * sefact;
        if (rt == 3 || rt == 4) *ener -= mee * NA * (qf - qr);
        if (rt == 5) {
            const double qq = C[E] * C[R->elidx] * k_f;
            *ener -= 3.0 * kB * (m->molwt[E] / m->molwt[R->elidx]) * (Te - T) * qq * NA;
        }
        if (rt == 7) *ener += mee * NA * (qf - qr);
    }
    for (int id = 0; id < ns; ++id) {
        C[id] *= 1.0e-6;
        wdot[id] *= 1.0e-6;
    }
}

/* Vidyut::update_rxnsrc_at_all_levels (Source/ScalarSolve.cpp:113-168) */
int or_update_rxnsrc(const vd_mech_desc* m, const vd_array4* sb, vd_array4* rxn,
                     const vd_box* box, const vd_params* prm, double time) {
    (void)time;
    int ns = m->nspec;
    double Tg = prm->gas_temperature;
#pragma omp parallel for collapse(2)
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                double Te = AT(sb, i, j, k, VD_ETEMP(ns));
                if (m->chem_form == VD_CHEM_TABLE) {
                    double n[VD_MAX_SPECIES], src[VD_MAX_SPECIES + 1];
                    for (int sp = 0; sp < ns; sp++) n[sp] = AT(sb, i, j, k, sp);
                    table_source(m, n, Te, Tg, src);
                    for (int sp = 0; sp <= ns; sp++) AT(rxn, i, j, k, sp) = src[sp];
                } else {
                    double C[VD_MAX_SPECIES], wdot[VD_MAX_SPECIES], ener = 0.0;
                    for (int sp = 0; sp < ns; sp++) C[sp] = AT(sb, i, j, k, sp) * 1.0e-6 / N_A;
                    if (m->chem_form == VD_CHEM_GEN)
                        gen_source(m, Tg, C, wdot, Te, &ener);
                    else
                        ckwc_source(m, Tg, C, wdot, Te, &ener);
                    for (int sp = 0; sp < ns; sp++) AT(rxn, i, j, k, sp) = wdot[sp] * N_A * 1.0e6;
                    AT(rxn, i, j, k, ns) = ener;
                }
                /* user_sources::add_user_react_sources (ScalarSolve.cpp:158-163) */
                for (int sp = 0; sp <= ns; sp++)
                    if (m->user_src[sp] != 0.0) AT(rxn, i, j, k, sp) += m->user_src[sp];
            }
    return VD_OK;
}

/* Vidyut::compute_elecenergy_source (Source/ElecEnergySource.cpp:16-173) */
int or_compute_elecenergy_source(const vd_mech_desc* m, const vd_array4* sb, const vd_array4* rxn,
                                 const vd_array4 efield[3], const vd_array4 gradne[3],
                                 vd_array4* dsdt, vd_array4* phi_new, const vd_box* box,
                                 const vd_geom* g, const vd_params* prm) {
    int ns = m->nspec, eidx = m->e_idx, dim = g->dim;
    double Tg = prm->gas_temperature;
#pragma omp parallel for collapse(2)
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                double charge = m->charge[eidx] * ECHARGE;
                double jheat = 0.0;
                for (int idim = 0; idim < dim; idim++) {
                    for (int f = 0; f < 2; f++) {
                        int face[3] = {i, j, k};
                        face[idim] += f;
                        if (face[idim] == g->dom_lo[idim]) face[idim] += 1;
                        if (face[idim] == g->dom_hi[idim] + 1) face[idim] -= 1;
                        int lc[3] = {face[0], face[1], face[2]}, rc[3] = {face[0], face[1], face[2]};
                        lc[idim] -= 1;
                        double etemp = 0.5 * (ATV(sb, lc, VD_ETEMP(ns)) + ATV(sb, rc, VD_ETEMP(ns)));
                        double Esum = 0.0, ef[3];
                        for (int q = 0; q < dim; q++) {
                            ef[q] = 0.5 * (ATV(sb, lc, VD_EFX(ns) + q) + ATV(sb, rc, VD_EFX(ns) + q));
                            Esum += pow(ef[q], 2.0);
                        }
                        double efield_mag = sqrt(Esum);
                        double ne = 0.5 * (ATV(sb, lc, eidx) + ATV(sb, rc, eidx));
                        double efield_face = ATV(&efield[idim], face, 0);
                        double gradne_face = ATV(&gradne[idim], face, 0);
                        double ndens = 0.0;
                        for (int sp = 0; sp < ns; sp++)
                            ndens += 0.5 * (ATV(sb, lc, sp) + ATV(sb, rc, sp));
                        double mu = prm->const_ele_trans ? prm->ele_mob / ndens
                                                         : spec_mob_st(m, eidx, etemp, Tg, ndens, efield_mag);
                        double dcoeff = prm->const_ele_trans
                                            ? prm->ele_diff / ndens
                                            : spec_diff_st(m, eidx, etemp, Tg, ndens, efield_mag);
                        double cur = charge * (mu * ne * efield_face - dcoeff * gradne_face);
                        jheat += cur * efield_face;
                    }
                }
                jheat *= 0.5;
                if (prm->floor_jh && jheat < 0.0) jheat = 0.0;
                AT(dsdt, i, j, k, 0) += jheat;
                AT(phi_new, i, j, k, VD_EJH(ns)) = jheat;
                AT(phi_new, i, j, k, VD_EIH(ns)) = AT(rxn, i, j, k, ns);
            }
    return VD_OK;
}

/* ----------------------------------------------------- CFL reduction
 * Vidyut::find_time_scales (Source/TimeStep.cpp:34-116) */
int or_find_time_scales(const vd_mech_desc* m, const vd_array4* phi, const vd_box* box,
                        const vd_geom* g, const vd_params* prm, double out[3]) {
    int ns = m->nspec, eidx = m->e_idx, dim = g->dim;
    double Tg = prm->gas_temperature;
    double mdrift = 0.0, mdiff = 0.0, mmuene = 0.0;
#pragma omp parallel for collapse(2) reduction(max : mdrift, mdiff, mmuene)
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                double etemp = AT(phi, i, j, k, VD_ETEMP(ns));
                double Esum = 0.0;
                for (int d = 0; d < dim; d++) Esum += pow(AT(phi, i, j, k, VD_EFX(ns) + d), 2.0);
                double emag = sqrt(Esum);
                double ndens = 0.0;
                for (int sp = 0; sp < ns; sp++) ndens += AT(phi, i, j, k, sp);
                double dcoeff = prm->const_ele_trans ? prm->ele_diff / ndens
                                                     : spec_diff_st(m, eidx, etemp, Tg, ndens, emag);
                double mu = prm->const_ele_trans ? prm->ele_mob / ndens
                                                 : spec_mob_st(m, eidx, etemp, Tg, ndens, emag);
                double a = fabs(fabs(mu) * emag), b = fabs(dcoeff),
                       c = fabs(fabs(mu) * AT(phi, i, j, k, eidx));
                if (a > mdrift) mdrift = a;
                if (b > mdiff) mdiff = b;
                if (c > mmuene) mmuene = c;
            }
    double dt_edrift = DBL_MAX, dt_ediff = DBL_MAX, dt_diel = DBL_MAX;
    if (mdrift > 0)
        for (int d = 0; d < dim; d++) dt_edrift = fmin(dt_edrift, g->dx[d] / mdrift);
    if (mdiff > 0)
        for (int d = 0; d < dim; d++)
            dt_ediff = fmin(dt_ediff, 0.5 * g->dx[d] * g->dx[d] / mdiff / dim);
    if (mmuene > 0) dt_diel = EPS0 / ECHARGE / mmuene;
    out[0] = dt_edrift;
    out[1] = dt_ediff;
    out[2] = dt_diel;
    return VD_OK;
}

/* ---------------------------------------------------------------- tags
 * state_based_refinement / stategrad_based_refinement (Source/Tagging.H:8-69) */
int or_error_est(const vd_array4* st, vd_tagarray4* tags, const vd_box* box, int ntag,
                 const int* comps, const double* thr, const double* gthr, char tagval) {
    int dim = (box->lo[2] == box->hi[2] && st->end[2] - st->begin[2] == 1) ? 2 : 3;
#pragma omp parallel for collapse(2)
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                char* t = tags->p + (i - tags->begin[0]) + (j - tags->begin[1]) * tags->jstride +
                          (k - tags->begin[2]) * tags->kstride;
                for (int c = 0; c < ntag; c++)
                    if (fabs(AT(st, i, j, k, comps[c])) > thr[c]) *t = tagval;
                for (int c = 0; c < ntag; c++) {
                    int q = comps[c];
                    double gv[3] = {0.0, 0.0, 0.0};
                    double gp = fabs(AT(st, i + 1, j, k, q) - AT(st, i, j, k, q));
                    double gm = fabs(AT(st, i, j, k, q) - AT(st, i - 1, j, k, q));
                    gv[0] = gp > gm ? gp : gm;
                    gp = fabs(AT(st, i, j + 1, k, q) - AT(st, i, j, k, q));
                    gm = fabs(AT(st, i, j, k, q) - AT(st, i, j - 1, k, q));
                    gv[1] = gp > gm ? gp : gm;
                    double mx = gv[0] > gv[1] ? gv[0] : gv[1];
                    if (dim == 3) {
                        gp = fabs(AT(st, i, j, k + 1, q) - AT(st, i, j, k, q));
                        gm = fabs(AT(st, i, j, k, q) - AT(st, i, j, k - 1, q));
                        gv[2] = gp > gm ? gp : gm;
                        mx = mx > gv[2] ? mx : gv[2];
                    }
                    if (mx > gthr[c]) *t = tagval;
                }
            }
    return VD_OK;
}

/* ============================================================ multigrid
 * Restatement of the MLABecLaplacian + MLMG solve used by
 * PotentialSolve.cpp:148-158,306-336 and ScalarSolve.cpp:420-437,579-596.
 * AMReX is not in the reference tree; the conventions adopted here (and
 * by the HIP library) are:
 *   operator   (L phi)_c = alpha_a*phi_c - beta*sum_d [b_hi (phi_hi-phi_c)
 *                          - b_lo (phi_c-phi_lo)]/dx_d^2, evaluated as
 *              dg*phi_c - of with dg = alpha_a + sum_d s_d*dd_d,
 *              of = sum_d s_d*oo_d, s_d = beta/(dx_d*dx_d)
 *   boundary   ghost = c0*phi_c + cinh:  Dirichlet (maxorder 2) c0=-1,
 *              cinh=2v; Neumann c0=1; inhomogeneous Neumann (value g =
 *              dphi/dx_d) c0=1, cinh=-/+dx*g (lo/hi); Robin a*phi+b*dphi/dn=f
 *              c0=-(a/2-b/dx)/(a/2+b/dx), cinh=f/(a/2+b/dx)
 *   cycle      correction form, V-cycle, red-black Gauss-Seidel (colour
 *              (i+j+k)&1, red first), cell-average restriction, linear
 *              (tensor 3/4-1/4) prolongation, face coefficients averaged to
 *              coarse faces, CG bottom solve on the coarsest level
 *   shortcut   when rho = sum_d 2|beta/dx_d^2| max|b_d| / |alpha| <= 1e-2
 *              an iteration is one finest-level red-black sweep on phi
 *   stop       ||r||_inf <= max(abstol, reltol*max(||b||_inf, ||r0||_inf)),
 *              b = rhs with the boundary data folded in (MLMG::solve's
 *              bnorm / resid0 choice)
 * ==================================================================== */

typedef struct {
    int n[3];
    double dx[3];
    double* bf[3];   /* face coefficients */
    double* c0[3][2];/* per boundary-face c0, [dir][side] */
    double *e, *r, *t;
} mglev;

static inline int64_t cidx(const mglev* L, int i, int j, int k) {
    return i + (int64_t)L->n[0] * (j + (int64_t)L->n[1] * k);
}
static inline int64_t fidx(const mglev* L, int d, int i, int j, int k) {
    int n0 = L->n[0] + (d == 0), n1 = L->n[1] + (d == 1);
    return i + (int64_t)n0 * (j + (int64_t)n1 * k);
}
static inline int64_t tidx(const mglev* L, int d, int i, int j, int k) {
    if (d == 0) return j + (int64_t)L->n[1] * k;
    if (d == 1) return i + (int64_t)L->n[0] * k;
    return i + (int64_t)L->n[0] * j;
}
static inline int64_t nface(const mglev* L, int d) {
    return (int64_t)(L->n[0] + (d == 0)) * (L->n[1] + (d == 1)) * (L->n[2] + (d == 2));
}
static inline int64_t ntan(const mglev* L, int d) {
    return d == 0 ? (int64_t)L->n[1] * L->n[2]
                  : (d == 1 ? (int64_t)L->n[0] * L->n[2] : (int64_t)L->n[0] * L->n[1]);
}

typedef struct {
    int dim, nlev, per[3];
    double alpha, beta;
    mglev L[24];
} mgctx;

static void stencil(const mgctx* M, const mglev* L, const double* x, int i, int j, int k,
                    double* dgo, double* ofo) {
    double dg = M->alpha, of = 0.0;
    int c[3] = {i, j, k};
    for (int d = 0; d < M->dim; d++) {
        double s = M->beta / (L->dx[d] * L->dx[d]);
        int nd = L->n[d];
        int cl[3] = {i, j, k}, ch[3] = {i, j, k};
        cl[d] -= 1;
        ch[d] += 1;
        double blo = L->bf[d][fidx(L, d, i, j, k)];
        double bhi = L->bf[d][fidx(L, d, ch[0], ch[1], ch[2])];
        double dd = 0.0, oo = 0.0;
        if (c[d] > 0) {
            dd += blo;
            oo += blo * x[cidx(L, cl[0], cl[1], cl[2])];
        } else if (M->per[d]) {
            cl[d] = nd - 1;
            dd += blo;
            oo += blo * x[cidx(L, cl[0], cl[1], cl[2])];
        } else {
            dd += blo * (1.0 - L->c0[d][0][tidx(L, d, i, j, k)]);
        }
        if (c[d] < nd - 1) {
            dd += bhi;
            oo += bhi * x[cidx(L, ch[0], ch[1], ch[2])];
        } else if (M->per[d]) {
            ch[d] = 0;
            dd += bhi;
            oo += bhi * x[cidx(L, ch[0], ch[1], ch[2])];
        } else {
            dd += bhi * (1.0 - L->c0[d][1][tidx(L, d, i, j, k)]);
        }
        dg += s * dd;
        of += s * oo;
    }
    *dgo = dg;
    *ofo = of;
}

static void gsrb(const mgctx* M, const mglev* L, double* x, const double* rhs) {
    for (int color = 0; color < 2; color++) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < L->n[2]; k++)
            for (int j = 0; j < L->n[1]; j++)
                for (int i = (color + j + k) & 1; i < L->n[0]; i += 2) {
                    double dg, of;
                    stencil(M, L, x, i, j, k, &dg, &of);
                    int64_t c = cidx(L, i, j, k);
                    x[c] = (rhs[c] + of) / dg;
                }
    }
}

static double residual(const mgctx* M, const mglev* L, const double* x, const double* rhs,
                       double* res) {
    double mx = 0.0;
#pragma omp parallel for collapse(2) reduction(max : mx)
    for (int k = 0; k < L->n[2]; k++)
        for (int j = 0; j < L->n[1]; j++)
            for (int i = 0; i < L->n[0]; i++) {
                double dg, of;
                stencil(M, L, x, i, j, k, &dg, &of);
                int64_t c = cidx(L, i, j, k);
                double r = rhs[c] - (dg * x[c] - of);
                res[c] = r;
                double a = fabs(r);
                if (a != a) a = INFINITY;
                if (a > mx) mx = a;
            }
    return mx;
}

static void apply_op(const mgctx* M, const mglev* L, const double* x, double* y) {
    for (int k = 0; k < L->n[2]; k++)
        for (int j = 0; j < L->n[1]; j++)
            for (int i = 0; i < L->n[0]; i++) {
                double dg, of;
                stencil(M, L, x, i, j, k, &dg, &of);
                int64_t c = cidx(L, i, j, k);
                y[c] = dg * x[c] - of;
            }
}

static void restrict_avg(const mgctx* M, const mglev* F, mglev* C) {
    double w = M->dim == 3 ? 0.125 : 0.25;
    int kz = M->dim == 3 ? 2 : 1;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                double s = 0.0;
                for (int kk = 0; kk < kz; kk++)
                    for (int jj = 0; jj < 2; jj++)
                        for (int ii = 0; ii < 2; ii++)
                            s += F->t[cidx(F, 2 * i + ii, 2 * j + jj, kz == 2 ? 2 * k + kk : 0)];
                C->r[cidx(C, i, j, k)] = s * w;
            }
}

/* e_f += P e_c, linear cell-centred prolongation */
static void prolong_add(const mgctx* M, const mglev* C, const mglev* F) {
#pragma omp parallel for collapse(2)
    for (int k = 0; k < F->n[2]; k++)
        for (int j = 0; j < F->n[1]; j++)
            for (int i = 0; i < F->n[0]; i++) {
                int f[3] = {i, j, k};
                int ci[3][2];
                double wt[3][2], fac[3][2];
                for (int d = 0; d < 3; d++) {
                    if (d >= M->dim) {
                        ci[d][0] = ci[d][1] = 0;
                        wt[d][0] = 1.0;
                        wt[d][1] = 0.0;
                        fac[d][0] = fac[d][1] = 1.0;
                        continue;
                    }
                    int I = f[d] >> 1, o = f[d] & 1;
                    int nb = o ? I + 1 : I - 1;
                    ci[d][0] = I;
                    wt[d][0] = 0.75;
                    wt[d][1] = 0.25;
                    fac[d][0] = 1.0;
                    fac[d][1] = 1.0;
                    if (nb < 0 || nb >= C->n[d]) {
                        if (M->per[d]) {
                            nb = nb < 0 ? C->n[d] - 1 : 0;
                        } else {
                            fac[d][1] = -2.0; /* marker: ghost of cell I */
                            nb = I;
                        }
                    }
                    ci[d][1] = nb;
                }
                double val = 0.0;
                int kn = M->dim == 3 ? 2 : 1;
                for (int kz = 0; kz < kn; kz++)
                    for (int jy = 0; jy < 2; jy++)
                        for (int ix = 0; ix < 2; ix++) {
                            int cc[3] = {ci[0][ix], ci[1][jy], ci[2][kz]};
                            int sel[3] = {ix, jy, kz};
                            double v = C->e[cidx(C, cc[0], cc[1], cc[2])];
                            for (int d = 0; d < M->dim; d++)
                                if (sel[d] == 1 && fac[d][1] == -2.0) {
                                    int side = (f[d] & 1) ? 1 : 0;
                                    v = v * C->c0[d][side][tidx(C, d, cc[0], cc[1], cc[2])];
                                }
                            double w = wt[0][ix] * wt[1][jy] * (M->dim == 3 ? wt[2][kz] : 1.0);
                            val += w * v;
                        }
                F->e[cidx(F, i, j, k)] += val;
            }
}

static double dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

/* CG on the coarsest level: e = A^{-1} r (A symmetric definite) */
static void bottom_cg(const mgctx* M, mglev* L) {
    int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double* r = (double*)malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; c++) {
        L->e[c] = 0.0;
        r[c] = L->r[c];
        p[c] = r[c];
    }
    double rr = dot(r, r, n), rr0 = rr;
    int maxit = (int)(4 * n + 20);
    for (int it = 0; it < maxit; it++) {
        if (rr <= 1e-28 * rr0 || rr == 0.0) break;
        apply_op(M, L, p, q);
        double pq = dot(p, q, n);
        if (pq == 0.0) break;
        double alpha = rr / pq;
        for (int64_t c = 0; c < n; c++) {
            L->e[c] += alpha * p[c];
            r[c] -= alpha * q[c];
        }
        double rn = dot(r, r, n);
        double beta = rn / rr;
        for (int64_t c = 0; c < n; c++) p[c] = r[c] + beta * p[c];
        rr = rn;
    }
    free(p);
    free(q);
    free(r);
}

static void vcycle(const mgctx* M, int l, int nu1, int nu2) {
    mglev* L = (mglev*)&M->L[l];
    int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
    if (l == M->nlev - 1) {
        bottom_cg(M, L);
        return;
    }
    memset(L->e, 0, sizeof(double) * n);
    for (int s = 0; s < nu1; s++) gsrb(M, L, L->e, L->r);
    residual(M, L, L->e, L->r, L->t);
    restrict_avg(M, L, (mglev*)&M->L[l + 1]);
    vcycle(M, l + 1, nu1, nu2);
    prolong_add(M, &M->L[l + 1], L);
    for (int s = 0; s < nu2; s++) gsrb(M, L, L->e, L->r);
}

static void coarsen_faces(const mgctx* M, const mglev* F, mglev* C) {
    int dim = M->dim;
    for (int d = 0; d < dim; d++) {
        int nn[3] = {C->n[0] + (d == 0), C->n[1] + (d == 1), C->n[2] + (d == 2)};
        double w = dim == 3 ? 0.25 : 0.5;
        for (int k = 0; k < nn[2]; k++)
            for (int j = 0; j < nn[1]; j++)
                for (int i = 0; i < nn[0]; i++) {
                    int b[3] = {2 * i, 2 * j, dim == 3 ? 2 * k : 0};
                    double s = 0.0;
                    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2; /* tangential dirs */
                    int nt2 = (dim == 3) ? 2 : 1;
                    for (int b2 = 0; b2 < nt2; b2++)
                        for (int b1 = 0; b1 < 2; b1++) {
                            int f[3] = {b[0], b[1], b[2]};
                            f[t1] += b1;
                            if (dim == 3) f[t2] += b2;
                            s += F->bf[d][fidx(F, d, f[0], f[1], f[2])];
                        }
                    C->bf[d][fidx(C, d, i, j, k)] = s * w;
                }
        if (!M->per[d])
            for (int side = 0; side < 2; side++) {
                int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                int nt1 = C->n[t1], nt2 = (dim == 3) ? C->n[t2] : 1;
                for (int b2 = 0; b2 < nt2; b2++)
                    for (int b1 = 0; b1 < nt1; b1++) {
                        double s = 0.0;
                        int nk = dim == 3 ? 2 : 1;
                        for (int q2 = 0; q2 < nk; q2++)
                            for (int q1 = 0; q1 < 2; q1++) {
                                int f[3] = {0, 0, 0};
                                f[t1] = 2 * b1 + q1;
                                if (dim == 3) f[t2] = 2 * b2 + q2;
                                s += F->c0[d][side][tidx(F, d, f[0], f[1], f[2])];
                            }
                        int c[3] = {0, 0, 0};
                        c[t1] = b1;
                        if (dim == 3) c[t2] = b2;
                        C->c0[d][side][tidx(C, d, c[0], c[1], c[2])] = s * w;
                    }
            }
    }
}

static int bc_type_of(int code, int periodic) {
    if (periodic) return VD_PERBC;
    return code;
}

/* ghost = c0*phi + cinh for one boundary face */
static void bc_coeffs(int type, int side, double dx, double val, double ra, double rb, double rf,
                      double* c0, double* cinh) {
    switch (type) {
        case VD_DIRCBC: *c0 = -1.0; *cinh = 2.0 * val; break;
        case VD_HNEUBC: *c0 = 1.0; *cinh = 0.0; break;
        case VD_IHNEUBC: *c0 = 1.0; *cinh = (side == 0 ? -dx : dx) * val; break;
        default: {
            double den = ra / 2.0 + rb / dx;
            *c0 = -(ra / 2.0 - rb / dx) / den;
            *cinh = rf / den;
        }
    }
}

static int build_hierarchy(mgctx* M, const vd_mlmg_problem* P, const vd_box* box,
                           const vd_geom* g) {
    M->dim = g->dim;
    M->alpha = P->alpha_a;
    M->beta = P->beta;
    for (int d = 0; d < 3; d++) M->per[d] = d < g->dim ? g->is_periodic[d] : 0;
    mglev* L0 = &M->L[0];
    memset(M->L, 0, sizeof(M->L));
    for (int d = 0; d < 3; d++) {
        L0->n[d] = d < g->dim ? box->hi[d] - box->lo[d] + 1 : 1;
        L0->dx[d] = d < g->dim ? g->dx[d] : 1.0;
    }
    int nlev = 1;
    int maxc = P->max_coarsening_level;
    while (nlev - 1 < maxc && nlev < 24) {
        mglev* F = &M->L[nlev - 1];
        int ok = 1;
        for (int d = 0; d < g->dim; d++)
            if (F->n[d] % 2 != 0 || F->n[d] < 4) ok = 0;
        if (!ok) break;
        mglev* C = &M->L[nlev];
        for (int d = 0; d < 3; d++) {
            C->n[d] = d < g->dim ? F->n[d] / 2 : 1;
            C->dx[d] = d < g->dim ? F->dx[d] * 2.0 : 1.0;
        }
        nlev++;
    }
    M->nlev = nlev;
    for (int l = 0; l < nlev; l++) {
        mglev* L = &M->L[l];
        int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
        L->e = (double*)calloc(n, sizeof(double));
        L->r = (double*)calloc(n, sizeof(double));
        L->t = (double*)calloc(n, sizeof(double));
        for (int d = 0; d < g->dim; d++) {
            L->bf[d] = (double*)calloc(nface(L, d), sizeof(double));
            for (int s = 0; s < 2; s++) L->c0[d][s] = (double*)calloc(ntan(L, d), sizeof(double));
        }
    }
    return VD_OK;
}

static void free_hierarchy(mgctx* M) {
    for (int l = 0; l < M->nlev; l++) {
        mglev* L = &M->L[l];
        free(L->e);
        free(L->r);
        free(L->t);
        for (int d = 0; d < M->dim; d++) {
            free(L->bf[d]);
            for (int s = 0; s < 2; s++) free(L->c0[d][s]);
        }
    }
}

/* fills level-0 face b, c0 and returns the boundary rhs shift in bshift */
static void setup_level0(mgctx* M, const vd_mlmg_problem* P, const vd_box* box, const vd_geom* g,
                         double* bshift) {
    mglev* L = &M->L[0];
    int dim = g->dim;
    for (int d = 0; d < dim; d++) {
        int nn[3] = {L->n[0] + (d == 0), L->n[1] + (d == 1), L->n[2] + (d == 2)};
        for (int k = 0; k < nn[2]; k++)
            for (int j = 0; j < nn[1]; j++)
                for (int i = 0; i < nn[0]; i++)
                    L->bf[d][fidx(L, d, i, j, k)] =
                        P->bface[d] ? AT(P->bface[d], i + box->lo[0], j + box->lo[1], k + box->lo[2], 0)
                                    : 1.0;
    }
    int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
    for (int64_t c = 0; c < n; c++) bshift[c] = 0.0;
    for (int d = 0; d < dim; d++) {
        if (M->per[d]) continue;
        double s = M->beta / (L->dx[d] * L->dx[d]);
        for (int side = 0; side < 2; side++) {
            int type = bc_type_of(side ? P->bc_hi[d] : P->bc_lo[d], 0);
            int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            int nt2 = dim == 3 ? L->n[t2] : 1;
            for (int b2 = 0; b2 < nt2; b2++)
                for (int b1 = 0; b1 < L->n[t1]; b1++) {
                    int c[3] = {0, 0, 0};
                    c[t1] = b1;
                    if (dim == 3) c[t2] = b2;
                    c[d] = side ? L->n[d] - 1 : 0;
                    int gc[3] = {c[0] + box->lo[0], c[1] + box->lo[1], c[2] + box->lo[2]};
                    gc[d] += side ? 1 : -1; /* ghost cell (absolute index) */
                    double val = P->bcval ? ATV(P->bcval, gc, 0) : 0.0;
                    double ra = 0, rb = 1, rf = 0;
                    if (type == VD_ROBINBC && P->robin_a) {
                        ra = ATV(P->robin_a, gc, 0);
                        rb = ATV(P->robin_b, gc, 0);
                        rf = ATV(P->robin_f, gc, 0);
                    }
                    double c0, cinh;
                    bc_coeffs(type, side, L->dx[d], val, ra, rb, rf, &c0, &cinh);
                    L->c0[d][side][tidx(L, d, c[0], c[1], c[2])] = c0;
                    int fc[3] = {c[0], c[1], c[2]};
                    fc[d] += side;
                    double b = L->bf[d][fidx(L, d, fc[0], fc[1], fc[2])];
                    bshift[cidx(L, c[0], c[1], c[2])] += s * b * cinh;
                }
        }
    }
}

int or_mlmg_assemble(const vd_mlmg_problem* P, const vd_box* box, const vd_geom* g, double* A,
                     double* bshift) {
    mgctx* M = (mgctx*)calloc(1, sizeof(mgctx));
    vd_mlmg_problem Q = *P;
    Q.max_coarsening_level = 0;
    build_hierarchy(M, &Q, box, g);
    setup_level0(M, &Q, box, g, bshift);
    mglev* L = &M->L[0];
    int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
    double* x = (double*)calloc(n, sizeof(double));
    double* y = (double*)calloc(n, sizeof(double));
    for (int64_t c = 0; c < n; c++) {
        x[c] = 1.0;
        apply_op(M, L, x, y);
        for (int64_t r = 0; r < n; r++) A[r * n + c] = y[r];
        x[c] = 0.0;
    }
    free(x);
    free(y);
    free_hierarchy(M);
    free(M);
    return VD_OK;
}

int or_mlmg_solve(const vd_mlmg_problem* P, vd_array4* phi, const vd_array4* rhs,
                  const vd_box* box, const vd_geom* g, vd_solve_info* info) {
    mgctx* M = (mgctx*)calloc(1, sizeof(mgctx));
    build_hierarchy(M, P, box, g);
    mglev* L = &M->L[0];
    int64_t n = (int64_t)L->n[0] * L->n[1] * L->n[2];
    double* f = (double*)malloc(sizeof(double) * n);
    double* x = (double*)malloc(sizeof(double) * n);
    setup_level0(M, P, box, g, f);
    for (int l = 1; l < M->nlev; l++) coarsen_faces(M, &M->L[l - 1], &M->L[l]);
    for (int k = 0; k < L->n[2]; k++)
        for (int j = 0; j < L->n[1]; j++)
            for (int i = 0; i < L->n[0]; i++) {
                int64_t c = cidx(L, i, j, k);
                f[c] = AT(rhs, i + box->lo[0], j + box->lo[1], k + box->lo[2], 0) + f[c];
                x[c] = AT(phi, i + box->lo[0], j + box->lo[1], k + box->lo[2], 0);
            }
    double r0 = residual(M, L, x, f, L->r);
    double bnorm = 0.0;
    for (int64_t c = 0; c < n; c++) {
        double a = fabs(f[c]);
        if (a != a) a = INFINITY;
        if (a > bnorm) bnorm = a;
    }
    double target = fmax(P->abstol, P->reltol * fmax(bnorm, r0));
    /* Jacobi bound rho = sum_d 2|beta/dx_d^2| max|b_d| / |alpha|: at
     * rho <= 1e-2 each red-black sweep contracts the error by rho and the
     * solve uses finest-level sweeps only */
    int smooth_only = 0;
    if (P->alpha_a != 0.0) {
        double rho = 0.0;
        for (int d = 0; d < g->dim; d++) {
            double bm = 0.0;
            for (int64_t q = 0; q < nface(L, d); q++) bm = fmax(bm, fabs(L->bf[d][q]));
            rho += 2.0 * fabs(P->beta / (L->dx[d] * L->dx[d])) * bm;
        }
        smooth_only = rho / fabs(P->alpha_a) <= 1e-2;
    }
    double rn = r0;
    int it = 0, st = VD_OK;
    if (!(r0 <= target)) {
        for (it = 1; it <= P->maxiter; it++) {
            if (smooth_only) {
                gsrb(M, L, x, f);
            } else {
                vcycle(M, 0, P->pre_smooth, P->post_smooth);
                for (int64_t c = 0; c < n; c++) x[c] += L->e[c];
            }
            rn = residual(M, L, x, f, L->r);
            if (rn <= target) break;
        }
        if (it > P->maxiter) {
            it = P->maxiter;
            st = VD_ENOCONV;
        }
    }
    for (int k = 0; k < L->n[2]; k++)
        for (int j = 0; j < L->n[1]; j++)
            for (int i = 0; i < L->n[0]; i++)
                AT(phi, i + box->lo[0], j + box->lo[1], k + box->lo[2], 0) = x[cidx(L, i, j, k)];
    if (info) {
        info->iterations = it;
        info->resnorm0 = r0;
        info->resnorm = rn;
        info->target = target;
    }
    free(f);
    free(x);
    free_hierarchy(M);
    free(M);
    return st;
}

/* face gradient of a solved field (MLMG::getGradSolution), same BC
 * conventions as the operator; grad[d] are face arrays on box. */
static void grad_solution(const vd_array4* x, vd_array4 grad[3], const vd_box* box,
                          const vd_geom* g, const int bclo[3], const int bchi[3],
                          const vd_array4* bcval, const vd_array4* ra, const vd_array4* rb,
                          const vd_array4* rf) {
    for (int d = 0; d < g->dim; d++) {
        int hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
        hi[d] += 1;
        double dx = g->dx[d];
        int nd = box->hi[d] - box->lo[d] + 1;
        for (int k = box->lo[2]; k <= hi[2]; k++)
            for (int j = box->lo[1]; j <= hi[1]; j++)
                for (int i = box->lo[0]; i <= hi[0]; i++) {
                    int f[3] = {i, j, k}, l[3] = {i, j, k}, r[3] = {i, j, k};
                    l[d] -= 1;
                    double gv;
                    int atlo = f[d] == box->lo[d], athi = f[d] == box->hi[d] + 1;
                    if (!atlo && !athi) {
                        gv = (ATV(x, r, 0) - ATV(x, l, 0)) / dx;
                    } else if (g->is_periodic[d]) {
                        if (atlo) l[d] += nd;
                        else r[d] -= nd;
                        gv = (ATV(x, r, 0) - ATV(x, l, 0)) / dx;
                    } else {
                        int side = athi;
                        int type = side ? bchi[d] : bclo[d];
                        int in[3] = {i, j, k}, gc[3] = {i, j, k};
                        if (side) in[d] -= 1;
                        else gc[d] -= 1;
                        double val = bcval ? ATV(bcval, gc, 0) : 0.0;
                        double xin = ATV(x, in, 0);
                        if (type == VD_DIRCBC) {
                            gv = side ? 2.0 * (val - xin) / dx : 2.0 * (xin - val) / dx;
                        } else if (type == VD_HNEUBC) {
                            gv = 0.0;
                        } else if (type == VD_IHNEUBC) {
                            gv = val;
                        } else {
                            double a = 0, b = 1, ff = 0, c0, cinh;
                            if (ra) {
                                a = ATV(ra, gc, 0);
                                b = ATV(rb, gc, 0);
                                ff = ATV(rf, gc, 0);
                            }
                            bc_coeffs(VD_ROBINBC, side, dx, val, a, b, ff, &c0, &cinh);
                            double xg = c0 * xin + cinh;
                            gv = side ? (xg - xin) / dx : (xin - xg) / dx;
                        }
                    }
                    AT(&grad[d], i, j, k, 0) = gv;
                }
    }
}

/* get_applied_potential (Source/HelperFuncs.H:75-92) */
static double applied_potential(double t, int domend, const vd_params* p) {
    double amp = domend == -1 ? p->voltage_amp_1 : p->voltage_amp_2;
    double v;
    if (p->voltage_profile == 1) {
        v = sin(2.0 * PI * p->voltage_freq * t) * amp;
    } else if (p->voltage_profile == 2) {
        double vc = p->voltage_center, vd = p->voltage_dur;
        if (t <= vc)
            v = (vc - t < vd / 2.0) ? (1.0 - (vc - t) / (vd / 2.0)) * amp : 0.0;
        else
            v = (t - vc < vd / 2.0) ? (1.0 - (t - vc) / (vd / 2.0)) * amp : 0.0;
    } else {
        v = amp;
    }
    return v;
}

static vd_array4 make_scratch(const vd_box* box, int ng, int ncomp, double fill) {
    vd_array4 a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; d++) {
        int gg = (box->lo[2] == box->hi[2] && d == 2) ? 0 : ng;
        a.begin[d] = box->lo[d] - gg;
        a.end[d] = box->hi[d] + 1 + gg;
    }
    a.jstride = a.end[0] - a.begin[0];
    a.kstride = a.jstride * (a.end[1] - a.begin[1]);
    a.nstride = a.kstride * (a.end[2] - a.begin[2]);
    a.ncomp = ncomp;
    a.p = (double*)malloc(sizeof(double) * a.nstride * ncomp);
    for (int64_t q = 0; q < a.nstride * ncomp; q++) a.p[q] = fill;
    return a;
}
static vd_array4 make_face(const vd_box* box, int d, double fill) {
    vd_box fb = *box;
    fb.hi[d] += 1;
    return make_scratch(&fb, 0, 1, fill);
}

/* Vidyut::solve_potential (Source/PotentialSolve.cpp:19-393) */
int or_solve_potential(const vd_mech_desc* m, const vd_array4* sb, vd_array4* phi_new,
                       vd_array4 efield[3], const vd_box* box, const vd_geom* g,
                       const vd_params* prm, double time, vd_solve_info* info) {
    int ns = m->nspec, dim = g->dim;
    vd_array4 rhs = make_scratch(box, 0, 1, 0.0);
    vd_array4 pot = make_scratch(box, 1, 1, 0.0);
    vd_array4 ra = make_scratch(box, 1, 1, 0.0), rb = make_scratch(box, 1, 1, 1.0),
              rf = make_scratch(box, 1, 1, 0.0);
    vd_array4 sol = make_scratch(box, 0, 1, 0.0);
    /* pot = Sborder POT with 1 ghost (PotentialSolve.cpp:166) */
    for (int k = pot.begin[2]; k < pot.end[2]; k++)
        for (int j = pot.begin[1]; j < pot.end[1]; j++)
            for (int i = pot.begin[0]; i < pot.end[0]; i++)
                AT(&pot, i, j, k, 0) = AT(sb, i, j, k, VD_POT(ns));
    if (prm->do_spacechrg) {
        for (int k = box->lo[2]; k <= box->hi[2]; k++)
            for (int j = box->lo[1]; j <= box->hi[1]; j++)
                for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                    double r = 0.0;
                    for (int sp = 0; sp < ns; sp++)
                        if (abs(m->charge[sp]) > 0) r += m->charge[sp] * AT(sb, i, j, k, sp);
                    r *= (-ECHARGE / EPS0);
                    AT(&rhs, i, j, k, 0) = r;
                }
    }
    /* potential_bc on every side (PlasmaChem.H:18-72) */
    for (int d = 0; d < dim; d++)
        for (int side = 0; side < 2; side++) {
            double V = applied_potential(time, side ? 1 : -1, prm);
            int hi[3] = {box->hi[0], box->hi[1], box->hi[2]}, lo[3] = {box->lo[0], box->lo[1], box->lo[2]};
            lo[d] = hi[d] = side ? box->hi[d] + 1 : box->lo[d] - 1;
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        AT(&ra, i, j, k, 0) = 1.0;
                        AT(&rb, i, j, k, 0) = 0.0;
                        AT(&rf, i, j, k, 0) = V;
                        AT(&pot, i, j, k, 0) = V;
                    }
        }
    vd_array4 bneg[3];
    const vd_array4* bptr[3] = {NULL, NULL, NULL};
    for (int d = 0; d < dim; d++) {
        bneg[d] = make_face(box, d, -1.0);
        bptr[d] = &bneg[d];
    }
    vd_mlmg_problem P;
    memset(&P, 0, sizeof(P));
    P.alpha_a = 0.0;
    P.beta = 1.0;
    for (int d = 0; d < 3; d++) {
        P.bface[d] = bptr[d];
        P.bc_lo[d] = prm->pot_bc_lo[d];
        P.bc_hi[d] = prm->pot_bc_hi[d];
    }
    P.bcval = &pot;
    P.robin_a = &ra;
    P.robin_b = &rb;
    P.robin_f = &rf;
    P.reltol = prm->linsolve_reltol;
    P.abstol = prm->linsolve_abstol;
    P.maxiter = prm->linsolve_maxiter;
    P.pre_smooth = prm->mg_pre_smooth;
    P.post_smooth = prm->mg_post_smooth;
    P.max_coarsening_level = prm->linsolve_max_coarsening_level;
    int st = or_mlmg_solve(&P, &sol, &rhs, box, g, info);
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++)
                AT(phi_new, i, j, k, VD_POT(ns)) = AT(&sol, i, j, k, 0);
    /* getGradSolution, E = -grad phi (PotentialSolve.cpp:347-357) */
    vd_array4 solg = make_scratch(box, 1, 1, 0.0);
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++)
                AT(&solg, i, j, k, 0) = AT(&sol, i, j, k, 0);
    int blo[3], bhi[3];
    for (int d = 0; d < 3; d++) {
        blo[d] = prm->pot_bc_lo[d];
        bhi[d] = prm->pot_bc_hi[d];
    }
    grad_solution(&solg, efield, box, g, blo, bhi, &pot, &ra, &rb, &rf);
    for (int d = 0; d < dim; d++) {
        int hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
        hi[d] += 1;
        for (int k = box->lo[2]; k <= hi[2]; k++)
            for (int j = box->lo[1]; j <= hi[1]; j++)
                for (int i = box->lo[0]; i <= hi[0]; i++) AT(&efield[d], i, j, k, 0) *= -1.0;
    }
    /* average_face_to_cellcenter + reduced field (PotentialSolve.cpp:359-379) */
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                for (int d = 0; d < dim; d++) {
                    int c[3] = {i, j, k};
                    double lo = ATV(&efield[d], c, 0);
                    c[d] += 1;
                    double hi = ATV(&efield[d], c, 0);
                    AT(phi_new, i, j, k, VD_EFX(ns) + d) = 0.5 * (lo + hi);
                }
                double Esum = 0.0, ndens = 0.0;
                for (int sp = 0; sp < ns; sp++) ndens += AT(phi_new, i, j, k, sp);
                for (int d = 0; d < dim; d++) {
                    double e = AT(phi_new, i, j, k, VD_EFX(ns) + d);
                    Esum += e * e;
                }
                AT(phi_new, i, j, k, VD_REF(ns)) = (pow(Esum, 0.5) / ndens) / 1.0e-21;
            }
    free(rhs.p);
    free(pot.p);
    free(ra.p);
    free(rb.p);
    free(rf.p);
    free(sol.p);
    free(solg.p);
    for (int d = 0; d < dim; d++) free(bneg[d].p);
    return st;
}

/* plasmachem_transport::species_bc (Source/PlasmaChem.H:74-184): value
 * written into the boundary ghost cell of bc_arr. */
static double species_bc_value(const vd_mech_desc* m, int spec, const vd_array4* phi,
                               const int face[3], int dir, int sgn, double Tg) {
    int ns = m->nspec, E = m->e_idx;
    int cell[3] = {face[0], face[1], face[2]};
    if (sgn == 1) cell[dir] -= 1;
    double on = (double)sgn;
    double ndens, emag;
    cell_nd_emag(phi, cell, ns, 3, &ndens, &emag);
    double Te = ATV(phi, cell, VD_ETEMP(ns));
    if (spec == E) {
        double ne = ATV(phi, cell, E);
        double D = spec_diff_st(m, spec, Te, Tg, ndens, emag);
        double cbar = sqrt(8.0 * K_B * Te / PI / ME);
        double flx = ne * cbar / 4.0 * on;
        return -flx / D;
    } else if (spec == ns) {
        double ne = ATV(phi, cell, E);
        double D = spec_diff_st(m, spec, Te, Tg, ndens, emag);
        double cbar = sqrt(8.0 * K_B * Te / PI / ME);
        double flx = (2.0 * K_B * Te) * ne * cbar / 4.0 * (double)sgn;
        return -flx / D;
    } else if (m->charge[spec] != 0) {
        double mu = spec_mob_st(m, spec, Te, Tg, ndens, emag);
        double D = spec_diff_st(m, spec, Te, Tg, ndens, emag);
        double efield = ATV(phi, cell, VD_EFX(ns) + dir);
        double idf = mu * ATV(phi, cell, spec) * efield;
        double dotn = idf * on;
        return (dotn >= 0.0) ? (-idf / D) : 0.0;
    } else {
        double D = spec_diff_st(m, spec, Te, Tg, ndens, emag);
        double cbar = sqrt(8.0 * K_B * Tg / PI / (m->molwt[spec] * M_AMU));
        double flx = ATV(phi, face, spec) * cbar / 4.0 * (double)sgn;
        return -flx / D;
    }
}

/* Vidyut::implicit_solve_scalar (Source/ScalarSolve.cpp:289-680) */
int or_implicit_solve_scalar(const vd_mech_desc* m, int spec, const vd_array4* sb,
                             const vd_array4* phi_old, const vd_array4* dsdt_expl,
                             vd_array4* phi_new, vd_array4 grad[3], const vd_box* box,
                             const vd_geom* g, const int bclo[3], const int bchi[3],
                             const vd_params* prm, double time, double dt, vd_solve_info* info) {
    (void)time;
    int ns = m->nspec, dim = g->dim, E = m->e_idx;
    int eflag = spec == E, enflag = spec == ns;
    double Tg = prm->gas_temperature;
    vd_array4 bcell = make_scratch(box, 1, 1, 0.0);
    vd_array4 specdata = make_scratch(box, 1, 1, 0.0);
    vd_array4 rhs = make_scratch(box, 0, 1, 0.0);
    vd_array4 sol = make_scratch(box, 1, 1, 0.0);
    /* D on the grown box (ScalarSolve.cpp:483-497) */
    for (int k = bcell.begin[2]; k < bcell.end[2]; k++)
        for (int j = bcell.begin[1]; j < bcell.end[1]; j++)
            for (int i = bcell.begin[0]; i < bcell.end[0]; i++) {
                int c[3] = {i, j, k};
                double nd, emag;
                cell_nd_emag(sb, c, ns, dim, &nd, &emag);
                AT(&bcell, i, j, k, 0) = spec_diff_st(m, spec, AT(sb, i, j, k, VD_ETEMP(ns)), Tg, nd, emag);
            }
    /* harmonic face average (ScalarSolve.cpp:511-512) */
    vd_array4 bf[3];
    const vd_array4* bptr[3] = {NULL, NULL, NULL};
    for (int d = 0; d < dim; d++) {
        bf[d] = make_face(box, d, 0.0);
        bptr[d] = &bf[d];
        int hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
        hi[d] += 1;
        for (int k = box->lo[2]; k <= hi[2]; k++)
            for (int j = box->lo[1]; j <= hi[1]; j++)
                for (int i = box->lo[0]; i <= hi[0]; i++) {
                    int l[3] = {i, j, k};
                    l[d] -= 1;
                    double a = ATV(&bcell, l, 0), b = AT(&bcell, i, j, k, 0);
                    AT(&bf[d], i, j, k, 0) = (a + b == 0.0) ? 0.0 : 2.0 * a * b / (a + b);
                }
    }
    /* specdata: Sborder comp with 1 ghost, then species_bc into boundary
     * ghosts of non-periodic directions (ScalarSolve.cpp:461-577) */
    for (int k = specdata.begin[2]; k < specdata.end[2]; k++)
        for (int j = specdata.begin[1]; j < specdata.end[1]; j++)
            for (int i = specdata.begin[0]; i < specdata.end[0]; i++)
                AT(&specdata, i, j, k, 0) = AT(sb, i, j, k, spec);
    for (int d = 0; d < dim; d++) {
        if (g->is_periodic[d]) continue;
        for (int side = 0; side < 2; side++) {
            int lo[3] = {box->lo[0], box->lo[1], box->lo[2]}, hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
            lo[d] = hi[d] = side ? box->hi[d] + 1 : box->lo[d];
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        int f[3] = {i, j, k}, gc[3] = {i, j, k};
                        if (!side) gc[d] -= 1;
                        ATV(&specdata, gc, 0) = species_bc_value(m, spec, sb, f, d, side ? 1 : -1, Tg);
                    }
        }
    }
    /* rhs = n_old/dt + dsdt; initial guess Sborder (ScalarSolve.cpp:443-465) */
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                AT(&rhs, i, j, k, 0) = (1.0 / dt) * AT(phi_old, i, j, k, spec) + 1.0 * AT(dsdt_expl, i, j, k, 0);
                AT(&sol, i, j, k, 0) = AT(sb, i, j, k, spec);
            }
    vd_mlmg_problem P;
    memset(&P, 0, sizeof(P));
    P.alpha_a = 1.0 / dt;
    P.beta = 1.0;
    for (int d = 0; d < 3; d++) {
        P.bface[d] = bptr[d];
        P.bc_lo[d] = bclo[d];
        P.bc_hi[d] = bchi[d];
    }
    P.bcval = &specdata;
    /* user_transport::species_bc Robin data (ScalarSolve.cpp:542-553) */
    vd_array4 rob[3];
    if (m->robin_set[spec]) {
        for (int q = 0; q < 3; q++) rob[q] = make_scratch(box, 1, 1, 0.0);
        for (int d = 0; d < dim; d++) {
            if (g->is_periodic[d]) continue;
            for (int side = 0; side < 2; side++) {
                int lo[3] = {box->lo[0], box->lo[1], box->lo[2]}, hi[3] = {box->hi[0], box->hi[1], box->hi[2]};
                lo[d] = hi[d] = side ? box->hi[d] + 1 : box->lo[d] - 1;
                for (int k = lo[2]; k <= hi[2]; k++)
                    for (int j = lo[1]; j <= hi[1]; j++)
                        for (int i = lo[0]; i <= hi[0]; i++)
                            for (int q = 0; q < 3; q++) AT(&rob[q], i, j, k, 0) = m->robin[spec][d][side][q];
            }
        }
        P.robin_a = &rob[0];
        P.robin_b = &rob[1];
        P.robin_f = &rob[2];
    }
    P.reltol = prm->linsolve_reltol;
    P.abstol = prm->linsolve_abstol;
    P.maxiter = prm->linsolve_maxiter;
    P.pre_smooth = prm->mg_pre_smooth;
    P.post_smooth = prm->mg_post_smooth;
    P.max_coarsening_level = prm->linsolve_max_coarsening_level;
    int st = or_mlmg_solve(&P, &sol, &rhs, box, g, info);
    /* floors (ScalarSolve.cpp:599-625) */
    if (prm->bound_specden && !enflag)
        for (int k = box->lo[2]; k <= box->hi[2]; k++)
            for (int j = box->lo[1]; j <= box->hi[1]; j++)
                for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                    double* v = a4p(&sol, i, j, k, 0);
                    double fl = eflag ? prm->min_electron_density : prm->min_species_density;
                    if (*v < fl) *v = fl;
                }
    if (eflag) grad_solution(&sol, grad, box, g, bclo, bchi, &specdata, NULL, NULL, NULL);
    for (int k = box->lo[2]; k <= box->hi[2]; k++)
        for (int j = box->lo[1]; j <= box->hi[1]; j++)
            for (int i = box->lo[0]; i <= box->hi[0]; i++)
                AT(phi_new, i, j, k, spec) = AT(&sol, i, j, k, 0);
    /* Te update (ScalarSolve.cpp:649-667) */
    if (enflag)
        for (int k = box->lo[2]; k <= box->hi[2]; k++)
            for (int j = box->lo[1]; j <= box->hi[1]; j++)
                for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                    double te = TWOTHIRD / K_B * AT(phi_new, i, j, k, VD_EEN(ns)) / AT(sb, i, j, k, E);
                    AT(phi_new, i, j, k, VD_ETEMP(ns)) = te;
                    if (te < prm->min_electron_temp) {
                        AT(phi_new, i, j, k, VD_ETEMP(ns)) = prm->min_electron_temp;
                        AT(phi_new, i, j, k, VD_EEN(ns)) =
                            1.5 * K_B * AT(phi_new, i, j, k, E) * prm->min_electron_temp;
                    }
                }
    free(bcell.p);
    free(specdata.p);
    free(rhs.p);
    free(sol.p);
    if (m->robin_set[spec])
        for (int q = 0; q < 3; q++) free(rob[q].p);
    for (int d = 0; d < dim; d++) free(bf[d].p);
    return st;
}

/* --------------------------------------------------------------- Evolve
 * Vidyut::Evolve (Source/Evolve.cpp:33-369), single level, fixed dt,
 * no output.  Sborder_old is only read in its valid region, so phi_old
 * is used directly. */
/* Vidyut::ComputeDt (Source/TimeStep.cpp:14-31) */
static double compute_dt(const vd_params* prm, double cur_time, double old_dt, const double ts[3]) {
    if (!(prm->adaptive_dt && cur_time > prm->adaptive_dt_delay)) return old_dt;
    double trans_min_dt = DBL_MAX;
    if (prm->do_transport) trans_min_dt = fmin(ts[0] * prm->advective_cfl, ts[1] * prm->diffusive_cfl);
    double adp = fmin(trans_min_dt, ts[2] * prm->dielectric_cfl);
    if (adp > prm->dt_max) adp = prm->dt_max;
    if (adp < prm->dt_min) adp = prm->dt_min;
    return fmin(old_dt * prm->dt_stretch, adp);
}

int or_evolve(const vd_mech_desc* m, vd_array4* phi_new, vd_array4* phi_old, const vd_box* box,
              const vd_geom* g, const vd_params* prm, int nsteps, double* dt_io, int ncorr,
              double* time) {
    double dt = *dt_io;
    int ns = m->nspec, nv = VD_NVAR(ns), dim = g->dim, E = m->e_idx;
    int ng = prm->hyp_order == 5 ? 3 : (prm->hyp_order == 2 ? 2 : 1);
    if (check_flux_args(prm)) return VD_EINVAL;
    if (!box_is_domain(box, g)) return VD_EINVAL;
    vd_array4 sb = make_scratch(box, ng, nv, 0.0);
    vd_array4 rxn = make_scratch(box, 0, ns + 1, 0.0);
    vd_array4 expl = make_scratch(box, 0, 1, 0.0);
    vd_array4 tmp = make_scratch(box, 0, nv, 0.0);
    vd_array4 flux[3], efield[3], gradne[3], gradfc[3];
    for (int d = 0; d < 3; d++) {
        int dd = d < dim ? d : 0;
        flux[d] = make_face(box, dd, 0.0);
        efield[d] = make_face(box, dd, 0.0);
        gradne[d] = make_face(box, dd, 0.0);
        gradfc[d] = make_face(box, dd, 0.0);
    }
    int st = VD_OK;
    double cur = *time;
    int64_t nvalid = (int64_t)(box->hi[0] - box->lo[0] + 1) * (box->hi[1] - box->lo[1] + 1) *
                     (box->hi[2] - box->lo[2] + 1);
    (void)nvalid;
    for (int step = 0; step < nsteps; step++) {
        double ts[3];
        or_find_time_scales(m, phi_new, box, g, prm, ts);
        dt = compute_dt(prm, cur, dt, ts);
        for (int q = 0; q < nv; q++)
            for (int k = box->lo[2]; k <= box->hi[2]; k++)
                for (int j = box->lo[1]; j <= box->hi[1]; j++)
                    for (int i = box->lo[0]; i <= box->hi[0]; i++)
                        AT(phi_old, i, j, k, q) = AT(phi_new, i, j, k, q);
        for (int niter = 0; niter < ncorr; niter++) {
            or_fill_patch(phi_new, &sb, box, g, 0, nv);
            vd_solve_info si;
            int s = or_solve_potential(m, &sb, phi_new, efield, box, g, prm, cur, &si);
            if (s && !st) st = s;
            or_fill_patch(phi_new, &sb, box, g, 0, nv);
            memset(rxn.p, 0, sizeof(double) * rxn.nstride * rxn.ncomp);
            if (prm->do_reactions) or_update_rxnsrc(m, &sb, &rxn, box, prm, cur);
            /* species loop order: electrons, energy, then others */
            int order[VD_MAX_SPECIES + 2], no = 0;
            order[no++] = E;
            if (prm->elecenergy_solve) order[no++] = ns;
            for (int ind = 0; ind < ns; ind++) order[no++] = ind + 1000; /* tagged */
            for (int q = 0; q < no; q++) {
                int spec = order[q];
                const int *blo, *bhi;
                vd_array4* gout = gradfc;
                if (spec >= 1000) {
                    int ind = spec - 1000;
                    if (ind == E || m->is_bg[ind]) {
                        if (prm->do_bg_reactions) {
                            for (int k = box->lo[2]; k <= box->hi[2]; k++)
                                for (int j = box->lo[1]; j <= box->hi[1]; j++)
                                    for (int i = box->lo[0]; i <= box->hi[0]; i++) {
                                        double* v = a4p(phi_new, i, j, k, ind);
                                        *v += AT(&rxn, i, j, k, ind) * dt;
                                        if (*v < prm->min_species_density && prm->bound_specden)
                                            *v = prm->min_species_density;
                                    }
                        }
                        continue;
                    }
                    spec = ind;
                    if (m->charge[ind] != 0) {
                        blo = prm->ion_bc_lo;
                        bhi = prm->ion_bc_hi;
                    } else {
                        blo = prm->neutral_bc_lo;
                        bhi = prm->neutral_bc_hi;
                    }
                } else if (spec == E) {
                    blo = prm->eden_bc_lo;
                    bhi = prm->eden_bc_hi;
                    gout = gradne;
                } else {
                    blo = prm->eenrg_bc_lo;
                    bhi = prm->eenrg_bc_hi;
                }
                /* update_explsrc_at_all_levels (ScalarSolve.cpp:57-111) */
                for (int d = 0; d < dim; d++)
                    memset(flux[d].p, 0, sizeof(double) * flux[d].nstride);
                if (prm->do_transport)
                    or_compute_scalar_transport_flux(m, &sb, flux, box, g, blo, bhi, prm, cur, dt, spec);
                or_compute_dsdt(flux, &rxn, &expl, box, g, spec);
                if (spec == ns)
                    or_compute_elecenergy_source(m, &sb, &rxn, efield, gradne, &expl, phi_new, box, g, prm);
                s = or_implicit_solve_scalar(m, spec, &sb, phi_old, &expl, phi_new, gout, box, g, blo,
                                             bhi, prm, cur, dt, &si);
                if (s && !st) st = s;
            }
            if (niter < ncorr - 1) {
                for (int q = 0; q < nv; q++)
                    for (int k = box->lo[2]; k <= box->hi[2]; k++)
                        for (int j = box->lo[1]; j <= box->hi[1]; j++)
                            for (int i = box->lo[0]; i <= box->hi[0]; i++)
                                AT(phi_new, i, j, k, q) =
                                    0.5 * AT(phi_old, i, j, k, q) + 0.5 * AT(phi_new, i, j, k, q);
            }
        }
        cur += dt;
    }
    *time = cur;
    *dt_io = dt;
    free(sb.p);
    free(rxn.p);
    free(expl.p);
    free(tmp.p);
    for (int d = 0; d < 3; d++) {
        free(flux[d].p);
        free(efield[d].p);
        free(gradne[d].p);
        free(gradfc[d].p);
    }
    return st;
}

/* ------------------------------------------------------ level transfers
 * Level l+1 refines level l by ratio r (2 or 4 per direction).  Fine cell
 * i lies in coarse cell floor(i / r). */
static int or_coarsen(int i, int r) { return i >= 0 ? i / r : -((-i - 1) / r) - 1; }

/* amrex::average_down as called by Vidyut::AverageDown
 * (Source/GridWork.cpp:94-108): coarse = (1/r^dim) * sum of the children,
 * summed k-outer, i-inner. */
int or_average_down(const vd_array4* fine, const vd_box* fbox, vd_array4* crse, int scomp, int ncomp,
                    const int ratio[3], int dim) {
    int r[3] = {ratio[0], ratio[1], dim == 3 ? ratio[2] : 1};
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = d < dim ? or_coarsen(fbox->lo[d], r[d]) : fbox->lo[d];
        hi[d] = d < dim ? or_coarsen(fbox->hi[d], r[d]) : fbox->hi[d];
    }
    double volfrac = 1.0 / (double)(r[0] * r[1] * r[2]);
    for (int n = scomp; n < scomp + ncomp; n++)
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    double sum = 0.0;
                    for (int kk = 0; kk < r[2]; kk++)
                        for (int jj = 0; jj < r[1]; jj++)
                            for (int ii = 0; ii < r[0]; ii++)
                                sum += AT(fine, i * r[0] + ii, j * r[1] + jj, k * r[2] + kk, n);
                    AT(crse, i, j, k, n) = volfrac * sum;
                }
    return VD_OK;
}

/* amrex::average_down_faces (Source/ScalarSolve.cpp:88-94): a coarse face
 * of direction d is the mean of the r^(dim-1) fine faces tiling it, the
 * higher tangential direction outermost. */
int or_average_down_faces(const vd_array4 fine[3], const vd_box* fbox, vd_array4 crse[3], const int ratio[3],
                          int dim) {
    int r[3] = {ratio[0], ratio[1], dim == 3 ? ratio[2] : 1};
    for (int d = 0; d < dim; d++) {
        int lo[3], hi[3];
        for (int e = 0; e < 3; e++) {
            int fl = fbox->lo[e], fh = fbox->hi[e] + (e == d);
            lo[e] = e < dim ? or_coarsen(fl, r[e]) : fl;
            hi[e] = e < dim ? or_coarsen(fh, r[e]) : fh;
        }
        /* tangential directions t1 < t2 (t2 absent in 2-D) */
        int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        double inv = 1.0 / (double)(r[t1] * r[t2]);
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    double sum = 0.0;
                    for (int b = 0; b < r[t2]; b++)
                        for (int a = 0; a < r[t1]; a++) {
                            int c[3] = {i * r[0], j * r[1], k * r[2]};
                            c[t1] += a;
                            c[t2] += b;
                            sum += ATV(&fine[d], c, 0);
                        }
                    AT(&crse[d], i, j, k, 0) = inv * sum;
                }
    }
    return VD_OK;
}

/* periodic wrap or clamp of index c into [lo, hi] */
static int or_wrap_clamp(int c, int lo, int hi, int per) {
    if (c >= lo && c <= hi) return c;
    if (per) {
        int n = hi - lo + 1, m = (c - lo) % n;
        if (m < 0) m += n;
        return lo + m;
    }
    return c < lo ? lo : hi;
}

/* cell_cons_interp (CellConservativeLinear without the cross-component
 * linear limiting) of coarse cell ic at fine offsets xo: MC-limited
 * central slopes, then a factor alpha <= 1 keeping every child within the
 * min/max of the 3^dim coarse neighbourhood.  Coarse neighbours wrap in
 * periodic directions and are clamped (foextrap) at domain faces. */
static double or_cons_lin(const vd_array4* C, const vd_geom* cg, int n, const int ic[3], const double xo[3],
                          const int r[3]) {
    int dim = cg->dim;
    double u0 = ATV(C, ic, n), s[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; d++) {
        int cp[3] = {ic[0], ic[1], ic[2]}, cm[3] = {ic[0], ic[1], ic[2]};
        cp[d] = or_wrap_clamp(ic[d] + 1, cg->dom_lo[d], cg->dom_hi[d], cg->is_periodic[d]);
        cm[d] = or_wrap_clamp(ic[d] - 1, cg->dom_lo[d], cg->dom_hi[d], cg->is_periodic[d]);
        double up = ATV(C, cp, n), um = ATV(C, cm, n);
        double cen = 0.5 * (up - um);
        double forw = 2.0 * (up - u0);
        double back = 2.0 * (u0 - um);
        double lim = 0.0;
        if (forw * back >= 0.0) lim = fabs(forw) < fabs(back) ? fabs(forw) : fabs(back);
        double mag = lim < fabs(cen) ? lim : fabs(cen);
        s[d] = (cen < 0.0 || (cen == 0.0 && signbit(cen)) ? -1.0 : 1.0) * mag;
    }
    double cmin = u0, cmax = u0;
    int kr = dim == 3 ? 1 : 0;
    for (int dk = -kr; dk <= kr; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                int c[3] = {or_wrap_clamp(ic[0] + di, cg->dom_lo[0], cg->dom_hi[0], cg->is_periodic[0]),
                            or_wrap_clamp(ic[1] + dj, cg->dom_lo[1], cg->dom_hi[1], cg->is_periodic[1]),
                            dim == 3 ? or_wrap_clamp(ic[2] + dk, cg->dom_lo[2], cg->dom_hi[2], cg->is_periodic[2])
                                     : ic[2]};
                double v = ATV(C, c, n);
                if (v < cmin) cmin = v;
                if (v > cmax) cmax = v;
            }
    double dlo = 0.0, dhi = 0.0;
    for (int c2 = 0; c2 < (dim == 3 ? r[2] : 1); c2++)
        for (int c1 = 0; c1 < r[1]; c1++)
            for (int c0 = 0; c0 < r[0]; c0++) {
                double y = ((c0 + 0.5) / r[0] - 0.5) * s[0] + ((c1 + 0.5) / r[1] - 0.5) * s[1];
                if (dim == 3) y = y + ((c2 + 0.5) / r[2] - 0.5) * s[2];
                if (y < dlo) dlo = y;
                if (y > dhi) dhi = y;
            }
    double alpha = 1.0;
    if (dlo != 0.0 && u0 + dlo < cmin) {
        double f = (cmin - u0) / dlo;
        if (f < alpha) alpha = f;
    }
    if (dhi != 0.0 && u0 + dhi > cmax) {
        double f = (cmax - u0) / dhi;
        if (f < alpha) alpha = f;
    }
    double v = u0 + xo[0] * (alpha * s[0]);
    v = v + xo[1] * (alpha * s[1]);
    if (dim == 3) v = v + xo[2] * (alpha * s[2]);
    return v;
}

/* FillPatchTwoLevels / InterpFromCoarseLevel (Source/GridWork.cpp:123-164):
 * fine valid data where a fine box covers the (wrapped) cell, coarse
 * interpolation elsewhere inside the domain, foextrap outside. */
int or_fill_patch_two_levels(const vd_array4* crse, const vd_geom* cg, const vd_array4* fine,
                             const vd_box* fboxes, int nfine, vd_array4* dst, const vd_geom* fg,
                             const int ratio[3], int scomp, int ncomp) {
    int dim = fg->dim;
    int r[3] = {ratio[0], ratio[1], dim == 3 ? ratio[2] : 1};
#pragma omp parallel for collapse(2)
    for (int k = dst->begin[2]; k < dst->end[2]; k++)
        for (int j = dst->begin[1]; j < dst->end[1]; j++)
            for (int i = dst->begin[0]; i < dst->end[0]; i++) {
                int c[3] = {i, j, k};
                for (int d = 0; d < dim; d++)
                    c[d] = or_wrap_clamp(c[d], fg->dom_lo[d], fg->dom_hi[d], fg->is_periodic[d]);
                int b = 0;
                for (; b < nfine; b++) {
                    int in = 1;
                    for (int d = 0; d < dim; d++)
                        if (c[d] < fboxes[b].lo[d] || c[d] > fboxes[b].hi[d]) in = 0;
                    if (in) break;
                }
                for (int n = scomp; n < scomp + ncomp; n++) {
                    if (b < nfine) {
                        AT(dst, i, j, k, n) = ATV(&fine[b], c, n);
                    } else {
                        int ic[3] = {c[0], c[1], c[2]};
                        double xo[3] = {0.0, 0.0, 0.0};
                        for (int d = 0; d < dim; d++) {
                            ic[d] = or_coarsen(c[d], r[d]);
                            xo[d] = (c[d] - ic[d] * r[d] + 0.5) / r[d] - 0.5;
                        }
                        AT(dst, i, j, k, n) = or_cons_lin(crse, cg, n, ic, xo, r);
                    }
                }
            }
    return VD_OK;
}

/* ------------------------------------------------ single-point hooks */
double or_species_bc_value(const vd_mech_desc* m, int spec, const vd_array4* phi, const int face[3], int dir,
                           int sgn, double Tg) {
    return species_bc_value(m, spec, phi, face, dir, sgn, Tg);
}
double or_gk(double w, double g) { return gk(w, g); }
double or_slope_lim(double c, double dl, double dm, double dr) { return slope_lim(c, dl, dm, dr); }
double or_applied_potential(double t, int domain_end, const vd_params* p) {
    return applied_potential(t, domain_end, p);
}
