// vd_chem.hip -- fused per-cell species + electron-energy sources and the
// Joule-heating source.  One thread per cell, FP64 throughout (rate
// exponentials included).  Per-thread source accumulators live in LDS
// columns (species-major, thread-minor: conflict-free, indexable by the
// table's species ids); densities are re-read through L1/L2.
#include <cstdlib>
#include <cstring>

#include "vd_internal.hpp"

namespace vd {

struct ChemArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb, rxn;
    double Tg;
    int3 lo;
    int nx, ny, nz;
    __device__ void rebase(int64_t o) {
        M = rb(M, o);
        vd::rebase(sb, o);
        vd::rebase(rxn, o);
    }
};

constexpr int CHEM_BLOCK = 128;

__device__ __forceinline__ double powi_nu(double n, double nu) {
    if (nu == 1.0) return n;
    if (nu == 2.0) return n * n;
    return pow(n, nu);
}

// Source/reaction_source.H:10-64 (table form)
__global__ void __launch_bounds__(CHEM_BLOCK) k_rxn_table(ChemArgs a, Batch bt) {
    a.rebase(case_off(bt));
    extern __shared__ double w[];
    const DevMech& M = *a.M;
    const int tid = threadIdx.x;
    const int ns = M.nspec;
    int64_t t = (int64_t)blockIdx.x * CHEM_BLOCK + tid;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    for (int sp = 0; sp <= ns; sp++) w[sp * CHEM_BLOCK + tid] = 0.0;
    const int o = a.sb.off(i, j, k);
    const double* __restrict__ sbp = a.sb.p;
    const int64_t nst = a.sb.ns;
    const double Te = sbp[(int64_t)VD_ETEMP(ns) * nst + o];
    const double lTe = log(Te), iTe = 1.0 / Te;
    const double lTg = log(a.Tg), iTg = 1.0 / a.Tg;
    const int nr = M.nreac;
    for (int r = 0; r < nr; r++) {
        const bool te = M.te[r] != 0;
        double kr = M.A[r] * exp(M.beta[r] * (te ? lTe : lTg) - M.Ea[r] * (te ? iTe : iTg));
        double conc = 1.0;
        const int nq = M.nreact[r];
        for (int q = 0; q < nq; q++) conc *= powi_nu(sbp[(int64_t)M.rsp[q][r] * nst + o], M.rnu[q][r]);
        const double rate = kr * conc;
        for (int q = 0; q < nq; q++) w[M.rsp[q][r] * CHEM_BLOCK + tid] -= M.rnu[q][r] * rate;
        const int np = M.nprod[r];
        for (int q = 0; q < np; q++) w[M.psp[q][r] * CHEM_BLOCK + tid] += M.pnu[q][r] * rate;
        w[ns * CHEM_BLOCK + tid] += M.energy[r] * rate;
    }
    const int orx = a.rxn.off(i, j, k);
    for (int sp = 0; sp <= ns; sp++) a.rxn.p[(int64_t)sp * a.rxn.ns + orx] = w[sp * CHEM_BLOCK + tid];
}

// CKWC / productionRate / comp_ener_exch of a generated Chemistry.H
// (test/models/ionelec_DC_1d/Chemistry.H:240-357) behind
// update_rxnsrc_at_all_levels (Source/ScalarSolve.cpp:143-165).
__global__ void __launch_bounds__(CHEM_BLOCK) k_rxn_ckwc(ChemArgs a, Batch bt) {
    a.rebase(case_off(bt));
    extern __shared__ double w[];
    const DevMech& M = *a.M;
    const int tid = threadIdx.x;
    const int ns = M.nspec, E = M.e_idx;
    int64_t t = (int64_t)blockIdx.x * CHEM_BLOCK + tid;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    for (int sp = 0; sp < ns; sp++) w[sp * CHEM_BLOCK + tid] = 0.0;
    const int o = a.sb.off(i, j, k);
    const double* __restrict__ sbp = a.sb.p;
    const int64_t nst = a.sb.ns;
    // C in mol/m^3: (n * 1e-6 / N_A) * 1e6
#define CONC(sp) ((sbp[(int64_t)(sp) * nst + o] * 1.0e-6 / VD_N_A) * 1.0e6)
    const double Te = sbp[(int64_t)VD_ETEMP(ns) * nst + o];
    const double T = a.Tg;
    const double lTe = log(Te), iTe = 1.0 / Te, lTg = log(T), iTg = 1.0 / T;
    const double CE = CONC(E);
    const double ne_pr = CE * 6.02214085774e23;
    const double Ue = 1.5 * Te * ne_pr * 1.380649e-23;
    const double eV2J = 1.602176620800000e-19, NA = 6.02214085774e23, kB = 1.380649e-23;
    const double mee = Ue / (CE * NA);
    const double sefact = (tanh((Te - T) / 0.5) + 1.0) / 2.0;
    double ener = 0.0;
    const int nr = M.nreac;
    for (int r = 0; r < nr; r++) {
        const bool te = M.te[r] != 0;
        double k_f = M.A[r] * exp(M.beta[r] * (te ? lTe : lTg) - M.Ea[r] * (te ? iTe : iTg));
        const int ord = M.order[r];
        for (int q = 1; q < ord; q++) k_f *= 6.02214076e23;
        double cprod = 1.0;
        bool first = true;
        const int nq = M.nreact[r];
        for (int q = 0; q < nq; q++) {
            const double c = CONC(M.rsp[q][r]);
            const int nu = (int)M.rnu[q][r];
            for (int e = 0; e < nu; e++) {
                cprod = first ? c : cprod * c;
                first = false;
            }
        }
        const double qf = k_f * cprod;
        const double qdot = qf - 0.0;
        const int nn = M.nnet[r];
        for (int q = 0; q < nn; q++) {
            const double net = M.netnu[q][r];
            double* wp = &w[M.netsp[q][r] * CHEM_BLOCK + tid];
            if (net == 1.0) *wp += qdot;
            else if (net == -1.0) *wp -= qdot;
            else *wp += net * qdot;
        }
        const int rt = M.rxntype[r];
        if (rt == 1 || rt == 2) ener -= (M.eexci[r] * eV2J) * NA * (qf - 0.0) * sefact;
        if (rt == 3) ener -= mee * NA * (qf - 0.0);
        if (rt == 5) {
            const int el = M.elidx[r];
            double qq = CE * CONC(el) * k_f;
            ener -= 3.0 * kB * (M.molwt[E] / M.molwt[el]) * (Te - T) * qq * NA;
        }
    }
    if (M.elastic_loss) {
        // test/models/ionelec_DC_1d/Chemistry.H:334-341
        double ndens = 0.0;
        for (int sp = 0; sp < ns; sp++) ndens += CONC(sp) * VD_N_A;
        double mu = spec_mob_st(M, E, Te, T, ndens, 0.0);
        double nu = -VD_ECHARGE / VD_ME / mu;
        double term = 3.0 / 2.0 * VD_K_B * CE * VD_N_A * (Te - T) * nu *
                      (2.0 * VD_ME / (M.elastic_partner_amu * VD_M_AMU));
        ener -= term;
    }
#undef CONC
    const int orx = a.rxn.off(i, j, k);
    for (int sp = 0; sp < ns; sp++)
        a.rxn.p[(int64_t)sp * a.rxn.ns + orx] = ((w[sp * CHEM_BLOCK + tid] * 1.0e-6) * VD_N_A) * 1.0e6;
    a.rxn.p[(int64_t)ns * a.rxn.ns + orx] = ener;
}

// VD_CHEM_GEN prologue: the k_f of every VD_RATE_ARRH reaction and g/RT of
// every species at the gas temperature T (one block; the generated
// productionRate computes them per cell from the same T,
// test/Reactor/CO2_H2_Ar_mechanism_reduced/Chemistry.H:913-950)
__global__ void __launch_bounds__(256) k_gen_tconst(const DevMech* __restrict__ Mp, double T,
                                                    double* __restrict__ tc, Batch bt) {
    Mp = rb(Mp, case_off(bt));
    tc = rb(tc, case_off(bt));
    const DevMech& M = *Mp;
    const double invT = 1.0 / T;
    const double logT = log(T / 300.0);
    for (int r = threadIdx.x; r < M.nreac; r += blockDim.x)
        tc[r] = M.rkind[r] == VD_RATE_ARRH ? M.A[r] * exp(M.beta[r] * logT - M.Ea[r] * invT) : 0.0;
    for (int k = threadIdx.x; k < M.nspec; k += blockDim.x) tc[VD_MAX_REACTIONS + k] = gen_gibbs(M, k, T);
}

// k_f of a VD_CHEM_GEN reaction at one cell (the T-only kinds from tc)
__device__ __forceinline__ double gen_kf(const DevMech& M, int r, const double* __restrict__ tc, double Te,
                                         double T, double TeeV, const double* Te_pow, const double* invTe_pow) {
    double k_f, sum = 0.0;
    switch (M.rkind[r]) {
    case VD_RATE_ARRH:
        k_f = tc[r];
        break;
    case VD_RATE_TEPOW:
        k_f = M.A[r] * pow(Te / M.tref[r], M.beta[r]);
        if (M.tgdiv[r]) k_f = k_f / T;
        break;
    case VD_RATE_TEARRH:
        k_f = M.A[r] * pow(Te, M.beta[r]) * exp(-M.Ea[r] / Te);
        break;
    case VD_RATE_JANEV:
        for (int j = 0; j < 9; j++) sum += M.fit[j][r] * Te_pow[j];
        k_f = M.A[r] * exp(sum) * 6.02214085774e23;
        break;
    case VD_RATE_FFIT:
        for (int j = 0; j < 4; j++) sum += M.fit[j][r] * invTe_pow[j];
        k_f = M.A[r] * exp(sum) * 6.02214085774e23;
        break;
    default:
        if (TeeV < M.fsw[r]) {
            for (int j = 0; j < 9; j++) sum += M.fit[j][r] * Te_pow[j];
            k_f = M.A[r] * exp(sum) * 6.02214085774e23;
        } else {
            for (int j = 0; j < 4; j++) sum += M.fit2[j][r] * invTe_pow[j];
            k_f = M.A2[r] * exp(sum) * 6.02214085774e23;
        }
        break;
    }
    if (M.kmax[r] > 0.0 && M.kmax[r] < k_f) k_f = M.kmax[r];
    return k_f;
}

// CKWC / productionRate / comp_ener_exch of a mechanism read from its
// generated Chemistry.H (test/Reactor/CO2_H2_Ar_mechanism_reduced/
// Chemistry.H:841-6399) behind update_rxnsrc_at_all_levels
// (Source/ScalarSolve.cpp:143-165): the generic form, reading the reaction
// records from the device descriptor; accumulators in LDS columns
__global__ void __launch_bounds__(CHEM_BLOCK) k_rxn_gen(ChemArgs a, const double* __restrict__ tc, Batch bt) {
    a.rebase(case_off(bt));
    tc = rb(tc, case_off(bt));
    extern __shared__ double w[];
    const DevMech& M = *a.M;
    const int tid = threadIdx.x;
    const int ns = M.nspec, E = M.e_idx;
    int64_t t = (int64_t)blockIdx.x * CHEM_BLOCK + tid;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    for (int sp = 0; sp < ns; sp++) w[sp * CHEM_BLOCK + tid] = 0.0;
    const int o = a.sb.off(i, j, k);
    const double* __restrict__ sbp = a.sb.p;
    const int64_t nst = a.sb.ns;
#define CONC(sp) ((sbp[(int64_t)(sp) * nst + o] * 1.0e-6 / VD_N_A) * 1.0e6)
    const double Te = sbp[(int64_t)VD_ETEMP(ns) * nst + o];
    const double T = a.Tg;
    const double invT = 1.0 / T;
    const double refC = 101325 / 8.31446 * invT;
    const double refCinv = 1 / refC;
    double mixture = 0.0;
    if (M.any_tb)
        for (int sp = 0; sp < ns; sp++) mixture += CONC(sp);
    const double CE = CONC(E);
    const double ne_pr = CE * 6.02214085774e23;
    const double Ue = 1.5 * Te * ne_pr * 1.380649e-23;
    const double TeeV = Te / 11595.0;
    double Te_pow[9], invTe_pow[4];
    if (M.any_fit) {
        const double logTe = log(TeeV);
        const double invTeeV = (Te == 0) ? 1.0 : 1.0 / (TeeV);
        for (int q = 0; q < 9; q++) Te_pow[q] = pow(logTe, q);
        for (int q = 0; q < 4; q++) invTe_pow[q] = pow(invTeeV, q + 1);
    }
    const double eV2J = 1.602176620800000e-19, NA = 6.02214085774e23, kB = 1.380649e-23;
    const double mee = Ue / (CE * NA);
    const double sefact = (tanh((Te - T) / 0.5) + 1.0) / 2.0;
    double ener = 0.0;
    const int nr = M.nreac;
    for (int r = 0; r < nr; r++) {
        const double k_f = gen_kf(M, r, tc, Te, T, TeeV, Te_pow, invTe_pow);
        const double pre = M.tb[r] ? mixture * k_f : k_f;
        double cprod = 1.0;
        bool first = true;
        for (int q = 0; q < M.nreact[r]; q++) {
            const double c = CONC(M.rsp[q][r]);
            const int nu = (int)M.rnu[q][r];
            for (int e = 0; e < nu; e++) {
                cprod = first ? c : cprod * c;
                first = false;
            }
        }
        const double qf = pre * cprod;
        double qr = 0.0;
        if (M.rev[r]) {
            double s = 0.0;
            for (int q = 0; q < M.ng[r]; q++) {
                const double g = tc[VD_MAX_REACTIONS + M.gsp[q][r]], nu = M.gnu[q][r];
                const double tm = nu == 1.0 ? g : (nu == -1.0 ? -g : nu * g);
                s = q == 0 ? tm : s + tm;
            }
            double v = pre * exp(-(s));
            for (int p = 0; p < M.refcp[r]; p++) v = v * refC;
            for (int p = 0; p < -M.refcp[r]; p++) v = v * refCinv;
            double rp = 1.0;
            first = true;
            for (int q = 0; q < M.nrev[r]; q++) {
                const double c = CONC(M.revsp[q][r]);
                const int nu = (int)M.revnu[q][r];
                for (int e = 0; e < nu; e++) {
                    rp = first ? c : rp * c;
                    first = false;
                }
            }
            qr = v * rp;
        }
        const double qdot = qf - qr;
        for (int q = 0; q < M.nwd[r]; q++) {
            const double nu = M.wdnu[q][r];
            double* wp = &w[M.wdsp[q][r] * CHEM_BLOCK + tid];
            if (nu == 1.0) *wp += qdot;
            else if (nu == -1.0) *wp -= qdot;
            else if (nu > 0.0) *wp += nu * qdot;
            else *wp -= (-nu) * qdot;
        }
        const int rt = M.rxntype[r];
        if (rt == 1 || rt == 2) ener -= (M.eexci[r] * eV2J) * NA * (qf - qr) * sefact;
        if (rt == 3 || rt == 4) ener -= mee * NA * (qf - qr);
        if (rt == 5) {
            const int el = M.elidx[r];
            const double qq = CE * CONC(el) * k_f;
            ener -= 3.0 * kB * (M.molwt[E] / M.molwt[el]) * (Te - T) * qq * NA;
        }
        if (rt == 7) ener += mee * NA * (qf - qr);
    }
#undef CONC
    const int orx = a.rxn.off(i, j, k);
    for (int sp = 0; sp < ns; sp++)
        a.rxn.p[(int64_t)sp * a.rxn.ns + orx] = ((w[sp * CHEM_BLOCK + tid] * 1.0e-6) * VD_N_A) * 1.0e6;
    a.rxn.p[(int64_t)ns * a.rxn.ns + orx] = ener;
}

// user_sources::add_user_react_sources (Source/ScalarSolve.cpp:158-163):
// the case's constant sources added to every component of rxn_src
__global__ void __launch_bounds__(256) k_user_src(ChemArgs a, Batch bt) {
    a.rebase(case_off(bt));
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    const DevMech& M = *a.M;
    const int o = a.rxn.off(a.lo.x + i, a.lo.y + j, a.lo.z + k);
    for (int sp = 0; sp <= M.nspec; sp++)
        if (M.user_src[sp] != 0.0) a.rxn.p[(int64_t)sp * a.rxn.ns + o] += M.user_src[sp];
}

static int add_user_src(vd_ctx* ctx, const vd_mech* m, const ChemArgs& a) {
    if (!m->hdev.any_user_src) return VD_OK;
    hipLaunchKernelGGL(k_user_src, bz(ctx, grid1d((int64_t)a.nx * a.ny * a.nz, 256)), dim3(256), 0, ctx->stream, a,
                       ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

int update_rxnsrc(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4* rxn,
                  const vd_box* box, const vd_params* prm) {
    if (rxn->ncomp < m->hdev.nspec + 1) {
        set_error("rxn_src needs NUM_SPECIES+1 components");
        return VD_EINVAL;
    }
    ChemArgs a;
    a.M = m->dev;
    a.sb = arr(sb);
    a.rxn = arr(rxn);
    a.Tg = prm->gas_temperature;
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    size_t lds = sizeof(double) * (m->hdev.nspec + 1) * CHEM_BLOCK;
    if (m->hdev.chem_form == VD_CHEM_GEN) {
        double* tc = nullptr;
        VD_TRY(scratch(ctx, "gen_tconst", sizeof(double) * GEN_TCONST_DOUBLES, (void**)&tc));
        kt_begin(ctx, VD_KT_CHEM);
        hipLaunchKernelGGL(k_gen_tconst, bz(ctx, dim3(1)), dim3(256), 0, ctx->stream, m->dev, a.Tg, tc, ctx->bt);
        if (m->use_spec && m->fn) {
            int st = chem_jit_launch(ctx, m, sb, rxn, box, prm->gas_temperature, tc);
            kt_end(ctx, VD_KT_CHEM, 8.0 * (double)tot * (2 * m->hdev.nspec + 3));
            if (st != VD_OK) return st;
        } else {
            hipLaunchKernelGGL(k_rxn_gen, bz(ctx, grid1d(tot, CHEM_BLOCK)), dim3(CHEM_BLOCK), lds, ctx->stream, a, tc,
                               ctx->bt);
            kt_end(ctx, VD_KT_CHEM, 8.0 * (double)tot * (2 * m->hdev.nspec + 3));
        }
        VD_CHECK_LAUNCH();
        return add_user_src(ctx, m, a);
    }
    kt_begin(ctx, VD_KT_CHEM);
    if (m->hdev.chem_form == VD_CHEM_TABLE && m->use_spec && m->fn) {
        int st = chem_jit_launch(ctx, m, sb, rxn, box, prm->gas_temperature, nullptr);
        kt_end(ctx, VD_KT_CHEM, 8.0 * (double)tot * (2 * m->hdev.nspec + 3));
        if (st != VD_OK) return st;
        return add_user_src(ctx, m, a);
    }
    if (m->hdev.chem_form == VD_CHEM_TABLE)
        hipLaunchKernelGGL(k_rxn_table, bz(ctx, grid1d(tot, CHEM_BLOCK)), dim3(CHEM_BLOCK), lds, ctx->stream, a,
                           ctx->bt);
    else
        hipLaunchKernelGGL(k_rxn_ckwc, bz(ctx, grid1d(tot, CHEM_BLOCK)), dim3(CHEM_BLOCK), lds, ctx->stream, a,
                           ctx->bt);
    kt_end(ctx, VD_KT_CHEM, 8.0 * (double)tot * (2 * m->hdev.nspec + 3));
    VD_CHECK_LAUNCH();
    return add_user_src(ctx, m, a);
}

struct JouleArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb, rxn, ef0, ef1, ef2, gn0, gn1, gn2, dsdt, phi;
    double Tg, ele_mob, ele_diff;
    int const_ele_trans, floor_jh, dim;
    int state_trans;  // the electron mobility or diffusivity is a fit of the state
    int3 dlo, dhi, lo;
    int nx, ny, nz;
    __device__ void rebase(int64_t o) {
        M = rb(M, o);
        vd::rebase(sb, o);
        vd::rebase(rxn, o);
        vd::rebase(ef0, o);
        vd::rebase(ef1, o);
        vd::rebase(ef2, o);
        vd::rebase(gn0, o);
        vd::rebase(gn1, o);
        vd::rebase(gn2, o);
        vd::rebase(dsdt, o);
        vd::rebase(phi, o);
    }
};

// the current density of the electrons on one face times e (charge *
// (mu n_e E - D grad n_e)) from the face's two cells ol, orr (offsets in
// the state) and its E and grad n_e (Source/ElecEnergySource.cpp:84-142);
// Te and n_e of the two cells come in as values
template <bool STATE, int DIM>
__device__ __forceinline__ double joule_cur(const JouleArgs& a, const DevMech& M, const double* S, int64_t N,
                                            double tl, double tr, double nl, double nr, double efield_face,
                                            double gradne_face, int ol, int orr, double charge) {
    const int ns = M.nspec, eidx = M.e_idx;
    double etemp = 0.5 * (tl + tr);
    double ne = 0.5 * (nl + nr);
    double mu, dcoeff;
    if (a.const_ele_trans || STATE) {
        double ndens = 0.0;
        for (int sp = 0; sp < ns; sp++) ndens += 0.5 * (S[(int64_t)sp * N + ol] + S[(int64_t)sp * N + orr]);
        if (!STATE || a.const_ele_trans) {
            mu = a.ele_mob / ndens;
            dcoeff = a.ele_diff / ndens;
        } else {
            // specMob / specDiff of the face-averaged state
            // (ElecEnergySource.cpp:102-131)
            double es = 0.0;
            for (int q = 0; q < DIM; q++) {
                const double e = 0.5 * (S[(int64_t)(VD_EFX(ns) + q) * N + ol] +
                                        S[(int64_t)(VD_EFX(ns) + q) * N + orr]);
                es += e * e;
            }
            mu = spec_mob_st(M, eidx, etemp, a.Tg, ndens, sqrt(es));
            dcoeff = spec_diff_st(M, eidx, etemp, a.Tg, ndens, sqrt(es));
        }
    } else {
        mu = spec_mob(M, eidx);
        dcoeff = spec_diff(M, eidx, etemp, a.Tg);
    }
    return charge * (mu * ne * efield_face - dcoeff * gradne_face);
}

// the face of direction idim and side f of cell (i, j, k) whose E and grad
// n_e a cell's Joule heating takes -- on a domain wall the face next to it
// (ElecEnergySource.cpp:74-83) -- and its left and right cells
__device__ __forceinline__ void joule_face_cells(const JouleArgs& a, int idim, int f, int i, int j, int k,
                                                 int face[3], int& ol, int& orr) {
    const int dlo[3] = {a.dlo.x, a.dlo.y, a.dlo.z}, dhi[3] = {a.dhi.x, a.dhi.y, a.dhi.z};
    face[0] = i;
    face[1] = j;
    face[2] = k;
    face[idim] += f;
    if (face[idim] == dlo[idim]) face[idim] += 1;
    if (face[idim] == dhi[idim] + 1) face[idim] -= 1;
    int lc[3] = {face[0], face[1], face[2]};
    lc[idim] -= 1;
    ol = a.sb.off(lc[0], lc[1], lc[2]);
    orr = a.sb.off(face[0], face[1], face[2]);
}

// Vidyut::compute_elecenergy_source (Source/ElecEnergySource.cpp:65-171)
// STATE: the electron coefficients are fits of the face state (else
// constants / Te only; a.state_trans mirrors it).  DIM: the dimension.
// The faces are taken one direction at a time (an unrolled face loop holds
// every face's addresses at once and runs at lower occupancy, no faster;
// a z-streamed form that keeps the column in registers ran slower still,
// its chain of dependent loads per thread exposed)
template <bool STATE, int DIM>
__global__ void __launch_bounds__(256) k_joule(JouleArgs a, Batch bt) {
    // a stays unmodified; the case's offset is applied where each array is read
    const int64_t co = case_off(bt);
    const DevMech& M = *rb(a.M, co);
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    const int ns = M.nspec, eidx = M.e_idx;
    const double charge = M.charge[eidx] * VD_ECHARGE;
    const double* S = rb(a.sb.p, co);
    const int64_t N = a.sb.ns;
    double jheat = 0.0;
#pragma unroll 1
    for (int idim = 0; idim < DIM; idim++) {
        const Arr<double>& EF = idim == 0 ? a.ef0 : (idim == 1 ? a.ef1 : a.ef2);
        const Arr<double>& GN = idim == 0 ? a.gn0 : (idim == 1 ? a.gn1 : a.gn2);
#pragma unroll
        for (int f = 0; f < 2; f++) {
            int face[3], ol, orr;
            joule_face_cells(a, idim, f, i, j, k, face, ol, orr);
            const double efield_face = at(EF, co, face[0], face[1], face[2]);
            const double gradne_face = at(GN, co, face[0], face[1], face[2]);
            const double cur = joule_cur<STATE, DIM>(a, M, S, N, S[VD_ETEMP(ns) * N + ol], S[VD_ETEMP(ns) * N + orr],
                                                     S[(int64_t)eidx * N + ol], S[(int64_t)eidx * N + orr],
                                                     efield_face, gradne_face, ol, orr, charge);
            jheat += cur * efield_face;
        }
    }
    jheat *= 0.5;
    if (a.floor_jh && jheat < 0.0) jheat = 0.0;
    at(a.dsdt, co, i, j, k) += jheat;
    at(a.phi, co, i, j, k, VD_EJH(ns)) = jheat;
    at(a.phi, co, i, j, k, VD_EIH(ns)) = at(a.rxn, co, i, j, k, ns);
}

int elecenergy_source(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, const vd_array4* rxn,
                      const vd_array4 ef[3], const vd_array4 gn[3], vd_array4* ds, vd_array4* phi,
                      const vd_box* box, const vd_geom* g, const vd_params* prm) {
    JouleArgs a;
    a.M = m->dev;
    a.sb = arr(sb);
    a.rxn = arr(rxn);
    a.ef0 = arr(&ef[0]);
    a.ef1 = g->dim > 1 ? arr(&ef[1]) : a.ef0;
    a.ef2 = g->dim > 2 ? arr(&ef[2]) : a.ef0;
    a.gn0 = arr(&gn[0]);
    a.gn1 = g->dim > 1 ? arr(&gn[1]) : a.gn0;
    a.gn2 = g->dim > 2 ? arr(&gn[2]) : a.gn0;
    a.dsdt = arr(ds);
    a.phi = arr(phi);
    a.Tg = prm->gas_temperature;
    a.ele_mob = prm->ele_mob;
    a.ele_diff = prm->ele_diff;
    a.const_ele_trans = prm->const_ele_trans;
    a.state_trans = m->hdev.mfit_kind[m->hdev.e_idx] == VD_TR_FIT || m->hdev.dfit_kind[m->hdev.e_idx] == VD_TR_FIT;
    a.floor_jh = prm->floor_jh;
    a.dim = g->dim;
    a.dlo = make_int3(g->dom_lo[0], g->dom_lo[1], g->dom_lo[2]);
    a.dhi = make_int3(g->dom_hi[0], g->dom_hi[1], g->dom_hi[2]);
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
#define VD_JOULE(ST, D) hipLaunchKernelGGL((k_joule<ST, D>), bz(ctx, grid1d(tot, 256)), dim3(256), 0, ctx->stream, a, ctx->bt)
    if (a.state_trans) {
        if (a.dim == 3) VD_JOULE(true, 3);
        else if (a.dim == 2) VD_JOULE(true, 2);
        else VD_JOULE(true, 1);
    } else {
        if (a.dim == 3) VD_JOULE(false, 3);
        else if (a.dim == 2) VD_JOULE(false, 2);
        else VD_JOULE(false, 1);
    }
#undef VD_JOULE
    VD_CHECK_LAUNCH();
    return VD_OK;
}

}  // namespace vd

using namespace vd;

extern "C" int vd_update_rxnsrc(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                vd_array4* rxn_src, const vd_box* box, const vd_params* prm,
                                double time) {
    (void)time;
    if (!ctx || !m || !sborder || !rxn_src || !box || !prm) return VD_EINVAL;
    VD_TRY(check_extents("update_rxnsrc", box, {sborder, rxn_src}));
    return update_rxnsrc(ctx, m, sborder, rxn_src, box, prm);
}

extern "C" int vd_compute_elecenergy_source(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                            const vd_array4* rxn_src, const vd_array4 efield[3],
                                            const vd_array4 gradne[3], vd_array4* dsdt,
                                            vd_array4* phi_new, const vd_box* box,
                                            const vd_geom* geom, const vd_params* prm) {
    if (!ctx || !m || !sborder || !rxn_src || !efield || !gradne || !dsdt || !phi_new || !box ||
        !geom || !prm)
        return VD_EINVAL;
    VD_TRY(check_extents("compute_elecenergy_source", box, {sborder, rxn_src, dsdt, phi_new}));
    VD_TRY(check_face_extents("compute_elecenergy_source", efield, geom->dim));
    VD_TRY(check_face_extents("compute_elecenergy_source", gradne, geom->dim));
    return elecenergy_source(ctx, m, sborder, rxn_src, efield, gradne, dsdt, phi_new, box, geom, prm);
}
