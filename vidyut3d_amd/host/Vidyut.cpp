// Vidyut.cpp -- host driver over the C ABI: the reference's member-function
// boundary and step loop (Source/Evolve.cpp, ScalarSolve.cpp,
// PotentialSolve.cpp, ElecEnergySource.cpp, TimeStep.cpp, Vidyut.cpp,
// GridWork.cpp, Plot.cpp) on device-resident single-box MultiFabs.
#include "Vidyut.H"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <sstream>

#include "../../include/vidyut3d_amd_host.h"

namespace vidyut_amd {

#define TRY(x)                 \
    do {                       \
        int s_ = (x);          \
        if (s_ != VD_OK) return s_; \
    } while (0)

// ------------------------------------------------------------- MultiFab
MultiFab::MultiFab(vd_ctx* ctx, const vd_box& valid, int dim, int ncomp, int ngrow, int face_dir) {
    define(ctx, valid, dim, ncomp, ngrow, face_dir);
}

MultiFab::~MultiFab() { release(); }

void MultiFab::release() {
    if (a_.p && ctx_) vd_free(ctx_, a_.p);
    a_.p = nullptr;
}

MultiFab::MultiFab(MultiFab&& o) noexcept {
    ctx_ = o.ctx_;
    a_ = o.a_;
    valid_ = o.valid_;
    ngrow_ = o.ngrow_;
    o.a_.p = nullptr;
}

MultiFab& MultiFab::operator=(MultiFab&& o) noexcept {
    if (this != &o) {
        release();
        ctx_ = o.ctx_;
        a_ = o.a_;
        valid_ = o.valid_;
        ngrow_ = o.ngrow_;
        o.a_.p = nullptr;
    }
    return *this;
}

void MultiFab::define(vd_ctx* ctx, const vd_box& valid, int dim, int ncomp, int ngrow, int face_dir) {
    release();
    ctx_ = ctx;
    valid_ = valid;
    ngrow_ = ngrow;
    if (face_dir >= 0) valid_.hi[face_dir] += 1;
    for (int d = 0; d < 3; d++) {
        int g = d < dim ? ngrow : 0;
        a_.begin[d] = valid_.lo[d] - g;
        a_.end[d] = valid_.hi[d] + 1 + g;
    }
    a_.jstride = a_.end[0] - a_.begin[0];
    a_.kstride = a_.jstride * (a_.end[1] - a_.begin[1]);
    a_.nstride = a_.kstride * (a_.end[2] - a_.begin[2]);
    a_.ncomp = ncomp;
    void* p = nullptr;
    if (vd_malloc(ctx, (size_t)a_.nstride * ncomp * sizeof(double), &p) == VD_OK) {
        a_.p = (double*)p;
        vd_memset_zero(ctx, p, (size_t)a_.nstride * ncomp * sizeof(double));
    }
}

int MultiFab::setVal(double v) {
    vd_box whole;
    for (int d = 0; d < 3; d++) {
        whole.lo[d] = a_.begin[d];
        whole.hi[d] = a_.end[d] - 1;
    }
    return vd_setval(ctx_, &a_, &whole, 0, a_.ncomp, v);
}

int MultiFab::copy_from(const MultiFab& src, int scomp, int dcomp, int n) {
    return vd_copy(ctx_, &src.array(), &a_, &valid_, scomp, dcomp, n);
}

// ---------------------------------------------------------------- Vidyut
Vidyut::Vidyut(vd_ctx* ctx, const vd_mech_desc& mech, const vd_geom& geom, const vd_params& prm)
    : ctx_(ctx), desc_(mech), geom_(geom), prm_(prm) {
    ns_ = mech.nspec;
    nvar_ = VD_NVAR(ns_);
    dim_ = geom.dim;
    for (int d = 0; d < 3; d++) {
        dom_.lo[d] = d < dim_ ? geom.dom_lo[d] : 0;
        dom_.hi[d] = d < dim_ ? geom.dom_hi[d] : 0;
    }
    // ngrow_for_fillpatch by hyp_order (Source/Vidyut.cpp:318-335)
    if (prm.hyp_order == 1) ngrow_fp_ = 1;
    else if (prm.hyp_order == 2) ngrow_fp_ = 2;
    else if (prm.hyp_order == 5) ngrow_fp_ = 3;
    else {
        status_ = VD_EINVAL;
        return;
    }
    status_ = vd_mech_create(ctx, &mech, &mech_);
    if (status_ != VD_OK) return;
    phi_new_.emplace_back(ctx, dom_, dim_, nvar_, 0);
    phi_old_.emplace_back(ctx, dom_, dim_, nvar_, 0);
    status_ = alloc_step_data();
}

Vidyut::~Vidyut() {
    if (mech_) vd_mech_destroy(mech_);
}

int Vidyut::alloc_step_data() {
    Sborder_.emplace_back(ctx_, dom_, dim_, nvar_, ngrow_fp_);
    Swork_.emplace_back(ctx_, dom_, dim_, nvar_, ngrow_fp_);
    expl_src_.emplace_back(ctx_, dom_, dim_, 1, 0);
    rxn_src_.emplace_back(ctx_, dom_, dim_, ns_ + 1, 0);
    for (auto* v : {&flux_, &efield_ec_, &gradne_fc_, &grad_fc_}) {
        v->emplace_back();
        for (int d = 0; d < 3; d++) (*v)[0][d].define(ctx_, dom_, dim_, 1, 0, d < dim_ ? d : 0);
    }
    if (!phi_new_[0].ok() || !Sborder_[0].ok() || !Swork_[0].ok()) return VD_ENOMEM;
    return VD_OK;
}

std::vector<std::string> Vidyut::varnames() const {
    static const char* pv[VD_NUM_PLASMAVARS] = {"Electron_energy",   "Electron_Temp", "Potential",
                                                "Efieldx",           "Efieldy",       "Efieldz",
                                                "Electron_Jheat",    "Electron_inelasticHeat",
                                                "Electron_elasticHeat", "ReducedEF"};
    std::vector<std::string> v;
    for (int s = 0; s < ns_; s++) v.emplace_back(desc_.names[s]);
    for (int q = 0; q < VD_NUM_PLASMAVARS; q++) v.emplace_back(pv[q]);
    return v;
}

int Vidyut::FillPatch(int lev, Real time, MultiFab& mf, int icomp, int ncomp) {
    (void)time;
    if (lev == 0) return vd_fill_patch(ctx_, &phi_new_[0].array(), &mf.array(), &dom_, &geom_, icomp, ncomp);
    if (lev != 1 || fine_new_.empty()) return VD_EINVAL;
    return fill_fine(mf, &fine_new_, &fine_ba_, icomp, ncomp);
}

// ------------------------------------------------------------ level 1
// the refined geometry of level 1 (refRatio(0) = ratio)
int Vidyut::set_fine_geom(const int ratio[3]) {
    fgeom_ = geom_;
    for (int d = 0; d < 3; d++) {
        ratio_[d] = d < dim_ ? ratio[d] : 1;
        if (d >= dim_) continue;
        if (ratio_[d] != 2 && ratio_[d] != 4) return VD_EINVAL;
        fgeom_.dom_lo[d] = geom_.dom_lo[d] * ratio_[d];
        fgeom_.dom_hi[d] = (geom_.dom_hi[d] + 1) * ratio_[d] - 1;
        fgeom_.dx[d] = geom_.dx[d] / ratio_[d];
    }
    return VD_OK;
}

// FillPatchTwoLevels into mf from the level-1 boxes src (none: from the
// coarse level alone, InterpFromCoarseLevel)
int Vidyut::fill_fine(MultiFab& mf, const std::vector<MultiFab>* src, const std::vector<vd_box>* src_ba, int icomp,
                      int ncomp) {
    std::vector<vd_array4> fa;
    if (src)
        for (const MultiFab& m : *src) fa.push_back(m.array());
    const int nf = src ? (int)src->size() : 0;
    return vd_fill_patch_two_levels(ctx_, &phi_new_[0].array(), &geom_, nf ? fa.data() : nullptr,
                                    nf ? src_ba->data() : nullptr, nf, &mf.array(), &fgeom_, ratio_, icomp, ncomp);
}

// Source/GridWork.cpp:143-164
int Vidyut::FillCoarsePatch(int lev, Real time, MultiFab& mf, int icomp, int ncomp) {
    (void)time;
    if (lev != 1) return VD_EINVAL;
    return fill_fine(mf, nullptr, nullptr, icomp, ncomp);
}

// Source/GridWork.cpp:17-29: phi_new of the new level from the coarse one
int Vidyut::MakeNewLevelFromCoarse(int lev, Real time, const std::vector<vd_box>& ba) {
    if (lev != 1) return VD_EINVAL;
    std::vector<MultiFab> nn(ba.size()), no(ba.size());
    for (size_t b = 0; b < ba.size(); b++) {
        nn[b].define(ctx_, ba[b], dim_, nvar_, 0);
        no[b].define(ctx_, ba[b], dim_, nvar_, 0);
        if (!nn[b].ok() || !no[b].ok()) return VD_ENOMEM;
        TRY(FillCoarsePatch(lev, time, nn[b], 0, nvar_));
    }
    fine_new_ = std::move(nn);
    fine_old_ = std::move(no);
    fine_ba_ = ba;
    t_new_fine_ = time;
    return VD_OK;
}

// Source/GridWork.cpp:33-49: the level on new grids, filled from its old
// data where the old grids cover a cell and from the coarse level elsewhere
int Vidyut::RemakeLevel(int lev, Real time, const std::vector<vd_box>& ba) {
    if (lev != 1 || fine_new_.empty()) return VD_EINVAL;
    std::vector<MultiFab> nn(ba.size()), no(ba.size());
    for (size_t b = 0; b < ba.size(); b++) {
        nn[b].define(ctx_, ba[b], dim_, nvar_, 0);
        no[b].define(ctx_, ba[b], dim_, nvar_, 0);
        if (!nn[b].ok() || !no[b].ok()) return VD_ENOMEM;
        TRY(fill_fine(nn[b], &fine_new_, &fine_ba_, 0, nvar_));
    }
    fine_new_ = std::move(nn);
    fine_old_ = std::move(no);
    fine_ba_ = ba;
    t_new_fine_ = time;
    return VD_OK;
}

// Source/GridWork.cpp:53-57
void Vidyut::ClearLevel(int lev) {
    if (lev != 1) return;
    fine_new_.clear();
    fine_old_.clear();
    fine_ba_.clear();
}

// Source/GridWork.cpp:94-108: covered coarse cells become the average of
// the fine cells over them
int Vidyut::AverageDownTo(int crse_lev) {
    if (crse_lev != 0) return VD_EINVAL;
    for (size_t b = 0; b < fine_new_.size(); b++)
        TRY(vd_average_down(ctx_, &fine_new_[b].array(), &fine_ba_[b], &phi_new_[0].array(), 0, nvar_, ratio_,
                            dim_));
    return VD_OK;
}

int Vidyut::AverageDown() { return fine_new_.empty() ? VD_OK : AverageDownTo(0); }

// regrid(lbase = 0): new level-1 grids from the tags of level 0
int Vidyut::regrid(int lbase, Real time, const vd_regrid_params& rp) {
    if (lbase != 0) return VD_EINVAL;
    if (!fine_new_.empty()) {
        for (int d = 0; d < dim_; d++)
            if (rp.ref_ratio[d] != ratio_[d]) return VD_EINVAL;  // refRatio is fixed for the run
    } else {
        TRY(set_fine_geom(rp.ref_ratio));
    }
    const size_t n = (size_t)(dom_.hi[0] - dom_.lo[0] + 1) * (dom_.hi[1] - dom_.lo[1] + 1) *
                     (dom_.hi[2] - dom_.lo[2] + 1);
    void* d = nullptr;
    TRY(vd_malloc(ctx_, n, &d));
    vd_tagarray4 tags;
    tags.p = (char*)d;
    for (int q = 0; q < 3; q++) {
        tags.begin[q] = dom_.lo[q];
        tags.end[q] = dom_.hi[q] + 1;
    }
    tags.jstride = tags.end[0] - tags.begin[0];
    tags.kstride = tags.jstride * (tags.end[1] - tags.begin[1]);
    std::vector<char> host(n);
    int st = vd_memset_zero(ctx_, d, n);
    if (st == VD_OK) st = ErrorEst(0, tags, time, 0);
    if (st == VD_OK) st = vd_memcpy_d2h(ctx_, host.data(), d, n);
    vd_free(ctx_, d);
    if (st != VD_OK) return st;
    const int per[3] = {geom_.is_periodic[0], geom_.is_periodic[1], geom_.is_periodic[2]};
    std::vector<vd_box> ba(64);
    int nb = 0;
    st = vd_make_new_grids(host.data(), &dom_, per, dim_, &rp, ba.data(), (int)ba.size(), &nb);
    if (st == VD_EINVAL && nb > (int)ba.size()) {
        ba.resize((size_t)nb);
        st = vd_make_new_grids(host.data(), &dom_, per, dim_, &rp, ba.data(), (int)ba.size(), &nb);
    }
    if (st != VD_OK) return st;
    ba.resize((size_t)nb);
    if (ba.empty()) {
        ClearLevel(1);
        return VD_OK;
    }
    return fine_new_.empty() ? MakeNewLevelFromCoarse(1, time, ba) : RemakeLevel(1, time, ba);
}

int Vidyut::compute_scalar_transport_flux(int lev, MultiFab& Sborder, FaceMF& flux, const int* bc_lo,
                                          const int* bc_hi, Real time, int specid) {
    (void)lev;
    vd_array4 f[3] = {flux[0].array(), flux[1].array(), flux[2].array()};
    return vd_compute_scalar_transport_flux(ctx_, mech_, &Sborder.array(), f, &dom_, &geom_, bc_lo, bc_hi,
                                            &prm_, time, 0.0, specid);
}

int Vidyut::compute_dsdt(int lev, int specid, FaceMF& flux, MultiFab& rxn_src, MultiFab& dsdt, Real time,
                         Real dt) {
    (void)lev;
    (void)time;
    (void)dt;
    vd_array4 f[3] = {flux[0].array(), flux[1].array(), flux[2].array()};
    return vd_compute_dsdt(ctx_, f, &rxn_src.array(), &dsdt.array(), &dom_, &geom_, specid);
}

int Vidyut::update_explsrc_at_all_levels(int specid, std::vector<MultiFab>& Sborder,
                                         std::vector<FaceMF>& flux, std::vector<MultiFab>& rxn_src,
                                         std::vector<MultiFab>& expl_src, const int* bc_lo,
                                         const int* bc_hi, Real cur_time) {
    (void)flux;
    // fluxes stay on chip (vd_transport_rhs == flux + dsdt)
    return vd_transport_rhs(ctx_, mech_, &Sborder[0].array(), &rxn_src[0].array(), &expl_src[0].array(),
                            &dom_, &geom_, bc_lo, bc_hi, &prm_, cur_time, 0.0, specid);
}

int Vidyut::update_rxnsrc_at_all_levels(std::vector<MultiFab>& Sborder, std::vector<MultiFab>& rxn_src,
                                        Real cur_time) {
    return vd_update_rxnsrc(ctx_, mech_, &Sborder[0].array(), &rxn_src[0].array(), &dom_, &prm_, cur_time);
}

int Vidyut::compute_elecenergy_source(int lev, MultiFab& Sborder, MultiFab& rxnsrc, FaceMF& efield,
                                      FaceMF& gradne, MultiFab& dsdt, Real time, Real dt, int floor_jh) {
    (void)lev;
    (void)time;
    (void)dt;
    vd_params p = prm_;
    p.floor_jh = floor_jh;
    vd_array4 ef[3] = {efield[0].array(), efield[1].array(), efield[2].array()};
    vd_array4 gn[3] = {gradne[0].array(), gradne[1].array(), gradne[2].array()};
    return vd_compute_elecenergy_source(ctx_, mech_, &Sborder.array(), &rxnsrc.array(), ef, gn,
                                        &dsdt.array(), &out().array(), &dom_, &geom_, &p);
}

int Vidyut::solve_potential(Real current_time, std::vector<MultiFab>& Sborder, const int* bc_lo,
                            const int* bc_hi, std::vector<FaceMF>& efield_ec) {
    vd_params p = prm_;
    for (int d = 0; d < 3; d++) {
        p.pot_bc_lo[d] = bc_lo[d];
        p.pot_bc_hi[d] = bc_hi[d];
    }
    vd_array4 ef[3] = {efield_ec[0][0].array(), efield_ec[0][1].array(), efield_ec[0][2].array()};
    int st = vd_solve_potential(ctx_, mech_, &Sborder[0].array(), &out().array(), ef, &dom_, &geom_, &p,
                                current_time, &pot_info_);
    mg_iters_ += pot_info_.iterations;
    return st;
}

int Vidyut::implicit_solve_scalar(Real current_time, Real dt, int spec_id, std::vector<MultiFab>& Sborder,
                                  std::vector<MultiFab>& Sborder_old, std::vector<MultiFab>& dsdt_expl,
                                  const int* bc_lo, const int* bc_hi, std::vector<FaceMF>& grad_fc) {
    vd_array4 g[3] = {grad_fc[0][0].array(), grad_fc[0][1].array(), grad_fc[0][2].array()};
    vd_solve_info info;
    int st = vd_implicit_solve_scalar(ctx_, mech_, spec_id, &Sborder[0].array(), &Sborder_old[0].array(),
                                      &dsdt_expl[0].array(), &out().array(), g, &dom_, &geom_, bc_lo,
                                      bc_hi, &prm_, current_time, dt, &info);
    mg_iters_ += info.iterations;
    return st;
}

int Vidyut::find_time_scales(int lev, Real& dt_edrift, Real& dt_ediff, Real& dt_diel_relax) {
    (void)lev;
    double out[3];
    TRY(vd_find_time_scales(ctx_, mech_, &state().array(), &dom_, &geom_, &prm_, out));
    dt_edrift = out[0];
    dt_ediff = out[1];
    dt_diel_relax = out[2];
    return VD_OK;
}

int Vidyut::set_tagging(const std::vector<int>& comps, const std::vector<double>& thr,
                        const std::vector<double>& gthr) {
    if (comps.size() != thr.size() || comps.size() != gthr.size()) return VD_EINVAL;
    for (int c : comps)
        if (c < 0 || c >= nvar_) return VD_ENOTFOUND;
    tag_comps_ = comps;
    tag_thr_ = thr;
    tag_gthr_ = gthr;
    return VD_OK;
}

int Vidyut::ErrorEst(int lev, vd_tagarray4& tags, Real time, int ngrow) {
    (void)ngrow;
    if (tag_comps_.empty()) return VD_OK;
    MultiFab S(ctx_, dom_, dim_, nvar_, 1);
    TRY(FillPatch(lev, time, S, 0, nvar_));
    return vd_error_est(ctx_, &S.array(), &tags, &dom_, (int)tag_comps_.size(), tag_comps_.data(),
                        tag_thr_.data(), tag_gthr_.data(), 1);
}

// Vidyut::ComputeDt (Source/TimeStep.cpp:14-31): the next step's dt from
// the time scales of this step's state (Source/Evolve.cpp:41-64)
int Vidyut::ComputeDt(Real cur_time, Real& dt) {
    Real ts[3];
    TRY(find_time_scales(0, ts[0], ts[1], ts[2]));
    return vd_compute_dt(&prm_, cur_time, dt, ts, &dt);
}

// Source/Evolve.cpp:33-369 on one level.  dt is vidyut.dt: every step's dt
// with a fixed step; with adaptive_dt the step before a simulation's first
// (later calls continue from the last step's dt, as the reference's dt[0])
int Vidyut::Evolve(int nsteps, Real dt, int ncorr) {
    if (!(prm_.adaptive_dt && dt_ > 0.0)) dt_ = dt;
    out_ = state_ = nullptr;
    const int st = ncorr == 1 ? evolve_one_corrector(nsteps) : evolve_corrected(nsteps, ncorr);
    out_ = state_ = nullptr;
    if (st != VD_OK) return st;
    return vd_ctx_sync(ctx_);
}

// The step body shared by both drivers, from the potential solve on:
// Sborder sb holds species, EEN and ETEMP of the state (ghost cells
// filled), results go to out(); Sborder_old is sb_old.
int Vidyut::step_body(std::vector<MultiFab>& sb, std::vector<MultiFab>& sb_old, Real cur_time, Real dt,
                      bool pot_in_sb) {
    const int E = desc_.e_idx;
    if (pot_in_sb) {
        // the potential, E and E/N are written into Sborder itself (its
        // species are the state's, which k_pot_post reads), whose ghost
        // cells are then filled in place
        MultiFab* o = out_;
        out_ = &sb[0];
        const int st = solve_potential(cur_time, sb, prm_.pot_bc_lo, prm_.pot_bc_hi, efield_ec_);
        out_ = o;
        TRY(st);
        TRY(vd_fill_patch(ctx_, &sb[0].array(), &sb[0].array(), &dom_, &geom_, VD_POT(ns_), 4));
    } else {
        TRY(solve_potential(cur_time, sb, prm_.pot_bc_lo, prm_.pot_bc_hi, efield_ec_));
        TRY(FillPatch(0, cur_time + dt, sb[0], VD_POT(ns_), 4));
    }
    if (prm_.do_reactions)
        TRY(update_rxnsrc_at_all_levels(sb, rxn_src_, cur_time));
    else
        TRY(rxn_src_[0].setVal(0.0));
    TRY(update_explsrc_at_all_levels(E, sb, flux_, rxn_src_, expl_src_, prm_.eden_bc_lo, prm_.eden_bc_hi,
                                     cur_time));
    int st = implicit_solve_scalar(cur_time, dt, E, sb, sb_old, expl_src_, prm_.eden_bc_lo, prm_.eden_bc_hi,
                                   gradne_fc_);
    if (st != VD_OK && st != VD_ENOCONV) return st;
    if (prm_.elecenergy_solve) {
        TRY(update_explsrc_at_all_levels(ns_, sb, flux_, rxn_src_, expl_src_, prm_.eenrg_bc_lo,
                                         prm_.eenrg_bc_hi, cur_time));
        TRY(compute_elecenergy_source(0, sb[0], rxn_src_[0], efield_ec_[0], gradne_fc_[0], expl_src_[0],
                                      cur_time, dt, prm_.floor_jh));
        st = implicit_solve_scalar(cur_time, dt, ns_, sb, sb_old, expl_src_, prm_.eenrg_bc_lo, prm_.eenrg_bc_hi,
                                   grad_fc_);
        if (st != VD_OK && st != VD_ENOCONV) return st;
    }
    for (int ind = 0; ind < ns_; ind++) {
        bool solve = !(desc_.is_bg[ind] || ind == E);
        if (solve) {
            const int* lo = desc_.charge[ind] != 0 ? prm_.ion_bc_lo : prm_.neutral_bc_lo;
            const int* hi = desc_.charge[ind] != 0 ? prm_.ion_bc_hi : prm_.neutral_bc_hi;
            TRY(update_explsrc_at_all_levels(ind, sb, flux_, rxn_src_, expl_src_, lo, hi, cur_time));
            st = implicit_solve_scalar(cur_time, dt, ind, sb, sb_old, expl_src_, lo, hi, grad_fc_);
            if (st != VD_OK && st != VD_ENOCONV) return st;
        } else if (prm_.do_bg_reactions) {
            // Source/Evolve.cpp:267-287 (also reached for the electron
            // index, whose solved value it then updates): n_new = n + dt *
            // rxn, a background species' n taken from Sborder, whose species
            // equal those of the state.  A run of consecutive background
            // species is one launch (each component's update reads only its
            // own n and rxn, so the order among them does not matter)
            int run = 1;
            if (ind != E)
                while (ind + run < ns_ && desc_.is_bg[ind + run] && ind + run != E) run++;
            const vd_array4& n_old = ind == E ? out().array() : sb[0].array();  // E: the solved value
            TRY(vd_lincomb(ctx_, &out().array(), 1.0, &n_old, ind, dt, &rxn_src_[0].array(), ind, ind, run, &dom_));
            if (prm_.bound_specden)
                for (int q = ind; q < ind + run; q++)
                    TRY(vd_clamp_min(ctx_, &out().array(), &dom_, q, prm_.min_species_density));
            ind += run - 1;
        }
    }
    return VD_OK;
}

// One corrector.  The state moves between phi_new and two ghosted buffers
// (Sborder_ and Swork_) so that no step copies the whole state: a step
// takes as its Sborder the buffer that holds the state -- its ghost cells
// filled in place; the first step of a call copies phi_new in, as
// FillPatch does -- and writes its results into the other buffer, or into
// phi_new in the last step of the call.  The components it does not write
// are copied over from the state (see below).  Sborder_old is Sborder: the
// implicit solves read only its species and EEN, which equal the state's.
int Vidyut::evolve_one_corrector(int nsteps) {
    MultiFab* P = &phi_new_[0];
    std::vector<MultiFab>* G[2] = {&Sborder_, &Swork_};
    MultiFab* cur = P;
    MultiFab* e_src = nullptr;  // the previous step's Sborder, which holds the state's E
    for (int step = 0; step < nsteps; step++) {
        const bool first = step == 0, last = step == nsteps - 1;
        std::vector<MultiFab>& sb = first ? *G[0] : (cur == &(*G[0])[0] ? *G[0] : *G[1]);
        std::vector<MultiFab>& other = &sb == G[0] ? *G[1] : *G[0];
        MultiFab* out = last ? P : &other[0];
        state_ = cur;
        out_ = out;
        phi_old_stale_ = false;
        const Real cur_time = t_new_;
        // the time scales read E where the last step left it
        if (e_src) TRY(vd_ctx_set_cfl_efield(ctx_, &e_src->array()));
        const int dst = ComputeDt(cur_time, dt_);
        if (e_src) TRY(vd_ctx_set_cfl_efield(ctx_, nullptr));
        TRY(dst);
        const Real dt = dt_;
        // species, EEN, ETEMP into Sborder (a copy in the first step, the
        // ghost cells only after)
        TRY(vd_fill_patch(ctx_, &cur->array(), &sb[0].array(), &dom_, &geom_, 0, ns_ + 2));
        TRY(step_body(sb, sb, cur_time, dt, true));
        // what the step wrote into Sborder: the potential, E and E/N go to
        // phi_new in the last step; within the call no step reads them from
        // the state but the next step's time scales, which read E from this
        // Sborder (it is the next step's output buffer, written only after
        // those time scales)
        const int nef = dim_;
        if (last) {
            TRY(out->copy_from(sb[0], VD_POT(ns_), VD_POT(ns_), 1 + nef));
            TRY(out->copy_from(sb[0], VD_REF(ns_), VD_REF(ns_), 1));
            e_src = nullptr;
        } else {
            e_src = &sb[0];
        }
        // the components no step writes keep their values, so each buffer
        // needs them once per call: the two ghosted buffers receive them in
        // the first two steps, and phi_new holds them already
        if (out != cur && step < 2) {
            if (nef < 3) TRY(out->copy_from(*cur, VD_EFX(ns_) + nef, VD_EFX(ns_) + nef, 3 - nef));
            TRY(out->copy_from(*cur, VD_EEH(ns_), VD_EEH(ns_), 1));
            if (!prm_.elecenergy_solve) {
                TRY(out->copy_from(*cur, VD_EEN(ns_), VD_EEN(ns_), 2));
                TRY(out->copy_from(*cur, VD_EJH(ns_), VD_EJH(ns_), 2));
            }
            if (!prm_.do_bg_reactions)
                for (int ind = 0; ind < ns_; ind++)
                    if (desc_.is_bg[ind] && ind != desc_.e_idx) TRY(out->copy_from(*cur, ind, ind, 1));
        }
        istep_++;
        t_new_ = cur_time + dt;
        sb_last_ = &sb[0];
        phi_old_stale_ = true;
        cur = out;
    }
    return VD_OK;
}

// More than one corrector (predictor-corrector averaging with phi_old).
int Vidyut::evolve_corrected(int nsteps, int ncorr) {
    for (int step = 0; step < nsteps; step++) {
        const Real cur_time = t_new_;
        TRY(ComputeDt(cur_time, dt_));
        const Real dt = dt_;
        TRY(phi_old_[0].copy_from(phi_new_[0], 0, 0, nvar_));
        phi_old_stale_ = false;
        for (int niter = 0; niter < ncorr; niter++) {
            // Sborder components read by the step: species, EEN, ETEMP (here)
            // and POT, EFX..EFZ (after the potential solve); EJH, EIH, EEH and
            // REF of Sborder are never read.
            TRY(FillPatch(0, cur_time + dt, Sborder_[0], 0, ns_ + 2));
            TRY(step_body(Sborder_, phi_old_, cur_time, dt, false));
            if (niter < ncorr - 1)
                TRY(vd_lincomb(ctx_, &phi_new_[0].array(), 0.5, &phi_old_[0].array(), 0, 0.5,
                               &phi_new_[0].array(), 0, 0, nvar_, &dom_));
        }
        istep_++;
        t_new_ = cur_time + dt;
    }
    return VD_OK;
}

// --------------------------------------------------------------- plotfile
static int mkdir_p(const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return VD_EINVAL;
    return VD_OK;
}

static std::string box_str(const vd_box& b, int dim) {
    std::ostringstream o;
    o << "((";
    for (int d = 0; d < dim; d++) o << b.lo[d] << (d + 1 < dim ? "," : "");
    o << ") (";
    for (int d = 0; d < dim; d++) o << b.hi[d] << (d + 1 < dim ? "," : "");
    o << ") (";
    for (int d = 0; d < dim; d++) o << 0 << (d + 1 < dim ? "," : "");
    o << "))";
    return o.str();
}

// One single-box VisMF MultiFab (version 1 header): <dir>/<prefix>_H and
// <dir>/<prefix>_D_00000 holding the FAB (little-endian IEEE doubles,
// component after component, x fastest), with per-component minima and
// maxima in the header.  h holds ncomp x ncell values.
// one FAB of a VisMF MultiFab: host data (ncomp x ncell, component-major)
// over a box
struct HostFab {
    std::vector<double> h;
    int64_t ncell;
    vd_box box;
};

// The VisMF MultiFab <prefix>: every FAB in <prefix>_D_00000 (a text
// header line, then the doubles), and <prefix>_H with the BoxArray, each
// FAB's file offset and the per-FAB, per-component minima and maxima.
static int write_vismf(const std::string& dir, const std::string& prefix, const std::vector<HostFab>& fabs,
                       int ncomp, int dim) {
    const std::string base = prefix.substr(prefix.find_last_of('/') + 1);
    std::vector<long long> off;
    {
        std::ofstream D(dir + "/" + prefix + "_D_00000", std::ios::binary);
        long long pos = 0;
        for (const HostFab& f : fabs) {
            std::ostringstream o;
            o << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_str(f.box, dim) << " " << ncomp
              << "\n";
            const std::string hdr = o.str();
            off.push_back(pos);
            D.write(hdr.data(), hdr.size());
            D.write((const char*)f.h.data(), (size_t)ncomp * f.ncell * sizeof(double));
            pos += (long long)hdr.size() + (long long)ncomp * f.ncell * (long long)sizeof(double);
        }
        if (!D.good()) return VD_EINVAL;
    }
    const size_t nf = fabs.size();
    std::vector<double> mn(nf * ncomp, DBL_MAX), mx(nf * ncomp, -DBL_MAX);
    for (size_t b = 0; b < nf; b++)
        for (int q = 0; q < ncomp; q++)
            for (int64_t c = 0; c < fabs[b].ncell; c++) {
                const double v = fabs[b].h[(size_t)q * fabs[b].ncell + c];
                mn[b * ncomp + q] = std::min(mn[b * ncomp + q], v);
                mx[b * ncomp + q] = std::max(mx[b * ncomp + q], v);
            }
    std::ofstream C(dir + "/" + prefix + "_H");
    C.precision(16);
    C << std::scientific;
    C << "1\n1\n" << ncomp << "\n0\n(" << nf << " 0\n";
    for (const HostFab& f : fabs) C << box_str(f.box, dim) << "\n";
    C << ")\n" << nf << "\n";
    for (size_t b = 0; b < nf; b++) C << "FabOnDisk: " << base << "_D_00000 " << off[b] << "\n";
    C << "\n" << nf << "," << ncomp << "\n";
    for (size_t b = 0; b < nf; b++) {
        for (int q = 0; q < ncomp; q++) C << mn[b * ncomp + q] << ",";
        C << "\n";
    }
    C << "\n" << nf << "," << ncomp << "\n";
    for (size_t b = 0; b < nf; b++) {
        for (int q = 0; q < ncomp; q++) C << mx[b * ncomp + q] << ",";
        C << "\n";
    }
    return C.good() ? VD_OK : VD_EINVAL;
}

// host copies of a level's phi_new FABs
static int level_fabs(vd_ctx* ctx, std::vector<MultiFab>& mfs, int nvar, std::vector<HostFab>& out) {
    for (MultiFab& m : mfs) {
        HostFab f;
        f.box = m.box();
        f.ncell = m.array().nstride;
        f.h.resize((size_t)f.ncell * nvar);
        TRY(vd_memcpy_d2h(ctx, f.h.data(), m.data(), f.h.size() * sizeof(double)));
        out.push_back(std::move(f));
    }
    return VD_OK;
}

// WritePlotFile -> amrex::WriteMultiLevelPlotfile (Source/Plot.cpp:29-38):
// HyperCLaw-V1.1 Header, Level_0/Cell_H (VisMF v1) and one Cell_D file
// holding the FAB (little-endian IEEE doubles, Fortran order per comp).
int Vidyut::WritePlotFile(const std::string& dir) const {
    Vidyut* self = const_cast<Vidyut*>(this);  // device reads only
    const int finest = finest_level();
    std::vector<std::vector<HostFab>> lv(finest + 1);
    TRY(level_fabs(ctx_, self->phi_new_, nvar_, lv[0]));
    if (finest == 1) TRY(level_fabs(ctx_, self->fine_new_, nvar_, lv[1]));
    TRY(mkdir_p(dir));
    const vd_geom* gl[2] = {&geom_, &fgeom_};
    std::vector<std::string> names = varnames();
    {
        std::ofstream H(dir + "/Header");
        H.precision(17);
        H << "HyperCLaw-V1.1\n" << nvar_ << "\n";
        for (auto& s : names) H << s << "\n";
        H << dim_ << "\n" << t_new_ << "\n" << finest << "\n";
        for (int d = 0; d < dim_; d++) H << geom_.prob_lo[d] << " ";
        H << "\n";
        for (int d = 0; d < dim_; d++) H << geom_.prob_hi[d] << " ";
        H << "\n";
        for (int l = 0; l < finest; l++) H << ratio_[0] << " ";
        H << "\n";
        for (int l = 0; l <= finest; l++) {
            vd_box db;
            for (int d = 0; d < 3; d++) {
                db.lo[d] = gl[l]->dom_lo[d];
                db.hi[d] = gl[l]->dom_hi[d];
            }
            H << box_str(db, dim_) << " ";
        }
        H << "\n";
        for (int l = 0; l <= finest; l++) H << istep_ << " ";
        H << "\n";
        for (int l = 0; l <= finest; l++) {
            for (int d = 0; d < dim_; d++) H << gl[l]->dx[d] << " ";
            H << "\n";
        }
        H << "0\n0\n";
        for (int l = 0; l <= finest; l++) {
            H << l << " " << lv[l].size() << " " << t_new_ << "\n" << istep_ << "\n";
            for (const HostFab& f : lv[l])
                for (int d = 0; d < dim_; d++) {
                    const vd_geom& g = *gl[l];
                    const double lo = g.prob_lo[d] + g.dx[d] * (f.box.lo[d] - g.dom_lo[d]);
                    const double hi = g.prob_lo[d] + g.dx[d] * (f.box.hi[d] + 1 - g.dom_lo[d]);
                    H << lo << " " << hi << "\n";
                }
            H << "Level_" << l << "/Cell\n";
        }
        if (!H.good()) return VD_EINVAL;
    }
    for (int l = 0; l <= finest; l++) {
        TRY(mkdir_p(dir + "/Level_" + std::to_string(l)));
        TRY(write_vismf(dir, "Level_" + std::to_string(l) + "/Cell", lv[l], nvar_, dim_));
    }
    return VD_OK;
}

// WriteCheckpointFile (Source/Plot.cpp:40-119): a "Checkpoint file for
// Vidyut" Header (finest level, istep, dt and t_new per level, 17
// significant digits, then the BoxArray) and phi_new as the VisMF MultiFab
// Level_0/phi.
int Vidyut::WriteCheckpointFile(const std::string& dir) const {
    Vidyut* self = const_cast<Vidyut*>(this);  // device reads only
    const int finest = finest_level();
    std::vector<std::vector<HostFab>> lv(finest + 1);
    TRY(level_fabs(ctx_, self->phi_new_, nvar_, lv[0]));
    if (finest == 1) TRY(level_fabs(ctx_, self->fine_new_, nvar_, lv[1]));
    TRY(mkdir_p(dir));
    {
        std::ofstream H(dir + "/Header");
        H.precision(17);
        H << "Checkpoint file for Vidyut\n" << finest << "\n";
        for (int l = 0; l <= finest; l++) H << istep_ << " ";
        H << "\n";
        for (int l = 0; l <= finest; l++) H << dt_ << " ";
        H << "\n";
        for (int l = 0; l <= finest; l++) H << t_new_ << " ";
        H << "\n";
        for (int l = 0; l <= finest; l++) {
            H << "(" << lv[l].size() << " 0\n";
            for (const HostFab& f : lv[l]) H << box_str(f.box, dim_) << "\n";
            H << ")\n";
        }
        if (!H.good()) return VD_EINVAL;
    }
    for (int l = 0; l <= finest; l++) {
        TRY(mkdir_p(dir + "/Level_" + std::to_string(l)));
        TRY(write_vismf(dir, "Level_" + std::to_string(l) + "/phi", lv[l], nvar_, dim_));
    }
    return VD_OK;
}

// ReadCheckpointFile (Source/Plot.cpp:121-209) for a single-level
// checkpoint of this domain: istep, dt, t_new and phi_new.  The FAB may be
// in either byte order.
// the FABs of the VisMF MultiFab <dir>/<prefix> (phi_H and its data file),
// in BoxArray order, as host doubles (ncomp x cells of each box)
static int read_vismf(const std::string& dir, const std::string& prefix, int ncomp,
                      std::vector<std::vector<double>>& out) {
    std::ifstream C(dir + "/" + prefix + "_H");
    if (!C.good()) return VD_ENOTFOUND;
    std::string ch((std::istreambuf_iterator<char>(C)), std::istreambuf_iterator<char>());
    const std::string sub = prefix.substr(0, prefix.find_last_of('/') + 1);
    size_t fp = 0;
    while ((fp = ch.find("FabOnDisk: ", fp)) != std::string::npos) {
        std::istringstream fl(ch.substr(fp + 11));
        fp += 11;
        std::string fname;
        long long off = 0;
        fl >> fname >> off;
        std::ifstream D(dir + "/" + sub + fname, std::ios::binary);
        if (!D.good()) return VD_ENOTFOUND;
        D.seekg(off);
        std::string fab;
        std::getline(D, fab);
        const bool le = fab.find("(8 7 6 5 4 3 2 1)") != std::string::npos;
        const bool be = fab.find("(1 2 3 4 5 6 7 8)") != std::string::npos;
        if (fab.compare(0, 4, "FAB ") != 0 || !(le || be)) return VD_EINVAL;
        const size_t nb = fab.rfind(' ');
        if (nb == std::string::npos || std::atoi(fab.c_str() + nb + 1) != ncomp) return VD_EINVAL;
        // the box ((lo) (hi) (type)) and ncomp after the real descriptor
        const size_t bp = fab.find(")))");
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        if (bp == std::string::npos) return VD_EINVAL;
        std::string bx = fab.substr(bp + 3);
        for (char& c : bx)
            if (c == '(' || c == ')' || c == ',') c = ' ';
        std::istringstream bs(bx);
        std::vector<long long> v;
        long long x;
        while (bs >> x) v.push_back(x);
        // lo[dim], hi[dim], type[dim], ncomp
        const int dim = (int)(v.size() - 1) / 3;
        if (dim < 1 || dim > 3) return VD_EINVAL;
        for (int d = 0; d < dim; d++) {
            lo[d] = (int)v[(size_t)d];
            hi[d] = (int)v[(size_t)(dim + d)];
        }
        size_t cells = 1;
        for (int d = 0; d < dim; d++) cells *= (size_t)(hi[d] - lo[d] + 1);
        std::vector<double> h(cells * (size_t)ncomp);
        D.read((char*)h.data(), h.size() * sizeof(double));
        if ((size_t)D.gcount() != h.size() * sizeof(double)) return VD_EINVAL;
        if (be)
            for (auto& w : h) {
                unsigned char* b = (unsigned char*)&w;
                std::reverse(b, b + 8);
            }
        out.push_back(std::move(h));
    }
    return out.empty() ? VD_EINVAL : VD_OK;
}

// the BoxArray "(n 0\n((lo) (hi) (t))\n...)" at the stream's position
static bool read_boxarray(std::istream& H, int dim, std::vector<vd_box>& ba) {
    std::string line;
    while (std::getline(H, line) && line.find('(') == std::string::npos) {
    }
    int n = 0;
    if (std::sscanf(line.c_str(), " (%d", &n) != 1 || n < 0) return false;
    for (int b = 0; b < n; b++) {
        if (!std::getline(H, line)) return false;
        for (char& c : line)
            if (c == '(' || c == ')' || c == ',') c = ' ';
        std::istringstream bs(line);
        vd_box x{};
        for (int d = 0; d < dim; d++) bs >> x.lo[d];
        for (int d = 0; d < dim; d++) bs >> x.hi[d];
        if (!bs) return false;
        ba.push_back(x);
    }
    return (bool)std::getline(H, line);  // ")"
}

// ReadCheckpointFile (Source/Plot.cpp:121-209): istep, dt and t_new, the
// level-0 state of this domain and, for a two-level checkpoint, level 1 on
// its BoxArray (refinement ratio: the last regrid's, else 2).  The FABs
// may be in either byte order.
int Vidyut::ReadCheckpointFile(const std::string& dir) {
    std::ifstream H(dir + "/Header");
    if (!H.good()) return VD_ENOTFOUND;
    std::string line;
    std::getline(H, line);  // title
    int finest = -1;
    H >> finest;
    if (finest != 0 && finest != 1) return VD_EINVAL;  // at most two levels
    long long st = 0;
    double dtv = 0.0, tv = 0.0, skip = 0.0;
    std::getline(H, line);
    std::getline(H, line);
    {
        std::istringstream l(line);
        l >> st;
    }
    std::getline(H, line);
    {
        std::istringstream l(line);
        l >> dtv;
    }
    std::getline(H, line);
    {
        std::istringstream l(line);
        l >> tv;
    }
    (void)skip;
    std::vector<vd_box> ba0, ba1;
    if (!read_boxarray(H, dim_, ba0) || ba0.size() != 1) return VD_EINVAL;
    for (int d = 0; d < dim_; d++)
        if (ba0[0].lo[d] != dom_.lo[d] || ba0[0].hi[d] != dom_.hi[d]) return VD_EINVAL;  // another domain
    if (finest == 1 && (!read_boxarray(H, dim_, ba1) || ba1.empty())) return VD_EINVAL;
    std::vector<std::vector<double>> f0, f1;
    TRY(read_vismf(dir, "Level_0/phi", nvar_, f0));
    const vd_array4& a = phi_new_[0].array();
    if (f0.size() != 1 || f0[0].size() != (size_t)a.nstride * nvar_) return VD_EINVAL;
    if (finest == 1) {
        TRY(read_vismf(dir, "Level_1/phi", nvar_, f1));
        if (f1.size() != ba1.size()) return VD_EINVAL;
        if (fine_new_.empty()) TRY(set_fine_geom(ratio_));
        std::vector<MultiFab> nn(ba1.size()), no(ba1.size());
        for (size_t b = 0; b < ba1.size(); b++) {
            nn[b].define(ctx_, ba1[b], dim_, nvar_, 0);
            no[b].define(ctx_, ba1[b], dim_, nvar_, 0);
            if (!nn[b].ok() || !no[b].ok()) return VD_ENOMEM;
            if (f1[b].size() * sizeof(double) != nn[b].bytes()) return VD_EINVAL;
            TRY(vd_memcpy_h2d(ctx_, nn[b].data(), f1[b].data(), nn[b].bytes()));
        }
        fine_new_ = std::move(nn);
        fine_old_ = std::move(no);
        fine_ba_ = ba1;
        t_new_fine_ = tv;
    } else {
        ClearLevel(1);
    }
    TRY(vd_memcpy_h2d(ctx_, a.p, f0[0].data(), f0[0].size() * sizeof(double)));
    istep_ = (int)st;
    dt_ = dtv;
    t_new_ = tv;
    phi_old_stale_ = false;
    return VD_OK;
}

}  // namespace vidyut_amd

// ------------------------------------------------------------ C handle
struct vd_sim {
    vidyut_amd::Vidyut* v;
    vd_ctx* ctx;
};

using vidyut_amd::Vidyut;

extern "C" {

int vd_sim_create(vd_ctx* ctx, const vd_mech_desc* mech, const vd_geom* geom, const vd_params* prm,
                  vd_sim** out) {
    if (!ctx || !mech || !geom || !prm || !out) return VD_EINVAL;
    if (geom->dim != 2 && geom->dim != 3) return VD_EINVAL;
    // the cases of a batched context share one step sequence and one
    // simulation time: each case's adaptive dt would differ, and a case that
    // joins later would see the batch's time in a time-dependent voltage
    if (vd_ctx_ncase(ctx) > 1 && (prm->adaptive_dt || prm->voltage_profile != 0)) return VD_EINVAL;
    Vidyut* v = new Vidyut(ctx, *mech, *geom, *prm);
    if (v->status() != VD_OK) {
        int st = v->status();
        delete v;
        return st;
    }
    *out = new vd_sim{v, ctx};
    return VD_OK;
}

int vd_sim_destroy(vd_sim* s) {
    if (!s) return VD_EINVAL;
    delete s->v;
    delete s;
    return VD_OK;
}

int vd_sim_set_state(vd_sim* s, const double* host, double time) {
    auto& mf = s->v->phi_new();
    s->v->set_time(time);
    return vd_memcpy_h2d(s->ctx, mf.data(), host, mf.bytes());
}

int vd_sim_get_state(vd_sim* s, double* host) {
    auto& mf = s->v->phi_new();
    return vd_memcpy_d2h(s->ctx, host, mf.data(), mf.bytes());
}

int vd_sim_set_state_case(vd_sim* s, int c, const double* host) {
    if (!s || !host || c < 0 || c >= vd_ctx_ncase(s->ctx)) return VD_EINVAL;
    auto& mf = s->v->phi_new();
    char* p = (char*)mf.data() + (size_t)c * (size_t)vd_ctx_case_stride(s->ctx);
    return vd_memcpy_h2d(s->ctx, p, host, mf.bytes());
}

int vd_sim_get_state_case(vd_sim* s, int c, double* host) {
    if (!s || !host || c < 0 || c >= vd_ctx_ncase(s->ctx)) return VD_EINVAL;
    auto& mf = s->v->phi_new();
    const char* p = (const char*)mf.data() + (size_t)c * (size_t)vd_ctx_case_stride(s->ctx);
    return vd_memcpy_d2h(s->ctx, host, p, mf.bytes());
}

int vd_sim_state_view(vd_sim* s, vd_array4* out) {
    *out = s->v->phi_new().array();
    return VD_OK;
}

int vd_sim_evolve(vd_sim* s, int nsteps, double dt, int ncorr) {
    if (nsteps < 0 || !(dt > 0) || ncorr < 1) return VD_EINVAL;
    return s->v->Evolve(nsteps, dt, ncorr);
}

double vd_sim_time(const vd_sim* s) { return s->v->time(); }
double vd_sim_dt(const vd_sim* s) { return s->v->dt(); }
long long vd_sim_mg_iterations(const vd_sim* s) { return s->v->mg_iterations(); }

int vd_sim_write_plotfile(vd_sim* s, const char* dir) { return s->v->WritePlotFile(dir); }

int vd_sim_write_checkpoint(vd_sim* s, const char* dir) {
    if (!s || !dir) return VD_EINVAL;
    return s->v->WriteCheckpointFile(dir);
}

int vd_sim_read_checkpoint(vd_sim* s, const char* dir) {
    if (!s || !dir) return VD_EINVAL;
    return s->v->ReadCheckpointFile(dir);
}

int vd_sim_step(const vd_sim* s) { return s->v->step(); }

int vd_sim_regrid(vd_sim* s, int ntag, const int* comps, const double* thr, const double* gthr,
                  const vd_regrid_params* rp, int* nboxes) {
    if (!s || !rp || !nboxes || ntag < 0 || (ntag > 0 && (!comps || !thr || !gthr))) return VD_EINVAL;
    std::vector<int> c(comps, comps + ntag);
    std::vector<double> t(thr, thr + ntag), g(gthr, gthr + ntag);
    int st = s->v->set_tagging(c, t, g);
    if (st) return st;
    st = s->v->regrid(0, s->v->time(), *rp);
    *nboxes = (int)s->v->fine_grids().size();
    return st;
}

int vd_sim_level_boxes(vd_sim* s, int lev, vd_box* out, int max_out, int* n) {
    if (!s || !n || (lev != 0 && lev != 1) || (max_out > 0 && !out)) return VD_EINVAL;
    if (lev == 0) {
        *n = 1;
        if (max_out < 1) return VD_EINVAL;
        out[0] = s->v->phi_new().box();
        return VD_OK;
    }
    const std::vector<vd_box>& ba = s->v->fine_grids();
    *n = (int)ba.size();
    if ((int)ba.size() > max_out) return VD_EINVAL;
    for (size_t b = 0; b < ba.size(); b++) out[b] = ba[b];
    return VD_OK;
}

int vd_sim_get_fine(vd_sim* s, int b, double* host) {
    if (!s || !host || b < 0 || b >= (int)s->v->fine_phi_new().size()) return VD_EINVAL;
    auto& mf = s->v->fine_phi_new()[(size_t)b];
    return vd_memcpy_d2h(s->ctx, host, mf.data(), mf.bytes());
}

int vd_sim_set_fine(vd_sim* s, int b, const double* host) {
    if (!s || !host || b < 0 || b >= (int)s->v->fine_phi_new().size()) return VD_EINVAL;
    auto& mf = s->v->fine_phi_new()[(size_t)b];
    return vd_memcpy_h2d(s->ctx, mf.data(), host, mf.bytes());
}

int vd_sim_average_down(vd_sim* s) {
    if (!s) return VD_EINVAL;
    const int st = s->v->AverageDown();
    return st != VD_OK ? st : vd_ctx_sync(s->ctx);
}

int vd_sim_error_est(vd_sim* s, int ntag, const int* comps, const double* thr, const double* gthr,
                     char* host_tags) {
    std::vector<int> c(comps, comps + ntag);
    std::vector<double> t(thr, thr + ntag), g(gthr, gthr + ntag);
    int st = s->v->set_tagging(c, t, g);
    if (st) return st;
    const vd_box& b = s->v->phi_new().box();
    size_t n = (size_t)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
    void* d;
    st = vd_malloc(s->ctx, n, &d);
    if (st) return st;
    vd_memset_zero(s->ctx, d, n);
    vd_tagarray4 tags;
    tags.p = (char*)d;
    for (int q = 0; q < 3; q++) {
        tags.begin[q] = b.lo[q];
        tags.end[q] = b.hi[q] + 1;
    }
    tags.jstride = tags.end[0] - tags.begin[0];
    tags.kstride = tags.jstride * (tags.end[1] - tags.begin[1]);
    st = s->v->ErrorEst(0, tags, s->v->time(), 0);
    if (st == VD_OK) st = vd_memcpy_d2h(s->ctx, host_tags, d, n);
    vd_free(s->ctx, d);
    return st;
}

}  // extern "C"
