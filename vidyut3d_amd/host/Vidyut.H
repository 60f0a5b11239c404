// Vidyut.H -- host C++ layer over the C ABI.  Keeps the member-function
// boundary of the reference's Vidyut class (Source/Vidyut.H:72-119) with
// MultiFab-in / MultiFab-out calls, single level, one box per level, all
// data resident in device memory.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "../../include/vidyut3d_amd_host.h"

namespace vidyut_amd {

using Real = double;

// One-box, device-resident MultiFab (component-major Array4 layout).
// face_dir >= 0 makes it face-centred in that direction (surroundingNodes).
class MultiFab {
   public:
    MultiFab() = default;
    MultiFab(vd_ctx* ctx, const vd_box& valid, int dim, int ncomp, int ngrow, int face_dir = -1);
    ~MultiFab();
    MultiFab(const MultiFab&) = delete;
    MultiFab& operator=(const MultiFab&) = delete;
    MultiFab(MultiFab&& o) noexcept;
    MultiFab& operator=(MultiFab&& o) noexcept;

    void define(vd_ctx* ctx, const vd_box& valid, int dim, int ncomp, int ngrow, int face_dir = -1);
    bool ok() const { return a_.p != nullptr; }
    vd_array4& array() { return a_; }
    const vd_array4& array() const { return a_; }
    const vd_box& box() const { return valid_; }
    int nComp() const { return a_.ncomp; }
    int nGrow() const { return ngrow_; }
    size_t bytes() const { return (size_t)a_.nstride * a_.ncomp * sizeof(double); }
    double* data() { return a_.p; }
    int setVal(double v);
    // copies comps [scomp, scomp+n) of src to [dcomp, ...) on the valid box
    int copy_from(const MultiFab& src, int scomp, int dcomp, int n);

   private:
    void release();
    vd_ctx* ctx_ = nullptr;
    vd_array4 a_{};
    vd_box valid_{};
    int ngrow_ = 0;
};

using FaceMF = std::array<MultiFab, 3>;

class Vidyut {
   public:
    Vidyut(vd_ctx* ctx, const vd_mech_desc& mech, const vd_geom& geom, const vd_params& prm);
    ~Vidyut();
    int status() const { return status_; }

    // ---- the reference's member-function boundary ----------------------
    int compute_scalar_transport_flux(int lev, MultiFab& Sborder, FaceMF& flux, const int* bc_lo,
                                      const int* bc_hi, Real time, int specid);
    int compute_dsdt(int lev, int specid, FaceMF& flux, MultiFab& rxn_src, MultiFab& dsdt, Real time,
                     Real dt);
    int update_explsrc_at_all_levels(int specid, std::vector<MultiFab>& Sborder,
                                     std::vector<FaceMF>& flux, std::vector<MultiFab>& rxn_src,
                                     std::vector<MultiFab>& expl_src, const int* bc_lo,
                                     const int* bc_hi, Real cur_time);
    int update_rxnsrc_at_all_levels(std::vector<MultiFab>& Sborder, std::vector<MultiFab>& rxn_src,
                                    Real cur_time);
    int compute_elecenergy_source(int lev, MultiFab& Sborder, MultiFab& rxnsrc, FaceMF& efield,
                                  FaceMF& gradne, MultiFab& dsdt, Real time, Real dt, int floor_jh);
    int solve_potential(Real current_time, std::vector<MultiFab>& Sborder, const int* bc_lo,
                        const int* bc_hi, std::vector<FaceMF>& efield_ec);
    int implicit_solve_scalar(Real current_time, Real dt, int spec_id, std::vector<MultiFab>& Sborder,
                              std::vector<MultiFab>& Sborder_old, std::vector<MultiFab>& dsdt_expl,
                              const int* bc_lo, const int* bc_hi, std::vector<FaceMF>& grad_fc);
    int find_time_scales(int lev, Real& dt_edrift, Real& dt_ediff, Real& dt_diel_relax);
    // ComputeDt (Source/TimeStep.cpp:14-31): dt in, the next step's dt out
    int ComputeDt(Real cur_time, Real& dt);
    int ErrorEst(int lev, vd_tagarray4& tags, Real time, int ngrow);
    int FillPatch(int lev, Real time, MultiFab& mf, int icomp, int ncomp);

    // ---- driver ---------------------------------------------------------
    // Evolve (Source/Evolve.cpp:33-369) for nsteps: at fixed dt, or with
    // adaptive_dt from dt (the first call) or the last step's dt
    int Evolve(int nsteps, Real dt, int ncorr);
    Real dt() const { return dt_; }
    int WritePlotFile(const std::string& dir) const;
    int WriteCheckpointFile(const std::string& dir) const;
    int ReadCheckpointFile(const std::string& dir);
    int set_tagging(const std::vector<int>& comps, const std::vector<double>& thr,
                    const std::vector<double>& gthr);

    // ---- AMR level 1 (GridWork.cpp:17-164): the regrid cycle and the
    // level transfers; the step driver advances level 0 only ---------------
    // regrid(0, time) as Evolve.cpp:66-72 runs it: ErrorEst on level 0,
    // MakeNewGrids (vd_make_new_grids), then MakeNewLevelFromCoarse for a
    // new level, RemakeLevel for an existing one, ClearLevel when no cell
    // is tagged
    int regrid(int lbase, Real time, const vd_regrid_params& rp);
    int MakeNewLevelFromCoarse(int lev, Real time, const std::vector<vd_box>& ba);
    int RemakeLevel(int lev, Real time, const std::vector<vd_box>& ba);
    void ClearLevel(int lev);
    int AverageDown();
    int AverageDownTo(int crse_lev);
    int FillCoarsePatch(int lev, Real time, MultiFab& mf, int icomp, int ncomp);
    int finest_level() const { return fine_new_.empty() ? 0 : 1; }
    const std::vector<vd_box>& fine_grids() const { return fine_ba_; }
    std::vector<MultiFab>& fine_phi_new() { return fine_new_; }

    MultiFab& phi_new() { return phi_new_[0]; }
    // after an Evolve with one corrector the start-of-step species and EEN
    // are taken from Sborder (which holds them unchanged) on first access
    MultiFab& phi_old() {
        if (phi_old_stale_ && sb_last_) {
            phi_old_[0].copy_from(*sb_last_, 0, 0, ns_ + 1);
            phi_old_stale_ = false;
        }
        return phi_old_[0];
    }
    Real time() const { return t_new_; }
    void set_time(Real t) { t_new_ = t; }
    int step() const { return istep_; }
    const vd_solve_info& last_potential_info() const { return pot_info_; }
    long long mg_iterations() const { return mg_iters_; }
    std::vector<std::string> varnames() const;

   private:
    int alloc_step_data();
    int step_body(std::vector<MultiFab>& sb, std::vector<MultiFab>& sb_old, Real cur_time, Real dt,
                  bool pot_in_sb);
    int evolve_one_corrector(int nsteps);
    int evolve_corrected(int nsteps, int ncorr);
    // where a step writes its results and reads the state (phi_new unless
    // the one-corrector driver rotates buffers)
    MultiFab& out() { return out_ ? *out_ : phi_new_[0]; }
    MultiFab& state() { return state_ ? *state_ : phi_new_[0]; }
    vd_ctx* ctx_;
    vd_mech* mech_ = nullptr;
    vd_mech_desc desc_;
    vd_geom geom_;
    vd_params prm_;
    vd_box dom_;
    int ns_, nvar_, dim_, ngrow_fp_;
    int status_ = VD_OK;
    Real t_new_ = 0.0;
    Real dt_ = 0.0;  // the last step's dt (checkpoint Header)
    int istep_ = 0;
    bool phi_old_stale_ = false;
    std::vector<MultiFab> phi_new_, phi_old_;
    // persistent step data (the reference allocates these every step);
    // Swork_ is the second ghosted state buffer of the one-corrector driver
    std::vector<MultiFab> Sborder_, Sborder_old_, Swork_, expl_src_, rxn_src_;
    MultiFab* out_ = nullptr;
    MultiFab* state_ = nullptr;
    MultiFab* sb_last_ = nullptr;  // Sborder of the last one-corrector step
    std::vector<FaceMF> flux_, efield_ec_, gradne_fc_, grad_fc_;
    vd_solve_info pot_info_{};
    long long mg_iters_ = 0;
    std::vector<int> tag_comps_;
    std::vector<double> tag_thr_, tag_gthr_;
    // level 1: geometry, BoxArray and per-box state (no ghost cells)
    vd_geom fgeom_{};
    int ratio_[3] = {2, 2, 1};
    std::vector<vd_box> fine_ba_;
    std::vector<MultiFab> fine_new_, fine_old_;
    Real t_new_fine_ = 0.0;
    int set_fine_geom(const int ratio[3]);
    int fill_fine(MultiFab& mf, const std::vector<MultiFab>* src, const std::vector<vd_box>* src_ba, int icomp,
                  int ncomp);
};

}  // namespace vidyut_amd
