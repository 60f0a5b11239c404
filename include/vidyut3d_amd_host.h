/*
 * vidyut3d_amd_host.h -- C handle over the C++ host driver (class
 * vidyut_amd::Vidyut in vidyut3d_amd/host/Vidyut.H), for callers that drive
 * whole steps (Source/main.cpp -> Vidyut::InitData / Evolve) instead of the
 * per-kernel entry points of vidyut3d_amd.h.
 */
#ifndef VIDYUT3D_AMD_HOST_H
#define VIDYUT3D_AMD_HOST_H

#include "vidyut3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vd_sim vd_sim;

/* Vidyut::Vidyut + the level-0 MultiFabs (single box covering the domain) */
int vd_sim_create(vd_ctx* ctx, const vd_mech_desc* mech, const vd_geom* geom, const vd_params* prm,
                  vd_sim** out);
int vd_sim_destroy(vd_sim* s);
/* whole state (NVAR comps, valid box, component-major, i fastest) */
int vd_sim_set_state(vd_sim* s, const double* host, double time);
int vd_sim_get_state(vd_sim* s, double* host);
/* A simulation on a batched context (vd_ctx_create_batched) holds one
 * case per region, all of one shape and vd_params (per-case voltages via
 * vd_ctx_set_case_voltage; adaptive_dt is refused), and vd_sim_evolve
 * steps the active cases (vd_ctx_set_active) together; the cases of the
 * serving loop's same-shape batches.  Case c's whole state: */
int vd_sim_set_state_case(vd_sim* s, int c, const double* host);
int vd_sim_get_state_case(vd_sim* s, int c, double* host);
/* device view of phi_new (no ghost cells; case 0's in a batched context) */
int vd_sim_state_view(vd_sim* s, vd_array4* out);
/* Vidyut::Evolve for nsteps with ncorr correctors.  dt is vidyut.dt: the
 * step of every step, or with prm->adaptive_dt the step before the
 * simulation's first (each step's dt then comes from ComputeDt of the
 * previous dt, continuing across calls and restarts) */
int vd_sim_evolve(vd_sim* s, int nsteps, double dt, int ncorr);
double vd_sim_time(const vd_sim* s);
/* the last step's dt (0 before the first step) */
double vd_sim_dt(const vd_sim* s);
long long vd_sim_mg_iterations(const vd_sim* s);
/* Vidyut::WritePlotFile (Source/Plot.cpp:29-38): AMReX HyperCLaw-V1.1 */
int vd_sim_write_plotfile(vd_sim* s, const char* dir);
/* Vidyut::WriteCheckpointFile (Source/Plot.cpp:40-119): Header (finest
 * level, istep, dt, t_new, BoxArray) + VisMF Level_0/phi of phi_new */
int vd_sim_write_checkpoint(vd_sim* s, const char* dir);
/* Vidyut::ReadCheckpointFile (Source/Plot.cpp:121-209): restores istep, dt,
 * t_new and phi_new from a single-level checkpoint of the same domain */
int vd_sim_read_checkpoint(vd_sim* s, const char* dir);
/* istep of level 0 */
int vd_sim_step(const vd_sim* s);
/* ErrorEst on phi_new (1-ghost FillPatch then the Tagging.H kernels);
 * tags is a host array over the valid box */
int vd_sim_error_est(vd_sim* s, int ntag, const int* comps, const double* thr, const double* gthr,
                     char* host_tags);

/* ------------------------------------------------------------- regrid
 * AmrMesh::MakeNewGrids for the level above a single-box level, as
 * Vidyut::Evolve's regrid(0, time) runs it every regrid_int steps
 * (Source/Evolve.cpp:66-72) on the tags of Vidyut::ErrorEst
 * (Source/Vidyut.cpp:171-255), with the amr.* keys below.  Host code: no
 * device work. */
typedef struct vd_regrid_params {
    int ref_ratio[3];       /* amr.ref_ratio: 2 or 4 */
    int blocking_factor[3]; /* amr.blocking_factor of the new level (multiple of ref_ratio) */
    int max_grid_size[3];   /* amr.max_grid_size of the new level (multiple of blocking_factor) */
    int n_error_buf;        /* amr.n_error_buf: cells of buffer around each tag (AMReX default 1) */
    double grid_eff;        /* amr.grid_eff: minimum tagged fraction of a box (AMReX default 0.7) */
} vd_regrid_params;

/* tags: host char array over `domain` (level lev, i fastest), nonzero =
 * tagged.  The tags are buffered by n_error_buf cells (periodic
 * directions wrap, cells beyond other faces are dropped), coarsened by
 * blocking_factor / ref_ratio, clustered (Berger-Rigoutsos: a box is kept
 * once its tagged fraction reaches grid_eff, else cut at a hole of the
 * tag signature, else at the strongest inflection of its Laplacian, else
 * in the middle of its longest side), refined to the new level and
 * chopped to max_grid_size.  Writes at most max_out boxes (fine index
 * space) to out and their number to *nout (VD_EINVAL, with the count
 * needed in *nout, when out is too small). */
int vd_make_new_grids(const char* tags, const vd_box* domain, const int is_periodic[3], int dim,
                      const vd_regrid_params* prm, vd_box* out, int max_out, int* nout);

/* AMR level 1 of a simulation (the step driver advances level 0 only).
 * Vidyut::regrid(0, time) (Source/Evolve.cpp:66-72): ErrorEst on level 0
 * with these tagging rules, vd_make_new_grids, then
 * MakeNewLevelFromCoarse (Source/GridWork.cpp:17-29) for a new level 1,
 * RemakeLevel (:33-49) for an existing one, or ClearLevel (:53-57) when no
 * cell is tagged.  *nboxes = the number of level-1 boxes afterwards. */
int vd_sim_regrid(vd_sim* s, int ntag, const int* comps, const double* thr, const double* gthr,
                  const vd_regrid_params* rp, int* nboxes);
/* the boxes of level lev (0: the domain; 1: the level-1 BoxArray) */
int vd_sim_level_boxes(vd_sim* s, int lev, vd_box* out, int max_out, int* n);
/* phi_new of level-1 box b (NVAR comps over the box, component-major) */
int vd_sim_get_fine(vd_sim* s, int b, double* host);
int vd_sim_set_fine(vd_sim* s, int b, const double* host);
/* Vidyut::AverageDown (Source/GridWork.cpp:94-101) */
int vd_sim_average_down(vd_sim* s);

#ifdef __cplusplus
}
#endif
#endif
