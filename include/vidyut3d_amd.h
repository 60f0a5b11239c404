/*
 * vidyut3d_amd.h -- C ABI of the MI355X (gfx950) implementation of Vidyut's
 * per-timestep hot path: drift-diffusion WENO face fluxes, species and
 * electron-energy sources, the variable-coefficient Poisson / implicit
 * species (Helmholtz) multigrid solves, the CFL reduction and the
 * refinement tags.
 *
 * Every hot-path entry point works on ONE box of a single-level MultiFab,
 * passed as an Array4-equivalent POD (vd_array4) whose pointer is DEVICE
 * memory.  The C++ host layer (vidyut3d_amd/host/Vidyut.H) keeps the
 * reference's Vidyut member signatures and calls these per box.
 *
 * All calls are asynchronous on the context's HIP stream unless stated.
 * No call aborts the process: failures come back as a VD_* status and
 * vd_last_error() holds a message.
 */
#ifndef VIDYUT3D_AMD_H
#define VIDYUT3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    VD_OK = 0,
    VD_EINVAL = 1,   /* bad argument: hyp_order, weno_scheme, bc code, shape */
    VD_ENOCONV = 2,  /* linear solver hit maxiter above its target        */
    VD_ENOMEM = 3,   /* device allocation failed / serving admission full   */
    VD_EHIP = 4,     /* HIP runtime error                                   */
    VD_ENOTFOUND = 5 /* named species / variable not found                 */
};

/* --------------------------------------------- state component layout
 * Source/VarDefines.H:5-16 : species 0..NS-1, then the plasma variables. */
#define VD_NUM_PLASMAVARS 10
#define VD_EEN(ns) ((ns) + 0)
#define VD_ETEMP(ns) ((ns) + 1)
#define VD_POT(ns) ((ns) + 2)
#define VD_EFX(ns) ((ns) + 3)
#define VD_EFY(ns) ((ns) + 4)
#define VD_EFZ(ns) ((ns) + 5)
#define VD_EJH(ns) ((ns) + 6)
#define VD_EIH(ns) ((ns) + 7)
#define VD_EEH(ns) ((ns) + 8)
#define VD_REF(ns) ((ns) + 9)
#define VD_NVAR(ns) ((ns) + VD_NUM_PLASMAVARS)

/* boundary-condition codes, Source/VarDefines.H:18-22 */
enum { VD_PERBC = 0, VD_DIRCBC = 1, VD_HNEUBC = 2, VD_IHNEUBC = 3, VD_ROBINBC = 4 };

/* ------------------------------------------------------------ data model
 * Array4 semantics: element (i,j,k,n) lives at
 *   p[(i-begin[0]) + (j-begin[1])*jstride + (k-begin[2])*kstride + n*nstride]
 * begin/end (exclusive) bound the allocated box INCLUDING ghost cells.
 * In 2-D, begin[2] = 0 and end[2] = 1. */
typedef struct vd_array4 {
    double* p;
    int begin[3];
    int end[3];
    int64_t jstride, kstride, nstride;
    int ncomp;
} vd_array4;

/* same layout for the char TagBox (Source/Tagging.H) */
typedef struct vd_tagarray4 {
    char* p;
    int begin[3];
    int end[3];
    int64_t jstride, kstride;
} vd_tagarray4;

/* index box (cell-centred, inclusive) */
typedef struct vd_box {
    int lo[3];
    int hi[3];
} vd_box;

/* amrex::Geometry of one level */
typedef struct vd_geom {
    int dim;            /* AMREX_SPACEDIM: 2 or 3 */
    int dom_lo[3];
    int dom_hi[3];
    double prob_lo[3];
    double prob_hi[3];
    double dx[3];
    int is_periodic[3];
} vd_geom;

/* The vidyut.* ParmParse keys of the hot path (Source/Vidyut.H:149-272) */
typedef struct vd_params {
    double gas_temperature; /* K  (vidyut.gas_temperature) */
    double gas_pressure;    /* Pa (vidyut.gas_pressure)    */
    int hyp_order;          /* 1 upwind, 2 WAF-MINBEE, 5 WENO */
    int weno_scheme;        /* 1 JS, 2 M, 3 Z */
    int elecenergy_solve;
    int do_reactions;
    int do_transport;
    int do_spacechrg;
    int do_bg_reactions;
    int bound_specden;
    double min_species_density;
    double min_electron_density;
    double min_electron_temp;
    int floor_jh;
    int const_ele_trans;
    double ele_mob;
    double ele_diff;
    int voltage_profile;    /* 0 DC, 1 sine, 2 triangle (HelperFuncs.H:75-92) */
    double voltage_amp_1, voltage_amp_2, voltage_freq, voltage_dur, voltage_center;
    double linsolve_reltol;
    double linsolve_abstol;
    int linsolve_maxiter;
    int linsolve_max_coarsening_level;
    int mg_pre_smooth;      /* GSRB sweeps before restriction  */
    int mg_post_smooth;     /* GSRB sweeps after prolongation  */
    int mg_precision;       /* 0: all-FP64 V-cycles; 1: FP32 V-cycles in FP64 IR */
    int pot_bc_lo[3], pot_bc_hi[3];
    int eden_bc_lo[3], eden_bc_hi[3];
    int eenrg_bc_lo[3], eenrg_bc_hi[3];
    int ion_bc_lo[3], ion_bc_hi[3];
    int neutral_bc_lo[3], neutral_bc_hi[3];
    /* Vidyut::ComputeDt (Source/TimeStep.cpp:14-31): vidyut.adaptive_dt and
     * its keys (Source/Vidyut.cpp:281-291, defaults Source/Vidyut.H:173-182) */
    int adaptive_dt;
    double advective_cfl, diffusive_cfl, dielectric_cfl;
    double dt_min, dt_max, adaptive_dt_delay, dt_stretch;
} vd_params;

/* fills the reference defaults of Source/Vidyut.H:149-272 */
void vd_params_default(vd_params* p);

/* ----------------------------------------------------- case plug-in ABI
 * Replaces the compile-time Chemistry.H / UserFunctions.H of a case:
 * species table (CKSYMS_STR, CKCHRG, mw), transport (specMob, specDiff)
 * and a table-driven reaction mechanism. */
#define VD_MAX_SPECIES 48
#define VD_MAX_REACTIONS 512
#define VD_MAX_REACTANTS 3
#define VD_MAX_PRODUCTS 4
#define VD_GEN_NG 8  /* g/RT terms of one reverse rate (VD_CHEM_GEN) */
#define VD_GEN_NW 8  /* wdot updates of one reaction (VD_CHEM_GEN) */

/* chemistry forms */
enum {
    VD_CHEM_TABLE = 0, /* Source/reaction_source.H:10-64: number densities,
                          rate = k*prod n^nu, energy += e_r*rate          */
    VD_CHEM_CKWC = 1,  /* generated Chemistry.H CKWC/productionRate form:
                          mol units round trip + comp_ener_exch           */
    VD_CHEM_GEN = 2    /* a generated productionRate read as data
                          (vidyut3d_amd/mechgen.py): the rate kinds,
                          third bodies, reverse rates from g/RT and wdot
                          updates below, in mol/m^3 and the file's units,
                          behind the same CKWC round trip                  */
};

/* VD_CHEM_GEN forward rate kinds, as productionRate writes them
 * (py_utils/ceptr_noThermo; test/Reactor/CO2_H2_Ar_mechanism_reduced/
 * Chemistry.H:913-6380): T = gas temperature, TeV = Te / 11595,
 * NA = 6.02214085774e23 */
enum {
    VD_RATE_ARRH = 0,      /* A exp(beta ln(T/300) - Ea / T)                */
    VD_RATE_TEPOW = 1,     /* A pow(Te / tref, beta), / T when tgas_div     */
    VD_RATE_TEARRH = 2,    /* A pow(Te, beta) exp(-Ea / Te)                 */
    VD_RATE_JANEV = 3,     /* A exp(sum_{j<9} fit[j] pow(ln TeV, j)) NA     */
    VD_RATE_FFIT = 4,      /* A exp(sum_{j<4} fit[j] pow(1/TeV, j+1)) NA    */
    VD_RATE_JANEV_FFIT = 5 /* JANEV (A, fit) for TeV < fit_switch_ev, else
                              FFIT (A2, fit2)                               */
};

/* diffusion models per species (specDiff) */
enum {
    VD_DIFF_EINSTEIN_TE = 0, /* D = (Te/eV)*|mu|   (electrons)          */
    VD_DIFF_EINSTEIN_TG = 1, /* D = (Tg/eV)*|mu|   (ions / neutrals)    */
    VD_DIFF_CONST = 2        /* D = diff[s]                             */
};

/* State-dependent transport coefficients, the form of a case Chemistry.H's
 * specMob / specDiff(specid, Te, ndens, |E|, Tg) (Source/PlasmaChem.H:
 * 186-206).  For kind VD_TR_FIT the coefficient is
 *   sign * exp( c0 + c1 ln|E| + c2 ln Te + c3 ln Tg + c4 ln N
 *             + c5 ln e + c6/e + c7/e^2 + c8/e^3
 *             + c9 ln(1 + c10 x^2 + c11 x^4) )
 * with N = sum of all species densities, e = 1.5 Te / (1 eV) the mean
 * electron energy in eV and x = |E| / N; terms whose coefficient is 0 are
 * skipped.  This covers the reference cases' constant, power-law
 * (test/verification/Streamer/Chemistry.H:165-200), hard-sphere
 * (test/models/Ar_DC_1d/Chemistry.H specMob) and fitted
 * (test/verification/He_RF_1d/Chemistry.H specMob) coefficients. */
#define VD_TR_NCOEF 12
enum {
    VD_TR_DEFAULT = 0, /* mob[] / diff_kind[] of the species (constant mobility) */
    VD_TR_FIT = 1      /* the fit above */
};
typedef struct vd_transport_fit {
    int kind;          /* VD_TR_* */
    double sign;       /* sign of the coefficient (a mobility carries the charge sign) */
    double c[VD_TR_NCOEF];
} vd_transport_fit;

typedef struct vd_reaction {
    double A;          /* k = A * T^beta * exp(-Ea/T), SI units */
    double beta;
    double Ea;         /* K */
    int temp_is_te;    /* T = Te (1) or Tg (0) */
    int nreact, nprod;
    int react_sp[VD_MAX_REACTANTS];
    double react_nu[VD_MAX_REACTANTS];
    int prod_sp[VD_MAX_PRODUCTS];
    double prod_nu[VD_MAX_PRODUCTS];
    double energy;     /* VD_CHEM_TABLE: J per event into the electron energy */
    int rxntype;       /* VD_CHEM_CKWC: 0 none, 1 excitation, 2 ionization,
                          3 recombination, 5 elastic exchange               */
    double eexci;      /* VD_CHEM_CKWC: eV */
    int elidx;         /* VD_CHEM_CKWC: collision partner for rxntype 5 */
    /* VD_CHEM_GEN (rxntype also 4 attachment and 7 detachment; reactants
     * in the order of the forward product, which multiplies them in turn) */
    int rate_kind;     /* VD_RATE_* */
    double tref;       /* VD_RATE_TEPOW */
    int tgas_div;      /* VD_RATE_TEPOW */
    double fit[9];
    double A2;
    double fit2[4];
    double fit_switch_ev;
    double kmax;       /* > 0: k_f = min(k_f, kmax) */
    int third_body;    /* qf and qr times the mixture concentration */
    int reversible;    /* qr = k_f exp(-(sum_q g_nu[q] g/RT(g_sp[q])))
                          refC^refc_pow prod_q C(rev_sp[q])^rev_nu[q],
                          refC = 101325 / 8.31446 / T */
    int ng;
    int g_sp[VD_GEN_NG];
    double g_nu[VD_GEN_NG];
    int refc_pow;
    int nrev;
    int rev_sp[VD_MAX_PRODUCTS];
    double rev_nu[VD_MAX_PRODUCTS];
    int nwd;           /* wdot[wd_sp[q]] += wd_nu[q] * (qf - qr), in order */
    int wd_sp[VD_GEN_NW];
    double wd_nu[VD_GEN_NW];
} vd_reaction;

typedef struct vd_mech_desc {
    int nspec;                       /* NUM_SPECIES            */
    int nreac;                       /* NUM_REACTIONS          */
    int e_idx;                       /* electron index (E_IDX) */
    char names[VD_MAX_SPECIES][16];  /* CKSYMS_STR            */
    int charge[VD_MAX_SPECIES];      /* CKCHRG                */
    double molwt[VD_MAX_SPECIES];    /* mw(), g/mol           */
    int is_bg[VD_MAX_SPECIES];       /* vidyut.bg_species_ids */
    double mob[VD_MAX_SPECIES];      /* signed mobility m^2/V/s (specMob) */
    int diff_kind[VD_MAX_SPECIES];   /* VD_DIFF_*             */
    double diff[VD_MAX_SPECIES];     /* for VD_DIFF_CONST     */
    int chem_form;                   /* VD_CHEM_*             */
    int elastic_loss;                /* VD_CHEM_CKWC: add the e-neutral elastic loss */
    double elastic_partner_amu;      /* ... with this partner mass */
    vd_reaction reac[VD_MAX_REACTIONS];
    /* specMob / specDiff, index nspec = the electron energy.  With
     * VD_TR_DEFAULT the energy takes 5/3 of the electron mobility and
     * (5/3)(Te/1 eV)|mu_e|; Einstein diffusion uses the species' own
     * (possibly fitted) mobility. */
    vd_transport_fit mob_fit[VD_MAX_SPECIES + 1];
    vd_transport_fit diff_fit[VD_MAX_SPECIES + 1];
    /* user_transport::compute_vel (selected by vidyut.user_defined_vel,
     * Source/ScalarSolve.cpp:237-243): a prescribed drift velocity per
     * direction replaces mobility * E in the face fluxes of species s */
    int vel_set[VD_MAX_SPECIES + 1];
    double vel[VD_MAX_SPECIES + 1][3];
    /* user_transport::species_bc Robin data (vidyut.user_defined_species,
     * Source/ScalarSolve.cpp:542-553): where species s has the VD_ROBINBC
     * code, a*n + b*dn/dn_out = f on the face, per direction and side (0
     * low, 1 high); without robin_set the reference default (0, 1, 0) */
    int robin_set[VD_MAX_SPECIES + 1];
    double robin[VD_MAX_SPECIES + 1][3][2][3];
    /* user_sources::add_user_react_sources: a constant added to each
     * species' (and, at nspec, the electron energy's) reaction source */
    double user_src[VD_MAX_SPECIES + 1];
    /* VD_CHEM_GEN: g/RT of each species (gibbs() of the generated file),
     * c0/T + c1 + c2 ln T + c3 T + c4 T^2 + c5 T^3 + c6 T^4 with gibbs_lo
     * below gibbs_tmid and gibbs_hi at and above it */
    double gibbs_tmid[VD_MAX_SPECIES];
    double gibbs_lo[VD_MAX_SPECIES][7];
    double gibbs_hi[VD_MAX_SPECIES][7];
} vd_mech_desc;

/* ---------------------------------------------------------- handles */
typedef struct vd_ctx vd_ctx;
typedef struct vd_mech vd_mech;

const char* vd_last_error(void);
const char* vd_version(void);

/* one context per host thread; owns a HIP stream, solver workspaces and
 * a device-memory pool reused across calls (no allocation after warm-up) */
int vd_ctx_create(int device, vd_ctx** out);
int vd_ctx_destroy(vd_ctx* ctx);
int vd_ctx_sync(vd_ctx* ctx);
/* frees the cached solver workspaces and scratch buffers (they are
 * re-created on demand); synchronises.  For long-lived contexts that
 * serve cases of many grid sizes. */
int vd_ctx_release_workspaces(vd_ctx* ctx);
void* vd_ctx_stream(vd_ctx* ctx); /* the hipStream_t, as void* */

/* ---------------------------------------------------- batched contexts
 * Same-shape cases stepped together, one launch per kernel for all of
 * them (the serving loop's batching; no counterpart in the reference,
 * which runs one case per executable).  A batched context holds ncase
 * identical regions of case_bytes device memory; every allocation the
 * context makes (vd_malloc, the solver workspaces, scratch, a mechanism's
 * device copy) is carved out of case 0's region, and case c's copy of it
 * sits case_bytes * c further on.  Every entry point of the step path
 * (transport, sources, Joule heating, the potential and implicit solves,
 * time scales, ghost fill, the bulk MultiFab operations) then runs on the
 * active cases at once: launches take one z-block per case, and the
 * solvers keep each case's convergence test, so every case takes exactly
 * the operations, and gets exactly the bits, it would alone.  Arrays
 * passed to those entry points must lie in case 0's region; the other
 * entry points (AMR, several boxes, tags, the no-allocation and generic
 * solver calls) reject a batched context with more than one active case.
 * A region that runs out gives VD_ENOMEM (vd_ctx_arena_used reports the
 * bytes one case has taken). */
#define VD_MAX_BATCH 16
struct vd_solve_info;
int vd_ctx_create_batched(int device, int ncase, size_t case_bytes, vd_ctx** out);
int vd_ctx_ncase(const vd_ctx* ctx);
/* byte distance between two cases' copies of an allocation */
long long vd_ctx_case_stride(const vd_ctx* ctx);
int vd_ctx_arena_used(const vd_ctx* ctx, size_t* used);
/* the cases (distinct, ascending or not, < ncase) the following calls run
 * on; n = 0 is refused */
int vd_ctx_set_active(vd_ctx* ctx, int n, const int* cases);
/* per-case applied voltage amplitudes (vidyut.voltage_amp_1/2) used by
 * the potential solve in place of vd_params' (until cleared with
 * use = 0) */
int vd_ctx_set_case_voltage(vd_ctx* ctx, int c, int use, double amp_1, double amp_2);
/* the iterations and norms of case c's last potential or implicit solve */
int vd_ctx_case_solve_info(const vd_ctx* ctx, int c, struct vd_solve_info* info);
/* case c's time scales from the last vd_find_time_scales */
int vd_ctx_case_time_scales(const vd_ctx* ctx, int c, double out[3]);

int vd_malloc(vd_ctx* ctx, size_t bytes, void** out);
int vd_free(vd_ctx* ctx, void* p);
int vd_memcpy_h2d(vd_ctx* ctx, void* dst, const void* src, size_t bytes);
int vd_memcpy_d2h(vd_ctx* ctx, void* dst, const void* src, size_t bytes);
int vd_memcpy_d2d(vd_ctx* ctx, void* dst, const void* src, size_t bytes);
int vd_memset_zero(vd_ctx* ctx, void* dst, size_t bytes);

/* replaces plasmachem::init (Source/PlasmaChem.cpp:12-19) and the case
 * Chemistry.H tables; copies the descriptor to device memory */
int vd_mech_create(vd_ctx* ctx, const vd_mech_desc* desc, vd_mech** out);
int vd_mech_destroy(vd_mech* m);
/* plasmachem::find_id (Source/PlasmaChem.cpp:21-30); -1 when absent */
int vd_mech_find_id(const vd_mech* m, const char* name);
/* Mechanism-specialised source kernel for table mechanisms: the reaction
 * table is compiled into straight-line gfx950 code, the run-time
 * counterpart of the reference compiling a case's generated Chemistry.H
 * into the executable (test/models/ionelec_DC_1d/Chemistry.H:240-357).
 * vd_mech_create builds it by default (environment VD_CHEM_SPECIALISE=0
 * turns that off); enable = 1 builds and uses it, 0 returns to the
 * generic table kernel.  VD_EINVAL for CKWC mechanisms. */
int vd_mech_specialise(vd_mech* m, int enable);
int vd_mech_is_specialised(const vd_mech* m);

/* ---------------------------------------------- bulk MultiFab operations
 * MultiFab::setVal / MultiFab::Copy / MultiFab::LinComb on the valid box
 * (Source/Evolve.cpp:105-171,290-300; Source/ScalarSolve.cpp:443-465). */
int vd_setval(vd_ctx* ctx, vd_array4* a, const vd_box* box, int scomp, int ncomp, double v);
int vd_copy(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box, int scomp,
            int dcomp, int ncomp);
/* dst[dcomp+n] = a*x[xcomp+n] + b*y[ycomp+n] */
int vd_lincomb(vd_ctx* ctx, vd_array4* dst, double a, const vd_array4* x, int xcomp, double b,
               const vd_array4* y, int ycomp, int dcomp, int ncomp, const vd_box* box);
/* a[comp] = max(a[comp], lo) on the valid box */
int vd_clamp_min(vd_ctx* ctx, vd_array4* a, const vd_box* box, int comp, double lo);

/* ------------------------------------------------------ hot path
 * `box` is the valid (cell-centred) box being processed; array views
 * must cover box grown by the ghost width each call needs. */

/* Vidyut::FillPatch single level (Source/GridWork.cpp:112-124 with the
 * foextrap / int_dir bcspec of Source/Vidyut.cpp:93-104): copies comps
 * [scomp, scomp+ncomp) of src (valid data) into dst and fills dst's
 * ghost cells by periodic wrap or first-order extrapolation. */
int vd_fill_patch(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box,
                  const vd_geom* geom, int scomp, int ncomp);

/* Vidyut::FillPatch single level over a BoxArray of nbox boxes (level 0
 * chopped by amr.max_grid_size; Source/GridWork.cpp:112-124, FillBoundary
 * between the boxes, periodic images and foextrap faces): src[b] holds the
 * valid data of boxes[b], dst[b] (grown) receives comps [scomp,
 * scomp+ncomp) on its whole extent, each cell from the domain cell found
 * by wrapping periodic indices and clamping the others, read from the box
 * that owns it.  The boxes must tile the domain without overlap; dst may
 * be src (ghost cells filled in place). */
int vd_fill_patch_boxes(vd_ctx* ctx, int nbox, const vd_array4* src, vd_array4* dst,
                        const vd_box* boxes, const vd_geom* geom, int scomp, int ncomp);
/* The valid cells of a level held in nbox boxes (tiling the domain) into
 * one array over the domain, and back: the potential and implicit solves
 * (vd_solve_potential, vd_implicit_solve_scalar, vd_mlmg_solve) take a
 * level as one box, as the MLMG of Source/PotentialSolve.cpp:148-336
 * takes the whole MultiFab. */
int vd_level_gather(vd_ctx* ctx, int nbox, const vd_array4* parts, const vd_box* boxes,
                    vd_array4* whole, const vd_geom* geom, int scomp, int ncomp);
int vd_level_scatter(vd_ctx* ctx, int nbox, const vd_array4* whole, vd_array4* parts,
                     const vd_box* boxes, const vd_geom* geom, int scomp, int ncomp);

/* Vidyut::compute_scalar_transport_flux (Source/ScalarSolve.cpp:170-255)
 * -> compute_flux (Source/compute_explicit_flux.H:200-326) on every face
 * of `box` in each direction.  sborder needs hyp_order-dependent ghosts
 * (1, 2, 3 for hyp_order 1, 2, 5).  flux[d] is the face box. */
int vd_compute_scalar_transport_flux(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                     vd_array4 flux[3], const vd_box* box, const vd_geom* geom,
                                     const int bc_lo[3], const int bc_hi[3], const vd_params* prm,
                                     double time, double dt, int specid);

/* Vidyut::compute_dsdt (Source/ScalarSolve.cpp:17-55) */
int vd_compute_dsdt(vd_ctx* ctx, const vd_array4 flux[3], const vd_array4* rxn_src,
                    vd_array4* dsdt, const vd_box* box, const vd_geom* geom, int specid);

/* compute_scalar_transport_flux + compute_dsdt fused (Source/ScalarSolve.cpp
 * :57-111 update_explsrc_at_all_levels for one level): face fluxes stay
 * on chip; only dsdt is written.  Same results as the two calls above. */
int vd_transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                     const vd_array4* rxn_src, vd_array4* dsdt, const vd_box* box,
                     const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                     const vd_params* prm, double time, double dt, int specid);

/* vd_transport_rhs without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Same arguments and results as
 * vd_transport_rhs; the face fluxes it keeps between its kernels live in a
 * fixed per-GPU array of VD_NOALLOC_SCRATCH_BYTES (held from the library's
 * load until the process ends) instead of the context's buffers.  A box
 * whose face fluxes do not fit (the x and, in 3-D, y faces of a WENO box,
 * every direction otherwise, each rounded up to 256 bytes; about 160^3
 * cells for 3-D WENO) is rejected with VD_EINVAL.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams or
 * host threads) race, so the caller orders them. */
#define VD_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int vd_transport_rhs_noalloc(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                             const vd_array4* rxn_src, vd_array4* dsdt, const vd_box* box,
                             const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                             const vd_params* prm, double time, double dt, int specid);

/* Vidyut::update_rxnsrc_at_all_levels (Source/ScalarSolve.cpp:113-168)
 * with the mechanism's CKWC / reaction_source.H source: rxn_src gets
 * NS species rates (1/m^3/s) and, at comp NS, the electron energy
 * source (J/m^3/s). */
int vd_update_rxnsrc(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder, vd_array4* rxn_src,
                     const vd_box* box, const vd_params* prm, double time);

/* Vidyut::compute_elecenergy_source (Source/ElecEnergySource.cpp:16-173):
 * Joule heating from face E and face grad(n_e); dsdt += heat; writes
 * phi_new EJH and EIH. */
int vd_compute_elecenergy_source(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                 const vd_array4* rxn_src, const vd_array4 efield[3],
                                 const vd_array4 gradne[3], vd_array4* dsdt, vd_array4* phi_new,
                                 const vd_box* box, const vd_geom* geom, const vd_params* prm);

/* solver report */
typedef struct vd_solve_info {
    int iterations;       /* outer (V-cycle / IR) iterations */
    double resnorm0;      /* max-norm of the initial residual */
    double resnorm;       /* max-norm of the final residual   */
    double target;        /* max(abstol, reltol*max(|b|, resnorm0)) */
} vd_solve_info;

/* Vidyut::solve_potential (Source/PotentialSolve.cpp:19-393): rhs
 * -e/eps0*sum z n, Dirichlet/Neumann/periodic/Robin BCs with applied
 * voltage, multigrid solve, E = -grad(phi) on faces (efield), cell
 * averages into phi_new EFX..EFZ, potential into POT and E/N into REF. */
int vd_solve_potential(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder, vd_array4* phi_new,
                       vd_array4 efield[3], const vd_box* box, const vd_geom* geom,
                       const vd_params* prm, double time, vd_solve_info* info);

/* Vidyut::implicit_solve_scalar (Source/ScalarSolve.cpp:289-680):
 * (1/dt) n - div(D grad n) = n_old/dt + dsdt_expl with harmonic face D,
 * species boundary conditions, floors, grad n_e on faces (electrons) and
 * the Te update (electron energy).  Writes phi_new comp `spec`. */
int vd_implicit_solve_scalar(vd_ctx* ctx, const vd_mech* m, int spec, const vd_array4* sborder,
                             const vd_array4* phi_old, const vd_array4* dsdt_expl,
                             vd_array4* phi_new, vd_array4 grad_fc[3], const vd_box* box,
                             const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                             const vd_params* prm, double time, double dt, vd_solve_info* info);

/* Vidyut::find_time_scales (Source/TimeStep.cpp:34-116): wave64
 * max-reductions; out = {dt_edrift, dt_ediff, dt_diel_relax}.  Blocks
 * until the three scalars are on the host. */
int vd_find_time_scales(vd_ctx* ctx, const vd_mech* m, const vd_array4* phi_new, const vd_box* box,
                        const vd_geom* geom, const vd_params* prm, double out[3]);

/* Vidyut::ComputeDt (Source/TimeStep.cpp:14-31) on host scalars: with
 * adaptive_dt and cur_time > adaptive_dt_delay, the next dt is
 * min(old_dt * dt_stretch, clamp(min(dt_edrift * advective_cfl (with
 * transport), dt_ediff * diffusive_cfl (with transport), dt_diel *
 * dielectric_cfl), dt_min, dt_max)); otherwise old_dt.  ts = the three
 * time scales of vd_find_time_scales. */
int vd_compute_dt(const vd_params* prm, double cur_time, double old_dt, const double ts[3], double* dt);

/* Vidyut::ErrorEst kernels (Source/Tagging.H:8-69, Source/Vidyut.cpp:
 * 221-253): state needs 1 ghost; tags[cell] = tagval where
 * |q| > refine_phi[c] or the one-sided gradient exceeds refine_phigrad[c]. */
int vd_error_est(vd_ctx* ctx, const vd_array4* state, vd_tagarray4* tags, const vd_box* box,
                 int ntagvars, const int* comps, const double* refine_phi,
                 const double* refine_phigrad, char tagval);

/* General cell-centred ABecLaplacian solve, the operator of
 * MLABecLaplacian as used by PotentialSolve.cpp / ScalarSolve.cpp:
 *   alpha_a*phi - beta*div(bface grad phi) = rhs
 * bface[d] are face arrays (NULL => constant 1), bc per side with values
 * in bcval (NULL => homogeneous): Dirichlet value, inhomogeneous-Neumann
 * d(phi)/dx_d, or Robin (robin_a, robin_b, robin_f; NULL => not Robin).
 * phi holds the initial guess on entry (valid region). */
typedef struct vd_mlmg_problem {
    double alpha_a;
    double beta;
    const vd_array4* bface[3];
    int bc_lo[3], bc_hi[3];
    const vd_array4* bcval;     /* ghost cells of a 1-ghost array hold BC data */
    const vd_array4* robin_a;
    const vd_array4* robin_b;
    const vd_array4* robin_f;
    double reltol, abstol;
    int maxiter;
    int precision;              /* 0 FP64, 1 FP32 V-cycles in FP64 IR */
    int pre_smooth, post_smooth;
    int max_coarsening_level;
} vd_mlmg_problem;

int vd_mlmg_solve(vd_ctx* ctx, const vd_mlmg_problem* prob, vd_array4* phi, const vd_array4* rhs,
                  const vd_box* box, const vd_geom* geom, vd_solve_info* info);

/* ---------------------------------------------- level transfers (AMR)
 * Level l+1 refines level l by ratio[d] (2 or 4; 1 in the unused third
 * direction of a 2-D run).  Fine index i lies in coarse cell
 * floor(i / ratio).  `dim` is 2 or 3. */

/* Vidyut::AverageDown / AverageDownTo (Source/GridWork.cpp:94-108) ->
 * amrex::average_down for one fine box: each coarse cell of crse inside
 * coarsen(fbox) becomes volfrac * (sum of its ratio^dim children, k
 * outermost, i innermost), volfrac = 1 / ratio^dim, components
 * [scomp, scomp+ncomp) of fine into the same components of crse. */
int vd_average_down(vd_ctx* ctx, const vd_array4* fine, const vd_box* fbox, vd_array4* crse, int scomp,
                    int ncomp, const int ratio[3], int dim);

/* average_down_faces of the species fluxes (Source/ScalarSolve.cpp:88-94)
 * for one fine box: each coarse face of direction d inside the coarsened
 * face box of fbox becomes the mean of the ratio^(dim-1) fine faces that
 * tile it (outer loop over the higher tangential direction).  fine[d] and
 * crse[d] are face arrays (1 component). */
int vd_average_down_faces(vd_ctx* ctx, const vd_array4 fine[3], const vd_box* fbox, vd_array4 crse[3],
                          const int ratio[3], int dim);

/* Vidyut::FillPatch at lev > 0 (Source/GridWork.cpp:123-139,
 * FillPatchTwoLevels with cell_cons_interp) and, with nfine = 0,
 * Vidyut::FillCoarsePatch (GridWork.cpp:143-164, InterpFromCoarseLevel).
 * Fills components [scomp, scomp+ncomp) of every cell of dst (a fine
 * box with its ghost cells, fine index space):
 *  - a cell inside the fine domain (periodic directions wrapped) that a
 *    fine valid box fboxes[b] covers takes fine[b]'s value there;
 *  - any other cell inside the domain takes the conservative linear
 *    interpolation of the coarse level crse (the whole coarse domain,
 *    valid data): MC-limited central slopes per direction, scaled by one
 *    factor per coarse cell and component so that no child leaves the
 *    min/max of the 3^dim coarse neighbourhood; coarse neighbours wrap
 *    in periodic directions and are clamped (foextrap) at domain faces;
 *  - a cell outside a non-periodic domain face takes the value of the
 *    nearest cell inside (foextrap, Source/Vidyut.cpp:93-104).
 * The fine boxes must not overlap. */
int vd_fill_patch_two_levels(vd_ctx* ctx, const vd_array4* crse, const vd_geom* cgeom, const vd_array4* fine,
                             const vd_box* fboxes, int nfine, vd_array4* dst, const vd_geom* fgeom,
                             const int ratio[3], int scomp, int ncomp);

/* ------------------------------------------------- kernel timing (bench)
 * When enabled for a kernel family, every launch of that family is
 * bracketed by hipEvents on the context's stream; the totals (time,
 * launches and the canonical algorithmic bytes of those launches) come
 * back with vd_ctx_kernel_timing (synchronises). */
enum {
    VD_KT_NONE = 0,
    VD_KT_GSRB = 1,       /* multigrid red-black colour sweep, finest level */
    VD_KT_TRANSPORT = 2,  /* fused WENO transport rhs */
    VD_KT_CHEM = 3,       /* per-cell chemistry source */
    VD_KT_RESID = 4,      /* multigrid residual, finest level */
    VD_KT_FILL = 5,       /* ghost fill */
    VD_KT_FUSED = 6,      /* fused multi-sweep red-black smoother, finest level */
    VD_KT_MGCOARSE = 7,   /* a V-cycle's legs below the finest level (sweeps,
                             restriction, prolongation, the coarse tail) */
    VD_KT_SPECIES = 8,    /* one whole implicit species solve */
    VD_KT_POTPOST = 9     /* the potential solve's E faces, cell E and E/N */
};
/* Where vd_find_time_scales reads E: the cell E components (EFX ..) of
 * `efield` instead of the state's, until called again with NULL.  No
 * counterpart in the reference (whose state always holds E): for a driver
 * that leaves the newest E in the Sborder of the last step instead of
 * copying it into the state between the steps of one call. */
int vd_ctx_set_cfl_efield(vd_ctx* ctx, const vd_array4* efield);
/* times one family (VD_KT_NONE: none); resets the totals */
int vd_ctx_set_kernel_timing(vd_ctx* ctx, int family);
/* times every family whose bit (1 << family) is set; resets the totals */
int vd_ctx_set_kernel_timing_mask(vd_ctx* ctx, unsigned mask);
/* totals of every timed family */
int vd_ctx_kernel_timing(vd_ctx* ctx, double* total_ms, long long* launches, double* bytes);
/* totals of one family */
int vd_ctx_kernel_timing_family(vd_ctx* ctx, int family, double* total_ms, long long* launches, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* VIDYUT3D_AMD_H */
