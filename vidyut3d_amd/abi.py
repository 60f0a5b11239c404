"""ctypes mirror of include/vidyut3d_amd.h (structs, constants, helpers).

The structs are the POD types of the C ABI: ``vd_array4`` (the Array4 view
of one box of a MultiFab), ``vd_geom`` (amrex::Geometry of a level),
``vd_params`` (the ``vidyut.*`` keys of Source/Vidyut.H:149-272) and the
case plug-in descriptor ``vd_mech_desc`` (Chemistry.H / UserFunctions.H).
"""
import ctypes as C

import numpy as np

VD_OK, VD_EINVAL, VD_ENOCONV, VD_ENOMEM, VD_EHIP, VD_ENOTFOUND = range(6)
PERBC, DIRCBC, HNEUBC, IHNEUBC, ROBINBC = range(5)
VD_CHEM_TABLE, VD_CHEM_CKWC, VD_CHEM_GEN = 0, 1, 2
VD_RATE_ARRH, VD_RATE_TEPOW, VD_RATE_TEARRH, VD_RATE_JANEV, VD_RATE_FFIT, VD_RATE_JANEV_FFIT = range(6)
VD_DIFF_EINSTEIN_TE, VD_DIFF_EINSTEIN_TG, VD_DIFF_CONST = 0, 1, 2
VD_MAX_SPECIES = 48
VD_MAX_REACTIONS = 512
VD_MAX_REACTANTS = 3
VD_MAX_PRODUCTS = 4
VD_GEN_NG = 8
VD_GEN_NW = 8
NUM_PLASMAVARS = 10
KT_NONE, KT_GSRB, KT_TRANSPORT, KT_CHEM, KT_RESID, KT_FILL, KT_FUSED, KT_MGCOARSE, KT_SPECIES, KT_POTPOST = range(10)
KT_NAMES = {KT_GSRB: "gsrb", KT_TRANSPORT: "transport", KT_CHEM: "chem", KT_RESID: "resid", KT_FILL: "fill",
            KT_FUSED: "fused", KT_MGCOARSE: "mg_coarse", KT_SPECIES: "species_solve", KT_POTPOST: "pot_post"}

# Source/UnivConstants.H:4-19
N_A = 6.02214076e23
K_B = 1.380649e-23
ECHARGE = 1.60217662e-19
EPS0 = 8.854187817e-12
EV = ECHARGE / K_B
ME = 9.10938188e-31
M_AMU = 1.66054e-27

STATUS_NAMES = {0: "VD_OK", 1: "VD_EINVAL", 2: "VD_ENOCONV", 3: "VD_ENOMEM",
                4: "VD_EHIP", 5: "VD_ENOTFOUND"}


def comp(ns):
    """Component indices of Source/VarDefines.H:5-16 for NUM_SPECIES=ns."""
    return dict(EEN=ns, ETEMP=ns + 1, POT=ns + 2, EFX=ns + 3, EFY=ns + 4, EFZ=ns + 5,
                EJH=ns + 6, EIH=ns + 7, EEH=ns + 8, REF=ns + 9, NVAR=ns + 10)


PLASMA_VAR_NAMES = ["Electron_energy", "Electron_Temp", "Potential", "Efieldx", "Efieldy",
                    "Efieldz", "Electron_Jheat", "Electron_inelasticHeat",
                    "Electron_elasticHeat", "ReducedEF"]  # Source/Vidyut.cpp:32-41


class vd_array4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("begin", C.c_int * 3), ("end", C.c_int * 3),
                ("jstride", C.c_int64), ("kstride", C.c_int64), ("nstride", C.c_int64),
                ("ncomp", C.c_int)]


class vd_tagarray4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("begin", C.c_int * 3), ("end", C.c_int * 3),
                ("jstride", C.c_int64), ("kstride", C.c_int64)]


class vd_box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class vd_geom(C.Structure):
    _fields_ = [("dim", C.c_int), ("dom_lo", C.c_int * 3), ("dom_hi", C.c_int * 3),
                ("prob_lo", C.c_double * 3), ("prob_hi", C.c_double * 3),
                ("dx", C.c_double * 3), ("is_periodic", C.c_int * 3)]


class vd_params(C.Structure):
    _fields_ = [
        ("gas_temperature", C.c_double), ("gas_pressure", C.c_double),
        ("hyp_order", C.c_int), ("weno_scheme", C.c_int),
        ("elecenergy_solve", C.c_int), ("do_reactions", C.c_int), ("do_transport", C.c_int),
        ("do_spacechrg", C.c_int), ("do_bg_reactions", C.c_int), ("bound_specden", C.c_int),
        ("min_species_density", C.c_double), ("min_electron_density", C.c_double),
        ("min_electron_temp", C.c_double), ("floor_jh", C.c_int), ("const_ele_trans", C.c_int),
        ("ele_mob", C.c_double), ("ele_diff", C.c_double), ("voltage_profile", C.c_int),
        ("voltage_amp_1", C.c_double), ("voltage_amp_2", C.c_double),
        ("voltage_freq", C.c_double), ("voltage_dur", C.c_double),
        ("voltage_center", C.c_double), ("linsolve_reltol", C.c_double),
        ("linsolve_abstol", C.c_double), ("linsolve_maxiter", C.c_int),
        ("linsolve_max_coarsening_level", C.c_int), ("mg_pre_smooth", C.c_int),
        ("mg_post_smooth", C.c_int), ("mg_precision", C.c_int),
        ("pot_bc_lo", C.c_int * 3), ("pot_bc_hi", C.c_int * 3),
        ("eden_bc_lo", C.c_int * 3), ("eden_bc_hi", C.c_int * 3),
        ("eenrg_bc_lo", C.c_int * 3), ("eenrg_bc_hi", C.c_int * 3),
        ("ion_bc_lo", C.c_int * 3), ("ion_bc_hi", C.c_int * 3),
        ("neutral_bc_lo", C.c_int * 3), ("neutral_bc_hi", C.c_int * 3),
        ("adaptive_dt", C.c_int), ("advective_cfl", C.c_double), ("diffusive_cfl", C.c_double),
        ("dielectric_cfl", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("adaptive_dt_delay", C.c_double), ("dt_stretch", C.c_double),
    ]


class vd_reaction(C.Structure):
    _fields_ = [("A", C.c_double), ("beta", C.c_double), ("Ea", C.c_double),
                ("temp_is_te", C.c_int), ("nreact", C.c_int), ("nprod", C.c_int),
                ("react_sp", C.c_int * VD_MAX_REACTANTS),
                ("react_nu", C.c_double * VD_MAX_REACTANTS),
                ("prod_sp", C.c_int * VD_MAX_PRODUCTS),
                ("prod_nu", C.c_double * VD_MAX_PRODUCTS),
                ("energy", C.c_double), ("rxntype", C.c_int), ("eexci", C.c_double),
                ("elidx", C.c_int),
                ("rate_kind", C.c_int), ("tref", C.c_double), ("tgas_div", C.c_int),
                ("fit", C.c_double * 9), ("A2", C.c_double), ("fit2", C.c_double * 4),
                ("fit_switch_ev", C.c_double), ("kmax", C.c_double), ("third_body", C.c_int),
                ("reversible", C.c_int), ("ng", C.c_int), ("g_sp", C.c_int * VD_GEN_NG),
                ("g_nu", C.c_double * VD_GEN_NG), ("refc_pow", C.c_int), ("nrev", C.c_int),
                ("rev_sp", C.c_int * VD_MAX_PRODUCTS), ("rev_nu", C.c_double * VD_MAX_PRODUCTS),
                ("nwd", C.c_int), ("wd_sp", C.c_int * VD_GEN_NW), ("wd_nu", C.c_double * VD_GEN_NW)]


VD_TR_NCOEF = 12
VD_TR_DEFAULT, VD_TR_FIT = 0, 1


class vd_transport_fit(C.Structure):
    _fields_ = [("kind", C.c_int), ("sign", C.c_double), ("c", C.c_double * VD_TR_NCOEF)]


class vd_mech_desc(C.Structure):
    _fields_ = [("nspec", C.c_int), ("nreac", C.c_int), ("e_idx", C.c_int),
                ("names", (C.c_char * 16) * VD_MAX_SPECIES),
                ("charge", C.c_int * VD_MAX_SPECIES),
                ("molwt", C.c_double * VD_MAX_SPECIES),
                ("is_bg", C.c_int * VD_MAX_SPECIES),
                ("mob", C.c_double * VD_MAX_SPECIES),
                ("diff_kind", C.c_int * VD_MAX_SPECIES),
                ("diff", C.c_double * VD_MAX_SPECIES),
                ("chem_form", C.c_int), ("elastic_loss", C.c_int),
                ("elastic_partner_amu", C.c_double),
                ("reac", vd_reaction * VD_MAX_REACTIONS),
                ("mob_fit", vd_transport_fit * (VD_MAX_SPECIES + 1)),
                ("diff_fit", vd_transport_fit * (VD_MAX_SPECIES + 1)),
                ("vel_set", C.c_int * (VD_MAX_SPECIES + 1)),
                ("vel", (C.c_double * 3) * (VD_MAX_SPECIES + 1)),
                ("robin_set", C.c_int * (VD_MAX_SPECIES + 1)),
                ("robin", (((C.c_double * 3) * 2) * 3) * (VD_MAX_SPECIES + 1)),
                ("user_src", C.c_double * (VD_MAX_SPECIES + 1)),
                ("gibbs_tmid", C.c_double * VD_MAX_SPECIES),
                ("gibbs_lo", (C.c_double * 7) * VD_MAX_SPECIES),
                ("gibbs_hi", (C.c_double * 7) * VD_MAX_SPECIES)]


def set_fit(f, sign, **coef):
    """Fills a vd_transport_fit: sign and the named coefficients c0..c11
    (the rest 0)."""
    f.kind = VD_TR_FIT
    f.sign = float(sign)
    for q in range(VD_TR_NCOEF):
        f.c[q] = float(coef.get("c%d" % q, 0.0))
    return f


class vd_solve_info(C.Structure):
    _fields_ = [("iterations", C.c_int), ("resnorm0", C.c_double), ("resnorm", C.c_double),
                ("target", C.c_double)]


class vd_mlmg_problem(C.Structure):
    _fields_ = [("alpha_a", C.c_double), ("beta", C.c_double),
                ("bface", C.POINTER(vd_array4) * 3),
                ("bc_lo", C.c_int * 3), ("bc_hi", C.c_int * 3),
                ("bcval", C.POINTER(vd_array4)), ("robin_a", C.POINTER(vd_array4)),
                ("robin_b", C.POINTER(vd_array4)), ("robin_f", C.POINTER(vd_array4)),
                ("reltol", C.c_double), ("abstol", C.c_double), ("maxiter", C.c_int),
                ("precision", C.c_int), ("pre_smooth", C.c_int), ("post_smooth", C.c_int),
                ("max_coarsening_level", C.c_int)]


class vd_regrid_params(C.Structure):
    _fields_ = [("ref_ratio", C.c_int * 3), ("blocking_factor", C.c_int * 3),
                ("max_grid_size", C.c_int * 3), ("n_error_buf", C.c_int), ("grid_eff", C.c_double)]


def ivec(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def make_box(lo, hi):
    b = vd_box()
    b.lo[:] = list(lo)
    b.hi[:] = list(hi)
    return b


def make_geom(n_cell, prob_lo, prob_hi, periodic):
    """Geometry of a single-level domain; len(n_cell) is the dimension."""
    dim = len(n_cell)
    g = vd_geom()
    g.dim = dim
    for d in range(3):
        if d < dim:
            g.dom_lo[d] = 0
            g.dom_hi[d] = n_cell[d] - 1
            g.prob_lo[d] = prob_lo[d]
            g.prob_hi[d] = prob_hi[d]
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d]
            g.is_periodic[d] = int(periodic[d])
        else:
            g.dom_lo[d] = g.dom_hi[d] = 0
            g.prob_lo[d], g.prob_hi[d] = 0.0, 1.0
            g.dx[d] = 1.0
            g.is_periodic[d] = 0
    return g


def domain_box(g):
    return make_box([g.dom_lo[d] for d in range(3)], [g.dom_hi[d] for d in range(3)])


def grown_extent(box, ng, dim, face_dir=None):
    """begin/end (exclusive) of a box grown by ng (not in z for 2-D);
    face_dir converts it to the face box of that direction."""
    begin, end = [], []
    for d in range(3):
        g = ng if d < dim else 0
        hi = box.hi[d] + (1 if face_dir == d else 0)
        begin.append(box.lo[d] - g)
        end.append(hi + 1 + g)
    return begin, end


def array4_view(ptr, begin, end, ncomp):
    a = vd_array4()
    a.p = ptr
    a.begin[:] = begin
    a.end[:] = end
    nx, ny, nz = (end[0] - begin[0]), (end[1] - begin[1]), (end[2] - begin[2])
    a.jstride = nx
    a.kstride = nx * ny
    a.nstride = nx * ny * nz
    a.ncomp = ncomp
    return a


def shape_of(begin, end, ncomp):
    """numpy shape (ncomp, nz, ny, nx) matching array4_view."""
    return (ncomp, end[2] - begin[2], end[1] - begin[1], end[0] - begin[0])


def host_array4(arr, begin, end):
    """vd_array4 over a C-contiguous float64 numpy array (ncomp, nz, ny, nx)."""
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    assert arr.shape[1:] == shape_of(begin, end, 1)[1:]
    return array4_view(arr.ctypes.data, begin, end, arr.shape[0])


def tag_view(ptr, begin, end):
    t = vd_tagarray4()
    t.p = ptr
    t.begin[:] = begin
    t.end[:] = end
    t.jstride = end[0] - begin[0]
    t.kstride = (end[0] - begin[0]) * (end[1] - begin[1])
    return t
