// vd_internal.hpp -- shared device/host helpers of the gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/vidyut3d_amd.h"

// Source/UnivConstants.H:4-19
#define VD_N_A 6.02214076e23
#define VD_K_B 1.380649e-23
#define VD_PI 3.14159265358979323846264338327950288
#define VD_ECHARGE 1.60217662e-19
#define VD_EPS0 8.854187817e-12
#define VD_EV (VD_ECHARGE / VD_K_B)
#define VD_ME 9.10938188e-31
#define VD_M_AMU 1.66054e-27
#define VD_FIVEBYTHREE (5.0 / 3.0)
#define VD_TWOTHIRD (2.0 / 3.0)

namespace vd {

void set_error(const char* fmt, ...);

#define VD_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::vd::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return VD_EHIP;                                                               \
        }                                                                                 \
    } while (0)

#define VD_CHECK_LAUNCH()                                                                 \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            ::vd::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return VD_EHIP;                                                               \
        }                                                                                 \
    } while (0)

#define VD_TRY(x)                  \
    do {                           \
        int s_ = (x);              \
        if (s_ != VD_OK) return s_;\
    } while (0)

// Array4 view for kernels: component base pointer + 32-bit cell offsets.
template <class T>
struct Arr {
    T* p;
    int bx, by, bz;
    int js, ks;
    int64_t ns;
    __host__ __device__ inline T* comp(int n) const { return p + (int64_t)n * ns; }
    __host__ __device__ inline int off(int i, int j, int k) const {
        return (i - bx) + (j - by) * js + (k - bz) * ks;
    }
    __host__ __device__ inline T& operator()(int i, int j, int k, int n = 0) const {
        return comp(n)[off(i, j, k)];
    }
};

// Same-shape cases in one launch.  A batched context places case c's copy
// of every device allocation at case 0's address plus c * stride (the
// allocations are carved out of identical per-case regions of one arena),
// so a kernel reaches its case's data by moving each of its pointers by
// one byte offset.  gridDim.z is the number of cases of a launch and idx
// names the case of each z-block; an unbatched launch has gridDim.z = 1
// and stride 0.
struct Batch {
    int64_t stride;
    union {
        unsigned char idx[VD_MAX_BATCH];
        unsigned idxw[VD_MAX_BATCH / 4];  // the same bytes as four words
    };
};
static_assert(VD_MAX_BATCH % 4 == 0, "Batch::idxw covers idx");
// doubles of the host-mapped norm read-back buffer (values of up to
// VD_MAX_BATCH cases x 5 slot groups, then a sequence number in the last)
constexpr int HNORM_DOUBLES = 256;
// the byte offset of this z-block's case.  The index is taken from its
// aligned word: a byte load from the kernel arguments at a block-uniform
// position has no scalar form and would be a vector load that every wave
// waits for before its first data load
__device__ __forceinline__ int64_t case_off(const Batch& b) {
    const unsigned z = blockIdx.z;
    return b.stride * (int64_t)((b.idxw[z >> 2] >> ((z & 3u) * 8u)) & 0xffu);
}
// p moved by o bytes (a null pointer stays null: it marks absent data)
template <class T>
__host__ __device__ __forceinline__ T* rb(T* p, int64_t o) {
    return p ? (T*)((char*)p + o) : p;
}
template <class T>
__host__ __device__ __forceinline__ void rebase(Arr<T>& a, int64_t o) {
    a.p = rb(a.p, o);
}
// element (i, j, k, n) of a's copy o bytes on, for kernels that keep their
// argument structs unmodified (a struct modified in a kernel and then
// selected or indexed at run time goes to private memory)
template <class T>
__device__ __forceinline__ T& at(const Arr<T>& a, int64_t o, int i, int j, int k, int n = 0) {
    return rb(a.p, o)[(int64_t)n * a.ns + a.off(i, j, k)];
}

inline Arr<double> arr(const vd_array4* a) {
    Arr<double> r;
    r.p = a->p;
    r.bx = a->begin[0];
    r.by = a->begin[1];
    r.bz = a->begin[2];
    r.js = (int)a->jstride;
    r.ks = (int)a->kstride;
    r.ns = a->nstride;
    return r;
}

// Kernels index a box, and every array with its ghost cells, with 32-bit
// offsets: the entry points reject boxes and arrays of 2^31 cells or more
// (checked first, before the context is used).
constexpr int64_t MAX_CELLS_32 = ((int64_t)1 << 31) - 1;

inline int64_t span_cells(const int* begin, const int* end) {
    int64_t c = 1;
    for (int d = 0; d < 3; d++) c *= (int64_t)end[d] - begin[d] > 0 ? (int64_t)end[d] - begin[d] : 0;
    return c;
}

inline int extent_error(const char* what, const char* kind, int64_t cells) {
    set_error("%s: %s of %lld cells: at most 2^31 - 1 (32-bit cell offsets)", what, kind, (long long)cells);
    return VD_EINVAL;
}

inline int check_box_extent(const char* what, const vd_box* box) {
    if (!box) return VD_OK;
    const int end[3] = {box->hi[0] + 1, box->hi[1] + 1, box->hi[2] + 1};
    const int64_t c = span_cells(box->lo, end);
    return c > MAX_CELLS_32 ? extent_error(what, "box", c) : VD_OK;
}

inline int check_array_extent(const char* what, const int* begin, const int* end, int64_t jstride,
                              int64_t kstride) {
    const int64_t c = span_cells(begin, end);
    if (c > MAX_CELLS_32 || jstride > MAX_CELLS_32 || kstride > MAX_CELLS_32)
        return extent_error(what, "array", c);
    return VD_OK;
}

inline int check_extents(const char* what, const vd_box* box, std::initializer_list<const vd_array4*> arrs) {
    VD_TRY(check_box_extent(what, box));
    for (const vd_array4* a : arrs)
        if (a) VD_TRY(check_array_extent(what, a->begin, a->end, a->jstride, a->kstride));
    return VD_OK;
}

// the first `dim` face arrays of a set of three (unused entries of a 2-D
// set are not read)
inline int check_face_extents(const char* what, const vd_array4* f, int dim) {
    if (!f) return VD_OK;
    for (int d = 0; d < dim && d < 3; d++) VD_TRY(check_array_extent(what, f[d].begin, f[d].end, f[d].jstride, f[d].kstride));
    return VD_OK;
}

// Device copy of the case plug-in descriptor (vd_mech_desc), reaction
// table in structure-of-arrays form for uniform (scalar) loads.
struct DevMech {
    int nspec, nreac, e_idx, chem_form, elastic_loss;
    double elastic_partner_amu;
    int charge[VD_MAX_SPECIES];
    double molwt[VD_MAX_SPECIES];
    double mob[VD_MAX_SPECIES];
    int diff_kind[VD_MAX_SPECIES];
    double diff[VD_MAX_SPECIES];
    // reactions
    double A[VD_MAX_REACTIONS], beta[VD_MAX_REACTIONS], Ea[VD_MAX_REACTIONS];
    double energy[VD_MAX_REACTIONS], eexci[VD_MAX_REACTIONS];
    int te[VD_MAX_REACTIONS], rxntype[VD_MAX_REACTIONS], elidx[VD_MAX_REACTIONS];
    int nreact[VD_MAX_REACTIONS], nprod[VD_MAX_REACTIONS];
    int rsp[VD_MAX_REACTANTS][VD_MAX_REACTIONS];
    double rnu[VD_MAX_REACTANTS][VD_MAX_REACTIONS];
    int psp[VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    double pnu[VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    // CKWC form: net stoichiometry per reaction as (species, net) lists
    int nnet[VD_MAX_REACTIONS];
    int netsp[VD_MAX_REACTANTS + VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    double netnu[VD_MAX_REACTANTS + VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    int order[VD_MAX_REACTIONS];
    // state-dependent specMob / specDiff (vd_transport_fit), index nspec =
    // the electron energy
    int mfit_kind[VD_MAX_SPECIES + 1], dfit_kind[VD_MAX_SPECIES + 1];
    double mfit_sign[VD_MAX_SPECIES + 1], dfit_sign[VD_MAX_SPECIES + 1];
    double mfit_c[VD_MAX_SPECIES + 1][VD_TR_NCOEF], dfit_c[VD_MAX_SPECIES + 1][VD_TR_NCOEF];
    // user_transport::compute_vel: prescribed drift velocities
    int vel_set[VD_MAX_SPECIES + 1];
    double vel[VD_MAX_SPECIES + 1][3];
    // user_sources::add_user_react_sources constants
    double user_src[VD_MAX_SPECIES + 1];
    int any_user_src;
    // VD_CHEM_GEN (a generated productionRate read as data)
    int rkind[VD_MAX_REACTIONS], tgdiv[VD_MAX_REACTIONS], tb[VD_MAX_REACTIONS], rev[VD_MAX_REACTIONS];
    double tref[VD_MAX_REACTIONS], A2[VD_MAX_REACTIONS], fsw[VD_MAX_REACTIONS], kmax[VD_MAX_REACTIONS];
    double fit[9][VD_MAX_REACTIONS], fit2[4][VD_MAX_REACTIONS];
    int ng[VD_MAX_REACTIONS], gsp[VD_GEN_NG][VD_MAX_REACTIONS];
    double gnu[VD_GEN_NG][VD_MAX_REACTIONS];
    int refcp[VD_MAX_REACTIONS], nrev[VD_MAX_REACTIONS], revsp[VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    double revnu[VD_MAX_PRODUCTS][VD_MAX_REACTIONS];
    int nwd[VD_MAX_REACTIONS], wdsp[VD_GEN_NW][VD_MAX_REACTIONS];
    double wdnu[VD_GEN_NW][VD_MAX_REACTIONS];
    double gtmid[VD_MAX_SPECIES], glo[VD_MAX_SPECIES][7], ghi[VD_MAX_SPECIES][7];
    int any_rev, any_tb, any_fit;
};

// VD_CHEM_GEN: the per-launch table of what depends on the gas temperature
// only (uniform over a launch): k_f of every VD_RATE_ARRH reaction, then
// g/RT of every species.  Computed once per launch by k_gen_tconst with the
// generated file's own arithmetic, so every cell sees the values its own
// evaluation would give.
constexpr int GEN_TCONST_DOUBLES = VD_MAX_REACTIONS + VD_MAX_SPECIES;

// g/RT of species k (gibbs() of a generated Chemistry.H): the terms summed
// left to right, the low range below gtmid
__host__ __device__ inline double gen_gibbs(const DevMech& m, int k, double T) {
#pragma clang fp contract(off)
    const double T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;
    const double invT = 1.0 / T, logT = log(T);
    const double* c = (m.gtmid[k] > 0.0 && !(T < m.gtmid[k])) ? m.ghi[k] : m.glo[k];
    return c[0] * invT + c[1] + c[2] * logT + c[3] * T + c[4] * T2 + c[5] * T3 + c[6] * T4;
}

// specMob / specDiff of the case plug-in (constant mobilities, Einstein
// diffusion; index nspec = electron energy with 5/3 of the electron values)
__host__ __device__ inline double spec_mob(const DevMech& m, int s) {
    return s == m.nspec ? VD_FIVEBYTHREE * m.mob[m.e_idx] : m.mob[s];
}
__host__ __device__ inline double spec_diff(const DevMech& m, int s, double Te, double Tg) {
    if (s == m.nspec) {
        double te_ev = Te / VD_EV;
        return VD_FIVEBYTHREE * te_ev * fabs(m.mob[m.e_idx]);
    }
    int k = m.diff_kind[s];
    if (k == VD_DIFF_EINSTEIN_TE) return (Te / VD_EV) * fabs(m.mob[s]);
    if (k == VD_DIFF_EINSTEIN_TG) return (Tg / VD_EV) * fabs(m.mob[s]);
    return m.diff[s];
}

// a vd_transport_fit coefficient: sign * exp(c0 + c1 ln|E| + c2 ln Te +
// c3 ln Tg + c4 ln N + c5 ln e + c6/e + c7/e^2 + c8/e^3 + c9 ln(1 + c10 x^2
// + c11 x^4)), e = 1.5 Te / 1 eV, x = |E| / N; zero terms skipped.  The
// same operations in the same order as the oracle, without contraction.
__host__ __device__ inline double tr_eval(const double* c, double sign, double Te, double Tg, double nd,
                                          double emag) {
#pragma clang fp contract(off)
    double l = c[0];
    if (c[1] != 0.0) l += c[1] * log(emag);
    if (c[2] != 0.0) l += c[2] * log(Te);
    if (c[3] != 0.0) l += c[3] * log(Tg);
    if (c[4] != 0.0) l += c[4] * log(nd);
    if (c[5] != 0.0 || c[6] != 0.0 || c[7] != 0.0 || c[8] != 0.0) {
        const double e = 1.5 * Te / VD_EV;
        if (c[5] != 0.0) l += c[5] * log(e);
        if (c[6] != 0.0) l += c[6] / e;
        if (c[7] != 0.0) l += c[7] / (e * e);
        if (c[8] != 0.0) l += c[8] / (e * e * e);
    }
    if (c[9] != 0.0) {
        const double x = emag / nd;
        l += c[9] * log(1.0 + c[10] * (x * x) + c[11] * (x * x * x * x));
    }
    return sign * exp(l);
}

// specMob / specDiff(specid, Te, ndens, |E|, Tg) of the case plug-in
__host__ __device__ inline double spec_mob_st(const DevMech& m, int s, double Te, double Tg, double nd,
                                              double emag) {
#pragma clang fp contract(off)
    if (m.mfit_kind[s] == VD_TR_FIT) return tr_eval(m.mfit_c[s], m.mfit_sign[s], Te, Tg, nd, emag);
    if (s == m.nspec && m.mfit_kind[m.e_idx] == VD_TR_FIT)
        return VD_FIVEBYTHREE * tr_eval(m.mfit_c[m.e_idx], m.mfit_sign[m.e_idx], Te, Tg, nd, emag);
    return spec_mob(m, s);
}
__host__ __device__ inline double spec_diff_st(const DevMech& m, int s, double Te, double Tg, double nd,
                                               double emag) {
#pragma clang fp contract(off)
    if (m.dfit_kind[s] == VD_TR_FIT) return tr_eval(m.dfit_c[s], m.dfit_sign[s], Te, Tg, nd, emag);
    const int e = m.e_idx;
    if (s == m.nspec) {
        if (m.dfit_kind[e] == VD_TR_FIT) return VD_FIVEBYTHREE * tr_eval(m.dfit_c[e], m.dfit_sign[e], Te, Tg, nd, emag);
        if (m.mfit_kind[e] == VD_TR_FIT)
            return VD_FIVEBYTHREE * (Te / VD_EV) * fabs(tr_eval(m.mfit_c[e], m.mfit_sign[e], Te, Tg, nd, emag));
        return spec_diff(m, s, Te, Tg);
    }
    if (m.mfit_kind[s] == VD_TR_FIT && m.diff_kind[s] != VD_DIFF_CONST) {
        const double mu = fabs(tr_eval(m.mfit_c[s], m.mfit_sign[s], Te, Tg, nd, emag));
        return (m.diff_kind[s] == VD_DIFF_EINSTEIN_TE ? Te / VD_EV : Tg / VD_EV) * mu;
    }
    return spec_diff(m, s, Te, Tg);
}

// whether species s's drift velocity needs the cell state (a fitted
// mobility) or is prescribed; else it is a constant mobility times E
inline bool drift_is_state(const DevMech& m, int s) {
    return m.vel_set[s] || m.mfit_kind[s] == VD_TR_FIT || (s == m.nspec && m.mfit_kind[m.e_idx] == VD_TR_FIT);
}
// whether species s's diffusion coefficient is the same in every cell
inline bool diff_is_uniform(const DevMech& m, int s) {
    return s < m.nspec && m.dfit_kind[s] != VD_TR_FIT && m.diff_kind[s] != VD_DIFF_EINSTEIN_TE &&
           !(m.diff_kind[s] == VD_DIFF_EINSTEIN_TG && m.mfit_kind[s] == VD_TR_FIT);
}

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace vd

struct vd_mech {
    vd_ctx* ctx;
    vd_mech_desc host;
    vd::DevMech hdev;     // host copy of the device struct
    vd::DevMech* dev;     // device memory
    // mechanism-specialised source kernel (table form), loaded on ctx's device
    void* mod = nullptr;  // hipModule_t
    void* fn = nullptr;   // hipFunction_t
    int use_spec = 0;
};

namespace vd {
int chem_jit_load(vd_mech* m);
void chem_jit_unload(vd_mech* m);
int chem_jit_launch(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4* rxn, const vd_box* box,
                    double Tg, const double* tc);
}

namespace vd {
struct MGWork;  // multigrid workspace (vd_mg.hip)
void mgwork_free(MGWork* w);
}

struct vd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // named grow-only scratch buffers (device memory), reused across calls
    std::map<std::string, vd::Buf> scratch;
    // the box list whose owner map is in scratch "level_owner"
    std::vector<int> owner_key;
    // multigrid workspaces keyed by shape + precision
    std::map<std::string, vd::MGWork*> mg;
    // small pinned host buffer for scalar read-backs
    double* pinned = nullptr;
    // host-mapped coherent buffer the multigrid norm kernels write into
    // (values, then a sequence number in the last slot the host waits on)
    double* hnorm = nullptr;
    double* hnorm_d = nullptr;  // its device address
    unsigned long long norm_seq = 0;
    // vd_ctx_set_cfl_efield: the array the CFL reduction reads E from
    vd_array4 cfl_ef{};
    bool cfl_ef_set = false;
    // kernel timing: families selected by bit, one event pair per launch
    // (families may nest: a species solve contains sweep launches)
    struct KtEvent {
        hipEvent_t a, b;
        int family;
    };
    unsigned kt_mask = 0;
    std::vector<KtEvent> kt_events;
    std::vector<hipEvent_t> kt_pool;
    double kt_bytes[32] = {0.0};
    long kt_open[32] = {0};  // index of the family's open pair + 1 (0: none)
    size_t kt_next = 0;
    // batching (vd_ctx_create_batched): ncase regions of arena_case bytes;
    // the context's allocations are carved out of region 0 (bump pointer,
    // released with the context).  bt / bn: the cases the next launches
    // run on (bn z-blocks); act / nact: the active cases of the caller
    int ncase = 1;
    char* arena = nullptr;
    size_t arena_case = 0, arena_used = 0;
    vd::Batch bt{};
    int bn = 1;
    int act[VD_MAX_BATCH] = {0};
    int nact = 1;
    // per-case overrides and results
    int case_volt_set[VD_MAX_BATCH] = {0};
    double case_volt[VD_MAX_BATCH][2] = {{0.0}};
    vd_solve_info case_info[VD_MAX_BATCH] = {};
    double case_ts[VD_MAX_BATCH][3] = {{0.0}};
};

namespace vd {

// the grid of a launch on the context's current cases (one z-block each)
inline dim3 bz(const vd_ctx* c, dim3 g) {
    g.z = (unsigned)c->bn;
    return g;
}
// device memory of the context: hipMalloc, or in a batched context the
// next piece of case 0's arena region (freed with the context)
int dev_alloc(vd_ctx* ctx, size_t bytes, void** out);
void dev_free(vd_ctx* ctx, void* p);
// the launch batch narrowed to `n` of the active cases for a scope
struct BatchScope {
    vd_ctx* c;
    Batch bt;
    int bn;
    BatchScope(vd_ctx* ctx, const int* cases, int n);
    ~BatchScope();
};
// entry points outside the batched step path: VD_EINVAL in a batched
// context (their pointers need not lie in the per-case regions)
int require_plain(vd_ctx* ctx, const char* what);

// returns a device buffer of at least `bytes` registered under `name`
int scratch(vd_ctx* ctx, const std::string& name, size_t bytes, void** out);

// takes the device address of the no-allocation entry points' fixed
// scratch array on `device` (current), once per device
int noalloc_setup(int device);

// resident workgroups of `kernel` (block threads) on the whole device
int resident_blocks(vd_ctx* ctx, const void* kernel, int block);
// chunk count in [nmin, nmax] that best fills whole rounds of `cap`
// resident workgroups with per_chunk workgroups per chunk
int round_filling_chunks(int64_t per_chunk, int cap, int nmin, int nmax);

// kernel-timing brackets (no-ops unless the family is selected)
void kt_begin(vd_ctx* ctx, int family);
void kt_end(vd_ctx* ctx, int family, double bytes);

inline dim3 grid1d(int64_t n, int block) { return dim3((unsigned)((n + block - 1) / block)); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): hardware block b runs on XCD b % 8.  Stencil kernels read the rows
// next to their own, so give every XCD a contiguous range of logical
// blocks: neighbouring rows then share an L2 instead of being fetched from
// HBM by up to three XCDs.
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned q = nb / NUM_XCD, r = nb % NUM_XCD;
    const unsigned x = b % NUM_XCD, idx = b / NUM_XCD;
    return x * q + (x < r ? x : r) + idx;
}
// logical (remapped) block index of a 1-D grid
__device__ __forceinline__ unsigned lblock() { return xcd_remap(blockIdx.x, gridDim.x); }
// cell (i, j, k), counted from 0, of linear index t in a box n0 cells wide
// and n1 deep.  Boxes and arrays hold fewer than 2^31 cells (every entry
// point checks), so the division is 32-bit unsigned, not the much longer
// 64-bit sequence
__device__ __forceinline__ void cell_of(unsigned t, int n0, int n1, int& i, int& j, int& k) {
    const unsigned q = t / (unsigned)n0;
    i = (int)(t - q * (unsigned)n0);
    j = (int)(q % (unsigned)n1);
    k = (int)(q / (unsigned)n1);
}

// a/b from the hardware reciprocal with two Newton steps and a final
// residual correction: within an ulp of the rounded quotient for normal,
// finite operands, at a quarter of the instructions of the IEEE division
// sequence (used where the divisor is a positive sum of coefficients)
__device__ __forceinline__ double fdiv(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ float fdiv(float a, float b) { return a / b; }

}  // namespace vd
