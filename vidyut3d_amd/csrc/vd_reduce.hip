// vd_reduce.hip -- CFL time-scale max-reductions (Source/TimeStep.cpp:
// 34-116) and the refinement tags (Source/Tagging.H:8-69).
#include <cfloat>
#include <cstring>

#include "vd_internal.hpp"

namespace vd {

struct CflArgs {
    const DevMech* __restrict__ M;
    Arr<double> phi;
    Arr<double> ef;  // holds E (EFX ..): phi, or the array vd_ctx_set_cfl_efield names
    int dim, const_ele_trans;
    int state_trans;  // the electron mobility or diffusivity is a fit of the state
    double Tg, ele_mob, ele_diff;
    int3 lo;
    int nx, ny, nz;
    unsigned long long* out;  // 3 accumulators (bit patterns of non-negative doubles)
    __device__ void rebase(int64_t o) {
        M = rb(M, o);
        vd::rebase(phi, o);
        vd::rebase(ef, o);
        out = rb(out, o);
    }
};

__device__ __forceinline__ double wave_max(double a) {
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    return a;
}

__device__ __forceinline__ unsigned long long bits(double v) {
    if (v != v) v = __builtin_huge_val();
    return (unsigned long long)__double_as_longlong(v);
}

// STATE: the electron coefficients are fits of the state (a.state_trans)
template <bool STATE>
__global__ void __launch_bounds__(256) k_cfl(CflArgs a, Batch bt) {
    a.rebase(case_off(bt));
    __shared__ double red[3][4];
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    const DevMech& M = *a.M;
    for (; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_of((unsigned)t, a.nx, a.ny, i, j, k);
        int o = a.phi.off(i + a.lo.x, j + a.lo.y, k + a.lo.z);
        const double* P = a.phi.p;
        const int64_t N = a.phi.ns;
        const int ns = M.nspec, e = M.e_idx;
        double etemp = P[VD_ETEMP(ns) * N + o];
        double es = 0.0;
        for (int d = 0; d < a.dim; d++) {
            double v = a.ef(i + a.lo.x, j + a.lo.y, k + a.lo.z, VD_EFX(ns) + d);
            es += v * v;
        }
        double emag = sqrt(es);
        double mu, dc;
        if (a.const_ele_trans) {
            double nd = 0.0;
            for (int sp = 0; sp < ns; sp++) nd += P[(int64_t)sp * N + o];
            dc = a.ele_diff / nd;
            mu = a.ele_mob / nd;
        } else if (STATE) {
            double nd = 0.0;
            for (int sp = 0; sp < ns; sp++) nd += P[(int64_t)sp * N + o];
            dc = spec_diff_st(M, e, etemp, a.Tg, nd, emag);
            mu = spec_mob_st(M, e, etemp, a.Tg, nd, emag);
        } else {
            dc = spec_diff(M, e, etemp, a.Tg);
            mu = spec_mob(M, e);
        }
        double v0 = fabs(fabs(mu) * emag), v1 = fabs(dc), v2 = fabs(fabs(mu) * P[(int64_t)e * N + o]);
        m0 = fmax(m0, v0 != v0 ? __builtin_huge_val() : v0);
        m1 = fmax(m1, v1 != v1 ? __builtin_huge_val() : v1);
        m2 = fmax(m2, v2 != v2 ? __builtin_huge_val() : v2);
    }
    m0 = wave_max(m0);
    m1 = wave_max(m1);
    m2 = wave_max(m2);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        red[0][w] = m0;
        red[1][w] = m1;
        red[2][w] = m2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) v = fmax(v, red[threadIdx.x][q]);
        atomicMax(&a.out[threadIdx.x], bits(v));
    }
}

__global__ void k_zero3(unsigned long long* p, Batch bt) {
    p = rb(p, case_off(bt));
    if (threadIdx.x < 3) p[threadIdx.x] = 0ull;
}

int find_time_scales(vd_ctx* ctx, const vd_mech* m, const vd_array4* phi, const vd_box* box,
                     const vd_geom* g, const vd_params* prm, double out[3]) {
    void* acc;
    VD_TRY(scratch(ctx, "cfl_acc", 64, &acc));
    CflArgs a;
    a.M = m->dev;
    a.phi = arr(phi);
    a.ef = ctx->cfl_ef_set ? arr(&ctx->cfl_ef) : a.phi;
    a.dim = g->dim;
    a.const_ele_trans = prm->const_ele_trans;
    a.state_trans = m->hdev.mfit_kind[m->hdev.e_idx] == VD_TR_FIT || m->hdev.dfit_kind[m->hdev.e_idx] == VD_TR_FIT;
    a.Tg = prm->gas_temperature;
    a.ele_mob = prm->ele_mob;
    a.ele_diff = prm->ele_diff;
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    a.out = (unsigned long long*)acc;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 2048) blocks = 2048;  // grid-stride: 8 blocks per CU
    hipLaunchKernelGGL(k_zero3, bz(ctx, dim3(1)), dim3(64), 0, ctx->stream, a.out, ctx->bt);
    if (a.state_trans)
        hipLaunchKernelGGL(k_cfl<true>, bz(ctx, dim3((unsigned)blocks)), dim3(256), 0, ctx->stream, a, ctx->bt);
    else
        hipLaunchKernelGGL(k_cfl<false>, bz(ctx, dim3((unsigned)blocks)), dim3(256), 0, ctx->stream, a, ctx->bt);
    VD_CHECK_LAUNCH();
    // each case's three maxima (at its own copy of the accumulators)
    for (int z = 0; z < ctx->bn; z++)
        VD_HIP(hipMemcpyAsync(ctx->pinned + 3 * z, (char*)acc + ctx->bt.stride * ctx->bt.idx[z], 3 * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
    VD_HIP(hipStreamSynchronize(ctx->stream));
    for (int z = ctx->bn - 1; z >= 0; z--) {
        double mx[3];
        memcpy(mx, ctx->pinned + 3 * z, sizeof(mx));
        double dt_edrift = DBL_MAX, dt_ediff = DBL_MAX, dt_diel = DBL_MAX;
        if (mx[0] > 0)
            for (int d = 0; d < g->dim; d++) dt_edrift = fmin(dt_edrift, g->dx[d] / mx[0]);
        if (mx[1] > 0)
            for (int d = 0; d < g->dim; d++) dt_ediff = fmin(dt_ediff, 0.5 * g->dx[d] * g->dx[d] / mx[1] / g->dim);
        if (mx[2] > 0) dt_diel = VD_EPS0 / VD_ECHARGE / mx[2];
        // out: the first active case's (every case's in vd_ctx_case_time_scales)
        out[0] = ctx->case_ts[ctx->bt.idx[z]][0] = dt_edrift;
        out[1] = ctx->case_ts[ctx->bt.idx[z]][1] = dt_ediff;
        out[2] = ctx->case_ts[ctx->bt.idx[z]][2] = dt_diel;
    }
    return VD_OK;
}

struct TagArgs {
    Arr<double> st;
    char* tag;
    int tbx, tby, tbz, tjs, tks;
    int ntag, dim;
    const int* comps;
    const double* thr;
    const double* gthr;
    char tagval;
    int3 lo;
    int nx, ny, nz;
};

// state_based_refinement + stategrad_based_refinement
__global__ void __launch_bounds__(256) k_tags(TagArgs a) {
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    char* tp = a.tag + (i - a.tbx) + (j - a.tby) * a.tjs + (k - a.tbz) * a.tks;
    const int o = a.st.off(i, j, k);
    const int sx = 1, sy = a.st.js, sz = a.st.ks;
    for (int c = 0; c < a.ntag; c++) {
        const double* S = a.st.comp(a.comps[c]);
        if (fabs(S[o]) > a.thr[c]) *tp = a.tagval;
    }
    for (int c = 0; c < a.ntag; c++) {
        const double* S = a.st.comp(a.comps[c]);
        double v = S[o];
        double gp = fabs(S[o + sx] - v), gm = fabs(v - S[o - sx]);
        double g0 = gp > gm ? gp : gm;
        gp = fabs(S[o + sy] - v);
        gm = fabs(v - S[o - sy]);
        double g1 = gp > gm ? gp : gm;
        double mx = g0 > g1 ? g0 : g1;
        if (a.dim == 3) {
            gp = fabs(S[o + sz] - v);
            gm = fabs(v - S[o - sz]);
            double g2 = gp > gm ? gp : gm;
            mx = mx > g2 ? mx : g2;
        }
        if (mx > a.gthr[c]) *tp = a.tagval;
    }
}

int error_est(vd_ctx* ctx, const vd_array4* st, vd_tagarray4* tags, const vd_box* box, int ntag,
              const int* comps, const double* thr, const double* gthr, char tagval) {
    if (ntag <= 0) return VD_OK;
    if (ntag > 64) {
        set_error("error_est: at most 64 tagged variables");
        return VD_EINVAL;
    }
    for (int c = 0; c < ntag; c++)
        if (comps[c] < 0 || comps[c] >= st->ncomp) {
            set_error("error_est: component %d not in the state (%d comps)", comps[c], st->ncomp);
            return VD_ENOTFOUND;
        }
    int dim = (st->end[2] - st->begin[2] == 1 && box->lo[2] == box->hi[2]) ? 2 : 3;
    for (int d = 0; d < dim; d++)
        if (st->begin[d] > box->lo[d] - 1 || st->end[d] < box->hi[d] + 2) {
            set_error("error_est: state needs one ghost cell");
            return VD_EINVAL;
        }
    void* tab;
    VD_TRY(scratch(ctx, "tag_tables", 64 * (sizeof(int) + 2 * sizeof(double)), &tab));
    std::vector<char> host(64 * (sizeof(int) + 2 * sizeof(double)));
    memcpy(host.data(), comps, ntag * sizeof(int));
    memcpy(host.data() + 64 * sizeof(int), thr, ntag * sizeof(double));
    memcpy(host.data() + 64 * sizeof(int) + 64 * sizeof(double), gthr, ntag * sizeof(double));
    VD_HIP(hipMemcpyAsync(tab, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    VD_HIP(hipStreamSynchronize(ctx->stream));
    TagArgs a;
    a.st = arr(st);
    a.tag = tags->p;
    a.tbx = tags->begin[0];
    a.tby = tags->begin[1];
    a.tbz = tags->begin[2];
    a.tjs = (int)tags->jstride;
    a.tks = (int)tags->kstride;
    a.ntag = ntag;
    a.dim = dim;
    a.comps = (const int*)tab;
    a.thr = (const double*)((char*)tab + 64 * sizeof(int));
    a.gthr = (const double*)((char*)tab + 64 * sizeof(int) + 64 * sizeof(double));
    a.tagval = tagval;
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL(k_tags, grid1d(tot, 256), dim3(256), 0, ctx->stream, a);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

}  // namespace vd

using namespace vd;

extern "C" int vd_find_time_scales(vd_ctx* ctx, const vd_mech* m, const vd_array4* phi_new,
                                   const vd_box* box, const vd_geom* geom, const vd_params* prm,
                                   double out[3]) {
    if (!ctx || !m || !phi_new || !box || !geom || !prm || !out) return VD_EINVAL;
    VD_TRY(check_extents("find_time_scales", box, {phi_new}));
    if (ctx->cfl_ef_set) VD_TRY(check_extents("find_time_scales (E)", box, {&ctx->cfl_ef}));
    return find_time_scales(ctx, m, phi_new, box, geom, prm, out);
}

extern "C" int vd_error_est(vd_ctx* ctx, const vd_array4* state, vd_tagarray4* tags, const vd_box* box,
                            int ntagvars, const int* comps, const double* refine_phi,
                            const double* refine_phigrad, char tagval) {
    if (!ctx || !state || !tags || !box || (ntagvars > 0 && (!comps || !refine_phi || !refine_phigrad)))
        return VD_EINVAL;
    VD_TRY(check_extents("error_est", box, {state}));
    VD_TRY(check_array_extent("error_est", tags->begin, tags->end, tags->jstride, tags->kstride));
    VD_TRY(require_plain(ctx, "error_est"));
    return error_est(ctx, state, tags, box, ntagvars, comps, refine_phi, refine_phigrad, tagval);
}
