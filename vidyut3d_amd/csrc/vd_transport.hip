// vd_transport.hip -- drift-diffusion face fluxes (WENO-5 JS/M/Z, WAF
// MINBEE, first-order upwind) and the flux divergence.
//
// Restates compute_flux (Source/compute_explicit_flux.H:200-326) with the
// constant-mobility compute_vel of Source/PlasmaChem.H:186-206.
#include "vd_internal.hpp"

namespace vd {

int fill_value(vd_ctx* ctx, double* p, int64_t n, double v);

#define VD_HEPS 1.0e-30

__device__ __forceinline__ double sq(double x) { return x * x; }

// gk (compute_explicit_flux.H:16-23)
__device__ __forceinline__ double gk(double wk, double g) {
    return fdiv(wk * (g + g * g - 3.0 * g * wk + wk * wk), g * g + wk * (1 - 2.0 * g));
}

// weno (compute_explicit_flux.H:25-112); pow(x,2.0) evaluated as x*x, and
// the normalised weights applied as (a1 p1 + a2 p2 + a3 p3) / (a1 + a2 + a3)
template <int S>
__device__ __forceinline__ double weno5(double im2, double im1, double i0, double ip1,
                                        double ip2) {
    const double eps = 1e-20;
    double p1 = (1.0 / 3.0) * im2 - (7.0 / 6.0) * im1 + (11.0 / 6.0) * i0;
    double p2 = -(1.0 / 6.0) * im1 + (5.0 / 6.0) * i0 + (1.0 / 3.0) * ip1;
    double p3 = (1.0 / 3.0) * i0 + (5.0 / 6.0) * ip1 - (1.0 / 6.0) * ip2;
    double b1 = (13.0 / 12.0) * sq(im2 - 2.0 * im1 + i0) + (1.0 / 4.0) * sq(im2 - 4.0 * im1 + 3.0 * i0);
    double b2 = (13.0 / 12.0) * sq(im1 - 2.0 * i0 + ip1) + (1.0 / 4.0) * sq(im1 - ip1);
    double b3 = (13.0 / 12.0) * sq(i0 - 2.0 * ip1 + ip2) + (1.0 / 4.0) * sq(3.0 * i0 - 4.0 * ip1 + ip2);
    const double g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0;
    double a1, a2, a3;
    if (S == 1) {
        a1 = fdiv(g1, sq(b1 + eps));
        a2 = fdiv(g2, sq(b2 + eps));
        a3 = fdiv(g3, sq(b3 + eps));
    } else if (S == 2) {
        a1 = fdiv(g1, sq(b1 + eps));
        a2 = fdiv(g2, sq(b2 + eps));
        a3 = fdiv(g3, sq(b3 + eps));
        const double is = fdiv(1.0, a1 + a2 + a3);
        a1 = gk(a1 * is, g1);
        a2 = gk(a2 * is, g2);
        a3 = gk(a3 * is, g3);
    } else {
        double tau = fabs(b1 - b3);
        a1 = g1 * (1.0 + sq(fdiv(tau, b1 + eps)));
        a2 = g2 * (1.0 + sq(fdiv(tau, b2 + eps)));
        a3 = g3 * (1.0 + sq(fdiv(tau, b3 + eps)));
    }
    return fdiv(a1 * p1 + a2 * p2 + a3 * p3, a1 + a2 + a3);
}

__device__ __forceinline__ double slope_lim(double c, double dl, double dm, double dr) {
    double r = (c >= 0.0) ? dl / (dm + VD_HEPS) : dr / (dm + VD_HEPS);
    double lim;
    if (r <= 0.0)
        lim = 1.0;
    else if (r > 0.0 && r <= 1.0)
        lim = 1.0 - (1.0 - fabs(c)) * r;
    else
        lim = fabs(c);
    return lim;
}

struct FluxArgs {
    Arr<double> sb;
    Arr<double> flux;
    int dir, specid, efcomp;
    double mob;          // spec_mob(specid)
    int lo_zero, hi_zero; // face index that is zeroed (IHNEU) or INT_MIN
    int3 flo;            // first face index
    int fnx, fny, fnz;   // face box extent
    double dxmin, dt;
};

// one thread per face of direction a.dir
template <int HYP, int S>
__global__ void __launch_bounds__(256) k_face_flux(FluxArgs a) {
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.fnx * a.fny * a.fnz;
    if (t >= tot) return;
    int i = (int)(t % a.fnx);
    int64_t r = t / a.fnx;
    int j = (int)(r % a.fny);
    int k = (int)(r / a.fny);
    i += a.flo.x;
    j += a.flo.y;
    k += a.flo.z;
    const int d = a.dir;
    const int st = d == 0 ? 1 : (d == 1 ? a.sb.js : a.sb.ks);
    const double* __restrict__ u = a.sb.comp(a.specid);
    const double* __restrict__ E = a.sb.comp(a.efcomp + d);
    int o = a.sb.off(i, j, k);  // right cell
    const int face_d = d == 0 ? i : (d == 1 ? j : k);
    double f;
    double uL = u[o - st], uR = u[o];
    double vL = a.mob * E[o - st], vR = a.mob * E[o];
    if (HYP == 1) {
        double fL = uL * vL, fR = uR * vR;
        double vm = 0.5 * (vL + vR);
        double sgn = copysign(1.0, vm);
        f = fL * 0.5 * (1.0 + sgn) + fR * 0.5 * (1.0 - sgn);
    } else if (HYP == 2) {
        double uLm1 = u[o - 2 * st], uRp1 = u[o + st];
        double dl = uL - uLm1, dm = uR - uL, dr = uRp1 - uR;
        double vm = 0.5 * (vL + vR);
        double sgn = copysign(1.0, vm);
        double c = vm * a.dt / a.dxmin;
        double fL = uL * vL, fR = uR * vR;
        double lim = slope_lim(c, dl, dm, dr);
        f = 0.5 * (fL + fR) - 0.5 * sgn * lim * (fR - fL);
    } else {
        double ulm2 = u[o - 3 * st], ulm1 = u[o - 2 * st], urp1 = u[o + st], urp2 = u[o + 2 * st];
        double vlm2 = a.mob * E[o - 3 * st], vlm1 = a.mob * E[o - 2 * st];
        double vrp1 = a.mob * E[o + st], vrp2 = a.mob * E[o + 2 * st];
        double flm2 = ulm2 * vlm2, flm1 = ulm1 * vlm1, fl = uL * vL, fr = uR * vR;
        double frp1 = urp1 * vrp1, frp2 = urp2 * vrp2;
        double lam = fabs(0.5 * (vL + vR));
        double fph = weno5<S>(flm2, flm1, fl, fr, frp1);
        double fmh = weno5<S>(frp2, frp1, fr, fl, flm1);
        double uph = weno5<S>(ulm2, ulm1, uL, uR, urp1);
        double umh = weno5<S>(urp2, urp1, uR, uL, ulm1);
        f = 0.5 * (fph + lam * uph) + 0.5 * (fmh - lam * umh);
    }
    if (face_d == a.lo_zero || face_d == a.hi_zero) f = 0.0;
    a.flux.p[a.flux.off(i, j, k)] = f;
}

struct DsdtArgs {
    Arr<double> f0, f1, f2, rxn, dsdt;
    int specid, dim;
    double dx0, dx1, dx2;
    int3 lo;
    int nx, ny, nz;
};

// compute_dsdt (Source/ScalarSolve.cpp:44-53)
__global__ void __launch_bounds__(256) k_dsdt(DsdtArgs a) {
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i = (int)(t % a.nx);
    int64_t r = t / a.nx;
    int j = (int)(r % a.ny);
    int k = (int)(r / a.ny);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    double v = (a.f0(i, j, k) - a.f0(i + 1, j, k)) / a.dx0 + a.rxn(i, j, k, a.specid);
    if (a.dim > 1) v += (a.f1(i, j, k) - a.f1(i, j + 1, k)) / a.dx1;
    if (a.dim == 3) v += (a.f2(i, j, k) - a.f2(i, j, k + 1)) / a.dx2;
    a.dsdt(i, j, k) = v;
}

template <int HYP, int S>
static void launch_flux(vd_ctx* ctx, const FluxArgs& a) {
    int64_t tot = (int64_t)a.fnx * a.fny * a.fnz;
    hipLaunchKernelGGL((k_face_flux<HYP, S>), grid1d(tot, 256), dim3(256), 0, ctx->stream, a);
}

int check_transport_args(const vd_params* prm, int specid, int nspec) {
    if (prm->hyp_order != 1 && prm->hyp_order != 2 && prm->hyp_order != 5) {
        set_error("hyp_order %d not implemented (Source/compute_explicit_flux.H:315)", prm->hyp_order);
        return VD_EINVAL;
    }
    if (prm->hyp_order == 5 && (prm->weno_scheme < 1 || prm->weno_scheme > 3)) {
        set_error("Unknown WENO scheme %d (Source/compute_explicit_flux.H:100)", prm->weno_scheme);
        return VD_EINVAL;
    }
    if (specid < 0 || specid > nspec) {
        set_error("species id %d out of range", specid);
        return VD_EINVAL;
    }
    return VD_OK;
}

int transport_flux(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4 flux[3],
                   const vd_box* box, const vd_geom* g, const int bc_lo[3], const int bc_hi[3],
                   const vd_params* prm, double dt, int specid) {
    VD_TRY(check_transport_args(prm, specid, m->hdev.nspec));
    int ng = prm->hyp_order == 5 ? 3 : prm->hyp_order;
    for (int d = 0; d < g->dim; d++)
        if (sb->begin[d] > box->lo[d] - ng || sb->end[d] < box->hi[d] + 1 + ng) {
            set_error("transport: Sborder needs %d ghost cells", ng);
            return VD_EINVAL;
        }
    double dxmin = g->dx[0];
    for (int d = 1; d < g->dim; d++) dxmin = fmin(dxmin, g->dx[d]);
    for (int d = 0; d < g->dim; d++) {
        FluxArgs a;
        a.sb = arr(sb);
        a.flux = arr(&flux[d]);
        a.dir = d;
        a.specid = specid;
        a.efcomp = VD_EFX(m->hdev.nspec);
        a.mob = spec_mob(m->hdev, specid);
        a.lo_zero = bc_lo[d] == VD_IHNEUBC ? g->dom_lo[d] : INT32_MIN;
        a.hi_zero = bc_hi[d] == VD_IHNEUBC ? g->dom_hi[d] + 1 : INT32_MIN;
        a.flo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
        a.fnx = box->hi[0] - box->lo[0] + 1 + (d == 0);
        a.fny = box->hi[1] - box->lo[1] + 1 + (d == 1);
        a.fnz = box->hi[2] - box->lo[2] + 1 + (d == 2);
        a.dxmin = dxmin;
        a.dt = dt;
        switch (prm->hyp_order) {
            case 1: launch_flux<1, 0>(ctx, a); break;
            case 2: launch_flux<2, 0>(ctx, a); break;
            default:
                if (prm->weno_scheme == 1) launch_flux<5, 1>(ctx, a);
                else if (prm->weno_scheme == 2) launch_flux<5, 2>(ctx, a);
                else launch_flux<5, 3>(ctx, a);
        }
        VD_CHECK_LAUNCH();
    }
    return VD_OK;
}

int dsdt(vd_ctx* ctx, const vd_array4 flux[3], const vd_array4* rxn, vd_array4* out,
         const vd_box* box, const vd_geom* g, int specid) {
    DsdtArgs a;
    a.f0 = arr(&flux[0]);
    a.f1 = g->dim > 1 ? arr(&flux[1]) : a.f0;
    a.f2 = g->dim > 2 ? arr(&flux[2]) : a.f0;
    a.rxn = arr(rxn);
    a.dsdt = arr(out);
    a.specid = specid;
    a.dim = g->dim;
    a.dx0 = g->dx[0];
    a.dx1 = g->dx[1];
    a.dx2 = g->dx[2];
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL(k_dsdt, grid1d(tot, 256), dim3(256), 0, ctx->stream, a);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

static vd_array4 face_view(double* p, const vd_box* box, int d) {
    vd_array4 a;
    a.p = p;
    for (int q = 0; q < 3; q++) {
        a.begin[q] = box->lo[q];
        a.end[q] = box->hi[q] + 1 + (q == d);
    }
    a.jstride = a.end[0] - a.begin[0];
    a.kstride = a.jstride * (a.end[1] - a.begin[1]);
    a.nstride = a.kstride * (a.end[2] - a.begin[2]);
    a.ncomp = 1;
    return a;
}

// update_explsrc_at_all_levels for one level (Source/ScalarSolve.cpp:57-111):
// fluxes zeroed, computed if do_transport, then the divergence.
int transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, const vd_array4* rxn,
                  vd_array4* out, const vd_box* box, const vd_geom* g, const int bc_lo[3],
                  const int bc_hi[3], const vd_params* prm, double dt, int specid) {
    vd_array4 fl[3];
    for (int d = 0; d < 3; d++) {
        int dd = d < g->dim ? d : 0;
        int64_t n = (int64_t)(box->hi[0] - box->lo[0] + 1 + (dd == 0)) *
                    (box->hi[1] - box->lo[1] + 1 + (dd == 1)) *
                    (box->hi[2] - box->lo[2] + 1 + (dd == 2));
        void* p;
        char name[32];
        snprintf(name, sizeof(name), "transport_flux%d", d);
        VD_TRY(scratch(ctx, name, n * sizeof(double), &p));
        fl[d] = face_view((double*)p, box, dd);
        if (!prm->do_transport && d < g->dim) VD_TRY(fill_value(ctx, (double*)p, n, 0.0));
    }
    double cells = (double)(box->hi[0] - box->lo[0] + 1) * (box->hi[1] - box->lo[1] + 1) *
                   (box->hi[2] - box->lo[2] + 1);
    kt_begin(ctx, VD_KT_TRANSPORT);
    if (prm->do_transport) VD_TRY(transport_flux(ctx, m, sb, fl, box, g, bc_lo, bc_hi, prm, dt, specid));
    int st = dsdt(ctx, fl, rxn, out, box, g, specid);
    kt_end(ctx, VD_KT_TRANSPORT, 8.0 * cells * (m->hdev.nspec + 6));
    return st;
}

}  // namespace vd

using namespace vd;

extern "C" int vd_compute_scalar_transport_flux(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                                vd_array4 flux[3], const vd_box* box,
                                                const vd_geom* geom, const int bc_lo[3],
                                                const int bc_hi[3], const vd_params* prm,
                                                double time, double dt, int specid) {
    (void)time;
    if (!ctx || !m || !sborder || !flux || !box || !geom || !prm) return VD_EINVAL;
    return transport_flux(ctx, m, sborder, flux, box, geom, bc_lo, bc_hi, prm, dt, specid);
}

extern "C" int vd_compute_dsdt(vd_ctx* ctx, const vd_array4 flux[3], const vd_array4* rxn_src,
                               vd_array4* dsdt_out, const vd_box* box, const vd_geom* geom,
                               int specid) {
    if (!ctx || !flux || !rxn_src || !dsdt_out || !box || !geom) return VD_EINVAL;
    return dsdt(ctx, flux, rxn_src, dsdt_out, box, geom, specid);
}

extern "C" int vd_transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                const vd_array4* rxn_src, vd_array4* dsdt_out, const vd_box* box,
                                const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                                const vd_params* prm, double time, double dt, int specid) {
    (void)time;
    if (!ctx || !m || !sborder || !rxn_src || !dsdt_out || !box || !geom || !prm) return VD_EINVAL;
    VD_TRY(check_transport_args(prm, specid, m->hdev.nspec));
    return transport_rhs(ctx, m, sborder, rxn_src, dsdt_out, box, geom, bc_lo, bc_hi, prm, dt, specid);
}
