// vd_transport.hip -- drift-diffusion face fluxes (WENO-5 JS/M/Z, WAF
// MINBEE, first-order upwind) and the flux divergence.
//
// Restates compute_flux (Source/compute_explicit_flux.H:200-326) with the
// constant-mobility compute_vel of Source/PlasmaChem.H:186-206.
#include <algorithm>
#include <cfloat>

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
// where that numerator is finite, else in the reference's order
// (w_k = a_k / sum first, then sum w_k p_k) so that large alpha_k times
// large p_k does not overflow
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
    const double num = a1 * p1 + a2 * p2 + a3 * p3, den = a1 + a2 + a3;
    if (__builtin_expect(fabs(num) <= DBL_MAX, 1)) return fdiv(num, den);
    return fdiv(a1, den) * p1 + fdiv(a2, den) * p2 + fdiv(a3, den) * p3;
}

// Both WENO reconstructions centred on cell c from (a, b, c, d, e) =
// u[c-2 .. c+2]: L = weno(a,b,c,d,e), the value at face c+1/2 from the
// left, and R = weno(e,d,c,b,a), the value at face c-1/2 from the right.
// The mirrored call has the same smoothness indicators in reverse order
// (beta'_1 = beta_3, beta'_2 = beta_2, beta'_3 = beta_1) and the same tau,
// so they and the weight ratios are computed once for both.
template <int S>
__device__ __forceinline__ void weno5_lr(double a, double b, double c, double d, double e, double& L,
                                         double& R) {
    const double eps = 1e-20;
    const double p1 = (1.0 / 3.0) * a - (7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double p2 = -(1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double p3 = (1.0 / 3.0) * c + (5.0 / 6.0) * d - (1.0 / 6.0) * e;
    const double q1 = (1.0 / 3.0) * e - (7.0 / 6.0) * d + (11.0 / 6.0) * c;
    const double q2 = -(1.0 / 6.0) * d + (5.0 / 6.0) * c + (1.0 / 3.0) * b;
    const double q3 = (1.0 / 3.0) * c + (5.0 / 6.0) * b - (1.0 / 6.0) * a;
    const double b1 = (13.0 / 12.0) * sq(a - 2.0 * b + c) + (1.0 / 4.0) * sq(a - 4.0 * b + 3.0 * c);
    const double b2 = (13.0 / 12.0) * sq(b - 2.0 * c + d) + (1.0 / 4.0) * sq(b - d);
    const double b3 = (13.0 / 12.0) * sq(c - 2.0 * d + e) + (1.0 / 4.0) * sq(3.0 * c - 4.0 * d + e);
    const double g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0;
    double x1, x2, x3, y1, y2, y3;
    if (S == 3) {
        const double tau = fabs(b1 - b3);
        // tau / (b_k + eps) for k = 1..3 from one reciprocal of the product
        // (within a few ulps of the three quotients); the separate quotients
        // where the product is not a finite normal number
        const double e1 = b1 + eps, e2 = b2 + eps, e3 = b3 + eps;
        const double e12 = e1 * e2, prod = e12 * e3;
        double t1, t2, t3;
        if (__builtin_expect(prod < 1e300 && prod > 1e-300, 1)) {
            const double ti = fdiv(tau, prod);
            t1 = ti * (e2 * e3);
            t2 = ti * (e1 * e3);
            t3 = ti * e12;
        } else {
            t1 = fdiv(tau, e1);
            t2 = fdiv(tau, e2);
            t3 = fdiv(tau, e3);
        }
        const double r1 = sq(t1), r2 = sq(t2), r3 = sq(t3);
        x1 = g1 * (1.0 + r1);
        x2 = g2 * (1.0 + r2);
        x3 = g3 * (1.0 + r3);
        y1 = g1 * (1.0 + r3);
        y2 = g2 * (1.0 + r2);
        y3 = g3 * (1.0 + r1);
    } else {
        const double i1 = fdiv(1.0, sq(b1 + eps)), i2 = fdiv(1.0, sq(b2 + eps)), i3 = fdiv(1.0, sq(b3 + eps));
        x1 = g1 * i1;
        x2 = g2 * i2;
        x3 = g3 * i3;
        y1 = g1 * i3;
        y2 = g2 * i2;
        y3 = g3 * i1;
        if (S == 2) {
            const double sx = fdiv(1.0, x1 + x2 + x3), sy = fdiv(1.0, y1 + y2 + y3);
            x1 = gk(x1 * sx, g1);
            x2 = gk(x2 * sx, g2);
            x3 = gk(x3 * sx, g3);
            y1 = gk(y1 * sy, g1);
            y2 = gk(y2 * sy, g2);
            y3 = gk(y3 * sy, g3);
        }
    }
    // Both normalisations from one reciprocal of the product of the weight
    // sums, or separately where that product leaves the normal range.  The
    // numerators sum x_k p_k are formed before the division (the reference
    // divides the weights first): the same to a few ulps while they stay
    // finite, which holds for |u| below ~1e50 (WENO-Z weights reach
    // (tau / eps)^2 ~ |u|^4 / 1e-40 beside a flat stencil).
    const double nl = x1 * p1 + x2 * p2 + x3 * p3, dl = x1 + x2 + x3;
    const double nr = y1 * q1 + y2 * q2 + y3 * q3, dr = y1 + y2 + y3;
    const double dlr = dl * dr;
    if (__builtin_expect(dlr < 1e300 && dlr > 1e-300, 1)) {
        const double ti = fdiv(1.0, dlr);
        L = nl * (dr * ti);
        R = nr * (dl * ti);
    } else {
        L = fdiv(nl, dl);
        R = fdiv(nr, dr);
    }
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
    // the x component of E (or of a precomputed drift velocity, then mob =
    // 1), laid out like sb, one sb component stride per direction
    const double* ev;
    double mob;          // spec_mob(specid)
    int lo_zero, hi_zero; // face index that is zeroed (IHNEU) or INT_MIN
    int3 flo;            // first face index
    int fnx, fny, fnz;   // face box extent
    double dxmin, dt;
    __device__ void rebase(int64_t o) {
        vd::rebase(sb, o);
        vd::rebase(flux, o);
        ev = rb(ev, o);
    }
};

// one thread per face of direction a.dir
template <int HYP, int S>
__global__ void __launch_bounds__(256) k_face_flux(FluxArgs a, Batch bt) {
    a.rebase(case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.fnx * a.fny * a.fnz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.fnx, a.fny, i, j, k);
    i += a.flo.x;
    j += a.flo.y;
    k += a.flo.z;
    const int d = a.dir;
    const int st = d == 0 ? 1 : (d == 1 ? a.sb.js : a.sb.ks);
    const double* __restrict__ u = a.sb.comp(a.specid);
    const double* __restrict__ E = a.ev + (int64_t)d * a.sb.ns;
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

// WENO-5 fluxes of direction D (1 = y, 2 = z) streamed along pencils: one
// thread per (i, other) column and chunk of faces, x fastest across lanes
// (coalesced).  Every cell of the pencil is reconstructed once (both
// one-sided values of u and of f = u v), and each face flux combines the
// left value of the cell below with the right value of the cell above
// (compute_explicit_flux.H:262-312), so the per-face work of k_face_flux is
// shared between neighbouring faces.
struct PencilArgs {
    Arr<double> sb, flux;
    int specid, efcomp;
    const double* ev;      // as FluxArgs::ev
    double mob;
    int lo_zero, hi_zero;  // face index zeroed (IHNEU) or INT_MIN
    int3 flo;              // first face index (box lo)
    int nx, no, nd;        // cells along x, the other direction, and D
    int nchunk;            // face chunks along D
    // divergence fused into the last direction (fuse = 1): dsdt of cell m
    // from the x (and y) face arrays, rxn and this pencil's own two faces,
    // in the order of k_dsdt; the D-direction faces are then not stored
    int fuse;
    Arr<double> fx, fy, rxn, ds;
    double dx0, dx1, dx2;
    __device__ void rebase(int64_t o) {
        vd::rebase(sb, o);
        vd::rebase(flux, o);
        ev = rb(ev, o);
        if (fuse) {
            vd::rebase(fx, o);
            vd::rebase(fy, o);
            vd::rebase(rxn, o);
            vd::rebase(ds, o);
        }
    }
};

template <int S, int D>
__global__ void __launch_bounds__(256) k_pencil_flux(PencilArgs a, Batch bt) {
    a.rebase(case_off(bt));
    const unsigned t = lblock() * blockDim.x + threadIdx.x;
    if (t >= (unsigned)a.nx * a.no * a.nchunk) return;
    const int i = (int)(t % a.nx);
    const unsigned r = t / a.nx;
    const int o = (int)(r % a.no), ch = (int)(r / a.no);
    // faces [f0, f1): the nd + 1 faces split as evenly as the chunk count allows
    const int f0 = (int)((int64_t)ch * (a.nd + 1) / a.nchunk);
    const int f1 = (int)((int64_t)(ch + 1) * (a.nd + 1) / a.nchunk);
    // cell (i, o, m) along D -> Array4 offset
    const int ig = a.flo.x + i;
    const int jg = D == 1 ? 0 : a.flo.y + o;
    const int kg = D == 1 ? a.flo.z + o : 0;
    const int st = D == 1 ? a.sb.js : a.sb.ks;
    const int mlo = D == 1 ? a.flo.y : a.flo.z;
    const int base = D == 1 ? a.sb.off(ig, mlo, kg) : a.sb.off(ig, jg, mlo);
    const double* __restrict__ u = a.sb.comp(a.specid);
    const double* __restrict__ E = a.ev + (int64_t)D * a.sb.ns;
    // window of cells c-2 .. c+2 (u, v and f = u v) for centre c
    double w_u[5], w_f[5], w_v[5];
    auto cell = [&](int m, double& uu, double& ff, double& vv) {
        const int off = base + m * st;
        uu = u[off];
        vv = a.mob * E[off];
        ff = uu * vv;
    };
#pragma unroll
    for (int q = 0; q < 5; q++) cell(f0 - 3 + q, w_u[q], w_f[q], w_v[q]);
    // with the fused divergence a chunk also forms its first face above the
    // chunk, so every cell of the chunk has both of its faces here
    const int fend = a.fuse && ch != a.nchunk - 1 ? f1 + 1 : f1;
    double Lu_prev = 0.0, Lf_prev = 0.0, v_prev = 0.0, F_prev = 0.0;
    for (int c = f0 - 1; c < fend; c++) {
        double Lu, Ru, Lf, Rf;
        weno5_lr<S>(w_u[0], w_u[1], w_u[2], w_u[3], w_u[4], Lu, Ru);
        weno5_lr<S>(w_f[0], w_f[1], w_f[2], w_f[3], w_f[4], Lf, Rf);
        const double vc = w_v[2];
        if (c >= f0) {
            // face c between cells c-1 and c
            const double lam = fabs(0.5 * (v_prev + vc));
            double fl = 0.5 * (Lf_prev + lam * Lu_prev) + 0.5 * (Rf - lam * Ru);
            const int fd = mlo + c;
            if (fd == a.lo_zero || fd == a.hi_zero) fl = 0.0;
            if (!a.fuse) {
                if (D == 1) a.flux.p[a.flux.off(ig, fd, kg)] = fl;
                else a.flux.p[a.flux.off(ig, jg, fd)] = fl;
            } else if (c > f0) {
                // cell m = c-1 has faces c-1 (F_prev) and c (fl)
                const int mg = fd - 1;
                const int jm = D == 1 ? mg : jg, km = D == 1 ? kg : mg;
                double v = (a.fx(ig, jm, km) - a.fx(ig + 1, jm, km)) / a.dx0 + a.rxn(ig, jm, km, a.specid);
                if (D == 2) v += (a.fy(ig, jm, km) - a.fy(ig, jm + 1, km)) / a.dx1;
                v += (F_prev - fl) / (D == 1 ? a.dx1 : a.dx2);
                a.ds(ig, jm, km) = v;
            }
            F_prev = fl;
        }
        Lu_prev = Lu;
        Lf_prev = Lf;
        v_prev = vc;
        if (c + 1 < fend) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                w_u[q] = w_u[q + 1];
                w_f[q] = w_f[q + 1];
                w_v[q] = w_v[q + 1];
            }
            cell(c + 3, w_u[4], w_f[4], w_v[4]);
        }
    }
}

// WENO-5 fluxes of the x direction: each wave reconstructs 64 consecutive
// cells of one row once (both one-sided values of u and f) and forms the
// fluxes of the 63 faces between them, taking the right-hand values of the
// next cell from the neighbouring lane (wave shuffle).  Waves of a row
// overlap by one cell.
struct XLineArgs {
    Arr<double> sb, flux;
    int specid, efx;
    const double* ev;  // as FluxArgs::ev
    double mob;
    int lo_zero, hi_zero;
    int3 flo;
    int nx, nrow, nseg;  // cells along x, rows (y, z); nseg unused
    int ny;
    __device__ void rebase(int64_t o) {
        vd::rebase(sb, o);
        vd::rebase(flux, o);
        ev = rb(ev, o);
    }
};

template <int S>
__global__ void __launch_bounds__(256) k_xline_flux(XLineArgs a, Batch bt) {
    a.rebase(case_off(bt));
    const unsigned wv = (lblock() * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    // the rows' reconstructed cells -1 .. nx laid end to end: wave wv
    // covers positions 63 wv .. 63 wv + 63, so waves overlap by one cell
    // and a wave may span the end of one row and the start of the next
    const int rl = a.nx + 2;
    const int64_t vtot = (int64_t)a.nrow * rl;
    if ((int64_t)wv * 63 >= vtot) return;  // whole waves exit together
    const int64_t v = (int64_t)wv * 63 + lane;
    const bool live = v < vtot;
    const int64_t vq = live ? v : vtot - 1;
    const int row = (int)(vq / rl);
    const int c = (int)(vq - (int64_t)row * rl) - 1;  // centre cell, relative to box lo
    const int jg = a.flo.y + row % a.ny, kg = a.flo.z + row / a.ny;
    const double* __restrict__ u = a.sb.comp(a.specid);
    const double* __restrict__ E = a.ev;
    const int base = a.sb.off(a.flo.x, jg, kg);
    double wu[5], wf[5], vc = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const int m = min(c - 2 + q, a.nx + 2);  // clamped for idle lanes
        const double uu = u[base + m], vv = a.mob * E[base + m];
        wu[q] = uu;
        wf[q] = uu * vv;
        if (q == 2) vc = vv;
    }
    double Lu = 0.0, Ru = 0.0, Lf = 0.0, Rf = 0.0;
    if (live) {
        weno5_lr<S>(wu[0], wu[1], wu[2], wu[3], wu[4], Lu, Ru);
        weno5_lr<S>(wf[0], wf[1], wf[2], wf[3], wf[4], Lf, Rf);
    }
    // face c+1 between this lane's cell and the next lane's
    const double Ru1 = __shfl_down(Ru, 1, 64), Rf1 = __shfl_down(Rf, 1, 64), v1 = __shfl_down(vc, 1, 64);
    const int f = c + 1;  // its right-hand cell is the next lane's when c < nx
    if (lane < 63 && live && c < a.nx) {
        const double lam = fabs(0.5 * (vc + v1));
        double fl = 0.5 * (Lf + lam * Lu) + 0.5 * (Rf1 - lam * Ru1);
        const int fd = a.flo.x + f;
        if (fd == a.lo_zero || fd == a.hi_zero) fl = 0.0;
        a.flux.p[a.flux.off(fd, jg, kg)] = fl;
    }
}

struct DsdtArgs {
    Arr<double> f0, f1, f2, rxn, dsdt;
    int specid, dim;
    double dx0, dx1, dx2;
    int3 lo;
    int nx, ny, nz;
    __device__ void rebase(int64_t o) {
        vd::rebase(f0, o);
        vd::rebase(f1, o);
        vd::rebase(f2, o);
        vd::rebase(rxn, o);
        vd::rebase(dsdt, o);
    }
};

// compute_dsdt (Source/ScalarSolve.cpp:44-53)
__global__ void __launch_bounds__(256) k_dsdt(DsdtArgs a, Batch bt) {
    a.rebase(case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
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
    hipLaunchKernelGGL((k_face_flux<HYP, S>), bz(ctx, grid1d(tot, 256)), dim3(256), 0, ctx->stream, a, ctx->bt);
}

// chunk of faces per pencil thread along D: long enough to amortise the
// 4-cell ramp, short enough to give the device >= ~8 waves per CU
static int launch_pencil(vd_ctx* ctx, const FluxArgs& f, int d, int scheme, const vd_box* box,
                         const vd_array4* fx = nullptr, const vd_array4* fy = nullptr,
                         const vd_array4* rxn = nullptr, vd_array4* ds = nullptr, const vd_geom* geom = nullptr) {
    PencilArgs a;
    a.fuse = ds != nullptr;
    if (a.fuse) {
        a.fx = arr(fx);
        a.fy = fy ? arr(fy) : a.fx;
        a.rxn = arr(rxn);
        a.ds = arr(ds);
        a.dx0 = geom->dx[0];
        a.dx1 = geom->dx[1];
        a.dx2 = geom->dx[2];
    }
    a.sb = f.sb;
    a.flux = f.flux;
    a.specid = f.specid;
    a.efcomp = f.efcomp;
    a.ev = f.ev;
    a.mob = f.mob;
    a.lo_zero = f.lo_zero;
    a.hi_zero = f.hi_zero;
    a.flo = f.flo;
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.nd = box->hi[d] - box->lo[d] + 1;
    a.no = d == 1 ? box->hi[2] - box->lo[2] + 1 : box->hi[1] - box->lo[1] + 1;
    const int64_t cols = (int64_t)a.nx * a.no;
    // chunks of 12-48 faces, as many as fill whole rounds of resident
    // workgroups (each thread streams its chunk, so a nearly empty last
    // round would run alone)
    const void* kfn = d == 1 ? (scheme == 1 ? (const void*)k_pencil_flux<1, 1>
                                            : (scheme == 2 ? (const void*)k_pencil_flux<2, 1>
                                                           : (const void*)k_pencil_flux<3, 1>))
                             : (scheme == 1 ? (const void*)k_pencil_flux<1, 2>
                                            : (scheme == 2 ? (const void*)k_pencil_flux<2, 2>
                                                           : (const void*)k_pencil_flux<3, 2>));
    const int nf = a.nd + 1;
    const int nmin = std::max(1, (nf + 47) / 48), nmax = std::max(nmin, nf / 12);
    a.nchunk = round_filling_chunks((cols + 255) / 256, resident_blocks(ctx, kfn, 256), nmin, nmax);
    const int64_t tot = cols * a.nchunk;
    if (tot >= ((int64_t)1 << 32)) {
        set_error("transport: box too large for one pencil launch");
        return VD_EINVAL;
    }
    dim3 g = grid1d(tot, 256), b(256);
    if (d == 1) {
        if (scheme == 1) hipLaunchKernelGGL((k_pencil_flux<1, 1>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
        else if (scheme == 2) hipLaunchKernelGGL((k_pencil_flux<2, 1>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
        else hipLaunchKernelGGL((k_pencil_flux<3, 1>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
    } else {
        if (scheme == 1) hipLaunchKernelGGL((k_pencil_flux<1, 2>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
        else if (scheme == 2) hipLaunchKernelGGL((k_pencil_flux<2, 2>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
        else hipLaunchKernelGGL((k_pencil_flux<3, 2>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

static int launch_xline(vd_ctx* ctx, const FluxArgs& f, int scheme, const vd_box* box) {
    XLineArgs a;
    a.sb = f.sb;
    a.flux = f.flux;
    a.specid = f.specid;
    a.efx = f.efcomp;
    a.ev = f.ev;
    a.mob = f.mob;
    a.lo_zero = f.lo_zero;
    a.hi_zero = f.hi_zero;
    a.flo = f.flo;
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nrow = a.ny * (box->hi[2] - box->lo[2] + 1);
    a.nseg = 0;  // unused: waves run over the rows laid end to end
    const int64_t threads = ((int64_t)a.nrow * (a.nx + 2) + 62) / 63 * 64;
    if (threads >= ((int64_t)1 << 32)) {
        set_error("transport: box too large for one flux launch");
        return VD_EINVAL;
    }
    dim3 g = grid1d(threads, 256), b(256);
    if (scheme == 1) hipLaunchKernelGGL((k_xline_flux<1>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
    else if (scheme == 2) hipLaunchKernelGGL((k_xline_flux<2>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
    else hipLaunchKernelGGL((k_xline_flux<3>), bz(ctx, g), b, 0, ctx->stream, a, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// The drift velocity of species s in every cell of sb's allocated box
// (ghost cells included), for species whose velocity is not a constant
// mobility times E: user_transport::compute_vel's prescribed velocity, or
// specMob(s, Te, ndens, |E|, Tg) * E_d (Source/PlasmaChem.H:186-206).
// vel has sb's layout, one component per direction.
struct DriftArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb;
    int spec, dim;
    int nx, ny, nz;  // sb's allocated extent
    double Tg;
    double* vel;
    __device__ void rebase(int64_t o) {
        M = rb(M, o);
        vd::rebase(sb, o);
        vel = rb(vel, o);
    }
};

__global__ void __launch_bounds__(256) k_drift_vel(DriftArgs a, Batch bt) {
    a.rebase(case_off(bt));
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    const DevMech& M = *a.M;
    const int ns = M.nspec;
    const int o = a.sb.off(a.sb.bx + i, a.sb.by + j, a.sb.bz + k);
    const double* S = a.sb.p;
    const int64_t N = a.sb.ns;
    if (M.vel_set[a.spec]) {
        for (int d = 0; d < a.dim; d++) a.vel[(int64_t)d * N + o] = M.vel[a.spec][d];
        return;
    }
    double nd = 0.0, es = 0.0;
    for (int sp = 0; sp < ns; sp++) nd += S[(int64_t)sp * N + o];
    for (int d = 0; d < 3; d++) {
        const double e = S[(int64_t)(VD_EFX(ns) + d) * N + o];
        es += e * e;
    }
    const double mu = spec_mob_st(M, a.spec, S[(int64_t)VD_ETEMP(ns) * N + o], a.Tg, nd, sqrt(es));
    for (int d = 0; d < a.dim; d++) a.vel[(int64_t)d * N + o] = mu * S[(int64_t)(VD_EFX(ns) + d) * N + o];
}

// doubles of the drift-velocity array k_drift_vel writes for sb
int64_t drift_vel_doubles(const vd_array4* sb, int dim) { return sb->nstride * (int64_t)dim; }

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

// with rxn and ds given (WENO-5 only), the divergence plus rxn is formed in
// the last direction's flux kernel and written to ds (that direction's face
// array is then left unwritten)
int transport_flux(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4 flux[3],
                   const vd_box* box, const vd_geom* g, const int bc_lo[3], const int bc_hi[3],
                   const vd_params* prm, double dt, int specid, const vd_array4* rxn = nullptr,
                   vd_array4* ds = nullptr, double* vbuf = nullptr) {
    VD_TRY(check_transport_args(prm, specid, m->hdev.nspec));
    int ng = prm->hyp_order == 5 ? 3 : prm->hyp_order;
    for (int d = 0; d < g->dim; d++)
        if (sb->begin[d] > box->lo[d] - ng || sb->end[d] < box->hi[d] + 1 + ng) {
            set_error("transport: Sborder needs %d ghost cells", ng);
            return VD_EINVAL;
        }
    // drift velocities: a constant mobility times E read in the flux
    // kernels, or (fitted mobility, prescribed velocity) precomputed per cell
    const double* ev = sb->p + (int64_t)VD_EFX(m->hdev.nspec) * sb->nstride;
    double mob = spec_mob(m->hdev, specid);
    if (drift_is_state(m->hdev, specid)) {
        const int64_t nv = drift_vel_doubles(sb, g->dim);
        if (!vbuf) {
            void* p;
            VD_TRY(scratch(ctx, "transport_vel", nv * sizeof(double), &p));
            vbuf = (double*)p;
        }
        DriftArgs v;
        v.M = m->dev;
        v.sb = arr(sb);
        v.spec = specid;
        v.dim = g->dim;
        v.nx = sb->end[0] - sb->begin[0];
        v.ny = sb->end[1] - sb->begin[1];
        v.nz = sb->end[2] - sb->begin[2];
        v.Tg = prm->gas_temperature;
        v.vel = vbuf;
        hipLaunchKernelGGL(k_drift_vel, bz(ctx, grid1d((int64_t)v.nx * v.ny * v.nz, 256)), dim3(256), 0, ctx->stream, v,
                           ctx->bt);
        VD_CHECK_LAUNCH();
        ev = vbuf;
        mob = 1.0;
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
        a.ev = ev;
        a.mob = mob;
        a.lo_zero = bc_lo[d] == VD_IHNEUBC ? g->dom_lo[d] : INT32_MIN;
        a.hi_zero = bc_hi[d] == VD_IHNEUBC ? g->dom_hi[d] + 1 : INT32_MIN;
        a.flo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
        a.fnx = box->hi[0] - box->lo[0] + 1 + (d == 0);
        a.fny = box->hi[1] - box->lo[1] + 1 + (d == 1);
        a.fnz = box->hi[2] - box->lo[2] + 1 + (d == 2);
        a.dxmin = dxmin;
        a.dt = dt;
        if (prm->hyp_order == 5) {
            if (d > 0 && ds && d == g->dim - 1)
                VD_TRY(launch_pencil(ctx, a, d, prm->weno_scheme, box, &flux[0], g->dim == 3 ? &flux[1] : nullptr,
                                     rxn, ds, g));
            else if (d > 0) VD_TRY(launch_pencil(ctx, a, d, prm->weno_scheme, box));
            else VD_TRY(launch_xline(ctx, a, prm->weno_scheme, box));
            continue;
        }
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
    hipLaunchKernelGGL(k_dsdt, bz(ctx, grid1d(tot, 256)), dim3(256), 0, ctx->stream, a, ctx->bt);
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

// face arrays transport_rhs writes: with the WENO divergence fused into
// the last direction's pencils that direction's faces stay on chip,
// otherwise every direction's array is written (and zeroed without
// transport)
int transport_rhs_nflux(const vd_params* prm, const vd_geom* g) {
    return prm->do_transport && prm->hyp_order == 5 && g->dim > 1 ? g->dim - 1 : g->dim;
}

int64_t transport_face_count(const vd_box* box, int d) {
    return (int64_t)(box->hi[0] - box->lo[0] + 1 + (d == 0)) * (box->hi[1] - box->lo[1] + 1 + (d == 1)) *
           (box->hi[2] - box->lo[2] + 1 + (d == 2));
}

// update_explsrc_at_all_levels for one level (Source/ScalarSolve.cpp:57-111):
// fluxes zeroed, computed if do_transport, then the divergence.  The face
// arrays are the context's scratch buffers, or with fbuf the caller's
// buffers for the first transport_rhs_nflux directions (no allocation and
// no kernel timing then); likewise the per-cell drift velocities of a
// species with a fitted mobility or a prescribed velocity, in vbuf
// (drift_vel_doubles) when given.
int transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, const vd_array4* rxn,
                  vd_array4* out, const vd_box* box, const vd_geom* g, const int bc_lo[3],
                  const int bc_hi[3], const vd_params* prm, double dt, int specid,
                  double* const* fbuf, double* vbuf) {
    vd_array4 fl[3];
    const int nused = transport_rhs_nflux(prm, g);
    for (int d = 0; d < 3; d++) {
        int dd = d < g->dim ? d : 0;
        int64_t n = transport_face_count(box, dd);
        void* p;
        if (fbuf) {
            // unwritten directions view the first array (never stored to)
            p = fbuf[d < nused ? d : 0];
        } else {
            char name[32];
            snprintf(name, sizeof(name), "transport_flux%d", d);
            VD_TRY(scratch(ctx, name, n * sizeof(double), &p));
        }
        fl[d] = face_view((double*)p, box, dd);
        if (!prm->do_transport && d < g->dim) VD_TRY(fill_value(ctx, (double*)p, n, 0.0));
    }
    double cells = (double)(box->hi[0] - box->lo[0] + 1) * (box->hi[1] - box->lo[1] + 1) *
                   (box->hi[2] - box->lo[2] + 1);
    if (!fbuf) kt_begin(ctx, VD_KT_TRANSPORT);
    int st = VD_OK;
    if (prm->do_transport && prm->hyp_order == 5 && g->dim > 1) {
        // WENO fluxes with the divergence fused into the last direction
        st = transport_flux(ctx, m, sb, fl, box, g, bc_lo, bc_hi, prm, dt, specid, rxn, out, vbuf);
    } else {
        if (prm->do_transport)
            VD_TRY(transport_flux(ctx, m, sb, fl, box, g, bc_lo, bc_hi, prm, dt, specid, nullptr, nullptr, vbuf));
        st = dsdt(ctx, fl, rxn, out, box, g, specid);
    }
    // what the launches must move, per cell: with the divergence fused into
    // the last direction, u and E_d read and the face flux written by each
    // of the other directions, and the last reading u, E_d, those faces and
    // rxn and writing dsdt (4 dim words); otherwise 3 per direction plus
    // dsdt's dim + 2; a per-cell drift velocity adds its pass over Sborder
    if (!fbuf) {
        const bool fused = prm->do_transport && prm->hyp_order == 5 && g->dim > 1;
        double words = fused ? 4.0 * g->dim : 4.0 * g->dim + 2.0;
        if (prm->do_transport && drift_is_state(m->hdev, specid))
            words += (double)sb->nstride / cells * (m->hdev.nspec + 4 + g->dim);
        kt_end(ctx, VD_KT_TRANSPORT, 8.0 * cells * words);
    }
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
    VD_TRY(check_extents("compute_scalar_transport_flux", box, {sborder}));
    VD_TRY(check_face_extents("compute_scalar_transport_flux", flux, geom->dim));
    return transport_flux(ctx, m, sborder, flux, box, geom, bc_lo, bc_hi, prm, dt, specid);
}

extern "C" int vd_compute_dsdt(vd_ctx* ctx, const vd_array4 flux[3], const vd_array4* rxn_src,
                               vd_array4* dsdt_out, const vd_box* box, const vd_geom* geom,
                               int specid) {
    if (!ctx || !flux || !rxn_src || !dsdt_out || !box || !geom) return VD_EINVAL;
    VD_TRY(check_extents("compute_dsdt", box, {rxn_src, dsdt_out}));
    VD_TRY(check_face_extents("compute_dsdt", flux, geom->dim));
    return dsdt(ctx, flux, rxn_src, dsdt_out, box, geom, specid);
}

extern "C" int vd_transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                const vd_array4* rxn_src, vd_array4* dsdt_out, const vd_box* box,
                                const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                                const vd_params* prm, double time, double dt, int specid) {
    (void)time;
    if (!ctx || !m || !sborder || !rxn_src || !dsdt_out || !box || !geom || !prm) return VD_EINVAL;
    VD_TRY(check_extents("transport_rhs", box, {sborder, rxn_src, dsdt_out}));
    VD_TRY(check_transport_args(prm, specid, m->hdev.nspec));
    return transport_rhs(ctx, m, sborder, rxn_src, dsdt_out, box, geom, bc_lo, bc_hi, prm, dt, specid, nullptr,
                         nullptr);
}
