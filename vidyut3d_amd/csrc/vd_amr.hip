// vd_amr.hip -- level transfers between an AMR level and the next finer
// one: conservative averaging of cells and of face fluxes onto the coarse
// level, and the fine-level ghost fill from fine boxes and coarse
// interpolation.
//
// Restates Vidyut::AverageDown / AverageDownTo, FillPatch (lev > 0) and
// FillCoarsePatch (Source/GridWork.cpp:94-164) and the flux
// average_down_faces of update_explsrc_at_all_levels
// (Source/ScalarSolve.cpp:88-94).  The AMReX routines they call
// (average_down, average_down_faces, FillPatchTwoLevels with
// cell_cons_interp) are restated from their published algorithms.
#include "vd_internal.hpp"

namespace vd {

// floor(i / r) for r > 0 (amrex::coarsen)
__host__ __device__ inline int crsn(int i, int r) { return i >= 0 ? i / r : -((-i - 1) / r) - 1; }

__device__ __forceinline__ int wrap_or_clamp(int c, int lo, int hi, int per) {
    if (c >= lo && c <= hi) return c;
    if (per) {
        const int n = hi - lo + 1;
        int r = (c - lo) % n;
        if (r < 0) r += n;
        return lo + r;
    }
    return c < lo ? lo : hi;
}

// ------------------------------------------------------------ average_down
// one thread per coarse cell of the region, blockIdx.y = component
__global__ void __launch_bounds__(256) k_average_down(Arr<double> fine, Arr<double> crse, int3 clo, int3 cn,
                                                       int3 r, double volfrac, int scomp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cn.x * cn.y * cn.z) return;
    const int n = scomp + (int)blockIdx.y;
    const int ic = clo.x + (int)(t % cn.x);
    const int64_t row = t / cn.x;
    const int jc = clo.y + (int)(row % cn.y), kc = clo.z + (int)(row / cn.y);
    const int ii = ic * r.x, jj = jc * r.y, kk = kc * r.z;
    const double* __restrict__ f = fine.comp(n);
    double c = 0.0;
    for (int kr = 0; kr < r.z; kr++)
        for (int jr = 0; jr < r.y; jr++)
            for (int ir = 0; ir < r.x; ir++) c += f[fine.off(ii + ir, jj + jr, kk + kr)];
    crse.comp(n)[crse.off(ic, jc, kc)] = volfrac * c;
}

// one thread per coarse face of direction d; the tangential children are
// summed with the higher direction outermost
__global__ void __launch_bounds__(256) k_average_down_faces(Arr<double> fine, Arr<double> crse, int d, int3 clo,
                                                             int3 cn, int3 r, double facinv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cn.x * cn.y * cn.z) return;
    const int ic = clo.x + (int)(t % cn.x);
    const int64_t row = t / cn.x;
    const int jc = clo.y + (int)(row % cn.y), kc = clo.z + (int)(row / cn.y);
    const int ii = ic * r.x, jj = jc * r.y, kk = kc * r.z;
    // tangential extents (a = lower tangential direction, b = higher)
    const int na = d == 0 ? r.y : r.x, nb = d == 2 ? r.y : r.z;
    double c = 0.0;
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            const int i = ii + (d == 0 ? 0 : a);
            const int j = jj + (d == 0 ? a : (d == 1 ? 0 : b));
            const int k = kk + (d == 2 ? 0 : b);
            c += fine.p[fine.off(i, j, k)];
        }
    crse.p[crse.off(ic, jc, kc)] = facinv * c;
}

// ------------------------------------------------ two-level ghost fill
struct FineBox {
    int lo[3], hi[3];
    Arr<double> a;
};

struct TwoLevArgs {
    Arr<double> dst, crse;
    const FineBox* __restrict__ boxes;
    int nbox;
    int dim, scomp, ncomp;
    int3 dbeg, dn;           // dst extent (begin, size)
    int3 flo, fhi, clo, chi; // fine and coarse domains
    int3 per, r;
};

// Conservative linear interpolation of coarse cell (ic) at fine offsets
// (xo) for component n: MC-limited central slopes per direction, then one
// factor alpha <= 1 so that no child of the coarse cell leaves the min/max
// of its 3^dim neighbourhood.  FP contraction is off: the arithmetic
// follows the statements one operation at a time.
template <int DIM>
__device__ double cons_lin(const TwoLevArgs& a, int n, const int ic[3], const double xo[3]) {
#pragma clang fp contract(off)
    const double* __restrict__ C = a.crse.comp(n);
    const int clo[3] = {a.clo.x, a.clo.y, a.clo.z}, chi[3] = {a.chi.x, a.chi.y, a.chi.z};
    const int per[3] = {a.per.x, a.per.y, a.per.z}, rr[3] = {a.r.x, a.r.y, a.r.z};
    auto at = [&](int di, int dj, int dk) {
        const int i = wrap_or_clamp(ic[0] + di, clo[0], chi[0], per[0]);
        const int j = wrap_or_clamp(ic[1] + dj, clo[1], chi[1], per[1]);
        const int k = DIM == 3 ? wrap_or_clamp(ic[2] + dk, clo[2], chi[2], per[2]) : ic[2];
        return C[a.crse.off(i, j, k)];
    };
    const double u0 = at(0, 0, 0);
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double up = at(d == 0, d == 1, d == 2), um = at(-(d == 0), -(d == 1), -(d == 2));
        const double cen = 0.5 * (up - um);
        const double forw = 2.0 * (up - u0);
        const double back = 2.0 * (u0 - um);
        const double slp = (forw * back >= 0.0) ? fmin(fabs(forw), fabs(back)) : 0.0;
        s[d] = copysign(1.0, cen) * fmin(slp, fabs(cen));
    }
    double cmn = u0, cmx = u0;
    for (int dk = (DIM == 3 ? -1 : 0); dk <= (DIM == 3 ? 1 : 0); dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const double v = at(di, dj, dk);
                cmn = fmin(cmn, v);
                cmx = fmax(cmx, v);
            }
    double dmn = 0.0, dmx = 0.0;
    for (int kr = 0; kr < (DIM == 3 ? rr[2] : 1); kr++)
        for (int jr = 0; jr < rr[1]; jr++)
            for (int ir = 0; ir < rr[0]; ir++) {
                const double ox = (ir + 0.5) / rr[0] - 0.5;
                const double oy = (jr + 0.5) / rr[1] - 0.5;
                double y = ox * s[0] + oy * s[1];
                if (DIM == 3) {
                    const double oz = (kr + 0.5) / rr[2] - 0.5;
                    y = y + oz * s[2];
                }
                dmn = fmin(dmn, y);
                dmx = fmax(dmx, y);
            }
    double alpha = 1.0;
    if (dmn != 0.0 && u0 + dmn < cmn) alpha = fmin(alpha, (cmn - u0) / dmn);
    if (dmx != 0.0 && u0 + dmx > cmx) alpha = fmin(alpha, (cmx - u0) / dmx);
    double v = u0 + xo[0] * (alpha * s[0]);
    v = v + xo[1] * (alpha * s[1]);
    if (DIM == 3) v = v + xo[2] * (alpha * s[2]);
    return v;
}

template <int DIM>
__global__ void __launch_bounds__(256) k_fill_two_levels(TwoLevArgs a) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.dn.x * a.dn.y * a.dn.z) return;
    const int i = a.dbeg.x + (int)(t % a.dn.x);
    const int64_t row = t / a.dn.x;
    const int j = a.dbeg.y + (int)(row % a.dn.y), k = a.dbeg.z + (int)(row / a.dn.y);
    // the cell inside the fine domain whose value this one takes: periodic
    // directions wrap, the others clamp (foextrap)
    int q[3] = {wrap_or_clamp(i, a.flo.x, a.fhi.x, a.per.x), wrap_or_clamp(j, a.flo.y, a.fhi.y, a.per.y),
                DIM == 3 ? wrap_or_clamp(k, a.flo.z, a.fhi.z, a.per.z) : k};
    int hit = -1;
    for (int b = 0; b < a.nbox && hit < 0; b++) {
        const FineBox& B = a.boxes[b];
        if (q[0] >= B.lo[0] && q[0] <= B.hi[0] && q[1] >= B.lo[1] && q[1] <= B.hi[1] &&
            (DIM == 2 || (q[2] >= B.lo[2] && q[2] <= B.hi[2])))
            hit = b;
    }
    const int o = a.dst.off(i, j, k);
    if (hit >= 0) {
        const Arr<double>& F = a.boxes[hit].a;
        const int fo = F.off(q[0], q[1], q[2]);
        for (int n = a.scomp; n < a.scomp + a.ncomp; n++) a.dst.comp(n)[o] = F.comp(n)[fo];
        return;
    }
    const int rr[3] = {a.r.x, a.r.y, a.r.z};
    int ic[3];
    double xo[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d < DIM) {
            ic[d] = crsn(q[d], rr[d]);
            xo[d] = (q[d] - ic[d] * rr[d] + 0.5) / rr[d] - 0.5;
        } else {
            ic[d] = q[d];
        }
    }
    for (int n = a.scomp; n < a.scomp + a.ncomp; n++) a.dst.comp(n)[o] = cons_lin<DIM>(a, n, ic, xo);
}

static bool inside(const vd_array4* A, const int lo[3], const int hi[3], int dim) {
    for (int d = 0; d < dim; d++)
        if (lo[d] < A->begin[d] || hi[d] >= A->end[d]) return false;
    return true;
}

static int check_ratio(const int ratio[3], int dim) {
    if (dim != 2 && dim != 3) {
        set_error("level transfer: dim %d", dim);
        return VD_EINVAL;
    }
    for (int d = 0; d < dim; d++)
        if (ratio[d] != 2 && ratio[d] != 4) {
            set_error("level transfer: refinement ratio %d (2 or 4)", ratio[d]);
            return VD_EINVAL;
        }
    return VD_OK;
}

static int check_aligned(const vd_box* fb, const int ratio[3], int dim) {
    for (int d = 0; d < dim; d++)
        if (fb->hi[d] < fb->lo[d] || fb->lo[d] % ratio[d] != 0 || (fb->hi[d] + 1) % ratio[d] != 0) {
            set_error("level transfer: fine box not coarsenable by the refinement ratio in direction %d", d);
            return VD_EINVAL;
        }
    return VD_OK;
}

}  // namespace vd

using namespace vd;

extern "C" int vd_average_down(vd_ctx* ctx, const vd_array4* fine, const vd_box* fbox, vd_array4* crse,
                               int scomp, int ncomp, const int ratio[3], int dim) {
    if (!ctx || !fine || !fbox || !crse || !ratio) return VD_EINVAL;
    VD_TRY(check_extents("average_down", fbox, {fine, crse}));
    VD_TRY(check_ratio(ratio, dim));
    VD_TRY(check_aligned(fbox, ratio, dim));
    VD_TRY(require_plain(ctx, "average_down"));
    if (ncomp <= 0) return VD_OK;
    if (scomp < 0 || scomp + ncomp > fine->ncomp || scomp + ncomp > crse->ncomp) {
        set_error("vd_average_down: components [%d, %d) beyond the arrays", scomp, scomp + ncomp);
        return VD_EINVAL;
    }
    const int r[3] = {ratio[0], ratio[1], dim == 3 ? ratio[2] : 1};
    int clo[3], chi[3];
    for (int d = 0; d < 3; d++) {
        clo[d] = d < dim ? crsn(fbox->lo[d], r[d]) : fbox->lo[d];
        chi[d] = d < dim ? crsn(fbox->hi[d], r[d]) : fbox->hi[d];
    }
    if (!inside(fine, fbox->lo, fbox->hi, dim) || !inside(crse, clo, chi, dim)) {
        set_error("vd_average_down: arrays do not cover the fine box or its coarsening");
        return VD_EINVAL;
    }
    const int3 cn = make_int3(chi[0] - clo[0] + 1, chi[1] - clo[1] + 1, dim == 3 ? chi[2] - clo[2] + 1 : 1);
    const int64_t tot = (int64_t)cn.x * cn.y * cn.z;
    hipLaunchKernelGGL(k_average_down, dim3((unsigned)((tot + 255) / 256), ncomp), dim3(256), 0, ctx->stream,
                       arr(fine), arr(crse), make_int3(clo[0], clo[1], clo[2]), cn, make_int3(r[0], r[1], r[2]),
                       1.0 / (double)(r[0] * r[1] * r[2]), scomp);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

extern "C" int vd_average_down_faces(vd_ctx* ctx, const vd_array4 fine[3], const vd_box* fbox, vd_array4 crse[3],
                                     const int ratio[3], int dim) {
    if (!ctx || !fine || !fbox || !crse || !ratio) return VD_EINVAL;
    VD_TRY(check_box_extent("average_down_faces", fbox));
    VD_TRY(check_face_extents("average_down_faces", fine, dim));
    VD_TRY(check_face_extents("average_down_faces", crse, dim));
    VD_TRY(check_ratio(ratio, dim));
    VD_TRY(check_aligned(fbox, ratio, dim));
    VD_TRY(require_plain(ctx, "average_down_faces"));
    const int r[3] = {ratio[0], ratio[1], dim == 3 ? ratio[2] : 1};
    for (int d = 0; d < dim; d++) {
        int flo[3], fhi[3], clo[3], chi[3];
        for (int e = 0; e < 3; e++) {
            flo[e] = fbox->lo[e];
            fhi[e] = fbox->hi[e] + (e == d ? 1 : 0);
            clo[e] = e < dim ? crsn(flo[e], r[e]) : flo[e];
            chi[e] = e < dim ? crsn(fhi[e], r[e]) : fhi[e];
        }
        if (!inside(&fine[d], flo, fhi, dim) || !inside(&crse[d], clo, chi, dim)) {
            set_error("vd_average_down_faces: face arrays of direction %d do not cover the fine box", d);
            return VD_EINVAL;
        }
        const int3 cn = make_int3(chi[0] - clo[0] + 1, chi[1] - clo[1] + 1, dim == 3 ? chi[2] - clo[2] + 1 : 1);
        const int64_t tot = (int64_t)cn.x * cn.y * cn.z;
        const double facinv = 1.0 / (double)(d == 0 ? r[1] * r[2] : (d == 1 ? r[0] * r[2] : r[0] * r[1]));
        hipLaunchKernelGGL(k_average_down_faces, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                           arr(&fine[d]), arr(&crse[d]), d, make_int3(clo[0], clo[1], clo[2]), cn,
                           make_int3(r[0], r[1], r[2]), facinv);
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

extern "C" int vd_fill_patch_two_levels(vd_ctx* ctx, const vd_array4* crse, const vd_geom* cg, const vd_array4* fine,
                                        const vd_box* fboxes, int nfine, vd_array4* dst, const vd_geom* fg,
                                        const int ratio[3], int scomp, int ncomp) {
    if (!ctx || !crse || !cg || !dst || !fg || !ratio || nfine < 0 || (nfine > 0 && (!fine || !fboxes)))
        return VD_EINVAL;
    VD_TRY(check_extents("fill_patch_two_levels", nullptr, {crse, dst}));
    for (int b = 0; b < nfine; b++) VD_TRY(check_extents("fill_patch_two_levels", &fboxes[b], {&fine[b]}));
    VD_TRY(require_plain(ctx, "fill_patch_two_levels"));
    const int dim = fg->dim;
    if (cg->dim != dim) {
        set_error("vd_fill_patch_two_levels: coarse and fine geometries differ in dim");
        return VD_EINVAL;
    }
    VD_TRY(check_ratio(ratio, dim));
    if (ncomp <= 0) return VD_OK;
    if (scomp < 0 || scomp + ncomp > crse->ncomp || scomp + ncomp > dst->ncomp) {
        set_error("vd_fill_patch_two_levels: components [%d, %d) beyond the arrays", scomp, scomp + ncomp);
        return VD_EINVAL;
    }
    for (int d = 0; d < dim; d++) {
        if (fg->dom_lo[d] != cg->dom_lo[d] * ratio[d] || fg->dom_hi[d] + 1 != (cg->dom_hi[d] + 1) * ratio[d] ||
            fg->is_periodic[d] != cg->is_periodic[d]) {
            set_error("vd_fill_patch_two_levels: fine domain is not the refined coarse domain (direction %d)", d);
            return VD_EINVAL;
        }
    }
    if (!inside(crse, cg->dom_lo, cg->dom_hi, dim)) {
        set_error("vd_fill_patch_two_levels: coarse array does not cover the coarse domain");
        return VD_EINVAL;
    }
    std::vector<FineBox> hb((size_t)nfine);
    for (int b = 0; b < nfine; b++) {
        const vd_box& B = fboxes[b];
        for (int d = 0; d < dim; d++)
            if (B.lo[d] < fg->dom_lo[d] || B.hi[d] > fg->dom_hi[d] || B.hi[d] < B.lo[d]) {
                set_error("vd_fill_patch_two_levels: fine box %d outside the fine domain", b);
                return VD_EINVAL;
            }
        if (!inside(&fine[b], B.lo, B.hi, dim) || scomp + ncomp > fine[b].ncomp) {
            set_error("vd_fill_patch_two_levels: fine array %d does not cover its box", b);
            return VD_EINVAL;
        }
        for (int d = 0; d < 3; d++) {
            hb[b].lo[d] = B.lo[d];
            hb[b].hi[d] = B.hi[d];
        }
        hb[b].a = arr(&fine[b]);
    }
    FineBox* db = nullptr;
    if (nfine > 0) {
        void* p = nullptr;
        VD_TRY(scratch(ctx, "amr_fine_boxes", sizeof(FineBox) * (size_t)nfine, &p));
        db = (FineBox*)p;
        // pageable source: the call returns once the list has been staged
        VD_HIP(hipMemcpyAsync(db, hb.data(), sizeof(FineBox) * (size_t)nfine, hipMemcpyHostToDevice, ctx->stream));
    }
    TwoLevArgs a;
    a.dst = arr(dst);
    a.crse = arr(crse);
    a.boxes = db;
    a.nbox = nfine;
    a.dim = dim;
    a.scomp = scomp;
    a.ncomp = ncomp;
    a.dbeg = make_int3(dst->begin[0], dst->begin[1], dst->begin[2]);
    a.dn = make_int3(dst->end[0] - dst->begin[0], dst->end[1] - dst->begin[1], dst->end[2] - dst->begin[2]);
    a.flo = make_int3(fg->dom_lo[0], fg->dom_lo[1], fg->dom_lo[2]);
    a.fhi = make_int3(fg->dom_hi[0], fg->dom_hi[1], fg->dom_hi[2]);
    a.clo = make_int3(cg->dom_lo[0], cg->dom_lo[1], cg->dom_lo[2]);
    a.chi = make_int3(cg->dom_hi[0], cg->dom_hi[1], cg->dom_hi[2]);
    a.per = make_int3(fg->is_periodic[0], fg->is_periodic[1], dim == 3 ? fg->is_periodic[2] : 0);
    a.r = make_int3(ratio[0], ratio[1], dim == 3 ? ratio[2] : 1);
    if (dim == 2 && (dst->begin[2] != 0 || dst->end[2] != 1 || crse->begin[2] > 0 || crse->end[2] < 1)) {
        set_error("vd_fill_patch_two_levels: 2-D arrays must span k = 0 only");
        return VD_EINVAL;
    }
    const int64_t tot = (int64_t)a.dn.x * a.dn.y * a.dn.z;
    if (tot <= 0) return VD_OK;
    if (dim == 3)
        hipLaunchKernelGGL(k_fill_two_levels<3>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_fill_two_levels<2>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, a);
    VD_CHECK_LAUNCH();
    return VD_OK;
}
