// vd_mfops.hip -- the bulk MultiFab operations of the step
// (MultiFab::setVal / Copy / LinComb, Source/Evolve.cpp:105-171,290-300).
#include "vd_internal.hpp"

namespace vd {

struct BoxIter {
    int3 lo;
    int nx, ny, nz;
};

// cell (i, j, k) of the box for the linear index t (x fastest); boxes hold
// fewer than 2^31 cells (32-bit cell offsets), so 32-bit unsigned division
// suffices
__device__ __forceinline__ void box_cell(const BoxIter& it, unsigned t, int& i, int& j, int& k) {
    const unsigned q = t / (unsigned)it.nx;
    i = (int)(t - q * (unsigned)it.nx) + it.lo.x;
    j = (int)(q % (unsigned)it.ny) + it.lo.y;
    k = (int)(q / (unsigned)it.ny) + it.lo.z;
}

static int box_too_large(int64_t cells) {
    set_error("box of %lld cells: at most 2^31 - 1 per box", (long long)cells);
    return VD_EINVAL;
}

static BoxIter box_iter(const vd_box* b, const vd_array4* a) {
    BoxIter it;
    it.lo = make_int3(b->lo[0], b->lo[1], b->lo[2]);
    it.nx = b->hi[0] - b->lo[0] + 1;
    it.ny = b->hi[1] - b->lo[1] + 1;
    it.nz = b->hi[2] - b->lo[2] + 1;
    (void)a;
    return it;
}

__global__ void k_setval(Arr<double> a, BoxIter it, int scomp, double v, Batch bt) {
    rebase(a, case_off(bt));
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)it.nx * it.ny * it.nz) return;
    int i, j, k;
    box_cell(it, (unsigned)t, i, j, k);
    a(i, j, k, scomp + (int)blockIdx.y) = v;
}

// dst = a*x + b*y  (b == 0 and y == x gives a scaled copy)
__global__ void k_lincomb(Arr<double> dst, Arr<double> x, Arr<double> y, BoxIter it, int xc, int yc,
                          int dc, double a, double b, int copy, Batch bt) {
    const int64_t co = case_off(bt);
    rebase(dst, co);
    rebase(x, co);
    rebase(y, co);
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)it.nx * it.ny * it.nz) return;
    int i, j, k;
    box_cell(it, (unsigned)t, i, j, k);
    int n = (int)blockIdx.y;
    if (copy)
        dst(i, j, k, dc + n) = x(i, j, k, xc + n);
    else
        dst(i, j, k, dc + n) = a * x(i, j, k, xc + n) + b * y(i, j, k, yc + n);
}

}  // namespace vd

using namespace vd;

extern "C" int vd_setval(vd_ctx* ctx, vd_array4* a, const vd_box* box, int scomp, int ncomp, double v) {
    if (!ctx || !a || !box || ncomp < 0 || scomp < 0 || scomp + ncomp > a->ncomp) return VD_EINVAL;
    VD_TRY(check_extents("setval", box, {a}));
    if (!ncomp) return VD_OK;
    BoxIter it = box_iter(box, a);
    int64_t tot = (int64_t)it.nx * it.ny * it.nz;
    if (tot >= ((int64_t)1 << 31)) return box_too_large(tot);
    hipLaunchKernelGGL(k_setval, bz(ctx, dim3((unsigned)((tot + 255) / 256), ncomp)), dim3(256), 0, ctx->stream,
                       arr(a), it, scomp, v, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

extern "C" int vd_copy(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box, int scomp,
                       int dcomp, int ncomp) {
    if (!ctx || !src || !dst || !box || ncomp < 0) return VD_EINVAL;
    VD_TRY(check_extents("copy", box, {src, dst}));
    if (scomp + ncomp > src->ncomp || dcomp + ncomp > dst->ncomp) return VD_EINVAL;
    if (!ncomp) return VD_OK;
    BoxIter it = box_iter(box, dst);
    int64_t tot = (int64_t)it.nx * it.ny * it.nz;
    if (tot >= ((int64_t)1 << 31)) return box_too_large(tot);
    hipLaunchKernelGGL(k_lincomb, bz(ctx, dim3((unsigned)((tot + 255) / 256), ncomp)), dim3(256), 0, ctx->stream,
                       arr(dst), arr(src), arr(src), it, scomp, scomp, dcomp, 1.0, 0.0, 1, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

extern "C" int vd_lincomb(vd_ctx* ctx, vd_array4* dst, double a, const vd_array4* x, int xcomp, double b,
                          const vd_array4* y, int ycomp, int dcomp, int ncomp, const vd_box* box) {
    if (!ctx || !dst || !x || !y || !box || ncomp < 0) return VD_EINVAL;
    VD_TRY(check_extents("lincomb", box, {dst, x, y}));
    if (!ncomp) return VD_OK;
    BoxIter it = box_iter(box, dst);
    int64_t tot = (int64_t)it.nx * it.ny * it.nz;
    if (tot >= ((int64_t)1 << 31)) return box_too_large(tot);
    hipLaunchKernelGGL(k_lincomb, bz(ctx, dim3((unsigned)((tot + 255) / 256), ncomp)), dim3(256), 0, ctx->stream,
                       arr(dst), arr(x), arr(y), it, xcomp, ycomp, dcomp, a, b, 0, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

namespace vd {
__global__ void k_clamp_min(Arr<double> a, BoxIter it, int comp, double lo, Batch bt) {
    rebase(a, case_off(bt));
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)it.nx * it.ny * it.nz) return;
    int i, j, k;
    box_cell(it, (unsigned)t, i, j, k);
    double& v = a(i, j, k, comp);
    if (v < lo) v = lo;
}
}  // namespace vd

// bounded background-species update floor (Source/Evolve.cpp:278-283)
extern "C" int vd_clamp_min(vd_ctx* ctx, vd_array4* a, const vd_box* box, int comp, double lo) {
    if (!ctx || !a || !box || comp < 0 || comp >= a->ncomp) return VD_EINVAL;
    VD_TRY(check_extents("clamp_min", box, {a}));
    BoxIter it = box_iter(box, a);
    int64_t tot = (int64_t)it.nx * it.ny * it.nz;
    if (tot >= ((int64_t)1 << 31)) return box_too_large(tot);
    hipLaunchKernelGGL(k_clamp_min, bz(ctx, dim3((unsigned)((tot + 255) / 256))), dim3(256), 0, ctx->stream, arr(a),
                       it, comp, lo, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}
