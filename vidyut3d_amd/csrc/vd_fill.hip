// vd_fill.hip -- ghost fill (Vidyut::FillPatch, single level) and the
// small fill/copy kernels used by the other entry points (no hipMemset in
// the step path, so a captured step replays correctly).
#include "vd_internal.hpp"

namespace vd {

__global__ void k_fill_value(double* __restrict__ p, int64_t n, double v) {
    int64_t i = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_value_f(float* __restrict__ p, int64_t n, float v) {
    int64_t i = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// zeroes n bytes at p: the bytes before the first 16-byte boundary and
// after the last one one per lane, the aligned middle in 16-byte stores
// (the grid has at least 16 lanes, so every head and tail byte has one)
__global__ void k_fill_bytes(unsigned char* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)lblock() * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t head = (int64_t)((16u - ((uintptr_t)p & 15u)) & 15u);
    if (head > n) head = n;
    const int64_t nv = (n - head) >> 4;
    uint4* __restrict__ q = reinterpret_cast<uint4*>(p + head);
    for (int64_t v = i; v < nv; v += stride) q[v] = make_uint4(0u, 0u, 0u, 0u);
    const int64_t tail0 = head + (nv << 4);
    if (i < head) p[i] = 0;
    if (i < n - tail0) p[tail0 + i] = 0;
}

static inline dim3 grid_stride(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

int fill_value(vd_ctx* ctx, double* p, int64_t n, double v) {
    if (n <= 0) return VD_OK;
    hipLaunchKernelGGL(k_fill_value, grid_stride(n, 256), dim3(256), 0, ctx->stream, p, n, v);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// several constant fills in one launch: the blocks of fill q are
// [first[q], first[q+1]) of a 1-D grid, FILL_CHUNK elements per block, so
// the grid is as large as the data and no block is idle
constexpr int FILL_BATCH = 32;
constexpr int FILL_CHUNK = 256 * 8;
struct FillBatch {
    double* p[FILL_BATCH];
    int64_t n[FILL_BATCH];
    double v[FILL_BATCH];
    unsigned first[FILL_BATCH + 1];
    int count;
};

__global__ void __launch_bounds__(256) k_fill_batch(FillBatch b) {
    const unsigned blk = blockIdx.x;
    int q = 0;
    while (q + 1 < b.count && blk >= b.first[q + 1]) q++;
    const int64_t base = (int64_t)(blk - b.first[q]) * FILL_CHUNK;
    const int64_t end = base + FILL_CHUNK < b.n[q] ? base + FILL_CHUNK : b.n[q];
    double* __restrict__ p = b.p[q];
    const double v = b.v[q];
    for (int64_t i = base + threadIdx.x; i < end; i += 256) p[i] = v;
}

int fill_values(vd_ctx* ctx, double* const* p, const int64_t* n, const double* v, int count) {
    for (int base = 0; base < count; base += FILL_BATCH) {
        FillBatch b;
        const int k = count - base < FILL_BATCH ? count - base : FILL_BATCH;
        int64_t nb = 0;
        b.count = 0;
        for (int q = 0; q < k; q++) {
            if (n[base + q] <= 0) continue;
            const int c = b.count++;
            b.p[c] = p[base + q];
            b.n[c] = n[base + q];
            b.v[c] = v[base + q];
            b.first[c] = (unsigned)nb;
            nb += (n[base + q] + FILL_CHUNK - 1) / FILL_CHUNK;
        }
        if (nb == 0) continue;
        if (nb > 0xffffffffLL / 256) {  // grid x extent times block size
            set_error("fill_values: %lld blocks", (long long)nb);
            return VD_EINVAL;
        }
        b.first[b.count] = (unsigned)nb;
        hipLaunchKernelGGL(k_fill_batch, dim3((unsigned)nb), dim3(256), 0, ctx->stream, b);
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

int fill_value_f(vd_ctx* ctx, float* p, int64_t n, float v) {
    if (n <= 0) return VD_OK;
    hipLaunchKernelGGL(k_fill_value_f, grid_stride(n, 256), dim3(256), 0, ctx->stream, p, n, v);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// FillPatchSingleLevel with foextrap / int_dir bcspec: every destination
// cell (valid or ghost) copies the source cell found by wrapping periodic
// indices and clamping the others into the domain.  One thread per
// destination cell, blockIdx.y = component.
__global__ void k_fill_patch(Arr<double> src, Arr<double> dst, int dim, int3 lo, int3 n,
                             int3 per, int nx, int ny, int nz, int scomp) {
    int q = blockIdx.y + scomp;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)nx * ny * nz;
    if (t >= tot) return;
    int i = (int)(t % nx);
    int64_t r = t / nx;
    int j = (int)(r % ny);
    int k = (int)(r / ny);
    i += dst.bx;
    j += dst.by;
    k += dst.bz;
    int c[3] = {i, j, k};
    int l[3] = {lo.x, lo.y, lo.z}, nn[3] = {n.x, n.y, n.z}, pp[3] = {per.x, per.y, per.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= dim) break;
        int v = c[d] - l[d];
        if (v < 0 || v >= nn[d]) {
            if (pp[d]) {
                v %= nn[d];
                if (v < 0) v += nn[d];
            } else {
                v = v < 0 ? 0 : nn[d] - 1;
            }
        }
        c[d] = v + l[d];
    }
    dst.comp(q)[dst.off(i, j, k)] = src.comp(q)[src.off(c[0], c[1], c[2])];
}

// FillPatch of an array into itself: its valid region already holds the
// data, so only the ghost cells are written, each from the valid cell
// k_fill_patch maps it to.  Threads enumerate the ghost cells slab by slab
// (z ghost planes, then y ghost rows of the valid planes, then x ghost
// cells of the valid rows); blockIdx.y = component.
__global__ void k_fill_ghosts(Arr<double> a, int dim, int3 lo, int3 n, int3 per, int gx, int gy, int gz,
                              int scomp) {
    const int q = blockIdx.y + scomp;
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    const int gxl = lo.x - a.bx, gyl = lo.y - a.by, gzl = lo.z - a.bz;
    const int64_t cz = (int64_t)(gz - n.z) * gy * gx, cy = (int64_t)n.z * (gy - n.y) * gx,
                  cx = (int64_t)n.z * n.y * (gx - n.x);
    int i, j, k;
    if (t < cz) {
        const int kk = (int)(t / ((int64_t)gy * gx));
        const int rem = (int)(t - (int64_t)kk * gy * gx);
        k = kk < gzl ? a.bz + kk : lo.z + n.z + (kk - gzl);
        j = a.by + rem / gx;
        i = a.bx + rem % gx;
    } else if (t < cz + cy) {
        const int64_t u = t - cz;
        const int per_plane = (gy - n.y) * gx;
        k = lo.z + (int)(u / per_plane);
        const int rem = (int)(u % per_plane), jj = rem / gx;
        j = jj < gyl ? a.by + jj : lo.y + n.y + (jj - gyl);
        i = a.bx + rem % gx;
    } else if (t < cz + cy + cx) {
        const int64_t u = t - cz - cy;
        const int per_plane = n.y * (gx - n.x);
        k = lo.z + (int)(u / per_plane);
        const int rem = (int)(u % per_plane), ii = rem % (gx - n.x);
        j = lo.y + rem / (gx - n.x);
        i = ii < gxl ? a.bx + ii : lo.x + n.x + (ii - gxl);
    } else {
        return;
    }
    int c[3] = {i, j, k};
    const int l[3] = {lo.x, lo.y, lo.z}, nn[3] = {n.x, n.y, n.z}, pp[3] = {per.x, per.y, per.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= dim) break;
        int v = c[d] - l[d];
        if (v < 0 || v >= nn[d]) {
            if (pp[d]) {
                v %= nn[d];
                if (v < 0) v += nn[d];
            } else {
                v = v < 0 ? 0 : nn[d] - 1;
            }
        }
        c[d] = v + l[d];
    }
    a.comp(q)[a.off(i, j, k)] = a.comp(q)[a.off(c[0], c[1], c[2])];
}

int fill_patch(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box,
               const vd_geom* g, int scomp, int ncomp) {
    for (int d = 0; d < g->dim; d++)
        if (box->lo[d] != g->dom_lo[d] || box->hi[d] != g->dom_hi[d]) {
            set_error("fill_patch: box must cover the domain (single box per level)");
            return VD_EINVAL;
        }
    if (ncomp <= 0) return VD_OK;
    int nx = dst->end[0] - dst->begin[0], ny = dst->end[1] - dst->begin[1],
        nz = dst->end[2] - dst->begin[2];
    int64_t tot = (int64_t)nx * ny * nz;
    int3 lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    int3 n = make_int3(box->hi[0] - box->lo[0] + 1, box->hi[1] - box->lo[1] + 1,
                       box->hi[2] - box->lo[2] + 1);
    int3 per = make_int3(g->is_periodic[0], g->is_periodic[1], g->dim > 2 ? g->is_periodic[2] : 0);
    const bool same = src->p == dst->p && src->jstride == dst->jstride && src->kstride == dst->kstride &&
                      src->nstride == dst->nstride && src->begin[0] == dst->begin[0] &&
                      src->begin[1] == dst->begin[1] && src->begin[2] == dst->begin[2];
    if (same) {
        // into itself: the ghost cells only
        const int64_t nv = (int64_t)n.x * n.y * n.z;
        const int64_t ng = tot - nv;
        if (ng <= 0) return VD_OK;
        kt_begin(ctx, VD_KT_FILL);
        hipLaunchKernelGGL(k_fill_ghosts, dim3((unsigned)((ng + 255) / 256), ncomp), dim3(256), 0, ctx->stream,
                           arr(dst), g->dim, lo, n, per, nx, ny, nz, scomp);
        kt_end(ctx, VD_KT_FILL, 16.0 * (double)ng * ncomp);
        VD_CHECK_LAUNCH();
        return VD_OK;
    }
    kt_begin(ctx, VD_KT_FILL);
    hipLaunchKernelGGL(k_fill_patch, dim3((unsigned)((tot + 255) / 256), ncomp), dim3(256), 0,
                       ctx->stream, arr(src), arr(dst), g->dim, lo, n, per, nx, ny, nz, scomp);
    kt_end(ctx, VD_KT_FILL, 16.0 * (double)tot * ncomp);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

}  // namespace vd

extern "C" int vd_fill_patch(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box,
                             const vd_geom* geom, int scomp, int ncomp) {
    if (!ctx || !src || !dst || !box || !geom) return VD_EINVAL;
    return vd::fill_patch(ctx, src, dst, box, geom, scomp, ncomp);
}

extern "C" int vd_memset_zero_kernel(vd_ctx* ctx, void* dst, size_t bytes) {
    if (!bytes) return VD_OK;
    hipLaunchKernelGGL(vd::k_fill_bytes, vd::grid_stride((int64_t)(bytes >> 4) + 1, 256), dim3(256), 0,
                       ctx->stream, (unsigned char*)dst, (int64_t)bytes);
    VD_CHECK_LAUNCH();
    return VD_OK;
}
