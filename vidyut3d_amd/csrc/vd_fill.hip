// vd_fill.hip -- ghost fill (Vidyut::FillPatch, single level) and the
// small fill/copy kernels used by the other entry points (no hipMemset in
// the step path, so a captured step replays correctly).
#include "vd_internal.hpp"

namespace vd {

__global__ void k_fill_value(double* __restrict__ p, int64_t n, double v, Batch bt) {
    p = rb(p, case_off(bt));
    int64_t i = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_value_f(float* __restrict__ p, int64_t n, float v, Batch bt) {
    p = rb(p, case_off(bt));
    int64_t i = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// zeroes n bytes at p: the bytes before the first 16-byte boundary and
// after the last one one per lane, the aligned middle in 16-byte stores
// (the grid has at least 16 lanes, so every head and tail byte has one)
__global__ void k_fill_bytes(unsigned char* __restrict__ p, int64_t n, Batch bt) {
    p = rb(p, case_off(bt));
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
    hipLaunchKernelGGL(k_fill_value, bz(ctx, grid_stride(n, 256)), dim3(256), 0, ctx->stream, p, n, v, ctx->bt);
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

__global__ void __launch_bounds__(256) k_fill_batch(FillBatch b, Batch bt) {
    const unsigned blk = blockIdx.x;
    int q = 0;
    while (q + 1 < b.count && blk >= b.first[q + 1]) q++;
    const int64_t base = (int64_t)(blk - b.first[q]) * FILL_CHUNK;
    const int64_t end = base + FILL_CHUNK < b.n[q] ? base + FILL_CHUNK : b.n[q];
    double* __restrict__ p = rb(b.p[q], case_off(bt));
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
        hipLaunchKernelGGL(k_fill_batch, bz(ctx, dim3((unsigned)nb)), dim3(256), 0, ctx->stream, b, ctx->bt);
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

int fill_value_f(vd_ctx* ctx, float* p, int64_t n, float v) {
    if (n <= 0) return VD_OK;
    hipLaunchKernelGGL(k_fill_value_f, bz(ctx, grid_stride(n, 256)), dim3(256), 0, ctx->stream, p, n, v, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// FillPatchSingleLevel with foextrap / int_dir bcspec: every destination
// cell (valid or ghost) copies the source cell found by wrapping periodic
// indices and clamping the others into the domain.  One thread per
// destination cell, blockIdx.y = component.
__global__ void k_fill_patch(Arr<double> src, Arr<double> dst, int dim, int3 lo, int3 n,
                             int3 per, int nx, int ny, int nz, int scomp, Batch bt) {
    const int64_t co = case_off(bt);
    rebase(src, co);
    rebase(dst, co);
    int q = blockIdx.y + scomp;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)nx * ny * nz;
    if (t >= tot) return;
    int i, j, k;
    cell_of((unsigned)t, nx, ny, i, j, k);
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
                              int scomp, Batch bt) {
    rebase(a, case_off(bt));
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
        hipLaunchKernelGGL(k_fill_ghosts, bz(ctx, dim3((unsigned)((ng + 255) / 256), ncomp)), dim3(256), 0,
                           ctx->stream, arr(dst), g->dim, lo, n, per, nx, ny, nz, scomp, ctx->bt);
        kt_end(ctx, VD_KT_FILL, 16.0 * (double)ng * ncomp);
        VD_CHECK_LAUNCH();
        return VD_OK;
    }
    kt_begin(ctx, VD_KT_FILL);
    hipLaunchKernelGGL(k_fill_patch, bz(ctx, dim3((unsigned)((tot + 255) / 256), ncomp)), dim3(256), 0,
                       ctx->stream, arr(src), arr(dst), g->dim, lo, n, per, nx, ny, nz, scomp, ctx->bt);
    kt_end(ctx, VD_KT_FILL, 16.0 * (double)tot * ncomp);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// ---------------------------------------------- levels held in several boxes
// Vidyut::FillPatch at level 0 over a BoxArray chopped by amr.max_grid_size
// (Source/GridWork.cpp:112-124): AMReX fills a box's ghost cells from the
// valid cells of the other boxes (FillBoundary, periodic images included)
// and then the physical faces by foextrap, which together give every
// destination cell the value of the domain cell found by wrapping periodic
// indices and clamping the others -- the single-box rule, with the source
// cell read from the box that owns it.
struct BoxPair {
    Arr<double> src, dst;
    int3 glo;           // first cell written (the grown box for a fill, the box for gather/scatter)
    int gx, gy, gz;     // extents of the cells written
    unsigned blk0;      // first block of this box in the batched grid
};

// the box of the batched grid that block `blk` belongs to
__device__ __forceinline__ int box_of_block(const BoxPair* __restrict__ t, int nbox, unsigned blk) {
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].blk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct OwnerMap {
    const int* owner;   // box id per (granule) of the domain
    int3 lo, n, g, m;   // domain lo and extent, granule size, map extent
};

__global__ void __launch_bounds__(256) k_fill_patch_boxes(const BoxPair* __restrict__ tab, int nbox, OwnerMap om,
                                                         int dim, int3 per, int scomp) {
    const int q = blockIdx.y + scomp;
    const int b = box_of_block(tab, nbox, blockIdx.x);
    const BoxPair& B = tab[b];
    const int64_t t = (int64_t)(blockIdx.x - B.blk0) * 256 + threadIdx.x;
    if (t >= (int64_t)B.gx * B.gy * B.gz) return;
    const int i = B.glo.x + (int)(t % B.gx);
    const int64_t r = t / B.gx;
    const int j = B.glo.y + (int)(r % B.gy), k = B.glo.z + (int)(r / B.gy);
    int c[3] = {i, j, k};
    const int l[3] = {om.lo.x, om.lo.y, om.lo.z}, nn[3] = {om.n.x, om.n.y, om.n.z}, pp[3] = {per.x, per.y, per.z};
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
    const int gi = (c[0] - om.lo.x) / om.g.x, gj = dim > 1 ? (c[1] - om.lo.y) / om.g.y : 0,
              gk = dim > 2 ? (c[2] - om.lo.z) / om.g.z : 0;
    const Arr<double>& S = tab[om.owner[(gk * om.m.y + gj) * om.m.x + gi]].src;
    B.dst.comp(q)[B.dst.off(i, j, k)] = S.comp(q)[S.off(c[0], c[1], c[2])];
}

// valid cells of each part into the whole level array (gather) or back
template <bool GATHER>
__global__ void __launch_bounds__(256) k_level_copy(const BoxPair* __restrict__ tab, int nbox, int scomp) {
    const int q = blockIdx.y + scomp;
    const int b = box_of_block(tab, nbox, blockIdx.x);
    const BoxPair& B = tab[b];
    const int64_t t = (int64_t)(blockIdx.x - B.blk0) * 256 + threadIdx.x;
    if (t >= (int64_t)B.gx * B.gy * B.gz) return;
    const int i = B.glo.x + (int)(t % B.gx);
    const int64_t r = t / B.gx;
    const int j = B.glo.y + (int)(r % B.gy), k = B.glo.z + (int)(r / B.gy);
    // src = the part, dst = the whole level
    if (GATHER) B.dst.comp(q)[B.dst.off(i, j, k)] = B.src.comp(q)[B.src.off(i, j, k)];
    else B.src.comp(q)[B.src.off(i, j, k)] = B.dst.comp(q)[B.dst.off(i, j, k)];
}

static int gcd_i(int a, int b) {
    while (b) {
        const int t = a % b;
        a = b;
        b = t;
    }
    return a < 0 ? -a : a;
}

static bool covers(const vd_array4* a, const int lo[3], const int hi[3], int dim) {
    for (int d = 0; d < dim; d++)
        if (a->begin[d] > lo[d] || a->end[d] <= hi[d]) return false;
    return true;
}

// the boxes must lie in the domain and tile it without overlap: checked on
// an owner map at the granularity of the box corners (which the fill
// kernel uses to find the box holding a cell); the map is cached per
// context for the last box list
static int level_owner(vd_ctx* ctx, int nbox, const vd_box* boxes, const vd_geom* g, OwnerMap* om) {
    const int dim = g->dim;
    int lo[3] = {g->dom_lo[0], g->dom_lo[1], g->dom_lo[2]};
    int n[3] = {1, 1, 1}, gr[3] = {1, 1, 1}, m[3] = {1, 1, 1};
    for (int d = 0; d < dim; d++) {
        n[d] = g->dom_hi[d] - g->dom_lo[d] + 1;
        gr[d] = n[d];
    }
    for (int b = 0; b < nbox; b++)
        for (int d = 0; d < dim; d++) {
            if (boxes[b].lo[d] < g->dom_lo[d] || boxes[b].hi[d] > g->dom_hi[d] || boxes[b].hi[d] < boxes[b].lo[d]) {
                set_error("level boxes: box %d is empty or leaves the domain", b);
                return VD_EINVAL;
            }
            gr[d] = gcd_i(gr[d], boxes[b].lo[d] - lo[d]);
            gr[d] = gcd_i(gr[d], boxes[b].hi[d] + 1 - lo[d]);
        }
    int64_t cells = 1;
    for (int d = 0; d < dim; d++) {
        m[d] = n[d] / gr[d];
        cells *= m[d];
    }
    if (cells > (1 << 24)) {
        set_error("level boxes: corners on a %d x %d x %d grid need an owner map of %lld entries", gr[0], gr[1], gr[2],
                  (long long)cells);
        return VD_EINVAL;
    }
    std::vector<int> key;
    key.reserve(6 * nbox + 9);
    for (int d = 0; d < 3; d++) key.push_back(d < dim ? g->dom_lo[d] : 0), key.push_back(d < dim ? g->dom_hi[d] : 0);
    key.push_back(dim), key.push_back(nbox), key.push_back(0);
    for (int b = 0; b < nbox; b++)
        for (int d = 0; d < 3; d++) key.push_back(d < dim ? boxes[b].lo[d] : 0), key.push_back(d < dim ? boxes[b].hi[d] : 0);
    void* dmap = nullptr;
    VD_TRY(scratch(ctx, "level_owner", sizeof(int) * (size_t)cells, &dmap));
    if (ctx->owner_key != key) {
        std::vector<int> own((size_t)cells, -1);
        for (int b = 0; b < nbox; b++) {
            int a[3] = {0, 0, 0}, e[3] = {1, 1, 1};
            for (int d = 0; d < dim; d++) {
                a[d] = (boxes[b].lo[d] - lo[d]) / gr[d];
                e[d] = (boxes[b].hi[d] + 1 - lo[d]) / gr[d];
            }
            for (int k = a[2]; k < e[2]; k++)
                for (int j = a[1]; j < e[1]; j++)
                    for (int i = a[0]; i < e[0]; i++) {
                        int& o = own[((size_t)k * m[1] + j) * m[0] + i];
                        if (o >= 0) {
                            set_error("level boxes: boxes %d and %d overlap", o, b);
                            return VD_EINVAL;
                        }
                        o = b;
                    }
        }
        for (size_t c = 0; c < own.size(); c++)
            if (own[c] < 0) {
                set_error("level boxes: the boxes leave part of the domain uncovered");
                return VD_EINVAL;
            }
        VD_HIP(hipMemcpy(dmap, own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice));
        ctx->owner_key = key;
    }
    om->owner = (const int*)dmap;
    om->lo = make_int3(lo[0], lo[1], lo[2]);
    om->n = make_int3(n[0], n[1], n[2]);
    om->g = make_int3(gr[0], gr[1], gr[2]);
    om->m = make_int3(m[0], m[1], m[2]);
    return VD_OK;
}

// the per-box table on the device, uploaded on the context's stream from
// pageable memory (the call returns once the runtime has taken the data)
static int upload_pairs(vd_ctx* ctx, const std::vector<BoxPair>& tab, BoxPair** out) {
    void* p = nullptr;
    VD_TRY(scratch(ctx, "level_pairs", sizeof(BoxPair) * tab.size(), &p));
    VD_HIP(hipMemcpyAsync(p, tab.data(), sizeof(BoxPair) * tab.size(), hipMemcpyHostToDevice, ctx->stream));
    *out = (BoxPair*)p;
    return VD_OK;
}

int fill_patch_boxes(vd_ctx* ctx, int nbox, const vd_array4* src, vd_array4* dst, const vd_box* boxes,
                     const vd_geom* g, int scomp, int ncomp) {
    if (nbox < 1 || nbox > 65536) {
        set_error("fill_patch_boxes: %d boxes", nbox);
        return VD_EINVAL;
    }
    OwnerMap om;
    VD_TRY(level_owner(ctx, nbox, boxes, g, &om));
    if (ncomp <= 0) return VD_OK;
    std::vector<BoxPair> tab((size_t)nbox);
    unsigned blk = 0;
    for (int b = 0; b < nbox; b++) {
        if (!covers(&src[b], boxes[b].lo, boxes[b].hi, g->dim) || src[b].ncomp < scomp + ncomp ||
            dst[b].ncomp < scomp + ncomp) {
            set_error("fill_patch_boxes: array %d does not hold box %d or components [%d, %d)", b, b, scomp,
                      scomp + ncomp);
            return VD_EINVAL;
        }
        BoxPair& P = tab[b];
        P.src = arr(&src[b]);
        P.dst = arr(&dst[b]);
        P.glo = make_int3(dst[b].begin[0], dst[b].begin[1], dst[b].begin[2]);
        P.gx = dst[b].end[0] - dst[b].begin[0];
        P.gy = dst[b].end[1] - dst[b].begin[1];
        P.gz = dst[b].end[2] - dst[b].begin[2];
        P.blk0 = blk;
        const int64_t nb = ((int64_t)P.gx * P.gy * P.gz + 255) / 256;
        if ((int64_t)blk + nb > 0x7fffffffLL) {
            set_error("fill_patch_boxes: too many cells for one launch");
            return VD_EINVAL;
        }
        blk += (unsigned)nb;
    }
    BoxPair* dtab = nullptr;
    VD_TRY(upload_pairs(ctx, tab, &dtab));
    const int3 per = make_int3(g->is_periodic[0], g->dim > 1 ? g->is_periodic[1] : 0, g->dim > 2 ? g->is_periodic[2] : 0);
    hipLaunchKernelGGL(k_fill_patch_boxes, dim3(blk, ncomp), dim3(256), 0, ctx->stream, dtab, nbox, om, g->dim, per,
                       scomp);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

int level_copy(vd_ctx* ctx, bool gather, int nbox, const vd_array4* parts, const vd_box* boxes, const vd_array4* whole,
               const vd_geom* g, int scomp, int ncomp) {
    if (nbox < 1 || nbox > 65536) {
        set_error("level gather/scatter: %d boxes", nbox);
        return VD_EINVAL;
    }
    OwnerMap om;
    VD_TRY(level_owner(ctx, nbox, boxes, g, &om));
    if (ncomp <= 0) return VD_OK;
    const int dom_lo[3] = {g->dom_lo[0], g->dom_lo[1], g->dom_lo[2]};
    const int dom_hi[3] = {g->dom_hi[0], g->dom_hi[1], g->dom_hi[2]};
    if (!covers(whole, dom_lo, dom_hi, g->dim) || whole->ncomp < scomp + ncomp) {
        set_error("level gather/scatter: the whole-level array does not hold the domain");
        return VD_EINVAL;
    }
    std::vector<BoxPair> tab((size_t)nbox);
    unsigned blk = 0;
    for (int b = 0; b < nbox; b++) {
        if (!covers(&parts[b], boxes[b].lo, boxes[b].hi, g->dim) || parts[b].ncomp < scomp + ncomp) {
            set_error("level gather/scatter: array %d does not hold box %d", b, b);
            return VD_EINVAL;
        }
        BoxPair& P = tab[b];
        P.src = arr(&parts[b]);
        P.dst = arr(whole);
        P.glo = make_int3(boxes[b].lo[0], boxes[b].lo[1], boxes[b].lo[2]);
        P.gx = boxes[b].hi[0] - boxes[b].lo[0] + 1;
        P.gy = boxes[b].hi[1] - boxes[b].lo[1] + 1;
        P.gz = boxes[b].hi[2] - boxes[b].lo[2] + 1;
        P.blk0 = blk;
        blk += (unsigned)(((int64_t)P.gx * P.gy * P.gz + 255) / 256);
    }
    BoxPair* dtab = nullptr;
    VD_TRY(upload_pairs(ctx, tab, &dtab));
    if (gather)
        hipLaunchKernelGGL(k_level_copy<true>, dim3(blk, ncomp), dim3(256), 0, ctx->stream, dtab, nbox, scomp);
    else
        hipLaunchKernelGGL(k_level_copy<false>, dim3(blk, ncomp), dim3(256), 0, ctx->stream, dtab, nbox, scomp);
    VD_CHECK_LAUNCH();
    return VD_OK;
}

}  // namespace vd

extern "C" int vd_fill_patch_boxes(vd_ctx* ctx, int nbox, const vd_array4* src, vd_array4* dst,
                                   const vd_box* boxes, const vd_geom* geom, int scomp, int ncomp) {
    if (!ctx || !src || !dst || !boxes || !geom) return VD_EINVAL;
    for (int b = 0; b < nbox; b++) VD_TRY(vd::check_extents("fill_patch_boxes", &boxes[b], {&src[b], &dst[b]}));
    VD_TRY(vd::require_plain(ctx, "fill_patch_boxes"));
    return vd::fill_patch_boxes(ctx, nbox, src, dst, boxes, geom, scomp, ncomp);
}

extern "C" int vd_level_gather(vd_ctx* ctx, int nbox, const vd_array4* parts, const vd_box* boxes,
                               vd_array4* whole, const vd_geom* geom, int scomp, int ncomp) {
    if (!ctx || !parts || !boxes || !whole || !geom) return VD_EINVAL;
    VD_TRY(vd::check_extents("level gather/scatter", nullptr, {whole}));
    for (int b = 0; b < nbox; b++) VD_TRY(vd::check_extents("level gather/scatter", &boxes[b], {&parts[b]}));
    VD_TRY(vd::require_plain(ctx, "level gather/scatter"));
    return vd::level_copy(ctx, true, nbox, parts, boxes, whole, geom, scomp, ncomp);
}

extern "C" int vd_level_scatter(vd_ctx* ctx, int nbox, const vd_array4* whole, vd_array4* parts,
                                const vd_box* boxes, const vd_geom* geom, int scomp, int ncomp) {
    if (!ctx || !parts || !boxes || !whole || !geom) return VD_EINVAL;
    VD_TRY(vd::check_extents("level gather/scatter", nullptr, {whole}));
    for (int b = 0; b < nbox; b++) VD_TRY(vd::check_extents("level gather/scatter", &boxes[b], {&parts[b]}));
    VD_TRY(vd::require_plain(ctx, "level gather/scatter"));
    return vd::level_copy(ctx, false, nbox, parts, boxes, whole, geom, scomp, ncomp);
}

extern "C" int vd_fill_patch(vd_ctx* ctx, const vd_array4* src, vd_array4* dst, const vd_box* box,
                             const vd_geom* geom, int scomp, int ncomp) {
    if (!ctx || !src || !dst || !box || !geom) return VD_EINVAL;
    VD_TRY(vd::check_extents("fill_patch", box, {src, dst}));
    return vd::fill_patch(ctx, src, dst, box, geom, scomp, ncomp);
}

extern "C" int vd_memset_zero_kernel(vd_ctx* ctx, void* dst, size_t bytes) {
    if (!bytes) return VD_OK;
    hipLaunchKernelGGL(vd::k_fill_bytes, vd::bz(ctx, vd::grid_stride((int64_t)(bytes >> 4) + 1, 256)), dim3(256),
                       0, ctx->stream, (unsigned char*)dst, (int64_t)bytes, ctx->bt);
    VD_CHECK_LAUNCH();
    return VD_OK;
}
