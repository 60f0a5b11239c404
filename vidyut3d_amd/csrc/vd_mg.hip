// vd_mg.hip -- geometric multigrid for the cell-centred ABecLaplacian
//   alpha_a*phi - beta*div(b grad phi) = f
// used by the potential solve (Source/PotentialSolve.cpp:148-336) and the
// implicit species solves (Source/ScalarSolve.cpp:420-596).
//
// Conventions (identical in the CPU oracle):
//  * stencil: dg = alpha_a + sum_d s_d*dd_d, of = sum_d s_d*oo_d,
//    s_d = beta/dx_d^2, (L x)_c = dg*x_c - of; boundary neighbours are
//    ghosts c0*x_c (+cinh on the finest level, folded into the rhs)
//  * Dirichlet (maxorder 2) c0=-1 cinh=2v; Neumann c0=1; inhomogeneous
//    Neumann c0=1 cinh=-/+dx*g; Robin c0=-(a/2-b/dx)/(a/2+b/dx),
//    cinh=f/(a/2+b/dx)
//  * correction-form V-cycle: red-black Gauss-Seidel (colour (i+j+k)&1,
//    red first), residual fused with cell-average restriction, linear
//    prolongation, face b and boundary c0 averaged to the coarse level,
//    CG bottom solve in one workgroup
//  * stop when ||r||_inf <= max(abstol, reltol*max(||b||_inf, ||r0||_inf)),
//    b = rhs with the boundary data folded in
//  * precision 1: the V-cycle (smoothing, restriction, prolongation)
//    runs in FP32 on FP32 copies of the hierarchy; the finest-level
//    residual, the update x += e, the stopping test and the bottom solve
//    (CG on the bottom operator coarsened in FP64) stay FP64 (iterative
//    refinement).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "vd_internal.hpp"

namespace vd {

int fill_value(vd_ctx* ctx, double* p, int64_t n, double v);
int fill_values(vd_ctx* ctx, double* const* p, const int64_t* n, const double* v, int count);
int fill_value_f(vd_ctx* ctx, float* p, int64_t n, float v);

// ------------------------------------------------------------ data
template <class T>
struct LevV {  // kernel view of one level
    int nx, ny, nz;
    T s0, s1, s2;        // beta/dx^2
    T alpha, bconst;
    const T* bf0;
    const T* bf1;
    const T* bf2;
    const T* dcell;      // BM 2: cell coefficient on the grown box
    const T* c0[3][2];   // per boundary face c0 (non-periodic sides)
    T c0u[3][2];         // c0 of sides whose c0 is uniform (no Robin side)
    int c0uni;           // 1: c0u holds every boundary c0 (no c0 loads)
    int per0, per1, per2;
    // the coefficient arrays moved by o bytes: only those the coefficient
    // mode BM reads (0 faces, 2 the cell coefficient; constant b none) and
    // the boundary c0 arrays when C0 (they are not read when every c0 is
    // uniform), so that no other pointer takes registers
    template <int BM, bool C0>
    __device__ __forceinline__ void rebase_used(int64_t o) {
        if constexpr (BM == 0) {
            bf0 = vd::rb(bf0, o);
            bf1 = vd::rb(bf1, o);
            bf2 = vd::rb(bf2, o);
        }
        if constexpr (BM == 2) dcell = vd::rb(dcell, o);
        if constexpr (C0) {
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int s = 0; s < 2; s++) c0[d][s] = vd::rb(c0[d][s], o);
        }
    }
    __device__ __forceinline__ void rebase(int64_t o) {
        bf0 = vd::rb(bf0, o);
        bf1 = vd::rb(bf1, o);
        bf2 = vd::rb(bf2, o);
        dcell = vd::rb(dcell, o);
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int s = 0; s < 2; s++) c0[d][s] = vd::rb(c0[d][s], o);
    }
};

template <class T>
struct Lev {
    int n[3];
    double dx[3];
    T* bf[3] = {nullptr, nullptr, nullptr};
    T* c0[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    T* e = nullptr;
    T* e2 = nullptr;  // second correction buffer of the fused smoother
    int fzc = 0;      // its plane chunk on this level (0: per-colour sweeps)
    T* r = nullptr;
    T* p = nullptr;  // CG scratch (bottom level)
    T* q = nullptr;
    T* w = nullptr;
};

// The implicit species solve's copy-back (ScalarSolve.cpp:632-647: the
// solution into the state's component, with the density floor), written by
// a sweep-plus-norm pass for every iterate it produces, so that the state
// holds the final iterate when the iterations stop and no copy pass
// follows.  p: the component at the level's cell (0,0,0); bytes: its
// extent from p (below 2^31, the pass's byte offsets).
struct SweepOut {
    double* p;
    int js, ks;
    unsigned bytes;
    int floor_on;
    double floor_val;
};

struct MGWork {
    vd_ctx* ctx = nullptr;  // the context whose memory holds the arrays
    int dim = 3, nlev = 0, prec = 0, constb = 0;
    int n[3] = {1, 1, 1};
    int per[3] = {0, 0, 0};
    double dx[3] = {1, 1, 1};
    std::vector<void*> allocs;
    // FP64 finest-level data
    double* x = nullptr;
    double* x2 = nullptr;  // second solution buffer of the fused smoother
    double* f = nullptr;
    double* res = nullptr;
    double* bf0[3] = {nullptr, nullptr, nullptr};
    double* c00[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    double* bcv[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    // the inhomogeneous ghost part of each side's faces, when the rhs pass
    // adds the boundary terms (beta/dx^2 * b * it) itself
    double* sh[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    // the initial residual and rhs norms are already in slot groups 0 and 1
    bool r0_done = false;
    double* rob[3][2][3];  // robin a, b, f per side
    unsigned long long* dnorm = nullptr;  // 8 groups of NSLOT norm slots
    double* dcell = nullptr;              // cell coefficient on the grown box
    double* part = nullptr;               // per-block partial norms
    int c0_types[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
    unsigned c0_mask = 0;                 // cases whose coarse c0 match c0_types
    // no Robin side: c0 is -1 (Dirichlet) or 1 (Neumann) on every boundary
    // face of every level, and the fused smoother takes it as a constant
    int c0uni = 0;
    // z-streamed sweeps, residual, restriction and prolongation on 3-D
    // levels (VD_STREAM_SWEEPS=0 at workspace creation: row-per-block)
    int stream = 1;
    double c0u[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    std::vector<Lev<double>> L64;
    std::vector<Lev<float>> L32;
    // FP32 V-cycles: the face and boundary coefficients of levels 1 ..
    // nlev-1 coarsened in FP64 (level 0 is bf0 / c00), and FP64 CG arrays
    // of the bottom level, for the FP64 bottom solve
    std::vector<Lev<double>> C64;
    // batched contexts: each case's copy of the coarse tail's arguments
    // (k_tail_args), the arguments they were made from, and the cases
    // whose copies are current
    mutable void* tail_dev = nullptr;
    mutable std::vector<unsigned char> tail_key;
    mutable unsigned tail_mask = 0;
    // an implicit species solve's copy-back, written by its fused sweeps
    // (post.p null: none), and the cases whose final iterate they wrote
    SweepOut post{};
    unsigned post_mask = 0;
};

void mgwork_free(MGWork* w) {
    if (!w) return;
    for (void* p : w->allocs) dev_free(w->ctx, p);
    delete w;
}

static int dalloc(MGWork* w, size_t bytes, void** out) {
    if (dev_alloc(w->ctx, bytes < 64 ? 64 : bytes, out) != VD_OK) return VD_ENOMEM;
    w->allocs.push_back(*out);
    return VD_OK;
}

static inline int64_t ncell(const int n[3]) { return (int64_t)n[0] * n[1] * n[2]; }
static inline int64_t nfacef(const int n[3], int d) {
    return (int64_t)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
}
static inline int64_t ntanf(const int n[3], int d) {
    return d == 0 ? (int64_t)n[1] * n[2] : (d == 1 ? (int64_t)n[0] * n[2] : (int64_t)n[0] * n[1]);
}

// ------------------------------------------------------------ stencil
// BM (coefficient mode): 0 face arrays, 1 constant b, 2 harmonic face
// value of a cell coefficient on the grown box (finest level of the
// implicit species solves: 8 B/cell read instead of 24 B/cell of faces)
template <class T>
__device__ __forceinline__ int fidx(const LevV<T>& L, int d, int i, int j, int k) {
    int n0 = L.nx + (d == 0), n1 = L.ny + (d == 1);
    return i + n0 * (j + n1 * k);
}
template <class T>
__device__ __forceinline__ int tidx(const LevV<T>& L, int d, int i, int j, int k) {
    return d == 0 ? j + L.ny * k : (d == 1 ? i + L.nx * k : i + L.nx * j);
}

// harmonic face average 2ab/(a+b) (ScalarSolve.cpp:511-512), 0 when
// a+b = 0; one formula for the face arrays and the on-the-fly faces
template <class T>
__device__ __forceinline__ T harm(T a, T b) {
    const T s = a + b;
    return (s == T(0)) ? T(0) : fdiv(T(2) * a * b, s);
}

// face coefficients of cell (i,j,k) in direction d
template <class T, int BM, int DIM>
__device__ __forceinline__ void face_coef(const LevV<T>& L, int d, int i, int j, int k, T& blo, T& bhi) {
    if (BM == 1) {
        blo = bhi = L.bconst;
    } else if (BM == 0) {
        const T* bf = d == 0 ? L.bf0 : (d == 1 ? L.bf1 : L.bf2);
        const int f = fidx(L, d, i, j, k);
        const int fs = d == 0 ? 1 : (d == 1 ? (L.nx) : (L.nx * L.ny));
        blo = bf[f];
        bhi = bf[f + fs];
    } else {
        const int gx = L.nx + 2, gxy = gx * (L.ny + 2);
        const int gc = (i + 1) + gx * (j + 1) + (DIM == 3 ? gxy * (k + 1) : 0);
        const int gs = d == 0 ? 1 : (d == 1 ? gx : gxy);
        const T dc = L.dcell[gc];
        blo = harm(L.dcell[gc - gs], dc);
        bhi = harm(dc, L.dcell[gc + gs]);
    }
}

// dg and of of cell (i,j,k); nb(d, side) returns the neighbour value on
// side 0 (low) / 1 (high) of direction d when that neighbour is a cell of
// the level (interior or periodic image); fc(d, blo, bhi) gives the face
// coefficients of direction d
template <class T, int DIM, class NB, class FC, bool C0U = false>
__device__ __forceinline__ void stencil_nbf(const LevV<T>& L, int i, int j, int k, NB nb, FC fc, T& dgo,
                                            T& ofo) {
    T dg = L.alpha, of = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int cd = d == 0 ? i : (d == 1 ? j : k);
        const int nd = d == 0 ? L.nx : (d == 1 ? L.ny : L.nz);
        const int per = d == 0 ? L.per0 : (d == 1 ? L.per1 : L.per2);
        const T s = d == 0 ? L.s0 : (d == 1 ? L.s1 : L.s2);
        T blo, bhi;
        fc(d, blo, bhi);
        T dd = T(0), oo = T(0);
        if (cd > 0 || per) {
            dd += blo;
            oo += blo * nb(d, 0);
        } else {
            dd += blo * (T(1) - (C0U ? L.c0u[d][0] : L.c0[d][0][tidx(L, d, i, j, k)]));
        }
        if (cd < nd - 1 || per) {
            dd += bhi;
            oo += bhi * nb(d, 1);
        } else {
            dd += bhi * (T(1) - (C0U ? L.c0u[d][1] : L.c0[d][1][tidx(L, d, i, j, k)]));
        }
        dg += s * dd;
        of += s * oo;
    }
    dgo = dg;
    ofo = of;
}

template <class T, int BM, int DIM, class NB, bool C0U = false>
__device__ __forceinline__ void stencil_nb(const LevV<T>& L, int i, int j, int k, NB nb, T& dgo, T& ofo) {
    auto fc = [&](int d, T& blo, T& bhi) { face_coef<T, BM, DIM>(L, d, i, j, k, blo, bhi); };
    stencil_nbf<T, DIM, NB, decltype(fc), C0U>(L, i, j, k, nb, fc, dgo, ofo);
}

// dg and of of cell (i,j,k) from x in memory.  Every load is issued
// unconditionally (neighbour index clamped to the cell itself on a
// non-periodic boundary, where the value is not used, wrapped on a periodic
// one) and the boundary terms are selected afterwards: loads under
// divergent branches would be waited for branch by branch.  The arithmetic
// is that of stencil_nb term for term.
template <class T, int BM, int DIM, bool FM = false>
__device__ __forceinline__ void stencil(const LevV<T>& L, const T* __restrict__ x, int i, int j,
                                        int k, T& dgo, T& ofo, T* fmax_d = nullptr) {
    const int c = i + L.nx * (j + L.ny * k);
    T xl[DIM], xh[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int cd = d == 0 ? i : (d == 1 ? j : k);
        const int nd = d == 0 ? L.nx : (d == 1 ? L.ny : L.nz);
        const int st = d == 0 ? 1 : (d == 1 ? L.nx : L.nx * L.ny);
        const int per = d == 0 ? L.per0 : (d == 1 ? L.per1 : L.per2);
        const int il = cd > 0 ? c - st : (per ? c + (nd - 1) * st : c);
        const int ih = cd < nd - 1 ? c + st : (per ? c - (nd - 1) * st : c);
        xl[d] = x[il];
        xh[d] = x[ih];
    }
    T dg = L.alpha, of = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int cd = d == 0 ? i : (d == 1 ? j : k);
        const int nd = d == 0 ? L.nx : (d == 1 ? L.ny : L.nz);
        const int per = d == 0 ? L.per0 : (d == 1 ? L.per1 : L.per2);
        const T s = d == 0 ? L.s0 : (d == 1 ? L.s1 : L.s2);
        T blo, bhi;
        face_coef<T, BM, DIM>(L, d, i, j, k, blo, bhi);
        if (FM) fmax_d[d] = fmax(fabs(blo), fabs(bhi));
        T c0l = T(1), c0h = T(1);
        if (!per) {
            if (L.c0uni) {
                c0l = L.c0u[d][0];
                c0h = L.c0u[d][1];
            } else {
                const int t = tidx(L, d, i, j, k);
                c0l = L.c0[d][0][t];
                c0h = L.c0[d][1][t];
            }
        }
        const bool lin = cd > 0 || per, hin = cd < nd - 1 || per;
        T dd = T(0), oo = T(0);
        dd = lin ? dd + blo : dd + blo * (T(1) - c0l);
        oo = lin ? oo + blo * xl[d] : oo;
        dd = hin ? dd + bhi : dd + bhi * (T(1) - c0h);
        oo = hin ? oo + bhi * xh[d] : oo;
        dg += s * dd;
        of += s * oo;
    }
    dgo = dg;
    ofo = of;
}

// ------------------------------------------------------------ kernels
// one colour of a red-black sweep: 1-D grid of gx blocks per row, rows
// ordered (j, k) and XCD-remapped
template <class T, int BM, int DIM>
__global__ void __launch_bounds__(128) k_gsrb(LevV<T> L, T* __restrict__ x, const T* __restrict__ rhs,
                                              int color, int gx, Batch bt) {
    const int64_t co = case_off(bt);
    L.template rebase_used<BM, true>(co);
    x = rb(x, co);
    rhs = rb(rhs, co);
    const unsigned lb = lblock();
    const unsigned row = lb / gx;
    const int j = (int)(row % L.ny), k = (int)(row / L.ny);
    const int ii = (int)(lb % gx) * blockDim.x + threadIdx.x;
    const int i = 2 * ii + ((color + j + k) & 1);
    if (i >= L.nx) return;
    T dg, of;
    stencil<T, BM, DIM>(L, x, i, j, k, dg, of);
    const int c = i + L.nx * (j + L.ny * k);
    x[c] = (rhs[c] + of) / dg;
}

// Fused red-black smoother for 3-D levels whose z direction is not
// periodic.  One launch applies S colour sweeps (S = 2: one red + black
// sweep, S = 4: two), optionally followed (RES) by the max-norm of the
// residual of the result, in a single pass over HBM instead of S launches
// that each re-read the level.
//  * A block owns an interior tile of FZ_IX x FZ_IY cells and zc planes;
//    it loads a halo of FZ_H cells / S+RES planes around it and streams up
//    the planes.  At pipeline step t, stage s updates plane t-s+1, so every
//    stage sees its z-neighbours after stage s-1 and before stage s+1 --
//    the values S separate colour launches would give, bit for bit.
//  * Each thread owns a red/black pair of x-adjacent cells of one row and
//    keeps its column of planes (and their rhs) in registers; in-plane
//    neighbours of the opposite colour go through LDS, one buffer per
//    stage, published once per step.
//  * Halo values go wrong by one cell per stage from the tile edge inward;
//    the halo is as wide as the stage count, so the interior is exact.
//  * Reads xin (zeros when xin is null), writes xout.  Halos overlap the
//    neighbouring blocks' interiors, so xin and xout must differ.
// Requires even nx, ny (colour parity of periodic images).
constexpr int FZ_IX = 64, FZ_IY = 16, FZ_H = 4;
constexpr int FZ_PX = (FZ_IX + 2 * FZ_H) / 2;  // 36 thread columns, 2 cells each
constexpr int FZ_RY = FZ_IY + 2 * FZ_H;        // 24 thread rows
constexpr int FZ_THREADS = FZ_PX * FZ_RY;      // 864
// the tile by coefficient mode: the cell coefficient (BM 2) keeps three
// more plane rings of face values in registers and takes a 32 x 16 tile
// with a 3-row y halo, enough for its one sweep and residual (440
// threads: two waves per SIMD, so up to 256 VGPRs each)
template <int BM>
struct FzTile {
    static constexpr int IX = BM == 2 ? 32 : FZ_IX;
    static constexpr int HY = BM == 2 ? 3 : FZ_H;
    static constexpr int PX = (IX + 2 * FZ_H) / 2;
    static constexpr int RY = FZ_IY + 2 * HY;
    static constexpr int THREADS = PX * RY;
};

// raw buffer resource over `bytes` of memory (gfx9 family config word);
// accesses at offsets >= bytes return 0 (loads) or are dropped (stores)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <class T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned off, T v);
template <>
__device__ __forceinline__ void buf_st<double>(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
    using U = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, (int)off, 0, 0);
}
template <>
__device__ __forceinline__ void buf_st<float>(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    using U = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(U, v), r, (int)off, 0, 0);
}

// f(integral_constant<K>, t + K) for K = K0 .. N-1 while t + K < tend
template <int K, int N, class F>
__device__ __forceinline__ void fused_unroll(F& f, int t, int tend) {
    if constexpr (K < N) {
        if (t + K < tend) {
            f(std::integral_constant<int, K>{}, t + K);
            fused_unroll<K + 1, N>(f, t, tend);
        }
    }
}

// ACC: instead of storing the smoothed x, add it to the FP64 array xacc
// (the correction of the last post-smooth of a V-cycle, x += e)
// BT: a batched launch (each z-block moves the arrays by its case's
// offset); unbatched launches take the BT = false kernel, whose registers
// the offsets do not take
// POST (FP64, BM 1, one sweep and its norm): the iterate also written to
// the state as SweepOut describes
template <class T, int BM, int S, int RES, int ZI, int ACC = 0, bool BT = false, int POST = 0>
__global__ void __launch_bounds__(FzTile<BM>::THREADS) k_gsrb_fused(LevV<T> L, const T* __restrict__ xin,
                                                           T* __restrict__ xout, const T* __restrict__ rhs,
                                                           int zc, int ntx, int nty,
                                                           double* __restrict__ part,
                                                           double* __restrict__ xacc, SweepOut so, Batch bt) {
    static_assert(!POST || (std::is_same<T, double>::value && BM == 1 && S == 2 && RES == 1 && !ZI && !ACC),
                  "POST: the FP64 constant-coefficient sweep and its norm");
    if constexpr (BT) {
        const int64_t co = case_off(bt);
        L.template rebase_used<BM, false>(co);
        xin = rb(xin, co);
        xout = rb(xout, co);
        rhs = rb(rhs, co);
        part = rb(part, co);
        xacc = rb(xacc, co);
        if constexpr (POST) so.p = rb(so.p, co);
    }
    // Register ring indexed by plane: plane p lives in slot (p - zlo) mod NV
    // (NV even), and the step loop is unrolled NV times, so every slot index
    // is a compile-time constant and no values move between registers.
    // Stage s of step t updates plane t-s+1 (slot k-s of unrolled step k);
    // values are kept by colour (red = (i+j+k) even) so that the colour a
    // stage updates is a compile-time constant too.
    constexpr int NV = ((S + 2 + RES) + 1) & ~1;
    constexpr int TZ_IX = FzTile<BM>::IX, TZ_PX = FzTile<BM>::PX, TZ_THREADS = FzTile<BM>::THREADS;
    constexpr int TZ_RY = FzTile<BM>::RY, TZ_HY = FzTile<BM>::HY;
    __shared__ T nbuf[S][TZ_RY][TZ_PX];
    // BM 1: reciprocal diagonal of each cell of the thread's pair by the
    // z class of its plane (bottom wall, interior, top wall), kept in LDS
    // rather than registers
    __shared__ T rdbuf[BM == 1 ? 3 : 1][2][BM == 1 ? TZ_THREADS : 1];
    // BM 2 (cell coefficient): the harmonic faces of a plane are formed once,
    // as the plane arrives, instead of six per stencil evaluation.  The low-x
    // face of cell 0 and the low-y faces of both cells go to LDS, where the
    // pair on the left and the row below read them as their high faces; the
    // face between the pair's cells and the low-z faces stay in registers.
    // Four planes are in use at a step (the arriving one and the three its
    // stages and residual read), ring-indexed by p & 3.
    static_assert(BM != 2 || (S == 2 && RES == 1 && !ZI && !ACC), "BM 2: one sweep and its residual");
    constexpr int FR = BM == 2 ? 4 : 1, FRY = BM == 2 ? TZ_RY : 1, FPX = BM == 2 ? TZ_PX : 1;
    __shared__ T fxs[FR][FRY][FPX];
    __shared__ T fys[FR][2][FRY][FPX];
    __shared__ T fbuf[RES ? 2 : 1][TZ_RY][TZ_PX];
    __shared__ double red[(TZ_THREADS + 63) / 64];

    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int x0 = (tile % ntx) * TZ_IX, y0 = (tile / ntx) * FZ_IY, z0 = zb * zc;
    const int tx = (int)threadIdx.x % TZ_PX, ty = (int)threadIdx.x / TZ_PX;
    // cells ia, ia+1 of row ja; global (periodically wrapped) indices
    const int ia = x0 - FZ_H + 2 * tx, ja = y0 - TZ_HY + ty;
    int ic0 = ia, ic1 = ia + 1, jc = ja;
    bool ok0, ok1, okj;
    if (L.per0) {
        ic0 %= L.nx;
        if (ic0 < 0) ic0 += L.nx;
        ic1 %= L.nx;
        if (ic1 < 0) ic1 += L.nx;
        ok0 = ok1 = true;
    } else {
        ok0 = ic0 >= 0 && ic0 < L.nx;
        ok1 = ic1 >= 0 && ic1 < L.nx;
    }
    if (L.per1) {
        jc %= L.ny;
        if (jc < 0) jc += L.ny;
        okj = true;
    } else {
        okj = jc >= 0 && jc < L.ny;
    }
    // interior cells of this tile that lie in the level (written / normed)
    const bool orow = tx >= FZ_H / 2 && tx < TZ_PX - FZ_H / 2 && ty >= TZ_HY && ty < TZ_RY - TZ_HY &&
                      ja < L.ny;
    const bool oc0 = orow && ia < L.nx, oc1 = orow && ia + 1 < L.nx;

    const int zlo = max(z0 - S - RES, 0), zhi = min(z0 + zc - 1 + S + RES, L.nz - 1);
    const int zout = min(z0 + zc, L.nz);
    const int pstr = L.nx * L.ny;
    const int par = ia + ja;  // colour of cell ia on plane p is (par + p) & 1
    // clamped indices: every lane evaluates the stencil (results of lanes
    // outside the level are discarded), so no memory access depends on a
    // branch and the compiler can count loads in flight exactly
    const int ic0c = min(max(ic0, 0), L.nx - 1), ic1c = min(max(ic1, 0), L.nx - 1);
    const int jcc = min(max(jc, 0), L.ny - 1);
    const int pc0 = ic0c + L.nx * jcc, pc1 = ic1c + L.nx * jcc;
    // range-checked buffer accesses: an offset with bit 31 set is outside
    // the buffer, so the load returns 0 and the store is dropped
    constexpr unsigned OOB = 0x80000000u;
    const unsigned lbytes = (unsigned)((int64_t)pstr * L.nz * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t rx = buf_rsrc(xin, lbytes), rf = buf_rsrc(rhs, lbytes),
                                 ro = buf_rsrc(xout, lbytes);
    auto offs = [&](int pc, bool ok, int p) -> unsigned {
        return ok ? (unsigned)(pc + pstr * p) * (unsigned)sizeof(T) : OOB;
    };
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(xacc, (unsigned)((int64_t)pstr * L.nz * 8));
    auto offd = [&](int pc, bool ok, int p) -> unsigned { return ok ? (unsigned)(pc + pstr * p) * 8u : OOB; };
    // POST: the state component (cells written are inside the level, so
    // their unwrapped and wrapped indices agree)
    const __amdgpu_buffer_rsrc_t rpo = buf_rsrc(POST ? (const void*)so.p : nullptr, POST ? so.bytes : 0u);
    const int po0 = POST ? ic0c + so.js * jcc : 0;  // cell 1 of a written pair is at po0 + 1

    T vr[NV], vb[NV], fr[NV], fb[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) vr[q] = vb[q] = fr[q] = fb[q] = T(0);
    // BM 2: the cell coefficient lives on the level grown by one cell.  A
    // lane reads it at its two cells, the cell left of cell 0 and the two
    // cells below; columns, rows and planes are clamped to the grown box
    // (wrapped on a periodic side, where the grown box's ghost cells hold
    // the periodic images), so the faces formed for the lanes just outside
    // a wall are the wall faces the cells inside it take as high faces.
    // By plane slot: fx1 the face between the pair's cells, fzr / fzb the
    // low-z faces of the red / black cell (the high-z face of a cell is the
    // low-z face of the opposite colour on the plane above), and ofb / dgb
    // the off-diagonal sum and diagonal of the black update, which are those
    // of the black cell's residual bit for bit (same neighbour values, same
    // arithmetic) and are not formed twice.
    T fx1[BM == 2 ? NV : 1], fzr[BM == 2 ? NV : 1], fzb[BM == 2 ? NV : 1];
    T ofb[BM == 2 ? NV : 1], dgb[BM == 2 ? NV : 1];
    int gA = 0, gB = 0, gL = 0, gAb = 0, gBb = 0;
    const int ggx = L.nx + 2, ggxy = ggx * (L.ny + 2);
    const __amdgpu_buffer_rsrc_t rdc =
        buf_rsrc(L.dcell, BM == 2 ? (unsigned)((int64_t)ggxy * (L.nz + 2) * (int64_t)sizeof(T)) : 0u);
    if constexpr (BM == 2) {
#pragma unroll
        for (int q = 0; q < NV; q++) fx1[q] = fzr[q] = fzb[q] = ofb[q] = dgb[q] = T(0);
        int c0, c1, cl, r0, rb;
        if (L.per0) {
            c0 = ic0;
            c1 = ic1;
            cl = ic0 - 1;
        } else {
            c0 = min(max(ia, -1), L.nx);
            c1 = min(max(ia + 1, -1), L.nx);
            cl = min(max(ia - 1, -1), L.nx);
        }
        if (L.per1) {
            r0 = jc;
            rb = jc - 1;
        } else {
            r0 = min(max(ja, -1), L.ny);
            rb = min(max(ja - 1, -1), L.ny);
        }
        gA = (c0 + 1) + ggx * (r0 + 1);
        gB = (c1 + 1) + ggx * (r0 + 1);
        gL = (cl + 1) + ggx * (r0 + 1);
        gAb = (c0 + 1) + ggx * (rb + 1);
        gBb = (c1 + 1) + ggx * (rb + 1);
    }
    struct PD {
        T a, b, l, ab, bb;  // cell 0, cell 1, left of cell 0, below cell 0, below cell 1
    };
    auto load_d = [&](PD& d, int p) {
        const unsigned o = (unsigned)(ggxy * (min(max(p, -1), L.nz) + 1));
        d.a = buf_ld<T>(rdc, (o + (unsigned)gA) * (unsigned)sizeof(T));
        d.b = buf_ld<T>(rdc, (o + (unsigned)gB) * (unsigned)sizeof(T));
        d.l = buf_ld<T>(rdc, (o + (unsigned)gL) * (unsigned)sizeof(T));
        d.ab = buf_ld<T>(rdc, (o + (unsigned)gAb) * (unsigned)sizeof(T));
        d.bb = buf_ld<T>(rdc, (o + (unsigned)gBb) * (unsigned)sizeof(T));
    };
    PD pd{};
    T dpa = T(0), dpb = T(0);  // the coefficient of the thread's cells on the previous plane
    // planes are prefetched two steps ahead into two alternating buffers
    // (no register moves of in-flight loads), issued at the end of a step
    // after that step's own loads
    struct PF {
        T x0, x1, f0, f1;  // by position
    };
    auto load = [&](PF& q, int p) {
        const bool okp = p >= zlo && p <= zhi && okj;
        const unsigned o0 = offs(pc0, okp && ok0, p), o1 = offs(pc1, okp && ok1, p);
        q.f0 = buf_ld<T>(rf, o0);
        q.f1 = buf_ld<T>(rf, o1);
        if constexpr (ZI) {
            q.x0 = q.x1 = T(0);
        } else {
            q.x0 = buf_ld<T>(rx, o0);
            q.x1 = buf_ld<T>(rx, o1);
        }
    };
    double amax = 0.0;
    const int txm = tx > 0 ? tx - 1 : tx, txp = tx < TZ_PX - 1 ? tx + 1 : tx;
    // BM 1 (constant coefficient b, uniform c0): stencil_nb's arithmetic with
    // the boundary selects hoisted.  A neighbour across a wall of the level
    // holds 0 here (lanes outside the level load 0 and are never written,
    // planes outside it are never loaded), so its term b * 0 adds nothing
    // and every neighbour term is taken unconditionally; the diagonal
    // depends only on which walls the cell touches, so its x and y parts are
    // fixed per lane and its z part per plane.
    const T bq = L.bconst;
    // bq * (1 - c0) of each wall (uniform values)
    T wx0 = T(0), wx1 = T(0), wy0 = T(0), wy1 = T(0), wz0 = T(0), wz1 = T(0);
    if constexpr (BM == 1) {
        wx0 = bq * (T(1) - L.c0u[0][0]);
        wx1 = bq * (T(1) - L.c0u[0][1]);
        wy0 = bq * (T(1) - L.c0u[1][0]);
        wy1 = bq * (T(1) - L.c0u[1][1]);
        wz0 = bq * (T(1) - L.c0u[2][0]);
        wz1 = bq * (T(1) - L.c0u[2][1]);
    }
    auto ddw = [&](bool lin, bool hin, T wl, T wh) -> T {
        T dd = T(0);
        dd = lin ? dd + bq : dd + wl;
        dd = hin ? dd + bq : dd + wh;
        return dd;
    };
    T dgxy0 = T(0), dgxy1 = T(0);
    if constexpr (BM == 1) {
        const T ddy = ddw(jcc > 0 || L.per1, jcc < L.ny - 1 || L.per1, wy0, wy1);
        dgxy0 = L.alpha + L.s0 * ddw(ic0c > 0 || L.per0, ic0c < L.nx - 1 || L.per0, wx0, wx1);
        dgxy0 = dgxy0 + L.s1 * ddy;
        dgxy1 = L.alpha + L.s0 * ddw(ic1c > 0 || L.per0, ic1c < L.nx - 1 || L.per0, wx0, wx1);
        dgxy1 = dgxy1 + L.s1 * ddy;
        // the z classes as st1 forms them for planes 0, 1 and nz-1 (nz >= 2)
        const T ddl = ddw(L.per2, true, wz0, wz1), ddi = ddw(true, true, wz0, wz1),
                ddh = ddw(true, L.per2, wz0, wz1);
        rdbuf[0][0][threadIdx.x] = T(1) / (dgxy0 + L.s2 * ddl);
        rdbuf[0][1][threadIdx.x] = T(1) / (dgxy1 + L.s2 * ddl);
        rdbuf[1][0][threadIdx.x] = T(1) / (dgxy0 + L.s2 * ddi);
        rdbuf[1][1][threadIdx.x] = T(1) / (dgxy1 + L.s2 * ddi);
        rdbuf[2][0][threadIdx.x] = T(1) / (dgxy0 + L.s2 * ddh);
        rdbuf[2][1][threadIdx.x] = T(1) / (dgxy1 + L.s2 * ddh);
    }
    auto st1 = [&](int u, int pc, T xl, T xh, T yl, T yh, T zl, T zh, T& dg, T& of) {
        const T ddz = ddw(pc > 0 || L.per2, pc < L.nz - 1 || L.per2, wz0, wz1);
        dg = (u ? dgxy1 : dgxy0) + L.s2 * ddz;
        T o0 = bq * xl;
        o0 = o0 + bq * xh;
        T o1 = bq * yl;
        o1 = o1 + bq * yh;
        T o2 = bq * zl;
        o2 = o2 + bq * zh;
        of = L.s0 * o0;
        of = of + L.s1 * o1;
        of = of + L.s2 * o2;
    };
    // BM 2 (cell coefficient, uniform c0): stencil_nbf's arithmetic term for
    // term without its branches.  A face enters the diagonal times 1 inside
    // the level and times (1 - c0) on a wall (times 1 is exact, so the
    // factor stands for stencil_nbf's two cases); the factors are per lane
    // (x, y) or per plane (z, never periodic here).  A neighbour across a
    // wall holds 0, so its product with the (finite) wall face adds nothing
    // and every neighbour term is taken.  f[6]: low / high faces of x, y, z.
    T mlx0 = T(1), mhx0 = T(1), mlx1 = T(1), mhx1 = T(1), mly = T(1), mhy = T(1);
    if constexpr (BM == 2) {
        const T w[2][2] = {{T(1) - L.c0u[0][0], T(1) - L.c0u[0][1]}, {T(1) - L.c0u[1][0], T(1) - L.c0u[1][1]}};
        mlx0 = ic0c > 0 || L.per0 ? T(1) : w[0][0];
        mhx0 = ic0c < L.nx - 1 || L.per0 ? T(1) : w[0][1];
        mlx1 = ic1c > 0 || L.per0 ? T(1) : w[0][0];
        mhx1 = ic1c < L.nx - 1 || L.per0 ? T(1) : w[0][1];
        mly = jcc > 0 || L.per1 ? T(1) : w[1][0];
        mhy = jcc < L.ny - 1 || L.per1 ? T(1) : w[1][1];
    }
    auto st2 = [&](int u, int pc, T xl, T xh, T yl, T yh, T zl, T zh, const T* f, T& dg, T& of) {
        const T mlx = u ? mlx1 : mlx0, mhx = u ? mhx1 : mhx0;
        const T mlz = pc > 0 ? T(1) : T(1) - L.c0u[2][0], mhz = pc < L.nz - 1 ? T(1) : T(1) - L.c0u[2][1];
        T ddx = f[0] * mlx;
        ddx = ddx + f[1] * mhx;
        T ddy = f[2] * mly;
        ddy = ddy + f[3] * mhy;
        T ddz = f[4] * mlz;
        ddz = ddz + f[5] * mhz;
        dg = L.alpha + L.s0 * ddx;
        dg = dg + L.s1 * ddy;
        dg = dg + L.s2 * ddz;
        T o0 = f[0] * xl;
        o0 = o0 + f[1] * xh;
        T o1 = f[2] * yl;
        o1 = o1 + f[3] * yh;
        T o2 = f[4] * zl;
        o2 = o2 + f[5] * zh;
        of = L.s0 * o0;
        of = of + L.s1 * o1;
        of = of + L.s2 * o2;
    };
    // ZI, BM 1: the first stage updates red cells from a zero correction,
    // so every neighbour term is b * 0 and its stencil sum is a constant
    // (formed once, in st1's order, for the same bits)
    T of0 = T(0);
    if constexpr (ZI && BM == 1) {
        T dg0;
        st1(0, 1, T(0), T(0), T(0), T(0), T(0), T(0), dg0, of0);
    }
    const int tym = ty > 0 ? ty - 1 : ty, typ = ty < TZ_RY - 1 ? ty + 1 : ty;
    const int tend = zout + S;
    PF qa, qb;
    auto step = [&](auto kc, int t) {
        constexpr int k = decltype(kc)::value;  // (t - zlo + 1) mod NV
        PF& q = (k & 1) ? qb : qa;
        double xa0 = 0.0, xa1 = 0.0;  // ACC: x of the plane finished this step
        if constexpr (ACC) {
            const int po = t - S;
            const bool outp = po >= z0 && po < zout;
            xa0 = buf_ld<double>(ra, offd(pc0, outp && oc0, po));
            xa1 = buf_ld<double>(ra, offd(pc1, outp && oc1, po));
        }
        {
            const bool c0red = ((par + t + 1) & 1) == 0;
            vr[k] = c0red ? q.x0 : q.x1;
            vb[k] = c0red ? q.x1 : q.x0;
            fr[k] = c0red ? q.f0 : q.f1;
            fb[k] = c0red ? q.f1 : q.f0;
        }
        if constexpr (BM == 2) {
            // faces of the arriving plane t+1; its LDS slot last held plane
            // t-3, whose last reader (the residual) ran before the previous
            // step's barrier
            fx1[k] = harm(pd.a, pd.b);
            {
                const T za = harm(dpa, pd.a), zb = harm(dpb, pd.b);
                const bool c0red = ((par + t + 1) & 1) == 0;
                fzr[k] = c0red ? za : zb;
                fzb[k] = c0red ? zb : za;
            }
            dpa = pd.a;
            dpb = pd.b;
            const int sl = (t + 1) & 3;
            fxs[sl][ty][tx] = harm(pd.l, pd.a);
            fys[sl][0][ty][tx] = harm(pd.ab, pd.a);
            fys[sl][1][ty][tx] = harm(pd.bb, pd.b);
        }
        // publish the opposite-colour value of each stage's plane
#pragma unroll
        for (int s = (ZI && BM == 1) ? 2 : 1; s <= S; s++) {
            const int q_s = (k - s + 2 * NV) % NV;  // plane t+1-s
            nbuf[s - 1][ty][tx] = ((s - 1) & 1) ? vr[q_s] : vb[q_s];
        }
        constexpr int qo = (k - S - 1 + 2 * NV) % NV;  // slot of plane t-S
        if constexpr (RES) {
            fbuf[0][ty][tx] = vr[qo];
            fbuf[1][ty][tx] = vb[qo];
        }
        __syncthreads();
#pragma unroll
        for (int s = 1; s <= S; s++) {
            const int p = t - s + 1;
            const int q_s = (k - s + 2 * NV) % NV;          // plane p
            const int q_lo = (k - s - 1 + 2 * NV) % NV;     // plane p-1
            const int q_hi = (k - s + 1 + 2 * NV) % NV;     // plane p+1
            const bool cs = ((s - 1) & 1) != 0;              // colour updated by this stage
            const int u = (par + p + (cs ? 1 : 0)) & 1;      // its position in the pair
            const bool valid = p >= zlo && p <= zhi && okj && (u ? ok1 : ok0);
            const T other = cs ? vr[q_s] : vb[q_s];
            const T zlo_v = cs ? vr[q_lo] : vb[q_lo];
            const T zhi_v = cs ? vr[q_hi] : vb[q_hi];
            // one LDS read: the x-neighbour outside the pair is on the left of
            // cell 0 (u = 0) or on the right of cell 1 (u = 1); an address
            // select, so the read is not predicated per lane
            const T xo = nbuf[s - 1][ty][u ? txp : txm];
            const T xm = u ? other : xo;
            const T xp = u ? xo : other;
            const T ym = nbuf[s - 1][tym][tx], yp = nbuf[s - 1][typ][tx];
            auto nb = [&](int d, int side) -> T {
                if (d == 0) return side == 0 ? xm : xp;
                if (d == 1) return side == 0 ? ym : yp;
                return side == 0 ? zlo_v : zhi_v;
            };
            T dg, of;
            if constexpr (BM == 1) {
                // times the reciprocal diagonal (within an ulp of the quotient)
                const int pc = min(max(p, 0), L.nz - 1);
                if (ZI && s == 1)
                    of = of0;  // all neighbours 0
                else
                    st1(u, pc, xm, xp, ym, yp, zlo_v, zhi_v, dg, of);
                const int zcl = pc == 0 ? 0 : (pc == L.nz - 1 ? 2 : 1);  // uniform
                const T rd = rdbuf[zcl][u][threadIdx.x];
                if (cs) {
                    const T nv = (fb[q_s] + of) * rd;
                    vb[q_s] = valid ? nv : vb[q_s];
                } else {
                    const T nv = (fr[q_s] + of) * rd;
                    vr[q_s] = valid ? nv : vr[q_s];
                }
            } else {
                if constexpr (BM == 2) {
                    // the faces of plane p: x and y from LDS (own / right pair,
                    // own / upper row), the pair's middle face and z from the ring
                    const int sl = p & 3;
                    const T fxo = fxs[sl][ty][u ? txp : tx];
                    const T fyl = fys[sl][u][ty][tx], fyh = fys[sl][u][typ][tx];
                    const T fxl = u ? fx1[q_s] : fxo, fxh = u ? fxo : fx1[q_s];
                    const T fzl = cs ? fzb[q_s] : fzr[q_s], fzh = cs ? fzr[q_hi] : fzb[q_hi];
                    const T fc[6] = {fxl, fxh, fyl, fyh, fzl, fzh};
                    st2(u, min(max(p, 0), L.nz - 1), xm, xp, ym, yp, zlo_v, zhi_v, fc, dg, of);
                    if (cs) {
                        ofb[q_s] = of;
                        dgb[q_s] = dg;
                    }
                } else {
                    stencil_nb<T, BM, 3, decltype(nb), true>(L, u ? ic1c : ic0c, jcc, min(max(p, 0), L.nz - 1),
                                                              nb, dg, of);
                }
                if (cs) {
                    const T nv = (fb[q_s] + of) / dg;
                    vb[q_s] = valid ? nv : vb[q_s];
                } else {
                    const T nv = (fr[q_s] + of) / dg;
                    vr[q_s] = valid ? nv : vr[q_s];
                }
            }
        }
        const int po = t - S;  // final from here on
        const bool outp = po >= z0 && po < zout;
        if constexpr (RES) {
            constexpr int qlo = (qo - 1 + NV) % NV, qhi = (qo + 1) % NV;
            const int poc = min(max(po, 0), L.nz - 1);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int u = (par + po + c) & 1;
                const bool valid = outp && (u ? oc1 : oc0);
                const T own = c ? vb[qo] : vr[qo];
                const T other = c ? vr[qo] : vb[qo];
                const T zlo_v = c ? vr[qlo] : vb[qlo];
                const T zhi_v = c ? vr[qhi] : vb[qhi];
                const T xo = fbuf[1 - c][ty][u ? txp : txm];
                const T xm = u ? other : xo;
                const T xp = u ? xo : other;
                const T ym = fbuf[1 - c][tym][tx], yp = fbuf[1 - c][typ][tx];
                auto nb = [&](int d, int side) -> T {
                    if (d == 0) return side == 0 ? xm : xp;
                    if (d == 1) return side == 0 ? ym : yp;
                    return side == 0 ? zlo_v : zhi_v;
                };
                T dg, of;
                if constexpr (BM == 1) {
                    st1(u, poc, xm, xp, ym, yp, zlo_v, zhi_v, dg, of);
                } else if constexpr (BM == 2) {
                    if (c) {
                        of = ofb[qo];
                        dg = dgb[qo];
                    } else {
                        const int sl = po & 3;
                        const T fxo = fxs[sl][ty][u ? txp : tx];
                        const T fyl = fys[sl][u][ty][tx], fyh = fys[sl][u][typ][tx];
                        const T fxl = u ? fx1[qo] : fxo, fxh = u ? fxo : fx1[qo];
                        const T fc[6] = {fxl, fxh, fyl, fyh, fzr[qo], fzb[qhi]};
                        st2(u, poc, xm, xp, ym, yp, zlo_v, zhi_v, fc, dg, of);
                    }
                } else {
                    stencil_nb<T, BM, 3, decltype(nb), true>(L, u ? ic1c : ic0c, jcc, poc, nb, dg, of);
                }
                const double r = (c ? fb[qo] : fr[qo]) - (dg * own - of);
                double a = fabs(r);
                if (a != a) a = __builtin_huge_val();
                amax = valid ? fmax(amax, a) : amax;
            }
        }
        {
            const bool c0red = ((par + po) & 1) == 0;
            const T o0 = c0red ? vr[qo] : vb[qo], o1 = c0red ? vb[qo] : vr[qo];
            if constexpr (ACC) {
                buf_st<double>(ra, offd(pc0, outp && oc0, po), xa0 + (double)o0);
                buf_st<double>(ra, offd(pc1, outp && oc1, po), xa1 + (double)o1);
            } else {
                buf_st<T>(ro, offs(pc0, outp && oc0, po), o0);
                buf_st<T>(ro, offs(pc1, outp && oc1, po), o1);
            }
            if constexpr (POST) {
                const T w0 = (so.floor_on && o0 < so.floor_val) ? T(so.floor_val) : o0;
                const T w1 = (so.floor_on && o1 < so.floor_val) ? T(so.floor_val) : o1;
                buf_st<T>(rpo, outp && oc0 ? (unsigned)(po0 + so.ks * po) * 8u : OOB, w0);
                buf_st<T>(rpo, outp && oc1 ? (unsigned)(po0 + 1 + so.ks * po) * 8u : OOB, w1);
            }
        }
        load(q, t + 3);
        if constexpr (BM == 2) load_d(pd, t + 2);  // one step ahead
        __syncthreads();
    };
    load(qa, zlo);
    load(qb, zlo + 1);
    if constexpr (BM == 2) {
        PD d0;
        load_d(d0, zlo - 1);
        dpa = d0.a;
        dpb = d0.b;
        load_d(pd, zlo);
    }
    for (int t = zlo - 1; t < tend; t += NV) fused_unroll<0, NV>(step, t, tend);
    if constexpr (RES) {
        for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = amax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++) m = fmax(m, red[w]);
            part[lb] = m;
        }
    }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    if (v != v) v = __builtin_huge_val();
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// norm accumulators: NSLOT slots, one per 128-byte line, blocks spread over
// them; the host takes the max of the slots
constexpr int NSLOT = 64;
constexpr int SLOT_STRIDE = 16;
constexpr int NORM_BLOCKS = 2048;

// block max (wave64 shuffles, then LDS), one atomic per block
__device__ __forceinline__ void block_max_atomic(double a, unsigned long long* slots) {
    __shared__ double red[16];
    __syncthreads();  // red may still be read by a previous call
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        atomic_max_nonneg(slots + (blockIdx.x % NSLOT) * SLOT_STRIDE, m);
    }
}

// r = f - L x (finest level, FP64), one block per gx-chunk of a row
// (XCD-remapped); each block writes the max |r| of its cells to part[lb].
// BN: also max |f| into the slot group bslots; FM (cell coefficient): also
// the max face coefficient of each direction into slot groups fslots[0..2]
template <int BM, int DIM, bool BN = false, bool FM = false>
__global__ void __launch_bounds__(256) k_resid0(LevV<double> L, const double* __restrict__ x,
                                                const double* __restrict__ f, double* __restrict__ r,
                                                double* __restrict__ part, int gx,
                                                unsigned long long* __restrict__ bslots,
                                                unsigned long long* __restrict__ fslots, Batch bt) {
    {
        const int64_t co = case_off(bt);
        L.template rebase_used<BM, true>(co);
        x = rb(x, co);
        f = rb(f, co);
        r = rb(r, co);
        part = rb(part, co);
        bslots = rb(bslots, co);
        fslots = rb(fslots, co);
    }
    constexpr int NV = 1 + (BN ? 1 : 0) + (FM ? DIM : 0);  // maxima reduced per block
    __shared__ double red[4][NV];
    const unsigned lb = lblock();
    const unsigned row = lb / gx;
    const int j = (int)(row % L.ny), k = (int)(row / L.ny);
    const int i = (int)(lb % gx) * blockDim.x + threadIdx.x;
    double m[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) m[q] = 0.0;
    if (i < L.nx) {
        double dg, of, fm[3] = {0.0, 0.0, 0.0};
        stencil<double, BM, DIM, FM>(L, x, i, j, k, dg, of, fm);
        const int c = i + L.nx * (int)row;
        const double fc = f[c];
        double v = fc - (dg * x[c] - of);
        if (r) r[c] = v;
        m[0] = fabs(v);
        if (BN) m[1] = fabs(fc);
        if (FM)
#pragma unroll
            for (int d = 0; d < DIM; d++) m[(BN ? 2 : 1) + d] = fm[d];
#pragma unroll
        for (int q = 0; q < NV; q++)
            if (m[q] != m[q]) m[q] = __builtin_huge_val();
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
        for (int o = 32; o > 0; o >>= 1) m[q] = fmax(m[q], __shfl_xor(m[q], o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = m[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) {
            mm[q] = red[0][q];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) mm[q] = fmax(mm[q], red[w][q]);
        }
        part[lb] = mm[0];
        if (BN) atomic_max_nonneg(bslots + (lb % NSLOT) * SLOT_STRIDE, mm[1]);
        if (FM)
#pragma unroll
            for (int d = 0; d < DIM; d++)
                atomic_max_nonneg(fslots + (size_t)d * NSLOT * SLOT_STRIDE + (lb % NSLOT) * SLOT_STRIDE,
                                  mm[(BN ? 2 : 1) + d]);
    }
}

// Residual of a 3-D level streamed along z.  A block owns RZ_X x RZ_Y
// columns and up to zc planes; each thread walks one (i, j) column keeping
// x -- and with BM 2 the cell coefficient -- of planes k-1, k, k+1 in
// registers, so every plane is fetched once per column rather than once per
// z-neighbour (the three planes of each array a row-per-block sweep keeps
// in flight overflow an XCD's L2 at 256^2-cell planes).  In-plane
// neighbours come from L1.  The arithmetic is stencil()'s term for term.
// Block maxima go straight to the norm slot groups: 0 the residual, 1 |f|
// (BN), 2 + d the face coefficients of direction d (FM).  C0U: every
// boundary c0 is uniform per side (L.c0uni).
constexpr int RZ_X = 64, RZ_Y = 4;
template <int BM, bool BN, bool FM, bool C0U>
__global__ void __launch_bounds__(RZ_X * RZ_Y) k_resid_z(LevV<double> L, const double* __restrict__ x,
                                                         const double* __restrict__ f, double* __restrict__ r,
                                                         int zc, int ntx, int nty,
                                                         unsigned long long* __restrict__ slots, Batch bt) {
    {
        const int64_t co = case_off(bt);
        L.template rebase_used<BM, !C0U>(co);
        x = rb(x, co);
        f = rb(f, co);
        r = rb(r, co);
        slots = rb(slots, co);
    }
    constexpr int NV = 1 + (BN ? 1 : 0) + (FM ? 3 : 0);
    __shared__ double red[RZ_X * RZ_Y / 64][NV];
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int i = (tile % ntx) * RZ_X + (int)threadIdx.x % RZ_X;
    const int j = (tile / ntx) * RZ_Y + (int)threadIdx.x / RZ_X;
    const int z0 = zb * zc, z1 = min(z0 + zc, L.nz);
    const bool own = i < L.nx && j < L.ny;
    const int ic = min(i, L.nx - 1), jc = min(j, L.ny - 1);
    const int pstr = L.nx * L.ny, c2 = ic + L.nx * jc;
    // neighbour offsets as stencil() takes them: wrapped on a periodic side,
    // the cell itself (value unused) on a wall
    const int oxl = ic > 0 ? -1 : (L.per0 ? L.nx - 1 : 0);
    const int oxh = ic < L.nx - 1 ? 1 : (L.per0 ? 1 - L.nx : 0);
    const int oyl = jc > 0 ? -L.nx : (L.per1 ? (L.ny - 1) * L.nx : 0);
    const int oyh = jc < L.ny - 1 ? L.nx : (L.per1 ? (1 - L.ny) * L.nx : 0);
    auto zw = [&](int k) {
        return k < 0 ? (L.per2 ? k + L.nz : 0) : (k >= L.nz ? (L.per2 ? k - L.nz : L.nz - 1) : k);
    };
    // the cell coefficient lives on the box grown by one cell
    const int gx = L.nx + 2, gxy = gx * (L.ny + 2), g2 = (ic + 1) + gx * (jc + 1);
    double xm = x[c2 + pstr * zw(z0 - 1)], xc = x[c2 + pstr * z0], xp = x[c2 + pstr * zw(z0 + 1)];
    double dm = 0.0, dc = 0.0, dp = 0.0;
    if constexpr (BM == 2) {
        dm = L.dcell[g2 + gxy * z0];
        dc = L.dcell[g2 + gxy * (z0 + 1)];
        dp = L.dcell[g2 + gxy * (z0 + 2)];
    }
    double m[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) m[q] = 0.0;
    for (int k = z0; k < z1; k++) {
        const int ck = c2 + pstr * k;
        const double xn = x[c2 + pstr * zw(k + 2)];  // next column value, two planes ahead
        const double xl0 = x[ck + oxl], xh0 = x[ck + oxh], xl1 = x[ck + oyl], xh1 = x[ck + oyh];
        const double fv = f[ck];
        double dn = 0.0, dl0 = 0.0, dh0 = 0.0, dl1 = 0.0, dh1 = 0.0;
        if constexpr (BM == 2) {
            const int gk = g2 + gxy * (k + 1);
            dn = L.dcell[g2 + gxy * min(k + 3, L.nz + 1)];
            dl0 = L.dcell[gk - 1];
            dh0 = L.dcell[gk + 1];
            dl1 = L.dcell[gk - gx];
            dh1 = L.dcell[gk + gx];
        }
        double fm[3] = {0.0, 0.0, 0.0};
        auto nb = [&](int d, int side) -> double {
            if (d == 0) return side ? xh0 : xl0;
            if (d == 1) return side ? xh1 : xl1;
            return side ? xp : xm;
        };
        auto fc = [&](int d, double& blo, double& bhi) {
            if constexpr (BM == 2) {
                blo = harm(d == 0 ? dl0 : (d == 1 ? dl1 : dm), dc);
                bhi = harm(dc, d == 0 ? dh0 : (d == 1 ? dh1 : dp));
            } else {
                face_coef<double, BM, 3>(L, d, ic, jc, k, blo, bhi);
            }
            if (FM) fm[d] = fmax(fabs(blo), fabs(bhi));
        };
        double dg, of;
        stencil_nbf<double, 3, decltype(nb), decltype(fc), C0U>(L, ic, jc, k, nb, fc, dg, of);
        const double v = fv - (dg * xc - of);
        if (r && own) r[ck] = v;
        if (own) {
            double a[NV];
            a[0] = fabs(v);
            if (BN) a[1] = fabs(fv);
            if (FM)
#pragma unroll
                for (int d = 0; d < 3; d++) a[(BN ? 2 : 1) + d] = fm[d];
#pragma unroll
            for (int q = 0; q < NV; q++) m[q] = fmax(m[q], a[q] != a[q] ? __builtin_huge_val() : a[q]);
        }
        xm = xc;
        xc = xp;
        xp = xn;
        dm = dc;
        dc = dp;
        dp = dn;
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
        for (int o = 32; o > 0; o >>= 1) m[q] = fmax(m[q], __shfl_xor(m[q], o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = m[q];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int q = (int)threadIdx.x;
        double mm = red[0][q];
        for (int w = 1; w < RZ_X * RZ_Y / 64; w++) mm = fmax(mm, red[w][q]);
        const int grp = q == 0 ? 0 : (BN && q == 1 ? 1 : 2 + (q - (BN ? 2 : 1)));
        atomic_max_nonneg(slots + ((size_t)grp * NSLOT + lb % NSLOT) * SLOT_STRIDE, mm);
    }
}

// The residual of a 3-D level with a constant coefficient (BM 1) streamed
// along z, one x-pair of cells per thread: the pair, the rows beside it and
// the rhs as 16-byte loads, the two outer x-neighbours as 8-byte loads --
// half the load instructions of one cell per thread.  Per cell the
// arithmetic of k_resid_z (stencil_nbf, term for term).  Even nx.  RT
// float stores the residual rounded to FP32, the finest rhs of an FP32
// V-cycle.
constexpr int RZ2_X = 32, RZ2_Y = 8;  // pairs x rows per block
template <bool BN, bool C0U, class RT = double>
__global__ void __launch_bounds__(RZ2_X * RZ2_Y) k_resid_z2(LevV<double> L, const double* __restrict__ x,
                                                            const double* __restrict__ f, RT* __restrict__ r,
                                                            int zc, int ntx, int nty,
                                                            unsigned long long* __restrict__ slots, Batch bt) {
    {
        const int64_t co = case_off(bt);
        L.template rebase_used<1, !C0U>(co);
        x = rb(x, co);
        f = rb(f, co);
        r = rb(r, co);
        slots = rb(slots, co);
    }
    constexpr int NV = 1 + (BN ? 1 : 0);
    __shared__ double red[RZ2_X * RZ2_Y / 64][NV];
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int i0 = 2 * ((tile % ntx) * RZ2_X + (int)threadIdx.x % RZ2_X);
    const int j = (tile / ntx) * RZ2_Y + (int)threadIdx.x / RZ2_X;
    const int z0 = zb * zc, z1 = min(z0 + zc, L.nz);
    const bool own = i0 < L.nx && j < L.ny;  // both cells of the pair (nx even)
    const int ic = min(i0, L.nx - 2), jc = min(j, L.ny - 1);
    const int pstr = L.nx * L.ny, c2 = ic + L.nx * jc;
    // the left neighbour of cell ic and the right one of ic+1 (offsets from
    // ic), and the rows: wrapped on a periodic side, a cell of the pair
    // (value unused) on a wall
    const int oxl = ic > 0 ? -1 : (L.per0 ? L.nx - 1 : 0);
    const int oxh = ic + 1 < L.nx - 1 ? 2 : (L.per0 ? 2 - L.nx : 1);
    const int oyl = jc > 0 ? -L.nx : (L.per1 ? (L.ny - 1) * L.nx : 0);
    const int oyh = jc < L.ny - 1 ? L.nx : (L.per1 ? (1 - L.ny) * L.nx : 0);
    auto zw = [&](int k) {
        return k < 0 ? (L.per2 ? k + L.nz : 0) : (k >= L.nz ? (L.per2 ? k - L.nz : L.nz - 1) : k);
    };
    auto ld2 = [](const double* p) { return *(const double2*)p; };  // even offsets: 16-byte aligned
    double2 xm = ld2(x + c2 + pstr * zw(z0 - 1)), xc = ld2(x + c2 + pstr * z0), xp = ld2(x + c2 + pstr * zw(z0 + 1));
    double m[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) m[q] = 0.0;
    for (int k = z0; k < z1; k++) {
        const int ck = c2 + pstr * k;
        const double2 xn = ld2(x + c2 + pstr * zw(k + 2));  // two planes ahead
        const double xl = x[ck + oxl], xh = x[ck + oxh];
        const double2 yl = ld2(x + ck + oyl), yh = ld2(x + ck + oyh);
        const double2 fv = ld2(f + ck);
        double v[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            auto nb = [&](int d, int side) -> double {
                if (d == 0) return c == 0 ? (side ? xc.y : xl) : (side ? xh : xc.x);
                if (d == 1) return side ? (c ? yh.y : yh.x) : (c ? yl.y : yl.x);
                return side ? (c ? xp.y : xp.x) : (c ? xm.y : xm.x);
            };
            auto fc = [&](int d, double& blo, double& bhi) { face_coef<double, 1, 3>(L, d, ic + c, jc, k, blo, bhi); };
            double dg, of;
            stencil_nbf<double, 3, decltype(nb), decltype(fc), C0U>(L, ic + c, jc, k, nb, fc, dg, of);
            v[c] = (c ? fv.y : fv.x) - (dg * (c ? xc.y : xc.x) - of);
        }
        if (r && own) {
            if constexpr (std::is_same<RT, float>::value)
                *(float2*)(r + ck) = make_float2((float)v[0], (float)v[1]);
            else
                *(double2*)(r + ck) = make_double2(v[0], v[1]);
        }
        if (own) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                double a[NV];
                a[0] = fabs(v[c]);
                if (BN) a[NV - 1] = fabs(c ? fv.y : fv.x);
#pragma unroll
                for (int q = 0; q < NV; q++) m[q] = fmax(m[q], a[q] != a[q] ? __builtin_huge_val() : a[q]);
            }
        }
        xm = xc;
        xc = xp;
        xp = xn;
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
        for (int o = 32; o > 0; o >>= 1) m[q] = fmax(m[q], __shfl_xor(m[q], o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = m[q];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int q = (int)threadIdx.x;
        double mm = red[0][q];
        for (int w = 1; w < RZ2_X * RZ2_Y / 64; w++) mm = fmax(mm, red[w][q]);
        atomic_max_nonneg(slots + ((size_t)q * NSLOT + lb % NSLOT) * SLOT_STRIDE, mm);
    }
}

// The rhs of an implicit species solve on the constant-coefficient stencil
// and the norms of its initial residual in one pass (3-D, even nx, the
// streamed layout of k_resid_z2), so that neither f nor x is read back:
//   f = (1/dt) n_old + dsdt (ScalarSolve.cpp:443-465), then on a wall cell
//       + beta/dx_d^2 * b * cinh of each wall face, directions in order
//       (the terms side_setup adds to f; cinh from side_setup's W.sh),
//   x = Sborder's component (the initial guess),
// and max |f - L x|, max |f| into slot groups 0 and 1, as
// k_resid_z2<true, C0U> leaves them.  Per cell the arithmetic of
// k_spec_rhs, side_setup and k_resid_z2, the contractions written out.
struct SpecR0Args {
    Arr<double> old, ds, sb;  // the species component of old and sb; ds component 0
    int spec;
    int3 lo;
    double idt;
    const double* sh[3][2];  // cinh by face of each wall side (null on periodic sides)
    double sbd[3];           // beta/dx_d^2 * b
    double* f;
    double* x;
    unsigned long long* slots;
    __device__ void rebase(int64_t o) {
        vd::rebase(old, o);
        vd::rebase(ds, o);
        vd::rebase(sb, o);
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int s = 0; s < 2; s++) sh[d][s] = vd::rb(sh[d][s], o);
        f = vd::rb(f, o);
        x = vd::rb(x, o);
        slots = vd::rb(slots, o);
    }
};

template <bool C0U>
__global__ void __launch_bounds__(RZ2_X * RZ2_Y) k_spec_rhs_r0(LevV<double> L, SpecR0Args a, int zc, int ntx,
                                                               int nty, Batch bt) {
    {
        const int64_t co = case_off(bt);
        L.template rebase_used<1, !C0U>(co);
        a.rebase(co);
    }
    __shared__ double red[RZ2_X * RZ2_Y / 64][2];
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int i0 = 2 * ((tile % ntx) * RZ2_X + (int)threadIdx.x % RZ2_X);
    const int j = (tile / ntx) * RZ2_Y + (int)threadIdx.x / RZ2_X;
    const int z0 = zb * zc, z1 = min(z0 + zc, L.nz);
    const bool own = i0 < L.nx && j < L.ny;  // both cells of the pair (nx even)
    const int ic = min(i0, L.nx - 2), jc = min(j, L.ny - 1);
    const int pstr = L.nx * L.ny, c2 = ic + L.nx * jc;
    // x from Sborder in its own layout: the pair on level plane 0, the
    // strides, and the neighbours as k_resid_z2 takes them (wrapped on a
    // periodic side, a cell of the pair on a wall)
    const double* S = a.sb.comp(a.spec);
    const int sj = a.sb.js, sk = a.sb.ks;
    const int s2 = a.sb.off(a.lo.x + ic, a.lo.y + jc, a.lo.z);
    const int oxl = ic > 0 ? -1 : (L.per0 ? L.nx - 1 : 0);
    const int oxh = ic + 1 < L.nx - 1 ? 2 : (L.per0 ? 2 - L.nx : 1);
    const int oyl = jc > 0 ? -sj : (L.per1 ? (L.ny - 1) * sj : 0);
    const int oyh = jc < L.ny - 1 ? sj : (L.per1 ? (1 - L.ny) * sj : 0);
    auto zw = [&](int k) {
        return k < 0 ? (L.per2 ? k + L.nz : 0) : (k >= L.nz ? (L.per2 ? k - L.nz : L.nz - 1) : k);
    };
    auto ld2 = [&](int o) { return make_double2(S[o], S[o + 1]); };
    const double* O = a.old.comp(a.spec);
    const int o2 = a.old.off(a.lo.x + ic, a.lo.y + jc, a.lo.z), ok = a.old.ks;
    const double* D = a.ds.p;
    const int d2 = a.ds.off(a.lo.x + ic, a.lo.y + jc, a.lo.z), dk = a.ds.ks;
    double2 xm = ld2(s2 + sk * zw(z0 - 1)), xc = ld2(s2 + sk * z0), xp = ld2(s2 + sk * zw(z0 + 1));
    double m[2] = {0.0, 0.0};
    for (int k = z0; k < z1; k++) {
        const int ck = c2 + pstr * k, sc = s2 + sk * k;
        const double2 xn = ld2(s2 + sk * zw(k + 2));  // two planes ahead
        const double xl = S[sc + oxl], xh = S[sc + oxh];
        const double2 yl = ld2(sc + oyl), yh = ld2(sc + oyh);
        double fv[2];
#pragma unroll
        for (int c = 0; c < 2; c++) fv[c] = fma(a.idt, O[o2 + ok * k + c], D[d2 + dk * k + c]);
        // the wall terms, direction by direction and low side first
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int i = ic + c;
            if (!L.per0) {
                if (i == 0) fv[c] = fma(a.sbd[0], a.sh[0][0][jc + L.ny * k], fv[c]);
                if (i == L.nx - 1) fv[c] = fma(a.sbd[0], a.sh[0][1][jc + L.ny * k], fv[c]);
            }
            if (!L.per1) {
                if (jc == 0) fv[c] = fma(a.sbd[1], a.sh[1][0][i + L.nx * k], fv[c]);
                if (jc == L.ny - 1) fv[c] = fma(a.sbd[1], a.sh[1][1][i + L.nx * k], fv[c]);
            }
            if (!L.per2) {
                if (k == 0) fv[c] = fma(a.sbd[2], a.sh[2][0][i + L.nx * jc], fv[c]);
                if (k == L.nz - 1) fv[c] = fma(a.sbd[2], a.sh[2][1][i + L.nx * jc], fv[c]);
            }
        }
        double v[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            auto nb = [&](int d, int side) -> double {
                if (d == 0) return c == 0 ? (side ? xc.y : xl) : (side ? xh : xc.x);
                if (d == 1) return side ? (c ? yh.y : yh.x) : (c ? yl.y : yl.x);
                return side ? (c ? xp.y : xp.x) : (c ? xm.y : xm.x);
            };
            auto fc = [&](int d, double& blo, double& bhi) { face_coef<double, 1, 3>(L, d, ic + c, jc, k, blo, bhi); };
            double dg, of;
            stencil_nbf<double, 3, decltype(nb), decltype(fc), C0U>(L, ic + c, jc, k, nb, fc, dg, of);
            v[c] = fv[c] - (dg * (c ? xc.y : xc.x) - of);
        }
        if (own) {
            *(double2*)(a.f + ck) = make_double2(fv[0], fv[1]);
            *(double2*)(a.x + ck) = xc;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const double r = fabs(v[c]), b = fabs(fv[c]);
                m[0] = fmax(m[0], r != r ? __builtin_huge_val() : r);
                m[1] = fmax(m[1], b != b ? __builtin_huge_val() : b);
            }
        }
        xm = xc;
        xc = xp;
        xp = xn;
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        for (int o = 32; o > 0; o >>= 1) m[q] = fmax(m[q], __shfl_xor(m[q], o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = m[q];
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int q = (int)threadIdx.x;
        double mm = red[0][q];
        for (int w = 1; w < RZ2_X * RZ2_Y / 64; w++) mm = fmax(mm, red[w][q]);
        atomic_max_nonneg(a.slots + ((size_t)q * NSLOT + lb % NSLOT) * SLOT_STRIDE, mm);
    }
}

// One colour of a red-black sweep of a 3-D level (even nx) streamed along
// z.  A thread owns the x-pair (i0, i0+1) of one row and walks up to zc
// planes; on plane k it updates the cell of the pair whose colour is
// `color`.  Every neighbour of that cell has the other colour, which this
// pass does not write, so the z-neighbours are the pair values of planes
// k-1 and k+1 held in registers from earlier loads, and the pair mate is
// the x-neighbour on one side.  Arithmetic as k_gsrb.
// ZI: the first colour of a sweep from x = 0 (no zeroing pass): every
// neighbour is 0, and the thread also writes the 0 of the other cell.
constexpr int GZ_X = 32, GZ_Y = 8;  // pairs x rows per block
template <class T, int BM, bool C0U, bool ZI = false, bool BT = false>
__global__ void __launch_bounds__(GZ_X * GZ_Y) k_gsrb_z(LevV<T> L, T* __restrict__ x, const T* __restrict__ rhs,
                                                        int color, int zc, int ntx, int nty, Batch bt) {
    if constexpr (BT) {
        const int64_t co = case_off(bt);
        L.template rebase_used<BM, !C0U>(co);
        x = rb(x, co);
        rhs = rb(rhs, co);
    }
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int i0 = 2 * ((tile % ntx) * GZ_X + (int)threadIdx.x % GZ_X);
    const int j = (tile / ntx) * GZ_Y + (int)threadIdx.x / GZ_X;
    const int z0 = zb * zc, z1 = min(z0 + zc, L.nz);
    const bool own = i0 < L.nx && j < L.ny;
    const int ic = min(i0, L.nx - 2), jc = min(j, L.ny - 1);
    const int pstr = L.nx * L.ny, c2 = ic + L.nx * jc;
    // x-neighbours outside the pair (left of ic, right of ic+1) and the
    // y-neighbours, as stencil() takes them: wrapped on a periodic side,
    // the cell itself (value unused) on a wall
    const int oxl = ic > 0 ? -1 : (L.per0 ? L.nx - 1 : 0);
    const int oxh = ic + 1 < L.nx - 1 ? 1 : (L.per0 ? 1 - L.nx : 0);
    const int oyl = jc > 0 ? -L.nx : (L.per1 ? (L.ny - 1) * L.nx : 0);
    const int oyh = jc < L.ny - 1 ? L.nx : (L.per1 ? (1 - L.ny) * L.nx : 0);
    auto zw = [&](int k) {
        return k < 0 ? (L.per2 ? k + L.nz : 0) : (k >= L.nz ? (L.per2 ? k - L.nz : L.nz - 1) : k);
    };
    const int gx = L.nx + 2, gxy = gx * (L.ny + 2), g2 = (ic + 1) + gx * (jc + 1);
    T am0 = T(0), am1 = T(0), ac0 = T(0), ac1 = T(0), ap0 = T(0), ap1 = T(0);  // x pair of planes k-1, k, k+1
    if constexpr (!ZI) {
        const int cm = c2 + pstr * zw(z0 - 1), cc = c2 + pstr * z0, cp = c2 + pstr * zw(z0 + 1);
        am0 = x[cm];
        am1 = x[cm + 1];
        ac0 = x[cc];
        ac1 = x[cc + 1];
        ap0 = x[cp];
        ap1 = x[cp + 1];
    }
    T dm0 = T(0), dm1 = T(0), dc0 = T(0), dc1 = T(0), dp0 = T(0), dp1 = T(0);
    if constexpr (BM == 2) {
        dm0 = L.dcell[g2 + gxy * z0];
        dm1 = L.dcell[g2 + 1 + gxy * z0];
        dc0 = L.dcell[g2 + gxy * (z0 + 1)];
        dc1 = L.dcell[g2 + 1 + gxy * (z0 + 1)];
        dp0 = L.dcell[g2 + gxy * (z0 + 2)];
        dp1 = L.dcell[g2 + 1 + gxy * (z0 + 2)];
    }
    for (int k = z0; k < z1; k++) {
        const int u = (color + j + k) & 1;  // position of the updated cell in the pair
        const int i = ic + u;
        const int ck = c2 + pstr * k;
        const int cn = c2 + pstr * zw(k + 2);
        T an0 = T(0), an1 = T(0), xo = T(0), yl = T(0), yh = T(0);
        if constexpr (!ZI) {
            an0 = x[cn];
            an1 = x[cn + 1];
            xo = x[ck + (u ? 1 + oxh : oxl)];
            yl = x[ck + u + oyl];
            yh = x[ck + u + oyh];
        }
        const T fv = rhs[ck + u];
        const T mate = u ? ac0 : ac1;
        const T xl0 = u ? mate : xo, xh0 = u ? xo : mate;
        const T zl = u ? am1 : am0, zh = u ? ap1 : ap0;
        T dn0 = T(0), dn1 = T(0), dxo = T(0), dyl = T(0), dyh = T(0);
        if constexpr (BM == 2) {
            const int gn = g2 + gxy * min(k + 3, L.nz + 1), gk = g2 + gxy * (k + 1);
            dn0 = L.dcell[gn];
            dn1 = L.dcell[gn + 1];
            dxo = L.dcell[gk + (u ? 2 : -1)];
            dyl = L.dcell[gk + u - gx];
            dyh = L.dcell[gk + u + gx];
        }
        auto nb = [&](int d, int side) -> T {
            if (d == 0) return side ? xh0 : xl0;
            if (d == 1) return side ? yh : yl;
            return side ? zh : zl;
        };
        auto fc = [&](int d, T& blo, T& bhi) {
            if constexpr (BM == 2) {
                const T dcc = u ? dc1 : dc0, dmate = u ? dc0 : dc1;
                const T lo = d == 0 ? (u ? dmate : dxo) : (d == 1 ? dyl : (u ? dm1 : dm0));
                const T hi = d == 0 ? (u ? dxo : dmate) : (d == 1 ? dyh : (u ? dp1 : dp0));
                blo = harm(lo, dcc);
                bhi = harm(dcc, hi);
            } else {
                face_coef<T, BM, 3>(L, d, i, jc, k, blo, bhi);
            }
        };
        T dg, of;
        stencil_nbf<T, 3, decltype(nb), decltype(fc), C0U>(L, i, jc, k, nb, fc, dg, of);
        if (own) {
            // the whole pair in one store: the other cell keeps its value (0
            // from zero), so every written line is written in full
            using V2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
            const T nv = (fv + of) / dg;
            const T keep = ZI ? T(0) : (u ? ac0 : ac1);
            V2 w;
            w.x = u ? keep : nv;
            w.y = u ? nv : keep;
            *reinterpret_cast<V2*>(x + ck) = w;
        }
        am0 = ac0;
        am1 = ac1;
        ac0 = ap0;
        ac1 = ap1;
        ap0 = an0;
        ap1 = an1;
        dm0 = dc0;
        dm1 = dc1;
        dc0 = dp0;
        dc1 = dp1;
        dp0 = dn0;
        dp1 = dn1;
    }
}

// z-chunk of the streamed kernels: the largest power of two <= 32 that
// still gives about 2048 blocks (8 per CU)
static int resid_zc(const int n[3], int ntile) {
    int zc = 32;
    while (zc > 2 && (int64_t)ntile * ((n[2] + zc - 1) / zc) < 2048) zc >>= 1;
    return zc;
}

// z-chunk of a streamed launch of `kern`: the power-of-two rule, unless
// on a level large enough to fill the device its last round of resident
// workgroups would be mostly empty (each workgroup streams its chunk, so
// that round would run nearly alone); then the chunk count (4..48 planes
// per chunk) that best fills whole rounds
static int stream_zc(vd_ctx* ctx, const void* kern, int block, const int n[3], int ntile) {
    const int zc0 = resid_zc(n, ntile);
    const int64_t b0 = (int64_t)ntile * ((n[2] + zc0 - 1) / zc0);
    if (b0 < 1024) return zc0;
    const int cap = resident_blocks(ctx, kern, block);
    if ((double)b0 / (double)(((b0 + cap - 1) / cap) * cap) >= 0.75) return zc0;
    const int nz = n[2];
    const int nmin = std::max(1, (nz + 47) / 48), nmax = std::max(nmin, nz / 4);
    const int nc = round_filling_chunks(ntile, cap, nmin, nmax);
    return (nz + nc - 1) / nc;
}

// max of n per-block partials into slot 0 of a slot group (one workgroup)
__global__ void __launch_bounds__(1024) k_reduce_partials(const double* __restrict__ part, int n,
                                                          unsigned long long* __restrict__ slots, Batch bt) {
    part = rb(part, case_off(bt));
    slots = rb(slots, case_off(bt));
    __shared__ double red[16];
    double a = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) a = fmax(a, part[t]);
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        atomic_max_nonneg(slots, m);
    }
}

__global__ void __launch_bounds__(256) k_maxabs(const double* __restrict__ f, unsigned n,
                                                unsigned long long* __restrict__ slots, Batch bt) {
    f = rb(f, case_off(bt));
    slots = rb(slots, case_off(bt));
    double a = 0.0;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        double b = fabs(f[t]);
        if (b != b) b = __builtin_huge_val();
        a = fmax(a, b);
    }
    block_max_atomic(a, slots);
}


// coarse r = average of the 2^DIM fine residuals rhs - L e (fused)
template <class T, int BM, int DIM>
__device__ __forceinline__ T resid_restrict_cell(const LevV<T>& F, const T* __restrict__ e,
                                                 const T* __restrict__ rhs, int I, int J, int K) {
    T s = T(0);
    const int kz = DIM == 3 ? 2 : 1;
    for (int kk = 0; kk < kz; kk++)
        for (int jj = 0; jj < 2; jj++)
            for (int ii = 0; ii < 2; ii++) {
                int i = 2 * I + ii, j = 2 * J + jj, k = DIM == 3 ? 2 * K + kk : 0;
                T dg, of;
                stencil<T, BM, DIM>(F, e, i, j, k, dg, of);
                int c = i + F.nx * (j + F.ny * k);
                s += rhs[c] - (dg * e[c] - of);
            }
    return s * (DIM == 3 ? T(0.125) : T(0.25));
}

// 1-D grid of gx blocks per coarse row, XCD-remapped
template <class T, int BM, int DIM, bool BT = false>
__global__ void __launch_bounds__(128) k_resid_restrict(LevV<T> F, const T* __restrict__ e,
                                                        const T* __restrict__ rhs, int cnx, int cny,
                                                        int cnz, T* __restrict__ rc, int gx, Batch bt) {
    if constexpr (BT) {
        const int64_t co = case_off(bt);
        F.template rebase_used<BM, true>(co);
        e = rb(e, co);
        rhs = rb(rhs, co);
        rc = rb(rc, co);
    }
    const unsigned lb = lblock();
    const unsigned row = lb / gx;
    const int J = (int)(row % cny), K = (int)(row / cny);
    const int I = (int)(lb % gx) * blockDim.x + threadIdx.x;
    if (I >= cnx) return;
    rc[I + cnx * (J + cny * K)] = resid_restrict_cell<T, BM, DIM>(F, e, rhs, I, J, K);
}

// e_f += P e_c (linear, tensor 3/4-1/4; ghosts of the coarse correction
// are c0 * interior, periodic wraps): the increment of fine cell (i,j,k)
template <class T, int DIM>
__device__ __forceinline__ T prolong_cell(const LevV<T>& C, const T* __restrict__ ec, int i, int j, int k,
                                          int64_t co = 0) {
    int f[3] = {i, j, k};
    int cn[3] = {C.nx, C.ny, C.nz};
    int pr[3] = {C.per0, C.per1, C.per2};
    int ci[3][2];
    bool ghost[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= DIM) {
            ci[d][0] = ci[d][1] = 0;
            ghost[d] = false;
            continue;
        }
        int I = f[d] >> 1, o = f[d] & 1;
        int nb = o ? I + 1 : I - 1;
        ghost[d] = false;
        if (nb < 0 || nb >= cn[d]) {
            if (pr[d]) nb = nb < 0 ? cn[d] - 1 : 0;
            else {
                ghost[d] = true;
                nb = I;
            }
        }
        ci[d][0] = I;
        ci[d][1] = nb;
    }
    T val = T(0);
    const int kn = DIM == 3 ? 2 : 1;
    for (int kz = 0; kz < kn; kz++)
        for (int jy = 0; jy < 2; jy++)
            for (int ix = 0; ix < 2; ix++) {
                int cc0 = ci[0][ix], cc1 = ci[1][jy], cc2 = ci[2][kz];
                int sel[3] = {ix, jy, kz};
                T v = ec[cc0 + C.nx * (cc1 + C.ny * cc2)];
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (sel[d] == 1 && ghost[d]) {
                        int side = f[d] & 1;
                        v = v * (C.c0uni ? C.c0u[d][side] : rb(C.c0[d][side], co)[tidx(C, d, cc0, cc1, cc2)]);
                    }
                T w = (ix == 0 ? T(0.75) : T(0.25)) * (jy == 0 ? T(0.75) : T(0.25)) *
                      (DIM == 3 ? (kz == 0 ? T(0.75) : T(0.25)) : T(1));
                val += w * v;
            }
    return val;
}

// 1-D grid of gx blocks per fine row
template <class T, int DIM, bool BT = false>
__global__ void __launch_bounds__(128) k_prolong(LevV<T> C, const T* __restrict__ ec, int fnx, int fny,
                                                 int fnz, T* __restrict__ ef, int gx, Batch bt) {
    int64_t cco = 0;  // C's c0 arrays: moved where read
    if constexpr (BT) {
        cco = case_off(bt);
        ec = rb(ec, cco);
        ef = rb(ef, cco);
    }
    const unsigned lb = lblock();
    const unsigned row = lb / gx;
    const int j = (int)(row % fny), k = (int)(row / fny);
    const int i = (int)(lb % gx) * blockDim.x + threadIdx.x;
    if (i >= fnx) return;
    ef[i + fnx * (j + fny * k)] += prolong_cell<T, DIM>(C, ec, i, j, k, cco);
}

// Residual-restriction of a 3-D level, streamed along z: a thread owns one
// coarse column (I, J) and walks coarse planes K; the 2 x 2 fine columns
// under it keep e of fine planes 2K-1 .. 2K+2 in registers, so only the
// in-plane neighbours outside the 2 x 2 block and rhs are loaded per
// plane.  Arithmetic (terms and their order) as resid_restrict_cell.
constexpr int RR_X = 64, RR_Y = 4;
template <class T, int BM, bool C0U, bool BT = false>
__global__ void __launch_bounds__(RR_X * RR_Y) k_resid_restrict_z(LevV<T> F, const T* __restrict__ e,
                                                                  const T* __restrict__ rhs, int cnx, int cny,
                                                                  int cnz, T* __restrict__ rc, int zc, int ntx,
                                                                  int nty, Batch bt) {
    if constexpr (BT) {
        const int64_t co = case_off(bt);
        F.template rebase_used<BM, !C0U>(co);
        e = rb(e, co);
        rhs = rb(rhs, co);
        rc = rb(rc, co);
    }
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int I = (tile % ntx) * RR_X + (int)threadIdx.x % RR_X;
    const int J = (tile / ntx) * RR_Y + (int)threadIdx.x / RR_X;
    const bool own = I < cnx && J < cny;
    const int Ic = min(I, cnx - 1), Jc = min(J, cny - 1);
    const int K0 = zb * zc, K1 = min(K0 + zc, cnz);
    const int fi = 2 * Ic, fj = 2 * Jc, fps = F.nx * F.ny;
    // in-plane offsets of the neighbours outside the 2 x 2 block (as
    // stencil() takes them: wrapped when periodic, the cell itself on a
    // wall, where the value is not used)
    const int oxl = fi > 0 ? -1 : (F.per0 ? F.nx - 1 : 0);
    const int oxh = fi + 1 < F.nx - 1 ? 1 : (F.per0 ? 1 - F.nx : 0);
    const int oyl = fj > 0 ? -F.nx : (F.per1 ? (F.ny - 1) * F.nx : 0);
    const int oyh = fj + 1 < F.ny - 1 ? F.nx : (F.per1 ? (1 - F.ny) * F.nx : 0);
    auto kw = [&](int k) {
        return k < 0 ? (F.per2 ? k + F.nz : 0) : (k >= F.nz ? (F.per2 ? k - F.nz : F.nz - 1) : k);
    };
    const int b00 = fi + F.nx * fj;  // fine (fi, fj) in a plane
    struct P4 {
        T v[2][2];  // [jj][ii]
    };
    auto ld4 = [&](int k) -> P4 {
        const int b = b00 + fps * kw(k);
        P4 q;
        q.v[0][0] = e[b];
        q.v[0][1] = e[b + 1];
        q.v[1][0] = e[b + F.nx];
        q.v[1][1] = e[b + F.nx + 1];
        return q;
    };
    P4 pm = ld4(2 * K0 - 1), p0 = ld4(2 * K0), p1 = ld4(2 * K0 + 1), p2 = ld4(2 * K0 + 2);
    for (int K = K0; K < K1; K++) {
        const P4 p3 = ld4(2 * K + 3), p4 = ld4(2 * K + 4);
        T s = T(0);
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            const int k = 2 * K + kk;
            const P4& zl = kk ? p0 : pm;
            const P4& zc_ = kk ? p1 : p0;
            const P4& zh = kk ? p2 : p1;
            const int bk = b00 + fps * k;
            T xo[2][2], yo[2][2], rv[2][2];  // [row or column][side]
#pragma unroll
            for (int q = 0; q < 2; q++) {
                xo[q][0] = e[bk + F.nx * q + oxl];
                xo[q][1] = e[bk + F.nx * q + 1 + oxh];
                yo[q][0] = e[bk + q + oyl];
                yo[q][1] = e[bk + F.nx + q + oyh];
                rv[q][0] = rhs[bk + F.nx * q];
                rv[q][1] = rhs[bk + F.nx * q + 1];
            }
#pragma unroll
            for (int jj = 0; jj < 2; jj++)
#pragma unroll
                for (int ii = 0; ii < 2; ii++) {
                    const int i = fi + ii, j = fj + jj;
                    auto nb = [&](int d, int side) -> T {
                        if (d == 0) return side ? (ii ? xo[jj][1] : zc_.v[jj][1]) : (ii ? zc_.v[jj][0] : xo[jj][0]);
                        if (d == 1) return side ? (jj ? yo[ii][1] : zc_.v[1][ii]) : (jj ? zc_.v[0][ii] : yo[ii][0]);
                        return side ? zh.v[jj][ii] : zl.v[jj][ii];
                    };
                    auto fc = [&](int d, T& blo, T& bhi) { face_coef<T, BM, 3>(F, d, i, j, k, blo, bhi); };
                    T dg, of;
                    stencil_nbf<T, 3, decltype(nb), decltype(fc), C0U>(F, i, j, k, nb, fc, dg, of);
                    s += rv[jj][ii] - (dg * zc_.v[jj][ii] - of);
                }
        }
        if (own) rc[Ic + cnx * (Jc + cny * K)] = s * T(0.125);
        pm = p1;
        p0 = p2;
        p1 = p3;
        p2 = p4;
    }
}

// e_f += P e_c on a 3-D level, streamed along z (uniform c0).  A thread
// owns one fine column (i, j) and walks coarse planes K, producing fine
// planes 2K and 2K+1.  The four coarse values a fine cell of the column
// draws on in a coarse plane (its coarse column and the neighbours on its
// side in x and y, x/y ghosts already scaled by c0) are kept for planes
// K-1, K, K+1, so each is loaded once per column.  Arithmetic (terms,
// order, ghost factors) as prolong_cell.
constexpr int PZ_X = 64, PZ_Y = 4;
template <class T>
__global__ void __launch_bounds__(PZ_X * PZ_Y) k_prolong_z(LevV<T> C, const T* __restrict__ ec, int fnx, int fny,
                                                           T* __restrict__ ef, int zc, int ntx, int nty, Batch bt) {
    {
        const int64_t co = case_off(bt);
        // C's coefficient arrays are not read (uniform c0 only)
        ec = rb(ec, co);
        ef = rb(ef, co);
    }
    const unsigned lb = lblock();
    const int ntile = ntx * nty;
    const int tile = (int)(lb % (unsigned)ntile), zb = (int)(lb / (unsigned)ntile);
    const int i = (tile % ntx) * PZ_X + (int)threadIdx.x % PZ_X;
    const int j = (tile / ntx) * PZ_Y + (int)threadIdx.x / PZ_X;
    const bool own = i < fnx && j < fny;
    const int ic = min(i, fnx - 1), jc = min(j, fny - 1);
    const int K0 = zb * zc, K1 = min(K0 + zc, C.nz);
    const int I = ic >> 1, sx = ic & 1, J = jc >> 1, sy = jc & 1;
    int nbx = sx ? I + 1 : I - 1, nby = sy ? J + 1 : J - 1;
    T cx = T(1), cy = T(1);
    if (nbx < 0 || nbx >= C.nx) {
        if (C.per0) {
            nbx = nbx < 0 ? C.nx - 1 : 0;
        } else {
            nbx = I;
            cx = C.c0u[0][sx];
        }
    }
    if (nby < 0 || nby >= C.ny) {
        if (C.per1) {
            nby = nby < 0 ? C.ny - 1 : 0;
        } else {
            nby = J;
            cy = C.c0u[1][sy];
        }
    }
    const int cps = C.nx * C.ny;
    const int o00 = I + C.nx * J, o10 = nbx + C.nx * J, o01 = I + C.nx * nby, o11 = nbx + C.nx * nby;
    // coarse plane of a plane index outside the level: wrapped (periodic)
    // or the edge plane (its values are then scaled by the wall's c0)
    auto kw = [&](int K) {
        return K < 0 ? (C.per2 ? K + C.nz : 0) : (K >= C.nz ? (C.per2 ? K - C.nz : C.nz - 1) : K);
    };
    struct Q {
        T a00, a10, a01, a11;
    };
    auto ld = [&](int K) -> Q {
        const int b = cps * kw(K);
        Q q;
        q.a00 = ec[b + o00];
        q.a10 = ec[b + o10] * cx;
        q.a01 = ec[b + o01] * cy;
        q.a11 = ec[b + o11] * cx * cy;
        return q;
    };
    // the eight terms of prolong_cell: plane Kc first (kz = 0), then the
    // neighbour plane (kz = 1, scaled by cz when it is a ghost)
    auto interp = [&](const Q& qc, const Q& qn, T cz) -> T {
        T v = T(0);
        v += T(0.75) * T(0.75) * T(0.75) * qc.a00;
        v += T(0.25) * T(0.75) * T(0.75) * qc.a10;
        v += T(0.75) * T(0.25) * T(0.75) * qc.a01;
        v += T(0.25) * T(0.25) * T(0.75) * qc.a11;
        v += T(0.75) * T(0.75) * T(0.25) * (qn.a00 * cz);
        v += T(0.25) * T(0.75) * T(0.25) * (qn.a10 * cz);
        v += T(0.75) * T(0.25) * T(0.25) * (qn.a01 * cz);
        v += T(0.25) * T(0.25) * T(0.25) * (qn.a11 * cz);
        return v;
    };
    const int fps = fnx * fny, fc = ic + fnx * jc;
    Q qm = ld(K0 - 1), qk = ld(K0), qp = ld(K0 + 1);
    for (int K = K0; K < K1; K++) {
        const Q qn = ld(K + 2);
        const T czl = (K == 0 && !C.per2) ? C.c0u[2][0] : T(1);
        const T czh = (K == C.nz - 1 && !C.per2) ? C.c0u[2][1] : T(1);
        const int c0 = fc + fps * (2 * K), c1 = c0 + fps;
        const T e0 = ef[c0], e1 = ef[c1];
        if (own) {
            ef[c0] = e0 + interp(qk, qm, czl);
            ef[c1] = e1 + interp(qk, qp, czh);
        }
        qm = qk;
        qk = qp;
        qp = qn;
    }
}

// CG bottom solve by one workgroup (all threads of the block call it);
// dots accumulated in double.  The iteration runs in T on the operator L;
// the rhs b comes in and the solution goes out in TB (e_out), so an FP32
// V-cycle can hand its bottom level to an FP64 solve; e is the T-typed
// solution array (e_out itself when T is TB).
template <class T, class TB, int BM, int DIM>
__device__ void bottom_cg(const LevV<T>& L, TB* e_out, const TB* __restrict__ b, T* __restrict__ e,
                          T* __restrict__ r, T* __restrict__ p, T* __restrict__ q) {
    __shared__ double red[1024 / 64];
    __shared__ double bc;
    const int n = L.nx * L.ny * L.nz;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto bdot = [&](const T* u, const T* v) -> double {
        double s = 0.0;
        for (int c = tid; c < n; c += nt) s += (double)u[c] * (double)v[c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        __syncthreads();
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            double z = 0.0;
            for (int w = 0; w < (nt + 63) / 64; w++) z += red[w];
            bc = z;
        }
        __syncthreads();
        return bc;
    };
    for (int c = tid; c < n; c += nt) {
        const T bv = T(b[c]);
        e[c] = T(0);
        r[c] = bv;
        p[c] = bv;
    }
    __syncthreads();
    double rr = bdot(r, r), rr0 = rr;
    const int maxit = 4 * n + 20;
    for (int it = 0; it < maxit; it++) {
        if (rr <= 1e-28 * rr0 || rr == 0.0) break;
        for (int c = tid; c < n; c += nt) {
            int i = c % L.nx, row = c / L.nx;
            int j = row % L.ny, k = row / L.ny;
            T dg, of;
            stencil<T, BM, DIM>(L, p, i, j, k, dg, of);
            q[c] = dg * p[c] - of;
        }
        __syncthreads();
        double pq = bdot(p, q);
        if (pq == 0.0) break;
        double alpha = rr / pq;
        for (int c = tid; c < n; c += nt) {
            e[c] += T(alpha) * p[c];
            r[c] -= T(alpha) * q[c];
        }
        __syncthreads();
        double rn = bdot(r, r);
        double beta = rn / rr;
        for (int c = tid; c < n; c += nt) p[c] = r[c] + T(beta) * p[c];
        __syncthreads();
        rr = rn;
    }
    if constexpr (!std::is_same<T, TB>::value) {
        for (int c = tid; c < n; c += nt) e_out[c] = TB(e[c]);
    }
}


// The coarse end of a V-cycle in one workgroup: from level l0 (whose rhs
// the finer level's restriction already holds) down to the bottom solve and
// back up to l0's corrected e, the operations and order of the per-level
// launches (zero, nu1 red/black sweeps, residual restriction; CG bottom;
// prolongation, nu2 sweeps), with workgroup barriers in place of kernel
// boundaries.  For levels of a few thousand cells, where a launch costs
// more than the work.
constexpr int VC_MAXL = 12;
constexpr int64_t VC_TAIL_CELLS = 4096;
template <class T>
struct VTail {
    LevV<T> v[VC_MAXL];  // levels l0 .. l0+nl-1, the last one the bottom
    T* e[VC_MAXL];
    T* r[VC_MAXL];
    T *cr, *cp, *cq;  // bottom CG scratch
    int off[VC_MAXL];   // first cell of each level in the LDS arrays
    int nl, nu1, nu2;
    // FP64 bottom solve of an FP32 V-cycle: the bottom level's operator
    // coarsened in FP64, and FP64 CG arrays for bottoms above
    // VC_WAVE_BOTTOM cells
    LevV<double> b64;
    double *be, *br, *bp, *bq;
    // every array moved by o bytes (case copies of the arguments)
    __device__ void rebase(int64_t o) {
        for (int l = 0; l < VC_MAXL; l++) {
            v[l].rebase(o);
            e[l] = vd::rb(e[l], o);
            r[l] = vd::rb(r[l], o);
        }
        cr = vd::rb(cr, o);
        cp = vd::rb(cp, o);
        cq = vd::rb(cq, o);
        b64.rebase(o);
        be = vd::rb(be, o);
        br = vd::rb(br, o);
        bp = vd::rb(bp, o);
        bq = vd::rb(bq, o);
    }
};

// f(i, j, k) for i < ni and every (j, k) row of a level, by one workgroup:
// each thread keeps one i and steps through the rows, so the row index is
// divided out once per call rather than once per cell
template <class F>
__device__ __forceinline__ void wg_rows(int ni, int ny, int nz, F f) {
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    if (ni <= nt) {
        const int rstep = nt / ni, i = tid % ni, r0 = tid / ni;
        if (r0 >= rstep) return;
        const int dk = rstep / ny, dj = rstep - dk * ny;
        int j = r0 % ny, k = r0 / ny;
        for (int row = r0; row < ny * nz; row += rstep) {
            f(i, j, k);
            j += dj;
            k += dk;
            if (j >= ny) {
                j -= ny;
                k++;
            }
        }
    } else {
        for (int t = tid; t < ni * ny * nz; t += nt) {
            const int row = t / ni;
            f(t - row * ni, row % ny, row / ny);
        }
    }
}

template <class T, int BM, int DIM>
__device__ __forceinline__ void wg_sweep(const LevV<T>& L, T* __restrict__ x, const T* __restrict__ rhs,
                                         int color) {
    wg_rows((L.nx + 1) / 2, L.ny, L.nz, [&](int ii, int j, int k) {
        const int i = 2 * ii + ((color + j + k) & 1);
        if (i >= L.nx) return;
        T dg, of;
        stencil<T, BM, DIM>(L, x, i, j, k, dg, of);
        const int c = i + L.nx * (j + L.ny * k);
        x[c] = (rhs[c] + of) / dg;
    });
    __syncthreads();
}

// CG bottom solve by the first wave alone for a bottom level of at most
// VC_WAVE_BOTTOM cells: e, r, p and q live in LDS, dots are wave
// reductions, and no workgroup barrier is needed between the phases of an
// iteration (a wave's LDS accesses complete in program order; the
// wavefront fences keep the compiler from reordering them).  The other
// waves wait at the caller's barrier.
constexpr int VC_WAVE_BOTTOM = 1024;
template <class T, class TB, int BM, int DIM>
__device__ void bottom_cg_wave(const LevV<T>& L, TB* __restrict__ e_out, const TB* __restrict__ b, T* lds) {
    const int n = L.nx * L.ny * L.nz;
    const int lane = threadIdx.x;
    T* e = lds;
    T* r = lds + VC_WAVE_BOTTOM;
    T* p = lds + 2 * VC_WAVE_BOTTOM;
    T* q = lds + 3 * VC_WAVE_BOTTOM;
    auto sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto wsum = [](double v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
    double rr = 0.0;
    for (int c = lane; c < n; c += 64) {
        const T bv = T(b[c]);
        e[c] = T(0);
        r[c] = bv;
        p[c] = bv;
        rr += (double)bv * (double)bv;
    }
    rr = wsum(rr);
    sync();
    const double rr0 = rr;
    const int maxit = 4 * n + 20;
    for (int it = 0; it < maxit; it++) {
        if (rr <= 1e-28 * rr0 || rr == 0.0) break;
        double pq = 0.0;
        for (int c = lane; c < n; c += 64) {
            const int i = c % L.nx, row = c / L.nx;
            T dg, of;
            stencil<T, BM, DIM>(L, p, i, row % L.ny, row / L.ny, dg, of);
            const T qv = dg * p[c] - of;
            q[c] = qv;
            pq += (double)p[c] * (double)qv;
        }
        pq = wsum(pq);
        if (pq == 0.0) break;
        const double alpha = rr / pq;
        double rn = 0.0;
        for (int c = lane; c < n; c += 64) {
            e[c] += T(alpha) * p[c];
            const T rv = r[c] - T(alpha) * q[c];
            r[c] = rv;
            rn += (double)rv * (double)rv;
        }
        rn = wsum(rn);
        const double beta = rn / rr;
        sync();
        for (int c = lane; c < n; c += 64) p[c] = r[c] + T(beta) * p[c];
        sync();
        rr = rn;
    }
    for (int c = lane; c < n; c += 64) e_out[c] = TB(e[c]);
}

// the tail levels' e and r live in LDS: levels of at most VC_TAIL_CELLS
// cells, each coarser one at most half as large per dimension, hold fewer
// than 4/3 VC_TAIL_CELLS cells in all (2-D; 8/7 in 3-D)
constexpr int VC_LDS_CELLS = 5464;

// LDS: the tail's levels hold at most VC_LDS_CELLS cells per array (always
// so when they coarsen; a bottom level that cannot coarsen may be larger,
// and then every level stays in memory)
template <class T, int BM, int DIM, bool LDS, bool B64>
__device__ __forceinline__ void vcycle_tail_body(const VTail<T>& a) {
    // with LDS, every sweep, transfer and the bottom solve of the tail read
    // and write LDS; only r of the first level comes from memory and only
    // its e goes back (the next finer level prolongs from it)
    __shared__ T le[LDS ? VC_LDS_CELLS : 1], lr[LDS ? VC_LDS_CELLS : 1];
    auto e = [&](int l) { return LDS ? le + a.off[l] : a.e[l]; };
    auto r = [&](int l) { return LDS ? lr + a.off[l] : a.r[l]; };
    if constexpr (LDS) {
        const int n0 = a.v[0].nx * a.v[0].ny * a.v[0].nz;
        for (int c = threadIdx.x; c < n0; c += blockDim.x) lr[c] = a.r[0][c];
    }
    __syncthreads();
    for (int l = 0; l < a.nl - 1; l++) {
        const LevV<T>& F = a.v[l];
        const int n = F.nx * F.ny * F.nz;
        for (int c = threadIdx.x; c < n; c += blockDim.x) e(l)[c] = T(0);
        __syncthreads();
        for (int s = 0; s < a.nu1; s++) {
            wg_sweep<T, BM, DIM>(F, e(l), r(l), 0);
            wg_sweep<T, BM, DIM>(F, e(l), r(l), 1);
        }
        const LevV<T>& Cv = a.v[l + 1];
        const int nc = Cv.nx * Cv.ny * Cv.nz;
        for (int t = threadIdx.x; t < nc; t += blockDim.x) {
            const int I = t % Cv.nx, row = t / Cv.nx;
            r(l + 1)[t] = resid_restrict_cell<T, BM, DIM>(F, e(l), r(l), I, row % Cv.ny, row / Cv.ny);
        }
        __syncthreads();
    }
    if constexpr (B64) {
        // FP64 bottom solve: the FP32 restriction is the rhs, the FP64
        // solution is rounded once into the FP32 correction
        const LevV<double>& B = a.b64;
        if (B.nx * B.ny * B.nz <= VC_WAVE_BOTTOM) {
            __shared__ double cg[4 * VC_WAVE_BOTTOM];
            if (threadIdx.x < 64) bottom_cg_wave<double, T, BM, DIM>(B, e(a.nl - 1), r(a.nl - 1), cg);
        } else {
            bottom_cg<double, T, BM, DIM>(B, e(a.nl - 1), r(a.nl - 1), a.be, a.br, a.bp, a.bq);
        }
    } else {
        const LevV<T>& B = a.v[a.nl - 1];
        if (B.nx * B.ny * B.nz <= VC_WAVE_BOTTOM) {
            __shared__ T cg[4 * VC_WAVE_BOTTOM];
            if (threadIdx.x < 64) bottom_cg_wave<T, T, BM, DIM>(B, e(a.nl - 1), r(a.nl - 1), cg);
        } else {
            bottom_cg<T, T, BM, DIM>(B, e(a.nl - 1), r(a.nl - 1), e(a.nl - 1), a.cr, a.cp, a.cq);
        }
    }
    __syncthreads();
    for (int l = a.nl - 2; l >= 0; l--) {
        const LevV<T>& F = a.v[l];
        const int n = F.nx * F.ny * F.nz;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = t % F.nx, row = t / F.nx;
            e(l)[t] += prolong_cell<T, DIM>(a.v[l + 1], e(l + 1), i, row % F.ny, row / F.ny);
        }
        __syncthreads();
        for (int s = 0; s < a.nu2; s++) {
            wg_sweep<T, BM, DIM>(F, e(l), r(l), 0);
            wg_sweep<T, BM, DIM>(F, e(l), r(l), 1);
        }
    }
    if constexpr (LDS) {
        const int n0 = a.v[0].nx * a.v[0].ny * a.v[0].nz;
        for (int c = threadIdx.x; c < n0; c += blockDim.x) a.e[0][c] = le[c];
    }
}

// one case: the arguments from the kernel argument block
template <class T, int BM, int DIM, bool LDS, bool B64>
__global__ void __launch_bounds__(1024) k_vcycle_tail(VTail<T> a) {
    vcycle_tail_body<T, BM, DIM, LDS, B64>(a);
}

// a batch of cases: each z-block reads its case's copy of the arguments
// (k_tail_args), whose arrays are already the case's -- the levels are
// indexed at run time, so the views are not moved inside this kernel
template <class T, int BM, int DIM, bool LDS, bool B64>
__global__ void __launch_bounds__(1024) k_vcycle_tail_mem(const VTail<T>* __restrict__ ap, Batch bt) {
    vcycle_tail_body<T, BM, DIM, LDS, B64>(*rb(ap, case_off(bt)));
}

// case c's copy of the tail arguments (case 0's moved by c's offset) into
// its copy of dst
template <class T>
__global__ void k_tail_args(VTail<T> a, VTail<T>* __restrict__ dst, Batch bt) {
    if (threadIdx.x != 0) return;
    const int64_t co = case_off(bt);
    a.rebase(co);
    *rb(dst, co) = a;
}

// coarse face b = average of the covering fine faces
template <class T>
__global__ void k_coarsen_faces(const T* __restrict__ fb, int fnx, int fny, int d, int dim, int cnx,
                                int cny, int cnz, T* __restrict__ cb, Batch bt) {
    fb = rb(fb, case_off(bt));
    cb = rb(cb, case_off(bt));
    int n0 = cnx + (d == 0), n1 = cny + (d == 1), n2 = cnz + (d == 2);
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n0 * n1 * n2) return;
    int i = (int)(t % n0);
    int row = (int)(t / n0);
    int j = row % n1, k = row / n1;
    int b[3] = {2 * i, 2 * j, dim == 3 ? 2 * k : 0};
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int nt2 = dim == 3 ? 2 : 1;
    int f0 = fnx + (d == 0), f1 = fny + (d == 1);
    T s = T(0);
    for (int b2 = 0; b2 < nt2; b2++)
        for (int b1 = 0; b1 < 2; b1++) {
            int f[3] = {b[0], b[1], b[2]};
            f[t1] += b1;
            if (dim == 3) f[t2] += b2;
            s += fb[f[0] + f0 * (f[1] + f1 * f[2])];
        }
    cb[t] = s * (dim == 3 ? T(0.25) : T(0.5));
}

template <class T>
__global__ void k_coarsen_c0(const T* __restrict__ fc, int fnx, int fny, int d, int dim, int cnx,
                             int cny, int cnz, T* __restrict__ cc, Batch bt) {
    fc = rb(fc, case_off(bt));
    cc = rb(cc, case_off(bt));
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int cn[3] = {cnx, cny, cnz}, fn[3] = {fnx, fny, 0};
    int nt1 = cn[t1], nt2 = dim == 3 ? cn[t2] : 1;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt1 * nt2) return;
    int b1 = (int)(t % nt1), b2 = (int)(t / nt1);
    T s = T(0);
    int nk = dim == 3 ? 2 : 1;
    for (int q2 = 0; q2 < nk; q2++)
        for (int q1 = 0; q1 < 2; q1++) {
            int f[3] = {0, 0, 0};
            f[t1] = 2 * b1 + q1;
            if (dim == 3) f[t2] = 2 * b2 + q2;
            int ti = d == 0 ? f[1] + fn[1] * f[2] : (d == 1 ? f[0] + fn[0] * f[2] : f[0] + fn[0] * f[1]);
            s += fc[ti];
        }
    int c[3] = {0, 0, 0};
    c[t1] = b1;
    if (dim == 3) c[t2] = b2;
    int ci = d == 0 ? c[1] + cny * c[2] : (d == 1 ? c[0] + cnx * c[2] : c[0] + cnx * c[1]);
    cc[ci] = s * (dim == 3 ? T(0.25) : T(0.5));
}

template <class A, class B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, int64_t n, Batch bt) {
    a = rb(a, case_off(bt));
    b = rb(b, case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (B)a[t];
}

template <class T>
__global__ void k_zero(T* __restrict__ p, int64_t n, Batch bt) {
    p = rb(p, case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t < n) p[t] = T(0);
}

template <class T>
__global__ void k_axpy_to_double(double* __restrict__ x, const T* __restrict__ e, int64_t n, Batch bt) {
    x = rb(x, case_off(bt));
    e = rb(e, case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t < n) x[t] += (double)e[t];
}


// BC coefficient for one face (ghost = c0*x + cinh)
__device__ __host__ inline void bc_coef(int type, int side, double dx, double val, double ra,
                                        double rb, double rf, double& c0, double& cinh) {
    if (type == VD_DIRCBC) {
        c0 = -1.0;
        cinh = 2.0 * val;
    } else if (type == VD_HNEUBC) {
        c0 = 1.0;
        cinh = 0.0;
    } else if (type == VD_IHNEUBC) {
        c0 = 1.0;
        cinh = (side == 0 ? -dx : dx) * val;
    } else {
        double den = ra / 2.0 + rb / dx;
        c0 = -(ra / 2.0 - rb / dx) / den;
        cinh = rf / den;
    }
}

struct SideArgs {
    int d, side, type, dim;
    int nx, ny, nz;
    double dx, s;      // dx_d, beta/dx_d^2
    const double* bcv;
    const double* ra;
    const double* rb;
    const double* rf;
    const double* bf;  // finest faces of dir d (null => dcell or bconst)
    const double* dcell;  // cell coefficient on the grown box (harmonic face)
    double bconst;
    double* c0;
    double* f;         // rhs to shift
    double* sh;        // or, when not null, the side's cinh stored by face
    __device__ void rebase(int64_t o) {
        sh = vd::rb(sh, o);
        bcv = vd::rb(bcv, o);
        ra = vd::rb(ra, o);
        rb = vd::rb(rb, o);
        rf = vd::rb(rf, o);
        bf = vd::rb(bf, o);
        dcell = vd::rb(dcell, o);
        c0 = vd::rb(c0, o);
        f = vd::rb(f, o);
    }
};

// finest-level c0 and rhs shift s*b*cinh for one boundary side
// both sides of one direction in one launch (blockIdx.y = side): they
// touch different cells, so their f updates do not meet
__global__ void k_side_setup(SideArgs a0, SideArgs a1, Batch bt) {
    SideArgs a = blockIdx.y ? a1 : a0;
    a.rebase(case_off(bt));
    int t1 = a.d == 0 ? 1 : 0, t2 = a.d == 2 ? 1 : 2;
    int n[3] = {a.nx, a.ny, a.nz};
    int nt1 = n[t1], nt2 = a.dim == 3 ? n[t2] : 1;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt1 * nt2) return;
    int b1 = (int)(t % nt1), b2 = (int)(t / nt1);
    int c[3] = {0, 0, 0};
    c[t1] = b1;
    if (a.dim == 3) c[t2] = b2;
    c[a.d] = a.side ? n[a.d] - 1 : 0;
    int ti = a.d == 0 ? c[1] + a.ny * c[2] : (a.d == 1 ? c[0] + a.nx * c[2] : c[0] + a.nx * c[1]);
    double val = a.bcv ? a.bcv[ti] : 0.0;
    double ra = 0.0, rb = 1.0, rf = 0.0;
    if (a.type == VD_ROBINBC && a.ra) {
        ra = a.ra[ti];
        rb = a.rb[ti];
        rf = a.rf[ti];
    }
    double c0, cinh;
    bc_coef(a.type, a.side, a.dx, val, ra, rb, rf, c0, cinh);
    a.c0[ti] = c0;
    int fc[3] = {c[0], c[1], c[2]};
    fc[a.d] += a.side;
    int n0 = a.nx + (a.d == 0), n1 = a.ny + (a.d == 1);
    double b;
    if (a.bf) {
        b = a.bf[fc[0] + n0 * (fc[1] + n1 * fc[2])];
    } else if (a.dcell) {
        const int gx = a.nx + 2, gxy = gx * (a.ny + 2);
        const int gs = a.d == 0 ? 1 : (a.d == 1 ? gx : gxy);
        const int gc = (fc[0] + 1) + gx * (fc[1] + 1) + (a.dim == 3 ? gxy * (fc[2] + 1) : 0);
        b = harm(a.dcell[gc - gs], a.dcell[gc]);
    } else {
        b = a.bconst;
    }
    if (a.sh) {
        // the rhs pass adds s * b * cinh itself (b is the constant bconst)
        a.sh[ti] = cinh;
    } else {
        const int cell = c[0] + a.nx * (c[1] + a.ny * c[2]);
        a.f[cell] = fma(a.s * b, cinh, a.f[cell]);
    }
}

// ------------------------------------------------------------ host side
template <class T>
static LevV<T> view(const MGWork& W, const Lev<T>& L, double alpha, double beta, double bconst) {
    LevV<T> v;
    v.nx = L.n[0];
    v.ny = L.n[1];
    v.nz = L.n[2];
    double s[3];
    for (int d = 0; d < 3; d++) s[d] = d < W.dim ? beta / (L.dx[d] * L.dx[d]) : 0.0;
    v.s0 = (T)s[0];
    v.s1 = (T)s[1];
    v.s2 = (T)s[2];
    v.alpha = (T)alpha;
    v.bconst = (T)bconst;
    v.bf0 = L.bf[0];
    v.bf1 = L.bf[1];
    v.bf2 = L.bf[2];
    v.dcell = nullptr;
    for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
            v.c0[d][sd] = L.c0[d][sd];
            v.c0u[d][sd] = (T)W.c0u[d][sd];
        }
    v.c0uni = W.c0uni;
    v.per0 = W.per[0];
    v.per1 = W.per[1];
    v.per2 = W.per[2];
    return v;
}

// planes per block of the fused smoother: 32-plane chunks when that gives
// at least half the CUs a block, else 16-plane chunks (more plane halo)
// when that fills every CU; 0 when the level is too small for the fused
// kernel to pay (per-colour launches are used).  Measured at 64^3 .. 256^3.
// VD_FUSED_ZC overrides.
static int fused_zc(const int n[3]) {
    const char* e = getenv("VD_FUSED_ZC");
    if (e && atoi(e) > 0) return std::min(atoi(e), n[2]);
    const int64_t tiles = (int64_t)((n[0] + FZ_IX - 1) / FZ_IX) * ((n[1] + FZ_IY - 1) / FZ_IY);
    if (tiles * ((n[2] + 31) / 32) >= 128) return 32;
    if (tiles * ((n[2] + 15) / 16) >= 256) return 16;
    return 0;
}

// plane chunk of the fused smoother on a coarse level of a V-cycle
// (VD_FUSED_COARSE=1; off by default): the largest of 32 / 16 planes that
// still gives 256 blocks, else 8 planes.  One launch of two red-black
// sweeps replaces four colour launches, but with few blocks per level its
// plane loop takes longer than the launches it saves (FP64 256^3
// headline: coarse legs 1.81 -> 1.90 ms per step).  0 when the level
// cannot take it (the conditions of fused_capable; levels of the
// one-workgroup coarse tail never sweep here).
static int fused_zc_coarse(const MGWork& W, const int n[3]) {
    const char* e = getenv("VD_FUSED_COARSE");
    if (!(e && strcmp(e, "1") == 0)) return 0;
    const char* f = getenv("VD_FUSED_SMOOTHER");
    if (f && strcmp(f, "0") == 0) return 0;
    if (W.dim != 3 || W.per[2] || n[0] % 2 || n[1] % 2 || n[0] < 8 || n[1] < 8 || n[2] < 2 ||
        ncell(n) <= VC_TAIL_CELLS)
        return 0;
    const int64_t tiles = (int64_t)((n[0] + FZ_IX - 1) / FZ_IX) * ((n[1] + FZ_IY - 1) / FZ_IY);
    if (tiles * ((n[2] + 31) / 32) >= 256) return 32;
    if (tiles * ((n[2] + 15) / 16) >= 256) return 16;
    return std::min(8, n[2]);
}

// finest levels the fused smoother handles: 3-D, z not periodic, even
// nx and ny, 32-bit byte offsets, enough blocks; VD_FUSED_SMOOTHER=0
// turns it off
static bool fused_capable(const MGWork& W) {
    const char* e = getenv("VD_FUSED_SMOOTHER");  // read when a workspace is created
    const bool env = !(e && strcmp(e, "0") == 0);
    return env && W.dim == 3 && !W.per[2] && W.n[0] % 2 == 0 && W.n[1] % 2 == 0 && W.n[0] >= 8 &&
           W.n[1] >= 8 && W.n[2] >= 2 && ncell(W.n) * 8 < ((int64_t)1 << 31) && fused_zc(W.n) > 0;
}

// the smoother-only solves on the cell coefficient (bm 2) run one fused
// sweep-plus-residual launch per iteration when the coefficient's grown box
// also has 32-bit byte offsets; VD_FUSED_CELLCOEF=0 keeps them on the
// per-colour launches and the streamed residual
static bool fused_cellcoef(const MGWork& W) {
    const char* e = getenv("VD_FUSED_CELLCOEF");
    if (e && strcmp(e, "0") == 0) return false;
    const int64_t ng = (int64_t)(W.n[0] + 2) * (W.n[1] + 2) * (W.n[2] + 2);
    return W.dim == 3 && ng * 8 < ((int64_t)1 << 31);
}

// plane chunk of the cell-coefficient sweep-plus-residual pass: the
// longest (at most 128 planes, at least 8) that still gives the 256 CUs a
// block each -- its blocks run one per CU, and each chunk pays three halo
// planes.  256^3 FP64 headline, 20 steps: 32-plane chunks 15.53 ms per
// step, 128-plane chunks 15.38.  VD_FUSED_CC_ZC overrides.
static int fused_zc_cc(const int n[3]) {
    const char* e = getenv("VD_FUSED_CC_ZC");
    if (e && atoi(e) > 0) return std::min(atoi(e), n[2]);
    constexpr int TIX = FzTile<2>::IX;
    const int64_t tiles = (int64_t)((n[0] + TIX - 1) / TIX) * ((n[1] + FZ_IY - 1) / FZ_IY);
    int zc = 128;
    while (zc > 8 && tiles * ((n[2] + zc - 1) / zc) < 256) zc /= 2;
    return std::min(zc, n[2]);
}

template <class T, int BM, int S, int RES>
static void launch_fused_k(vd_ctx* ctx, const LevV<T>& v, int zc, const T* xin, T* xout, const T* rhs,
                           double* part, int* nblk, double* xacc = nullptr, const SweepOut* so = nullptr) {
    // the kernel reads xin unless it starts from zero (ZI)
    constexpr int TIX = FzTile<BM>::IX, TT = FzTile<BM>::THREADS;
    const int ntx = (v.nx + TIX - 1) / TIX, nty = (v.ny + FZ_IY - 1) / FZ_IY;
    const int nzc = (v.nz + zc - 1) / zc;
    const int nb = ntx * nty * nzc;
    if (nblk) *nblk = nb;
    const bool bt = ctx->ncase > 1;
    const SweepOut sw = so ? *so : SweepOut{};
#define VD_FUSED_LAUNCH_P(ZIV, ACCV, XACC, POSTV)                                                                     \
    {                                                                                                                 \
        if (bt)                                                                                                       \
            hipLaunchKernelGGL((k_gsrb_fused<T, BM, S, RES, ZIV, ACCV, true, POSTV>), bz(ctx, dim3((unsigned)nb)),    \
                               dim3(TT), 0, ctx->stream, v, xin, xout, rhs, zc, ntx, nty, part, XACC, sw, ctx->bt);    \
        else                                                                                                          \
            hipLaunchKernelGGL((k_gsrb_fused<T, BM, S, RES, ZIV, ACCV, false, POSTV>), dim3((unsigned)nb), dim3(TT),  \
                               0, ctx->stream, v, xin, xout, rhs, zc, ntx, nty, part, XACC, sw, ctx->bt);             \
    }
#define VD_FUSED_LAUNCH(ZIV, ACCV, XACC) VD_FUSED_LAUNCH_P(ZIV, ACCV, XACC, 0)
    if constexpr (std::is_same<T, double>::value && BM == 1 && S == 2 && RES == 1) {
        if (xin && so && so->p) {
            VD_FUSED_LAUNCH_P(0, 0, (double*)nullptr, 1)
            return;
        }
    }
    if constexpr (RES == 0) {
        if (xin && xacc) {
            VD_FUSED_LAUNCH(0, 1, xacc)
            return;
        }
    }
    if (xin)
        VD_FUSED_LAUNCH(0, 0, (double*)nullptr)
    else if constexpr (RES == 0)
        VD_FUSED_LAUNCH(1, 0, (double*)nullptr)
#undef VD_FUSED_LAUNCH
#undef VD_FUSED_LAUNCH_P
}

// nsweep (1 or 2) red-black sweeps of x (zeros when xin is null) into
// xout in one launch, for face-array (bm 0) or constant (bm 1) coefficients;
// with res (FP64, one sweep) also the per-block
// max-norm of the residual of the result into part[0..*nblk)
// zc: the level's plane chunk (Lev::fzc); only finest-level launches are
// timed as the VD_KT_FUSED family
template <class T>
static void launch_fused(vd_ctx* ctx, const MGWork& W, int bm, const LevV<T>& v, int zc, const T* xin, T* xout,
                         const T* rhs, int nsweep, int res, double* part, int* nblk, double* xacc = nullptr,
                         const SweepOut* so = nullptr) {
    const bool fine = v.nx == W.n[0] && v.ny == W.n[1] && v.nz == W.n[2];
    if (fine) kt_begin(ctx, VD_KT_FUSED);
#define VD_FUSED_BM(BMV)                                                                          \
    if (nsweep == 2) launch_fused_k<T, BMV, 4, 0>(ctx, v, zc, xin, xout, rhs, part, nblk, xacc);     \
    else if constexpr (std::is_same<T, double>::value) {                                          \
        if (res) launch_fused_k<T, BMV, 2, 1>(ctx, v, zc, xin, xout, rhs, part, nblk, nullptr, so); \
        else launch_fused_k<T, BMV, 2, 0>(ctx, v, zc, xin, xout, rhs, part, nblk, xacc);        \
    } else launch_fused_k<T, BMV, 2, 0>(ctx, v, zc, xin, xout, rhs, part, nblk, xacc);
    if (bm == 1) {
        VD_FUSED_BM(1)
    } else if (bm == 2) {
        // the cell coefficient: one FP64 sweep and its residual norm only
        if constexpr (std::is_same<T, double>::value)
            if (nsweep == 1 && res && xin && !xacc) launch_fused_k<T, 2, 2, 1>(ctx, v, zc, xin, xout, rhs, part, nblk);
    } else {
        VD_FUSED_BM(0)
    }
#undef VD_FUSED_BM
    const int nb = bm == 1 ? 0 : (bm == 0 ? W.dim : 1);
    // one pass: read x (unless zero), rhs and the coefficient data; write x,
    // or read-modify-write the FP64 accumulation target
    const double N = (double)v.nx * v.ny * v.nz;
    if (fine)
        kt_end(ctx, VD_KT_FUSED,
               sizeof(T) * N * ((xin ? 1 : 0) + 1 + nb) + (xacc ? 16.0 * N : (double)sizeof(T) * N) +
                   (so && so->p && bm == 1 && res ? 8.0 * N : 0.0));
}

template <class T>
static int alloc_levels(MGWork* W, std::vector<Lev<T>>& L, bool share0) {
    L.resize(W->nlev);
    int n[3] = {W->n[0], W->n[1], W->n[2]};
    double dx[3] = {W->dx[0], W->dx[1], W->dx[2]};
    for (int l = 0; l < W->nlev; l++) {
        Lev<T>& v = L[l];
        for (int d = 0; d < 3; d++) {
            v.n[d] = n[d];
            v.dx[d] = dx[d];
        }
        int64_t nc = ncell(n);
        void* p;
        VD_TRY(dalloc(W, nc * sizeof(T), &p));
        v.e = (T*)p;
        v.fzc = l == 0 ? (fused_capable(*W) ? fused_zc(n) : 0) : fused_zc_coarse(*W, n);
        if (v.fzc > 0) {
            VD_TRY(dalloc(W, nc * sizeof(T), &p));
            v.e2 = (T*)p;
        }
        if (!(share0 && l == 0)) {
            VD_TRY(dalloc(W, nc * sizeof(T), &p));
            v.r = (T*)p;
        }
        if (l == W->nlev - 1) {
            VD_TRY(dalloc(W, nc * sizeof(T), &p));
            v.p = (T*)p;
            VD_TRY(dalloc(W, nc * sizeof(T), &p));
            v.q = (T*)p;
            VD_TRY(dalloc(W, nc * sizeof(T), &p));
            v.w = (T*)p;
        }
        for (int d = 0; d < W->dim; d++) {
            if (!W->constb && !(share0 && l == 0)) {
                VD_TRY(dalloc(W, nfacef(n, d) * sizeof(T), &p));
                v.bf[d] = (T*)p;
            }
            if (!W->per[d] && !(share0 && l == 0))
                for (int s = 0; s < 2; s++) {
                    VD_TRY(dalloc(W, ntanf(n, d) * sizeof(T), &p));
                    v.c0[d][s] = (T*)p;
                }
        }
        for (int d = 0; d < W->dim; d++) {
            n[d] /= 2;
            dx[d] *= 2.0;
        }
    }
    return VD_OK;
}

// the FP64 coefficient chain of an FP32 hierarchy (MGWork::C64): level 0
// views the FP64 finest data, levels 1 .. nlev-1 hold face b (variable
// coefficients) and boundary c0 (non-periodic sides), the bottom level
// also the FP64 CG arrays
static int alloc_coef64(MGWork* W) {
    std::vector<Lev<double>>& L = W->C64;
    L.resize(W->nlev);
    int n[3] = {W->n[0], W->n[1], W->n[2]};
    double dx[3] = {W->dx[0], W->dx[1], W->dx[2]};
    void* p;
    for (int l = 0; l < W->nlev; l++) {
        Lev<double>& v = L[l];
        for (int d = 0; d < 3; d++) {
            v.n[d] = n[d];
            v.dx[d] = dx[d];
        }
        for (int d = 0; d < W->dim; d++) {
            if (l == 0) {
                v.bf[d] = W->bf0[d];
                for (int sd = 0; sd < 2; sd++) v.c0[d][sd] = W->c00[d][sd];
                continue;
            }
            if (!W->constb) {
                VD_TRY(dalloc(W, nfacef(n, d) * sizeof(double), &p));
                v.bf[d] = (double*)p;
            }
            if (!W->per[d])
                for (int sd = 0; sd < 2; sd++) {
                    VD_TRY(dalloc(W, ntanf(n, d) * sizeof(double), &p));
                    v.c0[d][sd] = (double*)p;
                }
        }
        if (l == W->nlev - 1) {
            const int64_t nc = ncell(n);
            double** arr[4] = {&v.e, &v.r, &v.p, &v.q};
            for (double** a : arr) {
                VD_TRY(dalloc(W, nc * sizeof(double), &p));
                *a = (double*)p;
            }
        }
        for (int d = 0; d < W->dim; d++) {
            n[d] /= 2;
            dx[d] *= 2.0;
        }
    }
    return VD_OK;
}

// workspace for a shape / precision / coefficient kind, cached in the ctx
static int get_work(vd_ctx* ctx, const vd_geom* g, const vd_box* box, int prec, int constb,
                    int maxc, MGWork** out) {
    int n[3], per[3];
    for (int d = 0; d < 3; d++) {
        n[d] = d < g->dim ? box->hi[d] - box->lo[d] + 1 : 1;
        per[d] = d < g->dim ? g->is_periodic[d] : 0;
    }
    {
        // the workspace arrays hold the level with a ghost layer (the cell
        // coefficient), indexed with 32-bit offsets
        int64_t grown = 1;
        for (int d = 0; d < g->dim; d++) grown *= (int64_t)n[d] + 2;
        if (grown > MAX_CELLS_32) return extent_error("multigrid workspace", "level with its ghost cells", grown);
    }
    char key[160];
    snprintf(key, sizeof(key), "%d:%dx%dx%d:p%d%d%d:prec%d:cb%d:mc%d:dx%.17g,%.17g,%.17g", g->dim, n[0],
             n[1], n[2], per[0], per[1], per[2], prec, constb, maxc, g->dx[0], g->dx[1], g->dx[2]);
    auto it = ctx->mg.find(key);
    if (it != ctx->mg.end()) {
        *out = it->second;
        return VD_OK;
    }
    MGWork* W = new MGWork();
    W->ctx = ctx;
    W->dim = g->dim;
    W->prec = prec;
    W->constb = constb;
    {
        const char* e = getenv("VD_STREAM_SWEEPS");
        W->stream = !(e && strcmp(e, "0") == 0);
    }
    for (int d = 0; d < 3; d++) {
        W->n[d] = n[d];
        W->per[d] = per[d];
        W->dx[d] = d < g->dim ? g->dx[d] : 1.0;
        for (int s = 0; s < 2; s++)
            for (int q = 0; q < 3; q++) W->rob[d][s][q] = nullptr;
    }
    // coarsening: every dimension even and >= 4
    int nl = 1, m[3] = {n[0], n[1], n[2]};
    while (nl - 1 < maxc) {
        bool ok = true;
        for (int d = 0; d < g->dim; d++)
            if (m[d] % 2 != 0 || m[d] < 4) ok = false;
        if (!ok) break;
        for (int d = 0; d < g->dim; d++) m[d] /= 2;
        nl++;
    }
    W->nlev = nl;
    int64_t nc = ncell(n);
    void* p;
    int st = VD_OK;
    if ((st = dalloc(W, nc * 8, &p))) goto fail;
    W->x = (double*)p;
    if (fused_capable(*W)) {
        if ((st = dalloc(W, nc * 8, &p))) goto fail;
        W->x2 = (double*)p;
    }
    if ((st = dalloc(W, nc * 8, &p))) goto fail;
    W->f = (double*)p;
    if ((st = dalloc(W, nc * 8, &p))) goto fail;
    W->res = (double*)p;
    if ((st = dalloc(W, 8 * NSLOT * SLOT_STRIDE * sizeof(unsigned long long), &p))) goto fail;
    W->dnorm = (unsigned long long*)p;
    {
        // the norm slots start at zero, every case's (not only the active
        // ones'); k_norm_out clears what it reads
        int all[VD_MAX_BATCH];
        for (int c = 0; c < ctx->ncase; c++) all[c] = c;
        BatchScope every(ctx, all, ctx->ncase);
        hipLaunchKernelGGL(k_zero<unsigned long long>, bz(ctx, grid1d(8 * NSLOT * SLOT_STRIDE, 256)), dim3(256), 0,
                           ctx->stream, W->dnorm, (int64_t)8 * NSLOT * SLOT_STRIDE, ctx->bt);
    }
    {
        // per-block partial norms of k_resid0 and of the fused smoother
        int64_t np = ((n[0] + 255) / 256) * (int64_t)n[1] * n[2];
        constexpr int TIX = FzTile<2>::IX;  // the narrowest tile
        int64_t nf = (int64_t)((n[0] + TIX - 1) / TIX) * ((n[1] + FZ_IY - 1) / FZ_IY) * n[2];
        if ((st = dalloc(W, std::max(np, nf) * 8, &p))) goto fail;
    }
    W->part = (double*)p;
    if (!constb) {
        int64_t ng = (int64_t)(n[0] + 2) * (n[1] + 2) * (g->dim == 3 ? n[2] + 2 : 1);
        if ((st = dalloc(W, ng * 8, &p))) goto fail;
        W->dcell = (double*)p;
    }
    for (int d = 0; d < g->dim; d++) {
        if (!constb) {
            if ((st = dalloc(W, nfacef(n, d) * 8, &p))) goto fail;
            W->bf0[d] = (double*)p;
        }
        if (!per[d])
            for (int s = 0; s < 2; s++) {
                if ((st = dalloc(W, ntanf(n, d) * 8, &p))) goto fail;
                W->c00[d][s] = (double*)p;
                if ((st = dalloc(W, ntanf(n, d) * 8, &p))) goto fail;
                W->bcv[d][s] = (double*)p;
                if ((st = dalloc(W, ntanf(n, d) * 8, &p))) goto fail;
                W->sh[d][s] = (double*)p;
                for (int q = 0; q < 3; q++) {
                    if ((st = dalloc(W, ntanf(n, d) * 8, &p))) goto fail;
                    W->rob[d][s][q] = (double*)p;
                }
            }
    }
    if (ctx->ncase > 1 &&
        (st = dalloc(W, std::max(sizeof(VTail<double>), sizeof(VTail<float>)), &W->tail_dev)))
        goto fail;
    if (prec == 0) {
        if ((st = alloc_levels<double>(W, W->L64, true))) goto fail;
        W->L64[0].r = W->res;
        for (int d = 0; d < g->dim; d++) {
            W->L64[0].bf[d] = W->bf0[d];
            for (int s = 0; s < 2; s++) W->L64[0].c0[d][s] = W->c00[d][s];
        }
    } else {
        if ((st = alloc_levels<float>(W, W->L32, false))) goto fail;
        if ((st = alloc_coef64(W))) goto fail;
    }
    ctx->mg[key] = W;
    *out = W;
    return VD_OK;
fail:
    mgwork_free(W);
    return st;
}

// levels swept by k_gsrb_z (which can also start a sweep from zero): 3-D,
// even nx, streamed kernels on for the workspace
static bool streamed_sweep(const MGWork& W, int nx) { return W.dim == 3 && nx % 2 == 0 && W.stream; }

template <class T, int BM>
static void launch_gsrb_bm(vd_ctx* ctx, const MGWork& W, const LevV<T>& v, T* x, const T* rhs, int color,
                           bool zi = false) {
    const int dim = W.dim;
    if (streamed_sweep(W, v.nx)) {
        const int ntx = (v.nx / 2 + GZ_X - 1) / GZ_X, nty = (v.ny + GZ_Y - 1) / GZ_Y;
        const int n3[3] = {v.nx, v.ny, v.nz};
#define VD_GZ(C0UV, ZIV)                                                                                      \
    do {                                                                                                      \
        const int zc = stream_zc(ctx, (const void*)&k_gsrb_z<T, BM, C0UV, ZIV>, GZ_X * GZ_Y, n3, ntx * nty);   \
        const dim3 g((unsigned)(ntx * nty * ((v.nz + zc - 1) / zc)));                                        \
        if (ctx->ncase > 1)                                                                                   \
            hipLaunchKernelGGL((k_gsrb_z<T, BM, C0UV, ZIV, true>), bz(ctx, g), dim3(GZ_X * GZ_Y), 0, ctx->stream, v, \
                               x, rhs, color, zc, ntx, nty, ctx->bt);                                          \
        else                                                                                                  \
            hipLaunchKernelGGL((k_gsrb_z<T, BM, C0UV, ZIV, false>), g, dim3(GZ_X * GZ_Y), 0, ctx->stream, v, x, rhs, \
                               color, zc, ntx, nty, ctx->bt);                                                  \
    } while (0)
        if (v.c0uni) {
            if (zi) VD_GZ(true, true); else VD_GZ(true, false);
        } else {
            if (zi) VD_GZ(false, true); else VD_GZ(false, false);
        }
#undef VD_GZ
        return;
    }
    const int nh = (v.nx + 1) / 2;
    const int bx = nh > 64 ? 128 : 64;
    const int gx = (nh + bx - 1) / bx;
    dim3 g((unsigned)gx * v.ny * v.nz), b(bx);
    if (dim == 3) hipLaunchKernelGGL((k_gsrb<T, BM, 3>), bz(ctx, g), b, 0, ctx->stream, v, x, rhs, color, gx, ctx->bt);
    else hipLaunchKernelGGL((k_gsrb<T, BM, 2>), bz(ctx, g), b, 0, ctx->stream, v, x, rhs, color, gx, ctx->bt);
}

// bm: 0 face arrays, 1 constant, 2 cell coefficient (harmonic on the fly)
// zi: x is taken as 0 (streamed levels only: streamed_sweep)
template <class T>
static void launch_gsrb(vd_ctx* ctx, const MGWork& W, int bm, const LevV<T>& v, T* x, const T* rhs,
                        int color, bool fine, bool zi = false) {
    if (fine) kt_begin(ctx, VD_KT_GSRB);
    if (bm == 1) launch_gsrb_bm<T, 1>(ctx, W, v, x, rhs, color, zi);
    else if (bm == 0) launch_gsrb_bm<T, 0>(ctx, W, v, x, rhs, color, zi);
    else launch_gsrb_bm<T, 2>(ctx, W, v, x, rhs, color, zi);
    // one colour = half of a sweep that reads phi, rhs and the coefficient
    // data (none, DIM face arrays, or one cell array) and writes phi
    const int nb = bm == 1 ? 0 : (bm == 0 ? W.dim : 1);
    if (fine) kt_end(ctx, VD_KT_GSRB, 8.0 * (double)v.nx * v.ny * v.nz * (3 + nb) / 2.0);
}

template <class T, int BM>
static void launch_resid_restrict(vd_ctx* ctx, const MGWork& W, const LevV<T>& v, const Lev<T>& F, Lev<T>& Cc) {
    const int dim = W.dim;
    if (dim == 3 && W.stream) {
        const int ntx = (Cc.n[0] + RR_X - 1) / RR_X, nty = (Cc.n[1] + RR_Y - 1) / RR_Y;
        const void* kfn = v.c0uni ? (const void*)&k_resid_restrict_z<T, BM, true>
                                  : (const void*)&k_resid_restrict_z<T, BM, false>;
        const int zc = stream_zc(ctx, kfn, RR_X * RR_Y, Cc.n, ntx * nty);
        const dim3 g((unsigned)(ntx * nty * ((Cc.n[2] + zc - 1) / zc)));
#define VD_RRZ(C0UV, BTV)                                                                                          \
    hipLaunchKernelGGL((k_resid_restrict_z<T, BM, C0UV, BTV>), bz(ctx, g), dim3(RR_X * RR_Y), 0, ctx->stream, v, F.e, \
                       F.r, Cc.n[0], Cc.n[1], Cc.n[2], Cc.r, zc, ntx, nty, ctx->bt)
        if (ctx->ncase > 1) {
            if (v.c0uni) VD_RRZ(true, true);
            else VD_RRZ(false, true);
        } else {
            if (v.c0uni) VD_RRZ(true, false);
            else VD_RRZ(false, false);
        }
#undef VD_RRZ
        return;
    }
    const int bx = Cc.n[0] > 64 ? 128 : 64;
    const int gx = (Cc.n[0] + bx - 1) / bx;
    dim3 g((unsigned)gx * Cc.n[1] * Cc.n[2]), b(bx);
#define VD_RR(DIMV, BTV)                                                                                        \
    hipLaunchKernelGGL((k_resid_restrict<T, BM, DIMV, BTV>), bz(ctx, g), b, 0, ctx->stream, v, F.e, F.r, Cc.n[0], \
                       Cc.n[1], Cc.n[2], Cc.r, gx, ctx->bt)
    if (ctx->ncase > 1) {
        if (dim == 3) VD_RR(3, true);
        else VD_RR(2, true);
    } else {
        if (dim == 3) VD_RR(3, false);
        else VD_RR(2, false);
    }
#undef VD_RR
}


// first level of the coarse tail (levels l0 .. nlev-1 run in one
// workgroup launch)
template <class T>
static int vc_tail_level(const MGWork& W, const std::vector<Lev<T>>& L) {
    const int nl = W.nlev;
    int l0 = nl - 1;
    while (l0 > 0 && ncell(L[l0 - 1].n) <= VC_TAIL_CELLS && nl - l0 < VC_MAXL) l0--;
    return l0;
}

// the down leg on level l: pre-smooth from a zero correction, then the
// residual restricted to level l+1
template <class T>
static void vc_down(vd_ctx* ctx, const MGWork& W, int bm, std::vector<Lev<T>>& L, double alpha, double beta,
                    double bconst, int nu1, int l) {
    Lev<T>& F = L[l];
    LevV<T> v = view<T>(W, F, alpha, beta, bconst);
    int64_t nc = ncell(F.n);
    if (F.e2 && W.c0uni && bm != 2) {
        // from a zero correction: the first launch writes e without reading it
        for (int s = 0; s < nu1;) {
            const int k = nu1 - s >= 2 ? 2 : 1;
            if (s == 0) {
                launch_fused<T>(ctx, W, bm, v, F.fzc, nullptr, F.e, F.r, k, 0, nullptr, nullptr);
            } else {
                launch_fused<T>(ctx, W, bm, v, F.fzc, F.e, F.e2, F.r, k, 0, nullptr, nullptr);
                std::swap(F.e, F.e2);
            }
            s += k;
        }
        if (nu1 == 0) hipLaunchKernelGGL(k_zero<T>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, F.e, nc, ctx->bt);
    } else {
        // from a zero correction: the first colour writes e without reading it
        const bool zi = nu1 > 0 && streamed_sweep(W, F.n[0]);
        if (!zi) hipLaunchKernelGGL(k_zero<T>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, F.e, nc, ctx->bt);
        for (int s = 0; s < nu1; s++) {
            launch_gsrb<T>(ctx, W, bm, v, F.e, F.r, 0, l == 0, zi && s == 0);
            launch_gsrb<T>(ctx, W, bm, v, F.e, F.r, 1, l == 0);
        }
    }
    if (bm == 1) launch_resid_restrict<T, 1>(ctx, W, v, F, L[l + 1]);
    else launch_resid_restrict<T, 0>(ctx, W, v, F, L[l + 1]);
}

// the arguments of the coarse tail (levels l0 .. nlev-1); *lds: whether
// its levels fit in LDS
template <class T>
static VTail<T> tail_args(const MGWork& W, std::vector<Lev<T>>& L, double alpha, double beta, double bconst,
                          int nu1, int nu2, int l0, bool* lds) {
    const int nl = W.nlev;
    VTail<T> a;
    memset((void*)&a, 0, sizeof(a));  // defined bytes: the batched path compares them
    for (int l = l0; l < nl; l++) {
        a.v[l - l0] = view<T>(W, L[l], alpha, beta, bconst);
        a.e[l - l0] = L[l].e;
        a.r[l - l0] = L[l].r;
    }
    Lev<T>& B = L[nl - 1];
    a.cr = B.w;
    a.cp = B.p;
    a.cq = B.q;
    a.nl = nl - l0;
    int64_t tot = 0;
    for (int l = 0; l < a.nl; l++) {
        a.off[l] = (int)tot;
        tot += ncell(L[l0 + l].n);
    }
    *lds = tot <= VC_LDS_CELLS;
    a.nu1 = nu1;
    a.nu2 = nu2;
    a.be = a.br = a.bp = a.bq = nullptr;
    if (std::is_same<T, float>::value && !W.C64.empty()) {
        const Lev<double>& B64 = W.C64[nl - 1];
        a.b64 = view<double>(W, B64, alpha, beta, bconst);
        a.be = B64.e;
        a.br = B64.r;
        a.bp = B64.p;
        a.bq = B64.q;
    }
    return a;
}

// levels l0 .. nlev-1 (and the bottom solve) in one workgroup
template <class T>
static void vc_tail(vd_ctx* ctx, const MGWork& W, int bm, std::vector<Lev<T>>& L, double alpha, double beta,
                    double bconst, int nu1, int nu2, int l0) {
    bool lds;
    const VTail<T> a = tail_args<T>(W, L, alpha, beta, bconst, nu1, nu2, l0, &lds);
    // FP32 V-cycles solve their bottom level in FP64
    constexpr bool b64 = std::is_same<T, float>::value;
    const VTail<T>* ap = nullptr;
    if (ctx->ncase > 1) {
        // a batched context: the cases' own copies of the arguments, made
        // where they are missing or were made from other arguments
        const unsigned char* kb = (const unsigned char*)&a;
        if (W.tail_key.size() != sizeof(a) || memcmp(W.tail_key.data(), kb, sizeof(a)) != 0) {
            W.tail_key.assign(kb, kb + sizeof(a));
            W.tail_mask = 0;
        }
        int miss[VD_MAX_BATCH], nmiss = 0;
        for (int z = 0; z < ctx->bn; z++)
            if (!((W.tail_mask >> ctx->bt.idx[z]) & 1u)) miss[nmiss++] = ctx->bt.idx[z];
        if (nmiss > 0) {
            BatchScope sc(ctx, miss, nmiss);
            hipLaunchKernelGGL(k_tail_args<T>, bz(ctx, dim3(1)), dim3(64), 0, ctx->stream, a, (VTail<T>*)W.tail_dev,
                               ctx->bt);
            for (int q = 0; q < nmiss; q++) W.tail_mask |= 1u << miss[q];
        }
        ap = (const VTail<T>*)W.tail_dev;
    }
#define VD_TAIL(BMV, DIMV)                                                                                        \
    {                                                                                                             \
        if (ap && lds)                                                                                            \
            hipLaunchKernelGGL((k_vcycle_tail_mem<T, BMV, DIMV, true, b64>), bz(ctx, dim3(1)), dim3(1024), 0,    \
                               ctx->stream, ap, ctx->bt);                                                         \
        else if (ap)                                                                                              \
            hipLaunchKernelGGL((k_vcycle_tail_mem<T, BMV, DIMV, false, b64>), bz(ctx, dim3(1)), dim3(1024), 0,   \
                               ctx->stream, ap, ctx->bt);                                                         \
        else if (lds)                                                                                             \
            hipLaunchKernelGGL((k_vcycle_tail<T, BMV, DIMV, true, b64>), dim3(1), dim3(1024), 0, ctx->stream, a); \
        else                                                                                                      \
            hipLaunchKernelGGL((k_vcycle_tail<T, BMV, DIMV, false, b64>), dim3(1), dim3(1024), 0, ctx->stream, a);\
    }
    if (bm == 1) {
        if (W.dim == 3) VD_TAIL(1, 3) else VD_TAIL(1, 2)
    } else {
        if (W.dim == 3) VD_TAIL(0, 3) else VD_TAIL(0, 2)
    }
#undef VD_TAIL
}

// the up leg on level l: the correction of level l+1 prolonged into e, then
// the post-smooth (on the finest level with xacc, x += e in its last pass);
// returns whether x was updated
template <class T>
static bool vc_up(vd_ctx* ctx, const MGWork& W, int bm, std::vector<Lev<T>>& L, double alpha, double beta,
                  double bconst, int nu2, int l, double* xacc) {
    bool accumulated = false;
    Lev<T>& F = L[l];
    Lev<T>& Cc = L[l + 1];
    LevV<T> vc = view<T>(W, Cc, alpha, beta, bconst);
    LevV<T> v = view<T>(W, F, alpha, beta, bconst);
    if (W.dim == 3 && vc.c0uni && W.stream) {
        const int ntx = (F.n[0] + PZ_X - 1) / PZ_X, nty = (F.n[1] + PZ_Y - 1) / PZ_Y;
        const int zc = stream_zc(ctx, (const void*)&k_prolong_z<T>, PZ_X * PZ_Y, Cc.n, ntx * nty);
        const dim3 g((unsigned)(ntx * nty * ((Cc.n[2] + zc - 1) / zc)));
        hipLaunchKernelGGL((k_prolong_z<T>), bz(ctx, g), dim3(PZ_X * PZ_Y), 0, ctx->stream, vc, Cc.e, F.n[0], F.n[1], F.e,
                           zc, ntx, nty, ctx->bt);
    } else {
        const int bx = F.n[0] > 64 ? 128 : 64;
        const int gx = (F.n[0] + bx - 1) / bx;
        dim3 g((unsigned)gx * F.n[1] * F.n[2]), b(bx);
#define VD_PR(DIMV, BTV)                                                                                         \
    hipLaunchKernelGGL((k_prolong<T, DIMV, BTV>), bz(ctx, g), b, 0, ctx->stream, vc, Cc.e, F.n[0], F.n[1], F.n[2], F.e, \
                       gx, ctx->bt)
        if (ctx->ncase > 1) {
            if (W.dim == 3) VD_PR(3, true);
            else VD_PR(2, true);
        } else {
            if (W.dim == 3) VD_PR(3, false);
            else VD_PR(2, false);
        }
#undef VD_PR
    }
    if (F.e2 && W.c0uni && bm != 2) {
        for (int s = 0; s < nu2;) {
            const int k = nu2 - s >= 2 ? 2 : 1;
            if (l == 0 && xacc && s + k == nu2) {
                // last post-smooth of the finest level: x += e in the same pass
                launch_fused<T>(ctx, W, bm, v, F.fzc, F.e, F.e2, F.r, k, 0, nullptr, nullptr, xacc);
                accumulated = true;
            } else {
                launch_fused<T>(ctx, W, bm, v, F.fzc, F.e, F.e2, F.r, k, 0, nullptr, nullptr);
                std::swap(F.e, F.e2);
            }
            s += k;
        }
    } else {
        for (int s = 0; s < nu2; s++) {
            launch_gsrb<T>(ctx, W, bm, v, F.e, F.r, 0, l == 0);
            launch_gsrb<T>(ctx, W, bm, v, F.e, F.r, 1, l == 0);
        }
    }
    return accumulated;
}

// returns true when the finest post-smooth also added the correction to
// xacc (x += e), so the caller skips that update
template <class T>
static bool vcycle(vd_ctx* ctx, const MGWork& W, int bm, std::vector<Lev<T>>& L, double alpha,
                   double beta, double bconst, int nu1, int nu2, double* xacc = nullptr) {
    const int l0 = vc_tail_level<T>(W, L);
    // the legs below the finest level form one timing span (VD_KT_MGCOARSE)
    // whose canonical bytes are, per level of N cells: nu1 + nu2 sweeps of
    // (3 + n_b) arrays, the residual-restriction (2 + n_b, + N/8 written)
    // and the prolongation (2, + N/8 read)
    double cb = 0.0;
    const int nb = bm == 1 ? 0 : W.dim;
    for (int l = 1; l < W.nlev; l++)
        cb += (double)sizeof(T) * (double)ncell(L[l].n) * ((nu1 + nu2) * (3 + nb) + (2 + nb) + 2 + 0.25);
    for (int l = 0; l < l0; l++) {
        vc_down<T>(ctx, W, bm, L, alpha, beta, bconst, nu1, l);
        if (l == 0) kt_begin(ctx, VD_KT_MGCOARSE);
    }
    if (l0 == 0) kt_begin(ctx, VD_KT_MGCOARSE);
    vc_tail<T>(ctx, W, bm, L, alpha, beta, bconst, nu1, nu2, l0);
    bool accumulated = false;
    for (int l = l0 - 1; l >= 0; l--) {
        if (l == 0) kt_end(ctx, VD_KT_MGCOARSE, cb);
        accumulated |= vc_up<T>(ctx, W, bm, L, alpha, beta, bconst, nu2, l, xacc);
    }
    if (l0 == 0) kt_end(ctx, VD_KT_MGCOARSE, cb);
    return accumulated;
}

template <class T>
static void coarsen_hierarchy(vd_ctx* ctx, MGWork& W, std::vector<Lev<T>>& L, bool faces, bool c0) {
    for (int l = 1; l < W.nlev; l++) {
        Lev<T>& F = L[l - 1];
        Lev<T>& Cc = L[l];
        for (int d = 0; d < W.dim; d++) {
            if (faces && !W.constb) {
                int64_t nf = nfacef(Cc.n, d);
                hipLaunchKernelGGL(k_coarsen_faces<T>, bz(ctx, grid1d(nf, 256)), dim3(256), 0, ctx->stream,
                                   F.bf[d], F.n[0], F.n[1], d, W.dim, Cc.n[0], Cc.n[1], Cc.n[2], Cc.bf[d], ctx->bt);
            }
            if (c0 && !W.per[d])
                for (int s = 0; s < 2; s++) {
                    int64_t nt = ntanf(Cc.n, d);
                    hipLaunchKernelGGL(k_coarsen_c0<T>, bz(ctx, grid1d(nt, 256)), dim3(256), 0, ctx->stream,
                                       F.c0[d][s], F.n[0], F.n[1], d, W.dim, Cc.n[0], Cc.n[1], Cc.n[2],
                                       Cc.c0[d][s], ctx->bt);
                }
        }
    }
}

static inline dim3 norm_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return dim3((unsigned)(g < NORM_BLOCKS ? (g < 1 ? 1 : g) : NORM_BLOCKS));
}

// slot group q of the norm buffer
static inline unsigned long long* slot_group(MGWork& W, int q) { return W.dnorm + (size_t)q * NSLOT * SLOT_STRIDE; }

// the max of each slot group into the host-mapped norm buffer, then the
// call's sequence number into its last slot (one workgroup).  Slots hold the bit
// patterns of non-negative doubles (NaN is stored as +inf), so their
// maximum as doubles is exact.  Each slot is cleared by the lane that read
// it, so the groups start at zero for the next norm (every accumulation
// into slot groups is read back by exactly one read_groups call).
// In a batched context wave z of the one workgroup does this for case z of
// the launch batch (its groups at hout[z * ngroups ..]), and the sequence
// number follows every case's values (after the workgroup barrier).
constexpr int HNORM_SEQ = HNORM_DOUBLES - 1;
static_assert(VD_MAX_BATCH * 5 < HNORM_SEQ, "norm read-back buffer too small");
__global__ void __launch_bounds__(64 * VD_MAX_BATCH) k_norm_out(unsigned long long* __restrict__ slots0, int ngroups,
                                                                double* __restrict__ hout, unsigned long long seq,
                                                                Batch bt, int ncase) {
    const int s = (int)threadIdx.x & 63, z = (int)threadIdx.x >> 6;
    if (z < ncase) {
        unsigned long long* __restrict__ slots = rb(slots0, bt.stride * (int64_t)bt.idx[z]);
        for (int q = 0; q < ngroups; q++) {
            double v = 0.0;
            if (s < NSLOT) {
                const size_t o = ((size_t)q * NSLOT + s) * SLOT_STRIDE;
                v = __longlong_as_double((long long)slots[o]);
                slots[o] = 0ull;
            }
            for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
            if (s == 0) hout[z * ngroups + q] = v;
        }
        __threadfence_system();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        ((volatile unsigned long long*)hout)[HNORM_SEQ] = seq;
    }
}

// max |a| over n values into a slot group (grid-stride, 2 values per
// thread and step)
__global__ void __launch_bounds__(256) k_absmax(const double* __restrict__ a, int64_t n,
                                                unsigned long long* __restrict__ slots, Batch bt) {
    a = rb(a, case_off(bt));
    slots = rb(slots, case_off(bt));
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * 2;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < n; i += stride) {
        m = fmax(m, fabs(a[i]));
        if (i + 1 < n) m = fmax(m, fabs(a[i + 1]));
    }
    block_max_atomic(m, slots);
}

// the norms of up to 5 slot groups on the host, for each case of the
// launch batch (out[z * ngroups + q]): the host waits for the sequence
// number k_norm_out writes last, rather than for a copy and a stream
// synchronisation, and checks the stream for errors while it waits
static int read_groups(vd_ctx* ctx, MGWork& W, int ngroups, double* out) {
    const unsigned long long seq = ++ctx->norm_seq;
    hipLaunchKernelGGL(k_norm_out, dim3(1), dim3(64 * ctx->bn), 0, ctx->stream, W.dnorm, ngroups, ctx->hnorm_d, seq,
                       ctx->bt, ctx->bn);
    VD_CHECK_LAUNCH();
    volatile unsigned long long* flag = (volatile unsigned long long*)(ctx->hnorm + HNORM_SEQ);
    for (unsigned long it = 1; *flag != seq; it++) {
        if ((it & 1023) == 0) {
            const hipError_t e = hipStreamQuery(ctx->stream);
            if (e == hipSuccess) {
                // the stream is done: the value is there unless the write
                // did not reach host memory, which is an error
                if (*flag != seq) {
                    set_error("multigrid norm read-back: the device write did not arrive");
                    return VD_EHIP;
                }
                break;
            }
            if (e != hipErrorNotReady) {
                set_error("multigrid norm read-back: %s", hipGetErrorString(e));
                return VD_EHIP;
            }
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int q = 0; q < ngroups * ctx->bn; q++) out[q] = ((volatile double*)ctx->hnorm)[q];
    return VD_OK;
}

// residual of the finest level into W.res (unless write_res is false) and
// its max-norm; with bnorm also the max-norm of f (the rhs with boundary
// data), with fmax (cell coefficient, bm 2) the max face coefficient per
// direction -- all from the one pass over the level.  Per case of the
// launch batch: nrm_out[z], bnorm[z], fmax[3 z + d].
// With res32 the residual goes straight into res32 in FP32 instead, where
// the kernel allows (*in32 says whether it did).
static int resid0(vd_ctx* ctx, MGWork& W, int bm, const LevV<double>& v0, double* nrm_out,
                  double* bnorm = nullptr, bool write_res = true, double* fmax = nullptr,
                  float* res32 = nullptr, bool* in32 = nullptr, bool read = true) {
    if (in32) *in32 = false;
    int64_t nc = ncell(W.n);
    const int gx = (W.n[0] + 255) / 256;
    const int nblk = gx * W.n[1] * W.n[2];
    dim3 g((unsigned)nblk), b(256);
    unsigned long long* sl = slot_group(W, 0);
    unsigned long long* bs = slot_group(W, 1);
    unsigned long long* fs = slot_group(W, 2);
    double* part = W.part;
    double* res = write_res ? W.res : nullptr;  // the norm alone when no V-cycle follows
    kt_begin(ctx, VD_KT_RESID);
    if (W.dim == 3 && W.stream && bm == 1 && !fmax && W.n[0] % 2 == 0) {
        const int ntx = (W.n[0] / 2 + RZ2_X - 1) / RZ2_X, nty = (W.n[1] + RZ2_Y - 1) / RZ2_Y;
#define VD_RESIDZ2(BNV, C0UV)                                                                                 \
    {                                                                                                         \
        const int zc = stream_zc(ctx, (const void*)&k_resid_z2<BNV, C0UV>, RZ2_X * RZ2_Y, W.n, ntx * nty);    \
        const dim3 gz((unsigned)(ntx * nty * ((W.n[2] + zc - 1) / zc)));                                      \
        if (res && res32) {                                                                                   \
            hipLaunchKernelGGL((k_resid_z2<BNV, C0UV, float>), bz(ctx, gz), dim3(RZ2_X * RZ2_Y), 0, ctx->stream, v0,   \
                               W.x, W.f, res32, zc, ntx, nty, sl, ctx->bt);                                            \
            if (in32) *in32 = true;                                                                           \
        } else {                                                                                              \
            hipLaunchKernelGGL((k_resid_z2<BNV, C0UV>), bz(ctx, gz), dim3(RZ2_X * RZ2_Y), 0, ctx->stream, v0, W.x,     \
                               W.f, res, zc, ntx, nty, sl, ctx->bt);                                                   \
        }                                                                                                     \
    }
        if (bnorm) {
            if (v0.c0uni) VD_RESIDZ2(true, true) else VD_RESIDZ2(true, false)
        } else {
            if (v0.c0uni) VD_RESIDZ2(false, true) else VD_RESIDZ2(false, false)
        }
#undef VD_RESIDZ2
    } else if (W.dim == 3 && W.stream) {
        const int ntx = (W.n[0] + RZ_X - 1) / RZ_X, nty = (W.n[1] + RZ_Y - 1) / RZ_Y;
#define VD_RESIDZ3(BMV, BNV, FMV)                                                                             \
    {                                                                                                         \
        const int zc = stream_zc(ctx,                                                                         \
                                 v0.c0uni ? (const void*)&k_resid_z<BMV, BNV, FMV, true>                      \
                                          : (const void*)&k_resid_z<BMV, BNV, FMV, false>,                    \
                                 RZ_X * RZ_Y, W.n, ntx * nty);                                                \
        const dim3 gz((unsigned)(ntx * nty * ((W.n[2] + zc - 1) / zc)));                                      \
        if (v0.c0uni)                                                                                         \
            hipLaunchKernelGGL((k_resid_z<BMV, BNV, FMV, true>), bz(ctx, gz), dim3(RZ_X * RZ_Y), 0, ctx->stream, v0,   \
                               W.x, W.f, res, zc, ntx, nty, sl, ctx->bt);                                              \
        else                                                                                                  \
            hipLaunchKernelGGL((k_resid_z<BMV, BNV, FMV, false>), bz(ctx, gz), dim3(RZ_X * RZ_Y), 0, ctx->stream, v0,  \
                               W.x, W.f, res, zc, ntx, nty, sl, ctx->bt);                                              \
    }
#define VD_RESIDZ(BMV)                                                                                        \
    {                                                                                                         \
        if (fmax && bnorm)                                                                                    \
            VD_RESIDZ3(BMV, true, true)                                                                       \
        else if (bnorm)                                                                                       \
            VD_RESIDZ3(BMV, true, false)                                                                      \
        else                                                                                                  \
            VD_RESIDZ3(BMV, false, false)                                                                     \
    }
        if (bm == 1)
            VD_RESIDZ(1)
        else if (bm == 0)
            VD_RESIDZ(0)
        else
            VD_RESIDZ(2)
#undef VD_RESIDZ
#undef VD_RESIDZ3
    } else {  // one row per block
#define VD_RESID(BMV, DIMV)                                                                                   \
    {                                                                                                         \
        if (fmax && bnorm)                                                                                    \
            hipLaunchKernelGGL((k_resid0<BMV, DIMV, true, true>), bz(ctx, g), b, 0, ctx->stream, v0, W.x, W.f, res,     \
                               part, gx, bs, fs, ctx->bt);                                                             \
        else if (bnorm)                                                                                       \
            hipLaunchKernelGGL((k_resid0<BMV, DIMV, true, false>), bz(ctx, g), b, 0, ctx->stream, v0, W.x, W.f, res,    \
                               part, gx, bs, fs, ctx->bt);                                                             \
        else                                                                                                  \
            hipLaunchKernelGGL((k_resid0<BMV, DIMV, false, false>), bz(ctx, g), b, 0, ctx->stream, v0, W.x, W.f, res,   \
                               part, gx, bs, fs, ctx->bt);                                                             \
    }
        if (bm == 1) {
            if (W.dim == 3) VD_RESID(1, 3) else VD_RESID(1, 2)
        } else if (bm == 0) {
            if (W.dim == 3) VD_RESID(0, 3) else VD_RESID(0, 2)
        } else {
            if (W.dim == 3) VD_RESID(2, 3) else VD_RESID(2, 2)
        }
#undef VD_RESID
        hipLaunchKernelGGL(k_reduce_partials, bz(ctx, dim3(1)), dim3(1024), 0, ctx->stream, part, nblk, sl, ctx->bt);
    }
    const int nb = bm == 1 ? 0 : (bm == 0 ? W.dim : 1);
    kt_end(ctx, VD_KT_RESID, 8.0 * (double)nc * ((write_res ? 4 : 3) + nb));
    VD_CHECK_LAUNCH();
    if (!read) return VD_OK;  // the caller reads the slot groups
    double m[VD_MAX_BATCH * 5];
    const int ng = fmax ? 2 + W.dim : (bnorm ? 2 : 1);
    VD_TRY(read_groups(ctx, W, ng, m));
    for (int z = 0; z < ctx->bn; z++) {
        nrm_out[z] = m[z * ng];
        if (bnorm) bnorm[z] = m[z * ng + 1];
        if (fmax)
            for (int d = 0; d < W.dim; d++) fmax[3 * z + d] = m[z * ng + 2 + d];
    }
    return VD_OK;
}

// harmonic face coefficient of direction d from the cell coefficient on the
// grown box (ScalarSolve.cpp:511-512), plus the block max of the faces into
// a norm slot group (Jacobi bound of the implicit operator)
__global__ void __launch_bounds__(256) k_harmonic_faces(const double* __restrict__ dcell, int nx, int ny,
                                                        int nz, int d, int dim, double* __restrict__ bf,
                                                        unsigned long long* __restrict__ slots, Batch bt) {
    dcell = rb(dcell, case_off(bt));
    bf = rb(bf, case_off(bt));
    slots = rb(slots, case_off(bt));
    int n0 = nx + (d == 0), n1 = ny + (d == 1), n2 = nz + (d == 2);
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < (unsigned)n0 * n1 * n2) {
        int i = (int)(t % n0);
        unsigned r = t / n0;
        int j = (int)(r % n1), k = (int)(r / n1);
        const int gx = nx + 2, gxy = gx * (ny + 2);
        const int gs = d == 0 ? 1 : (d == 1 ? gx : gxy);
        const int gc = (i + 1) + gx * (j + 1) + (dim == 3 ? gxy * (k + 1) : 0);
        const double b = harm(dcell[gc - gs], dcell[gc]);
        bf[t] = b;
        m = fabs(b);
        if (m != m) m = __builtin_huge_val();
    }
    if (slots) block_max_atomic(m, slots);
}

// Jacobi bound of the finest operator, rho = sum_d 2|beta/dx_d^2| bmax_d / |alpha|.
// With rho <= SMOOTH_ONLY_RHO every red-black sweep contracts the error by
// at least rho, the coarse-grid correction buys nothing, and the solve
// runs finest-level sweeps only (the implicit species solves at small dt).
constexpr double SMOOTH_ONLY_RHO = 1e-2;

static int jacobi_rho(vd_ctx* ctx, MGWork& W, int bm, double alpha, double beta, double bconst, double* rho) {
    *rho = INFINITY;
    if (alpha == 0.0) return VD_OK;
    double bmax[3] = {0.0, 0.0, 0.0};
    if (bm == 1) {
        for (int d = 0; d < W.dim; d++) bmax[d] = fabs(bconst);
    } else {
        for (int d = 0; d < W.dim; d++) {
            int64_t nf = nfacef(W.n, d);
            hipLaunchKernelGGL(k_maxabs, bz(ctx, norm_grid(nf)), dim3(256), 0, ctx->stream, W.bf0[d], (unsigned)nf,
                               slot_group(W, d), ctx->bt);
        }
        VD_CHECK_LAUNCH();
        VD_TRY(read_groups(ctx, W, W.dim, bmax));
    }
    double r = 0.0;
    for (int d = 0; d < W.dim; d++) r += 2.0 * fabs(beta / (W.dx[d] * W.dx[d])) * bmax[d];
    *rho = r / fabs(alpha);
    return VD_OK;
}

// Solve with the finest-level data already in W (x, f with BC shift,
// bf0 / bconst / dcell, c00).  Builds the coarse hierarchy unless the
// operator is diagonally dominant enough for sweeps alone, then iterates.
// smooth_mode: 0 V-cycles, 1 finest-level sweeps only, 2 (bm 2, the cell
// coefficient) decided from the Jacobi bound of the face maxima that the
// first residual pass also returns; when V-cycles follow, the face arrays
// are built and the solve continues on them (bm 0).
// x_zero: the solution array holds zeros on entry (a solve from a zero
// initial guess), so the initial residual is the rhs itself: its norm needs
// one read of f, and the first V-cycle takes f as its finest rhs
// the finest level's operator view for these boundary types (sets the
// uniform-c0 flags of W)
static LevV<double> fine_view(MGWork& W, const int tlo[3], const int thi[3], double alpha, double beta,
                              double bconst, const double* dcell) {
    W.c0uni = 1;
    for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
            const int ty = sd ? thi[d] : tlo[d];
            W.c0u[d][sd] = ty == VD_DIRCBC ? -1.0 : 1.0;
            if (d < W.dim && !W.per[d] && ty == VD_ROBINBC) W.c0uni = 0;
        }
    Lev<double> fine;
    for (int d = 0; d < 3; d++) {
        fine.n[d] = W.n[d];
        fine.dx[d] = W.dx[d];
        fine.bf[d] = W.bf0[d];
        for (int s = 0; s < 2; s++) fine.c0[d][s] = W.c00[d][s];
    }
    LevV<double> v0 = view<double>(W, fine, alpha, beta, bconst);
    v0.dcell = dcell;
    return v0;
}

static int mg_iterate(vd_ctx* ctx, MGWork& W, int bm, double alpha, double beta, double bconst,
                      const double* dcell, const int tlo[3], const int thi[3], int smooth_mode,
                      double reltol, double abstol, int maxiter, int nu1, int nu2, vd_solve_info* info,
                      bool x_zero = false) {
    int64_t nc = ncell(W.n);
    const LevV<double> v0 = fine_view(W, tlo, thi, alpha, beta, bconst, dcell);
    // the caller's rhs pass already left the initial norms in slot groups 0, 1
    const bool r0_done = W.r0_done;
    W.r0_done = false;
    // Every case of the launch batch (ctx->bt, ctx->bn) takes the
    // operations it would take alone: its own initial norms, its own
    // choice of finest-level sweeps or V-cycles, its own stopping
    // iteration.  Launches run on the cases that are still iterating.
    const int ncs = ctx->bn;
    int cid[VD_MAX_BATCH];
    for (int z = 0; z < ncs; z++) cid[z] = ctx->bt.idx[z];
    double r0[VD_MAX_BATCH], rn[VD_MAX_BATCH], bnorm[VD_MAX_BATCH], target[VD_MAX_BATCH];
    bool sonly[VD_MAX_BATCH];
    int iters[VD_MAX_BATCH];
    for (int z = 0; z < ncs; z++) {
        sonly[z] = smooth_mode == 1;
        iters[z] = 0;
    }
    bool initial_done = false;
    W.post_mask = 0;
    const SweepOut* post = W.post.p && bm == 1 ? &W.post : nullptr;
    // constant coefficient, sweep-only iterations on the fused pass: the
    // first sweep (and its norm, into slot group 2) is launched right after
    // the initial residual, and one read-back returns both norms; a case
    // whose initial residual already meets its target keeps its initial
    // guess, and the sweep's result in the other buffer is not used
    bool spec1 = false;
    double r1[VD_MAX_BATCH];
    if (smooth_mode == 2 && bm == 1) {
        // constant coefficient: every face is b, so the bound needs no pass,
        // and the residual is stored only when V-cycles follow
        double rho = 0.0;
        for (int d = 0; d < W.dim; d++) rho += 2.0 * fabs(beta / (W.dx[d] * W.dx[d])) * fabs(bconst);
        const bool so = alpha != 0.0 && rho / fabs(alpha) <= SMOOTH_ONLY_RHO;
        for (int z = 0; z < ncs; z++) sonly[z] = so;
        if (so && W.x2 && W.c0uni && maxiter >= 1) {
            if (!r0_done) VD_TRY(resid0(ctx, W, bm, v0, r0, bnorm, false, nullptr, nullptr, nullptr, false));
            int nblk = 0;
            launch_fused<double>(ctx, W, bm, v0, std::max(1, fused_zc(W.n)), W.x, W.x2, W.f, 1, 1, W.part, &nblk,
                                 nullptr, post);
            hipLaunchKernelGGL(k_reduce_partials, bz(ctx, dim3(1)), dim3(1024), 0, ctx->stream, W.part, nblk,
                               slot_group(W, 2), ctx->bt);
            VD_CHECK_LAUNCH();
            double m[VD_MAX_BATCH * 5];
            VD_TRY(read_groups(ctx, W, 3, m));
            for (int z = 0; z < ncs; z++) {
                r0[z] = m[3 * z];
                bnorm[z] = m[3 * z + 1];
                r1[z] = m[3 * z + 2];
            }
            spec1 = true;
        } else {
            VD_TRY(resid0(ctx, W, bm, v0, r0, bnorm, !so));
        }
        initial_done = true;
    } else if (smooth_mode == 2) {
        double bmax[3 * VD_MAX_BATCH] = {0.0};
        VD_TRY(resid0(ctx, W, bm, v0, r0, bnorm, true, bmax));
        for (int z = 0; z < ncs; z++) {
            sonly[z] = false;
            if (alpha != 0.0) {
                double rho = 0.0;
                for (int d = 0; d < W.dim; d++) rho += 2.0 * fabs(beta / (W.dx[d] * W.dx[d])) * bmax[3 * z + d];
                sonly[z] = rho / fabs(alpha) <= SMOOTH_ONLY_RHO;
            }
        }
        initial_done = true;
    }
    // the V-cycle cases first (their solution stays in W.x), then the
    // sweep-only ones (whose fused path alternates W.x and W.x2)
    int st = VD_OK;
    for (int pass = 0; pass < 2; pass++) {
        const bool smooth_only = pass == 1;
        int grp[VD_MAX_BATCH], gz[VD_MAX_BATCH], ng = 0;  // case ids and their z in the caller's batch
        for (int z = 0; z < ncs; z++)
            if (sonly[z] == smooth_only) {
                gz[ng] = z;
                grp[ng++] = cid[z];
            }
        if (ng == 0) continue;
        BatchScope scope(ctx, grp, ng);
        int bmg = bm;
        if (!smooth_only && smooth_mode == 2 && bm == 2) {
            // V-cycles on the harmonic face arrays of the cell coefficient
            for (int d = 0; d < W.dim; d++)
                hipLaunchKernelGGL(k_harmonic_faces, bz(ctx, grid1d(nfacef(W.n, d), 256)), dim3(256), 0, ctx->stream,
                                   dcell, W.n[0], W.n[1], W.n[2], d, W.dim, W.bf0[d], (unsigned long long*)nullptr,
                                   ctx->bt);
            bmg = 0;
        }
        if (!smooth_only) {
            // coarse c0 depends only on the BC types unless a side is Robin:
            // rebuilt when a side is Robin, the types changed, or a case of
            // this group has not had it built for these types
            bool robin = false, same = true;
            for (int d = 0; d < W.dim; d++) {
                if (W.per[d]) continue;
                robin |= tlo[d] == VD_ROBINBC || thi[d] == VD_ROBINBC;
                same &= W.c0_types[d][0] == tlo[d] && W.c0_types[d][1] == thi[d];
            }
            unsigned gmask = 0;
            for (int q = 0; q < ng; q++) gmask |= 1u << grp[q];
            const bool c0 = robin || !same || (W.c0_mask & gmask) != gmask;
            if (W.prec == 0) {
                coarsen_hierarchy<double>(ctx, W, W.L64, true, c0);
            } else {
                Lev<float>& L0 = W.L32[0];
                for (int d = 0; d < W.dim; d++) {
                    if (!W.constb) {
                        int64_t nf = nfacef(W.n, d);
                        hipLaunchKernelGGL((k_convert<double, float>), bz(ctx, grid1d(nf, 256)), dim3(256), 0,
                                           ctx->stream, W.bf0[d], L0.bf[d], nf, ctx->bt);
                    }
                    if (c0 && !W.per[d])
                        for (int s = 0; s < 2; s++) {
                            int64_t nt = ntanf(W.n, d);
                            hipLaunchKernelGGL((k_convert<double, float>), bz(ctx, grid1d(nt, 256)), dim3(256), 0,
                                               ctx->stream, W.c00[d][s], L0.c0[d][s], nt, ctx->bt);
                        }
                }
                coarsen_hierarchy<float>(ctx, W, W.L32, true, c0);
                // the FP64 bottom solve's operator, coarsened in FP64
                coarsen_hierarchy<double>(ctx, W, W.C64, true, c0);
            }
            if (c0) {
                if (!same) W.c0_mask = 0;
                if (robin) W.c0_mask &= ~gmask;
                else W.c0_mask |= gmask;
            }
            for (int d = 0; d < 3; d++) {
                W.c0_types[d][0] = tlo[d];
                W.c0_types[d][1] = thi[d];
            }
            VD_CHECK_LAUNCH();
        }
        // the first V-cycle's finest rhs: the stored residual, or f when x is 0;
        // for FP32 V-cycles the residual pass may store it in FP32 directly
        bool rhs_is_f = false;
        float* r32p = W.prec == 1 && !W.L32.empty() ? W.L32[0].r : nullptr;
        double gr0[VD_MAX_BATCH], gb[VD_MAX_BATCH];
        bool r_in32 = false;
        if (!initial_done) {
            if (x_zero && !smooth_only) {
                // r = f - L 0 = f bit for bit (every stencil term is 0)
                hipLaunchKernelGGL(k_absmax, bz(ctx, dim3(NORM_BLOCKS)), dim3(256), 0, ctx->stream, W.f, nc,
                                   slot_group(W, 0), ctx->bt);
                VD_CHECK_LAUNCH();
                VD_TRY(read_groups(ctx, W, 1, gb));
                for (int q = 0; q < ng; q++) gr0[q] = gb[q];
                rhs_is_f = true;
            } else {
                VD_TRY(resid0(ctx, W, bmg, v0, gr0, gb, !smooth_only, nullptr, r32p, &r_in32));
            }
            for (int q = 0; q < ng; q++) {
                r0[gz[q]] = gr0[q];
                bnorm[gz[q]] = gb[q];
            }
        }
        // the cases still iterating (ids, and their z in the caller's batch)
        int act[VD_MAX_BATCH], az[VD_MAX_BATCH], na = 0;
        for (int q = 0; q < ng; q++) {
            const int z = gz[q];
            target[z] = std::max(abstol, reltol * std::max(bnorm[z], r0[z]));
            rn[z] = r0[z];
            if (!(r0[z] <= target[z])) {
                act[na] = grp[q];
                az[na++] = z;
            }
        }
        // fused sweep-only path: where each case's solution ended up
        double* xfinal[VD_MAX_BATCH];
        for (int q = 0; q < ng; q++) xfinal[gz[q]] = W.x;
        const bool fused_so =
            smooth_only && W.x2 && W.c0uni && (bmg != 2 || fused_cellcoef(W));
        for (int it = 1; it <= maxiter && na > 0; it++) {
            BatchScope iscope(ctx, act, na);
            double m[VD_MAX_BATCH];
            if (fused_so && it == 1 && spec1) {
                // the sweep launched with the initial residual
                std::swap(W.x, W.x2);
                for (int q = 0; q < na; q++) m[q] = r1[az[q]];
            } else if (fused_so) {
                // one red-black sweep and the residual norm of its result in one pass
                int nblk = 0;
                launch_fused<double>(ctx, W, bmg, v0, bmg == 2 ? fused_zc_cc(W.n) : std::max(1, fused_zc(W.n)), W.x,
                                     W.x2, W.f, 1, 1, W.part, &nblk, nullptr, bmg == 1 ? post : nullptr);
                std::swap(W.x, W.x2);
                hipLaunchKernelGGL(k_reduce_partials, bz(ctx, dim3(1)), dim3(1024), 0, ctx->stream, W.part, nblk,
                                   slot_group(W, 0), ctx->bt);
                VD_CHECK_LAUNCH();
                VD_TRY(read_groups(ctx, W, 1, m));
            } else {
                if (smooth_only) {
                    launch_gsrb<double>(ctx, W, bmg, v0, W.x, W.f, 0, true);
                    launch_gsrb<double>(ctx, W, bmg, v0, W.x, W.f, 1, true);
                } else if (W.prec == 0) {
                    if (rhs_is_f) W.L64[0].r = W.f;  // read only by the V-cycle
                    const bool acc = vcycle<double>(ctx, W, bmg, W.L64, alpha, beta, bconst, nu1, nu2, W.x);
                    W.L64[0].r = W.res;
                    if (!acc)
                        hipLaunchKernelGGL(k_axpy_to_double<double>, bz(ctx, grid1d(nc, 256)), dim3(256), 0,
                                           ctx->stream, W.x, W.L64[0].e, nc, ctx->bt);
                } else {
                    if (!r_in32)
                        hipLaunchKernelGGL((k_convert<double, float>), bz(ctx, grid1d(nc, 256)), dim3(256), 0,
                                           ctx->stream, rhs_is_f ? W.f : W.res, W.L32[0].r, nc, ctx->bt);
                    if (!vcycle<float>(ctx, W, bmg, W.L32, alpha, beta, bconst, nu1, nu2, W.x))
                        hipLaunchKernelGGL(k_axpy_to_double<float>, bz(ctx, grid1d(nc, 256)), dim3(256), 0,
                                           ctx->stream, W.x, W.L32[0].e, nc, ctx->bt);
                }
                VD_CHECK_LAUNCH();
                rhs_is_f = false;
                VD_TRY(resid0(ctx, W, bmg, v0, m, nullptr, !smooth_only, nullptr, r32p, &r_in32));
            }
            int nna = 0;
            for (int q = 0; q < na; q++) {
                const int z = az[q];
                rn[z] = m[q];
                xfinal[z] = W.x;
                if (rn[z] <= target[z]) {
                    iters[z] = it;
                } else {
                    act[nna] = act[q];
                    az[nna++] = z;
                }
            }
            na = nna;
        }
        for (int q = 0; q < na; q++) {
            const int z = az[q];
            iters[z] = maxiter;
            st = VD_ENOCONV;
            set_error("multigrid: no convergence in %d iterations (|r|=%g target=%g)", maxiter, rn[z], target[z]);
        }
        if (fused_so) {
            // the cases whose last iterate the sweeps also wrote to the state
            if (post && bmg == 1)
                for (int q = 0; q < ng; q++)
                    if (iters[gz[q]] >= 1) W.post_mask |= 1u << grp[q];
            // the cases that stopped with their solution in the other
            // buffer copy it into W.x, where every case's solution is read
            int cp[VD_MAX_BATCH], ncp = 0;
            for (int q = 0; q < ng; q++)
                if (xfinal[gz[q]] != W.x) cp[ncp++] = grp[q];
            if (ncp > 0) {
                BatchScope cscope(ctx, cp, ncp);
                hipLaunchKernelGGL((k_convert<double, double>), bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream,
                                   W.x2, W.x, nc, ctx->bt);
                VD_CHECK_LAUNCH();
            }
        }
    }
    // per case, and for the call: the first case's figures with the most
    // iterations of any case
    int imax = 0;
    for (int z = 0; z < ncs; z++) {
        vd_solve_info& ci = ctx->case_info[cid[z]];
        ci.iterations = iters[z];
        ci.resnorm0 = r0[z];
        ci.resnorm = rn[z];
        ci.target = target[z];
        imax = std::max(imax, iters[z]);
    }
    if (info) {
        info->iterations = imax;
        info->resnorm0 = r0[0];
        info->resnorm = rn[0];
        info->target = target[0];
    }
    return st;
}

// ------------------------------------------------ array4 <-> flat helpers
struct FlatArgs {
    Arr<double> a;
    int comp;
    int3 lo;
    int nx, ny, nz;
};

__global__ void k_a4_to_flat(FlatArgs f, double* __restrict__ out) {
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)f.nx * f.ny * f.nz) return;
    int i, j, k;
    cell_of((unsigned)t, f.nx, f.ny, i, j, k);
    out[t] = f.a(i + f.lo.x, j + f.lo.y, k + f.lo.z, f.comp);
}

__global__ void k_flat_to_a4(FlatArgs f, const double* __restrict__ in) {
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)f.nx * f.ny * f.nz) return;
    int i, j, k;
    cell_of((unsigned)t, f.nx, f.ny, i, j, k);
    f.a(i + f.lo.x, j + f.lo.y, k + f.lo.z, f.comp) = in[t];
}

static FlatArgs flat_args(const vd_array4* a, int comp, const int lo[3], const int n[3]) {
    FlatArgs f;
    f.a = arr(a);
    f.comp = comp;
    f.lo = make_int3(lo[0], lo[1], lo[2]);
    f.nx = n[0];
    f.ny = n[1];
    f.nz = n[2];
    return f;
}

// ghost-cell values of a 1-ghost array on one boundary side -> bcv (tangential)
__global__ void k_side_from_ghost(Arr<double> a, int d, int side, int3 lo, int nx, int ny, int nz, int dim,
                                  double* __restrict__ out) {
    int n[3] = {nx, ny, nz};
    int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    int nt1 = n[t1], nt2 = dim == 3 ? n[t2] : 1;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt1 * nt2) return;
    int c[3] = {0, 0, 0};
    c[t1] = (int)(t % nt1);
    if (dim == 3) c[t2] = (int)(t / nt1);
    int g[3] = {c[0] + lo.x, c[1] + lo.y, c[2] + lo.z};
    g[d] = side ? (d == 0 ? lo.x : (d == 1 ? lo.y : lo.z)) + n[d] : (d == 0 ? lo.x : (d == 1 ? lo.y : lo.z)) - 1;
    int ti = d == 0 ? c[1] + ny * c[2] : (d == 1 ? c[0] + nx * c[2] : c[0] + nx * c[1]);
    out[ti] = a(g[0], g[1], g[2]);
}

// face gradient (MLMG::getGradSolution) with the solver's BC conventions,
// scaled by `sign` (-1 turns grad phi into E)
struct GradArgs {
    const double* x;
    Arr<double> out;
    int d, dim, nx, ny, nz;
    int3 lo;
    int per, tlo, thi;
    double dx, sign;
    const double* bcv[2];
    const double* ra[2];
    const double* rb[2];
    const double* rf[2];
    __device__ void rebase(int64_t o) {
        x = vd::rb(x, o);
        vd::rebase(out, o);
#pragma unroll
        for (int s = 0; s < 2; s++) {
            bcv[s] = vd::rb(bcv[s], o);
            ra[s] = vd::rb(ra[s], o);
            rb[s] = vd::rb(rb[s], o);
            rf[s] = vd::rb(rf[s], o);
        }
    }
};

// the gradient on face (i, j, k) of direction a.d (face f[d] lies between
// cells f[d]-1 and f[d]), times a.sign
__device__ __forceinline__ double grad_face(const GradArgs& a, int i, int j, int k) {
    int f[3] = {i, j, k};
    int n[3] = {a.nx, a.ny, a.nz};
    const int st = a.d == 0 ? 1 : (a.d == 1 ? a.nx : a.nx * a.ny);
    int fd = f[a.d];
    double gv;
    if (fd > 0 && fd < n[a.d]) {
        int c = i + a.nx * (j + a.ny * k);
        gv = (a.x[c] - a.x[c - st]) / a.dx;
    } else if (a.per) {
        // faces 0 and n both separate cell n-1 (left) from cell 0 (right)
        int cr[3] = {i, j, k};
        cr[a.d] = 0;
        int c0i = cr[0] + a.nx * (cr[1] + a.ny * cr[2]);
        gv = (a.x[c0i] - a.x[c0i + (n[a.d] - 1) * st]) / a.dx;
    } else {
        int side = fd == 0 ? 0 : 1;
        int type = side ? a.thi : a.tlo;
        int ci[3] = {i, j, k};
        if (side) ci[a.d] -= 1;
        int c = ci[0] + a.nx * (ci[1] + a.ny * ci[2]);
        int ti = a.d == 0 ? ci[1] + a.ny * ci[2] : (a.d == 1 ? ci[0] + a.nx * ci[2] : ci[0] + a.nx * ci[1]);
        double val = a.bcv[side] ? a.bcv[side][ti] : 0.0;
        double xin = a.x[c];
        if (type == VD_DIRCBC) {
            gv = side ? 2.0 * (val - xin) / a.dx : 2.0 * (xin - val) / a.dx;
        } else if (type == VD_HNEUBC) {
            gv = 0.0;
        } else if (type == VD_IHNEUBC) {
            gv = val;
        } else {
            double ra = 0.0, rb = 1.0, rf = 0.0, c0, cinh;
            if (a.ra[side]) {
                ra = a.ra[side][ti];
                rb = a.rb[side][ti];
                rf = a.rf[side][ti];
            }
            bc_coef(VD_ROBINBC, side, a.dx, val, ra, rb, rf, c0, cinh);
            double xg = c0 * xin + cinh;
            gv = side ? (xg - xin) / a.dx : (xin - xg) / a.dx;
        }
    }
    if (a.sign != 1.0) gv *= a.sign;
    return gv;
}

__global__ void k_grad_faces(GradArgs a, Batch bt) {
    a.rebase(case_off(bt));
    int n0 = a.nx + (a.d == 0), n1 = a.ny + (a.d == 1), n2 = a.nz + (a.d == 2);
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n0 * n1 * n2) return;
    int i, j, k;
    cell_of((unsigned)t, n0, n1, i, j, k);
    a.out(i + a.lo.x, j + a.lo.y, k + a.lo.z) = grad_face(a, i, j, k);
}

static GradArgs grad_args(const MGWork& W, const int tlo[3], const int thi[3], vd_array4* out, int d,
                          const vd_box* box, double sign) {
    GradArgs a;
    a.x = W.x;
    a.out = arr(out);
    a.d = d;
    a.dim = W.dim;
    a.nx = W.n[0];
    a.ny = W.n[1];
    a.nz = W.n[2];
    a.lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
    a.per = W.per[d];
    a.tlo = tlo[d];
    a.thi = thi[d];
    a.dx = W.dx[d];
    a.sign = sign;
    for (int s = 0; s < 2; s++) {
        a.bcv[s] = W.bcv[d][s];
        bool rob = (s ? thi[d] : tlo[d]) == VD_ROBINBC;
        a.ra[s] = rob ? W.rob[d][s][0] : nullptr;
        a.rb[s] = rob ? W.rob[d][s][1] : nullptr;
        a.rf[s] = rob ? W.rob[d][s][2] : nullptr;
    }
    return a;
}

static int grad_faces(vd_ctx* ctx, MGWork& W, const int tlo[3], const int thi[3], vd_array4 out[3],
                      const vd_box* box, double sign) {
    for (int d = 0; d < W.dim; d++) {
        GradArgs a = grad_args(W, tlo, thi, &out[d], d, box, sign);
        int64_t nf = nfacef(W.n, d);
        hipLaunchKernelGGL(k_grad_faces, bz(ctx, grid1d(nf, 256)), dim3(256), 0, ctx->stream, a, ctx->bt);
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

// c0 and rhs shift for every non-periodic side (sequential sides: no races
// on corner cells).  bcv / rob must be filled.
// to_sh: each face's cinh goes to W.sh (for a rhs pass that adds the
// terms, constant b only) instead of its term into f
static int side_setup(vd_ctx* ctx, MGWork& W, const int tlo[3], const int thi[3], double beta,
                      double bconst, const double* dcell = nullptr, bool use_const = false, bool to_sh = false) {
    for (int d = 0; d < W.dim; d++) {
        if (W.per[d]) continue;
        SideArgs sa[2];
        for (int s = 0; s < 2; s++) {
            SideArgs& a = sa[s];
            a.d = d;
            a.side = s;
            a.type = s ? thi[d] : tlo[d];
            a.dim = W.dim;
            a.nx = W.n[0];
            a.ny = W.n[1];
            a.nz = W.n[2];
            a.dx = W.dx[d];
            a.s = beta / (W.dx[d] * W.dx[d]);
            a.bcv = W.bcv[d][s];
            bool rob = a.type == VD_ROBINBC;
            a.ra = rob ? W.rob[d][s][0] : nullptr;
            a.rb = rob ? W.rob[d][s][1] : nullptr;
            a.rf = rob ? W.rob[d][s][2] : nullptr;
            a.bf = (W.constb || dcell || use_const) ? nullptr : W.bf0[d];
            a.dcell = dcell;
            a.bconst = bconst;
            a.c0 = W.c00[d][s];
            a.f = to_sh ? nullptr : W.f;
            a.sh = to_sh ? W.sh[d][s] : nullptr;
        }
        // a direction one cell thick has both sides on the same cells:
        // keep them in order there
        dim3 gs = grid1d(ntanf(W.n, d), 256);
        if (W.n[d] > 1) {
            gs.y = 2;
            hipLaunchKernelGGL(k_side_setup, bz(ctx, gs), dim3(256), 0, ctx->stream, sa[0], sa[1], ctx->bt);
        } else {
            hipLaunchKernelGGL(k_side_setup, bz(ctx, gs), dim3(256), 0, ctx->stream, sa[0], sa[0], ctx->bt);
            hipLaunchKernelGGL(k_side_setup, bz(ctx, gs), dim3(256), 0, ctx->stream, sa[1], sa[1], ctx->bt);
        }
    }
    VD_CHECK_LAUNCH();
    return VD_OK;
}

static int check_bc(const vd_geom* g, const int lo[3], const int hi[3]) {
    for (int d = 0; d < g->dim; d++) {
        if (g->is_periodic[d]) continue;
        for (int v : {lo[d], hi[d]})
            if (v < VD_DIRCBC || v > VD_ROBINBC) {
                set_error("bc code %d invalid for non-periodic direction %d", v, d);
                return VD_EINVAL;
            }
    }
    return VD_OK;
}

// ------------------------------------------------ potential solve kernels
struct PotRhsArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb;
    int3 lo;
    int nx, ny, nz, spacechrg;
    __device__ void rebase(int64_t o) {
        M = vd::rb(M, o);
        vd::rebase(sb, o);
    }
};

// rhs = -e/eps0 * sum z n (PotentialSolve.cpp:203-223)
__global__ void k_pot_rhs(PotRhsArgs a, double* __restrict__ f, Batch bt) {
    a.rebase(case_off(bt));
    f = rb(f, case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    double r = 0.0;
    if (a.spacechrg) {
        int i, j, k;
        cell_of((unsigned)t, a.nx, a.ny, i, j, k);
        int o = a.sb.off(i + a.lo.x, j + a.lo.y, k + a.lo.z);
        const DevMech& M = *a.M;
        for (int sp = 0; sp < M.nspec; sp++)
            if (abs(M.charge[sp]) > 0) r += M.charge[sp] * a.sb.p[(int64_t)sp * a.sb.ns + o];
        r *= (-VD_ECHARGE / VD_EPS0);
    }
    f[t] = r;
}

struct PotPostArgs {
    const double* x;
    GradArgs g[3];  // the face gradients times -1 (E) and their arrays
    Arr<double> phi;
    int ns, dim;
    int3 lo;
    int nx, ny, nz;
    __device__ void rebase(int64_t o) {
        x = vd::rb(x, o);
#pragma unroll
        for (int d = 0; d < 3; d++) g[d].rebase(o);
        vd::rebase(phi, o);
    }
};

// E on the faces (MLMG::getGradSolution times -1), the POT copy, E cell
// averages (average_face_to_cellcenter) and E/N in Td
// (PotentialSolve.cpp:340-379) in one pass over the cells: a cell writes
// its low face in each direction, the last cell of a line also the high
// face, and the cell average takes both faces as computed, so the face
// arrays are written once and not read back
template <int D>
__device__ __forceinline__ double pot_face_pair(const GradArgs& g, const int c[3], const int n[3], int gi, int gj,
                                                int gk) {
    const double flo = grad_face(g, c[0], c[1], c[2]);
    const double fhi = grad_face(g, c[0] + (D == 0), c[1] + (D == 1), c[2] + (D == 2));
    g.out(gi, gj, gk) = flo;
    if (c[D] == n[D] - 1) g.out(gi + (D == 0), gj + (D == 1), gk + (D == 2)) = fhi;
    return 0.5 * (flo + fhi);
}

// both faces of direction D of a cell inside the level (no wall, no
// periodic wrap): grad_face's interior branch for each, from the three
// values of the line
template <int D>
__device__ __forceinline__ double pot_face_pair_in(const GradArgs& g, double xl, double xc, double xh, int gi, int gj,
                                                   int gk) {
    double flo = (xc - xl) / g.dx;
    double fhi = (xh - xc) / g.dx;
    if (g.sign != 1.0) {
        flo *= g.sign;
        fhi *= g.sign;
    }
    g.out(gi, gj, gk) = flo;
    return 0.5 * (flo + fhi);
}

template <bool BT>
__global__ void k_pot_post(PotPostArgs a, Batch bt) {
    if constexpr (BT) a.rebase(case_off(bt));
    // the box holds fewer than 2^31 cells (32-bit cell offsets)
    const int t = (int)((int64_t)lblock() * blockDim.x + threadIdx.x);
    if (t >= a.nx * a.ny * a.nz) return;
    const int q = t / a.nx;
    const int c[3] = {t - q * a.nx, q % a.ny, q / a.ny};
    const int n[3] = {a.nx, a.ny, a.nz};
    const int i = c[0] + a.lo.x, j = c[1] + a.lo.y, k = c[2] + a.lo.z;
    const int ns = a.ns;
    // the number density first (the species are not written here), its
    // loads four at a time, summed in species order
    double nd = 0.0;
    {
        const double* P = &a.phi(i, j, k, 0);
        const int64_t N = a.phi.ns;
        int sp = 0;
        for (; sp + 4 <= ns; sp += 4) {
            const double v0 = P[sp * N], v1 = P[(sp + 1) * N], v2 = P[(sp + 2) * N], v3 = P[(sp + 3) * N];
            nd += v0;
            nd += v1;
            nd += v2;
            nd += v3;
        }
        for (; sp < ns; sp++) nd += P[sp * N];
    }
    double e[3] = {0.0, 0.0, 0.0};
    const bool in = c[0] > 0 && c[0] < n[0] - 1 && (a.dim < 2 || (c[1] > 0 && c[1] < n[1] - 1)) &&
                    (a.dim < 3 || (c[2] > 0 && c[2] < n[2] - 1));
    double xc;
    if (in) {
        // a cell with every face inside the level: the seven values of its
        // stencil loaded together (a.x is the solution the gradients read)
        const double* X = a.x;
        const int sy = a.nx, sz = a.nx * a.ny;
        xc = X[t];
        const double xl = X[t - 1], xh = X[t + 1];
        const double yl = a.dim > 1 ? X[t - sy] : 0.0, yh = a.dim > 1 ? X[t + sy] : 0.0;
        const double zl = a.dim > 2 ? X[t - sz] : 0.0, zh = a.dim > 2 ? X[t + sz] : 0.0;
        e[0] = pot_face_pair_in<0>(a.g[0], xl, xc, xh, i, j, k);
        if (a.dim > 1) e[1] = pot_face_pair_in<1>(a.g[1], yl, xc, yh, i, j, k);
        if (a.dim > 2) e[2] = pot_face_pair_in<2>(a.g[2], zl, xc, zh, i, j, k);
    } else {
        xc = a.x[t];
        e[0] = pot_face_pair<0>(a.g[0], c, n, i, j, k);
        if (a.dim > 1) e[1] = pot_face_pair<1>(a.g[1], c, n, i, j, k);
        if (a.dim > 2) e[2] = pot_face_pair<2>(a.g[2], c, n, i, j, k);
    }
    a.phi(i, j, k, VD_POT(ns)) = xc;
    for (int d = 0; d < a.dim; d++) a.phi(i, j, k, VD_EFX(ns) + d) = e[d];
    double es = 0.0;
    for (int d = 0; d < a.dim; d++) es += e[d] * e[d];
    a.phi(i, j, k, VD_REF(ns)) = (sqrt(es) / nd) / 1.0e-21;
}

// applied voltage (Source/HelperFuncs.H:75-92)
static double applied_potential(double t, int domend, const vd_params* p, const double* amp_case = nullptr) {
    double amp = domend == -1 ? p->voltage_amp_1 : p->voltage_amp_2;
    if (amp_case) amp = domend == -1 ? amp_case[0] : amp_case[1];
    if (p->voltage_profile == 1) return sin(2.0 * VD_PI * p->voltage_freq * t) * amp;
    if (p->voltage_profile == 2) {
        double vc = p->voltage_center, vd = p->voltage_dur;
        if (t <= vc) return (vc - t < vd / 2.0) ? (1.0 - (vc - t) / (vd / 2.0)) * amp : 0.0;
        return (t - vc < vd / 2.0) ? (1.0 - (t - vc) / (vd / 2.0)) * amp : 0.0;
    }
    return amp;
}

static void dims_of(const vd_box* box, const vd_geom* g, int n[3], int lo[3]) {
    for (int d = 0; d < 3; d++) {
        n[d] = d < g->dim ? box->hi[d] - box->lo[d] + 1 : 1;
        lo[d] = box->lo[d];
    }
}

constexpr int FILL_SIDES = 25;  // 6 sides x (value, Robin a, b, f) + initial guess

int solve_potential(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4* phi,
                    vd_array4 ef[3], const vd_box* box, const vd_geom* g, const vd_params* prm,
                    double time, vd_solve_info* info) {
    VD_TRY(check_bc(g, prm->pot_bc_lo, prm->pot_bc_hi));
    MGWork* W;
    VD_TRY(get_work(ctx, g, box, prm->mg_precision, 1, prm->linsolve_max_coarsening_level, &W));
    int n[3], lo[3];
    dims_of(box, g, n, lo);
    int64_t nc = ncell(n);
    PotRhsArgs ra;
    ra.M = m->dev;
    ra.sb = arr(sb);
    ra.lo = make_int3(lo[0], lo[1], lo[2]);
    ra.nx = n[0];
    ra.ny = n[1];
    ra.nz = n[2];
    ra.spacechrg = prm->do_spacechrg;
    hipLaunchKernelGGL(k_pot_rhs, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, ra, W->f, ctx->bt);
    // potential_bc (PlasmaChem.H:18-72): Dirichlet / Robin(1,0,V) value V,
    // every side's arrays in one launch -- one per case when the cases of a
    // batched context have voltages of their own
    bool per_case = false;
    for (int z = 0; z < ctx->bn; z++) per_case |= ctx->case_volt_set[ctx->bt.idx[z]] != 0;
    for (int z = 0; z < (per_case ? ctx->bn : 1); z++) {
        const int c = ctx->bt.idx[z];
        const double* amp = per_case && ctx->case_volt_set[c] ? ctx->case_volt[c] : nullptr;
        double* fp[FILL_SIDES];
        int64_t fn[FILL_SIDES];
        double fv[FILL_SIDES];
        int nf = 0;
        for (int d = 0; d < g->dim; d++) {
            if (W->per[d]) continue;
            for (int s = 0; s < 2; s++) {
                const double V = applied_potential(time, s ? 1 : -1, prm, amp);
                const int64_t nt = ntanf(n, d);
                double* arrs[4] = {W->bcv[d][s], W->rob[d][s][0], W->rob[d][s][1], W->rob[d][s][2]};
                const double vals[4] = {V, 1.0, 0.0, V};
                for (int q = 0; q < 4; q++) {
                    fp[nf] = arrs[q];
                    fn[nf] = nt;
                    fv[nf] = vals[q];
                    nf++;
                }
            }
        }
        // zero initial guess (PotentialSolve.cpp:162-168)
        fp[nf] = W->x;
        fn[nf] = nc;
        fv[nf] = 0.0;
        nf++;
        if (per_case) {
            BatchScope one(ctx, &c, 1);
            VD_TRY(fill_values(ctx, fp, fn, fv, nf));
        } else {
            VD_TRY(fill_values(ctx, fp, fn, fv, nf));
        }
    }
    const double bconst = -1.0, alpha = 0.0, beta = 1.0;
    VD_TRY(side_setup(ctx, *W, prm->pot_bc_lo, prm->pot_bc_hi, beta, bconst));
    // W->x was set to zero above (the batched fill)
    int st = mg_iterate(ctx, *W, 1, alpha, beta, bconst, nullptr, prm->pot_bc_lo, prm->pot_bc_hi, false,
                        prm->linsolve_reltol, prm->linsolve_abstol, prm->linsolve_maxiter,
                        prm->mg_pre_smooth, prm->mg_post_smooth, info, true);
    if (st != VD_OK && st != VD_ENOCONV) return st;
    PotPostArgs pa;
    pa.x = W->x;
    for (int d = 0; d < 3; d++)
        pa.g[d] = grad_args(*W, prm->pot_bc_lo, prm->pot_bc_hi, &ef[d < g->dim ? d : 0], d < g->dim ? d : 0, box,
                            -1.0);
    pa.phi = arr(phi);
    pa.ns = m->hdev.nspec;
    pa.dim = g->dim;
    pa.lo = make_int3(lo[0], lo[1], lo[2]);
    pa.nx = n[0];
    pa.ny = n[1];
    pa.nz = n[2];
    // reads the solution and the species (E/N), writes the potential, E
    // cell and face components and E/N
    kt_begin(ctx, VD_KT_POTPOST);
    if (ctx->ncase > 1)
        hipLaunchKernelGGL(k_pot_post<true>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, pa, ctx->bt);
    else
        hipLaunchKernelGGL(k_pot_post<false>, grid1d(nc, 256), dim3(256), 0, ctx->stream, pa, ctx->bt);
    kt_end(ctx, VD_KT_POTPOST, 8.0 * (double)nc * (1 + m->hdev.nspec + 2 + 2 * g->dim));
    VD_CHECK_LAUNCH();
    return st;
}

// ------------------------------------------------ implicit species solve
struct CellDArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb;
    int spec, ns, dim;
    int state;          // the coefficient is a fit of the state (else it depends on Te only)
    int3 lo;
    int gnx, gny, gnz;  // grown box extent
    double Tg;
    __device__ void rebase(int64_t o) {
        M = vd::rb(M, o);
        vd::rebase(sb, o);
    }
};

// cell specDiff on the box grown by one (ScalarSolve.cpp:483-497)
template <bool STATE>
__global__ void __launch_bounds__(256) k_cell_diff(CellDArgs a, double* __restrict__ dcell, Batch bt) {
    a.rebase(case_off(bt));
    dcell = rb(dcell, case_off(bt));
    const int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.gnx * a.gny * a.gnz) return;
    int i, j, k;
    cell_of((unsigned)t, a.gnx, a.gny, i, j, k);
    const int gz = a.dim == 3 ? 1 : 0;
    const int o = a.sb.off(a.lo.x + i - 1, a.lo.y + j - 1, a.lo.z + k - gz);
    const double* S = a.sb.p;
    const int64_t N = a.sb.ns;
    const double te = S[(int64_t)VD_ETEMP(a.ns) * N + o];
    double D;
    if (STATE) {
        // ndens and |E| of the cell (ScalarSolve.cpp:484-490)
        double nd = 0.0, es = 0.0;
        for (int sp = 0; sp < a.ns; sp++) nd += S[(int64_t)sp * N + o];
        for (int d = 0; d < a.dim; d++) {
            const double e = S[(int64_t)(VD_EFX(a.ns) + d) * N + o];
            es += e * e;
        }
        D = spec_diff_st(*a.M, a.spec, te, a.Tg, nd, sqrt(es));
    } else {
        D = spec_diff(*a.M, a.spec, te, a.Tg);
    }
    dcell[i + a.gnx * (j + a.gny * k)] = D;
}

// harmonic face average of the cell coefficient (ScalarSolve.cpp:511-512)
struct SpecBcArgs {
    const DevMech* __restrict__ M;
    Arr<double> sb;
    int spec, d, side, ns, dim;
    int3 lo;
    int nx, ny, nz;
    double Tg;
    double* out;
    __device__ void rebase(int64_t o) {
        M = vd::rb(M, o);
        vd::rebase(sb, o);
        out = vd::rb(out, o);
    }
};

// the non-periodic sides of one solve, one launch (blockIdx.y = side)
struct SpecBcSides {
    int nside;
    int d[6], side[6];
    double* out[6];
};

// plasmachem_transport::species_bc (Source/PlasmaChem.H:74-184)
// STATE: the species' coefficients are fits of the state (then ndens and
// |E| of the boundary cell are formed)
template <bool STATE>
__global__ void k_species_bc(SpecBcArgs a, SpecBcSides sides, Batch bt) {
    a.d = sides.d[blockIdx.y];
    a.side = sides.side[blockIdx.y];
    a.out = sides.out[blockIdx.y];
    a.rebase(case_off(bt));
    int n[3] = {a.nx, a.ny, a.nz};
    int t1 = a.d == 0 ? 1 : 0, t2 = a.d == 2 ? 1 : 2;
    int nt1 = n[t1], nt2 = a.dim == 3 ? n[t2] : 1;
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt1 * nt2) return;
    int c[3] = {0, 0, 0};
    c[t1] = (int)(t % nt1);
    if (a.dim == 3) c[t2] = (int)(t / nt1);
    int ti = a.d == 0 ? c[1] + a.ny * c[2] : (a.d == 1 ? c[0] + a.nx * c[2] : c[0] + a.nx * c[1]);
    int lo[3] = {a.lo.x, a.lo.y, a.lo.z};
    int face[3] = {c[0] + lo[0], c[1] + lo[1], c[2] + lo[2]};
    face[a.d] = a.side ? lo[a.d] + n[a.d] : lo[a.d];
    int cell[3] = {face[0], face[1], face[2]};
    if (a.side) cell[a.d] -= 1;
    const DevMech& M = *a.M;
    const int ns = a.ns, E = M.e_idx;
    const int oc = a.sb.off(cell[0], cell[1], cell[2]);
    const double* S = a.sb.p;
    const int64_t N = a.sb.ns;
    double sgn = a.side ? 1.0 : -1.0;
    double Te = S[VD_ETEMP(ns) * N + oc];
    // ndens and |E| of the boundary cell (PlasmaChem.H:107-108,114-116)
    double nd = 0.0, emag = 0.0;
    if (STATE) {
        double es = 0.0;
        for (int sp = 0; sp < ns; sp++) nd += S[(int64_t)sp * N + oc];
        for (int d = 0; d < 3; d++) {
            const double e = S[(int64_t)(VD_EFX(ns) + d) * N + oc];
            es += e * e;
        }
        emag = sqrt(es);
    }
    double v;
    if (a.spec == E) {
        double ne = S[(int64_t)E * N + oc];
        double D = STATE ? spec_diff_st(M, a.spec, Te, a.Tg, nd, emag) : spec_diff(M, a.spec, Te, a.Tg);
        double cbar = sqrt(8.0 * VD_K_B * Te / VD_PI / VD_ME);
        double flx = ne * cbar / 4.0 * sgn;
        v = -flx / D;
    } else if (a.spec == ns) {
        double ne = S[(int64_t)E * N + oc];
        double D = STATE ? spec_diff_st(M, a.spec, Te, a.Tg, nd, emag) : spec_diff(M, a.spec, Te, a.Tg);
        double cbar = sqrt(8.0 * VD_K_B * Te / VD_PI / VD_ME);
        double flx = (2.0 * VD_K_B * Te) * ne * cbar / 4.0 * sgn;
        v = -flx / D;
    } else if (M.charge[a.spec] != 0) {
        double mu = STATE ? spec_mob_st(M, a.spec, Te, a.Tg, nd, emag) : spec_mob(M, a.spec);
        double D = STATE ? spec_diff_st(M, a.spec, Te, a.Tg, nd, emag) : spec_diff(M, a.spec, Te, a.Tg);
        double efield = S[(int64_t)(VD_EFX(ns) + a.d) * N + oc];
        double idf = mu * S[(int64_t)a.spec * N + oc] * efield;
        double dotn = idf * sgn;
        v = (dotn >= 0.0) ? (-idf / D) : 0.0;
    } else {
        double D = STATE ? spec_diff_st(M, a.spec, Te, a.Tg, nd, emag) : spec_diff(M, a.spec, Te, a.Tg);
        double cbar = sqrt(8.0 * VD_K_B * a.Tg / VD_PI / (M.molwt[a.spec] * VD_M_AMU));
        double flx = S[(int64_t)a.spec * N + a.sb.off(face[0], face[1], face[2])] * cbar / 4.0 * sgn;
        v = -flx / D;
    }
    a.out[ti] = v;
}

struct SpecRhsArgs {
    Arr<double> old, ds, sb;
    int spec;
    int3 lo;
    int nx, ny, nz;
    double idt;
    __device__ void rebase(int64_t o) {
        vd::rebase(old, o);
        vd::rebase(ds, o);
        vd::rebase(sb, o);
    }
};

// rhs = (1/dt) n_old + dsdt; initial guess = Sborder (ScalarSolve.cpp:443-465)
__global__ void k_spec_rhs(SpecRhsArgs a, double* __restrict__ f, double* __restrict__ x, Batch bt) {
    a.rebase(case_off(bt));
    f = rb(f, case_off(bt));
    x = rb(x, case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    f[t] = a.idt * a.old(i, j, k, a.spec) + 1.0 * a.ds(i, j, k);
    x[t] = a.sb(i, j, k, a.spec);
}

struct SpecPostArgs {
    double* x;
    Arr<double> phi, sb;
    int spec, ns, E, floor_on, energy;
    double floor_val, min_te;
    int3 lo;
    int nx, ny, nz;
    __device__ void rebase(int64_t o) {
        x = vd::rb(x, o);
        vd::rebase(phi, o);
        vd::rebase(sb, o);
    }
};

__global__ void k_spec_floor(SpecPostArgs a, Batch bt) {
    a.rebase(case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    if (a.floor_on && a.x[t] < a.floor_val) a.x[t] = a.floor_val;
}

// copy to phi_new and the Te update (ScalarSolve.cpp:632-667); FLOOR: the
// density floor of k_spec_floor applied on the way (when nothing reads the
// floored solution in between)
template <bool FLOOR>
__global__ void k_spec_post(SpecPostArgs a, Batch bt) {
    a.rebase(case_off(bt));
    int64_t t = (int64_t)lblock() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nx * a.ny * a.nz) return;
    int i, j, k;
    cell_of((unsigned)t, a.nx, a.ny, i, j, k);
    i += a.lo.x;
    j += a.lo.y;
    k += a.lo.z;
    double xv = a.x[t];
    if (FLOOR && a.floor_on && xv < a.floor_val) xv = a.floor_val;
    a.phi(i, j, k, a.spec) = xv;
    if (a.energy) {
        const int ns = a.ns;
        double te = VD_TWOTHIRD / VD_K_B * a.phi(i, j, k, VD_EEN(ns)) / a.sb(i, j, k, a.E);
        a.phi(i, j, k, VD_ETEMP(ns)) = te;
        if (te < a.min_te) {
            a.phi(i, j, k, VD_ETEMP(ns)) = a.min_te;
            a.phi(i, j, k, VD_EEN(ns)) = 1.5 * VD_K_B * a.phi(i, j, k, a.E) * a.min_te;
        }
    }
}

static int implicit_solve_body(vd_ctx* ctx, const vd_mech* m, int spec, const vd_array4* sb,
                               const vd_array4* old, const vd_array4* ds, vd_array4* phi, vd_array4 grad[3],
                               const vd_box* box, const vd_geom* g, const int bclo[3], const int bchi[3],
                               const vd_params* prm, double dt, vd_solve_info* info);

// one implicit species solve, timed as a whole (VD_KT_SPECIES) with the
// canonical bytes: n_old, dsdt and the initial guess read, the solution
// written, the cell coefficient once (n_b = 1, or 0 for a uniform D), each
// smoother iteration (3 + n_b) arrays, and for the energy the Te update
// (EEN, n_e read, Te written)
int implicit_solve(vd_ctx* ctx, const vd_mech* m, int spec, const vd_array4* sb, const vd_array4* old,
                   const vd_array4* ds, vd_array4* phi, vd_array4 grad[3], const vd_box* box,
                   const vd_geom* g, const int bclo[3], const int bchi[3], const vd_params* prm,
                   double dt, vd_solve_info* info) {
    vd_solve_info local;
    if (!info) info = &local;
    info->iterations = 0;
    kt_begin(ctx, VD_KT_SPECIES);
    const int st = implicit_solve_body(ctx, m, spec, sb, old, ds, phi, grad, box, g, bclo, bchi, prm, dt, info);
    const double nc = (double)(box->hi[0] - box->lo[0] + 1) * (box->hi[1] - box->lo[1] + 1) *
                      (box->hi[2] - box->lo[2] + 1);
    const int nb = spec < m->hdev.nspec && diff_is_uniform(m->hdev, spec) ? 0 : 1;
    kt_end(ctx, VD_KT_SPECIES,
           8.0 * nc * (4 + nb + info->iterations * (3 + nb) + (spec == m->hdev.nspec ? 3 : 0)));
    return st;
}

static int implicit_solve_body(vd_ctx* ctx, const vd_mech* m, int spec, const vd_array4* sb,
                               const vd_array4* old, const vd_array4* ds, vd_array4* phi, vd_array4 grad[3],
                               const vd_box* box, const vd_geom* g, const int bclo[3], const int bchi[3],
                               const vd_params* prm, double dt, vd_solve_info* info) {
    const int ns = m->hdev.nspec, E = m->hdev.e_idx;
    if (spec < 0 || spec > ns) {
        set_error("implicit solve: species %d out of range", spec);
        return VD_EINVAL;
    }
    VD_TRY(check_bc(g, bclo, bchi));
    MGWork* W;
    VD_TRY(get_work(ctx, g, box, prm->mg_precision, 0, prm->linsolve_max_coarsening_level, &W));
    int n[3], lo[3];
    dims_of(box, g, n, lo);
    int64_t nc = ncell(n);
    const double Tg = prm->gas_temperature;
    // a diffusion coefficient that does not depend on Te (constant or
    // Einstein at the gas temperature) is uniform, and so is every harmonic
    // face average of it: the solve then runs on the constant-coefficient
    // stencil (no cell-coefficient array, no per-face averages)
    const bool constd = diff_is_uniform(m->hdev, spec);
    // the coefficients of this solve depend on more of the state than Te
    // (fits; the energy follows the electron's)
    bool cd_state;
    {
        const DevMech& h = m->hdev;
        const int ef = spec == ns ? h.e_idx : spec;
        cd_state = h.dfit_kind[spec] == VD_TR_FIT || h.mfit_kind[spec] == VD_TR_FIT ||
                   h.dfit_kind[ef] == VD_TR_FIT || h.mfit_kind[ef] == VD_TR_FIT;
    }
    double bc = 0.0;
    if (constd) {
        const double D = spec_diff(m->hdev, spec, 0.0, Tg);
        bc = (D + D == 0.0) ? 0.0 : (2.0 * D * D) / (D + D);
    } else {
        CellDArgs cd;
        cd.M = m->dev;
        cd.sb = arr(sb);
        cd.spec = spec;
        cd.ns = ns;
        cd.dim = g->dim;
        cd.lo = make_int3(lo[0], lo[1], lo[2]);
        cd.gnx = n[0] + 2;
        cd.gny = n[1] + 2;
        cd.gnz = g->dim == 3 ? n[2] + 2 : 1;
        cd.Tg = Tg;
        cd.state = cd_state;
        const dim3 gcd = grid1d((int64_t)cd.gnx * cd.gny * cd.gnz, 256);
        if (cd.state) hipLaunchKernelGGL(k_cell_diff<true>, bz(ctx, gcd), dim3(256), 0, ctx->stream, cd, W->dcell, ctx->bt);
        else hipLaunchKernelGGL(k_cell_diff<false>, bz(ctx, gcd), dim3(256), 0, ctx->stream, cd, W->dcell, ctx->bt);
    }
    {
        SpecBcArgs b;
        b.M = m->dev;
        b.sb = arr(sb);
        b.spec = spec;
        b.d = b.side = 0;
        b.ns = ns;
        b.dim = g->dim;
        b.lo = make_int3(lo[0], lo[1], lo[2]);
        b.nx = n[0];
        b.ny = n[1];
        b.nz = n[2];
        b.Tg = Tg;
        b.out = nullptr;
        SpecBcSides sides;
        sides.nside = 0;
        int64_t ntmax = 0;
        for (int d = 0; d < g->dim; d++) {
            if (W->per[d]) continue;
            for (int s = 0; s < 2; s++) {
                sides.d[sides.nside] = d;
                sides.side[sides.nside] = s;
                sides.out[sides.nside] = W->bcv[d][s];
                sides.nside++;
            }
            ntmax = std::max(ntmax, ntanf(n, d));
        }
        if (sides.nside > 0) {
            dim3 gb = grid1d(ntmax, 256);
            gb.y = (unsigned)sides.nside;
            if (cd_state) hipLaunchKernelGGL(k_species_bc<true>, bz(ctx, gb), dim3(256), 0, ctx->stream, b, sides, ctx->bt);
            else hipLaunchKernelGGL(k_species_bc<false>, bz(ctx, gb), dim3(256), 0, ctx->stream, b, sides, ctx->bt);
        }
    }
    // Robin species sides: the case's user_transport::species_bc data, or
    // the reference default (a, b, f) = (0, 1, 0)
    for (int d = 0; d < g->dim; d++) {
        if (W->per[d]) continue;
        for (int s = 0; s < 2; s++)
            if ((s ? bchi[d] : bclo[d]) == VD_ROBINBC) {
                int64_t nt = ntanf(n, d);
                const bool user = m->host.robin_set[spec] != 0;
                for (int q = 0; q < 3; q++)
                    VD_TRY(fill_value(ctx, W->rob[d][s][q], nt,
                                      user ? m->host.robin[spec][d][s][q] : (q == 1 ? 1.0 : 0.0)));
            }
    }
    const double alpha = 1.0 / dt, beta = 1.0;
    const double* dcell = constd ? nullptr : W->dcell;
    // constant coefficient in 3-D on the streamed layout: the rhs, its
    // boundary terms and the initial residual's norms in one pass (the
    // boundary setup first, keeping each face's cinh); else the rhs pass
    // and then the boundary terms added to it
    if (constd && g->dim == 3 && W->stream && n[0] % 2 == 0) {
        VD_TRY(side_setup(ctx, *W, bclo, bchi, beta, bc, dcell, constd, true));
        const LevV<double> v0 = fine_view(*W, bclo, bchi, alpha, beta, bc, dcell);
        SpecR0Args r;
        r.old = arr(old);
        r.ds = arr(ds);
        r.sb = arr(sb);
        r.spec = spec;
        r.lo = make_int3(lo[0], lo[1], lo[2]);
        r.idt = 1.0 / dt;
        for (int d = 0; d < 3; d++) {
            r.sbd[d] = d < g->dim ? beta / (W->dx[d] * W->dx[d]) * bc : 0.0;
            for (int sd = 0; sd < 2; sd++) r.sh[d][sd] = W->per[d] || d >= g->dim ? nullptr : W->sh[d][sd];
        }
        r.f = W->f;
        r.x = W->x;
        r.slots = slot_group(*W, 0);
        const int ntx = (n[0] / 2 + RZ2_X - 1) / RZ2_X, nty = (n[1] + RZ2_Y - 1) / RZ2_Y;
        const void* kp = v0.c0uni ? (const void*)&k_spec_rhs_r0<true> : (const void*)&k_spec_rhs_r0<false>;
        const int zc = stream_zc(ctx, kp, RZ2_X * RZ2_Y, n, ntx * nty);
        const dim3 gz((unsigned)(ntx * nty * ((n[2] + zc - 1) / zc)));
        if (v0.c0uni)
            hipLaunchKernelGGL(k_spec_rhs_r0<true>, bz(ctx, gz), dim3(RZ2_X * RZ2_Y), 0, ctx->stream, v0, r, zc, ntx,
                               nty, ctx->bt);
        else
            hipLaunchKernelGGL(k_spec_rhs_r0<false>, bz(ctx, gz), dim3(RZ2_X * RZ2_Y), 0, ctx->stream, v0, r, zc, ntx,
                               nty, ctx->bt);
        VD_CHECK_LAUNCH();
        W->r0_done = true;
    } else {
        SpecRhsArgs r;
        r.old = arr(old);
        r.ds = arr(ds);
        r.sb = arr(sb);
        r.spec = spec;
        r.lo = make_int3(lo[0], lo[1], lo[2]);
        r.nx = n[0];
        r.ny = n[1];
        r.nz = n[2];
        r.idt = 1.0 / dt;
        hipLaunchKernelGGL(k_spec_rhs, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, r, W->f, W->x, ctx->bt);
        VD_CHECK_LAUNCH();
        VD_TRY(side_setup(ctx, *W, bclo, bchi, beta, bc, dcell, constd));
    }
    const bool floor_on = prm->bound_specden && spec != ns;
    const double floor_val = spec == E ? prm->min_electron_density : prm->min_species_density;
    // a constant-coefficient solve whose solution nothing reads before the
    // copy-back: its fused sweeps write each iterate to the state as well
    // (when the component's extent has 32-bit byte offsets)
    SweepOut so{};
    if (constd && spec != ns && !(spec == E && grad)) {
        const Arr<double> P = arr(phi);
        const int64_t ext = (int64_t)(n[0] - 1) + (int64_t)P.js * (n[1] - 1) + (int64_t)P.ks * (n[2] - 1) + 1;
        if (ext * 8 < ((int64_t)1 << 31)) {
            so.p = &P(lo[0], lo[1], lo[2], spec);
            so.js = P.js;
            so.ks = P.ks;
            so.bytes = (unsigned)(ext * 8);
            so.floor_on = floor_on;
            so.floor_val = floor_val;
        }
    }
    W->post = so;
    // smoother-only iterations when the operator is diagonally dominant
    // enough, else V-cycles (on harmonic face arrays of the cell coefficient)
    int st = mg_iterate(ctx, *W, constd ? 1 : 2, alpha, beta, bc, dcell, bclo, bchi, 2, prm->linsolve_reltol,
                        prm->linsolve_abstol, prm->linsolve_maxiter, prm->mg_pre_smooth, prm->mg_post_smooth,
                        info);
    W->post = SweepOut{};
    if (st != VD_OK && st != VD_ENOCONV) return st;
    SpecPostArgs p;
    p.x = W->x;
    p.phi = arr(phi);
    p.sb = arr(sb);
    p.spec = spec;
    p.ns = ns;
    p.E = E;
    p.floor_on = floor_on;
    p.floor_val = floor_val;
    p.energy = spec == ns;
    p.min_te = prm->min_electron_temp;
    p.lo = make_int3(lo[0], lo[1], lo[2]);
    p.nx = n[0];
    p.ny = n[1];
    p.nz = n[2];
    if (spec == E && grad) {
        // the electron gradient faces see the floored density
        hipLaunchKernelGGL(k_spec_floor, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, p, ctx->bt);
        VD_TRY(grad_faces(ctx, *W, bclo, bchi, grad, box, 1.0));
        hipLaunchKernelGGL(k_spec_post<false>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, p, ctx->bt);
    } else {
        // the cases whose final iterate the fused sweeps did not write
        int pc[VD_MAX_BATCH], npc = 0;
        for (int z = 0; z < ctx->bn; z++)
            if (!((W->post_mask >> ctx->bt.idx[z]) & 1u)) pc[npc++] = ctx->bt.idx[z];
        if (npc == ctx->bn) {
            hipLaunchKernelGGL(k_spec_post<true>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, p, ctx->bt);
        } else if (npc > 0) {
            BatchScope some(ctx, pc, npc);
            hipLaunchKernelGGL(k_spec_post<true>, bz(ctx, grid1d(nc, 256)), dim3(256), 0, ctx->stream, p, ctx->bt);
        }
    }
    VD_CHECK_LAUNCH();
    return st;
}

// ------------------------------------------------ generic solve
int mlmg_solve(vd_ctx* ctx, const vd_mlmg_problem* P, vd_array4* phi, const vd_array4* rhs,
               const vd_box* box, const vd_geom* g, vd_solve_info* info) {
    VD_TRY(check_bc(g, P->bc_lo, P->bc_hi));
    int constb = (P->bface[0] == nullptr);
    MGWork* W;
    VD_TRY(get_work(ctx, g, box, P->precision, constb, P->max_coarsening_level, &W));
    int n[3], lo[3];
    dims_of(box, g, n, lo);
    int64_t nc = ncell(n);
    FlatArgs fa = flat_args(rhs, 0, lo, n);
    hipLaunchKernelGGL(k_a4_to_flat, grid1d(nc, 256), dim3(256), 0, ctx->stream, fa, W->f);
    fa = flat_args(phi, 0, lo, n);
    hipLaunchKernelGGL(k_a4_to_flat, grid1d(nc, 256), dim3(256), 0, ctx->stream, fa, W->x);
    if (!constb)
        for (int d = 0; d < g->dim; d++) {
            int fn[3] = {n[0] + (d == 0), n[1] + (d == 1), n[2] + (d == 2)};
            FlatArgs fb = flat_args(P->bface[d], 0, lo, fn);
            hipLaunchKernelGGL(k_a4_to_flat, grid1d(ncell(fn), 256), dim3(256), 0, ctx->stream, fb,
                               W->bf0[d]);
        }
    for (int d = 0; d < g->dim; d++) {
        if (W->per[d]) continue;
        for (int s = 0; s < 2; s++) {
            int64_t nt = ntanf(n, d);
            dim3 gg = grid1d(nt, 256);
            int3 l3 = make_int3(lo[0], lo[1], lo[2]);
            if (P->bcval)
                hipLaunchKernelGGL(k_side_from_ghost, gg, dim3(256), 0, ctx->stream, arr(P->bcval), d, s,
                                   l3, n[0], n[1], n[2], g->dim, W->bcv[d][s]);
            else
                VD_TRY(fill_value(ctx, W->bcv[d][s], nt, 0.0));
            int type = s ? P->bc_hi[d] : P->bc_lo[d];
            if (type == VD_ROBINBC) {
                const vd_array4* rr[3] = {P->robin_a, P->robin_b, P->robin_f};
                const double dflt[3] = {0.0, 1.0, 0.0};
                for (int q = 0; q < 3; q++) {
                    if (rr[q])
                        hipLaunchKernelGGL(k_side_from_ghost, gg, dim3(256), 0, ctx->stream, arr(rr[q]), d,
                                           s, l3, n[0], n[1], n[2], g->dim, W->rob[d][s][q]);
                    else
                        VD_TRY(fill_value(ctx, W->rob[d][s][q], nt, dflt[q]));
                }
            }
        }
    }
    VD_CHECK_LAUNCH();
    VD_TRY(side_setup(ctx, *W, P->bc_lo, P->bc_hi, P->beta, 1.0));
    const int bm = constb ? 1 : 0;
    double rho;
    VD_TRY(jacobi_rho(ctx, *W, bm, P->alpha_a, P->beta, 1.0, &rho));
    int st = mg_iterate(ctx, *W, bm, P->alpha_a, P->beta, 1.0, nullptr, P->bc_lo, P->bc_hi,
                        rho <= SMOOTH_ONLY_RHO, P->reltol, P->abstol, P->maxiter, P->pre_smooth,
                        P->post_smooth, info);
    if (st != VD_OK && st != VD_ENOCONV) return st;
    fa = flat_args(phi, 0, lo, n);
    hipLaunchKernelGGL(k_flat_to_a4, grid1d(nc, 256), dim3(256), 0, ctx->stream, fa, W->x);
    VD_CHECK_LAUNCH();
    return st;
}

}  // namespace vd

using namespace vd;

extern "C" int vd_solve_potential(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                  vd_array4* phi_new, vd_array4 efield[3], const vd_box* box,
                                  const vd_geom* geom, const vd_params* prm, double time,
                                  vd_solve_info* info) {
    if (!ctx || !m || !sborder || !phi_new || !efield || !box || !geom || !prm) return VD_EINVAL;
    VD_TRY(check_extents("solve_potential", box, {sborder, phi_new}));
    VD_TRY(check_face_extents("solve_potential", efield, geom->dim));
    if (prm->mg_precision != 0 && prm->mg_precision != 1) return VD_EINVAL;
    return solve_potential(ctx, m, sborder, phi_new, efield, box, geom, prm, time, info);
}

extern "C" int vd_implicit_solve_scalar(vd_ctx* ctx, const vd_mech* m, int spec, const vd_array4* sborder,
                                        const vd_array4* phi_old, const vd_array4* dsdt_expl,
                                        vd_array4* phi_new, vd_array4 grad_fc[3], const vd_box* box,
                                        const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                                        const vd_params* prm, double time, double dt,
                                        vd_solve_info* info) {
    (void)time;
    if (!ctx || !m || !sborder || !phi_old || !dsdt_expl || !phi_new || !box || !geom || !prm)
        return VD_EINVAL;
    VD_TRY(check_extents("implicit_solve_scalar", box, {sborder, phi_old, dsdt_expl, phi_new}));
    VD_TRY(check_face_extents("implicit_solve_scalar", grad_fc, geom->dim));
    if (prm->mg_precision != 0 && prm->mg_precision != 1) return VD_EINVAL;
    if (!(dt > 0.0)) {
        set_error("implicit solve: dt must be positive");
        return VD_EINVAL;
    }
    return implicit_solve(ctx, m, spec, sborder, phi_old, dsdt_expl, phi_new, grad_fc, box, geom, bc_lo,
                          bc_hi, prm, dt, info);
}

extern "C" int vd_mlmg_solve(vd_ctx* ctx, const vd_mlmg_problem* prob, vd_array4* phi,
                             const vd_array4* rhs, const vd_box* box, const vd_geom* geom,
                             vd_solve_info* info) {
    if (!ctx || !prob || !phi || !rhs || !box || !geom) return VD_EINVAL;
    VD_TRY(check_extents("mlmg_solve", box, {phi, rhs, prob->bcval, prob->robin_a, prob->robin_b, prob->robin_f}));
    for (int d = 0; d < geom->dim && d < 3; d++) VD_TRY(check_extents("mlmg_solve", nullptr, {prob->bface[d]}));
    VD_TRY(require_plain(ctx, "mlmg_solve"));
    if (prob->precision != 0 && prob->precision != 1) return VD_EINVAL;
    return mlmg_solve(ctx, prob, phi, rhs, box, geom, info);
}
