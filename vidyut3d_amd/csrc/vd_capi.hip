// vd_capi.hip -- context, memory, case plug-in upload and kernel timing.
#include <algorithm>
#include <cfloat>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "vd_internal.hpp"

namespace vd {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// Workgroups of `kernel` (block threads) one CU keeps resident, times the
// CU count: the device's capacity for that launch shape.  Cached per kernel
// and device (contexts on several host threads share the cache).
int resident_blocks(vd_ctx* ctx, const void* kernel, int block) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    const std::pair<const void*, int> key(kernel, ctx->device * 4096 + block);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu <= 0)
        ncu = 256;
    const int cap = per_cu * ncu;
    std::lock_guard<std::mutex> g(mu);
    cache[key] = cap;
    return cap;
}

// The chunk count n in [nmin, nmax] whose launch of n * per_chunk blocks
// fills the device's resident capacity `cap` best over whole rounds (a
// last round that is mostly empty leaves most CUs idle while the long
// streaming workgroups of the round before finish); the fewest chunks
// among equally good counts.
int round_filling_chunks(int64_t per_chunk, int cap, int nmin, int nmax) {
    if (nmax < nmin) nmax = nmin;
    int best = nmin;
    double beff = -1.0;
    for (int n = nmin; n <= nmax; n++) {
        const int64_t b = per_chunk * n;
        const int64_t rounds = (b + cap - 1) / cap;
        const double eff = (double)b / (double)(rounds * cap);
        if (eff > beff + 1e-9) {
            beff = eff;
            best = n;
        }
    }
    return best;
}

int dev_alloc(vd_ctx* ctx, size_t bytes, void** out) {
    if (ctx->arena) {
        // 256-byte pieces of case 0's region; every case's copy of a piece
        // has the same alignment (regions are whole multiples of 4 KiB)
        const size_t sz = ((bytes ? bytes : 8) + 255) & ~(size_t)255;
        if (ctx->arena_used + sz > ctx->arena_case) {
            set_error("batched context: %zu more bytes per case needed, %zu of %zu used", sz, ctx->arena_used,
                      ctx->arena_case);
            return VD_ENOMEM;
        }
        *out = ctx->arena + ctx->arena_used;
        ctx->arena_used += sz;
        return VD_OK;
    }
    hipError_t e = hipMalloc(out, bytes ? bytes : 8);
    if (e != hipSuccess) {
        set_error("hipMalloc %zu: %s", bytes, hipGetErrorString(e));
        *out = nullptr;
        return VD_ENOMEM;
    }
    return VD_OK;
}

void dev_free(vd_ctx* ctx, void* p) {
    if (p && !ctx->arena) (void)hipFree(p);
}

BatchScope::BatchScope(vd_ctx* ctx, const int* cases, int n) : c(ctx), bt(ctx->bt), bn(ctx->bn) {
    for (int i = 0; i < n; i++) ctx->bt.idx[i] = (unsigned char)cases[i];
    ctx->bn = n;
}
BatchScope::~BatchScope() {
    c->bt = bt;
    c->bn = bn;
}

int require_plain(vd_ctx* ctx, const char* what) {
    if (ctx->arena) {
        set_error("%s: not available in a batched context", what);
        return VD_EINVAL;
    }
    return VD_OK;
}

int scratch(vd_ctx* ctx, const std::string& name, size_t bytes, void** out) {
    Buf& b = ctx->scratch[name];
    if (b.bytes < bytes) {
        if (b.p) {
            VD_HIP(hipStreamSynchronize(ctx->stream));
            dev_free(ctx, b.p);
        }
        b.p = nullptr;
        b.bytes = 0;
        size_t sz = bytes < 256 ? 256 : bytes;
        if (dev_alloc(ctx, sz, &b.p) != VD_OK) {
            b.p = nullptr;  // dev_alloc's message names the size
            return VD_ENOMEM;
        }
        b.bytes = sz;
    }
    *out = b.p;
    return VD_OK;
}

static hipEvent_t kt_event(vd_ctx* ctx) {
    if (ctx->kt_next < ctx->kt_pool.size()) return ctx->kt_pool[ctx->kt_next++];
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->kt_pool.push_back(e);
    ctx->kt_next++;
    return e;
}

static bool kt_on(const vd_ctx* ctx, int family) {
    return family > 0 && family < 32 && (ctx->kt_mask >> family) & 1u;
}

void kt_begin(vd_ctx* ctx, int family) {
    if (!kt_on(ctx, family)) return;
    hipEvent_t a = kt_event(ctx), b = kt_event(ctx);
    if (!a || !b) return;
    (void)hipEventRecord(a, ctx->stream);
    ctx->kt_events.push_back({a, b, family});
    ctx->kt_open[family] = (long)ctx->kt_events.size();
}

void kt_end(vd_ctx* ctx, int family, double bytes) {
    if (!kt_on(ctx, family) || ctx->kt_open[family] == 0) return;
    (void)hipEventRecord(ctx->kt_events[ctx->kt_open[family] - 1].b, ctx->stream);
    ctx->kt_open[family] = 0;
    ctx->kt_bytes[family] += bytes;
}

}  // namespace vd

using namespace vd;

extern "C" {

const char* vd_last_error(void) { return g_err; }
const char* vd_version(void) { return "vidyut3d_amd 0.1.0 gfx950"; }

void vd_params_default(vd_params* p) {
    memset(p, 0, sizeof(*p));
    p->gas_temperature = 300.0;
    p->gas_pressure = 101325.0;
    p->hyp_order = 1;
    p->weno_scheme = 3;
    p->do_reactions = 1;
    p->do_transport = 1;
    p->do_spacechrg = 1;
    p->min_species_density = 1e-20;
    p->min_electron_density = 1e12;
    p->min_electron_temp = 300.0;
    p->ele_mob = -1.0e24;
    p->ele_diff = 1.0e24;
    p->voltage_freq = 1.0;
    p->linsolve_reltol = 1e-10;
    p->linsolve_abstol = 1e-10;
    p->linsolve_maxiter = 100;
    p->linsolve_max_coarsening_level = 10;
    p->mg_pre_smooth = 2;
    p->mg_post_smooth = 2;
    p->adaptive_dt = 0;
    p->advective_cfl = 0.7;
    p->diffusive_cfl = 1.0;
    p->dielectric_cfl = 1.0;
    p->dt_min = 1e-14;
    p->dt_max = 1e-9;
    p->adaptive_dt_delay = 0.0;
    p->dt_stretch = DBL_MAX;
    for (int d = 0; d < 3; d++) {
        p->pot_bc_lo[d] = p->pot_bc_hi[d] = VD_ROBINBC;
        p->eden_bc_lo[d] = p->eden_bc_hi[d] = VD_ROBINBC;
        p->eenrg_bc_lo[d] = p->eenrg_bc_hi[d] = VD_ROBINBC;
        p->ion_bc_lo[d] = p->ion_bc_hi[d] = VD_ROBINBC;
        p->neutral_bc_lo[d] = p->neutral_bc_hi[d] = VD_ROBINBC;
    }
}

int vd_compute_dt(const vd_params* prm, double cur_time, double old_dt, const double ts[3], double* dt) {
    if (!prm || !ts || !dt) return VD_EINVAL;
    double v = old_dt;
    if (prm->adaptive_dt && cur_time > prm->adaptive_dt_delay) {
        double trans_min_dt = DBL_MAX;
        if (prm->do_transport) trans_min_dt = std::min(ts[0] * prm->advective_cfl, ts[1] * prm->diffusive_cfl);
        double adp = std::min(trans_min_dt, ts[2] * prm->dielectric_cfl);
        if (adp > prm->dt_max) adp = prm->dt_max;
        if (adp < prm->dt_min) adp = prm->dt_min;
        v = std::min(old_dt * prm->dt_stretch, adp);
    }
    *dt = v;
    return VD_OK;
}

int vd_ctx_create(int device, vd_ctx** out) {
    if (!out) return VD_EINVAL;
    *out = nullptr;
    int n = 0;
    VD_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("device %d not present (%d devices)", device, n);
        return VD_EINVAL;
    }
    VD_HIP(hipSetDevice(device));
    vd_ctx* c = new vd_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return VD_EHIP;
    }
    e = hipHostMalloc((void**)&c->pinned, 8192 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->hnorm, HNORM_DOUBLES * sizeof(double),
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&c->hnorm_d, c->hnorm, 0);
    if (e != hipSuccess) {
        if (c->hnorm) (void)hipHostFree(c->hnorm);
        if (c->pinned) (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("hipHostMalloc: %s", hipGetErrorString(e));
        return VD_EHIP;
    }
    memset(c->hnorm, 0, HNORM_DOUBLES * sizeof(double));
    c->bt.stride = 0;
    memset(c->bt.idx, 0, sizeof(c->bt.idx));
    // the device address of the no-allocation scratch on this device (once)
    const int st = noalloc_setup(device);
    if (st != VD_OK) {
        (void)hipHostFree(c->hnorm);
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return st;
    }
    *out = c;
    return VD_OK;
}

int vd_ctx_destroy(vd_ctx* ctx) {
    if (!ctx) return VD_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->scratch) dev_free(ctx, kv.second.p);
    for (auto& kv : ctx->mg) mgwork_free(kv.second);
    for (auto e : ctx->kt_pool) (void)hipEventDestroy(e);
    if (ctx->arena) (void)hipFree(ctx->arena);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->hnorm) (void)hipHostFree(ctx->hnorm);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VD_OK;
}

int vd_ctx_sync(vd_ctx* ctx) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    return VD_OK;
}

int vd_ctx_release_workspaces(vd_ctx* ctx) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    // a batched context's regions hold the cases' state as well: its
    // workspaces live until the context is destroyed
    if (ctx->arena) return VD_OK;
    for (auto& kv : ctx->scratch)
        if (kv.second.p) VD_HIP(hipFree(kv.second.p));
    ctx->scratch.clear();
    for (auto& kv : ctx->mg) mgwork_free(kv.second);
    ctx->mg.clear();
    return VD_OK;
}

void* vd_ctx_stream(vd_ctx* ctx) { return (void*)ctx->stream; }

int vd_malloc(vd_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return VD_EINVAL;
    VD_HIP(hipSetDevice(ctx->device));
    return dev_alloc(ctx, bytes, out);
}

int vd_free(vd_ctx* ctx, void* p) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, p);
    return VD_OK;
}

int vd_ctx_create_batched(int device, int ncase, size_t case_bytes, vd_ctx** out) {
    if (!out) return VD_EINVAL;
    *out = nullptr;
    if (ncase < 1 || ncase > VD_MAX_BATCH || case_bytes == 0) {
        set_error("vd_ctx_create_batched: ncase %d (1 .. %d), case_bytes %zu", ncase, VD_MAX_BATCH, case_bytes);
        return VD_EINVAL;
    }
    vd_ctx* c;
    VD_TRY(vd_ctx_create(device, &c));
    const size_t region = (case_bytes + 4095) & ~(size_t)4095;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, region * (size_t)ncase);
    if (e != hipSuccess) {
        vd_ctx_destroy(c);
        set_error("vd_ctx_create_batched: hipMalloc %d x %zu bytes: %s", ncase, region, hipGetErrorString(e));
        return VD_ENOMEM;
    }
    c->arena = (char*)p;
    c->arena_case = region;
    c->ncase = ncase;
    c->bt.stride = (int64_t)region;
    for (int i = 0; i < ncase; i++) {
        c->bt.idx[i] = (unsigned char)i;
        c->act[i] = i;
    }
    c->bn = c->nact = ncase;
    *out = c;
    return VD_OK;
}

int vd_ctx_ncase(const vd_ctx* ctx) { return ctx ? ctx->ncase : 0; }

long long vd_ctx_case_stride(const vd_ctx* ctx) { return ctx ? (long long)ctx->bt.stride : 0; }

int vd_ctx_arena_used(const vd_ctx* ctx, size_t* used) {
    if (!ctx || !used) return VD_EINVAL;
    *used = ctx->arena_used;
    return VD_OK;
}

int vd_ctx_set_active(vd_ctx* ctx, int n, const int* cases) {
    if (!ctx || n < 1 || n > ctx->ncase || !cases) {
        set_error("vd_ctx_set_active: %d cases of %d", n, ctx ? ctx->ncase : 0);
        return VD_EINVAL;
    }
    int seen[VD_MAX_BATCH] = {0};
    for (int i = 0; i < n; i++) {
        if (cases[i] < 0 || cases[i] >= ctx->ncase || seen[cases[i]]) {
            set_error("vd_ctx_set_active: case %d repeated or outside 0 .. %d", cases[i], ctx->ncase - 1);
            return VD_EINVAL;
        }
        seen[cases[i]] = 1;
    }
    for (int i = 0; i < n; i++) {
        ctx->act[i] = cases[i];
        ctx->bt.idx[i] = (unsigned char)cases[i];
    }
    ctx->nact = ctx->bn = n;
    return VD_OK;
}

int vd_ctx_set_case_voltage(vd_ctx* ctx, int c, int use, double amp_1, double amp_2) {
    if (!ctx || c < 0 || c >= ctx->ncase) return VD_EINVAL;
    ctx->case_volt_set[c] = use != 0;
    ctx->case_volt[c][0] = amp_1;
    ctx->case_volt[c][1] = amp_2;
    return VD_OK;
}

int vd_ctx_case_solve_info(const vd_ctx* ctx, int c, vd_solve_info* info) {
    if (!ctx || !info || c < 0 || c >= ctx->ncase) return VD_EINVAL;
    *info = ctx->case_info[c];
    return VD_OK;
}

int vd_ctx_case_time_scales(const vd_ctx* ctx, int c, double out[3]) {
    if (!ctx || !out || c < 0 || c >= ctx->ncase) return VD_EINVAL;
    for (int q = 0; q < 3; q++) out[q] = ctx->case_ts[c][q];
    return VD_OK;
}

int vd_memcpy_h2d(vd_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    VD_HIP(hipStreamSynchronize(ctx->stream));
    return VD_OK;
}

int vd_memcpy_d2h(vd_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VD_HIP(hipStreamSynchronize(ctx->stream));
    return VD_OK;
}

int vd_memcpy_d2d(vd_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return VD_OK;
}

int vd_memset_zero_kernel(vd_ctx* ctx, void* dst, size_t bytes);

// a fill kernel, not hipMemsetAsync, so captured steps replay correctly
int vd_memset_zero(vd_ctx* ctx, void* dst, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    return vd_memset_zero_kernel(ctx, dst, bytes);
}

// the VD_CHEM_GEN fields of reaction r, checked, into the device copy
static int gen_reaction(const vd_mech_desc* d, int r, DevMech& h) {
    const vd_reaction& R = d->reac[r];
    const int ns = d->nspec;
    auto bad_sp = [&](int s) { return s < 0 || s >= ns; };
    if (R.rate_kind < VD_RATE_ARRH || R.rate_kind > VD_RATE_JANEV_FFIT) {
        set_error("reaction %d: rate kind %d unknown", r, R.rate_kind);
        return VD_EINVAL;
    }
    if (R.rate_kind == VD_RATE_TEPOW && !(R.tref > 0.0)) {
        set_error("reaction %d: Te power law needs tref > 0", r);
        return VD_EINVAL;
    }
    if (R.ng < 0 || R.ng > VD_GEN_NG || R.nrev < 0 || R.nrev > VD_MAX_PRODUCTS || R.nwd < 0 ||
        R.nwd > VD_GEN_NW || R.refc_pow < -3 || R.refc_pow > 3 || (R.reversible && R.ng == 0)) {
        set_error("reaction %d: bad reverse-rate or wdot counts", r);
        return VD_EINVAL;
    }
    for (int q = 0; q < R.ng; q++)
        if (bad_sp(R.g_sp[q])) {
            set_error("reaction %d: g/RT species index %d", r, R.g_sp[q]);
            return VD_EINVAL;
        }
    for (int q = 0; q < R.nrev; q++)
        if (bad_sp(R.rev_sp[q]) || !(R.rev_nu[q] >= 1.0) || R.rev_nu[q] != (double)(int)R.rev_nu[q]) {
            set_error("reaction %d: reverse factor %d", r, q);
            return VD_EINVAL;
        }
    for (int q = 0; q < R.nreact; q++)
        if (!(R.react_nu[q] >= 1.0) || R.react_nu[q] != (double)(int)R.react_nu[q]) {
            set_error("reaction %d: reactant order must be a positive integer", r);
            return VD_EINVAL;
        }
    for (int q = 0; q < R.nwd; q++)
        if (bad_sp(R.wd_sp[q])) {
            set_error("reaction %d: wdot species index %d", r, R.wd_sp[q]);
            return VD_EINVAL;
        }
    if (R.rxntype == 5 && bad_sp(R.elidx)) {
        set_error("reaction %d: elastic partner %d", r, R.elidx);
        return VD_EINVAL;
    }
    h.rkind[r] = R.rate_kind;
    h.tref[r] = R.tref;
    h.tgdiv[r] = R.tgas_div != 0;
    for (int q = 0; q < 9; q++) h.fit[q][r] = R.fit[q];
    h.A2[r] = R.A2;
    for (int q = 0; q < 4; q++) h.fit2[q][r] = R.fit2[q];
    h.fsw[r] = R.fit_switch_ev;
    h.kmax[r] = R.kmax;
    h.tb[r] = R.third_body != 0;
    h.rev[r] = R.reversible != 0;
    h.ng[r] = R.ng;
    for (int q = 0; q < R.ng; q++) {
        h.gsp[q][r] = R.g_sp[q];
        h.gnu[q][r] = R.g_nu[q];
    }
    h.refcp[r] = R.refc_pow;
    h.nrev[r] = R.nrev;
    for (int q = 0; q < R.nrev; q++) {
        h.revsp[q][r] = R.rev_sp[q];
        h.revnu[q][r] = R.rev_nu[q];
    }
    h.nwd[r] = R.nwd;
    for (int q = 0; q < R.nwd; q++) {
        h.wdsp[q][r] = R.wd_sp[q];
        h.wdnu[q][r] = R.wd_nu[q];
    }
    h.any_rev |= h.rev[r];
    h.any_tb |= h.tb[r];
    h.any_fit |= R.rate_kind >= VD_RATE_JANEV;
    return VD_OK;
}

int vd_mech_create(vd_ctx* ctx, const vd_mech_desc* d, vd_mech** out) {
    if (!ctx || !d || !out) return VD_EINVAL;
    if (d->nspec < 1 || d->nspec > VD_MAX_SPECIES || d->nreac < 0 || d->nreac > VD_MAX_REACTIONS ||
        d->e_idx < 0 || d->e_idx >= d->nspec) {
        set_error("mechanism: bad sizes nspec=%d nreac=%d e_idx=%d", d->nspec, d->nreac, d->e_idx);
        return VD_EINVAL;
    }
    if (d->chem_form != VD_CHEM_TABLE && d->chem_form != VD_CHEM_CKWC && d->chem_form != VD_CHEM_GEN) {
        set_error("mechanism: unknown chem_form %d", d->chem_form);
        return VD_EINVAL;
    }
    vd_mech* m = new vd_mech();
    m->ctx = ctx;
    m->host = *d;
    DevMech& h = m->hdev;
    memset(&h, 0, sizeof(h));
    h.nspec = d->nspec;
    h.nreac = d->nreac;
    h.e_idx = d->e_idx;
    h.chem_form = d->chem_form;
    h.elastic_loss = d->elastic_loss;
    h.elastic_partner_amu = d->elastic_partner_amu;
    for (int s = 0; s < d->nspec; s++) {
        h.charge[s] = d->charge[s];
        h.molwt[s] = d->molwt[s];
        h.mob[s] = d->mob[s];
        h.diff_kind[s] = d->diff_kind[s];
        h.diff[s] = d->diff[s];
    }
    for (int s = 0; s <= d->nspec; s++) {
        const vd_transport_fit* f[2] = {&d->mob_fit[s], &d->diff_fit[s]};
        for (int q = 0; q < 2; q++)
            if (f[q]->kind != VD_TR_DEFAULT && f[q]->kind != VD_TR_FIT) {
                delete m;
                set_error("mechanism: species %d %s fit kind %d unknown", s, q ? "diffusion" : "mobility",
                          f[q]->kind);
                return VD_EINVAL;
            }
        h.mfit_kind[s] = d->mob_fit[s].kind;
        h.mfit_sign[s] = d->mob_fit[s].sign;
        h.dfit_kind[s] = d->diff_fit[s].kind;
        h.dfit_sign[s] = d->diff_fit[s].sign;
        for (int c = 0; c < VD_TR_NCOEF; c++) {
            h.mfit_c[s][c] = d->mob_fit[s].c[c];
            h.dfit_c[s][c] = d->diff_fit[s].c[c];
        }
        h.vel_set[s] = d->vel_set[s] != 0;
        for (int q = 0; q < 3; q++) h.vel[s][q] = d->vel[s][q];
        h.user_src[s] = d->user_src[s];
        if (d->user_src[s] != 0.0) h.any_user_src = 1;
    }
    for (int r = 0; r < d->nreac; r++) {
        const vd_reaction& R = d->reac[r];
        if (R.nreact < 0 || R.nreact > VD_MAX_REACTANTS || R.nprod < 0 || R.nprod > VD_MAX_PRODUCTS) {
            delete m;
            set_error("reaction %d: bad reactant/product count", r);
            return VD_EINVAL;
        }
        h.A[r] = R.A;
        h.beta[r] = R.beta;
        h.Ea[r] = R.Ea;
        h.te[r] = R.temp_is_te;
        h.energy[r] = R.energy;
        h.rxntype[r] = R.rxntype;
        h.eexci[r] = R.eexci;
        h.elidx[r] = R.elidx;
        h.nreact[r] = R.nreact;
        h.nprod[r] = R.nprod;
        int order = 0;
        for (int q = 0; q < R.nreact; q++) {
            if (R.react_sp[q] < 0 || R.react_sp[q] >= d->nspec) {
                delete m;
                set_error("reaction %d: reactant index %d", r, R.react_sp[q]);
                return VD_EINVAL;
            }
            h.rsp[q][r] = R.react_sp[q];
            h.rnu[q][r] = R.react_nu[q];
            order += (int)R.react_nu[q];
        }
        h.order[r] = order;
        for (int q = 0; q < R.nprod; q++) {
            if (R.prod_sp[q] < 0 || R.prod_sp[q] >= d->nspec) {
                delete m;
                set_error("reaction %d: product index %d", r, R.prod_sp[q]);
                return VD_EINVAL;
            }
            h.psp[q][r] = R.prod_sp[q];
            h.pnu[q][r] = R.prod_nu[q];
        }
        // net stoichiometry (ascending species), generated-code form
        int nn = 0;
        for (int sp = 0; sp < d->nspec; sp++) {
            double net = 0.0;
            int hit = 0;
            for (int q = 0; q < R.nreact; q++)
                if (R.react_sp[q] == sp) {
                    net -= R.react_nu[q];
                    hit = 1;
                }
            for (int q = 0; q < R.nprod; q++)
                if (R.prod_sp[q] == sp) {
                    net += R.prod_nu[q];
                    hit = 1;
                }
            if (!hit || net == 0.0) continue;
            h.netsp[nn][r] = sp;
            h.netnu[nn][r] = net;
            nn++;
        }
        h.nnet[r] = nn;
        if (d->chem_form == VD_CHEM_GEN) {
            const int bad = gen_reaction(d, r, h);
            if (bad) {
                delete m;
                return bad;
            }
        }
    }
    if (d->chem_form == VD_CHEM_GEN)
        for (int k = 0; k < d->nspec; k++) {
            h.gtmid[k] = d->gibbs_tmid[k];
            for (int q = 0; q < 7; q++) {
                h.glo[k][q] = d->gibbs_lo[k][q];
                h.ghi[k][q] = d->gibbs_hi[k][q];
            }
        }
    // in a batched context every case holds a copy (at its own offset), so
    // kernels may move any pointer they take by the case's offset
    if (dev_alloc(ctx, sizeof(DevMech), (void**)&m->dev) != VD_OK) {
        delete m;
        return VD_ENOMEM;
    }
    hipError_t e = hipSuccess;
    for (int c = 0; c < ctx->ncase && e == hipSuccess; c++)
        e = hipMemcpy((char*)m->dev + (size_t)c * ctx->bt.stride, &h, sizeof(DevMech), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dev_free(ctx, m->dev);
        delete m;
        set_error("mechanism upload: %s", hipGetErrorString(e));
        return VD_EHIP;
    }
    if ((h.chem_form == VD_CHEM_TABLE || h.chem_form == VD_CHEM_GEN) && h.nreac > 0) {
        const char* env = getenv("VD_CHEM_SPECIALISE");
        if (!(env && strcmp(env, "0") == 0)) {
            if (chem_jit_load(m) == VD_OK) m->use_spec = 1;
            else fprintf(stderr, "vidyut3d_amd: specialised source kernel unavailable (%s); "
                                 "using the table kernel\n", vd_last_error());
        }
    }
    *out = m;
    return VD_OK;
}

int vd_mech_destroy(vd_mech* m) {
    if (!m) return VD_EINVAL;
    chem_jit_unload(m);
    dev_free(m->ctx, m->dev);
    delete m;
    return VD_OK;
}

int vd_mech_specialise(vd_mech* m, int enable) {
    if (!m) return VD_EINVAL;
    if (!enable) {
        m->use_spec = 0;
        return VD_OK;
    }
    VD_TRY(chem_jit_load(m));
    m->use_spec = 1;
    return VD_OK;
}

int vd_mech_is_specialised(const vd_mech* m) { return m && m->use_spec && m->fn ? 1 : 0; }

int vd_mech_find_id(const vd_mech* m, const char* name) {
    for (int s = 0; s < m->host.nspec; s++)
        if (strncmp(m->host.names[s], name, 16) == 0) return s;
    return -1;
}

int vd_ctx_set_kernel_timing_mask(vd_ctx* ctx, unsigned mask) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->kt_mask = mask & ~1u;
    ctx->kt_events.clear();
    ctx->kt_next = 0;
    for (int f = 0; f < 32; f++) {
        ctx->kt_bytes[f] = 0.0;
        ctx->kt_open[f] = 0;
    }
    return VD_OK;
}

int vd_ctx_set_cfl_efield(vd_ctx* ctx, const vd_array4* efield) {
    if (!ctx) return VD_EINVAL;
    ctx->cfl_ef_set = efield != nullptr;
    if (efield) ctx->cfl_ef = *efield;
    return VD_OK;
}

int vd_ctx_set_kernel_timing(vd_ctx* ctx, int family) {
    if (!ctx || family < 0 || family >= 32) return VD_EINVAL;
    return vd_ctx_set_kernel_timing_mask(ctx, family == VD_KT_NONE ? 0u : 1u << family);
}

int vd_ctx_kernel_timing_family(vd_ctx* ctx, int family, double* total_ms, long long* launches,
                                double* bytes) {
    if (!ctx) return VD_EINVAL;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    long long n = 0;
    double by = 0.0;
    for (auto& e : ctx->kt_events) {
        if (family != VD_KT_NONE && e.family != family) continue;
        float ms = 0.f;
        VD_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        tot += ms;
        n++;
    }
    for (int f = 1; f < 32; f++)
        if (family == VD_KT_NONE || f == family) by += ctx->kt_bytes[f];
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    if (bytes) *bytes = by;
    return VD_OK;
}

int vd_ctx_kernel_timing(vd_ctx* ctx, double* total_ms, long long* launches, double* bytes) {
    return vd_ctx_kernel_timing_family(ctx, VD_KT_NONE, total_ms, launches, bytes);
}

}  // extern "C"
