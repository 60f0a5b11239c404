// vd_noalloc.hip -- entry points that allocate no device memory.
//
// vd_transport_rhs_noalloc is vd_transport_rhs with its face-flux scratch
// arrays carved from `scratch`, a file-scope device array of fixed size,
// instead of the context's grow-only buffers (hipMalloc on first use and on
// growth).  For callers that cannot allocate: code under HIP graph capture,
// tight serving loops.  Every GPU holds its own copy of `scratch` from the
// library's load until the process ends.  Host code never uses the array's
// name as an address: its device address on each GPU is taken once, with
// that GPU current, by noalloc_setup (called from vd_ctx_create) and kept
// per device.
#include <cstdint>
#include <mutex>

#include "vd_internal.hpp"

// 64 MiB: the x (and, in 3-D, y) face fluxes of a WENO box up to about
// 160^3 cells (with per-cell drift velocities too, smaller boxes)
static __device__ __align__(256) double scratch[VD_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch_ptr[MAX_DEVICES] = {nullptr};
}  // namespace

namespace vd {

int transport_rhs(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, const vd_array4* rxn, vd_array4* out,
                  const vd_box* box, const vd_geom* g, const int bc_lo[3], const int bc_hi[3],
                  const vd_params* prm, double dt, int specid, double* const* fbuf, double* vbuf);
int transport_rhs_nflux(const vd_params* prm, const vd_geom* g);
int64_t drift_vel_doubles(const vd_array4* sb, int dim);
int64_t transport_face_count(const vd_box* box, int d);
int check_transport_args(const vd_params* prm, int specid, int nspec);

// the device address of `scratch` on `device`, which must be current;
// once per device
int noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("no-allocation scratch: device %d beyond the %d supported", device, MAX_DEVICES);
        return VD_EINVAL;
    }
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch_ptr[device]) return VD_OK;
    void* p = nullptr;
    // ::scratch, the array (vd::scratch is the context's buffer function)
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p) {
        set_error("no-allocation scratch: hipGetSymbolAddress on device %d: %s", device,
                  hipGetErrorString(e));
        return VD_EHIP;
    }
    g_scratch_ptr[device] = p;
    return VD_OK;
}

static void* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_scratch_ptr[device];
}

}  // namespace vd

using namespace vd;

extern "C" int vd_transport_rhs_noalloc(vd_ctx* ctx, const vd_mech* m, const vd_array4* sborder,
                                        const vd_array4* rxn_src, vd_array4* dsdt_out, const vd_box* box,
                                        const vd_geom* geom, const int bc_lo[3], const int bc_hi[3],
                                        const vd_params* prm, double time, double dt, int specid) {
    (void)time;
    if (!ctx || !m || !sborder || !rxn_src || !dsdt_out || !box || !geom || !prm || !bc_lo || !bc_hi)
        return VD_EINVAL;
    VD_TRY(check_extents("transport_rhs_noalloc", box, {sborder, rxn_src, dsdt_out}));
    VD_TRY(check_transport_args(prm, specid, m->hdev.nspec));
    VD_TRY(require_plain(ctx, "transport_rhs_noalloc"));
    for (int d = 0; d < 3; d++)
        if (box->hi[d] < box->lo[d]) {
            set_error("transport (no allocation): empty box");
            return VD_EINVAL;
        }
    // the face arrays written and, for a species whose drift velocity is
    // computed per cell, the velocities; each starts on a 256-byte boundary
    const int nflux = transport_rhs_nflux(prm, geom);
    int64_t need = 0, first[3] = {0, 0, 0};
    for (int d = 0; d < nflux; d++) {
        first[d] = need;
        need += (transport_face_count(box, d) + 31) / 32 * 32;
    }
    const bool drift = prm->do_transport && drift_is_state(m->hdev, specid);
    const int64_t vfirst = need;
    if (drift) need += (drift_vel_doubles(sborder, geom->dim) + 31) / 32 * 32;
    const int64_t cap = (int64_t)(VD_NOALLOC_SCRATCH_BYTES / sizeof(double));
    if (need > cap) {
        set_error("transport (no allocation): the box's face fluxes need %lld bytes, beyond the %lld-byte "
                  "no-allocation scratch; use vd_transport_rhs",
                  (long long)need * 8, (long long)cap * 8);
        return VD_EINVAL;
    }
    double* base = (double*)noalloc_scratch(ctx->device);
    if (!base) {
        set_error("transport (no allocation): scratch not set up on device %d", ctx->device);
        return VD_EHIP;
    }
    double* fbuf[3] = {base + first[0], base + first[1], base + first[2]};
    return transport_rhs(ctx, m, sborder, rxn_src, dsdt_out, box, geom, bc_lo, bc_hi, prm, dt, specid, fbuf,
                         drift ? base + vfirst : nullptr);
}
