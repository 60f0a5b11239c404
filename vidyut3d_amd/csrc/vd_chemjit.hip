// vd_chemjit.hip -- mechanism-specialised source kernel (table form).
//
// The reference compiles each case's generated Chemistry.H into the
// executable, so every species index, stoichiometric coefficient and
// Arrhenius constant is a literal (test/models/ionelec_DC_1d/Chemistry.H:
// 240-357).  The table kernel (k_rxn_table) reads those from the device
// descriptor instead, re-reads the reactant densities through the cache
// for every reaction and keeps the NS+1 accumulators in LDS, which caps
// occupancy at 1-2 waves per SIMD and leaves the FP64 pipes idle on load
// latency.  Here the reaction table is printed as straight-line HIP source
// (densities and accumulators in registers, one exp per reaction) and
// compiled for gfx950 with hiprtc once per distinct mechanism in the
// process; each context loads the code object as a module on its device.
//
// Arithmetic order per reaction is that of Source/reaction_source.H:10-64
// and of k_rxn_table: k = A*exp(beta*log T - Ea/T), conc = product over
// reactants of n^nu (n*n for nu = 2) in table order, rate = k*conc,
// reactants -= nu*rate, products += nu*rate, energy += e_r*rate.
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "vd_internal.hpp"

namespace vd {

#define VD_GEN_ARGS_BODY                                                        \
    const double* sb;                                                           \
    double* rx;                                                                 \
    const double* tc;                                                           \
    long long sbn, rxn;                                                         \
    int sbo, rxo, sbj, sbk, rxj, rxk, nx, ny, nz, pad;                          \
    double Tg;                                                                  \
    long long cstride;                                                          \
    unsigned cidxw[4];
#define VD_STRX(...) #__VA_ARGS__
#define VD_STR(...) VD_STRX(__VA_ARGS__)

struct GenArgs {
    VD_GEN_ARGS_BODY
};
static_assert(VD_MAX_BATCH == 16, "GenArgs::cidxw holds one byte per batched case");

// the generated kernel's first lines: its case's arrays (the batch
// offset of blockIdx.z, as case_off gives it)
static const char* GEN_PROLOGUE =
    "  const unsigned z_ = blockIdx.z;\n"
    "  const long long co_ = a0.cstride * (long long)((a0.cidxw[z_ >> 2] >> ((z_ & 3u) * 8u)) & 0xffu);\n"
    "  GenArgs a = a0;\n"
    "  a.sb = (const double*)((const char*)a.sb + co_);\n"
    "  a.rx = (double*)((char*)a.rx + co_);\n"
    "  if (a.tc) a.tc = (const double*)((const char*)a.tc + co_);\n";

constexpr int GEN_BLOCK = 256;

// double literal that round-trips exactly and always parses as a double
static bool lit(std::string& out, double v) {
    if (!std::isfinite(v)) return false;
    char b[64];
    snprintf(b, sizeof b, "%.17g", v);
    out += "(";
    out += b;
    if (!strpbrk(b, ".eE")) out += ".0";
    out += ")";
    return true;
}

static bool chem_source(const DevMech& h, std::string& s) {
    const int ns = h.nspec;
    s.reserve(64 * 1024);
    s += "struct GenArgs { " VD_STR(VD_GEN_ARGS_BODY) " };\n";
    s += "extern \"C\" __global__ void __launch_bounds__(256) vd_chem_spec(GenArgs a0) {\n";
    s += GEN_PROLOGUE;
    // XCD-aware block order (same mapping as xcd_remap)
    s += "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
         "  const unsigned q8 = nb >> 3, r8 = nb & 7u, x8 = b & 7u;\n"
         "  const unsigned lb = x8 * q8 + (x8 < r8 ? x8 : r8) + (b >> 3);\n"
         "  const unsigned t = lb * 256u + threadIdx.x;\n"
         "  const unsigned nx = (unsigned)a.nx, ny = (unsigned)a.ny;\n"
         "  if (t >= nx * ny * (unsigned)a.nz) return;\n"
         "  const unsigned i = t % nx, r0 = t / nx, j = r0 % ny, k = r0 / ny;\n"
         "  const double* __restrict__ s = a.sb + (a.sbo + (int)i + (int)j * a.sbj + (int)k * a.sbk);\n"
         "  double* __restrict__ o = a.rx + (a.rxo + (int)i + (int)j * a.rxj + (int)k * a.rxk);\n"
         "  const long long n = a.sbn;\n";
    s += "  const double Te = s[" + std::to_string(VD_ETEMP(ns)) + " * n];\n";
    s += "  const double lTe = log(Te), iTe = 1.0 / Te, lTg = log(a.Tg), iTg = 1.0 / a.Tg;\n";
    std::vector<int> used(ns, 0);
    for (int r = 0; r < h.nreac; r++)
        for (int q = 0; q < h.nreact[r]; q++) used[h.rsp[q][r]] = 1;
    for (int sp = 0; sp < ns; sp++)
        if (used[sp]) s += "  const double c" + std::to_string(sp) + " = s[" + std::to_string(sp) + " * n];\n";
    for (int sp = 0; sp <= ns; sp++) s += "  double w" + std::to_string(sp) + " = 0.0;\n";
    for (int r = 0; r < h.nreac; r++) {
        const bool te = h.te[r] != 0;
        s += "  { const double k_ = ";
        if (!lit(s, h.A[r])) return false;
        s += " * exp(";
        if (!lit(s, h.beta[r])) return false;
        s += te ? " * lTe - " : " * lTg - ";
        if (!lit(s, h.Ea[r])) return false;
        s += te ? " * iTe);\n" : " * iTg);\n";
        // conc = ((p0 * p1) * p2), p = n, n*n or pow(n, nu)
        std::string conc;
        for (int q = 0; q < h.nreact[r]; q++) {
            const std::string c = "c" + std::to_string(h.rsp[q][r]);
            const double nu = h.rnu[q][r];
            std::string p;
            if (nu == 1.0) p = c;
            else if (nu == 2.0) p = "(" + c + " * " + c + ")";
            else {
                p = "pow(" + c + ", ";
                if (!lit(p, nu)) return false;
                p += ")";
            }
            conc = q == 0 ? p : "(" + conc + " * " + p + ")";
        }
        s += "    const double q_ = k_";
        if (!conc.empty()) s += " * " + conc;
        s += ";\n";
        for (int q = 0; q < h.nreact[r]; q++) {
            s += "    w" + std::to_string(h.rsp[q][r]) + " -= ";
            if (h.rnu[q][r] != 1.0) {
                if (!lit(s, h.rnu[q][r])) return false;
                s += " * ";
            }
            s += "q_;\n";
        }
        for (int q = 0; q < h.nprod[r]; q++) {
            s += "    w" + std::to_string(h.psp[q][r]) + " += ";
            if (h.pnu[q][r] != 1.0) {
                if (!lit(s, h.pnu[q][r])) return false;
                s += " * ";
            }
            s += "q_;\n";
        }
        if (h.energy[r] != 0.0) {
            s += "    w" + std::to_string(ns) + " += ";
            if (!lit(s, h.energy[r])) return false;
            s += " * q_;\n";
        }
        s += "  }\n";
    }
    for (int sp = 0; sp <= ns; sp++)
        s += "  o[" + std::to_string(sp) + " * a.rxn] = w" + std::to_string(sp) + ";\n";
    s += "}\n";
    return true;
}

static std::string num(double v) {
    std::string o;
    lit(o, v);
    return o;
}

// VD_CHEM_GEN: the generated productionRate of a mechanism read as data
// (test/Reactor/CO2_H2_Ar_mechanism_reduced/Chemistry.H:913-6380) printed
// back as straight-line code, operation for operation as k_rxn_gen and the
// oracle evaluate it; the gas-temperature-only k_f and g/RT come from the
// per-launch table a.tc (k_gen_tconst)
static bool chem_source_gen(const DevMech& h, std::string& s) {
    const int ns = h.nspec, E = h.e_idx;
    for (int r = 0; r < h.nreac; r++)
        if (!std::isfinite(h.A[r]) || !std::isfinite(h.A2[r]) || !std::isfinite(h.beta[r]) || !std::isfinite(h.Ea[r]))
            return false;
    s.reserve(256 * 1024);
    s += "struct GenArgs { " VD_STR(VD_GEN_ARGS_BODY) " };\n";
    s += "extern \"C\" __global__ void __launch_bounds__(256) vd_chem_spec(GenArgs a0) {\n";
    s += GEN_PROLOGUE;
    s += "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
         "  const unsigned q8 = nb >> 3, r8 = nb & 7u, x8 = b & 7u;\n"
         "  const unsigned lb = x8 * q8 + (x8 < r8 ? x8 : r8) + (b >> 3);\n"
         "  const unsigned t = lb * 256u + threadIdx.x;\n"
         "  const unsigned nx = (unsigned)a.nx, ny = (unsigned)a.ny;\n"
         "  if (t >= nx * ny * (unsigned)a.nz) return;\n"
         "  const unsigned i = t % nx, r0 = t / nx, j = r0 % ny, k = r0 / ny;\n"
         "  const double* __restrict__ s = a.sb + (a.sbo + (int)i + (int)j * a.sbj + (int)k * a.sbk);\n"
         "  double* __restrict__ o = a.rx + (a.rxo + (int)i + (int)j * a.rxj + (int)k * a.rxk);\n"
         "  const double* __restrict__ tc = a.tc;\n"
         "  const long long n = a.sbn;\n";
    s += "  const double Te = s[" + std::to_string(VD_ETEMP(ns)) + " * n];\n";
    s += "  const double T = a.Tg;\n  const double invT = 1.0 / T;\n"
         "  const double refC = 101325 / 8.31446 * invT;\n  const double refCinv = 1 / refC;\n";
    const std::string NA_UNIV = num(VD_N_A);
    for (int sp = 0; sp < ns; sp++)
        s += "  const double c" + std::to_string(sp) + " = ((s[" + std::to_string(sp) + " * n] * 1.0e-6 / " + NA_UNIV +
             ") * 1.0e6);\n";
    if (h.any_tb) {
        s += "  double mixture = 0.0;\n";
        for (int sp = 0; sp < ns; sp++) s += "  mixture += c" + std::to_string(sp) + ";\n";
    }
    const std::string cE = "c" + std::to_string(E);
    s += "  const double ne_pr = " + cE + " * 6.02214085774e23;\n"
         "  const double Ue = 1.5 * Te * ne_pr * 1.380649e-23;\n"
         "  const double TeeV = Te / 11595.0;\n";
    if (h.any_fit) {
        s += "  const double logTe = log(TeeV);\n  const double invTeeV = (Te == 0) ? 1.0 : 1.0 / (TeeV);\n";
        for (int q = 0; q < 9; q++) s += "  const double tp" + std::to_string(q) + " = pow(logTe, " + std::to_string(q) + ");\n";
        for (int q = 0; q < 4; q++)
            s += "  const double itp" + std::to_string(q) + " = pow(invTeeV, " + std::to_string(q + 1) + ");\n";
    }
    s += "  const double eV2J = 1.602176620800000e-19, NA = 6.02214085774e23, kB = 1.380649e-23;\n"
         "  const double mee = Ue / (" + cE + " * NA);\n"
         "  const double sefact = (tanh((Te - T) / 0.5) + 1.0) / 2.0;\n"
         "  double ener = 0.0;\n";
    for (int sp = 0; sp < ns; sp++) s += "  double w" + std::to_string(sp) + " = 0.0;\n";
    auto fit_sum = [&](const double* c, int nterm, const char* pw) {
        std::string o = "double s_ = 0.0;";
        for (int q = 0; q < nterm; q++) o += " s_ += " + num(c[q]) + " * " + pw + std::to_string(q) + ";";
        return o;
    };
    for (int r = 0; r < h.nreac; r++) {
        s += "  { double k_;\n";
        switch (h.rkind[r]) {
        case VD_RATE_ARRH:
            s += "    k_ = tc[" + std::to_string(r) + "];\n";
            break;
        case VD_RATE_TEPOW:
            s += "    k_ = " + num(h.A[r]) + " * pow(Te / " + num(h.tref[r]) + ", " + num(h.beta[r]) + ");\n";
            if (h.tgdiv[r]) s += "    k_ = k_ / T;\n";
            break;
        case VD_RATE_TEARRH:
            s += "    k_ = " + num(h.A[r]) + " * pow(Te, " + num(h.beta[r]) + ") * exp(" + num(-h.Ea[r]) + " / Te);\n";
            break;
        case VD_RATE_JANEV:
        case VD_RATE_FFIT: {
            double c[9];
            for (int q = 0; q < 9; q++) c[q] = h.fit[q][r];
            const bool j = h.rkind[r] == VD_RATE_JANEV;
            s += "    { " + fit_sum(c, j ? 9 : 4, j ? "tp" : "itp") + " k_ = " + num(h.A[r]) +
                 " * exp(s_) * 6.02214085774e23; }\n";
            break;
        }
        default: {
            double c[9], c2[4];
            for (int q = 0; q < 9; q++) c[q] = h.fit[q][r];
            for (int q = 0; q < 4; q++) c2[q] = h.fit2[q][r];
            s += "    if (TeeV < " + num(h.fsw[r]) + ") { " + fit_sum(c, 9, "tp") + " k_ = " + num(h.A[r]) +
                 " * exp(s_) * 6.02214085774e23; }\n    else { " + fit_sum(c2, 4, "itp") + " k_ = " + num(h.A2[r]) +
                 " * exp(s_) * 6.02214085774e23; }\n";
            break;
        }
        }
        if (h.kmax[r] > 0.0) s += "    k_ = (" + num(h.kmax[r]) + " < k_) ? " + num(h.kmax[r]) + " : k_;\n";
        s += std::string("    const double pre_ = ") + (h.tb[r] ? "mixture * k_;\n" : "k_;\n");
        auto prod = [&](const int* sp, const double* nu, int cnt, int stride_r) -> std::string {
            std::string p;
            for (int q = 0; q < cnt; q++)
                for (int e = 0; e < (int)nu[q * stride_r]; e++) {
                    const std::string c = "c" + std::to_string(sp[q * stride_r]);
                    p = p.empty() ? c : "(" + p + " * " + c + ")";
                }
            return p.empty() ? "1.0" : p;
        };
        s += "    const double qf_ = pre_ * " + prod(&h.rsp[0][r], &h.rnu[0][r], h.nreact[r], VD_MAX_REACTIONS) + ";\n";
        if (h.rev[r]) {
            std::string gs;
            for (int q = 0; q < h.ng[r]; q++) {
                const std::string g = "tc[" + std::to_string(VD_MAX_REACTIONS + h.gsp[q][r]) + "]";
                const double nu = h.gnu[q][r];
                const std::string t = nu == 1.0 ? g : (nu == -1.0 ? "(-" + g + ")" : "(" + num(nu) + " * " + g + ")");
                gs = q == 0 ? t : "(" + gs + " + " + t + ")";
            }
            std::string v = "(pre_ * exp(-(" + gs + ")))";
            for (int p = 0; p < h.refcp[r]; p++) v = "(" + v + " * refC)";
            for (int p = 0; p < -h.refcp[r]; p++) v = "(" + v + " * refCinv)";
            s += "    const double qr_ = " + v + " * " + prod(&h.revsp[0][r], &h.revnu[0][r], h.nrev[r], VD_MAX_REACTIONS) +
                 ";\n";
        } else {
            s += "    const double qr_ = 0.0;\n";
        }
        s += "    const double qd_ = qf_ - qr_;\n";
        for (int q = 0; q < h.nwd[r]; q++) {
            const double nu = h.wdnu[q][r];
            const std::string w = "    w" + std::to_string(h.wdsp[q][r]);
            if (nu == 1.0) s += w + " += qd_;\n";
            else if (nu == -1.0) s += w + " -= qd_;\n";
            else if (nu > 0.0) s += w + " += " + num(nu) + " * qd_;\n";
            else s += w + " -= " + num(-nu) + " * qd_;\n";
        }
        const int rt = h.rxntype[r];
        if (rt == 1 || rt == 2) s += "    ener -= (" + num(h.eexci[r]) + " * eV2J) * NA * (qf_ - qr_) * sefact;\n";
        if (rt == 3 || rt == 4) s += "    ener -= mee * NA * (qf_ - qr_);\n";
        if (rt == 5) {
            const int el = h.elidx[r];
            s += "    { const double qq = " + cE + " * c" + std::to_string(el) + " * k_;\n"
                 "      ener -= 3.0 * kB * (" + num(h.molwt[E]) + " / " + num(h.molwt[el]) + ") * (Te - T) * qq * NA; }\n";
        }
        if (rt == 7) s += "    ener += mee * NA * (qf_ - qr_);\n";
        s += "  }\n";
    }
    for (int sp = 0; sp < ns; sp++)
        s += "  o[" + std::to_string(sp) + " * a.rxn] = ((w" + std::to_string(sp) + " * 1.0e-6) * " + NA_UNIV + ") * 1.0e6;\n";
    s += "  o[" + std::to_string(ns) + " * a.rxn] = ener;\n}\n";
    return true;
}

// code objects by generated source, shared by every context in the process
static std::mutex g_jit_mu;
static std::map<std::string, std::vector<char>> g_jit_cache;

static int jit_compile(const std::string& src, std::vector<char>& code) {
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        auto it = g_jit_cache.find(src);
        if (it != g_jit_cache.end()) {
            code = it->second;
            return VD_OK;
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "vd_chem_spec.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        set_error("hiprtcCreateProgram failed");
        return VD_EHIP;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult rc = hiprtcCompileProgram(prog, 6, opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string lg(ls + 1, '\0');
        hiprtcGetProgramLog(prog, &lg[0]);
        set_error("hiprtc compile of the specialised chemistry kernel: %s: %.600s",
                  hiprtcGetErrorString(rc), lg.c_str());
        hiprtcDestroyProgram(&prog);
        return VD_EHIP;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_jit_mu);
    g_jit_cache[src] = code;
    return VD_OK;
}

int chem_jit_load(vd_mech* m) {
    if (m->fn) return VD_OK;
    if (m->hdev.chem_form != VD_CHEM_TABLE && m->hdev.chem_form != VD_CHEM_GEN) {
        set_error("the specialised source kernel exists for table and generated mechanisms only");
        return VD_EINVAL;
    }
    std::string src;
    if (!(m->hdev.chem_form == VD_CHEM_GEN ? chem_source_gen(m->hdev, src) : chem_source(m->hdev, src))) {
        set_error("mechanism has a non-finite constant; no specialised kernel");
        return VD_EINVAL;
    }
    std::vector<char> code;
    VD_TRY(jit_compile(src, code));
    VD_HIP(hipSetDevice(m->ctx->device));
    hipModule_t mod;
    VD_HIP(hipModuleLoadData(&mod, code.data()));
    hipFunction_t fn;
    hipError_t e = hipModuleGetFunction(&fn, mod, "vd_chem_spec");
    if (e != hipSuccess) {
        (void)hipModuleUnload(mod);
        set_error("hipModuleGetFunction(vd_chem_spec): %s", hipGetErrorString(e));
        return VD_EHIP;
    }
    m->mod = (void*)mod;
    m->fn = (void*)fn;
    return VD_OK;
}

void chem_jit_unload(vd_mech* m) {
    if (m->mod) (void)hipModuleUnload((hipModule_t)m->mod);
    m->mod = nullptr;
    m->fn = nullptr;
}

int chem_jit_launch(vd_ctx* ctx, const vd_mech* m, const vd_array4* sb, vd_array4* rxn, const vd_box* box,
                    double Tg, const double* tc) {
    GenArgs a;
    a.sb = sb->p;
    a.rx = rxn->p;
    a.tc = tc;
    a.sbn = sb->nstride;
    a.rxn = rxn->nstride;
    a.sbj = (int)sb->jstride;
    a.sbk = (int)sb->kstride;
    a.rxj = (int)rxn->jstride;
    a.rxk = (int)rxn->kstride;
    a.sbo = (box->lo[0] - sb->begin[0]) + (box->lo[1] - sb->begin[1]) * a.sbj + (box->lo[2] - sb->begin[2]) * a.sbk;
    a.rxo = (box->lo[0] - rxn->begin[0]) + (box->lo[1] - rxn->begin[1]) * a.rxj +
            (box->lo[2] - rxn->begin[2]) * a.rxk;
    a.nx = box->hi[0] - box->lo[0] + 1;
    a.ny = box->hi[1] - box->lo[1] + 1;
    a.nz = box->hi[2] - box->lo[2] + 1;
    a.pad = 0;
    a.Tg = Tg;
    a.cstride = ctx->bt.stride;
    memcpy(a.cidxw, ctx->bt.idx, sizeof(a.cidxw));
    const int64_t tot = (int64_t)a.nx * a.ny * a.nz;
    if (tot <= 0) return VD_OK;
    if (tot > 0xffffffffLL - GEN_BLOCK) {
        set_error("box of %lld cells is too large for one source launch", (long long)tot);
        return VD_EINVAL;
    }
    void* params[] = {&a};
    VD_HIP(hipModuleLaunchKernel((hipFunction_t)m->fn, (unsigned)((tot + GEN_BLOCK - 1) / GEN_BLOCK), 1,
                                 (unsigned)ctx->bn, GEN_BLOCK, 1, 1, 0, ctx->stream, params, nullptr));
    return VD_OK;
}

}  // namespace vd
