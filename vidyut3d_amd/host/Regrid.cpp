// Regrid.cpp -- new grids for the level above a single-box level from its
// refinement tags: AmrMesh::MakeNewGrids as Vidyut's regrid(0, time) runs
// it every regrid_int steps (Source/Evolve.cpp:66-72) on the tags of
// Vidyut::ErrorEst (Source/Vidyut.cpp:171-255).
//
// Steps (AMReX's, restated): buffer the tags by n_error_buf cells, coarsen
// them by blocking_factor / ref_ratio so that the boxes come out aligned to
// the blocking factor, cluster them (Berger & Rigoutsos, IEEE Trans. SMC
// 21(5), 1991: accept a box once its tagged fraction reaches grid_eff,
// else cut it at a hole of the tag signature, else at the strongest
// inflection of the signature's Laplacian, else in the middle of its
// longest side), refine the boxes to the new level and chop them to
// max_grid_size.  Host code only.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/vidyut3d_amd_host.h"

namespace vd {
void set_error(const char* fmt, ...);
}

namespace {

using IV = std::array<int, 3>;

struct Bx {
    IV lo, hi;
    long long vol() const { return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1); }
    int len(int d) const { return hi[d] - lo[d] + 1; }
};

Bx bounding(const std::vector<IV>& pts) {
    Bx b;
    b.lo = pts[0];
    b.hi = pts[0];
    for (const IV& p : pts)
        for (int d = 0; d < 3; d++) {
            b.lo[d] = std::min(b.lo[d], p[d]);
            b.hi[d] = std::max(b.hi[d], p[d]);
        }
    return b;
}

// Where to cut a cluster box: direction and the first index of the upper
// part.  Priority: a hole (zero of the signature) nearest the middle of its
// side, then the strongest sign change of the signature's Laplacian, then
// the middle of the longest side.
struct Cut {
    int dir = -1, at = 0;
};

Cut find_cut(const Bx& b, const std::vector<IV>& pts, int dim) {
    std::vector<long long> sig[3];
    for (int d = 0; d < dim; d++) sig[d].assign((size_t)b.len(d), 0);
    for (const IV& p : pts)
        for (int d = 0; d < dim; d++) sig[d][(size_t)(p[d] - b.lo[d])]++;
    // 1. holes: the box is the points' bounding box, so holes are interior
    Cut best;
    double bestoff = 1e300;
    int bestlen = -1;
    for (int d = 0; d < dim; d++) {
        const int n = b.len(d);
        const double mid = 0.5 * (n - 1);
        for (int i = 1; i < n - 1; i++)
            if (sig[d][(size_t)i] == 0) {
                const double off = std::fabs(i - mid) / n;
                if (off < bestoff || (off == bestoff && n > bestlen)) {
                    bestoff = off;
                    bestlen = n;
                    best.dir = d;
                    best.at = b.lo[d] + i;
                }
            }
    }
    if (best.dir >= 0) return best;
    // 2. inflections: sign changes of lap[i] = s[i-1] - 2 s[i] + s[i+1]
    long long beststr = 0;
    bestoff = 1e300;
    for (int d = 0; d < dim; d++) {
        const int n = b.len(d);
        if (n < 4) continue;
        const double mid = 0.5 * n;
        std::vector<long long> lap((size_t)n, 0);
        for (int i = 1; i < n - 1; i++) lap[(size_t)i] = sig[d][(size_t)i - 1] - 2 * sig[d][(size_t)i] + sig[d][(size_t)i + 1];
        for (int i = 1; i < n - 2; i++) {
            const long long a = lap[(size_t)i], c = lap[(size_t)i + 1];
            if ((a < 0 && c > 0) || (a > 0 && c < 0)) {
                const long long str = std::llabs(c - a);
                const double off = std::fabs((i + 1) - mid) / n;
                if (str > beststr || (str == beststr && off < bestoff)) {
                    beststr = str;
                    bestoff = off;
                    best.dir = d;
                    best.at = b.lo[d] + i + 1;  // between i and i+1
                }
            }
        }
    }
    if (best.dir >= 0) return best;
    // 3. the middle of the longest side
    int dl = 0;
    for (int d = 1; d < dim; d++)
        if (b.len(d) > b.len(dl)) dl = d;
    if (b.len(dl) < 2) return best;  // a single cell: nothing to cut
    best.dir = dl;
    best.at = b.lo[dl] + b.len(dl) / 2;
    return best;
}

void cluster(std::vector<IV> pts, int dim, double eff, std::vector<Bx>& out) {
    // explicit stack: (points) -> accept or split
    std::vector<std::vector<IV>> work;
    work.push_back(std::move(pts));
    while (!work.empty()) {
        std::vector<IV> p = std::move(work.back());
        work.pop_back();
        if (p.empty()) continue;
        const Bx b = bounding(p);
        if ((double)p.size() >= eff * (double)b.vol()) {
            out.push_back(b);
            continue;
        }
        const Cut c = find_cut(b, p, dim);
        if (c.dir < 0) {
            out.push_back(b);
            continue;
        }
        std::vector<IV> lo, hi;
        for (const IV& q : p) (q[c.dir] < c.at ? lo : hi).push_back(q);
        // upper part first on the stack so the lower one is emitted first
        work.push_back(std::move(hi));
        work.push_back(std::move(lo));
    }
}

// pieces of [lo, hi] (a multiple of `unit` long, starting on a multiple)
// of at most `mgs` cells each, as equal as the unit allows
void chop_side(int lo, int hi, int unit, int mgs, std::vector<std::pair<int, int>>& out) {
    const int len = hi - lo + 1;
    if (len <= mgs) {
        out.push_back({lo, hi});
        return;
    }
    const int units = len / unit;
    const int nblk = (len + mgs - 1) / mgs;
    const int base = units / nblk, extra = units % nblk;
    int s = lo;
    for (int k = 0; k < nblk; k++) {
        const int u = base + (k < extra ? 1 : 0);
        out.push_back({s, s + u * unit - 1});
        s += u * unit;
    }
}

}  // namespace

extern "C" int vd_make_new_grids(const char* tags, const vd_box* domain, const int is_periodic[3], int dim,
                                 const vd_regrid_params* prm, vd_box* out, int max_out, int* nout) {
    if (!tags || !domain || !is_periodic || !prm || !nout || (max_out > 0 && !out)) return VD_EINVAL;
    if (dim != 2 && dim != 3) {
        vd::set_error("vd_make_new_grids: dim %d", dim);
        return VD_EINVAL;
    }
    IV n{1, 1, 1}, cf{1, 1, 1};
    for (int d = 0; d < dim; d++) {
        const int r = prm->ref_ratio[d], bf = prm->blocking_factor[d], mgs = prm->max_grid_size[d];
        n[d] = domain->hi[d] - domain->lo[d] + 1;
        if ((r != 2 && r != 4) || bf < r || bf % r != 0 || mgs < bf || mgs % bf != 0 || n[d] <= 0) {
            vd::set_error("vd_make_new_grids: ref_ratio %d, blocking_factor %d, max_grid_size %d in direction %d",
                          r, bf, mgs, d);
            return VD_EINVAL;
        }
        cf[d] = bf / r;
        if (n[d] % cf[d] != 0) {
            vd::set_error("vd_make_new_grids: domain length %d not a multiple of blocking_factor/ref_ratio %d", n[d],
                          cf[d]);
            return VD_EINVAL;
        }
    }
    if (prm->n_error_buf < 0 || !(prm->grid_eff > 0.0 && prm->grid_eff <= 1.0)) {
        vd::set_error("vd_make_new_grids: n_error_buf %d, grid_eff %g", prm->n_error_buf, prm->grid_eff);
        return VD_EINVAL;
    }
    const long long ncell = (long long)n[0] * n[1] * n[2];
    // 1. buffered tags
    std::vector<char> buf((size_t)ncell, 0);
    const int nb = prm->n_error_buf;
    const int nbz = dim == 3 ? nb : 0;
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                if (!tags[(size_t)(i + (long long)n[0] * (j + (long long)n[1] * k))]) continue;
                for (int dk = -nbz; dk <= nbz; dk++)
                    for (int dj = -nb; dj <= nb; dj++)
                        for (int di = -nb; di <= nb; di++) {
                            IV c{i + di, j + dj, k + dk};
                            bool keep = true;
                            for (int d = 0; d < dim; d++) {
                                if (c[d] >= 0 && c[d] < n[d]) continue;
                                if (is_periodic[d])
                                    c[d] = ((c[d] % n[d]) + n[d]) % n[d];
                                else
                                    keep = false;
                            }
                            if (keep) buf[(size_t)(c[0] + (long long)n[0] * (c[1] + (long long)n[1] * c[2]))] = 1;
                        }
            }
    // 2. coarsened by cf (a coarse cell is tagged when any child is)
    const IV m{n[0] / cf[0], n[1] / cf[1], n[2] / cf[2]};
    std::vector<char> ct((size_t)m[0] * m[1] * m[2], 0);
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++)
                if (buf[(size_t)(i + (long long)n[0] * (j + (long long)n[1] * k))])
                    ct[(size_t)(i / cf[0] + (long long)m[0] * (j / cf[1] + (long long)m[1] * (k / cf[2])))] = 1;
    std::vector<IV> pts;
    for (int k = 0; k < m[2]; k++)
        for (int j = 0; j < m[1]; j++)
            for (int i = 0; i < m[0]; i++)
                if (ct[(size_t)(i + (long long)m[0] * (j + (long long)m[1] * k))]) pts.push_back({i, j, k});
    // 3. clusters in the coarsened space
    std::vector<Bx> cl;
    if (!pts.empty()) cluster(std::move(pts), dim, prm->grid_eff, cl);
    // 4. refine to the new level (domain offset kept) and chop
    std::vector<vd_box> res;
    for (const Bx& b : cl) {
        std::vector<std::pair<int, int>> side[3];
        for (int d = 0; d < 3; d++) {
            if (d >= dim) {
                side[d].push_back({0, 0});
                continue;
            }
            const int r = prm->ref_ratio[d];
            const int flo = (domain->lo[d] + b.lo[d] * cf[d]) * r;
            const int fhi = (domain->lo[d] + (b.hi[d] + 1) * cf[d]) * r - 1;
            chop_side(flo, fhi, prm->blocking_factor[d], prm->max_grid_size[d], side[d]);
        }
        for (auto& z : side[2])
            for (auto& y : side[1])
                for (auto& x : side[0]) {
                    vd_box v;
                    v.lo[0] = x.first;
                    v.hi[0] = x.second;
                    v.lo[1] = y.first;
                    v.hi[1] = y.second;
                    v.lo[2] = z.first;
                    v.hi[2] = z.second;
                    res.push_back(v);
                }
    }
    *nout = (int)res.size();
    if ((int)res.size() > max_out) {
        vd::set_error("vd_make_new_grids: %d boxes, room for %d", (int)res.size(), max_out);
        return VD_EINVAL;
    }
    std::copy(res.begin(), res.end(), out);
    return VD_OK;
}
