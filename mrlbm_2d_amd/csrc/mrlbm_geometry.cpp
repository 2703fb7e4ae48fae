// mrlbm_geometry.cpp -- host-side boundary geometry (see mrlbm_geometry.h).
// Compiled without FMA contraction so every value equals the reference's setup
// arithmetic (2d_mrlbm.f90:946-1134, 183-197, 809-811), which is what keeps the FP64
// step bit-identical to the reference.
#include "mrlbm_geometry.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>

namespace mrlbm {

static const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

void label_masks(int label_policy, uint16_t incs[16], uint16_t outs[16])
{
    // Closed rule behind the reference's table (2d_mrlbm.f90:2011-2057): a node takes
    // in the populations arriving from each fluid quadrant (NW -> 1,4,8; NE -> 3,4,7;
    // SW -> 1,2,5; SE -> 2,3,6) plus the rest population, and sends out their opposites.
    static const int quad[4][3] = {{1, 4, 8}, {3, 4, 7}, {1, 2, 5}, {2, 3, 6}};
    static const int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    for (int l = 0; l < 16; l++) {
        uint16_t in = 1, out = 1;
        for (int q = 0; q < 4; q++)
            if (l & (1 << q))
                for (int t = 0; t < 3; t++) {
                    in |= (uint16_t)(1u << quad[q][t]);
                    out |= (uint16_t)(1u << opp[quad[q][t]]);
                }
        const bool set_by_reference = !(l == 0 || l == 6 || l == 9 || l == 15);
        if (!set_by_reference && label_policy == 0) {
            in = 0; // Incs_b/Outs_b stay zero for these labels (2d_mrlbm.f90:2008-2009)
            out = 0;
        }
        incs[l] = in;
        outs[l] = out;
    }
}

std::string build_geometry(const GeometryInput &in, Geometry &g)
{
    const int nx = in.nx, ny = in.ny;
    if (nx < 4 || ny < 4) return "grid must be at least 4x4";
    g.nx = nx;
    g.ny = ny;
    g.x_c = in.x_c;
    g.y_c = in.y_c;
    g.r_c = in.r_c;
    const size_t N = (size_t)nx * ny;
    g.cls.assign(N, 1);
    g.cth.assign(N, 0.0);
    g.sth.assign(N, 0.0);
    auto id = [nx](int i, int j) { return (size_t)i + (size_t)nx * (size_t)j; };
    const double xc = in.x_c, yc = in.y_c, rc = in.r_c;
    const double eps = 0.5; // REAL(4) 0.5 (2d_mrlbm.f90:935)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const double x = (double)(i + 1), y = (double)(j + 1);
            const double radi = std::sqrt(((x - xc) * (x - xc)) + ((y - yc) * (y - yc)));
            g.cth[id(i, j)] = (x - xc) / radi;
            g.sth[id(i, j)] = (y - yc) / radi;
            const double d = radi;
            if (std::fabs(d - rc) <= eps) g.cls[id(i, j)] = 3;
            else if (std::fabs(d) <= rc) g.cls[id(i, j)] = 2;
            else g.cls[id(i, j)] = 1;
        }
    // Non-fluid 8-neighbours of interior fluid nodes become boundary (2d_mrlbm.f90:982-1002).
    std::vector<int8_t> promote(N, 0);
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++) {
            if (g.cls[id(i, j)] != 1) continue;
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++)
                    if (g.cls[id(i + di, j + dj)] == 2) promote[id(i + di, j + dj)] = 1;
        }
    for (size_t n = 0; n < N; n++)
        if (promote[n]) g.cls[n] = 3;

    auto fluid = [&](int i, int j) { return g.cls[id(i, j)] == 1 ? 1 : 0; };
    g.bnodes.clear();
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            if (g.cls[id(i, j)] != 3) continue;
            if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) return "cylinder touches the domain edge";
            BoundaryNode b{};
            b.i = i;
            b.j = j;
            const int f0 = !(fluid(i - 1, j) == 0 && fluid(i - 1, j + 1) == 0 && fluid(i, j + 1) == 0);
            const int f1 = !(fluid(i + 1, j) == 0 && fluid(i + 1, j + 1) == 0 && fluid(i, j + 1) == 0);
            const int f2 = !(fluid(i - 1, j) == 0 && fluid(i - 1, j - 1) == 0 && fluid(i, j - 1) == 0);
            const int f3 = !(fluid(i + 1, j) == 0 && fluid(i + 1, j - 1) == 0 && fluid(i, j - 1) == 0);
            b.label = f0 * 1 + f1 * 2 + f2 * 4 + f3 * 8;
            const double x = (double)(i + 1), y = (double)(j + 1);
            b.radius = std::sqrt(((x - xc) * (x - xc)) + ((y - yc) * (y - yc)));
            g.bnodes.push_back(b);
        }
    if (g.bnodes.empty()) return "no boundary nodes: radius too small";
    // r_cyl: the largest boundary-node radius, over the full list (2d_mrlbm.f90:1038-1050)
    double rmax = g.bnodes[0].radius;
    for (size_t k = 1; k < g.bnodes.size(); k++)
        if (g.bnodes[k].radius > rmax) rmax = g.bnodes[k].radius;
    g.r_cyl = rmax;
    if (in.label_policy == 1) {
        // Extension: nodes whose label the reference leaves unset are bulk fluid.
        std::vector<BoundaryNode> keep;
        for (auto &b : g.bnodes) {
            const int l = b.label;
            if (l == 0 || l == 6 || l == 9 || l == 15) g.cls[id(b.i, b.j)] = 1;
            else keep.push_back(b);
        }
        g.bnodes.swap(keep);
    }
    const double delx = std::sqrt(2.0);
    const double PI = 4.0 * std::atan(1.0);
    const double tiny = (double)1.0e-12f;
    for (auto &b : g.bnodes) { // 2d_mrlbm.f90:1056-1134
        const double x = (double)(b.i + 1), y = (double)(b.j + 1);
        const double rr = std::sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
        double ux = 0, uy = 0, xb = xc, yb = yc;
        if (rr > tiny) {
            ux = (x - xc) / rr;
            uy = (y - yc) / rr;
            xb = xc + (g.r_cyl * ux);
            yb = yc + (g.r_cyl * uy);
        }
        b.xref[0] = xb;
        b.yref[0] = yb;
        b.xref[1] = xb + (delx * ux);
        b.yref[1] = yb + (delx * uy);
        b.xref[2] = xb + ((2.0 * delx) * ux);
        b.yref[2] = yb + ((2.0 * delx) * uy);
        b.xref[3] = xb + ((3.0 * delx) * ux);
        b.yref[3] = yb + ((3.0 * delx) * uy);
        const double dx = x - xb, dy = y - yb;
        b.delta = std::sqrt(dx * dx + dy * dy);
        for (int r = 0; r < 3; r++) {
            const int fx = (int)std::floor(b.xref[r + 1]), fy = (int)std::floor(b.yref[r + 1]);
            const int oi[4] = {0, 1, 0, 1}, oj[4] = {0, 0, 1, 1};
            for (int c = 0; c < 4; c++) {
                b.ci[r][c] = fx + oi[c] - 1;
                b.cj[r][c] = fy + oj[c] - 1;
                if (b.ci[r][c] < 1 || b.ci[r][c] > nx - 2 || b.cj[r][c] < 1 || b.cj[r][c] > ny - 2)
                    return "interpolation stencil leaves the domain interior";
            }
        }
        double th = std::atan2(yb - yc, xb - xc);
        if (th < 0) th = th + 2.0 * PI;
        b.theta = th;
    }
    const int nb = (int)g.bnodes.size();
    g.idx.resize(nb);
    for (int k = 0; k < nb; k++) g.idx[k] = k;
    for (int a = 0; a < nb - 1; a++) // exchange sort, as 2d_mrlbm.f90:1142-1150
        for (int c = a + 1; c < nb; c++)
            if (g.bnodes[g.idx[a]].theta > g.bnodes[g.idx[c]].theta) std::swap(g.idx[a], g.idx[c]);
    g.thetas.resize(nb);
    for (int k = 0; k < nb; k++) g.thetas[k] = g.bnodes[g.idx[k]].theta;
    return "";
}

std::string build_tables(const Geometry &g, int pitch, BoundaryTables &t)
{
    const int nb = (int)g.bnodes.size();
    const int nx = g.nx, ny = g.ny;
    t = BoundaryTables{};
    t.nbct = nb;
    auto node = [pitch](int i, int j) { return (int32_t)(i + pitch * j); };
    auto gid = [nx](int i, int j) { return (size_t)i + (size_t)nx * (size_t)j; };
    std::map<int64_t, int> umap;  // (j, i) -> U index
    std::map<int64_t, int> bmap;  // node -> boundary index
    for (int k = 0; k < nb; k++) bmap[(int64_t)g.bnodes[k].j * nx + g.bnodes[k].i] = k;
    auto uidx = [&](int i, int j) {
        const int64_t key = (int64_t)j * nx + i;
        auto it = umap.find(key);
        if (it != umap.end()) return it->second;
        const int u = (int)umap.size();
        umap[key] = u;
        return u;
    };
    // assign U indices in a first pass, then renumber in node order for locality
    for (int k = 0; k < nb; k++) {
        const auto &b = g.bnodes[k];
        uidx(b.i, b.j);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) uidx(b.ci[r][c], b.cj[r][c]);
    }
    std::vector<int> remap(umap.size());
    {
        int n = 0;
        for (auto &kv : umap) remap[kv.second] = n++; // std::map iterates in (j, i) order
        for (auto &kv : umap) kv.second = remap[kv.second];
    }
    t.nU = (int)umap.size();
    t.unode.resize(t.nU);
    t.uij.resize(2 * t.nU);
    t.urot.resize(3 * t.nU);
    for (auto &kv : umap) {
        const int i = (int)(kv.first % nx), j = (int)(kv.first / nx), u = kv.second;
        t.unode[u] = node(i, j);
        t.uij[2 * u] = i;
        t.uij[2 * u + 1] = j;
        const double c = g.cth[gid(i, j)], s = g.sth[gid(i, j)];
        t.urot[3 * u + 0] = c * c;
        t.urot[3 * u + 1] = s * s;
        t.urot[3 * u + 2] = (2.0 * s) * c; // s2th_glb (2d_mrlbm.f90:964)
    }
    t.node.resize(nb);
    t.ij.resize(2 * nb);
    t.label.resize(nb);
    t.uself.resize(nb);
    t.uc.resize(12 * nb);
    t.dep.resize(8 * nb);
    t.real.assign((size_t)BT_NREAL * nb, 0.0);
    const double delx = std::sqrt(2.0);
    std::vector<int> level(nb, 0);
    for (int k = 0; k < nb; k++) {
        const auto &b = g.bnodes[k];
        t.node[k] = node(b.i, b.j);
        t.ij[2 * k] = b.i;
        t.ij[2 * k + 1] = b.j;
        t.label[k] = b.label;
        t.uself[k] = umap[(int64_t)b.j * nx + b.i];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) {
                t.uc[12 * k + 4 * r + c] = umap[(int64_t)b.cj[r][c] * nx + b.ci[r][c]];
                if (r < 2) {
                    // In-place velocity interpolation (2d_mrlbm.f90:773-815): a corner that
                    // is an earlier boundary node in the list has already been rewritten.
                    auto it = bmap.find((int64_t)b.cj[r][c] * nx + b.ci[r][c]);
                    const int m = (it != bmap.end() && it->second < k) ? it->second : -1;
                    t.dep[8 * k + 4 * r + c] = m;
                    if (m >= 0) level[k] = std::max(level[k], level[m] + 1);
                }
            }
        double *R = &t.real[(size_t)BT_NREAL * k];
        for (int r = 0; r < 3; r++) { // fractional offsets (2d_mrlbm.f90:783-784)
            R[BT_XD2 + 2 * r] = (b.xref[r + 1] - (double)(b.ci[r][0] + 1)) / 1.0;
            R[BT_YD2 + 2 * r] = (b.yref[r + 1] - (double)(b.cj[r][0] + 1)) / 1.0;
        }
        const double dl = b.delta; // 2d_mrlbm.f90:809-811
        R[BT_T1] = ((((2.0 * delx) * delx) - (dl * dl)) + ((3.0 * dl) * delx)) / (2.0 * (delx * delx));
        R[BT_T2] = (dl * (dl - 2.0 * delx)) / (delx * delx);
        R[BT_T3] = (dl * (dl - delx)) / (2.0 * (delx * delx));
        const double c = g.cth[gid(b.i, b.j)], s = g.sth[gid(b.i, b.j)];
        const double s2 = (2.0 * s) * c, c2 = (c * c) - (s * s);
        R[BT_CTH] = c;
        R[BT_STH] = s;
        R[BT_S2TH] = s2;
        R[BT_C2TH] = c2;
        // radii of the extrapolation points relative to the centre (2d_mrlbm.f90:661-669, 750-753)
        auto rad = [&](int r) {
            const double xr = b.xref[r] - g.x_c, yr = b.yref[r] - g.y_c;
            return std::sqrt((xr * xr) + (yr * yr));
        };
        R[BT_R2] = rad(1);
        R[BT_R3] = rad(2);
        R[BT_R4] = rad(3);
        R[BT_RS] = rad(0);
        static const double HXX[9] = {0.0 - 1.0 / 3.0, 1.0 - 1.0 / 3.0, 0.0 - 1.0 / 3.0, 1.0 - 1.0 / 3.0,
                                      0.0 - 1.0 / 3.0, 1.0 - 1.0 / 3.0, 1.0 - 1.0 / 3.0, 1.0 - 1.0 / 3.0,
                                      1.0 - 1.0 / 3.0};
        for (int q = 0; q < 9; q++) { // rotated lattice tables (2d_mrlbm.f90:187-192)
            const double hxx = (double)(EX[q] * EX[q]) - 1.0 / 3.0;
            const double hyy = (double)(EY[q] * EY[q]) - 1.0 / 3.0;
            const double hxy = (double)(EX[q] * EY[q]);
            (void)HXX;
            R[BT_CXP + q] = (double)EX[q] * c + (double)EY[q] * s;
            R[BT_CYP + q] = -((double)EX[q] * s) + (double)EY[q] * c;
            R[BT_HXXP + q] = ((hxx * (c * c)) + (hyy * (s * s))) + (hxy * s2);
            R[BT_HYYP + q] = ((hxx * (s * s)) + (hyy * (c * c))) - (hxy * s2);
            R[BT_HXYP + q] = (((hyy - hxx) * 0.50) * s2) + (hxy * c2);
        }
    }
    int nlev = 0;
    for (int k = 0; k < nb; k++) nlev = std::max(nlev, level[k] + 1);
    t.nlevels = nlev;
    t.level_start.assign(nlev + 1, 0);
    for (int k = 0; k < nb; k++) t.level_start[level[k] + 1]++;
    for (int l = 0; l < nlev; l++) t.level_start[l + 1] += t.level_start[l];
    t.level_order.resize(nb);
    {
        std::vector<int> fill(t.level_start.begin(), t.level_start.end() - 1);
        for (int k = 0; k < nb; k++) t.level_order[fill[level[k]]++] = k;
    }
    t.idx.assign(g.idx.begin(), g.idx.end());
    t.thetas.assign(g.thetas.begin(), g.thetas.end());
    // box of boundary and solid nodes
    int x0 = nx, y0 = ny, x1 = -1, y1 = -1;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const int8_t c = g.cls[gid(i, j)];
            if (c == 2 || c == 3) {
                x0 = std::min(x0, i); y0 = std::min(y0, j);
                x1 = std::max(x1, i); y1 = std::max(y1, j);
            }
        }
    if (x1 < 0) {
        t.bx0 = t.by0 = 0;
        t.bw = t.bh = 0;
        return "";
    }
    if (x0 < 1 || y0 < 1 || x1 > nx - 2 || y1 > ny - 2) return "solid nodes touch the domain edge";
    t.bx0 = x0;
    t.by0 = y0;
    t.bw = x1 - x0 + 1;
    t.bh = y1 - y0 + 1;
    t.boxcls.assign((size_t)t.bw * t.bh, 0);
    for (int j = y0; j <= y1; j++)
        for (int i = x0; i <= x1; i++) {
            const int8_t c = g.cls[gid(i, j)];
            t.boxcls[(size_t)(j - y0) * t.bw + (i - x0)] = c == 3 ? 1 : (c == 2 ? 2 : 0);
        }
    return "";
}

bool build_wall_groups(const BoundaryTables &t, int group, int maxu, WallGroups &w)
{
    w = WallGroups{};
    const int nb = t.nbct;
    if (nb == 0 || t.nlevels > 1) return false;
    w.node.assign(t.idx.begin(), t.idx.end()); // theta order: neighbours share stencil nodes
    w.slot.assign((size_t)13 * nb, 0);
    w.off.push_back(0);
    for (int g0 = 0; g0 < nb; g0 += group) {
        std::vector<int32_t> us; // the group's stencil nodes, sorted
        for (int q = g0; q < std::min(nb, g0 + group); q++) {
            const int k = w.node[q];
            for (int c = 0; c < 12; c++) us.push_back(t.uc[12 * k + c]);
            us.push_back(t.uself[k]);
        }
        std::sort(us.begin(), us.end());
        us.erase(std::unique(us.begin(), us.end()), us.end());
        if ((int)us.size() > maxu) {
            w = WallGroups{};
            return false;
        }
        for (int q = g0; q < std::min(nb, g0 + group); q++) {
            const int k = w.node[q];
            for (int c = 0; c < 13; c++) {
                const int u = c < 12 ? t.uc[12 * k + c] : t.uself[k];
                w.slot[13 * k + c] = (int16_t)(std::lower_bound(us.begin(), us.end(), u) - us.begin());
            }
        }
        w.list.insert(w.list.end(), us.begin(), us.end());
        w.off.push_back((int32_t)w.list.size());
    }
    return true;
}

} // namespace mrlbm
