// mrlbm_adjoint.hip -- discrete adjoint of the step for the time-averaged drag, with respect
// to a per-node Brinkman solid fraction chi and the inlet velocity uo (at fixed omega).
//
// Forward: the ordinary step (bulk nodes additionally scale u by 1 - chi before the
// collision), moment-field checkpoints every K steps.  Reverse, segment by segment: the K
// states of the segment are recomputed into a buffer, then each step is transposed.
//
// The adjoint state lam is the gradient with respect to the conserved form of the stored
// moments, (rho, j = rho u, P = rho m).  In those coordinates the reconstruction
// f_k = w_k (rho + 3 c_k.j + 4.5 Pxx Hxx + 4.5 Pyy Hyy + 9 Pxy Hxy) is linear with constant
// coefficients, so one reverse step of a bulk node y is a gather over its 9 downstream
// neighbours x = y + c_k of the population adjoints a_k(x) = dJ/df_k(x):
//     lam_t(y) = sum_k w_k a_k(y + c_k) (1, 3 c_k, 4.5 Hxx_k, 4.5 Hyy_k, 9 Hxy_k),
// and a_k(x) of a bulk node follows in closed form from lam_{t+1}(x), rho(x), u(x) and chi(x)
// (collision and Brinkman transposed by hand).  Curved-wall and edge nodes get a_k from
// their local Jacobian, evaluated by running the forward BC code on dual numbers (one
// seeded input per thread); the linear velocity and moment interpolations of the wall BC
// and the outlet's copy of column nx-1 are transposed through precomputed reverse lists.
// Special rows E (9 population adjoints per node) carry every non-bulk contribution.
#include "mrlbm_device.h"
#include "mrlbm_dual.h"
#include "mrlbm_handle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace mrlbm {

struct AdjointWork {
    bool f32 = false;
    int K = 0, nseg_cap = 0;
    void *ckpt = nullptr;      // nseg_cap states of 6 planes
    void *seg[2] = {nullptr, nullptr}; // K + 1 states each; seg[1] only when memory allows
    void *U2 = nullptr, *ub2 = nullptr; // wall scratch of the recompute stream
    hipEvent_t evR = nullptr, evV = nullptr; // recompute done / reverse done
    void *lam[2] = {nullptr, nullptr};
    double *G = nullptr;       // dJ/dchi, one plane
    double *guin = nullptr;    // dJ/d(uin_j), ny
    int32_t *spec = nullptr;   // row of each node in E, -1 if none (one plane)
    int ns = 0;
    bool flat = false;         // one dependency level and no stencil node in column nx-2
    void *E = nullptr;         // ns * 9
    void *E2 = nullptr;        // the same for the other step parity (a step's prep runs beside the previous bulk)
    int nedge = 0;
    int32_t *edge_ij = nullptr;
    void *Ufull = nullptr;     // 6 * nU pre-collision moments of the stencil nodes
    void *outm = nullptr;      // 3 * ny pre-collision rho, ux, uy of column nx-2
    void *lub_own = nullptr, *lub = nullptr, *lMi = nullptr; // 2 * nbct each
    void *lout = nullptr;      // 3 * ny
    int32_t *rev_dep_start = nullptr, *rev_dep = nullptr;
    int32_t *rev_u_start = nullptr, *rev_u = nullptr;
    std::vector<void *> allocs;
};

template <class S, class T> struct AdjArgs {
    const S *m1;         // state after the step being transposed
    const T *lam1;       // lam of that state (6 planes)
    T *lam0;             // lam of the state before the step
    double *G, *guin;
    const int32_t *spec;
    T *E;
    int ns, nedge;
    const int32_t *edge_ij;
    T *Ufull, *outm, *lub_own, *lub, *lMi, *lout;
    const int32_t *rev_dep_start, *rev_dep, *rev_u_start, *rev_u;
    T wJ;                // objective weight of this step's drag (0 outside the window)
    int stencil;         // adj_prep_kernel also fills the forward's stencil table (b.U)
};

// Conserved-coordinate weights v_k of the reconstruction: d f_k / d(rho, jx, jy, Pxx, Pyy, Pxy) / w_k.
template <class T> __device__ __forceinline__ void recon_row(int k, T v[6])
{
    v[0] = (T)1.0;
    v[1] = (T)3.0 * (T)EXk(k);
    v[2] = (T)3.0 * (T)EYk(k);
    v[3] = (T)4.5 * Hxxk<T>(k);
    v[4] = (T)4.5 * Hyyk<T>(k);
    v[5] = (T)9.0 * Hxyk<T>(k);
}

template <class S, class T> __device__ __forceinline__ bool on_ring(const StepArgs<S, T> &a, int i, int j)
{
    return i == 0 || i == a.nx - 1 || j == 0 || j == a.ny - 1;
}

template <class S, class T> __device__ __forceinline__ bool is_wall(const StepArgs<S, T> &a, int i, int j)
{
    const int bi = i - a.bx0, bj = j - a.by0;
    return (unsigned)bi < (unsigned)a.bw && (unsigned)bj < (unsigned)a.bh && (a.boxcls[bj * a.bw + bi] & 1);
}

// Transposed Brinkman + collision of a bulk node: from lam (conserved, post-collision) to the
// adjoint of the pre-collision conserved moments, and d/dchi.  Forward, with
// u' = (1 - chi) j / rho:  rho+ = rho, j+ = (1 - chi) j, P+ = (1 - om) P + om rho u' u'.
template <class T>
__device__ __forceinline__ void bulk_pre(const T l[6], T rho, T ux, T uy, T chi, T om, T lp[6], T &lchi)
{
    const T keep = (T)1.0 - chi;
    lp[0] = l[0] - om * ((l[3] * ux) * ux + (l[4] * uy) * uy + (l[5] * ux) * uy);
    lp[1] = keep * (l[1] + om * ((T)2.0 * l[3] * ux + l[5] * uy));
    lp[2] = keep * (l[2] + om * ((T)2.0 * l[4] * uy + l[5] * ux));
    lp[3] = ((T)1.0 - om) * l[3];
    lp[4] = ((T)1.0 - om) * l[4];
    lp[5] = ((T)1.0 - om) * l[5];
    lchi = -(rho / keep) * ((l[1] * ux + l[2] * uy) + (T)2.0 * om * ((l[3] * ux) * ux + (l[4] * uy) * uy + (l[5] * ux) * uy));
}

// a_k from the pre-collision adjoint: d(rho, j, P)/d f_k = (1, c_k, H_k).
template <class T> __device__ __forceinline__ T pop_adjoint(const T lp[6], int k)
{
    return ((((lp[0] + lp[1] * (T)EXk(k)) + lp[2] * (T)EYk(k)) + lp[3] * Hxxk<T>(k)) + lp[4] * Hyyk<T>(k))
           + lp[5] * Hxyk<T>(k);
}

template <class S, class T>
__device__ __forceinline__ void load_lam(const StepArgs<S, T> &a, const T *lam, long long n, T l[6])
{
#pragma unroll
    for (int p = 0; p < 6; p++) l[p] = lam[p * a.plane + n];
}

// ---------------------------------------------------------------------------------
// Pre-collision moments of the stencil nodes and of column nx-2 at the step's start
// state; clears the special rows.
template <class S, class T> __device__ void adj_prep_body(const StepArgs<S, T> &a, const AdjArgs<S, T> &d, int t)
{
    using F = typename FormOf<T>::F;
    if (t < d.ns * Q) d.E[t] = (T)0.0;
    const BoundaryDev &b = a.b;
    if (t < b.nU) {
        T f[Q], m[6];
        gather_interior(a, b.uij[2 * t], b.uij[2 * t + 1], f);
        F::moments(f, m);
        if (d.stencil) store_stencil_moments(a, t, m); // what stencil_moments_kernel writes
        m[0] = F::rho_of(m[0]);
#pragma unroll
        for (int p = 0; p < 6; p++) d.Ufull[p * b.nU + t] = m[p];
    }
    if (t < a.ny) {
        T f[Q], m[6];
        gather_edge(a, a.nx - 2, t, f);
        F::moments(f, m);
        d.outm[3 * t] = F::rho_of(m[0]);
        d.outm[3 * t + 1] = m[1];
        d.outm[3 * t + 2] = m[2];
    }
}

template <class S, class T> __global__ void __launch_bounds__(256) adj_prep_kernel(StepArgs<S, T> a, AdjArgs<S, T> d)
{
    adj_prep_body(a, d, blockIdx.x * blockDim.x + threadIdx.x);
}

template <class S, class T> __global__ void __launch_bounds__(256) adj_wallvel_kernel(StepArgs<S, T> a)
{
    wall_velocities(a);
}

// Post-collision conserved moments of a node from its (rho, u, m) values.
template <class D> __device__ __forceinline__ void conserved(const D m[6], D o[6])
{
    o[0] = m[0];
    o[1] = m[0] * m[1];
    o[2] = m[0] * m[2];
    o[3] = m[0] * m[3];
    o[4] = m[0] * m[4];
    o[5] = m[0] * m[5];
}

// Local Jacobians of the curved-wall and edge nodes, contracted with lam: one thread per
// (node, input).  Wall inputs: 9 streamed populations, interpolated u (2), interpolated
// rotated moments (2).  Edge inputs: 9 populations, the outlet's copied rho, ux, uy of
// column nx-2 (3), the inlet velocity (1).
template <class S, class T> __global__ void __launch_bounds__(256) adj_local_kernel(StepArgs<S, T> a, AdjArgs<S, T> d)
{
    using F = typename FormOf<T>::F;
    using D = Dual<T>;
    const BoundaryDev &b = a.b;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int node = t / 13, q = t % 13;
    if (node >= b.nbct + d.nedge) return;
    const T om = a.om, om1 = a.om1;
    if (node < b.nbct) {
        const int k = node;
        const int i = b.ij[2 * k], j = b.ij[2 * k + 1];
        const long long n = (long long)j * a.pitch + i;
        const double *R = b.real + (long long)k * BT_NREAL;
        // loads that do not depend on the arithmetic first: lam of the node, its special row,
        // the interpolated wall moments' inputs, then the gather
        T l[6];
        load_lam(a, d.lam1, n, l);
        const int srow = d.spec[n];
        T Mxx0 = (T)0.0, Myy0 = (T)0.0;
        if (b.rotated) wall_moments(b, (const T *)b.U, k, Mxx0, Myy0);
        T ux0, uy0; // the wall velocity: interpolated here with one dependency level, else by adj_wallvel
        if (b.nlevels <= 1) {
            wall_velocity_one(b, (const T *)b.U, k, ux0, uy0);
        } else {
            ux0 = ((const T *)b.ub)[2 * k];
            uy0 = ((const T *)b.ub)[2 * k + 1];
        }
        T fp[Q];
        gather_interior(a, i, j, fp);
        D ff[Q];
#pragma unroll
        for (int p = 0; p < Q; p++) ff[p] = D(F::dev ? fp[p] + F::base(p) : fp[p], p == q ? (T)1.0 : (T)0.0);
        const D ux(ux0, q == 9 ? (T)1.0 : (T)0.0), uy(uy0, q == 10 ? (T)1.0 : (T)0.0);
        const int lab = b.label[k];
        const uint16_t in = b.incs[lab], out = b.outs[lab];
        D m[6];
        if (b.rotated) {
            const D Mxx(Mxx0, q == 11 ? (T)1.0 : (T)0.0), Myy(Myy0, q == 12 ? (T)1.0 : (T)0.0);
            const D c = (D)R[BT_CTH], s = (D)R[BT_STH];
            const D uxp = ux * c + uy * s, uyp = uy * c - ux * s;
            wall_bc_rotated<D>(ff, uxp, uyp, R, in, out, b.mass_bc, D(om), Mxx, Myy, m);
        } else {
            if (q >= 11) return;
            wall_bc_unrotated<D>(ff, ux, uy, in, out, D(om), m);
        }
        m[1] = ux;
        m[2] = uy;
        collide<D>(m, D(om), D(om1), a.incomp_bulk);
        D o[6];
        conserved(m, o);
        if (d.wJ != (T)0.0) { // drag of this step: - sum_out fout_q c_qx, fout = recon(post state)
            for (int qq = 0; qq < Q; qq++)
                if (((out >> qq) & 1) && EXk(qq) != 0) {
                    T v[6];
                    recon_row<T>(qq, v);
                    const T c = d.wJ * wk<T>(qq) * (T)EXk(qq);
#pragma unroll
                    for (int p = 0; p < 6; p++) l[p] = l[p] - c * v[p];
                }
        }
        T g = (T)0.0;
#pragma unroll
        for (int p = 0; p < 6; p++) g = g + l[p] * o[p].d;
        if (q < Q) {
            if (d.wJ != (T)0.0 && ((in >> q) & 1)) g = g + d.wJ * (T)EXk(q); // + sum_in f_q c_qx
            d.E[srow * Q + q] = g;
        } else if (q < 11) {
            d.lub_own[2 * k + (q - 9)] = g;
        } else {
            d.lMi[2 * k + (q - 11)] = g;
        }
        return;
    }
    // edge node, as edge_update + collide
    const int e = node - b.nbct;
    const int i = d.edge_ij[2 * e], j = d.edge_ij[2 * e + 1];
    const bool out = i == a.nx - 1, in = i == 0, top = j == a.ny - 1, bot = j == 0;
    if (q >= 9 && q <= 11 && !out) return;
    if (q == 12 && !in) return;
    const long long n = (long long)j * a.pitch + i;
    T l[6];
    load_lam(a, d.lam1, n, l);
    const int srow = d.spec[n];
    T fp[Q];
    gather_edge(a, i, j, fp);
    D ff[Q];
#pragma unroll
    for (int p = 0; p < Q; p++) ff[p] = D(F::dev ? fp[p] + F::base(p) : fp[p], p == q ? (T)1.0 : (T)0.0);
    D mm[6];
    if (out) {
        mm[0] = D(d.outm[3 * j], q == 9 ? (T)1.0 : (T)0.0);
        mm[1] = D(d.outm[3 * j + 1], q == 10 ? (T)1.0 : (T)0.0);
        mm[2] = D(d.outm[3 * j + 2], q == 11 ? (T)1.0 : (T)0.0);
    }
    if (in) {
        mm[1] = D(a.uin[j], q == 12 ? (T)1.0 : (T)0.0);
        mm[2] = D(a.vin);
    }
    if (top || bot) {
        mm[1] = D((T)0.0);
        mm[2] = D((T)0.0);
    }
    const int side = bot ? 3 : (top ? 2 : (in ? 1 : 0));
    edge_bc<D>(side, a.incomp_bc, ff, mm, D(om));
    collide<D>(mm, D(om), D(om1), a.incomp_bulk);
    D o[6];
    conserved(mm, o);
    T g = (T)0.0;
#pragma unroll
    for (int p = 0; p < 6; p++) g = g + l[p] * o[p].d;
    if (q < Q) d.E[srow * Q + q] = g;
    else if (q < 12) d.lout[3 * j + (q - 9)] = g;
    else d.guin[j] += (double)g;
}

// Transposed interpolations (one workgroup): velocity levels in reverse, then the stencil
// nodes' pre-collision moments and the outlet's copied column, folded into their rows.
template <class S, class T> __global__ void __launch_bounds__(256) adj_interp_kernel(StepArgs<S, T> a, AdjArgs<S, T> d)
{
    const BoundaryDev &b = a.b;
    const int tid = threadIdx.x, nt = blockDim.x;
    const bool interp_u = b.vel_interp || !b.rotated;
    for (int L = b.nlevels - 1; L >= 0; L--) {
        for (int t = b.level_start[L] + tid; t < b.level_start[L + 1]; t += nt) {
            const int k = b.level_order[t];
            T gx = d.lub_own[2 * k], gy = d.lub_own[2 * k + 1];
            if (interp_u)
                for (int e = d.rev_dep_start[k]; e < d.rev_dep_start[k + 1]; e++) {
                    const int kk = d.rev_dep[e] >> 3, c = d.rev_dep[e] & 7;
                    const T w = interp_weight<T>(b.real + (long long)kk * BT_NREAL, c);
                    gx = gx + w * d.lub[2 * kk];
                    gy = gy + w * d.lub[2 * kk + 1];
                }
            d.lub[2 * k] = gx;
            d.lub[2 * k + 1] = gy;
        }
        __syncthreads();
    }
    const T *Uf = d.Ufull;
    for (int u = tid; u < b.nU; u += nt) {
        T lux = 0, luy = 0, lMx = 0, lMy = 0;
        for (int e = d.rev_u_start[u]; e < d.rev_u_start[u + 1]; e++) {
            const int kk = d.rev_u[e] >> 3, c = d.rev_u[e] & 7;
            const T w = interp_weight<T>(b.real + (long long)kk * BT_NREAL, c);
            if (interp_u && b.dep[8 * kk + c] < 0) {
                lux = lux + w * d.lub[2 * kk];
                luy = luy + w * d.lub[2 * kk + 1];
            }
            if (b.rotated && b.mom_interp) {
                lMx = lMx + w * d.lMi[2 * kk];
                lMy = lMy + w * d.lMi[2 * kk + 1];
            }
        }
        if (!interp_u)
            for (int kk = 0; kk < b.nbct; kk++)
                if (b.uself[kk] == u) {
                    lux = lux + d.lub[2 * kk];
                    luy = luy + d.lub[2 * kk + 1];
                }
        const T c2 = (T)b.urot[3 * u], s2 = (T)b.urot[3 * u + 1], sc = (T)b.urot[3 * u + 2];
        const T lm3 = c2 * lMx + s2 * lMy, lm4 = s2 * lMx + c2 * lMy, lm5 = sc * (lMx - lMy);
        const int nU = b.nU;
        const T rho = Uf[u], ux = Uf[nU + u], uy = Uf[2 * nU + u];
        const T mxx = Uf[3 * nU + u], myy = Uf[4 * nU + u], mxy = Uf[5 * nU + u];
        const int row = d.spec[b.unode[u]];
        for (int q = 0; q < Q; q++) {
            const T v = ((((lux * ((T)EXk(q) - ux) + luy * ((T)EYk(q) - uy)) + lm3 * (Hxxk<T>(q) - mxx))
                          + lm4 * (Hyyk<T>(q) - myy)) + lm5 * (Hxyk<T>(q) - mxy)) / rho;
            d.E[row * Q + q] = d.E[row * Q + q] + v;
        }
    }
    __syncthreads();
    for (int j = tid; j < a.ny; j += nt) {
        const T lr = d.lout[3 * j], lx = d.lout[3 * j + 1], ly = d.lout[3 * j + 2];
        const T rho = d.outm[3 * j], ux = d.outm[3 * j + 1], uy = d.outm[3 * j + 2];
        const int row = d.spec[(long long)j * a.pitch + (a.nx - 2)];
        for (int q = 0; q < Q; q++)
            d.E[row * Q + q] = d.E[row * Q + q] + (lr + (lx * ((T)EXk(q) - ux) + ly * ((T)EYk(q) - uy)) / rho);
    }
}

// The same for walls of one dependency level with no stencil node in column nx-2 (the two
// phases then write disjoint special rows): lub = lub_own, one thread per stencil node and
// per outlet row across the grid.
template <class S, class T> __global__ void __launch_bounds__(256) adj_interp_flat_kernel(StepArgs<S, T> a, AdjArgs<S, T> d)
{
    const BoundaryDev &b = a.b;
    const bool interp_u = b.vel_interp || !b.rotated;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < b.nU) {
        const int u = t, nU = b.nU;
        const T *Uf = d.Ufull;
        T lux = 0, luy = 0, lMx = 0, lMy = 0;
        for (int e = d.rev_u_start[u]; e < d.rev_u_start[u + 1]; e++) {
            const int kk = d.rev_u[e] >> 3, c = d.rev_u[e] & 7;
            const T w = interp_weight<T>(b.real + (long long)kk * BT_NREAL, c);
            if (interp_u) {
                lux = lux + w * d.lub_own[2 * kk];
                luy = luy + w * d.lub_own[2 * kk + 1];
            }
            if (b.rotated && b.mom_interp) {
                lMx = lMx + w * d.lMi[2 * kk];
                lMy = lMy + w * d.lMi[2 * kk + 1];
            }
        }
        if (!interp_u)
            for (int kk = 0; kk < b.nbct; kk++)
                if (b.uself[kk] == u) {
                    lux = lux + d.lub_own[2 * kk];
                    luy = luy + d.lub_own[2 * kk + 1];
                }
        const T c2 = (T)b.urot[3 * u], s2 = (T)b.urot[3 * u + 1], sc = (T)b.urot[3 * u + 2];
        const T lm3 = c2 * lMx + s2 * lMy, lm4 = s2 * lMx + c2 * lMy, lm5 = sc * (lMx - lMy);
        const T rho = Uf[u], ux = Uf[nU + u], uy = Uf[2 * nU + u];
        const T mxx = Uf[3 * nU + u], myy = Uf[4 * nU + u], mxy = Uf[5 * nU + u];
        const int row = d.spec[b.unode[u]];
        for (int q = 0; q < Q; q++) {
            const T v = ((((lux * ((T)EXk(q) - ux) + luy * ((T)EYk(q) - uy)) + lm3 * (Hxxk<T>(q) - mxx))
                          + lm4 * (Hyyk<T>(q) - myy)) + lm5 * (Hxyk<T>(q) - mxy)) / rho;
            d.E[row * Q + q] = d.E[row * Q + q] + v;
        }
        return;
    }
    const int j = t - b.nU;
    if (j >= a.ny) return;
    const T lr = d.lout[3 * j], lx = d.lout[3 * j + 1], ly = d.lout[3 * j + 2];
    const T rho = d.outm[3 * j], ux = d.outm[3 * j + 1], uy = d.outm[3 * j + 2];
    const int row = d.spec[(long long)j * a.pitch + (a.nx - 2)];
    for (int q = 0; q < Q; q++)
        d.E[row * Q + q] = d.E[row * Q + q] + (lr + (lx * ((T)EXk(q) - ux) + ly * ((T)EYk(q) - uy)) / rho);
}

// Reverse step of every node: gather of the downstream population adjoints, plus dJ/dchi.
template <class S, class T> __device__ void adj_bulk_body(const StepArgs<S, T> &a, const AdjArgs<S, T> &d, long long t)
{
    using F = typename FormOf<T>::F;
    if (t >= (long long)a.nx * a.ny) return;
    const int i = (int)(t % a.nx), j = (int)(t / a.nx);
    const long long ns_ = (long long)j * a.pitch + i; // this node
    const T om = a.om;
    // Everything the 9 downstream neighbours contribute is loaded before any arithmetic (one
    // memory round trip, then the special-row values); a neighbour outside the domain (its
    // f_k is a stale value) reads this node instead and is skipped below.
    T l[Q][6], rho[Q], ux[Q], uy[Q], chi[Q];
    int row[Q];
    bool inside[Q], bulk[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int xi = i + EXk(k), xj = j + EYk(k);
        inside[k] = xi >= 0 && xi < a.nx && xj >= 0 && xj < a.ny;
        const int ci = inside[k] ? xi : i, cj = inside[k] ? xj : j;
        const long long n = (long long)cj * a.pitch + ci;
        bulk[k] = inside[k] && !on_ring(a, ci, cj) && !is_wall(a, ci, cj);
        load_lam(a, d.lam1, n, l[k]);
        rho[k] = F::rho_of((T)d.m1[n]);
        ux[k] = (T)d.m1[a.plane + n];
        uy[k] = (T)d.m1[2 * a.plane + n];
        chi[k] = a.chi ? a.chi[n] : (T)0.0;
        row[k] = d.spec[n];
    }
    T acc[6] = {0, 0, 0, 0, 0, 0};
    T lchi0 = (T)0.0; // this node's own chi term (k = 0 is the node itself)
#pragma unroll
    for (int k = 0; k < Q; k++) {
        if (!inside[k]) continue;
        T ak = (T)0.0;
        if (bulk[k]) {
            T lp[6], lchi;
            bulk_pre(l[k], rho[k], ux[k], uy[k], chi[k], om, lp, lchi);
            ak = pop_adjoint(lp, k);
            if (k == 0) lchi0 = lchi;
        }
        if (row[k] >= 0) ak = ak + d.E[row[k] * Q + k];
        T v[6];
        recon_row<T>(k, v);
        const T wa = wk<T>(k) * ak;
#pragma unroll
        for (int p = 0; p < 6; p++) acc[p] = acc[p] + wa * v[p];
    }
#pragma unroll
    for (int p = 0; p < 6; p++) d.lam0[p * a.plane + ns_] = acc[p];
    if (a.chi && bulk[0]) d.G[ns_] += (double)lchi0;
}

template <class S, class T> __global__ void __launch_bounds__(256) adj_bulk_kernel(StepArgs<S, T> a, AdjArgs<S, T> d)
{
    adj_bulk_body(a, d, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The bulk gather of one reverse step and, in the workgroups past nbulk, the prep of the
// next one (an, dn: the state before; its own special-row buffer): nothing the gather reads
// is written by that prep.
template <class S, class T>
__global__ void __launch_bounds__(256) adj_bulk_prep_kernel(StepArgs<S, T> a, AdjArgs<S, T> d, StepArgs<S, T> an,
                                                           AdjArgs<S, T> dn, int nbulk)
{
    if ((int)blockIdx.x < nbulk) adj_bulk_body(a, d, (long long)blockIdx.x * blockDim.x + threadIdx.x);
    else adj_prep_body(an, dn, ((int)blockIdx.x - nbulk) * (int)blockDim.x + (int)threadIdx.x);
}

template <class S, class T>
static int adjoint_step(const StepArgs<S, T> &a, const AdjArgs<S, T> &d, bool flat, hipStream_t st,
                        const StepArgs<S, T> *an = nullptr, const AdjArgs<S, T> *dn = nullptr, bool prepped = false)
{
    const BoundaryDev &b = a.b;
    // one dependency level: the stencil table comes with the prep (same gathers), and the
    // prep of the next step (an, dn) with this step's bulk gather; otherwise the wall
    // velocities need the stencil table first, in launches of their own
    const bool one = b.nlevels <= 1;
    const int nprep = std::max(std::max(d.ns * Q, b.nU), a.ny);
    if (!prepped) {
        AdjArgs<S, T> dp = d;
        dp.stencil = one ? 1 : 0;
        if (!one) {
            hipError_t e = launch_stencil_moments<S, T>(a, st);
            if (e != hipSuccess) return MRLBM_E_HIP;
            hipLaunchKernelGGL((adj_wallvel_kernel<S, T>), dim3(1), dim3(256), 0, st, a);
        }
        hipLaunchKernelGGL((adj_prep_kernel<S, T>), dim3((nprep + 255) / 256), dim3(256), 0, st, a, dp);
    }
    const int nloc = (b.nbct + d.nedge) * 13;
    hipLaunchKernelGGL((adj_local_kernel<S, T>), dim3((nloc + 255) / 256), dim3(256), 0, st, a, d);
    if (flat)
        hipLaunchKernelGGL((adj_interp_flat_kernel<S, T>), dim3((b.nU + a.ny + 255) / 256), dim3(256), 0, st, a, d);
    else
        hipLaunchKernelGGL((adj_interp_kernel<S, T>), dim3(1), dim3(256), 0, st, a, d);
    const long long N = (long long)a.nx * a.ny;
    const int nbulk = (int)((N + 255) / 256);
    if (an && one) {
        AdjArgs<S, T> dp = *dn;
        dp.stencil = 1;
        const int nnext = std::max(std::max(dn->ns * Q, an->b.nU), an->ny);
        hipLaunchKernelGGL((adj_bulk_prep_kernel<S, T>), dim3((unsigned)(nbulk + (nnext + 255) / 256)), dim3(256), 0, st,
                           a, d, *an, dp, nbulk);
    } else {
        hipLaunchKernelGGL((adj_bulk_kernel<S, T>), dim3((unsigned)nbulk), dim3(256), 0, st, a, d);
    }
    return hipGetLastError() == hipSuccess ? 0 : MRLBM_E_HIP;
}

// ---------------------------------------------------------------------------------
// Host side.

template <class T> static T *walloc(AdjointWork *w, size_t n, hipError_t &e)
{
    T *p = dmalloc<T>(n, e);
    if (e == hipSuccess) w->allocs.push_back(p);
    return p;
}

void adjoint_free(mrlbm_handle *h)
{
    if (!h || !h->adj) return;
    for (void *p : h->adj->allocs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {h->adj->evR, h->adj->evV})
        if (e) (void)hipEventDestroy(e);
    delete h->adj;
    h->adj = nullptr;
}

template <class T> static int upload_vec(mrlbm_handle *h, AdjointWork *w, const std::vector<T> &v, T **out)
{
    hipError_t e;
    *out = walloc<T>(w, std::max<size_t>(v.size(), 1), e);
    HIPCHK(e);
    if (!v.empty()) HIPCHK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// Static tables (special rows, edge list, reverse interpolation lists) and buffers.
template <class S, class T> static int adjoint_setup(mrlbm_handle *h, int K, int nseg)
{
    if (h->adj && h->adj->K == K && h->adj->nseg_cap >= nseg) return 0;
    adjoint_free(h);
    auto *w = new AdjointWork();
    h->adj = w;
    w->f32 = h->f32;
    w->K = K;
    w->nseg_cap = nseg;
    const BoundaryTables &bt = h->bt;
    const int nx = h->nx, ny = h->ny, pitch = h->pitch;
    const long long plane = h->plane;
    hipError_t e;
    const size_t st_words = (size_t)6 * plane;
    w->ckpt = walloc<S>(w, st_words * nseg, e);
    HIPCHK(e);
    w->seg[0] = walloc<S>(w, st_words * (K + 1), e);
    HIPCHK(e);
    { // a second segment buffer lets segment s-1 be recomputed while segment s is reversed
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const size_t seg_b = st_words * (K + 1) * sizeof(S);
        if (free_b > 2 * seg_b + ((size_t)1 << 30)) {
            w->seg[1] = walloc<S>(w, st_words * (K + 1), e);
            HIPCHK(e);
            w->U2 = walloc<T>(w, (size_t)5 * std::max(bt.nU, 1), e);
            HIPCHK(e);
            w->ub2 = walloc<T>(w, (size_t)2 * std::max(bt.nbct, 1), e);
            HIPCHK(e);
            HIPCHK(hipEventCreateWithFlags(&w->evR, hipEventDisableTiming | hipEventReleaseToDevice));
            HIPCHK(hipEventCreateWithFlags(&w->evV, hipEventDisableTiming | hipEventReleaseToDevice));
        }
    }
    for (int b = 0; b < 2; b++) {
        w->lam[b] = walloc<T>(w, st_words, e);
        HIPCHK(e);
    }
    w->G = walloc<double>(w, plane, e);
    HIPCHK(e);
    w->guin = walloc<double>(w, ny, e);
    HIPCHK(e);
    // special rows: wall nodes, stencil nodes, the outer ring, column nx-2
    std::vector<int32_t> spec((size_t)plane, -1);
    int ns = 0;
    auto add = [&](long long n) {
        if (spec[n] < 0) spec[n] = ns++;
    };
    for (int k = 0; k < bt.nbct; k++) add(bt.node[k]);
    for (int u = 0; u < bt.nU; u++) add(bt.unode[u]);
    std::vector<int32_t> edge;
    for (int i = 0; i < nx; i++) { edge.push_back(i); edge.push_back(0); }
    for (int i = 0; i < nx; i++) { edge.push_back(i); edge.push_back(ny - 1); }
    for (int j = 1; j < ny - 1; j++) { edge.push_back(0); edge.push_back(j); }
    for (int j = 1; j < ny - 1; j++) { edge.push_back(nx - 1); edge.push_back(j); }
    w->nedge = (int)edge.size() / 2;
    for (int e2 = 0; e2 < w->nedge; e2++) add((long long)edge[2 * e2 + 1] * pitch + edge[2 * e2]);
    for (int j = 0; j < ny; j++) add((long long)j * pitch + (nx - 2));
    w->ns = ns;
    w->flat = bt.nlevels <= 1;
    for (int u = 0; u < bt.nU; u++)
        if (bt.uij[2 * u] == nx - 2) w->flat = false;
    int rc = upload_vec(h, w, spec, &w->spec);
    rc |= upload_vec(h, w, edge, &w->edge_ij);
    // reverse lists of the interpolation corners: entries k*8 + c
    std::vector<int32_t> ds(bt.nbct + 1, 0), us(bt.nU + 1, 0);
    for (int k = 0; k < bt.nbct; k++)
        for (int c = 0; c < 8; c++) {
            if (bt.dep[8 * k + c] >= 0) ds[bt.dep[8 * k + c] + 1]++;
            us[bt.uc[12 * k + c] + 1]++;
        }
    for (int k = 0; k < bt.nbct; k++) ds[k + 1] += ds[k];
    for (int u = 0; u < bt.nU; u++) us[u + 1] += us[u];
    std::vector<int32_t> dl(std::max(ds[bt.nbct], 1)), ul(std::max(us[bt.nU], 1));
    std::vector<int32_t> dfill(ds.begin(), ds.end() - 1), ufill(us.begin(), us.end() - 1);
    for (int k = 0; k < bt.nbct; k++)
        for (int c = 0; c < 8; c++) {
            if (bt.dep[8 * k + c] >= 0) dl[dfill[bt.dep[8 * k + c]]++] = 8 * k + c;
            ul[ufill[bt.uc[12 * k + c]]++] = 8 * k + c;
        }
    rc |= upload_vec(h, w, ds, &w->rev_dep_start);
    rc |= upload_vec(h, w, dl, &w->rev_dep);
    rc |= upload_vec(h, w, us, &w->rev_u_start);
    rc |= upload_vec(h, w, ul, &w->rev_u);
    if (rc) return rc;
    w->E = walloc<T>(w, (size_t)ns * Q, e);
    if (e == hipSuccess) w->E2 = walloc<T>(w, (size_t)ns * Q, e);
    HIPCHK(e);
    w->Ufull = walloc<T>(w, (size_t)6 * std::max(bt.nU, 1), e);
    HIPCHK(e);
    w->outm = walloc<T>(w, (size_t)3 * ny, e);
    HIPCHK(e);
    w->lout = walloc<T>(w, (size_t)3 * ny, e);
    HIPCHK(e);
    w->lub_own = walloc<T>(w, (size_t)2 * std::max(bt.nbct, 1), e);
    HIPCHK(e);
    w->lub = walloc<T>(w, (size_t)2 * std::max(bt.nbct, 1), e);
    HIPCHK(e);
    w->lMi = walloc<T>(w, (size_t)2 * std::max(bt.nbct, 1), e);
    HIPCHK(e);
    return 0;
}

template <class S, class T>
static int adjoint_run(mrlbm_handle *h, const mrlbm_adjoint_opts *o, double *dJ_dchi, mrlbm_adjoint_result *res)
{
    using clock = std::chrono::steady_clock;
    const int nsteps = o->nsteps, window = o->window, K = o->checkpoint_interval;
    const int nseg = (nsteps + K - 1) / K;
    const long long plane = h->plane;
    const size_t st_words = (size_t)6 * plane;
    if (!h->chi) { // zero field: the gradient is still wanted
        const std::vector<double> zero((size_t)h->nx * h->ny, 0.0);
        const int rc = mrlbm_set_brinkman(h, zero.data());
        if (rc) return rc;
    }
    {
        const int rc = adjoint_setup<S, T>(h, K, nseg);
        if (rc) return rc;
    }
    AdjointWork *w = h->adj;
    // ---- forward with checkpoints; drag recorded over the window
    HIPCHK(hipStreamSynchronize(h->st));
    const auto t0 = clock::now();
    mrlbm_step_opts so;
    mrlbm_default_step_opts(&so);
    so.post_process = 1;
    so.statsbegin = o->iter_first + nsteps - window;
    so.statsend = o->iter_first + nsteps - 1;
    const size_t q0 = h->q_fd.size();
    for (int s = 0; s < nseg; s++) {
        HIPCHK(hipMemcpyAsync((S *)w->ckpt + st_words * s, h->S[h->cur], st_words * sizeof(S),
                              hipMemcpyDeviceToDevice, h->st));
        const int len = std::min(K, nsteps - s * K);
        const int rc = mrlbm_step(h, o->iter_first + s * K, len, &so);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->st));
    const auto t1 = clock::now();
    if (h->q_fd.size() - q0 != (size_t)window) FAIL(MRLBM_E_STATE, "drag series has the wrong length");
    double fsum = 0.0;
    for (size_t s = q0; s < h->q_fd.size(); s++) fsum += h->q_fd[s];
    h->q_fd.resize(q0);
    h->q_fl.resize(q0);
    h->q_force_iter.resize(q0);
    const double uref = h->p.inlet_profile == MRLBM_INLET_PARABOLIC ? (2.0 / 3.0) * h->p.uo : h->p.uo;
    const double norm = 0.5 * 1.0 * uref * uref * (2.0 * h->geo.r_cyl);
    const double J = fsum / window / norm;
    // ---- reverse sweep
    HIPCHK(hipMemsetAsync(w->lam[0], 0, st_words * sizeof(T), h->st));
    HIPCHK(hipMemsetAsync(w->lam[1], 0, st_words * sizeof(T), h->st));
    HIPCHK(hipMemsetAsync(w->G, 0, plane * sizeof(double), h->st));
    HIPCHK(hipMemsetAsync(w->guin, 0, h->ny * sizeof(double), h->st));
    HIPCHK(hipMemsetAsync(w->lub, 0, (size_t)2 * std::max(h->bt.nbct, 1) * sizeof(T), h->st));
    int lc = 0;
    const T wJ = (T)(1.0 / ((double)window * norm));
    const int simple_blocks = (int)(((long long)h->nx * h->ny + 255) / 256);
    // Segment s's states are recomputed from its checkpoint before it is reversed.  With two
    // segment buffers the recompute of segment s-1 runs on stream B (its own wall scratch)
    // while stream A reverses segment s; B waits for the reverse that last used its buffer.
    const bool pipe = w->seg[1] != nullptr && !o->serial_recompute;
    auto recompute = [&](int s, hipStream_t q) -> int {
        const int len = std::min(K, nsteps - s * K);
        S *seg = (S *)w->seg[pipe ? (s & 1) : 0];
        HIPCHK(hipMemcpyAsync(seg, (S *)w->ckpt + st_words * s, st_words * sizeof(S), hipMemcpyDeviceToDevice, q));
        for (int t = 0; t < len; t++) {
            auto a = base_args<S, T>(h);
            a.src = seg + st_words * t;
            a.dst = seg + st_words * (t + 1);
            a.pops = nullptr;
            if (pipe) {
                a.b.U = w->U2;
                a.b.ub = w->ub2;
            }
            HIPCHK(launch_bulk<S, T>(a, simple_blocks, q));
            HIPCHK(launch_edges<S, T>(a, q));
            HIPCHK(launch_boundary<S, T>(a, q));
        }
        return 0;
    };
    if (pipe) {
        HIPCHK(hipEventRecord(h->evA, h->st)); // the memsets above
        HIPCHK(hipStreamWaitEvent(h->st2, h->evA, 0));
        const int rc = recompute(nseg - 1, h->st2);
        if (rc) return rc;
        HIPCHK(hipEventRecord(w->evR, h->st2));
    }
    for (int s = nseg - 1; s >= 0; s--) {
        const int len = std::min(K, nsteps - s * K);
        S *seg = (S *)w->seg[pipe ? (s & 1) : 0];
        if (pipe) {
            HIPCHK(hipStreamWaitEvent(h->st, w->evR, 0)); // segment s recomputed
            if (s > 0) { // segment s-1 into the buffer the reverse of segment s+1 used
                if (s + 1 < nseg) HIPCHK(hipStreamWaitEvent(h->st2, w->evV, 0));
                const int rc = recompute(s - 1, h->st2);
                if (rc) return rc;
                HIPCHK(hipEventRecord(w->evR, h->st2));
            }
        } else {
            const int rc = recompute(s, h->st);
            if (rc) return rc;
        }
        // arguments of the reverse of step s*K + t (state t of the segment -> t + 1), with the
        // lam and special-row buffers of parity lcv
        auto make = [&](int t, int lcv, StepArgs<S, T> &a, AdjArgs<S, T> &d) {
            const int step = s * K + t;
            a = base_args<S, T>(h);
            a.src = seg + st_words * t;
            a.dst = nullptr;
            a.pops = nullptr;
            d = AdjArgs<S, T>{};
            d.m1 = seg + st_words * (t + 1);
            d.lam1 = (const T *)w->lam[lcv];
            d.lam0 = (T *)w->lam[lcv ^ 1];
            d.G = w->G;
            d.guin = w->guin;
            d.spec = w->spec;
            d.E = (T *)(lcv ? w->E2 : w->E);
            d.ns = w->ns;
            d.nedge = w->nedge;
            d.edge_ij = w->edge_ij;
            d.Ufull = (T *)w->Ufull;
            d.outm = (T *)w->outm;
            d.lub_own = (T *)w->lub_own;
            d.lub = (T *)w->lub;
            d.lMi = (T *)w->lMi;
            d.lout = (T *)w->lout;
            d.rev_dep_start = w->rev_dep_start;
            d.rev_dep = w->rev_dep;
            d.rev_u_start = w->rev_u_start;
            d.rev_u = w->rev_u;
            d.wJ = step >= nsteps - window ? wJ : (T)0.0;
        };
        // within the segment, each step's bulk gather also does the next step's prep
        bool prepped = false;
        for (int t = len - 1; t >= 0; t--) {
            StepArgs<S, T> a, an;
            AdjArgs<S, T> d, dn;
            make(t, lc, a, d);
            int rc;
            if (t >= 1) {
                make(t - 1, lc ^ 1, an, dn);
                rc = adjoint_step<S, T>(a, d, w->flat, h->st, &an, &dn, prepped);
                prepped = h->bt.nlevels <= 1;
            } else {
                rc = adjoint_step<S, T>(a, d, w->flat, h->st, nullptr, nullptr, prepped);
                prepped = false;
            }
            if (rc) FAIL(rc, "adjoint kernel launch failed");
            lc ^= 1;
        }
        if (pipe) HIPCHK(hipEventRecord(w->evV, h->st)); // buffer s & 1 is free again
    }
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipStreamSynchronize(h->st2));
    const auto t2 = clock::now();
    // ---- outputs
    std::vector<double> guin(h->ny);
    HIPCHK(hipMemcpy(guin.data(), w->guin, h->ny * sizeof(double), hipMemcpyDeviceToHost));
    double duin = 0.0;
    for (int j = 0; j < h->ny; j++) {
        double phi = 1.0;
        if (h->p.inlet_profile == MRLBM_INLET_PARABOLIC) {
            const double sj = (double)j / (double)(h->ny - 1);
            phi = 4.0 * sj * (1.0 - sj);
        }
        duin += guin[j] * phi;
    }
    if (dJ_dchi) {
        std::vector<double> g((size_t)plane);
        HIPCHK(hipMemcpy(g.data(), w->G, plane * sizeof(double), hipMemcpyDeviceToHost));
        for (int j = 0; j < h->ny; j++)
            for (int i = 0; i < h->nx; i++) dJ_dchi[(size_t)j * h->nx + i] = g[(size_t)j * h->pitch + i];
    }
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->J = J;
        res->dJ_duo = duin - 2.0 * J / h->p.uo;
        res->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        res->adjoint_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return 0;
}

} // namespace mrlbm

extern "C" void mrlbm_default_adjoint_opts(mrlbm_adjoint_opts *o)
{
    std::memset(o, 0, sizeof *o);
    o->iter_first = 2;
    o->nsteps = 4000;
    o->window = 2000;
    o->checkpoint_interval = 200;
}

// The adjoint's forward sweep steps the handle with its own statistics window: the running
// inlet-density mean and its pending-sample bookkeeping are saved before and restored after,
// and a zero Brinkman field installed for the sweep is removed again, so later mrlbm_step
// calls see the handle as the caller left it (and keep the two-step path).
template <class S, class T>
static int adjoint_typed(mrlbm_handle *h, const mrlbm_adjoint_opts *o, double *dJ_dchi, mrlbm_adjoint_result *res)
{
    const bool had_chi = h->chi != nullptr;
    double rho_mean = 0.0;
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(&rho_mean, h->d_rho_mean, sizeof(double), hipMemcpyDeviceToHost));
    const int valid = h->prev_inlet_valid, buf = h->prev_inlet_buf;
    const double c = h->prev_inlet_c;
    const int rc = adjoint_run<S, T>(h, o, dJ_dchi, res);
    (void)hipStreamSynchronize(h->st);
    (void)hipStreamSynchronize(h->st2);
    const hipError_t e = hipMemcpy(h->d_rho_mean, &rho_mean, sizeof(double), hipMemcpyHostToDevice);
    h->prev_inlet_valid = valid;
    h->prev_inlet_buf = buf;
    h->prev_inlet_c = c;
    if (!had_chi && h->chi) {
        const int rc2 = mrlbm_set_brinkman(h, nullptr);
        if (!rc && rc2) return rc2;
    }
    if (rc) return rc;
    HIPCHK(e);
    return 0;
}

extern "C" int mrlbm_adjoint_drag(mrlbm_handle *h, const mrlbm_adjoint_opts *o, double *dJ_dchi,
                                  mrlbm_adjoint_result *res)
{
    if (!h || !o) return MRLBM_E_ARG;
    if (o->nsteps < 1 || o->window < 1 || o->window > o->nsteps || o->checkpoint_interval < 1)
        FAIL(MRLBM_E_ARG, "need nsteps >= 1, 1 <= window <= nsteps, checkpoint_interval >= 1");
    if (!h->has_b) FAIL(MRLBM_E_ARG, "the drag adjoint needs a cylinder");
    if (h->p.x_periodic || h->p.y_periodic) FAIL(MRLBM_E_ARG, "the drag adjoint needs walls on every side (no periodic axis)");
    if (h->p.incomp_bulk) FAIL(MRLBM_E_ARG, "the drag adjoint does not cover incomp_bulk");
    if (h->from_pops) FAIL(MRLBM_E_STATE, "the drag adjoint starts from moments: step once after mrlbm_set_populations");
    {
        const BoundaryTables &bt = h->bt;
        const int box_r = bt.bx0 + bt.bw; // the outlet's copy column must not touch the wall stencil
        if (box_r + 2 >= h->nx - 2 || bt.bx0 < 3 || bt.by0 < 3 || bt.by0 + bt.bh + 3 > h->ny)
            FAIL(MRLBM_E_ARG, "the drag adjoint needs the cylinder at least 3 nodes from every edge");
    }
    const size_t st_bytes = (size_t)6 * h->plane * (h->f32 ? 4 : 8);
    const size_t need = st_bytes * ((size_t)o->checkpoint_interval + 1 + (o->nsteps + o->checkpoint_interval - 1) / o->checkpoint_interval);
    if (need > ((size_t)200 << 30)) FAIL(MRLBM_E_NOMEM, "checkpoint and segment buffers exceed 200 GiB");
    return h->f32 ? adjoint_typed<float, float>(h, o, dJ_dchi, res) : adjoint_typed<double, double>(h, o, dJ_dchi, res);
}
