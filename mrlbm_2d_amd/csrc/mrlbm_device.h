// mrlbm_device.h -- per-node device functions shared by the step kernels and the adjoint:
// moment loads, the pull rule with the reference's edge conventions, the node update
// (moments, edge BC, collision, store) and the curved-wall BC of one boundary node.
#pragma once
#include "mrlbm_geometry.h"
#include "mrlbm_kernels.h"
#include "mrlbm_lattice.h"

namespace mrlbm {

template <class T> struct FormOf;
template <> struct FormOf<double> { using F = FormF64; };
template <> struct FormOf<float> { using F = FormF32; };

template <class S, class T>
__device__ __forceinline__ void load_moments(const StepArgs<S, T> &a, int i, int j, S m[6])
{
    const long long n = (long long)j * a.pitch + i;
#pragma unroll
    for (int p = 0; p < 6; p++) m[p] = a.src[p * a.plane + n];
}

// Rest state in stored units: rho = 1 (FP64) or rho-1 = 0 (FP32), everything else 0.
template <class S> __device__ __forceinline__ void rest_moments(S m[6])
{
    m[0] = FormOf<S>::F::dev ? (S)0 : (S)1;
#pragma unroll
    for (int p = 1; p < 6; p++) m[p] = (S)0;
}

// Population k of the regularized reconstruction of stored moments m: the same
// expression as recon9_f64 / recon9_dev for that k.
template <class S, class T> __device__ __forceinline__ T recon_one(const S m[6], int k)
{
    T f[Q];
    FormOf<T>::F::recon(m, f);
    T v = f[0];
#pragma unroll
    for (int q = 1; q < Q; q++)
        if (q == k) v = f[q];
    return v;
}

// Population k arriving at (i, j) after streaming (2d_mrlbm.f90:508-535): pulled from
// the upstream node; otherwise the shifted periodic wrap (same column for y, same row
// for x, y applied last) or, without periodicity, the stale initial value.
template <class S, class T> __device__ T pull_rule(const StepArgs<S, T> &a, int i, int j, int k)
{
    if (a.pops) return a.pops[(long long)k * a.plane + (long long)j * a.pitch + i];
    int si = i - EXk(k), sj = j - EYk(k);
    const bool xin = si >= 0 && si < a.nx, yin = sj >= 0 && sj < a.ny;
    if (!(xin && yin)) {
        if (!yin && a.yper) {
            si = i;
            sj = sj < 0 ? a.ny - 1 : 0;
        } else if (!xin && a.xper) {
            si = si < 0 ? a.nx - 1 : 0;
            sj = j;
        } else {
            const T *g = a.ghost;
            if (si < 0) return g[j * Q + k];
            if (si >= a.nx) return g[(a.ny + j) * Q + k];
            if (sj < 0) return g[(2 * a.ny + i) * Q + k];
            return g[(2 * a.ny + a.nx + i) * Q + k];
        }
    }
    S m[6];
    load_moments(a, si, sj, m);
    return recon_one<S, T>(m, k);
}

template <class S, class T> __device__ void gather_generic(const StepArgs<S, T> &a, int i, int j, T f[Q])
{
    for (int k = 0; k < Q; k++) f[k] = pull_rule(a, i, j, k);
}

// Gather at any node, with the same rules as pull_rule, but every source (upstream node,
// periodic wrap or stale-population table entry) is located first and all loads are issued
// before any arithmetic: a thread of the outer ring waits for one memory round trip.
template <class S, class T> __device__ void gather_edge(const StepArgs<S, T> &a, int i, int j, T f[Q])
{
    if (a.pops) {
        gather_generic(a, i, j, f);
        return;
    }
    S m[Q][6];
    T gv[Q];
    bool stale[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        int si = i - EXk(k), sj = j - EYk(k), gi = 0;
        const bool xin = si >= 0 && si < a.nx, yin = sj >= 0 && sj < a.ny;
        stale[k] = false;
        if (!(xin && yin)) {
            if (!yin && a.yper) {
                si = i;
                sj = sj < 0 ? a.ny - 1 : 0;
            } else if (!xin && a.xper) {
                si = si < 0 ? a.nx - 1 : 0;
                sj = j;
            } else { // the table entry; (i, j) itself stands in for the unused moment load
                stale[k] = true;
                gi = si < 0 ? j * Q + k
                   : si >= a.nx ? (a.ny + j) * Q + k
                   : sj < 0 ? (2 * a.ny + i) * Q + k
                            : (2 * a.ny + a.nx + i) * Q + k;
                si = i;
                sj = j;
            }
        }
        load_moments(a, si, sj, m[k]);
        gv[k] = a.ghost[gi];
    }
#pragma unroll
    for (int k = 0; k < Q; k++) {
        T g[Q];
        FormOf<T>::F::recon(m[k], g);
        f[k] = stale[k] ? gv[k] : g[k];
    }
}

// Gather at a node whose 8 neighbours are all inside the domain (the cylinder's stencil
// and boundary nodes): all 54 loads are issued before any arithmetic, so one thread
// waits for one memory round trip instead of nine.
template <class S, class T> __device__ void gather_interior(const StepArgs<S, T> &a, int i, int j, T f[Q])
{
    if (a.pops) {
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = a.pops[(long long)k * a.plane + (long long)j * a.pitch + i];
        return;
    }
    S m[Q][6];
#pragma unroll
    for (int k = 0; k < Q; k++) load_moments(a, i - EXk(k), j - EYk(k), m[k]);
#pragma unroll
    for (int k = 0; k < Q; k++) {
        T g[Q];
        FormOf<T>::F::recon(m[k], g);
        f[k] = g[k];
    }
}

// The five stencil values of stencil node u (moments m as F::moments gives them): rho, ux,
// uy and, for the rotated wall BC with moment interpolation, the rotated Mxx', Myy' (else 0;
// 2d_mrlbm.f90:840-869 for the rotation).
template <class T> __device__ __forceinline__ void stencil_values(const BoundaryDev &b, int u, const T m[6], T v[5])
{
    v[0] = m[0];
    v[1] = m[1];
    v[2] = m[2];
    v[3] = v[4] = (T)0.0;
    if (b.rotated && b.mom_interp) {
        const T c2 = (T)b.urot[3 * u], s2 = (T)b.urot[3 * u + 1], sc = (T)b.urot[3 * u + 2];
        v[3] = ((m[3] * c2) + (m[4] * s2)) + (m[5] * sc);
        v[4] = ((m[3] * s2) + (m[4] * c2)) - (m[5] * sc);
    }
}

// Pre-collision values of stencil node u into the wall's stencil table.
template <class S, class T> __device__ __forceinline__ void store_stencil_moments(const StepArgs<S, T> &a, int u, const T m[6])
{
    const BoundaryDev &b = a.b;
    T *U = (T *)b.U;
    const int nU = b.nU;
    T v[5];
    stencil_values(b, u, m, v);
    U[u] = v[0];
    U[nU + u] = v[1];
    U[2 * nU + u] = v[2];
    if (b.rotated && b.mom_interp) {
        U[3 * nU + u] = v[3];
        U[4 * nU + u] = v[4];
    }
}

// Stencil values read by boundary node k: plane p (0 rho, 1 ux, 2 uy, 3 Mxx', 4 Myy') of its
// reference c (0..11 the corners uc[12k + c], 12 the node itself), from the stencil table
// in memory (UGlobal) or from a wall group's LDS slots (ULocal).
template <class T> struct UGlobal {
    const T *U;
    int nU;
    const int32_t *uc, *uself;
    __device__ __forceinline__ T operator()(int k, int p, int c) const
    {
        return U[p * nU + (c < 12 ? uc[12 * k + c] : uself[k])];
    }
};
template <class T> struct ULocal {
    const T (*sU)[WALL_GROUP_MAXU];
    const int16_t *slot;
    __device__ __forceinline__ T operator()(int k, int p, int c) const { return sU[p][slot[13 * k + c]]; }
};

// Dirichlet values and the regularized edge BC, in the reference's order
// outlet -> inlet -> top -> bottom with the last writer winning (2d_mrlbm.f90:247-351).
template <class S, class T>
__device__ void edge_update(const StepArgs<S, T> &a, int c, int j, const T fpop[Q], T m[6], const T *gout = nullptr)
{
    using F = typename FormOf<T>::F;
    const bool out = !a.xper && c == a.nx - 1, in = !a.xper && c == 0;
    const bool top = !a.yper && j == a.ny - 1, bot = !a.yper && j == 0;
    if (!(out || in || top || bot)) return;
    T ff[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) ff[k] = F::dev ? fpop[k] + F::base(k) : fpop[k];
    T mm[6];
#pragma unroll
    for (int p = 0; p < 6; p++) mm[p] = m[p];
    T d_keep = m[0];
    mm[0] = F::rho_of(m[0]);
    if (out) { // rho, u copied from column nx-1 before any BC (2d_mrlbm.f90:249-253)
        T g[Q], mo[6];
        if (gout) { // the caller's gather at (nx-2, j)
#pragma unroll
            for (int k = 0; k < Q; k++) g[k] = gout[k];
        } else {
            gather_edge(a, a.nx - 2, j, g);
        }
        F::moments(g, mo);
        d_keep = mo[0];
        mm[0] = F::rho_of(mo[0]);
        mm[1] = mo[1];
        mm[2] = mo[2];
    }
    if (in) {
        mm[1] = a.uin[j];
        mm[2] = a.vin;
    }
    if (top || bot) {
        mm[1] = (T)0.0;
        mm[2] = (T)0.0;
    }
    const int side = bot ? 3 : (top ? 2 : (in ? 1 : 0));
    edge_bc<T>(side, a.incomp_bc, ff, mm, a.om);
    const bool rho_written = side != 0; // the outlet BC keeps the copied density
#pragma unroll
    for (int p = 1; p < 6; p++) m[p] = mm[p];
    m[0] = rho_written ? F::store_rho(mm[0]) : d_keep;
}

// Output nodes of an NS-step pass: NS or more nodes from the outer ring and outside the
// cylinder box dilated by NS-1 (everything else is finished by the band sub-steps).
template <class S, class T, int NS> __device__ __forceinline__ bool n_step_node(const StepArgs<S, T> &a, int c, int j)
{
    if (c < NS || c > a.nx - 1 - NS || j < NS || j > a.ny - 1 - NS) return false;
    return !(a.bw > 0 && c >= a.bx0 - (NS - 1) && c <= a.bx0 + a.bw + (NS - 2) && j >= a.by0 - (NS - 1) &&
             j <= a.by0 + a.bh + (NS - 2));
}

// Whether this sub-step records the samples and probes of node (c, j): in a four-step pass
// (skip_ns = 4) the four-step bulk records those of the nodes it writes, at every level.
template <class S, class T> __device__ __forceinline__ bool records_node(const StepArgs<S, T> &a, int c, int j)
{
    return a.skip_ns != 4 || !n_step_node<S, T, 4>(a, c, j);
}

// Everything after the gather for one bulk node: moments, edge BC, the Brinkman
// penalization u <- (1 - chi) u of non-edge nodes (CHI, not in the reference), solid-node
// zeroing on plot steps (2d_mrlbm.f90:375-384), recording, collision, store.
template <class S, class T, bool CHI = false>
__device__ __forceinline__ void node_update(const StepArgs<S, T> &a, int c, int j, const T f[Q], bool edge,
                                            const T *gout = nullptr)
{
    using F = typename FormOf<T>::F;
    int cls = 0;
    const int bi = c - a.bx0, bj = j - a.by0;
    if ((unsigned)bi < (unsigned)a.bw && (unsigned)bj < (unsigned)a.bh) cls = a.boxcls[bj * a.bw + bi];
    if (cls & 1) return; // boundary node: the boundary workgroup writes it
    const long long n = (long long)j * a.pitch + c;
    T m[6];
    F::moments(f, m);
    if (edge) edge_update(a, c, j, f, m, gout);
    if constexpr (CHI) {
        if (a.chi && !edge) {
            const T keep = (T)1.0 - a.chi[n];
            m[1] = m[1] * keep;
            m[2] = m[2] * keep;
        }
    }
    if (a.zero_solid && (cls & 2)) {
        m[1] = (T)0.0;
        m[2] = (T)0.0;
    }
    if (a.inlet_out && c == 0) a.inlet_out[j] = F::dev ? 1.0 + (double)m[0] : (double)m[0];
    const bool own = records_node(a, c, j);
    if (a.rec_slot >= 0 && own) {
        for (int p = 0; p < a.nprobe; p++)
            if (n == a.probe_node[p]) {
                double *r = a.rec + (long long)a.rec_slot * a.rec_stride + 2 + 3 * p;
                r[0] = F::dev ? 1.0 + (double)m[0] : (double)m[0];
                r[1] = m[1];
                r[2] = m[2];
            }
    }
    if (a.sample && own) {
        const double ux = m[1], uy = m[2];
        a.acc[n] += ux;
        a.acc[a.plane + n] += uy;
        a.acc[2 * a.plane + n] += ux * ux;
        a.acc[3 * a.plane + n] += uy * uy;
    }
    collide<T>(m, a.om, a.om1, a.incomp_bulk);
#pragma unroll
    for (int p = 0; p < 6; p++) a.dst[p * a.plane + n] = (S)m[p];
}

// ---------------------------------------------------------------------------------
// Boundary workgroup.

__device__ inline double block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; w++) t += sh[w];
    __syncthreads();
    return t;
}

// Rotated moment-based wall BC (2d_mrlbm.f90:1635-1765; weak :1767-1871; rhoeq :1873-1972).
// f: full streamed populations at the node; (uxp, uyp): velocity in the wall frame;
// Mxx, Myy: interpolated rotated moments.  Writes rho and the moments rotated back.
template <class T>
__device__ void wall_bc_rotated(const T f[Q], T uxp, T uyp, const double *R, uint16_t in, uint16_t out,
                                int mode, T om, T Mxx, T Myy, T m[6])
{
    T rhoI = 0, mxyI = 0, Ap = 0, Ep = 0, B11p = 0, B22p = 0, B12p = 0, Dxy = 0;
    T F11xy = 0, F22xy = 0, F12xy = 0;
    for (int k = 0; k < Q; k++) {
        const T cxp = (T)R[BT_CXP + k], cyp = (T)R[BT_CYP + k];
        const T hxx = (T)R[BT_HXXP + k], hyy = (T)R[BT_HYYP + k], hxy = (T)R[BT_HXYP + k];
        const T w = wk<T>(k);
        const T A = w * (((T)1.0 + ((T)3.0 * uxp) * cxp) + ((T)3.0 * uyp) * cyp);
        const T E = w * (((((T)1.0 + ((T)3.0 * uxp) * cxp) + ((T)3.0 * uyp) * cyp)
                          + ((T)4.50 * (uxp * uxp)) * hxx + ((T)4.50 * (uyp * uyp)) * hyy)
                         + ((T)9.0 * (uxp * uyp)) * hxy);
        const T B11 = ((T)4.50 * w) * hxx, B22 = ((T)4.50 * w) * hyy, B12 = ((T)4.50 * w) * hxy;
        if (mode == 0 && ((out >> k) & 1)) {
            Ap = Ap + A;
            Ep = Ep + E;
            B11p = B11p + B11;
            B22p = B22p + B22;
            B12p = B12p + B12;
        }
        if ((in >> k) & 1) {
            rhoI = rhoI + f[k];
            mxyI = mxyI + f[k] * hxy;
            if (mode == 2) {
                Ap = Ap + A;
                B11p = B11p + B11;
                B22p = B22p + B22;
                B12p = B12p + B12;
            }
            if (mode == 1) Ep = Ep + E;
            Dxy = Dxy + A * hxy;
            F11xy = F11xy + B11 * hxy;
            F22xy = F22xy + B22 * hxy;
            F12xy = F12xy + B12 * hxy;
        }
    }
    mxyI = mxyI * ((T)1.0 / rhoI);
    T Mxy, rho;
    if (mode == 0) {
        const T G = ((T)1.0 - om) * Ap + om * Ep;
        const T J11 = ((T)1.0 - om) * B11p, J22 = ((T)1.0 - om) * B22p, J12 = ((T)1.0 - om) * B12p;
        const T L11 = J11 * mxyI - F11xy, L22 = J22 * mxyI - F22xy, L12 = (T)2.0 * (J12 * mxyI - F12xy);
        const T Rxy = Dxy - G * mxyI;
        Mxy = ((Rxy - Mxx * L11) - Myy * L22) / L12;
        const T den = (((T)1.0 - om) * Ap + ((T)1.0 - om) * (((Mxx * B11p) + (Myy * B22p)) + (((T)2.0 * Mxy) * B12p)))
                      + om * Ep;
        rho = rhoI / den;
    } else if (mode == 2) {
        const T L11 = B11p * mxyI - F11xy, L22 = B22p * mxyI - F22xy, L12 = (T)2.0 * (B12p * mxyI - F12xy);
        const T Rxy = Dxy - Ap * mxyI;
        Mxy = ((Rxy - Mxx * L11) - Myy * L22) / L12;
        const T den = Ap + (((Mxx * B11p) + (Myy * B22p)) + (((T)2.0 * Mxy) * B12p));
        rho = rhoI / den;
    } else {
        const T Rxy = Ep * mxyI - Dxy;
        Mxy = ((Rxy - Mxx * F11xy) - Myy * F22xy) / ((T)2.0 * F12xy);
        rho = rhoI / Ep;
    }
    const T c = (T)R[BT_CTH], s = (T)R[BT_STH], s2 = (T)R[BT_S2TH], c2 = (T)R[BT_C2TH];
    m[0] = rho;
    m[3] = ((Mxx * (c * c)) + (Myy * (s * s))) - (Mxy * s2);
    m[4] = ((Mxx * (s * s)) + (Myy * (c * c))) + (Mxy * s2);
    m[5] = (((Mxx - Myy) * (T)0.50) * s2) + (Mxy * c2);
}

// Unrotated moment-based wall BC with Cramer's rule (2d_mrlbm.f90:1496-1633, 1974-2003).
template <class T>
__device__ void wall_bc_unrotated(const T f[Q], T ux, T uy, uint16_t in, uint16_t out, T om, T m[6])
{
    T rhoI = 0, mxxI = 0, myyI = 0, mxyI = 0, Ap = 0, Ep = 0, B11p = 0, B22p = 0, B12p = 0;
    T Dxx = 0, Dyy = 0, Dxy = 0;
    T F11xx = 0, F22xx = 0, F12xx = 0, F11yy = 0, F22yy = 0, F12yy = 0, F11xy = 0, F22xy = 0, F12xy = 0;
    for (int k = 0; k < Q; k++) {
        const T ex = (T)EXk(k), ey = (T)EYk(k);
        const T hxx = Hxxk<T>(k), hyy = Hyyk<T>(k), hxy = Hxyk<T>(k);
        const T w = wk<T>(k);
        const T A = w * (((T)1.0 + ((T)3.0 * ux) * ex) + ((T)3.0 * uy) * ey);
        const T E = w * (((((T)1.0 + ((T)3.0 * ux) * ex) + ((T)3.0 * uy) * ey) + ((T)4.50 * (ux * ux)) * hxx
                          + ((T)4.50 * (uy * uy)) * hyy) + ((T)9.0 * (ux * uy)) * hxy);
        const T B11 = ((T)4.50 * w) * hxx, B22 = ((T)4.50 * w) * hyy, B12 = ((T)4.50 * w) * hxy;
        if ((out >> k) & 1) {
            Ap = Ap + A; Ep = Ep + E;
            B11p = B11p + B11; B22p = B22p + B22; B12p = B12p + B12;
        }
        if ((in >> k) & 1) {
            rhoI = rhoI + f[k];
            mxxI = mxxI + f[k] * hxx;
            myyI = myyI + f[k] * hyy;
            mxyI = mxyI + f[k] * hxy;
            Dxx = Dxx + A * hxx; Dyy = Dyy + A * hyy; Dxy = Dxy + A * hxy;
            F11xx = F11xx + B11 * hxx; F22xx = F22xx + B22 * hxx; F12xx = F12xx + B12 * hxx;
            F11yy = F11yy + B11 * hyy; F22yy = F22yy + B22 * hyy; F12yy = F12yy + B12 * hyy;
            F11xy = F11xy + B11 * hxy; F22xy = F22xy + B22 * hxy; F12xy = F12xy + B12 * hxy;
        }
    }
    const T inv = (T)1.0 / rhoI;
    mxxI = mxxI * inv; myyI = myyI * inv; mxyI = mxyI * inv;
    const T G = ((T)1.0 - om) * Ap + om * Ep;
    const T J11 = ((T)1.0 - om) * B11p, J22 = ((T)1.0 - om) * B22p, J12 = ((T)1.0 - om) * B12p;
    const T a1 = J11 * mxxI - F11xx, b1 = J22 * mxxI - F22xx, c1 = (T)2.0 * (J12 * mxxI - F12xx);
    const T a2 = J11 * myyI - F11yy, b2 = J22 * myyI - F22yy, c2 = (T)2.0 * (J12 * myyI - F12yy);
    const T a3 = J11 * mxyI - F11xy, b3 = J22 * mxyI - F22xy, c3 = (T)2.0 * (J12 * mxyI - F12xy);
    const T d1 = Dxx - G * mxxI, d2 = Dyy - G * myyI, d3 = Dxy - G * mxyI;
    const T den = (((((a3 * b2) * c1 - (a2 * b3) * c1) - (a3 * b1) * c2) + (a1 * b3) * c2) + (a2 * b1) * c3)
                  - (a1 * b2) * c3;
    const T x11 = -((((((-b3 * c2) * d1 + (b2 * c3) * d1) + (b3 * c1) * d2) - (b1 * c3) * d2) - (b2 * c1) * d3)
                    + (b1 * c2) * d3) / den;
    const T x22 = -((((((a3 * c2) * d1 - (a2 * c3) * d1) - (a3 * c1) * d2) + (a1 * c3) * d2) + (a2 * c1) * d3)
                    - (a1 * c2) * d3) / den;
    const T x12 = -((((((-a3 * b2) * d1 + (a2 * b3) * d1) + (a3 * b1) * d2) - (a1 * b3) * d2) - (a2 * b1) * d3)
                    + (a1 * b2) * d3) / den;
    const T dn = (((T)1.0 - om) * Ap + ((T)1.0 - om) * (((x11 * B11p) + (x22 * B22p)) + (((T)2.0 * x12) * B12p))) + om * Ep;
    m[0] = rhoI / dn;
    m[3] = x11;
    m[4] = x22;
    m[5] = x12;
}

template <class T> __device__ __forceinline__ T bilin(T xd, T yd, T c0, T c1, T c2, T c3)
{ // 2d_mrlbm.f90:786-795
    const T v0 = ((T)1.0 - xd) * c0 + xd * c1;
    const T v1 = ((T)1.0 - xd) * c2 + xd * c3;
    return ((T)1.0 - yd) * v0 + yd * v1;
}

__device__ inline double quadratic_fit(double r1, double p1, double r2, double p2, double r3, double p3, double rs)
{ // quadratic_interpolation (2d_mrlbm.f90:743-766) with precomputed radii
    const double den = ((r1 - r2) * (r1 - r3)) * (r2 - r3);
    const double a0 = (((((r1 * r3) * p2) * (r3 - r1)) + ((r2 * r2) * ((r3 * p1) - (r1 * p3))))
                       + (r2 * (((r1 * r1) * p3) - ((r3 * r3) * p1)))) / den;
    const double a1 = ((((r3 * r3) * (p1 - p2)) + ((r1 * r1) * (p2 - p3))) + ((r2 * r2) * (p3 - p1))) / den;
    const double a2 = (((r3 * (p2 - p1)) + (r2 * (p1 - p3))) + (r1 * (p3 - p2))) / den;
    return (a0 + (a1 * rs)) + (a2 * (rs * rs));
}

// Pre-collision moments at the stencil nodes of the curved wall (the reference's moments
// sweep at those nodes, 2d_mrlbm.f90:231-245; same arithmetic as the bulk, so identical).
// Bilinear weights of the 8 interpolation corners of boundary node k (reference points
// 2 and 3), folded with the T8 extrapolation u = t1*0 + t2*u2 - t3*u3 (2d_mrlbm.f90:783-815).
template <class T> __device__ __forceinline__ T interp_weight(const double *R, int c)
{
    const int r = c >> 2, q = c & 3;
    const T xd = (T)R[BT_XD2 + 2 * r], yd = (T)R[BT_YD2 + 2 * r];
    const T wx = (q & 1) ? xd : (T)1.0 - xd, wy = (q & 2) ? yd : (T)1.0 - yd;
    return r == 0 ? (T)R[BT_T2] * (wy * wx) : -((T)R[BT_T3] * (wy * wx));
}

// Velocity at the boundary nodes from the stencil nodes' pre-collision moments U, level by
// level so that a node reads the already-interpolated velocity of an earlier node, as the
// in-place loop does (2d_mrlbm.f90:773-815).  One workgroup; writes b.ub.
// Quadratic extrapolation to the wall from bilinear values at the second and third
// reference points of boundary node R (2d_mrlbm.f90:783-815): c[0..3] the corners around
// point 2, c[4..7] around point 3.
template <class T> __device__ __forceinline__ T wall_interp(const double *R, const T c[8])
{
    const T xd2 = (T)R[BT_XD2], yd2 = (T)R[BT_YD2], xd3 = (T)R[BT_XD3], yd3 = (T)R[BT_YD3];
    const T t1 = (T)R[BT_T1], t2 = (T)R[BT_T2], t3 = (T)R[BT_T3];
    return (t1 * (T)0.0 + t2 * bilin(xd2, yd2, c[0], c[1], c[2], c[3])) - t3 * bilin(xd3, yd3, c[4], c[5], c[6], c[7]);
}

// Velocity of boundary node k when no corner is another boundary node (one dependency
// level): interpolated from the stencil nodes' u, or the node's own u without velocity
// interpolation in the rotated BC (2d_mrlbm.f90:768-819).
template <class T> __device__ __forceinline__ void wall_velocity_one(const BoundaryDev &b, const T *U, int k, T &vx, T &vy)
{
    const int nU = b.nU;
    if (b.vel_interp || !b.rotated) {
        const double *R = b.real + (long long)k * BT_NREAL;
        T cx[8], cy[8];
        for (int c = 0; c < 8; c++) {
            const int u = b.uc[12 * k + c];
            cx[c] = U[nU + u];
            cy[c] = U[2 * nU + u];
        }
        vx = wall_interp(R, cx);
        vy = wall_interp(R, cy);
    } else {
        vx = U[nU + b.uself[k]];
        vy = U[2 * nU + b.uself[k]];
    }
}

template <class S, class T> __device__ void wall_velocities(const StepArgs<S, T> &a)
{
    const BoundaryDev &b = a.b;
    T *U = (T *)b.U, *ub = (T *)b.ub;
    const int tid = threadIdx.x, nt = blockDim.x, nU = b.nU;
    const bool interp_u = b.vel_interp || !b.rotated;
    for (int L = 0; L < b.nlevels; L++) {
        for (int t = b.level_start[L] + tid; t < b.level_start[L + 1]; t += nt) {
            const int k = b.level_order[t];
            T vx, vy;
            if (interp_u) {
                const double *R = b.real + (long long)k * BT_NREAL;
                T cx[8], cy[8];
                for (int c = 0; c < 8; c++) {
                    const int dp = b.dep[8 * k + c], u = b.uc[12 * k + c];
                    cx[c] = dp >= 0 ? ub[2 * dp] : U[nU + u];
                    cy[c] = dp >= 0 ? ub[2 * dp + 1] : U[2 * nU + u];
                }
                vx = wall_interp(R, cx);
                vy = wall_interp(R, cy);
            } else {
                vx = U[nU + b.uself[k]];
                vy = U[2 * nU + b.uself[k]];
            }
            ub[2 * k] = vx;
            ub[2 * k + 1] = vy;
        }
        __syncthreads();
    }
}

// Interpolated rotated moments Mxx', Myy' of boundary node k (2d_mrlbm.f90:821-928); zero
// when moment interpolation is off (the reference then leaves mxxp unset).
template <class T> __device__ __forceinline__ void wall_moments(const BoundaryDev &b, const T *U, int k, T &Mxx, T &Myy)
{
    Mxx = (T)0.0;
    Myy = (T)0.0;
    if (!b.mom_interp) return;
    const int nU = b.nU;
    const double *R = b.real + (long long)k * BT_NREAL;
    T cx[8], cy[8];
    for (int c = 0; c < 8; c++) {
        const int u = b.uc[12 * k + c];
        cx[c] = U[3 * nU + u];
        cy[c] = U[4 * nU + u];
    }
    Mxx = wall_interp(R, cx);
    Myy = wall_interp(R, cy);
}

} // namespace mrlbm
