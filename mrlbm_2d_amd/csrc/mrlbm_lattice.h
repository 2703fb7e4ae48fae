// mrlbm_lattice.h -- D2Q9 lattice arithmetic shared by the bulk kernel and the boundary
// workgroup.  The FP64 functions reproduce the reference's operand order exactly (the
// FP64 path is compiled with -ffp-contract=off), so a step is bit-identical to
// 2d_mrlbm.f90 built without FMA.  The FP32 form stores and computes deviations from
// the rest state (rho-1, and the Hermite moments, which are already zero at rest).
#pragma once
#include <hip/hip_runtime.h>

namespace mrlbm {

constexpr int Q = 9;
// Velocity set in the reference's order (2d_mrlbm.f90:130-131).
__host__ __device__ constexpr int EXk(int k) { return k == 1 || k == 5 || k == 8 ? 1 : (k == 3 || k == 6 || k == 7 ? -1 : 0); }
__host__ __device__ constexpr int EYk(int k) { return k == 2 || k == 5 || k == 6 ? 1 : (k == 4 || k == 7 || k == 8 ? -1 : 0); }
__host__ __device__ constexpr int OPPk(int k) { return k == 0 ? 0 : (k <= 4 ? ((k + 1) % 4) + 1 : ((k - 3) % 4) + 5); }

// The weights are REAL(4) literals widened to real(8) (2d_mrlbm.f90:129): exactly
// c * (4/9, 1/9, 1/36) with c = 1 + 2^-27, so sum(w) = 1 + 2^-27.
constexpr float W0F = 16.0f / 36.0f, W1F = 4.0f / 36.0f, W5F = 1.0f / 36.0f;
constexpr float DEV_MASS_F = 7.450580596923828125e-9f; // sum(w) - 1 = 2^-27

template <class T> struct Lat;
template <> struct Lat<double> {
    static constexpr double w0 = (double)W0F, w1 = (double)W1F, w5 = (double)W5F;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double HA = 0.0 - cs2; // c_a^2 - cs2 for c_a = 0 (2d_mrlbm.f90:172)
    static constexpr double HB = 1.0 - cs2; // for c_a = +-1
};
template <> struct Lat<float> {
    static constexpr float w0 = W0F, w1 = W1F, w5 = W5F;
    static constexpr float cs2 = 1.0f / 3.0f;
    static constexpr float HA = 0.0f - cs2;
    static constexpr float HB = 1.0f - cs2;
};

template <class T> __host__ __device__ inline T wk(int k) { return k == 0 ? Lat<T>::w0 : (k <= 4 ? Lat<T>::w1 : Lat<T>::w5); }
template <class T> __host__ __device__ inline T Hxxk(int k) { return EXk(k) != 0 ? Lat<T>::HB : Lat<T>::HA; }
template <class T> __host__ __device__ inline T Hyyk(int k) { return EYk(k) != 0 ? Lat<T>::HB : Lat<T>::HA; }
template <class T> __host__ __device__ inline T Hxyk(int k) { return (T)(EXk(k) * EYk(k)); }

// ---------------------------------------------------------------------------------
// FP64, reference-exact.
// Regularized reconstruction fout_k = (w_k rho)(1 + 3 c.u + 4.5 mxx Hxx + 4.5 myy Hyy
// + 9 mxy Hxy) evaluated as 2d_mrlbm.f90:450-453 does, term by term; zero Hermite
// factors only add signed zeros, which leave the sums unchanged.
__device__ __forceinline__ void recon9_f64(double r, double ux, double uy, double mxx, double myy,
                                           double mxy, double f[Q])
{
    constexpr double HA = Lat<double>::HA, HB = Lat<double>::HB;
    const double X = 4.50 * mxx, Y = 4.50 * myy, Z = 9.00 * mxy;
    const double XA = X * HA, XB = X * HB, YA = Y * HA, YB = Y * HB;
    const double c1 = 3.0 * ux, c2 = 3.0 * uy;
    const double c5 = 3.0 * (ux + uy), c6 = 3.0 * (uy - ux), c8 = 3.0 * (ux - uy);
    const double w0r = Lat<double>::w0 * r, w1r = Lat<double>::w1 * r, w5r = Lat<double>::w5 * r;
    f[0] = w0r * ((1.0 + XA) + YA);
    f[1] = w1r * (((1.0 + c1) + XB) + YA);
    f[2] = w1r * (((1.0 + c2) + XA) + YB);
    f[3] = w1r * (((1.0 - c1) + XB) + YA);
    f[4] = w1r * (((1.0 - c2) + XA) + YB);
    f[5] = w5r * ((((1.0 + c5) + XB) + YB) + Z);
    f[6] = w5r * ((((1.0 + c6) + XB) + YB) - Z);
    f[7] = w5r * ((((1.0 - c5) + XB) + YB) + Z);
    f[8] = w5r * ((((1.0 + c8) + XB) + YB) - Z);
}

// Moments of the streamed populations, sequential sums over k = 0..8 (2d_mrlbm.f90:235-242).
__device__ __forceinline__ void moments_f64(const double f[Q], double &r, double &ux, double &uy,
                                            double &mxx, double &myy, double &mxy)
{
    constexpr double HA = Lat<double>::HA, HB = Lat<double>::HB;
    r = ((((((((f[0] + f[1]) + f[2]) + f[3]) + f[4]) + f[5]) + f[6]) + f[7]) + f[8]);
    const double sx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
    const double sy = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
    const double sxx = ((((((((f[0] * HA) + f[1] * HB) + f[2] * HA) + f[3] * HB) + f[4] * HA) + f[5] * HB)
                         + f[6] * HB) + f[7] * HB) + f[8] * HB;
    const double syy = ((((((((f[0] * HA) + f[1] * HA) + f[2] * HB) + f[3] * HA) + f[4] * HB) + f[5] * HB)
                         + f[6] * HB) + f[7] * HB) + f[8] * HB;
    const double sxy = ((f[5] - f[6]) + f[7]) - f[8];
    const double inv = 1.0 / r;
    ux = sx * inv;
    uy = sy * inv;
    mxx = sxx * inv;
    myy = syy * inv;
    mxy = sxy * inv;
}

// ---------------------------------------------------------------------------------
// FP32 deviation form.  e_k = f_k - w_k = w_k (d + (1+d) g_k), d = rho - 1, g_k the
// bracket of the reconstruction minus one.  Moments come from e_k alone, because
// sum_k w_k c_k = 0 and sum_k w_k H_k = 0; sum_k w_k = 1 + 2^-27 adds DEV_MASS_F to d.
__device__ __forceinline__ void recon9_dev(float d, float ux, float uy, float mxx, float myy, float mxy,
                                           float e[Q])
{
    // e_k = w_k d + r w_k T_k, r = 1 + d, T_k = 3 c_k.u + 4.5 H_k:m.  Opposite directions
    // (1,3), (2,4), (5,7), (6,8) share the even part w_k (d + r 4.5 H_k:m) and differ in the
    // sign of the odd part 3 w_k r c_k.u, so each pair costs one even and one odd term; the
    // weights are folded into the constants (35 operations for the nine values).
    constexpr float w0 = Lat<float>::w0, w1 = Lat<float>::w1, w5 = Lat<float>::w5;
    constexpr float KA = 4.5f * Lat<float>::HA, KB = 4.5f * Lat<float>::HB;
    const float r = 1.0f + d, t = mxx + myy;
    const float w1d = w1 * d, w5d = w5 * d;
    e[0] = __builtin_fmaf(r, (w0 * KA) * t, w0 * d);
    const float o1 = (3.0f * w1) * r, o5 = (3.0f * w5) * r;
    const float s13 = __builtin_fmaf(r, __builtin_fmaf(w1 * KB, mxx, (w1 * KA) * myy), w1d);
    const float s24 = __builtin_fmaf(r, __builtin_fmaf(w1 * KA, mxx, (w1 * KB) * myy), w1d);
    const float v1 = o1 * ux, v2 = o1 * uy;
    e[1] = s13 + v1;
    e[3] = s13 - v1;
    e[2] = s24 + v2;
    e[4] = s24 - v2;
    const float b = (w5 * KB) * t, z = (9.0f * w5) * mxy;
    const float s57 = __builtin_fmaf(r, b + z, w5d), s68 = __builtin_fmaf(r, b - z, w5d);
    const float v5 = o5 * (ux + uy), v6 = o5 * (uy - ux);
    e[5] = s57 + v5;
    e[7] = s57 - v5;
    e[6] = s68 + v6;
    e[8] = s68 - v6;
}

__device__ __forceinline__ void moments_dev(const float e[Q], float &d, float &ux, float &uy, float &mxx,
                                            float &myy, float &mxy)
{
    constexpr float HA = Lat<float>::HA;
    const float a13 = e[1] + e[3], a24 = e[2] + e[4], a57 = e[5] + e[7], a68 = e[6] + e[8];
    const float diag = a57 + a68;
    const float s0 = (e[0] + (a13 + a24)) + diag;
    d = s0 + DEV_MASS_F;
    const float d57 = e[5] - e[7], d68 = e[6] - e[8]; // e8 - e6 == -d68 exactly
    const float sx = (e[1] - e[3]) + (d57 - d68);
    const float sy = (e[2] - e[4]) + (d57 + d68);
    // sum_k e_k (c_ka^2 - cs2) = (sum of e_k with c_ka != 0) - cs2 sum_k e_k
    const float sxx = __builtin_fmaf(HA, s0, a13 + diag);
    const float syy = __builtin_fmaf(HA, s0, a24 + diag);
    const float sxy = a57 - a68;
    const float inv = __builtin_amdgcn_rcpf(1.0f + d); // 1 ulp reciprocal
    ux = sx * inv;
    uy = sy * inv;
    mxx = sxx * inv;
    myy = syy * inv;
    mxy = sxy * inv;
}

// ---------------------------------------------------------------------------------
// Precision forms.  S is the stored type; "pop" values travelling between nodes are
// fout_k (FP64) or e_k = fout_k - w_k (FP32).  Full populations for the boundary
// formulas are pop + base(k).
struct FormF64 {
    using S = double;
    using T = double;
    static constexpr bool dev = false;
    __device__ static void recon(const S m[6], T f[Q]) { recon9_f64(m[0], m[1], m[2], m[3], m[4], m[5], f); }
    __device__ static void moments(const T f[Q], T m[6]) { moments_f64(f, m[0], m[1], m[2], m[3], m[4], m[5]); }
    __host__ __device__ static T base(int) { return 0.0; }
    __host__ __device__ static T rho_of(T s0) { return s0; }
    __host__ __device__ static T store_rho(T rho) { return rho; }
};
struct FormF32 {
    using S = float;
    using T = float;
    static constexpr bool dev = true;
    __device__ static void recon(const S m[6], T f[Q]) { recon9_dev(m[0], m[1], m[2], m[3], m[4], m[5], f); }
    __device__ static void moments(const T f[Q], T m[6]) { moments_dev(f, m[0], m[1], m[2], m[3], m[4], m[5]); }
    __host__ __device__ static T base(int k) { return wk<float>(k); }
    __host__ __device__ static T rho_of(T s0) { return 1.0f + s0; }
    __host__ __device__ static T store_rho(T rho) { return rho - 1.0f; }
};

// BGK relaxation of the Hermite moments (2d_mrlbm.f90:410-412), then the optional
// incompressible trace projection whose result feeds the reconstruction (:419-428).
template <class T>
__device__ __forceinline__ void collide(T m[6], T om, T om1, int incomp_bulk);

// FP32 form: the same relaxation with fused multiply-adds.
template <>
__device__ __forceinline__ void collide<float>(float m[6], float om, float om1, int incomp_bulk)
{
    const float ux = m[1], uy = m[2];
    m[3] = __builtin_fmaf(om * ux, ux, om1 * m[3]);
    m[4] = __builtin_fmaf(om * uy, uy, om1 * m[4]);
    m[5] = __builtin_fmaf(om * ux, uy, om1 * m[5]);
    if (incomp_bulk) {
        const float tr = 0.5f * (((m[3] + m[4]) - ux * ux) - uy * uy);
        m[3] = m[3] - tr;
        m[4] = m[4] - tr;
    }
}

template <class T>
__device__ __forceinline__ void collide(T m[6], T om, T om1, int incomp_bulk)
{
    const T ux = m[1], uy = m[2];
    m[3] = om1 * m[3] + (om * ux) * ux;
    m[4] = om1 * m[4] + (om * uy) * uy;
    m[5] = om1 * m[5] + (om * ux) * uy;
    if (incomp_bulk) {
        const T tr = (T)0.50 * (((m[3] + m[4]) - ux * ux) - uy * uy);
        m[3] = m[3] - tr;
        m[4] = m[4] - tr;
    }
}

// Regularized edge BCs (apply_bc 2d_mrlbm.f90:1237-1366 and apply_bc_incomp :1368-1494).
// side 0 outlet (i = nx), 1 inlet (i = 1), 2 top (j = ny), 3 bottom (j = 1).  f holds the
// full streamed populations; m[0] (rho), m[1], m[2] hold the Dirichlet values.
__host__ __device__ constexpr int edge_in(int side, int k)
{
    // incoming sets Is (2d_mrlbm.f90:1254, 1283, 1311, 1339)
    return side == 0 ? (k != 3 && k != 6 && k != 7)
         : side == 1 ? (k != 1 && k != 5 && k != 8)
         : side == 2 ? (k != 4 && k != 7 && k != 8)
                     : (k != 2 && k != 5 && k != 6);
}

template <class T>
__device__ __forceinline__ void edge_bc(int side, int incomp, const T f[Q], T m[6], T om)
{
    T rhoI = 0, mxxI = 0, myyI = 0, mxyI = 0;
#pragma unroll
    for (int k = 0; k < Q; k++)
        if (edge_in(side, k)) {
            rhoI = rhoI + f[k];
            mxxI = mxxI + f[k] * Hxxk<T>(k);
            myyI = myyI + f[k] * Hyyk<T>(k);
            mxyI = mxyI + f[k] * Hxyk<T>(k);
        }
    const T inv = (T)1.0 / rhoI;
    mxxI = mxxI * inv;
    myyI = myyI * inv;
    mxyI = mxyI * inv;
    const T ux = m[1], uy = m[2];
    T rp;
    if (!incomp) {
        if (side == 0) {
            rp = m[0];
            m[3] = ((rp + ((T)9.0 * rhoI) * mxxI) - ((T)3.0 * rp) * ux) / ((T)6.0 * rp);
            m[4] = (((T)6.0 * rhoI) * myyI) / ((T)5.0 * rp);
            m[5] = (((T)6.0 * rhoI) * mxyI - rp * uy) / ((T)3.0 * rp);
        } else if (side == 1) {
            m[0] = (((T)4.0 * rhoI) + ((T)3.0 * rhoI) * mxxI) / ((T)3.0 - (T)3.0 * ux);
            rp = m[0];
            m[3] = ((rp + ((T)9.0 * rhoI) * mxxI) + ((T)3.0 * rp) * ux) / ((T)6.0 * rp);
            m[4] = (((T)6.0 * rhoI) * myyI) / ((T)5.0 * rp);
            m[5] = (((T)2.0 * rhoI) * mxyI) / rp;
        } else {
            m[0] = (((T)3.0 * rhoI) * ((T)4.0 + ((T)3.0 * ((T)1.0 - om)) * myyI)) / ((T)9.0 + om);
            rp = m[0];
            m[3] = (((T)6.0 * rhoI) * mxxI) / ((T)5.0 * rp);
            m[4] = (rp + ((T)9.0 * rhoI) * myyI) / ((T)6.0 * rp);
            m[5] = (((T)2.0 * rhoI) * mxyI) / rp;
        }
    } else {
        if (side == 0) {
            rp = m[0];
            m[3] = ux * ux;
            m[4] = uy * uy;
            m[5] = (((T)6.0 * rhoI) * mxyI - rp * uy) / ((T)3.0 * rp);
        } else if (side == 1) {
            m[0] = ((T)6.0 * rhoI) / (((T)5.0 - (T)3.0 * ux) - (T)3.0 * (ux * ux));
            rp = m[0];
            m[3] = ux * ux;
            m[4] = (T)0.0;
            m[5] = (((T)2.0 * rhoI) * mxyI) / rp;
        } else {
            m[0] = ((T)6.0 * rhoI) / (T)5.0;
            m[3] = (T)0.0;
            m[4] = (T)0.0;
            m[5] = ((T)5.0 * mxyI) / (T)3.0;
        }
    }
}

} // namespace mrlbm
