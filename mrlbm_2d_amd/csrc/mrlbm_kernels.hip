// mrlbm_kernels.hip -- gfx950 kernels of the D2Q9 moments-based regularized LBM step.
//
// One time step = the bulk kernel on the main stream and, concurrently on a side stream,
// the curved-wall kernels: stencil_moments_kernel recomputes the pre-collision moments
// the wall needs from the previous state, boundary_kernel applies the wall BC, collides
// and writes the boundary nodes, and reduces the momentum-exchange forces and the
// surface-pressure feed.  Both read only the previous state and write disjoint nodes.
// In the bulk each wave marches up a strip of rows over 64 columns (62 written, one halo
// column per side), reconstructs the 9 outgoing populations of its own node from the 6
// stored moments, exchanges the x-shifted ones with cross-lane shuffles, keeps the
// y-shifted ones in registers from row to row, and writes the 6 new moments.  HBM
// traffic is 6 words read + 6 words written per node; populations never touch memory.
// Edge nodes (inlet, outlet, channel walls) are skipped by the bulk and finished by
// edge_kernel on the side stream.
#include "mrlbm_device.h"

#include <algorithm>
#include <type_traits>

#include <hip/hip_ext.h>

namespace mrlbm {

// Workgroups are dealt to the 8 XCDs round-robin (block b -> XCD b % 8), each XCD with
// its own L2.  With xcd_order the blocks an XCD receives are consecutive strips of one
// column of waves, so the halo rows shared by vertically adjacent strips are read once
// into that XCD's L2.  Returns (column block, strip).
template <class S, class T> __device__ __forceinline__ int2 bulk_tile(const StepArgs<S, T> &a, int bb)
{
    if (!a.xcd_order) return make_int2(bb % a.ncb, bb / a.ncb);
    const int nb = a.ncb * a.nstrips;
    const int x = bb & 7, k = bb >> 3;
    const int L = x * (nb >> 3) + min(x, nb & 7) + k;
    return make_int2(L / a.nstrips, L % a.nstrips);
}

// Row bits of the strip for this lane's column: boundary nodes (skipped by the bulk) and
// solid nodes (u zeroed on plot steps).  Loaded once per strip, so the row loop touches
// no mask memory.
template <class S, class T>
__device__ __forceinline__ void strip_masks(const StepArgs<S, T> &a, int c, int st, unsigned long long &bm,
                                            unsigned long long &sm)
{
    bm = 0;
    sm = 0;
    const int bi = c - a.bx0;
    if ((unsigned)bi < (unsigned)a.bw) {
        const unsigned long long *p = a.boxbits + ((long long)bi * a.nstrips + st) * 2;
        bm = p[0];
        sm = p[1];
    }
}
__device__ __forceinline__ int row_cls(unsigned long long bm, unsigned long long sm, int t)
{
    return (int)((bm >> t) & 1ull) | ((int)((sm >> t) & 1ull) << 1);
}

// Whole-wave shifts by one lane, as DPP moves in the VALU (wave_shr:1 / wave_shl:1) rather
// than LDS permutes: lane i receives lane i-1 (shup) or lane i+1 (shdn); lane 0 (shup) and
// lane 63 (shdn) receive 0 (bound_ctrl: a lane without a source reads 0, so no register
// has to be cleared first), and the marches never use those lanes' shifted values.
template <int CTRL> __device__ __forceinline__ int dpp_shift(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <class T, int CTRL> __device__ __forceinline__ T dpp_shift_t(T v)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_shift<CTRL>(__builtin_bit_cast(int, v)));
    } else {
        const long long x = __builtin_bit_cast(long long, v);
        const int lo = dpp_shift<CTRL>((int)(x & 0xffffffffll)), hi = dpp_shift<CTRL>((int)(x >> 32));
        return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(unsigned)lo);
    }
}
// Raw buffer access for the marches: one resource per plane and wave (scalar registers), a
// 32-bit byte offset per lane and one per row (scalar), so a row load or store needs no
// vector address arithmetic.  Word 3 is the gfx9 raw-buffer format (32-bit data); the
// record count is the largest, so no access is dropped by the range check.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, -1, 0x00020000);
}
template <class S> __device__ __forceinline__ S buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(S) == 4) return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <class S> __device__ __forceinline__ void buf_store(S v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(S) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
    }
}
template <class T> __device__ __forceinline__ T shup(T v) { return dpp_shift_t<T, 0x138>(v); }
template <class T> __device__ __forceinline__ T shdn(T v) { return dpp_shift_t<T, 0x130>(v); }

// One row of the bulk march for the wave's 64 columns: rebuild the populations of the
// row above (moments in mN), assemble the 9 pulled populations of row j, finish the
// interior nodes of row j, and roll the register window up by one row.  PRE: the caller
// loaded the node's four u-statistics accumulators (acc_pre) a row ahead, with the state's
// row; loaded here, after the row's stores, the wave would wait for those stores and for
// every load in flight before it could add the sample.
template <class S, class T, bool REC, bool PRE = false>
__device__ __forceinline__ void march_row(const StepArgs<S, T> &a, int c, bool valid, int j, const S mN[6],
                                          T up_prev[3], T up_cur[3], T mid_cur[3], int cls,
                                          const double *acc_pre = nullptr)
{
    using F = typename FormOf<T>::F;
    T F9[Q];
    F::recon(mN, F9);
    T f[Q];
    f[0] = mid_cur[0];
    f[1] = shup(mid_cur[1]);
    f[3] = shdn(mid_cur[2]);
    f[2] = up_prev[0];
    f[5] = shup(up_prev[1]);
    f[6] = shdn(up_prev[2]);
    f[4] = F9[4];
    f[7] = shdn(F9[7]);
    f[8] = shup(F9[8]);
    if (valid) {
        if (!(cls & 1)) { // boundary nodes are written by the wall kernel
            T m[6];
            F::moments(f, m);
            if (a.zero_solid && (cls & 2)) { // plot steps zero u on solid nodes (2d_mrlbm.f90:375-384)
                m[1] = (T)0.0;
                m[2] = (T)0.0;
            }
            const long long n = (long long)j * a.pitch + c;
            if (REC && records_node(a, c, j)) {
                if (a.rec_slot >= 0)
                    for (int p = 0; p < a.nprobe; p++)
                        if (n == a.probe_node[p]) {
                            double *r = a.rec + (long long)a.rec_slot * a.rec_stride + 2 + 3 * p;
                            r[0] = F::dev ? 1.0 + (double)m[0] : (double)m[0];
                            r[1] = m[1];
                            r[2] = m[2];
                        }
                if (a.sample) {
                    const double ux = m[1], uy = m[2];
                    if constexpr (PRE) {
                        a.acc[n] = acc_pre[0] + ux;
                        a.acc[a.plane + n] = acc_pre[1] + uy;
                        a.acc[2 * a.plane + n] = acc_pre[2] + ux * ux;
                        a.acc[3 * a.plane + n] = acc_pre[3] + uy * uy;
                    } else {
                        a.acc[n] += ux;
                        a.acc[a.plane + n] += uy;
                        a.acc[2 * a.plane + n] += ux * ux;
                        a.acc[3 * a.plane + n] += uy * uy;
                    }
                }
            }
            collide<T>(m, a.om, a.om1, a.incomp_bulk);
#pragma unroll
            for (int p = 0; p < 6; p++) a.dst[p * a.plane + n] = (S)m[p];
        }
    }
    up_prev[0] = up_cur[0]; up_prev[1] = up_cur[1]; up_prev[2] = up_cur[2];
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
}

// Bulk (interior nodes 1..nx-2 x 1..ny-2): one wave = 64 columns (62 written, one halo
// column per side) marching up a strip of rows.  The loop is unrolled by two with two
// load buffers, so the loads of row j+2 are in flight while row j is computed and no
// register copy ever waits on them.
template <class S, class T, bool REC> __device__ __forceinline__ void bulk_march(const StepArgs<S, T> &a, int bb)
{
    using F = typename FormOf<T>::F;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int2 tile = bulk_tile(a, bb);
    const int cbi = tile.x, st = tile.y;
    const int g = cbi * 4 + wave;
    if (g >= a.ncg) return;
    const int c = g * 62 - 1 + lane;
    const int cl = min(max(c, 0), a.nx - 1);
    const bool valid = lane >= 1 && lane <= 62 && c >= 1 && c <= a.nx - 2;
    const int j0 = 1 + st * a.strip_rows, j1 = min(a.ny - 1, j0 + a.strip_rows);
    S mA[6], mB[6];
    T F9[Q], up_prev[3], up_cur[3], mid_cur[3];
    load_moments(a, cl, j0 - 1, mA);
    F::recon(mA, F9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    load_moments(a, cl, j0, mA);
    F::recon(mA, F9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    load_moments(a, cl, j0 + 1, mA);
    unsigned long long bm, sm;
    strip_masks(a, c, st, bm, sm);
    for (int j = j0; j < j1; j += 2) {
        if (j + 2 <= j1) load_moments(a, cl, j + 2, mB);
        march_row<S, T, REC>(a, c, valid, j, mA, up_prev, up_cur, mid_cur, row_cls(bm, sm, j - j0));
        if (j + 1 >= j1) break;
        if (j + 3 <= j1) load_moments(a, cl, j + 3, mA);
        march_row<S, T, REC>(a, c, valid, j + 1, mB, up_prev, up_cur, mid_cur, row_cls(bm, sm, j + 1 - j0));
    }
}

// Same march with the strip height a compile-time constant and the row loop fully
// unrolled: every load of the strip is independent straight-line code, so the compiler
// keeps many rows in flight without loop-carried copies of loaded registers.
template <class S, class T, bool REC, int H> __device__ __forceinline__ void bulk_march_u(const StepArgs<S, T> &a, int bb)
{
    using F = typename FormOf<T>::F;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int2 tile = bulk_tile(a, bb);
    const int cbi = tile.x, st = tile.y;
    const int g = cbi * 4 + wave;
    if (g >= a.ncg) return;
    const int c = g * 62 - 1 + lane;
    const int cl = min(max(c, 0), a.nx - 1);
    const bool valid = lane >= 1 && lane <= 62 && c >= 1 && c <= a.nx - 2;
    const int j0 = 1 + st * H;
    S m[H + 2][6];
#pragma unroll
    for (int t = 0; t < H + 2; t++) load_moments(a, cl, min(j0 - 1 + t, a.ny - 1), m[t]);
    T F9[Q], up_prev[3], up_cur[3], mid_cur[3];
    F::recon(m[0], F9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    F::recon(m[1], F9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    unsigned long long bm, sm;
    strip_masks(a, c, st, bm, sm);
#pragma unroll
    for (int t = 0; t < H; t++)
        march_row<S, T, REC>(a, c, valid && j0 + t <= a.ny - 2, j0 + t, m[t + 2], up_prev, up_cur, mid_cur,
                             row_cls(bm, sm, t));
}

// Aligned march: a wave owns 64 aligned columns and writes all of them with full-wave
// stores (a partially written 64-column span costs ~25% of HBM bandwidth here).  The two
// halo columns (x0-1 and x0+64) are loaded by every lane at a per-lane address (the lane's
// own column for lanes 1..62, so those loads coalesce with the main ones), rebuilt, and
// selected in by lanes 0 and 63 in place of the cross-lane shuffle.
template <class S, class T, bool REC>
__device__ __forceinline__ void march_row_a(const StepArgs<S, T> &a, int c, bool valid, int j, const S mN[6],
                                            const S hN[6], bool left, bool right, T up_prev[3], T up_cur[3],
                                            T mid_cur[3], T hu_prev[1], T hu_cur[1], T hm_cur[1], int cls)
{
    using F = typename FormOf<T>::F;
    T F9[Q], H9[Q];
    F::recon(mN, F9);
    F::recon(hN, H9);
    // halo populations that move into the wave: left halo sends 1, 5, 8; right sends 3, 6, 7
    const T hm_next = left ? H9[1] : H9[3];
    const T hu_next = left ? H9[5] : H9[6];
    const T hd = left ? H9[8] : H9[7];
    T f[Q];
    f[0] = mid_cur[0];
    const T s1 = shup(mid_cur[1]), s3 = shdn(mid_cur[2]);
    const T s5 = shup(up_prev[1]), s6 = shdn(up_prev[2]);
    const T s7 = shdn(F9[7]), s8 = shup(F9[8]);
    f[1] = left ? hm_cur[0] : s1;
    f[3] = right ? hm_cur[0] : s3;
    f[2] = up_prev[0];
    f[5] = left ? hu_prev[0] : s5;
    f[6] = right ? hu_prev[0] : s6;
    f[4] = F9[4];
    f[7] = right ? hd : s7;
    f[8] = left ? hd : s8;
    if (valid && !(cls & 1)) {
        T m[6];
        F::moments(f, m);
        if (a.zero_solid && (cls & 2)) {
            m[1] = (T)0.0;
            m[2] = (T)0.0;
        }
        const long long n = (long long)j * a.pitch + c;
        if (REC) {
            if (a.rec_slot >= 0)
                for (int p = 0; p < a.nprobe; p++)
                    if (n == a.probe_node[p]) {
                        double *r = a.rec + (long long)a.rec_slot * a.rec_stride + 2 + 3 * p;
                        r[0] = F::dev ? 1.0 + (double)m[0] : (double)m[0];
                        r[1] = m[1];
                        r[2] = m[2];
                    }
            if (a.sample) {
                const double ux = m[1], uy = m[2];
                a.acc[n] += ux;
                a.acc[a.plane + n] += uy;
                a.acc[2 * a.plane + n] += ux * ux;
                a.acc[3 * a.plane + n] += uy * uy;
            }
        }
        collide<T>(m, a.om, a.om1, a.incomp_bulk);
#pragma unroll
        for (int p = 0; p < 6; p++) a.dst[p * a.plane + n] = (S)m[p];
    }
    up_prev[0] = up_cur[0]; up_prev[1] = up_cur[1]; up_prev[2] = up_cur[2];
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    hu_prev[0] = hu_cur[0];
    hu_cur[0] = hu_next;
    hm_cur[0] = hm_next;
}

template <class S, class T, bool REC, int H> __device__ __forceinline__ void bulk_march_a(const StepArgs<S, T> &a, int bb)
{
    using F = typename FormOf<T>::F;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int2 tile = bulk_tile(a, bb);
    const int cbi = tile.x, st = tile.y;
    const int g = cbi * 4 + wave;
    if (g >= a.ncg) return;
    const int x0 = g * 64, c = x0 + lane;          // c < pitch: the pitch is a multiple of 64
    const bool left = lane == 0, right = lane == 63;
    const int hc = left ? max(x0 - 1, 0) : (right ? min(x0 + 64, a.pitch - 1) : c);
    const bool valid = c >= 1 && c <= a.nx - 2;
    const int j0 = 1 + st * H;
    S m[H + 2][6], h[H + 2][6];
#pragma unroll
    for (int t = 0; t < H + 2; t++) {
        const int r = min(j0 - 1 + t, a.ny - 1);
        load_moments(a, c, r, m[t]);
        load_moments(a, hc, r, h[t]);
    }
    T F9[Q], H9[Q], up_prev[3], up_cur[3], mid_cur[3], hu_prev[1], hu_cur[1], hm_cur[1];
    F::recon(m[0], F9);
    F::recon(h[0], H9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    hu_prev[0] = left ? H9[5] : H9[6];
    F::recon(m[1], F9);
    F::recon(h[1], H9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    hu_cur[0] = left ? H9[5] : H9[6];
    hm_cur[0] = left ? H9[1] : H9[3];
    unsigned long long bm, sm;
    strip_masks(a, c, st, bm, sm);
#pragma unroll
    for (int t = 0; t < H; t++)
        march_row_a<S, T, REC>(a, c, valid && j0 + t <= a.ny - 2, j0 + t, m[t + 2], h[t + 2], left, right, up_prev,
                               up_cur, mid_cur, hu_prev, hu_cur, hm_cur, row_cls(bm, sm, t));
}

// ---------------------------------------------------------------------------------
// Two steps per pass (temporal blocking).  A wave owns 64 aligned columns x0..x0+63 and
// marches up a strip of H output rows: per row it advances t -> t+1 for its own columns,
// then t+1 -> t+2 for its own columns, with full-wave aligned stores.  What lanes 0 and 63
// need from beyond the wave is computed once per strip, before the march, one lane per
// value: pass 1 rebuilds the populations of the nodes in the three columns at each side
// (x0-2..x0, x0+63..x0+65) over the strip's rows of state t; pass 2 advances the halo nodes
// (x0-1, x0+64) of every t+1 row, gathering their pulled populations from pass 1.  The
// march then takes the three populations that cross into the wave from those lanes with
// one cross-lane read each.  Every node update is the single-step arithmetic, so two passes
// of this kernel give the bits of four steps.

// Output nodes of the two-step pass: two or more nodes from the outer ring and outside the
// cylinder box dilated by one (everything else is finished by the band sub-steps).
template <class S, class T> __device__ __forceinline__ bool two_step_node(const StepArgs<S, T> &a, int c, int j)
{
    if (c < 2 || c > a.nx - 3 || j < 2 || j > a.ny - 3) return false;
    return !(a.bw > 0 && c >= a.bx0 - 1 && c <= a.bx0 + a.bw && j >= a.by0 - 1 && j <= a.by0 + a.bh);
}

// Pulled populations of a row from the carried rows below (up_prev: k = 2, 5, 6), the row
// itself (mid: k = 0, 1, 3) and the row above (N9); x-shifts from the neighbour lanes, or in
// lanes 0 / 63 the values crossing in from outside the wave (in0, in1, in2: the row's own,
// the row below's and the row above's populations 1/3, 5/6, 8/7).
template <class T>
__device__ __forceinline__ void pull_row(bool left, bool right, const T up_prev[3], const T mid[3], const T N9[Q],
                                         T in0, T in1, T in2, T f[Q])
{
    const T s1 = shup(mid[1]), s3 = shdn(mid[2]), s5 = shup(up_prev[1]), s6 = shdn(up_prev[2]);
    const T s8 = shup(N9[8]), s7 = shdn(N9[7]);
    f[0] = mid[0];
    f[1] = left ? in0 : s1;
    f[3] = right ? in0 : s3;
    f[2] = up_prev[0];
    f[5] = left ? in1 : s5;
    f[6] = right ? in1 : s6;
    f[4] = N9[4];
    f[8] = left ? in2 : s8;
    f[7] = right ? in2 : s7;
}

// u statistics of one node (mean and rms accumulators, FP64), as in node_update.
template <class S, class T> __device__ __forceinline__ void sample_u(const StepArgs<S, T> &a, long long n, T u, T v)
{
    const double ux = u, uy = v;
    a.acc[n] += ux;
    a.acc[a.plane + n] += uy;
    a.acc[2 * a.plane + n] += ux * ux;
    a.acc[3 * a.plane + n] += uy * uy;
}

// Both samples of a pair at one node in one read-modify-write of the accumulators: the
// t+1 velocity (u1, v1) then the t+2 one (u2, v2) added in that order, the rounding of two
// single-step samples.
template <class S, class T>
__device__ __forceinline__ void sample_u_pair(const StepArgs<S, T> &a, long long n, T u1, T v1, T u2, T v2)
{
    const double x1 = u1, y1 = v1, x2 = u2, y2 = v2;
    double s0 = a.acc[n], s1 = a.acc[a.plane + n], s2 = a.acc[2 * a.plane + n], s3 = a.acc[3 * a.plane + n];
    s0 += x1;
    s1 += y1;
    s2 += x1 * x1;
    s3 += y1 * y1;
    s0 += x2;
    s1 += y2;
    s2 += x2 * x2;
    s3 += y2 * y2;
    a.acc[n] = s0;
    a.acc[a.plane + n] = s1;
    a.acc[2 * a.plane + n] = s2;
    a.acc[3 * a.plane + n] = s3;
}

// Whether sub-step 1 of a two-step pair writes node (c, j) itself: the outer ring, the band
// rectangles a.rect (sub-step 1's) and the wall's boundary nodes.  Columns past nx (row
// padding) count as such too: nobody samples them.
template <class S, class T> __device__ __forceinline__ bool side_node1(const StepArgs<S, T> &a, int c, int j)
{
    bool side = c <= 0 || c >= a.nx - 1 || j <= 0 || j >= a.ny - 1;
    for (int q = 0; q < a.nrect; q++)
        side = side || (c >= a.rect[q][0] && c < a.rect[q][0] + a.rect[q][2] && j >= a.rect[q][1] &&
                        j < a.rect[q][1] + a.rect[q][3]);
    const int bi = c - a.bx0, bj = j - a.by0;
    if ((unsigned)bi < (unsigned)a.bw && (unsigned)bj < (unsigned)a.bh) side = side || (a.boxcls[bj * a.bw + bi] & 1);
    return side;
}

// SAMPLE: also accumulate the u statistics: at t+1 (a.sample) of the strip's own rows where
// sub-step 1 does not write the node, at t+2 (a.sample2) of the nodes the pass writes (with
// t+1 sampled as well, only those that sub-step 1 does not write).  Where the pass takes
// both samples of a node, the t+1 velocity is carried one row in registers and both are
// added in one read-modify-write (sample_u_pair).
template <class S, class T, int H, bool SAMPLE> __device__ __forceinline__ void bulk_march2(const StepArgs<S, T> &a, int bb)
{
    using F = typename FormOf<T>::F;
    constexpr int NQ = H + 4;             // state-t rows j0-2 .. j0+H+1
    constexpr int R1 = (6 * NQ + 63) / 64; // pass-1 rounds of 64 lanes
    static_assert(R1 <= 2 && 2 * (H + 2) <= 64, "at most two pass-1 rounds, one pass-2 round");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int2 tile = bulk_tile(a, bb);
    const int cbi = tile.x, st = tile.y;
    const int g = cbi * 4 + wave;
    if (g >= a.ncg) return;
    const int x0 = g * 64, c = x0 + lane; // c < pitch: the pitch is a multiple of 64
    const bool left = lane == 0, right = lane == 63;
    const int side = right ? 1 : 0;
    const int j0 = 2 + st * H; // output rows j0 .. j0+H-1 at t+2
    // pass-1 node n: the halo columns first (n = side*NQ + q: column x0-1 / x0+64), then the
    // outer and inner columns (n = 2NQ + side*2NQ + (col/2)*NQ + q, col 0 or 2); node n is
    // rebuilt by lane n % 64 in round n / 64.  Its loads are issued before the strip's own,
    // so waiting for them does not wait for the strip (loads complete in issue order for the
    // wait counter).
    auto node_of = [&](int n, int &x, int &q) {
        const int s = n < 2 * NQ ? n / NQ : (n - 2 * NQ) / (2 * NQ);
        int col = 1;
        if (n < 2 * NQ) q = n - s * NQ;
        else {
            const int r = n - 2 * NQ - s * 2 * NQ;
            col = r < NQ ? 0 : 2;
            q = r < NQ ? r : r - NQ;
        }
        x = s ? x0 + 63 + col : x0 - 2 + col;
    };
    auto node_index = [&](int s, int col, int q) {
        return col == 1 ? s * NQ + q : 2 * NQ + s * 2 * NQ + (col == 0 ? 0 : NQ) + q;
    };
    S pm[R1][6];
#pragma unroll
    for (int r = 0; r < R1; r++) {
        int x, q;
        node_of(min(r * 64 + lane, 6 * NQ - 1), x, q);
        load_moments(a, min(max(x, 0), a.pitch - 1), min(max(j0 - 2 + q, 0), a.ny - 1), pm[r]);
    }
    // State-t rows of the strip: FP32 loads all of them here; FP64 loads the first 3 + LA and
    // then row p + 3 + LA at the top of march step p (measured ~3% faster in FP64, slower in
    // FP32, where the earlier loads keep more memory requests in flight per wave).
    constexpr int LA = sizeof(S) == 8 ? 3 : NQ;
    S m[NQ][6];
#pragma unroll
    for (int t = 0; t < NQ; t++)
        if (t < 3 + LA) load_moments(a, c, min(j0 - 2 + t, a.ny - 1), m[t]);
    T P9[R1][Q];
#pragma unroll
    for (int r = 0; r < R1; r++) F::recon(pm[r], P9[r]);
    // populations of the halo columns that move into the wave (round 0, lanes 0 .. 2NQ-1)
    const bool s0 = lane >= NQ;
    const T av0 = s0 ? P9[0][3] : P9[0][1], av1 = s0 ? P9[0][6] : P9[0][5], av2 = s0 ? P9[0][7] : P9[0][8];
    // pass 2: lane = 2p + side advances the halo node of t+1 row j0-1+p (t row p+1)
    T bv0, bv1, bv2;
    {
        const int s2 = lane & 1, p = lane >> 1;
        T f[Q], mh[6], K9[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) {
            const int n = node_index(s2, 1 - EXk(k), min(p + 1 - EYk(k), NQ - 1));
            const T v0 = __shfl(P9[0][k], n & 63, 64);
            if constexpr (R1 > 1) {
                const T v1 = __shfl(P9[R1 - 1][k], n & 63, 64);
                f[k] = n < 64 ? v0 : v1;
            } else {
                f[k] = v0;
            }
        }
        F::moments(f, mh);
        collide<T>(mh, a.om, a.om1, a.incomp_bulk);
        F::recon(mh, K9);
        bv0 = s2 ? K9[3] : K9[1];
        bv1 = s2 ? K9[6] : K9[5];
        bv2 = s2 ? K9[7] : K9[8];
    }
    auto lane_a = [&](int q) { return side * NQ + q; };          // pass-1 lane of halo column, t row q
    auto lane_b = [&](int p) { return 2 * p + side; };           // pass-2 lane of halo node, t+1 row p
    T F9[Q], up_prev[3], up_cur[3], mid_cur[3];
    F::recon(m[0], F9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    F::recon(m[1], F9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    T vup_prev[3], vup_cur[3], vmid_cur[3];
    // t+1 velocity of the node the next level-2 row writes, when both of its samples are
    // this pass's (SAMPLE only)
    bool pf = false;
    T pu = (T)0, pv = (T)0;
#pragma unroll
    for (int p = 0; p < H + 2; p++) {
        if (p + 3 + LA < NQ) load_moments(a, c, min(j0 - 2 + p + 3 + LA, a.ny - 1), m[p + 3 + LA]);
        // level 1: t+1 row j0-1+p from t rows p, p+1, p+2
        F::recon(m[p + 2], F9);
        const T in0 = __shfl(av0, lane_a(p + 1), 64), in1 = __shfl(av1, lane_a(p), 64), in2 = __shfl(av2, lane_a(p + 2), 64);
        T f[Q], m1[6], G9[Q];
        pull_row(left, right, up_prev, mid_cur, F9, in0, in1, in2, f);
        F::moments(f, m1);
        bool nf = false;
        T nu_ = (T)0, nv_ = (T)0;
        if constexpr (SAMPLE) {
            const int j1 = j0 - 1 + p;
            if (a.sample && p >= 1 && p <= H && !side_node1(a, c, j1)) {
                nf = a.sample2 && two_step_node(a, c, j1); // its t+2 sample comes next row
                if (nf) {
                    nu_ = m1[1];
                    nv_ = m1[2];
                } else {
                    sample_u(a, (long long)j1 * a.pitch + c, m1[1], m1[2]);
                }
            }
        }
        collide<T>(m1, a.om, a.om1, a.incomp_bulk);
        up_prev[0] = up_cur[0]; up_prev[1] = up_cur[1]; up_prev[2] = up_cur[2];
        up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
        mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
        F::recon(m1, G9);
        if (p >= 2) {
            // level 2: t+2 row j0+p-2 from t+1 rows p-2, p-1, p
            const T i0 = __shfl(bv0, lane_b(p - 1), 64), i1 = __shfl(bv1, lane_b(p - 2), 64), i2 = __shfl(bv2, lane_b(p), 64);
            T g2[Q], m2[6];
            pull_row(left, right, vup_prev, vmid_cur, G9, i0, i1, i2, g2);
            const int j = j0 + p - 2;
            if (two_step_node(a, c, j)) {
                F::moments(g2, m2);
                const long long n = (long long)j * a.pitch + c;
                // when t+1 is sampled too, sub-step 1's side kernel adds t+1 into the nodes
                // it writes concurrently: those are left to pair_samples_kernel
                if constexpr (SAMPLE) {
                    if (pf) sample_u_pair(a, n, pu, pv, m2[1], m2[2]);
                    else if (a.sample2 && !(a.sample && side_node1(a, c, j))) sample_u(a, n, m2[1], m2[2]);
                }
                collide<T>(m2, a.om, a.om1, a.incomp_bulk);
#pragma unroll
                for (int q = 0; q < 6; q++) a.dst[q * a.plane + n] = (S)m2[q];
            }
        }
        pf = nf;
        pu = nu_;
        pv = nv_;
        if (p >= 1) {
            vup_prev[0] = p == 1 ? vup_prev[0] : vup_cur[0];
            vup_prev[1] = p == 1 ? vup_prev[1] : vup_cur[1];
            vup_prev[2] = p == 1 ? vup_prev[2] : vup_cur[2];
            vup_cur[0] = G9[2]; vup_cur[1] = G9[5]; vup_cur[2] = G9[6];
            vmid_cur[0] = G9[0]; vmid_cur[1] = G9[1]; vmid_cur[2] = G9[3];
        } else {
            vup_prev[0] = G9[2]; vup_prev[1] = G9[5]; vup_prev[2] = G9[6];
        }
    }
}

template <class S, class T, int H, bool SAMPLE = false>
__global__ void __launch_bounds__(256) bulk2_kernel(StepArgs<S, T> a)
{
    bulk_march2<S, T, H, SAMPLE>(a, blockIdx.x);
}

// NS steps in one pass over the plain interior, src = state t, dst = state t+NS.  A wave holds
// 64 consecutive columns and writes the middle 64 - 2 NS of them: state t+s computed in all
// lanes is right in lanes s .. 63-s (each level loses one column per side to the lanes whose
// x-shifted inputs come from outside the wave), so waves share no data and overlap by 2 NS
// columns instead.  The wave marches up a strip of H output rows: per row of state t loaded
// (LA rows ahead), level s = 1 .. NS pulls its next row from the three rows of level s-1 it
// carries (k = 2, 5, 6 of the row below, 0, 1, 3 of the row itself, 4, 7, 8 of the row just
// made above) and collides it; levels below NS rebuild the row's populations for the level
// above, level NS stores its moments.  Level s makes the rows j0-NS+s .. j0+H+NS-1-s, the
// first of them at march step 2s.  Per node and level the arithmetic is the single step's, so
// the result is bit-identical to NS single steps.
// REC: also accumulate the u statistics (levels with a bit in a.qsample) and record the probes
// (levels with a.qslot >= 0) of the nodes the pass writes, at every level, as single steps
// would; the sub-steps of the pass leave those nodes to it (records_node).  A node's level-s
// velocity is made NS-s march steps before its level-NS one, so the wave carries it that
// long, then adds the sampled levels in order in one read-modify-write of the accumulators.
template <class S, class T, int NS, int H, bool REC> __device__ __forceinline__ void bulk_march_n(const StepArgs<S, T> &a, int bb)
{
    using F = typename FormOf<T>::F;
    // rows of state t loaded ahead: FP32 4, FP64 1 (FP64 at 2 spills at two waves per SIMD;
    // both measured fastest at 8192 x 2048)
    constexpr int W = 64 - 2 * NS, NP = H + 2 * NS, LA = sizeof(S) == 8 ? 1 : 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = (bb % a.ncb) * 4 + wave, st = bb / a.ncb;
    if (g >= a.ncg) return;
    const int c = g * W + lane, cl = min(c, a.nx - 1); // lanes past the last column load it again
    const bool outc = lane >= NS && lane < 64 - NS;
    const int j0 = NS + st * H; // output rows j0 .. j0+H-1; state t rows j0-NS .. j0+H+NS-1
    const S *__restrict__ src = a.src; // the pass writes another state than it reads
    S *__restrict__ dst = a.dst;
    const long long plane = a.plane;
    // Buffer resources per plane based at the strip's first state-t row (which is <= ny-1):
    // row offsets within the strip and lane offsets within the row fit 32 bits at any grid size.
    const long long r0 = (long long)(j0 - NS) * a.pitch;
    __amdgpu_buffer_rsrc_t rs[6], rd[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        rs[k] = plane_rsrc(src + k * plane + r0);
        rd[k] = plane_rsrc(dst + k * plane + r0);
    }
    const unsigned lb = (unsigned)cl * (unsigned)sizeof(S);
    const unsigned rowb = (unsigned)a.pitch * (unsigned)sizeof(S);
    auto load_row = [&](int p, S m[6]) {
        const unsigned ro = (unsigned)(min(j0 - NS + p, a.ny - 1) - (j0 - NS)) * rowb;
#pragma unroll
        for (int k = 0; k < 6; k++) m[k] = buf_load<S>(rs[k], lb, ro);
    };
    S q[LA + 1][6];
#pragma unroll
    for (int i = 0; i <= LA; i++) load_row(i, q[i]);
    T upp[NS + 1][3], upc[NS + 1][3], mid[NS + 1][3];
    // REC: velocities of levels 1 .. NS-1 in LDS, a ring of NS-s+1 entries per level s: per
    // wave, level s at ring entries base(s) = sum over l < s of (NS-l+1), 2 values of 64 lanes
    // each (in registers they would cost FP32 its third wave per SIMD once the accumulators
    // are loaded ahead, and FP64 its second).
    constexpr int DLN = NS * (NS + 1) / 2 - 1; // ring entries of levels 1 .. NS-1
    __shared__ T dl[REC ? 4 : 1][REC ? DLN : 1][2][64];
    auto dl_at = [](int l, int e) { return (l - 1) * (NS + 1) - (l - 1) * l / 2 + e; }; // base(l) + e
    // REC, FP32: the accumulators of the row the last level finishes in this march step are
    // loaded at the step's start, before the next state row, and used at its end: the wave
    // then waits for nothing younger than the previous step's stores.  Loaded after this
    // step's stores, the wait would also take in those stores and the rows loaded ahead.
    // FP64 has no registers to spare at two waves per SIMD and loads them at the end.
    constexpr bool ACC_AHEAD = REC && sizeof(S) == 4;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        T Fin[Q]; // populations of the newest row of the level below
        F::recon(q[0], Fin);
#pragma unroll
        for (int i = 0; i < LA; i++)
#pragma unroll
            for (int k = 0; k < 6; k++) q[i][k] = q[i + 1][k];
        double av[4] = {};
        if constexpr (ACC_AHEAD) {
            const int j = j0 + p - 2 * NS;
            if (p >= 2 * NS && a.qsample && outc && n_step_node<S, T, NS>(a, c, j)) {
                const long long n = (long long)j * a.pitch + c;
#pragma unroll
                for (int k = 0; k < 4; k++) av[k] = a.acc[k * plane + n];
            }
        }
        if (p + LA + 1 < NP) load_row(p + LA + 1, q[LA]);
#pragma unroll
        for (int s = 1; s <= NS; s++) {
            if (p < 2 * s - 2) break; // level s starts two march steps after level s-1
            T f[Q];
            const bool make = p >= 2 * s;
            if (make) {
                f[0] = mid[s][0];
                f[1] = shup(mid[s][1]);
                f[3] = shdn(mid[s][2]);
                f[2] = upp[s][0];
                f[5] = shup(upp[s][1]);
                f[6] = shdn(upp[s][2]);
                f[4] = Fin[4];
                f[7] = shdn(Fin[7]);
                f[8] = shup(Fin[8]);
            }
            upp[s][0] = upc[s][0]; upp[s][1] = upc[s][1]; upp[s][2] = upc[s][2];
            upc[s][0] = Fin[2]; upc[s][1] = Fin[5]; upc[s][2] = Fin[6];
            mid[s][0] = Fin[0]; mid[s][1] = Fin[1]; mid[s][2] = Fin[3];
            if (!make) break; // its window fills at steps 2s-2 and 2s-1
            T m[6];
            F::moments(f, m);
            if constexpr (REC) {
                if (s < NS) {
                    dl[wave][dl_at(s, p % (NS - s + 1))][0][lane] = m[1];
                    dl[wave][dl_at(s, p % (NS - s + 1))][1][lane] = m[2];
                }
            }
            collide<T>(m, a.om, a.om1, a.incomp_bulk);
            if (s < NS) {
                F::recon(m, Fin);
            } else {
                const int j = j0 + p - 2 * NS;
                if (outc && n_step_node<S, T, NS>(a, c, j)) {
                    const long long n = (long long)j * a.pitch + c;
                    const unsigned ro = (unsigned)(j - (j0 - NS)) * rowb;
#pragma unroll
                    for (int k = 0; k < 6; k++) buf_store<S>((S)m[k], rd[k], lb, ro);
                    if constexpr (REC) {
                        if (a.qsample) {
                            double s0, s1, s2, s3;
                            if constexpr (ACC_AHEAD) {
                                s0 = av[0]; s1 = av[1]; s2 = av[2]; s3 = av[3];
                            } else {
                                s0 = a.acc[n]; s1 = a.acc[plane + n]; s2 = a.acc[2 * plane + n];
                                s3 = a.acc[3 * plane + n];
                            }
#pragma unroll
                            for (int l = 1; l <= NS; l++) {
                                if (!((a.qsample >> (l - 1)) & 1)) continue;
                                // level l made this row NS - l march steps ago
                                const int e = (p - (NS - l)) % (NS - l + 1);
                                double x = m[1], y = m[2];
                                if (l < NS) {
                                    x = dl[wave][dl_at(l, e)][0][lane];
                                    y = dl[wave][dl_at(l, e)][1][lane];
                                }
                                s0 += x;
                                s1 += y;
                                s2 += x * x;
                                s3 += y * y;
                            }
                            a.acc[n] = s0;
                            a.acc[plane + n] = s1;
                            a.acc[2 * plane + n] = s2;
                            a.acc[3 * plane + n] = s3;
                        }
                    }
                }
            }
        }
    }
}

// at least two waves per SIMD (left alone, the FP64 kernel takes 256 VGPRs plus AGPRs: one)
template <class S, class T, int NS, int H, bool REC = false>
__global__ void __launch_bounds__(256, 2) bulkn_kernel(StepArgs<S, T> a)
{
    bulk_march_n<S, T, NS, H, REC>(a, blockIdx.x);
}

// The side stream's node work of one step in one launch: the outer ring (edge update),
// the band rectangles of a two-step pair (plain update) and the curved wall's stencil
// nodes (pre-collision moments for boundary_kernel).  All three read only the source
// state and write disjoint nodes.
// Whether node (i, j) lies in one of the band rectangles a.rect[0 .. q-1].
template <class S, class T> __device__ __forceinline__ bool in_rects(const StepArgs<S, T> &a, int q, int i, int j)
{
    bool in = false;
    for (int r = 0; r < q; r++)
        in = in || (i >= a.rect[r][0] && i < a.rect[r][0] + a.rect[r][2] && j >= a.rect[r][1] &&
                    j < a.rect[r][1] + a.rect[r][3]);
    return in;
}

template <class S, class T> __device__ void side_body(const StepArgs<S, T> &a, long long t, int nring, int nband)
{
    if (t < nring) {
        edge_body(a, (int)t);
        return;
    }
    t -= nring;
    if (t < nband) {
        for (int q = 0; q < a.nrect; q++) {
            if ((a.band_march >> q) & 1) continue; // marched (band_march)
            const long long n = (long long)a.rect[q][2] * a.rect[q][3];
            if (t < n) { // every rectangle lies in [1, nx-2] x [1, ny-2]: all 8 neighbours inside
                const int i = a.rect[q][0] + (int)(t % a.rect[q][2]), j = a.rect[q][1] + (int)(t / a.rect[q][2]);
                if (in_rects(a, q, i, j)) return; // an earlier rectangle has it (the box near an edge)
                T f[Q];
                gather_interior(a, i, j, f);
                node_update(a, i, j, f, false);
                return;
            }
            t -= n;
        }
        return;
    }
    t -= nband;
    if (a.has_boundary) stencil_moments_body(a, (int)t);
}

// Stand-ins read in place of the accumulators on steps that do not sample and of the node
// class table on grids without a cylinder (never written).
__device__ double acc_standin[4];
__device__ unsigned char cls_standin[1];

// The four u-statistics accumulators of node (c, j) (a marched row's, one row ahead of its
// update; c and j lie in the grid), on steps that do not sample the stand-ins.  The loads
// are issued either way: the compiler's waits count the loads issued after the one waited
// for, and a load issued on one path only leaves a count that holds on every path (none),
// so the wait for this row would take in the next row's loads too.
template <class S, class T> __device__ __forceinline__ void load_acc_row(const StepArgs<S, T> &a, int c, int j, double v[4])
{
    const bool smp = a.sample;
    const double *base = smp ? a.acc : acc_standin;
    const long long n = smp ? (long long)j * a.pitch + c : 0, stride = smp ? a.plane : 1;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = base[k * stride + n];
}

// Class of node (c, j) from the cylinder's box table (0 outside it), the table read at a
// clamped index whether or not the node lies in the box (see load_acc_row).
template <class S, class T> __device__ __forceinline__ int box_class(const StepArgs<S, T> &a, int c, int j)
{
    const int bi = c - a.bx0, bj = j - a.by0;
    const bool in = (unsigned)bi < (unsigned)a.bw && (unsigned)bj < (unsigned)a.bh;
    const unsigned char *t = a.boxcls ? a.boxcls : cls_standin;
    const int v = t[in ? bj * a.bw + bi : 0];
    return in ? v : 0;
}

template <class S, class T> __device__ __forceinline__ int band_nodes(const StepArgs<S, T> &a)
{
    return (int)band_items(a.rect, a.nrect, a.band_march);
}

// Marched band rectangle: workgroup blk of the marched rectangles (in rectangle order, each
// cut into 62-column groups x BAND_STRIP-row strips), one single wave: lanes 1..62 write
// columns x0+62g .. x0+62g+61 of rows y0 .. y1-1, lanes 0 and 63 carry the halo columns, and
// each row is the bulk march's single-step row update (march_row) from rows loaded one
// ahead.  Node classes come from the cylinder's box (boundary nodes are the wall's, solid
// nodes are zeroed on plot steps); nodes of an earlier rectangle are left to it.  A row's
// class byte and u-statistics accumulators are loaded with the moments a row ahead: the
// wait for a value loaded just before its use also waits for every load issued before it,
// the next row's among them.
template <class S, class T> __device__ void band_march_segments(const StepArgs<S, T> &a, int q, int blk);
template <class S, class T> __device__ void band_march(const StepArgs<S, T> &a, int blk)
{
    using F = typename FormOf<T>::F;
    int q = 0;
    for (; q < a.nrect; q++) { // the rectangle of this workgroup
        if (!((a.band_march >> q) & 1)) continue;
        const int nb = band_march_rect_blocks(a.rect[q], q);
        if (blk < nb) break;
        blk -= nb;
    }
    if (band_column_rect(q)) {
        band_march_segments(a, q, blk);
        return;
    }
    const int x0 = a.rect[q][0], w = a.rect[q][2], per = (w + 61) / 62;
    const int g = blk % per, st = blk / per;
    const int y0 = a.rect[q][1] + st * BAND_STRIP, y1 = min(y0 + BAND_STRIP, a.rect[q][1] + a.rect[q][3]);
    const int lane = threadIdx.x, c = x0 - 1 + 62 * g + lane, cl = min(c, a.nx - 1);
    const bool lanev = lane >= 1 && lane <= 62 && c < x0 + w; // rows y0-1 .. y1 lie in the grid
    auto cls_of = [&](int j) { return box_class(a, c, j); };
    auto valid = [&](int j) { return lanev && !in_rects(a, q, c, j); };
    S mA[6], mB[6];
    double vA[4] = {}, vB[4] = {}; // accumulators of the row that mA / mB finish (sampling steps)
    T F9[Q], up_prev[3], up_cur[3], mid_cur[3];
    load_moments(a, cl, y0 - 1, mA);
    F::recon(mA, F9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    load_moments(a, cl, y0, mA);
    F::recon(mA, F9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    load_moments(a, cl, y0 + 1, mA);
    load_acc_row(a, cl, y0, vA);
    int kA = cls_of(y0), kB;
    for (int j = y0; j < y1; j += 2) { // loads past row y1 read row y1 again (unused)
        load_moments(a, cl, min(j + 2, y1), mB);
        load_acc_row(a, cl, j + 1, vB);
        kB = cls_of(j + 1);
        march_row<S, T, true, true>(a, c, valid(j), j, mA, up_prev, up_cur, mid_cur, kA, vA);
        if (j + 1 >= y1) break;
        load_moments(a, cl, min(j + 3, y1), mA);
        load_acc_row(a, cl, min(j + 2, y1), vA);
        kA = cls_of(min(j + 2, y1));
        march_row<S, T, true, true>(a, c, valid(j + 1), j + 1, mB, up_prev, up_cur, mid_cur, kB, vB);
    }
}

// Marched narrow column rectangle q (at most 6 columns, next to the inlet or outlet): the
// wave's lanes in 8 segments of 8, segment s of workgroup blk taking rows y0 + (8 blk + s) R
// .. + R-1 (R = BAND_SEG_ROWS) of the rectangle; within a segment, lane i holds column
// x0 - 1 + i (lanes 1..w write, lanes 0 and w+1 carry the halo columns).  Each segment
// marches its rows with march_row from rows loaded one ahead, as band_march does.  The
// whole-wave lane shifts hand the halo lanes values of the neighbouring segment, which only
// reach the halo lanes' own (unwritten) nodes.  Node classes from the cylinder's box; nodes
// of an earlier rectangle are left to it.
template <class S, class T> __device__ void band_march_segments(const StepArgs<S, T> &a, int q, int blk)
{
    using F = typename FormOf<T>::F;
    constexpr int R = BAND_SEG_ROWS;
    const int x0 = a.rect[q][0], w = a.rect[q][2], yr = a.rect[q][1], h = a.rect[q][3];
    const int lane = threadIdx.x, seg = lane >> 3, i = lane & 7;
    const int c = x0 - 1 + i, cl = min(c, a.nx - 1); // lanes past the halo load a column again
    const int y0 = yr + (8 * blk + seg) * R, y1 = min(y0 + R, yr + h); // rows y0 .. y1-1
    const bool lanev = i >= 1 && i <= w;
    auto row = [&](int j) { return min(j, a.ny - 1); }; // segments past the rectangle end
    auto cls_of = [&](int j) { return box_class(a, c, j); };
    auto valid = [&](int j) { return lanev && j < y1 && !in_rects(a, q, c, j); };
    S mA[6], mB[6];
    double vA[4] = {}, vB[4] = {}; // accumulators of the row that mA / mB finish (sampling steps)
    T F9[Q], up_prev[3], up_cur[3], mid_cur[3];
    load_moments(a, cl, row(y0 - 1), mA);
    F::recon(mA, F9);
    up_prev[0] = F9[2]; up_prev[1] = F9[5]; up_prev[2] = F9[6];
    load_moments(a, cl, row(y0), mA);
    F::recon(mA, F9);
    up_cur[0] = F9[2]; up_cur[1] = F9[5]; up_cur[2] = F9[6];
    mid_cur[0] = F9[0]; mid_cur[1] = F9[1]; mid_cur[2] = F9[3];
    load_moments(a, cl, row(y0 + 1), mA);
    load_acc_row(a, cl, row(y0), vA);
    int kA = cls_of(y0), kB;
    for (int k = 0; k < R; k += 2) { // every segment marches R rows (uniform trip count)
        const int j = y0 + k;
        load_moments(a, cl, row(j + 2), mB);
        load_acc_row(a, cl, row(j + 1), vB);
        kB = cls_of(j + 1);
        march_row<S, T, true, true>(a, c, valid(j), j, mA, up_prev, up_cur, mid_cur, kA, vA);
        load_moments(a, cl, row(j + 3), mA); // the last one unused
        load_acc_row(a, cl, row(j + 2), vA);
        kA = cls_of(j + 2);
        march_row<S, T, true, true>(a, c, valid(j + 1), j + 1, mB, up_prev, up_cur, mid_cur, kB, vB);
    }
}

// The inlet and outlet columns of the outer ring (rows 1 .. ny-2) in tiles of 64 rows, one
// wave each.  Down a column every gather load touches a cache line per lane (the outlet
// gathers twice: its own populations and column nx-2's, which it copies), so the wave first
// stages the moments its gathers read in LDS: rows j0-1 .. j0+64 of columns 0..1 (inlet)
// or nx-3 .. nx-1 (outlet).  Sources outside the domain follow gather_edge (periodic wrap
// from memory, else the stale-population table); the arithmetic is edge_body's.
template <class S, class T> __device__ void ring_column_tile(const StepArgs<S, T> &a, int side, int j0)
{
    constexpr int R = 66;
    __shared__ S sm[6][R][3];
    const int lane = threadIdx.x;
    const int xb = side == 0 ? -1 : a.nx - 3; // column of slot 0 (the inlet's is outside)
    for (int idx = lane; idx < R * 3; idx += 64) {
        const int r = idx / 3, k = idx - 3 * r, x = xb + k;
        if (x < 0) continue;
        const int j = min(max(j0 - 1 + r, 0), a.ny - 1); // clamped rows are never read
        const long long n = (long long)j * a.pitch + x;
#pragma unroll
        for (int p = 0; p < 6; p++) sm[p][r][k] = a.src[p * a.plane + n];
    }
    __syncthreads();
    const int j = j0 + lane;
    if (j > a.ny - 2) return;
    auto gather = [&](int i, T f[Q]) { // gather_edge at (i, j), 1 <= j <= ny-2
#pragma unroll
        for (int k = 0; k < Q; k++) {
            const int si = i - EXk(k), sj = j - EYk(k);
            S m[6];
            if (si < 0 || si >= a.nx) {
                if (!a.xper) {
                    f[k] = a.ghost[(si < 0 ? j : a.ny + j) * Q + k];
                    continue;
                }
                load_moments(a, si < 0 ? a.nx - 1 : 0, j, m);
            } else {
#pragma unroll
                for (int p = 0; p < 6; p++) m[p] = sm[p][sj - (j0 - 1)][si - xb];
            }
            T g[Q];
            FormOf<T>::F::recon(m, g);
            f[k] = g[k];
        }
    };
    T f[Q];
    if (side == 0) {
        gather(0, f);
        node_update(a, 0, j, f, true);
    } else {
        T g[Q];
        gather(a.nx - 1, f);
        gather(a.nx - 2, g);
        node_update(a, a.nx - 1, j, f, true, g);
    }
}

template <class S, class T> __device__ void wall_group(const StepArgs<S, T> &a, int g);

// Side workgroups of one step (single waves): the wall groups (WG, a.wall_groups of them), the
// ring columns as tiles when the column path applies (ring_column_tiles), then 64 items each
// of the ring rows (or the whole ring), the band rectangles and (without wall groups) the
// stencil nodes.
template <class S, class T, bool WG = false> __device__ void side_block(const StepArgs<S, T> &a, int blk, int nband)
{
    if constexpr (WG) {
        if (blk < a.wall_groups) {
            wall_group(a, blk);
            return;
        }
        blk -= a.wall_groups;
    }
    if (a.band_march) {
        const int nbm = band_march_blocks(a.rect, a.nrect, a.band_march);
        if (blk < nbm) {
            band_march(a, blk);
            return;
        }
        blk -= nbm;
    }
    const int nct = ring_column_tiles(a.nx, a.ny, a.pops != nullptr);
    if (blk < nct) {
        ring_column_tile(a, blk & 1, 1 + (blk >> 1) * 64);
        return;
    }
    const int nring = nct ? 2 * a.nx : 2 * a.nx + 2 * (a.ny - 2);
    side_body(a, (long long)(blk - nct) * 64 + threadIdx.x, nring, nband);
}

template <class S, class T, bool WG> __global__ void __launch_bounds__(64) side_kernel(StepArgs<S, T> a, int nband)
{
    side_block<S, T, WG>(a, blockIdx.x, nband);
}

// Bulk, one node per thread: the first step after mrlbm_set_populations (populations in
// memory), steps with a Brinkman field, and a checking path.
template <class S, class T> __device__ void bulk_simple(const StepArgs<S, T> &a, int bb)
{
    const long long t = (long long)bb * blockDim.x + threadIdx.x;
    if (t >= (long long)a.nx * a.ny) return;
    const int c = (int)(t % a.nx), j = (int)(t / a.nx);
    if (c == 0 || c == a.nx - 1 || j == 0 || j == a.ny - 1) return; // edge kernel
    T f[Q];
    gather_interior(a, c, j, f);
    node_update<S, T, true>(a, c, j, f, false);
}

// Edge nodes (the domain's outer ring): inlet, outlet, channel walls or periodic wraps,
// stale populations, in the reference's BC order (2d_mrlbm.f90:247-351).
template <class S, class T> __device__ void edge_body(const StepArgs<S, T> &a, int r)
{
    const int nx = a.nx, ny = a.ny;
    if (r >= 2 * nx + 2 * (ny - 2)) return;
    int c, j;
    if (r < nx) { c = r; j = 0; }
    else if (r < 2 * nx) { c = r - nx; j = ny - 1; }
    else if (r < 2 * nx + ny - 2) { c = 0; j = 1 + r - 2 * nx; }
    else { c = nx - 1; j = 1 + r - 2 * nx - (ny - 2); }
    T f[Q];
    gather_edge(a, c, j, f);
    node_update(a, c, j, f, true);
}

template <class S, class T> __global__ void __launch_bounds__(256) edge_kernel(StepArgs<S, T> a)
{
    edge_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

template <class S, class T> __device__ void stencil_moments_body(const StepArgs<S, T> &a, int u)
{
    using F = typename FormOf<T>::F;
    const BoundaryDev &b = a.b;
    if (u >= b.nU) return;
    T f[Q], m[6];
    gather_interior(a, b.uij[2 * u], b.uij[2 * u + 1], f);
    F::moments(f, m);
    store_stencil_moments(a, u, m);
}

template <class S, class T> __global__ void __launch_bounds__(256) stencil_moments_kernel(StepArgs<S, T> a)
{
    stencil_moments_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// Wall BC, collision, store, momentum exchange and surface-pressure feed of boundary node k,
// after stencil_moments_kernel of the same step.  ONE_PASS (one dependency level: no
// boundary velocity is interpolated from another boundary node's): the node interpolates its
// own wall velocity, its stencil loads issued together with the gather so that the thread
// waits for one memory round trip; otherwise ub holds the velocities from wall_velocities.
template <class S, class T, bool ONE_PASS, class UA = UGlobal<T>>
__device__ __forceinline__ void wall_node(const StepArgs<S, T> &a, int k, UA ua = UA{})
{
    using F = typename FormOf<T>::F;
    const BoundaryDev &b = a.b;
    T *U = (T *)b.U, *ub = (T *)b.ub;
    if constexpr (std::is_same_v<UA, UGlobal<T>>) ua = UGlobal<T>{U, b.nU, b.uc, b.uself};
    const T om = a.om;
    const bool interp_u = b.vel_interp || !b.rotated;
    const int i = b.ij[2 * k], j = b.ij[2 * k + 1];
    const double *R = b.real + (long long)k * BT_NREAL;
    T cu[8], cv[8], cm[8], cn[8], su = (T)0.0, sv = (T)0.0;
    if constexpr (ONE_PASS) { // stencil u, v, Mxx', Myy' (every uc index is a stencil node)
#pragma unroll
        for (int c = 0; c < 8; c++) {
            cu[c] = ua(k, 1, c);
            cv[c] = ua(k, 2, c);
            cm[c] = ua(k, 3, c);
            cn[c] = ua(k, 4, c);
        }
        su = ua(k, 1, 12);
        sv = ua(k, 2, 12);
    }
    T fp[Q], ff[Q];
    gather_interior(a, i, j, fp);
#pragma unroll
    for (int q = 0; q < Q; q++) ff[q] = F::dev ? fp[q] + F::base(q) : fp[q];
    T ux, uy;
    if constexpr (ONE_PASS) {
        ux = interp_u ? wall_interp(R, cu) : su;
        uy = interp_u ? wall_interp(R, cv) : sv;
        ub[2 * k] = ux;
        ub[2 * k + 1] = uy;
    } else {
        ux = ub[2 * k];
        uy = ub[2 * k + 1];
    }
    const int lab = b.label[k];
    const uint16_t in = b.incs[lab], out = b.outs[lab];
    T m[6];
    if (b.rotated) {
        T Mxx = (T)0.0, Myy = (T)0.0;
        if constexpr (!ONE_PASS) {
            wall_moments(b, U, k, Mxx, Myy);
        } else if (b.mom_interp) {
            Mxx = wall_interp(R, cm);
            Myy = wall_interp(R, cn);
        }
        const T c = (T)R[BT_CTH], s = (T)R[BT_STH];
        const T uxp = ux * c + uy * s, uyp = uy * c - ux * s; // :304-305
        wall_bc_rotated<T>(ff, uxp, uyp, R, in, out, b.mass_bc, om, Mxx, Myy, m);
    } else {
        wall_bc_unrotated<T>(ff, ux, uy, in, out, om, m);
    }
    m[1] = ux;
    m[2] = uy;
    const long long n = (long long)j * a.pitch + i;
    if (a.rec_slot >= 0)
        for (int p = 0; p < a.nprobe; p++)
            if (n == a.probe_node[p]) {
                double *r = a.rec + (long long)a.rec_slot * a.rec_stride + 2 + 3 * p;
                r[0] = m[0];
                r[1] = m[1];
                r[2] = m[2];
            }
    if (a.sample) {
        const double vx = m[1], vy = m[2];
        a.acc[n] += vx;
        a.acc[a.plane + n] += vy;
        a.acc[2 * a.plane + n] += vx * vx;
        a.acc[3 * a.plane + n] += vy * vy;
    }
    collide<T>(m, om, a.om1, a.incomp_bulk);
    S ms[6];
    ms[0] = (S)F::store_rho(m[0]);
#pragma unroll
    for (int p = 1; p < 6; p++) ms[p] = (S)m[p];
#pragma unroll
    for (int p = 0; p < 6; p++) a.dst[p * a.plane + n] = ms[p];
    if (a.rec_slot >= 0) { // momentum exchange (:558-591)
        T fo[Q];
        F::recon(ms, fo);
        double fxi = 0, fyi = 0, fxo = 0, fyo = 0;
        for (int q = 0; q < Q; q++) {
            if ((out >> q) & 1) {
                fxo = fxo + (double)fo[q] * (double)EXk(q);
                fyo = fyo + (double)fo[q] * (double)EYk(q);
            }
            if ((in >> q) & 1) {
                fxi = fxi + (double)fp[q] * (double)EXk(q);
                fyi = fyi + (double)fp[q] * (double)EYk(q);
            }
        }
        b.fxy[2 * k] = (fxi - fxo) + (F::dev ? b.wfx[lab] : 0.0);
        b.fxy[2 * k + 1] = (fyi - fyo) + (F::dev ? b.wfy[lab] : 0.0);
        if (a.record_p) { // extrapolated surface pressure (:620-676), p = rho*cs2 pre-BC
            double pv[3];
            for (int r = 0; r < 3; r++) {
                double pc[4];
                for (int c = 0; c < 4; c++) {
                    const double s0 = (double)ua(k, 0, 4 * r + c);
                    pc[c] = (F::dev ? 1.0 + s0 : s0) * (1.0 / 3.0);
                }
                pv[r] = bilin<double>(R[BT_XD2 + 2 * r], R[BT_YD2 + 2 * r], pc[0], pc[1], pc[2], pc[3]);
            }
            b.pcy[k] = quadratic_fit(R[BT_R2], pv[0], R[BT_R3], pv[1], R[BT_R4], pv[2], R[BT_RS]);
        }
    }
}

// Wall group g (single wave, walls of one dependency level): the pre-collision moments of the
// stencil nodes its boundary nodes read, into LDS (gather and moments as stencil_moments_body),
// then WALL_GROUP boundary nodes as wall_nodes_kernel does them.  Reading the source state
// only, a group needs no other workgroup of the step, so the groups run inside the side launch.
template <class S, class T> __device__ void wall_group(const StepArgs<S, T> &a, int g)
{
    using F = typename FormOf<T>::F;
    const BoundaryDev &b = a.b;
    __shared__ T sU[5][WALL_GROUP_MAXU];
    const int lane = threadIdx.x, off = b.wg_off[g], cnt = b.wg_off[g + 1] - off;
    for (int t = lane; t < cnt; t += 64) {
        const int u = b.wg_list[off + t];
        T f[Q], m[6], v[5];
        gather_interior(a, b.uij[2 * u], b.uij[2 * u + 1], f);
        F::moments(f, m);
        stencil_values(b, u, m, v);
#pragma unroll
        for (int p = 0; p < 5; p++) sU[p][t] = v[p];
    }
    __syncthreads();
    const int q = g * WALL_GROUP + lane;
    if (lane < WALL_GROUP && q < b.nbct) wall_node<S, T, true>(a, b.wg_node[q], ULocal<T>{sU, b.wl_slot});
}

// Weight of wall node q in the smoothed surface pressure at node t (2d_mrlbm.f90:722-741):
// a Gaussian in the angle with sigma = 0.1.
__device__ __forceinline__ double smooth_weight(double th_t, double th_q)
{
    const double den = 2.0 * (0.1 * 0.1);
    const double d = th_t - th_q;
    return exp((-(d * d)) / den);
}

// The weights depend on the wall angles only: computed once, stored column-major (q*n + t,
// so that the per-step sums read them coalesced), with their sums in the reference's order.
__global__ void __launch_bounds__(256) smooth_weights_kernel(BoundaryDev b, double *w, double *sw)
{
    const int n = b.nbct;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < n; q++) {
            const double v = smooth_weight(b.thetas[t], b.thetas[q]);
            w[(long long)q * n + t] = v;
            s = s + v;
        }
        sw[t] = s;
    }
}

// Gaussian-smoothed surface pressure at sorted wall node t (2d_mrlbm.f90:722-741), one wave:
// lane l sums the terms q = l, l+64, ... in order, then a fixed shuffle tree adds the 64
// partial sums, so the value does not depend on which wave computes it.  The weight sum is
// the precomputed one (the reference's sequential order) when the weights are stored.
// The reference sums all nbct terms in order; this differs from it by rounding only.
template <class S, class T>
__device__ void smooth_output(const StepArgs<S, T> &a, const double *pcy, double *rec, int t)
{
    const BoundaryDev &b = a.b;
    const int lane = threadIdx.x & 63, n = b.nbct;
    double sv = 0.0, sw = 0.0;
    for (int q = lane; q < n; q += 64) {
        const double w = b.gw ? b.gw[(long long)q * n + t] : smooth_weight(b.thetas[t], b.thetas[q]);
        sv = sv + w * pcy[b.idx[q]];
        sw = sw + w;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sv += __shfl_down(sv, o, 64);
        sw += __shfl_down(sw, o, 64);
    }
    if (lane == 0) rec[14 + n + t] = sv / (b.gsw ? b.gsw[t] : sw);
}

// The smoothing of one recording step spread over single-wave workgroups (one output per
// wave at a time): in one workgroup it read the nbct^2 weights with a few waves' memory
// parallelism and took longer than the rest of the step.
template <class S, class T> __global__ void __launch_bounds__(64) smooth_kernel(StepArgs<S, T> a)
{
    double *rec = a.rec + (long long)a.rec_slot * a.rec_stride;
    for (int t = blockIdx.x; t < a.b.nbct; t += gridDim.x) smooth_output(a, a.b.pcy, rec, t);
}

// The smoothing of the recording steps of a four-step pass (blockIdx.y: the step), after its
// records kernel: step k's pressures are in slot r.wslot[k] of the per-node array.
template <class S, class T> __global__ void __launch_bounds__(64) smooth4_kernel(StepArgs<S, T> a, RecSteps r)
{
    const int k = blockIdx.y;
    if (!(r.record_p[k] && r.rec_slot[k] >= 0)) return;
    double *rec = a.rec + (long long)r.rec_slot[k] * a.rec_stride;
    const double *pcy = a.b.pcy + (long long)r.wslot[k] * a.b.nbct;
    for (int t = blockIdx.x; t < a.b.nbct; t += gridDim.x) smooth_output(a, pcy, rec, t);
}

// Net force, records and the inlet-density running mean of one step (one workgroup), after
// every boundary node of the step has been written.
template <class S, class T>
__device__ void wall_records_step(const StepArgs<S, T> &a, int rec_slot, int record_p, int smooth_apart,
                                  const double *fxy, const double *pcy, const double *inlet_prev,
                                  double inlet_prev_c)
{
    const BoundaryDev &b = a.b;
    __shared__ double sh[16];
    const int tid = threadIdx.x, nt = blockDim.x;
    if (rec_slot >= 0) {
        double sx = 0.0, sy = 0.0;
        for (int k = tid; k < b.nbct; k += nt) {
            sx += fxy[2 * k];
            sy += fxy[2 * k + 1];
        }
        sx = block_sum(sx, sh);
        sy = block_sum(sy, sh);
        double *rec = a.rec + (long long)rec_slot * a.rec_stride;
        if (tid == 0) {
            rec[0] = sx;
            rec[1] = sy;
            if (a.facc) {
                a.facc[0] += sx;
                a.facc[1] += sy;
                a.facc[2] += sy * sy;
                a.facc[3] += 1.0;
            }
        }
        if (record_p) { // theta-sorted (:602-604); smoothed here unless a launch of its own does it
            double *ps = rec + 14;
            for (int t = tid; t < b.nbct; t += nt) ps[t] = pcy[b.idx[t]];
            if (!smooth_apart)
                for (int t = tid >> 6; t < b.nbct; t += nt >> 6) smooth_output(a, pcy, rec, t);
        }
    }
    if (inlet_prev) { // running mean of the inlet density of the previous step (:357-360)
        double s = 0.0;
        for (int j = tid; j < a.ny; j += nt) s += inlet_prev[j];
        s = block_sum(s, sh);
        if (tid == 0) {
            const double r = s / (double)a.ny, c = inlet_prev_c;
            *a.rho_inf_mean = ((c * *a.rho_inf_mean) + r) / (c + 1.0);
        }
        __syncthreads(); // the next step's fold reads the mean
    }
}
template <class S, class T> __device__ void wall_records(const StepArgs<S, T> &a)
{
    wall_records_step(a, a.rec_slot, a.record_p, a.smooth_apart, a.b.fxy, a.b.pcy, a.inlet_prev, a.inlet_prev_c);
}

// All boundary nodes of one step in one workgroup: velocities (by dependency level), nodes.
template <class S, class T> __device__ void wall_nodes_block(const StepArgs<S, T> &a)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    if (a.b.nlevels <= 1) {
        for (int k = tid; k < a.b.nbct; k += nt) wall_node<S, T, true>(a, k);
    } else {
        wall_velocities(a);
        for (int k = tid; k < a.b.nbct; k += nt) wall_node<S, T, false>(a, k);
    }
}

// ... and the records of the step.
template <class S, class T> __device__ void boundary_body(const StepArgs<S, T> &a)
{
    wall_nodes_block(a);
    __syncthreads();
    wall_records(a);
}

// One workgroup of up to 512 threads (one boundary node per thread up to nbct = 512).
template <class S, class T> __global__ void __launch_bounds__(512) boundary_kernel(StepArgs<S, T> a)
{
    boundary_body(a);
}

// One boundary node per thread in single-wave workgroups (one dependency level only): a
// workgroup this small finds room on a CU beside a running bulk launch as soon as one of the
// bulk's waves retires.
template <class S, class T> __global__ void __launch_bounds__(64) wall_nodes_kernel(StepArgs<S, T> a)
{
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k < a.b.nbct) wall_node<S, T, true>(a, k);
}

// The records of the steps of a four-step pass in step order (one workgroup), after the pass's
// side work: step k's forces and pressures from slot r.wslot[k] of the per-node arrays, and
// the inlet-density fold of the step before it.
template <class S, class T> __global__ void __launch_bounds__(256) wall_records4_kernel(StepArgs<S, T> a, RecSteps r)
{
    for (int k = 0; k < r.n; k++)
        wall_records_step(a, r.rec_slot[k], r.record_p[k], 1, a.b.fxy + (long long)r.wslot[k] * 2 * a.b.nbct,
                          a.b.pcy + (long long)r.wslot[k] * a.b.nbct, r.inlet_prev[k], r.inlet_prev_c[k]);
}
template <class S, class T> __global__ void __launch_bounds__(256) wall_records_kernel(StepArgs<S, T> a)
{
    wall_records(a);
}

// H: 0 row loop over a runtime strip height, 4/8/16 unrolled strip of H rows (62 written
// columns per wave), 304/308/316 aligned 64-column waves with an unrolled strip of H-300 rows.
template <class S, class T, bool REC, int H> __device__ __forceinline__ void bulk_body(const StepArgs<S, T> &a, int bb)
{
    if constexpr (H >= 300) bulk_march_a<S, T, REC, H - 300>(a, bb);
    else if constexpr (H == 0) bulk_march<S, T, REC>(a, bb);
    else bulk_march_u<S, T, REC, H>(a, bb);
}

template <class S, class T, bool REC, int H> __global__ void __launch_bounds__(256) bulk_kernel(StepArgs<S, T> a)
{
    bulk_body<S, T, REC, H>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------------
// Batched runs: per-run argument blocks in device memory, one launch per kernel for all
// runs.  The bulk walks a flattened tile table (run, block of that run), persistent or one
// block per tile; the side kernel takes the run from blockIdx.y, the boundary workgroup of
// run r is block r.
template <class S, class T, int H>
__global__ void __launch_bounds__(256) bulk_batch_kernel(const StepArgs<S, T> *__restrict__ args, const int2 *__restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        const StepArgs<S, T> a = args[tl.x];
        bulk_body<S, T, false, H>(a, tl.y);
    }
}

// The wall of every run, one workgroup per run, velocities by dependency level.
template <class S, class T> __global__ void __launch_bounds__(256) boundary_batch_kernel(const StepArgs<S, T> *__restrict__ args)
{
    const StepArgs<S, T> &a = args[blockIdx.x];
    if (!a.has_boundary) return;
    boundary_body(a);
}

// Side work of every run from a compact table of (run, block): wall groups, band marches, ring
// column tiles, then blocks of 64 ring, band and stencil nodes, in single-wave workgroups,
// none of them empty (side_block's order, per run).
template <class S, class T> __global__ void __launch_bounds__(64) side_tiles_kernel(const StepArgs<S, T> *__restrict__ args, const int2 *__restrict__ tiles)
{
    const int2 t = tiles[blockIdx.x];
    const StepArgs<S, T> &a = args[t.x];
    side_block<S, T, true>(a, t.y, band_nodes(a));
}

// Wall nodes of every run from a table of (run, block of 64 boundary nodes), for walls of one
// dependency level; the records follow in wall_records_batch_kernel.
template <class S, class T> __global__ void __launch_bounds__(64) wall_tiles_kernel(const StepArgs<S, T> *__restrict__ args, const int2 *__restrict__ tiles)
{
    const int2 t = tiles[blockIdx.x];
    const StepArgs<S, T> &a = args[t.x];
    const int k = t.y * 64 + threadIdx.x;
    if (k < a.b.nbct) wall_node<S, T, true>(a, k);
}

template <class S, class T> __global__ void __launch_bounds__(256) wall_records_batch_kernel(const StepArgs<S, T> *__restrict__ args)
{
    const StepArgs<S, T> &a = args[blockIdx.x];
    if (!a.has_boundary) return;
    wall_records(a);
}

template <class S, class T, int H>
__global__ void __launch_bounds__(256) bulk2_batch_kernel(const StepArgs<S, T> *__restrict__ args, const int2 *__restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        const StepArgs<S, T> a = args[tl.x];
        bulk_march2<S, T, H, false>(a, tl.y);
    }
}

// Four steps per pass for every run from the flattened tile table (run, four-step block).
template <class S, class T, int H>
__global__ void __launch_bounds__(256, 2) bulkn_batch_kernel(const StepArgs<S, T> *__restrict__ args, const int2 *__restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        const StepArgs<S, T> a = args[tl.x];
        bulk_march_n<S, T, 4, H, false>(a, tl.y);
    }
}

template <class S, class T> __global__ void __launch_bounds__(256) bulk_simple_kernel(StepArgs<S, T> a)
{
    bulk_simple(a, blockIdx.x);
}

// Probes inside a two-step pair, one thread per probe.  The ring, band and wall kernels of
// each sub-step record the probes on the nodes they write; this records the others, which
// the two-step bulk carries through t+1 without storing them: at t+1 the moments gathered
// from state t (what the bulk computes there; the collision leaves rho and u unchanged) for
// a node that sub-step 1's side work does not write, and at t+2 the stored rho and u of a
// node that the bulk writes.  a: sub-step 1 (src = state t, its band rectangles, the record
// slot of t+1 or -1); s2, slot2: state t+2 and its record slot (or -1).
template <class S, class T> __global__ void __launch_bounds__(64) pair_probes_kernel(StepArgs<S, T> a, const S *s2, int slot2)
{
    using F = typename FormOf<T>::F;
    const int p = threadIdx.x;
    if (p >= a.nprobe) return;
    const long long n = a.probe_node[p];
    const int c = (int)(n % a.pitch), j = (int)(n / a.pitch);
    if (a.rec_slot >= 0) {
        bool side = c == 0 || c == a.nx - 1 || j == 0 || j == a.ny - 1;
        for (int q = 0; q < a.nrect; q++)
            side = side || (c >= a.rect[q][0] && c < a.rect[q][0] + a.rect[q][2] && j >= a.rect[q][1] &&
                            j < a.rect[q][1] + a.rect[q][3]);
        const int bi = c - a.bx0, bj = j - a.by0;
        if ((unsigned)bi < (unsigned)a.bw && (unsigned)bj < (unsigned)a.bh) side = side || (a.boxcls[bj * a.bw + bi] & 1);
        if (!side) {
            T f[Q], m[6];
            gather_interior(a, c, j, f);
            F::moments(f, m);
            double *r = a.rec + (long long)a.rec_slot * a.rec_stride + 2 + 3 * p;
            r[0] = F::dev ? 1.0 + (double)m[0] : (double)m[0];
            r[1] = m[1];
            r[2] = m[2];
        }
    }
    if (slot2 >= 0 && two_step_node(a, c, j)) {
        const S r0 = s2[n], u = s2[a.plane + n], v = s2[2 * a.plane + n];
        double *r = a.rec + (long long)slot2 * a.rec_stride + 2 + 3 * p;
        r[0] = F::dev ? 1.0 + (double)r0 : (double)r0;
        r[1] = u;
        r[2] = v;
    }
}

// The t+2 samples of a pair with both steps sampled on the nodes that sub-step 1 writes and
// the two-step bulk writes at t+2 (the bands' rows and columns beyond the ones sub-step 2
// redoes), after both: from the stored rho and u of state t+2 (s2).  a: sub-step 1, with its
// band rectangles.
template <class S, class T> __global__ void __launch_bounds__(256) pair_samples_kernel(StepArgs<S, T> a, const S *s2)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int q = 0; q < a.nrect; q++) {
        const long long n = (long long)a.rect[q][2] * a.rect[q][3];
        if (t < n) {
            const int i = a.rect[q][0] + (int)(t % a.rect[q][2]), j = a.rect[q][1] + (int)(t / a.rect[q][2]);
            if (two_step_node(a, i, j) && !in_rects(a, q, i, j)) {
                const long long m = (long long)j * a.pitch + i;
                sample_u(a, m, (T)s2[a.plane + m], (T)s2[2 * a.plane + m]);
            }
            return;
        }
        t -= n;
    }
}

// Probes of a four-step pass on the nodes the four-step bulk writes, one workgroup per
// probe, after the bulk.  The bulk carries t+1 .. t+3 in registers, so this recomputes them
// at the probe node from state t: level l = 1..3 on the (7-2l)^2 nodes around it, each a
// plain interior update with the single step's arithmetic (such a node's cone of radius 3
// lies in the plain interior: it is four or more from the ring and outside the cylinder box
// dilated by three), then t+4 from the stored state.  a: the bulk's arguments (src = state
// t, dst = state t+4, qslot = the record slot of t+1 .. t+4 or -1).
template <class S, class T> __global__ void __launch_bounds__(64) quad_probes_kernel(StepArgs<S, T> a)
{
    using F = typename FormOf<T>::F;
    const int p = blockIdx.x, t = threadIdx.x;
    const long long n0 = a.probe_node[p];
    const int c = (int)(n0 % a.pitch), j = (int)(n0 / a.pitch);
    if (!n_step_node<S, T, 4>(a, c, j)) return; // the sub-steps record it (uniform per block)
    __shared__ T mom[49][6];  // moments of the level below, on its (9-2l)^2 nodes
    __shared__ T pop[49][Q];  // their populations
    auto record = [&](int slot, T r0, T u, T v) {
        double *r = a.rec + (long long)slot * a.rec_stride + 2 + 3 * p;
        r[0] = F::dev ? 1.0 + (double)r0 : (double)r0;
        r[1] = u;
        r[2] = v;
    };
    if (t < 49) { // level 0: state t at the 7 x 7 nodes around the probe
        const long long n = (long long)(j - 3 + t / 7) * a.pitch + (c - 3 + t % 7);
#pragma unroll
        for (int k = 0; k < 6; k++) mom[t][k] = a.src[k * a.plane + n];
    }
    __syncthreads();
    for (int l = 1; l <= 3; l++) {
        const int wb = 9 - 2 * l, w = wb - 2; // widths of the level below and of level l
        if (t < wb * wb) F::recon(mom[t], pop[t]);
        __syncthreads();
        T m[6];
        if (t < w * w) {
            const int x = t % w + 1, y = t / w + 1; // in the level below's coordinates
            T f[Q];
#pragma unroll
            for (int k = 0; k < Q; k++) f[k] = pop[(y - EYk(k)) * wb + (x - EXk(k))][k];
            F::moments(f, m);
            if (t == w * w / 2 && a.qslot[l - 1] >= 0) record(a.qslot[l - 1], m[0], m[1], m[2]);
            collide<T>(m, a.om, a.om1, a.incomp_bulk);
        }
        __syncthreads(); // every read of mom by recon is done
        if (t < w * w)
#pragma unroll
            for (int k = 0; k < 6; k++) mom[t][k] = m[k];
        __syncthreads();
    }
    if (t == 0 && a.qslot[3] >= 0) record(a.qslot[3], (T)a.dst[n0], (T)a.dst[a.plane + n0], (T)a.dst[2 * a.plane + n0]);
}

// A pitched state plane to a dense (nx, ny) FP64 array (i fastest) and back, shifted by
// `shift` (the FP32 form stores rho - 1): the host copies of the C-ABI are then one
// contiguous transfer.  Padding columns of the pitched plane are written as zeros.
template <class S>
__global__ void __launch_bounds__(256) dense_from_plane_kernel(const S *src, double *dst, int nx, int ny, int pitch, double shift)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nx * ny) return;
    const int i = (int)(t % nx), j = (int)(t / nx);
    dst[t] = shift + (double)src[(long long)j * pitch + i];
}
template <class S>
__global__ void __launch_bounds__(256) plane_from_dense_kernel(const double *src, S *dst, int nx, int ny, int pitch, double shift)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)pitch * ny) return;
    const int i = (int)(t % pitch), j = (int)(t / pitch);
    dst[t] = i < nx ? (S)(src[(long long)j * nx + i] - shift) : (S)0;
}

// State scan for the log step (2d_mrlbm.f90:546) and the finite check: per block, the
// largest ux over the block's nodes (NaN ignored) and whether any of the six stored moments
// of its nodes is NaN or Inf; scan_finish_kernel folds the block results.  part holds
// (max, flag) per block.
template <class S>
__global__ void __launch_bounds__(256) state_scan_kernel(const S *src, long long plane, int nx, int ny, int pitch,
                                                         double *part)
{
    __shared__ double smax[4];
    __shared__ int sbad[4];
    const long long N = (long long)nx * ny;
    double mx = -INFINITY;
    int bad = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)(t / nx);
        const long long n = (long long)j * pitch + i;
        const double u = (double)src[plane + n];
        mx = fmax(mx, u);
#pragma unroll
        for (int p = 0; p < 6; p++) bad |= !isfinite((double)src[p * plane + n]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_down(mx, o, 64));
        bad |= __shfl_down(bad, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smax[w] = mx;
        sbad[w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
            mx = fmax(mx, smax[k]);
            bad |= sbad[k];
        }
        part[2 * blockIdx.x] = mx;
        part[2 * blockIdx.x + 1] = bad ? 1.0 : 0.0;
    }
}

__global__ void __launch_bounds__(64) scan_finish_kernel(double *part, int nblocks, double *out)
{
    double mx = -INFINITY, bad = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 64) {
        mx = fmax(mx, part[2 * b]);
        bad = fmax(bad, part[2 * b + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_down(mx, o, 64));
        bad = fmax(bad, __shfl_down(bad, o, 64));
    }
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = bad;
    }
}

template <class S, class T> __global__ void __launch_bounds__(256) inlet_fold_kernel(StepArgs<S, T> a)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int j = threadIdx.x; j < a.ny; j += blockDim.x) s += a.inlet_prev[j];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double r = s / (double)a.ny, c = a.inlet_prev_c;
        *a.rho_inf_mean = ((c * *a.rho_inf_mean) + r) / (c + 1.0);
    }
}

// Post-stream populations of rows j0 .. j0+rows-1 as a dense band f(nx, rows, 0:8) (x fastest,
// then the band's rows, then k); the whole grid when j0 = 0 and rows = ny.
template <class S, class T>
__global__ void __launch_bounds__(256) get_populations_kernel(StepArgs<S, T> a, double *f, int j0, int rows)
{
    using F = typename FormOf<T>::F;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)a.nx * rows;
    if (t >= N) return;
    const int i = (int)(t % a.nx), j = j0 + (int)(t / a.nx);
    for (int k = 0; k < Q; k++) {
        const T v = pull_rule(a, i, j, k);
        f[k * N + t] = F::dev ? (double)F::base(k) + (double)v : (double)v;
    }
}

// Stale populations of a state given by moments: the reconstruction at the edge node itself.
template <class S, class T>
__global__ void __launch_bounds__(256) ghost_state_kernel(StepArgs<S, T> a, T *ghost)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = 2 * a.ny + 2 * a.nx;
    if (t >= L) return;
    int i, j;
    if (t < a.ny) { i = 0; j = t; }
    else if (t < 2 * a.ny) { i = a.nx - 1; j = t - a.ny; }
    else if (t < 2 * a.ny + a.nx) { i = t - 2 * a.ny; j = 0; }
    else { i = t - 2 * a.ny - a.nx; j = a.ny - 1; }
    S m[6];
    load_moments(a, i, j, m);
    T f[Q];
    FormOf<T>::F::recon(m, f);
    for (int k = 0; k < Q; k++) ghost[t * Q + k] = f[k];
}

template <class S, class T>
__global__ void __launch_bounds__(256) ghost_pops_kernel(StepArgs<S, T> a, T *ghost)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = 2 * a.ny + 2 * a.nx;
    if (t >= L) return;
    int i, j;
    if (t < a.ny) { i = 0; j = t; }
    else if (t < 2 * a.ny) { i = a.nx - 1; j = t - a.ny; }
    else if (t < 2 * a.ny + a.nx) { i = t - 2 * a.ny; j = 0; }
    else { i = t - 2 * a.ny - a.nx; j = a.ny - 1; }
    for (int k = 0; k < Q; k++) ghost[t * Q + k] = a.pops[(long long)k * a.plane + (long long)j * a.pitch + i];
}

// ---------------------------------------------------------------------------------
// Launch on stream st; t0/t1 (either may be null) take the kernel's own start and end times
// (no packets of their own between kernels): for timing, and t1 as the launch's completion
// event that another stream can wait on.
template <class K, class... A>
static void launch_timed(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t t0, hipEvent_t t1, A... args)
{
    if (t0 || t1) hipExtLaunchKernelGGL(kernel, g, b, 0, st, t0, t1, 0, args...);
    else hipLaunchKernelGGL(kernel, g, b, 0, st, args...);
}

// Four-step strip heights built: FP32 12..32 (multiples of 4), FP64 8, 16, 20, 24, 32, 40 (at
// 12 and 52 rows the compiler moves the FP64 march's carried rows to scratch memory).
template <class S> bool bulkn_rows_ok(int H)
{
    if constexpr (sizeof(S) == 4) return H == 12 || H == 16 || H == 20 || H == 24 || H == 28 || H == 32;
    else return H == 8 || H == 16 || H == 20 || H == 24 || H == 32 || H == 40;
}
// calls f with std::integral_constant<int, H> for a built strip height H (false otherwise)
template <class S, class F> static bool with_bulkn_rows(int H, F &&f)
{
    if constexpr (sizeof(S) == 4) {
        switch (H) {
        case 12: f(std::integral_constant<int, 12>{}); return true;
        case 16: f(std::integral_constant<int, 16>{}); return true;
        case 20: f(std::integral_constant<int, 20>{}); return true;
        case 24: f(std::integral_constant<int, 24>{}); return true;
        case 28: f(std::integral_constant<int, 28>{}); return true;
        case 32: f(std::integral_constant<int, 32>{}); return true;
        }
    } else {
        switch (H) {
        case 8: f(std::integral_constant<int, 8>{}); return true;
        case 16: f(std::integral_constant<int, 16>{}); return true;
        case 20: f(std::integral_constant<int, 20>{}); return true;
        case 24: f(std::integral_constant<int, 24>{}); return true;
        case 32: f(std::integral_constant<int, 32>{}); return true;
        case 40: f(std::integral_constant<int, 40>{}); return true;
        }
    }
    return false;
}
template <class S, class T>
hipError_t launch_bulk(const StepArgs<S, T> &a, int nblocks, hipStream_t st, hipEvent_t t0, hipEvent_t t1)
{
    const bool rec = a.sample || (a.rec_slot >= 0 && a.nprobe > 0);
    const dim3 g(nblocks), b(256);
#define MRLBM_BULK(R, H) launch_timed(bulk_kernel<S, T, R, H>, g, b, st, t0, t1, a)
    if (a.simple || a.pops || a.chi) launch_timed(bulk_simple_kernel<S, T>, g, b, st, t0, t1, a);
    else if (rec && a.variant >= 300) MRLBM_BULK(true, 304);
    else if (rec) MRLBM_BULK(true, 0);
    else if (a.variant == 4) MRLBM_BULK(false, 4);
    else if (a.variant == 8) MRLBM_BULK(false, 8);
    else if (a.variant == 16) MRLBM_BULK(false, 16);
    else if (a.variant == 304) MRLBM_BULK(false, 304);
    else if (a.variant == 308) MRLBM_BULK(false, 308);
    else if (a.variant == 316) MRLBM_BULK(false, 316);
    else MRLBM_BULK(false, 0);
#undef MRLBM_BULK
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_edges(const StepArgs<S, T> &a, hipStream_t st)
{
    const int n = 2 * a.nx + 2 * (a.ny - 2);
    hipLaunchKernelGGL((edge_kernel<S, T>), dim3((n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_boundary(const StepArgs<S, T> &a, hipStream_t st)
{
    hipLaunchKernelGGL((stencil_moments_kernel<S, T>), dim3((a.b.nU + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL((boundary_kernel<S, T>), dim3(1), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_bulk_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int variant, int grid, hipStream_t st)
{
    const dim3 g(grid), b(256);
    switch (variant) {
    case 0: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 0>), g, b, 0, st, args, tiles, ntiles); break;
    case 4: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 4>), g, b, 0, st, args, tiles, ntiles); break;
    case 8: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 8>), g, b, 0, st, args, tiles, ntiles); break;
    case 16: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 16>), g, b, 0, st, args, tiles, ntiles); break;
    case 304: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 304>), g, b, 0, st, args, tiles, ntiles); break;
    case 308: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 308>), g, b, 0, st, args, tiles, ntiles); break;
    case 316: hipLaunchKernelGGL((bulk_batch_kernel<S, T, 316>), g, b, 0, st, args, tiles, ntiles); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <class S, class T> int bulk_batch_resident_blocks(int variant)
{
    int per_cu = 0, dev = 0, ncu = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (variant) {
    case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 0>, 256, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 4>, 256, 0); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 8>, 256, 0); break;
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 16>, 256, 0); break;
    case 304: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 304>, 256, 0); break;
    case 308: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 308>, 256, 0); break;
    case 316: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bulk_batch_kernel<S, T, 316>, 256, 0); break;
    default: break;
    }
    if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * ncu;
}
template <class S, class T>
hipError_t launch_side_tiles(const StepArgs<S, T> *args, const int2 *stiles, int nstiles, const int2 *wtiles,
                             int nwtiles, int nruns, bool wall_levels, bool walls, bool records, hipStream_t st)
{
    if (nstiles > 0) hipLaunchKernelGGL((side_tiles_kernel<S, T>), dim3(nstiles), dim3(64), 0, st, args, stiles);
    if (wall_levels) { // some wall interpolates in place by level: one workgroup per run
        hipLaunchKernelGGL((boundary_batch_kernel<S, T>), dim3(nruns), dim3(256), 0, st, args);
    } else if (walls) { // the walls outside wall groups, then the records of every wall
        if (nwtiles > 0) hipLaunchKernelGGL((wall_tiles_kernel<S, T>), dim3(nwtiles), dim3(64), 0, st, args, wtiles);
        if (records) hipLaunchKernelGGL((wall_records_batch_kernel<S, T>), dim3(nruns), dim3(256), 0, st, args);
    }
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_bulk2_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int H, hipStream_t st)
{
    const dim3 g(ntiles), b(256);
    switch (H) {
    case 2: hipLaunchKernelGGL((bulk2_batch_kernel<S, T, 2>), g, b, 0, st, args, tiles, ntiles); break;
    case 4: hipLaunchKernelGGL((bulk2_batch_kernel<S, T, 4>), g, b, 0, st, args, tiles, ntiles); break;
    case 6: hipLaunchKernelGGL((bulk2_batch_kernel<S, T, 6>), g, b, 0, st, args, tiles, ntiles); break;
    case 8: hipLaunchKernelGGL((bulk2_batch_kernel<S, T, 8>), g, b, 0, st, args, tiles, ntiles); break;
    case 10: hipLaunchKernelGGL((bulk2_batch_kernel<S, T, 10>), g, b, 0, st, args, tiles, ntiles); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_bulkn_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int H, hipStream_t st)
{
    const dim3 g(ntiles), b(256);
    const bool ok = with_bulkn_rows<S>(H, [&](auto hh) {
        constexpr int HH = decltype(hh)::value;
        hipLaunchKernelGGL((bulkn_batch_kernel<S, T, HH>), g, b, 0, st, args, tiles, ntiles);
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}
template <class S, class T>
hipError_t launch_bulk2(const StepArgs<S, T> &a, int nblocks, int H, hipStream_t st, hipEvent_t t0, hipEvent_t t1)
{
    const dim3 g(nblocks), b(256);
    const bool smp = a.sample || a.sample2;
#define MRLBM_BULK2(HH)                                                                    \
    case HH:                                                                               \
        if (smp) launch_timed(bulk2_kernel<S, T, HH, true>, g, b, st, t0, t1, a);          \
        else launch_timed(bulk2_kernel<S, T, HH, false>, g, b, st, t0, t1, a);             \
        break;
    switch (H) {
        MRLBM_BULK2(2)
        MRLBM_BULK2(4)
        MRLBM_BULK2(6)
        MRLBM_BULK2(8)
        MRLBM_BULK2(10)
        MRLBM_BULK2(12)
        MRLBM_BULK2(14)
        MRLBM_BULK2(16)
    default: return hipErrorInvalidValue;
    }
#undef MRLBM_BULK2
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_bulkn(const StepArgs<S, T> &a, int nblocks, int H, hipStream_t st, hipEvent_t t0, hipEvent_t t1)
{
    const dim3 g(nblocks), b(256);
    // the recording variant when a level is sampled
    const bool rec = a.qsample != 0;
    const bool ok = with_bulkn_rows<S>(H, [&](auto hh) {
        constexpr int HH = decltype(hh)::value;
        if (rec) launch_timed(bulkn_kernel<S, T, 4, HH, true>, g, b, st, t0, t1, a);
        else launch_timed(bulkn_kernel<S, T, 4, HH>, g, b, st, t0, t1, a);
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}
template <class S, class T> hipError_t launch_side(const StepArgs<S, T> &a, hipStream_t st, bool stencil)
{
    const int nring = 2 * a.nx + 2 * (a.ny - 2);
    const long long nband = band_items(a.rect, a.nrect, a.band_march);
    const int nct = ring_column_tiles(a.nx, a.ny, a.pops != nullptr);
    const long long n = (nct ? 2 * a.nx : nring) + nband + (stencil && a.has_boundary ? a.b.nU : 0);
    // single-wave workgroups: they fit beside a running bulk launch (see wall_nodes_kernel)
    const dim3 g((unsigned)(a.wall_groups + band_march_blocks(a.rect, a.nrect, a.band_march) + nct + (n + 63) / 64));
    if (a.wall_groups) hipLaunchKernelGGL((side_kernel<S, T, true>), g, dim3(64), 0, st, a, (int)nband);
    else hipLaunchKernelGGL((side_kernel<S, T, false>), g, dim3(64), 0, st, a, (int)nband);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_records4(const StepArgs<S, T> &a, const RecSteps &r, hipStream_t st)
{
    if (r.n <= 0) return hipSuccess;
    hipLaunchKernelGGL((wall_records4_kernel<S, T>), dim3(1), dim3(256), 0, st, a, r);
    bool smooth = false;
    for (int k = 0; k < r.n; k++) smooth = smooth || (r.record_p[k] && r.rec_slot[k] >= 0);
    if (smooth)
        hipLaunchKernelGGL((smooth4_kernel<S, T>), dim3(std::min(a.b.nbct, 1024), r.n), dim3(64), 0, st, a, r);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_wall(const StepArgs<S, T> &a, hipStream_t st, bool records)
{
    if (a.b.nlevels <= 1) {
        if (a.b.nbct > 0 && !a.wall_groups) // (else the side launch did the wall nodes)
            hipLaunchKernelGGL((wall_nodes_kernel<S, T>), dim3((a.b.nbct + 63) / 64), dim3(64), 0, st, a);
        if (records && (a.rec_slot >= 0 || a.inlet_prev)) {
            StepArgs<S, T> r = a;
            r.smooth_apart = a.record_p && a.rec_slot >= 0 ? 1 : 0;
            hipLaunchKernelGGL((wall_records_kernel<S, T>), dim3(1), dim3(256), 0, st, r);
            if (r.smooth_apart)
                hipLaunchKernelGGL((smooth_kernel<S, T>), dim3(std::min(a.b.nbct, 1024)), dim3(64), 0, st, r);
        }
        return hipGetLastError();
    }
    const int threads = std::min(512, std::max(256, (a.b.nbct + 63) / 64 * 64));
    hipLaunchKernelGGL((boundary_kernel<S, T>), dim3(1), dim3(threads), 0, st, a);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_stencil_moments(const StepArgs<S, T> &a, hipStream_t st)
{
    hipLaunchKernelGGL((stencil_moments_kernel<S, T>), dim3((a.b.nU + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_smooth_weights(const BoundaryDev &b, double *w, double *sw, hipStream_t st)
{
    hipLaunchKernelGGL(smooth_weights_kernel, dim3((b.nbct + 255) / 256), dim3(256), 0, st, b, w, sw);
    return hipGetLastError();
}

template <class S, class T> hipError_t launch_pair_samples(const StepArgs<S, T> &a, const S *s2, hipStream_t st)
{
    long long n = 0;
    for (int q = 0; q < a.nrect; q++) n += (long long)a.rect[q][2] * a.rect[q][3];
    if (n > 0) hipLaunchKernelGGL((pair_samples_kernel<S, T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, s2);
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_pair_probes(const StepArgs<S, T> &a, const S *s2, int slot2, hipStream_t st)
{
    hipLaunchKernelGGL((pair_probes_kernel<S, T>), dim3(1), dim3(64), 0, st, a, s2, slot2);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_quad_probes(const StepArgs<S, T> &a, hipStream_t st)
{
    if (a.nprobe > 0) hipLaunchKernelGGL((quad_probes_kernel<S, T>), dim3(a.nprobe), dim3(64), 0, st, a);
    return hipGetLastError();
}
template <class S>
hipError_t launch_dense_from_plane(const S *src, double *dst, int nx, int ny, int pitch, double shift, hipStream_t st)
{
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((dense_from_plane_kernel<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, nx, ny,
                       pitch, shift);
    return hipGetLastError();
}
template <class S>
hipError_t launch_plane_from_dense(const double *src, S *dst, int nx, int ny, int pitch, double shift, hipStream_t st)
{
    const long long n = (long long)pitch * ny;
    hipLaunchKernelGGL((plane_from_dense_kernel<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, nx, ny,
                       pitch, shift);
    return hipGetLastError();
}
template hipError_t launch_dense_from_plane<float>(const float *, double *, int, int, int, double, hipStream_t);
template hipError_t launch_dense_from_plane<double>(const double *, double *, int, int, int, double, hipStream_t);
template hipError_t launch_plane_from_dense<float>(const double *, float *, int, int, int, double, hipStream_t);
template hipError_t launch_plane_from_dense<double>(const double *, double *, int, int, int, double, hipStream_t);

template <class S>
hipError_t launch_state_scan(const S *src, long long plane, int nx, int ny, int pitch, double *scratch,
                             int scratch_words, hipStream_t st)
{
    // scratch: the 2 result words, then 2 words per block
    const long long N = (long long)nx * ny;
    const int nb = (int)std::max(1LL, std::min<long long>((N + 255) / 256, (scratch_words - 2) / 2));
    if (nb < 1 || 2 * nb + 2 > scratch_words) return hipErrorInvalidValue;
    hipLaunchKernelGGL((state_scan_kernel<S>), dim3(nb), dim3(256), 0, st, src, plane, nx, ny, pitch, scratch + 2);
    hipLaunchKernelGGL(scan_finish_kernel, dim3(1), dim3(64), 0, st, scratch + 2, nb, scratch);
    return hipGetLastError();
}
template hipError_t launch_state_scan<float>(const float *, long long, int, int, int, double *, int, hipStream_t);
template hipError_t launch_state_scan<double>(const double *, long long, int, int, int, double *, int, hipStream_t);

template <class S, class T> hipError_t launch_inlet_fold(const StepArgs<S, T> &a, hipStream_t st)
{
    hipLaunchKernelGGL((inlet_fold_kernel<S, T>), dim3(1), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <class S, class T>
hipError_t launch_get_populations(const StepArgs<S, T> &a, double *f, int j0, int rows, hipStream_t st)
{
    if (j0 < 0 || rows <= 0 || j0 + rows > a.ny) return hipErrorInvalidValue;
    const long long N = (long long)a.nx * rows;
    hipLaunchKernelGGL((get_populations_kernel<S, T>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, a, f, j0,
                       rows);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_ghost_from_state(const StepArgs<S, T> &a, T *g, hipStream_t st)
{
    const int L = 2 * (a.nx + a.ny);
    hipLaunchKernelGGL((ghost_state_kernel<S, T>), dim3((L + 255) / 256), dim3(256), 0, st, a, g);
    return hipGetLastError();
}
template <class S, class T> hipError_t launch_ghost_from_pops(const StepArgs<S, T> &a, T *g, hipStream_t st)
{
    const int L = 2 * (a.nx + a.ny);
    hipLaunchKernelGGL((ghost_pops_kernel<S, T>), dim3((L + 255) / 256), dim3(256), 0, st, a, g);
    return hipGetLastError();
}

#define MRLBM_INSTANTIATE(S, T)                                                                  \
    template hipError_t launch_bulk<S, T>(const StepArgs<S, T> &, int, hipStream_t, hipEvent_t, hipEvent_t); \
    template hipError_t launch_boundary<S, T>(const StepArgs<S, T> &, hipStream_t);              \
    template hipError_t launch_edges<S, T>(const StepArgs<S, T> &, hipStream_t);                 \
    template hipError_t launch_stencil_moments<S, T>(const StepArgs<S, T> &, hipStream_t);       \
    template hipError_t launch_bulk2<S, T>(const StepArgs<S, T> &, int, int, hipStream_t, hipEvent_t, hipEvent_t); \
    template hipError_t launch_bulkn<S, T>(const StepArgs<S, T> &, int, int, hipStream_t, hipEvent_t, hipEvent_t); \
    template hipError_t launch_side<S, T>(const StepArgs<S, T> &, hipStream_t, bool);            \
    template hipError_t launch_wall<S, T>(const StepArgs<S, T> &, hipStream_t, bool);            \
    template hipError_t launch_records4<S, T>(const StepArgs<S, T> &, const RecSteps &, hipStream_t); \
    template hipError_t launch_bulk_batch<S, T>(const StepArgs<S, T> *, const int2 *, int, int, int, hipStream_t); \
    template hipError_t launch_side_tiles<S, T>(const StepArgs<S, T> *, const int2 *, int, const int2 *, int, int, bool, \
                                                bool, bool, hipStream_t);                            \
    template hipError_t launch_bulk2_batch<S, T>(const StepArgs<S, T> *, const int2 *, int, int, hipStream_t); \
    template hipError_t launch_bulkn_batch<S, T>(const StepArgs<S, T> *, const int2 *, int, int, hipStream_t); \
    template bool bulkn_rows_ok<S>(int);                                                         \
    template int bulk_batch_resident_blocks<S, T>(int);                                          \
    template hipError_t launch_inlet_fold<S, T>(const StepArgs<S, T> &, hipStream_t);            \
    template hipError_t launch_pair_probes<S, T>(const StepArgs<S, T> &, const S *, int, hipStream_t); \
    template hipError_t launch_pair_samples<S, T>(const StepArgs<S, T> &, const S *, hipStream_t); \
    template hipError_t launch_quad_probes<S, T>(const StepArgs<S, T> &, hipStream_t);            \
    template hipError_t launch_get_populations<S, T>(const StepArgs<S, T> &, double *, int, int, hipStream_t); \
    template hipError_t launch_ghost_from_state<S, T>(const StepArgs<S, T> &, T *, hipStream_t);  \
    template hipError_t launch_ghost_from_pops<S, T>(const StepArgs<S, T> &, T *, hipStream_t);
MRLBM_INSTANTIATE(double, double)
MRLBM_INSTANTIATE(float, float)

} // namespace mrlbm
