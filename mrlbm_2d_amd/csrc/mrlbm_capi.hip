// mrlbm_capi.hip -- the C-ABI of include/mrlbm.h: handles, device memory, the step loop.
#include "mrlbm_handle.h"
#include "mrlbm_lattice.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {
thread_local std::string g_create_error;
} // namespace

// Staging array of mrlbm_get_populations_noalloc: one per GPU, in the library's code object,
// held until the process ends.  Host code never uses its name as an address (in host code
// the name is a stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once per device at mrlbm_create with that device current.
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)1 << 20; // 8 MiB of FP64
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_WORDS];
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `dev`, which must be the current device.
hipError_t scratch_for_device(int dev, double **out)
{
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
} // namespace

extern "C" int mrlbm_abi_version(void) { return MRLBM_ABI_VERSION; }

extern "C" void mrlbm_default_params(mrlbm_params *p)
{
    std::memset(p, 0, sizeof *p);
    // defaults of the reference's input.dat (input.dat:1-45)
    p->x_periodic = 0;
    p->y_periodic = 1;
    p->channel_with_cylinder = 1;
    p->Re = 100.0;
    p->uo = 0.1;
    p->omega = 0.0;
    p->inlet_profile = MRLBM_INLET_UNIFORM;
    p->mass_bc = MRLBM_MASS_STRONG;
    p->rotated_coordinate = 1;
    p->vel_interp = 1;
    p->mom_interp = 1;
    p->incomp_bc = 0;
    p->incomp_bulk = 0;
    p->precision = MRLBM_FP64;
    p->label_policy = MRLBM_LABEL_REFERENCE;
    p->time_blocking = -1;
}

extern "C" void mrlbm_default_step_opts(mrlbm_step_opts *o)
{
    std::memset(o, 0, sizeof *o);
    o->post_process = 0;
    o->statsbegin = 1 << 30;
    o->statsend = -1;
    o->iplot = 0;
}

extern "C" const char *mrlbm_last_error(const mrlbm_handle *h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

template <class T> static int upload(mrlbm_handle *h, void **dst, const std::vector<T> &v)
{
    hipError_t e;
    T *p = dmalloc<T>(v.size(), e);
    HIPCHK(e);
    h->allocs.push_back(p);
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = p;
    return 0;
}

static int ensure_rec(mrlbm_handle *h)
{
    if (h->d_rec) return 0;
    h->rec_stride = 14 + 2 * h->bt.nbct;
    h->rec_cap = 256;
    hipError_t e;
    h->d_rec = dmalloc<double>((size_t)h->rec_cap * h->rec_stride, e);
    HIPCHK(e);
    HIPCHK(hipMemset(h->d_rec, 0, (size_t)h->rec_cap * h->rec_stride * sizeof(double)));
    return 0;
}

template <class S, class T> static int init_rest_state(mrlbm_handle *h)
{
    // rho = rho0 = 1, u = 0 (2d_mrlbm.f90:199-204): stored moments (1 or 0, 0, 0, 0, 0, 0)
    std::vector<S> row((size_t)h->plane, (S)0);
    for (int b = 0; b < 2; b++) {
        const S r0 = h->f32 ? (S)0 : (S)1;
        std::vector<S> rp((size_t)h->plane, r0);
        HIPCHK(hipMemcpy(h->S[b], rp.data(), h->plane * sizeof(S), hipMemcpyHostToDevice));
        for (int q = 1; q < 6; q++)
            HIPCHK(hipMemcpy((S *)h->S[b] + q * h->plane, row.data(), h->plane * sizeof(S), hipMemcpyHostToDevice));
    }
    std::vector<T> g((size_t)Q * 2 * (h->nx + h->ny));
    for (size_t t = 0; t < g.size() / Q; t++)
        for (int k = 0; k < Q; k++) g[t * Q + k] = h->f32 ? (T)0 : (T)wk<double>(k);
    HIPCHK(hipMemcpy(h->ghost, g.data(), g.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <class S, class T> static int create_typed(mrlbm_handle *h)
{
    hipError_t e;
    const size_t state_bytes = (size_t)6 * h->plane * sizeof(S);
    for (int b = 0; b < 2; b++) {
        HIPCHK(hipMalloc(&h->S[b], state_bytes));
    }
    HIPCHK(hipMalloc(&h->ghost, (size_t)Q * 2 * (h->nx + h->ny) * sizeof(T)));
    // inlet profile (uniform uo in the reference, 2d_mrlbm.f90:258)
    std::vector<T> uin(h->ny);
    for (int j = 0; j < h->ny; j++) {
        if (h->p.inlet_profile == MRLBM_INLET_PARABOLIC) {
            const double s = (double)j / (double)(h->ny - 1);
            uin[j] = (T)(4.0 * h->p.uo * s * (1.0 - s));
        } else {
            uin[j] = (T)h->p.uo;
        }
    }
    HIPCHK(hipMalloc(&h->uin, h->ny * sizeof(T)));
    HIPCHK(hipMemcpy(h->uin, uin.data(), h->ny * sizeof(T), hipMemcpyHostToDevice));
    for (int b = 0; b < kInletRing; b++) {
        h->d_inlet[b] = dmalloc<double>(h->ny, e);
        HIPCHK(e);
    }
    h->d_rho_mean = dmalloc<double>(1, e);
    HIPCHK(e);
    HIPCHK(hipMemset(h->d_rho_mean, 0, sizeof(double)));
    h->d_red = dmalloc<double>(kRedWords, e);
    HIPCHK(e);
    if (h->has_b) {
        const BoundaryTables &t = h->bt;
        BoundaryDev &b = h->bd;
        b.nbct = t.nbct;
        b.nU = t.nU;
        b.nlevels = t.nlevels;
        int rc = 0;
        void *p;
        rc |= upload(h, &p, t.node); b.node = (const int32_t *)p;
        rc |= upload(h, &p, t.ij); b.ij = (const int32_t *)p;
        rc |= upload(h, &p, t.label); b.label = (const int32_t *)p;
        rc |= upload(h, &p, t.uself); b.uself = (const int32_t *)p;
        rc |= upload(h, &p, t.uc); b.uc = (const int32_t *)p;
        rc |= upload(h, &p, t.dep); b.dep = (const int32_t *)p;
        rc |= upload(h, &p, t.level_order); b.level_order = (const int32_t *)p;
        rc |= upload(h, &p, t.level_start); b.level_start = (const int32_t *)p;
        rc |= upload(h, &p, t.real); b.real = (const double *)p;
        rc |= upload(h, &p, t.unode); b.unode = (const int32_t *)p;
        rc |= upload(h, &p, t.uij); b.uij = (const int32_t *)p;
        rc |= upload(h, &p, t.urot); b.urot = (const double *)p;
        rc |= upload(h, &p, t.idx); b.idx = (const int32_t *)p;
        rc |= upload(h, &p, t.thetas); b.thetas = (const double *)p;
        rc |= upload(h, &p, t.boxcls); h->d_box = (uint8_t *)p;
        WallGroups wg;
        if (build_wall_groups(t, WALL_GROUP, WALL_GROUP_MAXU, wg)) {
            rc |= upload(h, &p, wg.off); b.wg_off = (const int32_t *)p;
            rc |= upload(h, &p, wg.list); b.wg_list = (const int32_t *)p;
            rc |= upload(h, &p, wg.node); b.wg_node = (const int32_t *)p;
            rc |= upload(h, &p, wg.slot); b.wl_slot = (const int16_t *)p;
            b.nwg = (int)wg.off.size() - 1;
        }
        if (rc) return rc;
        label_masks(h->p.label_policy, b.incs, b.outs);
        for (int l = 0; l < 16; l++) { // rest-population part of the momentum exchange (FP32 form)
            double fx = 0, fy = 0;
            for (int k = 0; k < Q; k++) {
                const double w = wk<double>(k);
                if ((b.incs[l] >> k) & 1) { fx += w * EXk(k); fy += w * EYk(k); }
                if ((b.outs[l] >> k) & 1) { fx -= w * EXk(k); fy -= w * EYk(k); }
            }
            b.wfx[l] = fx;
            b.wfy[l] = fy;
        }
        b.mass_bc = h->p.mass_bc;
        b.rotated = h->p.rotated_coordinate;
        b.vel_interp = h->p.vel_interp;
        b.mom_interp = h->p.mom_interp;
        b.U = dmalloc<T>((size_t)5 * t.nU, e);
        HIPCHK(e);
        h->allocs.push_back(b.U);
        HIPCHK(hipMemset(b.U, 0, (size_t)5 * t.nU * sizeof(T)));
        b.ub = dmalloc<T>((size_t)2 * t.nbct, e);
        HIPCHK(e);
        h->allocs.push_back(b.ub);
        b.pcy = dmalloc<double>((size_t)kWallSlots * t.nbct, e); // slot 0 unless in a four-step pass
        HIPCHK(e);
        h->allocs.push_back(b.pcy);
        b.fxy = dmalloc<double>((size_t)kWallSlots * 2 * t.nbct, e);
        HIPCHK(e);
        h->allocs.push_back(b.fxy);
        // the surface-pressure smoothing weights (nbct^2 doubles; up to 1 GiB, else computed
        // in every recording step)
        if ((size_t)t.nbct * t.nbct * sizeof(double) <= ((size_t)1 << 30)) {
            double *gw = dmalloc<double>((size_t)t.nbct * t.nbct, e);
            HIPCHK(e);
            h->allocs.push_back(gw);
            double *gsw = dmalloc<double>(t.nbct, e);
            HIPCHK(e);
            h->allocs.push_back(gsw);
            HIPCHK(launch_smooth_weights(b, gw, gsw, h->st));
            HIPCHK(hipStreamSynchronize(h->st));
            b.gw = gw;
            b.gsw = gsw;
        }
    }
    return init_rest_state<S, T>(h);
}

static void free_all(mrlbm_handle *h)
{
    adjoint_free(h);
    for (void *p : {h->S[0], h->S[1], h->pops, h->ghost, h->uin, (void *)h->d_boxbits, (void *)h->d_rec,
                    (void *)h->d_rho_mean, (void *)h->d_acc, (void *)h->d_red, h->chi, h->S2, h->S3, (void *)h->d_dense})
        if (p) (void)hipFree(p);
    for (double *p : h->d_inlet)
        if (p) (void)hipFree(p);
    for (void *p : h->allocs)
        if (p) (void)hipFree(p);
    h->allocs.clear();
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
    h->tev.clear();
    for (hipEvent_t e : {h->evA, h->evB, h->evK})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t q : {h->st, h->st2})
        if (q) (void)hipStreamDestroy(q);
}

int mrlbm::set_bulk_shape(mrlbm_handle *h, int kernel_in, int strip_rows, std::string &why)
{
    // 0 auto picks the variant measured fastest at 8192 x 2048 on MI355X
    const int kernel = kernel_in == 0 ? (h->f32 ? 304 : 8) : kernel_in;
    const bool aligned = kernel == 304 || kernel == 308 || kernel == 316;
    int H = strip_rows;
    if (kernel == 4 || kernel == 8 || kernel == 16) {
        H = kernel; // unrolled strip of that many rows
    } else if (aligned) {
        H = kernel - 300; // aligned 64-column waves, unrolled strip of kernel-300 rows
    } else if (kernel != 1 && kernel != 2) {
        why = "kernel must be 0, 1, 2, 4, 8, 16, 304, 308 or 316";
        return MRLBM_E_ARG;
    }
    // waves of 64 (aligned) or 62 written columns, strips of rows
    const int ncg = aligned ? (h->nx + 63) / 64 : (h->nx + 61) / 62;
    if (H <= 0) {
        // aim for >= 8 bulk waves per SIMD slot of the 256 CUs, strips of 8..64 rows
        const long long target = 256LL * 4 * 8;
        H = 64;
        while (H > 8 && (long long)ncg * ((h->ny - 2 + H - 1) / H) < target) H /= 2;
    }
    if (H > 64) {
        why = "strip_rows must be <= 64";
        return MRLBM_E_ARG;
    }
    h->p.kernel = kernel_in;
    h->p.strip_rows = strip_rows;
    h->variant = (kernel == 1 || kernel == 2) ? 0 : kernel;
    h->ncg = ncg;
    h->ncb = (ncg + 3) / 4;
    h->strip_rows = H;
    h->nstrips = (h->ny - 2 + H - 1) / H; // interior rows 1..ny-2; the edge kernel does the ring
    if (h->bt.bw > 0) { // per box column and bulk strip: 64-bit row masks of boundary, solid nodes
        std::vector<unsigned long long> bits((size_t)h->bt.bw * h->nstrips * 2, 0ull);
        for (int bi = 0; bi < h->bt.bw; bi++)
            for (int bj = 0; bj < h->bt.bh; bj++) {
                const int j = h->bt.by0 + bj;
                if (j < 1 || j > h->ny - 2) continue;
                const int st = (j - 1) / h->strip_rows, t = (j - 1) % h->strip_rows;
                const uint8_t c = h->bt.boxcls[(size_t)bj * h->bt.bw + bi];
                if (c & 1) bits[((size_t)bi * h->nstrips + st) * 2] |= 1ull << t;
                if (c & 2) bits[((size_t)bi * h->nstrips + st) * 2 + 1] |= 1ull << t;
            }
        if (h->d_boxbits) (void)hipFree(h->d_boxbits);
        h->d_boxbits = nullptr;
        if (hipMalloc(&h->d_boxbits, bits.size() * sizeof(unsigned long long)) != hipSuccess ||
            hipMemcpy(h->d_boxbits, bits.data(), bits.size() * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess) {
            why = "box mask upload failed";
            return MRLBM_E_HIP;
        }
    }
    return 0;
}

// Band rectangles of a sub-step that writes the L layers inside the outer ring and the cylinder
// box dilated by L, the latter grown to hold the wall's stencil nodes dilated by G (G = 0: not
// grown).  A sub-step's wall kernels read the previous sub-step's state on the stencil nodes
// and their neighbours, and the plain updates there read one layer further out in the
// sub-step before: with n sub-steps, sub-step s holds the stencil nodes dilated by n - s.
// Returns how many rectangles.
static int band_rects(const mrlbm_handle *h, int L, int G, int r[5][4])
{
    const int nx = h->nx, ny = h->ny;
    auto put = [](int q[4], int x0, int y0, int x1, int y1) { q[0] = x0; q[1] = y0; q[2] = x1 - x0 + 1; q[3] = y1 - y0 + 1; };
    put(r[0], 1, 1, nx - 2, L);
    put(r[1], 1, ny - 1 - L, nx - 2, ny - 2);
    put(r[2], 1, L + 1, L, ny - 2 - L);
    put(r[3], nx - 1 - L, L + 1, nx - 2, ny - 2 - L);
    if (h->bt.bw <= 0) return 4;
    const BoundaryTables &t = h->bt;
    int x0 = t.bx0 - L, x1 = t.bx0 + t.bw - 1 + L, y0 = t.by0 - L, y1 = t.by0 + t.bh - 1 + L;
    if (G > 0)
        for (int u = 0; u < t.nU; u++) {
            x0 = std::min(x0, t.uij[2 * u] - G);
            x1 = std::max(x1, t.uij[2 * u] + G);
            y0 = std::min(y0, t.uij[2 * u + 1] - G);
            y1 = std::max(y1, t.uij[2 * u + 1] + G);
        }
    put(r[4], std::max(x0, 1), std::max(y0, 1), std::min(x1, nx - 2), std::min(y1, ny - 2));
    return 5;
}

// Two-step pass: the scratch state and the band rectangles.  Sub-step 1 (t -> t+1 into the
// scratch) covers the two layers inside the ring and the cylinder box grown to hold every
// node the wall kernels and sub-step 2 read; sub-step 2 (t+1 -> t+2) covers the layer
// inside the ring and the cylinder box grown by one.  The ring and the boundary nodes are
// done by the edge and wall kernels in each sub-step; the two-step bulk does the rest.
static int setup_time_blocking(mrlbm_handle *h, int H)
{
    const int nx = h->nx, ny = h->ny;
    HIPCHK(hipMalloc(&h->S2, (size_t)6 * h->plane * (h->f32 ? sizeof(float) : sizeof(double))));
    h->tb_H = H;
    h->nstrips2 = (ny - 4 + H - 1) / H; // output rows 2..ny-3
    h->ncg2 = (nx + 63) / 64;            // aligned 64-column waves whatever the single-step variant
    h->ncb2 = (h->ncg2 + 3) / 4;
    h->nrect1 = band_rects(h, 2, 1, h->rect1);
    h->nrect2 = band_rects(h, 1, 0, h->rect2);
    return 0;
}

// Four-step pass: sub-step s = 1..4 writes state t+s on the 7-s layers inside the ring and
// the cylinder box dilated by 7-s (grown to the stencil nodes dilated by 4-s), into S2, S3, S2
// and the destination; the four-step bulk writes the nodes four or more from the ring and
// the box.  Strips of 20 rows in FP32 (at 8192 x 2048 1.2% faster per step than the 24 rows
// of a sweep over 12..52), 40 in FP64 (5% faster than 20; a sweep over 8..48), each taken
// while it leaves at least 4096 waves, then 20, then 12 (FP32) or 8 (FP64) rows; or the
// caller's quad_rows.
static int setup_four_step(mrlbm_handle *h, int quad_rows)
{
    const int nx = h->nx, ny = h->ny;
    HIPCHK(hipMalloc(&h->S3, (size_t)6 * h->plane * (h->f32 ? sizeof(float) : sizeof(double))));
    const int tallest = h->f32 ? 20 : 40, narrow = h->f32 ? 12 : 8;
    h->ncgN = (nx - 8 + 55) / 56; // output columns 4..nx-5, 56 per wave
    h->ncbN = (h->ncgN + 3) / 4;
    auto waves = [&](int H) { return (long long)h->ncgN * ((ny - 8 + H - 1) / H); };
    h->ms_H = waves(tallest) >= 4096 ? tallest : waves(20) >= 4096 ? 20 : narrow;
    if (quad_rows > 0) h->ms_H = quad_rows;
    h->nstripsN = (ny - 8 + h->ms_H - 1) / h->ms_H; // output rows 4..ny-5
    for (int s = 1; s <= 4; s++) h->nrectq[s - 1] = band_rects(h, 7 - s, 4 - s, h->rectq[s - 1]);
    return 0;
}

extern "C" int mrlbm_create(const mrlbm_params *p, mrlbm_handle **out)
{
    g_create_error.clear();
    if (!p || !out) {
        g_create_error = "null argument";
        return MRLBM_E_ARG;
    }
    *out = nullptr;
    auto *h = new mrlbm_handle();
    h->p = *p;
    auto fail = [&](int code, const std::string &m) {
        g_create_error = m.empty() ? h->err : m;
        free_all(h);
        delete h;
        return code;
    };
    if (p->nx < 4 || p->ny < 4) return fail(MRLBM_E_ARG, "nx and ny must be >= 4");
    if (p->precision != MRLBM_FP64 && p->precision != MRLBM_FP32_DEV) return fail(MRLBM_E_ARG, "unknown precision");
    if (p->mass_bc < 0 || p->mass_bc > 2) return fail(MRLBM_E_ARG, "unknown mass_bc");
    // boundary tables and probes hold node indices i + pitch * j (pitch = nx rounded up to 64)
    // as 32-bit integers
    if (((long long)p->nx + 63) / 64 * 64 * p->ny > (1LL << 31) - 1)
        return fail(MRLBM_E_ARG, "grid too large for 32-bit node index");
    {
        const int k = p->kernel;
        if (k != 0 && k != 1 && k != 2 && k != 4 && k != 8 && k != 16 && k != 304 && k != 308 && k != 316)
            return fail(MRLBM_E_ARG, "kernel must be 0, 1, 2, 4, 8, 16, 304, 308 or 316");
        const int tb = p->time_blocking;
        if (tb != -1 && (tb < 0 || tb > 16 || tb % 2))
            return fail(MRLBM_E_ARG, "time_blocking must be -1 or an even number 0..16");
        if (p->strip_rows < 0 || p->strip_rows > 64) return fail(MRLBM_E_ARG, "strip_rows must be 0..64");
    }
    h->nx = p->nx;
    h->ny = p->ny;
    h->pitch = (p->nx + 63) / 64 * 64;
    h->plane = (long long)h->pitch * p->ny;
    h->f32 = p->precision == MRLBM_FP32_DEV;
    if (p->channel_with_cylinder) {
        GeometryInput gi{p->nx, p->ny, p->x_c, p->y_c, p->r_c, p->label_policy};
        std::string m = build_geometry(gi, h->geo);
        if (!m.empty()) return fail(MRLBM_E_GEOMETRY, m);
        m = build_tables(h->geo, h->pitch, h->bt);
        if (!m.empty()) return fail(MRLBM_E_GEOMETRY, m);
        h->has_b = h->bt.nbct > 0;
    }
    // viscosity and relaxation (2d_mrlbm.f90:133-135)
    const double uref = p->inlet_profile == MRLBM_INLET_PARABOLIC ? (2.0 / 3.0) * p->uo : p->uo;
    if (p->channel_with_cylinder) h->nu = uref * (2.0 * h->geo.r_cyl) / p->Re;
    else h->nu = uref * (double)(p->ny - 1) / p->Re;
    h->omega = p->omega > 0 ? p->omega : 1.0 / ((3.0 * h->nu) + (1.0 / 2.0));
    if (p->omega > 0) h->nu = (1.0 / p->omega - 0.5) / 3.0;
    if (hipGetDevice(&h->device) != hipSuccess) return fail(MRLBM_E_HIP, "hipGetDevice failed");
    {
        double *sp = nullptr;
        const hipError_t e = scratch_for_device(h->device, &sp);
        if (e != hipSuccess)
            return fail(MRLBM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    if (hipStreamCreate(&h->st) != hipSuccess || hipStreamCreate(&h->st2) != hipSuccess)
        return fail(MRLBM_E_HIP, "hipStreamCreate failed");
    // the stream-ordering events order work on this device only: a device-scope release
    // (no system-scope cache writeback each time one is recorded); so does the bulk launches'
    // stop event evK, which the side stream waits on
    const unsigned order_flags = hipEventDisableTiming | hipEventReleaseToDevice;
    if (hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess ||
        hipEventCreateWithFlags(&h->evA, order_flags) != hipSuccess ||
        hipEventCreateWithFlags(&h->evK, hipEventReleaseToDevice) != hipSuccess ||
        hipEventCreateWithFlags(&h->evB, order_flags) != hipSuccess)
        return fail(MRLBM_E_HIP, "hipEventCreate failed");
    h->bulk_done = h->evA;
    if (hipEventRecord(h->evA, h->st) != hipSuccess || hipEventRecord(h->evB, h->st2) != hipSuccess)
        return fail(MRLBM_E_HIP, "hipEventRecord failed");
    const int rc = h->f32 ? create_typed<float, float>(h) : create_typed<double, double>(h);
    if (rc) return fail(rc, "");
    {
        std::string why;
        const int rc2 = set_bulk_shape(h, p->kernel, p->strip_rows, why);
        if (rc2) return fail(rc2, why);
    }
    {
        // -1 auto: the strip height measured fastest at 8192 x 2048 on MI355X
        const int tb = p->time_blocking < 0 ? (h->f32 ? 10 : 12) : p->time_blocking;
        if (tb < 0 || tb > 16 || tb % 2) return fail(MRLBM_E_ARG, "time_blocking must be -1 or an even number 0..16");
        if (p->steps_per_pass != 0 && p->steps_per_pass != 2 && p->steps_per_pass != 4)
            return fail(MRLBM_E_ARG, "steps_per_pass must be 0, 2 or 4");
        if (p->quad_rows != 0 && !(h->f32 ? bulkn_rows_ok<float>(p->quad_rows) : bulkn_rows_ok<double>(p->quad_rows)))
            return fail(MRLBM_E_ARG, h->f32 ? "quad_rows must be 0, 12, 16, 20, 24, 28 or 32 (FP32)"
                                            : "quad_rows must be 0, 8, 16, 20, 24, 32 or 40 (FP64)");
        if (tb && h->nx >= 16 && h->ny >= 16) {
            const int rc3 = setup_time_blocking(h, tb);
            if (rc3) return fail(rc3, "");
        }
        if (tb && p->steps_per_pass != 2 && h->nx >= 32 && h->ny >= 32) {
            const int rc4 = setup_four_step(h, p->quad_rows);
            if (rc4) return fail(rc4, "");
        }
    }
    *out = h;
    return MRLBM_OK;
}

extern "C" int mrlbm_destroy(mrlbm_handle *h)
{
    if (!h) return MRLBM_E_ARG;
    for (hipStream_t q : {h->st, h->st2})
        if (q) (void)hipStreamSynchronize(q);
    free_all(h);
    delete h;
    return MRLBM_OK;
}

extern "C" int mrlbm_get_info(const mrlbm_handle *h, mrlbm_info *o)
{
    if (!h || !o) return MRLBM_E_ARG;
    std::memset(o, 0, sizeof *o);
    o->nx = h->nx;
    o->ny = h->ny;
    o->nbct = h->bt.nbct;
    o->precision = h->p.precision;
    o->r_cyl = h->geo.r_cyl;
    o->nu = h->nu;
    o->omega = h->omega;
    o->x_c = h->p.x_c;
    o->y_c = h->p.y_c;
    o->r_c = h->p.r_c;
    o->bytes_per_step = (int64_t)12 * h->nx * h->ny * (h->f32 ? 4 : 8);
    o->box[0] = h->bt.bx0;
    o->box[1] = h->bt.by0;
    o->box[2] = h->bt.bw;
    o->box[3] = h->bt.bh;
    o->n_stencil = h->bt.nU;
    o->levels = h->bt.nlevels;
    o->pair_rows = h->tb_H;
    o->quad_rows = h->ms_H;
    return MRLBM_OK;
}

extern "C" int mrlbm_get_boundary(const mrlbm_handle *h, int32_t *bi, int32_t *bj, int32_t *label,
                                  double *theta, int32_t *idx)
{
    if (!h) return MRLBM_E_ARG;
    const auto &bn = h->geo.bnodes;
    for (size_t k = 0; k < bn.size(); k++) {
        if (bi) bi[k] = bn[k].i + 1;
        if (bj) bj[k] = bn[k].j + 1;
        if (label) label[k] = bn[k].label;
        if (theta) theta[k] = bn[k].theta;
        if (idx) idx[k] = h->geo.idx[k] + 1;
    }
    return MRLBM_OK;
}

extern "C" int mrlbm_get_node_class(const mrlbm_handle *h, int8_t *cls)
{
    if (!h || !cls) return MRLBM_E_ARG;
    if (h->geo.cls.empty()) std::memset(cls, 1, (size_t)h->nx * h->ny);
    else std::memcpy(cls, h->geo.cls.data(), h->geo.cls.size());
    return MRLBM_OK;
}

// dense (nx, ny) host <-> pitched device plane
// (the layout change runs on the device: one contiguous host transfer per plane)
static double *dense_staging(mrlbm_handle *h)
{
    if (!h->d_dense && hipMalloc(&h->d_dense, (size_t)h->nx * h->ny * sizeof(double)) != hipSuccess) h->d_dense = nullptr;
    return h->d_dense;
}
template <class S> static int put_plane(mrlbm_handle *h, S *dplane, const double *src, double shift)
{
    double *d = dense_staging(h);
    if (!d) FAIL(MRLBM_E_NOMEM, "no device memory for the host-copy staging plane");
    HIPCHK(hipMemcpyAsync(d, src, (size_t)h->nx * h->ny * sizeof(double), hipMemcpyHostToDevice, h->st));
    HIPCHK(launch_plane_from_dense<S>(d, dplane, h->nx, h->ny, h->pitch, shift, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}
template <class S> static int get_plane(const mrlbm_handle *hc, const S *dplane, double *dst, double shift)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    double *d = dense_staging(h);
    if (!d) FAIL(MRLBM_E_NOMEM, "no device memory for the host-copy staging plane");
    HIPCHK(launch_dense_from_plane<S>(dplane, d, h->nx, h->ny, h->pitch, shift, h->st));
    HIPCHK(hipMemcpyAsync(dst, d, (size_t)h->nx * h->ny * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

template <class S, class T>
static int set_moments_typed(mrlbm_handle *h, const double *rho, const double *ux, const double *uy,
                             const double *mxx, const double *myy, const double *mxy)
{
    const size_t N = (size_t)h->nx * h->ny;
    std::vector<double> m[3];
    for (int q = 0; q < 3; q++) m[q].resize(N);
    for (size_t n = 0; n < N; n++) {
        m[0][n] = mxx ? mxx[n] : ux[n] * ux[n];
        m[1][n] = myy ? myy[n] : uy[n] * uy[n];
        m[2][n] = mxy ? mxy[n] : ux[n] * uy[n];
    }
    S *d = (S *)h->S[h->cur];
    int rc = put_plane<S>(h, d, rho, h->f32 ? 1.0 : 0.0);
    rc |= put_plane<S>(h, d + h->plane, ux, 0.0);
    rc |= put_plane<S>(h, d + 2 * h->plane, uy, 0.0);
    for (int q = 0; q < 3; q++) rc |= put_plane<S>(h, d + (3 + q) * h->plane, m[q].data(), 0.0);
    if (rc) return rc;
    h->from_pops = false;
    auto a = base_args<S, T>(h);
    HIPCHK(launch_ghost_from_state<S, T>(a, (T *)h->ghost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mrlbm_set_moments(mrlbm_handle *h, const double *rho, const double *ux, const double *uy,
                                 const double *mxx, const double *myy, const double *mxy)
{
    if (!h || !rho || !ux || !uy) return MRLBM_E_ARG;
    return h->f32 ? set_moments_typed<float, float>(h, rho, ux, uy, mxx, myy, mxy)
                  : set_moments_typed<double, double>(h, rho, ux, uy, mxx, myy, mxy);
}

template <class S, class T> static int set_pops_typed(mrlbm_handle *h, const double *f)
{
    if (!h->pops) HIPCHK(hipMalloc(&h->pops, (size_t)Q * h->plane * sizeof(T)));
    const size_t N = (size_t)h->nx * h->ny;
    std::vector<T> tmp((size_t)h->pitch * h->ny, (T)0);
    for (int k = 0; k < Q; k++) {
        const double base = h->f32 ? wk<double>(k) : 0.0;
        for (int j = 0; j < h->ny; j++)
            for (int i = 0; i < h->nx; i++)
                tmp[(size_t)j * h->pitch + i] = (T)(f[k * N + (size_t)j * h->nx + i] - base);
        HIPCHK(hipMemcpy((T *)h->pops + k * h->plane, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    h->from_pops = true;
    auto a = base_args<S, T>(h);
    HIPCHK(launch_ghost_from_pops<S, T>(a, (T *)h->ghost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

template <class T> static int set_chi_typed(mrlbm_handle *h, const double *chi)
{
    std::vector<T> tmp((size_t)h->plane, (T)0);
    for (int j = 0; j < h->ny; j++)
        for (int i = 0; i < h->nx; i++) {
            const double v = chi[(size_t)j * h->nx + i];
            if (!(v < 1.0) || !std::isfinite(v)) FAIL(MRLBM_E_ARG, "Brinkman fraction must be finite and below 1");
            tmp[(size_t)j * h->pitch + i] = (T)v;
        }
    if (!h->chi) HIPCHK(hipMalloc(&h->chi, (size_t)h->plane * sizeof(T)));
    HIPCHK(hipMemcpy(h->chi, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int mrlbm_set_brinkman(mrlbm_handle *h, const double *chi)
{
    if (!h) return MRLBM_E_ARG;
    if (!chi) {
        if (h->chi) {
            HIPCHK(hipStreamSynchronize(h->st));
            HIPCHK(hipFree(h->chi));
        }
        h->chi = nullptr;
        return 0;
    }
    return h->f32 ? set_chi_typed<float>(h, chi) : set_chi_typed<double>(h, chi);
}

extern "C" int mrlbm_set_populations(mrlbm_handle *h, const double *f)
{
    if (!h || !f) return MRLBM_E_ARG;
    return h->f32 ? set_pops_typed<float, float>(h, f) : set_pops_typed<double, double>(h, f);
}

static void drain_records(mrlbm_handle *h, const std::vector<int32_t> &iters, int nprobe, int record_p)
{
    const int n = (int)iters.size();
    if (!n) return;
    std::vector<double> buf((size_t)n * h->rec_stride);
    if (hipStreamWaitEvent(h->st, h->evB, 0) != hipSuccess) return; // forces come from the wall stream
    if (hipMemcpyAsync(buf.data(), h->d_rec, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess)
        return;
    (void)hipStreamSynchronize(h->st);
    const int nb = h->bt.nbct;
    for (int s = 0; s < n; s++) {
        const double *r = &buf[(size_t)s * h->rec_stride];
        if (h->has_b) {
            h->q_force_iter.push_back(iters[s]);
            h->q_fd.push_back(r[0]);
            h->q_fl.push_back(r[1]);
        }
        if (nprobe > 0) {
            h->q_probe_iter.push_back(iters[s]);
            for (int q = 0; q < 12; q++) h->q_probe.push_back(r[2 + q]);
        }
        if (record_p && h->has_b) {
            h->q_surf_iter.push_back(iters[s]);
            h->q_ps.insert(h->q_ps.end(), r + 14, r + 14 + nb);
            h->q_pfit.insert(h->q_pfit.end(), r + 14 + nb, r + 14 + 2 * nb);
            h->last_p_iter = iters[s];
            h->last_ps.assign(r + 14, r + 14 + nb);
            h->last_pfit.assign(r + 14 + nb, r + 14 + 2 * nb);
        }
    }
}

template <class S, class T> static int step_typed(mrlbm_handle *h, int iter_first, int nsteps, const mrlbm_step_opts *o)
{
    mrlbm_step_opts opts;
    if (o) opts = *o;
    else mrlbm_default_step_opts(&opts);
    if (opts.nprobe < 0 || opts.nprobe > 4) FAIL(MRLBM_E_ARG, "nprobe must be 0..4");
    int probe_node[4] = {-1, -1, -1, -1};
    for (int p = 0; p < opts.nprobe; p++) {
        const int i = opts.probe_i[p] - 1, j = opts.probe_j[p] - 1;
        if (i < 0 || i >= h->nx || j < 0 || j >= h->ny) FAIL(MRLBM_E_ARG, "probe outside the grid");
        probe_node[p] = i + h->pitch * j;
    }
    if (opts.post_process && opts.stats_interval > 0 && !h->d_acc) {
        HIPCHK(hipMalloc(&h->d_acc, (size_t)kAccPlanes * h->plane * sizeof(double)));
        HIPCHK(hipMemset(h->d_acc, 0, (size_t)kAccPlanes * h->plane * sizeof(double)));
    }
    if (opts.post_process) {
        const int rc = ensure_rec(h);
        if (rc) return rc;
    }
    const int march_blocks = h->ncb * h->nstrips;
    const long long N = (long long)h->nx * h->ny;
    const int simple_blocks = (int)((N + 255) / 256);
    std::vector<int32_t> rec_iters;
    auto is_window = [&](int iter) { return opts.post_process && iter >= opts.statsbegin && iter <= opts.statsend; };
    auto plot_step = [&](int iter) { return opts.post_process && opts.iplot > 0 && iter % opts.iplot == 0; };
    auto sample_step = [&](int iter) {
        return is_window(iter) && opts.stats_interval > 0 && (iter - opts.statsbegin) % opts.stats_interval == 0;
    };
    // kernel arguments of iteration iter reading state src and writing dst
    auto step_args = [&](int iter, const void *src, void *dst, int rec_slot) {
        const bool window = is_window(iter);
        auto a = base_args<S, T>(h);
        a.src = (const S *)src;
        a.dst = (S *)dst;
        a.zero_solid = plot_step(iter) ? 1 : 0;
        a.sample = sample_step(iter) ? 1 : 0;
        a.rec_slot = window ? rec_slot : -1;
        a.record_p = window && opts.record_surface_p ? 1 : 0;
        a.nprobe = opts.nprobe;
        for (int p = 0; p < 4; p++) a.probe_node[p] = probe_node[p];
        a.inlet_out = window ? h->d_inlet[iter & (kInletRing - 1)] : nullptr;
        a.inlet_prev = (h->prev_inlet_valid && h->has_b) ? h->d_inlet[h->prev_inlet_buf] : nullptr;
        a.inlet_prev_c = h->prev_inlet_c;
        a.wall_groups = h->bd.nwg; // the wall nodes as wall groups inside the side launch
        return a;
    };
    // Side work of one iteration on stream B: the ring, band and stencil-node kernel, then the
    // curved wall (wall nodes and, on recording steps, the records) or, without a cylinder,
    // the inlet-density fold.  The wall chain used to run on a third stream beside the ring
    // kernel; every launch beside the bulk costs it time, and the fork and join with their
    // extra launch cost more than the overlap gained (8192 x 2048: 2.2% FP32, 1% FP64).
    auto side_work = [&](const StepArgs<S, T> &a, bool records = true) -> hipError_t {
        hipError_t e = launch_side<S, T>(a, h->st2, a.wall_groups == 0);
        if (e != hipSuccess) return e;
        if (h->has_b) return launch_wall<S, T>(a, h->st2, records);
        if (h->prev_inlet_valid) {
            auto f = a;
            f.inlet_prev = h->d_inlet[h->prev_inlet_buf];
            return launch_inlet_fold<S, T>(f, h->st2);
        }
        return hipSuccess;
    };
    // inlet-density bookkeeping after iteration iter
    auto after_iter = [&](int iter) {
        const bool window = is_window(iter);
        h->prev_inlet_valid = window ? 1 : 0;
        h->prev_inlet_buf = iter & (kInletRing - 1);
        h->prev_inlet_c = (double)(float)(iter - opts.statsbegin); // real() is REAL(4) (:358)
        if (sample_step(iter)) h->nsamples++;
    };
    auto record_iter = [&](int iter) {
        if (!is_window(iter)) return;
        rec_iters.push_back(iter);
        if ((int)rec_iters.size() == h->rec_cap) {
            drain_records(h, rec_iters, opts.nprobe, opts.record_surface_p);
            rec_iters.clear();
        }
    };
    // a pair of iterations may share one two-step bulk pass when neither zeroes the solid
    // nodes' velocity (plot steps) and the records fit
    auto pair_ok = [&](int iter) {
        if (!h->tb_H || h->from_pops || h->chi || h->p.kernel == 1) return false;
        for (int it = iter; it <= iter + 1; it++) {
            if (plot_step(it)) return false;
        }
        const int nrec = (is_window(iter) ? 1 : 0) + (is_window(iter + 1) ? 1 : 0);
        return (int)rec_iters.size() + nrec <= std::max(h->rec_cap, 0) || nrec == 0;
    };
    // four iterations may share one four-step bulk pass when none of them zeroes the solid
    // nodes' velocity (plot steps) and the records fit; the bulk samples the u statistics and
    // records the probes of the nodes it writes, the sub-steps those of the others
    auto quad_ok = [&](int iter) {
        if (!h->ms_H || !h->tb_H || h->from_pops || h->chi || h->p.kernel == 1) return false;
        int nrec = 0;
        for (int it = iter; it <= iter + 3; it++) {
            if (plot_step(it)) return false;
            if (is_window(it)) nrec++;
        }
        return (int)rec_iters.size() + nrec <= std::max(h->rec_cap, 0) || nrec == 0;
    };
    // timing mode: the start and end times of every h->timing-th bulk launch (the first
    // included) into an event pair, at most kMaxTimed launches
    constexpr int kMaxTimed = 4096;
    int ntimed = 0, nlaunch = 0;
    h->last_ms = h->last_bulk_ms = 0.0;
    h->last_bulk_n = 0;
    h->tstat = mrlbm_timing{};
    std::vector<int> timed_steps; // steps advanced by each timed launch (1, 2 or 4)
    if (h->timing) {
        const size_t want = 2 * (size_t)std::min(nsteps / h->timing + 1, kMaxTimed);
        while (h->tev.size() < want) {
            // read after e1 has synchronized with the work; a timed launch's stop event is
            // also the side stream's wait target, so a device-scope release as evK's
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
            h->tev.push_back(e);
        }
        HIPCHK(hipEventRecord(h->e0, h->st));
    }
    // the event pair of the next bulk launch (null when it is not timed)
    auto timed_pair = [&](hipEvent_t &t0, hipEvent_t &t1, int steps) {
        t0 = t1 = nullptr;
        if (h->timing && nlaunch % h->timing == 0 && ntimed < kMaxTimed) {
            t0 = h->tev[2 * ntimed];
            t1 = h->tev[2 * ntimed + 1];
            ntimed++;
            timed_steps.push_back(steps);
        }
        nlaunch++;
        h->tstat.bulk_launches++;
        h->tstat.bulk_steps += steps;
    };
    // A bulk launch ends in its own stop event (a timing event, else evK): the side work of
    // the next step waits on it, with no marker packet between two bulk launches.  The wait
    // on the previous launch's event is enqueued before the next launch reuses evK.
    auto launch_bulk_pair = [&](const StepArgs<S, T> &bulk) -> hipError_t {
        hipEvent_t t0, t1;
        timed_pair(t0, t1, 2);
        if (!t1) t1 = h->evK;
        h->bulk_done = t1;
        return launch_bulk2<S, T>(bulk, h->ncb2 * h->nstrips2, h->tb_H, h->st, t0, t1);
    };
    auto launch_bulk_quad = [&](const StepArgs<S, T> &bulk) -> hipError_t {
        hipEvent_t t0, t1;
        timed_pair(t0, t1, 4);
        if (!t1) t1 = h->evK;
        h->bulk_done = t1;
        return launch_bulkn<S, T>(bulk, h->ncbN * h->nstripsN, h->ms_H, h->st, t0, t1);
    };
    auto launch_bulk_one = [&](const StepArgs<S, T> &a, int nb) -> hipError_t {
        hipEvent_t t0, t1;
        timed_pair(t0, t1, 1);
        if (!t1) t1 = h->evK;
        h->bulk_done = t1;
        return launch_bulk<S, T>(a, nb, h->st, t0, t1);
    };
    for (int s = 0; s < nsteps; s++) {
        const int iter = iter_first + s;
        if (s + 3 < nsteps && quad_ok(iter)) {
            // the bulk: state t -> t+4 on the interior; stream B: sub-steps 1..4 through the
            // scratch states, each the ring, bands, stencil and wall work of its iteration
            void *A = h->S[h->cur], *C = h->S[h->cur ^ 1];
            void *src[4] = {A, h->S2, h->S3, h->S2}, *dst[4] = {h->S2, h->S3, h->S2, C};
            auto bulk = step_args(iter, A, C, -1);
            bulk.ncg = h->ncgN;
            bulk.ncb = h->ncbN;
            bulk.nstrips = h->nstripsN;
            const int slot0 = (int)rec_iters.size();
            for (int k = 0, slot = slot0; k < 4; k++) { // the bulk's samples and probe records
                bulk.qsample |= sample_step(iter + k) ? 1 << k : 0;
                bulk.qslot[k] = is_window(iter + k) ? slot++ : -1;
            }
            HIPCHK(hipStreamWaitEvent(h->st2, h->bulk_done, 0));
            HIPCHK(hipStreamWaitEvent(h->st, h->evB, 0));
            HIPCHK(launch_bulk_quad(bulk));
            if (opts.nprobe > 0 && std::any_of(bulk.qslot, bulk.qslot + 4, [](int q) { return q >= 0; })) {
                // the probes on the bulk's nodes; the next pass's side work (which overwrites
                // state t's band nodes) then waits for this kernel too
                HIPCHK(launch_quad_probes<S, T>(bulk, h->st));
                HIPCHK(hipEventRecord(h->evA, h->st));
                h->bulk_done = h->evA;
            }
            int slot = slot0;
            // a wall of one level records after the pass's side work: sub-step k's wall writes
            // slot k of the per-node force and pressure arrays, and one launch then takes the
            // records of the four steps in order (the inlet densities wait in their ring)
            const bool defer = h->has_b && h->bd.nlevels <= 1;
            RecSteps rs{};
            StepArgs<S, T> last{};
            for (int k = 0; k < 4; k++) {
                auto a = step_args(iter + k, src[k], dst[k], slot);
                if (is_window(iter + k)) slot++;
                a.skip_ns = 4;
                a.nrect = h->nrectq[k];
                std::memcpy(a.rect, h->rectq[k], sizeof h->rectq[k]);
                a.band_march = band_march_rects(a.rect, a.nrect);
                if (defer) {
                    a.b.fxy = h->bd.fxy + (size_t)k * 2 * h->bd.nbct;
                    a.b.pcy = h->bd.pcy + (size_t)k * h->bd.nbct;
                    if (a.rec_slot >= 0 || a.inlet_prev) {
                        rs.rec_slot[rs.n] = a.rec_slot;
                        rs.record_p[rs.n] = a.record_p;
                        rs.wslot[rs.n] = k;
                        rs.inlet_prev[rs.n] = a.inlet_prev;
                        rs.inlet_prev_c[rs.n] = a.inlet_prev_c;
                        rs.n++;
                    }
                    last = a;
                }
                HIPCHK(side_work(a, !defer));
                after_iter(iter + k);
            }
            if (defer && rs.n > 0) {
                last.b.fxy = h->bd.fxy; // the records kernels add the slot offsets
                last.b.pcy = h->bd.pcy;
                HIPCHK(launch_records4<S, T>(last, rs, h->st2));
            }
            HIPCHK(hipEventRecord(h->evB, h->st2));
            h->cur ^= 1;
            for (int k = 0; k < 4; k++) record_iter(iter + k);
            s += 3;
            continue;
        }
        if (s + 1 < nsteps && pair_ok(iter)) {
            void *A = h->S[h->cur], *C = h->S[h->cur ^ 1];
            const int slot1 = (int)rec_iters.size();
            auto a1 = step_args(iter, A, h->S2, slot1);
            auto bulk = a1;
            bulk.dst = (S *)C;
            bulk.nstrips = h->nstrips2;
            bulk.ncg = h->ncg2;
            bulk.ncb = h->ncb2;
            bulk.sample2 = sample_step(iter + 1) ? 1 : 0; // (bulk.sample: of iter)
            bulk.nrect = h->nrect1;                        // what sub-step 1 writes itself
            std::memcpy(bulk.rect, h->rect1, sizeof h->rect1);
            HIPCHK(hipStreamWaitEvent(h->st2, h->bulk_done, 0));
            HIPCHK(hipStreamWaitEvent(h->st, h->evB, 0));
            HIPCHK(launch_bulk_pair(bulk));
            a1.nrect = h->nrect1;
            std::memcpy(a1.rect, h->rect1, sizeof h->rect1);
            a1.band_march = band_march_rects(a1.rect, a1.nrect);
            const int slot2 = is_window(iter + 1) ? slot1 + (is_window(iter) ? 1 : 0) : -1;
            if (opts.nprobe > 0 && (a1.rec_slot >= 0 || slot2 >= 0)) {
                // the probes the bulk carries; the next pair's side work (which overwrites
                // state t's band nodes) then waits for this kernel too
                HIPCHK(launch_pair_probes<S, T>(a1, (const S *)C, slot2, h->st));
                HIPCHK(hipEventRecord(h->evA, h->st));
                h->bulk_done = h->evA;
            }
            HIPCHK(side_work(a1));
            after_iter(iter);
            auto a2 = step_args(iter + 1, h->S2, C, slot1 + (is_window(iter) ? 1 : 0));
            a2.nrect = h->nrect2;
            std::memcpy(a2.rect, h->rect2, sizeof h->rect2);
            a2.band_march = band_march_rects(a2.rect, a2.nrect);
            HIPCHK(side_work(a2));
            if (a1.sample && bulk.sample2) { // after the bulk and sub-step 1 (see the kernel)
                HIPCHK(hipStreamWaitEvent(h->st2, h->bulk_done, 0));
                HIPCHK(launch_pair_samples<S, T>(a1, (const S *)C, h->st2));
            }
            after_iter(iter + 1);
            HIPCHK(hipEventRecord(h->evB, h->st2));
            h->cur ^= 1;
            record_iter(iter);
            record_iter(iter + 1);
            s++;
            continue;
        }
        auto a = step_args(iter, h->S[h->cur], h->S[h->cur ^ 1], (int)rec_iters.size());
        if (h->from_pops) a.pops = (const T *)h->pops;
        const bool pops_step = h->from_pops;
        const int nb = (pops_step || a.simple || a.chi) ? simple_blocks : march_blocks;
        // bulk_n after wall_{n-1}; wall_n after bulk_{n-1}: both read state n-1 only
        HIPCHK(hipStreamWaitEvent(h->st2, h->bulk_done, 0));
        HIPCHK(hipStreamWaitEvent(h->st, h->evB, 0));
        HIPCHK(launch_bulk_one(a, nb));
        HIPCHK(side_work(a));
        HIPCHK(hipEventRecord(h->evB, h->st2));
        after_iter(iter);
        h->cur ^= 1;
        h->from_pops = false;
        record_iter(iter);
    }
    HIPCHK(hipStreamWaitEvent(h->st, h->evB, 0)); // join the wall stream
    if (h->prev_inlet_valid) { // fold the last window step's inlet density now
        auto f = base_args<S, T>(h);
        f.inlet_prev = h->d_inlet[h->prev_inlet_buf];
        f.inlet_prev_c = h->prev_inlet_c;
        HIPCHK(launch_inlet_fold<S, T>(f, h->st));
        h->prev_inlet_valid = 0;
    }
    if (h->timing) {
        HIPCHK(hipEventRecord(h->e1, h->st));
        HIPCHK(hipEventSynchronize(h->e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
        h->last_ms = nsteps > 0 ? ms / nsteps : 0.0;
        double tot = 0;
        std::vector<double> d(ntimed);
        for (int k = 0; k < ntimed; k++) {
            HIPCHK(hipEventElapsedTime(&ms, h->tev[2 * k], h->tev[2 * k + 1]));
            d[k] = ms;
            tot += ms;
            h->tstat.timed_steps += timed_steps[k];
        }
        h->last_bulk_n = ntimed;
        h->last_bulk_ms = ntimed > 0 ? tot / ntimed : 0.0;
        h->tstat.step_ms = h->last_ms;
        h->tstat.timed_launches = ntimed;
        h->tstat.timed_ms = tot;
        if (ntimed > 0) {
            std::sort(d.begin(), d.end());
            h->tstat.bulk_mean_ms = tot / ntimed;
            h->tstat.bulk_median_ms = ntimed % 2 ? d[ntimed / 2] : 0.5 * (d[ntimed / 2 - 1] + d[ntimed / 2]);
            h->tstat.bulk_min_ms = d.front();
            h->tstat.bulk_max_ms = d.back();
        }
    }
    drain_records(h, rec_iters, opts.nprobe, opts.record_surface_p);
    h->last_nprobe = opts.nprobe;
    return 0;
}

extern "C" int mrlbm_step(mrlbm_handle *h, int32_t iter_first, int32_t nsteps, const mrlbm_step_opts *o)
{
    if (!h || nsteps < 0) return MRLBM_E_ARG;
    return h->f32 ? step_typed<float, float>(h, iter_first, nsteps, o)
                  : step_typed<double, double>(h, iter_first, nsteps, o);
}

extern "C" int mrlbm_synchronize(mrlbm_handle *h)
{
    if (!h) return MRLBM_E_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mrlbm_get_moments(const mrlbm_handle *hc, double *rho, double *ux, double *uy, double *mxx,
                                 double *myy, double *mxy)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h) return MRLBM_E_ARG;
    if (h->from_pops) FAIL(MRLBM_E_STATE, "no step since mrlbm_set_populations");
    HIPCHK(hipStreamSynchronize(h->st));
    double *outs[6] = {rho, ux, uy, mxx, myy, mxy};
    for (int q = 0; q < 6; q++) {
        if (!outs[q]) continue;
        const double shift = (q == 0 && h->f32) ? 1.0 : 0.0;
        const int rc = h->f32 ? get_plane<float>(h, (const float *)h->S[h->cur] + q * h->plane, outs[q], shift)
                              : get_plane<double>(h, (const double *)h->S[h->cur] + q * h->plane, outs[q], shift);
        if (rc) FAIL(rc, "device to host copy failed");
    }
    return 0;
}

extern "C" int mrlbm_get_fields(const mrlbm_handle *h, double *rho, double *ux, double *uy)
{
    return mrlbm_get_moments(h, rho, ux, uy, nullptr, nullptr, nullptr);
}

template <class S, class T> static int get_pops_typed(mrlbm_handle *h, double *f)
{
    const size_t N = (size_t)h->nx * h->ny;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)Q * N * sizeof(double)));
    auto a = base_args<S, T>(h);
    hipError_t e = launch_get_populations<S, T>(a, d, 0, h->ny, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e == hipSuccess) e = hipMemcpy(f, d, (size_t)Q * N * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIPCHK(e);
    return 0;
}

extern "C" int mrlbm_get_populations(const mrlbm_handle *hc, double *f)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h || !f) return MRLBM_E_ARG;
    return h->f32 ? get_pops_typed<float, float>(h, f) : get_pops_typed<double, double>(h, f);
}

// Same populations as mrlbm_get_populations, staged through this GPU's `scratch` array in
// bands of whole rows: per band one kernel writes f(nx, rows, 0:8) into the array, then nine
// contiguous copies (one per k) land in f.  Every staged word is written before it is read,
// so nothing needs clearing between calls.  No device or host allocation.
template <class S, class T> static int get_pops_noalloc_typed(mrlbm_handle *h, double *f)
{
    const size_t row_words = (size_t)Q * h->nx;
    if (row_words > NOALLOC_SCRATCH_WORDS)
        FAIL(MRLBM_E_ARG, "nx too large for mrlbm_get_populations_noalloc: one row of 9 populations must fit "
                          "its 8 MiB staging array (nx <= 116508)");
    double *sp = nullptr;
    {
        const hipError_t e = scratch_for_device(h->device, &sp);
        if (e != hipSuccess) FAIL(MRLBM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    const int band = (int)std::min<size_t>(NOALLOC_SCRATCH_WORDS / row_words, (size_t)h->ny);
    const size_t N = (size_t)h->nx * h->ny;
    auto a = base_args<S, T>(h);
    for (int j0 = 0; j0 < h->ny; j0 += band) {
        const int rows = std::min(band, h->ny - j0);
        const size_t bn = (size_t)h->nx * rows;
        HIPCHK(launch_get_populations<S, T>(a, sp, j0, rows, h->st));
        for (int k = 0; k < Q; k++)
            HIPCHK(hipMemcpyAsync(f + k * N + (size_t)j0 * h->nx, sp + k * bn, bn * sizeof(double),
                                  hipMemcpyDeviceToHost, h->st));
        // the next band's kernel overwrites the array: its copies must have landed
        HIPCHK(hipStreamSynchronize(h->st));
    }
    return 0;
}

extern "C" int mrlbm_get_populations_noalloc(mrlbm_handle *h, double *f)
{
    if (!h || !f) return MRLBM_E_ARG;
    return h->f32 ? get_pops_noalloc_typed<float, float>(h, f) : get_pops_noalloc_typed<double, double>(h, f);
}

extern "C" int mrlbm_get_forces(mrlbm_handle *h, int32_t *iters, double *fd, double *fl, int32_t max,
                                int32_t *count)
{
    if (!h || max < 0) return MRLBM_E_ARG;
    const int n = std::min<int>(max, (int)h->q_fd.size());
    for (int s = 0; s < n; s++) {
        if (iters) iters[s] = h->q_force_iter[s];
        if (fd) fd[s] = h->q_fd[s];
        if (fl) fl[s] = h->q_fl[s];
    }
    h->q_force_iter.erase(h->q_force_iter.begin(), h->q_force_iter.begin() + n);
    h->q_fd.erase(h->q_fd.begin(), h->q_fd.begin() + n);
    h->q_fl.erase(h->q_fl.begin(), h->q_fl.begin() + n);
    if (count) *count = n;
    return 0;
}

extern "C" int mrlbm_get_surface_p(const mrlbm_handle *h, int32_t *iter, double *ps, double *pfit)
{
    if (!h) return MRLBM_E_ARG;
    if (h->last_p_iter < 0) return MRLBM_E_STATE;
    if (iter) *iter = h->last_p_iter;
    if (ps) std::copy(h->last_ps.begin(), h->last_ps.end(), ps);
    if (pfit) std::copy(h->last_pfit.begin(), h->last_pfit.end(), pfit);
    return 0;
}

extern "C" int mrlbm_get_surface_p_series(mrlbm_handle *h, int32_t *iters, double *ps, double *pfit,
                                          int32_t max, int32_t *count)
{
    if (!h || max < 0) return MRLBM_E_ARG;
    const int nb = h->bt.nbct;
    const int n = std::min<int>(max, (int)h->q_surf_iter.size());
    for (int s = 0; s < n; s++) {
        if (iters) iters[s] = h->q_surf_iter[s];
        if (ps) std::copy(h->q_ps.begin() + (size_t)s * nb, h->q_ps.begin() + (size_t)(s + 1) * nb, ps + (size_t)s * nb);
        if (pfit)
            std::copy(h->q_pfit.begin() + (size_t)s * nb, h->q_pfit.begin() + (size_t)(s + 1) * nb, pfit + (size_t)s * nb);
    }
    h->q_surf_iter.erase(h->q_surf_iter.begin(), h->q_surf_iter.begin() + n);
    h->q_ps.erase(h->q_ps.begin(), h->q_ps.begin() + (size_t)n * nb);
    h->q_pfit.erase(h->q_pfit.begin(), h->q_pfit.begin() + (size_t)n * nb);
    if (count) *count = n;
    return 0;
}

extern "C" int mrlbm_get_probes(mrlbm_handle *h, int32_t *iters, double *vals, int32_t max, int32_t *count)
{
    if (!h || max < 0) return MRLBM_E_ARG;
    const int n = std::min<int>(max, (int)h->q_probe_iter.size());
    for (int s = 0; s < n; s++) {
        if (iters) iters[s] = h->q_probe_iter[s];
        if (vals) std::copy(h->q_probe.begin() + 12 * s, h->q_probe.begin() + 12 * (s + 1), vals + 12 * s);
    }
    h->q_probe_iter.erase(h->q_probe_iter.begin(), h->q_probe_iter.begin() + n);
    h->q_probe.erase(h->q_probe.begin(), h->q_probe.begin() + 12 * (size_t)n);
    if (count) *count = n;
    return 0;
}

extern "C" int mrlbm_get_rho_inf_mean(const mrlbm_handle *hc, double *v)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h || !v) return MRLBM_E_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(v, h->d_rho_mean, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mrlbm_get_stats(const mrlbm_handle *hc, double *mux, double *muy, double *rux, double *ruy,
                               int64_t *ns)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h) return MRLBM_E_ARG;
    if (ns) *ns = h->nsamples;
    if (!h->d_acc || h->nsamples == 0) FAIL(MRLBM_E_STATE, "no statistics samples");
    HIPCHK(hipStreamSynchronize(h->st));
    const size_t N = (size_t)h->nx * h->ny;
    std::vector<double> s[kAccPlanes];
    for (int q = 0; q < kAccPlanes; q++) {
        s[q].resize(N);
        const int rc = get_plane<double>(h, h->d_acc + q * h->plane, s[q].data(), 0.0);
        if (rc) FAIL(rc, "device to host copy failed");
    }
    const double inv = 1.0 / (double)h->nsamples;
    for (size_t n = 0; n < N; n++) {
        const double mx = s[0][n] * inv, my = s[1][n] * inv;
        if (mux) mux[n] = mx;
        if (muy) muy[n] = my;
        if (rux) rux[n] = std::sqrt(std::max(0.0, s[2][n] * inv - mx * mx));
        if (ruy) ruy[n] = std::sqrt(std::max(0.0, s[3][n] * inv - my * my));
    }
    return 0;
}

extern "C" int mrlbm_reset_stats(mrlbm_handle *h)
{
    if (!h) return MRLBM_E_ARG;
    h->nsamples = 0;
    if (h->d_acc) HIPCHK(hipMemsetAsync(h->d_acc, 0, (size_t)kAccPlanes * h->plane * sizeof(double), h->st));
    return 0;
}

// Largest ux and the non-finite flag of the current state, reduced on the device into the
// handle's reduction scratch (out[0], out[1]).
static int scan_state(mrlbm_handle *h, double out[2])
{
    if (h->from_pops) FAIL(MRLBM_E_STATE, "no step since mrlbm_set_populations");
    HIPCHK(hipStreamWaitEvent(h->st, h->evB, 0)); // the side stream's last writes
    const hipError_t e = h->f32 ? launch_state_scan<float>((const float *)h->S[h->cur], h->plane, h->nx, h->ny,
                                                           h->pitch, h->d_red, kRedWords, h->st)
                                : launch_state_scan<double>((const double *)h->S[h->cur], h->plane, h->nx,
                                                            h->ny, h->pitch, h->d_red, kRedWords, h->st);
    HIPCHK(e);
    HIPCHK(hipMemcpyAsync(out, h->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mrlbm_max_ux(const mrlbm_handle *hc, double *v)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h || !v) return MRLBM_E_ARG;
    double r[2];
    const int rc = scan_state(h, r);
    if (rc) return rc;
    *v = r[0];
    return 0;
}

extern "C" int mrlbm_check_finite(const mrlbm_handle *hc)
{
    auto *h = const_cast<mrlbm_handle *>(hc);
    if (!h) return MRLBM_E_ARG;
    double r[2];
    const int rc = scan_state(h, r);
    if (rc) return rc;
    if (r[1] != 0.0) FAIL(MRLBM_E_NONFINITE, "non-finite moment in the state");
    return 0;
}

extern "C" int mrlbm_set_timing(mrlbm_handle *h, int32_t on)
{
    if (!h) return MRLBM_E_ARG;
    if (on < 0) FAIL(MRLBM_E_ARG, "timing must be 0 (off) or a launch stride >= 1");
    h->timing = on;
    return 0;
}

extern "C" int mrlbm_last_step_ms(const mrlbm_handle *h, double *ms)
{
    if (!h || !ms) return MRLBM_E_ARG;
    *ms = h->last_ms;
    return 0;
}

extern "C" int mrlbm_last_timing(const mrlbm_handle *h, mrlbm_timing *out)
{
    if (!h || !out) return MRLBM_E_ARG;
    *out = h->tstat;
    return 0;
}

extern "C" int mrlbm_last_bulk_ms(const mrlbm_handle *h, double *ms_per_launch, int32_t *launches)
{
    if (!h || !ms_per_launch) return MRLBM_E_ARG;
    *ms_per_launch = h->last_bulk_ms;
    if (launches) *launches = h->last_bulk_n;
    return 0;
}
