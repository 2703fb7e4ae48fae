// mrlbm_handle.h -- the handle behind the C-ABI (internal to the library).
#pragma once
#include "mrlbm.h"
#include "mrlbm_geometry.h"
#include "mrlbm_kernels.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace mrlbm {
struct AdjointWork;
void adjoint_free(mrlbm_handle *h);
// Bulk decomposition for a kernel choice (see mrlbm_params::kernel): variant, strip height,
// wave columns, strips, and the per-strip boundary/solid row masks of the cylinder box.
int set_bulk_shape(mrlbm_handle *h, int kernel, int strip_rows, std::string &why);
} // namespace mrlbm

using namespace mrlbm;

template <class T> inline T *dmalloc(size_t n, hipError_t &e)
{
    void *p = nullptr;
    e = hipMalloc(&p, n ? n * sizeof(T) : sizeof(T));
    return (T *)p;
}
constexpr int kRedWords = 4096;
// u statistics accumulators (FP64 planes): sum ux, sum uy, sum ux^2, sum uy^2
constexpr int kAccPlanes = 4;
// Inlet-density buffers of the window steps (by iteration modulo the count): a four-step
// pass folds them into the running mean after its side work, so five must be held at once.
constexpr int kInletRing = 8;
static_assert((kInletRing & (kInletRing - 1)) == 0, "the ring index is iter & (kInletRing - 1)");
// Slots of the wall's per-node force and pressure arrays: one per step of a four-step pass.
constexpr int kWallSlots = 4;

struct mrlbm_handle {
    mrlbm_params p{};
    std::string err;
    int nx = 0, ny = 0, pitch = 0;
    long long plane = 0;
    bool f32 = false;
    Geometry geo;
    BoundaryTables bt;
    bool has_b = false;
    double nu = 0, omega = 0;
    int device = 0;             // HIP device current at mrlbm_create (the handle's memory lives there)
    hipStream_t st = nullptr;   // bulk kernel and every host-facing operation
    hipStream_t st2 = nullptr;  // side work (ring, bands, stencil, wall), concurrent with the bulk
    hipEvent_t evA = nullptr, evB = nullptr; // end of the last bulk / side work
    hipEvent_t evK = nullptr;      // stop event given to untimed bulk launches
    hipEvent_t bulk_done = nullptr; // the last bulk launch's stop event (evA before any)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int timing = 0; // 0 off, else time every timing-th bulk launch
    double last_ms = 0;
    // timing mode: an event pair around each bulk launch on stream st (reused, grown on demand)
    std::vector<hipEvent_t> tev;
    double last_bulk_ms = 0;
    int last_bulk_n = 0;
    mrlbm_timing tstat{}; // bulk launches of the last mrlbm_step call
    void *S[2] = {nullptr, nullptr};
    int cur = 0;
    void *pops = nullptr;
    bool from_pops = false;
    void *ghost = nullptr;
    void *uin = nullptr;
    uint8_t *d_box = nullptr;
    unsigned long long *d_boxbits = nullptr;
    std::vector<void *> allocs; // boundary tables and scratch
    BoundaryDev bd{};
    double *d_rec = nullptr;
    int rec_cap = 0, rec_stride = 0;
    double *d_inlet[kInletRing] = {}; // inlet densities of window steps, by iter % kInletRing
    double *d_rho_mean = nullptr;
    double *d_acc = nullptr;
    int64_t nsamples = 0;
    double *d_red = nullptr; // reduction scratch of kRedWords doubles
    int prev_inlet_valid = 0, prev_inlet_buf = 0;
    double prev_inlet_c = 0;
    int strip_rows = 32, ncg = 0, ncb = 0, nstrips = 0;
    int variant = 0; // bulk kernel variant (see StepArgs::variant)
    // host-side queues of recorded series
    std::vector<int32_t> q_force_iter;
    std::vector<double> q_fd, q_fl;
    std::vector<int32_t> q_probe_iter;
    std::vector<double> q_probe;
    std::vector<int32_t> q_surf_iter;
    std::vector<double> q_ps, q_pfit;
    int last_p_iter = -1;
    std::vector<double> last_ps, last_pfit;
    int last_nprobe = 0;
    void *chi = nullptr;      // Brinkman solid fraction per node (step type), or null
    void *S2 = nullptr;       // scratch state of the two-step pass (state t+1 on the bands)
    int tb_H = 0, nstrips2 = 0, ncg2 = 0, ncb2 = 0; // two-step pass: strip height, strips, 64-column waves, blocks per strip
    double *d_dense = nullptr;              // nx*ny FP64 staging for the host copies (on first use)
    int rect1[5][4] = {}, rect2[5][4] = {}; // band rectangles of sub-steps 1 and 2
    int nrect1 = 0, nrect2 = 0;
    // four-step pass: a second scratch state (sub-steps 1..3 alternate between S2 and S3), the
    // strip height, strips, waves (56 written columns each), blocks per strip, and the band
    // rectangles of sub-steps 1..4
    void *S3 = nullptr;
    int ms_H = 0, nstripsN = 0, ncgN = 0, ncbN = 0;
    int rectq[4][5][4] = {}, nrectq[4] = {};
    AdjointWork *adj = nullptr; // buffers of the discrete adjoint, allocated on first use
};

#define FAIL(code, msg)            \
    do {                           \
        h->err = (msg);            \
        return (code);             \
    } while (0)
#define HIPCHK(...)                                                                     \
    do {                                                                                \
        hipError_t e_ = (__VA_ARGS__);                                                  \
        if (e_ != hipSuccess) FAIL(MRLBM_E_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class S, class T> inline StepArgs<S, T> base_args(const mrlbm_handle *h)
{
    StepArgs<S, T> a{};
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.plane = h->plane;
    a.xper = h->p.x_periodic;
    a.yper = h->p.y_periodic;
    a.src = (const S *)h->S[h->cur];
    a.dst = (S *)h->S[h->cur ^ 1];
    a.pops = h->from_pops ? (const T *)h->pops : nullptr;
    a.ghost = (const T *)h->ghost;
    a.om = (T)h->omega;
    a.om1 = (T)(1.0 - h->omega);
    a.vin = (T)0.0;
    a.uin = (const T *)h->uin;
    a.incomp_bc = h->p.incomp_bc;
    a.incomp_bulk = h->p.incomp_bulk;
    a.bx0 = h->bt.bx0;
    a.by0 = h->bt.by0;
    a.bw = h->bt.bw;
    a.bh = h->bt.bh;
    a.boxcls = h->d_box;
    a.boxbits = h->d_boxbits;
    a.strip_rows = h->strip_rows;
    a.ncg = h->ncg;
    a.ncb = h->ncb;
    a.nstrips = h->nstrips;
    a.has_boundary = h->has_b ? 1 : 0;
    a.simple = h->p.kernel == 1 ? 1 : 0;
    a.variant = h->variant;
    a.xcd_order = (h->p.block_order == 1) ? 1 : 0;
    a.rec_slot = -1;
    a.rec = h->d_rec;
    a.rec_stride = h->rec_stride;
    a.acc = h->d_acc;
    a.rho_inf_mean = h->d_rho_mean;
    a.b = h->bd;
    a.chi = (const T *)h->chi;
    return a;
}

