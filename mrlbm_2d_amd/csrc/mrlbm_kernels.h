// mrlbm_kernels.h -- kernel argument blocks and launch entry points (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mrlbm {

// Boundary workgroup tables and scratch (device pointers), see mrlbm_geometry.h.
struct BoundaryDev {
    int nbct, nU, nlevels;
    const int32_t *node, *ij, *label, *uself, *uc, *dep, *level_order, *level_start;
    const double *real;
    const int32_t *unode, *uij;
    const double *urot;
    const int32_t *idx;
    const double *thetas;
    const double *gw;  // surface-pressure smoothing weights of (t, q) at [q*nbct + t], or null
    const double *gsw; // their sums over q (sequential), per t
    uint16_t incs[16], outs[16];
    double wfx[16], wfy[16]; // rest-population force per label (FP32 form only)
    int mass_bc, rotated, vel_interp, mom_interp;
    // wall groups (walls of one dependency level): group g takes the boundary nodes
    // wg_node[g*WALL_GROUP ...] (theta order) and the stencil nodes wg_list[wg_off[g] ..
    // wg_off[g+1]) it reads; wl_slot[13*k + c] is the group-local slot of boundary node k's
    // stencil reference c (0..11 the uc corners, 12 the node itself)
    int nwg;
    const int32_t *wg_off, *wg_list, *wg_node;
    const int16_t *wl_slot;
    void *U;      // 5*nU values of the step type: rho, ux, uy, Mxx', Myy' (pre-collision)
    void *ub;     // 2*nbct interpolated velocities
    double *pcy;  // nbct surface pressure, list order
    double *fxy;  // 2*nbct per-node forces
};

template <class S, class T> struct StepArgs {
    int nx, ny, pitch;
    long long plane;
    int xper, yper;
    const S *src;
    S *dst;
    const T *pops;             // non-null: this step reads streamed populations [k*plane + node]
    const T *ghost;            // stale populations: left[ny*9], right[ny*9], bottom[nx*9], top[nx*9]
    T om, om1, vin;
    const T *uin;              // inlet ux per row (ny)
    int incomp_bc, incomp_bulk;
    int bx0, by0, bw, bh;
    const uint8_t *boxcls;     // 1 boundary node (skipped here), 2 solid node
    const unsigned long long *boxbits; // per box column and bulk strip: row bits of boundary, solid nodes
    int zero_solid;
    int strip_rows, ncg, ncb, nstrips, has_boundary, simple;
    int variant;               // bulk kernel: 0 row loop (strip_rows), 4/8/16 unrolled strip of that height,
                               // 304/308/316 aligned 64-column waves, unrolled strip of variant-300 rows
    int xcd_order;             // 1: remap blocks so that each XCD takes consecutive strips of a column
    // recording
    int sample;                // accumulate mean/rms u (two-step bulk: of t+1)
    int sample2;               // two-step bulk: accumulate those of t+2
    int qsample;               // four-step bulk: bit s-1 set = accumulate those of t+s
    int qslot[4];              // four-step bulk: record slot of t+s for the probes, or -1
    int skip_ns;               // sub-step of a four-step pass: leave the samples and probe records of
                               // the nodes the four-step bulk writes (n_step_node, NS = skip_ns) to it
    double *acc;               // 4 planes: ux, uy, ux^2, uy^2
    int rec_slot;              // >= 0: record forces/probes/surface p into rec + slot*rec_stride
    double *rec;
    int rec_stride, record_p;
    int smooth_apart;          // records: the surface-pressure smoothing runs in smooth_kernel
    int nprobe, probe_node[4];
    double *inlet_out;         // post-BC rho at i = 1 (ny), or null
    const double *inlet_prev;  // previous window step's inlet rho, or null
    double inlet_prev_c;       // its mean counter
    double *rho_inf_mean;
    double *facc;              // batched runs: running sums of Fd, Fl, Fl^2 and the step count
    int wall_groups;           // side kernel: the first wall_groups blocks are wall groups (stencil
                               // moments in LDS, then the wall nodes); 0: the stencil nodes are
                               // side items and the wall nodes a launch of their own
    int band_march;            // side kernel: bit q set = band rectangle q goes by row marches, not
                               // node by node
    int nrect;                 // side kernel: band rectangles x0, y0, w, h
    int rect[6][4];
    const T *chi;              // Brinkman solid fraction (per-node bulk kernel only), or null
    BoundaryDev b;
};

// The band rectangles of a multi-step sub-step: 0 and 1 the rows next to the channel walls
// (rows 1..L and ny-1-L..ny-2, columns 1..nx-2), 2 and 3 the columns next to the inlet and
// outlet (L wide), 4 the cylinder's rectangle.  The side launch marches 0, 1 and 4 row by
// row (62 columns and up to BAND_STRIP rows per single-wave workgroup), and 2 and 3 (when at
// most 6 columns wide) row by row in packed segments: 8 segments of 8 lanes per wave, each
// the L columns and their two halo columns of its own BAND_SEG_ROWS rows.  The marched
// rectangles (a bit mask), the workgroups rectangle q's march takes, all marched workgroups,
// and the nodes taken node by node.
constexpr int BAND_STRIP = 16, BAND_SEG_ROWS = 16;
__host__ __device__ constexpr bool band_column_rect(int q) { return q == 2 || q == 3; }
__host__ __device__ inline int band_march_rects(const int (*rc)[4], int n)
{
    const int cols = n >= 4 && rc[2][2] <= 6 && rc[3][2] <= 6 ? 12 : 0;
    return (n >= 2 ? 3 : 0) | cols | (n >= 5 ? 16 : 0);
}
__host__ __device__ inline int band_march_rect_blocks(const int rc[4], int q)
{
    if (band_column_rect(q)) return (rc[3] + 8 * BAND_SEG_ROWS - 1) / (8 * BAND_SEG_ROWS);
    return ((rc[2] + 61) / 62) * ((rc[3] + BAND_STRIP - 1) / BAND_STRIP);
}
__host__ __device__ inline int band_march_blocks(const int rc[][4], int n, int marched)
{
    int b = 0;
    for (int q = 0; q < n; q++)
        if ((marched >> q) & 1) b += band_march_rect_blocks(rc[q], q);
    return b;
}
__host__ __device__ inline long long band_items(const int rc[][4], int n, int marched)
{
    long long t = 0;
    for (int q = 0; q < n; q++)
        if (!((marched >> q) & 1)) t += (long long)rc[q][2] * rc[q][3];
    return t;
}

// Boundary nodes per wall group, and the most stencil nodes a group may read (LDS slots).
constexpr int WALL_GROUP = 32, WALL_GROUP_MAXU = 256;

// Host launchers (mrlbm.hip).
template <class S, class T>
hipError_t launch_bulk(const StepArgs<S, T> &a, int nblocks, hipStream_t st, hipEvent_t t0 = nullptr,
                       hipEvent_t t1 = nullptr);
template <class S, class T>
hipError_t launch_boundary(const StepArgs<S, T> &a, hipStream_t st);
template <class S, class T>
hipError_t launch_edges(const StepArgs<S, T> &a, hipStream_t st);
template <class S, class T>
hipError_t launch_inlet_fold(const StepArgs<S, T> &a, hipStream_t st);
// Pitched state plane <-> dense (nx, ny) FP64 array (with a shift), on the device.
template <class S>
hipError_t launch_dense_from_plane(const S *src, double *dst, int nx, int ny, int pitch, double shift, hipStream_t st);
template <class S>
hipError_t launch_plane_from_dense(const double *src, S *dst, int nx, int ny, int pitch, double shift, hipStream_t st);
// The surface-pressure smoothing weights and their sums, once per handle (they depend on
// the wall angles only).
hipError_t launch_smooth_weights(const BoundaryDev &b, double *w, double *sw, hipStream_t st);
// Probes of a two-step pair on the nodes the two-step bulk carries (a: sub-step 1 with its
// band rectangles; s2, slot2: state t+2 and its record slot, or -1).
template <class S, class T>
hipError_t launch_pair_probes(const StepArgs<S, T> &a, const S *s2, int slot2, hipStream_t st);
// t+2 velocity samples of a pair with both steps sampled, on the nodes of sub-step 1's band
// rectangles that the two-step bulk writes (a: sub-step 1; s2: state t+2); after both.
template <class S, class T>
hipError_t launch_pair_samples(const StepArgs<S, T> &a, const S *s2, hipStream_t st);
// Probes of a four-step pass on the nodes the four-step bulk writes (a: the bulk's arguments,
// with qslot), after the bulk: t+1 .. t+3 recomputed from state t, t+4 from the stored state.
template <class S, class T>
hipError_t launch_quad_probes(const StepArgs<S, T> &a, hipStream_t st);
// Largest ux (NaN ignored) and a non-finite flag (1.0 or 0.0) over the six stored moments
// of the nx x ny nodes of one state, into scratch[0] and scratch[1]; the rest of the
// scratch_words doubles holds per-block partial results.
template <class S>
hipError_t launch_state_scan(const S *src, long long plane, int nx, int ny, int pitch, double *scratch,
                             int scratch_words, hipStream_t st);
// Post-stream populations of rows j0 .. j0+rows-1 into a dense band f(nx, rows, 0:8).
template <class S, class T>
hipError_t launch_get_populations(const StepArgs<S, T> &a, double *f_band, int j0, int rows, hipStream_t st);
template <class S, class T>
hipError_t launch_ghost_from_state(const StepArgs<S, T> &a, T *ghost, hipStream_t st);
template <class S, class T>
hipError_t launch_ghost_from_pops(const StepArgs<S, T> &a, T *ghost, hipStream_t st);
template <class S, class T>
hipError_t launch_stencil_moments(const StepArgs<S, T> &a, hipStream_t st);
// Two steps in one pass over the plain interior (nodes two or more away from the edges and
// from the dilated cylinder box), src = state t, dst = state t+2; strips of H rows.
template <class S, class T>
hipError_t launch_bulk2(const StepArgs<S, T> &a, int nblocks, int H, hipStream_t st, hipEvent_t t0 = nullptr,
                        hipEvent_t t1 = nullptr);
// Four steps in one pass over the plain interior (nodes four or more away from the edges and
// from the cylinder box), src = state t, dst = state t+4; waves of 64 columns writing the
// middle 56, strips of H rows (FP32 12 or 24, FP64 8 or 16).
template <class S, class T>
hipError_t launch_bulkn(const StepArgs<S, T> &a, int nblocks, int H, hipStream_t st, hipEvent_t t0 = nullptr,
                        hipEvent_t t1 = nullptr);
// Whether the four-step kernels are built for strips of H rows in this precision.
template <class S> bool bulkn_rows_ok(int H);
// Workgroups of the ring's inlet and outlet columns when they go by column tiles (64 rows
// each, staged in LDS): grids at least 8 wide stepping from moments; 0 otherwise (the
// columns are then ordinary ring items).
__host__ __device__ constexpr int ring_column_tiles(int nx, int ny, bool from_pops)
{
    return (!from_pops && nx >= 8 && ny >= 3) ? 2 * ((ny - 2 + 63) / 64) : 0;
}
// Side work of one step: ring + band rectangles (+ the wall's stencil nodes when stencil),
// and the boundary workgroup.
template <class S, class T>
hipError_t launch_side(const StepArgs<S, T> &a, hipStream_t st, bool stencil);
template <class S, class T>
hipError_t launch_wall(const StepArgs<S, T> &a, hipStream_t st, bool records = true);
// The steps of a four-step pass whose records run after the pass's side work (one workgroup
// for all of them, then one smoothing launch): per step its record slot, whether it records
// surface pressure, the slot of the per-node force and pressure arrays its wall wrote, and
// the inlet densities of the step before it (with that step's count) to fold into the mean.
struct RecSteps {
    int n;
    int rec_slot[4], record_p[4], wslot[4];
    const double *inlet_prev[4];
    double inlet_prev_c[4];
};
template <class S, class T> hipError_t launch_records4(const StepArgs<S, T> &a, const RecSteps &r, hipStream_t st);
// Batched runs: args[r] per run in device memory; tiles[t] = (run, bulk block of that run).
template <class S, class T>
hipError_t launch_bulk_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int variant, int grid,
                             hipStream_t st);
// Side work of every run of a batch (single-wave blocks from the (run, block) tables; a wall
// of one level in wall groups is among them), then the walls: per-run workgroups when some
// wall has several dependency levels, else the wall blocks of the walls outside wall groups
// plus, on recording steps, one records workgroup per run.
template <class S, class T>
hipError_t launch_side_tiles(const StepArgs<S, T> *args, const int2 *stiles, int nstiles, const int2 *wtiles,
                             int nwtiles, int nruns, bool wall_levels, bool walls, bool records, hipStream_t st);
template <class S, class T>
hipError_t launch_bulk2_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int H, hipStream_t st);
// Four steps per pass for every run of a batch (tiles: (run, four-step block)); strips of H
// rows (FP32 12 or 24, FP64 8 or 16).
template <class S, class T>
hipError_t launch_bulkn_batch(const StepArgs<S, T> *args, const int2 *tiles, int ntiles, int H, hipStream_t st);
// Resident blocks of the batched bulk kernel of a variant on the whole device (0 on error).
template <class S, class T> int bulk_batch_resident_blocks(int variant);

} // namespace mrlbm
