// mrlbm_geometry.h -- host-side construction of the curved-wall boundary tables.
// Follows get_coord (2d_mrlbm.f90:931-1157) and the rotated Hermite tables
// (2d_mrlbm.f90:183-197) for a general centre and radius, and compacts everything the
// boundary workgroup needs into flat arrays uploaded once at setup.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mrlbm {

struct GeometryInput {
    int nx, ny;
    double x_c, y_c, r_c; // 1-based lattice coordinates of the centre, nominal radius
    int label_policy;     // 0: reference masks (labels it leaves unset give NaN), 1: treat as fluid
};

// Per boundary node, in the reference's list order (i outer, j inner).
struct BoundaryNode {
    int i, j;             // 0-based
    int label;
    double radius;
    double xref[4], yref[4]; // reference points 1..4 (wall point and three fluid points)
    double delta;            // distance node -> wall point (2d_mrlbm.f90:1089-1091)
    int ci[3][4], cj[3][4];  // bilinear corners of reference points 2..4, 0-based
    double theta;
};

struct Geometry {
    int nx = 0, ny = 0;
    double x_c = 0, y_c = 0, r_c = 0, r_cyl = 0;
    std::vector<int8_t> cls;            // 1 fluid, 2 solid, 3 boundary; index i + nx*j
    std::vector<BoundaryNode> bnodes;
    std::vector<int> idx;               // theta-sorted order, 0-based
    std::vector<double> thetas;         // sorted theta
    std::vector<double> cth, sth;       // per node (i + nx*j), as 2d_mrlbm.f90:958-966
};

// Builds the node classes, boundary list, labels, r_cyl, reference points, corners and
// theta order.  Returns an empty string on success, else the reason.
std::string build_geometry(const GeometryInput &in, Geometry &g);

// Tables of the boundary workgroup (flat, ready to upload).
struct BoundaryTables {
    int nbct = 0, nU = 0, nlevels = 0;
    // per boundary node k
    std::vector<int32_t> node;          // i + pitch*j
    std::vector<int32_t> ij;            // 2 per node: i, j (0-based)
    std::vector<int32_t> label;
    std::vector<int32_t> uself;         // U index of the node itself
    std::vector<int32_t> uc;            // 12 per node: U index of corners of ref points 2,3,4
    std::vector<int32_t> dep;           // 8 per node: boundary index whose interpolated u a
                                        // corner of ref points 2,3 reads (-1: raw value)
    std::vector<int32_t> level_order;   // nodes sorted by dependency level
    std::vector<int32_t> level_start;   // nlevels+1 offsets into level_order
    std::vector<double> real;           // NREAL per node, see BT_* offsets
    // per unique stencil node u
    std::vector<int32_t> unode;         // i + pitch*j
    std::vector<int32_t> uij;           // 2 per u
    std::vector<double> urot;           // 3 per u: cos^2, sin^2, sin(2 theta) of that node
    // theta order
    std::vector<int32_t> idx;
    std::vector<double> thetas;
    // box of boundary and solid nodes, with per-node class bits (1 boundary, 2 solid)
    int bx0 = 0, by0 = 0, bw = 0, bh = 0;
    std::vector<uint8_t> boxcls;
};

// Offsets into BoundaryTables::real (NREAL doubles per boundary node).
enum {
    BT_XD2 = 0, BT_YD2, BT_XD3, BT_YD3, BT_XD4, BT_YD4,
    BT_T1, BT_T2, BT_T3,
    BT_CTH, BT_STH, BT_S2TH, BT_C2TH,
    BT_R2, BT_R3, BT_R4, BT_RS,
    BT_CXP = 17, BT_CYP = 26, BT_HXXP = 35, BT_HYYP = 44, BT_HXYP = 53,
    BT_NREAL = 62
};

std::string build_tables(const Geometry &g, int pitch, BoundaryTables &t);

// Boundary nodes in groups of `group` (theta order) with the stencil nodes each group reads:
// off (groups + 1 offsets into list), list (stencil node indices), node (boundary node of each
// group position), slot (13 per boundary node: group-local slot of its uc corners 0..11 and of
// itself).  False (and empty) for a wall of several dependency levels or when a group would
// read more than maxu stencil nodes.
struct WallGroups {
    std::vector<int32_t> off, list, node;
    std::vector<int16_t> slot;
};
bool build_wall_groups(const BoundaryTables &t, int group, int maxu, WallGroups &w);

// Incoming / outgoing masks per label (2d_mrlbm.f90:2005-2057) as 9-bit masks.
void label_masks(int label_policy, uint16_t incs[16], uint16_t outs[16]);

} // namespace mrlbm
