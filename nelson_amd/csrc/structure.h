// Structure stage of a BA problem on the host (the equivalent of
// DoubleSection::parametersReady / addEdge / structureReady,
// nelson/DoubleSection.hpp:30-246, 338-365): variable numbering, edge
// orders, W blocks and the S block pattern with its (W, W) pair lists.
#pragma once

#include <vector>

#include "nelson_amd.h"

namespace nsn {

struct HostStructure {
  int C = 0, P = 0, E = 0, NU = 0, NV = 0, nW = 0, Ec = 0, nsb = 0;
  long long npairs = 0;
  std::vector<int> cam_var, pt_var;            // user -> variable index or -1
  std::vector<int> var_cam_user, var_pt_user;  // variable -> user index
  // variable camera (solver order) -> its index in camera order among the
  // variable cameras (the order of the interface); the identity unless the
  // cameras are reordered
  std::vector<int> var_cam_api;
  // edges in point-major order (user point, then camera)
  std::vector<int> pe;  // user edge index of each slot
  std::vector<int> pt_ptr, pe_cam, pe_wblk;
  std::vector<double> pe_obs, pe_w;
  // W blocks (one per variable (camera, point) pair), point-major
  std::vector<int> wptr, w_cam, w_pt;
  bool w_dup = false;  // some block has more than one edge
  // edges of the variable cameras in camera-major order (camera, then point)
  std::vector<int> ce;  // user edge index of each slot
  std::vector<int> cam_ptr, ce_pt;
  std::vector<double> ce_obs, ce_w;
  // W blocks grouped by camera; w_ref = the reference's row-major W order
  // (camera in interface order, then point)
  std::vector<int> cw_ptr, cw_idx, w_ref;
  // S blocks (i <= j; the diagonal ones with empty pair lists) and, for the
  // off-diagonal ones, the (W_i, W_j) block pairs sharing a point
  std::vector<int> sb_i, sb_j, sb_ptr, sb_pairs;
};

// fills s from the problem description; NSN_OK or NSN_EINVAL (with the
// library's error message) for edges out of range.  allow_order: honour
// d->camera_order (otherwise the variable cameras keep camera order)
int host_structure(const nsn_problem_desc* d, HostStructure& s, bool allow_order = false);

// pieces shared by the two: variable numbering (cam_var, pt_var, their
// inverses, var_cam_api = identity) with the edge range check; whether the
// camera order applies; applying it on a camera graph (sorted adjacency
// lists); the camera graph of the S blocks; the reference order of W
int number_variables(const nsn_problem_desc* d, HostStructure& s);
bool wants_camera_order(const nsn_problem_desc* d, const HostStructure& s);
void apply_camera_order(const nsn_problem_desc* d, HostStructure& s, const std::vector<std::vector<int>>& adj);
std::vector<std::vector<int>> camera_graph_from_blocks(const HostStructure& s);
void w_ref_order(HostStructure& s);

// reverse Cuthill-McKee order of a graph given by sorted adjacency lists:
// order[k] = the node placed k-th (each connected component from a
// pseudo-peripheral node of least degree, neighbours by increasing degree)
std::vector<int> rcm_order(const std::vector<std::vector<int>>& adj);

// nested-dissection order of a graph given by sorted adjacency lists, for
// the tile Cholesky's independent diagonal chains: a part is laid out as
// [A, B, S] with S the nodes outside A adjacent to A (so no edge joins A and
// B), A a prefix of the part's reverse Cuthill-McKee order whose size is a
// multiple of `align` (so B starts on a tile boundary when the part does),
// chosen to minimise max(|A|, |B|) + |S|; A and B are split again the same
// way, at most `depth` levels, while a split shortens that critical length
// to at most 85 % of the part's.  Unsplit parts keep their reverse
// Cuthill-McKee order.  order[k] = the node placed k-th.
std::vector<int> nd_order(const std::vector<std::vector<int>>& adj, int align, int depth);

// the nested-dissection parameters of NSN_ORDER_ND: parts of multiples of
// 32 cameras (6 * 32 = 192 rows, three 64-row tiles), at most four levels
// (16 independent chains)
constexpr int kNdAlign = 32;
constexpr int kNdDepth = 4;

}  // namespace nsn
