// Host-side structure of a bundle-adjustment problem on the device.
//
// This replaces g2o's per-iteration pointer structures with flat index arrays
// built once, in the role of SparseOptimizer::buildIndexMapping
// (g2o/core/sparse_optimizer.cpp:175-200) and BlockSolver::buildStructure
// (g2o/core/block_solver.hpp:110-269):
//  * free (non-fixed) cameras get indices 0..Cf-1 and free points 0..Pf-1 in
//    input order, fixed vertices -1, exactly the reference's hessianIndex;
//  * active edges (not both vertices fixed) become "slots", ordered point-major
//    (points in a locality order, then by free camera index). A slot owns its
//    6x3 Hpl block, so the per-edge scatter is conflict-free;
//  * the Hschur block pattern is the set of free-camera pairs co-observing a
//    free point plus the Hpp diagonal (block_solver.hpp:237-266);
//  * "chunks" group consecutive points so that one workgroup can reduce their
//    camera-side terms in LDS and write one partial per (chunk, camera) or
//    (chunk, Hschur block); fixed CSR lists then sum those partials in a fixed
//    order, so every reduction is deterministic.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace g2ohip {

struct LinPlanCaps {
  int max_slots = 1024;
  int max_cams = 128;
  int max_pts = 256;
};

// Row-owner plan of the explicit Schur complement: one workgroup per tile of
// one Hschur block row (a free camera). The tile walks the camera's
// observations of free points and accumulates the row's blocks
// [k0, k1) (indices into the upper block list) in LDS; the first tile of a row
// also owns the diagonal block and the right-hand-side coefficient.
struct RowPlan {
  int max_cols = 64;                 // Hschur blocks per row tile
  int chunk = 256;                   // pair products per work item
  std::vector<int> tile_row, tile_k0, tile_k1;
  std::vector<int> tile_order;       // blockIdx -> tile, XCD-aware
  // pair products of each upper Hschur block (i, j): positions (qa, qb) of
  // the point's observations in camera i's and camera j's observation lists
  // (the camera-major order of Hpl), grouped by block, points in processing
  // order; CSR over the nS upper blocks
  std::vector<int> blk_pair_ptr;     // nS+1 (pair count < 2^31)
  std::vector<int> pair_qa, pair_qb;
  long long pair_instances = 0;
  int max_row_cols = 0;
};

struct BAStructure {
  int C = 0, P = 0, E = 0;
  int Cf = 0, Pf = 0;
  std::vector<int> camH, ptH;        // free index or -1
  std::vector<int> free_cam_raw;     // Cf -> raw camera
  // processing positions: points with at least one active edge
  int npos = 0;
  std::vector<int> pos_pt;           // raw point id
  std::vector<int> pos_h;            // free point index or -1
  std::vector<int> pos_ptr;          // npos+1 slot CSR
  std::vector<int> pos_nfree;        // slots with a free camera (they come first)
  // slots
  int ns = 0;
  std::vector<int> slot_edge, slot_cam, slot_pos;
  std::vector<int> slot_camh;        // free index of the slot's camera or -1
  std::vector<int> slot_ph;          // free index of the slot's point or -1
  // camera-major observation lists: slots with a free camera and a free
  // point, camera by camera, points in processing order. Hpl is stored in
  // this order ("q" index).
  int nq = 0;
  std::vector<int> cam_qptr;         // Cf+1
  std::vector<int> q_slot, q_ph;     // nq
  std::vector<int> slot_q;           // ns -> q or -1
  std::vector<int> edge_slot;        // E -> slot or -1 (inactive edge)
  // linearisation chunks
  LinPlanCaps lcaps;
  int nlc = 0;
  std::vector<int> lc_pos0;          // nlc+1
  std::vector<int> lc_cam_ptr;       // nlc+1
  std::vector<int> lc_cams;          // free cameras per chunk, sorted
  std::vector<int> camp_ptr, camp_idx;  // per free camera: rows of lc_cams
  // Hschur pattern: upper blocks sorted by (row, col)
  int nS = 0;
  std::vector<int> S_row, S_col, S_diag;   // S_diag: Cf -> block
  std::vector<int> S_rowptr;               // Cf+1
  // full symmetric block-CSR used by the PCG: both triangles
  int nF = 0;
  std::vector<int> F_ptr, F_col, F_up, F_trans;
  std::vector<int> up_pos, up_posT;  // upper block -> F rows (posT = -1 on diag)
  RowPlan rows;
};

// Builds everything except the Schur plans. Returns an empty string on
// success, else an error message.
std::string build_structure(BAStructure& st, int C, int P, int E, const int* edge_cam, const int* edge_pt,
                            const uint8_t* cam_fixed, const uint8_t* pt_fixed, bool schur);

// Builds st.rows; num_xcds spreads consecutive block rows over the same XCD.
std::string build_row_plan(BAStructure& st, int max_cols, int chunk, int num_xcds);

// 3D pose graph (VertexSE3 / EdgeSE3, 6x6 blocks, no Schur): free poses get
// indices 0..Nf-1 in input order; the Hessian pattern is the diagonal plus one
// upper block per distinct free pair joined by an edge (block_solver.hpp:
// 110-269 without marginalised vertices). Per-edge contributions are summed
// per block through fixed CSR lists, so the assembly is deterministic.
struct PGStructure {
  int N = 0, E = 0, Nf = 0;
  std::vector<int> H;          // free index per pose or -1
  std::vector<int> act;        // active edges (not both ends fixed)
  int nU = 0;                  // upper blocks (diagonal first, then pairs)
  std::vector<int> U_row, U_col;
  std::vector<int> edge_blk;   // per active edge: its off-diagonal upper block or -1
  std::vector<int> edge_swap;  // 1 if the edge's "from" pose has the higher free index
  // sums: diagonal block i = sum over end_ptr[i]..end_ptr[i+1] of end_ref
  // (2 * active position + 0 for "from", + 1 for "to"); off-diagonal block k
  // (k >= Nf) = sum over blk_ptr[k - Nf].. of active edge positions
  std::vector<int> end_ptr, end_ref, blk_ptr, blk_ref;
  // full block rows (both triangles): F_ptr / F_col, position of each upper
  // block and of its transpose (-1 on the diagonal), diagonal positions
  int nF = 0;
  std::vector<int> F_ptr, F_col, up_pos, up_posT, F_diag;
};

std::string build_pg_structure(PGStructure& st, int N, int E, const int* from, const int* to, const uint8_t* fixed);

}  // namespace g2ohip
