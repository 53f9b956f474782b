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

// Packing targets of the linearisation chunks: a chunk closes before it
// would exceed them. A single point larger than a target gets a chunk of its
// own; only max_point_cams (the LDS camera accumulators of one chunk) is a
// hard limit.
struct LinPlanCaps {
  int max_slots = 1024;
  int max_cams = 128;
  int max_pts = 256;
  int max_point_cams = 400;
};


// Window plan of the off-diagonal blocks of the explicit Schur complement.
// The points, in processing order, are cut into windows of at most obs_max
// observations with a free camera and a free point. One workgroup per window
// stages the window's Hpl blocks in LDS as G = Hpl L^-T (L L^T = Hll +
// lambda I of the observation's point), so that the point's contribution to
// Hschur block (i, j) is G_a G_b^T for its observations a by camera i and b by
// camera j. Every off-diagonal upper block (i < j) the window touches gets one
// partial sum per run of at most pch such pairs; one lane forms one partial.
// Partials are grouped 64 to a lane group, longest first, and their pair
// lists are stored lane-interleaved (row r holds pair r of the group's 64
// partials). Each partial owns a slot; slots are block-major (window order
// within a block) so that a second pass sums every block's slots in a fixed
// order. Diagonal blocks are formed per camera in FP64 and are not part of
// the plan.
struct WinPlan {
  int obs_max = 512;
  int pch = 32;
  long long pairs = 0;               // off-diagonal pair products
  std::vector<int> win_obs0;         // nwin+1: ranges of win_q
  std::vector<int> win_q;            // nq: each window's observations (Hpl index q), ascending q
  std::vector<int> win_grp0;         // nwin+1: lane groups of each window
  std::vector<int> grp_len;          // ngrp: pair rows of the group (its longest list)
  std::vector<int> grp_row0;         // ngrp+1: first pair row of the group
  std::vector<uint32_t> pair_rows;   // 64 per row: a | b << 16 (window-local), 0xFFFFFFFF = none
  std::vector<int> grp_slot;         // 64 per group: partial slot of the lane, -1 = none
  std::vector<int> od_blk;           // off-diagonal upper blocks (i < j)
  std::vector<int> od_ptr;           // their slot ranges (size + 1)
};

struct BAStructure {
  int C = 0, P = 0, E = 0;
  int Cf = 0, Pf = 0;
  std::vector<int> camH, ptH;        // free index or -1 (points: in processing order)
  std::vector<int> ph_pub;           // Pf: device point index -> hessianIndex (input order)
  std::vector<int> free_cam_raw;     // Cf -> raw camera
  // processing positions: points with at least one active edge
  int npos = 0;
  std::vector<int> pos_pt;           // raw point id
  std::vector<int> pt_dev;           // P: raw point -> index in the device point state (positions first)
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
  std::vector<int> q_pt;             // nq: position of each observation's point
  std::vector<int> slot_q;           // ns -> q or -1
  std::vector<int> edge_slot;        // E -> slot or -1 (inactive edge)
  // linearisation chunks
  LinPlanCaps lcaps;
  int nlc = 0;
  int lc_maxc = 0;  // most free cameras in one chunk
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
  std::vector<int> F_head;  // row_heads(F_ptr, F_col), uploaded for k_pcg_fused
  std::vector<int> up_pos, up_posT;  // upper block -> F rows (posT = -1 on diag)
  WinPlan win[2];  // Schur window plans: [0] FP32 blocks, [1] FP64 blocks
};

// Builds everything except the Schur plans. Returns an empty string on
// success, else an error message. caps (optional): the linearisation chunks'
// packing targets and per-point limit, instead of the defaults.
// The row-head table of k_pcg_fused (kernels.h PCG_HEAD_STRIDE per row):
// first block, length and the first PCG_HEAD_COLS columns of each block row.
std::vector<int> row_heads(const std::vector<int>& F_ptr, const std::vector<int>& F_col);

std::string build_structure(BAStructure& st, int C, int P, int E, const int* edge_cam, const int* edge_pt,
                            const uint8_t* cam_fixed, const uint8_t* pt_fixed, bool schur,
                            const LinPlanCaps* caps = nullptr);

// Builds a Schur window plan (windows of at most obs_max observations,
// partials of at most pch pairs; obs_max <= 65535).
std::string build_win_plan(const BAStructure& st, WinPlan& wp, int obs_max, int pch);
// Largest number of observations with a free camera of one free point.
int max_point_obs(const BAStructure& st);
// Off-diagonal Schur pair products: over free points, pairs of their
// observations with two different free cameras.
long long schur_pair_count(const BAStructure& st);

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
  std::vector<int> F_head;  // row_heads(F_ptr, F_col), uploaded for k_pcg_fused
};

std::string build_pg_structure(PGStructure& st, int N, int E, const int* from, const int* to, const uint8_t* fixed);

}  // namespace g2ohip
