// Launch interface of the HIP kernels (kernels.hip). Plain device pointers;
// every launcher enqueues on the given stream and returns the launch status.
#pragma once
#include <hip/hip_runtime.h>

namespace g2ohip {

constexpr int LIN_THREADS = 256;
constexpr int LIN_PTS = 256;   // points per linearisation chunk
// The chunk kernels size their LDS camera accumulators by the plan's largest
// chunk. lds_fits: the kernel's static LDS plus `dynamic` bytes stay within
// what one workgroup may allocate on the current device; the launchers
// return hipErrorInvalidValue instead of launching above it. The kernel's
// static size is queried at each call (kernels of one signature share the
// template instance, so it is not cached); the device limit is the same for
// every CDNA4 device and looked up once.
template <typename Kernel>
inline bool lds_fits(Kernel kernel, size_t dynamic) {
  static int limit = -1;
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kernel)) != hipSuccess) return false;
  if (limit < 0) {
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return false;
    limit = lim;
  }
  return a.sharedSizeBytes + dynamic <= static_cast<size_t>(limit);
}
constexpr int RED_BLOCKS = 512;  // fixed grid of the grid-stride reductions
// fixed grid of the chi2 pass: each thread's slots form a chain of dependent
// gathers, so the pass wants more waves in flight than the other reductions
constexpr int CHI2_BLOCKS = 2048;

// Device view of the structure (all arrays device-resident, read-only).
struct DevGraph {
  int C, P, Cf, Pf, npos, ns, nlc, nS, nF;
  int lc_maxc;  // most free cameras in one linearisation chunk
  double huber;
  const int *camH, *ptH;
  const int *pos_pt, *pos_h, *pos_ptr, *pos_nfree;
  const int *slot_cam, *slot_pos, *slot_camh, *slot_ph;
  int nq;                                    // observations with a free camera and a free point
  const int *slot_q, *cam_qptr, *q_ph;       // camera-major order of Hpl
  const int *q_pt, *cam_raw;                 // point position per q, raw camera per free camera
  // per slot in one load: {position (also the point's index in the
  // position-ordered point state), free point index or -1, raw camera, free
  // camera or -1}: the indices above without their dependent lookups
  const int4* slot_rec;
  const double *slot_meas, *slot_w;
  // per slot the upper-triangular factor [s0 s1 s2] of the edge's information
  // Omega = S^T S (devmath.h whiten_vec); null: Omega = I for every edge
  const double* slot_S;
  const int *lc_pos0, *lc_cam_ptr, *lc_cams, *camp_ptr, *camp_idx;
  const int *S_row, *S_col, *S_diag, *up_pos, *up_posT, *F_ptr, *F_col, *F_diag;
  const double* intr;
};

// Device view of the window plan of the off-diagonal Schur blocks
// (structure.h, WinPlan).
struct DevWinPlan {
  int nwin, nod, obs_max;
  const int *win_obs0, *win_q, *win_grp0, *grp_len, *grp_row0, *grp_slot, *od_blk, *od_ptr;
  const unsigned* pair_rows;
};

// Scalars of one PCG solve, kept on the device so iterations need no host
// round trip. dn is double-buffered by iteration parity.
struct PcgScalars {
  double dn[2];
  double d0;
  double dn0;
  int done;
  int iters;
  int err;
  int pad;
};

// Host-mapped record of one solve: the iteration heads store the iteration
// reached in `it`; the head that ends the solve stores the final values and
// then `done` = 1 with release semantics at system scope, so a host that
// reads done == 1 (acquire) reads the final values without a copy.
struct PcgProgress {
  int it;
  int done;
  int iters;   // PcgScalars::iters
  int err;     // PcgScalars::err
  int abort;   // the abort word seen by the final head
  int pad;
  double dn;   // r.M^-1 r at exit
  double d0;
  double dn0;
};

// Stop rule of one solve: d0 = tol * dn0 floored by resid_mem (the previous
// solve's residual memory, ignored when <= 0), or exactly d0_abs when >= 0.
// prog (device address of a host-mapped PcgProgress, or null).
// abort (device address, or null): a nonzero word stops the solve at its
// next iteration head (an earlier kernel of the same solve reported a
// failure).
struct PcgControl {
  double tol;
  double resid_mem;
  double d0_abs;
  PcgProgress* prog = nullptr;
  const int* abort = nullptr;
};

// Host-mapped copy of the LM trial's scalars (launch_publish_scalars):
// v = d_scal[0..3] (chi2, linearisation chi2, max diagonal, scale), err the
// error word; seq stored last (release, system scope).
struct TrialScalars {
  double v[4];
  int err;
  int seq;
};
hipError_t launch_publish_scalars(const double* scal, const int* err, TrialScalars* out, int seq, hipStream_t st);

// ---- state / residuals
// err4 (optional): four error words zeroed by the same launch.
hipError_t launch_cam_table(const DevGraph& g, const double* cams7, double* camRt, hipStream_t st,
                            int* err4 = nullptr);
hipError_t launch_chi2(const DevGraph& g, const double* camRt, const double* pts, double* partial, double* out,
                       double* err_out, hipStream_t st);
hipError_t launch_update(const DevGraph& g, const double* dx, double* cams7, double* pts, hipStream_t st);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st);

// ---- linear system
template <typename HT>
hipError_t launch_linearize(const DevGraph& g, const double* camRt, const double* pts, HT* Hpl, double* Hll,
                            double* b, double* hpp_part, double* chunk_chi, hipStream_t st);
hipError_t launch_reduce_hpp(const DevGraph& g, const double* hpp_part, double* Hpp, double* b, hipStream_t st);
hipError_t launch_sum(const double* partial, int n, double* out, hipStream_t st);
hipError_t launch_max_diag(const DevGraph& g, const double* Hpp, const double* Hll, double* partial, double* out,
                           hipStream_t st);
hipError_t launch_scale(int n, const double* x, const double* b, double lambda, double* partial, double* out,
                        hipStream_t st);

// ---- explicit Schur complement
// observations per window (LDS-staged G blocks): FP32 blocks 512 (40 KiB of
// LDS); FP64 blocks 256 (36 KiB, four workgroups per CU) unless a point has
// more, then 512
constexpr int SCHUR_WIN_OBS32 = 512;
constexpr int SCHUR_WIN_OBS64 = 256;
constexpr int SCHUR_WIN_PCH = 32;   // pair products per partial
// L^-1 of L L^T = Hll + lambda I per free point (6 doubles, lower triangle
// row by row); err |= 1 if not positive definite.
hipError_t launch_point_linv(const DevGraph& g, const double* Hll, double lambda, double* Linv, int* err,
                             hipStream_t st);
// Off-diagonal partials: part[slot*36] = sum of G_a G_b^T over the slot's
// pairs, G = Hpl L^-T with the point's L^-1 from launch_point_linv.
template <typename HT>
hipError_t launch_schur_win(const DevGraph& g, const DevWinPlan& wp, const HT* Hpl, const double* Linv, HT* part,
                            hipStream_t st);
// Sums every off-diagonal upper block's partials (fixed order, FP64) and
// writes -sum and its transpose into both triangles of F_val.
template <typename HT>
hipError_t launch_schur_offdiag(const DevGraph& g, const DevWinPlan& wp, const HT* part, HT* F_val, hipStream_t st);
// Off-diagonal Schur blocks by camera row (no staging in HBM): items[i] =
// {row a, first and end upper block k0..k1 (off-diagonal, S order), 1 if the
// item covers the whole row}; max_blk the longest item (its LDS size).
constexpr int SCHUR_ROW_CAP = 320;  // most blocks one row item accumulates (FP64: 146 KB of LDS)
constexpr int SCHUR_LONG_TRACK = 64;  // see g2ohip_ctx::use_schur_rows
template <typename HT>
hipError_t launch_schur_rows(const DevGraph& g, const int4* items, int nitems, int max_blk, const int* q_slot,
                             const HT* Hpl, const double* Linv, HT* F_val, hipStream_t st);

// ---- implicit Schur operator
template <typename HT>
hipError_t launch_point_dinv(const DevGraph& g, const double* Hll, double lambda, HT* Dinv, int* err, hipStream_t st);
template <typename HT>
hipError_t launch_implicit_pass(const DevGraph& g, const HT* Hpl, const HT* Dinv, const double* p, const double* b,
                                double* cam_part, const PcgScalars* sc, int it, hipStream_t st,
                                bool exclude_own = false, const double* Linv = nullptr);
// Sdiag (optional): FP64 diagonal blocks of the damped reduced system,
// applied in place of (Hpp + lambda I) when the passes excluded each
// camera's own terms (exclude_own)
template <typename HT>
hipError_t launch_implicit_reduce(const DevGraph& g, const double* cam_part, const double* Hpp, double lambda,
                                  const double* p, double* q, double* dq_part, PcgScalars* sc, int it,
                                  hipStream_t st, const double* Sdiag = nullptr);
hipError_t launch_implicit_rhs_reduce(const DevGraph& g, const double* cam_part, const double* b, double* bschur,
                                      hipStream_t st);

// ---- FP64 matrix-free residual of the reduced system and back-substitution
hipError_t launch_refine(const DevGraph& g, const double* camRt, const double* pts, double lambda, const double* xp,
                         double* xl, double* cam_part, int* err, hipStream_t st);
hipError_t launch_refine_reduce(const DevGraph& g, const double* cam_part, const double* xp, double lambda,
                                double* r, hipStream_t st);

// ---- block-Jacobi PCG (LinearSolverPCG semantics), block size BD
template <typename HT, int BD>
hipError_t launch_jacobi(int nb, const HT* F_val, const int* F_diag, HT* Jinv, int* err, hipStream_t st);
// Schur diagonal blocks in FP64, one workgroup per free camera: S_ii =
// Hpp_ii + lambda I - sum_p V_ip Dinv_p V_ip^T with V the Hpl block,
// re-linearised in FP64 from the state (Hpl == nullptr) or read from Hpl.
// Writes F_val's diagonal blocks (if F_val), their inverses Jinv (err |= 2
// if not PD) and, if bschur, bschur_i = b_i - sum_p V_ip Dinv_p b_p.
// q_slot: slot of each camera-major observation.
// The same for the LM path, point-major (V always re-linearised in FP64 from
// the state): chunk partials into part (27 per linearisation chunk camera
// row), then per camera the block, its inverse and bschur.
// Sdiag (optional): the blocks S_ii in FP64 [Cf][36]; F_val's diagonal
// blocks are then zeroed and the SpMV takes them from Sdiag.
template <typename HT>
hipError_t launch_schur_diag_pm(const DevGraph& g, const double* camRt, const double* pts, const double* Hll,
                                const double* Hpp, const double* b, double lambda, double* part, HT* F_val, HT* Jinv,
                                double* bschur, double* Sdiag, int* err, hipStream_t st);
template <typename HT>
hipError_t launch_schur_diag(const DevGraph& g, const int* q_slot, const double* camRt, const double* pts,
                             const HT* Hpl, const double* Hll, const double* Hpp, const double* b, double lambda,
                             HT* F_val, HT* Jinv, double* bschur, int* err, hipStream_t st);
// Two launches per iteration: pcg_dir_spmv (or pcg_dir + the matrix-free
// operator) then pcg_update. Vectors are FP64, the matrix and the
// preconditioner blocks MT. Partials: part_rs has pcg_row_blocks(nb)
// entries, part_dq pcg_spmv_blocks(nb) (explicit) or
// implicit_reduce_blocks(Cf) (implicit); both are at most nb. d and d_old
// swap every iteration.
int pcg_row_blocks(int nb, int bd = 6);
int pcg_spmv_blocks(int nb);
int implicit_reduce_blocks(int Cf);
template <typename MT, int BD>
hipError_t launch_pcg_init(int nb, const double* b, const MT* Jinv, double* x, double* r, double* s,
                           double* part_rs, PcgScalars* sc, hipStream_t st);
// Ddiag (optional): FP64 diagonal blocks [nb][BD*BD] applied in place of
// F_val's (zeroed) diagonal blocks. nnz: the number of blocks F_ptr[nb]
// (selects the launch shape for short or long rows; < 0 = not known, the
// long-row shape, which is correct for every row length).
template <typename MT, int BD>
hipError_t launch_pcg_dir_spmv(int nb, const int* F_ptr, const int* F_col, const MT* F_val, const double* s,
                               const double* d_old, double* d, double* q, const double* part_rs, double* part_dq,
                               PcgScalars* sc, const PcgControl& ctl, int it, int last, hipStream_t st,
                               const double* Ddiag = nullptr, long long nnz = -1);
// One launch per PCG iteration on the explicit operator (k_pcg_fused): the
// update of iteration it-1 and the direction + product of iteration it.
// F_head: the row heads (row_heads in structure.h).
// s, w, d: two buffers each (launch it writes index it & 1); part_init:
// k_pcg_init's partials; part: 2 x 4 x pcg_fused_blocks(nb) doubles.
// Row heads of a block-row matrix for k_pcg_fused: per row, PCG_HEAD_STRIDE
// ints = first block, length, then the first PCG_HEAD_COLS column indices
// (padded with the row's own index).
constexpr int PCG_HEAD_COLS = 128;
constexpr int PCG_HEAD_STRIDE = PCG_HEAD_COLS + 2;
template <typename MT, int BD>
hipError_t launch_pcg_fused(int nb, const int* F_head, const int* F_col, const MT* F_val, const double* Ddiag,
                            const MT* J, double* x, double* r, double* q, double* const s[2], double* const w[2],
                            double* const d[2], const double* part_init, double* part, PcgScalars* sc,
                            const PcgControl& ctl, int it, int last, hipStream_t st, long long nnz);
int pcg_fused_blocks(int nb);
// Every workgroup of a launch sums all partials of the previous one, so the
// one-launch iteration pays off while the grid is small: up to 256
// workgroups (2,048 block rows); larger systems take the two launches.
bool pcg_fused_fits(int nb);
template <int BD>
hipError_t launch_pcg_dir(int nb, const double* s, const double* d_old, double* d, const double* part_rs,
                          PcgScalars* sc, const PcgControl& ctl, int it, int last, hipStream_t st);
template <typename MT, int BD>
hipError_t launch_pcg_update(int nb, const MT* Jinv, double* x, double* r, const double* d, const double* q,
                             double* s, const double* part_dq, int n_dq, double* part_rs, PcgScalars* sc, int it,
                             hipStream_t st);
// y = A x over 6x6 blocks (block-CSR, or block diagonal when F_ptr is null)
hipError_t launch_block_mv6(int nb, const int* F_ptr, const int* F_col, const double* F_val, const double* x,
                            double* y, hipStream_t st);
// ---- exact Hessian (adjoint, Newton polish), adjoint.hip
hipError_t launch_linearize_exact(const DevGraph& g, const double* camRt, const double* pts, double* Hpl,
                                  double* Hll, double* hpp_part, hipStream_t st);
template <typename HT>
hipError_t launch_backsub_blocks(const DevGraph& g, const HT* Hpl, const HT* Dinv, const double* b, double* x,
                                 hipStream_t st);
hipError_t launch_adjoint_out(const DevGraph& g, const double* camRt, const double* pts, const double* lam,
                              double* out, hipStream_t st);
// ---- 3D pose graph, posegraph.hip
constexpr int PG_REC = 120;  // doubles per active edge record
struct DevPG {
  int N, Nf, A, nU;
  const int *H, *act, *from, *to, *edge_swap;
  const int *end_ptr, *end_ref, *blk_ptr, *blk_ref, *up_pos, *up_posT, *F_diag;
  const double *meas, *info;  // 7 and 36 per edge
};
int pg_edge_blocks(int A);
hipError_t launch_pg_edges(const DevPG& pg, const double* poses, double* rec, double* part, hipStream_t st);
hipError_t launch_pg_assemble(const DevPG& pg, const double* rec, double* F_val, double* Hdiag, double* b,
                              hipStream_t st);
hipError_t launch_pg_damp(const DevPG& pg, const double* Hdiag, double lambda, double* F_val, hipStream_t st);
hipError_t launch_pg_update(const DevPG& pg, const double* x, double* poses, hipStream_t st);
hipError_t launch_pg_chi2(const DevPG& pg, const double* poses, double* part, double* err_out, hipStream_t st);
hipError_t launch_pg_max_diag(int Nf, const double* Hdiag, double* out, hipStream_t st);

// ---- BAL (9-DoF cameras, bal_example.cpp), bal.hip
constexpr int BAL_TAB = 24;  // doubles per camera-table row
struct DevBAL {
  int C, P, Cf, Pf, ns, npos, nq, nO;
  double huber;
  const int *camH, *ptH, *cam_raw;
  const int4* slot_rec;                 // {raw camera, raw point, q or -1, free camera or -1}
  const double *meas, *w;               // per slot
  const int *pos_ptr, *pos_h;           // point positions: slot ranges, free point index
  const int *cs_ptr, *cs_slot;          // per free camera: its slots (any point)
  const int *cam_qptr, *q_ph;           // per free camera: its Hpl blocks; free point per block
  const int *od_ptr, *od_blk;           // off-diagonal upper blocks: pair ranges, upper block index
  const int2* od_pair;                  // (q of the row camera, q of the column camera)
  const int *F_diag, *up_pos, *up_posT;
};
hipError_t launch_bal_cam_table(int C, const double* cams, double* tab, hipStream_t st);
int bal_chi2_blocks();
int bal_point_blocks(int npos);
// chi2 into *out; err_out (optional) 2 per slot
hipError_t launch_bal_chi2(const DevBAL& g, const double* tab, const double* pts, double* part, double* out,
                           double* err_out, hipStream_t st);
// Hpp [Cf][81], Hll [Pf][6], Hpl [nq][27], b; the chi2 of the state into
// *chi_out (part: bal_point_blocks(npos) partials)
hipError_t launch_bal_linearize(const DevBAL& g, const double* tab, const double* pts, double* Hpp, double* Hll,
                                double* Hpl, double* b, double* part, double* chi_out, hipStream_t st);
// part: 256 partial maxima
hipError_t launch_bal_max_diag(const DevBAL& g, const double* Hpp, const double* Hll, double* part, double* out,
                               hipStream_t st);
// Dinv [Pf][9] and Linv [Pf][6] of the damped point blocks (err |= 1 unless
// positive definite), the diagonal blocks of the reduced system into F_val,
// bschur [9 Cf], and (G not null) G = Hpl L^-T [nq][27].
hipError_t launch_bal_schur_diag(const DevBAL& g, const double* Hpp, const double* Hll, const double* Hpl,
                                 const double* b, double lambda, double* Dinv, double* Linv, double* G,
                                 double* F_val, double* bschur, int* err, hipStream_t st);
// The off-diagonal blocks into F_val (both triangles): on the matrix cores
// from G (mfma), else on the vector units from Hpl and Dinv.
hipError_t launch_bal_schur_offdiag(const DevBAL& g, const double* Hpl, const double* Dinv, const double* G,
                                    double* F_val, bool mfma, hipStream_t st);
hipError_t launch_bal_backsub(const DevBAL& g, const double* Hpl, const double* Dinv, const double* b, double* x,
                              hipStream_t st);
hipError_t launch_bal_update(const DevBAL& g, const double* x, double* cams, double* pts, hipStream_t st);

template <typename HT>
hipError_t launch_to_double(int n, const HT* src, double* dst, double scale_add, hipStream_t st);
hipError_t launch_axpy_d(int n, double a, const double* x, double* y, hipStream_t st);

}  // namespace g2ohip
