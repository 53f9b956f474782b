// LinearSolverPCG::solve on a device-resident block matrix with no allocation:
// the variant of g2ohip_linear_solve_pcg for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops). Its
// scratch - block-Jacobi inverses, the PCG vectors, the per-workgroup dot
// partials, the solve's scalars and an error word - lives in one file-scope
// device array per GPU, `scratch`, whose device address is looked up once per
// device (g2ohip_create calls noalloc_init_device) and kept per device.
//
// Calls that overlap on one GPU share that array and race; the caller
// serialises them (one context per stream, calls in order). Every call
// rewrites what it reads back, so the previous call's contents do not leak.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.h"
#include "pcg_drive.h"

namespace g2ohip {

namespace {

// 6 MiB of doubles: 5 vectors of n plus nb * bd^2 inverse entries plus the
// partials, for n = nb * bd <= G2OHIP_NOALLOC_MAX_N unknowns.
constexpr size_t kScratchDoubles = (6u << 20) / sizeof(double);
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// 256-byte aligned carve-out of `count` elements of T from the scratch array
template <typename T>
T* carve(char*& cur, size_t count) {
  T* p = reinterpret_cast<T*>(cur);
  cur += (count * sizeof(T) + 255) / 256 * 256;
  return p;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Device address of this device's scratch array; the device must be current.
hipError_t noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[device] = static_cast<double*>(p);
  return hipSuccess;
}

// Scratch bytes one solve of nb blocks of size bd needs.
size_t noalloc_bytes(int nb, int bd) {
  const size_t n = static_cast<size_t>(nb) * bd;
  return 5 * align256(n * sizeof(double)) + align256(static_cast<size_t>(nb) * bd * bd * sizeof(double)) +
         align256(static_cast<size_t>(pcg_row_blocks(nb)) * sizeof(double)) +
         align256(static_cast<size_t>(pcg_spmv_blocks(nb)) * sizeof(double)) + align256(sizeof(PcgScalars)) +
         align256(sizeof(int));
}

}  // namespace g2ohip

using namespace g2ohip;

// HIP status of a call inside the no-allocation solve, reported without
// building a message string
#define NK(...)                                                                                      \
  do {                                                                                               \
    const hipError_t e_ = (__VA_ARGS__);                                                             \
    if (e_ != hipSuccess)                                                                            \
      return fail_nothrow(e_ == hipErrorOutOfMemory ? G2OHIP_ERR_OOM : G2OHIP_ERR_HIP,               \
                          "HIP call failed in the no-allocation PCG solve");                         \
  } while (0)

// Full block rows (both triangles) in device memory: F_ptr[nb+1], F_col[nnz],
// F_val[nnz * bd * bd] row-major blocks, F_diag[nb] the position of each
// diagonal block; b and x device vectors of nb * bd doubles.
int g2ohip_ctx::linear_solve_pcg_device(int nb, int bd, const int* F_ptr, const int* F_col, const double* F_val,
                                        const int* F_diag, const double* b, double* x, double tol, int max_iter,
                                        double* residual_io, int* iterations) {
  if (device < 0) return fail_nothrow(G2OHIP_ERR_NO_DEVICE, "context has no device");
  // error messages are string literals: no host allocation on any path
  if (bd != 3 && bd != 6) return fail_nothrow(G2OHIP_ERR_UNSUPPORTED, "block size must be 3 or 6");
  if (nb <= 0 || !F_ptr || !F_col || !F_val || !F_diag || !b || !x)
    return fail_nothrow(G2OHIP_ERR_INVALID, "invalid linear system");
  if (static_cast<long long>(nb) * bd > G2OHIP_NOALLOC_MAX_N || noalloc_bytes(nb, bd) > kScratchDoubles * sizeof(double))
    return fail_nothrow(G2OHIP_ERR_UNSUPPORTED, "system larger than the no-allocation limit (65536 unknowns)");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = device < kMaxDevices ? g_ptr[device] : nullptr;
  }
  if (!base) return fail_nothrow(G2OHIP_ERR_HIP, "scratch array of this device was not set up");
  const size_t n = static_cast<size_t>(nb) * bd;
  char* cur = reinterpret_cast<char*>(base);
  double* r = carve<double>(cur, n);
  double* d = carve<double>(cur, n);
  double* d2 = carve<double>(cur, n);
  double* q = carve<double>(cur, n);
  double* s = carve<double>(cur, n);
  double* J = carve<double>(cur, static_cast<size_t>(nb) * bd * bd);
  double* part_rs = carve<double>(cur, pcg_row_blocks(nb));
  double* part_dq = carve<double>(cur, pcg_spmv_blocks(nb));
  PcgScalars* sc = carve<PcgScalars>(cur, 1);
  int* err = carve<int>(cur, 1);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, err);
  NK(hipGetLastError());
  const PcgControl ctl = with_progress(PcgControl{tol, residual_io ? *residual_io : -1.0, -1.0});
  double* const dirs[2] = {d, d2};
  const int mi = max_iter < 0 ? static_cast<int>(n) : max_iter;
  PcgRun pr;
  bool prog_failed = false;
  auto run = [&](auto bdc) -> int {
    constexpr int BD = decltype(bdc)::value;
    NK((launch_jacobi<double, BD>(nb, F_val, F_diag, J, err, stream)));
    auto op = [&](int it, int last, const double* d_old, double* dnew) -> hipError_t {
      return launch_pcg_dir_spmv<double, BD>(nb, F_ptr, F_col, F_val, s, d_old, dnew, q, part_rs, part_dq, sc, ctl,
                                             it, last, stream);
    };
    NK((pcg_drive<double, BD>(stream, h_pinned, progress(), nb, b, J, x, r, s, dirs, q, part_rs, part_dq,
                              pcg_spmv_blocks(nb), sc, mi, op, &pr, &prog_failed)));
    return G2OHIP_OK;
  };
  if (bd == 3) CK(run(std::integral_constant<int, 3>()));
  else CK(run(std::integral_constant<int, 6>()));
  if (prog_failed) prog_ok = false;
  // the error word through the context's pinned buffer (no allocation)
  NK(hipMemcpyAsync(h_pinned, err, sizeof(int), hipMemcpyDeviceToHost, stream));
  NK(hipStreamSynchronize(stream));
  int e = 0;
  std::memcpy(&e, h_pinned, sizeof(int));
  if (iterations) *iterations = pr.iters;
  if (residual_io) *residual_io = 0.5 * pr.dn_final;
  if (e) return fail_nothrow(G2OHIP_ERR_NOT_PD, "singular diagonal block");
  return G2OHIP_OK;
}
