// Measured HBM ceiling of the device: a streaming copy (b -> a) and triad
// (a = b + s * c) over FP64 arrays far larger than the L2s and the MALL, with
// 16-byte coalesced accesses and enough workgroups to fill all 256 CUs.  The
// bench reports the triad's rate beside the 8 TB/s datasheet peak, and the copy
// of known size calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cato_types.hpp"

namespace cato {
int fail(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_copy(v2d *__restrict__ a, const v2d *__restrict__ b, size_t n2) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (size_t)gridDim.x * 256)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&b[i]), &a[i]);
}

__global__ __launch_bounds__(256) void k_triad(v2d *__restrict__ a, const v2d *__restrict__ b,
                                               const v2d *__restrict__ c, double s, size_t n2) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (size_t)gridDim.x * 256) {
    const v2d x = __builtin_nontemporal_load(&b[i]);
    const v2d y = __builtin_nontemporal_load(&c[i]);
    __builtin_nontemporal_store(x + s * y, &a[i]);
  }
}

__global__ void k_fill(double *a, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

struct Dev {
  void *p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace
}  // namespace cato

using namespace cato;

extern "C" int cato_hbm_probe(int kind, unsigned long long bytes_per_array, int reps, double *gb_per_s,
                              double *ms_per_launch) {
  if ((kind != 0 && kind != 1) || reps < 1 || bytes_per_array < (1ull << 20) || !gb_per_s)
    return fail(CATO_ERR_INVALID, "hbm probe: kind 0 (copy) or 1 (triad), reps >= 1, arrays >= 1 MiB");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CATO_ERR_NO_DEVICE, "no HIP device");
  const size_t n2 = (size_t)(bytes_per_array / 16);
  const size_t n = 2 * n2;
  Dev a, b, c;
  int rc;
  if ((rc = hip_check(hipMalloc(&a.p, n * 8), "hbm probe alloc")) ||
      (rc = hip_check(hipMalloc(&b.p, n * 8), "hbm probe alloc")) ||
      (kind == 1 && (rc = hip_check(hipMalloc(&c.p, n * 8), "hbm probe alloc"))))
    return rc;
  hipStream_t st;
  if ((rc = hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hbm probe stream"))) return rc;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, st, (double *)b.p, n, 1.0);
  if (kind == 1) hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, st, (double *)c.p, n, 2.0);
  // 8 workgroups per CU (2048): each thread streams several 16-byte elements
  const dim3 grid(2048);
  auto launch = [&] {
    if (kind == 0)
      hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, (v2d *)a.p, (const v2d *)b.p, n2);
    else
      hipLaunchKernelGGL(k_triad, grid, dim3(256), 0, st, (v2d *)a.p, (const v2d *)b.p,
                         (const v2d *)c.p, 0.5, n2);
  };
  for (int k = 0; k < 3; ++k) launch();
  std::vector<float> ms((size_t)reps);
  for (int k = 0; k < reps; ++k) {
    (void)hipEventRecord(e0, st);
    launch();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms[(size_t)k], e0, e1);
  }
  rc = hip_check(hipGetLastError(), "hbm probe launch");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  if (rc) return rc;
  std::sort(ms.begin(), ms.end());
  const double med = ms[ms.size() / 2];
  const double bytes = (double)n * 8.0 * (kind == 0 ? 2.0 : 3.0);
  *gb_per_s = bytes / (med * 1e-3) / 1e9;
  if (ms_per_launch) *ms_per_launch = med;
  return CATO_OK;
}
