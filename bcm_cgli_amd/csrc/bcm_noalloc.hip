// Flux/RK stage on caller-owned device arrays with no device allocation:
// for callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The kernels are those of the drop-in BCM_Flux_XYZ_Viscous_Runge_kutta (the
// marching flux kernel with the RK update fused in place, then the per-case
// norm reduction at RK 3).  What that drop-in keeps in engine allocations --
// the per-cube norm partials, the cube-to-case map, the flow constants and
// the case descriptor of the reduction -- lives here in one file-scope device
// array, `scratch`, of fixed size.  Each device holds its own copy until the
// process ends; its address on the current device is looked up once per
// device with hipGetSymbolAddress and only that address is handed to kernels.
// Every call on a device shares that device's array: calls that overlap on
// one device (two host threads, or two streams) race on it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/bcm_cgli.h"
#include "bcm_engine.h"

namespace bcm {
namespace {

// Layout of the scratch array (bytes): norm partials, 8 doubles per cube;
// cube-to-case map, one int per cube slot (slot 0 included); the flow
// constants and the reduction's case descriptor.
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t PARTIALS_OFF = 0;
constexpr size_t CUBE_CASE_OFF = PARTIALS_OFF + (size_t)BCM_NOALLOC_MAX_CUBES * 8 * sizeof(double);
constexpr size_t PHYS_OFF = align256(CUBE_CASE_OFF + (size_t)(BCM_NOALLOC_MAX_CUBES + 1) * sizeof(int));
constexpr size_t DESC_OFF = PHYS_OFF + 128;
constexpr size_t SCRATCH_BYTES = DESC_OFF + 128;
static_assert(SCRATCH_BYTES == BCM_NOALLOC_SCRATCH_BYTES, "scratch size as declared");
static_assert(sizeof(DevPhys) <= 128 && sizeof(CaseDesc) <= 128, "tables fit their slots");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// Per-device address of `scratch` (the array's name is a host-side stand-in).
constexpr int MAX_DEVICES = 64;
unsigned char* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// Writes the tables of one case covering cubes 1..ncube-1 into scratch: the
// cube-to-case map (slot 0 free), the constants and the case descriptor.
// A kernel rather than copies, so that nothing is staged from the host.
__global__ __launch_bounds__(256) void k_noalloc_tables(unsigned char* s, int ncube, DevPhys ph,
                                                        CaseDesc cd) {
  int* cube_case = reinterpret_cast<int*>(s + CUBE_CASE_OFF);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncube; c += gridDim.x * blockDim.x)
    cube_case[c] = c == 0 ? -1 : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *reinterpret_cast<DevPhys*>(s + PHYS_OFF) = ph;
    *reinterpret_cast<CaseDesc*>(s + DESC_OFF) = cd;
  }
}

int scratch_on_current_device(unsigned char** out, std::string& err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    err = "no current HIP device";
    return 1;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    err = "device index beyond the no-alloc table";
    return 1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc);
      return 1;
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace
}  // namespace bcm

using namespace bcm;

extern "C" int bcm_flux_rk_noalloc(int ncube, int n_cubes_norm, int RK, double deltaT,
                                   double deltaTau, double e, const signed char* FWS,
                                   const double* cube_size, double* U1_, const double* U1,
                                   const double* U1q, double* U1p1, double* U1p2, double* Rku1,
                                   double* Residual1, double* er, const double* phys7,
                                   void* stream) {
  if (!FWS || !cube_size || !U1_ || !U1 || !U1q || !U1p1 || !U1p2 || !Rku1 || !Residual1 ||
      !er || !phys7) {
    set_last_error("bcm_flux_rk_noalloc: null argument");
    return 1;
  }
  if (ncube < 2 || ncube - 1 > BCM_NOALLOC_MAX_CUBES) {
    set_last_error("bcm_flux_rk_noalloc: ncube - 1 must be in 1.." +
                   std::to_string(BCM_NOALLOC_MAX_CUBES) + " (the size of the scratch array)");
    return 1;
  }
  if (RK < 1 || RK > 3) {
    set_last_error("bcm_flux_rk_noalloc: RK must be 1, 2 or 3");
    return 1;
  }
  if (n_cubes_norm < 1) {
    set_last_error("bcm_flux_rk_noalloc: n_cubes_norm must be positive");
    return 1;
  }
  std::string err;
  unsigned char* s = nullptr;
  if (scratch_on_current_device(&s, err)) {
    set_last_error("bcm_flux_rk_noalloc: " + err);
    return 1;
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const double rho0 = phys7[0], P0 = phys7[5];
  const DevPhys ph{rho0, phys7[1], phys7[2], phys7[3], phys7[4], P0, P0 / rho0 / 287.0, phys7[6]};
  const CaseDesc cd{1, ncube - 1, n_cubes_norm, 0};
  hipLaunchKernelGGL(k_noalloc_tables, dim3((ncube + 255) / 256), dim3(256), 0, st, s, ncube, ph,
                     cd);
  double* partials = reinterpret_cast<double*>(s + PARTIALS_OFF);
  const PhysTab pt{reinterpret_cast<const DevPhys*>(s + PHYS_OFF),
                   reinterpret_cast<const int*>(s + CUBE_CASE_OFF)};
  launch_flux_march(st, U1_, U1, U1q, FWS, cube_size, Rku1, Residual1, ncube, pt, deltaT, deltaTau,
                    e, 1, RK, 1, U1p1, U1p2, partials);
  if (RK == 3)
    launch_reduce(st, partials, 1, reinterpret_cast<const CaseDesc*>(s + DESC_OFF), 1, pt.phys,
                  er);
  const hipError_t rc = hipGetLastError();
  if (rc != hipSuccess) {
    set_last_error(std::string("bcm_flux_rk_noalloc: launch failed: ") + hipGetErrorString(rc));
    return 1;
  }
  return 0;
}
