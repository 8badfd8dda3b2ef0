// noalloc.hip -- sw4hip_bcfortsg_adjoint_noalloc: the transposed boundary
// closure of sw4hip_bcfortsg_adjoint without any device allocation, for
// callers that cannot allocate (code under stream capture, serving loops).
//
// The kernels need two pieces of device memory beyond the caller's arrays:
// the SBP closure weights sbop (read through an SbpTables block) and one
// plane of 3 (nx+4)(ny+4) doubles that holds the adjoint of the free-surface
// ghost plane between its two kernels.  Both live in the file-scope array
// `scratch` below: a fixed block per GPU, held until the process ends.  Its
// device address is looked up once per device with hipGetSymbolAddress (the
// array's name is not a device address in host code) and only that address
// is handed to kernels.  Every call on a GPU shares the one array: calls that
// overlap on a GPU (on different streams, or from several threads) race.
// The plane is written whole by the first kernel before the second reads it
// and the weights are written by a kernel on the caller's stream before
// either, so nothing of an earlier call's contents is read.
#include <mutex>

#include "../../include/sw4hip.h"
#include "kernels_adj.h"
#include "sw4hip_internal.h"

namespace sw4amd {

constexpr long kTableDoubles = (sizeof(SbpTables) + 255) / 256 * (256 / sizeof(double));
static_assert((long)sizeof(SbpTables) <= kTableDoubles * (long)sizeof(double), "table block too small");

// [SbpTables block | plane of 3 * SW4HIP_NOALLOC_MAX_PLANE_POINTS doubles]
static __device__ __align__(256) double scratch[kTableDoubles + 3 * SW4HIP_NOALLOC_MAX_PLANE_POINTS];

namespace {

std::mutex g_scratch_mutex;
double* g_scratch[64] = {nullptr};  // device address of `scratch`, per device

// Device address of this device's copy of `scratch`, looked up on the first
// call on each device (with that device current).
int scratch_address(double** out)
{
   int dev = 0;
   int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
   if (rc) return rc;
   if (dev < 0 || dev >= 64) return set_error(E_RANGE, "device ordinal beyond 63");
   std::lock_guard<std::mutex> lock(g_scratch_mutex);
   if (!g_scratch[dev]) {
      void* p = nullptr;
      if ((rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))) return rc;
      if (!p) return set_error(E_HIP, "hipGetSymbolAddress(scratch) gave a null address");
      g_scratch[dev] = static_cast<double*>(p);
   }
   *out = g_scratch[dev];
   return OK;
}

struct Sbop5 {
   double v[5];
};

__global__ void k_put_sbop(SbpTables* __restrict__ t, Sbop5 s)
{
   if (threadIdx.x < 5) t->sbop[threadIdx.x] = s.v[threadIdx.x];
}

}  // namespace
}  // namespace sw4amd

using namespace sw4amd;

extern "C" int sw4hip_bcfortsg_adjoint_noalloc(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx,
                                               int ny, int nz, double* y, const double* fwd, double h,
                                               const int* bccnd, const double* sbop, const double* mu,
                                               const double* lambda, const double* strx, const double* stry,
                                               double* gmu, double* glambda, void* stream)
{
   int rc;
   if ((rc = check_bcfortsg_adjoint(ib, ie, jb, je, kb, ke, wind, nx, ny, nz, y, fwd, bccnd, sbop, mu, lambda, strx,
                                    stry, gmu, glambda)))
      return rc;
   const Box b = make_box(ib, ie, jb, je, kb, ke);
   if (b.nij > SW4HIP_NOALLOC_MAX_PLANE_POINTS)
      return set_error(E_RANGE, "bcfortsg_adjoint_noalloc: a plane of " + std::to_string(b.nij) +
                                    " points exceeds the scratch limit of " +
                                    std::to_string((long)SW4HIP_NOALLOC_MAX_PLANE_POINTS) + " points");
   double* base = nullptr;
   if ((rc = scratch_address(&base))) return rc;
   SbpTables* tables = reinterpret_cast<SbpTables*>(base);
   double* plane = base + kTableDoubles;
   const hipStream_t s = (hipStream_t)stream;
   Sbop5 sv;
   for (int m = 0; m < 5; m++) sv.v[m] = sbop[m];
   k_put_sbop<<<1, 64, 0, s>>>(tables, sv);
   if ((rc = check_hip(hipGetLastError(), "bcfortsg_adjoint_noalloc: table launch"))) return rc;
   return launch_bcfortsg_T(b, wind, nx, ny, nz, bccnd, y, fwd, h, tables, mu, lambda, strx, stry, gmu, glambda, plane,
                            s);
}
