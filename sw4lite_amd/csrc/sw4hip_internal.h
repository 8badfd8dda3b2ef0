// sw4hip_internal.h -- shared definitions for the gfx950 kernels and the
// host library.  Not part of the public C ABI (that is include/sw4hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

namespace sw4amd {

// Index box of one grid including its ghost points (Sarray::define,
// Sarray.C:282-304).  ni, nij, npts are element counts of one component.
struct Box {
   int ib, ie, jb, je, kb, ke;
   long ni, nij, npts;
};

inline Box make_box(int ib, int ie, int jb, int je, int kb, int ke)
{
   Box b{ib, ie, jb, je, kb, ke, 0, 0, 0};
   b.ni = ie - ib + 1;
   b.nij = b.ni * (je - jb + 1);
   b.npts = b.nij * (ke - kb + 1);
   return b;
}

// Scalar-field offset of (i,j,k).
__host__ __device__ __forceinline__ long sidx(const Box& b, int i, int j, int k)
{
   return (i - b.ib) + b.ni * (j - b.jb) + b.nij * (k - b.kb);
}

// Vector-field offset of component c (0..2) at (i,j,k).  CO=1: component
// blocked, i fastest (Sarray m_corder=true); CO=0: component interleaved.
template <int CO>
__host__ __device__ __forceinline__ long vidx(const Box& b, int c, int i, int j, int k)
{
   const long p = sidx(b, i, j, k);
   return CO ? p + b.npts * c : 3 * p + c;
}

// SBP coefficient block kept in device memory; uniform loads become scalar
// (constant-cache) loads.  Layout is the reference's: acof(k,q,m) at
// (k-1)+6(q-1)+48(m-1), bope(k,q) at (k-1)+6(q-1).
template <class R>
struct SbpTablesT {
   R acof[384];
   R bope[48];
   R ghcof[6];
   R sbop[5];
};
using SbpTables = SbpTablesT<double>;

// Status codes of the C ABI (mirrored in include/sw4hip.h).
enum Status : int {
   OK = 0,
   E_INVAL = 1,
   E_HIP = 2,
   E_NOMEM = 3,
   E_NODEV = 4,
   E_RANGE = 5,
   E_STATE = 6,
};

int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// Host copy of the exact SBP coefficients (GetStencilCoefficients,
// ew-cfromfort.C:1487-1944).
void fill_stencil_coefficients(double* acof, double* ghcof, double* bope, double* sbop);

}  // namespace sw4amd
