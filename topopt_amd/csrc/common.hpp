// common.hpp — shared types and helpers of the MI355X voxel-SIMP engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace topohip {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define TOPOHIP_CHECK(x)                                                          \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess)                                                         \
      throw ::topohip::HipError(std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

// Structured level geometry (element counts per axis; nodes = elements + 1).
struct Geom {
  int ex = 0, ey = 0, ez = 0;
  int nnx = 0, nny = 0, nnz = 0;
  __host__ __device__ long long nodes() const {
    return static_cast<long long>(nnx) * nny * nnz;
  }
  __host__ __device__ long long elems() const {
    return static_cast<long long>(ex) * ey * ez;
  }
  __host__ __device__ long long ndof() const { return 3 * nodes(); }
};

inline Geom make_geom(int ex, int ey, int ez) {
  Geom g;
  g.ex = ex; g.ey = ey; g.ez = ez;
  g.nnx = ex + 1; g.nny = ey + 1; g.nnz = ez + 1;
  return g;
}

// The 45 non-zeros of the hex8 element stiffness in the Hadamard (modal)
// basis, as (k, row, col) with modal index 3*m + comp and m = px + 2py + 4pz
// the mode's x/y/z parity bits. Derived from the isotropic element's cube
// symmetry; the host verifies every entry outside this set is zero for the
// problem's Poisson ratio before using it.
#define TOPOHIP_MODAL_PATTERN(X)                                                  \
  X(0, 3, 3) X(1, 3, 7) X(2, 3, 14) X(3, 4, 4) X(4, 4, 6) X(5, 5, 5) X(6, 5, 12)  \
  X(7, 6, 4) X(8, 6, 6) X(9, 7, 3) X(10, 7, 7) X(11, 7, 14) X(12, 8, 8)           \
  X(13, 8, 13) X(14, 9, 9) X(15, 9, 20) X(16, 10, 10) X(17, 10, 17) X(18, 11, 11) \
  X(19, 11, 16) X(20, 11, 18) X(21, 12, 5) X(22, 12, 12) X(23, 13, 8)             \
  X(24, 13, 13) X(25, 14, 3) X(26, 14, 7) X(27, 14, 14) X(28, 15, 15)             \
  X(29, 15, 19) X(30, 16, 11) X(31, 16, 16) X(32, 16, 18) X(33, 17, 10)           \
  X(34, 17, 17) X(35, 18, 11) X(36, 18, 16) X(37, 18, 18) X(38, 19, 15)           \
  X(39, 19, 19) X(40, 20, 9) X(41, 20, 20) X(42, 21, 21) X(43, 22, 22)            \
  X(44, 23, 23)
constexpr int kModalNnz = 45;

// The 45 modal entries take 8 values, by the cubic symmetry of the isotropic
// element (index r = 3*mode + component, mode bits x + 2y + 4z):
//   normal modes {3, 7, 14}:            [[a,b,b],[b,a,b],[b,b,a]]
//   shear pairs {4,6} {5,12} {8,13}:    d * [[1,1],[1,1]]
//   pairs {9,20} {10,17} {15,19}:       [[e,f],[f,e]]
//   triple {11,16,18}:                  [[g,h,h],[h,g,h],[h,h,g]]
//   modes {21,22,23}:                   k * I
// stored as (a - b, b, d, e, f, g - h, h, k), already divided by 64.
template <typename T>
struct ModalGroups {
  T amb, b, d, e, f, gmh, h, k;
};

template <typename T>
struct Modal {
  T c[kModalNnz];  // already divided by 64 (the Hadamard normalisation)
};

// Reference local node n (core/src/fea/assembly.cpp:45-54: bottom CCW then top
// CCW) -> parity-bit index m = x + 2y + 4z used on the device, and back.
constexpr int kBitOfRef[8] = {0, 1, 3, 2, 4, 5, 7, 6};
constexpr int kRefOfBit[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// Device-side state of one PCG solve (scalars live in device memory so the
// iteration runs without host round trips).
struct CgState {
  double rz, rz_new, pAp, rr, alpha, beta;
  double bnorm2, thresh2;
  double alpha_prev;  // alpha of the last even iteration (its x update deferred)
  int done;       // recurrence residual below threshold (or failure)
  int fail;       // breakdown: p^T A p <= 0 or non-finite
  int iters;      // completed iterations
  int cap;        // iteration cap (fail = 2 when reached unconverged)
  int pfresh;     // set at a restart: the next A.p takes p = z (beta = 0)
  int it0;        // iters at the last restart: iteration iters - it0 is even or odd
};

// MI355X places consecutive workgroup ids on its 8 XCDs round-robin, and each
// XCD has its own L2. Kernels whose neighbouring tiles share input (halos,
// transposed stencil blocks) launch a 1-D grid padded to a multiple of 8 and
// map the hardware id to a logical tile so that each XCD walks one contiguous
// eighth of the tiles: neighbours then meet in the same L2.
constexpr int kXcds = 8;
__host__ __device__ inline unsigned xcd_grid(long long tiles) {
  return static_cast<unsigned>((tiles + kXcds - 1) / kXcds * kXcds);
}
__device__ inline long long xcd_tile(unsigned hw_block, unsigned padded_grid) {
  const unsigned per = padded_grid / kXcds;
  return static_cast<long long>(hw_block % kXcds) * per + hw_block / kXcds;
}

// Load cases solved together (multi-RHS PCG): case r's vectors start at
// base + r * stride, its state is CgState[r], and kernels that take a `done`
// pointer (&st[0].done) skip case r once done[r * kDoneStride] is set. The
// case index is a grid dimension, so one launch covers every case.
constexpr int kDoneStride = static_cast<int>(sizeof(CgState) / sizeof(int));
static_assert(sizeof(CgState) % sizeof(int) == 0, "CgState must be int-strided");
__host__ __device__ inline bool case_done(const int* done, int r) {
  return done && done[r * kDoneStride];
}

}  // namespace topohip
