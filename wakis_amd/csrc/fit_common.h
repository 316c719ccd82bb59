// Shared device helpers of the FIT engine: grid descriptor, face-mask rule,
// flat-index neighbours with the reference operator's wrap/truncation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wf {

struct GridDev {
  int Nx, Ny, Nz;
  int64_t N, sy, sz;   // N = Nx*Ny*Nz, sy = Nx, sz = Nx*Ny
  int pecz, pmcz;      // zero-face bit words (bit f set => face value 0)
  int m1desc;          // H-row sum order
};

template <typename S>
struct MetricDev {
  const S *Lx, *Ly, *Lz, *tLx, *tLy, *tLz;
};

// Value of the reference's int8 face mask Field for component c at (i,j,k).
// boundaries.py:90-103: for component d the faces of the two other axes are
// assigned in x, y, z order (low then high), later assignments win.
__device__ __forceinline__ int face_mask(int zbits, int c, int i, int j, int k, int Nx,
                                         int Ny, int Nz) {
  if (zbits == 0) return 1;
  int f = -1;
  if (c == 0) {
    f = (k == Nz - 1) ? 5 : (k == 0) ? 2 : (j == Ny - 1) ? 4 : (j == 0) ? 1 : -1;
  } else if (c == 1) {
    f = (k == Nz - 1) ? 5 : (k == 0) ? 2 : (i == Nx - 1) ? 3 : (i == 0) ? 0 : -1;
  } else {
    f = (j == Ny - 1) ? 4 : (j == 0) ? 1 : (i == Nx - 1) ? 3 : (i == 0) ? 0 : -1;
  }
  return (f < 0) ? 1 : (((zbits >> f) & 1) ? 0 : 1);
}

// Flat-index neighbours of diags([-1,1],[0,s]) (solverFIT3D.py:207-209):
// n+s wraps into the next row/plane and is dropped past the array end; n-s
// likewise for the transposed operator. Each returns false when truncated.
struct Cell {
  int i, j, k;
};

__device__ __forceinline__ bool fwd_x(Cell &c, const GridDev &g) {
  if (c.i < g.Nx - 1) { c.i++; return true; }
  c.i = 0;
  if (c.j < g.Ny - 1) { c.j++; return true; }
  c.j = 0;
  if (c.k < g.Nz - 1) { c.k++; return true; }
  return false;
}
__device__ __forceinline__ bool fwd_y(Cell &c, const GridDev &g) {
  if (c.j < g.Ny - 1) { c.j++; return true; }
  c.j = 0;
  if (c.k < g.Nz - 1) { c.k++; return true; }
  return false;
}
__device__ __forceinline__ bool fwd_z(Cell &c, const GridDev &g) {
  if (c.k < g.Nz - 1) { c.k++; return true; }
  return false;
}
__device__ __forceinline__ bool bwd_x(Cell &c, const GridDev &g) {
  if (c.i > 0) { c.i--; return true; }
  c.i = g.Nx - 1;
  if (c.j > 0) { c.j--; return true; }
  c.j = g.Ny - 1;
  if (c.k > 0) { c.k--; return true; }
  return false;
}
__device__ __forceinline__ bool bwd_y(Cell &c, const GridDev &g) {
  if (c.j > 0) { c.j--; return true; }
  c.j = g.Ny - 1;
  if (c.k > 0) { c.k--; return true; }
  return false;
}
__device__ __forceinline__ bool bwd_z(Cell &c, const GridDev &g) {
  if (c.k > 0) { c.k--; return true; }
  return false;
}

__device__ __forceinline__ int64_t flat(const Cell &c, const GridDev &g) {
  return (int64_t)c.i + g.sy * c.j + g.sz * c.k;
}

// Edge length of component d at cell c from the per-axis vectors.
template <typename S>
__device__ __forceinline__ S len_of(const S *Lx, const S *Ly, const S *Lz, int d,
                                    const Cell &c) {
  return d == 0 ? Lx[c.i] : (d == 1 ? Ly[c.j] : Lz[c.k]);
}

}  // namespace wf
