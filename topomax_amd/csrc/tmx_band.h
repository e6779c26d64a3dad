// Tiled multi-sweep kernels for the middle levels of the multigrid V-cycle.
//
// Below the finest level the grids are small enough that a V-cycle there is bound by launch
// latency, not bandwidth: each level used to cost seven to nine dependent launches (smoothing
// sweeps, residual, restriction, prolongation) of a few microseconds each.  A band kernel
// does one half of a level's visit in one launch: each workgroup owns a tile of fine nodes,
// stages the tile plus a halo in LDS, and runs every sweep of the pass there with a barrier
// between sweeps.  The halo is as wide as the number of sweeps that read neighbours, so the
// owned nodes come out exactly as the sweep-by-sweep launches would compute them.
//
//   band_down: x = S^nu(0; b) (nu damped block-Jacobi sweeps from zero), r = b - A x,
//              bc = P^T r on the coarse nodes the tile owns; writes x (owned fine nodes) and bc.
//   band_up:   x <- x + P ec, then nu sweeps; writes the result (owned fine nodes) and,
//              optionally, per-workgroup partials of b.y (the finest level's r.z in PCG).
//
// A level's operator is either an element level (P1 element matrices of the two 'right'
// triangle types times per-cell weights) or an explicit 9-point block stencil.  When the
// extended tile has at most one node per thread (the small levels), each thread assembles or
// loads its node's stencil row and inverse diagonal block once and keeps them in registers
// for every sweep of the launch; otherwise the row is rebuilt per sweep.
#pragma once

#include <hip/hip_runtime.h>

#include "tmx_kernels.h"

namespace tmx {

constexpr int kBandThreads = 256;
// workgroup size of the wide variant: elasticity cell-weight levels whose extended tile has
// more nodes than kBandThreads but at most this many keep one node per thread (rows built
// once per launch instead of once per pass)
constexpr int kBandThreadsWide = 1024;
// workgroup size of the mid variant (option): the same tiles as kBandThreads, twice the threads
constexpr int kBandThreadsMid = 512;
constexpr size_t kBandLdsBytes = 64 * 1024;

// where a level's operator comes from
enum BandSrc : int {
  kSrcStencil = 0,  // explicit 9-point block stencil in global memory
  kSrcCells = 1,    // element blocks times per-cell triangle weights (elasticity)
  kSrcUnit = 2      // element blocks with unit weight on every in-grid cell (filter)
};

template <typename T>
struct BandLevel {
  int nx, ny;            // fine level cells
  int nxc, nyc;          // coarse level cells
  int nested;            // transfer: nested P1 (1) or tensor-product linear (0)
  unsigned fixed;        // Dirichlet sides (SideBits)
  int src;               // BandSrc
  const T* cw;           // kSrcCells: per-cell (wA, wB) weights
  const T* st;           // kSrcStencil: 9-point block stencil (SoA, as stencil_op)
  T kt[2][3][3][4];      // element sources: block (i, j) of triangle type t (A, B), BS*BS used
  T urow[9 * 4];         // kSrcUnit: the row of an interior node (all six triangles present)
  T udinv[4];            //   and the inverse of its diagonal block
  int nu_down, nu_up;    // sweeps of the pre- and post-smoothing passes, 1..kMaxNu
  T omega_down[kMaxNu];  // step size of each sweep
  T omega_up[kMaxNu];
  int tx, ty;            // owned fine tile (nodes), even
  int wide;              // 1: kSrcCells levels may use the kBandThreadsWide variant;
                         // 2 / 3: kSrcCells levels run kBandThreadsMid / kBandThreadsWide
                         // threads on the standard tiles
};

// element blocks of the two triangle types: elasticity (lam, mu) or the filter (eps)
template <typename T>
void band_elastic_table(BandLevel<T>& L, double hx, double hy, double lam, double mu);
template <typename T>
void band_filter_table(BandLevel<T>& L, double hx, double hy, double eps, int bs);

// LDS bytes of the down (down = true) or up launch of a level with tile (tx, ty)
size_t band_lds_bytes(int tx, int ty, int nu, int bs, size_t elem, int src, bool down);
// nodes along one side of the extended tile of a t-node tile
size_t band_ext(int t, int nu, bool down);
// tiles (workgroups) of a launch
int band_tiles(int nx, int ny, int tx, int ty);

template <typename T, int BS>
hipError_t band_down(const BandLevel<T>& L, const T* b, T* x, T* bc, hipStream_t s);
// dot: when given, receives one partial of b.y per workgroup (dot->nblocks is set)
template <typename T, int BS>
hipError_t band_up(const BandLevel<T>& L, const T* x, const T* ec, const T* b, T* y,
                   RedSlot* dot, hipStream_t s);

}  // namespace tmx
