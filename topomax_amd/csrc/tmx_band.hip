// Tiled multi-sweep V-cycle kernels for the middle multigrid levels (gfx950, wave64).
//
// Layout of one workgroup's LDS (T = float or double, BS = 1 or 2 values per node):
//   bl[EX*EY*BS]  right-hand side of the extended tile
//   xa, xb        ping-pong iterates of the extended tile (out-of-grid entries stay zero)
//   cwl           kSrcCells only: per-cell triangle weights of cells (OX-1 .. OX+EX-1) x
//                 (OY-1 .. OY+EY-1), (EX+1)*(EY+1) cells x 2 (zero outside the grid)
// The extended tile covers grid nodes (OX .. OX+EX-1) x (OY .. OY+EY-1) with OX = X0 - H.
// Sweep k of a pass updates the nodes at distance >= k (down) or >= k+1 (up) from the
// extended tile's edge, so each sweep only reads neighbours that the previous sweep
// computed; the owned tile (distance H) is exact after the last sweep.  Fixed (Dirichlet)
// nodes keep the value zero throughout, which is what the identity rows of the explicit
// stencils produce for a zero right-hand side there.
#include "tmx_band.h"

#include <cmath>

#include "tmx_device.h"

namespace tmx {

// The six triangles around a node: the cell they belong to (offset (dcx, dcy) from the node's
// lower-left cell), the triangle type (0 = A = (v0,v1,v3), 1 = B = (v0,v2,v3)), the node's
// local vertex index and the node offsets of the triangle's three local vertices.
struct BandTri {
  int dcx, dcy, t, li, ox[3], oy[3];
};

__device__ constexpr BandTri kBandTris[6] = {
    {0, 0, 0, 2, {-1, 0, 0}, {-1, -1, 0}},   // cell (x-1, y-1), A: v3 is the node
    {0, 0, 1, 2, {-1, -1, 0}, {-1, 0, 0}},   // cell (x-1, y-1), B: v3
    {1, 0, 1, 1, {0, 0, 1}, {-1, 0, 0}},     // cell (x, y-1),   B: v2
    {0, 1, 0, 1, {-1, 0, 0}, {0, 0, 1}},     // cell (x-1, y),   A: v1
    {1, 1, 0, 0, {0, 1, 1}, {0, 0, 1}},      // cell (x, y),     A: v0
    {1, 1, 1, 0, {0, 0, 1}, {0, 1, 1}},      // cell (x, y),     B: v0
};

// Structural zeros of the elasticity element blocks: entry (p, r) of block (i, j) of triangle
// type t is lam gi_p gj_r + mu gi_r gj_p + [p == r] mu gi.gj (gradients in units 1/hx, 1/hy),
// zero whenever every product of gradient components in it is zero.  Skipping those terms
// at compile time removes about a third of the multiply-adds of an element-level row.
__host__ __device__ constexpr int kBandGrad(int t, int i, int d) {
  return t == 0 ? (i == 0 ? (d == 0 ? -1 : 0) : i == 1 ? (d == 0 ? 1 : -1) : (d == 0 ? 0 : 1))
                : (i == 0 ? (d == 0 ? 0 : -1) : i == 1 ? (d == 0 ? -1 : 1) : (d == 0 ? 1 : 0));
}
__host__ __device__ constexpr bool band_nz(int bs, int t, int i, int j, int e) {
  if (bs == 1) return true;  // the filter's blocks carry the mass term everywhere
  const int p = e / 2, r = e % 2;
  const bool lam = kBandGrad(t, i, p) * kBandGrad(t, j, r) != 0;
  const bool mu = kBandGrad(t, i, r) * kBandGrad(t, j, p) != 0;
  const bool dot = p == r && (kBandGrad(t, i, 0) * kBandGrad(t, j, 0) != 0 ||
                              kBandGrad(t, i, 1) * kBandGrad(t, j, 1) != 0);
  return lam || mu || dot;
}

// floor(t / d) for 0 <= t <= 1024 and 1 <= d <= 1024: a reciprocal estimate and one correction
// step, a handful of instructions instead of a full integer division
__device__ __forceinline__ int div_small(int t, int d) {
  int q = (int)((float)t * __builtin_amdgcn_rcpf((float)d));
  if ((q + 1) * d <= t) ++q;
  if (q * d > t) --q;
  return q;
}

// f(rx, ry) for every point of an RX x RY region, points spread over the workgroup's threads
// in row-major order; coordinates are stepped by NT (the workgroup size) without a division
// per point
template <int NT, typename F>
__device__ __forceinline__ void for_region(int RX, int RY, F&& f) {
  const int sy = div_small(NT, RX), sx = NT - sy * RX;
  int ry = div_small((int)threadIdx.x, RX), rx = (int)threadIdx.x - ry * RX;
  while (ry < RY) {
    f(rx, ry);
    rx += sx;
    ry += sy;
    if (rx >= RX) {
      rx -= RX;
      ++ry;
    }
  }
}

// The same traversal NB points at a time, for staging from global memory: load(u, rx, ry) for
// the batch's NB points first, then store(u, rx, ry), so each thread has NB global loads in
// flight rather than waiting out the memory latency once per point.  valid(u) says whether
// slot u of the batch holds a point.
template <int NT, int NB, typename FL, typename FS>
__device__ __forceinline__ void for_region_batched(int RX, int RY, FL&& load, FS&& store) {
  const int sy = div_small(NT, RX), sx = NT - sy * RX;
  int ry = div_small((int)threadIdx.x, RX), rx = (int)threadIdx.x - ry * RX;
  while (ry < RY) {
    int px[NB], py[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      px[u] = rx;
      py[u] = ry;
      rx += sx;
      ry += sy;
      if (rx >= RX) {
        rx -= RX;
        ++ry;
      }
    }
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (py[u] < RY) load(u, px[u], py[u]);
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (py[u] < RY) store(u, px[u], py[u]);
  }
}

constexpr int kStageBatch = 4;

// one node's 9-point block row a[kk*BS*BS + a*BS + c], kk = (dy+1)*3 + (dx+1), and the inverse
// of its diagonal block
template <typename T, int BS>
struct BandRow {
  T a[9 * BS * BS];
  T dinv[BS * BS];
};

// extended-tile node (lx, ly), grid node (gx, gy) = g (in the grid, not fixed)
template <typename T, int BS, int SRC>
__device__ __forceinline__ void band_row(const BandLevel<T>& L, const T* __restrict__ cwl, int EX,
                                         int lx, int ly, int gx, int gy, unsigned g, unsigned nn,
                                         BandRow<T, BS>& R) {
  if constexpr (SRC == kSrcStencil) {
    const auto st = glob(L.st) + g;
#pragma unroll
    for (int e = 0; e < 9 * BS * BS; ++e) R.a[e] = st[(size_t)e * nn];
  } else {
    if constexpr (SRC == kSrcUnit) {
      if (gx >= 1 && gy >= 1 && gx < L.nx && gy < L.ny) {  // interior: the precomputed row
#pragma unroll
        for (int e = 0; e < 9 * BS * BS; ++e) R.a[e] = L.urow[e];
#pragma unroll
        for (int e = 0; e < BS * BS; ++e) R.dinv[e] = L.udinv[e];
        return;
      }
    }
#pragma unroll
    for (int e = 0; e < 9 * BS * BS; ++e) R.a[e] = T(0);
    const int c0 = ly * (EX + 1) + lx;  // LDS cell of grid cell (x-1, y-1)
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const BandTri& TR = kBandTris[r];
      T w;
      if constexpr (SRC == kSrcCells) {
        w = cwl[2 * (c0 + TR.dcy * (EX + 1) + TR.dcx) + TR.t];
      } else {
        const int cx = gx - 1 + TR.dcx, cy = gy - 1 + TR.dcy;
        w = (cx >= 0 && cy >= 0 && cx < L.nx && cy < L.ny) ? T(1) : T(0);
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int kk = (TR.oy[q] + 1) * 3 + (TR.ox[q] + 1);
#pragma unroll
        for (int e = 0; e < BS * BS; ++e)
          if (band_nz(BS, TR.t, TR.li, q, e)) R.a[kk * BS * BS + e] += w * L.kt[TR.t][TR.li][q][e];
      }
    }
  }
  const T* D = R.a + 4 * BS * BS;
  if constexpr (BS == 1) {
    R.dinv[0] = recip(D[0]);
  } else {
    const T inv = recip(D[0] * D[3] - D[1] * D[2]);
    R.dinv[0] = D[3] * inv;
    R.dinv[1] = -D[1] * inv;
    R.dinv[2] = -D[2] * inv;
    R.dinv[3] = D[0] * inv;
  }
}

// res = b - A x at extended-tile node i (ZERO: x = 0)
template <typename T, int BS, int SRC, bool ZERO>
__device__ __forceinline__ void band_res(const BandRow<T, BS>& R, const T* __restrict__ x,
                                         const T* __restrict__ bl, int EX, int i, T (&res)[BS]) {
#pragma unroll
  for (int q = 0; q < BS; ++q) res[q] = bl[BS * i + q];
  if constexpr (!ZERO) {
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      if (SRC != kSrcStencil && (kk == 2 || kk == 6)) continue;  // no (1,-1), (-1,1) edges
      const int j = i + (kk / 3 - 1) * EX + (kk % 3 - 1);
#pragma unroll
      for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int c = 0; c < BS; ++c) res[a] -= R.a[kk * BS * BS + a * BS + c] * x[BS * j + c];
    }
  }
}

// the node update of one pass: a damped block-Jacobi sweep (ZERO: from x = 0) or, with RES,
// the residual b - A x
template <typename T, int BS, int SRC, bool ZERO, bool RES>
__device__ __forceinline__ void band_update(const BandRow<T, BS>& R, const T* x, const T* bl,
                                            T* y, T omega, int EX, int i) {
  T res[BS];
  band_res<T, BS, SRC, ZERO>(R, x, bl, EX, i, res);
  if constexpr (RES) {
#pragma unroll
    for (int q = 0; q < BS; ++q) y[BS * i + q] = res[q];
  } else {
#pragma unroll
    for (int a = 0; a < BS; ++a) {
      T d = T(0);
#pragma unroll
      for (int c = 0; c < BS; ++c) d += R.dinv[a * BS + c] * res[c];
      y[BS * i + a] = (ZERO ? T(0) : x[BS * i + a]) + omega * d;
    }
  }
}

// Per-thread node of a launch.  ONE: thread t owns extended-tile node t for the whole launch
// (the tile has at most as many nodes as the workgroup threads) and keeps its row in
// registers; otherwise the
// nodes of each pass are spread over the threads and rows are rebuilt per pass.
template <typename T, int BS, int SRC, bool ONE>
struct BandNode {
  int lx = 0, ly = 0, gx = 0, gy = 0;
  bool live = false;   // in the extended tile and in the grid
  bool fixed = false;
  BandRow<T, BS> row;
  __device__ __forceinline__ void init(const BandLevel<T>& L, const T* cwl, int EX, int EY,
                                       int OX, int OY) {
    if constexpr (ONE) {
      const int t = (int)threadIdx.x;
      ly = div_small(t, EX);
      lx = t - ly * EX;
      gx = OX + lx;
      gy = OY + ly;
      live = ly < EY && gx >= 0 && gy >= 0 && gx <= L.nx && gy <= L.ny;
      fixed = live && node_fixed(L.fixed, gx, gy, L.nx, L.ny);
      if (live && !fixed) {
        const unsigned nvx = (unsigned)L.nx + 1u;
        band_row<T, BS, SRC>(L, cwl, EX, lx, ly, gx, gy, (unsigned)gy * nvx + (unsigned)gx,
                             nvx * ((unsigned)L.ny + 1u), row);
      }
    }
  }
};

// Interior pass of a scalar level whose tile lies strictly inside the grid (every node has
// the precomputed row, none is fixed): each thread updates two horizontally adjacent nodes
// per step, walking down a column pair, so the two nodes share their neighbour loads and the
// address arithmetic (one multiply per pass instead of one per node).
template <typename T, bool ZERO, bool RES, int NT>
__device__ __forceinline__ void band_pass_unit_pairs(const BandLevel<T>& L, const T* x,
                                                     const T* bl, T* y, T omega, int EX, int EY,
                                                     int s) {
  const int RX = EX - 2 * s, RY = EY - 2 * s;
  const int NP = (RX + 1) >> 1;          // column pairs of the region (<= NT)
  const int rows = div_small(NT, NP);    // rows in flight across the workgroup
  const int t = (int)threadIdx.x;
  const int r0 = div_small(t, NP), cp = t - r0 * NP;
  if (r0 >= rows) return;
  const bool two = 2 * cp + 1 < RX;      // the pair's second node is in the region
  const T a0 = L.urow[0], a1 = L.urow[1], a3 = L.urow[3], a4 = L.urow[4], a5 = L.urow[5];
  const T a7 = L.urow[7], a8 = L.urow[8], dv = L.udinv[0];
  const int step = rows * EX;
  int i = (s + r0) * EX + s + 2 * cp;
  for (int ry = r0; ry < RY; ry += rows, i += step) {
    const int j2 = two ? i + 2 : i;      // (the second node's right column, kept in the tile)
    T ra = bl[i], rb = two ? bl[i + 1] : T(0);
    if constexpr (!ZERO) {
      const T xs = x[i - EX - 1], xt = x[i - EX], xu = x[i - EX + 1];
      const T xl = x[i - 1], xa = x[i], xb = x[i + 1], xr = x[j2];
      const T xv = x[i + EX], xw = x[i + EX + 1], xz = x[j2 + EX];
      ra -= a0 * xs + a1 * xt + a3 * xl + a4 * xa + a5 * xb + a7 * xv + a8 * xw;
      rb -= a0 * xt + a1 * xu + a3 * xa + a4 * xb + a5 * xr + a7 * xw + a8 * xz;
      if constexpr (!RES) {
        ra = xa + omega * dv * ra;
        rb = xb + omega * dv * rb;
      }
    } else {
      ra = omega * dv * ra;
      rb = omega * dv * rb;
    }
    y[i] = ra;
    if (two) y[i + 1] = rb;
  }
}

// one pass over the extended-tile nodes at distance >= s from its edge
template <typename T, int BS, int SRC, bool ONE, bool ZERO, bool RES, int NT>
__device__ __forceinline__ void band_pass(const BandLevel<T>& L, BandNode<T, BS, SRC, ONE>& me,
                                          const T* x, const T* bl, const T* cwl, T* y, T omega,
                                          int EX, int EY, int OX, int OY, int s) {
  if constexpr (ONE) {
    if (me.live && me.lx >= s && me.lx < EX - s && me.ly >= s && me.ly < EY - s) {
      const int i = me.ly * EX + me.lx;
      if (me.fixed) {
#pragma unroll
        for (int q = 0; q < BS; ++q) y[BS * i + q] = T(0);
      } else {
        band_update<T, BS, SRC, ZERO, RES>(me.row, x, bl, y, omega, EX, i);
      }
    }
  } else {
    if constexpr (SRC == kSrcUnit) {
      // extended tile strictly inside the grid: every node has the interior row and none is
      // fixed, so the pass needs no per-node tests
      if (OX >= 1 && OY >= 1 && OX + EX <= L.nx && OY + EY <= L.ny) {
        if constexpr (BS == 1) {
          band_pass_unit_pairs<T, ZERO, RES, NT>(L, x, bl, y, omega, EX, EY, s);
          return;
        }
        BandRow<T, BS> U;
#pragma unroll
        for (int e = 0; e < 9 * BS * BS; ++e) U.a[e] = L.urow[e];
#pragma unroll
        for (int e = 0; e < BS * BS; ++e) U.dinv[e] = L.udinv[e];
        for_region<NT>(EX - 2 * s, EY - 2 * s, [&](int rx, int ry) {
          band_update<T, BS, SRC, ZERO, RES>(U, x, bl, y, omega, EX, (s + ry) * EX + (s + rx));
        });
        return;
      }
    }
    const unsigned nvx = (unsigned)L.nx + 1u;
    const unsigned nn = nvx * ((unsigned)L.ny + 1u);
    for_region<NT>(EX - 2 * s, EY - 2 * s, [&](int rx, int ry) {
      const int lx = s + rx, ly = s + ry;
      const int gx = OX + lx, gy = OY + ly;
      if (gx < 0 || gy < 0 || gx > L.nx || gy > L.ny) return;
      const int i = ly * EX + lx;
      if (node_fixed(L.fixed, gx, gy, L.nx, L.ny)) {
#pragma unroll
        for (int q = 0; q < BS; ++q) y[BS * i + q] = T(0);
        return;
      }
      BandRow<T, BS> R;
      band_row<T, BS, SRC>(L, cwl, EX, lx, ly, gx, gy, (unsigned)gy * nvx + (unsigned)gx, nn, R);
      band_update<T, BS, SRC, ZERO, RES>(R, x, bl, y, omega, EX, i);
    });
  }
}

// coarse nodes whose restriction centre lies in fine nodes [X0, X1) of one direction
__device__ __forceinline__ void band_coarse_range(int X0, int X1, int nf, int nc, int& lo,
                                                  int& hi) {
  if (nf == nc) {
    lo = X0;
    hi = min(X1, nf + 1);
  } else {
    lo = (X0 + 1) >> 1;
    hi = X1 > nf ? nc + 1 : (X1 + 1) >> 1;
  }
}

template <typename T, int SRC, int NT>
__device__ __forceinline__ void band_stage_cells(const BandLevel<T>& L, T* cwl, int EX, int EY,
                                                 int OX, int OY) {
  if constexpr (SRC == kSrcCells) {
    const int CX = EX + 1, CY = EY + 1;
    T w[kStageBatch][2];
    for_region_batched<NT, kStageBatch>(
        CX, CY,
        [&](int u, int rx, int ry) {
          const int cx = OX - 1 + rx, cy = OY - 1 + ry;
          w[u][0] = w[u][1] = T(0);
          if (cx >= 0 && cy >= 0 && cx < L.nx && cy < L.ny) {
            const unsigned c = (unsigned)cy * (unsigned)L.nx + (unsigned)cx;
            w[u][0] = glob(L.cw)[2 * c];
            w[u][1] = glob(L.cw)[2 * c + 1];
          }
        },
        [&](int u, int rx, int ry) {
          const int k = ry * CX + rx;
          cwl[2 * k] = w[u][0];
          cwl[2 * k + 1] = w[u][1];
        });
  }
}

// tile origin of this workgroup
template <typename T>
__device__ __forceinline__ void band_tile(const BandLevel<T>& L, unsigned bid, int& X0, int& Y0) {
  const unsigned ntx = (unsigned)((L.nx + L.tx) / L.tx);  // = ceil((nx + 1) / tx)
  X0 = (int)(bid % ntx) * L.tx;
  Y0 = (int)(bid / ntx) * L.ty;
}

template <typename T, int BS, int SRC, bool ONE, int NT>
__device__ __forceinline__ void band_down_body(const BlkId blk, const BandLevel<T>& L,
                                               const T* b_, T* xo_, T* bc_) {
  const auto b = glob(b_);
  const auto xo = glob(xo_);
  const auto bc = glob(bc_);
  extern __shared__ __align__(16) unsigned char band_lds[];
  const int nu = L.nu_down, H = nu + 1;
  const int EX = L.tx + 2 * H, EY = L.ty + 2 * H, NE = EX * EY;
  int X0, Y0;
  band_tile(L, blk.id, X0, Y0);
  const int OX = X0 - H, OY = Y0 - H;
  T* bl = reinterpret_cast<T*>(band_lds);
  T* xa = bl + BS * NE;
  T* xb = xa + BS * NE;
  T* cwl = xb + BS * NE;
  const unsigned nvx = (unsigned)L.nx + 1u;
  T bv[kStageBatch][BS];
  for_region_batched<NT, kStageBatch>(
      EX, EY,
      [&](int u, int rx, int ry) {
        const int gx = OX + rx, gy = OY + ry;
#pragma unroll
        for (int q = 0; q < BS; ++q) bv[u][q] = T(0);
        if (gx >= 0 && gy >= 0 && gx <= L.nx && gy <= L.ny) {
          const unsigned g = (unsigned)gy * nvx + (unsigned)gx;
#pragma unroll
          for (int q = 0; q < BS; ++q) bv[u][q] = b[BS * g + q];
        }
      },
      [&](int u, int rx, int ry) {
        const int k = ry * EX + rx;
#pragma unroll
        for (int q = 0; q < BS; ++q) {
          bl[BS * k + q] = bv[u][q];
          xa[BS * k + q] = T(0);
          xb[BS * k + q] = T(0);
        }
      });
  band_stage_cells<T, SRC, NT>(L, cwl, EX, EY, OX, OY);
  __syncthreads();
  BandNode<T, BS, SRC, ONE> me;
  me.init(L, cwl, EX, EY, OX, OY);
  T* cur = xa;
  T* nxt = xb;
  band_pass<T, BS, SRC, ONE, true, false, NT>(L, me, nullptr, bl, cwl, cur, L.omega_down[0], EX, EY,
                                          OX, OY, 0);
  __syncthreads();
  for (int k = 1; k < nu; ++k) {
    band_pass<T, BS, SRC, ONE, false, false, NT>(L, me, cur, bl, cwl, nxt, L.omega_down[k], EX, EY,
                                             OX, OY, k);
    __syncthreads();
    T* t = cur;
    cur = nxt;
    nxt = t;
  }
  // residual of the pre-smoothed iterate into the spare buffer
  band_pass<T, BS, SRC, ONE, false, true, NT>(L, me, cur, bl, cwl, nxt, T(0), EX, EY, OX, OY, nu);
  // owned fine nodes of the pre-smoothed iterate
  const int X1 = min(X0 + L.tx, L.nx + 1), Y1 = min(Y0 + L.ty, L.ny + 1);
  for_region<NT>(X1 - X0, Y1 - Y0, [&](int rx, int ry) {
    const int gx = X0 + rx, gy = Y0 + ry;
    const int i = (gy - OY) * EX + (gx - OX);
    const unsigned g = (unsigned)gy * nvx + (unsigned)gx;
#pragma unroll
    for (int q = 0; q < BS; ++q) xo[BS * g + q] = cur[BS * i + q];
  });
  __syncthreads();
  // restriction of the residual to the owned coarse nodes
  int I0, I1, J0, J1;
  band_coarse_range(X0, X0 + L.tx, L.nx, L.nxc, I0, I1);
  band_coarse_range(Y0, Y0 + L.ty, L.ny, L.nyc, J0, J1);
  const T* r = nxt;
  if (I1 > I0) for_region<NT>(I1 - I0, J1 - J0, [&](int rx, int ry) {
    const int I = I0 + rx, J = J0 + ry;
    T acc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) acc[q] = T(0);
    if (!node_fixed(L.fixed, I, J, L.nxc, L.nyc)) {
      if (L.nested) {
        const int i = (2 * J - OY) * EX + (2 * I - OX);
        const int nb[6] = {1, -1, EX, -EX, EX + 1, -EX - 1};
#pragma unroll
        for (int q = 0; q < BS; ++q) {
          T sum = T(0);
#pragma unroll
          for (int t = 0; t < 6; ++t) sum += r[BS * (i + nb[t]) + q];
          acc[q] = r[BS * i + q] + T(0.5) * sum;
        }
      } else {
        const int px = cpos(I, L.nx, L.nxc), py = cpos(J, L.ny, L.nyc);
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
          const int fy = py + dy;
          const double wy = (fy < 0 || fy > L.ny) ? 0.0 : interp1_weight(fy, L.ny, L.nyc, J);
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            const int fx = px + dx;
            const double w = (fx < 0 || fx > L.nx) ? 0.0 : wy * interp1_weight(fx, L.nx, L.nxc, I);
            const int i = (fy - OY) * EX + (fx - OX);
#pragma unroll
            for (int q = 0; q < BS; ++q) acc[q] += (T)w * r[BS * i + q];
          }
        }
      }
    }
    const unsigned c = (unsigned)J * ((unsigned)L.nxc + 1u) + (unsigned)I;
#pragma unroll
    for (int q = 0; q < BS; ++q) bc[BS * c + q] = acc[q];
  });
}

template <typename T, int BS, int SRC, bool ONE, int NT>
__global__ __launch_bounds__(NT) void k_band_down(BandLevel<T> L,
                                                            const T* __restrict__ b,
                                                            T* __restrict__ xo,
                                                            T* __restrict__ bc) {
  band_down_body<T, BS, SRC, ONE, NT>(hw_blk(), L, b, xo, bc);
}

template <typename T, int BS, int SRC, bool ONE, int NT>
__device__ __forceinline__ void band_up_body(const BlkId blk, const BandLevel<T>& L, const T* x_,
                                             const T* ec_, const T* b_, T* y_, double* dot) {
  const auto x = glob(x_);
  const auto ec = glob(ec_);
  const auto b = glob(b_);
  const auto y = glob(y_);
  extern __shared__ __align__(16) unsigned char band_lds[];
  const int nu = L.nu_up, H = nu;
  const int EX = L.tx + 2 * H, EY = L.ty + 2 * H, NE = EX * EY;
  int X0, Y0;
  band_tile(L, blk.id, X0, Y0);
  const int OX = X0 - H, OY = Y0 - H;
  T* bl = reinterpret_cast<T*>(band_lds);
  T* xa = bl + BS * NE;
  T* xb = xa + BS * NE;
  T* cwl = xb + BS * NE;
  const unsigned nvx = (unsigned)L.nx + 1u, nvc = (unsigned)L.nxc + 1u;
  T bv[kStageBatch][BS], xv[kStageBatch][BS];
  for_region_batched<NT, kStageBatch>(
      EX, EY,
      [&](int u, int rx, int ry) {
        const int gx = OX + rx, gy = OY + ry;
#pragma unroll
        for (int q = 0; q < BS; ++q) bv[u][q] = xv[u][q] = T(0);
        if (gx >= 0 && gy >= 0 && gx <= L.nx && gy <= L.ny) {
          const unsigned g = (unsigned)gy * nvx + (unsigned)gx;
#pragma unroll
          for (int q = 0; q < BS; ++q) {
            bv[u][q] = b[BS * g + q];
            xv[u][q] = x[BS * g + q];
          }
          if (!node_fixed(L.fixed, gx, gy, L.nx, L.ny)) {
            if (L.nested) {
              const unsigned I0 = (unsigned)gx >> 1, J0 = (unsigned)gy >> 1;
              const unsigned I1 = I0 + (gx & 1), J1 = J0 + (gy & 1);
              const unsigned c0 = J0 * nvc + I0, c1 = J1 * nvc + I1;
#pragma unroll
              for (int q = 0; q < BS; ++q)
                xv[u][q] += T(0.5) * (ec[BS * c0 + q] + ec[BS * c1 + q]);
            } else {
              const Interp1 ix = interp1(gx, L.nx, L.nxc), iy = interp1(gy, L.ny, L.nyc);
              const unsigned cx1 = ix.n > 1 ? ix.c[1] : ix.c[0];
              const unsigned cy1 = iy.n > 1 ? iy.c[1] : iy.c[0];
              const T wx0 = (T)ix.w[0], wx1 = ix.n > 1 ? (T)ix.w[1] : T(0);
              const T wy0 = (T)iy.w[0], wy1 = iy.n > 1 ? (T)iy.w[1] : T(0);
              const unsigned r0 = (unsigned)iy.c[0] * nvc, r1 = cy1 * nvc;
#pragma unroll
              for (int q = 0; q < BS; ++q)
                xv[u][q] += wx0 * wy0 * ec[BS * (r0 + ix.c[0]) + q] +
                            wx1 * wy0 * ec[BS * (r0 + cx1) + q] +
                            wx0 * wy1 * ec[BS * (r1 + ix.c[0]) + q] +
                            wx1 * wy1 * ec[BS * (r1 + cx1) + q];
            }
          }
        }
      },
      [&](int u, int rx, int ry) {
        const int k = ry * EX + rx;
#pragma unroll
        for (int q = 0; q < BS; ++q) {
          bl[BS * k + q] = bv[u][q];
          xa[BS * k + q] = xv[u][q];
          xb[BS * k + q] = T(0);
        }
      });
  band_stage_cells<T, SRC, NT>(L, cwl, EX, EY, OX, OY);
  __syncthreads();
  BandNode<T, BS, SRC, ONE> me;
  me.init(L, cwl, EX, EY, OX, OY);
  T* cur = xa;
  T* nxt = xb;
  for (int k = 0; k < nu; ++k) {
    band_pass<T, BS, SRC, ONE, false, false, NT>(L, me, cur, bl, cwl, nxt, L.omega_up[k], EX, EY, OX,
                                             OY, k + 1);
    __syncthreads();
    T* t = cur;
    cur = nxt;
    nxt = t;
  }
  const int X1 = min(X0 + L.tx, L.nx + 1), Y1 = min(Y0 + L.ty, L.ny + 1);
  double part = 0.0;
  for_region<NT>(X1 - X0, Y1 - Y0, [&](int rx, int ry) {
    const int gx = X0 + rx, gy = Y0 + ry;
    const int i = (gy - OY) * EX + (gx - OX);
    const unsigned g = (unsigned)gy * nvx + (unsigned)gx;
#pragma unroll
    for (int q = 0; q < BS; ++q) {
      y[BS * g + q] = cur[BS * i + q];
      part += (double)bl[BS * i + q] * (double)cur[BS * i + q];
    }
  });
  if constexpr (NT <= kBlock) {  // (block_partials holds one slot per wave of kBlock)
    if (dot) {
      const double v[1] = {part};
      double* const o[1] = {dot};
      block_partials<1>(v, o, blk.id);
    }
  }
}

template <typename T, int BS, int SRC, bool ONE, int NT>
__global__ __launch_bounds__(NT) void k_band_up(BandLevel<T> L, const T* __restrict__ x,
                                                          const T* __restrict__ ec,
                                                          const T* __restrict__ b,
                                                          T* __restrict__ y,
                                                          double* __restrict__ dot) {
  band_up_body<T, BS, SRC, ONE, NT>(hw_blk(), L, x, ec, b, y, dot);
}

// ---------------------------------------------------------------- host side
// gradient factors of the local vertices of the two triangle types, in units 1/hx, 1/hy
static const double kGrad[2][3][2] = {{{-1, 0}, {1, -1}, {0, 1}}, {{0, -1}, {-1, 1}, {1, 0}}};

template <typename T>
void band_elastic_table(BandLevel<T>& L, double hx, double hy, double lam, double mu) {
  const double area = 0.5 * hx * hy;
  for (int t = 0; t < 2; ++t)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double gi[2] = {kGrad[t][i][0] / hx, kGrad[t][i][1] / hy};
        const double gj[2] = {kGrad[t][j][0] / hx, kGrad[t][j][1] / hy};
        const double dotg = gi[0] * gj[0] + gi[1] * gj[1];
        for (int p = 0; p < 2; ++p)
          for (int r = 0; r < 2; ++r) {
            double v = lam * gi[p] * gj[r] + mu * gi[r] * gj[p];
            if (p == r) v += mu * dotg;
            L.kt[t][i][j][p * 2 + r] = band_nz(2, t, i, j, p * 2 + r) ? (T)(area * v) : T(0);
          }
      }
}

// the six triangles around a node (as kBandTris): triangle type, local index of the node and
// node offsets of the three local vertices
static const int kTrisHost[6][8] = {
    {0, 2, -1, 0, 0, -1, -1, 0}, {1, 2, -1, -1, 0, -1, 0, 0}, {1, 1, 0, 0, 1, -1, 0, 0},
    {0, 1, -1, 0, 0, 0, 0, 1},   {0, 0, 0, 1, 1, 0, 0, 1},    {1, 0, 0, 0, 1, 0, 1, 1}};

template <typename T>
void band_filter_table(BandLevel<T>& L, double hx, double hy, double eps, int bs) {
  const double area = 0.5 * hx * hy;
  double k[2][3][3];
  for (int t = 0; t < 2; ++t)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double gx = kGrad[t][i][0] * kGrad[t][j][0] / (hx * hx);
        const double gy = kGrad[t][i][1] * kGrad[t][j][1] / (hy * hy);
        k[t][i][j] = eps * eps * area * (gx + gy) + area / 12.0 * (i == j ? 2.0 : 1.0);
        for (int e = 0; e < 4; ++e) L.kt[t][i][j][e] = (T)(e == 0 ? k[t][i][j] : 0.0);
      }
  // interior row: every triangle present, summed in the device's order, in T
  T row[9] = {};
  for (const auto& tr : kTrisHost)
    for (int q = 0; q < 3; ++q) {
      const int kk = (tr[5 + q] + 1) * 3 + (tr[2 + q] + 1);
      row[kk] += T(1) * (T)k[tr[0]][tr[1]][q];
    }
  for (int e = 0; e < 9 * 4; ++e) L.urow[e] = T(0);
  for (int kk = 0; kk < 9; ++kk) L.urow[kk * bs * bs] = row[kk];
  for (int e = 0; e < 4; ++e) L.udinv[e] = T(0);
  L.udinv[0] = T(1) / row[4];
  if (bs == 2) L.udinv[3] = T(1) / row[4];  // (the filter is scalar; kept well-defined)
}

template void band_elastic_table<double>(BandLevel<double>&, double, double, double, double);
template void band_elastic_table<float>(BandLevel<float>&, double, double, double, double);
template void band_filter_table<double>(BandLevel<double>&, double, double, double, int);
template void band_filter_table<float>(BandLevel<float>&, double, double, double, int);

size_t band_ext(int t, int nu, bool down) {
  return (size_t)t + 2 * ((size_t)nu + (down ? 1 : 0));
}

size_t band_lds_bytes(int tx, int ty, int nu, int bs, size_t elem, int src, bool down) {
  const size_t EX = band_ext(tx, nu, down), EY = band_ext(ty, nu, down);
  size_t n = 3 * (size_t)bs * EX * EY;
  if (src == kSrcCells) n += 2 * (EX + 1) * (EY + 1);
  return n * elem;
}

int band_tiles(int nx, int ny, int tx, int ty) { return ((nx + tx) / tx) * ((ny + ty) / ty); }

template <typename T>
static bool band_valid(const BandLevel<T>& L) {
  for (int nu : {L.nu_down, L.nu_up})
    if (nu < 1 || nu > kMaxNu) return false;
  if (L.tx < 2 || L.ty < 2 || (L.tx & 1) || (L.ty & 1) || L.tx > 256 || L.ty > 256) return false;
  if (L.nx < 1 || L.ny < 1 || L.nxc < 1 || L.nyc < 1) return false;
  if ((size_t)(L.nx + 1) * (L.ny + 1) >= (1ull << 31)) return false;
  if (L.src == kSrcStencil ? L.st == nullptr
                           : L.src == kSrcCells ? L.cw == nullptr : L.src != kSrcUnit)
    return false;
  if (L.nested) return L.nx == 2 * L.nxc && L.ny == 2 * L.nyc;
  const bool okx = L.nxc == L.nx ? L.nx == 1 : L.nxc == (L.nx + 1) / 2;
  const bool oky = L.nyc == L.ny ? L.ny == 1 : L.nyc == (L.ny + 1) / 2;
  return okx && oky;
}

template <typename T, int BS, int SRC, bool ONE, int NT = kBandThreads>
static void launch_down(const BandLevel<T>& L, const T* b, T* x, T* bc, size_t lds,
                        hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(NT, (band_down_body<T, BS, SRC, ONE, NT>), band_tiles(L.nx, L.ny, L.tx, L.ty), lds, s, (k_band_down<T, BS, SRC, ONE, NT>), L, b, x, bc);
}

template <typename T, int BS, int SRC, bool ONE, int NT = kBandThreads>
static void launch_up(const BandLevel<T>& L, const T* x, const T* ec, const T* b, T* y,
                      double* dot, size_t lds, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(NT, (band_up_body<T, BS, SRC, ONE, NT>), band_tiles(L.nx, L.ny, L.tx, L.ty), lds, s, (k_band_up<T, BS, SRC, ONE, NT>), L, x, ec, b, y, dot);
}

template <typename T, int BS>
hipError_t band_down(const BandLevel<T>& L, const T* b, T* x, T* bc, hipStream_t s) {
  if (!band_valid(L)) return hipErrorInvalidValue;
  const size_t lds = band_lds_bytes(L.tx, L.ty, L.nu_down, BS, sizeof(T), L.src, true);
  if (lds > kBandLdsBytes) return hipErrorInvalidValue;
  const size_t ne = band_ext(L.tx, L.nu_down, true) * band_ext(L.ty, L.nu_down, true);
  const bool one = ne <= (size_t)kBandThreads;
  if (L.wide == 2 && L.src == kSrcCells) {
    if (ne <= (size_t)kBandThreadsMid)
      launch_down<T, BS, kSrcCells, true, kBandThreadsMid>(L, b, x, bc, lds, s);
    else
      launch_down<T, BS, kSrcCells, false, kBandThreadsMid>(L, b, x, bc, lds, s);
  } else if (L.wide == 3 && L.src == kSrcCells) {
    if (ne <= (size_t)kBandThreadsWide)
      launch_down<T, BS, kSrcCells, true, kBandThreadsWide>(L, b, x, bc, lds, s);
    else
      launch_down<T, BS, kSrcCells, false, kBandThreadsWide>(L, b, x, bc, lds, s);
  } else if (L.wide == 1 && L.src == kSrcCells && !one && ne <= (size_t)kBandThreadsWide) {
    launch_down<T, BS, kSrcCells, true, kBandThreadsWide>(L, b, x, bc, lds, s);
  } else if (L.src == kSrcStencil) {
    if (one) launch_down<T, BS, kSrcStencil, true>(L, b, x, bc, lds, s);
    else launch_down<T, BS, kSrcStencil, false>(L, b, x, bc, lds, s);
  } else if (L.src == kSrcCells) {
    if (one) launch_down<T, BS, kSrcCells, true>(L, b, x, bc, lds, s);
    else launch_down<T, BS, kSrcCells, false>(L, b, x, bc, lds, s);
  } else {
    if (one) launch_down<T, BS, kSrcUnit, true>(L, b, x, bc, lds, s);
    else launch_down<T, BS, kSrcUnit, false>(L, b, x, bc, lds, s);
  }
  return hipGetLastError();
}

template <typename T, int BS>
hipError_t band_up(const BandLevel<T>& L, const T* x, const T* ec, const T* b, T* y, RedSlot* dot,
                   hipStream_t s) {
  if (!band_valid(L)) return hipErrorInvalidValue;
  const size_t lds = band_lds_bytes(L.tx, L.ty, L.nu_up, BS, sizeof(T), L.src, false);
  if (lds > kBandLdsBytes) return hipErrorInvalidValue;
  const int tiles = band_tiles(L.nx, L.ny, L.tx, L.ty);
  double* part = nullptr;
  if (dot) {
    if (tiles > kMaxPartials) return hipErrorInvalidValue;
    dot->nblocks = tiles;
    part = dot->partials;
  }
  const size_t ne = band_ext(L.tx, L.nu_up, false) * band_ext(L.ty, L.nu_up, false);
  const bool one = ne <= (size_t)kBandThreads;
  if (L.wide == 2 && L.src == kSrcCells && !part) {
    if (ne <= (size_t)kBandThreadsMid)
      launch_up<T, BS, kSrcCells, true, kBandThreadsMid>(L, x, ec, b, y, part, lds, s);
    else
      launch_up<T, BS, kSrcCells, false, kBandThreadsMid>(L, x, ec, b, y, part, lds, s);
  } else if (L.wide == 3 && L.src == kSrcCells && !part) {
    if (ne <= (size_t)kBandThreadsWide)
      launch_up<T, BS, kSrcCells, true, kBandThreadsWide>(L, x, ec, b, y, part, lds, s);
    else
      launch_up<T, BS, kSrcCells, false, kBandThreadsWide>(L, x, ec, b, y, part, lds, s);
  } else if (L.wide == 1 && L.src == kSrcCells && !one && ne <= (size_t)kBandThreadsWide &&
             !part) {
    launch_up<T, BS, kSrcCells, true, kBandThreadsWide>(L, x, ec, b, y, part, lds, s);
  } else if (L.src == kSrcStencil) {
    if (one) launch_up<T, BS, kSrcStencil, true>(L, x, ec, b, y, part, lds, s);
    else launch_up<T, BS, kSrcStencil, false>(L, x, ec, b, y, part, lds, s);
  } else if (L.src == kSrcCells) {
    if (one) launch_up<T, BS, kSrcCells, true>(L, x, ec, b, y, part, lds, s);
    else launch_up<T, BS, kSrcCells, false>(L, x, ec, b, y, part, lds, s);
  } else {
    if (one) launch_up<T, BS, kSrcUnit, true>(L, x, ec, b, y, part, lds, s);
    else launch_up<T, BS, kSrcUnit, false>(L, x, ec, b, y, part, lds, s);
  }
  return hipGetLastError();
}

#define TMX_INST_BAND(T, BS)                                                                  \
  template hipError_t band_down<T, BS>(const BandLevel<T>&, const T*, T*, T*, hipStream_t);   \
  template hipError_t band_up<T, BS>(const BandLevel<T>&, const T*, const T*, const T*, T*,   \
                                     RedSlot*, hipStream_t);
TMX_INST_BAND(double, 1)
TMX_INST_BAND(double, 2)
TMX_INST_BAND(float, 1)
TMX_INST_BAND(float, 2)

}  // namespace tmx
