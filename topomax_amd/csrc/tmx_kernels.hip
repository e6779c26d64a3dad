// gfx950 kernels for the topomax elasticity hot path: matrix-free SIMP-P1 elasticity and
// Helmholtz-filter operators on the structured 'right'-diagonal triangle mesh, geometric
// multigrid transfers, Galerkin coarse stencils, and the vector kernels of PCG.
//
// Column-sweep design (elasticity and filter operators): one wavefront owns a strip of 64
// node columns (62 output columns, one halo column each side) and walks a chunk of node
// rows bottom to top.  Each lane computes the cell whose lower-left node is its own column:
// it loads only its own column's values (coalesced across the wave), obtains the right
// neighbour column by a cross-lane move, and hands the forces of the cell's right nodes to
// the right lane the same way.  Every HBM value is read once per sweep (plus one halo row
// per chunk and two halo columns per 64), no LDS is used, and a node row is written as soon
// as the two cell rows around it are done.
//
// Several sweeps stack in one pass the same way (k_elast_sweep_pair, k_elast_sweep_pair_rr,
// k_filter_pipe_*): each stage trails the one before by one cell row, so intermediate iterates
// never leave registers, and each stage costs one more halo column on each side of the strip.

#include "tmx_device.h"
#include "tmx_kernels.h"

#include <algorithm>

namespace tmx {

// ============================================================================ shapes
SweepShape SweepShape::make(int nx, int ny, int rows, int stripw) {
  SweepShape s;
  s.nx = nx;
  s.ny = ny;
  s.rows = rows < 1 ? 1 : rows;
  s.stripw = stripw;
  s.nstrips = (nx + 1 + stripw - 1) / stripw;
  s.nchunks = (ny + 1 + s.rows - 1) / s.rows;
  return s;
}

int SweepShape::blocks() const {
  const int waves = nstrips * nchunks;
  const int wpb = kBlock / kWave;
  return (waves + wpb - 1) / wpb;
}

template <typename T>
ElastConst<T> ElastConst<T>::make(double lam, double mu, double hx, double hy) {
  const double a = hy / hx, b = hx / hy;
  ElastConst<T> k;
  k.l2m_a = T((lam + 2 * mu) * a / 2);
  k.lam_h = T(lam / 2);
  k.l2m_b = T((lam + 2 * mu) * b / 2);
  k.mu_h = T(mu / 2);
  k.mu_a = T(mu * a / 2);
  k.mu_b = T(mu * b / 2);
  k.Pxx = T((lam + 2 * mu) * a / 2);
  k.Pyy = T(mu * a / 2);
  k.Qxx = T(mu * b / 2);
  k.Qyy = T((lam + 2 * mu) * b / 2);
  k.Rxx = T((lam + 2 * mu) * a / 2 + mu * b / 2);
  k.Rxy = T(-(lam + mu) / 2);
  k.Ryy = T((lam + 2 * mu) * b / 2 + mu * a / 2);
  return k;
}
template struct ElastConst<double>;
template struct ElastConst<float>;

template <typename T>
FilterConst<T> FilterConst<T>::make(double eps, double hx, double hy) {
  FilterConst<T> k;
  k.ka = T(eps * eps * (hy / hx) / 2);
  k.kb = T(eps * eps * (hx / hy) / 2);
  k.mA = T(hx * hy / 2 / 12);
  return k;
}
template struct FilterConst<double>;
template struct FilterConst<float>;

static inline unsigned grid1d(size_t n, int block = kBlock) {
  size_t g = (n + block - 1) / block;
  return (unsigned)std::max<size_t>(1, g);
}

// ============================================================================ elasticity
// Forces of one cell (triangles A = (v0,v1,v3) and B = (v0,v2,v3)) on its four nodes,
// given the triangles' mean SIMP factors wA, wB.  K_T = wT |T| B^T D B with
// D = [[l+2m, l, 0], [l, l+2m, 0], [0, 0, m]], written out for the two right triangles.
template <typename T>
__device__ __forceinline__ void cell_forces(const ElastConst<T>& k, T wA, T wB, V2<T> u0, V2<T> u1,
                                            V2<T> u2, V2<T> u3, V2<T>& f0, V2<T>& f1, V2<T>& f2,
                                            V2<T>& f3) {
  {  // triangle A: grad l0 = (-1/hx, 0), grad l1 = (1/hx, -1/hy), grad l3 = (0, 1/hy)
    const T e1 = u1.x - u0.x, e2 = u3.y - u1.y, g1 = u3.x - u1.x, g2 = u1.y - u0.y;
    const T Sx = wA * (k.l2m_a * e1 + k.lam_h * e2);
    const T Sy = wA * (k.lam_h * e1 + k.l2m_b * e2);
    const T T1 = wA * (k.mu_h * g1 + k.mu_a * g2);
    const T T2 = wA * (k.mu_b * g1 + k.mu_h * g2);
    f0 = {-Sx, -T1};
    f1 = {Sx - T2, T1 - Sy};
    f3 = {T2, Sy};
  }
  {  // triangle B: grad l0 = (0, -1/hy), grad l2 = (-1/hx, 1/hy), grad l3 = (1/hx, 0)
    const T e1 = u3.x - u2.x, e2 = u2.y - u0.y, g1 = u2.x - u0.x, g2 = u3.y - u2.y;
    const T Sx = wB * (k.l2m_a * e1 + k.lam_h * e2);
    const T Sy = wB * (k.lam_h * e1 + k.l2m_b * e2);
    const T T1 = wB * (k.mu_h * g1 + k.mu_a * g2);
    const T T2 = wB * (k.mu_b * g1 + k.mu_h * g2);
    f0 = f0 + V2<T>{-T2, -Sy};
    f2 = {T2 - Sx, Sy - T1};
    f3 = f3 + V2<T>{Sx, T1};
  }
}

template <typename T>
struct Diag3 {
  T xx, xy, yy;
};

// One row's inputs of a column sweep, loaded ahead of the row's turn: nodal xi (or a cell's
// weights) and a right-hand side / iterate value.  Rows in flight per wave:
constexpr int kPrefetchRows = 4;
template <typename T>
struct RowIn {
  T z = T(0);
  V2<T> w = {T(0), T(0)};
  V2<T> b = {T(0), T(0)};
  V2<T> x = {T(0), T(0)}, e = {T(0), T(0)}, e1 = {T(0), T(0)};  // iterate, coarse correction
};

template <typename T, bool XI, int OP>
__device__ __forceinline__ void elast_sweep_body(const BlkId blk, const ElastOpArgs<T>& a,
                                                 const T* x, const T* b, T* y, T omega,
                                                 double* dpart, const SweepExtra<T>& ex) {
  using V = V2<T>;
  constexpr bool RR = (OP == kResidualRestrict);  // residual -> nested restriction (y coarse)
  constexpr bool PS = (OP == kProlongSmooth);     // x + P e, then one Jacobi sweep
  constexpr bool AU = (OP == kApplyUpdate);       // p = z + beta p_old, y = A p
  constexpr bool SMOOTH = (OP == kSmooth || OP == kSmoothZero || PS);
  constexpr int H = RR ? 2 : 1;                   // halo lanes on each side of the strip
  const SweepShape sh = a.sh;
  const bool want_dot = dpart != nullptr;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blk.id);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1;
  const int nxc = nx >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const unsigned fixed = a.fixed;
  const bool colfix = ((fixed & kLeft) && c == 0) || ((fixed & kRight) && c == nx);
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);  // owned node rows [r0, rend)
  // cell rows swept: a node row is complete once the cell rows below and above it are done
  const int cy_begin = RR ? r0 - 2 : r0 - 1;
  const int cy_end = !wave_ok ? cy_begin : (RR ? min(rend, ny) + 1 : rend);
  const auto X = glob(reinterpret_cast<const V*>(x));
  const auto B = glob(reinterpret_cast<const V*>(b));
  const auto Y = glob(reinterpret_cast<V*>(y));
  const auto CW = glob(reinterpret_cast<const V*>(a.cellw));
  const auto XIP = glob(a.xi);
  const auto EC = glob(reinterpret_cast<const V*>(ex.ec));
  const auto PN = glob(reinterpret_cast<V*>(ex.pnew));
  const auto ZL = glob(reinterpret_cast<const float2*>(ex.zlow));
  const T beta = AU ? (T)*glob(ex.beta) : T(0);
  const auto XK = glob(reinterpret_cast<V*>(ex.xk));
  const T alpha = (AU && ex.xk) ? (T)*glob(ex.alpha) : T(0);
  const ElastConst<T> k = a.k;

  auto row_fixed = [&](int iy) {
    return colfix || ((fixed & kBottom) && iy == 0) || ((fixed & kTop) && iy == ny);
  };
  V po_next = {T(0), T(0)};  // AU: raw p_old of the row ld_x last loaded (for the deferred x)
  auto ld_x = [&](int iy, V& raw) -> V {
    V v = {T(0), T(0)};
    const bool ok = colv && iy >= 0 && iy <= ny;
    if (OP != kSmoothZero && ok) v = X[(unsigned)(iy * nvx + c)];
    if constexpr (AU) {
      po_next = v;
      if (ok) {
        V z;
        if (ZL) {
          const float2 zl = ZL[(unsigned)(iy * nvx + c)];
          z = {(T)zl.x, (T)zl.y};
        } else {
          z = B[(unsigned)(iy * nvx + c)];
        }
        v = {z.x + beta * v.x, z.y + beta * v.y};
      }
    }
    if constexpr (PS) {
      // nested P1 interpolation of the coarse correction at (c, iy)
      if (ok && !row_fixed(iy)) {
        const int I0 = c >> 1, J0 = iy >> 1, ox = c & 1, oy = iy & 1;
        V e = EC[(size_t)J0 * (nxc + 1) + I0];
        if (ox | oy) {
          const V e1 = EC[(size_t)(J0 + oy) * (nxc + 1) + I0 + ox];
          e = {T(0.5) * (e.x + e1.x), T(0.5) * (e.y + e1.y)};
        }
        v += e;
      }
    }
    raw = v;
    if (row_fixed(iy)) v = {T(0), T(0)};
    return v;
  };
  auto ld_xi = [&](int iy) -> T {
    if (XI && colv && iy >= 0 && iy <= ny) return XIP[(unsigned)(iy * nvx + c)];
    return T(0);
  };

  int cy = cy_begin;
  V ub_raw;
  V ub = ld_x(cy, ub_raw);
  V po_b = po_next;
  T zb = ld_xi(cy);
  V ubr = lane_down(ub);
  T zbr = lane_down(zb);
  V acc_b = {T(0), T(0)};
  Diag3<T> d_b = {T(0), T(0), T(0)};
  double dot = 0.0;
  V rA = {T(0), T(0)}, rB = {T(0), T(0)};  // residual rows cy-2, cy-1 (RR)

  for (; cy < cy_end; ++cy) {
    V ut_raw;
    const V ut = ld_x(cy + 1, ut_raw);
    const V po_t = po_next;
    const T zt = ld_xi(cy + 1);
    const V utr = lane_down(ut);
    const T ztr = lane_down(zt);
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    T wA = T(0), wB = T(0);
    if (cell_ok) {
      if (XI) {
        wA = pc_weight(a.m + a.cp * hsym3(zb, zbr, ztr, a.p), a.wmin);
        wB = pc_weight(a.m + a.cp * hsym3(zb, zt, ztr, a.p), a.wmin);
      } else {
        const V w = CW[(unsigned)(cy * nx + c)];
        wA = w.x;
        wB = w.y;
      }
    }
    V acc_t = {T(0), T(0)};
    if constexpr (OP != kSmoothZero) {
      V f0, f1, f2, f3;
      cell_forces(k, wA, wB, ub, ubr, ut, utr, f0, f1, f2, f3);
      acc_b += f0;
      acc_t = f2;
      acc_b += lane_up(f1);
      acc_t += lane_up(f3);
    }
    Diag3<T> d_t = {T(0), T(0), T(0)};
    if constexpr (SMOOTH) {
      const T wAl = lane_up(wA), wBl = lane_up(wB);
      // own bottom node = v0 of this cell (A and B) and v1 of the left cell (A)
      d_b.xx += wA * k.Pxx + wB * k.Qxx + wAl * k.Rxx;
      d_b.xy += wAl * k.Rxy;
      d_b.yy += wA * k.Pyy + wB * k.Qyy + wAl * k.Ryy;
      // own top node = v2 of this cell (B) and v3 of the left cell (A and B)
      d_t.xx = wB * k.Rxx + wAl * k.Qxx + wBl * k.Pxx;
      d_t.xy = wB * k.Rxy;
      d_t.yy = wB * k.Ryy + wAl * k.Qyy + wBl * k.Pyy;
    }
    // node row cy is complete
    const bool rowv = colv && cy >= 0 && cy <= ny;
    const unsigned idx = (unsigned)(cy * nvx + c);
    const bool fx = row_fixed(cy);
    if constexpr (RR) {
      const V Ax = fx ? ub_raw : acc_b;
      V rC = {T(0), T(0)};
      if (rowv) rC = B[idx] - Ax;
      // coarse row J: centre fine row 2J = cy-1 (cy odd), or 2J = ny at the top (ny even)
      const bool odd_emit = (cy & 1) && cy - 1 >= r0 && cy - 1 < rend;
      const bool top_emit = cy == ny && ny >= r0 && ny < rend;
      if (odd_emit || top_emit) {
        const V lo = odd_emit ? rA : rB, mid = odd_emit ? rB : rC;
        const V hi = odd_emit ? rC : V{T(0), T(0)};
        const V mr = lane_down(mid), ml = lane_up(mid), hr = lane_down(hi), ll = lane_up(lo);
        const int J = odd_emit ? (cy - 1) >> 1 : ny >> 1;
        if (outl && !(c & 1)) {
          const int I = c >> 1;
          V o = {T(0), T(0)};
          if (!node_fixed(fixed, I, J, nxc, ny >> 1)) {
            o.x = mid.x + T(0.5) * (mr.x + ml.x + hi.x + lo.x + hr.x + ll.x);
            o.y = mid.y + T(0.5) * (mr.y + ml.y + hi.y + lo.y + hr.y + ll.y);
          }
          Y[(size_t)J * (nxc + 1) + I] = o;
        }
      }
      rA = rB;
      rB = rC;
    } else if (cy >= r0 && outl) {
      const V Ax = fx ? ub_raw : acc_b;
      if constexpr (OP == kApply || AU) {
        Y[idx] = Ax;
        if constexpr (AU) {
          PN[idx] = ub_raw;
          if (XK) {  // deferred x update of the previous iteration (p_old kept from the load)
            V xo = XK[idx];
            xo.x += alpha * po_b.x;
            xo.y += alpha * po_b.y;
            XK[idx] = xo;
          }
        }
        if (want_dot) dot += (double)ub_raw.x * (double)Ax.x + (double)ub_raw.y * (double)Ax.y;
      } else if constexpr (OP == kResidual) {
        Y[idx] = B[idx] - Ax;
      } else {
        const V bb = B[idx];
        const V res = bb - Ax;
        V d;
        if (fx) {
          d = res;
        } else {
          const T det = d_b.xx * d_b.yy - d_b.xy * d_b.xy;
          const T inv = recip(det);
          d = {(d_b.yy * res.x - d_b.xy * res.y) * inv, (d_b.xx * res.y - d_b.xy * res.x) * inv};
        }
        const V yo = {ub_raw.x + omega * d.x, ub_raw.y + omega * d.y};
        Y[idx] = yo;
        if (want_dot) dot += (double)bb.x * (double)yo.x + (double)bb.y * (double)yo.y;
      }
    }
    acc_b = acc_t;
    d_b = d_t;
    ub = ut;
    ub_raw = ut_raw;
    po_b = po_t;
    ubr = utr;
    zb = zt;
    zbr = ztr;
  }
  if constexpr (!RR) {
    if (want_dot) {
      const double v[1] = {dot};
      double* const o[1] = {dpart};
      block_partials<1>(v, o, blk.id);
    }
  }
}

template <typename T, bool XI, int OP>
__global__ __launch_bounds__(kBlock) void k_elast_sweep(ElastOpArgs<T> a, const T* __restrict__ x,
                                                        const T* __restrict__ b, T* __restrict__ y,
                                                        T omega, double* __restrict__ dpart,
                                                        SweepExtra<T> ex) {
  elast_sweep_body<T, XI, OP>(hw_blk(), a, x, b, y, omega, dpart, ex);
}

template <typename T>
hipError_t elast_sweep(const ElastOpArgs<T>& a0, int op, const T* x, const T* b, T* y, T omega,
                       RedSlot* dot_xy, hipStream_t s, const SweepExtra<T>& ex) {
  const T* ec = ex.ec;
  ElastOpArgs<T> a = a0;
  if (op == kResidualRestrict) {
    if ((a.sh.nx & 1) || (a.sh.ny & 1)) return hipErrorInvalidValue;
    a.sh = SweepShape::make(a.sh.nx, a.sh.ny, std::max(2, a.sh.rows & ~1), kStripCols - 2);
  }
  if (op == kProlongSmooth && (!ec || (a.sh.nx & 1) || (a.sh.ny & 1))) return hipErrorInvalidValue;
  if (op == kApplyUpdate && (!ex.pnew || !ex.beta || (!b && !ex.zlow))) return hipErrorInvalidValue;
  if (op == kResidualRestrict && dot_xy) return hipErrorInvalidValue;
  if ((size_t)(a.sh.nx + 1) * (a.sh.ny + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  const int nb = a.sh.blocks();
  double* dp = nullptr;
  if (dot_xy) {
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    dp = dot_xy->partials;
    dot_xy->nblocks = nb;
  }
  const bool xi = a.xi != nullptr;
#define TMX_LAUNCH(XIV, OPV)                                                              \
  TMX_LAUNCH_OR_RECORD(kBlock, (elast_sweep_body<T, XIV, OPV>), nb, 0, s,                   \
                       (k_elast_sweep<T, XIV, OPV>), a, x, b, y, omega, dp, ex)
#define TMX_CASE(OPV) \
  case OPV: if (xi) TMX_LAUNCH(true, OPV); else TMX_LAUNCH(false, OPV); break;
  switch (op) {
    TMX_CASE(kApply)
    TMX_CASE(kResidual)
    TMX_CASE(kSmooth)
    TMX_CASE(kSmoothZero)
    TMX_CASE(kResidualRestrict)
    TMX_CASE(kProlongSmooth)
    TMX_CASE(kApplyUpdate)
    default: return hipErrorInvalidValue;
  }
#undef TMX_CASE
#undef TMX_LAUNCH
  return hipGetLastError();
}
template hipError_t elast_sweep<double>(const ElastOpArgs<double>&, int, const double*,
                                        const double*, double*, double, RedSlot*, hipStream_t,
                                        const SweepExtra<double>&);
template hipError_t elast_sweep<float>(const ElastOpArgs<float>&, int, const float*, const float*,
                                       float*, float, RedSlot*, hipStream_t,
                                       const SweepExtra<float>&);

// Two consecutive block-Jacobi sweeps in one pass over the grid: the first is kSmoothZero
// (from x = 0) or kProlongSmooth (from x + P e), the second a plain sweep of its result.  The
// wave's 64 lanes carry two halo columns on each side (60 output columns: the first sweep's
// result is right in lanes 1..63, the second's in lanes 2..61).  The second sweep trails the
// first by one cell row, so the intermediate iterate never leaves registers: per cell row cy
// the first sweep forms the cell's weights, diagonal blocks and (kProlongSmooth) forces and
// completes its node row cy; the second sweep then takes cell row cy - 1 between the first
// sweep's node rows cy - 1 and cy and completes its node row cy - 1.  A chunk of R node rows
// sweeps R + 3 cell rows (node rows r0 - 1 .. rend of the first sweep).
template <typename T, bool XI, int OP1>
__device__ __forceinline__ void elast_sweep_pair_body(const BlkId blk, const ElastOpArgs<T>& a,
                                                      const T* x, const T* b, T* y, T om1, T om2,
                                                      double* dpart, const T* ec) {
  using V = V2<T>;
  constexpr bool PS = (OP1 == kProlongSmooth);
  constexpr int H = 2;
  const SweepShape sh = a.sh;
  const bool want_dot = dpart != nullptr;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blk.id);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1;
  const int nxc = nx >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const unsigned fixed = a.fixed;
  const bool colfix = ((fixed & kLeft) && c == 0) || ((fixed & kRight) && c == nx);
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);  // owned node rows [r0, rend)
  const int cy_begin = r0 - 2;
  const int cy_end = wave_ok ? rend + 1 : cy_begin;
  const int first_row = max(0, r0 - 1);  // first node row of the first sweep with a full diagonal
  const auto X = glob(reinterpret_cast<const V*>(x));
  const auto B = glob(reinterpret_cast<const V*>(b));
  const auto Y = glob(reinterpret_cast<V*>(y));
  const auto CW = glob(reinterpret_cast<const V*>(a.cellw));
  const auto XIP = glob(a.xi);
  const auto EC = glob(reinterpret_cast<const V*>(ec));
  const ElastConst<T> k = a.k;
  const V zero = {T(0), T(0)};

  auto row_fixed = [&](int iy) {
    return colfix || ((fixed & kBottom) && iy == 0) || ((fixed & kTop) && iy == ny);
  };
  auto ld_x = [&](int iy, V& raw) -> V {  // x + P e (kProlongSmooth)
    V v = zero;
    const bool ok = colv && iy >= 0 && iy <= ny;
    if (ok) v = X[(unsigned)(iy * nvx + c)];
    if (ok && !row_fixed(iy)) {
      const int I0 = c >> 1, J0 = iy >> 1, ox = c & 1, oy = iy & 1;
      V e = EC[(size_t)J0 * (nxc + 1) + I0];
      if (ox | oy) {
        const V e1 = EC[(size_t)(J0 + oy) * (nxc + 1) + I0 + ox];
        e = {T(0.5) * (e.x + e1.x), T(0.5) * (e.y + e1.y)};
      }
      v += e;
    }
    raw = v;
    if (row_fixed(iy)) v = zero;
    return v;
  };
  auto ld_xi = [&](int iy) -> T {
    if (XI && colv && iy >= 0 && iy <= ny) return XIP[(unsigned)(iy * nvx + c)];
    return T(0);
  };

  int cy = cy_begin;
  // first sweep: cell row cy between node rows cy (u_b) and cy + 1
  V ub = zero, ub_raw = zero, ubr = zero;
  if constexpr (PS) {
    ub = ld_x(cy, ub_raw);
    ubr = lane_down(ub);
  }
  T zb = ld_xi(cy);
  T zbr = lane_down(zb);
  V acc1 = zero;                            // forces on node row cy from cell row cy - 1
  Diag3<T> d_b = {T(0), T(0), T(0)};        // its diagonal block from cell row cy - 1
  // second sweep: cell row cy - 1 between node rows cy - 1 (v_b) and cy
  V vb = zero, vb_raw = zero, vbr = zero;   // first-sweep iterate at node row cy - 1
  V acc2 = zero;                            // forces on node row cy - 1 from cell row cy - 2
  V bp = zero;                              // b at node row cy - 1
  Diag3<T> dp = {T(0), T(0), T(0)};         // diagonal block of node row cy - 1
  T invp = T(0);
  T wAp = T(0), wBp = T(0);                 // weights of cell row cy - 1
  double dot = 0.0;

  // Loads of cell row cy (x + P e and xi of node row cy + 1, the cell weights, b of node row
  // cy) are issued kPrefetchRows rows ahead from clamped in-grid addresses; validity, fixed
  // nodes and the interpolation are applied at the row's turn (see k_elast_sweep_pair_rr)
  const int cn = min(max(c, 0), nx), cc = min(max(c, 0), nx - 1);
  const int I0 = cn >> 1, ox = cn & 1;
  auto fetch = [&](int row) {
    RowIn<T> r;
    const int nt = min(max(row + 1, 0), ny), nb = min(max(row, 0), ny);
    const int cr = min(max(row, 0), ny - 1);
    if constexpr (PS) {
      const int J0 = nt >> 1, oy = nt & 1;
      r.x = X[(unsigned)(nt * nvx + cn)];
      r.e = EC[(size_t)J0 * (nxc + 1) + I0];
      r.e1 = EC[(size_t)(J0 + oy) * (nxc + 1) + I0 + ox];
    }
    if constexpr (XI) r.z = XIP[(unsigned)(nt * nvx + cn)];
    else r.w = CW[(unsigned)(cr * nx + cc)];
    r.b = B[(unsigned)(nb * nvx + cn)];
    return r;
  };
  // x + P e at node row iy from its prefetched values (as ld_x)
  auto form_x = [&](int iy, const RowIn<T>& in, V& raw) -> V {
    V v = zero;
    const bool ok = colv && iy >= 0 && iy <= ny;
    if (ok) v = in.x;
    if (ok && !row_fixed(iy)) {
      V e = in.e;
      if (ox | (iy & 1)) e = {T(0.5) * (e.x + in.e1.x), T(0.5) * (e.y + in.e1.y)};
      v += e;
    }
    raw = v;
    if (row_fixed(iy)) v = zero;
    return v;
  };
  RowIn<T> pf[kPrefetchRows];
  if (wave_ok) {
#pragma unroll
    for (int q = 0; q < kPrefetchRows; ++q) pf[q] = fetch(cy_begin + q);
  }
  for (int cy0 = cy_begin; cy0 < cy_end; cy0 += kPrefetchRows) {
#pragma unroll
   for (int q = 0; q < kPrefetchRows; ++q) {
    cy = cy0 + q;
    if (cy >= cy_end) break;
    const RowIn<T> in = pf[q];
    // ---- first sweep, cell row cy
    V ut = zero, ut_raw = zero, utr = zero;
    if constexpr (PS) {
      ut = form_x(cy + 1, in, ut_raw);
      utr = lane_down(ut);
    }
    const T zt = (XI && colv && cy + 1 >= 0 && cy + 1 <= ny) ? in.z : T(0);
    const T ztr = lane_down(zt);
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    T wA = T(0), wB = T(0);
    if (cell_ok) {
      if (XI) {
        wA = pc_weight(a.m + a.cp * hsym3(zb, zbr, ztr, a.p), a.wmin);
        wB = pc_weight(a.m + a.cp * hsym3(zb, zt, ztr, a.p), a.wmin);
      } else {
        wA = in.w.x;
        wB = in.w.y;
      }
    }
    V acc1t = zero;
    if constexpr (PS) {
      V f0, f1, f2, f3;
      cell_forces(k, wA, wB, ub, ubr, ut, utr, f0, f1, f2, f3);
      acc1 += f0;
      acc1t = f2;
      acc1 += lane_up(f1);
      acc1t += lane_up(f3);
    }
    const T wAl = lane_up(wA), wBl = lane_up(wB);
    d_b.xx += wA * k.Pxx + wB * k.Qxx + wAl * k.Rxx;
    d_b.xy += wAl * k.Rxy;
    d_b.yy += wA * k.Pyy + wB * k.Qyy + wAl * k.Ryy;
    Diag3<T> d_t;
    d_t.xx = wB * k.Rxx + wAl * k.Qxx + wBl * k.Pxx;
    d_t.xy = wB * k.Rxy;
    d_t.yy = wB * k.Ryy + wAl * k.Qyy + wBl * k.Pyy;
    // node row cy of the first sweep is complete
    const bool rowv = colv && cy >= first_row && cy <= ny;
    const bool fx = row_fixed(cy);
    V bb = zero, v_raw = zero;
    T inv = T(0);
    if (rowv) {
      bb = in.b;
      const V Ax = PS ? (fx ? ub_raw : acc1) : zero;
      const V res = bb - Ax;
      V d;
      if (fx) {
        d = res;
      } else {
        inv = recip(d_b.xx * d_b.yy - d_b.xy * d_b.xy);
        d = {(d_b.yy * res.x - d_b.xy * res.y) * inv, (d_b.xx * res.y - d_b.xy * res.x) * inv};
      }
      v_raw = {ub_raw.x + om1 * d.x, ub_raw.y + om1 * d.y};
    }
    const V vt = fx ? zero : v_raw;
    const V vtr = lane_down(vt);
    // ---- second sweep, cell row cy - 1
    {
      V f0, f1, f2, f3;
      cell_forces(k, wAp, wBp, vb, vbr, vt, vtr, f0, f1, f2, f3);
      acc2 += f0;
      acc2 += lane_up(f1);
      const int ry = cy - 1;  // node row ry of the second sweep is complete
      if (ry >= r0 && ry < rend && outl) {
        const bool fx2 = row_fixed(ry);
        const V Ax = fx2 ? vb_raw : acc2;
        const V res = bp - Ax;
        V d;
        if (fx2) {
          d = res;
        } else {
          d = {(dp.yy * res.x - dp.xy * res.y) * invp, (dp.xx * res.y - dp.xy * res.x) * invp};
        }
        const V yo = {vb_raw.x + om2 * d.x, vb_raw.y + om2 * d.y};
        Y[(unsigned)(ry * nvx + c)] = yo;
        if (want_dot) dot += (double)bp.x * (double)yo.x + (double)bp.y * (double)yo.y;
      }
      acc2 = f2;
      acc2 += lane_up(f3);
    }
    vb = vt;
    vb_raw = v_raw;
    vbr = vtr;
    bp = bb;
    dp = d_b;
    invp = inv;
    wAp = wA;
    wBp = wB;
    acc1 = acc1t;
    d_b = d_t;
    ub = ut;
    ub_raw = ut_raw;
    ubr = utr;
    zb = zt;
    zbr = ztr;
    pf[q] = fetch(cy + kPrefetchRows);
   }
  }
  if (want_dot) {
    const double v[1] = {dot};
    double* const o[1] = {dpart};
    block_partials<1>(v, o, blk.id);
  }
}

template <typename T, bool XI, int OP1>
__global__ __launch_bounds__(kBlock) void k_elast_sweep_pair(ElastOpArgs<T> a,
                                                             const T* __restrict__ x,
                                                             const T* __restrict__ b,
                                                             T* __restrict__ y, T om1, T om2,
                                                             double* __restrict__ dpart,
                                                             const T* __restrict__ ec) {
  elast_sweep_pair_body<T, XI, OP1>(hw_blk(), a, x, b, y, om1, om2, dpart, ec);
}

template <typename T>
hipError_t elast_sweep_pair(const ElastOpArgs<T>& a0, int op1, int rows, const T* x, const T* b,
                            T* y, T om1, T om2, RedSlot* dot_by, hipStream_t s, const T* ec) {
  if (op1 != kSmoothZero && op1 != kProlongSmooth) return hipErrorInvalidValue;
  if (op1 == kProlongSmooth && (!ec || !x || (a0.sh.nx & 1) || (a0.sh.ny & 1)))
    return hipErrorInvalidValue;
  if (!b || !y || (!a0.xi && !a0.cellw)) return hipErrorInvalidValue;
  if ((size_t)(a0.sh.nx + 1) * (a0.sh.ny + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  ElastOpArgs<T> a = a0;
  a.sh = SweepShape::make(a0.sh.nx, a0.sh.ny, std::max(1, rows), kWave - 2 * 2);
  const int nb = a.sh.blocks();
  double* dp = nullptr;
  if (dot_by) {
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    dp = dot_by->partials;
    dot_by->nblocks = nb;
  }
  const bool xi = a.xi != nullptr;
#define TMX_LAUNCH(XIV, OPV)                                                              \
  TMX_LAUNCH_OR_RECORD(kBlock, (elast_sweep_pair_body<T, XIV, OPV>), nb, 0, s,              \
                       (k_elast_sweep_pair<T, XIV, OPV>), a, x, b, y, om1, om2, dp, ec)
  if (op1 == kSmoothZero) {
    if (xi) TMX_LAUNCH(true, kSmoothZero);
    else TMX_LAUNCH(false, kSmoothZero);
  } else {
    if (xi) TMX_LAUNCH(true, kProlongSmooth);
    else TMX_LAUNCH(false, kProlongSmooth);
  }
#undef TMX_LAUNCH
  return hipGetLastError();
}
template hipError_t elast_sweep_pair<double>(const ElastOpArgs<double>&, int, int, const double*,
                                             const double*, double*, double, double, RedSlot*,
                                             hipStream_t, const double*);
template hipError_t elast_sweep_pair<float>(const ElastOpArgs<float>&, int, int, const float*,
                                            const float*, float*, float, float, RedSlot*,
                                            hipStream_t, const float*);

// The zero-start pair with the residual and its nested restriction fused behind it: three
// stages, each trailing the one before by a cell row.  At cell row cy the first sweep completes
// its node row cy (from zero: the diagonal blocks only), the second sweep takes cell row cy - 1
// and completes node row cy - 1 (written out: the pre-smoothed iterate), and the residual stage
// takes cell row cy - 2 of the second sweep's iterate and completes the residual of node row
// cy - 2; once the residual rows 2J - 1 .. 2J + 1 exist, the lanes of even columns write
// coarse node row J.  The lanes carry four halo columns on each side (56 output columns); a
// chunk of R (even) node rows sweeps R + 6 cell rows.
template <typename T, bool XI>
__device__ __forceinline__ void elast_sweep_pair_rr_body(const BlkId blk, const ElastOpArgs<T>& a,
                                                         const T* b, T* y, T* yc, T om1, T om2) {
  using V = V2<T>;
  constexpr int H = 4;
  const SweepShape sh = a.sh;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blk.id);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1;
  const int nxc = nx >> 1, nyc = ny >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const unsigned fixed = a.fixed;
  const bool colfix = ((fixed & kLeft) && c == 0) || ((fixed & kRight) && c == nx);
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);  // owned node rows [r0, rend)
  const int cy_begin = r0 - 4;
  const int cy_end = wave_ok ? rend + 2 : cy_begin;
  const int first_row = max(0, r0 - 3);  // first node row of the first sweep with a full diagonal
  const auto B = glob(reinterpret_cast<const V*>(b));
  const auto Y = glob(reinterpret_cast<V*>(y));
  const auto YC = glob(reinterpret_cast<V*>(yc));
  const auto CW = glob(reinterpret_cast<const V*>(a.cellw));
  const auto XIP = glob(a.xi);
  const ElastConst<T> k = a.k;
  const V zero = {T(0), T(0)};

  auto row_fixed = [&](int iy) {
    return colfix || ((fixed & kBottom) && iy == 0) || ((fixed & kTop) && iy == ny);
  };
  auto ld_xi = [&](int iy) -> T {
    if (XI && colv && iy >= 0 && iy <= ny) return XIP[(unsigned)(iy * nvx + c)];
    return T(0);
  };

  int cy = cy_begin;
  T zb = ld_xi(cy);
  T zbr = lane_down(zb);
  Diag3<T> d_b = {T(0), T(0), T(0)};  // first sweep: diagonal block of node row cy so far
  // second sweep: cell row cy - 1 between node rows cy - 1 (v_b) and cy of the first sweep
  V vb = zero, vb_raw = zero, vbr = zero;
  V acc2 = zero;
  V bp = zero;  // b at node row cy - 1
  Diag3<T> dp = {T(0), T(0), T(0)};
  T invp = T(0);
  T wAp = T(0), wBp = T(0);
  // residual stage: cell row cy - 2 between node rows cy - 2 (s_b) and cy - 1 of the second sweep
  V sb = zero, sb_raw = zero, sbr = zero;
  V acc3 = zero;
  V bpp = zero;  // b at node row cy - 2
  T wApp = T(0), wBpp = T(0);
  V rA = zero, rB = zero;  // residual rows cy - 4, cy - 3

  // The loads of cell row cy (xi of node row cy + 1 or the cell weights, b of node row cy) are
  // issued kPrefetchRows rows before the row is swept, from clamped in-grid addresses (validity
  // is applied where the values are used), so a wave waits out the memory latency once per
  // chunk rather than once per row.  The row loop is unrolled by kPrefetchRows so that each
  // ring slot keeps its registers (a rotating ring would force a wait at every move).
  const int cn = min(max(c, 0), nx), cc = min(max(c, 0), nx - 1);
  auto fetch = [&](int row) {
    RowIn<T> r;
    const int nt = min(max(row + 1, 0), ny), nb = min(max(row, 0), ny);
    const int cr = min(max(row, 0), ny - 1);
    if constexpr (XI) r.z = XIP[(unsigned)(nt * nvx + cn)];
    else r.w = CW[(unsigned)(cr * nx + cc)];
    r.b = B[(unsigned)(nb * nvx + cn)];
    return r;
  };
  RowIn<T> pf[kPrefetchRows];
  if (wave_ok) {
#pragma unroll
    for (int q = 0; q < kPrefetchRows; ++q) pf[q] = fetch(cy_begin + q);
  }
  for (int cy0 = cy_begin; cy0 < cy_end; cy0 += kPrefetchRows) {
#pragma unroll
   for (int q = 0; q < kPrefetchRows; ++q) {
    cy = cy0 + q;
    if (cy >= cy_end) break;
    const RowIn<T> in = pf[q];
    // ---- first sweep, cell row cy
    const T zt = (XI && colv && cy + 1 >= 0 && cy + 1 <= ny) ? in.z : T(0);
    const T ztr = lane_down(zt);
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    T wA = T(0), wB = T(0);
    if (cell_ok) {
      if (XI) {
        wA = pc_weight(a.m + a.cp * hsym3(zb, zbr, ztr, a.p), a.wmin);
        wB = pc_weight(a.m + a.cp * hsym3(zb, zt, ztr, a.p), a.wmin);
      } else {
        wA = in.w.x;
        wB = in.w.y;
      }
    }
    const T wAl = lane_up(wA), wBl = lane_up(wB);
    d_b.xx += wA * k.Pxx + wB * k.Qxx + wAl * k.Rxx;
    d_b.xy += wAl * k.Rxy;
    d_b.yy += wA * k.Pyy + wB * k.Qyy + wAl * k.Ryy;
    Diag3<T> d_t;
    d_t.xx = wB * k.Rxx + wAl * k.Qxx + wBl * k.Pxx;
    d_t.xy = wB * k.Rxy;
    d_t.yy = wB * k.Ryy + wAl * k.Qyy + wBl * k.Pyy;
    const bool rowv = colv && cy >= first_row && cy <= ny;
    const bool fx = row_fixed(cy);
    V bb = zero, v_raw = zero;
    T inv = T(0);
    if (rowv) {
      bb = in.b;
      V d;
      if (fx) {
        d = bb;
      } else {
        inv = recip(d_b.xx * d_b.yy - d_b.xy * d_b.xy);
        d = {(d_b.yy * bb.x - d_b.xy * bb.y) * inv, (d_b.xx * bb.y - d_b.xy * bb.x) * inv};
      }
      v_raw = {om1 * d.x, om1 * d.y};
    }
    const V vt = fx ? zero : v_raw;
    const V vtr = lane_down(vt);
    // ---- second sweep, cell row cy - 1: node row cy - 1 complete
    V y2_raw = zero;
    const int ry = cy - 1;
    const bool fx2 = row_fixed(ry);
    {
      V f0, f1, f2, f3;
      cell_forces(k, wAp, wBp, vb, vbr, vt, vtr, f0, f1, f2, f3);
      acc2 += f0;
      acc2 += lane_up(f1);
      if (colv && ry >= 0 && ry <= ny) {
        const V Ax = fx2 ? vb_raw : acc2;
        const V res = bp - Ax;
        V d;
        if (fx2) {
          d = res;
        } else {
          d = {(dp.yy * res.x - dp.xy * res.y) * invp, (dp.xx * res.y - dp.xy * res.x) * invp};
        }
        y2_raw = {vb_raw.x + om2 * d.x, vb_raw.y + om2 * d.y};
        if (ry >= r0 && ry < rend && outl) Y[(unsigned)(ry * nvx + c)] = y2_raw;
      }
      acc2 = f2;
      acc2 += lane_up(f3);
    }
    const V y2 = fx2 ? zero : y2_raw;
    const V y2r = lane_down(y2);
    // ---- residual stage, cell row cy - 2: residual of node row cy - 2 complete
    V rC = zero;
    const int rk = cy - 2;
    {
      V g0, g1, g2, g3;
      cell_forces(k, wApp, wBpp, sb, sbr, y2, y2r, g0, g1, g2, g3);
      acc3 += g0;
      acc3 += lane_up(g1);
      if (colv && rk >= 0 && rk <= ny) {
        const V Ax = row_fixed(rk) ? sb_raw : acc3;
        rC = bpp - Ax;
      }
      acc3 = g2;
      acc3 += lane_up(g3);
    }
    // coarse row J: centre fine row 2J = rk - 1 (rk odd), or 2J = ny at the top (ny even)
    const bool odd_emit = (rk & 1) && rk - 1 >= r0 && rk - 1 < rend;
    const bool top_emit = rk == ny && ny >= r0 && ny < rend;
    if (odd_emit || top_emit) {
      const V lo = odd_emit ? rA : rB, mid = odd_emit ? rB : rC;
      const V hi = odd_emit ? rC : zero;
      const V mr = lane_down(mid), ml = lane_up(mid), hr = lane_down(hi), ll = lane_up(lo);
      const int J = odd_emit ? (rk - 1) >> 1 : ny >> 1;
      if (outl && !(c & 1)) {
        const int I = c >> 1;
        V o = zero;
        if (!node_fixed(fixed, I, J, nxc, nyc)) {
          o.x = mid.x + T(0.5) * (mr.x + ml.x + hi.x + lo.x + hr.x + ll.x);
          o.y = mid.y + T(0.5) * (mr.y + ml.y + hi.y + lo.y + hr.y + ll.y);
        }
        YC[(size_t)J * (nxc + 1) + I] = o;
      }
    }
    rA = rB;
    rB = rC;
    // ---- shifts
    sb = y2;
    sb_raw = y2_raw;
    sbr = y2r;
    bpp = bp;
    wApp = wAp;
    wBpp = wBp;
    vb = vt;
    vb_raw = v_raw;
    vbr = vtr;
    bp = bb;
    dp = d_b;
    invp = inv;
    wAp = wA;
    wBp = wB;
    d_b = d_t;
    zb = zt;
    zbr = ztr;
    pf[q] = fetch(cy + kPrefetchRows);
   }
  }
}

template <typename T, bool XI>
__global__ __launch_bounds__(kBlock) void k_elast_sweep_pair_rr(ElastOpArgs<T> a,
                                                                const T* __restrict__ b,
                                                                T* __restrict__ y,
                                                                T* __restrict__ yc, T om1,
                                                                T om2) {
  elast_sweep_pair_rr_body<T, XI>(hw_blk(), a, b, y, yc, om1, om2);
}

template <typename T>
hipError_t elast_sweep_pair_rr(const ElastOpArgs<T>& a0, int rows, const T* b, T* y, T* yc,
                               T om1, T om2, hipStream_t s) {
  if (!b || !y || !yc || (!a0.xi && !a0.cellw)) return hipErrorInvalidValue;
  if ((a0.sh.nx & 1) || (a0.sh.ny & 1)) return hipErrorInvalidValue;
  if ((size_t)(a0.sh.nx + 1) * (a0.sh.ny + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  ElastOpArgs<T> a = a0;
  a.sh = SweepShape::make(a0.sh.nx, a0.sh.ny, std::max(2, rows & ~1), kWave - 2 * 4);
  const int nb = a.sh.blocks();
  if (a.xi)
    TMX_LAUNCH_OR_RECORD(kBlock, (elast_sweep_pair_rr_body<T, true>), nb, 0, s,
                         (k_elast_sweep_pair_rr<T, true>), a, b, y, yc, om1, om2);
  else
    TMX_LAUNCH_OR_RECORD(kBlock, (elast_sweep_pair_rr_body<T, false>), nb, 0, s,
                         (k_elast_sweep_pair_rr<T, false>), a, b, y, yc, om1, om2);
  return hipGetLastError();
}
template hipError_t elast_sweep_pair_rr<double>(const ElastOpArgs<double>&, int, const double*,
                                                double*, double*, double, double, hipStream_t);
template hipError_t elast_sweep_pair_rr<float>(const ElastOpArgs<float>&, int, const float*,
                                               float*, float*, float, float, hipStream_t);

template <typename T>
__device__ __forceinline__ void elast_cell_weights_body(const BlkId blk, int nx,
                                                        int ny,
                                                        const T* xi,
                                                        int p,
                                                        T m,
                                                        T cp,
                                                        T wmin,
                                                        V2<T>* cw, V2<float>* cw32) {
  const size_t i = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int cx = (int)(i % nx), cy = (int)(i / nx);
  const size_t v0 = (size_t)cy * (nx + 1) + cx;
  const T a0 = xi[v0], a1 = xi[v0 + 1], a2 = xi[v0 + nx + 1], a3 = xi[v0 + nx + 2];
  const V2<T> w = {pc_weight(m + cp * hsym3(a0, a1, a3, p), wmin),
                   pc_weight(m + cp * hsym3(a0, a2, a3, p), wmin)};
  cw[i] = w;
  if (cw32) cw32[i] = {(float)w.x, (float)w.y};
}

template <typename T>
__global__ void k_elast_cell_weights(int nx, int ny, const T* __restrict__ xi, int p, T m, T cp,
                                     T wmin, V2<T>* __restrict__ cw, V2<float>* __restrict__ cw32) {
  elast_cell_weights_body<T>(hw_blk(), nx, ny, xi, p, m, cp, wmin, cw, cw32);
}

template <typename T>
hipError_t elast_cell_weights(int nx, int ny, const T* xi, int p, T m, T cp, T wmin, T* cellw,
                              hipStream_t s, float* cellw32) {
  TMX_LAUNCH_OR_RECORD(kBlock, (elast_cell_weights_body<T>), grid1d((size_t)nx * ny), 0, s, (k_elast_cell_weights<T>), nx, ny, xi, p, m, cp, wmin, reinterpret_cast<V2<T>*>(cellw), reinterpret_cast<V2<float>*>(cellw32));
  return hipGetLastError();
}
template hipError_t elast_cell_weights<double>(int, int, const double*, int, double, double,
                                               double, double*, hipStream_t, float*);
template hipError_t elast_cell_weights<float>(int, int, const float*, int, float, float, float,
                                              float*, hipStream_t, float*);

// coarse triangle A (lower right) = fine (0,0).A + (1,0).A + (1,0).B + (1,1).A
// coarse triangle B (upper left)  = fine (0,0).B + (0,1).A + (0,1).B + (1,1).B
template <typename T>
__device__ __forceinline__ void restrict_weights_body(const BlkId blk, int nxc,
                                                      int nyc,
                                                      const V2<T>* f,
                                                      V2<T>* cw, V2<float>* cw32) {
  const size_t i = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (i >= (size_t)nxc * nyc) return;
  const int cx = (int)(i % nxc), cy = (int)(i / nxc);
  const int nxf = 2 * nxc;
  const V2<T> f00 = f[(size_t)(2 * cy) * nxf + 2 * cx];
  const V2<T> f10 = f[(size_t)(2 * cy) * nxf + 2 * cx + 1];
  const V2<T> f01 = f[(size_t)(2 * cy + 1) * nxf + 2 * cx];
  const V2<T> f11 = f[(size_t)(2 * cy + 1) * nxf + 2 * cx + 1];
  const V2<T> w = {T(0.25) * (f00.x + f10.x + f10.y + f11.x),
                   T(0.25) * (f00.y + f01.x + f01.y + f11.y)};
  cw[i] = w;
  if (cw32) cw32[i] = {(float)w.x, (float)w.y};
}

template <typename T>
__global__ void k_restrict_weights(int nxc, int nyc, const V2<T>* __restrict__ f,
                                   V2<T>* __restrict__ cw, V2<float>* __restrict__ cw32) {
  restrict_weights_body<T>(hw_blk(), nxc, nyc, f, cw, cw32);
}

template <typename T>
hipError_t elast_restrict_weights(int nxc, int nyc, const T* fine, T* coarse, hipStream_t s,
                                  float* coarse32) {
  TMX_LAUNCH_OR_RECORD(kBlock, (restrict_weights_body<T>), grid1d((size_t)nxc * nyc), 0, s, (k_restrict_weights<T>), nxc, nyc, reinterpret_cast<const V2<T>*>(fine), reinterpret_cast<V2<T>*>(coarse), reinterpret_cast<V2<float>*>(coarse32));
  return hipGetLastError();
}
template hipError_t elast_restrict_weights<double>(int, int, const double*, double*, hipStream_t,
                                                   float*);
template hipError_t elast_restrict_weights<float>(int, int, const float*, float*, hipStream_t,
                                                  float*);

// ---------------------------------------------------------------- explicit stencils
// local vertex offsets (dx, dy) and gradient factors (gx in 1/hx, gy in 1/hy) of the two
// triangle types: A = (v0, v1, v3), B = (v0, v2, v3)
// (compile-time tables: with the loops below unrolled, every stencil index is a constant and
// the accumulators stay in registers)
__device__ constexpr int c_tri_off[2][3][2] = {{{0, 0}, {1, 0}, {1, 1}}, {{0, 0}, {0, 1}, {1, 1}}};
__device__ constexpr double c_tri_grad[2][3][2] = {{{-1, 0}, {1, -1}, {0, 1}},
                                                   {{0, -1}, {-1, 1}, {1, 0}}};

__device__ inline int st_index(int dx, int dy) { return (dy + 1) * 3 + (dx + 1); }

template <typename T>
__device__ __forceinline__ void elast_to_stencil_body(const BlkId blk, const ElastOpArgs<T>& a,
                                                      T* st,
                                                      double hx,
                                                      double hy,
                                                      double lam,
                                                      double mu) {
  const int nx = a.sh.nx, ny = a.sh.ny;
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t v = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int ix = (int)(v % (nx + 1)), iy = (int)(v / (nx + 1));
  double acc[9][4];
#pragma unroll
  for (int kk = 0; kk < 9; ++kk)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[kk][e] = 0.0;
  const bool vf = node_fixed(a.fixed, ix, iy, nx, ny);
  if (vf) {
    acc[4][0] = acc[4][3] = 1.0;
  } else {
    const double area = 0.5 * hx * hy;
#pragma unroll
    for (int cyo = -1; cyo <= 0; ++cyo)
#pragma unroll
      for (int cxo = -1; cxo <= 0; ++cxo) {
        const int cx = ix + cxo, cy = iy + cyo;
        if (cx < 0 || cy < 0 || cx >= nx || cy >= ny) continue;
        double w[2];
        if (a.xi) {
          const size_t v0 = (size_t)cy * (nx + 1) + cx;
          const double x0 = a.xi[v0], x1 = a.xi[v0 + 1], x2 = a.xi[v0 + nx + 1],
                       x3 = a.xi[v0 + nx + 2];
          w[0] = pc_weight((double)a.m + (double)a.cp * hsym3(x0, x1, x3, a.p), (double)a.wmin);
          w[1] = pc_weight((double)a.m + (double)a.cp * hsym3(x0, x2, x3, a.p), (double)a.wmin);
        } else {
          w[0] = a.cellw[2 * ((size_t)cy * nx + cx)];
          w[1] = a.cellw[2 * ((size_t)cy * nx + cx) + 1];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          int li = -1;  // (cx - ix = cxo, cy - iy = cyo: a compile-time search)
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (cxo + c_tri_off[t][q][0] == 0 && cyo + c_tri_off[t][q][1] == 0) li = q;
          if (li < 0) continue;
          const double gix = c_tri_grad[t][li][0] / hx, giy = c_tri_grad[t][li][1] / hy;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const int jx = cx + c_tri_off[t][q][0], jy = cy + c_tri_off[t][q][1];
            if (node_fixed(a.fixed, jx, jy, nx, ny)) continue;
            const double gjx = c_tri_grad[t][q][0] / hx, gjy = c_tri_grad[t][q][1] / hy;
            const double gi[2] = {gix, giy}, gj[2] = {gjx, gjy};
            const double dotg = gix * gjx + giy * gjy;
            const int kk = st_index(cxo + c_tri_off[t][q][0], cyo + c_tri_off[t][q][1]);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
              for (int r = 0; r < 2; ++r) {
                double val = lam * gi[p] * gj[r] + mu * gi[r] * gj[p];
                if (p == r) val += mu * dotg;
                acc[kk][p * 2 + r] += w[t] * area * val;
              }
          }
        }
      }
  }
#pragma unroll
  for (int kk = 0; kk < 9; ++kk)
#pragma unroll
    for (int e = 0; e < 4; ++e) st[(size_t)(kk * 4 + e) * nn + v] = (T)acc[kk][e];
}

template <typename T>
__global__ void k_elast_to_stencil(ElastOpArgs<T> a, T* __restrict__ st,
                                   double hx, double hy, double lam, double mu) {
  elast_to_stencil_body<T>(hw_blk(), a, st, hx, hy, lam, mu);
}

// lam, mu and a = hy/hx are recovered from the force constants (lam_h = l/2, mu_h = m/2,
// mu_a = m a/2); the element matrices depend on hx, hy only through a.
template <>
hipError_t elast_to_stencil<double>(const ElastOpArgs<double>& a, double* st, hipStream_t s) {
  // the constants are scale free (a = hy/hx); use hx = 1, hy = a.
  const double aa = a.k.mu_a / a.k.mu_h;  // = hy/hx
  const double mu = 2 * a.k.mu_h, lam = 2 * a.k.lam_h;
  const size_t nn = (size_t)(a.sh.nx + 1) * (a.sh.ny + 1);
  TMX_LAUNCH_OR_RECORD(kBlock, (elast_to_stencil_body<double>), grid1d(nn), 0, s, (k_elast_to_stencil<double>), a, st, 1.0, aa, lam, mu);
  return hipGetLastError();
}

// ---------------------------------------------------------------- sensitivity
// b_v = -p(1-m) sum_{T∋v} e_T int_T xi^q lambda_v,  q = p-1,
// int_T xi^q lambda_v = 2|T| q!/(q+3)! H_q(xi_a, xi_b, xi_c, xi_v)  (complete homogeneous)
constexpr int kMaxQ = 8;

__device__ __forceinline__ void hsym3_all(double a, double b, double c, int q, double (&H)[kMaxQ + 1]) {
  double pa = 1.0, hab = 1.0, h = 1.0;
  H[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= kMaxQ; ++k) {
    if (k <= q) {
      pa *= a;
      hab = pa + b * hab;
      h = hab + c * h;
    }
    H[k] = h;
  }
}

__device__ __forceinline__ double hsym4_from3(const double (&H)[kMaxQ + 1], int q, double d) {
  double acc = 1.0;  // H3_0
#pragma unroll
  for (int j = 1; j <= kMaxQ; ++j)
    if (j <= q) acc = acc * d + H[j];
  return acc;
}

struct SensArgs {
  SweepShape sh;
  int q;
  double coef;  // -p(1-m) * 2|T| q!/(q+3)!
  double lam, mu, hx, hy;
};

__device__ __forceinline__ void elast_sensitivity_body(const BlkId blk, const SensArgs& a,
                                                       const double* xi,
                                                       const double* u,
                                                       double* out) {
  using V = V2<double>;
  const SweepShape sh = a.sh;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blk.id);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1;
  const int c = strip * sh.stripw - 1 + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= 1 && lane <= sh.stripw;
  const int r0 = chunk * sh.rows;
  const int r1 = wave_ok ? min(r0 + sh.rows, ny + 1) : r0 - 1;
  const V* U = reinterpret_cast<const V*>(u);
  auto ld_u = [&](int iy) -> V {
    if (colv && iy >= 0 && iy <= ny) return U[(size_t)iy * nvx + c];
    return V{0.0, 0.0};
  };
  auto ld_xi = [&](int iy) -> double {
    if (colv && iy >= 0 && iy <= ny) return xi[(size_t)iy * nvx + c];
    return 0.0;
  };
  const double lam = a.lam, mu = a.mu, ihx = 1.0 / a.hx, ihy = 1.0 / a.hy;
  auto energy = [&](double exx, double eyy, double gam) {
    const double tr = exx + eyy;
    return lam * tr * tr + 2 * mu * (exx * exx + eyy * eyy + 0.5 * gam * gam);
  };
  int cy = r0 - 1;
  V ub = ld_u(cy), ubr = lane_down(ub);
  double zb = ld_xi(cy), zbr = lane_down(zb);
  double acc_b = 0.0;
  for (; cy < r1; ++cy) {
    const V ut = ld_u(cy + 1), utr = lane_down(ut);
    const double zt = ld_xi(cy + 1), ztr = lane_down(zt);
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
    if (cell_ok) {
      double H[kMaxQ + 1];
      // triangle A = (v0 = ub, v1 = ubr, v3 = utr)
      const double eA = energy((ubr.x - ub.x) * ihx, (utr.y - ubr.y) * ihy,
                               (utr.x - ubr.x) * ihy + (ubr.y - ub.y) * ihx);
      hsym3_all(zb, zbr, ztr, a.q, H);
      f0 += eA * hsym4_from3(H, a.q, zb);
      f1 += eA * hsym4_from3(H, a.q, zbr);
      f3 += eA * hsym4_from3(H, a.q, ztr);
      // triangle B = (v0 = ub, v2 = ut, v3 = utr)
      const double eB = energy((utr.x - ut.x) * ihx, (ut.y - ub.y) * ihy,
                               (ut.x - ub.x) * ihy + (utr.y - ut.y) * ihx);
      hsym3_all(zb, zt, ztr, a.q, H);
      f0 += eB * hsym4_from3(H, a.q, zb);
      f2 += eB * hsym4_from3(H, a.q, zt);
      f3 += eB * hsym4_from3(H, a.q, ztr);
    }
    acc_b += f0 + lane_up(f1);
    const double acc_t = f2 + lane_up(f3);
    if (cy >= r0 && outl) out[(size_t)cy * nvx + c] = a.coef * acc_b;
    acc_b = acc_t;
    ub = ut;
    ubr = utr;
    zb = zt;
    zbr = ztr;
  }
}

__global__ __launch_bounds__(kBlock) void k_elast_sensitivity(SensArgs a, const double* __restrict__ xi,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ out) {
  elast_sensitivity_body(hw_blk(), a, xi, u, out);
}

hipError_t elast_sensitivity(const SweepShape& sh, const double* xi, const double* u, int p,
                             double m, double lam, double mu, double hx, double hy, double* out,
                             hipStream_t s) {
  SensArgs a;
  a.sh = sh;
  a.q = p - 1;
  double fq = 1.0;
  for (int i = 2; i <= a.q; ++i) fq *= i;
  double f3 = 1.0;
  for (int i = 2; i <= a.q + 3; ++i) f3 *= i;
  a.coef = -p * (1.0 - m) * (hx * hy) * fq / f3;
  a.lam = lam;
  a.mu = mu;
  a.hx = hx;
  a.hy = hy;
  TMX_LAUNCH_OR_RECORD(kBlock, (elast_sensitivity_body), sh.blocks(), 0, s, (k_elast_sensitivity), a, xi, u, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- load vector
// Body force / traction rules of the reference evaluated at the P2 interpolation points
// (vertices and edge midpoints 0.5 xa + 0.5 xb) with IEEE semantics (no contraction), then
// integrated exactly against the P1 test functions:
//   int_T N_vertex(j) l_i = |T|/30 (j = i), -|T|/60 (j != i)
//   int_T N_edge(jk) l_i  = 2|T|/15 (i in {j,k}), |T|/15 (otherwise)
//   int_e N_a phi_a = L/6, int_e N_b phi_a = 0, int_e N_mid phi_a = L/3 (edge e = (a,b))
__global__ void k_assemble_load(LoadDesc d, double* __restrict__ F) {
#pragma clang fp contract(off)
  const int nx = d.nx, ny = d.ny;
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int ix = (int)(v % (nx + 1)), iy = (int)(v / (nx + 1));
  double sx = 0.0, sy = 0.0;
  if (d.has_body) {
    for (int cyo = -1; cyo <= 0; ++cyo)
      for (int cxo = -1; cxo <= 0; ++cxo) {
        const int cx = ix + cxo, cy = iy + cyo;
        if (cx < 0 || cy < 0 || cx >= nx || cy >= ny) continue;
        for (int t = 0; t < 2; ++t) {
          int li = -1;
          int px[3], py[3];
          for (int q = 0; q < 3; ++q) {
            px[q] = cx + c_tri_off[t][q][0];
            py[q] = cy + c_tri_off[t][q][1];
            if (px[q] == ix && py[q] == iy) li = q;
          }
          if (li < 0) continue;
          const double area = 0.5 * fabs((vx(d, px[1]) - vx(d, px[0])) * (vy(d, py[2]) - vy(d, py[0])) -
                                         (vx(d, px[2]) - vx(d, px[0])) * (vy(d, py[1]) - vy(d, py[0])));
          double ax = 0.0, ay = 0.0;
          for (int q = 0; q < 3; ++q) {
            double fx, fy;
            body_value(d, vx(d, px[q]), vy(d, py[q]), fx, fy);
            const double wv = (q == li) ? area / 30.0 : -area / 60.0;
            ax += wv * fx;
            ay += wv * fy;
          }
          for (int q = 0; q < 3; ++q) {  // edge opposite vertex q
            const int e0 = (q + 1) % 3, e1 = (q + 2) % 3;
            double fx, fy;
            body_value(d, mid(vx(d, px[e0]), vx(d, px[e1])), mid(vy(d, py[e0]), vy(d, py[e1])), fx, fy);
            const double we = (q == li) ? area / 15.0 : 2.0 * area / 15.0;
            ax += we * fx;
            ay += we * fy;
          }
          sx += ax;
          sy += ay;
        }
      }
  }
  if (d.ntr > 0) {
    // boundary edges containing v: (dir, other node) along each side v lies on
    const double x0 = vx(d, ix), y0 = vy(d, iy);
    double tvx, tvy;
    traction_value(d, x0, y0, tvx, tvy);
    auto edge = [&](int jx, int jy) {
      const double x1 = vx(d, jx), y1 = vy(d, jy);
      double tmx, tmy;
      traction_value(d, mid(x0, x1), mid(y0, y1), tmx, tmy);
      const double len = (jx != ix) ? fabs(x1 - x0) : fabs(y1 - y0);
      sx += len * (tvx / 6.0 + tmx / 3.0);
      sy += len * (tvy / 6.0 + tmy / 3.0);
    };
    if (iy == 0 || iy == ny) {
      if (ix > 0) edge(ix - 1, iy);
      if (ix < nx) edge(ix + 1, iy);
    }
    if (ix == 0 || ix == nx) {
      if (iy > 0) edge(ix, iy - 1);
      if (iy < ny) edge(ix, iy + 1);
    }
  }
  F[2 * v] = sx;
  F[2 * v + 1] = sy;
}

hipError_t elast_assemble_load(const LoadDesc& d, double* F, hipStream_t s) {
  const size_t nn = (size_t)(d.nx + 1) * (d.ny + 1);
  k_assemble_load<<<grid1d(nn), kBlock, 0, s>>>(d, F);
  return hipGetLastError();
}

// ============================================================================ filter
template <typename T, int OP>
__device__ __forceinline__ void filter_sweep_body(const BlkId blk, const SweepShape& sh,
                                                  const FilterConst<T>& k, const T* x_,
                                                  const T* b_, T* y_, T omega, double* dpart,
                                                  const SweepExtra<T>& ex) {
  constexpr bool RR = (OP == kResidualRestrict);
  constexpr bool PS = (OP == kProlongSmooth);
  constexpr bool AU = (OP == kApplyUpdate);
  const auto x = glob(x_);
  const auto b = glob(b_);
  const auto y = glob(y_);
  const auto ec = glob(ex.ec);
  const auto zlow = glob(ex.zlow);
  const auto pnew = glob(ex.pnew);
  const auto xk = glob(ex.xk);
  const T beta = AU ? (T)*glob(ex.beta) : T(0);
  const T alpha = (AU && ex.xk) ? (T)*glob(ex.alpha) : T(0);
  constexpr bool SMOOTH = (OP == kSmooth || OP == kSmoothZero || PS);
  constexpr int H = RR ? 2 : 1;
  const bool want_dot = dpart != nullptr;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blk.id);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1;
  const int nxc = nx >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);
  const int cy_begin = RR ? r0 - 2 : r0 - 1;
  const int cy_end = !wave_ok ? cy_begin : (RR ? min(rend, ny) + 1 : rend);
  T po_next = T(0);  // AU: raw p_old of the row ld last loaded (for the deferred x)
  auto ld = [&](int iy) -> T {
    T v = T(0);
    const bool ok = colv && iy >= 0 && iy <= ny;
    if (OP != kSmoothZero && ok) v = x[(size_t)iy * nvx + c];
    if constexpr (AU) {
      po_next = v;
      if (ok) v = (ex.zlow ? (T)zlow[(size_t)iy * nvx + c] : b[(size_t)iy * nvx + c]) + beta * v;
    }
    if constexpr (PS) {
      if (ok) {
        const int I0 = c >> 1, J0 = iy >> 1, ox = c & 1, oy = iy & 1;
        T e = ec[(size_t)J0 * (nxc + 1) + I0];
        if (ox | oy) e = T(0.5) * (e + ec[(size_t)(J0 + oy) * (nxc + 1) + I0 + ox]);
        v += e;
      }
    }
    return v;
  };
  const T ka = k.ka, kb = k.kb, mA = k.mA;
  const T dfull = ka + kb + T(4) * mA, dhalf = ka + kb + T(2) * mA;
  int cy = cy_begin;
  T ub = ld(cy), ubr = lane_down(ub);
  T po_b = po_next;
  T acc_b = T(0), d_b = T(0);
  T rA = T(0), rB = T(0);
  double dot = 0.0;
  for (; cy < cy_end; ++cy) {
    const T ut = ld(cy + 1), utr = lane_down(ut);
    const T po_t = po_next;
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    const T w = cell_ok ? T(1) : T(0);
    T acc_t = T(0);
    if constexpr (OP != kSmoothZero) {
      const T u0 = ub, u1 = ubr, u2 = ut, u3 = utr;
      const T SA = u0 + u1 + u3, SB = u0 + u2 + u3;
      const T f0 = w * (ka * (u0 - u1) + mA * (u0 + SA) + kb * (u0 - u2) + mA * (u0 + SB));
      const T f1 = w * (ka * (u1 - u0) + kb * (u1 - u3) + mA * (u1 + SA));
      const T f2 = w * (kb * (u2 - u0) + ka * (u2 - u3) + mA * (u2 + SB));
      const T f3 = w * (kb * (u3 - u1) + mA * (u3 + SA) + ka * (u3 - u2) + mA * (u3 + SB));
      acc_b += f0 + lane_up(f1);
      acc_t = f2 + lane_up(f3);
    }
    T d_t = T(0);
    if constexpr (SMOOTH) {
      const T wl = lane_up(w);
      d_b += w * dfull + wl * dhalf;
      d_t = w * dhalf + wl * dfull;
    }
    const bool rowv = colv && cy >= 0 && cy <= ny;
    const size_t idx = (size_t)cy * nvx + c;
    if constexpr (RR) {
      T rC = T(0);
      if (rowv) rC = b[idx] - acc_b;
      const bool odd_emit = (cy & 1) && cy - 1 >= r0 && cy - 1 < rend;
      const bool top_emit = cy == ny && ny >= r0 && ny < rend;
      if (odd_emit || top_emit) {
        const T lo = odd_emit ? rA : rB, mid = odd_emit ? rB : rC, hi = odd_emit ? rC : T(0);
        const T mr = lane_down(mid), ml = lane_up(mid), hr = lane_down(hi), ll = lane_up(lo);
        const int J = odd_emit ? (cy - 1) >> 1 : ny >> 1;
        if (outl && !(c & 1))
          y[(size_t)J * (nxc + 1) + (c >> 1)] = mid + T(0.5) * (mr + ml + hi + lo + hr + ll);
      }
      rA = rB;
      rB = rC;
    } else if (cy >= r0 && outl) {
      if constexpr (OP == kApply || AU) {
        y[idx] = acc_b;
        if constexpr (AU) {
          pnew[idx] = ub;
          if (ex.xk) xk[idx] += alpha * po_b;  // deferred x update (see SweepExtra)
        }
        if (want_dot) dot += (double)ub * (double)acc_b;
      } else if constexpr (OP == kResidual) {
        y[idx] = b[idx] - acc_b;
      } else {
        const T bb = b[idx];
        const T yo = ub + omega * (bb - acc_b) / d_b;
        y[idx] = yo;
        if (want_dot) dot += (double)bb * (double)yo;
      }
    }
    acc_b = acc_t;
    d_b = d_t;
    ub = ut;
    ubr = utr;
    po_b = po_t;
  }
  if constexpr (!RR) {
    if (want_dot) {
      const double v[1] = {dot};
      double* const o[1] = {dpart};
      block_partials<1>(v, o, blk.id);
    }
  }
}

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_filter_sweep(SweepShape sh, FilterConst<T> k,
                                                         const T* __restrict__ x,
                                                         const T* __restrict__ b,
                                                         T* __restrict__ y, T omega,
                                                         double* __restrict__ dpart,
                                                         SweepExtra<T> ex) {
  filter_sweep_body<T, OP>(hw_blk(), sh, k, x, b, y, omega, dpart, ex);
}

template <typename T>
hipError_t filter_sweep(const SweepShape& sh0, const FilterConst<T>& k, int op, const T* x,
                        const T* b, T* y, T omega, RedSlot* dot_xy, hipStream_t s,
                        const SweepExtra<T>& ex) {
  const T* ec = ex.ec;
  SweepShape sh = sh0;
  if (op == kResidualRestrict) {
    if ((sh.nx & 1) || (sh.ny & 1)) return hipErrorInvalidValue;
    sh = SweepShape::make(sh.nx, sh.ny, std::max(2, sh.rows & ~1), kStripCols - 2);
  }
  if (op == kProlongSmooth && (!ec || (sh.nx & 1) || (sh.ny & 1))) return hipErrorInvalidValue;
  if (op == kApplyUpdate && (!ex.pnew || !ex.beta || (!b && !ex.zlow))) return hipErrorInvalidValue;
  if (op == kResidualRestrict && dot_xy) return hipErrorInvalidValue;
  const int nb = sh.blocks();
  double* dp = nullptr;
  if (dot_xy) {
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    dp = dot_xy->partials;
    dot_xy->nblocks = nb;
  }
#define TMX_CASE(OPV) \
  case OPV: TMX_LAUNCH_OR_RECORD(kBlock, (filter_sweep_body<T, OPV>), nb, 0, s, (k_filter_sweep<T, OPV>), sh, k, x, b, y, omega, dp, ex); break;
  switch (op) {
    TMX_CASE(kApply)
    TMX_CASE(kResidual)
    TMX_CASE(kSmooth)
    TMX_CASE(kSmoothZero)
    TMX_CASE(kResidualRestrict)
    TMX_CASE(kProlongSmooth)
    TMX_CASE(kApplyUpdate)
    default: return hipErrorInvalidValue;
  }
#undef TMX_CASE
  return hipGetLastError();
}
template hipError_t filter_sweep<double>(const SweepShape&, const FilterConst<double>&, int,
                                         const double*, const double*, double*, double,
                                         RedSlot*, hipStream_t, const SweepExtra<double>&);
template hipError_t filter_sweep<float>(const SweepShape&, const FilterConst<float>&, int,
                                        const float*, const float*, float*, float, RedSlot*,
                                        hipStream_t, const SweepExtra<float>&);

// ---------------------------------------------------------------------------- filter pipelines
// The filter's finest level as two column pipelines (no LDS): NU block-Jacobi sweeps stacked
// in one pass over the grid, sweep s trailing sweep s - 1 by one cell row so that the
// intermediate iterates stay in registers.  Each stage loses one lane of its neighbours on
// each side, hence the halo columns: NU + 2 for the down pass (NU sweeps from zero, the
// residual, the nested restriction), NU for the up pass (x + P e, NU sweeps, b.y).  The
// right-hand side, the inverse diagonal and the cell mask of the rows a stage still needs
// ride along in short register histories.  The filter operator has natural boundary
// conditions only, so no node is fixed.
template <typename T>
__device__ __forceinline__ void filter_cell(T ka, T kb, T mA, T w, T u0, T u1, T u2, T u3, T& f0,
                                            T& f1, T& f2, T& f3) {
  const T SA = u0 + u1 + u3, SB = u0 + u2 + u3;
  f0 = w * (ka * (u0 - u1) + mA * (u0 + SA) + kb * (u0 - u2) + mA * (u0 + SB));
  f1 = w * (ka * (u1 - u0) + kb * (u1 - u3) + mA * (u1 + SA));
  f2 = w * (kb * (u2 - u0) + ka * (u2 - u3) + mA * (u2 + SB));
  f3 = w * (kb * (u3 - u1) + mA * (u3 + SA) + ka * (u3 - u2) + mA * (u3 + SB));
}

template <typename T, int NU>
__global__ __launch_bounds__(kBlock) void k_filter_pipe_down(SweepShape sh, FilterConst<T> k,
                                                             PipeOmega<T> om,
                                                             const T* __restrict__ b,
                                                             T* __restrict__ y,
                                                             T* __restrict__ yc) {
  constexpr int H = NU + 2;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blockIdx.x);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1, nxc = nx >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);  // owned node rows [r0, rend)
  const int first_row = max(0, r0 - NU - 1);    // first sweep: first row with a full diagonal
  const int cy_begin = r0 - NU - 2;
  const int cy_end = wave_ok ? rend + NU : cy_begin;
  const T ka = k.ka, kb = k.kb, mA = k.mA;
  const T dfull = ka + kb + T(4) * mA, dhalf = ka + kb + T(2) * mA;
  T bh[NU + 1], dh[NU + 1], wh[NU + 2];  // rows / cell rows cy, cy - 1, ...
  T prev[NU + 2], prevr[NU + 2], acc[NU + 2];
#pragma unroll
  for (int i = 0; i < NU + 2; ++i) {
    if (i < NU + 1) bh[i] = dh[i] = T(0);
    wh[i] = prev[i] = prevr[i] = acc[i] = T(0);
  }
  T d_b = T(0), rA = T(0), rB = T(0);
  for (int cy = cy_begin; cy < cy_end; ++cy) {
#pragma unroll
    for (int i = NU + 1; i > 0; --i) {
      if (i <= NU) {
        bh[i] = bh[i - 1];
        dh[i] = dh[i - 1];
      }
      wh[i] = wh[i - 1];
    }
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    const T w = cell_ok ? T(1) : T(0);
    wh[0] = w;
    const T wl = lane_up(w);
    d_b += w * dfull + wl * dhalf;
    const T d_t = w * dhalf + wl * dfull;
    const bool rowv = colv && cy >= first_row && cy <= ny;
    T bb = T(0), di = T(0);
    if (rowv) {
      bb = b[(size_t)cy * nvx + c];
      di = T(1) / d_b;
    }
    bh[0] = bb;
    dh[0] = di;
    T cur = om.w[0] * bb * di;  // first sweep from zero, node row cy
    if (NU == 1 && cy >= r0 && cy < rend && outl) y[(size_t)cy * nvx + c] = cur;
    d_b = d_t;
#pragma unroll
    for (int s = 2; s <= NU + 1; ++s) {  // sweeps 2..NU, then the residual
      const int row = cy - s + 1;
      const T curr = lane_down(cur);
      T f0, f1, f2, f3;
      filter_cell(ka, kb, mA, wh[s - 1], prev[s], prevr[s], cur, curr, f0, f1, f2, f3);
      acc[s] += f0;
      acc[s] += lane_up(f1);
      const bool ok = colv && row >= 0 && row <= ny;
      T out = T(0);
      if (s <= NU) {
        if (ok) out = prev[s] + om.w[s - 1] * (bh[s - 1] - acc[s]) * dh[s - 1];
        if (s == NU && row >= r0 && row < rend && outl) y[(size_t)row * nvx + c] = out;
      } else if (ok) {
        out = bh[s - 1] - acc[s];
      }
      acc[s] = f2;
      acc[s] += lane_up(f3);
      prev[s] = cur;
      prevr[s] = curr;
      cur = out;
    }
    // cur: residual of node row rk; coarse row J from fine rows 2J - 1 .. 2J + 1
    const int rk = cy - NU;
    const bool odd_emit = (rk & 1) && rk - 1 >= r0 && rk - 1 < rend;
    const bool top_emit = rk == ny && ny >= r0 && ny < rend;
    if (odd_emit || top_emit) {
      const T lo = odd_emit ? rA : rB, mid = odd_emit ? rB : cur, hi = odd_emit ? cur : T(0);
      const T mr = lane_down(mid), ml = lane_up(mid), hr = lane_down(hi), ll = lane_up(lo);
      const int J = odd_emit ? (rk - 1) >> 1 : ny >> 1;
      if (outl && !(c & 1))
        yc[(size_t)J * (nxc + 1) + (c >> 1)] = mid + T(0.5) * (mr + ml + hi + lo + hr + ll);
    }
    rA = rB;
    rB = cur;
  }
}

template <typename T, int NU>
__global__ __launch_bounds__(kBlock) void k_filter_pipe_up(SweepShape sh, FilterConst<T> k,
                                                           PipeOmega<T> om,
                                                           const T* __restrict__ x,
                                                           const T* __restrict__ ec,
                                                           const T* __restrict__ b,
                                                           T* __restrict__ y,
                                                           double* __restrict__ dpart) {
  constexpr int H = NU;
  const bool want_dot = dpart != nullptr;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = wave_index(blockIdx.x);
  const int strip = wv % sh.nstrips, chunk = wv / sh.nstrips;
  const bool wave_ok = chunk < sh.nchunks;
  const int nx = sh.nx, ny = sh.ny, nvx = nx + 1, nxc = nx >> 1;
  const int c = strip * sh.stripw - H + lane;
  const bool colv = wave_ok && c >= 0 && c <= nx;
  const bool cellc = wave_ok && c >= 0 && c < nx;
  const bool outl = colv && lane >= H && lane < kWave - H;
  const int r0 = chunk * sh.rows;
  const int rend = min(r0 + sh.rows, ny + 1);
  const int first_row = max(0, r0 - NU + 1);
  const int cy_begin = r0 - NU;
  const int cy_end = wave_ok ? rend + NU - 1 : cy_begin;
  const T ka = k.ka, kb = k.kb, mA = k.mA;
  const T dfull = ka + kb + T(4) * mA, dhalf = ka + kb + T(2) * mA;
  auto ldx = [&](int iy) -> T {  // x + P e
    T v = T(0);
    if (colv && iy >= 0 && iy <= ny) {
      const int I0 = c >> 1, J0 = iy >> 1, ox = c & 1, oy = iy & 1;
      T e = ec[(size_t)J0 * (nxc + 1) + I0];
      if (ox | oy) e = T(0.5) * (e + ec[(size_t)(J0 + oy) * (nxc + 1) + I0 + ox]);
      v = x[(size_t)iy * nvx + c] + e;
    }
    return v;
  };
  T bh[NU], dh[NU], wh[NU + 1];
  T prev[NU + 1], prevr[NU + 1], acc[NU + 1];
#pragma unroll
  for (int i = 0; i < NU + 1; ++i) {
    if (i < NU) bh[i] = dh[i] = T(0);
    wh[i] = prev[i] = prevr[i] = acc[i] = T(0);
  }
  T ub = ldx(cy_begin);
  T ubr = lane_down(ub);
  T d_b = T(0), acc1 = T(0);
  double dot = 0.0;
  for (int cy = cy_begin; cy < cy_end; ++cy) {
#pragma unroll
    for (int i = NU; i > 0; --i) {
      if (i < NU) {
        bh[i] = bh[i - 1];
        dh[i] = dh[i - 1];
      }
      wh[i] = wh[i - 1];
    }
    const T ut = ldx(cy + 1);
    const T utr = lane_down(ut);
    const bool cell_ok = cellc && cy >= 0 && cy < ny;
    const T w = cell_ok ? T(1) : T(0);
    wh[0] = w;
    T f0, f1, f2, f3;
    filter_cell(ka, kb, mA, w, ub, ubr, ut, utr, f0, f1, f2, f3);
    acc1 += f0;
    acc1 += lane_up(f1);
    const T wl = lane_up(w);
    d_b += w * dfull + wl * dhalf;
    const T d_t = w * dhalf + wl * dfull;
    const bool rowv = colv && cy >= first_row && cy <= ny;
    T bb = T(0), di = T(0), cur = T(0);
    if (rowv) {
      bb = b[(size_t)cy * nvx + c];
      di = T(1) / d_b;
      cur = ub + om.w[0] * (bb - acc1) * di;
    }
    bh[0] = bb;
    dh[0] = di;
    if (NU == 1 && cy >= r0 && cy < rend && outl) {
      y[(size_t)cy * nvx + c] = cur;
      if (want_dot) dot += (double)bb * (double)cur;
    }
    acc1 = f2;
    acc1 += lane_up(f3);
    d_b = d_t;
    ub = ut;
    ubr = utr;
#pragma unroll
    for (int s = 2; s <= NU; ++s) {
      const int row = cy - s + 1;
      const T curr = lane_down(cur);
      T g0, g1, g2, g3;
      filter_cell(ka, kb, mA, wh[s - 1], prev[s], prevr[s], cur, curr, g0, g1, g2, g3);
      acc[s] += g0;
      acc[s] += lane_up(g1);
      T out = T(0);
      if (colv && row >= 0 && row <= ny)
        out = prev[s] + om.w[s - 1] * (bh[s - 1] - acc[s]) * dh[s - 1];
      if (s == NU && row >= r0 && row < rend && outl) {
        y[(size_t)row * nvx + c] = out;
        if (want_dot) dot += (double)bh[s - 1] * (double)out;
      }
      acc[s] = g2;
      acc[s] += lane_up(g3);
      prev[s] = cur;
      prevr[s] = curr;
      cur = out;
    }
  }
  if (want_dot) {
    const double v[1] = {dot};
    double* const o[1] = {dpart};
    block_partials<1>(v, o);
  }
}

template <typename T>
hipError_t filter_pipe_down(int nx, int ny, const FilterConst<T>& k, int nu, const T* omega,
                            int rows, const T* b, T* y, T* yc, hipStream_t s) {
  if ((nx & 1) || (ny & 1) || nu < 1 || nu > kMaxPipeNu || !b || !y || !yc)
    return hipErrorInvalidValue;
  if ((size_t)(nx + 1) * (ny + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  PipeOmega<T> om{};
  for (int i = 0; i < nu; ++i) om.w[i] = omega[i];
  const SweepShape sh = SweepShape::make(nx, ny, std::max(2, rows & ~1), kWave - 2 * (nu + 2));
  const int nb = sh.blocks();
  switch (nu) {
    case 1: k_filter_pipe_down<T, 1><<<nb, kBlock, 0, s>>>(sh, k, om, b, y, yc); break;
    case 2: k_filter_pipe_down<T, 2><<<nb, kBlock, 0, s>>>(sh, k, om, b, y, yc); break;
    case 3: k_filter_pipe_down<T, 3><<<nb, kBlock, 0, s>>>(sh, k, om, b, y, yc); break;
    default: k_filter_pipe_down<T, 4><<<nb, kBlock, 0, s>>>(sh, k, om, b, y, yc); break;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t filter_pipe_up(int nx, int ny, const FilterConst<T>& k, int nu, const T* omega,
                          int rows, const T* x, const T* ec, const T* b, T* y, RedSlot* dot,
                          hipStream_t s) {
  if ((nx & 1) || (ny & 1) || nu < 1 || nu > kMaxPipeNu || !x || !ec || !b || !y)
    return hipErrorInvalidValue;
  if ((size_t)(nx + 1) * (ny + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  PipeOmega<T> om{};
  for (int i = 0; i < nu; ++i) om.w[i] = omega[i];
  const SweepShape sh = SweepShape::make(nx, ny, std::max(1, rows), kWave - 2 * nu);
  const int nb = sh.blocks();
  double* dp = nullptr;
  if (dot) {
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    dp = dot->partials;
    dot->nblocks = nb;
  }
  switch (nu) {
    case 1: k_filter_pipe_up<T, 1><<<nb, kBlock, 0, s>>>(sh, k, om, x, ec, b, y, dp); break;
    case 2: k_filter_pipe_up<T, 2><<<nb, kBlock, 0, s>>>(sh, k, om, x, ec, b, y, dp); break;
    case 3: k_filter_pipe_up<T, 3><<<nb, kBlock, 0, s>>>(sh, k, om, x, ec, b, y, dp); break;
    default: k_filter_pipe_up<T, 4><<<nb, kBlock, 0, s>>>(sh, k, om, x, ec, b, y, dp); break;
  }
  return hipGetLastError();
}
template hipError_t filter_pipe_down<double>(int, int, const FilterConst<double>&, int,
                                             const double*, int, const double*, double*, double*,
                                             hipStream_t);
template hipError_t filter_pipe_down<float>(int, int, const FilterConst<float>&, int,
                                            const float*, int, const float*, float*, float*,
                                            hipStream_t);
template hipError_t filter_pipe_up<double>(int, int, const FilterConst<double>&, int,
                                           const double*, int, const double*, const double*,
                                           const double*, double*, RedSlot*, hipStream_t);
template hipError_t filter_pipe_up<float>(int, int, const FilterConst<float>&, int, const float*,
                                          int, const float*, const float*, const float*, float*,
                                          RedSlot*, hipStream_t);

template <typename T>
__global__ void k_filter_to_stencil(int nx, int ny, FilterConst<T> k, T* __restrict__ st) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int ix = (int)(v % (nx + 1)), iy = (int)(v / (nx + 1));
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // element matrices in units: stiffness eps^2 |T| gi.gj with gx in 1/hx, gy in 1/hy gives
  // ka (x-x products) and kb (y-y products); mass |T|/12 (1 + delta_ij)
  for (int cyo = -1; cyo <= 0; ++cyo)
    for (int cxo = -1; cxo <= 0; ++cxo) {
      const int cx = ix + cxo, cy = iy + cyo;
      if (cx < 0 || cy < 0 || cx >= nx || cy >= ny) continue;
      for (int t = 0; t < 2; ++t) {
        int li = -1;
        for (int q = 0; q < 3; ++q)
          if (cx + c_tri_off[t][q][0] == ix && cy + c_tri_off[t][q][1] == iy) li = q;
        if (li < 0) continue;
        for (int q = 0; q < 3; ++q) {
          const int jx = cx + c_tri_off[t][q][0], jy = cy + c_tri_off[t][q][1];
          const double kx = c_tri_grad[t][li][0] * c_tri_grad[t][q][0];
          const double ky = c_tri_grad[t][li][1] * c_tri_grad[t][q][1];
          // eps^2 |T| (gix gjx/hx^2 + giy gjy/hy^2) = ka kx + kb ky
          double val = (double)k.ka * kx + (double)k.kb * ky;
          val += (double)k.mA * (q == li ? 2.0 : 1.0);
          acc[st_index(jx - ix, jy - iy)] += val;
        }
      }
    }
  for (int kk = 0; kk < 9; ++kk) st[(size_t)kk * nn + v] = (T)acc[kk];
}

template <typename T>
hipError_t filter_to_stencil(int nx, int ny, const FilterConst<T>& k, T* st, hipStream_t s) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  k_filter_to_stencil<T><<<grid1d(nn), kBlock, 0, s>>>(nx, ny, k, st);
  return hipGetLastError();
}
template hipError_t filter_to_stencil<double>(int, int, const FilterConst<double>&, double*,
                                              hipStream_t);

// ============================================================================ transfers
template <typename T, int BS>
struct Blk {
  T v[BS];
};

template <typename T, int BS>
__device__ __forceinline__ Blk<T, BS> ldb(const T* p, size_t node) {
  Blk<T, BS> r;
#pragma unroll
  for (int i = 0; i < BS; ++i) r.v[i] = p[BS * node + i];
  return r;
}

template <typename T, int BS>
__device__ __forceinline__ void restrict_nested_body(const BlkId blk, int nxc,
                                                     int nyc,
                                                     unsigned fixed,
                                                     const T* rf,
                                                     T* bc) {
  const size_t nn = (size_t)(nxc + 1) * (nyc + 1);
  const size_t i = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int I = (int)(i % (nxc + 1)), J = (int)(i / (nxc + 1));
  const int nxf = 2 * nxc, nyf = 2 * nyc, nvf = nxf + 1;
  T acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = T(0);
  if (!node_fixed(fixed, I, J, nxc, nyc)) {
    const int fx = 2 * I, fy = 2 * J;
    const int off[7][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}};
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      const int gx = fx + off[t][0], gy = fy + off[t][1];
      if (gx < 0 || gy < 0 || gx > nxf || gy > nyf) continue;
      const T w = (t == 0) ? T(1) : T(0.5);
      const size_t g = (size_t)gy * nvf + gx;
#pragma unroll
      for (int q = 0; q < BS; ++q) acc[q] += w * rf[BS * g + q];
    }
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) bc[BS * i + q] = acc[q];
}

template <typename T, int BS>
__global__ void k_restrict_nested(int nxc, int nyc, unsigned fixed, const T* __restrict__ rf,
                                  T* __restrict__ bc) {
  restrict_nested_body<T, BS>(hw_blk(), nxc, nyc, fixed, rf, bc);
}

template <typename T, int BS>
__device__ __forceinline__ void prolong_nested_body(const BlkId blk, int nxc,
                                                    int nyc,
                                                    unsigned fixed,
                                                    const T* ec,
                                                    T* xf) {
  const int nxf = 2 * nxc, nyf = 2 * nyc;
  const size_t nn = (size_t)(nxf + 1) * (nyf + 1);
  const size_t i = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int fx = (int)(i % (nxf + 1)), fy = (int)(i / (nxf + 1));
  if (node_fixed(fixed, fx, fy, nxf, nyf)) return;
  const int nvc = nxc + 1;
  const int ox = fx & 1, oy = fy & 1;
  const int I0 = fx >> 1, J0 = fy >> 1;
  T add[BS];
  if (!ox && !oy) {
#pragma unroll
    for (int q = 0; q < BS; ++q) add[q] = ec[BS * ((size_t)J0 * nvc + I0) + q];
  } else {
    const int I1 = I0 + ox, J1 = J0 + oy;  // (odd,even): x-neighbour; (even,odd): y; (odd,odd): diagonal
#pragma unroll
    for (int q = 0; q < BS; ++q)
      add[q] = T(0.5) * (ec[BS * ((size_t)J0 * nvc + I0) + q] + ec[BS * ((size_t)J1 * nvc + I1) + q]);
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) xf[BS * i + q] += add[q];
}

template <typename T, int BS>
__global__ void k_prolong_nested(int nxc, int nyc, unsigned fixed, const T* __restrict__ ec,
                                 T* __restrict__ xf) {
  prolong_nested_body<T, BS>(hw_blk(), nxc, nyc, fixed, ec, xf);
}

template <typename T, int BS>
hipError_t restrict_nested(int nxc, int nyc, unsigned fixed, const T* rf, T* bc, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (restrict_nested_body<T, BS>), grid1d((size_t)(nxc + 1) * (nyc + 1)), 0, s, (k_restrict_nested<T, BS>), nxc, nyc, fixed, rf, bc);
  return hipGetLastError();
}
template <typename T, int BS>
hipError_t prolong_nested(int nxc, int nyc, unsigned fixed, const T* ec, T* xf, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (prolong_nested_body<T, BS>), grid1d((size_t)(2 * nxc + 1) * (2 * nyc + 1)), 0, s, (k_prolong_nested<T, BS>), nxc, nyc, fixed, ec, xf);
  return hipGetLastError();
}

template <typename T, int BS>
__device__ __forceinline__ void restrict_bilinear_node(int nxf, int nyf, int nxc, int nyc,
                                                       unsigned fixed, const T* __restrict__ rf,
                                                       T* __restrict__ bc, size_t i);
template <typename T, int BS>
__device__ __forceinline__ void prolong_bilinear_node(int nxf, int nyf, int nxc, int nyc,
                                                      unsigned fixed, const T* __restrict__ ec,
                                                      T* __restrict__ xf, size_t i);

template <typename T, int BS>
__device__ __forceinline__ void restrict_bilinear_body(const BlkId blk, int nxf,
                                                       int nyf,
                                                       int nxc,
                                                       int nyc,
                                                       unsigned fixed,
                                                       const T* rf,
                                                       T* bc) {
  const unsigned nn = (unsigned)(nxc + 1) * (unsigned)(nyc + 1);
  const unsigned i = blk.id * blockDim.x + threadIdx.x;
  if (i < nn) restrict_bilinear_node<T, BS>(nxf, nyf, nxc, nyc, fixed, rf, bc, i);
}

template <typename T, int BS>
__global__ void k_restrict_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                                    const T* __restrict__ rf, T* __restrict__ bc) {
  restrict_bilinear_body<T, BS>(hw_blk(), nxf, nyf, nxc, nyc, fixed, rf, bc);
}

template <typename T, int BS>
__device__ __forceinline__ void prolong_bilinear_body(const BlkId blk, int nxf,
                                                      int nyf,
                                                      int nxc,
                                                      int nyc,
                                                      unsigned fixed,
                                                      const T* ec,
                                                      T* xf) {
  const unsigned nn = (unsigned)(nxf + 1) * (unsigned)(nyf + 1);
  const unsigned i = blk.id * blockDim.x + threadIdx.x;
  if (i < nn) prolong_bilinear_node<T, BS>(nxf, nyf, nxc, nyc, fixed, ec, xf, i);
}

template <typename T, int BS>
__global__ void k_prolong_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                                   const T* __restrict__ ec, T* __restrict__ xf) {
  prolong_bilinear_body<T, BS>(hw_blk(), nxf, nyf, nxc, nyc, fixed, ec, xf);
}

template <typename T, int BS>
hipError_t restrict_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* rf,
                             T* bc, hipStream_t s) {
  if ((size_t)(nxf + 1) * (nyf + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  TMX_LAUNCH_OR_RECORD(kBlock, (restrict_bilinear_body<T, BS>), grid1d((size_t)(nxc + 1) * (nyc + 1)), 0, s, (k_restrict_bilinear<T, BS>), nxf, nyf, nxc, nyc, fixed, rf, bc);
  return hipGetLastError();
}
template <typename T, int BS>
hipError_t prolong_bilinear(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* ec,
                            T* xf, hipStream_t s) {
  if ((size_t)(nxf + 1) * (nyf + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  TMX_LAUNCH_OR_RECORD(kBlock, (prolong_bilinear_body<T, BS>), grid1d((size_t)(nxf + 1) * (nyf + 1)), 0, s, (k_prolong_bilinear<T, BS>), nxf, nyf, nxc, nyc, fixed, ec, xf);
  return hipGetLastError();
}

#define TMX_INST_TRANSFER(T, BS)                                                                   \
  template hipError_t restrict_nested<T, BS>(int, int, unsigned, const T*, T*, hipStream_t);      \
  template hipError_t prolong_nested<T, BS>(int, int, unsigned, const T*, T*, hipStream_t);       \
  template hipError_t restrict_bilinear<T, BS>(int, int, int, int, unsigned, const T*, T*,        \
                                               hipStream_t);                                       \
  template hipError_t prolong_bilinear<T, BS>(int, int, int, int, unsigned, const T*, T*,         \
                                              hipStream_t);
TMX_INST_TRANSFER(double, 1)
TMX_INST_TRANSFER(double, 2)
TMX_INST_TRANSFER(float, 1)
TMX_INST_TRANSFER(float, 2)

// ============================================================================ 9-point levels
// st layout (SoA): st[(k*BS*BS + a*BS + b) * nn + node], k = (dy+1)*3 + (dx+1).  Boundary
// nodes carry zero coefficients for missing neighbours, so neighbour indices are clamped
// and every load is issued unconditionally (no divergent branches around loads).
template <typename T, int BS, int OP>
__device__ __forceinline__ void s9_node(int nx, int ny, const T* __restrict__ st,
                                        const T* __restrict__ x, const T* __restrict__ b,
                                        T* __restrict__ y, T omega, size_t v) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const int ix = (int)(v % (nx + 1)), iy = (int)(v / (nx + 1));
  T acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = T(0);
  if (OP != kSmoothZero) {
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const int jx = min(max(ix + kk % 3 - 1, 0), nx), jy = min(max(iy + kk / 3 - 1, 0), ny);
      const size_t j = (size_t)jy * (nx + 1) + jx;
#pragma unroll
      for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc[a] += st[(size_t)(kk * BS * BS + a * BS + c) * nn + v] * x[BS * j + c];
    }
  }
  if (OP == kApply) {
#pragma unroll
    for (int q = 0; q < BS; ++q) y[BS * v + q] = acc[q];
    return;
  }
  T res[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) res[q] = b[BS * v + q] - acc[q];
  if (OP == kResidual) {
#pragma unroll
    for (int q = 0; q < BS; ++q) y[BS * v + q] = res[q];
    return;
  }
  T d[BS];
  if constexpr (BS == 1) {
    d[0] = res[0] / st[(size_t)4 * nn + v];
  } else {
    const T a00 = st[(size_t)(16 + 0) * nn + v], a01 = st[(size_t)(16 + 1) * nn + v];
    const T a10 = st[(size_t)(16 + 2) * nn + v], a11 = st[(size_t)(16 + 3) * nn + v];
    const T inv = T(1) / (a00 * a11 - a01 * a10);
    d[0] = (a11 * res[0] - a01 * res[1]) * inv;
    d[1] = (a00 * res[1] - a10 * res[0]) * inv;
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) {
    const T xo = (OP == kSmoothZero) ? T(0) : x[BS * v + q];
    y[BS * v + q] = xo + omega * d[q];
  }
}

template <typename T, int BS, int OP>
__device__ __forceinline__ void stencil_op_body(const BlkId blk, int nx,
                                                int ny,
                                                const T* st,
                                                const T* x,
                                                const T* b,
                                                T* y,
                                                T omega) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t v = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (v < nn) s9_node<T, BS, OP>(nx, ny, st, x, b, y, omega, v);
}

template <typename T, int BS, int OP>
__global__ void k_stencil_op(int nx, int ny, const T* __restrict__ st, const T* __restrict__ x,
                             const T* __restrict__ b, T* __restrict__ y, T omega) {
  stencil_op_body<T, BS, OP>(hw_blk(), nx, ny, st, x, b, y, omega);
}

template <typename T, int BS>
hipError_t stencil_op(int nx, int ny, const T* st, int op, const T* x, const T* b, T* y, T omega,
                      hipStream_t s) {
  const unsigned g = grid1d((size_t)(nx + 1) * (ny + 1), 64);
  switch (op) {
    case kApply: TMX_LAUNCH_OR_RECORD(64, (stencil_op_body<T, BS, kApply>), g, 0, s, (k_stencil_op<T, BS, kApply>), nx, ny, st, x, b, y, omega); break;
    case kResidual: TMX_LAUNCH_OR_RECORD(64, (stencil_op_body<T, BS, kResidual>), g, 0, s, (k_stencil_op<T, BS, kResidual>), nx, ny, st, x, b, y, omega); break;
    case kSmooth: TMX_LAUNCH_OR_RECORD(64, (stencil_op_body<T, BS, kSmooth>), g, 0, s, (k_stencil_op<T, BS, kSmooth>), nx, ny, st, x, b, y, omega); break;
    default: TMX_LAUNCH_OR_RECORD(64, (stencil_op_body<T, BS, kSmoothZero>), g, 0, s, (k_stencil_op<T, BS, kSmoothZero>), nx, ny, st, x, b, y, omega); break;
  }
  return hipGetLastError();
}
template hipError_t stencil_op<double, 1>(int, int, const double*, int, const double*,
                                          const double*, double*, double, hipStream_t);
template hipError_t stencil_op<double, 2>(int, int, const double*, int, const double*,
                                          const double*, double*, double, hipStream_t);
template hipError_t stencil_op<float, 1>(int, int, const float*, int, const float*, const float*,
                                         float*, float, hipStream_t);
template hipError_t stencil_op<float, 2>(int, int, const float*, int, const float*, const float*,
                                         float*, float, hipStream_t);

// per-node bilinear transfers (bodies of k_restrict_bilinear / k_prolong_bilinear and of the
// tail kernel).  Branch-free: every load is issued before any is used (out-of-range and
// non-contributing positions get zero weight at a clamped address), so a launch costs one
// memory round trip rather than one per contributing node.
template <typename T, int BS>
__device__ __forceinline__ void restrict_bilinear_node(int nxf, int nyf, int nxc, int nyc,
                                                       unsigned fixed, const T* __restrict__ rf,
                                                       T* __restrict__ bc, size_t i) {
  const int I = (int)((unsigned)i % (unsigned)(nxc + 1)), J = (int)((unsigned)i / (unsigned)(nxc + 1));
  double acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = 0.0;
  const int px = cpos(I, nxf, nxc), py = cpos(J, nyf, nyc);
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy) {
    const int fy = py + dy;
    const double wy = (fy < 0 || fy > nyf) ? 0.0 : interp1_weight(fy, nyf, nyc, J);
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      const int fx = px + dx;
      const double w = (fx < 0 || fx > nxf) ? 0.0 : wy * interp1_weight(fx, nxf, nxc, I);
      const size_t g = (size_t)min(max(fy, 0), nyf) * (nxf + 1) + min(max(fx, 0), nxf);
#pragma unroll
      for (int q = 0; q < BS; ++q) acc[q] += w * (double)rf[BS * g + q];
    }
  }
  const bool fx0 = node_fixed(fixed, I, J, nxc, nyc);
#pragma unroll
  for (int q = 0; q < BS; ++q) bc[BS * i + q] = fx0 ? T(0) : (T)acc[q];
}

template <typename T, int BS>
__device__ __forceinline__ void prolong_bilinear_node(int nxf, int nyf, int nxc, int nyc,
                                                      unsigned fixed, const T* __restrict__ ec,
                                                      T* __restrict__ xf, size_t i) {
  const int fx = (int)((unsigned)i % (unsigned)(nxf + 1)), fy = (int)((unsigned)i / (unsigned)(nxf + 1));
  if (node_fixed(fixed, fx, fy, nxf, nyf)) return;
  const Interp1 ix = interp1(fx, nxf, nxc), iy = interp1(fy, nyf, nyc);
  const int cx0 = ix.c[0], cx1 = ix.n > 1 ? ix.c[1] : ix.c[0];
  const int cy0 = iy.c[0], cy1 = iy.n > 1 ? iy.c[1] : iy.c[0];
  const double wx0 = ix.w[0], wx1 = ix.n > 1 ? ix.w[1] : 0.0;
  const double wy0 = iy.w[0], wy1 = iy.n > 1 ? iy.w[1] : 0.0;
  const size_t r0 = (size_t)cy0 * (nxc + 1), r1 = (size_t)cy1 * (nxc + 1);
  T add[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q)
    add[q] = (T)(wx0 * wy0) * ec[BS * (r0 + cx0) + q] + (T)(wx1 * wy0) * ec[BS * (r0 + cx1) + q] +
             (T)(wx0 * wy1) * ec[BS * (r1 + cx0) + q] + (T)(wx1 * wy1) * ec[BS * (r1 + cx1) + q];
#pragma unroll
  for (int q = 0; q < BS; ++q) xf[BS * i + q] += add[q];
}

// fused: coarse node i receives P^T (b - A x) of the fine 9-point level
template <typename T, int BS>
__device__ __forceinline__ void s9_resrestrict_node(int nxf, int nyf, int nxc, int nyc,
                                                    unsigned fixed, const T* __restrict__ st,
                                                    const T* __restrict__ x,
                                                    const T* __restrict__ b, T* __restrict__ bc,
                                                    size_t i) {
  const size_t nnf = (size_t)(nxf + 1) * (nyf + 1);
  const int I = (int)(i % (nxc + 1)), J = (int)(i / (nxc + 1));
  double acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = 0.0;
  if (!node_fixed(fixed, I, J, nxc, nyc)) {
    const int px = cpos(I, nxf, nxc), py = cpos(J, nyf, nyc);
    for (int dy = -1; dy <= 1; ++dy) {
      const int fy = py + dy;
      if (fy < 0 || fy > nyf) continue;
      const double wy = interp1_weight(fy, nyf, nyc, J);
      for (int dx = -1; dx <= 1; ++dx) {
        const int fx = px + dx;
        if (fx < 0 || fx > nxf) continue;
        const double w = wy * interp1_weight(fx, nxf, nxc, I);
        if (w == 0.0) continue;
        const size_t f = (size_t)fy * (nxf + 1) + fx;
        T r[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) r[q] = b[BS * f + q];
#pragma unroll
        for (int kk = 0; kk < 9; ++kk) {
          const int jx = min(max(fx + kk % 3 - 1, 0), nxf), jy = min(max(fy + kk / 3 - 1, 0), nyf);
          const size_t j = (size_t)jy * (nxf + 1) + jx;
#pragma unroll
          for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int c = 0; c < BS; ++c)
              r[a] -= st[(size_t)(kk * BS * BS + a * BS + c) * nnf + f] * x[BS * j + c];
        }
#pragma unroll
        for (int q = 0; q < BS; ++q) acc[q] += w * (double)r[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) bc[BS * i + q] = (T)acc[q];
}

// x + P e (bilinear) at fine node (fx, fy); fixed nodes are not corrected
template <typename T, int BS>
__device__ __forceinline__ void corrected(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                                          const T* __restrict__ x, const T* __restrict__ ec,
                                          int fx, int fy, T (&out)[BS]) {
  const size_t f = (size_t)fy * (nxf + 1) + fx;
#pragma unroll
  for (int q = 0; q < BS; ++q) out[q] = x[BS * f + q];
  if (node_fixed(fixed, fx, fy, nxf, nyf)) return;
  const Interp1 ix = interp1(fx, nxf, nxc), iy = interp1(fy, nyf, nyc);
  for (int bb = 0; bb < iy.n; ++bb)
    for (int aa = 0; aa < ix.n; ++aa) {
      const T w = (T)(ix.w[aa] * iy.w[bb]);
      const size_t cidx = (size_t)iy.c[bb] * (nxc + 1) + ix.c[aa];
#pragma unroll
      for (int q = 0; q < BS; ++q) out[q] += w * ec[BS * cidx + q];
    }
}

// fused: one Jacobi sweep of the fine 9-point level applied to x + P e
template <typename T, int BS>
__device__ __forceinline__ void s9_prolong_smooth_node(int nxf, int nyf, int nxc, int nyc,
                                                       unsigned fixed, const T* __restrict__ st,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ ec,
                                                       const T* __restrict__ b,
                                                       T* __restrict__ y, T omega, size_t v) {
  const size_t nn = (size_t)(nxf + 1) * (nyf + 1);
  const int ix = (int)(v % (nxf + 1)), iy = (int)(v / (nxf + 1));
  T acc[BS], xv[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = T(0);
#pragma unroll
  for (int kk = 0; kk < 9; ++kk) {
    const int jx = min(max(ix + kk % 3 - 1, 0), nxf), jy = min(max(iy + kk / 3 - 1, 0), nyf);
    T xj[BS];
    corrected<T, BS>(nxf, nyf, nxc, nyc, fixed, x, ec, jx, jy, xj);
    if (kk == 4) {
#pragma unroll
      for (int q = 0; q < BS; ++q) xv[q] = xj[q];
    }
#pragma unroll
    for (int a = 0; a < BS; ++a)
#pragma unroll
      for (int c = 0; c < BS; ++c) acc[a] += st[(size_t)(kk * BS * BS + a * BS + c) * nn + v] * xj[c];
  }
  T res[BS], d[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) res[q] = b[BS * v + q] - acc[q];
  if constexpr (BS == 1) {
    d[0] = res[0] / st[(size_t)4 * nn + v];
  } else {
    const T a00 = st[(size_t)(16 + 0) * nn + v], a01 = st[(size_t)(16 + 1) * nn + v];
    const T a10 = st[(size_t)(16 + 2) * nn + v], a11 = st[(size_t)(16 + 3) * nn + v];
    const T inv = T(1) / (a00 * a11 - a01 * a10);
    d[0] = (a11 * res[0] - a01 * res[1]) * inv;
    d[1] = (a00 * res[1] - a10 * res[0]) * inv;
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) y[BS * v + q] = xv[q] + omega * d[q];
}

template <typename T, int BS>
__global__ void k_stencil_resrestrict(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                                      const T* __restrict__ st, const T* __restrict__ x,
                                      const T* __restrict__ b, T* __restrict__ bc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)(nxc + 1) * (nyc + 1))
    s9_resrestrict_node<T, BS>(nxf, nyf, nxc, nyc, fixed, st, x, b, bc, i);
}

template <typename T, int BS>
__global__ void k_stencil_prolong_smooth(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                                         const T* __restrict__ st, const T* __restrict__ x,
                                         const T* __restrict__ ec, const T* __restrict__ b,
                                         T* __restrict__ y, T omega) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < (size_t)(nxf + 1) * (nyf + 1))
    s9_prolong_smooth_node<T, BS>(nxf, nyf, nxc, nyc, fixed, st, x, ec, b, y, omega, v);
}

template <typename T, int BS>
hipError_t stencil_resrestrict(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* st,
                               const T* x, const T* b, T* bc, hipStream_t s) {
  k_stencil_resrestrict<T, BS><<<grid1d((size_t)(nxc + 1) * (nyc + 1), 64), 64, 0, s>>>(
      nxf, nyf, nxc, nyc, fixed, st, x, b, bc);
  return hipGetLastError();
}
template <typename T, int BS>
hipError_t stencil_prolong_smooth(int nxf, int nyf, int nxc, int nyc, unsigned fixed, const T* st,
                                  const T* x, const T* ec, const T* b, T* y, T omega,
                                  hipStream_t s) {
  k_stencil_prolong_smooth<T, BS><<<grid1d((size_t)(nxf + 1) * (nyf + 1), 64), 64, 0, s>>>(
      nxf, nyf, nxc, nyc, fixed, st, x, ec, b, y, omega);
  return hipGetLastError();
}
#define TMX_INST_S9F(T, BS)                                                                       \
  template hipError_t stencil_resrestrict<T, BS>(int, int, int, int, unsigned, const T*, const T*, \
                                                 const T*, T*, hipStream_t);                       \
  template hipError_t stencil_prolong_smooth<T, BS>(int, int, int, int, unsigned, const T*,        \
                                                    const T*, const T*, const T*, T*, T,           \
                                                    hipStream_t);
TMX_INST_S9F(double, 1)
TMX_INST_S9F(double, 2)
TMX_INST_S9F(float, 1)
TMX_INST_S9F(float, 2)

// ---------------------------------------------------------------- tail V-cycle
// The bottom of the hierarchy (9-point levels of at most kTailThreads nodes, then the dense
// coarsest solve) runs as a single workgroup that steps through the whole V-cycle with a
// barrier between steps: one launch instead of ~7 per level.  Vectors stay in LDS; each
// thread owns one node and holds that node's stencil row in registers while its level is
// being smoothed, so a sweep costs LDS traffic only.

size_t tail_lds_bytes(const size_t* nn, int nlev, int bs, size_t elem) {
  size_t n = 0;
  for (int l = 0; l + 1 < nlev; ++l) n += 3 * (size_t)bs * nn[l];
  n += 2 * (size_t)bs * nn[nlev - 1];
  return n * elem;
}

// prefetched values of the coarsest inverse per thread (8 threads per row)
constexpr int kTailDensePre = 8;

template <typename T, int BS>
struct TailRow {
  T a[9 * BS * BS];  // a[k*BS*BS + e], k = (dy+1)*3 + (dx+1)
};

template <typename T, int BS>
__device__ __forceinline__ void tail_load_row(const TailLevel<T>& L, int v, TailRow<T, BS>& r) {
  const int nn = (L.nx + 1) * (L.ny + 1);
  if (v < nn) {
#pragma unroll
    for (int e = 0; e < 9 * BS * BS; ++e) r.a[e] = L.st[(size_t)e * nn + v];
  }
}

// one 9-point operation at node v of a level held in LDS (x, b, y LDS pointers); nb: the
// node's nine neighbours (clamped into the grid: the stencil entries towards missing
// neighbours are zero), inv: 1 / det of the 2x2 diagonal block (BS == 2)
template <typename T, int BS, int OP>
__device__ __forceinline__ void tail_node(const TailRow<T, BS>& r, const int (&nb)[9], T inv,
                                          const T* x, const T* b, T* y, T omega, int v) {
  T acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = T(0);
  if (OP != kSmoothZero) {
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const int j = nb[kk];
#pragma unroll
      for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc[a] += r.a[kk * BS * BS + a * BS + c] * x[BS * j + c];
    }
  }
  T res[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) res[q] = b[BS * v + q] - acc[q];
  if (OP == kResidual) {
#pragma unroll
    for (int q = 0; q < BS; ++q) y[BS * v + q] = res[q];
    return;
  }
  T d[BS];
  if constexpr (BS == 1) {
    d[0] = res[0] / r.a[4];
  } else {
    const T a00 = r.a[16], a01 = r.a[17], a10 = r.a[18], a11 = r.a[19];
    d[0] = (a11 * res[0] - a01 * res[1]) * inv;
    d[1] = (a00 * res[1] - a10 * res[0]) * inv;
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) {
    const T xo = (OP == kSmoothZero) ? T(0) : x[BS * v + q];
    y[BS * v + q] = xo + omega * d[q];
  }
}

// The same operation with the neighbour indices and the diagonal block's inverse determinant
// formed from the row on the spot (for rows re-read per sweep, where registers are short)
template <typename T, int BS, int OP>
__device__ __forceinline__ void tail_node_direct(const TailLevel<T>& L, const TailRow<T, BS>& r,
                                                 const T* x, const T* b, T* y, T omega, int v) {
  const int nx = L.nx, ny = L.ny;
  if (v >= (nx + 1) * (ny + 1)) return;
  const int ix = v % (nx + 1), iy = v / (nx + 1);
  T acc[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) acc[q] = T(0);
  if (OP != kSmoothZero) {
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const int jx = min(max(ix + kk % 3 - 1, 0), nx), jy = min(max(iy + kk / 3 - 1, 0), ny);
      const int j = jy * (nx + 1) + jx;
#pragma unroll
      for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc[a] += r.a[kk * BS * BS + a * BS + c] * x[BS * j + c];
    }
  }
  T res[BS];
#pragma unroll
  for (int q = 0; q < BS; ++q) res[q] = b[BS * v + q] - acc[q];
  if (OP == kResidual) {
#pragma unroll
    for (int q = 0; q < BS; ++q) y[BS * v + q] = res[q];
    return;
  }
  T d[BS];
  if constexpr (BS == 1) {
    d[0] = res[0] / r.a[4];
  } else {
    const T a00 = r.a[16], a01 = r.a[17], a10 = r.a[18], a11 = r.a[19];
    const T inv = T(1) / (a00 * a11 - a01 * a10);
    d[0] = (a11 * res[0] - a01 * res[1]) * inv;
    d[1] = (a00 * res[1] - a10 * res[0]) * inv;
  }
#pragma unroll
  for (int q = 0; q < BS; ++q) {
    const T xo = (OP == kSmoothZero) ? T(0) : x[BS * v + q];
    y[BS * v + q] = xo + omega * d[q];
  }
}

// A node's stencil row stays in registers across a level's sweeps when it is small enough
// (36 floats, 9 doubles) or the workgroup is at most 256 threads; otherwise (the FP64 2x2-block
// row, 36 doubles, at 1024 threads) it is re-read from global memory (an L2 hit) for each
// sweep, which keeps the kernel within 128 VGPRs without spills.  (A 256-thread tail with the
// 47x16 level moved to the band kernels measured slower at N=1000.)  The node's neighbour
// indices and its diagonal block's inverse determinant are formed once per level.
template <typename T, int BS>
struct TailRowHolder {
  static constexpr bool kKeep = sizeof(T) * 9 * BS * BS <= 144 || kTailThreads <= 256;
  TailRow<T, BS> row;
  int nb[9];  // (kKeep only)
  T inv = T(0);
  bool live = false;
  __device__ __forceinline__ static void neighbours(int ix, int iy, int nx, int ny,
                                                    int (&nb)[9]) {
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const int jx = min(max(ix + kk % 3 - 1, 0), nx), jy = min(max(iy + kk / 3 - 1, 0), ny);
      nb[kk] = jy * (nx + 1) + jx;
    }
  }
  __device__ __forceinline__ void level(const TailLevel<T>& L, int v) {
    if constexpr (kKeep) {
      const int nx = L.nx, ny = L.ny;
      live = v < (nx + 1) * (ny + 1);
      if (!live) return;
      neighbours(v % (nx + 1), v / (nx + 1), nx, ny, nb);
      tail_load_row<T, BS>(L, v, row);
    }
  }
  // the diagonal block's inverse determinant, formed at the level's first sweep (after the
  // phase that the row loads overlap)
  __device__ __forceinline__ void finish() {
    if constexpr (kKeep && BS == 2)
      if (live) inv = diag_inv(row);
  }
  __device__ __forceinline__ static T diag_inv(const TailRow<T, BS>& r) {
    return T(1) / (r.a[16] * r.a[19] - r.a[17] * r.a[18]);
  }
  template <int OP>
  __device__ __forceinline__ void sweep(const TailLevel<T>& L, const T* x, const T* b, T* y,
                                        T omega, int v) {
    if constexpr (kKeep) {
      if (live) tail_node<T, BS, OP>(row, nb, inv, x, b, y, omega, v);
    } else {
      tail_load_row<T, BS>(L, v, row);
      tail_node_direct<T, BS, OP>(L, row, x, b, y, omega, v);
    }
  }
};

template <typename T, int BS>
__device__ __forceinline__ void tail_vcycle_body(const BlkId blk, const TailDesc<T>& d) {
  extern __shared__ __align__(16) unsigned char tail_lds[];
  T* lds = reinterpret_cast<T*>(tail_lds);
  const int tid = threadIdx.x;
  const int nl = d.nlev;  // levels 0..nl-2 are 9-point, level nl-1 is the dense coarsest
  // ping-pong parity: the first sweep writes buffer (S-1)%2 so that after S sweeps the
  // iterate is in buffer 0; after the pre-sweeps it is in buffer nu_post%2 on every level
  const int w_first = (d.nu_pre + d.nu_post - 1) & 1;
  const int w_mid = d.nu_post & 1;
  // this thread's part of its row of the coarsest inverse, loaded now so that its latency
  // passes during the down leg (systems of at most 8 kPre unknowns; FP32 only: the FP64
  // kernel has no registers to spare)
  constexpr int kPre = sizeof(T) == 4 ? kTailDensePre : 1;
  double dpre[kPre];
  {
    const int row = tid >> 3, part = tid & 7, n = d.cn;
#pragma unroll
    for (int k = 0; k < kPre; ++k)
      dpre[k] = (sizeof(T) == 4 && n <= 8 * kPre && row < n && part + 8 * k < n)
                    ? d.dense[(size_t)row * n + part + 8 * k]
                    : 0.0;
  }
  // The holder keeps one level's rows.  Each level's rows are requested one phase before
  // their first sweep (before the right-hand side's load, the restriction or the
  // prolongation), so their load latency passes during that phase; the deepest 9-point
  // level's rows serve both legs.
  TailRowHolder<T, BS> rh;
  rh.level(d.lev[0], tid);
  {  // top-level right-hand side into LDS
    const TailLevel<T>& L = d.lev[0];
    const int n = BS * (L.nx + 1) * (L.ny + 1);
    T* b = lds + L.off + 2 * n;
    for (int i = tid; i < n; i += kTailThreads) b[i] = d.b_in[i];
  }
  __syncthreads();
  for (int l = 0; l + 1 < nl; ++l) {
    const TailLevel<T>& L = d.lev[l];
    const int nn = (L.nx + 1) * (L.ny + 1);
    T* const b0 = lds + L.off;
    T* const b1 = b0 + BS * nn;
    const T* b = b0 + 2 * BS * nn;
    int w = w_first;
    rh.finish();
    rh.template sweep<kSmoothZero>(L, nullptr, b, w ? b1 : b0, d.omega[0], tid);
    __syncthreads();
    for (int k = 1; k < d.nu_pre; ++k) {
      rh.template sweep<kSmooth>(L, w ? b1 : b0, b, w ? b0 : b1, d.omega[k], tid);
      __syncthreads();
      w ^= 1;
    }
    T* const spare = w ? b0 : b1;  // the residual goes to the free ping-pong buffer
    rh.template sweep<kResidual>(L, w ? b1 : b0, b, spare, T(0), tid);
    if (l + 2 < nl) rh.level(d.lev[l + 1], tid);  // (the dense coarsest has no rows)
    __syncthreads();
    const TailLevel<T>& C = d.lev[l + 1];
    const int nnc = (C.nx + 1) * (C.ny + 1);
    T* cb = lds + C.off + (l + 2 == nl ? 0 : 2 * BS * nnc);
    if (tid < nnc) restrict_bilinear_node<T, BS>(L.nx, L.ny, C.nx, C.ny, d.fixed, spare, cb, tid);
    __syncthreads();
  }
  {  // dense coarsest: x = Ainv b, 8 lanes per row (cn <= 128 rows)
    const TailLevel<T>& C = d.lev[nl - 1];
    const int n = d.cn;
    const T* cb = lds + C.off;
    T* cx = lds + C.off + n;
    const int part = tid & 7;
    for (int row0 = 0; row0 < n; row0 += kTailThreads / 8) {  // (all lanes loop alike)
      const int row = row0 + (tid >> 3);
      double s = 0.0;
      if (row < n) {
        if (sizeof(T) == 4 && n <= 8 * kPre && row0 == 0) {  // the row was loaded at the start
#pragma unroll
          for (int k = 0; k < kPre; ++k)
            if (part + 8 * k < n) s += dpre[k] * (double)cb[part + 8 * k];
        } else {
          const double* arow = d.dense + (size_t)row * n;
          for (int j = part; j < n; j += 8) s += arow[j] * (double)cb[j];
        }
      }
      s += __shfl_down(s, 4, 8);
      s += __shfl_down(s, 2, 8);
      s += __shfl_down(s, 1, 8);
      if (row < n && part == 0) cx[row] = (T)s;
    }
    __syncthreads();
  }
  for (int l = nl - 2; l >= 0; --l) {
    const TailLevel<T>& L = d.lev[l];
    const TailLevel<T>& C = d.lev[l + 1];
    const int nn = (L.nx + 1) * (L.ny + 1);
    T* const b0 = lds + L.off;
    T* const b1 = b0 + BS * nn;
    const T* b = b0 + 2 * BS * nn;
    const T* cx = lds + C.off + (l + 2 == nl ? d.cn : 0);
    int w = w_mid;
    if (tid < nn) prolong_bilinear_node<T, BS>(L.nx, L.ny, C.nx, C.ny, d.fixed, cx, w ? b1 : b0, tid);
    if (d.nu_post == 0 && l > 0) rh.level(d.lev[l - 1], tid);
    __syncthreads();
    if (l + 2 < nl) rh.finish();  // (the deepest level's is still set from the down leg)
    for (int k = 0; k < d.nu_post; ++k) {
      rh.template sweep<kSmooth>(L, w ? b1 : b0, b, w ? b0 : b1, d.omega[k], tid);
      if (k + 1 == d.nu_post && l > 0) rh.level(d.lev[l - 1], tid);
      __syncthreads();
      w ^= 1;
    }
    if (l == 0) {
      const int n = BS * nn;
      for (int i = tid; i < n; i += kTailThreads) d.x_out[i] = b0[i];
    }
  }
}

template <typename T, int BS>
__global__ __launch_bounds__(kTailThreads) void k_tail_vcycle(TailDesc<T> d) {
  tail_vcycle_body<T, BS>(hw_blk(), d);
}

template <typename T, int BS>
hipError_t tail_vcycle(const TailDesc<T>& d, size_t lds_bytes, hipStream_t s) {
  if (d.nlev < 2 || d.nlev > kMaxTail || d.cn > 128 || lds_bytes > kTailLdsBytes ||
      d.nu_pre < 1 || d.nu_pre > kMaxNu || d.nu_post < 0 || d.nu_post > kMaxNu)
    return hipErrorInvalidValue;
  for (int l = 0; l + 1 < d.nlev; ++l)
    if ((d.lev[l].nx + 1) * (d.lev[l].ny + 1) > kTailThreads) return hipErrorInvalidValue;
  TMX_LAUNCH_OR_RECORD(kTailThreads, (tail_vcycle_body<T, BS>), 1, lds_bytes, s, (k_tail_vcycle<T, BS>), d);
  return hipGetLastError();
}
template hipError_t tail_vcycle<double, 1>(const TailDesc<double>&, size_t, hipStream_t);
template hipError_t tail_vcycle<double, 2>(const TailDesc<double>&, size_t, hipStream_t);
template hipError_t tail_vcycle<float, 1>(const TailDesc<float>&, size_t, hipStream_t);
template hipError_t tail_vcycle<float, 2>(const TailDesc<float>&, size_t, hipStream_t);

// Galerkin product R A P for tensor-product linear (bilinear) coarsening, one thread per
// coarse node (I, J) and coarse stencil direction kc (nine threads per node): the thread forms
// the coarse block A_c(I,J)[kc] = sum over fine rows f near (I, J) and their stencil neighbours
// g of P(f,I) A_f(f,g) P(g, (I,J)+kc).  The fine nodes f that interpolate from (I, J) lie
// within one node of its fine position p, and their stencil neighbours g within two; the 1-D
// interpolation weights from those five fine positions per axis to the three coarse positions
// I-1..I+1 (and J-1..J+1) are tabulated, so every loop has compile-time trip counts and
// indices: the 9 x 9 stencil entries of the 3 x 3 fine rows are loaded unconditionally (rows
// clamped into the grid; their weight is zero then).  The nine threads of a node read the
// same fine entries (cache hits), and the per-thread work is a ninth of a node's.
template <int BS>
__device__ __forceinline__ void stencil_rap_body(const BlkId blk, int nxf,
                                                 int nyf,
                                                 int nxc,
                                                 int nyc,
                                                 unsigned fixed,
                                                 const double* stf,
                                                 double* stc, float* stc32) {
  const unsigned nnf = (unsigned)(nxf + 1) * (unsigned)(nyf + 1);
  const unsigned nnc = (unsigned)(nxc + 1) * (unsigned)(nyc + 1);
  const unsigned t = blk.id * blockDim.x + threadIdx.x;
  const unsigned i = t / 9u;
  const int kc = (int)(t - 9u * i);
  if (i >= nnc) return;
  const int I = (int)(i % (unsigned)(nxc + 1)), J = (int)(i / (unsigned)(nxc + 1));
  constexpr int B2 = BS * BS;
  double acc[B2];
#pragma unroll
  for (int e = 0; e < B2; ++e) acc[e] = 0.0;
  const int px = cpos(I, nxf, nxc), py = cpos(J, nyf, nyc);
  const int dxc = kc % 3, dyc = kc / 3;  // coarse neighbour (I-1+dxc, J-1+dyc)
  // wx[a]: weight of coarse column I-1+dxc at fine column px-2+a (zero outside the grid);
  // wx1[a]: weight of coarse column I at it (the row side of R)
  float wx[5], wy[5], wx1[5], wy1[5];
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    const int gx = px - 2 + a, gy = py - 2 + a;
    const bool ox = gx < 0 || gx > nxf, oy = gy < 0 || gy > nyf;
    // one interpolation stencil per fine position, read for both coarse columns (rows)
    const Interp1 rx = interp1(gx, nxf, nxc), ry = interp1(gy, nyf, nyc);
    auto wof = [](const Interp1& r, int c) {
      return (float)((r.c[0] == c ? r.w[0] : 0.0) + (r.n > 1 && r.c[1] == c ? r.w[1] : 0.0));
    };
    wx[a] = ox ? 0.f : wof(rx, I - 1 + dxc);
    wy[a] = oy ? 0.f : wof(ry, J - 1 + dyc);
    wx1[a] = ox ? 0.f : wof(rx, I);
    wy1[a] = oy ? 0.f : wof(ry, J);
  }
  const bool self_fixed = node_fixed(fixed, I, J, nxc, nyc);
  const bool live = !node_fixed(fixed, I - 1 + dxc, J - 1 + dyc, nxc, nyc);
  if (!self_fixed && live) {
#pragma unroll
    for (int fyi = 0; fyi < 3; ++fyi)
#pragma unroll
      for (int fxi = 0; fxi < 3; ++fxi) {
        // fine row f = (px-1+fxi, py-1+fyi): table index fxi+1, fyi+1
        const double wf = (double)(wx1[fxi + 1] * wy1[fyi + 1]);
        const int fx = min(max(px - 1 + fxi, 0), nxf), fy = min(max(py - 1 + fyi, 0), nyf);
        const unsigned f = (unsigned)fy * (unsigned)(nxf + 1) + (unsigned)fx;
        // the fine row's 9 B2 entries, all loads in flight before the first is used
        double v[9 * B2];
#pragma unroll
        for (int q = 0; q < 9 * B2; ++q) v[q] = stf[(size_t)q * nnf + f];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const int ax = fxi + k % 3, ay = fyi + k / 3;  // neighbour g at table index (ax, ay)
          const double w = wf * (double)(wx[ax] * wy[ay]);
#pragma unroll
          for (int e = 0; e < B2; ++e) acc[e] += w * v[k * B2 + e];
        }
      }
  }
  if (self_fixed && kc == 4) {
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a * BS + a] = 1.0;
  }
#pragma unroll
  for (int e = 0; e < B2; ++e) {
    stc[(size_t)(kc * B2 + e) * nnc + i] = acc[e];
    if (stc32) stc32[(size_t)(kc * B2 + e) * nnc + i] = (float)acc[e];
  }
}

template <int BS>
__global__ __launch_bounds__(64) void k_stencil_rap(int nxf, int nyf, int nxc, int nyc, unsigned fixed,
                              const double* __restrict__ stf, double* __restrict__ stc,
                              float* __restrict__ stc32) {
  stencil_rap_body<BS>(hw_blk(), nxf, nyf, nxc, nyc, fixed, stf, stc, stc32);
}

template <int BS>
hipError_t stencil_rap(int nxf, int nyf, int nxc, int nyc, int nested, unsigned fixed,
                       const double* stf, double* stc, hipStream_t s, float* stc32) {
  if (nested) return hipErrorInvalidValue;  // 9-point levels are always bilinear coarsenings
  if ((size_t)(nxf + 1) * (nyf + 1) >= (1ull << 31)) return hipErrorInvalidValue;
  TMX_LAUNCH_OR_RECORD(64, (stencil_rap_body<BS>), grid1d(9 * (size_t)(nxc + 1) * (nyc + 1), 64), 0, s, (k_stencil_rap<BS>), nxf, nyf, nxc, nyc, fixed, stf, stc, stc32);
  return hipGetLastError();
}
template hipError_t stencil_rap<1>(int, int, int, int, int, unsigned, const double*, double*,
                                   hipStream_t, float*);
template hipError_t stencil_rap<2>(int, int, int, int, int, unsigned, const double*, double*,
                                   hipStream_t, float*);

template <int BS>
__device__ __forceinline__ void stencil_to_dense_body(const BlkId blk, int nx,
                                                      int ny,
                                                      const double* st,
                                                      double* dense) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t n = BS * nn;
  const size_t v = (size_t)blk.id * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int ix = (int)(v % (nx + 1)), iy = (int)(v / (nx + 1));
  for (int k = 0; k < 9; ++k) {
    const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
    if (jx < 0 || jy < 0 || jx > nx || jy > ny) continue;
    const size_t j = (size_t)jy * (nx + 1) + jx;
    for (int a = 0; a < BS; ++a)
      for (int b = 0; b < BS; ++b)
        dense[(BS * v + a) * n + BS * j + b] = st[(size_t)(k * BS * BS + a * BS + b) * nn + v];
  }
}

template <int BS>
__global__ void k_stencil_to_dense(int nx, int ny, const double* __restrict__ st,
                                   double* __restrict__ dense) {
  stencil_to_dense_body<BS>(hw_blk(), nx, ny, st, dense);
}

template <int BS>
hipError_t stencil_to_dense(int nx, int ny, const double* st, double* dense, hipStream_t s) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  // (a fill kernel, not a memset: the merged batch step records kernels only)
  const hipError_t e = fill_zero_bytes(dense, sizeof(double) * BS * nn * BS * nn, s);
  if (e != hipSuccess) return e;
  TMX_LAUNCH_OR_RECORD(kBlock, (stencil_to_dense_body<BS>), grid1d(nn), 0, s, (k_stencil_to_dense<BS>), nx, ny, st, dense);
  return hipGetLastError();
}
template hipError_t stencil_to_dense<1>(int, int, const double*, double*, hipStream_t);
template hipError_t stencil_to_dense<2>(int, int, const double*, double*, hipStream_t);

// In-place Gauss-Jordan inversion of an SPD matrix (no pivoting), one workgroup.
__device__ __forceinline__ void dense_invert_body(const BlkId blk, int n,
                                                  double* a) {
  extern __shared__ double col[];
  for (int k = 0; k < n; ++k) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = a[(size_t)i * n + k];
    __syncthreads();
    const double piv = col[k];
    for (int j = threadIdx.x; j < n; j += blockDim.x)
      a[(size_t)k * n + j] = (j == k) ? 1.0 / piv : a[(size_t)k * n + j] / piv;
    __syncthreads();
    for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) {
      const int i = (int)(e / n), j = (int)(e % n);
      if (i == k) continue;
      if (j == k)
        a[e] = -col[i] * a[(size_t)k * n + k];
      else
        a[e] -= col[i] * a[(size_t)k * n + j];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_dense_invert(int n, double* __restrict__ a) {
  dense_invert_body(hw_blk(), n, a);
}

// The same elimination with the whole matrix in LDS (n^2 + n doubles <= 64 KB, n <= 90).
__device__ __forceinline__ void dense_invert_lds_body(const BlkId blk, int n,
                                                      double* a) {
  extern __shared__ double lds_m[];
  double* m = lds_m;
  double* col = lds_m + n * n;
  const int nn = n * n;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) m[e] = a[e];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = m[i * n + k];
    __syncthreads();
    const double rp = 1.0 / col[k];
    for (int j = threadIdx.x; j < n; j += blockDim.x) m[k * n + j] = (j == k) ? rp : m[k * n + j] * rp;
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
      const int i = (int)((unsigned)e / (unsigned)n), j = e - i * n;
      if (i == k) continue;
      m[e] = (j == k) ? -col[i] * rp : m[e] - col[i] * m[k * n + j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nn; e += blockDim.x) a[e] = m[e];
}

__global__ __launch_bounds__(1024) void k_dense_invert_lds(int n, double* __restrict__ a) {
  dense_invert_lds_body(hw_blk(), n, a);
}

// The same Gauss-Jordan steps with the matrix in registers, for n <= 4 G: thread (ti, tj) of a
// G x G arrangement holds the 4 x 4 block of rows 4 ti.. and columns 4 tj..; step k publishes
// column k and row k through LDS (double-buffered, so one barrier per step) and every thread
// updates its block from them.  Per step each thread reads 8 values from LDS for its 16
// updates, a quarter of the LDS traffic of the LDS-resident form, with the same arithmetic.
// G is the smallest of 8, 16, 32 that covers n: a barrier of fewer waves per step (one wave
// for n <= 32).
constexpr int kDenseRegMax = 128;
template <int G>
__device__ __forceinline__ void dense_invert_reg_body(const BlkId, int n, double* a) {
  __shared__ double colb[2][4 * G], rowb[2][4 * G];
  const int tid = (int)threadIdx.x, ti = tid / G, tj = tid % G;
  const int i0 = 4 * ti, j0 = 4 * tj;
  double m[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      m[r][c] = (i0 + r < n && j0 + c < n) ? a[(size_t)(i0 + r) * n + j0 + c] : 0.0;
  for (int k = 0; k < n; ++k) {
    const int buf = k & 1;
    if (tj == (k >> 2)) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if ((k & 3) == 0) colb[buf][i0 + r] = m[r][0];
        else if ((k & 3) == 1) colb[buf][i0 + r] = m[r][1];
        else if ((k & 3) == 2) colb[buf][i0 + r] = m[r][2];
        else colb[buf][i0 + r] = m[r][3];
    }
    if (ti == (k >> 2)) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if ((k & 3) == 0) rowb[buf][j0 + c] = m[0][c];
        else if ((k & 3) == 1) rowb[buf][j0 + c] = m[1][c];
        else if ((k & 3) == 2) rowb[buf][j0 + c] = m[2][c];
        else rowb[buf][j0 + c] = m[3][c];
    }
    __syncthreads();
    const double rp = 1.0 / colb[buf][k];
    double cl[4], rw[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) cl[r] = colb[buf][i0 + r];
#pragma unroll
    for (int c = 0; c < 4; ++c) rw[c] = rowb[buf][j0 + c];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = i0 + r, j = j0 + c;
        const double mkj = (j == k) ? rp : rw[c] * rp;  // the scaled pivot row
        if (i == k) m[r][c] = mkj;
        else m[r][c] = (j == k) ? -cl[r] * rp : m[r][c] - cl[r] * mkj;
      }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (i0 + r < n && j0 + c < n) a[(size_t)(i0 + r) * n + j0 + c] = m[r][c];
}

template <int G>
__global__ __launch_bounds__(G * G) void k_dense_invert_reg(int n, double* __restrict__ a) {
  dense_invert_reg_body<G>(hw_blk(), n, a);
}

hipError_t dense_invert(int n, double* a, double* /*work*/, hipStream_t s) {
  if (n >= 1 && n <= kDenseRegMax) {
    if (n <= 32)
      TMX_LAUNCH_OR_RECORD(64, (dense_invert_reg_body<8>), 1, 0, s, (k_dense_invert_reg<8>), n, a);
    else if (n <= 64)
      TMX_LAUNCH_OR_RECORD(256, (dense_invert_reg_body<16>), 1, 0, s, (k_dense_invert_reg<16>), n, a);
    else
      TMX_LAUNCH_OR_RECORD(1024, (dense_invert_reg_body<32>), 1, 0, s, (k_dense_invert_reg<32>), n, a);
    return hipGetLastError();
  }
  // the matrix in LDS up to 160 KiB (n <= 144 unknowns; a 64-node coarsest elasticity level
  // has 128); larger ones stream it from global memory
  constexpr size_t kLdsMax = 160 * 1024;
  const size_t lds = sizeof(double) * ((size_t)n * n + n);
  if (lds > 64 * 1024 && lds <= kLdsMax) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dense_invert_lds),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (lds <= kLdsMax) TMX_LAUNCH_OR_RECORD(1024, (dense_invert_lds_body), 1, lds, s, (k_dense_invert_lds), n, a);
  else TMX_LAUNCH_OR_RECORD(1024, (dense_invert_body), 1, sizeof(double) * n, s, (k_dense_invert), n, a);
  return hipGetLastError();
}

template <typename T>
__device__ __forceinline__ void dense_apply_body(const BlkId blk, int n,
                                                 const double* ainv,
                                                 const T* b,
                                                 T* x) {
  // 8 lanes per row (rows of the symmetric inverse read contiguously)
  const int t = blk.id * blockDim.x + threadIdx.x;
  const int row = t >> 3, part = t & 7;
  double s = 0.0;
  if (row < n)
    for (int j = part; j < n; j += 8) s += ainv[(size_t)row * n + j] * (double)b[j];
  s += __shfl_down(s, 4, 8);
  s += __shfl_down(s, 2, 8);
  s += __shfl_down(s, 1, 8);
  if (row < n && part == 0) x[row] = (T)s;
}

template <typename T>
__global__ void k_dense_apply(int n, const double* __restrict__ ainv, const T* __restrict__ b,
                              T* __restrict__ x) {
  dense_apply_body<T>(hw_blk(), n, ainv, b, x);
}

template <typename T>
hipError_t dense_apply(int n, const double* ainv, const T* b, T* x, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(256, (dense_apply_body<T>), (8 * n + 255) / 256, 0, s, (k_dense_apply<T>), n, ainv, b, x);
  return hipGetLastError();
}
template hipError_t dense_apply<double>(int, const double*, const double*, double*, hipStream_t);
template hipError_t dense_apply<float>(int, const double*, const float*, float*, hipStream_t);

template <typename TI, typename TO>
__device__ __forceinline__ void convert_body(const BlkId blk, size_t n, const TI* in_, TO* out_) {
  const auto in = glob(in_);
  const auto out = glob(out_);
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x)
    out[i] = (TO)in[i];
}

template <typename TI, typename TO>
__global__ void k_convert(size_t n, const TI* __restrict__ in, TO* __restrict__ out) {
  convert_body<TI, TO>(hw_blk(), n, in, out);
}
template <typename TI, typename TO>
hipError_t convert(size_t n, const TI* in, TO* out, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (convert_body<TI, TO>), std::min<unsigned>(grid1d(n), 8192), 0, s, (k_convert<TI, TO>), n, in, out);
  return hipGetLastError();
}
template hipError_t convert<double, float>(size_t, const double*, float*, hipStream_t);
template hipError_t convert<float, double>(size_t, const float*, double*, hipStream_t);

__device__ __forceinline__ void copy_split_body(const BlkId blk, size_t n, const double* in_,
                                                double* out64_, float* out32_) {
  const auto in = glob(in_);
  const auto out64 = glob(out64_);
  const auto out32 = glob(out32_);
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x) {
    const double v = in[i];
    out64[i] = v;
    out32[i] = (float)v;
  }
}

__global__ void k_copy_split(size_t n, const double* __restrict__ in, double* __restrict__ out64,
                             float* __restrict__ out32) {
  copy_split_body(hw_blk(), n, in, out64, out32);
}

hipError_t copy_split(size_t n, const double* in, double* out64, float* out32, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (copy_split_body), std::min<unsigned>(grid1d(n), 8192), 0, s, (k_copy_split), n, in, out64, out32);
  return hipGetLastError();
}

// ============================================================================ vectors
constexpr int kVecBlocks = 1024;

static inline unsigned vec_grid(size_t n) {
  return (unsigned)std::min<size_t>(kVecBlocks, std::max<size_t>(1, (n + kBlock - 1) / kBlock));
}

template <typename T>
__device__ __forceinline__ void dot_body(const BlkId blk, size_t n, const T* a_, const T* b_,
                                         double* part) {
  const auto a = glob(a_);
  const auto b = glob(b_);
  double s = 0.0;
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x)
    s += (double)a[i] * (double)b[i];
  const double v[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(v, o, blk.id);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(size_t n, const T* __restrict__ a,
                                                const T* __restrict__ b, double* __restrict__ part) {
  dot_body<T>(hw_blk(), n, a, b, part);
}

template <typename T>
hipError_t dot_t(size_t n, const T* a, const T* b, RedSlot& r, hipStream_t s) {
  const unsigned g = vec_grid(n);
  r.nblocks = (int)g;
  TMX_LAUNCH_OR_RECORD(kBlock, (dot_body<T>), g, 0, s, (k_dot<T>), n, a, b, r.partials);
  return hipGetLastError();
}
template hipError_t dot_t<double>(size_t, const double*, const double*, RedSlot&, hipStream_t);
template hipError_t dot_t<float>(size_t, const float*, const float*, RedSlot&, hipStream_t);

// Reliable-update PCG (FP32 Krylov vectors, FP64 iterate).  Flush: x64 += xs + alpha p and
// xs = 0 (xs: the FP32 iterate increments since the last flush, alpha p: the increment of the
// current iteration, not yet added).  Replace: r = (float) r64 and partials of |r64|^2, so the
// recursive FP32 residual restarts from the true FP64 one and the stop test sees the true norm.
__global__ __launch_bounds__(kBlock) void k_ru_flush(size_t n, const double* __restrict__ alpha,
                                                    const float* __restrict__ p,
                                                    float* __restrict__ xs,
                                                    double* __restrict__ x) {
  const double a = *alpha;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    x[i] += (double)xs[i] + a * (double)p[i];
    xs[i] = 0.0f;
  }
}

__global__ __launch_bounds__(kBlock) void k_ru_replace(size_t n, const double* __restrict__ r64,
                                                      float* __restrict__ r,
                                                      double* __restrict__ part) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = r64[i];
    r[i] = (float)v;
    s += v * v;
  }
  const double v[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(v, o);
}

__global__ __launch_bounds__(kBlock) void k_add_vec(size_t n, const double* __restrict__ a,
                                                    double* __restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += a[i];
}

hipError_t add_vec(size_t n, const double* a, double* y, hipStream_t s) {
  k_add_vec<<<vec_grid(n), kBlock, 0, s>>>(n, a, y);
  return hipGetLastError();
}

hipError_t ru_flush(size_t n, const double* alpha, const float* p, float* xs, double* x,
                    hipStream_t s) {
  k_ru_flush<<<vec_grid(n), kBlock, 0, s>>>(n, alpha, p, xs, x);
  return hipGetLastError();
}

hipError_t ru_replace(size_t n, const double* r64, float* r, RedSlot& rr, hipStream_t s) {
  const unsigned g = vec_grid(n);
  rr.nblocks = (int)g;
  k_ru_replace<<<g, kBlock, 0, s>>>(n, r64, r, rr.partials);
  return hipGetLastError();
}

__device__ __forceinline__ void reduce_to_body(const BlkId blk, const double* part,
                                               int n,
                                               double* out, double* host_out) {
  const double t = sum_partials(part, n);
  if (threadIdx.x == 0) {
    *out = t;
    if (host_out) {  // (as publish_scalar: the value to mapped host memory)
      host_out[0] = t;
      __threadfence_system();
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_to(const double* __restrict__ part, int n,
                                                      double* __restrict__ out,
                                                      double* __restrict__ host_out) {
  reduce_to_body(hw_blk(), part, n, out, host_out);
}

template <typename T>
__device__ __forceinline__ void axpy_dev_body(const BlkId blk, size_t n, const double* a,
                                              const T* x_, T* y_, const double* gate) {
  const auto x = glob(x_);
  const auto y = glob(y_);
  if (gate && *glob(gate) == 0.0) return;
  const T av = (T)*glob(a);
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n;
       i += (size_t)blk.n * blockDim.x)
    y[i] += av * x[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_dev(size_t n, const double* __restrict__ a,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     const double* __restrict__ gate) {
  axpy_dev_body<T>(hw_blk(), n, a, x, y, gate);
}

template <typename T>
hipError_t axpy_dev(size_t n, const double* a, const T* x, T* y, hipStream_t s,
                    const double* gate) {
  TMX_LAUNCH_OR_RECORD(kBlock, (axpy_dev_body<T>), vec_grid(n), 0, s, (k_axpy_dev<T>), n, a, x, y, gate);
  return hipGetLastError();
}
template hipError_t axpy_dev<double>(size_t, const double*, const double*, double*, hipStream_t,
                                     const double*);
template hipError_t axpy_dev<float>(size_t, const double*, const float*, float*, hipStream_t,
                                    const double*);

hipError_t reduce_to(const RedSlot& r, double* out, hipStream_t s, double* host_out) {
  TMX_LAUNCH_OR_RECORD(kBlock, (reduce_to_body), 1, 0, s, (k_reduce_to), r.partials, r.nblocks, out, host_out);
  return hipGetLastError();
}

hipError_t dot(size_t n, const double* a, const double* b, RedSlot& r, double* out,
               hipStream_t s, double* host_out) {
  hipError_t e = dot_t<double>(n, a, b, r, s);
  if (e != hipSuccess) return e;
  return reduce_to(r, out, s, host_out);
}

// PCG scalar steps (one block).  sc layout: see PcgSlots.
__global__ __launch_bounds__(kBlock) void k_pcg_alpha(const double* __restrict__ part, int n,
                                                      double* __restrict__ sc, PcgSlots q) {
  const double pq = sum_partials(part, n);
  if (threadIdx.x == 0) {
    sc[q.pq] = pq;
    const bool stop = sc[q.flag] != 0.0;
    if (!stop && !(pq > 0.0)) sc[q.flag] = (pq != pq) ? 3.0 : 2.0;  // NaN : not SPD
    sc[q.alpha] = (stop || !(pq > 0.0)) ? 0.0 : sc[q.rz] / pq;
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_rr(const double* __restrict__ part, int n,
                                                   double* __restrict__ sc, PcgSlots q) {
  const double rr = sum_partials(part, n);
  if (threadIdx.x == 0) {
    if (sc[q.flag] == 0.0) {
      sc[q.rr] = rr;
      sc[q.iters] += 1.0;
      if (!(rr == rr) || rr > 1e300) sc[q.flag] = 3.0;
      else if (rr <= sc[q.tol2]) sc[q.flag] = 1.0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_beta(const double* __restrict__ part, int n,
                                                     double* __restrict__ sc, PcgSlots q) {
  const double rzn = sum_partials(part, n);
  if (threadIdx.x == 0) {
    const bool stop = sc[q.flag] != 0.0;
    sc[q.beta] = stop ? 0.0 : rzn / sc[q.rz];
    if (!stop) sc[q.rz] = rzn;
  }
}

hipError_t pcg_alpha(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s) {
  k_pcg_alpha<<<1, kBlock, 0, s>>>(r.partials, r.nblocks, sc, q);
  return hipGetLastError();
}
hipError_t pcg_rr(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s) {
  k_pcg_rr<<<1, kBlock, 0, s>>>(r.partials, r.nblocks, sc, q);
  return hipGetLastError();
}
hipError_t pcg_beta(const RedSlot& r, double* sc, const PcgSlots& q, hipStream_t s) {
  k_pcg_beta<<<1, kBlock, 0, s>>>(r.partials, r.nblocks, sc, q);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update_xr(size_t n, const double* __restrict__ sc,
                                                          int i_alpha, T* __restrict__ x,
                                                          T* __restrict__ r, const T* __restrict__ p,
                                                          const T* __restrict__ q,
                                                          double* __restrict__ part) {
  const T al = (T)sc[i_alpha];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    x[i] += al * p[i];
    const T ri = r[i] - al * q[i];
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  const double v[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(v, o);
}

template <typename T>
hipError_t pcg_update_xr(size_t n, const double* sc, int i_alpha, T* x, T* r, const T* p,
                         const T* q, RedSlot& rr, hipStream_t s) {
  const unsigned g = vec_grid(n);
  rr.nblocks = (int)g;
  k_pcg_update_xr<T><<<g, kBlock, 0, s>>>(n, sc, i_alpha, x, r, p, q, rr.partials);
  return hipGetLastError();
}
template hipError_t pcg_update_xr<double>(size_t, const double*, int, double*, double*,
                                          const double*, const double*, RedSlot&, hipStream_t);
template hipError_t pcg_update_xr<float>(size_t, const double*, int, float*, float*,
                                         const float*, const float*, RedSlot&, hipStream_t);

template <typename T>
__device__ __forceinline__ void pcg_update_xr_fused_body(const BlkId blk, size_t n, double* sc,
                                                         PcgSlots q, const double* pqpart, int npq,
                                                         T* x_, T* r_, const T* p_, const T* qv_,
                                                         double* part, float* rlow_) {
  const auto x = glob(x_);
  const auto r = glob(r_);
  const auto p = glob(p_);
  const auto qv = glob(qv_);
  const auto rlow = glob(rlow_);
  __shared__ double alpha_sh;
  const double pq = sum_partials(pqpart, npq);
  if (threadIdx.x == 0) {
    const bool stop = sc[q.flag] != 0.0;
    alpha_sh = (stop || !(pq > 0.0)) ? 0.0 : sc[q.rz] / pq;
  }
  __syncthreads();
  const T al = (T)alpha_sh;
  if (!x && blk.id == 0 && threadIdx.x == 0) sc[q.alpha] = alpha_sh;  // for the deferred x
  double s = 0.0;
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x) {
    if (x) x[i] += al * p[i];
    const T ri = r[i] - al * qv[i];
    r[i] = ri;
    if (rlow) rlow[i] = (float)ri;
    s += (double)ri * (double)ri;
  }
  const double v[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(v, o, blk.id);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update_xr_fused(
    size_t n, double* __restrict__ sc, PcgSlots q, const double* __restrict__ pqpart, int npq,
    T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ qv,
    double* __restrict__ part, float* __restrict__ rlow) {
  pcg_update_xr_fused_body<T>(hw_blk(), n, sc, q, pqpart, npq, x, r, p, qv, part, rlow);
}

template <typename T>
hipError_t pcg_update_xr_fused(size_t n, double* sc, const PcgSlots& q, const RedSlot& pq,
                               T* x, T* r, const T* p, const T* qv, RedSlot& rr, hipStream_t s,
                               float* rlow) {
  const unsigned g = vec_grid(n);
  rr.nblocks = (int)g;
  TMX_LAUNCH_OR_RECORD(kBlock, (pcg_update_xr_fused_body<T>), g, 0, s, (k_pcg_update_xr_fused<T>), n, sc, q, pq.partials, pq.nblocks, x, r, p, qv, rr.partials, rlow);
  return hipGetLastError();
}
template hipError_t pcg_update_xr_fused<double>(size_t, double*, const PcgSlots&,
                                                const RedSlot&, double*, double*, const double*,
                                                const double*, RedSlot&, hipStream_t, float*);
template hipError_t pcg_update_xr_fused<float>(size_t, double*, const PcgSlots&,
                                               const RedSlot&, float*, float*, const float*,
                                               const float*, RedSlot&, hipStream_t, float*);

// One workgroup of kFinThreads: the three reductions share one pass over the partials (each
// thread keeps its loads in flight) and one barrier, so the kernel costs about one memory
// round trip plus the pinned-memory report.
constexpr int kFinThreads = 1024;
__device__ __forceinline__ void pcg_finalize_body(const BlkId blk, const double* pqp,
                                                  int npq,
                                                  const double* rrp,
                                                  int nrr,
                                                  const double* rzp,
                                                  int nrz,
                                                  double* sc,
                                                  PcgSlots q,
                                                  double* hout) {
  __shared__ double sh[3][kFinThreads / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  // the scalars the update reads, loaded before the partial sums so that their latency
  // passes meanwhile (thread 0 alone uses them)
  double flag0 = 0.0, tol2 = 0.0, rz_old = 0.0, seq0 = 0.0, iters0 = 0.0, rr_old = 0.0;
  if (tid == 0) {
    flag0 = sc[q.flag];
    tol2 = sc[q.tol2];
    rz_old = sc[q.rz];
    seq0 = sc[q.seq];
    iters0 = sc[q.iters];
    rr_old = sc[q.rr];
  }
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  const int n = max(npq, max(nrr, nrz));
  for (int i = tid; i < n; i += kFinThreads) {
    const double a = i < npq ? pqp[i] : 0.0;
    const double b = i < nrr ? rrp[i] : 0.0;
    const double c = i < nrz ? rzp[i] : 0.0;
    v0 += a;
    v1 += b;
    v2 += c;
  }
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  v2 = wave_sum(v2);
  if (lane == 0) {
    sh[0][w] = v0;
    sh[1][w] = v1;
    sh[2][w] = v2;
  }
  __syncthreads();
  if (tid == 0) {
    double pq = 0.0, rr = 0.0, rz = 0.0;
    for (int k = 0; k < kFinThreads / kWave; ++k) {
      pq += sh[0][k];
      rr += sh[1][k];
      rz += sh[2][k];
    }
    const bool was_stopped = flag0 != 0.0;
    sc[q.skip] = was_stopped ? 1.0 : 0.0;
    double flag = flag0, rr_now = rr_old, iters = iters0;
    if (!was_stopped) {
      sc[q.pq] = pq;
      rr_now = rr;
      iters = iters0 + 1.0;
      if (!(pq == pq) || !(rr == rr) || !(rz == rz) || rr > 1e300) flag = 3.0;
      else if (!(pq > 0.0)) flag = 2.0;
      else if (rr <= tol2) flag = 1.0;
      sc[q.rr] = rr_now;
      sc[q.iters] = iters;
      sc[q.flag] = flag;
    }
    const bool stop = flag != 0.0;
    sc[q.beta] = stop ? 0.0 : rz / rz_old;
    if (!stop) sc[q.rz] = rz;
    const double seq = seq0 + 1.0;
    sc[q.seq] = seq;
    if (hout) {  // stop flag, r.r and iteration count straight to pinned host memory, then
                 // the sequence number the host polls for, once the rest is visible
      hout[0] = flag;
      hout[1] = rr_now;
      hout[2] = iters;
      __threadfence_system();
      hout[3] = seq;
    }
  }
}

__global__ __launch_bounds__(kFinThreads) void k_pcg_finalize(const double* __restrict__ pqp, int npq,
                                                              const double* __restrict__ rrp, int nrr,
                                                              const double* __restrict__ rzp, int nrz,
                                                              double* __restrict__ sc, PcgSlots q,
                                                              double* __restrict__ hout) {
  pcg_finalize_body(hw_blk(), pqp, npq, rrp, nrr, rzp, nrz, sc, q, hout);
}

hipError_t pcg_finalize(const RedSlot& pq, const RedSlot& rr, const RedSlot& rz, double* sc,
                        const PcgSlots& q, hipStream_t s, double* host_out) {
  TMX_LAUNCH_OR_RECORD(kFinThreads, (pcg_finalize_body), 1, 0, s, (k_pcg_finalize), pq.partials, pq.nblocks, rr.partials, rr.nblocks, rz.partials, rz.nblocks, sc, q, host_out);
  return hipGetLastError();
}

__device__ __forceinline__ void pcg_start_body(const BlkId blk, const double* rzp,
                                               int nrz,
                                               double* sc,
                                               PcgSlots q) {
  const double rz = sum_partials(rzp, nrz);
  if (threadIdx.x == 0) {
    sc[q.rz] = rz;
    sc[q.beta] = 0.0;
    // keep a start flag set by pcg_begin_final (b = 0 or non-finite)
    if (sc[q.flag] == 0.0 && !(rz == rz)) sc[q.flag] = 3.0;
    sc[q.iters] = 0.0;
    sc[q.seq] = 0.0;
    sc[q.skip] = 0.0;
    sc[q.alpha] = 0.0;  // no deferred x update before the first direction
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_start(const double* __restrict__ rzp, int nrz,
                                                      double* __restrict__ sc, PcgSlots q) {
  pcg_start_body(hw_blk(), rzp, nrz, sc, q);
}

template <typename T>
__device__ __forceinline__ void pcg_begin_body(const BlkId blk, size_t n, const T* b_, T* x_,
                                               T* r_, float* rlow_, double* part) {
  const auto b = glob(b_);
  const auto x = glob(x_);
  const auto r = glob(r_);
  const auto rlow = glob(rlow_);
  double s = 0.0;
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n;
       i += (size_t)blk.n * blockDim.x) {
    const T v = b[i];
    r[i] = v;
    x[i] = T(0);
    if (rlow) rlow[i] = (float)v;
    s += (double)v * (double)v;
  }
  const double v1[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(v1, o, blk.id);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_begin(size_t n, const T* __restrict__ b,
                                                      T* __restrict__ x, T* __restrict__ r,
                                                      float* __restrict__ rlow,
                                                      double* __restrict__ part) {
  pcg_begin_body<T>(hw_blk(), n, b, x, r, rlow, part);
}

template <typename T>
hipError_t pcg_begin(size_t n, const T* b, T* x, T* r, float* rlow, RedSlot& bb, hipStream_t s) {
  const unsigned g = vec_grid(n);
  bb.nblocks = (int)g;
  TMX_LAUNCH_OR_RECORD(kBlock, (pcg_begin_body<T>), g, 0, s, (k_pcg_begin<T>), n, b, x, r, rlow, bb.partials);
  return hipGetLastError();
}
template hipError_t pcg_begin<double>(size_t, const double*, double*, double*, float*, RedSlot&,
                                      hipStream_t);
template hipError_t pcg_begin<float>(size_t, const float*, float*, float*, float*, RedSlot&,
                                     hipStream_t);

__device__ __forceinline__ void pcg_begin_final_body(const BlkId blk, const double* part,
                                                     int n,
                                                     double* sc,
                                                     PcgSlots q,
                                                     double rtol2,
                                                     int bb_slot,
                                                     double* host_bb) {
  const double bb = sum_partials(part, n);
  if (threadIdx.x == 0) {
    sc[bb_slot] = bb;
    sc[q.rr] = bb;
    sc[q.tol2] = rtol2 * bb;
    sc[q.flag] = !(bb == bb) || bb > 1e300 ? 3.0 : (bb == 0.0 ? 1.0 : 0.0);
    if (host_bb) {
      host_bb[0] = bb;
      __threadfence_system();
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_begin_final(const double* __restrict__ part, int n,
                                                            double* __restrict__ sc, PcgSlots q,
                                                            double rtol2, int bb_slot,
                                                            double* __restrict__ host_bb) {
  pcg_begin_final_body(hw_blk(), part, n, sc, q, rtol2, bb_slot, host_bb);
}

hipError_t pcg_begin_final(const RedSlot& bb, double* sc, const PcgSlots& q, double rtol2,
                           int bb_slot, double* host_bb, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (pcg_begin_final_body), 1, 0, s, (k_pcg_begin_final), bb.partials, bb.nblocks, sc, q, rtol2, bb_slot, host_bb);
  return hipGetLastError();
}

hipError_t pcg_start(const RedSlot& rz, double* sc, const PcgSlots& q, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (pcg_start_body), 1, 0, s, (k_pcg_start), rz.partials, rz.nblocks, sc, q);
  return hipGetLastError();
}

template <typename T>
__global__ void k_pcg_update_p(size_t n, const double* __restrict__ sc, int i_beta,
                               const T* __restrict__ z, T* __restrict__ p) {
  const T beta = (T)sc[i_beta];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

template <typename T>
hipError_t pcg_update_p(size_t n, const double* sc, int i_beta, const T* z, T* p, hipStream_t s) {
  k_pcg_update_p<T><<<std::min<unsigned>(grid1d(n), 8192), kBlock, 0, s>>>(n, sc, i_beta, z, p);
  return hipGetLastError();
}
template hipError_t pcg_update_p<double>(size_t, const double*, int, const double*, double*, hipStream_t);
template hipError_t pcg_update_p<float>(size_t, const double*, int, const float*, float*, hipStream_t);

// count independent copies in one launch: blockIdx.y picks the task
__global__ __launch_bounds__(kBlock) void k_copy_tasks(const CopyTask* __restrict__ tasks) {
  const CopyTask t = tasks[blockIdx.y];
  const double* __restrict__ src = t.src;
  double* __restrict__ dst = t.dst;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < t.n; i += (size_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

hipError_t copy_tasks(const CopyTask* tasks, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  k_copy_tasks<<<dim3(64, (unsigned)count), kBlock, 0, s>>>(tasks);
  return hipGetLastError();
}

// one device scalar to pinned host memory (the host reads it after the stream has drained)
__device__ __forceinline__ void publish_body(const BlkId, const double* src, double* host_dst) {
  if (threadIdx.x == 0) {
    host_dst[0] = src[0];
    __threadfence_system();
  }
}

__global__ __launch_bounds__(64) void k_publish(const double* __restrict__ src,
                                                double* __restrict__ host_dst) {
  publish_body(hw_blk(), src, host_dst);
}

hipError_t publish_scalar(const double* src, double* host_dst, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(64, (publish_body), 1, 0, s, (k_publish), src, host_dst);
  return hipGetLastError();
}

template <typename T>
__device__ __forceinline__ void copy_body(const BlkId blk, size_t n, const T* a_, T* b_) {
  const auto a = glob(a_);
  const auto b = glob(b_);
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x)
    b[i] = a[i];
}

template <typename T>
__global__ void k_copy(size_t n, const T* __restrict__ a, T* __restrict__ b) {
  copy_body<T>(hw_blk(), n, a, b);
}
template <typename T>
hipError_t copy_vec(size_t n, const T* a, T* b, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (copy_body<T>), std::min<unsigned>(grid1d(n), 8192), 0, s, (k_copy<T>), n, a, b);
  return hipGetLastError();
}
template hipError_t copy_vec<double>(size_t, const double*, double*, hipStream_t);
template hipError_t copy_vec<float>(size_t, const float*, float*, hipStream_t);

template <typename T>
__global__ void k_axpy_scaled(size_t n, const double* __restrict__ sc, int i, const T* __restrict__ d,
                              double* __restrict__ x) {
  const double scale = sqrt(sc[i]);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    x[k] += scale * (double)d[k];
}
template <typename T>
hipError_t axpy_scaled(size_t n, const double* sc, int i, const T* d, double* x, hipStream_t s) {
  k_axpy_scaled<T><<<std::min<unsigned>(grid1d(n), 8192), kBlock, 0, s>>>(n, sc, i, d, x);
  return hipGetLastError();
}
template hipError_t axpy_scaled<float>(size_t, const double*, int, const float*, double*, hipStream_t);
template hipError_t axpy_scaled<double>(size_t, const double*, int, const double*, double*, hipStream_t);

template <typename T>
__global__ void k_scale_to(size_t n, const double* __restrict__ sc, int i, const double* __restrict__ in,
                           T* __restrict__ out) {
  const double inv = 1.0 / sqrt(sc[i]);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    out[k] = (T)(in[k] * inv);
}
template <typename T>
hipError_t scale_to(size_t n, const double* sc, int i, const double* in, T* out, hipStream_t s) {
  k_scale_to<T><<<std::min<unsigned>(grid1d(n), 8192), kBlock, 0, s>>>(n, sc, i, in, out);
  return hipGetLastError();
}
template hipError_t scale_to<float>(size_t, const double*, int, const double*, float*, hipStream_t);
template hipError_t scale_to<double>(size_t, const double*, int, const double*, double*, hipStream_t);

__device__ __forceinline__ void fill_zero_body(const BlkId blk, size_t n,
                                               unsigned* p) {
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x)
    p[i] = 0u;
}

__global__ void k_fill_zero(size_t n, unsigned* __restrict__ p) {
  fill_zero_body(hw_blk(), n, p);
}
hipError_t fill_zero_bytes(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 4;
  TMX_LAUNCH_OR_RECORD(kBlock, (fill_zero_body), std::min<unsigned>(grid1d(n), 8192), 0, s, (k_fill_zero), n, static_cast<unsigned*>(p));
  return hipGetLastError();
}

__global__ void k_mask_fixed(int nx, int ny, unsigned fixed, double* __restrict__ v) {
  const size_t nn = (size_t)(nx + 1) * (ny + 1);
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int ix = (int)(i % (nx + 1)), iy = (int)(i / (nx + 1));
  if (node_fixed(fixed, ix, iy, nx, ny)) {
    v[2 * i] = 0.0;
    v[2 * i + 1] = 0.0;
  }
}
hipError_t mask_fixed(int nx, int ny, unsigned fixed, double* v, hipStream_t s) {
  k_mask_fixed<<<grid1d((size_t)(nx + 1) * (ny + 1)), kBlock, 0, s>>>(nx, ny, fixed, v);
  return hipGetLastError();
}

__global__ void k_set_scalar(double* __restrict__ sc, int i, double v) {
  if (threadIdx.x == 0) sc[i] = v;
}
hipError_t set_scalar(double* sc, int i, double v, hipStream_t s) {
  k_set_scalar<<<1, 64, 0, s>>>(sc, i, v);
  return hipGetLastError();
}

__device__ __forceinline__ void fill_body(const BlkId blk, size_t n,
                                          double val,
                                          double* p) {
  for (size_t i = (size_t)blk.id * blockDim.x + threadIdx.x; i < n; i += (size_t)blk.n * blockDim.x)
    p[i] = val;
}

__global__ void k_fill(size_t n, double val, double* __restrict__ p) {
  fill_body(hw_blk(), n, val, p);
}
hipError_t fill(size_t n, double val, double* p, hipStream_t s) {
  TMX_LAUNCH_OR_RECORD(kBlock, (fill_body), std::min<unsigned>(grid1d(n), 8192), 0, s, (k_fill), n, val, p);
  return hipGetLastError();
}


// ============================================================================ EMD
// Entropic mirror descent on nodal designs (vertex order); integrals use the lumped CG1 mass
// m_i = int phi_i, so sum_i m_i f_i is the integral of the CG1 interpolant of f.
__device__ __forceinline__ double expit_d(double x) { return 1.0 / (1.0 + exp(-x)); }

__global__ __launch_bounds__(kBlock) void k_emd_half_step(size_t n, const double* __restrict__ psi,
                                                          const double* __restrict__ w, double step,
                                                          double* __restrict__ half) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    half[i] = psi[i] - step * w[i];
}

// partials of int expit(half + c) and int expit'(half + c)
__global__ __launch_bounds__(kBlock) void k_emd_volume(size_t n, const double* __restrict__ half,
                                                       double c, const double* __restrict__ m,
                                                       double* __restrict__ p0,
                                                       double* __restrict__ p1) {
  double s0 = 0.0, s1 = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double e = expit_d(half[i] + c);
    s0 += m[i] * e;
    s1 += m[i] * (e * (1.0 - e));
  }
  const double v[2] = {s0, s1};
  double* const o[2] = {p0, p1};
  block_partials<2>(v, o);
}

__global__ __launch_bounds__(kBlock) void k_emd_apply(size_t n, const double* __restrict__ half,
                                                      double c, double* __restrict__ psi,
                                                      double* __restrict__ rho) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double p = half[i] + c;
    psi[i] = p;
    rho[i] = expit_d(p);
  }
}

// partials of int (rho - expit(psi_prev))^2
__global__ __launch_bounds__(kBlock) void k_emd_change(size_t n, const double* __restrict__ rho,
                                                       const double* __restrict__ psi_prev,
                                                       const double* __restrict__ m,
                                                       double* __restrict__ p0) {
  double s0 = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double d = rho[i] - expit_d(psi_prev[i]);
    s0 += m[i] * d * d;
  }
  const double v[1] = {s0};
  double* const o[1] = {p0};
  block_partials<1>(v, o);
}

__global__ __launch_bounds__(kBlock) void k_logit(size_t n, const double* __restrict__ rho,
                                                  double* __restrict__ psi) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    psi[i] = log(rho[i] / (1.0 - rho[i]));
}

// two partial arrays -> sc[i0], sc[i1]
__global__ __launch_bounds__(kBlock) void k_reduce_pair(const double* __restrict__ p0,
                                                        const double* __restrict__ p1, int n,
                                                        double* __restrict__ o0,
                                                        double* __restrict__ o1) {
  const double a = sum_partials(p0, n);
  const double b = sum_partials(p1, n);
  if (threadIdx.x == 0) {
    *o0 = a;
    *o1 = b;
  }
}

hipError_t emd_half_step(size_t n, const double* psi, const double* w, double step, double* half,
                         hipStream_t s) {
  k_emd_half_step<<<vec_grid(n), kBlock, 0, s>>>(n, psi, w, step, half);
  return hipGetLastError();
}

hipError_t emd_volume(size_t n, const double* half, double c, const double* m, RedSlot& r0,
                      RedSlot& r1, double* out0, double* out1, hipStream_t s) {
  const unsigned g = vec_grid(n);
  r0.nblocks = r1.nblocks = (int)g;
  k_emd_volume<<<g, kBlock, 0, s>>>(n, half, c, m, r0.partials, r1.partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_reduce_pair<<<1, kBlock, 0, s>>>(r0.partials, r1.partials, (int)g, out0, out1);
  return hipGetLastError();
}

hipError_t emd_apply(size_t n, const double* half, double c, double* psi, double* rho,
                     hipStream_t s) {
  k_emd_apply<<<vec_grid(n), kBlock, 0, s>>>(n, half, c, psi, rho);
  return hipGetLastError();
}

hipError_t emd_change(size_t n, const double* rho, const double* psi_prev, const double* m,
                      RedSlot& r, double* out, hipStream_t s) {
  const unsigned g = vec_grid(n);
  r.nblocks = (int)g;
  k_emd_change<<<g, kBlock, 0, s>>>(n, rho, psi_prev, m, r.partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return reduce_to(r, out, s);
}

hipError_t logit(size_t n, const double* rho, double* psi, hipStream_t s) {
  k_logit<<<vec_grid(n), kBlock, 0, s>>>(n, rho, psi);
  return hipGetLastError();
}

// ============================================================================ MINRES
// Preconditioned MINRES (Paige & Saunders 1975, the Lanczos form with Givens rotations) for
// K(xi) when the filtered design dips below zero and an odd penalty makes K indefinite.
// Scalars live in sc[q.base + k] (MrSlots); every kernel reads what the previous ones wrote.
__global__ __launch_bounds__(kBlock) void k_mr_lanczos1(size_t n, const double* __restrict__ sc,
                                                        MrSlots q, const double* __restrict__ v,
                                                        const double* __restrict__ r1,
                                                        double* __restrict__ y,
                                                        double* __restrict__ part) {
  const double beta = sc[q.beta], oldb = sc[q.oldb];
  const double c = oldb != 0.0 ? beta / oldb : 0.0;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double yi = y[i] - c * r1[i];
    y[i] = yi;
    s += v[i] * yi;
  }
  const double vv[1] = {s};
  double* const o[1] = {part};
  block_partials<1>(vv, o);
}

__global__ __launch_bounds__(kBlock) void k_mr_lanczos2(size_t n, double* __restrict__ sc,
                                                        MrSlots q, const double* __restrict__ part,
                                                        int npart, const double* __restrict__ r2,
                                                        double* __restrict__ y) {
  const double alfa = sum_partials(part, npart);
  const double c = alfa / sc[q.beta];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] -= c * r2[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[q.alfa] = alfa;
}

// start: beta1 = sqrt(r1 . M r1) from partials; Lanczos and rotation state reset
__global__ __launch_bounds__(kBlock) void k_mr_start(const double* __restrict__ part, int npart,
                                                     double* __restrict__ sc, MrSlots q,
                                                     double tol) {
  const double rz = sum_partials(part, npart);
  if (threadIdx.x == 0) {
    const double b1 = rz > 0.0 ? sqrt(rz) : 0.0;
    sc[q.beta1] = b1;
    sc[q.beta] = b1;
    sc[q.oldb] = 0.0;
    sc[q.dbar] = 0.0;
    sc[q.epsln] = 0.0;
    sc[q.phibar] = b1;
    sc[q.cs] = -1.0;
    sc[q.sn] = 0.0;
    sc[q.iters] = 0.0;
    sc[q.tol] = tol;
    sc[q.flag] = !(rz == rz) ? 3.0 : rz < 0.0 ? 2.0 : rz == 0.0 ? 1.0 : 0.0;
    sc[q.skip] = 0.0;
    sc[q.seq] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_mr_step(const double* __restrict__ part, int npart,
                                                    double* __restrict__ sc, MrSlots q,
                                                    double* __restrict__ hout) {
  const double rz = sum_partials(part, npart);
  if (threadIdx.x == 0) {
    const double seq = sc[q.seq] + 1.0;
    sc[q.seq] = seq;
    if (sc[q.flag] != 0.0) {  // stopped before this iteration, which then changes nothing
      sc[q.skip] = 1.0;
      hout[0] = sc[q.flag];
      hout[1] = sc[q.phibar] / sc[q.beta1];
      hout[2] = sc[q.iters];
      __threadfence_system();
      hout[3] = seq;
      return;
    }
    sc[q.skip] = 0.0;
    const double alfa = sc[q.alfa];
    const double beta = rz > 0.0 ? sqrt(rz) : 0.0;
    const double cs = sc[q.cs], sn = sc[q.sn], dbar = sc[q.dbar];
    const double oldeps = sc[q.epsln];
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    const double gamma = fmax(sqrt(gbar * gbar + beta * beta), 1e-300);
    const double cs2 = gbar / gamma, sn2 = beta / gamma;
    const double phibar = sc[q.phibar];
    sc[q.oldb] = sc[q.beta];
    sc[q.beta] = beta;
    sc[q.oldeps] = oldeps;
    sc[q.delta] = delta;
    sc[q.epsln] = sn * beta;
    sc[q.dbar] = -cs * beta;
    sc[q.cs] = cs2;
    sc[q.sn] = sn2;
    sc[q.phi] = cs2 * phibar;
    sc[q.phibar] = sn2 * phibar;
    sc[q.denom] = 1.0 / gamma;
    sc[q.iters] += 1.0;
    double flag = 0.0;
    if (!(rz == rz) || !(alfa == alfa)) flag = 3.0;
    else if (rz < 0.0) flag = 2.0;  // preconditioner not positive definite
    else if (sn2 * phibar <= sc[q.tol] * sc[q.beta1] || rz == 0.0) flag = 1.0;
    sc[q.flag] = flag;
    // the report, then the sequence number the host polls for, once the rest is visible
    hout[0] = flag;
    hout[1] = sc[q.phibar] / sc[q.beta1];
    hout[2] = sc[q.iters];
    __threadfence_system();
    hout[3] = seq;
  }
}

// w = (v - oldeps w_older - delta w_newer) / gamma into w_out; x += phi w; v = z / beta
__global__ __launch_bounds__(kBlock) void k_mr_update(size_t n, const double* __restrict__ sc,
                                                      MrSlots q, double* __restrict__ v,
                                                      const double* __restrict__ z,
                                                      double* __restrict__ w_out,
                                                      const double* __restrict__ w_older,
                                                      const double* __restrict__ w_newer,
                                                      double* __restrict__ x) {
  if (sc[q.skip] != 0.0) return;  // (uniform: an iteration after the stop)
  const double oldeps = sc[q.oldeps], delta = sc[q.delta], denom = sc[q.denom];
  const double phi = sc[q.phi], beta = sc[q.beta];
  const double rb = beta > 0.0 ? 1.0 / beta : 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double w = (v[i] - oldeps * w_older[i] - delta * w_newer[i]) * denom;
    w_out[i] = w;
    x[i] += phi * w;
    v[i] = z[i] * rb;
  }
}

__global__ __launch_bounds__(kBlock) void k_mr_scale(size_t n, const double* __restrict__ sc,
                                                     MrSlots q, const double* __restrict__ z,
                                                     double* __restrict__ v) {
  const double beta = sc[q.beta];
  const double rb = beta > 0.0 ? 1.0 / beta : 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v[i] = z[i] * rb;
}

hipError_t mr_start(const RedSlot& rz, double* sc, const MrSlots& q, double tol, hipStream_t s) {
  k_mr_start<<<1, kBlock, 0, s>>>(rz.partials, rz.nblocks, sc, q, tol);
  return hipGetLastError();
}
hipError_t mr_scale(size_t n, const double* sc, const MrSlots& q, const double* z, double* v,
                    hipStream_t s) {
  k_mr_scale<<<vec_grid(n), kBlock, 0, s>>>(n, sc, q, z, v);
  return hipGetLastError();
}
hipError_t mr_lanczos(size_t n, double* sc, const MrSlots& q, const double* v, const double* r1,
                      const double* r2, double* y, RedSlot& part, hipStream_t s) {
  const unsigned g = vec_grid(n);
  part.nblocks = (int)g;
  k_mr_lanczos1<<<g, kBlock, 0, s>>>(n, sc, q, v, r1, y, part.partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_mr_lanczos2<<<g, kBlock, 0, s>>>(n, sc, q, part.partials, part.nblocks, r2, y);
  return hipGetLastError();
}
hipError_t mr_lanczos2(size_t n, double* sc, const MrSlots& q, const RedSlot& part,
                       const double* r2, double* y, hipStream_t s) {
  k_mr_lanczos2<<<vec_grid(n), kBlock, 0, s>>>(n, sc, q, part.partials, part.nblocks, r2, y);
  return hipGetLastError();
}
hipError_t mr_step(const RedSlot& rz, double* sc, const MrSlots& q, double* host_out,
                   hipStream_t s) {
  k_mr_step<<<1, kBlock, 0, s>>>(rz.partials, rz.nblocks, sc, q, host_out);
  return hipGetLastError();
}
hipError_t mr_update(size_t n, const double* sc, const MrSlots& q, double* v, const double* z,
                     double* w_out, const double* w_older, const double* w_newer, double* x,
                     hipStream_t s) {
  k_mr_update<<<vec_grid(n), kBlock, 0, s>>>(n, sc, q, v, z, w_out, w_older, w_newer, x);
  return hipGetLastError();
}

__global__ void k_sub_vec(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                          double* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}
hipError_t sub_vec(size_t n, const double* a, const double* b, double* out, hipStream_t s) {
  k_sub_vec<<<vec_grid(n), kBlock, 0, s>>>(n, a, b, out);
  return hipGetLastError();
}

// ============================================================================ design transfer
// Value at every fine vertex of the CG1 function given by nodal values on a coarser 'right'
// mesh of the same rectangle (dolfin interpolate of a CG1 function onto another mesh's CG1
// space): locate the coarse cell, then the triangle (A below the diagonal v0-v3, B above),
// and evaluate the linear interpolant there.
__global__ void k_interp_p1(int nxc, int nyc, double hxc, double hyc, int nxf, int nyf,
                            double hxf, double hyf, const double* __restrict__ rc,
                            double* __restrict__ rf) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)(nxf + 1) * (nyf + 1)) return;
  const int ix = (int)(v % (nxf + 1)), iy = (int)(v / (nxf + 1));
  const double x = 0.0 + hxf * ix, y = 0.0 + hyf * iy;
  const double fx = x / hxc, fy = y / hyc;
  const int cx = min(max((int)floor(fx), 0), nxc - 1), cy = min(max((int)floor(fy), 0), nyc - 1);
  const double s = fmin(fmax(fx - cx, 0.0), 1.0), t = fmin(fmax(fy - cy, 0.0), 1.0);
  const size_t v0 = (size_t)cy * (nxc + 1) + cx;
  const double r0 = rc[v0], r1 = rc[v0 + 1], r2 = rc[v0 + nxc + 1], r3 = rc[v0 + nxc + 2];
  rf[v] = s >= t ? r0 + s * (r1 - r0) + t * (r3 - r1) : r0 + t * (r2 - r0) + s * (r3 - r2);
}

hipError_t interp_p1(int nxc, int nyc, double hxc, double hyc, int nxf, int nyf, double hxf,
                     double hyf, const double* rc, double* rf, hipStream_t s) {
  const size_t n = (size_t)(nxf + 1) * (nyf + 1);
  k_interp_p1<<<grid1d(n), kBlock, 0, s>>>(nxc, nyc, hxc, hyc, nxf, nyf, hxf, hyf, rc, rf);
  return hipGetLastError();
}

// deterministic pseudo-random start vector in [-1, 1] (integer hash of the index)
__global__ void k_fill_hash(size_t n, double* __restrict__ v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    unsigned h = (unsigned)(i * 2654435761u) ^ 0x9e3779b9u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    v[i] = (double)h * (2.0 / 4294967295.0) - 1.0;
  }
}
hipError_t fill_hash(size_t n, double* v, hipStream_t s) {
  k_fill_hash<<<std::min<unsigned>(grid1d(n), 8192), kBlock, 0, s>>>(n, v);
  return hipGetLastError();
}

__global__ void k_scale_copy(size_t n, double c, const double* __restrict__ in,
                             double* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = c * in[i];
}
hipError_t scale_copy(size_t n, double c, const double* in, double* out, hipStream_t s) {
  k_scale_copy<<<std::min<unsigned>(grid1d(n), 8192), kBlock, 0, s>>>(n, c, in, out);
  return hipGetLastError();
}

}  // namespace tmx
