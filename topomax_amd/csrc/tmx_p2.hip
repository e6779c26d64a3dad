// P2 displacement element kernels (see tmx_p2.h for the lattice layout).
//
// The element operator is written in barycentric moments.  With l_k the barycentric
// coordinates of a triangle and d_k = grad l_k, every P2 basis gradient is linear in l:
// grad phi_a = sum_l l_l g_al with g_kl = (4 delta_kl - 1) d_k for vertex k and
// g_ml = 4 (delta_li d_j + delta_lj d_i) for the midpoint m of edge (i, j).  So
// grad u = sum_l l_l G_l, sigma(u) = sum_l l_l S_l, and the element force on node a is
// y_a = int r(xi) sigma(u) grad phi_a = sum_k T_k g_ak with T_k = sum_l W_lk S_l and
// W_lk = int_T r(xi_h) l_l l_k.  For integer p the moments are exact:
//   int_T (a.l)^p l^beta = 2|T| beta! p! / (|beta| + p + 2)! h_p(a with a_i repeated beta_i
//   more times),
// h_p the complete homogeneous symmetric polynomial (the Dirichlet moment identity).
// Every kernel gathers per node over the triangles that contain it (the operator kernels from
// per-triangle forces staged in LDS), so results are deterministic.
#include <cmath>

#include "tmx_device.h"
#include "tmx_p2.h"

namespace tmx {

namespace {

constexpr int kMaxP = 8;

// complete homogeneous symmetric polynomial of degree p (0 <= p <= kMaxP) in N values
template <int N>
__device__ __forceinline__ double hsym_multi(const double (&x)[N], int p) {
  double h[kMaxP + 1];
#pragma unroll
  for (int k = 0; k <= kMaxP; ++k) h[k] = k == 0 ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int k = 1; k <= kMaxP; ++k) h[k] += x[i] * h[k - 1];
  double r = h[0];
#pragma unroll
  for (int k = 1; k <= kMaxP; ++k)
    if (k == p) r = h[k];
  return r;
}

__device__ __forceinline__ double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// local vertex offsets (cells) of triangle t: A = (v0, v1, v3), B = (v0, v2, v3)
__device__ __forceinline__ void tri_vertex(int t, int k, int& ox, int& oy) {
  // A: (0,0) (1,0) (1,1); B: (0,0) (0,1) (1,1)
  ox = k == 0 ? 0 : (k == 2 ? 1 : (t == 0 ? 1 : 0));
  oy = k == 0 ? 0 : (k == 2 ? 1 : (t == 0 ? 0 : 1));
}

// lattice coordinates of the 6 local nodes of triangle t in cell (cx, cy)
__device__ __forceinline__ void tri_nodes(int t, int cx, int cy, int (&I)[6], int (&J)[6]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int ox, oy;
    tri_vertex(t, k, ox, oy);
    I[k] = 2 * (cx + ox);
    J[k] = 2 * (cy + oy);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // midpoint opposite vertex k
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    I[3 + k] = (I[i] + I[j]) / 2;
    J[3 + k] = (J[i] + J[j]) / 2;
  }
}

// barycentric gradients d_k of triangle t, given ihx = 1/hx and ihy = 1/hy (formed once per
// thread by the caller: the same values as dividing here, without a division per triangle)
template <typename T>
__device__ __forceinline__ void tri_grads(int t, T ihx, T ihy, T (&d)[3][2]) {
  if (t == 0) {
    d[0][0] = -ihx; d[0][1] = T(0);
    d[1][0] = ihx;  d[1][1] = -ihy;
    d[2][0] = T(0); d[2][1] = ihy;
  } else {
    d[0][0] = T(0); d[0][1] = -ihy;
    d[1][0] = -ihx; d[1][1] = ihy;
    d[2][0] = ihx;  d[2][1] = T(0);
  }
}

template <typename T>
__device__ __forceinline__ T wlk(const T (&W)[6], int l, int k) {
  const int a = l < k ? l : k, b = l < k ? k : l;
  // packed 00 01 02 11 12 22
  const int idx = a == 0 ? b : (a == 1 ? 2 + b : 5);
  return W[idx];
}

// strain-stress coefficients S_l (xx, xy, yy) of sigma(u) = sum_l l_l S_l, and eps_l
template <typename T>
__device__ __forceinline__ void stress_coeffs(const T (&d)[3][2], const T (&u)[6][2], T lam, T mu,
                                              T (&S)[3][3], T (&E)[3][3]) {
  T U0[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < 2; ++j) U0[c][j] += u[k][c] * d[k][j];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T G[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < 2; ++j) G[c][j] = T(4) * u[l][c] * d[l][j] - U0[c][j];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      if (jj == l) continue;
      const int m = 3 + (3 - l - jj);  // midpoint of edge (l, jj)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 2; ++j) G[c][j] += T(4) * u[m][c] * d[jj][j];
    }
    const T exx = G[0][0], eyy = G[1][1], exy = T(0.5) * (G[0][1] + G[1][0]);
    const T tr = exx + eyy;
    E[l][0] = exx;
    E[l][1] = exy;
    E[l][2] = eyy;
    S[l][0] = lam * tr + T(2) * mu * exx;
    S[l][1] = T(2) * mu * exy;
    S[l][2] = lam * tr + T(2) * mu * eyy;
  }
}

// force of one triangle on its local node a (0..5) for local displacements u
__device__ __forceinline__ void tri_force(const double (&W)[6], const double (&d)[3][2],
                                          const double (&u)[6][2], double lam, double mu, int a,
                                          double (&y)[2]) {
  double S[3][3], E[3][3];
  stress_coeffs(d, u, lam, mu, S, E);
  double T[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int e = 0; e < 3; ++e)
      T[k][e] = wlk(W, 0, k) * S[0][e] + wlk(W, 1, k) * S[1][e] + wlk(W, 2, k) * S[2][e];
  auto tv = [&](int k, int dk, double& ox, double& oy) {  // T_k d_dk
    ox = T[k][0] * d[dk][0] + T[k][1] * d[dk][1];
    oy = T[k][1] * d[dk][0] + T[k][2] * d[dk][1];
  };
  if (a < 3) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double ox, oy;
      tv(k, a, ox, oy);
      const double c = k == a ? 3.0 : -1.0;  // 4 delta - 1
      sx += c * ox;
      sy += c * oy;
    }
    y[0] = sx;
    y[1] = sy;
  } else {
    const int k = a - 3, i = (k + 1) % 3, j = (k + 2) % 3;
    double ax, ay, bx, by;
    tv(i, j, ax, ay);
    tv(j, i, bx, by);
    y[0] = 4.0 * (ax + bx);
    y[1] = 4.0 * (ay + by);
  }
}

// forces of one triangle on all six local nodes (the stress coefficients formed once)
template <typename R>
__device__ __forceinline__ void tri_forces_all(const R (&W)[6], const R (&d)[3][2],
                                               const R (&u)[6][2], R lam, R mu, R (&y)[6][2]) {
  R S[3][3], E[3][3];
  stress_coeffs(d, u, lam, mu, S, E);
  R T[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int e = 0; e < 3; ++e)
      T[k][e] = wlk(W, 0, k) * S[0][e] + wlk(W, 1, k) * S[1][e] + wlk(W, 2, k) * S[2][e];
  R tv[3][3][2];  // T_k d_j
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      tv[k][j][0] = T[k][0] * d[j][0] + T[k][1] * d[j][1];
      tv[k][j][1] = T[k][1] * d[j][0] + T[k][2] * d[j][1];
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R sx = R(0), sy = R(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const R c = k == a ? R(3) : R(-1);  // 4 delta - 1
      sx += c * tv[k][a][0];
      sy += c * tv[k][a][1];
    }
    y[a][0] = sx;
    y[a][1] = sy;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    y[3 + k][0] = R(4) * (tv[i][j][0] + tv[j][i][0]);
    y[3 + k][1] = R(4) * (tv[i][j][1] + tv[j][i][1]);
  }
}

__device__ __forceinline__ bool p2_fixed(unsigned fixed, int I, int J, int nX, int nY) {
  return ((fixed & kLeft) && I == 0) || ((fixed & kRight) && I == nX - 1) ||
         ((fixed & kBottom) && J == 0) || ((fixed & kTop) && J == nY - 1);
}

template <typename T>
__device__ __forceinline__ void load_W(const T* __restrict__ W, int nx, int cx, int cy, int t,
                                       T (&w)[6]) {
  const T* src = W + 12 * ((size_t)cy * nx + cx) + 6 * t;
#pragma unroll
  for (int e = 0; e < 6; ++e) w[e] = src[e];
}

// visit every triangle containing lattice node (I, J): f(t, cx, cy, local index, nodes I, J)
template <class F>
__device__ __forceinline__ void for_each_triangle(const P2Geom& g, int I, int J, F&& f) {
  const int cx0 = (I & 1) ? (I - 1) / 2 : I / 2 - 1, cx1 = (I & 1) ? (I - 1) / 2 : I / 2;
  const int cy0 = (J & 1) ? (J - 1) / 2 : J / 2 - 1, cy1 = (J & 1) ? (J - 1) / 2 : J / 2;
  for (int cy = cy0; cy <= cy1; ++cy) {
    if (cy < 0 || cy >= g.ny) continue;
    for (int cx = cx0; cx <= cx1; ++cx) {
      if (cx < 0 || cx >= g.nx) continue;
      for (int t = 0; t < 2; ++t) {
        int NI[6], NJ[6];
        tri_nodes(t, cx, cy, NI, NJ);
        int a = -1;
#pragma unroll
        for (int q = 0; q < 6; ++q)
          if (NI[q] == I && NJ[q] == J) a = q;
        if (a >= 0) f(t, cx, cy, a, NI, NJ);
      }
    }
  }
}

__global__ void k_p2_weights(P2Geom g, const double* __restrict__ xi, int p, double m,
                             int positive, double* __restrict__ W) {
  const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= (size_t)g.nx * g.ny) return;
  const int cx = (int)(cell % g.nx), cy = (int)(cell / g.nx);
  const double area = 0.5 * g.hx * g.hy;
  const double cp = (1.0 - m) * 2.0 * area * factorial(p) / factorial(p + 4);
  for (int t = 0; t < 2; ++t) {
    double a[3];
    for (int k = 0; k < 3; ++k) {
      int ox, oy;
      tri_vertex(t, k, ox, oy);
      const double v = xi[(size_t)(cy + oy) * (g.nx + 1) + cx + ox];
      a[k] = positive ? fmax(v, 0.0) : v;
    }
    double* out = W + 12 * cell + 6 * t;
    int e = 0;
    for (int l = 0; l < 3; ++l)
      for (int k = l; k < 3; ++k, ++e) {
        const double beta_f = l == k ? 2.0 : 1.0;
        const double xs[5] = {a[0], a[1], a[2], a[l], a[k]};
        out[e] = m * area * beta_f / 12.0 + cp * beta_f * hsym_multi<5>(xs, p);
      }
  }
}

// Tiled operator: a workgroup owns the lattice nodes of kP2TileC x kP2TileC cells.  It stages
// the (masked) displacement of the cells touching those nodes (one more cell row and column
// below and to the left) in LDS, forms each of their triangles' forces on all six local nodes
// once (one thread per triangle), and then every owned node sums its triangles' forces from
// LDS (cells bottom to top, left to right, triangle A before B).  Forming the forces per node
// instead would form each triangle's stresses once for each of its six nodes.
//   MODE 0: y = K u (fixed nodes: y = u);  1: y = b - K u (fixed: b - u);
//   2: y = x + omega D^-1 (b - K x) (x null: from zero; fixed nodes 0);
//   3: MINRES's operator step y = K u - (*mb / *mo) b (0 when *mo is 0), with the partial of
//      u.y of the workgroup's owned nodes in part[workgroup] (FP64).
// T: FP64, or FP32 for the mixed scheme's preconditioner (weights, diagonal blocks and vectors
// in float, the diagonal blocks inverted by the hardware reciprocal).  TB / TY: the types of b
// and y (a double right-hand side read, or a double result written, by the float cycle's
// first and last sweeps); with bcopy, b rounded to T is also written there.
constexpr int kP2TileC = 8;                  // owned cells per side
constexpr int kP2TileThreads = 256;
constexpr int kP2ExtC = kP2TileC + 1;        // cells whose triangles reach the owned nodes
constexpr int kP2ExtN = 2 * kP2TileC + 3;    // lattice nodes of those cells per side

// The triangles that hold a lattice node (I, J), by its parity class (I & 1) + 2 (J & 1): the
// offset (dx, dy) of their cell from cell (I >> 1, J >> 1), the triangle t and the node's local
// index q, listed cells bottom to top, left to right, triangle A before B.  A vertex lies in six
// triangles, an edge midpoint in two.  Derived at compile time from the local node layout of
// tri_nodes (translation invariant), so a node's gather is a fixed list of LDS reads.
struct P2Hold {
  int dx, dy, t, q;
};
struct P2HoldSet {
  int n;
  P2Hold e[6];
};
constexpr P2HoldSet p2_holds(int cls) {
  P2HoldSet hs{0, {}};
  const int pi = cls & 1, pj = cls >> 1;
  for (int dy = -1; dy <= 0; ++dy)
    for (int dx = -1; dx <= 0; ++dx)
      for (int t = 0; t < 2; ++t) {
        int vi[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {  // (tri_vertex, relative to lattice node (2 ci, 2 cj))
          const int ox = k == 0 ? 0 : (k == 2 ? 1 : (t == 0 ? 1 : 0));
          const int oy = k == 0 ? 0 : (k == 2 ? 1 : (t == 0 ? 0 : 1));
          vi[k] = 2 * (dx + ox);
          vj[k] = 2 * (dy + oy);
        }
        for (int q = 0; q < 6; ++q) {
          int ni = 0, nj = 0;
          if (q < 3) {
            ni = vi[q];
            nj = vj[q];
          } else {  // midpoint opposite vertex q - 3
            const int a = (q - 3 + 1) % 3, b = (q - 3 + 2) % 3;
            ni = (vi[a] + vi[b]) / 2;
            nj = (vj[a] + vj[b]) / 2;
          }
          if (ni == pi && nj == pj && hs.n < 6) hs.e[hs.n++] = P2Hold{dx, dy, t, q};
        }
      }
  return hs;
}
static_assert(p2_holds(0).n == 6 && p2_holds(1).n == 2 && p2_holds(2).n == 2 &&
                  p2_holds(3).n == 2,
              "a vertex lies in six triangles, a midpoint in two");

// the sum over the triangles holding owned node (I, J) of their forces on it (LDS fs)
template <int CLS, typename T>
__device__ __forceinline__ void p2_gather(const T (&fs)[kP2ExtC * kP2ExtC][2][6][2],
                                          const P2Geom& g, int I, int J, int X0, int Y0, T& f0,
                                          T& f1) {
  constexpr P2HoldSet H = p2_holds(CLS);
  const int ci = I >> 1, cj = J >> 1;
#pragma unroll
  for (int k = 0; k < H.n; ++k) {
    const int cx = ci + H.e[k].dx, cy = cj + H.e[k].dy;
    if (cx < 0 || cy < 0 || cx >= g.nx || cy >= g.ny) continue;
    const int c = (cy - Y0 + 1) * kP2ExtC + (cx - X0 + 1);
    f0 += fs[c][H.e[k].t][H.e[k].q][0];
    f1 += fs[c][H.e[k].t][H.e[k].q][1];
  }
}

template <int MODE, typename T, typename TB = T, typename TY = T>
__global__ __launch_bounds__(kP2TileThreads) void k_p2_tile(P2Geom g, const T* __restrict__ W,
                                                            const T* __restrict__ D,
                                                            const T* __restrict__ x,
                                                            const TB* __restrict__ b,
                                                            TY* __restrict__ y, T omega,
                                                            T* __restrict__ bcopy = nullptr,
                                                            const double* __restrict__ mb = nullptr,
                                                            const double* __restrict__ mo = nullptr,
                                                            double* __restrict__ part = nullptr) {
  __shared__ T us[kP2ExtN * kP2ExtN][2];
  __shared__ T fs[kP2ExtC * kP2ExtC][2][6][2];
  const int nX = 2 * g.nx + 1, nY = 2 * g.ny + 1;
  const int ntx = (g.nx + kP2TileC - 1) / kP2TileC;
  const int X0 = (int)(blockIdx.x % (unsigned)ntx) * kP2TileC;
  const int Y0 = (int)(blockIdx.x / (unsigned)ntx) * kP2TileC;
  const int LI0 = 2 * X0 - 2, LJ0 = 2 * Y0 - 2;  // lattice origin of the staged region
  const bool forces = MODE != 2 || x != nullptr;
  if (forces) {
    for (int k = (int)threadIdx.x; k < kP2ExtN * kP2ExtN; k += kP2TileThreads) {
      const int lj = k / kP2ExtN, li = k - lj * kP2ExtN;
      const int I = LI0 + li, J = LJ0 + lj;
      T u0 = T(0), u1 = T(0);
      if (I >= 0 && J >= 0 && I < nX && J < nY && !p2_fixed(g.fixed, I, J, nX, nY)) {
        const size_t n = (size_t)J * nX + I;
        u0 = x[2 * n];
        u1 = x[2 * n + 1];
      }
      us[k][0] = u0;
      us[k][1] = u1;
    }
    __syncthreads();
    for (int k = (int)threadIdx.x; k < kP2ExtC * kP2ExtC * 2; k += kP2TileThreads) {
      const int t = k & 1, c = k >> 1;
      const int lcy = c / kP2ExtC, lcx = c - lcy * kP2ExtC;
      const int cx = X0 - 1 + lcx, cy = Y0 - 1 + lcy;
      if (cx < 0 || cy < 0 || cx >= g.nx || cy >= g.ny) continue;  // (no owned node reads it)
      T w[6], d[3][2], ul[6][2], f[6][2];
      load_W(W, g.nx, cx, cy, t, w);
      tri_grads(t, T(1) / (T)g.hx, T(1) / (T)g.hy, d);
      int NI[6], NJ[6];
      tri_nodes(t, cx, cy, NI, NJ);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int li = (NI[q] - LI0) + (NJ[q] - LJ0) * kP2ExtN;
        ul[q][0] = us[li][0];
        ul[q][1] = us[li][1];
      }
      tri_forces_all(w, d, ul, (T)g.lam, (T)g.mu, f);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        fs[c][t][q][0] = f[q][0];
        fs[c][t][q][1] = f[q][1];
      }
    }
    __syncthreads();
  }
  // owned lattice nodes [2 X0, IE) x [2 Y0, JE); the last tile of a row or column also owns
  // the lattice's last line
  const int IE = X0 + kP2TileC >= g.nx ? nX : 2 * (X0 + kP2TileC);
  const int JE = Y0 + kP2TileC >= g.ny ? nY : 2 * (Y0 + kP2TileC);
  const int OW = IE - 2 * X0, OH = JE - 2 * Y0;
  double dot_part = 0.0;  // (MODE 3)
  double cl = 0.0;
  if constexpr (MODE == 3) cl = *mo != 0.0 ? *mb / *mo : 0.0;
  for (int k = (int)threadIdx.x; k < OW * OH; k += kP2TileThreads) {
    const int oj = k / OW, oi = k - oj * OW;
    const int I = 2 * X0 + oi, J = 2 * Y0 + oj;
    const size_t n = (size_t)J * nX + I;
    const bool fx = p2_fixed(g.fixed, I, J, nX, nY);
    T f0 = T(0), f1 = T(0);
    if (forces && !fx) {
      switch ((I & 1) + 2 * (J & 1)) {
        case 0: p2_gather<0>(fs, g, I, J, X0, Y0, f0, f1); break;
        case 1: p2_gather<1>(fs, g, I, J, X0, Y0, f0, f1); break;
        case 2: p2_gather<2>(fs, g, I, J, X0, Y0, f0, f1); break;
        default: p2_gather<3>(fs, g, I, J, X0, Y0, f0, f1); break;
      }
    }
    if constexpr (MODE == 0) {
      y[2 * n] = fx ? x[2 * n] : f0;
      y[2 * n + 1] = fx ? x[2 * n + 1] : f1;
    } else if constexpr (MODE == 3) {
      const T u0 = x[2 * n], u1 = x[2 * n + 1];
      const T y0 = (fx ? u0 : f0) - cl * b[2 * n], y1 = (fx ? u1 : f1) - cl * b[2 * n + 1];
      y[2 * n] = y0;
      y[2 * n + 1] = y1;
      dot_part += u0 * y0 + u1 * y1;
    } else if constexpr (MODE == 1) {
      y[2 * n] = b[2 * n] - (fx ? x[2 * n] : f0);
      y[2 * n + 1] = b[2 * n + 1] - (fx ? x[2 * n + 1] : f1);
    } else {
      const T b0 = (T)b[2 * n], b1 = (T)b[2 * n + 1];
      if (bcopy) {
        bcopy[2 * n] = b0;
        bcopy[2 * n + 1] = b1;
      }
      if (fx) {
        y[2 * n] = y[2 * n + 1] = TY(0);
      } else {
        const T rx = b0 - f0, ry = b1 - f1;
        const T a = D[3 * n], c = D[3 * n + 1], e = D[3 * n + 2];
        const T inv = recip(a * e - c * c);
        const T x0 = x ? x[2 * n] : T(0), x1 = x ? x[2 * n + 1] : T(0);
        y[2 * n] = (TY)(x0 + omega * (e * rx - c * ry) * inv);
        y[2 * n + 1] = (TY)(x1 + omega * (a * ry - c * rx) * inv);
      }
    }
  }
  if constexpr (MODE == 3) {
    const double vv[1] = {dot_part};
    double* const o[1] = {part};
    block_partials<1, kP2TileThreads>(vv, o, blockIdx.x);
  }
}

__global__ void k_p2_diag(P2Geom g, const double* __restrict__ W, double* __restrict__ D) {
  const int nX = 2 * g.nx + 1;
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (size_t)nX * (2 * g.ny + 1)) return;
  const int I = (int)(n % nX), J = (int)(n / nX);
  double dxx = 0.0, dxy = 0.0, dyy = 0.0;
  const double ihx = 1.0 / g.hx, ihy = 1.0 / g.hy;
  for_each_triangle(g, I, J, [&](int t, int cx, int cy, int a, const int (&)[6],
                                 const int (&)[6]) {
    double w[6], d[3][2], ul[6][2], y[2];
    load_W(W, g.nx, cx, cy, t, w);
    tri_grads(t, ihx, ihy, d);
    for (int c = 0; c < 2; ++c) {
#pragma unroll
      for (int q = 0; q < 6; ++q) ul[q][0] = ul[q][1] = 0.0;
      ul[a][c] = 1.0;
      tri_force(w, d, ul, g.lam, g.mu, a, y);
      if (c == 0) {
        dxx += y[0];
        dxy += y[1];
      } else {
        dyy += y[1];
      }
    }
  });
  D[3 * n] = dxx;
  D[3 * n + 1] = dxy;
  D[3 * n + 2] = dyy;
}

// r1 at P1 vertex (i, j) = r2(2i, 2j) + 1/2 (the six adjacent midpoints)
template <typename T>
__global__ void k_p2_restrict(P2Geom g, const T* __restrict__ r2, T* __restrict__ r1) {
  const int nx = g.nx, ny = g.ny, nX = 2 * nx + 1;
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)(nx + 1) * (ny + 1)) return;
  const int i = (int)(v % (nx + 1)), j = (int)(v / (nx + 1));
  T sx = T(0), sy = T(0);
  if (!node_fixed(g.fixed, i, j, nx, ny)) {
    const int I = 2 * i, J = 2 * j;
    sx = r2[2 * ((size_t)J * nX + I)];
    sy = r2[2 * ((size_t)J * nX + I) + 1];
    const int off[6][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}};
    for (int q = 0; q < 6; ++q) {
      const int mI = I + off[q][0], mJ = J + off[q][1];
      if (mI < 0 || mJ < 0 || mI > 2 * nx || mJ > 2 * ny) continue;
      const size_t m = (size_t)mJ * nX + mI;
      sx += T(0.5) * r2[2 * m];
      sy += T(0.5) * r2[2 * m + 1];
    }
  }
  r1[2 * v] = sx;
  r1[2 * v + 1] = sy;
}

template <typename T>
__global__ void k_p2_prolong_add(P2Geom g, const T* __restrict__ e1, T* __restrict__ x2) {
  const int nx = g.nx, nX = 2 * nx + 1, nY = 2 * g.ny + 1;
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (size_t)nX * nY) return;
  const int I = (int)(n % nX), J = (int)(n / nX);
  if (p2_fixed(g.fixed, I, J, nX, nY)) return;
  // the two P1 vertices whose average this node is (the same vertex twice at a vertex)
  const int i0 = I / 2, j0 = J / 2;
  const int i1 = (I + 1) / 2, j1 = (J + 1) / 2;
  const size_t a = (size_t)j0 * (nx + 1) + i0, b = (size_t)j1 * (nx + 1) + i1;
  x2[2 * n] += T(0.5) * (e1[2 * a] + e1[2 * b]);
  x2[2 * n + 1] += T(0.5) * (e1[2 * a + 1] + e1[2 * b + 1]);
}

// sensitivity at P1 vertex v: sum over its triangles of
//   -p (1-m) sum_lk Q_lk int (a.l)^(p-1) l_l l_k l_i,  Q_lk = S_l : eps_k
__global__ void k_p2_sensitivity(P2Geom g, const double* __restrict__ xi,
                                 const double* __restrict__ u, int p, double m,
                                 double* __restrict__ out) {
  const int nx = g.nx, ny = g.ny, nX = 2 * nx + 1;
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)(nx + 1) * (ny + 1)) return;
  const int iv = (int)(v % (nx + 1)), jv = (int)(v / (nx + 1));
  const double area = 0.5 * g.hx * g.hy;
  const double coef = -(double)p * (1.0 - m) * 2.0 * area * factorial(p - 1) / factorial(p + 4);
  double acc = 0.0;
  const double ihx = 1.0 / g.hx, ihy = 1.0 / g.hy;
  for_each_triangle(g, 2 * iv, 2 * jv, [&](int t, int cx, int cy, int li, const int (&NI)[6],
                                           const int (&NJ)[6]) {
    double d[3][2], ul[6][2], a[3];
    tri_grads(t, ihx, ihy, d);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const size_t n = (size_t)NJ[q] * nX + NI[q];
      ul[q][0] = u[2 * n];
      ul[q][1] = u[2 * n + 1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = xi[(size_t)(NJ[k] / 2) * (nx + 1) + NI[k] / 2];
    double S[3][3], E[3][3];
    stress_coeffs(d, ul, g.lam, g.mu, S, E);
    double s = 0.0;
    for (int l = 0; l < 3; ++l)
      for (int k = 0; k < 3; ++k) {
        const double Q = S[l][0] * E[k][0] + 2.0 * S[l][1] * E[k][1] + S[l][2] * E[k][2];
        const int c0 = (l == 0) + (k == 0) + (li == 0), c1 = (l == 1) + (k == 1) + (li == 1),
                  c2 = (l == 2) + (k == 2) + (li == 2);
        const double bf = factorial(c0) * factorial(c1) * factorial(c2);
        const double xs[6] = {a[0], a[1], a[2], a[l], a[k], a[li]};
        s += Q * bf * hsym_multi<6>(xs, p - 1);
      }
    acc += coef * s;
  });
  out[v] = acc;
}

// ---------------------------------------------------------------- loads
__device__ __forceinline__ double lat_x(const LoadDesc& d, int I) {
  return (I & 1) ? mid(vx(d, I / 2), vx(d, I / 2 + 1)) : vx(d, I / 2);
}
__device__ __forceinline__ double lat_y(const LoadDesc& d, int J) {
  return (J & 1) ? mid(vy(d, J / 2), vy(d, J / 2 + 1)) : vy(d, J / 2);
}

// P2 mass matrix * 180 / |T| (local order: vertices, then midpoints opposite vertex 0, 1, 2)
__constant__ double c_p2_mass[6][6] = {{6, -1, -1, -4, 0, 0},   {-1, 6, -1, 0, -4, 0},
                                       {-1, -1, 6, 0, 0, -4},   {-4, 0, 0, 32, 16, 16},
                                       {0, -4, 0, 16, 32, 16},  {0, 0, -4, 16, 16, 32}};

__global__ void k_p2_load(LoadDesc d, double* __restrict__ F) {
  const int nx = d.nx, ny = d.ny, nX = 2 * nx + 1, nY = 2 * ny + 1;
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (size_t)nX * nY) return;
  const int I = (int)(n % nX), J = (int)(n / nX);
  double sx = 0.0, sy = 0.0;
  if (d.has_body) {
    P2Geom g{nx, ny, d.ab, d.cd, 0.0, 0.0, 0u};
    for_each_triangle(g, I, J, [&](int t, int cx, int cy, int a, const int (&NI)[6],
                                   const int (&NJ)[6]) {
      int px[3], py[3];
      for (int k = 0; k < 3; ++k) {
        px[k] = NI[k] / 2;
        py[k] = NJ[k] / 2;
      }
      const double area = 0.5 * fabs((vx(d, px[1]) - vx(d, px[0])) * (vy(d, py[2]) - vy(d, py[0])) -
                                     (vx(d, px[2]) - vx(d, px[0])) * (vy(d, py[1]) - vy(d, py[0])));
      for (int q = 0; q < 6; ++q) {
        double fx, fy;
        body_value(d, lat_x(d, NI[q]), lat_y(d, NJ[q]), fx, fy);
        const double w = c_p2_mass[a][q] * area / 180.0;
        sx += w * fx;
        sy += w * fy;
      }
    });
  }
  if (d.ntr > 0 && (I == 0 || I == nX - 1 || J == 0 || J == nY - 1)) {
    // boundary edges (lattice end nodes e0, e1 even; midpoint between) containing this node
    auto edge = [&](int I0, int J0, int I1, int J1) {
      const int IM = (I0 + I1) / 2, JM = (J0 + J1) / 2;
      double t0x, t0y, t1x, t1y, tmx, tmy;
      traction_value(d, lat_x(d, I0), lat_y(d, J0), t0x, t0y);
      traction_value(d, lat_x(d, I1), lat_y(d, J1), t1x, t1y);
      traction_value(d, lat_x(d, IM), lat_y(d, JM), tmx, tmy);
      const double len = (I0 != I1) ? fabs(lat_x(d, I1) - lat_x(d, I0))
                                    : fabs(lat_y(d, J1) - lat_y(d, J0));
      const double s = len / 30.0;
      if (I == I0 && J == J0) {
        sx += s * (4 * t0x - t1x + 2 * tmx);
        sy += s * (4 * t0y - t1y + 2 * tmy);
      } else if (I == I1 && J == J1) {
        sx += s * (-t0x + 4 * t1x + 2 * tmx);
        sy += s * (-t0y + 4 * t1y + 2 * tmy);
      } else {
        sx += s * (2 * t0x + 2 * t1x + 16 * tmx);
        sy += s * (2 * t0y + 2 * t1y + 16 * tmy);
      }
    };
    if (J == 0 || J == nY - 1) {  // horizontal boundary: edges [2k, 2k+2] along I
      if (I & 1) {
        edge(I - 1, J, I + 1, J);
      } else {
        if (I > 0) edge(I - 2, J, I, J);
        if (I < nX - 1) edge(I, J, I + 2, J);
      }
    }
    if (I == 0 || I == nX - 1) {  // vertical boundary
      if (J & 1) {
        edge(I, J - 1, I, J + 1);
      } else {
        if (J > 0) edge(I, J - 2, I, J);
        if (J < nY - 1) edge(I, J, I, J + 2);
      }
    }
  }
  F[2 * n] = sx;
  F[2 * n + 1] = sy;
}

__global__ void k_p2_mask(P2Geom g, double* __restrict__ v) {
  const int nX = 2 * g.nx + 1, nY = 2 * g.ny + 1;
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (size_t)nX * nY) return;
  if (p2_fixed(g.fixed, (int)(n % nX), (int)(n / nX), nX, nY)) v[2 * n] = v[2 * n + 1] = 0.0;
}

unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); }

}  // namespace

hipError_t p2_weights(const P2Geom& g, const double* xi, int p, double m, bool positive_part,
                      double* W, hipStream_t s) {
  if (p < 1 || p > kMaxP) return hipErrorInvalidValue;
  k_p2_weights<<<grid_for((size_t)g.nx * g.ny), 256, 0, s>>>(g, xi, p, m, positive_part ? 1 : 0, W);
  return hipGetLastError();
}

static unsigned p2_tiles(const P2Geom& g) {
  return (unsigned)(((g.nx + kP2TileC - 1) / kP2TileC) * ((g.ny + kP2TileC - 1) / kP2TileC));
}
bool p2_apply_lanczos_ok(const P2Geom& g) { return p2_tiles(g) <= (unsigned)kMaxPartials; }

hipError_t p2_apply(const P2Geom& g, const double* W, const double* u, double* y, hipStream_t s) {
  if (!u || !y) return hipErrorInvalidValue;
  k_p2_tile<0, double, double, double><<<p2_tiles(g), kP2TileThreads, 0, s>>>(g, W, nullptr, u, nullptr, y, 0.0);
  return hipGetLastError();
}

hipError_t p2_apply_lanczos(const P2Geom& g, const double* W, const double* v, const double* r1,
                            double* y, const double* beta, const double* oldb, double* part,
                            int* nparts, hipStream_t s) {
  if (!v || !r1 || !y || !part || p2_tiles(g) > (unsigned)kMaxPartials) return hipErrorInvalidValue;
  *nparts = (int)p2_tiles(g);
  k_p2_tile<3, double, double, double><<<p2_tiles(g), kP2TileThreads, 0, s>>>(
      g, W, nullptr, v, r1, y, 0.0, nullptr, beta, oldb, part);
  return hipGetLastError();
}

template <typename T>
hipError_t p2_residual(const P2Geom& g, const T* W, const T* x, const T* b, T* r, hipStream_t s) {
  if (!x || !b || !r) return hipErrorInvalidValue;
  k_p2_tile<1, T><<<p2_tiles(g), kP2TileThreads, 0, s>>>(g, W, nullptr, x, b, r, T(0));
  return hipGetLastError();
}
template hipError_t p2_residual<double>(const P2Geom&, const double*, const double*,
                                        const double*, double*, hipStream_t);
template hipError_t p2_residual<float>(const P2Geom&, const float*, const float*, const float*,
                                       float*, hipStream_t);

hipError_t p2_diag(const P2Geom& g, const double* W, double* D, hipStream_t s) {
  k_p2_diag<<<grid_for(g.nodes()), 256, 0, s>>>(g, W, D);
  return hipGetLastError();
}

template <typename T>
hipError_t p2_smooth(const P2Geom& g, const T* W, const T* D, const T* x, const T* b, T* y,
                     double omega, hipStream_t s) {
  if (!D || !b || !y) return hipErrorInvalidValue;
  k_p2_tile<2, T><<<p2_tiles(g), kP2TileThreads, 0, s>>>(g, W, D, x, b, y, (T)omega);
  return hipGetLastError();
}

hipError_t p2_smooth_first32(const P2Geom& g, const float* W, const float* D, const double* b,
                             float* b32, float* y, double omega, hipStream_t s) {
  if (!D || !b || !b32 || !y) return hipErrorInvalidValue;
  k_p2_tile<2, float, double, float><<<p2_tiles(g), kP2TileThreads, 0, s>>>(
      g, W, D, nullptr, b, y, (float)omega, b32);
  return hipGetLastError();
}

hipError_t p2_smooth_last32(const P2Geom& g, const float* W, const float* D, const float* x,
                            const float* b, double* y, double omega, hipStream_t s) {
  if (!D || !x || !b || !y) return hipErrorInvalidValue;
  k_p2_tile<2, float, float, double><<<p2_tiles(g), kP2TileThreads, 0, s>>>(
      g, W, D, x, b, y, (float)omega);
  return hipGetLastError();
}
template hipError_t p2_smooth<double>(const P2Geom&, const double*, const double*, const double*,
                                      const double*, double*, double, hipStream_t);
template hipError_t p2_smooth<float>(const P2Geom&, const float*, const float*, const float*,
                                     const float*, float*, double, hipStream_t);

template <typename T>
hipError_t p2_restrict(const P2Geom& g, const T* r2, T* r1, hipStream_t s) {
  k_p2_restrict<T><<<grid_for((size_t)(g.nx + 1) * (g.ny + 1)), 256, 0, s>>>(g, r2, r1);
  return hipGetLastError();
}
template hipError_t p2_restrict<double>(const P2Geom&, const double*, double*, hipStream_t);
template hipError_t p2_restrict<float>(const P2Geom&, const float*, float*, hipStream_t);

template <typename T>
hipError_t p2_prolong_add(const P2Geom& g, const T* e1, T* x2, hipStream_t s) {
  k_p2_prolong_add<T><<<grid_for(g.nodes()), 256, 0, s>>>(g, e1, x2);
  return hipGetLastError();
}
template hipError_t p2_prolong_add<double>(const P2Geom&, const double*, double*, hipStream_t);
template hipError_t p2_prolong_add<float>(const P2Geom&, const float*, float*, hipStream_t);

hipError_t p2_assemble_load(const LoadDesc& d, double* F, hipStream_t s) {
  k_p2_load<<<grid_for((size_t)(2 * d.nx + 1) * (2 * d.ny + 1)), 256, 0, s>>>(d, F);
  return hipGetLastError();
}

hipError_t p2_mask_fixed(const P2Geom& g, double* v, hipStream_t s) {
  k_p2_mask<<<grid_for(g.nodes()), 256, 0, s>>>(g, v);
  return hipGetLastError();
}

hipError_t p2_sensitivity(const P2Geom& g, const double* xi, const double* u, int p, double m,
                          double* b, hipStream_t s) {
  if (p < 1 || p > kMaxP) return hipErrorInvalidValue;
  k_p2_sensitivity<<<grid_for((size_t)(g.nx + 1) * (g.ny + 1)), 256, 0, s>>>(g, xi, u, p, m, b);
  return hipGetLastError();
}

}  // namespace tmx
