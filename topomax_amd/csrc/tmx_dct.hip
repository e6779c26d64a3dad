// DCT-I / tridiagonal solver of the filter's tensor-lumped operator (see tmx_dct.h).
//
// Data layout: vectors are node-major rows, v = iy (nx + 1) + ix, so a y-column is strided and a
// mode row (fixed y-mode, all x) is contiguous.  The transforms run on y-columns held in LDS
// (k_dct1_cols: a group of 2 or 4 adjacent columns per workgroup, every row read and written as
// one short contiguous segment); the tridiagonal solves run on mode rows (k_dct_tri).
#include "tmx_dct.h"

#include <cmath>
#include <vector>

#include "tmx_device.h"

namespace tmx {

bool fft_plan(int L, FftPlan& plan) {
  plan = FftPlan();
  plan.L = L;
  if (L < 2) return false;
  int m = L;
  auto push = [&](int r) {
    if (plan.nstages >= FftPlan::kMaxStages) return false;
    plan.radix[plan.nstages++] = r;
    m /= r;
    return true;
  };
  while (m % 4 == 0)
    if (!push(4)) return false;
  while (m % 2 == 0)
    if (!push(2)) return false;
  for (int r : {3, 5, 7})
    while (m % r == 0)
      if (!push(r)) return false;
  return m == 1;
}

// ---------------------------------------------------------------- complex FFT in LDS
// Threads per FFT workgroup by its column count: four columns (two complex sequences) take
// 512 threads, two take 256.  FP32 stages the twiddle table in LDS; FP64 reads it from global
// memory and always takes two columns, as its butterflies take twice the registers and its
// columns twice the LDS (four columns on 512 threads with the table in LDS measured 164 us per
// FP64 application at N=1000, against 103 us).
template <int C>
constexpr int kFftThreads = C == 4 ? 512 : 256;
template <typename T>
constexpr bool kTwiddlesInLds = sizeof(T) == 4;
constexpr int kFftMaxPoints = 4096;  // points of the complex sequences of one workgroup

template <typename T>
__device__ __forceinline__ V2<T> cmul(V2<T> a, V2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// cos and sin of 2 pi q / R, q = 0 .. R-1, for the odd radices
template <int R>
struct RootsOdd;
template <>
struct RootsOdd<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct RootsOdd<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <>
struct RootsOdd<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};

// forward DFT of R points in place: y_k = sum_j v_j exp(-2 pi i j k / R)
template <typename T, int R>
__device__ __forceinline__ void dft(V2<T> (&v)[R]) {
  if constexpr (R == 2) {
    const V2<T> a = v[0];
    v[0] = a + v[1];
    v[1] = a - v[1];
  } else if constexpr (R == 4) {
    const V2<T> s02 = v[0] + v[2], d02 = v[0] - v[2], s13 = v[1] + v[3], d13 = v[1] - v[3];
    v[0] = s02 + s13;
    v[2] = s02 - s13;
    v[1] = {d02.x + d13.y, d02.y - d13.x};  // d02 - i d13
    v[3] = {d02.x - d13.y, d02.y + d13.x};  // d02 + i d13
  } else {
    // odd R: pair v_k with v_{R-k}; y_m = M_m - i N_m, y_{R-m} = M_m + i N_m with
    // M_m = v_0 + sum_k cos(2 pi m k / R) (v_k + v_{R-k}), N_m = sum_k sin(2 pi m k / R) (v_k - v_{R-k})
    constexpr int H = R / 2;
    V2<T> t[H + 1], u[H + 1];
    V2<T> y0 = v[0];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      t[k] = v[k] + v[R - k];
      u[k] = v[k] - v[R - k];
      y0 = y0 + t[k];
    }
    V2<T> out[R];
    out[0] = y0;
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      V2<T> M = v[0], N = {T(0), T(0)};
#pragma unroll
      for (int k = 1; k <= H; ++k) {
        const T c = (T)RootsOdd<R>::c[(m * k) % R], sn = (T)RootsOdd<R>::s[(m * k) % R];
        M = {M.x + c * t[k].x, M.y + c * t[k].y};
        N = {N.x + sn * u[k].x, N.y + sn * u[k].y};
      }
      out[m] = {M.x + N.y, M.y - N.x};
      out[R - m] = {M.x - N.y, M.y + N.x};
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
}

// q = floor(j / d) for 0 <= j, d <= 8192 from a float reciprocal and one correction each way
__device__ __forceinline__ int div_fft(int j, int d, float dinv) {
  int q = (int)((float)j * dinv);
  if ((q + 1) * d <= j) ++q;
  if (q * d > j) --q;
  return q;
}

// One Stockham radix-R stage over nseq sequences of L points in LDS, Ns = product of the
// radices of the stages before it: butterfly j of a sequence takes the points j + r L/R,
// twiddles them by exp(-2 pi i (j mod Ns) r / (Ns R)), transforms them and writes them to
// (j div Ns) Ns R + (j mod Ns) + r Ns.  Every thread first reads and transforms all its
// butterflies into registers, then (after a barrier) writes them, so the stage runs in place.
template <typename T, int R, int NT>
__device__ __forceinline__ void fft_stage(V2<T>* buf, int L, int nseq, int Ns,
                                          const V2<T>* __restrict__ tw) {
  constexpr int MAXB = (kFftMaxPoints / R + NT - 1) / NT;
  const int nb = L / R;
  const int total = nseq * nb;
  const int span = L / (Ns * R);
  const float ns_inv = 1.0f / (float)Ns;
  V2<T> v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = (int)threadIdx.x + q * NT;
    dst[q] = -1;
    if (b < total) {
      const int f = b >= nb ? 1 : 0;  // at most two sequences per workgroup
      const int j = b - f * nb;
      const int jm = j - div_fft(j, Ns, ns_inv) * Ns;
      const V2<T>* src = buf + f * L + j;
      v[q][0] = src[0];
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = cmul(src[r * nb], tw[jm * r * span]);
      dft<T, R>(v[q]);
      dst[q] = f * L + (j - jm) * R + jm;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q)
    if (dst[q] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns] = v[q][r];
    }
  __syncthreads();
}

// DCT-I of the y-columns x0 .. x0 + C - 1 (x0 = C * group, C = 2 or 4), all n + 1 rows:
//   dst[m][x] = sum_j src[j][x] cos(pi j m / n),  m = 0 .. n,
// with (INV) partials of sum rdot . dst per workgroup.  Workgroups b, b + 8, b + 16, .. (one
// XCD each, as workgroups are dealt out to the XCDs in turn) take consecutive column groups,
// so the workgroups that share a row's cache line share that XCD's L2.
template <typename T, bool INV, int C>
__device__ __forceinline__ void dct1_cols_body(const BlkId blk, int n, int mx, int groups,
                                               const FftPlan& plan, const V2<T>* tw_,
                                               const T* src_, T* dst_, const T* rdot_,
                                               double* part) {
  const auto tw = glob(tw_);
  const auto src = glob(src_);
  const auto dst = glob(dst_);
  const auto rdot = glob(rdot_);
  constexpr int NS = C / 2;  // complex sequences (C = 2 or 4 columns)
  constexpr int NT = kFftThreads<C>;
  extern __shared__ __align__(16) unsigned char dct_lds[];
  V2<T>* z = reinterpret_cast<V2<T>*>(dct_lds);
  __shared__ V2<T> ends[2][2];  // each sequence's points 0 and n before the transform
  const int L = 2 * n;
  V2<T>* twl = z + NS * L;  // the twiddle table, staged with the columns (FP32)
  const int per = (int)blk.n / 8;
  const int g = ((int)blk.id % 8) * per + (int)blk.id / 8;
  double acc = 0.0;
  if (g < groups) {
    const int x0 = g * C;
    const int total = (n + 1) * C;
    T* zs = reinterpret_cast<T*>(z);
    // column c is the real (c even) or imaginary part of sequence c / 2; point j is stored
    // at j and (0 < j < n) at its mirror L - j.  Eight loads in flight per thread.
    constexpr int NB = 8;
    constexpr int lc = C == 4 ? 2 : 1;  // log2 C
    if (kTwiddlesInLds<T>)
    for (int k0 = (int)threadIdx.x; k0 < L; k0 += 4 * NT) {
      V2<T> w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k0 + u * NT < L) w[u] = tw[k0 + u * NT];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k0 + u * NT < L) twl[k0 + u * NT] = w[u];
    }
    for (int i0 = (int)threadIdx.x; i0 < total; i0 += NB * NT) {
      T v[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = i0 + u * NT;
        const int j = i >> lc, c = i & (C - 1);
        v[u] = (i < total && x0 + c < mx) ? src[(size_t)j * mx + x0 + c] : T(0);
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = i0 + u * NT;
        if (i < total) {
          const int j = i >> lc, c = i & (C - 1);
          const int base = 2 * (c / 2) * L + (c & 1);
          zs[base + 2 * j] = v[u];
          if (j > 0 && j < n) zs[base + 2 * (L - j)] = v[u];
        }
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < NS) {
      ends[threadIdx.x][0] = z[threadIdx.x * L];
      ends[threadIdx.x][1] = z[threadIdx.x * L + n];
    }
    // (the first stage's barrier orders these reads before any write of the transform)
    int Ns = 1;
    for (int st = 0; st < plan.nstages; ++st) {
      const int R = plan.radix[st];
      const V2<T>* twp = kTwiddlesInLds<T> ? static_cast<const V2<T>*>(twl) : tw;
      switch (R) {
        case 2: fft_stage<T, 2, NT>(z, L, NS, Ns, twp); break;
        case 3: fft_stage<T, 3, NT>(z, L, NS, Ns, twp); break;
        case 4: fft_stage<T, 4, NT>(z, L, NS, Ns, twp); break;
        case 5: fft_stage<T, 5, NT>(z, L, NS, Ns, twp); break;
        default: fft_stage<T, 7, NT>(z, L, NS, Ns, twp); break;
      }
      Ns *= R;
    }
    // T[m] = (Y[m] + r_0 + (-1)^m r_n) / 2; INV reads r for the r.z partials NO points at a
    // time, so its loads are in flight together (FP32; FP64 has no registers to spare)
    constexpr int NO = sizeof(T) == 4 ? NB : 1;
    for (int i0 = (int)threadIdx.x; i0 < total; i0 += NO * NT) {
      T rv[NO];
      if constexpr (INV) {
#pragma unroll
        for (int u = 0; u < NO; ++u) {
          const int i = i0 + u * NT;
          const int m = i >> lc, x = x0 + (i & (C - 1));
          rv[u] = (i < total && x < mx) ? rdot[(size_t)m * mx + x] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < NO; ++u) {
        const int i = i0 + u * NT;
        const int m = i >> lc, c = i & (C - 1), x = x0 + c;
        if (i < total && x < mx) {
          const int f = c / 2;
          const V2<T> Y = z[f * L + m];
          const T y = (c & 1) ? Y.y : Y.x;
          const T r0 = (c & 1) ? ends[f][0].y : ends[f][0].x;
          const T rn = (c & 1) ? ends[f][1].y : ends[f][1].x;
          const T val = T(0.5) * (y + r0 + ((m & 1) ? -rn : rn));
          dst[(size_t)m * mx + x] = val;
          if constexpr (INV) acc += (double)rv[u] * (double)val;
        }
      }
    }
  }
  if constexpr (INV) {
    if (part) {
      const double v[1] = {acc};
      double* const out[1] = {part};
      block_partials<1, NT>(v, out, blk.id);
    }
  }
}

template <typename T, bool INV, int C>
__global__ __launch_bounds__(kFftThreads<C>) void k_dct1_cols(int n, int mx, int groups,
                                                          FftPlan plan,
                                                          const V2<T>* __restrict__ tw,
                                                          const T* __restrict__ src,
                                                          T* __restrict__ dst,
                                                          const T* __restrict__ rdot,
                                                          double* __restrict__ part) {
  dct1_cols_body<T, INV, C>(hw_blk(), n, mx, groups, plan, tw, src, dst, rdot, part);
}

// one DCT-I launch over the column groups (C = 2 or 4 columns per workgroup)
template <typename T, bool INV>
static void launch_dct1_cols(int cols, unsigned nblk, size_t lds, hipStream_t s, int n, int mx,
                             int groups, const FftPlan& plan, const V2<T>* tw, const T* src,
                             T* dst, const T* rdot, double* part) {
  if constexpr (sizeof(T) == 4) {  // (FP64 always takes two columns)
    if (cols == 4) {
      TMX_LAUNCH_OR_RECORD(kFftThreads<4>, (dct1_cols_body<T, INV, 4>), nblk, lds, s, (k_dct1_cols<T, INV, 4>), n, mx, groups, plan, tw, src, dst, rdot, part);
      return;
    }
  }
  TMX_LAUNCH_OR_RECORD(kFftThreads<2>, (dct1_cols_body<T, INV, 2>), nblk, lds, s, (k_dct1_cols<T, INV, 2>), n, mx, groups, plan, tw, src, dst, rdot, part);
}

// One tridiagonal solve along x per mode row: row m of `modes` is mode m (workgroup m).  Each
// solves
// (a T_x + c_m W_x) u = f / D_m in place, one workgroup of kTriThreads per mode, by the
// partition method: thread t owns the k = ceil(mx / kTriThreads) consecutive rows s..e of one
// block.  A downward elimination writes each of its rows as
//   x_i + al_i x_{s-1} + ga_i x_{i+1} = ph_i,
// an upward one as x_i = P_i - A_i x_{s-1} - G_i x_e (P_e = A_e = 0, G_e = -1).  Substituting
// the next block's first row into this block's last one leaves a tridiagonal system in the
// blocks' last unknowns y_t = x_{e_t}:
//   al_e y_{t-1} + (1 - ga_e A'_s) y_t - ga_e G'_s y_{t+1} = ph_e - ga_e P'_s   (' : block t+1),
// solved by parallel cyclic reduction in LDS (log2(kTriThreads) steps, one equation per
// thread); then x_i = P_i - A_i y_{t-1} - G_i y_t.  The operator is diagonally dominant, so the
// eliminations need no pivoting, and FP32 takes the hardware reciprocal (1 ulp; the solver is a
// preconditioner) where FP64 divides.  ph / P stay in the staged row; al / A and ga / G sit in LDS
// transposed (index j * nblk + t), so a wave's accesses fall in distinct banks.  LDS is sized
// to the row (40 KB in FP32 at 3001 nodes: four workgroups per CU).
constexpr int kTriThreads = 256;

// LDS of k_dct_tri for rows of mx nodes: the row, two coefficient arrays of k * nblk values
// (k rows per block, nblk blocks) and four arrays of the interface system
template <typename T>
static size_t dct_tri_lds(int mx) {
  const int k = (mx + kTriThreads - 1) / kTriThreads;
  const int nblk = (mx + k - 1) / k;
  return sizeof(T) * ((size_t)mx + 2 * (size_t)k * nblk + 4 * (size_t)kTriThreads);
}

template <typename T>
__device__ __forceinline__ void dct_tri_body(const BlkId blk, int mx, int n, T a_off,
                                             double hxhy, double cy, T* modes_) {
  const auto modes = glob(modes_);
  extern __shared__ __align__(16) unsigned char tri_lds[];
  const int k = (mx + kTriThreads - 1) / kTriThreads;
  const int nblk = (mx + k - 1) / k;
  T* lrow = reinterpret_cast<T*>(tri_lds);  // right-hand side, then ph, P and x
  T* lal = lrow + mx;                       // al then A, at j * nblk + t
  T* lga = lal + (size_t)k * nblk;          // ga then G, at j * nblk + t
  T* ra = lga + (size_t)k * nblk;           // interface system; ra..rc first hold the first
  T* rb = ra + kTriThreads;                 // rows' P, A, G
  T* rc = rb + kTriThreads;
  T* rf = rc + kTriThreads;
  const int mode = (int)blk.id;
  T* row = modes + (size_t)mode * mx;
  const double lam = 2.0 - 2.0 * std::cos(M_PI * (double)mode / (double)n);
  const T c = (T)(hxhy + cy * lam);
  const T inv_d = (mode == 0 || mode == n) ? T(1) / (T)n : T(2) / (T)n;
  const T b_in = a_off * T(2) + c, b_edge = a_off + c * T(0.5);
  const int tid = threadIdx.x;
  // the row's loads all in flight before the first use (up to kMaxRow / kTriThreads each)
  {
    constexpr int NB = (DctPrecond<T>::kMaxRow + kTriThreads - 1) / kTriThreads;
    T v[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = tid + u * kTriThreads;
      v[u] = i < mx ? row[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = tid + u * kTriThreads;
      if (i < mx) lrow[i] = v[u] * inv_d;
    }
  }
  const int s0 = tid * k;
  const int jl = tid < nblk ? min(s0 + k, mx) - 1 - s0 : -1;  // last row of the block, local
  __syncthreads();
  // downward over rows s0 .. s0 + jl
  T al = T(0), ga = T(0), ph = T(0);
  for (int j = 0; j <= jl; ++j) {
    const int i = s0 + j;
    const T ai = i > 0 ? -a_off : T(0);
    const T ci = i < mx - 1 ? -a_off : T(0);
    const T bi = (i == 0 || i == mx - 1) ? b_edge : b_in;
    const T inv = recip(bi - ai * ga);  // ga = 0 at j = 0
    al = j == 0 ? ai * inv : -ai * al * inv;
    ph = (lrow[i] - ai * ph) * inv;
    ga = ci * inv;
    lrow[i] = ph;
    lal[j * nblk + tid] = al;
    lga[j * nblk + tid] = ga;
  }
  const T ale = al, gae = ga, phe = ph;  // the block's last row
  // upward: (ph, al, ga) -> (P, A, G) in place
  T P = T(0), A = T(0), G = T(-1);
  if (jl >= 0) {
    lrow[s0 + jl] = P;
    lal[jl * nblk + tid] = A;
    lga[jl * nblk + tid] = G;
  }
  for (int j = jl - 1; j >= 0; --j) {
    const int q = j * nblk + tid;
    const T g = lga[q];
    P = lrow[s0 + j] - g * P;
    A = lal[q] - g * A;
    G = -g * G;
    lrow[s0 + j] = P;
    lal[q] = A;
    lga[q] = G;
  }
  ra[tid] = P;
  rb[tid] = A;
  rc[tid] = G;
  __syncthreads();
  // interface system in the blocks' last unknowns (identity rows past the last block)
  {
    T ai = T(0), bi = T(1), ci = T(0), fi = T(0);
    if (tid < nblk) {
      ai = ale;
      fi = phe;
      if (tid + 1 < nblk) {
        bi = T(1) - gae * rb[tid + 1];
        ci = -gae * rc[tid + 1];
        fi = phe - gae * ra[tid + 1];
      }
    }
    __syncthreads();  // ra..rc are overwritten below
    ra[tid] = ai;
    rb[tid] = bi;
    rc[tid] = ci;
    rf[tid] = fi;
  }
  __syncthreads();
  for (int st = 1; st < kTriThreads; st *= 2) {
    const int im = tid - st, ip = tid + st;
    T ai = T(0), bi = rb[tid], ci = T(0), fi = rf[tid];
    if (im >= 0) {
      const T k1 = ra[tid] * recip(rb[im]);
      ai = -ra[im] * k1;
      bi -= rc[im] * k1;
      fi -= rf[im] * k1;
    }
    if (ip < kTriThreads) {
      const T k2 = rc[tid] * recip(rb[ip]);
      ci = -rc[ip] * k2;
      bi -= ra[ip] * k2;
      fi -= rf[ip] * k2;
    }
    __syncthreads();
    ra[tid] = ai;
    rb[tid] = bi;
    rc[tid] = ci;
    rf[tid] = fi;
    __syncthreads();
  }
  const T y = rf[tid] * recip(rb[tid]);  // y_t
  __syncthreads();
  ra[tid] = y;
  __syncthreads();
  const T ym = tid > 0 ? ra[tid - 1] : T(0);
  for (int j = 0; j <= jl; ++j) {
    const int q = j * nblk + tid;
    lrow[s0 + j] = lrow[s0 + j] - lal[q] * ym - lga[q] * y;
  }
  __syncthreads();
  for (int i = tid; i < mx; i += kTriThreads) row[i] = lrow[i];
}

template <typename T>
__global__ __launch_bounds__(kTriThreads) void k_dct_tri(int mx, int n, T a_off, double hxhy,
                                                         double cy, T* __restrict__ modes) {
  dct_tri_body<T>(hw_blk(), mx, n, a_off, hxhy, cy, modes);
}

// LDS of k_dct1_cols: the complex sequences of `cols` columns (2 n points each per pair) and
// the twiddle table of 2 n points
template <typename T>
static size_t dct_cols_lds(int cols, int n) {
  return ((size_t)(cols / 2) + (kTwiddlesInLds<T> ? 1 : 0)) * 2 * (size_t)n * sizeof(V2<T>);
}

// columns per FFT workgroup: in FP32 four (two complex sequences, 16-byte row segments, on
// 512 threads) for columns of 512 to 1024 cells, whose FFTs give each thread a few butterflies
// per stage; else two (on 256 threads: shorter columns leave most of 512 threads idle, and
// longer ones exceed the LDS); two in FP64
template <typename T>
static int dct_cols(int n) {
  return sizeof(T) == 4 && n >= 512 && (4 / 2) * 2 * n <= kFftMaxPoints &&
                 dct_cols_lds<T>(4, n) <= 64 * 1024
             ? 4
             : 2;
}

template <typename T>
bool DctPrecond<T>::supported(int nx, int ny) {
  FftPlan p;
  return ny >= 1 && nx >= 1 && nx + 1 <= kMaxRow && fft_plan(2 * ny, p) &&
         (kCols / 2) * 2 * ny <= kFftMaxPoints;
}

template <typename T>
int DctPrecond<T>::init(int nx, int ny, double hx, double hy, double eps, hipStream_t s) {
  if (!supported(nx, ny)) {
    set_error("internal: DCT preconditioner needs 2 ny with prime factors <= 7, ny <= " +
              std::to_string(kFftMaxPoints / kCols) + " and nx + 1 <= " +
              std::to_string(kMaxRow));
    return -1;
  }
  nx_ = nx;
  ny_ = ny;
  mx_ = nx + 1;
  hx_ = hx;
  hy_ = hy;
  eps_ = eps;
  const int L = 2 * ny;
  (void)fft_plan(L, plan_);
  // twiddles exp(-2 pi i k / L) from the exact integer angle, evaluated in double
  std::vector<T> tw(2 * (size_t)L);
  for (int k = 0; k < L; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)L;
    tw[2 * (size_t)k] = (T)std::cos(a);
    tw[2 * (size_t)k + 1] = (T)-std::sin(a);
  }
  TMX_TRY(tw_.alloc(tw.size()));
  TMX_TRY(t_.alloc((size_t)(ny + 1) * mx_));
  TMX_HIP(hipMemcpyAsync(tw_.get(), tw.data(), tw.size() * sizeof(T), hipMemcpyHostToDevice, s));
  TMX_HIP(hipStreamSynchronize(s));
  // rows up to 3072 nodes need more than 64 KB of dynamic LDS in FP64
  TMX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dct_tri<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)dct_tri_lds<T>(kMaxRow)));
  std::vector<const void*> fns = {reinterpret_cast<const void*>(&k_dct1_cols<T, false, 2>),
                                   reinterpret_cast<const void*>(&k_dct1_cols<T, true, 2>)};
  if constexpr (sizeof(T) == 4) {
    fns.push_back(reinterpret_cast<const void*>(&k_dct1_cols<T, false, 4>));
    fns.push_back(reinterpret_cast<const void*>(&k_dct1_cols<T, true, 4>));
  }
  for (const void* f : fns)
    TMX_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)dct_cols_lds<T>(kCols, kFftMaxPoints / 2)));
  return 0;
}

template <typename T>
int DctPrecond<T>::apply(const T* r, T* z, RedSlot* rz, hipStream_t s) {
  const int n = ny_, mx = mx_;
  const int cols = dct_cols<T>(n);
  const int groups = (mx + cols - 1) / cols;
  const int nblk = 8 * ((groups + 7) / 8);
  const size_t lds = dct_cols_lds<T>(cols, n);
  const V2<T>* tw = reinterpret_cast<const V2<T>*>(tw_.get());
  T* t = t_.get();
  // forward DCT-I along y: t[m] = sum_j r[j] cos(pi j m / n)
  launch_dct1_cols<T, false>(cols, nblk, lds, s, n, mx, groups, plan_, tw, r, t, nullptr, nullptr);
  TMX_HIP(hipGetLastError());
  // tridiagonal solves along x, one workgroup per mode (scaled by 1 / D_m there)
  const T a_off = (T)(eps_ * eps_ * hy_ / hx_);
  const double cy = eps_ * eps_ * hx_ / hy_;
  TMX_LAUNCH_OR_RECORD(kTriThreads, (dct_tri_body<T>), n + 1, dct_tri_lds<T>(mx), s, (k_dct_tri<T>), mx, n, a_off, hx_ * hy_, cy, t);
  TMX_HIP(hipGetLastError());
  // inverse: z[j] = sum_m t[m] cos(pi j m / n), with the partials of r.z
  double* part = nullptr;
  if (rz) {
    if (nblk > kMaxPartials) {
      set_error("internal: too many DCT workgroups for the reduction slots");
      return -1;
    }
    rz->nblocks = nblk;
    part = rz->partials;
  }
  launch_dct1_cols<T, true>(cols, nblk, lds, s, n, mx, groups, plan_, tw, t, z, r, part);
  TMX_HIP(hipGetLastError());
  return 0;
}

template class DctPrecond<float>;
template class DctPrecond<double>;

}  // namespace tmx
