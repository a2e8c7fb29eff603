// The device library, part four: batched symmetric eigendecomposition in FP64 for the split of the
// two-site tensor (the Gram matrices of its blocks), in four launches for a whole batch of
// matrices of different sizes:
//
//   1. Householder tridiagonalisation, one workgroup per matrix.  The rank-2 update of step k is
//      deferred into step k+1's symmetric matrix-vector pass, so each step reads and writes the
//      trailing matrix once.  The reflectors stay in the matrix's lower part.
//   2. Cuppen divide and conquer on the tridiagonal: leaves of 32 rows by implicit QL (one wave
//      each), then merges bottom up, one workgroup per matrix: deflation of small coupling components and of close poles (Givens rotations), the
//      secular roots by rational ("middle way") steps from the nearer pole (one thread per root), eigenvectors
//      by the Gu-Eisenstat formula (orthogonal to working precision however close the roots), and
//      the merge product through LDS tiles.
//   3. Back-transformation of the eigenvectors by the reflectors, one workgroup per 16-column
//      panel held in LDS.
//
// Eigenvalues come out ascending; eigenvector j is row j of the matrix's n x n output.  gfx950 only.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

using ttdev::as_stream;
using ttdev::fail;

namespace {

constexpr int kEigThreads = 256;
constexpr int kEigWaves = kEigThreads / 64;
constexpr int kEigMaxN = 1024;      // largest matrix (LDS vectors are sized for it)
constexpr int kEigBatch = 64;       // problems per launch (kernel-argument table)
constexpr int kPanel = 16;          // columns per back-transformation workgroup
constexpr int kLeaf = 32;           // order of the divide-and-conquer leaves
constexpr double kEps = 2.220446049250313e-16;

struct Problem {
  int64_t matrix;       // offset of the n x n matrix (and of its eigenvectors)
  int64_t values;       // offset of its n eigenvalues
  int64_t work;         // offset of its workspace (3 n^2 + 3 n doubles)
  int32_t n;
  int32_t first_panel;  // index of its first back-transformation panel in the launch
  int32_t first_leaf;   // index of its first divide-and-conquer leaf in the launch
  int32_t pad;
};

struct Batch {
  Problem p[kEigBatch];
  int32_t count;
};

__device__ double wave_sum(double v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

// ---------------------------------------------------------------------------------------------
// 1. Tridiagonalisation.  A is column-major n x n (symmetric, so either order), in place.
//    Output: diag[n], off[n-1], tau[n-1]; reflector k is A[k+1.., k] with A[k+1, k] = 1.
//    1024 threads: in the trailing pass thread t takes rows k+1 + t % 256 (+ 256 r) of every fourth
//    column (group t / 256), eight columns at a time (four with four rows), so each thread keeps 16
//    loads in flight;
//    the four groups' partial products are summed in LDS in group order.  kTriRows (rows per thread:
//    1, 2 or 4) is the smallest that covers the batch's largest matrix, to keep the loads in registers.

constexpr int kTriThreads = 1024;
constexpr int kTriGroups = kTriThreads / 256;

__device__ double tri_block_sum(double v, double *scratch) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  double total = 0.0;
  for (int w = 0; w < kTriThreads / 64; ++w) total += scratch[w];
  return total;
}

template <int kTriRows>
__global__ __launch_bounds__(kTriThreads) void tridiag_kernel(Batch batch, double *__restrict__ matrices,
                                                               double *__restrict__ work) {
  constexpr int kTriUnroll = kTriRows == 4 ? 4 : 8;   // columns per thread per batch of loads
  const Problem pr = batch.p[blockIdx.x];
  const int n = pr.n;
  double *__restrict__ A = matrices + pr.matrix;
  double *diag = work + pr.work + 3 * int64_t(n) * n, *off = diag + n, *tau = off + n;
  // (v, w) of the previous step and of this one, swapped by index after every step
  __shared__ double vbuf[2][kEigMaxN], wbuf[2][kEigMaxN];
  __shared__ double part[kTriGroups][kEigMaxN];
  __shared__ double red[kTriThreads / 64];
  __shared__ double s_alpha, s_diag;
  const int t = threadIdx.x, row0 = t & 255, group = t >> 8;
  for (int i = t; i < n; i += kTriThreads) vbuf[0][i] = wbuf[0][i] = 0.0;
  __syncthreads();
  int cur = 0;
  for (int k = 0; k + 1 < n; ++k) {
    const double *vp = vbuf[cur], *wp = wbuf[cur];
    double *vc = vbuf[cur ^ 1], *wc = wbuf[cur ^ 1];
    double *col = A + int64_t(k) * n;
    // (a) column k, rows k.., brought up to date with the previous step's update
    double sigma = 0.0;
    for (int i = k + t; i < n; i += kTriThreads) {
      const double a = col[i] - (vp[i] * wp[k] + wp[i] * vp[k]);
      if (i == k) s_diag = a;
      if (i == k + 1) s_alpha = a;
      if (i >= k + 2) {
        col[i] = a;
        sigma += a * a;
      }
    }
    sigma = tri_block_sum(sigma, red);
    // every thread forms the reflector's scalars from the same values
    const double alpha = s_alpha;
    double beta = alpha, tk = 0.0, scale = 0.0;
    if (sigma != 0.0) {
      const double norm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -norm : norm;
      tk = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    if (t == 0) {
      diag[k] = s_diag;
      off[k] = beta;
      tau[k] = tk;
      col[k + 1] = 1.0;
    }
    for (int i = k + 1 + t; i < n; i += kTriThreads) {
      double v = 1.0;
      if (i >= k + 2) {
        v = col[i] * scale;
        col[i] = v;
      }
      vc[i] = v;
    }
    __syncthreads();
    // (b) one pass over the trailing matrix: apply the previous update, form p = tau A v
    double p[kTriRows], vpi[kTriRows], wpi[kTriRows];
#pragma unroll
    for (int r = 0; r < kTriRows; ++r) {
      const int i = k + 1 + row0 + r * 256;
      p[r] = 0.0;
      vpi[r] = i < n ? vp[i] : 0.0;
      wpi[r] = i < n ? wp[i] : 0.0;
    }
    for (int j0 = k + 1 + group; j0 < n; j0 += kTriUnroll * kTriGroups) {
      double a[kTriUnroll][kTriRows];
#pragma unroll
      for (int u = 0; u < kTriUnroll; ++u) {
        const int j = j0 + u * kTriGroups;
#pragma unroll
        for (int r = 0; r < kTriRows; ++r) {
          const int i = k + 1 + row0 + r * 256;
          a[u][r] = (j < n && i < n) ? A[i + int64_t(j) * n] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < kTriUnroll; ++u) {
        const int j = j0 + u * kTriGroups;
        if (j >= n) continue;
        const double vpj = vp[j], wpj = wp[j], vcj = vc[j];
#pragma unroll
        for (int r = 0; r < kTriRows; ++r) {
          const int i = k + 1 + row0 + r * 256;
          if (i < n) {
            const double x = a[u][r] - (vpi[r] * wpj + wpi[r] * vpj);
            A[i + int64_t(j) * n] = x;
            p[r] += x * vcj;
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kTriRows; ++r) {
      const int i = k + 1 + row0 + r * 256;
      if (i < n) part[group][i] = p[r];
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = k + 1 + t; i < n; i += kTriThreads) {
      double total = 0.0;
      for (int g = 0; g < kTriGroups; ++g) total += part[g][i];
      total *= tk;
      wc[i] = total;   // p, for now
      pv += total * vc[i];
    }
    pv = tri_block_sum(pv, red);
    for (int i = k + 1 + t; i < n; i += kTriThreads) wc[i] = wc[i] - 0.5 * tk * pv * vc[i];
    __syncthreads();
    cur ^= 1;
  }
  if (t == 0) diag[n - 1] = A[int64_t(n - 1) * n + n - 1] - 2.0 * vbuf[cur][n - 1] * wbuf[cur][n - 1];
}

// ---------------------------------------------------------------------------------------------
// 2a. Leaves of the divide and conquer: every block of kLeaf rows of the tridiagonal (with the
//     couplings of its two cuts taken off its end diagonals, Cuppen's split) solved by implicit QL
//     with Wilkinson-type shifts and eigenvector accumulation, one wave per leaf.  Every lane runs
//     the scalar recurrence on the wave's copy of the leaf in LDS; lane r applies the rotations to
//     row r of the leaf's eigenvectors.  The eigenpairs are then sorted ascending and written to the
//     leaf's diagonal block of Q, whose other entries in the leaf's rows are zeroed (and the whole
//     rows of the second buffer too).

__global__ __launch_bounds__(kEigThreads) void leaf_kernel(Batch batch, double *__restrict__ values,
                                                            double *__restrict__ work, int32_t leaves) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int leaf = static_cast<int>(blockIdx.x) * kEigWaves + wave;
  __shared__ double dl[kEigWaves][kLeaf], el[kEigWaves][kLeaf], Z[kEigWaves][kLeaf][kLeaf + 1];
  if (leaf >= leaves) return;
  int b = 0;
  while (b + 1 < batch.count && batch.p[b + 1].first_leaf <= leaf) ++b;
  const Problem pr = batch.p[b];
  const int n = pr.n;
  const int lo = (leaf - pr.first_leaf) * kLeaf, hi = min(lo + kLeaf, n), m = hi - lo;
  double *Q = work + pr.work, *Qt = Q + int64_t(n) * n;
  const double *diag = Q + 3 * int64_t(n) * n, *off = diag + n;
  double *d = dl[wave], *e = el[wave];
  double (*z)[kLeaf + 1] = Z[wave];
  if (lane < m) {
    double v = diag[lo + lane];
    if (lane == 0 && lo > 0) v -= off[lo - 1];
    if (lane == m - 1 && hi < n) v -= off[hi - 1];
    d[lane] = v;
    e[lane] = lane + 1 < m ? off[lo + lane] : 0.0;
    for (int c = 0; c < m; ++c) z[lane][c] = lane == c ? 1.0 : 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  for (int l = 0; l < m; ++l) {
    for (int iter = 0; iter < 60; ++iter) {
      int mm = l;
      for (; mm < m - 1; ++mm)
        if (fabs(e[mm]) <= kEps * (fabs(d[mm]) + fabs(d[mm + 1]))) break;
      if (mm == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
      double sn = 1.0, cs = 1.0, p = 0.0;
      bool split = false;
      for (int i = mm - 1; i >= l; --i) {
        const double f = sn * e[i], bb = cs * e[i];
        r = hypot(f, g);
        __builtin_amdgcn_wave_barrier();
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[mm] = 0.0;
          split = true;
          break;
        }
        sn = f / r;
        cs = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * cs * bb;
        p = sn * r;
        __builtin_amdgcn_wave_barrier();
        d[i + 1] = g + p;
        g = cs * r - bb;
        if (lane < m) {
          const double zi = z[lane][i], zi1 = z[lane][i + 1];
          z[lane][i + 1] = sn * zi + cs * zi1;
          z[lane][i] = cs * zi - sn * zi1;
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[mm] = 0.0;
      __builtin_amdgcn_wave_barrier();
    }
  }
  // ascending order, vectors along
  for (int i = 0; i < m; ++i) {
    int best = i;
    for (int j = i + 1; j < m; ++j)
      if (d[j] < d[best]) best = j;
    __builtin_amdgcn_wave_barrier();
    if (best != i) {
      const double di = d[i], db = d[best];
      __builtin_amdgcn_wave_barrier();
      d[i] = db;
      d[best] = di;
      if (lane < m) {
        const double a = z[lane][i];
        z[lane][i] = z[lane][best];
        z[lane][best] = a;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  double *lam = values + pr.values;
  if (lane < m) lam[lo + lane] = d[lane];
  // the leaf's rows of both buffers: zero, then the leaf's eigenvectors in its diagonal block
  for (int c = 0; c < n; ++c)
    if (lane < m) {
      Q[lo + lane + int64_t(c) * n] = (c >= lo && c < hi) ? z[lane][c - lo] : 0.0;
      Qt[lo + lane + int64_t(c) * n] = 0.0;
    }
}

// ---------------------------------------------------------------------------------------------
// 2b. Divide and conquer.  Q and Qt are column-major n x n (Q[r + c n]); U is a merge's k x k.

struct Rotation {
  int32_t a, b;  // local columns
  double c, s;
};

// The secular function f(x) = 1 + psi(x) + phi(x) of root j, with x measured from pole o: psi sums
// over the kept poles 0..j (at or left of the root's interval), phi over the rest; both with their
// derivatives.  One thread evaluates it serially; every thread of the workgroup reads the same pole
// at the same time, an LDS broadcast.
struct Secular {
  double f, psi, phi, dpsi, dphi;
};

// 1/d from the hardware reciprocal estimate and one Newton-Raphson correction r + r (1 - d r).
__device__ inline double reciprocal(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

__device__ Secular secular(const double *dk, const double *z2, int k, int j, int o, double rho, double x) {
  const double origin = dk[o];
  double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
  // one loop for both sums, so lanes working on different roots do not diverge
  for (int i = 0; i < k; ++i) {
    const double inv = reciprocal((dk[i] - origin) - x);
    const double w = z2[i] * inv;
    const bool left = i <= j;
    psi += left ? w : 0.0;
    dpsi += left ? w * inv : 0.0;
    phi += left ? 0.0 : w;
    dphi += left ? 0.0 : w * inv;
  }
  return Secular{1.0 + rho * (psi + phi), rho * psi, rho * phi, rho * dpsi, rho * dphi};
}

__global__ __launch_bounds__(kEigThreads) void dc_kernel(Batch batch, double *__restrict__ values,
                                                          double *__restrict__ work) {
  const Problem pr = batch.p[blockIdx.x];
  const int n = pr.n;
  double *const base = work + pr.work;
  double *__restrict__ U = base + 2 * int64_t(n) * n;
  const double *diag = base + 3 * int64_t(n) * n, *off = diag + n;
  // each level reads its blocks from Q and writes the merged blocks to Qt; the two then swap
  double *Q = base, *Qt = base + int64_t(n) * n;
  double *lam = values + pr.values;
  __shared__ double dloc[kEigMaxN], zloc[kEigMaxN], pole[kEigMaxN], zsort[kEigMaxN];
  __shared__ double dk[kEigMaxN], z2[kEigMaxN], zk[kEigMaxN], tau[kEigMaxN], zh[kEigMaxN], val[kEigMaxN];
  __shared__ int32_t perm[kEigMaxN], defl[kEigMaxN], keep[kEigMaxN], org[kEigMaxN], src[kEigMaxN], rnk[kEigMaxN];
  __shared__ Rotation rot[kEigMaxN];  // each rotation deflates one pole: at most s - 1
  __shared__ double tile_a[32][33], tile_b[32][33];
  __shared__ int s_k, s_rot;
  __shared__ double s_rho;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  // the leaves (blocks of kLeaf rows) were solved by leaf_kernel, which also cleared both buffers
  for (int width = kLeaf; width < n; width *= 2) {
    // a block with no partner at this level moves over unchanged
    {
      const int lo = ((n - 1) / (2 * width)) * (2 * width);
      if (lo + width >= n) {
        const int s = n - lo;
        for (int64_t e = t; e < int64_t(s) * s; e += kEigThreads)
          Qt[lo + e % s + int64_t(lo + e / s) * n] = Q[lo + e % s + int64_t(lo + e / s) * n];
      }
    }
    for (int lo = 0; lo + width < n; lo += 2 * width) {
      const int mid = lo + width, hi = min(lo + 2 * width, n), s = hi - lo, left = mid - lo;
      const double beta = off[mid - 1];
      const double sign = beta < 0.0 ? -1.0 : 1.0;
      // (a) coupling vector (last row of the left block, first row of the right) and poles
      for (int i = t; i < s; i += kEigThreads) {
        const double q = i < left ? Q[(mid - 1) + int64_t(lo + i) * n] : Q[mid + int64_t(lo + i) * n];
        zloc[i] = q * 0.7071067811865476;
        dloc[i] = sign * lam[lo + i];
      }
      __syncthreads();
      // (b) one thread: merge the sorted halves, deflate small z and close poles
      if (t == 0) {
        const double rho = 2.0 * beta * sign;
        int a = 0, b = 0;
        for (int p = 0; p < s; ++p) {
          // ascending in dloc: a negated half is descending, so it is walked from its end
          const int ia = sign > 0 ? a : left - 1 - a, ib = sign > 0 ? left + b : s - 1 - b;
          const bool take_a = b >= s - left || (a < left && dloc[ia] <= dloc[ib]);
          perm[p] = take_a ? ia : ib;
          if (take_a) ++a; else ++b;
        }
        double dmax = 0.0;
        for (int p = 0; p < s; ++p) {
          pole[p] = dloc[perm[p]];
          zsort[p] = zloc[perm[p]];
          dmax = fmax(dmax, fabs(pole[p]));
        }
        const double tol = 8.0 * kEps * fmax(dmax, rho);
        int nrot = 0, prev = -1;
        for (int p = 0; p < s; ++p) {
          defl[p] = rho * fabs(zsort[p]) <= tol ? 1 : 0;
          if (defl[p]) continue;
          if (prev >= 0) {
            const double r = hypot(zsort[p], zsort[prev]);
            const double c = zsort[p] / r, sn = -zsort[prev] / r;
            if (fabs((pole[p] - pole[prev]) * c * sn) <= tol) {
              zsort[p] = r;
              zsort[prev] = 0.0;
              rot[nrot++] = Rotation{perm[prev], perm[p], c, sn};
              const double dp = pole[prev] * c * c + pole[p] * sn * sn;
              pole[p] = pole[prev] * sn * sn + pole[p] * c * c;
              pole[prev] = dp;
              defl[prev] = 1;
            }
          }
          prev = p;
        }
        int k = 0;
        for (int p = 0; p < s; ++p)
          if (!defl[p]) keep[k++] = p;
        s_k = k;
        s_rot = nrot;
        s_rho = rho;
      }
      __syncthreads();
      const int k = s_k, nrot = s_rot;
      const double rho = s_rho;
      // (c) the deflating rotations on Q's columns, in order; each thread owns its rows
      for (int r = lo + t; r < hi; r += kEigThreads)
        for (int q = 0; q < nrot; ++q) {
          double *ca = Q + r + int64_t(lo + rot[q].a) * n, *cb = Q + r + int64_t(lo + rot[q].b) * n;
          const double x = *ca, y = *cb;
          *ca = rot[q].c * x + rot[q].s * y;
          *cb = -rot[q].s * x + rot[q].c * y;
        }
      for (int j = t; j < k; j += kEigThreads) {
        dk[j] = pole[keep[j]];
        zk[j] = zsort[keep[j]];
        z2[j] = zk[j] * zk[j];
      }
      __syncthreads();
      // (d) secular roots, one thread per root, from the nearer pole of its interval.  Each step
      //     models the sums left and right of the interval by one pole each, matching value and
      //     slope at the iterate, and takes the root of that model (the rational "middle way"
      //     step); steps that leave the bracket bisect it.
      for (int j = t; j < k; j += kEigThreads) {
        int o = j;
        double a = 0.0, b;
        if (j + 1 < k) {
          const double gap = dk[j + 1] - dk[j];
          if (secular(dk, z2, k, j, j, rho, 0.5 * gap).f > 0.0) {
            b = 0.5 * gap;
          } else {
            o = j + 1;
            a = -0.5 * gap;
            b = 0.0;
          }
        } else {
          double zs = 0.0;
          for (int i = 0; i < k; ++i) zs += z2[i];
          b = rho * zs;
        }
        const double lo_pole = dk[j] - dk[o];
        const double hi_pole = j + 1 < k ? dk[j + 1] - dk[o] : 0.0;
        double x = 0.5 * (a + b);
        for (int it = 0; it < 80; ++it) {
          const Secular v = secular(dk, z2, k, j, o, rho, x);
          if (v.f > 0.0) b = x; else a = x;
          // converged when f is within its own rounding error, or the bracket is at the
          // resolution of x
          if (fabs(v.f) <= 8.0 * kEps * k * (1.0 + fabs(v.psi) + fabs(v.phi)) ||
              b - a <= 4.0 * kEps * fmax(fabs(a), fabs(b)))
            break;
          const double dlo = lo_pole - x;
          double eta;
          if (j + 1 < k) {
            const double dhi = hi_pole - x;
            const double S = dlo * dlo * v.dpsi, T = dhi * dhi * v.dphi;
            const double c = v.f - dlo * v.dpsi - dhi * v.dphi;
            const double B = c * (dlo + dhi) + S + T, C = c * dlo * dhi + S * dhi + T * dlo;
            const double root = sqrt(fmax(B * B - 4.0 * c * C, 0.0));
            if (B >= 0.0)
              eta = 2.0 * C / (B + root);
            else if (c != 0.0)
              eta = (B - root) / (2.0 * c);
            else
              eta = C / B;
          } else {
            const double c = v.f - dlo * v.dpsi;
            eta = c != 0.0 ? dlo + dlo * dlo * v.dpsi / c : 0.5 * (b - a);
          }
          double next = x + eta;
          if (!(next > a && next < b)) next = 0.5 * (a + b);
          if (fabs(next - x) <= 2.0 * kEps * fabs(x)) {
            x = next;
            break;
          }
          x = next;
        }
        org[j] = o;
        tau[j] = x;
      }
      __syncthreads();
      // (e) Gu-Eisenstat: zh_i^2 = (lam_i - d_i)/rho prod_{j != i} (lam_j - d_i)/(d_j - d_i)
      for (int i = t; i < k; i += kEigThreads) {
        double w = ((dk[org[i]] - dk[i]) + tau[i]) / rho;
        for (int j = 0; j < k; ++j)
          if (j != i) w *= ((dk[org[j]] - dk[i]) + tau[j]) / (dk[j] - dk[i]);
        zh[i] = copysign(sqrt(fabs(w)), zk[i]);
      }
      __syncthreads();
      // (f) U[:, j] = zh / (d - lam_j), normalised (k x k, column-major): one wave per column
      for (int j = wave; j < k; j += kEigWaves) {
        const double dj = dk[org[j]], tj = tau[j];
        double norm = 0.0;
        for (int i = lane; i < k; i += 64) {
          const double u = zh[i] / ((dk[i] - dj) - tj);
          norm += u * u;
        }
        const double inv = 1.0 / sqrt(wave_sum(norm));
        for (int i = lane; i < k; i += 64) U[i + int64_t(j) * k] = inv * zh[i] / ((dk[i] - dj) - tj);
      }
      // (g) slots: 0..k-1 the roots, then the deflated poles; their values and sources
      if (t == 0) {
        int slot = k;
        for (int p = 0; p < s; ++p)
          if (defl[p]) {
            val[slot] = pole[p];
            src[slot] = perm[p];
            ++slot;
          }
      }
      for (int j = t; j < k; j += kEigThreads) {
        val[j] = dk[org[j]] + tau[j];
        src[j] = -1;
      }
      __syncthreads();
      // (h) ascending rank of every slot (ties by slot), and its final column
      for (int x = t; x < s; x += kEigThreads) {
        int r = 0;
        for (int y = 0; y < s; ++y) r += (val[y] < val[x] || (val[y] == val[x] && y < x)) ? 1 : 0;
        rnk[x] = sign > 0 ? r : s - 1 - r;
        lam[lo + rnk[x]] = sign * val[x];
      }
      __syncthreads();
      // (i) deflated columns copied, root columns = Q[:, kept] U, into Qt
      for (int64_t e = t; e < int64_t(s - k) * s; e += kEigThreads) {
        const int x = k + static_cast<int>(e / s), r = static_cast<int>(e % s);
        Qt[lo + r + int64_t(lo + rnk[x]) * n] = Q[lo + r + int64_t(lo + src[x]) * n];
      }
      const int ty = t >> 3, tx = (t & 7) * 4;
      for (int r0 = 0; r0 < s; r0 += 32)
        for (int j0 = 0; j0 < k; j0 += 32) {
          double acc[4] = {0.0, 0.0, 0.0, 0.0};
          for (int i0 = 0; i0 < k; i0 += 32) {
            __syncthreads();
            for (int e = t; e < 32 * 32; e += kEigThreads) {
              const int fast = e & 31, slow = e >> 5;   // consecutive threads walk down a column
              const int row = r0 + fast, kk = i0 + slow;
              tile_a[fast][slow] = (row < s && kk < k) ? Q[lo + row + int64_t(lo + perm[keep[kk]]) * n] : 0.0;
              const int kb = i0 + fast, jb = j0 + slow;
              tile_b[fast][slow] = (kb < k && jb < k) ? U[kb + int64_t(jb) * k] : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int ii = 0; ii < 32; ++ii) {
              const double a = tile_a[ty][ii];
#pragma unroll
              for (int c = 0; c < 4; ++c) acc[c] += a * tile_b[ii][tx + c];
            }
          }
          const int row = r0 + ty;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = j0 + tx + c;
            if (row < s && j < k) Qt[lo + row + int64_t(lo + rnk[j]) * n] = acc[c];
          }
        }
      __syncthreads();
    }
    double *swap = Q;
    Q = Qt;
    Qt = swap;
    __syncthreads();
  }
  // the result must end in the first buffer, where the back-transformation reads it
  if (Q != base)
    for (int64_t e = t; e < int64_t(n) * n; e += kEigThreads) base[e] = Q[e];
}

// ---------------------------------------------------------------------------------------------
// 3. Back-transformation: Y = H_0 H_1 ... H_{n-2} Z on a panel of 16 columns of Z held in LDS.
//    Eigenvector c is written as row c of the output.

__global__ __launch_bounds__(kEigThreads) void back_kernel(Batch batch, const double *__restrict__ matrices,
                                                            double *__restrict__ vectors, const double *__restrict__ work) {
  // the problem this panel belongs to
  int b = 0;
  while (b + 1 < batch.count && batch.p[b + 1].first_panel <= static_cast<int>(blockIdx.x)) ++b;
  const Problem pr = batch.p[b];
  const int n = pr.n;
  const int c0 = (static_cast<int>(blockIdx.x) - pr.first_panel) * kPanel;
  const int width = min(kPanel, n - c0);
  const double *A = matrices + pr.matrix;
  const double *Z = work + pr.work;
  const double *tau = Z + 3 * int64_t(n) * n + 2 * n;
  __shared__ double Y[kEigMaxN][kPanel];
  __shared__ double part[kEigThreads / kPanel][kPanel];
  __shared__ double w[kPanel];
  const int t = threadIdx.x, c = t % kPanel, g = t / kPanel, groups = kEigThreads / kPanel;
  for (int i = g; i < n; i += groups) Y[i][c] = c < width ? Z[i + int64_t(c0 + c) * n] : 0.0;
  __syncthreads();
  for (int k = n - 2; k >= 0; --k) {
    const double tk = tau[k];
    if (tk == 0.0) continue;
    const double *v = A + int64_t(k) * n;   // v[k+1] = 1, v[k+2..] stored
    double sum = 0.0;
    for (int i = k + 1 + g; i < n; i += groups) sum += v[i] * Y[i][c];
    part[g][c] = sum;
    __syncthreads();
    if (t < kPanel) {
      double total = 0.0;
      for (int q = 0; q < groups; ++q) total += part[q][t];
      w[t] = tk * total;
    }
    __syncthreads();
    for (int i = k + 1 + g; i < n; i += groups) Y[i][c] -= v[i] * w[c];
    __syncthreads();
  }
  double *out = vectors + pr.matrix;
  if (c < width)
    for (int i = g; i < n; i += groups) out[int64_t(c0 + c) * n + i] = Y[i][c];
}

}  // namespace

extern "C" int64_t tt_eigh_max_order(void) { return kEigMaxN; }

extern "C" int64_t tt_eigh_workspace(const tt_eigh_problem *problems, int64_t count) {
  if (count < 0 || (count > 0 && !problems)) {
    fail("tt_eigh_workspace: bad problem list");
    return -1;
  }
  int64_t total = 0;
  for (int64_t b = 0; b < count; ++b) total += 3 * problems[b].n * problems[b].n + 3 * problems[b].n;
  return total;
}

extern "C" int tt_eigh_batched(const tt_eigh_problem *problems, int64_t count, double *matrices, double *values,
                               double *vectors, double *workspace, void *stream) {
  if (count < 0 || (count > 0 && !problems)) return fail("tt_eigh_batched: bad problem list");
  if (count == 0) return 0;
  if (!matrices || !values || !vectors || !workspace) return fail("tt_eigh_batched: null buffer");
  for (int64_t b = 0; b < count; ++b) {
    const tt_eigh_problem &p = problems[b];
    if (p.n < 1 || p.n > kEigMaxN)
      return fail("tt_eigh_batched: matrix order " + std::to_string(p.n) + " outside 1.." + std::to_string(kEigMaxN));
    if (p.matrix_offset < 0 || p.value_offset < 0) return fail("tt_eigh_batched: negative offset");
  }
  int64_t work = 0;
  for (int64_t first = 0; first < count; first += kEigBatch) {
    Batch batch{};
    batch.count = static_cast<int32_t>(std::min<int64_t>(kEigBatch, count - first));
    int32_t panels = 0, leaves = 0;
    for (int32_t b = 0; b < batch.count; ++b) {
      const tt_eigh_problem &p = problems[first + b];
      batch.p[b] = Problem{p.matrix_offset, p.value_offset, work, static_cast<int32_t>(p.n), panels, leaves, 0};
      work += 3 * p.n * p.n + 3 * p.n;
      panels += static_cast<int32_t>((p.n + kPanel - 1) / kPanel);
      leaves += static_cast<int32_t>((p.n + kLeaf - 1) / kLeaf);
    }
    hipStream_t s = as_stream(stream);
    int64_t largest = 0;
    for (int32_t b = 0; b < batch.count; ++b) largest = std::max<int64_t>(largest, batch.p[b].n);
    if (largest <= 257)
      hipLaunchKernelGGL(tridiag_kernel<1>, dim3(batch.count), dim3(kTriThreads), 0, s, batch, matrices, workspace);
    else if (largest <= 513)
      hipLaunchKernelGGL(tridiag_kernel<2>, dim3(batch.count), dim3(kTriThreads), 0, s, batch, matrices, workspace);
    else
      hipLaunchKernelGGL(tridiag_kernel<4>, dim3(batch.count), dim3(kTriThreads), 0, s, batch, matrices, workspace);
    TT_LAUNCHED();
    hipLaunchKernelGGL(leaf_kernel, dim3((leaves + kEigWaves - 1) / kEigWaves), dim3(kEigThreads), 0, s, batch, values,
                       workspace, leaves);
    TT_LAUNCHED();
    hipLaunchKernelGGL(dc_kernel, dim3(batch.count), dim3(kEigThreads), 0, s, batch, values, workspace);
    TT_LAUNCHED();
    hipLaunchKernelGGL(back_kernel, dim3(panels), dim3(kEigThreads), 0, s, batch, matrices, vectors, workspace);
    TT_LAUNCHED();
  }
  return 0;
}
