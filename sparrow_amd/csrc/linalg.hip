// Batched dense linear algebra of the SCF on gfx950 (FP64):
//   * symmetric eigensolver: Householder tridiagonalisation (one workgroup per molecule,
//     LAPACK dsytd2 formulation), Sturm-sequence bisection for every eigenvalue (one thread
//     per eigenvalue, Kahan / LAPACK dstebz), inverse iteration for the wanted eigenvectors
//     (LAPACK dstein: tridiagonal LU with partial pivoting, re-orthogonalisation inside clusters
//     of nearly equal eigenvalues) and back-transformation with the stored reflectors;
//   * GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4f64), used for P = 2 C_occ C_occ^T and
//     for the DIIS commutator F P;
//   * small SCF helpers: commutator, DIIS dot products / extrapolation, density change (RMSD).
// Every kernel takes the batch description and works on all molecules at once.
#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <string>

#include "linalg.hpp"

namespace sparrow_amd {

namespace {

constexpr int kMaxDiis = 16;  // largest DIIS subspace

__device__ inline double blockSum(double v, double* red) {
  // red: >= blockDim.x/64 doubles of LDS
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < nw; ++k) s += red[k];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------------------------
// Householder tridiagonalisation of the full symmetric A (column-major, both triangles kept).
// On exit: d (diagonal), e (sub-diagonal), tau and the reflector vectors below the sub-diagonal
// of A: H_k = I - tau_k u_k u_k^T, u_k = (0..0, 1, A[k+2:n, k]).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_sytrd(EigJob* jobs, int njobs) {
  const EigJob J = jobs[blockIdx.x];
  const int n = J.n;
  if (n <= 0) return;
  double* A = J.A;
  double* v = J.v;
  double* w = J.w;
  const int tid = threadIdx.x, T = blockDim.x;
  __shared__ double red[16];
  __shared__ double sc[3];
  __shared__ double part[1024];
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;
    double loc = 0.0;
    for (int i = k + 2 + tid; i < n; i += T) {
      double a = A[(long long)k * n + i];
      loc += a * a;
    }
    const double sigma = blockSum(loc, red);
    if (tid == 0) {
      const double alpha = A[(long long)k * n + k + 1];
      double t = 0.0, beta = alpha, scal = 0.0;
      if (sigma != 0.0) {
        beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
        t = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      sc[0] = t;
      sc[1] = scal;
      J.d[k] = A[(long long)k * n + k];
      J.e[k] = beta;
      J.tau[k] = t;
    }
    __syncthreads();
    const double tau = sc[0], scal = sc[1];
    if (tau == 0.0) continue;
    for (int i = k + 2 + tid; i < n; i += T) {
      double val = A[(long long)k * n + i] * scal;
      A[(long long)k * n + i] = val;
      v[i] = val;
    }
    if (tid == 0) v[k + 1] = 1.0;
    __syncthreads();
    // p = tau * A22 v  (rows k+1..n-1), split over column groups when the trailing block is small
    const int RB = min(T, ((m + 63) / 64) * 64);
    const int G = T / RB;
    if (G == 1) {
      for (int i = k + 1 + tid; i < n; i += T) {
        double s = 0.0;
        for (int j = k + 1; j < n; ++j) s += A[(long long)j * n + i] * v[j];
        w[i] = tau * s;
      }
    } else {
      const int ri = tid % RB, g = tid / RB;
      const int i = k + 1 + ri;
      double s = 0.0;
      if (i < n && g < G)
        for (int j = k + 1 + g; j < n; j += G) s += A[(long long)j * n + i] * v[j];
      if (g < G) part[g * RB + ri] = s;
      __syncthreads();
      if (g == 0 && i < n) {
        double t = 0.0;
        for (int q = 0; q < G; ++q) t += part[q * RB + ri];
        w[i] = tau * t;
      }
    }
    __syncthreads();
    double dl = 0.0;
    for (int i = k + 1 + tid; i < n; i += T) dl += w[i] * v[i];
    const double dot = blockSum(dl, red);
    const double K = -0.5 * tau * dot;
    for (int i = k + 1 + tid; i < n; i += T) w[i] += K * v[i];
    __syncthreads();
    // A22 -= v w^T + w v^T
    const long long mm = (long long)m * m;
    for (long long q = tid; q < mm; q += T) {
      const int jj = (int)(q / m), ii = (int)(q - (long long)jj * m);
      const int i = k + 1 + ii, j = k + 1 + jj;
      A[(long long)j * n + i] -= v[i] * w[j] + w[i] * v[j];
    }
    __syncthreads();
  }
  if (tid == 0) {
    J.d[n - 1] = A[(long long)(n - 1) * n + n - 1];
    J.e[n - 1] = 0.0;
    J.tau[n - 1] = 0.0;
  }
}

// ---------------------------------------------------------------------------------------------
// Fused variant: the rank-2 update of step k and the matrix-vector product of step k+1 share
// one pass over the trailing matrix (read + write once per element and step instead of
// read, read, write).  Order inside step k (column k already final):
//   1. update column k+1 with (v_k, w_k); 2. Householder vector v_{k+1} from it;
//   3. for columns j >= k+2: A[:,j] -= v w_j + w v_j and p += A[:,j] v_{k+1}[j];
//   4. w_{k+1} = tau p - 1/2 tau^2 (p.v) v.
// v, w, v2, p: four length-n scratch vectors (J.v = [v | v2], J.w = [w | p]).
// ---------------------------------------------------------------------------------------------
__device__ bool householderColumn(double* A, int n, int k, double* vout, double* red, double* sc, EigJob& J) {
  // reflector for column k (rows k+1..n-1); writes scaled vector into A and vout; returns tau != 0
  const int tid = threadIdx.x, T = blockDim.x;
  double loc = 0.0;
  for (int i = k + 2 + tid; i < n; i += T) {
    double a = A[(long long)k * n + i];
    loc += a * a;
  }
  const double sigma = blockSum(loc, red);
  if (tid == 0) {
    const double alpha = A[(long long)k * n + k + 1];
    double t = 0.0, beta = alpha, scal = 0.0;
    if (sigma != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    sc[0] = t;
    sc[1] = scal;
    J.d[k] = A[(long long)k * n + k];
    J.e[k] = beta;
    J.tau[k] = t;
  }
  __syncthreads();
  const double tau = sc[0], scal = sc[1];
  if (tau != 0.0) {
    for (int i = k + 2 + tid; i < n; i += T) {
      double val = A[(long long)k * n + i] * scal;
      A[(long long)k * n + i] = val;
      vout[i] = val;
    }
    if (tid == 0) vout[k + 1] = 1.0;
  }
  __syncthreads();
  return tau != 0.0;
}

__global__ __launch_bounds__(1024) void k_sytrd_fused(EigJob* jobs, int njobs) {
  const EigJob Jc = jobs[blockIdx.x];
  EigJob J = Jc;
  const int n = J.n;
  if (n <= 0) return;
  double* A = J.A;
  double* va = J.v;      // current reflector v_k
  double* vb = J.v + n;  // next reflector v_{k+1}
  double* wa = J.w;      // w_k
  double* pb = J.w + n;  // accumulating p_{k+1}
  const int tid = threadIdx.x, T = blockDim.x;
  __shared__ double red[16];
  __shared__ double sc[3];
  __shared__ double part[1024];
  if (n == 1) {
    if (tid == 0) {
      J.d[0] = A[0];
      J.e[0] = 0.0;
      J.tau[0] = 0.0;
    }
    return;
  }
  // prologue: reflector 0 and p_0 = A22 v_0 by a plain product
  bool haveTau = householderColumn(A, n, 0, va, red, sc, J);
  double tau = sc[0];
  if (haveTau) {
    const int m = n - 1;
    const int RB = min(T, ((m + 63) / 64) * 64), G = T / RB;
    const int ri = tid % RB, g = tid / RB;
    for (int i0 = 1; i0 < n; i0 += RB) {
      const int i = i0 + ri;
      double s = 0.0;
      if (i < n && g < G)
        for (int j = 1 + g; j < n; j += G) s += A[(long long)j * n + i] * va[j];
      if (G > 1) {
        if (g < G) part[g * RB + ri] = s;
        __syncthreads();
        if (g == 0 && i < n) {
          double t = 0.0;
          for (int q = 0; q < G; ++q) t += part[q * RB + ri];
          wa[i] = tau * t;
        }
        __syncthreads();
      } else if (i < n) {
        wa[i] = tau * s;
      }
    }
    __syncthreads();
  }
  for (int k = 0; k + 1 < n; ++k) {
    // finish w_k = p - 1/2 tau (p.v) v
    if (haveTau) {
      double dl = 0.0;
      for (int i = k + 1 + tid; i < n; i += T) dl += wa[i] * va[i];
      const double dot = blockSum(dl, red);
      const double K = -0.5 * tau * dot;
      for (int i = k + 1 + tid; i < n; i += T) wa[i] += K * va[i];
      __syncthreads();
    }
    if (k + 2 >= n) {
      // last step: 1x1 trailing block
      if (haveTau && tid == 0) {
        const int i = n - 1;
        A[(long long)i * n + i] -= 2.0 * va[i] * wa[i];
      }
      __syncthreads();
      break;
    }
    // 1. update column k+1 (rows k+1..n-1)
    if (haveTau) {
      const double vk1 = va[k + 1], wk1 = wa[k + 1];
      for (int i = k + 1 + tid; i < n; i += T) A[(long long)(k + 1) * n + i] -= va[i] * wk1 + wa[i] * vk1;
    }
    __syncthreads();
    // 2. next reflector from column k+1
    const bool nextTau = householderColumn(A, n, k + 1, vb, red, sc, J);
    const double tauN = sc[0];
    // 3. fused update of columns >= k+2 and p = A' v_{k+1} over rows >= k+2
    const int m = n - k - 2;
    const int RB = min(T, ((m + 63) / 64) * 64), G = T / RB;
    const int ri = tid % RB, g = tid / RB;
    for (int i0 = k + 2; i0 < n; i0 += RB) {
      const int i = i0 + ri;
      double s = 0.0;
      if (i < n && g < G) {
        const double vi = haveTau ? va[i] : 0.0, wi = haveTau ? wa[i] : 0.0;
        for (int j = k + 2 + g; j < n; j += G) {
          double* a = A + (long long)j * n + i;
          double x = *a;
          if (haveTau) {
            x -= vi * wa[j] + wi * va[j];
            *a = x;
          }
          if (nextTau) s += x * vb[j];
        }
      }
      if (G > 1) {
        if (g < G) part[g * RB + ri] = s;
        __syncthreads();
        if (g == 0 && i < n) {
          double t = 0.0;
          for (int q = 0; q < G; ++q) t += part[q * RB + ri];
          pb[i] = tauN * t;
        }
        __syncthreads();
      } else if (i < n) {
        pb[i] = tauN * s;
      }
    }
    __syncthreads();
    // rotate buffers: v_{k+1} -> current, p -> current w
    double* t1 = va;
    va = vb;
    vb = t1;
    double* t2 = wa;
    wa = pb;
    pb = t2;
    haveTau = nextTau;
    tau = tauN;
  }
  if (tid == 0) {
    J.d[n - 1] = A[(long long)(n - 1) * n + n - 1];
    J.e[n - 1] = 0.0;
    J.tau[n - 1] = 0.0;
  }
}

// ---------------------------------------------------------------------------------------------
// The same tridiagonalisation for one large matrix, each column step split into two launches so
// that the trailing matrix is swept by many workgroups (the host loops over r = 0 .. n-2):
//   col(r), one workgroup: p_{r-1} = tau_{r-1} * (the partial products of upd(r-1) summed in
//     tile order, so the result is deterministic), w_{r-1} = p - 1/2 tau (p.v) v (rows >= r);
//     column r -= v w_r + w v_r (rows >= r) with (v, w) of reflector r - 1; reflector r from
//     column r (d, e, tau, v_r into A and V[r % 2]).
//   upd(r), 64 x 64 tiles of rows and columns >= r + 1: A -= v w^T + w v^T with (v, w) of
//     reflector r - 1, and the partial products sum_j A(i, j) v_r(j) of the tile's 64 columns
//     into part[tile column][i].
// Same operations as k_sytrd_fused in the same order per element; V = J.v (2n), W = J.w (n),
// part = J.work (ceil((n - 1) / 64) x n).  No workgroup waits for another: a device-scope
// hand-over inside one launch would need an L2 write-back per workgroup, which costs more than
// the kernel boundary it saves.
// ---------------------------------------------------------------------------------------------
constexpr int kTrdTile = 64;

__device__ void sytrdColumn(EigJob& J, int r, double* red, double* sc) {
  const int n = J.n;
  double* A = J.A;
  double* vPrev = J.v + ((r + 1) % 2) * n;  // v_{r-1}
  double* vCur = J.v + (r % 2) * n;         // v_r (written here)
  double* W = J.w;                          // w_{r-1}
  const double* part = J.work;
  const int tid = threadIdx.x, T = blockDim.x;
  if (r >= 1 && J.tau[r - 1] != 0.0) {
    const double tau = J.tau[r - 1];
    const int R = (n - r + kTrdTile - 1) / kTrdTile;  // tile columns of upd(r - 1), rows >= r
    double dl = 0.0;
    for (int i = r + tid; i < n; i += T) {
      double t = 0.0;
#pragma unroll 4
      for (int q = 0; q < R; ++q) t += part[(long long)q * n + i];
      const double p = tau * t;
      W[i] = p;
      dl += p * vPrev[i];
    }
    const double dot = blockSum(dl, red);
    const double K = -0.5 * tau * dot;
    for (int i = r + tid; i < n; i += T) W[i] += K * vPrev[i];
    __syncthreads();
    const double vr = vPrev[r], wr = W[r];
    for (int i = r + tid; i < n; i += T) A[(long long)r * n + i] -= vPrev[i] * wr + W[i] * vr;
    __syncthreads();
  }
  householderColumn(A, n, r, vCur, red, sc, J);
}

__global__ __launch_bounds__(1024) void k_sytrd_col(EigJob* jobs, int r) {
  EigJob J = jobs[0];
  __shared__ double red[16];
  __shared__ double sc[3];
  sytrdColumn(J, r, red, sc);
}

__global__ __launch_bounds__(1024) void k_sytrd_upd(EigJob* jobs, int r) {
  EigJob J = jobs[0];
  const int n = J.n;
  double* __restrict__ A = J.A;
  double* part = J.work;
  const bool upd = r >= 1 && J.tau[r - 1] != 0.0, prod = J.tau[r] != 0.0;
  const int lane = threadIdx.x & 63;
  const int i = r + 1 + blockIdx.x * kTrdTile + lane;
  const int g = threadIdx.x >> 6;  // column group: tile columns g, g + 16, g + 32, g + 48
  const int j0 = r + 1 + blockIdx.y * kTrdTile;
  __shared__ double sv[3][kTrdTile];  // v_{r-1}, w_{r-1}, v_r at the tile's columns
  __shared__ double sp[16][64];
  if (threadIdx.x < kTrdTile) {
    const int j = j0 + threadIdx.x;
    const int jv = (r + 1) % 2, jc = r % 2;
    sv[0][threadIdx.x] = j < n ? J.v[jv * n + j] : 0.0;
    sv[1][threadIdx.x] = j < n ? J.w[j] : 0.0;
    sv[2][threadIdx.x] = j < n ? J.v[jc * n + j] : 0.0;
  }
  const double vi = (upd && i < n) ? J.v[((r + 1) % 2) * n + i] : 0.0;
  const double wi = (upd && i < n) ? J.w[i] : 0.0;
  double x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = j0 + g + 16 * q;
    x[q] = (i < n && j < n) ? A[(long long)j * n + i] : 0.0;
  }
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int jj = g + 16 * q, j = j0 + jj;
    if (i < n && j < n) {
      if (upd) {
        x[q] -= vi * sv[1][jj] + wi * sv[0][jj];
        A[(long long)j * n + i] = x[q];
      }
      if (prod) s += x[q] * sv[2][jj];
    }
  }
  sp[g][lane] = s;
  __syncthreads();
  if (g == 0 && i < n) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += sp[q][lane];
    part[(long long)blockIdx.y * n + i] = t;
  }
}

__global__ void k_sytrd_final(EigJob* jobs) {
  const EigJob J = jobs[0];
  const int n = J.n;
  if (threadIdx.x == 0) {
    J.d[n - 1] = J.A[(long long)(n - 1) * n + n - 1];
    J.e[n - 1] = 0.0;
    J.tau[n - 1] = 0.0;
  }
}

// ---------------------------------------------------------------------------------------------
// Eigenvalues of the tridiagonal (d, e) by bisection on the Sturm count; one thread per index.
// ---------------------------------------------------------------------------------------------
__device__ int sturmCount(const double* d, const double* e, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// Multisection: one wave per eigenvalue index j; each step the 64 lanes take the Sturm count at
// 64 interior points of [a, b] (spacing (b - a) / 65) and the wave keeps the sub-interval that
// holds eigenvalue j (count(a) <= j < count(b) throughout), so ~9 steps reach the stopping width
// max(2 eps ||T||, 2 eps max(|a|, |b|)) that plain bisection (LAPACK dstebz) reaches in ~55.
__global__ __launch_bounds__(256) void k_stebz(EigJob* jobs, int njobs) {
  const EigJob J = jobs[blockIdx.y];
  const int n = J.n;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;  // whole waves
  const double* __restrict__ d = J.d;
  const double* __restrict__ e = J.e;
  double lo = DBL_MAX, hi = -DBL_MAX, emax2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i + 1 < n) emax2 = fmax(emax2, e[i] * e[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    emax2 = fmax(emax2, __shfl_xor(emax2, o, 64));
  }
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  lo -= 2.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;
  hi += 2.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;
  const double atol = 2.0 * DBL_EPSILON * tnorm;
  double a = lo, b = hi;
  for (int it = 0; it < 64; ++it) {
    if (b - a <= fmax(atol, 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b)))) break;
    const double x = a + (double)(lane + 1) * ((b - a) * (1.0 / 65.0));
    const bool above = x >= b || (x > a && sturmCount(d, e, n, x, pivmin) > j);
    const unsigned long long m = __ballot(above);
    double na = a, nb = b;
    if (m) {
      const int k = __ffsll((long long)m) - 1;  // lowest lane with count > j
      nb = fmin(b, __shfl(x, k, 64));
      if (k > 0) na = fmax(a, __shfl(x, k - 1, 64));
    } else {
      na = fmax(a, __shfl(x, 63, 64));
    }
    if (na == a && nb == b) break;  // no representable progress
    a = na;
    b = nb;
  }
  if (lane == 0) J.wval[j] = 0.5 * (a + b);
}

// ---------------------------------------------------------------------------------------------
// Inverse iteration for eigenvectors 0..nvec-1 of the tridiagonal (LAPACK dstein scheme).
// One thread per cluster of eigenvalues closer than ortol (see below); inside a cluster the
// eigenvalues are separated by at least pertol = 10 eps ||T|| and the vectors are
// re-orthogonalised (modified Gram-Schmidt).
// Output: Z (n x nvec, row-major: Z[i*nvec + j]).
// ---------------------------------------------------------------------------------------------
__device__ inline double startValue(unsigned j, unsigned i) {
  unsigned x = j * 2654435761u ^ (i + 0x9e3779b9u) * 2246822519u;
  x ^= x >> 15;
  x *= 2246822519u;
  x ^= x >> 13;
  x *= 3266489917u;
  x ^= x >> 16;
  return (double)(x & 0xffffff) / (double)0x800000 - 1.0;
}

__global__ __launch_bounds__(256) void k_stein(EigJob* jobs, int njobs) {
  const EigJob J = jobs[blockIdx.y];
  const int n = J.n, nv = J.nvec;
  const int j0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (j0 >= nv) return;
  const double* __restrict__ d = J.d;
  const double* __restrict__ e = J.e;
  const double* wv = J.wval;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) tnorm = fmax(tnorm, fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0));
  // Inverse-iteration vectors of eigenvalues separated by g are orthogonal to ~eps*||T||/g, so
  // only gaps below 1e-7 ||T|| (orthogonality loss up to ~2e-9) need re-orthogonalisation.
  // LAPACK's 1e-3 ||T|| would chain the dense spectra of large molecules into serial clusters.
  const double ortol = 1e-7 * tnorm;
  const double pertol = 10.0 * DBL_EPSILON * tnorm;
  const double pivtol = DBL_EPSILON * tnorm > 0 ? DBL_EPSILON * tnorm : DBL_MIN;
  if (j0 > 0 && wv[j0] - wv[j0 - 1] <= ortol) return;  // not the first of its cluster
  int j1 = j0 + 1;
  while (j1 < nv && wv[j1] - wv[j1 - 1] <= ortol) ++j1;
  // work arrays (slot j0): sub/diag/super/super2 of the LU factors, pivots, iterate
  const long long ld = nv;
  double* fl = J.work + 0 * (long long)n * ld;
  double* fd = J.work + 1 * (long long)n * ld;
  double* fu = J.work + 2 * (long long)n * ld;
  double* fu2 = J.work + 3 * (long long)n * ld;
  double* x = J.work + 4 * (long long)n * ld;
  unsigned char* piv = J.ipiv;
  auto W = [&](double* a, int i) -> double& { return a[(long long)i * ld + j0]; };
  double* Z = J.Z;
  double lamPrev = 0.0;
  for (int jj = j0; jj < j1; ++jj) {
    double lam = wv[jj];
    if (jj > j0 && lam - lamPrev < pertol) lam = lamPrev + pertol;
    lamPrev = lam;
    // LU factorisation with partial pivoting of T - lam I (LAPACK dgttrf); the running diagonal
    // and superdiagonal entries of the row being eliminated stay in registers
    {
      double dc = d[0] - lam, uc = n > 1 ? e[0] : 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        const double li = e[i], dn = d[i + 1] - lam, un = i + 2 < n ? e[i + 1] : 0.0;
        if (fabs(dc) >= fabs(li)) {
          piv[(long long)i * ld + j0] = 0;
          const double f = dc != 0.0 ? li / dc : 0.0;
          W(fl, i) = f;
          W(fd, i) = fabs(dc) < pivtol ? (dc < 0.0 ? -pivtol : pivtol) : dc;
          W(fu, i) = uc;
          W(fu2, i) = 0.0;
          dc = dn - f * uc;
          uc = un;
        } else {
          piv[(long long)i * ld + j0] = 1;
          const double f = dc / li;
          W(fl, i) = f;
          W(fd, i) = fabs(li) < pivtol ? (li < 0.0 ? -pivtol : pivtol) : li;
          W(fu, i) = dn;
          W(fu2, i) = un;
          dc = uc - f * dn;
          uc = -f * un;
        }
      }
      W(fd, n - 1) = fabs(dc) < pivtol ? (dc < 0.0 ? -pivtol : pivtol) : dc;
    }
    for (int i = 0; i < n; ++i) W(x, i) = startValue(jj, i);
    for (int iter = 0; iter < 3; ++iter) {
      // solve L U y = x (LAPACK dgtts2), the carried component in a register
      {
        double xc = W(x, 0);
        for (int i = 0; i + 1 < n; ++i) {
          const double xn = W(x, i + 1), f = W(fl, i);
          if (piv[(long long)i * ld + j0] == 0) {
            W(x, i) = xc;
            xc = xn - f * xc;
          } else {
            W(x, i) = xn;
            xc = xc - f * xn;
          }
        }
        W(x, n - 1) = xc;
      }
      {
        double x1 = W(x, n - 1) / W(fd, n - 1), x2 = 0.0;
        W(x, n - 1) = x1;
        if (n > 1) {
          x2 = x1;
          x1 = (W(x, n - 2) - W(fu, n - 2) * x2) / W(fd, n - 2);
          W(x, n - 2) = x1;
        }
        for (int i = n - 3; i >= 0; --i) {
          const double xi = (W(x, i) - W(fu, i) * x1 - W(fu2, i) * x2) / W(fd, i);
          W(x, i) = xi;
          x2 = x1;
          x1 = xi;
        }
      }
      // re-orthogonalise against the previous vectors of the cluster
      for (int kk = j0; kk < jj; ++kk) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += Z[(long long)i * nv + kk] * W(x, i);
        for (int i = 0; i < n; ++i) W(x, i) -= s * Z[(long long)i * nv + kk];
      }
      double nrm = 0.0, big = 0.0;
      for (int i = 0; i < n; ++i) big = fmax(big, fabs(W(x, i)));
      if (big == 0.0) big = 1.0;
      for (int i = 0; i < n; ++i) {
        double t = W(x, i) / big;
        nrm += t * t;
      }
      const double inv = 1.0 / (big * sqrt(nrm));
      for (int i = 0; i < n; ++i) W(x, i) *= inv;
    }
    // sign convention: the component of largest magnitude is positive
    int imax = 0;
    double amax = -1.0;
    for (int i = 0; i < n; ++i)
      if (fabs(W(x, i)) > amax) {
        amax = fabs(W(x, i));
        imax = i;
      }
    const double sgn = W(x, imax) < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) Z[(long long)i * nv + jj] = sgn * W(x, i);
  }
}

// ---------------------------------------------------------------------------------------------
// Back-transformation Z <- H_0 H_1 ... H_{n-3} Z (LAPACK dormtr, unblocked), one thread per column.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_ormtr(EigJob* jobs, int njobs) {
  const EigJob J = jobs[blockIdx.y];
  const int n = J.n, nv = J.nvec;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const double* A = J.A;
  double* Z = J.Z;
  for (int k = n - 3; k >= 0; --k) {
    const double tau = J.tau[k];
    if (tau == 0.0) continue;
    double s = Z[(long long)(k + 1) * nv + j];
    for (int i = k + 2; i < n; ++i) s += A[(long long)k * n + i] * Z[(long long)i * nv + j];
    s *= tau;
    Z[(long long)(k + 1) * nv + j] -= s;
    for (int i = k + 2; i < n; ++i) Z[(long long)i * nv + j] -= A[(long long)k * n + i] * s;
  }
}

// ---------------------------------------------------------------------------------------------
// Blocked back-transformation (compact WY form of LAPACK dormtr / dlarft, forward, columnwise):
// the reflectors k0 .. k1-1 of block b (kOrmBlock of them) multiply to
//   H_k0 H_k0+1 ... H_k1-1 = I - V T V^T,
// V(:, j) = u_{k0+j} (1 at row k0+j+1, A[k0+j+2:n, k0+j] below, zero above), T upper triangular:
//   T(j, j) = tau_j,  T(0:j, j) = -tau_j T(0:j, 0:j) V(:, 0:j)^T V(:, j).
// Z <- H_0 ... H_{n-3} Z applies the blocks last to first, Z <- Z - V (T (V^T Z)); one launch
// per block, each workgroup a tile of kOrmCols columns of Z (row-major n x nvec) of one job.
// Launch `step` applies, for each job, its block nblocks - 1 - step (nothing if negative).
// ---------------------------------------------------------------------------------------------
constexpr int kOrmBlock = 32, kOrmCols = 32;

__global__ __launch_bounds__(256) void k_ormtr_blocked(EigJob* jobs, int njobs, int step) {
  const EigJob J = jobs[blockIdx.y];
  const int n = J.n, nv = J.nvec;
  const int nrefl = n - 2;  // reflectors 0 .. n-3
  if (nrefl <= 0) return;
  const int nblk = (nrefl + kOrmBlock - 1) / kOrmBlock;
  const int b = nblk - 1 - step;
  const int c0 = blockIdx.x * kOrmCols;
  if (b < 0 || c0 >= nv) return;
  const int k0 = b * kOrmBlock, m = min(kOrmBlock, nrefl - k0);
  const int nc = min(kOrmCols, nv - c0);
  const double* A = J.A;
  double* Z = J.Z;
  const int tid = threadIdx.x;
  __shared__ double G[kOrmBlock][kOrmBlock + 1];  // V^T V (strictly upper part used)
  __shared__ double T[kOrmBlock][kOrmBlock + 1];
  __shared__ double W[kOrmBlock][kOrmCols + 1];   // V^T Z, then T V^T Z
  __shared__ double W2[kOrmBlock][kOrmCols + 1];
  __shared__ double tau[kOrmBlock];
  // V(r, j): implicit unit at r = k0 + j + 1, the stored tail below it
  auto V = [&](int r, int j) -> double {
    const int top = k0 + j + 1;
    if (r < top) return 0.0;
    if (r == top) return 1.0;
    return A[(long long)(k0 + j) * n + r];
  };
  if (tid < m) tau[tid] = J.tau[k0 + tid];
  // 1. G(i, j) = V(:, i)^T V(:, j) for i < j (rows from k0 + j + 1: V(:, j) is zero above)
  for (int q = tid; q < m * m; q += blockDim.x) {
    const int i = q / m, j = q - i * m;
    if (i >= j) continue;
    double acc = 0.0;
    for (int r = k0 + j + 1; r < n; ++r) acc += V(r, i) * V(r, j);
    G[i][j] = acc;
  }
  __syncthreads();
  // 2. T column by column (tau_j = 0 gives a zero column: H_j = I)
  for (int j = 0; j < m; ++j) {
    if (tid < j) {  // T(tid, j) = -tau_j sum_{l < j} T(tid, l) G(l, j), T upper triangular
      double acc = 0.0;
      for (int l = tid; l < j; ++l) acc += T[tid][l] * G[l][j];
      T[tid][j] = -tau[j] * acc;
    } else if (tid < m) {
      T[tid][j] = tid == j ? tau[j] : 0.0;
    }
    __syncthreads();
  }
  // 3. W = V^T Z(:, c0 : c0 + nc)
  for (int q = tid; q < m * nc; q += blockDim.x) {
    const int j = q / nc, c = q - j * nc;
    double acc = 0.0;
    for (int r = k0 + j + 1; r < n; ++r) acc += V(r, j) * Z[(long long)r * nv + c0 + c];
    W[j][c] = acc;
  }
  __syncthreads();
  // 4. W2 = T W
  for (int q = tid; q < m * nc; q += blockDim.x) {
    const int i = q / nc, c = q - i * nc;
    double acc = 0.0;
    for (int l = i; l < m; ++l) acc += T[i][l] * W[l][c];
    W2[i][c] = acc;
  }
  __syncthreads();
  // 5. Z(r, c) -= sum_j V(r, j) W2(j, c) for rows r >= k0 + 1
  for (int q = tid; q < (n - k0 - 1) * nc; q += blockDim.x) {
    const int r = k0 + 1 + q / nc, c = q % nc;
    double acc = 0.0;
    const int jmax = min(m, r - k0);  // V(r, j) = 0 for j > r - k0 - 1
    for (int j = 0; j < jmax; ++j) acc += V(r, j) * W2[j][c];
    Z[(long long)r * nv + c0 + c] -= acc;
  }
}

// ---------------------------------------------------------------------------------------------
// Back-transformation of one large job on the matrix cores: blocks of kOrmGemm reflectors, each
// Z <- Z - (V T) (V^T Z) by two FP64-MFMA GEMMs.  k_ormtr_prep (one workgroup per block, all
// blocks at once) writes the explicit reflector columns row-major into Vr (Vr[r n + k] = u_k(r),
// unit at r = k + 1, zero above) and V T column-major into VT (VT[k n + r], rows r >= k0 + 1),
// with T from the same recurrence as k_ormtr_blocked.
// ---------------------------------------------------------------------------------------------
constexpr int kOrmGemm = 64;

__global__ __launch_bounds__(256) void k_ormtr_prep(EigJob* jobs, double* __restrict__ Vr, double* __restrict__ VT) {
  const EigJob J = jobs[0];
  const int n = J.n, nrefl = n - 2;
  const int k0 = blockIdx.x * kOrmGemm, m = min(kOrmGemm, nrefl - k0);
  if (m <= 0) return;
  const double* A = J.A;
  const int tid = threadIdx.x;
  // one array: T (upper triangle with the diagonal) and G = V^T V transposed in the strict lower
  // triangle (GT[j][i] = G(i, j), i < j)
  __shared__ double T[kOrmGemm][kOrmGemm + 1];
  __shared__ double tau[kOrmGemm];
  auto V = [&](int r, int j) -> double {
    const int top = k0 + j + 1;
    if (r < top) return 0.0;
    if (r == top) return 1.0;
    return A[(long long)(k0 + j) * n + r];
  };
  for (long long q = tid; q < (long long)n * m; q += blockDim.x) {
    const int r = (int)(q / m), j = (int)(q - (long long)r * m);
    Vr[(long long)r * n + k0 + j] = V(r, j);
  }
  if (tid < m) tau[tid] = J.tau[k0 + tid];
  for (int q = tid; q < m * m; q += blockDim.x) {
    const int i = q / m, j = q - i * m;
    if (i >= j) continue;
    double acc = 0.0;
    for (int r = k0 + j + 1; r < n; ++r) acc += V(r, i) * V(r, j);
    T[j][i] = acc;
  }
  __syncthreads();
  for (int j = 0; j < m; ++j) {  // column j of T reads columns < j of T and row j of GT
    if (tid < j) {
      double acc = 0.0;
      for (int l = tid; l < j; ++l) acc += T[tid][l] * T[j][l];
      T[tid][j] = -tau[j] * acc;
    } else if (tid == j) {
      T[j][j] = tau[j];
    }
    __syncthreads();
  }
  const int rows = n - k0 - 1;
  for (long long q = tid; q < (long long)rows * m; q += blockDim.x) {
    const int i = (int)(q / rows), r = k0 + 1 + (int)(q - (long long)i * rows);
    const int lmax = min(i, r - k0 - 1);  // V(r, l) = 0 for l > r - k0 - 1
    double acc = 0.0;
    for (int l = 0; l <= lmax; ++l) acc += V(r, l) * T[l][i];
    VT[(long long)(k0 + i) * n + r] = acc;
  }
}

// ---------------------------------------------------------------------------------------------
// GEMM C = alpha * A * B^T + beta * D with arbitrary element strides, FP64 matrix cores.
// A(i,k) = A[i*sAi + k*sAk], B(j,k) = B[j*sBj + k*sBk], C(i,j) = C[i*sCi + j*sCj], D likewise.
// alpha == 0 skips the products (C = beta * D).  D may alias C.
// TILE x TILE output per workgroup of 4 waves; each wave owns a (TILE/2)^2 quarter made of
// (TILE/32)^2 MFMA 16x16 blocks.
// MFMA f64 16x16x4 operands: A lane l -> (row l%16, k l/16); B lane l -> (k l/16, col l%16);
// D lane l register r -> (row l/16 + 4*r, col l%16)  (verified against numpy on gfx950).
// ---------------------------------------------------------------------------------------------
typedef double double4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const double gdouble;  // global-memory double
typedef __attribute__((address_space(1))) double gdoubleW;       // (written)
typedef __attribute__((address_space(1))) const unsigned short gbf16c;  // bf16 bits, read
typedef __attribute__((address_space(1))) unsigned short gbf16w;        // bf16 bits, written
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const bf16x8 gbf16x8;  // 8 bf16, one 16-byte load
typedef unsigned short ushort8 __attribute__((ext_vector_type(8)));  // 8 bf16 bit patterns
typedef __attribute__((address_space(1))) ushort8 gushort8;        // one 16-byte store

// v = hi + lo in bf16 (about 16 significant bits), from its FP32 rounding
__device__ inline void splitBf16(double v, unsigned short& hi, unsigned short& lo) {
  const float f = (float)v;
  const __bf16 h = (__bf16)f;
  const __bf16 l = (__bf16)(f - (float)h);
  hi = __builtin_bit_cast(unsigned short, h);
  lo = __builtin_bit_cast(unsigned short, l);
}

__device__ inline double bf16Value(unsigned short b) { return (double)__builtin_bit_cast(float, (unsigned)b << 16); }

// X(i, j) of a matrix held as split copies (element index idx = j ldp + i): hi + lo, exact in FP64
__device__ inline double splitValue(const gbf16c* h, const gbf16c* l, long long idx) {
  return bf16Value(h[idx]) + bf16Value(l[idx]);
}

// stores split(v) and returns the value the copies hold
__device__ inline double splitStore(double v, gbf16w* h, gbf16w* l) {
  unsigned short a, b;
  splitBf16(v, a, b);
  *h = a;
  *l = b;
  return bf16Value(a) + bf16Value(b);
}

template <int TILE>
__global__ __launch_bounds__(256) void k_gemm_f64(const GemmJob* __restrict__ jobs, const int* __restrict__ tileStart, int njobs) {
  constexpr int NB = TILE / 32;  // MFMA blocks per wave per dimension
  int lo = 0, hi = njobs - 1;
  const int t = blockIdx.x;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tileStart[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const GemmJob G = jobs[lo];
  const double alpha = G.alpha, beta = G.beta;
  const int tilesN = (G.N + TILE - 1) / TILE;
  const int local = t - tileStart[lo];
  const int ti = local / tilesN, tj = local - ti * tilesN;
  const int row0 = ti * TILE, col0 = tj * TILE;
  __shared__ double sA[TILE][17];
  __shared__ double sB[TILE][17];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * (TILE / 2), wc = (wv & 1) * (TILE / 2);
  double4v acc[NB][NB];
  for (int a = 0; a < NB; ++a)
    for (int c = 0; c < NB; ++c) acc[a][c] = (double4v){0.0, 0.0, 0.0, 0.0};
  if (alpha != 0.0) {
    for (int k0 = 0; k0 < G.K; k0 += 16) {
      for (int q = tid; q < TILE * 16; q += 256) {
        const int r = q % TILE, kk = q / TILE;
        const int gi = row0 + r, gk = k0 + kk;
        sA[r][kk] = (gi < G.M && gk < G.K) ? G.A[(long long)gi * G.sAi + (long long)gk * G.sAk] : 0.0;
        const int gj = col0 + r;
        sB[r][kk] = (gj < G.N && gk < G.K) ? G.B[(long long)gj * G.sBj + (long long)gk * G.sBk] : 0.0;
      }
      __syncthreads();
      for (int ks = 0; ks < 16; ks += 4) {
        double af[NB], bf[NB];
        for (int a = 0; a < NB; ++a) af[a] = sA[wr + a * 16 + (lane & 15)][ks + (lane >> 4)];
        for (int c = 0; c < NB; ++c) bf[c] = sB[wc + c * 16 + (lane & 15)][ks + (lane >> 4)];
        for (int a = 0; a < NB; ++a)
          for (int c = 0; c < NB; ++c) acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[c], acc[a][c], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  for (int a = 0; a < NB; ++a)
    for (int c = 0; c < NB; ++c)
      for (int r = 0; r < 4; ++r) {
        const int gi = row0 + wr + a * 16 + (lane >> 4) + 4 * r;
        const int gj = col0 + wc + c * 16 + (lane & 15);
        if (gi < G.M && gj < G.N) {
          double v = alpha * acc[a][c][r];
          if (beta != 0.0) v += beta * G.D[(long long)gi * G.sDi + (long long)gj * G.sDj];
          G.C[(long long)gi * G.sCi + (long long)gj * G.sCj] = v;
        }
      }
}

// ---------------------------------------------------------------------------------------------
// SCF helpers.  All operate on the molecule list `mols` with matrices at matOff.
// ---------------------------------------------------------------------------------------------
// dst = sum_s coef[mol*slots + s] * hist_s
__global__ __launch_bounds__(256) void k_diis_extrapolate(const MolDev* __restrict__ mols, const double* __restrict__ hist,
                                                          int slots, const double* __restrict__ coef, double* __restrict__ dst) {
  const int mol = blockIdx.y;
  const MolDev m = mols[mol];
  const long long nn = (long long)m.n * m.n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (long long)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int s = 0; s < slots; ++s) {
      const double c = coef[m.id * slots + s];
      if (c != 0.0) v += c * hist[m.histOff + s * nn + q];
    }
    dst[m.matOff + q] = v;
  }
}

// copy per-molecule n x n matrices: src/dst located at matOff (mode 0) or histOff + slot*n*n (mode 1)
__global__ __launch_bounds__(256) void k_copy_mats(const MolDev* __restrict__ mols, const double* __restrict__ src, int srcMode,
                                                   double* __restrict__ dst, int dstMode, const ScfMolDev* __restrict__ scf,
                                                   int slots) {
  const int mol = blockIdx.y;
  const MolDev m = mols[mol];
  const long long nn = (long long)m.n * m.n;
  const int slot = (srcMode || dstMode) ? (scf[m.id].iter - 1) % slots : 0;
  const double* s = src + (srcMode ? m.histOff + slot * nn : m.matOff);
  double* d = dst + (dstMode ? m.histOff + slot * nn : m.matOff);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (long long)gridDim.x * blockDim.x) d[q] = s[q];
}

// ---------------------------------------------------------------------------------------------
// Density matrix by second-order spectral projection (SP2; A. M. N. Niklasson, Phys. Rev. B 66,
// 155115 (2002)).  X0 = (e_max I - F) / (e_max - e_min) maps the spectrum of F into [0, 1]
// with the occupied states near 1; each step replaces X by X^2 or 2X - X^2, whichever moves
// tr X towards N_occ, until X is idempotent.  Then P = 2 X, the same projector as
// 2 C_occ C_occ^T of a diagonalisation, built from GEMMs only.  Gershgorin discs bound the
// spectrum.  The controller runs on the device, one thread per molecule, so that steps of all
// molecules can be issued without host round trips; converged molecules turn into copies.
// ---------------------------------------------------------------------------------------------
// Stopping rule (after E. H. Rubensson's parameterless criterion for recursive expansions): the
// idempotency error e = tr X - tr X^2 of two consecutive steps with different polynomials falls
// quadratically (each eigenvalue's distance from 0 or 1 is squared, times 2 or 4) once e < 1/4,
// so e_k >= e_{k-2} after such a pair means rounding noise dominates: X_k is idempotent to
// working precision.  Pairs of equal polynomials are not compared: they double one side's
// distances.  e < 1e-2 keeps the rule out of the first steps, where eigenvalues near 1/2 can
// make e grow.
// Low-precision densities (early SCF iterations of the mixed form): the steps run in split-BF16
// x3 (products good to about 2^-16) until the same rule sees that precision's noise floor
// (e < 1e-2 + 1e-4 n) or e < 1e-4, then exactly two FP64 steps purify X (eigenvalue errors
// d -> ~d^2 each) and the projection ends; the occupation count must then be right.
__device__ inline void sp2Decide(Sp2Dev& S, double t1, double t2) {
  const double e = t1 - t2;
  S.t1 = t1;
  S.t2 = t2;
  if (!S.done) {
    const double eFlat = S.lowp ? 1e-2 + 1e-4 * (double)S.n : 1e-2;
    const bool flat = S.step >= 2 && S.poly != S.prevPoly && e < eFlat && e >= S.ePrev2;
    const bool exact = fabs(e) <= 1e-14 * (double)S.n;
    if (S.lowp && (flat || exact || e < 1e-4 || S.lowpSteps >= 60)) {
      S.lowp = 0;
      S.tail = 2;
    } else if (S.tail == 0) {
      // after the two FP64 purification steps the occupation and the idempotency must be right
      // to 1e-3 (measured: ~1e-6); otherwise the iteration is diagonalised instead
      S.done = 1;
      S.converged = (fabs(t1 - S.nocc) < 1e-3 && fabs(e) < 1e-3) ? 1 : 0;
      S.doneStep = S.step;
    } else if (S.tail < 0 && (flat || exact || S.step >= S.maxSteps)) {
      S.done = 1;
      S.converged = ((flat || exact) && fabs(t1 - S.nocc) < 1e-6) ? 1 : 0;
      S.doneStep = S.step;
    } else if (S.step >= S.maxSteps) {
      S.done = 1;
      S.converged = 0;
      S.doneStep = S.step;
    }
  }
  S.ePrev2 = S.ePrev1;
  S.ePrev1 = e;
  S.prevPoly = S.poly;
  if (S.done) {
    S.ab[0] = 0.0;
    S.ab[1] = 1.0;
    S.poly = 2;
  } else if (fabs(t2 - S.nocc) < fabs(2.0 * t1 - t2 - S.nocc)) {
    S.ab[0] = 1.0;  // X <- X^2
    S.ab[1] = 0.0;
    S.poly = 0;
  } else {
    S.ab[0] = -1.0;  // X <- 2X - X^2
    S.ab[1] = 2.0;
    S.poly = 1;
  }
}

// State after evaluating the next X (traces t1, t2): counts the evaluation and, unless the
// projection is finished, decides the next polynomial.
__device__ inline Sp2Dev sp2Advance(Sp2Dev S, double t1, double t2) {
  S.xSplit = S.lowp;  // a split-BF16 x3 step (or X0 of a low-precision start) left X in split form
  S.step += 1;
  if (S.lowp) S.lowpSteps += 1;
  if (S.tail > 0 && !S.done) S.tail -= 1;  // an FP64 purification step was taken
  if (!S.done) sp2Decide(S, t1, t2);
  return S;
}

// Traces of a molecule's X from its per-tile partials part[2q], part[2q+1], q in [q0, q1).
// Always summed by the first 4 waves with stride 256 and combined in wave order, so the value
// is the same in every workgroup size that calls it (all threads must call; red: >= 8 doubles).
__device__ inline void sumPartials(const double* __restrict__ part, int q0, int q1, double* red, double& t1, double& t2) {
  double a = 0.0, b = 0.0;
  if (threadIdx.x < 256)
    for (int q = q0 + (int)threadIdx.x; q < q1; q += 256) {
      a += part[2 * (long long)q];
      b += part[2 * (long long)q + 1];
    }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && w < 4) {
    red[w] = a;
    red[4 + w] = b;
  }
  __syncthreads();
  t1 = ((red[0] + red[1]) + red[2]) + red[3];
  t2 = ((red[4] + red[5]) + red[6]) + red[7];
  __syncthreads();
}

// decision on the X whose partials are in partIn, for the host's convergence check; one
// workgroup per molecule (the next step kernel takes the same decision itself)
__global__ __launch_bounds__(256) void k_sp2_decide(const Sp2StepJob* __restrict__ jobs, const int* __restrict__ tileStart,
                                                    const double* __restrict__ partIn, const Sp2Dev* __restrict__ stateIn,
                                                    Sp2Dev* __restrict__ stateOut) {
  const int k = blockIdx.x, mol = jobs[k].mol;
  __shared__ double red[8];
  double t1, t2;
  sumPartials(partIn, tileStart[k], tileStart[k + 1], red, t1, t2);
  if (threadIdx.x == 0) stateOut[mol] = sp2Advance(stateIn[mol], t1, t2);
}

// P = 2 X with X the molecule's final SP2 iterate, X_s in buffer s % 2 (X0 or X1)
__global__ __launch_bounds__(256) void k_p_from_x(const MolDev* __restrict__ mols, const Sp2Dev* __restrict__ state,
                                                  const double* __restrict__ X0, const double* __restrict__ X1,
                                                  double* __restrict__ P) {
  const MolDev m = mols[blockIdx.y];
  const double* X = (state[m.id].doneStep & 1) ? X1 : X0;
  const long long nn = (long long)m.n * m.n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (long long)gridDim.x * blockDim.x)
    P[m.matOff + q] = 2.0 * X[m.matOff + q];
}

// Pulay DIIS coefficients for every molecule, one wave per molecule.  The history holds the
// molecule's last min(iter - start + 1, slots) iterations (start: its first iteration with FP64
// densities in the mixed form, else 1); slot (it - 1) % slots belongs to iteration it.
// B(a, b) = <E_a, E_b>, kept in Bmat across iterations; the bordered system
//   [B/scale  -1] [c]   [ 0]
//   [ -1       0] [l] = [-1]     (scale = largest diagonal of B)
// is solved by Gaussian elimination with partial pivoting.  With fewer than two vectors, or if
// the system is singular, the newest Fock matrix is used alone.
__global__ __launch_bounds__(64) void k_diis_solve(const MolDev* __restrict__ mols, int nmols,
                                                   const ScfMolDev* __restrict__ scf, const double* __restrict__ dots,
                                                   double* __restrict__ Bmat, double* __restrict__ coef, int slots) {
  // one wave per molecule: the bordered system lives in LDS, lane r owns row r of the elimination
  const int k = blockIdx.x, lane = threadIdx.x;
  const int id = mols[k].id;
  const int it = scf[id].iter;
  // the subspace restarts with the first FP64 iteration of a mixed-precision molecule (the
  // low-precision iterations' errors carry that precision's noise); fp64Start is 0 otherwise
  const int slot = (it - 1) % slots, cnt = min(it - max(scf[id].fp64Start, 1) + 1, slots);
  double* B = Bmat + (long long)id * slots * slots;
  double* c = coef + (long long)id * slots;
  const double* d = dots + (long long)id * slots;
  __shared__ double A[(kMaxDiis + 1) * (kMaxDiis + 1)], rhs[kMaxDiis + 1], x[kMaxDiis + 1];
  __shared__ int stored[kMaxDiis], sPiv, sOk;
  if (lane < cnt) stored[lane] = (it - cnt + lane) % slots;  // oldest first
  __syncthreads();
  // new row/column of B: <E_slot, E_q> for the stored q
  if (lane < cnt) {
    const int q = stored[lane];
    B[slot * slots + q] = d[q];
    B[q * slots + slot] = d[q];
  }
  auto bAt = [&](int a, int b) {  // B(stored[a], stored[b]) with this iteration's row taken from dots
    const int qa = stored[a], qb = stored[b];
    if (qa == slot) return d[qb];
    if (qb == slot) return d[qa];
    return B[qa * slots + qb];
  };
  if (cnt < 2) {
    if (lane < slots) c[lane] = lane == slot ? 1.0 : 0.0;
    return;
  }
  const int dim = cnt + 1;
  double scale = 0.0;
  for (int a = 0; a < cnt; ++a) scale = fmax(scale, fabs(bAt(a, a)));
  if (scale <= 0.0) scale = 1.0;
  if (lane < dim) {
    const int a = lane;
    rhs[a] = a == cnt ? -1.0 : 0.0;
    for (int b = 0; b < dim; ++b) A[a * dim + b] = (a < cnt && b < cnt) ? bAt(a, b) / scale : (a == b ? 0.0 : -1.0);
  }
  if (lane == 0) sOk = 1;
  __syncthreads();
  for (int p = 0; p < dim; ++p) {  // Gaussian elimination with partial pivoting
    if (lane == 0) {
      int piv = p;
      for (int r = p + 1; r < dim; ++r)
        if (fabs(A[r * dim + p]) > fabs(A[piv * dim + p])) piv = r;
      sPiv = piv;
      if (A[piv * dim + p] == 0.0) sOk = 0;
    }
    __syncthreads();
    if (!sOk) break;
    const int piv = sPiv;
    if (piv != p) {
      if (lane < dim) {
        const double tmp = A[p * dim + lane];
        A[p * dim + lane] = A[piv * dim + lane];
        A[piv * dim + lane] = tmp;
      }
      if (lane == 0) {
        const double tmp = rhs[p];
        rhs[p] = rhs[piv];
        rhs[piv] = tmp;
      }
    }
    __syncthreads();
    if (lane > p && lane < dim) {
      const int r = lane;
      const double f = A[r * dim + p] / A[p * dim + p];
      for (int q = p; q < dim; ++q) A[r * dim + q] -= f * A[p * dim + q];
      rhs[r] -= f * rhs[p];
    }
    __syncthreads();
  }
  if (lane == 0 && sOk) {
    for (int p = dim - 1; p >= 0; --p) {
      double v = rhs[p];
      for (int q = p + 1; q < dim; ++q) v -= A[p * dim + q] * x[q];
      x[p] = v / A[p * dim + p];
    }
    for (int a = 0; a < cnt; ++a)
      if (!isfinite(x[a])) sOk = 0;
  }
  __syncthreads();
  if (lane < slots) {  // coefficient of history slot `lane`; the newest Fock matrix alone on failure
    double v = (!sOk && lane == slot) ? 1.0 : 0.0;
    if (sOk)
      for (int a = 0; a < cnt; ++a)
        if (stored[a] == lane) v = x[a];
    c[lane] = v;
  }
}

// End of an SCF iteration, one thread per molecule of the Fock list: the convergence test
// (|E - E_old| < energyTol and RMSD(dP) < rmsdTol after the first iteration) for molecules still
// iterating, the stop at the iteration cap, the final energy for those that stopped iterating in
// the previous iteration, and the next iteration number.
// Mixed precision: a molecule starts with FP64 densities ("armed"); once |E - E_old| < lowpStart
// its densities come from split-BF16 x3 projections, until |E - E_old| < lowpTol, then FP64
// again.  The first iterations from the standard guess are where DIIS decides which SCF solution
// it reaches, and projection noise of ~1e-5 there can lead it to another one; after the energy
// change has fallen below lowpStart the path is settled.  The convergence test can pass only
// from the second iteration with FP64 densities after a low-precision window (E of iteration it
// comes from the density of iteration it - 1).
__global__ __launch_bounds__(64) void k_scf_check(const int* __restrict__ molList, int nmols, ScfMolDev* __restrict__ scf,
                                                  const double* __restrict__ eMol, const double* __restrict__ rmsd,
                                                  double energyTol, double rmsdTol, double lowpStart, double lowpTol,
                                                  int maxIterations) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nmols) return;
  const int id = molList[k];
  ScfMolDev S = scf[id];
  const double E = eMol[id];
  if (!S.converged && !S.stopped) {
    const double dE = fabs(E - S.Eold);
    if (S.iter > 1 && !S.lowp && S.iter >= S.fp64Start + 2 && dE < energyTol && rmsd[id] < rmsdTol) {
      S.converged = 1;
      S.convIter = S.iter;
    } else if (S.lowp && S.iter > 1 && dE < lowpTol) {
      S.lowp = 0;
      S.fp64Start = S.iter + 1;
    } else if (S.lowpArmed && S.iter > 1 && dE < lowpStart) {
      S.lowpArmed = 0;
      if (dE >= lowpTol) S.lowp = 1;  // the next iteration's density is a low-precision one
    }
    if (!S.converged && S.iter >= maxIterations) S.stopped = 1;
    S.Eold = E;
  } else if (!S.haveEnergy) {
    S.Eel = E;
    S.haveEnergy = 1;
  }
  S.iter += 1;
  scf[id] = S;
}



// ---------------------------------------------------------------------------------------------
// Element-wise passes over the per-molecule n x n matrices in 32 x 32 tiles.  tiles32[i] is the
// first tile of molecule i (prefix sums of ceil(n/32)^2); each workgroup of 256 threads takes
// one tile (thread: row tid % 32, columns tid / 32 + 8c) and writes its partial sums to
// part[tile]; a finishing kernel adds a molecule's partials in a fixed order, so results do not
// depend on scheduling or on the other molecules of the batch.
// ---------------------------------------------------------------------------------------------
struct Tile32 {
  int mol, n, r0, c0;
};

__device__ inline int findSegment(const int* __restrict__ start, int nseg, int t) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ inline Tile32 tile32Of(const MolDev* __restrict__ mols, const int* __restrict__ tiles32, int nmols, int t) {
  Tile32 T;
  T.mol = findSegment(tiles32, nmols, t);
  T.n = mols[T.mol].n;
  const int tn = (T.n + 31) >> 5;
  const int local = t - tiles32[T.mol];
  T.c0 = (local / tn) * 32;
  T.r0 = (local - (local / tn) * tn) * 32;
  return T;
}

// E = FP - (FP)^T into history slot `slot`, and part[t*slots + s] = <E, E_s> over the tile for
// every slot s (slots <= kMaxDiis)

__global__ __launch_bounds__(256) void k_commutator_dots(const MolDev* __restrict__ mols, const int* __restrict__ tiles32,
                                                         int nmols, const double* __restrict__ FP, double* __restrict__ hist,
                                                         int slots, const ScfMolDev* __restrict__ scf,
                                                         double* __restrict__ part) {
  const Tile32 T = tile32Of(mols, tiles32, nmols, blockIdx.x);
  const MolDev m = mols[T.mol];
  const int slot = (scf[m.id].iter - 1) % slots;
  const long long n = T.n, nn = n * n;
  // The error FP - PF = FP - (FP)^T is exactly antisymmetric (a - b = -(b - a) in floating point)
  // with a zero diagonal, so every dot is twice its sum over the elements above the diagonal:
  // only tiles on and above the diagonal are formed, stored and read (the history's lower
  // triangles are never used).
  if (T.r0 > T.c0) {
    if (threadIdx.x < slots) part[(long long)blockIdx.x * slots + threadIdx.x] = 0.0;
    return;
  }
  __shared__ double sT[32][33];  // sT[a][b] = FP(c0 + a, r0 + b): the mirrored tile
  const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
  const double* fp = FP + m.matOff;
  // coalesced reads down the columns of the mirrored tile
  for (int k = 0; k < 4; ++k) {
    const int j = tc + 8 * k;                // column offset within the mirrored tile
    const long long gr = T.c0 + tr, gc = T.r0 + j;  // element (gr, gc) of FP: row c0+tr, column r0+j
    sT[tr][j] = (gr < n && gc < n) ? fp[gc * n + gr] : 0.0;
  }
  __syncthreads();
  double acc[kMaxDiis];
#pragma unroll
  for (int q = 0; q < kMaxDiis; ++q) acc[q] = 0.0;
  double* dst = hist + m.histOff + (long long)slot * nn;
  for (int k = 0; k < 4; ++k) {
    const int c = tc + 8 * k;
    const long long gr = T.r0 + tr, gc = T.c0 + c;
    if (gr < gc && gc < n) {  // strictly above the diagonal (gr < gc < n)
      const long long q = gc * n + gr;
      const double e = fp[q] - sT[c][tr];  // FP(gr, gc) - FP(gc, gr)
      dst[q] = e;
#pragma unroll
      for (int sl = 0; sl < kMaxDiis; ++sl)
        if (sl < slots) acc[sl] += e * (sl == slot ? e : hist[m.histOff + (long long)sl * nn + q]);
    }
  }
  __shared__ double red[4];
  for (int sl = 0; sl < slots; ++sl) {
    const double v = blockSum(acc[sl], red);
    if (threadIdx.x == 0) part[(long long)blockIdx.x * slots + sl] = 2.0 * v;
  }
}

// out[id*slots + s] = sum over the molecule's tiles of part[t*slots + s]; grid (slots, nmols)
__global__ __launch_bounds__(256) void k_dots_finish(const MolDev* __restrict__ mols, const int* __restrict__ tiles32,
                                                     const double* __restrict__ part, int slots, double* __restrict__ out) {
  const int sl = blockIdx.x, mol = blockIdx.y;
  double v = 0.0;
  for (int t = tiles32[mol] + threadIdx.x; t < tiles32[mol + 1]; t += blockDim.x) v += part[(long long)t * slots + sl];
  __shared__ double red[4];
  v = blockSum(v, red);
  if (threadIdx.x == 0) out[mols[mol].id * slots + sl] = v;
}

// part[t] = sum over the tile of (P - Pold)^2; Pold <- P
__global__ __launch_bounds__(256) void k_rmsd_tiles(const MolDev* __restrict__ mols, const int* __restrict__ tiles32, int nmols,
                                                    const double* __restrict__ P, double* __restrict__ Pold,
                                                    double* __restrict__ part) {
  const Tile32 T = tile32Of(mols, tiles32, nmols, blockIdx.x);
  const long long n = T.n, off = mols[T.mol].matOff;
  const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
  double v = 0.0;
  for (int k = 0; k < 4; ++k) {
    const long long gr = T.r0 + tr, gc = T.c0 + tc + 8 * k;
    if (gr < n && gc < n) {
      const long long q = off + gc * n + gr;
      const double p = P[q];
      const double d = p - Pold[q];
      v += d * d;
      Pold[q] = p;
    }
  }
  __shared__ double red[4];
  v = blockSum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// rmsd[mol] = sqrt(sum of the molecule's partials / n^2); one workgroup per molecule
__global__ __launch_bounds__(256) void k_rmsd_finish(const MolDev* __restrict__ mols, const int* __restrict__ tiles32,
                                                     const double* __restrict__ part, double* __restrict__ rmsd) {
  const int mol = blockIdx.x;
  double v = 0.0;
  for (int t = tiles32[mol] + threadIdx.x; t < tiles32[mol + 1]; t += blockDim.x) v += part[t];
  __shared__ double red[4];
  v = blockSum(v, red);
  const double nn = (double)mols[mol].n * (double)mols[mol].n;
  if (threadIdx.x == 0) rmsd[mols[mol].id] = nn > 0 ? sqrt(v / nn) : 0.0;
}

// ---------------------------------------------------------------------------------------------
// EDIIS (Kudin, Scuseria, Cancès, J. Chem. Phys. 116, 8255 (2002)).  With P_i the density a Fock
// matrix F_i was built from and E_i its energy, the energy of the interpolated density is, for
// G linear in P,
//   E(c) = sum_i c_i E_i - 1/4 sum_ij c_i c_j tr[(P_i - P_j)(F_i - F_j)],
// minimised over c_i >= 0, sum_i c_i = 1.  The minimum lies in the relative interior of some face
// of the simplex, where it is a stationary point of E on that face: every non-empty subset of the
// stored iterations is tried (at most 2^16 - 1; 255 for 8) and the feasible candidate with the
// lowest E(c) kept.  EDIIS+DIIS (Garza, Scuseria, J. Chem. Phys. 137, 054110 (2012)) blends the
// two coefficient sets by the error norm e = ||FP - PF||_F: EDIIS for e > 1e-1, DIIS for e < 1e-4,
// c = 10 e c_EDIIS + (1 - 10 e) c_DIIS in between.
// ---------------------------------------------------------------------------------------------
// part[t * 2 slots + s] = <P_cur, F_s> and part[t * 2 slots + slots + s] = <P_s, F_cur> over tile t
__global__ __launch_bounds__(256) void k_hist_pf_dots(const MolDev* __restrict__ mols, const int* __restrict__ tiles32,
                                                      int nmols, const double* __restrict__ histP,
                                                      const double* __restrict__ histF, int slots,
                                                      const ScfMolDev* __restrict__ scf, double* __restrict__ part) {
  const Tile32 T = tile32Of(mols, tiles32, nmols, blockIdx.x);
  const MolDev m = mols[T.mol];
  const long long n = T.n, nn = n * n;
  const int cur = (scf[m.id].iter - 1) % slots;
  const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
  double a[kMaxDiis], b[kMaxDiis];
#pragma unroll
  for (int q = 0; q < kMaxDiis; ++q) a[q] = b[q] = 0.0;
  const double* Pc = histP + m.histOff + (long long)cur * nn;
  const double* Fc = histF + m.histOff + (long long)cur * nn;
  for (int k = 0; k < 4; ++k) {
    const long long gr = T.r0 + tr, gc = T.c0 + tc + 8 * k;
    if (gr < n && gc < n) {
      const long long q = gc * n + gr;
      const double pc = Pc[q], fc = Fc[q];
#pragma unroll
      for (int sl = 0; sl < kMaxDiis; ++sl)
        if (sl < slots) {
          a[sl] += pc * histF[m.histOff + (long long)sl * nn + q];
          b[sl] += fc * histP[m.histOff + (long long)sl * nn + q];
        }
    }
  }
  __shared__ double red[4];
  for (int sl = 0; sl < slots; ++sl) {
    const double va = blockSum(a[sl], red);
    const double vb = blockSum(b[sl], red);
    if (threadIdx.x == 0) {
      part[(long long)blockIdx.x * 2 * slots + sl] = va;
      part[(long long)blockIdx.x * 2 * slots + slots + sl] = vb;
    }
  }
}

// energies of the history: eHist[id*slots + cur] = eMol[id]
__global__ void k_hist_energy(const MolDev* __restrict__ mols, int nmols, const ScfMolDev* __restrict__ scf,
                              const double* __restrict__ eMol, double* __restrict__ eHist, int slots) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nmols) return;
  const int id = mols[k].id;
  eHist[(long long)id * slots + (scf[id].iter - 1) % slots] = eMol[id];
}

// EDIIS coefficients (mixer 2) or the EDIIS+DIIS blend with the DIIS coefficients already in
// coef (mixer 3); one thread per molecule.  pf[id * 2 slots + ...]: the new overlaps of
// k_hist_pf_dots; Mmat[id][i][j] = tr(P_i F_j), kept across iterations; dots[id*slots + cur] =
// ||E_cur||^2 from the DIIS error overlaps.
__global__ __launch_bounds__(64) void k_ediis_solve(const MolDev* __restrict__ mols, int nmols,
                                                   const ScfMolDev* __restrict__ scf, const double* __restrict__ pf,
                                                   const double* __restrict__ eHist, const double* __restrict__ dots,
                                                   double* __restrict__ Mmat, double* __restrict__ coef, int slots,
                                                   int blend) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nmols) return;
  const int id = mols[k].id;
  const int it = scf[id].iter;
  const int cur = (it - 1) % slots, cnt = min(it - max(scf[id].fp64Start, 1) + 1, slots);  // as in k_diis_solve
  double* M = Mmat + (long long)id * slots * slots;
  const double* pfm = pf + (long long)id * 2 * slots;
  int stored[kMaxDiis];
  for (int a = 0; a < cnt; ++a) stored[a] = (it - cnt + a) % slots;
  for (int a = 0; a < cnt; ++a) {
    const int q = stored[a];
    M[cur * slots + q] = pfm[q];          // tr(P_cur F_q)
    M[q * slots + cur] = pfm[slots + q];  // tr(P_q F_cur)
  }
  double* c = coef + (long long)id * slots;
  double wE = 1.0;  // weight of the EDIIS coefficients
  if (blend) {
    const double err = sqrt(fmax(dots[(long long)id * slots + cur], 0.0));
    wE = err > 1e-1 ? 1.0 : (err < 1e-4 ? 0.0 : 10.0 * err);
  }
  if (wE == 0.0) return;  // DIIS coefficients stay
  // A(a, b) = tr[(P_a - P_b)(F_a - F_b)] over the stored iterations, E(a) their energies
  double A[kMaxDiis * kMaxDiis], E[kMaxDiis];
  for (int a = 0; a < cnt; ++a) {
    const int qa = stored[a];
    E[a] = eHist[(long long)id * slots + qa];
    for (int b = 0; b < cnt; ++b) {
      const int qb = stored[b];
      A[a * kMaxDiis + b] = M[qa * slots + qa] + M[qb * slots + qb] - M[qa * slots + qb] - M[qb * slots + qa];
    }
  }
  double best = DBL_MAX, cbest[kMaxDiis];
  for (int a = 0; a < cnt; ++a) cbest[a] = (a == cnt - 1) ? 1.0 : 0.0;  // newest alone if nothing better
  for (unsigned mask = 1; mask < (1u << cnt); ++mask) {
    int idx[kMaxDiis], m = 0;
    for (int a = 0; a < cnt; ++a)
      if (mask & (1u << a)) idx[m++] = a;
    // stationary point on the face: E_S - 1/2 A_SS c = lambda 1, sum c = 1
    const int dim = m + 1;
    double L[(kMaxDiis + 1) * (kMaxDiis + 1)], r[kMaxDiis + 1], x[kMaxDiis + 1];
    for (int a = 0; a < m; ++a) {
      for (int b = 0; b < m; ++b) L[a * dim + b] = -0.5 * A[idx[a] * kMaxDiis + idx[b]];
      L[a * dim + m] = -1.0;
      L[m * dim + a] = 1.0;
      r[a] = -E[idx[a]];
    }
    L[m * dim + m] = 0.0;
    r[m] = 1.0;
    bool ok = true;
    for (int p = 0; p < dim && ok; ++p) {  // Gaussian elimination, partial pivoting
      int piv = p;
      for (int q = p + 1; q < dim; ++q)
        if (fabs(L[q * dim + p]) > fabs(L[piv * dim + p])) piv = q;
      if (fabs(L[piv * dim + p]) < 1e-300) {
        ok = false;
        break;
      }
      if (piv != p) {
        for (int q = 0; q < dim; ++q) {
          const double t = L[p * dim + q];
          L[p * dim + q] = L[piv * dim + q];
          L[piv * dim + q] = t;
        }
        const double t = r[p];
        r[p] = r[piv];
        r[piv] = t;
      }
      for (int q = p + 1; q < dim; ++q) {
        const double f = L[q * dim + p] / L[p * dim + p];
        for (int w = p; w < dim; ++w) L[q * dim + w] -= f * L[p * dim + w];
        r[q] -= f * r[p];
      }
    }
    if (!ok) continue;
    for (int p = dim - 1; p >= 0; --p) {
      double v = r[p];
      for (int q = p + 1; q < dim; ++q) v -= L[p * dim + q] * x[q];
      x[p] = v / L[p * dim + p];
    }
    bool feasible = true;
    for (int a = 0; a < m; ++a) feasible = feasible && isfinite(x[a]) && x[a] >= 0.0;
    if (!feasible) continue;
    double cc[kMaxDiis];
    for (int a = 0; a < cnt; ++a) cc[a] = 0.0;
    for (int a = 0; a < m; ++a) cc[idx[a]] = x[a];
    double f = 0.0;
    for (int a = 0; a < cnt; ++a) {
      f += cc[a] * E[a];
      for (int b = 0; b < cnt; ++b) f -= 0.25 * cc[a] * cc[b] * A[a * kMaxDiis + b];
    }
    if (f < best) {
      best = f;
      for (int a = 0; a < cnt; ++a) cbest[a] = cc[a];
    }
  }
  if (wE >= 1.0) {
    for (int q = 0; q < slots; ++q) c[q] = 0.0;
    for (int a = 0; a < cnt; ++a) c[stored[a]] = cbest[a];
  } else {
    for (int q = 0; q < slots; ++q) c[q] *= 1.0 - wE;
    for (int a = 0; a < cnt; ++a) c[stored[a]] += wE * cbest[a];
  }
}

// Gershgorin discs of F, one wave per column (F is symmetric: column sums = row sums):
// lo[j] = F_jj - sum_{i != j} |F_ij|, hi[j] = F_jj + sum_{i != j} |F_ij|; grid (columns / 4, molecules)
__global__ __launch_bounds__(256) void k_sp2_discs(const MolDev* __restrict__ mols, const double* __restrict__ F,
                                                   double* __restrict__ lo, double* __restrict__ hi) {
  const MolDev m = mols[blockIdx.y];
  const long long j = (long long)blockIdx.x * 4 + (threadIdx.x >> 6), n = m.n;
  const int lane = threadIdx.x & 63;
  if (j >= n) return;  // whole wave leaves together
  const double* col = F + m.matOff + j * n;
  double r = 0.0;
  for (long long i = lane; i < n; i += 64)
    if (i != j) r += fabs(col[i]);
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
  if (lane == 0) {
    const double d = col[j];
    lo[m.vecOff + j] = d - r;
    hi[m.vecOff + j] = d + r;
  }
}

// spectral range of each molecule (widened by a margin so that no eigenvalue of X0 sits exactly
// on the fixed points 0 or 1); one workgroup per molecule
__global__ __launch_bounds__(256) void k_sp2_range(const MolDev* __restrict__ mols, const double* __restrict__ lo,
                                                   const double* __restrict__ hi, Sp2Dev* __restrict__ state, int maxSteps,
                                                   const ScfMolDev* __restrict__ scf) {
  const MolDev m = mols[blockIdx.x];
  double a = DBL_MAX, b = -DBL_MAX;
  for (int i = threadIdx.x; i < m.n; i += blockDim.x) {
    a = fmin(a, lo[m.vecOff + i]);
    b = fmax(b, hi[m.vecOff + i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = fmin(a, __shfl_xor(a, o, 64));
    b = fmax(b, __shfl_xor(b, o, 64));
  }
  __shared__ double sa[4], sb[4];
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      a = fmin(a, sa[w]);
      b = fmax(b, sb[w]);
    }
    const double margin = 1e-3 * (b - a) + 1e-6;
    Sp2Dev S{};
    S.n = m.n;
    S.nocc = m.nocc;
    S.maxSteps = maxSteps;
    S.emin = a - margin;
    S.emax = b + margin;
    S.ePrev1 = S.ePrev2 = DBL_MAX;
    S.step = -1;  // no X evaluated yet
    S.poly = S.prevPoly = -1;
    S.lowp = scf ? scf[m.id].lowp : 0;
    S.tail = -1;
    state[m.id] = S;
  }
}

// ---------------------------------------------------------------------------------------------
// Spectral bounds of F from a short Lanczos run, to narrow the Gershgorin interval of the SP2
// start (the discs of an NDDO Fock matrix overestimate the top of its spectrum several times,
// and every halving of the interval saves about one purification step).
// Lanczos pool of a molecule: 5 n doubles at 5 vecOff: q_j in slot j % 3, y_j = F q_j in slot
// 3 + j % 2.  coef, 2 kLzSteps + 1 doubles per molecule: alpha_j at [j], beta_(j+1) at
// [kLzSteps + j], [2 kLzSteps] = 1 when the Krylov space was exhausted (bounds then unused).
// Launch j (grid: blocks of kLzCols columns x molecules) forms q_j from q_(j-1), q_(j-2) and
// y_(j-1) of the earlier launches (every block the whole vector in the same order, so all blocks
// hold the same q_j in LDS) and y_j(i) = F(:, i) . q_j for its columns (F symmetric).
// ---------------------------------------------------------------------------------------------
constexpr int kLzSteps = 8;
constexpr int kLzCols = 16;    // columns per block: 4 waves x 4 columns
constexpr int kLzMinN = 32;    // smallest molecule for which the bounds are ever taken
constexpr int kLzMaxN = 8000;  // q_j held in LDS (64000 bytes)

__device__ inline double lzStart(int l) { return sin(0.7548776662466927 * l + 0.3) + 0.5; }

__device__ inline void blockSum2(double& a, double& b, double* red) {
  // red: >= 2 blockDim.x/64 doubles of LDS; both sums in a fixed order
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = b = 0.0;
  for (int k = 0; k < nw; ++k) {
    a += red[2 * k];
    b += red[2 * k + 1];
  }
  __syncthreads();
}

// q_j of molecule m into LDS q (all n entries); returns false when the Krylov space is exhausted
// (q_j = 0 then).  cf: the molecule's coefficients; the block with writer set records alpha_(j-1),
// beta_j and the exhaustion flag.
__device__ inline void lzNextVector(const MolDev& m, const double* base, double* cf, int j, double* q, double* red,
                                    bool writer) {
  const int n = m.n;
  if (j == 0) {
    double ss = 0.0, unused = 0.0;
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
      const double v = lzStart(l);
      q[l] = v;
      ss += v * v;
    }
    blockSum2(ss, unused, red);
    const double inv = 1.0 / sqrt(ss);
    for (int l = threadIdx.x; l < n; l += blockDim.x) q[l] *= inv;
    if (writer && threadIdx.x == 0) cf[2 * kLzSteps] = 0.0;
    __syncthreads();
    return;
  }
  const double* qa = base + (long long)((j - 1) % 3) * n;      // q_(j-1)
  const double* qb = base + (long long)((j + 1) % 3) * n;      // q_(j-2)
  const double* y = base + (long long)(3 + (j - 1) % 2) * n;   // y_(j-1) = F q_(j-1)
  double d = 0.0, yy = 0.0;
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    d += qa[l] * y[l];
    yy += y[l] * y[l];
  }
  blockSum2(d, yy, red);
  const double alpha = d;
  const double bPrev = j >= 2 ? cf[kLzSteps + j - 2] : 0.0;  // beta_(j-1)
  double ss = 0.0, unused = 0.0;
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    double u = y[l] - alpha * qa[l];
    if (j >= 2) u -= bPrev * qb[l];
    q[l] = u;
    ss += u * u;
  }
  blockSum2(ss, unused, red);
  const double beta = sqrt(ss);
  const bool exhausted = !(beta > 1e-10 * sqrt(yy));  // also true for NaN
  const double inv = exhausted ? 0.0 : 1.0 / beta;
  for (int l = threadIdx.x; l < n; l += blockDim.x) q[l] *= inv;
  if (writer && threadIdx.x == 0) {
    if (j - 1 < kLzSteps) cf[j - 1] = alpha;
    cf[kLzSteps + j - 1] = beta;
    if (exhausted) cf[2 * kLzSteps] = 1.0;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_lanczos_step(const MolDev* __restrict__ mols, const double* __restrict__ F,
                                                      double* __restrict__ lz, double* __restrict__ coef, int j, int minN) {
  extern __shared__ double q[];
  __shared__ double red[8];
  const MolDev m = mols[blockIdx.y];
  const int n = m.n;
  if (n < minN || n > kLzMaxN) return;
  const int c0 = blockIdx.x * kLzCols;
  if (c0 >= n) return;  // whole block leaves together
  double* base = lz + 5 * m.vecOff;
  double* cf = coef + (long long)m.id * (2 * kLzSteps + 1);
  lzNextVector(m, base, cf, j, q, red, blockIdx.x == 0);
  double* qOut = base + (long long)(j % 3) * n;
  double* yOut = base + (long long)(3 + j % 2) * n;
  const int c1 = min(c0 + kLzCols, n);
  for (int i = c0 + threadIdx.x; i < c1; i += blockDim.x) qOut[i] = q[i];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double* Fm = F + m.matOff;
  for (int pass = 0; pass < kLzCols / 16; ++pass) {
    const int cb = c0 + (pass * 4 + w) * 4;
    if (cb >= c1) break;  // per wave; no block barrier follows
    const int nc = min(4, c1 - cb);
    const double* col = Fm + (long long)cb * n;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (nc == 4) {
      for (int l = lane; l < n; l += 64) {
        const double ql = q[l];
        a0 += col[l] * ql;
        a1 += col[n + l] * ql;
        a2 += col[2 * (long long)n + l] * ql;
        a3 += col[3 * (long long)n + l] * ql;
      }
    } else {
      for (int l = lane; l < n; l += 64) {
        const double ql = q[l];
        a0 += col[l] * ql;
        if (nc > 1) a1 += col[n + l] * ql;
        if (nc > 2) a2 += col[2 * (long long)n + l] * ql;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      a0 += __shfl_xor(a0, o, 64);
      a1 += __shfl_xor(a1, o, 64);
      a2 += __shfl_xor(a2, o, 64);
      a3 += __shfl_xor(a3, o, 64);
    }
    if (lane == 0) {
      yOut[cb] = a0;
      if (nc > 1) yOut[cb + 1] = a1;
      if (nc > 2) yOut[cb + 2] = a2;
      if (nc > 3) yOut[cb + 3] = a3;
    }
  }
}

// number of eigenvalues of the tridiagonal T (kLzSteps x kLzSteps: diagonal a, couplings b[i]
// between i and i + 1) below x, from the signs of the LDL^T pivots of T - x I (loops over the
// compile-time order, so a and b stay in registers)
__device__ inline int lzSturm(const double (&a)[kLzSteps], const double (&b)[kLzSteps], double x) {
  int c = 0;
  double d = a[0] - x;
#pragma unroll
  for (int i = 0; i < kLzSteps; ++i) {
    if (i > 0) d = a[i] - x - b[i - 1] * b[i - 1] / d;
    if (d == 0.0) d = -1e-300;
    c += d < 0.0 ? 1 : 0;
  }
  return c;
}

// the idx-th eigenvalue of T by bisection in [lo, hi]
__device__ inline double lzEig(const double (&a)[kLzSteps], const double (&b)[kLzSteps], int idx, double lo, double hi) {
  for (int it = 0; it < 36; ++it) {  // 2^-36 of the Gershgorin range of T
    const double mid = 0.5 * (lo + hi);
    if (lzSturm(a, b, mid) > idx) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// |last component| of T's normalised eigenvector for the eigenvalue t:
// s^2 = p_(K-1)(t) / p_K'(t) with p_k(t) = det(t I - T_k) of the leading k x k block
__device__ inline double lzLastComponent(const double (&a)[kLzSteps], const double (&b)[kLzSteps], double t) {
  double p0 = 1.0, p1 = t - a[0], d0 = 0.0, d1 = 1.0;
#pragma unroll
  for (int i = 1; i < kLzSteps; ++i) {
    const double bb = b[i - 1] * b[i - 1];
    const double p2 = (t - a[i]) * p1 - bb * p0;
    const double d2 = p1 + (t - a[i]) * d1 - bb * d0;
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d2;
  }
  return sqrt(fabs(p0 / d1));
}

// End of the Lanczos run (one block per molecule): alpha_(K-1), beta_K from the last launch, then
// the extreme Ritz values theta of T, widened by their residual norms beta_K |s_K| and by 1 % of
// theta_max - theta_min, narrow the molecule's Gershgorin interval in state (the intersection of
// the two; unchanged when the run broke down).
__global__ __launch_bounds__(256) void k_lanczos_bounds(const MolDev* __restrict__ mols, double* __restrict__ lz,
                                                        double* __restrict__ coef, Sp2Dev* __restrict__ state, int minN) {
  __shared__ double red[8];
  const MolDev m = mols[blockIdx.x];
  const int n = m.n;
  if (n < minN || n > kLzMaxN) return;
  double* base = lz + 5 * m.vecOff;
  double* cf = coef + (long long)m.id * (2 * kLzSteps + 1);
  const int K = kLzSteps;
  const double* qa = base + (long long)((K - 1) % 3) * n;
  const double* qb = base + (long long)((K + 1) % 3) * n;
  const double* y = base + (long long)(3 + (K - 1) % 2) * n;
  const double bPrev = cf[kLzSteps + K - 2];
  double d = 0.0, unused = 0.0;
  for (int l = threadIdx.x; l < n; l += blockDim.x) d += qa[l] * y[l];
  blockSum2(d, unused, red);
  const double alpha = d;
  double ss = 0.0;
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    const double u = y[l] - alpha * qa[l] - bPrev * qb[l];
    ss += u * u;
  }
  unused = 0.0;
  blockSum2(ss, unused, red);
  if (threadIdx.x != 0 || cf[2 * kLzSteps] != 0.0) return;
  const double betaK = sqrt(ss);
  double a[kLzSteps], b[kLzSteps];
#pragma unroll
  for (int i = 0; i < kLzSteps; ++i) {
    a[i] = i < K - 1 ? cf[i] : alpha;
    b[i] = i < K - 1 ? cf[kLzSteps + i] : 0.0;  // beta_(i+1): coupling of i and i + 1
  }
  double lo = DBL_MAX, hi = -DBL_MAX;
#pragma unroll
  for (int i = 0; i < kLzSteps; ++i) {
    const double r = fabs(b[i]) + (i > 0 ? fabs(b[i - 1]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  const double tmin = lzEig(a, b, 0, lo, hi), tmax = lzEig(a, b, K - 1, lo, hi);
  const double S = tmax - tmin;
  const double eLo = tmin - betaK * lzLastComponent(a, b, tmin) - 0.01 * S;
  const double eHi = tmax + betaK * lzLastComponent(a, b, tmax) + 0.01 * S;
  if (!(eLo < eHi) || !isfinite(eLo) || !isfinite(eHi)) return;
  Sp2Dev& st = state[m.id];
  const double nLo = fmax(st.emin, eLo), nHi = fmin(st.emax, eHi);
  if (nLo < nHi) {
    st.emin = nLo;
    st.emax = nHi;
  }
}

// X0 = (e_max I - F) / (e_max - e_min) over the tiles on and above the diagonal, stored to both
// triangles (and, for molecules starting in split-BF16 x3, as the split copies over the padded
// ldp x ldp tiles, zero outside n); part[2t], part[2t+1] = trace and squared norm of tile t's
// share of X0
__global__ __launch_bounds__(256) void k_sp2_x0(const Sp2StepJob* __restrict__ jobs, const int* __restrict__ tileStart,
                                                int nmols, const double* __restrict__ F, const MolDev* __restrict__ mols,
                                                const Sp2Dev* __restrict__ state, double* __restrict__ part) {
  const int t = blockIdx.x;
  const int mol = findSegment(tileStart, nmols, t);
  const Sp2StepJob J = jobs[mol];
  const int T = J.tile, n = J.n, tn = (n + T - 1) / T;
  int local = t - tileStart[mol], ti = 0;
  while (local >= tn - ti) {
    local -= tn - ti;
    ++ti;
  }
  const int tj = ti + local, r0 = ti * T, c0 = tj * T;
  const double* Fm = F + mols[mol].matOff;
  double* Xm = const_cast<double*>(J.X);
  const double emin = state[J.mol].emin, emax = state[J.mol].emax, inv = 1.0 / (emax - emin);
  const bool lowp = state[J.mol].lowp != 0;
  unsigned short* Xh = const_cast<unsigned short*>(J.Xh);
  unsigned short* Xl = const_cast<unsigned short*>(J.Xl);
  const long long ldp = J.ldp;
  const bool diagTile = ti == tj;
  double t1 = 0.0, t2 = 0.0;
  for (int e = threadIdx.x; e < T * T; e += 256) {  // lower/own elements, down columns
    const int row = e % T, cl = e / T;
    const int gi = r0 + row, gj = c0 + cl;
    if (diagTile && row < cl) continue;
    double x = 0.0;
    if (gi < n && gj < n) {
      x = ((gi == gj ? emax : 0.0) - Fm[(long long)gj * n + gi]) * inv;
      Xm[(long long)gj * n + gi] = x;
      t2 += x * x;
      if (gi == gj) t1 += x;
    }
    if (lowp && gi < ldp && gj < ldp) splitBf16(x, Xh[gj * ldp + gi], Xl[gj * ldp + gi]);
  }
  for (int e = threadIdx.x; e < T * T; e += 256) {  // mirror, along rows
    const int cl = e % T, row = e / T;
    const int gi = r0 + row, gj = c0 + cl;
    if (diagTile && row <= cl) continue;
    double x = 0.0;
    if (gi < n && gj < n) {
      x = (0.0 - Fm[(long long)gj * n + gi]) * inv;  // same expression as its mirror
      Xm[(long long)gi * n + gj] = x;
      t2 += x * x;
    }
    if (lowp && gi < ldp && gj < ldp) splitBf16(x, Xh[gi * ldp + gj], Xl[gi * ldp + gj]);
  }
  __shared__ double red[4];
  t1 = blockSum(t1, red);
  t2 = blockSum(t2, red);
  if (threadIdx.x == 0) {
    part[2 * (long long)t] = t1;
    part[2 * (long long)t + 1] = t2;
  }
}

// ---------------------------------------------------------------------------------------------
// SP2 step Y = alpha X X + beta X for symmetric X (alpha, beta from the controller decision).
// Only the tiles on and above the diagonal are computed, and each result is stored to both
// triangles, so Y is exactly symmetric.  One 32 x 32 tile per workgroup of kStepWaves waves that
// split K; small molecules give few tiles, so the K split is what keeps enough loads in flight
// per CU.  The partial tiles are added in LDS in a fixed order; per-tile trace and squared norm
// go to part[2t], part[2t+1].
//  - FP64: wave w takes k = 4 w + 64 s .. +3 and loads its `v_mfma_f64_16x16x4f64` operands
//    straight from global memory (X(i,k) = X[k n + i] and, by symmetry, X(k,j) = X[k n + j]:
//    16 consecutive doubles per quarter-wave), two steps ahead of the products.
//  - split-BF16 x3: X = Xh + Xl in bf16; X X ~ Xh Xh + Xh Xl + Xl Xh with FP32 accumulation
//    (`v_mfma_f32_16x16x32_bf16`, operand A(i, k..k+7) = X(k.., i) and B(k..k+7, j) = X(k.., j):
//    16 contiguous bytes of a column).  Wave w takes the 32-wide k chunks w, w + 16, ...; the
//    padded copies (zero outside n) need no masks.  The beta X term and the stored Y are FP64;
//    Y is also split into Yh + Yl for the next low-precision step.
// ---------------------------------------------------------------------------------------------
constexpr int kStepWaves = 16;

// Workgroup b of a grid of g is dispatched to XCD b % 8 (round robin over the eight XCDs, each
// with its own L2).  Mapping b to the logical tile (b % 8) q + b / 8 (q = g / 8; the last g % 8
// keep their index) gives each XCD a contiguous range of tiles, so the tiles of one molecule
// share an L2 and their operand panels are fetched from HBM about once.
__device__ inline int xcdTile(int b, int g) {
  const int q = g >> 3;
  return b < 8 * q ? (b & 7) * q + (b >> 3) : b;
}

__device__ inline void mmaF64(double4v (&acc)[2][2], double a0, double a1, double b0, double b1) {
  acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
  acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
  acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
  acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
}

// acc += A(r0:r0+32, k) B(k, c0:c0+32) over the k of wave w (k = 4 w + 64 s + lane/16), for
// symmetric n x n column-major operands read down their columns: A(i, k) = Ag[k n + i] and
// B(k, j) = B(j, k) = Bg[k n + j].  Operand rows/columns are clamped into the matrix: rows i >= n
// of A (columns j >= n of B) only feed outputs that are never stored, so they need no zeroing,
// and every load is valid and unconditional.  Only k >= n must contribute zero: A is zeroed there
// in the last two steps (a zero A adds exactly 0 to the finite accumulator).  Two steps of
// operands are in flight ahead of the MFMAs.
// the operands of a wave's first two k-steps (registers R: even steps, Q: odd steps)
struct F64Lead {
  double ra0, ra1, rb0, rb1, qa0, qa1, qb0, qb1;
};
__device__ inline F64Lead kSplitF64Lead(const gdouble* Ag, const gdouble* Bg, int n, int r0, int c0, int w, int li, int lk) {
  const int ca0 = min(r0 + li, n - 1), ca1 = min(r0 + 16 + li, n - 1);
  const int cb0 = min(c0 + li, n - 1), cb1 = min(c0 + 16 + li, n - 1);
  constexpr int KS = 4 * kStepWaves;
  const int k0 = 4 * w;
  F64Lead L;
  const gdouble* ca = Ag + (long long)min(k0 + lk, n - 1) * n;
  const gdouble* cb = Bg + (long long)min(k0 + lk, n - 1) * n;
  L.ra0 = ca[ca0];
  L.ra1 = ca[ca1];
  L.rb0 = cb[cb0];
  L.rb1 = cb[cb1];
  ca = Ag + (long long)min(k0 + KS + lk, n - 1) * n;
  cb = Bg + (long long)min(k0 + KS + lk, n - 1) * n;
  L.qa0 = ca[ca0];
  L.qa1 = ca[ca1];
  L.qb0 = cb[cb0];
  L.qb1 = cb[cb1];
  return L;
}

// L: the operands of the first two k-steps, from kSplitF64Lead with the same arguments (the SP2
// step loads them before its controller decision)
__device__ inline void kSplitF64(double4v (&acc)[2][2], const gdouble* Ag, const gdouble* Bg, int n, int r0, int c0,
                                 int w, int li, int lk, const F64Lead& L) {
  const int ca0 = min(r0 + li, n - 1), ca1 = min(r0 + 16 + li, n - 1);
  const int cb0 = min(c0 + li, n - 1), cb1 = min(c0 + 16 + li, n - 1);
  constexpr int KS = 4 * kStepWaves;  // k stride between a wave's steps
  const int S = 4 * w < n ? (n - 4 * w + KS - 1) / KS : 0;  // steps of this wave
  int k0 = 4 * w;
  const gdouble *ca, *cb;
  // step registers: R (even steps) and Q (odd steps)
  double ra0 = L.ra0, ra1 = L.ra1, rb0 = L.rb0, rb1 = L.rb1;
  double qa0 = L.qa0, qa1 = L.qa1, qb0 = L.qb0, qb1 = L.qb1;
  int s = 0;
  for (; s + 2 < S; s += 2, k0 += 2 * KS) {  // steps s and s + 1 are full (not the last one)
    mmaF64(acc, ra0, ra1, rb0, rb1);
    ca = Ag + (long long)min(k0 + 2 * KS + lk, n - 1) * n;
    cb = Bg + (long long)min(k0 + 2 * KS + lk, n - 1) * n;
    ra0 = ca[ca0];
    ra1 = ca[ca1];
    rb0 = cb[cb0];
    rb1 = cb[cb1];
    mmaF64(acc, qa0, qa1, qb0, qb1);
    ca = Ag + (long long)min(k0 + 3 * KS + lk, n - 1) * n;
    cb = Bg + (long long)min(k0 + 3 * KS + lk, n - 1) * n;
    qa0 = ca[ca0];
    qa1 = ca[ca1];
    qb0 = cb[cb0];
    qb1 = cb[cb1];
  }
  // the remaining (at most two) steps, masked: step s if s < S, step s + 1 if s + 1 < S
  const bool m0 = s < S && k0 + lk < n, m1 = s + 1 < S && k0 + KS + lk < n;
  mmaF64(acc, m0 ? ra0 : 0.0, m0 ? ra1 : 0.0, rb0, rb1);
  mmaF64(acc, m1 ? qa0 : 0.0, m1 ? qa1 : 0.0, qb0, qb1);
}

// kSplitF64 for an X held as split copies (the first FP64 step after split-BF16 x3 steps):
// A(i, k) = B(k, i) = Xh[k ldp + i] + Xl[k ldp + i], converted exactly; same k partition.
__device__ inline void kSplitF64FromSplit(double4v (&acc)[2][2], const gbf16c* Xh, const gbf16c* Xl,
                                          long long ldp, int n, int r0, int c0, int w, int li, int lk) {
  const int ca0 = min(r0 + li, n - 1), ca1 = min(r0 + 16 + li, n - 1);
  const int cb0 = min(c0 + li, n - 1), cb1 = min(c0 + 16 + li, n - 1);
  for (int k0 = 4 * w; k0 < n; k0 += 4 * kStepWaves) {
    const int k = k0 + lk;
    const long long col = (long long)min(k, n - 1) * ldp;
    const bool valid = k < n;
    const double a0 = valid ? splitValue(Xh, Xl, col + ca0) : 0.0;
    const double a1 = valid ? splitValue(Xh, Xl, col + ca1) : 0.0;
    mmaF64(acc, a0, a1, splitValue(Xh, Xl, col + cb0), splitValue(Xh, Xl, col + cb1));
  }
}

// one 32-wide k chunk of split operands: hi/lo fragments of rows (A) and columns (B)
struct Bf16Chunk {
  bf16x8 ah0, ah1, al0, al1, bh0, bh1, bl0, bl1;
};

__device__ inline void mmaBf16x3(floatx4 (&acc)[2][2], const Bf16Chunk& c) {
  const bf16x8 ah[2] = {c.ah0, c.ah1}, al[2] = {c.al0, c.al1}, bh[2] = {c.bh0, c.bh1}, bl[2] = {c.bl0, c.bl1};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
    }
}

__global__ __launch_bounds__(kStepWaves * 64) void k_sp2_step(const Sp2StepJob* __restrict__ jobs,
                                                              const int* __restrict__ tileStart, int njobs,
                                                              const double* __restrict__ partIn, double* __restrict__ partOut,
                                                              const Sp2Dev* __restrict__ stateIn,
                                                              Sp2Dev* __restrict__ stateOut) {
  const int t = xcdTile(blockIdx.x, gridDim.x);  // molecules with 32-wide tiles come first in the list
  const int mol = findSegment(tileStart, njobs, t);
  const Sp2StepJob J = jobs[mol];
  const int n = J.n;
  __shared__ double red[kStepWaves];
  const int tn = (n + 31) >> 5;
  int local = t - tileStart[mol], ti = 0;
  while (local >= tn - ti) {
    local -= tn - ti;
    ++ti;
  }
  const int tj = ti + local;
  const int r0 = ti * 32, c0 = tj * 32;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const double* __restrict__ X = J.X;
  // the first operands of an FP64 product from the FP64 X, issued before the controller's decision
  // (which does not change them) so that their latency overlaps the partial-trace sums; unused
  // when the step is split-BF16, reads split copies, or the projection is final (X is always an
  // n x n FP64 buffer, so the loads stay in bounds)
  const F64Lead lead = kSplitF64Lead((const gdouble*)X, (const gdouble*)X, n, r0, c0, w, li, lk);
  // controller: every workgroup of the molecule takes the same decision from the partials of X
  // (written by the previous launch); the first one records it for the next launch
  double t1, t2;
  sumPartials(partIn, tileStart[mol], tileStart[mol + 1], red, t1, t2);
  const Sp2Dev S = sp2Advance(stateIn[J.mol], t1, t2);
  if (t == tileStart[mol] && threadIdx.x == 0) stateOut[J.mol] = S;
  if (S.done) return;  // X of this molecule is final (it stays in the buffer it was written to)
  // the decision made on X_s says which kind of product computes X_{s+1}; S.lowp == 1 means
  // this step is a low-precision one
  const bool lowp = S.lowp != 0;
  const double alpha = S.ab[0], beta = S.ab[1];
  const gbf16c* Xh16 = (const gbf16c*)J.Xh;
  const gbf16c* Xl16 = (const gbf16c*)J.Xl;
  const gdouble* Xd = (const gdouble*)X;
  const long long ldp = J.ldp;
  const bool diagTile = ti == tj;
  // X(gi, gj) for the beta term: FP64, or the split copies after a split-BF16 x3 step
  const bool xs = S.xSplit != 0;
  auto xAt = [&](int gi, int gj) {
    return xs ? splitValue(Xh16, Xl16, (long long)gj * ldp + gi) : Xd[(long long)gj * n + gi];
  };
  // the beta-term values of this thread's two result elements (the tile element and its mirror
  // written by the epilogue's two passes), loaded before the products, which they do not need
  double xb1 = 0.0, xb2 = 0.0;
  if (beta != 0.0) {
    const int e = threadIdx.x;
    if (!(diagTile && (e & 31) < (e >> 5)) && r0 + (e & 31) < n && c0 + (e >> 5) < n) xb1 = xAt(r0 + (e & 31), c0 + (e >> 5));
    if (!(diagTile && (e >> 5) <= (e & 31)) && r0 + (e >> 5) < n && c0 + (e & 31) < n) xb2 = xAt(r0 + (e >> 5), c0 + (e & 31));
  }
  // partial tiles of the waves: [wave][row][col], doubles (FP64) or floats (split-BF16)
  __shared__ double redW[kStepWaves][32][33];
  float(*redF)[32][33] = reinterpret_cast<float(*)[32][33]>(&redW[0][0][0]);
  if (!lowp) {
    double4v acc[2][2];
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) acc[a][c] = (double4v){0.0, 0.0, 0.0, 0.0};
    if (alpha != 0.0) {
      if (S.xSplit) kSplitF64FromSplit(acc, (const gbf16c*)J.Xh, (const gbf16c*)J.Xl, J.ldp, n, r0, c0, w, li, lk);
      else kSplitF64(acc, (const gdouble*)(X), (const gdouble*)(X), n, r0, c0, w, li, lk, lead);
    }
    // f64 16x16 result: lane l, register r -> row l/16 + 4r, column l%16
    for (int a = 0; a < 2; ++a)
      for (int cc = 0; cc < 2; ++cc)
        for (int r = 0; r < 4; ++r) redW[w][16 * a + lk + 4 * r][16 * cc + li] = acc[a][cc][r];
  } else {
    floatx4 acc[2][2];
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) acc[a][c] = (floatx4){0.f, 0.f, 0.f, 0.f};
    if (alpha != 0.0) {
      const long long ldp = J.ldp;
      const gbf16x8* Xh = (const gbf16x8*)(J.Xh);
      const gbf16x8* Xl = (const gbf16x8*)(J.Xl);
      // element offsets (multiples of 8) of this lane's fragments in chunk 0
      const long long oa0 = (r0 + li) * ldp + 8 * lk, oa1 = (r0 + 16 + li) * ldp + 8 * lk;
      const long long ob0 = (c0 + li) * ldp + 8 * lk, ob1 = (c0 + 16 + li) * ldp + 8 * lk;
      const int nch = (int)(ldp >> 5);
      auto load = [&](int ch, Bf16Chunk& c) {
        const long long k = 32LL * ch;
        c.ah0 = Xh[(oa0 + k) >> 3];
        c.ah1 = Xh[(oa1 + k) >> 3];
        c.al0 = Xl[(oa0 + k) >> 3];
        c.al1 = Xl[(oa1 + k) >> 3];
        c.bh0 = Xh[(ob0 + k) >> 3];
        c.bh1 = Xh[(ob1 + k) >> 3];
        c.bl0 = Xl[(ob0 + k) >> 3];
        c.bl1 = Xl[(ob1 + k) >> 3];
      };
      if (w < nch) {
        // two register sets used in turn (no copies, which the compiler would turn back into
        // loads at the top of the loop): the next chunk's loads are in flight during the products
        Bf16Chunk fa, fb;
        load(w, fa);
        int ch = w;
        for (; ch + kStepWaves < nch; ch += 2 * kStepWaves) {
          load(ch + kStepWaves, fb);
          mmaBf16x3(acc, fa);
          load(min(ch + 2 * kStepWaves, nch - 1), fa);  // valid even past the end (then unused)
          mmaBf16x3(acc, fb);
        }
        if (ch < nch) mmaBf16x3(acc, fa);
      }
    }
    // f32 16x16 result: lane l, register r -> row 4 (l/16) + r, column l%16
    for (int a = 0; a < 2; ++a)
      for (int cc = 0; cc < 2; ++cc)
        for (int r = 0; r < 4; ++r) redF[w][16 * a + 4 * lk + r][16 * cc + li] = acc[a][cc][r];
  }
  __syncthreads();
  gdoubleW* Y = (gdoubleW*)J.Y;
  gbf16w* Yh = (gbf16w*)J.Yh;
  gbf16w* Yl = (gbf16w*)J.Yl;
  double tr = 0.0, sq = 0.0;
  auto tileSum = [&](int row, int cl) {
    if (lowp) {
      float v = redF[0][row][cl];
      for (int q = 1; q < kStepWaves; ++q) v += redF[q][row][cl];
      return (double)v;
    }
    double v = redW[0][row][cl];
    for (int q = 1; q < kStepWaves; ++q) v += redW[q][row][cl];
    return v;
  };
  // A split-BF16 x3 step leaves Y as its split copies only (with the traces of what they hold);
  // an FP64 step stores Y in FP64.  One element per thread in each pass (1024 threads).
  // pass 1: Y(r0 + row, c0 + col), consecutive threads down a column
  {
    const int e = threadIdx.x;
    const int row = e & 31, cl = e >> 5;
    const int gi = r0 + row, gj = c0 + cl;
    const bool in = gi < n && gj < n;
    if (!(diagTile && row < cl)) {
      const double v = in ? alpha * tileSum(row, cl) + beta * xb1 : 0.0;
      if (lowp) {
        const double r = splitStore(v, &Yh[gj * ldp + gi], &Yl[gj * ldp + gi]);
        sq += r * r;
        if (gi == gj) tr += r;
      } else if (in) {
        Y[(long long)gj * n + gi] = v;
        sq += v * v;
        if (gi == gj) tr += v;
      }
    }
  }
  // pass 2: the mirror Y(c0 + col, r0 + row), consecutive threads along a row of the tile
  {
    const int e = threadIdx.x;
    const int cl = e & 31, row = e >> 5;
    const int gi = r0 + row, gj = c0 + cl;
    const bool in = gi < n && gj < n;
    if (!(diagTile && row <= cl)) {
      const double v = in ? alpha * tileSum(row, cl) + beta * xb2 : 0.0;
      if (lowp) {
        const double r = splitStore(v, &Yh[gi * ldp + gj], &Yl[gi * ldp + gj]);
        sq += r * r;
      } else if (in) {
        Y[(long long)gi * n + gj] = v;
        sq += v * v;
      }
    }
  }
  tr = blockSum(tr, red);
  sq = blockSum(sq, red);
  if (threadIdx.x == 0) {
    partOut[2 * (long long)t] = tr;
    partOut[2 * (long long)t + 1] = sq;
  }
}

// SP2 step for larger molecules: one TILE x TILE tile (TILE = 32 NB) per workgroup of four waves
// in a 2 x 2 arrangement, each wave owning a 16 NB x 16 NB quarter ((NB x NB) MFMA 16 x 16 blocks)
// over the whole k range, with the operands of the next k steps in flight during the products.
// Without a k split there is no cross-wave reduction: results go from the accumulators straight
// to Y (and its split copies), each element of the tile and its mirror.  Same controller,
// decisions, stored layout and partial traces as k_sp2_step; the products differ from it only in
// summation order (a molecule uses one of the two kernels for all its steps).
//  - FP64: v_mfma_f64_16x16x4f64, steps of 4 k, 7 (64-wide tiles) or 3 (128-wide) steps of
//    operands in flight;
//    D lane l register r -> (row l/16 + 4 r, column l%16).
//  - split-BF16 x3: v_mfma_f32_16x16x32_bf16 over 32-wide k chunks of the padded split copies
//    (rows/columns past ldp clamped to ldp - 1: they only feed outputs that are never stored);
//    D lane l register r -> (row 4 (l/16) + r, column l%16).
// Offset (bf16 elements) of 16-byte k group q (8 k) of a row in the LDS operand stages of the
// split products: [stage][panel][plane][row][32 k].  Group q of a row sits in slot
// q ^ ((row / 4) % 4), so the 16 rows a fragment load reads at one k group, and the 4 rows 16
// consecutive threads store, fall on 16 distinct 16-byte bank groups.
template <int TILE>
__device__ inline int stageOffset(int buf, int panel, int plane, int row, int q) {
  return ((((buf * 2 + panel) * 2 + plane) * TILE + row) << 5) + ((q ^ ((row >> 2) & 3)) << 3);
}

template <int NB>
__global__ __launch_bounds__(256) void k_sp2_step_tiled(const Sp2StepJob* __restrict__ jobs,
                                                        const int* __restrict__ tileStart, int njobs, int tileBase,
                                                        const double* __restrict__ partIn, double* __restrict__ partOut,
                                                        const Sp2Dev* __restrict__ stateIn, Sp2Dev* __restrict__ stateOut) {
  constexpr int TILE = 32 * NB, WT = 16 * NB;
  const int t = tileBase + xcdTile(blockIdx.x, gridDim.x);
  const int mol = findSegment(tileStart, njobs, t);
  const Sp2StepJob J = jobs[mol];
  const int n = J.n;
  __shared__ double red[8];
  double t1, t2;
  sumPartials(partIn, tileStart[mol], tileStart[mol + 1], red, t1, t2);
  const Sp2Dev S = sp2Advance(stateIn[J.mol], t1, t2);
  if (t == tileStart[mol] && threadIdx.x == 0) stateOut[J.mol] = S;
  if (S.done) return;
  const bool lowp = S.lowp != 0;
  const double alpha = S.ab[0], beta = S.ab[1];
  const int tn = (n + TILE - 1) / TILE;
  int local = t - tileStart[mol], ti = 0;
  while (local >= tn - ti) {
    local -= tn - ti;
    ++ti;
  }
  const int tj = ti + local;
  const bool diagTile = ti == tj;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = ti * TILE + (w >> 1) * WT, c0 = tj * TILE + (w & 1) * WT;  // this wave's quarter
  const double* __restrict__ X = J.X;
  const gdouble* Xd = (const gdouble*)X;
  gdoubleW* Y = (gdoubleW*)J.Y;
  gbf16w* Yh = (gbf16w*)J.Yh;
  gbf16w* Yl = (gbf16w*)J.Yl;
  const gbf16c* Xh16 = (const gbf16c*)J.Xh;
  const gbf16c* Xl16 = (const gbf16c*)J.Xl;
  const long long ldp = J.ldp;
  double tr = 0.0, sq = 0.0;
  // one result element (gi, gj) of this wave: stored with its mirror when on or below the
  // diagonal of a diagonal tile (every element of an off-diagonal tile, where gi < gj)
  // (a split-BF16 x3 step leaves Y as its split copies only, with the traces of what they hold;
  // X(gi, gj) of the beta term comes from the split copies after such a step)
  const bool xs = S.xSplit != 0;
  // X(gi, gj) = X(gj, gi) for the beta term (0 where it is not needed)
  auto xAt = [&](int gi, int gj) {
    if (beta == 0.0 || (diagTile && gi < gj) || gi >= n || gj >= n) return 0.0;
    return xs ? splitValue(Xh16, Xl16, (long long)gi * ldp + gj) : Xd[(long long)gi * n + gj];
  };
  auto store = [&](int gi, int gj, double prod, double x) {
    if (diagTile && gi < gj) return;
    const bool in = gi < n && gj < n;
    const double v = in ? alpha * prod + beta * x : 0.0;
    if (lowp) {
      if (gi < ldp && gj < ldp) {
        const double r = splitStore(v, &Yh[gj * ldp + gi], &Yl[gj * ldp + gi]);
        if (gi != gj) splitStore(v, &Yh[gi * ldp + gj], &Yl[gi * ldp + gj]);
        sq += gi != gj ? 2.0 * (r * r) : r * r;
        if (gi == gj) tr += r;
      }
    } else if (in) {
      Y[(long long)gj * n + gi] = v;
      if (gi != gj) Y[(long long)gi * n + gj] = v;
      sq += gi != gj ? 2.0 * (v * v) : v * v;
      if (gi == gj) tr += v;
    }
  };
  if (!lowp) {
    double4v acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int c = 0; c < NB; ++c) acc[a][c] = (double4v){0.0, 0.0, 0.0, 0.0};
    if (alpha != 0.0) {
      const gdouble* Xg = (const gdouble*)(X);
      int ra[NB], cb[NB];  // operand rows (A) and columns (B), clamped into the matrix
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        ra[a] = min(r0 + 16 * a + li, n - 1);
        cb[a] = min(c0 + 16 * a + li, n - 1);
      }
      auto load = [&](int k0, double (&va)[NB], double (&vb)[NB]) {
        if (xs) {  // the first FP64 step after split-BF16 x3 steps: X = Xh + Xl
          const long long col = (long long)min(k0 + lk, n - 1) * ldp;
#pragma unroll
          for (int a = 0; a < NB; ++a) {
            va[a] = splitValue(Xh16, Xl16, col + ra[a]);
            vb[a] = splitValue(Xh16, Xl16, col + cb[a]);
          }
          return;
        }
        const gdouble* col = Xg + (long long)min(k0 + lk, n - 1) * n;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          va[a] = col[ra[a]];
          vb[a] = col[cb[a]];
        }
      };
      auto mma = [&](const double (&va)[NB], const double (&vb)[NB], bool valid) {
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
          for (int c = 0; c < NB; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(valid ? va[a] : 0.0, vb[c], acc[a][c], 0, 0, 0);
      };
      // D register sets used in turn: the loads of the next D - 1 steps are in flight during the
      // products of one step (k past the matrix is clamped in the loads and zeroed in the products).
      // Measured with 64-wide tiles, 8 sets against 4: the 2000-atom aggregate's solve 2.15 -> 2.10 s,
      // the 256-molecule 10-300-atom batch unchanged (1.58 s FP64).
      constexpr int D = NB == 2 ? 8 : 4;
      double va[D][NB], vb[D][NB];
#pragma unroll
      for (int d = 0; d < D; ++d) load(4 * d, va[d], vb[d]);
      const int S4 = (n + 3) / 4;  // steps of 4 k
      int s = 0;
      for (; s + D < S4; s += D) {  // steps s .. s + D - 1 are full
#pragma unroll
        for (int d = 0; d < D; ++d) {
          mma(va[d], vb[d], true);
          load(4 * (s + d + D), va[d], vb[d]);
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) mma(va[d], vb[d], s + d < S4 && 4 * (s + d) + lk < n);
    }
    // by row of blocks: the X values of the beta term are all loaded before any store (the
    // stores could alias them for the compiler, which would otherwise wait on every load)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double xv[NB][4];
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) xv[c][r] = xAt(r0 + 16 * a + lk + 4 * r, c0 + 16 * c + li);
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) store(r0 + 16 * a + lk + 4 * r, c0 + 16 * c + li, acc[a][c][r], xv[c][r]);
    }
  } else {
    floatx4 acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int c = 0; c < NB; ++c) acc[a][c] = (floatx4){0.f, 0.f, 0.f, 0.f};
    // LDS of the split branch: the operand stages of the products, then the split result tile
    constexpr int kStageBytes = 2 * 2 * 2 * TILE * 32 * 2;  // [stage][panel][plane][row][32 k] bf16
    constexpr int kSplitBytes = 2 * 4 * WT * (WT + 8) * 2;  // [plane][wave][row][col + 8] bf16
    __shared__ __attribute__((aligned(16))) unsigned char sMem[kStageBytes > kSplitBytes ? kStageBytes : kSplitBytes];
    if (alpha != 0.0) {
      // The split operand panels of the tile's rows (panel 0) and columns (panel 1; a diagonal
      // tile has only panel 0), 32 k at a time, are loaded by the whole workgroup into one of two
      // LDS stages while the four waves multiply out of the other.  Every wave accumulates its
      // products in the same order as with operands read straight from memory: chunk by chunk,
      // and within a chunk Xh Xh, Xh Xl, Xl Xh per 16 x 16 block.
      constexpr int kPieces = TILE * 8 / 256;  // 16-byte pieces per thread and panel (2 planes x TILE rows x 4)
      const gbf16x8* Xh = (const gbf16x8*)(J.Xh);
      const gbf16x8* Xl = (const gbf16x8*)(J.Xl);
      unsigned short* sStage = reinterpret_cast<unsigned short*>(sMem);
      const int npan = diagTile ? 1 : 2, pb = diagTile ? 0 : 1;
      const int nch = (int)(ldp >> 5);
      bf16x8 stg[2 * kPieces];
      // piece p = threadIdx.x + 256 j: k group p & 3 of row (p >> 2) % TILE of plane (p >> 2) / TILE
      // (h, l in turn) of panel j / kPieces; rows past the matrix are clamped into it (their
      // products are not stored)
      auto gload = [&](int ch) {
#pragma unroll
        for (int j = 0; j < 2 * kPieces; ++j) {
          if (j < npan * kPieces) {
            const int p = (int)threadIdx.x + 256 * j, rest = p >> 2, row = rest % TILE, pp = rest / TILE;
            const long long grow = min((long long)(((pp >> 1) ? tj : ti) * TILE + row), ldp - 1);
            const long long e = grow * ldp + 32LL * ch + 8 * (p & 3);
            stg[j] = (pp & 1) ? Xl[e >> 3] : Xh[e >> 3];
          }
        }
      };
      auto sstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 2 * kPieces; ++j) {
          if (j < npan * kPieces) {
            const int p = (int)threadIdx.x + 256 * j, rest = p >> 2, row = rest % TILE, pp = rest / TILE;
            *(bf16x8*)(sStage + stageOffset<TILE>(buf, pp >> 1, pp & 1, row, p & 3)) = stg[j];
          }
        }
      };
      auto mma = [&](int buf) {
        bf16x8 bh[NB], bl[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          const int row = (w & 1) * WT + 16 * c + li;
          bh[c] = *(const bf16x8*)(sStage + stageOffset<TILE>(buf, pb, 0, row, lk));
          bl[c] = *(const bf16x8*)(sStage + stageOffset<TILE>(buf, pb, 1, row, lk));
        }
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          const int row = (w >> 1) * WT + 16 * a + li;
          const bf16x8 ah = *(const bf16x8*)(sStage + stageOffset<TILE>(buf, 0, 0, row, lk));
          const bf16x8 al = *(const bf16x8*)(sStage + stageOffset<TILE>(buf, 0, 1, row, lk));
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[c], acc[a][c], 0, 0, 0);
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[c], acc[a][c], 0, 0, 0);
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[c], acc[a][c], 0, 0, 0);
          }
        }
      };
      gload(0);
      sstore(0);
      __syncthreads();
      for (int ch = 0; ch < nch; ++ch) {
        // the loads of chunk ch + 1 are in flight during the products of chunk ch; the stage they
        // go to was last read in chunk ch - 1, which every wave finished before the barrier
        const bool more = ch + 1 < nch;
        if (more) gload(ch + 1);
        mma(ch & 1);
        if (more) sstore((ch + 1) & 1);
        __syncthreads();
      }
    }
    // Y in split form, staged in LDS per wave and written with 16-byte stores down the columns
    // of the quarter and (off-diagonal tiles) of its mirror.  A diagonal tile is computed whole:
    // its (i, j) and (j, i) products are the same products of a symmetric X, so it is stored
    // without mirroring.  (The operand stages are no longer read: the loop ends on a barrier.)
    auto sSplit = reinterpret_cast<unsigned short(*)[4][WT][WT + 8]>(sMem);
    unsigned short(*sH)[WT + 8] = sSplit[0][w];
    unsigned short(*sL)[WT + 8] = sSplit[1][w];
    auto xAll = [&](int gi, int gj) {  // X(gj, gi) = X(gi, gj), read along the lanes' row
      if (beta == 0.0 || gi >= n || gj >= n) return 0.0;
      return splitValue(Xh16, Xl16, (long long)gi * ldp + gj);
    };
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double xv[NB][4];
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) xv[c][r] = xAll(r0 + 16 * a + 4 * lk + r, c0 + 16 * c + li);
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rl = 16 * a + 4 * lk + r, cl = 16 * c + li, gi = r0 + rl, gj = c0 + cl;
          const double v = (gi < n && gj < n) ? alpha * (double)acc[a][c][r] + beta * xv[c][r] : 0.0;
          unsigned short h, l;
          splitBf16(v, h, l);
          sH[rl][cl] = h;
          sL[rl][cl] = l;
          const double rv = bf16Value(h) + bf16Value(l);  // the value the next step sees
          sq += diagTile ? rv * rv : 2.0 * (rv * rv);
          if (gi == gj) tr += rv;
        }
    }
    __syncthreads();
    constexpr int CH = WT / 8;  // 8-element chunks per column of the quarter
#pragma unroll
    for (int q = lane; q < WT * CH; q += 64) {
      const int col = q / CH, rc = q - col * CH;  // column col of the quarter, rows 8 rc .. 8 rc + 7
      if (c0 + col < ldp && r0 + 8 * rc < ldp) {
        ushort8 vh, vl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          vh[e] = sH[8 * rc + e][col];
          vl[e] = sL[8 * rc + e][col];
        }
        const long long o = (long long)(c0 + col) * ldp + r0 + 8 * rc;
        *(gushort8*)(Yh + o) = vh;
        *(gushort8*)(Yl + o) = vl;
      }
      if (!diagTile && r0 + col < ldp && c0 + 8 * rc < ldp) {  // mirror: column r0 + col, rows c0 + 8 rc ..
        ushort8 vh, vl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          vh[e] = sH[col][8 * rc + e];
          vl[e] = sL[col][8 * rc + e];
        }
        const long long o = (long long)(r0 + col) * ldp + c0 + 8 * rc;
        *(gushort8*)(Yh + o) = vh;
        *(gushort8*)(Yl + o) = vl;
      }
    }
  }
  tr = blockSum(tr, red);
  sq = blockSum(sq, red);
  if (threadIdx.x == 0) {
    partOut[2 * (long long)t] = tr;
    partOut[2 * (long long)t + 1] = sq;
  }
}

// T = F P for the DIIS commutator, one 32 x 32 tile of T per workgroup (every tile of every
// molecule of the list; F and P symmetric, column-major), the k range split over 16 waves and the
// partial tiles summed through LDS in wave order.
__global__ __launch_bounds__(kStepWaves * 64) void k_fp_product(const MolDev* __restrict__ mols,
                                                                const int* __restrict__ tiles32, int nmols,
                                                                const double* __restrict__ F,
                                                                const double* __restrict__ P, double* __restrict__ T) {
  const Tile32 tl = tile32Of(mols, tiles32, nmols, blockIdx.x);
  const MolDev m = mols[tl.mol];
  const int n = tl.n, r0 = tl.r0, c0 = tl.c0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  __shared__ double redW[kStepWaves][32][33];
  double4v acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = (double4v){0.0, 0.0, 0.0, 0.0};
  const F64Lead L = kSplitF64Lead((const gdouble*)(F + m.matOff), (const gdouble*)(P + m.matOff), n, r0, c0, w, li, lk);
  kSplitF64(acc, (const gdouble*)(F + m.matOff), (const gdouble*)(P + m.matOff), n, r0, c0, w, li, lk, L);
  for (int a = 0; a < 2; ++a)
    for (int cc = 0; cc < 2; ++cc)
      for (int r = 0; r < 4; ++r) redW[w][16 * a + lk + 4 * r][16 * cc + li] = acc[a][cc][r];
  __syncthreads();
  double* __restrict__ Tm = T + m.matOff;
  for (int e = threadIdx.x; e < 1024; e += kStepWaves * 64) {
    const int row = e & 31, cl = e >> 5;
    const int gi = r0 + row, gj = c0 + cl;
    if (gi < n && gj < n) {
      double v = redW[0][row][cl];
      for (int q = 1; q < kStepWaves; ++q) v += redW[q][row][cl];
      Tm[(long long)gj * n + gi] = v;
    }
  }
}

// T = F P for molecules with 64-wide SP2 tiles (whole k per wave, no reduction): one 64 x 64 tile
// of T per workgroup of four waves, each wave a 32 x 32 quarter; F and P symmetric, so both
// operands are read down columns (F(i, k) = F[k n + i], P(k, j) = P[k n + j]); four steps of
// operands in flight; rows/columns past n clamped into the matrix (their results are not
// stored), k past n zeroed in the products.  tiles64: prefix sums of ceil(n/64)^2 over mols.
__global__ __launch_bounds__(256) void k_fp_product_tiled(const MolDev* __restrict__ mols,
                                                          const int* __restrict__ tiles64, int nmols,
                                                          const double* __restrict__ F,
                                                          const double* __restrict__ P, double* __restrict__ T) {
  constexpr int NB = 2, TILE = 64, WT = 32, D = 4;
  const int t = xcdTile(blockIdx.x, gridDim.x);
  const int mol = findSegment(tiles64, nmols, t);
  const MolDev m = mols[mol];
  const int n = m.n, tn = (n + TILE - 1) / TILE, local = t - tiles64[mol];
  const int tj = local / tn, ti = local - tj * tn;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = ti * TILE + (w >> 1) * WT, c0 = tj * TILE + (w & 1) * WT;
  const gdouble* Fg = (const gdouble*)(F + m.matOff);
  const gdouble* Pg = (const gdouble*)(P + m.matOff);
  int ra[NB], cb[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    ra[a] = min(r0 + 16 * a + li, n - 1);
    cb[a] = min(c0 + 16 * a + li, n - 1);
  }
  double4v acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int c = 0; c < NB; ++c) acc[a][c] = (double4v){0.0, 0.0, 0.0, 0.0};
  auto load = [&](int k0, double (&va)[NB], double (&vb)[NB]) {
    const long long col = (long long)min(k0 + lk, n - 1) * n;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      va[a] = Fg[col + ra[a]];
      vb[a] = Pg[col + cb[a]];
    }
  };
  auto mma = [&](const double (&va)[NB], const double (&vb)[NB], bool valid) {
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int c = 0; c < NB; ++c)
        acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(valid ? va[a] : 0.0, vb[c], acc[a][c], 0, 0, 0);
  };
  double va[D][NB], vb[D][NB];
#pragma unroll
  for (int d = 0; d < D; ++d) load(4 * d, va[d], vb[d]);
  const int S4 = (n + 3) / 4;
  int s = 0;
  for (; s + D < S4; s += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      mma(va[d], vb[d], true);
      load(4 * (s + d + D), va[d], vb[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) mma(va[d], vb[d], s + d < S4 && 4 * (s + d) + lk < n);
  // f64 16x16 result: lane l, register r -> row l/16 + 4r, column l%16
  double* __restrict__ Tm = T + m.matOff;
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = r0 + 16 * a + lk + 4 * r, gj = c0 + 16 * c + li;
        if (gi < n && gj < n) Tm[(long long)gj * n + gi] = acc[a][c][r];
      }
}

}  // namespace

void launchSytrd(EigJob* jobs, int njobs, hipStream_t s) {
  // SPARROW_AMD_SYTRD=plain selects the two-pass kernel (kept for comparison)
  static const bool plain = [] {
    const char* e = getenv("SPARROW_AMD_SYTRD");
    return e && std::string(e) == "plain";
  }();
  if (njobs <= 0) return;
  if (plain) hipLaunchKernelGGL(k_sytrd, dim3(njobs), dim3(1024), 0, s, jobs, njobs);
  else hipLaunchKernelGGL(k_sytrd_fused, dim3(njobs), dim3(1024), 0, s, jobs, njobs);
}
void launchSytrdWide(EigJob* job, int n, hipStream_t s) {
  if (n <= 1) {
    hipLaunchKernelGGL(k_sytrd_fused, dim3(1), dim3(1024), 0, s, job, 1);
    return;
  }
  for (int r = 0; r + 1 < n; ++r) {
    hipLaunchKernelGGL(k_sytrd_col, dim3(1), dim3(1024), 0, s, job, r);
    const int R = (n - r - 1 + kTrdTile - 1) / kTrdTile;
    hipLaunchKernelGGL(k_sytrd_upd, dim3(R, R), dim3(1024), 0, s, job, r);
  }
  hipLaunchKernelGGL(k_sytrd_final, dim3(1), dim3(64), 0, s, job);
}
void launchStebz(EigJob* jobs, int njobs, int maxN, hipStream_t s) {
  if (njobs > 0 && maxN > 0) hipLaunchKernelGGL(k_stebz, dim3((maxN + 3) / 4, njobs), dim3(256), 0, s, jobs, njobs);
}
void launchStein(EigJob* jobs, int njobs, int maxNvec, hipStream_t s) {
  if (njobs > 0 && maxNvec > 0) hipLaunchKernelGGL(k_stein, dim3((maxNvec + 255) / 256, njobs), dim3(256), 0, s, jobs, njobs);
}
void launchOrmtr(EigJob* jobs, int njobs, int maxN, int maxNvec, hipStream_t s) {
  if (njobs <= 0 || maxNvec <= 0) return;
  // SPARROW_AMD_ORMTR=plain selects the unblocked kernel (kept for comparison)
  static const bool plain = [] {
    const char* e = getenv("SPARROW_AMD_ORMTR");
    return e && std::string(e) == "plain";
  }();
  if (plain) {
    hipLaunchKernelGGL(k_ormtr, dim3((maxNvec + 63) / 64, njobs), dim3(64), 0, s, jobs, njobs);
    return;
  }
  const int nblk = maxN > 2 ? (maxN - 2 + kOrmBlock - 1) / kOrmBlock : 0;
  for (int step = 0; step < nblk; ++step)
    hipLaunchKernelGGL(k_ormtr_blocked, dim3((maxNvec + kOrmCols - 1) / kOrmCols, njobs), dim3(256), 0, s, jobs, njobs,
                       step);
}
int ormtrGemmJobs(int n) { return n > 2 ? 2 * ((n - 2 + kOrmGemm - 1) / kOrmGemm) : 0; }
void launchOrmtrGemm(const EigJob& J, EigJob* dJ, GemmJob* hG, GemmJob* dG, const int* dZero, hipStream_t s) {
  const int n = J.n, nv = J.nvec, nrefl = n - 2;
  if (nrefl <= 0 || nv <= 0) return;
  const int nblk = (nrefl + kOrmGemm - 1) / kOrmGemm;
  const long long nn = (long long)n * n;
  double* Vr = J.work;
  double* VT = J.work + nn;
  double* Y = J.work + 2 * nn;  // m x nv of the current block
  for (int b = 0; b < nblk; ++b) {
    const int k0 = b * kOrmGemm, m = std::min(kOrmGemm, nrefl - k0), rows = n - k0 - 1;
    GemmJob& g1 = hG[2 * b];  // Y = V^T Z(k0+1:, :)
    g1 = GemmJob{};
    g1.M = m;
    g1.N = nv;
    g1.K = rows;
    g1.A = Vr + (long long)(k0 + 1) * n + k0;  // A(i, k) = u_{k0+i}(k0+1+k)
    g1.sAi = 1;
    g1.sAk = n;
    g1.B = J.Z + (long long)(k0 + 1) * nv;  // B(c, k) = Z(k0+1+k, c)
    g1.sBj = 1;
    g1.sBk = nv;
    g1.C = Y;
    g1.D = Y;
    g1.sCi = g1.sDi = nv;
    g1.sCj = g1.sDj = 1;
    g1.alpha = 1.0;
    g1.beta = 0.0;
    GemmJob& g2 = hG[2 * b + 1];  // Z(k0+1:, :) -= (V T) Y
    g2 = GemmJob{};
    g2.M = rows;
    g2.N = nv;
    g2.K = m;
    g2.A = VT + (long long)k0 * n + k0 + 1;  // A(r, i) = (V T)(k0+1+r, i)
    g2.sAi = 1;
    g2.sAk = n;
    g2.B = Y;  // B(c, i) = Y(i, c)
    g2.sBj = 1;
    g2.sBk = nv;
    g2.C = J.Z + (long long)(k0 + 1) * nv;
    g2.D = g2.C;
    g2.sCi = g2.sDi = nv;
    g2.sCj = g2.sDj = 1;
    g2.alpha = -1.0;
    g2.beta = 1.0;
  }
  hipMemcpyAsync(dG, hG, sizeof(GemmJob) * 2 * nblk, hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(k_ormtr_prep, dim3(nblk), dim3(256), 0, s, dJ, Vr, VT);
  auto tiles = [](int M, int N, int t) { return ((M + t - 1) / t) * ((N + t - 1) / t); };
  for (int b = nblk - 1; b >= 0; --b) {
    const GemmJob& g1 = hG[2 * b];
    const GemmJob& g2 = hG[2 * b + 1];
    launchGemm(dG + 2 * b, dZero, 1, tiles(g1.M, g1.N, 32), 32, s);
    const int t2 = tiles(g2.M, g2.N, 64) >= 256 ? 64 : 32;
    launchGemm(dG + 2 * b + 1, dZero, 1, tiles(g2.M, g2.N, t2), t2, s);
  }
}
void launchGemm(const GemmJob* jobs, const int* tileStart, int njobs, int ntiles, int tile, hipStream_t s) {
  if (ntiles <= 0) return;
  if (tile == 32) hipLaunchKernelGGL(k_gemm_f64<32>, dim3(ntiles), dim3(256), 0, s, jobs, tileStart, njobs);
  else hipLaunchKernelGGL(k_gemm_f64<64>, dim3(ntiles), dim3(256), 0, s, jobs, tileStart, njobs);
}
void launchFPProduct(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* F, const double* P,
                     double* T, hipStream_t s, int nmols32, const int* tiles64, int ntiles64) {
  if (nmols32 < 0) nmols32 = nmols;  // every molecule on 32-wide k-split tiles
  if (nmols32 > 0 && ntiles32 > 0)
    hipLaunchKernelGGL(k_fp_product, dim3(ntiles32), dim3(kStepWaves * 64), 0, s, mols, tiles32, nmols32, F, P, T);
  if (nmols > nmols32 && ntiles64 > 0)
    hipLaunchKernelGGL(k_fp_product_tiled, dim3(ntiles64), dim3(256), 0, s, mols + nmols32, tiles64, nmols - nmols32, F,
                       P, T);
}
void launchCommutatorDots(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* FP, double* hist,
                          int slots, const ScfMolDev* scf, double* part, double* out, hipStream_t s) {
  if (slots > kMaxDiis) throw std::runtime_error("DIIS subspace larger than 16");
  if (nmols <= 0) return;
  hipLaunchKernelGGL(k_commutator_dots, dim3(ntiles32), dim3(256), 0, s, mols, tiles32, nmols, FP, hist, slots, scf, part);
  hipLaunchKernelGGL(k_dots_finish, dim3(slots, nmols), dim3(256), 0, s, mols, tiles32, part, slots, out);
}
void launchEdiis(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* histP, const double* histF,
                 int slots, const ScfMolDev* scf, const double* eMol, double* eHist, double* part, double* pf,
                 const double* dots, double* Mmat, double* coef, bool blend, hipStream_t s) {
  if (nmols <= 0) return;
  if (slots > kMaxDiis) throw std::runtime_error("DIIS subspace larger than 16");
  hipLaunchKernelGGL(k_hist_energy, dim3((nmols + 63) / 64), dim3(64), 0, s, mols, nmols, scf, eMol, eHist, slots);
  hipLaunchKernelGGL(k_hist_pf_dots, dim3(ntiles32), dim3(256), 0, s, mols, tiles32, nmols, histP, histF, slots, scf, part);
  hipLaunchKernelGGL(k_dots_finish, dim3(2 * slots, nmols), dim3(256), 0, s, mols, tiles32, part, 2 * slots, pf);
  hipLaunchKernelGGL(k_ediis_solve, dim3((nmols + 63) / 64), dim3(64), 0, s, mols, nmols, scf, pf, eHist, dots, Mmat, coef,
                     slots, blend ? 1 : 0);
}
void launchDiisSolve(const MolDev* mols, int nmols, const ScfMolDev* scf, const double* dots, double* Bmat, double* coef,
                     int slots, hipStream_t s) {
  if (nmols > 0) hipLaunchKernelGGL(k_diis_solve, dim3(nmols), dim3(64), 0, s, mols, nmols, scf, dots, Bmat, coef, slots);
}
void launchScfCheck(const int* molList, int nmols, ScfMolDev* scf, const double* eMol, const double* rmsd, double energyTol,
                    double rmsdTol, double lowpStart, double lowpTol, int maxIterations, hipStream_t s) {
  if (nmols > 0)
    hipLaunchKernelGGL(k_scf_check, dim3((nmols + 63) / 64), dim3(64), 0, s, molList, nmols, scf, eMol, rmsd, energyTol,
                       rmsdTol, lowpStart, lowpTol, maxIterations);
}
// blocks per molecule of an elementwise pass over n x n matrices: about 4 elements per thread,
// at most 16384 blocks in all
static int matBlocks(int maxN, int nmols) {
  const long long want = ((long long)maxN * maxN + 1023) / 1024;
  return (int)std::max(1LL, std::min({want, 1024LL, 16384LL / std::max(1, nmols)}));
}
void launchDiisExtrapolate(const MolDev* mols, int nmols, int maxN, const double* hist, int slots, const double* coef,
                           double* dst, hipStream_t s) {
  if (nmols > 0)
    hipLaunchKernelGGL(k_diis_extrapolate, dim3(matBlocks(maxN, nmols), nmols), dim3(256), 0, s, mols, hist, slots, coef, dst);
}
void launchCopyMats(const MolDev* mols, int nmols, int maxN, const double* src, bool srcHist, double* dst, bool dstHist,
                    const ScfMolDev* scf, int slots, hipStream_t s) {
  if (nmols > 0)
    hipLaunchKernelGGL(k_copy_mats, dim3(matBlocks(maxN, nmols), nmols), dim3(256), 0, s, mols, src, srcHist ? 1 : 0, dst,
                       dstHist ? 1 : 0, scf, slots);
}
void launchSp2Init(const MolDev* mols, int nmols, int maxN, const double* F, const Sp2StepJob* jobs, const int* tileStart,
                   int ntiles, double* lo, double* hi, double* part, Sp2Dev* state, int maxSteps, const ScfMolDev* scf,
                   hipStream_t s, double* lz, double* lzCoef, int lzMinN) {
  if (nmols <= 0) return;
  hipLaunchKernelGGL(k_sp2_discs, dim3((maxN + 3) / 4, nmols), dim3(256), 0, s, mols, F, lo, hi);
  const int minN = std::max(lzMinN, kLzMinN);
  const bool lanczos = lz && lzCoef && maxN >= minN;
  if (lanczos) {
    const int nl = std::min(maxN, kLzMaxN);
    for (int j = 0; j < kLzSteps; ++j)
      hipLaunchKernelGGL(k_lanczos_step, dim3((nl + kLzCols - 1) / kLzCols, nmols), dim3(256), nl * sizeof(double), s, mols,
                         F, lz, lzCoef, j, minN);
  }
  hipLaunchKernelGGL(k_sp2_range, dim3(nmols), dim3(256), 0, s, mols, lo, hi, state, maxSteps, scf);
  if (lanczos) hipLaunchKernelGGL(k_lanczos_bounds, dim3(nmols), dim3(256), 0, s, mols, lz, lzCoef, state, minN);
  hipLaunchKernelGGL(k_sp2_x0, dim3(ntiles), dim3(256), 0, s, jobs, tileStart, nmols, F, mols, state, part);
}
int sp2LanczosCoefs() { return 2 * kLzSteps + 1; }
void launchSp2Decide(const Sp2StepJob* jobs, const int* tileStart, int njobs, const double* partIn, const Sp2Dev* stateIn,
                     Sp2Dev* stateOut, hipStream_t s) {
  if (njobs > 0) hipLaunchKernelGGL(k_sp2_decide, dim3(njobs), dim3(256), 0, s, jobs, tileStart, partIn, stateIn, stateOut);
}
void launchPFromX(const MolDev* mols, int nmols, const Sp2Dev* state, const double* X0, const double* X1, double* P,
                  hipStream_t s) {
  if (nmols > 0) hipLaunchKernelGGL(k_p_from_x, dim3(64, nmols), dim3(256), 0, s, mols, state, X0, X1, P);
}
void launchSp2Step(const Sp2StepJob* jobs, const int* tileStart, int njobs, int nt32, int nt64, int nt128,
                   const double* partIn, double* partOut, const Sp2Dev* stateIn, Sp2Dev* stateOut, hipStream_t s) {
  if (nt32 > 0)
    hipLaunchKernelGGL(k_sp2_step, dim3(nt32), dim3(kStepWaves * 64), 0, s, jobs, tileStart, njobs, partIn, partOut,
                       stateIn, stateOut);
  if (nt64 > 0)
    hipLaunchKernelGGL(k_sp2_step_tiled<2>, dim3(nt64), dim3(256), 0, s, jobs, tileStart, njobs, nt32, partIn, partOut,
                       stateIn, stateOut);
  if (nt128 > 0)
    hipLaunchKernelGGL(k_sp2_step_tiled<4>, dim3(nt128), dim3(256), 0, s, jobs, tileStart, njobs, nt32 + nt64, partIn,
                       partOut, stateIn, stateOut);
}
void launchRmsdUpdate(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* P, double* Pold,
                      double* part, double* rmsd, hipStream_t s) {
  if (nmols <= 0) return;
  hipLaunchKernelGGL(k_rmsd_tiles, dim3(ntiles32), dim3(256), 0, s, mols, tiles32, nmols, P, Pold, part);
  hipLaunchKernelGGL(k_rmsd_finish, dim3(nmols), dim3(256), 0, s, mols, tiles32, part, rmsd);
}

}  // namespace sparrow_amd
