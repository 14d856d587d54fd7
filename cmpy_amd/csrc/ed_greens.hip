// Green's-function kernels.
//
// Reference counterparts:
//   _apply_{creation,annihilation}_{up,dn}   cmpy/operators.py:652-703
//   CreationOperator / AnnihilationOperator  cmpy/operators.py:706-791
//   _accumulate_sum / accumulate_gf          cmpy/exactdiag.py:110-134
//   solve_sector (np.linalg.eigh)            cmpy/exactdiag.py:46-57
// The continued fraction has no reference counterpart (cmpy has none).
#include "ed_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdint>
#include <cstdlib>
#include <mutex>

namespace ed {

// ---------------------------------------------------------------------------
// c^dagger / c as gathers over the target sector.  The reference scatters
// source blocks into a zeroed target (operators.py:652-703); a gather writes
// every target element exactly once, so no atomics and no separate memset.
// ---------------------------------------------------------------------------

__global__ void k_unrank_states(uint32_t* out, int64_t count, int num_sites, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = comb_unrank(i, num_sites, n);
}

// sigma = up: one block per target row (grid-stride), lanes over columns;
// blockIdx.y picks the vector (source stride xs, target stride ys).
__global__ __launch_bounds__(256) void k_ladder_up(int num_sites, int n_tgt, int pos, int create, int jw,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   int64_t rows_tgt, int64_t cols, int64_t xs, int64_t ys) {
  x += (int64_t)blockIdx.y * xs;
  y += (int64_t)blockIdx.y * ys;
  const uint32_t bit = 1u << pos;
  for (int64_t it = blockIdx.x; it < rows_tgt; it += gridDim.x) {
    const uint32_t st = comb_unrank(it, num_sites, n_tgt);
    const bool has = (st & bit) != 0;
    const bool valid = create ? has : !has;
    int64_t is = 0;
    double sg = 1.0;
    if (valid) {
      const uint32_t ss = st ^ bit;
      is = comb_rank(ss);
      if (jw && (popc32(ss & (bit - 1u)) & 1)) sg = -1.0;
    }
    for (int64_t c = threadIdx.x; c < cols; c += blockDim.x)
      y[it * cols + c] = valid ? sg * x[is * cols + c] : 0.0;
  }
}

// sigma = dn: lanes over target columns, grid.y over row chunks, grid.z over
// vectors (source stride xs, target stride ys).
__global__ __launch_bounds__(256) void k_ladder_dn(int num_sites, int n_tgt, int pos, int create, int jw,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   int64_t rows, int64_t cols_src, int64_t cols_tgt,
                                                   int64_t rows_per_block, int64_t xs, int64_t ys) {
  x += (int64_t)blockIdx.z * xs;
  y += (int64_t)blockIdx.z * ys;
  const int64_t jt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jt >= cols_tgt) return;
  const uint32_t bit = 1u << pos;
  const uint32_t st = comb_unrank(jt, num_sites, n_tgt);
  const bool has = (st & bit) != 0;
  const bool valid = create ? has : !has;
  int64_t js = 0;
  double sg = 1.0;
  if (valid) {
    const uint32_t ss = st ^ bit;
    js = comb_rank(ss);
    if (jw && (popc32(ss & (bit - 1u)) & 1)) sg = -1.0;
  }
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int64_t r = r0; r < r1; ++r) y[r * cols_tgt + jt] = valid ? sg * x[r * cols_src + js] : 0.0;
}

static int ladder(int num_sites, int n_up, int n_dn, int pos, int sigma, int sign_mode, int create,
                  const double* x, double* y, int64_t nvec, hipStream_t s) {
  ED_REQUIRE(num_sites >= 1 && num_sites <= ED_MAX_SITES, "num_sites=%d outside [1, %d]", num_sites,
             ED_MAX_SITES);
  ED_REQUIRE(n_up >= 0 && n_up <= num_sites && n_dn >= 0 && n_dn <= num_sites, "sector (%d, %d) invalid",
             n_up, n_dn);
  ED_REQUIRE(pos >= 0 && pos < num_sites, "pos=%d outside [0, %d)", pos, num_sites);
  ED_REQUIRE(sigma == ED_UP || sigma == ED_DN, "sigma must be ED_UP or ED_DN");
  ED_REQUIRE(sign_mode == ED_SIGN_REFERENCE || sign_mode == ED_SIGN_JORDAN_WIGNER, "bad sign_mode");
  ED_REQUIRE(x && y, "NULL buffer");
  ED_REQUIRE(nvec >= 0, "nvec < 0");
  const int d = create ? 1 : -1;
  const int nu_t = sigma == ED_UP ? n_up + d : n_up;
  const int nd_t = sigma == ED_DN ? n_dn + d : n_dn;
  ED_REQUIRE(nu_t >= 0 && nu_t <= num_sites && nd_t >= 0 && nd_t <= num_sites,
             "target sector (%d, %d) does not exist", nu_t, nd_t);
  const int64_t su = (int64_t)binom_u(num_sites, n_up), sd = (int64_t)binom_u(num_sites, n_dn);
  const int64_t tu = (int64_t)binom_u(num_sites, nu_t), td = (int64_t)binom_u(num_sites, nd_t);
  const int jw = sign_mode == ED_SIGN_JORDAN_WIGNER;
  // vectors in launches of at most 65535 (the y / z grid limit)
  for (int64_t v0 = 0; v0 < nvec; v0 += 65535) {
    const unsigned nv = (unsigned)std::min<int64_t>(nvec - v0, 65535);
    const double* xv = x + v0 * su * sd;
    double* yv = y + v0 * tu * td;
    if (sigma == ED_UP) {
      const unsigned g = (unsigned)std::min<int64_t>(tu, std::max<int64_t>(1, 65535 / nv));
      hipLaunchKernelGGL(k_ladder_up, dim3(g, nv), dim3(256), 0, s, num_sites, nu_t, pos, create, jw, xv, yv, tu,
                         td, su * sd, tu * td);
    } else {
      const int64_t rpb = 64;
      const int64_t chunks = (su + rpb - 1) / rpb;
      ED_REQUIRE(chunks <= 65535, "sector too large for the dn ladder grid");
      const dim3 g((unsigned)((td + 255) / 256), (unsigned)chunks, nv);
      hipLaunchKernelGGL(k_ladder_dn, g, dim3(256), 0, s, num_sites, nd_t, pos, create, jw, xv, yv, su, sd, td, rpb,
                         su * sd, tu * td);
    }
  }
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// ---------------------------------------------------------------------------
// Continued fraction and pole sums (complex FP64, one lane per frequency).
// ---------------------------------------------------------------------------
struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  } else {
    const double r = b.re / b.im, den = b.re * r + b.im;
    return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
  }
}

__global__ __launch_bounds__(256) void k_cf_eval(const double* __restrict__ alpha, const double* __restrict__ b2,
                                                 int m, double norm2, double shift, int sign,
                                                 const double* __restrict__ z, int64_t nz, double* __restrict__ g,
                                                 int accumulate) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  const cplx w = {sign * z[2 * k] + shift, sign * z[2 * k + 1]};
  cplx t = {0.0, 0.0};
  for (int j = m - 1; j >= 1; --j) t = cdiv({b2[j - 1], 0.0}, {w.re - alpha[j] - t.re, w.im - t.im});
  cplx r = cdiv({sign * norm2, 0.0}, {w.re - alpha[0] - t.re, w.im - t.im});
  if (accumulate) {
    g[2 * k] += r.re;
    g[2 * k + 1] += r.im;
  } else {
    g[2 * k] = r.re;
    g[2 * k + 1] = r.im;
  }
}

__global__ __launch_bounds__(256) void k_pole_sum(const double* __restrict__ w, const double* __restrict__ p,
                                                  int64_t m, const double* __restrict__ z, int64_t nz,
                                                  double* __restrict__ g, int accumulate) {
  __shared__ double sw[1024], sp[1024];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  cplx acc = {0.0, 0.0};
  const cplx zz = k < nz ? cplx{z[2 * k], z[2 * k + 1]} : cplx{0.0, 1.0};
  for (int64_t base = 0; base < m; base += 1024) {
    const int64_t cnt = m - base < 1024 ? m - base : 1024;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < cnt; i += blockDim.x) {
      sw[i] = w[base + i];
      sp[i] = p[base + i];
    }
    __syncthreads();
    for (int64_t i = 0; i < cnt; ++i) {
      const cplx q = cdiv({sw[i], 0.0}, {zz.re - sp[i], zz.im});
      acc.re += q.re;
      acc.im += q.im;
    }
  }
  if (k < nz) {
    if (accumulate) {
      g[2 * k] += acc.re;
      g[2 * k + 1] += acc.im;
    } else {
      g[2 * k] = acc.re;
      g[2 * k + 1] = acc.im;
    }
  }
}

// Time-domain sum over the Gauss quadrature of a vector's spectral measure:
// nodes theta[j] (Ritz values of the Lanczos tridiagonal), weights
// vecs[j * ldc]^2 (first components of its eigenvectors, column-major):
//   g[k] (+)= -i s norm2 sum_j w_j exp(-i s (theta_j - e0) t_k),  s = sign.
constexpr int kTevoMaxSteps = 8192;  // the m x m eigenvectors take 512 MiB at the limit
__global__ __launch_bounds__(256) void k_tevo_sum(const double* __restrict__ theta, const double* __restrict__ vecs,
                                                  int m, int ldc, double norm2, double e0, int sign,
                                                  const double* __restrict__ t, int64_t nt, double* __restrict__ g,
                                                  int accumulate) {
  __shared__ double sw[1024], sf[1024];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double tk = k < nt ? t[k] : 0.0;
  double re = 0.0, im = 0.0;
  for (int base = 0; base < m; base += 1024) {
    const int cnt = m - base < 1024 ? m - base : 1024;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const double c0 = vecs[(int64_t)(base + i) * ldc];
      sw[i] = c0 * c0;
      sf[i] = sign * (theta[base + i] - e0);
    }
    __syncthreads();
    for (int i = 0; i < cnt; ++i) {
      double sn, cs;
      sincos(sf[i] * tk, &sn, &cs);
      re = fma(sw[i], cs, re);
      im = fma(-sw[i], sn, im);
    }
  }
  if (k < nt) {
    // (-i s)(re + i im) = s im - i s re
    const double gr = sign * im * norm2, gi = -sign * re * norm2;
    if (accumulate) {
      g[2 * k] += gr;
      g[2 * k + 1] += gi;
    } else {
      g[2 * k] = gr;
      g[2 * k + 1] = gi;
    }
  }
}

// Finite-T pair sum, stage 1: block (z tile, pair group) -> partial[group][z].
// Pairs p = m * n + ni of the group are staged in LDS kPairTile at a time as
// weight w = ov^2 (e^{-beta(E_n - emin)} + e^{-beta(E_m - emin)}) and pole
// E_m - E_n, computed once per block; each lane then sums w / (z - pole) for
// its frequency as w (dx - i y) / (dx^2 + y^2), dx = Re z - pole, with the
// reciprocal from the hardware estimate (v_rcp_f64) refined by two Newton
// steps r <- r + r (1 - d r), each of which squares the relative error.
// With a second overlap ov2 (off-diagonal G_ij) the weight takes ov * ov2
// in place of ov^2.
//
// Screening (finite beta): a pair whose two Boltzmann factors are both below
// e^-kPairScreen (1e-20) is skipped.  keep[0] / keep[1] hold one past the
// last n / m whose factor reaches it, so the pairs summed are those with
// ni < keep[0] (any mi) and those with mi < keep[1] (ni >= keep[0]).  Since
// sum_m |<m|c+|n>|^2 <= 1, the skipped terms of one sector pair add up to at
// most (N + M) 1e-20 / Im z before the division by Z >= 1.
constexpr int kPairTile = 1024;
constexpr int64_t kPairsPerGroup = 8192;
constexpr double kPairScreen = 46.0;
__global__ __launch_bounds__(256) void k_pairs_keep(const double* __restrict__ en, int64_t n,
                                                    const double* __restrict__ em, int64_t m, double beta,
                                                    double emin, int screen, int64_t* __restrict__ keep) {
  __shared__ int64_t sh[2][4];
  int64_t kn = 0, km = 0;
  if (screen && isfinite(beta)) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
      if (beta * (en[i] - emin) <= kPairScreen) kn = i + 1;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x)
      if (beta * (em[i] - emin) <= kPairScreen) km = i + 1;
  } else {
    kn = n;
    km = m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    kn = max(kn, (int64_t)__shfl_down(kn, off, 64));
    km = max(km, (int64_t)__shfl_down(km, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = kn;
    sh[1][threadIdx.x >> 6] = km;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      kn = max(kn, sh[0][w]);
      km = max(km, sh[1][w]);
    }
    keep[0] = max(kn, sh[0][0]);
    keep[1] = max(km, sh[1][0]);
  }
}

__global__ __launch_bounds__(256) void k_pairs(const double* __restrict__ ov, const double* __restrict__ ov2,
                                               const double* __restrict__ en, int64_t n,
                                               const double* __restrict__ em, int64_t m, double beta, double emin,
                                               const int64_t* __restrict__ keep, const double* __restrict__ z,
                                               int64_t nz, double* __restrict__ partial) {
  __shared__ double sw[kPairTile], sp[kPairTile];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t grp = blockIdx.y;
  const int64_t nk = keep[0], mk = keep[1], rest = n - nk;
  const int64_t head = m * nk;
  const int64_t np = head + mk * rest;
  const int64_t p0 = grp * kPairsPerGroup;
  const int64_t p1 = p0 + kPairsPerGroup < np ? p0 + kPairsPerGroup : np;
  const double zr = k < nz ? z[2 * k] : 0.0, zi = k < nz ? z[2 * k + 1] : 1.0;
  const double zi2 = zi * zi;
  const bool fin = isfinite(beta);
  double re = 0.0, sr = 0.0;
  for (int64_t t0 = p0; t0 < p1; t0 += kPairTile) {
    const int cnt = (int)(p1 - t0 < kPairTile ? p1 - t0 : kPairTile);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const int64_t p = t0 + i;
      int64_t mi, ni;
      if (p < head) {
        mi = p / nk;
        ni = p - mi * nk;
      } else {
        const int64_t q = p - head;
        mi = q / rest;
        ni = nk + (q - mi * rest);
      }
      const int64_t at = mi * n + ni;
      const double o = ov[at] * (ov2 ? ov2[at] : ov[at]);
      const double e_m = em[mi], e_n = en[ni];
      sw[i] = o * (fin ? exp(-beta * (e_n - emin)) + exp(-beta * (e_m - emin)) : 2.0);
      sp[i] = e_m - e_n;
    }
    __syncthreads();
    for (int i = 0; i < cnt; ++i) {
      const double dx = zr - sp[i];
      const double d = fma(dx, dx, zi2);
      const double r0 = __builtin_amdgcn_rcp(d);
      const double r1 = fma(r0, fma(-d, r0, 1.0), r0);
      const double r = fma(r1, fma(-d, r1, 1.0), r1);
      const double wr = sw[i] * r;
      re = fma(wr, dx, re);
      sr += wr;
    }
  }
  if (k < nz) {
    partial[2 * (grp * nz + k)] = re;
    partial[2 * (grp * nz + k) + 1] = -zi * sr;
  }
}

__global__ __launch_bounds__(256) void k_pairs_reduce(const double* __restrict__ partial, int64_t ng, int64_t nz,
                                                      double* __restrict__ g) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double re = 0.0, im = 0.0;
  for (int64_t gi = 0; gi < ng; ++gi) {
    re += partial[2 * (gi * nz + k)];
    im += partial[2 * (gi * nz + k) + 1];
  }
  g[2 * k] += re;
  g[2 * k + 1] += im;
}

// C[M x N] = A B^T with A: [M][K], B: [N][K] (vectors stored one after another),
// 64x64 output tiles, 16-wide K slabs staged in LDS, 4x4 outputs per lane.
__global__ __launch_bounds__(256) void k_gemm_nt(const double* __restrict__ A, const double* __restrict__ B,
                                                 double* __restrict__ Cm, int64_t M, int64_t N, int64_t K) {
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64 + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t m0 = (int64_t)blockIdx.y * 64, n0 = (int64_t)blockIdx.x * 64;
  double acc[4][4] = {};
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int i = threadIdx.x; i < 16 * 64; i += 256) {
      const int row = i / 16, kk = i % 16;
      const int64_t kg = k0 + kk;
      As[kk][row] = (kg < K && m0 + row < M) ? A[(m0 + row) * K + kg] : 0.0;
      Bs[kk][row] = (kg < K && n0 + row < N) ? B[(n0 + row) * K + kg] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t r = m0 + ty * 4 + i, c = n0 + tx * 4 + j;
      if (r < M && c < N) Cm[r * N + c] = acc[i][j];
    }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// C[M x N] = A B^T as k_gemm_nt, on the FP64 matrix cores: a 64x64 output
// tile per block, each of the four waves a 32x32 quarter of it as 2x2
// v_mfma_f64_16x16x4 tiles.  16-wide K slabs are staged in LDS (zero-padded
// past M, N and K).  Operand lanes: lane l gives A[l % 16][l / 16] and
// B^T[l / 16][l % 16] of a 16x4 / 4x16 step; result element r of lane l is
// C[l / 16 + 4 r][l % 16].
__global__ __launch_bounds__(256) void k_gemm_nt_mfma(const double* __restrict__ A, const double* __restrict__ B,
                                                      double* __restrict__ Cm, int64_t M, int64_t N, int64_t K) {
  __shared__ double As[64][16 + 1];
  __shared__ double Bs[64][16 + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.y * 64, n0 = (int64_t)blockIdx.x * 64;
  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int i = threadIdx.x; i < 64 * 16; i += 256) {
      const int row = i >> 4, kk = i & 15;
      const int64_t kg = k0 + kk;
      As[row][kk] = (kg < K && m0 + row < M) ? A[(m0 + row) * K + kg] : 0.0;
      Bs[row][kk] = (kg < K && n0 + row < N) ? B[(n0 + row) * K + kg] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      const double a0 = As[wm + li][ks + lk], a1 = As[wm + 16 + li][ks + lk];
      const double b0 = Bs[wn + li][ks + lk], b1 = Bs[wn + 16 + li][ks + lk];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + ti * 16 + lk + 4 * r, col = n0 + wn + tj * 16 + li;
        if (row < M && col < N) Cm[row * N + col] = acc[ti][tj][r];
      }
}

// Occupation partials: block (c, k) sums w_k |V_k[i]|^2 over the indices i of
// chunk c whose up/dn/both spins occupy `pos`.
__global__ __launch_bounds__(256) void k_occ(const double* __restrict__ V, int64_t D, int64_t num_dn,
                                             const uint32_t* __restrict__ st_up,
                                             const uint32_t* __restrict__ st_dn, uint32_t bit, int kind,
                                             const double* __restrict__ w, int64_t chunk, double* partial) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.x, k = blockIdx.y;
  const double* v = V + k * D;
  double acc = 0.0;
  const int64_t i1 = (c + 1) * chunk < D ? (c + 1) * chunk : D;
  for (int64_t i = c * chunk + threadIdx.x; i < i1; i += blockDim.x) {
    const int64_t iu = i / num_dn, id = i - iu * num_dn;
    const bool up = (st_up[iu] & bit) != 0, dn = (st_dn[id] & bit) != 0;
    const bool sel = kind == 0 ? up : (kind == 1 ? dn : (up && dn));
    if (sel) acc += v[i] * v[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[k * gridDim.x + c] = w[k] * ((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

__global__ void k_sum_serial(const double* __restrict__ partial, int64_t n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += partial[i];
  *out = s;
}

// rocBLAS handles, one per (device, stream): calls on different streams may
// run concurrently from different host threads, and a handle's stream (and
// its workspace) must not be switched under a call in flight.
static std::mutex g_handle_mu;
struct HandleSlot {
  int dev;
  hipStream_t stream;
  rocblas_handle h;
};
static std::vector<HandleSlot> g_handles;

static int get_handle(rocblas_handle* h, hipStream_t s) {
  int dev = 0;
  ED_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_handle_mu);
  for (const HandleSlot& e : g_handles)
    if (e.dev == dev && e.stream == s) {
      *h = e.h;
      return ED_OK;
    }
  ED_REQUIRE(g_handles.size() < 256, "too many (device, stream) pairs for the solver handles");
  rocblas_handle nh;
  if (rocblas_create_handle(&nh) != rocblas_status_success) return set_error(ED_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(nh, s) != rocblas_status_success) {
    (void)rocblas_destroy_handle(nh);
    return set_error(ED_ERR_HIP, "rocblas_set_stream failed");
  }
  g_handles.push_back({dev, s, nh});
  *h = nh;
  return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" {

int ed_apply_cdag(int num_sites, int n_up, int n_dn, int pos, int sigma, int sign_mode, const double* x_dev,
                  double* y_dev, int64_t nvec, void* stream) {
  clear_error();
  return ladder(num_sites, n_up, n_dn, pos, sigma, sign_mode, 1, x_dev, y_dev, nvec, as_stream(stream));
}

int ed_apply_c(int num_sites, int n_up, int n_dn, int pos, int sigma, int sign_mode, const double* x_dev,
               double* y_dev, int64_t nvec, void* stream) {
  clear_error();
  return ladder(num_sites, n_up, n_dn, pos, sigma, sign_mode, 0, x_dev, y_dev, nvec, as_stream(stream));
}

int ed_cf_eval(const double* alpha, const double* beta, int32_t m, double norm2, double shift, int32_t sign,
               const double* z_dev, int64_t nz, double* g_dev, int32_t accumulate, void* stream) {
  clear_error();
  ED_REQUIRE(m >= 0 && m <= 1000000, "m=%d invalid", m);
  ED_REQUIRE(sign == 1 || sign == -1, "sign must be +1 or -1");
  ED_REQUIRE(nz >= 0 && (nz == 0 || (z_dev && g_dev)), "NULL buffer");
  ED_REQUIRE(m == 0 || alpha, "alpha is NULL");
  ED_REQUIRE(m <= 1 || beta, "beta is NULL");
  hipStream_t s = as_stream(stream);
  if (nz == 0) return ED_OK;
  if (m == 0 || norm2 == 0.0) {
    if (!accumulate) ED_HIP(hipMemsetAsync(g_dev, 0, sizeof(double) * 2 * nz, s));
    return ED_OK;
  }
  std::vector<double> h(2 * (size_t)m);
  for (int j = 0; j < m; ++j) h[j] = alpha[j];
  for (int j = 0; j + 1 < m; ++j) h[m + j] = beta[j] * beta[j];
  // stream-ordered buffer: no device-wide synchronisation, so evaluations on
  // other streams keep running
  double* d = nullptr;
  ED_HIP(hipMallocAsync((void**)&d, sizeof(double) * 2 * m, s));
  int rc = ED_OK;
  if (hipMemcpyAsync(d, h.data(), sizeof(double) * 2 * m, hipMemcpyHostToDevice, s) != hipSuccess) {
    rc = set_error(ED_ERR_HIP, "copying CF coefficients failed");
  } else {
    hipLaunchKernelGGL(k_cf_eval, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, d, d + m, m, norm2, shift,
                       sign, z_dev, nz, g_dev, accumulate);
    if (hipGetLastError() != hipSuccess) rc = set_error(ED_ERR_HIP, "k_cf_eval launch failed");
  }
  (void)hipFreeAsync(d, s);
  // h (pageable host memory) must outlive the copy
  if (hipStreamSynchronize(s) != hipSuccess && rc == ED_OK) rc = set_error(ED_ERR_HIP, "stream sync failed");
  return rc;
}

int ed_tevo_eval(const double* alpha, const double* beta, int32_t m, double norm2, double e0, int32_t sign,
                 const double* t_dev, int64_t nt, double* g_dev, int32_t accumulate, void* stream) {
  clear_error();
  ED_REQUIRE(m >= 0 && m <= kTevoMaxSteps, "m=%d invalid (at most %d)", m, kTevoMaxSteps);
  ED_REQUIRE(sign == 1 || sign == -1, "sign must be +1 or -1");
  ED_REQUIRE(nt >= 0 && (nt == 0 || (t_dev && g_dev)), "NULL buffer");
  ED_REQUIRE(m == 0 || alpha, "alpha is NULL");
  ED_REQUIRE(m <= 1 || beta, "beta is NULL");
  hipStream_t s = as_stream(stream);
  if (nt == 0) return ED_OK;
  if (m == 0 || norm2 == 0.0) {
    if (!accumulate) ED_HIP(hipMemsetAsync(g_dev, 0, sizeof(double) * 2 * nt, s));
    return ED_OK;
  }
  rocblas_handle h;
  ED_TRY(get_handle(&h, s));
  // one allocation: diagonal [m], off-diagonal [m], eigenvectors [m * m], info
  const size_t nd = (size_t)m * (m + 2);
  double* d = nullptr;
  ED_HIP(hipMalloc(&d, sizeof(double) * nd + sizeof(rocblas_int)));
  double* off = d + m;
  double* vec = d + 2 * (size_t)m;
  rocblas_int* info = (rocblas_int*)(d + nd);
  std::vector<double> hd(2 * (size_t)m, 0.0);
  for (int j = 0; j < m; ++j) hd[j] = alpha[j];
  for (int j = 0; j + 1 < m; ++j) hd[m + j] = beta[j];
  int rc = ED_OK;
  rocblas_int hinfo = 0;
  if (hipMemcpyAsync(d, hd.data(), sizeof(double) * 2 * m, hipMemcpyHostToDevice, s) != hipSuccess) {
    rc = set_error(ED_ERR_HIP, "copying Lanczos coefficients failed");
  } else {
    const rocblas_status st = rocsolver_dstedc(h, rocblas_evect_tridiagonal, (rocblas_int)m, d, off, vec,
                                               (rocblas_int)m, info);
    if (st != rocblas_status_success) {
      rc = set_error(ED_ERR_HIP, "rocsolver_dstedc failed with status %d", (int)st);
    } else if (hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(ED_ERR_HIP, "reading rocsolver info failed");
    } else if (hinfo != 0) {
      rc = set_error(ED_ERR_NUMERIC, "rocsolver_dstedc did not converge (info=%d)", (int)hinfo);
    } else {
      hipLaunchKernelGGL(k_tevo_sum, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, d, vec, m, m, norm2, e0,
                         sign, t_dev, nt, g_dev, accumulate);
      if (hipGetLastError() != hipSuccess) rc = set_error(ED_ERR_HIP, "k_tevo_sum launch failed");
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == ED_OK) rc = set_error(ED_ERR_HIP, "stream sync failed");
  (void)hipFree(d);
  return rc;
}

int ed_pole_sum(const double* w_dev, const double* p_dev, int64_t m, const double* z_dev, int64_t nz,
                double* g_dev, int32_t accumulate, void* stream) {
  clear_error();
  ED_REQUIRE(m >= 0 && nz >= 0, "negative size");
  ED_REQUIRE(nz == 0 || (z_dev && g_dev), "NULL buffer");
  ED_REQUIRE(m == 0 || (w_dev && p_dev), "NULL buffer");
  if (nz == 0) return ED_OK;
  hipLaunchKernelGGL(k_pole_sum, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, as_stream(stream), w_dev, p_dev,
                     m, z_dev, nz, g_dev, accumulate);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// Scratch of the pair sum, in doubles: the per-group partials (2 per group
// and frequency) and the two screening bounds (int64).  Every entry is
// written by the kernels of the call that uses it before it is read.
static int64_t pair_scratch_doubles(int64_t ng, int64_t nz) { return 2 * ng * nz + 2; }

// `scratch`: device memory of pair_scratch_doubles(ng, nz) doubles, or null
// for stream-ordered scratch allocated (and freed) here.
static int lehmann_pairs(const double* ov, const double* ov2, const double* evals_dev, int64_t n,
                         const double* evals_p1_dev, int64_t m, double beta, double emin, const double* z_dev,
                         int64_t nz, double* g_dev, hipStream_t s, double* scratch = nullptr,
                         int64_t scratch_doubles = 0) {
  ED_REQUIRE(n >= 0 && m >= 0 && nz >= 0, "negative size");
  if (n == 0 || m == 0 || nz == 0) return ED_OK;
  ED_REQUIRE(ov && evals_dev && evals_p1_dev && z_dev && g_dev, "NULL buffer");
  ED_REQUIRE(!std::isnan(beta) && std::isfinite(emin), "beta/emin invalid");
  const int64_t ng = (n * m + kPairsPerGroup - 1) / kPairsPerGroup;
  ED_REQUIRE(ng <= 65535, "sector pair too large for the pair-sum grid (%lld groups)", (long long)ng);
  if (scratch && pair_scratch_doubles(ng, nz) > scratch_doubles)
    return set_error(ED_ERR_LIMIT,
                     "pair sum of %lld x %lld pairs on %lld frequencies needs %lld scratch doubles; the "
                     "fixed scratch holds %lld (2 x groups x frequencies + 2, 8192 pairs per group)",
                     (long long)m, (long long)n, (long long)nz, (long long)pair_scratch_doubles(ng, nz),
                     (long long)scratch_doubles);
  // CMPY_AMD_PAIR_SCREEN=0 sums every pair
  const char* sv = getenv("CMPY_AMD_PAIR_SCREEN");
  const int screen = !(sv && *sv && atoi(sv) == 0);
  // stream-ordered scratch: no device-wide synchronisation per sector pair
  double* part = scratch;
  if (!part) ED_HIP(hipMallocAsync((void**)&part, sizeof(double) * pair_scratch_doubles(ng, nz), s));
  int64_t* keep = reinterpret_cast<int64_t*>(part + 2 * ng * nz);
  int rc = ED_OK;
  hipLaunchKernelGGL(k_pairs_keep, dim3(1), dim3(256), 0, s, evals_dev, n, evals_p1_dev, m, beta, emin, screen, keep);
  hipLaunchKernelGGL(k_pairs, dim3((unsigned)((nz + 255) / 256), (unsigned)ng), dim3(256), 0, s, ov, ov2, evals_dev,
                     n, evals_p1_dev, m, beta, emin, keep, z_dev, nz, part);
  hipLaunchKernelGGL(k_pairs_reduce, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, part, ng, nz, g_dev);
  if (hipGetLastError() != hipSuccess) rc = set_error(ED_ERR_HIP, "pair-sum launch failed");
  if (!scratch && hipFreeAsync(part, s) != hipSuccess && rc == ED_OK) rc = set_error(ED_ERR_HIP, "hipFreeAsync failed");
  return rc;
}

// Fixed scratch of ed_lehmann_pairs_noalloc: 4 MiB per device, resident for
// the life of the process.  Host code never passes the symbol itself to a
// kernel (on the host it is only a stand-in): the device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr int64_t kPairScratchDoubles = int64_t(1) << 19;
static __device__ __align__(256) double scratch[kPairScratchDoubles];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device; null until resolved

static int pair_scratch_ptr(double** out) {
  int dev = 0;
  ED_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(ED_ERR_HIP, "hipGetSymbolAddress of the pair-sum scratch failed on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ED_OK;
}

int ed_lehmann_pairs(const double* overlap_dev, const double* evals_dev, int64_t n, const double* evals_p1_dev,
                     int64_t m, double beta, double emin, const double* z_dev, int64_t nz, double* g_dev,
                     void* stream) {
  clear_error();
  return lehmann_pairs(overlap_dev, nullptr, evals_dev, n, evals_p1_dev, m, beta, emin, z_dev, nz, g_dev,
                       as_stream(stream));
}

int ed_lehmann_pairs_noalloc(const double* overlap_dev, const double* evals_dev, int64_t n,
                             const double* evals_p1_dev, int64_t m, double beta, double emin, const double* z_dev,
                             int64_t nz, double* g_dev, void* stream) {
  clear_error();
  if (n <= 0 || m <= 0 || nz <= 0)
    return lehmann_pairs(overlap_dev, nullptr, evals_dev, n, evals_p1_dev, m, beta, emin, z_dev, nz, g_dev,
                         as_stream(stream));  // validates and returns before any scratch is needed
  double* p = nullptr;
  ED_TRY(pair_scratch_ptr(&p));
  return lehmann_pairs(overlap_dev, nullptr, evals_dev, n, evals_p1_dev, m, beta, emin, z_dev, nz, g_dev,
                       as_stream(stream), p, kPairScratchDoubles);
}

int ed_lehmann_pairs_ij(const double* overlap_i_dev, const double* overlap_j_dev, const double* evals_dev, int64_t n,
                        const double* evals_p1_dev, int64_t m, double beta, double emin, const double* z_dev,
                        int64_t nz, double* g_dev, void* stream) {
  clear_error();
  ED_REQUIRE(n == 0 || m == 0 || nz == 0 || overlap_j_dev, "NULL buffer");
  return lehmann_pairs(overlap_i_dev, overlap_j_dev, evals_dev, n, evals_p1_dev, m, beta, emin, z_dev, nz, g_dev,
                       as_stream(stream));
}

int ed_gemm_nt(const double* a_dev, const double* b_dev, double* c_dev, int64_t m, int64_t n, int64_t k,
               void* stream) {
  clear_error();
  ED_REQUIRE(k >= 0 && m >= 0 && n >= 0, "negative size");
  if (m == 0 || n == 0) return ED_OK;
  ED_REQUIRE(c_dev && (k == 0 || (a_dev && b_dev)), "NULL buffer");
  ED_REQUIRE((m + 63) / 64 <= 65535 && (n + 63) / 64 <= (int64_t)INT32_MAX, "m=%lld n=%lld too large",
             (long long)m, (long long)n);
  hipStream_t s = as_stream(stream);
  // CMPY_AMD_GEMM: 0 VALU kernel, 1 matrix-core kernel, 2 rocBLAS dgemm;
  // unset: the matrix-core kernel up to kGemmLibFlops multiply-adds, rocBLAS
  // above (measured: 1225^3 in 82 us against 258 us, 245 x 49 x 245 in 10 us
  // against 31 us; 49 x 7 x 49 in 8 us against 10 us)
  constexpr double kGemmLibFlops = 1 << 17;
  const char* v = getenv("CMPY_AMD_GEMM");
  const int mode = v && *v ? atoi(v) : ((double)m * n * k > kGemmLibFlops ? 2 : 1);
  ED_REQUIRE(mode >= 0 && mode <= 2, "CMPY_AMD_GEMM=%d is not 0, 1 or 2", mode);
  if (mode == 2 && k > 0) {
    ED_REQUIRE(m <= INT32_MAX && n <= INT32_MAX && k <= INT32_MAX, "sizes beyond 32-bit BLAS indices");
    rocblas_handle h;
    ED_TRY(get_handle(&h, s));
    const double one = 1.0, zero = 0.0;
    // row-major C = A B^T is column-major C^T = B^T' A' with A', B' the
    // column-major views (K x M, K x N) of the same memory
    const rocblas_status st = rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)n,
                                            (rocblas_int)m, (rocblas_int)k, &one, b_dev, (rocblas_int)k, a_dev,
                                            (rocblas_int)k, &zero, c_dev, (rocblas_int)n);
    if (st != rocblas_status_success) return set_error(ED_ERR_HIP, "rocblas_dgemm failed (status %d)", (int)st);
    return ED_OK;
  }
  const dim3 g((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64));
  if (mode == 0)
    hipLaunchKernelGGL(k_gemm_nt, g, dim3(256), 0, s, a_dev, b_dev, c_dev, m, n, k);
  else
    hipLaunchKernelGGL(k_gemm_nt_mfma, g, dim3(256), 0, s, a_dev, b_dev, c_dev, m, n, k);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// Occupation sum over caller-held spin-state lists (device uint32 arrays):
// partials per chunk and vector, then one serial sum, all stream-ordered.
static int occupation_sum(const double* vecs_dev, int64_t nvec, const uint32_t* su, int64_t nu,
                          const uint32_t* sd, int64_t nd, int pos, int kind, const double* w_dev, double* out_dev,
                          hipStream_t s) {
  const int64_t D = nu * nd;
  const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>((D + 4095) / 4096, 1024));
  const int64_t chunk = (D + nchunk - 1) / nchunk;
  double* part = nullptr;
  if (hipMallocAsync((void**)&part, sizeof(double) * std::max<int64_t>(nchunk * nvec, 1), s) != hipSuccess)
    return set_error(ED_ERR_NOMEM, "hipMallocAsync failed");
  int rc = ED_OK;
  if (nvec > 0 && D > 0)
    hipLaunchKernelGGL(k_occ, dim3((unsigned)nchunk, (unsigned)nvec), dim3(256), 0, s, vecs_dev, D, nd, su, sd,
                       1u << pos, kind, w_dev, chunk, part);
  hipLaunchKernelGGL(k_sum_serial, dim3(1), dim3(64), 0, s, part, (nvec > 0 && D > 0) ? nchunk * nvec : 0,
                     out_dev);
  if (hipGetLastError() != hipSuccess) rc = set_error(ED_ERR_HIP, "occupation launch failed");
  (void)hipFreeAsync(part, s);
  return rc;
}

int ed_occupation(const double* vecs_dev, int64_t nvec, int num_sites, int n_up, int n_dn, int pos, int kind,
                  const double* w_dev, double* out_dev, void* stream) {
  clear_error();
  ED_REQUIRE(num_sites >= 1 && num_sites <= ED_MAX_SITES, "num_sites=%d invalid", num_sites);
  ED_REQUIRE(n_up >= 0 && n_up <= num_sites && n_dn >= 0 && n_dn <= num_sites, "sector invalid");
  ED_REQUIRE(pos >= 0 && pos < num_sites, "pos=%d invalid", pos);
  ED_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (up), 1 (dn) or 2 (double)");
  ED_REQUIRE(nvec >= 0 && nvec <= 65535, "nvec=%lld outside [0, 65535]", (long long)nvec);
  ED_REQUIRE(out_dev && (nvec == 0 || (vecs_dev && w_dev)), "NULL buffer");
  hipStream_t s = as_stream(stream);
  const int64_t nu = (int64_t)binom_u(num_sites, n_up), nd = (int64_t)binom_u(num_sites, n_dn);
  uint32_t *su = nullptr, *sd = nullptr;
  int rc = ED_OK;
  // stream-ordered scratch, freed in stream order (no synchronisation)
  if (hipMallocAsync((void**)&su, sizeof(uint32_t) * std::max<int64_t>(nu, 1), s) != hipSuccess ||
      hipMallocAsync((void**)&sd, sizeof(uint32_t) * std::max<int64_t>(nd, 1), s) != hipSuccess) {
    rc = set_error(ED_ERR_NOMEM, "hipMallocAsync failed");
  } else {
    hipLaunchKernelGGL(k_unrank_states, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, s, su, nu, num_sites, n_up);
    hipLaunchKernelGGL(k_unrank_states, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, sd, nd, num_sites, n_dn);
    if (hipGetLastError() != hipSuccess)
      rc = set_error(ED_ERR_HIP, "occupation launch failed");
    else
      rc = occupation_sum(vecs_dev, nvec, su, nu, sd, nd, pos, kind, w_dev, out_dev, s);
  }
  if (su) (void)hipFreeAsync(su, s);
  if (sd) (void)hipFreeAsync(sd, s);
  return rc;
}

int ed_occupation_states(const double* vecs_dev, int64_t nvec, const uint32_t* up_states_dev, int64_t num_up,
                         const uint32_t* dn_states_dev, int64_t num_dn, int pos, int kind, const double* w_dev,
                         double* out_dev, void* stream) {
  clear_error();
  ED_REQUIRE(num_up >= 0 && num_dn >= 0, "negative state count");
  ED_REQUIRE(num_dn == 0 || num_up <= INT64_MAX / num_dn, "num_up * num_dn overflows");
  ED_REQUIRE(pos >= 0 && pos < 32, "pos=%d outside [0, 32)", pos);
  ED_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (up), 1 (dn) or 2 (double)");
  ED_REQUIRE(nvec >= 0 && nvec <= 65535, "nvec=%lld outside [0, 65535]", (long long)nvec);
  const bool any = nvec > 0 && num_up > 0 && num_dn > 0;
  ED_REQUIRE(out_dev && (!any || (vecs_dev && w_dev && up_states_dev && dn_states_dev)), "NULL buffer");
  return occupation_sum(vecs_dev, nvec, up_states_dev, num_up, dn_states_dev, num_dn, pos, kind, w_dev, out_dev,
                        as_stream(stream));
}

int ed_dense_eigh(double* a_dev, int64_t n, double* w_dev, void* stream) {
  clear_error();
  ED_REQUIRE(n >= 0 && n <= 65536, "n=%lld invalid", (long long)n);
  if (n == 0) return ED_OK;
  ED_REQUIRE(a_dev && w_dev, "NULL buffer");
  hipStream_t s = as_stream(stream);
  rocblas_handle h;
  ED_TRY(get_handle(&h, s));
  double* e = nullptr;
  rocblas_int* info = nullptr;
  ED_HIP(hipMalloc(&e, sizeof(double) * n));
  if (hipMalloc(&info, sizeof(rocblas_int)) != hipSuccess) {
    (void)hipFree(e);
    return set_error(ED_ERR_NOMEM, "hipMalloc failed");
  }
  int rc = ED_OK;
  const rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, a_dev,
                                             (rocblas_int)n, w_dev, e, info);
  rocblas_int hinfo = 0;
  if (st != rocblas_status_success) {
    rc = set_error(ED_ERR_HIP, "rocsolver_dsyevd failed with status %d", (int)st);
  } else if (hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess) {
    rc = set_error(ED_ERR_HIP, "reading rocsolver info failed");
  } else if (hinfo != 0) {
    rc = set_error(ED_ERR_NUMERIC, "rocsolver_dsyevd did not converge (info=%d)", (int)hinfo);
  }
  (void)hipFree(e);
  (void)hipFree(info);
  return rc;
}

int ed_dense_eigh_batched(double* a_dev, int64_t n, int64_t batch, double* w_dev, void* stream) {
  clear_error();
  ED_REQUIRE(n >= 0 && n <= 65536, "n=%lld invalid", (long long)n);
  ED_REQUIRE(batch >= 0 && batch <= (1 << 30), "batch=%lld invalid", (long long)batch);
  if (n == 0 || batch == 0) return ED_OK;
  ED_REQUIRE(a_dev && w_dev, "NULL buffer");
  hipStream_t s = as_stream(stream);
  rocblas_handle h;
  ED_TRY(get_handle(&h, s));
  double* e = nullptr;
  ED_HIP(hipMalloc(&e, sizeof(double) * n * batch + sizeof(rocblas_int) * batch));
  rocblas_int* info = (rocblas_int*)(e + n * batch);
  int rc = ED_OK;
  const rocblas_status st = rocsolver_dsyevd_strided_batched(
      h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, a_dev, (rocblas_int)n, (rocblas_stride)(n * n),
      w_dev, (rocblas_stride)n, e, (rocblas_stride)n, info, (rocblas_int)batch);
  std::vector<rocblas_int> hinfo((size_t)batch, 0);
  if (st != rocblas_status_success) {
    rc = set_error(ED_ERR_HIP, "rocsolver_dsyevd_strided_batched failed with status %d", (int)st);
  } else if (hipMemcpyAsync(hinfo.data(), info, sizeof(rocblas_int) * batch, hipMemcpyDeviceToHost, s) !=
                 hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess) {
    rc = set_error(ED_ERR_HIP, "reading rocsolver info failed");
  } else {
    for (int64_t k = 0; k < batch; ++k)
      if (hinfo[k] != 0) {
        rc = set_error(ED_ERR_NUMERIC, "rocsolver_dsyevd did not converge for matrix %lld (info=%d)", (long long)k,
                       (int)hinfo[k]);
        break;
      }
  }
  (void)hipFree(e);
  return rc;
}

}  // extern "C"
