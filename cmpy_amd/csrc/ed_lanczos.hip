// FP64 Lanczos on the device.
//
// Replaces the ARPACK ground-state call of compute_groundstate
// (cmpy/exactdiag.py:24-43, one sector) and the unnormalised coefficient
// recurrence iter_lanczos_coeffs / lanczos_ground_state (cmpy/exactdiag.py:324-375).
// Here the recurrence is normalised, does one H.v per step, starts from a seeded
// vector and keeps every scalar on the device, so a whole solve runs without a
// host round trip except at convergence checks.
//
// Stored vectors are unnormalised: v_j = s_j u_j with s_j = 1/beta_j kept on the
// device.  Step j:
//   w      = s_j H u_j - beta_j s_{j-1} u_{j-1}         (fused into the H.v sweep,
//   alpha_j = s_j <w, u_j>                               with the block partials)
//   w     -= alpha_j s_j u_j,  beta_{j+1} = |w|          (one vector pass)
//   u_{j+1} = w,  s_{j+1} = 1/beta_{j+1}
// All reductions are two-stage with a fixed summation order, so repeated runs
// reproduce every vector bit for bit.  The eigenvector's second pass
// regenerates the vectors from the stored coefficients with alpha_j folded
// into the H.v (no separate update), so its vectors equal the first pass's
// to rounding.
#include "ed_internal.h"

#include <algorithm>
#include <chrono>

namespace ed {

enum RedKind { RED_NORM0 = 0, RED_ALPHA = 1, RED_BETA = 2, RED_SUM_TO = 3, RED_SQRT_TO = 4 };

struct LzCoef {
  double* alpha;  // [M]
  double* beta;   // [M+1]
  double* scale;  // [M+1]
  double* cprev;  // [M+1]  beta_j * s_{j-1}
  double* misc;   // [16]   0: breakdown flag, 1: one, 2..: scratch scalars
};

constexpr int kRedThreads = 1024;
constexpr int kVecThreads = 256;

// beta_{j+1} = sqrt(S) from S = |w|^2, zero (and the breakdown flag set)
// where it falls below 1e-12 of |alpha_j| + beta_j or is not finite; then
// scale_{j+1} = 1/beta_{j+1} and cprev_{j+1} = beta_{j+1} scale_j.
__device__ __forceinline__ void finish_beta(LzCoef c, int j, double S) {
  double b = sqrt(fmax(S, 0.0));
  const double ref = fabs(c.alpha[j]) + (j > 0 ? c.beta[j] : 0.0);
  if (c.misc[0] != 0.0 || !(b > 1e-12 * ref) || !isfinite(b)) {
    if (!isfinite(b)) c.misc[3] = 1.0;
    b = 0.0;
    c.misc[0] = 1.0;
  }
  c.beta[j + 1] = b;
  c.scale[j + 1] = b > 0.0 ? 1.0 / b : 0.0;
  c.cprev[j + 1] = b * c.scale[j];
}

__global__ __launch_bounds__(kRedThreads) void k_reduce(const double* __restrict__ partial, int n,
                                                        int kind, LzCoef c, int j, double* out) {
  __shared__ double sh[kRedThreads / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) v += partial[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S = 0.0;
  for (int w = 0; w < kRedThreads / 64; ++w) S += sh[w];
  switch (kind) {
    case RED_NORM0: {
      if (!isfinite(S)) c.misc[3] = 1.0;
      const double b = isfinite(S) ? sqrt(fmax(S, 0.0)) : 0.0;
      c.beta[0] = b;
      c.scale[0] = b > 0.0 ? 1.0 / b : 0.0;
      c.cprev[0] = 0.0;
      c.misc[0] = b > 0.0 ? 0.0 : 1.0;
      break;
    }
    case RED_ALPHA:
      c.alpha[j] = c.scale[j] * S;
      break;
    case RED_BETA:
      finish_beta(c, j, S);
      break;
    case RED_SUM_TO:
      *out = S;
      break;
    case RED_SQRT_TO:
      *out = sqrt(fmax(S, 0.0));
      break;
  }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  return r;
}

// alpha_j = s_j S with S the sum of the np partials of <w, u_j> that the
// H.v sweep left, summed by every block in the same order (so every block
// holds the same bits); block 0 stores alpha_j.  Returns alpha_j s_j, the
// factor of u_j in the update.  sh holds blockDim.x / 64 doubles.
__device__ __forceinline__ double fused_alpha(const double* __restrict__ hv_part, int np, double* alpha,
                                              const double* scale, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += hv_part[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double S = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) S += sh[w];
  __syncthreads();  // sh is reused by the block's own sum
  const double sj = *scale;
  const double al = sj * S;
  if (blockIdx.x == 0 && threadIdx.x == 0) *alpha = al;
  return al * sj;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// u = v0 (or seeded uniform [-1, 1)), partial |u|^2.  u is stored in rows
// of ld >= nd doubles (zeros past nd); v0 and the seeded values are indexed
// by the contiguous state index iu * nd + id.
__global__ __launch_bounds__(kVecThreads) void k_start(double* __restrict__ u, const double* __restrict__ v0,
                                                       int64_t nu, int64_t nd, int64_t ld, uint64_t seed,
                                                       double* partial) {
  __shared__ double sh[kVecThreads / 64];
  double acc = 0.0;
  const int64_t n = nu * ld;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    const int64_t iu = p / ld, id = p - iu * ld;
    double v = 0.0;
    if (id < nd) {
      const int64_t k = iu * nd + id;
      if (v0)
        v = v0[k];
      else
        v = (double)(splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(k + 1))) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    }
    u[p] = v;
    acc += v * v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// dst[v][iu * nd + id] = s_v * src[v * src_stride + iu * ld + id] for vectors
// v = blockIdx.y (rows of ld doubles back to the contiguous layout); s_v =
// scale[v] (or 1/scale[v] with invert), or 1 when scale is null.
__global__ __launch_bounds__(kVecThreads) void k_compact(double* __restrict__ dst, const double* __restrict__ src,
                                                         int64_t src_stride, int64_t nu, int64_t nd, int64_t ld,
                                                         const double* scale, int invert) {
  double s = scale ? scale[blockIdx.y] : 1.0;
  if (scale && invert) s = s > 0.0 ? 1.0 / s : 0.0;
  const int64_t D = nu * nd;
  const double* sv = src + (int64_t)blockIdx.y * src_stride;
  double* dv = dst + (int64_t)blockIdx.y * D;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    const int64_t iu = k / nd, id = k - iu * nd;
    dv[k] = s * sv[iu * ld + id];
  }
}

// w -= alpha_j s_j u;  psi += yj s_j u (optional);  partial |w|^2 (optional).
// With hv_part, alpha_j is first formed from the H.v partials (fused_alpha)
// and stored; without, it is read from *alpha.  partial and hv_part must not
// overlap (blocks write the one while others still read the other).
__global__ __launch_bounds__(kVecThreads) void k_update(double* __restrict__ w, const double* __restrict__ u,
                                                        int64_t D, double* alpha, const double* scale,
                                                        double* __restrict__ psi, double yj, double* partial,
                                                        const double* __restrict__ hv_part, int np) {
  __shared__ double sh[kVecThreads / 64];
  const double a = hv_part ? fused_alpha(hv_part, np, alpha, scale, sh) : (*alpha) * (*scale);
  const double ys = yj * (*scale);
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    const double uk = u[k];
    const double wk = w[k] - a * uk;
    w[k] = wk;
    acc += wk * wk;
    if (psi) psi[k] += ys * uk;
  }
  if (partial) {
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

typedef double v2d __attribute__((ext_vector_type(2)));

// k_update on pairs of doubles (16-byte loads and stores; all pointers
// 16-byte aligned, D2 = D / 2 with D even).  NT: u and w streamed with
// nontemporal hints, for vectors far larger than the caches (the next step
// reads both from memory anyway); smaller ones keep them cache-resident.
template <bool NT>
__global__ __launch_bounds__(kVecThreads) void k_update2(double2* __restrict__ w, const double2* __restrict__ u,
                                                         int64_t D2, double* alpha, const double* scale,
                                                         double2* __restrict__ psi, double yj, double* partial,
                                                         const double* __restrict__ hv_part, int np) {
  __shared__ double sh[kVecThreads / 64];
  const double a = hv_part ? fused_alpha(hv_part, np, alpha, scale, sh) : (*alpha) * (*scale);
  const double ys = yj * (*scale);
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D2; k += stride) {
    const v2d uk = NT ? __builtin_nontemporal_load((const v2d*)(u + k)) : *(const v2d*)(u + k);
    v2d wk = NT ? __builtin_nontemporal_load((const v2d*)(w + k)) : *(const v2d*)(w + k);
    wk.x -= a * uk.x;
    wk.y -= a * uk.y;
    if (NT)
      __builtin_nontemporal_store(wk, (v2d*)(w + k));
    else
      *(v2d*)(w + k) = wk;
    acc += wk.x * wk.x + wk.y * wk.y;
    if (psi) {
      double2 pk = psi[k];
      pk.x += ys * uk.x;
      pk.y += ys * uk.y;
      psi[k] = pk;
    }
  }
  if (partial) {
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// dst = scale * src (scale from device), partial |dst|^2 optional
__global__ __launch_bounds__(kVecThreads) void k_scale(double* dst, const double* src,
                                                       int64_t D, const double* scale, int invert,
                                                       double* partial) {
  __shared__ double sh[kVecThreads / 64];
  double s = *scale;
  if (invert) s = s > 0.0 ? 1.0 / s : 0.0;
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    const double v = s * src[k];
    dst[k] = v;
    acc += v * v;
  }
  if (partial) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// partial |w|^2; gate (optional): skip the launch unless *gate != 0
__global__ __launch_bounds__(kVecThreads) void k_norm2(const double* __restrict__ w, int64_t D, double* partial,
                                                       const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ double sh[kVecThreads / 64];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) acc += w[k] * w[k];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial |w - e psi|^2
__global__ __launch_bounds__(kVecThreads) void k_resid(const double* __restrict__ w, const double* __restrict__ psi,
                                                       int64_t D, double e, double* partial) {
  __shared__ double sh[kVecThreads / 64];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    const double r = w[k] - e * psi[k];
    acc += r * r;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Block (c0, g) sums <V_i, w> for the kDotGroup vectors i of group g (then
// g + gridDim.y, ...) over the chunks c = c0, c0 + gridDim.x, ... < nchunks,
// into partial[i * gridDim.x + c0]: one read of w serves a group, and each
// thread keeps kDotGroup + 1 loads in flight (a bounded grid, so a gated-off
// launch costs one small dispatch).  gate (optional): skip the whole launch
// unless *gate != 0 (partial reorthogonalisation decides on the device
// whether a step needs it).
constexpr int kMultidotRows = 32;    // blocks along the vector groups
constexpr int kMultidotChunks = 32;  // blocks along the chunks
constexpr int kDotGroup = 4;
__global__ __launch_bounds__(kVecThreads) void k_multidot(const double* __restrict__ V, int64_t D,
                                                          const double* __restrict__ w, int64_t chunk,
                                                          double* partial, int nchunks, int nv, const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ double sh[kDotGroup][kVecThreads / 64];
  const int ngroups = (nv + kDotGroup - 1) / kDotGroup;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int g = blockIdx.y; g < ngroups; g += gridDim.y) {
    const int i0 = g * kDotGroup;
    const int nq = min(kDotGroup, nv - i0);
    const double* v0 = V + (int64_t)i0 * D;
    double acc[kDotGroup];
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) acc[q] = 0.0;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
      const int64_t k0 = (int64_t)c * chunk, k1 = std::min<int64_t>(D, k0 + chunk);
#pragma unroll 2
      for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const double wk = w[k];
#pragma unroll
        for (int q = 0; q < kDotGroup; ++q)
          if (q < nq) acc[q] += v0[(int64_t)q * D + k] * wk;
      }
    }
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) {
      double v = acc[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) sh[q][wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nq) {
      double t = 0.0;
      for (int wv = 0; wv < kVecThreads / 64; ++wv) t += sh[threadIdx.x][wv];
      partial[(int64_t)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();  // sh is reused by the next group
  }
}

// One classical Gram-Schmidt pass against the stored basis v_i = scale_i V_i:
// coef_i = scale_i^2 sum_c dots[i][c] (the ng partials of <V_i, w> per vector
// that k_multidot left, summed by every block in the same order), then
// w -= sum_i coef_i V_i, and the partial |w|^2 of this block into
// wpart[blockIdx.x].  gate as in k_multidot.
__global__ __launch_bounds__(kVecThreads) void k_gs_update(const double* __restrict__ V, int nv, int64_t D,
                                                           const double* __restrict__ dots, int ng,
                                                           const double* __restrict__ scale, double* __restrict__ w,
                                                           double* __restrict__ wpart, const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ double cs[];
  __shared__ double sh[kVecThreads / 64];
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c < ng; ++c) acc += dots[(int64_t)i * ng + c];
    cs[i] = acc * scale[i] * scale[i];
  }
  __syncthreads();
  double n2 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    double acc = 0.0;
    for (int i = 0; i < nv; ++i) acc += cs[i] * V[(int64_t)i * D + k];
    const double wk = w[k] - acc;
    w[k] = wk;
    n2 += wk * wk;
  }
  const double t = block_sum(n2, sh);
  if (threadIdx.x == 0) wpart[blockIdx.x] = t;
}

// beta_{j+1} from n partials of |w|^2 (finish_beta), unless *gate == 0.
__global__ __launch_bounds__(kRedThreads) void k_beta_gated(const double* __restrict__ partial, int n, LzCoef c,
                                                            int j, const int* gate) {
  if (*gate == 0) return;
  __shared__ double sh[kRedThreads / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) v += partial[i];
  const double S = block_sum(v, sh);
  if (threadIdx.x == 0) finish_beta(c, j, S);
}

// w -= sum_i coef_i V_i   (sign = -1)   or   w = sum_i coef_i V_i (sign = 0)
__global__ __launch_bounds__(kVecThreads) void k_axpy_multi(const double* __restrict__ V, int nv, int64_t D,
                                                            const double* __restrict__ coef,
                                                            double* __restrict__ w, int assign, const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ double cs[];
  for (int i = threadIdx.x; i < nv; i += blockDim.x) cs[i] = coef[i];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D; k += stride) {
    double acc = 0.0;
    for (int i = 0; i < nv; ++i) acc += cs[i] * V[(int64_t)i * D + k];
    w[k] = assign ? acc : w[k] - acc;
  }
}

constexpr int kAxpyBlocks = 512;  // grid of the (gated) basis combination
static int vec_blocks(int64_t D) { return (int)std::max<int64_t>(1, std::min<int64_t>((D + 255) / 256, 2048)); }
// Grid of a Lanczos step's vector update: every block first sums all of the
// H.v partials for alpha_j, so fewer, longer blocks (12-site chain sector,
// D = 731,808: 512 blocks 5% faster per solve than 2048).
static int update_blocks(int64_t D) { return std::min(vec_blocks(D), 512); }

// k_update, as 16-byte accesses where D is even and every pointer is 16-byte
// aligned; vb partials either way.  hv_part (np partials, optional): form
// and store alpha_j first (see k_update).
static void launch_update(int vb, hipStream_t s, double* w, const double* u, int64_t D, double* alpha,
                          const double* scale, double* psi, double yj, double* partial,
                          const double* hv_part = nullptr, int np = 0) {
  const uintptr_t al = (uintptr_t)w | (uintptr_t)u | (uintptr_t)psi;
  if (D % 2 == 0 && (al & 15) == 0)
    hipLaunchKernelGGL(D >= kStreamStates ? k_update2<true> : k_update2<false>, dim3(vb), dim3(kVecThreads), 0, s,
                       (double2*)w, (const double2*)u, D / 2, alpha, scale, (double2*)psi, yj, partial, hv_part, np);
  else
    hipLaunchKernelGGL(k_update, dim3(vb), dim3(kVecThreads), 0, s, w, u, D, alpha, scale, psi, yj, partial,
                       hv_part, np);
}

int launch_reduce(const double* partial, int n, double* out, int kind, hipStream_t stream) {
  LzCoef c{};
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, stream, partial, n, kind, c, 0, out);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// ---------------------------------------------------------------------------
// Host tridiagonal solver: lowest eigenpair of T (diag a[0..m-1], off b[0..m-2])
// by Sturm bisection and inverse iteration.
// ---------------------------------------------------------------------------
static int sturm_count(const double* a, const double* b, int m, double x) {
  int cnt = 0;
  double q = a[0] - x;
  const double tiny = 1e-300;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    if (q == 0.0) q = tiny;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

double tri_lowest_value(const double* a, const double* b, int m) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(a, b, m, mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Solve (T - lam I) x = r with partial pivoting (tridiagonal Gaussian elimination).
static void tri_solve_shifted(const double* a, const double* b, int m, double lam, std::vector<double>& x) {
  std::vector<double> d(m), du(m, 0.0), du2(m, 0.0), dl(m, 0.0);
  for (int i = 0; i < m; ++i) d[i] = a[i] - lam;
  for (int i = 0; i < m - 1; ++i) {
    du[i] = b[i];
    dl[i] = b[i];
  }
  std::vector<int> piv(m, 0);
  const double tiny = 1e-300;
  for (int i = 0; i < m - 1; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] == 0.0) d[i] = tiny;
      const double f = dl[i] / d[i];
      d[i + 1] -= f * du[i];
      if (i < m - 2) du2[i] = 0.0;
      dl[i] = f;
      piv[i] = 0;
    } else {
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      const double tmp = d[i + 1];
      d[i + 1] = du[i] - f * tmp;
      if (i < m - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du2[i];
      }
      du[i] = tmp;
      dl[i] = f;
      piv[i] = 1;
    }
  }
  if (d[m - 1] == 0.0) d[m - 1] = tiny;
  // forward
  for (int i = 0; i < m - 1; ++i) {
    if (piv[i] == 0) {
      x[i + 1] -= dl[i] * x[i];
    } else {
      const double t = x[i];
      x[i] = x[i + 1];
      x[i + 1] = t - dl[i] * x[i + 1];
    }
  }
  // backward
  x[m - 1] /= d[m - 1];
  if (m > 1) x[m - 2] = (x[m - 2] - du[m - 2] * x[m - 1]) / d[m - 2];
  for (int i = m - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
}

void tri_lowest_vector(const double* a, const double* b, int m, double lam, std::vector<double>& y) {
  y.assign(m, 1.0);
  if (m == 1) return;
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = std::max(scale, fabs(a[i]) + (i < m - 1 ? fabs(b[i]) : 0.0));
  const double shift = lam - 1e-13 * std::max(scale, 1.0);
  for (int i = 0; i < m; ++i) y[i] = 1.0 + 0.01 * (double)((i * 7919) % 13) / 13.0;
  for (int it = 0; it < 3; ++it) {
    tri_solve_shifted(a, b, m, shift, y);
    double n = 0.0;
    for (double v : y) n += v * v;
    n = sqrt(n);
    for (double& v : y) v /= n;
  }
  if (y[0] < 0)
    for (double& v : y) v = -v;
}

// ---------------------------------------------------------------------------
// Workspace layout
// ---------------------------------------------------------------------------
// Partials: region A (npart doubles: the H.v sweep's and the Gram-Schmidt
// dot products'), M + 2 coefficients, region B (npart2 doubles: the vector
// updates' |w|^2), then with reorthogonalisation the orthogonality estimate
// (3 (M + 2) doubles and two flags).
struct LzLayout {
  int64_t coef_off, part_off, part2_off, pro_off, vec_off, total;
  int64_t D;           // doubles per stored vector: nu rows of ld (zeros past nd)
  int64_t nu, nd, ld;  // rows, states per row, row stride
  int M;
  int nvec;
  int npart, npart2;
};

static int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

static LzLayout lz_layout(const ed_op* op, int M, int reorth, int want_vector, int basis_vectors) {
  LzLayout L{};
  L.nu = op->num_up;
  L.nd = op->num_dn;
  L.ld = vector_ld(op);
  L.D = L.nu * L.ld;
  L.M = M;
  const int P1 = matvec_num_partials(op, L.ld);
  const int P2 = vec_blocks(L.D);
  int64_t chunks = std::min<int64_t>((L.D + 4095) / 4096, 256);
  int64_t P3 = reorth ? (int64_t)(M + 1) * chunks : 0;
  L.npart = (int)std::max<int64_t>({(int64_t)P1, (int64_t)P2, P3, (int64_t)(M + 2)});
  L.npart2 = P2;
  L.nvec = reorth ? (M + 2) : (3 + (want_vector ? 1 : 0) + basis_vectors);
  L.coef_off = 0;
  L.part_off = align256(L.coef_off + sizeof(double) * (4 * (int64_t)M + 4 + 16));
  L.part2_off = align256(L.part_off + sizeof(double) * (int64_t)(L.npart + M + 2));
  L.pro_off = align256(L.part2_off + sizeof(double) * (int64_t)L.npart2);
  const int64_t pro_bytes = reorth ? (int64_t)sizeof(double) * (3 * ((int64_t)M + 2) + 1) : 0;
  L.vec_off = align256(L.pro_off + pro_bytes);
  L.total = L.vec_off + align256(sizeof(double) * L.D) * L.nvec;
  return L;
}

static LzCoef coef_ptrs(char* base, int M) {
  double* c = reinterpret_cast<double*>(base);
  LzCoef k;
  k.alpha = c;
  k.beta = c + M;
  k.scale = c + 2 * M + 1;
  k.cprev = c + 3 * M + 2;
  k.misc = c + 4 * M + 3;
  return k;
}

static double* vec_ptr(char* base, const LzLayout& L, int i) {
  return reinterpret_cast<double*>(base + L.vec_off + align256(sizeof(double) * L.D) * i);
}

// u = the start vector (v0 in the contiguous layout, or the seeded one) in
// the padded layout; partial |u|^2 per block (vec_blocks(L.D) of them).
static void launch_start(const LzLayout& L, hipStream_t s, double* u, const double* v0, uint64_t seed,
                         double* part) {
  hipLaunchKernelGGL(k_start, dim3(vec_blocks(L.D)), dim3(kVecThreads), 0, s, u, v0, L.nu, L.nd, L.ld, seed, part);
}

// nvec vectors stored in the padded layout (src_stride doubles apart) into
// the contiguous layout, vector v scaled by scale[v] (1/scale[v] with invert;
// 1 if null).
static void launch_compact(const LzLayout& L, hipStream_t s, double* dst, const double* src, int64_t src_stride,
                           int nvec, const double* scale, int invert) {
  hipLaunchKernelGGL(k_compact, dim3(vec_blocks(L.nu * L.nd), (unsigned)nvec), dim3(kVecThreads), 0, s, dst, src,
                     src_stride, L.nu, L.nd, L.ld, scale, invert);
}

__global__ void k_init_misc(double* misc) {
  if (threadIdx.x < 16) misc[threadIdx.x] = 0.0;
  if (threadIdx.x == 1) misc[1] = 1.0;
}

struct EventTimer {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  bool on = true;
  ~EventTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  hipEvent_t next() {
    if (used == ev.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        on = false;
        return nullptr;
      }
      ev.push_back(e);
    }
    return ev[used++];
  }
  double total_ms() {
    double t = 0.0;
    for (size_t i = 0; i + 1 < used; i += 2) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) t += ms;
    }
    return t;
  }
};

// One step of the three-vector recurrence: H.v with the epilogue (partials
// of <w, u_j> into part), the update with alpha_j formed in it (|w|^2
// partials into part2), beta_{j+1}.  Without reduce (the eigenvector's
// second pass) alpha_j and beta_{j+1} are already known, and the H.v's
// epilogue applies alpha_j and accumulates psi itself: no update pass.
static int run_step(ed_op* op, const LzLayout& L, LzCoef c, double* part, double* part2, double* u_prev,
                    double* u_cur, double* w, int j, hipStream_t s, EventTimer* tm, double* psi, double yj,
                    bool reduce) {
  HvEpilogue e;
  e.xprev = j > 0 ? u_prev : nullptr;
  e.sx = c.scale + j;
  e.cprev = c.cprev + j;
  e.partial = reduce ? part : nullptr;
  if (!reduce) {
    e.cself = c.alpha + j;
    e.psi = psi;
    e.psi_y = yj;
  }
  if (tm && tm->on) {
    hipEvent_t a = tm->next();
    if (a) ED_HIP(hipEventRecord(a, s));
  }
  int np = 0;
  ED_TRY(launch_matvec(op, u_cur, w, L.ld, e, s, &np));
  if (tm && tm->on) {
    hipEvent_t b = tm->next();
    if (b) ED_HIP(hipEventRecord(b, s));
  }
  if (reduce) {
    const int ub = update_blocks(L.D);
    launch_update(ub, s, w, u_cur, L.D, c.alpha + j, c.scale + j, psi, yj, part2, part, np);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part2, ub, (int)RED_BETA, c, j, nullptr);
  }
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// Full reorthogonalisation.  V[0..M] holds normalised basis vectors.
struct ReorthCtx {
  double* V;
  double* w;
  double* coefs;
  double* part2;  // region B
  int64_t chunks, chunk;
  int gx;         // k_multidot blocks along the chunks (dot partials per vector)
};

static ReorthCtx reorth_ctx(char* base, const LzLayout& L, double* part) {
  ReorthCtx r;
  r.V = vec_ptr(base, L, 0);
  r.w = vec_ptr(base, L, L.M + 1);
  r.coefs = part + L.npart;
  r.part2 = reinterpret_cast<double*>(base + L.part2_off);
  r.chunks = std::min<int64_t>((L.D + 4095) / 4096, 256);
  r.chunk = (L.D + r.chunks - 1) / r.chunks;
  r.gx = (int)std::min<int64_t>(r.chunks, kMultidotChunks);
  return r;
}

static int reorth_start(const LzLayout& L, LzCoef c, double* part, const ReorthCtx& r, const double* v0,
                        uint64_t seed, hipStream_t s) {
  const int vb = vec_blocks(L.D);
  // V_0 = the start vector, unnormalised (v_0 = scale_0 V_0)
  launch_start(L, s, r.V, v0, seed, part);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part, vb, (int)RED_NORM0, c, 0, nullptr);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// Partial reorthogonalisation (H. D. Simon, Math. Comp. 42 (1984) 115-142):
// omega[j][k] estimates <v_j, v_k> through the recurrence the Lanczos
// vectors themselves obey,
//   beta_{j+1} w_{j+1,k} = beta_{k+1} w_{j,k+1} + (alpha_k - alpha_j) w_{j,k}
//                          + beta_k w_{j,k-1} - beta_j w_{j-1,k} (+ rounding term),
// with w_{k,k} = 1 and w_{k+1,k} = eps.  When max_k |w_{j+1,k}| passes
// sqrt(eps) the new vector is orthogonalised against all previous ones (and
// so is the next one), and its row is reset to eps.  Rows live in om[3][M+2]
// (row r at (r % 3)); flags[0] = orthogonalise this step, flags[1] = the
// next step must too.
struct ProState {
  double* om;   // [3][M + 2]
  int* flags;   // [2]
  int M;
};

// The estimate's rows and flags in the workspace (LzLayout::pro_off), reset.
static int pro_init(ProState& pro, char* base, const LzLayout& L, hipStream_t s);

__global__ __launch_bounds__(256) void k_pro_init(ProState p) {
  for (int i = threadIdx.x; i < 3 * (p.M + 2); i += blockDim.x) p.om[i] = 0.0;
  if (threadIdx.x == 0) {
    p.om[0] = 1.0;  // w_{0,0}
    p.flags[0] = 0;
    p.flags[1] = 0;
  }
}

// bt = |w| before orthogonalisation, summed here from the np partials of
// |w|^2 that the vector update left.  On a step that needs no Gram-Schmidt
// this |w| is final: beta_{j+1} is set here (finish_beta), and the gated
// kernels after it do nothing.  One block of 256 threads.
__global__ __launch_bounds__(256) void k_pro_omega(ProState p, LzCoef c, int j, const double* __restrict__ wpart,
                                                   int np) {
  const double* __restrict__ alpha = c.alpha;
  const double* __restrict__ beta = c.beta;
  __shared__ double sh[256 / 64];
  __shared__ int decide;
  __shared__ double bt_sh;
  const int W = p.M + 2;
  const double* cur = p.om + (int64_t)(j % 3) * W;
  const double* prev = p.om + (int64_t)((j + 2) % 3) * W;
  double* nxt = p.om + (int64_t)((j + 1) % 3) * W;
  const double eps = 2.220446049250313e-16;
  double w2 = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) w2 += wpart[i];
  w2 = block_sum(w2, sh);
  if (threadIdx.x == 0) bt_sh = sqrt(fmax(w2, 0.0));
  __syncthreads();
  const double bt = bt_sh;
  double mx = 0.0;
  if (bt > 0.0) {
    for (int k = threadIdx.x; k < j; k += blockDim.x) {
      double t = beta[k + 1] * cur[k + 1] + (alpha[k] - alpha[j]) * cur[k] - beta[j] * prev[k];
      if (k > 0) t += beta[k] * cur[k - 1];
      t += (t >= 0.0 ? 1.0 : -1.0) * eps * (beta[k + 1] + bt) * 0.3;
      const double o = t / bt;
      nxt[k] = o;
      mx = fmax(mx, fabs(o));
    }
  }
  // block maximum
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < 4; ++w) m = fmax(m, sh[w]);
    const int force = p.flags[1];
    const int re = bt > 0.0 && (m > 1.4901161193847656e-08 || force);
    p.flags[0] = re;
    p.flags[1] = re && !force;
    decide = re;
    nxt[j] = eps;
    nxt[j + 1] = 1.0;
    if (!re) finish_beta(c, j, w2);
  }
  __syncthreads();
  if (decide)
    for (int k = threadIdx.x; k < j; k += blockDim.x) nxt[k] = eps;
}

static int pro_init(ProState& pro, char* base, const LzLayout& L, hipStream_t s) {
  double* p = reinterpret_cast<double*>(base + L.pro_off);
  pro.om = p;
  pro.flags = reinterpret_cast<int*>(p + 3 * (size_t)(pro.M + 2));
  hipLaunchKernelGGL(k_pro_init, dim3(1), dim3(256), 0, s, pro);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

// Step j, with the basis kept unnormalised (v_i = scale_i V_i, as the
// three-vector recurrence keeps its vectors): W = V_{j+1} = s_j H V_j -
// beta_j s_{j-1} V_{j-1}; alpha_j = s_j <W, V_j> (formed in the update);
// W -= alpha_j s_j V_j; classical Gram-Schmidt against v_0..v_j: two passes on
// every step, or, with a ProState, one pass where the orthogonality estimate
// asks for it (the basis is then semi-orthogonal, |<v_i, v_k>| <= sqrt(eps),
// and W's components along it are O(sqrt(eps) |W|), so one pass leaves
// O(eps)); beta_{j+1} = |W|, scale_{j+1} = 1/beta_{j+1}.  Launches per step:
// the two H.v sweeps, the update, then with a ProState the estimate (which
// sets beta_{j+1} when no pass is due) and three gated kernels (dots,
// combination, beta); without, two passes of (dots, combination) and beta.
static int reorth_step(ed_op* op, const LzLayout& L, LzCoef c, double* part, const ReorthCtx& r, int j,
                       hipStream_t s, EventTimer* tm, const ProState* pro = nullptr) {
  const int64_t D = L.D;
  const int vb = vec_blocks(D);
  double* vj = r.V + (int64_t)j * D;
  double* w = r.V + (int64_t)(j + 1) * D;
  HvEpilogue e;
  e.xprev = j > 0 ? r.V + (int64_t)(j - 1) * D : nullptr;
  e.sx = c.scale + j;
  e.cprev = c.cprev + j;
  e.partial = part;
  if (tm && tm->on) {
    hipEvent_t a = tm->next();
    if (a) ED_HIP(hipEventRecord(a, s));
  }
  int np = 0;
  ED_TRY(launch_matvec(op, vj, w, L.ld, e, s, &np));
  if (tm && tm->on) {
    hipEvent_t b = tm->next();
    if (b) ED_HIP(hipEventRecord(b, s));
  }
  const int ub = update_blocks(D);
  launch_update(ub, s, w, vj, D, c.alpha + j, c.scale + j, nullptr, 0.0, r.part2, part, np);
  const int* gate = nullptr;
  if (pro) {
    hipLaunchKernelGGL(k_pro_omega, dim3(1), dim3(256), 0, s, *pro, c, j, r.part2, ub);
    gate = pro->flags;
  }
  const int gb = std::min(vb, kAxpyBlocks);
  for (int pass = 0; pass < (pro ? 1 : 2); ++pass) {
    hipLaunchKernelGGL(k_multidot,
                       dim3((unsigned)r.gx, (unsigned)std::min((j + kDotGroup) / kDotGroup, kMultidotRows)),
                       dim3(kVecThreads), 0, s, r.V, D, w, r.chunk, part, (int)r.chunks, j + 1, gate);
    hipLaunchKernelGGL(k_gs_update, dim3((unsigned)gb), dim3(kVecThreads), sizeof(double) * (j + 1), s, r.V, j + 1,
                       D, part, r.gx, c.scale, w, r.part2, gate);
  }
  if (pro)
    hipLaunchKernelGGL(k_beta_gated, dim3(1), dim3(kRedThreads), 0, s, r.part2, gb, c, j, gate);
  else
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, r.part2, gb, (int)RED_BETA, c, j, nullptr);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

static int validate_cfg(const ed_lanczos_cfg* cfg) {
  ED_REQUIRE(cfg != nullptr, "cfg is NULL");
  ED_REQUIRE(cfg->max_iter >= 1 && cfg->max_iter <= 100000, "max_iter=%d outside [1, 100000]", cfg->max_iter);
  ED_REQUIRE(cfg->check_every >= 1, "check_every must be >= 1");
  ED_REQUIRE(cfg->tol >= 0.0 && std::isfinite(cfg->tol), "tol must be finite and >= 0");
  // the reorthogonalisation kernels stage one coefficient per basis vector in LDS
  ED_REQUIRE(cfg->reorth >= 0 && cfg->reorth <= 2, "reorth must be 0 (none), 1 (full) or 2 (partial)");
  ED_REQUIRE(!cfg->reorth || cfg->max_iter <= 4096, "reorth supports max_iter <= 4096 (got %d)", cfg->max_iter);
  return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_lanczos_workspace(const ed_op* op, const ed_lanczos_cfg* cfg, int64_t* bytes) {
  clear_error();
  ED_REQUIRE(op && bytes, "NULL argument");
  ED_TRY(validate_cfg(cfg));
  const int M = std::min<int64_t>(cfg->max_iter, op->size);
  *bytes = lz_layout(op, std::max(M, 1), cfg->reorth, cfg->want_vector, 0).total;
  return ED_OK;
}

extern "C" int ed_lanczos_groundstate(ed_op* op, const ed_lanczos_cfg* cfg, const double* v0_dev,
                                      double* psi_dev, void* work_dev, int64_t work_bytes,
                                      ed_lanczos_result* res, void* stream) {
  clear_error();
  ED_REQUIRE(op && res && work_dev, "NULL argument");
  ED_TRY(validate_cfg(cfg));
  ED_REQUIRE(!cfg->want_vector || psi_dev, "want_vector requires psi_dev");
  const auto t0 = std::chrono::steady_clock::now();
  const int M = (int)std::min<int64_t>(cfg->max_iter, op->size);
  const LzLayout L = lz_layout(op, M, cfg->reorth, cfg->want_vector, 0);
  ED_REQUIRE(work_bytes >= L.total, "workspace too small: %lld < %lld bytes", (long long)work_bytes,
             (long long)L.total);
  hipStream_t s = as_stream(stream);
  char* base = static_cast<char*>(work_dev);
  LzCoef c = coef_ptrs(base, M);
  double* part = reinterpret_cast<double*>(base + L.part_off);
  const int64_t D = L.D;
  const int vb = vec_blocks(D);
  *res = ed_lanczos_result{};

  hipLaunchKernelGGL(k_init_misc, dim3(1), dim3(64), 0, s, c.misc);
  EventTimer tm;
  tm.on = cfg->time_matvec != 0;
  double* part2 = reinterpret_cast<double*>(base + L.part2_off);
  std::vector<double> ha(M), hb(M + 1), y;
  int m = 0;
  bool converged = false, breakdown = false;
  double e0 = 0.0, ritz = 0.0;

  auto check = [&](int steps) -> int {
    ED_HIP(hipMemcpyAsync(ha.data(), c.alpha, sizeof(double) * steps, hipMemcpyDeviceToHost, s));
    ED_HIP(hipMemcpyAsync(hb.data(), c.beta, sizeof(double) * (steps + 1), hipMemcpyDeviceToHost, s));
    double flag[4];
    ED_HIP(hipMemcpyAsync(flag, c.misc, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    ED_HIP(hipStreamSynchronize(s));
    if (flag[3] != 0.0 || !std::isfinite(ha[steps - 1]))
      return set_error(ED_ERR_NUMERIC, "non-finite value in the Lanczos recurrence at step %d", steps);
    if (hb[0] == 0.0) return set_error(ED_ERR_INVALID, "the Lanczos start vector is zero");
    int mm = steps;
    breakdown = false;
    for (int k = 1; k <= steps; ++k) {
      if (hb[k] == 0.0) {
        mm = k;
        breakdown = true;
        break;
      }
    }
    m = mm;
    e0 = tri_lowest_value(ha.data(), hb.data() + 1, m);
    tri_lowest_vector(ha.data(), hb.data() + 1, m, e0, y);
    ritz = breakdown ? 0.0 : fabs(hb[m]) * fabs(y[m - 1]);
    return ED_OK;
  };

  if (!cfg->reorth) {
    double* u[3] = {vec_ptr(base, L, 0), vec_ptr(base, L, 1), vec_ptr(base, L, 2)};
    double* psi = cfg->want_vector ? vec_ptr(base, L, 3) : nullptr;
    launch_start(L, s, u[0], v0_dev, cfg->seed, part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part, vb, (int)RED_NORM0, c, 0, nullptr);
    ED_HIP(hipGetLastError());
    int ip = 2, ic = 0, iw = 1;  // u_prev, u_cur, w
    int steps = 0;
    for (int j = 0; j < M; ++j) {
      ED_TRY(run_step(op, L, c, part, part2, u[ip], u[ic], u[iw], j, s, &tm, nullptr, 0.0, true));
      steps = j + 1;
      const int old_p = ip;
      ip = ic;
      ic = iw;
      iw = old_p;
      const bool last = (j + 1 == M);
      const bool due = !cfg->fixed_iter && steps >= cfg->min_iter && (steps % cfg->check_every == 0);
      if (due || last) {
        ED_TRY(check(steps));
        if (breakdown || ritz <= cfg->tol) {
          converged = true;
          break;
        }
      }
    }
    res->iterations = steps;
    if (cfg->want_vector) {
      // second pass: regenerate u_0..u_{m-1} and accumulate psi = sum_j y_j s_j u_j
      ED_HIP(hipMemsetAsync(psi, 0, sizeof(double) * D, s));
      launch_start(L, s, u[0], v0_dev, cfg->seed, part);
      ip = 2;
      ic = 0;
      iw = 1;
      for (int j = 0; j < m; ++j) {
        if (j + 1 < m) {
          ED_TRY(run_step(op, L, c, part, part2, u[ip], u[ic], u[iw], j, s, nullptr, psi, y[j], false));
        } else {
          // last vector: accumulate only (w is scratch)
          launch_update(vb, s, u[iw], u[ic], D, c.misc + 2, c.scale + j, psi, y[j], nullptr);
        }
        const int old_p = ip;
        ip = ic;
        ic = iw;
        iw = old_p;
      }
      ED_HIP(hipGetLastError());
    }
  } else {
    // full or partial reorthogonalisation: V holds normalised basis vectors
    const ReorthCtx rc = reorth_ctx(base, L, part);
    double* V = rc.V;
    double* coefs = rc.coefs;
    ED_TRY(reorth_start(L, c, part, rc, v0_dev, cfg->seed, s));
    ProState pro{nullptr, nullptr, M};
    if (cfg->reorth == 2) ED_TRY(pro_init(pro, base, L, s));
    int steps = 0;
    for (int j = 0; j < M; ++j) {
      ED_TRY(reorth_step(op, L, c, part, rc, j, s, &tm, cfg->reorth == 2 ? &pro : nullptr));
      steps = j + 1;
      const bool last = (j + 1 == M);
      const bool due = !cfg->fixed_iter && steps >= cfg->min_iter && (steps % cfg->check_every == 0);
      if (due || last) {
        ED_TRY(check(steps));
        if (breakdown || ritz <= cfg->tol) {
          converged = true;
          break;
        }
      }
    }
    res->iterations = steps;
    if (cfg->want_vector) {
      // psi = sum_j y_j v_j = sum_j (y_j / beta_j) V_j  (beta_0 = |v0|)
      std::vector<double> yc(m);
      for (int k = 0; k < m; ++k) yc[k] = y[k] / hb[k];
      ED_HIP(hipMemcpyAsync(coefs, yc.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_axpy_multi, dim3(vb), dim3(kVecThreads), sizeof(double) * m, s, V, m, D, coefs,
                         vec_ptr(base, L, M + 1), 1, nullptr);
      ED_HIP(hipStreamSynchronize(s));  // yc goes out of scope
      ED_HIP(hipGetLastError());
    }
  }

  res->e0 = e0;
  res->ritz_residual = ritz;
  res->converged = converged ? 1 : 0;
  res->breakdown = breakdown ? 1 : 0;

  if (cfg->want_vector) {
    // normalise psi (in place, padded layout), compute the explicit residual
    // |H psi - e0 psi|, and write psi into psi_dev in the contiguous layout
    double* psi = cfg->reorth ? vec_ptr(base, L, M + 1) : vec_ptr(base, L, 3);
    hipLaunchKernelGGL(k_norm2, dim3(vb), dim3(kVecThreads), 0, s, psi, D, part, nullptr);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part, vb, (int)RED_SQRT_TO, c, 0, c.misc + 4);
    hipLaunchKernelGGL(k_scale, dim3(vb), dim3(kVecThreads), 0, s, psi, psi, D, c.misc + 4, 1, nullptr);
    double* w = vec_ptr(base, L, 0);
    HvEpilogue none{nullptr, nullptr, nullptr, nullptr};
    ED_TRY(launch_matvec(op, psi, w, L.ld, none, s, nullptr));
    hipLaunchKernelGGL(k_resid, dim3(vb), dim3(kVecThreads), 0, s, w, psi, D, e0, part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part, vb, (int)RED_SQRT_TO, c, 0, c.misc + 5);
    launch_compact(L, s, psi_dev, psi, 0, 1, nullptr, 0);
    ED_HIP(hipGetLastError());
    double r[2];
    ED_HIP(hipMemcpyAsync(r, c.misc + 4, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    ED_HIP(hipStreamSynchronize(s));
    res->residual = r[1];
    if (!std::isfinite(r[1])) return set_error(ED_ERR_NUMERIC, "non-finite ground-state residual");
  } else {
    res->residual = ritz;
  }
  ED_HIP(hipStreamSynchronize(s));
  res->matvec_ms = tm.total_ms();
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ED_OK;
}

extern "C" int ed_lanczos_coeffs_workspace(const ed_op* op, int32_t nsteps, int32_t reorth, int64_t* bytes) {
  clear_error();
  ED_REQUIRE(op && bytes, "NULL argument");
  ED_REQUIRE(nsteps >= 1 && nsteps <= 100000, "nsteps=%d outside [1, 100000]", nsteps);
  ED_REQUIRE(reorth >= 0 && reorth <= 2, "reorth must be 0 (none), 1 (full) or 2 (partial)");
  ED_REQUIRE(!reorth || nsteps <= 4096, "reorth supports nsteps <= 4096 (got %d)", nsteps);
  *bytes = lz_layout(op, nsteps, reorth ? 1 : 0, 0, 0).total;
  return ED_OK;
}

extern "C" int ed_lanczos_coeffs(ed_op* op, const double* v0_dev, int32_t nsteps, int32_t reorth, double* alpha,
                                 double* beta, double* norm2, int32_t* steps_out, double* basis_dev, void* work_dev,
                                 int64_t work_bytes, void* stream) {
  clear_error();
  ED_REQUIRE(op && v0_dev && alpha && beta && norm2 && steps_out && work_dev, "NULL argument");
  ED_REQUIRE(nsteps >= 1 && nsteps <= 100000, "nsteps=%d outside [1, 100000]", nsteps);
  ED_REQUIRE(reorth >= 0 && reorth <= 2, "reorth must be 0 (none), 1 (full) or 2 (partial)");
  ED_REQUIRE(!reorth || nsteps <= 4096, "reorth supports nsteps <= 4096 (got %d)", nsteps);
  const int M = nsteps;
  const LzLayout L = lz_layout(op, M, reorth ? 1 : 0, 0, 0);
  ED_REQUIRE(work_bytes >= L.total, "workspace too small: %lld < %lld bytes", (long long)work_bytes,
             (long long)L.total);
  hipStream_t s = as_stream(stream);
  char* base = static_cast<char*>(work_dev);
  LzCoef c = coef_ptrs(base, M);
  double* part = reinterpret_cast<double*>(base + L.part_off);
  const int64_t D = L.D;
  const int vb = vec_blocks(D);
  hipLaunchKernelGGL(k_init_misc, dim3(1), dim3(64), 0, s, c.misc);
  if (reorth) {
    const ReorthCtx rc = reorth_ctx(base, L, part);
    ED_TRY(reorth_start(L, c, part, rc, v0_dev, 0, s));
    ProState pro{nullptr, nullptr, M};
    if (reorth == 2) ED_TRY(pro_init(pro, base, L, s));
    for (int j = 0; j < M; ++j) ED_TRY(reorth_step(op, L, c, part, rc, j, s, nullptr, reorth == 2 ? &pro : nullptr));
    // basis vector j is rc.V + j D, scaled by scale_j
    if (basis_dev) launch_compact(L, s, basis_dev, rc.V, D, M, c.scale, 0);
  } else {
    double* u[3] = {vec_ptr(base, L, 0), vec_ptr(base, L, 1), vec_ptr(base, L, 2)};
    launch_start(L, s, u[0], v0_dev, 0, part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedThreads), 0, s, part, vb, (int)RED_NORM0, c, 0, nullptr);
    ED_HIP(hipGetLastError());
    int ip = 2, ic = 0, iw = 1;
    for (int j = 0; j < M; ++j) {
      if (basis_dev) launch_compact(L, s, basis_dev + (int64_t)j * L.nu * L.nd, u[ic], 0, 1, c.scale + j, 0);
      ED_TRY(run_step(op, L, c, part, reinterpret_cast<double*>(base + L.part2_off), u[ip], u[ic], u[iw], j, s,
                      nullptr, nullptr, 0.0, true));
      const int old_p = ip;
      ip = ic;
      ic = iw;
      iw = old_p;
    }
  }
  std::vector<double> ha(M), hb(M + 1);
  double flag[4];
  ED_HIP(hipMemcpyAsync(ha.data(), c.alpha, sizeof(double) * M, hipMemcpyDeviceToHost, s));
  ED_HIP(hipMemcpyAsync(hb.data(), c.beta, sizeof(double) * (M + 1), hipMemcpyDeviceToHost, s));
  ED_HIP(hipMemcpyAsync(flag, c.misc, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  if (flag[3] != 0.0) return set_error(ED_ERR_NUMERIC, "non-finite value in the Lanczos recurrence");
  for (int k = 0; k < M; ++k)
    if (!std::isfinite(ha[k])) return set_error(ED_ERR_NUMERIC, "non-finite alpha at step %d", k);
  *norm2 = hb[0] * hb[0];
  if (hb[0] == 0.0) {
    *steps_out = 0;
    return ED_OK;
  }
  int m = M;
  for (int k = 1; k <= M; ++k)
    if (hb[k] == 0.0) {
      m = k;
      break;
    }
  for (int k = 0; k < m; ++k) alpha[k] = ha[k];
  for (int k = 0; k + 1 < m; ++k) beta[k] = hb[k + 1];
  *steps_out = m;
  return ED_OK;
}
