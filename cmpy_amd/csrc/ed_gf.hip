// T = 0 Green's functions in one call each: ed_greens_cf (Lanczos continued
// fractions) and ed_greens_lehmann_t0 (dense Lehmann sums), composed on the
// device from the ladder operators, the Lanczos coefficient recurrence, the
// dense eigensolver and the continued-fraction / pole sums, so a binder in
// any language gets G(z) of a site from the ground state with one call.
//
// Both take the ground state psi0 of sector (n_up, n_dn) of num_sites sites
// and the operators of the sectors one particle above (op_p) and below (op_m)
// in spin sigma (NULL where that sector does not exist), and return
//   G(z) = <psi0| c (z + E0 - H)^-1 c+ |psi0> + <psi0| c+ (z - E0 + H)^-1 c |psi0>,
// the T = 0 limit of the Lehmann sum that cmpy accumulates in
// _accumulate_sum (cmpy/exactdiag.py:110-134) for c = c_{pos,sigma}.
//
// Also here: ed_fourier_t2z, the real-time to complex-frequency transform of
// fourier_t2z (cmpy/exactdiag.py:311-316).
#include "ed_internal.h"

#include <algorithm>
#include <vector>

namespace ed {

static int64_t align256_b(int64_t b) { return (b + 255) & ~int64_t(255); }

// Checks that op (if present) holds sector (n_up + d, n_dn) (sigma up) or
// (n_up, n_dn + d) (sigma dn) of num_sites sites.
static int check_ladder_op(const ed_op* op, int num_sites, int n_up, int n_dn, int sigma, int d, const char* which) {
  if (!op) return ED_OK;
  const int nu = n_up + (sigma == ED_UP ? d : 0), nd = n_dn + (sigma == ED_DN ? d : 0);
  ED_REQUIRE(op->num_sites == num_sites && op->n_up == nu && op->n_dn == nd,
             "%s holds sector (%d, %d) of %d sites; expected (%d, %d) of %d", which, op->n_up, op->n_dn,
             op->num_sites, nu, nd, num_sites);
  return ED_OK;
}

static int check_common(const ed_op* op_p, const ed_op* op_m, int num_sites, int n_up, int n_dn, int pos, int sigma,
                        int sign_mode) {
  ED_REQUIRE(num_sites >= 1 && num_sites <= 32, "num_sites=%d outside [1, 32]", num_sites);
  ED_REQUIRE(n_up >= 0 && n_up <= num_sites && n_dn >= 0 && n_dn <= num_sites, "sector (%d, %d) invalid", n_up, n_dn);
  ED_REQUIRE(pos >= 0 && pos < num_sites, "pos=%d outside [0, %d)", pos, num_sites);
  ED_REQUIRE(sigma == ED_UP || sigma == ED_DN, "sigma must be ED_UP or ED_DN");
  ED_REQUIRE(sign_mode == ED_SIGN_REFERENCE || sign_mode == ED_SIGN_JORDAN_WIGNER, "sign_mode invalid");
  ED_TRY(check_ladder_op(op_p, num_sites, n_up, n_dn, sigma, +1, "op_p"));
  ED_TRY(check_ladder_op(op_m, num_sites, n_up, n_dn, sigma, -1, "op_m"));
  return ED_OK;
}

// weights[k] = ov[k]^2, poles[k] = sign (w[k] - e0)
__global__ void k_lehmann_terms(const double* __restrict__ ov, const double* __restrict__ w, int64_t n, double e0,
                                double sign, double* __restrict__ weights, double* __restrict__ poles) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    weights[k] = ov[k] * ov[k];
    poles[k] = sign * (w[k] - e0);
  }
}

struct c64 {
  double re, im;
};
__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 cdivc(c64 a, c64 b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ c64 cexp_(c64 x) {
  double sn, cs;
  sincos(x.im, &sn, &cs);
  const double e = exp(x.re);
  return {e * cs, e * sn};
}

// G(z_j) = int_0^T e^{i z_j t} G(t) dt with G linear between the samples
// t_0 < ... < t_{nt-1}: per interval of length h, with a = i z and x = a h,
//   e^{a t_k} h [g_k E1(x) + (g_{k+1} - g_k) E2(x)],
//   E1 = (e^x - 1)/x = sum_n x^n/(n+1)!,  E2 = (x e^x - e^x + 1)/x^2 = sum_n (n+1) x^n/(n+2)!,
// the series (12 terms) for |x| < 0.1.  One thread per frequency, the
// intervals summed in order.
__global__ __launch_bounds__(256) void k_fourier_t2z(const double* __restrict__ t, const double* __restrict__ g,
                                                     int64_t nt, const double* __restrict__ z, int64_t nz,
                                                     double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nz) return;
  const c64 a = {-z[2 * j + 1], z[2 * j]};  // i z
  c64 acc = {0.0, 0.0};
  for (int64_t k = 0; k + 1 < nt; ++k) {
    const double h = t[k + 1] - t[k];
    const c64 x = {a.re * h, a.im * h};
    c64 e1, e2;
    if (x.re * x.re + x.im * x.im < 0.01) {
      e1 = {0.0, 0.0};
      e2 = {0.0, 0.0};
      double f1 = 1.0, f2 = 1.0;  // 1/(n+1)!, 1/(n+2)! built downward from n = 11
      for (int n = 1; n <= 13; ++n) f2 /= (double)n;  // 1/13!
      f1 = f2 * 13.0;                                  // 1/12!
      for (int n = 11; n >= 0; --n) {
        e1 = cmul(e1, x);
        e1.re += f1;
        e2 = cmul(e2, x);
        e2.re += (double)(n + 1) * f2;
        f2 = f1;              // 1/(n+1)! becomes the next 1/(n+2)!
        f1 *= (double)(n + 1);  // 1/n!
      }
    } else {
      const c64 ex = cexp_(x);
      const c64 em1 = {ex.re - 1.0, ex.im};
      e1 = cdivc(em1, x);
      const c64 xe = cmul(x, ex);
      e2 = cdivc({xe.re - em1.re, xe.im - em1.im}, cmul(x, x));
    }
    const c64 ph = cexp_({a.re * t[k], a.im * t[k]});
    const c64 gk = {g[2 * k], g[2 * k + 1]};
    const c64 dg = {g[2 * k + 2] - gk.re, g[2 * k + 3] - gk.im};
    const c64 p1 = cmul(gk, e1), p2 = cmul(dg, e2);
    const c64 term = cmul(ph, {h * (p1.re + p2.re), h * (p1.im + p2.im)});
    acc.re += term.re;
    acc.im += term.im;
  }
  out[2 * j] = acc.re;
  out[2 * j + 1] = acc.im;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_fourier_t2z(const double* t_dev, const double* g_dev, int64_t nt, const double* z_dev, int64_t nz,
                              double* out_dev, void* stream) {
  clear_error();
  ED_REQUIRE(nt >= 0 && nz >= 0, "negative size");
  ED_REQUIRE(nz == 0 || (z_dev && out_dev), "NULL buffer");
  ED_REQUIRE(nt == 0 || (t_dev && g_dev), "NULL buffer");
  if (nz == 0) return ED_OK;
  hipLaunchKernelGGL(k_fourier_t2z, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, as_stream(stream), t_dev,
                     g_dev, nt, z_dev, nz, out_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

extern "C" int ed_greens_cf_workspace(const ed_op* op_p, const ed_op* op_m, int32_t nsteps, int32_t reorth,
                                      int64_t* bytes) {
  clear_error();
  ED_REQUIRE(bytes, "NULL argument");
  int64_t dmax = 0, lw = 0;
  for (const ed_op* op : {op_p, op_m}) {
    if (!op || op->size == 0) continue;
    int64_t b = 0;
    ED_TRY(ed_lanczos_coeffs_workspace(op, std::max<int32_t>(1, (int32_t)std::min<int64_t>(nsteps, op->size)),
                                       reorth, &b));
    lw = std::max(lw, b);
    dmax = std::max(dmax, op->size);
  }
  *bytes = align256_b(8 * dmax) + lw;
  return ED_OK;
}

extern "C" int ed_greens_cf(ed_op* op_p, ed_op* op_m, int32_t num_sites, int32_t n_up, int32_t n_dn, int32_t pos,
                            int32_t sigma, int32_t sign_mode, const double* psi_dev, double e0, int32_t nsteps,
                            int32_t reorth, const double* z_dev, int64_t nz, double* g_dev, ed_cf_parts* parts,
                            void* work_dev, int64_t work_bytes, void* stream) {
  clear_error();
  ED_TRY(check_common(op_p, op_m, num_sites, n_up, n_dn, pos, sigma, sign_mode));
  ED_REQUIRE(psi_dev && (nz == 0 || (z_dev && g_dev)), "NULL buffer");
  ED_REQUIRE(nz >= 0, "nz=%lld invalid", (long long)nz);
  ED_REQUIRE(nsteps >= 1 && nsteps <= 100000, "nsteps=%d outside [1, 100000]", nsteps);
  int64_t need = 0;
  ED_TRY(ed_greens_cf_workspace(op_p, op_m, nsteps, reorth, &need));
  ED_REQUIRE(work_dev && work_bytes >= need, "workspace too small: %lld < %lld bytes", (long long)work_bytes,
             (long long)need);
  hipStream_t s = as_stream(stream);
  if (parts) {
    parts->norm2[0] = parts->norm2[1] = 0.0;
    parts->steps[0] = parts->steps[1] = 0;
  }
  if (nz > 0) ED_HIP(hipMemsetAsync(g_dev, 0, sizeof(double) * 2 * (size_t)nz, s));
  char* base = static_cast<char*>(work_dev);
  double* phi = reinterpret_cast<double*>(base);
  int64_t dmax = 0;
  for (const ed_op* op : {op_p, op_m})
    if (op) dmax = std::max(dmax, op->size);
  void* lws = base + align256_b(8 * dmax);
  const int64_t lbytes = work_bytes - align256_b(8 * dmax);
  for (int part = 0; part < 2; ++part) {
    ed_op* op = part == 0 ? op_p : op_m;
    if (!op || op->size == 0) continue;
    // phi = c+ psi0 (particle, sector N+1) or c psi0 (hole, sector N-1)
    if (part == 0)
      ED_TRY(ed_apply_cdag(num_sites, n_up, n_dn, pos, sigma, sign_mode, psi_dev, phi, 1, stream));
    else
      ED_TRY(ed_apply_c(num_sites, n_up, n_dn, pos, sigma, sign_mode, psi_dev, phi, 1, stream));
    const int32_t ns = (int32_t)std::min<int64_t>(nsteps, op->size);
    std::vector<double> a(ns), b(std::max(ns, 1));
    double n2 = 0.0;
    int32_t m = 0;
    ED_TRY(ed_lanczos_coeffs(op, phi, ns, reorth, a.data(), b.data(), &n2, &m, nullptr, lws, lbytes, stream));
    if (parts) {
      parts->norm2[part] = n2;
      parts->steps[part] = m;
      if (parts->alpha[part]) std::copy(a.begin(), a.begin() + m, parts->alpha[part]);
      if (parts->beta[part] && m > 1) std::copy(b.begin(), b.begin() + (m - 1), parts->beta[part]);
    }
    if (m > 0 && n2 > 0.0 && nz > 0)
      ED_TRY(ed_cf_eval(a.data(), b.data(), m, n2, e0, part == 0 ? 1 : -1, z_dev, nz, g_dev, 1, stream));
  }
  ED_HIP(hipGetLastError());
  return ED_OK;
}

extern "C" int ed_greens_lehmann_t0_workspace(const ed_op* op_p, const ed_op* op_m, int64_t* bytes) {
  clear_error();
  ED_REQUIRE(bytes, "NULL argument");
  int64_t dmax = 0;
  for (const ed_op* op : {op_p, op_m})
    if (op) dmax = std::max(dmax, op->size);
  ED_REQUIRE(dmax <= 65536, "sector of %lld states too large for the dense Lehmann sum", (long long)dmax);
  // dense H / eigenvectors, eigenvalues, phi, overlaps, weights, poles
  *bytes = align256_b(8 * dmax * dmax) + 5 * align256_b(8 * dmax);
  return ED_OK;
}

extern "C" int ed_greens_lehmann_t0(ed_op* op_p, ed_op* op_m, int32_t num_sites, int32_t n_up, int32_t n_dn,
                                    int32_t pos, int32_t sigma, int32_t sign_mode, const double* psi_dev, double e0,
                                    const double* z_dev, int64_t nz, double* g_dev, void* work_dev,
                                    int64_t work_bytes, void* stream) {
  clear_error();
  ED_TRY(check_common(op_p, op_m, num_sites, n_up, n_dn, pos, sigma, sign_mode));
  ED_REQUIRE(psi_dev && (nz == 0 || (z_dev && g_dev)), "NULL buffer");
  ED_REQUIRE(nz >= 0, "nz=%lld invalid", (long long)nz);
  int64_t need = 0;
  ED_TRY(ed_greens_lehmann_t0_workspace(op_p, op_m, &need));
  ED_REQUIRE(work_dev && work_bytes >= need, "workspace too small: %lld < %lld bytes", (long long)work_bytes,
             (long long)need);
  hipStream_t s = as_stream(stream);
  if (nz > 0) ED_HIP(hipMemsetAsync(g_dev, 0, sizeof(double) * 2 * (size_t)nz, s));
  int64_t dmax = 0;
  for (const ed_op* op : {op_p, op_m})
    if (op) dmax = std::max(dmax, op->size);
  char* base = static_cast<char*>(work_dev);
  double* h = reinterpret_cast<double*>(base);
  char* p = base + align256_b(8 * dmax * dmax);
  double* w = reinterpret_cast<double*>(p);
  double* phi = reinterpret_cast<double*>(p + align256_b(8 * dmax));
  double* ov = reinterpret_cast<double*>(p + 2 * align256_b(8 * dmax));
  double* wt = reinterpret_cast<double*>(p + 3 * align256_b(8 * dmax));
  double* pl = reinterpret_cast<double*>(p + 4 * align256_b(8 * dmax));
  for (int part = 0; part < 2; ++part) {
    ed_op* op = part == 0 ? op_p : op_m;
    if (!op || op->size == 0) continue;
    const int64_t D = op->size;
    if (part == 0)
      ED_TRY(ed_apply_cdag(num_sites, n_up, n_dn, pos, sigma, sign_mode, psi_dev, phi, 1, stream));
    else
      ED_TRY(ed_apply_c(num_sites, n_up, n_dn, pos, sigma, sign_mode, psi_dev, phi, 1, stream));
    // eigenpairs of the sector (vectors stored one after another), overlaps
    // <m|phi>, then the poles E_m - E0 (particle) or E0 - E_n (hole)
    ED_TRY(ed_op_dense(op, h, stream));
    ED_TRY(ed_dense_eigh(h, D, w, stream));
    ED_TRY(ed_gemm_nt(h, phi, ov, D, 1, D, stream));
    hipLaunchKernelGGL(k_lehmann_terms, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, s, ov, w, D, e0,
                       part == 0 ? 1.0 : -1.0, wt, pl);
    ED_HIP(hipGetLastError());
    if (nz > 0) ED_TRY(ed_pole_sum(wt, pl, D, z_dev, nz, g_dev, 1, stream));
  }
  return ED_OK;
}
