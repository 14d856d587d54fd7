// Sector H.v sweep (replaces HamiltonOperator._matvec, cmpy/operators.py:626-630,
// which loops over the COO triples emitted by project_* in cmpy/operators.py).
//
// Layout: x and y are X[num_up][num_dn] row-major (dn fastest).  One workgroup
// owns a tile of RB consecutive rows x 256 consecutive columns; lane = column.
//   diagonal   (d_up[iu] + d_dn[id] + U(up & dn)) x[iu, id]      fused
//   dn hops    sum_k val_k x[iu, col_k(id)]                        per-lane ELL list,
//                                                                   reused for RB rows
//   up hops    sum_k val_k x[col_k(iu), id]                        wave-uniform list,
//                                                                   coalesced row reads
// The Lanczos epilogue y = s*Hx - c*xprev and the block partial of <y, x> are
// fused into the same pass (no extra traffic: x is already in registers).
#include "ed_internal.h"

namespace ed {

__device__ __forceinline__ double u_term_dev(const HvArgs& a, uint32_t both) {
  if (a.u_mode == 1) return a.u_uniform * (double)popc32(both);
  if (a.u_mode == 2) {
    double acc = 0.0;
    uint32_t m = both & a.u_mask;
    while (m) {
      const int p = __builtin_ctz(m);
      acc += a.u_site[p];
      m &= m - 1;
    }
    return acc;
  }
  return 0.0;
}

__device__ __forceinline__ double block_sum_256(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (red[0] + red[1]) + (red[2] + red[3]);
  return r;
}

template <int RB, bool LZ>
__global__ __launch_bounds__(256) void k_hv(HvArgs a, const double* __restrict__ x,
                                            double* __restrict__ y, HvEpilogue e, int order,
                                            int64_t nrb, int64_t nct) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  int64_t rb, ct;
  if (order == 0) {
    rb = b % nrb;
    ct = b / nrb;
  } else {
    ct = b % nct;
    rb = b / nct;
  }
  const int64_t nd = a.num_dn;
  const int64_t id = ct * 256 + threadIdx.x;
  const int64_t r0 = rb * RB;
  const bool act = id < nd;
  double dot = 0.0;
  if (act) {
    double acc[RB], xo[RB];
    const uint32_t sdn = a.st_dn[id];
    const double ddn = a.d_dn[id];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int64_t iu = r0 + r;
      if (iu < a.num_up) {
        xo[r] = x[iu * nd + id];
        acc[r] = (a.d_up[iu] + ddn + u_term_dev(a, a.st_up[iu] & sdn)) * xo[r];
      } else {
        xo[r] = 0.0;
        acc[r] = 0.0;
      }
    }
    // dn hops: per-lane list, loaded once for the RB rows
    const int cd = a.dn_cnt[id];
    for (int k = 0; k < cd; ++k) {
      const int64_t c = a.dn_col[(int64_t)k * nd + id];
      const double v = a.dn_val[(int64_t)k * nd + id];
#pragma unroll
      for (int r = 0; r < RB; ++r)
        if (r0 + r < a.num_up) acc[r] += v * x[(r0 + r) * nd + c];
    }
    // up hops: uniform list per row, coalesced source rows
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int64_t iu = r0 + r;
      if (iu < a.num_up) {
        const int cu = a.up_cnt[iu];
        const int32_t* cp = a.up_col + iu * a.up_w;
        const double* vp = a.up_val + iu * a.up_w;
        for (int k = 0; k < cu; ++k) acc[r] += vp[k] * x[(int64_t)cp[k] * nd + id];
      }
    }
    if (LZ) {
      const double sx = e.sx ? *e.sx : 1.0;
      const double cp = e.cprev ? *e.cprev : 0.0;
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int64_t iu = r0 + r;
        if (iu < a.num_up) {
          const int64_t k = iu * nd + id;
          double v = sx * acc[r];
          if (e.xprev) v -= cp * e.xprev[k];
          y[k] = v;
          dot += v * xo[r];
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int64_t iu = r0 + r;
        if (iu < a.num_up) y[iu * nd + id] = acc[r];
      }
    }
  }
  if (LZ && e.partial) {
    const double s = block_sum_256(dot, red);
    if (threadIdx.x == 0) e.partial[b] = s;
  }
}

static int pick_rb(const ed_op* op) {
  if (op->num_up >= 4096) return 4;
  if (op->num_up >= 64) return 2;
  return 1;
}

static void grid_shape(const ed_op* op, int rb, int64_t* nrb, int64_t* nct) {
  *nrb = (op->num_up + rb - 1) / rb;
  *nct = (op->num_dn + 255) / 256;
}

int matvec_num_partials(const ed_op* op) {
  int64_t nrb, nct;
  grid_shape(op, pick_rb(op), &nrb, &nct);
  return (int)(nrb * nct);
}

template <bool LZ>
static void launch_rb(int rb, const HvArgs& a, const double* x, double* y, const HvEpilogue& e,
                      int64_t nrb, int64_t nct, hipStream_t s) {
  const dim3 grid((unsigned)(nrb * nct)), block(256);
  const int order = 0;
  switch (rb) {
    case 1: hipLaunchKernelGGL((k_hv<1, LZ>), grid, block, 0, s, a, x, y, e, order, nrb, nct); break;
    case 2: hipLaunchKernelGGL((k_hv<2, LZ>), grid, block, 0, s, a, x, y, e, order, nrb, nct); break;
    case 4: hipLaunchKernelGGL((k_hv<4, LZ>), grid, block, 0, s, a, x, y, e, order, nrb, nct); break;
    default: hipLaunchKernelGGL((k_hv<8, LZ>), grid, block, 0, s, a, x, y, e, order, nrb, nct); break;
  }
}

int launch_matvec(const ed_op* op, const double* x, double* y, const HvEpilogue& epi, hipStream_t s,
                  int* num_partials) {
  const HvArgs a = make_hv_args(op);
  const int rb = pick_rb(op);
  int64_t nrb, nct;
  grid_shape(op, rb, &nrb, &nct);
  ED_REQUIRE(nrb * nct < (1ll << 31), "grid too large");
  const bool lz = epi.xprev || epi.sx || epi.cprev || epi.partial;
  if (lz)
    launch_rb<true>(rb, a, x, y, epi, nrb, nct, s);
  else
    launch_rb<false>(rb, a, x, y, epi, nrb, nct, s);
  if (num_partials) *num_partials = (int)(nrb * nct);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

}  // namespace ed

namespace ed {
// y[t] = sum_{k in ptr[t]..ptr[t+1]} val[k] * x[src[k]], summed in stored order.
__global__ __launch_bounds__(256) void k_csr(const int64_t* __restrict__ ptr, const int64_t* __restrict__ src,
                                             const double* __restrict__ val, int64_t n,
                                             const double* __restrict__ x, double* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double acc = 0.0;
  for (int64_t k = ptr[t]; k < ptr[t + 1]; ++k) acc += val[k] * x[src[k]];
  y[t] = acc;
}
}  // namespace ed

using namespace ed;

extern "C" int ed_csr_matvec(const int64_t* ptr_dev, const int64_t* src_dev, const double* val_dev, int64_t n,
                             const double* x_dev, double* y_dev, void* stream) {
  clear_error();
  ED_REQUIRE(n >= 0, "n < 0");
  if (n == 0) return ED_OK;
  ED_REQUIRE(ptr_dev && src_dev && val_dev && x_dev && y_dev, "NULL buffer");
  hipLaunchKernelGGL(k_csr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), ptr_dev, src_dev,
                     val_dev, n, x_dev, y_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

extern "C" int ed_op_matvec(ed_op* op, const double* x_dev, double* y_dev, int64_t nvec, void* stream) {
  clear_error();
  ED_REQUIRE(op && x_dev && y_dev, "NULL argument");
  ED_REQUIRE(nvec >= 0, "nvec < 0");
  const int64_t D = op->size;
  const char* xb = (const char*)x_dev;
  const char* yb = (const char*)y_dev;
  ED_REQUIRE(xb + sizeof(double) * D * nvec <= yb || yb + sizeof(double) * D * nvec <= xb,
             "x and y overlap");
  HvEpilogue none{nullptr, nullptr, nullptr, nullptr};
  for (int64_t v = 0; v < nvec; ++v)
    ED_TRY(launch_matvec(op, x_dev + v * D, y_dev + v * D, none, as_stream(stream), nullptr));
  return ED_OK;
}
