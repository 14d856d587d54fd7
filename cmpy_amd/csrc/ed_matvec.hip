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

#include <cstdlib>
#include <mutex>

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
  const int64_t nd = a.num_dn, ld = a.ld;
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
        xo[r] = x[iu * ld + id];
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
        if (r0 + r < a.num_up) acc[r] += v * x[(r0 + r) * ld + c];
    }
    // up hops: uniform list per row, coalesced source rows
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int64_t iu = r0 + r;
      if (iu < a.num_up) {
        const int cu = a.up_cnt[iu];
        const int32_t* cp = a.up_col + iu * a.up_w;
        const double* vp = a.up_val + iu * a.up_w;
        for (int k = 0; k < cu; ++k) acc[r] += vp[k] * x[(int64_t)cp[k] * ld + id];
      }
    }
    if (LZ) {
      const double sx = e.sx ? *e.sx : 1.0;
      const double cp = e.cprev ? *e.cprev : 0.0;
      const double cs = e.cself ? *e.cself : 0.0;
      const double ps = e.psi_y * sx;
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int64_t iu = r0 + r;
        if (iu < a.num_up) {
          const int64_t k = iu * ld + id;
          double v = sx * (acc[r] - cs * xo[r]);
          if (e.xprev) v -= cp * e.xprev[k];
          y[k] = v;
          if (e.psi) e.psi[k] += ps * xo[r];
          dot += v * xo[r];
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int64_t iu = r0 + r;
        if (iu < a.num_up) y[iu * ld + id] = acc[r];
      }
    }
  } else if (id < ld) {
    // row padding: zeros
#pragma unroll
    for (int r = 0; r < RB; ++r)
      if (r0 + r < a.num_up) y[(r0 + r) * ld + id] = 0.0;
  }
  if (LZ && e.partial) {
    const double s = block_sum_256(dot, red);
    if (threadIdx.x == 0) e.partial[b] = s;
  }
}

// ---------------------------------------------------------------------------
// Two-pass sweep for large sectors.
//
// Pass C (column slabs): y_up = s*(up hops).  Column tiles of TW columns are
// assigned to XCDs round-robin and each XCD sweeps every row of a tile before
// moving on, so the tile's slab of x (num_up x TW doubles) stays in that
// XCD's L2 while ~17 source rows per output row are read from it.
// Pass R (rows): a workgroup stages R whole rows of x in LDS and writes
//   y = s*(diag + dn hops) + y_up - c*xprev   (+ block partial of <y, x>)
// so every dn-hop gather is an LDS read and y_up is read back as whole
// contiguous rows.  With the uniform-hopping tables pass C runs first and
// writes y; otherwise pass R runs first and pass C adds to y.
// ---------------------------------------------------------------------------
struct Hv2Args {
  int64_t nu, nd;
  int64_t ld;  // row stride of x and y in doubles (>= nd)
  const uint32_t* st_up;
  const uint32_t* st_dn;
  const double* d_up;
  const double* d_dn;
  const int32_t* up_cnt;
  const uint32_t* up_pk;  // [nu][up_w]
  int32_t up_w;
  const int32_t* dn_cnt;
  const uint32_t* dn_pk;  // [dn_w][nd]
  int32_t dn_w;
  // sign-split tables (uniform hopping; see Tables)
  const uint32_t* dn_off_pos;   // [wpos][nd]
  const uint32_t* dn_off_neg;   // [wneg][nd]
  const int32_t* dn_cmax_pos;   // [ceil(nd / 64)]
  const int32_t* dn_cmax_neg;
  const uint32_t* dn_h16;       // [hwords][nd] half-width index pairs
  const uint32_t* dn_h16q;      // the same, lane-contiguous groups of four words (see Tables::h16q)
  int32_t dn_hgroups;
  const int32_t* dn_hwords;     // [2 * ceil(nd / 64)] positive and total words per chunk
  const int32_t* up_tgt;        // [nu][up_w]
  const uint32_t* up_cnt_pn;    // [nu]
  const uint32_t* up_pair;      // [nu][up_pair_stride] pair lists (see Tables::pair16)
  int32_t up_pair_stride;
  const uint32_t* up_quad;      // [nu][up_quad_stride] quad lists (see Tables::quad16)
  int32_t up_quad_stride;
  const double* hop;
  int32_t nb;
  int32_t u_mode;
  double u_uniform;
  uint32_t u_mask;
  const double* u_site;
};

__device__ __forceinline__ double u_term2(const Hv2Args& a, uint32_t both) {
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

template <int NT, typename P>
__device__ __forceinline__ double block_sum_n(double v, P red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; ++w) r += red[w];
  return r;
}

// Zeros into the row padding (entries nd .. ld-1) of rows r0 .. r0+R-1 of y.
__device__ __forceinline__ void zero_row_pads(double* y, int64_t r0, int R, int64_t nd, int64_t ld) {
  const int64_t np = ld - nd;
  if (np <= 0) return;
  for (int64_t k = threadIdx.x; k < (int64_t)R * np; k += blockDim.x) {
    const int64_t r = k / np;
    y[(r0 + r) * ld + nd + (k - r * np)] = 0.0;
  }
}

// Load from an address that is the same for every lane of the wave, through
// the scalar cache (the tables it reads are never written by these kernels).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

constexpr int kRowThreads = 1024;

// Hop value of a packed entry: uniform models (every bond the same hop t) need
// only the sign bit; otherwise the bond's value comes from the LDS copy.
template <bool UNI>
__device__ __forceinline__ double hv_val(uint32_t ent, double t, const double* hv) {
  if (UNI) return (ent & 0x80000000u) ? -t : t;
  return hop_value(ent, hv);
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Copies n doubles into LDS with every thread holding U loads in flight
// (one HBM round trip per NT*U elements instead of one per element).
template <int NT, int U, typename P>
__device__ __forceinline__ void stage_rows(P lds, const double* __restrict__ src, int64_t n) {
  for (int64_t base = 0; base < n; base += (int64_t)NT * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + (int64_t)u * NT + threadIdx.x;
      v[u] = __builtin_nontemporal_load(src + (k < n ? k : 0));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + (int64_t)u * NT + threadIdx.x;
      if (k < n) lds[k] = v[u];
    }
  }
}

constexpr int kHopBatch = 16;  // hop terms whose loads are issued together (column pass)
constexpr int kRowBatch = 8;   // same for the row pass

// Row pass, general tables.  Hops run in batches of kRowBatch: entries first,
// then independent LDS gathers, then FMAs; slots past a column's hop count
// read index 0 with value 0, so no branch separates the loads of a batch.
template <bool UNI>
__global__ __launch_bounds__(kRowThreads) void k_hv_rows(Hv2Args a, int rows_per_block,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y, HvEpilogue e) {
  extern __shared__ double lds[];
  __shared__ double red[kRowThreads / 64];
  const int64_t nd = a.nd, ld = a.ld;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int R = (int)(a.nu - r0 < rows_per_block ? a.nu - r0 : rows_per_block);
  double* hv = lds + (int64_t)rows_per_block * nd;
  if (ld == nd)
    stage_rows<kRowThreads, 8>(lds, x + r0 * nd, (int64_t)R * nd);
  else
    for (int r = 0; r < R; ++r) stage_rows<kRowThreads, 8>(lds + (int64_t)r * nd, x + (r0 + r) * ld, nd);
  for (int b = threadIdx.x; b < a.nb; b += kRowThreads) hv[b] = a.hop[b];
  zero_row_pads(y, r0, R, nd, ld);
  __syncthreads();
  const double sx = e.sx ? *e.sx : 1.0;
  const double cp = e.cprev ? *e.cprev : 0.0;
  const double cs = e.cself ? *e.cself : 0.0;
  const double ps = e.psi_y * sx;
  const double t = a.nb > 0 ? hv[0] : 0.0;
  double dot = 0.0;
  const int64_t nd_pad = (nd + 63) & ~int64_t(63);
  for (int64_t id = threadIdx.x; id < nd_pad; id += kRowThreads) {
    const bool act = id < nd;
    const int64_t ids = act ? id : 0;
    const int cnt0 = a.dn_cnt[ids];
    const uint32_t sdn = a.st_dn[ids];
    const double ddn = a.d_dn[ids];
    const int cnt = act ? cnt0 : 0;
    const int cmax = wave_max_int(cnt);
    for (int r = 0; r < R; ++r) {
      const int64_t iu = r0 + r;
      const int64_t gi = iu * ld + id;
      const double xp = (e.xprev && act) ? e.xprev[gi] : 0.0;
      const double* row = lds + (int64_t)r * nd;
      const double xo = act ? row[ids] : 0.0;
      double acc = (a.d_up[iu] + ddn + u_term2(a, a.st_up[iu] & sdn)) * xo;
      for (int k0 = 0; k0 < cmax; k0 += kRowBatch) {
        uint32_t tg[kRowBatch];
        double vl[kRowBatch];
        {
          uint32_t en[kRowBatch];
#pragma unroll
          for (int q = 0; q < kRowBatch; ++q) {
            const int kk = k0 + q;
            en[q] = a.dn_pk[(kk < cnt) ? (int64_t)kk * nd + ids : 0];
          }
#pragma unroll
          for (int q = 0; q < kRowBatch; ++q) {
            const bool ok = k0 + q < cnt;
            tg[q] = ok ? (uint32_t)hop_target(en[q]) : 0u;
            vl[q] = ok ? hv_val<UNI>(en[q], t, hv) : 0.0;
          }
        }
        double g[kRowBatch];
#pragma unroll
        for (int q = 0; q < kRowBatch; ++q) g[q] = row[tg[q]];
#pragma unroll
        for (int q = 0; q < kRowBatch; ++q) acc += vl[q] * g[q];
      }
      if (act) {
        const double v = sx * (acc - cs * xo) - cp * xp;
        __builtin_nontemporal_store(v, y + gi);
        if (e.psi) e.psi[gi] += ps * xo;
        dot += v * xo;
      }
    }
  }
  if (e.partial) {
    const double s = block_sum_n<kRowThreads>(dot, red);
    if (threadIdx.x == 0) e.partial[blockIdx.x] = s;
  }
}

// Sums one sign list of the row pass over RB staged rows: 8 table words per
// batch (coalesced over the wave's 64 columns), reused for every row of the
// group; padding slots address the row's zero pad, so the loop runs the
// chunk's slot count, rounded up to 8, with no per-lane test.
template <int RB>
__device__ __forceinline__ void gather_list(const uint32_t* __restrict__ tab, int kmax, int64_t nd, int64_t ids,
                                            int rg, int R, uint32_t row_bytes, const double* lds, double* acc) {
  const char* base = (const char*)lds;
  for (int k0 = 0; k0 < kmax; k0 += 8) {
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) w[q] = tab[(int64_t)(k0 + q) * nd + ids];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (rg + r < R) {
        const uint32_t rb = (uint32_t)(rg + r) * row_bytes;
        double g[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) g[q] = *(const double*)(base + (rb + w[q]));
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[r] += g[q];
      }
    }
  }
}

// Row pass for uniform hopping t: R rows of x staged in LDS, each followed by
// a zero pad; y = s*(diag*x + t*(P - N)) - c*xprev with P (N) the sum of the
// dn-hop sources of sign +1 (-1), gathered through the sign-split offset
// tables.  The loop over a 64-column chunk runs that chunk's largest count.
template <int RB, bool ADDUP>
__global__ __launch_bounds__(kRowThreads) void k_hv_rows_off(Hv2Args a, int rows_per_block,
                                                             const double* __restrict__ x,
                                                             double* __restrict__ y, HvEpilogue e) {
  extern __shared__ double lds[];
  __shared__ double red[kRowThreads / 64];
  const int64_t nd = a.nd, ld = a.ld;
  const uint32_t stride = (uint32_t)nd + 1u;  // doubles per staged row
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int R = (int)(a.nu - r0 < rows_per_block ? a.nu - r0 : rows_per_block);
  for (int r = 0; r < R; ++r) stage_rows<kRowThreads, 8>(lds + (int64_t)r * stride, x + (r0 + r) * ld, nd);
  if ((int)threadIdx.x < R) lds[(int64_t)threadIdx.x * stride + nd] = 0.0;
  zero_row_pads(y, r0, R, nd, ld);
  __syncthreads();
  const double sx = e.sx ? *e.sx : 1.0;
  const double cp = e.cprev ? *e.cprev : 0.0;
  const double cs = e.cself ? *e.cself : 0.0;
  const double ps = e.psi_y * sx;
  const double t = a.hop[0];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double dot = 0.0;
  for (int64_t c0 = (int64_t)wave * 64; c0 < nd; c0 += kRowThreads) {
    const int64_t id = c0 + lane;
    const bool act = id < nd;
    const int64_t ids = act ? id : c0;
    const int64_t chunk = c0 >> 6;
    const int kp = (uniform_load(a.dn_cmax_pos + chunk) + 7) & ~7;
    const int kn = (uniform_load(a.dn_cmax_neg + chunk) + 7) & ~7;
    const uint32_t sdn = a.st_dn[ids];
    const double ddn = a.d_dn[ids];
    for (int rg = 0; rg < R; rg += RB) {
      double xp[RB], yu[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        xp[r] = (e.xprev && act && rg + r < R) ? e.xprev[(r0 + rg + r) * ld + id] : 0.0;
        yu[r] = (ADDUP && act && rg + r < R) ? __builtin_nontemporal_load(y + (r0 + rg + r) * ld + id) : 0.0;
      }
      double accp[RB], accn[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) accp[r] = accn[r] = 0.0;
      gather_list<RB>(a.dn_off_pos, kp, nd, ids, rg, R, stride * 8u, lds, accp);
      gather_list<RB>(a.dn_off_neg, kn, nd, ids, rg, R, stride * 8u, lds, accn);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        if (act && rg + r < R) {
          const int64_t iu = r0 + rg + r;
          const double xo = lds[(int64_t)(rg + r) * stride + id];
          const double dg = a.d_up[iu] + ddn + u_term2(a, a.st_up[iu] & sdn) - cs;
          const double v = sx * (dg * xo + t * (accp[r] - accn[r])) + yu[r] - cp * xp[r];
          __builtin_nontemporal_store(v, y + iu * ld + id);
          if (e.psi) e.psi[iu * ld + id] += ps * xo;
          dot += v * xo;
        }
      }
    }
  }
  if (e.partial) {
    const double s = block_sum_n<kRowThreads>(dot, red);
    if (threadIdx.x == 0) e.partial[blockIdx.x] = s;
  }
}

// P - N over one chunk's half-width word list of one column: NW words, word
// k at byte tb + k * nd4 + lane4 (uniform base plus the lane's offset), the
// first nwp of sign +1.  Straight-line code: every word is requested before
// the first gather, so a chunk costs one table round trip and one LDS round
// trip.
// Workgroup LDS addressed from offset 0 (for kernels without static LDS, whose
// dynamic LDS therefore starts at 0).
using lds_f64 = __attribute__((address_space(3))) double;

// One word's two hops of sign sg into a0 (low entry) and a1 (high entry).
__device__ __forceinline__ void add_word(uint32_t x, double sg, const lds_f64* row, double& a0, double& a1) {
  a0 = fma(sg, row[x & 0xffffu], a0);
  a1 = fma(sg, row[x >> 16], a1);
}

template <int NW>
__device__ __forceinline__ double sum_words(const char* tb, uint32_t nd4, uint32_t lane4, int nwp,
                                            const lds_f64* row) {
  uint32_t w[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = *(const uint32_t*)(tb + (size_t)k * nd4 + lane4);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int k = 0; k < NW; ++k) add_word(w[k], k < nwp ? 1.0 : -1.0, row, a0, a1);
  return a0 + a1;
}

__device__ __forceinline__ double sum_words_n(int nw, const char* tb, uint32_t nd4, uint32_t lane4, int nwp,
                                              const lds_f64* row) {
  switch (nw) {
#define ED_SUM_CASE(n) \
  case n: return sum_words<n>(tb, nd4, lane4, nwp, row);
    ED_SUM_CASE(1) ED_SUM_CASE(2) ED_SUM_CASE(3) ED_SUM_CASE(4) ED_SUM_CASE(5) ED_SUM_CASE(6)
    ED_SUM_CASE(7) ED_SUM_CASE(8) ED_SUM_CASE(9) ED_SUM_CASE(10) ED_SUM_CASE(11) ED_SUM_CASE(12)
    ED_SUM_CASE(13) ED_SUM_CASE(14) ED_SUM_CASE(15) ED_SUM_CASE(16) ED_SUM_CASE(17) ED_SUM_CASE(18)
    ED_SUM_CASE(19) ED_SUM_CASE(20) ED_SUM_CASE(21) ED_SUM_CASE(22) ED_SUM_CASE(23) ED_SUM_CASE(24)
#undef ED_SUM_CASE
    default: return 0.0;  // nw == 0 (kHalfMaxWords bounds it above)
  }
}

// Row pass for uniform hopping t and num_dn <= kHalfMaxStates, one row per
// block: the row is staged in LDS with 32 zero pads after it (indices nd ..
// nd+31), and the dn hops gather through the bank-ordered half-width table,
// whose slots each read 32 distinct bank pairs per half-wave.
template <bool ADDUP>
__global__ __launch_bounds__(kRowThreads) void k_hv_rows_half(Hv2Args a, int rows_per_block,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ y, HvEpilogue e) {
  // all LDS is dynamic and addressed from offset 0 (a gather address is
  // index * 8): row [nd], pads [32], reduction slots [kRowThreads / 64]
  extern __shared__ double dyn_lds[];
  (void)dyn_lds;
  lds_f64* row = (lds_f64*)(uintptr_t)0;
  const int64_t nd = a.nd, ld = a.ld;
  lds_f64* red = row + nd + 32;
  const int64_t iu = blockIdx.x;
  stage_rows<kRowThreads, 16>(row, x + iu * ld, nd);
  if (threadIdx.x < 32) row[nd + threadIdx.x] = 0.0;
  zero_row_pads(y, iu, 1, nd, ld);
  __syncthreads();
  const double sx = e.sx ? *e.sx : 1.0;
  const double cp = e.cprev ? *e.cprev : 0.0;
  const double cs = e.cself ? *e.cself : 0.0;
  const double ps = e.psi_y * sx;
  const double t = a.hop[0];
  const double dup = a.d_up[iu] - cs;
  const uint32_t sup = a.st_up[iu];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double dot = 0.0;
  for (int64_t c0 = (int64_t)wave * 64; c0 < nd; c0 += kRowThreads) {
    const int64_t id = c0 + lane;
    const bool act = id < nd;
    const int64_t ids = act ? id : c0;
    const int64_t chunk = c0 >> 6;
    const int nwp = uniform_load(a.dn_hwords + 2 * chunk);
    const int nw = uniform_load(a.dn_hwords + 2 * chunk + 1);
    const int64_t gi = iu * ld + ids;
    // psi read before the gathers, so its HBM latency overlaps them
    const double pso = e.psi && act ? __builtin_nontemporal_load(e.psi + gi) : 0.0;
    const double xp = e.xprev ? e.xprev[gi] : 0.0;
    const double yu = ADDUP ? __builtin_nontemporal_load(y + gi) : 0.0;
    const uint32_t sdn = a.st_dn[ids];
    const double ddn = a.d_dn[ids];
    const double pn = sum_words_n(nw, (const char*)a.dn_h16, (uint32_t)nd * 4u, (uint32_t)ids * 4u, nwp, row);
    if (act) {
      const double xo = row[id];
      const double dg = dup + ddn + u_term2(a, sup & sdn);
      const double v = sx * (dg * xo + t * pn) + yu - cp * xp;
      __builtin_nontemporal_store(v, y + gi);
      if (e.psi) __builtin_nontemporal_store(pso + ps * xo, e.psi + gi);
      dot += v * xo;
    }
  }
  if (e.partial) {
    const double s = block_sum_n<kRowThreads>(dot, red);
    if (threadIdx.x == 0) e.partial[blockIdx.x] = s;
  }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// P - N over a chunk's words in the lane-contiguous layout: NG 16-byte loads
// (four words each) per lane, all requested before the first gather; words
// 0 .. nwp - 1 have sign +1, the rest -1 (pads gather zeros).
template <int NG>
__device__ __forceinline__ double sum_words_q(const u32x4* tq, int nwp, const lds_f64* row) {
  u32x4 w[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) w[g] = tq[g * 64];
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int j = 0; j < 4; ++j) add_word(w[g][j], 4 * g + j < nwp ? 1.0 : -1.0, row, a0, a1);
  return a0 + a1;
}

__device__ __forceinline__ double sum_words_q_n(int ng, const u32x4* tq, int nwp, const lds_f64* row) {
  switch (ng) {
    case 1: return sum_words_q<1>(tq, nwp, row);
    case 2: return sum_words_q<2>(tq, nwp, row);
    case 3: return sum_words_q<3>(tq, nwp, row);
    case 4: return sum_words_q<4>(tq, nwp, row);
    case 5: return sum_words_q<5>(tq, nwp, row);
    case 6: return sum_words_q<6>(tq, nwp, row);
    default: return 0.0;  // no words (kHalfMaxWords = 24 bounds ng by 6)
  }
}

// k_hv_rows_half with the row's y_up (ADDUP) and x_prev values loaded
// together with the staging of x, held in registers (one per column chunk of
// the thread, indexed by the uniform chunk counter), so no chunk waits for
// HBM: only table words (L2) and LDS gathers remain on each chunk's path.
// For num_dn <= kRowHeld * kRowThreads.  NT: x_prev read with a nontemporal
// hint (vectors of at least kStreamStates states, read once per step).
constexpr int kRowHeld = 16;

template <bool ADDUP, bool NT = false>
__global__ __launch_bounds__(kRowThreads) void k_hv_rows_held(Hv2Args a, int rows_per_block,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ y, HvEpilogue e) {
  extern __shared__ double dyn_lds[];
  (void)dyn_lds;
  lds_f64* row = (lds_f64*)(uintptr_t)0;
  const int64_t nd = a.nd, ld = a.ld;
  lds_f64* red = row + nd + 32;
  const int64_t iu = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const double cp = e.cprev ? *e.cprev : 0.0;
  // held[k] = y_up - c x_prev of the thread's chunk k (both loads issued
  // before the staging loads, so one HBM round trip covers all of them)
  double held[kRowHeld], ph[kRowHeld];
#pragma unroll
  for (int k = 0; k < kRowHeld; ++k) {
    const int64_t c0 = (int64_t)wave * 64 + (int64_t)k * kRowThreads;
    held[k] = 0.0;
    ph[k] = 0.0;
    if (c0 < nd) {
      const int64_t id = c0 + lane;
      const int64_t gi = iu * ld + (id < nd ? id : c0);
      if (ADDUP) held[k] = __builtin_nontemporal_load(y + gi);
      if (e.xprev) ph[k] = NT ? __builtin_nontemporal_load(e.xprev + gi) : e.xprev[gi];
    }
  }
  stage_rows<kRowThreads, 16>(row, x + iu * ld, nd);
  if (threadIdx.x < 32) row[nd + threadIdx.x] = 0.0;
  zero_row_pads(y, iu, 1, nd, ld);
#pragma unroll
  for (int k = 0; k < kRowHeld; ++k) held[k] = fma(-cp, ph[k], held[k]);
  __syncthreads();
  const double sx = e.sx ? *e.sx : 1.0;
  const double ps = e.psi_y * sx;
  const double t = a.hop[0];
  const double dup = a.d_up[iu] - (e.cself ? *e.cself : 0.0);
  const uint32_t sup = a.st_up[iu];
  double dot = 0.0;
  for (int k = 0; k < kRowHeld; ++k) {
    const int64_t c0 = (int64_t)wave * 64 + (int64_t)k * kRowThreads;
    if (c0 >= nd) break;
    const int64_t id = c0 + lane;
    const bool act = id < nd;
    const int64_t ids = act ? id : c0;
    const int64_t chunk = c0 >> 6;
    const int nwp = uniform_load(a.dn_hwords + 2 * chunk);
    const int nw = uniform_load(a.dn_hwords + 2 * chunk + 1);
    const int64_t gi = iu * ld + ids;
    // psi read before the gathers, so its HBM latency overlaps them
    const double pso = e.psi && act ? __builtin_nontemporal_load(e.psi + gi) : 0.0;
    const uint32_t sdn = a.st_dn[ids];
    const double ddn = a.d_dn[ids];
    const double pn =
        a.dn_h16q ? sum_words_q_n((nw + 3) >> 2, (const u32x4*)a.dn_h16q + (chunk * a.dn_hgroups) * 64 + lane, nwp, row)
                  : sum_words_n(nw, (const char*)a.dn_h16, (uint32_t)nd * 4u, (uint32_t)ids * 4u, nwp, row);
    if (act) {
      const double xo = row[id];
      const double dg = dup + ddn + u_term2(a, sup & sdn);
      const double v = sx * (dg * xo + t * pn) + held[k];
      __builtin_nontemporal_store(v, y + gi);
      if (e.psi) __builtin_nontemporal_store(pso + ps * xo, e.psi + gi);
      dot += v * xo;
    }
  }
  if (e.partial) {
    const double s = block_sum_n<kRowThreads>(dot, red);
    if (threadIdx.x == 0) e.partial[blockIdx.x] = s;
  }
}

constexpr int kColMaxRows = 64;

// Column pass, general tables: the block stages its rows' up-hop entries in
// LDS; each row's hops run in batches of independent slab loads.
template <int TW, int MAXH, bool UNI>
__global__ __launch_bounds__(256) void k_hv_cols(Hv2Args a, const double* __restrict__ x,
                                                 double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                 int64_t nrc, int rch, double* partial) {
  constexpr int RPI = 256 / TW;
  __shared__ double hv[128];
  __shared__ double red[4];
  __shared__ uint32_t sent[kColMaxRows * MAXH];
  __shared__ int scnt[kColMaxRows];
  const int64_t b = blockIdx.x;
  const int64_t xcd = b & 7, j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + xcd;
  const int64_t rc = j % nrc;
  const int64_t nd = a.nd, ld = a.ld;
  const int64_t row0 = rc * rch;
  for (int i = threadIdx.x; i < a.nb; i += 256) hv[i] = a.hop[i];
  for (int i = threadIdx.x; i < rch * MAXH; i += 256) {
    const int r = i / MAXH, k = i - r * MAXH;
    const int64_t iu = row0 + r;
    uint32_t v = 0;
    if (iu < a.nu && k < a.up_cnt[iu]) v = a.up_pk[iu * a.up_w + k];
    sent[i] = v;
  }
  for (int r = threadIdx.x; r < rch; r += 256) scnt[r] = row0 + r < a.nu ? a.up_cnt[row0 + r] : 0;
  __syncthreads();
  const double t0 = a.nb > 0 ? hv[0] : 0.0;
  double dot = 0.0;
  if (tile < ntiles) {
    const int64_t id = tile * TW + (threadIdx.x % TW);
    const bool act = id < nd;
    const int64_t ids = act ? id : 0;
    const int rl = threadIdx.x / TW;
    const double sx = e.sx ? *e.sx : 1.0;
    for (int it = rl; it < rch; it += RPI) {
      const int64_t iu = row0 + it;
      const int cnt = scnt[it];
      const int cmax = wave_max_int(cnt);
      const uint32_t* ent = sent + it * MAXH;
      // the output's read-modify-write operands are requested before the gathers
      const bool wr = act && iu < a.nu;
      const int64_t gw = wr ? iu * ld + id : ids;
      const double yold = __builtin_nontemporal_load(y + gw);
      const double xown = x[gw];
      double acc = 0.0;
      // B divides MAXH (both powers of two) and cmax <= MAXH, so a batch never
      // reads past the row's MAXH staged slots; slots at or past cnt gather
      // from the lane's own column of row 0 with value 0.
      constexpr int B = kHopBatch < MAXH ? kHopBatch : MAXH;
      for (int k0 = 0; k0 < cmax; k0 += B) {
        uint32_t en[B];
#pragma unroll
        for (int q = 0; q < B; ++q) en[q] = ent[k0 + q];
        double g[B];
#pragma unroll
        for (int q = 0; q < B; ++q) {
          const bool ok = k0 + q < cnt;
          g[q] = x[ok ? (int64_t)hop_target(en[q]) * ld + ids : ids];
        }
#pragma unroll
        for (int q = 0; q < B; ++q) acc += (k0 + q < cnt ? hv_val<UNI>(en[q], t0, hv) : 0.0) * g[q];
      }
      if (wr) {
        const double up = sx * acc;
        __builtin_nontemporal_store(yold + up, y + gw);
        dot += up * xown;
      }
    }
  }
  if (partial) {
    const double s = block_sum_256(dot, red);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// Column pass for uniform hopping t: y += s*t*(P - N) over the up hops, one
// wave per output row of a TW-column tile.  The row's sign-sorted target list
// (positive first; MAXL entries, unused ones 0) comes in one scalar batch;
// lane group q (lanes q*TW .. q*TW+TW-1) of load j takes entry j*HP + q, so
// a load instruction covers HP hops, and all of a row's loads are issued
// before the first is consumed.  Tiles go to XCDs round-robin and each XCD
// sweeps every row of its tile, so the tile's slab of x (nu x TW doubles)
// stays in that XCD's L2.  Requires D * 8 <= 2^32 (32-bit offsets).
template <int TW, int MAXL>
__global__ __launch_bounds__(256) void k_hv_cols_split(Hv2Args a, const double* __restrict__ x,
                                                       double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                       int64_t nrc, int rch, double* partial) {
  constexpr int HP = 64 / TW;
  constexpr int NJ = MAXL / HP;
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  const int64_t xcd = b & 7, j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + xcd;
  const int64_t rc = j % nrc;
  const int64_t nd = a.nd, ld = a.ld;
  const int64_t row0 = rc * rch;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int q = lane / TW;
  double dot = 0.0;
  if (tile < ntiles) {
    const int64_t col = tile * TW + (lane % TW);
    const bool act = col < nd;
    const int64_t colc = act ? col : tile * TW;
    const uint32_t laneoff = (uint32_t)colc * 8u;
    const uint32_t nd8 = (uint32_t)ld * 8u;  // bytes per row of x
    const char* xb = (const char*)x;
    const double sx = e.sx ? *e.sx : 1.0;
    const double st = sx * a.hop[0];
    const bool wr = act && q == 0;
    for (int it = wave; it < rch; it += 4) {
      const int64_t iu = row0 + it;
      if (iu >= a.nu) break;
      const int64_t gw = iu * ld + colc;
      double yold = 0.0, xown = 0.0;
      if (wr) {
        yold = __builtin_nontemporal_load(y + gw);
        xown = x[gw];
      }
      const uint32_t cw = uniform_load(a.up_cnt_pn + iu);
      const int np = (int)(cw & 0xffffu), n = np + (int)(cw >> 16);
      const int32_t* tg = a.up_tgt + iu * MAXL;
      uint32_t tv[MAXL];
#pragma unroll
      for (int k = 0; k < MAXL; ++k) tv[k] = (uint32_t)uniform_load(tg + k) * nd8;
      double g[NJ];
#pragma unroll
      for (int jj = 0; jj < NJ; ++jj) {
        if (jj * HP < n) {
          uint32_t o = tv[jj * HP];
#pragma unroll
          for (int h = 1; h < HP; ++h) o = q == h ? tv[jj * HP + h] : o;
          g[jj] = *(const double*)(xb + (laneoff + o));
        }
      }
      double acc = 0.0;
#pragma unroll
      for (int jj = 0; jj < NJ; ++jj) {
        if (jj * HP < n) {
          const int sl = jj * HP + q;
          const double v = sl < np ? g[jj] : -g[jj];
          acc += sl < n ? v : 0.0;
        }
      }
#pragma unroll
      for (int off = TW; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
      if (wr) {
        const double up = st * acc;
        __builtin_nontemporal_store(yold + up, y + gw);
        dot += up * xown;
      }
    }
  }
  if (partial) {
    const double s = block_sum_256(dot, red);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// Buffer resource over `bytes` bytes from p: a load at a byte offset at or
// past `bytes` returns zero without touching memory (the pad entries of the
// pair lists address one row past the vector).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ double buffer_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// Lane constants of a column tile in the pair-list column pass.
struct PairLane {
  uint32_t laneoff;  // byte offset of the lane's column within a row
  uint32_t hsh;      // 0 or 16: which entry of a pair the lane's half-wave takes
  int64_t colc;      // the lane's column (clamped into the sector)
  bool wr;           // the lane writes y
};

template <int TW>
__device__ __forceinline__ PairLane pair_lane(int64_t tile, int64_t nd) {
  const int lane = threadIdx.x & 63;
  const int half = TW == 32 ? lane >> 5 : 0;
  const int64_t col = tile * TW + (lane % TW);
  const bool act = col < nd;
  PairLane L;
  L.colc = act ? col : tile * TW;
  L.laneoff = (uint32_t)L.colc * 8u;
  L.hsh = (uint32_t)half * 16u;
  L.wr = act && half == 0;
  return L;
}

// One row's pair list: header and all kPairMaxPairs words (the table carries
// kPairMaxPairs words of slack after its last row), read through the scalar
// cache.
struct PairList {
  uint32_t hdr;
  uint32_t w[kPairMaxPairs];
};

__device__ __forceinline__ void load_pair_list(const Hv2Args& a, int64_t iu, PairList& l) {
  const uint32_t* wl = a.up_pair + iu * a.up_pair_stride;
  l.hdr = uniform_load(wl);
#pragma unroll
  for (int m = 0; m < kPairMaxPairs; ++m) l.w[m] = uniform_load(wl + 1 + m);
}

// sum_pairs with a hook between issuing the loads and the first add: the
// caller's prefetch of the next row's list goes there, after this row's list
// has been consumed, so waiting for it never delays this row.
template <int TW, int NPR, typename Pre>
__device__ __forceinline__ double sum_pairs_pre(const uint32_t (&w)[kPairMaxPairs], int pp, uint32_t hsh,
                                                uint32_t nd8, uint32_t laneoff, __amdgpu_buffer_rsrc_t rx,
                                                Pre&& pre) {
  constexpr int NL = TW == 32 ? NPR : 2 * NPR;
  double g[NL];
#pragma unroll
  for (int m = 0; m < NPR; ++m) {
    if (TW == 32) {
      g[m] = buffer_load_f64(rx, __umul24((w[m] >> hsh) & 0xffffu, nd8) + laneoff);
    } else {
      g[2 * m] = buffer_load_f64(rx, __umul24(w[m] & 0xffffu, nd8) + laneoff);
      g[2 * m + 1] = buffer_load_f64(rx, __umul24(w[m] >> 16, nd8) + laneoff);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  pre();
  __builtin_amdgcn_sched_barrier(0);
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < NPR; ++m) {
    const double sg = m < pp ? 1.0 : -1.0;
    if (TW == 32) {
      acc = fma(sg, g[m], acc);
    } else {
      acc = fma(sg, g[2 * m], acc);
      acc = fma(sg, g[2 * m + 1], acc);
    }
  }
  return acc;
}

template <int TW, typename Pre>
__device__ __forceinline__ double sum_pairs_pre_n(int npr, const uint32_t (&w)[kPairMaxPairs], int pp,
                                                  uint32_t hsh, uint32_t nd8, uint32_t laneoff,
                                                  __amdgpu_buffer_rsrc_t rx, Pre&& pre) {
  switch (npr) {
#define ED_PRE_CASE(n) \
  case n: return sum_pairs_pre<TW, n>(w, pp, hsh, nd8, laneoff, rx, pre);
    ED_PRE_CASE(1) ED_PRE_CASE(2) ED_PRE_CASE(3) ED_PRE_CASE(4) ED_PRE_CASE(5) ED_PRE_CASE(6)
    ED_PRE_CASE(7) ED_PRE_CASE(8) ED_PRE_CASE(9) ED_PRE_CASE(10) ED_PRE_CASE(11) ED_PRE_CASE(12)
    ED_PRE_CASE(13) ED_PRE_CASE(14) ED_PRE_CASE(15) ED_PRE_CASE(16)
#undef ED_PRE_CASE
    default: pre(); return 0.0;
  }
}

// Column pass from the pair lists, after the row pass (y += up hops).  Block b
// runs on XCD b & 7 and takes row chunk j % nrc of tile (j / nrc) * 8 +
// (b & 7), j = b >> 3; the dispatcher issues blocks in order, so the blocks
// resident on an XCD cover a window of consecutive row chunks of (mostly)
// one tile and that tile's slab (num_up x TW doubles) is what its L2 holds.
// The four waves of a block take the chunk's rows in turn, each prefetching
// its next row's list while the current row's gathers are in flight.
// Requires (num_up + 1) * num_dn * 8 < 2^32 (32-bit buffer offsets, the pad
// row included).
template <int TW>
__global__ __launch_bounds__(256) void k_hv_cols_pipe(Hv2Args a, const double* __restrict__ x,
                                                      double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                      int64_t nrc, int rch, double* partial) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  const int64_t j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + (b & 7);
  const int64_t row0 = (j % nrc) * rch;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double dot = 0.0;
  if (tile < ntiles) {
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, (uint32_t)(a.nu * a.ld * 8));
    const double st = (e.sx ? *e.sx : 1.0) * a.hop[0];
    const PairLane L = pair_lane<TW>(tile, a.nd);
    const uint32_t nd8 = (uint32_t)a.ld * 8u;  // bytes per row of x
    const int64_t rend = row0 + rch < a.nu ? row0 + rch : a.nu;
    int64_t iu = row0 + wave;
    PairList nxt;
    if (iu < rend) load_pair_list(a, iu, nxt);
    // the previous row's y store is issued after this row's gathers, so
    // reusing its data register never waits for it while loads are pending
    double pend = 0.0;
    double* pend_at = nullptr;
    while (iu < rend) {
      const PairList cur = nxt;
      const int64_t inx = iu + 4;
      const int pp = (int)(cur.hdr & 0xffu), npr = (int)((cur.hdr >> 8) & 0xffu);
      const int64_t gw = iu * a.ld + L.colc;
      double yold = 0.0, xown = 0.0;
      if (L.wr) {
        yold = __builtin_nontemporal_load(y + gw);
        if (partial) xown = x[gw];
      }
      double acc = sum_pairs_pre_n<TW>(npr, cur.w, pp, L.hsh, nd8, L.laneoff, rx, [&]() {
        if (L.wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
        if (inx < rend) load_pair_list(a, inx, nxt);
      });
      if (TW == 32) acc += __shfl_xor(acc, 32, 64);
      const double up = st * acc;
      pend = yold + up;
      pend_at = y + gw;
      dot += up * xown;
      iu = inx;
    }
    if (L.wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
  }
  if (partial) {
    const double s = block_sum_256(dot, red);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// Column pass run before the row pass: y = s * t * (up hops) over each tile,
// written without reading y (the row pass then adds the diagonal, the dn
// hops and the epilogue from whole contiguous rows, and forms the dot).  Grid,
// row chunks and list prefetch as in k_hv_cols_pipe.
template <int TW>
__global__ __launch_bounds__(256) void k_hv_cols_first(Hv2Args a, const double* __restrict__ x,
                                                       double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                       int64_t nrc, int rch, double*) {
  const int64_t b = blockIdx.x;
  const int64_t j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + (b & 7);
  const int64_t row0 = (j % nrc) * rch;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (tile >= ntiles) return;
  const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, (uint32_t)(a.nu * a.ld * 8));
  const double st = (e.sx ? *e.sx : 1.0) * a.hop[0];
  const PairLane L = pair_lane<TW>(tile, a.nd);
  const uint32_t nd8 = (uint32_t)a.ld * 8u;  // bytes per row of x
  const int64_t rend = row0 + rch < a.nu ? row0 + rch : a.nu;
  int64_t iu = row0 + wave;
  PairList nxt;
  if (iu < rend) load_pair_list(a, iu, nxt);
  // the previous row's store is issued after this row's gathers (see
  // k_hv_cols_pipe)
  double pend = 0.0;
  double* pend_at = nullptr;
  while (iu < rend) {
    const PairList cur = nxt;
    const int64_t inx = iu + 4;
    const int pp = (int)(cur.hdr & 0xffu), npr = (int)((cur.hdr >> 8) & 0xffu);
    double acc = sum_pairs_pre_n<TW>(npr, cur.w, pp, L.hsh, nd8, L.laneoff, rx, [&]() {
      if (L.wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
      if (inx < rend) load_pair_list(a, inx, nxt);
    });
    if (TW == 32) acc += __shfl_xor(acc, 32, 64);
    pend = st * acc;
    pend_at = y + iu * a.ld + L.colc;
    iu = inx;
  }
  if (L.wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
}

typedef double f64x2 __attribute__((ext_vector_type(2)));

// One row's quad list: header and kQuadWords - 1 words (the table carries
// that much slack after its last row), through the scalar cache.
struct QuadList {
  uint32_t hdr;
  uint32_t w[2 * kQuadMaxQuads];
};

__device__ __forceinline__ void load_quad_list(const Hv2Args& a, int64_t iu, QuadList& l) {
  const uint32_t* wl = a.up_quad + iu * a.up_quad_stride;
  l.hdr = uniform_load(wl);
#pragma unroll
  for (int m = 0; m < 2 * kQuadMaxQuads; ++m) l.w[m] = uniform_load(wl + 1 + m);
}

// Up hops of one row over a 32-column tile from its quad list (NQ quads, the
// first qp of sign +1): quarter-wave q takes entry q of every quad and each
// lane reads two adjacent columns (16 bytes), so one load serves four hops.
// pre() runs between issuing the loads and the first add.
// psel: byte selector of the lane's quarter (entry q of a quad is bytes
// 2q, 2q + 1 of the word pair, zero-extended).
template <int NQ, typename Pre>
__device__ __forceinline__ void sum_quads(const uint32_t (&w)[2 * kQuadMaxQuads], int qp, uint32_t psel,
                                          uint32_t nd8, uint32_t laneoff, __amdgpu_buffer_rsrc_t rx, double& s0,
                                          double& s1, Pre&& pre) {
  f64x2 g[NQ];
#pragma unroll
  for (int m = 0; m < NQ; ++m) {
    const uint32_t ent = __builtin_amdgcn_perm(w[2 * m + 1], w[2 * m], psel);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rx, __umul24(ent, nd8) + laneoff, 0, 0);
    g[m] = __builtin_bit_cast(f64x2, v);
  }
  __builtin_amdgcn_sched_barrier(0);
  pre();
  __builtin_amdgcn_sched_barrier(0);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int m = 0; m < NQ; ++m) {
    const double sg = m < qp ? 1.0 : -1.0;
    a0 = fma(sg, g[m].x, a0);
    a1 = fma(sg, g[m].y, a1);
  }
  s0 = a0;
  s1 = a1;
}

template <typename Pre>
__device__ __forceinline__ void sum_quads_n(int nq, const uint32_t (&w)[2 * kQuadMaxQuads], int qp, uint32_t psel,
                                            uint32_t nd8, uint32_t laneoff, __amdgpu_buffer_rsrc_t rx, double& s0,
                                            double& s1, Pre&& pre) {
  switch (nq) {
#define ED_QUAD_CASE(n) \
  case n: sum_quads<n>(w, qp, psel, nd8, laneoff, rx, s0, s1, pre); return;
    ED_QUAD_CASE(1) ED_QUAD_CASE(2) ED_QUAD_CASE(3) ED_QUAD_CASE(4) ED_QUAD_CASE(5) ED_QUAD_CASE(6)
    ED_QUAD_CASE(7) ED_QUAD_CASE(8) ED_QUAD_CASE(9) ED_QUAD_CASE(10) ED_QUAD_CASE(11) ED_QUAD_CASE(12)
#undef ED_QUAD_CASE
    default: pre(); s0 = 0.0; s1 = 0.0; return;
  }
}

// Column pass run before the row pass (as k_hv_cols_first) from the quad
// lists: 32-column tiles, 16-byte loads, four hops per load.  Requires an
// even num_dn and (num_up + 1) * num_dn * 8 < 2^32.
__global__ __launch_bounds__(256) void k_hv_cols_quad(Hv2Args a, const double* __restrict__ x,
                                                      double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                      int64_t nrc, int rch, double*) {
  const int64_t b = blockIdx.x;
  const int64_t j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + (b & 7);
  const int64_t row0 = (j % nrc) * rch;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4;
  // bytes 2q and 2q + 1 of {w[2m + 1], w[2m]} into the low half, zeros above
  const uint32_t psel = 0x0c0c0000u | ((uint32_t)(2 * q + 1) << 8) | (uint32_t)(2 * q);
  const int64_t nd = a.nd;
  const int64_t col = tile * 32 + 2 * (lane & 15);
  const bool act = col < nd;
  const int64_t colc = act ? col : nd - 2;
  const uint32_t laneoff = (uint32_t)colc * 8u;
  const bool wr = act && q == 0;
  const int64_t ld = a.ld;
  const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, (uint32_t)(a.nu * ld * 8));
  const double st = (e.sx ? *e.sx : 1.0) * a.hop[0];
  const uint32_t nd8 = (uint32_t)ld * 8u;  // bytes per row of x
  const int64_t rend = row0 + rch < a.nu ? row0 + rch : a.nu;
  int64_t iu = row0 + wave;
  QuadList nxt;
  if (iu < rend) load_quad_list(a, iu, nxt);
  // the previous row's store is issued after this row's gathers (see
  // k_hv_cols_pipe)
  f64x2 pend = {0.0, 0.0};
  f64x2* pend_at = nullptr;
  while (iu < rend) {
    const QuadList cur = nxt;
    const int64_t inx = iu + 4;
    const int qp = (int)(cur.hdr & 0xffu), nq = (int)((cur.hdr >> 8) & 0xffu);
    double s0, s1;
    sum_quads_n(nq, cur.w, qp, psel, nd8, laneoff, rx, s0, s1, [&]() {
      if (wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
      if (inx < rend) load_quad_list(a, inx, nxt);
    });
    s0 += __shfl_xor(s0, 16, 64);
    s1 += __shfl_xor(s1, 16, 64);
    s0 += __shfl_xor(s0, 32, 64);
    s1 += __shfl_xor(s1, 32, 64);
    pend = f64x2{st * s0, st * s1};
    pend_at = (f64x2*)(y + iu * ld + colc);
    iu = inx;
  }
  if (wr && pend_at) __builtin_nontemporal_store(pend, pend_at);
}

// Rows per block of the oct column pass: their quad lists are staged in LDS.
constexpr int kOctRows = 16;

// Word k of staged row r's quad list (header at k = 0).
__device__ __forceinline__ uint32_t oct_word(const uint32_t* lists, int r, int stride, int k) {
  return lists[r * stride + k];
}

// Two rows' up hops over a 16-column tile, two quads (an oct) per load:
// eighth ei of the wave (lanes 8 ei .. 8 ei + 7) takes entry ei of the oct,
// the half ei & 1 of its word ei >> 1, and each lane reads two adjacent
// columns, so one load serves eight hops.  Entries of quad 2m + (ei >> 2)
// have sign +1 below qp quads; quads from nq on (the second half of a row's
// last oct when nq is odd, and every oct past the row's own when the other
// row has more) read the pad row, which the buffer range turns into zeros.
// psel picks entry ei & 3 of a word pair.  NO octs per row, all loads of
// both rows in flight together.
template <int NO>
__device__ __forceinline__ void sum_octs2(const uint32_t* lists, int stride, int ra, int rb, const int (&qp)[2],
                                          const int (&nq)[2], uint32_t pad, int ei, uint32_t psel, uint32_t nd8,
                                          uint32_t laneoff, __amdgpu_buffer_rsrc_t rx, double (&s)[2][2]) {
  f64x2 g[2][NO];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = h ? rb : ra;
#pragma unroll
    for (int m = 0; m < NO; ++m) {
      const uint32_t w0 = oct_word(lists, r, stride, 1 + 4 * m), w1 = oct_word(lists, r, stride, 2 + 4 * m);
      const uint32_t w2 = oct_word(lists, r, stride, 3 + 4 * m), w3 = oct_word(lists, r, stride, 4 + 4 * m);
      const uint32_t lo = __builtin_amdgcn_perm(w1, w0, psel);
      const uint32_t hi = __builtin_amdgcn_perm(w3, w2, psel);
      const uint32_t ent = 2 * m + (ei >> 2) < nq[h] ? (ei >= 4 ? hi : lo) : pad;
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rx, __umul24(ent, nd8) + laneoff, 0, 0);
      g[h][m] = __builtin_bit_cast(f64x2, v);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int m = 0; m < NO; ++m) {
      const double sg = 2 * m + (ei >> 2) < qp[h] ? 1.0 : -1.0;
      a0 = fma(sg, g[h][m].x, a0);
      a1 = fma(sg, g[h][m].y, a1);
    }
    s[h][0] = a0;
    s[h][1] = a1;
  }
}

// Column pass over 16-column tiles (one 128-byte line per row of x) from the
// quad lists, run first like k_hv_cols_quad: a tile's slab of x is half as
// large (1.65 MB for 12,870 rows), so more of it stays in the XCD's L2 while
// the tile's rows gather from it.  The block's rows' lists are staged in LDS
// and each wave takes two rows at a time.  Requires rch <= kOctRows, an even
// num_dn and (num_up + 1) * num_dn * 8 < 2^32.
__global__ __launch_bounds__(256) void k_hv_cols_oct(Hv2Args a, const double* __restrict__ x,
                                                     double* __restrict__ y, HvEpilogue e, int64_t ntiles,
                                                     int64_t nrc, int rch, double*) {
  __shared__ uint32_t lists[kOctRows * kQuadWords];
  const int64_t b = blockIdx.x;
  const int64_t j = b >> 3;
  const int64_t tile = (j / nrc) * 8 + (b & 7);
  const int64_t row0 = (j % nrc) * rch;
  if (tile >= ntiles) return;  // the whole block
  const int64_t rend = row0 + rch < a.nu ? row0 + rch : a.nu;
  const int nrows = (int)(rend - row0);
  const int stride = a.up_quad_stride;
  for (int t = threadIdx.x; t < nrows * stride; t += blockDim.x) lists[t] = a.up_quad[row0 * stride + t];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int ei = lane >> 3;
  const int q = ei & 3;
  const uint32_t psel = 0x0c0c0000u | ((uint32_t)(2 * q + 1) << 8) | (uint32_t)(2 * q);
  const int64_t nd = a.nd;
  const int64_t col = tile * 16 + 2 * (lane & 7);
  const bool act = col < nd;
  const int64_t colc = act ? col : nd - 2;
  const uint32_t laneoff = (uint32_t)colc * 8u;
  const bool wr = act && ei == 0;
  const int64_t ld = a.ld;
  const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, (uint32_t)(a.nu * ld * 8));
  const double st = (e.sx ? *e.sx : 1.0) * a.hop[0];
  const uint32_t nd8 = (uint32_t)ld * 8u;
  const uint32_t pad = (uint32_t)a.nu;
  for (int ra = wave; ra < nrows; ra += 8) {
    const int rb = ra + 4;
    const bool hb = rb < nrows;  // uniform
    const uint32_t h0 = oct_word(lists, ra, stride, 0);
    const uint32_t h1 = hb ? oct_word(lists, rb, stride, 0) : 0u;
    const int qp[2] = {(int)(h0 & 0xffu), (int)(h1 & 0xffu)};
    const int nq[2] = {(int)((h0 >> 8) & 0xffu), (int)((h1 >> 8) & 0xffu)};
    const int no = ((nq[0] > nq[1] ? nq[0] : nq[1]) + 1) >> 1;
    double sm[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    const int rbc = hb ? rb : ra;  // a missing second row reads the first row's words, all masked
    switch (no) {
#define ED_OCT_CASE(n) \
  case n: sum_octs2<n>(lists, stride, ra, rbc, qp, nq, pad, ei, psel, nd8, laneoff, rx, sm); break;
      ED_OCT_CASE(1) ED_OCT_CASE(2) ED_OCT_CASE(3) ED_OCT_CASE(4) ED_OCT_CASE(5) ED_OCT_CASE(6)
#undef ED_OCT_CASE
      default: break;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double v = sm[h][c];
        v += __shfl_xor(v, 8, 64);
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        sm[h][c] = v;
      }
    if (wr) {
      __builtin_nontemporal_store(f64x2{st * sm[0][0], st * sm[0][1]}, (f64x2*)(y + (row0 + ra) * ld + colc));
      if (hb)
        __builtin_nontemporal_store(f64x2{st * sm[1][0], st * sm[1][1]}, (f64x2*)(y + (row0 + rb) * ld + colc));
    }
  }
}

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

static Hv2Args make_hv2_args(const ed_op* op, int64_t ld) {
  Hv2Args a;
  a.nu = op->num_up;
  a.nd = op->num_dn;
  a.ld = ld;
  a.st_up = op->st_up;
  a.st_dn = op->st_dn;
  a.d_up = op->d_up;
  a.d_dn = op->d_dn;
  a.up_cnt = op->up.cnt;
  a.up_pk = op->up.pk_rm;
  a.up_w = op->up.width;
  a.dn_cnt = op->dn.cnt;
  a.dn_pk = op->dn.pk_cm;
  a.dn_w = op->dn.width;
  a.dn_off_pos = op->dn.off_pos;
  a.dn_off_neg = op->dn.off_neg;
  a.dn_cmax_pos = op->dn.cmax_pos;
  a.dn_cmax_neg = op->dn.cmax_neg;
  a.dn_h16 = op->dn.h16;
  a.dn_h16q = op->dn.h16q;
  a.dn_hgroups = op->dn.hgroups;
  a.dn_hwords = op->dn.hwords_chunk;
  a.up_tgt = op->up.tgt_sorted;
  a.up_cnt_pn = op->up.cnt_pn;
  a.up_pair = op->up.pair16;
  a.up_pair_stride = op->up.pair_stride;
  a.up_quad = op->up.quad16;
  a.up_quad_stride = op->up.quad_stride;
  a.hop = op->hop_dev;
  a.nb = op->num_bonds;
  a.u_mode = op->u_mode;
  a.u_uniform = op->u_uniform;
  a.u_mask = op->u_mask;
  a.u_site = op->u_site;
  return a;
}

using RowKernel = void (*)(Hv2Args, int, const double*, double*, HvEpilogue);
using ColKernel = void (*)(Hv2Args, const double*, double*, HvEpilogue, int64_t, int64_t, int, double*);

struct TwoPassPlan {
  bool ok;
  int rows_per_block;
  size_t lds_bytes;
  int64_t grid_r;
  int tw, rch;
  int64_t ntiles, nrc, grid_c;
  RowKernel rk;
  ColKernel ck;
  bool col_first;  // column pass writes y first, the row pass adds to it
};

// Smallest compiled hop-list length >= m (0 if none fits).
static int maxh_for(int m) {
  if (m <= 8) return 8;
  if (m <= 16) return 16;
  if (m <= 32) return 32;
  if (m <= 64) return 64;
  return 0;
}

// The half-width row kernels address their dynamic LDS from offset 0
// (lds_f64* row = 0), which is right only while they declare no static LDS:
// true if the compiled kernel has none (checked once per kernel).
static bool lds_starts_at_zero(RowKernel k) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, bool>> seen;
  std::lock_guard<std::mutex> lk(mu);
  for (auto& v : seen)
    if (v.first == (const void*)k) return v.second;
  hipFuncAttributes attr{};
  const bool ok = hipFuncGetAttributes(&attr, (const void*)k) == hipSuccess && attr.sharedSizeBytes == 0;
  if (!ok) (void)hipGetLastError();
  seen.emplace_back((const void*)k, ok);
  return ok;
}

// Grants a row kernel `bytes` of dynamic LDS (needed above 64 KiB); false if refused.
static bool ensure_row_lds(RowKernel k, size_t bytes) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> granted;
  if (bytes <= 65536) return true;
  std::lock_guard<std::mutex> lk(mu);
  for (auto& g : granted)
    if (g.first == (const void*)k && g.second >= bytes) return true;
  const hipError_t err = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  granted.emplace_back((const void*)k, bytes);
  return true;
}

template <int TW>
static ColKernel col_kernel_tw(int maxh, bool uni) {
  switch (maxh) {
    case 8: return uni ? k_hv_cols<TW, 8, true> : k_hv_cols<TW, 8, false>;
    case 16: return uni ? k_hv_cols<TW, 16, true> : k_hv_cols<TW, 16, false>;
    case 32: return uni ? k_hv_cols<TW, 32, true> : k_hv_cols<TW, 32, false>;
    default: return uni ? k_hv_cols<TW, 64, true> : k_hv_cols<TW, 64, false>;
  }
}

template <int TW>
static ColKernel col_kernel_split_tw(int stride) {
  switch (stride) {
    case 8: return k_hv_cols_split<TW, 8>;
    case 16: return k_hv_cols_split<TW, 16>;
    case 24: return k_hv_cols_split<TW, 24>;
    default: return k_hv_cols_split<TW, 32>;
  }
}

static ColKernel col_kernel_split(int tw, int maxl) {
  if (tw == 16) return col_kernel_split_tw<16>(maxl);
  if (tw == 64) return col_kernel_split_tw<64>(maxl);
  return col_kernel_split_tw<32>(maxl);
}

constexpr int64_t kMinRowBlocks = 512;

// Below this many states the single fused pass (k_hv) is the faster sweep:
// its rows gather their up-hop sources straight from x, which at this size
// stays in the caches, and it is one launch instead of two.  Measured
// back-to-back on MI355X: 12-site chain (6,6), D = 853,776: 19.0 us fused
// against 26.5 us in two passes; 14-site chain (7,7), D = 11.8 M: 208 against
// 177 us; 4x4 (5,5), D = 19.1 M: 1063 against 307 us (break-even near 3 M).
constexpr int64_t kTwoPassMinStates = int64_t(1) << 21;

// Two-pass plan.  Kernels: sign-split tables when hopping is uniform and the
// tables exist (CMPY_AMD_HV_SPLIT=0 turns them off), general packed tables
// otherwise.  CMPY_AMD_HV_MODE: 0 single fused pass, 1 two passes, unset =
// two passes from kTwoPassMinStates up.
static TwoPassPlan plan_two_pass(const ed_op* op, int64_t ld) {
  TwoPassPlan p{};
  p.ok = false;
  const int mode = op->hv_mode >= 0 ? op->hv_mode : env_int("CMPY_AMD_HV_MODE", -1);
  if (!op->packed || mode == 0) return p;
  if (mode < 0 && op->size < kTwoPassMinStates) return p;
  const int maxh_up = maxh_for(op->max_hops_up);
  if (!maxh_up || !maxh_for(op->max_hops_dn)) return p;
  bool uni = op->num_bonds > 0;
  for (double h : op->hop)
    if (h != op->hop[0]) uni = false;
  const bool split = uni && op->dn.off_pos && op->up.tgt_sorted && op->num_up * ld * 8 <= (1ll << 32) &&
                     op->up.tgt_stride <= 32 && env_int("CMPY_AMD_HV_SPLIT", 1) != 0;
  int dev = 0, max_lds = 65536;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) max_lds = 65536;
  // leave room for the kernel's static LDS and the hop-value slots
  const size_t budget = (size_t)max_lds - 8 * 128 - 4096;
  const size_t row_bytes = sizeof(double) * (size_t)(op->num_dn + (split ? 1 : 0));
  int R = (int)std::min<size_t>(budget / row_bytes, 32);
  if (R < 1) return p;
  R = (int)std::min<int64_t>(R, op->num_up);
  // keep at least ~2 row blocks per CU: short rows (small num_dn) would
  // otherwise pack so many rows per block that most of the chip idles
  R = (int)std::max<int64_t>(1, std::min<int64_t>(R, op->num_up / kMinRowBlocks));
  p.rows_per_block = R;
  // one-row blocks with the half-width tables where at most two rows fit
  const bool half = split && R <= 2 && op->dn.h16 &&
                    sizeof(double) * (size_t)(op->num_dn + 32 + kRowThreads / 64) <= budget &&
                    env_int("CMPY_AMD_HV_HALF", 1) != 0 && lds_starts_at_zero(k_hv_rows_half<false>) &&
                    lds_starts_at_zero(k_hv_rows_half<true>) && lds_starts_at_zero(k_hv_rows_held<true>) &&
                    lds_starts_at_zero(k_hv_rows_held<true, true>);
  if (half) {
    p.rows_per_block = R = 1;
    p.lds_bytes = sizeof(double) * (size_t)(op->num_dn + 32 + kRowThreads / 64);
    p.rk = k_hv_rows_half<false>;
  } else if (split) {
    p.lds_bytes = row_bytes * R;
    p.rk = R == 1 ? k_hv_rows_off<1, false> : k_hv_rows_off<4, false>;
  } else {
    p.lds_bytes = row_bytes * R + 8 * 128;
    p.rk = uni ? k_hv_rows<true> : k_hv_rows<false>;
  }
  if (!ensure_row_lds(p.rk, p.lds_bytes)) return p;
  p.grid_r = (op->num_up + R - 1) / R;
  if (split) {
    // 64-column tiles: one hop per load instruction, so the source row and
    // the sign are wave-uniform
    p.tw = env_int("CMPY_AMD_HV_TW", 64);
    if (p.tw != 16 && p.tw != 32) p.tw = 64;
    p.rch = env_int("CMPY_AMD_HV_RCH", 32);
    p.rch = std::max(4, std::min(256, p.rch)) / 4 * 4;
    p.ck = col_kernel_split(p.tw, op->up.tgt_stride);
  } else {
    p.tw = env_int("CMPY_AMD_HV_TW", 32);
    if (p.tw != 16 && p.tw != 32) p.tw = 32;
    // rows per column-pass block: a multiple of the rows one pass of the block
    // covers (256 / TW), so every wave runs the same number of row iterations
    const int rpi = 256 / p.tw;
    p.rch = env_int("CMPY_AMD_HV_RCH", 32);
    p.rch = std::max(rpi, std::min(kColMaxRows, p.rch));
    p.rch = (p.rch / rpi) * rpi;
    p.ck = p.tw == 16 ? col_kernel_tw<16>(maxh_up, uni) : col_kernel_tw<32>(maxh_up, uni);
  }
  p.ntiles = (op->num_dn + p.tw - 1) / p.tw;
  p.nrc = (op->num_up + p.rch - 1) / p.rch;
  p.grid_c = 8 * ((p.ntiles + 7) / 8) * p.nrc;
  // column pass from the pair or quad lists (32-bit buffer offsets with the
  // pad row included); CMPY_AMD_HV_PAIR=0 turns it off, CMPY_AMD_HV_PTW picks
  // 32- (default) or 64-column tiles for the pair lists, CMPY_AMD_HV_PRCH
  // rows per block, CMPY_AMD_HV_FIRST=0 runs it after the row pass
  if (split && op->up.pair16 && (op->num_up + 1) * ld * 8 < (1ll << 32) &&
      env_int("CMPY_AMD_HV_PAIR", 1) != 0) {
    p.tw = env_int("CMPY_AMD_HV_PTW", 32) == 64 ? 64 : 32;
    p.rch = std::max(4, std::min(256, env_int("CMPY_AMD_HV_PRCH", 16))) / 4 * 4;
    p.ntiles = (op->num_dn + p.tw - 1) / p.tw;
    if (env_int("CMPY_AMD_HV_FIRST", 1) != 0) {
      // column pass first, writing y; the row pass adds to it.  The quad
      // lists with 16-byte loads where num_dn is even (CMPY_AMD_HV_QUAD=0
      // turns them off)
      p.col_first = true;
      if (op->up.quad16 && op->num_dn % 2 == 0 && ld % 2 == 0 && env_int("CMPY_AMD_HV_QUAD", 1) != 0) {
        // CMPY_AMD_HV_OCT=1: 16-column tiles from the same lists
        const bool oct = env_int("CMPY_AMD_HV_OCT", 0) != 0;
        if (oct) p.rch = std::min(p.rch, kOctRows);
        p.tw = oct ? 16 : 32;
        p.ntiles = (op->num_dn + p.tw - 1) / p.tw;
        p.ck = oct ? k_hv_cols_oct : k_hv_cols_quad;
      } else {
        p.ck = p.tw == 64 ? k_hv_cols_first<64> : k_hv_cols_first<32>;
      }
      p.nrc = (op->num_up + p.rch - 1) / p.rch;
      if (p.rk == k_hv_rows_half<false>)
        p.rk = op->num_dn <= (int64_t)kRowHeld * kRowThreads && env_int("CMPY_AMD_HV_HELD", 1) != 0
                   ? (op->num_up * ld >= kStreamStates ? k_hv_rows_held<true, true> : k_hv_rows_held<true>)
                   : k_hv_rows_half<true>;
      else if (p.rk == k_hv_rows_off<1, false>)
        p.rk = k_hv_rows_off<1, true>;
      else
        p.rk = k_hv_rows_off<4, true>;
      if (!ensure_row_lds(p.rk, p.lds_bytes)) {
        p.ok = false;
        return p;
      }
    } else {
      p.ck = p.tw == 64 ? k_hv_cols_pipe<64> : k_hv_cols_pipe<32>;
      p.nrc = (op->num_up + p.rch - 1) / p.rch;
    }
    p.grid_c = 8 * ((p.ntiles + 7) / 8) * p.nrc;
  }
  p.ok = p.grid_c + p.grid_r < (1ll << 31);
  return p;
}

static int launch_two_pass(const ed_op* op, const TwoPassPlan& p, const double* x, double* y, int64_t ld,
                           const HvEpilogue& epi, hipStream_t s, int* num_partials) {
  const Hv2Args a = make_hv2_args(op, ld);
  if (p.col_first) {
    hipLaunchKernelGGL(p.ck, dim3((unsigned)p.grid_c), dim3(256), 0, s, a, x, y, epi, p.ntiles, p.nrc, p.rch,
                       nullptr);
    hipLaunchKernelGGL(p.rk, dim3((unsigned)p.grid_r), dim3(kRowThreads), p.lds_bytes, s, a,
                       p.rows_per_block, x, y, epi);
    ED_HIP(hipGetLastError());
    if (num_partials) *num_partials = (int)p.grid_r;
    return ED_OK;
  }
  hipLaunchKernelGGL(p.rk, dim3((unsigned)p.grid_r), dim3(kRowThreads), p.lds_bytes, s, a,
                     p.rows_per_block, x, y, epi);
  double* part_c = epi.partial ? epi.partial + p.grid_r : nullptr;
  hipLaunchKernelGGL(p.ck, dim3((unsigned)p.grid_c), dim3(256), 0, s, a, x, y, epi, p.ntiles, p.nrc, p.rch, part_c);
  ED_HIP(hipGetLastError());
  if (num_partials) *num_partials = (int)(p.grid_r + p.grid_c);
  return ED_OK;
}

static int pick_rb(const ed_op* op) {
  if (op->num_up >= 4096) return 4;
  if (op->num_up >= 64) return 2;
  return 1;
}

static void grid_shape(const ed_op* op, int rb, int64_t ld, int64_t* nrb, int64_t* nct) {
  *nrb = (op->num_up + rb - 1) / rb;
  *nct = (ld + 255) / 256;  // column tiles cover the row padding too (written as zeros)
}

int matvec_num_partials(const ed_op* op, int64_t ld) {
  const TwoPassPlan p = plan_two_pass(op, ld);
  if (p.ok) return (int)(p.col_first ? p.grid_r : p.grid_r + p.grid_c);
  int64_t nrb, nct;
  grid_shape(op, pick_rb(op), ld, &nrb, &nct);
  return (int)(nrb * nct);
}

int64_t vector_ld(const ed_op* op) {
  const int64_t nd = op->num_dn;
  const int64_t ld = (nd + 15) / 16 * 16;
  if (ld == nd || env_int("CMPY_AMD_HV_PITCH", 1) == 0) return nd;
  // pad only where the two-pass sweep runs on the padded rows (its column
  // slabs then cover whole 128-byte lines)
  return plan_two_pass(op, ld).ok ? ld : nd;
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

int launch_matvec(const ed_op* op, const double* x, double* y, int64_t ld, const HvEpilogue& epi, hipStream_t s,
                  int* num_partials) {
  const TwoPassPlan p = plan_two_pass(op, ld);
  if (p.ok) return launch_two_pass(op, p, x, y, ld, epi, s, num_partials);
  HvArgs a = make_hv_args(op);
  a.ld = ld;
  const int rb = pick_rb(op);
  int64_t nrb, nct;
  grid_shape(op, rb, ld, &nrb, &nct);
  ED_REQUIRE(nrb * nct < (1ll << 31), "grid too large");
  const bool lz = epi.xprev || epi.sx || epi.cprev || epi.partial || epi.cself || epi.psi;
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
    ED_TRY(launch_matvec(op, x_dev + v * D, y_dev + v * D, op->num_dn, none, as_stream(stream), nullptr));
  return ED_OK;
}

extern "C" int ed_op_matvec_ld(ed_op* op, const double* x_dev, double* y_dev, int64_t ld, int64_t nvec,
                               void* stream) {
  clear_error();
  ED_REQUIRE(op && x_dev && y_dev, "NULL argument");
  ED_REQUIRE(nvec >= 0, "nvec < 0");
  ED_REQUIRE(ld >= op->num_dn && ld <= op->num_dn + 4096, "ld=%lld outside [num_dn, num_dn + 4096]",
             (long long)ld);
  const int64_t n = op->num_up * ld;  // doubles per vector
  const char* xb = (const char*)x_dev;
  const char* yb = (const char*)y_dev;
  ED_REQUIRE(xb + sizeof(double) * n * nvec <= yb || yb + sizeof(double) * n * nvec <= xb, "x and y overlap");
  HvEpilogue none{nullptr, nullptr, nullptr, nullptr};
  for (int64_t v = 0; v < nvec; ++v)
    ED_TRY(launch_matvec(op, x_dev + v * n, y_dev + v * n, ld, none, as_stream(stream), nullptr));
  return ED_OK;
}

extern "C" int ed_op_vector_ld(const ed_op* op, int64_t* ld) {
  clear_error();
  ED_REQUIRE(op && ld, "NULL argument");
  *ld = vector_ld(op);
  return ED_OK;
}
