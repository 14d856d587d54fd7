// Basis, projector tables and the sector-operator handle.
//
// Reference counterparts:
//   Basis.generate_states / get_states        cmpy/basis.py:655-676
//   bisect_left                               cmpy/operators.py:276-299
//   bit_count / _hopping_sign                 cmpy/operators.py:253-273, 425-433
//   _compute_hopping_term / project_hopping   cmpy/operators.py:436-527
//   weighted_element / project_onsite_energy  cmpy/operators.py:226-250, 359-422
//   project_hubbard_inter                     cmpy/operators.py:305-356
//   AbstractManyBodyModel.hamilton_operator   cmpy/models/abc.py:242-256
#include "ed_internal.h"
#include "ed_colouring.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace ed {

static thread_local std::string g_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

void clear_error() { g_error.clear(); }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void k_unrank(uint32_t* out, int64_t count, int num_sites, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = comb_unrank(i, num_sites, n);
}

__global__ void k_rank(const uint32_t* states, int64_t count, int num_sites, int n, int64_t* ranks) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = states[i];
  const bool ok = popc32(s) == n && (num_sites >= 32 || (s >> num_sites) == 0);
  ranks[i] = ok ? comb_rank(s) : -1;
}

// One hop term per state of the spin sector for a single bond.
__global__ void k_hop_terms(int64_t count, int num_sites, int n, int width, int site1, int site2,
                            double hop, int64_t* target, double* value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = comb_unrank(i, num_sites, n);
  const uint32_t op1 = 1u << site1, op2 = 1u << site2;
  const bool occ1 = (s & op1) != 0, occ2 = (s & op2) != 0;
  if (occ1 != occ2) {
    const uint32_t t = s ^ op1 ^ op2;
    target[i] = comb_rank(t);
    value[i] = (double)hop_sign(s, width, site1, site2) * hop;
  } else {
    target[i] = -1;
    value[i] = 0.0;
  }
}

// ELL hop table of one spin sector over all bonds; both layouts.
__global__ void k_hop_table(const uint32_t* states, int64_t count, int width, const int32_t* bonds,
                            const double* hops, int nb, int32_t* cnt, int32_t* col_rm,
                            double* val_rm, int32_t* col_cm, double* val_cm, uint32_t* pk_rm,
                            uint32_t* pk_cm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = states[i];
  int k = 0;
  for (int b = 0; b < nb; ++b) {
    const int s1 = bonds[2 * b], s2 = bonds[2 * b + 1];
    const uint32_t op1 = 1u << s1, op2 = 1u << s2;
    const bool occ1 = (s & op1) != 0, occ2 = (s & op2) != 0;
    if (occ1 == occ2) continue;
    const int32_t t = (int32_t)comb_rank(s ^ op1 ^ op2);
    const int sg = hop_sign(s, width, s1, s2);
    const double v = (double)sg * hops[b];
    col_rm[i * nb + k] = t;
    val_rm[i * nb + k] = v;
    col_cm[(int64_t)k * count + i] = t;
    val_cm[(int64_t)k * count + i] = v;
    if (pk_rm) {
      pk_rm[i * nb + k] = pack_hop(t, b, sg);
      pk_cm[(int64_t)k * count + i] = pack_hop(t, b, sg);
    }
    ++k;
  }
  for (int kk = k; kk < nb; ++kk) {
    col_rm[i * nb + kk] = 0;
    val_rm[i * nb + kk] = 0.0;
    col_cm[(int64_t)kk * count + i] = 0;
    val_cm[(int64_t)kk * count + i] = 0.0;
    if (pk_rm) {
      pk_rm[i * nb + kk] = 0;
      pk_cm[(int64_t)kk * count + i] = 0;
    }
  }
  cnt[i] = k;
}

// Counts of each state's hops with sign +1 and -1 (packed column-major table).
__global__ void k_sign_split(const uint32_t* pk_cm, const int32_t* cnt, int64_t count, int32_t* np,
                             int32_t* nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int p = 0, m = 0;
  for (int k = 0; k < cnt[i]; ++k) {
    if (pk_cm[(int64_t)k * count + i] & 0x80000000u)
      ++m;
    else
      ++p;
  }
  np[i] = p;
  nn[i] = m;
}

// Per 64-state chunk: the largest positive and negative count.
__global__ void k_chunk_max(const int32_t* np, const int32_t* nn, int64_t count, int32_t* cmp,
                            int32_t* cmn) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (count + 63) / 64) return;
  int mp = 0, mn = 0;
  for (int64_t i = c * 64; i < count && i < (c + 1) * 64; ++i) {
    mp = max(mp, (int)np[i]);
    mn = max(mn, (int)nn[i]);
  }
  cmp[c] = mp;
  cmn[c] = mn;
}

// Sign-split byte-offset tables (see Tables::off_pos), bond order kept in each.
__global__ void k_offset_tables(const uint32_t* pk_cm, const int32_t* cnt, int64_t count, int hp, int hn,
                                uint32_t* off_pos, uint32_t* off_neg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t pad = (uint32_t)(count * 8);
  int p = 0, m = 0;
  for (int k = 0; k < cnt[i]; ++k) {
    const uint32_t e = pk_cm[(int64_t)k * count + i];
    const uint32_t off = (uint32_t)hop_target(e) * 8u;
    if (e & 0x80000000u)
      off_neg[(int64_t)(m++) * count + i] = off;
    else
      off_pos[(int64_t)(p++) * count + i] = off;
  }
  for (; p < hp; ++p) off_pos[(int64_t)p * count + i] = pad;
  for (; m < hn; ++m) off_neg[(int64_t)m * count + i] = pad;
}

// Sign-sorted target lists (see Tables::tgt_sorted), rows of `stride` entries.
__global__ void k_sorted_targets(const uint32_t* pk_rm, const int32_t* cnt, int64_t count, int width, int stride,
                                 int32_t* tgt, uint32_t* cnt_pn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t* row = pk_rm + i * width;
  int32_t* out = tgt + i * stride;
  const int c = cnt[i];
  int k = 0;
  for (int j = 0; j < c; ++j)
    if (!(row[j] & 0x80000000u)) out[k++] = hop_target(row[j]);
  const int p = k;
  for (int j = 0; j < c; ++j)
    if (row[j] & 0x80000000u) out[k++] = hop_target(row[j]);
  for (; k < stride; ++k) out[k] = 0;
  cnt_pn[i] = (uint32_t)p | ((uint32_t)(c - p) << 16);
}

// Pair form of the sign-sorted lists (see Tables::pair16).
__global__ void k_pair_table(const int32_t* tgt, const uint32_t* cnt_pn, int64_t count, int tstride, int pstride,
                             uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t c = cnt_pn[i];
  const int np = (int)(c & 0xffffu), nn = (int)(c >> 16);
  const int32_t* t = tgt + i * tstride;
  uint32_t* o = out + i * pstride;
  const uint32_t pad = (uint32_t)count;
  int w = 1;
  for (int k = 0; k < np; k += 2) o[w++] = (uint32_t)t[k] | ((k + 1 < np ? (uint32_t)t[k + 1] : pad) << 16);
  const int pp = w - 1;
  for (int k = 0; k < nn; k += 2)
    o[w++] = (uint32_t)t[np + k] | ((k + 1 < nn ? (uint32_t)t[np + k + 1] : pad) << 16);
  const int pt = w - 1;
  for (; w < pstride; ++w) o[w] = pad | (pad << 16);
  o[0] = (uint32_t)pp | ((uint32_t)pt << 8);
}

// Quad form of the sign-sorted lists (see Tables::quad16).
__global__ void k_quad_table(const int32_t* tgt, const uint32_t* cnt_pn, int64_t count, int tstride, int qstride,
                             uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t c = cnt_pn[i];
  const int np = (int)(c & 0xffffu), nn = (int)(c >> 16);
  const int32_t* t = tgt + i * tstride;
  uint32_t* o = out + i * qstride;
  const uint32_t pad = (uint32_t)count;
  int w = 1;
  for (int g = 0; g < 2; ++g) {
    const int n = g == 0 ? np : nn;
    const int32_t* tg = g == 0 ? t : t + np;
    for (int k = 0; k < n; k += 2)
      o[w++] = (uint32_t)tg[k] | ((k + 1 < n ? (uint32_t)tg[k + 1] : pad) << 16);
    if (((n + 1) / 2) & 1) o[w++] = pad | (pad << 16);  // complete the last quad
    if (g == 0) o[0] = (uint32_t)((w - 1) / 2);
  }
  o[0] |= (uint32_t)((w - 1) / 2) << 8;
  for (; w < qstride; ++w) o[w] = pad | (pad << 16);
}

// Per-spin diagonal: sum_i eps_i over set bits, in ascending site order
// (the summation order of weighted_element, cmpy/operators.py:246-250).
__global__ void k_weighted(const uint32_t* states, int64_t count, const double* eps, int num_sites,
                           double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = states[i];
  double acc = 0.0;
  for (int b = 0; b < num_sites; ++b)
    if (s & (1u << b)) acc += eps[b];
  out[i] = acc;
}

__device__ inline double u_term(const HvArgs& a, uint32_t both) {
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

__global__ void k_diagonal(HvArgs a, double* diag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.num_up * a.num_dn) return;
  const int64_t iu = k / a.num_dn, id = k - iu * a.num_dn;
  diag[k] = a.d_up[iu] + a.d_dn[id] + u_term(a, a.st_up[iu] & a.st_dn[id]);
}

// Dense H (row-major, H[target][origin]).  One thread per row: diagonal plus the
// row's hops; every hop list holds both directions, so H is filled exactly.
__global__ void k_dense(HvArgs a, double* h) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t D = a.num_up * a.num_dn;
  if (row >= D) return;
  const int64_t iu = row / a.num_dn, id = row - iu * a.num_dn;
  double* hr = h + row * D;
  for (int64_t c = 0; c < D; ++c) hr[c] = 0.0;
  hr[row] = a.d_up[iu] + a.d_dn[id] + u_term(a, a.st_up[iu] & a.st_dn[id]);
  const int cu = a.up_cnt[iu];
  for (int k = 0; k < cu; ++k) hr[(int64_t)a.up_col[iu * a.up_w + k] * a.num_dn + id] += a.up_val[iu * a.up_w + k];
  const int cd = a.dn_cnt[id];
  for (int k = 0; k < cd; ++k)
    hr[iu * a.num_dn + a.dn_col[(int64_t)k * a.num_dn + id]] += a.dn_val[(int64_t)k * a.num_dn + id];
}

__global__ void k_copy_i32(const int32_t* src, int32_t* dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

HvArgs make_hv_args(const ed_op* op) {
  HvArgs a;
  a.num_up = op->num_up;
  a.num_dn = op->num_dn;
  a.ld = op->num_dn;
  a.st_up = op->st_up;
  a.st_dn = op->st_dn;
  a.d_up = op->d_up;
  a.d_dn = op->d_dn;
  a.up_cnt = op->up.cnt;
  a.up_col = op->up.col_rm;
  a.up_val = op->up.val_rm;
  a.up_w = op->up.width;
  a.dn_cnt = op->dn.cnt;
  a.dn_col = op->dn.col_cm;
  a.dn_val = op->dn.val_cm;
  a.dn_w = op->dn.width;
  a.u_mode = op->u_mode;
  a.u_uniform = op->u_uniform;
  a.u_mask = op->u_mask;
  a.u_site = op->u_site;
  return a;
}

static void free_tables(Tables& t) {
  if (t.cnt) (void)hipFree(t.cnt);
  if (t.col_rm) (void)hipFree(t.col_rm);
  if (t.val_rm) (void)hipFree(t.val_rm);
  if (t.col_cm) (void)hipFree(t.col_cm);
  if (t.val_cm) (void)hipFree(t.val_cm);
  if (t.pk_rm) (void)hipFree(t.pk_rm);
  if (t.pk_cm) (void)hipFree(t.pk_cm);
  if (t.off_pos) (void)hipFree(t.off_pos);
  if (t.off_neg) (void)hipFree(t.off_neg);
  if (t.h16) (void)hipFree(t.h16);
  if (t.h16q) (void)hipFree(t.h16q);
  if (t.hwords_chunk) (void)hipFree(t.hwords_chunk);
  if (t.cmax_pos) (void)hipFree(t.cmax_pos);
  if (t.cmax_neg) (void)hipFree(t.cmax_neg);
  if (t.tgt_sorted) (void)hipFree(t.tgt_sorted);
  if (t.cnt_pn) (void)hipFree(t.cnt_pn);
  if (t.pair16) (void)hipFree(t.pair16);
  if (t.quad16) (void)hipFree(t.quad16);
  t = Tables{};
}

static void free_op(ed_op* op) {
  if (!op) return;
  if (op->st_up) (void)hipFree(op->st_up);
  if (op->st_dn) (void)hipFree(op->st_dn);
  if (op->d_up) (void)hipFree(op->d_up);
  if (op->d_dn) (void)hipFree(op->d_dn);
  if (op->u_site) (void)hipFree(op->u_site);
  if (op->hop_dev) (void)hipFree(op->hop_dev);
  free_tables(op->up);
  free_tables(op->dn);
  delete op;
}

template <typename T>
static int dev_alloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(ED_ERR_NOMEM, "hipMalloc of %lld bytes failed: %s",
                     (long long)(sizeof(T) * count), hipGetErrorString(e));
  }
  return ED_OK;
}


// Pair table for the column pass when every row's pair list fits the unrolled
// maximum; otherwise none is built and the column pass uses the plain lists.
static int build_pair_table(Tables& tab, int64_t count, hipStream_t s) {
  std::vector<uint32_t> cp(count);
  ED_HIP(hipMemcpyAsync(cp.data(), tab.cnt_pn, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  int mx = 0;
  for (uint32_t c : cp) mx = std::max(mx, (int)(((c & 0xffffu) + 1) / 2 + ((c >> 16) + 1) / 2));
  if (mx > kPairMaxPairs) return ED_OK;
  tab.pair_stride = 1 + mx;
  // kPairMaxPairs words of slack: the column pass reads that many after any
  // row's header
  ED_TRY(dev_alloc(&tab.pair16, count * tab.pair_stride + kPairMaxPairs));
  ED_HIP(hipMemsetAsync(tab.pair16 + count * tab.pair_stride, 0, sizeof(uint32_t) * kPairMaxPairs, s));
  hipLaunchKernelGGL(k_pair_table, dim3(grid_for(count, 256)), dim3(256), 0, s, tab.tgt_sorted, tab.cnt_pn, count,
                     tab.tgt_stride, tab.pair_stride, tab.pair16);
  ED_HIP(hipGetLastError());
  // quads
  int mq = 0;
  for (uint32_t c : cp) mq = std::max(mq, (int)(((c & 0xffffu) + 3) / 4 + ((c >> 16) + 3) / 4));
  if (mq <= kQuadMaxQuads) {
    tab.quad_stride = 1 + 2 * mq;
    ED_TRY(dev_alloc(&tab.quad16, count * tab.quad_stride + tab.quad_stride + kQuadWords));
    ED_HIP(hipMemsetAsync(tab.quad16 + count * tab.quad_stride, 0,
                          sizeof(uint32_t) * (tab.quad_stride + kQuadWords), s));
    hipLaunchKernelGGL(k_quad_table, dim3(grid_for(count, 256)), dim3(256), 0, s, tab.tgt_sorted, tab.cnt_pn,
                       count, tab.tgt_stride, tab.quad_stride, tab.quad16);
    ED_HIP(hipGetLastError());
  }
  return ED_OK;
}

// Builds the half-width table (see Tables::h16) on the host from the packed
// column-major table and uploads it.
static int build_half_tables(Tables& tab, int64_t count, int width, hipStream_t s) {
  constexpr int L = kColourLanes;
  std::vector<uint32_t> pk((size_t)count * width);
  std::vector<int32_t> cnt(count);
  ED_HIP(hipMemcpyAsync(pk.data(), tab.pk_cm, sizeof(uint32_t) * pk.size(), hipMemcpyDeviceToHost, s));
  ED_HIP(hipMemcpyAsync(cnt.data(), tab.cnt, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  const int64_t nchunk = (count + 63) / 64;
  // colourings per sign and half-wave: [chunk][half]
  std::vector<ChunkColouring> cp(2 * nchunk), cn(2 * nchunk);
  std::vector<int32_t> words(2 * nchunk);  // per chunk: positive words, total words
  int maxw = 0;
  for (int64_t c = 0; c < nchunk; ++c) {
    for (int h = 0; h < 2; ++h) {
      std::vector<std::vector<int32_t>> lp(L), ln(L);
      for (int l = 0; l < L; ++l) {
        const int64_t i = c * 64 + h * L + l;
        if (i >= count) break;
        for (int k = 0; k < cnt[i]; ++k) {
          const uint32_t e = pk[(size_t)k * count + i];
          ((e & 0x80000000u) ? ln : lp)[l].push_back(hop_target(e));
        }
      }
      cp[2 * c + h] = colour_chunk(lp, count);
      cn[2 * c + h] = colour_chunk(ln, count);
      if (cp[2 * c + h].colours < 0 || cn[2 * c + h].colours < 0)
        return set_error(ED_ERR_INVALID, "hop table entry out of range while ordering the row-pass slots");
    }
    const int kp = std::max(cp[2 * c].colours, cp[2 * c + 1].colours);
    const int kn = std::max(cn[2 * c].colours, cn[2 * c + 1].colours);
    words[2 * c] = (kp + 1) / 2;
    words[2 * c + 1] = (kp + 1) / 2 + (kn + 1) / 2;
    maxw = std::max(maxw, words[2 * c + 1]);
  }
  if (maxw > kHalfMaxWords) return ED_OK;  // row pass uses the 32-bit tables
  tab.hwords = maxw;
  std::vector<uint32_t> out((size_t)std::max(maxw, 1) * count, 0);
  for (int64_t i = 0; i < count; ++i) {
    const int64_t c = i / 64;
    const int half = (int)((i % 64) / L), l = (int)(i % L);
    const uint32_t lane_pad = (uint32_t)(count + ((l - count % L + L) % L));  // zero pad in bank pair l
    const int wp = words[2 * c];
    for (int w = 0; w < maxw; ++w) {
      const ChunkColouring& col = w < wp ? cp[2 * c + half] : cn[2 * c + half];
      const int k0 = 2 * (w < wp ? w : w - wp);
      uint32_t v[2];
      for (int j = 0; j < 2; ++j) {
        const int k = k0 + j;
        v[j] = k < col.colours ? (uint32_t)col.target[(size_t)k * L + l] : lane_pad;
      }
      out[(size_t)w * count + i] = v[0] | (v[1] << 16);
    }
  }
  // lane-contiguous copy: chunk c, group g, lane l, word 4g + j
  tab.hgroups = (maxw + 3) / 4;
  std::vector<uint32_t> outq((size_t)nchunk * tab.hgroups * 256);
  for (int64_t c = 0; c < nchunk; ++c)
    for (int g = 0; g < tab.hgroups; ++g)
      for (int l = 0; l < 64; ++l) {
        const int64_t i = c * 64 + l;
        const uint32_t lane_pad = (uint32_t)(count + ((l % L) - count % L + L) % L);
        for (int j = 0; j < 4; ++j) {
          const int w = 4 * g + j;
          uint32_t v = lane_pad | (lane_pad << 16);
          if (i < count && w < maxw) v = out[(size_t)w * count + i];
          outq[(((size_t)c * tab.hgroups + g) * 64 + l) * 4 + j] = v;
        }
      }
  ED_TRY(dev_alloc(&tab.h16q, (int64_t)outq.size()));
  ED_HIP(hipMemcpyAsync(tab.h16q, outq.data(), sizeof(uint32_t) * outq.size(), hipMemcpyHostToDevice, s));
  ED_TRY(dev_alloc(&tab.h16, (int64_t)out.size()));
  ED_TRY(dev_alloc(&tab.hwords_chunk, 2 * nchunk));
  ED_HIP(hipMemcpyAsync(tab.h16, out.data(), sizeof(uint32_t) * out.size(), hipMemcpyHostToDevice, s));
  ED_HIP(hipMemcpyAsync(tab.hwords_chunk, words.data(), sizeof(int32_t) * words.size(), hipMemcpyHostToDevice, s));
  ED_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  return ED_OK;
}

static int build_spin(ed_op* op, int n, uint32_t** states, double** diag, Tables& tab,
                      const int32_t* d_bonds, const double* d_hops, const double* d_eps,
                      hipStream_t s, int* max_hops) {
  const int64_t count = (int64_t)binom_u(op->num_sites, n);
  ED_TRY(dev_alloc(states, count));
  ED_TRY(dev_alloc(diag, count));
  const int nb = std::max(op->num_bonds, 1);
  tab.width = nb;
  ED_TRY(dev_alloc(&tab.cnt, count));
  ED_TRY(dev_alloc(&tab.col_rm, count * nb));
  ED_TRY(dev_alloc(&tab.val_rm, count * nb));
  ED_TRY(dev_alloc(&tab.col_cm, count * nb));
  ED_TRY(dev_alloc(&tab.val_cm, count * nb));
  if (op->packed) {
    ED_TRY(dev_alloc(&tab.pk_rm, count * nb));
    ED_TRY(dev_alloc(&tab.pk_cm, count * nb));
  }
  hipLaunchKernelGGL(k_unrank, dim3(grid_for(count, 256)), dim3(256), 0, s, *states, count,
                     op->num_sites, n);
  hipLaunchKernelGGL(k_weighted, dim3(grid_for(count, 256)), dim3(256), 0, s, *states, count, d_eps,
                     op->num_sites, *diag);
  if (op->num_bonds > 0) {
    hipLaunchKernelGGL(k_hop_table, dim3(grid_for(count, 256)), dim3(256), 0, s, *states, count,
                       op->width, d_bonds, d_hops, op->num_bonds, tab.cnt, tab.col_rm, tab.val_rm,
                       tab.col_cm, tab.val_cm, tab.pk_rm, tab.pk_cm);
  } else {
    if (op->packed) {
      ED_HIP(hipMemsetAsync(tab.pk_rm, 0, sizeof(uint32_t) * count, s));
      ED_HIP(hipMemsetAsync(tab.pk_cm, 0, sizeof(uint32_t) * count, s));
    }
    ED_HIP(hipMemsetAsync(tab.cnt, 0, sizeof(int32_t) * count, s));
    ED_HIP(hipMemsetAsync(tab.col_rm, 0, sizeof(int32_t) * count, s));
    ED_HIP(hipMemsetAsync(tab.val_rm, 0, sizeof(double) * count, s));
    ED_HIP(hipMemsetAsync(tab.col_cm, 0, sizeof(int32_t) * count, s));
    ED_HIP(hipMemsetAsync(tab.val_cm, 0, sizeof(double) * count, s));
  }
  // sign-split tables for uniform hopping; the row pass addresses a row of
  // doubles plus its zero pad with 32-bit byte offsets
  if (op->packed && op->uniform_hop && op->num_bonds > 0 && count * 8 < (1ll << 31)) {
    const int64_t nchunk = (count + 63) / 64;
    int32_t *np = nullptr, *nn = nullptr;
    int rc = dev_alloc(&np, count);
    if (rc == ED_OK) rc = dev_alloc(&nn, count);
    if (rc == ED_OK) rc = dev_alloc(&tab.cmax_pos, nchunk);
    if (rc == ED_OK) rc = dev_alloc(&tab.cmax_neg, nchunk);
    std::vector<int32_t> cmp(nchunk), cmn(nchunk);
    if (rc == ED_OK) {
      hipLaunchKernelGGL(k_sign_split, dim3(grid_for(count, 256)), dim3(256), 0, s, tab.pk_cm, tab.cnt, count, np,
                         nn);
      hipLaunchKernelGGL(k_chunk_max, dim3(grid_for(nchunk, 256)), dim3(256), 0, s, np, nn, count, tab.cmax_pos,
                         tab.cmax_neg);
      if (hipMemcpyAsync(cmp.data(), tab.cmax_pos, sizeof(int32_t) * nchunk, hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipMemcpyAsync(cmn.data(), tab.cmax_neg, sizeof(int32_t) * nchunk, hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        rc = set_error(ED_ERR_HIP, "copy of hop-count maxima failed");
    }
    if (np) (void)hipFree(np);
    if (nn) (void)hipFree(nn);
    ED_TRY(rc);
    for (int64_t c = 0; c < nchunk; ++c) {
      tab.wpos = std::max(tab.wpos, cmp[c]);
      tab.wneg = std::max(tab.wneg, cmn[c]);
    }
    const int hp8 = (tab.wpos + 7) & ~7, hn8 = (tab.wneg + 7) & ~7;
    ED_TRY(dev_alloc(&tab.off_pos, hp8 * count));
    ED_TRY(dev_alloc(&tab.off_neg, hn8 * count));
    tab.tgt_stride = (nb + 7) & ~7;
    ED_TRY(dev_alloc(&tab.tgt_sorted, count * tab.tgt_stride));
    ED_TRY(dev_alloc(&tab.cnt_pn, count));
    hipLaunchKernelGGL(k_offset_tables, dim3(grid_for(count, 256)), dim3(256), 0, s, tab.pk_cm, tab.cnt, count,
                       hp8, hn8, tab.off_pos, tab.off_neg);
    // half-width tables where one row (with its pads) fits an LDS plane; sizes
    // from 4096 states up, where rows are too long for several per block
    if (count <= kHalfMaxStates && count >= 4096) ED_TRY(build_half_tables(tab, count, nb, s));
    hipLaunchKernelGGL(k_sorted_targets, dim3(grid_for(count, 256)), dim3(256), 0, s, tab.pk_rm, tab.cnt, count, nb,
                       tab.tgt_stride, tab.tgt_sorted, tab.cnt_pn);
    if (count < 65536) ED_TRY(build_pair_table(tab, count, s));
  }
  ED_HIP(hipGetLastError());
  // host-side maximum hop count (used by kernels that size per-lane lists)
  std::vector<int32_t> cnt(count);
  ED_HIP(hipMemcpyAsync(cnt.data(), tab.cnt, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  int mx = 0;
  for (int32_t c : cnt) mx = std::max(mx, (int)c);
  *max_hops = mx;
  return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" {

const char* ed_last_error(void) { return g_error.c_str(); }

int ed_version(void) { return 10000; }

int ed_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int64_t ed_binom(int n, int k) {
  if (n < 0 || k < 0 || k > n || n > 64) return 0;
  if (n <= kMaxSites) return (int64_t)kHostBinom.v[n][k];
  // 33..64: multiplicative formula (fits int64 for every k when n <= 62)
  __int128 r = 1;
  if (k > n - k) k = n - k;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return (int64_t)r;
}

int ed_sector_states_host(int num_sites, int n, uint32_t* out) {
  clear_error();
  ED_REQUIRE(num_sites >= 0 && num_sites <= ED_MAX_SITES, "num_sites=%d outside [0, %d]", num_sites,
             ED_MAX_SITES);
  ED_REQUIRE(n >= 0 && n <= num_sites, "filling n=%d outside [0, %d]", n, num_sites);
  ED_REQUIRE(out != nullptr, "out is NULL");
  const int64_t count = (int64_t)binom_u(num_sites, n);
  if (n == 0) {
    out[0] = 0;
    return ED_OK;
  }
  // Gosper's hack: next larger integer with the same popcount.
  uint32_t s = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
  for (int64_t i = 0; i < count; ++i) {
    out[i] = s;
    if (i + 1 == count) break;
    const uint32_t c = s & (0u - s);
    const uint32_t r = s + c;
    s = (((r ^ s) >> 2) / c) | r;
  }
  return ED_OK;
}

int ed_sector_states(int num_sites, int n, uint32_t* out_dev, void* stream) {
  clear_error();
  ED_REQUIRE(num_sites >= 0 && num_sites <= ED_MAX_SITES, "num_sites=%d outside [0, %d]", num_sites,
             ED_MAX_SITES);
  ED_REQUIRE(n >= 0 && n <= num_sites, "filling n=%d outside [0, %d]", n, num_sites);
  ED_REQUIRE(out_dev != nullptr, "out_dev is NULL");
  const int64_t count = (int64_t)binom_u(num_sites, n);
  hipLaunchKernelGGL(k_unrank, dim3(grid_for(count, 256)), dim3(256), 0, as_stream(stream), out_dev,
                     count, num_sites, n);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

int ed_sector_rank(int num_sites, int n, const uint32_t* states_dev, int64_t count, int64_t* ranks_dev,
                   void* stream) {
  clear_error();
  ED_REQUIRE(num_sites >= 0 && num_sites <= ED_MAX_SITES, "num_sites=%d outside [0, %d]", num_sites,
             ED_MAX_SITES);
  ED_REQUIRE(count >= 0, "count < 0");
  if (count == 0) return ED_OK;
  ED_REQUIRE(states_dev && ranks_dev, "NULL buffer");
  hipLaunchKernelGGL(k_rank, dim3(grid_for(count, 256)), dim3(256), 0, as_stream(stream), states_dev,
                     count, num_sites, n, ranks_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

int ed_hop_terms(int num_sites, int n, int width, int site1, int site2, double hop,
                 int64_t* target_dev, double* value_dev, void* stream) {
  clear_error();
  ED_REQUIRE(num_sites >= 1 && num_sites <= ED_MAX_SITES, "num_sites=%d outside [1, %d]", num_sites,
             ED_MAX_SITES);
  ED_REQUIRE(n >= 0 && n <= num_sites, "filling n=%d outside [0, %d]", n, num_sites);
  ED_REQUIRE(site1 >= 0 && site1 < site2 && site2 < num_sites,
             "bond (%d, %d) must satisfy 0 <= site1 < site2 < num_sites", site1, site2);
  ED_REQUIRE(width >= 0 && width <= 32, "width=%d outside [0, 32]", width);
  ED_REQUIRE(target_dev && value_dev, "NULL buffer");
  const int64_t count = (int64_t)binom_u(num_sites, n);
  hipLaunchKernelGGL(k_hop_terms, dim3(grid_for(count, 256)), dim3(256), 0, as_stream(stream), count,
                     num_sites, n, width, site1, site2, hop, target_dev, value_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

int ed_op_create(ed_op** out, const ed_model_desc* d, void* stream) {
  clear_error();
  ED_REQUIRE(out != nullptr && d != nullptr, "NULL argument");
  *out = nullptr;
  ED_REQUIRE(d->num_sites >= 1 && d->num_sites <= ED_MAX_SITES, "num_sites=%d outside [1, %d]",
             d->num_sites, ED_MAX_SITES);
  ED_REQUIRE(d->n_up >= 0 && d->n_up <= d->num_sites && d->n_dn >= 0 && d->n_dn <= d->num_sites,
             "sector (%d, %d) outside [0, %d]", d->n_up, d->n_dn, d->num_sites);
  ED_REQUIRE(d->num_bonds >= 0 && (d->num_bonds == 0 || (d->bonds && d->hop)), "bad bond arrays");
  ED_REQUIRE(d->eps && d->u, "eps/u are NULL");
  ED_REQUIRE(d->width >= 0 && d->width <= 32, "width=%d outside [0, 32]", d->width);
  for (int b = 0; b < d->num_bonds; ++b) {
    const int i = d->bonds[2 * b], j = d->bonds[2 * b + 1];
    ED_REQUIRE(0 <= i && i < j && j < d->num_sites,
               "bond %d = (%d, %d) must satisfy 0 <= i < j < num_sites", b, i, j);
    ED_REQUIRE(std::isfinite(d->hop[b]), "hop[%d] is not finite", b);
  }
  for (int i = 0; i < d->num_sites; ++i)
    ED_REQUIRE(std::isfinite(d->eps[i]) && std::isfinite(d->u[i]), "eps/u[%d] not finite", i);
  const int64_t nu = (int64_t)binom_u(d->num_sites, d->n_up);
  const int64_t nd = (int64_t)binom_u(d->num_sites, d->n_dn);
  ED_REQUIRE(nu < (1ll << 31) && nd < (1ll << 31), "spin sector too large");

  ed_op* op = new ed_op();
  op->num_sites = d->num_sites;
  op->n_up = d->n_up;
  op->n_dn = d->n_dn;
  op->num_bonds = d->num_bonds;
  op->width = d->width;
  op->num_up = nu;
  op->num_dn = nd;
  op->size = nu * nd;
  op->bonds.assign(d->bonds, d->bonds + 2 * d->num_bonds);
  op->hop.assign(d->hop, d->hop + d->num_bonds);
  op->eps.assign(d->eps, d->eps + d->num_sites);
  op->u.assign(d->u, d->u + d->num_sites);
  bool uniform = true, any = false;
  for (int i = 0; i < d->num_sites; ++i) {
    if (d->u[i] != d->u[0]) uniform = false;
    if (d->u[i] != 0.0) {
      any = true;
      op->u_mask |= 1u << i;
    }
  }
  op->u_mode = !any ? 0 : (uniform ? 1 : 2);
  op->u_uniform = d->u[0];
  (void)hipGetDevice(&op->device);
  op->packed = nu < (1ll << 24) && nd < (1ll << 24) && d->num_bonds <= 127;
  op->uniform_hop = true;
  for (int b = 1; b < d->num_bonds; ++b)
    if (d->hop[b] != d->hop[0]) op->uniform_hop = false;

  hipStream_t s = as_stream(stream);
  int32_t* d_bonds = nullptr;
  double* d_hops = nullptr;
  double* d_eps = nullptr;
  int rc = ED_OK;
  do {
    if ((rc = dev_alloc(&d_bonds, 2 * std::max(d->num_bonds, 1))) != ED_OK) break;
    if ((rc = dev_alloc(&d_hops, std::max(d->num_bonds, 1))) != ED_OK) break;
    if ((rc = dev_alloc(&d_eps, d->num_sites)) != ED_OK) break;
    if ((rc = dev_alloc(&op->u_site, d->num_sites)) != ED_OK) break;
    if (d->num_bonds > 0) {
      if (hipMemcpyAsync(d_bonds, d->bonds, sizeof(int32_t) * 2 * d->num_bonds, hipMemcpyHostToDevice, s) !=
              hipSuccess ||
          hipMemcpyAsync(d_hops, d->hop, sizeof(double) * d->num_bonds, hipMemcpyHostToDevice, s) !=
              hipSuccess) {
        rc = set_error(ED_ERR_HIP, "copying bonds to the device failed");
        break;
      }
    }
    if (hipMemcpyAsync(d_eps, d->eps, sizeof(double) * d->num_sites, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(op->u_site, d->u, sizeof(double) * d->num_sites, hipMemcpyHostToDevice, s) !=
            hipSuccess) {
      rc = set_error(ED_ERR_HIP, "copying site arrays to the device failed");
      break;
    }
    if ((rc = build_spin(op, d->n_up, &op->st_up, &op->d_up, op->up, d_bonds, d_hops, d_eps, s,
                         &op->max_hops_up)) != ED_OK)
      break;
    if ((rc = build_spin(op, d->n_dn, &op->st_dn, &op->d_dn, op->dn, d_bonds, d_hops, d_eps, s,
                         &op->max_hops_dn)) != ED_OK)
      break;
  } while (0);
  if (d_bonds) (void)hipFree(d_bonds);
  op->hop_dev = d_hops;  // kept: packed hop entries index it
  if (d_eps) (void)hipFree(d_eps);
  if (rc != ED_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return ED_OK;
}

int ed_op_destroy(ed_op* op) {
  clear_error();
  free_op(op);
  return ED_OK;
}

int ed_op_shape(const ed_op* op, int64_t* num_up, int64_t* num_dn) {
  clear_error();
  ED_REQUIRE(op != nullptr, "op is NULL");
  if (num_up) *num_up = op->num_up;
  if (num_dn) *num_dn = op->num_dn;
  return ED_OK;
}

int ed_op_diagonal(ed_op* op, double* diag_dev, void* stream) {
  clear_error();
  ED_REQUIRE(op && diag_dev, "NULL argument");
  HvArgs a = make_hv_args(op);
  hipLaunchKernelGGL(k_diagonal, dim3(grid_for(op->size, 256)), dim3(256), 0, as_stream(stream), a,
                     diag_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

int ed_op_dense(ed_op* op, double* h_dev, void* stream) {
  clear_error();
  ED_REQUIRE(op && h_dev, "NULL argument");
  ED_REQUIRE(op->size <= 32768, "dense H of dimension %lld is too large", (long long)op->size);
  HvArgs a = make_hv_args(op);
  hipLaunchKernelGGL(k_dense, dim3(grid_for(op->size, 64)), dim3(64), 0, as_stream(stream), a, h_dev);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

int ed_op_hop_table(const ed_op* op, int sigma, int32_t* max_hops, int32_t* cnt_dev, int32_t* col_dev,
                    double* val_dev, void* stream) {
  clear_error();
  ED_REQUIRE(op != nullptr, "op is NULL");
  ED_REQUIRE(sigma == ED_UP || sigma == ED_DN, "sigma must be ED_UP or ED_DN");
  const Tables& t = sigma == ED_UP ? op->up : op->dn;
  const int64_t count = sigma == ED_UP ? op->num_up : op->num_dn;
  if (max_hops) *max_hops = t.width;
  if (!cnt_dev && !col_dev && !val_dev) return ED_OK;
  ED_REQUIRE(cnt_dev && col_dev && val_dev, "all of cnt/col/val must be given");
  hipStream_t s = as_stream(stream);
  ED_HIP(hipMemcpyAsync(cnt_dev, t.cnt, sizeof(int32_t) * count, hipMemcpyDeviceToDevice, s));
  ED_HIP(hipMemcpyAsync(col_dev, t.col_rm, sizeof(int32_t) * count * t.width, hipMemcpyDeviceToDevice, s));
  ED_HIP(hipMemcpyAsync(val_dev, t.val_rm, sizeof(double) * count * t.width, hipMemcpyDeviceToDevice, s));
  return ED_OK;
}

}  // extern "C"
