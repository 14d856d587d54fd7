// Internal declarations shared by the libcmpy_amd translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cmpy_amd.h"

namespace ed {

// ---------------------------------------------------------------------------
// Errors: every C entry point returns a status and records a thread-local
// message; nothing throws across the ABI.
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
void clear_error();

#define ED_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::ed::set_error(ED_ERR_HIP, "%s failed: %s (%s:%d)", #call,          \
                             hipGetErrorString(e_), __FILE__, __LINE__);          \
  } while (0)

#define ED_TRY(call)              \
  do {                            \
    int s_ = (call);              \
    if (s_ != ED_OK) return s_;   \
  } while (0)

#define ED_REQUIRE(cond, ...)                                          \
  do {                                                                 \
    if (!(cond)) return ::ed::set_error(ED_ERR_INVALID, __VA_ARGS__);  \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Binomial coefficients, usable on host and device (constant memory).
// ---------------------------------------------------------------------------
constexpr int kMaxSites = 32;
struct BinomTable {
  uint64_t v[kMaxSites + 1][kMaxSites + 1];
};
constexpr BinomTable make_binom() {
  BinomTable t{};
  for (int n = 0; n <= kMaxSites; ++n) {
    t.v[n][0] = 1;
    for (int k = 1; k <= n; ++k) t.v[n][k] = t.v[n - 1][k - 1] + (k <= n - 1 ? t.v[n - 1][k] : 0);
  }
  return t;
}
constexpr BinomTable kHostBinom = make_binom();
static __constant__ BinomTable kDevBinom = make_binom();

__host__ __device__ inline uint64_t binom_u(int n, int k) {
  if (n < 0 || k < 0 || k > n || n > kMaxSites) return 0;
#ifdef __HIP_DEVICE_COMPILE__
  return kDevBinom.v[n][k];
#else
  return kHostBinom.v[n][k];
#endif
}

// Combinadic rank of a bit pattern: r(s) = sum_k C(p_k, k), k 1-based over the
// set bits p_1 < p_2 < ...; equals the position of s in the ascending list of
// states with the same popcount.
__host__ __device__ inline int64_t comb_rank(uint32_t s) {
  int64_t r = 0;
  int k = 0;
  while (s) {
    const int p = __builtin_ctz(s);
    ++k;
    r += (int64_t)binom_u(p, k);
    s &= s - 1;
  }
  return r;
}

// Inverse of comb_rank within the (num_sites, n) sector: greedy from the top.
__host__ __device__ inline uint32_t comb_unrank(int64_t r, int num_sites, int n) {
  uint32_t s = 0;
  int p = num_sites - 1;
  for (int k = n; k >= 1; --k) {
    while (p >= 0 && (int64_t)binom_u(p, k) > r) --p;
    s |= 1u << p;
    r -= (int64_t)binom_u(p, k);
    --p;
  }
  return s;
}

__host__ __device__ inline int popc32(uint32_t v) { return __builtin_popcount(v); }

// Hopping sign (-1)^popcount(state & bits strictly between site1 and site2,
// restricted to the low `width` bits): cmpy/operators.py:253-273, 425-433.
__host__ __device__ inline int hop_sign(uint32_t state, int width, int site1, int site2) {
  if (width <= 0) return 1;
  uint32_t between = 0;
  if (site2 - site1 > 1) between = ((site2 >= 32 ? 0u : (1u << site2)) - (1u << (site1 + 1)));
  const uint32_t low = width >= 32 ? 0xffffffffu : ((1u << width) - 1u);
  return (popc32(state & between & low) & 1) ? -1 : 1;
}

// ---------------------------------------------------------------------------
// Sector operator handle.
// ---------------------------------------------------------------------------
struct Tables {
  // ELL hop table for one spin species, valid entries first in bond order.
  int32_t width = 0;         // slots per state (= number of bonds)
  int32_t* cnt = nullptr;    // [D]
  int32_t* col_rm = nullptr; // [D][width]
  double* val_rm = nullptr;  // [D][width]
  int32_t* col_cm = nullptr; // [width][D]
  double* val_cm = nullptr;  // [width][D]
  // Packed entries: target (bits 0-23) | bond (bits 24-30) | sign (bit 31);
  // the value is sign * hop[bond].  Present when D < 2^24 and bonds <= 127.
  uint32_t* pk_rm = nullptr; // [D][width]
  uint32_t* pk_cm = nullptr; // [width][D]
  // Sign-split offset tables (uniform hopping, row pass): state i's hops with
  // sign +1 are off_pos[k][i] = target * 8 (a byte offset into a row of
  // doubles), k < its positive count; the remaining slots up to wpos hold D * 8,
  // the offset of a zero pad stored after the row.  off_neg likewise for the
  // hops with sign -1.  cmax_pos/cmax_neg: per 64-state chunk, the largest
  // count; wpos/wneg: the largest over all chunks.
  uint32_t* off_pos = nullptr;  // [roundup(wpos, 8)][D]; slots past wpos are pads
  uint32_t* off_neg = nullptr;  // [roundup(wneg, 8)][D]
  int32_t* cmax_pos = nullptr;  // [ceil(D / 64)]
  int32_t* cmax_neg = nullptr;
  int32_t wpos = 0, wneg = 0;
  // Half-width form of the same lists for D <= kHalfMaxStates: 16-bit entries
  // p, the index of a double in a staged row, where p is a target or one of 32
  // zero pads D .. D+31; slots 2w and 2w+1 sit in the low and high half of
  // word w.  In each 64-state chunk the first hwords_chunk[2c] words hold the
  // hops of sign +1 and the words up to hwords_chunk[2c+1] those of sign -1;
  // within each 32-state half-wave the slots of either list are ordered so
  // that every slot reads 32 distinct LDS bank pairs (p mod 32).  [hwords][D].
  uint32_t* h16 = nullptr;
  int32_t* hwords_chunk = nullptr;  // [2 * ceil(D / 64)]
  int32_t hwords = 0;
  // The same words with each lane's words of a chunk contiguous, four per
  // 16-byte group: chunk c's group g is 64 lanes x 4 words (1 KiB) at
  // (c * hgroups + g) KiB; words past a chunk's count are the lane's pads.
  uint32_t* h16q = nullptr;
  int32_t hgroups = 0;  // groups per chunk, ceil(hwords / 4)
  // Sign-sorted target lists (uniform hopping, column pass): row i holds its
  // targets with sign +1 first, then those with sign -1; cnt_pn[i] = number
  // positive | number negative << 16.  [D][tgt_stride], tgt_stride = width
  // rounded up to 8, unused entries 0.
  int32_t* tgt_sorted = nullptr;
  uint32_t* cnt_pn = nullptr;
  int32_t tgt_stride = 0;
  // Pair form of the sign-sorted lists (column pass over 32-column tiles),
  // for D < 65536: row i is pair_stride words; word 0 = number of sign +1
  // pairs | total pairs << 8, then the pairs, each two 16-bit targets (low
  // half first).  The +1 targets fill the first pairs and the -1 targets the
  // rest; a group of odd length ends in the pad entry D.  Words past the
  // row's pairs hold pads, and kPairMaxPairs zero words follow the last row.
  uint32_t* pair16 = nullptr;
  int32_t pair_stride = 0;
  // Quad form (column pass with 16-byte loads): row i is quad_stride words;
  // word 0 = number of sign +1 quads | total quads << 8, then the quads, each
  // four 16-bit targets in two words (entry 4m + q is half q & 1 of word
  // 2m + (q >> 1)).  Each sign group is padded to a multiple of four with
  // the pad entry D; words past the row's quads hold pads, and quad_stride
  // words of slack follow the last row.
  uint32_t* quad16 = nullptr;
  int32_t quad_stride = 0;
};

constexpr int kPairMaxPairs = 16;  // longest pair list the column pass unrolls
constexpr int kQuadMaxQuads = 12;  // longest quad list the column pass unrolls
constexpr int kQuadWords = 1 + 2 * kQuadMaxQuads;

// Largest spin-sector size with half-width row-pass tables: entries and the 32
// pads D .. D+31 must fit 16 bits, and the staged row with its pads must fit
// the LDS of one workgroup (checked when the kernel is chosen).
constexpr int64_t kHalfMaxStates = 65535 - 32;
// Vectors of at least this many states (128 MiB) are streamed with
// nontemporal hints where a kernel reads or writes them once per step; below
// it they stay cache-resident between the launches of a step.
constexpr int64_t kStreamStates = int64_t(1) << 24;
constexpr int kHalfMaxWords = 24;  // longest per-chunk word list the row pass unrolls

__host__ __device__ inline uint32_t pack_hop(int32_t target, int bond, int sign) {
  return (uint32_t)target | ((uint32_t)bond << 24) | (sign < 0 ? 0x80000000u : 0u);
}
__host__ __device__ inline int32_t hop_target(uint32_t e) { return (int32_t)(e & 0x00ffffffu); }
__host__ __device__ inline int hop_bond(uint32_t e) { return (int)((e >> 24) & 0x7fu); }
__host__ __device__ inline double hop_value(uint32_t e, const double* hop) {
  const double v = hop[hop_bond(e)];
  return (e & 0x80000000u) ? -v : v;
}

}  // namespace ed

struct ed_op {
  int device = 0;
  int num_sites = 0, n_up = 0, n_dn = 0, num_bonds = 0, width = 0;
  int64_t num_up = 0, num_dn = 0, size = 0;
  std::vector<int32_t> bonds;
  std::vector<double> hop, eps, u;
  int u_mode = 0;        // 0: no interaction, 1: uniform (popcount), 2: per site
  double u_uniform = 0.0;
  uint32_t u_mask = 0;
  // device data
  uint32_t* st_up = nullptr;
  uint32_t* st_dn = nullptr;
  double* d_up = nullptr;
  double* d_dn = nullptr;
  double* u_site = nullptr;  // [num_sites]
  ed::Tables up, dn;
  int max_hops_up = 0, max_hops_dn = 0;
  double* hop_dev = nullptr;  // [num_bonds]
  bool packed = false;        // packed tables present
  bool uniform_hop = false;   // every bond has the same hop value
  int hv_mode = -1;           // H.v kernel choice (-1 = automatic)
};

namespace ed {

// Plain matvec arguments (device pointers from the op).
struct HvArgs {
  int64_t num_up, num_dn;
  int64_t ld;  // row stride of x and y in doubles (>= num_dn)
  const uint32_t* st_up;
  const uint32_t* st_dn;
  const double* d_up;
  const double* d_dn;
  const int32_t* up_cnt;
  const int32_t* up_col;  // row-major [num_up][up_w]
  const double* up_val;
  int32_t up_w;
  const int32_t* dn_cnt;
  const int32_t* dn_col;  // column-major [dn_w][num_dn]
  const double* dn_val;
  int32_t dn_w;
  int32_t u_mode;
  double u_uniform;
  uint32_t u_mask;
  const double* u_site;
};

HvArgs make_hv_args(const ed_op* op);

// Lanczos epilogue: y = s_j * H x - c_j * xprev, partial <y, x>.
struct HvEpilogue {
  const double* xprev = nullptr;  // may be null
  const double* sx = nullptr;     // device scalar: scale of x (null = 1)
  const double* cprev = nullptr;  // device scalar: coefficient of xprev (null = 0)
  double* partial = nullptr;      // per-block partial dots (null = none)
  // Recurrence with known coefficients (the eigenvector's second pass):
  // y = H (sx x) - cself sx x - cprev xprev, and psi += psi_y sx x over the
  // sector entries (row padding untouched).  Applied by the row pass.
  const double* cself = nullptr;  // device scalar (null = 0)
  double* psi = nullptr;          // null = none
  double psi_y = 0.0;
};

// Launch helpers implemented in ed_matvec.hip.  Vectors are stored row by
// row at a stride of ld >= num_dn doubles; the H.v writes zeros into y's
// entries num_dn .. ld-1 of every row and never reads those of x.
int launch_matvec(const ed_op* op, const double* x, double* y, int64_t ld, const HvEpilogue& epi,
                  hipStream_t stream, int* num_partials);
int matvec_num_partials(const ed_op* op, int64_t ld);
// Row stride the library's own vectors use: num_dn rounded up to 16 doubles
// (128-byte rows) where the two-pass sweep runs, num_dn otherwise.
int64_t vector_ld(const ed_op* op);

// Deterministic reductions (ed_lanczos.hip).
int launch_reduce(const double* partial, int n, double* out, int op_kind, hipStream_t stream);

}  // namespace ed
