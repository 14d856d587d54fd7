// Batched Anderson-impurity solver: many SingleImpurityAndersonModel instances
// of uneven bath size in one persistent launch.
//
// No reference counterpart: cmpy solves one model per call
// (cmpy/models/anderson.py:163-164 -> cmpy/exactdiag.py:137-245).  Per model
// this computes what the single-model path does -- ground state of the
// (n_up, n_dn) sector by Lanczos (cmpy/exactdiag.py:24-43), then the impurity
// Green's function G(i w_n) = particle + hole continued fraction from
// c+_{0,up} psi0 and c_{0,up} psi0 (the T = 0 limit of
// cmpy/exactdiag.py:110-134) -- with the model's Hamiltonian built as
// SingleImpurityAndersonModel does (cmpy/models/anderson.py:147-158): site 0
// is the impurity (eps_imp, U), sites 1..nb the bath (eps_k), bonds (0, k)
// with hopping V_k, hopping signs of width `sign_width` (0 = the reference's
// all-positive signs).
//
// Work layout.  Models are split by sector size into two queues, both sorted
// by decreasing cost.  Workgroups [0, wave_blocks) run "wave mode": each wave
// takes one small model at a time; the other workgroups run "block mode": the
// whole workgroup takes one large model.  Every group pops tasks from its
// queue with one atomic add until the queue is empty, so every wave reaches
// the kernel's end.  A solve is two persistent launches over the same queues:
// the ground-state tasks (one per model), then the continued-fraction tasks
// (particle and hole of each model, two tasks), so a model's two fractions
// run on two groups at once; a last small kernel sums G from each model's
// coefficient record.  All vectors and spin tables of a group live in its slot
// of a global workspace (L1/L2-resident for these sizes); Lanczos scalars are
// reduced in a fixed order, so results are deterministic for a given
// configuration.  The ground-state Lanczos vectors of every model are kept
// in a store of their own ((max_iter + 1) D doubles per model), so psi is
// one combination of them instead of a second Lanczos pass; a batch whose
// store would not fit in 60% of the free device memory is solved in chunks
// of models, one launch each.
#include "ed_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <numeric>

namespace ed {
namespace batch {

constexpr int kMaxBath = ED_BATCH_MAX_BATH;
constexpr int kMaxSites = kMaxBath + 1;
constexpr int kBlock = 1024;      // threads per workgroup of the backward kernel
// Solve kernels, per phase (ground-state tasks, continued-fraction tasks):
// threads per workgroup (the VGPR budget is 512 / waves per SIMD) and whether
// block-mode sweeps take the wave-private row path.  Measured on the
// 512-model batch (bath sizes 5..9): the ground-state phase is fastest with
// 16 waves and workgroup rows (the wave-private path needs more than the 128
// VGPRs of 16 waves: 27-30 ms against 32-40 ms on 8 waves), the fraction phase
// with 12 waves and wave-private rows (54-57 ms against 70 ms).
constexpr int kBlockGround = 1024;
constexpr bool kFastGround = true;
constexpr int kBlockFrac = 768;
constexpr bool kFastFrac = true;
constexpr int kBlockSolveMax = kBlockGround > kBlockFrac ? kBlockGround : kBlockFrac;

struct ModelParams {
  int32_t nb, n_up, n_dn, pad;
  double u, eps_imp;  // eps_imp already minus mu
  double eps[kMaxBath];
  double v[kMaxBath];
};

struct Config {
  int64_t num_models;
  const ModelParams* params;
  const int32_t* queue_wave;  // model indices, decreasing cost
  const int32_t* queue_block;
  int32_t n_wave, n_block;
  int32_t* counters;          // [4]: next position in each queue, ground-state then fraction tasks
  int32_t wave_blocks;        // workgroups [0, wave_blocks) run wave mode
  // slot workspace
  char* ws_wave;
  char* ws_block;
  int64_t slot_bytes_wave, slot_bytes_block;
  int32_t list_cap_wave, list_cap_block;   // spin-list capacity (states)
  int64_t vec_cap_wave, vec_cap_block;     // vector capacity (doubles)
  // algorithm
  int32_t max_iter, cf_steps, check_every, sign_width;
  int32_t staged;    // block mode: 1 wave-private row staging (else workgroup rows), 2 workgroup rows, 0 none
  double tol, cf_tol;
  uint64_t seed;
  const double* wn;  // [nw] Matsubara frequencies (imaginary parts)
  int32_t nw;
  double wmin;       // smallest |w_n|: where the continued fraction converges last
  // outputs
  double* e0;     // [num_models]
  double* g;      // [num_models][nw][2]
  int32_t* iters; // [num_models][3]: ground-state, particle and hole Lanczos steps
  double* resid;  // [num_models][2]: Ritz residual, explicit residual
  // per model coefficient record [a+ | b+ | a- | b- (kcap each) | e0, n2+,
  // n2-, m+, m-] (kcap = cf_steps); kept for the adjoint backward (keep != 0):
  // per model psi [D] and the continued-fraction Krylov vectors state-major
  // Qt[D_part][kcap] (particle, hole)
  int32_t keep, kcap;
  const int64_t* off_psi;  // [num_models]
  const int64_t* off_q;    // [num_models][2]
  double* store_psi;
  double* store_q;
  double* store_coef;      // [num_models][4 * kcap + 8]
  double* gs_store;        // ground-state Krylov vectors, (max_iter + 1) D per model
  const int64_t* off_gs;   // [num_models]
};

// ---------------------------------------------------------------------------
// Group: one wave (wave mode) or the whole workgroup (block mode).
// ---------------------------------------------------------------------------
using lds_f64 = __attribute__((address_space(3))) double;
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using lds_i32 = __attribute__((address_space(3))) int32_t;

constexpr int kLdsChunks = 64;  // 64-state chunks of the dn sector the LDS cache covers
// LDS sizes: staged rows (solve kernel: one row of at most 256 columns per
// wave, 4096 doubles; backward kernel: 2048 doubles, one row of at most 128
// columns per wave or two buffers of R rows of at most 64 * 16 / R columns),
// cached dn and up hop slots (12 bytes each) and up-state counts
constexpr int kLdsRows = 2048;
constexpr int kLdsRowsSolve = 4096;
constexpr int kLdsDnSolve = 4096;   // solve kernel: 48 KiB
constexpr int kLdsUpSolve = 4096;   // 48 KiB
constexpr int kLdsUpRowsSolve = 4096;
constexpr int kLdsDiagSolve = 256;  // dn-sector diagonal and states: 3 KiB
constexpr int kLdsDnBwd = 2048;     // backward kernel (its LDS also holds a W chunk): 24 KiB
constexpr int kLdsUpBwd = 2048;     // 24 KiB
constexpr int kLdsUpRowsBwd = 2048;

// LDS working set of a block-mode group (null in wave mode): rows of u staged
// for the dn-hop gathers, and the dn sector's hop lists transposed
// ([slot][state]), slots past a state's count holding target 0 and value 0.
// meta[0] = states cached (0: none), meta[1] = slots per state, meta[2 + c] =
// largest count in 64-state chunk c.
struct BlockLds {
  lds_f64* rows;
  int rows_cap;  // doubles
  lds_u32* tgt;
  lds_f64* val;
  int dn_cap;    // slots
  lds_i32* meta;
  // the up sector's hop lists, row-major ([state][slot], counts in ucnt);
  // umeta[0] = states cached (0: none), umeta[1] = slots per state,
  // umeta[2..3] = the cached list's address (which sector it is)
  lds_u32* utgt;
  lds_f64* uval;
  lds_i32* ucnt;
  int u_cap;     // slots
  int u_rows;    // states
  lds_i32* umeta;
  // the dn sector's diagonal and states (filled with the dn hop cache when
  // the sector has at most ddn_cap states; solve kernel only)
  lds_f64* ddn = nullptr;
  lds_u32* sdn = nullptr;
  int ddn_cap = 0;
};

template <bool BLOCK>
struct Group {
  int rank, size;
  double* red;  // LDS: 2 * 16 doubles (block mode)
  BlockLds lds{};
  bool has_lds = false;
  bool wave_rows = false;  // block mode: wave-private row staging where the dn sector fits

  __device__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (BLOCK)
      __syncthreads();
    else
      __builtin_amdgcn_wave_barrier();
  }
  // Sum over the group, the same value on every member, fixed order.
  __device__ double sum(double v) const {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (!BLOCK) return v;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
  }
  __device__ double max(double v) const {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if (!BLOCK) return v;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
  }
};

// ---------------------------------------------------------------------------
// Per-group slot of the workspace: spin lists of the three up sectors
// (n_up - 1, n_up, n_up + 1) and the dn sector, four vectors and the
// coefficient arrays.  Only the base and the capacities are kept; every
// pointer is recomputed from them when used (cheap scalar arithmetic, and it
// keeps the solver's live register set small).
// ---------------------------------------------------------------------------
struct SpinList {
  const uint32_t* st;
  const double* dg;
  const int32_t* cnt;
  const int32_t* tgt;  // [n][kMaxBath] hop target | bond << 16 | sign << 31
  const double* val;   // [n][kMaxBath] hop values sign * V_k
  int32_t n;
};

__host__ __device__ inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct SlotRef {
  char* base;
  int32_t lcap;   // spin-list capacity (states)
  int64_t vcap;   // vector capacity (doubles)
  int32_t miter;  // max_iter
  int32_t csteps; // cf_steps

  // byte sizes of the pieces, each rounded to 256
  __host__ __device__ int64_t list_bytes() const {
    return align_up(4ll * lcap, 256) + align_up(8ll * lcap, 256) + align_up(4ll * lcap, 256) +
           align_up(4ll * lcap * kMaxBath, 256) + align_up(8ll * lcap * kMaxBath, 256);
  }
  __host__ __device__ int64_t vec_bytes() const { return align_up(8 * vcap, 256); }
  __host__ __device__ int64_t coef_bytes(int32_t n) const { return align_up(8ll * (n + 2), 256); }
  __host__ __device__ int64_t bytes() const {
    return 4 * list_bytes() + 4 * vec_bytes() + 4 * coef_bytes(miter) + 4 * coef_bytes(csteps) + 256;
  }
  // list l: 0..2 up sectors n_up-1, n_up, n_up+1; 3 dn sector
  __device__ char* list_base(int l) const { return base + l * list_bytes(); }
  __device__ uint32_t* st(int l) const { return (uint32_t*)list_base(l); }
  __device__ double* dg(int l) const { return (double*)(list_base(l) + align_up(4ll * lcap, 256)); }
  __device__ int32_t* cnt(int l) const {
    return (int32_t*)(list_base(l) + align_up(4ll * lcap, 256) + align_up(8ll * lcap, 256));
  }
  __device__ int32_t* tgt(int l) const { return (int32_t*)((char*)cnt(l) + align_up(4ll * lcap, 256)); }
  __device__ double* val(int l) const { return (double*)((char*)tgt(l) + align_up(4ll * lcap * kMaxBath, 256)); }
  // vectors 0..3: uprev, u, w, psi
  __device__ double* vec(int i) const { return (double*)(base + 4 * list_bytes() + i * vec_bytes()); }
  __device__ char* coef_base() const { return base + 4 * list_bytes() + 4 * vec_bytes(); }
  __device__ double* alpha() const { return (double*)coef_base(); }
  __device__ double* beta() const { return (double*)(coef_base() + coef_bytes(miter)); }
  __device__ double* sv() const { return (double*)(coef_base() + 2 * coef_bytes(miter)); }
  __device__ double* tc() const { return (double*)(coef_base() + 3 * coef_bytes(miter)); }
  // continued-fraction coefficients: 0 ap, 1 bp, 2 am, 3 bm
  __device__ double* cfc(int i) const { return (double*)(coef_base() + 4 * coef_bytes(miter) + i * coef_bytes(csteps)); }
  __device__ double* scal() const { return (double*)(coef_base() + 4 * coef_bytes(miter) + 4 * coef_bytes(csteps)); }
  __device__ SpinList list(int l, int32_t n) const { return SpinList{st(l), dg(l), cnt(l), tgt(l), val(l), n}; }
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Spin list of the sector (ns, n): states, diagonal sum_i eps_i n_i (ascending
// site order), hops over the bonds (0, k+1).
template <bool BLOCK>
__device__ int build_list(const Group<BLOCK>& g, const ModelParams& p, int ns, int n, int width, const SlotRef& sl,
                          int l) {
  const int count = (n < 0 || n > ns) ? 0 : (int)binom_u(ns, n);
  uint32_t* st = sl.st(l);
  double* dg = sl.dg(l);
  int32_t* cnt = sl.cnt(l);
  int32_t* tgt = sl.tgt(l);
  double* val = sl.val(l);
  for (int i = g.rank; i < count; i += g.size) {
    const uint32_t s = comb_unrank(i, ns, n);
    st[i] = s;
    double d = (s & 1u) ? p.eps_imp : 0.0;
    for (int k = 0; k < p.nb; ++k)
      if (s & (2u << k)) d += p.eps[k];
    dg[i] = d;
    int c = 0;
    for (int k = 0; k < p.nb; ++k) {
      const uint32_t m = 1u | (2u << k);
      const uint32_t b = s & m;
      if (b == 0 || b == m) continue;
      const bool neg = hop_sign(s, width, 0, k + 1) < 0;
      tgt[i * kMaxBath + c] = (int32_t)((uint32_t)comb_rank(s ^ m) | ((uint32_t)k << 16) | (neg ? 0x80000000u : 0u));
      val[i * kMaxBath + c] = neg ? -p.v[k] : p.v[k];
      ++c;
    }
    cnt[i] = c;
  }
  return count;
}

// Fills the group's LDS dn-list cache for the sector Dn when it fits (else
// marks it empty).  Every member must call it.
__device__ __noinline__ void cache_dn(const Group<true>& g, const SpinList& Dn) {
  if (!g.has_lds) return;
  const BlockLds& L = g.lds;
  int mc = 0;
  for (int i = g.rank; i < Dn.n; i += g.size) mc = mc > Dn.cnt[i] ? mc : Dn.cnt[i];
  const int slots = (int)g.max((double)mc);
  const int nch = (Dn.n + 63) >> 6;
  const bool fit = Dn.n > 0 && nch <= kLdsChunks && (int64_t)Dn.n * (slots > 0 ? slots : 1) <= L.dn_cap;
  if (fit) {
    for (int64_t k = g.rank; k < (int64_t)slots * Dn.n; k += g.size) {
      const int e = (int)(k / Dn.n), i = (int)(k - (int64_t)e * Dn.n);
      const bool on = e < Dn.cnt[i];
      L.tgt[k] = on ? (uint32_t)Dn.tgt[i * kMaxBath + e] & 0xffffu : 0u;
      L.val[k] = on ? Dn.val[i * kMaxBath + e] : 0.0;
    }
    for (int c = g.rank; c < nch; c += g.size) {
      int m = 0;
      for (int i = c * 64; i < Dn.n && i < c * 64 + 64; ++i) m = m > Dn.cnt[i] ? m : Dn.cnt[i];
      L.meta[2 + c] = m;
    }
    if (Dn.n <= L.ddn_cap)
      for (int i = g.rank; i < Dn.n; i += g.size) {
        L.ddn[i] = Dn.dg[i];
        L.sdn[i] = Dn.st[i];
      }
  }
  if (g.rank == 0) {
    L.meta[0] = fit ? Dn.n : 0;
    L.meta[1] = slots;
  }
  g.sync();
}

// Fills the group's LDS up-list cache with the sector U when it fits (else
// marks it empty).  Every member must call it.
__device__ __noinline__ void cache_up(const Group<true>& g, const SpinList& U) {
  const BlockLds& L = g.lds;
  int mc = 0;
  for (int i = g.rank; i < U.n; i += g.size) mc = mc > U.cnt[i] ? mc : U.cnt[i];
  const int slots = (int)g.max((double)mc);
  const bool fit = U.n > 0 && U.n <= L.u_rows && (int64_t)U.n * (slots > 0 ? slots : 1) <= L.u_cap;
  if (fit) {
    for (int i = g.rank; i < U.n; i += g.size) L.ucnt[i] = U.cnt[i];
    for (int64_t k = g.rank; k < (int64_t)slots * U.n; k += g.size) {
      const int i = (int)(k / slots), e = (int)(k - (int64_t)i * slots);
      const bool on = e < U.cnt[i];
      L.utgt[k] = on ? (uint32_t)U.tgt[i * kMaxBath + e] & 0xffffu : 0u;
      L.uval[k] = on ? U.val[i * kMaxBath + e] : 0.0;
    }
  }
  if (g.rank == 0) {
    const uint64_t a = (uint64_t)(uintptr_t)U.cnt;
    L.umeta[0] = fit ? U.n : 0;
    L.umeta[1] = slots;
    L.umeta[2] = (int32_t)(uint32_t)a;
    L.umeta[3] = (int32_t)(uint32_t)(a >> 32);
  }
  g.sync();
}

__device__ __forceinline__ bool up_cached(const Group<true>& g, const SpinList& U) {
  const uint64_t a = (uint64_t)(uintptr_t)U.cnt;
  return g.lds.umeta[0] == U.n && g.lds.umeta[2] == (int32_t)(uint32_t)a &&
         g.lds.umeta[3] == (int32_t)(uint32_t)(a >> 32);
}

// Wave-uniform copy of a value every lane reads from the same address (the
// spin lists are written by this kernel, so they are read with vector loads,
// not through the scalar cache).
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double rfl_d(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

constexpr int kUpPre = 6;  // up hops whose loads start before the staging barrier

__device__ __forceinline__ double readlane_d(double v, int l) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}


// Block-mode sweep 1 of lanczos_step with staged rows, for dn sectors of at
// most 64 * 16 states: w = sj H u - bj sp uprev, returning this thread's part
// of <w, u>.  Each step stages R = 16 / (64-column chunks per row) rows of u
// in LDS (one barrier per step, two row buffers); wave wv handles chunk wv % chunks of row wv / chunks: diagonal, up
// hops from the row's list (uniform over the wave, each hop a contiguous
// 64-column piece of another row, the first kUpPre loads issued before the
// staging barrier) and dn hops gathered from the staged row through the
// cached lists.
// LEAN: the row's up-hop list is read once (one entry per lane) and broadcast
// hop by hop, hop sources are addressed from wave-uniform row bases, and the
// dn hops go four slots at a time (table reads, then gathers, then the sums
// in slot order); the same arithmetic in the same order.
template <bool LEAN>
__device__ double hv_rows_block(const Group<true>& g, const ModelParams& p, const SpinList& U, const SpinList& Dn,
                                const double* __restrict__ u, const double* __restrict__ uprev,
                                double* __restrict__ w, double sj, double sp, double bj) {
  const int nd = Dn.n, nu = U.n;
  const int nch = (nd + 63) >> 6;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int R = (int)(blockDim.x >> 6) / nch;
  const int rr = wave / nch;
  const int id = (wave - rr * nch) * 64 + (int)(threadIdx.x & 63);
  const bool act = id < nd;
  const int idc = act ? id : nd - 1;
  const double bsp = bj * sp;
  // per-column constants of the dn sector
  const double ddn = Dn.dg[idc];
  const uint32_t sdn = Dn.st[idc];
  const int cm = rfl(g.lds.meta[2 + (wave - rr * nch)]);
  const int us = g.lds.umeta[1];  // up slots per state
  double dot = 0.0;
  int buf = 0;
  for (int r0 = 0; r0 < nu; r0 += R) {
    // two row buffers: a wave writes the other one next step only after
    // every wave has passed this step's barrier, i.e. finished the previous
    // step's gathers from it
    lds_f64* const row = g.lds.rows + (buf * R + rr) * nd;
    buf ^= 1;
    const int iu = r0 + rr;
    const bool on = rr < R && iu < nu;  // uniform over the wave
    double gu[kUpPre];
    double vu[kUpPre];
    int cu = 0;
    double xp = 0.0, dgu = 0.0;
    uint32_t stu = 0;
    const int ro = iu * nd;
    const int lane = (int)(threadIdx.x & 63);
    int lt = 0;
    double lv = 0.0;
    if (on) {
      cu = rfl(g.lds.ucnt[iu]);
      if constexpr (LEAN) {
        lt = lane < cu ? (int)g.lds.utgt[iu * us + lane] : 0;
        lv = lane < cu ? g.lds.uval[iu * us + lane] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < kUpPre; ++e) {
        if (e < cu) {
          if constexpr (LEAN) {
            vu[e] = readlane_d(lv, e);
            gu[e] = (u + (int64_t)__builtin_amdgcn_readlane(lt, e) * nd)[idc];
          } else {
            const int t = rfl((int)g.lds.utgt[iu * us + e]);
            vu[e] = rfl_d(g.lds.uval[iu * us + e]);
            gu[e] = u[t * nd + idc];
          }
        }
      }
      if (bsp != 0.0) xp = uprev[ro + idc];
      if (act) row[id] = u[ro + id];
      dgu = U.dg[iu];
      stu = (uint32_t)rfl((int)U.st[iu]);
    }
    __syncthreads();
    if (on) {
      double acc = 0.0;
#pragma unroll
      for (int e = 0; e < kUpPre; ++e)
        if (e < cu) acc = fma(vu[e], gu[e], acc);
      // the rest in batches of kUpPre loads in flight
      for (int e0 = kUpPre; e0 < cu; e0 += kUpPre) {
#pragma unroll
        for (int e = 0; e < kUpPre; ++e) {
          if (e0 + e < cu) {
            if constexpr (LEAN) {
              vu[e] = readlane_d(lv, e0 + e);
              gu[e] = (u + (int64_t)__builtin_amdgcn_readlane(lt, e0 + e) * nd)[idc];
            } else {
              const int t = rfl((int)g.lds.utgt[iu * us + e0 + e]);
              vu[e] = rfl_d(g.lds.uval[iu * us + e0 + e]);
              gu[e] = u[t * nd + idc];
            }
          }
        }
#pragma unroll
        for (int e = 0; e < kUpPre; ++e)
          if (e0 + e < cu) acc = fma(vu[e], gu[e], acc);
      }
      int e = 0;
      if constexpr (LEAN) {
        for (; e + 4 <= cm; e += 4) {
          uint32_t tg[4];
          double vl[4], gr[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            tg[r] = g.lds.tgt[(e + r) * nd + idc];
            vl[r] = g.lds.val[(e + r) * nd + idc];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) gr[r] = row[tg[r]];
#pragma unroll
          for (int r = 0; r < 4; ++r) acc = fma(vl[r], gr[r], acc);
        }
      }
      for (; e < cm; ++e) acc = fma(g.lds.val[e * nd + idc], row[g.lds.tgt[e * nd + idc]], acc);
      const double xk = row[idc];
      acc = fma(ddn + dgu + ((stu & sdn & 1u) ? p.u : 0.0), xk, acc);
      const double wk = sj * acc - bsp * xp;
      if (act) {
        w[ro + id] = wk;
        dot += wk * xk;
      }
    }
  }
  __syncthreads();  // the rows are free for the next sweep
  return dot;
}

// Block-mode sweep 1 of lanczos_step with wave-private rows, for dn sectors of
// at most 64 C states: w = sj H u - bj sp uprev, returning this thread's part
// of <w, u>.  Wave wv takes rows wv, wv + waves, ...; lane l holds columns
// l + 64 c (c < C), clamped to the last column so every load is in bounds
// (a clamped lane computes the last column's value again and stores it to the
// same address; only its dot term is masked).  Per row: the row's up-hop
// list is read once (one entry per lane) and broadcast hop by hop, the row of
// u and HB hops' sources (contiguous pieces of other rows, addressed from a
// wave-uniform row base) are loaded together, the row is staged in the
// wave's own LDS buffer, and the dn hops are gathered slot by slot with the C
// columns' table reads and gathers in flight together.  No workgroup
// barrier is taken: every wave streams its rows at its own pace.
template <int C>
__device__ double hv_rows_wave(const Group<true>& g, const ModelParams& p, const SpinList& U, const SpinList& Dn,
                               const double* __restrict__ u, const double* __restrict__ uprev,
                               double* __restrict__ w, double sj, double sp, double bj) {
  constexpr int HB = C >= 4 ? 2 : 4;  // up hops whose loads (C columns each) are in flight together
  const int nd = Dn.n, nu = U.n;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int nwv = (int)(blockDim.x >> 6);
  lds_f64* const row = g.lds.rows + wave * (64 * C);
  const double bsp = bj * sp;
  const int us = g.lds.umeta[1];      // up slots per state
  const int ds = rfl(g.lds.meta[1]);  // dn slots per state
  int idc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) idc[c] = lane + 64 * c < nd ? lane + 64 * c : nd - 1;
  double dot = 0.0;
  for (int iu = wave; iu < nu; iu += nwv) {
    const int64_t ro = (int64_t)iu * nd;
    const double* __restrict__ xr = u + ro;
    const int cu = rfl(g.lds.ucnt[iu]);
    const int lt = lane < cu ? (int)g.lds.utgt[iu * us + lane] : 0;
    const double lv = lane < cu ? g.lds.uval[iu * us + lane] : 0.0;
    double x[C], acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      x[c] = xr[idc[c]];
      acc[c] = 0.0;
    }
    for (int e0 = 0; e0 < cu; e0 += HB) {
      double gv[HB][C];
      double hv[HB];
#pragma unroll
      for (int e = 0; e < HB; ++e) {
        const bool on = e0 + e < cu;  // uniform over the wave; an off hop re-reads hop e0 with value 0
        const int ee = on ? e0 + e : e0;
        const double* __restrict__ src = u + (int64_t)__builtin_amdgcn_readlane(lt, ee) * nd;
        hv[e] = on ? readlane_d(lv, ee) : 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) gv[e][c] = src[idc[c]];
      }
#pragma unroll
      for (int e = 0; e < HB; ++e)
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] = fma(hv[e], gv[e][c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < C; ++c) row[idc[c]] = x[c];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double xp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) xp[c] = bsp != 0.0 ? uprev[ro + idc[c]] : 0.0;
    // dn hops slot by slot (slots past a state's count hold target 0 and value 0)
    for (int e = 0; e < ds; ++e) {
      uint32_t tg[C];
      double vl[C], gr[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        tg[c] = g.lds.tgt[e * nd + idc[c]];
        vl[c] = g.lds.val[e * nd + idc[c]];
      }
#pragma unroll
      for (int c = 0; c < C; ++c) gr[c] = row[tg[c]];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = fma(vl[c], gr[c], acc[c]);
    }
    const double dgu = U.dg[iu];
    const uint32_t stu = (uint32_t)rfl((int)U.st[iu]);
    const bool dl = nd <= g.lds.ddn_cap;  // the dn diagonal and states cached in LDS
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const double ddn = dl ? (double)g.lds.ddn[idc[c]] : Dn.dg[idc[c]];
      const uint32_t sdn = dl ? (uint32_t)g.lds.sdn[idc[c]] : Dn.st[idc[c]];
      const double a = fma(ddn + dgu + ((stu & sdn & 1u) ? p.u : 0.0), x[c], acc[c]);
      const double wk = sj * a - bsp * xp[c];
      w[ro + idc[c]] = wk;
      if (lane + 64 * c < nd) dot += wk * x[c];
    }
    // every lane's gathers from the row precede the next row's staging
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  return dot;
}

// One Lanczos step on sector (U, Dn), lazily scaled: v_j = sj * u, v_{j-1} =
// sp * uprev.  w = sj H u - bj * sp * uprev; then (if alpha < 0 on entry
// coefficient unknown) a = <w, v_j>; w -= a v_j; returns |w|^2 through *nrm2.
// With known coefficients (regeneration passes) a_in is used as given.
template <bool BLOCK, bool FAST = false, bool LEAN = false>
__device__ void lanczos_step(const Group<BLOCK>& g, const ModelParams& p, const SpinList& U, const SpinList& Dn,
                             const double* __restrict__ u, const double* __restrict__ uprev, double* __restrict__ w,
                             double sj, double sp, double bj, bool known, double a_in, double* a_out, double* nrm2) {
  const int nd = Dn.n;
  const int64_t D = (int64_t)U.n * nd;
  double dot = 0.0;
  // sweep 1: w = sj H u - bj sp uprev, dot <w, u>
  bool staged = false;
  if constexpr (BLOCK) {
    const int nch = (nd + 63) >> 6;
    const int nwv = (int)(blockDim.x >> 6);
    if (g.has_lds && g.lds.meta[0] == nd && nd > 0 && nch <= nwv && 2ll * (nwv / nch) * nd <= g.lds.rows_cap &&
        D < (1ll << 31)) {
      if (up_cached(g, U)) {
        const int rc = g.lds.rows_cap;
        if (FAST && g.wave_rows && nd <= 64 && nwv * 64 <= rc)
          dot = hv_rows_wave<1>(g, p, U, Dn, u, uprev, w, sj, sp, bj);
        else if (FAST && g.wave_rows && nd <= 128 && nwv * 128 <= rc)
          dot = hv_rows_wave<2>(g, p, U, Dn, u, uprev, w, sj, sp, bj);
        else if (FAST && g.wave_rows && nd <= 256 && nwv * 256 <= rc)
          dot = hv_rows_wave<4>(g, p, U, Dn, u, uprev, w, sj, sp, bj);
        else
          dot = hv_rows_block<LEAN>(g, p, U, Dn, u, uprev, w, sj, sp, bj);
        staged = true;
      }
    }
  }
  if (!staged) {
    int64_t k = g.rank;
    int iu = (int)(k / (nd > 0 ? nd : 1)), id = (int)(k - (int64_t)iu * nd);
    for (; k < D; k += g.size) {
      const double xk = u[k];
      double acc = (U.dg[iu] + Dn.dg[id] + ((U.st[iu] & Dn.st[id] & 1u) ? p.u : 0.0)) * xk;
      const int cu = U.cnt[iu];
      for (int e = 0; e < cu; ++e)
        acc += U.val[iu * kMaxBath + e] * u[(int64_t)(U.tgt[iu * kMaxBath + e] & 0xffff) * nd + id];
      const int cd = Dn.cnt[id];
      for (int e = 0; e < cd; ++e)
        acc += Dn.val[id * kMaxBath + e] * u[(int64_t)iu * nd + (Dn.tgt[id * kMaxBath + e] & 0xffff)];
      double wk = sj * acc;
      if (bj != 0.0) wk -= bj * sp * uprev[k];
      w[k] = wk;
      dot += wk * xk;
      id += g.size;
      while (id >= nd) {
        id -= nd;
        ++iu;
      }
    }
  }
  double a = a_in;
  if (!known) a = sj * g.sum(dot);
  *a_out = a;
  g.sync();
  // sweep 2: w -= a sj u, |w|^2 (four states per member in flight; the
  // same order of operations as one state at a time)
  double n2 = 0.0;
  const double as = a * sj;
  const int64_t gs = g.size;
  int64_t k = g.rank;
  for (; FAST && k + 3 * gs < D; k += 4 * gs) {
    double wv[4], uv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      wv[r] = w[k + r * gs];
      uv[r] = u[k + r * gs];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double wk = wv[r] - as * uv[r];
      w[k + r * gs] = wk;
      n2 += wk * wk;
    }
  }
  for (; k < D; k += gs) {
    const double wk = w[k] - as * u[k];
    w[k] = wk;
    n2 += wk * wk;
  }
  *nrm2 = known ? 0.0 : g.sum(n2);
  g.sync();
}

// lanczos_step as a call of its own: its registers are allocated apart from
// the caller's (the ground-state task keeps more state live across the step
// than the 768-thread budget leaves beside the wave-private row sweep).
template <bool BLOCK, bool FAST, bool LEAN>
__device__ __noinline__ void lanczos_step_call(const Group<BLOCK>& g, const ModelParams& p, const SpinList& U,
                                               const SpinList& Dn, const double* __restrict__ u,
                                               const double* __restrict__ uprev, double* __restrict__ w, double sj,
                                               double sp, double bj, bool known, double a_in, double* a_out,
                                               double* nrm2) {
  lanczos_step<BLOCK, FAST, LEAN>(g, p, U, Dn, u, uprev, w, sj, sp, bj, known, a_in, a_out, nrm2);
}

// Lowest eigenpair of the m x m tridiagonal (alpha, beta): multisection over
// the group's lanes on the Sturm count, then inverse iteration at a shift just
// below the eigenvalue (T - shift is positive definite there, so the Thomas
// solve needs no pivoting).  Writes the unit eigenvector to sv (tc is scratch)
// and returns theta and |last component| through theta / slast.
template <bool BLOCK>
__device__ __noinline__ void tri_lowest(const Group<BLOCK>& g, const double* alpha, const double* beta, int m, double* sv,
                           double* tc, double* scal, double* theta, double* slast) {
  double lo = 1e300, hi = -1e300;
  for (int i = g.rank; i < m; i += g.size) {
    const double r = (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i < m - 1 ? fabs(beta[i]) : 0.0);
    lo = fmin(lo, alpha[i] - r);
    hi = fmax(hi, alpha[i] + r);
  }
  lo = -g.max(-lo);
  hi = g.max(hi);
  constexpr int P = 64;  // shifts tested per round, one per lane of the first wave
  for (int round = 0; round < 64; ++round) {
    if (hi - lo <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 1e-300) break;
    // lane r < P tests x_r = lo + (hi - lo) (r + 1) / (P + 1): is an eigenvalue below it?
    bool below = false;
    if (g.rank < P) {
      const double xr = lo + (hi - lo) * (double)(g.rank + 1) / (double)(P + 1);
      double q = alpha[0] - xr;
      int c = q < 0.0 ? 1 : 0;
      for (int i = 1; i < m; ++i) {
        const double qq = q != 0.0 ? q : 1e-300;
        q = alpha[i] - xr - beta[i - 1] * beta[i - 1] / qq;
        c += q < 0.0 ? 1 : 0;
      }
      below = c >= 1;
    }
    // r1 = first shift with an eigenvalue below it (P if none): the lowest
    // eigenvalue lies in [x_{r1-1}, x_{r1}] with x_{-1} = lo, x_P = hi
    const int r1 = (int)(-g.max(below ? -(double)g.rank : -(double)P));
    const double nlo = lo + (hi - lo) * (double)r1 / (double)(P + 1);
    const double nhi = r1 >= P ? hi : lo + (hi - lo) * (double)(r1 + 1) / (double)(P + 1);
    lo = nlo;
    hi = nhi;
  }
  const double th = 0.5 * (lo + hi);
  if (g.rank == 0) {
    const double shift = th - fmax(1e-10 * fabs(th), 1e-12);
    for (int i = 0; i < m; ++i) sv[i] = 1.0;
    for (int it = 0; it < 2; ++it) {
      // Thomas algorithm on (T - shift) x = sv: forward elimination, then back substitution
      double c_prev = 0.0, d_prev = 0.0;
      for (int i = 0; i < m; ++i) {
        const double bl = i > 0 ? beta[i - 1] : 0.0;
        const double den = (alpha[i] - shift) - bl * c_prev;
        const double ci = i < m - 1 ? beta[i] / den : 0.0;
        const double di = (sv[i] - bl * d_prev) / den;
        tc[i] = ci;
        sv[i] = di;
        c_prev = ci;
        d_prev = di;
      }
      for (int i = m - 2; i >= 0; --i) sv[i] -= tc[i] * sv[i + 1];
      double nrm = 0.0;
      for (int i = 0; i < m; ++i) nrm += sv[i] * sv[i];
      nrm = 1.0 / sqrt(nrm);
      for (int i = 0; i < m; ++i) sv[i] *= nrm;
    }
    if (sv[0] < 0.0)
      for (int i = 0; i < m; ++i) sv[i] = -sv[i];
    scal[0] = th;
    scal[1] = fabs(sv[m - 1]);
  }
  g.sync();
  *theta = scal[0];
  *slast = scal[1];
  g.sync();
}

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// s * norm2 / (s z + e0 - a0 - b0^2 / (s z + e0 - a1 - ...)), z = i w.
// The coefficients are loaded eight steps ahead of the recursion (the same
// steps in the same order), so their latency is paid once per eight steps.
__device__ cplx cf_value(const double* a, const double* b, int m, double norm2, double e0, double s, double w) {
  const cplx z = {e0, s * w};
  cplx t = {0.0, 0.0};
  int j = m - 1;
  for (; j >= 8; j -= 8) {
    double aa[8], bb[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      aa[r] = a[j - r];
      bb[r] = b[j - r - 1];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) t = cdiv({bb[r] * bb[r], 0.0}, {z.re - aa[r] - t.re, z.im - t.im});
  }
  for (; j >= 1; --j) t = cdiv({b[j - 1] * b[j - 1], 0.0}, {z.re - a[j] - t.re, z.im - t.im});
  const cplx r = cdiv({s * norm2, 0.0}, {z.re - a[0] - t.re, z.im - t.im});
  return r;
}

// Lanczos coefficients of the sector (U, Dn) from the start vector in vec(1)
// (unnormalised, |u|^2 = n2), stopping at breakdown, at cap steps, or when the
// continued fraction s n2 / (s z + e0 - a0 - ...) at z = i wmin has changed by
// at most cf_tol * max(|G|, n2) over each of the last two 8-step checks (the
// fraction converges last at the frequency closest to the real axis).
// Returns the step count.
template <bool BLOCK, bool FAST>
__device__ int cf_coeffs(const Group<BLOCK>& g, const ModelParams& p, const SpinList& U, const SpinList& Dn,
                         const SlotRef& sl, double n2, int cap, double e0, double sgn, double wmin, double cf_tol,
                         double* a, double* b, double* qt, int kcap) {
  double* u = sl.vec(1);
  double* up = sl.vec(0);
  double* w = sl.vec(2);
  double sj = 1.0 / sqrt(n2), sp = 0.0, bj = 0.0;
  int m = 0, calm = 0;
  cplx prev = {0.0, 0.0};
  // Continuants of the fraction at z = i wmin, carried step by step:
  // y_m = d_0 - b_0^2 / (d_1 - b_1^2 / (... d_{m-1})) = A_m / B_m with
  // d_k = (e0 - a_k) + i sgn wmin, A_n = d_n A_{n-1} - b_{n-1}^2 A_{n-2}
  // (A_{-1} = 1, A_0 = d_0; B_{-1} = 0, B_0 = 1), and f_m = sgn n2 B_m / A_m
  // is the value cf_value(a, b, m, n2, e0, sgn, wmin) takes.  Every member
  // holds the same values (rescaled by the size of A_m each step), so a check
  // costs O(1) and no barrier.
  cplx A = {1.0, 0.0}, Ap = {0.0, 0.0}, B = {0.0, 0.0}, Bp = {1.0, 0.0};
  const int64_t D = (int64_t)U.n * Dn.n;
  for (int j = 0; j < cap; ++j) {
    double aj, nn;
    if (qt)  // keep the normalised Krylov vector v_j = sj u, state-major
      for (int64_t k = g.rank; k < D; k += g.size) qt[k * kcap + j] = sj * u[k];
    lanczos_step<BLOCK, FAST, true>(g, p, U, Dn, u, up, w, sj, sp, bj, false, 0.0, &aj, &nn);
    a[j] = aj;
    m = j + 1;
    const double bn = sqrt(nn);
    if (j == cap - 1) break;
    if (!(bn > 1e-12 * (fabs(aj) + bj))) break;  // invariant subspace
    b[j] = bn;
    {
      const cplx d = {e0 - aj, sgn * wmin};
      const double c = -bj * bj;  // -b_{j-1}^2 (unused at j = 0)
      cplx An, Bn;
      if (j == 0) {
        An = d;
        Bn = {1.0, 0.0};
        Ap = {1.0, 0.0};
        Bp = {0.0, 0.0};
      } else {
        An = {d.re * A.re - d.im * A.im + c * Ap.re, d.re * A.im + d.im * A.re + c * Ap.im};
        Bn = {d.re * B.re - d.im * B.im + c * Bp.re, d.re * B.im + d.im * B.re + c * Bp.im};
        Ap = A;
        Bp = B;
      }
      A = An;
      B = Bn;
      const double sz = fmax(fabs(A.re), fabs(A.im));
      if (sz > 0.0) {
        const double r = 1.0 / sz;
        A = {A.re * r, A.im * r};
        B = {B.re * r, B.im * r};
        Ap = {Ap.re * r, Ap.im * r};
        Bp = {Bp.re * r, Bp.im * r};
      }
    }
    if (m % 8 == 0) {
      const cplx f = cdiv({sgn * n2 * B.re, sgn * n2 * B.im}, A);
      // change measured against the part's spectral weight n2 as well, so a
      // small |G(i wmin)| does not make the test unreachable
      const double dd = hypot(f.re - prev.re, f.im - prev.im);
      calm = dd <= cf_tol * fmax(hypot(f.re, f.im), n2) ? calm + 1 : 0;
      prev = f;
      if (calm >= 2) break;
    }
    double* t = up;
    up = u;
    u = w;
    w = t;
    sp = sj;
    sj = 1.0 / bn;
    bj = bn;
  }
  return m;
}

__device__ __forceinline__ double start_value(uint64_t seed, int64_t k) {
  return (double)(mix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(k + 1))) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

// Ground-state task of model mi: Lanczos over the (n_up, n_dn) sector with
// every vector kept in the model's store, psi = sum_j sv_j v_j written over the
// store's first vector, its explicit residual, E0 (also into the model's
// coefficient record) and the step count.
// psi[k] = sum_j c_j V[j D + k] over V_0 (each state's entries of every V_j
// are read by the one member that then overwrites V_0 there), eight vectors'
// loads in flight at a time, summed in step order.  A call of its own, so its
// registers are allocated apart from the ground-state loop's.
template <bool BLOCK>
__device__ __noinline__ void combine_psi(const Group<BLOCK>& g, double* V, const double* c, int m, int64_t D) {
  for (int64_t k = g.rank; k < D; k += g.size) {
    double acc = 0.0;
    int j = 0;
    for (; j + 8 <= m; j += 8) {
      double cv[8], vv[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        cv[r] = c[j + r];
        vv[r] = V[(int64_t)(j + r) * D + k];
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) acc += cv[r] * vv[r];
    }
    for (; j < m; ++j) acc += c[j] * V[(int64_t)j * D + k];
    V[k] = acc;
  }
}

template <bool BLOCK, bool FAST>
__device__ void solve_ground(const Group<BLOCK>& g, const Config& cf, int mi, const SlotRef& sl) {
  const ModelParams& p = cf.params[mi];
  const int ns = p.nb + 1;
  const int n_0 = build_list<BLOCK>(g, p, ns, p.n_up, cf.sign_width, sl, 1);
  const int nd = build_list<BLOCK>(g, p, ns, p.n_dn, cf.sign_width, sl, 3);
  g.sync();
  const SpinList U = sl.list(1, n_0);
  const SpinList Dn = sl.list(3, nd);
  if constexpr (BLOCK) {
    if (g.has_lds) {
      cache_dn(g, Dn);
      cache_up(g, U);
    }
  }
  const int64_t D = (int64_t)n_0 * nd;
  double* const alpha = sl.alpha();
  double* const beta = sl.beta();

  // ---- Lanczos with every vector kept (u_j = V_j, v_j = s_j u_j) ----
  const uint64_t seed = cf.seed ^ (0xA0761D6478BD642Full * (uint64_t)(mi + 1));
  double* const V = cf.gs_store + cf.off_gs[mi];
  double n2 = 0.0;
  for (int64_t k = g.rank; k < D; k += g.size) {
    const double r = start_value(seed, k);
    V[k] = r;
    n2 += r * r;
  }
  n2 = g.sum(n2);
  g.sync();
  const double s0 = 1.0 / sqrt(n2);
  double sj = s0, sp = 0.0, bj = 0.0;
  int m = 0;
  double theta = 0.0, slast = 0.0, ritz = 1e300;
  const int cap = (int)((int64_t)cf.max_iter < D ? (int64_t)cf.max_iter : D);
  for (int j = 0; j < cap; ++j) {
    double aj, nn;
    const double* u = V + (int64_t)j * D;
    lanczos_step_call<BLOCK, FAST, true>(g, p, U, Dn, u, j > 0 ? u - D : u, V + (int64_t)(j + 1) * D, sj, sp, bj, false, 0.0, &aj, &nn);
    m = j + 1;
    const double bn = sqrt(nn);
    if (g.rank == 0) {
      alpha[j] = aj;
      beta[j] = bn;
    }
    g.sync();
    const bool breakdown = !(bn > 1e-12 * (fabs(aj) + bj));
    const bool check = breakdown || m == cap || (m >= cf.check_every && m % cf.check_every == 0);
    if (check) {
      bool done = false;
      if constexpr (BLOCK) {
        // block mode: the tridiagonal solve runs on copies in the LDS row
        // buffer (idle between sweeps); the eigenvector goes back to sv
        if (g.has_lds && 4 * m + 8 <= g.lds.rows_cap) {
          double* const st = (double*)g.lds.rows;
          for (int i = g.rank; i < m; i += g.size) {
            st[i] = alpha[i];
            st[m + i] = beta[i];
          }
          g.sync();
          tri_lowest<BLOCK>(g, st, st + m, m, st + 2 * m, st + 3 * m, st + 4 * m, &theta, &slast);
          for (int i = g.rank; i < m; i += g.size) sl.sv()[i] = st[2 * m + i];
          g.sync();
          done = true;
        }
      }
      if (!done) tri_lowest<BLOCK>(g, alpha, beta, m, sl.sv(), sl.tc(), sl.scal(), &theta, &slast);
      ritz = breakdown ? 0.0 : bn * slast;
      if (ritz <= cf.tol || breakdown) break;
    }
    sp = sj;
    sj = 1.0 / bn;
    bj = bn;
  }
  // psi = sum_j sv_j s_j V_j in step order, over V_0: each state's entries of
  // every V_j are read by the one member that then overwrites V_0 there
  double* const psi = V;
  {
    double* const c = sl.tc();
    for (int j = g.rank; j < m; j += g.size) c[j] = sl.sv()[j] * (j == 0 ? s0 : 1.0 / beta[j - 1]);
    g.sync();
    combine_psi<BLOCK>(g, V, c, m, D);
  }
  g.sync();
  // normalise psi and take the explicit residual |H psi - theta psi|
  double pn = 0.0;
  for (int64_t k = g.rank; k < D; k += g.size) pn += psi[k] * psi[k];
  pn = g.sum(pn);
  g.sync();
  {
    const double inv = 1.0 / sqrt(pn);
    for (int64_t k = g.rank; k < D; k += g.size) psi[k] *= inv;
  }
  g.sync();
  double res2;
  {
    double aj, nn;
    // w = H psi - theta psi (unit scale, no previous vector, known coefficient)
    double* wr = sl.vec(2);
    lanczos_step_call<BLOCK, FAST, true>(g, p, U, Dn, psi, sl.vec(0), wr, 1.0, 0.0, 0.0, true, theta, &aj, &nn);
    double r2 = 0.0;
    for (int64_t k = g.rank; k < D; k += g.size) r2 += wr[k] * wr[k];
    res2 = g.sum(r2);
    g.sync();
  }
  if (cf.keep) {
    double* ps = cf.store_psi + cf.off_psi[mi];
    for (int64_t k = g.rank; k < D; k += g.size) ps[k] = psi[k];
  }
  if (g.rank == 0) {
    cf.store_coef[(int64_t)mi * (4 * cf.kcap + 8) + 4 * cf.kcap] = theta;
    cf.e0[mi] = theta;
    cf.iters[3 * mi] = m;
    cf.resid[2 * mi] = ritz;
    cf.resid[2 * mi + 1] = sqrt(res2);
  }
  g.sync();
}

// Continued-fraction task of model mi, after its ground-state task (an
// earlier launch): part 0 the particle fraction from c+_{0,up} psi (sector
// n_up + 1), part 1 the hole fraction from c_{0,up} psi (n_up - 1).  The
// coefficients (zero past the step count), |phi|^2 and the step count go to
// the part's half of the model's coefficient record
// [a+ | b+ | a- | b- (kcap each) | e0, n2+, n2-, m+, m-].
template <bool BLOCK, bool FAST>
__device__ void solve_fraction(const Group<BLOCK>& g, const Config& cf, int mi, int part, const SlotRef& sl) {
  const ModelParams& p = cf.params[mi];
  const int ns = p.nb + 1;
  const int l = part == 0 ? 2 : 0;
  const int nt = build_list<BLOCK>(g, p, ns, p.n_up + (part == 0 ? 1 : -1), cf.sign_width, sl, l);
  const int nd = build_list<BLOCK>(g, p, ns, p.n_dn, cf.sign_width, sl, 3);
  g.sync();
  const SpinList T = sl.list(l, nt);
  const SpinList Dn = sl.list(3, nd);
  const int kc = cf.kcap;
  double* const co = cf.store_coef + (int64_t)mi * (4 * kc + 8);
  double* const a = co + 2 * part * kc;
  double* const b = a + kc;
  const double theta = cf.e0[mi];
  const double* const psi = cf.gs_store + cf.off_gs[mi];
  const int64_t Dt = (int64_t)nt * nd;
  double* const phi = sl.vec(1);
  double nn = 0.0;
  // gather: target row it takes source row rank(st ^ 1) when its bit 0 fits
  for (int64_t k = g.rank; k < Dt; k += g.size) {
    const int it = (int)(k / nd), id = (int)(k - (int64_t)it * nd);
    const uint32_t st = T.st[it];
    const bool ok = part == 0 ? (st & 1u) != 0 : (st & 1u) == 0;
    double v = 0.0;
    if (ok) v = psi[(int64_t)comb_rank(st ^ 1u) * nd + id];
    phi[k] = v;
    nn += v * v;
  }
  nn = g.sum(nn);
  g.sync();
  int m = 0;
  if (nn > 0.0) {
    if constexpr (BLOCK) {
      if (g.has_lds) {
        cache_dn(g, Dn);
        cache_up(g, T);
      }
    }
    double* qt = cf.keep ? cf.store_q + cf.off_q[2 * mi + part] : nullptr;
    m = cf_coeffs<BLOCK, FAST>(g, p, T, Dn, sl, nn, cf.cf_steps, theta, part == 0 ? 1.0 : -1.0, cf.wmin, cf.cf_tol, a, b,
                         qt, kc);
  }
  // coefficient arrays were written by every member with the same values
  g.sync();
  for (int j = g.rank; j < kc; j += g.size) {
    if (j >= m) a[j] = 0.0;
    if (j >= m - 1) b[j] = 0.0;
  }
  if (g.rank == 0) {
    co[4 * kc + 1 + part] = nn;
    co[4 * kc + 3 + part] = m;
    cf.iters[3 * mi + 1 + part] = m;
  }
  g.sync();
}

// PHASE 0: ground-state tasks, one per model of the queues.  PHASE 1: the
// fraction tasks, two per model (queue position q >> 1, part q & 1), in the
// same order.  Each phase pops its own pair of counters.
template <int PHASE>
__global__ __launch_bounds__(PHASE ? kBlockFrac : kBlockGround) void k_batch_solve(Config cf) {
  __shared__ double red[32];
  __shared__ int popped;
  int32_t* const counters = cf.counters + 2 * PHASE;
  if ((int)blockIdx.x < cf.wave_blocks) {
    // wave mode
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    Group<false> g{(int)(threadIdx.x & 63), 64, nullptr};
    const int64_t slot = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    const SlotRef sl{cf.ws_wave + slot * cf.slot_bytes_wave, cf.list_cap_wave, cf.vec_cap_wave, cf.max_iter,
                     cf.cf_steps};
    const int n = PHASE ? 2 * cf.n_wave : cf.n_wave;
    while (true) {
      int q = 0;
      if (g.rank == 0) q = atomicAdd(&counters[0], 1);
      q = __builtin_amdgcn_readfirstlane(__shfl(q, 0, 64));  // uniform task index
      if (q >= n) break;
      if (PHASE == 0)
        solve_ground<false, kFastGround>(g, cf, __builtin_amdgcn_readfirstlane(cf.queue_wave[q]), sl);
      else
        solve_fraction<false, kFastFrac>(g, cf, __builtin_amdgcn_readfirstlane(cf.queue_wave[q >> 1]), q & 1, sl);
    }
  } else {
    __shared__ double s_rows[kLdsRowsSolve];
    __shared__ uint32_t s_tgt[kLdsDnSolve];
    __shared__ double s_val[kLdsDnSolve];
    __shared__ int32_t s_meta[2 + kLdsChunks];
    __shared__ uint32_t s_utgt[kLdsUpSolve];
    __shared__ double s_uval[kLdsUpSolve];
    __shared__ int32_t s_ucnt[kLdsUpRowsSolve];
    __shared__ int32_t s_umeta[4];
    __shared__ double s_ddn[kLdsDiagSolve];
    __shared__ uint32_t s_sdn[kLdsDiagSolve];
    Group<true> g{(int)threadIdx.x, (int)blockDim.x, red};
    g.lds = BlockLds{(lds_f64*)s_rows, kLdsRowsSolve,     (lds_u32*)s_tgt,  (lds_f64*)s_val,
                     kLdsDnSolve,      (lds_i32*)s_meta,  (lds_u32*)s_utgt, (lds_f64*)s_uval,
                     (lds_i32*)s_ucnt, kLdsUpSolve,       kLdsUpRowsSolve,  (lds_i32*)s_umeta,
                     (lds_f64*)s_ddn,  (lds_u32*)s_sdn,   kLdsDiagSolve};
    if (threadIdx.x == 0) s_umeta[0] = 0;
    g.has_lds = cf.staged != 0;
    g.wave_rows = cf.staged == 1;
    const int64_t slot = (int64_t)blockIdx.x - cf.wave_blocks;
    const SlotRef sl{cf.ws_block + slot * cf.slot_bytes_block, cf.list_cap_block, cf.vec_cap_block, cf.max_iter,
                     cf.cf_steps};
    const int n = PHASE ? 2 * cf.n_block : cf.n_block;
    while (true) {
      if (threadIdx.x == 0) popped = atomicAdd(&counters[1], 1);
      __syncthreads();
      const int q = __builtin_amdgcn_readfirstlane(popped);
      __syncthreads();
      if (q >= n) break;
      if (PHASE == 0)
        solve_ground<true, kFastGround>(g, cf, __builtin_amdgcn_readfirstlane(cf.queue_block[q]), sl);
      else
        solve_fraction<true, kFastFrac>(g, cf, __builtin_amdgcn_readfirstlane(cf.queue_block[q >> 1]), q & 1, sl);
    }
  }
}

// G(i w_n) of every model from its coefficient record: the particle plus the
// hole continued fraction, one thread per (model, frequency).
__global__ void k_batch_g(Config cf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cf.num_models * cf.nw) return;
  const int64_t mi = t / cf.nw;
  const int n = (int)(t - mi * cf.nw);
  const int kc = cf.kcap;
  const double* co = cf.store_coef + mi * (4 * kc + 8);
  const double e0 = co[4 * kc];
  const int mp = (int)co[4 * kc + 3], mm = (int)co[4 * kc + 4];
  cplx gsum = {0.0, 0.0};
  if (mp > 0) {
    const cplx c = cf_value(co, co + kc, mp, co[4 * kc + 1], e0, 1.0, cf.wn[n]);
    gsum.re += c.re;
    gsum.im += c.im;
  }
  if (mm > 0) {
    // hole part: -norm2 / (-z + e0 - a0 - ...) = norm2 / (z - e0 + a0 + ...)
    const cplx c = cf_value(co + 2 * kc, co + 3 * kc, mm, co[4 * kc + 2], e0, -1.0, cf.wn[n]);
    gsum.re += c.re;
    gsum.im += c.im;
  }
  cf.g[2 * t] = gsum.re;
  cf.g[2 * t + 1] = gsum.im;
}

// ===========================================================================
// Adjoint backward of the batch (bath fit).  For L = sum_n |G_n - T_n|^2 the
// caller passes dL/dRe G and dL/dIm G; with w_n = dL/dRe G_n - i dL/dIm G_n,
//   dL/dtheta = S+ - S- + (B - A) dE0/dtheta - 2 psi^T (dH/dtheta) zeta,
// where, for the particle (+) and hole (-) parts with Krylov bases Q and
// resolvent vectors x_n = Q u_n (u_n = sigma |phi| (sigma z_n + E0 - T)^-1 e1):
//   K = Q Re(W) Q^T,  W = sum_n w_n u_n u_n^T,
//   S = tr(dH/dtheta K) over the part's sector,  A (B) = tr K of the particle (hole) part,
//   dE0/dtheta = psi^T (dH/dtheta) psi  (Hellmann-Feynman),
//   zeta: (H - E0) zeta = P r, P = 1 - psi psi^T, r = Re(c X + c+ Y),
//   X = sum_n w_n x_n (particle), Y = sum_n w_n y_n (hole).
// theta = (eps_bath, v): dH/d eps_k = n_k,up + n_k,dn and dH/d v_k = the bond
// (0, k) hopping with unit amplitude (and the model's hopping signs).
// One workgroup per model; no orthogonality of Q is assumed (everything is
// evaluated through K in the sector's state space).
// ===========================================================================
// A continued-fraction part whose K products run device-wide (its Krylov
// basis is long): P = Qt Re W by the grouped matrix-core GEMM, then diag(K)
// and the bond sums of K over 64-row blocks.
struct BwdJob {
  int32_t model, part, m, dt;  // dt: states of the part's sector
  int64_t q_off;               // Krylov basis in the basis store (rows of stride kcap)
  int64_t a_off, b_off;        // A = (Re V, -Im V) and B = (Re U, Im U) interleaved, m x 2 nw each
  int64_t w_off;               // Re W = A B^T (m x m, row-major) in the job store
  int64_t p_off;               // P (dt x m, row-major) in the job store
  int32_t blk0, nblk;          // its 64-row blocks in the block list
};
// One product C = A B^T of the backward's device-wide stage (row-major,
// C with row stride N); A from the basis store (a_in_q) or the job store.
struct BwdGemm {
  int64_t a_off, b_off, c_off;
  int32_t lda, ldb, M, N, K, a_in_q;
};
struct BwdTile {  // a 64 x 64 tile of one product
  int32_t job, r0, c0, pad;
};
struct BwdBlock {  // 64 rows of one job's sector
  int32_t job, r0;
};
constexpr int kRecPart = 1 + 2 * kMaxBath;           // per part: tr K, S_eps[k], S_v[k] (before the part's sign)
constexpr int kRec = 2 * kRecPart + 4 * kMaxBath;    // + dE0/deps_k, psi dH/deps_k zeta, dE0/dv_k, psi dH/dv_k zeta

struct BwdConfig {
  const ModelParams* params;
  int32_t num_models;
  int32_t* counter;
  char* ws;
  int64_t slot_bytes;
  int32_t lcap;
  int64_t vcap;
  int32_t kcap, nw, sign_width, max_bath;
  int32_t staged;     // 1 wave-private row staging where it fits (else workgroup rows), 2 workgroup rows, 0 none
  const double* wn;
  const double* gre;  // [num_models][nw] dL/dRe G
  const double* gim;  // [num_models][nw] dL/dIm G
  const int64_t* off_psi;
  const int64_t* off_q;
  const double* store_psi;
  const double* store_q;
  const double* store_coef;
  int32_t cg_max;
  double cg_tol;
  double* grad_eps;   // [num_models][max_bath]
  double* grad_v;     // [num_models][max_bath]
  int32_t* cg_iters;  // [num_models]
  // parts whose K products run device-wide (job_of[2 model + part] >= 0)
  const int32_t* job_of;
  const BwdJob* jobs;
  const BwdBlock* blocks;
  double* jstore;
  double* rec;        // [num_models][kRec]
  double* partial;    // [blocks][kRecPart]
};

// Backward slot: the four spin lists (same layout as SlotRef), then
// N-sector vectors zeta, res, pvec, q; X re/im, Y re/im; diag d; U, U2
// ([nw][kcap] complex each); ReW [kcap][kcap]; ubar (complex [kcap]);
// P [vcap][kcap]; scalars.
struct BwdSlot {
  SlotRef lists;  // only the list accessors are used
  __host__ __device__ int64_t vb() const { return align_up(8 * lists.vcap, 256); }
  __host__ __device__ int64_t ub(int nw) const { return align_up(16ll * nw * lists.csteps, 256); }
  __host__ __device__ int64_t wb() const { return align_up(8ll * lists.csteps * lists.csteps, 256); }
  __host__ __device__ int64_t pb() const { return align_up(8ll * lists.vcap * lists.csteps, 256); }
  __host__ __device__ int64_t bytes(int nw) const {
    return 4 * lists.list_bytes() + 9 * vb() + 2 * ub(nw) + wb() + align_up(16ll * lists.csteps, 256) + pb() +
           align_up(8 * 64, 256);
  }
  __device__ char* after_lists() const { return lists.base + 4 * lists.list_bytes(); }
  // 0 zeta, 1 res, 2 pvec, 3 q, 4 Xre, 5 Xim, 6 Yre, 7 Yim, 8 diag
  __device__ double* vec(int i) const { return (double*)(after_lists() + i * vb()); }
  __device__ double* U(int which, int nw) const { return (double*)(after_lists() + 9 * vb() + which * ub(nw)); }
  __device__ double* ReW(int nw) const { return (double*)(after_lists() + 9 * vb() + 2 * ub(nw)); }
  __device__ double* ubar(int nw) const { return (double*)((char*)ReW(nw) + wb()); }
  __device__ double* P(int nw) const { return (double*)((char*)ubar(nw) + align_up(16ll * lists.csteps, 256)); }
  __device__ double* scal(int nw) const { return (double*)((char*)P(nw) + pb()); }
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// Occupation of bath site k (bond k couples sites 0 and k + 1) in a state pair.
__device__ __forceinline__ double occ_k(uint32_t su, uint32_t sd, int k) {
  return (double)(((su >> (k + 1)) & 1u) + ((sd >> (k + 1)) & 1u));
}

// sum over directed hops a -> t on bond k of the sector (U, Dn): sign * f(a, t)
template <typename F>
__device__ __forceinline__ double bond_hops(const SpinList& U, const SpinList& Dn, int64_t a, int k, F f) {
  const int nd = Dn.n;
  const int iu = (int)(a / nd), id = (int)(a - (int64_t)iu * nd);
  double s = 0.0;
  for (int e = 0; e < U.cnt[iu]; ++e) {
    const uint32_t t = (uint32_t)U.tgt[iu * kMaxBath + e];
    if ((int)((t >> 16) & 0xff) != k) continue;
    const double sg = (t & 0x80000000u) ? -1.0 : 1.0;
    s += sg * f(a, (int64_t)(t & 0xffff) * nd + id);
  }
  for (int e = 0; e < Dn.cnt[id]; ++e) {
    const uint32_t t = (uint32_t)Dn.tgt[id * kMaxBath + e];
    if ((int)((t >> 16) & 0xff) != k) continue;
    const double sg = (t & 0x80000000u) ? -1.0 : 1.0;
    s += sg * f(a, (int64_t)iu * nd + (t & 0xffff));
  }
  return s;
}

constexpr int kPc = 16;             // W columns per LDS chunk in the backward
constexpr int kBwdLds = 512 * kPc;  // doubles of the backward's LDS staging buffer (64 KiB)
constexpr int kBwdMaxSteps = 1024;  // continued-fraction steps a kept basis may hold
constexpr int kBwdSplitM = 96;      // parts with this many steps run their K products device-wide

// sum_j x[j] y[j] over j < n in order, eight entries' loads in flight at a
// time (the same sum as one term at a time).
__device__ __forceinline__ double row_dot(const double* __restrict__ x, const double* __restrict__ y, int n) {
  double acc = 0.0;
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    double a[8], b[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      a[r] = x[j + r];
      b[r] = y[j + r];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) acc += a[r] * b[r];
  }
  for (; j < n; ++j) acc += x[j] * y[j];
  return acc;
}

// <p, y1> and <p, y2> (y2 may be null: its sum stays 0) in order, p's loads
// shared, eight entries of each in flight at a time.
__device__ __forceinline__ void row_dot2(const double* __restrict__ x, const double* __restrict__ y1,
                                         const double* __restrict__ y2, int n, double* s1, double* s2) {
  double a1 = 0.0, a2 = 0.0;
  int j = 0;
  if (y2) {
    for (; j + 8 <= n; j += 8) {
      double xv[8], v1[8], v2[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        xv[r] = x[j + r];
        v1[r] = y1[j + r];
        v2[r] = y2[j + r];
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        a1 += xv[r] * v1[r];
        a2 += xv[r] * v2[r];
      }
    }
    for (; j < n; ++j) {
      a1 += x[j] * y1[j];
      a2 += x[j] * y2[j];
    }
  } else {
    a1 = row_dot(x, y1, n);
  }
  *s1 = a1;
  *s2 = a2;
}

__device__ void backward_model(const Group<true>& g, const BwdConfig& bc, int mi, const BwdSlot& bs, double* wl) {
  const ModelParams& p = bc.params[mi];
  const int ns = p.nb + 1;
  const int nb = p.nb;
  const SlotRef& sl = bs.lists;
  const int n_m = build_list<true>(g, p, ns, p.n_up - 1, bc.sign_width, sl, 0);
  const int n_0 = build_list<true>(g, p, ns, p.n_up, bc.sign_width, sl, 1);
  const int n_p = build_list<true>(g, p, ns, p.n_up + 1, bc.sign_width, sl, 2);
  const int nd = build_list<true>(g, p, ns, p.n_dn, bc.sign_width, sl, 3);
  g.sync();
  const SpinList U0 = sl.list(1, n_0);
  const SpinList Dn = sl.list(3, nd);
  if (g.has_lds) {
    cache_dn(g, Dn);
    cache_up(g, U0);
  }
  const int64_t D = (int64_t)n_0 * nd;
  const int kcap = bc.kcap, nw = bc.nw;
  const double* co = bc.store_coef + (int64_t)mi * (4 * kcap + 8);
  const double e0 = co[4 * kcap];
  const double* psi = bc.store_psi + bc.off_psi[mi];
  double* const rec = bc.rec + (int64_t)mi * kRec;
  for (int t = g.rank; t < kRec; t += g.size) rec[t] = 0.0;
  g.sync();
  for (int part = 0; part < 2; ++part) {
    const int m = (int)co[4 * kcap + 3 + part];
    if (m <= 0) continue;
    const int job = bc.job_of[2 * mi + part];
    const double sigma = part == 0 ? 1.0 : -1.0;
    const double* a = co + (part == 0 ? 0 : 2 * kcap);
    const double* b = co + (part == 0 ? kcap : 3 * kcap);
    const double nphi = sqrt(co[4 * kcap + 1 + part]);
    const SpinList T = sl.list(part == 0 ? 2 : 0, part == 0 ? n_p : n_m);
    const int64_t Dt = (int64_t)T.n * nd;
    const double* qt = bc.store_q + bc.off_q[2 * mi + part];
    double* Uc = bs.U(0, nw);   // c' then x
    double* Ur = bs.U(1, nw);   // rhs'
    // (a) u_n = sigma |phi| (sigma z_n + e0 - T)^-1 e1, complex Thomas, one
    // thread per frequency; U is stored frequency-fastest ([i][n], complex),
    // so a step's stores and loads are coalesced over the frequencies.  The
    // denominators are O(|z| + |T|), far from overflow, so 1/den is taken as
    // conj(den) / |den|^2 (one division per step)
    auto ux = [&](int n, int i) { return 2 * ((int64_t)i * nw + n); };
    for (int n = g.rank; n < nw; n += g.size) {
      const cplx zs = {e0, sigma * bc.wn[n]};
      cplx cp = {0.0, 0.0}, rp = {0.0, 0.0};
      for (int i = 0; i < m; ++i) {
        const double offl = i > 0 ? -b[i - 1] : 0.0;
        const cplx den = {zs.re - a[i] - offl * cp.re, zs.im - offl * cp.im};
        const double s2 = 1.0 / (den.re * den.re + den.im * den.im);
        const cplx inv = {den.re * s2, -den.im * s2};
        const cplx ci = i < m - 1 ? cplx{-b[i] * inv.re, -b[i] * inv.im} : cplx{0.0, 0.0};
        const cplx ri = cmul({(i == 0 ? 1.0 : 0.0) - offl * rp.re, -offl * rp.im}, inv);
        Uc[ux(n, i)] = ci.re;
        Uc[ux(n, i) + 1] = ci.im;
        Ur[ux(n, i)] = ri.re;
        Ur[ux(n, i) + 1] = ri.im;
        cp = ci;
        rp = ri;
      }
      cplx x = rp;
      Uc[ux(n, m - 1)] = sigma * nphi * x.re;
      Uc[ux(n, m - 1) + 1] = sigma * nphi * x.im;
      for (int i = m - 2; i >= 0; --i) {
        const cplx ci = {Uc[ux(n, i)], Uc[ux(n, i) + 1]};
        const cplx ri = {Ur[ux(n, i)], Ur[ux(n, i) + 1]};
        const cplx t = cmul(ci, x);
        x = {ri.re - t.re, ri.im - t.im};
        Uc[ux(n, i)] = sigma * nphi * x.re;
        Uc[ux(n, i) + 1] = sigma * nphi * x.im;
      }
    }
    g.sync();
    // (b) Re W = Re(U diag(gre - i gim) U^T), upper triangle in 4 x 4 tiles
    // (one per thread and pass), frequencies staged through LDS nc at a
    // time as v_in = (gre_n - i gim_n) u_in and u_jn; then ubar
    double* W = bs.ReW(nw);
    double* ub = bs.ubar(nw);
    const double* gre = bc.gre + (int64_t)mi * nw;
    const double* gim = bc.gim + (int64_t)mi * nw;
    if (job >= 0) {
      // long basis: A and B of Re W = A B^T to the job store (the product
      // and everything after it run device-wide)
      const BwdJob& jb = bc.jobs[job];
      double* const Aj = bc.jstore + jb.a_off;
      double* const Bj = bc.jstore + jb.b_off;
      for (int64_t t = g.rank; t < (int64_t)m * nw; t += g.size) {
        const int i = (int)(t / nw), n = (int)(t - (int64_t)i * nw);
        const double ur = Uc[ux(n, i)], ui = Uc[ux(n, i) + 1];
        const double vr = gre[n] * ur + gim[n] * ui, vi = gre[n] * ui - gim[n] * ur;
        Aj[2 * t] = vr;
        Aj[2 * t + 1] = -vi;
        Bj[2 * t] = ur;
        Bj[2 * t + 1] = ui;
      }
    } else {
      const int nbk = (m + 3) / 4;
      const int ntile = nbk * (nbk + 1) / 2;
      const int ncap = kBwdLds / (4 * m);
      const int nc = ncap < nw ? (ncap > 0 ? ncap : 1) : nw;
      double* vs = wl;              // [nc][m] complex
      double* us = wl + 2 * nc * m; // [nc][m] complex
      for (int t0 = 0; t0 < ntile; t0 += g.size) {
        const int t = t0 + g.rank;
        const bool on = t < ntile;
        // tile (bi, bj), bi <= bj, row-major over the upper triangle
        int bi = 0, bj = 0;
        if (on) {
          const double q = 2.0 * nbk + 1.0;
          bi = (int)((q - sqrt(q * q - 8.0 * t)) * 0.5);
          while (bi > 0 && bi * nbk - bi * (bi - 1) / 2 > t) --bi;
          while ((bi + 1) * nbk - (bi + 1) * bi / 2 <= t) ++bi;
          bj = bi + (t - (bi * nbk - bi * (bi - 1) / 2));
        }
        double acc[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] = 0.0;
        for (int n0 = 0; n0 < nw; n0 += nc) {
          const int nn = nw - n0 < nc ? nw - n0 : nc;
          __syncthreads();
          for (int k = g.rank; k < nn * m; k += g.size) {
            const int i = k / nn, nl = k - i * nn;
            const int n = n0 + nl;
            const int o = nl * m + i;
            const double ur = Uc[ux(n, i)], ui = Uc[ux(n, i) + 1];
            const double cr = gre[n], ci = -gim[n];
            vs[2 * o] = cr * ur - ci * ui;
            vs[2 * o + 1] = cr * ui + ci * ur;
            us[2 * o] = ur;
            us[2 * o + 1] = ui;
          }
          __syncthreads();
          if (on) {
            for (int nl = 0; nl < nn; ++nl) {
              double vr[4], vi[4], wr[4], wi[4];
#pragma unroll
              for (int a = 0; a < 4; ++a) {
                const int i = bi * 4 + a < m ? bi * 4 + a : m - 1;
                const int j = bj * 4 + a < m ? bj * 4 + a : m - 1;
                vr[a] = vs[2 * (nl * m + i)];
                vi[a] = vs[2 * (nl * m + i) + 1];
                wr[a] = us[2 * (nl * m + j)];
                wi[a] = us[2 * (nl * m + j) + 1];
              }
#pragma unroll
              for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[a * 4 + c] = fma(vr[a], wr[c], fma(-vi[a], wi[c], acc[a * 4 + c]));
            }
          }
        }
        if (on) {
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int i = bi * 4 + a, j = bj * 4 + c;
              if (i < m && j < m && i <= j) {
                W[(int64_t)i * kcap + j] = acc[a * 4 + c];
                W[(int64_t)j * kcap + i] = acc[a * 4 + c];
              }
            }
        }
      }
      __syncthreads();
    }
    for (int i = g.rank; i < m; i += g.size) {
      double re = 0.0, im = 0.0;
      for (int n = 0; n < nw; ++n) {
        const double ur = Uc[ux(n, i)], ui = Uc[ux(n, i) + 1];
        re += gre[n] * ur + gim[n] * ui;
        im += gre[n] * ui - gim[n] * ur;
      }
      ub[2 * i] = re;
      ub[2 * i + 1] = im;
    }
    g.sync();
    // (c) P = Qt Re W (rows of the sector's states), (e) X = Qt ubar
    double* P = bs.P(nw);
    double* Xr = bs.vec(part == 0 ? 4 : 6);
    double* Xi = bs.vec(part == 0 ? 5 : 7);
    for (int64_t r = g.rank; r < Dt; r += g.size) {
      // eight entries' loads in flight at a time, summed in order
      const double* q = qt + r * kcap;
      double xr = 0.0, xi = 0.0;
      int i = 0;
      for (; i + 8 <= m; i += 8) {
        double qv[8], ur[8], ui[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          qv[t] = q[i + t];
          ur[t] = ub[2 * (i + t)];
          ui[t] = ub[2 * (i + t) + 1];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          xr += qv[t] * ur[t];
          xi += qv[t] * ui[t];
        }
      }
      for (; i < m; ++i) {
        xr += q[i] * ub[2 * i];
        xi += q[i] * ub[2 * i + 1];
      }
      Xr[r] = xr;
      Xi[r] = xi;
    }
    if (job >= 0) {
      g.sync();
      continue;
    }
    // P = Qt Re W with FP64 matrix cores (16 x 16 x 4), in groups of at most
    // one 32-column block per wave: within a group starting at column c0,
    // wave (wr, wc) accumulates the 32 x 32 block of P at rows
    // rb = R0 + 32 wr, columns c0 + 32 wc, as 2 x 2 tiles; the group's
    // columns of Re W are staged through LDS kc rows at a time.  Operand
    // layout of the instruction: lane l holds A[l % 16][l / 16] and
    // B[l / 16][l % 16], and its four results are D[l / 16 + 4 r][l % 16].
    // Krylov indices at or past m contribute zeros (Qt holds no data there).
    typedef double f64x4 __attribute__((ext_vector_type(4)));
    const int nwv = (int)(g.size >> 6);
    const int ncb_all = (m + 31) / 32;
    for (int cb0 = 0; cb0 < ncb_all; cb0 += nwv) {
      const int ncb = ncb_all - cb0 < nwv ? ncb_all - cb0 : nwv;
      const int c0 = 32 * cb0;
      const int wpr = nwv / ncb;
      const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
      const int lane = threadIdx.x & 63;
      const int wr = wave / ncb, wc = wave - wr * ncb;
      const bool won = wr < wpr;
      const int ldw = 32 * ncb;
      const int kc = (kBwdLds / ldw) / 4 * 4;
      for (int64_t R0 = 0; R0 < Dt; R0 += 32 * (int64_t)wpr) {
        const int64_t rb = R0 + 32 * (int64_t)wr;
        f64x4 acc[2][2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = f64x4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < m; k0 += kc) {
          const int kn = m - k0 < kc ? m - k0 : kc;
          __syncthreads();
          for (int t = g.rank; t < kc * ldw; t += g.size) {
            const int kk = t / ldw, c = t - kk * ldw;
            wl[t] = (kk < kn && c0 + c < m) ? W[(int64_t)(k0 + kk) * kcap + c0 + c] : 0.0;
          }
          __syncthreads();
          if (won && rb < Dt) {
            for (int ks = 0; ks < kn; ks += 4) {
              const int kk = ks + lane / 16;
              const bool kon = kk < kn;
              double av[2], bv[2];
#pragma unroll
              for (int ti = 0; ti < 2; ++ti) {
                const int64_t row = rb + 16 * ti + (lane % 16);
                const int64_t rowc = row < Dt ? row : Dt - 1;
                av[ti] = kon ? qt[rowc * kcap + k0 + kk] : 0.0;
              }
#pragma unroll
              for (int tj = 0; tj < 2; ++tj) bv[tj] = wl[kk * ldw + 32 * wc + 16 * tj + (lane % 16)];
#pragma unroll
              for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                  acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], bv[tj], acc[ti][tj], 0, 0, 0);
            }
          }
        }
        if (won) {
#pragma unroll
          for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int64_t row = rb + 16 * ti + lane / 16 + 4 * r;
                const int col = c0 + 32 * wc + 16 * tj + (lane % 16);
                if (row < Dt && col < m) P[row * kcap + col] = acc[ti][tj][r];
              }
        }
      }
      __syncthreads();
    }
    g.sync();
    // (d) diag of K, its trace, and the per-bond sums
    double* dg = bs.vec(8);
    double tr = 0.0;
    for (int64_t r = g.rank; r < Dt; r += g.size) {
      const double acc = row_dot(P + r * kcap, qt + r * kcap, m);
      dg[r] = acc;
      tr += acc;
    }
    tr = g.sum(tr);
    if (g.rank == 0) rec[part * kRecPart] = tr;
    g.sync();
    for (int k = 0; k < nb; ++k) {
      double se = 0.0, sv = 0.0;
      for (int64_t r = g.rank; r < Dt; r += g.size) {
        const int iu = (int)(r / nd), id = (int)(r - (int64_t)iu * nd);
        se += occ_k(T.st[iu], Dn.st[id], k) * dg[r];
        // the bond's up hop and dn hop from r (at most one each), their dots
        // with P's row r together; summed as one hop after the other
        int64_t t0 = -1, t1 = -1;
        double g0 = 0.0, g1 = 0.0;
        for (int e = 0; e < T.cnt[iu]; ++e) {
          const uint32_t t = (uint32_t)T.tgt[iu * kMaxBath + e];
          if ((int)((t >> 16) & 0xff) != k) continue;
          g0 = (t & 0x80000000u) ? -1.0 : 1.0;
          t0 = (int64_t)(t & 0xffff) * nd + id;
        }
        for (int e = 0; e < Dn.cnt[id]; ++e) {
          const uint32_t t = (uint32_t)Dn.tgt[id * kMaxBath + e];
          if ((int)((t >> 16) & 0xff) != k) continue;
          g1 = (t & 0x80000000u) ? -1.0 : 1.0;
          t1 = (int64_t)iu * nd + (t & 0xffff);
        }
        if (t0 < 0) {  // dn hop only: it takes the first slot
          t0 = t1;
          g0 = g1;
          t1 = -1;
        }
        if (t0 >= 0) {
          double d0, d1;
          row_dot2(P + r * kcap, qt + t0 * kcap, t1 >= 0 ? qt + t1 * kcap : nullptr, m, &d0, &d1);
          double sb = g0 * d0;
          if (t1 >= 0) sb += g1 * d1;
          sv += sb;
        }
      }
      se = g.sum(se);
      sv = g.sum(sv);
      if (g.rank == 0) {
        rec[part * kRecPart + 1 + k] = se;
        rec[part * kRecPart + 1 + kMaxBath + k] = sv;
      }
      g.sync();
    }
  }
  // (f) r = Re(c X + c+ Y) in the N sector, projected off psi
  double* zeta = bs.vec(0);
  double* res = bs.vec(1);
  double* pv = bs.vec(2);
  double* qv = bs.vec(3);
  const double* Xr = bs.vec(4);
  const double* Yr = bs.vec(6);
  const bool has_p = co[4 * kcap + 3] > 0, has_m = co[4 * kcap + 4] > 0;
  double pr = 0.0;
  for (int64_t k = g.rank; k < D; k += g.size) {
    const int iu = (int)(k / nd), id = (int)(k - (int64_t)iu * nd);
    const uint32_t st = U0.st[iu];
    double r = 0.0;
    if (!(st & 1u) && has_p) r += Xr[(int64_t)comb_rank(st | 1u) * nd + id];
    if ((st & 1u) && has_m) r += Yr[(int64_t)comb_rank(st ^ 1u) * nd + id];
    res[k] = r;
    pr += r * psi[k];
  }
  pr = g.sum(pr);
  g.sync();
  double rho = 0.0;
  for (int64_t k = g.rank; k < D; k += g.size) {
    const double r = res[k] - pr * psi[k];
    res[k] = r;
    pv[k] = r;
    zeta[k] = 0.0;
    rho += r * r;
  }
  rho = g.sum(rho);
  g.sync();
  // (g) projected conjugate gradients on (H - E0) over the complement of psi
  const double rho0 = rho;
  int it = 0;
  for (; it < bc.cg_max && rho > bc.cg_tol * bc.cg_tol * rho0 && rho > 0.0; ++it) {
    double aj, nn;
    lanczos_step<true>(g, p, U0, Dn, pv, bs.vec(8), qv, 1.0, 0.0, 0.0, true, e0, &aj, &nn);  // q = (H - E0) p
    double qp = 0.0, qpsi = 0.0;
    for (int64_t k = g.rank; k < D; k += g.size) {
      qp += qv[k] * pv[k];
      qpsi += qv[k] * psi[k];
    }
    qp = g.sum(qp);
    qpsi = g.sum(qpsi);
    const double alpha = rho / qp;
    double rn = 0.0;
    for (int64_t k = g.rank; k < D; k += g.size) {
      const double q = qv[k] - qpsi * psi[k];
      zeta[k] += alpha * pv[k];
      const double r = res[k] - alpha * q;
      res[k] = r;
      rn += r * r;
    }
    rn = g.sum(rn);
    g.sync();
    const double beta = rn / rho;
    rho = rn;
    for (int64_t k = g.rank; k < D; k += g.size) pv[k] = res[k] + beta * pv[k];
    g.sync();
  }
  // (h) dE0/dtheta and psi^T dH zeta, (i) the gradient
  for (int k = 0; k < nb; ++k) {
    double de = 0.0, te = 0.0, dv = 0.0, tv = 0.0;
    for (int64_t r = g.rank; r < D; r += g.size) {
      const int iu = (int)(r / nd), id = (int)(r - (int64_t)iu * nd);
      const double o = occ_k(U0.st[iu], Dn.st[id], k);
      de += o * psi[r] * psi[r];
      te += o * psi[r] * zeta[r];
      dv += bond_hops(U0, Dn, r, k, [&](int64_t aa, int64_t bb) { return psi[aa] * psi[bb]; });
      tv += bond_hops(U0, Dn, r, k, [&](int64_t aa, int64_t bb) { return psi[aa] * zeta[bb]; });
    }
    de = g.sum(de);
    te = g.sum(te);
    dv = g.sum(dv);
    tv = g.sum(tv);
    if (g.rank == 0) {
      double* const rt = rec + 2 * kRecPart;
      rt[k] = de;
      rt[kMaxBath + k] = te;
      rt[2 * kMaxBath + k] = dv;
      rt[3 * kMaxBath + k] = tv;
    }
  }
  if (g.rank == 0) bc.cg_iters[mi] = it;
  g.sync();
}

__global__ __launch_bounds__(kBlock) void k_batch_backward(BwdConfig bc) {
  __shared__ double red[32];
  __shared__ int popped;
  __shared__ double wl[kBwdLds];  // 64 KiB: a chunk of Re W
  __shared__ double s_rows[kLdsRows];
  __shared__ uint32_t s_tgt[kLdsDnBwd];
  __shared__ double s_val[kLdsDnBwd];
  __shared__ int32_t s_meta[2 + kLdsChunks];
  __shared__ uint32_t s_utgt[kLdsUpBwd];
  __shared__ double s_uval[kLdsUpBwd];
  __shared__ int32_t s_ucnt[kLdsUpRowsBwd];
  __shared__ int32_t s_umeta[4];
  Group<true> g{(int)threadIdx.x, (int)blockDim.x, red};
  g.lds = BlockLds{(lds_f64*)s_rows, kLdsRows,          (lds_u32*)s_tgt,  (lds_f64*)s_val,
                   kLdsDnBwd,        (lds_i32*)s_meta,  (lds_u32*)s_utgt, (lds_f64*)s_uval,
                   (lds_i32*)s_ucnt, kLdsUpBwd,         kLdsUpRowsBwd,    (lds_i32*)s_umeta};
  if (threadIdx.x == 0) s_umeta[0] = 0;
  g.has_lds = bc.staged != 0;
  g.wave_rows = bc.staged == 1;
  const BwdSlot bs{SlotRef{bc.ws + (int64_t)blockIdx.x * bc.slot_bytes, bc.lcap, bc.vcap, 0, bc.kcap}};
  while (true) {
    if (threadIdx.x == 0) popped = atomicAdd(bc.counter, 1);
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(popped);
    __syncthreads();
    if (q >= bc.num_models) break;
    backward_model(g, bc, q, bs, wl);
  }
}

// The backward's device-wide products C = A B^T (Re W = A B^T, then
// P = Qt Re W = Qt Re W^T, Re W being symmetric): one 64 x 64 tile of one
// product per block from the tile list, each of the four
// waves a 32 x 32 quarter as 2 x 2 v_mfma_f64_16x16x4 tiles, 16-wide K slabs
// staged in LDS (zero past the job's rows, columns and m).  Operand lanes:
// lane l gives A[l % 16][l / 16] and B^T[l / 16][l % 16] of a 16x4 / 4x16
// step; result element r of lane l is C[l / 16 + 4 r][l % 16].
__global__ __launch_bounds__(256) void k_bwd_gemm(const BwdGemm* __restrict__ gemms, const BwdTile* __restrict__ tiles,
                                                  const double* __restrict__ store_q, double* __restrict__ jstore) {
  __shared__ double As[64][16 + 1];
  __shared__ double Bs[64][16 + 1];
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  const BwdTile tl = tiles[blockIdx.x];
  const BwdGemm gm = gemms[tl.job];
  const double* __restrict__ A = (gm.a_in_q ? store_q : jstore) + gm.a_off;
  const double* __restrict__ B = jstore + gm.b_off;
  double* __restrict__ Cm = jstore + gm.c_off;
  const int64_t M = gm.M;
  const int N = gm.N, K = gm.K, lda = gm.lda, ldb = gm.ldb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t m0 = tl.r0;
  const int n0 = tl.c0;
  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int i = threadIdx.x; i < 64 * 16; i += 256) {
      const int row = i >> 4, kk = i & 15;
      const int kg = k0 + kk;
      As[row][kk] = (kg < K && m0 + row < M) ? A[(m0 + row) * lda + kg] : 0.0;
      Bs[row][kk] = (kg < K && n0 + row < N) ? B[(int64_t)(n0 + row) * ldb + kg] : 0.0;
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
        const int64_t row = m0 + wm + ti * 16 + lk + 4 * r;
        const int col = n0 + wn + tj * 16 + li;
        if (row < M && col < N) Cm[row * N + col] = acc[ti][tj][r];
      }
}

// <x, y> over n entries, lanes striding the entries, the same value on every
// lane (fixed order).
__device__ __forceinline__ double wave_dot(const double* __restrict__ x, const double* __restrict__ y, int n) {
  double s = 0.0;
  for (int j = (int)(threadIdx.x & 63); j < n; j += 64) s += x[j] * y[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// diag(K) = rows of P . rows of Qt, its trace, S_eps[k] = sum occ_k diag(K)
// and S_v[k] = sum over the hops of bond k of K[a, t] (with the hop signs)
// over one 64-row block of one job: a wave per row in turn, states and hops
// regenerated from the row index (the same hops as the spin lists hold);
// the block's sums, wave by wave in order, go to its partial record.
__global__ __launch_bounds__(256) void k_bwd_sums(BwdConfig bc) {
  __shared__ double red[4][kRecPart];
  const BwdBlock bl = bc.blocks[blockIdx.x];
  const BwdJob jb = bc.jobs[bl.job];
  const ModelParams& p = bc.params[jb.model];
  const int ns = p.nb + 1, nb = p.nb;
  const int n_t = p.n_up + (jb.part == 0 ? 1 : -1);
  const int nd = (int)binom_u(ns, p.n_dn);
  const int m = jb.m, kcap = bc.kcap;
  const double* __restrict__ P = bc.jstore + jb.p_off;
  const double* __restrict__ Q = bc.store_q + jb.q_off;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[kRecPart];  // uniform over the wave
#pragma unroll
  for (int t = 0; t < kRecPart; ++t) acc[t] = 0.0;
  for (int rr = wave; rr < 64; rr += 4) {
    const int64_t r = (int64_t)bl.r0 + rr;
    if (r >= jb.dt) break;
    const int it = (int)(r / nd), id = (int)(r - (int64_t)it * nd);
    const uint32_t st = comb_unrank(it, ns, n_t), sd = comb_unrank(id, ns, p.n_dn);
    const double* pr = P + r * m;
    const double dg = wave_dot(pr, Q + r * kcap, m);
    acc[0] += dg;
    for (int k = 0; k < nb; ++k) {
      acc[1 + k] += occ_k(st, sd, k) * dg;
      const uint32_t mk = 1u | (2u << k);
      double sv = 0.0;
      const uint32_t bu = st & mk;
      if (bu != 0 && bu != mk) {
        const double sg = hop_sign(st, bc.sign_width, 0, k + 1) < 0 ? -1.0 : 1.0;
        sv += sg * wave_dot(pr, Q + ((int64_t)comb_rank(st ^ mk) * nd + id) * kcap, m);
      }
      const uint32_t bd = sd & mk;
      if (bd != 0 && bd != mk) {
        const double sg = hop_sign(sd, bc.sign_width, 0, k + 1) < 0 ? -1.0 : 1.0;
        sv += sg * wave_dot(pr, Q + ((int64_t)it * nd + comb_rank(sd ^ mk)) * kcap, m);
      }
      acc[1 + kMaxBath + k] += sv;
    }
  }
  if (lane == 0)
    for (int t = 0; t < kRecPart; ++t) red[wave][t] = acc[t];
  __syncthreads();
  for (int t = threadIdx.x; t < kRecPart; t += blockDim.x)
    bc.partial[(int64_t)blockIdx.x * kRecPart + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// Gradient of every model from its record and, for its device-wide parts,
// the blocks' partial sums (in block order):
//   dL/deps_k = sum_parts s_part S_eps[k] + (trK_hole - trK_particle) dE0/deps_k - 2 psi dH/deps_k zeta,
// the same for v_k with S_v.
__global__ void k_bwd_combine(BwdConfig bc) {
  const int mi = blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= bc.num_models) return;
  const ModelParams& p = bc.params[mi];
  const double* rec = bc.rec + (int64_t)mi * kRec;
  const double* rt = rec + 2 * kRecPart;
  const double* co = bc.store_coef + (int64_t)mi * (4 * bc.kcap + 8);
  double tr[2] = {0.0, 0.0};
  for (int part = 0; part < 2; ++part) {
    if ((int)co[4 * bc.kcap + 3 + part] <= 0) continue;
    const int job = bc.job_of[2 * mi + part];
    if (job < 0) {
      tr[part] = rec[part * kRecPart];
    } else {
      const BwdJob jb = bc.jobs[job];
      double t = 0.0;
      for (int b = 0; b < jb.nblk; ++b) t += bc.partial[(int64_t)(jb.blk0 + b) * kRecPart];
      tr[part] = t;
    }
  }
  for (int k = 0; k < p.nb; ++k) {
    double se = 0.0, sv = 0.0;
    for (int part = 0; part < 2; ++part) {
      if ((int)co[4 * bc.kcap + 3 + part] <= 0) continue;
      const double sigma = part == 0 ? 1.0 : -1.0;
      const int job = bc.job_of[2 * mi + part];
      double e, v;
      if (job < 0) {
        e = rec[part * kRecPart + 1 + k];
        v = rec[part * kRecPart + 1 + kMaxBath + k];
      } else {
        const BwdJob jb = bc.jobs[job];
        e = 0.0;
        v = 0.0;
        for (int b = 0; b < jb.nblk; ++b) {
          e += bc.partial[(int64_t)(jb.blk0 + b) * kRecPart + 1 + k];
          v += bc.partial[(int64_t)(jb.blk0 + b) * kRecPart + 1 + kMaxBath + k];
        }
      }
      se += sigma * e;
      sv += sigma * v;
    }
    const double ba = tr[1] - tr[0];
    bc.grad_eps[(int64_t)mi * bc.max_bath + k] = se + ba * rt[k] - 2.0 * rt[kMaxBath + k];
    bc.grad_v[(int64_t)mi * bc.max_bath + k] = sv + ba * rt[2 * kMaxBath + k] - 2.0 * rt[3 * kMaxBath + k];
  }
}

__global__ void k_reset_counters(int32_t* c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

}  // namespace batch
}  // namespace ed

// ===========================================================================
// Host side
// ===========================================================================
using namespace ed;
using namespace ed::batch;

static int env_flag(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

struct ed_batch {
  int device = 0;
  int64_t num_models = 0;
  std::vector<ModelParams> params;
  std::vector<int64_t> dim;   // sector size per model
  std::vector<int64_t> dmax;  // largest of the three sectors per model
  std::vector<int32_t> order; // models by decreasing cost
  int32_t max_iter = 300, cf_steps = 64, check_every = 8, sign_width = 0;
  double tol = 1e-10, cf_tol = 1e-13;
  int32_t max_bath = 1;
  uint64_t seed = 0x5EED;
  // device
  ModelParams* d_params = nullptr;
  int32_t* d_queues = nullptr;  // [num_models]: wave queue then block queue
  int32_t* d_counters = nullptr;
  double *d_e0 = nullptr, *d_g = nullptr, *d_resid = nullptr, *d_wn = nullptr;
  int32_t* d_iters = nullptr;
  char *d_ws = nullptr;
  // adjoint backward (keep_basis)
  bool keep = false;
  int32_t kcap = 0;
  int64_t *d_off_psi = nullptr, *d_off_q = nullptr;
  // ground-state vector store: per-model offsets (within the model's chunk),
  // chunks of models in solve order (first model, count)
  double* d_gs_store = nullptr;
  int64_t* d_off_gs = nullptr;
  std::vector<int64_t> gs_need;                 // doubles per model
  std::vector<std::pair<int64_t, int64_t>> chunks;
  double *d_store_psi = nullptr, *d_store_q = nullptr, *d_store_coef = nullptr;
  double *d_gre = nullptr, *d_gim = nullptr, *d_grad_eps = nullptr, *d_grad_v = nullptr;
  int32_t* d_cg = nullptr;
  int32_t* d_bcounter = nullptr;
  // backward: parts whose K products run device-wide (tables and stores)
  std::vector<int64_t> h_off_q;
  int32_t* d_job_of = nullptr;
  BwdJob* d_jobs = nullptr;
  BwdGemm* d_gemms = nullptr;
  BwdTile* d_tiles = nullptr;
  BwdBlock* d_blocks = nullptr;
  double *d_jstore = nullptr, *d_rec = nullptr, *d_partial = nullptr;
  int64_t cap_jobs = 0, cap_gemms = 0, cap_tiles = 0, cap_blocks = 0, cap_jstore = 0;
  char* d_bws = nullptr;
  int64_t bws_bytes = 0;
  int32_t solved_nw = -1;  // frequencies of the last solve (the backward uses them)
  int64_t ws_bytes = 0;
  int32_t nw_cap = 0;
  // launch configuration
  int64_t wave_max_dim = -1;  // models with dim <= this run in wave mode (-1: not chosen)
  int32_t grid = 0, wave_blocks = 0;
  bool tuned = false;
  double tune_ms = 0.0;
  Config cfg{};
  // per chunk: wave and block queue bounds in d_queues
  std::vector<std::array<int32_t, 4>> chunk_q;
  // captured graph of (reset counters, solve) for repeated solves
  hipGraphExec_t graph = nullptr;
  hipStream_t graph_stream = nullptr;
};

static int64_t sector_dim(int ns, int nu, int nd) {
  return (int64_t)binom_u(ns, nu) * (int64_t)binom_u(ns, nd);
}

static void free_batch(ed_batch* b) {
  if (!b) return;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (void* p : {(void*)b->d_params, (void*)b->d_queues, (void*)b->d_counters, (void*)b->d_e0, (void*)b->d_g,
                  (void*)b->d_resid, (void*)b->d_wn, (void*)b->d_iters, (void*)b->d_ws, (void*)b->d_off_psi,
                  (void*)b->d_off_q, (void*)b->d_store_psi, (void*)b->d_store_q, (void*)b->d_store_coef,
                  (void*)b->d_gre, (void*)b->d_gim, (void*)b->d_grad_eps, (void*)b->d_grad_v, (void*)b->d_cg,
                  (void*)b->d_bcounter, (void*)b->d_bws, (void*)b->d_gs_store, (void*)b->d_off_gs, (void*)b->d_job_of,
                  (void*)b->d_jobs, (void*)b->d_gemms, (void*)b->d_tiles, (void*)b->d_blocks, (void*)b->d_jstore, (void*)b->d_rec,
                  (void*)b->d_partial})
    if (p) (void)hipFree(p);
  delete b;
}

template <typename T>
static int balloc(T** p, int64_t n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  const hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(ED_ERR_NOMEM, "hipMalloc of %lld bytes failed: %s", (long long)(sizeof(T) * n),
                     hipGetErrorString(e));
  }
  return ED_OK;
}

// Lays out queues and workspace for a wave/block split at wave_max_dim and a
// grid size; (re)allocates the workspace when it grows.
static int configure(ed_batch* b, int64_t wave_max_dim, int grid_blocks, int wave_blocks, hipStream_t s) {
  std::vector<int32_t> qw, qb;
  int64_t vw = 1, vb = 1;
  int32_t lw = 1, lb = 1;
  // queues of every chunk, [wave | block] one chunk after another
  std::vector<int32_t> q;
  b->chunk_q.clear();
  for (const auto& ch : b->chunks) {
    std::array<int32_t, 4> cq{};
    cq[0] = (int32_t)q.size();
    for (int64_t k = ch.first; k < ch.first + ch.second; ++k)
      if (b->dim[b->order[k]] <= wave_max_dim) q.push_back(b->order[k]);
    cq[1] = (int32_t)q.size() - cq[0];
    cq[2] = (int32_t)q.size();
    for (int64_t k = ch.first; k < ch.first + ch.second; ++k)
      if (b->dim[b->order[k]] > wave_max_dim) q.push_back(b->order[k]);
    cq[3] = (int32_t)q.size() - cq[2];
    b->chunk_q.push_back(cq);
  }
  for (int32_t mi : b->order) {
    const ModelParams& p = b->params[mi];
    const int ns = p.nb + 1;
    int32_t lcap = 1;
    for (int d = -1; d <= 1; ++d) lcap = std::max<int32_t>(lcap, (int32_t)binom_u(ns, std::max(0, p.n_up + d)));
    lcap = std::max<int32_t>(lcap, (int32_t)binom_u(ns, p.n_dn));
    if (b->dim[mi] <= wave_max_dim) {
      qw.push_back(mi);
      vw = std::max(vw, b->dmax[mi]);
      lw = std::max(lw, lcap);
    } else {
      qb.push_back(mi);
      vb = std::max(vb, b->dmax[mi]);
      lb = std::max(lb, lcap);
    }
  }
  if (qw.empty()) wave_blocks = 0;
  if (qb.empty()) wave_blocks = grid_blocks;
  if (!qw.empty() && wave_blocks == 0) wave_blocks = 1;
  if (!qb.empty() && wave_blocks >= grid_blocks) grid_blocks = wave_blocks + 1;
  const int64_t sw = SlotRef{nullptr, lw, vw, b->max_iter, b->cf_steps}.bytes();
  const int64_t sb = SlotRef{nullptr, lb, vb, b->max_iter, b->cf_steps}.bytes();
  const int64_t n_wave_slots = (int64_t)wave_blocks * (kBlockSolveMax / 64);
  const int64_t n_block_slots = grid_blocks - wave_blocks;
  const int64_t bytes = n_wave_slots * sw + n_block_slots * sb + 256;
  if (bytes > b->ws_bytes) {
    if (b->d_ws) (void)hipFree(b->d_ws);
    b->d_ws = nullptr;
    b->ws_bytes = 0;
    ED_TRY(balloc(&b->d_ws, bytes));
    b->ws_bytes = bytes;
  }
  ED_HIP(hipMemcpyAsync(b->d_queues, q.data(), sizeof(int32_t) * q.size(), hipMemcpyHostToDevice, s));
  ED_HIP(hipStreamSynchronize(s));
  Config& c = b->cfg;
  c.num_models = b->num_models;
  c.params = b->d_params;
  c.queue_wave = b->d_queues + b->chunk_q[0][0];
  c.n_wave = b->chunk_q[0][1];
  c.queue_block = b->d_queues + b->chunk_q[0][2];
  c.n_block = b->chunk_q[0][3];
  c.counters = b->d_counters;
  c.wave_blocks = wave_blocks;
  c.ws_wave = b->d_ws;
  c.ws_block = b->d_ws + n_wave_slots * sw;
  c.slot_bytes_wave = sw;
  c.slot_bytes_block = sb;
  c.list_cap_wave = lw;
  c.list_cap_block = lb;
  c.vec_cap_wave = vw;
  c.vec_cap_block = vb;
  c.max_iter = b->max_iter;
  c.cf_steps = b->cf_steps;
  c.check_every = b->check_every;
  c.sign_width = b->sign_width;
  c.staged = env_flag("CMPY_AMD_BATCH_STAGED", 1);
  c.tol = b->tol;
  c.cf_tol = b->cf_tol;
  c.seed = b->seed;
  c.e0 = b->d_e0;
  c.g = b->d_g;
  c.iters = b->d_iters;
  c.resid = b->d_resid;
  c.keep = b->keep ? 1 : 0;
  c.kcap = b->kcap;
  c.off_psi = b->d_off_psi;
  c.off_q = b->d_off_q;
  c.store_psi = b->d_store_psi;
  c.store_q = b->d_store_q;
  c.store_coef = b->d_store_coef;
  c.gs_store = b->d_gs_store;
  c.off_gs = b->d_off_gs;
  b->wave_max_dim = wave_max_dim;
  b->grid = grid_blocks;
  b->wave_blocks = wave_blocks;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return ED_OK;
}

// Per chunk of models (one in all but very large batches): the ground-state
// launch, then the fraction launch; then G of every model.
static int launch(ed_batch* b, hipStream_t s) {
  for (const auto& cq : b->chunk_q) {
    Config c = b->cfg;
    c.queue_wave = b->d_queues + cq[0];
    c.n_wave = cq[1];
    c.queue_block = b->d_queues + cq[2];
    c.n_block = cq[3];
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, b->d_counters);
    hipLaunchKernelGGL(k_batch_solve<0>, dim3((unsigned)b->grid), dim3(kBlockGround), 0, s, c);
    hipLaunchKernelGGL(k_batch_solve<1>, dim3((unsigned)b->grid), dim3(kBlockFrac), 0, s, c);
  }
  const int64_t nt = b->num_models * (int64_t)b->cfg.nw;
  if (nt > 0) hipLaunchKernelGGL(k_batch_g, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, b->cfg);
  ED_HIP(hipGetLastError());
  return ED_OK;
}

static double time_launch(ed_batch* b, hipStream_t s, int* rc) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    *rc = set_error(ED_ERR_HIP, "hipEventCreate failed");
    return 0.0;
  }
  (void)hipEventRecord(e0, s);
  *rc = launch(b, s);
  (void)hipEventRecord(e1, s);
  float ms = 0.0f;
  if (*rc == ED_OK) {
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      *rc = set_error(ED_ERR_HIP, "batch solve failed: %s", hipGetErrorString(hipGetLastError()));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

// Autotuner: tries wave/block splits and grid sizes on the batch itself and
// keeps the fastest (results are recomputed by the caller's solve).
static int autotune(ed_batch* b, hipStream_t s) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  std::vector<int64_t> dims(b->dim);
  std::sort(dims.begin(), dims.end());
  dims.erase(std::unique(dims.begin(), dims.end()), dims.end());
  std::vector<int64_t> splits = {-1};  // all block mode
  for (int64_t d : dims)
    if (d <= 65536) splits.push_back(d);
  double best = 1e300;
  int64_t best_split = -1;
  int best_grid = cus, best_wave = 0;
  for (int64_t split : splits) {
    int64_t n_small = 0, n_large = 0;
    for (int64_t d : b->dim) (d <= split ? n_small : n_large)++;
    for (int per_cu : {1, 2}) {
      const int grid = cus * per_cu;
      // wave blocks in proportion to the small models' share of the work
      double w_small = 0.0, w_all = 0.0;
      for (size_t i = 0; i < b->dim.size(); ++i) {
        const double c = (double)b->dim[i];
        w_all += c;
        if (b->dim[i] <= split) w_small += c;
      }
      int wb = n_small == 0 ? 0 : std::max(1, (int)std::lround(grid * w_small / std::max(w_all, 1.0)));
      if (n_large > 0) wb = std::min(wb, grid - std::max<int>(1, (int)std::min<int64_t>(n_large, grid / 2)));
      wb = std::max(wb, n_small > 0 ? 1 : 0);
      ED_TRY(configure(b, split, grid, wb, s));
      int rc = ED_OK;
      const double ms = time_launch(b, s, &rc);
      ED_TRY(rc);
      if (ms < best) {
        best = ms;
        best_split = split;
        best_grid = b->grid;
        best_wave = b->wave_blocks;
      }
    }
  }
  ED_TRY(configure(b, best_split, best_grid, best_wave, s));
  b->tuned = true;
  b->tune_ms = best;
  return ED_OK;
}

extern "C" int ed_batch_create(ed_batch** out, const ed_batch_desc* d, void* stream) {
  clear_error();
  ED_REQUIRE(out && d, "NULL argument");
  *out = nullptr;
  ED_REQUIRE(d->num_models > 0, "num_models must be positive");
  ED_REQUIRE(d->num_bath && d->n_up && d->n_dn && d->u && d->eps_imp && d->eps_bath && d->v, "NULL model array");
  ED_REQUIRE(d->max_bath >= 1, "max_bath must be positive");
  ED_REQUIRE(d->max_iter >= 1 && d->max_iter <= 2000, "max_iter out of range (1..2000)");
  ED_REQUIRE(d->cf_steps >= 1 && d->cf_steps <= 2000, "cf_steps out of range (1..2000)");
  ED_REQUIRE(d->tol >= 0.0 && d->cf_tol >= 0.0, "tol and cf_tol must be non-negative");
  ed_batch* b = new ed_batch();
  (void)hipGetDevice(&b->device);
  b->num_models = d->num_models;
  b->max_iter = d->max_iter;
  b->cf_steps = d->cf_steps;
  b->check_every = d->check_every > 0 ? d->check_every : 8;
  b->sign_width = d->sign_width;
  b->tol = d->tol;
  b->cf_tol = d->cf_tol;
  b->seed = d->seed;
  b->params.resize(d->num_models);
  b->dim.resize(d->num_models);
  b->dmax.resize(d->num_models);
  for (int64_t i = 0; i < d->num_models; ++i) {
    ModelParams& p = b->params[i];
    std::memset(&p, 0, sizeof(p));
    p.nb = d->num_bath[i];
    p.n_up = d->n_up[i];
    p.n_dn = d->n_dn[i];
    const int ns = p.nb + 1;
    if (p.nb < 1 || p.nb > kMaxBath || p.nb > d->max_bath || p.n_up < 0 || p.n_up > ns || p.n_dn < 0 ||
        p.n_dn > ns) {
      free_batch(b);
      return set_error(ED_ERR_INVALID, "model %lld: num_bath must be 1..%d (and <= max_bath) with 0 <= n <= sites",
                       (long long)i, kMaxBath);
    }
    if (d->sign_width < 0 || d->sign_width > 32) {
      free_batch(b);
      return set_error(ED_ERR_INVALID, "sign_width must be 0..32");
    }
    p.u = d->u[i];
    p.eps_imp = d->eps_imp[i];
    bool finite = std::isfinite(p.u) && std::isfinite(p.eps_imp);
    for (int k = 0; k < p.nb; ++k) {
      p.eps[k] = d->eps_bath[i * d->max_bath + k];
      p.v[k] = d->v[i * d->max_bath + k];
      finite = finite && std::isfinite(p.eps[k]) && std::isfinite(p.v[k]);
    }
    if (!finite) {
      free_batch(b);
      return set_error(ED_ERR_INVALID, "model %lld: parameters must be finite", (long long)i);
    }
    b->dim[i] = sector_dim(ns, p.n_up, p.n_dn);
    int64_t dm = b->dim[i];
    if (p.n_up + 1 <= ns) dm = std::max(dm, sector_dim(ns, p.n_up + 1, p.n_dn));
    if (p.n_up >= 1) dm = std::max(dm, sector_dim(ns, p.n_up - 1, p.n_dn));
    b->dmax[i] = dm;
    if (b->dim[i] > (int64_t)1 << 24) {
      free_batch(b);
      return set_error(ED_ERR_LIMIT, "model %lld: sector of %lld states exceeds the batched solver's limit",
                       (long long)i, (long long)b->dim[i]);
    }
  }
  b->kcap = d->cf_steps;
  b->order.resize(d->num_models);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int32_t a, int32_t c) { return b->dim[a] > b->dim[c]; });
  hipStream_t s = as_stream(stream);
  int rc = ED_OK;
  if ((rc = balloc(&b->d_params, d->num_models)) != ED_OK || (rc = balloc(&b->d_queues, d->num_models)) != ED_OK ||
      (rc = balloc(&b->d_counters, 4)) != ED_OK || (rc = balloc(&b->d_e0, d->num_models)) != ED_OK ||
      (rc = balloc(&b->d_iters, 3 * d->num_models)) != ED_OK || (rc = balloc(&b->d_resid, 2 * d->num_models)) != ED_OK ||
      (rc = balloc(&b->d_store_coef, d->num_models * (4 * (int64_t)d->cf_steps + 8))) != ED_OK) {
    free_batch(b);
    return rc;
  }
  if (hipMemcpyAsync(b->d_params, b->params.data(), sizeof(ModelParams) * d->num_models, hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    free_batch(b);
    return set_error(ED_ERR_HIP, "upload of model parameters failed");
  }
  if (d->keep_basis) {
    if (d->cf_steps > kBwdMaxSteps) {
      free_batch(b);
      return set_error(ED_ERR_INVALID, "keep_basis needs cf_steps <= %d", kBwdMaxSteps);
    }
    // per-model stores for the adjoint backward
    b->keep = true;
    const int64_t M = d->num_models;
    std::vector<int64_t> off_psi(M), off_q(2 * M);
    int64_t np = 0, nq = 0;
    for (int64_t i = 0; i < M; ++i) {
      const ModelParams& p = b->params[i];
      const int ns = p.nb + 1;
      off_psi[i] = np;
      np += b->dim[i];
      off_q[2 * i] = nq;
      if (p.n_up + 1 <= ns) nq += sector_dim(ns, p.n_up + 1, p.n_dn) * b->kcap;
      off_q[2 * i + 1] = nq;
      if (p.n_up >= 1) nq += sector_dim(ns, p.n_up - 1, p.n_dn) * b->kcap;
    }
    if ((rc = balloc(&b->d_off_psi, M)) != ED_OK || (rc = balloc(&b->d_off_q, 2 * M)) != ED_OK ||
        (rc = balloc(&b->d_store_psi, np)) != ED_OK || (rc = balloc(&b->d_store_q, nq)) != ED_OK ||
        (rc = balloc(&b->d_grad_eps, M * d->max_bath)) != ED_OK || (rc = balloc(&b->d_grad_v, M * d->max_bath)) != ED_OK ||
        (rc = balloc(&b->d_cg, M)) != ED_OK || (rc = balloc(&b->d_bcounter, 1)) != ED_OK ||
        (rc = balloc(&b->d_rec, M * kRec)) != ED_OK || (rc = balloc(&b->d_job_of, 2 * M)) != ED_OK) {
      free_batch(b);
      return rc;
    }
    b->h_off_q = off_q;
    if (hipMemcpyAsync(b->d_off_psi, off_psi.data(), sizeof(int64_t) * M, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->d_off_q, off_q.data(), sizeof(int64_t) * 2 * M, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      free_batch(b);
      return set_error(ED_ERR_HIP, "upload of store offsets failed");
    }
  }
  // ground-state vector store: (min(max_iter, D) + 1) D doubles per model,
  // chunks of models (in solve order) within 60% of the free device memory
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)16 << 30;
    const int64_t budget = (int64_t)(0.6 * (double)free_b / sizeof(double));
    const int64_t M = d->num_models;
    b->gs_need.resize(M);
    for (int64_t i = 0; i < M; ++i) b->gs_need[i] = (std::min<int64_t>(b->max_iter, b->dim[i]) + 1) * b->dim[i];
    std::vector<int64_t> off(M, 0);
    int64_t first = 0, used = 0, largest = 0;
    for (int64_t k = 0; k < M; ++k) {
      const int64_t need = b->gs_need[b->order[k]];
      if (need > budget) {
        free_batch(b);
        return set_error(ED_ERR_NOMEM, "model %lld: ground-state store of %lld bytes exceeds the free device memory",
                         (long long)b->order[k], (long long)(need * 8));
      }
      if (used + need > budget && k > first) {
        b->chunks.emplace_back(first, k - first);
        largest = std::max(largest, used);
        first = k;
        used = 0;
      }
      off[b->order[k]] = used;
      used += need;
    }
    b->chunks.emplace_back(first, M - first);
    largest = std::max(largest, used);
    if ((rc = balloc(&b->d_gs_store, largest)) != ED_OK || (rc = balloc(&b->d_off_gs, M)) != ED_OK) {
      free_batch(b);
      return rc;
    }
    if (hipMemcpyAsync(b->d_off_gs, off.data(), sizeof(int64_t) * M, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      free_batch(b);
      return set_error(ED_ERR_HIP, "upload of store offsets failed");
    }
  }
  b->max_bath = d->max_bath;
  *out = b;
  return ED_OK;
}

extern "C" int ed_batch_destroy(ed_batch* b) {
  clear_error();
  free_batch(b);
  return ED_OK;
}

extern "C" int ed_batch_config(ed_batch* b, int64_t wave_max_dim, int32_t grid_blocks, int32_t wave_blocks,
                               void* stream) {
  clear_error();
  ED_REQUIRE(b, "NULL batch");
  ED_REQUIRE(grid_blocks >= 1 && grid_blocks <= 65536, "grid_blocks out of range");
  ED_REQUIRE(wave_blocks >= 0 && wave_blocks <= grid_blocks, "wave_blocks out of range");
  b->tuned = true;
  return configure(b, wave_max_dim, grid_blocks, wave_blocks, as_stream(stream));
}

extern "C" int ed_batch_solve(ed_batch* b, const double* wn, int32_t nw, double* e0, double* g, int32_t* iters,
                              double* resid, double* elapsed_ms, void* stream) {
  clear_error();
  ED_REQUIRE(b, "NULL batch");
  ED_REQUIRE(nw >= 0 && (nw == 0 || wn), "bad frequency list");
  hipStream_t s = as_stream(stream);
  if (nw > b->nw_cap) {
    if (b->d_g) (void)hipFree(b->d_g);
    if (b->d_wn) (void)hipFree(b->d_wn);
    b->d_g = nullptr;
    b->d_wn = nullptr;
    b->nw_cap = 0;
    ED_TRY(balloc(&b->d_g, 2 * b->num_models * (int64_t)nw));
    ED_TRY(balloc(&b->d_wn, nw));
    b->nw_cap = nw;
    if (b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  double wmin = 1.0;
  for (int32_t k = 0; k < nw; ++k) {
    ED_REQUIRE(std::isfinite(wn[k]) && wn[k] != 0.0, "frequencies must be finite and non-zero");
    wmin = k == 0 ? fabs(wn[k]) : std::min(wmin, fabs(wn[k]));
  }
  if (nw > 0) ED_HIP(hipMemcpyAsync(b->d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice, s));
  if (!b->tuned) {
    b->cfg.g = b->d_g;  // the tuner's configure() calls set the rest of cfg
    b->cfg.wn = b->d_wn;
    b->cfg.nw = nw;
    b->cfg.wmin = wmin;
    ED_TRY(autotune(b, s));
  }
  if (b->cfg.nw != nw || b->cfg.g != b->d_g || b->cfg.wn != b->d_wn || b->cfg.wmin != wmin) {
    b->cfg.g = b->d_g;
    b->cfg.wn = b->d_wn;
    b->cfg.nw = nw;
    if (b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  b->cfg.g = b->d_g;
  b->cfg.wn = b->d_wn;
  b->cfg.nw = nw;
  b->cfg.wmin = wmin;
  // capture (reset counters, solve) once per configuration; replay afterwards
  if (!b->graph || b->graph_stream != s) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    hipStream_t cs;
    ED_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    int rc = ED_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return set_error(ED_ERR_HIP, "hipStreamBeginCapture failed");
    }
    rc = launch(b, cs);
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    ED_TRY(rc);
    if (ec != hipSuccess) return set_error(ED_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      b->graph = nullptr;
      return set_error(ED_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    b->graph_stream = s;
  }
  hipEvent_t ev0, ev1;
  ED_HIP(hipEventCreate(&ev0));
  ED_HIP(hipEventCreate(&ev1));
  (void)hipEventRecord(ev0, s);
  const hipError_t el = hipGraphLaunch(b->graph, s);
  (void)hipEventRecord(ev1, s);
  if (el != hipSuccess) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return set_error(ED_ERR_HIP, "hipGraphLaunch failed: %s", hipGetErrorString(el));
  }
  const hipError_t es = hipEventSynchronize(ev1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (es != hipSuccess) return set_error(ED_ERR_HIP, "batch solve failed: %s", hipGetErrorString(es));
  if (elapsed_ms) *elapsed_ms = ms;
  b->solved_nw = nw;
  const int64_t M = b->num_models;
  if (e0) ED_HIP(hipMemcpyAsync(e0, b->d_e0, sizeof(double) * M, hipMemcpyDeviceToHost, s));
  if (g && nw > 0) ED_HIP(hipMemcpyAsync(g, b->d_g, sizeof(double) * 2 * M * nw, hipMemcpyDeviceToHost, s));
  if (iters) ED_HIP(hipMemcpyAsync(iters, b->d_iters, sizeof(int32_t) * 3 * M, hipMemcpyDeviceToHost, s));
  if (resid) ED_HIP(hipMemcpyAsync(resid, b->d_resid, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  return ED_OK;
}

extern "C" int ed_batch_info(const ed_batch* b, int64_t* wave_max_dim, int32_t* grid_blocks, int32_t* wave_blocks,
                             double* tune_ms) {
  clear_error();
  ED_REQUIRE(b, "NULL batch");
  if (wave_max_dim) *wave_max_dim = b->wave_max_dim;
  if (grid_blocks) *grid_blocks = b->grid;
  if (wave_blocks) *wave_blocks = b->wave_blocks;
  if (tune_ms) *tune_ms = b->tune_ms;
  return ED_OK;
}

__global__ void k_reset_one(int32_t* c) {
  if (threadIdx.x == 0) *c = 0;
}

extern "C" int ed_batch_backward(ed_batch* b, const double* dl_dre, const double* dl_dim, double* grad_eps,
                                 double* grad_v, int32_t* cg_iters, double* elapsed_ms, void* stream) {
  clear_error();
  ED_REQUIRE(b, "NULL batch");
  ED_REQUIRE(b->keep, "the batch was created without keep_basis");
  ED_REQUIRE(b->solved_nw > 0, "call ed_batch_solve before ed_batch_backward");
  ED_REQUIRE(dl_dre && dl_dim, "NULL gradient input");
  hipStream_t s = as_stream(stream);
  const int64_t M = b->num_models;
  const int32_t nw = b->solved_nw;
  if (!b->d_gre) {
    ED_TRY(balloc(&b->d_gre, M * nw));
    ED_TRY(balloc(&b->d_gim, M * nw));
  }
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  const int grid = (int)std::min<int64_t>(M, 2 * (int64_t)cus);
  int32_t lcap = 1;
  int64_t vcap = 1;
  for (int64_t i = 0; i < M; ++i) {
    const ModelParams& p = b->params[i];
    const int ns = p.nb + 1;
    for (int d = -1; d <= 1; ++d) lcap = std::max<int32_t>(lcap, (int32_t)binom_u(ns, std::max(0, p.n_up + d)));
    lcap = std::max<int32_t>(lcap, (int32_t)binom_u(ns, p.n_dn));
    vcap = std::max(vcap, b->dmax[i]);
  }
  const int64_t slot = BwdSlot{SlotRef{nullptr, lcap, vcap, 0, b->kcap}}.bytes(nw);
  if (slot * grid > b->bws_bytes) {
    if (b->d_bws) (void)hipFree(b->d_bws);
    b->d_bws = nullptr;
    b->bws_bytes = 0;
    ED_TRY(balloc(&b->d_bws, slot * grid));
    b->bws_bytes = slot * grid;
  }
  ED_HIP(hipMemcpyAsync(b->d_gre, dl_dre, sizeof(double) * M * nw, hipMemcpyHostToDevice, s));
  ED_HIP(hipMemcpyAsync(b->d_gim, dl_dim, sizeof(double) * M * nw, hipMemcpyHostToDevice, s));
  // parts with at least split_m continued-fraction steps run their K
  // products device-wide (CMPY_AMD_BWD_SPLIT_M; the per-model workgroup
  // would spend O(D m^2) on them alone)
  const int split_m = env_flag("CMPY_AMD_BWD_SPLIT_M", kBwdSplitM);
  std::vector<int32_t> iters(3 * M);
  ED_HIP(hipMemcpy(iters.data(), b->d_iters, sizeof(int32_t) * 3 * M, hipMemcpyDeviceToHost));
  std::vector<int32_t> job_of(2 * M, -1);
  std::vector<BwdJob> jobs;
  std::vector<BwdGemm> gemms;                 // Re W products, then P products
  std::vector<BwdTile> tiles_w, tiles;        // their tiles
  std::vector<BwdBlock> blocks;
  int64_t jst = 0;
  for (int64_t i = 0; i < M; ++i) {
    const ModelParams& p = b->params[i];
    const int ns = p.nb + 1;
    for (int part = 0; part < 2; ++part) {
      const int m = iters[3 * i + 1 + part];
      const int n_t = p.n_up + (part == 0 ? 1 : -1);
      if (m <= 0 || m < split_m || n_t < 0 || n_t > ns) continue;
      BwdJob jb{};
      jb.model = (int32_t)i;
      jb.part = part;
      jb.m = m;
      jb.dt = (int32_t)sector_dim(ns, n_t, p.n_dn);
      jb.q_off = b->h_off_q[2 * i + part];
      jb.a_off = jst;
      jst += 2ll * m * nw;
      jb.b_off = jst;
      jst += 2ll * m * nw;
      jb.w_off = jst;
      jst += (int64_t)m * m;
      jb.p_off = jst;
      jst += (int64_t)jb.dt * m;
      jb.blk0 = (int32_t)blocks.size();
      jb.nblk = (jb.dt + 63) / 64;
      const int32_t j = (int32_t)jobs.size();
      for (int32_t r0 = 0; r0 < jb.dt; r0 += 64) blocks.push_back(BwdBlock{j, r0});
      job_of[2 * i + part] = j;
      jobs.push_back(jb);
    }
  }
  // Re W = A B^T (m x m, K = 2 nw) for every job, then P = Qt Re W^T
  const int32_t nj = (int32_t)jobs.size();
  for (int32_t j = 0; j < nj; ++j) {
    const BwdJob& jb = jobs[j];
    gemms.push_back(BwdGemm{jb.a_off, jb.b_off, jb.w_off, 2 * nw, 2 * nw, jb.m, jb.m, 2 * nw, 0});
    for (int32_t r0 = 0; r0 < jb.m; r0 += 64)
      for (int32_t c0 = 0; c0 < jb.m; c0 += 64) tiles_w.push_back(BwdTile{j, r0, c0, 0});
  }
  for (int32_t j = 0; j < nj; ++j) {
    const BwdJob& jb = jobs[j];
    gemms.push_back(BwdGemm{jb.q_off, jb.w_off, jb.p_off, b->kcap, jb.m, jb.dt, jb.m, jb.m, 1});
    for (int32_t r0 = 0; r0 < jb.dt; r0 += 64)
      for (int32_t c0 = 0; c0 < jb.m; c0 += 64) tiles.push_back(BwdTile{nj + j, r0, c0, 0});
  }
  auto grow = [](auto** ptr, int64_t* cap, int64_t n) -> int {
    if (n <= *cap) return ED_OK;
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    *cap = 0;
    ED_TRY(balloc(ptr, n));
    *cap = n;
    return ED_OK;
  };
  ED_TRY(grow(&b->d_jobs, &b->cap_jobs, (int64_t)jobs.size()));
  ED_TRY(grow(&b->d_gemms, &b->cap_gemms, (int64_t)gemms.size()));
  ED_TRY(grow(&b->d_tiles, &b->cap_tiles, (int64_t)(tiles_w.size() + tiles.size())));
  if ((int64_t)blocks.size() > b->cap_blocks) {
    if (b->d_partial) (void)hipFree(b->d_partial);
    b->d_partial = nullptr;
    ED_TRY(grow(&b->d_blocks, &b->cap_blocks, (int64_t)blocks.size()));
    ED_TRY(balloc(&b->d_partial, b->cap_blocks * kRecPart));
  }
  ED_TRY(grow(&b->d_jstore, &b->cap_jstore, jst));
  ED_HIP(hipMemcpy(b->d_job_of, job_of.data(), sizeof(int32_t) * 2 * M, hipMemcpyHostToDevice));
  if (!jobs.empty()) {
    ED_HIP(hipMemcpy(b->d_jobs, jobs.data(), sizeof(BwdJob) * jobs.size(), hipMemcpyHostToDevice));
    ED_HIP(hipMemcpy(b->d_gemms, gemms.data(), sizeof(BwdGemm) * gemms.size(), hipMemcpyHostToDevice));
    ED_HIP(hipMemcpy(b->d_tiles, tiles_w.data(), sizeof(BwdTile) * tiles_w.size(), hipMemcpyHostToDevice));
    ED_HIP(hipMemcpy(b->d_tiles + tiles_w.size(), tiles.data(), sizeof(BwdTile) * tiles.size(),
                     hipMemcpyHostToDevice));
    ED_HIP(hipMemcpy(b->d_blocks, blocks.data(), sizeof(BwdBlock) * blocks.size(), hipMemcpyHostToDevice));
  }
  BwdConfig bc{};
  bc.params = b->d_params;
  bc.num_models = (int32_t)M;
  bc.counter = b->d_bcounter;
  bc.ws = b->d_bws;
  bc.slot_bytes = slot;
  bc.lcap = lcap;
  bc.vcap = vcap;
  bc.kcap = b->kcap;
  bc.nw = nw;
  bc.sign_width = b->sign_width;
  bc.staged = env_flag("CMPY_AMD_BATCH_STAGED", 1);
  bc.max_bath = b->max_bath;
  bc.wn = b->d_wn;
  bc.gre = b->d_gre;
  bc.gim = b->d_gim;
  bc.off_psi = b->d_off_psi;
  bc.off_q = b->d_off_q;
  bc.store_psi = b->d_store_psi;
  bc.store_q = b->d_store_q;
  bc.store_coef = b->d_store_coef;
  bc.cg_max = 4000;
  bc.cg_tol = 1e-12;
  bc.grad_eps = b->d_grad_eps;
  bc.grad_v = b->d_grad_v;
  bc.cg_iters = b->d_cg;
  bc.job_of = b->d_job_of;
  bc.jobs = b->d_jobs;
  bc.blocks = b->d_blocks;
  bc.jstore = b->d_jstore;
  bc.rec = b->d_rec;
  bc.partial = b->d_partial;
  hipEvent_t e0, e1;
  ED_HIP(hipEventCreate(&e0));
  ED_HIP(hipEventCreate(&e1));
  (void)hipEventRecord(e0, s);
  hipLaunchKernelGGL(k_reset_one, dim3(1), dim3(64), 0, s, b->d_bcounter);
  hipLaunchKernelGGL(k_batch_backward, dim3((unsigned)grid), dim3(kBlock), 0, s, bc);
  if (!jobs.empty()) {
    hipLaunchKernelGGL(k_bwd_gemm, dim3((unsigned)tiles_w.size()), dim3(256), 0, s, (const BwdGemm*)b->d_gemms,
                       (const BwdTile*)b->d_tiles, (const double*)b->d_store_q, b->d_jstore);
    hipLaunchKernelGGL(k_bwd_gemm, dim3((unsigned)tiles.size()), dim3(256), 0, s, (const BwdGemm*)b->d_gemms,
                       (const BwdTile*)(b->d_tiles + tiles_w.size()), (const double*)b->d_store_q, b->d_jstore);
    hipLaunchKernelGGL(k_bwd_sums, dim3((unsigned)blocks.size()), dim3(256), 0, s, bc);
  }
  hipLaunchKernelGGL(k_bwd_combine, dim3((unsigned)((M + 63) / 64)), dim3(64), 0, s, bc);
  const hipError_t el = hipGetLastError();
  (void)hipEventRecord(e1, s);
  const hipError_t es = hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (el != hipSuccess || es != hipSuccess)
    return set_error(ED_ERR_HIP, "batch backward failed: %s", hipGetErrorString(el != hipSuccess ? el : es));
  if (elapsed_ms) *elapsed_ms = ms;
  if (grad_eps)
    ED_HIP(hipMemcpyAsync(grad_eps, b->d_grad_eps, sizeof(double) * M * b->max_bath, hipMemcpyDeviceToHost, s));
  if (grad_v) ED_HIP(hipMemcpyAsync(grad_v, b->d_grad_v, sizeof(double) * M * b->max_bath, hipMemcpyDeviceToHost, s));
  if (cg_iters) ED_HIP(hipMemcpyAsync(cg_iters, b->d_cg, sizeof(int32_t) * M, hipMemcpyDeviceToHost, s));
  ED_HIP(hipStreamSynchronize(s));
  return ED_OK;
}

extern "C" int ed_batch_set_bath(ed_batch* b, const double* eps_bath, const double* v, void* stream) {
  clear_error();
  ED_REQUIRE(b && eps_bath && v, "NULL argument");
  for (int64_t i = 0; i < b->num_models; ++i) {
    ModelParams& p = b->params[i];
    for (int k = 0; k < p.nb; ++k) {
      const double e = eps_bath[i * b->max_bath + k], x = v[i * b->max_bath + k];
      ED_REQUIRE(std::isfinite(e) && std::isfinite(x), "model %lld: parameters must be finite", (long long)i);
      p.eps[k] = e;
      p.v[k] = x;
    }
  }
  hipStream_t s = as_stream(stream);
  ED_HIP(hipMemcpyAsync(b->d_params, b->params.data(), sizeof(ModelParams) * b->num_models, hipMemcpyHostToDevice,
                        s));
  ED_HIP(hipStreamSynchronize(s));
  b->solved_nw = -1;  // the kept solve no longer matches the parameters
  return ED_OK;
}
