// CPU check of operator construction (ed_op_create and the table builders of
// ed_core.hip: states, diagonal, ELL and packed hop tables, sign-split offset
// tables, half-width row tables with their host-side bank colouring, sorted
// target, pair and quad lists) under AddressSanitizer / UBSan, with the
// library's kernels run on the device emulation of hip/hip_runtime.h.
//
// Input (text, argv[1]): M, then per model
//   num_sites n_up n_dn width num_bonds / bonds (pairs) / hop[num_bonds] /
//   eps[num_sites] / u[num_sites]
// Output per model: "model m D num_up num_dn", then "dense v..." (row-major)
// where D <= 400, and "hop sigma k cnt col val col val ..." per state of each
// spin species with at most 10,000 states.
#include <algorithm>
#include <fstream>
#include <string>

#include "cmpy_amd/csrc/ed_core.cpp"

// Decodes a species' half-width row tables back into each state's hop list
// and compares it, as a multiset of (target, sign), with the packed table;
// checks that every slot of a half-wave reads 32 distinct bank pairs and that
// the lane-contiguous copy holds the same words.  Prints "half sigma words"
// and returns 0, or reports and returns 1.
static int check_half(const ed::Tables& t, int64_t count, int sigma) {
  if (!t.h16) return 0;
  constexpr int L = ed::kColourLanes;
  const int64_t nchunk = (count + 63) / 64;
  const int nw = t.hwords;
  for (int64_t i = 0; i < count; ++i) {
    const int64_t c = i / 64;
    const int nwp = t.hwords_chunk[2 * c], nwc = t.hwords_chunk[2 * c + 1];
    std::vector<std::pair<int32_t, int>> got, want;
    for (int w = 0; w < nw; ++w) {
      const uint32_t x = t.h16[(size_t)w * count + i];
      const uint32_t q = t.h16q[((size_t)(c * t.hgroups + w / 4) * 64 + i % 64) * 4 + w % 4];
      if (q != x) {
        std::fprintf(stderr, "sigma %d state %lld word %d: lane-contiguous copy differs\n", sigma, (long long)i, w);
        return 1;
      }
      for (int j = 0; j < 2; ++j) {
        const int32_t p = (int32_t)((x >> (16 * j)) & 0xffffu);
        const int sg = w < nwp ? 1 : -1;
        if (p >= count + L) {
          std::fprintf(stderr, "sigma %d state %lld: entry %d past the pads\n", sigma, (long long)i, p);
          return 1;
        }
        if (p >= count) continue;
        if (w >= nwc) {
          std::fprintf(stderr, "sigma %d state %lld: hop past the chunk's words\n", sigma, (long long)i);
          return 1;
        }
        got.emplace_back(p, sg);
      }
    }
    for (int k = 0; k < t.cnt[i]; ++k) {
      const uint32_t e = t.pk_cm[(size_t)k * count + i];
      want.emplace_back(ed::hop_target(e), (e & 0x80000000u) ? -1 : 1);
    }
    std::sort(got.begin(), got.end());
    std::sort(want.begin(), want.end());
    if (got != want) {
      std::fprintf(stderr, "sigma %d state %lld: decoded hops differ from the packed table\n", sigma, (long long)i);
      return 1;
    }
  }
  for (int64_t c = 0; c < nchunk; ++c)
    for (int h = 0; h < 2; ++h)
      for (int w = 0; w < nw; ++w)
        for (int j = 0; j < 2; ++j) {
          std::vector<char> used(L, 0);
          for (int l = 0; l < L; ++l) {
            const int64_t i = c * 64 + h * L + l;
            if (i >= count) break;
            const int32_t p = (int32_t)((t.h16[(size_t)w * count + i] >> (16 * j)) & 0xffffu);
            if (used[p % L]++) {
              std::fprintf(stderr, "sigma %d chunk %lld half %d slot %d: bank pair %d twice\n", sigma,
                           (long long)c, h, 2 * w + j, p % L);
              return 1;
            }
          }
        }
  std::printf("half %d %d\n", sigma, nw);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  int M = 0;
  in >> M;
  for (int m = 0; m < M; ++m) {
    int ns, nu, nd, width, nb;
    in >> ns >> nu >> nd >> width >> nb;
    std::vector<int32_t> bonds(2 * nb);
    std::vector<double> hop(nb), eps(ns), u(ns);
    for (auto& b : bonds) in >> b;
    for (auto& h : hop) in >> h;
    for (auto& e : eps) in >> e;
    for (auto& x : u) in >> x;
    if (!in) return 2;
    ed_model_desc d{};
    d.num_sites = ns;
    d.n_up = nu;
    d.n_dn = nd;
    d.num_bonds = nb;
    d.bonds = bonds.data();
    d.hop = hop.data();
    d.eps = eps.data();
    d.u = u.data();
    d.width = width;
    ed_op* op = nullptr;
    if (ed_op_create(&op, &d, nullptr) != ED_OK) {
      std::fprintf(stderr, "model %d: %s\n", m, ed_last_error());
      return 1;
    }
    int64_t a = 0, b = 0;
    ed_op_shape(op, &a, &b);
    const int64_t D = a * b;
    std::printf("model %d %lld %lld %lld\n", m, (long long)D, (long long)a, (long long)b);
    if (check_half(op->up, a, 1) || check_half(op->dn, b, 2)) return 1;
    if (D <= 400) {
      std::vector<double> h((size_t)(D * D));
      double* hd = nullptr;
      (void)hipMalloc(&hd, sizeof(double) * D * D);
      if (ed_op_dense(op, hd, nullptr) != ED_OK) return 1;
      std::memcpy(h.data(), hd, sizeof(double) * D * D);
      (void)hipFree(hd);
      std::printf("dense");
      for (double v : h) std::printf(" %.17g", v);
      std::printf("\n");
    }
    for (int sigma = 1; sigma <= 2; ++sigma) {
      const int64_t n = sigma == 1 ? a : b;
      if (n > 10000) continue;
      int32_t w = 0;
      if (ed_op_hop_table(op, sigma, &w, nullptr, nullptr, nullptr, nullptr) != ED_OK) return 1;
      int32_t *cnt = nullptr, *col = nullptr;
      double* val = nullptr;
      (void)hipMalloc(&cnt, sizeof(int32_t) * n);
      (void)hipMalloc(&col, sizeof(int32_t) * n * std::max(w, 1));
      (void)hipMalloc(&val, sizeof(double) * n * std::max(w, 1));
      if (ed_op_hop_table(op, sigma, &w, cnt, col, val, nullptr) != ED_OK) return 1;
      for (int64_t k = 0; k < n; ++k) {
        std::printf("hop %d %lld %d", sigma, (long long)k, cnt[k]);
        for (int e = 0; e < cnt[k]; ++e) std::printf(" %d %.17g", col[k * w + e], val[k * w + e]);
        std::printf("\n");
      }
      (void)hipFree(cnt);
      (void)hipFree(col);
      (void)hipFree(val);
    }
    ed_op_destroy(op);
  }
  emu::release_fibers();
  return 0;
}
