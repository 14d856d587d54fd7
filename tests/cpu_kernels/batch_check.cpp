// CPU check of the batched impurity solver's kernels (k_batch_solve,
// k_batch_backward) under AddressSanitizer / UBSan: the library's own
// ed_batch.hip, compiled for the host against the device emulation in
// hip/hip_runtime.h (build_harness.py makes the copy and builds this file).
//
// Input (text, from the file named by argv[1]):
//   M max_bath wave_max_dim grid_blocks wave_blocks max_iter cf_steps keep nw
//   w_1 .. w_nw
//   then M lines: nb n_up n_dn u eps_imp eps_1 .. eps_nb v_1 .. v_nb
// Output (stdout): "e0 i value", "it i a b c", "res i r1 r2", "g i n re im",
// and with keep: "grad i k d_eps d_v" for dL = sum_n Re G_i(n) + Im G_i(n).
//
// Between the workspace slots of the groups (one per wave in wave mode, one
// per workgroup in block mode) the harness leaves poisoned guard bytes, so an
// access past a group's own slot is reported; every LDS array and private
// array has its exact size; every device buffer is a separate allocation.
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "cmpy_amd/csrc/ed_batch.cpp"

namespace ed {
static thread_local std::string g_err;
int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void clear_error() { g_err.clear(); }
}  // namespace ed

// Rebuilds the solver workspace with kGuard poisoned bytes after every slot.
static char* guard_slots(ed_batch* b, std::vector<std::pair<char*, size_t>>& guards) {
  constexpr int64_t kGuard = 4096;
  Config& c = b->cfg;
  const int64_t n_wave = (int64_t)c.wave_blocks * (kBlock / 64);
  const int64_t n_block = (int64_t)b->grid - c.wave_blocks;
  const int64_t sw = c.slot_bytes_wave + kGuard, sb = c.slot_bytes_block + kGuard;
  const size_t bytes = (size_t)(n_wave * sw + n_block * sb);
  char* ws = static_cast<char*>(std::aligned_alloc(256, (bytes + 255) / 256 * 256));
  std::memset(ws, 0xFF, bytes);
  for (int64_t i = 0; i < n_wave; ++i) guards.emplace_back(ws + i * sw + c.slot_bytes_wave, (size_t)kGuard);
  for (int64_t i = 0; i < n_block; ++i)
    guards.emplace_back(ws + n_wave * sw + i * sb + c.slot_bytes_block, (size_t)kGuard);
  for (auto& g : guards) ASAN_POISON_MEMORY_REGION(g.first, g.second);
  c.ws_wave = ws;
  c.ws_block = ws + n_wave * sw;
  c.slot_bytes_wave = sw;
  c.slot_bytes_block = sb;
  return ws;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s models.txt\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  int64_t M, wave_max_dim;
  int max_bath, grid, wave_blocks, max_iter, cf_steps, keep, nw;
  in >> M >> max_bath >> wave_max_dim >> grid >> wave_blocks >> max_iter >> cf_steps >> keep >> nw;
  std::vector<double> wn(nw);
  for (auto& w : wn) in >> w;
  std::vector<int32_t> nb(M), nu(M), nd(M);
  std::vector<double> u(M), ei(M), eb(M * max_bath, 0.0), v(M * max_bath, 0.0);
  for (int64_t i = 0; i < M; ++i) {
    in >> nb[i] >> nu[i] >> nd[i] >> u[i] >> ei[i];
    for (int k = 0; k < nb[i]; ++k) in >> eb[i * max_bath + k];
    for (int k = 0; k < nb[i]; ++k) in >> v[i * max_bath + k];
  }
  if (!in) {
    std::fprintf(stderr, "bad input\n");
    return 2;
  }
  ed_batch_desc d{};
  d.num_models = M;
  d.num_bath = nb.data();
  d.n_up = nu.data();
  d.n_dn = nd.data();
  d.u = u.data();
  d.eps_imp = ei.data();
  d.eps_bath = eb.data();
  d.v = v.data();
  d.max_bath = max_bath;
  d.sign_width = 0;
  d.max_iter = max_iter;
  d.cf_steps = cf_steps;
  d.check_every = 8;
  d.keep_basis = keep;
  d.tol = 1e-10;
  d.seed = 0x5EED;
  d.cf_tol = 1e-13;
  ed_batch* b = nullptr;
  if (ed_batch_create(&b, &d, nullptr) != ED_OK || ed_batch_config(b, wave_max_dim, grid, wave_blocks, nullptr) != ED_OK) {
    std::fprintf(stderr, "setup failed: %s\n", ed::g_err.c_str());
    return 1;
  }
  std::vector<std::pair<char*, size_t>> guards;
  char* ws = guard_slots(b, guards);
  std::vector<double> e0(M), g(2 * M * nw), resid(2 * M);
  std::vector<int32_t> iters(3 * M);
  double ms = 0.0;
  if (ed_batch_solve(b, wn.data(), nw, e0.data(), g.data(), iters.data(), resid.data(), &ms, nullptr) != ED_OK) {
    std::fprintf(stderr, "solve failed: %s\n", ed::g_err.c_str());
    return 1;
  }
  std::printf("switches %llu\n", (unsigned long long)emu::E().switches);
  std::printf("chunks %zu\n", b->chunks.size());
  for (int64_t i = 0; i < M; ++i) {
    std::printf("e0 %lld %.17g\n", (long long)i, e0[i]);
    std::printf("it %lld %d %d %d\n", (long long)i, iters[3 * i], iters[3 * i + 1], iters[3 * i + 2]);
    std::printf("res %lld %.6g %.6g\n", (long long)i, resid[2 * i], resid[2 * i + 1]);
    for (int n = 0; n < nw; ++n)
      std::printf("g %lld %d %.17g %.17g\n", (long long)i, n, g[2 * (i * nw + n)], g[2 * (i * nw + n) + 1]);
  }
  if (keep) {
    std::vector<double> dre(M * nw, 1.0), dim(M * nw, 1.0), ge(M * max_bath), gv(M * max_bath);
    std::vector<int32_t> cg(M);
    if (ed_batch_backward(b, dre.data(), dim.data(), ge.data(), gv.data(), cg.data(), &ms, nullptr) != ED_OK) {
      std::fprintf(stderr, "backward failed: %s\n", ed::g_err.c_str());
      return 1;
    }
    for (int64_t i = 0; i < M; ++i)
      for (int k = 0; k < nb[i]; ++k)
        std::printf("grad %lld %d %.17g %.17g\n", (long long)i, k, ge[i * max_bath + k], gv[i * max_bath + k]);
  }
  std::fflush(stdout);
  for (auto& gd : guards) ASAN_UNPOISON_MEMORY_REGION(gd.first, gd.second);
  b->cfg.ws_wave = b->cfg.ws_block = nullptr;
  std::free(ws);
  ed_batch_destroy(b);
  emu::release_fibers();
  return 0;
}
