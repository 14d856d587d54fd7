// Bank-conflict-free slot order for the half-width row-pass tables (host code,
// no HIP dependency).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

// ---------------------------------------------------------------------------
// Bank-conflict-free slot order for the half-width row-pass tables.
//
// The row pass gathers doubles from an LDS row; a 64-bit LDS read serves a
// half-wave (32 lanes) at a time, and the double at index t occupies bank
// pair (t mod 32).  Lane l of a half-wave chunk gathers its hop sources one
// slot at a time, so a slot is conflict-free when its 32 lanes use 32
// distinct bank pairs.  The entries of a half-wave chunk are the edge set of
// the bipartite multigraph lanes x bank pairs, and a proper edge colouring
// with K colours gives K conflict-free slots.  Bipartite graphs need only
// K = max degree colours (Konig); the colouring below finds one by
// alternating-path swaps.  Lanes without an entry in a slot read one of the
// zero pads D + j (j < L) placed in a bank pair that slot leaves free.
// ---------------------------------------------------------------------------
namespace ed {

constexpr int kColourLanes = 32;  // lanes served together, = bank pairs

struct ChunkColouring {
  int colours = 0;
  std::vector<int32_t> target;  // [colours][kColourLanes]: index read by lane l in slot k
};

// lanes: per lane (at most kColourLanes), its targets (0 <= t < D); pad
// indices are D + j, j < kColourLanes.  colours = -1 flags invalid input.
inline ChunkColouring colour_chunk(const std::vector<std::vector<int32_t>>& lanes, int64_t D) {
  constexpr int L = kColourLanes, B = kColourLanes;
  ChunkColouring out;
  if (lanes.size() > (size_t)L) {
    out.colours = -1;
    return out;
  }
  for (const auto& v : lanes)
    for (int32_t t : v)
      if (t < 0 || t >= D) {
        out.colours = -1;
        return out;
      }
  int K = 0;
  std::vector<int> bank_deg(B, 0);
  for (const auto& v : lanes) {
    K = std::max(K, (int)v.size());
    for (int32_t t : v) ++bank_deg[t % B];
  }
  for (int d : bank_deg) K = std::max(K, d);
  out.colours = K;
  if (K == 0) return out;
  // at_lane[l*K + k] / at_bank[b*K + k]: edge id using colour k, or -1
  std::vector<int> at_lane(L * K, -1), at_bank(B * K, -1), col;
  std::vector<int> e_lane, e_bank;
  for (int l = 0; l < (int)lanes.size(); ++l)
    for (int32_t t : lanes[l]) {
      e_lane.push_back(l);
      e_bank.push_back(t % B);
    }
  col.assign(e_lane.size(), -1);
  std::vector<int> path;
  for (size_t e = 0; e < e_lane.size(); ++e) {
    const int l = e_lane[e], b = e_bank[e];
    int a = 0, c = 0;
    while (at_lane[l * K + a] >= 0) ++a;  // free at the lane
    while (at_bank[b * K + c] >= 0) ++c;  // free at the bank
    if (at_bank[b * K + a] >= 0 && at_lane[l * K + c] < 0) a = c;  // c is free at both ends
    if (at_bank[b * K + a] >= 0) {
      // swap colours a and c along the path from bank b that starts with
      // colour a; afterwards a is free at b (and still free at l)
      path.clear();
      int v = b, k = a;
      bool on_bank = true;
      while (true) {
        const int ed = on_bank ? at_bank[v * K + k] : at_lane[v * K + k];
        if (ed < 0) break;
        path.push_back(ed);
        v = on_bank ? e_lane[ed] : e_bank[ed];
        on_bank = !on_bank;
        k = (k == a) ? c : a;
      }
      for (int ed : path) {
        at_lane[e_lane[ed] * K + col[ed]] = -1;
        at_bank[e_bank[ed] * K + col[ed]] = -1;
      }
      for (int ed : path) {
        col[ed] = (col[ed] == a) ? c : a;
        at_lane[e_lane[ed] * K + col[ed]] = ed;
        at_bank[e_bank[ed] * K + col[ed]] = ed;
      }
    }
    col[e] = a;
    at_lane[l * K + a] = (int)e;
    at_bank[b * K + a] = (int)e;
  }
  out.target.assign((size_t)K * L, -1);
  std::vector<int32_t> tg_of(e_lane.size());
  {
    size_t e = 0;
    for (int l = 0; l < (int)lanes.size(); ++l)
      for (int32_t t : lanes[l]) tg_of[e++] = t;
  }
  for (size_t e = 0; e < e_lane.size(); ++e) out.target[(size_t)col[e] * L + e_lane[e]] = tg_of[e];
  // pads in the banks each slot leaves free
  for (int k = 0; k < K; ++k) {
    std::vector<char> used(B, 0);
    for (int l = 0; l < L; ++l)
      if (out.target[(size_t)k * L + l] >= 0) used[out.target[(size_t)k * L + l] % B] = 1;
    int fb = 0;
    for (int l = 0; l < L; ++l) {
      if (out.target[(size_t)k * L + l] >= 0) continue;
      while (used[fb]) ++fb;
      used[fb] = 1;
      out.target[(size_t)k * L + l] = (int32_t)(D + ((fb - D % B + B) % B));
    }
  }
  return out;
}

}  // namespace ed
