// Forward-path work plan (host only): the degree-skew bins of the forward
// counting kernels (k_forward.hip) and the planner that fills them.  Plain
// C++ with no HIP dependency, so that the CPU tests compile the same planner
// (tests/test_pack_plan.py) and check its bounds against an emulation of the
// wave-pack kernel's index arithmetic.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <utility>
#include <vector>

namespace accfwd {
constexpr int kWG = 512;                              // threads per workgroup (8 waves)
constexpr int kWaves = kWG / 64;
constexpr int kTableBits = 12;                        // workgroup LDS hash table of 4096 keys
constexpr uint32_t kLdsCap = 1u << (kTableBits - 1);  // |N<(b)| up to 2048 stays in LDS
constexpr uint32_t kChunk = 512;                      // upper slots per workgroup unit
constexpr uint32_t kBitmapIds = 32u << kTableBits;    // b - min N(b) <= 2^17: N<(b) as an LDS bitmap of [min N(b), b)
constexpr uint32_t kBitmapMaxLow = 65535;             // word prefixes are 16-bit
constexpr int kWaveTableBits = 9;                     // wave-private table of 512 keys
constexpr uint32_t kWaveCap = 1u << (kWaveTableBits - 1);  // |N<(b)| up to 256 in a wave unit
constexpr uint32_t kWaveSlots = 64;                   // upper slots of a wave unit
constexpr int kBigWG = 256;
}  // namespace accfwd

// One forward-path work unit: source b, its upper slots [j0, j1), nlow = |N<(b)|.
struct FwdUnit {
  int32_t b;
  uint32_t j0;
  uint32_t j1;
  uint32_t nlow;
};

// Forward plan, binned by degree skew: a source b with upper slots (neighbors
// > b) is a small source when it has at most wave_slots upper slots and
// |N<(b)| <= wave_low; runs of consecutive small sources form wave packs
// (FwdUnit {b0, b1, upper slots, lower entries}: at most 64 sources, 64
// upper slots, 256 lower entries and 512 hash-table entries, each source's
// table being the smallest power of two >= 2 |N<(b)|, at least 16).  Other
// sources become workgroup units of at most `chunk` upper slots (N<(b) as an
// LDS bitmap when b - lo[b] <= kBitmapIds, else an LDS hash table); sources
// whose N<(b) fits neither go to the global-memory variant.  hs[b] is the
// first upper slot of b, lo[b] its smallest neighbor.
inline void build_forward_range(uint32_t b_lo, uint32_t b_hi, const uint32_t *off, const uint32_t *hs,
                                const int32_t *lo, uint32_t wave_slots, uint32_t wave_low, uint32_t chunk,
                                uint32_t max_sources, std::vector<FwdUnit> &wave, std::vector<FwdUnit> &block,
                                std::vector<FwdUnit> &big) {
  auto table = [](uint32_t nl) -> uint32_t {
    if (!nl) return 0;
    uint32_t t = 16;
    while (t < 2 * nl) t *= 2;
    return t;
  };
  FwdUnit pack{0, 0, 0, 0};
  uint32_t pack_tbl = 0, pack_sources = 0;
  bool open = false;
  auto close = [&]() {
    if (open) wave.push_back(pack);
    open = false;
  };
  for (uint32_t b = b_lo; b < b_hi; b++) {
    const uint32_t nl = hs[b] - off[b];
    const uint32_t nup = off[b + 1] - hs[b];
    if (nup == 0) continue;
    if (nup <= wave_slots && nl <= wave_low) {
      const uint32_t t = table(nl);
      if (open && (b + 1 - static_cast<uint32_t>(pack.b) > 64 || pack_sources == max_sources || pack.j1 + nup > 64 ||
                   pack.nlow + nl > accfwd::kWaveCap || pack_tbl + t > (1u << accfwd::kWaveTableBits)))
        close();
      if (!open) {
        pack = FwdUnit{static_cast<int32_t>(b), b + 1, 0, 0};
        pack_tbl = 0;
        pack_sources = 0;
        open = true;
      }
      pack_sources++;
      pack.j0 = b + 1;
      pack.j1 += nup;
      pack.nlow += nl;
      pack_tbl += t;
      continue;
    }
    close();
    for (uint32_t j0 = hs[b]; j0 < off[b + 1]; j0 += chunk) {
      FwdUnit u{static_cast<int32_t>(b), j0, std::min(j0 + chunk, off[b + 1]), nl};
      const uint32_t span = b - static_cast<uint32_t>(lo[b]);
      const bool lds = nl <= accfwd::kLdsCap || (span <= accfwd::kBitmapIds && nl <= accfwd::kBitmapMaxLow);
      (lds ? block : big).push_back(u);
    }
  }
  close();
}

// Builds the plan of sources [b_lo, b_hi) in kPlanThreads contiguous pieces
// on as many host threads (packs do not cross a piece boundary) and joins the
// pieces' lists in order.  Sources per pack: just enough packs to give every
// wave slot of the device one (a small graph keeps one source per wave), at
// most 64.
constexpr uint32_t kPlanThreads = 8;
inline void build_forward_plan(uint32_t b_lo, uint32_t b_hi, const uint32_t *off, const uint32_t *hs,
                               const int32_t *lo, uint32_t wave_slots, uint32_t wave_low, uint32_t chunk,
                               uint32_t wave_slots_on_device, std::vector<FwdUnit> &wave, std::vector<FwdUnit> &block,
                               std::vector<FwdUnit> &big) {
  const uint32_t span = b_hi - b_lo;
  const uint32_t pieces = span < (1u << 16) ? 1 : kPlanThreads;
  std::vector<uint64_t> small(pieces, 0);
  std::vector<std::vector<FwdUnit>> w(pieces), bl(pieces), bg(pieces);
  auto piece = [&](uint32_t t) {
    return std::pair<uint32_t, uint32_t>(b_lo + static_cast<uint32_t>(uint64_t(span) * t / pieces),
                                         b_lo + static_cast<uint32_t>(uint64_t(span) * (t + 1) / pieces));
  };
  auto run = [&](auto &&fn) {
    std::vector<std::thread> th;
    uint32_t t = 1;
    try {
      for (; t < pieces; t++) th.emplace_back(fn, t);
    } catch (...) {  // no thread to spare: the rest run here
    }
    fn(0);
    for (uint32_t r = t; r < pieces; r++) fn(r);
    for (auto &x : th) x.join();
  };
  run([&](uint32_t t) {
    const auto [a, e] = piece(t);
    uint64_t c = 0;
    for (uint32_t b = a; b < e; b++) {
      const uint32_t nl = hs[b] - off[b], nup = off[b + 1] - hs[b];
      c += nup && nup <= wave_slots && nl <= wave_low;
    }
    small[t] = c;
  });
  uint64_t n_small = 0;
  for (uint64_t c : small) n_small += c;
  const uint64_t per = (n_small + wave_slots_on_device - 1) / std::max<uint32_t>(wave_slots_on_device, 1);
  const uint32_t max_sources = static_cast<uint32_t>(std::min<uint64_t>(64, std::max<uint64_t>(1, per)));
  run([&](uint32_t t) {
    const auto [a, e] = piece(t);
    build_forward_range(a, e, off, hs, lo, wave_slots, wave_low, chunk, max_sources, w[t], bl[t], bg[t]);
  });
  for (uint32_t t = 0; t < pieces; t++) {
    wave.insert(wave.end(), w[t].begin(), w[t].end());
    block.insert(block.end(), bl[t].begin(), bl[t].end());
    big.insert(big.end(), bg[t].begin(), bg[t].end());
  }
}

