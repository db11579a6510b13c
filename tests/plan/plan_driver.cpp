// C entry point over the forward planner (acctricnt_amd/csrc/fwd_plan.h) for
// the CPU tests: the same header the HIP library compiles, built here with
// g++ alone.
#include <cstring>

#include "fwd_plan.h"

extern "C" int plan_forward(uint32_t n, const uint32_t *off, const uint32_t *hs, const int32_t *lo, uint32_t wave_slots,
                            uint32_t wave_low, uint32_t chunk, uint32_t device_waves, uint32_t cap, FwdUnit *wave,
                            uint32_t *n_wave, FwdUnit *block, uint32_t *n_block, FwdUnit *big, uint32_t *n_big) {
  std::vector<FwdUnit> w, bl, bg;
  build_forward_plan(0, n, off, hs, lo, wave_slots, wave_low, chunk, device_waves, w, bl, bg);
  *n_wave = static_cast<uint32_t>(w.size());
  *n_block = static_cast<uint32_t>(bl.size());
  *n_big = static_cast<uint32_t>(bg.size());
  if (w.size() > cap || bl.size() > cap || bg.size() > cap) return -1;
  if (!w.empty()) std::memcpy(wave, w.data(), sizeof(FwdUnit) * w.size());
  if (!bl.empty()) std::memcpy(block, bl.data(), sizeof(FwdUnit) * bl.size());
  if (!bg.empty()) std::memcpy(big, bg.data(), sizeof(FwdUnit) * bg.size());
  return 0;
}
