// Split arithmetic of the SH gather, shared by the device planner and the
// host code that sizes the partial-sum buffer (also compiled by the CPU
// tests, so it includes no HIP header).
//
// The gather divides the VPL range of a light into `s` splits (a power of
// two) and writes one partial of 27 sums per 64-cache tile and group of four
// splits: ceil(s / 4) groups of 27 rows of `pitch` floats.  A buffer of
// 27 x `columns` floats therefore holds a gather with s splits exactly when
// ceil(s / 4) x pitch <= columns.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#define DRV_HOST_DEVICE __host__ __device__
#else
#define DRV_HOST_DEVICE
#endif

namespace drvk {

// Splits the device planner picks: the smallest power of two that gives the
// persistent grid `target_items` (tile, split) items, at most `max_splits`
// (a power of two) and keeping every split at least `quantum` VPLs long.
DRV_HOST_DEVICE inline uint32_t plan_split_count(uint32_t live, uint32_t n_vpl, uint32_t quantum,
                                                 uint32_t max_splits, uint32_t target_items) {
  const uint32_t tiles64 = (live + 63u) / 64u;
  uint32_t s = 1u;
  while (s < max_splits && tiles64 * s < target_items && (n_vpl / (2u * s)) >= quantum) s *= 2u;
  return s;
}

// Partial columns (floats per coefficient row) a gather with `splits` splits
// writes at partial pitch `pitch`.
DRV_HOST_DEVICE inline uint64_t gather_partial_columns(uint32_t splits, uint32_t pitch) {
  return static_cast<uint64_t>((splits + 3u) / 4u) * pitch;
}

// The largest power of two s <= cap_splits (a power of two) whose partials
// fit `columns`; 0 when not even one group of splits fits.
DRV_HOST_DEVICE inline uint32_t max_fitting_splits(uint64_t columns, uint32_t pitch, uint32_t cap_splits) {
  uint32_t s = cap_splits;
  while (s >= 1u && gather_partial_columns(s, pitch) > columns) s >>= 1;
  return s;
}

}  // namespace drvk
