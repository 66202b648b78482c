// Process-wide tuning table by size bucket: source-range parts per tile and
// the force-kernel variant per arithmetic mode.
#pragma once
#include <cstdint>

namespace nbody {

constexpr int kTuningBuckets = 32;  // bucket b holds n with floor(log2 n) == b

// Parts for a system of n bodies (table entry of its bucket).
int tuning_split(int64_t n);
// Replace / read the table (kTuningBuckets entries, each 1..kMaxSplit).
void tuning_set(const int* parts, int count);
void tuning_get(int* parts, int count);
// Force-kernel variant (0 .. kForceVariants-1) for a system of n bodies in
// an arithmetic mode (Mode; MODE_F64_STRICT has one kernel: always 0).
int tuning_variant(int mode, int64_t n);
// Replace / read a mode's variant table (kTuningBuckets entries).
void tuning_set_variant(int mode, const int* v, int count);
void tuning_get_variant(int mode, int* v, int count);
// Load "split" and "variant" entries from a tuning file; returns false if
// absent or without any entry.
bool tuning_load(const char* path);

}  // namespace nbody
