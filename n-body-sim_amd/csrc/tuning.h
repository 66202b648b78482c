// Process-wide tuning table: source-range parts per tile by size bucket.
#pragma once
#include <cstdint>

namespace nbody {

constexpr int kTuningBuckets = 32;  // bucket b holds n with floor(log2 n) == b

// Parts for a system of n bodies (table entry of its bucket).
int tuning_split(int64_t n);
// Replace / read the table (kTuningBuckets entries, each 1..kMaxSplit).
void tuning_set(const int* parts, int count);
void tuning_get(int* parts, int count);
// Load "split" entries from a tuning file; returns false if absent/invalid.
bool tuning_load(const char* path);

}  // namespace nbody
