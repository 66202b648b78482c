// Tuning table: source-range parts per target tile, by size bucket.
//
// The default reproduces the rule measured on MI355X: aim for ~16384 waves per
// system with at least 256 sources per part, at most kMaxSplit parts.  An
// autotuned table (n-body-sim_amd/tuning/gfx950.txt, written by
// nbody_amd/autotune.py from timings on the GPU) replaces it when present next
// to the library or named by NBODY_AMD_TUNING.  The table is read once per
// process; results depend on it (summation order), so a given table always
// gives the same bits.
#include "tuning.h"

#include <dlfcn.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "nbody_types.h"

namespace nbody {
namespace {

std::once_flag g_once;
int g_split[kTuningBuckets];
int g_variant[4][kTuningBuckets];  // by Mode; zero (the default kernel) at start

int default_split(int b) {
  const int64_t n = (int64_t)1 << b;
  const int64_t tiles = (n + kWave - 1) / kWave;
  int64_t k = std::min<int64_t>(16384 / std::max<int64_t>(tiles, 1), n / 256);
  return (int)std::max<int64_t>(1, std::min<int64_t>(k, kMaxSplit));
}

std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&default_split, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
  }
  return ".";
}

void init() {
  for (int b = 0; b < kTuningBuckets; ++b) g_split[b] = default_split(b);
  const char* env = std::getenv("NBODY_AMD_TUNING");
  if (env && *env) {
    tuning_load(env);
  } else {
    tuning_load((library_dir() + "/tuning/gfx950.txt").c_str());
  }
}

// The whitespace-separated integers of `s` after its first word `word`
// (at most 3); returns how many were read, or -1 when the word differs, a
// token is not an integer or one leaves the int range (strtol, not sscanf's
// %d, whose overflow is undefined).
int parse_entry(const char* s, const char* word, int* v) {
  const size_t wl = std::strlen(word);
  if (std::strncmp(s, word, wl) != 0 || (s[wl] != ' ' && s[wl] != '\t'))
    return -1;
  const char* p = s + wl;
  int k = 0;
  for (;;) {
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\0' || *p == '\n' || *p == '\r' || *p == '#') return k;
    if (k == 3) return -1;
    char* end = nullptr;
    errno = 0;
    const long x = std::strtol(p, &end, 10);
    if (end == p || errno == ERANGE || x < INT_MIN || x > INT_MAX) return -1;
    v[k++] = (int)x;
    p = end;
  }
}

int bucket_of(int64_t n) {
  int b = 0;
  while (b + 1 < kTuningBuckets && ((int64_t)1 << (b + 1)) <= n) ++b;
  return b;
}

}  // namespace

int tuning_split(int64_t n) {
  std::call_once(g_once, init);
  return g_split[bucket_of(std::max<int64_t>(n, 1))];
}

void tuning_set(const int* parts, int count) {
  std::call_once(g_once, init);
  for (int b = 0; b < kTuningBuckets && b < count; ++b)
    g_split[b] = std::max(1, std::min(parts[b], kMaxSplit));
}

void tuning_get(int* parts, int count) {
  std::call_once(g_once, init);
  for (int b = 0; b < kTuningBuckets && b < count; ++b) parts[b] = g_split[b];
}

int tuning_variant(int mode, int64_t n) {
  std::call_once(g_once, init);
  if (mode <= MODE_F64_STRICT || mode > MODE_F32R) return 0;
  return g_variant[mode][bucket_of(std::max<int64_t>(n, 1))];
}

void tuning_set_variant(int mode, const int* v, int count) {
  std::call_once(g_once, init);
  if (mode <= MODE_F64_STRICT || mode > MODE_F32R) return;
  for (int b = 0; b < kTuningBuckets && b < count; ++b)
    g_variant[mode][b] = std::max(0, std::min(v[b], kForceVariants - 1));
}

void tuning_get_variant(int mode, int* v, int count) {
  std::call_once(g_once, init);
  for (int b = 0; b < kTuningBuckets && b < count; ++b)
    v[b] = (mode <= MODE_F64_STRICT || mode > MODE_F32R) ? 0 : g_variant[mode][b];
}

// File format: lines "split <bucket> <parts>" and "variant <mode> <bucket>
// <variant>" (mode: 1 f64, 2 f32c, 3 f32r); '#' starts a comment.
bool tuning_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char line[256];
  int loaded = 0;
  int tmp[kTuningBuckets];
  int tv[4][kTuningBuckets];
  std::memcpy(tmp, g_split, sizeof(tmp));
  std::memcpy(tv, g_variant, sizeof(tv));
  while (std::fgets(line, sizeof(line), f)) {
    // a line longer than the buffer is dropped whole
    const bool whole = std::strchr(line, '\n') || std::feof(f);
    if (!whole) {
      int c;
      while ((c = std::fgetc(f)) != EOF && c != '\n') {
      }
      continue;
    }
    if (line[0] == '#') continue;
    int v[3];
    if (parse_entry(line, "split", v) == 2 && v[0] >= 0 &&
        v[0] < kTuningBuckets && v[1] >= 1) {
      tmp[v[0]] = std::min(v[1], kMaxSplit);
      ++loaded;
    } else if (parse_entry(line, "variant", v) == 3 && v[0] > MODE_F64_STRICT &&
               v[0] <= MODE_F32R && v[1] >= 0 && v[1] < kTuningBuckets &&
               v[2] >= 0 && v[2] < kForceVariants) {
      tv[v[0]][v[1]] = v[2];
      ++loaded;
    }
  }
  std::fclose(f);
  if (loaded == 0) return false;
  std::memcpy(g_split, tmp, sizeof(tmp));
  std::memcpy(g_variant, tv, sizeof(tv));
  return true;
}

}  // namespace nbody
