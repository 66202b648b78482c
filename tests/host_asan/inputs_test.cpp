// Host tests of the code that reads user input or writes files, built with
// -fsanitize=address,undefined by tests/test_host_inputs_asan.py:
//   floats     format_f64 / parse_f64: round trips of random bit patterns,
//              the reference layout of listed values, random strings fed to
//              the parser (accepted ones agree with strtod)
//   csv        load_initial_conditions on well-formed and malformed files:
//              column order, extra columns, quoting, CRLF, a byte-order
//              mark, empty lines and files, missing / duplicate columns,
//              ragged records, bad floats, a 200 kB field
//   writer     SnapshotWriter: .csv and .nbody output larger than its
//              buffer, read back exactly; the error paths
//   scenarios  the four generators at small and odd sizes
//   tuning     tuning_load on well-formed, malformed and over-long tables
// Exit status 0 when every check passes; each failure prints a line.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/nbody.h"
#include "../../n-body-sim_amd/csrc/engine.h"
#include "../../n-body-sim_amd/csrc/formats.h"
#include "../../n-body-sim_amd/csrc/tuning.h"

using nbody::Error;

static int g_fail = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
      ++g_fail;                                                    \
    }                                                              \
  } while (0)

static std::string tmpdir() {
  const char* t = std::getenv("HOST_TEST_DIR");
  return t ? t : "/tmp";
}

static void write_file(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::printf("FAIL cannot write %s\n", path.c_str());
    ++g_fail;
    return;
  }
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
}

static std::string read_file(const std::string& path) {
  std::string s;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return s;
  char b[4096];
  size_t n;
  while ((n = std::fread(b, 1, sizeof(b), f)) > 0) s.append(b, n);
  std::fclose(f);
  return s;
}

// ---- floats --------------------------------------------------------------

static bool same_bits(double a, double b) {
  return (std::isnan(a) && std::isnan(b)) ||
         std::memcmp(&a, &b, sizeof(double)) == 0;
}

static void test_floats() {
  std::mt19937_64 rng(7);
  char buf[32];
  for (int k = 0; k < 400000; ++k) {
    uint64_t bits = rng();
    if (k % 4 == 1) bits &= 0x800fffffffffffffull;  // subnormals and zeros
    if (k % 4 == 2) bits = (bits & 0xffffffffffffull) | 0x3ff0000000000000ull;
    double v;
    std::memcpy(&v, &bits, 8);
    const int len = nbody::format_f64(v, buf);
    CHECK(len > 0 && len < 32);
    double back = 0.0;
    const bool ok = nbody::parse_f64(buf, (size_t)len, &back);
    CHECK(ok);
    if (!ok || !same_bits(v, back)) {
      std::printf("  round trip of %016llx: '%.*s'\n", (unsigned long long)bits,
                  len, buf);
      ++g_fail;
      if (g_fail > 20) return;
    }
  }
  const struct {
    double v;
    const char* text;
  } listed[] = {{1.0, "1.0"},       {0.25, "0.25"},    {12345.678, "12345.678"},
                {0.00001, "0.00001"}, {1e-6, "1e-6"},  {1.5e16, "1.5e16"},
                {-9.184518199699633e-6, "-9.184518199699633e-6"},
                {-0.0, "-0.0"},     {1e30, "1e30"},    {1e16, "1e16"},
                {1e15, "1000000000000000.0"}};
  for (const auto& c : listed) {
    const int len = nbody::format_f64(c.v, buf);
    if (std::string(buf, (size_t)len) != c.text) {
      std::printf("FAIL format %.17g: '%.*s' against '%s'\n", c.v, len, buf,
                  c.text);
      ++g_fail;
    }
  }
  // random strings over the grammar's alphabet: never a crash; an accepted
  // string reads as strtod reads it
  const char alphabet[] = "0123456789+-.eEinfatyINFATY x";
  for (int k = 0; k < 200000; ++k) {
    const size_t len = rng() % 12;
    std::string s;
    for (size_t i = 0; i < len; ++i) s.push_back(alphabet[rng() % (sizeof(alphabet) - 1)]);
    double v = 0.0;
    if (nbody::parse_f64(s.data(), s.size(), &v)) {
      const double w = std::strtod(s.c_str(), nullptr);
      CHECK(same_bits(v, w) || (std::isnan(v) && std::isnan(w)));
    }
  }
  const char* bad[] = {"", "+", "-", ".", "e5", "1e", "1e+", "0x10", " 1", "1 ",
                       "nan(1)", "in", "infinit", "1.2.3", "--1"};
  for (const char* s : bad) {
    double v;
    CHECK(!nbody::parse_f64(s, std::strlen(s), &v));
  }
  const char* good[] = {"5.", ".5", "+1", "-0", "1E5", "1e-400", "1e400",
                        "INF", "-Infinity", "NaN", "+nan"};
  for (const char* s : good) {
    double v;
    CHECK(nbody::parse_f64(s, std::strlen(s), &v));
  }
}

// ---- initial conditions ------------------------------------------------------

static int load_code(const std::string& text, size_t* bodies = nullptr) {
  const std::string p = tmpdir() + "/ic.csv";
  write_file(p, text);
  try {
    const auto ic = nbody::load_initial_conditions(p);
    if (bodies) *bodies = ic.size();
    return 0;
  } catch (const Error& e) {
    return e.code;
  }
}

static void test_csv() {
  const std::string hdr = "mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n";
  size_t n = 99;
  CHECK(load_code(hdr + "1,2,3,4,5,6,7\n", &n) == 0 && n == 1);
  CHECK(load_code("", &n) == 0 && n == 0);
  CHECK(load_code(hdr, &n) == 0 && n == 0);
  CHECK(load_code("\xEF\xBB\xBF" + hdr + "1,2,3,4,5,6,7", &n) == 0 && n == 1);
  CHECK(load_code("vel_z,x,mass,pos_x,pos_y,pos_z,vel_x,vel_y\r\n"
                  "1,junk,2,3,4,5,6,7\r\n\r\n8,\"a,b\",9,10,11,12,13,14\r\n",
                  &n) == 0 && n == 2);
  CHECK(load_code(hdr + "\"1\",\"2\",3,4,5,6,\"7\"\n", &n) == 0 && n == 1);
  CHECK(load_code("note,mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n"
                  "\"say \"\"hi\"\"\nthere\",1,2,3,4,5,6,7\n", &n) == 0 && n == 1);
  CHECK(load_code("mass,pos_x,pos_y,pos_z,vel_x,vel_y\n1,2,3,4,5,6\n") ==
        NBODY_E_INVALID);                                       // missing column
  CHECK(load_code("mass,mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n") ==
        NBODY_E_INVALID);                                       // duplicate
  CHECK(load_code(hdr + "1,2,3,4,5,6\n") == NBODY_E_INVALID);   // ragged
  CHECK(load_code(hdr + "1,2,3,4,5,6,7,8\n") == NBODY_E_INVALID);
  CHECK(load_code(hdr + "1,2,3,x,5,6,7\n") == NBODY_E_INVALID); // bad float
  CHECK(load_code(hdr + "1,2,3, 4,5,6,7\n") == NBODY_E_INVALID);
  CHECK(load_code(hdr + "1,2,3,4,5,6,\"7\n") == NBODY_E_INVALID);  // open quote
  CHECK(load_code(hdr + std::string(200000, '9') + ",2,3,4,5,6,7\n", &n) == 0 &&
        n == 1);
  CHECK(load_code(hdr + "1,2,3," + std::string(200000, 'x') + ",5,6,7\n") ==
        NBODY_E_INVALID);
  std::string nul = hdr + "1,2,3,4,5,6,7\n";
  nul[hdr.size() + 2] = '\0';
  CHECK(load_code(nul) == NBODY_E_INVALID);
  try {
    nbody::load_initial_conditions(tmpdir() + "/no_such_dir/ic.csv");
    CHECK(false);
  } catch (const Error& e) {
    CHECK(e.code == NBODY_E_NOTFOUND);
  }
  // random bytes: an error code or bodies, never a crash
  std::mt19937_64 rng(11);
  const char alphabet[] = "0123456789.,\"\r\n-e x";
  for (int k = 0; k < 300; ++k) {
    std::string t = hdr;
    const size_t len = rng() % 400;
    for (size_t i = 0; i < len; ++i) t.push_back(alphabet[rng() % (sizeof(alphabet) - 1)]);
    const int c = load_code(t);
    CHECK(c == 0 || c == NBODY_E_INVALID);
  }
}

// ---- snapshot writer --------------------------------------------------------

static void test_writer() {
  const int64_t n = 50000;  // 3 snapshots: beyond the 1 MiB buffer
  std::vector<double> x(n), y(n), z(n);
  std::mt19937_64 rng(3);
  std::normal_distribution<double> g(0.0, 3.0);
  for (int64_t i = 0; i < n; ++i) x[i] = g(rng), y[i] = g(rng), z[i] = g(rng);
  const double times[3] = {0.0, 0.001, 0.0025};
  {
    const std::string p = tmpdir() + "/out.nbody";
    nbody::SnapshotWriter w(p);
    for (double t : times) w.append(t, n, x.data(), y.data(), z.data());
    w.close();
    CHECK(w.records() == (uint64_t)(3 * n));
    const std::string b = read_file(p);
    CHECK(b.size() == 8 + 40 * 3 * (size_t)n && b.compare(0, 8, "NBODY001") == 0);
    if (b.size() == 8 + 40 * 3 * (size_t)n) {
      for (int s = 0; s < 3; ++s)
        for (int64_t i = 0; i < n; i += 997) {
          double rec[5];
          std::memcpy(rec, b.data() + 8 + 40 * ((size_t)s * n + i), 40);
          uint64_t id;
          std::memcpy(&id, &rec[1], 8);
          CHECK(rec[0] == times[s] && id == (uint64_t)i && rec[2] == x[i] &&
                rec[3] == y[i] && rec[4] == z[i]);
        }
    }
  }
  {
    const std::string p = tmpdir() + "/out.CSV";
    nbody::SnapshotWriter w(p);
    for (double t : times) w.append(t, n, x.data(), y.data(), z.data());
    w.close();
    const std::string t = read_file(p);
    CHECK(t.compare(0, 14, "time,id,x,y,z\n") == 0);
    size_t pos = 14, lines = 0;
    bool exact = true;
    while (pos < t.size()) {
      const size_t e = t.find('\n', pos);
      if (e == std::string::npos) break;
      const std::string line = t.substr(pos, e - pos);
      double f[5];
      size_t a = 0;
      for (int c = 0; c < 5; ++c) {
        const size_t b2 = c < 4 ? line.find(',', a) : line.size();
        if (b2 == std::string::npos ||
            !nbody::parse_f64(line.data() + a, b2 - a, &f[c])) {
          exact = false;
          break;
        }
        a = b2 + 1;
      }
      const size_t s = lines / (size_t)n, i = lines % (size_t)n;
      if (exact)
        exact = f[0] == times[s] && f[1] == (double)i && f[2] == x[i] &&
                f[3] == y[i] && f[4] == z[i];
      ++lines;
      pos = e + 1;
    }
    CHECK(exact && lines == 3 * (size_t)n);
  }
  {
    const std::string p = tmpdir() + "/empty.csv";
    nbody::SnapshotWriter w(p);
    w.append(0.0, 0, nullptr, nullptr, nullptr);
    w.close();
    CHECK(read_file(p).empty());  // no record, no header
  }
  int code = 0;
  try {
    nbody::SnapshotWriter w(tmpdir() + "/out.txt");
  } catch (const Error& e) {
    code = e.code;
  }
  CHECK(code == NBODY_E_INVALID);
  code = 0;
  try {
    nbody::SnapshotWriter w(tmpdir() + "/no_such_dir/out.csv");
  } catch (const Error& e) {
    code = e.code;
  }
  CHECK(code == NBODY_E_NOTFOUND);
  nbody::SnapshotWriter w(tmpdir() + "/late.nbody");
  code = 0;
  try {
    w.append(0.0, -1, x.data(), y.data(), z.data());
  } catch (const Error& e) {
    code = e.code;
  }
  CHECK(code == NBODY_E_INVALID);
  code = 0;
  try {
    w.append(0.0, 2, nullptr, y.data(), z.data());
  } catch (const Error& e) {
    code = e.code;
  }
  CHECK(code == NBODY_E_INVALID);
  w.close();
  w.close();  // a second close is a no-op
  code = 0;
  try {
    w.append(0.0, 1, x.data(), y.data(), z.data());
  } catch (const Error& e) {
    code = e.code;
  }
  CHECK(code == NBODY_E_STATE);
  CHECK(nbody::output_format("a/b.c/run.NBODY") == "nbody");
  CHECK(nbody::output_format(".csv") == "");
  CHECK(nbody::output_format("dir.csv/run") == "");
  CHECK(nbody::output_format("run.csv.bak") == "");
}

// ---- scenarios ------------------------------------------------------------------

static bool finite_all(const std::vector<std::vector<double>>& a, int64_t n) {
  for (const auto& v : a)
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(v[i])) return false;
  return true;
}

static void test_scenarios() {
  for (int64_t n : {1, 2, 3, 63, 64, 65, 1000}) {
    std::vector<std::vector<double>> a(7, std::vector<double>(n));
    double* p[7];
    for (int q = 0; q < 7; ++q) p[q] = a[q].data();
    CHECK(nbody_scenario_plummer(n, 5, 1.0, 1.0, 1.0, p[0], p[1], p[2], p[3],
                                 p[4], p[5], p[6]) == NBODY_OK);
    CHECK(finite_all(a, n));
    CHECK(nbody_scenario_star_system(n, 5, 1.0, 1e-7, 1e-4, 0.5, 10.0, 5.0, 1.0,
                                     p[0], p[1], p[2], p[3], p[4], p[5],
                                     p[6]) == NBODY_OK);
    CHECK(finite_all(a, n));
    CHECK(nbody_scenario_disc(n, 5, 1.0, 1e-3, 0.5, 0.0, 1.0, p[0], p[1], p[2],
                              p[3], p[4], p[5], p[6]) == NBODY_OK);
    CHECK(finite_all(a, n));
    const int rc = nbody_scenario_colliding(n, 5, 10.0, 0.5, 1.0, p[0], p[1],
                                            p[2], p[3], p[4], p[5], p[6]);
    CHECK(n < 2 ? rc == NBODY_E_INVALID : rc == NBODY_OK);
    if (n >= 2) CHECK(finite_all(a, n));
  }
}

// ---- tuning table -------------------------------------------------------------------

static void test_tuning() {
  const std::string p = tmpdir() + "/tuning.txt";
  int before[nbody::kTuningBuckets], after[nbody::kTuningBuckets];
  nbody::tuning_get(before, nbody::kTuningBuckets);
  write_file(p, "# nothing\n\n");
  CHECK(!nbody::tuning_load(p.c_str()));
  write_file(p, "split -1 4\nsplit 99 4\nsplit 3 0\nvariant 0 3 1\nvariant 9 3 1\n"
                "variant 2 3 99\nsplit x y\nsplit 2147483647 2147483647\n");
  CHECK(!nbody::tuning_load(p.c_str()));  // no valid entry
  nbody::tuning_get(after, nbody::kTuningBuckets);
  CHECK(std::memcmp(before, after, sizeof(before)) == 0);
  std::string longline(5000, 'x');
  write_file(p, longline + "\nsplit 5 2147483647\nvariant 3 5 1\n" + longline +
                    "split 6 3\n");
  CHECK(nbody::tuning_load(p.c_str()));
  CHECK(nbody::tuning_split(32) >= 1 && nbody::tuning_split(32) <= nbody::kMaxSplit);
  CHECK(nbody::tuning_variant(3, 40) == 1);
  CHECK(nbody::tuning_variant(0, 40) == 0 && nbody::tuning_variant(7, 40) == 0);
  CHECK(nbody::tuning_split(-5) >= 1 && nbody::tuning_split(INT64_MAX) >= 1);
  CHECK(!nbody::tuning_load((tmpdir() + "/no_such_dir/t.txt").c_str()));
}

int main(int argc, char** argv) {
  const std::string part = argc > 1 ? argv[1] : "all";
  if (part == "floats" || part == "all") test_floats();
  if (part == "csv" || part == "all") test_csv();
  if (part == "writer" || part == "all") test_writer();
  if (part == "scenarios" || part == "all") test_scenarios();
  if (part == "tuning" || part == "all") test_tuning();
  if (g_fail) {
    std::printf("%d failure(s)\n", g_fail);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
