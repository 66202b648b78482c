// Snapshot and initial-condition file formats (formats.h).
#include "formats.h"

#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "../../include/nbody.h"
#include "engine.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__,
              "the .nbody records are written as host bytes");

namespace nbody {

// ---- float text ---------------------------------------------------------

// std::to_chars without a precision gives the shortest digit string that
// reads back as v (the nearest such string, ties to even: the same digits as
// Ryu, Adams 2018).  Only the layout is ours: with the digits D (length L)
// and value D * 10^k, kk = L + k, the Rust ryu crate's Buffer::format writes
//   0 <= k, kk <= 16      D, (kk - L) zeros, ".0"        1234e7 -> 12340000000.0
//   0 < kk <= 16          D[0:kk] "." D[kk:]             1234e-2 -> 12.34
//   -5 < kk <= 0          "0." (-kk zeros) D             1234e-6 -> 0.001234
//   L == 1                D "e" (kk - 1)                 1e30
//   otherwise             D[0] "." D[1:] "e" (kk - 1)    1234e30 -> 1.234e33
// and "0.0", "NaN", "inf", "-inf"; a negative sign (of -0.0 too) leads.
int format_f64(double v, char* buf) {
  char* p = buf;
  if (std::isnan(v)) {
    std::memcpy(buf, "NaN", 3);
    return 3;
  }
  if (std::signbit(v)) *p++ = '-';
  const double a = std::fabs(v);
  if (std::isinf(a)) {
    std::memcpy(p, "inf", 3);
    return (int)(p - buf) + 3;
  }
  if (a == 0.0) {
    std::memcpy(p, "0.0", 3);
    return (int)(p - buf) + 3;
  }
  char sci[40];
  const auto r = std::to_chars(sci, sci + sizeof(sci), a,
                               std::chars_format::scientific);
  // sci = d[.ddd]e(+|-)XX
  char digits[24];
  int L = 0;
  const char* q = sci;
  for (; q < r.ptr && *q != 'e'; ++q)
    if (*q != '.') digits[L++] = *q;
  int e10 = 0;
  std::from_chars(q + 1 + (q[1] == '+'), r.ptr, e10);
  const int kk = e10 + 1;
  const int k = kk - L;
  if (k >= 0 && kk <= 16) {
    std::memcpy(p, digits, L);
    p += L;
    for (int i = L; i < kk; ++i) *p++ = '0';
    *p++ = '.';
    *p++ = '0';
  } else if (kk > 0 && kk <= 16) {
    std::memcpy(p, digits, kk);
    p += kk;
    *p++ = '.';
    std::memcpy(p, digits + kk, L - kk);
    p += L - kk;
  } else if (kk > -5 && kk <= 0) {
    *p++ = '0';
    *p++ = '.';
    for (int i = 0; i < -kk; ++i) *p++ = '0';
    std::memcpy(p, digits, L);
    p += L;
  } else {
    *p++ = digits[0];
    if (L > 1) {
      *p++ = '.';
      std::memcpy(p, digits + 1, L - 1);
      p += L - 1;
    }
    *p++ = 'e';
    const auto w = std::to_chars(p, p + 8, kk - 1);
    p = w.ptr;
  }
  return (int)(p - buf);
}

static bool ieq(const char* s, size_t len, const char* word) {
  if (std::strlen(word) != len) return false;
  for (size_t i = 0; i < len; ++i) {
    char c = s[i];
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    if (c != word[i]) return false;
  }
  return true;
}

bool parse_f64(const char* s, size_t len, double* out) {
  size_t i = 0;
  if (i < len && (s[i] == '+' || s[i] == '-')) ++i;
  const char* body = s + i;
  const size_t blen = len - i;
  bool special = ieq(body, blen, "inf") || ieq(body, blen, "infinity") ||
                 ieq(body, blen, "nan");
  if (!special) {
    size_t nd = 0;
    while (i < len && s[i] >= '0' && s[i] <= '9') ++i, ++nd;
    if (i < len && s[i] == '.') {
      ++i;
      while (i < len && s[i] >= '0' && s[i] <= '9') ++i, ++nd;
    }
    if (nd == 0) return false;
    if (i < len && (s[i] == 'e' || s[i] == 'E')) {
      ++i;
      if (i < len && (s[i] == '+' || s[i] == '-')) ++i;
      size_t ne = 0;
      while (i < len && s[i] >= '0' && s[i] <= '9') ++i, ++ne;
      if (ne == 0) return false;
    }
    if (i != len) return false;
  }
  // strtod rounds correctly (glibc); the grammar check above keeps it to
  // the forms Rust accepts.  Out-of-range values give +-inf or 0 as there.
  std::string z(s, len);
  if (special) {
    const bool neg = len > 0 && s[0] == '-';
    if (ieq(body, blen, "nan"))
      *out = std::nan("");
    else
      *out = neg ? -HUGE_VAL : HUGE_VAL;
    return true;
  }
  errno = 0;
  char* end = nullptr;
  *out = std::strtod(z.c_str(), &end);
  return end == z.c_str() + z.size();
}

// ---- initial conditions ---------------------------------------------------

namespace {

// RFC 4180 records as the csv crate reads them: ',' separates fields; "\n",
// "\r" or "\r\n" ends a record; empty lines are skipped; a field may be
// quoted, with "" for a quote inside.  A UTF-8 byte-order mark is skipped.
struct CsvRecord {
  std::vector<std::string> fields;
  int64_t line = 0;
};

std::vector<CsvRecord> read_csv(const std::string& text) {
  std::vector<CsvRecord> out;
  size_t i = 0;
  const size_t n = text.size();
  if (n >= 3 && (unsigned char)text[0] == 0xEF &&
      (unsigned char)text[1] == 0xBB && (unsigned char)text[2] == 0xBF)
    i = 3;
  int64_t line = 1;
  while (i < n) {
    if (text[i] == '\n' || text[i] == '\r') {  // empty line
      if (text[i] == '\n') ++line;
      ++i;
      continue;
    }
    CsvRecord rec;
    rec.line = line;
    std::string field;
    bool end_of_record = false;
    while (!end_of_record) {
      field.clear();
      if (i < n && text[i] == '"') {
        ++i;
        while (i < n) {
          if (text[i] == '"') {
            if (i + 1 < n && text[i + 1] == '"') {
              field.push_back('"');
              i += 2;
            } else {
              ++i;
              break;
            }
          } else {
            if (text[i] == '\n') ++line;
            field.push_back(text[i++]);
          }
        }
      }
      while (i < n && text[i] != ',' && text[i] != '\n' && text[i] != '\r')
        field.push_back(text[i++]);
      rec.fields.push_back(field);
      if (i >= n) {
        end_of_record = true;
      } else if (text[i] == ',') {
        ++i;
      } else {
        if (text[i] == '\r' && i + 1 < n && text[i + 1] == '\n') ++i;
        if (text[i] == '\n') ++line;
        ++i;
        end_of_record = true;
      }
    }
    out.push_back(std::move(rec));
  }
  return out;
}

}  // namespace

InitialConditions load_initial_conditions(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f)
    throw Error(NBODY_E_NOTFOUND, "cannot open initial conditions file '" +
                                      path + "': " + std::strerror(errno));
  std::string text;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0)
    text.append(chunk, got);
  const bool bad = std::ferror(f);
  std::fclose(f);
  if (bad) throw Error(NBODY_E_INVALID, "read error on '" + path + "'");

  const std::vector<CsvRecord> recs = read_csv(text);
  InitialConditions ic;
  if (recs.empty()) return ic;  // no header: no bodies
  static const char* kCols[7] = {"mass",  "pos_x", "pos_y", "pos_z",
                                 "vel_x", "vel_y", "vel_z"};
  int col[7];
  const CsvRecord& hdr = recs[0];
  for (int c = 0; c < 7; ++c) {
    col[c] = -1;
    for (size_t k = 0; k < hdr.fields.size(); ++k) {
      if (hdr.fields[k] != kCols[c]) continue;
      if (col[c] >= 0)
        throw Error(NBODY_E_INVALID,
                    std::string("duplicate column `") + kCols[c] + "`");
      col[c] = (int)k;
    }
    if (col[c] < 0)
      throw Error(NBODY_E_INVALID,
                  std::string("missing field `") + kCols[c] + "`");
  }
  std::vector<double>* dst[7] = {&ic.m,  &ic.x,  &ic.y, &ic.z,
                                 &ic.vx, &ic.vy, &ic.vz};
  const size_t width = hdr.fields.size();
  for (size_t r = 1; r < recs.size(); ++r) {
    const CsvRecord& rec = recs[r];
    if (rec.fields.size() != width)
      throw Error(NBODY_E_INVALID,
                  "line " + std::to_string(rec.line) + ": found record with " +
                      std::to_string(rec.fields.size()) +
                      " fields, but the previous record has " +
                      std::to_string(width) + " fields");
    for (int c = 0; c < 7; ++c) {
      const std::string& s = rec.fields[col[c]];
      double v;
      if (!parse_f64(s.data(), s.size(), &v))
        throw Error(NBODY_E_INVALID,
                    "line " + std::to_string(rec.line) + ", field `" +
                        kCols[c] + "`: invalid float literal '" + s + "'");
      dst[c]->push_back(v);
    }
  }
  return ic;
}

// ---- snapshots ------------------------------------------------------------

std::string output_format(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  const std::string name =
      slash == std::string::npos ? path : path.substr(slash + 1);
  const size_t dot = name.find_last_of('.');
  if (dot == std::string::npos || dot == 0) return "";
  std::string ext = name.substr(dot + 1);
  for (char& c : ext)
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
  return ext == "csv" || ext == "nbody" ? ext : "";
}

SnapshotWriter::SnapshotWriter(const std::string& path) {
  const std::string fmt = output_format(path);
  if (fmt.empty())
    throw Error(NBODY_E_INVALID,
                "unsupported output file format '" + path +
                    "': the extension must be .csv or .nbody");
  csv_ = fmt == "csv";
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_)
    throw Error(NBODY_E_NOTFOUND, "cannot create output file '" + path +
                                      "': " + std::strerror(errno));
  buf_.resize(1 << 20);
  if (!csv_) put("NBODY001", 8);
}

SnapshotWriter::~SnapshotWriter() {
  if (f_) std::fclose(f_);
}

void SnapshotWriter::drain() {
  if (used_ && std::fwrite(buf_.data(), 1, used_, f_) != used_)
    throw Error(NBODY_E_INVALID, "write error on snapshot file");
  used_ = 0;
}

void SnapshotWriter::put(const char* p, size_t len) {
  if (used_ + len > buf_.size()) drain();
  std::memcpy(buf_.data() + used_, p, len);
  used_ += len;
}

void SnapshotWriter::append(double time, int64_t n, const double* x,
                            const double* y, const double* z) {
  if (!f_) throw Error(NBODY_E_STATE, "snapshot file is closed");
  if (n < 0) throw Error(NBODY_E_INVALID, "n < 0");
  if (n > 0 && (!x || !y || !z))
    throw Error(NBODY_E_INVALID, "null position array");
  if (csv_) {
    if (n > 0 && !header_done_) {
      put("time,id,x,y,z\n", 14);
      header_done_ = true;
    }
    char t[32];
    const int tl = format_f64(time, t);
    char line[160];
    for (int64_t i = 0; i < n; ++i) {
      char* p = line;
      std::memcpy(p, t, tl);
      p += tl;
      *p++ = ',';
      p = std::to_chars(p, p + 24, (uint64_t)i).ptr;
      *p++ = ',';
      p += format_f64(x[i], p);
      *p++ = ',';
      p += format_f64(y[i], p);
      *p++ = ',';
      p += format_f64(z[i], p);
      *p++ = '\n';
      put(line, (size_t)(p - line));
    }
  } else {
    char rec[40];
    for (int64_t i = 0; i < n; ++i) {
      const uint64_t id = (uint64_t)i;
      std::memcpy(rec, &time, 8);
      std::memcpy(rec + 8, &id, 8);
      std::memcpy(rec + 16, x + i, 8);
      std::memcpy(rec + 24, y + i, 8);
      std::memcpy(rec + 32, z + i, 8);
      put(rec, 40);
    }
  }
  records_ += (uint64_t)n;
}

void SnapshotWriter::close() {
  if (!f_) return;
  drain();
  const bool bad = std::fflush(f_) != 0 || std::ferror(f_);
  const bool bad_close = std::fclose(f_) != 0;
  f_ = nullptr;
  if (bad || bad_close) throw Error(NBODY_E_INVALID, "write error on snapshot file");
}

}  // namespace nbody
