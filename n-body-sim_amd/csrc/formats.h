// Snapshot and initial-condition file formats of the n-body simulator.
//
// Output (src/output.rs:73-139 of the reference): the format follows the
// file extension, compared case-insensitively.
//   .csv    header "time,id,x,y,z", then one line per body per snapshot,
//           "\n" line ends, floats in shortest round-trip form laid out as
//           the Rust csv crate writes them (ryu), ids in decimal.  The
//           header goes out with the first record, so a run that records
//           nothing leaves an empty file.
//   .nbody  8-byte magic "NBODY001", then 40-byte little-endian records
//           f64 time, u64 id, f64 x, f64 y, f64 z (src/output.rs:124-138).
// Input (src/input.rs:8-33): CSV with a header that names the columns
// mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z in any order (other columns are
// ignored), one body per record, fields parsed with Rust's f64 grammar.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace nbody {

// Shortest decimal that reads back as v, in the Rust ryu layout:
//   1.0, 0.25, 12345.678, 0.00001, 1e-6, 1.5e16, -9.184518199699633e-6,
//   NaN, inf, -inf, -0.0.
// Writes into buf (at least 32 bytes); returns the length.
int format_f64(double v, char* buf);

// Parses one field with Rust's `f64::from_str` grammar: optional sign, then
// inf / infinity / nan (any case) or digits with an optional '.' and an
// optional exponent.  No surrounding spaces, no hex.  Returns false when
// the text is not such a number.
bool parse_f64(const char* s, size_t len, double* out);

struct InitialConditions {
  std::vector<double> m, x, y, z, vx, vy, vz;
  size_t size() const { return m.size(); }
};

// Reads initial conditions; throws Error(NBODY_E_INVALID) with a message
// naming the line on malformed input, Error(NBODY_E_NOTFOUND) when the file
// cannot be opened.
InitialConditions load_initial_conditions(const std::string& path);

// "csv", "nbody" or "" (other / no extension), after Rust's
// Path::extension: the text after the last '.' of the file name, none when
// the name has no '.' or its only '.' is the first character.
std::string output_format(const std::string& path);

class SnapshotWriter {
 public:
  // Creates the file; throws Error(NBODY_E_INVALID) for an unsupported
  // extension, Error(NBODY_E_NOTFOUND) when the file cannot be created.
  explicit SnapshotWriter(const std::string& path);
  ~SnapshotWriter();
  SnapshotWriter(const SnapshotWriter&) = delete;
  SnapshotWriter& operator=(const SnapshotWriter&) = delete;

  // One snapshot: records (time, id, x[id], y[id], z[id]) for id 0..n-1.
  void append(double time, int64_t n, const double* x, const double* y,
              const double* z);
  // Flushes and closes; throws on a write error.
  void close();
  uint64_t records() const { return records_; }

 private:
  void put(const char* p, size_t len);
  void drain();
  FILE* f_ = nullptr;
  bool csv_ = false;
  bool header_done_ = false;
  uint64_t records_ = 0;
  std::vector<char> buf_;
  size_t used_ = 0;
};

}  // namespace nbody
