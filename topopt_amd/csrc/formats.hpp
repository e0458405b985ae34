// formats.hpp — the run records a design loop leaves behind, byte-compatible
// with the reference: iterations.csv, density snapshots (.f16 + .json),
// design.bin v1, fields.bin v1 and the CLI's PROGRESS / VARIANT stdout lines.
// Host code only; the arrays they carry are formed on the device (fields.hip).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/topopt_hip.h"

namespace topohip {

// kIterationCsvHeader, core/src/simp/observability.cpp:252-262.
extern const char kIterationCsvHeader[];

// IterationCsvWriter::append_at's row, core/src/simp/observability.cpp:282-330.
std::string iterations_csv_row(std::size_t rung, const topohip_iteration_obs& o,
                               long long wall_ms);

// mkdir -p.
void make_dirs(const std::string& dir);

// Epoch milliseconds from the system clock (the CSV's wall_ms column).
long long wall_clock_ms();

// process_memory's Linux branch (core/src/simp/observability.cpp:116-205) into
// the memory columns of o.
void sample_process_memory(topohip_iteration_obs& o);

class CsvWriter {
 public:
  explicit CsvWriter(const std::string& path);
  ~CsvWriter();
  CsvWriter(const CsvWriter&) = delete;
  CsvWriter& operator=(const CsvWriter&) = delete;
  void append_at(std::size_t rung, const topohip_iteration_obs& o, long long wall_ms);
  std::size_t rows() const { return rows_; }

 private:
  std::FILE* f_ = nullptr;
  std::string path_;
  std::size_t rows_ = 0;
};

void write_density_snapshot(const std::string& dir, const std::string& basename, int nx, int ny,
                            int nz, double spacing, const double origin[3],
                            const std::vector<std::uint16_t>& halves, std::size_t rung,
                            int iteration, bool boundary);
std::vector<float> read_density_f16(const std::string& path, std::size_t count);

// SnapshotCapture (core/src/simp/observability.cpp:426-455): rung boundaries
// always written and never evicted; per-iteration snapshots every every_n
// iterations, the oldest evicted once more than cap (> 0) are kept.
class SnapshotCapture {
 public:
  SnapshotCapture(std::string dir, int every_n, int cap)
      : dir_(std::move(dir)), every_n_(every_n > 0 ? every_n : 1), cap_(cap) {}
  // Whether this event is written; the basename it would take.
  bool wants(int iteration, bool boundary) const {
    return boundary || iteration % every_n_ == 0;
  }
  static std::string basename(std::size_t rung, int iteration, bool boundary);
  // Record a written snapshot and apply the cap.
  void written(const std::string& base, bool boundary);
  const std::string& dir() const { return dir_; }
  std::size_t n_written() const { return written_; }
  std::size_t n_evicted() const { return evicted_; }

 private:
  std::string dir_;
  int every_n_;
  int cap_;
  std::size_t written_ = 0, evicted_ = 0;
  std::vector<std::string> iter_snaps_;
};
float half_to_float(std::uint16_t h);

std::uint64_t design_fingerprint(const double* d, std::size_t n);

void write_design_file(const std::string& path, const topohip_grid_frame& g, int n,
                       const topohip_design_variant* v);

struct DesignFile {
  topohip_grid_frame grid{};
  std::vector<topohip_design_variant> meta;
  std::vector<std::vector<double>> density;
};
DesignFile read_design_file(const std::string& path);

void write_fields_file(const std::string& path, const topohip_grid_frame& g, int n,
                       const topohip_fields_variant* v);

std::string progress_line(std::size_t rung, std::size_t rungs, int iteration);
// run_info.json (the subset this library's solve knows), reference key order.
std::string run_info_json(const topohip_run_info& info);
std::string variant_line(double vf, double achieved, double printed, double margin,
                         const std::string& mesh);

}  // namespace topohip
