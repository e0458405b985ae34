// formats.cpp — run records byte-compatible with the reference's writers.
//
// Each writer emits the reference's byte layout (cited per function); the
// layouts are fixed because existing readers (the app, the LAN worker, the
// reference's own read_design_file) consume them.
#include "formats.hpp"

#include <sys/resource.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cinttypes>
#include <cstring>

#include "engine_buf.hpp"

namespace topohip {

const char kIterationCsvHeader[] =
    "rung,iter,wall_ms,compliance,achieved_vf,plateau,cg_iters,cg_multigrid,"
    "beta,hier_built,mg_cycles_attempted,infeasible,recycle_dim,"
    "active_fraction,"
    "total_ms,tail_prev_ms,filter_ms,project_ms,solve_ms,fea_ms,sens_ms,"
    "update_ms,analysis_ms,observe_ms,residual_ms,"
    "solver_build_ms,mg_build_ms,mg_ms,cg_ms,geneo_setup_ms,geneo_apply_ms,"
    "recycle_ms,"
    "fea_solves,matvecs,geneo_dim,geneo_action,geneo_burn,geneo_threshold,"
    "rss_mb,peak_rss_mb,compressed_mb,available_mb,major_faults,swapins";

namespace {

[[noreturn]] void bad(const std::string& m) { throw Error(TOPOHIP_EINVAL, m); }
[[noreturn]] void io(const std::string& m) { throw Error(TOPOHIP_EINTERNAL, m); }

// Little-endian byte sink / source (the host is little-endian x86-64; the
// explicit byte order keeps the files independent of struct layout).
struct LeOut {
  std::vector<std::uint8_t> b;
  template <typename T>
  void put(T v) {
    std::uint8_t tmp[sizeof(T)];
    std::memcpy(tmp, &v, sizeof(T));
    b.insert(b.end(), tmp, tmp + sizeof(T));
  }
  void zeros(std::size_t n) { b.insert(b.end(), n, 0); }
  template <typename T>
  void array(const T* p, std::size_t n) {
    const auto* q = reinterpret_cast<const std::uint8_t*>(p);
    b.insert(b.end(), q, q + n * sizeof(T));
  }
};

struct LeIn {
  const std::vector<std::uint8_t>& b;
  std::size_t at = 0;
  std::string what;
  template <typename T>
  T get() {
    if (at + sizeof(T) > b.size()) bad(what + ": truncated");
    T v;
    std::memcpy(&v, b.data() + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  void skip(std::size_t n) {
    if (at + n > b.size()) bad(what + ": truncated");
    at += n;
  }
};

// Write the whole container beside its final name, then rename, so a reader
// sees the complete file or the previous one (core/src/io/design_store.cpp:145-163).
void publish(const std::string& path, const std::vector<std::uint8_t>& bytes) {
  const std::string tmp = path + ".part";
  std::FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) io("cannot open for writing: " + tmp);
  const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
  if (std::fclose(f) != 0 || !ok) {
    std::remove(tmp.c_str());
    io("failed writing: " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    io("failed publishing: " + path);
  }
}

std::vector<std::uint8_t> slurp(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) bad("cannot open: " + path);
  std::vector<std::uint8_t> out;
  std::uint8_t buf[1 << 16];
  std::size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err) io("failed reading: " + path);
  return out;
}

std::string join(const std::string& dir, const std::string& name) {
  return dir.empty() ? name : dir + "/" + name;
}

}  // namespace

void make_dirs(const std::string& dir) {
  if (dir.empty()) return;
  std::string cur;
  std::size_t pos = 0;
  while (pos != std::string::npos) {
    pos = dir.find('/', pos + 1);
    cur = dir.substr(0, pos);
    if (cur.empty()) continue;
    if (::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
      io("cannot create directory " + cur + ": " + std::strerror(errno));
  }
}

// core/src/simp/observability.cpp:282-330: the column order and the printf
// conversions of append_at (%.10g objective, %.6f fractions, %.6g beta, %.3f
// phase spans, %.2f memory) define the file, so they are kept.
std::string iterations_csv_row(std::size_t rung, const topohip_iteration_obs& o,
                               long long wall_ms) {
  char buf[1024];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "%zu,%d,%lld,%.10g,%.6f,%d,%d,%d,%.6g,%d,%d,%d,%d,%.6f,"
      "%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,"
      "%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,"
      "%lld,%lld,%d,%d,%d,%d,"
      "%.2f,%.2f,%.2f,%.2f,%lld,%lld\n",
      rung, o.iteration, wall_ms, o.compliance, o.volume_fraction, o.plateau ? 1 : 0,
      o.cg_iterations, o.cg_used_multigrid ? 1 : 0, o.beta, o.cg_hier_built ? 1 : 0,
      o.cg_mg_cycles_attempted, o.infeasible ? 1 : 0, o.recycle_dim, o.active_fraction,
      o.total_ms, o.tail_prev_ms, o.filter_ms, o.project_ms, o.solve_ms, o.fea_ms, o.sens_ms,
      o.update_ms, o.analysis_ms, o.observe_ms, o.residual_ms, o.solver_build_ms,
      o.solver_mg_build_ms, o.solver_mg_ms, o.solver_cg_ms, o.solver_geneo_setup_ms,
      o.solver_geneo_apply_ms, o.solver_recycle_ms, o.fea_solves, o.matvecs, o.geneo_dim,
      o.geneo_action, o.geneo_burn, o.geneo_threshold, o.rss_mb, o.peak_rss_mb,
      o.compressed_mb, o.available_mb, o.major_faults, o.swapins);
  if (n < 0 || n >= static_cast<int>(sizeof(buf))) bad("iterations.csv: row does not fit");
  return std::string(buf, static_cast<std::size_t>(n));
}

long long wall_clock_ms() {
  using namespace std::chrono;
  return duration_cast<milliseconds>(system_clock::now().time_since_epoch()).count();
}

void sample_process_memory(topohip_iteration_obs& o) {
  constexpr double kMb = 1024.0 * 1024.0;
  o.rss_mb = o.peak_rss_mb = o.compressed_mb = o.available_mb = -1.0;
  o.major_faults = o.swapins = -1;
  struct rusage ru;
  if (getrusage(RUSAGE_SELF, &ru) == 0) {
    o.peak_rss_mb = static_cast<double>(ru.ru_maxrss) / 1024.0;  // KiB on Linux
    o.major_faults = static_cast<long long>(ru.ru_majflt);
  }
  if (std::FILE* f = std::fopen("/proc/self/statm", "r")) {
    long long total = 0, resident = 0;
    if (std::fscanf(f, "%lld %lld", &total, &resident) == 2)
      o.rss_mb = static_cast<double>(resident) * static_cast<double>(sysconf(_SC_PAGESIZE)) / kMb;
    std::fclose(f);
  }
  if (std::FILE* f = std::fopen("/proc/meminfo", "r")) {
    char key[64];
    long long kb = 0;
    while (std::fscanf(f, "%63s %lld kB\n", key, &kb) == 2)
      if (std::strcmp(key, "MemAvailable:") == 0) {
        o.available_mb = static_cast<double>(kb) / 1024.0;
        break;
      }
    std::fclose(f);
  }
}

CsvWriter::CsvWriter(const std::string& path) : path_(path) {
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_) io("IterationCsvWriter: cannot open " + path);
  if (std::fprintf(f_, "%s\n", kIterationCsvHeader) < 0 || std::fflush(f_) != 0) {
    std::fclose(f_);
    f_ = nullptr;
    io("IterationCsvWriter: failed writing header to " + path);
  }
}

CsvWriter::~CsvWriter() {
  if (f_) std::fclose(f_);
}

void CsvWriter::append_at(std::size_t rung, const topohip_iteration_obs& o, long long wall_ms) {
  const std::string row = iterations_csv_row(rung, o, wall_ms);
  if (std::fwrite(row.data(), 1, row.size(), f_) != row.size() || std::fflush(f_) != 0)
    io("IterationCsvWriter: failed writing row to " + path_);
  ++rows_;
}

// core/src/simp/observability.cpp:362-408: raw halves in grid order, then the
// one-line JSON sidecar with the same keys, order and conversions.
void write_density_snapshot(const std::string& dir, const std::string& basename, int nx, int ny,
                            int nz, double spacing, const double origin[3],
                            const std::vector<std::uint16_t>& halves, std::size_t rung,
                            int iteration, bool boundary) {
  const std::size_t voxels = static_cast<std::size_t>(nx) * ny * nz;
  if (halves.size() != voxels) bad("write_density_snapshot: density size != grid voxel count");
  make_dirs(dir);
  const std::string raw = join(dir, basename + ".f16");
  {
    std::FILE* f = std::fopen(raw.c_str(), "wb");
    if (!f) io("write_density_snapshot: cannot open " + raw);
    const bool ok = std::fwrite(halves.data(), 2, voxels, f) == voxels;
    if (std::fclose(f) != 0 || !ok) io("write_density_snapshot: failed writing " + raw);
  }
  const std::string js = join(dir, basename + ".json");
  char buf[512];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "{\"dims\":[%d,%d,%d],\"spacing\":%.10g,"
      "\"origin\":[%.10g,%.10g,%.10g],\"rung\":%zu,\"iter\":%d,"
      "\"boundary\":%s,\"voxels\":%zu,\"dtype\":\"float16\","
      "\"order\":\"x-fastest\"}\n",
      nx, ny, nz, spacing, origin[0], origin[1], origin[2], rung, iteration,
      boundary ? "true" : "false", voxels);
  std::FILE* f = std::fopen(js.c_str(), "wb");
  if (!f) io("write_density_snapshot: cannot open " + js);
  const bool ok = std::fwrite(buf, 1, static_cast<std::size_t>(n), f) == static_cast<std::size_t>(n);
  if (std::fclose(f) != 0 || !ok) io("write_density_snapshot: failed writing " + js);
}

std::string SnapshotCapture::basename(std::size_t rung, int iteration, bool boundary) {
  char base[96];
  std::snprintf(base, sizeof(base), "snap_rung%zu_iter%04d%s", rung, iteration,
                boundary ? "_boundary" : "");
  return base;
}

void SnapshotCapture::written(const std::string& base, bool boundary) {
  ++written_;
  if (boundary) return;
  iter_snaps_.push_back(base);
  while (cap_ > 0 && static_cast<int>(iter_snaps_.size()) > cap_) {
    const std::string victim = iter_snaps_.front();
    iter_snaps_.erase(iter_snaps_.begin());
    std::remove(join(dir_, victim + ".f16").c_str());
    std::remove(join(dir_, victim + ".json").c_str());
    ++evicted_;
  }
}

// IEEE binary16 -> float, exact (every half is a float).
float half_to_float(std::uint16_t h) {
  const std::uint32_t sign = static_cast<std::uint32_t>(h & 0x8000u) << 16;
  const int e = (h >> 10) & 0x1F;
  const std::uint32_t m = h & 0x3FFu;
  std::uint32_t bits;
  if (e == 31) {
    bits = sign | 0x7F800000u | (m << 13);
  } else if (e == 0) {
    if (m == 0) {
      bits = sign;
    } else {  // subnormal: m * 2^-24, exact in float
      float v = static_cast<float>(m) * 5.9604644775390625e-8f;
      std::memcpy(&bits, &v, 4);
      bits |= sign;
    }
  } else {
    bits = sign | (static_cast<std::uint32_t>(e - 15 + 127) << 23) | (m << 13);
  }
  float out;
  std::memcpy(&out, &bits, 4);
  return out;
}

std::vector<float> read_density_f16(const std::string& path, std::size_t count) {
  const std::vector<std::uint8_t> b = slurp(path);
  if (b.size() != 2 * count) bad("read_density_f16: " + path + " is not exactly the expected size");
  std::vector<float> out(count);
  for (std::size_t i = 0; i < count; ++i) {
    std::uint16_t h;
    std::memcpy(&h, b.data() + 2 * i, 2);
    out[i] = half_to_float(h);
  }
  return out;
}

// core/src/io/design_store.cpp:75-88: FNV-1a 64 over the doubles' bytes,
// started from the reference's offset basis 1469598103934665603 (one digit
// shorter than the published FNV-1a basis): the fingerprints in the design.bin
// files the reference writes are computed from it, so this one is too.
std::uint64_t design_fingerprint(const double* d, std::size_t n) {
  std::uint64_t h = 1469598103934665603ULL;
  const auto* b = reinterpret_cast<const std::uint8_t*>(d);
  for (std::size_t i = 0; i < n * sizeof(double); ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

// design.bin v1, core/include/topopt/design_store.hpp:46-86.
void write_design_file(const std::string& path, const topohip_grid_frame& g, int n,
                       const topohip_design_variant* v) {
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || n < 0 || (n > 0 && !v))
    bad("design.bin: invalid grid or variant list");
  const std::size_t voxels = static_cast<std::size_t>(g.nx) * g.ny * g.nz;
  LeOut w;
  w.put<std::uint8_t>(1);
  w.zeros(3);
  w.put<std::int32_t>(g.nx);
  w.put<std::int32_t>(g.ny);
  w.put<std::int32_t>(g.nz);
  for (int c = 0; c < 3; ++c) w.put<double>(g.origin[c]);
  w.put<double>(g.spacing);
  w.put<std::int32_t>(n);
  w.zeros(4);
  for (int i = 0; i < n; ++i) {
    const topohip_design_variant& d = v[i];
    if (!d.density) bad("design.bin: variant " + std::to_string(i) + " has no density");
    w.put<double>(d.requested_volume_fraction);
    w.put<double>(d.achieved_volume_fraction);
    w.put<double>(d.margin_worst_case);
    w.put<double>(d.margin_effective);
    w.put<double>(d.max_von_mises_mpa);
    w.put<std::int32_t>(d.accepted ? 1 : 0);
    w.put<std::int32_t>(d.iterations);
    for (int c = 0; c < 3; ++c) w.put<double>(d.applied_build_dir[c]);
    w.put<std::int32_t>(d.build_direction_auto_applied ? 1 : 0);
    w.put<std::int32_t>(d.export_baked ? 1 : 0);
    w.put<std::uint64_t>(design_fingerprint(d.density, voxels));
    w.put<std::int64_t>(static_cast<std::int64_t>(voxels));
    w.array(d.density, voxels);
  }
  publish(path, w.b);
}

// core/src/io/design_store.cpp:167-243: version first, header sanity, per
// variant count and fingerprint.
DesignFile read_design_file(const std::string& path) {
  const std::vector<std::uint8_t> bytes = slurp(path);
  LeIn r{bytes, 0, "design.bin"};
  const int version = r.get<std::uint8_t>();
  if (version != 1)
    bad("design.bin: unrecognised version " + std::to_string(version) +
        " (this build reads version 1)");
  r.skip(3);
  DesignFile out;
  out.grid.nx = r.get<std::int32_t>();
  out.grid.ny = r.get<std::int32_t>();
  out.grid.nz = r.get<std::int32_t>();
  for (int c = 0; c < 3; ++c) out.grid.origin[c] = r.get<double>();
  out.grid.spacing = r.get<double>();
  const int count = r.get<std::int32_t>();
  r.skip(4);
  if (out.grid.nx <= 0 || out.grid.ny <= 0 || out.grid.nz <= 0 || count < 0)
    bad("design.bin: nonsensical header");
  const std::size_t voxels = static_cast<std::size_t>(out.grid.nx) * out.grid.ny * out.grid.nz;
  for (int i = 0; i < count; ++i) {
    topohip_design_variant d{};
    d.requested_volume_fraction = r.get<double>();
    d.achieved_volume_fraction = r.get<double>();
    d.margin_worst_case = r.get<double>();
    d.margin_effective = r.get<double>();
    d.max_von_mises_mpa = r.get<double>();
    d.accepted = r.get<std::int32_t>() != 0;
    d.iterations = r.get<std::int32_t>();
    for (int c = 0; c < 3; ++c) d.applied_build_dir[c] = r.get<double>();
    d.build_direction_auto_applied = r.get<std::int32_t>() != 0;
    d.export_baked = r.get<std::int32_t>() != 0;
    d.fingerprint = r.get<std::uint64_t>();
    const std::int64_t cnt = r.get<std::int64_t>();
    if (cnt < 0 || static_cast<std::size_t>(cnt) != voxels)
      bad("design.bin: variant " + std::to_string(i) + " density count " + std::to_string(cnt) +
          " != header voxel count " + std::to_string(voxels));
    std::vector<double> dens(voxels);
    if (r.at + voxels * 8 > bytes.size()) bad("design.bin: truncated");
    std::memcpy(dens.data(), bytes.data() + r.at, voxels * 8);
    r.at += voxels * 8;
    const std::uint64_t fp = design_fingerprint(dens.data(), voxels);
    if (fp != d.fingerprint)
      bad("design.bin: variant " + std::to_string(i) +
          " density does not match its recorded fingerprint");
    out.meta.push_back(d);
    out.density.push_back(std::move(dens));
  }
  return out;
}

// fields.bin v1, core/include/topopt/fields.hpp:34-60.
void write_fields_file(const std::string& path, const topohip_grid_frame& g, int n,
                       const topohip_fields_variant* v) {
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || n < 0 || (n > 0 && !v))
    bad("fields.bin: invalid grid or variant list");
  const std::size_t voxels = static_cast<std::size_t>(g.nx) * g.ny * g.nz;
  const std::size_t nodes = static_cast<std::size_t>(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  LeOut w;
  w.put<std::uint8_t>(1);
  w.zeros(3);
  w.put<std::int32_t>(g.nx);
  w.put<std::int32_t>(g.ny);
  w.put<std::int32_t>(g.nz);
  for (int c = 0; c < 3; ++c) w.put<double>(g.origin[c]);
  w.put<double>(g.spacing);
  w.put<double>(g.spacing * g.spacing * g.spacing);
  w.put<std::int32_t>(n);
  w.zeros(4);
  for (int i = 0; i < n; ++i) {
    const topohip_fields_variant& f = v[i];
    // core/src/io/fields.cpp:93-105: a variant's fields must match the grid
    if (!f.von_mises || !f.displacement)
      bad("fields.bin: variant " + std::to_string(i) + " lacks its von Mises or displacement field");
    const std::size_t nvm = voxels, nd = 3 * nodes;
    w.put<double>(f.requested_volume_fraction);
    w.put<double>(f.mass_grams);
    w.put<std::int32_t>(f.support_volume_voxels);
    w.zeros(4);
    w.put<std::int64_t>(static_cast<std::int64_t>(nvm));
    w.put<std::int64_t>(0);  // stress tensor: empty in v1
    w.put<std::int64_t>(static_cast<std::int64_t>(nd));
    w.array(f.von_mises, nvm);
    w.array(f.displacement, nd);
  }
  publish(path, w.b);
}

// core/src/cli/run_job.cpp:10225 and :10256-10259.
std::string progress_line(std::size_t rung, std::size_t rungs, int iteration) {
  char buf[128];
  std::snprintf(buf, sizeof(buf), "PROGRESS rung=%zu rungs=%zu iter=%d\n", rung, rungs, iteration);
  return buf;
}

// run_info_json (core/src/simp/observability.cpp:476-1284): `"key": value`
// lines of two-space indent in the reference's order, strings with `"` and `\`
// escaped, reals as %.10g, 64-bit integers as %lld, the multigrid outcome null
// until observed; the last key carries no comma.
std::string run_info_json(const topohip_run_info& in) {
  std::string s = "{\n";
  char nb[64];
  auto esc = [](const char* v) {
    std::string o;
    for (const char* c = v ? v : ""; *c; ++c) {
      if (*c == '"' || *c == '\\') o.push_back('\\');
      o.push_back(*c);
    }
    return o;
  };
  auto num = [&](const char* k, const std::string& v, bool comma = true) {
    s += "  \"";
    s += k;
    s += "\": " + v + (comma ? ",\n" : "\n");
  };
  auto str = [&](const char* k, const char* v) { num(k, "\"" + esc(v) + "\""); };
  auto fmt = [&](double d) {
    std::snprintf(nb, sizeof(nb), "%.10g", d);
    return std::string(nb);
  };
  auto fmt_ll = [&](long long v) {
    std::snprintf(nb, sizeof(nb), "%lld", v);
    return std::string(nb);
  };
  auto fmt_i = [&](int v) {
    std::snprintf(nb, sizeof(nb), "%d", v);
    return std::string(nb);
  };
  auto b = [](int v) { return std::string(v ? "true" : "false"); };
  str("cli_version", in.cli_version);
  str("fingerprint", in.fingerprint);
  str("mode", in.mode);
  str("material", in.material);
  str("source_format", in.source_format);
  num("resolution", fmt_i(in.resolution));
  str("load_source", in.load_source);
  str("solver", in.solver);
  num("cg_multigrid", in.cg_multigrid_observed ? b(in.cg_multigrid) : "null");
  num("mg_levels", in.cg_multigrid_observed ? fmt_i(in.mg_levels) : "null");
  num("mg_mode", in.mg_mode && *in.mg_mode ? "\"" + esc(in.mg_mode) + "\"" : "null");
  num("galerkin_block_cache", b(1));  // the level-1 product always uses the 8 S_c blocks
  num("mixed_precision", b(in.mixed_precision));
  num("matfree_threads", fmt_i(in.matfree_threads));
  num("krylov_recycling", b(0));
  num("krylov_recycle_dim", fmt_i(0));
  num("warm_start_inherit", b(in.warm_start_inherit));
  num("solved_grid_dofs", fmt_ll(in.solved_grid_dofs));
  num("projection", b(in.projection));
  std::string nc = "[", ni = "[", nr = "[";
  for (int i = 0; i < in.n_rungs; ++i) {
    const char* sep = i ? ", " : "";
    nc += sep + b(in.rung_non_convergent ? in.rung_non_convergent[i] : 0);
    ni += sep + fmt_i(in.rung_non_convergent_iteration ? in.rung_non_convergent_iteration[i] : 0);
    nr += sep + fmt(in.rung_non_convergent_residual ? in.rung_non_convergent_residual[i] : 0.0);
  }
  num("rung_non_convergent", nc + "]");
  num("rung_non_convergent_iteration", ni + "]");
  num("rung_non_convergent_residual", nr + "]");
  std::string ladder = "[";
  for (int i = 0; i < in.n_ladder; ++i) ladder += (i ? ", " : "") + fmt(in.ladder[i]);
  num("ladder", ladder + "]");
  str("ladder_mode", in.ladder_mode);
  num("created_wall_ms", fmt_ll(in.created_wall_ms));
  num("iteration_csv", b(in.iteration_csv));
  num("density_snapshots", b(in.density_snapshots));
  num("snapshot_every", fmt_i(in.snapshot_every));
  num("snapshot_cap", fmt_i(in.snapshot_cap), false);
  s += "}\n";
  return s;
}

std::string variant_line(double vf, double achieved, double printed, double margin,
                         const std::string& mesh) {
  char buf[256];
  const int n = std::snprintf(buf, sizeof(buf),
                              "VARIANT vf=%.6f achieved=%.6f printed=%.6f margin=%.6g accepted=1 mesh=",
                              vf, achieved, printed, margin);
  return std::string(buf, static_cast<std::size_t>(n)) + mesh + "\n";
}

}  // namespace topohip
