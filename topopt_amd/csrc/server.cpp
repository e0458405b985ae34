// server.cpp — persistent continuous-batching job loop.
//
// The reference serves optimisation jobs one at a time through a FIFO worker
// (tools/topopt-worker/topopt_worker.py:356-459). Here a scheduler admits
// queued jobs into slots whenever a slot and enough of the device-memory budget
// are free; each admitted job gets its own problem handle (own HIP stream,
// device-resident density, displacement and hierarchy) and runs on its own
// host thread, so the kernels of concurrently running jobs interleave on the
// GPU. A job's kernels, launch shapes and reductions do not depend on what else
// runs, so its result is the same alone and in the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "engine.hpp"
#include "formats.hpp"

using namespace topohip;

namespace {
double wall_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void job_bcs(const topohip_job& j, std::vector<int32_t>& bn, std::vector<int32_t>& bcomp,
             std::vector<double>& bv, std::vector<int32_t>& ln, std::vector<int32_t>& lc,
             std::vector<double>& lv) {
  const int nx = j.nx, ny = j.ny, nz = j.nz;
  auto node = [&](int a, int b, int c) { return (c * (ny + 1) + b) * (nx + 1) + a; };
  if (j.bc_kind == 0) {  // cantilever: x=0 face clamped, -z line load at x=nx, z=0
    for (int c = 0; c <= nz; ++c)
      for (int b = 0; b <= ny; ++b)
        for (int q = 0; q < 3; ++q) { bn.push_back(node(0, b, c)); bcomp.push_back(q); bv.push_back(0.0); }
    for (int b = 0; b <= ny; ++b) { ln.push_back(node(nx, b, 0)); lc.push_back(2); lv.push_back(-1.0 / (ny + 1)); }
  } else {  // MBB half-beam: symmetry u_x=0 at x=0, roller (u_y,u_z) at x=nx,y=0, -y load at x=0,y=ny
    for (int c = 0; c <= nz; ++c)
      for (int b = 0; b <= ny; ++b) { bn.push_back(node(0, b, c)); bcomp.push_back(0); bv.push_back(0.0); }
    for (int c = 0; c <= nz; ++c)
      for (int q = 1; q < 3; ++q) { bn.push_back(node(nx, 0, c)); bcomp.push_back(q); bv.push_back(0.0); }
    for (int c = 0; c <= nz; ++c) { ln.push_back(node(0, ny, c)); lc.push_back(1); lv.push_back(-1.0 / (nz + 1)); }
  }
}

// Per-job run records (optional): iterations.csv fed by the design loop's
// observer, the CLI's PROGRESS lines in stdout.txt as the iterations complete,
// then design.bin, fields.bin and the VARIANT line once the job is done.
struct JobOut {
  std::unique_ptr<CsvWriter> csv;
  std::FILE* log = nullptr;
  int used_multigrid = 0, mg_levels = 0;  // of the last solve the observer saw
  ~JobOut() {
    if (log) std::fclose(log);
  }
};

void observe_job(const topohip_iteration_obs* o, void* user) {
  auto* out = static_cast<JobOut*>(user);
  out->used_multigrid = o->cg_used_multigrid;
  out->mg_levels = o->cg_mg_levels;
  out->csv->append_at(0, *o, wall_clock_ms());
  const std::string line = progress_line(0, 1, o->iteration);
  std::fputs(line.c_str(), out->log);
  std::fflush(out->log);
}

// design.bin and fields.bin of the finished job (one variant, the job's volume
// fraction; the device forms the von Mises field of the final solve) and the
// VARIANT line. No margin gate runs here: the margin fields are 0 and the
// variant counts as accepted; no mesh is exported.
void write_job_records(Problem& P, const topohip_job& j, const topohip_result& r,
                       const std::vector<double>& xphys, const std::string& dir, JobOut& jo,
                       long long created_ms) {
  // run_info.json: the record of how this job was solved
  {
    const int nc = r.non_convergent ? 1 : 0;
    const double res = 0.0;
    const double ladder = j.volfrac;
    topohip_run_info ri{};
    ri.cli_version = topohip_version();
    ri.fingerprint = "dev";
    ri.mode = "minimize_plastic";
    ri.load_source = "loadcase";
    ri.solver = "Hip";
    ri.cg_multigrid_observed = r.iterations > 0 ? 1 : 0;
    ri.cg_multigrid = jo.used_multigrid;
    ri.mg_levels = jo.mg_levels;
    ri.mixed_precision = j.precision == TOPOHIP_PREC_MIXED ? 1 : 0;
    ri.solved_grid_dofs = P.g.ndof();
    ri.n_rungs = 1;
    ri.rung_non_convergent = &nc;
    const int nci = nc ? r.iterations : 0;
    ri.rung_non_convergent_iteration = &nci;
    ri.rung_non_convergent_residual = &res;
    ri.n_ladder = 1;
    ri.ladder = &ladder;
    ri.ladder_mode = "reduction";
    ri.created_wall_ms = created_ms;
    ri.iteration_csv = 1;
    const std::string js = run_info_json(ri);
    std::FILE* f = std::fopen((dir + "/run_info.json").c_str(), "wb");
    if (f) {
      std::fwrite(js.data(), 1, js.size(), f);
      std::fclose(f);
    }
  }
  const size_t nv = xphys.size(), nd = static_cast<size_t>(P.g.ndof());
  // A run whose final solve did not converge has no certified field: it gets
  // design.bin only (the reference's run leaves such a rung out of fields.bin).
  const bool have_u = P.u_valid && !r.non_convergent;
  std::vector<float> vm, du;
  float vmax = 0.f;
  if (have_u) {
    vm.resize(nv);
    du.resize(nd);
    result_fields(P, 1.0, nullptr, nullptr, vm.data(), du.data());
    for (float v : vm) vmax = std::max(vmax, v);
  }
  const topohip_grid_frame g{j.nx, j.ny, j.nz, {0.0, 0.0, 0.0}, 1.0};
  topohip_design_variant d{};
  d.requested_volume_fraction = j.volfrac;
  d.achieved_volume_fraction = r.volume_fraction;
  d.max_von_mises_mpa = vmax;
  d.accepted = have_u ? 1 : 0;
  d.iterations = r.iterations;
  d.applied_build_dir[2] = 1.0;
  d.density = xphys.data();
  write_design_file(dir + "/design.bin", g, 1, &d);
  if (!have_u) return;
  topohip_fields_variant f{};
  f.requested_volume_fraction = j.volfrac;
  f.von_mises = vm.data();
  f.displacement = du.data();
  write_fields_file(dir + "/fields.bin", g, 1, &f);
  const std::string line =
      variant_line(j.volfrac, r.volume_fraction, r.volume_fraction, 0.0, "");
  std::fputs(line.c_str(), jo.log);
  std::fflush(jo.log);
}

struct JobRec {
  int64_t id = 0;
  topohip_job job{};
  topohip_job_result res{};
  size_t footprint = 0;
  std::vector<double> xphys;
  std::string error;  // message of a failed job
};
}  // namespace

struct topohip_server {
  std::string out_dir;  // empty: no run records
  size_t budget = 0;
  int max_slots = 1;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<int64_t> queue;
  std::map<int64_t, JobRec> jobs;
  int64_t next_id = 1;
  size_t in_use = 0;
  int running = 0;
  bool stop = false;
  double t0 = 0.0;
  std::thread sched;
  std::vector<std::thread> workers;

  void run_job(int64_t id) {
    JobRec* rec;
    {
      std::lock_guard<std::mutex> lk(mu);
      rec = &jobs[id];
    }
    topohip_job j = rec->job;
    int status = TOPOHIP_OK;
    std::string err;
    topohip_result r{};
    std::vector<double> xp(static_cast<size_t>(j.nx) * j.ny * j.nz);
    try {
      (void)hipSetDevice(device);
      std::vector<int32_t> bn, bc, ln, lc;
      std::vector<double> bv, lv;
      job_bcs(j, bn, bc, bv, ln, lc, lv);
      auto P = create_problem(j.nx, j.ny, j.nz, 1.0, 0.3, nullptr, static_cast<int>(bn.size()),
                              bn.data(), bc.data(), bv.data(), j.precision, device);
      const int32_t cnt = static_cast<int32_t>(ln.size());
      set_loads(*P, 1, &cnt, ln.data(), lc.data(), lv.data());
      topohip_options o{};
      o.volfrac = j.volfrac; o.rmin = j.rmin; o.move = j.move; o.penalty = j.penalty;
      o.E0 = 1.0; o.rho_min = 1e-3; o.change_tol = 0.0; o.cg_tol = j.cg_tol;
      o.max_iterations = j.iterations; o.cg_max_iterations = 0;
      std::unique_ptr<JobOut> jo;
      std::string dir;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!out_dir.empty()) dir = out_dir + "/job_" + std::to_string(id);
      }
      const long long created_ms = wall_clock_ms();
      if (!dir.empty()) {
        jo = std::make_unique<JobOut>();
        make_dirs(dir);
        jo->csv = std::make_unique<CsvWriter>(dir + "/iterations.csv");
        jo->log = std::fopen((dir + "/stdout.txt").c_str(), "wb");
        if (!jo->log) throw Error(TOPOHIP_EINTERNAL, "cannot open " + dir + "/stdout.txt");
        P->observer = observe_job;
        P->observer_user = jo.get();
      }
      optimize(*P, o, nullptr, xp.data(), nullptr, &r);
      if (jo) write_job_records(*P, j, r, xp, dir, *jo, created_ms);
    } catch (const Error& e) {
      status = e.code;
      err = e.what();
    } catch (const std::exception& e) {
      status = TOPOHIP_EDEVICE;
      err = e.what();
    }
    std::lock_guard<std::mutex> lk(mu);
    rec->res.iterations = r.iterations;
    rec->res.compliance = r.compliance;
    rec->res.volume_fraction = r.volume_fraction;
    rec->res.status = status;
    rec->res.t_end_ms = wall_ms() - t0;
    rec->res.state = status == TOPOHIP_OK ? 2 : 3;
    rec->xphys = std::move(xp);
    rec->error = std::move(err);
    in_use -= rec->footprint;
    --running;
    cv.notify_all();
    done_cv.notify_all();
  }

  void schedule() {
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
      bool admitted = false;
      while (!queue.empty() && running < max_slots) {
        JobRec& rec = jobs[queue.front()];
        // Admit when the job fits next to the running ones; a job larger than
        // the whole budget runs alone.
        if (running > 0 && in_use + rec.footprint > budget) break;
        queue.pop_front();
        in_use += rec.footprint;
        ++running;
        rec.res.state = 1;
        rec.res.t_start_ms = wall_ms() - t0;
        workers.emplace_back(&topohip_server::run_job, this, rec.id);
        admitted = true;
      }
      if (!admitted) cv.wait(lk);
    }
  }
};

extern "C" {

size_t topohip_job_footprint(const topohip_job* j) {
  if (!j) return 0;
  const double nv = static_cast<double>(j->nx) * j->ny * j->nz;
  const double nd = 3.0 * (j->nx + 1.0) * (j->ny + 1.0) * (j->nz + 1.0);
  // fine vectors (FP64 x, r, p, b, rt, u, Ap, z + FP32 shadows), per-voxel
  // fields, coarse element matrices (FP64 build + apply copy, ~1/8 of the fine
  // element count per level) and the dense FP64 coarsest inverse (<= 6000^2).
  double bytes = nd * (8.0 * 8 + 4.0 * 6) + nv * (8.0 * 10 + 4.0);
  bytes += (nv / 8.0) * 576.0 * 12.0 * 8.0 / 7.0;
  const double ncoarse = std::min(nd, 6100.0);
  bytes += ncoarse * ncoarse * 8.0 + ncoarse * 100.0 * 8.0;
  return static_cast<size_t>(bytes * 1.1) + (64u << 20);
}

int topohip_server_create(size_t budget_bytes, int max_slots, topohip_server** out) {
  if (!out || max_slots < 1) return TOPOHIP_EINVAL;
  auto* s = new topohip_server();
  s->budget = budget_bytes;
  s->max_slots = max_slots;
  (void)hipGetDevice(&s->device);
  s->t0 = wall_ms();
  s->sched = std::thread(&topohip_server::schedule, s);
  *out = s;
  return TOPOHIP_OK;
}

int topohip_server_submit(topohip_server* s, const topohip_job* job, int64_t* job_id) {
  if (!s || !job || !job_id) return TOPOHIP_EINVAL;
  if (job->nx < 1 || job->ny < 1 || job->nz < 1 || job->iterations < 1 ||
      !(job->volfrac > 0.0 && job->volfrac <= 1.0) || !(job->rmin > 0.0) ||
      (job->bc_kind != 0 && job->bc_kind != 1))
    return TOPOHIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t id = s->next_id++;
  JobRec& rec = s->jobs[id];
  rec.id = id;
  rec.job = *job;
  rec.footprint = topohip_job_footprint(job);
  rec.res.state = 0;
  rec.res.t_submit_ms = wall_ms() - s->t0;
  s->queue.push_back(id);
  *job_id = id;
  s->cv.notify_all();
  return TOPOHIP_OK;
}

int topohip_server_poll(topohip_server* s, int64_t job_id, topohip_job_result* out) {
  if (!s || !out) return TOPOHIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return TOPOHIP_EINVAL;
  *out = it->second.res;
  return TOPOHIP_OK;
}

int topohip_server_drain(topohip_server* s) {
  if (!s) return TOPOHIP_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->done_cv.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
  return TOPOHIP_OK;
}

int topohip_server_fetch_density(topohip_server* s, int64_t job_id, double* xphys) {
  if (!s || !xphys) return TOPOHIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end() || it->second.res.state != 2) return TOPOHIP_EINVAL;
  std::copy(it->second.xphys.begin(), it->second.xphys.end(), xphys);
  return TOPOHIP_OK;
}

int topohip_server_job_error(topohip_server* s, int64_t job_id, char* buf, size_t n) {
  if (!s || !buf || n == 0) return TOPOHIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return TOPOHIP_EINVAL;
  std::strncpy(buf, it->second.error.c_str(), n - 1);
  buf[n - 1] = '\0';
  return TOPOHIP_OK;
}

int topohip_server_destroy(topohip_server* s) {
  if (!s) return TOPOHIP_EINVAL;
  topohip_server_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  s->sched.join();
  for (auto& w : s->workers) w.join();
  delete s;
  return TOPOHIP_OK;
}

int topohip_server_set_output(topohip_server* s, const char* dir) {
  if (!s) return TOPOHIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->out_dir = dir ? dir : "";
  return TOPOHIP_OK;
}

}  // extern "C"
