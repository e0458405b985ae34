// oracle.cpp — FP64 CPU restatement of the reference voxel-SIMP inner loop.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp). Built with -ffp-contract=off so the
// element apply is mul+add exactly as the reference's axpy24 (no FMA).
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>

namespace oracle {

namespace {
double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Fork-join pool for the colour-partitioned apply. Each colour's element range
// is split into contiguous chunks; within a colour no two elements share a
// node, so results are bit-identical for any thread count — the same argument
// as core/src/fea/matfree.cpp:219-346.
class Pool {
 public:
  explicit Pool(int n) : n_(n) {
    for (int t = 1; t < n_; ++t) workers_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int size() const { return n_; }
  // Run body(lo,hi) over [b,e) split in n_ contiguous chunks.
  void run(int b, int e, const std::function<void(int, int)>& body) {
    if (e <= b) return;
    if (n_ == 1 || e - b < 64) { body(b, e); return; }
    {
      std::lock_guard<std::mutex> lk(mu_);
      body_ = &body; b_ = b; e_ = e; pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    chunk(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  void chunk(int t) {
    const long long len = e_ - b_;
    const int lo = b_ + static_cast<int>(len * t / n_);
    const int hi = b_ + static_cast<int>(len * (t + 1) / n_);
    if (hi > lo) (*body_)(lo, hi);
  }
  void loop(int t) {
    long long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      chunk(t);
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (--pending_ == 0) done_cv_.notify_one();
      }
    }
  }
  int n_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int, int)>* body_ = nullptr;
  int b_ = 0, e_ = 0, pending_ = 0;
  long long gen_ = 0;
  bool stop_ = false;
};

int g_threads = 1;
std::unique_ptr<Pool> g_pool;
Pool& pool() {
  if (!g_pool || g_pool->size() != g_threads) {
    g_pool.reset();
    g_pool = std::make_unique<Pool>(g_threads);
  }
  return *g_pool;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double norm2(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }
}  // namespace

void set_threads(int n) {
  if (n <= 0) n = static_cast<int>(std::thread::hardware_concurrency());
  if (n <= 0) n = 1;
  g_threads = n;
}
int threads() { return g_threads; }

std::size_t Grid::solid_count() const {
  if (solid.empty()) return voxels();
  std::size_t n = 0;
  for (auto s : solid) n += (s != 0);
  return n;
}

// ---------------------------------------------------------------------------
// core/src/fea/hex_element.cpp:14-16 (corner natural coordinates) and :23-114.
void hex8_stiffness(double E, double nu, double h, double* K) {
  if (!(E > 0.0)) throw std::invalid_argument("hex8_stiffness: youngs_modulus must be > 0");
  if (!(h > 0.0)) throw std::invalid_argument("hex8_stiffness: element_size must be > 0");
  if (!(nu > -1.0 && nu < 0.5))
    throw std::invalid_argument("hex8_stiffness: poisson must be in (-1, 0.5)");
  static const double xs[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
  static const double ys[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
  static const double zs[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double D[6][6] = {};
  D[0][0] = D[1][1] = D[2][2] = c * (1.0 - nu);
  D[0][1] = D[0][2] = D[1][0] = D[1][2] = D[2][0] = D[2][1] = c * nu;
  D[3][3] = D[4][4] = D[5][5] = c * (1.0 - 2.0 * nu) / 2.0;
  const double g = 1.0 / std::sqrt(3.0);
  const double q[2] = {-g, +g};
  const double half = h / 2.0, detJ = half * half * half, s = 2.0 / h;
  for (int i = 0; i < kDof * kDof; ++i) K[i] = 0.0;
  for (int ig = 0; ig < 2; ++ig)
    for (int jg = 0; jg < 2; ++jg)
      for (int kg = 0; kg < 2; ++kg) {
        const double xi = q[ig], et = q[jg], ze = q[kg];
        double dx[8], dy[8], dz[8];
        for (int a = 0; a < 8; ++a) {
          dx[a] = s * 0.125 * xs[a] * (1.0 + et * ys[a]) * (1.0 + ze * zs[a]);
          dy[a] = s * 0.125 * ys[a] * (1.0 + xi * xs[a]) * (1.0 + ze * zs[a]);
          dz[a] = s * 0.125 * zs[a] * (1.0 + xi * xs[a]) * (1.0 + et * ys[a]);
        }
        double B[6][24] = {};
        for (int a = 0; a < 8; ++a) {
          B[0][3 * a] = dx[a];
          B[1][3 * a + 1] = dy[a];
          B[2][3 * a + 2] = dz[a];
          B[3][3 * a] = dy[a];
          B[3][3 * a + 1] = dx[a];
          B[4][3 * a + 1] = dz[a];
          B[4][3 * a + 2] = dy[a];
          B[5][3 * a] = dz[a];
          B[5][3 * a + 2] = dx[a];
        }
        double DB[6][24];
        for (int r = 0; r < 6; ++r)
          for (int col = 0; col < 24; ++col) {
            double acc = 0.0;
            for (int m = 0; m < 6; ++m) acc += D[r][m] * B[m][col];
            DB[r][col] = acc;
          }
        for (int r = 0; r < 24; ++r)
          for (int col = 0; col < 24; ++col) {
            double acc = 0.0;
            for (int m = 0; m < 6; ++m) acc += B[m][r] * DB[m][col];
            K[r * 24 + col] += acc * detJ;
          }
      }
}

// core/src/fea/hex_element.cpp:242-316. The full 6x24 B and 6x6 D products, in
// the reference's order, at the given natural coordinates.
Hex8Stress hex8_stress(double E, double nu, double h, const double* ue, double xi, double et,
                       double ze) {
  if (!(E > 0.0)) throw std::invalid_argument("hex8_stress: youngs_modulus must be > 0");
  if (!(h > 0.0)) throw std::invalid_argument("hex8_stress: element_size must be > 0");
  if (!(nu > -1.0 && nu < 0.5))
    throw std::invalid_argument("hex8_stress: poisson must be in (-1, 0.5)");
  static const double xs[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
  static const double ys[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
  static const double zs[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double D[6][6] = {};
  D[0][0] = D[1][1] = D[2][2] = c * (1.0 - nu);
  D[0][1] = D[0][2] = D[1][0] = D[1][2] = D[2][0] = D[2][1] = c * nu;
  D[3][3] = D[4][4] = D[5][5] = c * (1.0 - 2.0 * nu) / 2.0;
  const double s = 2.0 / h;
  double B[6][24] = {};
  for (int a = 0; a < 8; ++a) {
    const double dx = s * 0.125 * xs[a] * (1.0 + et * ys[a]) * (1.0 + ze * zs[a]);
    const double dy = s * 0.125 * ys[a] * (1.0 + xi * xs[a]) * (1.0 + ze * zs[a]);
    const double dz = s * 0.125 * zs[a] * (1.0 + xi * xs[a]) * (1.0 + et * ys[a]);
    B[0][3 * a] = dx;
    B[1][3 * a + 1] = dy;
    B[2][3 * a + 2] = dz;
    B[3][3 * a] = dy;
    B[3][3 * a + 1] = dx;
    B[4][3 * a + 1] = dz;
    B[4][3 * a + 2] = dy;
    B[5][3 * a] = dz;
    B[5][3 * a + 2] = dx;
  }
  double eps[6];
  for (int r = 0; r < 6; ++r) {
    double acc = 0.0;
    for (int col = 0; col < 24; ++col) acc += B[r][col] * ue[col];
    eps[r] = acc;
  }
  Hex8Stress out{};
  for (int r = 0; r < 6; ++r) {
    double acc = 0.0;
    for (int m = 0; m < 6; ++m) acc += D[r][m] * eps[m];
    out.sigma[r] = acc;
  }
  const double* q = out.sigma;
  out.von_mises = std::sqrt(0.5 * ((q[0] - q[1]) * (q[0] - q[1]) + (q[1] - q[2]) * (q[1] - q[2]) +
                                   (q[2] - q[0]) * (q[2] - q[0])) +
                            3.0 * (q[3] * q[3] + q[4] * q[4] + q[5] * q[5]));
  return out;
}

std::array<int, 8> element_nodes(const Grid& g, int i, int j, int k) {
  return {g.node_index(i, j, k),         g.node_index(i + 1, j, k),
          g.node_index(i + 1, j + 1, k), g.node_index(i, j + 1, k),
          g.node_index(i, j, k + 1),     g.node_index(i + 1, j, k + 1),
          g.node_index(i + 1, j + 1, k + 1), g.node_index(i, j + 1, k + 1)};
}

std::vector<NodalLoad> self_weight_loads(const Grid& g, double density,
                                         double gravity, const double dir[3]) {
  if (density < 0.0 || gravity < 0.0)
    throw std::invalid_argument("self_weight_loads: density/gravity must be >= 0");
  const double len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (!(len > 1e-12)) throw std::invalid_argument("self_weight_loads: zero direction");
  const double d[3] = {dir[0] / len, dir[1] / len, dir[2] / len};
  const double per_node = gravity * density * g.h * g.h * g.h / 8.0;
  std::vector<double> nf(static_cast<std::size_t>(g.nodes()) * 3, 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        for (int n : element_nodes(g, i, j, k))
          for (int c = 0; c < 3; ++c) nf[3 * static_cast<std::size_t>(n) + c] += per_node * d[c];
      }
  std::vector<NodalLoad> out;
  for (int n = 0; n < g.nodes(); ++n)
    for (int c = 0; c < 3; ++c) {
      const double v = nf[3 * static_cast<std::size_t>(n) + c];
      if (v != 0.0) out.push_back({n, c, v});
    }
  return out;
}

// ---------------------------------------------------------------------------
std::vector<MfElem> build_elems(const Grid& g, const std::vector<double>* youngs,
                                std::vector<int>* coff) {
  if (youngs && youngs->size() != g.voxels())
    throw std::invalid_argument("fea_solve_mgcg_matfree: per-voxel modulus vector size != voxel_count");
  std::vector<MfElem> buckets[kNumColors];
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        MfElem el;
        if (youngs) {
          el.factor = (*youngs)[g.index(i, j, k)];
          if (!(el.factor > 0.0))
            throw std::invalid_argument(
                "fea_solve_mgcg_matfree: per-voxel Young's modulus must be > 0 on a solid voxel");
        }
        const auto en = element_nodes(g, i, j, k);
        for (int a = 0; a < 8; ++a)
          for (int c = 0; c < 3; ++c) el.edof[3 * a + c] = 3 * en[a] + c;
        const int color = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
        buckets[color].push_back(el);
      }
  std::vector<MfElem> elems;
  elems.reserve(g.solid_count());
  if (coff) coff->assign(kNumColors + 1, 0);
  for (int c = 0; c < kNumColors; ++c) {
    if (coff) (*coff)[c] = static_cast<int>(elems.size());
    elems.insert(elems.end(), buckets[c].begin(), buckets[c].end());
  }
  if (coff) (*coff)[kNumColors] = static_cast<int>(elems.size());
  return elems;
}

void apply_full(const std::vector<MfElem>& elems, const std::vector<int>& coff,
                const double* Ke, const double* x, double* y, int ndof) {
  std::fill(y, y + ndof, 0.0);
  double KeCM[kDof * kDof];
  for (int r = 0; r < kDof; ++r)
    for (int c = 0; c < kDof; ++c) KeCM[c * kDof + r] = Ke[r * kDof + c];
  const std::function<void(int, int)> body = [&](int lo, int hi) {
    double ul[kDof], res[kDof];
    for (int e = lo; e < hi; ++e) {
      const MfElem& el = elems[static_cast<std::size_t>(e)];
      for (int r = 0; r < kDof; ++r) { ul[r] = x[el.edof[r]]; res[r] = 0.0; }
      for (int c = 0; c < kDof; ++c) {
        const double w = ul[c];
        const double* col = &KeCM[c * kDof];
        for (int r = 0; r < kDof; ++r) res[r] += w * col[r];
      }
      for (int r = 0; r < kDof; ++r) y[el.edof[r]] += el.factor * res[r];
    }
  };
  for (int c = 0; c < kNumColors; ++c) pool().run(coff[c], coff[c + 1], body);
}

std::vector<double> matfree_apply(const Grid& g, const std::vector<double>& youngs,
                                  double nu, const std::vector<double>& u) {
  const int ndof = 3 * g.nodes();
  if (static_cast<int>(u.size()) != ndof)
    throw std::invalid_argument("fea_matfree_apply: u size != 3*node_count");
  double Ke[kDof * kDof];
  hex8_stiffness(1.0, nu, g.h, Ke);
  std::vector<int> coff;
  const auto elems = build_elems(g, &youngs, &coff);
  std::vector<double> y(static_cast<std::size_t>(ndof));
  apply_full(elems, coff, Ke, u.data(), y.data(), ndof);
  return y;
}

// ---------------------------------------------------------------------------
void Reduced::apply_kgg(const double* xg, double* yg) const {
  ++matvecs;
  std::fill(xfull.begin(), xfull.end(), 0.0);
  for (int k = 0; k < ng; ++k) xfull[static_cast<std::size_t>(kept_global[k])] = xg[k];
  apply_full(elems, color_offsets, Ke, xfull.data(), yfull.data(), ndof);
  for (int k = 0; k < ng; ++k) yg[k] = yfull[static_cast<std::size_t>(kept_global[k])];
}

Reduced build_reduced(const Grid& g, const std::vector<double>& youngs, double nu,
                      const std::vector<DirichletBC>& bcs,
                      const std::vector<NodalLoad>& loads, CgInfo* info) {
  const char* who = "fea_solve_mgcg_matfree";
  const int nn = g.nodes();
  const int ndof = 3 * nn;
  Reduced m;
  m.ndof = ndof;
  hex8_stiffness(1.0, nu, g.h, m.Ke);
  m.elems = build_elems(g, &youngs, &m.color_offsets);
  m.xfull.assign(static_cast<std::size_t>(ndof), 0.0);
  m.yfull.assign(static_cast<std::size_t>(ndof), 0.0);

  std::vector<char> fixed(static_cast<std::size_t>(ndof), 0);
  m.up.assign(static_cast<std::size_t>(ndof), 0.0);
  bool nonzero_bc = false;
  for (const auto& bc : bcs) {
    if (bc.node < 0 || bc.node >= nn || bc.comp < 0 || bc.comp > 2)
      throw std::invalid_argument(std::string(who) + ": BC index out of range");
    const int d = 3 * bc.node + bc.comp;
    fixed[static_cast<std::size_t>(d)] = 1;
    m.up[static_cast<std::size_t>(d)] = bc.value;
    if (bc.value != 0.0) nonzero_bc = true;
  }
  std::vector<double> f(static_cast<std::size_t>(ndof), 0.0);
  for (const auto& l : loads) {
    if (l.node < 0 || l.node >= nn || l.comp < 0 || l.comp > 2)
      throw std::invalid_argument(std::string(who) + ": load index out of range");
    f[3 * static_cast<std::size_t>(l.node) + l.comp] += l.value;
  }
  std::vector<double> rhs = f;
  if (nonzero_bc) {
    std::vector<double> kup(static_cast<std::size_t>(ndof));
    apply_full(m.elems, m.color_offsets, m.Ke, m.up.data(), kup.data(), ndof);
    for (int d = 0; d < ndof; ++d) rhs[d] -= kup[d];
  }
  std::vector<int> free_of(static_cast<std::size_t>(ndof), -1);
  int nf = 0;
  for (int d = 0; d < ndof; ++d)
    if (!fixed[d]) free_of[d] = nf++;
  std::vector<char> touched(static_cast<std::size_t>(nf), 0);
  std::vector<double> diag(static_cast<std::size_t>(ndof), 0.0);
  for (const auto& el : m.elems)
    for (int r = 0; r < kDof; ++r) {
      const int fr = free_of[el.edof[r]];
      if (fr >= 0) touched[fr] = 1;
      diag[el.edof[r]] += el.factor * m.Ke[r * kDof + r];
    }
  if (nf > 0) {
    bool any = false;
    for (int i = 0; i < nf; ++i) if (touched[i]) { any = true; break; }
    if (!any) {
      if (info) { info->converged = false; info->iterations = 0; info->residual = 0.0; }
      throw std::runtime_error(std::string(who) +
                               ": singular system (no stiffness — every free DOF is void)");
    }
  }
  double rmax = 0.0;
  for (int d = 0; d < ndof; ++d)
    if (free_of[d] >= 0) rmax = std::max(rmax, std::fabs(rhs[d]));
  const double load_tol = 1e-9 * rmax;
  for (int d = 0; d < ndof; ++d) {
    const int fr = free_of[d];
    if (fr < 0 || touched[fr]) continue;
    if (std::fabs(rhs[d]) > load_tol) {
      if (info) { info->converged = false; info->iterations = 0; info->residual = 0.0; }
      throw std::runtime_error(std::string(who) +
                               ": under-constrained system (load applied to a void DOF with no "
                               "stiffness — no equilibrium possible)");
    }
  }
  for (int d = 0; d < ndof; ++d) {
    const int fr = free_of[d];
    if (fr < 0 || !touched[fr]) continue;
    m.kept_global.push_back(d);
  }
  m.ng = static_cast<int>(m.kept_global.size());
  m.rg.resize(static_cast<std::size_t>(m.ng));
  m.invdiag.resize(static_cast<std::size_t>(m.ng));
  for (int k = 0; k < m.ng; ++k) {
    const int d = m.kept_global[k];
    m.rg[k] = rhs[d];
    m.invdiag[k] = 1.0 / diag[d];
  }
  return m;
}

void cg_solve(const Reduced& m, double tol, int max_it, std::vector<double>& x,
              int& iters, double& err, bool& converged) {
  const int n = m.ng;
  const int maxIters = max_it > 0 ? max_it : 2 * n;
  const double rhsNorm2 = dot(m.rg, m.rg);
  if (rhsNorm2 == 0.0) {
    std::fill(x.begin(), x.end(), 0.0);
    iters = 0; err = 0.0; converged = true;
    return;
  }
  double threshold = tol * tol * rhsNorm2;
  const double tiny = std::numeric_limits<double>::min();
  if (threshold < tiny) threshold = tiny;
  std::vector<double> r(n), tmp(n), p(n), z(n);
  m.apply_kgg(x.data(), tmp.data());
  for (int i = 0; i < n; ++i) r[i] = m.rg[i] - tmp[i];
  double rn2 = dot(r, r);
  if (rn2 < threshold) {
    iters = 0; err = std::sqrt(rn2 / rhsNorm2); converged = err <= tol;
    return;
  }
  for (int i = 0; i < n; ++i) p[i] = m.invdiag[i] * r[i];
  double absNew = dot(r, p);
  int i = 0;
  while (i < maxIters) {
    m.apply_kgg(p.data(), tmp.data());
    const double alpha = absNew / dot(p, tmp);
    for (int q = 0; q < n; ++q) { x[q] += alpha * p[q]; r[q] -= alpha * tmp[q]; }
    rn2 = dot(r, r);
    if (rn2 < threshold) break;
    for (int q = 0; q < n; ++q) z[q] = m.invdiag[q] * r[q];
    const double absOld = absNew;
    absNew = dot(r, z);
    const double beta = absNew / absOld;
    for (int q = 0; q < n; ++q) p[q] = z[q] + beta * p[q];
    ++i;
  }
  err = std::sqrt(rn2 / rhsNorm2);
  iters = i;
  bool finite = true;
  for (double v : x) if (!std::isfinite(v)) { finite = false; break; }
  converged = (err <= tol) && finite;
}

// ---------------------------------------------------------------------------
// Sparse helpers.
void Csr::spmv(const double* x, double* y) const {
  for (int r = 0; r < rows; ++r) {
    double s = 0.0;
    for (int p = ptr[r]; p < ptr[r + 1]; ++p) s += val[p] * x[col[p]];
    y[r] = s;
  }
}
void Csr::spmv_t(const double* x, double* y) const {
  std::fill(y, y + cols, 0.0);
  for (int r = 0; r < rows; ++r) {
    const double xr = x[r];
    for (int p = ptr[r]; p < ptr[r + 1]; ++p) y[col[p]] += val[p] * xr;
  }
}

namespace {
Csr transpose(const Csr& A) {
  Csr T;
  T.rows = A.cols; T.cols = A.rows;
  T.ptr.assign(static_cast<std::size_t>(T.rows) + 1, 0);
  for (int c : A.col) ++T.ptr[c + 1];
  for (int r = 0; r < T.rows; ++r) T.ptr[r + 1] += T.ptr[r];
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  std::vector<int> cur(T.ptr.begin(), T.ptr.end() - 1);
  for (int r = 0; r < A.rows; ++r)
    for (int p = A.ptr[r]; p < A.ptr[r + 1]; ++p) {
      const int q = cur[A.col[p]]++;
      T.col[q] = r;
      T.val[q] = A.val[p];
    }
  return T;
}
// C = A * B with a dense sparse-accumulator row (columns sorted ascending).
Csr multiply(const Csr& A, const Csr& B) {
  Csr C;
  C.rows = A.rows; C.cols = B.cols;
  C.ptr.assign(static_cast<std::size_t>(C.rows) + 1, 0);
  std::vector<double> acc(static_cast<std::size_t>(B.cols), 0.0);
  std::vector<int> mark(static_cast<std::size_t>(B.cols), -1), idx;
  for (int r = 0; r < A.rows; ++r) {
    idx.clear();
    for (int p = A.ptr[r]; p < A.ptr[r + 1]; ++p) {
      const int k = A.col[p];
      const double a = A.val[p];
      for (int q = B.ptr[k]; q < B.ptr[k + 1]; ++q) {
        const int c = B.col[q];
        if (mark[c] != r) { mark[c] = r; acc[c] = 0.0; idx.push_back(c); }
        acc[c] += a * B.val[q];
      }
    }
    std::sort(idx.begin(), idx.end());
    for (int c : idx) { C.col.push_back(c); C.val.push_back(acc[c]); }
    C.ptr[r + 1] = static_cast<int>(C.col.size());
  }
  return C;
}
// COO -> CSR, summing duplicates in insertion order.
struct Trip { int r, c; double v; };
Csr from_trips(int rows, int cols, std::vector<Trip>& t) {
  std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  Csr A;
  A.rows = rows; A.cols = cols;
  A.ptr.assign(static_cast<std::size_t>(rows) + 1, 0);
  for (std::size_t i = 0; i < t.size();) {
    std::size_t j = i;
    double s = 0.0;
    while (j < t.size() && t[j].r == t[i].r && t[j].c == t[i].c) s += t[j++].v;
    A.col.push_back(t[i].c);
    A.val.push_back(s);
    ++A.ptr[t[i].r + 1];
    i = j;
  }
  for (int r = 0; r < rows; ++r) A.ptr[r + 1] += A.ptr[r];
  return A;
}
std::vector<double> inverse_diagonal(const Csr& A) {
  std::vector<double> d(static_cast<std::size_t>(A.rows), 0.0);
  for (int r = 0; r < A.rows; ++r)
    for (int p = A.ptr[r]; p < A.ptr[r + 1]; ++p)
      if (A.col[p] == r) { d[r] = A.val[p] > 0.0 ? 1.0 / A.val[p] : 0.0; break; }
  return d;
}
// Exact coarsest solve (stands in for SimplicialLDLT, core/src/fea/multigrid.cpp:
// 690-694): envelope (skyline) Cholesky under a node ordering that puts the
// longest grid axis slowest, which keeps the envelope a few grid planes wide.
using Skyline = Level::Factor;
bool skyline_cholesky(const Csr& A, const std::vector<int>& perm, Skyline& S) {
  const int n = A.rows;
  S.n = n;
  S.perm = perm;
  std::vector<int> inv(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) inv[perm[i]] = i;
  S.first.assign(static_cast<std::size_t>(n), 0);
  for (int i = 0; i < n; ++i) {
    const int r = perm[i];
    int f = i;
    for (int p = A.ptr[r]; p < A.ptr[r + 1]; ++p) f = std::min(f, inv[A.col[p]]);
    S.first[i] = f;
  }
  S.off.assign(static_cast<std::size_t>(n) + 1, 0);
  for (int i = 0; i < n; ++i) S.off[i + 1] = S.off[i] + static_cast<std::size_t>(i - S.first[i] + 1);
  S.L.assign(S.off[n], 0.0);
  for (int i = 0; i < n; ++i) {
    const int r = perm[i];
    for (int p = A.ptr[r]; p < A.ptr[r + 1]; ++p) {
      const int j = inv[A.col[p]];
      if (j <= i) S.L[S.off[i] + (j - S.first[i])] = A.val[p];
    }
  }
  for (int i = 0; i < n; ++i) {
    double* Li = &S.L[S.off[i]];
    const int fi = S.first[i];
    for (int j = fi; j <= i; ++j) {
      const int fj = S.first[j];
      const double* Lj = &S.L[S.off[j]];
      const int k0 = std::max(fi, fj);
      double s = Li[j - fi];
      for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
      if (j < i) {
        Li[j - fi] = s / Lj[j - fj];
      } else {
        if (!(s > 0.0) || !std::isfinite(s)) return false;
        Li[i - fi] = std::sqrt(s);
      }
    }
  }
  return true;
}
void skyline_solve(const Skyline& S, const double* b, double* x) {
  const int n = S.n;
  std::vector<double> y(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) {
    const double* Li = &S.L[S.off[i]];
    const int fi = S.first[i];
    double s = b[S.perm[i]];
    for (int k = fi; k < i; ++k) s -= Li[k - fi] * y[k];
    y[i] = s / Li[i - fi];
  }
  for (int i = n - 1; i >= 0; --i) {
    const double* Li = &S.L[S.off[i]];
    const int fi = S.first[i];
    y[i] /= Li[i - fi];
    const double yi = y[i];
    for (int k = fi; k < i; ++k) y[k] -= Li[k - fi] * yi;
  }
  for (int i = 0; i < n; ++i) x[S.perm[i]] = y[i];
}
// Ordering of a level's active DOFs with the longest node axis slowest.
std::vector<int> coarse_ordering(int nx, int ny, int nz, const std::vector<int>& active, int n) {
  int dims[3] = {nx, ny, nz};
  int order[3] = {0, 1, 2};  // fastest .. slowest
  std::sort(order, order + 3, [&](int a, int b) { return dims[a] < dims[b]; });
  std::vector<int> perm;
  perm.reserve(static_cast<std::size_t>(n));
  int idx[3];
  for (idx[order[2]] = 0; idx[order[2]] < dims[order[2]]; ++idx[order[2]])
    for (idx[order[1]] = 0; idx[order[1]] < dims[order[1]]; ++idx[order[1]])
      for (idx[order[0]] = 0; idx[order[0]] < dims[order[0]]; ++idx[order[0]]) {
        const int node = (idx[2] * ny + idx[1]) * nx + idx[0];
        for (int c = 0; c < 3; ++c) {
          const int d = active[static_cast<std::size_t>(node) * 3 + c];
          if (d >= 0) perm.push_back(d);
        }
      }
  return perm;
}

constexpr double kJacobiOmega = 0.6;  // core/src/fea/multigrid.cpp:72
constexpr int kMinCoarseElems = 2;    // core/include/topopt/coarsen.hpp:76
constexpr int kCoarseDofCap = 6000;   // core/include/topopt/coarsen.hpp:80
constexpr int kMinLevels = 2;         // core/include/topopt/coarsen.hpp:82
constexpr int kMgIterBudget = 300;    // core/src/fea/multigrid.cpp:108

// axis_weights: core/src/fea/multigrid.cpp:494-506.
inline int axis_weights(int f, int (&ci)[2], double (&cw)[2]) {
  if ((f & 1) == 0) { ci[0] = f / 2; cw[0] = 1.0; return 1; }
  ci[0] = (f - 1) / 2; cw[0] = 0.5;
  ci[1] = (f + 1) / 2; cw[1] = 0.5;
  return 2;
}

// Trilinear P between an active fine node grid and the coincident coarse grid
// (core/src/fea/multigrid.cpp:586-654 and :1376-1442). Also returns the per-row
// (col, w) lists when `rows_out` is given.
Csr build_prolong(int fnx, int fny, int fnz, const std::vector<int>& factive, int fn,
                  int cnx, int cny, int cnz, const std::vector<int>& cactive, int nc,
                  std::vector<std::vector<std::pair<int, double>>>* rows_out) {
  std::vector<Trip> trips;
  trips.reserve(static_cast<std::size_t>(fn) * 8);
  if (rows_out) rows_out->assign(static_cast<std::size_t>(fn), {});
  for (int fc = 0; fc < fnz; ++fc)
    for (int fb = 0; fb < fny; ++fb)
      for (int fa = 0; fa < fnx; ++fa) {
        const int fnode = (fc * fny + fb) * fnx + fa;
        int ia[2], ib[2], ic[2];
        double wa[2], wb[2], wc[2];
        const int na = axis_weights(fa, ia, wa);
        const int nb = axis_weights(fb, ib, wb);
        const int ncz = axis_weights(fc, ic, wc);
        for (int comp = 0; comp < 3; ++comp) {
          const int row = factive[static_cast<std::size_t>(fnode) * 3 + comp];
          if (row < 0) continue;
          for (int x = 0; x < na; ++x)
            for (int y = 0; y < nb; ++y)
              for (int z = 0; z < ncz; ++z) {
                const int cnode = (ic[z] * cny + ib[y]) * cnx + ia[x];
                const int col = cactive[static_cast<std::size_t>(cnode) * 3 + comp];
                if (col < 0) continue;
                const double w = wa[x] * wb[y] * wc[z];
                trips.push_back({row, col, w});
                if (rows_out) (*rows_out)[static_cast<std::size_t>(row)].emplace_back(col, w);
              }
        }
      }
  return from_trips(fn, nc, trips);
}

// Coarse active map: active iff the coincident fine node-DOF is active.
int coarse_active(int fnx, int fny, const std::vector<int>& factive, int cnx, int cny,
                  int cnz, std::vector<int>& cactive) {
  cactive.assign(static_cast<std::size_t>(cnx) * cny * cnz * 3, -1);
  int nc = 0;
  for (int c = 0; c < cnz; ++c)
    for (int b = 0; b < cny; ++b)
      for (int a = 0; a < cnx; ++a) {
        const int fnode = ((2 * c) * fny + 2 * b) * fnx + 2 * a;
        const int cnode = (c * cny + b) * cnx + a;
        for (int comp = 0; comp < 3; ++comp)
          if (factive[static_cast<std::size_t>(fnode) * 3 + comp] >= 0)
            cactive[static_cast<std::size_t>(cnode) * 3 + comp] = nc++;
      }
  return nc;
}

// build_hierarchy: core/src/fea/multigrid.cpp:557-716 (scalar Jacobi path).
std::vector<Level> build_hierarchy(const Csr& A0, int nx0, int ny0, int nz0,
                                   std::vector<int> active0) {
  std::vector<Level> levels;
  Level fine;
  fine.nx = nx0; fine.ny = ny0; fine.nz = nz0;
  fine.n = A0.rows;
  fine.A = A0;
  fine.Dinv = inverse_diagonal(A0);
  fine.active = std::move(active0);
  levels.push_back(std::move(fine));
  while (true) {
    const Level& f = levels.back();
    const int fex = f.nx - 1, fey = f.ny - 1, fez = f.nz - 1;
    if ((fex & 1) || (fey & 1) || (fez & 1)) break;
    const int cex = fex / 2, cey = fey / 2, cez = fez / 2;
    if (cex < kMinCoarseElems || cey < kMinCoarseElems || cez < kMinCoarseElems) break;
    const int cnx = cex + 1, cny = cey + 1, cnz = cez + 1;
    std::vector<int> cact;
    const int nc = coarse_active(f.nx, f.ny, f.active, cnx, cny, cnz, cact);
    if (nc == 0) break;
    Csr P = build_prolong(f.nx, f.ny, f.nz, f.active, f.n, cnx, cny, cnz, cact, nc, nullptr);
    const Csr AP = multiply(f.A, P);
    Csr Ac = multiply(transpose(P), AP);
    levels.back().P = std::move(P);
    Level coarse;
    coarse.nx = cnx; coarse.ny = cny; coarse.nz = cnz;
    coarse.n = nc;
    coarse.A = std::move(Ac);
    coarse.Dinv = inverse_diagonal(coarse.A);
    coarse.active = std::move(cact);
    levels.push_back(std::move(coarse));
    if (nc <= kCoarseDofCap) break;
  }
  if (static_cast<int>(levels.size()) < kMinLevels) return {};
  if (levels.back().n > kCoarseDofCap) return {};
  Level& bottom = levels.back();
  bottom.coarsest = true;
  if (!skyline_cholesky(bottom.A, coarse_ordering(bottom.nx, bottom.ny, bottom.nz, bottom.active, bottom.n), bottom.chol)) return {};
  return levels;
}

// galerkin_a1_from_prolong (geometric stencil, colour block cache):
// core/src/fea/multigrid.cpp:1543-1868. A1 is accumulated into a fixed
// 81-slot row layout (27 neighbour coarse nodes x 3 comps) in element order,
// so each entry receives the same terms in the same order as the reference.
Csr galerkin_a1(const Reduced& m, const std::vector<int>& active,
                const std::vector<std::vector<std::pair<int, double>>>& prolong, int nc,
                int cnx, int cny, const std::vector<int>& cnode_of) {
  const int nelems = static_cast<int>(m.elems.size());
  std::vector<double> slots(static_cast<std::size_t>(nc) * 81, 0.0);
  std::vector<char> used(static_cast<std::size_t>(nc) * 81, 0);
  double W[kDof * kDof], KW[kDof * kDof], S[kDof * kDof];
  double cacheS[kNumColors][kDof * kDof];
  bool cache_valid[kNumColors] = {};
  int kg[kDof];
  int cds[kDof];
  int color = 0;
  for (int e = 0; e < nelems; ++e) {
    while (color + 1 < kNumColors && e >= m.color_offsets[color + 1]) ++color;
    const MfElem& el = m.elems[static_cast<std::size_t>(e)];
    int mloc = 0;
    for (int r = 0; r < kDof; ++r) {
      kg[r] = active[el.edof[r]];
      if (kg[r] < 0) continue;
      for (const auto& pr : prolong[static_cast<std::size_t>(kg[r])]) {
        bool found = false;
        for (int t = 0; t < mloc; ++t) if (cds[t] == pr.first) { found = true; break; }
        if (!found) cds[mloc++] = pr.first;
      }
    }
    if (mloc == 0) continue;
    bool generic = mloc == kDof;
    if (generic)
      for (int r = 0; r < kDof; ++r) if (kg[r] < 0) { generic = false; break; }
    if (generic && cache_valid[color]) {
      std::copy(cacheS[color], cacheS[color] + kDof * kDof, S);
    } else {
      for (int r = 0; r < kDof; ++r)
        for (int cl = 0; cl < mloc; ++cl) W[r * kDof + cl] = 0.0;
      for (int r = 0; r < kDof; ++r) {
        if (kg[r] < 0) continue;
        for (const auto& pr : prolong[static_cast<std::size_t>(kg[r])]) {
          int idx = 0;
          while (cds[idx] != pr.first) ++idx;
          W[r * kDof + idx] += pr.second;
        }
      }
      for (int r = 0; r < kDof; ++r)
        for (int cl = 0; cl < mloc; ++cl) {
          double s = 0.0;
          for (int c = 0; c < kDof; ++c) s += m.Ke[r * kDof + c] * W[c * kDof + cl];
          KW[r * kDof + cl] = s;
        }
      for (int cl = 0; cl < mloc; ++cl)
        for (int dl = 0; dl < mloc; ++dl) {
          double s = 0.0;
          for (int r = 0; r < kDof; ++r) s += W[r * kDof + cl] * KW[r * kDof + dl];
          S[cl * kDof + dl] = s;
        }
      if (generic) {
        std::copy(S, S + kDof * kDof, cacheS[color]);
        cache_valid[color] = true;
      }
    }
    for (int cl = 0; cl < mloc; ++cl) {
      const int i = cds[cl];
      const int ni = cnode_of[static_cast<std::size_t>(i)];
      const int ia = ni % cnx, ib = (ni / cnx) % cny, ic = ni / (cnx * cny);
      for (int dl = 0; dl < mloc; ++dl) {
        const double v = el.factor * S[cl * kDof + dl];
        if (v == 0.0) continue;
        const int j = cds[dl];
        const int nj = cnode_of[static_cast<std::size_t>(j)];
        const int ja = nj % cnx, jb = (nj / cnx) % cny, jc = nj / (cnx * cny);
        const int comp = cnode_of[static_cast<std::size_t>(nc) + j];
        const int s = (((jc - ic + 1) * 3 + (jb - ib + 1)) * 3 + (ja - ia + 1)) * 3 + comp;
        const std::size_t at = static_cast<std::size_t>(i) * 81 + s;
        slots[at] += v;
        used[at] = 1;
      }
    }
  }
  // Compress to CSR with ascending column (coarse DOF id) order.
  Csr A;
  A.rows = nc; A.cols = nc;
  A.ptr.assign(static_cast<std::size_t>(nc) + 1, 0);
  std::vector<std::pair<int, double>> cols;
  for (int i = 0; i < nc; ++i) {
    cols.clear();
    const int ni = cnode_of[static_cast<std::size_t>(i)];
    const int ia = ni % cnx, ib = (ni / cnx) % cny, ic = ni / (cnx * cny);
    for (int s = 0; s < 81; ++s) {
      const std::size_t at = static_cast<std::size_t>(i) * 81 + s;
      if (!used[at]) continue;
      const int comp = s % 3, nb = s / 3;
      const int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
      const int nj = ((ic + dz) * cny + (ib + dy)) * cnx + (ia + dx);
      const int j = cnode_of[static_cast<std::size_t>(nc) * 2 + static_cast<std::size_t>(nj) * 3 + comp];
      cols.emplace_back(j, slots[at]);
    }
    std::sort(cols.begin(), cols.end());
    for (auto& cv : cols) { A.col.push_back(cv.first); A.val.push_back(cv.second); }
    A.ptr[i + 1] = static_cast<int>(A.col.size());
  }
  return A;
}

int g_smoother = 0, g_cheb_degree = 1;

// lambda_max(D^-1 A) by kChebPowerSteps power steps from cheb_start (the
// level DOF of index i sits at node-major position gid[i]): y = A v, then
// v = D^-1 y / ||D^-1 y|| for all but the last step, whose Rayleigh quotient
// (v.Av) / (v.Dv) is the estimate.
template <class Apply>
double power_lambda(int n, const std::vector<long long>& gid, const std::vector<double>& dinv,
                    Apply apply) {
  std::vector<double> v(static_cast<std::size_t>(n)), y(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) v[i] = cheb_start(gid[i]);
  for (int k = 0;; ++k) {
    apply(v.data(), y.data());
    if (k + 1 == kChebPowerSteps) break;
    for (int i = 0; i < n; ++i) v[i] = dinv[i] * y[i];
    const double nv = norm2(v);
    for (int i = 0; i < n; ++i) v[i] = v[i] / nv;
  }
  double vav = 0.0, vdv = 0.0;
  for (int i = 0; i < n; ++i) {
    vav += v[i] * y[i];
    vdv += v[i] * v[i] / dinv[i];
  }
  return vav / vdv;
}

ChebLevel cheb_level(double lambda) {
  ChebLevel c;
  c.lambda = lambda;
  const double hi = kChebHigh * lambda, lo = kChebLow * lambda;
  c.theta = 0.5 * (hi + lo);
  c.delta = 0.5 * (hi - lo);
  return c;
}

// Chebyshev smoothing of A x = b on one level (degree operator applies when
// starting from x, degree - 1 from zero): d = D^-1 r / theta, x += d, then
// rho' = 1 / (2 sigma - rho), d = rho' rho d + (2 rho' / delta) D^-1 r,
// x += d, with sigma = theta / delta, rho = 1 / sigma at the start.
template <class Apply>
void cheb_smooth(int n, const std::vector<double>& dinv, const ChebLevel& c, int degree,
                 const std::vector<double>& b, std::vector<double>& x, bool zero_start,
                 Apply apply) {
  std::vector<double> r(static_cast<std::size_t>(n)), d(static_cast<std::size_t>(n));
  if (zero_start) {
    x.assign(static_cast<std::size_t>(n), 0.0);
    r = b;
  } else {
    apply(x.data(), r.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  }
  for (int i = 0; i < n; ++i) {
    d[i] = dinv[i] * r[i] / c.theta;
    x[i] += d[i];
  }
  const double sigma = c.theta / c.delta;
  double rho = 1.0 / sigma;
  for (int k = 1; k < degree; ++k) {
    apply(x.data(), r.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    const double c1 = rho_new * rho, c2 = 2.0 * rho_new / c.delta;
    for (int i = 0; i < n; ++i) {
      d[i] = c1 * d[i] + c2 * (dinv[i] * r[i]);
      x[i] += d[i];
    }
    rho = rho_new;
  }
}

void jacobi_sweep(const Level& L, const std::vector<double>& b, std::vector<double>& x) {
  std::vector<double> Ax(static_cast<std::size_t>(L.n));
  L.A.spmv(x.data(), Ax.data());
  for (int i = 0; i < L.n; ++i) x[i] += kJacobiOmega * (L.Dinv[i] * (b[i] - Ax[i]));
}

// v_cycle: core/src/fea/multigrid.cpp:798-819 (gamma 1, 1+1 sweeps).
std::vector<double> v_cycle(const std::vector<Level>& levels, int l,
                            const std::vector<double>& b, const std::vector<ChebLevel>& cheb) {
  const Level& L = levels[static_cast<std::size_t>(l)];
  std::vector<double> x(static_cast<std::size_t>(L.n), 0.0);
  if (L.coarsest) {
    skyline_solve(L.chol, b.data(), x.data());
    return x;
  }
  auto spmv = [&](const double* in, double* out) { L.A.spmv(in, out); };
  if (g_smoother == 1)
    cheb_smooth(L.n, L.Dinv, cheb[static_cast<std::size_t>(l) + 1], g_cheb_degree, b, x, true, spmv);
  else
    jacobi_sweep(L, b, x);
  std::vector<double> r(static_cast<std::size_t>(L.n));
  L.A.spmv(x.data(), r.data());
  for (int i = 0; i < L.n; ++i) r[i] = b[i] - r[i];
  std::vector<double> bc(static_cast<std::size_t>(L.P.cols));
  L.P.spmv_t(r.data(), bc.data());
  const std::vector<double> ec = v_cycle(levels, l + 1, bc, cheb);
  std::vector<double> pe(static_cast<std::size_t>(L.n));
  L.P.spmv(ec.data(), pe.data());
  for (int i = 0; i < L.n; ++i) x[i] += pe[i];
  if (g_smoother == 1)
    cheb_smooth(L.n, L.Dinv, cheb[static_cast<std::size_t>(l) + 1], g_cheb_degree, b, x, false, spmv);
  else
    jacobi_sweep(L, b, x);
  return x;
}
}  // namespace

bool build_mf_hierarchy(const Reduced& m, int nnx, int nny, int nnz, Hierarchy& H) {
  const int fex = nnx - 1, fey = nny - 1, fez = nnz - 1;
  if ((fex & 1) || (fey & 1) || (fez & 1)) return false;
  const int cex = fex / 2, cey = fey / 2, cez = fez / 2;
  if (cex < kMinCoarseElems || cey < kMinCoarseElems || cez < kMinCoarseElems) return false;
  const int cnx = cex + 1, cny = cey + 1, cnz = cez + 1;
  std::vector<int> active(static_cast<std::size_t>(m.ndof), -1);
  for (int k = 0; k < m.ng; ++k) active[static_cast<std::size_t>(m.kept_global[k])] = k;
  std::vector<int> cact;
  const int nc = coarse_active(nnx, nny, active, cnx, cny, cnz, cact);
  if (nc == 0) return false;
  std::vector<std::vector<std::pair<int, double>>> prolong;
  H.P0 = build_prolong(nnx, nny, nnz, active, m.ng, cnx, cny, cnz, cact, nc, &prolong);
  // cnode_of: [0,nc) coarse dof -> node; [nc,2nc) -> comp; [2nc, 2nc+3*nodes) node*3+comp -> dof.
  std::vector<int> cnode_of(static_cast<std::size_t>(nc) * 2 + cact.size(), -1);
  for (std::size_t q = 0; q < cact.size(); ++q) {
    cnode_of[static_cast<std::size_t>(nc) * 2 + q] = cact[q];
    if (cact[q] >= 0) {
      cnode_of[static_cast<std::size_t>(cact[q])] = static_cast<int>(q / 3);
      cnode_of[static_cast<std::size_t>(nc) + cact[q]] = static_cast<int>(q % 3);
    }
  }
  Csr A1 = galerkin_a1(m, active, prolong, nc, cnx, cny, cnode_of);
  std::vector<Level> coarse = build_hierarchy(A1, cnx, cny, cnz, cact);
  if (coarse.empty()) {
    if (A1.rows > kCoarseDofCap) return false;
    Level only;
    only.nx = cnx; only.ny = cny; only.nz = cnz;
    only.n = A1.rows;
    only.A = A1;
    only.Dinv = inverse_diagonal(A1);
    only.active = cact;
    only.coarsest = true;
    if (!skyline_cholesky(only.A, coarse_ordering(only.nx, only.ny, only.nz, only.active, only.n), only.chol)) return false;
    coarse.push_back(std::move(only));
  }
  H.m = &m;
  H.coarse = std::move(coarse);
  H.cheb.clear();
  if (g_smoother == 1) {
    // level 0: the matrix-free fine operator over the kept DOFs
    std::vector<long long> gid(m.kept_global.begin(), m.kept_global.end());
    H.cheb.push_back(cheb_level(power_lambda(
        m.ng, gid, m.invdiag, [&](const double* in, double* out) { m.apply_kgg(in, out); })));
    for (const Level& L : H.coarse) {
      if (L.coarsest) break;
      std::vector<long long> lg(static_cast<std::size_t>(L.n));
      for (std::size_t q = 0; q < L.active.size(); ++q)
        if (L.active[q] >= 0) lg[static_cast<std::size_t>(L.active[q])] = static_cast<long long>(q);
      H.cheb.push_back(cheb_level(power_lambda(
          L.n, lg, L.Dinv, [&](const double* in, double* out) { L.A.spmv(in, out); })));
    }
  }
  return true;
}

void set_smoother(int kind, int degree) {
  if (kind < 0 || kind > 1 || degree < 1) throw std::invalid_argument("set_smoother: bad arguments");
  g_smoother = kind;
  g_cheb_degree = degree;
}

void mf_v_cycle(const Hierarchy& H, const std::vector<double>& b, std::vector<double>& x) {
  const Reduced& m = *H.m;
  const int ng = m.ng;
  std::vector<double> Ax(static_cast<std::size_t>(ng));
  x.assign(static_cast<std::size_t>(ng), 0.0);
  auto sweep = [&]() {
    m.apply_kgg(x.data(), Ax.data());
    for (int i = 0; i < ng; ++i) x[i] += kJacobiOmega * (m.invdiag[i] * (b[i] - Ax[i]));
  };
  auto kgg = [&](const double* in, double* out) { m.apply_kgg(in, out); };
  if (g_smoother == 1)
    cheb_smooth(ng, m.invdiag, H.cheb[0], g_cheb_degree, b, x, true, kgg);
  else
    sweep();
  m.apply_kgg(x.data(), Ax.data());
  std::vector<double> vr(static_cast<std::size_t>(ng));
  for (int i = 0; i < ng; ++i) vr[i] = b[i] - Ax[i];
  std::vector<double> bc(static_cast<std::size_t>(H.P0.cols));
  H.P0.spmv_t(vr.data(), bc.data());
  const std::vector<double> ec = v_cycle(H.coarse, 0, bc, H.cheb);
  std::vector<double> pe(static_cast<std::size_t>(ng));
  H.P0.spmv(ec.data(), pe.data());
  for (int i = 0; i < ng; ++i) x[i] += pe[i];
  if (g_smoother == 1)
    cheb_smooth(ng, m.invdiag, H.cheb[0], g_cheb_degree, b, x, false, kgg);
  else
    sweep();
}

bool mf_mgpcg(const Hierarchy& H, const std::vector<double>& b, double tol, int max_it,
              std::vector<double>& x, int& iters, double& resid) {
  const int n = H.m->ng;
  const double bnorm = norm2(b);
  iters = 0;
  resid = 0.0;
  if (!(bnorm > 0.0)) { x.assign(static_cast<std::size_t>(n), 0.0); return true; }
  const double threshold = tol * bnorm;
  std::vector<double> r(static_cast<std::size_t>(n)), Ap(static_cast<std::size_t>(n));
  const bool warm = static_cast<int>(x.size()) == n;
  if (warm) {
    H.m->apply_kgg(x.data(), Ap.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  } else {
    x.assign(static_cast<std::size_t>(n), 0.0);
    r = b;
  }
  std::vector<double> z;
  mf_v_cycle(H, r, z);
  std::vector<double> p = z;
  double rz = dot(r, z);
  if (!std::isfinite(rz)) return false;
  for (int k = 1; k <= max_it; ++k) {
    H.m->apply_kgg(p.data(), Ap.data());
    const double pAp = dot(p, Ap);
    if (!(pAp > 0.0) || !std::isfinite(pAp)) return false;
    const double alpha = rz / pAp;
    for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
    iters = k;
    const double rn = norm2(r);
    resid = rn / bnorm;
    if (rn <= threshold) {
      for (double v : x) if (!std::isfinite(v)) return false;
      return true;
    }
    mf_v_cycle(H, r, z);
    const double rznew = dot(r, z);
    if (!std::isfinite(rznew)) return false;
    const double beta = rznew / rz;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rznew;
  }
  return false;
}

std::vector<double> solve_mgcg_matfree(const Grid& g, const std::vector<double>& youngs,
                                       double nu, const std::vector<DirichletBC>& bcs,
                                       const std::vector<NodalLoad>& loads, double tol,
                                       int max_it, CgInfo* info,
                                       const std::vector<double>* guess) {
  const double t0 = now_ms();
  Reduced m = build_reduced(g, youngs, nu, bcs, loads, info);
  CgInfo diag;
  diag.converged = true;
  diag.t_build_ms = now_ms() - t0;
  std::vector<double> u = m.up;
  if (m.ng > 0) {
    const int cap = max_it > 0 ? max_it : std::max(1000, 2 * m.ng);
    Hierarchy H;
    const double th = now_ms();
    const bool have_h = build_mf_hierarchy(m, g.nnx(), g.nny(), g.nnz(), H);
    diag.t_mg_build_ms = now_ms() - th;
    diag.hier_built = have_h;
    std::vector<double> xkept(static_cast<std::size_t>(m.ng), 0.0);
    bool warm = false;
    if (guess && guess->size() == u.size()) {
      warm = true;
      for (int k = 0; k < m.ng; ++k) {
        const double v = (*guess)[static_cast<std::size_t>(m.kept_global[k])];
        if (!std::isfinite(v)) { warm = false; break; }
        xkept[k] = v;
      }
      if (!warm) std::fill(xkept.begin(), xkept.end(), 0.0);
    }
    bool solved = false;
    const double tm = now_ms();
    if (have_h) {
      std::vector<double> xg;
      if (warm) xg = xkept;
      int it = 0;
      double res = 0.0;
      const bool ok = mf_mgpcg(H, m.rg, tol, std::min(cap, kMgIterBudget), xg, it, res);
      diag.mg_cycles_attempted = it;
      if (ok) {
        xkept = xg;
        diag.iterations = it;
        diag.residual = res;
        diag.converged = true;
        diag.used_multigrid = true;
        diag.mg_levels = H.levels();
        solved = true;
      }
    }
    diag.t_mg_ms = have_h ? now_ms() - tm : 0.0;
    if (!solved) {
      const double tc = now_ms();
      cg_solve(m, tol, cap, xkept, diag.iterations, diag.residual, diag.converged);
      diag.used_multigrid = false;
      diag.mg_levels = 0;
      diag.t_cg_ms = now_ms() - tc;
      diag.t_total_ms = now_ms() - t0;
      diag.matvecs = m.matvecs;
      if (info) *info = diag;
      if (!diag.converged)
        throw SolverNonConvergence(
            "fea_solve_mgcg_matfree: CG did not reach the requested tolerance within "
            "max_iterations",
            diag.iterations, diag.residual);
    }
    for (int k = 0; k < m.ng; ++k) u[static_cast<std::size_t>(m.kept_global[k])] = xkept[k];
  }
  diag.t_total_ms = now_ms() - t0;
  diag.matvecs = m.matvecs;
  if (info) *info = diag;
  return u;
}

std::vector<double> solve_cg_matfree(const Grid& g, const std::vector<double>& youngs,
                                     double nu, const std::vector<DirichletBC>& bcs,
                                     const std::vector<NodalLoad>& loads, double tol,
                                     int max_it, CgInfo* info) {
  Reduced m = build_reduced(g, youngs, nu, bcs, loads, info);
  CgInfo diag;
  diag.converged = true;
  std::vector<double> u = m.up;
  if (m.ng > 0) {
    std::vector<double> x(static_cast<std::size_t>(m.ng), 0.0);
    cg_solve(m, tol, max_it, x, diag.iterations, diag.residual, diag.converged);
    if (info) *info = diag;
    if (!diag.converged)
      throw SolverNonConvergence("fea_solve_cg_matfree: CG did not converge", diag.iterations,
                                 diag.residual);
    for (int k = 0; k < m.ng; ++k) u[static_cast<std::size_t>(m.kept_global[k])] = x[k];
  }
  diag.matvecs = m.matvecs;
  if (info) *info = diag;
  return u;
}

// ---------------------------------------------------------------------------
double simp_youngs(const SimpParams& p, double rho) {
  const double r = std::min(1.0, std::max(p.density_min, rho));
  return std::pow(r, p.penalty) * p.E0;
}

SimpCompliance simp_compliance(const Grid& g, const SimpParams& p,
                               const std::vector<double>& density,
                               const std::vector<DirichletBC>& bcs,
                               const std::vector<NodalLoad>& loads, double tol, int max_it,
                               bool use_mg, const std::vector<double>* initial_guess) {
  if (density.size() != g.voxels())
    throw std::invalid_argument("simp_compliance: density vector size != voxel_count");
  std::vector<double> E(g.voxels(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        const std::size_t e = g.index(i, j, k);
        const double rho = std::min(1.0, std::max(p.density_min, density[e]));
        E[e] = std::pow(rho, p.penalty) * p.E0;
      }
  SimpCompliance out;
  out.u = use_mg ? solve_mgcg_matfree(g, E, p.nu, bcs, loads, tol, max_it, &out.cg, initial_guess)
                 : solve_cg_matfree(g, E, p.nu, bcs, loads, tol, max_it, &out.cg);
  double Ku[kDof * kDof];
  hex8_stiffness(1.0, p.nu, g.h, Ku);
  out.dc.assign(g.voxels(), 0.0);
  double c = 0.0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        const std::size_t e = g.index(i, j, k);
        const auto en = element_nodes(g, i, j, k);
        double ue[kDof];
        for (int a = 0; a < 8; ++a)
          for (int comp = 0; comp < 3; ++comp)
            ue[3 * a + comp] = out.u[3 * static_cast<std::size_t>(en[a]) + comp];
        const double rho = std::min(1.0, std::max(p.density_min, density[e]));
        double q = 0.0;
        for (int r = 0; r < kDof; ++r) {
          double kr = 0.0;
          for (int cc = 0; cc < kDof; ++cc) kr += Ku[r * kDof + cc] * ue[cc];
          q += ue[r] * kr;
        }
        c += std::pow(rho, p.penalty) * p.E0 * q;
        out.dc[e] = -p.penalty * std::pow(rho, p.penalty - 1.0) * p.E0 * q;
      }
  out.compliance = c;
  return out;
}

DensityFilter make_density_filter(const Grid& g, double radius,
                                  const std::vector<std::uint8_t>& mask) {
  if (!(radius > 0.0)) throw std::invalid_argument("make_density_filter: radius must be > 0");
  if (mask.size() != g.voxels())
    throw std::invalid_argument("make_density_filter: mask size != voxel_count");
  auto is_active = [&](int i, int j, int k) {
    return g.is_solid(i, j, k) && mask[g.index(i, j, k)] == kActive;
  };
  DensityFilter f;
  f.voxel_count = g.voxels();
  f.row_start.assign(f.voxel_count + 1, 0);
  f.weight_sum.assign(f.voxel_count, 0.0);
  const int R = static_cast<int>(std::ceil(radius));
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const std::size_t e = g.index(i, j, k);
        f.row_start[e] = f.neighbor.size();
        if (!is_active(i, j, k)) continue;
        double hs = 0.0;
        for (int dk = -R; dk <= R; ++dk)
          for (int dj = -R; dj <= R; ++dj)
            for (int di = -R; di <= R; ++di) {
              const int ii = i + di, jj = j + dj, kk = k + dk;
              if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
              if (!is_active(ii, jj, kk)) continue;
              const double dist = std::sqrt(static_cast<double>(di * di + dj * dj + dk * dk));
              const double hw = radius - dist;
              if (hw <= 0.0) continue;
              f.neighbor.push_back(g.index(ii, jj, kk));
              f.weight.push_back(hw);
              hs += hw;
            }
        f.weight_sum[e] = hs;
      }
  f.row_start[f.voxel_count] = f.neighbor.size();
  return f;
}

// The plain filter is the all-Active mask (simp.cpp:892-896).
DensityFilter make_density_filter(const Grid& g, double radius) {
  return make_density_filter(g, radius, std::vector<std::uint8_t>(g.voxels(), kActive));
}

std::vector<double> DensityFilter::filter_density(const std::vector<double>& x) const {
  std::vector<double> out(voxel_count, 0.0);
  for (std::size_t e = 0; e < voxel_count; ++e) {
    if (!(weight_sum[e] > 0.0)) continue;
    double acc = 0.0;
    for (std::size_t p = row_start[e]; p < row_start[e + 1]; ++p) acc += weight[p] * x[neighbor[p]];
    out[e] = acc / weight_sum[e];
  }
  return out;
}

std::vector<double> DensityFilter::filter_sensitivity(const std::vector<double>& gsens) const {
  std::vector<double> out(voxel_count, 0.0);
  for (std::size_t e = 0; e < voxel_count; ++e) {
    if (!(weight_sum[e] > 0.0)) continue;
    const double se = gsens[e] / weight_sum[e];
    for (std::size_t p = row_start[e]; p < row_start[e + 1]; ++p) out[neighbor[p]] += weight[p] * se;
  }
  return out;
}

std::vector<double> oc_update(const Grid& g, const DensityFilter& f,
                              const std::vector<double>& x, const std::vector<double>& dcomp,
                              double volfrac, double move, double density_min, int* nbis,
                              double* lambda_out) {
  const std::size_t N = g.voxels();
  if (x.size() != N || dcomp.size() != N)
    throw std::invalid_argument("oc_update: vector size != voxel_count");
  if (!(volfrac > 0.0 && volfrac <= 1.0))
    throw std::invalid_argument("oc_update: volume_fraction must be in (0, 1]");
  if (!(move > 0.0)) throw std::invalid_argument("oc_update: move must be > 0");
  const std::vector<double> dc = f.filter_sensitivity(dcomp);
  std::vector<double> ones(N, 0.0);
  std::size_t n_design = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (g.is_solid(i, j, k)) { ones[g.index(i, j, k)] = 1.0; ++n_design; }
  const std::vector<double> dv = f.filter_sensitivity(ones);
  const double target = volfrac * static_cast<double>(n_design);
  auto candidate = [&](double lambda) {
    std::vector<double> xn(N, 0.0);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (!g.is_solid(i, j, k)) continue;
          const std::size_t e = g.index(i, j, k);
          const double xe = x[e];
          double be = 0.0;
          if (dv[e] > 0.0) be = -dc[e] / (lambda * dv[e]);
          if (be < 0.0) be = 0.0;
          double xt = xe * std::sqrt(be);
          const double lo = std::max(density_min, xe - move);
          const double hi = std::min(1.0, xe + move);
          if (xt < lo) xt = lo;
          if (xt > hi) xt = hi;
          xn[e] = xt;
        }
    return xn;
  };
  auto phys_vol = [&](const std::vector<double>& xn) {
    const std::vector<double> xp = f.filter_density(xn);
    double v = 0.0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
          if (g.is_solid(i, j, k)) v += xp[g.index(i, j, k)];
    return v;
  };
  double l1 = 0.0, l2 = 1.0;
  while (l2 < 1e30 && phys_vol(candidate(l2)) > target) l2 *= 2.0;
  int it = 0;
  for (; it < 100 && (l2 - l1) > 1e-6 * (l1 + l2); ++it) {
    const double lmid = 0.5 * (l1 + l2);
    if (phys_vol(candidate(lmid)) > target) l1 = lmid; else l2 = lmid;
  }
  if (nbis) *nbis = it;
  if (lambda_out) *lambda_out = 0.5 * (l1 + l2);
  return candidate(0.5 * (l1 + l2));
}

double penalty_at_iteration(const std::vector<PenaltyStage>& s, int iteration) {
  if (s.empty()) throw std::invalid_argument("penalty_at_iteration: schedule must be non-empty");
  if (iteration < 1) throw std::invalid_argument("penalty_at_iteration: iteration must be >= 1");
  long long remaining = iteration;
  for (const auto& st : s) {
    remaining -= st.iterations;
    if (remaining <= 0) return st.penalty;
  }
  return s.back().penalty;
}

std::vector<PenaltyStage> penalty_continuation_ramp(double p0, double p1, double step,
                                                    int per_stage) {
  if (!(std::isfinite(p0) && p0 > 0.0) || !(std::isfinite(p1) && p1 >= p0) ||
      !(std::isfinite(step) && step > 0.0) || per_stage < 1)
    throw std::invalid_argument("penalty_continuation_ramp: bad arguments");
  const long long n = static_cast<long long>((p1 - p0) / step + 1e-9);
  std::vector<PenaltyStage> out;
  for (long long k = 0; k <= n; ++k) out.push_back({p0 + static_cast<double>(k) * step, per_stage});
  return out;
}

std::vector<ProjStage> heaviside_continuation_schedule() {
  return {{1.0, 0.2, 50}, {2.0, 0.2, 50}, {4.0, 0.2, 50},
          {8.0, 0.2, 50}, {16.0, 0.1, 50}, {32.0, 0.05, 50}};
}

namespace {
// project_solid (simp.cpp:1637-1645): design voxels only.
void project_solid(const Grid& g, std::vector<double>& x, double beta, double eta) {
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (g.is_solid(i, j, k)) {
          const std::size_t e = g.index(i, j, k);
          x[e] = heaviside_project(x[e], beta, eta);
        }
}
}  // namespace

std::vector<double> oc_update_projected(const Grid& g, const DensityFilter& f,
                                        const std::vector<double>& x,
                                        const std::vector<double>& xtilde,
                                        const std::vector<double>& dcomp, double beta,
                                        double eta, double volfrac, double move,
                                        double density_min, int* nbis, double* lambda_out) {
  const std::size_t N = g.voxels();
  std::vector<double> dc_phys(N, 0.0), dproj(N, 0.0);
  std::size_t n_design = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        const std::size_t e = g.index(i, j, k);
        dproj[e] = heaviside_project_derivative(xtilde[e], beta, eta);
        dc_phys[e] = dcomp[e] * dproj[e];
        ++n_design;
      }
  const std::vector<double> dc = f.filter_sensitivity(dc_phys);
  const std::vector<double> dv = f.filter_sensitivity(dproj);
  const double target = volfrac * static_cast<double>(n_design);
  auto candidate = [&](double lambda) {
    std::vector<double> xn(N, 0.0);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (!g.is_solid(i, j, k)) continue;
          const std::size_t e = g.index(i, j, k);
          const double xe = x[e];
          double be = -dc[e] / (lambda * dv[e]);
          if (!std::isfinite(be)) be = 1.0;
          if (be < 0.0) be = 0.0;
          double xt = xe * std::sqrt(be);
          const double lo = std::max(density_min, xe - move);
          const double hi = std::min(1.0, xe + move);
          if (xt < lo) xt = lo;
          if (xt > hi) xt = hi;
          xn[e] = xt;
        }
    return xn;
  };
  auto proj_vol = [&](const std::vector<double>& xn) {
    const std::vector<double> xp = f.filter_density(xn);
    double v = 0.0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
          if (g.is_solid(i, j, k)) v += heaviside_project(xp[g.index(i, j, k)], beta, eta);
    return v;
  };
  double l1 = 0.0, l2 = 1.0;
  while (l2 < 1e30 && proj_vol(candidate(l2)) > target) l2 *= 2.0;
  int it = 0;
  for (; it < 100 && (l2 - l1) > 1e-6 * (l1 + l2); ++it) {
    const double lmid = 0.5 * (l1 + l2);
    if (proj_vol(candidate(lmid)) > target) l1 = lmid; else l2 = lmid;
  }
  if (nbis) *nbis = it;
  if (lambda_out) *lambda_out = 0.5 * (l1 + l2);
  return candidate(0.5 * (l1 + l2));
}

// validate_mma_projection_options (simp.cpp:112-135).
void validate_mma_projection(const SimpOptions& o) {
  // validate_adaptive_move_options (simp.cpp:140-163)
  if (o.adaptive_move) {
    if (!o.mma) throw std::invalid_argument("simp_optimize: adaptive_move requires the MMA updater");
    if (!(o.adaptive_move_grow > 1.0) || !std::isfinite(o.adaptive_move_grow))
      throw std::invalid_argument("simp_optimize: adaptive_move_grow must be finite and > 1");
    if (!(o.adaptive_move_shrink > 0.0 && o.adaptive_move_shrink < 1.0))
      throw std::invalid_argument("simp_optimize: adaptive_move_shrink must be in (0, 1)");
    if (!(o.adaptive_move_osc_lo >= 0.0 && o.adaptive_move_osc_lo <= o.adaptive_move_osc_hi &&
          o.adaptive_move_osc_hi <= 1.0))
      throw std::invalid_argument(
          "simp_optimize: adaptive_move_osc thresholds must satisfy 0 <= lo <= hi <= 1");
    if (!(o.adaptive_move_min > 0.0) || !(o.adaptive_move_max >= o.adaptive_move_min) ||
        !(o.adaptive_move_max < 1.0))
      throw std::invalid_argument(
          "simp_optimize: adaptive_move bounds must satisfy 0 < min <= max < 1");
  }
  if (!o.mma_projection) return;
  if (!o.mma) throw std::invalid_argument("simp_optimize: mma_projection requires the MMA updater");
  if (!o.projection.empty())
    throw std::invalid_argument("simp_optimize: mma_projection and a Heaviside projection "
                                "schedule are mutually exclusive");
  if (!(o.projection_eta > 0.0 && o.projection_eta < 1.0))
    throw std::invalid_argument("simp_optimize: projection_eta must be in (0, 1)");
  if (!(o.mma_projection_beta0 > 0.0) || !std::isfinite(o.mma_projection_beta0))
    throw std::invalid_argument("simp_optimize: mma_projection_beta0 must be finite and > 0");
  if (!(o.mma_projection_beta_max >= o.mma_projection_beta0) ||
      !std::isfinite(o.mma_projection_beta_max))
    throw std::invalid_argument(
        "simp_optimize: mma_projection_beta_max must be finite and >= mma_projection_beta0");
}

SimpResult simp_optimize(const Grid& g, const SimpParams& p,
                         const std::vector<DirichletBC>& bcs,
                         const std::vector<NodalLoad>& loads, const SimpOptions& o) {
  if (!(o.volume_fraction > 0.0 && o.volume_fraction <= 1.0))
    throw std::invalid_argument("simp_optimize: volume_fraction must be in (0, 1]");
  if (!(o.move > 0.0)) throw std::invalid_argument("simp_optimize: move must be > 0");
  if (o.max_iterations < 1) throw std::invalid_argument("simp_optimize: max_iterations must be >= 1");
  const DensityFilter f = make_density_filter(g, o.filter_radius);
  const double n_design = static_cast<double>(g.solid_count());
  auto volfrac = [&](const std::vector<double>& xp) {
    double v = 0.0;
    for (double q : xp) v += q;
    return n_design > 0.0 ? v / n_design : 0.0;
  };
  // Stage plan (build_stage_plan, simp.cpp:1053-1068).
  // (cont: the single dynamic stage of MMA with projection, :1036-1068)
  struct Stage { bool project; double beta, move; int iterations; bool cont; };
  std::vector<Stage> plan;
  validate_mma_projection(o);
  if (o.mma_projection) plan.push_back({false, 0.0, o.move, o.max_iterations, true});
  else if (o.projection.empty()) plan.push_back({false, 0.0, o.move, o.max_iterations, false});
  else
    for (const ProjStage& ps : o.projection)
      plan.push_back({true, ps.beta, ps.move, ps.iterations, false});
  double mma_beta = o.mma_projection_beta0;
  std::size_t stage_start = 0;  // history index where the current beta stage began
  double adaptive_move_val = 0.0;
  std::vector<std::size_t> design_dof;  // the adaptive move's voxels: the solid ones
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (g.is_solid(i, j, k)) design_dof.push_back(g.index(i, j, k));
  const double eta = o.projection_eta;
  std::vector<double> x(g.voxels(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (g.is_solid(i, j, k)) x[g.index(i, j, k)] = o.volume_fraction;
  SimpResult res;
  if (o.mma && !o.projection.empty())
    throw std::invalid_argument("simp_optimize: projection requires the OC updater");
  MmaState mma_state;
  int mma_iter = 0;
  std::vector<double> curve;  // compliance history for the MMA plateau stop
  for (const Stage& st : plan) {
    bool stage_converged = false;
    for (int it = 0; it < st.iterations; ++it) {
      const double cur_beta = st.cont ? mma_beta : st.beta;
      double cur_move = st.cont ? mma_continuation_move(st.move, cur_beta) : st.move;
      if (o.adaptive_move) {
        // seeded at the reference move for the first two MMA steps (and after a
        // beta advance restarts the state), then evolved (:2179-2194)
        if (mma_iter < 2 || mma_state.xold2.empty()) adaptive_move_val = cur_move;
        else
          adaptive_move_val = adapt_move_value(
              adaptive_move_val,
              oscillation_fraction(x, mma_state.xold1, mma_state.xold2, design_dof), o);
        cur_move = adaptive_move_val;
      }
      std::vector<double> xphys = f.filter_density(x);
      if (st.project || st.cont) project_solid(g, xphys, cur_beta, eta);
      SimpParams tp = p;
      if (!o.penalty_continuation.empty())
        tp.penalty = penalty_at_iteration(o.penalty_continuation, res.iterations + 1);
      SimpCompliance c;
      try {
        c = simp_compliance(g, tp, xphys, bcs, loads, o.cg_tolerance, o.cg_max_iterations,
                            o.use_mg);
      } catch (const SolverNonConvergence&) {
        res.non_convergent = true;
        break;
      }
      std::vector<double> xn;
      if (st.cont) {
        // mma_update_projected at the current continuation beta (:2283-2293)
        const std::vector<double> xtilde = f.filter_density(x);
        xn = mma_update(mma_state, ++mma_iter, g, f, x, c.dc, o.volume_fraction, cur_move,
                        p.density_min, nullptr, nullptr, &xtilde, cur_beta, eta);
      } else if (st.project) {
        const std::vector<double> xtilde = f.filter_density(x);
        xn = oc_update_projected(g, f, x, xtilde, c.dc, st.beta, eta, o.volume_fraction, st.move,
                                 p.density_min);
      } else if (o.mma) {
        xn = mma_update(mma_state, ++mma_iter, g, f, x, c.dc, o.volume_fraction, cur_move,
                        p.density_min);
      } else {
        xn = oc_update(g, f, x, c.dc, o.volume_fraction, st.move, p.density_min);
      }
      double change = 0.0;
      for (std::size_t e = 0; e < x.size(); ++e) change = std::max(change, std::fabs(xn[e] - x[e]));
      x = xn;
      if (res.iterations == 0) res.initial_compliance = c.compliance;
      ++res.iterations;
      std::vector<double> xa = f.filter_density(x);
      if (st.project) project_solid(g, xa, st.beta, eta);
      res.history.push_back({c.compliance, change, volfrac(xa), c.cg.iterations});
      curve.push_back(c.compliance);
      // stage_should_stop, core/src/simp/simp.cpp:1219-1232
      if (st.cont) {
        // beta continuation (:2461-2491): a plateau of the current beta stage
        // doubles beta (capped) and restarts the MMA state; a plateau at the
        // cap ends the run. The progress gate guards the first stage only.
        const std::vector<double> sc(curve.begin() + static_cast<std::ptrdiff_t>(stage_start),
                                     curve.end());
        const double min_drop = stage_start == 0 ? o.mma_plateau_min_drop : 0.0;
        if (mma_objective_plateau(sc, o.mma_plateau_window, o.mma_plateau_tol, min_drop,
                                  o.mma_plateau_flat_windows)) {
          if (mma_beta < o.mma_projection_beta_max) {
            mma_beta = std::min(mma_beta * 2.0, o.mma_projection_beta_max);
            stage_start = curve.size();
            mma_state = MmaState{};
            mma_iter = 0;
          } else {
            stage_converged = true;
            break;
          }
        }
        continue;
      }
      const bool stop = (o.mma && o.mma_plateau_window > 0)
                            ? mma_objective_plateau(curve, o.mma_plateau_window, o.mma_plateau_tol,
                                                    o.mma_plateau_min_drop,
                                                    o.mma_plateau_flat_windows)
                            : change < o.change_tol;
      if (stop) { stage_converged = true; break; }
    }
    res.converged = stage_converged;
    if (res.non_convergent) break;
  }
  res.design = x;
  res.physical_density = f.filter_density(x);
  if (plan.back().project) project_solid(g, res.physical_density, plan.back().beta, eta);
  else if (plan.back().cont) project_solid(g, res.physical_density, mma_beta, eta);
  if (res.non_convergent) {
    res.converged = false;
    res.compliance = res.history.empty() ? 0.0 : res.history.back().compliance;
  } else {
    try {
      const SimpCompliance fc = simp_compliance(g, p, res.physical_density, bcs, loads,
                                                o.cg_tolerance, o.cg_max_iterations, o.use_mg);
      res.compliance = fc.compliance;
    } catch (const SolverNonConvergence&) {
      res.non_convergent = true;
      res.converged = false;
      res.compliance = res.history.empty() ? 0.0 : res.history.back().compliance;
    }
  }
  res.volume_fraction = volfrac(res.physical_density);
  return res;
}

std::vector<std::uint8_t> effective_design_mask(const Grid& g,
                                                const std::vector<std::uint8_t>& mask) {
  if (mask.size() != g.voxels())
    throw std::invalid_argument("effective_design_mask: mask size != grid.voxel_count()");
  std::vector<std::uint8_t> eff = mask;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (!g.is_solid(i, j, k)) eff[g.index(i, j, k)] = kFrozenVoid;
  return eff;
}

Grid analysis_grid_for_mask(const Grid& g, const std::vector<std::uint8_t>& eff) {
  Grid a = g;
  if (a.solid.empty()) a.solid.assign(g.voxels(), 1);
  for (std::size_t e = 0; e < a.solid.size(); ++e)
    if (eff[e] == kFrozenVoid) a.solid[e] = 0;
  return a;
}

void apply_mask_pins(const std::vector<std::uint8_t>& eff, std::vector<double>& xphys) {
  for (std::size_t e = 0; e < xphys.size(); ++e) {
    if (eff[e] == kFrozenSolid) xphys[e] = 1.0;
    else if (eff[e] == kFrozenVoid) xphys[e] = 0.0;
  }
}

std::vector<double> oc_update_masked(const Grid& g, const DensityFilter& f,
                                     const std::vector<std::uint8_t>& eff,
                                     const std::vector<double>& x,
                                     const std::vector<double>& dcomp, double volfrac,
                                     double move, double density_min, int* nbis,
                                     double* lambda_out) {
  const std::size_t N = g.voxels();
  if (x.size() != N || dcomp.size() != N || eff.size() != N)
    throw std::invalid_argument("oc_update_masked: vector size != voxel_count");
  const std::vector<double> dc = f.filter_sensitivity(dcomp);
  std::vector<double> ones(N, 0.0);
  std::size_t n_active = 0;
  for (std::size_t e = 0; e < N; ++e)
    if (eff[e] == kActive) { ones[e] = 1.0; ++n_active; }
  const std::vector<double> dv = f.filter_sensitivity(ones);
  const double target = volfrac * static_cast<double>(n_active);
  auto candidate = [&](double lambda) {
    std::vector<double> xn = x;  // frozen voxels keep their x
    for (std::size_t e = 0; e < N; ++e) {
      if (eff[e] != kActive) continue;
      const double xe = x[e];
      double be = 0.0;
      if (dv[e] > 0.0) be = -dc[e] / (lambda * dv[e]);
      if (be < 0.0) be = 0.0;
      double xt = xe * std::sqrt(be);
      const double lo = std::max(density_min, xe - move);
      const double hi = std::min(1.0, xe + move);
      if (xt < lo) xt = lo;
      if (xt > hi) xt = hi;
      xn[e] = xt;
    }
    return xn;
  };
  auto phys_vol = [&](const std::vector<double>& xn) {
    const std::vector<double> xp = f.filter_density(xn);
    double v = 0.0;
    for (double q : xp) v += q;
    return v;
  };
  double l1 = 0.0, l2 = 1.0;
  while (l2 < 1e30 && phys_vol(candidate(l2)) > target) l2 *= 2.0;
  int it = 0;
  for (; it < 100 && (l2 - l1) > 1e-6 * (l1 + l2); ++it) {
    const double lmid = 0.5 * (l1 + l2);
    if (phys_vol(candidate(lmid)) > target) l1 = lmid; else l2 = lmid;
  }
  if (nbis) *nbis = it;
  if (lambda_out) *lambda_out = 0.5 * (l1 + l2);
  return candidate(0.5 * (l1 + l2));
}

SimpResult simp_optimize_masked(const Grid& g, const SimpParams& p,
                                const std::vector<DirichletBC>& bcs,
                                const std::vector<NodalLoad>& loads, const SimpOptions& o,
                                const std::vector<std::uint8_t>& mask) {
  if (!(o.volume_fraction > 0.0 && o.volume_fraction <= 1.0))
    throw std::invalid_argument("simp_optimize: volume_fraction must be in (0, 1]");
  if (!(o.move > 0.0)) throw std::invalid_argument("simp_optimize: move must be > 0");
  if (o.max_iterations < 1) throw std::invalid_argument("simp_optimize: max_iterations must be >= 1");
  if (mask.size() != g.voxels()) throw std::invalid_argument("simp_optimize: mask size != voxel_count");
  const std::vector<std::uint8_t> eff = effective_design_mask(g, mask);
  const DensityFilter f = make_density_filter(g, o.filter_radius, eff);
  const Grid analysis = analysis_grid_for_mask(g, eff);
  double n_active = 0.0;
  for (std::uint8_t m : eff)
    if (m == kActive) n_active += 1.0;
  const double vf_target = o.volume_fraction;
  auto active_volfrac = [&](const std::vector<double>& xp) {
    double v = 0.0;
    for (double q : xp) v += q;
    return n_active > 0.0 ? v / n_active : 0.0;
  };
  std::vector<double> x(g.voxels(), 0.0);
  for (std::size_t e = 0; e < x.size(); ++e) {
    if (eff[e] == kActive) x[e] = vf_target;
    else if (eff[e] == kFrozenSolid) x[e] = 1.0;
  }
  SimpResult res;
  std::vector<std::uint8_t> active(g.voxels(), 0);
  for (std::size_t e = 0; e < eff.size(); ++e) active[e] = eff[e] == kActive;
  // Stage plan (build_stage_plan, simp.cpp:1053-1068), as in simp_optimize.
  // (cont: the single dynamic stage of MMA with projection, :1036-1068)
  struct Stage { bool project; double beta, move; int iterations; bool cont; };
  std::vector<Stage> plan;
  validate_mma_projection(o);
  if (o.mma_projection) plan.push_back({false, 0.0, o.move, o.max_iterations, true});
  else if (o.projection.empty()) plan.push_back({false, 0.0, o.move, o.max_iterations, false});
  else
    for (const ProjStage& ps : o.projection)
      plan.push_back({true, ps.beta, ps.move, ps.iterations, false});
  double mma_beta = o.mma_projection_beta0;
  std::size_t stage_start = 0;  // history index where the current beta stage began
  double adaptive_move_val = 0.0;
  std::vector<std::size_t> design_dof;  // the adaptive move's voxels: the Active ones
  for (std::size_t e = 0; e < eff.size(); ++e)
    if (eff[e] == kActive) design_dof.push_back(e);
  if (o.mma && !o.projection.empty())
    throw std::invalid_argument("simp_optimize: projection requires the OC updater");
  const double eta = o.projection_eta;
  MmaState mma_state;
  int mma_iter = 0;
  std::vector<double> curve;
  for (const Stage& st : plan) {
    bool stage_converged = false;
    for (int it = 0; it < st.iterations; ++it) {
      // physical density: filter, project the Active voxels, pin (:3196-3204)
      const double cur_beta = st.cont ? mma_beta : st.beta;
      double cur_move = st.cont ? mma_continuation_move(st.move, cur_beta) : st.move;
      if (o.adaptive_move) {
        // seeded at the reference move for the first two MMA steps (and after a
        // beta advance restarts the state), then evolved (:2179-2194)
        if (mma_iter < 2 || mma_state.xold2.empty()) adaptive_move_val = cur_move;
        else
          adaptive_move_val = adapt_move_value(
              adaptive_move_val,
              oscillation_fraction(x, mma_state.xold1, mma_state.xold2, design_dof), o);
        cur_move = adaptive_move_val;
      }
      std::vector<double> xphys = f.filter_density(x);
      if (st.project || st.cont) project_active(eff, xphys, cur_beta, eta);
      apply_mask_pins(eff, xphys);
      SimpParams tp = p;
      if (!o.penalty_continuation.empty())
        tp.penalty = penalty_at_iteration(o.penalty_continuation, res.iterations + 1);
      SimpCompliance c;
      try {
        c = simp_compliance(analysis, tp, xphys, bcs, loads, o.cg_tolerance, o.cg_max_iterations,
                            o.use_mg);
      } catch (const SolverNonConvergence&) {
        res.non_convergent = true;
        break;
      }
      // core/src/simp/simp.cpp:3306-3313 (projected OC), 3351-3357 (MMA) or the OC update
      std::vector<double> xn;
      if (st.cont) {
        // mma_update_masked_projected at the current continuation beta (:3296-3307)
        const std::vector<double> xtilde = f.filter_density(x);
        xn = mma_update(mma_state, ++mma_iter, g, f, x, c.dc, vf_target, cur_move,
                        p.density_min, nullptr, &active, &xtilde, cur_beta, eta);
      } else if (st.project)
        xn = oc_update_masked_projected(g, f, eff, x, f.filter_density(x), c.dc, st.beta, eta,
                                        vf_target, st.move, p.density_min);
      else if (o.mma)
        xn = mma_update(mma_state, ++mma_iter, g, f, x, c.dc, vf_target, cur_move, p.density_min,
                        nullptr, &active);
      else
        xn = oc_update_masked(g, f, eff, x, c.dc, vf_target, st.move, p.density_min);
      double change = 0.0;
      for (std::size_t e = 0; e < x.size(); ++e) change = std::max(change, std::fabs(xn[e] - x[e]));
      x = xn;
      if (res.iterations == 0) res.initial_compliance = c.compliance;
      ++res.iterations;
      std::vector<double> xa = f.filter_density(x);
      if (st.project) project_active(eff, xa, st.beta, eta);
      res.history.push_back({c.compliance, change, active_volfrac(xa), c.cg.iterations});
      curve.push_back(c.compliance);
      if (st.cont) {
        // beta continuation (:2461-2491): a plateau of the current beta stage
        // doubles beta (capped) and restarts the MMA state; a plateau at the
        // cap ends the run. The progress gate guards the first stage only.
        const std::vector<double> sc(curve.begin() + static_cast<std::ptrdiff_t>(stage_start),
                                     curve.end());
        const double min_drop = stage_start == 0 ? o.mma_plateau_min_drop : 0.0;
        if (mma_objective_plateau(sc, o.mma_plateau_window, o.mma_plateau_tol, min_drop,
                                  o.mma_plateau_flat_windows)) {
          if (mma_beta < o.mma_projection_beta_max) {
            mma_beta = std::min(mma_beta * 2.0, o.mma_projection_beta_max);
            stage_start = curve.size();
            mma_state = MmaState{};
            mma_iter = 0;
          } else {
            stage_converged = true;
            break;
          }
        }
        continue;
      }
      const bool stop = (o.mma && o.mma_plateau_window > 0)
                            ? mma_objective_plateau(curve, o.mma_plateau_window, o.mma_plateau_tol,
                                                    o.mma_plateau_min_drop,
                                                    o.mma_plateau_flat_windows)
                            : change < o.change_tol;
      if (stop) { stage_converged = true; break; }
    }
    res.converged = stage_converged;
    if (res.non_convergent) break;
  }
  res.design = x;
  res.physical_density = f.filter_density(x);
  if (plan.back().project) project_active(eff, res.physical_density, plan.back().beta, eta);
  else if (plan.back().cont) project_active(eff, res.physical_density, mma_beta, eta);
  const std::vector<double> unpinned = res.physical_density;
  apply_mask_pins(eff, res.physical_density);
  if (res.non_convergent) {
    res.compliance = res.history.empty() ? 0.0 : res.history.back().compliance;
  } else {
    try {
      const SimpCompliance fc = simp_compliance(analysis, p, res.physical_density, bcs, loads,
                                                o.cg_tolerance, o.cg_max_iterations, o.use_mg);
      res.compliance = fc.compliance;
    } catch (const SolverNonConvergence&) {
      res.non_convergent = true;
      res.converged = false;
      res.compliance = res.history.empty() ? 0.0 : res.history.back().compliance;
    }
  }
  res.volume_fraction = active_volfrac(unpinned);
  return res;
}

double heaviside_project(double rt, double beta, double eta) {
  const double den = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta));
  return (std::tanh(beta * eta) + std::tanh(beta * (rt - eta))) / den;
}
double heaviside_project_derivative(double rt, double beta, double eta) {
  const double den = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta));
  const double ch = std::cosh(beta * (rt - eta));
  return beta / (ch * ch * den);
}

void project_active(const std::vector<std::uint8_t>& eff, std::vector<double>& x, double beta,
                    double eta) {
  for (std::size_t e = 0; e < x.size(); ++e)
    if (eff[e] == kActive) x[e] = heaviside_project(x[e], beta, eta);
}

std::vector<double> oc_update_masked_projected(const Grid& g, const DensityFilter& f,
                                               const std::vector<std::uint8_t>& eff,
                                               const std::vector<double>& x,
                                               const std::vector<double>& xtilde,
                                               const std::vector<double>& dcomp, double beta,
                                               double eta, double volfrac, double move,
                                               double density_min) {
  const std::size_t N = g.voxels();
  if (x.size() != N || xtilde.size() != N || dcomp.size() != N || eff.size() != N)
    throw std::invalid_argument("oc_update_masked_projected: vector size != voxel_count");
  // chain rule through the projection on the Active set, 0 elsewhere
  std::vector<double> dcp(N, 0.0), dpr(N, 0.0);
  std::size_t n_active = 0;
  for (std::size_t e = 0; e < N; ++e) {
    if (eff[e] != kActive) continue;
    dpr[e] = heaviside_project_derivative(xtilde[e], beta, eta);
    dcp[e] = dcomp[e] * dpr[e];
    ++n_active;
  }
  const std::vector<double> dc = f.filter_sensitivity(dcp);
  const std::vector<double> dv = f.filter_sensitivity(dpr);
  const double target = volfrac * static_cast<double>(n_active);
  auto candidate = [&](double lambda) {
    std::vector<double> xn = x;  // FrozenSolid keeps 1, FrozenVoid 0
    for (std::size_t e = 0; e < N; ++e) {
      if (eff[e] != kActive) continue;
      const double xe = x[e];
      double be = -dc[e] / (lambda * dv[e]);  // unclamped dv
      if (!std::isfinite(be)) be = 1.0;       // 0/0 holds the voxel
      if (be < 0.0) be = 0.0;
      double xt = xe * std::sqrt(be);
      const double lo = std::max(density_min, xe - move);
      const double hi = std::min(1.0, xe + move);
      if (xt < lo) xt = lo;
      if (xt > hi) xt = hi;
      xn[e] = xt;
    }
    return xn;
  };
  // projected Active volume of a candidate (the filter is 0 off the Active set)
  auto proj_vol = [&](const std::vector<double>& xn) {
    const std::vector<double> xp = f.filter_density(xn);
    double v = 0.0;
    for (std::size_t e = 0; e < N; ++e)
      if (eff[e] == kActive) v += heaviside_project(xp[e], beta, eta);
    return v;
  };
  double l1 = 0.0, l2 = 1.0;
  while (l2 < 1e30 && proj_vol(candidate(l2)) > target) l2 *= 2.0;
  for (int it = 0; it < 100 && (l2 - l1) > 1e-6 * (l1 + l2); ++it) {
    const double lmid = 0.5 * (l1 + l2);
    if (proj_vol(candidate(lmid)) > target) l1 = lmid; else l2 = lmid;
  }
  return candidate(0.5 * (l1 + l2));
}

}  // namespace oracle

namespace oracle {

// core/src/fea/assembly.cpp:976-996.
std::vector<double> von_mises_field(const Grid& g, double E, double nu,
                                    const std::vector<double>& u) {
  if (u.size() != 3 * static_cast<std::size_t>(g.nodes()))
    throw std::invalid_argument("von_mises_field: displacement size != 3 * nodes");
  std::vector<double> field(g.voxels(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.is_solid(i, j, k)) continue;
        const auto en = element_nodes(g, i, j, k);
        double ue[24];
        for (int a = 0; a < 8; ++a)
          for (int c = 0; c < 3; ++c) ue[3 * a + c] = u[3 * static_cast<std::size_t>(en[a]) + c];
        field[g.index(i, j, k)] = hex8_stress(E, nu, g.h, ue).von_mises;
      }
  return field;
}

// core/src/simp/observability.cpp:35-75: sign, exponent re-bias, mantissa with
// round-to-nearest-even on the dropped bits; subnormals by shifting in the
// implicit bit; overflow to Inf; NaN keeps only the quiet bit.
std::uint16_t float_to_half(float v) {
  std::uint32_t x;
  std::memcpy(&x, &v, 4);
  const std::uint32_t sign = (x >> 16) & 0x8000u;
  const int biased = static_cast<int>((x >> 23) & 0xFFu);
  const std::uint32_t mant = x & 0x7FFFFFu;
  if (biased == 255) return static_cast<std::uint16_t>(sign | 0x7C00u | (mant ? 0x200u : 0u));
  const int e = biased - 112;  // half's biased exponent
  if (e >= 31) return static_cast<std::uint16_t>(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return static_cast<std::uint16_t>(sign);
    const std::uint32_t m = mant | 0x800000u;
    const int sh = 14 - e;
    std::uint32_t hm = m >> sh;
    const std::uint32_t rem = m & ((1u << sh) - 1u), mid = 1u << (sh - 1);
    if (rem > mid || (rem == mid && (hm & 1u))) ++hm;
    return static_cast<std::uint16_t>(sign | hm);
  }
  std::uint32_t hv = sign | (static_cast<std::uint32_t>(e) << 10) | (mant >> 13);
  const std::uint32_t rem = mant & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (hv & 1u))) ++hv;
  return static_cast<std::uint16_t>(hv);
}

// core/src/io/design_store.cpp:75-88: FNV-1a with the 64-bit prime and the
// reference's offset basis, 1469598103934665603 (:78; the published basis has
// one more digit, 14695981039346656037). The reference's design.bin files
// carry fingerprints under its own basis.
std::uint64_t fnv1a64(const void* data, std::size_t n, std::uint64_t basis) {
  const auto* b = static_cast<const std::uint8_t*>(data);
  std::uint64_t h = basis;
  for (std::size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

}  // namespace oracle

namespace oracle {

// core/src/simp/simp.cpp:1433-1488 and core/src/simp/mma_joint.cpp:18-118
// (asyinit 0.5, asyincr 1.2, asydecr 0.7, albefa 0.1, raa0 1e-5, move as a
// fraction of xrange, the 1-D dual bisection).
double mma_continuation_move(double move, double beta) {
  return beta > 8.0 ? move * (8.0 / beta) : move;
}

double oscillation_fraction(const std::vector<double>& x, const std::vector<double>& xold1,
                            const std::vector<double>& xold2,
                            const std::vector<std::size_t>& dof) {
  long long moved = 0, reversed = 0;
  for (std::size_t e : dof) {
    const double d1 = x[e] - xold1[e], d2 = xold1[e] - xold2[e];
    if (d1 == 0.0 || d2 == 0.0) continue;  // at a bound: no direction
    ++moved;
    if (d1 * d2 < 0.0) ++reversed;
  }
  return moved > 0 ? static_cast<double>(reversed) / static_cast<double>(moved) : 0.0;
}

double adapt_move_value(double cur, double f_osc, const SimpOptions& o) {
  if (f_osc > o.adaptive_move_osc_hi) cur *= o.adaptive_move_shrink;
  else if (f_osc < o.adaptive_move_osc_lo) cur *= o.adaptive_move_grow;
  return std::min(std::max(cur, o.adaptive_move_min), o.adaptive_move_max);
}

std::vector<double> mma_update(MmaState& st, int mma_iter, const Grid& g, const DensityFilter& f,
                               const std::vector<double>& x, const std::vector<double>& dcomp,
                               double volfrac, double move, double density_min,
                               double* lambda_out, const std::vector<std::uint8_t>* dset,
                               const std::vector<double>* xtilde, double beta, double eta) {
  const std::size_t N = g.voxels();
  if (x.size() != N || dcomp.size() != N || (xtilde && xtilde->size() != N))
    throw std::invalid_argument("mma_update: vector size != voxel_count");
  std::vector<double> ones(N, 0.0);
  std::vector<std::size_t> dof;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const std::size_t e = g.index(i, j, k);
        if (dset ? (*dset)[e] != 0 : g.is_solid(i, j, k)) {
          ones[e] = 1.0;
          dof.push_back(e);
        }
      }
  std::vector<double> dc, dv;
  double vol = 0.0;
  if (xtilde) {
    // projected: dproj on the design set, dc * dproj and dproj through the
    // filter transpose, volume = sum of the projected filtered design
    std::vector<double> dcp(N, 0.0), dpr(N, 0.0);
    for (std::size_t e : dof) {
      dpr[e] = heaviside_project_derivative((*xtilde)[e], beta, eta);
      dcp[e] = dcomp[e] * dpr[e];
      vol += heaviside_project((*xtilde)[e], beta, eta);
    }
    dc = f.filter_sensitivity(dcp);
    dv = f.filter_sensitivity(dpr);
  } else {
    dc = f.filter_sensitivity(dcomp);
    dv = f.filter_sensitivity(ones);
    const std::vector<double> xphys = f.filter_density(x);
    for (double v : xphys) vol += v;
  }
  const double target = volfrac * static_cast<double>(dof.size());
  double scale = 0.0;
  for (std::size_t e : dof) scale = std::max(scale, std::fabs(dc[e]));
  if (scale > 0.0)
    for (std::size_t e : dof) dc[e] /= scale;
  const double g0 = vol - target;

  const double asyinit = 0.5, asyincr = 1.2, asydecr = 0.7, albefa = 0.1, raa0 = 1e-5;
  const double xmin = density_min, xmax = 1.0, xrange = xmax - xmin;
  if (st.low.size() != N) {
    st.low.assign(N, 0.0);
    st.upp.assign(N, 0.0);
  }
  for (std::size_t e : dof) {
    const double xe = x[e];
    if (mma_iter <= 2) {
      st.low[e] = xe - asyinit * xrange;
      st.upp[e] = xe + asyinit * xrange;
    } else {
      const double s = (xe - st.xold1[e]) * (st.xold1[e] - st.xold2[e]);
      const double gam = s < 0.0 ? asydecr : (s > 0.0 ? asyincr : 1.0);
      double L = xe - gam * (st.xold1[e] - st.low[e]);
      double U = xe + gam * (st.upp[e] - st.xold1[e]);
      L = std::min(L, xe - 0.01 * xrange);
      L = std::max(L, xe - 10.0 * xrange);
      U = std::max(U, xe + 0.01 * xrange);
      U = std::min(U, xe + 10.0 * xrange);
      st.low[e] = L;
      st.upp[e] = U;
    }
  }
  std::vector<double> p0(N, 0.0), q0(N, 0.0), p1(N, 0.0), q1(N, 0.0), al(N, 0.0), be(N, 0.0);
  double b = -g0;
  for (std::size_t e : dof) {
    const double xe = x[e], L = st.low[e], U = st.upp[e];
    const double ux1 = U - xe, xl1 = xe - L;
    const double dcp = std::max(dc[e], 0.0), dcm = std::max(-dc[e], 0.0);
    const double dvp = std::max(dv[e], 0.0), dvm = std::max(-dv[e], 0.0);
    p0[e] = ux1 * ux1 * (1.001 * dcp + 0.001 * dcm + raa0 / xrange);
    q0[e] = xl1 * xl1 * (0.001 * dcp + 1.001 * dcm + raa0 / xrange);
    p1[e] = ux1 * ux1 * (1.001 * dvp + 0.001 * dvm + raa0 / xrange);
    q1[e] = xl1 * xl1 * (0.001 * dvp + 1.001 * dvm + raa0 / xrange);
    b += p1[e] / ux1 + q1[e] / xl1;
    al[e] = std::max({xmin, L + albefa * (xe - L), xe - move * xrange});
    be[e] = std::min({xmax, U - albefa * (U - xe), xe + move * xrange});
  }
  auto candidate = [&](double lam) {
    std::vector<double> xn(N, 0.0);
    for (std::size_t e : dof) {
      const double pl = std::sqrt(p0[e] + lam * p1[e]);
      const double ql = std::sqrt(q0[e] + lam * q1[e]);
      double xt = (pl * st.low[e] + ql * st.upp[e]) / (pl + ql);
      if (xt < al[e]) xt = al[e];
      if (xt > be[e]) xt = be[e];
      xn[e] = xt;
    }
    return xn;
  };
  auto gval = [&](const std::vector<double>& xc) {
    double s = 0.0;
    for (std::size_t e : dof) s += p1[e] / (st.upp[e] - xc[e]) + q1[e] / (xc[e] - st.low[e]);
    return s - b;
  };
  double lambda = 0.0;
  if (gval(candidate(0.0)) > 0.0) {
    double l1 = 0.0, l2 = 1.0;
    while (l2 < 1e30 && gval(candidate(l2)) > 0.0) l2 *= 2.0;
    for (int it = 0; it < 100 && (l2 - l1) > 1e-9 * (1.0 + l1 + l2); ++it) {
      const double lm = 0.5 * (l1 + l2);
      if (gval(candidate(lm)) > 0.0) l1 = lm;
      else l2 = lm;
    }
    lambda = 0.5 * (l1 + l2);
  }
  std::vector<double> xn = candidate(lambda);
  st.xold2 = st.xold1;
  st.xold1 = x;
  if (lambda_out) *lambda_out = lambda;
  return xn;
}

// core/src/simp/simp.cpp:1081-1134: running-minimum improvement over the
// trailing window below rel_tol, behind the progress gate (or its flatness
// escape).
bool mma_objective_plateau(const std::vector<double>& c, int window, double rel_tol,
                           double min_drop, int min_flat_windows) {
  if (window <= 0) return false;
  const int n = static_cast<int>(c.size());
  if (n < window + 1) return false;
  double best_now = c[0];
  for (int i = 1; i < n; ++i) best_now = std::min(best_now, c[i]);
  if (min_drop > 0.0 && !(best_now <= c[0] * (1.0 - min_drop))) {
    bool escape = false;
    if (min_flat_windows > 0) {
      const int span = min_flat_windows * window;
      if (n >= span + 1) {
        double best_before = c[0];
        for (int i = 1; i < n - span; ++i) best_before = std::min(best_before, c[i]);
        if (best_before > 0.0) escape = (best_before - best_now) / best_before < rel_tol;
      }
    }
    if (!escape) return false;
  }
  double best_prev = c[0];
  for (int i = 1; i < n - window; ++i) best_prev = std::min(best_prev, c[i]);
  if (!(best_prev > 0.0)) return false;
  return (best_prev - best_now) / best_prev < rel_tol;
}

}  // namespace oracle
