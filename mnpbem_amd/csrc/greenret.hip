// Retarded Green function of a closed flat-triangle particle on gfx950:
// face geometry, refinement selection, Taylor refinement tables (self terms
// by polar integration, near field by the 28-point rule), the closed-surface
// sum rule, and dense assembly of G, F, H1, H2 for a wavenumber k.
//
// Algorithm of MNPBEM Greenfun/@greenret/private/init.m, eval1.m and
// Greenfun/@compgreenret/private/initclosed.m, restructured for the GPU:
// one workgroup per source column for the O(n^2) scans (refinement
// selection, sum rule), one thread per refined pair for the quadrature, and
// a coalesced column-major streaming kernel for the O(n^2) assembly.
#include <algorithm>
#include <cmath>
#include <vector>

#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>

#include "mnp_greenret.hpp"
#include "mnp_linalg.hpp"
#include "mnp_nearfield.hpp"

namespace mnp {

// ------------------------------------------------------------ library state
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
static hipStream_t g_stream = nullptr;
static thread_local hipStream_t t_stream = nullptr;  // a worker's own stream
static thread_local bool t_in_worker = false;
hipStream_t stream() { return t_stream ? t_stream : g_stream; }
hipStream_t thread_stream() { return t_stream; }
void set_thread_stream(hipStream_t s) { t_stream = s; }

// ------------------------------------------------------------------ timers
namespace {
struct TimerRec {
  int kind;
  double work, bytes;
  hipEvent_t a, b;
};
bool g_timers_on = false;
std::mutex g_timer_mu;  // timed launches may come from worker threads
std::vector<TimerRec> g_timer_recs;
std::vector<hipEvent_t> g_event_pool;  // events are reused: creating one per launch costs host time
mnp_timers g_timer_acc{};
hipEvent_t event_get() {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  MNP_HIP(hipEventCreate(&e));
  return e;
}
void timers_flush() {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  if (g_timer_recs.empty()) return;
  for (auto& r : g_timer_recs) {
    float ms = 0;
    MNP_HIP(hipEventSynchronize(r.b));
    MNP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    g_timer_acc.launches[r.kind] += 1;
    g_timer_acc.ms[r.kind] += ms;
    g_timer_acc.work[r.kind] += r.work;
    g_timer_acc.bytes[r.kind] += r.bytes;
    g_event_pool.push_back(r.a);
    g_event_pool.push_back(r.b);
  }
  g_timer_recs.clear();
}
}  // namespace

static thread_local bool t_timers_paused = false;  // inside a stream capture
void pause_timers(bool on) { t_timers_paused = on; }
TimerScope::TimerScope(int kind_, double work_, double bytes_) : kind(kind_), work(work_), bytes(bytes_) {
  if (!g_timers_on || kind < 0 || t_timers_paused) return;
  hipEvent_t a = event_get();
  MNP_HIP(hipEventRecord(a, stream()));
  ev0 = a;
}
TimerScope::~TimerScope() {
  if (!ev0) return;
  hipEvent_t b;
  try {
    b = event_get();
  } catch (...) {
    return;
  }
  (void)hipEventRecord(b, stream());
  size_t nrec;
  {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    g_timer_recs.push_back(TimerRec{kind, work, bytes, (hipEvent_t)ev0, b});
    nrec = g_timer_recs.size();
  }
  if (nrec > 16384) {
    try {
      timers_flush();
    } catch (...) {
    }
  }
}

// ----------------------------------------------------- concurrent tasks
namespace {
// Three persistent worker threads (detached: they idle until the process
// ends, and their thread-local workspaces live as long).
struct WorkerPool {
  struct W {
    std::mutex mu;
    std::condition_variable cv;
    std::function<void()> job;
    bool has = false, done = true;
  };
  W w[3];
  WorkerPool() {
    for (auto& x : w) std::thread([&x] { loop(x); }).detach();
  }
  static void loop(W& x) {
    t_in_worker = true;
    for (;;) {
      std::function<void()> j;
      {
        std::unique_lock<std::mutex> lk(x.mu);
        x.cv.wait(lk, [&] { return x.has; });
        j = std::move(x.job);
        x.has = false;
      }
      j();  // never throws: run_concurrent's jobs catch everything
      {
        std::lock_guard<std::mutex> lk(x.mu);
        x.done = true;
      }
      x.cv.notify_all();
    }
  }
  void start(int i, std::function<void()> j) {
    {
      std::lock_guard<std::mutex> lk(w[i].mu);
      w[i].job = std::move(j);
      w[i].has = true;
      w[i].done = false;
    }
    w[i].cv.notify_all();
  }
  void wait(int i) {
    std::unique_lock<std::mutex> lk(w[i].mu);
    w[i].cv.wait(lk, [&] { return w[i].done; });
  }
};
WorkerPool& pool() {
  static WorkerPool* p = new WorkerPool;  // never destroyed (detached threads)
  return *p;
}
// one stream per worker thread and device, default priority
hipStream_t worker_stream(int dev) {
  static thread_local std::vector<hipStream_t> s;
  if ((int)s.size() <= dev) s.resize(dev + 1, nullptr);
  if (!s[dev]) MNP_HIP(hipStreamCreateWithFlags(&s[dev], hipStreamNonBlocking));
  return s[dev];
}
}  // namespace

static thread_local bool t_in_run = false;  // this thread is inside run_concurrent

void run_concurrent(std::vector<std::function<void()>>& tasks) {
  if (tasks.size() <= 1 || tasks.size() > 4 || t_in_worker || t_in_run) {
    for (auto& t : tasks) t();
    return;
  }
  // one caller at a time owns the pool
  static std::mutex run_mu;
  std::lock_guard<std::mutex> run_lk(run_mu);
  struct InRun {
    InRun() { t_in_run = true; }
    ~InRun() { t_in_run = false; }
  } in_run;
  int dev = 0;
  MNP_HIP(hipGetDevice(&dev));
  const hipStream_t caller = stream();
  const int nt = (int)tasks.size();
  std::vector<hipEvent_t> ev(nt, nullptr);  // ev[0]: caller's start; ev[i]: task i's end
  for (auto& e : ev) MNP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  MNP_HIP(hipEventRecord(ev[0], caller));
  const int split = split_bf16();
  std::vector<std::exception_ptr> err(nt);
  std::vector<char> ended(nt, 0);
  WorkerPool& P = pool();
  for (int i = 1; i < nt; ++i)
    P.start(i - 1, [&, i] {
      hipStream_t s = nullptr;
      try {
        MNP_HIP(hipSetDevice(dev));
        s = worker_stream(dev);
        t_stream = s;
        set_split_bf16(split);
        MNP_HIP(hipStreamWaitEvent(s, ev[0], 0));
        tasks[i]();
        MNP_HIP(hipEventRecord(ev[i], s));
        ended[i] = 1;
      } catch (...) {
        err[i] = std::current_exception();
        if (s) (void)hipStreamSynchronize(s);  // nothing of this task stays in flight
      }
      t_stream = nullptr;
      set_split_bf16(0);
    });
  try {
    tasks[0]();
  } catch (...) {
    err[0] = std::current_exception();
  }
  for (int i = 1; i < nt; ++i) P.wait(i - 1);
  for (int i = 1; i < nt; ++i)
    if (ended[i]) (void)hipStreamWaitEvent(caller, ev[i], 0);
  if (err[0]) (void)hipStreamSynchronize(caller);
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// ------------------------------------------------------------- quadrature
// Quadrature tables live in constant memory; the 28-point rule is the
// Lyness-Jespersen rule 18 (Misc/integration/@quadface/private/
// triangle_unit_set.m:471-505), the polar rule follows
// Misc/integration/@quadface/private/init.m:25-53.
__constant__ double c_q18[3][28];
__constant__ double c_qpol[3][MNP_MAX_POLAR];

void rule18(double* x, double* y, double* w) {
  const double a = 1.0 / 3.0, b = 0.9480217181434233, c = 0.02598914092828833,
               d = 0.8114249947041546, e = 0.09428750264792270, f = 0.01072644996557060,
               g = 0.4946367750172147, p = 0.5853132347709715, q = 0.2073433826145142,
               r = 0.1221843885990187, s = 0.4389078057004907, t = 0.6779376548825902,
               u = 0.04484167758913055, v = 0.27722066752827925, W = 0.8588702812826364,
               X = 0.0, Y = 0.1411297187173636;
  const double w1 = 0.08797730116222190, w2 = 0.008744311553736190, w3 = 0.03808157199393533,
               w4 = 0.01885544805613125, w5 = 0.07215969754474100, w6 = 0.06932913870553720,
               w7 = 0.04105631542928860, w8 = 0.007362383783300573;
  const double xs[28] = {a, b, c, c, d, e, e, f, g, g, p, q, q, r,
                         s, s, t, t, u, u, v, v, W, W, X, X, Y, Y};
  const double ys[28] = {a, c, b, c, e, d, e, g, f, g, q, p, q, s,
                         r, s, u, v, t, v, t, u, X, Y, W, Y, W, X};
  const double ws[28] = {w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w6,
                         w6, w6, w7, w7, w7, w7, w7, w7, w8, w8, w8, w8, w8, w8};
  for (int i = 0; i < 28; ++i) {
    x[i] = xs[i];
    y[i] = ys[i];
    w[i] = ws[i];
  }
}

// Legendre-Gauss-Lobatto nodes (n+1 points, +1 .. -1) by Newton iteration
// from the Chebyshev-Gauss-Lobatto points (Misc/integration/lglnodes.m).
static void lgl(int n, std::vector<double>& x, std::vector<double>& w) {
  const int n1 = n + 1;
  x.assign(n1, 0.0);
  w.assign(n1, 0.0);
  for (int i = 0; i < n1; ++i) x[i] = std::cos(M_PI * i / n);
  std::vector<double> p(n1 * n1), xold(n1, 2.0);
  for (int it = 0; it < 100; ++it) {
    double diff = 0;
    for (int i = 0; i < n1; ++i) diff = std::max(diff, std::fabs(x[i] - xold[i]));
    if (!(diff > 2.220446049250313e-16)) break;
    xold = x;
    for (int i = 0; i < n1; ++i) {
      double* pi = &p[i * n1];
      pi[0] = 1.0;
      pi[1] = x[i];
      for (int k = 2; k <= n; ++k) pi[k] = ((2 * k - 1) * x[i] * pi[k - 1] - (k - 1) * pi[k - 2]) / k;
      x[i] = xold[i] - (x[i] * pi[n1 - 1] - pi[n - 1]) / (n1 * pi[n1 - 1]);
    }
  }
  for (int i = 0; i < n1; ++i) w[i] = 2.0 / (n * n1 * p[i * n1 + n1 - 1] * p[i * n1 + n1 - 1]);
}

// Polar rule on the unit triangle: radial LGL x angular LGL in each of the
// three sub-triangles around the centroid of the equilateral reference
// triangle, mapped to unit-triangle coordinates; weights sum to one.
static int polar_rule(int nr, int np, double* x, double* y, double* w) {
  std::vector<double> x1, w1, x2, w2;
  lgl(nr, x1, w1);
  lgl(np, x2, w2);
  const int n1 = nr + 1, n2 = np + 1, m = n1 * n2;
  if (3 * m > MNP_MAX_POLAR) return -1;
  const double phi0 = 120.0 / 180.0 * M_PI;
  double wsum = 0;
  for (int s = 0; s < 3; ++s)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {  // radius fastest (ndgrid order)
        const double rho = 0.5 * (x1[i] + 1 + 1e-6);
        const double phi = (270 + 60 * x2[j]) / 180 * M_PI;
        const double rad = 1.0 / std::fabs(2 * std::sin(phi));
        const double ph = phi + s * phi0;
        const double xx = std::cos(ph) * (rho * rad), yy = std::sin(ph) * (rho * rad);
        const int q = s * m + j * n1 + i;
        x[q] = (1 - std::sqrt(3.0) * xx - yy) / 3;
        y[q] = (1 + std::sqrt(3.0) * xx - yy) / 3;
        w[q] = w1[i] * w2[j] * rho * rad * rad;
        wsum += w[q];
      }
  for (int q = 0; q < 3 * m; ++q) w[q] /= wsum;
  return 3 * m;
}

// ------------------------------------------------------------------ kernels
// Face geometry (Particles/@particle/private/norm_flat.m:17-50), the
// enclosing radius (Misc/+misc/bradius.m) and the quadrature area
// 0.5|(v2-v1)x(v3-v1)| (quad_flat.m:21-24).
__global__ void k_geometry(int64_t n, const double* __restrict__ verts, int64_t nv,
                           const int64_t* __restrict__ faces, Geo geo) {
#pragma clang fp contract(off)
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[3][3];
  for (int a = 0; a < 3; ++a) {
    int64_t vi = faces[i + a * n];
    for (int c = 0; c < 3; ++c) v[a][c] = verts[vi + c * nv];
  }
  double pos[3], e1[3], e2[3], nv3[3];
  for (int c = 0; c < 3; ++c) {
    pos[c] = (v[0][c] + v[1][c] + v[2][c]) / 3;
    e1[c] = v[0][c] - v[1][c];
    e2[c] = v[2][c] - v[1][c];
  }
  nv3[0] = e1[1] * e2[2] - e1[2] * e2[1];
  nv3[1] = e1[2] * e2[0] - e1[0] * e2[2];
  nv3[2] = e1[0] * e2[1] - e1[1] * e2[0];
  const double nn = sqrt(nv3[0] * nv3[0] + nv3[1] * nv3[1] + nv3[2] * nv3[2]);
  const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  double t1[3], nu[3], t2[3];
  for (int c = 0; c < 3; ++c) {
    t1[c] = e1[c] / l1;
    nu[c] = nv3[c] / nn;
  }
  t2[0] = nu[1] * t1[2] - nu[2] * t1[1];
  t2[1] = nu[2] * t1[0] - nu[0] * t1[2];
  t2[2] = nu[0] * t1[1] - nu[1] * t1[0];
  double f1[3], f2[3], qn[3];
  for (int c = 0; c < 3; ++c) {
    f1[c] = v[1][c] - v[0][c];
    f2[c] = v[2][c] - v[0][c];
  }
  qn[0] = f1[1] * f2[2] - f1[2] * f2[1];
  qn[1] = f1[2] * f2[0] - f1[0] * f2[2];
  qn[2] = f1[0] * f2[1] - f1[1] * f2[0];
  double rad = 0;
  for (int a = 0; a < 3; ++a) {
    double dx = pos[0] - v[a][0], dy = pos[1] - v[a][1], dz = pos[2] - v[a][2];
    rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz));
  }
  for (int c = 0; c < 3; ++c) {
    geo.pos[c * n + i] = pos[c];
    geo.nvec[c * n + i] = nu[c];
    geo.t1[c * n + i] = t1[c];
    geo.t2[c * n + i] = t2[c];
    for (int a = 0; a < 3; ++a) geo.vtx[(a * 3 + c) * n + i] = v[a][c];
  }
  geo.area[i] = 0.5 * nn;
  geo.qarea[i] = 0.5 * sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2]);
  geo.rad[i] = rad;
}

// One workgroup per source column j: number of refined rows.
__global__ void __launch_bounds__(256) k_refine_count(int64_t n, Geo geo, double rel,
                                                      double absc, int64_t* count) {
  const int64_t j = blockIdx.x;
  int c = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) c += refine_kind(geo, n, i, j, rel, absc) != 0;
  __shared__ int red[256];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) count[j] = red[0];
}

// Exclusive prefix sum of count[0..n) into off[0..n], one workgroup.
__global__ void __launch_bounds__(1024) k_scan(int64_t n, const int64_t* count, int64_t* off) {
  __shared__ int64_t s[1024];
  int64_t run = 0;
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < n ? count[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      int64_t t = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) off[i] = run + s[threadIdx.x] - v;
    const int64_t tot = s[1023];
    __syncthreads();
    run += tot;
  }
  if (threadIdx.x == 0) off[n] = run;
}

// Ordered fill of the refined pairs of column j (rows ascending), so the
// pair list is in column-major order like Matlab's find(ir).
__global__ void __launch_bounds__(256) k_refine_fill(int64_t n, Geo geo, double rel, double absc,
                                                     const int64_t* off, Pairs pr) {
  const int64_t j = blockIdx.x;
  __shared__ int wsum[4];
  int64_t run = off[j];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t base = 0; base < n; base += 256) {
    const int64_t i = base + threadIdx.x;
    const int kind = i < n ? refine_kind(geo, n, i, j, rel, absc) : 0;
    const unsigned long long m = __ballot(kind != 0);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wid] = __popcll(m);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      woff += w < wid ? wsum[w] : 0;
      tot += wsum[w];
    }
    if (kind) {
      const int64_t p = run + woff + before;
      pr.row[p] = (int32_t)i;
      pr.col[p] = (int32_t)j;
      pr.kind[p] = (int8_t)kind;
      if (kind == 2 && i == j) pr.diag[j] = p;
    }
    run += tot;
    __syncthreads();
  }
}

// Largest distance between a self-integrated face centroid and its polar
// points; rr = max(r, 1e-4 max r) uses this maximum over all self faces
// (greenret/private/init.m:77).
__global__ void k_self_rmax(int64_t nnz, Geo geo, int64_t n, Pairs pr, int npol,
                            unsigned long long* rmax_bits) {
#pragma clang fp contract(off)
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double rm = 0;
  if (p < nnz && pr.kind[p] == 2) {
    const int64_t i = pr.row[p], j = pr.col[p];
    for (int q = 0; q < npol; ++q) {
      const double x = c_qpol[0][q], y = c_qpol[1][q], z = 1 - x - y;
      double r2 = 0;
      for (int c = 0; c < 3; ++c) {
        const double pt = x * geo.vtx[(0 * 3 + c) * n + j] + y * geo.vtx[(1 * 3 + c) * n + j] +
                          z * geo.vtx[(2 * 3 + c) * n + j];
        const double dv = geo.pos[c * n + i] - pt;
        r2 += dv * dv;
      }
      rm = fmax(rm, sqrt(r2));
    }
  }
  // wave-level max then one atomic per wave
  for (int s = 32; s > 0; s >>= 1) rm = fmax(rm, __shfl_xor(rm, s));
  if ((threadIdx.x & 63) == 0 && rm > 0) atomicMax(rmax_bits, (unsigned long long)__double_as_longlong(rm));
}

// Taylor refinement tables, one thread per refined pair (row i, source j):
//   self (kind 2, greenret/private/init.m:44-95): polar rule over face j,
//     g_n = sum w r^(n-1)/n!, f in the (n, t1, t2) frame, rotated to xyz;
//   near field (kind 1, init.m:116-176): 28-point rule over face j with
//     r = |pos_i - y_q|, r0 = |pos_i - pos_j|,
//     g_n = sum w (r-r0)^n/(n! r),  f_0 = -sum w x/r^3,
//     f_n = sum w x [-(r-r0)^n/(r^3 n!) + (r-r0)^(n-1)/(r^2 (n-1)!)].
// Tables are column-major: g[o*nnz+p], f[(o*3+c)*nnz+p].
__global__ void __launch_bounds__(256) k_tables(int64_t nnz, Geo geo, int64_t n, Pairs pr,
                                                int order, int npol, const unsigned long long* rmax_bits,
                                                double* __restrict__ gt, double* __restrict__ ft) {
  // Without FMA contraction the quadrature points and distances round as in
  // the reference's vectorised expressions (x v1 + y v2 + (1-x-y) v3 and
  // |pos - pt|); for the innermost polar ring, 5e-7 face radii from the
  // centroid, that subtraction cancels about nine digits.
#pragma clang fp contract(off)
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int64_t i = pr.row[p], j = pr.col[p];
  double fact[MNP_MAX_ORDER + 2];
  fact[0] = 1;
  for (int o = 1; o <= order + 1; ++o) fact[o] = fact[o - 1] * o;
  double g[MNP_MAX_ORDER + 1], f[3][MNP_MAX_ORDER + 1];
  for (int o = 0; o <= order; ++o) g[o] = f[0][o] = f[1][o] = f[2][o] = 0;
  double pi3[3], vj[3][3];
  for (int c = 0; c < 3; ++c) {
    pi3[c] = geo.pos[c * n + i];
    for (int a = 0; a < 3; ++a) vj[a][c] = geo.vtx[(a * 3 + c) * n + j];
  }
  if (pr.kind[p] == 2) {
    const double rmax = __longlong_as_double((long long)*rmax_bits);
    double nu[3], t1[3], t2[3];
    for (int c = 0; c < 3; ++c) {
      nu[c] = geo.nvec[c * n + i];
      t1[c] = geo.t1[c * n + i];
      t2[c] = geo.t2[c * n + i];
    }
    double fn[MNP_MAX_ORDER + 1], f1[MNP_MAX_ORDER + 1], f2[MNP_MAX_ORDER + 1];
    for (int o = 0; o <= order; ++o) fn[o] = f1[o] = f2[o] = 0;
    const double area = geo.area[j];
    for (int q = 0; q < npol; ++q) {
      const double x = c_qpol[0][q], y = c_qpol[1][q], z = 1 - x - y, w = c_qpol[2][q] * area;
      double vec[3];
      for (int c = 0; c < 3; ++c) vec[c] = pi3[c] - (x * vj[0][c] + y * vj[1][c] + z * vj[2][c]);
      const double r = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
      const double in = vec[0] * nu[0] + vec[1] * nu[1] + vec[2] * nu[2];
      const double in1 = vec[0] * t1[0] + vec[1] * t1[1] + vec[2] * t1[2];
      const double in2 = vec[0] * t2[0] + vec[1] * t2[1] + vec[2] * t2[2];
      const double rr = fmax(r, 1e-4 * rmax);
      const double rr3 = rr * rr * rr;
      double rp = 1.0 / r;                   // r^(o-1)
      double rm3 = 1.0 / (r * r * r);        // r^(o-3)
      double ro = 1.0;                       // r^o
      for (int o = 0; o <= order; ++o) {
        g[o] += w * rp;
        fn[o] += w * ((o - 1) * in * rm3);
        f1[o] += w * ((o - 1) * in1 * ro / rr3);
        f2[o] += w * ((o - 1) * in2 * ro / rr3);
        rp *= r;
        rm3 *= r;
        ro *= r;
      }
    }
    for (int o = 0; o <= order; ++o) {
      g[o] /= fact[o];
      const double a = fn[o] / fact[o], b = f1[o] / fact[o], c2 = f2[o] / fact[o];
      for (int c = 0; c < 3; ++c) f[c][o] = nu[c] * a + t1[c] * b + t2[c] * c2;
    }
  } else {
    near_taylor(pi3, geo, n, j, order, fact, &c_q18[0][0], g, f);
  }
  for (int o = 0; o <= order; ++o) {
    gt[o * nnz + p] = g[o];
    for (int c = 0; c < 3; ++c) ft[(o * 3 + c) * nnz + p] = f[c][o];
  }
}

// Closed-surface sum rule, one workgroup per source column i:
//   fsum_i = sum_row area_row F_static(row, i) / area_i
// with the refined static F = n_row . f_0 where (row, i) is refined
// (compgreenret/private/initclosed.m:85-119, k = 0).
// For a compound (fpart set) only the rows of the parts of column i's closed
// surface are summed (initclosed.m:34-47 with closedparticle.m).
__global__ void __launch_bounds__(256) k_closed_sum(int64_t n, Geo geo, Pairs pr, const int64_t* off,
                                                    int64_t nnz, const double* __restrict__ ft,
                                                    const int32_t* __restrict__ fpart,
                                                    const uint32_t* __restrict__ cmask, double* fsum) {
  extern __shared__ unsigned int bits[];
  const int64_t i = blockIdx.x;
  const uint32_t rows = fpart ? cmask[fpart[i]] : 0xffffffffu;
  auto summed = [&](int64_t r) { return !fpart || ((rows >> fpart[r]) & 1u); };
  const int64_t nw = (n + 31) / 32;
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) bits[w] = 0;
  __syncthreads();
  const int64_t p0 = off[i], p1 = off[i + 1];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int r = pr.row[p];
    atomicOr(&bits[r >> 5], 1u << (r & 31));
  }
  __syncthreads();
  const double xi = geo.pos[i], yi = geo.pos[n + i], zi = geo.pos[2 * n + i];
  const double ai = geo.area[i], inva = 1.0 / ai;
  double s = 0;
  for (int64_t r = threadIdx.x; r < n; r += blockDim.x) {
    if ((bits[r >> 5] & (1u << (r & 31))) || !summed(r)) continue;
    const double dx = geo.pos[r] - xi, dy = geo.pos[n + r] - yi, dz = geo.pos[2 * n + r] - zi;
    const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
    const double in = geo.nvec[r] * dx + geo.nvec[n + r] * dy + geo.nvec[2 * n + r] * dz;
    const double F = in * (-1.0 / d) / (d * d) * ai;
    s += geo.area[r] * (F * inva);
  }
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int64_t r = pr.row[p];
    if (!summed(r)) continue;
    const double F = geo.nvec[r] * ft[0 * nnz + p] + geo.nvec[n + r] * ft[1 * nnz + p] +
                     geo.nvec[2 * n + r] * ft[2 * nnz + p];
    s += geo.area[r] * (F * inva);
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) fsum[i] = red[0];
}

// f_ii,0 += n_i (-2 pi - fsum_i)  (initclosed.m:78, greenret/diag.m:146-147);
// faces of a compound part without a closed surface keep theirs.
__global__ void k_closed_apply(int64_t n, Geo geo, Pairs pr, int64_t nnz, const double* fsum,
                               const int32_t* __restrict__ fpart, const uint32_t* __restrict__ cmask, double* ft) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = pr.diag[i];
  if (p < 0 || (fpart && cmask[fpart[i]] == 0)) return;
  const double v = -2 * M_PI - fsum[i];
  for (int c = 0; c < 3; ++c) ft[c * nnz + p] += geo.nvec[c * n + i] * v;
}

// Dense G(k), F(k) +- 2 pi delta (eval1.m:24-40, :110-124, :140-143) for
// rows i < nrows and all columns j, column-major with leading dimension ld.  Each thread writes
// one element of up to two double-precision matrices and up to two
// single-precision mirrors (the mixed chain's inputs); NULL outputs are
// skipped.  Refined elements are overwritten by k_assemble_refined.
// With DK the kernel writes the derivatives with respect to k instead:
// dG/dk = i area_j e^{ikd}, dF/dk = -k (n_i . (pos_i - pos_j)) area_j e^{ikd} / d
// (the +-2 pi term does not depend on k).
template <class TG, bool DK = false>
__global__ void __launch_bounds__(256) k_assemble(int64_t n, int64_t nrows, Geo geo, double2 k, double sgn2pi,
                                                  TG outG, TG outH, int64_t ld) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= nrows) return;
  const double dx = geo.pos[i] - geo.pos[j];
  const double dy = geo.pos[n + i] - geo.pos[n + j];
  const double dz = geo.pos[2 * n + i] - geo.pos[2 * n + j];
  const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
  const double aj = geo.area[j];
  const double2 ph = cexp_ikd(k, d);
  const int64_t idx = j * ld + i;
  if constexpr (DK) {
    if (outG.any()) outG.store(idx, mk<double>(-aj * ph.y, aj * ph.x));
    if (outH.any()) {
      const double in = dx * geo.nvec[i] + dy * geo.nvec[n + i] + dz * geo.nvec[2 * n + i];
      const double2 kp = cmul<double>(k, ph);
      const double s = -in * aj / d;
      outH.store(idx, mk<double>(s * kp.x, s * kp.y));
    }
    return;
  }
  if (outG.any()) {
    const double gr = (1.0 / d) * aj;
    outG.store(idx, mk<double>(gr * ph.x, gr * ph.y));
  }
  if (outH.any()) {
    // f = (ik - 1/d)/d^2 ; F = (f x n1 + f y n2 + f z n3) area_j
    // (i k - 1/d)/d^2 with i k = (-k.y, k.x)
    const double id2 = 1.0 / (d * d);
    const double2 fk = mk<double>((-k.y - 1.0 / d) * id2, k.x * id2);
    const double in = dx * geo.nvec[i] + dy * geo.nvec[n + i] + dz * geo.nvec[2 * n + i];
    double2 F = mk<double>(fk.x * in * aj, fk.y * in * aj);
    F = cmul<double>(F, ph);
    if (i == j) F.x += sgn2pi;  // d == 0 only on the diagonal for distinct centroids
    outH.store(idx, F);
  }
}

// Overwrites refined elements in rows i < nrows: G = (sum_n g_n (ik)^n) e^{ikd},
// F = (n_i . sum_n f_n (ik)^n) e^{ikd}, H = F +- 2 pi on coincident pairs.
// With DK, the derivatives with respect to k: d/dk [P(ik) e^{ikd}] =
// (i P'(ik) + i d P(ik)) e^{ikd} for P = sum_n c_n z^n.
template <class TG, bool DK = false>
__global__ void __launch_bounds__(256) k_assemble_refined(int64_t nnz, int64_t n, int64_t nrows, Geo geo,
                                                          Pairs pr, int order, const double* __restrict__ gt,
                                                          const double* __restrict__ ft, double2 k,
                                                          double sgn2pi, TG outG, TG outH, int64_t ld) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int64_t i = pr.row[p], j = pr.col[p];
  if (i >= nrows) return;
  const double dx = geo.pos[i] - geo.pos[j];
  const double dy = geo.pos[n + i] - geo.pos[n + j];
  const double dz = geo.pos[2 * n + i] - geo.pos[2 * n + j];
  const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
  const double2 ph = cexp_ikd(k, d);
  const double2 ik = mk<double>(-k.y, k.x);
  double2 pw = mk<double>(1, 0), G = mk<double>(0, 0), F = mk<double>(0, 0);
  double2 dG = mk<double>(0, 0), dF = mk<double>(0, 0), pwm = mk<double>(0, 0);  // pwm = (ik)^(o-1)
  const double n0 = geo.nvec[i], n1 = geo.nvec[n + i], n2 = geo.nvec[2 * n + i];
  for (int o = 0; o <= order; ++o) {
    const double go = gt[o * nnz + p];
    const double fo = n0 * ft[(o * 3 + 0) * nnz + p] + n1 * ft[(o * 3 + 1) * nnz + p] +
                      n2 * ft[(o * 3 + 2) * nnz + p];
    G = cadd<double>(G, cscale<double>(pw, go));
    F = cadd<double>(F, cscale<double>(pw, fo));
    if constexpr (DK) {
      dG = cadd<double>(dG, cscale<double>(pwm, o * go));
      dF = cadd<double>(dF, cscale<double>(pwm, o * fo));
      pwm = pw;
    }
    pw = cmul<double>(pw, ik);
  }
  const int64_t idx = j * ld + i;
  if constexpr (DK) {
    // i P'(ik) + i d P(ik)
    const double2 i1 = mk<double>(0, 1), id = mk<double>(0, d);
    const double2 gk = cadd<double>(cmul<double>(i1, dG), cmul<double>(id, G));
    const double2 fk = cadd<double>(cmul<double>(i1, dF), cmul<double>(id, F));
    if (outG.any()) outG.store(idx, cmul<double>(gk, ph));
    if (outH.any()) outH.store(idx, cmul<double>(fk, ph));
    return;
  }
  if (outG.any()) outG.store(idx, cmul<double>(G, ph));
  if (outH.any()) {
    double2 H = cmul<double>(F, ph);
    if (pr.kind[p] == 2) H.x += sgn2pi;
    outH.store(idx, H);
  }
}

// Compound assembly of one side (eval1.m:16-51 through initmat.m:27-31):
// column j takes the wavenumber of its side's medium m = colmed[j]; row i
// enters with + when its same side faces m (g{s,s}), with - when its other
// side does (g{3-s,s}), else the element is zero.  The +-2 pi of H sits on
// the diagonal, which is always a same-side element.
__device__ __forceinline__ double conn_sign(int32_t m, int32_t same, int32_t other) {
  return same == m ? 1.0 : (other == m ? -1.0 : 0.0);
}
template <class TG>
__global__ void __launch_bounds__(256) k_assemble_conn(int64_t n, Geo geo, const int32_t* __restrict__ colmed,
                                                       const int32_t* __restrict__ same,
                                                       const int32_t* __restrict__ other, KTab kt, double sgn2pi,
                                                       TG outG, TG outH, int64_t ld) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= n) return;
  const int32_t m = colmed[j];
  const double sg = conn_sign(m, same[i], other[i]);
  const int64_t idx = j * ld + i;
  if (sg == 0) {
    if (outG.any()) outG.store(idx, mk<double>(0, 0));
    if (outH.any()) outH.store(idx, mk<double>(0, 0));
    return;
  }
  const double2 k = kt.k[m - 1];
  const double dx = geo.pos[i] - geo.pos[j];
  const double dy = geo.pos[n + i] - geo.pos[n + j];
  const double dz = geo.pos[2 * n + i] - geo.pos[2 * n + j];
  const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
  const double aj = geo.area[j];
  const double2 ph = cexp_ikd(k, d);
  if (outG.any()) {
    const double gr = sg * (1.0 / d) * aj;
    outG.store(idx, mk<double>(gr * ph.x, gr * ph.y));
  }
  if (outH.any()) {
    const double id2 = 1.0 / (d * d);
    const double2 fk = mk<double>((-k.y - 1.0 / d) * id2, k.x * id2);
    const double in = sg * (dx * geo.nvec[i] + dy * geo.nvec[n + i] + dz * geo.nvec[2 * n + i]);
    double2 F = cmul<double>(mk<double>(fk.x * in * aj, fk.y * in * aj), ph);
    if (i == j) F.x += sgn2pi;
    outH.store(idx, F);
  }
}
template <class TG>
__global__ void __launch_bounds__(256) k_assemble_refined_conn(int64_t nnz, int64_t n, Geo geo, Pairs pr, int order,
                                                               const double* __restrict__ gt,
                                                               const double* __restrict__ ft,
                                                               const int32_t* __restrict__ colmed,
                                                               const int32_t* __restrict__ same,
                                                               const int32_t* __restrict__ other, KTab kt,
                                                               double sgn2pi, TG outG, TG outH, int64_t ld) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int64_t i = pr.row[p], j = pr.col[p];
  const int32_t m = colmed[j];
  const double sg = conn_sign(m, same[i], other[i]);
  if (sg == 0) return;  // k_assemble_conn wrote the zero
  const double2 k = kt.k[m - 1];
  const double dx = geo.pos[i] - geo.pos[j];
  const double dy = geo.pos[n + i] - geo.pos[n + j];
  const double dz = geo.pos[2 * n + i] - geo.pos[2 * n + j];
  const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
  const double2 ph = cexp_ikd(k, d);
  const double2 ik = mk<double>(-k.y, k.x);
  double2 pw = mk<double>(1, 0), G = mk<double>(0, 0), F = mk<double>(0, 0);
  const double n0 = geo.nvec[i], n1 = geo.nvec[n + i], n2 = geo.nvec[2 * n + i];
  for (int o = 0; o <= order; ++o) {
    const double go = gt[o * nnz + p];
    const double fo = n0 * ft[(o * 3 + 0) * nnz + p] + n1 * ft[(o * 3 + 1) * nnz + p] + n2 * ft[(o * 3 + 2) * nnz + p];
    G = cadd<double>(G, cscale<double>(pw, go));
    F = cadd<double>(F, cscale<double>(pw, fo));
    pw = cmul<double>(pw, ik);
  }
  const int64_t idx = j * ld + i;
  if (outG.any()) outG.store(idx, cscale<double>(cmul<double>(G, ph), sg));
  if (outH.any()) {
    double2 H = cscale<double>(cmul<double>(F, ph), sg);
    if (pr.kind[p] == 2) H.x += sgn2pi;
    outH.store(idx, H);
  }
}

// ------------------------------------------------------------- host side
void GreenRet::assemble_side(int side, const KTab& kt, OutPair g, OutPair h, int64_t ld) const {
  MNP_CHECK(compound(), MNP_ERR_STATE, "greenret: assemble_side needs a compound");
  MNP_CHECK(side == 1 || side == 2, MNP_ERR_ARG, "greenret: side must be 1 or 2");
  MNP_CHECK(n <= 65535, MNP_ERR_ARG, "greenret: dense assembly supports at most 65535 faces");
  if (!g.any() && !h.any()) return;
  const int32_t* colmed = side == 1 ? fin : fout;
  const int32_t* same = side == 1 ? fin : fout;
  const int32_t* other = side == 1 ? fout : fin;
  const double sgn2pi = side == 1 ? 2 * M_PI : -2 * M_PI;
  const double bytes = (double)n * n * ((g.d ? 16 : 0) + (g.f ? 8 : 0) + (h.d ? 16 : 0) + (h.f ? 8 : 0));
  TimerScope ts(MNP_TIMER_ASSEMBLE, bytes, bytes);
  hipLaunchKernelGGL(k_assemble_conn<OutPair>, dim3((unsigned)cdiv_up(n, 256), (unsigned)n), dim3(256), 0, stream(),
                     n, geo, colmed, same, other, kt, sgn2pi, g, h, ld);
  MNP_HIP(hipGetLastError());
  if (nnz) {
    hipLaunchKernelGGL(k_assemble_refined_conn<OutPair>, dim3((unsigned)cdiv_up(nnz, 256)), dim3(256), 0, stream(),
                       nnz, n, geo, pairs, order, gtab, ftab, colmed, same, other, kt, sgn2pi, g, h, ld);
    MNP_HIP(hipGetLastError());
  }
}

void GreenRet::assemble(double2 k, double sgn2pi, OutPair g, OutPair h, int64_t ld) const {
  assemble_rows(k, sgn2pi, g, h, n, ld);
}

void GreenRet::assemble_dk(double2 k, OutPair g, OutPair h, int64_t ld) const {
  if (!g.any() && !h.any()) return;
  MNP_CHECK(n <= 65535, MNP_ERR_ARG, "greenret: dense assembly supports at most 65535 faces");
  hipLaunchKernelGGL((k_assemble<OutPair, true>), dim3((unsigned)cdiv_up(n, 256), (unsigned)n), dim3(256), 0,
                     stream(), n, n, geo, k, 0.0, g, h, ld);
  MNP_HIP(hipGetLastError());
  if (nnz) {
    hipLaunchKernelGGL((k_assemble_refined<OutPair, true>), dim3((unsigned)cdiv_up(nnz, 256)), dim3(256), 0,
                       stream(), nnz, n, n, geo, pairs, order, gtab, ftab, k, 0.0, g, h, ld);
    MNP_HIP(hipGetLastError());
  }
}

void GreenRet::assemble_rows(double2 k, double sgn2pi, OutPair g, OutPair h, int64_t nrows, int64_t ld) const {
  if (!g.any() && !h.any()) return;
  MNP_CHECK(nrows >= 1 && nrows <= n && ld >= nrows, MNP_ERR_ARG, "greenret: bad row range");
  MNP_CHECK(n <= 65535, MNP_ERR_ARG, "greenret: dense assembly supports at most 65535 faces");
  dim3 grid((unsigned)cdiv_up(nrows, 256), (unsigned)n);
  const double bytes = (double)nrows * n *
                       ((g.d ? 16 : 0) + (g.f ? 8 : 0) + (h.d ? 16 : 0) + (h.f ? 8 : 0));
  TimerScope ts(MNP_TIMER_ASSEMBLE, bytes, bytes);
  hipLaunchKernelGGL(k_assemble<OutPair>, grid, dim3(256), 0, stream(), n, nrows, geo, k, sgn2pi, g, h, ld);
  MNP_HIP(hipGetLastError());
  if (nnz) {
    hipLaunchKernelGGL(k_assemble_refined<OutPair>, dim3((unsigned)cdiv_up(nnz, 256)), dim3(256), 0,
                       stream(), nnz, n, nrows, geo, pairs, order, gtab, ftab, k, sgn2pi, g, h, ld);
    MNP_HIP(hipGetLastError());
  }
}

GreenRet::GreenRet(const double* verts, int64_t nv, const int64_t* faces, int64_t nf,
                   const mnp_green_options& op_, const GreenRet* like, const CompoundSpec* cs)
    : n(nf), op(op_) {
  MNP_CHECK(nf > 0 && nv >= 3, MNP_ERR_ARG, "greenret: empty mesh");
  MNP_CHECK(op.order >= 0 && op.order <= MNP_MAX_ORDER, MNP_ERR_ARG, "greenret: order out of range");
  MNP_CHECK(n < (int64_t)1 << 31, MNP_ERR_ARG, "greenret: too many faces");
  for (int64_t t = 0; t < 3 * nf; ++t)
    MNP_CHECK(faces[t] >= 0 && faces[t] < nv, MNP_ERR_ARG, "greenret: face index out of range");
  order = op.order;
  // quadrature tables (the polar rule depends on npol)
  {
    double q18[3][28];
    rule18(q18[0], q18[1], q18[2]);
    std::vector<double> qp(3 * MNP_MAX_POLAR, 0.0);
    npolq = polar_rule(op.npol_rho, op.npol_phi, &qp[0], &qp[MNP_MAX_POLAR], &qp[2 * MNP_MAX_POLAR]);
    MNP_CHECK(npolq > 0, MNP_ERR_ARG, "greenret: polar rule too large");
    MNP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_q18), q18, sizeof(q18)));
    MNP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_qpol), qp.data(), sizeof(double) * 3 * MNP_MAX_POLAR));
    this->q18 = dev_alloc<double>(3 * 28);
    this->qpol = dev_alloc<double>(3 * MNP_MAX_POLAR);
    MNP_HIP(hipMemcpy(this->q18, q18, sizeof(q18), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(this->qpol, qp.data(), sizeof(double) * 3 * MNP_MAX_POLAR, hipMemcpyHostToDevice));
  }
  nverts = nv;
  faces_h.assign(faces, faces + 3 * nf);
  // device copies of the mesh (column-major n x 3, as passed)
  double* dverts = dev_alloc<double>(nv * 3);
  int64_t* dfaces = dev_alloc<int64_t>(nf * 3);
  MNP_HIP(hipMemcpy(dverts, verts, sizeof(double) * nv * 3, hipMemcpyHostToDevice));
  MNP_HIP(hipMemcpy(dfaces, faces, sizeof(int64_t) * nf * 3, hipMemcpyHostToDevice));
  geo.pos = dev_alloc<double>(3 * n);
  geo.nvec = dev_alloc<double>(3 * n);
  geo.t1 = dev_alloc<double>(3 * n);
  geo.t2 = dev_alloc<double>(3 * n);
  geo.vtx = dev_alloc<double>(9 * n);
  geo.area = dev_alloc<double>(n);
  geo.qarea = dev_alloc<double>(n);
  geo.rad = dev_alloc<double>(n);
  hipStream_t s = stream();
  hipLaunchKernelGGL(k_geometry, dim3((unsigned)cdiv_up(n, 256)), dim3(256), 0, s, n, dverts, nv,
                     dfaces, geo);
  MNP_HIP(hipGetLastError());
  // refinement selection (or the pairs of `like`)
  int64_t* cnt = dev_alloc<int64_t>(n);
  off = dev_alloc<int64_t>(n + 1);
  if (like) {
    MNP_CHECK(like->n == n && like->faces_h == faces_h, MNP_ERR_ARG,
              "greenret: the template Green function has other faces");
    nnz = like->nnz;
    pairs.row = dev_alloc<int32_t>(nnz);
    pairs.col = dev_alloc<int32_t>(nnz);
    pairs.kind = dev_alloc<int8_t>(nnz);
    pairs.diag = dev_alloc<int64_t>(n);
    MNP_HIP(hipMemcpyAsync(off, like->off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s));
    MNP_HIP(hipMemcpyAsync(pairs.row, like->pairs.row, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, s));
    MNP_HIP(hipMemcpyAsync(pairs.col, like->pairs.col, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, s));
    MNP_HIP(hipMemcpyAsync(pairs.kind, like->pairs.kind, sizeof(int8_t) * nnz, hipMemcpyDeviceToDevice, s));
    MNP_HIP(hipMemcpyAsync(pairs.diag, like->pairs.diag, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
  } else {
    hipLaunchKernelGGL(k_refine_count, dim3((unsigned)n), dim3(256), 0, s, n, geo, op.rel_cutoff,
                       op.abs_cutoff, cnt);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, n, cnt, off);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipMemcpyAsync(&nnz, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MNP_HIP(hipStreamSynchronize(s));
    pairs.row = dev_alloc<int32_t>(nnz);
    pairs.col = dev_alloc<int32_t>(nnz);
    pairs.kind = dev_alloc<int8_t>(nnz);
    pairs.diag = dev_alloc<int64_t>(n);
    MNP_HIP(hipMemsetAsync(pairs.diag, 0xff, sizeof(int64_t) * n, s));
    hipLaunchKernelGGL(k_refine_fill, dim3((unsigned)n), dim3(256), 0, s, n, geo, op.rel_cutoff,
                       op.abs_cutoff, off, pairs);
    MNP_HIP(hipGetLastError());
  }
  if (cs) {
    const int np = cs->nparts;
    MNP_CHECK(np >= 1 && np <= MNP_MAX_PARTS && cs->part_off && cs->inout && cs->cmask, MNP_ERR_ARG,
              "greenret: compound needs 1..32 parts");
    MNP_CHECK(cs->part_off[0] == 0 && cs->part_off[np] == nf, MNP_ERR_ARG, "greenret: part offsets must cover the faces");
    nparts = np;
    fin_h.resize(nf);
    fout_h.resize(nf);
    std::vector<int32_t> fp(nf);
    for (int q = 0; q < np; ++q) {
      MNP_CHECK(cs->part_off[q] <= cs->part_off[q + 1], MNP_ERR_ARG, "greenret: part offsets must ascend");
      const int32_t a = cs->inout[2 * q], b = cs->inout[2 * q + 1];
      MNP_CHECK(a >= 1 && b >= 1 && a <= MNP_MAX_MED && b <= MNP_MAX_MED && a != b, MNP_ERR_ARG,
                "greenret: media must be distinct indices in 1..16");
      nmed = std::max(nmed, (int)std::max(a, b));
      for (int64_t i = cs->part_off[q]; i < cs->part_off[q + 1]; ++i) {
        fin_h[i] = a;
        fout_h[i] = b;
        fp[i] = q;
      }
      for (int r = 0; r < np; ++r) con12 = con12 || a == cs->inout[2 * r + 1];
    }
    fin = dev_alloc<int32_t>(nf);
    fout = dev_alloc<int32_t>(nf);
    fpart = dev_alloc<int32_t>(nf);
    cmask = dev_alloc<uint32_t>(np);
    MNP_HIP(hipMemcpy(fin, fin_h.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(fout, fout_h.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(fpart, fp.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(cmask, cs->cmask, sizeof(uint32_t) * np, hipMemcpyHostToDevice));
  }
  // refinement tables
  gtab = dev_alloc<double>(nnz * (order + 1));
  ftab = dev_alloc<double>(nnz * 3 * (order + 1));
  unsigned long long* rmax = dev_alloc<unsigned long long>(1);
  MNP_HIP(hipMemsetAsync(rmax, 0, sizeof(unsigned long long), s));
  if (nnz) {
    const unsigned nb = (unsigned)cdiv_up(nnz, 256);
    hipLaunchKernelGGL(k_self_rmax, dim3(nb), dim3(256), 0, s, nnz, geo, n, pairs, npolq, rmax);
    hipLaunchKernelGGL(k_tables, dim3(nb), dim3(256), 0, s, nnz, geo, n, pairs, order, npolq, rmax,
                       gtab, ftab);
    MNP_HIP(hipGetLastError());
  }
  if (op.closed && nnz) {
    double* fsum = dev_alloc<double>(n);
    const size_t lds = sizeof(unsigned int) * (size_t)((n + 31) / 32);
    MNP_CHECK(lds <= 64 * 1024, MNP_ERR_ARG, "greenret: mesh too large for the sum-rule kernel");
    hipLaunchKernelGGL(k_closed_sum, dim3((unsigned)n), dim3(256), lds, s, n, geo, pairs, off, nnz,
                       ftab, (const int32_t*)fpart, (const uint32_t*)cmask, fsum);
    hipLaunchKernelGGL(k_closed_apply, dim3((unsigned)cdiv_up(n, 256)), dim3(256), 0, s, n, geo,
                       pairs, nnz, fsum, (const int32_t*)fpart, (const uint32_t*)cmask, ftab);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(s));
    MNP_HIP(hipFree(fsum));
  }
  MNP_HIP(hipStreamSynchronize(s));
  MNP_HIP(hipFree(rmax));
  MNP_HIP(hipFree(cnt));
  MNP_HIP(hipFree(dverts));
  MNP_HIP(hipFree(dfaces));
}

GreenRet::~GreenRet() {
  void* ptrs[] = {geo.pos, geo.nvec, geo.t1, geo.t2, geo.vtx,   geo.area, geo.qarea, geo.rad, off, pairs.row,
                  pairs.col, pairs.kind, pairs.diag, gtab, ftab, q18, qpol, fin, fout, fpart, cmask};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

extern "C" {

const char* mnp_last_error(void) { return mnp::g_last_error.c_str(); }
int mnp_version(void) { return 100; }

int mnp_set_device(int device) {
  return guarded([&] { MNP_HIP(hipSetDevice(device)); });
}
int mnp_set_stream(void* s) {
  return guarded([&] { mnp::g_stream = (hipStream_t)s; });
}
void* mnp_get_stream(void) { return (void*)mnp::g_stream; }
int mnp_timers_enable(int on) {
  return guarded([&] {
    mnp::timers_flush();
    mnp::g_timers_on = on != 0;
    mnp::g_timer_acc = mnp_timers{};
  });
}
int mnp_timers_read(mnp_timers* out) {
  return guarded([&] {
    MNP_CHECK(out, MNP_ERR_ARG, "timers_read: NULL argument");
    mnp::timers_flush();
    *out = mnp::g_timer_acc;
  });
}
int mnp_synchronize(void) {
  return guarded([&] { MNP_HIP(hipStreamSynchronize(mnp::g_stream)); });
}

void mnp_green_options_default(mnp_green_options* op) {
  if (!op) return;
  op->order = 5;
  op->rel_cutoff = 3.0;
  op->abs_cutoff = 0.0;
  op->npol_rho = 7;
  op->npol_phi = 5;
  op->closed = 1;
}

int mnp_greenret_create(const double* verts, int64_t nverts, const int64_t* faces, int64_t nfaces,
                        const mnp_green_options* op, mnp_greenret_t* out) {
  return guarded([&] {
    MNP_CHECK(verts && faces && out, MNP_ERR_ARG, "greenret_create: NULL argument");
    mnp_green_options o;
    if (op)
      o = *op;
    else
      mnp_green_options_default(&o);
    auto* h = new mnp_greenret_s{nullptr};
    try {
      h->g = new GreenRet(verts, nverts, faces, nfaces, o);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_greenret_create_compound(const double* verts, int64_t nverts, const int64_t* faces, int64_t nfaces,
                                 int32_t nparts, const int64_t* part_offsets, const int32_t* inout,
                                 const uint32_t* closed_mask, const mnp_green_options* op, mnp_greenret_t* out) {
  return guarded([&] {
    MNP_CHECK(verts && faces && out && part_offsets && inout && closed_mask, MNP_ERR_ARG,
              "greenret_create_compound: NULL argument");
    mnp_green_options o;
    if (op)
      o = *op;
    else
      mnp_green_options_default(&o);
    // inout arrives column-major (nparts x 2); the spec takes (in, out) pairs
    std::vector<int32_t> io(2 * (size_t)std::max(nparts, 0));
    for (int32_t q = 0; q < nparts; ++q) {
      io[2 * q] = inout[q];
      io[2 * q + 1] = inout[nparts + q];
    }
    bool any_closed = false;
    for (int32_t q = 0; q < nparts; ++q) any_closed = any_closed || closed_mask[q] != 0;
    o.closed = any_closed ? 1 : 0;
    const CompoundSpec cs{nparts, part_offsets, io.data(), closed_mask};
    auto* h = new mnp_greenret_s{nullptr};
    try {
      h->g = new GreenRet(verts, nverts, faces, nfaces, o, nullptr, &cs);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_greenret_create_like(const double* verts, int64_t nverts, const int64_t* faces, int64_t nfaces,
                             const mnp_green_options* op, mnp_greenret_t like, mnp_greenret_t* out) {
  return guarded([&] {
    MNP_CHECK(verts && faces && out && like, MNP_ERR_ARG, "greenret_create_like: NULL argument");
    mnp_green_options o;
    if (op)
      o = *op;
    else
      mnp_green_options_default(&o);
    auto* h = new mnp_greenret_s{nullptr};
    try {
      h->g = new GreenRet(verts, nverts, faces, nfaces, o, like->g);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_greenret_destroy(mnp_greenret_t g) {
  return guarded([&] {
    if (!g) return;
    delete g->g;
    delete g;
  });
}

int mnp_greenret_size(mnp_greenret_t g, int64_t* nfaces, int64_t* nrefined) {
  return guarded([&] {
    MNP_CHECK(g, MNP_ERR_ARG, "greenret_size: NULL handle");
    if (nfaces) *nfaces = g->g->n;
    if (nrefined) *nrefined = g->g->nnz;
  });
}

int mnp_greenret_geometry(mnp_greenret_t g, double* pos, double* nvec, double* tvec1, double* tvec2,
                          double* area) {
  return guarded([&] {
    MNP_CHECK(g, MNP_ERR_ARG, "greenret_geometry: NULL handle");
    const GreenRet& G = *g->g;
    MNP_HIP(hipStreamSynchronize(stream()));
    const size_t b3 = sizeof(double) * 3 * G.n;
    if (pos) MNP_HIP(hipMemcpy(pos, G.geo.pos, b3, hipMemcpyDeviceToHost));
    if (nvec) MNP_HIP(hipMemcpy(nvec, G.geo.nvec, b3, hipMemcpyDeviceToHost));
    if (tvec1) MNP_HIP(hipMemcpy(tvec1, G.geo.t1, b3, hipMemcpyDeviceToHost));
    if (tvec2) MNP_HIP(hipMemcpy(tvec2, G.geo.t2, b3, hipMemcpyDeviceToHost));
    if (area) MNP_HIP(hipMemcpy(area, G.geo.area, b3 / 3, hipMemcpyDeviceToHost));
  });
}

int mnp_greenret_refined(mnp_greenret_t g, int64_t* rows, int64_t* cols, double* gtab, double* ftab) {
  return guarded([&] {
    MNP_CHECK(g, MNP_ERR_ARG, "greenret_refined: NULL handle");
    const GreenRet& G = *g->g;
    MNP_HIP(hipStreamSynchronize(stream()));
    const int64_t nnz = G.nnz;
    if (rows || cols) {
      std::vector<int32_t> tmp(nnz);
      if (rows) {
        MNP_HIP(hipMemcpy(tmp.data(), G.pairs.row, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < nnz; ++p) rows[p] = tmp[p];
      }
      if (cols) {
        MNP_HIP(hipMemcpy(tmp.data(), G.pairs.col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < nnz; ++p) cols[p] = tmp[p];
      }
    }
    if (gtab) MNP_HIP(hipMemcpy(gtab, G.gtab, sizeof(double) * nnz * (G.order + 1), hipMemcpyDeviceToHost));
    if (ftab)
      MNP_HIP(hipMemcpy(ftab, G.ftab, sizeof(double) * nnz * 3 * (G.order + 1), hipMemcpyDeviceToHost));
  });
}

int mnp_greenret_eval(mnp_greenret_t g, mnp_complex k, int key, int precision, void* out, int64_t ld) {
  return guarded([&] {
    MNP_CHECK(g && out, MNP_ERR_ARG, "greenret_eval: NULL argument");
    const GreenRet& G = *g->g;
    MNP_CHECK(ld >= G.n, MNP_ERR_ARG, "greenret_eval: ld < n");
    MNP_CHECK(key >= MNP_KEY_G && key <= MNP_KEY_H2, MNP_ERR_ARG, "greenret_eval: bad key");
    MNP_CHECK(precision == MNP_FP64 || precision == MNP_FP32, MNP_ERR_ARG, "greenret_eval: bad precision");
    OutPair og{}, oh{};
    OutPair& o = key == MNP_KEY_G ? og : oh;
    if (precision == MNP_FP64)
      o.d = (double2*)out;
    else
      o.f = (float2*)out;
    const double sgn = key == MNP_KEY_H1 ? 2 * M_PI : key == MNP_KEY_H2 ? -2 * M_PI : 0.0;
    G.assemble(to_d2(k), sgn, og, oh, ld);
  });
}

}  // extern "C"
