// Flow-CG's restricted-master QP on the host, native: min 1/2 lam'Q lam - h'lam over the simplex
// (lam >= 0, sum lam = 1; Q symmetric), by the reference's active set (solver/common/qp_native.py:37-100):
//   on the working set W: lam_W = Q_WW^-1 (h_W - mu 1), mu = (1'Q_WW^-1 h_W - 1) / (1'Q_WW^-1 1)
//   (Q_WW + reg I, Cholesky); drop the most negative weight, else add the inactive column with
//   the most negative gradient (below -tol), else stop; then a Euclidean simplex projection and
//   support = weights above 1e-9.
// The Cholesky factor of the working set is extended by one row when a column enters and downdated
// when one leaves (both O(k^2)). Where it breaks down (a non-positive pivot of a nearly dependent set),
// that step takes the minimum-norm least-squares solution from a symmetric eigendecomposition
// (cyclic Jacobi), as the reference falls back to numpy's lstsq.
#include "rumgpu_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// Host threads for the gradient of the inactive columns, the one loop of a master solve that grows
// with K x k. The columns are cut into contiguous parts and every g_j is computed whole by one
// thread, in the same order as the single-threaded loop, so the results do not depend on the
// thread count. Threads: OMP_NUM_THREADS if set, else the hardware's, at most 16; created at
// the first large gradient and kept until the process ends. A worker polls for the next run for
// about a millisecond before it sleeps (a column generation's masters come a fraction of a
// millisecond apart, and waking a sleeping thread costs about as much as one gradient).
// set on a thread while it runs one part of a pool run that is itself a whole solve (a batch of
// masters): that solve's gradient then stays on its thread instead of starting a run of its own
thread_local bool tl_in_pool_run = false;

class HostPool {
 public:
  static HostPool& get() {
    static HostPool pool(threads_wanted());
    return pool;
  }
  int size() const { return (int)workers_.size() + 1; }
  // f(part) for every part in [0, parts), on the pool's threads and the caller's
  void run(int parts, const std::function<void(int)>& f) {
    std::lock_guard<std::mutex> one(run_mu_);
    std::unique_lock<std::mutex> lk(mu_);
    fn_ = &f;
    parts_ = parts;
    next_ = 0;
    remaining_ = parts;
    gen_.fetch_add(1, std::memory_order_release);
    cv_work_.notify_all();
    drain(lk);
    cv_done_.wait(lk, [&] { return remaining_ == 0; });
    fn_ = nullptr;
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& w : workers_) w.join();
  }

 private:
  explicit HostPool(int nt) {
    for (int t = 1; t < nt; ++t) workers_.emplace_back([this] { loop(); });
  }
  static int threads_wanted() {
    int nt = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
      const int v = std::atoi(e);
      if (v > 0) nt = v;
    }
    return std::max(1, std::min(nt, 16));
  }
  void loop() {
    unsigned long long seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      if (!stop_ && gen_.load(std::memory_order_relaxed) == seen) {
        lk.unlock();
        const auto t0 = std::chrono::steady_clock::now();
        for (int spin = 0; gen_.load(std::memory_order_acquire) == seen; ++spin) {
          if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(1000)) break;
          std::this_thread::yield();
        }
        lk.lock();
      }
      cv_work_.wait(lk, [&] { return stop_ || gen_.load(std::memory_order_relaxed) != seen; });
      if (stop_) return;
      seen = gen_.load(std::memory_order_relaxed);
      drain(lk);
    }
  }
  // take parts of the current run until none is left (lock held on entry and exit)
  void drain(std::unique_lock<std::mutex>& lk) {
    while (next_ < parts_) {
      const int p = next_++;
      lk.unlock();
      (*fn_)(p);
      lk.lock();
      if (--remaining_ == 0) cv_done_.notify_all();
    }
  }
  std::mutex run_mu_, mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::thread> workers_;
  const std::function<void(int)>* fn_ = nullptr;
  int parts_ = 0, next_ = 0, remaining_ = 0;
  std::atomic<unsigned long long> gen_{0};
  bool stop_ = false;
};

// g_j for the inactive columns j in [j0, j1): mu - h_j + sum_i Q[j, work_i] lam_i, by whole rows of
// Q (symmetric), whichever are fewer: the inactive rows (a dot product with lam spread over all K
// columns, in four partial sums) or the k working rows (axpys into every g_j, in working-set order)
template <typename TQ>
void grad_range(int j0, int j1, int K, int k, const TQ* Q, int64_t ldq, const double* h, const char* in,
                const int* work, const double* lam_w, const double* lam_full, double mu, double* grad) {
  if (K - k < k) {
    for (int j = j0; j < j1; ++j) {
      if (in[j]) continue;
      const TQ* qj = Q + (int64_t)j * ldq;
      double pp[4] = {0.0, 0.0, 0.0, 0.0};
      int t = 0;
      for (; t + 4 <= K; t += 4)
        for (int u = 0; u < 4; ++u) pp[u] += (double)qj[t + u] * lam_full[t + u];
      for (int u = 0; t < K; ++t, ++u) pp[u] += (double)qj[t] * lam_full[t];
      grad[j] = (mu - h[j]) + ((pp[0] + pp[1]) + (pp[2] + pp[3]));
    }
  } else {
    double* __restrict__ g = grad;
    for (int j = j0; j < j1; ++j) g[j] = mu - h[j];
    // four working rows per sweep over g: each g_j still takes its products in working-set order,
    // one rounded add at a time, but is loaded and stored once per four rows
    int i = 0;
    for (; i + 4 <= k; i += 4) {
      const TQ* __restrict__ q0 = Q + (int64_t)work[i] * ldq;
      const TQ* __restrict__ q1 = Q + (int64_t)work[i + 1] * ldq;
      const TQ* __restrict__ q2 = Q + (int64_t)work[i + 2] * ldq;
      const TQ* __restrict__ q3 = Q + (int64_t)work[i + 3] * ldq;
      const double l0 = lam_w[i], l1 = lam_w[i + 1], l2 = lam_w[i + 2], l3 = lam_w[i + 3];
      for (int j = j0; j < j1; ++j) {
        double t = g[j];
        t += (double)q0[j] * l0;
        t += (double)q1[j] * l1;
        t += (double)q2[j] * l2;
        t += (double)q3[j] * l3;
        g[j] = t;
      }
    }
    for (; i < k; ++i) {
      const TQ* __restrict__ qi = Q + (int64_t)work[i] * ldq;
      const double li = lam_w[i];
      for (int j = j0; j < j1; ++j) g[j] += (double)qi[j] * li;
    }
  }
}

struct Chol {
  int k = 0;
  std::vector<double> L;   // row-major lower factor, row stride cap (grown by doubling, plus 8 so that
                           // the stride is no power of two and rows do not alias in the caches)
  int cap = 0;
  bool ok = false;
  void reserve(int need) {
    if (need <= cap) return;
    cap = std::max(need, 2 * cap) + 8;
    L.assign((size_t)cap * cap, 0.0);
    ok = false;
  }
  // room for `need` rows keeping the factor: its k rows move to the wider stride (O(k^2))
  void grow(int need) {
    if (need <= cap) return;
    const int nc = std::max(need, 2 * cap) + 8;
    std::vector<double> nl((size_t)nc * nc, 0.0);
    for (int i = 0; i < k; ++i)
      std::copy(L.begin() + (size_t)i * cap, L.begin() + (size_t)i * cap + i + 1, nl.begin() + (size_t)i * nc);
    L.swap(nl);
    cap = nc;
  }
};

// L L' = A (k x k, row-major in a, leading dimension k); false on a non-positive pivot
bool chol_full(Chol& c, const std::vector<double>& a, int k) {
  c.reserve(k);
  c.k = k;
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = a[(size_t)i * k + j];
      for (int t = 0; t < j; ++t) s -= c.L[(size_t)i * c.cap + t] * c.L[(size_t)j * c.cap + t];
      if (i == j) {
        if (!(s > 0.0)) return c.ok = false;
        c.L[(size_t)i * c.cap + i] = std::sqrt(s);
      } else {
        c.L[(size_t)i * c.cap + j] = s / c.L[(size_t)j * c.cap + j];
      }
    }
  }
  return c.ok = true;
}

// append one row: q = A(new, old) (k values), d0 = A(new, new)
bool chol_append(Chol& c, const double* q, double d0) {
  const int k = c.k;
  double* row = &c.L[(size_t)k * c.cap];
  double s = d0;
  for (int j = 0; j < k; ++j) {
    const double* lj = &c.L[(size_t)j * c.cap];
    double pp[4] = {0.0, 0.0, 0.0, 0.0};    // the dot product in four partial sums
    int t = 0;
    for (; t + 4 <= j; t += 4)
      for (int u = 0; u < 4; ++u) pp[u] += row[t + u] * lj[t + u];
    for (int u = 0; t < j; ++t, ++u) pp[u] += row[t] * lj[t];
    const double v = (q[j] - ((pp[0] + pp[1]) + (pp[2] + pp[3]))) / lj[j];
    row[j] = v;
    s -= v * v;
  }
  if (!(s > 0.0)) return c.ok = false;
  row[k] = std::sqrt(s);
  c.k = k + 1;
  return true;
}

// drop row and column i of the factored matrix: the trailing block takes the rank-1 update by
// the dropped column (the standard stable rotation form, rotation p acting on column p of the rows
// below it), then rows and columns after i move up. The rotations run row by row, so every access
// walks a row; each element still takes rotations 0, 1, ... in order.
void chol_delete(Chol& c, int i) {
  const int k = c.k, m = k - i - 1, cap = c.cap;
  std::vector<double> cs(m > 0 ? m : 0), sn(m > 0 ? m : 0);
  for (int q = 0; q < m; ++q) {
    double* row = &c.L[(size_t)(i + 1 + q) * cap + (i + 1)];
    double xq = row[-1];
    for (int p = 0; p < q; ++p) {
      row[p] = (row[p] + sn[p] * xq) / cs[p];
      xq = cs[p] * xq - sn[p] * row[p];
    }
    const double d = row[q], r = std::sqrt(d * d + xq * xq);
    cs[q] = r / d, sn[q] = xq / d;
    row[q] = r;
  }
  for (int r = i; r < k - 1; ++r) {
    double* dst = &c.L[(size_t)r * cap];
    const double* src = &c.L[(size_t)(r + 1) * cap];
    for (int col = 0; col < i; ++col) dst[col] = src[col];
    for (int col = i; col <= r; ++col) dst[col] = src[col + 1];
  }
  c.k = k - 1;
}

// Solves with the working set's factor for two right-hand sides at once (h_W and 1).
// Forward substitution L y = b by rows, row i's dot product in four partial sums per right-hand
// side (eight independent chains). Rows 0..i-1 of L do not change when a row is appended, so the
// y of a grown factor is the old y plus one fwd_row step.
void fwd_row(const Chol& c, int i, const double* __restrict__ b1, const double* __restrict__ b2,
             double* __restrict__ y1, double* __restrict__ y2) {
  const double* __restrict__ row = &c.L[(size_t)i * c.cap];
  double p1[4] = {0.0, 0.0, 0.0, 0.0}, p2[4] = {0.0, 0.0, 0.0, 0.0};
  int t = 0;
  for (; t + 4 <= i; t += 4)
    for (int u = 0; u < 4; ++u) {
      p1[u] += row[t + u] * y1[t + u];
      p2[u] += row[t + u] * y2[t + u];
    }
  for (int u = 0; t < i; ++t, ++u) {
    p1[u] += row[t] * y1[t];
    p2[u] += row[t] * y2[t];
  }
  y1[i] = (b1[i] - ((p1[0] + p1[1]) + (p1[2] + p1[3]))) / row[i];
  y2[i] = (b2[i] - ((p2[0] + p2[1]) + (p2[2] + p2[3]))) / row[i];
}

// L' x = y by rows of L: x_i is final once the rows below have been applied
void bwd2(const Chol& c, const double* __restrict__ y1, const double* __restrict__ y2, double* __restrict__ x1,
          double* __restrict__ x2) {
  const int k = c.k;
  for (int i = 0; i < k; ++i) x1[i] = y1[i], x2[i] = y2[i];
  for (int i = k - 1; i >= 0; --i) {
    const double* __restrict__ row = &c.L[(size_t)i * c.cap];
    const double xi1 = x1[i] / row[i], xi2 = x2[i] / row[i];
    x1[i] = xi1;
    x2[i] = xi2;
    for (int t = 0; t < i; ++t) {
      x1[t] -= row[t] * xi1;
      x2[t] -= row[t] * xi2;
    }
  }
}

// minimum-norm least squares A x = b for symmetric A (k x k) through cyclic Jacobi eigenpairs;
// singular values below eps * k * max |eigenvalue| count as zero (numpy lstsq's rcond=None)
void pinv_solve(std::vector<double> a, int k, const double* b1, const double* b2, double* x1, double* x2) {
  std::vector<double> v((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) v[(size_t)i * k + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) off += a[(size_t)p * k + q] * a[(size_t)p * k + q];
    if (off < 1e-30) break;
    for (int p = 0; p < k; ++p) {
      for (int q = p + 1; q < k; ++q) {
        const double apq = a[(size_t)p * k + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (a[(size_t)q * k + q] - a[(size_t)p * k + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int r = 0; r < k; ++r) {
          const double arp = a[(size_t)r * k + p], arq = a[(size_t)r * k + q];
          a[(size_t)r * k + p] = cs * arp - sn * arq;
          a[(size_t)r * k + q] = sn * arp + cs * arq;
        }
        for (int r = 0; r < k; ++r) {
          const double apr = a[(size_t)p * k + r], aqr = a[(size_t)q * k + r];
          a[(size_t)p * k + r] = cs * apr - sn * aqr;
          a[(size_t)q * k + r] = sn * apr + cs * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const double vrp = v[(size_t)r * k + p], vrq = v[(size_t)r * k + q];
          v[(size_t)r * k + p] = cs * vrp - sn * vrq;
          v[(size_t)r * k + q] = sn * vrp + cs * vrq;
        }
      }
    }
  }
  double emax = 0.0;
  for (int i = 0; i < k; ++i) emax = std::max(emax, std::fabs(a[(size_t)i * k + i]));
  const double cut = 2.220446049250313e-16 * k * emax;
  for (int i = 0; i < k; ++i) x1[i] = x2[i] = 0.0;
  for (int e = 0; e < k; ++e) {
    const double lam = a[(size_t)e * k + e];
    if (std::fabs(lam) <= cut) continue;
    double c1 = 0.0, c2 = 0.0;
    for (int r = 0; r < k; ++r) {
      c1 += v[(size_t)r * k + e] * b1[r];
      c2 += v[(size_t)r * k + e] * b2[r];
    }
    c1 /= lam, c2 /= lam;
    for (int r = 0; r < k; ++r) {
      x1[r] += c1 * v[(size_t)r * k + e];
      x2[r] += c2 * v[(size_t)r * k + e];
    }
  }
}

// Euclidean projection onto the simplex by sorting (qp_native.py:20-33)
void project_simplex(std::vector<double>& lam) {
  const int K = (int)lam.size();
  std::vector<double> u(lam);
  std::sort(u.begin(), u.end(), [](double a, double b) { return a > b; });
  double cs = 0.0, thr = 0.0;
  int r = -1;
  double run = 0.0;
  for (int i = 0; i < K; ++i) {
    run += u[i];
    const double c = run - 1.0;
    if (u[i] - c / (double)(i + 1) > 0.0) r = i, cs = c;
  }
  if (r < 0) {
    for (auto& x : lam) x = 1.0 / K;
    return;
  }
  thr = cs / (double)(r + 1);
  for (auto& x : lam) x = std::max(x - thr, 0.0);
}

}  // namespace

// A column generation's master QPs in sequence: the last working set and its factor. Q only grows
// (new columns append rows and columns; old entries stay), so a call that starts from that
// working set reuses the factor instead of refactoring it.
struct rumgpu_qp {
  std::vector<int> work;
  Chol ch;
  // the last solve on that working set and factor: its right-hand side h_W, Q_WW^-1 (h_W, 1), and
  // the forward solutions L^-1 h_W and L^-1 1
  std::vector<double> bh, qh, qe, yh, ye;
  // Q in float32 for the gradient's row reads (half the bytes). A column generation's Q holds
  // integer overlap counts, which float32 holds exactly, so the products are the same doubles;
  // the copy is off for good once an entry is not a float32. Rows and columns 0..k32-1 are copied
  // from the Q at q_src (ldq q_ld); a call with another Q, or whose last copied row differs from
  // the copy, copies it all again. New rows are copied whole; the new columns of the old rows are
  // taken from the new rows (Q is symmetric), so every read of Q is a contiguous row.
  std::vector<float> q32;
  int64_t ld32 = 0;
  int k32 = 0;
  const double* q_src = nullptr;
  int64_t q_ld = 0;
  bool q32_ok = true;
};

// the state's float32 copy of Q brought up to K x K (null when Q is not all float32)
static const float* sync_q32(rumgpu_qp* s, int K, const double* Q, int64_t ldq) {
  if (!s->q32_ok) return nullptr;
  int from = (s->q_src == Q && s->q_ld == ldq && s->k32 <= K) ? s->k32 : 0;
  if (from > 0) {
    const float* cr = s->q32.data() + (int64_t)(from - 1) * s->ld32;
    const double* qr = Q + (int64_t)(from - 1) * ldq;
    for (int j = 0; j < from; ++j)
      if ((double)cr[j] != qr[j]) {
        from = 0;
        break;
      }
  }
  if (K > s->ld32) {
    const int64_t nl = std::max<int64_t>(K, 2 * s->ld32) + 16;
    std::vector<float> nq((size_t)(nl * nl));
    for (int i = 0; i < from; ++i)
      std::copy(s->q32.begin() + (size_t)(i * s->ld32), s->q32.begin() + (size_t)(i * s->ld32 + from),
                nq.begin() + (size_t)(i * nl));
    s->q32.swap(nq);
    s->ld32 = nl;
  }
  float* c = s->q32.data();
  bool exact = true;
  for (int i = from; i < K; ++i) {
    const double* qi = Q + (int64_t)i * ldq;
    float* ci = c + (int64_t)i * s->ld32;
    for (int j = 0; j < K; ++j) {
      const float f = (float)qi[j];
      exact &= (double)f == qi[j];
      ci[j] = f;
    }
  }
  for (int i = 0; i < from; ++i) {
    float* ci = c + (int64_t)i * s->ld32;
    for (int j = from; j < K; ++j) ci[j] = c[(int64_t)j * s->ld32 + i];
  }
  s->q_src = Q;
  s->q_ld = ldq;
  s->k32 = K;
  if (!exact) {
    s->q32_ok = false;
    std::vector<float>().swap(s->q32);
    s->ld32 = s->k32 = 0;
    return nullptr;
  }
  return c;
}

extern "C" int rumgpu_qp_create(rumgpu_qp** out) {
  RG_CHECK(out, RG_EINVAL, "qp_create: null argument");
  *out = new rumgpu_qp();
  return RG_OK;
}

extern "C" int rumgpu_qp_destroy(rumgpu_qp* q) {
  delete q;
  return RG_OK;
}

static int qp_simplex_impl(int K, const double* Q, int64_t ldq, const double* h, const int32_t* A_init, int nA,
                           int max_outer, double tol, double reg, double* lam_out, int32_t* support, int* n_support,
                           rumgpu_qp* state);

// below this many multiply-adds a gradient stays on the calling thread (a pool round costs ~10 us)
static constexpr double kParallelFma = 131072.0;

extern "C" int rumgpu_qp_simplex(int K, const double* Q, int64_t ldq, const double* h, const int32_t* A_init, int nA,
                                 int max_outer, double tol, double reg, double* lam_out, int32_t* support,
                                 int* n_support) {
  return qp_simplex_impl(K, Q, ldq, h, A_init, nA, max_outer, tol, reg, lam_out, support, n_support, nullptr);
}

extern "C" int rumgpu_qp_simplex_warm(rumgpu_qp* state, int K, const double* Q, int64_t ldq, const double* h,
                                      const int32_t* A_init, int nA, int max_outer, double tol, double reg,
                                      double* lam_out, int32_t* support, int* n_support) {
  RG_CHECK(state, RG_EINVAL, "qp_simplex_warm: null state");
  return qp_simplex_impl(K, Q, ldq, h, A_init, nA, max_outer, tol, reg, lam_out, support, n_support, state);
}

static int qp_simplex_impl(int K, const double* Q, int64_t ldq, const double* h, const int32_t* A_init, int nA,
                           int max_outer, double tol, double reg, double* lam_out, int32_t* support, int* n_support,
                           rumgpu_qp* state) {
  RG_CHECK(K >= 1 && Q && h && lam_out && support && n_support && ldq >= K, RG_EINVAL, "qp_simplex: bad argument");
  const float* q32 = state ? sync_q32(state, K, Q, ldq) : nullptr;
  const int64_t ld32 = q32 ? state->ld32 : 0;
  std::vector<int> work;
  std::vector<char> in(K, 0);
  for (int i = 0; i < nA; ++i) {
    const int a = A_init ? A_init[i] : -1;
    if (a >= 0 && a < K && !in[a]) work.push_back(a), in[a] = 1;
  }
  if (work.empty()) {
    int j = 0;
    for (int i = 1; i < K; ++i)
      if (h[i] > h[j]) j = i;
    work.push_back(j), in[j] = 1;
  }
  Chol ch;
  std::vector<double> a, qh(K), qe(K), bh(K), be(K, 1.0), lam_w, grad, lam_full;
  std::vector<int> solved;     // the working set lam_w belongs to
  auto rebuild = [&]() {
    const int k = (int)work.size();
    a.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) a[(size_t)i * k + j] = Q[(int64_t)work[i] * ldq + work[j]] + (i == j ? reg : 0.0);
    return chol_full(ch, a, k);
  };
  // With a state whose last working set holds every warm-start column (the usual case: the
  // warm start is that set's support), start from the whole last set and its factor; its other
  // members leave again at the first negative weight.
  bool reuse = false;
  if (state && state->ch.ok && state->ch.k == (int)state->work.size() && !state->work.empty()) {
    reuse = true;
    std::vector<char> inw(K, 0);
    for (int w : state->work) {
      if (w < 0 || w >= K) { reuse = false; break; }
      inw[w] = 1;
    }
    for (size_t i = 0; reuse && i < work.size(); ++i)
      if (!inw[work[i]]) reuse = false;
    if (reuse) {
      work = state->work;
      std::fill(in.begin(), in.end(), 0);
      for (int w : work) in[w] = 1;
    }
  }
  // the last call ended on a solve with this working set and factor: when h_W is also the same, its
  // solution is this call's first
  bool cached = false;
  bool fwd = false;                            // yh, ye hold L^-1 h_W and L^-1 1 for work and ch
  std::vector<double> yh(K + 1), ye(K + 1);
  if (reuse) {
    ch = std::move(state->ch);                 // the factor of this working set, from the last call
    const size_t k = work.size();
    cached = state->bh.size() == k && state->qh.size() == k && state->qe.size() == k;
    for (size_t i = 0; cached && i < k; ++i) cached = state->bh[i] == h[work[i]];
    if (cached && state->yh.size() == k && state->ye.size() == k) {
      std::copy(state->yh.begin(), state->yh.end(), yh.begin());
      std::copy(state->ye.begin(), state->ye.end(), ye.begin());
      fwd = true;
    }
  } else {
    ch.reserve(std::min(K, std::max(64, (int)work.size() + 16)));   // room for the columns to come
    rebuild();
  }
  double mu = 0.0;
  for (int outer = 0; outer < max_outer; ++outer) {
    const int k = (int)work.size();
    for (int i = 0; i < k; ++i) bh[i] = h[work[i]];
    if (outer == 0 && cached) {
      std::copy(state->qh.begin(), state->qh.end(), qh.begin());
      std::copy(state->qe.begin(), state->qe.end(), qe.begin());
    } else if (ch.ok && ch.k == k) {
      if (!fwd) {
        for (int i = 0; i < k; ++i) fwd_row(ch, i, bh.data(), be.data(), yh.data(), ye.data());
        fwd = true;
      }
      bwd2(ch, yh.data(), ye.data(), qh.data(), qe.data());
    } else {
      a.assign((size_t)k * k, 0.0);
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) a[(size_t)i * k + j] = Q[(int64_t)work[i] * ldq + work[j]] + (i == j ? reg : 0.0);
      pinv_solve(a, k, bh.data(), be.data(), qh.data(), qe.data());
    }
    double se = 0.0, sh = 0.0;
    for (int i = 0; i < k; ++i) se += qe[i], sh += qh[i];
    mu = std::fabs(se) > 1e-30 ? (sh - 1.0) / se : 0.0;
    lam_w.assign(k, 0.0);
    solved = work;
    int imin = 0;
    for (int i = 0; i < k; ++i) {
      lam_w[i] = qh[i] - mu * qe[i];
      if (lam_w[i] < lam_w[imin]) imin = i;
    }
    if (k && lam_w[imin] < -tol) {
      in[work[imin]] = 0;
      work.erase(work.begin() + imin);
      if (ch.ok && ch.k == k) chol_delete(ch, imin);
      else rebuild();
      fwd = false;
      continue;
    }
    if (K > k) {
      // gradient of the inactive columns: g_j = mu - h_j + sum_i Q[j, work_i] lam_i. Q is symmetric,
      // so it is read by whole rows, whichever are fewer: the K - k inactive rows (a dot product
      // with lam spread over all K columns) or the k working rows (axpys into every g_j)
      grad.resize(K);
      if (K - k < k) {
        lam_full.assign(K, 0.0);
        for (int i = 0; i < k; ++i) lam_full[work[i]] = lam_w[i];
      }
      const double work_fma = (double)K * (double)std::min(k, K - k);
      const int parts =
          (work_fma < kParallelFma || tl_in_pool_run) ? 1 : std::min(HostPool::get().size(), K / 128);
      auto part = [&](int j0, int j1) {
        if (q32)
          grad_range<float>(j0, j1, K, k, q32, ld32, h, in.data(), work.data(), lam_w.data(), lam_full.data(), mu,
                            grad.data());
        else
          grad_range<double>(j0, j1, K, k, Q, ldq, h, in.data(), work.data(), lam_w.data(), lam_full.data(), mu,
                             grad.data());
      };
      if (parts <= 1) {
        part(0, K);
      } else {
        // part boundaries on multiples of 16 columns (whole cache lines of grad and of the rows)
        const int chunk = ((K + parts - 1) / parts + 15) & ~15;
        HostPool::get().run(parts, [&](int p) {
          const int j0 = std::min(K, p * chunk), j1 = std::min(K, j0 + chunk);
          part(j0, j1);
        });
      }
      int jbest = -1;
      double gbest = 0.0;
      for (int j = 0; j < K; ++j) {
        if (in[j]) continue;
        if (jbest < 0 || grad[j] < gbest) jbest = j, gbest = grad[j];
      }
      if (jbest >= 0 && gbest < -tol) {
        std::vector<double> q(k);
        for (int i = 0; i < k; ++i) q[i] = Q[(int64_t)jbest * ldq + work[i]];
        work.push_back(jbest), in[jbest] = 1;
        if (ch.ok && ch.k == k) ch.grow(k + 1);
        if (ch.ok && ch.k == k && chol_append(ch, q.data(), Q[(int64_t)jbest * ldq + jbest] + reg)) {
          if (fwd) {                           // the grown factor's forward solutions: one more row
            bh[k] = h[jbest];
            fwd_row(ch, k, bh.data(), be.data(), yh.data(), ye.data());
          }
        } else {
          rebuild();
          fwd = false;
        }
        continue;
      }
    }
    break;
  }
  if (state) {
    state->work = work;
    if (ch.ok && ch.k == (int)work.size()) state->ch = std::move(ch);
    else state->ch.ok = false;
    // the loop's last solve was on the final working set unless it ended at max_outer after a change
    const int k = (int)work.size();
    if (state->ch.ok && solved == work) {
      state->bh.assign(bh.begin(), bh.begin() + k);
      state->qh.assign(qh.begin(), qh.begin() + k);
      state->qe.assign(qe.begin(), qe.begin() + k);
      if (fwd) {
        state->yh.assign(yh.begin(), yh.begin() + k);
        state->ye.assign(ye.begin(), ye.begin() + k);
      } else {
        state->yh.clear(), state->ye.clear();
      }
    } else {
      state->bh.clear(), state->qh.clear(), state->qe.clear(), state->yh.clear(), state->ye.clear();
    }
  }
  std::vector<double> lam(K, 0.0);
  for (size_t i = 0; i < solved.size(); ++i) lam[solved[i]] = lam_w[i];
  project_simplex(lam);
  int ns = 0;
  for (int w : work)
    if (lam[w] > 1e-9) support[ns++] = w;
  *n_support = ns;
  for (int i = 0; i < K; ++i) lam_out[i] = lam[i];
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// The host side of one column-generation run between two pricing sweeps (the loop body of
// flow_cg/solver.py:92-240 of the reference, as the engine's cg_steps runs it): the ranking
// columns as bitsets over the x-major rows, their integer Gram matrix Q and h = V'p, the master QP
// warm from the last working set and factor, p_out = V lam over lam's support, and the float32
// scores of the next sweep, in one call per round. A batched bootstrap runs tens of these per
// round beside one batched sweep, so these steps bound it once the sweeps are shared.
struct rumgpu_cgr {
  int n = 0, W = 0;
  int64_t M = 0;
  std::vector<double> p;          // x-major rows: float32-rounded p as double
  std::vector<uint32_t> menu;     // the rows' menus
  std::vector<int32_t> alt;       // the rows' alternatives
  std::vector<uint64_t> bits;     // cap x W
  std::vector<double> Q;          // cap x cap (row i at i * cap)
  std::vector<double> h, lam;     // cap
  std::vector<double> p_out;      // M
  std::vector<int32_t> work;      // the last master's support, the next one's start
  rumgpu_qp qp;
  int K = 0, cap = 0, max_K = 0;
};

static void cgr_grow(rumgpu_cgr* c, int cap) {
  cap = std::min(std::max(cap, 1), c->max_K);
  if (cap <= c->cap) return;
  std::vector<uint64_t> bits((size_t)cap * c->W, 0);
  std::vector<double> Q((size_t)cap * cap, 0.0), h(cap, 0.0), lam(cap, 0.0);
  std::copy(c->bits.begin(), c->bits.begin() + (size_t)c->K * c->W, bits.begin());
  for (int i = 0; i < c->K; ++i)
    std::copy(c->Q.begin() + (size_t)i * c->cap, c->Q.begin() + (size_t)i * c->cap + c->K, Q.begin() + (size_t)i * cap);
  std::copy(c->h.begin(), c->h.begin() + c->K, h.begin());
  std::copy(c->lam.begin(), c->lam.begin() + c->K, lam.begin());
  c->bits.swap(bits), c->Q.swap(Q), c->h.swap(h), c->lam.swap(lam);
  c->cap = cap;
}

// append a column given as a bitset: its Gram row (integer overlaps), h (its rows' p in ascending
// row order) and lam = 0
static int cgr_append(rumgpu_cgr* c, const uint64_t* b) {
  RG_CHECK(c->K < c->max_K, RG_EINVAL, "cg rounds: max_K = %d columns exceeded", c->max_K);
  if (c->K >= c->cap) cgr_grow(c, 2 * c->cap);
  const int k = c->K, W = c->W, cap = c->cap;
  uint64_t* bk = c->bits.data() + (size_t)k * W;
  std::copy(b, b + W, bk);
  double* Qk = c->Q.data() + (size_t)k * cap;
  for (int j = 0; j < k; ++j) {
    const uint64_t* bj = c->bits.data() + (size_t)j * W;
    int64_t cnt = 0;
    for (int w = 0; w < W; ++w) cnt += __builtin_popcountll(bj[w] & bk[w]);
    Qk[j] = (double)cnt;
    c->Q[(size_t)j * cap + k] = (double)cnt;
  }
  int64_t self = 0;
  double hs = 0.0;
  for (int w = 0; w < W; ++w) {
    self += __builtin_popcountll(bk[w]);
    uint64_t m = bk[w];
    while (m) {
      const int t = __builtin_ctzll(m);
      m &= m - 1;
      hs += c->p[(size_t)w * 64 + t];
    }
  }
  Qk[k] = (double)self;
  c->h[k] = hs;
  c->lam[k] = 0.0;
  c->K = k + 1;
  return RG_OK;
}

// p_out = sum over lam's nonzero entries (ascending column) of lam_k times column k
static void cgr_pout(rumgpu_cgr* c) {
  std::fill(c->p_out.begin(), c->p_out.end(), 0.0);
  for (int k = 0; k < c->K; ++k) {
    const double l = c->lam[k];
    if (l == 0.0) continue;
    const uint64_t* bk = c->bits.data() + (size_t)k * c->W;
    for (int w = 0; w < c->W; ++w) {
      uint64_t m = bk[w];
      while (m) {
        const int t = __builtin_ctzll(m);
        m &= m - 1;
        c->p_out[(size_t)w * 64 + t] += l;
      }
    }
  }
}

extern "C" int rumgpu_cgr_create(int n, int64_t M, const uint64_t* menu, const int32_t* alt, const double* p,
                                 int max_K, rumgpu_cgr** out) {
  RG_CHECK(out && menu && alt && p && M >= 1 && max_K >= 1, RG_EINVAL, "cgr_create: bad argument");
  RG_CHECK(n >= 1 && n <= 32, RG_EUNSUP, "cgr_create: n = %d outside 1..32", n);
  rumgpu_cgr* c = new rumgpu_cgr();
  c->n = n;
  c->M = M;
  c->W = (int)((M + 63) / 64);
  c->max_K = max_K;
  c->p.assign(p, p + M);
  c->menu.resize(M);
  c->alt.assign(alt, alt + M);
  for (int64_t r = 0; r < M; ++r) {
    c->menu[r] = (uint32_t)menu[r];
    if (alt[r] < 0 || alt[r] >= n || !((menu[r] >> alt[r]) & 1ull)) {
      delete c;
      RG_CHECK(false, RG_EINVAL, "cgr_create: row %lld's alternative is not in its menu", (long long)r);
    }
  }
  c->p_out.assign(M, 0.0);
  cgr_grow(c, 64);
  *out = c;
  return RG_OK;
}

extern "C" int rumgpu_cgr_destroy(rumgpu_cgr* c) {
  delete c;
  return RG_OK;
}

// the column of a ranking (rank_pos[x] = position of x, 0 = most preferred): row r is 1 iff its
// alternative is the best-ranked member of its menu (cg_core.py:63-80 of the reference)
extern "C" int rumgpu_cgr_add_ranking(rumgpu_cgr* c, const int32_t* rank_pos) {
  RG_CHECK(c && rank_pos, RG_EINVAL, "cgr_add_ranking: null argument");
  std::vector<uint64_t> b(c->W, 0);
  for (int64_t r = 0; r < c->M; ++r) {
    uint32_t m = c->menu[r];
    int top = -1, best = 1 << 30;
    while (m) {
      const int x = __builtin_ctz(m);
      m &= m - 1u;
      if (rank_pos[x] < best) {
        best = rank_pos[x];
        top = x;
      }
    }
    if (top == c->alt[r]) b[r >> 6] |= 1ull << (r & 63);
  }
  return cgr_append(c, b.data());
}

// a column given as 0/1 bytes over the rows
extern "C" int rumgpu_cgr_add_column(rumgpu_cgr* c, const uint8_t* col) {
  RG_CHECK(c && col, RG_EINVAL, "cgr_add_column: null argument");
  std::vector<uint64_t> b(c->W, 0);
  for (int64_t r = 0; r < c->M; ++r)
    if (col[r]) b[r >> 6] |= 1ull << (r & 63);
  return cgr_append(c, b.data());
}

// dst's columns = src's (same rows): the bitsets and Gram matrix copied, h from dst's p
extern "C" int rumgpu_cgr_copy_columns(rumgpu_cgr* dst, const rumgpu_cgr* src) {
  RG_CHECK(dst && src && dst->M == src->M, RG_EINVAL, "cgr_copy_columns: incompatible runs");
  RG_CHECK(src->K <= dst->max_K, RG_EINVAL, "cgr_copy_columns: %d columns exceed max_K = %d", src->K, dst->max_K);
  dst->K = 0;
  cgr_grow(dst, src->K + 64);
  const int k = src->K, W = dst->W;
  std::copy(src->bits.begin(), src->bits.begin() + (size_t)k * W, dst->bits.begin());
  for (int i = 0; i < k; ++i)
    std::copy(src->Q.begin() + (size_t)i * src->cap, src->Q.begin() + (size_t)i * src->cap + k,
              dst->Q.begin() + (size_t)i * dst->cap);
  for (int i = 0; i < k; ++i) {
    const uint64_t* bk = dst->bits.data() + (size_t)i * W;
    double hs = 0.0;
    for (int w = 0; w < W; ++w) {
      uint64_t m = bk[w];
      while (m) {
        const int t = __builtin_ctzll(m);
        m &= m - 1;
        hs += dst->p[(size_t)w * 64 + t];
      }
    }
    dst->h[i] = hs;
    dst->lam[i] = 0.0;
  }
  dst->K = k;
  return RG_OK;
}

// lam of the first K columns (the rest 0), the next master's starting working set, and p_out
extern "C" int rumgpu_cgr_set_lambda(rumgpu_cgr* c, const double* lam, int K, const int32_t* A, int nA) {
  RG_CHECK(c && lam && K >= 0 && K <= c->K && (A || nA == 0), RG_EINVAL, "cgr_set_lambda: bad argument");
  std::fill(c->lam.begin(), c->lam.end(), 0.0);
  std::copy(lam, lam + K, c->lam.begin());
  c->work.assign(A, A + nA);
  cgr_pout(c);
  return RG_OK;
}

// One master: the QP over the K columns from the last working set (its factor kept), then p_out,
// out[0] = lam'Q lam and out[1] = h'lam (sums over the support, ascending), and the next sweep's
// scores float32(p - p_out).
extern "C" int rumgpu_cgr_master(rumgpu_cgr* c, int max_outer, double tol, double reg, double* out, float* scores) {
  RG_CHECK(c && out && scores && c->K >= 1, RG_EINVAL, "cgr_master: bad argument");
  const int K = c->K;
  std::vector<double> lam(K);
  std::vector<int32_t> sup(K);
  int ns = 0;
  RG_TRY(qp_simplex_impl(K, c->Q.data(), c->cap, c->h.data(), c->work.data(), (int)c->work.size(), max_outer, tol, reg,
                         lam.data(), sup.data(), &ns, &c->qp));
  c->work.assign(sup.begin(), sup.begin() + ns);
  std::copy(lam.begin(), lam.end(), c->lam.begin());
  cgr_pout(c);
  double lql = 0.0, hl = 0.0;
  for (int i = 0; i < K; ++i) {
    const double li = lam[i];
    if (li == 0.0) continue;
    const double* Qi = c->Q.data() + (size_t)i * c->cap;
    double s = 0.0;
    for (int j = 0; j < K; ++j)
      if (lam[j] != 0.0) s += Qi[j] * lam[j];
    lql += li * s;
    hl += c->h[i] * li;
  }
  out[0] = lql;
  out[1] = hl;
  for (int64_t r = 0; r < c->M; ++r) scores[r] = (float)(c->p[r] - c->p_out[r]);
  return RG_OK;
}

// The masters of `count` runs with the same rows, one per host thread of the pool at a time: run i's
// out[2 i .. 2 i + 1] and scores[i M ..] are those of rumgpu_cgr_master on that run alone, status[i]
// its return code (the last error message belongs to one of the failed runs). With ranks (count x n,
// may be null), a run whose has_rank[i] is set first takes that ranking's column
// (rumgpu_cgr_add_ranking), as the column generation does before each master.
extern "C" int rumgpu_cgr_master_batch(rumgpu_cgr* const* runs, int count, const int32_t* ranks, const int32_t* has_rank,
                                       int max_outer, double tol, double reg, double* out, float* scores, int* status) {
  RG_CHECK(runs && count >= 0 && out && scores && status && (!ranks || has_rank), RG_EINVAL,
           "cgr_master_batch: bad argument");
  if (count == 0) return RG_OK;
  RG_CHECK(runs[0], RG_EINVAL, "cgr_master_batch: null run");
  const int64_t M = runs[0]->M;
  const int n = runs[0]->n;
  for (int i = 0; i < count; ++i)
    RG_CHECK(runs[i] && runs[i]->M == M && runs[i]->n == n, RG_EINVAL, "cgr_master_batch: runs of different rows");
  HostPool::get().run(count, [&](int i) {
    tl_in_pool_run = true;
    int st = RG_OK;
    if (ranks && has_rank[i]) st = rumgpu_cgr_add_ranking(runs[i], ranks + (size_t)i * n);
    if (st == RG_OK)
      st = rumgpu_cgr_master(runs[i], max_outer, tol, reg, out + 2 * (size_t)i, scores + (size_t)i * M);
    status[i] = st;
    tl_in_pool_run = false;
  });
  for (int i = 0; i < count; ++i)
    if (status[i] != RG_OK) return status[i];
  return RG_OK;
}

// lam (K), p_out (M) and the columns as 0/1 bytes (K x M); any output may be null
extern "C" int rumgpu_cgr_get(rumgpu_cgr* c, int* K, double* lam, double* p_out, uint8_t* cols) {
  RG_CHECK(c, RG_EINVAL, "cgr_get: null handle");
  if (K) *K = c->K;
  if (lam) std::copy(c->lam.begin(), c->lam.begin() + c->K, lam);
  if (p_out) std::copy(c->p_out.begin(), c->p_out.end(), p_out);
  if (cols)
    for (int k = 0; k < c->K; ++k) {
      const uint64_t* bk = c->bits.data() + (size_t)k * c->W;
      for (int64_t r = 0; r < c->M; ++r) cols[(size_t)k * c->M + r] = (uint8_t)((bk[r >> 6] >> (r & 63)) & 1ull);
    }
  return RG_OK;
}
