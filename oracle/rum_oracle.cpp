// CPU oracle: a plain restatement of the reference's BM-ADMM / Flow-CG arithmetic.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links or loads this file: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as the checker
// (or as the timed CPU baseline).
//
// Built with -O2 -ffp-contract=off so every a*X + b*Y is two roundings of products and one of
// the sum, the same IEEE operations the HIP kernels perform (they are built the same way).
//
// Each function names the reference file:line it restates (paths relative to the reference
// repository root, solver/ prefix omitted where obvious).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>
#include <algorithm>

typedef int64_t i64;
typedef uint64_t u64;

static inline u64 compress_bit(u64 m, int i) {            // bm_admm/kernels.py:30-33
  return (m & ((1ull << i) - 1ull)) | ((m >> (i + 1)) << i);
}
static inline u64 expand_bit(u64 rest, int i) {           // arcflow_kernels.py:310 (tail of an arc)
  return (rest & ((1ull << i) - 1ull)) | ((rest >> i) << (i + 1));
}

// ---------------------------------------------------------------------------------------------
// Generalized 2x2 butterfly over the n-1 rest bits of every item block, bit stages ascending.
// common/butterfly.py:22-48 (bfly_blk), bm_admm/butterfly2.py:73-102 (bfly_blk2): per stage
// (lo, hi) -> (a*lo + b*hi, c*lo + d*hi).
template <typename T>
static void butterfly(T* v, int n, T a, T b, T c, T d) {
  const i64 Nloc = 1ll << (n - 1);
#pragma omp parallel for schedule(static)
  for (int x = 0; x < n; ++x) {
    T* col = v + (i64)x * Nloc;
    for (int j = 0; j < n - 1; ++j) {
      const i64 h = 1ll << j;
      for (i64 l = 0; l < Nloc; ++l) {
        if (l & h) continue;
        T X = col[l], Y = col[l | h];
        T lo = a * X + b * Y;
        T hi = c * X + d * Y;
        col[l] = lo;
        col[l | h] = hi;
      }
    }
  }
}

// Threads the parallel loops use (OMP_NUM_THREADS, else every core the process may run on).
extern "C" int orc_threads() { return omp_get_max_threads(); }

extern "C" void orc_butterfly_f64(double* v, int n, double a, double b, double c, double d) {
  butterfly<double>(v, n, a, b, c, d);
}
extern "C" void orc_butterfly_f32(float* v, int n, float a, float b, float c, float d) {
  butterfly<float>(v, n, a, b, c, d);
}

// Slab positions of the observed rows: transforms.py:59-65 (build_obs_idx).
extern "C" void orc_obs_index(int n, i64 M, const u64* menu, const int32_t* alt, i64* out) {
  const i64 Nloc = 1ll << (n - 1);
  for (i64 j = 0; j < M; ++j) {
    int x = alt[j];
    u64 m = menu[j] & ~(1ull << x);
    out[j] = (i64)x * Nloc + (i64)compress_bit(m, x);
  }
}

// ---------------------------------------------------------------------------------------------
// Exact per-menu prox: bm_admm/solver.py:66-139 (prox_all), one menu S at a time, all 2^n - 1
// menus. Values are `real` (T); the sort and the threshold run in double.
template <typename T>
static void prox_all(int n, const T* slab_in, T* slab_out, const T* xhat, const T* wmask, T beta2) {
  const i64 Nloc = 1ll << (n - 1);
  const u64 NN = 1ull << n;
#pragma omp parallel for schedule(dynamic, 4096)
  for (u64 S = 1; S < NN; ++S) {
    T vals[64];
    i64 pos_arr[64];
    double srt[64];
    int k = __builtin_popcountll(S);
    u64 SS = S;
    int kk = 0;
    T w0 = (T)0;
    bool mixed = false;
    bool single = false;
    while (SS) {
      u64 bit = SS & (0ull - SS);
      int i = __builtin_ctzll(bit);
      u64 rest = compress_bit(S & ~(1ull << i), i);
      i64 pos = (i64)i * Nloc + (i64)rest;
      if (k == 1) { slab_out[pos] = (T)1; single = true; break; }
      T v = slab_in[pos], w = wmask[pos], xh = xhat[pos];
      if (kk == 0) w0 = w; else if (w != w0) mixed = true;
      vals[kk] = (beta2 * v + (T)2 * w * xh) / (beta2 + (T)2 * w);
      pos_arr[kk] = pos;
      SS ^= bit;
      kk++;
    }
    if (single) continue;
    if (mixed) {
      for (int a2 = 0; a2 < k; ++a2) {
        i64 q = pos_arr[a2];
        srt[a2] = (double)beta2 * (double)slab_in[q] + 2.0 * (double)wmask[q] * (double)xhat[q];
      }
      for (int a2 = 1; a2 < k; ++a2) {               // stable insertion sort, descending
        double key = srt[a2];
        i64 kp = pos_arr[a2];
        int b = a2 - 1;
        while (b >= 0 && srt[b] < key) { srt[b + 1] = srt[b]; pos_arr[b + 1] = pos_arr[b]; --b; }
        srt[b + 1] = key;
        pos_arr[b + 1] = kp;
      }
      double stc = 0.0, sic = 0.0, mu = 0.0;
      for (int a2 = 0; a2 < k; ++a2) {
        double c = (double)beta2 + 2.0 * (double)wmask[pos_arr[a2]];
        stc += srt[a2] / c;
        sic += 1.0 / c;
        double th = (stc - 1.0) / sic;
        if (srt[a2] - th > 0.0) mu = th;
      }
      for (int a2 = 0; a2 < k; ++a2) {
        double c = (double)beta2 + 2.0 * (double)wmask[pos_arr[a2]];
        double yv = (srt[a2] - mu) / c;
        slab_out[pos_arr[a2]] = yv > 0.0 ? (T)yv : (T)0.0;
      }
      continue;
    }
    for (int a2 = 0; a2 < k; ++a2) srt[a2] = (double)vals[a2];
    for (int a2 = 1; a2 < k; ++a2) {
      double key = srt[a2];
      int b = a2 - 1;
      while (b >= 0 && srt[b] < key) { srt[b + 1] = srt[b]; --b; }
      srt[b + 1] = key;
    }
    double css = 0.0, tau = 0.0;
    for (int a2 = 0; a2 < k; ++a2) {
      css += srt[a2];
      double th = (css - 1.0) / (double)(a2 + 1);
      if (srt[a2] - th > 0.0) tau = th;
    }
    for (int a2 = 0; a2 < k; ++a2) {
      double yv = (double)vals[a2] - tau;
      slab_out[pos_arr[a2]] = yv > 0.0 ? (T)yv : (T)0.0;
    }
  }
}

extern "C" void orc_prox_f64(int n, const double* in, double* out, const double* xhat, const double* wmask,
                             double beta2) {
  prox_all<double>(n, in, out, xhat, wmask, beta2);
}
extern "C" void orc_prox_f32(int n, const float* in, float* out, const float* xhat, const float* wmask,
                             float beta2) {
  prox_all<float>(n, in, out, xhat, wmask, beta2);
}

// inv_gl[rest] = 1 / (gamma + prod over bits of evals[bit]): bm_admm/solver.py:207-213 (the same
// multiplication order as numpy's loop over bits).
extern "C" void orc_inv_gl(int n, double e0, double e1, double gamma, double* out) {
  const i64 Nloc = 1ll << (n - 1);
  for (i64 r = 0; r < Nloc; ++r) {
    double lam = 1.0;
    for (int bit = 0; bit < n - 1; ++bit) lam *= ((r >> bit) & 1) ? e1 : e0;
    out[r] = 1.0 / (gamma + lam);
  }
}

// ---------------------------------------------------------------------------------------------
// `iters` literal BM-ADMM iterations: bm_admm/solver.py:240-254 (one_iter) with the fused
// elementwise kernels of :33-53. State (x2, s, y, z) is updated in place. Q is row-major
// [Q00, Q01, Q10, Q11] of numpy's eigh of [[1,-1],[-1,1+p^2]]; inv_gl is in the slab dtype.
template <typename T>
static void admm_iters(int n, i64 iters, T* x2, T* s, T* y, T* z, const T* xhat, const T* wmask,
                       const T* inv_gl, double beta1, double beta2, double alpha_relax, double p,
                       const double* Q) {
  const i64 Nloc = 1ll << (n - 1);
  const i64 m2 = (i64)n * Nloc;
  std::vector<T> w(m2), AT(m2), rhs(m2), x(m2), Ax(m2), xk(m2), x2n(m2);
  const T b1 = (T)beta1, b2 = (T)beta2, ib1 = (T)(1.0 / beta1), ib2 = (T)(1.0 / beta2), a = (T)alpha_relax;
  const T one = (T)1;
  for (i64 it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) w[i] = y[i] - b1 * s[i];                                  // fuse_w
    std::memcpy(AT.data(), w.data(), m2 * sizeof(T));
    butterfly<T>(AT.data(), n, (T)1.0, (T)0.0, (T)-1.0, (T)p);                              // raw A'
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) rhs[i] = (z[i] + b2 * x2[i] - AT[i]) * ib1;               // fuse_rhs
    std::memcpy(x.data(), rhs.data(), m2 * sizeof(T));
    butterfly<T>(x.data(), n, (T)Q[0], (T)Q[2], (T)Q[1], (T)Q[3]);                          // Q^T
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) x[i] *= inv_gl[i % Nloc];                                  // diag_mul
    butterfly<T>(x.data(), n, (T)Q[0], (T)Q[1], (T)Q[2], (T)Q[3]);                          // Q
    std::memcpy(Ax.data(), x.data(), m2 * sizeof(T));
    butterfly<T>(Ax.data(), n, (T)1.0, (T)-1.0, (T)0.0, (T)p);                              // raw A
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) xk[i] = a * x[i] + (one - a) * x2[i] - ib2 * z[i];        // fuse_xk
    prox_all<T>(n, xk.data(), x2n.data(), xhat, wmask, b2);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) {                                                          // fuse_z
      T xo = a * x[i] + (one - a) * x2[i];
      z[i] -= b2 * (xo - x2n[i]);
    }
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m2; ++i) {                                                          // fuse_sy
      T axor = -a * Ax[i] - (one - a) * s[i];
      T sn = ib1 * y[i] - axor;
      sn = sn > (T)0 ? sn : (T)0;
      y[i] -= b1 * (axor + sn);
      s[i] = sn;
    }
    std::memcpy(x2, x2n.data(), m2 * sizeof(T));
  }
}

extern "C" void orc_admm_iters_f64(int n, i64 iters, double* x2, double* s, double* y, double* z,
                                   const double* xhat, const double* wmask, const double* inv_gl,
                                   double beta1, double beta2, double alpha, double p, const double* Q) {
  admm_iters<double>(n, iters, x2, s, y, z, xhat, wmask, inv_gl, beta1, beta2, alpha, p, Q);
}
extern "C" void orc_admm_iters_f32(int n, i64 iters, float* x2, float* s, float* y, float* z,
                                   const float* xhat, const float* wmask, const float* inv_gl,
                                   double beta1, double beta2, double alpha, double p, const double* Q) {
  admm_iters<float>(n, iters, x2, s, y, z, xhat, wmask, inv_gl, beta1, beta2, alpha, p, Q);
}

// ---------------------------------------------------------------------------------------------
// Neumaier-compensated sum (the accurate reduction the oracle reports).
struct KSum {
  double s = 0.0, c = 0.0;
  void add(double v) {
    double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) c += (s - t) + v; else c += (v - t) + s;
    s = t;
  }
  double get() const { return s + c; }
};

// Nodes sorted by popcount (stable, ascending mask within a layer) and layer bounds:
// arcflow_kernels.py:329-337 (build_layers).
static void build_layers(int n, std::vector<i64>& order, std::vector<i64>& bounds) {
  const i64 NN = 1ll << n;
  order.resize(NN);
  bounds.assign(n + 2, 0);
  std::vector<i64> cnt(n + 2, 0);
  for (i64 v = 0; v < NN; ++v) cnt[__builtin_popcountll(v)]++;
  for (int c = 0; c <= n; ++c) bounds[c + 1] = bounds[c] + cnt[c];
  std::vector<i64> fill(bounds.begin(), bounds.end());
  for (i64 v = 0; v < NN; ++v) order[fill[__builtin_popcountll(v)]++] = v;
}

// The float64 feasibility push of any slab iterate: transforms.py:94-98 (push_to_ub64) =
// mobius (:47-50) -> make_r_feasible_gpu64 (arcflow_kernels.py:415-434: wmax64 :354, outwt64
// :357-365, massg64 :367-376 per layer, extract64 :378-386) -> zeta (:53-56) -> phi_at (:68-71).
// x is the iterate in float64 (the caller casts). Returns phi_ub; x_ub_obs[j] = x_ub[obs[j]].
extern "C" double orc_push_ub64(int n, const double* x, i64 M, const i64* obs, const double* phat,
                                double* x_ub_obs) {
  const i64 Nloc = 1ll << (n - 1);
  const i64 m2 = (i64)n * Nloc;
  const i64 NN = 1ll << n;
  std::vector<double> w(x, x + m2);
  butterfly<double>(w.data(), n, 1.0, -1.0, 0.0, 1.0);                   // superset Mobius
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < m2; ++i) w[i] = (w[i] > 0.0 ? w[i] : 0.0) + 1e-12;   // wmax64
  std::vector<double> outw(NN), mass(NN, 0.0), rf(m2);
#pragma omp parallel for schedule(static)
  for (i64 D = 0; D < NN; ++D) {                                          // outwt64
    double sacc = 0.0;
    for (int xx = 0; xx < n; ++xx) {
      i64 bit = 1ll << xx;
      if (!(D & bit)) sacc += w[(i64)xx * Nloc + (i64)compress_bit((u64)D, xx)];
    }
    outw[D] = sacc;
  }
  std::vector<i64> order, bounds;
  build_layers(n, order, bounds);
  mass[0] = 1.0;
  for (int c = 1; c <= n; ++c) {                                          // massg64 per layer
#pragma omp parallel for schedule(static)
    for (i64 t = bounds[c]; t < bounds[c + 1]; ++t) {
      i64 C = order[t];
      double sacc = 0.0;
      for (int xx = 0; xx < n; ++xx) {
        i64 bit = 1ll << xx;
        if (C & bit) {
          i64 P = C ^ bit;
          double ow = outw[P];
          if (ow > 0.0) sacc += mass[P] * w[(i64)xx * Nloc + (i64)compress_bit((u64)P, xx)] / ow;
        }
      }
      mass[C] = sacc;
    }
  }
#pragma omp parallel for schedule(static)
  for (i64 a = 0; a < m2; ++a) {                                          // extract64
    i64 xx = a / Nloc, rest = a - xx * Nloc;
    i64 T = (i64)expand_bit((u64)rest, (int)xx);
    double ow = outw[T];
    rf[a] = (ow > 0.0) ? mass[T] * w[a] / ow : 0.0;
  }
  butterfly<double>(rf.data(), n, 1.0, 1.0, 0.0, 1.0);                   // superset zeta
  KSum acc;
  for (i64 j = 0; j < M; ++j) {
    double v = rf[obs[j]];
    if (x_ub_obs) x_ub_obs[j] = v;
    double d = v - phat[j];
    acc.add(d * d);
  }
  return std::sqrt(acc.get());
}

// The float64 Frank-Wolfe lower bound at a point given by its observed values y:
// transforms.py:101-114 (fw_lb_rows_gpu) with longest_path_gpu64 (arcflow_kernels.py:437-448,
// lpath64 :388-397). Returns lb; rc_out (if non-null) gets rc.
extern "C" double orc_fw_lb64(int n, i64 M, const i64* obs, const double* y, const double* phat, double* rc_out) {
  const i64 Nloc = 1ll << (n - 1);
  const i64 m2 = (i64)n * Nloc;
  const i64 NN = 1ll << n;
  std::vector<double> g(m2, 0.0), r(M);
  KSum f, ry;
  for (i64 j = 0; j < M; ++j) {
    r[j] = phat[j] - y[j];
    f.add(r[j] * r[j]);
    ry.add(r[j] * y[j]);
    g[obs[j]] = r[j];
  }
  butterfly<double>(g.data(), n, 1.0, 0.0, 1.0, 1.0);                    // sum over subsets
  std::vector<i64> order, bounds;
  build_layers(n, order, bounds);
  std::vector<double> dp(NN);
  dp[0] = 0.0;
  for (int c = 1; c <= n; ++c) {
#pragma omp parallel for schedule(static)
    for (i64 t = bounds[c]; t < bounds[c + 1]; ++t) {
      i64 C = order[t];
      double best = -1e300;
      for (int xx = 0; xx < n; ++xx) {
        i64 bit = 1ll << xx;
        if (C & bit) {
          i64 P = C ^ bit;
          double v = dp[P] + g[(i64)xx * Nloc + (i64)compress_bit((u64)P, xx)];
          if (v > best) best = v;
        }
      }
      dp[C] = best;
    }
  }
  double rc = dp[NN - 1] - ry.get();
  if (rc_out) *rc_out = rc;
  double lb2 = f.get() - 2.0 * rc;
  return std::sqrt(lb2 > 0.0 ? lb2 : 0.0);
}

// ---------------------------------------------------------------------------------------------
// Mask-indexed Bellman-Ford pricing: flow_cg/bf_pricing.py:8-47 (bf_update) swept over the
// layers k = 1..n in ascending mask order (flow_cg/solver.py:171-186), then the host backtrace
// (:193-200). Rows are x-major: offsets[x]..offsets[x+1] are alternative x's rows. V and pred
// are float32 / uint8 per node, zero-filled first. Returns V[full] and rank_pos.
extern "C" float orc_price_mask(int n, const u64* masks, const float* scores, const u64* offsets, float* V,
                                uint8_t* pred, int32_t* rank_pos) {
  const i64 NN = 1ll << n;
  std::memset(V, 0, NN * sizeof(float));
  std::memset(pred, 0, NN);
  std::vector<i64> order, bounds;
  build_layers(n, order, bounds);
  for (int c = 1; c <= n; ++c) {
    for (i64 t = bounds[c]; t < bounds[c + 1]; ++t) {
      u64 D = (u64)order[t];
      float best = 0.0f;
      bool seen = false;
      u64 S = D;
      while (S) {
        u64 bit = S & (0ull - S);
        int x = __builtin_ctzll(bit);
        u64 prev = D ^ bit;
        float w = 0.0f;
        for (u64 j = offsets[x]; j < offsets[x + 1]; ++j) {
          u64 Dp = masks[j];
          if ((Dp & D) == Dp) w += scores[j];
        }
        float cand = V[prev] + w;
        if (!seen || cand > best) { best = cand; seen = true; pred[D] = (uint8_t)x; }
        S ^= bit;
      }
      V[D] = best;
    }
  }
  u64 D = (u64)NN - 1;
  for (int i = 0; i < n; ++i) {
    int cur = pred[D];
    rank_pos[cur] = i;
    D ^= (1ull << cur);
  }
  return V[NN - 1];
}

// Colex unranking / ranking: flow_cg/cg_core.py:28-60, pricing_stream.py:254-264.
static void binom_table(int n, std::vector<u64>& tab) {
  tab.assign((size_t)(n + 1) * (n + 1), 0);
  for (int b = 0; b <= n; ++b) {
    u64 c = 1;
    for (int t = 0; t <= b; ++t) {
      tab[(size_t)b * (n + 1) + t] = c;
      c = c * (u64)(b - t) / (u64)(t + 1);
    }
  }
}

extern "C" u64 orc_colex_rank(u64 mask) {
  u64 r = 0;
  int t = 0, b = 0;
  while (mask) {
    if (mask & 1) {
      t += 1;
      u64 c = 1;                       // C(b, t)
      if (t <= b) { for (int q = 0; q < t; ++q) c = c * (u64)(b - q) / (u64)(q + 1); } else c = 0;
      r += c;
    }
    mask >>= 1;
    b += 1;
  }
  return r;
}

// The two-live-layer rank-indexed sweep with block quantization: flow_cg/pricing_stream.py:61-182
// (BODY, bf_stream_f32/_u16/_u24) and StreamPricer.sweep (:302-345). mode 0 = f32, 1 = u16,
// 2 = u24. BLOCK = 1024 ranks per quantization block. The u16/u24 decode is written as
// fmaf(q, e, m) (the contraction NVRTC applies by default to m + q * e). pred_top receives the
// predecessor bytes of the layers k >= k_pred concatenated (layer k at offset sum_{k' < k} C(n,k')
// minus the same sum at k_pred). Returns Vmax in double (decoded as the host does) and the margin.
extern "C" double orc_price_stream(int n, int mode, int k_pred, const u64* masks, const float* scores,
                                   const u64* offsets, uint8_t* pred_top, double* margin_out) {
  const int BLOCK = 1024;
  std::vector<u64> binom;
  binom_table(n, binom);
  auto C = [&](int b, int t) -> u64 { return binom[(size_t)b * (n + 1) + t]; };
  u64 cap = 0;
  for (int k = 0; k <= n; ++k) cap = std::max(cap, C(n, k));
  u64 nblk = (cap + BLOCK - 1) / BLOCK;
  std::vector<float> vf[2];
  std::vector<uint32_t> vq[2];
  std::vector<float> mb[2], eb[2];
  for (int i = 0; i < 2; ++i) {
    vf[i].assign(cap, 0.0f); vq[i].assign(cap, 0); mb[i].assign(nblk, 0.0f); eb[i].assign(nblk, 0.0f);
  }
  std::vector<float> layer_step(n + 1, 0.0f);
  int cur = 0, prev = 1;
  u64 pred_off = 0;
  const float levels = mode == 1 ? 65535.0f : 16777215.0f;
  for (int k = 1; k <= n; ++k) {
    u64 count = C(n, k);
    u64 grid = (count + BLOCK - 1) / BLOCK;
    bool store = k >= k_pred;
    std::vector<float> vals(count);
    std::vector<uint8_t> preds(count);
    for (u64 r = 0; r < count; ++r) {
      u64 rr = r;
      int b = n - 1;
      int elems[64];
      u64 D = 0;
      for (int t = k; t >= 1; --t) {
        while (C(b, t) > rr) --b;
        elems[t - 1] = b;
        D |= 1ull << b;
        rr -= C(b, t);
        --b;
      }
      float best = 0.0f;
      bool seen = false;
      int bestx = 0;
      u64 suf = 0, remain = r;
      for (int t = k - 1; t >= 0; --t) {
        int x = elems[t];
        remain -= C(x, t + 1);
        u64 rp = remain + suf;
        suf += C(x, t);
        float vp;
        if (mode == 0) vp = vf[prev][rp];
        else vp = fmaf((float)vq[prev][rp], eb[prev][rp / BLOCK], mb[prev][rp / BLOCK]);
        float w = 0.0f;
        for (u64 i = offsets[x]; i < offsets[x + 1]; ++i) {
          u64 Dp = masks[i];
          if ((Dp & D) == Dp) w += scores[i];
        }
        float cand = vp + w;
        if (!seen || cand >= best) { best = cand; seen = true; bestx = x; }
      }
      vals[r] = best;
      preds[r] = (uint8_t)bestx;
    }
    if (mode == 0) {
      for (u64 r = 0; r < count; ++r) vf[cur][r] = vals[r];
    } else {
      float smax_layer = 0.0f;
      for (u64 g = 0; g < grid; ++g) {
        float bmin = 3.0e38f, bmax = -3.0e38f;
        for (u64 r = g * BLOCK; r < std::min(count, (g + 1) * BLOCK); ++r) {
          bmin = std::min(bmin, vals[r]);
          bmax = std::max(bmax, vals[r]);
        }
        float step = (bmax - bmin) / levels;
        mb[cur][g] = bmin;
        eb[cur][g] = step;
        if (g == 0 || step > smax_layer) smax_layer = step;
        for (u64 r = g * BLOCK; r < std::min(count, (g + 1) * BLOCK); ++r) {
          float q = (step > 0.0f) ? rintf((vals[r] - bmin) / step) : 0.0f;
          q = std::fmin(std::fmax(q, 0.0f), levels);
          vq[cur][r] = (uint32_t)q;
        }
      }
      layer_step[k] = smax_layer;
    }
    if (store && pred_top) {
      std::memcpy(pred_top + pred_off, preds.data(), count);
      pred_off += count;
    }
    std::swap(cur, prev);
  }
  double vmax, margin = 0.0;
  if (mode == 0) {
    vmax = (double)vf[prev][0];
  } else {
    vmax = (double)mb[prev][0] + (double)vq[prev][0] * (double)eb[prev][0];
    double sacc = 0.0;
    for (int k = 1; k <= n; ++k) sacc += (double)layer_step[k];
    margin = mode == 1 ? 0.5 * sacc : sacc;
  }
  if (margin_out) *margin_out = margin;
  return vmax;
}

// The 0/1 column of a ranking: flow_cg/cg_core.py:63-80 (map_ranking_to_choice).
extern "C" void orc_ranking_column(int n, i64 M, const u64* menu, const int32_t* alt, const int32_t* rank_pos,
                                   uint8_t* out) {
  for (i64 s = 0; s < M; ++s) {
    u64 mask = menu[s];
    int best_alt = -1, best_rank = n + 1;
    for (int a = 0; a < n; ++a) {
      if ((mask >> a) & 1) {
        int r = rank_pos[a];
        if (r < best_rank) { best_rank = r; best_alt = a; }
      }
    }
    out[s] = best_alt == alt[s] ? 1 : 0;
  }
}
