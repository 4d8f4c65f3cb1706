// BM-ADMM on the Block-Marschak slab, MI355X (gfx950).
//
// One ADMM iteration of the reference (solver/bm_admm/solver.py:240-254) is
//   w = y - b1 s;  AT = A'w;  rhs = (z + b2 x2 - AT)/b1;  x = Q diag(inv_gl) Q^T rhs;  Ax = A x
//   xk = a x + (1-a) x2 - z/b2;  x2n = prox(xk);  z -= b2 (a x + (1-a) x2 - x2n)
//   s, y from Ax (fuse_sy);  x2 = x2n
// with A', Q^T, Q, A Kronecker transforms over the n-1 rest bits. Here the transforms are cut
// into windows (bfly.hpp) and the pointwise updates ride inside the window passes, so with two
// windows an iteration is five passes over the slab:
//
//   SY  (window 1)  A on window 1 -> Ax; s, y update; w; A' on window 1      r tb s y  w tb s y
//   RHS (window k)  A' on window k; rhs; Q^T on window k                     r tb z x2 w tb
//   D   (window 1)  Q^T on window 1; * inv_gl; Q on window 1                  r tb      w tb
//   XK  (window k)  Q on window k -> x; xk; A on window k                     r tb x2 z w xk tb
//   prox            per menu: x2 = prox(xk), z = b2 (x2 - xk)                 r xk      w x2 z
//
// 20 slab transfers (the reference's schedule moves about 45). The A update of s, y is
// deferred to the next iteration's first pass ("pending"). z = b2 (x2n - xk) equals the
// reference's z -= b2 (xo - x2n) in exact arithmetic (xk = xo - z/b2).
//
// The certificates run in float64 whatever the iterate's dtype: the push to a feasible point
// (common/transforms.py:94-98, arcflow_kernels.py:415-434) and the Frank-Wolfe lower bound
// (transforms.py:101-114, arcflow_kernels.py:437-448).
#include "bfly.hpp"
#include "rumgpu_internal.hpp"
#include "transform.hpp"

#include <hip/hip_ext.h>

#include <cmath>
#include <algorithm>
#include <mutex>

enum PassKind { K_FIRST, K_SY, K_SYONLY, K_RHS, K_D, K_XK, K_FULL, K_FULL_FIRST, K_T };

template <typename T>
struct PassArgs {
  T* tb;
  T* s;
  T* y;
  const T* z;
  const T* x2;
  T* xk;
  const T* inv_gl;
  T b1, b2, ib1, ib2, a;
  Coef<T> cA, cAp, cQ, cQt, cT;
  WinGeom g;
  const int* fresh;   // per slot: 1 = state just loaded, skip the deferred s, y update once
  int nt;             // 1: non-temporal slab accesses (a working set the caches cannot hold)
};

// slab element access of the fused passes and the prox; NT: non-temporal
template <bool NT, typename T>
__device__ __forceinline__ T ldg(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void stg(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename T, int RL, int KIND, int BT = 0, int CT = -1, bool NT = false>
__global__ __launch_bounds__(512) void bm_pass(PassArgs<T> P) {
  extern __shared__ __align__(16) unsigned char smem[];
  // NT (slabs too large for the caches): non-temporal accesses for every slab but xk, which the
  // prox reads right after this pass and finds in the Infinity Cache when it is stored normally
  // (at n = 20: 342 us per iteration, 353 with xk non-temporal too, 386 with no NT at all)
  Tile<T, RL, BT, CT> t;
  t.init(P.g, reinterpret_cast<T*>(smem), true, KIND == K_D);
  const T one = (T)1;
  constexpr int R = Tile<T, RL, BT, CT>::R;
  // every global access happens in a lane-contiguous layout (t.coalesce() before it)
  if constexpr (KIND == K_FIRST || KIND == K_FULL_FIRST) {
    if (t.active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const i64 i = t.gidx(j);
        t.v[j] = ldg<NT>(P.y + i) - P.b1 * ldg<NT>(P.s + i);                // fuse_w
      }
    }
  } else {
    t.template load<NT>(P.tb);
  }
  if constexpr (KIND == K_SY || KIND == K_SYONLY || KIND == K_FULL) {
    t.template transform<true>(P.cA);                                                      // Ax (raw A)
    t.coalesce();
    if (t.active) {
      const bool fresh = P.fresh && P.fresh[t.item / P.g.nper];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const i64 i = t.gidx(j);
        const T sj = ldg<NT>(P.s + i), yj = ldg<NT>(P.y + i);
        if (fresh) {                                                        // a refilled slot
          t.v[j] = yj - P.b1 * sj;
          continue;
        }
        const T axor = -P.a * t.v[j] - (one - P.a) * sj;                     // fuse_sy
        T sn = P.ib1 * yj - axor;
        sn = sn > (T)0 ? sn : (T)0;
        const T yn = yj - P.b1 * (axor + sn);
        stg<NT>(P.s + i, sn);
        stg<NT>(P.y + i, yn);
        t.v[j] = yn - P.b1 * sn;                                            // next fuse_w
      }
    }
    if constexpr (KIND == K_SYONLY) return;
  }
  if constexpr (KIND == K_FIRST || KIND == K_SY || KIND == K_RHS || KIND == K_FULL || KIND == K_FULL_FIRST)
    t.template transform<true>(P.cAp);                                                     // raw A'
  if constexpr (KIND == K_RHS || KIND == K_FULL || KIND == K_FULL_FIRST) {
    t.coalesce();
    if (t.active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const i64 i = t.gidx(j);
        t.v[j] = (ldg<NT>(P.z + i) + P.b2 * ldg<NT>(P.x2 + i) - t.v[j]) * P.ib1;  // fuse_rhs
      }
    }
    t.template transform<true>(P.cQt);
  }
  if constexpr (KIND == K_D) t.template transform<true>(P.cQt);
  if constexpr (KIND == K_D || KIND == K_FULL || KIND == K_FULL_FIRST) {
    // inv_gl is one item block (2^(n-1) values, L2-resident): read it in any layout
    if (t.active) {
#pragma unroll
      for (int j = 0; j < R; ++j) t.v[j] *= P.inv_gl[t.rest(j)];            // diag_mul
    }
    t.template transform<true>(P.cQ);
  }
  if constexpr (KIND == K_XK) t.template transform<true>(P.cQ);
  if constexpr (KIND == K_XK || KIND == K_FULL || KIND == K_FULL_FIRST) {
    t.coalesce();
    if (t.active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const i64 i = t.gidx(j);
        P.xk[i] = P.a * t.v[j] + (one - P.a) * ldg<NT>(P.x2 + i) - P.ib2 * ldg<NT>(P.z + i);  // fuse_xk
      }
    }
  }
  if constexpr (KIND == K_XK) t.template transform<true>(P.cA);
  if constexpr (KIND == K_T) t.template transform<true>(P.cT);
  t.coalesce();
  t.template store<NT>(P.tb);
}

// ---------------------------------------------------------------------------------------------
// Per-menu prox (bm_admm/solver.py:66-139). Thread = menu S (natural mask order, so for item i
// the lanes of a wavefront touch a contiguous run of item i's block). Observed members come from
// obsmask[S] (bit i set = row (S, i) observed) and p at pc[obsoff[S] + rank of i in obsmask].
// The sort is a Batcher network on N registers, descending, ties to the lower item (the stable
// insertion sort's order); the threshold sums run in double in sorted order as in the reference.
template <int N>
__device__ __forceinline__ void ce_desc(double (&key)[N], int (&pay)[N], int a, int b) {
  const bool sw = (key[b] > key[a]) || (key[b] == key[a] && (pay[b] & 255) < (pay[a] & 255));
  const double ka = key[a], kb = key[b];
  const int pa = pay[a], pb = pay[b];
  key[a] = sw ? kb : ka;
  key[b] = sw ? ka : kb;
  pay[a] = sw ? pb : pa;
  pay[b] = sw ? pa : pb;
}

template <int N>
__device__ __forceinline__ void sort_desc(double (&key)[N], int (&pay)[N]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p)) ce_desc<N>(key, pay, i + j, i + j + k);
        }
      }
    }
  }
}

// Iteration prox (the hot one): x2 = prox(xk) and z = b2 (x2 - xk) for every menu. The
// threshold comes from Michelot's active-set iteration: start from every member, drop those at or
// below the threshold, recompute, until the set is stable. It reaches the same support and
// threshold as the reference's sorted scan without sorting, so only the members' values (T) and
// keys (double) stay in registers.
template <typename T, int N>
__global__ __launch_bounds__(256, 4) void bm_prox_iter(const T* __restrict__ xk, T* __restrict__ out, T* __restrict__ z,
                                                    const u32* __restrict__ obsmask, const int* __restrict__ obsoff,
                                                    const double* __restrict__ pc, T b2, i64 m2slot, i64 Mslot,
                                                    int* __restrict__ fresh, const int* __restrict__ on) {
  if (on && !on[blockIdx.y]) return;      // an idle draw slot
  // slot = blockIdx.y: its slab and its p
  xk += (i64)blockIdx.y * m2slot;
  out += (i64)blockIdx.y * m2slot;
  z += (i64)blockIdx.y * m2slot;
  pc += (i64)blockIdx.y * Mslot;
  const u32 S = blockIdx.x * blockDim.x + threadIdx.x;
  if (S == 0) {
    if (fresh) fresh[blockIdx.y] = 0;        // this iteration's SY pass has consumed the flag
    return;
  }
  if (S >= (1u << N)) return;
  constexpr u32 Nloc = 1u << (N - 1);     // slab positions fit 32 bits for n <= 26
  // float64 slabs from n = 21 on: non-temporal accesses (at n = 20 the prox's x2 and z stores are
  // still found in the caches by the next passes: 342 against 369 us per iteration)
  constexpr bool NT = sizeof(T) == 8 && N >= 21;
  const int k = __popc(S);
  if (k == 1) {
    const u32 pos = (u32)(__ffs((int)S) - 1) * Nloc;
    out[pos] = (T)1;
    z[pos] = b2 * ((T)1 - xk[pos]);
    return;
  }
  // the offset is loaded with the mask (not after it), so an observed menu's p follows one round trip
  const u32 om = obsmask[S];
  const int off_all = obsoff[S];
  const int off = om ? off_all : 0;
  const bool mixed = (om != 0) && (om != S);
  const double c_obs = (double)b2 + 2.0, c_un = (double)b2;
  // every member's value first, with no use in between, so all the loads are in flight together;
  // key = v when the menu has no observation ((b2 v) / b2 = v).
  // The per-member arrays hold at least 17 elements: the AMDGPU backend turns a private array of
  // up to 16 elements that is written under branches into one vector value, and every branch
  // then copies the whole vector (n = 9-16 spilled 0.1-1.5 KB per lane to scratch and ran the
  // prox about 30 times slower); from 17 elements on the array is split into scalar registers.
  constexpr int NA = N < 17 ? 17 : N;
  double key[NA];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    key[i] = 0.0;
    if ((S >> i) & 1u) key[i] = (double)ldg<NT>(xk + (u32)i * Nloc + (u32)compress_bit(S & ~(1u << i), i));
  }
  // n <= 19: an observed menu keeps its members' xk in LDS for z = b2 (x2 - xk) at the end (four
  // blocks of 256 threads fit a CU's LDS); larger n read xk again
  constexpr bool XO_LDS = N <= 19;
  __shared__ T sxo[XO_LDS ? N : 1][256];
  if (om != 0) {
    // observed menus: the observed members' p, loaded without branches (a clamped index for the
    // others, whose value is not used) so the loads are in flight together, then the keys
    const int nob = __popc(om);
    double xh[NA];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int rk = __popc(om & ((1u << i) - 1u));
      const double v = pc[off + (rk < nob ? rk : nob - 1)];
      xh[i] = ((om >> i) & 1u) ? v : 0.0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if ((S >> i) & 1u) {
        const T xv = (T)key[i];
        if constexpr (XO_LDS) sxo[i][threadIdx.x] = xv;
        const bool ob = (om >> i) & 1u;
        const T xhv = ob ? (T)xh[i] : (T)0;
        key[i] = mixed ? (double)b2 * (double)xv + 2.0 * (ob ? 1.0 : 0.0) * (double)xhv
                       : (double)((b2 * xv + (T)2 * xhv) / (b2 + (T)2));
      }
    }
  }
  u32 act = S;
  double thr = 0.0;
  // curvature c_i = b2 + 2 w_i enters through its reciprocal (two constants per call)
  const double r_obs = 1.0 / c_obs, r_un = 1.0 / c_un;
  const u32 wobs = mixed ? om : 0u;
#pragma unroll 1
  for (int it = 0; it <= N; ++it) {
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if ((act >> i) & 1u) {
        const double r = mixed ? (((wobs >> i) & 1u) ? r_obs : r_un) : 1.0;
        s1 += key[i] * r;
        s2 += r;
      }
    }
    thr = (s1 - 1.0) / s2;
    u32 nact = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (((act >> i) & 1u) && key[i] > thr) nact |= 1u << i;
    if (nact == act || nact == 0) break;
    act = nact;
  }
  if (om == 0) {
    // no observation: r = 1 and xk at pos is the key itself
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if ((S >> i) & 1u) {
        const double yv = (key[i] - thr) * 1.0;
        const T o = yv > 0.0 ? (T)yv : (T)0.0;
        const u32 pos = (u32)i * Nloc + (u32)compress_bit(S & ~(1u << i), i);
        stg<NT>(out + pos, o);
        stg<NT>(z + pos, b2 * (o - (T)key[i]));
      }
    }
    return;
  }
  // observed menus: the outputs replace the keys; xk from LDS, or read again (all loads in flight together)
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((S >> i) & 1u) {
      const double r = mixed ? (((wobs >> i) & 1u) ? r_obs : r_un) : 1.0;
      const double yv = (key[i] - thr) * r;
      key[i] = (double)(yv > 0.0 ? (T)yv : (T)0.0);
    }
  }
  T xo[NA];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    xo[i] = (T)0;
    if ((S >> i) & 1u) {
      if constexpr (XO_LDS) xo[i] = sxo[i][threadIdx.x];
      else xo[i] = ldg<NT>(xk + (u32)i * Nloc + (u32)compress_bit(S & ~(1u << i), i));
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((S >> i) & 1u) {
      const T o = (T)key[i];
      const u32 pos = (u32)i * Nloc + (u32)compress_bit(S & ~(1u << i), i);
      stg<NT>(out + pos, o);
      stg<NT>(z + pos, b2 * (o - xo[i]));
    }
  }
}

// MODE 0: iteration prox, writes x2 and z = b2 (x2 - xk).  MODE 1: cold start, input = xhat,
// writes x2 only.  MODE 2: plain (reference schedule), writes x2n only.
template <typename T, int N, int MODE>
__global__ __launch_bounds__(256) void bm_prox(const T* __restrict__ xk, T* __restrict__ out, T* __restrict__ z,
                                               const u32* __restrict__ obsmask, const int* __restrict__ obsoff,
                                               const double* __restrict__ pc, T b2, i64 m2slot, i64 Mslot) {
  if (xk) xk += (i64)blockIdx.y * m2slot;
  out += (i64)blockIdx.y * m2slot;
  if (z) z += (i64)blockIdx.y * m2slot;
  pc += (i64)blockIdx.y * Mslot;
  const u64 S = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (S == 0 || S >= (1ull << N)) return;
  constexpr i64 Nloc = 1ll << (N - 1);
  const int k = __popcll(S);
  if (k == 1) {
    const int i = __ffsll((long long)S) - 1;
    const i64 pos = (i64)i * Nloc;
    out[pos] = (T)1;
    if constexpr (MODE == 0) z[pos] = b2 * ((T)1 - xk[pos]);
    return;
  }
  const u32 om = obsmask[S];
  const int off = om ? obsoff[S] : 0;
  const bool mixed = (om != 0) && ((u64)om != S);
  const T w0 = om ? (T)1 : (T)0;
  T v[N], vals[N], xh[N];
  double key[N];
  int pay[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const bool in = (S >> i) & 1ull;
    const bool ob = (om >> i) & 1u;
    T xv = (T)0, xhv = (T)0;
    if (ob) xhv = (T)pc[off + __popc(om & ((1u << i) - 1u))];
    if (in) {
      const i64 pos = (i64)i * Nloc + (i64)compress_bit(S & ~(1ull << i), i);
      xv = (MODE == 1) ? xhv : xk[pos];
    }
    v[i] = xv;
    xh[i] = xhv;
    if (mixed) {
      key[i] = in ? (double)b2 * (double)xv + 2.0 * (ob ? 1.0 : 0.0) * (double)xhv : -HUGE_VAL;
    } else {
      vals[i] = (b2 * xv + (T)2 * w0 * xhv) / (b2 + (T)2 * w0);
      key[i] = in ? (double)vals[i] : -HUGE_VAL;
    }
    pay[i] = i | (ob ? 256 : 0);
  }
  double thr = 0.0;
  {
    sort_desc<N>(key, pay);
    if (mixed) {
      double stc = 0.0, sic = 0.0;
      const double c_obs = (double)b2 + 2.0, c_un = (double)b2 + 0.0;
#pragma unroll
      for (int a2 = 0; a2 < N; ++a2) {
        if (a2 < k) {
          const double c = (pay[a2] & 256) ? c_obs : c_un;
          stc += key[a2] / c;
          sic += 1.0 / c;
          const double th = (stc - 1.0) / sic;
          if (key[a2] - th > 0.0) thr = th;
        }
      }
    } else {
      double css = 0.0;
#pragma unroll
      for (int a2 = 0; a2 < N; ++a2) {
        if (a2 < k) {
          css += key[a2];
          const double th = (css - 1.0) / (double)(a2 + 1);
          if (key[a2] - th > 0.0) thr = th;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((S >> i) & 1ull) {
      const bool ob = (om >> i) & 1u;
      double yv;
      if (mixed) {
        const double ti = (double)b2 * (double)v[i] + 2.0 * (ob ? 1.0 : 0.0) * (double)xh[i];
        const double c = (double)b2 + 2.0 * (ob ? 1.0 : 0.0);
        yv = (ti - thr) / c;
      } else {
        yv = (double)vals[i] - thr;
      }
      const T o = yv > 0.0 ? (T)yv : (T)0.0;
      const i64 pos = (i64)i * Nloc + (i64)compress_bit(S & ~(1ull << i), i);
      out[pos] = o;
      if constexpr (MODE == 0) z[pos] = b2 * (o - v[i]);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Elementwise kernels of the plain (reference-schedule) iteration, bm_admm/solver.py:33-53.
template <typename T>
__global__ void ew_w(const T* y, const T* s, T b1, T* w, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x)
    w[i] = y[i] - b1 * s[i];
}
template <typename T>
__global__ void ew_rhs(const T* AT, const T* z, const T* x2, T b2, T ib1, T* rhs, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x)
    rhs[i] = (z[i] + b2 * x2[i] - AT[i]) * ib1;
}
template <typename T>
__global__ void ew_diag(const T* g, i64 Nloc, T* v, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x)
    v[i] *= g[i & (Nloc - 1)];
}
template <typename T>
__global__ void ew_xk(const T* x, const T* x2, const T* z, T a, T ib2, T* xk, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x)
    xk[i] = a * x[i] + ((T)1 - a) * x2[i] - ib2 * z[i];
}
template <typename T>
__global__ void ew_z(const T* x, const T* x2, const T* x2n, T a, T b2, T* z, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x) {
    T xo = a * x[i] + ((T)1 - a) * x2[i];
    z[i] -= b2 * (xo - x2n[i]);
  }
}
template <typename T>
__global__ void ew_sy(const T* Ax, T a, T b1, T ib1, T* s, T* y, i64 m2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x) {
    T axor = -a * Ax[i] - ((T)1 - a) * s[i];
    T sn = ib1 * y[i] - axor;
    sn = sn > (T)0 ? sn : (T)0;
    y[i] -= b1 * (axor + sn);
    s[i] = sn;
  }
}
template <typename TA, typename TB>
__global__ void ew_copy(const TA* a, TB* b, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) b[i] = (TB)a[i];
}
template <typename T>
__global__ void ew_fill(T* a, T v, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) a[i] = v;
}

// a[d * stride] = v for d < D, and out[d] = a[d * stride]
__global__ void k_set_strided(double* a, i64 stride, int D, double v) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < D) a[(i64)d * stride] = v;
}
__global__ void k_get_strided(const double* a, i64 stride, int D, double* out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < D) out[d] = a[(i64)d * stride];
}

// Warm-start corrections at the observed cells of one slot: x2[obs] += dx, z[obs] += dz.
template <typename T>
__global__ void k_add_obs(T* x2, T* z, const i64* __restrict__ obs, const double* __restrict__ dx,
                          const double* __restrict__ dz, i64 M) {
  const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const i64 q = obs[j];
  if (dx) x2[q] = (T)((double)x2[q] + dx[j]);
  if (dz) z[q] = (T)((double)z[q] + dz[j]);
}

// x2 and z at the observed cells of one slot, as float64.
template <typename T>
__global__ void k_get_obs(const T* x2, const T* z, const i64* __restrict__ obs, double* ox, double* oz, i64 M) {
  const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  ox[j] = (double)x2[obs[j]];
  oz[j] = (double)z[obs[j]];
}

// inv_gl[rest] = 1 / (gamma + prod over bits of evals[bit]) in bit order (solver.py:207-213).
template <typename T>
__global__ void k_inv_gl(T* out, int nb, double e0, double e1, double gamma) {
  const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (1ll << nb)) return;
  double lam = 1.0;
  for (int bit = 0; bit < nb; ++bit) lam *= ((r >> bit) & 1) ? e1 : e0;
  out[r] = (T)(1.0 / (gamma + lam));
}

// ---------------------------------------------------------------------------------------------
// Certificates (float64).
// outwt64 (arcflow_kernels.py:357-365): node out-weight, x ascending.
__global__ void k_outw(const double* __restrict__ w, double* __restrict__ outw, int n, i64 Nloc) {
  w += (i64)blockIdx.y * n * Nloc;
  outw += (i64)blockIdx.y << n;
  const i64 D = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (D >= (1ll << n)) return;
  double s = 0.0;
  for (int x = 0; x < n; ++x) {
    if (!((D >> x) & 1)) s += w[(i64)x * Nloc + (i64)compress_bit((u64)D, x)];
  }
  outw[D] = s;
}

// massg64 (arcflow_kernels.py:367-376): one cardinality layer of the forward push.
__global__ void k_mass(const double* __restrict__ w, const double* __restrict__ outw, double* mass,
                       const int* __restrict__ nodes, i64 ns, int n, i64 Nloc) {
  w += (i64)blockIdx.y * n * Nloc;
  outw += (i64)blockIdx.y << n;
  mass += (i64)blockIdx.y << n;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const i64 C = nodes[i];
  double s = 0.0;
  for (int x = 0; x < n; ++x) {
    if ((C >> x) & 1) {
      const i64 P = C ^ (1ll << x);
      const double ow = outw[P];
      if (ow > 0.0) s += mass[P] * w[(i64)x * Nloc + (i64)compress_bit((u64)P, x)] / ow;
    }
  }
  mass[C] = s;
}

// lpath64 (arcflow_kernels.py:388-397): one layer of the longest path.
__global__ void k_lpath(const double* __restrict__ R, double* dp, const int* __restrict__ nodes, i64 ns, int n,
                        i64 Nloc) {
  R += (i64)blockIdx.y * n * Nloc;
  dp += (i64)blockIdx.y << n;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const i64 C = nodes[i];
  double best = -1e300;
  for (int x = 0; x < n; ++x) {
    if ((C >> x) & 1) {
      const i64 P = C ^ (1ll << x);
      const double v = dp[P] + R[(i64)x * Nloc + (i64)compress_bit((u64)P, x)];
      if (v > best) best = v;
    }
  }
  dp[C] = best;
}

// The same three layer kernels with n fixed at compile time: the member loop unrolls and every
// load is unconditional (a non-member x reads the values at the node itself, a valid index, and
// is masked out afterwards), so all loads of a node are in flight before the first is used (the
// runtime-n loops wait for each member's loads in turn, and k_mass waited for outw before loading
// the rest). Same arithmetic, same order of the sums and of the max, so the results are bit for
// bit those of the kernels above.
template <int N>
__global__ void k_outw_n(const double* __restrict__ w, double* __restrict__ outw, i64 Nloc) {
  w += (i64)blockIdx.y * N * Nloc;
  outw += (i64)blockIdx.y << N;
  const i64 D = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (D >= (1ll << N)) return;
  double v[N];
#pragma unroll
  for (int x = 0; x < N; ++x) v[x] = w[(i64)x * Nloc + (i64)compress_bit((u64)D, x)];
  double s = 0.0;
#pragma unroll
  for (int x = 0; x < N; ++x)
    if (!((D >> x) & 1)) s += v[x];
  outw[D] = s;
}

template <int N>
__global__ void k_mass_n(const double* __restrict__ w, const double* __restrict__ outw, double* mass,
                         const int* __restrict__ nodes, i64 ns, i64 Nloc) {
  w += (i64)blockIdx.y * N * Nloc;
  outw += (i64)blockIdx.y << N;
  mass += (i64)blockIdx.y << N;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const i64 C = nodes[i];
  double ow[N], mv[N], wv[N];
#pragma unroll
  for (int x = 0; x < N; ++x) {
    const i64 P = ((C >> x) & 1) ? C ^ (1ll << x) : C;
    ow[x] = outw[P];
    mv[x] = mass[P];
    wv[x] = w[(i64)x * Nloc + (i64)compress_bit((u64)P, x)];
  }
  double s = 0.0;
#pragma unroll
  for (int x = 0; x < N; ++x)
    if (((C >> x) & 1) && ow[x] > 0.0) s += mv[x] * wv[x] / ow[x];
  mass[C] = s;
}

template <int N>
__global__ void k_lpath_n(const double* __restrict__ R, double* dp, const int* __restrict__ nodes, i64 ns,
                          i64 Nloc) {
  R += (i64)blockIdx.y * N * Nloc;
  dp += (i64)blockIdx.y << N;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const i64 C = nodes[i];
  double v[N];
#pragma unroll
  for (int x = 0; x < N; ++x) {
    const i64 P = ((C >> x) & 1) ? C ^ (1ll << x) : C;
    v[x] = dp[P] + R[(i64)x * Nloc + (i64)compress_bit((u64)P, x)];
  }
  double best = -1e300;
#pragma unroll
  for (int x = 0; x < N; ++x)
    if (((C >> x) & 1) && v[x] > best) best = v[x];
  dp[C] = best;
}

// launchers: the compile-time-n kernels for n = 2..26, the runtime-n ones otherwise
#define RG_N_CASES(M)                                                                                    \
  M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12) M(13) M(14) M(15) M(16) M(17) M(18) M(19)  \
  M(20) M(21) M(22) M(23) M(24) M(25) M(26)
static void launch_outw(int n, dim3 grid, hipStream_t st, const double* w, double* outw, i64 Nloc) {
  switch (n) {
#define RG_C(NV) case NV: hipLaunchKernelGGL(k_outw_n<NV>, grid, dim3(256), 0, st, w, outw, Nloc); return;
    RG_N_CASES(RG_C)
#undef RG_C
    default: hipLaunchKernelGGL(k_outw, grid, dim3(256), 0, st, w, outw, n, Nloc);
  }
}
static void launch_mass(int n, dim3 grid, hipStream_t st, const double* w, const double* outw, double* mass,
                        const int* nodes, i64 ns, i64 Nloc) {
  switch (n) {
#define RG_C(NV) case NV: hipLaunchKernelGGL(k_mass_n<NV>, grid, dim3(256), 0, st, w, outw, mass, nodes, ns, Nloc); return;
    RG_N_CASES(RG_C)
#undef RG_C
    default: hipLaunchKernelGGL(k_mass, grid, dim3(256), 0, st, w, outw, mass, nodes, ns, n, Nloc);
  }
}
static void launch_lpath(int n, dim3 grid, hipStream_t st, const double* R, double* dp, const int* nodes, i64 ns,
                         i64 Nloc) {
  switch (n) {
#define RG_C(NV) case NV: hipLaunchKernelGGL(k_lpath_n<NV>, grid, dim3(256), 0, st, R, dp, nodes, ns, Nloc); return;
    RG_N_CASES(RG_C)
#undef RG_C
    default: hipLaunchKernelGGL(k_lpath, grid, dim3(256), 0, st, R, dp, nodes, ns, n, Nloc);
  }
}
#undef RG_N_CASES

// Compensated (Neumaier) block sums. Each block writes (sum, comp) pairs per quantity.
__device__ __forceinline__ void two_sum(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v; else c += (v - t) + s;
  s = t;
}

template <int Q>
__device__ __forceinline__ void block_reduce_store(double (&s)[Q], double (&c)[Q], double* part) {
  __shared__ double sh[2 * Q][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Q; ++q) { sh[2 * q][t] = s[q]; sh[2 * q + 1][t] = c[q]; }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double ss = sh[2 * q][t], cc = sh[2 * q + 1][t] + sh[2 * q + 1][t + h];
        two_sum(ss, cc, sh[2 * q][t + h]);
        sh[2 * q][t] = ss;
        sh[2 * q + 1][t] = cc;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;     // (slot, block)
#pragma unroll
    for (int q = 0; q < Q; ++q) { part[b * 2 * Q + 2 * q] = sh[2 * q][0]; part[b * 2 * Q + 2 * q + 1] = sh[2 * q + 1][0]; }
  }
}

// phi: sum (x[obs] - p)^2; optionally the gathered x[obs] (float64) into fit.
template <typename T>
__global__ __launch_bounds__(256) void k_phi(const T* __restrict__ x, const i64* __restrict__ obs,
                                             const double* __restrict__ p, i64 M, double* fit, double* part,
                                             i64 m2slot) {
  x += (i64)blockIdx.y * m2slot;
  p += (i64)blockIdx.y * M;
  if (fit) fit += (i64)blockIdx.y * M;
  double s[1] = {0.0}, c[1] = {0.0};
  for (i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (i64)gridDim.x * blockDim.x) {
    const double v = (double)x[obs[j]];
    if (fit) fit[j] = v;
    const double d = v - p[j];
    two_sum(s[0], c[0], d * d);
  }
  block_reduce_store<1>(s, c, part);
}

// LB preparation: r = p - y at the observed rows, g[obs] = r (g zero-filled before),
// f = sum r^2, ry = sum r y.
template <typename T>
__global__ __launch_bounds__(256) void k_lb_prep(const T* __restrict__ x, const i64* __restrict__ obs,
                                                 const double* __restrict__ p, i64 M, double* g, double* part,
                                                 i64 m2slot) {
  x += (i64)blockIdx.y * m2slot;
  g += (i64)blockIdx.y * m2slot;
  p += (i64)blockIdx.y * M;
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  for (i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (i64)gridDim.x * blockDim.x) {
    const double y = (double)x[obs[j]];
    const double r = p[j] - y;
    g[obs[j]] = r;
    two_sum(s[0], c[0], r * r);
    two_sum(s[1], c[1], r * y);
  }
  block_reduce_store<2>(s, c, part);
}

// The same with y given per row (float64) instead of read from a slab.
__global__ __launch_bounds__(256) void k_lb_rows(const double* __restrict__ yv, const i64* __restrict__ obs,
                                                 const double* __restrict__ p, i64 M, double* g, double* part) {
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  for (i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (i64)gridDim.x * blockDim.x) {
    const double y = yv[j];
    const double r = p[j] - y;
    g[obs[j]] = r;
    two_sum(s[0], c[0], r * r);
    two_sum(s[1], c[1], r * y);
  }
  block_reduce_store<2>(s, c, part);
}

// Final reduction of block partials: out[q] = sum over blocks of (sum + comp).
// nblk <= 256 partial pairs per quantity, Q <= 2 quantities (every caller). The block's 64 threads
// stage the partials in LDS first, so the serial compensated sums below, the same adds in the same
// block order as a loop over global memory, do not wait on a load per step.
__global__ __launch_bounds__(64) void k_final(const double* part, int nblk, int Q, double* out) {
  __shared__ double sp[256 * 2 * 2];
  part += (size_t)blockIdx.x * nblk * 2 * Q;        // slot = blockIdx.x
  out += (size_t)blockIdx.x * Q;
  const int tot = nblk * 2 * Q;
  for (int i = threadIdx.x; i < tot; i += 64) sp[i] = part[i];
  __syncthreads();
  const int q = threadIdx.x;
  if (q >= Q) return;
  double s = 0.0, c = 0.0;
  for (int b = 0; b < nblk; ++b) {
    two_sum(s, c, sp[b * 2 * Q + 2 * q]);
    c += sp[b * 2 * Q + 2 * q + 1];
  }
  out[q] = s + c;
}

// ---------------------------------------------------------------------------------------------
// Host side.
namespace {

// Start and stop events for the next pass launch (the per-pass profiler, rumgpu_bm_profile): with
// them the launch records both events in its own dispatch (hipExtLaunchKernelGGL), so the interval
// is the kernel's execution, not the gaps between eager launches. Consumed by that launch.
thread_local hipEvent_t g_pass_ev0 = nullptr, g_pass_ev1 = nullptr;

template <typename F, typename... Args>
void pass_launch(F kernel, const dim3& grid, const dim3& block, size_t shm, hipStream_t st, Args... args) {
  if (g_pass_ev0) {
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)shm, st, g_pass_ev0, g_pass_ev1, 0, args...);
    g_pass_ev0 = g_pass_ev1 = nullptr;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shm, st, args...);
  }
}

// A fused pass; with P.nt the non-temporal variant, which exists for float64 on the geometries of
// slabs too large for the caches (8-13 bits at bit 0, 4-8 bits with 16 columns above: single draws
// from n = 20 on and batches of draw slots; and the runtime-geometry kernels); elsewhere P.nt is
// ignored (the results are the same either way).
template <typename T, int RL, int KIND, int BT = 0, int CT = -1>
void launch_bm(const PassArgs<T>& P, const dim3& grid, const dim3& block, size_t shm, hipStream_t st) {
  if constexpr (sizeof(T) == 8 && (BT == 0 || (CT == 0 && BT >= 8) || (CT == 4 && BT >= 4))) {
    if (P.nt) {
      pass_launch((bm_pass<T, RL, KIND, BT, CT, true>), grid, block, shm, st, P);
      return;
    }
  }
  pass_launch((bm_pass<T, RL, KIND, BT, CT, false>), grid, block, shm, st, P);
}

template <typename T, int KIND>
int launch_pass_rl(int RL, const PassArgs<T>& P, hipStream_t st) {
  const WinGeom& g = P.g;
  const size_t shm = (size_t)g.TPB * tile_lds_elems(g) * sizeof(T);
  const dim3 grid((unsigned)((g.ntiles + g.TPB - 1) / g.TPB)), block((unsigned)(g.TT * g.TPB));
  // the window at bit 0 of the fused passes with 3 register bits: geometry fixed at compile time
  if constexpr (KIND == K_FIRST || KIND == K_SY || KIND == K_SYONLY || KIND == K_D || KIND == K_FULL ||
                KIND == K_FULL_FIRST) {
    // 13 bits at bit 0 (4 or 5 register bits, 512 or 256 threads): with 8 bits above it, two windows cover
    // n = 22; the float64 tile needs 68 KiB of LDS, which gfx950's 160 KiB per workgroup allows
    if ((RL == 4 || RL == 5) && g.s == 0 && g.C == 0 && g.B == 13) {
      if (RL == 4) launch_bm<T, 4, KIND, 13, 0>(P, grid, block, shm, st);
      else launch_bm<T, 5, KIND, 13, 0>(P, grid, block, shm, st);
      RG_LAUNCH_CHECK();
      return RG_OK;
    }
    // 4 register bits on 10-12 bits at bit 0: one LDS exchange fewer per transform than 3
    if (RL == 4 && g.s == 0 && g.C == 0 && g.B >= 10 && g.B <= 12) {
      switch (g.B) {
        case 10: launch_bm<T, 4, KIND, 10, 0>(P, grid, block, shm, st); break;
        case 11: launch_bm<T, 4, KIND, 11, 0>(P, grid, block, shm, st); break;
        default: launch_bm<T, 4, KIND, 12, 0>(P, grid, block, shm, st); break;
      }
      RG_LAUNCH_CHECK();
      return RG_OK;
    }
    if (RL == 3 && g.s == 0 && g.C == 0 && g.B >= 8 && g.B <= 12) {
      switch (g.B) {
        case 8: launch_bm<T, 3, KIND, 8, 0>(P, grid, block, shm, st); break;
        case 9: launch_bm<T, 3, KIND, 9, 0>(P, grid, block, shm, st); break;
        case 10: launch_bm<T, 3, KIND, 10, 0>(P, grid, block, shm, st); break;
        case 11: launch_bm<T, 3, KIND, 11, 0>(P, grid, block, shm, st); break;
        default: launch_bm<T, 3, KIND, 12, 0>(P, grid, block, shm, st); break;
      }
      RG_LAUNCH_CHECK();
      return RG_OK;
    }
  }
  // the window above it (16 float64 or 32 float32 columns) of the fused passes, likewise
  if constexpr (KIND == K_RHS || KIND == K_XK) {
    constexpr int CW = sizeof(T) == 8 ? 4 : 5;
    if ((RL == 3 || RL == 4) && g.C == CW && g.B >= 4 && g.B <= 8) {
#define RG_HI_CASE(BB)                                                                                   \
  case BB:                                                                                               \
    if (RL == 3) launch_bm<T, 3, KIND, BB, CW>(P, grid, block, shm, st);             \
    else launch_bm<T, 4, KIND, BB, CW>(P, grid, block, shm, st);                     \
    break;
      switch (g.B) {
        RG_HI_CASE(4) RG_HI_CASE(5) RG_HI_CASE(6) RG_HI_CASE(7)
        default: RG_HI_CASE(8)
      }
#undef RG_HI_CASE
      RG_LAUNCH_CHECK();
      return RG_OK;
    }
  }
  switch (RL) {
    case 1: launch_bm<T, 1, KIND>(P, grid, block, shm, st); break;
    case 2: launch_bm<T, 2, KIND>(P, grid, block, shm, st); break;
    case 3: launch_bm<T, 3, KIND>(P, grid, block, shm, st); break;
    case 4: launch_bm<T, 4, KIND>(P, grid, block, shm, st); break;
    case 5: launch_bm<T, 5, KIND>(P, grid, block, shm, st); break;
    default: rg_set_error("bad RL %d", RL); return RG_EINVAL;
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

template <typename T>
int launch_pass(int kind, int RL, const PassArgs<T>& P, hipStream_t st) {
  switch (kind) {
    case K_FIRST: return launch_pass_rl<T, K_FIRST>(RL, P, st);
    case K_SY: return launch_pass_rl<T, K_SY>(RL, P, st);
    case K_SYONLY: return launch_pass_rl<T, K_SYONLY>(RL, P, st);
    case K_RHS: return launch_pass_rl<T, K_RHS>(RL, P, st);
    case K_D: return launch_pass_rl<T, K_D>(RL, P, st);
    case K_XK: return launch_pass_rl<T, K_XK>(RL, P, st);
    case K_FULL: return launch_pass_rl<T, K_FULL>(RL, P, st);
    case K_FULL_FIRST: return launch_pass_rl<T, K_FULL_FIRST>(RL, P, st);
    case K_T: return launch_pass_rl<T, K_T>(RL, P, st);
  }
  rg_set_error("bad pass kind %d", kind);
  return RG_EINVAL;
}

template <typename T, int MODE>
int launch_prox(int n, const T* xk, T* out, T* z, const u32* om, const int* oo, const double* pc, T b2,
                hipStream_t st, int D = 1, i64 m2 = 0, i64 M = 0, int* fresh = nullptr, const int* on = nullptr) {
  const i64 NN = 1ll << n;
  const dim3 grid(grid_for(NN, 256), (unsigned)D), block(256);
#define RG_PROX_CASE(NV)                                                                                   \
  case NV:                                                                                                 \
    if constexpr (MODE == 0)                                                                               \
      pass_launch((bm_prox_iter<T, NV>), grid, block, 0, st, xk, out, z, om, oo, pc, b2, m2, M, fresh, on); \
    else                                                                                                   \
      hipLaunchKernelGGL((bm_prox<T, NV, MODE>), grid, block, 0, st, xk, out, z, om, oo, pc, b2, m2, M);  \
    break;
  switch (n) {
    RG_PROX_CASE(2) RG_PROX_CASE(3) RG_PROX_CASE(4) RG_PROX_CASE(5) RG_PROX_CASE(6) RG_PROX_CASE(7)
    RG_PROX_CASE(8) RG_PROX_CASE(9) RG_PROX_CASE(10) RG_PROX_CASE(11) RG_PROX_CASE(12) RG_PROX_CASE(13)
    RG_PROX_CASE(14) RG_PROX_CASE(15) RG_PROX_CASE(16) RG_PROX_CASE(17) RG_PROX_CASE(18) RG_PROX_CASE(19)
    RG_PROX_CASE(20) RG_PROX_CASE(21) RG_PROX_CASE(22) RG_PROX_CASE(23) RG_PROX_CASE(24) RG_PROX_CASE(25)
    RG_PROX_CASE(26)
    default: rg_set_error("bm_admm prox: n = %d outside 2..26", n); return RG_EUNSUP;
  }
#undef RG_PROX_CASE
  RG_LAUNCH_CHECK();
  return RG_OK;
}

inline unsigned ew_grid(i64 m) {
  i64 g = (m + 255) / 256;
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

}  // namespace

// The BM-ADMM object behind the C-ABI handle.
struct rumgpu_bm {
  rumgpu_ctx* ctx = nullptr;
  int n = 0, nb = 0, isz = 8, plain = 0;
  int D = 1;                  // draw slots: every slab, p, fit and node array holds D blocks
  int* fresh = nullptr;       // per slot: state loaded since the last iteration
  int* on = nullptr;          // per slot: 1 = the iterations advance it (the draw queue idles empty slots)
  int nt = 0;                 // non-temporal slab accesses in the fused passes
  double* dwarm = nullptr;    // 2 M float64: warm-start corrections at the observed cells
  i64 Nloc = 0, m2 = 0, M = 0, NN = 0;
  double beta = 1, alpha = 1.95, pp = 1.4142135623730951;
  double e0 = 0, e1 = 0, gamma = 1, Q[4] = {0, 0, 0, 0};
  std::vector<WinPlan> wins;
  std::vector<WinGeom> geoms;
  // dtype slabs
  void *x2 = nullptr, *s = nullptr, *y = nullptr, *z = nullptr, *tb = nullptr, *xk = nullptr, *inv_gl = nullptr;
  // kept state (Session keep / warm)
  void *kx2 = nullptr, *ks = nullptr, *ky = nullptr, *kz = nullptr;
  int kept_valid = 0;
  // plain-schedule work slabs
  void *pw = nullptr, *pAT = nullptr, *prhs = nullptr, *px = nullptr, *pAx = nullptr, *px2n = nullptr;
  // observation structures
  u32* obsmask = nullptr;
  int* obsoff = nullptr;
  double* pc = nullptr;       // p in (menu, ascending alternative) order
  i64* pcpos = nullptr;       // host-side permutation input row -> pc index (device copy unused)
  std::vector<i64> h_pcidx;   // input row j -> index in pc
  i64* obs = nullptr;         // slab position per input row
  double* p = nullptr;        // p per input row
  // certificates
  double* cert = nullptr;     // float64 slab
  double* outw = nullptr;
  double* mass = nullptr;     // also the longest-path dp
  int* order = nullptr;
  std::vector<i64> bounds;
  double* part = nullptr;     // block partials
  double* res = nullptr;      // device scalars
  double* fit = nullptr;      // gathered x at obs (float64, input order)
  double* h_res = nullptr;    // pinned host scalars
  int pending = 0;            // s, y update of the last iteration still owed (tb holds partial A x)
  i64 iters = 0;
  // per-pass event profiler (rumgpu_bm_profile): events recorded between the launches of the
  // fused passes, pass kind of each interval
  int prof_on = 0;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_kind;
  size_t prof_n = 0;
  // graph of `graph_len` steady-state iterations
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  i64 graph_len = 0;
  int graph_gen = 0;          // bumped whenever the iteration graph is rebuilt (beta, plan, length)
  struct BmQueue* queue = nullptr;   // the device draw queue of a batch object (rumgpu_bm_queue_*)
};

static void queue_free(rumgpu_bm* b, bool reset_on = true);

static void bm_free_all(rumgpu_bm* b) {
  void* bufs[] = {b->x2, b->s, b->y, b->z, b->tb, b->xk, b->inv_gl, b->kx2, b->ks, b->ky, b->kz,
                  b->pw, b->pAT, b->prhs, b->px, b->pAx, b->px2n, b->obsmask, b->obsoff, b->pc, b->obs,
                  b->p, b->cert, b->outw, b->mass, b->order, b->part, b->res, b->fit, b->fresh, b->dwarm, b->on};
  queue_free(b, false);
  for (void* q : bufs) rg_free(q);
  if (b->h_res) (void)hipHostFree(b->h_res);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->prof_ev) (void)hipEventDestroy(e);
}

// profiler kinds: 0 SY/FIRST, 1 RHS, 2 D, 3 XK, 4 prox, 5 middle-window transform, 6 FULL (one window)
// arm the next pass launch with an event pair charged to `kind`
static inline void prof_arm(rumgpu_bm* b, int kind) {
  if (!b->prof_on || b->prof_n + 2 > b->prof_ev.size()) return;
  g_pass_ev0 = b->prof_ev[b->prof_n];
  g_pass_ev1 = b->prof_ev[b->prof_n + 1];
  b->prof_kind[b->prof_n] = kind;
  b->prof_n += 2;
}

template <typename T>
static PassArgs<T> pass_args(rumgpu_bm* b, int w) {
  PassArgs<T> P;
  P.tb = (T*)b->tb;
  P.s = (T*)b->s;
  P.y = (T*)b->y;
  P.z = (const T*)b->z;
  P.x2 = (const T*)b->x2;
  P.xk = (T*)b->xk;
  P.inv_gl = (const T*)b->inv_gl;
  P.b1 = (T)b->beta;
  P.b2 = (T)b->beta;
  P.ib1 = (T)(1.0 / b->beta);
  P.ib2 = (T)(1.0 / b->beta);
  P.a = (T)b->alpha;
  P.cA = coef<T>(1.0, -1.0, 0.0, b->pp);
  P.cAp = coef<T>(1.0, 0.0, -1.0, b->pp);
  P.cQ = coef<T>(b->Q[0], b->Q[1], b->Q[2], b->Q[3]);
  P.cQt = coef<T>(b->Q[0], b->Q[2], b->Q[1], b->Q[3]);
  P.cT = P.cA;
  P.g = b->geoms[w];
  P.g.on = b->on;
  P.fresh = b->fresh;
  P.nt = b->nt;
  return P;
}

template <typename T>
static int run_pass(rumgpu_bm* b, int kind, int w, const Coef<T>* ct = nullptr) {
  PassArgs<T> P = pass_args<T>(b, w);
  if (ct) P.cT = *ct;
  return launch_pass<T>(kind, b->wins[w].RL, P, b->ctx->stream);
}

// One fused iteration (see the header comment). With the profiler on, every launch carries a start
// and stop event pair charged to its pass kind.
template <typename T>
static int fused_iter(rumgpu_bm* b) {
  const int k = (int)b->wins.size();
  hipStream_t st = b->ctx->stream;
  const PassArgs<T> P0 = pass_args<T>(b, 0);
  if (k == 1) {
    prof_arm(b, 6);
    RG_TRY(run_pass<T>(b, b->pending ? K_FULL : K_FULL_FIRST, 0));
  } else {
    prof_arm(b, 0);
    RG_TRY(run_pass<T>(b, b->pending ? K_SY : K_FIRST, 0));
    for (int w = 1; w < k - 1; ++w) { prof_arm(b, 5); RG_TRY(run_pass<T>(b, K_T, w, &P0.cAp)); }
    prof_arm(b, 1);
    RG_TRY(run_pass<T>(b, K_RHS, k - 1));
    for (int w = k - 2; w >= 1; --w) { prof_arm(b, 5); RG_TRY(run_pass<T>(b, K_T, w, &P0.cQt)); }
    prof_arm(b, 2);
    RG_TRY(run_pass<T>(b, K_D, 0));
    for (int w = 1; w < k - 1; ++w) { prof_arm(b, 5); RG_TRY(run_pass<T>(b, K_T, w, &P0.cQ)); }
    prof_arm(b, 3);
    RG_TRY(run_pass<T>(b, K_XK, k - 1));
  }
  prof_arm(b, 4);
  RG_TRY((launch_prox<T, 0>(b->n, (const T*)b->xk, (T*)b->x2, (T*)b->z, b->obsmask, b->obsoff, b->pc, (T)b->beta, st,
                            b->D, b->m2, b->M, b->fresh, b->on)));
  for (int w = k - 2; w >= 1; --w) { prof_arm(b, 5); RG_TRY(run_pass<T>(b, K_T, w, &P0.cA)); }
  b->pending = 1;
  b->iters += 1;
  return RG_OK;
}

// Settle the deferred s, y update so (x2, s, y, z) is the reference's state.
template <typename T>
static int finalize(rumgpu_bm* b) {
  if (!b->pending) return RG_OK;
  RG_TRY(run_pass<T>(b, K_SYONLY, 0));
  b->pending = 0;
  return RG_OK;
}

// Full transform over every window: dst = T(src), first window from src, the rest in place.
template <typename TIN, typename T>
static int full_transform(rumgpu_bm* b, const TIN* src, T* dst, Coef<T> k, int epi_wmax, int pro_extract,
                          int slots = 1) {
  hipStream_t st = b->ctx->stream;
  std::vector<WinPlan> wins = plan_windows(b->nb, sizeof(T));
  const int nw = (int)wins.size();
  for (int w = 0; w < nw; ++w) {
    const WinGeom g = make_geom(b->n, wins[w], slots);
    const int RL = wins[w].RL;
    const bool last = (w == nw - 1);
    if (w == 0 && pro_extract) {
      if (last && epi_wmax) RG_TRY((launch_tr_rl<double, T, 1, 1>(RL, (const double*)src, dst, k, g, b->outw, b->mass, st)));
      else RG_TRY((launch_tr_rl<double, T, 1, 0>(RL, (const double*)src, dst, k, g, b->outw, b->mass, st)));
    } else if (w == 0) {
      if (last && epi_wmax) RG_TRY((launch_tr_rl<TIN, T, 0, 1>(RL, src, dst, k, g, nullptr, nullptr, st)));
      else RG_TRY((launch_tr_rl<TIN, T, 0, 0>(RL, src, dst, k, g, nullptr, nullptr, st)));
    } else {
      if (last && epi_wmax) RG_TRY((launch_tr_rl<T, T, 0, 1>(RL, dst, dst, k, g, nullptr, nullptr, st)));
      else RG_TRY((launch_tr_rl<T, T, 0, 0>(RL, dst, dst, k, g, nullptr, nullptr, st)));
    }
  }
  return RG_OK;
}

// One iteration in the reference's own schedule (the plain baseline): separate elementwise
// kernels, four full butterflies, the literal z update and a copy.
template <typename T>
static int plain_iter(rumgpu_bm* b) {
  hipStream_t st = b->ctx->stream;
  const i64 m2 = b->m2;
  const unsigned gr = ew_grid(m2);
  T *x2 = (T*)b->x2, *s = (T*)b->s, *y = (T*)b->y, *z = (T*)b->z, *xk = (T*)b->xk;
  T *w = (T*)b->pw, *AT = (T*)b->pAT, *rhs = (T*)b->prhs, *x = (T*)b->px, *Ax = (T*)b->pAx, *x2n = (T*)b->px2n;
  const T b1 = (T)b->beta, b2 = (T)b->beta, ib1 = (T)(1.0 / b->beta), ib2 = (T)(1.0 / b->beta), a = (T)b->alpha;
  hipLaunchKernelGGL(ew_w<T>, dim3(gr), dim3(256), 0, st, y, s, b1, w, m2);
  RG_TRY((full_transform<T, T>(b, w, AT, coef<T>(1.0, 0.0, -1.0, b->pp), 0, 0)));
  hipLaunchKernelGGL(ew_rhs<T>, dim3(gr), dim3(256), 0, st, AT, z, x2, b2, ib1, rhs, m2);
  RG_TRY((full_transform<T, T>(b, rhs, x, coef<T>(b->Q[0], b->Q[2], b->Q[1], b->Q[3]), 0, 0)));
  hipLaunchKernelGGL(ew_diag<T>, dim3(gr), dim3(256), 0, st, (const T*)b->inv_gl, b->Nloc, x, m2);
  RG_TRY((full_transform<T, T>(b, x, x, coef<T>(b->Q[0], b->Q[1], b->Q[2], b->Q[3]), 0, 0)));
  RG_TRY((full_transform<T, T>(b, x, Ax, coef<T>(1.0, -1.0, 0.0, b->pp), 0, 0)));
  hipLaunchKernelGGL(ew_xk<T>, dim3(gr), dim3(256), 0, st, x, x2, z, a, ib2, xk, m2);
  RG_TRY((launch_prox<T, 2>(b->n, xk, x2n, nullptr, b->obsmask, b->obsoff, b->pc, b2, st)));
  hipLaunchKernelGGL(ew_z<T>, dim3(gr), dim3(256), 0, st, x, x2, x2n, a, b2, z, m2);
  hipLaunchKernelGGL(ew_sy<T>, dim3(gr), dim3(256), 0, st, Ax, a, b1, ib1, s, y, m2);
  hipLaunchKernelGGL((ew_copy<T, T>), dim3(gr), dim3(256), 0, st, x2n, x2, m2);
  RG_LAUNCH_CHECK();
  b->iters += 1;
  return RG_OK;
}

template <typename T>
static int do_iters(rumgpu_bm* b, i64 iters) {
  if (b->plain) {
    for (i64 i = 0; i < iters; ++i) RG_TRY(plain_iter<T>(b));
    return RG_OK;
  }
  i64 i = 0;
  if (!b->pending && iters > 0) {
    RG_TRY(fused_iter<T>(b));
    ++i;
  }
  // steady-state iterations replay the captured graph of graph_len iterations
  if (b->gexec && b->graph_len > 0 && !b->prof_on) {
    const i64 reps = (iters - i) / b->graph_len;
    for (i64 r = 0; r < reps; ++r) RG_HIP(hipGraphLaunch(b->gexec, b->ctx->stream));
    i += reps * b->graph_len;
    b->iters += reps * b->graph_len;
  }
  for (; i < iters; ++i) RG_TRY(fused_iter<T>(b));
  return RG_OK;
}

template <typename T>
static int build_graph(rumgpu_bm* b, i64 len) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->graph_len = 0;
  b->graph_gen += 1;
  if (len <= 0 || b->plain) return RG_OK;
  hipStream_t st = b->ctx->stream;
  const int pend = b->pending;
  const i64 it0 = b->iters;
  b->pending = 1;                               // the graph holds steady-state iterations
  RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = RG_OK;
  for (i64 i = 0; i < len && rc == RG_OK; ++i) rc = fused_iter<T>(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  b->pending = pend;
  b->iters = it0;
  if (rc != RG_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) { rg_set_error("graph capture failed: %s", hipGetErrorString(e)); return RG_EHIP; }
  b->graph = g;
  RG_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  b->graph_len = len;
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// Certificates.
// The certificates of slots 0..D-1 (D = 1: the single-solve entry points, which act on slot 0).
template <typename T>
static int phi_launch(rumgpu_bm* b, const void* x, int want_fit, int D, double* out) {
  const int nblk = (int)std::min<i64>(256, std::max<i64>(1, (b->M + 255) / 256));
  hipLaunchKernelGGL(k_phi<T>, dim3(nblk, D), dim3(256), 0, b->ctx->stream, (const T*)x, b->obs, b->p, b->M,
                     want_fit ? b->fit : nullptr, b->part, b->m2);
  hipLaunchKernelGGL(k_final, dim3(D), dim3(64), 0, b->ctx->stream, b->part, nblk, 1, out);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

template <typename T>
static int phi_impl(rumgpu_bm* b, const void* x, double* out_sq, int want_fit, int D = 1) {
  hipStream_t st = b->ctx->stream;
  RG_TRY(phi_launch<T>(b, x, want_fit, D, b->res));
  RG_HIP(hipMemcpyAsync(b->h_res, b->res, (size_t)D * sizeof(double), hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < D; ++d) out_sq[d] = b->h_res[d];
  return RG_OK;
}

// The pushed point's squared distance of every slot into out[D] (device), its gathered values
// into fit: Mobius of x2 in float64, w = max(q, 0) + 1e-12, the layered mass push, extract and
// zeta, then phi of the pushed point. No host synchronisation (the draw queue captures it).
template <typename T>
static int push_launch(rumgpu_bm* b, int D, double* out) {
  hipStream_t st = b->ctx->stream;
  const int n = b->n;
  RG_TRY((full_transform<T, double>(b, (const T*)b->x2, b->cert, coef<double>(1.0, -1.0, 0.0, 1.0), 1, 0, D)));
  launch_outw(n, dim3(grid_for(b->NN, 256), D), st, b->cert, b->outw, b->Nloc);
  hipLaunchKernelGGL(k_set_strided, dim3(grid_for(D, 64)), dim3(64), 0, st, b->mass, b->NN, D, 1.0);
  for (int c = 1; c <= n; ++c) {
    const i64 lo = b->bounds[c], ns = b->bounds[c + 1] - lo;
    if (ns <= 0) continue;
    launch_mass(n, dim3(grid_for(ns, 256), D), st, b->cert, b->outw, b->mass, b->order + lo, ns, b->Nloc);
  }
  RG_LAUNCH_CHECK();
  // rf = extract, x_ub = zeta(rf) (in place)
  RG_TRY((full_transform<double, double>(b, b->cert, b->cert, coef<double>(1.0, 1.0, 0.0, 1.0), 0, 1, D)));
  return phi_launch<double>(b, b->cert, 1, D, out);
}

template <typename T>
static int push_impl(rumgpu_bm* b, double* ub, int D = 1) {
  hipStream_t st = b->ctx->stream;
  RG_TRY(push_launch<T>(b, D, b->res));
  RG_HIP(hipMemcpyAsync(b->h_res, b->res, (size_t)D * sizeof(double), hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < D; ++d) ub[d] = std::sqrt(b->h_res[d]);
  return RG_OK;
}

// The Frank-Wolfe bound's parts of every slot, on the device: res[2d] = f = |r|^2,
// res[2d + 1] = r.y, res[2D + d] = the longest path (vmax) to the full set.
template <typename T>
static int lb_launch(rumgpu_bm* b, int D) {
  hipStream_t st = b->ctx->stream;
  const int n = b->n;
  hipLaunchKernelGGL((ew_fill<double>), dim3(ew_grid(b->m2 * D)), dim3(256), 0, st, b->cert, 0.0, b->m2 * D);
  const int nblk = (int)std::min<i64>(256, std::max<i64>(1, (b->M + 255) / 256));
  hipLaunchKernelGGL(k_lb_prep<T>, dim3(nblk, D), dim3(256), 0, st, (const T*)b->x2, b->obs, b->p, b->M, b->cert,
                     b->part, b->m2);
  hipLaunchKernelGGL(k_final, dim3(D), dim3(64), 0, st, b->part, nblk, 2, b->res);
  RG_LAUNCH_CHECK();
  RG_TRY((full_transform<double, double>(b, b->cert, b->cert, coef<double>(1.0, 0.0, 1.0, 1.0), 0, 0, D)));
  hipLaunchKernelGGL(k_set_strided, dim3(grid_for(D, 64)), dim3(64), 0, st, b->mass, b->NN, D, 0.0);
  for (int c = 1; c <= n; ++c) {
    const i64 lo = b->bounds[c], ns = b->bounds[c + 1] - lo;
    if (ns <= 0) continue;
    launch_lpath(n, dim3(grid_for(ns, 256), D), st, b->cert, b->mass, b->order + lo, ns, b->Nloc);
  }
  hipLaunchKernelGGL(k_get_strided, dim3(grid_for(D, 64)), dim3(64), 0, st, b->mass + (b->NN - 1), b->NN, D,
                     b->res + 2 * D);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

// lb = sqrt(max(0, f - 2 (vmax - r.y))): the same expression on the host and in the draw queue.
__host__ __device__ inline double fw_lb_value(double f, double ry, double vmax) {
  const double rc = vmax - ry;
  const double l2 = f - 2.0 * rc;
  return sqrt(l2 > 0.0 ? l2 : 0.0);
}

template <typename T>
static int lb_impl(rumgpu_bm* b, double* lb, int D = 1) {
  hipStream_t st = b->ctx->stream;
  RG_TRY(lb_launch<T>(b, D));
  RG_HIP(hipMemcpyAsync(b->h_res, b->res, (size_t)3 * D * sizeof(double), hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < D; ++d) lb[d] = fw_lb_value(b->h_res[2 * d], b->h_res[2 * d + 1], b->h_res[2 * D + d]);
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// Draw queue of a batch object: retirement and refill on the device.
//
// The host-driven slots go back to the host after every check (certificates copied out, the
// stop tests run per slot, the next draw's p and start uploaded). Here a cycle is `check` fused
// iterations, the float64 certificates of every slot, a decision kernel and the refill kernels,
// all enqueued on the stream and replayed as one HIP graph (two: with and without the lower
// bound). The decision kernel runs the host driver's stop tests in the same order with the same
// float64 arithmetic, so every draw retires at the same check with the same bracket; the next
// draws come from a ring of pending draws (p, thresholds, optional warm-start corrections) that
// the host tops up between runs of many cycles.

enum { QS_NONE = 0, QS_THRESH = 1, QS_TOL = 2, QS_TIME = 3, QS_ITERS = 4 };
enum { QC_HEAD = 0, QC_TAIL = 1, QC_RETIRED = 2, QC_LIVE = 3, QC_CHECKS = 4, QC_NREF = 5, QC_N = 8 };

struct BmQueue {
  i64 cap = 0, total = 0, tail = 0;   // ring entries; draws of the run; draws pushed so far
  int check = 1, lb_every = 1, corr = 0, primed = 0;
  double abs_tol = 0, rel_tol = 0;
  i64 max_iter = 0;
  rumgpu_bm* src = nullptr;           // the kept state every refill starts from
  double *rP = nullptr, *rDx = nullptr, *rDz = nullptr, *rBelow = nullptr, *rAbove = nullptr;  // ring
  i64* ctr = nullptr;                 // QC_* counters
  i64* sdraw = nullptr;               // per slot: the draw's sequence number, -1 when idle
  i64* sit = nullptr;                 // per slot: iterations of its draw
  double *sub = nullptr, *slb = nullptr, *sbelow = nullptr, *sabove = nullptr;
  int* skill = nullptr;               // per slot: the host's time stop
  i64* refill = nullptr;              // (slot, ring entry) pairs of the last decision
  double* ubsq = nullptr;             // per slot: squared distance of the pushed point
  i64* pcidx = nullptr;               // input row -> p position in (menu, alternative) order
  double *qlb = nullptr, *qub = nullptr;   // per draw at retirement: running max lb, running min ub
  i64 *qit = nullptr, *qorder = nullptr;   // per draw: iterations, retirement rank
  int* qstop = nullptr;               // per draw: QS_* (QS_NONE while open)
  i64* h_ctr = nullptr;               // pinned copy of ctr after each run
  hipGraph_t g[2] = {nullptr, nullptr};
  hipGraphExec_t gx[2] = {nullptr, nullptr};
  int gen = -1;                       // the handle's graph_gen when the cycle graphs were captured
  i64 cycles = 0;                     // cycles run; the lower bound runs when cycles % lb_every == 0
};

struct QueueArgs {
  i64 *ctr, *sdraw, *sit, *refill, *qit, *qorder;
  double *sub, *slb, *sbelow, *sabove, *qlb, *qub;
  const double *rBelow, *rAbove, *ubsq, *res;
  int *skill, *qstop, *on;
  i64 cap, max_iter;
  double abs_tol, rel_tol;
  int D, check;
};

// numpy's minimum / maximum: NaN if either operand is NaN.
__device__ inline double nan_min(double a, double b) {
  return (isnan(a) || isnan(b)) ? __longlong_as_double(0x7ff8000000000000ll) : (b < a ? b : a);
}
__device__ inline double nan_max(double a, double b) {
  return (isnan(a) || isnan(b)) ? __longlong_as_double(0x7ff8000000000000ll) : (b > a ? b : a);
}

// One thread walks the slots in ascending order, as the host driver does. A live slot adds the
// cycle's iterations, folds the certificate into its running bracket and takes the stop tests
// (thresholds, tolerance, the host's time stop, max_iter); a slot that retires (every slot when
// priming) takes the next pending draw, or goes idle.
__global__ void k_queue_decide(QueueArgs a, int want_lb, int prime) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  i64 head = a.ctr[QC_HEAD], retired = a.ctr[QC_RETIRED], nref = 0, live = 0;
  const i64 tail = a.ctr[QC_TAIL];
  for (int d = 0; d < a.D; ++d) {
    const i64 g = a.sdraw[d];
    bool open_slot = prime != 0;
    if (!prime && g >= 0) {
      const i64 it = a.sit[d] + a.check;
      a.sit[d] = it;
      const double ub = nan_min(a.sub[d], sqrt(a.ubsq[d]));
      a.sub[d] = ub;
      double lb = a.slb[d];
      if (want_lb) {
        lb = nan_max(lb, fw_lb_value(a.res[2 * d], a.res[2 * d + 1], a.res[2 * a.D + d]));
        a.slb[d] = lb;
      }
      const double below = a.sbelow[d], above = a.sabove[d];
      int stop = QS_NONE;
      if (!isnan(below) && below > 0 && ub <= below) stop = QS_THRESH;
      else if (!isnan(above) && lb >= above) stop = QS_THRESH;
      else {
        const double gap = ub - lb;
        if (gap <= a.abs_tol || (lb > 0 && gap / lb <= a.rel_tol)) stop = QS_TOL;
        else if (a.skill[d]) stop = QS_TIME;
        else if (it >= a.max_iter) stop = QS_ITERS;
      }
      if (stop != QS_NONE) {
        a.qlb[g] = lb;
        a.qub[g] = ub;
        a.qit[g] = it;
        a.qorder[g] = retired++;
        a.qstop[g] = stop;
        open_slot = true;
      }
    }
    if (open_slot) {
      a.skill[d] = 0;
      if (head < tail) {
        const i64 e = head % a.cap;
        a.sdraw[d] = head++;
        a.sit[d] = 0;
        a.sub[d] = __longlong_as_double(0x7ff0000000000000ll);
        a.slb[d] = 0.0;
        a.sbelow[d] = a.rBelow[e];
        a.sabove[d] = a.rAbove[e];
        a.refill[2 * nref] = d;
        a.refill[2 * nref + 1] = e;
        ++nref;
      } else {
        a.sdraw[d] = -1;
      }
    }
    a.on[d] = a.sdraw[d] >= 0 ? 1 : 0;     // an idle slot's tiles move no data until it is refilled
    if (a.sdraw[d] >= 0) ++live;
  }
  a.ctr[QC_HEAD] = head;
  a.ctr[QC_RETIRED] = retired;
  a.ctr[QC_LIVE] = live;
  a.ctr[QC_NREF] = nref;
  if (!prime) a.ctr[QC_CHECKS] += 1;
}

// p of every refilled slot from its ring entry, in input order and in (menu, alternative) order.
__global__ __launch_bounds__(256) void k_queue_load_p(const i64* __restrict__ ctr, const i64* __restrict__ refill,
                                                      const double* __restrict__ rP, const i64* __restrict__ pcidx,
                                                      double* p, double* pc, i64 M) {
  const i64 total = ctr[QC_NREF] * M;
  for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (i64)gridDim.x * blockDim.x) {
    const i64 k = t / M, j = t - k * M;
    const i64 d = refill[2 * k], e = refill[2 * k + 1];
    const double v = rP[e * M + j];
    p[d * M + j] = v;
    pc[d * M + pcidx[j]] = v;
  }
}

// The kept state into every refilled slot (the conversions of rumgpu_bm_slot_load), marked fresh.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_queue_load_state(const i64* __restrict__ ctr, const i64* __restrict__ refill,
                                                          const TA* __restrict__ kx2, const TA* __restrict__ ks,
                                                          const TA* __restrict__ ky, const TA* __restrict__ kz,
                                                          TB* x2, TB* s, TB* y, TB* z, int* fresh, i64 m2) {
  const i64 total = ctr[QC_NREF] * m2;
  for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (i64)gridDim.x * blockDim.x) {
    const i64 k = t / m2, i = t - k * m2;
    const i64 d = refill[2 * k], o = d * m2 + i;
    x2[o] = (TB)kx2[i];
    s[o] = (TB)ks[i];
    y[o] = (TB)ky[i];
    z[o] = (TB)kz[i];
    if (i == 0) fresh[d] = 1;
  }
}

// Warm-start corrections of every refilled slot at its observed cells (rumgpu_bm_slot_add_obs).
template <typename T>
__global__ __launch_bounds__(256) void k_queue_add_obs(const i64* __restrict__ ctr, const i64* __restrict__ refill,
                                                       T* x2, T* z, const i64* __restrict__ obs,
                                                       const double* __restrict__ rDx, const double* __restrict__ rDz,
                                                       i64 M, i64 m2) {
  const i64 total = ctr[QC_NREF] * M;
  for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (i64)gridDim.x * blockDim.x) {
    const i64 k = t / M, j = t - k * M;
    const i64 d = refill[2 * k], e = refill[2 * k + 1];
    const i64 q = d * m2 + obs[j];
    if (rDx) x2[q] = (T)((double)x2[q] + rDx[e * M + j]);
    if (rDz) z[q] = (T)((double)z[q] + rDz[e * M + j]);
  }
}

__global__ void k_set_i64(i64* a, i64 v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *a = v;
}

static void queue_free(rumgpu_bm* b, bool reset_on) {
  BmQueue* q = b->queue;
  if (!q) return;
  if (reset_on && b->on)      // outside a queue every slot runs
    hipLaunchKernelGGL((ew_fill<int>), dim3(grid_for((i64)b->D, 256)), dim3(256), 0, b->ctx->stream, b->on, 1,
                       (i64)b->D);
  for (int k = 0; k < 2; ++k) {
    if (q->gx[k]) (void)hipGraphExecDestroy(q->gx[k]);
    if (q->g[k]) (void)hipGraphDestroy(q->g[k]);
  }
  void* bufs[] = {q->rP, q->rDx, q->rDz, q->rBelow, q->rAbove, q->ctr, q->sdraw, q->sit, q->sub, q->slb,
                  q->sbelow, q->sabove, q->skill, q->refill, q->ubsq, q->pcidx, q->qlb, q->qub, q->qit,
                  q->qorder, q->qstop};
  for (void* v : bufs) rg_free(v);
  if (q->h_ctr) (void)hipHostFree(q->h_ctr);
  delete q;
  b->queue = nullptr;
}

static QueueArgs queue_args(rumgpu_bm* b) {
  BmQueue* q = b->queue;
  QueueArgs a;
  a.ctr = q->ctr; a.sdraw = q->sdraw; a.sit = q->sit; a.refill = q->refill; a.qit = q->qit; a.qorder = q->qorder;
  a.sub = q->sub; a.slb = q->slb; a.sbelow = q->sbelow; a.sabove = q->sabove; a.qlb = q->qlb; a.qub = q->qub;
  a.rBelow = q->rBelow; a.rAbove = q->rAbove; a.ubsq = q->ubsq; a.res = b->res;
  a.skill = q->skill; a.qstop = q->qstop; a.on = b->on;
  a.cap = q->cap; a.max_iter = q->max_iter;
  a.abs_tol = q->abs_tol; a.rel_tol = q->rel_tol;
  a.D = b->D; a.check = q->check;
  return a;
}

static const unsigned QUEUE_GRID = 1024;

// Decision and refill kernels (enqueued only).
template <typename T>
static int queue_decide_refill(rumgpu_bm* b, int want_lb, int prime) {
  hipStream_t st = b->ctx->stream;
  BmQueue* q = b->queue;
  hipLaunchKernelGGL(k_queue_decide, dim3(1), dim3(64), 0, st, queue_args(b), want_lb, prime);
  hipLaunchKernelGGL(k_queue_load_p, dim3(QUEUE_GRID), dim3(256), 0, st, (const i64*)q->ctr, (const i64*)q->refill,
                     (const double*)q->rP, (const i64*)q->pcidx, b->p, b->pc, b->M);
  rumgpu_bm* s = q->src;
  if (s->isz == 8)
    hipLaunchKernelGGL((k_queue_load_state<double, T>), dim3(QUEUE_GRID), dim3(256), 0, st, (const i64*)q->ctr,
                       (const i64*)q->refill, (const double*)s->kx2, (const double*)s->ks, (const double*)s->ky,
                       (const double*)s->kz, (T*)b->x2, (T*)b->s, (T*)b->y, (T*)b->z, b->fresh, b->m2);
  else
    hipLaunchKernelGGL((k_queue_load_state<float, T>), dim3(QUEUE_GRID), dim3(256), 0, st, (const i64*)q->ctr,
                       (const i64*)q->refill, (const float*)s->kx2, (const float*)s->ks, (const float*)s->ky,
                       (const float*)s->kz, (T*)b->x2, (T*)b->s, (T*)b->y, (T*)b->z, b->fresh, b->m2);
  if (q->corr)
    hipLaunchKernelGGL(k_queue_add_obs<T>, dim3(QUEUE_GRID), dim3(256), 0, st, (const i64*)q->ctr,
                       (const i64*)q->refill, (T*)b->x2, (T*)b->z, (const i64*)b->obs,
                       (const double*)((q->corr & 1) ? q->rDx : nullptr),
                       (const double*)((q->corr & 2) ? q->rDz : nullptr), b->M, b->m2);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

// One cycle: `check` fused iterations, the pushed upper bound (and the lower bound when want_lb)
// of every slot, the decision and the refills. Enqueued only; captured by the cycle graphs.
template <typename T>
static int queue_cycle(rumgpu_bm* b, int want_lb) {
  BmQueue* q = b->queue;
  for (int i = 0; i < q->check; ++i) RG_TRY(fused_iter<T>(b));
  RG_TRY(push_launch<T>(b, b->D, q->ubsq));
  if (want_lb) RG_TRY(lb_launch<T>(b, b->D));
  return queue_decide_refill<T>(b, want_lb, 0);
}

template <typename T>
static int queue_graphs(rumgpu_bm* b) {
  BmQueue* q = b->queue;
  if (q->gen == b->graph_gen && q->gx[0] && q->gx[1]) return RG_OK;
  hipStream_t st = b->ctx->stream;
  for (int k = 0; k < 2; ++k) {
    if (q->gx[k]) { (void)hipGraphExecDestroy(q->gx[k]); q->gx[k] = nullptr; }
    if (q->g[k]) { (void)hipGraphDestroy(q->g[k]); q->g[k] = nullptr; }
  }
  const int pend = b->pending;
  const i64 it0 = b->iters;
  for (int k = 0; k < 2; ++k) {
    b->pending = 1;                             // steady state: every slot owes its s, y update
    RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = queue_cycle<T>(b, k);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    b->pending = pend;
    b->iters = it0;
    if (rc != RG_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { rg_set_error("queue graph capture failed: %s", hipGetErrorString(e)); return RG_EHIP; }
    q->g[k] = g;
    RG_HIP(hipGraphInstantiate(&q->gx[k], g, nullptr, nullptr, 0));
  }
  q->gen = b->graph_gen;
  return RG_OK;
}

template <typename T>
static int queue_run_t(rumgpu_bm* b, i64 cycles) {
  BmQueue* q = b->queue;
  hipStream_t st = b->ctx->stream;
  if (!q->primed) {
    RG_TRY(queue_decide_refill<T>(b, 0, 1));
    q->primed = 1;
  }
  for (i64 c = 0; c < cycles; ++c) {
    const int want_lb = (q->cycles % q->lb_every) == 0;
    if (!b->pending) {
      RG_TRY(queue_cycle<T>(b, want_lb));       // the first iteration after a load runs eagerly
    } else {
      RG_TRY(queue_graphs<T>(b));
      RG_HIP(hipGraphLaunch(q->gx[want_lb], st));
      b->iters += q->check;
    }
    q->cycles += 1;
  }
  RG_HIP(hipMemcpyAsync(q->h_ctr, q->ctr, QC_N * sizeof(i64), hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// C-ABI.
#define BM_DISPATCH(b, fn, ...) ((b)->isz == 8 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

static int bm_create_impl(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, const double* p, int64_t M,
                          int dtype_bytes, double beta, double alpha, double precond_p, double e0, double e1,
                          double gamma, const double* Q, int flags, int slots, rumgpu_bm** out) {
  RG_CHECK(ctx && out && menu && alt && p && Q, RG_EINVAL, "bm_create: null argument");
  RG_CHECK(slots >= 1 && slots <= 65535, RG_EINVAL, "bm_create: slots = %d outside 1..65535", slots);
  RG_CHECK(slots == 1 || !(flags & 1), RG_EINVAL, "bm_create: the plain schedule runs one slot");
  RG_CHECK(n >= 2 && n <= 26, RG_EUNSUP, "bm_admm handles 2 <= n <= 26 here (n = %d)", n);
  RG_CHECK(dtype_bytes == 4 || dtype_bytes == 8, RG_EINVAL, "dtype must be 4 (float32) or 8 (float64) bytes");
  RG_CHECK(M >= 1, RG_EINVAL, "bm_create: no rows");
  RG_CHECK(beta > 0, RG_EINVAL, "beta must be > 0");
  RG_HIP(hipSetDevice(ctx->device));
  rumgpu_bm* b = new rumgpu_bm();
  b->ctx = ctx;
  b->n = n;
  b->nb = n - 1;
  b->isz = dtype_bytes;
  b->plain = (flags & 1) ? 1 : 0;
  b->D = slots;
  b->Nloc = 1ll << (n - 1);
  b->m2 = (i64)n * b->Nloc;
  // a float64 object whose six iteration slabs (all draw slots) pass the 256 MiB Infinity Cache
  // streams them with non-temporal accesses: measured faster for one draw from n = 20 on (n = 19,
  // 239 MB, is slower) and for 16-64 slots at n = 16 (403-1,611 MB)
  b->nt = (dtype_bytes == 8 && 6 * b->m2 * 8 * slots > (256ll << 20)) ? 1 : 0;
  b->NN = 1ll << n;
  b->M = M;
  b->beta = beta;
  b->alpha = alpha;
  b->pp = precond_p;
  b->e0 = e0;
  b->e1 = e1;
  b->gamma = gamma;
  for (int i = 0; i < 4; ++i) b->Q[i] = Q[i];
  b->wins = plan_windows(b->nb, dtype_bytes);
  for (auto& w : b->wins) b->geoms.push_back(make_geom(n, w, slots));
  int rc = RG_OK;
  auto fail = [&](int code) {
    bm_free_all(b);
    delete b;
    return code;
  };
  // host-side observation structures
  std::vector<u32> hmask(b->NN, 0);
  std::vector<int> hoff(b->NN, 0);
  std::vector<i64> hobs(M);
  for (i64 j = 0; j < M; ++j) {
    const u64 m = menu[j];
    const int a = alt[j];
    if (a < 0 || a >= n || !((m >> a) & 1ull) || m >= (1ull << n) || (m & (m - 1)) == 0) {
      rg_set_error("bm_create: row %lld (menu %llu, alternative %d) is not a non-singleton cell of n = %d",
                   (long long)j, (unsigned long long)m, a, n);
      return fail(RG_EINVAL);
    }
    if ((hmask[m] >> a) & 1u) {
      rg_set_error("bm_create: row %lld repeats (menu %llu, alternative %d)", (long long)j, (unsigned long long)m, a);
      return fail(RG_EINVAL);
    }
    hmask[m] |= 1u << a;
    hobs[j] = (i64)a * b->Nloc + (i64)compress_bit(m & ~(1ull << a), a);
  }
  // compact p order: menus ascending, alternatives ascending
  {
    i64 acc = 0;
    for (i64 S = 0; S < b->NN; ++S) {
      hoff[S] = (int)acc;
      acc += __builtin_popcount(hmask[S]);
    }
    b->h_pcidx.resize(M);
    for (i64 j = 0; j < M; ++j) {
      const u64 m = menu[j];
      const int a = alt[j];
      b->h_pcidx[j] = hoff[m] + __builtin_popcount(hmask[m] & ((1u << a) - 1u));
    }
  }
  const size_t T = dtype_bytes;
  const size_t D = (size_t)slots;
  const size_t slab = (size_t)b->m2 * T * D;
  void** slabs[] = {&b->x2, &b->s, &b->y, &b->z, &b->tb, &b->xk};
  for (void** q : slabs)
    if ((rc = rg_malloc(q, slab, "bm_admm state slab")) != RG_OK) return fail(rc);
  if (b->plain) {
    void** ps[] = {&b->pw, &b->pAT, &b->prhs, &b->px, &b->pAx, &b->px2n};
    for (void** q : ps)
      if ((rc = rg_malloc(q, slab, "bm_admm plain-schedule slab")) != RG_OK) return fail(rc);
  }
  if ((rc = rg_malloc(&b->inv_gl, (size_t)b->Nloc * T, "inv_gl")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->obsmask, (size_t)b->NN * 4, "obsmask")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->obsoff, (size_t)b->NN * 4, "obsoff")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->pc, (size_t)M * 8 * D, "p (menu order)")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->obs, (size_t)M * 8, "obs index")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->p, (size_t)M * 8 * D, "p")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->fit, (size_t)M * 8 * D, "p_fit")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->dwarm, (size_t)M * 8 * 2, "warm-start corrections")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->cert, (size_t)b->m2 * 8 * D, "float64 certificate slab")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->outw, (size_t)b->NN * 8 * D, "outw")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->mass, (size_t)b->NN * 8 * D, "mass")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->order, (size_t)b->NN * 4, "layers")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->part, (size_t)256 * 4 * 8 * D, "partials")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->res, (16 + 4 * D) * 8, "scalars")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->fresh, D * 4, "slot flags")) != RG_OK) return fail(rc);
  if ((rc = rg_malloc((void**)&b->on, D * 4, "slot flags")) != RG_OK) return fail(rc);
  if (hipHostMalloc((void**)&b->h_res, (16 + 4 * D) * 8, hipHostMallocDefault) != hipSuccess) {
    rg_set_error("hipHostMalloc of %zu bytes failed", (size_t)(16 + 4 * D) * 8);
    return fail(RG_ENOMEM);
  }
  hipStream_t st = ctx->stream;
  // layers: nodes by popcount, ascending within a layer
  {
    std::vector<int> hord(b->NN);
    b->bounds.assign(n + 2, 0);
    std::vector<i64> cnt(n + 2, 0);
    for (i64 v = 0; v < b->NN; ++v) cnt[__builtin_popcountll(v)]++;
    for (int c = 0; c <= n; ++c) b->bounds[c + 1] = b->bounds[c] + cnt[c];
    std::vector<i64> fill(b->bounds.begin(), b->bounds.end());
    for (i64 v = 0; v < b->NN; ++v) hord[fill[__builtin_popcountll(v)]++] = (int)v;
    if (rg_copy(ctx->stream, b->order, hord.data(), (size_t)b->NN * 4, hipMemcpyHostToDevice) != hipSuccess) return fail(RG_EHIP);
  }
  if (rg_copy(ctx->stream, b->obsmask, hmask.data(), (size_t)b->NN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      rg_copy(ctx->stream, b->obsoff, hoff.data(), (size_t)b->NN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      rg_copy(ctx->stream, b->obs, hobs.data(), (size_t)M * 8, hipMemcpyHostToDevice) != hipSuccess) {
    rg_set_error("bm_create: upload failed");
    return fail(RG_EHIP);
  }
  hipLaunchKernelGGL((ew_fill<int>), dim3(grid_for((i64)D, 256)), dim3(256), 0, st, b->fresh, 0, (i64)D);
  hipLaunchKernelGGL((ew_fill<int>), dim3(grid_for((i64)D, 256)), dim3(256), 0, st, b->on, 1, (i64)D);
  if (dtype_bytes == 8)
    hipLaunchKernelGGL(k_inv_gl<double>, dim3(grid_for(b->Nloc, 256)), dim3(256), 0, st, (double*)b->inv_gl, b->nb, e0, e1, gamma);
  else
    hipLaunchKernelGGL(k_inv_gl<float>, dim3(grid_for(b->Nloc, 256)), dim3(256), 0, st, (float*)b->inv_gl, b->nb, e0, e1, gamma);
  *out = b;
  rc = rumgpu_bm_set_p(b, p);
  if (rc != RG_OK) {
    *out = nullptr;
    return fail(rc);
  }
  rc = rumgpu_bm_init_cold(b);
  if (rc != RG_OK) {
    *out = nullptr;
    return fail(rc);
  }
  return RG_OK;
}

extern "C" int rumgpu_bm_create(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, const double* p,
                                int64_t M, int dtype_bytes, double beta, double alpha, double precond_p, double e0,
                                double e1, double gamma, const double* Q, int flags, rumgpu_bm** out) {
  return bm_create_impl(ctx, n, menu, alt, p, M, dtype_bytes, beta, alpha, precond_p, e0, e1, gamma, Q, flags, 1, out);
}

extern "C" int rumgpu_bm_create_batch(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt,
                                      const double* p, int64_t M, int dtype_bytes, double beta, double alpha,
                                      double precond_p, double e0, double e1, double gamma, const double* Q,
                                      int slots, rumgpu_bm** out) {
  return bm_create_impl(ctx, n, menu, alt, p, M, dtype_bytes, beta, alpha, precond_p, e0, e1, gamma, Q, 0, slots, out);
}

extern "C" int rumgpu_bm_destroy(rumgpu_bm* b) {
  if (!b) return RG_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  bm_free_all(b);
  delete b;
  return RG_OK;
}

extern "C" int rumgpu_bm_set_p(rumgpu_bm* b, const double* p) {
  RG_CHECK(b && p, RG_EINVAL, "bm_set_p: null argument");
  std::vector<double> hpc(b->M);
  for (i64 j = 0; j < b->M; ++j) hpc[b->h_pcidx[j]] = p[j];
  hipStream_t st = b->ctx->stream;
  for (int d = 0; d < b->D; ++d) {
    RG_HIP(hipMemcpyAsync(b->p + (i64)d * b->M, p, (size_t)b->M * 8, hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(b->pc + (i64)d * b->M, hpc.data(), (size_t)b->M * 8, hipMemcpyHostToDevice, st));
  }
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

template <typename T>
static int init_cold_t(rumgpu_bm* b) {
  hipStream_t st = b->ctx->stream;
  const i64 m = b->m2 * b->D;
  const unsigned gr = ew_grid(m);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->s, (T)0, m);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->y, (T)0, m);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->z, (T)0, m);
  hipLaunchKernelGGL((ew_fill<int>), dim3(grid_for((i64)b->D, 256)), dim3(256), 0, st, b->fresh, 0, (i64)b->D);
  RG_LAUNCH_CHECK();
  RG_TRY((launch_prox<T, 1>(b->n, (const T*)b->xk, (T*)b->x2, nullptr, b->obsmask, b->obsoff, b->pc, (T)b->beta, st,
                            b->D, b->m2, b->M)));
  b->pending = 0;
  b->iters = 0;
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_init_cold(rumgpu_bm* b) {
  RG_CHECK(b, RG_EINVAL, "bm_init_cold: null handle");
  return BM_DISPATCH(b, init_cold_t, b);
}

template <typename T>
static int set_state_t(rumgpu_bm* b, const double* x2, const double* s, const double* y, const double* z) {
  hipStream_t st = b->ctx->stream;
  const double* src[4] = {x2, s, y, z};
  void* dst[4] = {b->x2, b->s, b->y, b->z};
  std::vector<T> tmp((size_t)b->m2);
  for (int q = 0; q < 4; ++q) {
    for (i64 i = 0; i < b->m2; ++i) tmp[i] = (T)src[q][i];
    RG_HIP(hipMemcpyAsync(dst[q], tmp.data(), (size_t)b->m2 * sizeof(T), hipMemcpyHostToDevice, st));
    RG_HIP(hipStreamSynchronize(st));
  }
  b->pending = 0;
  return RG_OK;
}

extern "C" int rumgpu_bm_set_state(rumgpu_bm* b, const double* x2, const double* s, const double* y, const double* z) {
  RG_CHECK(b && x2 && s && y && z, RG_EINVAL, "bm_set_state: null argument");
  RG_CHECK(b->D == 1, RG_EINVAL, "bm_set_state: a batch handle loads slots with bm_slot_load");
  return BM_DISPATCH(b, set_state_t, b, x2, s, y, z);
}

template <typename T>
static int get_state_t(rumgpu_bm* b, double* x2, double* s, double* y, double* z) {
  RG_TRY(finalize<T>(b));
  hipStream_t st = b->ctx->stream;
  double* dst[4] = {x2, s, y, z};
  void* src[4] = {b->x2, b->s, b->y, b->z};
  std::vector<T> tmp((size_t)b->m2);
  for (int q = 0; q < 4; ++q) {
    if (!dst[q]) continue;
    RG_HIP(hipMemcpyAsync(tmp.data(), src[q], (size_t)b->m2 * sizeof(T), hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    for (i64 i = 0; i < b->m2; ++i) dst[q][i] = (double)tmp[i];
  }
  return RG_OK;
}

extern "C" int rumgpu_bm_get_state(rumgpu_bm* b, double* x2, double* s, double* y, double* z) {
  RG_CHECK(b, RG_EINVAL, "bm_get_state: null handle");
  return BM_DISPATCH(b, get_state_t, b, x2, s, y, z);
}

// Copy the state of src into dst on the device (any dtypes, same n and rows).
template <typename TA, typename TB>
static int copy_state_t(rumgpu_bm* dst, rumgpu_bm* src) {
  RG_TRY(finalize<TA>(src));
  hipStream_t st = dst->ctx->stream;
  const unsigned gr = ew_grid(dst->m2);
  void* a[4] = {src->x2, src->s, src->y, src->z};
  void* b[4] = {dst->x2, dst->s, dst->y, dst->z};
  for (int q = 0; q < 4; ++q) hipLaunchKernelGGL((ew_copy<TA, TB>), dim3(gr), dim3(256), 0, st, (const TA*)a[q], (TB*)b[q], dst->m2);
  RG_LAUNCH_CHECK();
  dst->pending = 0;
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_copy_state(rumgpu_bm* dst, rumgpu_bm* src) {
  RG_CHECK(dst && src && dst->n == src->n && dst->M == src->M, RG_EINVAL, "bm_copy_state: incompatible handles");
  RG_CHECK(dst->D == 1, RG_EINVAL, "bm_copy_state: a batch handle loads slots with bm_slot_load");
  if (src->isz == 8 && dst->isz == 8) return copy_state_t<double, double>(dst, src);
  if (src->isz == 8 && dst->isz == 4) return copy_state_t<double, float>(dst, src);
  if (src->isz == 4 && dst->isz == 8) return copy_state_t<float, double>(dst, src);
  return copy_state_t<float, float>(dst, src);
}

template <typename T>
static int keep_t(rumgpu_bm* b, int load) {
  RG_TRY(finalize<T>(b));
  hipStream_t st = b->ctx->stream;
  void** live[4] = {&b->x2, &b->s, &b->y, &b->z};
  void** kept[4] = {&b->kx2, &b->ks, &b->ky, &b->kz};
  if (!load) {
    for (int q = 0; q < 4; ++q) {
      if (!*kept[q]) RG_TRY(rg_malloc(kept[q], (size_t)b->m2 * sizeof(T), "kept state slab"));
      RG_HIP(hipMemcpyAsync(*kept[q], *live[q], (size_t)b->m2 * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    b->kept_valid = 1;
  } else {
    RG_CHECK(b->kept_valid, RG_EINVAL, "bm_load_kept: no kept state");
    for (int q = 0; q < 4; ++q)
      RG_HIP(hipMemcpyAsync(*live[q], *kept[q], (size_t)b->m2 * sizeof(T), hipMemcpyDeviceToDevice, st));
    b->pending = 0;
    b->iters = 0;
  }
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_keep(rumgpu_bm* b) {
  RG_CHECK(b, RG_EINVAL, "bm_keep: null handle");
  return BM_DISPATCH(b, keep_t, b, 0);
}
extern "C" int rumgpu_bm_load_kept(rumgpu_bm* b) {
  RG_CHECK(b, RG_EINVAL, "bm_load_kept: null handle");
  RG_CHECK(b->D == 1, RG_EINVAL, "bm_load_kept: a batch handle loads slots with bm_slot_load");
  return BM_DISPATCH(b, keep_t, b, 1);
}

extern "C" int rumgpu_bm_run(rumgpu_bm* b, int64_t iters) {
  RG_CHECK(b && iters >= 0, RG_EINVAL, "bm_run: bad argument");
  RG_HIP(hipSetDevice(b->ctx->device));
  int rc = BM_DISPATCH(b, do_iters, b, iters);
  if (rc != RG_OK) return rc;
  RG_HIP(hipStreamSynchronize(b->ctx->stream));
  return RG_OK;
}

extern "C" int rumgpu_bm_set_graph(rumgpu_bm* b, int64_t len) {
  RG_CHECK(b && len >= 0, RG_EINVAL, "bm_set_graph: bad argument");
  return BM_DISPATCH(b, build_graph, b, len);
}

// A window plan for the fused passes (the certificate transforms keep the default plan): the
// owed s, y update is settled first, since the pass buffer holds a transform split by the old plan.
extern "C" int rumgpu_bm_set_plan(rumgpu_bm* b, int lo_bits, int lo_rl, int hi_rl) {
  RG_CHECK(b, RG_EINVAL, "bm_set_plan: null handle");
  RG_CHECK(!b->plain, RG_EINVAL, "bm_set_plan: the plain schedule has no window plan");
  RG_CHECK(lo_bits >= 1 && lo_bits <= 13 && lo_rl >= 1 && lo_rl <= 5 && hi_rl >= 1 && hi_rl <= 5, RG_EINVAL,
           "bm_set_plan: lo_bits 1..13, register bits 1..5");
  std::vector<WinPlan> wins = plan_windows_explicit(b->nb, b->isz, lo_bits, lo_rl, hi_rl);
  std::vector<WinGeom> geoms;
  for (auto& w : wins) {
    const WinGeom g = make_geom(b->n, w, b->D);
    const size_t lds = (size_t)g.TPB * tile_lds_elems(g) * b->isz;
    RG_CHECK(g.TT * g.TPB <= 512 && lds <= b->ctx->lds_per_block, RG_EINVAL,
             "bm_set_plan: a tile of %d bits with %d register bits needs %d threads and %zu bytes of LDS per block",
             w.B + w.C, w.RL, g.TT * g.TPB, lds);
    geoms.push_back(g);
  }
  RG_TRY(BM_DISPATCH(b, finalize, b));
  b->wins = wins;
  b->geoms = geoms;
  b->graph_gen += 1;
  if (b->gexec) return BM_DISPATCH(b, build_graph, b, b->graph_len);
  return RG_OK;
}

// A new penalty for the next iterations (b1 = b2 = beta, so gamma = b2/b1 and inv_gl stay). The
// multipliers y, z are unscaled, so the state carries over as it is once the deferred s, y update
// is settled with the old penalty; a captured graph is rebuilt.
extern "C" int rumgpu_bm_set_beta(rumgpu_bm* b, double beta) {
  RG_CHECK(b && beta > 0 && std::isfinite(beta), RG_EINVAL, "bm_set_beta: beta must be finite and > 0");
  if (beta == b->beta) return RG_OK;
  RG_TRY(BM_DISPATCH(b, finalize, b));       // the owed s, y update still uses the old penalty
  b->beta = beta;
  b->graph_gen += 1;
  if (b->gexec) return BM_DISPATCH(b, build_graph, b, b->graph_len);
  return RG_OK;
}

extern "C" int rumgpu_bm_finalize(rumgpu_bm* b) {
  RG_CHECK(b, RG_EINVAL, "bm_finalize: null handle");
  RG_TRY(BM_DISPATCH(b, finalize, b));
  RG_HIP(hipStreamSynchronize(b->ctx->stream));
  return RG_OK;
}

extern "C" int rumgpu_bm_phi(rumgpu_bm* b, double* phi) {
  RG_CHECK(b && phi, RG_EINVAL, "bm_phi: null argument");
  double sq = 0.0;
  RG_TRY(b->isz == 8 ? phi_impl<double>(b, b->x2, &sq, 0) : phi_impl<float>(b, b->x2, &sq, 0));
  *phi = std::sqrt(sq);
  return RG_OK;
}

extern "C" int rumgpu_bm_push_ub64(rumgpu_bm* b, double* ub, double* p_fit) {
  RG_CHECK(b && ub, RG_EINVAL, "bm_push_ub64: null argument");
  RG_TRY(BM_DISPATCH(b, push_impl, b, ub));
  if (p_fit) {
    RG_HIP(rg_copy(b->ctx->stream, p_fit, b->fit, (size_t)b->M * 8, hipMemcpyDeviceToHost));
  }
  return RG_OK;
}

extern "C" int rumgpu_bm_get_fit(rumgpu_bm* b, double* p_fit) {
  RG_CHECK(b && p_fit, RG_EINVAL, "bm_get_fit: null argument");
  RG_HIP(rg_copy(b->ctx->stream, p_fit, b->fit, (size_t)b->M * 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

extern "C" int rumgpu_bm_lb64(rumgpu_bm* b, double* lb) {
  RG_CHECK(b && lb, RG_EINVAL, "bm_lb64: null argument");
  return BM_DISPATCH(b, lb_impl, b, lb);
}

extern "C" int rumgpu_bm_x_obs(rumgpu_bm* b, double* out) {
  RG_CHECK(b && out, RG_EINVAL, "bm_x_obs: null argument");
  double sq = 0.0;
  RG_TRY(b->isz == 8 ? phi_impl<double>(b, b->x2, &sq, 1) : phi_impl<float>(b, b->x2, &sq, 1));
  RG_HIP(rg_copy(b->ctx->stream, out, b->fit, (size_t)b->M * 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

extern "C" int rumgpu_bm_info(rumgpu_bm* b, int64_t* iters, int* nwin, int* win_bits) {
  RG_CHECK(b, RG_EINVAL, "bm_info: null handle");
  if (iters) *iters = b->iters;
  if (nwin) *nwin = (int)b->wins.size();
  if (win_bits)
    for (size_t i = 0; i < b->wins.size(); ++i) win_bits[i] = b->wins[i].B;
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// Draw slots of a batch handle. Slot d owns block d of every slab; the iterations run every slot
// at once (rumgpu_bm_run), each slot is loaded, corrected and certified on its own.
#define RG_SLOT_CHECK(b, slot, what)                                                                    \
  RG_CHECK((b) && (slot) >= 0 && (slot) < (b)->D, RG_EINVAL, what ": slot %d outside 0..%d", (int)(slot), \
           (b) ? (b)->D - 1 : -1)

extern "C" int rumgpu_bm_slot_set_p(rumgpu_bm* b, int slot, const double* p) {
  RG_SLOT_CHECK(b, slot, "bm_slot_set_p");
  RG_CHECK(p, RG_EINVAL, "bm_slot_set_p: null p");
  std::vector<double> hpc(b->M);
  for (i64 j = 0; j < b->M; ++j) hpc[b->h_pcidx[j]] = p[j];
  hipStream_t st = b->ctx->stream;
  RG_HIP(hipMemcpyAsync(b->p + (i64)slot * b->M, p, (size_t)b->M * 8, hipMemcpyHostToDevice, st));
  RG_HIP(hipMemcpyAsync(b->pc + (i64)slot * b->M, hpc.data(), (size_t)b->M * 8, hipMemcpyHostToDevice, st));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

static int mark_fresh(rumgpu_bm* b, int slot) {
  hipLaunchKernelGGL((ew_fill<int>), dim3(1), dim3(64), 0, b->ctx->stream, b->fresh + slot, 1, (i64)1);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

template <typename TA, typename TB>
static int slot_load_t(rumgpu_bm* dst, int slot, rumgpu_bm* src) {
  hipStream_t st = dst->ctx->stream;
  const unsigned gr = ew_grid(dst->m2);
  const i64 off = (i64)slot * dst->m2 * (i64)sizeof(TB);
  void* a[4] = {src->kx2, src->ks, src->ky, src->kz};
  void* b[4] = {dst->x2, dst->s, dst->y, dst->z};
  for (int q = 0; q < 4; ++q)
    hipLaunchKernelGGL((ew_copy<TA, TB>), dim3(gr), dim3(256), 0, st, (const TA*)a[q], (TB*)((char*)b[q] + off), dst->m2);
  RG_LAUNCH_CHECK();
  RG_TRY(mark_fresh(dst, slot));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_slot_load(rumgpu_bm* b, int slot, rumgpu_bm* src) {
  RG_SLOT_CHECK(b, slot, "bm_slot_load");
  RG_CHECK(src && src->n == b->n && src->M == b->M, RG_EINVAL, "bm_slot_load: incompatible source handle");
  RG_CHECK(src->kept_valid, RG_EINVAL, "bm_slot_load: the source handle keeps no state (bm_keep)");
  RG_CHECK(src->ctx->stream == b->ctx->stream, RG_EINVAL, "bm_slot_load: handles of different contexts");
  if (src->isz == 8 && b->isz == 8) return slot_load_t<double, double>(b, slot, src);
  if (src->isz == 8 && b->isz == 4) return slot_load_t<double, float>(b, slot, src);
  if (src->isz == 4 && b->isz == 8) return slot_load_t<float, double>(b, slot, src);
  return slot_load_t<float, float>(b, slot, src);
}

template <typename T>
static int slot_cold_t(rumgpu_bm* b, int slot) {
  hipStream_t st = b->ctx->stream;
  const i64 o = (i64)slot * b->m2;
  const unsigned gr = ew_grid(b->m2);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->s + o, (T)0, b->m2);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->y + o, (T)0, b->m2);
  hipLaunchKernelGGL((ew_fill<T>), dim3(gr), dim3(256), 0, st, (T*)b->z + o, (T)0, b->m2);
  RG_LAUNCH_CHECK();
  RG_TRY((launch_prox<T, 1>(b->n, nullptr, (T*)b->x2 + o, nullptr, b->obsmask, b->obsoff, b->pc + (i64)slot * b->M,
                            (T)b->beta, st)));
  RG_TRY(mark_fresh(b, slot));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_slot_cold(rumgpu_bm* b, int slot) {
  RG_SLOT_CHECK(b, slot, "bm_slot_cold");
  return BM_DISPATCH(b, slot_cold_t, b, slot);
}

template <typename T>
static int slot_obs_t(rumgpu_bm* b, int slot, const double* dx, const double* dz, double* ox, double* oz) {
  hipStream_t st = b->ctx->stream;
  const i64 o = (i64)slot * b->m2, M = b->M;
  double *ddx = b->dwarm, *ddz = b->dwarm + M;
  if (ox || oz) {
    hipLaunchKernelGGL(k_get_obs<T>, dim3(grid_for(M, 256)), dim3(256), 0, st, (const T*)b->x2 + o, (const T*)b->z + o,
                       b->obs, ddx, ddz, M);
    RG_LAUNCH_CHECK();
    if (ox) RG_HIP(rg_copy(st, ox, ddx, (size_t)M * 8, hipMemcpyDeviceToHost));
    if (oz) RG_HIP(rg_copy(st, oz, ddz, (size_t)M * 8, hipMemcpyDeviceToHost));
    return RG_OK;
  }
  if (dx) RG_HIP(hipMemcpyAsync(ddx, dx, (size_t)M * 8, hipMemcpyHostToDevice, st));
  if (dz) RG_HIP(hipMemcpyAsync(ddz, dz, (size_t)M * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_add_obs<T>, dim3(grid_for(M, 256)), dim3(256), 0, st, (T*)b->x2 + o, (T*)b->z + o, b->obs,
                     dx ? (const double*)ddx : nullptr, dz ? (const double*)ddz : nullptr, M);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

extern "C" int rumgpu_bm_slot_add_obs(rumgpu_bm* b, int slot, const double* dx2, const double* dz) {
  RG_SLOT_CHECK(b, slot, "bm_slot_add_obs");
  return BM_DISPATCH(b, slot_obs_t, b, slot, dx2, dz, nullptr, nullptr);
}

extern "C" int rumgpu_bm_slot_get_obs(rumgpu_bm* b, int slot, double* x2_obs, double* z_obs) {
  RG_SLOT_CHECK(b, slot, "bm_slot_get_obs");
  RG_CHECK(x2_obs || z_obs, RG_EINVAL, "bm_slot_get_obs: no output");
  return BM_DISPATCH(b, slot_obs_t, b, slot, nullptr, nullptr, x2_obs, z_obs);
}

extern "C" int rumgpu_bm_certify(rumgpu_bm* b, int want_lb, double* ub, double* lb, double* phi) {
  RG_CHECK(b && ub && phi && (lb || !want_lb), RG_EINVAL, "bm_certify: null argument");
  RG_HIP(hipSetDevice(b->ctx->device));
  const int D = b->D;
  if (b->isz == 8) {
    RG_TRY(phi_impl<double>(b, b->x2, phi, 0, D));
    RG_TRY(push_impl<double>(b, ub, D));
    if (want_lb) RG_TRY(lb_impl<double>(b, lb, D));
  } else {
    RG_TRY(phi_impl<float>(b, b->x2, phi, 0, D));
    RG_TRY(push_impl<float>(b, ub, D));
    if (want_lb) RG_TRY(lb_impl<float>(b, lb, D));
  }
  for (int d = 0; d < D; ++d) phi[d] = std::sqrt(phi[d]);
  return RG_OK;
}

// ---- the device draw queue (see the comment above k_queue_decide)
extern "C" int rumgpu_bm_queue_start(rumgpu_bm* b, rumgpu_bm* src, int64_t total, int64_t capacity, int check,
                                     int lb_every, int64_t max_iter, double abs_tol, double rel_tol, int corr) {
  RG_CHECK(b && src, RG_EINVAL, "bm_queue_start: null handle");
  RG_CHECK(!b->plain, RG_EINVAL, "bm_queue_start: the plain schedule has no draw slots");
  RG_CHECK(src->n == b->n && src->M == b->M && src->ctx->stream == b->ctx->stream, RG_EINVAL,
           "bm_queue_start: incompatible source handle");
  RG_CHECK(src->kept_valid, RG_EINVAL, "bm_queue_start: the source handle keeps no state (bm_keep)");
  RG_CHECK(total >= 1 && capacity >= 1 && check >= 1 && lb_every >= 1 && max_iter >= 1 && corr >= 0 && corr <= 3,
           RG_EINVAL, "bm_queue_start: bad argument");
  RG_CHECK(abs_tol >= 0 && rel_tol >= 0, RG_EINVAL, "bm_queue_start: tolerances must be >= 0");
  RG_HIP(hipSetDevice(b->ctx->device));
  queue_free(b);
  BmQueue* q = new BmQueue();
  b->queue = q;
  q->cap = capacity < total ? capacity : total;
  q->total = total;
  q->check = check;
  q->lb_every = lb_every;
  q->max_iter = max_iter;
  q->abs_tol = abs_tol;
  q->rel_tol = rel_tol;
  q->corr = corr;
  q->src = src;
  const i64 D = b->D, M = b->M, cap = q->cap;
  int rc = RG_OK;
  struct { void** p; size_t bytes; const char* what; } al[] = {
      {(void**)&q->rP, (size_t)(cap * M * 8), "queue ring p"},
      {(void**)&q->rBelow, (size_t)cap * 8, "queue ring thresholds"},
      {(void**)&q->rAbove, (size_t)cap * 8, "queue ring thresholds"},
      {(void**)&q->ctr, QC_N * 8, "queue counters"},
      {(void**)&q->sdraw, (size_t)D * 8, "queue slots"},
      {(void**)&q->sit, (size_t)D * 8, "queue slots"},
      {(void**)&q->sub, (size_t)D * 8, "queue slots"},
      {(void**)&q->slb, (size_t)D * 8, "queue slots"},
      {(void**)&q->sbelow, (size_t)D * 8, "queue slots"},
      {(void**)&q->sabove, (size_t)D * 8, "queue slots"},
      {(void**)&q->skill, (size_t)D * 4, "queue slots"},
      {(void**)&q->refill, (size_t)D * 16, "queue refills"},
      {(void**)&q->ubsq, (size_t)D * 8, "queue bounds"},
      {(void**)&q->pcidx, (size_t)M * 8, "queue row order"},
      {(void**)&q->qlb, (size_t)total * 8, "queue results"},
      {(void**)&q->qub, (size_t)total * 8, "queue results"},
      {(void**)&q->qit, (size_t)total * 8, "queue results"},
      {(void**)&q->qorder, (size_t)total * 8, "queue results"},
      {(void**)&q->qstop, (size_t)total * 4, "queue results"}};
  for (auto& a : al)
    if (rc == RG_OK) rc = rg_malloc(a.p, a.bytes, a.what);
  if (rc == RG_OK && (corr & 1)) rc = rg_malloc((void**)&q->rDx, (size_t)(cap * M * 8), "queue ring corrections");
  if (rc == RG_OK && (corr & 2)) rc = rg_malloc((void**)&q->rDz, (size_t)(cap * M * 8), "queue ring corrections");
  if (rc == RG_OK && hipHostMalloc((void**)&q->h_ctr, QC_N * 8) != hipSuccess) {
    rg_set_error("bm_queue_start: pinned counters");
    rc = RG_ENOMEM;
  }
  if (rc != RG_OK) { queue_free(b); return rc; }
  hipStream_t st = b->ctx->stream;
  hipLaunchKernelGGL((ew_fill<i64>), dim3(1), dim3(64), 0, st, q->ctr, (i64)0, (i64)QC_N);
  hipLaunchKernelGGL((ew_fill<i64>), dim3(grid_for(D, 256)), dim3(256), 0, st, q->sdraw, (i64)-1, D);
  hipLaunchKernelGGL((ew_fill<int>), dim3(grid_for(D, 256)), dim3(256), 0, st, q->skill, 0, D);
  hipLaunchKernelGGL((ew_fill<int>), dim3(ew_grid(total)), dim3(256), 0, st, q->qstop, 0, (i64)total);
  RG_LAUNCH_CHECK();
  for (int k = 0; k < QC_N; ++k) q->h_ctr[k] = 0;
  rc = rg_copy(st, q->pcidx, b->h_pcidx.data(), (size_t)M * 8, hipMemcpyHostToDevice) == hipSuccess ? RG_OK : RG_EHIP;
  if (rc != RG_OK) { rg_set_error("bm_queue_start: copy of the row order failed"); queue_free(b); }
  return rc;
}

// Appends k pending draws: p (k x M, input row order), the thresholds on the non-singleton part
// (NaN: none) and, when the queue was started with corrections, dx2 / dz (k x M). The ring holds
// `capacity` draws that no slot has taken yet.
extern "C" int rumgpu_bm_queue_push(rumgpu_bm* b, int64_t k, const double* P, const double* below,
                                    const double* above, const double* dx2, const double* dz) {
  RG_CHECK(b && b->queue, RG_EINVAL, "bm_queue_push: no queue (bm_queue_start)");
  BmQueue* q = b->queue;
  RG_CHECK(k >= 0 && P && below && above, RG_EINVAL, "bm_queue_push: bad argument");
  RG_CHECK(!(q->corr & 1) || dx2, RG_EINVAL, "bm_queue_push: the queue takes dx2 corrections");
  RG_CHECK(!(q->corr & 2) || dz, RG_EINVAL, "bm_queue_push: the queue takes dz corrections");
  RG_CHECK(q->tail + k <= q->total, RG_EINVAL, "bm_queue_push: %lld draws exceed the run's %lld",
           (long long)(q->tail + k), (long long)q->total);
  RG_CHECK(q->tail + k - q->h_ctr[QC_HEAD] <= q->cap, RG_EINVAL,
           "bm_queue_push: the ring holds %lld pending draws (%lld taken of %lld pushed)", (long long)q->cap,
           (long long)q->h_ctr[QC_HEAD], (long long)q->tail);
  if (k == 0) return RG_OK;
  RG_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  const i64 M = b->M;
  i64 done = 0;
  while (done < k) {
    const i64 e = (q->tail + done) % q->cap;
    const i64 m = std::min(k - done, q->cap - e);
    RG_HIP(hipMemcpyAsync(q->rP + e * M, P + done * M, (size_t)(m * M * 8), hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(q->rBelow + e, below + done, (size_t)m * 8, hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(q->rAbove + e, above + done, (size_t)m * 8, hipMemcpyHostToDevice, st));
    if (q->corr & 1) RG_HIP(hipMemcpyAsync(q->rDx + e * M, dx2 + done * M, (size_t)(m * M * 8), hipMemcpyHostToDevice, st));
    if (q->corr & 2) RG_HIP(hipMemcpyAsync(q->rDz + e * M, dz + done * M, (size_t)(m * M * 8), hipMemcpyHostToDevice, st));
    done += m;
  }
  q->tail += k;
  hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, st, q->ctr + QC_TAIL, q->tail);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

// Runs `cycles` cycles (the first call first gives every slot a draw) and returns, once they are
// done, status[0..4] = draws taken, draws pushed, draws retired, live slots, checks run.
extern "C" int rumgpu_bm_queue_run(rumgpu_bm* b, int64_t cycles, int64_t* status) {
  RG_CHECK(b && b->queue && cycles >= 0 && status, RG_EINVAL, "bm_queue_run: bad argument");
  RG_HIP(hipSetDevice(b->ctx->device));
  RG_TRY(BM_DISPATCH(b, queue_run_t, b, cycles));
  for (int k = 0; k < 5; ++k) status[k] = b->queue->h_ctr[k];
  return RG_OK;
}

// Results of the draws with sequence numbers first .. first + count - 1: lb and ub of the
// non-singleton part (running max / min at retirement), iterations, retirement rank and the stop
// (0 open, 1 threshold, 2 tolerance, 3 time, 4 max_iter). Any output may be null.
extern "C" int rumgpu_bm_queue_results(rumgpu_bm* b, int64_t first, int64_t count, double* lb, double* ub,
                                       int64_t* iters, int64_t* order, int32_t* stop) {
  RG_CHECK(b && b->queue, RG_EINVAL, "bm_queue_results: no queue");
  BmQueue* q = b->queue;
  RG_CHECK(first >= 0 && count >= 0 && first + count <= q->total, RG_EINVAL, "bm_queue_results: range outside the run");
  if (count == 0) return RG_OK;
  hipStream_t st = b->ctx->stream;
  if (lb) RG_HIP(hipMemcpyAsync(lb, q->qlb + first, (size_t)count * 8, hipMemcpyDeviceToHost, st));
  if (ub) RG_HIP(hipMemcpyAsync(ub, q->qub + first, (size_t)count * 8, hipMemcpyDeviceToHost, st));
  if (iters) RG_HIP(hipMemcpyAsync(iters, q->qit + first, (size_t)count * 8, hipMemcpyDeviceToHost, st));
  if (order) RG_HIP(hipMemcpyAsync(order, q->qorder + first, (size_t)count * 8, hipMemcpyDeviceToHost, st));
  if (stop) RG_HIP(hipMemcpyAsync(stop, q->qstop + first, (size_t)count * 4, hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

// Per slot: the draw it holds (-1 idle) and that draw's iterations so far.
extern "C" int rumgpu_bm_queue_slots(rumgpu_bm* b, int64_t* draw, int64_t* iters) {
  RG_CHECK(b && b->queue && draw && iters, RG_EINVAL, "bm_queue_slots: bad argument");
  hipStream_t st = b->ctx->stream;
  RG_HIP(hipMemcpyAsync(draw, b->queue->sdraw, (size_t)b->D * 8, hipMemcpyDeviceToHost, st));
  RG_HIP(hipMemcpyAsync(iters, b->queue->sit, (size_t)b->D * 8, hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

// The host's time stop: slot d's draw stops with stop 3 at the next decision when kill[d] != 0
// (unless a threshold or the tolerance stops it first, as in the host driver).
extern "C" int rumgpu_bm_queue_expire(rumgpu_bm* b, const int32_t* kill) {
  RG_CHECK(b && b->queue && kill, RG_EINVAL, "bm_queue_expire: bad argument");
  RG_HIP(rg_copy(b->ctx->stream, b->queue->skill, kill, (size_t)b->D * 4, hipMemcpyHostToDevice));
  return RG_OK;
}

extern "C" int rumgpu_bm_queue_end(rumgpu_bm* b) {
  RG_CHECK(b, RG_EINVAL, "bm_queue_end: null handle");
  (void)hipStreamSynchronize(b->ctx->stream);
  queue_free(b);
  return RG_OK;
}

extern "C" int rumgpu_bm_slot_fit(rumgpu_bm* b, int slot, double* p_fit) {
  RG_SLOT_CHECK(b, slot, "bm_slot_fit");
  RG_CHECK(p_fit, RG_EINVAL, "bm_slot_fit: null p_fit");
  RG_HIP(rg_copy(b->ctx->stream, p_fit, b->fit + (i64)slot * b->M, (size_t)b->M * 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

// Standalone transform of a float64 host slab (the butterfly as a building block and its test
// surface): v = T(v) with (a, b, c, d) on every rest bit.
extern "C" int rumgpu_butterfly64(rumgpu_ctx* ctx, int n, double* v, double a, double bcoef, double c, double d) {
  RG_CHECK(ctx && v && n >= 2 && n <= 30, RG_EINVAL, "butterfly64: bad argument");
  const i64 m2 = (i64)n << (n - 1);
  double* dv = nullptr;
  RG_TRY(rg_malloc((void**)&dv, (size_t)m2 * 8, "butterfly slab"));
  hipStream_t st = ctx->stream;
  int rc = RG_OK;
  if (rg_copy(ctx->stream, dv, v, (size_t)m2 * 8, hipMemcpyHostToDevice) != hipSuccess) rc = RG_EHIP;
  std::vector<WinPlan> wins = plan_windows(n - 1, 8);
  for (size_t w = 0; w < wins.size() && rc == RG_OK; ++w) {
    const WinGeom g = make_geom(n, wins[w]);
    rc = launch_tr_rl<double, double, 0, 0>(wins[w].RL, dv, dv, coef<double>(a, bcoef, c, d), g, nullptr, nullptr, st);
  }
  if (rc == RG_OK && rg_copy(ctx->stream, v, dv, (size_t)m2 * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = RG_EHIP;
  rg_free(dv);
  if (rc == RG_EHIP) rg_set_error("butterfly64: HIP error");
  return rc;
}

extern "C" int rumgpu_butterfly32(rumgpu_ctx* ctx, int n, float* v, float a, float bcoef, float c, float d) {
  RG_CHECK(ctx && v && n >= 2 && n <= 30, RG_EINVAL, "butterfly32: bad argument");
  const i64 m2 = (i64)n << (n - 1);
  float* dv = nullptr;
  RG_TRY(rg_malloc((void**)&dv, (size_t)m2 * 4, "butterfly slab"));
  hipStream_t st = ctx->stream;
  int rc = RG_OK;
  if (rg_copy(ctx->stream, dv, v, (size_t)m2 * 4, hipMemcpyHostToDevice) != hipSuccess) rc = RG_EHIP;
  std::vector<WinPlan> wins = plan_windows(n - 1, 4);
  for (size_t w = 0; w < wins.size() && rc == RG_OK; ++w) {
    const WinGeom g = make_geom(n, wins[w]);
    rc = launch_tr_rl<float, float, 0, 0>(wins[w].RL, dv, dv, coef<float>(a, bcoef, c, d), g, nullptr, nullptr, st);
  }
  if (rc == RG_OK && rg_copy(ctx->stream, v, dv, (size_t)m2 * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = RG_EHIP;
  rg_free(dv);
  if (rc == RG_EHIP) rg_set_error("butterfly32: HIP error");
  return rc;
}

// Prox of a host slab (test surface for the per-menu kernel): out = prox(xk) for rows (menu,
// alt, p) with penalty beta2, dtype by dtype_bytes. The values are float64 on the host.
extern "C" int rumgpu_prox_slab(rumgpu_bm* b, const double* xk, double* out) {
  RG_CHECK(b && xk && out, RG_EINVAL, "prox_slab: null argument");
  RG_TRY(BM_DISPATCH(b, finalize, b));       // tb is used as scratch below
  hipStream_t st = b->ctx->stream;
  const i64 m2 = b->m2;
  int rc = RG_OK;
  if (b->isz == 8) {
    RG_HIP(rg_copy(b->ctx->stream, b->xk, xk, (size_t)m2 * 8, hipMemcpyHostToDevice));
    rc = launch_prox<double, 2>(b->n, (const double*)b->xk, (double*)b->tb, nullptr, b->obsmask, b->obsoff, b->pc, b->beta, st);
    if (rc == RG_OK) RG_HIP(rg_copy(b->ctx->stream, out, b->tb, (size_t)m2 * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(m2);
    for (i64 i = 0; i < m2; ++i) tmp[i] = (float)xk[i];
    RG_HIP(rg_copy(b->ctx->stream, b->xk, tmp.data(), (size_t)m2 * 4, hipMemcpyHostToDevice));
    rc = launch_prox<float, 2>(b->n, (const float*)b->xk, (float*)b->tb, nullptr, b->obsmask, b->obsoff, b->pc, (float)b->beta, st);
    if (rc == RG_OK) {
      RG_HIP(rg_copy(b->ctx->stream, tmp.data(), b->tb, (size_t)m2 * 4, hipMemcpyDeviceToHost));
      for (i64 i = 0; i < m2; ++i) out[i] = tmp[i];
    }
  }
  return rc;
}

// Standalone float64 lower bound at a point y on the rows (its own scratch): the same kernels
// as rumgpu_bm_lb64.
extern "C" int rumgpu_lb64(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, int64_t M,
                           const double* y, const double* p, double* lb) {
  RG_CHECK(ctx && menu && alt && y && p && lb && M >= 1, RG_EINVAL, "lb64: bad argument");
  RG_CHECK(n >= 2 && n <= 30, RG_EUNSUP, "lb64: n = %d outside 2..30", n);
  RG_HIP(hipSetDevice(ctx->device));
  const i64 Nloc = 1ll << (n - 1), m2 = (i64)n * Nloc, NN = 1ll << n;
  std::vector<i64> hobs(M);
  for (i64 j = 0; j < M; ++j) {
    const int a = alt[j];
    RG_CHECK(a >= 0 && a < n && ((menu[j] >> a) & 1ull), RG_EINVAL, "lb64: row %lld alternative not in menu", (long long)j);
    hobs[j] = (i64)a * Nloc + (i64)compress_bit(menu[j] & ~(1ull << a), a);
  }
  double *g = nullptr, *dp = nullptr, *dy = nullptr, *dpv = nullptr, *part = nullptr, *res = nullptr;
  i64* obs = nullptr;
  int* order = nullptr;
  int rc = RG_OK;
  if ((rc = rg_malloc((void**)&g, (size_t)m2 * 8, "lb64 slab")) == RG_OK &&
      (rc = rg_malloc((void**)&dp, (size_t)NN * 8, "lb64 dp")) == RG_OK &&
      (rc = rg_malloc((void**)&dy, (size_t)M * 8, "lb64 y")) == RG_OK &&
      (rc = rg_malloc((void**)&dpv, (size_t)M * 8, "lb64 p")) == RG_OK &&
      (rc = rg_malloc((void**)&obs, (size_t)M * 8, "lb64 obs")) == RG_OK &&
      (rc = rg_malloc((void**)&order, (size_t)NN * 4, "lb64 layers")) == RG_OK &&
      (rc = rg_malloc((void**)&part, (size_t)256 * 4 * 8, "lb64 partials")) == RG_OK &&
      (rc = rg_malloc((void**)&res, 16 * 8, "lb64 scalars")) == RG_OK) {
    hipStream_t st = ctx->stream;
    std::vector<int> hord(NN);
    std::vector<i64> bounds(n + 2, 0), cnt(n + 2, 0);
    for (i64 v = 0; v < NN; ++v) cnt[__builtin_popcountll(v)]++;
    for (int c = 0; c <= n; ++c) bounds[c + 1] = bounds[c] + cnt[c];
    std::vector<i64> fill(bounds.begin(), bounds.end());
    for (i64 v = 0; v < NN; ++v) hord[fill[__builtin_popcountll(v)]++] = (int)v;
    // y is a float64 "slab" of M values read through obs: gather index = identity
    std::vector<i64> ident(M);
    for (i64 j = 0; j < M; ++j) ident[j] = j;
    if (rg_copy(ctx->stream, order, hord.data(), (size_t)NN * 4, hipMemcpyHostToDevice) != hipSuccess ||
        rg_copy(ctx->stream, obs, hobs.data(), (size_t)M * 8, hipMemcpyHostToDevice) != hipSuccess ||
        rg_copy(ctx->stream, dy, y, (size_t)M * 8, hipMemcpyHostToDevice) != hipSuccess ||
        rg_copy(ctx->stream, dpv, p, (size_t)M * 8, hipMemcpyHostToDevice) != hipSuccess) {
      rc = RG_EHIP;
      rg_set_error("lb64: upload failed");
    }
    if (rc == RG_OK) {
      hipLaunchKernelGGL((ew_fill<double>), dim3(ew_grid(m2)), dim3(256), 0, st, g, 0.0, m2);
      const int nblk = (int)std::min<i64>(256, std::max<i64>(1, (M + 255) / 256));
      // y per row is scattered into g directly by k_lb_scatter below
      hipLaunchKernelGGL(k_lb_rows, dim3(nblk), dim3(256), 0, st, dy, obs, dpv, (i64)M, g, part);
      hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, part, nblk, 2, res);
      std::vector<WinPlan> wins = plan_windows(n - 1, 8);
      for (size_t w = 0; w < wins.size() && rc == RG_OK; ++w)
        rc = launch_tr_rl<double, double, 0, 0>(wins[w].RL, g, g, coef<double>(1.0, 0.0, 1.0, 1.0),
                                                make_geom(n, wins[w]), nullptr, nullptr, st);
      if (rc == RG_OK) {
        hipLaunchKernelGGL((ew_fill<double>), dim3(1), dim3(64), 0, st, dp, 0.0, (i64)1);
        for (int c = 1; c <= n; ++c) {
          const i64 lo = bounds[c], ns = bounds[c + 1] - lo;
          if (ns <= 0) continue;
          launch_lpath(n, dim3(grid_for(ns, 256)), st, g, dp, order + lo, ns, Nloc);
        }
        double h[3];
        if (hipMemcpyAsync(h, res, 2 * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(h + 2, dp + (NN - 1), 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
          rc = RG_EHIP;
          rg_set_error("lb64: download failed");
        } else {
          const double l2 = h[0] - 2.0 * (h[2] - h[1]);
          *lb = std::sqrt(l2 > 0.0 ? l2 : 0.0);
        }
      }
    }
  }
  rg_free(g); rg_free(dp); rg_free(dy); rg_free(dpv); rg_free(obs); rg_free(order); rg_free(part); rg_free(res);
  return rc;
}

// ---------------------------------------------------------------------------------------------
// The same float64 lower bound without any device allocation (rumgpu_lb64_noalloc): rows, y and p
// are device arrays, and every buffer rumgpu_lb64 allocates per call lives in one fixed device
// array of this library. Layout (byte offsets, sized for n = LB64_NOALLOC_MAX_N, each a multiple of
// 256): the float64 slab g, dp[2^n], the node order int[2^n], the block partials (256 blocks x 2
// quantities x (sum, compensation)), the two reduced sums, and a bad-row flag.
#define LB64_NOALLOC_MAX_N 18
static constexpr size_t kNaSlab = (size_t)LB64_NOALLOC_MAX_N << (LB64_NOALLOC_MAX_N - 1);   // doubles
static constexpr size_t kNaNodes = (size_t)1 << LB64_NOALLOC_MAX_N;
static constexpr size_t kNaOffDp = kNaSlab * 8;
static constexpr size_t kNaOffOrder = kNaOffDp + kNaNodes * 8;
static constexpr size_t kNaOffPart = kNaOffOrder + kNaNodes * 4;
static constexpr size_t kNaOffRes = kNaOffPart + 256 * 4 * 8;
static constexpr size_t kNaOffFlag = kNaOffRes + 256;
static constexpr size_t kNaBytes = kNaOffFlag + 256;                                      // 22,028,800

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Device address of `scratch` per device, taken once with that device current (hipGetSymbolAddress);
// host code uses only these addresses, never the array's name.
static std::mutex g_na_mu;
static void* g_na_ptr[64] = {};
static hipError_t g_na_err[64] = {};

int rg_noalloc_setup(int device) {
  if (device < 0 || device >= 64) return RG_EUNSUP;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[device]) return RG_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_na_err[device] = e;
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    rg_set_error("lb64_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return RG_EHIP;
  }
  g_na_ptr[device] = p;
  return RG_OK;
}

// r = p - y at row j (the slab index from its menu and alternative), g[index] = r, f = sum r^2,
// ry = sum r y: the arithmetic and the order of k_lb_rows. A row whose alternative is outside 0..n-1
// or not in its menu, or whose menu has bits at n or above, is skipped and raises the flag.
__global__ __launch_bounds__(256) void k_lb_rows_dev(const u64* __restrict__ menu, const int32_t* __restrict__ alt,
                                                     int n, i64 Nloc, const double* __restrict__ yv,
                                                     const double* __restrict__ p, i64 M, double* g, double* part,
                                                     int* flag) {
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  for (i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (i64)gridDim.x * blockDim.x) {
    const u64 m = menu[j];
    const int a = alt[j];
    if (a < 0 || a >= n || !((m >> a) & 1ull) || (m >> n) != 0ull) {
      flag[0] = 1;
      continue;
    }
    const i64 o = (i64)a * Nloc + (i64)compress_bit(m & ~(1ull << a), a);
    const double y = yv[j];
    const double r = p[j] - y;
    g[o] = r;
    two_sum(s[0], c[0], r * r);
    two_sum(s[1], c[1], r * y);
  }
  block_reduce_store<2>(s, c, part);
}

// Node order by popcount, ascending masks within a layer: node v sits at (number of masks of lower
// popcount) + (colex rank of v among the masks of its popcount), the order rumgpu_lb64 builds on the host.
__device__ inline i64 na_binom(int a, int b) {
  if (b < 0 || b > a) return 0;
  i64 r = 1;
  for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
  return r;
}
__global__ void k_layer_order(int* order, int n) {
  const i64 v = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (1ll << n)) return;
  const int k = __popcll((u64)v);
  i64 pos = 0;
  for (int j = 0; j < k; ++j) pos += na_binom(n, j);
  int i = 0;
  for (int b = 0; b < n; ++b)
    if ((v >> b) & 1) pos += na_binom(b, ++i);
  order[pos] = (int)v;
}

// lb = sqrt(max(0, f - 2 (dp[full] - ry))), NaN when a row was bad (the flag is cleared at the start of
// every call).
__global__ void k_lb_finish(const double* res, const double* dp, i64 full, int* flag, double* lb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double l2 = res[0] - 2.0 * (dp[full] - res[1]);
  lb[0] = flag[0] ? __longlong_as_double(0x7ff8000000000000ll) : sqrt(l2 > 0.0 ? l2 : 0.0);
}

extern "C" int rumgpu_lb64_noalloc(rumgpu_ctx* ctx, int n, const uint64_t* menu_dev, const int32_t* alt_dev,
                                   int64_t M, const double* y_dev, const double* p_dev, double* lb_dev) {
  RG_CHECK(ctx && menu_dev && alt_dev && y_dev && p_dev && lb_dev && M >= 1, RG_EINVAL,
           "lb64_noalloc: bad argument");
  RG_CHECK(n >= 2 && n <= LB64_NOALLOC_MAX_N, RG_EUNSUP,
           "lb64_noalloc: n = %d outside 2..%d (its fixed scratch holds %zu bytes per device)", n,
           LB64_NOALLOC_MAX_N, kNaBytes);
  RG_HIP(hipSetDevice(ctx->device));
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (ctx->device >= 0 && ctx->device < 64) base = g_na_ptr[ctx->device];
  }
  RG_CHECK(base, RG_EHIP, "lb64_noalloc: the scratch of device %d has no address (set-up failed: %s)", ctx->device,
           hipGetErrorString(ctx->device >= 0 && ctx->device < 64 ? g_na_err[ctx->device] : hipErrorInvalidDevice));
  unsigned char* sb = (unsigned char*)base;
  double* g = (double*)sb;
  double* dp = (double*)(sb + kNaOffDp);
  int* order = (int*)(sb + kNaOffOrder);
  double* part = (double*)(sb + kNaOffPart);
  double* res = (double*)(sb + kNaOffRes);
  int* flag = (int*)(sb + kNaOffFlag);
  const i64 Nloc = 1ll << (n - 1), m2 = (i64)n * Nloc, NN = 1ll << n;
  hipStream_t st = ctx->stream;
  // the scratch keeps the last call's contents: clear the slab, dp[0] and the flag
  hipLaunchKernelGGL((ew_fill<double>), dim3(ew_grid(m2)), dim3(256), 0, st, g, 0.0, m2);
  hipLaunchKernelGGL((ew_fill<int>), dim3(1), dim3(64), 0, st, flag, 0, (i64)1);
  hipLaunchKernelGGL(k_layer_order, dim3(grid_for(NN, 256)), dim3(256), 0, st, order, n);
  const int nblk = (int)std::min<i64>(256, std::max<i64>(1, (M + 255) / 256));
  hipLaunchKernelGGL(k_lb_rows_dev, dim3(nblk), dim3(256), 0, st, (const u64*)menu_dev, alt_dev, n, Nloc, y_dev,
                     p_dev, (i64)M, g, part, flag);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, part, nblk, 2, res);
  RG_LAUNCH_CHECK();
  WinPlan wins[kMaxWindows];
  const int nw = plan_windows_into(n - 1, 8, wins);
  for (int w = 0; w < nw; ++w)
    RG_TRY((launch_tr_rl<double, double, 0, 0>(wins[w].RL, g, g, coef<double>(1.0, 0.0, 1.0, 1.0),
                                                make_geom(n, wins[w]), nullptr, nullptr, st)));
  hipLaunchKernelGGL((ew_fill<double>), dim3(1), dim3(64), 0, st, dp, 0.0, (i64)1);
  i64 lo = 1;                                   // layer c starts after the C(n, j) masks of j < c
  i64 ns = n;
  for (int c = 1; c <= n; ++c) {
    launch_lpath(n, dim3(grid_for(ns, 256)), st, g, dp, order + lo, ns, Nloc);
    lo += ns;
    ns = ns * (n - c) / (c + 1);
  }
  hipLaunchKernelGGL(k_lb_finish, dim3(1), dim3(64), 0, st, res, dp, NN - 1, flag, lb_dev);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  return RG_OK;
}

// Per-pass device time of `iters` fused iterations, measured with HIP events on the context stream
// between consecutive launches (the graph is bypassed). ms_by_kind[7] and launches_by_kind[7]:
// 0 SY, 1 RHS, 2 D, 3 XK, 4 prox, 5 middle-window transform, 6 one-window FULL pass.
extern "C" int rumgpu_bm_profile(rumgpu_bm* b, int64_t iters, double* ms_by_kind, int64_t* launches_by_kind) {
  RG_CHECK(b && ms_by_kind && launches_by_kind && iters > 0, RG_EINVAL, "bm_profile: bad argument");
  RG_CHECK(!b->plain, RG_EINVAL, "bm_profile: the plain schedule has no fused passes");
  RG_HIP(hipSetDevice(b->ctx->device));
  for (int q = 0; q < 7; ++q) { ms_by_kind[q] = 0.0; launches_by_kind[q] = 0; }
  const int per_iter = 2 * (2 + 2 * (int)b->wins.size() + 4);     // an event pair per launch
  const i64 batch = 16;
  if (b->prof_ev.size() < (size_t)(per_iter * batch)) {
    for (auto e : b->prof_ev) (void)hipEventDestroy(e);
    b->prof_ev.assign((size_t)per_iter * batch, nullptr);
    for (auto& e : b->prof_ev) RG_HIP(hipEventCreate(&e));
    b->prof_kind.assign(b->prof_ev.size(), -1);
  }
  i64 done = 0;
  int rc = RG_OK;
  while (done < iters && rc == RG_OK) {
    const i64 now = std::min<i64>(batch, iters - done);
    b->prof_on = 1;
    b->prof_n = 0;
    for (i64 i = 0; i < now && rc == RG_OK; ++i) rc = BM_DISPATCH(b, fused_iter, b);
    b->prof_on = 0;
    if (rc != RG_OK) break;
    RG_HIP(hipStreamSynchronize(b->ctx->stream));
    for (size_t e = 0; e + 1 < b->prof_n; e += 2) {
      const int kind = b->prof_kind[e];
      float ms = 0.0f;
      RG_HIP(hipEventElapsedTime(&ms, b->prof_ev[e], b->prof_ev[e + 1]));
      ms_by_kind[kind] += ms;
      launches_by_kind[kind] += 1;
    }
    done += now;
  }
  b->prof_on = 0;
  g_pass_ev0 = g_pass_ev1 = nullptr;       // an armed pair that no launch consumed (an error path)
  return rc;
}
