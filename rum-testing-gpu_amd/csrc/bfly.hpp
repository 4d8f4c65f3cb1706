// Window passes of the generalized 2x2 butterfly over the subset lattice.
//
// A slab holds n item blocks of Nloc = 2^nb values (nb = n - 1 rest bits). A Kronecker
// transform applies the same 2x2 matrix [[a, b], [c, d]] on every rest bit. Its bit stages
// commute, so the nb bits are cut into windows and each window is one pass over the slab:
//
//   window [s, s+B)   the bits the pass transforms
//   columns [0, C)    for a window above bit 0 (s >= C): 2^C contiguous values per row, so a
//                     wavefront reads 2^C-element runs of memory (128 B) instead of one value
//   outer bits        the rest: one tile per (item, outer bits), tile = 2^(B+C) values
//
// Inside a tile each thread holds R = 2^RL values that differ only in RL window bits, so RL
// stages run in registers; the next RL bits are reached by one exchange through LDS. A window
// of B bits costs ceil(B/RL) register rounds and ceil(B/RL) - 1 exchanges. Several transforms
// and the pointwise ADMM updates between them run in one pass while the tile is resident
// (the fused passes are in bm_admm.hip).
#pragma once
#include "common.hpp"

struct WinGeom {
  int nb;        // rest bits
  int s;         // first window bit
  int B;         // window bits
  int C;         // column bits (low contiguous bits, 0 for the window at s = 0)
  int TT;        // threads per tile = 2^(B + C - RL)
  int TPB;       // tiles per block
  i64 Nloc;      // 2^nb
  i64 ntiles;    // items * 2^(nb - B - C)
  int nper;      // items per slot (n): a batch of D draws is n * D item blocks
  const int* on; // per slot: 0 = idle, its tiles move no data (null: every slot runs)
};

// Host-side window plan for a slab of itemsize bytes: list of (s, B, C, RL).
struct WinPlan {
  int s, B, C, RL;
};

// An explicit plan: the window at bit 0 takes lo_bits (C = 0, lo_rl register bits), the bits above
// it go to windows of at most 8 bits (hi_rl register bits) with 16 (float64) or 32 (float32)
// columns. The autotuner picks among these per n.
inline std::vector<WinPlan> plan_windows_explicit(int nb, int itemsize, int lo_bits, int lo_rl, int hi_rl) {
  std::vector<WinPlan> out;
  if (nb <= 0) return out;
  const int lo = lo_bits < nb ? lo_bits : nb;
  out.push_back({0, lo, 0, lo_rl < lo ? lo_rl : lo});
  int s = lo, rem = nb - lo;
  const int k = (rem + 7) / 8;
  for (int i = 0; i < k; ++i) {
    const int left = k - i;
    const int B = (rem + left - 1) / left;
    int Cw = itemsize == 8 ? 4 : 5;
    if (Cw > s) Cw = s;
    out.push_back({s, B, Cw, hi_rl < B ? hi_rl : B});
    s += B;
    rem -= B;
  }
  return out;
}

// At most 4 windows for nb <= 36 (12 + 3 x 8 bits).
constexpr int kMaxWindows = 4;

// The built-in plan written into out[kMaxWindows]; returns the number of windows (no allocation).
inline int plan_windows_into(int nb, int itemsize, WinPlan* out) {
  if (nb <= 0) return 0;
  // bit 0 window up to 12 bits (a 4096-value contiguous tile, 33 KiB of float64 LDS); windows
  // above it up to 8 bits with 16-32 columns. As few windows as that allows: each window is one
  // fused pass per transform.
  int k = nb <= 12 ? 1 : 1 + (nb - 12 + 7) / 8;
  if (k > kMaxWindows) k = kMaxWindows;
  int lo = (nb + k - 1) / k;
  if (lo < nb - 8 * (k - 1)) lo = nb - 8 * (k - 1);
  if (lo > 12) lo = 12;
  int rem = nb - lo;
  out[0] = {0, lo, 0, lo < 3 ? lo : 3};
  int s = lo;
  for (int i = 1; i < k; ++i) {
    int left = k - i;
    int B = (rem + left - 1) / left;
    int Cw = itemsize == 8 ? 4 : 5;
    if (Cw > s) Cw = s;
    int RL = B < 4 ? B : 4;
    out[i] = {s, B, Cw, RL};
    s += B;
    rem -= B;
  }
  return k;
}

inline std::vector<WinPlan> plan_windows(int nb, int itemsize) {
  WinPlan w[kMaxWindows];
  const int k = plan_windows_into(nb, itemsize, w);
  return std::vector<WinPlan>(w, w + k);
}

inline WinGeom make_geom(int n, const WinPlan& w, int slots = 1) {
  WinGeom g;
  g.nper = n;
  g.nb = n - 1;
  g.s = w.s;
  g.B = w.B;
  g.C = w.C;
  g.TT = 1 << (w.B + w.C - w.RL);
  g.TPB = g.TT >= 256 ? 1 : 256 / g.TT;
  g.Nloc = 1ll << g.nb;
  g.ntiles = ((i64)n * slots) << (g.nb - w.B - w.C);
  g.on = nullptr;
  if (g.ntiles < g.TPB) g.TPB = (int)g.ntiles;
  if (g.TPB < 1) g.TPB = 1;
  return g;
}

__host__ __device__ inline size_t tile_lds_elems(const WinGeom& g) {
  size_t e = (size_t)1 << (g.B + g.C);
  return e + (e >> 4) + 1;
}

template <typename T>
struct Coef {
  T a, b, c, d;
};

// One tile of a window pass held in registers. The layout index `cur` names the register
// round whose RL window bits the thread currently holds. BT > 0 fixes the window bits and CT >= 0
// the column bits at compile time (they must equal the geometry's): the rounds then unroll and
// the register and LDS index arithmetic folds to constants; BT = 0, CT = -1 read both from the
// geometry.
template <typename T, int RL, int BT = 0, int CT = -1>
struct Tile {
  static constexpr int R = 1 << RL;
  T v[R];
  int s, B, C, nr, cur;
  int ot, col;        // thread's window bits outside the round, and its column
  i64 base;           // slab index of the tile's (w = 0, col = 0) element
  i64 restbase;       // base minus the item offset (the rest index of that element)
  int item;           // the tile's item (alternative)
  T* lds;
  bool active;

  // start_last: begin in the last round's layout, where a window at bit 0 reads contiguous memory
  // across the lanes (the first round's layout gives each thread R consecutive values instead);
  // the first transform then runs its rounds in descending order.
  //
  // item_inner: the tiles are taken in (rest tile, item) order with the item fastest, and each of
  // the 8 XCDs (which receive workgroups round robin) gets one contiguous run of that order, so
  // the items of one rest tile run back to back on one XCD and share its L2 for data indexed by
  // the rest alone (the D pass's inv_gl).
  __device__ __forceinline__ void init(const WinGeom& g, T* lds_all, bool start_last = false,
                                       bool item_inner = false) {
    s = g.s;
    B = BT > 0 ? BT : g.B;
    C = CT >= 0 ? CT : g.C;
    nr = (B + RL - 1) / RL;
    cur = (start_last && C == 0) ? nr - 1 : 0;
    const int tl = threadIdx.x / g.TT;
    const int t = threadIdx.x - tl * g.TT;
    const int outer_bits = g.nb - B - C;
    i64 x, o;
    if (item_inner) {
      const u32 per = gridDim.x >> 3;
      u32 blk = blockIdx.x;
      if (blk < per * 8u) blk = (blk & 7u) * per + (blk >> 3);
      const i64 lin = (i64)blk * g.TPB + tl;
      const i64 nitems = g.ntiles >> outer_bits;
      active = lin < g.ntiles;
      o = lin / nitems;
      x = lin - o * nitems;
    } else {
      const i64 tile = (i64)blockIdx.x * g.TPB + tl;
      active = tile < g.ntiles;
      x = tile >> outer_bits;
      o = tile & ((1ll << outer_bits) - 1);
    }
    if (active && g.on && !g.on[x / g.nper]) active = false;   // an idle draw slot
    const int lowo = s - C;
    const i64 lowpart = o & ((1ll << lowo) - 1);
    const i64 highpart = o >> lowo;
    restbase = (lowpart << C) | (highpart << (s + B));
    base = x * g.Nloc + restbase;
    item = (int)x;
    col = t & ((1 << C) - 1);
    ot = t >> C;
    lds = lds_all + (size_t)tl * tile_lds_elems(g);
  }
  __device__ __forceinline__ int lo_of(int r) const {
    int l = r * RL;
    return l < B - RL ? l : B - RL;
  }
  __device__ __forceinline__ int w_of(int r, int j) const {
    const int lo = lo_of(r);
    return ((ot >> lo) << (lo + RL)) | (j << lo) | (ot & ((1 << lo) - 1));
  }
  // slab index / rest index of register j in the current layout
  __device__ __forceinline__ i64 gidx(int j) const { return base + ((i64)w_of(cur, j) << s) + col; }
  __device__ __forceinline__ i64 rest(int j) const { return restbase + ((i64)w_of(cur, j) << s) + col; }
  // LDS slot of register j in round r's layout: the tile index with bits 4-7 folded into bits 0-3
  // (a bijection on the tile), which spreads the strided rounds of a window at bit 0 over the banks
  __device__ __forceinline__ int lidx(int r, int j) const {
    const int L = (w_of(r, j) << C) + col;
    return L ^ ((L >> 4) & 15);
  }
  __device__ __forceinline__ void exchange(int r) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) lds[lidx(cur, j)] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = lds[lidx(r, j)];
    cur = r;
  }
  // FMA: fuse each output's second product and the sum (one rounding fewer); the plain passes
  // keep separately rounded products, the oracle's arithmetic
  template <bool FMA = false>
  __device__ __forceinline__ void stages(int r, const Coef<T>& k) {
    const int lo = lo_of(r);
    const int q0 = r * RL - lo;
    const int q1e = (r + 1) * RL < B ? (r + 1) * RL : B;
    const int q1 = q1e - lo;
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      if (q >= q0 && q < q1) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (!(j & (1 << q))) {
            const T X = v[j], Y = v[j | (1 << q)];
            if constexpr (FMA) {
              v[j] = __builtin_fma(k.a, X, k.b * Y);
              v[j | (1 << q)] = __builtin_fma(k.c, X, k.d * Y);
            } else {
              v[j] = k.a * X + k.b * Y;
              v[j | (1 << q)] = k.c * X + k.d * Y;
            }
          }
        }
      }
    }
  }
  // For a window at bit 0, move to the last round's layout, where a wavefront's lanes hold
  // consecutive slab positions for every register (one contiguous 512 B run per access for
  // float64); windows with columns are contiguous along the lanes in every layout already.
  __device__ __forceinline__ void coalesce() {
    if (C == 0 && cur != nr - 1) exchange(nr - 1);
  }
  // The whole window transform; starts from the current layout and runs the rounds in the
  // direction that needs no exchange before the first one.
  template <bool FMA = false>
  __device__ __forceinline__ void transform(const Coef<T>& k) {
    const bool fwd = (cur == 0);
#pragma unroll
    for (int i = 0; i < nr; ++i) {
      const int r = fwd ? i : nr - 1 - i;
      if (r != cur) exchange(r);
      stages<FMA>(r, k);
    }
  }
  // NT: non-temporal accesses, for slabs too large for the caches to give anything back
  template <bool NT = false, typename S>
  __device__ __forceinline__ void load(const S* __restrict__ p) {
    if (active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if constexpr (NT) v[j] = (T)__builtin_nontemporal_load(p + gidx(j));
        else v[j] = (T)p[gidx(j)];
      }
    }
  }
  template <bool NT = false, typename S>
  __device__ __forceinline__ void store(S* __restrict__ p) const {
    if (active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if constexpr (NT) __builtin_nontemporal_store((S)v[j], p + gidx(j));
        else p[gidx(j)] = (S)v[j];
      }
    }
  }
};
