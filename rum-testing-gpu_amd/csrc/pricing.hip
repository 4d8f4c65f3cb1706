// Flow-CG pricing on MI355X: the longest path through the subset lattice,
//   V[D] = max over x in D of V[D minus x] + w(x, D),   w(x, D) = sum of the scores of x's rows
//   whose menu is a subset of D,   pred[D] = the maximizing x (lowest x on ties),
// and the ranking read back along pred from the full set (flow_cg/bf_pricing.py:8-47,
// flow_cg/solver.py:168-200, flow_cg/pricing_stream.py:61-412).
//
// Three layouts:
//   mask    one float32 V and one pred byte per subset; a thread per subset of a cardinality
//           layer scans its members' rows in row order, exactly the reference's float32
//           additions, so V and the ranking are bit-identical to bf_update for any scores.
//           The row loop is uniform across a wavefront (scalar loads of the rows).
//   slab    the scores are scattered on the Block-Marschak slab and summed over subsets by the
//           float32 butterfly (1,0,1,1), which gives w(x, D) for every arc at once
//           (common/transforms.py:108-113 uses the same identity for the lower bound); the
//           layered max then reads one arc value per member. HBM traffic is O(n 2^n) instead
//           of the mask scan's O(mp 2^(n-1)) tests; the float32 sums are in butterfly order,
//           so V equals the mask layout's wherever every partial sum is exact (quantized duals).
//   stream  two live layers addressed by colex rank, with float32, uint16 or uint24 block
//           quantization (BLOCK = 1024 ranks per (min, step) pair) and predecessors stored
//           for the top layers only; the rest of the path comes from an exact float32 mask DP on
//           the sub-lattice of the node reached (pricing_stream.py:348-405).
#include "rumgpu_internal.hpp"
#include "transform.hpp"

#include <cmath>
#include <algorithm>
#include <chrono>

// ---------------------------------------------------------------------------------------------
// mask layout: one cardinality layer
__global__ __launch_bounds__(256) void cg_mask_layer(float* __restrict__ V, uint8_t* __restrict__ pred,
                                                     const u32* __restrict__ masks, const float* __restrict__ scores,
                                                     const u64* __restrict__ offsets, const u32* __restrict__ nodes,
                                                     i64 ns, int n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < ns;
  const u32 D = act ? nodes[i] : 0u;
  float best = 0.0f;
  bool seen = false;
  int bx = 0;
  for (int x = 0; x < n; ++x) {
    const bool in = act && ((D >> x) & 1u);
    float w = 0.0f;
    const u64 lo = offsets[x], hi = offsets[x + 1];
    for (u64 j = lo; j < hi; ++j) {
      const u32 Dp = masks[j];
      if ((Dp & D) == Dp) w += scores[j];
    }
    if (in) {
      const float cand = V[D ^ (1u << x)] + w;
      if (!seen || cand > best) {
        best = cand;
        seen = true;
        bx = x;
      }
    }
  }
  if (act) {
    V[D] = best;
    pred[D] = (uint8_t)bx;
  }
}

// mask layout, DB draws per sweep: the subset tests of a node are shared by the draws, each draw
// keeps its own sums, values and choices. V and pred are node-major ([node][DB]), scores row-major
// ([row][DB]), so the DB values of one node or row are one contiguous line. Per draw the arithmetic
// and its order are cg_mask_layer's (row order within x, x ascending, strict > after the first
// member), so every draw's V and pred equal a single-draw sweep on its scores bit for bit.
template <int DB>
__global__ __launch_bounds__(256) void cg_mask_layer_batch(float* __restrict__ V, uint8_t* __restrict__ pred,
                                                           const u32* __restrict__ masks,
                                                           const float* __restrict__ scores,
                                                           const u64* __restrict__ offsets,
                                                           const u32* __restrict__ nodes, i64 ns, int n,
                                                           u32 zmask) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const u32 D = nodes[i];
  // no member has rows: every arc weighs 0, so by induction V = 0 here and the first member
  // (ascending, strict >) is the choice. Such a node is not stored: its successors read the
  // empty set's row (0, set before the sweep, cache-resident) in its place, and the backtrace
  // takes its lowest member.
  if ((D & zmask) == 0u) return;
  float best[DB];
  uint8_t bx[DB];
#pragma unroll
  for (int d = 0; d < DB; ++d) {
    best[d] = 0.0f;
    bx[d] = 0;
  }
  bool seen = false;
  for (int x = 0; x < n; ++x) {
    if (!((D >> x) & 1u)) continue;
    float w[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) w[d] = 0.0f;
    // subset tests 32 rows at a time (uniform loads), then each lane adds only its own matches
    // (a lane-private loop, ascending rows: the same order as the single-draw sweep)
    const u64 lo = offsets[x], hi = offsets[x + 1];
    for (u64 base = lo; base < hi; base += 32) {
      const int cnt = (hi - base) < 32 ? (int)(hi - base) : 32;
      u32 mm = 0;
      for (int t = 0; t < cnt; ++t) {
        const u32 Dp = masks[base + t];
        mm |= ((Dp & D) == Dp ? 1u : 0u) << t;
      }
      while (mm) {
        const int t = __ffs((int)mm) - 1;
        mm &= mm - 1u;
        const float* sj = scores + (base + (u64)t) * DB;
#pragma unroll
        for (int d = 0; d < DB; ++d) w[d] += sj[d];
      }
    }
    const u32 P = D ^ (1u << x);
    const float* vp = V + (size_t)((P & zmask) ? P : 0u) * DB;     // a row-free P reads V[empty] = 0
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      const float cand = vp[d] + w[d];
      if (!seen || cand > best[d]) {
        best[d] = cand;
        bx[d] = (uint8_t)x;
      }
    }
    seen = true;
  }
  float* vo = V + (size_t)D * DB;
  uint8_t* po = pred + (size_t)D * DB;
#pragma unroll
  for (int d = 0; d < DB; ++d) {
    vo[d] = best[d];
    po[d] = bx[d];
  }
}

// float64 value of the same longest path (no choices kept): the Frank-Wolfe lower bound's
// max over rankings of sum_j r_j [the ranking picks row j's alternative in its menu]
// (fw_lb_rows_gpu, common/transforms.py:101-114, which reaches it through a float64 subset-sum
// slab and a layered longest path; here the arc weights are the rows' subset tests, in float64)
__global__ __launch_bounds__(256) void cg_mask_layer64(double* __restrict__ V, const u32* __restrict__ masks,
                                                       const double* __restrict__ r, const u64* __restrict__ offsets,
                                                       const u32* __restrict__ nodes, i64 ns, int n, u32 zmask) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const u32 D = nodes[i];
  if ((D & zmask) == 0u) return;       // no member has rows: V = 0, not stored (as in cg_mask_layer_batch)
  double best = 0.0;
  bool seen = false;
  for (int x = 0; x < n; ++x) {
    if (!((D >> x) & 1u)) continue;
    double w = 0.0;
    const u64 lo = offsets[x], hi = offsets[x + 1];
    for (u64 j = lo; j < hi; ++j) {
      const u32 Dp = masks[j];
      if ((Dp & D) == Dp) w += r[j];
    }
    const u32 P = D ^ (1u << x);
    const double cand = V[(P & zmask) ? P : 0u] + w;
    if (!seen || cand > best) best = cand;
    seen = true;
  }
  V[D] = best;
}

// the float64 longest path for DB draws at once (the lower bounds of draws that finish in the same
// round): node-major V ([node][DB], 8 B per draw), row-major residuals ([row][DB]); per draw the
// sums and the max run in cg_mask_layer64's order, so each draw's value equals its own sweep's
template <int DB>
__global__ __launch_bounds__(256) void cg_mask_layer64_batch(double* __restrict__ V, const u32* __restrict__ masks,
                                                             const double* __restrict__ r,
                                                             const u64* __restrict__ offsets,
                                                             const u32* __restrict__ nodes, i64 ns, int n,
                                                             u32 zmask) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const u32 D = nodes[i];
  if ((D & zmask) == 0u) return;       // no member has rows: V = 0, not stored (as in cg_mask_layer_batch)
  double best[DB];
#pragma unroll
  for (int d = 0; d < DB; ++d) best[d] = 0.0;
  bool seen = false;
  for (int x = 0; x < n; ++x) {
    if (!((D >> x) & 1u)) continue;
    double w[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) w[d] = 0.0;
    const u64 lo = offsets[x], hi = offsets[x + 1];
    for (u64 base = lo; base < hi; base += 32) {
      const int cnt = (hi - base) < 32 ? (int)(hi - base) : 32;
      u32 mm = 0;
      for (int t = 0; t < cnt; ++t) {
        const u32 Dp = masks[base + t];
        mm |= ((Dp & D) == Dp ? 1u : 0u) << t;
      }
      while (mm) {
        const int t = __ffs((int)mm) - 1;
        mm &= mm - 1u;
        const double* rj = r + (base + (u64)t) * DB;
#pragma unroll
        for (int d = 0; d < DB; ++d) w[d] += rj[d];
      }
    }
    const u32 P = D ^ (1u << x);
    const double* vp = V + (size_t)((P & zmask) ? P : 0u) * DB;     // a row-free P reads V[empty] = 0
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      const double cand = vp[d] + w[d];
      if (!seen || cand > best[d]) best[d] = cand;
    }
    seen = true;
  }
  double* vo = V + (size_t)D * DB;
#pragma unroll
  for (int d = 0; d < DB; ++d) vo[d] = best[d];
}

// one thread per draw: walk the choices down from the full set; vmax = V[full]
// (nodes without a row-carrying member are not stored: V = 0 and the lowest member is the choice)
__global__ void cg_backtrace_batch(const uint8_t* __restrict__ pred, const float* __restrict__ V, int n, int DB,
                                   int nd, int32_t* rank_pos, float* vmax, u64 full, u32 zmask) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  u64 S = full;
  vmax[d] = (full & zmask) ? V[full * DB + d] : 0.0f;
  for (int i = 0; i < n; ++i) {
    const int cur = (S & zmask) ? (int)pred[S * DB + d] : __ffsll((long long)S) - 1;
    rank_pos[d * n + cur] = i;
    S ^= 1ull << cur;
  }
}

// ---------------------------------------------------------------------------------------------
// mask layout when every row belongs to one alternative a (a panel of one alternative's choice
// shares). Every arc of another alternative weighs exactly 0, and V of a node without a is 0, so
// for D containing a, with T = compress(D minus a, a) on the other n-1 alternatives,
//   V[D] = max(0 + W[T], max over x in D, x != a, of V[D minus x] + 0) = max over T' subset of T of W[T'],
// where W[T] = w(a, D) is the float32 sum, in row order, of a's rows whose menu lies in D. A max
// returns one of its operands and adding 0.0f to a value that is never -0 returns it unchanged,
// so these are the layered sweep's values bit for bit. Along the path from the full set V stays
// at vmax = max over all T of W[T] until a is chosen, so the choice at D is the first member x
// (ascending) with: x = a and W[T] = vmax, or x != a and some T' subset of T minus x has
// W[T'] = vmax. The sweep therefore keeps one bit per node and draw, "W[T] = vmax", closed
// under supersets by OR, instead of V and pred for every node: one pass computes W with a max
// per tile, a second settles the global max, the closure runs in windows, and a thread per draw
// walks the path.
#define CL_TILE 256
#define CL_MAX_N 20
#define CG_BATCH 32

// dynamic LDS: the rows' scores, [row][DB] float32 (a chunk of at most rows_lds rows)
template <int DB>
__global__ __launch_bounds__(256) void cl_wmax(const u32* __restrict__ cm, const float* __restrict__ sc, int ra,
                                               int rows_lds, u32 NT, int tpb, u32* __restrict__ F,
                                               float* __restrict__ tmax, float* __restrict__ bmax) {
  extern __shared__ float s_sc[];
  __shared__ float red[4][DB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const bool one_chunk = ra <= rows_lds;
  if (one_chunk) {
    for (int i = tid; i < ra * DB; i += 256) s_sc[i] = sc[i];
    __syncthreads();
  }
  float brun = -INFINITY;
  for (int it = 0; it < tpb; ++it) {
    const u32 tile = (u32)blockIdx.x * (u32)tpb + (u32)it;
    if ((u64)tile * CL_TILE >= NT) break;
    const u32 T = tile * CL_TILE + (u32)tid;
    const bool valid = T < NT;
    float w[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) w[d] = 0.0f;
    for (int base = 0; base < ra; base += rows_lds) {
      const int cnt = min(rows_lds, ra - base);
      if (!one_chunk) {
        __syncthreads();
        for (int i = tid; i < cnt * DB; i += 256) s_sc[i] = sc[(size_t)base * DB + i];
        __syncthreads();
      }
      for (int sub = 0; sub < cnt; sub += 32) {
        const int c32 = min(32, cnt - sub);
        u32 mm = 0;
        for (int t = 0; t < c32; ++t) {
          const u32 m = cm[base + sub + t];
          mm |= ((m & T) == m ? 1u : 0u) << t;
        }
        // each lane adds its own matches, ascending rows
        while (mm) {
          const int t = __ffs((int)mm) - 1;
          mm &= mm - 1u;
          const float* sj = s_sc + (size_t)(sub + t) * DB;
#pragma unroll
          for (int d = 0; d < DB; ++d) w[d] += sj[d];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      float v = valid ? w[d] : -INFINITY;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
      if (lane == 0) red[wv][d] = v;
    }
    __syncthreads();
    u32 bits = 0;
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      const float tm = fmaxf(fmaxf(red[0][d], red[1][d]), fmaxf(red[2][d], red[3][d]));
      bits |= (valid && w[d] == tm ? 1u : 0u) << d;
    }
    if (valid) F[T] = bits;
    if (tid < DB) {
      const float tm = fmaxf(fmaxf(red[0][tid], red[1][tid]), fmaxf(red[2][tid], red[3][tid]));
      tmax[(size_t)tile * DB + tid] = tm;
      brun = fmaxf(brun, tm);
    }
    __syncthreads();
  }
  if (tid < DB) bmax[(size_t)blockIdx.x * DB + tid] = brun;
}

// float64 max over T of W64[T] (the lower bound's longest path in this case): no choices, so no bits
__global__ __launch_bounds__(256) void cl_wmax64(const u32* __restrict__ cm, const double* __restrict__ r, int ra,
                                                 int rows_lds, u32 NT, int tpb, double* __restrict__ bmax) {
  extern __shared__ double s_r[];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const bool one_chunk = ra <= rows_lds;
  if (one_chunk) {
    for (int i = tid; i < ra; i += 256) s_r[i] = r[i];
    __syncthreads();
  }
  double run = -INFINITY;
  for (int it = 0; it < tpb; ++it) {
    const u32 tile = (u32)blockIdx.x * (u32)tpb + (u32)it;
    if ((u64)tile * CL_TILE >= NT) break;
    const u32 T = tile * CL_TILE + (u32)tid;
    double w = 0.0;
    for (int base = 0; base < ra; base += rows_lds) {
      const int cnt = min(rows_lds, ra - base);
      if (!one_chunk) {
        __syncthreads();
        for (int i = tid; i < cnt; i += 256) s_r[i] = r[base + i];
        __syncthreads();
      }
      for (int sub = 0; sub < cnt; sub += 32) {
        const int c32 = min(32, cnt - sub);
        u32 mm = 0;
        for (int t = 0; t < c32; ++t) {
          const u32 m = cm[base + sub + t];
          mm |= ((m & T) == m ? 1u : 0u) << t;
        }
        while (mm) {
          const int t = __ffs((int)mm) - 1;
          mm &= mm - 1u;
          w += s_r[sub + t];
        }
      }
    }
    if (T < NT) run = fmax(run, w);
  }
  double v = run;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (tid == 0) bmax[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// vmax[d] = max over the blocks' maxima (one block)
template <typename T>
__global__ __launch_bounds__(256) void cl_gmax(const T* __restrict__ bmax, int nb, int DB, T* __restrict__ vmax) {
  __shared__ T red[256];
  const int tid = threadIdx.x, G = 256 / DB, d = tid % DB, g = tid / DB;
  T v = -INFINITY;
  if (g < G)
    for (int b = g; b < nb; b += G) v = v > bmax[(size_t)b * DB + d] ? v : bmax[(size_t)b * DB + d];
  red[tid] = v;
  __syncthreads();
  if (tid < DB) {
    T m = red[tid];
    for (int q = 1; q < G; ++q) m = m > red[q * DB + tid] ? m : red[q * DB + tid];
    vmax[tid] = m;
  }
}

// first closure window: 2^w consecutive nodes per block. A tile's bit for draw d stands only if
// the tile's max is the global max; then F[T] |= F[T minus b] for the window's bits b.
__global__ __launch_bounds__(256) void cl_close_lo(u32* __restrict__ F, const float* __restrict__ tmax,
                                                   const float* __restrict__ vmax, int DB, u32 NT, int w) {
  __shared__ u32 s[4096];
  __shared__ u32 keep[16];
  const int tid = threadIdx.x;
  const u32 span = 1u << w, base = (u32)blockIdx.x * span;
  const u32 ntile = (span + CL_TILE - 1) / CL_TILE;
  if (tid < (int)ntile) {
    const u32 tile = base / CL_TILE + (u32)tid;
    u32 k = 0;
    for (int d = 0; d < DB; ++d) k |= (tmax[(size_t)tile * DB + d] == vmax[d] ? 1u : 0u) << d;
    keep[tid] = k;
  }
  __syncthreads();
  for (u32 e = tid; e < span; e += 256) s[e] = F[base + e] & keep[e / CL_TILE];
  for (int b = 0; b < w; ++b) {
    __syncthreads();
    for (u32 e = tid; e < span; e += 256)
      if ((e >> b) & 1u) s[e] |= s[e ^ (1u << b)];
  }
  __syncthreads();
  for (u32 e = tid; e < span; e += 256) F[base + e] = s[e];
}

// later windows: bits b0 .. b0+w-1, a block holds 2^w x 32 nodes (32 consecutive low indexes,
// so every load and store is a 128-byte run)
__global__ __launch_bounds__(256) void cl_close_hi(u32* __restrict__ F, u32 NT, int b0, int w) {
  __shared__ u32 s[256 * 32];
  const int tid = threadIdx.x;
  const u32 L = 32, groups = (1u << b0) / L;
  const u32 g = (u32)blockIdx.x % groups, hi = (u32)blockIdx.x / groups;
  const u32 origin = g * L + (hi << (b0 + w));
  const u32 cnt = (1u << w) * L;
  for (u32 e = tid; e < cnt; e += 256) s[e] = F[origin + (e % L) + ((e / L) << b0)];
  for (int b = 0; b < w; ++b) {
    __syncthreads();
    for (u32 e = tid; e < cnt; e += 256)
      if (((e / L) >> b) & 1u) s[e] |= s[e ^ (L << b)];
  }
  __syncthreads();
  for (u32 e = tid; e < cnt; e += 256) F[origin + (e % L) + ((e / L) << b0)] = s[e];
}

// one thread per draw: the path from the full set. W[T] of the node itself (for x = a) is summed
// again here, in the same row order. err is set if no member qualifies (it cannot when the sums
// match the sweep's).
__global__ void cl_backtrace(const u32* __restrict__ F, const u32* __restrict__ cm, const float* __restrict__ sc,
                             int ra, int DB, int nd, int n, int a, const float* __restrict__ vmax,
                             int32_t* __restrict__ rank_pos, int* __restrict__ err) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const float vm = vmax[d];
  u32 D = (n >= 32) ? 0xffffffffu : ((1u << n) - 1u);
  u32 R = (1u << (n - 1)) - 1u;
  bool has_a = true;
  for (int i = 0; i < n; ++i) {
    int x = -1;
    if (has_a) {
      for (int y = 0; y < n && x < 0; ++y) {
        if (!((D >> y) & 1u)) continue;
        if (y == a) {
          float wv = 0.0f;
          for (int j = 0; j < ra; ++j)
            if ((cm[j] & R) == cm[j]) wv += sc[(size_t)j * DB + d];
          if (wv == vm) x = y;
        } else {
          const int cb = y < a ? y : y - 1;
          if ((F[R & ~(1u << cb)] >> d) & 1u) x = y;
        }
      }
      if (x < 0) {
        atomicOr(err, 1);
        x = a;
      }
    } else {
      x = __ffs((int)D) - 1;
    }
    rank_pos[(size_t)d * n + x] = i;
    D &= ~(1u << x);
    if (x == a) has_a = false;
    else if (has_a) R &= ~(1u << (x < a ? x : x - 1));
  }
}

// ---------------------------------------------------------------------------------------------
// One-alternative sweep in window passes (sweep mode 2). It rests on the identity and the path
// rule above (V[D] = max over T' subset of T of W[T'] for D containing a; along the path from the
// full set V stays at vmax until a is chosen; the choice at D is the first member x, ascending,
// with x = a and W[T] = vmax, or x != a and some T' subset of T minus x with W[T'] = vmax), in
// kernels of their own:
//   oa_wbits   W for 256 consecutive T per block (a thread per T, a's menus through uniform loads,
//              the rows' scores staged in LDS with a 33-float stride, each lane's matches summed in
//              row order per draw), the block's maximum per draw, and per T the bits "W = block max"
//   oa_gmax    the global maximum per draw (two levels)
//   oa_or_lo   per 4096 T: drop the bits of blocks whose maximum is not the global one, then close the
//              bits under supersets over T bits 0..11 (OR, in LDS)
//   oa_or_hi   the closure over 8 more bits at a time, 32 consecutive T (128 bytes) per row
//   oa_bt      a block per draw walks the path (W of the node summed again in row order)
// 4 bytes per subset of the other alternatives (64 MB at n = 25) for 32 draws. Every value compared
// is a float32 sum in the layered sweep's order and a maximum returns an operand, so the rankings
// and vmax are the layered sweep's bit for bit.
#define OA_DB 32         // draws per sweep (a batch is padded to 32)
#define OA_ROWS 256      // a's rows staged in LDS at a time
#define OA_LO 12         // T bits closed by oa_or_lo (4096 T per block)
#define OA_HI 8          // T bits per oa_or_hi pass
#define OA_BT_ROWS 4096  // a's rows the path walk holds in LDS (the sweep mode is refused above)

// W of this lane's T for 32 draws (w, float32) and, with L > 0, for L more draws in float64 (w64,
// the float64 bound's row sums of the residuals r, [ra][L]): the same matches, each added in row
// order, so w64 equals oa_wmax64's sums bit for bit
template <int L>
__device__ __forceinline__ void oa_w_rows(const u32* __restrict__ cm, const float* __restrict__ sc,
                                          const double* __restrict__ r, int ra, u32 T, bool own, float* srow,
                                          double* srow64, float (&w)[OA_DB], double (&w64)[L > 0 ? L : 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int d = 0; d < OA_DB; ++d) w[d] = 0.0f;
#pragma unroll
  for (int l = 0; l < (L > 0 ? L : 1); ++l) w64[l] = 0.0;
  for (int base = 0; base < ra; base += OA_ROWS) {
    const int cnt = min(OA_ROWS, ra - base);
    __syncthreads();
    for (int i = tid; i < cnt * OA_DB; i += blockDim.x)
      srow[(i / OA_DB) * (OA_DB + 1) + (i % OA_DB)] = sc[(size_t)base * OA_DB + i];
    if constexpr (L > 0)
      for (int i = tid; i < cnt * L; i += blockDim.x) srow64[i] = r[(size_t)base * L + i];
    __syncthreads();
    for (int sub = 0; sub < cnt; sub += 32) {
      u32 mm = 0;
#pragma unroll
      for (int t = 0; t < 32; ++t) {                  // padded to 32: no T contains a padding menu
        const u32 m = cm[base + sub + t];
        mm |= ((m & T) == m ? 1u : 0u) << t;
      }
      if (!own) mm = 0;
      while (mm) {                                   // this lane's matches, ascending rows
        const int t = __ffs((int)mm) - 1;
        mm &= mm - 1u;
        const float* sj = srow + (sub + t) * (OA_DB + 1);
#pragma unroll
        for (int d = 0; d < OA_DB; ++d) w[d] += sj[d];
        if constexpr (L > 0) {
          const double* rj = srow64 + (sub + t) * L;
#pragma unroll
          for (int l = 0; l < L; ++l) w64[l] += rj[l];
        }
      }
    }
  }
}

// L > 0: also the float64 maxima of L more draws' row sums over the block's T (bmax64[block][L]),
// the float64 bound of the draws that finish beside this sweep's pricing, on the same subset tests
template <int L>
__global__ __launch_bounds__(256) void oa_wbits(const u32* __restrict__ cm, const float* __restrict__ sc,
                                                const double* __restrict__ r, int ra, u32 NT, u32* __restrict__ F,
                                                float* __restrict__ bmax, double* __restrict__ bmax64) {
  __shared__ float srow[OA_ROWS * (OA_DB + 1)];      // the staged rows, then the block's W rows
  __shared__ double srow64[OA_ROWS * (L > 0 ? L : 1)];
  __shared__ double red64[4][L > 0 ? L : 1];
  __shared__ float part[8][OA_DB];
  __shared__ float bm[OA_DB];
  const int tid = threadIdx.x;
  const u32 T = (u32)blockIdx.x * 256u + (u32)tid;
  const bool own = T < NT;
  float w[OA_DB];
  double w64[L > 0 ? L : 1];
  oa_w_rows<L>(cm, sc, r, ra, T, own, srow, srow64, w, w64);
  if constexpr (L > 0) {
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      double v = own ? w64[l] : -INFINITY;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double u = __shfl_xor(v, off);
        v = u > v ? u : v;
      }
      if (lane == 0) red64[wv][l] = v;
    }
  }
  // the block's maximum per draw: W rows through LDS, thread (g, d) takes rows 32 g .. 32 g + 31
  __syncthreads();
#pragma unroll
  for (int d = 0; d < OA_DB; ++d) srow[tid * (OA_DB + 1) + d] = own ? w[d] : -INFINITY;
  __syncthreads();
  {
    const int d = tid % OA_DB, g = tid / OA_DB;
    float m = -INFINITY;
    for (int k = 0; k < 32; ++k) {
      const float v = srow[(g * 32 + k) * (OA_DB + 1) + d];
      m = v > m ? v : m;
    }
    part[g][d] = m;
  }
  __syncthreads();
  if (tid < OA_DB) {
    float m = part[0][tid];
    for (int g = 1; g < 8; ++g) m = part[g][tid] > m ? part[g][tid] : m;
    bm[tid] = m;
    bmax[(size_t)blockIdx.x * OA_DB + tid] = m;
  }
  if constexpr (L > 0) {
    if (tid >= 64 && tid < 64 + L) {               // red64 was written before the syncs above
      const int l = tid - 64;
      double m = red64[0][l];
      for (int q = 1; q < 4; ++q) m = red64[q][l] > m ? red64[q][l] : m;
      bmax64[(size_t)blockIdx.x * L + l] = m;
    }
  }
  __syncthreads();
  u32 bits = 0;
#pragma unroll
  for (int d = 0; d < OA_DB; ++d) bits |= (w[d] == bm[d] ? 1u : 0u) << d;
  if (own) F[T] = bits;
}

// maxima per draw over blocks: block g of the grid takes blocks g, g + G, ... (one thread per
// (group of 8, draw)); called twice, the second time with one block
__global__ __launch_bounds__(256) void oa_gmax(const float* __restrict__ in, int nin, float* __restrict__ out) {
  __shared__ float red[256];
  const int tid = threadIdx.x, d = tid % OA_DB, q = tid / OA_DB;       // 8 groups of 32
  float m = -INFINITY;
  for (int b = (int)blockIdx.x * 8 + q; b < nin; b += (int)gridDim.x * 8) {
    const float v = in[(size_t)b * OA_DB + d];
    m = v > m ? v : m;
  }
  red[tid] = m;
  __syncthreads();
  if (tid < OA_DB) {
    for (int k = 1; k < 8; ++k) m = red[k * OA_DB + tid] > m ? red[k * OA_DB + tid] : m;
    out[(size_t)blockIdx.x * OA_DB + tid] = m;
  }
}

// per span = 2^lo consecutive T (lo = min(12, n - 1)): a T's bit for draw d stands only where its
// block's maximum is the global one; then F[T] |= F[T minus b] for the bits b < lo
__global__ __launch_bounds__(256) void oa_or_lo(u32* __restrict__ F, const float* __restrict__ bmax,
                                                const float* __restrict__ vmax, int lo) {
  __shared__ u32 s[1 << OA_LO];
  __shared__ u32 keep[(1 << OA_LO) / 256];
  const int tid = threadIdx.x;
  const u32 span = 1u << lo, base = (u32)blockIdx.x * span;
  const u32 nsub = (span + 255u) / 256u;
  if (tid < (int)nsub) {
    const size_t blk = base / 256u + (u32)tid;
    u32 k = 0;
    for (int d = 0; d < OA_DB; ++d) k |= (bmax[blk * OA_DB + d] == vmax[d] ? 1u : 0u) << d;
    keep[tid] = k;
  }
  __syncthreads();
  for (u32 e = tid; e < span; e += 256u) s[e] = F[base + e] & keep[e / 256u];
  for (int b = 0; b < lo; ++b) {
    __syncthreads();
    for (u32 e = tid; e < span; e += 256u)
      if ((e >> b) & 1u) s[e] |= s[e ^ (1u << b)];
  }
  __syncthreads();
  for (u32 e = tid; e < span; e += 256u) F[base + e] = s[e];
}

// the closure over T bits b0 .. b0+w-1 (b0 >= 12, w <= 8): a block holds 2^w rows of 32 consecutive
// T (128 bytes each): T = g 32 + j + k 2^b0 + hi 2^(b0 + w); grid = 2^(n - 1 - w - 5) blocks
__global__ __launch_bounds__(256) void oa_or_hi(u32* __restrict__ F, int b0, int w) {
  __shared__ u32 s[(1 << OA_HI) * 32];
  const int tid = threadIdx.x;
  const u32 groups = 1u << (b0 - 5);
  const u32 g = (u32)blockIdx.x & (groups - 1u), hi = (u32)blockIdx.x >> (b0 - 5);
  const u32 origin = g * 32u + (hi << (b0 + w));
  const int cnt = (1 << w) * 32;
  for (int e = tid; e < cnt; e += 256) s[e] = F[origin + (u32)(e & 31) + ((u32)(e >> 5) << b0)];
  for (int b = 0; b < w; ++b) {
    __syncthreads();
    for (int e = tid; e < cnt; e += 256)
      if (((e >> 5) >> b) & 1) s[e] |= s[e ^ (32 << b)];
  }
  __syncthreads();
  for (int e = tid; e < cnt; e += 256) F[origin + (u32)(e & 31) + ((u32)(e >> 5) << b0)] = s[e];
}

// a block of 64 lanes per draw: the walk from the full set. Per step the lanes read the closed
// bits of the predecessors (lane y for member y != a), test a's rows for the node (64 rows per
// ballot; lane 0 sums the matches in row order and compares with vmax), and lane 0 takes the
// first member in ascending order that qualifies. err is set if none does (it cannot when the
// sums are the sweep's).
__global__ __launch_bounds__(64) void oa_bt(const u32* __restrict__ F, const u32* __restrict__ cm,
                                            const float* __restrict__ sc, int ra, int n, int a,
                                            const float* __restrict__ vmax, int32_t* __restrict__ rank_pos,
                                            int* __restrict__ err) {
  __shared__ u32 s_cm[OA_BT_ROWS];
  __shared__ float s_sc[OA_BT_ROWS];
  __shared__ int ok[32];
  __shared__ int pick;
  const int d = blockIdx.x, lane = threadIdx.x;
  for (int j = lane; j < ra; j += 64) {
    s_cm[j] = cm[j];
    s_sc[j] = sc[(size_t)j * OA_DB + d];
  }
  const float vm = vmax[d];
  u32 T = (1u << (n - 1)) - 1u, D = (1u << n) - 1u;
  __syncthreads();
  bool has_a = true;
  for (int i = 0; i < n; ++i) {
    int x;
    if (has_a) {
      if (lane < n && lane != a)
        ok[lane] = ((D >> lane) & 1u) ? (int)((F[T & ~(1u << (lane < a ? lane : lane - 1))] >> d) & 1u) : 0;
      float wv = 0.0f;                                 // lane 0: W of the node, rows in order
      for (int base = 0; base < ra; base += 64) {
        const int j = base + lane;
        const bool hit = j < ra && (s_cm[j] & T) == s_cm[j];
        u64 mm = __ballot(hit);
        if (lane == 0) {
          while (mm) {
            const int t = __ffsll((unsigned long long)mm) - 1;
            mm &= mm - 1ull;
            wv += s_sc[base + t];
          }
        }
      }
      if (lane == 0) ok[a] = wv == vm ? 1 : 0;
      __syncthreads();
      if (lane == 0) {
        int bx = -1;
        for (int y = 0; y < n && bx < 0; ++y)
          if (((D >> y) & 1u) && ok[y]) bx = y;
        if (bx < 0) {
          atomicOr(err, 1);
          bx = a;
        }
        pick = bx;
      }
      __syncthreads();
      x = pick;
      __syncthreads();
    } else {
      x = __ffs((int)D) - 1;
    }
    if (lane == 0) rank_pos[(size_t)d * n + x] = i;
    D &= ~(1u << x);
    if (x == a) has_a = false;
    else if (has_a) T &= ~(1u << (x < a ? x : x - 1));
  }
}

// float64: max over T of W64[T] for DB <= 8 draws (the lower bound's longest path with one row
// alternative); each block writes its maxima, oa_max64_final reduces them
template <int DB>
__global__ __launch_bounds__(256) void oa_wmax64(const u32* __restrict__ cm, const double* __restrict__ r, int ra,
                                                 u32 NT, double* __restrict__ bmax) {
  __shared__ double srow[OA_ROWS * (DB + 1)];
  __shared__ double red[4][DB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double run[DB];
#pragma unroll
  for (int d = 0; d < DB; ++d) run[d] = -INFINITY;
  const bool one_chunk = ra <= OA_ROWS;
  if (one_chunk) {
    for (int i = tid; i < ra * DB; i += 256) srow[(i / DB) * (DB + 1) + (i % DB)] = r[i];
    __syncthreads();
  }
  for (u32 T0 = (u32)blockIdx.x * 256u; T0 < NT; T0 += gridDim.x * 256u) {
    const u32 T = T0 + (u32)tid;
    const bool own = T < NT;
    double w[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) w[d] = 0.0;
    for (int base = 0; base < ra; base += OA_ROWS) {
      const int cnt = min(OA_ROWS, ra - base);
      if (!one_chunk) {
        __syncthreads();
        for (int i = tid; i < cnt * DB; i += 256) srow[(i / DB) * (DB + 1) + (i % DB)] = r[(size_t)base * DB + i];
        __syncthreads();
      }
      for (int sub = 0; sub < cnt; sub += 32) {
        u32 mm = 0;
#pragma unroll
        for (int t = 0; t < 32; ++t) {
          const u32 m = cm[base + sub + t];
          mm |= ((m & T) == m ? 1u : 0u) << t;
        }
        if (!own) mm = 0;
        while (mm) {
          const int t = __ffs((int)mm) - 1;
          mm &= mm - 1u;
#pragma unroll
          for (int d = 0; d < DB; ++d) w[d] += srow[(sub + t) * (DB + 1) + d];
        }
      }
    }
    if (own) {
#pragma unroll
      for (int d = 0; d < DB; ++d) run[d] = w[d] > run[d] ? w[d] : run[d];
    }
  }
#pragma unroll
  for (int d = 0; d < DB; ++d) {
    double v = run[d];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double u = __shfl_xor(v, off);
      v = u > v ? u : v;
    }
    if (lane == 0) red[wv][d] = v;
  }
  __syncthreads();
  if (tid < DB) {
    double m = red[0][tid];
    for (int q = 1; q < 4; ++q) m = red[q][tid] > m ? red[q][tid] : m;
    bmax[(size_t)blockIdx.x * DB + tid] = m;
  }
}

// one block: thread (g, d) takes every (256 / DB)-th block's maximum of draw d, then LDS
// block g of the grid: the maxima over inputs g, g + G, ... per draw (DB <= 8), out[g][DB]; a grid
// of G blocks before oa_max64_final over many block maxima
__global__ __launch_bounds__(256) void oa_max64_part(const double* __restrict__ in, int nin, int DB,
                                                     double* __restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x, G = 256 / DB, d = tid % DB, q = tid / DB;
  double m = -INFINITY;
  if (q < G)
    for (int b = (int)blockIdx.x * G + q; b < nin; b += (int)gridDim.x * G)
      m = in[(size_t)b * DB + d] > m ? in[(size_t)b * DB + d] : m;
  red[tid] = m;
  __syncthreads();
  if (tid < DB) {
    double v = red[tid];
    for (int k = 1; k < G; ++k) v = red[k * DB + tid] > v ? red[k * DB + tid] : v;
    out[(size_t)blockIdx.x * DB + tid] = v;
  }
}

__global__ __launch_bounds__(256) void oa_max64_final(const double* __restrict__ bmax, int nblk, int DB,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x, G = 256 / DB, d = tid % DB, g = tid / DB;
  double m = -INFINITY;
  if (g < G)
    for (int b = g; b < nblk; b += G) m = bmax[(size_t)b * DB + d] > m ? bmax[(size_t)b * DB + d] : m;
  red[tid] = m;
  __syncthreads();
  if (tid < DB) {
    double v = red[tid];
    for (int q = 1; q < G; ++q) v = red[q * DB + tid] > v ? red[q * DB + tid] : v;
    out[tid] = v;
  }
}

// Slab layout, first window of the subset-sum butterfly (at bit 0, 2^B consecutive positions per
// tile): the tile starts at zero with the rows whose slab position falls inside it (rows sorted by
// position, zoff[tile] .. zoff[tile + 1]), then takes the window's stages. The same values as
// scattering the scores on a zeroed slab and running the plain pass on it.
template <int RL>
__global__ __launch_bounds__(512) void cg_zeta_first(float* __restrict__ R, WinGeom g, const i64* __restrict__ zoff,
                                                     const i64* __restrict__ zpos, const int32_t* __restrict__ zrow,
                                                     const float* __restrict__ scores) {
  extern __shared__ __align__(16) unsigned char smem[];
  Tile<float, RL> t;
  t.init(g, reinterpret_cast<float*>(smem));
  constexpr int NR = Tile<float, RL>::R;
  const int tl = threadIdx.x / g.TT, tt = threadIdx.x - tl * g.TT;
  const i64 tile = (i64)blockIdx.x * g.TPB + tl;
  const int E = 1 << g.B;
  for (int e = tt; e < E; e += g.TT) t.lds[e ^ ((e >> 4) & 15)] = 0.0f;
  __syncthreads();
  if (t.active) {
    for (i64 q = zoff[tile] + tt; q < zoff[tile + 1]; q += g.TT) {
      const int e = (int)(zpos[q] - t.base);
      t.lds[e ^ ((e >> 4) & 15)] = scores[zrow[q]];
    }
  }
  __syncthreads();
  if (t.active) {
#pragma unroll
    for (int j = 0; j < NR; ++j) t.v[j] = t.lds[t.lidx(t.cur, j)];
  }
  t.transform(Coef<float>{1.0f, 0.0f, 1.0f, 1.0f});
  t.store(R);
}

// Slab layout's layered max, tile by tile. V[D] = max over members x of D of V[D - x] + R_x[D - x]
// (R_x at slab position x Nloc + compress(D - x, x)), the lowest x on ties, for every subset D. The n
// bits are split into L low bits and n - L high bits; tile h holds the 2^L subsets h | l, which are
// contiguous in V and pred. A member x < L leads to a subset of the same tile, and its arc weights
// for the tile are one contiguous run R[x Nloc + (h >> 1) + c], c < 2^(L-1); a member x >= L leads
// to tile h - x, at the same l, with contiguous arc weights too. Tiles are launched by the popcount of
// h, so the tiles they read are finished; inside a tile the subsets go by the popcount of l in LDS,
// with the tile's low arcs staged in LDS first. One wavefront works one tile, so the steps between
// the tile's popcount layers cost no more than a wavefront's own LDS ordering, and many tiles are
// resident per CU. Every load from HBM is a contiguous run, each arc weight is read once, and V and
// pred are written once, in order. A node's value and choice are those of the one-node-per-thread
// form: each candidate is V + R rounded once, the members are taken in ascending order with a
// strict >, and the high members (all above the low ones) are folded in after the low ones with a
// strict >, which selects the same member.
#define CG_SLAB_TILE_THREADS 64
template <int L>
__global__ __launch_bounds__(CG_SLAB_TILE_THREADS) void cg_slab_tile(float* __restrict__ V, uint8_t* __restrict__ pred,
                                                                     const float* __restrict__ R,
                                                                     const u32* __restrict__ hlist,
                                                                     const uint16_t* __restrict__ lord, i64 Nloc) {
  constexpr int T = 1 << L, H = T >> 1, NT = CG_SLAB_TILE_THREADS;
  __shared__ float vt[T];
  __shared__ float hv[T];
  __shared__ uint8_t hx[T];
  __shared__ float rl[L][H];
  const int tid = threadIdx.x;
  const u32 h = hlist[blockIdx.x];
  const u32 hm = h << L;
  for (int e = tid; e < L * H; e += NT) {
    const int x = e / H, cidx = e - x * H;
    rl[x][cidx] = R[(i64)x * Nloc + (i64)(hm >> 1) + cidx];
  }
  // the high members, ascending, for every subset of the tile. They are the tile's high part, the
  // same for every lane, so each step takes four members and issues the V and arc-weight loads of
  // all the lane's subsets (a contiguous run across the lanes for each member) before folding them
  // in member order (strict >: ties keep the lowest member)
  constexpr int PER = T > NT ? T / NT : 1;
  float hb[PER];
  int hbx[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) hb[q] = 0.0f, hbx[q] = 0;
  const bool on = T >= NT || tid < T;
  bool seen = false;
  for (u32 S = h; S;) {
    int xs[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      xs[m] = S ? L + __builtin_ctz(S) : -1;
      S &= S - 1;
    }
    float cand[4][PER];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        cand[m][q] = 0.0f;
        if (xs[m] >= 0 && on) {
          const u32 P = (hm | (u32)(tid + q * NT)) ^ (1u << xs[m]);
          cand[m][q] = V[P] + R[(i64)xs[m] * Nloc + (i64)compress_bit(P, xs[m])];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (xs[m] < 0) continue;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (!seen || cand[m][q] > hb[q]) {
          hb[q] = cand[m][q];
          hbx[q] = xs[m];
        }
      }
      seen = true;
    }
  }
  if (on) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      hv[tid + q * NT] = hb[q];
      hx[tid + q * NT] = (uint8_t)hbx[q];
    }
  }
  __syncthreads();
  if (tid == 0) vt[0] = h ? hv[0] : 0.0f;       // the tile's own empty low part (V[empty set] = 0)
  __syncthreads();
  int lo = 1, cnt = L;
  for (int c = 1; c <= L; ++c) {
    for (int i = tid; i < cnt; i += NT) {
      const u32 l = lord[lo + i];
      float b = 0.0f;
      int bx = 0;
      bool seen = false;
      for (u32 S = l; S; S &= S - 1) {
        const int x = __builtin_ctz(S);
        const u32 pl = l ^ (1u << x);
        const float cand = vt[pl] + rl[x][compress_bit(pl, x)];
        if (!seen || cand > b) {
          b = cand;
          bx = x;
          seen = true;
        }
      }
      if (h && hv[l] > b) {
        b = hv[l];
        bx = hx[l];
      }
      vt[l] = b;
      hx[l] = (uint8_t)bx;
    }
    __syncthreads();
    lo += cnt;
    cnt = cnt * (L - c) / (c + 1);
  }
  for (int l = tid; l < T; l += NT) {
    if ((hm | (u32)l) == 0u) continue;           // the empty set keeps V = 0 and has no choice
    V[hm | (u32)l] = vt[l];
    pred[hm | (u32)l] = hx[l];
  }
}

// walk pred from the full set (one thread): rank_pos[x] = position of x, 0 = best
__global__ void cg_backtrace(const uint8_t* __restrict__ pred, int n, int32_t* rank_pos, u64 full) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  u64 D = full;
  for (int i = 0; i < n; ++i) {
    const int cur = pred[D];
    rank_pos[cur] = i;
    D ^= 1ull << cur;
  }
}

__global__ void cg_fill_f32(float* a, float v, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void cg_fill_f64(double* a, double v, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void cg_fill_u8(uint8_t* a, uint8_t v, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void cg_fill_i32(int* a, int v, i64 m) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) a[i] = v;
}

// ---------------------------------------------------------------------------------------------
// stream layout: one layer k of colex ranks, 1024 ranks per block (256 threads x 4).
#define CG_BLOCK 1024
#define CG_THREADS 256
#define CG_VPT 4

// float64 value of one colex layer (sweep_max64, pricing_stream.py:190-248): the same unranking,
// predecessor ranks and block-uniform element loop as cg_stream_layer (one rank per thread), values
// and row sums in float64, no choices kept.
__global__ __launch_bounds__(CG_THREADS) void cg_stream_layer64(const u64* __restrict__ binom, int n, int k,
                                                              u64 count, const double* __restrict__ vprev,
                                                              double* __restrict__ vcur,
                                                              const u64* __restrict__ masks,
                                                              const double* __restrict__ r,
                                                              const u64* __restrict__ offsets) {
  const u64 rk = (u64)blockIdx.x * CG_THREADS + threadIdx.x;
  const int np1 = n + 1;
  u64 D = 0;
  if (rk < count) {
    u64 rr = rk;
    int b = n - 1;
    for (int t = k; t >= 1; --t) {
      while (binom[b * np1 + t] > rr) --b;
      D |= 1ull << b;
      rr -= binom[b * np1 + t];
      --b;
    }
  }
  double best = 0.0;
  bool seen = false;
  u64 suf = 0, remain = rk;
  int t = k - 1;
  for (int x = n - 1; x >= 0; --x) {
    double w = 0.0;
    const u64 hi = offsets[x + 1];
#pragma unroll 8
    for (u64 i = offsets[x]; i < hi; ++i) {
      const u64 Dp = masks[i];
      if ((Dp & D) == Dp) w += r[i];
    }
    if ((D >> x) & 1ull) {
      remain -= binom[x * np1 + (t + 1)];
      const u64 rp = remain + suf;
      suf += binom[x * np1 + t];
      --t;
      const double cand = vprev[rp] + w;
      if (!seen || cand > best) best = cand;
      seen = true;
    }
  }
  if (rk < count) vcur[rk] = best;
}

template <int MODE>
__device__ __forceinline__ float cg_load_prev(const void* vprev, const uint8_t* lprev, const float* mprev,
                                              const float* eprev, u64 rp) {
  if constexpr (MODE == 0) {
    return ((const float*)vprev)[rp];
  } else if constexpr (MODE == 1) {
    const float q = (float)((const uint16_t*)vprev)[rp];
    return fmaf(q, eprev[rp / CG_BLOCK], mprev[rp / CG_BLOCK]);
  } else {
    const unsigned qi = ((unsigned)((const uint16_t*)vprev)[rp] << 8) | (unsigned)lprev[rp];
    return fmaf((float)qi, eprev[rp / CG_BLOCK], mprev[rp / CG_BLOCK]);
  }
}

// One colex layer of the row-order sweep. Each thread owns CG_VPT ranks; the element x runs from
// n-1 down to 0 in step across the block, so the rows of alternative x (masks, scores) are the same
// for every lane: they come through scalar loads once per wave and each is tested against the
// thread's CG_VPT subsets in registers. A lane meets its own members in the same order as a walk
// from its highest member down, so predecessor ranks, row-order sums and the tie rule (a later,
// lower member wins ties) are those of that walk.
template <int MODE>
__global__ __launch_bounds__(CG_THREADS) void cg_stream_layer(const u64* __restrict__ binom, int n, int k, u64 count,
                                                            const void* vprev, const uint8_t* lprev, const float* mprev,
                                                            const float* eprev, void* vcur, uint8_t* lcur, float* mcur,
                                                            float* ecur, uint8_t* pred, int store_pred,
                                                            const u64* __restrict__ masks,
                                                            const float* __restrict__ scores,
                                                            const u64* __restrict__ offsets) {
  const u64 base = (u64)blockIdx.x * CG_BLOCK;
  const int np1 = n + 1;
  u64 D[CG_VPT], remain[CG_VPT], suf[CG_VPT];
  int t[CG_VPT];
  float vals[CG_VPT];
  uint8_t preds[CG_VPT];
  bool seen[CG_VPT];
#pragma unroll
  for (int j = 0; j < CG_VPT; ++j) {
    const u64 r = base + (u64)j * CG_THREADS + threadIdx.x;
    D[j] = 0;
    remain[j] = r;
    suf[j] = 0;
    t[j] = k - 1;
    vals[j] = 0.0f;
    preds[j] = 0;
    seen[j] = false;
    if (r < count) {
      // colex unranking: the largest element is the largest b with C(b, t) <= r
      u64 rr = r;
      int b = n - 1;
      for (int tt = k; tt >= 1; --tt) {
        while (binom[b * np1 + tt] > rr) --b;
        D[j] |= 1ull << b;
        rr -= binom[b * np1 + tt];
        --b;
      }
    }
  }
  for (int x = n - 1; x >= 0; --x) {
    float w[CG_VPT];
#pragma unroll
    for (int j = 0; j < CG_VPT; ++j) w[j] = 0.0f;
    const u64 hi = offsets[x + 1];
#pragma unroll 8
    for (u64 i = offsets[x]; i < hi; ++i) {
      const u64 Dp = masks[i];
      const float sc = scores[i];
#pragma unroll
      for (int j = 0; j < CG_VPT; ++j)
        if ((Dp & D[j]) == Dp) w[j] += sc;
    }
#pragma unroll
    for (int j = 0; j < CG_VPT; ++j) {
      if ((D[j] >> x) & 1ull) {
        // predecessor rank from prefix/suffix binomial sums
        remain[j] -= binom[x * np1 + (t[j] + 1)];
        const u64 rp = remain[j] + suf[j];
        suf[j] += binom[x * np1 + t[j]];
        --t[j];
        const float cand = cg_load_prev<MODE>(vprev, lprev, mprev, eprev, rp) + w[j];
        if (!seen[j] || cand >= vals[j]) {
          vals[j] = cand;
          seen[j] = true;
          preds[j] = (uint8_t)x;
        }
      }
    }
  }
  float tmin = 3.0e38f, tmax = -3.0e38f;
#pragma unroll
  for (int j = 0; j < CG_VPT; ++j) {
    if (base + (u64)j * CG_THREADS + threadIdx.x < count) {
      tmin = vals[j] < tmin ? vals[j] : tmin;
      tmax = vals[j] > tmax ? vals[j] : tmax;
    }
  }
  if constexpr (MODE == 0) {
#pragma unroll
    for (int j = 0; j < CG_VPT; ++j) {
      const u64 r = base + (u64)j * CG_THREADS + threadIdx.x;
      if (r < count) {
        ((float*)vcur)[r] = vals[j];
        if (store_pred) pred[r] = preds[j];
      }
    }
  } else {
    __shared__ float smin[CG_THREADS / 64], smax[CG_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) {
      tmin = fminf(tmin, __shfl_xor(tmin, o, 64));
      tmax = fmaxf(tmax, __shfl_xor(tmax, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      smin[threadIdx.x >> 6] = tmin;
      smax[threadIdx.x >> 6] = tmax;
    }
    __syncthreads();
    float bmin = smin[0], bmax = smax[0];
    for (int i = 1; i < CG_THREADS / 64; ++i) {
      bmin = fminf(bmin, smin[i]);
      bmax = fmaxf(bmax, smax[i]);
    }
    const float levels = MODE == 1 ? 65535.0f : 16777215.0f;
    const float step = (bmax - bmin) / levels;     // 0 when the block is constant: q = 0
    if (threadIdx.x == 0) {
      mcur[blockIdx.x] = bmin;
      ecur[blockIdx.x] = step;
    }
#pragma unroll
    for (int j = 0; j < CG_VPT; ++j) {
      const u64 r = base + (u64)j * CG_THREADS + threadIdx.x;
      if (r < count) {
        float q = (step > 0.0f) ? rintf((vals[j] - bmin) / step) : 0.0f;
        q = fminf(fmaxf(q, 0.0f), levels);
        const unsigned qi = (unsigned)q;
        if constexpr (MODE == 1) {
          ((uint16_t*)vcur)[r] = (uint16_t)qi;
        } else {
          ((uint16_t*)vcur)[r] = (uint16_t)(qi >> 8);
          lcur[r] = (uint8_t)(qi & 255u);
        }
        if (store_pred) pred[r] = preds[j];
      }
    }
  }
}

// largest quantization step of a layer: step[k] = max over blocks (one block of 256)
__global__ void cg_layer_step(const float* __restrict__ e, u64 nblk, float* out) {
  __shared__ float sh[256];
  float m = 0.0f;
  for (u64 i = threadIdx.x; i < nblk; i += 256) m = fmaxf(m, e[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// ---------------------------------------------------------------------------------------------
struct rumgpu_cg {
  rumgpu_ctx* ctx = nullptr;
  int n = 0, layout = 0, dtype = 0, k_pred = 1;
  i64 M = 0, NN = 0, Nloc = 0;
  std::vector<u64> h_masks, h_offsets;
  u32 zmask = 0;             // alternatives with at least one row (mask layout, n <= 31)
  // one-alternative sweep (mask layout, every row belongs to alternative single_a)
  int sweep = 0;             // 0: layer by layer for any rows; 1: the closure sweep, 2: the window-pass
                             // sweep, when every row belongs to one alternative
  int single_a = -1;
  int ra = 0;                // rows of single_a
  u32* clcm = nullptr;       // their menus without a, compressed to the other n-1 bits
  u32* clF = nullptr;        // per node T: bit d = "W[T] = vmax" for draw d, then closed under supersets
  float* cltmax = nullptr;   // per tile of 256 nodes and draw: the tile's max of W
  float* clbmax = nullptr;   // per block and draw
  float* clvmax = nullptr;
  double* clbmax64 = nullptr;
  double* clv64 = nullptr;
  int* clerr = nullptr;
  u32 clNT = 0;
  int cl_grid = 0, cl_tpb = 0;
  // window-pass one-alternative sweep (sweep mode 2)
  u32* oaF = nullptr;        // per T: the closed bits of 32 draws (4 bytes per subset of the others)
  float* oabm = nullptr;     // per block of 256 T and draw: the block's maximum of W
  float* oagp = nullptr;     // per block of the first oa_gmax level and draw
  int* oaerr = nullptr;
  float* oavmax = nullptr;   // 32 draws
  double* oab64 = nullptr;   // per block and draw (8) float64 maxima
  double* oafb64 = nullptr;  // per block of 256 T and draw (8): float64 maxima of a fused sweep
  double* oav64 = nullptr;   // 8 draws
  std::vector<i64> bounds;
  // device rows
  u32* masks32 = nullptr;
  u64* masks64 = nullptr;
  u64* offsets = nullptr;
  float* scores = nullptr;
  // slab layout: the rows' slab positions sorted ascending (zpos) with their rows (zrow), and per
  // tile of the first butterfly window the range of them inside it (zoff)
  i64* zpos = nullptr;
  int32_t* zrow = nullptr;
  i64* zoff = nullptr;
  int zB = 0;                // bits of the first window (its tiles are 2^zB consecutive positions)
  // mask / slab
  float* V = nullptr;
  uint8_t* pred = nullptr;
  u32* nodes = nullptr;
  // batched mask sweeps (allocated at the first rumgpu_cg_price_batch)
  float* Vb = nullptr;
  uint8_t* predb = nullptr;
  float* scoresb = nullptr;
  float* vmaxb = nullptr;
  int32_t* rankb = nullptr;
  // float64 lower-bound sweep (allocated at the first rumgpu_cg_lb64)
  double* V64 = nullptr;       // mask: V per node; stream: two layers of cap64 values
  double* r64 = nullptr;
  double* V64b = nullptr;      // batched float64 bound: [node][CG_LB_BATCH] (allocated at first use)
  double* r64b = nullptr;
  u64 cap64 = 0;
  float* R = nullptr;        // slab
  // tiles of the slab layout's layered max: high subsets by popcount, local subsets by popcount
  int tL = 0;
  u32* thlist = nullptr;
  uint16_t* tlord = nullptr;
  std::vector<i64> thbnd;
  int32_t* rank = nullptr;
  // column store of one column-generation run (rumgpu_cg_cols_*)
  u64* cbits = nullptr;      // [ccap][cW]: bit r % 64 of word r / 64 = x-major row r is in the column
  i64 ccap = 0, cK = 0, cW = 0;
  double* cp = nullptr;      // observed p per x-major row
  double* cpout = nullptr;   // V lam of the last rumgpu_cg_cols_scores
  unsigned long long* cq = nullptr;   // inner products of the last added column [ccap]
  int32_t* cidx = nullptr;   // support of lam [ccap]
  double* clam = nullptr;    // lam on that support [ccap]
  int32_t* crank = nullptr;  // rank positions of the column being added
  double* cpart = nullptr;   // per block of cg_col_bits: the sum of p over the new column's rows
  bool scores_resident = false;   // c->scores holds float(p - V lam) from rumgpu_cg_cols_scores
  // stream
  u64* binom = nullptr;
  std::vector<u64> h_binom;
  void* vbuf[2] = {nullptr, nullptr};
  uint8_t* lbuf[2] = {nullptr, nullptr};
  float* mbuf[2] = {nullptr, nullptr};
  float* ebuf[2] = {nullptr, nullptr};
  float* lstep = nullptr;
  std::vector<uint8_t*> predk;   // per layer k (k >= k_pred)
  // stream sub-lattice (mask DP on 2^k_pred nodes)
  float* subV = nullptr;
  uint8_t* subpred = nullptr;
  u32* subnodes = nullptr;
  std::vector<i64> subbounds;
  u32* submasks = nullptr;
  float* subscores = nullptr;
  u64* suboffsets = nullptr;
  size_t device_bytes = 0;
  double last_ms = 0.0;
  // the layered sweep (mask or slab layout) captured once and replayed per sweep; rebuilt when the
  // slab tiles change
  hipGraph_t sgraph = nullptr;
  hipGraphExec_t sgexec = nullptr;
};

static void sweep_graph_drop(rumgpu_cg* c) {
  if (c->sgexec) (void)hipGraphExecDestroy(c->sgexec);
  if (c->sgraph) (void)hipGraphDestroy(c->sgraph);
  c->sgexec = nullptr;
  c->sgraph = nullptr;
}

static void cg_free_all(rumgpu_cg* c) {
  sweep_graph_drop(c);
  void* bufs[] = {c->masks32, c->masks64, c->offsets, c->scores, c->zpos, c->zrow, c->zoff, c->V, c->pred, c->nodes, c->R, c->rank,
                  c->binom, c->vbuf[0], c->vbuf[1], c->lbuf[0], c->lbuf[1], c->mbuf[0], c->mbuf[1], c->ebuf[0],
                  c->ebuf[1], c->lstep, c->subV, c->subpred, c->subnodes, c->submasks, c->subscores, c->suboffsets,
                  c->Vb, c->predb, c->scoresb, c->vmaxb, c->rankb, c->V64, c->r64, c->V64b, c->r64b, c->clcm, c->clF, c->cltmax,
                  c->clbmax, c->clvmax, c->clbmax64, c->clv64, c->clerr, c->oaF, c->oabm, c->oagp, c->oaerr, c->oavmax, c->oab64, c->oav64, c->oafb64, c->cbits, c->cp, c->cpout, c->cq,
                  c->cidx, c->clam, c->crank, c->cpart, c->thlist, c->tlord};
  for (void* q : bufs) rg_free(q);
  for (auto* p : c->predk) rg_free(p);
}

static int cg_alloc(rumgpu_cg* c, void** p, size_t bytes, const char* what) {
  int rc = rg_malloc(p, bytes, what);
  if (rc == RG_OK) c->device_bytes += bytes;
  return rc;
}

static void layer_order(int n, std::vector<u32>& ord, std::vector<i64>& bounds) {
  const i64 NN = 1ll << n;
  ord.resize(NN);
  bounds.assign(n + 2, 0);
  std::vector<i64> cnt(n + 2, 0);
  for (i64 v = 0; v < NN; ++v) cnt[__builtin_popcountll(v)]++;
  for (int k = 0; k <= n; ++k) bounds[k + 1] = bounds[k] + cnt[k];
  std::vector<i64> fill(bounds.begin(), bounds.end());
  for (i64 v = 0; v < NN; ++v) ord[fill[__builtin_popcountll(v)]++] = (u32)v;
}

// Tile bits of the slab layout's layered max for n alternatives: 7 (128-subset tiles, 2.6 KiB of LDS,
// one wavefront each), n when smaller. Measured on whole sweeps over L = 5..10: n20_d1pct 0.31 ms at
// L = 7 (0.41 at 6, 0.40 at 8), config3 (n = 22) 0.57 ms (0.54 at 6), n25_sparse 3.60 ms (3.68 at 6,
// 3.78 at 8); larger tiles leave the wavefront too much serial work per tile.
static int slab_tile_bits(int n) { return n < 7 ? n : 7; }

// The tile tables for L tile bits: the high subsets by popcount (thlist, thbnd) and the subsets of a
// tile by popcount (tlord). Replaces earlier tables.
static int slab_tile_tables(rumgpu_cg* c, int L) {
  std::vector<u32> hord, lord32;
  std::vector<i64> hb, lb;
  layer_order(c->n - L, hord, hb);
  layer_order(L, lord32, lb);
  std::vector<uint16_t> lord(lord32.begin(), lord32.end());
  u32* th = nullptr;
  uint16_t* tl = nullptr;
  int rc = RG_OK;
  if ((rc = rg_malloc((void**)&th, hord.size() * 4, "slab tiles")) != RG_OK) return rc;
  if ((rc = rg_malloc((void**)&tl, lord.size() * 2, "slab tile order")) != RG_OK) {
    rg_free(th);
    return rc;
  }
  if (rg_copy(c->ctx->stream, th, hord.data(), hord.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      rg_copy(c->ctx->stream, tl, lord.data(), lord.size() * 2, hipMemcpyHostToDevice) != hipSuccess) {
    rg_free(th);
    rg_free(tl);
    rg_set_error("slab tile tables: upload failed");
    return RG_EHIP;
  }
  if (c->thlist) c->device_bytes -= (((size_t)1 << (c->n - c->tL)) * 4 + ((size_t)1 << c->tL) * 2);
  rg_free(c->thlist);
  rg_free(c->tlord);
  c->thlist = th;
  c->tlord = tl;
  c->thbnd = hb;
  c->tL = L;
  c->device_bytes += hord.size() * 4 + lord.size() * 2;
  return RG_OK;
}

static u64 binom_u64(int b, int t) {
  if (t < 0 || t > b) return 0;
  u64 c = 1;
  for (int q = 0; q < t; ++q) c = c * (u64)(b - q) / (u64)(q + 1);
  return c;
}

static u64 colex_rank(u64 mask) {
  u64 r = 0;
  int t = 0, b = 0;
  while (mask) {
    if (mask & 1) {
      t += 1;
      r += binom_u64(b, t);
    }
    mask >>= 1;
    b += 1;
  }
  return r;
}

extern "C" int rumgpu_cg_create(rumgpu_ctx* ctx, int n, const uint64_t* masks_x, const uint64_t* offsets, int64_t M,
                                int layout, int dtype, int k_pred, rumgpu_cg** out) {
  RG_CHECK(ctx && masks_x && offsets && out, RG_EINVAL, "cg_create: null argument");
  RG_CHECK(layout >= 0 && layout <= 2, RG_EINVAL, "cg_create: layout must be 0 (mask), 1 (slab) or 2 (stream)");
  RG_CHECK(dtype >= 0 && dtype <= 2, RG_EINVAL, "cg_create: dtype must be 0 (f32), 1 (u16) or 2 (u24)");
  RG_CHECK(layout == 2 || dtype == 0, RG_EINVAL, "cg_create: the mask and slab layouts store float32 only");
  RG_CHECK(n >= 1 && n <= (layout == 2 ? 39 : 31), RG_EUNSUP, "cg_create: n = %d too large for this layout", n);
  RG_CHECK(M >= 1 && (i64)offsets[n] == M, RG_EINVAL, "cg_create: offsets[n] must equal M");
  RG_HIP(hipSetDevice(ctx->device));
  rumgpu_cg* c = new rumgpu_cg();
  c->ctx = ctx;
  c->n = n;
  c->layout = layout;
  c->dtype = dtype;
  c->M = M;
  c->NN = 1ll << (layout == 2 ? 0 : n);
  c->Nloc = 1ll << (n - 1);
  c->h_masks.assign(masks_x, masks_x + M);
  c->h_offsets.assign(offsets, offsets + n + 1);
  for (int x = 0; x < n && x < 32; ++x)
    if (offsets[x + 1] > offsets[x]) c->zmask |= 1u << x;
  if (layout == 0 && n >= 2 && __builtin_popcount(c->zmask) == 1) {
    c->single_a = __builtin_ctz(c->zmask);
    c->ra = (int)(offsets[c->single_a + 1] - offsets[c->single_a]);
  }
  auto fail = [&](int code) {
    cg_free_all(c);
    delete c;
    return code;
  };
  int rc = RG_OK;
  hipStream_t st = ctx->stream;
  if ((rc = cg_alloc(c, (void**)&c->offsets, (size_t)(n + 1) * 8, "offsets")) != RG_OK) return fail(rc);
  if ((rc = cg_alloc(c, (void**)&c->scores, (size_t)M * 4, "scores")) != RG_OK) return fail(rc);
  if ((rc = cg_alloc(c, (void**)&c->rank, (size_t)64 * 4, "rank")) != RG_OK) return fail(rc);
  RG_HIP(rg_copy(ctx->stream, c->offsets, offsets, (size_t)(n + 1) * 8, hipMemcpyHostToDevice));
  if (layout == 0 || layout == 1) {
    c->NN = 1ll << n;
    std::vector<u32> m32(M);
    for (i64 j = 0; j < M; ++j) m32[j] = (u32)masks_x[j];
    if ((rc = cg_alloc(c, (void**)&c->masks32, (size_t)M * 4, "menu masks")) != RG_OK) return fail(rc);
    RG_HIP(rg_copy(ctx->stream, c->masks32, m32.data(), (size_t)M * 4, hipMemcpyHostToDevice));
    if ((rc = cg_alloc(c, (void**)&c->V, (size_t)c->NN * 4, "V")) != RG_OK) return fail(rc);
    if ((rc = cg_alloc(c, (void**)&c->pred, (size_t)c->NN, "pred")) != RG_OK) return fail(rc);
    std::vector<u32> ord;
    layer_order(n, ord, c->bounds);
    if ((rc = cg_alloc(c, (void**)&c->nodes, (size_t)c->NN * 4, "layers")) != RG_OK) return fail(rc);
    RG_HIP(rg_copy(ctx->stream, c->nodes, ord.data(), (size_t)c->NN * 4, hipMemcpyHostToDevice));
    if (c->single_a >= 0) {
      const int a = c->single_a;
      // padded to a multiple of 32 with bit 31 set: a menu no subset of the other n - 1 <= 25
      // alternatives contains (the window-pass sweep reads 32 menus at a time)
      const int rpad = (c->ra + 31) / 32 * 32;
      std::vector<u32> cm(rpad, 0x80000000u);
      for (int j = 0; j < c->ra; ++j)
        cm[j] = (u32)compress_bit(masks_x[offsets[a] + j] & ~(1ull << a), a);
      if ((rc = cg_alloc(c, (void**)&c->clcm, (size_t)rpad * 4, "one-alternative menus")) != RG_OK) return fail(rc);
      RG_HIP(rg_copy(ctx->stream, c->clcm, cm.data(), (size_t)rpad * 4, hipMemcpyHostToDevice));
    }
    if (layout == 1) {
      RG_CHECK(n >= 2, RG_EUNSUP, "slab pricing needs n >= 2");
      std::vector<i64> pos(M);
      for (int x = 0; x < n; ++x)
        for (u64 j = offsets[x]; j < offsets[x + 1]; ++j)
          pos[j] = (i64)x * c->Nloc + (i64)compress_bit(masks_x[j] & ~(1ull << x), x);
      // rows by slab position, and the range of them in each tile of the first window
      std::vector<int32_t> row(M);
      for (i64 j = 0; j < M; ++j) row[j] = (int32_t)j;
      std::sort(row.begin(), row.end(), [&](int32_t a, int32_t b2) { return pos[a] < pos[b2]; });
      std::vector<i64> zp(M);
      for (i64 j = 0; j < M; ++j) zp[j] = pos[row[j]];
      WinPlan w0[kMaxWindows];
      plan_windows_into(n - 1, 4, w0);
      c->zB = w0[0].B;
      const i64 ntl = ((i64)n * c->Nloc) >> c->zB;
      std::vector<i64> zo(ntl + 1, 0);
      for (i64 j = 0; j < M; ++j) zo[(zp[j] >> c->zB) + 1]++;
      for (i64 t = 0; t < ntl; ++t) zo[t + 1] += zo[t];
      if ((rc = cg_alloc(c, (void**)&c->zpos, (size_t)M * 8, "slab positions")) != RG_OK) return fail(rc);
      if ((rc = cg_alloc(c, (void**)&c->zrow, (size_t)M * 4, "slab position rows")) != RG_OK) return fail(rc);
      if ((rc = cg_alloc(c, (void**)&c->zoff, (size_t)(ntl + 1) * 8, "slab tile ranges")) != RG_OK) return fail(rc);
      RG_HIP(rg_copy(ctx->stream, c->zpos, zp.data(), (size_t)M * 8, hipMemcpyHostToDevice));
      RG_HIP(rg_copy(ctx->stream, c->zrow, row.data(), (size_t)M * 4, hipMemcpyHostToDevice));
      RG_HIP(rg_copy(ctx->stream, c->zoff, zo.data(), (size_t)(ntl + 1) * 8, hipMemcpyHostToDevice));
      if ((rc = cg_alloc(c, (void**)&c->R, (size_t)n * c->Nloc * 4, "pricing slab")) != RG_OK) return fail(rc);
      if ((rc = slab_tile_tables(c, slab_tile_bits(n))) != RG_OK) return fail(rc);
    }
  } else {
    c->k_pred = k_pred <= 0 ? n - 10 : k_pred;
    if (c->k_pred < 1) c->k_pred = 1;
    if (c->k_pred > n) c->k_pred = n;
    if ((rc = cg_alloc(c, (void**)&c->masks64, (size_t)M * 8, "menu masks")) != RG_OK) return fail(rc);
    RG_HIP(rg_copy(ctx->stream, c->masks64, masks_x, (size_t)M * 8, hipMemcpyHostToDevice));
    c->h_binom.assign((size_t)(n + 1) * (n + 1), 0);
    for (int b = 0; b <= n; ++b)
      for (int t = 0; t <= b; ++t) c->h_binom[(size_t)b * (n + 1) + t] = binom_u64(b, t);
    if ((rc = cg_alloc(c, (void**)&c->binom, c->h_binom.size() * 8, "binomials")) != RG_OK) return fail(rc);
    RG_HIP(rg_copy(ctx->stream, c->binom, c->h_binom.data(), c->h_binom.size() * 8, hipMemcpyHostToDevice));
    u64 cap = 0;
    for (int k = 0; k <= n; ++k) cap = std::max(cap, binom_u64(n, k));
    const u64 nblk = (cap + CG_BLOCK - 1) / CG_BLOCK;
    const size_t vsz = dtype == 0 ? 4 : 2;
    for (int i = 0; i < 2; ++i) {
      if ((rc = cg_alloc(c, &c->vbuf[i], (size_t)cap * vsz, "stream layer")) != RG_OK) return fail(rc);
      if (dtype == 2 && (rc = cg_alloc(c, (void**)&c->lbuf[i], (size_t)cap, "stream layer low bytes")) != RG_OK)
        return fail(rc);
      if ((rc = cg_alloc(c, (void**)&c->mbuf[i], (size_t)nblk * 4, "block offsets")) != RG_OK) return fail(rc);
      if ((rc = cg_alloc(c, (void**)&c->ebuf[i], (size_t)nblk * 4, "block steps")) != RG_OK) return fail(rc);
    }
    if ((rc = cg_alloc(c, (void**)&c->lstep, (size_t)(n + 1) * 4, "layer steps")) != RG_OK) return fail(rc);
    c->predk.assign(n + 1, nullptr);
    for (int k = c->k_pred; k <= n; ++k)
      if ((rc = cg_alloc(c, (void**)&c->predk[k], (size_t)binom_u64(n, k), "stored predecessors")) != RG_OK)
        return fail(rc);
    // sub-lattice scaffolding
    const int K = c->k_pred;
    std::vector<u32> ord;
    layer_order(K, ord, c->subbounds);
    if ((rc = cg_alloc(c, (void**)&c->subV, (size_t)4 << K, "sub-lattice V")) != RG_OK) return fail(rc);
    if ((rc = cg_alloc(c, (void**)&c->subpred, (size_t)1 << K, "sub-lattice pred")) != RG_OK) return fail(rc);
    if ((rc = cg_alloc(c, (void**)&c->subnodes, (size_t)4 << K, "sub-lattice layers")) != RG_OK) return fail(rc);
    RG_HIP(rg_copy(ctx->stream, c->subnodes, ord.data(), (size_t)4 << K, hipMemcpyHostToDevice));
    if ((rc = cg_alloc(c, (void**)&c->submasks, (size_t)M * 4, "sub-lattice masks")) != RG_OK) return fail(rc);
    if ((rc = cg_alloc(c, (void**)&c->subscores, (size_t)M * 4, "sub-lattice scores")) != RG_OK) return fail(rc);
    if ((rc = cg_alloc(c, (void**)&c->suboffsets, (size_t)(K + 1) * 8, "sub-lattice offsets")) != RG_OK)
      return fail(rc);
  }
  RG_HIP(hipStreamSynchronize(st));
  *out = c;
  return RG_OK;
}

extern "C" int rumgpu_cg_destroy(rumgpu_cg* c) {
  if (!c) return RG_OK;
  (void)hipStreamSynchronize(c->ctx->stream);
  cg_free_all(c);
  delete c;
  return RG_OK;
}

// mask DP over a stored-layer lattice of n nodes' bits (used for the mask layout and the stream
// sub-lattice)
static int mask_sweep(hipStream_t st, int n, float* V, uint8_t* pred, const u32* masks, const float* scores,
                      const u64* offsets, const u32* nodes, const std::vector<i64>& bounds) {
  const i64 NN = 1ll << n;
  hipLaunchKernelGGL(cg_fill_f32, dim3(grid_for(NN, 256) > 4096 ? 4096 : grid_for(NN, 256)), dim3(256), 0, st, V, 0.0f, NN);
  hipLaunchKernelGGL(cg_fill_u8, dim3(grid_for(NN, 256) > 4096 ? 4096 : grid_for(NN, 256)), dim3(256), 0, st, pred,
                     (uint8_t)0, NN);
  for (int k = 1; k <= n; ++k) {
    const i64 lo = bounds[k], ns = bounds[k + 1] - lo;
    if (ns <= 0) continue;
    hipLaunchKernelGGL(cg_mask_layer, dim3(grid_for(ns, 256)), dim3(256), 0, st, V, pred, masks, scores, offsets,
                       nodes + lo, ns, n);
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

static int zeta_first(rumgpu_cg* c, hipStream_t st, const WinGeom& g, int RL) {
  RG_CHECK(g.s == 0 && g.C == 0 && g.B == c->zB, RG_EINVAL, "slab pricing: first window geometry");
  const size_t shm = (size_t)g.TPB * tile_lds_elems(g) * 4;
  const dim3 grid((unsigned)((g.ntiles + g.TPB - 1) / g.TPB)), block((unsigned)(g.TT * g.TPB));
  switch (RL) {
#define RG_Z(RV) case RV: hipLaunchKernelGGL(cg_zeta_first<RV>, grid, block, shm, st, c->R, g, c->zoff, c->zpos, c->zrow, c->scores); break;
    RG_Z(1) RG_Z(2) RG_Z(3) RG_Z(4) RG_Z(5)
#undef RG_Z
    default: RG_CHECK(false, RG_EINVAL, "slab pricing: %d register bits", RL);
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

// the tiles of the slab layout's layered max, one launch per popcount of the high part
static int slab_tiles(rumgpu_cg* c, hipStream_t st) {
  const int L = c->tL;
  for (int j = 0; j <= c->n - L; ++j) {
    const i64 lo = c->thbnd[j], ns = c->thbnd[j + 1] - lo;
    if (ns <= 0) continue;
    const dim3 g((unsigned)ns), b(CG_SLAB_TILE_THREADS);
    switch (L) {
#define RG_T(LV) case LV: hipLaunchKernelGGL(cg_slab_tile<LV>, g, b, 0, st, c->V, c->pred, c->R, c->thlist + lo, c->tlord, c->Nloc); break;
      RG_T(2) RG_T(3) RG_T(4) RG_T(5) RG_T(6) RG_T(7) RG_T(8) RG_T(9) RG_T(10)
#undef RG_T
      default: RG_CHECK(false, RG_EUNSUP, "slab pricing: tile of %d bits", L);
    }
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

static bool use_closure(const rumgpu_cg* c) { return c->layout == 0 && c->sweep == 1 && c->single_a >= 0; }

// buffers of the closure sweep, at its first use: 4 bytes per node plus per-tile maxima
static int cl_ready(rumgpu_cg* c) {
  if (c->clF) return RG_OK;
  const u32 NT = (u32)c->Nloc;
  const u32 tiles = NT >= CL_TILE ? NT / CL_TILE : 1u;
  const int tpb = tiles > 2048 ? (int)(tiles / 2048) : 1;
  const int grid = (int)((tiles + (u32)tpb - 1) / (u32)tpb);
  int rc = RG_OK;
  if ((rc = cg_alloc(c, (void**)&c->clF, (size_t)NT * 4, "closure bits")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->cltmax, (size_t)tiles * CG_BATCH * 4, "tile maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->clbmax, (size_t)grid * CG_BATCH * 4, "block maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->clvmax, (size_t)CG_BATCH * 4, "closure vmax")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->clbmax64, (size_t)grid * 8, "float64 block maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->clv64, 8, "float64 vmax")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->clerr, 4, "closure status")) != RG_OK) {
    void** bufs[] = {(void**)&c->clF, (void**)&c->cltmax, (void**)&c->clbmax, (void**)&c->clvmax,
                     (void**)&c->clbmax64, (void**)&c->clv64, (void**)&c->clerr};
    for (void** b : bufs) {
      rg_free(*b);
      *b = nullptr;
    }
    return rc;
  }
  c->clNT = NT;
  c->cl_tpb = tpb;
  c->cl_grid = grid;
  return RG_OK;
}

// the closure sweep for nd <= DB draws whose rows of single_a are sc ([row][DB]); writes
// rank (nd x n) and vmax (nd floats) on the device
template <int DB>
static int cl_sweep(rumgpu_cg* c, hipStream_t st, const float* sc, int nd, int32_t* rank, float* vmax) {
  RG_TRY(cl_ready(c));
  const int n = c->n, nb = n - 1, a = c->single_a, ra = c->ra;
  const u32 NT = c->clNT;
  const int rows_lds = std::max(1, std::min(ra, 12288 / DB));
  hipLaunchKernelGGL(cl_wmax<DB>, dim3(c->cl_grid), dim3(256), (size_t)rows_lds * DB * 4, st, c->clcm, sc, ra,
                     rows_lds, NT, c->cl_tpb, c->clF, c->cltmax, c->clbmax);
  hipLaunchKernelGGL(cl_gmax<float>, dim3(1), dim3(256), 0, st, c->clbmax, c->cl_grid, DB, vmax);
  const int w0 = std::min(nb, 12);
  hipLaunchKernelGGL(cl_close_lo, dim3(NT >> w0), dim3(256), 0, st, c->clF, c->cltmax, vmax, DB, NT, w0);
  for (int b0 = w0; b0 < nb; b0 += 8) {
    const int w = std::min(8, nb - b0);
    hipLaunchKernelGGL(cl_close_hi, dim3(1u << (nb - w - 5)), dim3(256), 0, st, c->clF, NT, b0, w);
  }
  RG_HIP(hipMemsetAsync(c->clerr, 0, 4, st));
  hipLaunchKernelGGL(cl_backtrace, dim3(1), dim3(64), 0, st, c->clF, c->clcm, sc, ra, DB, nd, n, a, vmax, rank,
                     c->clerr);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

static int cl_check(rumgpu_cg* c) {
  int err = 0;
  RG_HIP(rg_copy(c->ctx->stream, &err, c->clerr, 4, hipMemcpyDeviceToHost));
  RG_CHECK(err == 0, RG_EHIP, "cg pricing: the one-alternative path found no choice matching vmax");
  return RG_OK;
}

static bool use_window(const rumgpu_cg* c) { return c->layout == 0 && c->sweep == 2 && c->single_a >= 0; }

// one draw's scores of a's rows (x-major, on the device) as column 0 of a [ra][32] table
__global__ void oa_pad1(const float* __restrict__ src, int ra, float* __restrict__ dst) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ra * OA_DB; e += gridDim.x * blockDim.x)
    dst[e] = (e % OA_DB) == 0 ? src[e / OA_DB] : 0.0f;
}

static int oa_nblk64(const rumgpu_cg* c) {
  const i64 tiles = c->Nloc / 256;
  return (int)std::max<i64>(1, std::min<i64>(2048, tiles));
}

// buffers of the window-pass sweep, at its first use
static int oa_ready(rumgpu_cg* c) {
  if (c->oaF) return RG_OK;
  const size_t nblk = (size_t)((c->Nloc + 255) / 256);
  int rc = RG_OK;
  if ((rc = cg_alloc(c, (void**)&c->oaF, (size_t)c->Nloc * 4, "one-alternative bits")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oabm, nblk * OA_DB * 4, "one-alternative block maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oagp, (size_t)256 * OA_DB * 4, "one-alternative maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oaerr, 4, "one-alternative status")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oavmax, (size_t)OA_DB * 4, "one-alternative vmax")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oab64, (size_t)2048 * 8 * 8, "float64 block maxima")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oav64, (size_t)8 * 8, "float64 vmax")) != RG_OK ||
      (rc = cg_alloc(c, (void**)&c->oafb64, nblk * 8 * 8, "fused float64 block maxima")) != RG_OK) {
    void** bufs[] = {(void**)&c->oaF, (void**)&c->oabm, (void**)&c->oagp, (void**)&c->oaerr, (void**)&c->oavmax,
                     (void**)&c->oab64, (void**)&c->oav64, (void**)&c->oafb64};
    for (void** q : bufs) {
      rg_free(*q);
      *q = nullptr;
    }
    return rc;
  }
  return RG_OK;
}

// the window-pass sweep for nd <= 32 draws whose scores of a's rows are sc ([ra][32], unused
// columns 0): rank (nd x n) and vmax (nd floats) on the device; check oaerr afterwards. With L64 > 0
// (1, 2, 4 or 8) the same sweep also takes the float64 maxima of L64 more draws whose residuals of
// a's rows are r64 ([ra][L64], device) into oav64 (the values oa_vmax64 gives)
static int oa_sweep(rumgpu_cg* c, hipStream_t st, const float* sc, int nd, int32_t* rank, float* vmax,
                    const double* r64 = nullptr, int L64 = 0) {
  RG_TRY(oa_ready(c));
  const int n = c->n, nb = n - 1;
  const u32 NT = (u32)c->Nloc;
  const int nblk = (int)((NT + 255u) / 256u), g1 = std::min(256, nblk), lo = std::min(OA_LO, nb);
  switch (L64) {
    case 0:
      hipLaunchKernelGGL(oa_wbits<0>, dim3(nblk), dim3(256), 0, st, c->clcm, sc, nullptr, c->ra, NT, c->oaF, c->oabm,
                         nullptr);
      break;
    case 1:
      hipLaunchKernelGGL(oa_wbits<1>, dim3(nblk), dim3(256), 0, st, c->clcm, sc, r64, c->ra, NT, c->oaF, c->oabm,
                         c->oafb64);
      break;
    case 2:
      hipLaunchKernelGGL(oa_wbits<2>, dim3(nblk), dim3(256), 0, st, c->clcm, sc, r64, c->ra, NT, c->oaF, c->oabm,
                         c->oafb64);
      break;
    case 4:
      hipLaunchKernelGGL(oa_wbits<4>, dim3(nblk), dim3(256), 0, st, c->clcm, sc, r64, c->ra, NT, c->oaF, c->oabm,
                         c->oafb64);
      break;
    case 8:
      hipLaunchKernelGGL(oa_wbits<8>, dim3(nblk), dim3(256), 0, st, c->clcm, sc, r64, c->ra, NT, c->oaF, c->oabm,
                         c->oafb64);
      break;
    default: rg_set_error("cg pricing: %d float64 draws in a sweep (1, 2, 4 or 8)", L64); return RG_EINVAL;
  }
  if (L64 > 0) {                               // block maxima -> 256 partial maxima -> the maxima
    hipLaunchKernelGGL(oa_max64_part, dim3(g1), dim3(256), 0, st, c->oafb64, nblk, L64, c->oab64);
    hipLaunchKernelGGL(oa_max64_final, dim3(1), dim3(256), 0, st, c->oab64, g1, L64, c->oav64);
  }
  hipLaunchKernelGGL(oa_gmax, dim3(g1), dim3(256), 0, st, c->oabm, nblk, c->oagp);
  hipLaunchKernelGGL(oa_gmax, dim3(1), dim3(256), 0, st, c->oagp, g1, vmax);
  hipLaunchKernelGGL(oa_or_lo, dim3(NT >> lo), dim3(256), 0, st, c->oaF, c->oabm, vmax, lo);
  for (int b0 = lo; b0 < nb; b0 += OA_HI) {
    const int w = std::min(OA_HI, nb - b0);
    hipLaunchKernelGGL(oa_or_hi, dim3(1u << (nb - w - 5)), dim3(256), 0, st, c->oaF, b0, w);
  }
  hipLaunchKernelGGL(cg_fill_i32, dim3(1), dim3(64), 0, st, c->oaerr, 0, (i64)1);
  hipLaunchKernelGGL(oa_bt, dim3(nd), dim3(64), 0, st, c->oaF, c->clcm, sc, c->ra, n, c->single_a, vmax, rank,
                     c->oaerr);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

static int oa_check(rumgpu_cg* c) {
  int err = 0;
  RG_HIP(rg_copy(c->ctx->stream, &err, c->oaerr, 4, hipMemcpyDeviceToHost));
  RG_CHECK(err == 0, RG_EHIP, "cg pricing: the window-pass sweep found no choice matching vmax");
  return RG_OK;
}

// float64 max over T of W64[T] for DB draws whose residuals of a's rows are r ([ra][DB], device)
static int oa_vmax64(rumgpu_cg* c, hipStream_t st, const double* r, int DB, double* out_host, int nd) {
  RG_TRY(oa_ready(c));
  const int nblk = oa_nblk64(c);
  const u32 NT = (u32)c->Nloc;
  switch (DB) {
    case 1: hipLaunchKernelGGL(oa_wmax64<1>, dim3(nblk), dim3(256), 0, st, c->clcm, r, c->ra, NT, c->oab64); break;
    case 2: hipLaunchKernelGGL(oa_wmax64<2>, dim3(nblk), dim3(256), 0, st, c->clcm, r, c->ra, NT, c->oab64); break;
    case 4: hipLaunchKernelGGL(oa_wmax64<4>, dim3(nblk), dim3(256), 0, st, c->clcm, r, c->ra, NT, c->oab64); break;
    default: hipLaunchKernelGGL(oa_wmax64<8>, dim3(nblk), dim3(256), 0, st, c->clcm, r, c->ra, NT, c->oab64); break;
  }
  hipLaunchKernelGGL(oa_max64_final, dim3(1), dim3(256), 0, st, c->oab64, nblk, DB, c->oav64);
  RG_LAUNCH_CHECK();
  RG_HIP(rg_copy(st, out_host, c->oav64, (size_t)nd * 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

extern "C" int rumgpu_cg_set_sweep(rumgpu_cg* c, int mode) {
  RG_CHECK(c && mode >= 0 && mode <= 2, RG_EINVAL,
           "cg_set_sweep: mode must be 0 (layers), 1 (closure) or 2 (window passes)");
  RG_CHECK(mode != 2 || (c->n >= 2 && c->n <= 26 && c->ra <= OA_BT_ROWS), RG_EUNSUP,
           "cg_set_sweep: the window-pass sweep handles 2 <= n <= 26 and up to %d rows (n = %d, %d rows)",
           OA_BT_ROWS, c->n, c->ra);
  // The closure sweep's first run at n = 25 ended in a GPU fault whose cause has not been shown; it
  // passed at n <= 20, so it is refused above that until the cause is known.
  RG_CHECK(mode != 1 || c->n <= CL_MAX_N, RG_EUNSUP,
           "cg_set_sweep: the one-alternative closure sweep is not run above n = %d (n = %d)", CL_MAX_N, c->n);
  c->sweep = mode;
  return RG_OK;
}

extern "C" int rumgpu_cg_set_slab_tile(rumgpu_cg* c, int bits) {
  RG_CHECK(c, RG_EINVAL, "cg_set_slab_tile: null handle");
  RG_CHECK(c->layout == 1, RG_EINVAL, "cg_set_slab_tile: tiles belong to the slab layout");
  const int L = bits == 0 ? slab_tile_bits(c->n) : bits;
  RG_CHECK(L >= 2 && L <= 10 && L <= c->n, RG_EINVAL, "cg_set_slab_tile: %d tile bits outside 2..min(10, n = %d)", L,
           c->n);
  RG_HIP(hipSetDevice(c->ctx->device));
  RG_HIP(hipStreamSynchronize(c->ctx->stream));
  if (L == c->tL) return RG_OK;
  sweep_graph_drop(c);
  return slab_tile_tables(c, L);
}

extern "C" int rumgpu_cg_sweep_info(rumgpu_cg* c, int* one_alternative) {
  RG_CHECK(c && one_alternative, RG_EINVAL, "cg_sweep_info: null argument");
  *one_alternative = (use_closure(c) || use_window(c)) ? c->single_a : -1;
  return RG_OK;
}

// the layered sweep from the scores on the device to the ranking in c->rank (V of the full set in
// c->V[NN - 1]): mask layout, or the slab layout's butterfly and tiles
static int layered_sweep(rumgpu_cg* c, hipStream_t st) {
  const int n = c->n;
  if (c->layout == 0) {
    RG_TRY(mask_sweep(st, n, c->V, c->pred, c->masks32, c->scores, c->offsets, c->nodes, c->bounds));
  } else {
    // arc weights: the subset-sum butterfly of the scores scattered on the slab; the first window
    // pass builds its tiles from the rows inside them (no fill and scatter pass over the slab)
    WinPlan wins[kMaxWindows];
    const int nw = plan_windows_into(n - 1, 4, wins);
    RG_TRY(zeta_first(c, st, make_geom(n, wins[0]), wins[0].RL));
    for (int w = 1; w < nw; ++w)
      RG_TRY((launch_tr_rl<float, float, 0, 0>(wins[w].RL, c->R, c->R, coef<float>(1.0, 0.0, 1.0, 1.0),
                                               make_geom(n, wins[w]), nullptr, nullptr, st)));
    hipLaunchKernelGGL(cg_fill_f32, dim3(1), dim3(64), 0, st, c->V, 0.0f, (i64)1);
    RG_TRY(slab_tiles(c, st));
  }
  hipLaunchKernelGGL(cg_backtrace, dim3(1), dim3(64), 0, st, c->pred, n, c->rank, (u64)(c->NN - 1));
  RG_LAUNCH_CHECK();
  return RG_OK;
}

// 15-40 launches per sweep, most of them short at n <= 20: one graph launch instead (its kernels
// and arguments do not change between sweeps; only the scores buffer's contents do)
static int launch_layered_sweep(rumgpu_cg* c, hipStream_t st) {
  if (!c->sgexec) {
    RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = layered_sweep(c, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != RG_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    RG_CHECK(e == hipSuccess, RG_EHIP, "cg pricing: sweep capture failed: %s", hipGetErrorString(e));
    c->sgraph = g;
    const hipError_t ei = hipGraphInstantiate(&c->sgexec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      c->sgexec = nullptr;
      sweep_graph_drop(c);
      RG_CHECK(false, RG_EHIP, "cg pricing: sweep graph: %s", hipGetErrorString(ei));
    }
  }
  RG_HIP(hipGraphLaunch(c->sgexec, st));
  return RG_OK;
}

extern "C" int rumgpu_cg_price(rumgpu_cg* c, const float* scores_x, double* vmax, int32_t* rank_pos, double* margin) {
  RG_CHECK(c && vmax && rank_pos, RG_EINVAL, "cg_price: null argument");
  RG_CHECK(scores_x || c->scores_resident, RG_EINVAL,
           "cg_price: no scores (pass them, or compute them on the device with rumgpu_cg_cols_scores)");
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  const int n = c->n;
  auto t0 = std::chrono::steady_clock::now();
  if (scores_x) {
    RG_HIP(hipMemcpyAsync(c->scores, scores_x, (size_t)c->M * 4, hipMemcpyHostToDevice, st));
    c->scores_resident = false;
  }
  if (margin) *margin = 0.0;
  if (use_window(c)) {
    if (!c->scoresb) RG_TRY(cg_alloc(c, (void**)&c->scoresb, (size_t)c->M * CG_BATCH * 4, "batched scores"));
    hipLaunchKernelGGL(oa_pad1, dim3(grid_for((i64)c->ra * OA_DB, 256)), dim3(256), 0, st,
                       c->scores + c->h_offsets[c->single_a], c->ra, c->scoresb);
    RG_TRY(oa_sweep(c, st, c->scoresb, 1, c->rank, c->oavmax));
    float vf = 0.0f;
    RG_HIP(hipMemcpyAsync(&vf, c->oavmax, 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(rank_pos, c->rank, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RG_TRY(oa_check(c));
    *vmax = (double)vf;
    c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RG_OK;
  }
  if (use_closure(c)) {
    RG_TRY(cl_sweep<1>(c, st, c->scores + c->h_offsets[c->single_a], 1, c->rank, c->clvmax));
    float vf = 0.0f;
    RG_HIP(hipMemcpyAsync(&vf, c->clvmax, 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(rank_pos, c->rank, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RG_TRY(cl_check(c));
    *vmax = (double)vf;
    c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RG_OK;
  }
  if (c->layout == 0 || c->layout == 1) {
    RG_TRY(launch_layered_sweep(c, st));
    float vf = 0.0f;
    RG_HIP(hipMemcpyAsync(&vf, c->V + (c->NN - 1), 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(rank_pos, c->rank, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    *vmax = (double)vf;
    c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RG_OK;
  }
  // stream layout
  const int mode = c->dtype;
  int cur = 0, prev = 1;
  const size_t vsz = mode == 0 ? 4 : 2;
  RG_HIP(hipMemsetAsync(c->vbuf[prev], 0, vsz, st));
  if (c->lbuf[prev]) RG_HIP(hipMemsetAsync(c->lbuf[prev], 0, 1, st));
  RG_HIP(hipMemsetAsync(c->mbuf[prev], 0, 4, st));
  RG_HIP(hipMemsetAsync(c->ebuf[prev], 0, 4, st));
  RG_HIP(hipMemsetAsync(c->lstep, 0, (size_t)(n + 1) * 4, st));
  for (int k = 1; k <= n; ++k) {
    const u64 count = binom_u64(n, k);
    const u64 grid = (count + CG_BLOCK - 1) / CG_BLOCK;
    const int store = k >= c->k_pred;
    uint8_t* pr = store ? c->predk[k] : nullptr;
    switch (mode) {
      case 0:
        hipLaunchKernelGGL(cg_stream_layer<0>, dim3((unsigned)grid), dim3(CG_THREADS), 0, st, c->binom, n, k, count,
                           c->vbuf[prev], c->lbuf[prev], c->mbuf[prev], c->ebuf[prev], c->vbuf[cur], c->lbuf[cur],
                           c->mbuf[cur], c->ebuf[cur], pr, store, c->masks64, c->scores, c->offsets);
        break;
      case 1:
        hipLaunchKernelGGL(cg_stream_layer<1>, dim3((unsigned)grid), dim3(CG_THREADS), 0, st, c->binom, n, k, count,
                           c->vbuf[prev], c->lbuf[prev], c->mbuf[prev], c->ebuf[prev], c->vbuf[cur], c->lbuf[cur],
                           c->mbuf[cur], c->ebuf[cur], pr, store, c->masks64, c->scores, c->offsets);
        break;
      default:
        hipLaunchKernelGGL(cg_stream_layer<2>, dim3((unsigned)grid), dim3(CG_THREADS), 0, st, c->binom, n, k, count,
                           c->vbuf[prev], c->lbuf[prev], c->mbuf[prev], c->ebuf[prev], c->vbuf[cur], c->lbuf[cur],
                           c->mbuf[cur], c->ebuf[cur], pr, store, c->masks64, c->scores, c->offsets);
        break;
    }
    if (mode != 0) hipLaunchKernelGGL(cg_layer_step, dim3(1), dim3(256), 0, st, c->ebuf[cur], grid, c->lstep + k);
    RG_LAUNCH_CHECK();
    std::swap(cur, prev);
  }
  // top value and the per-layer steps
  float top[3] = {0, 0, 0};
  uint16_t q16 = 0;
  uint8_t q8 = 0;
  std::vector<float> steps(n + 1, 0.0f);
  if (mode == 0) {
    RG_HIP(hipMemcpyAsync(&top[0], c->vbuf[prev], 4, hipMemcpyDeviceToHost, st));
  } else {
    RG_HIP(hipMemcpyAsync(&q16, c->vbuf[prev], 2, hipMemcpyDeviceToHost, st));
    if (mode == 2) RG_HIP(hipMemcpyAsync(&q8, c->lbuf[prev], 1, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(&top[1], c->mbuf[prev], 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(&top[2], c->ebuf[prev], 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(steps.data(), c->lstep, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost, st));
  }
  RG_HIP(hipStreamSynchronize(st));
  double mg = 0.0;
  if (mode == 0) {
    *vmax = (double)top[0];
  } else if (mode == 1) {
    *vmax = (double)top[1] + (double)q16 * (double)top[2];
    for (int k = 1; k <= n; ++k) mg += (double)steps[k];
    mg *= 0.5;
  } else {
    const unsigned q = ((unsigned)q16 << 8) | (unsigned)q8;
    *vmax = (double)top[1] + (double)q * (double)top[2];
    for (int k = 1; k <= n; ++k) mg += (double)steps[k];
  }
  if (margin) *margin = mg;
  // backtrace: stored predecessors down to layer k_pred, then the exact sub-lattice, whose rows
  // are gathered on the host (from the device copy when the scores are resident)
  std::vector<float> resident;
  const float* sx = scores_x;
  if (!sx) {
    resident.resize((size_t)c->M);
    RG_HIP(rg_copy(st, resident.data(), c->scores, (size_t)c->M * 4, hipMemcpyDeviceToHost));
    sx = resident.data();
  }
  u64 D = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  int pos = 0;
  for (int k = n; k > c->k_pred; --k) {
    const u64 r = colex_rank(D);
    uint8_t x = 0;
    RG_HIP(rg_copy(c->ctx->stream, &x, c->predk[k] + r, 1, hipMemcpyDeviceToHost));
    rank_pos[x] = pos++;
    D ^= 1ull << x;
  }
  const int K = c->k_pred;
  std::vector<int> elems;
  for (int b = 0; b < n; ++b)
    if ((D >> b) & 1ull) elems.push_back(b);
  RG_CHECK((int)elems.size() == K, RG_EHIP, "stream backtrace reached a node of the wrong size");
  std::vector<int> posof(n, -1);
  for (int j = 0; j < K; ++j) posof[elems[j]] = j;
  // rows whose menu lies inside D, relabelled to 0..K-1, x-major (stable by new label)
  std::vector<std::vector<std::pair<u32, float>>> rows(K);
  for (int x = 0; x < n; ++x) {
    if (posof[x] < 0) continue;
    for (u64 j = c->h_offsets[x]; j < c->h_offsets[x + 1]; ++j) {
      const u64 m = c->h_masks[j];
      if ((m & D) != m) continue;
      u32 sm = 0;
      for (int e = 0; e < K; ++e)
        if ((m >> elems[e]) & 1ull) sm |= 1u << e;
      rows[posof[x]].push_back({sm, sx[j]});
    }
  }
  std::vector<u32> sm;
  std::vector<float> ss;
  std::vector<u64> so(K + 1, 0);
  for (int x = 0; x < K; ++x) {
    for (auto& pr : rows[x]) {
      sm.push_back(pr.first);
      ss.push_back(pr.second);
    }
    so[x + 1] = sm.size();
  }
  if (!sm.empty()) {
    RG_HIP(hipMemcpyAsync(c->submasks, sm.data(), sm.size() * 4, hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(c->subscores, ss.data(), ss.size() * 4, hipMemcpyHostToDevice, st));
  }
  RG_HIP(hipMemcpyAsync(c->suboffsets, so.data(), (size_t)(K + 1) * 8, hipMemcpyHostToDevice, st));
  RG_TRY(mask_sweep(st, K, c->subV, c->subpred, c->submasks, c->subscores, c->suboffsets, c->subnodes, c->subbounds));
  hipLaunchKernelGGL(cg_backtrace, dim3(1), dim3(64), 0, st, c->subpred, K, c->rank, (u64)((1ull << K) - 1ull));
  int32_t order[64];
  RG_HIP(hipMemcpyAsync(order, c->rank, (size_t)K * 4, hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  for (int j = 0; j < K; ++j) rank_pos[elems[j]] = pos + order[j];
  c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RG_OK;
}

template <int DB>
static int mask_sweep_batch(rumgpu_cg* c, hipStream_t st) {
  const int n = c->n;
  hipLaunchKernelGGL(cg_fill_f32, dim3(1), dim3(64), 0, st, c->Vb, 0.0f, (i64)DB);     // V[empty set] = 0
  for (int k = 1; k <= n; ++k) {
    const i64 lo = c->bounds[k], ns = c->bounds[k + 1] - lo;
    if (ns <= 0) continue;
    hipLaunchKernelGGL(cg_mask_layer_batch<DB>, dim3(grid_for(ns, 256)), dim3(256), 0, st, c->Vb, c->predb,
                       c->masks32, c->scoresb, c->offsets, c->nodes + lo, ns, n, c->zmask);
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

extern "C" int rumgpu_cg_price_batch(rumgpu_cg* c, int D, const float* scores, double* vmax, int32_t* rank_pos) {
  RG_CHECK(c && scores && vmax && rank_pos && D >= 1, RG_EINVAL, "cg_price_batch: bad argument");
  RG_CHECK(c->layout == 0, RG_EUNSUP, "cg_price_batch: batched sweeps use the mask layout");
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  const int n = c->n;
  const i64 M = c->M;
  auto t0 = std::chrono::steady_clock::now();
  const bool closure = use_closure(c), window = use_window(c);
  if (!c->scoresb) {
    int rc = RG_OK;
    if ((rc = cg_alloc(c, (void**)&c->scoresb, (size_t)M * CG_BATCH * 4, "batched scores")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->vmaxb, (size_t)CG_BATCH * 4, "batched vmax")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->rankb, (size_t)CG_BATCH * 64 * 4, "batched ranks")) != RG_OK) {
      rg_free(c->scoresb), rg_free(c->vmaxb), rg_free(c->rankb);
      c->scoresb = nullptr, c->vmaxb = nullptr, c->rankb = nullptr;
      return rc;
    }
  }
  if (!closure && !window && !c->Vb) {
    int rc = RG_OK;
    if ((rc = cg_alloc(c, (void**)&c->Vb, (size_t)c->NN * CG_BATCH * 4, "batched V")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->predb, (size_t)c->NN * CG_BATCH, "batched pred")) != RG_OK) {
      rg_free(c->Vb), rg_free(c->predb);
      c->Vb = nullptr, c->predb = nullptr;
      return rc;
    }
  }
  std::vector<float> tr((size_t)M * CG_BATCH), vf(CG_BATCH);
  for (int first = 0; first < D; first += CG_BATCH) {
    const int nd = std::min(CG_BATCH, D - first);
    const int DB = window ? OA_DB : nd <= 1 ? 1 : nd <= 2 ? 2 : nd <= 4 ? 4 : nd <= 8 ? 8 : nd <= 16 ? 16 : 32;
    for (i64 j = 0; j < M; ++j)
      for (int d = 0; d < DB; ++d) tr[(size_t)j * DB + d] = d < nd ? scores[(size_t)(first + d) * M + j] : 0.0f;
    RG_HIP(hipMemcpyAsync(c->scoresb, tr.data(), (size_t)M * DB * 4, hipMemcpyHostToDevice, st));
    if (window) {
      RG_TRY(oa_sweep(c, st, c->scoresb + (size_t)c->h_offsets[c->single_a] * OA_DB, nd, c->rankb, c->vmaxb));
      RG_HIP(hipMemcpyAsync(vf.data(), c->vmaxb, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
      RG_HIP(hipMemcpyAsync(rank_pos + (size_t)first * n, c->rankb, (size_t)nd * n * 4, hipMemcpyDeviceToHost, st));
      RG_TRY(oa_check(c));
      for (int d = 0; d < nd; ++d) vmax[first + d] = (double)vf[d];
      continue;
    }
    if (closure) {
      const float* sa = c->scoresb + (size_t)c->h_offsets[c->single_a] * DB;
      switch (DB) {
        case 1: RG_TRY(cl_sweep<1>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
        case 2: RG_TRY(cl_sweep<2>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
        case 4: RG_TRY(cl_sweep<4>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
        case 8: RG_TRY(cl_sweep<8>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
        case 16: RG_TRY(cl_sweep<16>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
        default: RG_TRY(cl_sweep<32>(c, st, sa, nd, c->rankb, c->vmaxb)); break;
      }
      RG_HIP(hipMemcpyAsync(vf.data(), c->vmaxb, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
      RG_HIP(hipMemcpyAsync(rank_pos + (size_t)first * n, c->rankb, (size_t)nd * n * 4, hipMemcpyDeviceToHost, st));
      RG_TRY(cl_check(c));
      for (int d = 0; d < nd; ++d) vmax[first + d] = (double)vf[d];
      continue;
    }
    switch (DB) {
      case 1: RG_TRY(mask_sweep_batch<1>(c, st)); break;
      case 2: RG_TRY(mask_sweep_batch<2>(c, st)); break;
      case 4: RG_TRY(mask_sweep_batch<4>(c, st)); break;
      case 8: RG_TRY(mask_sweep_batch<8>(c, st)); break;
      case 16: RG_TRY(mask_sweep_batch<16>(c, st)); break;
      default: RG_TRY(mask_sweep_batch<32>(c, st)); break;
    }
    hipLaunchKernelGGL(cg_backtrace_batch, dim3(1), dim3(64), 0, st, c->predb, c->Vb, n, DB, nd, c->rankb, c->vmaxb,
                       (u64)(c->NN - 1), c->zmask);
    RG_LAUNCH_CHECK();
    RG_HIP(hipMemcpyAsync(vf.data(), c->vmaxb, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipMemcpyAsync(rank_pos + (size_t)first * n, c->rankb, (size_t)nd * n * 4, hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < nd; ++d) vmax[first + d] = (double)vf[d];
  }
  c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RG_OK;
}

static int stream_vmax64(rumgpu_cg* c, const double* r_x, double* vmax) {
  hipStream_t st = c->ctx->stream;
  const int n = c->n;
  if (!c->V64) {
    u64 cap = 0;
    for (int k = 0; k <= n; ++k) cap = std::max(cap, binom_u64(n, k));
    int rc = RG_OK;
    if ((rc = cg_alloc(c, (void**)&c->V64, (size_t)cap * 8 * 2, "float64 stream layers")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->r64, (size_t)c->M * 8, "float64 residual")) != RG_OK) {
      rg_free(c->V64), rg_free(c->r64);
      c->V64 = nullptr, c->r64 = nullptr;
      return rc;
    }
    c->cap64 = cap;
  }
  RG_HIP(hipMemcpyAsync(c->r64, r_x, (size_t)c->M * 8, hipMemcpyHostToDevice, st));
  double* lay[2] = {c->V64, c->V64 + c->cap64};
  int cur = 0, prev = 1;
  hipLaunchKernelGGL(cg_fill_f64, dim3(1), dim3(64), 0, st, lay[prev], 0.0, (i64)1);     // layer 0: V = 0
  for (int k = 1; k <= n; ++k) {
    const u64 count = binom_u64(n, k);
    hipLaunchKernelGGL(cg_stream_layer64, dim3((unsigned)((count + CG_THREADS - 1) / CG_THREADS)), dim3(CG_THREADS),
                       0, st, c->binom, n, k, count, lay[prev], lay[cur], c->masks64, c->r64, c->offsets);
    std::swap(cur, prev);
  }
  RG_LAUNCH_CHECK();
  RG_HIP(rg_copy(st, vmax, lay[prev], 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

extern "C" int rumgpu_cg_vmax64(rumgpu_cg* c, const double* r_x, double* vmax) {
  RG_CHECK(c && r_x && vmax, RG_EINVAL, "cg_vmax64: null argument");
  RG_CHECK(c->layout == 0 || c->layout == 2, RG_EUNSUP, "cg_vmax64: the float64 sweep uses the mask or stream layout");
  RG_HIP(hipSetDevice(c->ctx->device));
  if (c->layout == 2) return stream_vmax64(c, r_x, vmax);
  hipStream_t st = c->ctx->stream;
  if (use_window(c)) {
    if (!c->r64) RG_TRY(cg_alloc(c, (void**)&c->r64, (size_t)c->M * 8, "float64 residual"));
    RG_HIP(hipMemcpyAsync(c->r64, r_x, (size_t)c->M * 8, hipMemcpyHostToDevice, st));
    return oa_vmax64(c, st, c->r64 + c->h_offsets[c->single_a], 1, vmax, 1);
  }
  if (use_closure(c)) {
    RG_TRY(cl_ready(c));
    if (!c->r64) RG_TRY(cg_alloc(c, (void**)&c->r64, (size_t)c->M * 8, "float64 residual"));
    RG_HIP(hipMemcpyAsync(c->r64, r_x, (size_t)c->M * 8, hipMemcpyHostToDevice, st));
    const int rows_lds = std::max(1, std::min(c->ra, 6144));
    hipLaunchKernelGGL(cl_wmax64, dim3(c->cl_grid), dim3(256), (size_t)rows_lds * 8, st, c->clcm,
                       c->r64 + c->h_offsets[c->single_a], c->ra, rows_lds, c->clNT, c->cl_tpb, c->clbmax64);
    hipLaunchKernelGGL(cl_gmax<double>, dim3(1), dim3(256), 0, st, c->clbmax64, c->cl_grid, 1, c->clv64);
    RG_LAUNCH_CHECK();
    RG_HIP(rg_copy(st, vmax, c->clv64, 8, hipMemcpyDeviceToHost));
    return RG_OK;
  }
  if (!c->V64) RG_TRY(cg_alloc(c, (void**)&c->V64, (size_t)c->NN * 8, "float64 V"));
  if (!c->r64) RG_TRY(cg_alloc(c, (void**)&c->r64, (size_t)c->M * 8, "float64 residual"));
  RG_HIP(hipMemcpyAsync(c->r64, r_x, (size_t)c->M * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(cg_fill_f64, dim3(1), dim3(64), 0, st, c->V64, 0.0, (i64)1);
  for (int k = 1; k <= c->n; ++k) {
    const i64 lo = c->bounds[k], ns = c->bounds[k + 1] - lo;
    if (ns <= 0) continue;
    hipLaunchKernelGGL(cg_mask_layer64, dim3(grid_for(ns, 256)), dim3(256), 0, st, c->V64, c->masks32, c->r64,
                       c->offsets, c->nodes + lo, ns, c->n, c->zmask);
  }
  RG_LAUNCH_CHECK();
  if (((u64)(c->NN - 1) & c->zmask) == 0) {       // no rows at all: every node is 0
    *vmax = 0.0;
    return RG_OK;
  }
  RG_HIP(rg_copy(st, vmax, c->V64 + (c->NN - 1), 8, hipMemcpyDeviceToHost));
  return RG_OK;
}

#define CG_LB_BATCH 8

template <int DB>
static int mask_sweep64_batch(rumgpu_cg* c, hipStream_t st) {
  hipLaunchKernelGGL(cg_fill_f64, dim3(1), dim3(64), 0, st, c->V64b, 0.0, (i64)DB);     // V[empty set] = 0
  for (int k = 1; k <= c->n; ++k) {
    const i64 lo = c->bounds[k], ns = c->bounds[k + 1] - lo;
    if (ns <= 0) continue;
    hipLaunchKernelGGL(cg_mask_layer64_batch<DB>, dim3(grid_for(ns, 256)), dim3(256), 0, st, c->V64b, c->masks32,
                       c->r64b, c->offsets, c->nodes + lo, ns, c->n, c->zmask);
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

extern "C" int rumgpu_cg_vmax64_batch(rumgpu_cg* c, int D, const double* r, double* vmax) {
  RG_CHECK(c && r && vmax && D >= 1, RG_EINVAL, "cg_vmax64_batch: bad argument");
  RG_CHECK(c->layout == 0, RG_EUNSUP, "cg_vmax64_batch: the batched float64 sweep uses the mask layout");
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  const i64 M = c->M;
  const bool window = use_window(c);
  if (window && !c->r64b)
    RG_TRY(cg_alloc(c, (void**)&c->r64b, (size_t)M * CG_LB_BATCH * 8, "batched float64 residuals"));
  if (!window && !c->V64b) {
    int rc = RG_OK;
    if ((rc = cg_alloc(c, (void**)&c->V64b, (size_t)c->NN * CG_LB_BATCH * 8, "batched float64 V")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->r64b, (size_t)M * CG_LB_BATCH * 8, "batched float64 residuals")) != RG_OK) {
      rg_free(c->V64b), rg_free(c->r64b);
      c->V64b = nullptr, c->r64b = nullptr;
      return rc;
    }
  }
  const u64 full = (u64)(c->NN - 1);
  std::vector<double> tr((size_t)M * CG_LB_BATCH), out(CG_LB_BATCH);
  for (int first = 0; first < D; first += CG_LB_BATCH) {
    const int nd = std::min(CG_LB_BATCH, D - first);
    const int DB = nd <= 1 ? 1 : nd <= 2 ? 2 : nd <= 4 ? 4 : 8;
    for (i64 j = 0; j < M; ++j)
      for (int d = 0; d < DB; ++d) tr[(size_t)j * DB + d] = d < nd ? r[(size_t)(first + d) * M + j] : 0.0;
    RG_HIP(hipMemcpyAsync(c->r64b, tr.data(), (size_t)M * DB * 8, hipMemcpyHostToDevice, st));
    if (window) {
      RG_TRY(oa_vmax64(c, st, c->r64b + (size_t)c->h_offsets[c->single_a] * DB, DB, out.data(), nd));
      for (int d = 0; d < nd; ++d) vmax[first + d] = out[d];
      continue;
    }
    switch (DB) {
      case 1: RG_TRY(mask_sweep64_batch<1>(c, st)); break;
      case 2: RG_TRY(mask_sweep64_batch<2>(c, st)); break;
      case 4: RG_TRY(mask_sweep64_batch<4>(c, st)); break;
      default: RG_TRY(mask_sweep64_batch<8>(c, st)); break;
    }
    if ((full & c->zmask) == 0) {              // no rows at all: every node is 0
      for (int d = 0; d < nd; ++d) vmax[first + d] = 0.0;
      continue;
    }
    RG_HIP(rg_copy(st, out.data(), c->V64b + full * DB, (size_t)nd * 8, hipMemcpyDeviceToHost));
    for (int d = 0; d < nd; ++d) vmax[first + d] = out[d];
  }
  return RG_OK;
}

// rumgpu_cg_price_batch for D draws and rumgpu_cg_vmax64_batch for L more in one call; on the
// one-alternative window-pass sweep (D <= 32, L <= 8) the float64 maxima ride in the pricing sweep,
// which tests every subset once for both; otherwise the two calls one after the other. The values
// are those of the two calls either way.
extern "C" int rumgpu_cg_price_batch_lb(rumgpu_cg* c, int D, const float* scores, double* vmax, int32_t* rank_pos,
                                        int L, const double* r, double* vmax64) {
  RG_CHECK(c && scores && vmax && rank_pos && D >= 1 && L >= 0 && (L == 0 || (r && vmax64)), RG_EINVAL,
           "cg_price_batch_lb: bad argument");
  if (L == 0) return rumgpu_cg_price_batch(c, D, scores, vmax, rank_pos);
  if (c->layout != 0 || !use_window(c) || D > CG_BATCH || L > 8) {
    RG_TRY(rumgpu_cg_price_batch(c, D, scores, vmax, rank_pos));
    return rumgpu_cg_vmax64_batch(c, L, r, vmax64);
  }
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  const int n = c->n;
  const i64 M = c->M;
  auto t0 = std::chrono::steady_clock::now();
  if (!c->scoresb) {
    int rc = RG_OK;
    if ((rc = cg_alloc(c, (void**)&c->scoresb, (size_t)M * CG_BATCH * 4, "batched scores")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->vmaxb, (size_t)CG_BATCH * 4, "batched vmax")) != RG_OK ||
        (rc = cg_alloc(c, (void**)&c->rankb, (size_t)CG_BATCH * 64 * 4, "batched ranks")) != RG_OK) {
      rg_free(c->scoresb), rg_free(c->vmaxb), rg_free(c->rankb);
      c->scoresb = nullptr, c->vmaxb = nullptr, c->rankb = nullptr;
      return rc;
    }
  }
  if (!c->r64b) RG_TRY(cg_alloc(c, (void**)&c->r64b, (size_t)M * CG_LB_BATCH * 8, "batched float64 residuals"));
  const int DL = L <= 1 ? 1 : L <= 2 ? 2 : L <= 4 ? 4 : 8;
  std::vector<float> tr((size_t)M * OA_DB), vf(OA_DB);
  for (i64 j = 0; j < M; ++j)
    for (int d = 0; d < OA_DB; ++d) tr[(size_t)j * OA_DB + d] = d < D ? scores[(size_t)d * M + j] : 0.0f;
  std::vector<double> tr64((size_t)M * DL), out64(8);
  for (i64 j = 0; j < M; ++j)
    for (int d = 0; d < DL; ++d) tr64[(size_t)j * DL + d] = d < L ? r[(size_t)d * M + j] : 0.0;
  RG_HIP(hipMemcpyAsync(c->scoresb, tr.data(), (size_t)M * OA_DB * 4, hipMemcpyHostToDevice, st));
  RG_HIP(hipMemcpyAsync(c->r64b, tr64.data(), (size_t)M * DL * 8, hipMemcpyHostToDevice, st));
  const size_t off = (size_t)c->h_offsets[c->single_a];
  RG_TRY(oa_sweep(c, st, c->scoresb + off * OA_DB, D, c->rankb, c->vmaxb, c->r64b + off * DL, DL));
  RG_HIP(hipMemcpyAsync(vf.data(), c->vmaxb, (size_t)D * 4, hipMemcpyDeviceToHost, st));
  RG_HIP(hipMemcpyAsync(rank_pos, c->rankb, (size_t)D * n * 4, hipMemcpyDeviceToHost, st));
  RG_HIP(hipMemcpyAsync(out64.data(), c->oav64, (size_t)L * 8, hipMemcpyDeviceToHost, st));
  RG_TRY(oa_check(c));                         // synchronizes the stream
  for (int d = 0; d < D; ++d) vmax[d] = (double)vf[d];
  for (int d = 0; d < L; ++d) vmax64[d] = out64[d];
  c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RG_OK;
}

extern "C" int rumgpu_cg_info(rumgpu_cg* c, size_t* device_bytes, double* last_kernel_ms) {
  RG_CHECK(c, RG_EINVAL, "cg_info: null handle");
  if (device_bytes) *device_bytes = c->device_bytes;
  if (last_kernel_ms) *last_kernel_ms = c->last_ms;
  return RG_OK;
}

// ---------------------------------------------------------------------------------------------
// Column store of one column-generation run. The ranking columns are bitsets over the x-major rows
// (bit r % 64 of word r / 64), built on the device from the ranking the sweep returned; the new
// column's Gram row and the residual scores float(p - V lam) of the next sweep are computed where
// the columns live, so a run's per-column host work is the master QP on the K x K Gram matrix.

// one word per wavefront: lane l of the wave that starts at row 64 w decides row 64 w + l; each
// block also writes the sum of p over its rows that are in the column (a fixed reduction order)
template <typename MT>
__global__ __launch_bounds__(256) void cg_col_bits(const MT* __restrict__ masks, const u64* __restrict__ offsets,
                                                   int n, i64 M, i64 W, const int32_t* __restrict__ rank_pos,
                                                   const double* __restrict__ p, u64* __restrict__ out,
                                                   double* __restrict__ part) {
  __shared__ int rk[64];
  __shared__ u64 off[65];
  if (threadIdx.x < (unsigned)n) rk[threadIdx.x] = rank_pos[threadIdx.x];
  if (threadIdx.x <= (unsigned)n) off[threadIdx.x] = offsets[threadIdx.x];
  __syncthreads();
  const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
  bool on = false;
  if (r < M) {
    int x = 0;                                   // the row's alternative: offsets[x] <= r < offsets[x + 1]
    while (off[x + 1] <= (u64)r) ++x;
    u64 S = (u64)masks[r];
    int best = 1 << 30, top = -1;                // the menu's top member under the ranking
    while (S) {
      const int b = __builtin_ctzll(S);
      S &= S - 1;
      if (rk[b] < best) {
        best = rk[b];
        top = b;
      }
    }
    on = top == x;
  }
  const u64 word = __ballot(on);
  if ((threadIdx.x & 63) == 0 && (r >> 6) < W) out[r >> 6] = word;
  __shared__ double wsum[4];
  double v = on ? p[r] : 0.0;                  // on implies r < M
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// q[j] = |col_j AND col_k| for j = 0..k (block j)
__global__ __launch_bounds__(256) void cg_col_gram(const u64* __restrict__ bits, i64 W, i64 k,
                                                   unsigned long long* __restrict__ q) {
  __shared__ unsigned long long part[4];
  const u64* a = bits + (i64)blockIdx.x * W;
  const u64* b = bits + k * W;
  unsigned long long s = 0;
  for (i64 w = threadIdx.x; w < W; w += 256) s += (unsigned long long)__popcll(a[w] & b[w]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) q[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// p_out[r] = sum_t lam[t] [r in col idx[t]] (in t order), scores[r] = float(p[r] - p_out[r]).
// A block owns 4 words (256 rows). Per chunk of 64 support columns its threads load the 4 words of
// each column at once (thread i: column i / 4, word i % 4) into LDS; then every lane adds, in t
// order, the weights of the columns whose bit for its row is set (broadcast LDS reads).
__global__ __launch_bounds__(256) void cg_col_scores(const u64* __restrict__ bits, i64 W, i64 M, int nA,
                                                     const int32_t* __restrict__ idx,
                                                     const double* __restrict__ lam, const double* __restrict__ p,
                                                     double* __restrict__ pout, float* __restrict__ scores) {
  __shared__ u64 wl[64][4];
  __shared__ double sl[64];
  const i64 w0 = (i64)blockIdx.x * 4;
  const int tid = threadIdx.x;
  const int myw = tid >> 6, lane = tid & 63;
  const i64 r = w0 * 64 + tid;
  double acc = 0.0;
  for (int a0 = 0; a0 < nA; a0 += 64) {
    const int cnt = nA - a0 < 64 ? nA - a0 : 64;
    __syncthreads();
    {
      const int tc = tid >> 2, wq = tid & 3;
      u64 word = 0;
      if (tc < cnt && w0 + wq < W) word = bits[(i64)idx[a0 + tc] * W + w0 + wq];
      wl[tc][wq] = word;
      if (tid < cnt) sl[tid] = lam[a0 + tid];
    }
    __syncthreads();
    for (int t = 0; t < cnt; ++t) acc += ((wl[t][myw] >> lane) & 1ull) ? sl[t] : 0.0;
  }
  if (r < M) {
    pout[r] = acc;
    scores[r] = (float)(p[r] - acc);
  }
}

static int cols_grow(rumgpu_cg* c, i64 need) {
  if (need <= c->ccap) return RG_OK;
  i64 cap = c->ccap ? 2 * c->ccap : 64;
  if (cap < need) cap = need;
  u64* nb = nullptr;
  unsigned long long* nq = nullptr;
  int32_t* ni = nullptr;
  double* nl = nullptr;
  int rc = RG_OK;
  if ((rc = rg_malloc((void**)&nb, (size_t)cap * c->cW * 8, "column bitsets")) != RG_OK) return rc;
  if ((rc = rg_malloc((void**)&nq, (size_t)cap * 8, "Gram row")) != RG_OK ||
      (rc = rg_malloc((void**)&ni, (size_t)cap * 4, "lambda support")) != RG_OK ||
      (rc = rg_malloc((void**)&nl, (size_t)cap * 8, "lambda")) != RG_OK) {
    rg_free(nb);
    rg_free(nq);
    rg_free(ni);
    return rc;
  }
  if (c->cK > 0)
    RG_HIP(hipMemcpyAsync(nb, c->cbits, (size_t)c->cK * c->cW * 8, hipMemcpyDeviceToDevice, c->ctx->stream));
  RG_HIP(hipStreamSynchronize(c->ctx->stream));
  rg_free(c->cbits);
  rg_free(c->cq);
  rg_free(c->cidx);
  rg_free(c->clam);
  c->device_bytes += (size_t)(cap - c->ccap) * ((size_t)c->cW * 8 + 20);
  c->cbits = nb;
  c->cq = nq;
  c->cidx = ni;
  c->clam = nl;
  c->ccap = cap;
  return RG_OK;
}

extern "C" int rumgpu_cg_cols_init(rumgpu_cg* c, const double* p_x, int64_t cap) {
  RG_CHECK(c && p_x, RG_EINVAL, "cg_cols_init: null argument");
  RG_CHECK(cap >= 1, RG_EINVAL, "cg_cols_init: cap must be >= 1");
  RG_CHECK(c->masks32 || c->masks64, RG_EINVAL, "cg_cols_init: pricer has no rows");
  RG_HIP(hipSetDevice(c->ctx->device));
  int rc = RG_OK;
  if (!c->cp) {
    c->cW = (c->M + 63) / 64;
    if ((rc = cg_alloc(c, (void**)&c->cp, (size_t)c->M * 8, "column store p")) != RG_OK) return rc;
    if ((rc = cg_alloc(c, (void**)&c->cpout, (size_t)c->M * 8, "column store p_out")) != RG_OK) return rc;
    if ((rc = cg_alloc(c, (void**)&c->crank, 64 * 4, "column ranking")) != RG_OK) return rc;
    if ((rc = cg_alloc(c, (void**)&c->cpart, (size_t)((c->cW * 64 + 255) / 256) * 8, "column sums")) != RG_OK)
      return rc;
  }
  c->cK = 0;
  c->scores_resident = false;
  RG_TRY(cols_grow(c, cap));
  RG_HIP(rg_copy(c->ctx->stream, c->cp, p_x, (size_t)c->M * 8, hipMemcpyHostToDevice));
  return RG_OK;
}

extern "C" int rumgpu_cg_cols_load(rumgpu_cg* c, int64_t k, const uint64_t* bits) {
  RG_CHECK(c && (bits || k == 0), RG_EINVAL, "cg_cols_load: null argument");
  RG_CHECK(c->cp, RG_EINVAL, "cg_cols_load: call rumgpu_cg_cols_init first");
  RG_CHECK(k >= 0, RG_EINVAL, "cg_cols_load: k must be >= 0");
  RG_HIP(hipSetDevice(c->ctx->device));
  c->cK = 0;
  RG_TRY(cols_grow(c, k));
  if (k) RG_HIP(rg_copy(c->ctx->stream, c->cbits, bits, (size_t)k * c->cW * 8, hipMemcpyHostToDevice));
  c->cK = k;
  c->scores_resident = false;
  return RG_OK;
}

extern "C" int rumgpu_cg_cols_add(rumgpu_cg* c, const int32_t* rank_pos, double* gram, double* h, uint64_t* bits) {
  RG_CHECK(c && rank_pos && gram && h, RG_EINVAL, "cg_cols_add: null argument");
  RG_CHECK(c->cp, RG_EINVAL, "cg_cols_add: call rumgpu_cg_cols_init first");
  const int n = c->n;
  std::vector<char> seen(n, 0);
  for (int b = 0; b < n; ++b) {
    RG_CHECK(rank_pos[b] >= 0 && rank_pos[b] < n && !seen[rank_pos[b]], RG_EINVAL,
             "cg_cols_add: rank_pos must be a permutation of 0..n-1");
    seen[rank_pos[b]] = 1;
  }
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  RG_TRY(cols_grow(c, c->cK + 1));
  const i64 k = c->cK;
  u64* col = c->cbits + k * c->cW;
  RG_HIP(hipMemcpyAsync(c->crank, rank_pos, (size_t)n * 4, hipMemcpyHostToDevice, st));
  const i64 blocks = (c->cW * 64 + 255) / 256;
  if (c->masks32)
    hipLaunchKernelGGL(cg_col_bits<u32>, dim3((unsigned)blocks), dim3(256), 0, st, c->masks32, c->offsets, n, c->M,
                       c->cW, c->crank, c->cp, col, c->cpart);
  else
    hipLaunchKernelGGL(cg_col_bits<u64>, dim3((unsigned)blocks), dim3(256), 0, st, c->masks64, c->offsets, n, c->M,
                       c->cW, c->crank, c->cp, col, c->cpart);
  hipLaunchKernelGGL(cg_col_gram, dim3((unsigned)(k + 1)), dim3(256), 0, st, c->cbits, c->cW, k, c->cq);
  RG_LAUNCH_CHECK();
  std::vector<unsigned long long> q((size_t)k + 1);
  RG_HIP(hipMemcpyAsync(q.data(), c->cq, (size_t)(k + 1) * 8, hipMemcpyDeviceToHost, st));
  if (bits) RG_HIP(hipMemcpyAsync(bits, col, (size_t)c->cW * 8, hipMemcpyDeviceToHost, st));
  std::vector<double> part((size_t)blocks);
  RG_HIP(hipMemcpyAsync(part.data(), c->cpart, (size_t)blocks * 8, hipMemcpyDeviceToHost, st));
  RG_HIP(hipStreamSynchronize(st));
  for (i64 j = 0; j <= k; ++j) gram[j] = (double)q[j];
  double hs = 0.0;
  for (i64 b = 0; b < blocks; ++b) hs += part[(size_t)b];
  *h = hs;
  c->cK = k + 1;
  return RG_OK;
}

extern "C" int rumgpu_cg_cols_scores(rumgpu_cg* c, int32_t nA, const int32_t* idx, const double* lam) {
  RG_CHECK(c && (nA == 0 || (idx && lam)), RG_EINVAL, "cg_cols_scores: null argument");
  RG_CHECK(c->cp, RG_EINVAL, "cg_cols_scores: call rumgpu_cg_cols_init first");
  RG_CHECK(nA >= 0 && nA <= c->cK, RG_EINVAL, "cg_cols_scores: nA = %d outside 0..%lld", nA, (long long)c->cK);
  for (int t = 0; t < nA; ++t)
    RG_CHECK(idx[t] >= 0 && idx[t] < c->cK, RG_EINVAL, "cg_cols_scores: column %d outside 0..%lld", idx[t],
             (long long)c->cK - 1);
  RG_HIP(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  if (nA) {
    RG_HIP(hipMemcpyAsync(c->cidx, idx, (size_t)nA * 4, hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(c->clam, lam, (size_t)nA * 8, hipMemcpyHostToDevice, st));
  }
  hipLaunchKernelGGL(cg_col_scores, dim3((unsigned)((c->cW + 3) / 4)), dim3(256), 0, st, c->cbits, c->cW, c->M,
                     (int)nA, c->cidx, c->clam, c->cp, c->cpout, c->scores);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  c->scores_resident = true;
  return RG_OK;
}

extern "C" int rumgpu_cg_cols_pout(rumgpu_cg* c, double* p_out) {
  RG_CHECK(c && p_out, RG_EINVAL, "cg_cols_pout: null argument");
  RG_CHECK(c->cp && c->scores_resident, RG_EINVAL, "cg_cols_pout: no rumgpu_cg_cols_scores since the last change");
  RG_HIP(hipSetDevice(c->ctx->device));
  RG_HIP(rg_copy(c->ctx->stream, p_out, c->cpout, (size_t)c->M * 8, hipMemcpyDeviceToHost));
  return RG_OK;
}
