// The plain window-pass transform kernel (any window, any 2x2 coefficients), shared by the
// BM-ADMM certificates and the slab-form pricing.
#pragma once
#include "bfly.hpp"

// ---------------------------------------------------------------------------------------------
// Plain transform pass (any window): dst = T(src) on the window. PRO 1: the input is the
// extract of the push (rf = mass[tail] w / outw[tail], arcflow_kernels.py:378-386) read from dst
// itself; EPI 1: wmax (w = max(q, 0) + 1e-12, arcflow_kernels.py:354) before the store.
template <typename TIN, typename T, int RL, int PRO, int EPI>
__global__ __launch_bounds__(512) void tr_pass(const TIN* src, T* dst, Coef<T> k, WinGeom g,
                                                const double* outw, const double* mass) {
  extern __shared__ __align__(16) unsigned char smem[];
  Tile<T, RL> t;
  t.init(g, reinterpret_cast<T*>(smem));
  constexpr int R = Tile<T, RL>::R;
  if constexpr (PRO == 1) {
    if (t.active) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const i64 a = t.gidx(j);
        // item blocks of slot d are items [d n, (d + 1) n); its node arrays start at d 2^n
        const int d = t.item / g.nper, x = t.item - d * g.nper;
        const i64 tail = ((i64)d << g.nper) + (i64)expand_bit((u64)t.rest(j), x);
        const double ow = outw[tail];
        t.v[j] = (T)((ow > 0.0) ? mass[tail] * (double)src[a] / ow : 0.0);
      }
    }
  } else {
    t.load(src);
  }
  t.transform(k);
  if constexpr (EPI == 1) {
#pragma unroll
    for (int j = 0; j < R; ++j) t.v[j] = (t.v[j] > (T)0 ? t.v[j] : (T)0) + (T)1e-12;
  }
  t.store(dst);
}


template <typename T>
inline Coef<T> coef(double a, double b, double c, double d) {
  return Coef<T>{(T)a, (T)b, (T)c, (T)d};
}

template <typename TIN, typename T, int PRO, int EPI>
int launch_tr_rl(int RL, const TIN* src, T* dst, Coef<T> k, const WinGeom& g, const double* outw,
                 const double* mass, hipStream_t st) {
  const size_t shm = (size_t)g.TPB * tile_lds_elems(g) * sizeof(T);
  const dim3 grid((unsigned)((g.ntiles + g.TPB - 1) / g.TPB)), block((unsigned)(g.TT * g.TPB));
  switch (RL) {
    case 1: hipLaunchKernelGGL((tr_pass<TIN, T, 1, PRO, EPI>), grid, block, shm, st, src, dst, k, g, outw, mass); break;
    case 2: hipLaunchKernelGGL((tr_pass<TIN, T, 2, PRO, EPI>), grid, block, shm, st, src, dst, k, g, outw, mass); break;
    case 3: hipLaunchKernelGGL((tr_pass<TIN, T, 3, PRO, EPI>), grid, block, shm, st, src, dst, k, g, outw, mass); break;
    case 4: hipLaunchKernelGGL((tr_pass<TIN, T, 4, PRO, EPI>), grid, block, shm, st, src, dst, k, g, outw, mass); break;
    case 5: hipLaunchKernelGGL((tr_pass<TIN, T, 5, PRO, EPI>), grid, block, shm, st, src, dst, k, g, outw, mass); break;
    default: rg_set_error("bad RL %d", RL); return RG_EINVAL;
  }
  RG_LAUNCH_CHECK();
  return RG_OK;
}

