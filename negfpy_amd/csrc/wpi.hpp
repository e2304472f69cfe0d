// One (omega, k) instance per wavefront, products on the matrix cores (gfx950).
//
// The workgroup kernels give an instance a whole workgroup, so its waves wait on each
// other at every barrier.  Here a 64-lane workgroup is one wavefront and owns one
// instance: there is no cross-wave synchronisation (the __syncthreads below only order
// LDS traffic inside the wave), and the GPU is filled by many resident instances.
// Primitives: wave.hpp.

#pragma once

namespace negf {
namespace wpi {

// ---------------------------------------------------------------------------
// Sancho-Rubio decimation, one instance per wave (reference
// src/negfpy/core/surface_gf.py:18-81): g = (z - e)^-1, e_s += a g b,
// e += a g b + b g a, a <- a g a, b <- b g b until max(|a|, |b|) < tol.
// FP64: writes g_s = (z - e_s)^-1 and the surface DOS; FP32: writes e_s for the
// FP64 completion, like k_sancho_rubio<float>.
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ void __launch_bounds__(WAVE) k_sr_wave(SrArgs a) {
  constexpr int TT = (8 * R + 15) / 16;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int inst = blockIdx.x;
  if (a.redo && a.status[inst] == ST_OK) return;
  const int n = a.n, ld = n + 1;
  const size_t msz = (size_t)n * ld;
  T* base = reinterpret_cast<T*>(smem);
  const LM<T> A{base, base + msz, ld}, B{base + 2 * msz, base + 3 * msz, ld},
      G{base + 4 * msz, base + 5 * msz, ld}, U{base + 6 * msz, base + 7 * msz, ld};
  GjW<T>& gb = *reinterpret_cast<GjW<T>*>(smem + (lds_bytes<T>(n, 4) - sizeof(GjW<T>)));
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;
  lm_gload<T>(A, blk + nn * 2, n, lane);      // alpha = d01
  lm_gload<T>(B, blk + 2 * nn * 2, n, lane);  // beta = d10
  Acc<T, TT> es, d, t;
  acc_gload<T, TT>(es, blk, n, lane);
  double zrd, zid;
  zsq(a.omega[inst], a.eta[inst], zrd, zid);
  const T zr = (T)zrd, zi = (T)zid;
  const T tol = (T)a.tol;
  // e = e_s + d: the surface term e_s collects a g b and the bulk remainder d collects
  // b g a, so both are pure MFMA accumulators
  acc_zero(d);
  int status = ST_OK, it = 0;
  bool conv = false;
  T gprev = 0;  // max |a|^2, |b|^2 of the previous doubling (FP32 stall test)
  while (it < a.maxiter) {
    ++it;
    acc_zminus2(es, d, G, n, zr, zi, lane);
    __syncthreads();
    const int st = gj_wave<T, R>(G, n, gb, lane);
    if (st != ST_OK) {
      status = st;
      break;
    }
    acc_zero(t);
    mm<T, TT, OPN, OPN>(A, G, n, t, lane);  // T1 = a g
    acc_store(t, U, n, lane);
    acc_zero(t);
    mm<T, TT, OPN, OPN>(B, G, n, t, lane);  // T2 = b g
    __syncthreads();                          // every read of g is done
    acc_store(t, G, n, lane);
    __syncthreads();
    mm<T, TT, OPN, OPN>(G, A, n, d, lane);  // d += T2 a
    acc_zero(t);
    mm<T, TT, OPN, OPN>(U, A, n, t, lane);  // a' = T1 a
    bool bad = false;
    T amax = acc_maxsq(t, bad);
    __syncthreads();
    acc_store(t, A, n, lane);
    mm<T, TT, OPN, OPN>(U, B, n, es, lane);  // e_s += T1 b
    acc_zero(t);
    mm<T, TT, OPN, OPN>(G, B, n, t, lane);  // b' = T2 b
    amax = fmax(amax, acc_maxsq(t, bad));
    __syncthreads();
    acc_store(t, B, n, lane);
    const T gmax = wave_max(amax);
    if (__any(bad)) {
      status = ST_NONFINITE;
      break;
    }
    if (sqrt(gmax) < tol) {
      conv = true;
      break;
    }
    if constexpr (sizeof(T) == 4) {  // the mixed path's FP32 pass: give up on a stalled mode
      constexpr T r2 = (T)(NEGF_MIXED_FP32_STALL_RATIO * NEGF_MIXED_FP32_STALL_RATIO);
      if (it >= NEGF_MIXED_FP32_STALL_FROM && gmax > r2 * gprev) break;
      gprev = gmax;
    }
  }
  if (status == ST_OK && !conv) status = ST_NOCONV;
  if (lane == 0) {
    a.status[inst] = status;
    a.iters[inst] = it;
  }
  if (status != ST_OK) return;
  double* out = a.g_out + (size_t)inst * nn * 2;
  if constexpr (sizeof(T) == 4) {
    acc_gstore(es, out, n, lane);
  } else {
    __syncthreads();
    acc_apply<T, TT, true>(es, G, n, zr, zi, lane);
    __syncthreads();
    const int st = gj_wave<T, R>(G, n, gb, lane);
    if (st != ST_OK) {
      if (lane == 0) a.status[inst] = st;
      return;
    }
    lm_gstore<T>(G, out, n, lane);
    if (a.dos_out) {
      double tr = 0.0;
      for (int i = lane; i < n; i += WAVE) tr += -2.0 * (double)G.im[i * ld + i];
      tr = wave_sum(tr);
      if (lane == 0) a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
    }
  }
}

// ---------------------------------------------------------------------------
// Recursive Green's function sweep for G_1N, one instance per wave (replaces the
// sparse LU of src/negfpy/core/negf.py:350-354): g_1 = A_11^-1, Y_i = g_i H_i,
// S_{i+1} = H_i^H Y_i, g_{i+1} = (A_{i+1,i+1} - S_{i+1})^-1, Q_i = Q_{i-1} Y_i,
// G_1N = Q_{N-1} g_N.
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ void __launch_bounds__(WAVE) k_rgf_wave(RgfArgs a) {
  constexpr int TT = (8 * R + 15) / 16;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int inst = blockIdx.x;
  if (a.redo && a.status[inst] != 10 + ST_NONFINITE && a.status[inst] != 10 + ST_SINGULAR) return;
  if (a.status_in && a.status_in[inst] != ST_OK) {
    if (lane == 0) a.status[inst] = a.status_in[inst];
    return;
  }
  const int n = a.n, nl = a.n_layers, ld = n + 1;
  const size_t msz = (size_t)n * ld, nn = (size_t)n * n;
  T* base = reinterpret_cast<T*>(smem);
  const LM<T> G{base, base + msz, ld}, H{base + 2 * msz, base + 3 * msz, ld},
      Q{base + 4 * msz, base + 5 * msz, ld};
  GjW<T>& gb = *reinterpret_cast<GjW<T>*>(smem + (lds_bytes<T>(n, 3) - sizeof(GjW<T>)));
  const size_t set = (size_t)a.set_of[inst];
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  double zrd, zid;
  zsq(a.omega[inst], a.eta_dev[inst], zrd, zid);
  const T zr = (T)zrd, zi = (T)zid;
  Acc<T, TT> s, t;
  acc_zero(s);
  int status = ST_OK;
  for (int layer = 0; layer < nl; ++layer) {
    // A = ((z - D - S) - Sigma_L) - Sigma_R into G: z - D and the contact
    // self-energies in coalesced passes, S from the accumulator
    const double2* d = ons + (size_t)layer * nn;
    for (int e = lane; e < (int)nn; e += WAVE) {
      const int i = e / n, j = e - (e / n) * n;
      const double2 dv = d[e];
      G.re[i * ld + j] = (i == j ? zr : T(0)) - (T)dv.x;
      G.im[i * ld + j] = (i == j ? zi : T(0)) - (T)dv.y;
    }
    __syncthreads();
    acc_apply<T, TT, false>(s, G, n, zr, zi, lane);
    if (layer == 0 || layer == nl - 1) {
      __syncthreads();
      for (int e = lane; e < (int)nn; e += WAVE) {
        const int i = e / n, j = e - (e / n) * n;
        T vr = G.re[i * ld + j], vi = G.im[i * ld + j];
        if (layer == 0) {
          vr -= (T)sgl[e].x;
          vi -= (T)sgl[e].y;
        }
        if (layer == nl - 1) {
          vr -= (T)sgr[e].x;
          vi -= (T)sgr[e].y;
        }
        G.re[i * ld + j] = vr;
        G.im[i * ld + j] = vi;
      }
    }
    __syncthreads();
    const int st = gj_wave<T, R>(G, n, gb, lane);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    double* out = a.g1n_out + (size_t)inst * nn * 2;
    if (layer == nl - 1) {
      if (nl == 1) {
        lm_gstore<T>(G, out, n, lane);
      } else {
        acc_zero(t);
        mm<T, TT, OPN, OPN>(Q, G, n, t, lane);
        acc_gstore(t, out, n, lane);
      }
      break;
    }
    lm_gload<T>(H, cpl + (size_t)layer * nn * 2, n, lane);
    __syncthreads();
    acc_zero(t);
    mm<T, TT, OPN, OPN>(G, H, n, t, lane);  // Y = g H
    __syncthreads();
    acc_store(t, G, n, lane);
    if (layer == 0) acc_store(t, Q, n, lane);
    __syncthreads();
    acc_zero(s);
    mm<T, TT, OPH, OPN>(H, G, n, s, lane);  // S = H^H Y
    if (layer > 0) {
      acc_zero(t);
      mm<T, TT, OPN, OPN>(Q, G, n, t, lane);  // Q = Q Y
      __syncthreads();
      acc_store(t, Q, n, lane);
    }
    __syncthreads();
  }
  if (lane == 0) a.status[inst] = status;
}

}  // namespace wpi
}  // namespace negf
