// Wavefront-level complex matrix primitives for gfx950: MFMA products and a
// register-resident wave Gauss-Jordan inversion.
//
// Every routine here is executed by one 64-lane wavefront on data it alone touches, so
// ordering its LDS traffic needs no barrier: wave_sync() waits for the wave's own LDS
// operations (lgkmcnt) and stops the compiler from moving memory accesses across it.
// The one-instance-per-wave kernels (wpi.hpp) use them, and the workgroup kernels
// run their inversions on one wave with them.
//
// Matrices live in LDS in planar form (real plane, imaginary plane), n rows of
// LD = n + 1 entries.  Products C += op(A) op(B) run on
//   v_mfma_f32_16x16x4f32 (FP32)   or   v_mfma_f64_16x16x4f64 (FP64)
// as four real products per complex one, over 16 x 16 output tiles (TT x TT of
// them, n <= 16 TT), the operands loaded from LDS per k-step (rows/columns past n
// read as zero).  Accumulators stay in VGPRs in the MFMA output layout:
//   FP32: lane l, register v holds C[4 (l / 16) + v][l % 16]
//   FP64: lane l, register v holds C[4 v + l / 16][l % 16]
// (both checked on the device against a host product).
//
// Inversion: in-place Gauss-Jordan with partial pivoting (largest |re| + |im| among
// unused rows, first index on ties, as LAPACK's izamax) without row exchanges,
// register resident: the lanes form an 8 x 8 grid, lane (a, b) owns rows a + 8 m
// and columns b + 8 q (m, q < R, n <= 8 R).  A step needs column k (published by its
// 8 owner lanes, which also find the pivot with a DPP arg-max) and pivot row p
// (published by its 8 owner lanes) through two small LDS buffers; the result is
// scattered once: A^-1[q(i)][p(j)] = Y[i][j], p(k) = pivot row of step k, q = p^-1.

#pragma once

namespace negf {
namespace wpi {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<float> {
  using V = v4f;
  __device__ static __forceinline__ V fma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output-layout row of register v in lane l (column = l % 16)
  __device__ static __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mf<double> {
  using V = v4d;
  __device__ static __forceinline__ V fma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};

// Orders this wavefront's LDS accesses: its own earlier LDS operations have completed
// and the compiler keeps memory accesses on their side of the call.
__device__ __forceinline__ void wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// An opaque copy of a lane-dependent value: helpers re-derive their per-lane offsets
// and bounds masks from it at every call instead of the compiler hoisting dozens of
// them out of the iteration loops into registers.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// planar LDS matrix
template <typename T>
struct LM {
  T* re;
  T* im;
  int ld;
};

// complex accumulator in the MFMA output layout: TT x TT tiles of 16 x 16
template <typename T, int TT>
struct Acc {
  typename Mf<T>::V re[TT][TT], im[TT][TT];
};

template <typename T, int TT>
__device__ __forceinline__ void acc_zero(Acc<T, TT>& c) {
#pragma unroll
  for (int a = 0; a < TT; ++a)
#pragma unroll
    for (int b = 0; b < TT; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        c.re[a][b][v] = T(0);
        c.im[a][b][v] = T(0);
      }
}

// element (i, k) of op(X); rows/columns at or past n read as zero
template <typename T, int OP>
__device__ __forceinline__ void ldop(const LM<T>& X, int i, int k, bool ok, T& r, T& m) {
  const int ii = ok ? i : 0, kk = ok ? k : 0;
  T vr, vi;
  if constexpr (OP == OPN) {
    vr = X.re[ii * X.ld + kk];
    vi = X.im[ii * X.ld + kk];
  } else if constexpr (OP == OPH) {
    vr = X.re[kk * X.ld + ii];
    vi = -X.im[kk * X.ld + ii];
  } else if constexpr (OP == OPC) {
    vr = X.re[ii * X.ld + kk];
    vi = -X.im[ii * X.ld + kk];
  } else {
    vr = X.re[kk * X.ld + ii];
    vi = X.im[kk * X.ld + ii];
  }
  r = ok ? vr : T(0);
  m = ok ? vi : T(0);
}

// c += op(A) op(B), n x n complex, four real MFMA chains per output tile
template <typename T, int TT, int OPA, int OPB>
__device__ __forceinline__ void mm(const LM<T>& A, const LM<T>& B, int n, Acc<T, TT>& c, int lane_) {
  using M = Mf<T>;
  const int lane = opaque(lane_);
  const int r16 = lane & 15, q = lane >> 4;
  const int nks = (n + 3) >> 2;
#pragma unroll 2
  for (int s = 0; s < nks; ++s) {
    const int k = 4 * s + q;
    T ar[TT], ai[TT], na[TT], br[TT], bi[TT];
#pragma unroll
    for (int t = 0; t < TT; ++t) {
      const int i = 16 * t + r16;
      const bool ok = i < n && k < n;
      ldop<T, OPA>(A, i, k, ok, ar[t], ai[t]);
      ldop<T, OPB>(B, k, i, ok, br[t], bi[t]);
      na[t] = -ai[t];
    }
#pragma unroll
    for (int x = 0; x < TT; ++x)
#pragma unroll
      for (int y = 0; y < TT; ++y) c.re[x][y] = M::fma(ar[x], br[y], c.re[x][y]);
#pragma unroll
    for (int x = 0; x < TT; ++x)
#pragma unroll
      for (int y = 0; y < TT; ++y) c.im[x][y] = M::fma(ar[x], bi[y], c.im[x][y]);
#pragma unroll
    for (int x = 0; x < TT; ++x)
#pragma unroll
      for (int y = 0; y < TT; ++y) c.re[x][y] = M::fma(na[x], bi[y], c.re[x][y]);
#pragma unroll
    for (int x = 0; x < TT; ++x)
#pragma unroll
      for (int y = 0; y < TT; ++y) c.im[x][y] = M::fma(ai[x], br[y], c.im[x][y]);
  }
}

// store the n x n part of an accumulator into an LDS matrix
template <typename T, int TT>
__device__ __forceinline__ void acc_store(const Acc<T, TT>& c, const LM<T>& X, int n, int lane_) {
  const int lane = opaque(lane_);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * x + Mf<T>::row(v, lane), j = 16 * y + (lane & 15);
        if (i < n && j < n) {
          X.re[i * X.ld + j] = c.re[x][y][v];
          X.im[i * X.ld + j] = c.im[x][y][v];
        }
      }
}

// X <- z I - c (zdiag) or X <- X - c, on the n x n part
template <typename T, int TT, bool ZMINUS>
__device__ __forceinline__ void acc_apply(const Acc<T, TT>& c, const LM<T>& X, int n, T zr, T zi,
                                          int lane_) {
  const int lane = opaque(lane_);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * x + Mf<T>::row(v, lane), j = 16 * y + (lane & 15);
        if (i < n && j < n) {
          const int o = i * X.ld + j;
          if constexpr (ZMINUS) {
            X.re[o] = (i == j ? zr : T(0)) - c.re[x][y][v];
            X.im[o] = (i == j ? zi : T(0)) - c.im[x][y][v];
          } else {
            X.re[o] = X.re[o] - c.re[x][y][v];
            X.im[o] = X.im[o] - c.im[x][y][v];
          }
        }
      }
}

// X <- (z I - a) - b on the n x n part
template <typename T, int TT>
__device__ __forceinline__ void acc_zminus2(const Acc<T, TT>& a, const Acc<T, TT>& b, const LM<T>& X,
                                            int n, T zr, T zi, int lane_) {
  const int lane = opaque(lane_);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * x + Mf<T>::row(v, lane), j = 16 * y + (lane & 15);
        if (i < n && j < n) {
          const int o = i * X.ld + j;
          X.re[o] = ((i == j ? zr : T(0)) - a.re[x][y][v]) - b.re[x][y][v];
          X.im[o] = ((i == j ? zi : T(0)) - a.im[x][y][v]) - b.im[x][y][v];
        }
      }
}

// max |c_ij|^2 over the lane's entries, and whether any is not finite
template <typename T, int TT>
__device__ __forceinline__ T acc_maxsq(const Acc<T, TT>& c, bool& bad) {
  T amax = T(0);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T r = c.re[x][y][v], m = c.im[x][y][v];
        const T a2 = r * r + m * m;
        bad |= !(a2 < T(INFINITY));
        amax = fmax(amax, a2);
      }
  return amax;
}

// accumulator from a complex128 n x n global block (row-major)
template <typename T, int TT>
__device__ __forceinline__ void acc_gload(Acc<T, TT>& c, const double* __restrict__ g, int n,
                                          int lane_) {
  const int lane = opaque(lane_);
  const double2* s = reinterpret_cast<const double2*>(g);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * x + Mf<T>::row(v, lane), j = 16 * y + (lane & 15);
        double2 e = make_double2(0.0, 0.0);
        if (i < n && j < n) e = s[i * n + j];
        c.re[x][y][v] = (T)e.x;
        c.im[x][y][v] = (T)e.y;
      }
}

// complex128 global block from an accumulator
template <typename T, int TT>
__device__ __forceinline__ void acc_gstore(const Acc<T, TT>& c, double* __restrict__ g, int n,
                                           int lane_) {
  const int lane = opaque(lane_);
  double2* d = reinterpret_cast<double2*>(g);
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * x + Mf<T>::row(v, lane), j = 16 * y + (lane & 15);
        if (i < n && j < n) d[i * n + j] = make_double2((double)c.re[x][y][v], (double)c.im[x][y][v]);
      }
}

// LDS matrix <- complex128 global n x n block
template <typename T>
__device__ __forceinline__ void lm_gload(const LM<T>& X, const double* __restrict__ g, int n,
                                         int lane) {
  const double2* s = reinterpret_cast<const double2*>(g);
  const int nn = n * n;
  for (int e = lane; e < nn; e += WAVE) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 v = s[e];
    X.re[i * X.ld + j] = (T)v.x;
    X.im[i * X.ld + j] = (T)v.y;
  }
}

template <typename T>
__device__ __forceinline__ void lm_gstore(const LM<T>& X, double* __restrict__ g, int n, int lane) {
  double2* d = reinterpret_cast<double2*>(g);
  const int nn = n * n;
  for (int e = lane; e < nn; e += WAVE) {
    const int i = e / n, j = e - (e / n) * n;
    d[e] = make_double2((double)X.re[i * X.ld + j], (double)X.im[i * X.ld + j]);
  }
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// LDS exchange buffers of the wave Gauss-Jordan (n <= 32)
template <typename T>
struct GjW {
  T cr[32], ci[32];  // column k of the current state, by row
  T rr[32], ri[32];  // pivot row of the current state, by column
  int pv[32];        // pivot row of step k
  int qv[32];        // step at which row r was the pivot
};

// register (uniform index u) of a small array
template <typename T, int R>
__device__ __forceinline__ T pick(const T (&x)[R], int u) {
  T v = x[0];
#pragma unroll
  for (int t = 1; t < R; ++t) v = (u == t) ? x[t] : v;
  return v;
}

// In-place inversion of the n x n LDS matrix X (n <= 8 R) by one wavefront.
// Returns ST_OK, ST_SINGULAR or ST_NONFINITE (uniform); X is untouched on failure.
//
// Step k (pivot p, r = 1 / pivot): every entry takes x - c_i s'_j with c = column k
// and s' = r * row p, and then two fix-ups overwrite what that formula gets wrong
// -- the pivot row becomes s' (r in column k) and column k becomes -c_i r -- by
// direct register writes in their owner lanes (the register index is uniform), so
// the bulk update carries no per-entry selects.
template <typename T, int R>
__device__ __forceinline__ int gj_wave(const LM<T>& X, int n, GjW<T>& b, int lane_) {
  static_assert(R >= 1 && R <= 4, "the wave inversion covers n <= 32");
  const int lane = opaque(lane_);
  const int ra = lane & 7, cb = lane >> 3;
  T xr[R][R], xi[R][R];
  unsigned used = 0;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int i = ra + 8 * m;
    if (i >= n) used |= 1u << m;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int j = cb + 8 * q;
      const bool ok = i < n && j < n;
      const int off = ok ? i * X.ld + j : 0;
      const T vr = X.re[off], vi = X.im[off];
      xr[m][q] = ok ? vr : T(0);
      xi[m][q] = ok ? vi : T(0);
    }
  }
  wave_sync();
  for (int k = 0; k < n; ++k) {
    const int kb = k & 7, kq = k >> 3;
    T best = T(-1);
    int bidx = n;
    if (cb == kb) {
      bool bad = false;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        T cr = xr[m][0], ci = xi[m][0];
#pragma unroll
        for (int q = 1; q < R; ++q) {
          cr = (kq == q) ? xr[m][q] : cr;
          ci = (kq == q) ? xi[m][q] : ci;
        }
        b.cr[ra + 8 * m] = cr;
        b.ci[ra + 8 * m] = ci;
        const T mg = fabs(cr) + fabs(ci);
        const bool live = !((used >> m) & 1u);
        bad = bad || (live && mg != mg);
        const bool better = live && mg > best;
        best = better ? mg : best;
        bidx = better ? ra + 8 * m : bidx;
      }
      if (bad) best = T(INFINITY);
      group_argmax<T, 8>(best, bidx);
    }
    const int p = __builtin_amdgcn_readlane(bidx, 8 * kb);
    const T pbest = readlane(best, 8 * kb);
    const int flag = isinf(pbest) ? ST_NONFINITE : (pbest > T(0) ? ST_OK : ST_SINGULAR);
    if (flag != ST_OK) {
      wave_sync();
      return flag;
    }
    const int mp = p >> 3;
    if (ra == (p & 7)) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        T vr = xr[0][q], vi = xi[0][q];
#pragma unroll
        for (int m = 1; m < R; ++m) {
          vr = (mp == m) ? xr[m][q] : vr;
          vi = (mp == m) ? xi[m][q] : vi;
        }
        b.rr[cb + 8 * q] = vr;
        b.ri[cb + 8 * q] = vi;
      }
    }
    if (lane == 0) {
      b.pv[k] = p;
      b.qv[p] = k;
    }
    wave_sync();
    T pr_, pi_;
    crecip1(b.cr[p], b.ci[p], pr_, pi_);
    T sr[R], si[R], c_r[R], c_i[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T vr = b.rr[cb + 8 * q], vi = b.ri[cb + 8 * q];
      const bool jk = (cb + 8 * q) == k;
      sr[q] = jk ? pr_ : vr * pr_ - vi * pi_;
      si[q] = jk ? pi_ : vr * pi_ + vi * pr_;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
      c_r[m] = b.cr[ra + 8 * m];
      c_i[m] = b.ci[ra + 8 * m];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        T yr = xr[m][q], yi = xi[m][q];
        yr = fma(-c_r[m], sr[q], yr);
        yr = fma(c_i[m], si[q], yr);
        yi = fma(-c_r[m], si[q], yi);
        yi = fma(-c_i[m], sr[q], yi);
        xr[m][q] = yr;
        xi[m][q] = yi;
      }
      used |= (ra + 8 * m == p) ? (1u << m) : 0u;
    }
    // column k: -c_i r (r on the pivot row)
    if (cb == kb) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (kq == q) {
#pragma unroll
          for (int m = 0; m < R; ++m) {
            const bool isp = (ra + 8 * m) == p;
            xr[m][q] = isp ? pr_ : -(c_r[m] * pr_ - c_i[m] * pi_);
            xi[m][q] = isp ? pi_ : -(c_r[m] * pi_ + c_i[m] * pr_);
          }
        }
      }
    }
    // pivot row: s'
    if (ra == (p & 7)) {
#pragma unroll
      for (int m = 0; m < R; ++m) {
        if (mp == m) {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            xr[m][q] = sr[q];
            xi[m][q] = si[q];
          }
        }
      }
    }
    wave_sync();
  }
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int i = ra + 8 * m;
    if (i < n) {
      const int qi = b.qv[i];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int j = cb + 8 * q;
        if (j < n) {
          const int pj = b.pv[j];
          X.re[qi * X.ld + pj] = xr[m][q];
          X.im[qi * X.ld + pj] = xi[m][q];
        }
      }
    }
  }
  wave_sync();
  return ST_OK;
}

// LDS bytes of a wave kernel holding `nmat` n x n planar matrices
template <typename T>
__host__ __device__ constexpr size_t lds_bytes(int n, int nmat) {
  return ((sizeof(T) * (size_t)(2 * nmat) * n * (n + 1) + 15) & ~size_t(15)) + sizeof(GjW<T>);
}

}  // namespace wpi
}  // namespace negf
