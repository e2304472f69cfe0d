// Workgroup-cooperative complex matrix primitives for gfx950 (CDNA4).
//
// Every matrix of one (omega, k) instance lives in LDS in planar form: a real
// plane and an imaginary plane, each NMAX rows of LD = NMAX + 2 entries.  The
// odd-pair padding keeps column walks (pivot search, A-operand reads) spread
// over the LDS banks.  A thread "owns" a TM x TM output tile of every product
// and keeps tiles of state matrices (onsite energies, Schur accumulators) in
// VGPRs, so only operands that other threads must read go through LDS.
//
// Runtime block sizes n <= NMAX are handled by masking: only tiles that
// intersect [0, n)^2 are owned, the K loops run over [0, n), and entries of
// owned tiles outside [0, n) are never written back to live data.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace negf {

constexpr int WAVE = 64;

enum Op { OPN = 0, OPH = 1, OPC = 2, OPT = 3 };

// Ownership of a TM x TM tile of an m x p output by thread `tid`.
template <int TM>
struct Tile {
  int i0, j0;
  bool own;
  __device__ __forceinline__ Tile(int m, int p, int tid) {
    const int ntr = (m + TM - 1) / TM;
    const int ntc = (p + TM - 1) / TM;
    own = tid < ntr * ntc;
    const int t = own ? tid : 0;
    i0 = (t / ntc) * TM;
    j0 = (t % ntc) * TM;
  }
};

// A-operand element (i, kk) of op(A).
template <typename T, int OP>
__device__ __forceinline__ void lda(const T* __restrict__ re, const T* __restrict__ im, int ld, int i,
                                    int kk, T& r, T& m) {
  if constexpr (OP == OPN) {
    r = re[i * ld + kk];
    m = im[i * ld + kk];
  } else if constexpr (OP == OPH) {
    r = re[kk * ld + i];
    m = -im[kk * ld + i];
  } else if constexpr (OP == OPC) {
    r = re[i * ld + kk];
    m = -im[i * ld + kk];
  } else {
    r = re[kk * ld + i];
    m = im[kk * ld + i];
  }
}

// B-operand element (kk, j) of op(B).
template <typename T, int OP>
__device__ __forceinline__ void ldb(const T* __restrict__ re, const T* __restrict__ im, int ld, int kk,
                                    int j, T& r, T& m) {
  if constexpr (OP == OPN) {
    r = re[kk * ld + j];
    m = im[kk * ld + j];
  } else if constexpr (OP == OPH) {
    r = re[j * ld + kk];
    m = -im[j * ld + kk];
  } else if constexpr (OP == OPC) {
    r = re[kk * ld + j];
    m = -im[kk * ld + j];
  } else {
    r = re[j * ld + kk];
    m = im[j * ld + kk];
  }
}

template <typename T>
__device__ __forceinline__ void cmac(T& cr, T& ci, T ar, T ai, T br, T bi) {
  cr = fma(ar, br, cr);
  cr = fma(-ai, bi, cr);
  ci = fma(ar, bi, ci);
  ci = fma(ai, br, ci);
}

template <typename T, int TM>
__device__ __forceinline__ void tzero(T (&r)[TM][TM], T (&i)[TM][TM]) {
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      r[a][b] = T(0);
      i[a][b] = T(0);
    }
}

// C_tile += op(A)[i0:i0+TM, 0:k] * op(B)[0:k, j0:j0+TM]
template <typename T, int TM, int OPA, int OPB>
__device__ __forceinline__ void mac(const T* are, const T* aim, const T* bre, const T* bim, int ld, int k,
                                    int i0, int j0, T (&cr)[TM][TM], T (&ci)[TM][TM]) {
#pragma unroll 2
  for (int kk = 0; kk < k; ++kk) {
    T ar[TM], ai[TM], br[TM], bi[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) lda<T, OPA>(are, aim, ld, i0 + a, kk, ar[a], ai[a]);
#pragma unroll
    for (int b = 0; b < TM; ++b) ldb<T, OPB>(bre, bim, ld, kk, j0 + b, br[b], bi[b]);
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TM; ++b) cmac(cr[a][b], ci[a][b], ar[a], ai[a], br[b], bi[b]);
  }
}

// Two products sharing the B operand: C1 += A1 B, C2 += A2 B.
template <typename T, int TM>
__device__ __forceinline__ void mac2b(const T* a1r, const T* a1i, const T* a2r, const T* a2i,
                                      const T* bre, const T* bim, int ld, int k, int i0, int j0,
                                      T (&c1r)[TM][TM], T (&c1i)[TM][TM], T (&c2r)[TM][TM],
                                      T (&c2i)[TM][TM]) {
#pragma unroll 2
  for (int kk = 0; kk < k; ++kk) {
    T xr[TM], xi[TM], yr[TM], yi[TM], br[TM], bi[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      xr[a] = a1r[(i0 + a) * ld + kk];
      xi[a] = a1i[(i0 + a) * ld + kk];
      yr[a] = a2r[(i0 + a) * ld + kk];
      yi[a] = a2i[(i0 + a) * ld + kk];
    }
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      br[b] = bre[kk * ld + j0 + b];
      bi[b] = bim[kk * ld + j0 + b];
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TM; ++b) {
        cmac(c1r[a][b], c1i[a][b], xr[a], xi[a], br[b], bi[b]);
        cmac(c2r[a][b], c2i[a][b], yr[a], yi[a], br[b], bi[b]);
      }
  }
}

// The four decimation products of one Sancho-Rubio step, sharing operands:
//   P1 = T1 b, P2 = T2 a, P3 = T1 a, P4 = T2 b   with T1 = a g, T2 = b g.
template <typename T, int TM>
__device__ __forceinline__ void mac4(const T* t1r, const T* t1i, const T* t2r, const T* t2i,
                                     const T* ar_, const T* ai_, const T* br_, const T* bi_, int ld,
                                     int k, int i0, int j0, T (&p1r)[TM][TM], T (&p1i)[TM][TM],
                                     T (&p2r)[TM][TM], T (&p2i)[TM][TM], T (&p3r)[TM][TM],
                                     T (&p3i)[TM][TM], T (&p4r)[TM][TM], T (&p4i)[TM][TM]) {
#pragma unroll 2
  for (int kk = 0; kk < k; ++kk) {
    T xr[TM], xi[TM], yr[TM], yi[TM], ur[TM], ui[TM], vr[TM], vi[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      xr[a] = t1r[(i0 + a) * ld + kk];
      xi[a] = t1i[(i0 + a) * ld + kk];
      yr[a] = t2r[(i0 + a) * ld + kk];
      yi[a] = t2i[(i0 + a) * ld + kk];
    }
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      ur[b] = ar_[kk * ld + j0 + b];
      ui[b] = ai_[kk * ld + j0 + b];
      vr[b] = br_[kk * ld + j0 + b];
      vi[b] = bi_[kk * ld + j0 + b];
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TM; ++b) {
        cmac(p1r[a][b], p1i[a][b], xr[a], xi[a], vr[b], vi[b]);
        cmac(p2r[a][b], p2i[a][b], yr[a], yi[a], ur[b], ui[b]);
        cmac(p3r[a][b], p3i[a][b], xr[a], xi[a], ur[b], ui[b]);
        cmac(p4r[a][b], p4i[a][b], yr[a], yi[a], vr[b], vi[b]);
      }
  }
}

// Store the valid part of an owned tile into an LDS matrix.
template <typename T, int TM>
__device__ __forceinline__ void tstore(T* re, T* im, int ld, int m, int p, const Tile<TM>& t,
                                       const T (&cr)[TM][TM], const T (&ci)[TM][TM]) {
  if (!t.own) return;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      const int i = t.i0 + a, j = t.j0 + b;
      if (i < m && j < p) {
        re[i * ld + j] = cr[a][b];
        im[i * ld + j] = ci[a][b];
      }
    }
}

// Complex reciprocal by Smith's scaled division.
template <typename T>
__device__ __forceinline__ void crecip(T ar, T ai, T& rr, T& ri) {
  if (fabs(ar) >= fabs(ai)) {
    const T r = ai / ar;
    const T d = ar + ai * r;
    rr = T(1) / d;
    ri = -r / d;
  } else {
    const T r = ar / ai;
    const T d = ai + ar * r;
    rr = r / d;
    ri = T(-1) / d;
  }
}

// Per-thread coordinates of the elements e = tid + m*NT of an n x n matrix.
template <int NT, int MAXE>
struct ElemMap {
  int16_t i[MAXE], j[MAXE];
  int cnt;
  __device__ __forceinline__ ElemMap(int n, int tid) {
    const int nn = n * n;
    cnt = 0;
#pragma unroll
    for (int m = 0; m < MAXE; ++m) {
      const int e = tid + m * NT;
      const int r = e / n;
      i[m] = (int16_t)r;
      j[m] = (int16_t)(e - r * n);
      if (e < nn) cnt = m + 1;
    }
  }
};

// Status codes shared with the host (see include/negf_hip.h).
enum Status { ST_OK = 0, ST_NOCONV = 1, ST_NONFINITE = 2, ST_SINGULAR = 3, ST_REDO = 7 };

// In-place Gauss-Jordan inversion with partial pivoting (largest |re|+|im| in
// the column, first index on ties, as LAPACK's izamax), row interchanges during
// elimination and the inverse column permutation at the end.  All threads of
// the workgroup must call it.  Returns ST_OK, ST_SINGULAR (exact zero pivot) or
// ST_NONFINITE (NaN column).  Ends with a barrier.
template <typename T, int NT, int MAXE>
__device__ int gj_inverse(T* __restrict__ re, T* __restrict__ im, int ld, int n, T* rbr, T* rbi,
                          T* cbr, T* cbi, int* perm, const ElemMap<NT, MAXE>& em, int tid) {
  const int lane = tid & (WAVE - 1);
  for (int k = 0; k < n; ++k) {
    T best = T(-1);
    int bidx = n;
    bool nan_seen = false;
    for (int i = k + lane; i < n; i += WAVE) {
      const T mg = fabs(re[i * ld + k]) + fabs(im[i * ld + k]);
      if (mg != mg) nan_seen = true;
      if (mg > best) {
        best = mg;
        bidx = i;
      }
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
      const T ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      if (ob > best || (ob == best && oi < bidx)) {
        best = ob;
        bidx = oi;
      }
    }
    const bool any_nan = __any(nan_seen);
    if (any_nan) {
      __syncthreads();
      return ST_NONFINITE;
    }
    if (!(best > T(0))) {
      __syncthreads();
      return ST_SINGULAR;
    }
    const int p = bidx;
    T ivr, ivi;
    crecip(re[p * ld + k], im[p * ld + k], ivr, ivi);
    for (int j = tid; j < n; j += NT) {
      const T vr = re[p * ld + j], vi = im[p * ld + j];
      if (p != k && j != k) {
        re[p * ld + j] = re[k * ld + j];
        im[p * ld + j] = im[k * ld + j];
      }
      if (j == k) {
        rbr[j] = ivr;
        rbi[j] = ivi;
      } else {
        rbr[j] = vr * ivr - vi * ivi;
        rbi[j] = vr * ivi + vi * ivr;
      }
    }
    for (int i = tid; i < n; i += NT) {
      const int src = (i == p) ? k : i;
      cbr[i] = re[src * ld + k];
      cbi[i] = im[src * ld + k];
    }
    if (tid == 0) perm[k] = p;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXE; ++m) {
      if (m < em.cnt) {
        const int i = em.i[m], j = em.j[m];
        if (tid + m * NT < n * n) {
          if (i == k) {
            re[k * ld + j] = rbr[j];
            im[k * ld + j] = rbi[j];
          } else if (j == k) {
            const T ar = cbr[i], ai = cbi[i];
            re[i * ld + k] = -(ar * ivr - ai * ivi);
            im[i * ld + k] = -(ar * ivi + ai * ivr);
          } else {
            const T ar = cbr[i], ai = cbi[i], br = rbr[j], bi = rbi[j];
            T xr = re[i * ld + j], xi = im[i * ld + j];
            xr = fma(-ar, br, xr);
            xr = fma(ai, bi, xr);
            xi = fma(-ar, bi, xi);
            xi = fma(-ai, br, xi);
            re[i * ld + j] = xr;
            im[i * ld + j] = xi;
          }
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += NT) {
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        const T tr = re[i * ld + k], ti = im[i * ld + k];
        re[i * ld + k] = re[i * ld + p];
        im[i * ld + k] = im[i * ld + p];
        re[i * ld + p] = tr;
        im[i * ld + p] = ti;
      }
    }
  }
  __syncthreads();
  return ST_OK;
}

// Workgroup max-reduction of a non-negative value plus an "any" flag.
template <int NT>
__device__ __forceinline__ void block_max_any(double v, bool flag, double* sred, int* sflag, int tid,
                                              double& vmax, bool& fany) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  const bool wf = __any(flag);
  const int w = tid / WAVE;
  if ((tid & (WAVE - 1)) == 0) {
    sred[w] = v;
    sflag[w] = wf ? 1 : 0;
  }
  __syncthreads();
  vmax = sred[0];
  int f = sflag[0];
#pragma unroll
  for (int q = 1; q < NT / WAVE; ++q) {
    vmax = fmax(vmax, sred[q]);
    f |= sflag[q];
  }
  fany = f != 0;
  __syncthreads();
}

// Workgroup sum of doubles (result valid in all threads).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sred, int tid) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((tid & (WAVE - 1)) == 0) sred[tid / WAVE] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NT / WAVE; ++q) s += sred[q];
  __syncthreads();
  return s;
}

// z = (w + i eta)^2 evaluated as (w*w - eta*eta) + i (w*eta + eta*w), each
// operation rounded separately (no FMA contraction), like complex ** 2 in Python.
__device__ __forceinline__ void zsq(double w, double eta, double& zr, double& zi) {
  zr = __dsub_rn(__dmul_rn(w, w), __dmul_rn(eta, eta));
  zi = __dadd_rn(__dmul_rn(w, eta), __dmul_rn(eta, w));
}

}  // namespace negf
