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

// Status codes shared with the host (see include/negf_hip.h).
enum Status { ST_OK = 0, ST_NOCONV = 1, ST_NONFINITE = 2, ST_SINGULAR = 3, ST_REDO = 7 };

// 32-bit DPP lane moves used by the pivot search (row_mask = bank_mask = 0xF).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = dpp_i<CTRL>((int)(u & 0xffffffffLL));
  const int hi = dpp_i<CTRL>((int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

template <typename T, int CTRL>
__device__ __forceinline__ void argmax_step(T& best, int& idx) {
  const T ob = dpp_v<CTRL>(best);
  const int oi = dpp_i<CTRL>(idx);
  if (ob > best || (ob == best && oi < idx)) {
    best = ob;
    idx = oi;
  }
}

// Arg-max (largest value, smallest index on ties) over aligned groups of RG
// consecutive lanes, RG in {1, 2, 4, 8, 16}; every lane of a group ends with it.
template <typename T, int RG>
__device__ __forceinline__ void group_argmax(T& best, int& idx) {
  if constexpr (RG >= 2) argmax_step<T, 0xB1>(best, idx);   // quad_perm [1,0,3,2]: lane ^ 1
  if constexpr (RG >= 4) argmax_step<T, 0x4E>(best, idx);   // quad_perm [2,3,0,1]: lane ^ 2
  if constexpr (RG >= 8) argmax_step<T, 0x141>(best, idx);  // row_half_mirror: the other quad
  if constexpr (RG >= 16) argmax_step<T, 0x140>(best, idx); // row_mirror: the other half-row
}

// Pivot of one elimination step, double-buffered by step parity in LDS.
template <typename T>
struct Pivot {
  T ir, ii;  // reciprocal of the pivot element
  int p;     // pivot row
  int flag;  // ST_OK, ST_SINGULAR or ST_NONFINITE
};

template <typename T, int NMAX>
struct GjBuf {
  Pivot<T> piv[2];
  int pv[NMAX];  // pivot row of step k
  int qv[NMAX];  // step at which row r was the pivot
};

// 1 / (ar + i ai) = conj(a) / |a|^2; the real reciprocal is the hardware
// estimate refined by one Newton step (error well below one ulp of the result).
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  const float e = fmaf(-d, r, 1.0f);
  return fmaf(r, e, r);
}
template <typename T>
__device__ __forceinline__ void crecip1(T ar, T ai, T& rr, T& ri) {
  const T d = rcp_nr(ar * ar + ai * ai);
  rr = ar * d;
  ri = -ai * d;
}

// In-place inversion of the n x n LDS matrix (re, im) by Gauss-Jordan
// elimination with partial pivoting: the pivot of column k is the largest
// |re| + |im| among the rows not yet used (first index on ties, as LAPACK's
// izamax).  Rows are never exchanged; step k pivots on row p_k and the result
// is scattered once at the end, A^-1[q(r)][p_m] = Y[r][m] with q = p^-1.
//
// Thread t < n_cols * RG owns column j = t / RG and rows g, g + RG, ...
// (g = t % RG) in registers, so the RG owners of a column are consecutive lanes
// of one wave and the pivot search is an RG-lane DPP reduction.  The state
// after each step is written to LDS, alternating between (re, im) and the free
// matrix (wre, wim), so one barrier per step suffices: step k reads the pivot
// row and pivot column of the state written in step k-1.  All threads must
// call it; it ends with a barrier.  Returns ST_OK, ST_SINGULAR or ST_NONFINITE.
template <typename T, int NT, int NMAX, int RG>
__device__ __forceinline__ int gj_inverse(T* __restrict__ re, T* __restrict__ im,
                                          T* __restrict__ wre, T* __restrict__ wim, int ld, int n,
                                          GjBuf<T, NMAX>& b, int tid) {
  constexpr int MAXR = (NMAX + RG - 1) / RG;
  static_assert(NMAX * RG <= NT, "not enough threads for the inversion layout");
  const bool on = tid < n * RG;
  const int j = on ? tid / RG : 0;
  const int g = tid % RG;
  T xr[MAXR], xi[MAXR];
  bool used[MAXR];
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int i = g + RG * m;
    used[m] = !(on && i < n);
    xr[m] = T(0);
    xi[m] = T(0);
    if (!used[m]) {
      xr[m] = re[i * ld + j];
      xi[m] = im[i * ld + j];
    }
  }
  // pivot of column `col` from the owners' registers; the winner publishes it
  auto choose = [&](int col, int par) {
    if (on && j == col) {
      T best = T(-1);
      int bidx = n;
      bool bad = false;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        if (!used[m]) {
          const T mg = fabs(xr[m]) + fabs(xi[m]);
          if (mg != mg) bad = true;
          if (mg > best) {
            best = mg;
            bidx = g + RG * m;
          }
        }
      }
      if (bad) best = T(INFINITY);  // a NaN anywhere in the group wins and marks it
      group_argmax<T, RG>(best, bidx);
      if (g == 0 || (bidx < n && g == bidx % RG)) {
        const int flag = isinf(best) ? ST_NONFINITE : (best > T(0) ? ST_OK : ST_SINGULAR);
        if (g == 0 && flag != ST_OK) {
          b.piv[par].flag = flag;
          b.piv[par].p = 0;
        }
        if (flag == ST_OK && g == bidx % RG) {
          const int mw = bidx / RG;
#pragma unroll
          for (int m = 0; m < MAXR; ++m)
            if (m == mw) crecip1(xr[m], xi[m], b.piv[par].ir, b.piv[par].ii);
          b.piv[par].p = bidx;
          b.piv[par].flag = ST_OK;
        }
      }
    }
  };
  choose(0, 0);
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const int par = k & 1;
    const T* cre = par ? wre : re;
    const T* cim = par ? wim : im;
    T* nre = par ? re : wre;
    T* nim = par ? im : wim;
    // column k of the current state does not depend on the pivot: load it
    // together with the pivot record, then the pivot row entry, and only then
    // store (the loads must not wait behind stores to the other buffer)
    T cr[MAXR], ci[MAXR];
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      const int i = g + RG * m;
      cr[m] = (on && i < n) ? cre[i * ld + k] : T(0);
      ci[m] = (on && i < n) ? cim[i * ld + k] : T(0);
    }
    const Pivot<T> pk = b.piv[par];
    if (pk.flag != ST_OK) {
      __syncthreads();
      return pk.flag;
    }
    const int p = pk.p;
    if (on) {
      const T pr = cre[p * ld + j], pi = cim[p * ld + j];
      const T sr = pr * pk.ir - pi * pk.ii, si = pr * pk.ii + pi * pk.ir;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        const int i = g + RG * m;
        if (i < n) {
          if (i == p) {
            used[m] = true;
            xr[m] = (j == k) ? pk.ir : sr;
            xi[m] = (j == k) ? pk.ii : si;
          } else if (j == k) {
            xr[m] = -(cr[m] * pk.ir - ci[m] * pk.ii);
            xi[m] = -(cr[m] * pk.ii + ci[m] * pk.ir);
          } else {
            T yr = xr[m], yi = xi[m];
            yr = fma(-cr[m], sr, yr);
            yr = fma(ci[m], si, yr);
            yi = fma(-cr[m], si, yi);
            yi = fma(-ci[m], sr, yi);
            xr[m] = yr;
            xi[m] = yi;
          }
        }
      }
      if (k + 1 < n) {
#pragma unroll
        for (int m = 0; m < MAXR; ++m) {
          const int i = g + RG * m;
          if (i < n) {
            nre[i * ld + j] = xr[m];
            nim[i * ld + j] = xi[m];
          }
        }
      }
    }
    if (tid == 0) {
      b.pv[k] = p;
      b.qv[p] = k;
    }
    if (k + 1 < n) choose(k + 1, par ^ 1);
    __syncthreads();
  }
  if (on) {
    const int pj = b.pv[j];
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      const int i = g + RG * m;
      if (i < n) {
        const int qi = b.qv[i];
        re[qi * ld + pj] = xr[m];
        im[qi * ld + pj] = xi[m];
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
