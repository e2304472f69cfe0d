// Workgroup-cooperative complex matrix primitives for gfx950 (CDNA4).
//
// Every matrix of one (omega, k) instance lives in LDS in planar form: a real
// plane and an imaginary plane, each NMAX rows of LD = NMAX + 1 entries.  The
// odd padding keeps column walks (A-operand reads) spread over the LDS banks.  A thread "owns" a TM x TM output tile of every product
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

// Two products sharing the A operand: C1 += A B1, C2 += A B2.
template <typename T, int TM>
__device__ __forceinline__ void mac2a(const T* are, const T* aim, const T* b1r, const T* b1i,
                                      const T* b2r, const T* b2i, int ld, int k, int i0, int j0,
                                      T (&c1r)[TM][TM], T (&c1i)[TM][TM], T (&c2r)[TM][TM],
                                      T (&c2i)[TM][TM]) {
#pragma unroll 2
  for (int kk = 0; kk < k; ++kk) {
    T xr[TM], xi[TM], ur[TM], ui[TM], vr[TM], vi[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      xr[a] = are[(i0 + a) * ld + kk];
      xi[a] = aim[(i0 + a) * ld + kk];
    }
#pragma unroll
    for (int b = 0; b < TM; ++b) {
      ur[b] = b1r[kk * ld + j0 + b];
      ui[b] = b1i[kk * ld + j0 + b];
      vr[b] = b2r[kk * ld + j0 + b];
      vi[b] = b2i[kk * ld + j0 + b];
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TM; ++b) {
        cmac(c1r[a][b], c1i[a][b], xr[a], xi[a], ur[b], ui[b]);
        cmac(c2r[a][b], c2i[a][b], xr[a], xi[a], vr[b], vi[b]);
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
enum Status { ST_OK = 0, ST_NOCONV = 1, ST_NONFINITE = 2, ST_SINGULAR = 3, ST_EIGEN_ZERO_ETA = 4,
              ST_REDO = 7 };

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
  int pv[NMAX];       // pivot row of step k
  int qv[NMAX];       // step at which row r was the pivot
  T colr[2][NMAX];    // column k of the current state, by step parity
  T coli[2][NMAX];
};

// Lane shuffle within a wave (ds_bpermute), 32- or 64-bit.
template <typename T>
__device__ __forceinline__ T lane_shfl(T v, int src_lane) {
  if constexpr (sizeof(T) == 8) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(u & 0xffffffffLL));
    const int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(u >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  } else {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
  }
}

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
// An empty side effect inside a wave-uniform branch: the compiler then keeps the
// branch (a scalar jump) instead of turning its body into selects on every lane.
__device__ __forceinline__ void uniform_branch() { asm volatile(""); }

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
// Thread t < NMAX * RG owns column j = t / RG and rows g, g + RG, ... (g = t % RG)
// in registers for the whole elimination; the matrix is read from LDS once and
// written back once.  A step needs only two things from other lanes: column k
// of the current state, which its owners publish into a parity-buffered column
// buffer, and the pivot row's entry of the lane's own column, which the lane
// g = p % RG of the same (wave-contained) column group holds and hands over by a
// lane shuffle.  The pivot search is an RG-lane DPP arg-max among the owners of
// the next column, so one barrier per step suffices.  Requires NMAX % RG == 0:
// every lane then works on rows and columns < NMAX of the padded matrix, and the
// padding (rows/columns >= n, marked used) never feeds a live entry.  The
// update is new = x - c s on every row, with s the scaled pivot row (s_k =
// 1 / pivot) and x = 0 in column k (its owners clear their registers when they
// publish it); the pivot row is then set to s.  Row p sits in one register slot
// (wave-uniform) of one lane, so that costs scalar branches, not selects on
// every row.  All
// threads must call it; it ends with a barrier.  Returns ST_OK, ST_SINGULAR or
// ST_NONFINITE.
//
// LS (lockstep): several sub-blocks of one workgroup invert their own matrices
// of the same n side by side, so every sub-block must pass the same n + 2
// barriers: a failed pivot then keeps stepping through the remaining barriers
// instead of returning early (gj_inverse_ls).
// The same elimination with a shorter pivot step, for the workgroup classes and the
// pooled class (at most twelve rows per lane, column groups of 4, 8 or 16 lanes).  A step's cost is the
// instruction stream of the wave that publishes the next column, so that stream is
// cut to the essentials:
//  - the pivot is the largest |re| + |im| compared to 2^-14 relative (2^-13 above 64
//    rows: the magnitude's high word with its low 6 (7) bits replaced by 63 - row
//    (127 - row), maximised with v_max_u32 over
//    the group by DPP; the first row among magnitudes equal to that resolution), and
//    its failure flag travels in the pivot-row word (-flag);
//  - the column and the pivot are read with one LDS wait;
//  - the permutation stays in registers: the owner lanes of column k keep its pivot
//    row, and the lane holding row p keeps the step at which it was the pivot, so the
//    final scatter needs no LDS tables;
//  - a failed pivot does not leave the loop: the step goes on with row 0, the status
//    of the first failure is returned and nothing is scattered, so every call passes
//    n + 2 barriers (what the lockstep form needs) and the loop keeps its registers in
//    place (an exit from it made the compiler copy the whole state every step).
// Partial pivoting within that tolerance is as stable as the exact maximum; results
// differ from gj_core's only where two candidates agree to 2^-14.
template <typename T, int NMAX>
struct GjLeanBuf {
  T colr[2][NMAX], coli[2][NMAX];  // column k of the current state, by step parity
  T pir[2], pii[2];                // reciprocal of the pivot
  int pp[2];                       // pivot row, or -flag when the pivot failed
};

template <typename T, int NT, int NMAX, int RG, bool LS, typename PT>
__device__ __forceinline__ int gj_lean(PT __restrict__ re, PT __restrict__ im, int ld, int n,
                                       GjLeanBuf<T, NMAX>& b, int tid) {
  constexpr int MAXR = NMAX / RG;
  static_assert(NMAX * RG == NT || NMAX * RG <= NT, "not enough threads for the inversion layout");
  static_assert(NMAX <= 128, "row index in 7 bits");
  // the key's low RB bits carry (2^RB - 1 - row); the magnitude keeps 20 - RB mantissa bits
  constexpr int RB = NMAX <= 64 ? 6 : 7;
  constexpr unsigned RMASK = (1u << RB) - 1u;
  const bool on = tid < NMAX * RG;
  const int jc = on ? tid / RG : 0;
  const int j = on ? jc : 0x7fff;
  const int g = tid % RG;
  const int lane = tid & (WAVE - 1);
  const int gbase = lane & ~(RG - 1);
  const int wave = tid / WAVE;
  T xr[MAXR], xi[MAXR];
  int qrow[MAXR];     // step at which row g + RG m was the pivot
  int pcol = 0;       // pivot row of the step of this lane's column
  unsigned used = 0;  // bit m: row g + RG m is a padding row or has been a pivot
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int i = g + RG * m;
    if (i >= n) used |= 1u << m;
    xr[m] = re[i * ld + jc];
    xi[m] = im[i * ld + jc];
    qrow[m] = 0;
  }
  // the owners of column `col` publish it and choose its pivot
  auto publish = [&](int col, int par) {
    if (wave != (col * RG) / WAVE) return;  // wave-uniform
    if (j == col) {
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        b.colr[par][g + RG * m] = xr[m];
        b.coli[par][g + RG * m] = xi[m];
      }
      unsigned key = 0;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        const double mg = (double)(fabs(xr[m]) + fabs(xi[m]));
        const unsigned kb = ((unsigned)((unsigned long long)__double_as_longlong(mg) >> 32) & ~RMASK) |
                            (unsigned)((int)RMASK - (g + RG * m));
        key = ((used >> m) & 1u) ? key : max(key, kb);
      }
      // (mov_dpp with bound_ctrl lets the compiler fold each move into v_max_u32_dpp)
      key = max(key, (unsigned)__builtin_amdgcn_mov_dpp((int)key, 0xB1, 0xF, 0xF, true));
      if constexpr (RG >= 4)
        key = max(key, (unsigned)__builtin_amdgcn_mov_dpp((int)key, 0x4E, 0xF, 0xF, true));
      if constexpr (RG >= 8)
        key = max(key, (unsigned)__builtin_amdgcn_mov_dpp((int)key, 0x141, 0xF, 0xF, true));
      if constexpr (RG >= 16)
        key = max(key, (unsigned)__builtin_amdgcn_mov_dpp((int)key, 0x140, 0xF, 0xF, true));
      const int bidx = (int)RMASK - (int)(key & RMASK);
      // NaN and infinity sort above every finite magnitude; zero (or subnormal) is singular
      const int flag = key >= 0x7FF00000u ? ST_NONFINITE : (key <= RMASK ? ST_SINGULAR : ST_OK);
      if (g == (bidx & (RG - 1))) {  // the lane holding the pivot row
        const int mw = bidx / RG;
        T pr = xr[0], pi = xi[0];
#pragma unroll
        for (int m = 1; m < MAXR; ++m) {
          pr = (m == mw) ? xr[m] : pr;
          pi = (m == mw) ? xi[m] : pi;
        }
        T ir, ii;
        crecip1(pr, pi, ir, ii);
        b.pir[par] = ir;
        b.pii[par] = ii;
        b.pp[par] = flag == ST_OK ? bidx : -flag;
      }
      // the published column becomes -c s in its step: it starts from zero
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        xr[m] = T(0);
        xi[m] = T(0);
      }
    }
  };
  publish(0, 0);
  __syncthreads();
  int fail = 0;  // first failed pivot's status (wave-uniform)
  for (int k = 0; k < n; ++k) {
    const int par = k & 1;
    T cr[MAXR], ci[MAXR];
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      cr[m] = b.colr[par][g + RG * m];
      ci[m] = b.coli[par][g + RG * m];
    }
    T pir = b.pir[par], pii = b.pii[par];
    int pp = b.pp[par];
    // one LDS round trip: every load is issued before the branch on the pivot
#pragma unroll
    for (int m = 0; m < MAXR; ++m) asm volatile("" : "+v"(cr[m]), "+v"(ci[m]));
    asm volatile("" : "+v"(pir), "+v"(pii), "+v"(pp));
    // a failed pivot is remembered and the elimination steps on with row 0 (its values
    // are discarded): no exit from the loop, whose state then stays in place
    const int pw = __builtin_amdgcn_readfirstlane(pp);
    fail = (fail == 0 && pw < 0) ? -pw : fail;
    const int p = pw < 0 ? 0 : pw;
    const int mp = p / RG, gp = p - (p / RG) * RG;
    T vr = xr[0], vi = xi[0];
#pragma unroll
    for (int m = 1; m < MAXR; ++m)
      if (m == mp) {
        uniform_branch();
        vr = xr[m];
        vi = xi[m];
      }
    // pivot row entry of this lane's column, from the group lane holding row p
    const T pr = lane_shfl(vr, gbase + gp);
    const T pi = lane_shfl(vi, gbase + gp);
    const bool jk = (j == k);
    const T sr = jk ? pir : pr * pir - pi * pii;
    const T si = jk ? pii : pr * pii + pi * pir;
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      T yr = fma(-cr[m], sr, xr[m]);
      yr = fma(ci[m], si, yr);
      T yi = fma(-cr[m], si, xi[m]);
      yi = fma(-ci[m], sr, yi);
      xr[m] = yr;
      xi[m] = yi;
    }
    // the pivot row becomes s (the update above left rounding residue there)
#pragma unroll
    for (int m = 0; m < MAXR; ++m)
      if (m == mp) {
        uniform_branch();
        const bool own = (g == gp);
        xr[m] = own ? sr : xr[m];
        xi[m] = own ? si : xi[m];
        used |= own ? (1u << m) : 0u;
        qrow[m] = own ? k : qrow[m];
      }
    pcol = jk ? p : pcol;
    if (k + 1 < n) publish(k + 1, par ^ 1);
    __syncthreads();
  }
  if (on && jc < n && fail == 0) {
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      const int i = g + RG * m;
      if (i < n) {
        re[qrow[m] * ld + pcol] = xr[m];
        im[qrow[m] * ld + pcol] = xi[m];
      }
    }
  }
  __syncthreads();
  return fail ? fail : ST_OK;
}

template <typename T, int NT, int NMAX, int RG, bool LS, typename PT, bool SF = true>
__device__ __forceinline__ int gj_core(PT __restrict__ re, PT __restrict__ im, int ld, int n,
                                       GjBuf<T, NMAX>& b, int tid) {
  static_assert(NMAX % RG == 0, "row groups must tile NMAX");
  static_assert(WAVE % RG == 0, "a column group must sit inside one wave");
  constexpr int MAXR = NMAX / RG;
  static_assert(NMAX * RG <= NT, "not enough threads for the inversion layout");
  static_assert(MAXR <= 32, "row mask is 32 bits");
  // (the pooled class, twelve rows per lane, takes the short step too since its loop
  // has no exit: its decimation spills 44 VGPRs there against 73 in this step)
  if constexpr (SF && MAXR <= 12 && (RG == 4 || RG == 8 || RG == 16) && NMAX <= 128 &&
                sizeof(GjLeanBuf<T, NMAX>) <= sizeof(GjBuf<T, NMAX>)) {
    return gj_lean<T, NT, NMAX, RG, LS>(re, im, ld, n,
                                       *reinterpret_cast<GjLeanBuf<T, NMAX>*>(&b), tid);
  }
  const bool on = tid < NMAX * RG;
  const int j = on ? tid / RG : 0;
  const int g = tid % RG;
  const int gbase = (tid & (WAVE - 1)) & ~(RG - 1);
  T xr[MAXR], xi[MAXR];
  unsigned used = 0;  // bit m: row g + RG m is a padding row or has been a pivot
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int i = g + RG * m;
    if (i >= n) used |= 1u << m;
    xr[m] = re[i * ld + j];
    xi[m] = im[i * ld + j];
  }
  // the owners of column `col` publish it and choose its pivot
  auto publish = [&](int col, int par) {
    if (on && j == col) {
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        b.colr[par][g + RG * m] = xr[m];
        b.coli[par][g + RG * m] = xi[m];
      }
      T best = T(-1);
      int bidx = n;
      bool bad = false;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        const T mg = fabs(xr[m]) + fabs(xi[m]);
        const bool live = !((used >> m) & 1u);
        bad = bad || (live && mg != mg);
        const bool better = live && mg > best;
        best = better ? mg : best;
        bidx = better ? g + RG * m : bidx;
      }
      if (bad) best = T(INFINITY);  // a NaN anywhere in the column wins and marks it
      group_argmax<T, RG>(best, bidx);
      if (g == 0 || (bidx < n && g == bidx % RG)) {
        const int flag = isinf(best) ? ST_NONFINITE : (best > T(0) ? ST_OK : ST_SINGULAR);
        if (g == 0 && flag != ST_OK) {
          b.piv[par].flag = flag;
          b.piv[par].p = 0;
        }
        if (flag == ST_OK && g == bidx % RG) {
          const int mw = bidx / RG;
          T pr = xr[0], pi = xi[0];
#pragma unroll
          for (int m = 1; m < MAXR; ++m) {
            pr = (m == mw) ? xr[m] : pr;
            pi = (m == mw) ? xi[m] : pi;
          }
          crecip1(pr, pi, b.piv[par].ir, b.piv[par].ii);
          b.piv[par].p = bidx;
          b.piv[par].flag = ST_OK;
        }
      }
      // the published column becomes -c s in its step: it starts from zero
      if constexpr (SF) {
#pragma unroll
        for (int m = 0; m < MAXR; ++m) {
          xr[m] = T(0);
          xi[m] = T(0);
        }
      }
    }
  };
  publish(0, 0);
  __syncthreads();
  int fail = 0;  // first failed pivot's status (wave-uniform)
  for (int k = 0; k < n; ++k) {
    const int par = k & 1;
    T cr[MAXR], ci[MAXR];
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      cr[m] = b.colr[par][g + RG * m];
      ci[m] = b.coli[par][g + RG * m];
    }
    const Pivot<T> pk = b.piv[par];
    // a failed pivot (published with p = 0) is remembered and the step goes on; the
    // values are discarded (no exit from the loop, see gj_lean)
    const int flag = __builtin_amdgcn_readfirstlane(pk.flag);
    fail = (fail == ST_OK && flag != ST_OK) ? flag : fail;
    const int p = flag != ST_OK ? 0 : __builtin_amdgcn_readfirstlane(pk.p);
    if constexpr (SF) {
      // row p sits in register slot mp of the group lane gp: both are wave-uniform, so
      // the slot is chosen by scalar branches (uniform_branch keeps them branches
      // rather than per-row selects)
      const int mp = p / RG, gp = p - (p / RG) * RG;
      T vr = xr[0], vi = xi[0];
#pragma unroll
      for (int m = 1; m < MAXR; ++m)
        if (m == mp) {
          uniform_branch();
          vr = xr[m];
          vi = xi[m];
        }
      // pivot row entry of this lane's column, from the group lane holding row p
      const T pr = lane_shfl(vr, gbase + gp);
      const T pi = lane_shfl(vi, gbase + gp);
      const bool jk = (j == k);
      const T sr = jk ? pk.ir : pr * pk.ir - pi * pk.ii;
      const T si = jk ? pk.ii : pr * pk.ii + pi * pk.ir;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        T yr = fma(-cr[m], sr, xr[m]);
        yr = fma(ci[m], si, yr);
        T yi = fma(-cr[m], si, xi[m]);
        yi = fma(-ci[m], sr, yi);
        xr[m] = yr;
        xi[m] = yi;
      }
      // the pivot row becomes s (the update above left rounding residue there)
#pragma unroll
      for (int m = 0; m < MAXR; ++m)
        if (m == mp) {
          uniform_branch();
          const bool own = (g == gp);
          xr[m] = own ? sr : xr[m];
          xi[m] = own ? si : xi[m];
          used |= own ? (1u << m) : 0u;
        }
    } else {
      // the per-row select form: c = -1 and x' = 0 on the pivot row, x' = 0 in column k
      const int mp = p / RG;
      T vr = xr[0], vi = xi[0];
#pragma unroll
      for (int m = 1; m < MAXR; ++m) {
        vr = (m == mp) ? xr[m] : vr;
        vi = (m == mp) ? xi[m] : vi;
      }
      const T pr = lane_shfl(vr, gbase + (p % RG));
      const T pi = lane_shfl(vi, gbase + (p % RG));
      const bool jk = (j == k);
      const T sr = jk ? pk.ir : pr * pk.ir - pi * pk.ii;
      const T si = jk ? pk.ii : pr * pk.ii + pi * pk.ir;
#pragma unroll
      for (int m = 0; m < MAXR; ++m) {
        const bool isp = (g + RG * m == p);
        const T c_r = isp ? T(-1) : cr[m];
        const T c_i = isp ? T(0) : ci[m];
        T yr = (isp || jk) ? T(0) : xr[m];
        T yi = (isp || jk) ? T(0) : xi[m];
        yr = fma(-c_r, sr, yr);
        yr = fma(c_i, si, yr);
        yi = fma(-c_r, si, yi);
        yi = fma(-c_i, sr, yi);
        xr[m] = yr;
        xi[m] = yi;
        used |= isp ? (1u << m) : 0u;
      }
    }
    if (tid == 0) {
      b.pv[k] = p;
      b.qv[p] = k;
    }
    if (k + 1 < n) publish(k + 1, par ^ 1);
    __syncthreads();
  }
  if (on && j < n && fail == ST_OK) {
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
  return fail;
}

template <typename T, int NT, int NMAX, int RG, typename PT>
__device__ __forceinline__ int gj_inverse(PT __restrict__ re, PT __restrict__ im, int ld, int n,
                                          GjBuf<T, NMAX>& b, int tid) {
  return gj_core<T, NT, NMAX, RG, false>(re, im, ld, n, b, tid);
}

// The per-row select form of the same step (same results on the matrices where both
// were compared), kept for the eigenmode route's spectral-projector kernels: a
// 90-DOF case of that route stopped converging when they took the select-free step,
// for a reason not found.
template <typename T, int NT, int NMAX, int RG, typename PT>
__device__ __forceinline__ int gj_inverse_sel(PT __restrict__ re, PT __restrict__ im, int ld, int n,
                                              GjBuf<T, NMAX>& b, int tid) {
  return gj_core<T, NT, NMAX, RG, false, PT, false>(re, im, ld, n, b, tid);
}

// Lockstep form: a sub-block with live == false (finished, failed or without an
// instance) passes the n + 2 barriers of an inversion and touches nothing.
template <typename T, int NT, int NMAX, int RG, typename PT>
__device__ __forceinline__ int gj_inverse_ls(PT __restrict__ re, PT __restrict__ im, int ld, int n,
                                             GjBuf<T, NMAX>& b, int tid, bool live) {
  if (!live) {
    for (int r = 0; r < n + 2; ++r) __syncthreads();
    return ST_SINGULAR;
  }
  return gj_core<T, NT, NMAX, RG, true>(re, im, ld, n, b, tid);
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
