// Block-size class 96 (FP64): principal layers of 49..96 DOF.
//
// Three 96 x 97 planar complex matrices already need 447 KB, more than a CU's
// 160 KB of LDS, so in this class the matrices of an instance live in a slot of a
// device-memory pool (one slot per resident workgroup, reused instance after
// instance; the working set of the resident slots stays in the L2 / MALL).  The
// inversion is the same register-resident Gauss-Jordan as the LDS classes (its
// pivot buffers stay in LDS).  The decimation and the sweep run their products on
// the matrix cores with the operands streamed from the slot (big_mm); the
// self-energy, the trace and the eigenmode route keep the owner-tile loops, their
// operands staged through LDS in K chunks (big_mul).
// The kernels are persistent: the grid is at most the number of slots and each
// workgroup walks the instances blockIdx.x, blockIdx.x + gridDim.x, ...
//
// Each kernel follows its LDS counterpart step for step (same products, same
// order of the additions), so results agree with the LDS classes to rounding.

#pragma once

#include "cmat.hpp"
#include "mm.hpp"

namespace negf {

constexpr int BIG = 96;
constexpr int BIG_NT = 768;
constexpr int BIG_TM = 4;                         // (96 / 4)^2 = 576 owner tiles
constexpr int BIG_RG = 8;                         // inversion: 96 columns x 8 lanes
constexpr int BIG_LD = BIG + 1;
constexpr size_t BIG_PLANE = (size_t)BIG * BIG_LD;
constexpr size_t BIG_MAT = 2 * BIG_PLANE;         // doubles per planar matrix
constexpr int BIG_SLOT_MATS = 9;                  // the decimation needs the most
constexpr size_t BIG_SLOT = BIG_SLOT_MATS * BIG_MAT;

constexpr int BIG_KT = 16;                        // K chunk of the products staged in LDS

struct BigShared {
  GjBuf<double, BIG> gj;
  double sred[BIG_NT / WAVE];
  int sflag[BIG_NT / WAVE];
  // op(A)[:, k0:k0+KT] as a[kk][i] and op(B)[k0:k0+KT, :] as b[kk][j]
  double ar[BIG_KT][BIG], ai[BIG_KT][BIG];
  double br[BIG_KT][BIG], bi[BIG_KT][BIG];
};

// Slot matrices are addressed as global memory with 32-bit element offsets, so a
// load or store is one scalar base plus one vector offset (no 64-bit address per
// element kept live across the loops).
using gdouble = __attribute__((address_space(1))) double;

struct BigMat {
  gdouble* re;
  gdouble* im;
};

__device__ __forceinline__ BigMat big_mat(double* slot, int q) {
  gdouble* p = (gdouble*)(slot) + (size_t)q * BIG_MAT;
  return BigMat{p, p + BIG_PLANE};
}

__device__ __forceinline__ unsigned big_off(int i, int j) { return (unsigned)(i * BIG_LD + j); }

// The slot base as an opaque (uniform) value: matrix addresses derived from it are
// formed where they are used instead of being hoisted out of the instance and
// iteration loops and kept live in registers.
__device__ __forceinline__ double* big_fresh(double* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// A-operand element (i, kk) of op(A) and B-operand element (kk, j) of op(B)
template <int OP>
__device__ __forceinline__ void big_lda(const BigMat& A, int i, int kk, double& r, double& m) {
  if constexpr (OP == OPN) {
    r = A.re[big_off(i, kk)];
    m = A.im[big_off(i, kk)];
  } else if constexpr (OP == OPH) {
    r = A.re[big_off(kk, i)];
    m = -A.im[big_off(kk, i)];
  } else if constexpr (OP == OPC) {
    r = A.re[big_off(i, kk)];
    m = -A.im[big_off(i, kk)];
  } else {
    r = A.re[big_off(kk, i)];
    m = A.im[big_off(kk, i)];
  }
}
template <int OP>
__device__ __forceinline__ void big_ldb(const BigMat& B, int kk, int j, double& r, double& m) {
  if constexpr (OP == OPN) {
    r = B.re[big_off(kk, j)];
    m = B.im[big_off(kk, j)];
  } else if constexpr (OP == OPH) {
    r = B.re[big_off(j, kk)];
    m = -B.im[big_off(j, kk)];
  } else if constexpr (OP == OPC) {
    r = B.re[big_off(kk, j)];
    m = -B.im[big_off(kk, j)];
  } else {
    r = B.re[big_off(j, kk)];
    m = B.im[big_off(j, kk)];
  }
}

// m x p block (complex128, row-major) into a slot matrix, and a slot matrix out
__device__ __forceinline__ void big_gload(const double* __restrict__ src, int m, int p,
                                          const BigMat& D, int tid) {
  const double2* s = reinterpret_cast<const double2*>(src);
  const int mp = m * p;
  for (int e = tid; e < mp; e += BIG_NT) {
    const int i = e / p, j = e - (e / p) * p;
    const double2 v = s[e];
    D.re[big_off(i, j)] = v.x;
    D.im[big_off(i, j)] = v.y;
  }
}
__device__ __forceinline__ void big_gstore(double* __restrict__ dst, int n, const BigMat& S, int tid) {
  double2* d = reinterpret_cast<double2*>(dst);
  const int nn = n * n;
  for (int e = tid; e < nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    d[e] = make_double2(S.re[big_off(i, j)], S.im[big_off(i, j)]);
  }
}

using BigTile = Tile<BIG_TM>;

// C_tile = op(A)[0:m, 0:k] op(B)[0:k, 0:p] for the owner tiles of an m x p output.
// K is walked in chunks of BIG_KT staged in LDS by the whole workgroup (every
// thread must call); the accumulation order over k is the plain k = 0, 1, ... of
// the LDS classes.  Ends with a barrier, so the operands may be overwritten next.
template <int OPA, int OPB>
__device__ __forceinline__ void big_mul(const BigMat& A, const BigMat& B, int m, int k, int p,
                                        const BigTile& t, double (&cr)[BIG_TM][BIG_TM],
                                        double (&ci)[BIG_TM][BIG_TM], BigShared& sh, int tid) {
  tzero<double, BIG_TM>(cr, ci);
  for (int k0 = 0; k0 < k; k0 += BIG_KT) {
    const int kt = (k - k0) < BIG_KT ? (k - k0) : BIG_KT;
#pragma unroll 1
    for (int e = tid; e < BIG_KT * BIG; e += BIG_NT) {
      const int kk = e / BIG, i = e - (e / BIG) * BIG;
      double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
      if (kk < kt) {
        if (i < m) big_lda<OPA>(A, i, k0 + kk, xr, xi);
        if (i < p) big_ldb<OPB>(B, k0 + kk, i, yr, yi);
      }
      sh.ar[kk][i] = xr;
      sh.ai[kk][i] = xi;
      sh.br[kk][i] = yr;
      sh.bi[kk][i] = yi;
    }
    __syncthreads();
    if (t.own) {
#pragma unroll 1
      for (int kk = 0; kk < kt; ++kk) {
        double xr[BIG_TM], xi[BIG_TM], yr[BIG_TM], yi[BIG_TM];
#pragma unroll
        for (int q = 0; q < BIG_TM; ++q) {
          xr[q] = sh.ar[kk][t.i0 + q];
          xi[q] = sh.ai[kk][t.i0 + q];
          yr[q] = sh.br[kk][t.j0 + q];
          yi[q] = sh.bi[kk][t.j0 + q];
        }
#pragma unroll
        for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
          for (int y = 0; y < BIG_TM; ++y) cmac(cr[x][y], ci[x][y], xr[x], xi[x], yr[y], yi[y]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// products on the matrix cores (v_mfma_f64_16x16x4f64; 96 = 6 x 16 output tiles)
// ---------------------------------------------------------------------------
//
// The 36 output tiles of 16 x 16 go three to a wave: wave w (of 12) owns the tiles
// (R, C) = (w / 2, 3 (w % 2) + q), q < 3, so its A fragment is shared by its three
// tiles.  Lane l of the wave loads A[16 R + l % 16][4 s + l / 16] and, per tile,
// B[4 s + l / 16][16 C + l % 16] straight from the slot (L2) at K step s -- no LDS
// staging and no barrier inside a product -- the fragments of step s + 1 in flight
// while step s runs on the matrix cores.  Lane l holds element
// (16 R + 4 v + l / 16, 16 C + l % 16) of tile q in accumulator register v.
// Operand entries outside [0, m) x [0, k) and [0, k) x [0, p) are read as zero.
// A complex product is four real chains per tile, the minus of Ai Bi (and the
// conjugation of an OPH / OPC operand) in the negation bits of the instruction.

typedef double big_v4d __attribute__((ext_vector_type(4)));
constexpr int BIG_TPW = 3;  // output tiles per wave

struct BigOwn {
  int R, C0, l;
  __device__ __forceinline__ explicit BigOwn(int tid) {
    const int w = tid / WAVE;
    l = tid & (WAVE - 1);
    R = w >> 1;
    C0 = 3 * (w & 1);
  }
  __device__ __forceinline__ int row(int v) const { return 16 * R + 4 * v + (l >> 4); }
  __device__ __forceinline__ int col(int q) const { return 16 * (C0 + q) + (l & 15); }
};

template <int NEG>
__device__ __forceinline__ big_v4d bmf(double a, double b, big_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, NEG);
}

template <int OP>
__device__ __forceinline__ void big_frag(const BigMat& X, int r, int c, bool ok, double& xr,
                                         double& xi) {
  xr = 0.0;
  xi = 0.0;
  if (ok) {
    const unsigned o = (OP == OPN || OP == OPC) ? big_off(r, c) : big_off(c, r);
    xr = X.re[o];
    xi = X.im[o];
  }
}

template <int OPA, int OPB>
__device__ __forceinline__ void big_mm(const BigMat& A, const BigMat& B, int m, int k, int p,
                                       const BigOwn& o, big_v4d (&cr)[BIG_TPW],
                                       big_v4d (&ci)[BIG_TPW]) {
  constexpr int sa = imsign<OPA>(), sb = imsign<OPB>();
  const big_v4d z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < BIG_TPW; ++q) {
    cr[q] = z;
    ci[q] = z;
  }
  const int ar = 16 * o.R + (o.l & 15);
  const int kl = o.l >> 4;
  const int ns = (k + 3) / 4;
  double xr, xi, yr[BIG_TPW], yi[BIG_TPW];
  big_frag<OPA>(A, ar, kl, ar < m && kl < k, xr, xi);
#pragma unroll
  for (int q = 0; q < BIG_TPW; ++q) big_frag<OPB>(B, kl, o.col(q), kl < k && o.col(q) < p, yr[q], yi[q]);
  for (int s = 0; s < ns; ++s) {
    // the next step's fragments, in flight while this step's chains run
    const int kn = 4 * (s + 1) + kl;
    double nxr, nxi, nyr[BIG_TPW], nyi[BIG_TPW];
    big_frag<OPA>(A, ar, kn, s + 1 < ns && ar < m && kn < k, nxr, nxi);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q)
      big_frag<OPB>(B, kn, o.col(q), s + 1 < ns && kn < k && o.col(q) < p, nyr[q], nyi[q]);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q) {
      cr[q] = bmf<0>(xr, yr[q], cr[q]);
      ci[q] = bmf<(sb < 0) ? 1 : 0>(xr, yi[q], ci[q]);
      cr[q] = bmf<(sa * sb > 0) ? 1 : 0>(xi, yi[q], cr[q]);
      ci[q] = bmf<(sa < 0) ? 1 : 0>(xi, yr[q], ci[q]);
    }
    xr = nxr;
    xi = nxi;
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q) {
      yr[q] = nyr[q];
      yi[q] = nyi[q];
    }
  }
}

// owned elements inside [0, m) x [0, p) into a slot matrix
__device__ __forceinline__ void big_mstore(const BigMat& C, int m, int p, const BigOwn& o,
                                           const big_v4d (&cr)[BIG_TPW],
                                           const big_v4d (&ci)[BIG_TPW]) {
#pragma unroll
  for (int q = 0; q < BIG_TPW; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = o.row(v), j = o.col(q);
      if (i < m && j < p) {
        C.re[big_off(i, j)] = cr[q][v];
        C.im[big_off(i, j)] = ci[q][v];
      }
    }
}

__device__ __forceinline__ void big_tstore(const BigMat& C, int m, int p, const BigTile& t,
                                           const double (&cr)[BIG_TM][BIG_TM],
                                           const double (&ci)[BIG_TM][BIG_TM]) {
  if (!t.own) return;
#pragma unroll
  for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
    for (int y = 0; y < BIG_TM; ++y) {
      const int i = t.i0 + x, j = t.j0 + y;
      if (i < m && j < p) {
        C.re[big_off(i, j)] = cr[x][y];
        C.im[big_off(i, j)] = ci[x][y];
      }
    }
}

__device__ __forceinline__ void big_store(const BigMat& C, int n, const BigTile& t,
                                          const double (&cr)[BIG_TM][BIG_TM],
                                          const double (&ci)[BIG_TM][BIG_TM]) {
  big_tstore(C, n, n, t, cr, ci);
}

// ---------------------------------------------------------------------------
// decimation (surface_gf.py:18-81), slots: a, b, a', b', e, e_s, g, T1 = a g, T2 = b g
// ---------------------------------------------------------------------------

__device__ __forceinline__ void big_sr_one(const SrArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  if (a.redo && a.status[inst] == ST_OK) return;
  const int n = a.n;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;
  int qa = 0, qb = 1;  // slots of alpha and beta; the next ones go to 2 - qa, 4 - qb... swapped
  {
    double* sl = big_fresh(slot);
    big_gload(blk + nn * 2, n, n, big_mat(sl, 0), tid);
    big_gload(blk + 2 * nn * 2, n, n, big_mat(sl, 1), tid);
    big_gload(blk, n, n, big_mat(sl, 4), tid);
    big_gload(blk, n, n, big_mat(sl, 5), tid);
  }
  double zr, zi;
  zsq(a.omega[inst], a.eta[inst], zr, zi);
  const BigOwn o(tid);
  __syncthreads();
  int status = ST_OK;
  int it = 0;
  bool conv = false;
  while (it < a.maxiter) {
    ++it;
    double* sl = big_fresh(slot);
    const BigMat A = big_mat(sl, qa), B = big_mat(sl, qb), An = big_mat(sl, qa ^ 2),
                 Bn = big_mat(sl, qb ^ 2);
    const BigMat E = big_mat(sl, 4), ES = big_mat(sl, 5), G = big_mat(sl, 6);
    const BigMat T1 = big_mat(sl, 7), T2 = big_mat(sl, 8);
    for (int e = tid; e < (int)nn; e += BIG_NT) {
      const int i = e / n, j = e - (e / n) * n;
      G.re[big_off(i, j)] = (i == j ? zr : 0.0) - E.re[big_off(i, j)];
      G.im[big_off(i, j)] = (i == j ? zi : 0.0) - E.im[big_off(i, j)];
    }
    __syncthreads();
    const int st = gj_inverse<double, BIG_NT, BIG, BIG_RG>(G.re, G.im, BIG_LD, n, sh.gj, tid);
    if (st != ST_OK) {
      status = st;
      break;
    }
    big_v4d cr[BIG_TPW], ci[BIG_TPW];
    big_mm<OPN, OPN>(A, G, n, n, n, o, cr, ci);
    big_mstore(T1, n, n, o, cr, ci);
    big_mm<OPN, OPN>(B, G, n, n, n, o, cr, ci);
    big_mstore(T2, n, n, o, cr, ci);
    __syncthreads();
    bool bad = false;
    double emax = 0.0;
    // P1 = T1 b into e_s and e; P3 = T1 a is the next alpha
    big_mm<OPN, OPN>(T1, B, n, n, n, o, cr, ci);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = o.row(v), j = o.col(q);
        if (i < n && j < n) {
          const unsigned off = big_off(i, j);
          ES.re[off] = ES.re[off] + cr[q][v];
          ES.im[off] = ES.im[off] + ci[q][v];
          E.re[off] = E.re[off] + cr[q][v];
          E.im[off] = E.im[off] + ci[q][v];
          bad |= !(isfinite(ES.re[off]) && isfinite(ES.im[off]));
        }
      }
    big_mm<OPN, OPN>(T1, A, n, n, n, o, cr, ci);
    big_mstore(An, n, n, o, cr, ci);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (o.row(v) < n && o.col(q) < n) {
          bad |= !(isfinite(cr[q][v]) && isfinite(ci[q][v]));
          emax = fmax(emax, cr[q][v] * cr[q][v] + ci[q][v] * ci[q][v]);
        }
    // P2 = T2 a into e; P4 = T2 b is the next beta
    big_mm<OPN, OPN>(T2, A, n, n, n, o, cr, ci);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = o.row(v), j = o.col(q);
        if (i < n && j < n) {
          const unsigned off = big_off(i, j);
          E.re[off] = E.re[off] + cr[q][v];
          E.im[off] = E.im[off] + ci[q][v];
          bad |= !(isfinite(E.re[off]) && isfinite(E.im[off]));
        }
      }
    big_mm<OPN, OPN>(T2, B, n, n, n, o, cr, ci);
    big_mstore(Bn, n, n, o, cr, ci);
#pragma unroll
    for (int q = 0; q < BIG_TPW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (o.row(v) < n && o.col(q) < n) {
          bad |= !(isfinite(cr[q][v]) && isfinite(ci[q][v]));
          emax = fmax(emax, cr[q][v] * cr[q][v] + ci[q][v] * ci[q][v]);
        }
    double gmax;
    bool gbad;
    block_max_any<BIG_NT>(emax, bad, sh.sred, sh.sflag, tid, gmax, gbad);
    qa ^= 2;
    qb ^= 2;
    if (gbad) {
      status = ST_NONFINITE;
      break;
    }
    if (sqrt(gmax) < a.tol) {
      conv = true;
      break;
    }
  }
  if (status == ST_OK && !conv) status = ST_NOCONV;
  if (tid == 0) {
    a.status[inst] = status;
    a.iters[inst] = it;
  }
  if (status != ST_OK) return;
  double* sl = big_fresh(slot);
  const BigMat ES = big_mat(sl, 5), G = big_mat(sl, 6);
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    G.re[big_off(i, j)] = (i == j ? zr : 0.0) - ES.re[big_off(i, j)];
    G.im[big_off(i, j)] = (i == j ? zi : 0.0) - ES.im[big_off(i, j)];
  }
  __syncthreads();
  const int st = gj_inverse<double, BIG_NT, BIG, BIG_RG>(G.re, G.im, BIG_LD, n, sh.gj, tid);
  if (st != ST_OK) {
    if (tid == 0) a.status[inst] = st;
    return;
  }
  big_gstore(a.g_out + (size_t)inst * nn * 2, n, G, tid);
  if (tid == 0 && a.dos_out) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += -2.0 * G.im[big_off(i, i)];
    a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
  }
}

__global__ void __launch_bounds__(BIG_NT) k_sr_big(SrArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * BIG_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    big_sr_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// self-energy Sigma = V g V^H embedded in the end layer (negf.py:72-97)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void big_sigma_one(const SigmaArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  const int nl = a.n_lead, nc = a.n_c, nd = a.n_dev;
  double2* out = reinterpret_cast<double2*>(a.sigma_out + (size_t)inst * nd * nd * 2);
  if (a.status_in && a.status_in[inst] != ST_OK) return;
  double* sl = big_fresh(slot);
  const BigMat V = big_mat(sl, 0), Gm = big_mat(sl, 1), X = big_mat(sl, 2), S = big_mat(sl, 3);
  big_gload(a.v + (size_t)a.set_of[inst] * nc * nl * 2, nc, nl, V, tid);
  big_gload(a.g + (size_t)inst * nl * nl * 2, nl, nl, Gm, tid);
  int* pos = sh.gj.pv;
  for (int i = tid; i < nd; i += BIG_NT) pos[i] = -1;
  __syncthreads();
  for (int c = tid; c < nc; c += BIG_NT) pos[a.idx[c]] = c;
  __syncthreads();
  double xr[BIG_TM][BIG_TM], xi[BIG_TM][BIG_TM];
  {
    const BigTile t(nc, nl, tid);
    big_mul<OPN, OPN>(V, Gm, nc, nl, nl, t, xr, xi, sh, tid);
    big_tstore(X, nc, nl, t, xr, xi);
  }
  __syncthreads();
  {
    const BigTile t(nc, nc, tid);
    big_mul<OPN, OPH>(X, V, nc, nl, nc, t, xr, xi, sh, tid);
    big_tstore(S, nc, nc, t, xr, xi);
  }
  __syncthreads();
  for (int e = tid; e < nd * nd; e += BIG_NT) {
    const int i = e / nd, j = e - (e / nd) * nd;
    const int pi = pos[i], pj = pos[j];
    out[e] = (pi >= 0 && pj >= 0) ? make_double2(S.re[big_off(pi, pj)], S.im[big_off(pi, pj)])
                                  : make_double2(0.0, 0.0);
  }
}

__global__ void __launch_bounds__(BIG_NT) k_sigma_big(SigmaArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * BIG_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    big_sigma_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// recursive Green's function sweep for G_1N; slots: g (then Y), H, Q, S = H^H Y
// ---------------------------------------------------------------------------

__device__ __forceinline__ void big_rgf_one(const RgfArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  if (a.redo && a.status[inst] != 10 + ST_NONFINITE && a.status[inst] != 10 + ST_SINGULAR) return;
  if (a.status_in && a.status_in[inst] != ST_OK) {
    if (tid == 0) a.status[inst] = a.status_in[inst];
    return;
  }
  const int n = a.n, nl = a.n_layers;
  const size_t nn = (size_t)n * n;
  const size_t set = (size_t)a.set_of[inst];
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  const BigOwn o(tid);
  double zr, zi;
  zsq(a.omega[inst], a.eta_dev[inst], zr, zi);
  int status = ST_OK;
  for (int layer = 0; layer < nl; ++layer) {
    double* sl = big_fresh(slot);
    const BigMat Gm = big_mat(sl, 0), H = big_mat(sl, 1), Q = big_mat(sl, 2), S = big_mat(sl, 3);
    const double2* d = ons + (size_t)layer * nn;
    for (int e = tid; e < (int)nn; e += BIG_NT) {
      const int i = e / n, j = e - (e / n) * n;
      const unsigned o = big_off(i, j);
      const double2 dv = d[e];
      double vr = (i == j ? zr : 0.0) - dv.x;
      double vi = (i == j ? zi : 0.0) - dv.y;
      if (layer > 0) {
        vr -= S.re[o];
        vi -= S.im[o];
      } else {
        vr -= 0.0;
        vi -= 0.0;
      }
      if (layer == 0) {
        vr -= sgl[e].x;
        vi -= sgl[e].y;
      }
      if (layer == nl - 1) {
        vr -= sgr[e].x;
        vi -= sgr[e].y;
      }
      Gm.re[o] = vr;
      Gm.im[o] = vi;
    }
    __syncthreads();
    const int st = gj_inverse<double, BIG_NT, BIG, BIG_RG>(Gm.re, Gm.im, BIG_LD, n, sh.gj, tid);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    big_v4d cr[BIG_TPW], ci[BIG_TPW];
    if (layer == nl - 1) {
      // G_1N = Q_{N-1} g_N (g_1 itself for one layer)
      if (nl == 1) {
        big_gstore(a.g1n_out + (size_t)inst * nn * 2, n, Gm, tid);
      } else {
        big_mm<OPN, OPN>(Q, Gm, n, n, n, o, cr, ci);
        big_mstore(H, n, n, o, cr, ci);
        __syncthreads();
        big_gstore(a.g1n_out + (size_t)inst * nn * 2, n, H, tid);
      }
      break;
    }
    big_gload(cpl + (size_t)layer * nn * 2, n, n, H, tid);
    __syncthreads();
    big_mm<OPN, OPN>(Gm, H, n, n, n, o, cr, ci);
    __syncthreads();  // every read of g is done: Y replaces it (and starts Q at layer 0)
    big_mstore(Gm, n, n, o, cr, ci);
    if (layer == 0) big_mstore(Q, n, n, o, cr, ci);
    __syncthreads();
    big_mm<OPH, OPN>(H, Gm, n, n, n, o, cr, ci);
    big_mstore(S, n, n, o, cr, ci);
    if (layer > 0) {
      big_mm<OPN, OPN>(Q, Gm, n, n, n, o, cr, ci);
      __syncthreads();
      big_mstore(Q, n, n, o, cr, ci);
    }
    __syncthreads();
  }
  if (tid == 0) a.status[inst] = status;
}

__global__ void __launch_bounds__(BIG_NT) k_rgf_big(RgfArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * BIG_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    big_rgf_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// T = Re Tr[Gamma_L G Gamma_R G^H] (negf.py:115-116, :356-357); slots: Gamma_L,
// Gamma_R, G, P = Gamma_L G
// ---------------------------------------------------------------------------

__device__ __forceinline__ void big_trace_one(const TraceArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  if (a.status_in && a.status_in[inst] != ST_OK) {
    if (tid == 0) a.t_out[inst] = NAN;
    return;
  }
  const int n = a.n;
  const int nn = n * n;
  const double2* sl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  double* sb = big_fresh(slot);
  const BigMat GL = big_mat(sb, 0), GR = big_mat(sb, 1), Gm = big_mat(sb, 2), P = big_mat(sb, 3);
  for (int e = tid; e < nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 a1 = sl[i * n + j], a2 = sl[j * n + i];
    GL.re[big_off(i, j)] = -(a1.y + a2.y);
    GL.im[big_off(i, j)] = a1.x - a2.x;
    const double2 b1 = sr[i * n + j], b2 = sr[j * n + i];
    GR.re[big_off(i, j)] = -(b1.y + b2.y);
    GR.im[big_off(i, j)] = b1.x - b2.x;
  }
  big_gload(a.g1n + (size_t)inst * nn * 2, n, n, Gm, tid);
  __syncthreads();
  const BigTile t(n, n, tid);
  double cr[BIG_TM][BIG_TM], ci[BIG_TM][BIG_TM];
  // P = Gamma_L G (kept by the owner, read back by the same thread)
  big_mul<OPN, OPN>(GL, Gm, n, n, n, t, cr, ci, sh, tid);
  big_store(P, n, t, cr, ci);
  // Q_ij = (Gamma_R G^H)_ji = sum_k conj(G_ik) Gamma_R_jk
  big_mul<OPC, OPT>(Gm, GR, n, n, n, t, cr, ci, sh, tid);
  double acc = 0.0;
  if (t.own) {
#pragma unroll
    for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
      for (int y = 0; y < BIG_TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) acc += P.re[big_off(i, j)] * cr[x][y] - P.im[big_off(i, j)] * ci[x][y];
      }
  }
  const double tot = block_sum<BIG_NT>(acc, sh.sred, tid);
  if (tid == 0) a.t_out[inst] = tot;
}

__global__ void __launch_bounds__(BIG_NT) k_trace_big(TraceArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * BIG_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    big_trace_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

}  // namespace negf

namespace negf {

// ---------------------------------------------------------------------------
// Eigenmode surface GF where decimation failed (n <= 48, pencil 2n <= 96)
// ---------------------------------------------------------------------------
//
// The reference's eigenmode routes (surface_gf.py:84-199) build the transfer matrix
// from the n finite eigenmodes of smallest |lambda| of the pencil
//   A = [[-(d00 - z), -d10], [I, 0]],  B = [[d01, 0], [0, I]]
// (those inside the unit circle, or the n smallest when fewer are inside), as
// T = Phi Lambda Phi^-1 with Sigma = d01 T, g = (z - d00 - Sigma)^-1.  Only the
// invariant subspace of those n eigenvalues matters, so no eigensolver is needed:
// for a radius r, the Moebius map lambda -> (lambda - r)/(lambda + r) sends
// |lambda| < r to the left half-plane, the matrix sign S of
// C = (A + rB)^-1 (A - rB) (scaled Newton, S <- (s S + S^-1 / s) / 2) gives the
// spectral projector P = (I - S)/2 onto it, trace(P) counts it, and r is searched
// (bracket from 1 outwards, then bisection) until the count is n.  With a basis
// P[:, n:2n] of the subspace, T = P12 P22^-1.  The first block row of A and B is
// scaled by 1 / max(|d00|, |d01|, |z|) (eigenvalues unchanged) to balance the
// pencil.  Slots: A, B, Y = (A + rB)^-1 then T, X (work), S, d00, d01, Sigma/G.

struct EigArgs {
  int n, n_inst;
  const double* lead;  // [n_sets][3][n][n]
  const int32_t* set_of;
  const double* omega;
  const double* eta;
  double* g_io;
  double* dos_out;
  int32_t* status;     // instances with status != 0 are (re)computed; 0 on success
};

constexpr int EIG_SIGN_MAXIT = 100;

// Newton sign iteration of the matrix in S (2n x 2n) in place; X is work.  Returns
// true when converged (step below 1e-13, or below 1e-9 and no longer halving).
__device__ bool big_sign(const BigMat& S, const BigMat& X, int m, BigShared& sh, int tid) {
  double dprev = INFINITY;
  for (int it = 0; it < EIG_SIGN_MAXIT; ++it) {
    double ss = 0.0;
    for (int e = tid; e < m * m; e += BIG_NT) {
      const int i = e / m, j = e - (e / m) * m;
      const unsigned o = big_off(i, j);
      const double xr = S.re[o], xi = S.im[o];
      X.re[o] = xr;
      X.im[o] = xi;
      ss += xr * xr + xi * xi;
    }
    const double ns = block_sum<BIG_NT>(ss, sh.sred, tid);
    const int st = gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(X.re, X.im, BIG_LD, m, sh.gj, tid);
    if (st != ST_OK) return false;
    double xs = 0.0;
    for (int e = tid; e < m * m; e += BIG_NT) {
      const int i = e / m, j = e - (e / m) * m;
      const unsigned o = big_off(i, j);
      xs += X.re[o] * X.re[o] + X.im[o] * X.im[o];
    }
    const double nx = block_sum<BIG_NT>(xs, sh.sred, tid);
    const double sc = sqrt(sqrt(nx / ns));
    double dd = 0.0, nn = 0.0;
    for (int e = tid; e < m * m; e += BIG_NT) {
      const int i = e / m, j = e - (e / m) * m;
      const unsigned o = big_off(i, j);
      const double yr = 0.5 * (sc * S.re[o] + X.re[o] / sc);
      const double yi = 0.5 * (sc * S.im[o] + X.im[o] / sc);
      const double dr = yr - S.re[o], di = yi - S.im[o];
      dd += dr * dr + di * di;
      nn += yr * yr + yi * yi;
      S.re[o] = yr;
      S.im[o] = yi;
    }
    const double d = sqrt(block_sum<BIG_NT>(dd, sh.sred, tid) / block_sum<BIG_NT>(nn, sh.sred, tid));
    if (!(d == d)) return false;
    if (d < 1e-13 || (d < 1e-9 && d > 0.5 * dprev)) return true;
    dprev = d;
  }
  return dprev < 1e-9;
}

// Count of eigenvalues with |lambda| < r (-1 when the sign iteration failed); leaves
// the sign matrix in S.
__device__ int big_count_inside(double r, const BigMat& A, const BigMat& B, const BigMat& Y,
                                const BigMat& X, const BigMat& S, int m, BigShared& sh, int tid) {
  for (int e = tid; e < m * m; e += BIG_NT) {
    const int i = e / m, j = e - (e / m) * m;
    const unsigned o = big_off(i, j);
    Y.re[o] = A.re[o] + r * B.re[o];
    Y.im[o] = A.im[o] + r * B.im[o];
    X.re[o] = A.re[o] - r * B.re[o];
    X.im[o] = A.im[o] - r * B.im[o];
  }
  __syncthreads();
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(Y.re, Y.im, BIG_LD, m, sh.gj, tid) != ST_OK) return -1;
  const BigTile t(m, m, tid);
  double cr[BIG_TM][BIG_TM], ci[BIG_TM][BIG_TM];
  big_mul<OPN, OPN>(Y, X, m, m, m, t, cr, ci, sh, tid);
  big_store(S, m, t, cr, ci);
  __syncthreads();
  if (!big_sign(S, X, m, sh, tid)) return -1;
  double tr = 0.0;
  for (int i = tid; i < m; i += BIG_NT) tr += S.re[big_off(i, i)];
  const double trp = 0.5 * ((double)m - block_sum<BIG_NT>(tr, sh.sred, tid));
  const double c = rint(trp);
  return fabs(trp - c) < 1e-3 ? (int)c : -1;
}

__device__ void big_eig_one(const EigArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  if (a.status[inst] == ST_OK) return;
  // eta = 0 with the decimation failing: modes on the unit circle, whose selection the
  // reference leaves to LAPACK's ordering (surface_gf.py:118-124) -- not reproduced
  if (a.eta[inst] == 0.0) {
    if (tid == 0) a.status[inst] = ST_EIGEN_ZERO_ETA;
    return;
  }
  const int n = a.n, m = 2 * n;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;
  double* sl = big_fresh(slot);
  const BigMat A = big_mat(sl, 0), B = big_mat(sl, 1), Y = big_mat(sl, 2), X = big_mat(sl, 3);
  const BigMat S = big_mat(sl, 4), D0 = big_mat(sl, 5), D1 = big_mat(sl, 6), G = big_mat(sl, 7);
  const BigMat D10 = big_mat(sl, 8);
  big_gload(blk, n, n, D0, tid);
  big_gload(blk + nn * 2, n, n, D1, tid);
  big_gload(blk + 2 * nn * 2, n, n, D10, tid);
  double zr, zi;
  zsq(a.omega[inst], a.eta[inst], zr, zi);
  __syncthreads();
  double amax = 0.0;
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    amax = fmax(amax, fmax(hypot(D0.re[o], D0.im[o]), hypot(D1.re[o], D1.im[o])));
  }
  double gmax;
  bool gbad;
  block_max_any<BIG_NT>(amax, false, sh.sred, sh.sflag, tid, gmax, gbad);
  const double s = 1.0 / fmax(gmax, hypot(zr, zi));
  for (int e = tid; e < m * m; e += BIG_NT) {
    const int i = e / m, j = e - (e / m) * m;
    const unsigned o = big_off(i, j);
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    if (i < n && j < n) {
      const unsigned q = big_off(i, j);
      ar = -s * (D0.re[q] - (i == j ? zr : 0.0));
      ai = -s * (D0.im[q] - (i == j ? zi : 0.0));
      br = s * D1.re[q];
      bi = s * D1.im[q];
    } else if (i < n) {
      const unsigned q = big_off(i, j - n);
      ar = -s * D10.re[q];
      ai = -s * D10.im[q];
    } else if (j < n) {
      ar = (i - n == j) ? 1.0 : 0.0;
    } else {
      br = (i == j) ? 1.0 : 0.0;
    }
    A.re[o] = ar;
    A.im[o] = ai;
    B.re[o] = br;
    B.im[o] = bi;
  }
  __syncthreads();
  // radius with exactly n eigenvalues inside
  int c = big_count_inside(1.0, A, B, Y, X, S, m, sh, tid);
  if (c < 0) return;
  if (c != n) {
    const double sgn = c < n ? 1.0 : -1.0;
    double lo = 0.0, hi = -1.0, h = 1e-8;
    while (h < 0.5) {
      c = big_count_inside(1.0 + sgn * h, A, B, Y, X, S, m, sh, tid);
      if (c >= 0 && (sgn > 0 ? c >= n : c <= n)) {
        hi = h;
        break;
      }
      lo = h;
      h *= 4.0;
    }
    if (hi < 0.0) return;
    for (int b = 0; b < 40 && c != n; ++b) {
      h = lo > 0.0 ? sqrt(fmax(lo, 1e-12) * hi) : 0.25 * hi;
      c = big_count_inside(1.0 + sgn * h, A, B, Y, X, S, m, sh, tid);
      if (c < 0) return;
      if (sgn > 0 ? c >= n : c <= n)
        hi = h;
      else
        lo = h;
    }
    if (c != n) return;
  }
  // P22 = (I - S22)/2 -> X (n x n), inverted; T = P12 P22^-1 = -S12 P22^-1 / 2 -> Y
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j), q = big_off(n + i, n + j);
    X.re[o] = 0.5 * ((i == j ? 1.0 : 0.0) - S.re[q]);
    X.im[o] = -0.5 * S.im[q];
  }
  __syncthreads();
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(X.re, X.im, BIG_LD, n, sh.gj, tid) != ST_OK) return;
  const BigMat S12{S.re + n, S.im + n};
  const BigTile t(n, n, tid);
  double cr[BIG_TM][BIG_TM], ci[BIG_TM][BIG_TM];
  big_mul<OPN, OPN>(S12, X, n, n, n, t, cr, ci, sh, tid);
#pragma unroll
  for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
    for (int y = 0; y < BIG_TM; ++y) {
      cr[x][y] *= -0.5;
      ci[x][y] *= -0.5;
    }
  big_store(Y, n, t, cr, ci);
  __syncthreads();
  // Sigma = d01 T;  G = (z - d00 - Sigma)^-1
  big_mul<OPN, OPN>(D1, Y, n, n, n, t, cr, ci, sh, tid);
  if (t.own) {
#pragma unroll
    for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
      for (int y = 0; y < BIG_TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) {
          const unsigned o = big_off(i, j);
          G.re[o] = ((i == j ? zr : 0.0) - D0.re[o]) - cr[x][y];
          G.im[o] = ((i == j ? zi : 0.0) - D0.im[o]) - ci[x][y];
        }
      }
  }
  __syncthreads();
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(G.re, G.im, BIG_LD, n, sh.gj, tid) != ST_OK) return;
  bool bad = false;
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    bad |= !(isfinite(G.re[o]) && isfinite(G.im[o]));
  }
  block_max_any<BIG_NT>(0.0, bad, sh.sred, sh.sflag, tid, gmax, gbad);
  if (gbad) return;
  big_gstore(a.g_io + (size_t)inst * nn * 2, n, G, tid);
  if (tid == 0) {
    if (a.dos_out) {
      double tr = 0.0;
      for (int i = 0; i < n; ++i) tr += -2.0 * G.im[big_off(i, i)];
      a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
    }
    a.status[inst] = ST_OK;
  }
}

__global__ void __launch_bounds__(BIG_NT) k_sgf_eigen_big(EigArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * BIG_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    big_eig_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

}  // namespace negf

namespace negf {

// ---------------------------------------------------------------------------
// Eigenmode surface GF for 49..96-DOF blocks (pencil 2n <= 192)
// ---------------------------------------------------------------------------
//
// The same spectral projector as k_sgf_eigen_big with every 2n x 2n matrix held as
// 2 x 2 blocks of n x n slot matrices.  A 2n inverse is the block inverse through
// the Schur complement of the diagonal block with the smaller condition estimate
// ||M_ii||_F ||M_ii^-1||_F: with P = M11 (or M22 symmetrically), U = P^-1 M12,
// V = M21 P^-1, K = M22 - M21 U:
//   [[P^-1 + U K^-1 V, -U K^-1], [-K^-1 V, K^-1]].
// The first block row of the pencil is scaled as before; B = diag(s d01, I) is not
// stored (A +- rB only changes the diagonal blocks).  Slots (n x n): A 0-3, S 4-7,
// X 8-11 (inverse / work), Y 12-15 ((A + rB)^-1, then T), s d01 16, and work 17-22.

constexpr int HUGE_SLOT_MATS = 23;
constexpr size_t HUGE_SLOT = HUGE_SLOT_MATS * BIG_MAT;

struct Blk4 {
  BigMat b[4];  // 11, 12, 21, 22
};

__device__ __forceinline__ Blk4 blk4(double* slot, int q0) {
  return Blk4{{big_mat(slot, q0), big_mat(slot, q0 + 1), big_mat(slot, q0 + 2), big_mat(slot, q0 + 3)}};
}

__device__ __forceinline__ double big_fro2(const BigMat& M, int n, int tid) {
  double s = 0.0;
  for (int e = tid; e < n * n; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    s += M.re[o] * M.re[o] + M.im[o] * M.im[o];
  }
  return s;
}

__device__ __forceinline__ void big_copy(const BigMat& D, const BigMat& S, int n, int tid) {
  for (int e = tid; e < n * n; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    D.re[o] = S.re[o];
    D.im[o] = S.im[o];
  }
}

// C = alpha op(A) op(B) (+ beta_one * C when acc): result stored after the product's barrier
template <int OPA, int OPB>
__device__ void big_gemm(const BigMat& C, const BigMat& A, const BigMat& B, int n, double alpha,
                         bool acc, BigShared& sh, int tid) {
  const BigTile t(n, n, tid);
  double cr[BIG_TM][BIG_TM], ci[BIG_TM][BIG_TM];
  big_mul<OPA, OPB>(A, B, n, n, n, t, cr, ci, sh, tid);
  if (t.own) {
#pragma unroll
    for (int x = 0; x < BIG_TM; ++x)
#pragma unroll
      for (int y = 0; y < BIG_TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) {
          const unsigned o = big_off(i, j);
          const double r = alpha * cr[x][y], m = alpha * ci[x][y];
          C.re[o] = acc ? C.re[o] + r : r;
          C.im[o] = acc ? C.im[o] + m : m;
        }
      }
  }
  __syncthreads();
}

// Inverse of the 2n matrix M (blocks) into X; W[0..4] are work slots.  False when a
// pivot block or the Schur complement is singular.
__device__ bool blk_inverse(const Blk4& M, const Blk4& X, const BigMat* W, int n, BigShared& sh,
                            int tid) {
  // condition estimates of both diagonal blocks
  double cnd[2];
  for (int d = 0; d < 2; ++d) {
    const BigMat& Mi = M.b[d == 0 ? 0 : 3];
    big_copy(W[d], Mi, n, tid);
    const double nm = block_sum<BIG_NT>(big_fro2(Mi, n, tid), sh.sred, tid);
    if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(W[d].re, W[d].im, BIG_LD, n, sh.gj, tid) != ST_OK) {
      cnd[d] = INFINITY;
      continue;
    }
    cnd[d] = sqrt(nm * block_sum<BIG_NT>(big_fro2(W[d], n, tid), sh.sred, tid));
  }
  if (!(cnd[0] < INFINITY) && !(cnd[1] < INFINITY)) return false;
  const int p = cnd[0] <= cnd[1] ? 0 : 1;         // pivot block 11 (p = 0) or 22 (p = 1)
  const BigMat& Pi = W[p];
  const BigMat& Mpq = M.b[p == 0 ? 1 : 2];         // M12 (or M21)
  const BigMat& Mqp = M.b[p == 0 ? 2 : 1];         // M21 (or M12)
  const BigMat& Mqq = M.b[p == 0 ? 3 : 0];
  const BigMat &U = W[2], &V = W[3], &K = W[4];
  big_gemm<OPN, OPN>(U, Pi, Mpq, n, 1.0, false, sh, tid);   // U = P^-1 M_pq
  big_gemm<OPN, OPN>(V, Mqp, Pi, n, 1.0, false, sh, tid);   // V = M_qp P^-1
  big_copy(K, Mqq, n, tid);
  __syncthreads();
  big_gemm<OPN, OPN>(K, Mqp, U, n, -1.0, true, sh, tid);    // K = M_qq - M_qp U
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(K.re, K.im, BIG_LD, n, sh.gj, tid) != ST_OK) return false;
  // blocks of the result: (pp) = P^-1 + U K^-1 V, (pq) = -U K^-1, (qp) = -K^-1 V, (qq) = K^-1
  const BigMat& Xpp = X.b[p == 0 ? 0 : 3];
  const BigMat& Xpq = X.b[p == 0 ? 1 : 2];
  const BigMat& Xqp = X.b[p == 0 ? 2 : 1];
  const BigMat& Xqq = X.b[p == 0 ? 3 : 0];
  big_gemm<OPN, OPN>(Xpq, U, K, n, -1.0, false, sh, tid);
  big_gemm<OPN, OPN>(Xqp, K, V, n, -1.0, false, sh, tid);
  big_copy(Xpp, Pi, n, tid);
  big_copy(Xqq, K, n, tid);
  __syncthreads();
  big_gemm<OPN, OPN>(Xpp, Xpq, V, n, -1.0, true, sh, tid);  // P^-1 - (-U K^-1) V
  return true;
}

__device__ bool huge_sign(const Blk4& S, const Blk4& X, const BigMat* W, int n, BigShared& sh,
                          int tid) {
  double dprev = INFINITY;
  for (int it = 0; it < EIG_SIGN_MAXIT; ++it) {
    double ss = 0.0;
    for (int q = 0; q < 4; ++q) ss += big_fro2(S.b[q], n, tid);
    const double ns = block_sum<BIG_NT>(ss, sh.sred, tid);
    if (!blk_inverse(S, X, W, n, sh, tid)) return false;
    double xs = 0.0;
    for (int q = 0; q < 4; ++q) xs += big_fro2(X.b[q], n, tid);
    const double nx = block_sum<BIG_NT>(xs, sh.sred, tid);
    const double sc = sqrt(sqrt(nx / ns));
    double dd = 0.0, nn = 0.0;
    for (int q = 0; q < 4; ++q) {
      for (int e = tid; e < n * n; e += BIG_NT) {
        const int i = e / n, j = e - (e / n) * n;
        const unsigned o = big_off(i, j);
        const double yr = 0.5 * (sc * S.b[q].re[o] + X.b[q].re[o] / sc);
        const double yi = 0.5 * (sc * S.b[q].im[o] + X.b[q].im[o] / sc);
        const double dr = yr - S.b[q].re[o], di = yi - S.b[q].im[o];
        dd += dr * dr + di * di;
        nn += yr * yr + yi * yi;
        S.b[q].re[o] = yr;
        S.b[q].im[o] = yi;
      }
    }
    const double d = sqrt(block_sum<BIG_NT>(dd, sh.sred, tid) / block_sum<BIG_NT>(nn, sh.sred, tid));
    if (!(d == d)) return false;
    if (d < 1e-13 || (d < 1e-9 && d > 0.5 * dprev)) return true;
    dprev = d;
  }
  return dprev < 1e-9;
}

// eigenvalues of the pencil with |lambda| < r (-1 on failure); leaves the sign in S
__device__ int huge_count_inside(double r, const Blk4& A, const BigMat& D1s, const Blk4& S,
                                 const Blk4& X, const Blk4& Y, const BigMat* W, int n,
                                 BigShared& sh, int tid) {
  // Y <- A + rB (inverted below), S <- A - rB
  for (int q = 0; q < 4; ++q) {
    for (int e = tid; e < n * n; e += BIG_NT) {
      const int i = e / n, j = e - (e / n) * n;
      const unsigned o = big_off(i, j);
      double br = 0.0, bi = 0.0;
      if (q == 0) {
        br = D1s.re[o];
        bi = D1s.im[o];
      } else if (q == 3) {
        br = (i == j) ? 1.0 : 0.0;
      }
      Y.b[q].re[o] = A.b[q].re[o] + r * br;
      Y.b[q].im[o] = A.b[q].im[o] + r * bi;
      S.b[q].re[o] = A.b[q].re[o] - r * br;
      S.b[q].im[o] = A.b[q].im[o] - r * bi;
    }
  }
  __syncthreads();
  if (!blk_inverse(Y, X, W, n, sh, tid)) return -1;
  // C = (A + rB)^-1 (A - rB): block product X * S into Y
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj) {
      big_gemm<OPN, OPN>(Y.b[2 * bi + bj], X.b[2 * bi], S.b[bj], n, 1.0, false, sh, tid);
      big_gemm<OPN, OPN>(Y.b[2 * bi + bj], X.b[2 * bi + 1], S.b[2 + bj], n, 1.0, true, sh, tid);
    }
  for (int q = 0; q < 4; ++q) big_copy(S.b[q], Y.b[q], n, tid);
  __syncthreads();
  if (!huge_sign(S, X, W, n, sh, tid)) return -1;
  double tr = 0.0;
  for (int i = tid; i < n; i += BIG_NT) tr += S.b[0].re[big_off(i, i)] + S.b[3].re[big_off(i, i)];
  const double trp = 0.5 * (2.0 * n - block_sum<BIG_NT>(tr, sh.sred, tid));
  const double c = rint(trp);
  return fabs(trp - c) < 1e-3 ? (int)c : -1;
}

__device__ void huge_eig_one(const EigArgs& a, int inst, double* slot, BigShared& sh, int tid) {
  if (a.status[inst] == ST_OK) return;
  // eta = 0 with the decimation failing: modes on the unit circle, whose selection the
  // reference leaves to LAPACK's ordering (surface_gf.py:118-124) -- not reproduced
  if (a.eta[inst] == 0.0) {
    if (tid == 0) a.status[inst] = ST_EIGEN_ZERO_ETA;
    return;
  }
  const int n = a.n;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;
  double* sl = big_fresh(slot);
  const Blk4 A = blk4(sl, 0), S = blk4(sl, 4), X = blk4(sl, 8), Y = blk4(sl, 12);
  const BigMat D1s = big_mat(sl, 16);
  const BigMat W[5] = {big_mat(sl, 17), big_mat(sl, 18), big_mat(sl, 19), big_mat(sl, 20),
                       big_mat(sl, 21)};
  const BigMat D0 = big_mat(sl, 22);
  // raw blocks: d00 -> D0, d01 -> D1s (scaled below), d10 -> A12 (scaled below)
  big_gload(blk, n, n, D0, tid);
  big_gload(blk + nn * 2, n, n, D1s, tid);
  big_gload(blk + 2 * nn * 2, n, n, A.b[1], tid);
  double zr, zi;
  zsq(a.omega[inst], a.eta[inst], zr, zi);
  __syncthreads();
  double amax = 0.0;
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    amax = fmax(amax, fmax(hypot(D0.re[o], D0.im[o]), hypot(D1s.re[o], D1s.im[o])));
  }
  double gmax;
  bool gbad;
  block_max_any<BIG_NT>(amax, false, sh.sred, sh.sflag, tid, gmax, gbad);
  const double s = 1.0 / fmax(gmax, hypot(zr, zi));
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    A.b[0].re[o] = -s * (D0.re[o] - (i == j ? zr : 0.0));
    A.b[0].im[o] = -s * (D0.im[o] - (i == j ? zi : 0.0));
    A.b[1].re[o] = -s * A.b[1].re[o];
    A.b[1].im[o] = -s * A.b[1].im[o];
    A.b[2].re[o] = (i == j) ? 1.0 : 0.0;
    A.b[2].im[o] = 0.0;
    A.b[3].re[o] = 0.0;
    A.b[3].im[o] = 0.0;
    D1s.re[o] = s * D1s.re[o];
    D1s.im[o] = s * D1s.im[o];
  }
  __syncthreads();
  const int m = 2 * n;
  int c = huge_count_inside(1.0, A, D1s, S, X, Y, W, n, sh, tid);
  if (c < 0) return;
  if (c != n) {
    const double sgn = c < n ? 1.0 : -1.0;
    double lo = 0.0, hi = -1.0, h = 1e-8;
    while (h < 0.5) {
      c = huge_count_inside(1.0 + sgn * h, A, D1s, S, X, Y, W, n, sh, tid);
      if (c >= 0 && (sgn > 0 ? c >= n : c <= n)) {
        hi = h;
        break;
      }
      lo = h;
      h *= 4.0;
    }
    if (hi < 0.0) return;
    for (int b = 0; b < 40 && c != n; ++b) {
      h = lo > 0.0 ? sqrt(fmax(lo, 1e-12) * hi) : 0.25 * hi;
      c = huge_count_inside(1.0 + sgn * h, A, D1s, S, X, Y, W, n, sh, tid);
      if (c < 0) return;
      if (sgn > 0 ? c >= n : c <= n)
        hi = h;
      else
        lo = h;
    }
    if (c != n) return;
  }
  (void)m;
  // T = P12 P22^-1 = -S12 ((I - S22)/2)^-1 / 2 -> Y11;  Sigma = d01 T;  G = (z - d00 - Sigma)^-1
  const BigMat& P22 = W[0];
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    P22.re[o] = 0.5 * ((i == j ? 1.0 : 0.0) - S.b[3].re[o]);
    P22.im[o] = -0.5 * S.b[3].im[o];
  }
  __syncthreads();
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(P22.re, P22.im, BIG_LD, n, sh.gj, tid) != ST_OK) return;
  const BigMat& T = Y.b[0];
  big_gemm<OPN, OPN>(T, S.b[1], P22, n, -0.5, false, sh, tid);
  // d01 = D1s / s (unscaled) into W1, Sigma = d01 T into W2
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    W[1].re[o] = D1s.re[o] / s;
    W[1].im[o] = D1s.im[o] / s;
  }
  __syncthreads();
  big_gemm<OPN, OPN>(W[2], W[1], T, n, 1.0, false, sh, tid);
  const BigMat& G = W[3];
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    G.re[o] = ((i == j ? zr : 0.0) - D0.re[o]) - W[2].re[o];
    G.im[o] = ((i == j ? zi : 0.0) - D0.im[o]) - W[2].im[o];
  }
  __syncthreads();
  if (gj_inverse_sel<double, BIG_NT, BIG, BIG_RG>(G.re, G.im, BIG_LD, n, sh.gj, tid) != ST_OK) return;
  bool bad = false;
  for (int e = tid; e < (int)nn; e += BIG_NT) {
    const int i = e / n, j = e - (e / n) * n;
    const unsigned o = big_off(i, j);
    bad |= !(isfinite(G.re[o]) && isfinite(G.im[o]));
  }
  block_max_any<BIG_NT>(0.0, bad, sh.sred, sh.sflag, tid, gmax, gbad);
  if (gbad) return;
  big_gstore(a.g_io + (size_t)inst * nn * 2, n, G, tid);
  if (tid == 0) {
    if (a.dos_out) {
      double tr = 0.0;
      for (int i = 0; i < n; ++i) tr += -2.0 * G.im[big_off(i, i)];
      a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
    }
    a.status[inst] = ST_OK;
  }
}

__global__ void __launch_bounds__(BIG_NT) k_sgf_eigen_huge(EigArgs a, double* pool) {
  __shared__ BigShared sh;
  double* slot = pool + (size_t)blockIdx.x * HUGE_SLOT;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    huge_eig_one(a, inst, slot, sh, threadIdx.x);
    __syncthreads();
  }
}

}  // namespace negf
