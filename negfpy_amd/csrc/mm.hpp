// Complex FP64 block products on the matrix cores for the workgroup kernels.
//
// v_mfma_f64_4x4x4f64 computes four independent 4 x 4 x 4 products per wavefront
// at the full FP64 rate (256 multiply-adds per instruction, a quarter of the cycles
// of the 16 x 16 x 4 form), so blocks of any multiple of 4 DOF -- 24 included --
// run without padding waste.  Lane layout (measured on gfx950 with a probe):
//   A[blk][i][k] at lane 16 k + 4 blk + i
//   B[blk][k][j] at lane 16 k + 4 blk + j
//   C/D[blk][i][j] at lane 16 i + 4 blk + j
// and bit 0 of the blgp field negates A (bit 1 B, bit 2 C).
//
// An NMAX x NMAX output is cut into 8 x 8 tile-groups, each the four 4 x 4 blocks
// (2 x 2) of one instruction.  Wave w of the workgroup owns tile-groups
// w * TPW .. w * TPW + TPW - 1 in row-major order, and lane l holds one element of
// each: row 8 R + 4 (blk >> 1) + (l >> 4), column 8 C + 4 (blk & 1) + (l & 3), with
// blk = (l >> 2) & 3.  State matrices that products update (decimation iterates,
// the sweep's Schur term) are kept in that ownership, one element per tile-group.
//
// A complex product C += op(A) op(B) is four real chains per tile-group,
//   Cr += Ar Br - Ai Bi,   Ci += Ar Bi + Ai Br,
// the minus (and the conjugation of an OPH operand) folded into the negation bit.
// Operands come from the planar LDS matrices of cmat.hpp; entries of the padding
// (rows or columns in [n, 4 ceil(n / 4))) must be zero there, which holds because
// every matrix is cleared at kernel start and only its [0, n)^2 part is written.

#pragma once

#include "cmat.hpp"

namespace negf {

template <int NEG>
__device__ __forceinline__ double mf44(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, NEG);
}

template <int NMAX, int NT>
struct MOwn {
  static constexpr int TGR = NMAX / 8;          // tile-groups per row / column
  static constexpr int NTG = TGR * TGR;
  static constexpr int NW = NT / WAVE;          // waves of the workgroup
  static constexpr int TPW = NTG / NW;          // tile-groups per wave
  static constexpr int NS = NMAX / 4;           // K steps of a full product
  static constexpr bool SHARED_R = (TGR % TPW) == 0;  // a wave's tile-groups share R
  // K steps unrolled at once: all of them up to 24 DOF; 32 and 48 DOF in groups of 4, which
  // keeps the loads in flight within the registers of three waves per SIMD
  static constexpr int UNROLL = NS <= 6 ? NS : 4;
  static_assert(NMAX % 8 == 0 && NTG % NW == 0 && TPW >= 1, "tile-groups must tile the waves");
  int w, k, blk, i;
  __device__ __forceinline__ explicit MOwn(int tid) {
    const int l = tid & (WAVE - 1);
    w = tid / WAVE;
    k = l >> 4;
    blk = (l >> 2) & 3;
    i = l & 3;
  }
  __device__ __forceinline__ int tgr(int q) const { return (w * TPW + q) / TGR; }
  __device__ __forceinline__ int tgc(int q) const { return (w * TPW + q) % TGR; }
  // A-operand row and B-operand column of this lane in tile-group q
  __device__ __forceinline__ int arow(int q) const { return 8 * tgr(q) + 4 * (blk >> 1) + i; }
  __device__ __forceinline__ int bcol(int q) const { return 8 * tgc(q) + 4 * (blk & 1) + i; }
  // output element of this lane in tile-group q
  __device__ __forceinline__ int row(int q) const { return 8 * tgr(q) + 4 * (blk >> 1) + k; }
  __device__ __forceinline__ int col(int q) const { return 8 * tgc(q) + 4 * (blk & 1) + i; }
};

// Operand element of op(X) at (r, c) of a planar LDS matrix; the sign of the
// imaginary part of OPH / OPC is applied by the caller through the negation bits.
template <int OP>
__device__ __forceinline__ void mop(const double* __restrict__ re, const double* __restrict__ im,
                                    int ld, int r, int c, double& xr, double& xi) {
  if constexpr (OP == OPN || OP == OPC) {
    xr = re[r * ld + c];
    xi = im[r * ld + c];
  } else {
    xr = re[c * ld + r];
    xi = im[c * ld + r];
  }
}
template <int OP>
constexpr int imsign() {
  return (OP == OPH || OP == OPC) ? -1 : 1;
}

// One K step of C += op(A) op(B) for one tile-group: the four real chains.
template <int OPA, int OPB>
__device__ __forceinline__ void mstep(double ar, double ai, double br, double bi, double& cr,
                                      double& ci) {
  constexpr int sa = imsign<OPA>(), sb = imsign<OPB>();
  cr = mf44<0>(ar, br, cr);
  ci = mf44<(sb < 0) ? 1 : 0>(ar, bi, ci);
  cr = mf44<(sa * sb > 0) ? 1 : 0>(ai, bi, cr);
  ci = mf44<(sa < 0) ? 1 : 0>(ai, br, ci);
}

// C(q) += op(A) op(B) over K in [0, 4 ns) for every tile-group q of this wave.
template <int NMAX, int NT, int OPA, int OPB>
__device__ __forceinline__ void mm(const double* are, const double* aim, const double* bre,
                                   const double* bim, int ld, int ns, const MOwn<NMAX, NT>& o,
                                   double (&cr)[MOwn<NMAX, NT>::TPW],
                                   double (&ci)[MOwn<NMAX, NT>::TPW]) {
  using M = MOwn<NMAX, NT>;
#pragma unroll M::UNROLL
  for (int s = 0; s < M::NS; ++s) {
    if (s < ns) {
      const int kc = 4 * s + o.k;
      double xr = 0, xi = 0;
#pragma unroll
      for (int q = 0; q < M::TPW; ++q) {
        if (q == 0 || !M::SHARED_R) mop<OPA>(are, aim, ld, o.arow(q), kc, xr, xi);
        double yr, yi;
        mop<OPB>(bre, bim, ld, kc, o.bcol(q), yr, yi);
        mstep<OPA, OPB>(xr, xi, yr, yi, cr[q], ci[q]);
      }
    }
  }
}

// Two products sharing the B operand: C1 += op(A1) B, C2 += op(A2) B.
template <int NMAX, int NT, int OPA1, int OPA2>
__device__ __forceinline__ void mm2b(const double* a1r, const double* a1i, const double* a2r,
                                     const double* a2i, const double* bre, const double* bim,
                                     int ld, int ns, const MOwn<NMAX, NT>& o,
                                     double (&c1r)[MOwn<NMAX, NT>::TPW],
                                     double (&c1i)[MOwn<NMAX, NT>::TPW],
                                     double (&c2r)[MOwn<NMAX, NT>::TPW],
                                     double (&c2i)[MOwn<NMAX, NT>::TPW]) {
  using M = MOwn<NMAX, NT>;
#pragma unroll M::UNROLL
  for (int s = 0; s < M::NS; ++s) {
    if (s < ns) {
      const int kc = 4 * s + o.k;
      double ur = 0, ui = 0, vr = 0, vi = 0;
#pragma unroll
      for (int q = 0; q < M::TPW; ++q) {
        if (q == 0 || !M::SHARED_R) {
          mop<OPA1>(a1r, a1i, ld, o.arow(q), kc, ur, ui);
          mop<OPA2>(a2r, a2i, ld, o.arow(q), kc, vr, vi);
        }
        double yr, yi;
        mop<OPN>(bre, bim, ld, kc, o.bcol(q), yr, yi);
        mstep<OPA1, OPN>(ur, ui, yr, yi, c1r[q], c1i[q]);
        mstep<OPA2, OPN>(vr, vi, yr, yi, c2r[q], c2i[q]);
      }
    }
  }
}

// Two products sharing the A operand: C1 += A B1, C2 += A B2.
template <int NMAX, int NT>
__device__ __forceinline__ void mm2a(const double* are, const double* aim, const double* b1r,
                                     const double* b1i, const double* b2r, const double* b2i,
                                     int ld, int ns, const MOwn<NMAX, NT>& o,
                                     double (&c1r)[MOwn<NMAX, NT>::TPW],
                                     double (&c1i)[MOwn<NMAX, NT>::TPW],
                                     double (&c2r)[MOwn<NMAX, NT>::TPW],
                                     double (&c2i)[MOwn<NMAX, NT>::TPW]) {
  using M = MOwn<NMAX, NT>;
#pragma unroll M::UNROLL
  for (int s = 0; s < M::NS; ++s) {
    if (s < ns) {
      const int kc = 4 * s + o.k;
      double xr = 0, xi = 0;
#pragma unroll
      for (int q = 0; q < M::TPW; ++q) {
        if (q == 0 || !M::SHARED_R) mop<OPN>(are, aim, ld, o.arow(q), kc, xr, xi);
        double ur, ui, vr, vi;
        mop<OPN>(b1r, b1i, ld, kc, o.bcol(q), ur, ui);
        mop<OPN>(b2r, b2i, ld, kc, o.bcol(q), vr, vi);
        mstep<OPN, OPN>(xr, xi, ur, ui, c1r[q], c1i[q]);
        mstep<OPN, OPN>(xr, xi, vr, vi, c2r[q], c2i[q]);
      }
    }
  }
}

template <int TPW>
__device__ __forceinline__ void mzero(double (&r)[TPW], double (&i)[TPW]) {
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    r[q] = 0.0;
    i[q] = 0.0;
  }
}

// Store the owned elements inside [0, n)^2 into a planar LDS matrix.
template <int NMAX, int NT>
__device__ __forceinline__ void mstore(double* re, double* im, int ld, int n,
                                       const MOwn<NMAX, NT>& o,
                                       const double (&cr)[MOwn<NMAX, NT>::TPW],
                                       const double (&ci)[MOwn<NMAX, NT>::TPW]) {
#pragma unroll
  for (int q = 0; q < MOwn<NMAX, NT>::TPW; ++q) {
    const int r = o.row(q), c = o.col(q);
    if (r < n && c < n) {
      re[r * ld + c] = cr[q];
      im[r * ld + c] = ci[q];
    }
  }
}

}  // namespace negf
