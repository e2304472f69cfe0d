// Batched phonon-NEGF transport kernels for MI355X (gfx950) and their C ABI.
//
// One workgroup solves one (omega, k_par) instance; all per-instance matrices
// are held in LDS in planar complex form (cmat.hpp), except for the pooled 96-DOF
// class (big.hpp), whose matrices sit in a device-memory slot per workgroup.  Kernels:
//
//   k_sancho_rubio<T> / k_sr_mm   Lopez-Sancho decimation of a semi-infinite lead
//                      (reference src/negfpy/core/surface_gf.py:18-81).  In FP64
//                      mode it ends with g_s = (z - e_s)^-1 and the surface DOS;
//                      in FP32 mode it hands e_s to k_refine.
//   k_refine           FP64 g_s = (z - e_s)^-1, then Newton steps on
//                      g = (z - d00 - d01 g d10)^-1 (correction from a Stein equation
//                      solved in FP32 by doubling), certified by an FP64 residual;
//                      uncertified instances are marked for the FP64 decimation.
//   k_self_energy      Sigma = V g V^H embedded in the device end layer, FP64
//                      (negf.py:72-97, :245-248).
//   k_rgf<T> / k_rgf_mm  block recursive Green's function sweep for G_{1N} of
//                      A = z - D - Sigma_L - Sigma_R (replaces the sparse LU of
//                      negf.py:350-354; same matrix, same G_{1N}); k_rgf_mm<NMAX, true>
//                      forms Sigma_L, Sigma_R and the trace in the same launch.
//   k_trace            T = Re Tr[Gamma_L G_1N Gamma_R G_1N^H], Gamma = i(S - S^H), FP64
//                      (negf.py:115-116, :356-357).
//   k_adjoint          T and dT/d(every device block) by one forward and one adjoint
//                      sweep; k_lead_adjoint carries dT/dSigma into the lead blocks.
//   k_conductance      Landauer thermal conductance per temperature, FP64
//                      (workflows/ifc_bulk.py:408-438).
//   k_heat_current     finite-bias Landauer heat current, FP64
//                      (core/thermal_current.py:30-67).
// Block products run on the FP64 matrix cores (mm.hpp, v_mfma_f64_4x4x4f64; big.hpp,
// v_mfma_f64_16x16x4f64) or as register-tiled FMA loops (cmat.hpp mac).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "cmat.hpp"
#include "mm.hpp"
#include "wave.hpp"
#include "../../include/negf_hip.h"

namespace negf {

// ---------------------------------------------------------------------------
// configurations: (NMAX, threads, tile) per block-size class
// ---------------------------------------------------------------------------

template <int NMAX_>
struct Cfg;
template <>
struct Cfg<8> {
  static constexpr int NMAX = 8, NT = 64, TM = 1;
};
template <>
struct Cfg<16> {
  static constexpr int NMAX = 16, NT = 64, TM = 2;
};
template <>
struct Cfg<24> {
  static constexpr int NMAX = 24, NT = 192, TM = 2;
};
template <>
struct Cfg<32> {
  static constexpr int NMAX = 32, NT = 256, TM = 2;
};
template <>
struct Cfg<48> {
  static constexpr int NMAX = 48, NT = 768, TM = 2;
};
template <>
struct Cfg<64> {
  static constexpr int NMAX = 64, NT = 1024, TM = 2;
};
// matrices in a device-memory pool slot (the full Green's function and adjoint kernels)
template <>
struct Cfg<96> {
  static constexpr int NMAX = 96, NT = 768, TM = 4;
};

// Waves per SIMD the main kernels must allow: the LDS footprint of one
// instance admits about this many resident workgroups per CU, so capping
// registers below that would only leave LDS unused.  FP64 class 16 (one wave per
// instance, 13 KB of LDS) fits twelve instances per CU: at three waves per SIMD
// the decimation keeps its state in 168 VGPRs without spilling (the mirrored
// form spills 25) and runs 8-12 % faster than at two.
template <typename T, int NMAX>
struct Occ {
  static constexpr int value = (NMAX <= 32) ? (sizeof(T) == 8 ? 3 : 4) : 1;
};
// The recursive sweep keeps three matrices per instance and fits 128 VGPRs: at
// n = 24 in FP64 five workgroups (15 waves) fit a CU.  (The decimation also keeps
// three matrices but needs more registers than 128 without spilling.)
template <typename T, int NMAX>
struct Occ3 {
  static constexpr int value = (sizeof(T) == 8 && NMAX == 24) ? 4 : Occ<T, NMAX>::value;
};

// The decimation with matrix-core products keeps its state in fewer registers than
// the register-tiled form: at 24 DOF it runs at four waves per SIMD, so the five
// instances per CU that its LDS admits are all resident (a few VGPRs spill around
// the inversion; measured 3.36 -> 3.11 ms on the 4096-instance ribbon).  The
// mirrored form carries one more state matrix and stays at three.
template <int NMAX, bool MIRROR>
struct OccMM {
  static constexpr int value = (NMAX == 24 && !MIRROR) ? 4 : Occ<double, NMAX>::value;
};

template <typename T, int NMAX, int NT, int NMAT>
struct Lds {
  static constexpr int LD = NMAX + 1;
  static constexpr int MSZ = NMAX * LD;
  // rows per column owner group of the inversion: largest power of two <= NT / NMAX
  static constexpr int RG = (NT / NMAX >= 16) ? 16 : (NT / NMAX >= 8) ? 8 : (NT / NMAX >= 4) ? 4
                            : (NT / NMAX >= 2) ? 2 : 1;
  using Gj = GjBuf<T, NMAX>;
  static constexpr size_t mat_bytes() { return sizeof(T) * (size_t)(2 * NMAT * MSZ); }
  static constexpr size_t gj_off() { return (mat_bytes() + 15) & ~size_t(15); }
  static constexpr size_t red_off() { return (gj_off() + sizeof(Gj) + 15) & ~size_t(15); }
  static constexpr size_t bytes() {
    return red_off() + sizeof(double) * (NT / WAVE) + sizeof(int) * (NT / WAVE) + 16;
  }
  T* re[NMAT];
  T* im[NMAT];
  Gj* gj;
  double* sred;
  int* sflag;
  unsigned char* base_;
  __device__ __forceinline__ void zero(int tid) {
    uint32_t* p = reinterpret_cast<uint32_t*>(base_);
    for (size_t e = tid; e < mat_bytes() / 4; e += NT) p[e] = 0u;
  }
  __device__ __forceinline__ Lds(unsigned char* base) : base_(base) {
    T* p = reinterpret_cast<T*>(base);
#pragma unroll
    for (int q = 0; q < NMAT; ++q) {
      re[q] = p;
      p += MSZ;
      im[q] = p;
      p += MSZ;
    }
    gj = reinterpret_cast<Gj*>(base + gj_off());
    sred = reinterpret_cast<double*>(base + red_off());
    sflag = reinterpret_cast<int*>(sred + NT / WAVE);
  }
};

// The same interface with the matrices in a device-memory slot (NMAT planar
// matrices of NMAX x (NMAX + 1)) and the pivot buffers and reduction scratch in LDS:
// used where the matrices of one instance exceed the LDS (class 96).
template <typename T, int NMAX, int NT, int NMAT>
struct GmShared {
  GjBuf<T, NMAX> gj;
  double sred[NT / WAVE];
  int sflag[NT / WAVE];
};

template <typename T, int NMAX, int NT, int NMAT>
struct GmL {
  static constexpr int LD = NMAX + 1;
  static constexpr int MSZ = NMAX * LD;
  static constexpr int RG = Lds<T, NMAX, NT, NMAT>::RG;
  static constexpr size_t slot_elems() { return (size_t)2 * NMAT * MSZ; }
  T* re[NMAT];
  T* im[NMAT];
  GjBuf<T, NMAX>* gj;
  double* sred;
  int* sflag;
  T* base_;
  __device__ __forceinline__ void zero(int tid) {
    for (size_t e = tid; e < slot_elems(); e += NT) base_[e] = T(0);
  }
  __device__ __forceinline__ GmL(T* slot, GmShared<T, NMAX, NT, NMAT>* sh) : base_(slot) {
    T* p = slot;
#pragma unroll
    for (int q = 0; q < NMAT; ++q) {
      re[q] = p;
      p += MSZ;
      im[q] = p;
      p += MSZ;
    }
    gj = &sh->gj;
    sred = sh->sred;
    sflag = sh->sflag;
  }
};

template <typename T, int NT>
__device__ __forceinline__ void lds_zero(unsigned char* base, size_t nbytes, int tid) {
  uint32_t* p = reinterpret_cast<uint32_t*>(base);
  const size_t nw = nbytes / 4;
  for (size_t e = tid; e < nw; e += NT) p[e] = 0u;
}

// Load an m x p complex128 block (row-major, leading dim p) into planar LDS as T.
template <typename T, int NT>
__device__ __forceinline__ void gload(const double* __restrict__ src, int m, int p, T* re, T* im,
                                      int ld, int tid) {
  const double2* s = reinterpret_cast<const double2*>(src);
  const int mp = m * p;
  for (int e = tid; e < mp; e += NT) {
    const int i = e / p, j = e - (e / p) * p;
    const double2 v = s[e];
    re[i * ld + j] = (T)v.x;
    im[i * ld + j] = (T)v.y;
  }
}

// Store an LDS n x n planar matrix into complex128 global memory.
template <typename T, int NT>
__device__ __forceinline__ void gstore(double* __restrict__ dst, int n, const T* re, const T* im, int ld,
                                       int tid) {
  double2* d = reinterpret_cast<double2*>(dst);
  const int nn = n * n;
  for (int e = tid; e < nn; e += NT) {
    const int i = e / n, j = e - (e / n) * n;
    d[e] = make_double2((double)re[i * ld + j], (double)im[i * ld + j]);
  }
}

__device__ __forceinline__ bool finite_d(double x) { return isfinite(x); }
__device__ __forceinline__ bool finite_d(float x) { return isfinite(x); }

// ---------------------------------------------------------------------------
// Sancho-Rubio decimation
// ---------------------------------------------------------------------------

struct SrArgs {
  int n;
  int n_inst;
  const double* lead;  // [n_sets][3][n][n] complex128: d00, d01, d10
  const int32_t* set_of;
  const double* omega;
  const double* eta;
  double tol;
  int maxiter;
  double* g_out;  // [n_inst][n][n]: g_s (FP64 mode) or e_s (FP32 mode)
  double* dos_out;
  int32_t* iters;
  int32_t* status;
  int redo;  // 1: only instances whose status is nonzero are (re)computed
  // mirrored lead (d00, d10, d01), from the same decimation (k_sancho_rubio<T, NMAX, true>)
  double* g_out2;
  double* dos_out2;
  int32_t* iters2;
  int32_t* status2;
};


// g = (z - e_s)^-1 from the VGPR tiles of e_s into out (FP64), or e_s itself (FP32
// mode, for the FP64 refinement); the surface DOS (omega / pi) tr(-2 Im g).
template <typename T, int NMAX, int TM, typename L>
__device__ __forceinline__ void sr_write(const T (&er)[TM][TM], const T (&ei)[TM][TM],
                                         const Tile<TM>& t, L& s, int n, T zr, T zi, double omega,
                                         double* out, double* dos, int32_t* status, int inst,
                                         int tid) {
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  if constexpr (sizeof(T) == 4) {
    if (t.own) {
      double2* o = reinterpret_cast<double2*>(out);
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) o[i * n + j] = make_double2((double)er[x][y], (double)ei[x][y]);
        }
    }
  } else {
    __syncthreads();  // earlier reads of plane 2 are done
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            s.re[2][i * LD + j] = (i == j ? zr : T(0)) - er[x][y];
            s.im[2][i * LD + j] = (i == j ? zi : T(0)) - ei[x][y];
          }
        }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      if (tid == 0) status[inst] = st;
      return;
    }
    gstore<T, NT>(out, n, s.re[2], s.im[2], LD, tid);
    if (tid == 0 && dos) {
      double tr = 0.0;
      for (int i = 0; i < n; ++i) tr += -2.0 * (double)s.im[2][i * LD + i];
      dos[inst] = (omega / M_PI) * tr;
    }
  }
}

// MIRROR: the decimation also sums e~_s += b g a, the surface term of the lead with
// d01 and d10 exchanged (the same e, alpha and beta iterates; the reference decimates
// that lead on its own, surface_gf.py:39-79), and writes its g to g_out2.
template <typename T, int NMAX, bool MIRROR>
__global__ void __launch_bounds__(Cfg<NMAX>::NT)
    __attribute__((amdgpu_waves_per_eu(Occ<T, NMAX>::value))) k_sancho_rubio(SrArgs a) {
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  using L = Lds<T, NMAX, NT, 3>;
  constexpr int LD = L::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  // redo pass: only the surfaces whose status is nonzero (uniform per block)
  const bool do1 = !a.redo || a.status[inst] != ST_OK;
  bool do2 = false;
  if constexpr (MIRROR) do2 = !a.redo || a.status2[inst] != ST_OK;
  if (!do1 && !do2) return;
  const int n = a.n;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;

  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  // alpha = d01 -> s.re[0], beta = d10 -> s.re[1]; g, then T1 = alpha g, then T2 = beta g -> [2]
  gload<T, NT>(blk + nn * 2, n, n, s.re[0], s.im[0], LD, tid);
  gload<T, NT>(blk + 2 * nn * 2, n, n, s.re[1], s.im[1], LD, tid);

  const Tile<TM> t(n, n, tid);
  T er[TM][TM], ei[TM][TM], sr[TM][TM], si[TM][TM];
  tzero<T, TM>(er, ei);
  if (t.own) {
    const double2* d00 = reinterpret_cast<const double2*>(blk);
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) {
          const double2 v = d00[i * n + j];
          er[x][y] = (T)v.x;
          ei[x][y] = (T)v.y;
        }
      }
  }
  T mr[MIRROR ? TM : 1][MIRROR ? TM : 1], mi[MIRROR ? TM : 1][MIRROR ? TM : 1];
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TM; ++y) {
      sr[x][y] = er[x][y];
      si[x][y] = ei[x][y];
      if constexpr (MIRROR) {
        mr[x][y] = er[x][y];
        mi[x][y] = ei[x][y];
      }
    }
  double zrd, zid;
  zsq(a.omega[inst], a.eta[inst], zrd, zid);
  const T zr = (T)zrd, zi = (T)zid;
  const T tol = (T)a.tol;
  __syncthreads();

  int status = ST_OK;
  int it = 0;
  bool conv = false;
  double gprev = 0.0;  // max |alpha|^2, |beta|^2 of the previous doubling (FP32 stall test)
  while (it < a.maxiter) {
    ++it;
    // G = z - e
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            s.re[2][i * LD + j] = (i == j ? zr : T(0)) - er[x][y];
            s.im[2][i * LD + j] = (i == j ? zi : T(0)) - ei[x][y];
          }
        }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = st;
      break;
    }
    // T1 = a g, T2 = b g
    T t1r[TM][TM], t1i[TM][TM], t2r[TM][TM], t2i[TM][TM];
    tzero<T, TM>(t1r, t1i);
    tzero<T, TM>(t2r, t2i);
    if (t.own)
      mac2b<T, TM>(s.re[0], s.im[0], s.re[1], s.im[1], s.re[2], s.im[2], LD, n, t.i0, t.j0, t1r, t1i,
                   t2r, t2i);
    __syncthreads();
    tstore<T, TM>(s.re[2], s.im[2], LD, n, n, t, t1r, t1i);
    __syncthreads();
    // P1 = T1 b (into e_s and e at once), P3 = T1 a (the next alpha)
    T p1r[TM][TM], p1i[TM][TM], p3r[TM][TM], p3i[TM][TM];
    tzero<T, TM>(p1r, p1i);
    tzero<T, TM>(p3r, p3i);
    if (t.own)
      mac2a<T, TM>(s.re[2], s.im[2], s.re[1], s.im[1], s.re[0], s.im[0], LD, n, t.i0, t.j0, p1r, p1i,
                   p3r, p3i);
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TM; ++y) {
        sr[x][y] = sr[x][y] + p1r[x][y];
        si[x][y] = si[x][y] + p1i[x][y];
        er[x][y] = er[x][y] + p1r[x][y];
        ei[x][y] = ei[x][y] + p1i[x][y];
      }
    __syncthreads();
    tstore<T, TM>(s.re[2], s.im[2], LD, n, n, t, t2r, t2i);
    __syncthreads();
    // P2 = T2 a (into e), P4 = T2 b (the next beta)
    T p2r[TM][TM], p2i[TM][TM], p4r[TM][TM], p4i[TM][TM];
    tzero<T, TM>(p2r, p2i);
    tzero<T, TM>(p4r, p4i);
    if (t.own)
      mac2a<T, TM>(s.re[2], s.im[2], s.re[0], s.im[0], s.re[1], s.im[1], LD, n, t.i0, t.j0, p2r, p2i,
                   p4r, p4i);
    __syncthreads();
    bool bad = false;
    double emax = 0.0;
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            er[x][y] = er[x][y] + p2r[x][y];
            ei[x][y] = ei[x][y] + p2i[x][y];
            if constexpr (MIRROR) {
              mr[x][y] = mr[x][y] + p2r[x][y];
              mi[x][y] = mi[x][y] + p2i[x][y];
              bad |= !(finite_d(mr[x][y]) && finite_d(mi[x][y]));
            }
            s.re[0][i * LD + j] = p3r[x][y];
            s.im[0][i * LD + j] = p3i[x][y];
            s.re[1][i * LD + j] = p4r[x][y];
            s.im[1][i * LD + j] = p4i[x][y];
            bad |= !(finite_d(sr[x][y]) && finite_d(si[x][y]) && finite_d(er[x][y]) &&
                     finite_d(ei[x][y]) && finite_d(p3r[x][y]) && finite_d(p3i[x][y]) &&
                     finite_d(p4r[x][y]) && finite_d(p4i[x][y]));
            const T m3 = p3r[x][y] * p3r[x][y] + p3i[x][y] * p3i[x][y];
            const T m4 = p4r[x][y] * p4r[x][y] + p4i[x][y] * p4i[x][y];
            emax = fmax(emax, (double)fmax(m3, m4));
          }
        }
    }
    double gmax;
    bool gbad;
    block_max_any<NT>(emax, bad, s.sred, s.sflag, tid, gmax, gbad);
    if (gbad) {
      status = ST_NONFINITE;
      break;
    }
    if ((T)sqrt(gmax) < tol) {
      conv = true;
      break;
    }
    if constexpr (sizeof(T) == 4) {  // the mixed path's FP32 pass: give up on a stalled mode
      constexpr double r2 = NEGF_MIXED_FP32_STALL_RATIO * NEGF_MIXED_FP32_STALL_RATIO;
      if (it >= NEGF_MIXED_FP32_STALL_FROM && gmax > r2 * gprev) break;
      gprev = gmax;
    }
  }
  if (status == ST_OK && !conv) status = ST_NOCONV;
  if (tid == 0) {
    if (do1) {
      a.status[inst] = status;
      a.iters[inst] = it;
    }
    if constexpr (MIRROR) {
      if (do2) {
        a.status2[inst] = status;
        a.iters2[inst] = it;
      }
    }
  }
  if (status != ST_OK) return;
  const double w = a.omega[inst];
  if (do1)
    sr_write<T, NMAX, TM>(sr, si, t, s, n, zr, zi, w, a.g_out + (size_t)inst * nn * 2, a.dos_out,
                          a.status, inst, tid);
  if constexpr (MIRROR) {
    if (do2)
      sr_write<T, NMAX, TM>(mr, mi, t, s, n, zr, zi, w, a.g_out2 + (size_t)inst * nn * 2,
                            a.dos_out2, a.status2, inst, tid);
  }
}

// g = (z - e_s)^-1 from e_s held in the matrix-core ownership (FP64), and the
// surface DOS (omega / pi) tr(-2 Im g).
template <int NMAX, typename L>
__device__ __forceinline__ void sr_write_mm(const double (&er)[MOwn<NMAX, Cfg<NMAX>::NT>::TPW],
                                            const double (&ei)[MOwn<NMAX, Cfg<NMAX>::NT>::TPW],
                                            const MOwn<NMAX, Cfg<NMAX>::NT>& o, L& s, int n,
                                            double zr, double zi, double omega, double* out,
                                            double* dos, int32_t* status, int inst, int tid) {
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  __syncthreads();  // earlier reads of plane 2 are done
#pragma unroll
  for (int q = 0; q < MOwn<NMAX, NT>::TPW; ++q) {
    const int i = o.row(q), j = o.col(q);
    if (i < n && j < n) {
      s.re[2][i * LD + j] = (i == j ? zr : 0.0) - er[q];
      s.im[2][i * LD + j] = (i == j ? zi : 0.0) - ei[q];
    }
  }
  __syncthreads();
  const int st = gj_inverse<double, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid);
  if (st != ST_OK) {
    if (tid == 0) status[inst] = st;
    return;
  }
  gstore<double, NT>(out, n, s.re[2], s.im[2], LD, tid);
  if (tid == 0 && dos) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += -2.0 * s.im[2][i * LD + i];
    dos[inst] = (omega / M_PI) * tr;
  }
}

// The FP64 decimation with every block product on the matrix cores (mm.hpp): the
// same recurrence, step for step, as k_sancho_rubio; e, e_s (and the mirror's
// e~_s) are held in the matrix-core ownership.
template <int NMAX, bool MIRROR>
__global__ void __launch_bounds__(Cfg<NMAX>::NT)
    __attribute__((amdgpu_waves_per_eu(OccMM<NMAX, MIRROR>::value))) k_sr_mm(SrArgs a) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  using L = Lds<T, NMAX, NT, 3>;
  using M = MOwn<NMAX, NT>;
  constexpr int LD = L::LD, TPW = M::TPW;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  const bool do1 = !a.redo || a.status[inst] != ST_OK;
  bool do2 = false;
  if constexpr (MIRROR) do2 = !a.redo || a.status2[inst] != ST_OK;
  if (!do1 && !do2) return;
  const int n = a.n;
  const int ns = (n + 3) / 4;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;

  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  // alpha = d01 -> [0], beta = d10 -> [1]; g, then T1 = alpha g, then T2 = beta g -> [2]
  gload<T, NT>(blk + nn * 2, n, n, s.re[0], s.im[0], LD, tid);
  gload<T, NT>(blk + 2 * nn * 2, n, n, s.re[1], s.im[1], LD, tid);

  const M o(tid);
  T er[TPW], ei[TPW], sr[TPW], si[TPW];
  T mr[MIRROR ? TPW : 1], mi[MIRROR ? TPW : 1];
  {
    const double2* d00 = reinterpret_cast<const double2*>(blk);
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      double2 v = make_double2(0.0, 0.0);
      if (i < n && j < n) v = d00[i * n + j];
      er[q] = v.x;
      ei[q] = v.y;
      sr[q] = v.x;
      si[q] = v.y;
      if constexpr (MIRROR) {
        mr[q] = v.x;
        mi[q] = v.y;
      }
    }
  }
  double zr, zi;
  zsq(a.omega[inst], a.eta[inst], zr, zi);
  const T tol = a.tol;
  __syncthreads();

  int status = ST_OK;
  int it = 0;
  bool conv = false;
  while (it < a.maxiter) {
    ++it;
    // G = z - e
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      if (i < n && j < n) {
        s.re[2][i * LD + j] = (i == j ? zr : T(0)) - er[q];
        s.im[2][i * LD + j] = (i == j ? zi : T(0)) - ei[q];
      }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = st;
      break;
    }
    // T1 = a g, T2 = b g
    T t1r[TPW], t1i[TPW], t2r[TPW], t2i[TPW];
    mzero<TPW>(t1r, t1i);
    mzero<TPW>(t2r, t2i);
    mm2b<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], s.re[2], s.im[2], LD, ns, o, t1r,
                             t1i, t2r, t2i);
    __syncthreads();
    mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, t1r, t1i);
    __syncthreads();
    // P1 = T1 b (into e_s and e at once), P3 = T1 a (the next alpha)
    T p1r[TPW], p1i[TPW], p3r[TPW], p3i[TPW];
    mzero<TPW>(p1r, p1i);
    mzero<TPW>(p3r, p3i);
    mm2a<NMAX, NT>(s.re[2], s.im[2], s.re[1], s.im[1], s.re[0], s.im[0], LD, ns, o, p1r, p1i, p3r,
                   p3i);
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      sr[q] = sr[q] + p1r[q];
      si[q] = si[q] + p1i[q];
      er[q] = er[q] + p1r[q];
      ei[q] = ei[q] + p1i[q];
    }
    __syncthreads();
    mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, t2r, t2i);
    __syncthreads();
    // P2 = T2 a (into e), P4 = T2 b (the next beta)
    T p2r[TPW], p2i[TPW], p4r[TPW], p4i[TPW];
    mzero<TPW>(p2r, p2i);
    mzero<TPW>(p4r, p4i);
    mm2a<NMAX, NT>(s.re[2], s.im[2], s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, p2r, p2i, p4r,
                   p4i);
    __syncthreads();
    bool bad = false;
    double emax = 0.0;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      if (i < n && j < n) {
        er[q] = er[q] + p2r[q];
        ei[q] = ei[q] + p2i[q];
        if constexpr (MIRROR) {
          mr[q] = mr[q] + p2r[q];
          mi[q] = mi[q] + p2i[q];
          bad |= !(finite_d(mr[q]) && finite_d(mi[q]));
        }
        s.re[0][i * LD + j] = p3r[q];
        s.im[0][i * LD + j] = p3i[q];
        s.re[1][i * LD + j] = p4r[q];
        s.im[1][i * LD + j] = p4i[q];
        bad |= !(finite_d(sr[q]) && finite_d(si[q]) && finite_d(er[q]) && finite_d(ei[q]) &&
                 finite_d(p3r[q]) && finite_d(p3i[q]) && finite_d(p4r[q]) && finite_d(p4i[q]));
        const T m3 = p3r[q] * p3r[q] + p3i[q] * p3i[q];
        const T m4 = p4r[q] * p4r[q] + p4i[q] * p4i[q];
        emax = fmax(emax, fmax(m3, m4));
      }
    }
    double gmax;
    bool gbad;
    block_max_any<NT>(emax, bad, s.sred, s.sflag, tid, gmax, gbad);
    if (gbad) {
      status = ST_NONFINITE;
      break;
    }
    if (sqrt(gmax) < tol) {
      conv = true;
      break;
    }
  }
  if (status == ST_OK && !conv) status = ST_NOCONV;
  if (tid == 0) {
    if (do1) {
      a.status[inst] = status;
      a.iters[inst] = it;
    }
    if constexpr (MIRROR) {
      if (do2) {
        a.status2[inst] = status;
        a.iters2[inst] = it;
      }
    }
  }
  if (status != ST_OK) return;
  const double w = a.omega[inst];
  if (do1)
    sr_write_mm<NMAX>(sr, si, o, s, n, zr, zi, w, a.g_out + (size_t)inst * nn * 2, a.dos_out,
                      a.status, inst, tid);
  if constexpr (MIRROR) {
    if (do2)
      sr_write_mm<NMAX>(mr, mi, o, s, n, zr, zi, w, a.g_out2 + (size_t)inst * nn * 2, a.dos_out2,
                        a.status2, inst, tid);
  }
}

// ---------------------------------------------------------------------------
// FP64 completion of an FP32 decimation: Newton on the surface equation with FP32
// Stein corrections, and a certificate
// ---------------------------------------------------------------------------
//
// The retarded surface GF solves F(g) = I - (z - d00 - d01 g d10) g = 0, the fixed
// point of the decimation (src/negfpy/core/surface_gf.py:41-81).  From the FP32
// decimation's e_s, g = (z - e_s)^-1 in FP64, then each Newton step
//   R = F(g)                            (FP64: three products)
//   X - P X Q = g R,  P = g d01, Q = d10 g   (the linearised equation, a Stein equation)
//   g <- g + X
// where the Stein equation is solved in FP32 by doubling (Smith's method:
// X += P X Q, P <- P^2, Q <- Q^2) until |P| |Q| < 1e-10 or 64 doublings.  A final
// FP64 residual certifies the result: accepted when the relative residual
// |R| / (|M| |g|) < rtol, the last correction |X| / |g| < xtol (quadratic
// convergence reached) and g moved less than dtol from the FP32 decimation's g (the
// Newton root is the retarded one the decimation approached, not another solution of
// the quadratic equation).  Anything else -- a doubling that diverges because FP32
// cannot separate |lambda| = 1 - eta/v from the unit circle, a non-finite value, a
// singular pivot -- marks the instance ST_REDO for the FP64 decimation.
struct RefineArgs {
  int n;
  int n_inst;
  const double* lead;
  const int32_t* set_of;
  const double* omega;
  const double* eta;
  int steps;       // Newton steps
  double rtol, xtol, dtol;
  double* g_io;    // in: e_s, out: g_s (certified instances)
  double* dos_out;
  int32_t* status;
};

constexpr int STEIN_MAX_DOUBLINGS = 64;

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_refine(RefineArgs a) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  using L = Lds<T, NMAX, NT, 4>;  // [0] d01 [1] d10 [2] g [3] scratch (FP64)
  constexpr int LD = L::LD, MSZ = L::MSZ;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  // FP32 Stein matrices after the FP64 part: X, P, Q, W
  float* fb = reinterpret_cast<float*>(smem + ((L::bytes() + 15) & ~size_t(15)));
  float* fr[4];
  float* fi[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    fr[q] = fb + (2 * q) * MSZ;
    fi[q] = fb + (2 * q + 1) * MSZ;
  }
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  if (a.status[inst] != ST_OK) return;  // uniform per block
  const int n = a.n;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (size_t)a.set_of[inst] * 3 * nn * 2;
  double* io = a.g_io + (size_t)inst * nn * 2;
  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  lds_zero<float, NT>(reinterpret_cast<unsigned char*>(fb), sizeof(float) * 8 * MSZ, tid);
  __syncthreads();
  gload<T, NT>(blk + nn * 2, n, n, s.re[0], s.im[0], LD, tid);      // d01
  gload<T, NT>(blk + 2 * nn * 2, n, n, s.re[1], s.im[1], LD, tid);  // d10
  const Tile<TM> t(n, n, tid);
  double zr, zi;
  zsq(a.omega[inst], a.eta[inst], zr, zi);
  {
    const double2* es = reinterpret_cast<const double2*>(io);
    for (int e = tid; e < (int)nn; e += NT) {
      const int i = e / n, j = e - (e / n) * n;
      const double2 v = es[e];
      s.re[2][i * LD + j] = (i == j ? zr : 0.0) - v.x;
      s.im[2][i * LD + j] = (i == j ? zi : 0.0) - v.y;
    }
  }
  __syncthreads();
  int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid);
  if (st != ST_OK) {
    if (tid == 0) a.status[inst] = ST_REDO;
    return;
  }
  // the FP32 decimation's g, for the drift test at the end
  gstore<T, NT>(io, n, s.re[2], s.im[2], LD, tid);
  const double2* d00 = reinterpret_cast<const double2*>(blk);
  double rres = INFINITY, xrel = INFINITY;
  bool bad = false;
  for (int step = 0; step <= a.steps; ++step) {
    // T = d01 g -> [3]
    T xr[TM][TM], xi[TM][TM];
    tzero<T, TM>(xr, xi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, n, t.i0, t.j0, xr, xi);
    tstore<T, TM>(s.re[3], s.im[3], LD, n, n, t, xr, xi);
    __syncthreads();
    // M = (z - d00) - T d10 -> [3]
    tzero<T, TM>(xr, xi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[3], s.im[3], s.re[1], s.im[1], LD, n, t.i0, t.j0, xr, xi);
    double mn = 0.0;
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            const double2 d = d00[i * n + j];
            xr[x][y] = ((i == j ? zr : 0.0) - d.x) - xr[x][y];
            xi[x][y] = ((i == j ? zi : 0.0) - d.y) - xi[x][y];
            mn += xr[x][y] * xr[x][y] + xi[x][y] * xi[x][y];
          } else {
            xr[x][y] = 0.0;
            xi[x][y] = 0.0;
          }
        }
    }
    __syncthreads();
    tstore<T, TM>(s.re[3], s.im[3], LD, n, n, t, xr, xi);
    __syncthreads();
    // R = I - M g (tiles), norms of R and g
    tzero<T, TM>(xr, xi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[3], s.im[3], s.re[2], s.im[2], LD, n, t.i0, t.j0, xr, xi);
    double rn = 0.0, gn = 0.0;
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            xr[x][y] = (i == j ? 1.0 : 0.0) - xr[x][y];
            xi[x][y] = -xi[x][y];
            rn += xr[x][y] * xr[x][y] + xi[x][y] * xi[x][y];
            const double gr = s.re[2][i * LD + j], gi = s.im[2][i * LD + j];
            gn += gr * gr + gi * gi;
          }
        }
    }
    rn = block_sum<NT>(rn, s.sred, tid);
    mn = block_sum<NT>(mn, s.sred, tid);
    gn = block_sum<NT>(gn, s.sred, tid);
    rres = sqrt(rn / (mn * gn));
    if (!(rres < INFINITY)) {
      bad = true;
      break;
    }
    if (step == a.steps) break;
    // C = g R -> X (FP32); P = g d01, Q = d10 g (FP32)
    tstore<T, TM>(s.re[3], s.im[3], LD, n, n, t, xr, xi);
    __syncthreads();
    {
      T cr[TM][TM], ci[TM][TM];
      tzero<T, TM>(cr, ci);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[2], s.im[2], s.re[3], s.im[3], LD, n, t.i0, t.j0, cr, ci);
      float fcr[TM][TM], fci[TM][TM];
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          fcr[x][y] = (float)cr[x][y];
          fci[x][y] = (float)ci[x][y];
        }
      tstore<float, TM>(fr[0], fi[0], LD, n, n, t, fcr, fci);
      tzero<T, TM>(cr, ci);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, n, t.i0, t.j0, cr, ci);
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          fcr[x][y] = (float)cr[x][y];
          fci[x][y] = (float)ci[x][y];
        }
      tstore<float, TM>(fr[1], fi[1], LD, n, n, t, fcr, fci);
      tzero<T, TM>(cr, ci);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, n, t.i0, t.j0, cr, ci);
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          fcr[x][y] = (float)cr[x][y];
          fci[x][y] = (float)ci[x][y];
        }
      tstore<float, TM>(fr[2], fi[2], LD, n, n, t, fcr, fci);
    }
    __syncthreads();
    // Stein doubling in FP32: X += P X Q, P <- P^2, Q <- Q^2
    for (int k = 0; k < STEIN_MAX_DOUBLINGS; ++k) {
      float wr[TM][TM], wi[TM][TM];
      tzero<float, TM>(wr, wi);
      if (t.own) mac<float, TM, OPN, OPN>(fr[1], fi[1], fr[0], fi[0], LD, n, t.i0, t.j0, wr, wi);
      tstore<float, TM>(fr[3], fi[3], LD, n, n, t, wr, wi);  // W = P X
      __syncthreads();
      float vr[TM][TM], vi[TM][TM], pr[TM][TM], pi[TM][TM], qr[TM][TM], qi[TM][TM];
      tzero<float, TM>(vr, vi);
      tzero<float, TM>(pr, pi);
      tzero<float, TM>(qr, qi);
      if (t.own) {
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TM; ++y) {
            const int i = t.i0 + x, j = t.j0 + y;
            if (i < n && j < n) {
              vr[x][y] = fr[0][i * LD + j];
              vi[x][y] = fi[0][i * LD + j];
            }
          }
        mac<float, TM, OPN, OPN>(fr[3], fi[3], fr[2], fi[2], LD, n, t.i0, t.j0, vr, vi);
        mac<float, TM, OPN, OPN>(fr[1], fi[1], fr[1], fi[1], LD, n, t.i0, t.j0, pr, pi);
        mac<float, TM, OPN, OPN>(fr[2], fi[2], fr[2], fi[2], LD, n, t.i0, t.j0, qr, qi);
      }
      double pn = 0.0, qn = 0.0;
      bool nf = false;
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          pn += (double)pr[x][y] * pr[x][y] + (double)pi[x][y] * pi[x][y];
          qn += (double)qr[x][y] * qr[x][y] + (double)qi[x][y] * qi[x][y];
          nf |= !(isfinite(vr[x][y]) && isfinite(vi[x][y]));
        }
      __syncthreads();
      tstore<float, TM>(fr[0], fi[0], LD, n, n, t, vr, vi);
      tstore<float, TM>(fr[1], fi[1], LD, n, n, t, pr, pi);
      tstore<float, TM>(fr[2], fi[2], LD, n, n, t, qr, qi);
      double pq, dummy;
      bool anynf;
      block_max_any<NT>(0.0, nf, s.sred, s.sflag, tid, dummy, anynf);
      pq = sqrt(block_sum<NT>(pn, s.sred, tid) * block_sum<NT>(qn, s.sred, tid));
      if (anynf || !(pq < 1e30)) {
        bad = true;
        break;
      }
      if (pq < 1e-10) break;
    }
    if (bad) break;
    // g <- g + X
    double xn = 0.0;
    __syncthreads();
    if (t.own) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            const double dr = fr[0][i * LD + j], di = fi[0][i * LD + j];
            s.re[2][i * LD + j] += dr;
            s.im[2][i * LD + j] += di;
            xn += dr * dr + di * di;
          }
        }
    }
    xn = block_sum<NT>(xn, s.sred, tid);
    xrel = sqrt(xn / gn);
  }
  // drift from the FP32 decimation's g
  double dn = 0.0, gn = 0.0;
  if (!bad && t.own) {
    const double2* g0 = reinterpret_cast<const double2*>(io);
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) {
          const double gr = s.re[2][i * LD + j], gi = s.im[2][i * LD + j];
          const double2 v = g0[i * n + j];
          dn += (gr - v.x) * (gr - v.x) + (gi - v.y) * (gi - v.y);
          gn += gr * gr + gi * gi;
        }
      }
  }
  dn = block_sum<NT>(dn, s.sred, tid);
  gn = block_sum<NT>(gn, s.sred, tid);
  const bool ok = !bad && rres < a.rtol && xrel < a.xtol && sqrt(dn / gn) < a.dtol;
  if (!ok) {
    if (tid == 0) a.status[inst] = ST_REDO;
    return;
  }
  gstore<T, NT>(io, n, s.re[2], s.im[2], LD, tid);
  if (tid == 0 && a.dos_out) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += -2.0 * s.im[2][i * LD + i];
    a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
  }
}

// ---------------------------------------------------------------------------
// Sigma = V g V^H embedded into an n_dev x n_dev end-layer block (FP64)
// ---------------------------------------------------------------------------

struct SigmaArgs {
  int n_lead, n_c, n_dev, n_inst;
  const double* v;  // [n_sets][n_c][n_lead]
  const int32_t* set_of;
  const double* g;    // [n_inst][n_lead][n_lead]
  const int32_t* idx; // [n_c] positions inside the end layer
  const int32_t* status_in;  // may be null; skip instances with status != 0
  double* sigma_out;  // [n_inst][n_dev][n_dev]
};

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_self_energy(SigmaArgs a) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  using L = Lds<T, NMAX, NT, 3>;
  constexpr int LD = L::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  const int nl = a.n_lead, nc = a.n_c, nd = a.n_dev;
  double2* out = reinterpret_cast<double2*>(a.sigma_out + (size_t)inst * nd * nd * 2);
  if (a.status_in && a.status_in[inst] != ST_OK) return;
  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  gload<T, NT>(a.v + (size_t)a.set_of[inst] * nc * nl * 2, nc, nl, s.re[0], s.im[0], LD, tid);
  gload<T, NT>(a.g + (size_t)inst * nl * nl * 2, nl, nl, s.re[1], s.im[1], LD, tid);
  int* pos = s.gj->pv;  // inverse contact map: device-layer index -> contact row
  for (int i = tid; i < nd; i += NT) pos[i] = -1;
  __syncthreads();
  for (int c = tid; c < nc; c += NT) pos[a.idx[c]] = c;
  __syncthreads();
  {
    const Tile<TM> t(nc, nl, tid);
    T xr[TM][TM], xi[TM][TM];
    tzero<T, TM>(xr, xi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, nl, t.i0, t.j0, xr, xi);
    tstore<T, TM>(s.re[2], s.im[2], LD, nc, nl, t, xr, xi);
  }
  __syncthreads();
  {
    const Tile<TM> t(nc, nc, tid);
    T xr[TM][TM], xi[TM][TM];
    tzero<T, TM>(xr, xi);
    if (t.own) mac<T, TM, OPN, OPH>(s.re[2], s.im[2], s.re[0], s.im[0], LD, nl, t.i0, t.j0, xr, xi);
    __syncthreads();
    tstore<T, TM>(s.re[1], s.im[1], LD, nc, nc, t, xr, xi);
  }
  __syncthreads();
  for (int e = tid; e < nd * nd; e += NT) {
    const int i = e / nd, j = e - (e / nd) * nd;
    const int pi = pos[i], pj = pos[j];
    out[e] = (pi >= 0 && pj >= 0) ? make_double2(s.re[1][pi * LD + pj], s.im[1][pi * LD + pj])
                                  : make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Recursive Green's function sweep for G_{1N}
// ---------------------------------------------------------------------------

struct RgfArgs {
  int n, n_layers, n_inst;
  const double* onsite;    // [n_sets][L][n][n]
  const double* coupling;  // [n_sets][L-1][n][n]
  const int32_t* set_of;
  const double* omega;
  const double* eta_dev;
  const double* sigma_l;  // [n_inst][n][n]
  const double* sigma_r;
  const int32_t* status_in;  // lead status (skip when != 0); may be null
  double* g1n_out;           // [n_inst][n][n]
  int32_t* status;
  int redo;                  // 1: only instances whose device status is 12/13 are recomputed
};

template <typename T, int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT)
    __attribute__((amdgpu_waves_per_eu(Occ3<T, NMAX>::value))) k_rgf(RgfArgs a) {
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  using L = Lds<T, NMAX, NT, 3>;
  constexpr int LD = L::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
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
  // LDS: [0] g (inverted in place), then Y = g H; [1] H; [2] Q = G_{1,i} H_i
  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  const Tile<TM> t(n, n, tid);
  double zrd, zid;
  zsq(a.omega[inst], a.eta_dev[inst], zrd, zid);
  const T zr = (T)zrd, zi = (T)zid;
  T sr[TM][TM], si[TM][TM];
  tzero<T, TM>(sr, si);
  int status = ST_OK;
  for (int layer = 0; layer < nl; ++layer) {
    if (t.own) {
      const double2* d = ons + (size_t)layer * nn;
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            const double2 dv = d[i * n + j];
            s.re[0][i * LD + j] = (i == j ? zr : T(0)) - (T)dv.x - sr[x][y];
            s.im[0][i * LD + j] = (i == j ? zi : T(0)) - (T)dv.y - si[x][y];
          }
        }
    }
    // the contact self-energies of the end layers, in a coalesced pass (same
    // order of operations: ((z - D - S) - Sigma_L) - Sigma_R)
    if (layer == 0 || layer == nl - 1) {
      __syncthreads();
      for (int e = tid; e < (int)nn; e += NT) {
        const int i = e / n, j = e - (e / n) * n;
        T vr = s.re[0][i * LD + j], vi = s.im[0][i * LD + j];
        if (layer == 0) {
          vr -= (T)sgl[e].x;
          vi -= (T)sgl[e].y;
        }
        if (layer == nl - 1) {
          vr -= (T)sgr[e].x;
          vi -= (T)sgr[e].y;
        }
        s.re[0][i * LD + j] = vr;
        s.im[0][i * LD + j] = vi;
      }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    if (layer == nl - 1) {
      // G_1N = Q_{N-1} g_N (g_1 itself for one layer), straight to the output
      T gr[TM][TM], gi[TM][TM];
      tzero<T, TM>(gr, gi);
      if (nl == 1) {
        if (t.own) {
#pragma unroll
          for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int y = 0; y < TM; ++y) {
              const int i = t.i0 + x, j = t.j0 + y;
              if (i < n && j < n) {
                gr[x][y] = s.re[0][i * LD + j];
                gi[x][y] = s.im[0][i * LD + j];
              }
            }
        }
      } else if (t.own) {
        mac<T, TM, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, n, t.i0, t.j0, gr, gi);
      }
      // through LDS, so the output is written coalesced and its addresses are
      // formed only here
      tstore<T, TM>(s.re[1], s.im[1], LD, n, n, t, gr, gi);
      __syncthreads();
      gstore<T, NT>(a.g1n_out + (size_t)inst * nn * 2, n, s.re[1], s.im[1], LD, tid);
      break;
    }
    gload<T, NT>(cpl + (size_t)layer * nn * 2, n, n, s.re[1], s.im[1], LD, tid);
    __syncthreads();
    T yr[TM][TM], yi[TM][TM];
    tzero<T, TM>(yr, yi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, n, t.i0, t.j0, yr, yi);
    __syncthreads();  // every read of g is done: Y replaces it (and starts Q at layer 0)
    tstore<T, TM>(s.re[0], s.im[0], LD, n, n, t, yr, yi);
    if (layer == 0) tstore<T, TM>(s.re[2], s.im[2], LD, n, n, t, yr, yi);
    __syncthreads();
    tzero<T, TM>(sr, si);
    if (t.own) mac<T, TM, OPH, OPN>(s.re[1], s.im[1], s.re[0], s.im[0], LD, n, t.i0, t.j0, sr, si);
    if (layer > 0) {
      T qr[TM][TM], qi[TM][TM];
      tzero<T, TM>(qr, qi);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, n, t.i0, t.j0, qr, qi);
      __syncthreads();
      tstore<T, TM>(s.re[2], s.im[2], LD, n, n, t, qr, qi);
    } else {
      __syncthreads();
    }
  }
  if (tid == 0) a.status[inst] = status;
}

// Inputs of the contact self-energies and the trace when they run inside the sweep
// (k_rgf_mm<NMAX, true>): the self-energy arguments of each lead (their sigma_out is
// the end-layer block the sweep reads back, status_in the lead status) and T.
struct RgfFuse {
  SigmaArgs l, r;
  double* t_out;  // [n_inst]
};

// Sigma = V g V^H of one lead on the matrix cores, embedded into the n x n end-layer
// block at the contact positions idx and written to sigma_out (coalesced).  Uses all
// three LDS matrices and the pivot-row array as scratch and leaves them dirty; the
// lead block must fit the class (n_lead, n_c <= NMAX).  Every thread reaches every
// barrier; only `live` sub-blocks work (the persistent launch shares barriers between
// sub-blocks).  The persistent launch and the fused sweep share this arithmetic, so
// their self-energies are identical.
template <int NMAX, typename L>
__device__ __forceinline__ void sigma_mm(const SigmaArgs& sa, int inst, int n, L& s, int tid,
                                         bool live) {
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using M = MOwn<NMAX, NT>;
  constexpr int TPW = M::TPW;
  const M o(tid);
  const int nlead = sa.n_lead, nc = sa.n_c;
  const int nsl = (nlead + 3) / 4;
  int* pos = s.gj->pv;  // device-layer index -> contact row
  __syncthreads();  // the previous stage's last reads of these matrices are done
  if (live) s.zero(tid);
  __syncthreads();
  if (live) {
    gload<double, NT>(sa.v + (size_t)sa.set_of[inst] * nc * nlead * 2, nc, nlead, s.re[1],
                      s.im[1], LD, tid);
    gload<double, NT>(sa.g + (size_t)inst * nlead * nlead * 2, nlead, nlead, s.re[2], s.im[2],
                      LD, tid);
    for (int i = tid; i < n; i += NT) pos[i] = -1;
  }
  __syncthreads();
  double xr[TPW], xi[TPW];
  mzero<TPW>(xr, xi);
  if (live) {
    for (int c = tid; c < nc; c += NT) pos[sa.idx[c]] = c;
    mm<NMAX, NT, OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, nsl, o, xr, xi);  // V g
    // rows >= nc and columns >= nlead of V g are exact zeros (zero operand padding)
    mstore<NMAX, NT>(s.re[0], s.im[0], LD, nc > nlead ? nc : nlead, o, xr, xi);
  }
  __syncthreads();
  if (live) {
    mzero<TPW>(xr, xi);
    mm<NMAX, NT, OPN, OPH>(s.re[0], s.im[0], s.re[1], s.im[1], LD, nsl, o, xr, xi);  // (V g) V^H
    mstore<NMAX, NT>(s.re[2], s.im[2], LD, nc, o, xr, xi);
  }
  __syncthreads();
  if (live) {
    double2* out = reinterpret_cast<double2*>(sa.sigma_out + (size_t)inst * n * n * 2);
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e - (e / n) * n;
      const int pi = pos[i], pj = pos[j];
      out[e] = (pi >= 0 && pj >= 0) ? make_double2(s.re[2][pi * LD + pj], s.im[2][pi * LD + pj])
                                    : make_double2(0.0, 0.0);
    }
  }
  // the trace reads elements other threads of the workgroup wrote here: a workgroup
  // barrier orders them (one CU, one L1)
  __syncthreads();
}

// T = Re Tr[Gamma_L G Gamma_R G^H] = Re sum_ij U_ij conj(G_ij), W = G Gamma_R, U = Gamma_L W,
// Gamma = i (Sigma - Sigma^H), on the matrix cores.  G (n x n) is in LDS matrix [1] and in
// (gr, gi) in the matrix-core ownership; Gamma_L is built in [0] and Gamma_R in [2] from
// the self-energies in global memory, and [2] then takes W.  Returns T to every thread
// (0 where not live).  Shared by the fused sweep and the persistent launch.
template <int NMAX, typename L>
__device__ __forceinline__ double trace_mm(const double* sigma_l, const double* sigma_r, int inst,
                                           int n, L& s, const double (&gr)[MOwn<NMAX, Cfg<NMAX>::NT>::TPW],
                                           const double (&gi)[MOwn<NMAX, Cfg<NMAX>::NT>::TPW],
                                           int tid, bool live) {
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using M = MOwn<NMAX, NT>;
  constexpr int TPW = M::TPW;
  const M o(tid);
  const int ns = (n + 3) / 4;
  const int nn = n * n;
  if (live) {
    const double2* sl = reinterpret_cast<const double2*>(sigma_l + (size_t)inst * nn * 2);
    const double2* sr = reinterpret_cast<const double2*>(sigma_r + (size_t)inst * nn * 2);
    // Gamma_ij = i (S_ij - conj(S_ji)) = (-(Im S_ij + Im S_ji), Re S_ij - Re S_ji)
    for (int e = tid; e < nn; e += NT) {
      const int i = e / n, j = e - (e / n) * n;
      const double2 a1 = sl[e], a2 = sl[j * n + i];
      s.re[0][i * LD + j] = -(a1.y + a2.y);
      s.im[0][i * LD + j] = a1.x - a2.x;
      const double2 b1 = sr[e], b2 = sr[j * n + i];
      s.re[2][i * LD + j] = -(b1.y + b2.y);
      s.im[2][i * LD + j] = b1.x - b2.x;
    }
  }
  __syncthreads();
  double wr[TPW], wi[TPW];
  mzero<TPW>(wr, wi);
  if (live) mm<NMAX, NT, OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, ns, o, wr, wi);
  __syncthreads();
  if (live) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, wr, wi);
  __syncthreads();
  double acc = 0.0;
  if (live) {
    mzero<TPW>(wr, wi);
    mm<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, ns, o, wr, wi);
#pragma unroll
    for (int q = 0; q < TPW; ++q)
      if (o.row(q) < n && o.col(q) < n) acc = __fma_rn(wi[q], gi[q], __fma_rn(wr[q], gr[q], acc));
  }
  return block_sum<NT>(acc, s.sred, tid);
}

// The same sweep in FP64 with every block product on the matrix cores (mm.hpp):
// Y = g H, then S = H^H Y and Q = Q Y from one pass over Y.  The Schur term S
// stays in registers in the matrix-core ownership.  FUSE: the launch also forms
// both contact self-energies first (sigma_mm, into a.sigma_l / a.sigma_r) and, after
// G_1N, T = Re Tr[Gamma_L G_1N Gamma_R G_1N^H] as Re sum_ij U_ij conj(G_ij) with
// W = G_1N Gamma_R and U = Gamma_L W (two more matrix-core products), so the lead
// surface GFs go in and T comes out of one kernel; G_1N is still written when
// a.g1n_out is set.  The lead status is merged as negf_merge_status does (left first).
template <int NMAX, bool FUSE>
__global__ void __launch_bounds__(Cfg<NMAX>::NT)
    __attribute__((amdgpu_waves_per_eu(Occ3<double, NMAX>::value)))
    k_rgf_mm(RgfArgs a, RgfFuse f) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  using L = Lds<T, NMAX, NT, 3>;
  using M = MOwn<NMAX, NT>;
  constexpr int LD = L::LD, TPW = M::TPW;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  if (a.redo && a.status[inst] != 10 + ST_NONFINITE && a.status[inst] != 10 + ST_SINGULAR) return;
  if constexpr (FUSE) {
    const int stl = f.l.status_in[inst], str = f.r.status_in[inst];
    const int lead_st = stl != ST_OK ? stl : str;
    if (lead_st != ST_OK) {
      if (tid == 0) {
        a.status[inst] = lead_st;
        f.t_out[inst] = NAN;
      }
      return;
    }
  } else if (a.status_in && a.status_in[inst] != ST_OK) {
    if (tid == 0) a.status[inst] = a.status_in[inst];
    return;
  }
  const int n = a.n, nl = a.n_layers;
  const int ns = (n + 3) / 4;
  const size_t nn = (size_t)n * n;
  const size_t set = (size_t)a.set_of[inst];
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  // LDS: [0] g (inverted in place), then Y = g H; [1] H; [2] Q = G_{1,i} H_i
  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  if constexpr (FUSE) {
    sigma_mm<NMAX>(f.l, inst, n, s, tid, true);
    sigma_mm<NMAX>(f.r, inst, n, s, tid, true);
    lds_zero<T, NT>(smem, L::mat_bytes(), tid);
    __syncthreads();
  }
  const M o(tid);
  double zrd, zid;
  zsq(a.omega[inst], a.eta_dev[inst], zrd, zid);
  const T zr = zrd, zi = zid;
  T sr[TPW], si[TPW];
  mzero<TPW>(sr, si);
  int status = ST_OK;
  for (int layer = 0; layer < nl; ++layer) {
    {
      const double2* d = ons + (size_t)layer * nn;
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const int i = o.row(q), j = o.col(q);
        if (i < n && j < n) {
          const double2 dv = d[i * n + j];
          s.re[0][i * LD + j] = (i == j ? zr : T(0)) - dv.x - sr[q];
          s.im[0][i * LD + j] = (i == j ? zi : T(0)) - dv.y - si[q];
        }
      }
    }
    if (layer == 0 || layer == nl - 1) {
      __syncthreads();
      for (int e = tid; e < (int)nn; e += NT) {
        const int i = e / n, j = e - (e / n) * n;
        T vr = s.re[0][i * LD + j], vi = s.im[0][i * LD + j];
        if (layer == 0) {
          vr -= sgl[e].x;
          vi -= sgl[e].y;
        }
        if (layer == nl - 1) {
          vr -= sgr[e].x;
          vi -= sgr[e].y;
        }
        s.re[0][i * LD + j] = vr;
        s.im[0][i * LD + j] = vi;
      }
    }
    // this layer's coupling block: its loads stay in flight through the inversion
    constexpr int HPT = (NMAX * NMAX + NT - 1) / NT;
    double2 hv[HPT];
    if (layer < nl - 1) {
      const double2* h = reinterpret_cast<const double2*>(cpl) + (size_t)layer * nn;
#pragma unroll
      for (int r = 0; r < HPT; ++r) {
        const int e = tid + r * NT;
        hv[r] = e < (int)nn ? h[e] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    if (layer == nl - 1) {
      // G_1N = Q_{N-1} g_N (g_1 itself for one layer), through LDS to a coalesced store
      T gr[TPW], gi[TPW];
      mzero<TPW>(gr, gi);
      if (nl > 1) {
        mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o, gr, gi);
      } else {
#pragma unroll
        for (int q = 0; q < TPW; ++q) {
          const int i = o.row(q), j = o.col(q);
          if (i < n && j < n) {
            gr[q] = s.re[0][i * LD + j];
            gi[q] = s.im[0][i * LD + j];
          }
        }
        __syncthreads();
      }
      mstore<NMAX, NT>(s.re[1], s.im[1], LD, n, o, gr, gi);
      __syncthreads();
      if (a.g1n_out) gstore<T, NT>(a.g1n_out + (size_t)inst * nn * 2, n, s.re[1], s.im[1], LD, tid);
      break;
    }
#pragma unroll
    for (int r = 0; r < HPT; ++r) {
      const int e = tid + r * NT;
      if (e < (int)nn) {
        const int i = e / n, j = e - (e / n) * n;
        s.re[1][i * LD + j] = hv[r].x;
        s.im[1][i * LD + j] = hv[r].y;
      }
    }
    __syncthreads();
    T yr[TPW], yi[TPW];
    mzero<TPW>(yr, yi);
    mm<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, yr, yi);
    __syncthreads();  // every read of g is done: Y replaces it (and starts Q at layer 0)
    mstore<NMAX, NT>(s.re[0], s.im[0], LD, n, o, yr, yi);
    if (layer == 0) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, yr, yi);
    __syncthreads();
    mzero<TPW>(sr, si);
    if (layer > 0) {
      T qr[TPW], qi[TPW];
      mzero<TPW>(qr, qi);
      mm2b<NMAX, NT, OPH, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o, sr,
                               si, qr, qi);
      __syncthreads();
      mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, qr, qi);
    } else {
      mm<NMAX, NT, OPH, OPN>(s.re[1], s.im[1], s.re[0], s.im[0], LD, ns, o, sr, si);
      __syncthreads();
    }
  }
  if constexpr (FUSE) {
    // the trace, outside the layer loop so its registers do not add to the loop's: G_1N
    // back from LDS matrix [1] into the matrix-core ownership
    const bool ok = status == ST_OK;
    T gr[TPW], gi[TPW];
    mzero<TPW>(gr, gi);
    if (ok) {
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const int i = o.row(q), j = o.col(q);
        if (i < n && j < n) {
          gr[q] = s.re[1][i * LD + j];
          gi[q] = s.im[1][i * LD + j];
        }
      }
    }
    const double tot = trace_mm<NMAX>(a.sigma_l, a.sigma_r, inst, n, s, gr, gi, tid, ok);
    if (tid == 0) f.t_out[inst] = ok ? tot : NAN;
  }
  if (tid == 0) a.status[inst] = status;
}

// ---------------------------------------------------------------------------
// Full device Green's function G = A^-1 (FP64), all L^2 blocks
// ---------------------------------------------------------------------------

struct DgfArgs {
  int n, n_layers, n_inst;
  const double* onsite;
  const double* coupling;
  const int32_t* set_of;
  const double* omega;
  const double* eta_dev;
  const double* sigma_l;
  const double* sigma_r;
  double* g_out;  // [n_inst][L n][L n]
  int32_t* status;
};

// LDS <-> global block moves for an n x n block with global leading dim `gld` (complex).
template <typename T, int NT>
__device__ __forceinline__ void bload(const double* __restrict__ g, int gld, int n, T* re, T* im,
                                      int ld, int tid) {
  const double2* s = reinterpret_cast<const double2*>(g);
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 v = s[(size_t)i * gld + j];
    re[i * ld + j] = (T)v.x;
    im[i * ld + j] = (T)v.y;
  }
}

template <typename T, int TM>
__device__ __forceinline__ void tstore_glb(double* __restrict__ g, int gld, int n, const Tile<TM>& t,
                                           const T (&cr)[TM][TM], const T (&ci)[TM][TM]) {
  if (!t.own) return;
  double2* d = reinterpret_cast<double2*>(g);
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TM; ++y) {
      const int i = t.i0 + x, j = t.j0 + y;
      if (i < n && j < n) d[(size_t)i * gld + j] = make_double2((double)cr[x][y], (double)ci[x][y]);
    }
}

template <int NMAX, class L>
__device__ void dgf_body(const DgfArgs& a, int inst, L& s, int tid) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  constexpr int LD = L::LD;
  const int n = a.n, nl = a.n_layers, dim = n * nl;
  const size_t nn = (size_t)n * n;
  const size_t set = (size_t)a.set_of[inst];
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  double* G = a.g_out + (size_t)inst * dim * dim * 2;
  auto blk = [&](int bi, int bj) { return G + ((size_t)bi * n * dim + (size_t)bj * n) * 2; };
  s.zero(tid);
  __syncthreads();
  const Tile<TM> t(n, n, tid);
  double zr, zi;
  zsq(a.omega[inst], a.eta_dev[inst], zr, zi);
  T sr[TM][TM], si[TM][TM];
  tzero<T, TM>(sr, si);
  int status = ST_OK;
  // forward sweep: left-connected g_i stored in the diagonal blocks of G
  for (int layer = 0; layer < nl && status == ST_OK; ++layer) {
    if (t.own) {
      const double2* d = ons + (size_t)layer * nn;
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int i = t.i0 + x, j = t.j0 + y;
          if (i < n && j < n) {
            const double2 dv = d[i * n + j];
            T vr = (i == j ? zr : 0.0) - dv.x - sr[x][y];
            T vi = (i == j ? zi : 0.0) - dv.y - si[x][y];
            if (layer == 0) {
              vr -= sgl[i * n + j].x;
              vi -= sgl[i * n + j].y;
            }
            if (layer == nl - 1) {
              vr -= sgr[i * n + j].x;
              vi -= sgr[i * n + j].y;
            }
            s.re[0][i * LD + j] = vr;
            s.im[0][i * LD + j] = vi;
          }
        }
    }
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    {
      T gr[TM][TM], gi[TM][TM];
      tzero<T, TM>(gr, gi);
      if (t.own) {
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TM; ++y) {
            const int i = t.i0 + x, j = t.j0 + y;
            if (i < n && j < n) {
              gr[x][y] = s.re[0][i * LD + j];
              gi[x][y] = s.im[0][i * LD + j];
            }
          }
      }
      tstore_glb<T, TM>(blk(layer, layer), dim, n, t, gr, gi);
    }
    if (layer == nl - 1) break;
    gload<T, NT>(cpl + (size_t)layer * nn * 2, n, n, s.re[2], s.im[2], LD, tid);
    __syncthreads();
    T yr[TM][TM], yi[TM][TM];
    tzero<T, TM>(yr, yi);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, n, t.i0, t.j0, yr, yi);
    tstore<T, TM>(s.re[3], s.im[3], LD, n, n, t, yr, yi);
    __syncthreads();
    tzero<T, TM>(sr, si);
    if (t.own) mac<T, TM, OPH, OPN>(s.re[2], s.im[2], s.re[3], s.im[3], LD, n, t.i0, t.j0, sr, si);
    __syncthreads();
  }
  if (tid == 0) a.status[inst] = status;
  if (status != ST_OK) return;
  __threadfence_block();
  __syncthreads();
  // backward pass: rows/columns i from the finished blocks of i + 1
  for (int i = nl - 2; i >= 0; --i) {
    bload<T, NT>(blk(i, i), dim, n, s.re[0], s.im[0], LD, tid);  // g_i
    gload<T, NT>(cpl + (size_t)i * nn * 2, n, n, s.re[2], s.im[2], LD, tid);  // H_i
    __syncthreads();
    T ur[TM][TM], ui[TM][TM];
    tzero<T, TM>(ur, ui);
    if (t.own) mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, n, t.i0, t.j0, ur, ui);
    T vr[TM][TM], vi[TM][TM];
    tzero<T, TM>(vr, vi);
    if (t.own) mac<T, TM, OPH, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, n, t.i0, t.j0, vr, vi);
    tstore<T, TM>(s.re[1], s.im[1], LD, n, n, t, ur, ui);  // U = g_i H_i
    tstore<T, TM>(s.re[3], s.im[3], LD, n, n, t, vr, vi);  // V = H_i^H g_i
    __syncthreads();
    for (int j = i + 1; j < nl; ++j) {
      // G_{i,j} = U G_{i+1,j}
      bload<T, NT>(blk(i + 1, j), dim, n, s.re[2], s.im[2], LD, tid);
      __syncthreads();
      T cr[TM][TM], ci[TM][TM];
      tzero<T, TM>(cr, ci);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, n, t.i0, t.j0, cr, ci);
      tstore_glb<T, TM>(blk(i, j), dim, n, t, cr, ci);
      __syncthreads();
      // G_{j,i} = G_{j,i+1} V
      bload<T, NT>(blk(j, i + 1), dim, n, s.re[2], s.im[2], LD, tid);
      __syncthreads();
      tzero<T, TM>(cr, ci);
      if (t.own) mac<T, TM, OPN, OPN>(s.re[2], s.im[2], s.re[3], s.im[3], LD, n, t.i0, t.j0, cr, ci);
      tstore_glb<T, TM>(blk(j, i), dim, n, t, cr, ci);
      __threadfence_block();
      __syncthreads();
    }
    // G_{i,i} = g_i + U G_{i+1,i}
    bload<T, NT>(blk(i + 1, i), dim, n, s.re[2], s.im[2], LD, tid);
    __syncthreads();
    T cr[TM][TM], ci[TM][TM];
    tzero<T, TM>(cr, ci);
    if (t.own) {
      mac<T, TM, OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, n, t.i0, t.j0, cr, ci);
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TM; ++y) {
          const int r = t.i0 + x, c = t.j0 + y;
          if (r < n && c < n) {
            cr[x][y] += s.re[0][r * LD + c];
            ci[x][y] += s.im[0][r * LD + c];
          }
        }
    }
    tstore_glb<T, TM>(blk(i, i), dim, n, t, cr, ci);
    __threadfence_block();
    __syncthreads();
  }
}

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_device_green(DgfArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Lds<double, NMAX, Cfg<NMAX>::NT, 4> s(smem);
  dgf_body<NMAX>(a, blockIdx.x, s, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Adjoint: T and dT/dD for every device block (FP64)
// ---------------------------------------------------------------------------
//
// With A = z - D - Sigma_L - Sigma_R, G = A^-1 and W = Gamma_R G_1N^H Gamma_L,
// dT = 2 Re Tr[W dG_1N] and dG_1N = sum_ij G_1i dD_ij G_jN, so
//   dT/dD_ij (as d/dRe + i d/dIm) = 2 (G_jN W G_1i)^H.
// Row 1 of G comes from a right-connected sweep (v_0 = G_11, v_i = v_{i-1} R_i,
// R_i = C_{i-1} gR_i), column N from the left-connected sweep (u_{N-1} = G_NN,
// u_i = Y_i u_{i+1}, Y_i = gL_i C_i).  Onsite block i gets 2 w (u_i P_i)^H and
// coupling C_i (which enters D as C_i above and C_i^H below the diagonal) gets
// 2 w ((u_{i+1} P_i)^H + u_i P_{i+1}), with P_i = W v_i and w the caller's
// weight of instance T (its dL/dT).

struct AdjArgs {
  int n, n_layers, n_inst;
  const double* onsite;
  const double* coupling;
  const int32_t* set_of;
  const double* omega;
  const double* eta_dev;
  const double* sigma_l;
  const double* sigma_r;
  const int32_t* status_in;  // lead status; may be null
  const double* weight;      // [n_inst] dL/dT; may be null (weight 1)
  double* work;              // [n_inst][2L-1][n][n] complex
  double* t_out;             // [n_inst]
  double* grad_onsite;       // [n_inst][L][n][n] complex
  double* grad_coupling;     // [n_inst][L-1][n][n] complex
  int32_t* status;
  double* g1n_out;           // [n_inst][n][n] G_1N; may be null
  int ckpt;                  // 0: store every Y_i; K > 0: keep the left sweep's state every K
                             // layers and recompute Y_i per segment in the gradient pass
  int per_set;               // 1: gradients summed into [n_sets][..] blocks (atomic adds)
  double* grad_ends_out;     // [n_inst][2][n][n] onsite gradients of layers 1 and N; may be null
};

// Workspace blocks (n x n complex) of one adjoint instance: the L row-1 blocks (R_i, then
// v_i, then P_i) and either the L - 1 left-sweep blocks Y_i or, with checkpoints every K
// layers, ceil(L / K) carried states, a K-block segment buffer and one block for u.
__host__ __device__ inline size_t adj_blocks(int n_layers, int ckpt) {
  if (ckpt <= 0) return (size_t)(2 * n_layers - 1);
  return (size_t)n_layers + (size_t)((n_layers + ckpt - 1) / ckpt) + (size_t)ckpt + 1;
}

// Ownership of the n x n results of the adjoint's block products: a TM x TM register
// tile per thread with FMA loops on the vector ALUs (Tile, mac), or the matrix-core
// tile-groups of mm.hpp (MOwn, mm: every lane of the workgroup takes part, and the LDS
// operands must be zero outside [0, n)^2).  R owned elements per thread, element q at
// (i, j) when at(q, i, j).
template <int NMAX, bool MM>
struct AdjOwn;
template <int NMAX>
struct AdjOwn<NMAX, false> {
  static constexpr int TM = Cfg<NMAX>::TM, R = TM * TM;
  Tile<TM> t;
  int n;
  __device__ AdjOwn(int n_, int tid) : t(n_, n_, tid), n(n_) {}
  __device__ __forceinline__ bool at(int q, int& i, int& j) const {
    i = t.i0 + q / TM;
    j = t.j0 + q % TM;
    return t.own && i < n && j < n;
  }
  template <int OPA, int OPB>
  __device__ __forceinline__ void mul(const double* ar, const double* ai, const double* br,
                                      const double* bi, int ld, double (&cr)[R],
                                      double (&ci)[R]) const {
    if (t.own)
      mac<double, TM, OPA, OPB>(ar, ai, br, bi, ld, n, t.i0, t.j0,
                                reinterpret_cast<double (&)[TM][TM]>(cr),
                                reinterpret_cast<double (&)[TM][TM]>(ci));
  }
};
template <int NMAX>
struct AdjOwn<NMAX, true> {
  using M = MOwn<NMAX, Cfg<NMAX>::NT>;
  static constexpr int R = M::TPW;
  M o;
  int n, ns;
  __device__ AdjOwn(int n_, int tid) : o(tid), n(n_), ns((n_ + 3) / 4) {}
  __device__ __forceinline__ bool at(int q, int& i, int& j) const {
    i = o.row(q);
    j = o.col(q);
    return i < n && j < n;
  }
  template <int OPA, int OPB>
  __device__ __forceinline__ void mul(const double* ar, const double* ai, const double* br,
                                      const double* bi, int ld, double (&cr)[R],
                                      double (&ci)[R]) const {
    mm<NMAX, Cfg<NMAX>::NT, OPA, OPB>(ar, ai, br, bi, ld, ns, o, cr, ci);
  }
};

template <int NMAX, class L, bool MM>
__device__ void adj_body(const AdjArgs& a, int inst, L& s, int tid) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using O = AdjOwn<NMAX, MM>;
  constexpr int R = O::R;
  if (a.status_in && a.status_in[inst] != ST_OK) {
    if (tid == 0) {
      a.status[inst] = a.status_in[inst];
      a.t_out[inst] = NAN;
    }
    return;
  }
  const int n = a.n, nl = a.n_layers;
  const size_t nn = (size_t)n * n;
  const size_t set = (size_t)a.set_of[inst];
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  const int kc = a.ckpt;
  double* vw = a.work + (size_t)inst * adj_blocks(nl, kc) * nn * 2;  // R_i, then v_i, then P_i
  double* yw = vw + (size_t)nl * nn * 2;  // Y_i, or the left sweep's checkpoints
  double* ybuf = yw + (size_t)(kc > 0 ? (nl + kc - 1) / kc : 0) * nn * 2;  // Y of a segment
  double* usave = ybuf + (size_t)kc * nn * 2;                               // u across it
  const size_t gi = a.per_set ? (size_t)a.set_of[inst] : (size_t)inst;
  double* go = a.grad_onsite + gi * nl * nn * 2;
  double* gc = a.grad_coupling + gi * (nl > 1 ? nl - 1 : 0) * nn * 2;
  auto blk = [&](double* base, int i) { return base + (size_t)i * nn * 2; };

  const double w2 = 2.0 * (a.weight ? a.weight[inst] : 1.0);
  s.zero(tid);
  __syncthreads();
  const O t(n, tid);
  auto ozero = [](T (&xr)[R], T (&xi)[R]) {
#pragma unroll
    for (int q = 0; q < R; ++q) xr[q] = xi[q] = T(0);
  };
  auto put_lds = [&](T* re, T* im, const T (&xr)[R], const T (&xi)[R]) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i, j;
      if (t.at(q, i, j)) {
        re[i * LD + j] = xr[q];
        im[i * LD + j] = xi[q];
      }
    }
  };
  auto put_glb = [&](double* g, const T (&xr)[R], const T (&xi)[R]) {
    double2* d = reinterpret_cast<double2*>(g);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i, j;
      if (t.at(q, i, j)) d[(size_t)i * n + j] = make_double2(xr[q], xi[q]);
    }
  };
  // a gradient block: stored per instance, or added into its set's (instances of one set
  // run concurrently: atomic adds)
  auto gput = [&](double* g, const T (&cr)[R], const T (&ci)[R]) {
    if (!a.per_set) {
      put_glb(g, cr, ci);
      return;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i, j;
      if (t.at(q, i, j)) {
        atomicAdd(g + ((size_t)i * n + j) * 2, cr[q]);
        atomicAdd(g + ((size_t)i * n + j) * 2 + 1, ci[q]);
      }
    }
  };
  double zr, zi;
  zsq(a.omega[inst], a.eta_dev[inst], zr, zi);
  // A_ii - S into LDS matrix 0
  auto build = [&](int layer, const T (&sr)[R], const T (&si)[R]) {
    const double2* d = ons + (size_t)layer * nn;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i, j;
      if (t.at(q, i, j)) {
        const double2 dv = d[i * n + j];
        T vr = (i == j ? zr : 0.0) - dv.x - sr[q];
        T vi = (i == j ? zi : 0.0) - dv.y - si[q];
        if (layer == 0) {
          vr -= sgl[i * n + j].x;
          vi -= sgl[i * n + j].y;
        }
        if (layer == nl - 1) {
          vr -= sgr[i * n + j].x;
          vi -= sgr[i * n + j].y;
        }
        s.re[0][i * LD + j] = vr;
        s.im[0][i * LD + j] = vi;
      }
    }
  };
  int status = ST_OK;
  T sr[R], si[R];
  T cr[R], ci[R];

  // 1. right-connected sweep: gR_i, R_i = C_{i-1} gR_i (stored), ends with gR_0 = G_11
  ozero(sr, si);
  for (int i = nl - 1; i >= 0; --i) {
    build(i, sr, si);
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid);
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
    if (i == 0) break;
    gload<T, NT>(cpl + (size_t)(i - 1) * nn * 2, n, n, s.re[2], s.im[2], LD, tid);
    __syncthreads();
    ozero(cr, ci);
    t.template mul<OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, cr, ci);
    put_lds(s.re[3], s.im[3], cr, ci);
    put_glb(blk(vw, i), cr, ci);
    __syncthreads();
    ozero(sr, si);
    t.template mul<OPN, OPH>(s.re[3], s.im[3], s.re[2], s.im[2], LD, sr, si);
    __syncthreads();
  }
  if (status != ST_OK) {
    if (tid == 0) {
      a.status[inst] = status;
      a.t_out[inst] = NAN;
    }
    return;
  }
  // 2. row 1 of G: v_0 = G_11 (matrix 0), v_i = v_{i-1} R_i
  ozero(cr, ci);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    int r, c;
    if (t.at(q, r, c)) {
      cr[q] = s.re[0][r * LD + c];
      ci[q] = s.im[0][r * LD + c];
    }
  }
  put_glb(blk(vw, 0), cr, ci);
  __threadfence_block();
  __syncthreads();
  for (int i = 1; i < nl; ++i) {
    bload<T, NT>(blk(vw, i), n, n, s.re[1], s.im[1], LD, tid);
    __syncthreads();
    ozero(cr, ci);
    t.template mul<OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, cr, ci);
    __syncthreads();
    put_lds(s.re[0], s.im[0], cr, ci);
    put_glb(blk(vw, i), cr, ci);
    __threadfence_block();
    __syncthreads();
  }
  if (a.g1n_out) gstore<T, NT>(a.g1n_out + (size_t)inst * nn * 2, n, s.re[0], s.im[0], LD, tid);
  // 3. Gamma_L (matrix 1), Gamma_R (matrix 2); W = Gamma_R G_1N^H Gamma_L; T = Re Tr[G_1N W]
  for (int e = tid; e < (int)nn; e += NT) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 a1 = sgl[i * n + j], a2 = sgl[j * n + i];
    s.re[1][i * LD + j] = -(a1.y + a2.y);
    s.im[1][i * LD + j] = a1.x - a2.x;
    const double2 b1 = sgr[i * n + j], b2 = sgr[j * n + i];
    s.re[2][i * LD + j] = -(b1.y + b2.y);
    s.im[2][i * LD + j] = b1.x - b2.x;
  }
  __syncthreads();
  ozero(cr, ci);
  t.template mul<OPH, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, cr, ci);
  put_lds(s.re[3], s.im[3], cr, ci);
  __syncthreads();
  ozero(cr, ci);
  t.template mul<OPN, OPN>(s.re[2], s.im[2], s.re[3], s.im[3], LD, cr, ci);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    int r, c;
    if (t.at(q, r, c)) acc += cr[q] * s.re[0][c * LD + r] - ci[q] * s.im[0][c * LD + r];
  }
  const double tval = block_sum<NT>(acc, s.sred, tid);
  if (tid == 0) a.t_out[inst] = tval;
  put_lds(s.re[1], s.im[1], cr, ci);  // W
  __syncthreads();
  // 4. P_i = W v_i, in place
  for (int i = 0; i < nl; ++i) {
    bload<T, NT>(blk(vw, i), n, n, s.re[2], s.im[2], LD, tid);
    __syncthreads();
    ozero(cr, ci);
    t.template mul<OPN, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], LD, cr, ci);
    __syncthreads();
    put_glb(blk(vw, i), cr, ci);
  }
  __threadfence_block();
  __syncthreads();
  // 5. left-connected sweep: gL_i, Y_i = gL_i C_i (stored, or with checkpoints only the
  //    carried state S entering every K-th layer), ends with gL_{N-1} = G_NN
  // one step of it: layer i from the carried S (sr, si); Y_i to ydst (if set); the next S
  auto left_step = [&](int i, double* ydst) -> int {
    build(i, sr, si);
    __syncthreads();
    const int st = gj_inverse<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid);
    if (st != ST_OK || i == nl - 1) return st;
    gload<T, NT>(cpl + (size_t)i * nn * 2, n, n, s.re[2], s.im[2], LD, tid);
    __syncthreads();
    ozero(cr, ci);
    t.template mul<OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, cr, ci);
    put_lds(s.re[3], s.im[3], cr, ci);
    if (ydst) put_glb(ydst, cr, ci);
    __syncthreads();
    ozero(sr, si);
    t.template mul<OPH, OPN>(s.re[2], s.im[2], s.re[3], s.im[3], LD, sr, si);
    __syncthreads();
    return st;
  };
  ozero(sr, si);
  for (int i = 0; i < nl; ++i) {
    if (kc > 0 && i % kc == 0) put_glb(blk(yw, i / kc), sr, si);
    const int st = left_step(i, kc > 0 ? nullptr : blk(yw, i));
    if (st != ST_OK) {
      status = 10 + st;
      break;
    }
  }
  if (tid == 0) a.status[inst] = status;
  if (status != ST_OK) return;
  __threadfence_block();
  // 6. gradients, i = N-1 .. 0, with u_{i+1} and P_{i+1} kept in LDS
  T* u1r = s.re[0];  // u_{N-1} = G_NN
  T* u1i = s.im[0];
  T* p1r = s.re[1];
  T* p1i = s.im[1];
  T* f0r = s.re[2];
  T* f0i = s.im[2];
  T* f1r = s.re[3];
  T* f1i = s.im[3];
  __syncthreads();
  bload<T, NT>(blk(vw, nl - 1), n, n, p1r, p1i, LD, tid);
  __syncthreads();
  ozero(cr, ci);
  t.template mul<OPH, OPH>(p1r, p1i, u1r, u1i, LD, cr, ci);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    cr[q] *= w2;
    ci[q] *= w2;
  }
  auto gput_onsite = [&](int i, const T (&cr)[R], const T (&ci)[R]) {
    gput(blk(go, i), cr, ci);
    if (a.grad_ends_out) {
      double* e = a.grad_ends_out + (size_t)inst * 2 * nn * 2;
      if (i == 0) put_glb(e, cr, ci);
      if (i == nl - 1) put_glb(e + nn * 2, cr, ci);
    }
  };
  gput_onsite(nl - 1, cr, ci);
  int seg_a = nl;  // first layer of the segment whose Y are in ybuf (checkpoints)
  for (int i = nl - 2; i >= 0; --i) {
    if (kc > 0 && i < seg_a) {
      // recompute Y_a..Y_i of this segment from its checkpoint; u_{i+1} waits in usave
      // and P_{i+1} is reloaded from the row-1 blocks
      seg_a = (i / kc) * kc;
      __syncthreads();
      gstore<T, NT>(usave, n, u1r, u1i, LD, tid);
      __threadfence_block();
      __syncthreads();
      {
        const double2* ck = reinterpret_cast<const double2*>(blk(yw, i / kc));
#pragma unroll
        for (int q = 0; q < R; ++q) {
          int r, c;
          const double2 v = t.at(q, r, c) ? ck[r * n + c] : make_double2(0.0, 0.0);
          sr[q] = v.x;
          si[q] = v.y;
        }
      }
      for (int l = seg_a; l <= i; ++l) left_step(l, blk(ybuf, l - seg_a));
      __threadfence_block();
      __syncthreads();
      u1r = s.re[0];
      u1i = s.im[0];
      p1r = s.re[1];
      p1i = s.im[1];
      f0r = s.re[2];
      f0i = s.im[2];
      f1r = s.re[3];
      f1i = s.im[3];
      bload<T, NT>(usave, n, n, u1r, u1i, LD, tid);
      bload<T, NT>(blk(vw, i + 1), n, n, p1r, p1i, LD, tid);
      __syncthreads();
    }
    bload<T, NT>(kc > 0 ? blk(ybuf, i - seg_a) : blk(yw, i), n, n, f0r, f0i, LD, tid);  // Y_i
    bload<T, NT>(blk(vw, i), n, n, f1r, f1i, LD, tid);  // P_i
    __syncthreads();
    ozero(sr, si);  // u_i = Y_i u_{i+1}
    t.template mul<OPN, OPN>(f0r, f0i, u1r, u1i, LD, sr, si);
    __syncthreads();
    put_lds(f0r, f0i, sr, si);  // f0 = u_i
    __syncthreads();
    // coupling i: (u_{i+1} P_i)^H + u_i P_{i+1} = P_i^H u_{i+1}^H + u_i P_{i+1}
    ozero(cr, ci);
    t.template mul<OPH, OPH>(f1r, f1i, u1r, u1i, LD, cr, ci);
    t.template mul<OPN, OPN>(f0r, f0i, p1r, p1i, LD, cr, ci);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cr[q] *= w2;
      ci[q] *= w2;
    }
    gput(blk(gc, i), cr, ci);
    // onsite i: (u_i P_i)^H = P_i^H u_i^H
    ozero(cr, ci);
    t.template mul<OPH, OPH>(f1r, f1i, f0r, f0i, LD, cr, ci);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cr[q] *= w2;
      ci[q] *= w2;
    }
    gput_onsite(i, cr, ci);
    __syncthreads();
    // u_{i+1} <- u_i (f0), P_{i+1} <- P_i (f1); the old ones become free
    T* xr = u1r;
    T* xi = u1i;
    u1r = f0r;
    u1i = f0i;
    f0r = xr;
    f0i = xi;
    xr = p1r;
    xi = p1i;
    p1r = f1r;
    p1i = f1i;
    f1r = xr;
    f1i = xi;
  }
}

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_adjoint(AdjArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using L = Lds<double, NMAX, Cfg<NMAX>::NT, 4>;
  L s(smem);
  adj_body<NMAX, L, true>(a, blockIdx.x, s, threadIdx.x);
}

// ---------------------------------------------------------------------------
// T = Re Tr[Gamma_L G Gamma_R G^H]  (FP64)
// ---------------------------------------------------------------------------

struct TraceArgs {
  int n, n_inst;
  const double* sigma_l;
  const double* sigma_r;
  const double* g1n;
  const int32_t* status_in;
  double* t_out;
};

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_trace(TraceArgs a) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT, TM = Cfg<NMAX>::TM;
  using L = Lds<T, NMAX, NT, 3>;
  constexpr int LD = L::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  if (a.status_in && a.status_in[inst] != ST_OK) {
    if (tid == 0) a.t_out[inst] = NAN;
    return;
  }
  const int n = a.n;
  const int nn = n * n;
  const double2* sl = reinterpret_cast<const double2*>(a.sigma_l + (size_t)inst * nn * 2);
  const double2* sr = reinterpret_cast<const double2*>(a.sigma_r + (size_t)inst * nn * 2);
  lds_zero<T, NT>(smem, L::mat_bytes(), tid);
  __syncthreads();
  // Gamma = i (S - S^H):  Gamma_ij = i (S_ij - conj(S_ji)) = (-(Im S_ij + Im S_ji), Re S_ij - Re S_ji)
  for (int e = tid; e < nn; e += NT) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 a1 = sl[i * n + j], a2 = sl[j * n + i];
    s.re[0][i * LD + j] = -(a1.y + a2.y);
    s.im[0][i * LD + j] = a1.x - a2.x;
    const double2 b1 = sr[i * n + j], b2 = sr[j * n + i];
    s.re[1][i * LD + j] = -(b1.y + b2.y);
    s.im[1][i * LD + j] = b1.x - b2.x;
  }
  gload<T, NT>(a.g1n + (size_t)inst * nn * 2, n, n, s.re[2], s.im[2], LD, tid);
  __syncthreads();
  const Tile<TM> t(n, n, tid);
  T pr[TM][TM], pi[TM][TM], qr[TM][TM], qi[TM][TM];
  tzero<T, TM>(pr, pi);
  tzero<T, TM>(qr, qi);
  double acc = 0.0;
  if (t.own) {
    // P = Gamma_L G ; Q_ij = (Gamma_R G^H)_ji = sum_k conj(G_ik) Gamma_R_jk
    mac<T, TM, OPN, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, n, t.i0, t.j0, pr, pi);
    mac<T, TM, OPC, OPT>(s.re[2], s.im[2], s.re[1], s.im[1], LD, n, t.i0, t.j0, qr, qi);
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TM; ++y) {
        const int i = t.i0 + x, j = t.j0 + y;
        if (i < n && j < n) acc += pr[x][y] * qr[x][y] - pi[x][y] * qi[x][y];
      }
  }
  const double tot = block_sum<NT>(acc, s.sred, tid);
  if (tid == 0) a.t_out[inst] = tot;
}

// ---------------------------------------------------------------------------
// Adjoint of the lead side: dT/d(lead blocks d00, d01, d10) and dT/dV (FP64)
// ---------------------------------------------------------------------------
//
// For one lead (left: contact layer 1, right: layer N), with the device adjoint's
// weighted gradient E of that layer's onsite block, G = G_1N and w the instance weight:
//   dT/dSigma  = P^T (E - 2 i w X) P,  X = G Gamma_R G^H (left) or G^H Gamma_L G (right),
// P the contact selection (A = z - D - Sigma: Sigma enters G like the onsite block; the
// -2 i w X term is T's direct dependence through Gamma = i (Sigma - Sigma^H), X Hermitian).
// With S = dT/dSigma and Sigma = V g V^H:
//   dT/dg = V^H S V,   dT/dV = S V g^H + S^H V g.
// The surface GF solves g^-1 = z - d00 - d01 g d10, so dg - (g d01) dg (d10 g) =
// g (dd00 + dd01 g d10 + d01 g dd10) g; with Z the solution of the adjoint Stein equation
//   Z - (g d01)^H Z (d10 g)^H = dT/dg
// (by doubling: Z += P^H Z Q^H, P <- P^2, Q <- Q^2, P = g d01, Q = d10 g, until the
// increment is below rounding) and K = g^H Z g^H:
//   dT/dd00 = K,   dT/dd01 = K d10^H g^H,   dT/dd10 = g^H d01^H K.
// Gradients in torch's convention (d/dRe + i d/dIm), already multiplied by w like E.
// The reference has no gradients; this differentiates its Sigma = V g V^H
// (src/negfpy/core/negf.py:72-97) and surface GF (src/negfpy/core/surface_gf.py:18-81).

struct LeadAdjArgs {
  int n, n_inst, side;  // side 0: left lead (device layer 1), 1: right lead (layer N)
  int n_lead, n_c;
  const double* lead;      // [n_sets][3][n_lead][n_lead]  d00, d01, d10
  const double* v;         // [n_sets][n_c][n_lead]
  const int32_t* idx;      // [n_c] contact positions in the end layer
  const int32_t* set_of;
  const double* g;         // [n_inst][n_lead][n_lead] surface GF of this lead
  const double* g1n;       // [n_inst][n][n]
  const double* sigma_l;   // [n_inst][n][n] embedded self-energies
  const double* sigma_r;
  const double* grad_end;  // weighted dT/dD of the contact layer's onsite block (n x n) ..
  size_t grad_stride;      // .. of instance i at grad_end + i * grad_stride (doubles)
  const double* weight;    // [n_inst] or null (1)
  const int32_t* status_in;  // instances with status != 0 get zero gradients
  int maxiter;             // doublings of the Stein solve
  double* grad_lead;       // [n_inst][3][n_lead][n_lead]
  double* grad_v;          // [n_inst][n_c][n_lead]
  int32_t* status;         // 0, status_in, 1 (Stein doubling not converged), 2 (non-finite)
};

template <int NMAX>
__global__ void __launch_bounds__(Cfg<NMAX>::NT) k_lead_adjoint(LeadAdjArgs a) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  using L = Lds<T, NMAX, NT, 4>;
  using M = MOwn<NMAX, NT>;
  constexpr int LD = L::LD, TPW = M::TPW;
  extern __shared__ __align__(16) unsigned char smem[];
  L s(smem);
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  const int n = a.n, nl = a.n_lead, nc = a.n_c;
  const size_t nn = (size_t)n * n, ll = (size_t)nl * nl;
  double* gl_out = a.grad_lead + (size_t)inst * 3 * ll * 2;
  double* gv_out = a.grad_v + (size_t)inst * nc * nl * 2;
  if (a.status_in && a.status_in[inst] != ST_OK) {
    for (size_t e = tid; e < 3 * ll * 2; e += NT) gl_out[e] = 0.0;
    for (size_t e = tid; e < (size_t)nc * nl * 2; e += NT) gv_out[e] = 0.0;
    if (tid == 0) a.status[inst] = a.status_in[inst];
    return;
  }
  const M o(tid);
  const int ns = (n + 3) / 4, nsl = (nl + 3) / 4, nsc = (nc + 3) / 4;
  const size_t set = (size_t)a.set_of[inst];
  const double* d00 = a.lead + set * 3 * ll * 2;
  const double* d01 = d00 + ll * 2;
  const double* d10 = d01 + ll * 2;
  const double* vb = a.v + set * nc * nl * 2;
  const double* gb = a.g + (size_t)inst * ll * 2;
  const double w = a.weight ? a.weight[inst] : 1.0;
  const int ncl = nc > nl ? nc : nl;
  auto clear = [&]() {
    __syncthreads();
    s.zero(tid);
    __syncthreads();
  };
  // owned elements of a matrix-core result, to global memory ([rows][cols], row-major)
  auto put = [&](double* dst, int rows, int cols, const T (&xr)[TPW], const T (&xi)[TPW]) {
    double2* d = reinterpret_cast<double2*>(dst);
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int r = o.row(q), c = o.col(q);
      if (r < rows && c < cols) d[r * cols + c] = make_double2(xr[q], xi[q]);
    }
  };
  T xr[TPW], xi[TPW], yr[TPW], yi[TPW];

  // 1. X = G Gamma_R G^H (left) or G^H Gamma_L G (right): G in [0], Gamma in [1]
  clear();
  gload<T, NT>(a.g1n + (size_t)inst * nn * 2, n, n, s.re[0], s.im[0], LD, tid);
  {
    const double2* so =
        reinterpret_cast<const double2*>((a.side == 0 ? a.sigma_r : a.sigma_l) + (size_t)inst * nn * 2);
    for (int e = tid; e < (int)nn; e += NT) {
      const int i = e / n, j = e - (e / n) * n;
      const double2 s1 = so[e], s2 = so[j * n + i];
      s.re[1][i * LD + j] = -(s1.y + s2.y);
      s.im[1][i * LD + j] = s1.x - s2.x;
    }
  }
  __syncthreads();
  mzero<TPW>(xr, xi);
  if (a.side == 0)
    mm<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, xr, xi);
  else
    mm<NMAX, NT, OPH, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, xr, xi);
  mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, xr, xi);
  __syncthreads();
  mzero<TPW>(xr, xi);
  if (a.side == 0)
    mm<NMAX, NT, OPN, OPH>(s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o, xr, xi);
  else
    mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o, xr, xi);
  // 2. dT/dSigma in the end layer: E - 2 i w X, into [3]; its contact block S into [2]
  {
    const double2* ge = reinterpret_cast<const double2*>(a.grad_end + (size_t)inst * a.grad_stride);
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int r = o.row(q), c = o.col(q);
      if (r < n && c < n) {
        const double2 e = ge[r * n + c];
        s.re[3][r * LD + c] = e.x + 2.0 * w * xi[q];
        s.im[3][r * LD + c] = e.y - 2.0 * w * xr[q];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < L::MSZ; e += NT) {
    s.re[0][e] = s.im[0][e] = 0.0;
    s.re[1][e] = s.im[1][e] = 0.0;
    s.re[2][e] = s.im[2][e] = 0.0;
  }
  __syncthreads();
  for (int e = tid; e < nc * nc; e += NT) {
    const int i = e / nc, j = e - (e / nc) * nc;
    const int pi = a.idx[i], pj = a.idx[j];
    s.re[2][i * LD + j] = s.re[3][pi * LD + pj];
    s.im[2][i * LD + j] = s.im[3][pi * LD + pj];
  }
  __syncthreads();
  for (int e = tid; e < L::MSZ; e += NT) s.re[3][e] = s.im[3][e] = 0.0;
  gload<T, NT>(vb, nc, nl, s.re[0], s.im[0], LD, tid);
  gload<T, NT>(gb, nl, nl, s.re[1], s.im[1], LD, tid);
  __syncthreads();
  // 3. U1 = S V into [3]; U2 = S^H V; dT/dg = V^H U1; dT/dV = U1 g^H + U2 g
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, nsc, o, xr, xi);
  mstore<NMAX, NT>(s.re[3], s.im[3], LD, ncl, o, xr, xi);  // rows >= nc, cols >= nl: zeros
  mzero<TPW>(yr, yi);
  mm<NMAX, NT, OPH, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, nsc, o, yr, yi);
  __syncthreads();
  T gbr[TPW], gbi[TPW];  // dT/dg, the right side of the Stein equation
  mzero<TPW>(gbr, gbi);
  mm<NMAX, NT, OPH, OPN>(s.re[0], s.im[0], s.re[3], s.im[3], LD, nsc, o, gbr, gbi);
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPH>(s.re[3], s.im[3], s.re[1], s.im[1], LD, nsl, o, xr, xi);
  __syncthreads();
  mstore<NMAX, NT>(s.re[2], s.im[2], LD, ncl, o, yr, yi);  // U2 over S
  __syncthreads();
  mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[1], s.im[1], LD, nsl, o, xr, xi);
  put(gv_out, nc, nl, xr, xi);
  // 4. Stein equation by doubling: [0] P, [1] Q, [2] Z, [3] P^H Z
  clear();
  gload<T, NT>(gb, nl, nl, s.re[3], s.im[3], LD, tid);
  gload<T, NT>(d01, nl, nl, s.re[0], s.im[0], LD, tid);
  gload<T, NT>(d10, nl, nl, s.re[1], s.im[1], LD, tid);
  __syncthreads();
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPN>(s.re[3], s.im[3], s.re[0], s.im[0], LD, nsl, o, xr, xi);  // g d01
  mzero<TPW>(yr, yi);
  mm<NMAX, NT, OPN, OPN>(s.re[1], s.im[1], s.re[3], s.im[3], LD, nsl, o, yr, yi);  // d10 g
  __syncthreads();
  mstore<NMAX, NT>(s.re[0], s.im[0], LD, nl, o, xr, xi);
  mstore<NMAX, NT>(s.re[1], s.im[1], LD, nl, o, yr, yi);
  mstore<NMAX, NT>(s.re[2], s.im[2], LD, nl, o, gbr, gbi);
  __syncthreads();
  int status = ST_OK;
  bool conv = false;
  for (int it = 0; it < a.maxiter && !conv; ++it) {
    mzero<TPW>(xr, xi);
    mm<NMAX, NT, OPH, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, nsl, o, xr, xi);  // P^H Z
    mstore<NMAX, NT>(s.re[3], s.im[3], LD, nl, o, xr, xi);
    __syncthreads();
    mzero<TPW>(xr, xi);
    mm<NMAX, NT, OPN, OPH>(s.re[3], s.im[3], s.re[1], s.im[1], LD, nsl, o, xr, xi);  // .. Q^H
    T pr[TPW], pi[TPW], qr[TPW], qi[TPW];
    mzero<TPW>(pr, pi);
    mm<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[0], s.im[0], LD, nsl, o, pr, pi);  // P^2
    mzero<TPW>(qr, qi);
    mm<NMAX, NT, OPN, OPN>(s.re[1], s.im[1], s.re[1], s.im[1], LD, nsl, o, qr, qi);  // Q^2
    double dmax = 0.0, zmax = 0.0;
    bool bad = false;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      gbr[q] += xr[q];
      gbi[q] += xi[q];
      dmax = fmax(dmax, fabs(xr[q]) + fabs(xi[q]));
      zmax = fmax(zmax, fabs(gbr[q]) + fabs(gbi[q]));
      bad = bad || !isfinite(gbr[q]) || !isfinite(gbi[q]);
    }
    double dm, zm;
    bool any_bad, unused;
    block_max_any<NT>(dmax, bad, s.sred, s.sflag, tid, dm, any_bad);
    block_max_any<NT>(zmax, false, s.sred, s.sflag, tid, zm, unused);
    if (any_bad) {
      status = ST_NONFINITE;
      break;
    }
    conv = dm <= 1e-17 * zm;
    mstore<NMAX, NT>(s.re[0], s.im[0], LD, nl, o, pr, pi);
    mstore<NMAX, NT>(s.re[1], s.im[1], LD, nl, o, qr, qi);
    mstore<NMAX, NT>(s.re[2], s.im[2], LD, nl, o, gbr, gbi);
    __syncthreads();
  }
  if (status == ST_OK && !conv) status = ST_NOCONV;
  // 5. K = g^H Z g^H; dT/dd00 = K, dT/dd01 = K d10^H g^H, dT/dd10 = g^H d01^H K
  __syncthreads();
  gload<T, NT>(gb, nl, nl, s.re[0], s.im[0], LD, tid);
  __syncthreads();
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPH, OPN>(s.re[0], s.im[0], s.re[2], s.im[2], LD, nsl, o, xr, xi);  // g^H Z
  mstore<NMAX, NT>(s.re[3], s.im[3], LD, nl, o, xr, xi);
  __syncthreads();
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPH>(s.re[3], s.im[3], s.re[0], s.im[0], LD, nsl, o, xr, xi);  // K
  put(gl_out, nl, nl, xr, xi);
  mstore<NMAX, NT>(s.re[1], s.im[1], LD, nl, o, xr, xi);
  __syncthreads();
  gload<T, NT>(d10, nl, nl, s.re[3], s.im[3], LD, tid);
  __syncthreads();
  mzero<TPW>(yr, yi);
  mm<NMAX, NT, OPN, OPH>(s.re[1], s.im[1], s.re[3], s.im[3], LD, nsl, o, yr, yi);  // K d10^H
  mstore<NMAX, NT>(s.re[2], s.im[2], LD, nl, o, yr, yi);
  __syncthreads();
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPH>(s.re[2], s.im[2], s.re[0], s.im[0], LD, nsl, o, xr, xi);  // .. g^H
  put(gl_out + ll * 2, nl, nl, xr, xi);
  __syncthreads();
  gload<T, NT>(d01, nl, nl, s.re[3], s.im[3], LD, tid);
  __syncthreads();
  mzero<TPW>(yr, yi);
  mm<NMAX, NT, OPH, OPH>(s.re[0], s.im[0], s.re[3], s.im[3], LD, nsl, o, yr, yi);  // g^H d01^H
  mstore<NMAX, NT>(s.re[2], s.im[2], LD, nl, o, yr, yi);
  __syncthreads();
  mzero<TPW>(xr, xi);
  mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[1], s.im[1], LD, nsl, o, xr, xi);  // .. K
  put(gl_out + 2 * ll * 2, nl, nl, xr, xi);
  if (tid == 0) a.status[inst] = status;
}

// ---------------------------------------------------------------------------
// k-space block builder (FP64)
// ---------------------------------------------------------------------------
//
// Phi_ab(k) = sum_t Phi_ab,t e^{i theta_t(k)} (plus the onsite pinning on the
// diagonal of Phi_00), scaled as D = M_l^-1/2 Phi M_r^-1/2, for every k point and
// every block of the lead (d00, d01, d10) and of the device (L onsite, L-1
// coupling blocks).  Reference: src/negfpy/models/material_kspace.py:99-128
// (_sum_k_terms, _build_dynamical_blocks) and :166-193 (per-layer masses).  The
// phases and 1/sqrt(m) come from the host; each output element is one chain of
// separately rounded products and sums in the reference's order (no fused
// multiply-add), so the blocks equal the reference's to the last bit.

struct KspaceArgs {
  int n, n_layers, n_k;
  int t00, t01, t10;        // term counts; t10 = 0: Phi_10 = Phi_01^H
  const double* phi00;      // [t00][n][n] complex
  const double* phi01;      // [t01][n][n]
  const double* phi10;      // [t10][n][n]
  const double* ph00;       // [n_k][t00] complex phases
  const double* ph01;       // [n_k][t01]
  const double* ph10;       // [n_k][t10]
  double pinning;
  const double* s_lead;     // [n] 1/sqrt(m) of the lead cell (null: no lead output)
  const double* s_dev;      // [L][n] per device layer (null: no device output)
  double* lead_out;         // [n_k][3][n][n]
  double* onsite_out;       // [n_k][L][n][n]
  double* coupling_out;     // [n_k][L-1][n][n]
};

__device__ __forceinline__ double2 kspace_sum(const double* __restrict__ phi,
                                              const double* __restrict__ ph, int nt, size_t nn,
                                              size_t e) {
#pragma clang fp contract(off)
  double ar = 0.0, ai = 0.0;
  for (int q = 0; q < nt; ++q) {
    const double br = phi[(q * nn + e) * 2], bi = phi[(q * nn + e) * 2 + 1];
    const double c = ph[2 * q], sn = ph[2 * q + 1];
    const double pr = br * c - bi * sn;
    const double pi = br * sn + bi * c;
    ar = ar + pr;
    ai = ai + pi;
  }
  return make_double2(ar, ai);
}

__global__ void __launch_bounds__(256) k_kspace_blocks(KspaceArgs a) {
#pragma clang fp contract(off)
  const int k = blockIdx.x;
  const int b = blockIdx.y;  // 0..2 lead d00, d01, d10; then L onsite; then L-1 couplings
  const int n = a.n, nl = a.n_layers;
  const size_t nn = (size_t)n * n;
  int kind;      // 0: Phi00, 1: Phi01, 2: Phi10
  const double* sl;
  const double* sr;
  double* out;
  if (b < 3) {
    if (!a.s_lead) return;
    kind = b;
    sl = a.s_lead;
    sr = a.s_lead;
    out = a.lead_out + ((size_t)k * 3 + b) * nn * 2;
  } else if (b < 3 + nl) {
    if (!a.s_dev) return;
    const int i = b - 3;
    kind = 0;
    sl = a.s_dev + (size_t)i * n;
    sr = sl;
    out = a.onsite_out + ((size_t)k * nl + i) * nn * 2;
  } else {
    if (!a.s_dev) return;
    const int i = b - 3 - nl;
    kind = 1;
    sl = a.s_dev + (size_t)i * n;
    sr = a.s_dev + (size_t)(i + 1) * n;
    out = a.coupling_out + ((size_t)k * (nl - 1) + i) * nn * 2;
  }
  for (int e = threadIdx.x; e < (int)nn; e += 256) {
    const int i = e / n, j = e - (e / n) * n;
    double2 v;
    if (kind == 0) {
      v = kspace_sum(a.phi00, a.ph00 + (size_t)k * a.t00 * 2, a.t00, nn, e);
      if (a.pinning != 0.0 && i == j) v.x = v.x + a.pinning;
    } else if (kind == 1 || a.t10 > 0) {
      const bool is10 = (kind == 2);
      v = is10 ? kspace_sum(a.phi10, a.ph10 + (size_t)k * a.t10 * 2, a.t10, nn, e)
               : kspace_sum(a.phi01, a.ph01 + (size_t)k * a.t01 * 2, a.t01, nn, e);
    } else {
      // Phi_10(k) = Phi_01(k)^H
      const double2 w = kspace_sum(a.phi01, a.ph01 + (size_t)k * a.t01 * 2, a.t01, nn,
                                   (size_t)j * n + i);
      v = make_double2(w.x, -w.y);
    }
    // (M_l^-1/2 Phi) M_r^-1/2 for d00, d01; d10 = M_r^-1/2 Phi_10 M_l^-1/2 of the lead
    const double li = sl[i], rj = sr[j];
    double2 o;
    o.x = (li * v.x) * rj;
    o.y = (li * v.y) * rj;
    reinterpret_cast<double2*>(out)[e] = o;
  }
}

// ---------------------------------------------------------------------------
// omega integrals (FP64)
// ---------------------------------------------------------------------------

constexpr double HBAR_J_S = 1.054571817e-34;
constexpr double KB_J_K = 1.380649e-23;

// k_B T(w) x^2 e^-x / (1 - e^-x)^2 with x = hbar w / (k_B T); 1 - x^2/12 below 1e-6, 0 from 700.
__device__ __forceinline__ double conductance_integrand(double w, double tw, double temp) {
  const double x = (HBAR_J_S * w) / (KB_J_K * temp);
  double ker = 0.0;
  if (x < 1e-6) {
    ker = 1.0 - (x * x) / 12.0;
  } else if (x < 700.0) {
    const double y = exp(-x);
    const double d = (1.0 - y) * (1.0 - y);
    ker = (x * x) * y / d;
  }
  return KB_J_K * tw * ker;
}

__global__ void __launch_bounds__(256) k_conductance(int n_w, const double* __restrict__ w,
                                                      const double* __restrict__ tw, int n_t,
                                                      const double* __restrict__ temps,
                                                      double* __restrict__ out) {
  __shared__ double sred[4];
  const int tid = threadIdx.x;
  const int it = blockIdx.x;
  const double temp = temps[it];
  double acc = 0.0;
  if (temp > 0.0) {
    for (int i = tid; i + 1 < n_w; i += 256) {
      const double f0 = conductance_integrand(w[i], tw[i], temp);
      const double f1 = conductance_integrand(w[i + 1], tw[i + 1], temp);
      acc += (w[i + 1] - w[i]) * (f1 + f0) / 2.0;
    }
  }
  const double s = block_sum<256>(acc, sred, tid);
  if (tid == 0) out[it] = temp > 0.0 ? (1.0 / (2.0 * M_PI)) * s : 0.0;
}

// Full counting statistics (core/fcs.py:47-156): for each temperature pair,
// C1 = (1/2pi) int <T> (A_L - A_R) dw and
// C2 = (1/2pi) int <T> (B_L + B_R) + (<T> + <T^2>) (A_L - A_R)^2 dw
// (single-spectrum form T (B_L + B_R) + T (1 + T)(A_L - A_R)^2 when t2 is null),
// A = k_B T phi1(x), B = (k_B T)^2 phi2(x), x = hbar w / (k_B T); phi1, phi2 with
// the reference's series below |x| = 1e-6, its expm1 forms up to 700 (including
// their overflow to NaN) and 0 beyond; each term rounded separately, as numpy
// does, so equal temperatures give exactly C1 = 0.
__device__ __forceinline__ double fcs_phi1(double x) {
#pragma clang fp contract(off)
  const double ax = fabs(x);
  if (ax < 1e-6) return 1.0 - 0.5 * x + (x * x) / 12.0 - pow(x, 4.0) / 720.0;
  if (x < 700.0) return x / expm1(x);
  return 0.0;
}
__device__ __forceinline__ double fcs_phi2(double x) {
#pragma clang fp contract(off)
  const double ax = fabs(x);
  if (ax < 1e-6) return 1.0 - (x * x) / 12.0 + pow(x, 4.0) / 240.0;
  if (x < 700.0) {
    const double em1 = expm1(x);
    return (x * x) * (em1 + 1.0) / (em1 * em1);
  }
  return 0.0;
}
__device__ __forceinline__ void fcs_integrand(double w, double t1, const double* t2, int i,
                                              double tl, double tr, double& f1, double& f2) {
#pragma clang fp contract(off)
  const double sl = KB_J_K * tl, sr = KB_J_K * tr;
  const double xl = (HBAR_J_S * w) / sl, xr = (HBAR_J_S * w) / sr;
  const double ad = sl * fcs_phi1(xl) - sr * fcs_phi1(xr);
  const double bsum = (sl * sl) * fcs_phi2(xl) + (sr * sr) * fcs_phi2(xr);
  f1 = t1 * ad;
  f2 = t2 ? t1 * bsum + (t1 + t2[i]) * (ad * ad) : t1 * bsum + t1 * (1.0 + t1) * (ad * ad);
}

__global__ void __launch_bounds__(256) k_fcs_cumulants(int n_w, const double* __restrict__ w,
                                                        const double* __restrict__ t1,
                                                        const double* __restrict__ t2,
                                                        const double* __restrict__ temps_l,
                                                        const double* __restrict__ temps_r,
                                                        double* __restrict__ out) {
  __shared__ double sred[4];
  const int tid = threadIdx.x;
  const int p = blockIdx.x;
  const double tl = temps_l[p], tr = temps_r[p];
  double a1 = 0.0, a2 = 0.0;
  if (tl > 0.0 && tr > 0.0) {
    for (int i = tid; i + 1 < n_w; i += 256) {
      double f1a, f2a, f1b, f2b;
      fcs_integrand(w[i], t1[i], t2, i, tl, tr, f1a, f2a);
      fcs_integrand(w[i + 1], t1[i + 1], t2, i + 1, tl, tr, f1b, f2b);
      const double h = w[i + 1] - w[i];
      a1 += h * (f1b + f1a) / 2.0;
      a2 += h * (f2b + f2a) / 2.0;
    }
  }
  const double s1 = block_sum<256>(a1, sred, tid);
  const double s2 = block_sum<256>(a2, sred, tid);
  if (tid == 0) {
    const bool ok = tl > 0.0 && tr > 0.0;
    out[2 * p] = ok ? (1.0 / (2.0 * M_PI)) * s1 : NAN;
    out[2 * p + 1] = ok ? (1.0 / (2.0 * M_PI)) * s2 : NAN;
  }
}

// d kappa(T_t) / d T(w_i): the trapezoid weights of k_conductance times its integrand
// per unit transmission, so that kappa = weights @ T.
__global__ void __launch_bounds__(256) k_conductance_weights(int n_w, const double* __restrict__ w,
                                                              int n_t,
                                                              const double* __restrict__ temps,
                                                              double* __restrict__ out) {
  const int it = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_w) return;
  const double temp = temps[it];
  double v = 0.0;
  if (temp > 0.0 && n_w > 1) {
    double h = 0.0;
    if (i + 1 < n_w) h += w[i + 1] - w[i];
    if (i > 0) h += w[i] - w[i - 1];
    v = (1.0 / (2.0 * M_PI)) * conductance_integrand(w[i], 1.0, temp) * h / 2.0;
  }
  out[(size_t)it * n_w + i] = v;
}

// x / (e^x - 1) with the small-x series below 1e-6 and 0 from 700.
__device__ __forceinline__ double phi_bose(double x) {
  if (x < 1e-6) return 1.0 - 0.5 * x + (x * x) / 12.0;
  if (x < 700.0) return x / expm1(x);
  return 0.0;
}

__global__ void __launch_bounds__(256) k_heat_current(int n_w, const double* __restrict__ w,
                                                       const double* __restrict__ tw, double t_l,
                                                       double t_r, double* __restrict__ out) {
  __shared__ double sred[4];
  const int tid = threadIdx.x;
  double acc = 0.0;
  for (int i = tid; i + 1 < n_w; i += 256) {
    double f[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const double wq = w[i + q];
      const double xl = HBAR_J_S * wq / (KB_J_K * t_l);
      const double xr = HBAR_J_S * wq / (KB_J_K * t_r);
      f[q] = (KB_J_K * t_l * phi_bose(xl) - KB_J_K * t_r * phi_bose(xr)) * tw[i + q];
    }
    acc += (w[i + 1] - w[i]) * (f[1] + f[0]) / 2.0;
  }
  const double s = block_sum<256>(acc, sred, tid);
  if (tid == 0) out[0] = (1.0 / (2.0 * M_PI)) * s;
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------

}  // namespace negf

#include "big.hpp"
#include "persist.hpp"
#include "wpi.hpp"

namespace negf {

// full Green's function and adjoint for 49..96-DOF blocks: the same bodies with
// their four matrices in a pool slot, persistent over the instances
using GreenBig = GmL<double, BIG, BIG_NT, 4>;
static_assert(GreenBig::slot_elems() <= BIG_SLOT, "pool slot too small");

__global__ void __launch_bounds__(BIG_NT) k_device_green_big(DgfArgs a, double* pool) {
  __shared__ GmShared<double, BIG, BIG_NT, 4> sh;
  GreenBig s(pool + (size_t)blockIdx.x * BIG_SLOT, &sh);
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    dgf_body<BIG>(a, inst, s, threadIdx.x);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BIG_NT) k_adjoint_big(AdjArgs a, double* pool) {
  __shared__ GmShared<double, BIG, BIG_NT, 4> sh;
  GreenBig s(pool + (size_t)blockIdx.x * BIG_SLOT, &sh);
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) {
    adj_body<BIG, GreenBig, false>(a, inst, s, threadIdx.x);
    __syncthreads();
  }
}

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

static int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(NEGF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return NEGF_OK;
}

// Block-size class chosen for a block of n DOF: the smallest class >= n unless the
// autotuner recorded another one for (precision, n).  Entries are read on every
// launch and written rarely (negf_set_block_class), hence atomics.  FP64 classes:
// 8, 16, 24, 32, 48 (LDS-resident) and 96 (matrices in a device-memory pool);
// mixed precision: 8 .. 64 (its FP64 refinement bounds it at 48 DOF).
constexpr int FP64_MAX = 96;
constexpr int MIXED_MAX = 48;
constexpr int GREEN_MAX = 96;  // full Green's function and adjoint kernels
static std::atomic<int> g_class_override[2][FP64_MAX + 1];

static int default_nmax(int n, bool fp64) {
  if (n < 1) return 0;
  const int opts[] = {8, 16, 24, 32, 48};
  for (int o : opts)
    if (n <= o) return o;
  if (fp64) return n <= BIG ? BIG : 0;
  return n <= 64 ? 64 : 0;
}

// One-instance-per-wavefront MFMA kernels (wpi.hpp) for blocks of up to 32 DOF.  With
// the matrix-core products and the lean pivot step of the workgroup classes, the
// classes are faster at every size on MI355X (8192 instances, 16 layers, eta = 1e-6,
// ms per solve, class / wave): FP64 n = 6: 0.50 / 1.40, 12: 2.26 / 2.48, 16: 3.09 / 4.36,
// 24: 8.33 / 15.5; the FP32 decimation of the mixed path at n = 24: 1.46 / 1.61.  So mode 1
// (default) uses the wave family only where a per-size choice (negf_set_kernel_family,
// the autotuner's) asks for it; mode 2 uses it for every block up to 32 DOF, mode 0 never
// (negf_set_wave_kernels).  [0] of the per-size table: FP64 decimations and sweeps,
// [1]: FP32 decimations and sweeps.
constexpr int WAVE_MAX = 32;
static std::atomic<int> g_wave_kernels{1};
static std::atomic<int> g_family_override[2][WAVE_MAX + 1];
static bool use_wave(int n, bool fp64, bool sweep) {
  (void)sweep;
  if (n < 1 || n > WAVE_MAX) return false;
  const int mode = g_wave_kernels.load();
  if (mode == 2) return true;
  if (mode != 1) return false;
  const int f = g_family_override[fp64 ? 0 : 1][n].load(std::memory_order_relaxed);
  return f == NEGF_FAMILY_WAVE;
}
// Matrix-core products (mm.hpp) in the FP64 workgroup decimation and sweep of
// classes 8..48: mode 1 (default) on unless a per-size choice (negf_set_product_form,
// the autotuner's) asks for the vector-ALU tiles, mode 0 off
// (negf_set_matrix_core_products).
constexpr int FORM_MAX = 48;
static std::atomic<int> g_mm_products{1};
static std::atomic<int> g_product_form[FORM_MAX + 1];
static bool use_mm(int nmax, int n) {
  if (g_mm_products.load() != 1 || nmax < 8 || nmax > 48) return false;
  const int f = (n >= 1 && n <= FORM_MAX) ? g_product_form[n].load(std::memory_order_relaxed) : 0;
  return f != NEGF_PRODUCTS_VECTOR;
}
// rows per lane of the wave inversion: n <= 8 R
static int wave_rows(int n) { return n <= 16 ? 2 : (n <= 24 ? 3 : 4); }

static int pick_nmax(int n, bool fp64) {
  if (n >= 1 && n <= FP64_MAX) {
    const int o = g_class_override[fp64 ? 0 : 1][n].load(std::memory_order_relaxed);
    if (o) return o;
  }
  return default_nmax(n, fp64);
}

// Device-memory pool of the 96 class: one slot per resident workgroup, allocated on
// first use on a device and kept for the life of the process.  Eager calls on
// different streams are ordered through an event, so two calls never use the slots at
// once.  A call on a capturing stream neither waits on nor records that event (an
// event of eager work cannot order a graph node): it only needs the pool to exist,
// and the graph's replays are ordered by the stream they run on.
struct BigPool {
  double* ptr = nullptr;
  int slots = 0;
  hipEvent_t done = nullptr;
};
static std::mutex g_big_mu;
static BigPool g_big[64];    // slots of BIG_SLOT doubles
static BigPool g_huge[64];   // slots of HUGE_SLOT doubles (eigenmode route, 49..96 DOF)

class BigUse {
 public:
  explicit BigUse(hipStream_t st, BigPool* table = g_big, size_t slot = BIG_SLOT)
      : st_(st), lk_(g_big_mu) {
    int dev = 0;
    rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return;
    if (dev < 0 || dev >= 64) {
      rc = fail(NEGF_ERR_ARG, "device index %d out of range", dev);
      return;
    }
    p_ = &table[dev];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    rc = hip_check(hipStreamIsCapturing(st_, &cs), "stream capture status");
    if (rc) return;
    capturing_ = cs != hipStreamCaptureStatusNone;
    if (!p_->ptr && capturing_) {
      rc = fail(NEGF_ERR_ARG,
                "the 49..96 DOF class allocates its device pool on its first call on a "
                "device, which cannot be captured: make one call outside capture first");
      return;
    }
    if (!p_->ptr) {
      int cus = 0, per = 0;
      rc = hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                     "CU count");
      if (rc) return;
      rc = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per, reinterpret_cast<const void*>(&k_sr_big), BIG_NT, 0),
                     "occupancy");
      if (rc) return;
      const int slots = cus * (per > 0 ? per : 1);
      const size_t bytes = (size_t)slots * slot * sizeof(double);
      double* ptr = nullptr;
      rc = hip_check(hipMalloc(reinterpret_cast<void**>(&ptr), bytes), "96-class pool");
      if (rc) return;
      rc = hip_check(hipMemsetAsync(ptr, 0, bytes, st_), "96-class pool clear");
      if (rc) return;
      rc = hip_check(hipEventCreateWithFlags(&p_->done, hipEventDisableTiming), "pool event");
      if (rc) return;
      p_->ptr = ptr;
      p_->slots = slots;
    } else if (!capturing_) {
      rc = hip_check(hipStreamWaitEvent(st_, p_->done, 0), "pool order");
    }
  }
  // marks the end of this call's use of the pool on its stream
  int finish() {
    return capturing_ ? NEGF_OK : hip_check(hipEventRecord(p_->done, st_), "pool event record");
  }
  double* pool() const { return p_->ptr; }
  int grid(int n_inst) const { return n_inst < p_->slots ? n_inst : p_->slots; }
  int rc = NEGF_OK;

 private:
  hipStream_t st_;
  std::lock_guard<std::mutex> lk_;
  BigPool* p_ = nullptr;
  bool capturing_ = false;
};

#define NEGF_POOL_LAUNCH(KERNEL, STREAM, ARGS, TABLE, SLOT)                                 \
  do {                                                                                      \
    BigUse u_(STREAM, TABLE, SLOT);                                                         \
    if (u_.rc) return u_.rc;                                                                \
    hipLaunchKernelGGL(KERNEL, dim3(u_.grid((ARGS).n_inst)), dim3(BIG_NT), 0, (STREAM),     \
                       (ARGS), u_.pool());                                                  \
    const hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess) return hip_check(e_, #KERNEL " launch");                          \
    return u_.finish();                                                                     \
  } while (0)
#define NEGF_BIG_LAUNCH(KERNEL, STREAM, ARGS) NEGF_POOL_LAUNCH(KERNEL, STREAM, ARGS, g_big, BIG_SLOT)

#define NEGF_LAUNCH(KERNEL, GRID, NT, LDS, STREAM, ...)                                          \
  do {                                                                                          \
    hipError_t e_ = hipFuncSetAttribute(reinterpret_cast<const void*>(&KERNEL),                 \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)(LDS)); \
    if (e_ != hipSuccess) return hip_check(e_, #KERNEL " attribute");                           \
    hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(NT), (LDS), (STREAM), __VA_ARGS__);            \
    e_ = hipGetLastError();                                                                     \
    if (e_ != hipSuccess) return hip_check(e_, #KERNEL " launch");                              \
  } while (0)

// block classes of the mirrored-lead decimation
static bool sr_mirror_class(int nmax) { return nmax >= 8 && nmax <= 48; }

template <typename T, bool MIRROR = false>
static int run_sr(hipStream_t st, int nmax, const SrArgs& a) {
#define SR_CASE(NM)                                                                           \
  case NM: {                                                                                  \
    constexpr int NT = Cfg<NM>::NT;                                                           \
    NEGF_LAUNCH((k_sancho_rubio<T, NM, MIRROR>), a.n_inst, NT, (Lds<T, NM, NT, 3>::bytes()), st, \
                a);                                                                           \
    return NEGF_OK;                                                                           \
  }
  if constexpr (!MIRROR) {
    if (use_wave(a.n, sizeof(T) == 8, false)) {
      const size_t lds = wpi::lds_bytes<T>(a.n, 4);
      switch (wave_rows(a.n)) {
        case 2: NEGF_LAUNCH((wpi::k_sr_wave<T, 2>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
        case 3: NEGF_LAUNCH((wpi::k_sr_wave<T, 3>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
        default: NEGF_LAUNCH((wpi::k_sr_wave<T, 4>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
      }
    }
  }
  if (MIRROR && !sr_mirror_class(nmax))
    return fail(NEGF_ERR_ARG, "no mirrored decimation for block class %d", nmax);
  if constexpr (sizeof(T) == 8) {
    if (use_mm(nmax, a.n)) {
#define SRM_CASE(NM)                                                                        \
  case NM: {                                                                                \
    constexpr int NT = Cfg<NM>::NT;                                                         \
    NEGF_LAUNCH((k_sr_mm<NM, MIRROR>), a.n_inst, NT, (Lds<double, NM, NT, 3>::bytes()), st, a); \
    return NEGF_OK;                                                                         \
  }
      switch (nmax) {
        SRM_CASE(8)
        SRM_CASE(16)
        SRM_CASE(24)
        SRM_CASE(32)
        SRM_CASE(48)
      }
#undef SRM_CASE
    }
  }
  switch (nmax) {
    SR_CASE(8)
    SR_CASE(16)
    SR_CASE(24)
    SR_CASE(32)
    SR_CASE(48)
    case 64:
      if constexpr (sizeof(T) == 4 && !MIRROR) {
        constexpr int NT = Cfg<64>::NT;
        NEGF_LAUNCH((k_sancho_rubio<T, 64, false>), a.n_inst, NT, (Lds<T, 64, NT, 3>::bytes()), st,
                    a);
        return NEGF_OK;
      }
      break;
    case BIG:
      if constexpr (sizeof(T) == 8 && !MIRROR) NEGF_BIG_LAUNCH(k_sr_big, st, a);
      break;
  }
#undef SR_CASE
  return fail(NEGF_ERR_ARG, "unsupported block size class %d", nmax);
}

__global__ void k_mark_redo(int n_inst, int32_t* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_inst && status[i] == ST_OK) status[i] = ST_REDO;
}

// LDS of k_refine: four FP64 and four FP32 planar matrices plus the inversion buffers
template <int NM>
static constexpr size_t refine_lds() {
  using L = Lds<double, NM, Cfg<NM>::NT, 4>;
  return ((L::bytes() + 15) & ~size_t(15)) + sizeof(float) * 8 * L::MSZ;
}

static int run_refine(hipStream_t st, int nmax, const RefineArgs& a) {
#define RF_CASE(NM)                                                                 \
  case NM: {                                                                        \
    constexpr int NT = Cfg<NM>::NT;                                                 \
    NEGF_LAUNCH((k_refine<NM>), a.n_inst, NT, refine_lds<NM>(), st, a);             \
    return NEGF_OK;                                                                 \
  }
  switch (nmax) {
    RF_CASE(8)
    RF_CASE(16)
    RF_CASE(24)
    RF_CASE(32)
    case 48:
      // the Newton step's eight matrices of 48 DOF exceed a CU's LDS: those instances
      // are decimated in FP64 by the redo pass
      hipLaunchKernelGGL(k_mark_redo, dim3((a.n_inst + 255) / 256), dim3(256), 0, st, a.n_inst,
                         a.status);
      return hip_check(hipGetLastError(), "k_mark_redo launch");
  }
#undef RF_CASE
  return fail(NEGF_ERR_ARG, "block size %d exceeds the FP64 limit of 48", a.n);
}

static int run_sigma(hipStream_t st, int nmax, const SigmaArgs& a) {
#define SG_CASE(NM)                                                                          \
  case NM: {                                                                                 \
    constexpr int NT = Cfg<NM>::NT;                                                          \
    NEGF_LAUNCH((k_self_energy<NM>), a.n_inst, NT, (Lds<double, NM, NT, 3>::bytes()), st, a); \
    return NEGF_OK;                                                                          \
  }
  switch (nmax) {
    SG_CASE(8)
    SG_CASE(16)
    SG_CASE(24)
    SG_CASE(32)
    SG_CASE(48)
    case BIG:
      NEGF_BIG_LAUNCH(k_sigma_big, st, a);
  }
#undef SG_CASE
  return fail(NEGF_ERR_ARG, "self-energy block size class %d unsupported", nmax);
}

template <typename T>
static int run_rgf(hipStream_t st, int nmax, const RgfArgs& a) {
#define RG_CASE(NM)                                                                  \
  case NM: {                                                                         \
    constexpr int NT = Cfg<NM>::NT;                                                  \
    NEGF_LAUNCH((k_rgf<T, NM>), a.n_inst, NT, (Lds<T, NM, NT, 3>::bytes()), st, a);  \
    return NEGF_OK;                                                                  \
  }
  if (use_wave(a.n, sizeof(T) == 8, true)) {
    const size_t lds = wpi::lds_bytes<T>(a.n, 3);
    switch (wave_rows(a.n)) {
      case 2: NEGF_LAUNCH((wpi::k_rgf_wave<T, 2>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
      case 3: NEGF_LAUNCH((wpi::k_rgf_wave<T, 3>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
      default: NEGF_LAUNCH((wpi::k_rgf_wave<T, 4>), a.n_inst, WAVE, lds, st, a); return NEGF_OK;
    }
  }
  if constexpr (sizeof(T) == 8) {
    if (use_mm(nmax, a.n)) {
#define RGM_CASE(NM)                                                                       \
  case NM: {                                                                               \
    constexpr int NT = Cfg<NM>::NT;                                                        \
    NEGF_LAUNCH((k_rgf_mm<NM, false>), a.n_inst, NT, (Lds<double, NM, NT, 3>::bytes()), st, a, \
                RgfFuse{});                                                                \
    return NEGF_OK;                                                                        \
  }
      switch (nmax) {
        RGM_CASE(8)
        RGM_CASE(16)
        RGM_CASE(24)
        RGM_CASE(32)
        RGM_CASE(48)
      }
#undef RGM_CASE
    }
  }
  switch (nmax) {
    RG_CASE(8)
    RG_CASE(16)
    RG_CASE(24)
    RG_CASE(32)
    RG_CASE(48)
    case 64:
      if constexpr (sizeof(T) == 4) {
        constexpr int NT = Cfg<64>::NT;
        NEGF_LAUNCH((k_rgf<T, 64>), a.n_inst, NT, (Lds<T, 64, NT, 3>::bytes()), st, a);
        return NEGF_OK;
      }
      break;
    case BIG:
      if constexpr (sizeof(T) == 8) NEGF_BIG_LAUNCH(k_rgf_big, st, a);
      break;
  }
#undef RG_CASE
  return fail(NEGF_ERR_ARG, "unsupported block size class %d", nmax);
}

static int run_trace(hipStream_t st, int nmax, const TraceArgs& a) {
#define TR_CASE(NM)                                                                     \
  case NM: {                                                                            \
    constexpr int NT = Cfg<NM>::NT;                                                     \
    NEGF_LAUNCH((k_trace<NM>), a.n_inst, NT, (Lds<double, NM, NT, 3>::bytes()), st, a); \
    return NEGF_OK;                                                                     \
  }
  switch (nmax) {
    TR_CASE(8)
    TR_CASE(16)
    TR_CASE(24)
    TR_CASE(32)
    TR_CASE(48)
    case BIG:
      NEGF_BIG_LAUNCH(k_trace_big, st, a);
  }
#undef TR_CASE
  return fail(NEGF_ERR_ARG, "trace block size class %d unsupported", nmax);
}

static int run_dgf(hipStream_t st, int nmax, const DgfArgs& a) {
#define DG_CASE(NM)                                                                           \
  case NM: {                                                                                  \
    constexpr int NT = Cfg<NM>::NT;                                                           \
    NEGF_LAUNCH((k_device_green<NM>), a.n_inst, NT, (Lds<double, NM, NT, 4>::bytes()), st, a); \
    return NEGF_OK;                                                                           \
  }
  switch (nmax) {
    DG_CASE(8)
    DG_CASE(16)
    DG_CASE(24)
    DG_CASE(32)
    DG_CASE(48)
    case BIG:
      NEGF_BIG_LAUNCH(k_device_green_big, st, a);
  }
#undef DG_CASE
  return fail(NEGF_ERR_ARG, "block size class %d unsupported", nmax);
}

static int run_adj(hipStream_t st, int nmax, const AdjArgs& a) {
#define ADJ_CASE(NM)                                                                      \
  case NM: {                                                                              \
    constexpr int NT = Cfg<NM>::NT;                                                       \
    NEGF_LAUNCH((k_adjoint<NM>), a.n_inst, NT, (Lds<double, NM, NT, 4>::bytes()), st, a); \
    return NEGF_OK;                                                                       \
  }
  switch (nmax) {
    ADJ_CASE(8)
    ADJ_CASE(16)
    ADJ_CASE(24)
    ADJ_CASE(32)
    ADJ_CASE(48)
    case BIG:
      NEGF_BIG_LAUNCH(k_adjoint_big, st, a);
  }
#undef ADJ_CASE
  return fail(NEGF_ERR_ARG, "block size class %d unsupported", nmax);
}

}  // namespace negf

using namespace negf;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int negf_version(void) { return NEGF_ABI_VERSION; }

const char* negf_last_error(void) { return g_err; }

int negf_max_block(int precision) { return precision == NEGF_FP64 ? FP64_MAX : MIXED_MAX; }

int negf_max_block_green(void) { return GREEN_MAX; }

int negf_set_block_class(int precision, int n, int nmax) {
  if (precision != NEGF_FP64 && precision != NEGF_MIXED)
    return fail(NEGF_ERR_ARG, "unknown precision %d", precision);
  const bool fp64 = precision == NEGF_FP64;
  const int limit = fp64 ? FP64_MAX : 64;
  if (n < 1 || n > limit) return fail(NEGF_ERR_ARG, "block size %d outside 1..%d", n, limit);
  if (nmax != 0) {
    const bool known = nmax == 8 || nmax == 16 || nmax == 24 || nmax == 32 || nmax == 48 ||
                       (fp64 ? nmax == BIG : nmax == 64);
    if (!known || nmax < n || nmax > limit)
      return fail(NEGF_ERR_ARG, "class %d cannot hold blocks of %d (classes 8, 16, 24, 32, "
                  "48, %d)", nmax, n, fp64 ? BIG : 64);
  }
  g_class_override[fp64 ? 0 : 1][n].store(nmax, std::memory_order_relaxed);
  return NEGF_OK;
}

int negf_set_kernel_family(int precision, int n, int family) {
  if (precision != NEGF_FP64 && precision != NEGF_MIXED)
    return fail(NEGF_ERR_ARG, "unknown precision %d", precision);
  if (n < 1 || n > WAVE_MAX)
    return fail(NEGF_ERR_ARG, "kernel families are chosen for blocks of 1..%d DOF (got %d)",
                WAVE_MAX, n);
  if (family != 0 && family != NEGF_FAMILY_WORKGROUP && family != NEGF_FAMILY_WAVE)
    return fail(NEGF_ERR_ARG, "kernel family %d is not 0, %d or %d", family,
                NEGF_FAMILY_WORKGROUP, NEGF_FAMILY_WAVE);
  g_family_override[precision == NEGF_FP64 ? 0 : 1][n].store(family, std::memory_order_relaxed);
  return NEGF_OK;
}

int negf_kernel_family(int precision, int n) {
  if (precision != NEGF_FP64 && precision != NEGF_MIXED) return 0;
  return use_wave(n, precision == NEGF_FP64, false) ? NEGF_FAMILY_WAVE : NEGF_FAMILY_WORKGROUP;
}

int negf_set_product_form(int n, int form) {
  if (n < 1 || n > FORM_MAX)
    return fail(NEGF_ERR_ARG, "product forms are chosen for blocks of 1..%d DOF (got %d)",
                FORM_MAX, n);
  if (form != 0 && form != NEGF_PRODUCTS_MATRIX_CORE && form != NEGF_PRODUCTS_VECTOR)
    return fail(NEGF_ERR_ARG, "product form %d is not 0, %d or %d", form,
                NEGF_PRODUCTS_MATRIX_CORE, NEGF_PRODUCTS_VECTOR);
  g_product_form[n].store(form, std::memory_order_relaxed);
  return NEGF_OK;
}

int negf_product_form(int n) {
  if (n < 1 || n > FORM_MAX) return 0;
  const int nmax = pick_nmax(n, true);
  return use_mm(nmax, n) ? NEGF_PRODUCTS_MATRIX_CORE : NEGF_PRODUCTS_VECTOR;
}

int negf_set_matrix_core_products(int mode) {
  if (mode < 0 || mode > 1)
    return fail(NEGF_ERR_ARG, "matrix-core product mode %d is not 0 or 1", mode), -1;
  return g_mm_products.exchange(mode);
}

int negf_set_wave_kernels(int mode) {
  if (mode < 0 || mode > 2) return fail(NEGF_ERR_ARG, "wave-kernel mode %d is not 0, 1 or 2", mode), -1;
  return g_wave_kernels.exchange(mode);
}

// iteration limit of a mixed-precision FP32 decimation pass; the FP64 redo of what it
// leaves unconverged or uncertified runs the caller's limit
static int fp32_maxiter(int maxiter) {
  return maxiter < NEGF_MIXED_FP32_MAXITER ? maxiter : NEGF_MIXED_FP32_MAXITER;
}

int negf_block_class(int precision, int n) {
  if (precision != NEGF_FP64 && precision != NEGF_MIXED) return 0;
  return pick_nmax(n, precision == NEGF_FP64);
}

int negf_sancho_rubio_batched(void* stream, int precision, int n, int n_inst, const double* lead,
                              const int32_t* set_of, const double* omega, const double* eta,
                              double tol, int maxiter, int refine_steps, double* g_out,
                              double* dos_out, int32_t* iters, int32_t* status) {
  if (n <= 0 || n_inst < 0 || maxiter < 0)
    return fail(NEGF_ERR_ARG, "invalid sizes (n=%d n_inst=%d maxiter=%d)", n, n_inst, maxiter);
  if (!lead || !set_of || !omega || !eta || !g_out || !iters || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const bool fp64 = precision == NEGF_FP64;
  const int nmax = pick_nmax(n, fp64);
  if (!nmax) return fail(NEGF_ERR_ARG, "block size %d exceeds the supported maximum", n);
  SrArgs a{n, n_inst, lead, set_of, omega, eta, tol, maxiter, g_out, dos_out, iters, status, 0};
  int rc;
  if (!fp64) a.maxiter = fp32_maxiter(maxiter);
  if (fp64) {
    rc = run_sr<double>(st, nmax, a);
  } else if (precision == NEGF_MIXED) {
    if (n > MIXED_MAX)
      return fail(NEGF_ERR_ARG, "block size %d exceeds the mixed-precision limit of %d", n,
                  MIXED_MAX);
    rc = run_sr<float>(st, nmax, a);
    if (rc == NEGF_OK && refine_steps >= 0) {
      rc = negf_surface_refine_batched(stream, n, n_inst, lead, set_of, omega, eta, refine_steps,
                                       g_out, dos_out, status);
      if (rc == NEGF_OK) rc = negf_sancho_rubio_redo(stream, n, n_inst, lead, set_of, omega, eta,
                                                    tol, maxiter, g_out, dos_out, iters, status);
    }
  } else {
    return fail(NEGF_ERR_ARG, "unknown precision %d", precision);
  }
  return rc;
}

int negf_surface_gf_eigen_batched(void* stream, int n, int n_inst, const double* lead,
                                  const int32_t* set_of, const double* omega, const double* eta,
                                  double* g_io, double* dos_out, int32_t* status) {
  if (n <= 0 || n_inst < 0) return fail(NEGF_ERR_ARG, "invalid sizes (n=%d n_inst=%d)", n, n_inst);
  if (!lead || !set_of || !omega || !eta || !g_io || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0 || n > BIG) return NEGF_OK;
  EigArgs a{n, n_inst, lead, set_of, omega, eta, g_io, dos_out, status};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (2 * n <= BIG) NEGF_BIG_LAUNCH(k_sgf_eigen_big, st, a);
  NEGF_POOL_LAUNCH(k_sgf_eigen_huge, st, a, g_huge, HUGE_SLOT);
}

int negf_surface_refine_batched(void* stream, int n, int n_inst, const double* lead,
                                const int32_t* set_of, const double* omega, const double* eta,
                                int steps, double* g_io, double* dos_out, int32_t* status) {
  if (n <= 0 || n_inst < 0 || steps < 0) return fail(NEGF_ERR_ARG, "invalid refinement sizes");
  if (!lead || !set_of || !omega || !eta || !g_io || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const int nm64 = pick_nmax(n, true);
  if (!nm64 || nm64 > MIXED_MAX)
    return fail(NEGF_ERR_ARG, "block size %d exceeds the FP64 refinement limit of %d", n, MIXED_MAX);
  RefineArgs r{n,          n_inst,           lead,           set_of,  omega, eta,
               steps,      NEGF_NEWTON_RTOL, NEGF_NEWTON_XTOL, NEGF_NEWTON_DTOL, g_io, dos_out,
               status};
  return run_refine(reinterpret_cast<hipStream_t>(stream), nm64, r);
}

int negf_sancho_rubio_redo(void* stream, int n, int n_inst, const double* lead,
                           const int32_t* set_of, const double* omega, const double* eta,
                           double tol, int maxiter, double* g_out, double* dos_out, int32_t* iters,
                           int32_t* status) {
  if (n <= 0 || n_inst < 0) return fail(NEGF_ERR_ARG, "invalid sizes");
  if (n_inst == 0) return NEGF_OK;
  const int nm64 = pick_nmax(n, true);
  if (!nm64) return fail(NEGF_ERR_ARG, "block size %d exceeds the FP64 limit", n);
  SrArgs a{n, n_inst, lead, set_of, omega, eta, tol, maxiter, g_out, dos_out, iters, status, 1};
  return run_sr<double>(reinterpret_cast<hipStream_t>(stream), nm64, a);
}

int negf_self_energy_batched(void* stream, int n_lead, int n_c, int n_dev, int n_inst,
                             const double* v, const int32_t* set_of, const double* g,
                             const int32_t* idx, const int32_t* status_in, double* sigma_out) {
  if (n_lead <= 0 || n_c <= 0 || n_dev <= 0 || n_c > n_dev || n_inst < 0)
    return fail(NEGF_ERR_ARG, "invalid self-energy sizes (lead %d, contact %d, layer %d)", n_lead,
                n_c, n_dev);
  if (!v || !set_of || !g || !idx || !sigma_out) return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  int m = n_lead > n_dev ? n_lead : n_dev;
  const int nmax = pick_nmax(m, true);
  if (!nmax) return fail(NEGF_ERR_ARG, "self-energy block %d exceeds %d", m, FP64_MAX);
  SigmaArgs a{n_lead, n_c, n_dev, n_inst, v, set_of, g, idx, status_in, sigma_out};
  return run_sigma(reinterpret_cast<hipStream_t>(stream), nmax, a);
}

int negf_rgf_g1n_batched(void* stream, int precision, int n, int n_layers, int n_inst,
                         const double* onsite, const double* coupling, const int32_t* set_of,
                         const double* omega, const double* eta_dev, const double* sigma_l,
                         const double* sigma_r, const int32_t* status_in, double* g1n_out,
                         int32_t* status) {
  if (n <= 0 || n_layers <= 0 || n_inst < 0)
    return fail(NEGF_ERR_ARG, "invalid RGF sizes (n=%d layers=%d)", n, n_layers);
  if (!onsite || (n_layers > 1 && !coupling) || !set_of || !omega || !eta_dev || !sigma_l ||
      !sigma_r || !g1n_out || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const bool fp64 = precision == NEGF_FP64;
  const int nmax = pick_nmax(n, fp64);
  if (!nmax) return fail(NEGF_ERR_ARG, "block size %d exceeds the supported maximum", n);
  RgfArgs a{n, n_layers, n_inst, onsite, coupling, set_of, omega, eta_dev,
            sigma_l, sigma_r, status_in, g1n_out, status, 0};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (fp64) return run_rgf<double>(st, nmax, a);
  if (precision == NEGF_MIXED) {
    const int nm64 = pick_nmax(n, true);
    if (n > MIXED_MAX)
      return fail(NEGF_ERR_ARG, "block size %d exceeds the mixed-precision limit of %d", n,
                  MIXED_MAX);
    int rc = run_rgf<float>(st, nmax, a);
    if (rc != NEGF_OK) return rc;
    a.redo = 1;  // FP32 sweep failures (non-finite / singular) are recomputed in FP64
    return run_rgf<double>(st, nm64, a);
  }
  return fail(NEGF_ERR_ARG, "unknown precision %d", precision);
}

int negf_device_green_batched(void* stream, int n, int n_layers, int n_inst, const double* onsite,
                              const double* coupling, const int32_t* set_of, const double* omega,
                              const double* eta_dev, const double* sigma_l, const double* sigma_r,
                              double* g_out, int32_t* status) {
  if (n <= 0 || n_layers <= 0 || n_inst < 0)
    return fail(NEGF_ERR_ARG, "invalid sizes (n=%d layers=%d)", n, n_layers);
  if (!onsite || (n_layers > 1 && !coupling) || !set_of || !omega || !eta_dev || !sigma_l ||
      !sigma_r || !g_out || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const int nmax = pick_nmax(n, true);
  if (!nmax || nmax > GREEN_MAX)
    return fail(NEGF_ERR_ARG, "block size %d exceeds the full Green's function limit of %d", n,
                GREEN_MAX);
  DgfArgs a{n, n_layers, n_inst, onsite, coupling, set_of, omega, eta_dev, sigma_l, sigma_r, g_out,
            status};
  return run_dgf(reinterpret_cast<hipStream_t>(stream), nmax, a);
}

size_t negf_adjoint_workspace_bytes(int n, int n_layers, int n_inst, int checkpoint) {
  if (n <= 0 || n_layers <= 0 || n_inst <= 0 || checkpoint < 0) return 0;
  const int k = checkpoint > n_layers ? n_layers : checkpoint;
  return (size_t)n_inst * adj_blocks(n_layers, k) * (size_t)n * (size_t)n * 16u;
}

int negf_adjoint_batched(void* stream, int n, int n_layers, int n_inst, const double* onsite,
                         const double* coupling, const int32_t* set_of, const double* omega,
                         const double* eta_dev, const double* sigma_l, const double* sigma_r,
                         const int32_t* status_in, const double* weight, void* workspace,
                         size_t workspace_bytes, double* t_out, double* grad_onsite,
                         double* grad_coupling, int32_t* status, double* g1n_out,
                         int checkpoint, int per_set, double* grad_ends_out) {
  if (n <= 0 || n_layers <= 0 || n_inst < 0)
    return fail(NEGF_ERR_ARG, "invalid sizes (n=%d layers=%d)", n, n_layers);
  if (checkpoint < 0) return fail(NEGF_ERR_ARG, "checkpoint interval %d is negative", checkpoint);
  if (per_set != 0 && per_set != 1) return fail(NEGF_ERR_ARG, "per_set %d is not 0 or 1", per_set);
  if (checkpoint > n_layers) checkpoint = n_layers;
  if (!onsite || (n_layers > 1 && (!coupling || !grad_coupling)) || !set_of || !omega ||
      !eta_dev || !sigma_l || !sigma_r || !t_out || !grad_onsite || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const size_t need = negf_adjoint_workspace_bytes(n, n_layers, n_inst, checkpoint);
  if (!workspace || workspace_bytes < need)
    return fail(NEGF_ERR_ARG, "workspace of %zu bytes is smaller than the %zu needed",
                workspace_bytes, need);
  const int nmax = pick_nmax(n, true);
  if (!nmax || nmax > GREEN_MAX)
    return fail(NEGF_ERR_ARG, "block size %d exceeds the adjoint limit of %d", n, GREEN_MAX);
  AdjArgs a{n, n_layers, n_inst, onsite, coupling, set_of, omega, eta_dev, sigma_l, sigma_r,
            status_in, weight, reinterpret_cast<double*>(workspace), t_out, grad_onsite,
            grad_coupling, status, g1n_out, checkpoint, per_set, grad_ends_out};
  return run_adj(reinterpret_cast<hipStream_t>(stream), nmax, a);
}

int negf_lead_adjoint_batched(void* stream, int side, int n, int n_inst, int n_lead, int n_c,
                              const double* lead, const double* v, const int32_t* idx,
                              const int32_t* set_of, const double* g, const double* g1n,
                              const double* sigma_l, const double* sigma_r,
                              const double* grad_ends, const double* weight,
                              const int32_t* status_in, int maxiter, double* grad_lead,
                              double* grad_v, int32_t* status) {
  if (side != 0 && side != 1) return fail(NEGF_ERR_ARG, "lead side %d is not 0 or 1", side);
  if (n <= 0 || n_inst < 0 || n_lead <= 0 || n_c <= 0 || n_c > n)
    return fail(NEGF_ERR_ARG, "invalid lead adjoint sizes (n=%d lead=%d contact=%d)", n, n_lead,
                n_c);
  if (maxiter <= 0) return fail(NEGF_ERR_ARG, "maxiter must be positive (got %d)", maxiter);
  if (!lead || !v || !idx || !set_of || !g || !g1n || !sigma_l || !sigma_r || !grad_ends ||
      !grad_lead || !grad_v || !status)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const int m = n > n_lead ? n : n_lead;
  const int nmax = default_nmax(m, true);
  if (!nmax || nmax > 48)
    return fail(NEGF_ERR_ARG, "lead gradients take blocks of up to 48 DOF (got %d)", m);
  // the contact layer's onsite gradient: [0] (left) or [1] (right) of [n_inst][2][n][n]
  const size_t nn = (size_t)n * n;
  LeadAdjArgs a{n, n_inst, side, n_lead, n_c, lead, v, idx, set_of, g, g1n, sigma_l, sigma_r,
                grad_ends + (side ? nn * 2 : 0), 2 * nn * 2, weight, status_in, maxiter,
                grad_lead, grad_v, status};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define LA_CASE(NM)                                                                          \
  case NM: {                                                                                 \
    constexpr int NT = Cfg<NM>::NT;                                                          \
    NEGF_LAUNCH((k_lead_adjoint<NM>), n_inst, NT, (Lds<double, NM, NT, 4>::bytes()), st, a); \
    return NEGF_OK;                                                                          \
  }
  switch (nmax) {
    LA_CASE(8)
    LA_CASE(16)
    LA_CASE(24)
    LA_CASE(32)
    LA_CASE(48)
  }
#undef LA_CASE
  return fail(NEGF_ERR_ARG, "lead gradients take blocks of up to 48 DOF (got %d)", m);
}

int negf_fcs_cumulants(void* stream, int n_w, const double* omega_rad_s, const double* t_mean,
                       const double* t2_mean, int n_pairs, const double* temps_left,
                       const double* temps_right, double* out) {
  if (n_w < 2 || n_pairs < 0)
    return fail(NEGF_ERR_ARG, "invalid sizes (n_w=%d pairs=%d)", n_w, n_pairs);
  if (n_pairs == 0) return NEGF_OK;
  if (!omega_rad_s || !t_mean || !temps_left || !temps_right || !out)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  hipLaunchKernelGGL(k_fcs_cumulants, dim3(n_pairs), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n_w, omega_rad_s, t_mean, t2_mean,
                     temps_left, temps_right, out);
  return hip_check(hipGetLastError(), "k_fcs_cumulants launch");
}

int negf_kspace_blocks(void* stream, int n, int n_layers, int n_k, int t00, int t01, int t10,
                       const double* phi00, const double* phi01, const double* phi10,
                       const double* ph00, const double* ph01, const double* ph10,
                       double pinning, const double* s_lead, const double* s_dev,
                       double* lead_out, double* onsite_out, double* coupling_out) {
  if (n <= 0 || n_k < 0 || t00 <= 0 || t01 <= 0 || t10 < 0 || (s_dev && n_layers <= 0))
    return fail(NEGF_ERR_ARG, "invalid sizes (n=%d layers=%d k=%d terms=%d/%d/%d)", n, n_layers,
                n_k, t00, t01, t10);
  if (!phi00 || !phi01 || (t10 > 0 && (!phi10 || !ph10)) || !ph00 || !ph01)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if ((s_lead && !lead_out) || (s_dev && (!onsite_out || (n_layers > 1 && !coupling_out))))
    return fail(NEGF_ERR_ARG, "null output for a requested block set");
  if (!s_lead && !s_dev) return fail(NEGF_ERR_ARG, "neither lead nor device blocks requested");
  if (n_k == 0) return NEGF_OK;
  if (n_k > 2147483647 / 2) return fail(NEGF_ERR_ARG, "too many k points");
  const int nl = s_dev ? n_layers : 0;
  const int nblk = 3 + nl + (nl > 1 ? nl - 1 : 0);
  if (nblk > 65535) return fail(NEGF_ERR_ARG, "too many layers (%d)", n_layers);
  KspaceArgs a{n, nl, n_k, t00, t01, t10, phi00, phi01, phi10, ph00, ph01, ph10, pinning,
               s_lead, s_dev, lead_out, onsite_out, coupling_out};
  hipLaunchKernelGGL(k_kspace_blocks, dim3(n_k, nblk), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), a);
  return hip_check(hipGetLastError(), "k_kspace_blocks launch");
}

int negf_conductance_weights(void* stream, int n_w, const double* omega_rad_s, int n_t,
                             const double* temps, double* weights_out) {
  if (n_w < 0 || n_t < 0) return fail(NEGF_ERR_ARG, "invalid sizes (n_w=%d n_t=%d)", n_w, n_t);
  if (n_w == 0 || n_t == 0) return NEGF_OK;
  if (!omega_rad_s || !temps || !weights_out) return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_t > 65535) return fail(NEGF_ERR_ARG, "at most 65535 temperatures per call");
  hipLaunchKernelGGL(k_conductance_weights, dim3((n_w + 255) / 256, n_t), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n_w, omega_rad_s, n_t, temps,
                     weights_out);
  return hip_check(hipGetLastError(), "k_conductance_weights launch");
}

int negf_transmission_trace_batched(void* stream, int n, int n_inst, const double* sigma_l,
                                    const double* sigma_r, const double* g1n,
                                    const int32_t* status_in, double* t_out) {
  if (n <= 0 || n_inst < 0) return fail(NEGF_ERR_ARG, "invalid trace sizes");
  if (!sigma_l || !sigma_r || !g1n || !t_out) return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  const int nmax = pick_nmax(n, true);
  if (!nmax) return fail(NEGF_ERR_ARG, "trace block %d exceeds %d", n, FP64_MAX);
  TraceArgs a{n, n_inst, sigma_l, sigma_r, g1n, status_in, t_out};
  return run_trace(reinterpret_cast<hipStream_t>(stream), nmax, a);
}

int negf_transmission_sweep_batched(
    void* stream, int n, int n_layers, int n_inst, const double* onsite, const double* coupling,
    const int32_t* set_of, const double* omega, const double* eta_dev, int n_lead_l, int n_c_l,
    const double* v_l, const double* g_l, const int32_t* idx_l, int32_t* status_l, int n_lead_r,
    int n_c_r, const double* v_r, const double* g_r, const int32_t* idx_r,
    const int32_t* status_r, double* sigma_l, double* sigma_r, double* g1n_out, int32_t* status,
    double* t_out) {
  if (n <= 0 || n_layers <= 0 || n_inst < 0)
    return fail(NEGF_ERR_ARG, "invalid sweep sizes (n=%d layers=%d)", n, n_layers);
  if (n_lead_l <= 0 || n_c_l <= 0 || n_c_l > n || n_lead_r <= 0 || n_c_r <= 0 || n_c_r > n)
    return fail(NEGF_ERR_ARG, "invalid contact sizes (lead %d/%d, contact %d/%d, layer %d)",
                n_lead_l, n_lead_r, n_c_l, n_c_r, n);
  if (!onsite || (n_layers > 1 && !coupling) || !set_of || !omega || !eta_dev || !v_l || !g_l ||
      !idx_l || !status_l || !v_r || !g_r || !idx_r || !status_r || !sigma_l || !sigma_r ||
      !g1n_out || !status || !t_out)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_inst == 0) return NEGF_OK;
  if (!pick_nmax(n, true)) return fail(NEGF_ERR_ARG, "block size %d exceeds the supported maximum", n);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // one class holds the device block and both lead blocks (the persistent launch picks
  // its class the same way)
  int m = n > n_lead_l ? n : n_lead_l;
  m = m > n_lead_r ? m : n_lead_r;
  const int nmax = pick_nmax(m, true);
  const bool fused = nmax && !use_wave(n, true, true) && use_mm(nmax, n);
  if (fused) {
    RgfArgs a{n, n_layers, n_inst, onsite, coupling, set_of, omega, eta_dev,
              sigma_l, sigma_r, nullptr, g1n_out, status, 0};
    RgfFuse f{SigmaArgs{n_lead_l, n_c_l, n, n_inst, v_l, set_of, g_l, idx_l, status_l, sigma_l},
              SigmaArgs{n_lead_r, n_c_r, n, n_inst, v_r, set_of, g_r, idx_r, status_r, sigma_r},
              t_out};
#define RGF_CASE(NM)                                                                         \
  case NM: {                                                                                 \
    constexpr int NT = Cfg<NM>::NT;                                                          \
    NEGF_LAUNCH((k_rgf_mm<NM, true>), n_inst, NT, (Lds<double, NM, NT, 3>::bytes()), st, a, f); \
    return NEGF_OK;                                                                          \
  }
    switch (nmax) {
      RGF_CASE(8)
      RGF_CASE(16)
      RGF_CASE(24)
      RGF_CASE(32)
      RGF_CASE(48)
    }
#undef RGF_CASE
  }
  int rc = negf_self_energy_batched(stream, n_lead_l, n_c_l, n, n_inst, v_l, set_of, g_l, idx_l,
                                    status_l, sigma_l);
  if (rc) return rc;
  rc = negf_self_energy_batched(stream, n_lead_r, n_c_r, n, n_inst, v_r, set_of, g_r, idx_r,
                                status_r, sigma_r);
  if (rc) return rc;
  rc = negf_merge_status(stream, n_inst, status_l, status_r);
  if (rc) return rc;
  rc = negf_rgf_g1n_batched(stream, NEGF_FP64, n, n_layers, n_inst, onsite, coupling, set_of, omega,
                            eta_dev, sigma_l, sigma_r, status_l, g1n_out, status);
  if (rc) return rc;
  return negf_transmission_trace_batched(stream, n, n_inst, sigma_l, sigma_r, g1n_out, status,
                                         t_out);
}

// a mirrored lead pair (right lead = left lead with d01 and d10 exchanged) whose blocks
// fit the mirrored decimation: one decimation yields both surfaces
static bool sr_mirror_fits(const negf_transmission_problem* p) {
  if (p->share_leads != NEGF_LEADS_MIRRORED || p->n_lead_r != p->n_lead_l) return false;
  if (p->precision == NEGF_FP64) return sr_mirror_class(pick_nmax(p->n_lead_l, true));
  if (p->precision == NEGF_MIXED)
    return p->n_lead_l <= MIXED_MAX && sr_mirror_class(pick_nmax(p->n_lead_l, false)) &&
           sr_mirror_class(pick_nmax(p->n_lead_l, true));
  return false;
}

// g_L = (z - e_s)^-1 and g_R = (z - e~_s)^-1 from one decimation of the left lead;
// mixed precision refines each surface against its own lead and redoes the failures
// of either in FP64, as negf_sancho_rubio_batched does for one lead.
static int sr_mirrored(void* stream, const negf_transmission_problem* p, double* g_l,
                       double* g_r, int32_t* st_l, int32_t* st_r) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int n = p->n_lead_l;
  if (p->maxiter < 0) return fail(NEGF_ERR_ARG, "invalid maxiter %d", p->maxiter);
  if (!p->lead_l || !p->set_of || !p->omega || !p->eta_lead || !p->iters_l)
    return fail(NEGF_ERR_ARG, "null pointer argument");
  SrArgs a{n,        p->n_inst,   p->lead_l, p->set_of, p->omega, p->eta_lead, p->tol,
           p->maxiter, g_l,       p->dos_l,  p->iters_l, st_l,     0,           g_r,
           p->dos_r, p->iters_r,  st_r};
  if (p->precision == NEGF_FP64) return run_sr<double, true>(st, pick_nmax(n, true), a);
  a.maxiter = fp32_maxiter(p->maxiter);
  int rc = run_sr<float, true>(st, pick_nmax(n, false), a);
  a.maxiter = p->maxiter;
  if (rc || p->refine_steps < 0) return rc;
  rc = negf_surface_refine_batched(stream, n, p->n_inst, p->lead_l, p->set_of, p->omega,
                                   p->eta_lead, p->refine_steps, g_l, p->dos_l, st_l);
  if (rc) return rc;
  rc = negf_surface_refine_batched(stream, n, p->n_inst, p->lead_r, p->set_of, p->omega,
                                   p->eta_lead, p->refine_steps, g_r, p->dos_r, st_r);
  if (rc) return rc;
  a.redo = 1;
  return run_sr<double, true>(st, pick_nmax(n, true), a);
}

size_t negf_transmission_workspace_bytes(const negf_transmission_problem* p) {
  if (!p) return 0;
  const size_t ni = (size_t)p->n_inst;
  const size_t nll = (size_t)p->n_lead_l * p->n_lead_l, nlr = (size_t)p->n_lead_r * p->n_lead_r;
  const size_t nd = (size_t)p->n * p->n;
  size_t b = 0;
  b += ni * nll * 16;                              // g_L
  b += p->share_leads == NEGF_LEADS_SHARED ? 0 : ni * nlr * 16;  // g_R
  b += 2 * ni * nd * 16;                           // Sigma_L, Sigma_R
  b += ni * nd * 16;                               // G_1N
  b += 2 * ni * sizeof(int32_t);                   // lead status L/R
  return b + 256;
}

int negf_transmission_batched(void* stream, const negf_transmission_problem* p, void* workspace,
                              size_t workspace_bytes) {
  if (!p) return fail(NEGF_ERR_ARG, "null problem");
  if (p->n_inst == 0) return NEGF_OK;
  const size_t need = negf_transmission_workspace_bytes(p);
  if (!workspace || workspace_bytes < need)
    return fail(NEGF_ERR_ARG, "workspace too small (%zu < %zu bytes)", workspace_bytes, need);
  const size_t ni = (size_t)p->n_inst;
  unsigned char* w = reinterpret_cast<unsigned char*>(workspace);
  double* g_l = reinterpret_cast<double*>(w);
  w += ni * p->n_lead_l * p->n_lead_l * 16;
  double* g_r = g_l;
  if (p->share_leads != NEGF_LEADS_SHARED) {
    g_r = reinterpret_cast<double*>(w);
    w += ni * p->n_lead_r * p->n_lead_r * 16;
  }
  const size_t nd = (size_t)p->n * p->n;
  double* sig_l = reinterpret_cast<double*>(w);
  w += ni * nd * 16;
  double* sig_r = reinterpret_cast<double*>(w);
  w += ni * nd * 16;
  double* g1n = reinterpret_cast<double*>(w);
  w += ni * nd * 16;
  int32_t* st_l = reinterpret_cast<int32_t*>(w);
  w += ni * sizeof(int32_t);
  int32_t* st_r = reinterpret_cast<int32_t*>(w);

  if (p->surface_route != NEGF_SURFACE_DECIMATION && p->surface_route != NEGF_SURFACE_EIGEN)
    return fail(NEGF_ERR_ARG, "unknown surface route %d", p->surface_route);
  if (p->ladder)
    return fail(NEGF_ERR_ARG, "eta ladders are solved by negf_transmission_multi_batched");
  if (p->share_leads < 0 || p->share_leads > NEGF_LEADS_MIRRORED)
    return fail(NEGF_ERR_ARG, "share_leads %d is not 0, 1 or 2", p->share_leads);
  if (p->precision == NEGF_MIXED && p->refine_steps < 0)
    return fail(NEGF_ERR_ARG, "mixed precision needs refine_steps >= 0 (got %d)", p->refine_steps);
  if (p->share_leads != NEGF_LEADS_SHARED && (!p->lead_r || !p->iters_r))
    return fail(NEGF_ERR_ARG, "null pointer argument (right lead)");
  const bool eig = p->surface_route == NEGF_SURFACE_EIGEN;
  int rc;
  if (sr_mirror_fits(p)) {
    rc = sr_mirrored(stream, p, g_l, g_r, st_l, st_r);
    if (rc) return rc;
  } else {
    rc = negf_sancho_rubio_batched(stream, p->precision, p->n_lead_l, p->n_inst, p->lead_l,
                                   p->set_of, p->omega, p->eta_lead, p->tol, p->maxiter,
                                   p->refine_steps, g_l, p->dos_l, p->iters_l, st_l);
    if (rc) return rc;
    if (p->share_leads != NEGF_LEADS_SHARED) {
      rc = negf_sancho_rubio_batched(stream, p->precision, p->n_lead_r, p->n_inst, p->lead_r,
                                     p->set_of, p->omega, p->eta_lead, p->tol, p->maxiter,
                                     p->refine_steps, g_r, p->dos_r, p->iters_r, st_r);
      if (rc) return rc;
    }
  }
  if (eig) {
    rc = negf_surface_gf_eigen_batched(stream, p->n_lead_l, p->n_inst, p->lead_l, p->set_of,
                                       p->omega, p->eta_lead, g_l, p->dos_l, st_l);
    if (rc) return rc;
  }
  if (p->share_leads != NEGF_LEADS_SHARED) {
    if (eig) {
      rc = negf_surface_gf_eigen_batched(stream, p->n_lead_r, p->n_inst, p->lead_r, p->set_of,
                                         p->omega, p->eta_lead, g_r, p->dos_r, st_r);
      if (rc) return rc;
    }
  } else {
    hipError_t e = hipMemcpyAsync(st_r, st_l, ni * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                  reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_check(e, "status copy");
  }
  // self-energies, the device sweep (FP64 in both precisions) and the trace; the lead
  // status is merged left first, as the reference evaluates the left lead first
  return negf_transmission_sweep_batched(
      stream, p->n, p->n_layers, p->n_inst, p->onsite, p->coupling, p->set_of, p->omega,
      p->eta_dev, p->n_lead_l, p->n_c_l, p->v_l, g_l, p->idx_l, st_l, p->n_lead_r, p->n_c_r, p->v_r,
      g_r, p->idx_r, st_r, sig_l, sig_r, g1n, p->status, p->t_out);
}

// ---------------------------------------------------------------------------
// allocation-free transmission: the workspace is a file-scope device array
// ---------------------------------------------------------------------------
//
// negf_transmission_batched_noalloc runs exactly negf_transmission_batched with the
// workspace carved from `scratch` (NEGF_NOALLOC_SCRATCH_BYTES per GPU, held from the
// library's load on that device until the process ends) instead of a caller buffer.
// `scratch` names a host-side stand-in in host code; its device address on each GPU
// comes from hipGetSymbolAddress, looked up once per device on the first call with
// that device current and kept in g_scratch_ptr.  Blocks that need the 96-class
// device-memory pool (more than 48 DOF) or the eigenmode route (also pooled) are
// rejected, because that pool is allocated on first use.  Nothing in the workspace is
// read before the same call writes it, so no clearing is needed.  Calls that overlap
// on one GPU (different streams or threads) race on the array.
static __device__ __align__(256) unsigned char scratch[NEGF_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_on_current_device(void** out) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= 64) return fail(NEGF_ERR_ARG, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!ptr) return fail(NEGF_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return NEGF_OK;
}

int negf_transmission_batched_noalloc(void* stream, const negf_transmission_problem* p) {
  if (!p) return fail(NEGF_ERR_ARG, "null problem");
  if (p->n_inst == 0) return NEGF_OK;
  const size_t need = negf_transmission_workspace_bytes(p);
  if (need > (size_t)NEGF_NOALLOC_SCRATCH_BYTES)
    return fail(NEGF_ERR_ARG,
                "problem needs %zu workspace bytes, above the %d-byte scratch of the "
                "allocation-free entry point: split the instances into smaller calls",
                need, NEGF_NOALLOC_SCRATCH_BYTES);
  if (p->surface_route != NEGF_SURFACE_DECIMATION)
    return fail(NEGF_ERR_ARG, "the allocation-free entry point takes the decimation route only");
  // every FP64 class the staged path picks (leads, self-energies, sweep, trace) must be
  // an LDS class: the 96 class (also reachable through negf_set_block_class) is pooled
  const int sizes[5] = {p->n, p->n_lead_l, p->n_lead_r, p->n_lead_l > p->n ? p->n_lead_l : p->n,
                        p->n_lead_r > p->n ? p->n_lead_r : p->n};
  for (int q = 0; q < 5; ++q) {
    const int c = pick_nmax(sizes[q], true);
    if (sizes[q] < 1 || c == 0 || c > MIXED_MAX)
      return fail(NEGF_ERR_ARG,
                  "blocks of %d DOF run in block class %d, which uses the device-memory pool the "
                  "allocation-free entry point does not (limit: class %d)", sizes[q], c, MIXED_MAX);
  }
  void* ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  return negf_transmission_batched(stream, p, ws, (size_t)NEGF_NOALLOC_SCRATCH_BYTES);
}

// ---------------------------------------------------------------------------
// several problems of different block sizes in one persistent launch
// ---------------------------------------------------------------------------

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// workspace of one problem: the staged path's, plus the current rung's etas of a ladder
static size_t pl_problem_bytes(const negf_transmission_problem& p) {
  const size_t ladder = p.ladder ? align256(2 * (size_t)p.n_inst * sizeof(double)) : 0;
  return align256(negf_transmission_workspace_bytes(&p)) + ladder;
}

size_t negf_transmission_multi_workspace_bytes(int n_problems,
                                               const negf_transmission_problem* problems) {
  if (n_problems <= 0 || !problems) return 0;
  size_t b = align256(sizeof(PlArgs)) + 256;  // problem table, ticket counter
  for (int i = 0; i < n_problems; ++i) b += pl_problem_bytes(problems[i]);
  return b;
}

static int pl_check(const negf_transmission_problem& p, int i) {
  if (p.precision != NEGF_FP64 && p.precision != NEGF_MIXED)
    return fail(NEGF_ERR_ARG, "problem %d: unknown precision %d", i, p.precision);
  if (p.precision == NEGF_MIXED) {
    if (p.ladder)
      return fail(NEGF_ERR_ARG, "problem %d: eta ladders run in FP64 (the ladder is walked "
                  "inside the persistent launch, the mixed decimation is not)", i);
    if (p.refine_steps < 0)
      return fail(NEGF_ERR_ARG, "problem %d: mixed precision needs refine_steps >= 0", i);
    const int ml = p.n_lead_l > p.n_lead_r ? p.n_lead_l : p.n_lead_r;
    if (ml > MIXED_MAX)
      return fail(NEGF_ERR_ARG, "problem %d: lead blocks of %d DOF exceed the mixed-precision "
                  "limit of %d", i, ml, MIXED_MAX);
  }
  if (p.n <= 0 || p.n_layers <= 0 || p.n_inst < 0 || p.n_lead_l <= 0 || p.n_lead_r <= 0 ||
      p.n_c_l <= 0 || p.n_c_r <= 0 || p.n_c_l > p.n || p.n_c_r > p.n || p.maxiter < 0)
    return fail(NEGF_ERR_ARG,
                "problem %d: invalid sizes (n=%d layers=%d n_inst=%d leads %d/%d contacts %d/%d "
                "maxiter=%d)",
                i, p.n, p.n_layers, p.n_inst, p.n_lead_l, p.n_lead_r, p.n_c_l, p.n_c_r, p.maxiter);
  if (p.share_leads < 0 || p.share_leads > NEGF_LEADS_MIRRORED)
    return fail(NEGF_ERR_ARG, "problem %d: share_leads %d is not 0, 1 or 2", i, p.share_leads);
  if (p.share_leads && p.n_lead_r != p.n_lead_l)
    return fail(NEGF_ERR_ARG, "problem %d: shared leads of different sizes", i);
  if (p.surface_route != NEGF_SURFACE_DECIMATION && p.surface_route != NEGF_SURFACE_EIGEN)
    return fail(NEGF_ERR_ARG, "problem %d: unknown surface route %d", i, p.surface_route);
  int m = p.n > p.n_lead_l ? p.n : p.n_lead_l;
  m = m > p.n_lead_r ? m : p.n_lead_r;
  if (m > BIG) return fail(NEGF_ERR_ARG, "problem %d: block size %d exceeds %d", i, m, BIG);
  if (p.ladder) {
    const negf_eta_ladder& l = *p.ladder;
    if (p.surface_route != NEGF_SURFACE_DECIMATION)
      return fail(NEGF_ERR_ARG, "problem %d: eta ladders run with the decimation route only", i);
    if (l.max_rungs < 1 || !(l.nonneg_tol >= 0.0))
      return fail(NEGF_ERR_ARG, "problem %d: invalid eta ladder (max_rungs=%d)", i, l.max_rungs);
    if (p.n_inst > 0 && (!l.etas || !l.first || !l.count || !l.rung_out || !l.t_rungs ||
                         !l.status_rungs))
      return fail(NEGF_ERR_ARG, "problem %d: null pointer in the eta ladder", i);
  }
  if (p.n_inst == 0) return NEGF_OK;
  if (!p.lead_l || (p.share_leads != NEGF_LEADS_SHARED && !p.lead_r) || !p.v_l || !p.v_r || !p.idx_l || !p.idx_r ||
      !p.onsite || (p.n_layers > 1 && !p.coupling) || !p.set_of || !p.omega ||
      (!p.ladder && !p.eta_lead) || ((!p.ladder || p.ladder->eta_dev_fixed) && !p.eta_dev) ||
      !p.t_out || !p.iters_l || (p.share_leads != NEGF_LEADS_SHARED && !p.iters_r) || !p.status)
    return fail(NEGF_ERR_ARG, "problem %d: null pointer argument", i);
  return NEGF_OK;
}

static int pl_launch(hipStream_t st, const PlArgs* table, int* counter, int phase, bool reset,
                     double* pool, int grid, size_t lds) {
  if (reset) {
    hipLaunchKernelGGL(k_pl_reset, dim3(1), dim3(1), 0, st, counter);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_check(e, "k_pl_reset launch");
  }
  hipLaunchKernelGGL(k_persistent, dim3(grid), dim3(PL_NT), lds, st, table, phase, pool);
  return hip_check(hipGetLastError(), "k_persistent launch");
}

int negf_transmission_multi_batched(void* stream, int n_problems,
                                    const negf_transmission_problem* problems, void* workspace,
                                    size_t workspace_bytes) {
  if (n_problems < 0 || n_problems > NEGF_MULTI_MAX_PROBLEMS)
    return fail(NEGF_ERR_ARG, "%d problems (at most %d per call)", n_problems,
                NEGF_MULTI_MAX_PROBLEMS);
  if (n_problems == 0) return NEGF_OK;
  if (!problems) return fail(NEGF_ERR_ARG, "null problem list");
  for (int i = 0; i < n_problems; ++i) {
    const int rc = pl_check(problems[i], i);
    if (rc) return rc;
  }
  const size_t need = negf_transmission_multi_workspace_bytes(n_problems, problems);
  if (!workspace || workspace_bytes < need)
    return fail(NEGF_ERR_ARG, "workspace too small (%zu < %zu bytes)", workspace_bytes, need);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);

  // problems in order of decreasing ticket cost (class^3 x decimations and layers)
  int order[NEGF_MULTI_MAX_PROBLEMS], cls[NEGF_MULTI_MAX_PROBLEMS];
  double cost[NEGF_MULTI_MAX_PROBLEMS];
  int np = 0;
  for (int i = 0; i < n_problems; ++i) {
    const negf_transmission_problem& p = problems[i];
    if (p.n_inst == 0) continue;
    int m = p.n > p.n_lead_l ? p.n : p.n_lead_l;
    m = m > p.n_lead_r ? m : p.n_lead_r;
    cls[i] = default_nmax(m, true);
    const double c = (double)cls[i];
    cost[i] = c * c * c * ((p.share_leads == NEGF_LEADS_SHARED ? 1.0 : 2.0) * 30.0 * 7.0 + 4.0 * p.n_layers);
    int k = np++;
    while (k > 0 && cost[order[k - 1]] < cost[i]) {
      order[k] = order[k - 1];
      --k;
    }
    order[k] = i;
  }
  if (np == 0) return NEGF_OK;

  PlArgs a{};
  unsigned char* w = reinterpret_cast<unsigned char*>(workspace);
  PlArgs* table = reinterpret_cast<PlArgs*>(w);
  w += align256(sizeof(PlArgs));
  a.counter = reinterpret_cast<int*>(w);
  w += 256;
  a.n_prob = np;
  int tickets = 0;
  size_t lds = 0;
  bool big = false, eig = false, ladder = false;
  for (int i = 0; i < n_problems; ++i) {
    // workspace slices in the caller's order, laid out as in negf_transmission_batched
    const negf_transmission_problem& p = problems[i];
    unsigned char* base = w;
    unsigned char* lad_ws = w + align256(negf_transmission_workspace_bytes(&p));
    w += pl_problem_bytes(p);
    if (p.n_inst == 0) continue;
    int slot = 0;
    while (order[slot] != i) ++slot;
    PlProb& q = a.prob[slot];
    const size_t ni = (size_t)p.n_inst, nd = (size_t)p.n * p.n;
    q.g_l = reinterpret_cast<double*>(base);
    base += ni * p.n_lead_l * p.n_lead_l * 16;
    q.g_r = q.g_l;
    if (p.share_leads != NEGF_LEADS_SHARED) {
      q.g_r = reinterpret_cast<double*>(base);
      base += ni * p.n_lead_r * p.n_lead_r * 16;
    }
    q.sig_l = reinterpret_cast<double*>(base);
    base += ni * nd * 16;
    q.sig_r = reinterpret_cast<double*>(base);
    base += ni * nd * 16;
    q.g1n = reinterpret_cast<double*>(base);
    base += ni * nd * 16;
    q.st_l = reinterpret_cast<int32_t*>(base);
    base += ni * sizeof(int32_t);
    q.st_r = reinterpret_cast<int32_t*>(base);
    q.n = p.n;
    q.n_layers = p.n_layers;
    q.n_inst = p.n_inst;
    q.n_lead_l = p.n_lead_l;
    q.n_lead_r = p.n_lead_r;
    q.n_c_l = p.n_c_l;
    q.n_c_r = p.n_c_r;
    q.share = p.share_leads == NEGF_LEADS_SHARED ? 1 : 0;
    q.maxiter = p.maxiter;
    q.cls = cls[i];
    q.route = p.surface_route;
    q.tol = p.tol;
    q.lead_l = p.lead_l;
    q.lead_r = p.share_leads == NEGF_LEADS_SHARED ? p.lead_l : p.lead_r;
    q.v_l = p.v_l;
    q.v_r = p.v_r;
    q.idx_l = p.idx_l;
    q.idx_r = p.idx_r;
    q.onsite = p.onsite;
    q.coupling = p.coupling ? p.coupling : p.onsite;  // never read for one layer
    q.set_of = p.set_of;
    q.omega = p.omega;
    q.eta_lead = p.eta_lead;
    q.eta_dev = p.eta_dev;
    q.t_out = p.t_out;
    q.dos_l = p.dos_l;
    q.dos_r = p.dos_r;
    q.iters_l = p.iters_l;
    q.iters_r = p.iters_r;
    q.status = p.status;
    if (p.ladder) {
      const negf_eta_ladder& l = *p.ladder;
      q.lad_eta = l.etas;
      q.lad_first = l.first;
      q.lad_count = l.count;
      q.lad_tmax = l.t_max;
      q.lad_rung = l.rung_out;
      q.lad_t = l.t_rungs;
      q.lad_st = l.status_rungs;
      q.eta_cur = reinterpret_cast<double*>(lad_ws);
      q.eta_dev_cur = q.eta_cur + ni;
      q.lad_tol = l.nonneg_tol;
      q.lad_max = l.max_rungs;
      q.lad_dev_fixed = l.eta_dev_fixed ? 1 : 0;
      ladder = true;
    }
    big = big || cls[i] == BIG;
    eig = eig || p.surface_route == NEGF_SURFACE_EIGEN;
    const size_t l = pl_lds_bytes(cls[i]);
    lds = l > lds ? l : lds;
  }
  for (int k = 0; k < np; ++k) {
    PlProb& q = a.prob[k];
    q.group = q.cls == BIG ? 1
              : q.cls == 8 ? PlCfg<8>::G
              : q.cls == 16 ? PlCfg<16>::G
              : q.cls == 24 ? PlCfg<24>::G
              : q.cls == 32 ? PlCfg<32>::G
                            : PlCfg<48>::G;
    q.ticket0 = tickets;
    tickets += (q.n_inst + q.group - 1) / q.group;
  }
  a.n_tickets = tickets;
  if (ladder && eig)
    return fail(NEGF_ERR_ARG, "eta ladders and the eigenmode route cannot share a call");

  int dev = 0, cus = 0, per = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  if (rc) return rc;
  rc = hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_persistent),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                 "k_persistent attribute");
  if (rc) return rc;
  rc = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     &per, reinterpret_cast<const void*>(&k_persistent), PL_NT, lds),
                 "k_persistent occupancy");
  if (rc) return rc;
  if (per < 1) return fail(NEGF_ERR_HIP, "k_persistent cannot be resident with %zu B of LDS", lds);
  int grid = cus * per;
  grid = tickets < grid ? tickets : grid;

  hipLaunchKernelGGL(k_pl_setup, dim3(1), dim3(64), 0, st, a, table);
  rc = hip_check(hipGetLastError(), "k_pl_setup launch");
  if (rc) return rc;
  bool mixed = false;
  for (int i = 0; i < n_problems; ++i) mixed = mixed || problems[i].precision == NEGF_MIXED;
  // one launch per phase; the eigenmode fallback (instances whose decimation
  // failed, eigenmode-route problems only) runs between the lead and device phases
  auto run_phase = [&](int phase, bool reset) -> int {
    if (!big) return pl_launch(st, table, a.counter, phase, reset, nullptr, grid, lds);
    BigUse u(st);
    if (u.rc) return u.rc;
    const int r = pl_launch(st, table, a.counter, phase, reset, u.pool(), u.grid(grid), lds);
    if (r) return r;
    return u.finish();
  };
  if (!eig && !mixed) return run_phase(PL_ALL, false);
  if (mixed) {
    // mixed precision: every problem's surface GFs by the staged decimation of its own
    // precision (FP32 decimation, FP64 Newton certificate and FP64 redo for mixed
    // problems; one decimation per mirrored pair), into the workspace slices the
    // persistent device phase reads
    for (int i = 0; i < n_problems; ++i) {
      const negf_transmission_problem& p = problems[i];
      if (p.n_inst == 0) continue;
      int slot = 0;
      while (order[slot] != i) ++slot;
      const PlProb& q = a.prob[slot];
      if (sr_mirror_fits(&p)) {
        rc = sr_mirrored(stream, &p, q.g_l, q.g_r, q.st_l, q.st_r);
        if (rc) return rc;
        continue;
      }
      rc = negf_sancho_rubio_batched(stream, p.precision, p.n_lead_l, p.n_inst, p.lead_l, p.set_of,
                                     p.omega, p.eta_lead, p.tol, p.maxiter, p.refine_steps, q.g_l,
                                     p.dos_l, p.iters_l, q.st_l);
      if (rc) return rc;
      if (p.share_leads != NEGF_LEADS_SHARED) {
        rc = negf_sancho_rubio_batched(stream, p.precision, p.n_lead_r, p.n_inst, p.lead_r,
                                       p.set_of, p.omega, p.eta_lead, p.tol, p.maxiter,
                                       p.refine_steps, q.g_r, p.dos_r, p.iters_r, q.st_r);
        if (rc) return rc;
      }
    }
  } else {
    rc = run_phase(PL_LEADS, false);
    if (rc) return rc;
  }
  for (int k = 0; k < np; ++k) {
    const PlProb& q = a.prob[k];
    if (q.route != NEGF_SURFACE_EIGEN) continue;
    rc = negf_surface_gf_eigen_batched(stream, q.n_lead_l, q.n_inst, q.lead_l, q.set_of, q.omega,
                                       q.eta_lead, q.g_l, q.dos_l, q.st_l);
    if (rc) return rc;
    if (!q.share) {
      rc = negf_surface_gf_eigen_batched(stream, q.n_lead_r, q.n_inst, q.lead_r, q.set_of,
                                         q.omega, q.eta_lead, q.g_r, q.dos_r, q.st_r);
      if (rc) return rc;
    }
  }
  return run_phase(PL_DEVICE, true);
}

}  // extern "C"

namespace negf {
__global__ void k_merge_status(int n, int32_t* a, const int32_t* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && a[i] == ST_OK) a[i] = b[i];
}
}  // namespace negf

extern "C" int negf_merge_status(void* stream, int n_inst, int32_t* st_l, const int32_t* st_r) {
  if (n_inst <= 0) return NEGF_OK;
  hipLaunchKernelGGL(negf::k_merge_status, dim3((n_inst + 255) / 256), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n_inst, st_l, st_r);
  return hip_check(hipGetLastError(), "merge status");
}

extern "C" int negf_thermal_conductance(void* stream, int n_w, const double* omega_rad_s,
                                        const double* trans, int n_t, const double* temps,
                                        double* out) {
  if (n_w < 2 || n_t < 0) return fail(NEGF_ERR_ARG, "need >= 2 frequencies");
  if (!omega_rad_s || !trans || !temps || !out) return fail(NEGF_ERR_ARG, "null pointer argument");
  if (n_t == 0) return NEGF_OK;
  hipLaunchKernelGGL(negf::k_conductance, dim3(n_t), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n_w, omega_rad_s, trans, n_t, temps,
                     out);
  return hip_check(hipGetLastError(), "conductance launch");
}

extern "C" int negf_heat_current(void* stream, int n_w, const double* omega_rad_s,
                                 const double* trans, double temp_left, double temp_right,
                                 double* out) {
  if (n_w < 2) return fail(NEGF_ERR_ARG, "need >= 2 frequencies");
  if (!(temp_left > 0.0) || !(temp_right > 0.0))
    return fail(NEGF_ERR_ARG, "temperatures must be positive");
  if (!omega_rad_s || !trans || !out) return fail(NEGF_ERR_ARG, "null pointer argument");
  hipLaunchKernelGGL(negf::k_heat_current, dim3(1), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n_w, omega_rad_s, trans, temp_left,
                     temp_right, out);
  return hip_check(hipGetLastError(), "heat current launch");
}
