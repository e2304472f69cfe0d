// One persistent launch over several transmission problems of different block
// sizes (FP64): negf_transmission_multi_batched.
//
// Every (problem, instance) pair is solved by the whole transmission pipeline of
// the staged path (negf_transmission_batched: both surface GFs by decimation, both
// self-energies, the recursive sweep for G_1N and the trace), fused per instance
// and run by one 768-thread workgroup per CU that takes tickets from an atomic
// work queue until none are left.  A ticket is G consecutive instances of one
// problem, solved side by side by G sub-blocks of that problem's block-size class
// (G x the class's own workgroup size <= 768 threads; 12 x 8, 11 x 16, 4 x 24,
// 3 x 32, 1 x 48 DOF in LDS, 1 x 96 DOF in a device-memory pool slot).  Tickets of
// the most expensive problems come first, so the queue ends on short tickets.
//
// Sub-blocks share the workgroup barrier, so each pipeline stage is written in
// lockstep form: every sub-block passes the same barriers, and a sub-block that
// has no instance, has converged, or has failed keeps passing them with its work
// switched off.  The decimation loop runs until no sub-block of the ticket is
// still iterating (a per-instance convergence flag in LDS); each instance stops
// at its own step, so its result and step count are those of the staged kernel.
// The arithmetic of every stage is the staged kernels' own, in the same order: the
// decimation and the sweep with their block products on the matrix cores (mm.hpp), as
// the staged kernels run them by default.
// Consecutive stages reuse the same LDS matrices, so each stage opens with a
// barrier: the previous stage's last reads (its output stores from LDS) must be
// done before the matrices are cleared.

#pragma once

#include "big.hpp"

namespace negf {

constexpr int PL_NT = 768;
constexpr int PL_MAX_PROBLEMS = NEGF_MULTI_MAX_PROBLEMS;
enum { PL_ALL = 0, PL_LEADS = 1, PL_DEVICE = 2 };

// sub-blocks per workgroup of each LDS class (limited by the 160 KB of LDS for 16 DOF)
template <int NMAX>
struct PlCfg;
template <>
struct PlCfg<8> {
  static constexpr int G = 12;
};
template <>
struct PlCfg<16> {
  static constexpr int G = 11;
};
template <>
struct PlCfg<24> {
  static constexpr int G = 4;
};
template <>
struct PlCfg<32> {
  static constexpr int G = 3;
};
template <>
struct PlCfg<48> {
  static constexpr int G = 1;
};

template <int NMAX>
struct PlSub {
  static constexpr int NT = Cfg<NMAX>::NT;
  static constexpr int G = PlCfg<NMAX>::G;
  using L = Lds<double, NMAX, NT, 3>;
  static constexpr size_t STRIDE = (L::bytes() + 15) & ~size_t(15);
  static constexpr size_t BYTES = STRIDE * G;
  static_assert(G * NT <= PL_NT, "sub-blocks exceed the workgroup");
};

// One problem of the launch; the workspace slices are those of the staged path.
struct PlProb {
  int n, n_layers, n_inst, n_lead_l, n_lead_r, n_c_l, n_c_r, share, maxiter;
  int cls;      // block-size class: 8, 16, 24, 32, 48 or BIG
  int group;    // instances per ticket
  int ticket0;  // first ticket of this problem
  int route;    // NEGF_SURFACE_DECIMATION or NEGF_SURFACE_EIGEN
  double tol;
  const double *lead_l, *lead_r, *v_l, *v_r;
  const int32_t *idx_l, *idx_r;
  const double *onsite, *coupling;
  const int32_t* set_of;
  const double *omega, *eta_lead, *eta_dev;
  double *t_out, *dos_l, *dos_r;
  int32_t *iters_l, *iters_r, *status;
  double *g_l, *g_r, *sig_l, *sig_r, *g1n;
  int32_t *st_l, *st_r;
  // eta ladder (lad_eta null: one eta per instance from eta_lead / eta_dev)
  const double* lad_eta;
  const int32_t *lad_first, *lad_count;
  const double* lad_tmax;
  int32_t* lad_rung;
  double* lad_t;
  int32_t* lad_st;
  double *eta_cur, *eta_dev_cur;  // the current rung's etas (workspace)
  double lad_tol;
  int lad_max, lad_dev_fixed;
};

struct PlArgs {
  int n_prob, n_tickets;
  int* counter;  // ticket counter, 0 at launch
  PlProb prob[PL_MAX_PROBLEMS];
};

struct PlShared {
  int run[16];          // sub-block still iterating (decimation loop)
  int more[16];         // sub-block retrying at its next eta rung
  int ticket;
  double idle_red[PL_NT / WAVE];  // reduction scratch of threads outside every sub-block
  int idle_flag[PL_NT / WAVE];
};

// ---------------------------------------------------------------------------
// lockstep stages (the staged kernels k_sancho_rubio, k_self_energy, k_rgf and
// k_trace in FP64, with their work switched off where a sub-block has none)
// ---------------------------------------------------------------------------

// Decimation of one lead; returns the instance status.
template <int NMAX, class L>
__device__ int pl_sr(const SrArgs& a, int inst, bool active, L& s, int tid, PlShared& ps, int sub,
                     int nsub) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using M = MOwn<NMAX, NT>;
  constexpr int TPW = M::TPW;
  const int n = a.n;
  const int ns = (n + 3) / 4;
  const size_t nn = (size_t)n * n;
  const double* blk = a.lead + (active ? (size_t)a.set_of[inst] * 3 * nn * 2 : 0);
  __syncthreads();  // the previous stage's last reads of these matrices are done
  if (active) s.zero(tid);
  __syncthreads();
  if (active) {
    gload<T, NT>(blk + nn * 2, n, n, s.re[0], s.im[0], LD, tid);
    gload<T, NT>(blk + 2 * nn * 2, n, n, s.re[1], s.im[1], LD, tid);
  }
  const M o(tid);
  T er[TPW], ei[TPW], sr[TPW], si[TPW];
  {
    const double2* d00 = reinterpret_cast<const double2*>(blk);
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      double2 v = make_double2(0.0, 0.0);
      if (active && i < n && j < n) v = d00[i * n + j];
      er[q] = v.x;
      ei[q] = v.y;
      sr[q] = v.x;
      si[q] = v.y;
    }
  }
  double zr = 0.0, zi = 0.0;
  if (active) zsq(a.omega[inst], a.eta[inst], zr, zi);
  const T tol = a.tol;
  __syncthreads();

  int status = ST_OK;
  int it = 0;
  bool conv = false;
  bool run = active;
  for (;;) {
    if (run && it >= a.maxiter) run = false;
    if (tid == 0 && sub < nsub) ps.run[sub] = run ? 1 : 0;
    __syncthreads();
    int any = 0;
    for (int q = 0; q < nsub; ++q) any |= ps.run[q];
    any = __builtin_amdgcn_readfirstlane(any);  // the same in every lane
    // the next write to ps.run follows this step's inversion barriers
    if (!any) break;
    if (run) ++it;
    // G = z - e
    if (run) {
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const int i = o.row(q), j = o.col(q);
        if (i < n && j < n) {
          s.re[2][i * LD + j] = (i == j ? zr : T(0)) - er[q];
          s.im[2][i * LD + j] = (i == j ? zi : T(0)) - ei[q];
        }
      }
    }
    __syncthreads();
    const int st = gj_inverse_ls<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid, run);
    if (run && st != ST_OK) {
      status = st;
      run = false;
    }
    // T1 = a g, T2 = b g
    T t1r[TPW], t1i[TPW], t2r[TPW], t2i[TPW];
    mzero<TPW>(t1r, t1i);
    mzero<TPW>(t2r, t2i);
    if (run)
      mm2b<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], s.re[2], s.im[2], LD, ns, o,
                               t1r, t1i, t2r, t2i);
    __syncthreads();
    if (run) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, t1r, t1i);
    __syncthreads();
    // P1 = T1 b (into e_s and e at once), P3 = T1 a (the next alpha)
    T p1r[TPW], p1i[TPW], p3r[TPW], p3i[TPW];
    mzero<TPW>(p1r, p1i);
    mzero<TPW>(p3r, p3i);
    if (run) {
      mm2a<NMAX, NT>(s.re[2], s.im[2], s.re[1], s.im[1], s.re[0], s.im[0], LD, ns, o, p1r, p1i,
                     p3r, p3i);
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        sr[q] = sr[q] + p1r[q];
        si[q] = si[q] + p1i[q];
        er[q] = er[q] + p1r[q];
        ei[q] = ei[q] + p1i[q];
      }
    }
    __syncthreads();
    if (run) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, t2r, t2i);
    __syncthreads();
    // P2 = T2 a (into e), P4 = T2 b (the next beta)
    T p2r[TPW], p2i[TPW], p4r[TPW], p4i[TPW];
    mzero<TPW>(p2r, p2i);
    mzero<TPW>(p4r, p4i);
    if (run)
      mm2a<NMAX, NT>(s.re[2], s.im[2], s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, p2r, p2i,
                     p4r, p4i);
    __syncthreads();
    bool bad = false;
    double emax = 0.0;
    if (run) {
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const int i = o.row(q), j = o.col(q);
        if (i < n && j < n) {
          er[q] = er[q] + p2r[q];
          ei[q] = ei[q] + p2i[q];
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
    }
    double gmax;
    bool gbad;
    block_max_any<NT>(emax, bad, s.sred, s.sflag, tid, gmax, gbad);
    if (run) {
      if (gbad) {
        status = ST_NONFINITE;
        run = false;
      } else if (sqrt(gmax) < tol) {
        conv = true;
        run = false;
      }
    }
  }
  if (active && status == ST_OK && !conv) status = ST_NOCONV;
  if (active && tid == 0) {
    a.status[inst] = status;
    a.iters[inst] = it;
  }
  // g_s = (z - e_s)^-1
  const bool fin = active && status == ST_OK;
  __syncthreads();  // earlier reads of plane 2 are done
  if (fin) {
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      if (i < n && j < n) {
        s.re[2][i * LD + j] = (i == j ? zr : T(0)) - sr[q];
        s.im[2][i * LD + j] = (i == j ? zi : T(0)) - si[q];
      }
    }
  }
  __syncthreads();
  const int st = gj_inverse_ls<T, NT, NMAX, L::RG>(s.re[2], s.im[2], LD, n, *s.gj, tid, fin);
  if (fin && st != ST_OK) {
    status = st;
    if (tid == 0) a.status[inst] = st;
  }
  if (fin && st == ST_OK) {
    gstore<T, NT>(a.g_out + (size_t)inst * nn * 2, n, s.re[2], s.im[2], LD, tid);
    if (tid == 0 && a.dos_out) {
      double tr = 0.0;
      for (int i = 0; i < n; ++i) tr += -2.0 * s.im[2][i * LD + i];
      a.dos_out[inst] = (a.omega[inst] / M_PI) * tr;
    }
  }
  return status;
}

// Sigma = V g V^H embedded in the end layer (skipped unless live): the fused sweep's
// arithmetic (sigma_mm), so both launches give identical self-energies.
template <int NMAX, class L>
__device__ void pl_sigma(const SigmaArgs& a, int inst, bool live, L& s, int tid) {
  sigma_mm<NMAX>(a, inst, a.n_dev, s, tid, live);
}

// Recursive sweep for G_1N; returns the device status (st_in when nonzero).
template <int NMAX, class L>
__device__ int pl_rgf(const RgfArgs& a, int inst, bool active, int st_in, L& s, int tid) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using M = MOwn<NMAX, NT>;
  constexpr int TPW = M::TPW;
  bool live = active && st_in == ST_OK;
  const int n = a.n, nl = a.n_layers;
  const int ns = (n + 3) / 4;
  const size_t nn = (size_t)n * n;
  const size_t set = active ? (size_t)a.set_of[inst] : 0;
  const double2* ons = reinterpret_cast<const double2*>(a.onsite + set * nl * nn * 2);
  const double* cpl = a.coupling + set * (nl > 1 ? nl - 1 : 0) * nn * 2;
  const size_t io = active ? (size_t)inst * nn * 2 : 0;
  const double2* sgl = reinterpret_cast<const double2*>(a.sigma_l + io);
  const double2* sgr = reinterpret_cast<const double2*>(a.sigma_r + io);
  // LDS: [0] g (inverted in place), then Y = g H; [1] H; [2] Q = G_{1,i} H_i
  __syncthreads();  // the previous stage's last reads of these matrices are done
  if (live) s.zero(tid);
  __syncthreads();
  const M o(tid);
  double zr = 0.0, zi = 0.0;
  if (active) zsq(a.omega[inst], a.eta_dev[inst], zr, zi);
  T sr[TPW], si[TPW];
  mzero<TPW>(sr, si);
  int status = ST_OK;
  for (int layer = 0; layer < nl; ++layer) {
    if (live) {
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
      if (live) {
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
    }
    __syncthreads();
    const int st = gj_inverse_ls<T, NT, NMAX, L::RG>(s.re[0], s.im[0], LD, n, *s.gj, tid, live);
    if (live && st != ST_OK) {
      status = 10 + st;
      live = false;
    }
    if (layer == nl - 1) {
      // G_1N = Q_{N-1} g_N (g_1 itself for one layer), through LDS to a coalesced store
      if (nl > 1) {
        T gr[TPW], gi[TPW];
        mzero<TPW>(gr, gi);
        if (live) {
          mm<NMAX, NT, OPN, OPN>(s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o, gr, gi);
          mstore<NMAX, NT>(s.re[1], s.im[1], LD, n, o, gr, gi);
        }
        __syncthreads();
        if (live) gstore<T, NT>(a.g1n_out + (size_t)inst * nn * 2, n, s.re[1], s.im[1], LD, tid);
      } else if (live) {
        gstore<T, NT>(a.g1n_out + (size_t)inst * nn * 2, n, s.re[0], s.im[0], LD, tid);
      }
      break;
    }
    if (live) gload<T, NT>(cpl + (size_t)layer * nn * 2, n, n, s.re[1], s.im[1], LD, tid);
    __syncthreads();
    T yr[TPW], yi[TPW];
    mzero<TPW>(yr, yi);
    if (live) mm<NMAX, NT, OPN, OPN>(s.re[0], s.im[0], s.re[1], s.im[1], LD, ns, o, yr, yi);
    __syncthreads();  // every read of g is done: Y replaces it (and starts Q at layer 0)
    if (live) {
      mstore<NMAX, NT>(s.re[0], s.im[0], LD, n, o, yr, yi);
      if (layer == 0) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, yr, yi);
    }
    __syncthreads();
    mzero<TPW>(sr, si);
    if (layer > 0) {
      T qr[TPW], qi[TPW];
      mzero<TPW>(qr, qi);
      if (live)
        mm2b<NMAX, NT, OPH, OPN>(s.re[1], s.im[1], s.re[2], s.im[2], s.re[0], s.im[0], LD, ns, o,
                                 sr, si, qr, qi);
      __syncthreads();
      if (live) mstore<NMAX, NT>(s.re[2], s.im[2], LD, n, o, qr, qi);
    } else {
      if (live) mm<NMAX, NT, OPH, OPN>(s.re[1], s.im[1], s.re[0], s.im[0], LD, ns, o, sr, si);
      __syncthreads();
    }
  }
  return st_in != ST_OK ? st_in : status;
}

// T = Re Tr[Gamma_L G_1N Gamma_R G_1N^H] (NaN for a failed instance); returns it to
// every thread of the sub-block.  G_1N comes back from global memory into LDS matrix [1]
// and the registers of the matrix-core ownership; the product and sum are the fused
// sweep's (trace_mm), so both launches give identical T.
template <int NMAX, class L>
__device__ double pl_trace(const TraceArgs& a, int inst, bool active, int status, L& s, int tid) {
  using T = double;
  constexpr int NT = Cfg<NMAX>::NT;
  constexpr int LD = L::LD;
  using M = MOwn<NMAX, NT>;
  constexpr int TPW = M::TPW;
  const bool live = active && status == ST_OK;
  if (active && !live && tid == 0) a.t_out[inst] = NAN;
  const int n = a.n;
  __syncthreads();  // the previous stage's last reads of these matrices are done
  if (live) s.zero(tid);
  __syncthreads();
  if (live) gload<T, NT>(a.g1n + (size_t)inst * n * n * 2, n, n, s.re[1], s.im[1], LD, tid);
  __syncthreads();
  const M o(tid);
  T gr[TPW], gi[TPW];
  mzero<TPW>(gr, gi);
  if (live) {
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int i = o.row(q), j = o.col(q);
      if (i < n && j < n) {
        gr[q] = s.re[1][i * LD + j];
        gi[q] = s.im[1][i * LD + j];
      }
    }
  }
  const double tot = trace_mm<NMAX>(a.sigma_l, a.sigma_r, inst, n, s, gr, gi, tid, live);
  if (live && tid == 0) a.t_out[inst] = tot;
  return live ? tot : NAN;
}

// ---------------------------------------------------------------------------
// one ticket
// ---------------------------------------------------------------------------

__device__ __forceinline__ SrArgs pl_sr_args(const PlProb& P, bool right, const double* eta) {
  return right ? SrArgs{P.n_lead_r, P.n_inst, P.lead_r, P.set_of, P.omega, eta, P.tol,
                        P.maxiter, P.g_r, P.dos_r, P.iters_r, P.st_r, 0}
               : SrArgs{P.n_lead_l, P.n_inst, P.lead_l, P.set_of, P.omega, eta, P.tol,
                        P.maxiter, P.g_l, P.dos_l, P.iters_l, P.st_l, 0};
}

__device__ __forceinline__ SigmaArgs pl_sigma_args(const PlProb& P, bool right,
                                                   const int32_t* status_in) {
  return right ? SigmaArgs{P.n_lead_r, P.n_c_r, P.n, P.n_inst, P.v_r, P.set_of, P.g_r, P.idx_r,
                           status_in, P.sig_r}
               : SigmaArgs{P.n_lead_l, P.n_c_l, P.n, P.n_inst, P.v_l, P.set_of, P.g_l, P.idx_l,
                           status_in, P.sig_l};
}

// Acceptance of a rung as the reference's adaptive loop sees it (negf.py:446-514):
// status 0, finite T, T >= -nonneg_tol, T <= the channel limit.
__device__ __forceinline__ bool pl_accept(const PlProb& P, int inst, int st, double t) {
  const double tmax = P.lad_tmax ? P.lad_tmax[inst] : INFINITY;
  return st == ST_OK && isfinite(t) && t >= -P.lad_tol && t <= tmax;
}

// The eta of an instance's rung into the per-instance arrays the stages read.
__device__ __forceinline__ void pl_set_rung(const PlProb& P, int inst, int rung) {
  const double e = P.lad_eta[P.lad_first[inst] + rung];
  P.eta_cur[inst] = e;
  P.eta_dev_cur[inst] = P.lad_dev_fixed ? P.eta_dev[inst] : e;
}

__device__ __forceinline__ void pl_record_rung(const PlProb& P, int inst, int rung, int st,
                                               double t) {
  P.lad_t[(size_t)inst * P.lad_max + rung] = t;
  P.lad_st[(size_t)inst * P.lad_max + rung] = st;
}

// The pipeline of one ticket's instances at the given etas; returns T and the status.
template <int NMAX, class L>
__device__ void pl_pipeline(const PlProb& P, const double* eta, const double* eta_dev, int inst,
                            bool active, int phase, L& s, int lt, PlShared& ps, int sub,
                            double& t, int& st) {
  constexpr int G = PlSub<NMAX>::G;
  int st_l = ST_OK, st_r = ST_OK;
  t = NAN;
  st = ST_OK;
  if (phase != PL_DEVICE) {
    st_l = pl_sr<NMAX>(pl_sr_args(P, false, eta), inst, active, s, lt, ps, sub, G);
    st_r = st_l;
    if (!P.share) st_r = pl_sr<NMAX>(pl_sr_args(P, true, eta), inst, active, s, lt, ps, sub, G);
    if (phase == PL_LEADS) return;
  } else if (active) {
    st_l = P.st_l[inst];
    st_r = P.share ? st_l : P.st_r[inst];
  }
  pl_sigma<NMAX>(pl_sigma_args(P, false, nullptr), inst, active && st_l == ST_OK, s, lt);
  pl_sigma<NMAX>(pl_sigma_args(P, true, nullptr), inst, active && st_r == ST_OK, s, lt);
  const int merged = st_l != ST_OK ? st_l : st_r;
  const RgfArgs ra{P.n, P.n_layers, P.n_inst, P.onsite, P.coupling, P.set_of, P.omega, eta_dev,
                   P.sig_l, P.sig_r, nullptr, P.g1n, P.status, 0};
  st = pl_rgf<NMAX>(ra, inst, active, merged, s, lt);
  if (active && lt == 0) P.status[inst] = st;
  const TraceArgs ta{P.n, P.n_inst, P.sig_l, P.sig_r, P.g1n, nullptr, P.t_out};
  t = pl_trace<NMAX>(ta, inst, active, st, s, lt);
}

template <int NMAX>
__device__ void pl_ticket(const PlProb& P, int first, int phase, unsigned char* smem, PlShared& ps,
                          int tid) {
  using S = PlSub<NMAX>;
  using L = typename S::L;
  const int sub = tid / S::NT;
  const int lt = tid - sub * S::NT;
  const bool real = sub < S::G;
  L s(smem + (size_t)(real ? sub : 0) * S::STRIDE);
  if (!real) {  // threads past the last sub-block: barriers and reductions only
    s.sred = ps.idle_red;
    s.sflag = ps.idle_flag;
  }
  const int inst = first + sub;
  const bool active = real && inst < P.n_inst;
  // One pass of the pipeline per rung.  Without a ladder there is one rung at the
  // problem's etas; with one, every instance retries at its next rung until one is
  // accepted, and the ticket runs until none of its sub-blocks retries.
  const bool lad = P.lad_eta != nullptr;
  const double* eta = lad ? P.eta_cur : P.eta_lead;
  const double* eta_dev = lad ? P.eta_dev_cur : P.eta_dev;
  const int count = !active ? 0 : lad ? min(P.lad_count[inst], P.lad_max) : 1;
  bool go = count > 0;
  if (active && !go && lt == 0) P.lad_rung[inst] = 0;
  int rung = 0;
  for (;;) {
    if (lad && go && lt == 0) pl_set_rung(P, inst, rung);
    if (lt == 0 && real) ps.more[sub] = go ? 1 : 0;
    __syncthreads();
    int any = 0;
    for (int q = 0; q < S::G; ++q) any |= ps.more[q];
    any = __builtin_amdgcn_readfirstlane(any);  // the same in every lane
    if (!any) break;
    double t;
    int st;
    pl_pipeline<NMAX>(P, eta, eta_dev, inst, go, phase, s, lt, ps, sub, t, st);
    if (go) {
      bool ok = true;
      if (lad) {
        if (lt == 0) pl_record_rung(P, inst, rung, st, t);
        ok = pl_accept(P, inst, st, t);
        if ((ok || rung + 1 >= count) && lt == 0) P.lad_rung[inst] = ok ? rung : count;
      }
      if (ok || rung + 1 >= count) {
        go = false;
      } else {
        ++rung;
      }
    }
  }
}

// Class 96: one instance per ticket, the class's own per-instance stages on a pool
// slot; statuses travel through the workspace as in the staged path.
__device__ __forceinline__ void pl_pipeline_big(const PlProb& P, const double* eta,
                                                const double* eta_dev, int inst, int phase,
                                                double* slot, BigShared& sh, int tid) {
  if (phase != PL_DEVICE) {
    big_sr_one(pl_sr_args(P, false, eta), inst, slot, sh, tid);
    __syncthreads();
    if (!P.share) {
      big_sr_one(pl_sr_args(P, true, eta), inst, slot, sh, tid);
      __syncthreads();
    }
    if (phase == PL_LEADS) return;
  }
  const int32_t* st_r = P.share ? P.st_l : P.st_r;
  big_sigma_one(pl_sigma_args(P, false, P.st_l), inst, slot, sh, tid);
  __syncthreads();
  big_sigma_one(pl_sigma_args(P, true, st_r), inst, slot, sh, tid);
  __syncthreads();
  // left failure wins, as the reference evaluates the left lead first
  if (tid == 0 && P.st_l[inst] == ST_OK) P.st_l[inst] = st_r[inst];
  __syncthreads();
  const RgfArgs ra{P.n, P.n_layers, P.n_inst, P.onsite, P.coupling, P.set_of, P.omega, eta_dev,
                   P.sig_l, P.sig_r, P.st_l, P.g1n, P.status, 0};
  big_rgf_one(ra, inst, slot, sh, tid);
  __syncthreads();
  const TraceArgs ta{P.n, P.n_inst, P.sig_l, P.sig_r, P.g1n, P.status, P.t_out};
  big_trace_one(ta, inst, slot, sh, tid);
}

__device__ __forceinline__ void pl_ticket_big(const PlProb& P, int inst, int phase, double* slot,
                                              BigShared& sh, int tid) {
  // one pass per rung, as in pl_ticket (one instance per ticket here)
  const bool lad = P.lad_eta != nullptr;
  const double* eta = lad ? P.eta_cur : P.eta_lead;
  const double* eta_dev = lad ? P.eta_dev_cur : P.eta_dev;
  const int count =
      lad ? __builtin_amdgcn_readfirstlane(min(P.lad_count[inst], P.lad_max)) : 1;
  if (count <= 0 && tid == 0) P.lad_rung[inst] = 0;
  for (int rung = 0; rung < count; ++rung) {
    if (lad && tid == 0) pl_set_rung(P, inst, rung);
    __syncthreads();
    pl_pipeline_big(P, eta, eta_dev, inst, phase, slot, sh, tid);
    if (!lad) break;
    __syncthreads();
    const double t = P.t_out[inst];
    const int st = P.status[inst];
    __syncthreads();  // every thread has read them before the next rung writes them
    if (tid == 0) pl_record_rung(P, inst, rung, st, t);
    const bool ok = __builtin_amdgcn_readfirstlane(pl_accept(P, inst, st, t) ? 1 : 0) != 0;
    if (ok || rung + 1 >= count) {
      if (tid == 0) P.lad_rung[inst] = ok ? rung : count;
      break;
    }
  }
}

// Copies the problem table into the workspace and clears the ticket counter.
__global__ void k_pl_setup(PlArgs a, PlArgs* table) {
  if (threadIdx.x == 0) {
    *table = a;
    *a.counter = 0;
  }
}

__global__ void k_pl_reset(int* counter) { *counter = 0; }

__global__ void __launch_bounds__(PL_NT) k_persistent(const PlArgs* __restrict__ t, int phase,
                                                      double* pool) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ PlShared ps;
  const int tid = threadIdx.x;
  double* slot = pool ? pool + (size_t)blockIdx.x * BIG_SLOT : nullptr;
  const int n_prob = t->n_prob, n_tickets = t->n_tickets;
  int* counter = t->counter;
  for (;;) {
    if (tid == 0) ps.ticket = atomicAdd(counter, 1);
    __syncthreads();
    const int tk = __builtin_amdgcn_readfirstlane(ps.ticket);
    __syncthreads();
    if (tk >= n_tickets) break;
    // the ticket's problem; its fields are read from the table where they are used
    int q = 0;
    while (q + 1 < n_prob && tk >= t->prob[q + 1].ticket0) ++q;
    const PlProb& P = t->prob[q];
    const int first = (tk - P.ticket0) * P.group;
    switch (P.cls) {
      case 8:
        pl_ticket<8>(P, first, phase, smem, ps, tid);
        break;
      case 16:
        pl_ticket<16>(P, first, phase, smem, ps, tid);
        break;
      case 24:
        pl_ticket<24>(P, first, phase, smem, ps, tid);
        break;
      case 32:
        pl_ticket<32>(P, first, phase, smem, ps, tid);
        break;
      case 48:
        pl_ticket<48>(P, first, phase, smem, ps, tid);
        break;
      default:
        pl_ticket_big(P, first, phase, slot, *reinterpret_cast<BigShared*>(smem), tid);
        break;
    }
    __syncthreads();
  }
}

// dynamic LDS of k_persistent: the largest class in use
__host__ __device__ constexpr size_t pl_lds_bytes(int cls) {
  return cls == 8    ? PlSub<8>::BYTES
         : cls == 16 ? PlSub<16>::BYTES
         : cls == 24 ? PlSub<24>::BYTES
         : cls == 32 ? PlSub<32>::BYTES
         : cls == 48 ? PlSub<48>::BYTES
                     : sizeof(BigShared);
}

}  // namespace negf
