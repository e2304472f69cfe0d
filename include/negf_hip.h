/*
 * C ABI of libnegf_hip.so: batched phonon-NEGF transport on MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Every array argument is DEVICE memory
 * (hipMalloc'd or a PyTorch-ROCm tensor); complex arrays are complex128
 * interleaved (re, im), row-major.  Every call is stream-ordered on `stream`
 * (a hipStream_t, NULL = default stream) and returns before the work finishes.
 * Return value: NEGF_OK or an NEGF_ERR_* code; negf_last_error() (thread
 * local) describes the last failure.
 *
 * One "instance" is one (omega, k_par, eta) point.  Blocks that depend only on
 * k_par are stored once per "set" and instances point to their set through
 * set_of[].  Per-instance outcomes are reported in int32 status arrays:
 *   0 ok, 1 decimation did not converge within maxiter, 2 non-finite values,
 *   3 exactly singular pivot in a lead inverse,
 *   4 eigenmode route at eta = 0 where the decimation fails (propagating modes on the
 *     unit circle): the reference's choice among them is not reproduced,
 *   7 (transient, mixed precision only) marked for FP64 recomputation,
 *   12 / 13 non-finite / singular in the device sweep.
 *
 * The entry points replace the Python-level hot path of the reference package
 * negfpy (paths relative to its root); the ctypes binding is
 * negfpy_amd/_lib.py.
 */
#ifndef NEGF_HIP_H
#define NEGF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEGF_ABI_VERSION 1

/* Mixed precision: certificate of a Newton-refined FP32 decimation (relative Frobenius
 * norms).  Accepted when the FP64 residual |I - (z - d00 - d01 g d10) g| / (|M| |g|) is
 * below RTOL, the last Newton correction below XTOL and the drift from the FP32
 * decimation's g below DTOL; otherwise the instance is decimated again in FP64. */
#define NEGF_NEWTON_RTOL 1e-13
#define NEGF_NEWTON_XTOL 1e-8
#define NEGF_NEWTON_DTOL 1e-4
/* Mixed precision: iteration limit of the FP32 decimation (at most maxiter).  An instance
 * that has not converged after this many doublings (a chain of 2^16 layers) has a mode
 * with |lambda| above about 1 - 4e-4 (tol 1e-12), which FP32 cannot resolve to the certificate: it goes
 * to the FP64 decimation (with the caller's maxiter) without further FP32 work. */
#define NEGF_MIXED_FP32_MAXITER 16
/* Mixed precision: the FP32 decimation also stops (status 1, FP64 redo) from doubling
 * STALL_FROM on when one doubling shrinks max(|alpha|, |beta|) by less than STALL_RATIO.
 * The ratio of doubling k is about |lambda|^(2^(k-1)) for the slowest mode, so a ratio
 * above 0.9 at doubling 7 means 1 - |lambda| < 1.6e-3: the decimation would need at
 * least 15 of its NEGF_MIXED_FP32_MAXITER doublings to reach tol 1e-12, and such
 * instances (in-band modes at small eta) are the ones the FP64 certificate rejects. */
#define NEGF_MIXED_FP32_STALL_FROM 7
#define NEGF_MIXED_FP32_STALL_RATIO 0.9

enum { NEGF_FP64 = 0, NEGF_MIXED = 1 };
enum { NEGF_SURFACE_DECIMATION = 0, NEGF_SURFACE_EIGEN = 1 };
/* negf_transmission_problem.share_leads: independent leads; the right lead is the left
 * one (one decimation serves both); the right lead is the left one mirrored, lead_r =
 * (d00_L, d10_L, d01_L) for every set, as the interface contacts of a bulk force-constant
 * model are -- in negf_transmission_batched one decimation yields both surfaces where
 * the block fits (<= 48 DOF), two decimations otherwise and in
 * negf_transmission_multi_batched.  The caller guarantees the mirror relation; lead_r
 * must still be passed. */
enum { NEGF_LEADS_INDEPENDENT = 0, NEGF_LEADS_SHARED = 1, NEGF_LEADS_MIRRORED = 2 };
enum { NEGF_OK = 0, NEGF_ERR_ARG = 1, NEGF_ERR_HIP = 2 };

/* ABI version (NEGF_ABI_VERSION). */
int negf_version(void);

/* Description of the last error on this thread. */
const char* negf_last_error(void);

/* Largest supported principal-layer block size of the transmission, surface-GF and
 * self-energy entries for a precision (FP64: 96; mixed: 48). */
int negf_max_block(int precision);

/* Largest block size of negf_device_green_batched and negf_adjoint_batched (96). */
int negf_max_block_green(void);

/*
 * Block-size class (8, 16, 24, 32, 48 and, in FP64, 96; 64 in mixed precision:
 * workgroup size, rows per lane, tile size and LDS footprint of every kernel) used
 * for blocks of n DOF in the given precision.  Class 96 keeps the matrices of an
 * instance in a device-memory pool the library allocates on first use on a device
 * (about 1.3 MB per resident workgroup) and keeps until the process ends; eager calls
 * on different streams are ordered so they never share it at once.  Under stream
 * capture the pool must already exist (one eager call on the device first, else
 * NEGF_ERR_ARG); a captured call is ordered only by its own stream, so a graph that
 * holds class-96 work must not be replayed while class-96 work runs on another
 * stream.  Any class >= n is valid; nmax = 0 restores the default (the
 * smallest class >= n).  Set by the autotuner (negfpy_amd.core.autotune); the
 * table is process-wide and thread-safe.  No counterpart in the reference, whose
 * dense/sparse solves have no launch configuration.
 */
int negf_set_block_class(int precision, int n, int nmax);
int negf_block_class(int precision, int n);

/*
 * Kernel family for blocks of up to 32 DOF.  The one-wavefront family solves each
 * instance on one 64-lane wavefront with its products on the matrix cores
 * (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) and a register-resident wave
 * Gauss-Jordan; the other family gives an instance a workgroup (the block-size
 * classes above).  mode 1 (default): the wave family only for the sizes a per-size
 * choice (negf_set_kernel_family) asks for -- the workgroup classes measured faster
 * at every size on MI355X; mode 2: the wave family for every block up to 32 DOF;
 * mode 0: never.  Both give the same results to rounding.  Returns the previous mode
 * (-1 and NEGF_ERR_ARG text for a bad mode); process-wide.  No counterpart in the
 * reference (a launch-configuration switch, like negf_set_block_class).
 */
int negf_set_wave_kernels(int mode);

/*
 * Per-size kernel family (blocks of 1..32 DOF), for the autotuner: family
 * NEGF_FAMILY_WORKGROUP or NEGF_FAMILY_WAVE for the decimations and sweeps of that
 * precision's arithmetic (NEGF_FP64: the FP64 decimations and sweeps of both
 * precisions; NEGF_MIXED: the FP32 decimations), 0 for the default rule of mode 1.
 * Applies in mode 1 only.  negf_kernel_family returns the family a decimation of n-DOF
 * blocks uses now.  Process-wide, thread-safe.  No counterpart in the reference.
 */
enum { NEGF_FAMILY_WORKGROUP = 1, NEGF_FAMILY_WAVE = 2 };
int negf_set_kernel_family(int precision, int n, int family);
int negf_kernel_family(int precision, int n);

/*
 * Block products of the FP64 workgroup decimation and sweep (classes 8..48): mode 1
 * (default) runs them on the matrix cores (v_mfma_f64_4x4x4f64, complex products as
 * four real chains), mode 0 as register-tiled FMA loops on the vector ALUs.  Both give
 * the same results to rounding.  Returns the previous mode (-1 and NEGF_ERR_ARG text
 * for a bad mode); process-wide.  No counterpart in the reference (a
 * launch-configuration switch, like negf_set_wave_kernels).
 */
int negf_set_matrix_core_products(int mode);

/*
 * Per-size product form of the FP64 workgroup decimations and sweeps (blocks of
 * 1..48 DOF), for the autotuner: NEGF_PRODUCTS_MATRIX_CORE (v_mfma_f64_4x4x4f64 tiles,
 * four per 8 x 8 tile-group of a wave) or NEGF_PRODUCTS_VECTOR (a TM x TM register tile
 * of FMA loops per thread), 0 for the default (matrix cores).  Applies while
 * negf_set_matrix_core_products is 1.  negf_product_form returns the form n-DOF blocks
 * use now (0 outside 1..48 DOF).
 * Process-wide, thread-safe.  No counterpart in the reference.
 */
enum { NEGF_PRODUCTS_MATRIX_CORE = 1, NEGF_PRODUCTS_VECTOR = 2 };
int negf_set_product_form(int n, int form);
int negf_product_form(int n);

/*
 * Surface Green's function of semi-infinite leads by Lopez-Sancho decimation.
 * Replaces surface_gf_sancho_rubio, src/negfpy/core/surface_gf.py:18-81
 * (and the per-k loop of lead_surface_dos*, src/negfpy/core/dos.py:31-251).
 *   lead    [n_sets][3][n][n]  d00, d01, d10 of each set (d10 already resolved)
 *   set_of  [n_inst]           set index of each instance
 *   omega, eta [n_inst]        z = (omega + i eta)^2
 *   tol, maxiter               stop when max(|alpha|, |beta|) < tol
 *   refine_steps               NEGF_MIXED only: Newton steps of the FP64 completion after
 *                              the FP32 decimation (negf_surface_refine_batched, then
 *                              negf_sancho_rubio_redo for what it does not certify); < 0
 *                              leaves the FP32 decimation's eps_s in g_out (complex128) for
 *                              separate calls.  The FP32 pass stops after
 *                              min(maxiter, NEGF_MIXED_FP32_MAXITER) iterations
 *   g_out   [n_inst][n][n]     g_s = (z - eps_s)^-1
 *   dos_out [n_inst] or NULL   (omega/pi) Re Tr[i (g - g^H)]
 *   iters, status [n_inst]
 */
int negf_sancho_rubio_batched(void* stream, int precision, int n, int n_inst, const double* lead,
                              const int32_t* set_of, const double* omega, const double* eta,
                              double tol, int maxiter, int refine_steps, double* g_out,
                              double* dos_out, int32_t* iters, int32_t* status);

/*
 * FP64 completion of a mixed-precision decimation (blocks up to 32 DOF; 33..48 are all
 * marked for the FP64 redo): g = (z - eps_s)^-1 from the eps_s in g_io, then `steps`
 * Newton steps on F(g) = I - (z - d00 - d01 g d10) g = 0 (the fixed point of the
 * decimation, surface_gf.py:41-81): an FP64 residual, the correction X - P X Q = g R
 * with P = g d01, Q = d10 g solved in FP32 by doubling, g += X.  A final FP64 residual
 * and the NEGF_NEWTON_* tolerances certify the instance: certified ones get g and the
 * surface DOS written back, the others status ST_REDO (7) for negf_sancho_rubio_redo.
 * Instances with status != 0 on entry are skipped.  No counterpart in the reference
 * (which decimates in FP64 only).
 */
int negf_surface_refine_batched(void* stream, int n, int n_inst, const double* lead,
                                const int32_t* set_of, const double* omega, const double* eta,
                                int steps, double* g_io, double* dos_out, int32_t* status);

/*
 * FP64 decimation of exactly the instances whose status is nonzero (FP32
 * failures and refinements the Newton certificate rejected); others are
 * left untouched.  negf_sancho_rubio_batched(NEGF_MIXED) ends with this call.
 */
int negf_sancho_rubio_redo(void* stream, int n, int n_inst, const double* lead,
                           const int32_t* set_of, const double* omega, const double* eta,
                           double tol, int maxiter, double* g_out, double* dos_out, int32_t* iters,
                           int32_t* status);

/*
 * Contact self-energy Sigma = V g V^H scattered into an n_dev x n_dev end-layer
 * block at positions idx[0..n_c).  Replaces _self_energy_contact and the
 * contact embedding, src/negfpy/core/negf.py:72-97 and :245-248.
 *   v [n_sets][n_c][n_lead], g [n_inst][n_lead][n_lead], sigma_out [n_inst][n_dev][n_dev]
 *   status_in [n_inst] or NULL: instances with a nonzero status are skipped.
 */
int negf_self_energy_batched(void* stream, int n_lead, int n_c, int n_dev, int n_inst,
                             const double* v, const int32_t* set_of, const double* g,
                             const int32_t* idx, const int32_t* status_in, double* sigma_out);

/*
 * Block recursive Green's function sweep for G_{1N} of
 * A = (omega + i eta_dev)^2 - D - Sigma_L - Sigma_R (D block tridiagonal, lower
 * blocks = coupling^H).  Replaces the sparse assembly and SuperLU solve for the
 * right-contact columns, src/negfpy/core/negf.py:53-69, :229-250, :350-354.
 *   onsite [n_sets][n_layers][n][n], coupling [n_sets][n_layers-1][n][n]
 *   sigma_l/sigma_r [n_inst][n][n] (from negf_self_energy_batched)
 *   g1n_out [n_inst][n][n]; status [n_inst] (12/13 on device failure, or status_in)
 * NEGF_MIXED sweeps in FP32 and recomputes FP32 failures (12/13) in FP64.
 */
int negf_rgf_g1n_batched(void* stream, int precision, int n, int n_layers, int n_inst,
                         const double* onsite, const double* coupling, const int32_t* set_of,
                         const double* omega, const double* eta_dev, const double* sigma_l,
                         const double* sigma_r, const int32_t* status_in, double* g1n_out,
                         int32_t* status);

/*
 * Full device Green's function G = A^-1 of A = (omega + i eta_dev)^2 - D - Sigma_L - Sigma_R,
 * FP64, every block: a left-connected sweep followed by a backward pass over all
 * L^2 blocks.  Replaces device_green_function's dense solve,
 * src/negfpy/core/negf.py:253-303.
 *   g_out [n_inst][L n][L n]; status [n_inst] (12/13 on device failure)
 */
int negf_device_green_batched(void* stream, int n, int n_layers, int n_inst, const double* onsite,
                              const double* coupling, const int32_t* set_of, const double* omega,
                              const double* eta_dev, const double* sigma_l, const double* sigma_r,
                              double* g_out, int32_t* status);

/*
 * Adjoint of the transmission (FP64): T and, for every instance, the gradient of
 * w * T with respect to every onsite block D_ii and coupling block C_i of the
 * device (C_i enters D above the diagonal, C_i^H below), as dT/dRe + i dT/dIm:
 *   grad_onsite_i   = 2 w (G_iN W G_1i)^H,
 *   grad_coupling_i = 2 w ((G_{i+1,N} W G_1i)^H + G_iN W G_{1,i+1}),
 * W = Gamma_R G_1N^H Gamma_L, from a right-connected sweep (row 1 of G) and a
 * left-connected sweep (column N).  The reference has no gradients; this is the
 * reverse mode of its transmission, src/negfpy/core/negf.py:306-357, with respect
 * to the device blocks it assembles at :53-69.
 *   weight [n_inst] (dL/dT, null: 1); workspace >= negf_adjoint_workspace_bytes (same
 *   checkpoint);
 *   t_out [n_inst]; grad_onsite [n_inst][L][n][n]; grad_coupling [n_inst][L-1][n][n];
 *   status [n_inst] (lead status passed through, 12/13 on device failure);
 *   g1n_out [n_inst][n][n] or NULL: G_1N, for negf_lead_adjoint_batched;
 *   checkpoint: 0 stores every block of both sweeps ((2L - 1) n^2 complex per instance);
 *     K > 0 keeps only the left-connected sweep's carried state every K layers and
 *     recomputes each segment's blocks in the gradient pass (one more left sweep of
 *     GJ inversions and products; (L + ceil(L/K) + K + 1) n^2 per instance, least near
 *     K = sqrt(L)), same results;
 *   per_set: 0 gives per-instance gradients; 1 sums them into grad_onsite
 *     [n_sets][L][n][n] / grad_coupling [n_sets][L-1][n][n] with atomic adds (the caller
 *     zeroes them; the order of the sums varies, so results vary at rounding level);
 *   grad_ends_out [n_inst][2][n][n] or NULL: the onsite gradients of layers 1 and N of
 *     every instance (also with per_set), for negf_lead_adjoint_batched.
 */
size_t negf_adjoint_workspace_bytes(int n, int n_layers, int n_inst, int checkpoint);
int negf_adjoint_batched(void* stream, int n, int n_layers, int n_inst, const double* onsite,
                         const double* coupling, const int32_t* set_of, const double* omega,
                         const double* eta_dev, const double* sigma_l, const double* sigma_r,
                         const int32_t* status_in, const double* weight, void* workspace,
                         size_t workspace_bytes, double* t_out, double* grad_onsite,
                         double* grad_coupling, int32_t* status, double* g1n_out,
                         int checkpoint, int per_set, double* grad_ends_out);

/*
 * Lead side of the adjoint: weight * dT/d(lead blocks) and weight * dT/dV for one lead
 * (side 0: left, contact layer 1; 1: right, layer N), from the device adjoint's outputs.
 * No counterpart in the reference (no gradients); differentiates Sigma = V g V^H
 * (src/negfpy/core/negf.py:72-97) and the surface GF g^-1 = z - d00 - d01 g d10
 * (src/negfpy/core/surface_gf.py:18-81) by the implicit-function theorem: the adjoint
 * Stein equation Z - (g d01)^H Z (d10 g)^H = dT/dg is solved by doubling (at most
 * maxiter doublings) on the matrix cores.  Blocks of up to 48 DOF (max(n, n_lead)).
 *   lead  [n_sets][3][n_lead][n_lead]  d00, d01, d10 of this lead (as decimated)
 *   v     [n_sets][n_c][n_lead], idx [n_c]   its coupling and contact positions
 *   g     [n_inst][n_lead][n_lead]      its surface GF (the forward's)
 *   g1n   [n_inst][n][n]                G_1N (negf_adjoint_batched's g1n_out)
 *   sigma_l, sigma_r [n_inst][n][n]     the embedded self-energies of the forward
 *   grad_ends [n_inst][2][n][n]         negf_adjoint_batched's grad_ends_out (same weight)
 *   status_in [n_inst]                  the adjoint's status; != 0 gives zero gradients
 *   grad_lead [n_inst][3][n_lead][n_lead], grad_v [n_inst][n_c][n_lead]   written
 *   status [n_inst]                     0, status_in, 1 (not converged), 2 (non-finite)
 */
int negf_lead_adjoint_batched(void* stream, int side, int n, int n_inst, int n_lead, int n_c,
                              const double* lead, const double* v, const int32_t* idx,
                              const int32_t* set_of, const double* g, const double* g1n,
                              const double* sigma_l, const double* sigma_r,
                              const double* grad_ends, const double* weight,
                              const int32_t* status_in, int maxiter, double* grad_lead,
                              double* grad_v, int32_t* status);

/*
 * T = Re Tr[Gamma_L G_1N Gamma_R G_1N^H], Gamma = i (Sigma - Sigma^H), in FP64.
 * Replaces _broadening and the trace, src/negfpy/core/negf.py:115-116, :346-357.
 * Instances with status_in != 0 get NaN.
 */
int negf_transmission_trace_batched(void* stream, int n, int n_inst, const double* sigma_l,
                                    const double* sigma_r, const double* g1n,
                                    const int32_t* status_in, double* t_out);

/* Copy st_r into st_l wherever st_l is 0 (left-lead failure takes precedence). */
int negf_merge_status(void* stream, int n_inst, int32_t* st_l, const int32_t* st_r);

/*
 * The stages of negf_transmission_batched after the surface GFs, in one launch where
 * the block class allows it: Sigma_L and Sigma_R (as negf_self_energy_batched, into
 * sigma_l / sigma_r), the FP64 sweep for G_1N (as negf_rgf_g1n_batched, into g1n_out)
 * and T (as negf_transmission_trace_batched), with every block product on the matrix
 * cores.  Replaces the self-energies, the solve and the trace of
 * src/negfpy/core/negf.py:306-357 (with :115-116) in one kernel.  Fused for FP64
 * workgroup classes 8..48 with matrix-core products (negf_set_matrix_core_products 1):
 * the launch takes the class of the largest of the device block and the two lead
 * blocks.  Otherwise (a block above 48 DOF, vector-ALU products, the one-wavefront
 * family) the four kernels run in turn, and status_l then holds the merged lead
 * status.
 *   g_l, g_r [n_inst][n_lead][n_lead]  surface GFs; v_*, idx_*: as negf_self_energy_batched
 *   status_l, status_r [n_inst]        lead statuses, merged left first into status
 *   sigma_l, sigma_r, g1n_out [n_inst][n][n]   written
 *   status, t_out [n_inst]             0 / lead status / 12, 13; NaN T where status != 0
 */
int negf_transmission_sweep_batched(
    void* stream, int n, int n_layers, int n_inst, const double* onsite, const double* coupling,
    const int32_t* set_of, const double* omega, const double* eta_dev, int n_lead_l, int n_c_l,
    const double* v_l, const double* g_l, const int32_t* idx_l, int32_t* status_l, int n_lead_r,
    int n_c_r, const double* v_r, const double* g_r, const int32_t* idx_r,
    const int32_t* status_r, double* sigma_l, double* sigma_r, double* g1n_out, int32_t* status,
    double* t_out);

/*
 * An eta ladder per instance (negf_transmission_multi_batched only, decimation
 * route): each instance tries its rungs in order, eta_lead = the rung's eta, inside
 * the persistent launch, until one is accepted (status 0, finite T,
 * T >= -nonneg_tol, T <= t_max) or none is left.  Replaces the per-k eta loop of
 * transmission_kavg_adaptive, src/negfpy/core/negf.py:446-514, which retries one
 * point at a time.  The caller keeps every index inside etas.
 */
typedef struct negf_eta_ladder {
  const double* etas;      /* rungs of all instances, concatenated */
  const int32_t* first;    /* [n_inst] index in etas of the instance's first rung */
  const int32_t* count;    /* [n_inst] rungs of the instance */
  const double* t_max;     /* [n_inst] channel limit, or NULL for none */
  int32_t* rung_out;       /* [n_inst] accepted rung, or count when none was accepted */
  double* t_rungs;         /* [n_inst][max_rungs] T of every attempted rung */
  int32_t* status_rungs;   /* [n_inst][max_rungs] status of every attempted rung */
  double nonneg_tol;       /* acceptance: T >= -nonneg_tol */
  int max_rungs;           /* row length of t_rungs / status_rungs (>= every count) */
  int eta_dev_fixed;       /* 1: the problem's eta_dev[] at every rung; 0: eta_dev = rung eta */
} negf_eta_ladder;

/* One batched transmission problem (see negf_transmission_batched). */
typedef struct negf_transmission_problem {
  int precision;    /* NEGF_FP64 or NEGF_MIXED */
  int n;            /* device principal-layer block size */
  int n_layers;     /* device layers */
  int n_inst;       /* instances */
  int n_lead_l;     /* left lead block size */
  int n_lead_r;     /* right lead block size */
  int n_c_l;        /* left contact size (rows of v_l) */
  int n_c_r;        /* right contact size (rows of v_r) */
  int share_leads;  /* NEGF_LEADS_INDEPENDENT, _SHARED or _MIRRORED */
  int maxiter;      /* decimation iteration limit */
  int refine_steps; /* NEGF_MIXED: Newton steps per surface GF (>= 0) */
  double tol;       /* decimation tolerance */
  const double* lead_l;   /* [n_sets][3][n_lead_l][n_lead_l] */
  const double* lead_r;   /* [n_sets][3][n_lead_r][n_lead_r] */
  const double* v_l;      /* [n_sets][n_c_l][n_lead_l] device-to-lead coupling */
  const double* v_r;      /* [n_sets][n_c_r][n_lead_r] */
  const int32_t* idx_l;   /* [n_c_l] contact positions inside the first layer */
  const int32_t* idx_r;   /* [n_c_r] contact positions inside the last layer */
  const double* onsite;   /* [n_sets][n_layers][n][n] */
  const double* coupling; /* [n_sets][n_layers-1][n][n] */
  const int32_t* set_of;  /* [n_inst] */
  const double* omega;    /* [n_inst] */
  const double* eta_lead; /* [n_inst] */
  const double* eta_dev;  /* [n_inst] */
  double* t_out;          /* [n_inst] transmission */
  double* dos_l;          /* [n_inst] left-lead surface DOS, or NULL */
  double* dos_r;          /* [n_inst] right-lead surface DOS, or NULL (unused when shared) */
  int32_t* iters_l;       /* [n_inst] decimation steps, left lead */
  int32_t* iters_r;       /* [n_inst] decimation steps, right lead (unused when shared) */
  int32_t* status;        /* [n_inst] */
  int surface_route;      /* NEGF_SURFACE_DECIMATION, or NEGF_SURFACE_EIGEN: lead instances whose
                             decimation fails take the eigenmode route */
  const negf_eta_ladder* ladder; /* NULL, or an eta ladder per instance (replaces eta_lead;
                                    negf_transmission_multi_batched only) */
} negf_transmission_problem;

/*
 * Eigenmode surface GF for every instance whose status is nonzero (after
 * negf_sancho_rubio_batched): the transfer matrix of the n finite eigenmodes of
 * smallest |lambda| of the lead's quadratic eigenproblem, as the reference's
 * generalized_eigen[_svd] routes select them (src/negfpy/core/surface_gf.py:84-199),
 * from a spectral projector (matrix sign function, no eigensolver).  On success g,
 * the DOS and status (0) are written; instances it cannot do keep their status.
 * Blocks up to 96 DOF (pencils of 2n up to 96 directly, larger ones as 2 x 2 blocks
 * with a second device-memory pool of about 3.5 MB per resident workgroup).
 */
int negf_surface_gf_eigen_batched(void* stream, int n, int n_inst, const double* lead,
                                  const int32_t* set_of, const double* omega, const double* eta,
                                  double* g_io, double* dos_out, int32_t* status);

/* Device workspace needed by negf_transmission_batched. */
size_t negf_transmission_workspace_bytes(const negf_transmission_problem* p);

/*
 * Full (omega, k_par) transmission solve for a batch: both surface GFs, both
 * self-energies, the device sweep and the trace.  Replaces transmission(),
 * src/negfpy/core/negf.py:306-357, and the k loops of transmission_kavg*,
 * :360-688, which call it once per point.
 * NEGF_MIXED: FP32 decimation, FP64 Newton completion with FP32 Stein corrections and
 * a certificate, FP64 decimation of the uncertified instances; the device sweep,
 * self-energies, Gamma and the trace in FP64 (an FP32 sweep leaves about a quarter
 * of a disordered 32-layer ribbon's frequencies above 1e-6 relative error in T at
 * eta = 1e-6, and correcting it by FP64-residual refinement costs more than the FP64
 * sweep).  refine_steps must be >= 0 in NEGF_MIXED.
 */
int negf_transmission_batched(void* stream, const negf_transmission_problem* p, void* workspace,
                              size_t workspace_bytes);

/* Scratch of negf_transmission_batched_noalloc per GPU: 64 MiB. */
#define NEGF_NOALLOC_SCRATCH_BYTES (64 << 20)

/*
 * negf_transmission_batched without a workspace argument and without allocating:
 * no hipMalloc, hipFree or allocating library call, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  No counterpart in the reference.  The workspace is a file-scope device array
 * of NEGF_NOALLOC_SCRATCH_BYTES on each GPU, held from the library's load on that
 * device until the process ends; its address is looked up (hipGetSymbolAddress) on
 * the first call on a device.  Limits: negf_transmission_workspace_bytes(p) must not
 * exceed the scratch (24-DOF blocks: up to 1,800 instances per call; split larger
 * batches); blocks up to 48 DOF; decimation route only.  Errors are NEGF_ERR_ARG with
 * a message.  Results equal negf_transmission_batched's.  Every call on a GPU uses
 * that GPU's one array: calls that overlap on one GPU (other streams, other threads)
 * race and must be ordered by the caller.
 */
int negf_transmission_batched_noalloc(void* stream, const negf_transmission_problem* p);

/* Problems per negf_transmission_multi_batched call. */
#define NEGF_MULTI_MAX_PROBLEMS 8

/*
 * Several transmission problems of different block sizes (up to 96 DOF) solved
 * together: every instance of every problem goes through one persistent
 * kernel launch (two, split around the eigenmode fallback, when a problem asks for
 * NEGF_SURFACE_EIGEN) whose workgroups take tickets of same-problem instances from
 * an atomic work queue, most expensive problems first; per-instance convergence
 * flags let the instances of a ticket stop decimating at their own step.  Outputs
 * per problem as from negf_transmission_batched on that problem (the same
 * per-stage arithmetic; each problem's blocks use the smallest class that holds
 * max(n, n_lead_l, n_lead_r), whatever negf_set_block_class says).  Replaces the
 * per-point loops of transmission_kavg* / heat_current_* over k-points whose
 * layouts differ (src/negfpy/core/negf.py:360-688), which the reference runs one
 * point at a time.  Problems with n_inst = 0 are skipped.  When any problem of the
 * call is NEGF_MIXED (leads up to 48 DOF, refine_steps >= 0, no ladder), every
 * problem's surface GFs are computed first by the staged decimation of its own
 * precision (for mixed problems: FP32 decimation, FP64 Newton certificate, FP64
 * redo; one decimation per mirrored pair), and one persistent launch then runs the
 * self-energies, sweeps and traces of all problems in FP64.  Otherwise (FP64 only)
 * refine_steps is unused and the decimations run inside the persistent launch.
 * Capturable in a hipGraph when no problem needs blocks above 48 DOF (the class-96
 * pool orders calls through an event).
 * A problem with an eta ladder runs its rungs inside the launch (outputs: those of
 * the last attempted rung, and the ladder's arrays); ladders and the eigenmode route
 * cannot be mixed in one call.
 *   workspace >= negf_transmission_multi_workspace_bytes(n_problems, problems)
 */
size_t negf_transmission_multi_workspace_bytes(int n_problems,
                                               const negf_transmission_problem* problems);
int negf_transmission_multi_batched(void* stream, int n_problems,
                                    const negf_transmission_problem* problems, void* workspace,
                                    size_t workspace_bytes);

/*
 * Thermal conductance G(T_i) = (1/2pi) int k_B T(w) x^2 e^-x / (1 - e^-x)^2 dw,
 * x = hbar w / (k_B T_i), trapezoid rule over the given grid (w in rad/s);
 * 0 for T_i <= 0.  Replaces _thermal_conductance_from_spectrum,
 * src/negfpy/workflows/ifc_bulk.py:408-438.
 */
int negf_thermal_conductance(void* stream, int n_w, const double* omega_rad_s, const double* trans,
                             int n_t, const double* temps, double* out);

/*
 * Heat-current cumulants of coherent phonon transport for n_pairs temperature
 * pairs (FP64): out[p] = (C1 [J/s], C2 [J^2/s]) with
 *   C1 = (1/2pi) int <T> (A_L - A_R) dw,
 *   C2 = (1/2pi) int <T>(B_L + B_R) + (<T> + <T^2>)(A_L - A_R)^2 dw,
 * A = k_B T x/expm1(x), B = (k_B T)^2 x^2 e^x/expm1(x)^2, x = hbar w/(k_B T),
 * trapezoid rule on the given grid (w in rad/s).  t2_mean = null gives the
 * single-spectrum form T(B_L + B_R) + T(1 + T)(A_L - A_R)^2.  A pair with a
 * temperature <= 0 gets NaN (the Python layer raises first).  Replaces
 * heat_current_cumulants_from_spectrum / _from_k_moments,
 * src/negfpy/core/fcs.py:89-156.
 */
int negf_fcs_cumulants(void* stream, int n_w, const double* omega_rad_s, const double* t_mean,
                       const double* t2_mean, int n_pairs, const double* temps_left,
                       const double* temps_right, double* out);

/*
 * k-space blocks of a force-constant model for n_k k points at once (FP64):
 * Phi_ab(k) = sum_t Phi_ab,t ph_ab[k][t] (+ pinning on the diagonal of Phi_00),
 * D = M_l^-1/2 Phi M_r^-1/2.  Replaces the per-k host callbacks of
 * src/negfpy/models/material_kspace.py:99-193 (_sum_k_terms,
 * _build_dynamical_blocks, the lead and device builders).
 *   phi00/phi01/phi10 [t][n][n] complex128 terms (t10 = 0: Phi_10 = Phi_01^H);
 *   ph00/ph01/ph10 [n_k][t] complex phases e^{i (ky dy + kz dz)};
 *   s_lead [n] = 1/sqrt(m) of the lead cell (null: no lead blocks);
 *   s_dev [L][n] per device layer (null: no device blocks);
 *   lead_out [n_k][3][n][n] (d00, d01, d10); onsite_out [n_k][L][n][n];
 *   coupling_out [n_k][L-1][n][n].
 * Products and sums are rounded separately in the reference's order, so the
 * blocks equal the reference builder's for the same phases and masses.
 */
int negf_kspace_blocks(void* stream, int n, int n_layers, int n_k, int t00, int t01, int t10,
                       const double* phi00, const double* phi01, const double* phi10,
                       const double* ph00, const double* ph01, const double* ph10,
                       double pinning, const double* s_lead, const double* s_dev,
                       double* lead_out, double* onsite_out, double* coupling_out);

/*
 * d G(T_t) / d T(w_i) for the conductance above: weights_out [n_t][n_w] with
 * G(T_t) = sum_i weights_out[t][i] T(w_i).  Chain rule of
 * src/negfpy/workflows/ifc_bulk.py:408-438 for gradients of kappa.
 */
int negf_conductance_weights(void* stream, int n_w, const double* omega_rad_s, int n_t,
                             const double* temps, double* weights_out);

/*
 * Landauer heat current J = (1/2pi) int [k_B T_L phi(x_L) - k_B T_R phi(x_R)] T(w) dw,
 * phi(x) = x / (e^x - 1).  Replaces heat_current_from_spectrum,
 * src/negfpy/core/thermal_current.py:30-67.  out: one double.
 */
int negf_heat_current(void* stream, int n_w, const double* omega_rad_s, const double* trans,
                      double temp_left, double temp_right, double* out);

#ifdef __cplusplus
}
#endif

#endif /* NEGF_HIP_H */
