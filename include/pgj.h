/*
 * libpgj: MI355X (gfx950) kernels for the partial Green-Jacobi homogenization
 * hot path, behind a plain C ABI.
 *
 * Conventions shared by every entry point:
 *  - all array arguments are device pointers (HBM) unless marked "host";
 *  - grids are (nz, ny, nx); node = (z*ny + y)*nx + x; displacement-like
 *    vectors are AoS, (L, 3N) load-major, node-major, 3 components per node;
 *  - `stream` is a hipStream_t passed as void*; every call only enqueues work
 *    on that stream: no allocation, no synchronization, no memset, so every
 *    call is safe inside hip stream capture (hipGraph);
 *  - buffers named `ws`, `plan`, `state` are caller-owned device memory sized
 *    by the matching *_bytes() query; `state` arrays are documented below;
 *  - the return value is 0 on success, a PGJ_ERR_* code otherwise, and
 *    pgj_last_error() then returns a thread-local message.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference checkout).
 */
#ifndef PGJ_H
#define PGJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PGJ_OK = 0, PGJ_ERR_ARG = 1, PGJ_ERR_HIP = 2, PGJ_ERR_UNSUPPORTED = 3 };

/* Geometry and reference medium of a periodic grid (host struct). The
 * element material is a(x) * C0(lam, mu); the Green reference medium is C0. */
typedef struct pgj_grid {
  int32_t nz, ny, nx;
  int32_t reserved;
  double hx, hy, hz;
  double lam, mu;
} pgj_grid;

int pgj_version(void);
const char* pgj_last_error(void);

/* ---------------------------------------------------------------- operator */

/* out_l = K(a) u_l for each of L loads; loads with active[l] == 0 get zeros.
 * active may be NULL (all active). Replaces the scalar-a kernel `apply`
 * (exp/asg_sctop_reproduction_20261105/scalar_fi.cu:46-53,
 *  exp/asg_gpu_resident_20261103/scalar_fi.cu:46-53) and, for C = a*C0, the
 * general `apply` (solver/src/materials/hybrid_kernels.cu:117-127). */
int pgj_hex8_apply(const pgj_grid* g, int L, const double* a, const double* u, const uint8_t* active,
                   double* out, void* stream);

/* Same operator by an independent thread-per-node kernel; used as the
 * certification operator (the reference certifies with its general kernel,
 * exp/asg_gpu_resident_20261103/resident.py:100-102). */
int pgj_hex8_apply_reference(const pgj_grid* g, int L, const double* a, const double* u, double* out,
                             void* stream);
/* FP32 drop-in apply: the row-march on FP32 a, u and out (FP32 element
 * arithmetic), for the inner operator of iterative refinement
 * (solver/src/runtime/fi/source/femlib/batched/ir.py:33-35, op_inner in FP32).
 * Needs a grid the row-march supports. */
int pgj_hex8_apply32(const pgj_grid* g, int L, const float* a, const float* u, const uint8_t* active, float* out,
                     void* stream);
/* The same FP32 operator as a dense 24x24 Ke contraction on the matrix cores
 * (v_mfma_f32_16x16x4_f32 over 16 elements x 24 DOFs per tile): the element
 * product of `apply_scalar` (exp/asg_sctop_reproduction_20261105/scalar_fi.cu:14-36)
 * with the dense Ke of exp/asg_candidates_20261103/candidates.py:30-54 in place
 * of the matrix-free modal form. soa != 0: vectors are (L, 3, N) component-major
 * (the resident solver's layout), else (L, N, 3). Needs nx = 64 or 128 and ny
 * divisible by 4. */
int pgj_hex8_apply32_mfma(const pgj_grid* g, int L, const float* a, const float* u, const uint8_t* active,
                          float* out, int soa, void* stream);

/* f = -sum_e a_e Ke chi(macro), shape (1, 3N); macro is 6 host doubles in
 * Voigt order with engineering shear. Replaces `rhs`
 * (solver/src/materials/hybrid_kernels.cu:128-137). */
int pgj_hex8_rhs(const pgj_grid* g, const double* a, const double macro[6], double* out, void* stream);

/* Element energies e_l,e = (u_e + chi_l)^T Ke(C0) (u_e + chi_l) / V, out (L, N);
 * macros is L*6 host doubles. Replaces `evaluate`
 * (exp/asg_scenarios_20261105/energy.cu:26-35). */
int pgj_hex8_energy(const pgj_grid* g, int L, const double* u, const double* macros, double* out,
                    void* stream);

/* Homogenized tensor C^H (36 doubles, row-major) from the six unit-strain
 * solutions u6 (6, 3N): C_ij = < a (u_i + chi_i)^T Ke (u_j + chi_j) > / V.
 * Replaces PaperResident.effective_tensor
 * (exp/asg_sctop_reproduction_20261105/reproduction.py:166-176). */
size_t pgj_effective_tensor_ws_bytes(const pgj_grid* g);
int pgj_effective_tensor(const pgj_grid* g, const double* a, const double* u6, double* ws, double* out36,
                         void* stream);

/* C^H as pgj_effective_tensor, with no workspace argument and no allocation. No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The per-block partial sums live in a file-scope device array of
 * 2,752,512 bytes, one per GPU, held until the process ends; its address is
 * looked up on the first call on each device (make that call outside a
 * capture). Grids of more than pgj_effective_tensor_noalloc_max_elements()
 * (4,194,304) elements are rejected with PGJ_ERR_ARG. Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams or threads)
 * race; calls ordered on one stream are safe. */
int64_t pgj_effective_tensor_noalloc_max_elements(void);
int pgj_effective_tensor_noalloc(const pgj_grid* g, const double* a, const double* u6, double* out36,
                                 void* stream);

/* Mean stress <a C0 (eps(macro) + B_bar u)> (6 doubles). Replaces `stress`
 * (solver/src/materials/hybrid_kernels.cu:138-147). */
size_t pgj_avg_stress_ws_bytes(const pgj_grid* g);
int pgj_hex8_avg_stress(const pgj_grid* g, const double* a, const double* u, const double macro[6], double* ws,
                        double* out6, void* stream);

/* Nodal diagonal d of K(a) and the partial Green-Jacobi scale
 * s = exp(-alpha/2 (log d - mean log d)) over all 3N entries, both (1, 3N).
 * d may be NULL. Replaces `diagonal`/`set_alpha`
 * (solver/experiments/adaptive_green_scaling/study.py:80-101) and the inline
 * form in Resident.select (exp/asg_gpu_resident_20261103/resident.py:170-174). */
size_t pgj_scale_ws_bytes(const pgj_grid* g);
int pgj_diagonal_scale(const pgj_grid* g, const double* a, double alpha, double* d, double* s, double* ws,
                       void* stream);

/* Row-march launch shape of the HEX8 kernel for grid g with L loads per
 * launch in precision es (8: FP64, 4: FP32): R element rows per wave (1 or
 * 2), W row waves per block (1..8), ZC z-planes marched per block, plane = 1
 * for the plane-carry march (in-plane butterflies shared by the two element
 * layers of a node plane). set_shape records a tuned shape (R = W = ZC = 0
 * clears it); get_shape reports the shape in use and the blocks per load
 * (out5 = R, W, ZC, plane, blocks). Counterpart of the reference's tile/variant
 * sweep (solver/src/hex8_modal/modal_adapter.py:28-41, gpu_study.py:138-142). */
int pgj_hex8_set_shape(const pgj_grid* g, int L, int es, int R, int W, int ZC, int plane);
/* Radix plan of one FFT pass of the resident PCG iteration on the cube g in
 * precision es (8: FP64, 4: FP32): pass 0 forward x/y, 1 z with the Green
 * multiply, 2 inverse y/x; plan 0 takes radix-4 stages first (64 = 4 x 4 x 4),
 * plan 1 radix-8 stages first (64 = 8 x 8); for the z pass at 64^3, plan 2
 * keeps the transforms in registers (two radix-8 passes with one LDS transpose
 * each, LDS otherwise only for the 3x3 Green multiply). plan -1 restores the
 * default, plan 0. The
 * choice holds for every later launch of that pass on any n^3 grid of that
 * precision (a graph keeps the kernels of its capture; a solver reads the plan
 * when it sets up its launches). get_plan writes the three passes' plans.
 * Counterpart of the reference's variant sweep
 * (solver/src/hex8_modal/gpu_study.py:174-222) for the FFT radix schedule. */
int pgj_fft_set_plan(const pgj_grid* g, int es, int pass, int plan);
int pgj_fft_get_plan(const pgj_grid* g, int es, int* out3);
int pgj_hex8_get_shape(const pgj_grid* g, int L, int es, int* out5);

/* Layout conversions between AoS (L, N, 3) and SoA (L, 3, N) vectors. */
int pgj_aos_to_soa(const pgj_grid* g, int L, const double* in, double* out, void* stream);
int pgj_soa_to_aos(const pgj_grid* g, int L, const double* in, double* out, void* stream);

/* Zero-mean projection per load and component, in place.
 * Replaces `project_zero_mean` (solver/src/materials/hybrid_adapter.py:78-80). */
size_t pgj_project_ws_bytes(const pgj_grid* g, int L);
int pgj_project(const pgj_grid* g, int L, double* v, double* ws, void* stream);

/* ------------------------------------------------------------ Green / FFT */

/* Plan: FFT twiddles and the dense symmetric Green symbol (6 coefficients per
 * half-spectrum frequency, [G00,G11,G22,G01,G02,G12], 1/N folded in).
 * Replaces `fi_coeff` + `prepare`
 * (solver/src/reference_medium/formula_kernels.cu:199-203,
 *  solver/src/spectral/precision_kernels.cu:129-142) and the cuFFT plans
 * (solver/src/spectral/precision_fft.cu:34-56). Supported axis lengths:
 * products of 2, 3, 5 (and 7), each <= 256. */
size_t pgj_plan_bytes(const pgj_grid* g);
int pgj_plan_init(const pgj_grid* g, void* plan, void* stream);
/* Copies the symbol out of a plan as (6, nz, ny, nx/2+1) doubles. */
int pgj_plan_symbol(const pgj_grid* g, const void* plan, double* out, void* stream);
/* The factored FI Green symbol of the grid's (lam, mu): out4 (4, nz, ny, nx/2+1)
 * holds D_00, D_11, D_22 and beta with G = D - beta (D v)(D v)^T (no 1/N), and
 * vtab (nullable) the per-axis v tables, nx/2+1 then ny then nz values.
 * Replaces `formula.fi_coeff` and the v tables of `enable_fi_reference`
 * (solver/src/reference_medium/formula_kernels.cu:104-126, formula_adapter.py:49-54). */
int pgj_fi_coeff(const pgj_grid* g, double* out4, double* vtab, void* stream);

/* Spectrum of the scale s (3, nz, ny, nx/2+1) complex128, needed by the
 * fused projection of M_alpha; recompute whenever s changes. */
size_t pgj_precond_ws_bytes(const pgj_grid* g, int L);
int pgj_scale_spectrum(const pgj_grid* g, const void* plan, const double* s, void* s_hat, void* ws, void* stream);

/* z_l = P s G s P r_l (s NULL: z = P G P r), L loads. Replaces the native
 * Green apply plus projection/scaling
 * (solver/src/spectral/native_adapter.py:23-33,
 *  exp/asg_gpu_resident_20261103/resident.py:104-117,
 *  solver/experiments/adaptive_green_scaling/study.py:91-101). */
int pgj_precond_apply(const pgj_grid* g, int L, const void* plan, const double* r, const double* s,
                      const void* s_hat, double* z, void* ws, void* stream);
/* The same apply in FP32 arithmetic on FP64 (L, 3N) vectors: r is rounded to
 * FP32 as it is read, z is the FP32 result widened to FP64. plan32 from
 * pgj_plan32_init, s and s_hat FP32 / complex64 (nullable together; s_hat is
 * pgj_scale_spectrum's output rounded by pgj_to_f32), the
 * pgj_precond_ws_bytes workspace; cubic n in 16, 24, 32, 48, 64, 96, 128.
 * Replaces the 'mixed' (default) Green apply of enable_fi_reference
 * (solver/src/reference_medium/formula_adapter.py:40, :59-66). */
int pgj_precond_apply_mixed(const pgj_grid* g, int L, const void* plan32, const double* r, const float* s,
                            const void* s_hat, double* z, void* ws, void* stream);

/* ------------------------------------------------------ resident PCG (FP64) */

/* The resident solver keeps its vectors (u, r, p, q, z, rhs) and the scale s
 * in SoA layout, (L, 3, N) and (3, N): one contiguous plane per component, so
 * the FFT passes stream whole component planes. Convert with pgj_aos_to_soa /
 * pgj_soa_to_aos at the boundary. */

/* Per-load solver state, `state` of pgj_resident_state_bytes(L) bytes:
 *   double rz[L], target2[L], norm[L], absolute[L];
 *   int64  iterations[L];
 *   uint8  active[L], bad[L]   (8-byte aligned blocks, in this order),
 * followed by private fields. Offsets are given by pgj_resident_state_offsets. */
size_t pgj_resident_state_bytes(int L);
/* offsets[0..6] = byte offsets of rz, target2, norm, absolute, iterations, active, bad. */
int pgj_resident_state_offsets(int L, size_t* offsets7);
size_t pgj_resident_ws_bytes(const pgj_grid* g, int L);

/* Partial Green-Jacobi scale s (SoA, (3, N)) of the field a for exponent
 * alpha (alpha = 0 gives s = 1) and its spectrum s_hat (3, nz, ny, nx/2+1)
 * complex128. Replaces Resident.select (resident.py:156-174). */
int pgj_resident_scale(const pgj_grid* g, const void* plan, const double* a, double alpha, double* s, void* s_hat,
                       void* ws, void* stream);
/* pgj_resident_scale with alpha read from device memory (one double), as
 * written by pgj_alpha_select; selection and scaling then run without a host
 * round trip, as the reference's torch ops do (resident.py:156-174). */
int pgj_resident_scale_dev(const pgj_grid* g, const void* plan, const double* a, const double* alpha, double* s,
                           void* s_hat, void* ws, void* stream);

/* ------------------------------------------------------- alpha selectors */
/* On-device choice of alpha for the field a (N) and, for the gray selector,
 * the physical density rho (N). method: 0 fixed (params[0] = alpha; the eta
 * selector is this with the anchor of the Emin decade chosen on the host,
 * run.py:52), 1 green (0), 2 gj (1), 4 contrast (params = anchors b4, b6, b9;
 * run.py:28-33), 5 gray (params = b4, b6, b9, gamma, g_ref; run.py:37-43),
 * 6 compact8 and 7 compact8q (law features of a in FP32,
 * solver/experiments/global_alpha_compact_law/law.py:9-18, and the formula of
 * resident.py:163-168 / run.py:60-66). alpha (1 double, device) receives the
 * value; features (nullable, 7 doubles, device) receive log10 contrast
 * clamped to [0, 9], <rho>, <rho^2> (0 without rho) and area, jump, law
 * contrast, complexity (0 unless the method is 6 or 7).
 * ws holds pgj_alpha_ws_bytes(g) bytes. Replaces Resident.select /
 * SelectorResident.choose (resident.py:156-168, run.py:53-67). */
size_t pgj_alpha_ws_bytes(const pgj_grid* g);
int pgj_alpha_select(const pgj_grid* g, int method, const double* params, const double* a, const double* rho,
                     double* ws, double* features, double* alpha, void* stream);
/* Resident sets: the L loads of a resident set form L/group consecutive
 * designs of `group` loads each (group divides L, L <= 4096). Design d owns
 * a[d] (N), s[d] (3, N) and s_hat[d] (3, nz, ny, nx/2+1); group = L is one
 * design, the reference's case, and several designs of one grid in one
 * launch is the batched form. Vectors, `state` and `ws` are laid out for all
 * L loads; each call below works on the slice of loads [l0, l0 + nl), which
 * covers whole designs or lies within one design, and leaves every other load
 * untouched, so designs can join and leave a running set (continuous
 * batching) and the loads of one design can run as slices on several
 * streams. Pass l0 = 0, nl = L for the whole set. All pointers are to the
 * whole set's buffers, and every call on a set passes the set's L: the FFT
 * passes' block shapes, and so the layout of the partial sums one pass leaves
 * for the next, follow it. */

/* 1 if the z pass of a resident set of L loads (its workspace's load count) with
 * elements of es bytes runs on the wide kx chunks on the current device (64^3
 * FP64 where the default chunks need more than one round of blocks), else 0. */
int pgj_zpass_wide(const pgj_grid* g, int L, int es);

/* rhs_l for the L macro strains (host L*6), norms, targets
 * max(rtol |f|, 1e-13/sqrt(nx))^2 (NaN if |f| is NaN, as torch.maximum),
 * iterations = 0, bad = 0, and the per-load iteration cap `cap` (>= 1; the
 * reference's is 4000, exp/asg_gpu_resident_20261103/resident.py:50).
 * Replaces Resident._solve_batch set-up (resident.py:186-191). */
int pgj_resident_setup(const pgj_grid* g, int L, int group, int l0, int nl, const double* a, const double* macros,
                       double rtol, int64_t cap, double* rhs, void* state, void* ws, void* stream);
/* u <- P u; r = rhs - K u (reference operator); absolute = |r|.
 * Replaces Resident.certify (resident.py:133-136). */
int pgj_resident_certify(const pgj_grid* g, int L, int group, int l0, int nl, const double* a, double* u,
                         const double* rhs, double* r, void* state, void* ws, void* stream);
/* active = |r|^2 > target2 & iterations < cap & !bad; first direction
 * p = M^-1 r (written to p). Replaces Resident.initialize
 * (resident.py:119-123). */
int pgj_resident_init(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan, const double* s,
                      const void* s_hat, const double* r, double* p, void* state, void* ws, void* stream);
/* `iters` (even) masked PCG iterations; p0/p1 are ping-pong direction
 * buffers (the first iteration reads p0, as written by pgj_resident_init),
 * q is a scratch vector, all (L, 3N). The preconditioned residual z is not
 * stored: the inverse FFT pass writes p_next = z + beta p directly. Replaces
 * Resident.iteration with update_ur/update_p (resident.py:39-58, 125-131). */
int pgj_resident_iterate(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan, const double* a,
                         const double* s, const void* s_hat, double* u, double* r, double* p0, double* p1, double* q,
                         void* state, void* ws, int iters, void* stream);
/* One stage of one iteration (0: q = K p, 1: vector update + forward xy FFT,
 * 2: z pass + Green, 3: inverse xy FFT + projection + p_next = z + beta p),
 * with p the direction of this iteration; pgj_resident_iterate is these four
 * stages with p0/p1 alternating. Exposed so each kernel can be timed alone. */
int pgj_resident_stage(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan, const double* a,
                       const double* s, const void* s_hat, double* u, double* r, double* p, double* p_next, double* q,
                       void* state, void* ws, int stage, void* stream);

/* ------------------------------------------ mixed precision (FP32 inner) */
/* FP32 inner PCG inside FP64 iterative refinement. The inner solver is the
 * resident iteration in FP32 (row-march HEX8, compile-time Stockham FFT and
 * Green multiply; FP64 reductions and PCG scalars) on the same slices of a
 * resident set as the FP64 entry points; cubic n in 16, 24, 32, 48, 64, 96,
 * 128 only. One refinement step is: pgj_resident_certify (FP64 r = f - K u,
 * |r|), pgj_inner_begin, pgj_inner_init, pgj_inner_iterate until no inner
 * load is active, pgj_inner_end. No reference counterpart: the reference's
 * refinement precedent is solve_ir_batched
 * (solver/src/runtime/fi/source/femlib/batched/ir.py:23-107), which refines
 * with a generic inner solver. */
size_t pgj_plan32_bytes(const pgj_grid* g);
/* FP32 copy of a plan (twiddles and symbol rounded from the FP64 plan). */
int pgj_plan32_init(const pgj_grid* g, const void* plan, void* plan32, void* stream);
/* y = (float) x for n values (fields, scales; a complex spectrum is 2n reals). */
int pgj_to_f32(int64_t n, const double* x, float* y, void* stream);
/* Workspace of the inner solver for a set of L loads (FP32 spectrum). */
size_t pgj_inner_ws_bytes(const pgj_grid* g, int L);
/* Per load of [l0, l0+nl): s_l = |r_l| from the FP64 state of the last
 * certification, or 0 if that load is certified; r32 = r64 / s_l; u32 = 0;
 * inner target |r32| <= max(rtol, min(target_l / (2 s_l), 1/2)), so the last
 * step stops once the FP64 target can be met. state32 is a resident state
 * block of L loads. */
int pgj_inner_begin(const pgj_grid* g, int L, int l0, int nl, const double* r64, const void* state64, double rtol,
                    float* r32, float* u32, void* state32, void* ws32, void* stream);
/* pgj_resident_init / _stage / _iterate in FP32 (plan32, FP32 fields and vectors). */
int pgj_inner_init(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32, const float* s,
                   const void* s_hat, const float* r, float* p, void* state, void* ws, void* stream);
int pgj_inner_stage(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32, const float* a,
                    const float* s, const void* s_hat, float* u, float* r, float* p, float* p_next, float* q,
                    void* state, void* ws, int stage, void* stream);
int pgj_inner_iterate(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32, const float* a,
                      const float* s, const void* s_hat, float* u, float* r, float* p0, float* p1, float* q,
                      void* state, void* ws, int iters, void* stream);
/* u64 += s_l * u32 for the loads of [l0, l0+nl) (FP64 accumulation). */
int pgj_inner_end(const pgj_grid* g, int L, int l0, int nl, const float* u32, const void* ws32, double* u64,
                  void* stream);

/* ------------------------------------------------------------ OC updates */

/* Fields are (n, n, n) FP64. `ws` holds pgj_oc_ws_bytes(n) bytes of per-block
 * partials (volume sums or maxima). Products are evaluated without FMA
 * contraction, in the reference's operation order. */
size_t pgj_oc_ws_bytes(int n);

/* Clipped cone filter y_i = sum_j w_j x[i + o_j] over in-range neighbours,
 * offsets (m, 3) int32 in (dx, dy, dz). With conditional != 0 it does nothing
 * once bounds[2] == 0. Replaces `filter`
 * (exp/asg_sctop_reproduction_20261105/outer.cu:8-19, :218-224). */
int pgj_cone_filter(int n, const double* x, const int32_t* offsets, const double* weights, int m, double* y,
                    const double* bounds, int conditional, void* stream);
/* eta-projection of t/hs: ph, dp, a = Emin + (1-Emin) ph^3; ws (nullable)
 * receives volume partials. Replaces `project` (outer.cu:20-28, :225-229). */
int pgj_cone_project(int n, const double* t, const double* hs, double* ph, double* dp, double* a,
                     const double* bounds, double beta, double eta, double emin, int conditional, double* ws,
                     void* stream);
/* gin = dp (3 (1-Emin) ph^2) e / hs, dvin = dp / hs (reproduction.py:92-94). */
int pgj_cone_sensitivity(int n, const double* dp, const double* ph, const double* energy, const double* hs,
                         double emin, double* gin, double* dvin, void* stream);
/* g /= n^3, ratio = -g/dv, bounds = [0, 1e9, 1, 0] (reproduction.py:93-96). */
int pgj_cone_ratio(int n, double* g, const double* dv, double* ratio, double* bounds, void* stream);
/* OC candidate clip(x (ratio/mid)^0.3, x -+ move, [0.001, 1]) while bounds[2] != 0.
 * Replaces `candidate` (outer.cu:29-34) with the move limit as a parameter. */
int pgj_cone_candidate(int n, const double* x, const double* ratio, const double* bounds, double move, double* y,
                       void* stream);
/* One bisection step on the volume (mean of the projected candidate, from
 * pgj_cone_project's ws) with relative stop 1e-4; bounds[3] counts steps.
 * Replaces `bisect` (outer.cu:35-39). */
int pgj_cone_bisect(int n, const double* ws, double* bounds, double volume, void* stream);
/* out[0] = max |a - b| (the design change). */
int pgj_absmax_diff(int n, const double* a, const double* b, double* ws, double* out, void* stream);

/* Periodic separable Gaussian (axes z, y, x) in scipy.ndimage's symmetric
 * accumulation order; weights has 2*radius+1 entries. Replaces `filter`
 * (exp/asg_gpu_resident_20261103/outer.cu:7-17, :100-107). */
int pgj_gauss_filter(int n, const double* x, const double* weights, int radius, double* tmp1, double* tmp2,
                     double* y, void* stream);
/* tanh projection around 0.5 and a = Emin + (1-Emin) ph^3; ws (nullable)
 * receives volume partials. Replaces `physical` (outer.cu:18-23, :108-114). */
int pgj_gauss_project(int n, const double* t, double beta, double* ph, double* dp, double* a, double emin,
                      double* ws, void* stream);
/* gin = dp p (1-Emin) ph^2 e (outer.py:35). */
int pgj_gauss_sensitivity(int n, const double* dp, const double* ph, const double* energy, double emin,
                          double penal, double* gin, void* stream);
/* ratio = max(g,1e-30)/max(dv,1e-30), bounds = [0, 1e4 max ratio] (outer.py:37-38). */
int pgj_gauss_ratio(int n, const double* g, const double* dv, double* ratio, double* ws, double* bounds,
                    void* stream);
/* candidate clip(x sqrt(ratio/lam), [x-move, x+move] cap [0,1]) (outer.cu:24-29). */
int pgj_gauss_candidate(int n, const double* x, const double* ratio, const double* bounds, double move, double* y,
                        void* stream);
/* bisection on the mean of the projected candidate (outer.cu:30-33). */
int pgj_gauss_bisect(int n, const double* ws, double* bounds, double volume, void* stream);

/* --------------------------------------------- generic batched PCG vectors */
/* Vector side of the operator-agnostic best-iterate PCG on B batches of n
 * FP64 unknowns, (B, n) row-major; the caller supplies A p and M^-1 r. One
 * iteration is pgj_pcg_pre_step, then (unless stopping) M^-1 r and
 * pgj_pcg_post_step. Together they replace the eager/compiled/Jacobi update
 * paths of solve_pcg_batched
 * (solver/src/runtime/fi/source/femlib/batched/pcg.py:20-245): the fused
 * compiled steps (pcg.py:622-666, :741-745), the convergence latches
 * (pcg.py:677-712), the latched best copy (pcg.py:715-738) and the Triton/CUDA Jacobi vector
 * kernels (pcg.py:459-532). ws holds 3 B ceil(n/1024)-capped partial rows. */
size_t pgj_pcg_ws_bytes(int B, int64_t n);
/* r = rhs - Ax (Ax NULL: r = rhs) and rhs_norm = max(|rhs_b|, 1e-30) (pcg.py:79-85). */
int pgj_pcg_start(int B, int64_t n, const double* rhs, const double* Ax, double* r, double* rhs_norm, double* ws,
                  void* stream);
/* p = z, rz = r.z per batch (pcg.py:87-89). */
int pgj_pcg_init(int B, int64_t n, const double* r, const double* z, double* p, double* rz, double* ws,
                 void* stream);
/* alpha = rz/max(pAp,1e-30), 0 where pAp is not finite or <= 0; u += alpha p;
 * r -= alpha Ap; res = |r|/rhs_norm; improved = res < best_res (latched: and
 * state[0] == 0); best_res and best_u updated where improved. latched != 0:
 * state (int64[2]) gets state[0] = iteration when every res < tol first holds,
 * and state[1] = 1 when every res < near_tol first holds (near_tol > 0 only)
 * (pcg.py:181-199, :637-666, :677-738). improved is uint8[B]. */
int pgj_pcg_pre_step(int B, int64_t n, double* u, double* r, const double* p, const double* Ap, const double* rz,
                     const double* rhs_norm, double* best_u, double* best_res, double* res, uint8_t* improved,
                     int64_t* state, double tol, double near_tol, int64_t iteration, int latched, double* ws,
                     void* stream);
/* rz_new = r.z; p = z + rz_new/max(rz,1e-30) p; rz = rz_new (pcg.py:238-245,
 * :741-745). minv non-NULL: z = minv * r is formed first (Jacobi,
 * pcg.py:495-508); minv_stride is 0 (one (n) diagonal for every batch) or n. */
int pgj_pcg_post_step(int B, int64_t n, const double* r, double* z, const double* minv, int64_t minv_stride,
                      double* p, double* rz, double* ws, void* stream);
/* The same four steps on FP32 vectors (u, r, p, Ap, z, best_u, minv); sums,
 * rz, rhs_norm, res and best_res stay FP64. The inner PCG of
 * solve_ir_batched (solver/src/runtime/fi/source/femlib/batched/ir.py:84-92)
 * runs on these. */
int pgj_pcg32_start(int B, int64_t n, const float* rhs, const float* Ax, float* r, double* rhs_norm, double* ws,
                    void* stream);
int pgj_pcg32_init(int B, int64_t n, const float* r, const float* z, float* p, double* rz, double* ws, void* stream);
int pgj_pcg32_pre_step(int B, int64_t n, float* u, float* r, const float* p, const float* Ap, const double* rz,
                       const double* rhs_norm, float* best_u, double* best_res, double* res, uint8_t* improved,
                       int64_t* state, double tol, double near_tol, int64_t iteration, int latched, double* ws,
                       void* stream);
int pgj_pcg32_post_step(int B, int64_t n, const float* r, float* z, const float* minv, int64_t minv_stride, float* p,
                        double* rz, double* ws, void* stream);

/* ---- general-C21 material (hybrid storage) ----
 * Element e has the stiffness C_e given by its 21 upper-triangle Voigt
 * entries in row-major order ((0,0),(0,1),...,(0,5),(1,1),...,(5,5), as
 * solver/src/materials/material_field.py:14 PAIRS), read from
 * pool + 21*index[e] (index != NULL: shared phase rows and per-voxel rows,
 * solver/src/materials/hybrid_material.py:14-70) or pool + 21*e (index NULL,
 * the dense ablation). Only the grid's extents and voxel sizes are used.
 * Index entries must lie in [0, pool rows): the kernels read the row they
 * name without a bounds check (pgj.hybrid.HybridFIOperator checks once). */

/* out_l = K(C) u_l for L AoS (L, 3N) loads; active (uint8[L], nullable)
 * zeroes inactive loads. Each block owns a 16 x 16 node column over a z-chunk
 * and evaluates every element once (plus a one-element halo); node sums are
 * formed in a fixed order in LDS (no atomics). Replaces `apply`
 * (solver/src/materials/hybrid_kernels.cu:25-48, host :117-127). */
int pgj_hybrid_apply(const pgj_grid* g, int L, const double* pool, const int32_t* index, const double* u,
                     const uint8_t* active, double* out, void* stream);
/* The same K by an independent thread-per-node gather that evaluates each of
 * the node's eight elements (an eight-fold compute redundancy); a second code
 * path for certification and for checking pgj_hybrid_apply. */
int pgj_hybrid_apply_gather(const pgj_grid* g, int L, const double* pool, const int32_t* index, const double* u,
                            const uint8_t* active, double* out, void* stream);
/* f = -V sum_e Bbar_e^T C_e eps(macro), (1, 3N); macro is 6 host doubles.
 * Replaces `rhs` (solver/src/materials/hybrid_kernels.cu:49-68, host :128-137). */
int pgj_hybrid_rhs(const pgj_grid* g, const double* pool, const int32_t* index, const double macro[6], double* out,
                   void* stream);
/* Mean stress <C_e (eps(macro) + Bbar u_e)> (6 doubles on the device).
 * Replaces `stress` (solver/src/materials/hybrid_kernels.cu:69-106, host
 * :138-147); ws holds 6 partial sums per 256 elements. */
size_t pgj_hybrid_stress_ws_bytes(const pgj_grid* g);
int pgj_hybrid_avg_stress(const pgj_grid* g, const double* pool, const int32_t* index, const double* u,
                          const double macro[6], double* ws, double* out6, void* stream);

/* Resident PCG over the general-C21 operator: one design of L loads (group = L)
 * with the resident state, workspace (pgj_resident_state_bytes,
 * pgj_resident_ws_bytes), first direction (pgj_resident_init) and FFT stages of
 * the scalar solver. The grid's lam, mu are the Green reference medium (the
 * zeman medium of the pool); s/s_hat scale the preconditioner as in the scalar
 * solver (pass the unit scale for M = P G P). The reference solves this
 * operator with the generic PCG inside compute_effective_stiffness_batched
 * (solver/src/runtime/fi/source/femlib/batched/driver.py:13-75, pcg.py:20-252);
 * these run it as the resident recurrence with certification
 * (exp/asg_gpu_resident_20261103/resident.py:39-58, :186-209). SoA vectors.
 * setup: rhs_l = pgj_hybrid_rhs of the L macro strains (host L*6), norms,
 * targets and the iteration cap, as pgj_resident_setup. */
int pgj_hybrid_resident_setup(const pgj_grid* g, int L, int l0, int nl, const double* pool, const int32_t* index,
                              const double* macros, double rtol, int64_t cap, double* rhs, void* state, void* ws,
                              void* stream);
/* u <- P u; r = rhs - K u by the gather apply; absolute = |r| (pgj_resident_certify). */
int pgj_hybrid_resident_certify(const pgj_grid* g, int L, int l0, int nl, const double* pool, const int32_t* index,
                                double* u, const double* rhs, double* r, void* state, void* ws, void* stream);
/* Stage 0 is q = K p by the tiled apply (p.q partials per block); stages 1-3
 * and the iteration loop as pgj_resident_stage / pgj_resident_iterate. */
int pgj_hybrid_resident_stage(const pgj_grid* g, int L, int l0, int nl, const void* plan, const double* pool,
                              const int32_t* index, const double* s, const void* s_hat, double* u, double* r,
                              double* p, double* p_next, double* q, void* state, void* ws, int stage, void* stream);
int pgj_hybrid_resident_iterate(const pgj_grid* g, int L, int l0, int nl, const void* plan, const double* pool,
                                const int32_t* index, const double* s, const void* s_hat, double* u, double* r,
                                double* p0, double* p1, double* q, void* state, void* ws, int iters, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PGJ_H */
