/*
 * em3d_hip.h -- C-ABI of libem3d_hip.so, the MI355X (gfx950) implementation of
 * the FFT volume-integral operator and Krylov vector kernels of em3d.
 *
 * Conventions
 *   - Every function returns int: 0 = success, non-zero = failure with the
 *     message in em3d_last_error() (per calling thread).  No exception crosses.
 *   - All array pointers are DEVICE pointers owned by the caller (e.g. torch
 *     data_ptr()); the library never frees them.  Work is enqueued on the
 *     caller's hipStream_t and is asynchronous unless stated.
 *   - Complex arrays are interleaved (re, im): complex64 for precision 0,
 *     complex128 for precision 1.  Layouts are C order, z fastest:
 *       field  u, y   (batch, ncomp, Nx, Ny, Nz)          ncomp = 3 (EM) or 1
 *       contrast chi  CHI_ISO (Nx,Ny,Nz) | CHI_DIAG (3,Nx,Ny,Nz) |
 *                     CHI_FULL (3,3,Nx,Ny,Nz)   (acoustic: CHI_ISO only)
 *       khat          octant spectrum (K, Nx+1, Ny+1, Nz+1), K = 6 for EM
 *                     (xx,yy,zz,xy,xz,yz) or 1 for acoustic, pre-scaled by
 *                     1/(8 Nx Ny Nz)
 *   - A plan handle must not be used by two host threads at once.
 *
 * Reference interfaces replaced (qwerty29544/em3d 0.8.0, paths under src/em3d):
 * see the comment above each declaration.
 */
#ifndef EM3D_HIP_H
#define EM3D_HIP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct em3d_plan_s* em3d_plan_t;

enum { EM3D_SINGLE = 0, EM3D_DOUBLE = 1 };
enum { EM3D_CHI_NONE = 0, EM3D_CHI_ISO = 1, EM3D_CHI_DIAG = 2, EM3D_CHI_FULL = 3 };

/* Library identification string. */
const char* em3d_version(void);

/* Message of the last failure on this thread ("" if none). */
const char* em3d_last_error(void);

/* Plan for an (nx, ny, nz) grid, ncomp = 3 (EM, Operator) or 1 (acoustic,
 * AcousticOperator), `batch` right-hand sides per call, precision 0/1.
 * Holds the radix plans and twiddle tables of the three doubled axes.
 * Replaces the per-operator state set up in Operator.__init__
 * (operator.py:321-338) and AcousticOperator.__init__
 * (acoustics/operator.py:28-34).  Each doubled axis 2n must factor into
 * 2, 3, 5, 7.  Synchronous (copies the twiddle tables). */
int em3d_plan_create(int nx, int ny, int nz, int ncomp, int batch, int precision, em3d_plan_t* plan);
int em3d_plan_destroy(em3d_plan_t plan);

/* Byte sizes of the octant spectrum, of the per-call workspace of
 * em3d_matvec/em3d_rmatvec, and of the FP64 scratch of em3d_kernel_build. */
int em3d_plan_sizes(em3d_plan_t plan, size_t* khat_bytes, size_t* work_bytes, size_t* build_bytes);

/* Spectrum of the discrete kernel on the doubled grid, computed in FP64 on the
 * device (spatial kernel, 3-D FFT, octant extraction) and stored in the plan
 * precision.  EM: dyadic Green blocks with self term -I/3
 * (PreparedEMKernel.build, operator.py:171-223; _kernel_geometry and
 * _kernel_block_from_geometry, operator.py:28-84).  Acoustic: k0^2 dv G with
 * the equivalent-sphere self cell (prep_coeffs_acoustic,
 * acoustics/kernel.py:21-49).  h = cell sizes L/N per axis. */
int em3d_kernel_build(em3d_plan_t plan, double k0, double hx, double hy, double hz, void* khat,
                      void* build_work, hipStream_t stream);

/* y = u - B(chi u) for all batch fields (Operator.matvec, operator.py:353-363;
 * AcousticOperator.matvec, acoustics/operator.py:40-54).  `work` holds
 * em3d_plan_sizes().work_bytes; y must not alias u. */
int em3d_matvec(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                void* work, hipStream_t stream);

/* y = v - chi^H B^H v (Operator.rmatvec with the derived adjoint
 * conj(Khat[-k]), operator.py:365-398 and 305-317;
 * AcousticOperator.rmatvec, acoustics/operator.py:56-67). */
int em3d_rmatvec(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* v, void* y,
                 void* work, hipStream_t stream);

/* em3d_matvec whose forward transforms read only the box of cells
 * [lo, lo + n) (lo[a] >= 0, n[a] >= 1, lo[a] + n[a] <= N[a]): chi u is taken
 * as zero outside it, and y = u - B(chi u) is formed on the whole grid.  It
 * equals em3d_matvec whenever chi u vanishes outside the box, e.g. when the box
 * holds supp(chi), and then skips the forward work on the rest of the grid.
 * No counterpart in the reference, whose operator always transforms the whole
 * grid (operator.py:353-363); its caller here is the support-restricted
 * solve, whose final field is f - A u~ with u~ zero outside the contrast's box. */
int em3d_matvec_window(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                       void* work, const int* lo, const int* n, hipStream_t stream);

/* em3d_matvec without a caller workspace: the FFT workspace is a fixed
 * file-scope device array of EM3D_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends), so the call allocates nothing.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Fails with a message when the plan's workspace (em3d_plan_sizes
 * work_bytes) exceeds the array: 64 MiB holds EM grids up to 64^3 complex64
 * (37.7 MB) and acoustic grids up to 64^3 complex128.  The plan must belong
 * to the current device.  Calls on one GPU share the array: calls that
 * overlap on that GPU (two streams, two threads) race. */
#define EM3D_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int em3d_matvec_noalloc(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                        hipStream_t stream);

/* Per-launch timing of em3d_matvec/em3d_rmatvec on this plan, the measured
 * counterpart of the reference's matvec timer (benchmark_matvec,
 * experiments/scans.py:275-300).  While enabled, each apply call records a
 * hipEvent before every kernel launch and after the last one and WAITS for
 * them (one host sync per call); em3d_plan_timing returns the average
 * milliseconds of each launch slot (z pass, y pass, x pass, y pass, z pass)
 * over the calls since enabling.  Enabling resets the averages. */
int em3d_plan_set_timing(em3d_plan_t plan, int enable);
int em3d_plan_timing(em3d_plan_t plan, double* ms_per_launch, int max_launches, int* nlaunch, long long* calls);

/* Packed group of operators of different grid shapes (each with its own
 * plan, spectrum, contrast and buffers) applied by one hipGraph launch: the
 * first em3d_group_matvec captures every member's five passes as a parallel
 * branch of one graph (a private stream per member, forked from and joined
 * into a capture stream), later calls replay the graph on the caller's
 * stream, so the grids run concurrently.  y_i = A_i u_i
 * with the u_i, y_i, work_i given at creation (the graph holds these
 * pointers).  No counterpart in the reference (one operator per call,
 * operator.py:353-363). */
typedef struct em3d_group_s* em3d_group_t;
int em3d_group_create(int n, const em3d_plan_t* plans, const void* const* khats, const void* const* chis,
                      const int* chi_kinds, const void* const* inputs, void* const* outputs, void* const* works,
                      em3d_group_t* group);
int em3d_group_matvec(em3d_group_t group, hipStream_t stream);
int em3d_group_destroy(em3d_group_t group);

/* FFT autotuner (no counterpart in the reference, whose FFTs are
 * numpy/cuFFT calls, operator.py:353-363).  Each pass kind (0 strided y
 * passes, 1 x pass, 2 z passes) runs one of em3d_plan_variant_count() fast
 * kernel variants (radix plans, tile widths, pencils per block; variant 0 is
 * the default) or, with -1, the generic LDS kernels.  em3d_plan_autotune()
 * times whole matvecs on the caller's buffers (`reps` after one warm-up, per
 * candidate, coordinate search x -> strided -> z), keeps the fastest variant
 * of each kind in the plan and returns its time per matvec.  Synchronous;
 * y and work are overwritten. */
int em3d_plan_variant_count(em3d_plan_t plan, int kind);
int em3d_plan_set_variants(em3d_plan_t plan, const int* variants);
int em3d_plan_get_variants(em3d_plan_t plan, int* variants);
int em3d_plan_autotune(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                       void* work, int reps, double* best_ms, hipStream_t stream);

/* Device-resident BiCGStab on the operator of a batch-1 plan
 * (BiCGStab.solve, solvers/bicgstab.py:12-211, with SolverConfig fields of
 * solvers/base.py:7-25).  Same recurrence, early exit on ||s||, statuses and
 * history semantics as the reference; every scalar stays on the device and
 * the host reads the solver state back once per `chunk` iterations.
 * u (out) and rhs are device fields of the plan's shape.  `work` is the
 * plan's apply workspace, `solver_work` holds em3d_bicgstab_work_bytes().
 * history / action_counts are DEVICE arrays of history_cap entries that
 * receive the relative residuals and the action count of each entry.
 * Negative divergence_guard / max_operator_actions mean "not set". */
typedef struct {
  int max_iter;
  double rtol, atol;
  double divergence_guard;
  long long max_operator_actions;
  int chunk;
  /* 1: the p and s updates and the dot products run inside the matvecs' z
   * passes and the u/r update carries the next (r_hat, r) (isotropic or
   * diagonal contrast, no preconditioner); 0: stand-alone vector kernels */
  int fused;
} em3d_bicgstab_params;

/* status: 0 converged, 1 nonfinite, 2 max_operator_actions, 3 breakdown_rho,
 * 4 breakdown_omega, 5 breakdown_alpha, 6 breakdown_t, 7 divergence_guard,
 * 8 max_iter */
typedef struct {
  int status;
  int iterations;
  int matvec_count;
  int history_len;
  double true_final_residual;
} em3d_solver_result;

size_t em3d_bicgstab_work_bytes(em3d_plan_t plan);
int em3d_bicgstab(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* rhs, void* u,
                  void* work, void* solver_work, const em3d_bicgstab_params* params, double* history,
                  long long* action_counts, int history_cap, em3d_solver_result* result, hipStream_t stream);

/* Bytes of the reduction scratch used by em3d_dotc. */
size_t em3d_reduce_work_bytes(void);

/* out[0..1] (device, two doubles) = sum conj(x_i) y_i accumulated in FP64
 * (xp.vdot and xp.linalg.norm in the solvers, solvers/bicgstab.py:64-223,
 * solvers/sim.py:48, solvers/twostep.py:99-100). */
int em3d_dotc(int precision, long long n, const void* x, const void* y, void* out, void* red_work,
              hipStream_t stream);

/* out = a x + b y + c z (z may be NULL; out may alias x, y or z elementwise)
 * (the array updates of solvers/bicgstab.py:145-221, sim.py:78,
 * twostep.py:192-212). */
int em3d_axpbypcz(int precision, long long n, double a_re, double a_im, const void* x, double b_re,
                  double b_im, const void* y, double c_re, double c_im, const void* z, void* out,
                  hipStream_t stream);

/* dst = a src (+ dst when accumulate != 0), converting between complex64
 * (precision 0) and complex128 (1): the precision changes of the mixed
 * precision refinement, which the reference has no counterpart for (its
 * solvers run in one precision, solvers/bicgstab.py:12-211). */
int em3d_convert_axpy(long long n, int src_precision, const void* src, double a_re, double a_im, int dst_precision,
                      void* dst, int accumulate, hipStream_t stream);

/* ---- restriction to the support of the contrast --------------------------
 * A u = u - B(chi u) (operator.py:353-363) reads u only where chi != 0, so
 * for cells i, j of a box holding supp(chi), (A u)_i depends on u inside the
 * box alone, through the same kernel values (offsets (i - j) h): the operator
 * restricted to the box is the operator of a smaller grid with the same cell
 * size.  Outside the box the solution is explicit, u = f + B(chi u).  No
 * counterpart in the reference, which always solves on the whole grid.
 *
 * em3d_support_bounds: bounds (DEVICE int[6]) = x_lo, x_hi, y_lo, y_hi,
 * z_lo, z_hi (inclusive) of the cells where any entry of the packed contrast
 * (layouts as for em3d_matvec) is non-zero; (n, -1) on an axis when chi is
 * zero everywhere.
 * The box may take its own axis order: box axis a is grid axis perm[a], so
 * the box grid has sides box[perm[0]], box[perm[1]], box[perm[2]] (the
 * operator is covariant under axis permutations when vector components and
 * contrast indices permute with the axes).  grid, box, lo and perm are HOST
 * int[3] arrays in grid-axis order; planes come in groups of ncc = 1
 * (scalar), 3 (vector components) or 9 (3 x 3 contrast entries).
 * em3d_support_gather: dst (nplanes, box axes) = src (nplanes, nx, ny, nz)
 * cropped at offset lo, components permuted; nplanes = batch x ncc.
 * em3d_support_scatter: out (nplanes, nx, ny, nz) = the box field inside the
 * box and, outside it, f - y when f and y are given, else 0.  out may alias
 * f or y, not the box field.  The box must lie inside the grid. */
int em3d_support_bounds(int nx, int ny, int nz, int precision, const void* chi, int chi_kind, int* bounds,
                        hipStream_t stream);
int em3d_support_gather(int precision, long long nplanes, int ncc, const int* grid, const int* box, const int* lo,
                        const int* perm, const void* src, void* dst, hipStream_t stream);
int em3d_support_scatter(int precision, long long nplanes, int ncc, const int* grid, const int* box, const int* lo,
                         const int* perm, const void* boxed, const void* f, const void* y, void* out,
                         hipStream_t stream);

/* Far-field scatter integral F_b(d_m) = dv sum_q (chi u_b)(r_q)
 * exp(-i k0 d_m . r_q) (scatter_integral_direct, farfield/_core.py:8-71;
 * acoustics/farfield.py for ncomp = 1), accumulated in FP64.  u: (batch,
 * ncomp, N) fields, chi packed as for em3d_matvec (acoustic: chi = eta - 1,
 * kind 1); dirs: DEVICE (ndir, 3) doubles; out: DEVICE (batch, ndir, ncomp)
 * complex128; work: em3d_scatter_work_bytes(nx, ny, nz, batch, ndir, ncomp)
 * bytes (partial sums, per-axis phase tables, a collapsed field).
 * shared_axis: -1, or the axis (0 x, 1 y, 2 z) along which every direction
 * has the same component, exactly (a plane cut: rcs_plane/pattern_plane,
 * farfield/__init__.py:86-125); the sum over that axis is then taken once
 * for all directions.  The caller guarantees it: direction 0's component is
 * used for all. */
size_t em3d_scatter_work_bytes(int nx, int ny, int nz, int batch, int ndir, int ncomp);
int em3d_scatter_integral(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                          int batch, int ncomp, const void* chi, int chi_kind, const void* u, int ndir,
                          const double* dirs, int shared_axis, double k0, double dv, double* out, void* work,
                          hipStream_t stream);

/* Batch of plane waves f_b = A_b exp(i sign (k o_b . r + phi0_b)), out
 * (batch, 3, nx, ny, nz) in `precision`; ax/ay/az are the grid's cell-centre
 * axes on the device in the precision's real type; params is a DEVICE array of
 * batch x 10 doubles: o(3), Re A(3), Im A(3), phi0.  The phase is evaluated in
 * FP64.  Replaces one flat_wave_vec call per field (wave.py:9-37). */
int em3d_plane_waves(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision, int batch,
                     const double* params, double k, int sign, void* out, hipStream_t stream);

/* Contrast chi = eps - I of a background with ellipsoidal inclusions tested
 * at cell centres in FP64 with tolerance `tol` (center_mask and
 * sample_contrast_tensor, geometry.py:62-93, 233-252); later ellipsoids
 * overwrite earlier ones.  ellipsoids: DEVICE array of n x 24 doubles
 * (centre(3), radii(3), eps 3x3 complex row-major re/im); background: DEVICE
 * array of 18 doubles.  chi_kind selects the packed output: 1 ISO (eps[0][0]),
 * 2 DIAG, 3 FULL. */
int em3d_ellipsoid_contrast(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                            int n_ellipsoids, const double* ellipsoids, const double* background, double tol,
                            int chi_kind, void* chi, hipStream_t stream);

/* Acoustic eta of a background with ellipsoidal inclusions tested in the
 * axes' real type without tolerance (eta_ellipsoid,
 * acoustics/materials.py:62-83); later inclusions overwrite earlier ones.
 * inclusions: DEVICE array of n x 8 doubles (centre(3), radii(3), Re eta,
 * Im eta). */
int em3d_eta_inclusions(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                        int n_inclusions, const double* inclusions, double outside_re, double outside_im, void* eta,
                        hipStream_t stream);

/* ---- learned preconditioner: Fourier neural operator (FNO) ----------------
 * No counterpart in the reference (its solvers are unpreconditioned,
 * solvers/bicgstab.py:12-211).  A network on a grid^3 coarse grid with
 * `channels` hidden channels and `layers` Fourier layers keeping `modes`
 * modes per axis (k in [-modes/2, modes/2)), linear in its in_channels
 * inputs and conditioned on material_channels inputs through multiplicative
 * gates:  v0 = W_in x,  v_{l+1} = sigmoid(U_l m + d_l) * (W_l v_l + S_l v_l),
 * y = W_out v_L.  Pointwise channel mixing runs on bf16 MFMA with fp32
 * accumulation.  Parameters: fp32 master array of em3d_fno_param_count()
 * floats with a bf16 copy (unsigned short) the forward pass reads; inputs are
 * channels-last [samples][cells][in_channels + material_channels] fp32,
 * outputs [samples][cells][out_channels].  work: em3d_fno_work_bytes(),
 * initialised once by em3d_fno_init_work() (twiddle table); a forward call
 * keeps the activations em3d_fno_backward() of the same samples needs. */
typedef struct {
  int grid, channels, modes, layers, in_channels, out_channels, material_channels;
} em3d_fno_config;

long long em3d_fno_param_count(const em3d_fno_config* config);
size_t em3d_fno_work_bytes(const em3d_fno_config* config, int samples);
int em3d_fno_init_work(const em3d_fno_config* config, int samples, void* work, hipStream_t stream);
int em3d_fno_forward(const em3d_fno_config* config, const float* params, const unsigned short* params_bf16,
                     const float* x, float* y, void* work, int samples, hipStream_t stream);
/* grad (fp32, param layout) += dL/dparams for dL/dy = dy */
int em3d_fno_backward(const em3d_fno_config* config, const unsigned short* params_bf16, const float* x,
                      const float* dy, float* grad, void* work, int samples, hipStream_t stream);
/* loss (device double, accumulated) += sum_s |y_s - t_s|^2 / |t_s|^2;
 * dy = gradient of (that sum / samples); work: em3d_fno_loss_work_bytes() */
size_t em3d_fno_loss_work_bytes(int samples);
int em3d_fno_loss(int samples, long long per_sample, const float* y, const float* target, float* dy, double* loss,
                  double* work, hipStream_t stream);
int em3d_fno_adam(long long n, float* params, unsigned short* params_bf16, const float* grad, float* m, float* v,
                  double lr, double beta1, double beta2, double eps, double weight_decay, int step,
                  hipStream_t stream);
int em3d_fno_to_bf16(long long n, const float* params, unsigned short* params_bf16, hipStream_t stream);

/* Preconditioner M r = r + s P(y),  y = FNO(R r / s, chi features),
 * s = |r| / sqrt(ncomp N^3) per right-hand side, R = mean over factor^3
 * fine cells, P = trilinear interpolation back.  The network runs on two real samples per
 * right-hand side (real parts, imaginary parts), so M is complex-linear.
 * Grids are (grid * factor)^3; in_channels = out_channels = ncomp,
 * material_channels = chi_channels; work and features are sized for
 * 2 batch samples.
 * chi_features: device [cells][chi_channels] from em3d_fno_chi_features;
 * features: em3d_fno_features_bytes(config, batch); work as above for
 * `batch` samples. */
typedef struct {
  em3d_fno_config config;
  int factor;
  const float* params;
  const unsigned short* params_bf16;
  void* work;
  float* features;
  const float* chi_features;
  int chi_channels;
} em3d_fno_model;

size_t em3d_fno_features_bytes(const em3d_fno_config* config, int samples);
/* scale[b] = |field_b| / sqrt(n) (1 when zero) for batch fields of n complex
 * values; work: em3d_fno_scale_work_bytes(batch). */
size_t em3d_fno_scale_work_bytes(int batch);
int em3d_fno_scale(int batch, long long n, int precision, const void* field, double* scale, void* work,
                   hipStream_t stream);
/* out[2b + p][cell] = (mean over the cell's factor^3 fine cells of the real
 * (p = 0) or imaginary (p = 1) part of field_b / scale[b] per component, then
 * the chi_channels chi features) */
int em3d_fno_restrict(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* field,
                      const double* scale, const float* chi_features, int chi_channels, float* out,
                      hipStream_t stream);
/* chi features: means of the diagonal of chi over each coarse cell (re, im;
 * 3 entries EM, 1 acoustic), then, when chi_channels is odd, one constant
 * channel `extra` (the wave number feature). */
int em3d_fno_chi_features(int nx, int ny, int nz, int factor, int precision, const void* chi, int chi_kind,
                          int chi_channels, double extra, float* out, hipStream_t stream);
/* out = base + scale[b] P(y_2b + i y_2b+1) (base may be NULL): P is the
 * trilinear prolongation of the preconditioner; y [2 batch][cells][ncomp].
 * em3d_fno_prolong_adjoint: out[2b + p] = P^T (Re / Im g_b) / scale[b]. */
int em3d_fno_prolong(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* base,
                     const double* scale, const float* y, void* out, hipStream_t stream);
int em3d_fno_prolong_adjoint(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* g,
                             const double* scale, float* out, hipStream_t stream);
int em3d_fno_precondition(const em3d_fno_model* model, int batch, int ncomp, int nx, int ny, int nz, int precision,
                          const void* r, void* out, hipStream_t stream);

/* em3d_bicgstab with the FNO as right preconditioner (flexible form: the
 * preconditioned directions M p and M s are kept and enter the update, so
 * the preconditioner may be any fixed nonlinear map).  Same arguments,
 * statuses and histories as em3d_bicgstab; solver_work holds
 * em3d_bicgstab_pc_work_bytes(); the model's work and features are sized for
 * the plan's batch.  No counterpart in the reference. */
size_t em3d_bicgstab_pc_work_bytes(em3d_plan_t plan);
int em3d_bicgstab_pc(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* rhs, void* u,
                     void* work, void* solver_work, const em3d_bicgstab_params* params, double* history,
                     long long* action_counts, int history_cap, em3d_solver_result* result,
                     const em3d_fno_model* precond, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* EM3D_HIP_H */
