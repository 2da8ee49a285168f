/*
 * C ABI of libsgh_mi355x.so: hand-written gfx950 HIP kernels for the spectral
 * Galerkin ETD1 + Picard step of spectral_galerkin_heat.
 *
 * Conventions
 *   - Every array is a caller-owned device pointer to a C-contiguous [z, y, x]
 *     (x fastest) volume or [y, x] surface, in the precision named by `dtype`
 *     (SGH_F64 = double, SGH_F32 = float). Reductions and norms are always double.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream). Every
 *     call is asynchronous and stream-ordered; the caller keeps buffers alive until
 *     the stream has been synchronised.
 *   - Return 0 on success or a negative SGH_ERR_* code; sgh_last_error() gives a
 *     per-thread message. No C++ exception crosses the ABI. No call allocates
 *     device memory except sgh_batch_pool_create / sgh_batch_create (a plan owns its
 *     packed state).
 *
 * Each entry point cites the reference function it replaces (paths relative to
 * the reference checkout; sgh/ = src/spectral_galerkin_heat/).
 */
#ifndef SGH_MI355X_H
#define SGH_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SGH_F64 = 0, SGH_F32 = 1 };

/* Arithmetic of a 3-D axis contraction (sgh_axis_contract):
 *   SGH_XFORM_MFMA   matrix cores in the data type (f32 16x16x4 or f64 16x16x4);
 *   SGH_XFORM_BF16X3 split-BF16: x = hi + lo, three bf16 16x16x32 products
 *                    hi·hi + hi·lo + lo·hi with FP32 accumulation (FP32 data only);
 *   SGH_XFORM_VALU   LDS-tiled vector-ALU GEMM in the data type. */
enum { SGH_XFORM_MFMA = 0, SGH_XFORM_BF16X3 = 1, SGH_XFORM_VALU = 2 };
enum {
  SGH_OK = 0,
  SGH_ERR_ARG = -1,   /* bad shape, null pointer, unsupported dtype           */
  SGH_ERR_HIP = -2,   /* HIP runtime / launch failure                         */
  SGH_ERR_STATE = -3, /* misuse of a plan (wrong order, destroyed handle)     */
  SGH_ERR_NOMEM = -4  /* device allocation failed, or a plan's pool is full    */
};

/* ---- library ------------------------------------------------------------- */
int sgh_version(void);
const char* sgh_last_error(void);

/* ---- propagator tables ---------------------------------------------------
 * E[p,n,m] = exp(-αλΔt), KK = (1 - E)/(αλ)/(ρc) with λ = (πm/Lx)²+(πn/Ly)²+(πp/Lz)²,
 * evaluated in FP64 with expm1 and stored in `dtype`; KK[0,0,0] = Δt/(ρc).
 * Replaces precompute_K_KK (sgh/physics/spectral_state.py:374-421). */
int sgh_etd_tables(int dtype, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                   double alpha, double rho_cp, double dt, void* E, void* KK, void* stream);

/* ---- pointwise modal updates ---------------------------------------------
 * out = aK + KK·(wt[p]·St[n,m] + wb[p]·Sb[n,m] + Q[p,n,m]); (wb,Sb) and Q may be NULL.
 * Replaces update_modes_etd1 (sgh/physics/spectral_cpu_kernels.py:83-91; GPU K1
 * sgh/physics/spectral_gpu_kernels.py:59-62,129-131) fused with the accumulators below. */
int sgh_etd1_update(int dtype, int nx, int ny, int nz, const void* aK, const void* KK,
                    const void* wt, const void* St, const void* wb, const void* Sb,
                    const void* Q, void* out, void* stream);

/* a += KK·Q.  Replaces add_source_term_modes (spectral_cpu_kernels.py:93-103). */
int sgh_add_volume_source(int dtype, long long n, void* a, const void* KK, const void* Q,
                          void* stream);

/* a += KK·w[p]·B[n,m].  Replaces add_bottom_surface_source (spectral_cpu_kernels.py:105-113). */
int sgh_add_surface_source(int dtype, int nx, int ny, int nz, void* a, const void* KK,
                           const void* w, const void* B, void* stream);

/* a *= E.  Replaces xp.multiply(a, K, out=a) (sgh/solvers/spectral.py:283). */
int sgh_mul_inplace(int dtype, long long n, void* a, const void* E, void* stream);

/* parts[c][n,m] = Σ_{p in chunk c} w[p]·a[p,n,m]; the full contraction is the sum of the
 * `nchunks` planes.  Replaces the einsum of reconstruct_surface_temperature /
 * reconstruct_bottom_temperature (sgh/physics/spectral_ops.py:210-231). */
int sgh_zcontract(int dtype, int nx, int ny, int nz, const void* a, const void* w,
                  int nchunks, void* parts, void* stream);

/* parts[0][i] = Σ_q parts[q][i] for `nparts` planes of n elements, in place (the
 * finishing sum of the z-contraction above, so the 2-D transform reads one plane).
 * Part of the same einsum (sgh/physics/spectral_ops.py:210-231). */
int sgh_sum_parts(int dtype, long long n, int nparts, void* parts, void* stream);

/* ---- fused Picard iteration ----------------------------------------------
 * a_raw = aK + KK·(wt⊗St + wb⊗Sb + Q);  r = a_raw - a_old;  a_new = a_old + ω·r;
 * partials[2b] = Σ r², partials[2b+1] = Σ a_old² (FP64, one pair per block);
 * if wz != NULL also zc_parts[c] = Σ_{p in chunk c} wz[p]·a_new[p] (next iterate's
 * surface contraction).  Replaces spectral.py:431-465 (update, accumulators,
 * relaxation and both vdots).  `partials` must hold 2·sgh_picard_partials_count(). */
int sgh_picard_partials_count(int nx, int ny, int nz, int nchunks);
int sgh_picard_iter_fused(int dtype, int nx, int ny, int nz, const void* aK, const void* KK,
                          const void* a_old, const void* wt, const void* St, const void* wb,
                          const void* Sb, const void* Q, double omega, const void* wz,
                          int nchunks, void* zc_parts, void* a_new, double* partials,
                          void* stream);
/* The same pass with the norm reduction of sgh_reduce_norms folded in: the last
 * workgroup to finish (counted on `counter`, a device word that must start at 0 and is
 * left at 0) sums every pair in sgh_reduce_norms' fixed order and writes norms[3]
 * (spectral.py:464-468), so one launch replaces two.  If host_out (pinned, coherent
 * host memory from sgh_host_alloc) is given, it also writes the three norms there and
 * then host_out[3] = seq, so the host can poll for this iteration's residual instead of
 * waiting on the stream (the host read of spectral.py:468). */
int sgh_picard_iter_fused_norms(int dtype, int nx, int ny, int nz, const void* aK,
                                const void* KK, const void* a_old, const void* wt,
                                const void* St, const void* wb, const void* Sb, const void* Q,
                                double omega, const void* wz, int nchunks, void* zc_parts,
                                void* a_new, double* partials, unsigned int* counter,
                                double* norms, double* host_out, double seq, void* stream);

/* The same pass for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. Its scratch (the per-workgroup norm pairs and the last-workgroup counter)
 * is a fixed array of SGH_NOALLOC_MAX_PAIRS pairs (1 MiB) that the library holds on every
 * GPU it runs on until the process ends, so the caller passes only inputs and outputs
 * (norms: 3 device doubles; host_out and seq as above). Refuses
 * sgh_picard_partials_count(nx, ny, nz, nchunks) > SGH_NOALLOC_MAX_PAIRS with
 * SGH_ERR_ARG. Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams, no ordering between them) race and must not be issued. The
 * first call on a device also enqueues a one-thread kernel on `stream` that zeroes the
 * counter; the pass leaves it at zero. Results are bit-identical to
 * sgh_picard_iter_fused_norms with caller buffers. */
#define SGH_NOALLOC_MAX_PAIRS 65536
int sgh_picard_iter_fused_norms_noalloc(int dtype, int nx, int ny, int nz, const void* aK,
                                        const void* KK, const void* a_old, const void* wt,
                                        const void* St, const void* wb, const void* Sb,
                                        const void* Q, double omega, const void* wz,
                                        int nchunks, void* zc_parts, void* a_new,
                                        double* norms, double* host_out, double seq,
                                        void* stream);

/* Pinned, coherent host memory that kernels can write and the host can poll
 * (hipHostMalloc with hipHostMallocCoherent); no reference counterpart. */
int sgh_host_alloc(size_t bytes, void** out);
int sgh_host_free(void* p);

/* out[0] = √(Σr²/N), out[1] = √(Σa²/N), out[2] = out[0]/max(out[1], 1e-9), summed in a
 * fixed order.  Replaces the norm of spectral.py:464-468. */
int sgh_reduce_norms(const double* partials, int n_pairs, double n_elements, double* out,
                     void* stream);

/* out[0] = Σ x (op 0) or max x (op 1) over n elements, in FP64.  Replaces xp.sum /
 * xp.max on the path (spectral.py:278,537). */
int sgh_reduce(int dtype, long long n, const void* x, int op, double* out, void* stream);

/* ---- separable transforms ------------------------------------------------
 * out = scale · L · X · R for X = Σ_{c<nparts} Xparts[c] (ny×nx, plane stride ny·nx),
 * L (ny×ny) and R (nx×nx) row-major.  With L, R the ortho DCT-II matrix or its
 * transpose this is DCT_II / IDCT_II of a surface field (spectral_cpu_kernels.py:374-389;
 * GPU K9/K10 spectral_gpu_kernels.py:371-451); with the cell-centred cosine basis it is
 * the surface reconstruction / projection directly.  `batch` independent transforms
 * share L and R (stride ny·nx·nparts in, ny·nx out). */
int sgh_sep2d(int dtype, int batch, int ny, int nx, const void* L, const void* R,
              const void* X, int nparts, double scale, void* out, void* stream);

/* The inverse surface transform fused with the top-surface forcing: T = scale·L·X·R
 * (one surface), then q_evap = evaporation(T) and q_net = (q_las − q_evap − h_top(T − T0))
 * [· k̄/k(T) with a property program] in the same launch, as sgh_sep2d followed by
 * sgh_surface_forcing (spectral_ops.py:210-219 then spectral.py:392-406).  Surfaces the
 * one-round kernel cannot hold, and forcing with a property program (whose per-cell
 * interpreter is better spread over the separate forcing launch), run those two calls,
 * T landing in T_s (which the fused form leaves untouched). */
int sgh_sep2d_forcing(int dtype, int ny, int nx, const void* L, const void* R, const void* X,
                      double scale, const void* q_las, double P0, double T_boil, double dH,
                      double R_v, double T_liq, double h_top, double T0, const void* prog,
                      double kbar, void* T_s, void* q_evap, void* q_net, void* stream);

/* out[o,k,i] = scale · Σ_j M[k,j]·in[o,j,i] for an [outer, n, inner] view; M is n×n
 * row-major; variant = SGH_XFORM_*.  Three calls (x, y, z) make the 3-D DCT of
 * project_volume / reconstruct_volume (spectral_ops.py:83-102) and of DCT_II / IDCT_II
 * on volumes (spectral_cpu_kernels.py:374-389). in != out. */
int sgh_axis_contract(int dtype, int variant, long long outer, int n, long long inner,
                      const void* M, const void* in, double scale, void* out, void* stream);

/* The same with a rectangular n_out × n_in matrix: out[o,k,i] (k < n_out) =
 * scale · Σ_j M[k,j]·in[o,j,i] (j < n_in).  With the (n+1) × n node bases it evaluates
 * the cosine series at the n+1 nodes of each axis, the node-centred reconstruction of
 * reconstruct_temperature_DCT (sgh/physics/spectral_helpers.py:147-207). */
int sgh_axis_contract_rect(int dtype, int variant, long long outer, int n_in, int n_out,
                           long long inner, const void* M, const void* in, double scale,
                           void* out, void* stream);

/* The same with M read at row stride ldm >= n_in, so a window of columns of a wider
 * basis is used in place.  With the laser-following windows of the fine-box bases
 * (FineMeshState.update, sgh/physics/spectral_state.py:227-235) three calls make
 * reconstruct_temperature_box (sgh/physics/spectral_ops.py:173-179) and three make
 * project_box_to_modes (spectral_ops.py:53-74). */
int sgh_axis_contract_strided(int dtype, int variant, long long outer, int n_in, int n_out,
                              long long inner, const void* M, long long ldm, const void* in,
                              double scale, void* out, void* stream);

/* Box-mode latent source in one pass over the modes: T on the box cells from the modal
 * state a (nz × ny × nx) by three contractions, z (BzT, nz_box × nz), y (ByT, the window
 * rows of the transposed fine y basis, row stride ldy) and x (BxT likewise, stride ldx),
 * with Q = −ρL(f_l(T) − f_l(T_prev))/dt formed in the last contraction's epilogue, so T
 * itself never reaches HBM.  Replaces compute_latent_heat_source
 * (sgh/physics/spectral_ops.py:234-267) with its reconstruct_temperature_box (:173-179).
 * prog: the property program for the blended-ρ form (core/properties.py:375-430), or
 * NULL.  work: nz_box·ny·nx + nz_box·ny_box·nx elements. */
int sgh_box_latent_source(int dtype, int variant, int nx, int ny, int nz, const void* a,
                          int nxb, int nyb, int nzb, const void* BxT, long long ldx,
                          const void* ByT, long long ldy, const void* BzT, const void* T_prev,
                          double T_S, double T_L, double rho, double L_f, double dt,
                          const void* prog, void* work, void* Q, void* stream);

/* Point evaluation of the cosine series, after the x contraction U = a ·ₓ Bx (U is
 * nz × ny × P for P points): out[i] = Σ_p Bz[p][i] Σ_n By[n][i] U[p][n][i] with By
 * (ny × P) and Bz (nz × P) the bases at the points.  With sgh_axis_contract_rect for the
 * x step it replaces reconstruct_temperature_volume_at_points
 * (sgh/physics/spectral_helpers.py:209-255). */
int sgh_points_yz(int dtype, int nz, int ny, int P, const void* U, const void* By,
                  const void* Bz, void* out, void* stream);

/* ---- box-mode history ------------------------------------------------------
 * out[z] = shift(in[z]) for every z slice of an nz × ny × nx box, bilinear,
 * out[o] = in(o - s); outside the slice: edge value (nearest != 0) or zero.  The
 * T_prev (nearest) and Q_prev (constant) shifts of shift_latent_heat_history
 * (sgh/physics/spectral_ops.py:292-306, _shift_xy_3d spectral_cpu_kernels.py:302-322).
 * in != out. */
int sgh_shift3d_xy(int dtype, int nz, int ny, int nx, const void* in, double sy, double sx,
                   int nearest, void* out, void* stream);

/* Peaks of |Q| over a box for the truncation warning of
 * _check_latent_box_truncation (sgh/solvers/spectral.py:545-586): out[0] = max |Q|,
 * out[1] = max over the floor Q[0], out[2] = max over the x faces Q[:,:,0|nx-1],
 * out[3] = max over the y faces Q[:,0|ny-1,:] (FP64).  out holds 4·(nz + 1) doubles:
 * one partial row per z slice follows the result. */
int sgh_box_face_peaks(int dtype, int nz, int ny, int nx, const void* Q, double* out,
                       void* stream);

/* ---- surface forcing -----------------------------------------------------
 * Laser flux on cell centres x_i = (i+½)dx, y_j = (j+½)dy: I0·exp(-2ρⁿ) (order n), or
 * the cell-mean Gaussian when cell_integrated != 0.  Replaces LaserProfile.flux
 * (sgh/core/laser.py:42-53,65-90,122-135). */
int sgh_laser_flux(int dtype, int nx, int ny, double dx, double dy, double x0, double y0,
                   double rx, double ry, double order, double I0, int cell_integrated,
                   void* out, void* stream);

/* out = base + alpha · shift(in), bilinear, out[o] = in(o - s), zero outside
 * (scipy.ndimage.shift order=1 mode='constant'); base may be NULL.
 * Replaces shift_flux / _ndshift (spectral_cpu_kernels.py:281-396). */
int sgh_shift2d(int dtype, int ny, int nx, const void* in, double sy, double sx,
                const void* base, double alpha, void* out, void* stream);

/* q_evap = evaporation(T) (0 below T_liq);
 * q_net = (q_las - q_evap - h_top·(T - T0)) · (prog ? kbar/k(T) : 1).  q_las may be NULL.
 * Replaces compute_evaporation_flux (spectral_cpu_kernels.py:144-162; GPU K4) and the
 * top-flux assembly of spectral.py:390-407. `prog` is a packed property program. */
int sgh_surface_forcing(int dtype, long long S, const void* T, const void* q_las,
                        double P0, double T_boil, double dH, double R_v, double T_liq,
                        double h_top, double T0, const void* prog, double kbar,
                        void* q_evap, void* q_net, void* stream);

/* out = -h·(T - T0)·(prog ? kbar/k(T) : 1).  Convection flux (spectral.py:341,348,376-378). */
int sgh_convection_flux(int dtype, long long S, const void* T, double h, double T0,
                        const void* prog, double kbar, void* out, void* stream);

/* ---- volume forcing ------------------------------------------------------
 * Q = -ρ·L·(f_l(T) - f_l(T_prev))/dt, f_l clamped; ρ = rho, or the blended ρ(T) of
 * `prog` when prog != NULL.  Replaces compute_source_term_from_temperature
 * (spectral_cpu_kernels.py:115-140; GPU K5) and MaterialModel.latent_heat_source
 * (sgh/core/properties.py:375-430; GPU K7). */
int sgh_latent_source(int dtype, long long n, const void* T, const void* T_prev, double T_S,
                      double T_L, double rho, double L, double dt, const void* prog,
                      void* out, void* stream);

/* f = ∇·(k'∇T) - a'(T - T_prev)/dt with k' = k(T) - kbar, a' = ρ(T)c(T) - abar from
 * `prog`, numpy-gradient stencil (central inside, one-sided on faces).  `work` holds
 * 3·nz·ny·nx elements.  Replaces MaterialModel.k_prime_a_prime (properties.py:318-373;
 * GPU K6) + _correction_source (spectral_cpu_kernels.py:165-252; GPU K8). */
int sgh_property_correction_source(int dtype, int nx, int ny, int nz, const void* T,
                                   const void* T_prev, const void* prog, double kbar,
                                   double abar, double dt, double dx, double dy, double dz,
                                   void* work, void* out, void* stream);

/* Property program layout size in bytes (for validation on the host). */
int sgh_property_program_bytes(void);

/* ---- batched surface-collapsed engine ------------------------------------
 * Many independent cuboids with top-surface forcing only (laser + evaporation,
 * constant properties, no latent heat, no convection) advance together, each with the
 * reference's step semantics (sgh/solvers/spectral.py:219-543): warm start from the
 * shifted previous evaporation flux, relaxed Picard loop with its own ω, tol and cap,
 * one step per sgh_batch_run iteration. Every iterate is aE + KK·(w⊗G), so the loop
 * runs on surface-sized arrays (exact norm identities) and one volume pass per step
 * writes the state; E and KK are regenerated in FP64 from 1-D tables.
 * The laser state for every step comes from a host table copied at creation, so a
 * step replays as one captured hipGraph without parameter updates. */
typedef struct sgh_cuboid_spec {
  int nx, ny, nz;         /* modes; nx, ny <= 256                                  */
  int max_it;             /* Picard cap                                            */
  int n_steps;            /* rows of `laser`; the cuboid stops after that many steps */
  int cell_integrated;    /* Gaussian cell-mean flux (order 2 only)                */
  double Lx, Ly, Lz;      /* extents (m)                                           */
  double k, rho, cp;      /* constant conductivity, density, heat capacity         */
  double Pa, T_boil, dH, R_v, T_liq;  /* evaporation law                          */
  double absorptivity, rx, ry, order; /* beam                                     */
  double dt, omega, tol;  /* time step, relaxation, Picard tolerance               */
  const double* laser;    /* host: n_steps × {x, y, power, on, vx, vy}             */
} sgh_cuboid_spec;

/* A plan is a pool of cuboid slots over fixed-capacity packed device arrays (the only
 * allocating calls are the two constructors). Capacities: state elements, surface
 * elements, laser-table rows, largest nz; `metrics_rows` steps of per-cuboid metrics
 * {P_laser, T_surface_max, n_it, converged, rel_err} are kept per slot.
 * Replaces a loop of SpectralSolver.step calls over independent runs. */
int sgh_batch_pool_create(int dtype, int max_cuboids, long long state_capacity,
                          long long surface_capacity, long long laser_capacity, int max_nz,
                          int metrics_rows, void** plan);
/* Pool sized exactly for `n` cuboids, all added (slots 0..n-1). */
int sgh_batch_create(int dtype, int n, const sgh_cuboid_spec* specs, int metrics_rows,
                     void** plan, void* stream);
/* Continuous batching: place a cuboid in a free slot (state zero, step 0; copies the
 * laser table) or free a slot. Both synchronise the stream. Returns SGH_ERR_NOMEM when
 * the pool has no room (the caller keeps the request queued), SGH_ERR_ARG for a spec the
 * pool can never hold (nz above the pool's max_nz, a surface axis above 256). The cosine
 * matrices of a mode count are shared by the cuboids that use it and released with the
 * last of them. */
int sgh_batch_add(void* plan, const sgh_cuboid_spec* spec, int* slot, void* stream);
int sgh_batch_remove(void* plan, int slot, void* stream);
/* Distinct surface mode counts whose matrices the pool holds now: writes up to
 * `capacity` of them to out and returns how many there are. */
int sgh_batch_matrix_sizes(void* plan, int* out, int capacity);
int sgh_batch_destroy(void* plan);
/* Advance every live cuboid by n_steps steps (one captured graph per step if use_graph). */
int sgh_batch_run(void* plan, int n_steps, int use_graph, void* stream);
/* Device-to-device copy of cuboid c's modal state [nz, ny, nx] in the plan's dtype. */
int sgh_batch_set_state(void* plan, int c, const void* src, void* stream);
int sgh_batch_get_state(void* plan, int c, void* dst, void* stream);
int sgh_batch_state_layout(void* plan, int c, long long* a_offset, int* nx, int* ny, int* nz);
/* Copies rows × 5 metrics of slot c to host memory (synchronises the stream). */
int sgh_batch_read_metrics(void* plan, int c, double* host_out, int rows, void* stream);
/* Per-slot step counters (max_cuboids ints) to host memory (synchronises the stream). */
int sgh_batch_steps_done(void* plan, int* host_out, void* stream);
/* Event timing of the surface-iteration and volume kernels (disables graphs while on):
 * out4 = {surface ms, surface launches, volume ms, volume launches}. */
int sgh_batch_set_timing(void* plan, int enable);
int sgh_batch_timing(void* plan, double* out4);
/* out6 = {live state elements, surface capacity, surface items, volume items,
 *         max Picard cap, LDS bytes per surface workgroup}. */
int sgh_batch_info(void* plan, long long* out6);

/* ---- autotuning knobs of the batched surface transform (no reference counterpart;
 * the reference has no batched engine).  A variant is (rows per work item, threads per
 * workgroup) plus the dynamic LDS bytes per workgroup (4 KiB to 159 KiB); the default is
 * (16, 1024, 150 KiB).  sgh_batch_variants writes up to `capacity` (rows, threads) pairs
 * and returns how many exist.  sgh_batch_set_variant synchronises `stream`; the work
 * lists and the captured step graph are rebuilt by the next sgh_batch_run.  Results
 * are deterministic for a given variant; different variants sum per-item partials in
 * different groupings (last-bit differences).  rows = -1 (threads 512) is the resident
 * step kernel: FP32 plans only (refused for FP64), a team of 1-8 workgroups per cuboid
 * runs the whole surface part of a step in one launch with split-bf16 MFMA transforms;
 * cuboids with a surface side wider than 128 run on 16-row per-pass items with the given
 * LDS budget in the same step (beside the resident launch, on a second stream the plan
 * owns), so every cuboid's results are those of its run alone.  sgh_batch_get_variant
 * reports rows -1 while it is selected. */
int sgh_batch_variants(int* out_pairs, int capacity);
int sgh_batch_set_variant(void* plan, int rows, int threads, int lds_bytes, void* stream);
/* out3 = {rows, threads, LDS bytes}. */
int sgh_batch_get_variant(void* plan, int* out3);
/* Persistent step kernel (no reference counterpart; the reference has no batched engine).
 * With enable = 1 each step's surface work (warm start, every Picard iteration, step end)
 * runs as ONE launch in which the work items of a cuboid meet at device-scope barriers
 * between passes, instead of 2·cap + 2 launches.  It is used while the work list fits one
 * workgroup per CU and the variant has it (rows 16 with 512 or 1024 threads, LDS ≤ 158 KiB);
 * otherwise the per-pass launches run.  The iterates are identical to the per-pass form's.
 * A barrier wait is bounded: if one times out, sgh_batch_steps_done and
 * sgh_batch_read_metrics return an error.  Synchronises `stream`; the work lists and step
 * graph are rebuilt by the next run.  out2 = {requested, in use by the last work lists}.
 * Currently switched off: enable = 1 returns SGH_ERR_STATE (a GPU memory fault in a run that
 * used the kernel is not yet explained). */
int sgh_batch_set_loop(void* plan, int enable, void* stream);
/* Bound of one barrier wait of the persistent step kernel, in seconds (default 0.02; 0 makes
 * any wait time out at once, which tests use to exercise the error path). */
int sgh_batch_set_loop_timeout(void* plan, double seconds, void* stream);
int sgh_batch_loop_state(void* plan, int* out2);
/* Resident teams (no reference counterpart; the reference has no batched engine).  With
 * the resident step kernel, a cuboid's step runs on a team of 1 to team_max workgroups
 * (1..8, default 8; at most one per 16 rows of the surface) that split every stage by rows
 * and exchange the transformed fields through device memory twice per Picard iteration.
 * The sizes are chosen per work list so that the longest team step is as short as the CUs
 * allow (teams only while the batch fits one workgroup per CU).  Every sum is formed per
 * tile and combined in a fixed order, so results are bit-identical for every team size.
 * A team barrier's wait is bounded by the sgh_batch_set_loop_timeout bound and reported
 * like the persistent kernel's.  sgh_batch_set_team_max synchronises `stream`; the work
 * lists and step graph are rebuilt by the next run.  sgh_batch_team_sizes writes the team
 * size of each slot in the last work lists (0: free slot, or the resident kernel not in
 * use) to out[0 .. capacity) and returns the slot count. */
int sgh_batch_set_team_max(void* plan, int team_max, void* stream);
int sgh_batch_team_sizes(void* plan, int* out, int capacity);
/* Step counters, Picard status and surface fields of every live cuboid back to zero
 * (synchronises `stream`); the modal states are kept: set them with
 * sgh_batch_set_state before the next run to restart from an initial condition. */
int sgh_batch_rewind(void* plan, void* stream);

/* ---- one Picard iteration of the plain path in one call ----------------------
 * Enqueues, in the order of SpectralSolver.step's loop body (sgh/solvers/spectral.py:
 * 384-497), every kernel of one relaxed-Picard iteration on one cuboid: top-surface
 * temperature (z-partial sum + 2-D inverse), evaporation / convection / k̄/k(T) flux
 * forcing, its projection, the latent source (coarse grid: reconstruct, source, project;
 * fine box: the fused sgh_box_latent_source then the window projection), the property
 * correction when `rescale`, and the fused ETD1 + relaxation + norm pass whose last
 * workgroup writes the norms to norms_host[4·parity ..] followed by `seq` (pinned,
 * coherent host memory from sgh_host_alloc), which the host polls.  Buffers indexed by
 * `parity` (= iteration & 1) alternate so that the next iteration can be enqueued
 * before this one's residual is read: a_old = a_buf[parity], a_new = a_buf[parity ^ 1];
 * q_evap, Q_latent, Q_modes and norms_host by parity.  Every pointer is a device pointer
 * except norms_host, which both sides address. */
typedef struct SghPlainIter {
  int dtype, xform, nx, ny, nz, nchunks, n_pairs;
  int latent_mode;     /* 0 none, 1 coarse grid, 2 fine box */
  int rescale;         /* temperature-dependent properties */
  int have_bottom;     /* lagged bottom-convection source S_bot_raw */
  int nxb, nyb, nzb;   /* fine box */
  int ntx, nty;        /* fine-grid cells along x and y (row strides of the x/y bases) */
  int x_start, y_start;/* box window of this step */
  int pad0;
  double dxdy, dV, dV_fine, n_elements, omega;
  double Pa, T_boil, dH, R_v, T_liq, h_top, T0, kbar, abar;
  double T_S, T_L, rho, L_f, dt, dx, dy, dz;
  const void* aK; const void* KK; void* a_buf[2];
  const void* w_top; const void* w_bot;
  void* zc; void* T_s; const void* q_las; void* q_evap[2]; void* q_tmp; void* S_top;
  const void* S_zero; const void* S_bot_raw;
  const void* sByT; const void* sBx; const void* sBy; const void* sBxT;   /* surface bases */
  void* Q_latent[2]; void* Q_modes[2]; const void* T_prev_full; const void* prog;
  const void* vBxT; const void* vByT; const void* vBzT;   /* volume reconstruction */
  const void* vBx; const void* vBy; const void* vBz;      /* volume projection */
  void* vol_w1; void* vol_w2; void* vol_T;                /* N elements each */
  void* a_raw; void* corr_work; void* corr_f; void* C_modes;   /* corr_work: 3N */
  const void* fBxT; const void* fByT; const void* fBzT;   /* fine box bases */
  const void* fBx; const void* fBy; const void* fBz;
  const void* T_prev_box; void* box_work;  /* box_work: 2·nzb·max(ny, nyb)·nx */
  double* partials; double* norms; double* norms_host;
  unsigned int* counter;   /* device word at 0: the fused pass's last-workgroup count */
  double seq;              /* written after the norms; the host waits for it */
  void* side_stream;       /* NULL, or from sgh_plain_iter_fork_init */
  void* ev_fork; void* ev_join;
} SghPlainIter;

/* sizeof(SghPlainIter), so a binding can check its mirror of the layout. */
size_t sgh_plain_iter_size(void);
int sgh_plain_iterate(const SghPlainIter* d, int parity, void* stream);
/* A side stream (default priority) and two events for sgh_plain_iterate (no reference
 * counterpart): with them, an iteration that has a latent source enqueues its surface
 * chain (z-partial sum, 2-D inverse with the flux forcing, 2-D forward) on the side stream,
 * forked from `stream` at the iteration's start, while the latent source's reconstruction,
 * source and projection go on `stream`; `stream` waits for the side chain before the
 * ETD1 update that needs both.  The kernels and their inputs are those of the serial
 * order, so results are bit-identical.  It pays for a cuboid that has the GPU to itself
 * (measured: 128x64x32 with latent heat and T-dependent properties 3.44 -> 3.20 ms/step in
 * FP64, an 80x48x120 box-mode track 4.05 -> 3.55); cuboids stepped concurrently on
 * streams of their own run slower with it (the cross-stream waits stall the hardware
 * queues they share), as do small ones.  sgh_plain_iter_fork_release destroys them (after
 * the caller has synchronised) and sets the fields to NULL. */
int sgh_plain_iter_fork_init(SghPlainIter* d);
int sgh_plain_iter_fork_release(SghPlainIter* d);

#ifdef __cplusplus
}
#endif

#endif /* SGH_MI355X_H */
