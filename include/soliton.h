/*
 * C-ABI of libsoliton.so: the MI355X relaxation core of soliton_solver.
 *
 * The reference reaches its Numba-CUDA kernels from Python (Simulation in
 * core/simulation.py, theory kernels/observables modules); this library is the
 * native side of that same boundary.  Plain pointers and sizes only; every call
 * returns 0 on success and a non-zero status on failure, with the message in
 * ss_last_error(ctx) (or ss_global_error() when there is no context yet).
 * No exception crosses the ABI.  One context = one relaxation on one GPU, one
 * HIP stream; calls on one context must come from one host thread.
 *
 * Host arrays use the reference's flat layout (core/utils.py:16-42):
 * idx(a, x, y) = a*xlen*ylen + x*ylen + y, FP64.
 */
#ifndef SOLITON_H
#define SOLITON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ss_ctx ss_ctx;

/* Theory ids (registry names of soliton_solver/theories/<name>/__init__.py). */
enum {
  SS_THEORY_CHIRAL_MAGNET = 1,
  SS_THEORY_GINZBURG_LANDAU = 2,
  SS_THEORY_BABY_SKYRME = 3,
  SS_THEORY_ANISOTROPIC_SC = 4,
  SS_THEORY_BEC = 5,
  SS_THEORY_MAXWELL_CHERN_SIMONS_HIGGS = 6,
  SS_THEORY_FERROMAGNETIC_SC = 7,
  SS_THEORY_SPIN_TRIPLET = 8,
  SS_THEORY_LIQUID_CRYSTAL = 9
};

/* Storage / stencil precision.  Energy and charge sums and the arrest test are
 * FP64 in all three.
 * FP32C: FP32 storage and stencils as FP32, with a second FP32 word per field
 *   value (the field is hi + lo, nf*N more floats): the combine of each step
 *   adds its increment to hi + lo in FP64 and splits the sum again, so steps
 *   that move the field by less than an FP32 ulp accumulate instead of being
 *   rounded away.  Stencils, energies and charges read hi.  Reference flow of
 *   theories without the L2 normalisation; not in the fused 4+1 sweep; not
 *   in batches.  ss_upload / ss_download of the field split / join hi + lo;
 *   ss_init_field and ss_add_noise set lo to zero. */
enum { SS_PREC_FP64 = 0, SS_PREC_FP32 = 1, SS_PREC_FP32C = 2 };

/* Variation used by the flow and by ss_variation:
 * REFERENCE: the reference's Euler-Lagrange expression (theory
 *   kernels.do_gradient_step_point), masks and projections included;
 * EXACT_ADJOINT: the exact gradient dE_h/dphi of the discrete energy sum. */
enum { SS_VAR_REFERENCE = 0, SS_VAR_EXACT_ADJOINT = 1 };

/* Buffers for ss_upload / ss_download. */
enum { SS_BUF_FIELD = 0, SS_BUF_VELOCITY = 1 };

/* Library version (major*10000 + minor*100 + patch). */
int ss_version(void);

/* Message of the last failure that had no context (ss_ctx_create). */
const char* ss_global_error(void);

/* Number of GPUs visible to HIP (0 when none); never fails. */
int ss_device_count(void);

/* Replaces Simulation.__init__ buffer allocation (core/simulation.py:43-115)
 * and set_params (:144-166).  p_i / p_f are the packed arrays of the theory's
 * pack_device_params (e.g. chiral_magnet/params.py:333-378). */
int ss_ctx_create(int theory_id, const int32_t* p_i, int n_i, const double* p_f, int n_f,
                  int precision, int variation_mode, ss_ctx** out);
int ss_ctx_destroy(ss_ctx* ctx);

/* Simulation.set_params (core/simulation.py:144-166).  The lattice shape
 * (xlen, ylen, number_total_fields) must not change. */
int ss_set_params(ss_ctx* ctx, const int32_t* p_i, int n_i, const double* p_f, int n_f);

/* Switch storage precision in place (FP32 relaxation -> FP64 polish).
 * FP64 -> FP32C splits the field exactly into hi + lo; FP32C -> FP64 joins
 * it; FP32 -> FP32C starts lo at zero; FP32C -> FP32 rounds hi + lo to FP32. */
int ss_set_precision(ss_ctx* ctx, int precision);
int ss_set_variation_mode(ss_ctx* ctx, int variation_mode);

/* Copies of Field / Velocity between the host (FP64, reference layout) and
 * the device.  n must equal number_total_fields*xlen*ylen. */
int ss_upload(ss_ctx* ctx, int buf, const double* host, size_t n);
int ss_download(ss_ctx* ctx, int buf, double* host, size_t n);

/* nsteps x Simulation.step (core/simulation.py:388-409 ->
 * core/integrator.py:217-348): step i compares its new energy with the
 * previous step's (the first with prev_energy) using strict '>', and kills the
 * velocity when killkinen is set.  Returns the last step's energy, the signed
 * max of its stage-4 energy gradient (the reference's err, utils.py:397-420)
 * and max |stage-4 energy gradient|.  Any output pointer may be NULL.
 * Non-finite values are not dropped: a NaN or infinite energy gradient at any
 * site makes err_abs_max NaN or +inf (the error reductions propagate NaN, where
 * the reference's tree of '<' comparisons and numpy max usually do), so a
 * blown-up relaxation never compares below a tolerance; a non-finite field
 * also gives a non-finite e_new. */
int ss_step(ss_ctx* ctx, double prev_energy, int nsteps, double* e_new, double* err_signed_max,
            double* err_abs_max);

/* Per-step energies and errs of the last ss_step call (n <= its nsteps). */
int ss_step_history(ss_ctx* ctx, double* energies, double* errs, int n);

/* theory observables.compute_energy (e.g. chiral_magnet/observables.py:15-57). */
int ss_energy(ss_ctx* ctx, double* e);

/* Charges of the theory: kind indexes the theory's charge list (chiral magnet:
 * 0 skyrmion number, 1 magnetic charge; Ginzburg-Landau: 0 vortex number ...).
 * The reference's compute_vortex_number(..., which, ...) has no ``which``
 * here: each flux component it selects is a charge kind of its own.
 * Ferromagnetic superconductor (ferromagnetic_superconductor/kernels.py:291-367)
 * and spin-triplet superconducting ferromagnet
 * (spin_triplet_superconducting_magnet/kernels.py:301-377): kind 0 skyrmion
 * number, kind 1 = which 2 (d_x A2 - d_y A1), kind 2 = which 0 (d_y A3),
 * kind 3 = which 1 (-d_x A3).  Ginzburg-Landau, anisotropic superconductor and
 * Maxwell-Chern-Simons-Higgs ignore ``which`` in the reference (one flux
 * component): their vortex number is kind 0. */
int ss_charge(ss_ctx* ctx, int kind, double* q);
int ss_num_charges(ss_ctx* ctx);

/* Per-site density times grid volume (kind -1: energy, >= 0: charge kind),
 * n = xlen*ylen values in idx_field(0, x, y) order
 * (Simulation.compute_energy_density, core/simulation.py:194-209). */
int ss_density(ss_ctx* ctx, int kind, double* host_out, size_t n);

/* Initial conditions generated on the device (the reference's
 * create_ground_state_kernel / create_initial_configuration_kernel /
 * create_skyrmion_kernel / create_vortex_kernel, e.g.
 * chiral_magnet/initial_config.py:117-386): mode 0 ground state, 1 the
 * theory's centred ansatz, 2 solitons (x, y, radius) composed by stereographic
 * addition (unit-vector theories), 3 vortex psi = u tanh(scale r) e^{i N th}
 * with A = 0 (Ginzburg-Landau).  extra[] carries theory-specific values (BEC:
 * Thomas-Fermi radius, beta, vortex sign, core size; vortex seeds in
 * solitons).  Sets Field, zeroes Velocity; fails for a mode the theory does
 * not provide.  ss_add_noise adds U(-a, a) (or N(0, a^2)) from a
 * counter-based generator of ``seed`` to components [begin, end) and then
 * rescales the first renorm_ncomp components to norm renorm_norm (0: none);
 * its stream is not NumPy's, so fields meant to match a host-seeded oracle
 * are generated on the host and uploaded. */
typedef struct ss_init_desc {
  int mode;
  int ansatz;              /* 0 Bloch, 1 Neel, 2 anti */
  double skyrmion_number;
  double rotation;
  double scale;
  double vortex_number;
  double extra[4];
  int nsol;
  const double* solitons;  /* 3 * nsol values */
} ss_init_desc;
int ss_init_field(ss_ctx* ctx, const ss_init_desc* desc);
int ss_add_noise(ss_ctx* ctx, int comp_begin, int comp_end, int gaussian, double amplitude, uint64_t seed,
                 int renorm_ncomp, double renorm_norm);

/* Auxiliary per-site quantities of the reference's Simulation
 * (core/simulation.py:211-386; theory kernels compute_norm_higgs*_point,
 * compute_phase_difference_point, compute_supercurrent_point,
 * compute_magnetic_field_kernel): |psi|^2 (HIGGS, or HIGGS1/HIGGS2 for two
 * order parameters), cos of the phase difference, the supercurrent (2 or 3
 * components) and the magnetic field (1 or 3 components, each the flux
 * density times gv/(2 pi q) as the reference writes it).  ss_aux_components
 * is 0 where the theory does not define the kind.  n = xlen*ylen values in
 * idx_field(0, x, y) order, FP64, no grid-volume factor unless stated. */
enum { SS_AUX_HIGGS = 0, SS_AUX_HIGGS1 = 1, SS_AUX_HIGGS2 = 2, SS_AUX_PHASE_DIFFERENCE = 3,
       SS_AUX_SUPERCURRENT = 4, SS_AUX_MAGNETIC_FIELD = 5 };
int ss_aux_components(ss_ctx* ctx, int kind);
int ss_aux_density(ss_ctx* ctx, int kind, int comp, double* host_out, size_t n);

/* Energy gradient EG of the current Field (the gradient kernel with
 * Temp = Field, core/integrator.py:15-48), FP64, reference layout. */
int ss_variation(ss_ctx* ctx, int mode, double* host_eg, size_t n);

/* Device-pointer variation for the autograd boundary: field_dev and grad_dev
 * hold nf*xlen*ylen FP64 values in reference layout (the context must be in
 * SS_PREC_FP64; other precisions fail); runs on ``stream`` (a hipStream_t;
 * NULL is the null stream, as for any HIP call).  These calls use the
 * context's scratch and partial buffers, so each one first makes ``stream``
 * wait for the context's stream and then makes the context's stream wait for
 * the call: they are ordered against ss_step and against each other, also
 * when issued from different streams. */
int ss_variation_dev(ss_ctx* ctx, const void* field_dev, void* grad_dev, int mode, void* stream);

/* Device-pointer energy: writes the FP64 total into energy_dev[0]. */
int ss_energy_dev(ss_ctx* ctx, const void* field_dev, double* energy_dev, void* stream);

/* Parameter derivatives of the discrete energy at a fixed FP64 field (device
 * pointer, reference layout): grad_dev[j] = dE_h/dp_f[j] for the theory's
 * packed coupling constants p_f[6 ..] (n = the packed p_f length; entries 0..5,
 * the lattice geometry, are written as 0).  Exact: forward-mode dual numbers
 * seeded at p_f[j] run through the theory's own packed-parameter derivation
 * and energy density (anisotropic_superconductor/kernels.py:96-170 with the
 * vacuum constant of :129-168), summed in FP64.  The derivative with respect
 * to a user parameter chains through the host's ResolvedParams derivation
 * (e.g. anisotropic_superconductor/params.py:190-243: u1, u2, theta_diff,
 * beta3tilde).  Available for the anisotropic superconductor; other theories
 * fail (-1).  FP64 contexts only; ordered against the context's stream like
 * ss_energy_dev; allocates its partials on first use.  (The reference has no
 * parameter derivatives: it is the backward of the differentiable energy.) */
int ss_energy_param_grad(ss_ctx* ctx, const void* field_dev, double* grad_dev, int n, void* stream);

/* ss_energy_dev without any allocation and without the context's scratch:
 * same result (the same kernels and summation order), with the per-tile
 * energy partials in a library-held static device array of 2 MiB (262144
 * doubles per GPU, held until the process ends) instead of the context's
 * partials.  No counterpart in the reference; it is for callers that cannot
 * allocate: code under HIP graph capture and allocation-free serving loops.
 * Calls no hipMalloc / hipFree and creates no event; the array's device
 * address is resolved once per GPU by ss_ctx_create.  Fails (-1) for a
 * context not in SS_PREC_FP64 and for lattices that need more than 262144
 * partials (about 8192 x 8192 sites in FP64 and beyond).  Every call on one
 * GPU shares that GPU's array and it keeps the last call's partials: calls
 * that overlap on one GPU (concurrent streams, threads) race; calls ordered on
 * one stream do not.  It does not touch the context's device memory, so it
 * may overlap with ss_step on the same context. */
int ss_energy_dev_noalloc(ss_ctx* ctx, const void* field_dev, double* energy_dev, void* stream);

/* Capture steps_per_graph steps (rounded up to even) into one hipGraph and
 * replay it inside ss_step; 0 disables graphs. */
int ss_use_graphs(ss_ctx* ctx, int steps_per_graph);

/* Set up ahead of time what the first ss_step(ctx, _, nsteps) would otherwise
 * build inside the call: the step-history buffer for nsteps, the scratch of
 * the exact-adjoint and L2-normalised flows, and the step graphs of both
 * buffer parities (captured, instantiated and uploaded, not run).  Runs no
 * step; the relaxation state is unchanged.  (No counterpart in the
 * reference, which launches every kernel eagerly.) */
int ss_prepare_graphs(ss_ctx* ctx, int nsteps);

/* Wait for all work of the context. */
int ss_synchronize(ss_ctx* ctx);

/* Eager steps with a HIP event around every launch (on the context's
 * stream); avg_ms[0..5] = mean duration of pass 1, energy finalize, pass 2
 * (or 2+3), pass 3, pass 4 (or 4+1), error finalize (with 4+1: and the energy
 * finalize).  With fused 4+1 only the first step of a call runs pass 1.  State
 * advances like ss_step. */
int ss_time_passes(ss_ctx* ctx, double prev_energy, int nsteps, double* avg_ms);

/* Launch shapes of the four fused passes (no counterpart in the reference,
 * whose kernels use fixed 8x8/16x16 thread blocks).  Variants 0-2 stage one
 * LDS tile per workgroup: the largest that fits ~48 KiB, then its rows halved
 * and quartered (less LDS, more resident workgroups per CU).  Variants 3-5
 * march a workgroup down a strip of 64, 128 or 256 rows, 64 columns wide,
 * through an LDS ring with register prefetch (rows fetched once per strip);
 * variants 6-8 are the same strips 128 columns wide (half the y-halo reads
 * per site; where a theory's rings do not fit that wide, they run as 3-5).
 * ss_autotune_tiles times every pass in every variant ``reps`` times on the
 * live buffers, keeps the fastest per pass, and reports chosen[0..3] and
 * ms[9*(pass-1)+variant] (ms holds 72 entries, see ss_set_fused_stages;
 * variants a pass does not run in read 0);
 * it writes only scratch buffers, so the relaxation state is unchanged.
 * ss_get_tile_variants reports rows per tile or strip.  Fields after a step
 * do not depend on the variants; energies only through FP64 summation order. */
int ss_set_tile_variants(ss_ctx* ctx, const int* variants);
int ss_get_tile_variants(ss_ctx* ctx, int* variants, int* rows);
int ss_autotune_tiles(ss_ctx* ctx, int reps, int* chosen, double* ms);

/* Fused sweeps (no counterpart in the reference).  SS_FUSE_23: passes 2 and
 * 3 of a step read the same inputs (F, V, l1), so one sweep builds both stage
 * fields from one set of loads and writes l2 and l3 (as strips, or as LDS
 * tiles holding both fields where such a tile fits).  SS_FUSE_41:
 * the field after a step, F' = Temp4 + dt (l1 - 2 l2 + l3)/6 (then the unit
 * constraint), is pointwise in the inputs of pass 4, so one sweep builds Temp4
 * and F' from one set of loads, finishes the step (F', V', err partials) and
 * runs pass 1 of the next step on F' (its l1 and energy), followed by one
 * finalize launch for both; a call then runs pass 1 only in its first step,
 * and l1 alternates between two buffers.
 * Both sweeps also run as LDS tiles holding two fields where such a tile
 * fits.  ss_autotune_tiles times each sweep in every variant it runs in
 * (ms[36..44] for 2+3, ms[45..53] for 4+1, variants 0..8) and switches each
 * on when it beats the best
 * pair of separate passes; ss_set_fused_stages forces them (a fused sweep uses
 * the variant of pass 2 or pass 4 where it runs in it, else variant 3) and
 * fails when neither fits for the theory and precision.  Results equal the
 * separate passes' up to the FP64 summation order of the energy.
 * SS_FUSE_PAIRS (set alone; strip variants): the
 * stage field of pass 2, F + dt V/2, does not depend on l1, so a step runs as
 * two sweeps, passes 1+2 (F, V in; l1, l2 and the energy partials out) and
 * passes 3+4 (F, V, l1, l2 in; Temp3 and Temp4 in one ring, l3 kept in
 * registers; F', V' and err partials out): 10 words per site and component
 * instead of 14.  Passes 1+2 read V before the arrest decision; after an
 * arrest passes 3+4 take l2 = l1, the l2 of V = 0.  Pass 1+2 uses the strip
 * variant of pass 1, pass 3+4 that of pass 4 (variant 3 where the one set
 * does not run the sweep).  In the compensated form the 1+2 sweep also
 * carries the low words (a three-part ring: F, Temp2, lo; hi, lo, V in; l1
 * with the energy of hi + lo, and l2 out) and the 3+4 sweep ends with the
 * compensated combine: 5 + 8 = 13 words instead of 17; where the three-part
 * ring does not fit the theory, passes 1 and 2 run separately before the 3+4
 * sweep (3 + 3 + 8 = 14 words), and ss_get_fused_stages reports
 * SS_FUSE_COMP12 only when the single 1+2 sweep runs (setting it has no
 * effect).  ss_autotune_tiles times both sweeps in every
 * strip variant (ms[54..62] for 1+2, ms[63..71] for 3+4) and switches the
 * paired schedule on when its two sweeps beat the schedule it chose
 * otherwise by 3 %. */
#define SS_FUSE_23 1
#define SS_FUSE_41 2
#define SS_FUSE_PAIRS 4
#define SS_FUSE_COMP12 8
int ss_set_fused_stages(ss_ctx* ctx, int flags);
int ss_get_fused_stages(ss_ctx* ctx, int* flags);

/* Device memory held by the context, in bytes. */
int ss_memory_bytes(ss_ctx* ctx, size_t* bytes);

const char* ss_last_error(ss_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Batched form: many independent relaxations of one theory (each with its own
 * lattice shape, spacing, time step and couplings) advanced together, one
 * kernel launch per RK4 pass over the tiles of every active instance.  The
 * reference drives one Simulation per lattice (core/simulation.py:43-409);
 * ss_batch replaces a Python loop over such Simulations.  Each row of p_i_all
 * / p_f_all is one instance's pack_device_params output (all rows the same
 * length). */
typedef struct ss_batch ss_batch;

int ss_batch_create(int theory_id, int ninst, const int32_t* p_i_all, int n_i, const double* p_f_all, int n_f,
                    int precision, ss_batch** out);
int ss_batch_destroy(ss_batch* b);
const char* ss_batch_last_error(ss_batch* b);
int ss_batch_num_tiles(ss_batch* b);

/* Field / Velocity of one instance (FP64 host, reference layout, n = nf*xlen*ylen). */
int ss_batch_upload(ss_batch* b, int inst, int buf, const double* host, size_t n);
int ss_batch_download(ss_batch* b, int inst, int buf, double* host, size_t n);

/* Switch the batch's storage precision in place (SS_PREC_FP32 <-> SS_PREC_FP64):
 * fields and velocities of every instance are converted, the per-instance
 * arrest state is kept; the shape variant and schedule stay (a fused sweep the
 * new precision lacks falls back to separate passes).  For an FP64 retire
 * phase after an FP32 relaxation: switch, then ss_batch_begin to reactivate
 * every instance for the tighter tolerance.  (No counterpart in the reference,
 * which runs FP64 only.) */
int ss_batch_set_precision(ss_batch* b, int precision);

/* Start a relaxation: energy of every instance's field becomes its first
 * prev_energy (the reference's viewer loop seeds its E the same way,
 * chiral_magnet/render_gl.py:663-666); all instances active. */
int ss_batch_begin(ss_batch* b);

/* Up to nsteps arrested-Newton steps of every active instance.  An instance
 * retires (stops updating) after the first step whose max |stage-4 energy
 * gradient| is below tol, or, as failed, whose max |stage-4 energy gradient|
 * is NaN or infinite.  The host reads the activity flags every
 * check_every steps and returns early when none is active. */
int ss_batch_step(ss_batch* b, int nsteps, double tol, int check_every, int* n_active);

/* Capture steps_per_graph steps into one hipGraph replayed by ss_batch_step
 * (0 disables). */
int ss_batch_use_graphs(ss_batch* b, int steps_per_graph);

/* Per-instance energy of the current field, last max |EG|, steps taken and
 * the step at which it converged (-1 if not yet, -2 if it failed with a
 * non-finite error); arrays of ninst, any may be NULL. */
int ss_batch_state(ss_batch* b, double* energies, double* err_absmax, int* steps, int* converged_step);
int ss_batch_memory_bytes(ss_batch* b, size_t* bytes);

/* Launch shape of the batch's RK4 passes (no counterpart in the reference):
 * the single-lattice variants (0-2 LDS tiles of the default rows, halved,
 * quartered; 3-8 streaming strips of 64/128/256 rows, 64 or 128 columns
 * wide), one shape for every instance.  ss_batch_autotune_tiles times the
 * four passes in each variant ``reps`` times on the live buffers, and passes
 * 1, 2+3 (one sweep, SS_FUSE_23 as for a single lattice), 4 where the variant
 * has the fused sweep, and keeps the fastest variant and schedule; ms[0..8] =
 * mean time of the four passes per variant, ms[9..17] = the same with passes
 * 2+3 fused, ms[18..26] = the paired sweeps 1+2 and 3+4 (SS_FUSE_PAIRS as for
 * a single lattice; strip variants, kept when 3 % faster than the best other
 * schedule) (0 where the variant has no such sweep; ms holds 27 entries).
 * The relaxation state is unchanged; results do not depend on the variant or
 * schedule beyond FP64 summation order.  ss_batch_set_fused_stages (flags 0,
 * SS_FUSE_23 or SS_FUSE_PAIRS) fails where the current variant has no such sweep;
 * ss_batch_get_fused_stages reports whether steps use it. */
int ss_batch_autotune_tiles(ss_batch* b, int reps, int* chosen, double* ms);
int ss_batch_set_tile_variant(ss_batch* b, int variant);
int ss_batch_tile_variant(ss_batch* b, int* variant, int* rows);
int ss_batch_set_fused_stages(ss_batch* b, int flags);
int ss_batch_get_fused_stages(ss_batch* b, int* flags);

#ifdef __cplusplus
}
#endif

#endif /* SOLITON_H */
