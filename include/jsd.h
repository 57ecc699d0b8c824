/*
 * jsd.h - C ABI of the MI355X channel-flow step library (libjsd.so).
 *
 * The library advances the velocity-vorticity channel-flow equations of
 * dakling/jax-spectral-dns by its RK3/Crank-Nicolson step on one GPU, with
 * hand-written HIP kernels for gfx950.  All arrays handed over at run time
 * are device pointers in the reference's layout: complex128 (interleaved
 * re/im doubles) of shape (Nx_modes, Ny, Nz_modes) per velocity component,
 * components contiguous, wavenumbers stored [0..K, -K..-1], y index 0 at
 * y = +1.  Plain pointers and sizes only; no torch types cross this ABI.
 *
 * Every function returns JSD_OK (0) or a negative jsd_status; the message of
 * the last error of the calling thread is available from jsd_last_error().
 * A plan is immutable after creation and may be used by one stream at a
 * time.  Citations are reference file:line (jax_spectral_dns package).
 */
#ifndef JSD_H
#define JSD_H

#ifdef __cplusplus
extern "C" {
#endif

enum jsd_status {
  JSD_OK = 0,
  JSD_ERR_ARG = -1,    /* invalid argument or unsupported shape */
  JSD_ERR_HIP = -2,    /* a HIP runtime call failed */
  JSD_ERR_NOMEM = -3,  /* device allocation failed */
  JSD_ERR_STATE = -4   /* call not valid for this plan kind */
};

enum jsd_kind {
  JSD_KIND_NS = 0,            /* NavierStokesVelVort, skew-symmetric NL term */
  JSD_KIND_PERTURBATION = 1,  /* NavierStokesVelVortPerturbation */
  JSD_KIND_DUAL = 2           /* NavierStokesVelVortPerturbationDual */
};

typedef struct jsd_plan jsd_plan;

/* Host-side description; every array is copied at plan creation.  The
 * operator tables come from the host precompute (one entry per distinct
 * (|kx|, |kz|) pair, Ka = nxm/2+1, Kb = nzm/2+1, row-major). */
typedef struct jsd_plan_desc {
  int nxm, ny, nzm;   /* stored modes (odd, = 1 mod 4) and Chebyshev points (<= 130) */
  int px, pz;         /* the reference's physical grid: fixes the spectral scaling */
  int ppx, ppz;       /* transform grid of the FFT kernels: >= 3K+1, radices 2,3,4,5,7,8 */
  double lx, lz;      /* box lengths */
  double re_tau;      /* Reynolds number the tables were built with (negative for the dual) */
  double dt;          /* explicit-term time step (negative for the dual) */
  int kind;           /* jsd_kind */
  int linearise;      /* perturbation / dual: drop the u.grad u term (navier_stokes_perturbation.py:191) */
  const double* D;          /* ny*ny collocation derivative matrix */
  const double* kx;         /* nxm stored wavenumbers */
  const double* kz;         /* nzm */
  const double* lhs_inv;    /* 3*Ka*Kb*ny*ny inverse of hom_dir(I - beta_s dt L(k)) */
  const double* poisson;    /* Ka*Kb*ny*ny Poisson inverse */
  const double* v_a;        /* 3*Ka*Kb*ny wall-correction solution */
  const double* infl_inv;   /* 3*Ka*Kb*4 inverse influence matrix */
  const double* rhs_base;   /* 3*ny*ny hom_dir(I + alpha_s dt D^2 / Re) */
  const double* rhs_ns;     /* 3*(2ny)^2 mean-flow block operators */
  const double* lhs_inv_ns; /* 3*(2ny)^2 */
  const double* base_u;     /* ny base-flow profile U(y) (perturbation / dual) */
  const double* base_du;    /* ny U'(y) */
  const double* y_weights;  /* ny definite-integral weights */
  int adjoint;              /* nonzero: also build the tables of the discrete adjoint
                               (always built for linearised perturbation plans) */
  const double* y_eig;      /* optional (NULL: the solve applies lhs_inv / poisson):
                               eigen-decomposition of the interior D^2 (rows and
                               columns 1..ny-2) on its even / odd halves
                               (e_l = (x_l + x_l')/2, o_l = (x_l - x_l')/2,
                               l' = ny-1-l; half sizes ne = ny/2-1+ny%2, no = ny/2-1):
                               E_e (ne*ne), E_o (no*no), E_e^-1, E_o^-1 row-major,
                               then lam_e (ne), lam_o (no).  The Helmholtz and Poisson
                               inverses are then applied as E diag(.) E^-1 with
                               beta_s |dt/Re| from dt and re_tau; checked against
                               lhs_inv / poisson at plan creation */
  int no_coupling_term;     /* the reference's coupling_term = False, as the time-dependent
                               base flow uses (navier_stokes_perturbation_instationary.py:95-108,
                               navier_stokes_perturbation_instationary_dual.py:96-119):
                               perturbation plans add d_z hat(U' v) to h_g
                               (navier_stokes_perturbation.py:214-224), dual plans add
                               (d_x^2 + d_z^2) hat(U' v_x) to h_v
                               (navier_stokes_perturbation_dual.py:269-278) */
} jsd_plan_desc;

/* Replaces NavierStokesVelVort.__init__ / assemble_matrices
 * (navier_stokes.py:313-384, navier_stokes.py:437-502): uploads the operator
 * tables and allocates the step's workspaces. */
int jsd_plan_create(const jsd_plan_desc* desc, jsd_plan** out);
int jsd_plan_destroy(jsd_plan* plan);

/* Replaces the nonlinear_update_fn plug-in
 * (navier_stokes.py:324-332 -> :236-261; navier_stokes_perturbation.py:507-520
 * -> :160-239; navier_stokes_perturbation_dual.py:531-540 -> :214-294):
 * vel_hat (3 components) -> h_v, h_g (1 component each) and conv_ns
 * (3 components).  fwd_hat is the forward state for a nonlinear dual plan,
 * NULL otherwise. */
int jsd_nonlinear(jsd_plan* plan, const void* vel_hat, const void* fwd_hat,
                  void* h_v, void* h_g, void* conv_ns, void* stream);

/* Replaces perform_time_step / perform_runge_kutta_step
 * (navier_stokes.py:1778-1813, navier_stokes.py:1219-1775): advances vel_hat
 * in place by nsteps RK3/CN steps with constant pressure gradient dPdx. */
int jsd_step(jsd_plan* plan, void* vel_hat, double dPdx, int nsteps, void* stream);

/* One dual (adjoint) step of a JSD_KIND_DUAL plan
 * (navier_stokes_perturbation_dual.py:690-839 inner step): advances adj_hat in
 * place; fwd_hat is the forward state history[-1-t]
 * (navier_stokes_perturbation_dual.py:591-602), may be NULL when linearised. */
int jsd_dual_step(jsd_plan* plan, void* adj_hat, const void* fwd_hat, void* stream);

/* Replaces VectorField.energy of no_hat(vel_hat)
 * (field.py:607-611, field.py:1468-1473): writes the energy to *out (host). */
int jsd_energy(jsd_plan* plan, const void* vel_hat, double* out, void* stream);

/* Discrete adjoint (vector-Jacobian product) of nsteps linearised
 * perturbation steps: grad holds the adjoint of the state after the steps on
 * entry and the adjoint of the state before them on exit (real-pair view of
 * the complex128 layout).  Replaces reverse-mode AD through the scan,
 * jax.value_and_grad in optimiser.py:115-121 (navier_stokes.py:1839-1851).
 * Only for JSD_KIND_PERTURBATION plans with linearise = 1. */
int jsd_step_vjp(jsd_plan* plan, void* grad, int nsteps, void* stream);

/* One step of the discrete adjoint of a nonlinear step (JSD_KIND_NS, or
 * JSD_KIND_PERTURBATION with linearise = 0; the plan must be created with
 * adjoint = 1): grad holds the adjoint of the state after the step on entry
 * and the adjoint of the state before it on exit; state is the forward state
 * at the start of the step (unchanged) and dPdx the step's pressure
 * gradient.  The three substep states are recomputed from state and the
 * nonlinear term is transposed about each of them (transposed transforms and
 * products).  Replaces reverse-mode AD through the nonlinear scan
 * (jax.value_and_grad in optimiser.py:115-121).  The first call allocates
 * 9 spectral fields of recompute storage, kept by the plan.  Linearised
 * perturbation plans ignore state and run jsd_step_vjp. */
int jsd_step_vjp_state(jsd_plan* plan, const void* state, void* grad, double dPdx, void* stream);

/* Recorded steps: the forward keeps what the adjoint of a nonlinear step
 * would recompute, so that jsd_step_vjp_recorded recomputes nothing.  No
 * counterpart in the reference, whose reverse-mode AD (jax.value_and_grad
 * in optimiser.py:115-121) keeps or rematerialises the scan's residuals
 * itself; this is the same trade made explicit for callers with the memory
 * (a 96x129x80 step record is 0.42 GB).  Plans as for jsd_step_vjp_state.
 * jsd_step_record_bytes gives the size of one step's record;
 * jsd_step_record advances vel_hat by one step (as jsd_step) and writes the
 * record (device memory of that size: substep states 1 and 2, and the
 * inverse x transforms of substeps 0 and 1); jsd_step_vjp_recorded is
 * jsd_step_vjp_state with state the step's start state and record that
 * step's record.  The adjoint overwrites the record's transforms: one
 * adjoint per record. */
int jsd_step_record_bytes(jsd_plan* plan, size_t* bytes);
int jsd_step_record(jsd_plan* plan, void* vel_hat, void* record, double dPdx, void* stream);
int jsd_step_vjp_recorded(jsd_plan* plan, const void* state, const void* record, void* grad, double dPdx,
                          void* stream);

/* Replaces the base flow of a perturbation plan (ny values of U(y) and U'(y),
 * host arrays, copied before return): the time-dependent base flow of
 * NavierStokesVelVortPerturbationInstationary (vel_base_fn,
 * navier_stokes_perturbation_instationary.py:39-80) is set before each step.
 * Not for plans in a batch while it steps. */
int jsd_set_base_flow(jsd_plan* plan, const double* base_u, const double* base_du, void* stream);
/* As jsd_set_base_flow with device arrays: two device-to-device copies
 * queued on the stream, no synchronisation, so a loop of
 * (jsd_set_base_flow_device, jsd_step) over a device table of profiles
 * (one row per time step) never waits for the GPU and can be captured in a
 * graph.  The arrays must stay valid until the stream has run the copies.
 * Replaces the same reference interface (vel_base_fn evaluated per step,
 * navier_stokes_perturbation_instationary.py:39-80). */
int jsd_set_base_flow_device(jsd_plan* plan, const double* base_u, const double* base_du, void* stream);

/* One RK3/CN step like jsd_step (forward plans, fwd_hat NULL) or
 * jsd_dual_step (dual plans) with the optional source term of
 * get_source_term and the enforce_base switch.  source_hat (device,
 * complex128, 3 components, NULL for none) is the step's source, used the
 * way perform_runge_kutta_step uses it (navier_stokes.py:1697-1704, and the
 * mean-flow block source_x_00 / source_z_00, navier_stokes.py:1461-1490):
 * with Q = curl(source)_y, every kx row of h_v gains
 * (i kz) Q[kx row 0] - (i kx_1) Q[kx row 2] and every kx row of h_g gains
 * Q[kx row 1], as the reference's indexing of curl_source does; needs
 * nxm >= 3.  enforce_base != 0 drops the (0,0) convective terms
 * (navier_stokes.py:1581-1590; NavierStokesVelVortPerturbation enforce_base,
 * navier_stokes_perturbation.py:418).  The first call allocates 1 spectral
 * field plus 2 y-kz planes of workspace, kept by the plan. */
int jsd_step_source(jsd_plan* plan, void* vel_hat, const void* fwd_hat, double dPdx, const void* source_hat,
                    int enforce_base, void* stream);

/* One RK substep s (0, 1, 2) with the nonlinear terms supplied by the
 * caller: the nonlinear_update_fn plug-in (navier_stokes.py:324-332,
 * navier_stokes.py:1688-1696) returns h_v, h_g (1 component each) and
 * conv_ns (3 components) for vel_hat; the solve takes omega_y from vel_hat
 * itself (the reference's terms all return vort = curl(vel_hat)).  The
 * old terms are those of the previous substep call on this plan (s = 0
 * reads none).  source_hat and enforce_base as in jsd_step_source. */
int jsd_substep_terms(jsd_plan* plan, void* vel_hat, int s, double dPdx, const void* h_v, const void* h_g,
                      const void* conv_ns, const void* source_hat, int enforce_base, void* stream);

/* out = scale * laplacian(in) of ncomp spectral components
 * (Domain.laplacian of a Fourier field, domain.py:671-673; the dual's
 * dissipation source term, navier_stokes_perturbation_dual.py:962-988).
 * in and out must not overlap. */
int jsd_laplacian(jsd_plan* plan, const void* in, void* out, int ncomp, double scale, void* stream);

/* acc (1 spectral component) += scale * sum_ij (d_j u_i)^2 with the complex
 * squares of the spectral derivatives of vel_hat, the summand of the
 * dissipation objective (get_dissipation_average,
 * navier_stokes_perturbation_dual.py:1087-1104).  Uses the plan's
 * y-derivative workspace: not between the substeps of a step. */
int jsd_dissipation_accumulate(jsd_plan* plan, const void* vel_hat, void* acc, double scale, void* stream);

/* *out (host) = volume integral of no_hat(f_hat) for one spectral
 * component (PhysicalField.volume_integral, field.py:1456-1466). */
int jsd_volume_integral(jsd_plan* plan, const void* f_hat, double* out, void* stream);

/* No counterpart in the reference: jsd_energy for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Needs no plan and allocates nothing: vel_hat
 * (3 components, the layout above), y_weights (ny definite-integral
 * weights) and out (1 double) are device pointers; the energy is written
 * to *out on the stream, without a host copy or synchronisation.  The
 * 1024 partial sums live in a static 8 KB device array per GPU, shared by
 * every call on that GPU: calls that overlap on one GPU race.  ny <= 132,
 * odd nxm, nzm.  The same kernels and grid as jsd_energy: the results are
 * identical. */
int jsd_energy_noalloc(const void* vel_hat, int nxm, int ny, int nzm, int px, int pz, double lx, double lz,
                       const double* y_weights, double* out, void* stream);

/* out = scale * dE/d(vel_hat) on the real-pair view, E the energy of
 * jsd_energy (field.py:607-611). */
int jsd_energy_grad(jsd_plan* plan, const void* vel_hat, void* out, double scale, void* stream);

/* y-derivative of order 1 or 2 of ncomp spectral components
 * (FourierDomain.diff along y, domain.py:903-917). */
int jsd_ydiff(jsd_plan* plan, const void* in, void* out, int ncomp, int order, void* stream);

/* Instrumentation, no counterpart in the reference (its only timer is the
 * wall clock around solve_scan, equation.py:347-401).  enable is a bit mask
 * over the 7 kernel classes (bit 0 y-derivative GEMM, 1 inverse x FFT, 2 z
 * products pass, 3 forward x FFT, 4 h_v GEMM, 5 implicit solve, 6 linearised
 * term; 0x7f all, 0 none): every later launch of a selected class is
 * bracketed by HIP events on its stream.  A call with ms / count != NULL
 * first returns the accumulated milliseconds and launch counts of the
 * classes, then resets them. */
int jsd_profile(jsd_plan* plan, int enable, double* ms, long* count);

/* ---- batches of independent runs ------------------------------------------
 * Two kinds of batch.  Nonlinear perturbation runs (JSD_KIND_PERTURBATION
 * plans with linearise = 0, coupling term on; with adjoint buffers, plans
 * created with adjoint = 1) step concurrently: every run has a stream of its
 * own, forked from the caller's stream and joined back to it, so the stage
 * kernels of different runs overlap on the GPU; their adjoint takes each
 * run's forward state (jsd_batch_step_vjp_state).  Many linearised runs
 * (JSD_KIND_PERTURBATION or JSD_KIND_DUAL plans with
 * linearise = 1, uneven shapes, Re, dt and box sizes allowed) are advanced with
 * one launch per stage for all of them: each stage kernel walks the
 * concatenated per-run grids with a persistent grid whose size is picked by
 * timing candidates at creation (the autotuner), and one RK3 step (forward
 * and, for perturbation plans, adjoint) is captured once in a HIP graph and
 * replayed.  Replaces a Python loop of perform_time_step calls over runs
 * (navier_stokes.py:1778-1813, one launch sequence per run and step) for the
 * batched gradient evaluations of the optimisation drivers
 * (gradient_descent_solver.py:490-934); the batching itself has no
 * counterpart in the reference.
 *
 * states[r] / adjoints[r] are the runs' device state buffers (complex128,
 * each plan's layout), bound at creation; adjoints may be NULL (no VJP).
 * Creation runs the autotuner on the stream and leaves the states
 * unchanged.  A batch uses its plans' workspaces: do not step those plans
 * concurrently with the batch.  Each run needs its own plan and buffers: a
 * plan, state or adjoint buffer that appears twice is rejected (JSD_ERR_ARG). */
typedef struct jsd_batch jsd_batch;
int jsd_batch_create(jsd_plan* const* plans, void* const* states, void* const* adjoints, int nrun, void* stream,
                     jsd_batch** out);
int jsd_batch_destroy(jsd_batch* batch);
/* nsteps RK3 steps of every run (as jsd_step with dPdx = 0 on each). */
int jsd_batch_step(jsd_batch* batch, int nsteps, void* stream);
/* nsteps adjoint steps of every run (as jsd_step_vjp on each adjoint);
 * linearised batches only. */
int jsd_batch_step_vjp(jsd_batch* batch, int nsteps, void* stream);
/* One adjoint step of every run, as jsd_step_vjp_state(plan_r, states[r],
 * adjoint_r, 0) (states[r]: device pointer to run r's forward state at the
 * start of the step); on a linearised batch the states are not read and
 * this is jsd_batch_step_vjp with nsteps = 1. */
int jsd_batch_step_vjp_state(jsd_batch* batch, const void* const* states, void* stream);
/* Autotuned persistent grid sizes of the 7 stage kernels (forward y-GEMM,
 * linearised term, solve; adjoint solve, h_v/h_g transpose, linearised term
 * transpose, D^T accumulation) and their total logical block counts (all 0
 * for a batch of nonlinear runs, which launches each run's own kernels). */
int jsd_batch_info(jsd_batch* batch, int* grid, int* blocks);
/* Tile shapes, per stage as in jsd_batch_info: pencils and waves per
 * workgroup.  The autotuner times the forward solve's shapes together with
 * the grid size on the batch's own data: 4 pencils on 4 waves (one 16-column
 * MFMA tile per wave), 8 on 4 waves (two tiles per wave) or 8 on 8 waves
 * (one parity per wave); the last two for eigen-form plans only.  The
 * adjoint solve reports 8 / 8 (its eigen-form transpose) or 4 / 4 (table
 * form).  0 for stages with one shape and for batches of nonlinear runs.
 * waves may be NULL.  Replaces nothing in the reference (which has no
 * batching); it reports how the batch runs. */
int jsd_batch_tiles(jsd_batch* batch, int* pencils, int* waves);

const char* jsd_last_error(void);
int jsd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* JSD_H */
