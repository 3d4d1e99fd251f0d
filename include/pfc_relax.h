/*
 * pfc_relax.h -- C ABI of libpfcgpu, the MI355X (gfx950) engine behind the
 * pfc-relax drop-in binary.
 *
 * The reference (petenez/pfc, src/pfc-relax.c) is one C translation unit with
 * no library API: its "boundary" is the executable plus text files.  The
 * functions below are the seam a maintainer would bind to replace the solver
 * core of that file; each one names the reference function(s) it stands in
 * for.  All arguments are plain C types and host pointers; device memory is
 * private to the context.  Every call returns 0 on success or a negative
 * PFC_E* code and never exits the process.
 *
 * A context is not thread safe: use one per host thread.
 */
#ifndef PFC_RELAX_H
#define PFC_RELAX_H

#ifdef __cplusplus
extern "C" {
#endif

#define PFC_MAX_FIELDS 8   /* largest U accepted */
#define PFC_MAX_MODES 8    /* largest W[u] accepted */

/* Model parameters, as read from the "M" and "R" blocks of an input file
 * (reference: struct S, src/pfc-relax.c:34-120; parsed at :541-718).
 * Matrices are indexed [u][v] and only the upper triangle is read:
 * alpha/beta/gamma use v >= u, epsilon/mu2 use v > u.
 * nu2 and mu2 hold 1/nu^2 and 1/mu^2 (the file stores nu and mu, :654, :699);
 * mu2[u][v] is 0 when beta[u][v] == 0 (:698). */
typedef struct pfc_params {
    int U;                                        /* number of density fields */
    int W[PFC_MAX_FIELDS];                        /* modes per field */
    int energy_compat;                            /* see PFC_ENERGY_* below */
    int reserved;
    double c;                                     /* 0: nonconserved, else conserved */
    double sigma;                                 /* Gaussian width in k space */
    double dt;                                    /* time step */
    double lambda[PFC_MAX_FIELDS];
    double delta[PFC_MAX_FIELDS];
    double alpha[PFC_MAX_FIELDS][PFC_MAX_FIELDS];
    double beta[PFC_MAX_FIELDS][PFC_MAX_FIELDS];
    double gamma[PFC_MAX_FIELDS][PFC_MAX_FIELDS];
    double epsilon[PFC_MAX_FIELDS][PFC_MAX_FIELDS];
    double mu2[PFC_MAX_FIELDS][PFC_MAX_FIELDS];
    double nu2[PFC_MAX_FIELDS][PFC_MAX_MODES];
    double omega[PFC_MAX_FIELDS][PFC_MAX_MODES];
    double phi[PFC_MAX_FIELDS][PFC_MAX_MODES];
} pfc_params;

/* energy_compat: which coefficient the free energy's beta_uv gradient
 * coupling uses.  The reference's evaluate() aliases mu2 to nu2
 * (src/pfc-relax.c:1019) and then reads nu2[u][v] (:1064), past the end of
 * nu2[u] when v >= W[u].  PFC_ENERGY_MODEL evaluates the documented model
 * (mu2, :9); PFC_ENERGY_NU2 reproduces the alias, using nu2[u][v] when
 * v < W[u] and 0 otherwise.  Dynamics are unaffected (step() uses mu2). */
#define PFC_ENERGY_MODEL 0
#define PFC_ENERGY_NU2 1

/* precision */
#define PFC_FP64 0    /* all-FP64 FFTs, update and reductions */
#define PFC_MIXED 1   /* FP32 FFTs/update, FP64 reductions, k=0 mode and box */

/* error codes */
#define PFC_OK 0
#define PFC_EINVAL (-1)     /* bad argument (size, U, W, pointer) */
#define PFC_EDEVICE (-2)    /* HIP runtime error */
#define PFC_ENOMEM (-3)     /* device allocation failed */
#define PFC_ESIZE (-4)      /* grid size has no kernel (Nx, Ny: 2^k in [64, 8192] or 384, 768, 1536) */
#define PFC_ESTATE (-5)     /* call out of order (e.g. step before set_fields) */

typedef struct pfc_ctx pfc_ctx;

/* Allocates device state for an Nx x Ny grid of p->U fields and compiles no
 * plan: kernels and twiddle tables are fixed per size.
 * Replaces allocate_parameters/allocate_data/plan_FFTs (src/pfc-relax.c:123-239). */
int pfc_create(pfc_ctx** ctx, const pfc_params* p, int Nx, int Ny, int precision);

/* Uploads U density fields, each Ny rows of Nx values (row-major, x fastest,
 * the order of the density files), and synchronises the spectral state to
 * them (N = FFT(n)/NxNy).  Replaces read_data's store into m (:728-763). */
int pfc_set_fields(pfc_ctx* ctx, const double* fields);

/* Sets the discretisation; A and B operators follow from it.
 * Replaces the dx/dy assignment plus update_AB (:900-957). */
int pfc_set_box(pfc_ctx* ctx, double dx, double dy);

/* Runs nsteps semi-implicit steps t0, t0+1, ...; for each step t with
 * t % 100 == 0 the spectral state is first re-synchronised to the real-space
 * field.  A NaN in field 0 at grid point (0,0) at the start of a step is
 * recorded (see pfc_nan_step) and the remaining steps still run.
 * Replaces step() (:1133-1247). */
int pfc_step(pfc_ctx* ctx, int t0, int nsteps);

/* Free energy per unit area and mean densities (nave has U entries); the
 * energy density field is kept on the device (pfc_get_energy) and the
 * spectral state is re-synchronised, as in the reference.
 * Replaces evaluate() incl. its MPI_Allreduce (:987-1130). */
int pfc_evaluate(pfc_ctx* ctx, double* fave, double* nave);

/* The same fave and nave as pfc_evaluate (to rounding) and the same
 * re-synchronised state, without the energy density field: the terms with a
 * filtered field are summed in k space (Parseval), the local ones in real
 * space.  For the evaluates whose density is never written (relax prints,
 * :1299-1302, and optimize's samples, :1250-1289, unless a write_energy
 * follows before the next evaluate).  Afterwards pfc_get_energy fails with
 * PFC_ESTATE until the next pfc_evaluate.
 * Replaces evaluate() (:987-1130) where s->f is not read. */
int pfc_evaluate_sums(pfc_ctx* ctx, double* fave, double* nave);

/* pfc_evaluate_sums followed by pfc_step(ctx, t0, nsteps), nsteps >= 1, as
 * relax() runs a print and the steps after it when no write falls on t0
 * (:1299-1310).  Where step t0 refreshes the state (t0 % 100 == 0,
 * :1178-1184), its row and column passes also take the evaluate's sums of
 * the refreshed state, which is what the evaluate's own refresh (:1126-1128)
 * computes, so the evaluate costs no pass of its own; elsewhere the two run
 * one after the other.  fave and nave are those of the state at t0; the
 * call returns once they are known, with the later steps possibly still
 * running on the device.  Afterwards pfc_get_energy fails with PFC_ESTATE
 * until the next pfc_evaluate.
 * Replaces evaluate() (:987-1130) and step() (:1133-1247) of a print step. */
int pfc_step_evaluate(pfc_ctx* ctx, int t0, int nsteps, double* fave, double* nave);

/* Downloads the current density fields (layout as pfc_set_fields).
 * Feeds write_data (:766-813). */
int pfc_get_fields(pfc_ctx* ctx, double* fields);

/* Downloads the energy density of the last pfc_evaluate, Ny x Nx values
 * (zeros before the first evaluate); PFC_ESTATE if a pfc_evaluate_sums came
 * after it.  Feeds write_energy (:816-858). */
int pfc_get_energy(pfc_ctx* ctx, double* f);

/* First step index whose start saw a NaN at field 0, point (0,0), or -1.
 * Replaces the check at :1175 (fail code 39). */
int pfc_nan_step(pfc_ctx* ctx, int* step);

/* Measurement hooks; no counterpart in the reference (it only prints whole
 * seconds, :880).  While enabled, pfc_step brackets each of its two passes
 * with HIP events on the context's stream.  pfc_timing(ctx, 1) also clears
 * what was collected.  pfc_timing_read waits for the stream and returns the
 * summed milliseconds of the row pass (ms[0]) and the column pass (ms[1])
 * over *nsteps recorded steps. */
int pfc_timing(pfc_ctx* ctx, int enable);
int pfc_timing_read(pfc_ctx* ctx, double* ms, int* nsteps);

/* Measurement hook with no counterpart in the reference: multiplies every
 * real and imaginary part of the step state (the row transforms of the
 * density, and the spectral state except its columns kx = 0 and kx = Nx/2,
 * whose anti-Hermitian part the unstable modes amplify ~1e15-fold between
 * refreshes; the mean density is thus left exact) by 1 + rel * r, r a zero-mean,
 * unit-variance uniform deviate drawn from a counter hash of (seed, index).
 * Applied after each step it emulates a per-step round-off of relative size
 * rel in an otherwise FP64 run, the control that separates the mixed form's
 * FP32 round-off from the sensitivity of a relaxation to any such error. */
int pfc_perturb(pfc_ctx* ctx, double rel, unsigned long long seed);

/* Releases everything.  Replaces clean() (:1315-1389). */
int pfc_destroy(pfc_ctx* ctx);

/* Human-readable text for a PFC_E* code. */
const char* pfc_strerror(int code);

/* Gaussian smoothing of one Ny x Nx field on the GPU: the spectrum is
 * multiplied by exp(-k^2/(2 sigma^2)) (sigma > 0, in reciprocal units; dx, dy
 * the grid spacings) and transformed back; in and out may be the same array.
 * Replaces the computation of the reference's smooth tool
 * (src/smooth.c:57-117; its final division by sum(C2R(G)), which is 1 up to
 * round-off, is left out).  Uses a temporary context (PFC_FP64 or PFC_MIXED
 * arithmetic; same size limits as pfc_create). */
int pfc_smooth(const double* in, double* out, int Nx, int Ny, double dx, double dy,
               double sigma, int precision);

/* Orientation field, and optionally deformation field, of one H x W density
 * (row-major), the computation of the reference's fields tool
 * (src/fields.c:104-359, driven by main :368-424) on the GPU in FP64.
 *   phi: 2*W*H doubles, (re, im) per point, row-major, divided by max |phi|
 *        (fields.c:104-226);
 *   chi: NULL, or W*H doubles: the sum over scales n = 2, 4, 8, ... while
 *        n*4*pi/sqrt(3) < min(W dx, H dy)/2 of |grad phi|^power smoothed by a
 *        Gaussian of width 1/n, each divided by its maximum (fields.c:229-359).
 * fold: order of rotational symmetry (even); sk: width of the complex kernel
 * around |k| = 1; sb: width of the Gaussian smoothing of phi (reciprocal
 * units).  W and H: any sides whose transforms fit one block's LDS (the
 * generic transforms: prime factors up to 13 up to 4864, others up to 2048);
 * 2^k in [64, 8192] and 3 * 2^k run the tuned ones. */
int pfc_fields(const double* density, int W, int H, double dx, double dy, int fold, double sk,
               double sb, double power, double* phi, double* chi);

/* Gaussian smoothing as pfc_smooth, on device buffers and without
 * allocating anything.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  in and out are device pointers to
 * Ny x Nx doubles (they may be the same buffer); stream is a hipStream_t
 * (NULL: the null stream), on which the work is queued asynchronously.  All
 * work memory is one static device array per GPU of
 * PFC_NOALLOC_SCRATCH_BYTES (held from the library's load until the process
 * ends), which limits Nx * Ny to PFC_NOALLOC_MAX_POINTS (2048 x 2048); larger
 * grids return PFC_ESIZE.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other threads, other streams) race. */
#define PFC_NOALLOC_MAX_POINTS 4194304
#define PFC_NOALLOC_SCRATCH_BYTES 34603008
int pfc_smooth_noalloc(const double* in, double* out, int Nx, int Ny, double dx, double dy,
                       double sigma, int precision, void* stream);

/* ---- batch form: many independent systems stepped together ------------
 * No counterpart in the reference, which relaxes one system per run
 * (src/pfc-relax.c:1391-1434); a batch is that run for many systems at once.
 *
 * pfc_batch_create groups n existing contexts of one precision (U = 1 or 3;
 * Nx and Ny each in {256, 384, 512, 768, 1024, 1536, 2048}).  The contexts
 * keep working through the one-system calls (pfc_set_fields, pfc_get_fields,
 * pfc_evaluate, pfc_set_box, pfc_nan_step); pfc_batch_step advances all of
 * them by steps t0 ... t0 + nsteps - 1 exactly as pfc_step would.  Systems of
 * equal (Nx, Ny, U) are cut into chunks of a cache-sized working set; each
 * chunk's row and column passes are launches of one block per row group or
 * column block of all its systems, chained on one of four streams, and unless
 * PFC_BATCH_NO_GRAPH, blocks of 20 steps (single steps where t0 or nsteps do
 * not align) are replayed from captured HIP graphs.  PFC_BATCH_AUTOTUNE times the
 * FFT shapes of every (length, U) and pass present -- points per thread 8 or
 * 16 (U = 1) and 8 or 4 (U = 3) for 2^k lengths, 12 or 6 (U = 1) and 6 or 12
 * (U = 3) for 3*2^k, and the first of these with half the rows (columns) per
 * block; shapes whose kernel spills registers are left out where one without
 * spills exists -- once per process, and times chunk budgets of 8 to 64 MiB
 * and two ways of spreading the chunks over the streams for the batch;
 * without it the first shape is used unless this process has tuned the
 * class, with 16 MiB chunks spread by work.  The tuning passes write only the
 * contexts' scratch buffers.  Destroy the batch before its contexts.
 * (Environment: PFC_BATCH_CHUNK_MB and PFC_BATCH_POLICY (0 or 1) fix the
 * budget and the spreading; PFC_BATCH_VERBOSE=1 prints the choices.) */
#define PFC_BATCH_AUTOTUNE 1
#define PFC_BATCH_NO_GRAPH 2
typedef struct pfc_batch pfc_batch;
int pfc_batch_create(pfc_batch** batch, pfc_ctx* const* ctxs, int n, int flags);
int pfc_batch_step(pfc_batch* batch, int t0, int nsteps);
/* points per thread chosen for pass 0 (rows, length Nx) or 1 (columns,
 * length Ny) of the class (n, U) */
int pfc_batch_shape(const pfc_batch* batch, int pass, int n, int U, int* points);
int pfc_batch_destroy(pfc_batch* batch);

#ifdef __cplusplus
}
#endif

#endif
