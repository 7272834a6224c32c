/*
 * cpsd_abi.h -- C-ABI of the MI355X (gfx950) SD step, libcpsd_hip.so.
 *
 * Drop-in entry points (same symbols, arguments and semantics as the
 * reference; paths relative to the reference's Cpsd-mpi-bamboo/):
 *
 *   cpsd()       replaces ccpsdv3/cpsd.c:33-71 (prototype at
 *                ccpsdv3/main.c:18-56): `inner` steepest-descent steps of
 *                the orbitals followed by one energy evaluation.  All
 *                arrays are the caller's host arrays with the reference's
 *                layouts; on return c1, c2, cpsi (= c2 - c1), dn, hml, lmd,
 *                E[0..9], *deltae, *deltac, *deltar, r1 (and with move = 1
 *                r2, fion: ccpsdv3/cpsd.c:246-364, 401-411) hold what the
 *                reference leaves there.  cpsdwork is not written (the
 *                scratch lives on the GPU).
 *   cr_fft3b_()  replaces fftpack3d/fftpack3d.f:22-96 (Fortran ABI, all
 *                arguments by pointer): in-place unnormalised inverse 3D FFT
 *                of a half-complex (N/2+1) x N x N array into padded real
 *                rows of N+2 (pads set to zero).
 *   rc_fft3f_()  replaces fftpack3d/fftpack3d.f:107-196: the forward
 *                transform, Im(kx=0) = Im(kx=N/2) = 0.
 *   cpsd_rotate() replaces the three calls of the run epilogue at
 *                ccpsdv3/main.c:466-484 (esyev('V','U'), eigsrt, egemm):
 *                hml <- its eigenvectors, eig <- eigenvalues (descending),
 *                c1 = c2 hml, for ispin = 1.
 *
 * Supported inputs: ispin = 1, move = 0 or 1, ne[1] = 0, nx = ny = nz = nfft
 * with nfft in {8,12,16,20,24,32,40,48,64,80,96,128}, the reference's
 * increments, orbitals that vanish outside the spherical cutoff mask (the
 * reference keeps them there).  Anything else is reported on stderr and,
 * unless cpsd_set_abort_on_error(0) was called, aborts the process; the
 * library never falls back to a CPU path.
 *
 * The cpsd_dev_* functions have no counterpart in the reference: they keep
 * the orbitals resident in HBM across calls for callers that run many SD
 * steps (the run driver, benchmarks) and avoid the per-call host<->device
 * copies that the host-pointer contract of cpsd() implies.
 */
#ifndef CPSD_ABI_H
#define CPSD_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ccpsdv3/cpsd.c:33-71 */
void cpsd(const int inner, const int move, const int nion, const int ispin, const int *ne,
          const int nfft, const int nfft3d, const int n2ft3d, const double *unitg,
          const double omega, const double dt, const double fmass, const int *katm,
          const double *amass, const double *zv, const int nsh, const double *rcell,
          const double cewald, const double rcut, const int *masker, const double *G,
          const double *tg, const double *vc, const double *vg, int *lmmax, double *vnlnrm,
          double *vl, double *vnl, double *r1, double *r2, double *fion, double *c1,
          double *c2, double *cpsi, double *dn, double *hml, double *lmd, double *cpsdwork,
          double *E, double *deltae, double *deltac, double *deltar);

/* fftpack3d/fftpack3d.f:22-35 */
void cr_fft3b_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny,
               int *nz, double *tmp, int *ntmp);
/* fftpack3d/fftpack3d.f:107-135 */
void rc_fft3f_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny,
               int *nz, double *tmp, int *ntmp);

/* ccpsdv3/main.c:466-484 (ispin = 1): esyev('V','U', ne, hml) + eigsrt
 * (ccpsdv3/eigsrt.c:11-39) + egemm('N','N', n2ft3d, ne, ne, 1, c2, hml, 0,
 * c1).  On exit hml holds the eigenvectors in columns and eig the
 * eigenvalues in eigsrt's descending order; c1 = c2 hml.  Eigenvectors are
 * unique up to sign (and rotation within degenerate eigenvalues), so c1
 * matches the reference's up to the sign of each orbital.  Follows a cpsd()
 * call with the same nfft and ne, whose device tables it reuses. */
void cpsd_rotate(const int ispin, const int *ne, const int n2ft3d, double *hml, double *eig,
                 const double *c2, double *c1);

/* ---- error handling (no reference counterpart) ---------------------- */
const char *cpsd_last_error(void);
void cpsd_set_abort_on_error(int on);

/* ---- device-resident API (no reference counterpart) ------------------
 * Static inputs use the reference's layouts (ccpsdv3/main.c:145-320):
 * masker[n2ft3d], tg/vc/vg[nfft3d], vl[nkind*nfft3d], vnl[nkind*9*nfft3d],
 * vnlnrm[nkind*9], lmmax[nkind], katm[nion] (0-based), zv[nkind],
 * rcell[3*nsh], r[3*nion].  Returns 0 on success, nonzero on error
 * (message in cpsd_last_error()).                                        */
typedef struct cpsd_dev cpsd_dev;

int cpsd_dev_create(cpsd_dev **h, int nfft, int ne, int nion, const int *katm, const double *zv,
                    const int *lmmax, const double *vnlnrm, const double *vl, const double *vnl,
                    const int *masker, const double *tg, const double *vc, const double *vg,
                    const double *unitg, double omega, int nsh, const double *rcell,
                    double cewald, double rcut, const double *r);
/* orbitals c[ne*n2ft3d] in the reference's ELCIN layout */
int cpsd_dev_set_orbitals(cpsd_dev *h, const double *c);
/* ions for the following runs: positions r2[3*nion] (the r2 that cpsd()
 * takes); move = 1 also moves them every step with the forces of
 * ccpsdv3/cpsd.c:246-364 and masses amass[nkind] (as passed to cpsd()) */
int cpsd_dev_set_ions(cpsd_dev *h, int move, const double *amass, const double *r2);
/* positions at the start (r1) and end (r2) of the last step and, after a
 * moving-ion run, its forces fion[3*nion]; any pointer may be NULL */
int cpsd_dev_get_ions(cpsd_dev *h, double *r1, double *r2, double *fion);
/* inner SD steps + energy; E[0] holds the previous energy on entry
 * (as in the reference), E[0..9] on exit; deltas as in cpsd().         */
int cpsd_dev_run(cpsd_dev *h, int inner, double dt, double fmass, double *E, double *deltae,
                 double *deltac, double *deltar);
/* cpsd_dev_run() in two halves: _start enqueues the call (inner steps and
 * the energy evaluation, E[0] read as the previous energy) on the context's
 * stream and returns at once; _finish waits for it and returns what
 * cpsd_dev_run() returns.  Starting the calls of several contexts before
 * finishing any advances independent cells concurrently.  One started call
 * per context at a time.  No counterpart in the reference. */
int cpsd_dev_run_start(cpsd_dev *h, int inner, double dt, double fmass, const double *E);
int cpsd_dev_run_finish(cpsd_dev *h, double *E, double *deltae, double *deltac, double *deltar);
/* any pointer may be NULL; layouts as in cpsd() */
int cpsd_dev_get(cpsd_dev *h, double *c1, double *c2, double *cpsi, double *dn, double *hml,
                 double *lmd);
/* the run epilogue of ccpsdv3/main.c:451-484 on the resident state, after
 * a run: ortho[ne*ne] = gcdotc(c1_i, c1_j) of the last step's starting
 * orbitals (the orthonormality table of :456-464), eig[ne] and V[ne*ne] the
 * eigenpairs of hml as cpsd_rotate() returns them, and the device's c1
 * replaced by c2 V (read it with cpsd_dev_get); any output may be NULL */
int cpsd_dev_epilogue(cpsd_dev *h, double *ortho, double *eig, double *V);
/* precision of the orbital transforms: mode 0 runs every step in FP64 (the
 * default, the reference's float_double.h build); mode 1 runs the 3D FFTs of
 * the orbitals (G->r, the density sum, V psi and r->G) and the KB projection
 * and apply (FP32 MFMA products) on FP32 values in
 * every step of a cpsd_dev_run() call but its last, while the density
 * accumulation, the potentials, the overlap and constraint products, the
 * Lagrange iteration and the energies stay FP64.  The last
 * step, whose density and -H psi the energies are formed from, is FP64, so
 * E, deltae and deltac are FP64 values of the current orbitals.  Mode 2 is
 * mode 1 with the KB projection's inputs rounded to BF16 (FP32 sums), which
 * moves the converged energy by up to a few 1e-7 Ha per atom.  Mode 3 (ne =
 * 128 or 256) runs the FFTs as mode 1 and every product of those steps in
 * split BF16: each FP32 factor as hi + lo BF16 (hi = bf16(x), lo = bf16(x -
 * hi)), three BF16 MFMA products, FP32 sums, about 2^-16 relative per term.
 * That covers the KB projection and apply, the preconditioned update
 * (cpsd_dev_set_preconditioner) and the constraint, and the overlaps
 * (ccpsdv3/cpsd.c:415-437): those are formed as <c2|c1> = <c1|c1> + <d|c1>
 * and <c2|c2> = <c1|c1> + <d|c1> + <c1|d> + <d|d> from the FP64 <c1|c1>
 * carried from the previous step and the products of the step d = c2 - c1,
 * so that their error scales with |d|.  The Lagrange iteration stays FP64,
 * and the next FP64 step forms the full FP64 Gram, so a call returns
 * orthonormal orbitals as in mode 0.  Moving ions are refused with modes 2
 * and 3.  The reference has no such modes; its all-FP32 build (float_single.h
 * with sfftpack3d) is the nearest. */
int cpsd_dev_set_precision(cpsd_dev *h, int mode);
/* FP64 refresh period of the mixed modes (1-3): step k of a run (counted
 * from the last cpsd_dev_set_orbitals, k = 1, 2, ...) runs in FP64 when k is
 * a multiple of `period`, besides the last step of every call, which the
 * energies need in FP64 (so a call with inner = 1 runs only FP64 steps);
 * 0 (the default) refreshes on the last step of each call only.  Modes 2
 * and 3 are refused with moving ions (set_ions move = 1), and moving ions
 * with mode 2 or 3 set. */
int cpsd_dev_set_refresh(cpsd_dev *h, int period);
/* Safeguards of cpsd_dev_run(): a non-finite energy or delta is an error
 * (the orbitals are lost; set them again).  With the fallback on (the
 * default, rise = 1e-8), a call whose energy exceeds the previous call's by
 * more than rise * max(1, |E|) while a mixed mode or a preconditioner is set
 * turns both off, so the run continues with the reference's FP64 step;
 * _info reports how often that happened and the mode and preconditioner now
 * in force. */
int cpsd_dev_set_fallback(cpsd_dev *h, int enable, double rise);
int cpsd_dev_fallback_info(cpsd_dev *h, int *fallbacks, int *mode, int *precond);
/* the preconditioned SD step (no counterpart in the reference): with
 * K[nfft3d] set (one factor per half-complex G in the reference's layout,
 * e.g. the output of cpsd_precond_predict()), every step of a call becomes
 * c2 = c1 + K o dte (cpsi + c1 h) before the overlaps, the Lagrange
 * iteration and the constraint, with h the Lagrange matrix hml of the last
 * energy evaluation of these orbitals; the reference's step
 * c2 = c1 + dte cpsi (ccpsdv3/cpsd.c:399-400) is K = 1, h = 0.  At the
 * fixed point h = <c1|H|c1>, so the converged orbitals are the reference's.
 * Calls start preconditioning once an energy evaluation of the current
 * orbitals exists (the first call after cpsd_dev_set_orbitals() or the
 * epilogue runs plain steps).  cpsi, hml and the energies keep the
 * reference's definitions.  K = NULL turns it off. */
int cpsd_dev_set_preconditioner(cpsd_dev *h, const double *K);
/* ---- learned preconditioner (no counterpart in the reference) ----------
 * A per-G MLP (4 features -> 32 -> 32 -> 1, tanh, FP32 weights;
 * CPSD_MLP_NPARAM floats) predicts the factors of the preconditioned step
 * from features of each masked G: the kinetic diagonal |G|^2/2, the
 * diagonal of the KB operator and the local pseudopotential at G.
 * cpsd_dev_set_learned_preconditioner() evaluates it on the context's G and
 * sets K = clamp(alpha exp(y), kmin, kmax) as cpsd_dev_set_preconditioner()
 * would.  Training (cpsd_mlp_train) fits the least-squares step factor
 * K R ~ c_inf - c1 per G over states of the project's own SD runs, from the
 * per-G sums that cpsd_dev_stats_* collect; X[n][4] features, A[n], B[n]
 * sums, loss L = sum A K^2 - 2 B K.                                      */
#define CPSD_MLP_NPARAM 1249
int cpsd_mlp_nparam(void);
int cpsd_dev_set_learned_preconditioner(cpsd_dev *h, const float *params, double alpha, double kmin,
                                        double kmax);
/* features of the masked G in packed order (kx fastest, then ky, kz as in
 * the reference's grid), X[ng][4]; ng from cpsd_dev_sizes() */
int cpsd_dev_precond_features(cpsd_dev *h, float *X);
/* cinf: converged orbitals (ELCIN layout).  Between _begin and _end every
 * energy evaluation of cpsd_dev_run() adds its state c1 (the last step's):
 * A += |dte R|^2, B += Re(conj(dte R)(cinf - c1)) per G summed over the
 * orbitals, R = cpsi + c1 hml; _read: A[ng], B[ng] in packed order */
int cpsd_dev_stats_begin(cpsd_dev *h, const double *cinf);
int cpsd_dev_stats_end(cpsd_dev *h);
int cpsd_dev_stats_read(cpsd_dev *h, double *A, double *B);
/* the loss and its gradient for params (one forward and backward pass on
 * the GPU), `iters` Adam steps from params (updated in place; loss_hist[iters]
 * may be NULL), and the network output y[n] */
int cpsd_mlp_loss_grad(const float *X, const double *A, const double *B, long n, const float *params,
                       float *grad, double *loss);
int cpsd_mlp_train(const float *X, const double *A, const double *B, long n, float *params, int iters,
                   float lr, double *loss_hist);
int cpsd_mlp_predict(const float *X, long n, const float *params, float *y);
/* enable != 0: each cpsd_dev_run() call (its inner steps and the energy
 * evaluation) is captured once into a HIP graph and replayed on later calls
 * with the same inner, dt, fmass, precision mode and state parity, which
 * saves the per-kernel launch cost of small cells; results are those of the
 * uncaptured calls, bit for bit.  Calls with moving ions or profiling, and
 * the first call in each precision mode, run uncaptured.  No counterpart in
 * the reference. */
int cpsd_dev_set_graphs(cpsd_dev *h, int enable);
/* Many independent cells advanced together (no counterpart in the
 * reference).  Each cpsd_batch_run() call runs `inner` steps and the energy
 * evaluation of every cell (E[10 i .. 10 i + 9], E[10 i] the previous energy
 * on entry, deltae/deltac/deltar[i]).  A call is captured once into one HIP
 * graph holding every cell's work on its own stream, and replayed with one
 * hipGraphLaunch while every cell is in the state the capture started from;
 * other calls (a cell's first call in its precision mode, moving ions,
 * profiling, training statistics) run eagerly, one stream per cell.  The
 * cells must not be run on their own while the batch is in use.  _info:
 * graph launches, captures and eager calls so far. */
typedef struct cpsd_batch cpsd_batch;
int cpsd_batch_create(cpsd_batch **b, cpsd_dev *const *cells, int n);
int cpsd_batch_run(cpsd_batch *b, int inner, double dt, double fmass, double *E, double *deltae,
                   double *deltac, double *deltar);
int cpsd_batch_info(cpsd_batch *b, long *graph_launches, long *captures, long *eager_calls);
void cpsd_batch_destroy(cpsd_batch *b);
/* The launch-plan autotuner (no counterpart in the reference): measures
 * candidate plans on the context's own orbitals (split-K workgroup target
 * of the small-output products, orbitals per FFT launch pair, orbitals per
 * density partial plane), `steps` SD steps each, timed with HIP events, and
 * keeps the fastest; plans are remembered per size class (grid, orbitals,
 * projectors, precision mode, preconditioner) for the process, so later
 * contexts of the class take theirs without measuring.  The orbitals are
 * kept and the run state is reset as cpsd_dev_set_orbitals() resets it.
 * plan[5] (may be NULL): fb, orbitals per density plane (FP64, FP32 pass),
 * split-K target, 1 if taken from the process's table. */
int cpsd_dev_autotune(cpsd_dev *h, int steps, int *plan);
/* last step's Lagrange iteration count and final max|dX| */
int cpsd_dev_lagrange_info(cpsd_dev *h, int *iterations, double *adiff);
/* sizes: [0] ng (masked G vectors), [1] ngp, [2] nproj */
int cpsd_dev_sizes(cpsd_dev *h, int *sizes);
/* per-phase GPU time (ms) accumulated with HIP events on the library's
 * stream since the last reset; names separated by ';' */
int cpsd_dev_profile(cpsd_dev *h, int enable);
int cpsd_dev_profile_read(cpsd_dev *h, char *names, int names_len, double *ms, int *calls,
                          int max_phases);
void cpsd_dev_destroy(cpsd_dev *h);
/* copies internal buffer `what` (0 dn, 1 V, 2 dng, 3 CP, 4 P, 5 Gram,
 * 6-9 s11 s12 s21 s22, 10 lmd, 11 both orbital blocks, 12 xcp, 13 vall,
 * 14 KB coefficients, 15 projectors, 16 Lagrange info) into out[cap];
 * returns the number of doubles or -1 (test hook) */
long cpsd_dev_debug_fetch(cpsd_dev *h, int what, double *out, long cap);

/* ---- input generation on the device ---------------------------------
 * The Kleinman-Bylander transform of ckbppv3/dkbppv3.c:335-444 (the .vpp
 * generator) with the arguments of cpsdh_kbpp (include/cpsd_host.h): radial
 * table rho/vp/wp[(lmax+1)*nrho] of a PSP file, G[3*nfft3d] ->
 * vl[nfft3d], vnl[lmax^2 channels of nfft3d], vnlnrm[9]; vp is not modified. */
int cpsd_kbpp_device(double zv, int lmax, int nrho, double drho, const double *rho,
                     const double *vp, const double *wp, int nfft3d, const double *G,
                     double *vl, double *vnl, double *vnlnrm);

/* ---- kernel-level test hooks (no reference counterpart) -------------
 * Batched versions of the FFT entry points and the LDA kernel, used by
 * the parity tests.                                                     */
int cpsd_fft3d_batch(int dir, int nfft, int nbatch, double *A);  /* dir +1: cr, -1: rc */

/* ---- allocation-free entry point -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cpsd_fft3d_batch() on the caller's DEVICE array
 * A[nbatch][(nfft+2) nfft nfft] doubles, transformed in place with the same
 * kernels (dir +1: half-complex in, padded real out as cr_fft3b_; dir -1:
 * the reverse as rc_fft3f_), enqueued on `stream` (a hipStream_t; NULL is
 * the null stream) and not waited for.  It calls no hipMalloc, hipFree or
 * other allocating routine: the z-transformed intermediate and the kernels'
 * tables live in a file-scope device array of CPSD_NOALLOC_SCRATCH_BYTES
 * that each GPU holds until the process ends, so nbatch is limited to
 * cpsd_fft3d_batch_noalloc_max(nfft) (3 at 128^3, 9 at 96^3, 240 at 32^3);
 * larger requests return -1 with the reason in cpsd_last_error() and enqueue
 * nothing.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams or host threads) race and must be
 * serialised by the caller.  Returns 0 on success.                      */
#define CPSD_NOALLOC_SCRATCH_BYTES (64u << 20)
int cpsd_fft3d_batch_noalloc(int dir, int nfft, int nbatch, double *A, void *stream);
long cpsd_fft3d_batch_noalloc_max(int nfft);
int cpsd_lda_device(int n, const double *dn, double *xcp, double *xce);
int cpsd_erfc_device(int n, const double *w, double *out);
/* the host eigensolver of cpsd_rotate() and cpsd_dev_epilogue() (cyclic
 * Jacobi, no GPU work): a[n*n] column-major symmetric (upper triangle read)
 * <- eigenvectors in columns, eig[n] <- eigenvalues in eigsrt's descending
 * order (ccpsdv3/eigsrt.c:11-39 applied to dsyev's ascending output) */
int cpsd_syev_sorted(int n, double *a, double *eig);

#ifdef __cplusplus
}
#endif
#endif
