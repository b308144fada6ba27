/*
 * vie_op.h -- C ABI of the MI355X (gfx950) VIE operator library, libvie_hip.so.
 *
 * The library applies MARIE's FFT volume-integral-equation operators
 * (N and K, PWC and PWL voxel bases, their Hermitian adjoints, and the full
 * VIE operator J - (Mc/Mr) invG N J) and runs a device-resident GMRES on
 * them.  It replaces the operator handles that MARIE's MATLAB code binds in
 * src/src_utils/src_loaders/parse_inputs.m:6-40 and calls through
 *   Jout = func.mvp_N(Jin, T)            (src/src_solver/src_mvp/mvp_vie/mvp_N/mvp_N_{pwc,pwl}_tucker.m:1)
 *   Jout = func.mvp_K(Jin, T)            (.../mvp_K/mvp_K_{pwc,pwl}_tucker.m:1)
 *   Jout = func.mvp_herm_adj_N(Jin, T)   (.../mvp_N/mvp_N_herm_adj_{pwc,pwl}_tucker.m:1)
 *   Jout = func.mvp_herm_adj_K(Jin, T)   (.../mvp_K/mvp_K_herm_adj_{pwc,pwl}_tucker.m:1)
 *   Jout = mvp_vie(Jb, EP, func, dim, fN)(src/src_solver/src_mvp/mvp_vie/mvp_vie.m:1)
 *   [x,iter,resvec] = is_gmres_vie(A, b, restart, tol, maxit, x0)
 *                                        (src/src_solver/src_iterative_solvers/is_gmres_vie.m:1)
 *
 * Conventions
 *   - Complex numbers are interleaved (re, im) doubles, the layout of
 *     MATLAB's R2018a interleaved-complex API and of numpy complex128.
 *   - Voxel fields are column-major [n1, n2, n3, ql] (x fastest), as MATLAB
 *     stores Jin.  Packed vectors are component-major over idxS, as
 *     idx_solution orders them (src/src_geometry/body_geometry/geo_body_domain.m:20-35).
 *   - Batches of nrhs vectors are stored back to back.
 *   - Every call returns VIE_OK or a negative status; vie_last_error() gives
 *     the message.  Nothing throws across the ABI.
 *   - A context owns one HIP stream (or the one given with vie_set_stream);
 *     device-pointer calls are asynchronous on it, host-pointer calls are
 *     synchronous.  Calls on one context must not overlap.
 */
#ifndef VIE_OP_H
#define VIE_OP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define VIE_OK 0
#define VIE_ERR_ARG (-1)      /* invalid argument or unsupported size        */
#define VIE_ERR_HIP (-2)      /* HIP runtime error (no device, launch failure) */
#define VIE_ERR_OOM (-3)      /* device allocation failed                      */
#define VIE_ERR_STATE (-4)    /* operator/material not set before use           */

/* voxel basis (geo_assembly.m:6-8: PWX = 0 -> PWC, 1 -> PWL) */
#define VIE_BASIS_PWC 0
#define VIE_BASIS_PWL 1

/* operators */
#define VIE_OP_N 0     /* mvp_N_*_tucker            dense [n1,n2,n3,ql] in/out */
#define VIE_OP_K 1     /* mvp_K_*_tucker            dense                      */
#define VIE_OP_NH 2    /* mvp_N_herm_adj_*_tucker   dense                      */
#define VIE_OP_KH 3    /* mvp_K_herm_adj_*_tucker   dense                      */
#define VIE_OP_VIE 4   /* mvp_vie                   packed [ql*nS] in/out      */
#define VIE_OP_VIEH 5  /* (mvp_vie)^H               packed                     */

/* arithmetic of the FFT and spectral product */
#define VIE_PREC_FP64 0
#define VIE_PREC_FP32 1   /* FP32 complex FFT + product, FP64 in/out and identity term */

/* where the caller's pointers live */
#define VIE_MEM_HOST 0
#define VIE_MEM_DEVICE 1

typedef struct vie_ctx vie_ctx;

typedef struct vie_desc {
  int32_t n1, n2, n3;  /* voxel grid (MREDM.dimensions.n1..n3)                  */
  int32_t basis;       /* VIE_BASIS_PWC or VIE_BASIS_PWL                          */
  double res;          /* voxel edge [m] (MREDM.dimensions.res)                   */
  int32_t device;      /* HIP device ordinal                                      */
} vie_desc;

/* Create a context for one voxel grid.  Replaces the residency decision of
 * to_GPU.m:8-51: everything the operator needs is kept on the device. */
int vie_create(const vie_desc* desc, vie_ctx** out);
int vie_destroy(vie_ctx* ctx);
/* Message of the last failing call on ctx (or of the last failing vie_create
 * when ctx is NULL).  Valid until the next call. */
const char* vie_last_error(const vie_ctx* ctx);
/* Run on the caller's HIP stream (a hipStream_t) instead of the context's own. */
int vie_set_stream(vie_ctx* ctx, void* hip_stream);
/* Voxel edge of the Gram diagonal the VIE operators apply (invG), as the
 * handles T = func.mvp_G(T, res) / func.mvp_invG(T, res) take it
 * (mvp_G_pwc.m:3, mvp_invG_pwc.m:3, mvp_invG_pwl.m:3-7); vie_create sets it
 * from desc->res.  The N and K operators do not depend on it, so a context
 * built for mvp_N can serve mvp_vie after this call. */
int vie_set_gram(vie_ctx* ctx, double res);
/* Padded FFT lengths chosen for the grid (2*n_d when 2*n_d is 2,3,5,7-smooth). */
int vie_fft_dims(const vie_ctx* ctx, int32_t* P1, int32_t* P2, int32_t* P3);

/* Circulant Green spectrum, from the reference's Tucker T-struct.
 * Replaces the per-matvec re-expansion nmp(nmp(nmp(T.G,T.U1,1),T.U2,2),T.U3,3)
 * (mvp_N_pwl_tucker.m:31, nmp.m:1-21) by one FP64 expansion at set-up.
 *   op    VIE_OP_N or VIE_OP_K
 *   term  PWC: component j (0-based, T(j+1)); PWL: i*ncomp + j for T(i+1,j+1),
 *         i = (l,lp) pair 0..9, j = component 0..5 (N) or 0..2 (K)
 *   core  r1 x r2 x r3 complex, column-major (T.G)
 *   U1..U3  r_d x 2 n_d complex, column-major (T.U1..T.U3, assembly_fft_circ_tucker_*.m:22-27)
 * A term of rank 0 along some mode (an all-zero component) is the zero term;
 * its pointers are not read.  Every term of the operator must be set before
 * it is applied. */
int vie_set_green_tucker(vie_ctx* ctx, int32_t op, int32_t term, const double* core,
                         int32_t r1, int32_t r2, int32_t r3, const double* U1,
                         const double* U2, const double* U3);

/* Circulant Green spectrum from the real-space kernel with no truncation:
 * A is [n1, n2, n3, pwx, ncomp] complex column-major (the output of
 * assembly_N.m:16 / assembly_K.m:16; pwx = 1 PWC, 10 PWL; ncomp = 6 for N,
 * 3 for K).  Same embedding and parity as assembly_fft_circ_tucker_*.m at
 * tol -> 0.  mem says where A lives. */
int vie_set_green_dense(vie_ctx* ctx, int32_t op, const double* A, int32_t mem);

/* Material and scatterer support (em_assembly.m:14-16, geo_body_domain.m:13-35).
 * Mr, Mc: [n1, n2, n3] complex (host).  idxS: nS sorted 0-based column-major
 * voxel indices (MATLAB idxS - 1).  May be called again to change the
 * material (gradient / finite-difference loops); the support may change too. */
int vie_set_material(vie_ctx* ctx, const double* Mr, const double* Mc, const int64_t* idxS,
                     int64_t nS);
int64_t vie_num_scatterers(const vie_ctx* ctx);

/* Apply an operator to nrhs vectors.  Dense ops take/return [n1,n2,n3,ql]
 * fields; VIE ops take/return packed vectors of ql*nS (idx_solution order).
 * in and out must not overlap. */
int vie_apply(vie_ctx* ctx, int32_t op, int32_t precision, const double* in, double* out,
              int32_t nrhs, int32_t mem);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vie_apply on device pointers (in, out) that calls no
 * hipMalloc, hipFree or other allocating routine.  The five passes keep their
 * intermediate arrays in a fixed scratch array of VIE_NOALLOC_SCRATCH_MIB MiB
 * inside the library, one per device, held until the process ends; a call
 * whose workspace (2 ql nrhs n1 (n2 + P2) P3 complex values of the grid the
 * operator runs on, the support box for VIE/VIEH) does not fit returns
 * VIE_ERR_ARG.  Everything else must have been set up already: the spectrum,
 * the material, and for VIE_PREC_FP32 the FP32 spectrum (made by vie_autotune
 * or vie_apply at that precision); tile widths not tuned yet take the default.
 * Asynchronous on the context's stream.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (other streams, other contexts,
 * other threads) race and must be serialised by the caller.  Results equal
 * vie_apply's bitwise. */
#define VIE_NOALLOC_SCRATCH_MIB 128
int vie_apply_noalloc(vie_ctx* ctx, int32_t op, int32_t precision, const double* in, double* out,
                      int32_t nrhs);

/* GMRES on the VIE operator (op VIE or VIEH), device resident.
 * Plain form (precision VIE_PREC_FP64) mirrors is_gmres_vie.m /
 * is_iter_gmres_vie.m: GMRES(restart) with classical Gram-Schmidt, a
 * least-squares solve each step, restart from the recurrence residual,
 * stop when ||r|| <= tol ||b||.
 * Mixed form (VIE_PREC_FP32): FP64 iterative refinement whose correction
 * equations are solved by the same GMRES with FP64 Arnoldi vectors and the
 * FP32 operator; the outer residual b - A x uses the FP64 operator.
 *   b, x       nrhs packed vectors; x holds x0 on entry (zeros allowed)
 *   resvec     per rhs, up to restart*maxit+1 entries (may be NULL)
 *   nres       per rhs, number of resvec entries written (may be NULL)
 *   iters      per rhs, [p, outit] as is_gmres_vie.m:29 returns (may be NULL)
 *   nmatvec    total operator applications per rhs (may be NULL)
 * All pointer arguments use `mem` for b and x; the statistics are host memory. */
int vie_gmres(vie_ctx* ctx, int32_t op, int32_t precision, const double* b, double* x,
              int32_t nrhs, double tol, int32_t restart, int32_t maxit, int32_t mem,
              double* resvec, int32_t* nres, int32_t* iters, int32_t* nmatvec);

/* No counterpart in the reference (its solver is is_gmres_vie.m, vie_gmres
 * above): the short-recurrence solver of the mixed-precision batched form.
 * COCR (conjugate orthogonal conjugate residual) on the VIE system
 * A x = b, op VIE or VIEH, device resident, nrhs at once.  The operator
 * A = I - D N of mvp_vie.m (D = diag((Mc/Mr) invG), N the symmetric Galerkin
 * matrix of mvp_N) is symmetric in the bilinear form [u, v] = sum u_i v_i / d_i
 * (VIEH: sum u_i v_i conj(d_i)), so one operator application and a fixed
 * number of vector operations per iteration replace GMRES's Gram-Schmidt over
 * a growing basis.
 *   precision VIE_PREC_FP64: FP64 operator and vectors.
 *   precision VIE_PREC_FP32 (mixed): FP32 operator; recurrence scalars, x, b
 *     and the residual FP64, the direction vectors and the solution increment
 *     stored FP32.  x is accumulated in FP64 and the residual replaced by
 *     b - A x with the FP64 operator whenever the recurrence residual has
 *     fallen by 100x since the last replacement, and when it meets tol
 *     (reliable updates): a solve ends only when the FP64 residual satisfies
 *     ||b - A x|| < tol ||b||, or after maxit iterations.
 *   b, x       nrhs packed vectors (mem); x holds x0 on entry (zeros allowed)
 *   maxit      iterations (operator applications of the recurrence) per rhs
 *   resvec     per rhs up to 2*maxit+2 entries: ||r||/||b|| at the start,
 *              after every iteration (the recurrence residual) and after
 *              every replacement (the FP64 residual) (may be NULL)
 *   nres       per rhs, entries written (may be NULL)
 *   nmatvec    per rhs, operator applications at either precision (may be NULL)
 *   nreplace   per rhs, FP64 residual replacements (may be NULL) */
int vie_cocr(vie_ctx* ctx, int32_t op, int32_t precision, const double* b, double* x,
             int32_t nrhs, double tol, int32_t maxit, int32_t mem, double* resvec,
             int32_t* nres, int32_t* nmatvec, int32_t* nreplace);

/* ---- block COCG on the VIE operator --------------------------------------
 * No counterpart in the reference (its solver is GMRES one rhs at a time,
 * is_gmres_vie.m); for batches of rhs on one scatterer.  The rhs share one
 * block Krylov space (groups of up to 32 rhs; a larger nrhs runs group by
 * group): block conjugate gradients in the weighted bilinear form of
 * vie_cocr, with the residual block kept orthonormal (R = Q C, Cholesky QR).
 * One iteration applies the operator to every rhs of the group once.
 * Arguments, outputs and the stopping rule as vie_cocr: a solve ends when
 * every FP64 residual satisfies ||b_j - A x_j|| < tol ||b_j|| (checked with
 * the FP64 operator) or after maxit iterations.  precision VIE_PREC_FP32
 * runs the FP32 operator on an FP32 copy of the direction block (every other
 * block and all s x s algebra FP64) until the residual block has fallen by
 * 10x, then restarts from the FP64 residual in the FP64 form.  A breakdown
 * (numerically dependent residual columns, a singular Gram matrix) restarts
 * from the FP64 residual, at most 3 times; linearly dependent rhs end the
 * block recurrence, and the group then finishes with vie_cocr from the block
 * iterate (its counts included). */
int vie_bcocg(vie_ctx* ctx, int32_t op, int32_t precision, const double* b, double* x,
              int32_t nrhs, double tol, int32_t maxit, int32_t mem, double* resvec,
              int32_t* nres, int32_t* nmatvec, int32_t* nreplace);

/* ---- set-up of the real-space Green tensor (operator assembly) ----------
 * Far/medium-field Galerkin integrals of the N (op VIE_OP_N) or K (VIE_OP_K)
 * kernel with an nq-point Gauss-Legendre rule in all six coordinates, for
 * every offset in [0,e1) x [0,e2) x [0,e3); the rule of VV_Nop.m:1-49 /
 * VV_Kop.m:1-65 as assembly_N.m:20-52 and assembly_K.m apply it (far:
 * nq = 4 over the whole grid, medium: nq = 8 over the first 8^3 offsets).
 * A is [n1, n2, n3, pwx, ncomp] complex column-major (the input layout of
 * vie_set_green_dense); entries outside the box are left untouched.
 * nodes/weights are the nq Gauss-Legendre nodes and weights on [-1, 1].
 * The near offsets {0,1}^3 need singular integration and are filled by the
 * caller. */
int vie_green_quadrature(vie_ctx* ctx, int32_t op, double k0, int32_t nq, const double* nodes,
                         const double* weights, int32_t e1, int32_t e2, int32_t e3, double* A,
                         int32_t mem);

/* ---- tile-width autotuner of the FFT axis passes ------------------------
 * No counterpart in the reference (its FFTs are library calls).  Unless the
 * context has already fixed them for this precision and batch shape (one rhs /
 * several), times every compiled tile width (pencils per block) of the four
 * axis passes on scratch vectors of the given op / precision / batch size and
 * keeps the fastest per pass; tiles[4] (may be NULL) receives the widths of
 * the z forward, y forward, y inverse and z inverse passes.  The tiling
 * changes no arithmetic: results are bitwise the same for every width.  The
 * tuner also picks each pass's FFT schedule on the first tuning of a
 * precision (vie_axis_schedule).
 * Without this call the context tunes itself on the first vie_apply /
 * vie_gmres of each precision and batch shape (so calling it first only
 * moves the tuning out of a timed region); the environment variable
 * VIE_AUTOTUNE=0 keeps the default width, VIE_AXIS_TX=w fixes width w. */
int vie_autotune(vie_ctx* ctx, int32_t op, int32_t prec, int32_t nrhs, int32_t* tiles);
/* No reference counterpart.  FFT schedule of the four axis passes (z forward,
 * y forward, y inverse, z inverse) at precision prec, as the tuner fixed it:
 * 1 = Stockham mixed-radix stages with the largest power-of-two radix first,
 * 2 = two-step A x B register split, 3 = Stockham stages with the odd radices
 * (7, 5, 3) first, 0 = not fixed yet (the default: 1 for FP64, 2 for FP32).
 * The tuner picks the schedule with the tile width (vie_autotune); the
 * schedules agree to rounding, not bitwise, so the schedule of a precision is
 * fixed once per context.  VIE_AXIS_FFT=ct, rf or cto forces one. */
int vie_axis_schedule(const vie_ctx* ctx, int32_t prec, int32_t* sched);

/* No reference counterpart.  Support box of the packed operators: VIE_OP_VIE,
 * VIE_OP_VIEH and vie_gmres need only the kernel offsets inside the
 * scatterer's bounding box, so when a circulant of that box (FFT lengths
 * >= 2 w_d) is at most 0.9 of the grid's, they run on it (spectrum of the
 * N kernel cropped to offsets < w_d, from the dense kernel or the Tucker terms
 * rebuilt in real space; packed vectors unchanged).  Set up by
 * vie_set_material / vie_set_green_*; not used when the environment variable
 * VIE_BOX=0 or VIE_PRUNE=0.  Writes the box origin lo[3], extent w[3] and FFT
 * lengths P[3], all zero when no box is in use. */
int vie_box_dims(const vie_ctx* ctx, int32_t* lo, int32_t* w, int32_t* P);

#ifdef __cplusplus
}
#endif

#endif /* VIE_OP_H */
