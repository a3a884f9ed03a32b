/*
 * tsc_hotpath.h -- C-ABI of the MI355X (gfx950) hot path of the two-site
 * SU(2) DMRG for the t1-t2-J1-J2 t-J model on triangular cylinders.
 *
 * Everything here is plain C: pointers, sizes, int status codes (0 = ok).
 * Device buffers are owned by a context; host code hands over host arrays
 * (column-major complex128, Fortran layout) and gets host arrays back, or
 * keeps vectors resident on the device between calls through the
 * tsc_vec_* handles.
 *
 * Reference interfaces replaced (hyx15903/TSC_Tri_t1t2J1J2, DMRG_code/):
 *   model_wave        get_sys.f90:6915-6922  -> tsc_plan_* + tsc_hpsi
 *   hmerge            get_sys.f90:7398-7402  -> tsc_block_axpy (merged op)
 *   lanczos_memo_new  eigen.f90:4-12         -> tsc_lanczos
 *   wave_product / wave_norm / Gram_Schmidt_process
 *                     block.f90:1092-1150,1251-1289 -> tsc_vec_dot_re,
 *                                            tsc_vec_norm, used by tsc_lanczos
 *   Get_density_matrix_sys / _env
 *                     den_mat.f90:4-99       -> tsc_density_matrix
 *   Get_truncation_operator
 *                     den_mat.f90:184-418    -> tsc_truncation_operator
 *   update_trun_dia / update_trun_ndia
 *                     block.f90:3465-3683    -> tsc_update_trun
 *   W6JS              block.f90:3744-3795    -> tsc_w6j
 *
 * Threading: a context is single-stream and not re-entrant, as the
 * reference's globals are. Errors: nonzero return, message from
 * tsc_last_error(); the Fortran shim turns them into `error stop`.
 */
#ifndef TSC_HOTPATH_H
#define TSC_HOTPATH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TSC_OK 0
#define TSC_EINVAL 1   /* bad argument / inconsistent descriptor        */
#define TSC_EHIP 2     /* HIP runtime error                              */
#define TSC_ENOMEM 3   /* device allocation failed                       */
#define TSC_ESTATE 4   /* call out of order (e.g. no plan)               */

/* ---- precision policies for the grouped contraction -------------------- */
#define TSC_PREC_F64 0    /* complex FP64, FP64 MFMA                      */
#define TSC_PREC_F32 1    /* complex FP32 products, FP32 MFMA, FP32 acc   */
#define TSC_PREC_BF16X3 2 /* BF16x3 split products, FP32 accumulation     */

/* ---- operand modes of a contribution ----------------------------------- */
#define TSC_OP_N 0      /* A                   */
#define TSC_OP_T 1      /* A^T                 */
#define TSC_OP_C 2      /* A^H (conj-transpose)*/
#define TSC_OP_CONJ 3   /* conj(A)             */
#define TSC_OP_I 4      /* identity (no load)  */

/* ---- buffers a contribution can address -------------------------------- */
#define TSC_BUF_OPS 0   /* operator arena (uploaded with the plan)        */
#define TSC_BUF_IN 1    /* input vector psi                               */
#define TSC_BUF_MID 2   /* stage-1 intermediates (owned by the plan)      */
#define TSC_BUF_OUT 3   /* output vector H psi                            */

/*
 * One term of a grouped contraction:
 *     C[out block] (+)= coef * op_a(A) * op_b(B)
 * A is rows(C) x k after op_a, B is k x cols(C) after op_b; both stored
 * column-major at element offsets a_off / b_off of buffers a_buf / b_buf
 * with leading dimensions lda / ldb (complex elements). With op_a = I the
 * term is coef*op_b(B) (k = rows(C)); with op_b = I it is coef*op_a(A).
 */
typedef struct tsc_contrib {
  int64_t a_off, b_off;
  int32_t lda, ldb;
  int32_t k;
  int8_t a_buf, b_buf, a_op, b_op;
  double coef_re, coef_im;
} tsc_contrib;

/*
 * One output block: an m x n window (leading dimension ldc) of buffer
 * c_buf at element offset c_off, written as the sum of contributions
 * [cbeg, cend). mode 0 overwrites the window, mode 1 adds to it.
 */
typedef struct tsc_outblock {
  int64_t c_off;
  int32_t ldc, m, n;
  int32_t c_buf, cbeg, cend, mode;
} tsc_outblock;

typedef struct tsc_ctx tsc_ctx;
typedef struct tsc_plan tsc_plan;
typedef struct tsc_vec tsc_vec;

const char *tsc_last_error(void);
int tsc_version(void);

int tsc_ctx_create(int device, tsc_ctx **out);
int tsc_ctx_destroy(tsc_ctx *ctx);
int tsc_ctx_sync(tsc_ctx *ctx);
/*
 * MFMA accumulator layouts found by the probe run at context creation:
 * modes3[0], modes3[1] = row layout of the f64 / f32 16x16 MFMAs (0: row =
 * 4*(lane/16)+r, 1: row = lane/16+4r); modes3[2] = 0 if the 32x32x16 bf16
 * layout is row = 8*(r/4)+4*(lane/32)+r%4 (BF16x3 usable), -1 otherwise.
 * Diagnostic; has no counterpart in the reference.
 */
int tsc_ctx_mfma_layout(const tsc_ctx *ctx, int32_t *modes3);

/*
 * Free and total device memory (bytes) of the context's GPU, as the HIP
 * runtime reports them (every process's allocations count). Diagnostic for
 * HBM footprints; has no counterpart in the reference.
 */
int tsc_ctx_mem_info(const tsc_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- resident vectors (psi, Lanczos vectors) ---------------------------- */
int tsc_vec_create(tsc_ctx *ctx, int64_t n, tsc_vec **out);
int tsc_vec_destroy(tsc_vec *v);
int tsc_vec_upload(tsc_vec *v, const double *host_c128);
int tsc_vec_download(const tsc_vec *v, double *host_c128);
/* Re<u|v> as in wave_product (block.f90:1092-1108: real part only) */
int tsc_vec_dot_re(const tsc_vec *u, const tsc_vec *v, double *out);
int tsc_vec_norm(const tsc_vec *v, double *out);
/* device pointer of a vector (complex128, n elements) */
void *tsc_vec_dptr(tsc_vec *v);
/* v = seeded complex Gaussians (same generator as the operator fill) */
int tsc_vec_fill_gaussian(tsc_vec *v, uint64_t seed, uint64_t key, double sigma);
/* y = a*x + b*y (real a, b), elementwise */
int tsc_vec_axpby(tsc_vec *y, double a, const tsc_vec *x, double b);
/* y = x (same length) */
int tsc_vec_copy(tsc_vec *y, const tsc_vec *x);
/*
 * Element ranges, asynchronous on the context stream (host arrays must stay
 * valid until tsc_ctx_sync): y[y_off, y_off+n) = x[x_off, x_off+n) on the
 * device; v[off, off+n) = host; host = v[off, off+n). Block operators live
 * in device arenas between DMRG steps (the reference keeps them in
 * Total_Model%sub(:)%mat, common.f90:84-174); these move sub-ranges of them.
 */
int tsc_vec_copy_range(tsc_vec *y, int64_t y_off, const tsc_vec *x, int64_t x_off, int64_t n);
int tsc_vec_upload_range(tsc_vec *v, int64_t off, const double *host_c128, int64_t n);
int tsc_vec_download_range(const tsc_vec *v, int64_t off, double *host_c128, int64_t n);

/*
 * ---- grouped two-stage contraction plan (the H.psi engine) --------------
 * Stage 1 writes intermediates into the plan's MID buffer (mid_elems
 * complex elements), stage 2 writes the output vector. The operator arena
 * (ops_elems complex128 values) is uploaded once and stays resident.
 * Replaces the per-call sector scans of model_wave (get_sys.f90:6915-7393)
 * and the kernels it dispatches (get_sys.f90:1783-5204).
 */
int tsc_plan_create(tsc_ctx *ctx, const double *ops_host_c128, int64_t ops_elems,
                    int64_t mid_elems, const tsc_contrib *c1, int32_t nc1,
                    const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                    int32_t nc2, const tsc_outblock *o2, int32_t no2,
                    int64_t in_elems, int64_t out_elems, tsc_plan **out);
/*
 * Same, with the operator arena borrowed from a device vector (ops->n
 * elements, not copied; it must outlive the plan). The plan's output vector
 * may be that same vector when the output windows do not overlap the
 * operator blocks read (used to merge operators in place, as hmerge does,
 * get_sys.f90:7398-7626).
 */
int tsc_plan_create_dev(tsc_ctx *ctx, const tsc_vec *ops, int64_t mid_elems, const tsc_contrib *c1,
                        int32_t nc1, const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                        int32_t nc2, const tsc_outblock *o2, int32_t no2, int64_t in_elems,
                        int64_t out_elems, tsc_plan **out);
int tsc_plan_destroy(tsc_plan *p);
/* replace the operator arena values (same layout) */
int tsc_plan_update_ops(tsc_plan *p, const double *ops_host_c128);
/*
 * Fill arena elements [off, off+count) with seeded complex Gaussians
 * (E|z|^2 = sigma^2) from a counter-based generator, on the device. Used
 * to build large synthetic superblocks without a host copy; with
 * ops_host_c128 = NULL at plan creation the arena starts zeroed.
 */
int tsc_plan_fill_ops_gaussian(tsc_plan *p, int64_t off, int64_t count, uint64_t seed,
                               uint64_t key, double sigma);
/* algorithmic FLOP count of one application (8 flops per complex MAC) */
int tsc_plan_flops(const tsc_plan *p, double *flops);
/* out = H in, on the context stream; precision = TSC_PREC_* */
int tsc_hpsi(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int precision);
/* same with host arrays (uploads, runs, downloads) */
int tsc_hpsi_host(tsc_plan *p, const double *in_c128, double *out_c128, int precision);
/*
 * out = H in in FP64 as one rocBLAS ZGEMM call per contribution, output
 * block by output block, on the context stream: the reference's own
 * per-product structure (model_wave's ZGEMM per operator pair and sector,
 * get_sys.f90:1783-5204), kept as the baseline the grouped kernels are
 * measured against. Plans with elementwise terms (TSC_OP_I) or
 * TSC_OP_CONJ operands are rejected with TSC_EINVAL.
 */
int tsc_hpsi_loop(tsc_plan *p, const tsc_vec *in, tsc_vec *out);

/*
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP stream capture, tight serving loops, harnesses
 * that forbid hipMalloc.
 * out = H in in FP64, as tsc_hpsi does, on the context stream (enqueued
 * only; tsc_ctx_sync waits for it), with the plan's 64x64 grouped FP64
 * tiles. in_c128 / out_c128 are device pointers to at least the plan's
 * input / output lengths of complex128 values, and must differ. The stage-1
 * intermediates and the partial tiles of split work units go to a static
 * device buffer of TSC_NOALLOC_SCRATCH_BYTES, which each device the library
 * runs on holds until the process ends, instead of plan-owned memory: the
 * call allocates nothing. A plan whose intermediates and partial tiles
 * (each part rounded up to 256 bytes) need more than that is rejected with
 * TSC_EINVAL before anything is launched. Every call on one device shares
 * the buffer: calls that overlap on one GPU (other streams, threads or
 * contexts) race.
 */
#define TSC_NOALLOC_SCRATCH_BYTES (64LL << 20)
int tsc_hpsi_noalloc(tsc_plan *p, const double *in_c128, double *out_c128);
/* capture one application in a hipGraph and replay it from then on */
int tsc_hpsi_graph(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int precision);
/* device time (ms) of the last stage-1 / stage-2 launches, HIP events */
int tsc_plan_last_times(const tsc_plan *p, float *ms_stage1, float *ms_stage2);

/*
 * Autotuner of the work units and tile shape: cuts the plan's units for
 * `precision` (F64 or BF16X3) with each candidate target of the unit
 * builder (2048, 4096, 8192 units: how finely long contribution lists are
 * split; FP64 also weighs its second tile shape, 64x64 grouped tiles
 * instead of 128x64 LDS units, reported as target 1), times H.psi with each
 * on in -> out (out is overwritten), and keeps the fastest. The choice is
 * cached per (precision, log2 of the plan's flops, log2 of its output
 * blocks, sector-size histogram: tenths of the stage-2 flops in output
 * blocks whose larger side is < 64, < 128, < 256, < 512, >= 512); a plan of
 * a cached shape takes it without timing unless flags & 1. *chosen = the
 * target kept (0 for F32, whose units are fixed). flags & 2: no timing, cut
 * the units for the target given in *chosen (1 selects the FP64 64x64
 * tiles). Results agree with the untuned plan to rounding (the split
 * changes the order of partial sums). Has no counterpart in the reference.
 */
int tsc_plan_autotune(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int precision, int flags, double *chosen);

/*
 * ---- Lanczos (lanczos_memo_new, eigen.f90:4-199) -------------------------
 * Krylov vectors stay resident; alpha/beta/dots are FP64; tridiagonal
 * solves run on the host in FP64. psi holds the start vector on entry and
 * the normalised Ritz vector on exit. precision selects H.psi products
 * for the Krylov build; refine_iters > 0 appends a Lanczos pass of that
 * many steps (at least 2) restarted from the mixed-precision Ritz vector,
 * whose first H.psi is FP64 (exact <psi|H psi> and residual) and whose
 * later products use `precision`.
 * info[0] = total H.psi calls, info[1] = outer restarts used.
 */
int tsc_lanczos(tsc_plan *p, tsc_vec *psi, double tol, int max_inner, int max_outer,
                int precision, int refine_iters, double *energy, int32_t *info);
/*
 * H.psi products of the last tsc_lanczos call on p that ran as replays of a
 * captured hipGraph. Plans below 2e11 flops per H.psi replay every product
 * (the graph is captured once per precision on two vectors of the plan, the
 * Krylov vectors are copied in and out); larger plans launch eagerly.
 * Diagnostic; has no counterpart in the reference.
 */
int tsc_lanczos_replays(const tsc_plan *p, int32_t *replays);
/*
 * Precision record of the last tsc_lanczos call on p:
 *   out[0] the Ritz value of the working-precision pass,
 *   out[1] <psi|H psi> of its normalised Ritz vector psi, FP64 (the
 *          refinement pass's first product),
 *   out[2] |H psi - out[1] psi|, FP64 (the residual norm of that vector),
 *   out[3] the energy after the refinement pass (what tsc_lanczos returned),
 *   out[4] the weight of the refined Ritz vector outside psi (1 - z_1^2 of
 *          the pass's normalised tridiagonal eigenvector).
 * out[1..4] are NaN when the call ran no refinement pass (FP64 precision or
 * refine_iters = 0). out[0] - out[1] is the energy error of the mixed
 * arithmetic at the mixed solution. The refined energy meets the mixed
 * products of the pass (all but its first) only through the directions
 * outside psi, so to first order its mixed-arithmetic error is out[4]
 * times that of a mixed Rayleigh quotient, |out[0] - out[1]| being a
 * measured sample of the latter. Diagnostic; has no counterpart in the
 * reference (lanczos_memo_new, eigen.f90:4-199, runs in FP64 only).
 */
int tsc_lanczos_refinement(const tsc_plan *p, double *out);  /* out[5] */

/*
 * ---- density matrix, eigensolve, truncation (den_mat.f90) ---------------
 * psi sectors are column-major (sys_dim x env_dim) blocks laid end to end.
 * side 0: rho_s = psi_s psi_s^H (sys, den_mat.f90:41-43);
 * side 1: rho_s = psi_s^H psi_s (env, den_mat.f90:89-91).
 * rho_out holds sum_s d_s^2 complex128 values (d_s = rows of rho_s).
 */
int tsc_density_matrix(tsc_ctx *ctx, const tsc_vec *psi, int32_t nsec,
                       const int32_t *sys_dim, const int32_t *env_dim, int side,
                       double *rho_out_c128);

/*
 * Get_truncation_operator (den_mat.f90:184-418) on the device:
 * per-sector Hermitian eigensolve of rho (ascending), global descending
 * sort (the reference's quicksort order), kept-state count from
 * kept_min / kept_max / trun_err, per-sector kept counts, and U (the top
 * kept eigenvectors of each sector, ascending order, column-major
 * d_s x kept_s, laid end to end over sectors with kept_s > 0).
 *   twos[s]        2S of sector s (entropy weights 2S+1)
 *   eigval_out     sum_s d_s eigenvalues, ascending per sector
 *   kept_out[s]    kept states of sector s
 *   u_out          sum_s d_s*kept_s complex128 values; u_cap is its
 *                  size in complex values, checked before anything is
 *                  written (kept_s can exceed kept_max: the rule keeps
 *                  kept_min states and then at least one more, and all of
 *                  them when n <= kept_min; sum_s d_s^2 always suffices)
 *   scalars_out    [0]=num_state [1]=discarded weight [2]=entropy (kept
 *                  spectrum, Entropys) [3]=entropy1 (U(1)-resolved)
 */
int tsc_truncation_operator(tsc_ctx *ctx, int32_t nsec, const int32_t *dim,
                            const int32_t *twos, const double *rho_c128,
                            int32_t kept_min, int32_t kept_max, double trun_err,
                            double *eigval_out, int32_t *kept_out, double *u_out_c128,
                            int64_t u_cap, double *scalars_out);

/*
 * Get_density_matrix_sys/_env followed by Get_truncation_operator without
 * a host round trip of rho (den_mat.f90:4-99 then :184-418): rho is built
 * and diagonalised on the device from the resident psi; only eigenvalues
 * and the kept eigenvectors come back. Arguments as in the two calls above
 * (dims per sector from sys_dim or env_dim according to side).
 */
int tsc_rho_truncate(tsc_ctx *ctx, const tsc_vec *psi, int32_t nsec, const int32_t *sys_dim,
                     const int32_t *env_dim, const int32_t *twos, int side, int32_t kept_min,
                     int32_t kept_max, double trun_err, double *eigval_out, int32_t *kept_out,
                     double *u_out_c128, int64_t u_cap, double *scalars_out);

/*
 * Wall milliseconds of the last tsc_rho_truncate / tsc_truncation_operator
 * on this context: [0] rho (device allocation and build, or upload),
 * [1] per-sector eigensolves, [2] eigenvalues to the host and the kept-state
 * selection, [3] kept eigenvectors to the host. A measurement hook of
 * Get_truncation_operator's stages (den_mat.f90:229-398); no counterpart in
 * the reference.
 */
int tsc_rho_last_times(const tsc_ctx *ctx, float *ms4);

/*
 * The kept-state selection of Get_truncation_operator alone
 * (den_mat.f90:255-364), host code that needs no device: from the
 * eigenvalues of every sector (ascending per sector, laid end to end) it
 * sorts the spectrum in the reference's quicksort order and returns the
 * per-sector kept counts and scalars_out as tsc_truncation_operator does.
 * kept_min, kept_max and trun_err must be >= 0; the count never passes
 * the end of the spectrum.
 */
int tsc_select_states(int32_t nsec, const int32_t *dim, const int32_t *twos, const double *eigval,
                      int32_t kept_min, int32_t kept_max, double trun_err, int32_t *kept_out,
                      double *scalars_out);

/*
 * update_trun_dia / update_trun_ndia (block.f90:3465-3683):
 * eff_i = U_lhs(i)^H . O_i . U_rhs(i) for nblk operator sub-blocks.
 * For each i: O_i is row_i x col_i at ori_off[i] (complex elements in
 * ori_c128), U_lhs is row_i x kl_i at ul_off[i], U_rhs is col_i x kr_i at
 * ur_off[i] (offsets into u_c128), result kl_i x kr_i at eff_off[i].
 */
int tsc_update_trun(tsc_ctx *ctx, int32_t nblk, const int32_t *row, const int32_t *col,
                    const int32_t *kl, const int32_t *kr, const int64_t *ori_off,
                    const int64_t *ul_off, const int64_t *ur_off, const int64_t *eff_off,
                    const double *ori_c128, int64_t ori_elems, const double *u_c128,
                    int64_t u_elems, double *eff_c128, int64_t eff_elems, int precision);

/*
 * ---- native SU(2) superblock planner (host) -----------------------------
 * Builds the grouped-contraction task arrays of model_wave for one
 * superblock [sys block, sys site | env block, env site] from block
 * descriptors, replacing the sector scans and quantum-number searches of
 * model_wave and its kernels (get_sys.f90:1783-5204, 6915-7393). Operator
 * matrices (reduced elements, bra x ket, column-major) live in one host
 * arena; blocks refer to them by offset. Sectors are (N, 2S, dim).
 * Term kinds: 1 = block operator ops[op], 2 = site operator op (0 c+,
 * 1 S, 2 n); tilde = use the annihilation tensor c~ derived from c+.
 * Every bond product is one term [X(sys side) x Y(env side)]^0 with coef;
 * block-site products inside one side are pairs (scalar on that side).
 */
typedef struct tsc_sector {
  int32_t n, twos, dim;
} tsc_sector;

typedef struct tsc_op_block {
  int32_t bra, ket; /* sector indices of the block the operator acts on */
  int64_t off;      /* arena offset of the bra_dim x ket_dim matrix     */
} tsc_op_block;

typedef struct tsc_op_desc {
  int32_t side; /* 0 sys block, 1 env block */
  int32_t up_dif, k2, nblk;
  const tsc_op_block *blk;
} tsc_op_desc;

typedef struct tsc_pair_desc {
  double coef;
  int32_t block_op, block_tilde, site_op, site_tilde;
} tsc_pair_desc;

typedef struct tsc_term_desc {
  double coef;
  int32_t l_kind, l_op, l_tilde, r_kind, r_op, r_tilde;
} tsc_term_desc;

typedef struct tsc_sb_desc {
  int32_t nsys, nenv;
  const tsc_sector *sys, *env;
  int64_t arena_len;
  int32_t nops;
  const tsc_op_desc *ops;
  const tsc_op_desc *sys_ham, *env_ham;
  int32_t nlpairs;
  const tsc_pair_desc *lpairs;
  int32_t nrpairs;
  const tsc_pair_desc *rpairs;
  int32_t nterms;
  const tsc_term_desc *terms;
  int32_t ntot, s2tot;
} tsc_sb_desc;

typedef struct tsc_tasks tsc_tasks;
int tsc_superblock_tasks(const tsc_sb_desc *desc, tsc_tasks **out);
/* counts: nc1, no1, nc2, no2, mid_len, psi_len, n_super, n_lsec, n_rsec */
int tsc_tasks_info(const tsc_tasks *t, int64_t *counts9);
/* any output pointer may be NULL; super_xy: (x, y, dx, dy) per psi sector,
 * lsec/rsec: (N, 2J, dim) of the enlarged sys/env bases */
int tsc_tasks_copy(const tsc_tasks *t, tsc_contrib *c1, tsc_outblock *o1, tsc_contrib *c2,
                   tsc_outblock *o2, int32_t *super_xy, int32_t *lsec, int32_t *rsec);
/*
 * flops2[0] = flops of one application as planned (8 per complex MAC, each
 * stage-1 intermediate once; the planner forms O_L psi or psi O_R^T first,
 * per source window and term, whichever is cheaper with the window's
 * intermediates shared); flops2[1] = the algorithmic count with every
 * (L transition, R transition, source window) product charged at its own
 * cheaper order and nothing shared. The reference's block-block kernel
 * chooses the order per product by shape in the same way
 * (get_sys.f90:3385-3408).
 */
int tsc_tasks_flops(const tsc_tasks *t, double *flops2);
int tsc_tasks_plan(tsc_ctx *ctx, const tsc_tasks *t, const double *arena_c128, int64_t arena_len,
                   tsc_plan **plan);
/* the same plan over an operator arena already on the device (borrowed) */
int tsc_tasks_plan_dev(tsc_ctx *ctx, const tsc_tasks *t, const tsc_vec *arena, tsc_plan **plan);
int tsc_tasks_destroy(tsc_tasks *t);

/*
 * ---- block models: model_wave / hmerge and Get_Hamiltonian_Sys/Env -------
 * A block model is what the reference keeps in Total_Model
 * (common.f90:84-174): the block's sectors, its Hamiltonian and, for every
 * site with bonds leaving the block, that site's c+, S and n on the block
 * (reduced elements, bra x ket column-major in one host arena). Bonds are
 * the lattice's (i, j, t, J): hopping t, exchange J (the reference's
 * jt / jd / jn, dmrg_sub.f90:453-540).
 */
typedef struct tsc_site_ops {
  int32_t site;
  const tsc_op_desc *c, *S, *n; /* up_dif/k2: c+ 1/1, S 0/2, n 0/0 */
} tsc_site_ops;

typedef struct tsc_model_desc {
  int32_t nsec;
  const tsc_sector *sec;
  int32_t nsites;
  const int32_t *sites; /* lattice sites of the block */
  const tsc_op_desc *ham;
  int32_t nkeep;
  const tsc_site_ops *keep;
} tsc_model_desc;

typedef struct tsc_bond {
  int32_t i, j;
  double t, J;
} tsc_bond;

/*
 * model_wave's task arrays for the superblock [sys, s1 | env, s2]
 * (get_sys.f90:6915-7393) with hmerge's merging of partner operators
 * (get_sys.f90:7398-7626): the bonds are classified by the parts they join,
 * the operators of the partner sites of each class are summed with their
 * couplings into merged operators (appended to a copy of the arena held by
 * the tasks), and the planner above builds the contraction. Env-block
 * operators act as psi . O^T (the conjugates of the reference's, which
 * apply psi . O^H, get_sys.f90:1869). tsc_tasks_plan(ctx, t, NULL, 0, &p)
 * uploads the tasks' own arena.
 */
int tsc_model_tasks(const tsc_model_desc *sys, const tsc_model_desc *env, int32_t s1, int32_t s2,
                    int32_t nbonds, const tsc_bond *bonds, int32_t ntot, int32_t s2tot,
                    const double *arena_c128, int64_t arena_len, tsc_tasks **out);

/*
 * Get_Hamiltonian_Sys / _Env followed by update_trun_dia / _ndia
 * (get_sys.f90:4-592, block.f90:3465-3683) as one contraction plan: the
 * enlarged block blk (x) new_site (sectors in Get_basis order,
 * block.f90:2087-2189) with its Hamiltonian (block H + the merged
 * block-site bonds) and the c+, S, n of every site of keep_sites, each
 * truncated as U_x2^H O U_x. The truncation keeps the enlarged sectors
 * order[0..norder) (the new block's sector order) with kept[q] states;
 * the plan's input is U_x (dim_x x kept columns, column-major) laid end to
 * end in that order, its output the new block's operators. Their layout
 * comes from tsc_tasks_layout.
 */
int tsc_renorm_tasks(const tsc_model_desc *blk, int32_t new_site, int32_t nbonds, const tsc_bond *bonds,
                     int32_t nkeep, const int32_t *keep_sites, int32_t norder, const int32_t *order,
                     const int32_t *kept, const double *arena_c128, int64_t arena_len, tsc_tasks **out);

/*
 * Output layout of renormalisation tasks: counts3 = (windows, new sectors,
 * enlarged sectors); layout[3w..3w+2] = (operator, bra, ket) of window w
 * at output offset offsets[w] (operator -1: Hamiltonian; 3k + 0/1/2: c+, S,
 * n of keep_sites[k]; bra/ket: new sector indices); new_sec = (N, 2S,
 * kept) per new sector; enl_sec = (N, 2S, dim) per enlarged sector. Any
 * output pointer may be NULL.
 */
int tsc_tasks_layout(const tsc_tasks *t, int64_t *counts3, int32_t *layout, int64_t *offsets,
                     int32_t *new_sec, int32_t *enl_sec);
/* lens3 = (input length, output length, length of the tasks' own arena) */
int tsc_tasks_lengths(const tsc_tasks *t, int64_t *lens3);
/* the tasks' own arena (lens3[2] complex128 values) */
int tsc_tasks_copy_arena(const tsc_tasks *t, double *arena_c128);

/* Racah 6j symbol with doubled arguments (W6JS, block.f90:3744-3795) */
double tsc_w6j(int32_t j1, int32_t j2, int32_t j3, int32_t l1, int32_t l2, int32_t l3);

#ifdef __cplusplus
}
#endif
#endif /* TSC_HOTPATH_H */
