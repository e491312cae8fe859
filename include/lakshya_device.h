/*
 * Device layer of the MI355X interior-point path: a C ABI over the HIP runtime and the
 * hand-written gfx950 kernels.  Host C++ (the solver classes) includes only this header, never
 * a HIP header; every argument is a plain pointer, size or scalar.  Device pointers are plain
 * `void*`/typed pointers into device memory, `stream` is an opaque hipStream_t (NULL = default).
 *
 * Return value of every function: 0 on success, otherwise a HIP error code (lkd_error_string).
 * Kernel entry points take a precision selector where the panel storage type varies:
 *   LKD_F64 = 0 (double panels), LKD_F32 = 1 (float panels).  Vectors are always double.
 *
 * There is no counterpart of this layer in the reference (src/ipm/mehrotra_device.cu and
 * src/cuda/kkt.cu launch their kernels inline); each kernel cites the reference kernel or
 * library call it replaces.
 */
#ifndef LAKSHYA_DEVICE_H
#define LAKSHYA_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LKD_F64 0
#define LKD_F32 1

/* ---------------------------------------------------------------- runtime plumbing */
const char* lkd_error_string(int code);
int lkd_device_count(int* count);
int lkd_device_cus(int* cus);          /* compute units of the current device */
int lkd_set_device(int dev);
int lkd_malloc(void** ptr, size_t bytes);
int lkd_free(void* ptr);
int lkd_host_alloc(void** ptr, size_t bytes);     /* pinned host memory */
int lkd_host_free(void* ptr);
int lkd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int lkd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int lkd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int lkd_memset(void* dst, int value, size_t bytes, void* stream);
int lkd_stream_create(void** stream);
int lkd_stream_destroy(void* stream);
int lkd_stream_sync(void* stream);
int lkd_event_create(void** ev);
int lkd_event_destroy(void* ev);
int lkd_event_record(void* ev, void* stream);
int lkd_event_elapsed_ms(float* ms, void* start, void* stop);
int lkd_fill_f64(double* dst, double value, int64_t n, void* stream);   /* kernel fill */
int lkd_wallclock_khz(int* khz);   /* rate of the device wall clock (wall_clock64) */
/* streaming copy kernel (n even, 16-byte aligned): calibrates HBM traffic counters */
int lkd_copy_f64(double* dst, const double* src, int64_t n, void* stream);
/* Queue completeness after launches that deal tasks from per-XCD queues: for every queue q < nq
 * and XCD x < nx, counter cnt[q nx + x] (tickets taken) must have reached the queue's length
 * qoff[q (nx + 1) + x + 1] - qoff[q (nx + 1) + x]; otherwise `bit` is OR-ed into *err. */
int lkd_queue_check(const uint32_t* cnt, const int32_t* qoff, int32_t nq, int32_t nx,
                    int32_t bit, int32_t* err, void* stream);

/* ---------------------------------------------------------------- supernodal plan (device) */
/* One work item of the factorization / triangular-solve schedule.  Field meaning per kernel:
 *   off    panel offset (elements) of the supernode, ld = panel rows, ncol = panel columns
 *   k0/kb  block column start / width inside the panel
 *   r0/c0  tile row / column start inside the panel (kernel specific)
 *   kbeg/kend  K range of an UPDATE task, of a PART task, and of a STEP tile's own GEMM
 *   ws     DIAG / STEP: staging slot offset (STEP reads it)
 *   nsplit STEP: 1 on the tile that also factors the next diagonal block
 *   sn     supernode id, first = first (permuted) column of the supernode
 *   np/px  STEP: number of partial products (from the PART launch before it) added to the
 *          tile's own GEMM, stored at work + px + s * LKD_PART (s < np); PART: px = where its
 *          product goes
 *   prio   STEP: 1 on tiles of the critical path (their wavefronts get issue priority)
 *   part   STEPS: 1 on a PART task (product of K range kbeg..kend into work + px)          */
typedef struct {
    int64_t off;
    int64_t aux;     /* diagonal-copy offset (DIAG / STEP next block / DIAGWB / TRSV) */
    int64_t aux2;    /* row-list offset (UPDATE / TRSV); STEP: staging slot of the next block */
    int64_t ws;
    int32_t ld, ncol;
    int32_t k0, kb;
    int32_t r0, c0;
    int32_t kbeg, kend;
    int32_t nsplit, sn;
    int32_t first, np;
    int64_t px;
    int32_t prio, part;
} lkd_task;

typedef struct {
    int32_t nsuper;
    const int32_t* sn_first;    /* [nsuper+1] */
    const int64_t* sn_rowptr;   /* [nsuper+1] */
    const int32_t* sn_rows;     /* permuted row indices, sorted per supernode */
    const int64_t* sn_panel;    /* [nsuper+1] panel offsets */
    const int32_t* col_to_sn;   /* [m] */
} lkd_snplan;

/* Load balance of sparse row passes: rows (or pair-map segments) longer than the threshold are
 * listed once at set-up and each handled by a whole 256-thread block, in a fixed order.  Row
 * passes may split listed rows further: with split > 1, a pre-pass sums each listed row k in
 * `split` pieces of equal length (one block each) into part[k * split + p], and the pass adds
 * the pieces in order (about LKD_LONG_PIECE entries per piece, at most LKD_LONG_MAX_SPLIT). */
#define LKD_LONG_ROW 512
#define LKD_LONG_SEG 64
#define LKD_LONG_PIECE 1024
#define LKD_LONG_MAX_SPLIT 64
typedef struct {
    int32_t n;              /* number of listed rows / segments */
    const int32_t* idx;     /* their indices (device) */
    int32_t split;          /* row passes: pieces per listed row (0 or 1: none) */
    double* part;           /* row passes, split > 1: n * split piece sums (device scratch) */
} lkd_long;

/* ---------------------------------------------------------------- normal-equations assembly */
/* The pair map: every product a_ik * a_jk (i >= j, permuted rows, k a column of A) that adds
 * to the lower triangle of M0 = A Theta A^T, grouped by destination entry (segments, in panel
 * order).  Pair q stores the static product pp[q] = a_ik * a_jk and its column pk[q] = k, so an
 * entry is  M0_ij = sum_q pp[q] * theta[pk[q]]  and theta is the only value gathered per pair.
 * This replaces the Theta factor and row-pair loop of compute_M_numerics_kernel
 * (src/cuda/kkt.cu:177-215). */

/* A segment is summed in a fixed order that depends only on its length L:
 *   L <= LKD_SHORT_SEG: one thread, acc = fma(pp[q], theta[pk[q]], acc) over q in order;
 *   L <= LKD_LONG_SEG:  16 lanes, lane l sums q = beg+l, beg+l+16, ... in order, then a fixed
 *                       xor butterfly over the 16 lanes;
 *   longer:             a 256-thread block, thread-strided fma partials, then a fixed tree.
 * The mid and long classes are listed once at set-up (segment indices for the panel pass, row
 * indices for the diagonal pass). */
#define LKD_SHORT_SEG 4

/* Diagonal of M0 per permuted row from its pair-map segment (the same fixed order as
 * lkd_assemble_pairs, so the two agree bitwise): dseg[i] = segment of (i,i) or -1 (empty row);
 * mid_diag / long_diag list the rows whose diagonal segment is in the mid / long class.
 * D_i = 1/sqrt(max(|M0_ii|, 1e-12)) when equilibrate != 0, else 1 (src/cuda/kkt.cu:65-79). */
int lkd_assemble_diag(int32_t m, const int64_t* dseg, const int64_t* segptr, const double* pp,
                      const int32_t* pk, const double* theta, int equilibrate, double* D,
                      double* m0diag, lkd_long mid_diag, lkd_long long_diag, void* stream);
/* The whole panel (panel_elems entries, precision prec) from the M0 entries of the lower
 * triangle:
 *   mval[t] = M0(segment t) * D[segi[t]] * D[segj[t]] (+ delta on the diagonal of non-empty
 *             rows), each segment summed as above (mid_segs / long_segs list the segments of the
 *             mid / long class; mval: nseg device doubles of scratch);
 *   then panel chunk c (entries [c, c+1) * LKD_PANEL_CHUNK) is built in LDS, zeros plus the
 *   entries of segments chunk_seg[c] .. chunk_seg[c+1]-1 at segdst[t], and stored whole.
 * Writing each chunk once replaces a separate zero fill and scattered writes into HBM.
 * Replaces compute_M_numerics_kernel + build_equilibrated_M_kernel / add_regularization_kernel
 * (src/cuda/kkt.cu:81-94,157-171,177-215). */
#define LKD_PANEL_CHUNK 4096
int lkd_assemble_pairs(int prec, int64_t nseg, const int64_t* segptr, const int64_t* segdst,
                       const int32_t* segi, const int32_t* segj, const double* pp,
                       const int32_t* pk, const double* theta, const double* D,
                       const uint8_t* empty_row, double delta, lkd_long mid_segs,
                       lkd_long long_segs, double* mval, int64_t panel_elems,
                       const int64_t* chunk_seg, void* panel, void* stream);

/* ---------------------------------------------------------------- supernodal Cholesky */
/* Replaces cholesky_level_kernel and its level loop (src/cuda/kkt.cu:217-285,582-596).
 * Left-looking per 64-column block column, per etree level:
 *   DIAG    factor block column 0's diagonal block (ws = staging slot, aux = diag copy)
 *   PART    split K of a STEP launch with few tiles: one K range [kbeg, kend) of a tile's
 *           GEMM, written to work + px (LKD_PART elements per product, the tile's thread
 *           layout); in the same launch, tasks with nsplit = 2 factor the step's diagonal
 *           block k0 (as DIAG) for the one-launch-per-step form
 *   PSUM    (after PART) the np partial products of a tile summed in order into the first,
 *           task kbeg = slice of 512 elements (LKD_PART / 512 tasks per tile)
 *   STEP k  128-row tiles below the diagonal: GEMM update (own range kbeg..kend plus np
 *           partial products) + solve against the staged factor; in a STEPS launch the tile
 *           holding the next diagonal block (nsplit = 1, c0 rows) also factors it (staging
 *           slot aux2, diag copy aux)
 *   STEPS   (in place of the STEP / PART launches) every STEP tile of the level in one
 *           launch: a tile runs once the tiles it depends on are done (see lkd_chol_step_dag)
 *   UPDATE  ancestors (split K over kbeg..kend)
 * and DIAGWB (diagonal blocks into the panel) after the last level.
 * err: device int, bit 0 set when a pivot is not positive, bit 1 when an update falls outside
 * the L structure, bit 5 when a STEPS wait exceeded its poll limit; kernels return early once
 * set.
 * diag: row-major copies of the factored diagonal blocks (element offsets), with the
 * reciprocal on the diagonal, read by TRSV.  piv: the pivots L_jj per permuted column.
 * stage: staged diagonal factors, LKD_STAGE elements per slot (all panel precision). */
#define LKD_STAGE (64 * 64 + 4 * 256)
#define LKD_PART (128 * 64)
int lkd_chol_diag(int prec, int32_t ntask, const lkd_task* tasks, const void* panel, void* diag,
                  void* piv, void* stage, int32_t* err, void* stream);
int lkd_chol_part(int prec, int32_t ntask, const lkd_task* tasks, const void* panel, void* diag,
                  void* piv, void* stage, void* work, int32_t* err, void* stream);
int lkd_chol_psum(int prec, int32_t ntask, const lkd_task* tasks, void* work, void* stream);
int lkd_chol_step(int prec, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
                  void* piv, void* stage, const void* work, int32_t* err, void* stream);
/* STEPS: the STEP tiles of one level (np = 0).  deps: 4 task indices per task (-1 = none):
 * the tiles of the previous block column whose rows overlap its rows, and the tile that
 * factored its diagonal block (whose rows its GEMM reads); sdeps: that last tile (-1: block
 * staged before the launch).  Every supernode's tasks belong to one of nx queues (one per XCD):
 * queue x = queue[qoff[x] .. qoff[x+1]) (task indices, every task after its dependencies).
 * grid workgroups (at least nx times the longest queue when workgroups go to the XCDs in turn)
 * each read the XCD they run on (XCC_ID mod nx) and take tasks of its queue from counter
 * ctl[1 + q nx + x] until it is empty.  A task waits until flags[dep] == ctl[0] for each
 * dependency, runs its GEMM, waits until flags2[sdep] == ctl[0], solves, then sets
 * flags[task] = ctl[0]; a tile that factors the next diagonal block sets flags[task] once its
 * rows are stored and flags2[task] once the factor is staged.  A task with part = 1 is a PART
 * (its product over kbeg..kend into work + px, then flags[task]); a tile with np = 1 adds the
 * product at work + px (its PART is among its dependencies).  Staging slots must not be
 * shared by two block columns of the level.  lkd_chol_dag_begin(ctl, nq) starts a
 * factorization: ctl[0] += 1 and the nq counters ctl[1 .. nq] = 0. */
int lkd_chol_step_dag(int prec, int32_t ntask, const lkd_task* tasks, const int32_t* deps,
                      const int32_t* sdeps, const int32_t* queue, const int32_t* qoff,
                      int32_t nx, int32_t grid, uint32_t* flags, uint32_t* flags2, uint32_t* ctl,
                      int32_t q, void* panel, void* diag, void* piv, void* stage, void* work,
                      int32_t* err, void* stream);
int lkd_chol_dag_begin(uint32_t* ctl, int32_t nq, void* stream);
int lkd_chol_update(int prec, int32_t ntask, const lkd_task* tasks, void* panel,
                    lkd_snplan plan, int32_t* err, void* stream);
int lkd_chol_diagwb(int prec, int32_t ntask, const lkd_task* tasks, void* panel, const void* diag,
                    const void* piv, void* stream);

/* ---------------------------------------------------------------- triangular solves */
/* Replaces cusparseSpSV (L and L^T) (src/cuda/kkt.cu:731-734,788-789).  x is the permuted
 * vector, solved in place: L z = x (dir 0, levels leaves first) or L^T x = z (dir 1, root
 * first), one launch per etree level, g = task.c0 workgroups of nw (4 or 8) wavefronts per
 * supernode (task.r0 = workgroup index; g = 1, or with nw = 8: 4 for forward launches, 2 for
 * backward ones).  With
 * g = 4, grid workgroups each read the XCD they run on (XCC_ID mod nx) and take their task from
 * that XCD's queue, queue[qoff[x] .. qoff[x+1]) (task indices; a supernode's g tasks
 * consecutive), through counter cnt[x]; grid >= nx times the longest queue.  The workgroups of
 * a supernode then share an L2 for their hand-offs.
 * task.aux = offset of the supernode's first diagonal copy, task.aux2 = row-list offset,
 * task.kbeg = offset of its rows below in `below` (scratch, one double per row below a
 * supernode), task.ws = offset of its area in `pbuf` (nb * G * nw * 64 doubles, nb = 64-column
 * blocks of the supernode: forward z exchange, backward partial sums), task.px = offset of its
 * nb + nb * G flags in `flags`.  *epoch is the solve number, advanced by lkd_trsv_epoch before
 * every solve (flags hold it, so they are never reset); lkd_trsv_epoch also zeroes the ncnt
 * queue counters cnt.  Waits are bounded: a timeout sets err bit 0x10 and the result is
 * invalid. */
int lkd_trsv_epoch(uint32_t* epoch, uint32_t* cnt, int32_t ncnt, void* stream);
int lkd_trsv_level(int prec, int dir, int nw, int g, int32_t ntask, const lkd_task* tasks,
                   const void* panel, const void* diag, const int32_t* sn_rows, double* x,
                   double* below, double* pbuf, uint32_t* flags, const uint32_t* epoch,
                   const int32_t* queue, const int32_t* qoff, int32_t nx, uint32_t* cnt,
                   int32_t grid, int32_t* err, void* stream);
/* Wide level solve (dir 0 forward, 1 backward) for levels of few supernodes with many blocks:
 * one workgroup per row tile (forward) or block column (backward) of each supernode, tasks
 * taken in the order `order` from the ticket counter cnt (reset by lkd_trsv_epoch); task r0 =
 * tile / block index, px = the supernode's flags (nb forward, then nb backward).  Blocks are
 * handed between workgroups on any XCD through device-coherent stores and flags holding the
 * epoch.  Same error bits as lkd_trsv_level. */
int lkd_trsv_wide(int prec, int dir, int32_t ntask, const lkd_task* tasks, const int32_t* order,
                  uint32_t* cnt, const void* panel, const void* diag, const int32_t* sn_rows,
                  double* x, uint32_t* flags, const uint32_t* epoch, int32_t* err, void* stream);

/* ---------------------------------------------------------------- KKT solve vector kernels */
/* r[i] = rhs[P[i]]; req[i] = D[i] * r[i] */
int lkd_kkt_gather(int32_t m, const int32_t* P, const double* rhs, const double* D, double* r,
                   double* req, void* stream);
/* dy[i] = (accumulate ? dy[i] : 0) + D[i] * z[i];  optionally backup[i] = old dy[i] */
int lkd_kkt_update(int32_t m, const double* D, const double* z, double* dy, double* backup,
                   int accumulate, void* stream);
/* r = rhs_perm - A_p (theta o (A_p^T dy)), req = D r; long_rows / long_cols list the rows of A_p
 * and the columns with more than LKD_LONG_ROW entries.  kS (device doubles) receives
 * kS[0] = max|r|, kS[1] = max|rhs_perm|, kS[2] = max|dy|, kS[3] = max M0_ii (m0d); block
 * partials use kS + LKD_KS_PART.  The FP64 residual against the unregularised M0
 * (src/cuda/kkt.cu:110-123, computed without storing M0). */
#define LKD_KS_SLOTS 16
#define LKD_KS_PART 16
#define LKD_MAX_NBLK 4096   /* blocks of a reducing pass: partials need q * LKD_MAX_NBLK doubles */
int lkd_kkt_residual(int32_t m, int32_t n, const int32_t* colptr, const int32_t* prow,
                     const double* cval, const int32_t* prp, const int32_t* pci,
                     const double* pval, const double* theta, const double* dy,
                     const double* rhs_perm, const double* m0d, double* u, double* r,
                     const double* D, double* req, double* kS, lkd_long long_rows,
                     lkd_long long_cols, void* stream);
/* Preconditioned CG pieces on the permuted normal equations (kS slots 4..15):
 *   matvec:  q = M0 p (two passes), kS[pq] = p^T q
 *   update:  a = kS[rz]/kS[pq]; dy += a p; r -= a q; req = D r; kS[4] = max|r|, kS[2] = max|dy|
 *   precdot: z <- D z (after the triangular solves), kS[rz_out] = r^T z
 *   dir:     p = z + (kS[rz_new]/kS[rz_old]) p */
int lkd_pcg_matvec(int32_t m, int32_t n, const int32_t* colptr, const int32_t* prow,
                   const double* cval, const int32_t* prp, const int32_t* pci,
                   const double* pval, const double* theta, const double* p, double* u,
                   double* q, double* kS, int32_t slot_pq, lkd_long long_rows, lkd_long long_cols,
                   void* stream);
int lkd_pcg_update(int32_t m, double* kS, int32_t slot_rz, int32_t slot_pq, const double* p,
                   const double* q, double* dy, double* r, const double* D, double* req,
                   void* stream);
int lkd_pcg_precdot(int32_t m, const double* D, double* z, const double* r, double* kS,
                    int32_t slot_rz, void* stream);
int lkd_pcg_dir(int32_t m, const double* kS, int32_t slot_new, int32_t slot_old,
                const double* z, double* p, void* stream);
/* out[P[i]] = dy[i] (src/cuda/kkt.cu:105-108) */
int lkd_kkt_scatter(int32_t m, const int32_t* P, const double* dy, double* out, void* stream);
/* gather L (panel storage) into CSR values in the reference's L layout (diagonal last). */
int lkd_panel_to_csr(int prec, int64_t nnz, const int64_t* src_index, const void* panel,
                     double* vals, void* stream);

/* ---------------------------------------------------------------- IPM vector kernels */
/* Device scalar slots (doubles) shared by the IPM kernels. */
enum {
    LKS_RP2 = 0, LKS_RD2, LKS_XS, LKS_CX, LKS_BY, LKS_RPINF, LKS_RDINF,       /* residuals */
    LKS_R1, LKS_R2, LKS_R3, LKS_RXS2, LKS_APMAX, LKS_ADMAX, LKS_DX2, LKS_DS2, /* direction */
    LKS_MUAFF, LKS_SIGMAMU, LKS_AP, LKS_AD, LKS_DY2, LKS_KKTR,
    LKS_XSMAX,                                          /* max_j |x_j (s_j - rd_j)| */
    LKS_THMAX, LKS_THMIN,                               /* max_j, min_j x_j / s_j */
    LKS_NSLOTS
};
#define LKD_MAX_TRIALS 32
/* line-search trial outputs: 6 values per trial after the fixed slots */
#define LKS_TRIAL_BASE 32
#define LKS_TRIAL_STRIDE 8
#define LKS_TOTAL (LKS_TRIAL_BASE + LKS_TRIAL_STRIDE * LKD_MAX_TRIALS)

typedef struct {
    int32_t m, n;
    const int32_t *arp, *aci; const double* aval;     /* A in CSR (scaled) */
    const int32_t *trp, *tci; const double* tval;     /* A^T in CSR (= CSC of A) */
    const double *b, *c, *R, *C;                      /* scaled data, scale vectors */
    int32_t grp_rows, grp_cols;                       /* lanes per row of A / of A^T */
    lkd_long long_rows, long_cols;                    /* rows / columns > LKD_LONG_ROW entries */
    double* part;                                     /* block partials scratch */
    int32_t nblk_rows, nblk_cols;                     /* blocks used for row / column passes */
    double* S;                                        /* scalar slots [LKS_TOTAL] */
} lkd_ipm;

/* rp = A x - b, rd = A^T y + s - c, theta = x / s; slots RP2 RD2 XS CX BY RPINF RDINF
 * (compute_residuals + compute_norm/compute_mu/compute_objective, mehrotra_device.cu:589-624) */
int lkd_ipm_residuals(const lkd_ipm* P, const double* x, const double* y, const double* s,
                      double* rp, double* rd, double* theta, void* stream);
/* r_xs = -x s + S[sigma_slot] - (dxa ds a if dxa) ; v = -theta rd - r_xs / s ; rkkt = -rp + A v
 * (compute_r_xs_kernel, compute_v_kernel, compute_rkkt: mehrotra_device.cu:111-126,627-638) */
int lkd_ipm_rhs(const lkd_ipm* P, const double* x, const double* s, const double* theta,
                const double* rd, const double* rp, const double* dxa, const double* dsa,
                double sigma_mu, double* rxs, double* v, double* rkkt, void* stream);
/* ds = -rd - A^T dy ; dx = -theta ds + r_xs / s ; qd = A^T dy + ds ; slots R2 R3 RXS2 APMAX
 * ADMAX DX2 DS2 DY2; qp = A dx ; slot R1 = ||rp + A dx||^2
 * (compute_ds, compute_dx_kernel, verify_newton_direction, step_length_functor:
 *  mehrotra_device.cu:128-133,168-176,494-587,641-660) */
int lkd_ipm_direction(const lkd_ipm* P, const double* x, const double* s, const double* y,
                      const double* theta, const double* rd, const double* rp,
                      const double* rxs, const double* dy, double* dx, double* ds, double* qd,
                      double* qp, void* stream);
/* slot MUAFF = (x + ap dxa)^T (s + ad dsa)  (mehrotra_device.cu:831-835) */
int lkd_ipm_mu_aff(const lkd_ipm* P, const double* x, const double* s, const double* dxa,
                   const double* dsa, double ap, double ad, void* stream);
/* the same with ap = min(1, slot APMAX), ad = min(1, slot ADMAX) taken on the device, so the
 * affine step needs no host round trip (mehrotra_device.cu:824-835 computes them on the host) */
int lkd_ipm_mu_aff_step(const lkd_ipm* P, const double* x, const double* s, const double* dxa,
                        const double* dsa, void* stream);
/* All backtracking trials t = 0..ntrial-1 with (ap, ad) * 0.5^t evaluated in one pass:
 * per trial ||rp + ap_t qp||^2, ||rd + ad_t qd||^2, (x + ap_t dx)^T (s + ad_t ds),
 * min(x + ap_t dx), min(s + ad_t ds), ||y + ad_t dy||^2   (mehrotra_device.cu:920-995) */
int lkd_ipm_trials(const lkd_ipm* P, const double* x, const double* s, const double* y,
                   const double* dx, const double* ds, const double* dy, const double* rp,
                   const double* rd, const double* qp, const double* qd, double ap, double ad,
                   int32_t ntrial, void* stream);
/* x += ap dx ; s += ad ds ; y += ad dy  (mehrotra_device.cu:1000-1005) */
int lkd_ipm_commit(const lkd_ipm* P, double* x, double* s, double* y, const double* dx,
                   const double* ds, const double* dy, double ap, double ad, void* stream);
/* x = 1, s = 1, y = 0 (init_vars_kernel, mehrotra_device.cu:93-102) */
int lkd_ipm_init(int32_t m, int32_t n, double* x, double* y, double* s, void* stream);

/* ---------------------------------------------------------------- batched dense IPM */
/* Thousands of small LPs (min c^T x, A x = b, x >= 0) in one persistent launch: every
 * workgroup takes the next LP from a ticket counter (order[] gives the visit order) and runs
 * up to iters_per_visit IPM iterations on it -- Ruiz scaling, dense A Theta A^T assembly,
 * MFMA Cholesky, triangular solves, PCG and the Mehrotra step -- with all vectors of the LP
 * in its own work area.  LP state survives between launches in lkb_state.
 *
 * Layout of one LP (lkb_lp):
 *   idx[a ..]  CSR row pointers (m + 1), CSR columns (nnz), CSC position of each CSR entry
 *              (nnz), CSC column pointers (n + 1), CSC rows (nnz)             [int32]
 *   val0[v ..] original CSC values (nnz)
 *   valw[2 v ..] scaled CSR values (nnz), then scaled CSC values (nnz)
 *   rhs[bc ..] original b (m) then c (n)
 *   work[w ..] LKB_NVN n-vectors, LKB_NVM m-vectors, one staged 64x64 factor (LKB_STAGE)
 *   fac[f ..]  the dense m x ld factor (column-major, lower; doubles, or floats in the same
 *              bytes for an FP32 factor)
 *   x_out/s_out[on ..], y_out[om ..]  original-space solution */
#define LKB_MAX_M 1600
#define LKB_NVN 14
#define LKB_NVM 14
#define LKB_STAGE (64 * 65 + 1024)
typedef struct {
    int32_t m, n, nnz, ld;
    int64_t a, v, bc, w, f, on, om;
} lkb_lp;

typedef struct {
    int32_t phase;                 /* 0 not started, 1 iterating, 2 finished */
    int32_t status;                /* 0 Optimal, 3 IterationLimit */
    int32_t iterations, fail_code; /* fail_code as lk_ipm_result */
    int32_t factorizations, fp32_factorizations, solves, pcg_steps;
    double rp0, rd0, mu0, binf, cinf;
    double objective, dual_objective, rel_gap, rel_primal, rel_dual;
    double primal_residual, dual_residual, mu;
    /* wall-clock ticks spent by the LP's workgroup: scaling and start point, sparse passes and
     * vector work, assembly, dense factorization, dense triangular solves */
    double t_init, t_other, t_assemble, t_factor, t_solve;
    /* start point: 0 standard, 1 the warm point, 2 the warm point rejected (its original-space
     * relative infeasibility, warm_infeas, above warm_max_infeas) and the standard start taken */
    double warm_infeas;
    int32_t warm, warm_pad;
} lkb_state;

typedef struct {
    int32_t count, stop_rule, precision, max_iter;
    int32_t iters_per_visit, mixed_max_refine;     /* PCG budget with an FP32 factor */
    double tol, gap_tol, mixed_spread;
    const lkb_lp* lps;
    const int32_t* order;
    const int32_t* idx;
    const double* val0;
    double* valw;
    const double* rhs;
    double* work;
    void* fac;
    lkb_state* state;
    double *x_out, *y_out, *s_out;
    uint32_t* ctl;                 /* [0] ticket counter, [1] LPs finished */
    /* optional warm start (original space, the layout of x_out / y_out / s_out): LP i starts
     * from it when warm_flag[i] != 0 and its relative infeasibility
     * max(||A x0 - b||_inf / (1 + ||b||_inf), ||A^T y0 + s0 - c||_inf / (1 + ||c||_inf)) is at
     * most warm_max_infeas -- x = max(x0 / C, warm_floor), s = max(s0 C, warm_floor),
     * y = y0 / R after scaling, then, with warm_mu > 0, the smaller of x_j and s_j raised so
     * that x_j s_j >= warm_mu -- else from x = s = 1, y = 0 */
    const double *x0, *y0, *s0;
    const int32_t* warm_flag;
    double warm_floor, warm_mu, warm_max_infeas;
} lkb_args;

/* ctl[0] = 0; with `restart` also ctl[1] = 0 and the phase of LPs order[0 .. count) back to 0
 * (a new solve).  A kernel, not a memset, so a captured graph may hold it. */
int lkd_batch_reset(uint32_t* ctl, lkb_state* state, const int32_t* order, int32_t count,
                    int32_t restart, void* stream);
/* the persistent launch: `grid` workgroups of 256 threads; args (host struct) are passed to
 * the kernel by value */
int lkd_batch_solve(const lkb_args* args, int32_t grid, void* stream);

/* Crossover of every LP of a batch (one workgroup per LP, after a solve): from the interior
 * point x (x_out layout) to a basic solution.  The basis is built from the support of x (largest
 * values first; a Gauss-Jordan explicit inverse in the LP's factor slot), the remaining positive
 * nonbasic columns are purified along directions that do not increase c^T x, and x_B = B^-1 b.
 * Outputs: basis (y layout: the basic column of each row, -1 artificial), xv (x layout), status
 * (0 optimal vertex, 1 primal-feasible vertex that is not dual feasible, 2 rank deficient,
 * 3 unbounded direction, 4 basic solution infeasible) and pivots, per LP. */
typedef struct {
    const double* x;
    int32_t* basis;
    double* xv;
    int32_t* status;
    int32_t* pivots;
} lkx_args;
int lkd_batch_crossover(const lkb_args* args, const lkx_args* xo, void* stream);

/* Phase I / Phase II primal simplex of every LP of a batch (one workgroup per LP; the
 * reference's CPU route solve_with_phase1 + primal_simplex_phase2, src/simplex/phase1.cpp:124-449
 * and primal.cpp:27-423, with a dense explicit basis inverse in the LP's factor slot): Devex
 * pricing with Bland's rule after 50 consecutive degenerate steps, Harris two-pass ratio test,
 * at most max_iter iterations per phase.  Outputs per LP: status (0 Optimal, 1 Infeasible,
 * 2 Unbounded, 3 IterationLimit), basis (y layout; -1 for a row left artificial), x (x layout,
 * nonbasic 0), duals y = B^-T c_B of the final basis (y layout), objective c^T x,
 * iterations[2 i .. 2 i + 1] of Phase I and Phase II. */
#define LKS_MAX_ITER 10000
typedef struct {
    int32_t* basis;
    double* x;
    double* y;
    int32_t* status;
    int32_t* iterations;
    double* objective;
    int32_t max_iter;
} lks_args;
int lkd_batch_simplex(const lkb_args* args, const lks_args* sx, void* stream);

/* Device scratch of the no-allocation batch entry point: a fixed file-scope device array of
 * LKB_SCRATCH_BYTES (every GPU the library runs on holds one until the process ends).  *ptr is
 * its address on the current device, resolved once per device. */
#define LKB_SCRATCH_BYTES (64ull << 20)
int lkd_batch_scratch(void** ptr, size_t* bytes);

/* ---------------------------------------------------------------- launch graphs */
/* Stream capture of a launch sequence into an executable graph (kernels only: no memset or
 * copy nodes), its replay and destruction. */
int lkd_graph_begin(void* stream);
int lkd_graph_end(void* stream, void** exec);
int lkd_graph_launch(void* exec, void* stream);
int lkd_graph_destroy(void* exec);

/* ---------------------------------------------------------------- warm-start network */
/* C(m, n) = alpha sum_k A(m, k) B(n, k) + beta C(m, n) (+ bias[n]; then 0 where
 * mask[m ldc + n] <= 0), A(m, k) = A[m sam + k sak], B(n, k) = B[n sbn + k sbk], C(m, n) =
 * C[m ldc + n]; relu_a / relu_b read that operand as max(., 0).  Operands are rounded to BF16
 * and multiplied on MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.  work (optional,
 * work_elems floats): a product with few output tiles and a long K is split over K into it and
 * the splits are added in a fixed order (deterministic). */
int lkd_mlp_gemm(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                 int64_t sbn, int64_t sbk, float* C, int64_t ldc, float alpha, float beta,
                 const float* bias, const float* mask, int relu_a, int relu_b, float* work,
                 int64_t work_elems, void* stream);
/* out[n] = beta out[n] + sum_m G[m ld + n] (rows in order) */
int lkd_mlp_colsum(const float* G, int M, int N, int64_t ld, float beta, float* out, void* stream);
/* g = scale (Y - T) over `count` entries; *loss = sum (Y - T)^2 (block partials in `part`,
 * at most nparts, added in order) */
int lkd_mlp_mse(const float* Y, const float* T, int64_t count, float scale, float* g, float* part,
                int32_t nparts, float* loss, void* stream);
/* dst[r, :] = src[idx[r], :] (rows of `cols` floats) */
int lkd_mlp_gather(const float* src, const int32_t* idx, int rows, int cols, float* dst, void* stream);
/* Adam step `step` (>= 1) on n FP32 weights w with gradient g and moments m, v */
int lkd_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int64_t step, void* stream);
/* one v_mfma_f32_16x16x16_bf16 on 64 lanes: lane l supplies a[4l..4l+3], b[4l..4l+3] (rounded
 * to BF16); out[4 l + e] = element e of its result */
int lkd_probe_mfma_bf16(const float* a, const float* b, float* out, void* stream);

/* ---------------------------------------------------------------- probes (tests) */
/* One 16x16x4 MFMA (FP64 or FP32) on 64 lanes: lane l supplies a[l], b[l]; out[l*4+e] = D[e]. */
int lkd_probe_mfma(int prec, const void* a, const void* b, void* out, void* stream);
/* out[0, 8192) (128x64, column-major) = A (ra x K, ld lda, ra <= 128) * B (rb x K, ldb, rb <= 64)^T
 * from the NT GEMM core; out[8192, 16384) = the same from the transposed (TN) core */
int lkd_probe_tile_gemm(int prec, const void* a, int64_t lda, int ra, const void* b, int64_t ldb,
                        int rb, int K, void* out, void* stream);
/* FP64 direct-to-LDS TN products (the STEP / PART staging) against the register-staged TN core:
 * out (4 x 128 x 64 doubles, column-major): gemm_tn, the direct-to-LDS 128 x 64 product, and the
 * two 64-column halves of the 128 x 128 product (B rows 0..63, 64..127).  K a multiple of 16,
 * ra <= 128, rb <= 128, lda > ra, ldb > rb. */
int lkd_probe_tile_gemm_dma(const double* a, int64_t lda, int ra, const double* b, int64_t ldb, int rb,
                            int K, double* out, void* stream);
/* in-place Cholesky of one 64x64 column-major SPD block (lower part returned, upper zeroed);
 * *ok = 0 when a pivot is not positive */
int lkd_probe_potrf(int prec, void* a, int32_t* ok, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LAKSHYA_DEVICE_H */
