/*
 * mi355df -- C ABI of the MI355X (gfx950) density-fitted J/K library.
 *
 * Conventions for every entry point:
 *   - plain C, extern "C"; returns 0 on success, otherwise an MDF_E* code and
 *     a message readable with mdf_last_error() (thread-local);
 *   - every array argument is a DEVICE pointer unless said otherwise; row-major
 *     with an explicit leading dimension (in elements);
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); calls
 *     are asynchronous on it;
 *   - scratch memory is a caller-provided workspace (`work`, `work_bytes`),
 *     sized by the matching *_workspace() query; hot calls allocate nothing.
 *
 * The cderi is the reference's packed three-index tensor: (naux, npairs),
 * pair index contiguous, pair p <-> (i >= j) with address i*nao + j
 * (electrolites/fastdfcderi.py:446-449, 690).  Without a pair map the pairs
 * are the full lower triangle in row-major order, p = i(i+1)/2 + j.
 */
#ifndef MI355DF_H
#define MI355DF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDF_ABI_VERSION 12

/* arithmetic precision of a contraction */
#define MDF_PREC_FP64 0
#define MDF_PREC_FP32 1
#define MDF_PREC_BF16X3 2

/* storage type of a cderi: FP64, or FP32 storage with every value widened
 * to FP64 before arithmetic (the reference's FASTDF_FP32 mode,
 * electrolites/_dffp32.py:166-196; its widened blocks, fastdf.py:611-671) */
#define MDF_DTYPE_F64 0
#define MDF_DTYPE_F32 1

/* ---- library / device ------------------------------------------------- */
int mdf_abi_version(void);
const char *mdf_last_error(void);
/* cu_count, HBM bytes and gfx arch name of the current device (host pointers). */
int mdf_device_info(int *cu_count, long long *hbm_bytes, char *arch, int arch_len);

/* ---- J: Coulomb over the packed cderi -----------------------------------
 * Replaces electrolites/fastdf.py:348-386 (_vj_pass: rho = L.d, vj = L^T.rho)
 * as used by _get_jk (:936-946) and the J-only branch _get_j (:688-745).
 *   cderi     (naux, npairs), row stride ld, MDF_DTYPE_F64 or _F32 values
 *   dm_pairs  (n_dm, npairs), row stride ldd: the density gathered onto the
 *             pairs with the symmetry factors of fastdf.py:674-686
 *   rho       (n_dm, naux) out, row stride ldrho (the fitted density)
 *   vj_pairs  (n_dm, npairs) out, row stride ldvj; accumulate != 0 adds.
 */
size_t mdf_get_j_workspace(long naux, long npairs, int n_dm);
int mdf_get_j(const void *cderi, int cderi_dtype, long naux, long npairs, long ld,
              const double *dm_pairs, int n_dm, long ldd,
              double *rho, long ldrho, double *vj_pairs, long ldvj, int accumulate,
              void *work, size_t work_bytes, void *stream);
/* mdf_get_j with no allocation and no workspace argument: the partial sums
 * of both passes live in a file-scope device array of
 * MDF_NOALLOC_SCRATCH_BYTES (256 MiB) per device, held from library load to
 * process end.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Sizes whose partials exceed the
 * array are rejected (MDF_EINVAL, message names the need); the (H2O)64
 * tensor (naux 6976, 378,596 pairs) fits with any number of densities.
 * Calls that overlap on one device race on the array. */
#define MDF_NOALLOC_SCRATCH_BYTES (256L << 20)
int mdf_get_j_noalloc(const void *cderi, int cderi_dtype, long naux, long npairs, long ld,
                      const double *dm_pairs, int n_dm, long ldd, double *rho, long ldrho,
                      double *vj_pairs, long ldvj, int accumulate, void *stream);

/* The two halves of mdf_get_j for a three-centre block B (naux, width) that
 * is produced and consumed chunk by chunk -- the J build without a stored
 * cderi.  Replace GPU4PySCF's Int3c2eOpt.contract_dm / contract_auxvec as
 * the reference's RI-J path calls them (electrolites/fastdfj3c.py:388-530,
 * :726-784):
 *   mdf_j_rho: rho[n, P] (+)= sum_p B[P, p] d[n, p]     d (n_dm, width), ldd
 *   mdf_j_vj:  vj[n, p]  (+)= sum_P c[n, P] B[P, p]     c (n_dm, naux), ldc
 * Workspace: mdf_get_j_workspace(naux, width, n_dm). */
int mdf_j_rho(const void *B, int dtype, long naux, long width, long ld, const double *d, int n_dm, long ldd,
              double *rho, long ldrho, int accumulate, void *work, size_t work_bytes, void *stream);
int mdf_j_vj(const void *B, int dtype, long naux, long width, long ld, const double *c, int n_dm, long ldc,
             double *vj, long ldvj, int accumulate, void *work, size_t work_bytes, void *stream);

/* ---- K: exchange over the packed cderi (FP64 MFMA) ----------------------
 * Replaces electrolites/fastdf.py:974-1005 (fill_symmetric unpack, step 1
 * contract('Lij,njk->nikL'), step 2 _syrk / contract('nikL,njkL->nij')) and
 * _fill_from_triangle (:389-402), for one density:
 *   vk = sum_L (A_L fl)(A_L fr)^T,   fr == NULL: fr = fl (symmetric, mode 0)
 *   cderi     MDF_DTYPE_F64 or _F32 values (FP32 storage is widened on load)
 *   pair_map  (nao*nao) int32 pair index of (i,j) or -1 when screened; NULL
 *             when the cderi holds the full packed triangle
 *   fl, fr    (nao, nocc) density factors, row stride ldf
 *   vk        (nao, nao) out, row stride ldvk; accumulate != 0 adds.
 * prec: MDF_PREC_FP64, MDF_PREC_FP32 (both steps on FP32 MFMA with FP32
 *   accumulation, R stored in FP32) or MDF_PREC_BF16X3 (both steps on BF16
 *   MFMA with each operand split into hi + lo bfloat16 and three product
 *   terms, FP32 accumulation, R stored split) for the mixed-precision SCF
 *   cycles; inputs, the cross-split sums and vk stay FP64.  The reference's
 *   only narrow mode is FP32 *storage* widened before arithmetic
 *   (electrolites/_dffp32.py:166-196).
 * The aux rows are transformed in passes of at most 8 GiB of transformed
 * blocks (a function of naux, nao, nocc, the factor count and prec only);
 * max_block_rows > 0 lowers that bound for the workspace query, and a
 * smaller workspace makes mdf_get_k use shorter passes.
 */
size_t mdf_get_k_workspace(long naux, int nao, int nocc, int two_factors, int prec,
                           long max_block_rows);
int mdf_get_k(const void *cderi, int cderi_dtype, long naux, long npairs, long ld, const int *pair_map,
              int nao, const double *fl, const double *fr, int nocc, long ldf, int prec,
              double *vk, long ldvk, int accumulate, void *work, size_t work_bytes,
              void *stream);

/* ---- grouped J/K over many (molecule, density) items -------------------
 * No counterpart in the reference (each get_jk call there contracts one
 * molecule, electrolites/fastdf.py:805-1026); this batches many.  One item
 * per molecule (all pointers DEVICE memory):
 *   cderi (naux, npairs), row stride ld, pair_map as for mdf_get_k;
 *   n_dm (1..8) densities: dm_pairs (n_dm, npairs) packed as mdf_pack_dm
 *   writes, rho (n_dm, naux) and vj_pairs (n_dm, npairs) out (J);
 *   fl (n_dm, nao, nocc) [and fr, or NULL for symmetric] factors, vk
 *   (n_dm, nao, nao) out (K).
 * mdf_grouped_plan_create flattens the batch into tile tables kept in device
 * memory the plan owns (it allocates; the runs do not) and reports the
 * workspace a run needs; bn = 0 picks the step-1 tile width.  K runs in
 * waves of consecutive items whose transformed blocks and partials fit
 * work_limit bytes together (0: 32 GiB; a wave holds at least one item), so
 * the workspace is about max(work_limit, the largest item's need).
 * mdf_grouped_jk runs the batch as persistent launches only (capturable in a
 * HIP graph).  Each item is cut and computed exactly as the single-item
 * entries cut it (mdf_get_k given at least mdf_get_k_workspace(..., 0)
 * bytes), so its J/K are bit-identical alone and in a batch. */
typedef struct {
    const void *cderi;
    long ld;
    long naux;
    long npairs;
    int nao;
    const int *pair_map;
    int n_dm;
    const double *dm_pairs;
    double *rho;
    double *vj_pairs;
    const double *fl;
    const double *fr;
    int nocc;
    double *vk;
    int cderi_dtype;      /* MDF_DTYPE_F64 or _F32; one storage type per plan */
} mdf_jk_item;

typedef struct mdf_grouped_plan mdf_grouped_plan;
int mdf_grouped_plan_create(const mdf_jk_item *items, int nitems, int with_j, int with_k, int prec,
                            int bn, size_t work_limit, mdf_grouped_plan **plan, size_t *work_bytes);
int mdf_grouped_jk(const mdf_grouped_plan *plan, void *work, size_t work_bytes, void *stream);
int mdf_grouped_plan_destroy(mdf_grouped_plan *plan);
/* tiles[0..6]: J pass-1, J reduce-1, J pass-2, J reduce-2, K step-1, K step-2, K finalize;
 * with ntiles >= 9 also tiles[7] = K waves, tiles[8] = K launches */
int mdf_grouped_plan_info(const mdf_grouped_plan *plan, long *tiles, int ntiles);

/* ---- pair-vector helpers ----------------------------------------------
 * mdf_pack_dm replaces electrolites/fastdf.py:674-686 (_sparse_dm): the
 * density gathered onto the stored pairs, 2 D_ij (hermi 1) or D_ij + D_ji
 * (hermi 0), diagonal halved.  dms is (n_dm, nao, nao) with matrix stride ldm.
 * mdf_unpack_pairs replaces fastdf.py:1014-1015: out[n, i, j] = out[n, j, i]
 * = vj_pairs[n, p], zero where no pair is stored; matrix stride ldo.
 * pair_addr (int32, i*nao + j) may be NULL for the full lower triangle.
 * mdf_build_pair_map: map[i*nao + j] = p (both orders), -1 if not stored;
 * the table mdf_get_k reads for a screened cderi.
 */
int mdf_pack_dm(const double *dms, int n_dm, int nao, long ldm, const int *pair_addr,
                long npairs, int hermi, double *dm_pairs, long ldd, void *stream);
int mdf_unpack_pairs(const double *vj_pairs, int n_dm, long ldv, const int *pair_addr,
                     long npairs, int nao, double *out, long ldo, void *stream);
int mdf_build_pair_map(const int *pair_addr, long npairs, int nao, int *map, void *stream);
int mdf_fill_zero(double *x, long n, void *stream);
/* dst[r * ldd + c] = (float)src[r * lds + c] (round to nearest even): the
 * FP32 store of a cderi chunk, as the reference's FASTDF_FP32 build casts
 * each FP64 block on its way into the tensor (electrolites/_dffp32.py:67-198). */
int mdf_cast_f32(const double *src, long lds, long rows, long cols, float *dst, long ldd, void *stream);
/* dst[r * ldd + c] = (double)src[r * lds + c]: the widening of one block of
 * an FP32-stored cderi for the block iterator's consumers, which take FP64
 * blocks (electrolites/_dffp32.py:166-196 BlockWidener, used by
 * electrolites/fastdffp32.py:72-141 DF.loop). */
int mdf_widen_f32(const float *src, long lds, long rows, long cols, double *dst, long ldd, void *stream);
/* rows x width bytes from src (row pitch spitch bytes) to dst (pitch dpitch)
 * on the stream, either side device or page-locked host memory: the host
 * store's chunk write-out and block prefetch (electrolites/_dfstore.py, the
 * cudaMemcpy2DAsync of its 'stage' mode; electrolites/fastdf.py:611-671). */
int mdf_copy_2d(void *dst, long dpitch, const void *src, long spitch, long width, long rows, void *stream);
/* dst (cols, rows), row stride ldd, = the transpose of src (rows, cols), row
 * stride lds; MDF_DTYPE_F64 or _F32 on both sides.  Reads the reference's
 * aux-contiguous device cderi, the (naux, npairs) transpose of a C-order
 * (npairs, naux) buffer (electrolites/fastdfcderi.py:648, :690), into the
 * pair-contiguous rows the J and K kernels stream: once for a kept copy, or
 * a block of aux rows at a time; and writes the build's pair-contiguous
 * chunks out in that layout. */
int mdf_transpose(const void *src, int dtype, long rows, long cols, long lds, void *dst, long ldd, void *stream);

/* ---- Gaussian integrals (Rys quadrature, one kernel per angular class) --
 * A basis is a set of contracted shells; all arrays are DEVICE memory:
 *   nprim[s], poff[s] (offset into exps/coefs), aooff[s] (first spherical
 *   function), xyz[3 s] (bohr), exps[k], coefs[k] (normalised: gto_norm and
 *   the contraction normalisation folded in, PySCF convention).
 * Spherical functions follow PySCF (m = -l..l; p as x, y, z).
 */
typedef struct {
    int nbas;
    int lmax;                 /* largest l in the basis */
    const int *nprim;
    const int *poff;
    const int *aooff;
    const double *xyz;
    const double *exps;
    const double *coefs;
} mdf_basis;

/* Three-centre (ij|P) into one pair chunk: out[P * ldo + c] with
 * c = p - p0 (col_map NULL) or c = col_map[p - p0] (the chunk's columns in a
 * cderi that stores only the screened-in pairs), p = i(i+1)/2 + j, i >= j.  Replaces GPU4PySCF's int3c2e evaluator in the
 * reference's cderi build (electrolites/fastdfcderi.py:404-409, :604, :667).
 *   pairs          device int32 (ish, jsh) with ish >= jsh, sorted by class
 *                  (l_i, l_j); pair_class_off (HOST, 10 entries) bounds the
 *                  class la*3 + lb.  Every function of the listed pairs must
 *                  fall in [p0, p0 + ldo); unlisted (screened) pairs are left
 *                  untouched (zero-fill out first).
 *   aux_sorted     device int32 aux shells sorted by l; aux_l_off (HOST,
 *                  6 entries) bounds l = 0..4.
 * Supported: orbital l <= 2, auxiliary l <= 4. */
int mdf_int3c2e(const mdf_basis *ao, const mdf_basis *aux, const int *pairs,
                const long *pair_class_off, const int *aux_sorted, const long *aux_l_off,
                long p0, const int *col_map, double *out, long ldo, void *stream);

/* The Coulomb metric (P|Q), all naux x naux entries, out[P * ldo + Q].
 * Replaces electrolites/_dfj2c.py:234-262 (int2c2e). */
int mdf_int2c2e(const mdf_basis *aux, const int *aux_sorted, const long *aux_l_off,
                double *out, long ldo, void *stream);

/* The first step of mdf_get_k alone, FP64: R[L * strideL + i * ldr + k] =
 * (A_L F)[i, k] for the nL rows of one block (A_L the unpacked symmetric
 * row L of the cderi; pair_map and dtypes as for mdf_get_k).  The DF
 * gradient's rhok[L] = F_r^T A_L F_l (electrolites/fastdfgrad.py:16-45,
 * ``rhok`` from the cderi, :665-714) takes its inner product from R. */
int mdf_k_step1(const void *cderi, int cderi_dtype, long nL, long npairs, long ld, const int *pair_map,
                int nao, const double *f, long ldf, int nocc, double *R, long ldr, long strideL, void *stream);
/* mdf_k_step1_lower: as mdf_k_step1 with A_L replaced by its lower triangle,
 *   diagonal halved (T_L with A_L = T_L + T_L^T): R_L = T_L F, about half
 *   the work (each 128-row tile's K loop stops at its diagonal block).  For
 *   the symmetric bilinear form of the DF gradient, F^T A_L F = G + G^T with
 *   G = F^T R_L (electrolites/fastdfgrad.py:870-1182, rhok there from the
 *   full product). */
int mdf_k_step1_lower(const void *cderi, int cderi_dtype, long nL, long npairs, long ld, const int *pair_map,
                      int nao, const double *f, long ldf, int nocc, double *R, long ldr, long strideL,
                      void *stream);

/* Nuclear derivative of the DF gradient's three-centre term:
 *   grad[3a + x]     += sum_{P,i,j} W[P, (i,j)] d(ij|P)/dR_{a,x}  over the
 *                       orbital centres of the listed shell pairs (an
 *                       off-diagonal pair stands for both orders), and
 *   grad_aux[3a + x] += the aux centre's share (minus the orbital centres'
 *                       sum, by translational invariance).
 * W[P * ldw + c]: one value per packed pair p = i(i+1)/2 + j (i >= j) at
 * c = col_map[p] (the compact columns of a cderi that stores only the
 * screened-in pairs) or c = p (col_map NULL).  ao_atom / aux_atom: device
 * int32 atom (< natm) of each shell; grad, grad_aux: device (natm, 3) float64,
 * accumulated (zero them first; they may be the same array).  pairs and
 * aux_sorted as for mdf_int3c2e.  Replaces GPU4PySCF's
 * sum_ejk_int3c2e_ip1 in the reference's DF gradient
 * (electrolites/fastdfgrad.py:1014-1040). */
int mdf_int3c2e_ip(const mdf_basis *ao, const mdf_basis *aux, const int *pairs,
                   const long *pair_class_off, const int *aux_sorted, const long *aux_l_off,
                   const double *W, long ldw, const int *col_map, const int *ao_atom,
                   const int *aux_atom, int natm, double *grad, double *grad_aux, void *stream);

/* Nuclear derivative of the metric term: grad[3a + x] += sum_{P,Q} Z[P * ldz
 * + Q] d(P|Q)/dR_{a,x} over all ordered aux function pairs.  Replaces
 * GPU4PySCF's int2c2e_ip1_per_atom (electrolites/fastdfgrad.py:1172-1178). */
int mdf_int2c2e_ip(const mdf_basis *aux, const int *aux_sorted, const long *aux_l_off,
                   const double *Z, long ldz, const int *aux_atom, int natm, double *grad, void *stream);

/* Overlap, kinetic and nuclear attraction (natm point charges, charges[] and
 * atom_xyz[3*natm] device arrays) over the listed shell pairs, written to
 * both triangles of the (nao, nao) outputs; pairs as for mdf_int3c2e (screened
 * pairs untouched: zero-fill first).  The one-electron Hamiltonian of the SCF
 * driver (PySCF int1e_ovlp, int1e_kin, int1e_nuc); no reference counterpart
 * on the DF path itself. */
int mdf_int1e(const mdf_basis *ao, int nao, const int *pairs, const long *pair_class_off, int natm,
              const double *charges, const double *atom_xyz, double *S, double *T, double *V,
              void *stream);

/* ---- FP64 dense linear algebra of the metric (MFMA tile engine) --------
 * Replace the cuSOLVER/cuBLAS calls of electrolites/fastdfcderi.py:
 * potrf (:196), explicit L^-1 by solve_triangular (:479-481) and the
 * triangular multiply of _trmm_right_lowerT (:237-261).  Row-major.
 * mdf_gemm: C = alpha A op(B) + beta C; transb: B is (N, K); tri_a 1: A is
 *   lower triangular (K loop bounded per row tile), 2: upper triangular (the
 *   K loop of row tile m starts at its first row), either needing M == K;
 *   lower_c: only the lower 128x128 tiles of a square C are written.
 * mdf_gemm_batched: the same over `batch` problems, operand i at A + i
 *   strideA, B + i strideB, C + i strideC (a stride of 0 shares the
 *   operand); the DF gradient's contractions of the metric factor and the
 *   occupied-space blocks (electrolites/fastdfgrad.py:870-1182, there on
 *   cuBLAS/cuTENSOR).
 * mdf_potrf: A = L L^T in place (lower; the strict upper part of each
 *   diagonal 128-block is zeroed).  *info (device int, zero it first) gets
 *   the 1-based index of the first non-positive pivot.  The workspace keeps
 *   the diagonal-block inverses that mdf_trtri reads.
 * mdf_trtri: X = L^-1 (X must not alias L), D = mdf_potrf's workspace.
 * mdf_trmm: C = Linv B for lower-triangular Linv (n x n), B (n x ncols).
 */
int mdf_gemm(int transb, int tri_a, int lower_c, long M, long N, long K, double alpha,
             const double *A, long lda, const double *B, long ldb, double beta, double *C, long ldc,
             void *stream);
int mdf_gemm_batched(int transb, int tri_a, int lower_c, long M, long N, long K, double alpha,
                     const double *A, long lda, long strideA, const double *B, long ldb, long strideB,
                     double beta, double *C, long ldc, long strideC, long batch, void *stream);
/* mdf_gemm_pairs_batched: for each z < batch, P_z = A_z op(B_z) (n x n;
 *   A_z = A + z strideA is n x K, op(B_z) is K x n) is put at the stored
 *   pairs of a packed cderi without being written itself:
 *     W[z ldw + c] = base + alpha P_z[i][j],  c = pair_col[i (i + 1) / 2 + j],
 *   for n > i >= j where pair_col (device int32, n (n + 1) / 2 entries) is
 *   >= 0; base = init_scale[z] init_vals[c] when init_scale (batch device
 *   doubles) and init_vals (device doubles, one per column) are given (W's
 *   old contents are then not read), else W's own value (W += alpha P_z).
 *   pair_col must not name one column twice.  Only the lower 128-tiles of
 *   P_z are computed.  The DF gradient's exchange weights F_l dmoo_P F_r^T
 *   at the cderi's compact columns, with its Coulomb weights c_P D_c as the
 *   base (electrolites/fastdfgrad.py:870-1182, there a dense cuTENSOR
 *   product, a gather and a separate sum). */
int mdf_gemm_pairs_batched(int transb, long n, long K, double alpha, const double *A, long lda,
                           long strideA, const double *B, long ldb, long strideB, const int *pair_col,
                           double *W, long ldw, const double *init_scale, const double *init_vals,
                           long batch, void *stream);
/* mdf_sym_pack: for each z < batch, the packed upper triangle (row-major
 *   over a <= b; (a, b) at a n - a (a - 1) / 2 + b - a) of G_z + G_z^T,
 *   scaled by diag_scale on the diagonal and off_scale off it:
 *     P[z strideP + (a, b)] = s_ab (G[z strideG + a n + b] + G[z strideG + b n + a]).
 * mdf_sym_unpack: the inverse layout, G_z[a][b] = G_z[b][a] = s_ab P_z[(a, b)].
 *   The DF gradient's occupied-space blocks rhok_L = G_L + G_L^T and dmoo_P,
 *   symmetric for one factor, kept packed between its products
 *   (electrolites/fastdfgrad.py:870-1182, there full nocc x nocc blocks). */
int mdf_sym_pack(long n, const double *G, long strideG, double *P, long strideP, double diag_scale,
                 double off_scale, long batch, void *stream);
int mdf_sym_unpack(long n, const double *P, long strideP, double *G, long strideG, double diag_scale,
                   double off_scale, long batch, void *stream);
/* SP2 purification of the serving loop's SCF (no counterpart in the
 * reference, whose SCF driver is GPU4PySCF's): for a symmetric X (n x n,
 * row stride ldx) one step writes Y = X^2 when |tr(X^2) - nocc| >=
 * |2 tr(X) - tr(X^2) - nocc|, else Y = 2X - X^2, computing only the lower
 * 64-tiles of the product and mirroring them (Y symmetric to the bit).
 * tr(X) and tr(X^2) = ||X||_F^2 come from part_in, mdf_sp2_tiles(n) pairs of
 * partial sums in a fixed order, written by mdf_sp2_begin (for X itself) or
 * by the previous step's part_out (for its Y); the sums are taken in the
 * same order by every workgroup, so a step is reproducible.  Y and part_out
 * must not alias X and part_in. */
long mdf_sp2_tiles(long n);
int mdf_sp2_begin(long n, const double *X, long ldx, double *part, void *stream);
int mdf_sp2_step(long n, double nocc, const double *X, long ldx, const double *part_in, double *Y, long ldy,
                 double *part_out, void *stream);
size_t mdf_potrf_workspace(long n);
int mdf_potrf(double *A, long n, long lda, int *info, void *work, size_t work_bytes, void *stream);
size_t mdf_trtri_workspace(long n);
int mdf_trtri(const double *L, long n, long ldl, const double *D, double *X, long ldx, void *work,
              size_t work_bytes, void *stream);
int mdf_trmm(const double *Linv, long n, long ldl, const double *B, long ldb, long ncols, double *C,
             long ldc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MI355DF_H */
