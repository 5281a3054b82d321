/*
 * qsx.h -- C-ABI of the MI355X (gfx950) block-sparse symmetric tensor
 * contraction library (QSpace contractQS hot path).
 *
 * Plain C types only: pointers, sizes, error codes.  Every entry point
 * returns 0 on success or a QSX_ERR_* code and writes a message into the
 * caller's err buffer (errlen bytes, may be NULL/0).
 *
 * Reference interfaces replaced (citations into moondg/qspace):
 *   qsx_contract        <- contractQS usage #1, Source/contractQS.cc:505-632
 *                          (contract_plain + CONTRACT_QS), which calls
 *                          QSpace::contract, Source/QSpace.cc:4296-4420
 *   qsx_contract_cell   <- contractQS usage #2 (cell form, itag matching),
 *                          contractQS_itags, Source/contractQS.cc:634-808
 *   qsx_ctr_plan_*      <- the same QSpace::contract split into its structure
 *                          half (contract_getIdxSet, Source/QSpace.cc:4131-4294)
 *                          and its data half (contractDATA_group,
 *                          Source/QSpace_aux.cc:27-355); the split mirrors the
 *                          pair-list caching of CPAT (Source/cpat.hh:572-593)
 *   qsx_hpsi_*          <- two-site H*psi of Class/@Hamilton1D/private/get_HPsi.m:54-57
 *                          (full MPO) and :62-83 (pseudo MPO, qsx_hpsi_plan_create_pseudo)
 *                          (contractQS({HK_L,Psi},HK_R)) and the legacy C++
 *                          CPAT::HZTimes, Source/cpat.hh:487-527
 *   qsx_lanczos         <- RunEigs Lanczos loop, Source/eigs2SiteQS.cc:98-334
 *   qsx_davidson        <- Davidson loop, Class/@Hamilton1D/private/update_psi_2site.m:223-391
 *                          and RunDavidson, Source/david2SiteQS.cc:316-535
 *   qsx_block_zero_flags <- SkipZeroData's zero test, Source/QSpace.cc:3200-3270
 *   qsx_permute         <- permuteQS, Source/permuteQS.cc:25-87
 *   qsx_plus            <- plusQS, Source/plusQS.cc:17-87
 *   qsx_ortho           <- orthoQS core (SVD_Data::blockSVD + dmrgTruncate),
 *                          Source/mpsortho.cc:545-954,1031-1078; called by
 *                          mpsOrthoQS after the eigensolve, eigs2SiteQS.cc:481-484
 *   qsx_su2_cg / qsx_su2_x3 / qsx_su2_om_basis / qsx_su2_xom
 *                       <- SU(2) Clebsch-Gordan data and X-symbols of
 *                          X3Map::contractCGR, Source/clebsch_aux.cc:4073-4510
 */
#ifndef QSX_H
#define QSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSX_MAX_RANK 8
#define QSX_MAX_SYM 4

/* error codes */
#define QSX_OK 0
#define QSX_ERR_ARG 1         /* invalid argument, index, permutation */
#define QSX_ERR_ITAG 2        /* itag mismatch (contract_getIdxSet) */
#define QSX_ERR_QDIM 3        /* QDIM / qtype inconsistency */
#define QSX_ERR_STRUCT 4      /* invalid QSpace structure (dims, labels) */
#define QSX_ERR_UNSUPPORTED 5 /* outside the supported symmetry set (OM>1 etc.) */
#define QSX_ERR_DEVICE 6      /* HIP runtime failure / no device */
#define QSX_ERR_NOMEM 7
#define QSX_ERR_CG 8          /* Clebsch-Gordan / X-symbol inconsistency */

/* execution modes of the data half of a contraction */
#define QSX_MODE_BATCHED 0    /* one grouped FP64 MFMA launch per contraction */
#define QSX_MODE_PLAIN 1      /* one FP64 GEMM launch per (ia,ib) pair, no batching */
#define QSX_MODE_SPLIT_BF16 2 /* BF16 multi-slice operands on MFMA, FP32 accumulate, FP64 recombine */
#define QSX_MODE_SPLIT_F32 3  /* FP32 3-slice operands, FP32 accumulate, FP64 recombine */

/*
 * Symmetry-labelled block-sparse tensor in QSpace layout
 * (struct {Q, data, info}, Source/QSpace.hh:1455-1475, mx layout 3150-3364).
 *   Q     : nblk x (rank*qdim) int32 label records, row-major; a record is
 *           leg-major with the qdim ints of one leg in qtype order
 *           (QIDX of the reference).  SU(2) labels are 2S.
 *   dims  : nblk x rank block dimensions; blocks are column-major.
 *   data  : nblk pointers to the block data (host memory).  May be NULL
 *           for the structure-only plan calls.
 *   itags : rank itags, 8 chars packed little-endian into a uint64; the
 *           conj ("outgoing", trailing '*') flag is bit 7 of byte 0
 *           (itag_, Source/wbindex.hh:632-700).  May be NULL for purely
 *           abelian tensors.
 *   qtype : comma separated symmetry list, e.g. "A", "SU2", "A,SU2"
 *           ("" = abelian labels without type).
 *   cgw   : nblk x nsym outer-multiplicity weights (scalar per block and
 *           non-abelian symmetry; entries of abelian symmetries ignored).
 *           NULL = external normalisation |cgw| = sqrt(d) of
 *           CRef::normExt (Source/clebsch.cc:4738-4760).
 *   om    : nblk outer multiplicities (NULL = 1 for every block).  A block
 *           with om > 1 carries a trailing data dimension of that size, the
 *           coefficients of its data on the om basis tensors of
 *           qsx_su2_om_basis (the extra data dimension of QS_FULL_OM
 *           blocks, ExpandOM / FuseOM in Source/QSpace_aux.cc:242-345);
 *           om must equal the size of that basis.  Contractions accept such
 *           blocks (one SU(2) symmetry); the other entry points reject them.
 */
typedef struct qsx_tensor {
  int32_t nblk;
  int32_t rank;
  int32_t qdim;
  const int32_t* Q;
  const int64_t* dims;
  const double* const* data;
  const uint64_t* itags;
  const char* qtype;
  const double* cgw;
  /* imaginary parts of complex data: nblk pointers to blocks of the same
     layout as data, or NULL for real data (qsx_contract, qsx_permute,
     qsx_plus; the other entry points reject complex data) */
  const double* const* data_im;
  const int32_t* om;
} qsx_tensor;

/* Result tensor, owned by the library; release with qsx_tensor_free. */
typedef struct qsx_tensor_out {
  int32_t nblk, rank, qdim, nsym;
  int32_t* Q;        /* nblk x rank*qdim */
  int64_t* dims;     /* nblk x rank */
  int64_t* offs;     /* nblk+1 element offsets of the blocks in data */
  double* data;      /* offs[nblk] doubles, host memory (NULL for structure only) */
  uint64_t* itags;   /* rank */
  double* cgw;       /* nblk x nsym */
  char qtype[32];
  /* statistics of the contraction that produced it */
  int64_t npairs;        /* executed (ia,ib) pairs */
  double useful_flops;   /* sum over executed pairs of 2*M*N*K */
  double gpu_ms;         /* device time of the data half (events) */
  double* data_im;       /* imaginary parts (complex results), else NULL */
  int32_t* om;           /* nblk outer multiplicities, or NULL when all are 1 */
} qsx_tensor_out;

typedef struct qsx_opts {
  int32_t mode;          /* QSX_MODE_* */
  int32_t verbose;       /* 0 quiet, 1 warnings (default), 2 info */
  int32_t nslices;       /* split modes: slices per operand (0 = default) */
  int32_t skip_zero;     /* 1: SkipZeroData + NormCGW as CONTRACT_QS does (default) */
  int32_t hconj;         /* H*psi: also add the Hermitian-conjugate term (get_HPsi.m:55-56) */
  const double* fermA;   /* optional per-block sign factors of A (contractDATA_group afac);
                            for H*psi: per-block signs of HL */
  const double* fermB;   /* optional per-block sign factors of B */
} qsx_opts;

/* library / device ------------------------------------------------------ */
const char* qsx_version(void);
int qsx_device_init(int device, char* err, size_t errlen);
int qsx_malloc_device(void** ptr, size_t bytes, char* err, size_t errlen);
int qsx_free_device(void* ptr);
int qsx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int qsx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int qsx_memset_device(void* dst, int value, size_t bytes, void* stream);
int qsx_synchronize(void);
/* device copy kernel of n doubles (a plain streaming kernel; also the
   calibration reference for hardware byte counters) */
int qsx_copy_device(double* dst, const double* src, int64_t n, void* stream);

/* one-shot contraction: contractQS(A,ia,B,ib[,perm]) -------------------- */
/* ia/ib are 0-based leg indices; conjA/conjB are the trailing '*' flags;
   perm (0-based, may be shorter than the output rank or NULL) is completed
   with the identity as in Source/contractQS.cc:505-594.  Complex operands
   (data_im set) give a complex result: the real/complex dispatch of
   contractQS.cc:505-594 (zgemm in QSpace_aux.cc), a conj flag conjugating
   that operand's data as well as reversing its arrows. */
int qsx_contract(const qsx_tensor* A, const int32_t* ia, int32_t na, int32_t conjA,
                 const qsx_tensor* B, const int32_t* ib, int32_t nb, int32_t conjB,
                 const int32_t* perm, int32_t nperm, const qsx_opts* opts,
                 qsx_tensor_out** C, char* err, size_t errlen);
void qsx_tensor_free(qsx_tensor_out* t);

/* contractQS usage #2, the cell form with itag auto-contraction
   (contractQS_itags, Source/contractQS.cc:634-808; the options of icFlags,
   :159-293; the matching of iTags::getCtr, Source/wbindex.cc:1512-1590):
   contractQS({A, opts.., {B, opts.., C}, opts..}, D, opts.. [,perm]).
   An entry is a QSpace (T) or a nested cell (sub, nsub >= 2 entries), with
   the options that follow it in the reference's argument list: option
   strings ("*" conj, "23" / "!2" the 1-based legs that may / may not take
   part in the match, combined as "!2*") and at most one numeric index
   option (idx, 1-based).  Every level contracts right to left over all
   legs whose itags match (equal names with opposite conj flags, or equal
   flags when one side carries "*"; with no match, the match is retried
   with conj(A), the reference's missing-conj rule).  One top-level entry:
   its conj flag and index option give a conj / permutation of it.  perm
   (0-based, may be shorter than the rank) permutes the result; the result
   takes SkipZeroData unless opts->skip_zero == 0.  Real operands stay on
   the device for the whole fold (each distinct leaf uploaded once);
   complex leaves are contracted step by step through the complex path. */
typedef struct qsx_cell {
  const qsx_tensor* T;           /* a QSpace, or NULL for a nested cell */
  int32_t nsub;                  /* nested cell: its entries */
  const struct qsx_cell* sub;
  int32_t nopt;                  /* option strings following the entry */
  const char* const* opt;
  int32_t nidx;                  /* numeric index option (1-based legs), 0 if none */
  const int32_t* idx;
} qsx_cell;
int qsx_contract_cell(const qsx_cell* items, int32_t nitems, const int32_t* perm, int32_t nperm,
                      const qsx_opts* opts, qsx_tensor_out** C, char* err, size_t errlen);
/* one step of that fold on the host, structure only: the options of A and
   B (as in qsx_cell) and the itag matching give the contracted legs
   ia / ib (0-based, nctr of them; arrays of at least min(rank) entries)
   and the conj flags the contraction takes (conjA after the missing-conj
   rule); needs no device */
int qsx_itag_match(const qsx_tensor* A, int32_t nopta, const char* const* opta, const int32_t* idxa, int32_t nidxa,
                   const qsx_tensor* B, int32_t noptb, const char* const* optb, const int32_t* idxb, int32_t nidxb,
                   int32_t* ia, int32_t* ib, int32_t* nctr, int32_t* conjA, int32_t* conjB, char* err,
                   size_t errlen);

/* whole-QSpace block operations (device kernels over all blocks) -------- */
/* permuteQS(A, perm [,'conj']) <- Source/permuteQS.cc:25-87 (QSpace::permute,
   Source/QSpace.hh:2721-2780, then NormCGW, Source/QSpace.cc:3727-3805).
   perm is 0-based and completed with the identity when shorter than the
   rank; conj != 0 flips the conj flag of every leg and conjugates complex
   data (wbarray::permute, Source/wbarray.cc:2738).  An identity
   permutation without conj returns A unchanged.  B->gpu_ms = device time. */
int qsx_permute(const qsx_tensor* A, const int32_t* perm, int32_t nperm, int32_t conj,
                qsx_tensor_out** B, char* err, size_t errlen);
/* plusQS(A, B [,bfac]) <- Source/plusQS.cc:17-87: C = A + bfac*B
   (QSpace::plus_plain, Source/QSpace.cc:3593-3705, for abelian labels; Cat +
   MakeUnique, Source/QSpace.cc:3075-3140,3807-3905, with SU(2) labels), then
   SkipZeroData and NormCGW (opts->skip_zero = 0 keeps zero blocks; opts
   may be NULL).  Complex operands: both planes (a real operand's
   imaginary plane is zero; a block is zero when both planes are).
   C->gpu_ms = device time of the sum. */
int qsx_plus(const qsx_tensor* A, const qsx_tensor* B, double bfac, const qsx_opts* opts,
             qsx_tensor_out** C, char* err, size_t errlen);

/* statistics of a truncation (the Sout struct of dmrgTruncate,
   Source/mpsortho.cc:758-769); release with qsx_ortho_info_free */
typedef struct qsx_ortho_info {
  int64_t ntot;      /* singular values (all groups) */
  int64_t nkeep;     /* kept */
  int32_t nkmin;
  int32_t flag;      /* number of notes dmrgTruncate would print */
  double stol;
  double svd2tr;     /* discarded weight / total weight */
  double sv2tot;     /* total weight, sum of squared singular values */
  double sfac;       /* norm readjustment of the kept singular values */
  double* svd;       /* ntot singular values, descending */
  double gpu_ms;     /* device time (assembly, SVD, result blocks) */
  int32_t sweeps;    /* largest number of Jacobi sweeps of a group */
  int32_t ngroups;   /* symmetry groups (block matrices) */
  double* svd_qdim;  /* ntot multiplet dimensions of the svd records (S, qdim)
                        sorted descending; all 1 for abelian symmetries */
} qsx_ortho_info;

/* orthoQS(PSI, K, Nkmin, Nkeep, stol, itag) <- Source/mpsortho.cc:1031-1078
   (SVD_Data::blockSVD, :545-613, and dmrgTruncate, :615-954): PSI's blocks
   are grouped by the label sum over legs [0] (K=1), [r-1] (K=r-1) or
   [0..K-1]; each group's block matrix (legs 0..K-1 against the rest) is
   decomposed; the kept singular values (markSet 'desc': at most nkeep,
   those above stol, at least nkmin; nkeep < 0: no limit) give
   A1 = U (legs 0..K-1, new leg) and A2 = V S sfac (legs K..r-1, new leg),
   PSI ~ A1 . A2 over the new leg.  itag names the new leg (NULL or "":
   the reference's default).  Real data; abelian symmetries, and SU(2)
   (alone or with abelian factors) for K = 1 or r-1 splits of rank <= 3,
   where each block enters times prod cgw / sqrt(qdim) of the grouping
   multiplet and the weights count every singular value qdim times; for
   K = r-1 A1 keeps PSI's CGT and A2 is the rank-2 identity, for K = 1 of
   a rank-3 PSI A1 is the identity and A2 carries PSI's CGT permuted to
   (legs 1, 2, new); output weights normExt.
   QSX_ERR_UNSUPPORTED otherwise, and for a K = r-1 split whose last leg is
   incoming, which the reference answers with an identity-CG flip. */
int qsx_ortho(const qsx_tensor* psi, int32_t K, int32_t nkmin, int64_t nkeep, double stol, const char* itag,
              qsx_tensor_out** A1, qsx_tensor_out** A2, qsx_ortho_info** info, char* err, size_t errlen);
void qsx_ortho_info_free(qsx_ortho_info* info);

/* parse a ctr-idx string ("13", "1,3", "[1 3]*") into 0-based indices;
   returns the number of indices (>0) or a negative error */
int qsx_parse_ctridx(const char* s, int32_t* idx, int32_t maxlen, int32_t* conj);

/* device-resident contraction plans ------------------------------------- */
/* Device data of a tensor is one buffer holding the blocks back to back in
   block order, column-major, at the element offsets returned by
   qsx_packed_offsets (no padding). */
int qsx_packed_offsets(const qsx_tensor* T, int64_t* offs /* nblk+1 */);

typedef struct qsx_ctr_plan qsx_ctr_plan;
int qsx_ctr_plan_create(const qsx_tensor* A, const int32_t* ia, int32_t na, int32_t conjA,
                        const qsx_tensor* B, const int32_t* ib, int32_t nb, int32_t conjB,
                        const int32_t* perm, int32_t nperm, const qsx_opts* opts,
                        qsx_ctr_plan** plan, char* err, size_t errlen);
/* structure of the result (data==NULL); caller frees with qsx_tensor_free */
int qsx_ctr_plan_output(const qsx_ctr_plan* plan, qsx_tensor_out** C);
/* executed GEMM pairs grouped by output block: ia/ib/afac have npairs
   entries, seg has nout+1 entries (may pass NULL to query the sizes).  With
   outer multiplicity the groups are the output blocks' components and a
   pair is one (ia, ib, a, b) term of one component. */
int qsx_ctr_plan_pairs(const qsx_ctr_plan* plan, int64_t* npairs, int64_t* nout,
                       int32_t* ia, int32_t* ib, double* afac, int64_t* seg);
/* the index set of contract_getIdxSet (Source/QSpace.cc:4131-4294) as the
   plan holds it: ALL label-matched pairs (ia, ib) grouped by output block in
   the reference order (matchIndex / groupRecs), seg (nout+1 offsets), the
   output label records qc (nout x (rank_A + rank_B - 2k) * qdim, before the
   output permutation), and per pair afac = signs x X-symbols x cgw of the
   operands (before NormCGW; 0 for a pair below CG_SKIP_EPS2).  Any pointer
   may be NULL to query the sizes. */
int qsx_ctr_plan_idxset(const qsx_ctr_plan* plan, int64_t* npairs, int64_t* nout, int32_t* ia, int32_t* ib,
                        double* afac, int64_t* seg, int32_t* qc);
int qsx_ctr_execute(qsx_ctr_plan* plan, const double* dA, const double* dB, double* dC,
                    void* stream);
double qsx_ctr_plan_flops(const qsx_ctr_plan* plan);
/* SkipZeroData's test on device data (Source/QSpace.cc:3200-3270, eps = 0):
   flags[i] = 1 if block i (elements offs[i]..offs[i+1]-1 of ddata) is all
   exact zeros; offs / flags are host arrays, ddata device memory. */
int qsx_block_zero_flags(const int64_t* offs, int32_t nblk, const double* ddata, int32_t* flags, char* err,
                         size_t errlen);
/* qsx_block_zero_flags without any device allocation.  No counterpart in
   the reference: it is for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments and result as qsx_block_zero_flags; the
   device copy of offs and the flags live in a fixed per-device array of the
   library (1 MiB, held until the process ends), so nblk is at most 65536
   (larger lists: QSX_ERR_ARG).  Requires qsx_device_init, which takes the
   array's address on the device.  Calls that overlap on one device share
   the array and race; the call is synchronous. */
int qsx_block_zero_flags_noalloc(const int64_t* offs, int32_t nblk, const double* ddata, int32_t* flags, char* err,
                                 size_t errlen);
/* set the work decomposition policy (int32[QSX_POLICY_INTS], see below) */
int qsx_ctr_plan_set_policy(qsx_ctr_plan* plan, const int32_t* policy);
void qsx_ctr_plan_destroy(qsx_ctr_plan* plan);

/* two-site H*psi and eigensolvers ---------------------------------------- */
/* HL(a,a',w), psi(a',b') [rank 2, itag-matched], HR(b',b,w): H*psi =
   contractQS({HL,psi},HR) with itag auto-matching.  The environments are
   copied to the device once; psi/hpsi are device buffers in the packed
   layout of psi's structure. */
typedef struct qsx_hpsi_plan qsx_hpsi_plan;
int qsx_hpsi_plan_create(const qsx_tensor* HL, const qsx_tensor* psi, const qsx_tensor* HR,
                         const qsx_opts* opts, qsx_hpsi_plan** plan, char* err, size_t errlen);
/* H*psi of the pseudo MPO (Class/@Hamilton1D/private/get_HPsi.m:62-83, the
   form update_psi_2site.m:63-66 takes when the Hamiltonian has no full
   MPO): HK_L.psi + psi.HK_R + sum_j contractQS({OP_L(j),'*',psi},OP_R(j))
   [+ contractQS({OP_L(j),psi},OP_R(j),'*') when hconj[j]], summed in that
   order, all contractions by itag matching.  HKL / HKR (rank 2 block
   Hamiltonians) may be NULL; OPL / OPR are nop operator tensors (rank 3,
   OP_L(j) acting on psi's left leg, OP_R(j) on its right leg, their
   operator legs matched by itag); the caller passes the operators of the
   terms J it selects.  Same plan type as qsx_hpsi_plan_create (apply,
   Lanczos, Davidson, autotune); batched FP64 or plain mode. */
int qsx_hpsi_plan_create_pseudo(const qsx_tensor* HKL, const qsx_tensor* HKR, int32_t nop, const qsx_tensor* OPL,
                                const qsx_tensor* OPR, const int32_t* hconj, const qsx_tensor* psi,
                                const qsx_opts* opts, qsx_hpsi_plan** plan, char* err, size_t errlen);
int64_t qsx_hpsi_size(const qsx_hpsi_plan* plan);      /* doubles in psi */
double qsx_hpsi_flops(const qsx_hpsi_plan* plan);      /* useful flops per H*psi */
double qsx_hpsi_bytes(const qsx_hpsi_plan* plan);      /* algorithmic bytes per H*psi */
int qsx_hpsi_apply(qsx_hpsi_plan* plan, const double* dpsi, double* dhpsi, void* stream);
/* full-space norm weight of each element of psi's packed layout (product
   over non-abelian symmetries of cgw^2 of its block, CGT data normalised
   to 1): <x,y> = sum_i w_i x_i y_i is the inner product in which H*psi is
   symmetric; the eigensolvers work in the basis sqrt(w) x */
int qsx_hpsi_weights(const qsx_hpsi_plan* plan, double* w);
/* per-stage work of one H*psi, info[QSX_STAGE_INFO]: [0..1] useful flops
   of stage 1, 2; [2..3] algorithmic bytes; [4..5] work items; [6..7] pairs;
   [8..9] engine (0 workgroup tiles, 1 wave tiles); [10..11] workgroups of
   the main GEMM launch; [12..13] largest fragments per tile side (wave
   tiles); [14..15] reduction workgroups (0: no split tiles); [16..17]
   executed slice-product flops (split precision, else 0); engine 2 = the
   split-precision kernel; [18..19] flops of the FP64 MFMAs the wave engine
   issues (tile and k padding included; 0 for the other engines) */
#define QSX_STAGE_INFO 20
int qsx_hpsi_stage_info(const qsx_hpsi_plan* plan, double* info);
/* Work decomposition policy of one batched contraction, as
   int32[QSX_POLICY_INTS]: [0] engine (0 workgroup tiles, 1 wave tiles),
   [1] workgroup tiles: k-steps of 16 per chunk before a tile is split,
   [2] workgroup tiles: per-tile overhead used when choosing shapes,
   [3] workgroup tiles: persistent grid size, [4] wave tiles: max fragments
   (of 16) per tile side, [5] wave tiles: MFMAs per tile above which the 4
   waves of a workgroup share it, [6] wave tiles: MFMAs per workgroup before
   the k-range is split over workgroups, [7] wave tiles: XCD-aware
   workgroup order over this many L2 domains (0 or 1: longest first only),
   [8] wave tiles of at most 2x2 fragments: waves per workgroup = waves
   sharing one long tile (4, 8 or 16), [9..11] reserved (0).  Any policy gives
   the same result to rounding (k-splits change the summation grouping). */
#define QSX_POLICY_INTS 12
/* autotune the work decomposition of both stages by timing candidates on
   this plan; stage_ms[2] = best times (or -1 when taken from the
   per-process cache), policy[2*QSX_POLICY_INTS] = the chosen policies of
   stage 1 and stage 2 (either pointer may be NULL) */
int qsx_hpsi_autotune(qsx_hpsi_plan* plan, int32_t reps, double* stage_ms, int32_t* policy);
/* set the policies of both stages (policy[2*QSX_POLICY_INTS]) */
int qsx_hpsi_set_policy(qsx_hpsi_plan* plan, const int32_t* policy);
/* capture/replay the H*psi in a hipGraph (0 = direct launches) */
int qsx_hpsi_use_graph(qsx_hpsi_plan* plan, int32_t on);
/* per-kernel device timing of the last direct (non-graph) apply, HIP events
   on the launch stream: [grouped GEMM stage 1, grouped GEMM stage 2] in ms */
int qsx_hpsi_last_timing(const qsx_hpsi_plan* plan, double* ms2);
/* mean device time per launch of one stage's GEMM (stage 1 or 2; with its
   split-tile reduction, if any) over reps back-to-back launches between two
   HIP events on the launch stream, after one H*psi of dpsi; batched FP64
   mode.  Back-to-back launches overlap each launch's dispatch with the
   previous kernel, so this is the kernel's duration plus the gap between
   kernels. */
int qsx_hpsi_stage_ms(qsx_hpsi_plan* plan, int32_t stage, int32_t reps, const double* dpsi, double* dhpsi,
                      double* ms);
void qsx_hpsi_plan_destroy(qsx_hpsi_plan* plan);

/* Lanczos with full FP64 re-orthogonalisation; psi0 is a host vector of
   qsx_hpsi_size doubles; returns the ground energy, the Ritz values of each
   iteration (niter entries) and the ground state (host, may be NULL). */
int qsx_lanczos(qsx_hpsi_plan* plan, const double* psi0, int32_t niter, double* e0,
                double* ritz, double* gs, char* err, size_t errlen);
/* Davidson without preconditioner (update_psi_2site.m:223-391) */
int qsx_davidson(qsx_hpsi_plan* plan, const double* psi0, int32_t niter, double* e0,
                 double* ritz, double* gs, char* err, size_t errlen);

/* SU(2) coefficient engine (host, exact arithmetic) ---------------------- */
/* Clebsch-Gordan <j1 m1; j2 m2 | j3 m3> in Condon-Shortley convention,
   all arguments as two-j / two-m integers */
double qsx_su2_cg(int32_t tj1, int32_t tm1, int32_t tj2, int32_t tm2, int32_t tj3, int32_t tm3);
/* X-symbol x3 of contracting the normalised CGTs of two blocks into the
   CGT of the result block (X3Map::contractCGR, OM=1); labels are 2S per
   leg, dirs are +1 (in) / -1 (out).  Returns 0 and sets *x3. */
int qsx_su2_x3(int32_t ra, const int32_t* tja, const int32_t* dira,
               int32_t rb, const int32_t* tjb, const int32_t* dirb,
               const int32_t* ia, const int32_t* ib, int32_t nctr, double* x3,
               char* err, size_t errlen);
/* Outer-multiplicity basis of a block's SU(2) CGT space (rank r, labels
   2S, dirs +1/-1): the om unit basis tensors, om x prod(tj[k]+1) doubles,
   each column-major over the legs with index i = j - m, basis index
   slowest.  For rank <= 3 the single canonical CGT.  Replaces the OM
   bases of QS_FULL_OM CData (Source/clebsch.hh CData, RC_STORE fusion
   trees) by this library's fusion-chain convention.  out may be NULL
   (count only); returns om >= 0, or -code on error. */
int qsx_su2_om_basis(int32_t r, const int32_t* tj, const int32_t* dir, double* out, int64_t outlen,
                     char* err, size_t errlen);
/* canonical (sorted QSet) leg order of a block CGT (QSet::Sort,
   Source/clebsch.cc:1435-1505): ord[k] = block leg at position k */
int qsx_su2_canon_order(int32_t r, const int32_t* tj, const int32_t* dir, int32_t* ord);
/* Matrix-valued X-symbol of two blocks' bases (X3Map::contractCGR with
   outer multiplicity, Source/clebsch_aux.cc:4464-4471): x[(a*ob + b)*oc + c]
   = <C_c | contract(A_a, B_b)>, the kept legs [A | B] permuted by perm
   (new leg l = old leg perm[l]; NULL = identity).  Writes oa, ob, oc (oc = 0:
   no invariant output) and up to xlen coefficients. */
int qsx_su2_xom(int32_t ra, const int32_t* tja, const int32_t* dira, int32_t rb, const int32_t* tjb,
                const int32_t* dirb, const int32_t* ia, const int32_t* ib, int32_t nctr, const int32_t* perm,
                int32_t* dims3, double* x, int64_t xlen, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* QSX_H */
