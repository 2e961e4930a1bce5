/*
 * ORACLE -- test infrastructure only.  A plain-C FP64 restatement of the
 * DemystiFicatioN hot path, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product links,
 * loads or calls it.
 *
 * What pins each restated function (tests/ file in brackets):
 *   directIII / directII, eval_kernel_ModeIII / _ModeII_T / _ModeII_N, both
 *     friction laws, compute_P (crossings, both quirks, permeability_coupling = 1),
 *     rkqs / rkck, the solver loop, cluster and block cluster trees: bit-equal to
 *     the reference's own Fortran built by Makefile.ref, on committed fixtures and
 *     live [test_oracle_golden.py];
 *   geometry and loading projections (initialisation.f90:139-206): Mohr's-circle
 *     closed forms on faults of every strike, all five stress rates
 *     [test_oracle_setup_pins.py];
 *   find_fault_crossing (initialisation.f90:279-344): hand-built junctions and
 *     near misses at the single-precision threshold, brute force over node pairs,
 *     the generator's crossing count [test_oracle_setup_pins.py];
 *   full-pivot ACA and recompression rank rules: numpy SVD of factors with a known
 *     spectrum and of every built block [test_oracle_setup_pins.py]; the operator
 *     against the dense kernel to 20 eps_ACA [test_oracle_closed_forms.py];
 *   Hgemv_right: dense reconstruction of the block lists [test_oracle_closed_forms.py].
 * Parity unpinned: the loop order of the reference BLAS ZGEMV inside Hgemv_right
 * and the LAPACK QR/SVD inside the recompression (neither library is in the
 * image; only their products are pinned, as above).
 */
#ifndef DFN_ORACLE_INTERNAL_H
#define DFN_ORACLE_INTERNAL_H

#include <stdint.h>
#include "../include/dfn.h"

/* H-matrix in the reference's own storage layout (src/variables.f90:125-141 and
 * src/hmatrix/H_matrix_construction.f90): 1-based cluster/block arrays,
 * factor arrays with the 4x reservation per low-rank block, real FP64 instead of
 * complex128 with zero imaginary part. */
typedef struct {
  int n, nbclus, nbblclus;
  int *mat_corr;        /* [n] 1-based: H position k -> original element mat_corr[k-1] */
  int *ind_clusters;    /* [2*nbclus] first,last (1-based) */
  int *sons;            /* [2*nbclus] son ids (1-based), 0 for leaves */
  int *bl_type;         /* [nbblclus] 0 dense, 1 low-rank, 2 split */
  int *bl_row, *bl_col; /* [nbblclus] row / col cluster id */
  int64_t *info;        /* [nbblclus*6] info_blclusters(b, 1:6) row-major by block */
  double *A, *B, *F;
  int64_t dimA, dimB, dimF;
} ohmat;

typedef struct {
  int mode, kernel, nfault, N;
  int64_t *fnel;            /* faults(i)%nb_element */
  int64_t *map1, *map2;     /* mapping_fault(i,1:2), 1-based */
  double *node;             /* [2*(N+1)] */
  double *elem, *ds, *tang, *nrm, *nrm_left, *nrm_right; /* original order */
  double *a, *b, *Dc, *f0, *V0, *nld, *sld;
  double *perm, *perm_star, *perm_x, *kmin, *kmax, *Snk, *Lk, *Tk; /* perm_x [N+1] 1-based */
  double mu, cs, cp, por, fcomp, rcomp, visc, tol;
  double s11d, s12d, s22d, s31d, s32d;
  int coupling;
  int cg_extra;             /* test hook: CG iterations past the stopping test (0 = the reference) */
  int cg_reverse;           /* test hook: CG dot products summed from the last element (0 = the reference) */
  int nsrc;
  int64_t *src_fault, *src_idx;
  double *Q, *tb, *te;
  int ncross;
  int64_t *cross;           /* [ncross*6] crossing(c, 1:6) */
  int nfake;
  int64_t *fake;            /* 0-based element indices */
  /* H-matrix parameters */
  double eta, eps_aca;
  int nleaf, max_rank;
  ohmat *H[2];              /* DFN_OP_T, DFN_OP_N */
  /* state */
  double *Y, *P, *Pdot;
  double time;
  double *sT, *sN;          /* SigmaT_dot, SigmaN_dot globals */
  /* solver loop state */
  double t_try;
  int loop_started;
  dfn_run_stats st;
  int nthreads;
} oracle_t;

#endif
