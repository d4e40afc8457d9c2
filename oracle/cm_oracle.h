/*
 * cm_oracle.h -- CPU restatement of the reference's coarsening path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * in cumetis_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates (cuMetis.cu is
 * abbreviated "cu:").  Offsets are int64 throughout; the reference uses int,
 * which gives the same values for every graph it can represent.
 *
 * Pinning (what checks each function independently of its own formula):
 *   orc_rand / orc_randarrayofp  glibc rand() known-answer sequence (seed 1);
 *                                hand-worked permutations of both branches
 *                                (n < 10 swaps, n >= 10 block swaps)
 *   orc_number                   closed form: rank of representatives
 *   orc_cpu_contract             closed form P^T A P (diagonal dropped),
 *                                brute force over all matchings of tiny graphs,
 *                                first-seen row order from cu:762-819
 *   orc_cpu_match                matching invariants, HEM choice replay,
 *                                degree-bucket order of the visit permutation
 *   orc_handshake_match          invariants (symmetric, maximal, heavy-edge
 *                                local dominance), brute force on tiny graphs
 *   orc_project                  closed form where_f = where_c[cmap]
 *   orc_edgecut / orc_pwgts      closed forms
 *   orc_initpart                 a bisection traced by hand through
 *                                cu:1854-1875, 2536-2661, 2367-2531 with the
 *                                glibc seed-1 stream (BFS start, FM heap ties,
 *                                rollback), plus invariants
 *   orc_kway_refine              a 3-part pass traced by hand through
 *                                cu:3244-3309, 3401-3488 (bounds, tie to the
 *                                lower part, the two passes), plus invariants
 *   orc_kway_balance             hand-worked known answers (no reference
 *                                counterpart: the engine's own pass)
 */
#ifndef CM_ORACLE_H
#define CM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

void orc_srand(unsigned seed);
int orc_rand(void);
int orc_randinrange(int max);
void orc_randarrayofp(int n, int *p, int m, int flag);

int orc_log2(int a);
int orc_threshold(int n, int nparts);
int orc_maxvwgt(long long tvwgt, int threshold);

int orc_cpu_match(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                  const int *vwgt, int maxvwgt, int *match, int *cmap, int *perm_out);
int orc_number(int n, const int *match, int *cmap);
int64_t orc_cpu_contract(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                         const int *vwgt, const int *match, const int *cmap, int nc,
                         int64_t *cxadj, int *cadjncy, int *cadjwgt, int *cvwgt);
int orc_coarsen_continue(int nc, int nf, int64_t mc, int threshold);

uint32_t orc_edge_hash(uint64_t seed, uint32_t a, uint32_t b);
int orc_handshake_match(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                        const int *vwgt, int maxvwgt, uint64_t seed, int max_rounds,
                        int *match);

void orc_project(int n, const int *cmap, const int *where_c, int *where_f);
int64_t orc_edgecut(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                    const int *where);
void orc_pwgts(int n, const int *vwgt, const int *where, int nparts, int64_t *pwgts);

void orc_segsort_u64(int64_t nseg, const int64_t *segoff, uint64_t *keys, uint32_t *vals);

int orc_initpart(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                 const int *vwgt, int nparts, int *where);
void orc_kway_refine(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                     const int *vwgt, int nparts, int *where);
/* the engine's balancing pass (no reference counterpart); returns moves */
int orc_kway_balance(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                     const int *vwgt, int nparts, int *where);

#ifdef __cplusplus
}
#endif
#endif
