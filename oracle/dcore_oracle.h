/*
 * dcore_oracle.h -- CPU oracle for D-core ((k,l)-core) decomposition.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  It is linked or loaded only by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 * The product library never links it and has no CPU fallback to it.
 *
 * Inputs are compact CSR arrays (off[v+1]-off[v] == degree, no slack slots),
 * uint32 vertex ids and offsets (64-bit offsets are accepted so that
 * scale-24 graphs fit), int32 results -- the types of the reference
 * (src/common.h:59-62).
 *
 * Pinning of every function is recorded in the header of dcore_oracle.c.
 */
#ifndef DCORE_ORACLE_H
#define DCORE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Level-synchronous peel for one k (reference scan/process semantics,
 * main.cu:27-141, driven as in KListInplace main.cu:171-195).
 * Returns lmax_k (the last level that removed a vertex; -1 when V == 0). */
int32_t dco_klist(uint32_t V,
                  const uint64_t* out_off, const uint32_t* out_nbr,
                  const uint64_t* in_off, const uint32_t* in_nbr,
                  int32_t k, int32_t* l_out);

/* kmax pass: the same peel on the in/out-swapped graph with k = 0
 * (dcore main.cu:218-220).  Returns kmax (-1 when V == 0). */
int32_t dco_kmax(uint32_t V,
                 const uint64_t* out_off, const uint32_t* out_nbr,
                 const uint64_t* in_off, const uint32_t* in_nbr,
                 int32_t* kmaxes);

/* All rows k = 0..kmax (dcore main.cu:238-244).  `lmaxes` holds
 * (kmax+1)*V int32, row-major by k.  Rows are independent and run in
 * parallel over `nthreads` OpenMP threads (<= 0: OpenMP default).
 * Returns lmax. */
int32_t dco_lrows(uint32_t V,
                  const uint64_t* out_off, const uint32_t* out_nbr,
                  const uint64_t* in_off, const uint32_t* in_nbr,
                  int32_t kmax, int32_t* lmaxes, int nthreads);

/* Second exact path: the h-index fixpoint for one k (the l-value iteration
 * designed in the commented-out kList* family, main.cu:982-1242), Jacobi
 * updates.  Needs the kmaxes of the same graph.  Returns lmax_k over the
 * (k,0)-core (0 if it is empty); *iters receives the number of sweeps. */
int32_t dco_klist_hindex(uint32_t V,
                         const uint64_t* out_off, const uint32_t* out_nbr,
                         const uint64_t* in_off, const uint32_t* in_nbr,
                         const int32_t* kmaxes, int32_t k, int32_t* l_out,
                         int32_t* iters);

/* h-index fixpoint for the in-core numbers (kmaxRefineHIndex main.cu:589-663
 * started from the in-degree upper bound).  Returns kmax. */
int32_t dco_kmax_hindex(uint32_t V,
                        const uint64_t* in_off, const uint32_t* in_nbr,
                        int32_t* kmaxes, int32_t* iters);

/* Hash of the whole decomposition (the formula of dcx_state_hash in the
 * product's include/dcx.h) computed one row at a time on its (k,0)-core and
 * never stored: memory O(V + E) plus O(V) per thread, so scale-24 graphs fit.
 * *kmax_out, *lmax_out receive kmax and lmax; row_hash (may be NULL, else
 * kmax + 2 entries, sized from a previous dco_kmax) receives
 * the kmaxes' part at [0] and row k's at [k + 1]. */
uint64_t dco_state_hash(uint32_t V,
                        const uint64_t* out_off, const uint32_t* out_nbr,
                        const uint64_t* in_off, const uint32_t* in_nbr,
                        int nthreads, int32_t* kmax_out, int32_t* lmax_out,
                        uint64_t* row_hash);

/* Timed sample of one decomposition for a CPU baseline: the kmax pass, the
 * kmax-rank relabelling and the rows k = 0, stride, 2*stride, ... on their
 * cores.  t[3] receives the wall seconds of the three parts, w[2] the
 * algorithmic bytes (8 E_k + 20 V_k summed) of the sampled rows and of all
 * rows.  Returns kmax. */
int32_t dco_time_sample(uint32_t V,
                        const uint64_t* out_off, const uint32_t* out_nbr,
                        const uint64_t* in_off, const uint32_t* in_nbr,
                        int nthreads, int32_t stride, double* t, double* w);

#ifdef __cplusplus
}
#endif
#endif
