/*
 * fmm_oracle.h -- CPU restatement of the Tapas FMM path (FP64).
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never shipped
 * in the product path.  See fmm_oracle.c for the reference citations.
 */
#ifndef FMM_ORACLE_H_
#define FMM_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_P 10
#define ORC_NTERM 55

typedef struct orc_ctx orc_ctx;

/* geometry and keys */
void orc_bbox(const double *x, const double *y, const double *z, int64_t n, double *out_min,
              double *out_max);
uint64_t orc_finest_key(const int anchor[3]);
void orc_keys(const double *x, const double *y, const double *z, int64_t n, const double *rmin,
              const double *rmax, uint64_t *keys);
uint64_t orc_key_child(uint64_t k, int idx);
uint64_t orc_key_next(uint64_t k);
void orc_cell_region(uint64_t key, const double *rmin, const double *rmax, double *cmin,
                     double *cmax);

/* harmonics (P*P complex = 2*P*P doubles, interleaved re/im) */
void orc_eval_multipole(double rho, double alpha, double beta, double *ynm, double *ynm_theta);
void orc_eval_local(double rho, double alpha, double beta, double *ynm);

/* kernels on raw coefficient arrays (NTERM complex each) */
void orc_p2m(double *M, const double *cen, const double *x, const double *y, const double *z,
             const double *q, int64_t nb);
void orc_m2m(double *Mp, const double *Mc, const double *dX);
void orc_m2l(double *Li, const double *Mj, const double *dX);
void orc_l2l(double *Lc, const double *Lp, const double *dX);
void orc_l2p(double *trg, const double *bx, double q, const double *L, const double *cen);

/* Cartesian Taylor basis (orc_cart_terms() = P (P+1) (P+2) / 6 doubles per
 * expansion; term (x, y, z) at orc_cart_index(x, y, z)) */
int orc_cart_terms(void);
int orc_cart_index(int x, int y, int z);
void orc_cart_derivs(double *C, const double *dX);
void orc_cart_powers(double *C, const double *dX, double c0);
void orc_cart_p2m(double *M, const double *cen, const double *x, const double *y,
                  const double *z, const double *q, int64_t nb);
void orc_cart_m2m(double *Mp, const double *Mc, const double *dX);
void orc_cart_m2l(double *Li, const double *Mj, const double *dX);
void orc_cart_m2l_mutual(double *Li, double *Lj, const double *Mi, const double *Mj,
                         const double *dX);
void orc_cart_l2l(double *Lc, const double *Lp, const double *dX);
void orc_cart_l2p(double *trg, const double *bx, double q, const double *L, const double *cen);

/* full evaluation; phases: 0 tree, 1 +upward, 2 +traversal, 3 +downward */
orc_ctx *orc_run(const double *x, const double *y, const double *z, const double *q, int64_t n,
                 int ncrit, double theta, int phases, int record_lists);
/* the same with the reference's mutual interaction (mutual = 1) */
orc_ctx *orc_run_mutual(const double *x, const double *y, const double *z, const double *q,
                        int64_t n, int ncrit, double theta, int phases, int record_lists,
                        int mutual);
/* either basis: 0 spherical, 1 Cartesian */
orc_ctx *orc_run_basis(const double *x, const double *y, const double *z, const double *q,
                       int64_t n, int ncrit, double theta, int phases, int record_lists,
                       int mutual, int basis);
void orc_free(orc_ctx *c);
void orc_region(const orc_ctx *c, double *rmin, double *rmax);
int64_t orc_ncells(const orc_ctx *c);
int orc_max_depth(const orc_ctx *c);
void orc_sorted(const orc_ctx *c, uint64_t *keys, int64_t *perm);
void orc_cells(const orc_ctx *c, uint64_t *keys, int64_t *nb, int64_t *bid, int32_t *leaf,
               double *center, double *width);
void orc_expansions(const orc_ctx *c, double *M, double *L);
void orc_fields(const orc_ctx *c, double *trg4);
void orc_counts(const orc_ctx *c, uint64_t *out4);
void orc_lists(const orc_ctx *c, int64_t *m2l, int64_t *p2p);
void orc_list_sizes(const orc_ctx *c, int64_t *out2);

void orc_direct(const double *x, const double *y, const double *z, const double *q, int64_t n,
                const int64_t *targets, int64_t nt, double *trg4);
/* periodic images (images >= 1, period cycle): the FMM and the direct sum
 * over the (3^images)^3 copies of the bodies */
orc_ctx *orc_run_periodic(const double *x, const double *y, const double *z, const double *q,
                          int64_t n, int ncrit, double theta, int images, double cycle,
                          int phases, int record_lists);
orc_ctx *orc_run_periodic_basis(const double *x, const double *y, const double *z,
                                const double *q, int64_t n, int ncrit, double theta, int images,
                                double cycle, int phases, int record_lists, int basis);
void orc_direct_periodic(const double *x, const double *y, const double *z, const double *q,
                         int64_t n, const int64_t *targets, int64_t nt, int images, double cycle,
                         double *trg4);
void orc_rel_l2(const double *trg, const double *ref, int64_t n, double *err2);

#ifdef __cplusplus
}
#endif
#endif
