/*
 * oracle.h -- FP64 CPU restatement of the gpu_determ collision-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  The functions declared here are the checker that
 * the HIP library is compared against.  Only tests/, the smoke() entry of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load this code;
 * the product path (libgdt.so and everything that calls it) never does.
 *
 * Every array is host memory.  Velocity-space arrays of one cell are in the
 * box layout idx = (i*n_v + j)*n_w + k unless a name says "cmp" (compact
 * sphere-cut layout, node order i-j-k as velocity_grid.cu:114-126 emits it).
 * Several cells are stored one after the other: [cell][node].
 */
#ifndef GDT_ORACLE_H
#define GDT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ERR 1.0e-10 /* velocity.h:37 */

/* ---- velocity grid (velocity_grid.cu:6-14, 73-126) ---------------------- */
int orc_grid_nodes(const int n[3], const double vmin[3], const double vmax[3],
                   double R, int *box2c, int *cidx);
void orc_grid_centres(const int n[3], const double vmin[3],
                      const double vmax[3], int ref_float, double *u,
                      double *v, double *w);
void orc_init_f_reference(const int n[3], const double vmin[3],
                          const double vmax[3], double *f_box);

/* ---- analytic collision matrix (compute_matrix.cu:21-257, 505-562) ----- */
double orc_fterm(int kind, double p, double A);
double orc_box_area(double du, double dv, double dw, int in, int jn, int kn,
                    double g);
void orc_build_tables(const int n[3], const double vmin[3],
                      const double vmax[3], int mode, double *b, double *a);
long orc_glist_build(const int n[3], const double *a, int *gptr,
                     int *gbox, double *gval, long cap);

/* ---- loss and gain (collision_integral.cu:32-72, 75-188) ---------------- */
void orc_loss(const int n[3], const double *b, const double *f_box,
              double *nu_box, int cells);
void orc_gain(const int n[3], const double *a, const double *f_box,
              double *gain_box, int cells, const int *dest, int ndest);

/* ---- conservation correction (collision_integral.cu:217-526) ------------ */
int orc_correct(int npts, const double *xi, const double *f, const double *nu,
                const double *gain, double *nu_corr, double *Q, double *X,
                int cells);
int orc_correct_loss(int npts, const double *xi, const double *loss,
                     const double *Qin, double *Qout, double *X, int cells);
int orc_solve5(const double A[25], const double B[5], double X[5]);

/* ---- explicit step (collision_integral.cu:776-833) ---------------------- */
double orc_courant_dt(const double *nu, long count, double courant);
void orc_euler(double *f, const double *Q, long count, double dt);

/* ---- Korobov / projection method (no reference counterpart) ------------ */
uint32_t orc_korobov_pick(uint32_t M, uint64_t seed, int ncand,
                          double *best_p2);
double orc_korobov_p2(uint32_t M, uint32_t a, int dims);
void orc_sincos_2pi(double t, double *s, double *c);
long orc_korobov_stencil(const int n[3], const int *box2c, int npts,
                         uint32_t M, uint32_t a, double K, int32_t *idx6,
                         double *r, double *w, uint32_t *node, long cap);
void orc_korobov_collide(int npts, long nn, const int32_t *idx6,
                         const double *r, const double *w,
                         const double *f_cmp, double *loss, double *Q,
                         int cells);
void orc_korobov_dest_lists(int npts, long nn, const int32_t *idx6, const double *r,
                            long *qrow, int32_t *qnode, double *qcoef, long *lrow, int32_t *lnode);
void orc_korobov_collide_dest(int npts, long nn, const int32_t *idx6,
                              const double *r, const double *w,
                              const long *qrow, const int32_t *qnode, const double *qcoef,
                              const long *lrow, const int32_t *lnode,
                              const double *f_cmp, double *loss, double *Q,
                              int cells);

#ifdef __cplusplus
}
#endif
#endif
