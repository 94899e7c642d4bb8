/* gala_oracle.h -- types of the CPU oracle (test infrastructure only). */
#ifndef GALA_ORACLE_H
#define GALA_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
#define ORC_MAXC 16

enum {
    ORC_NULL = 0,
    ORC_HERNQUIST = 1,
    ORC_PLUMMER = 2,
    ORC_MIYAMOTONAGAI = 3,
    ORC_NFW = 4,
    ORC_SCF = 5
};

/* One potential component: pars[0] = G, then the parameters in the
 * PotentialParameter declaration order (cybuiltin.pyx:165-301 packs
 * [G] + parameters).  q0 / R / shift_rotate follow cpotential.pyx:82-102. */
typedef struct {
    int type;
    int shift_rotate;
    int nmax, lmax;       /* SCF only */
    double pars[8];
    double q0[3];
    double R[9];
    const double *S;      /* SCF only: (nmax+1)(lmax+1)(lmax+1), C order */
    const double *T;
} orc_comp;

typedef struct {
    int n;
    int pad;
    orc_comp c[ORC_MAXC];
} orc_pot;


double orc_component_value(const orc_comp *c, const double *q);
double orc_component_density(const orc_comp *c, const double *q);
void orc_component_gradient(const orc_comp *c, const double *q, double *g);
void orc_component_hessian(const orc_comp *c, const double *q, double *H);
void orc_dop853(const orc_pot *P, const double *w0, size_t N, const double *t, size_t ntimes,
                double atol, double rtol, long nmax, long nstiff, double dt_max, int save_all,
                int nbatch, double *out, int *status, long *stats);
#ifdef __cplusplus
}
#endif
#endif
