/* Internal declarations shared by the oracle translation units.
 * TEST INFRASTRUCTURE ONLY (see aak_oracle.h). */
#ifndef AAK_ORACLE_INT_H
#define AAK_ORACLE_INT_H
#include "aak_oracle.h"

/* Constants.h:4-20, evaluated with the same double expressions. */
#define O_G 6.67259e-11
#define O_C 299792458.0
#define O_MSUN 1.98892e30
#define O_MSUN_SEC (O_G * O_MSUN / (O_C * O_C * O_C))
#define O_PC 3.0856775807e16
#define O_GPCINSEC (O_PC * 1.e9)
#define O_GPC (O_GPCINSEC / O_C)
#define O_YEAR 31558149.8
#define O_AU 1.4959787066e11
#define O_AUSEC (O_AU / O_C)
/* KSParMap.h:17-18 (PI2 is a truncated pi^2 in the reference). */
#define O_EULERGAMMA 0.5772156649015329
#define O_PI2 9.86960440109

void o_set_error(const char *msg);

double ok_rplus(double a);
double ok_delta(double r, double a);
double ok_eeqpro(double r, double a);
double ok_eeqret(double r, double a);
double ok_lzeqpro(double r, double a);
double ok_lzeqret(double r, double a);
double ok_isco_pro(double a);
double ok_isco_ret(double a);

typedef struct {
  double p, ecc, cosiota, a;
  double peri, ap, tansqiota, taniota;
  double E, Lz, Q, alpha_w, beta_w, zedplus, zedminus, betazedplus;
  double r1, r2, r3, r4, p3, p4;
  int Stable;
} o_iekg_t;

typedef struct {
  double E, Lz, Q, Lovercosi, cosiota;
} o_ckg_t;

int o_iekg_init(o_iekg_t *g, double p, double ecc, double cosi, double a, int want_roots);
void o_iekg_frequencies(const o_iekg_t *g, double *Om);
int o_ckg_init(o_ckg_t *c, double cosi, double r, double a);
int o_gkr(double p, double e, double E, double Lz, double Q, double a, double *out);

#endif
