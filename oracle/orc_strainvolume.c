/* placeholder: replaced by the strain-volume restatement */
#include <math.h>
#include "ucy_oracle.h"
void orc_refsys_strainvolume(int ns, const double *x, const double *L, const double *W,
                             const double *strike, double *sv, double *dv, double *nv, double *xc)
{ for (int i = 0; i < 3 * ns; i++) sv[i] = dv[i] = nv[i] = xc[i] = NAN; }
int orc_traction_kernels_strainvolume(const double x[3], const double sv[3], const double dv[3],
                                      const double nv[3], int ns, const double *y,
                                      const double *L, const double *T, const double *W,
                                      const double *strike, const double epsp[6],
                                      double G, double lambda, double *ts, double *td, double *tn)
{ for (int i = 0; i < ns; i++) ts[i] = td[i] = tn[i] = NAN; return -1; }
int orc_stress_kernels_strainvolume(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const double *y,
                                    const double *L, const double *T, const double *W,
                                    const double *strike, const double epsp[6],
                                    double G, double lambda, double *s11, double *s12,
                                    double *s13, double *s22, double *s23, double *s33)
{ for (int i = 0; i < ns; i++) s11[i] = s12[i] = s13[i] = s22[i] = s23[i] = s33[i] = NAN; return -1; }
int orc_stress_strainvolume(double x1, double x2, double x3, double q1, double q2, double q3,
                            double L, double T, double W, double theta,
                            const double epsp[6], double G, double nu, double sig[6])
{ for (int k = 0; k < 6; k++) sig[k] = NAN; return -1; }
