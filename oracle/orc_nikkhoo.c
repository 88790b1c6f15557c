/* placeholder: replaced by the triangle restatement */
#include <math.h>
#include "ucy_oracle.h"
void orc_refsys_nikkhoo15(int nVe, const double *v, int ns, const int *i1, const int *i2,
                          const int *i3, double *sv, double *dv, double *nv, double *xc)
{ (void)nVe; (void)v; (void)i1; (void)i2; (void)i3; for (int i = 0; i < 3 * ns; i++) sv[i] = dv[i] = nv[i] = xc[i] = NAN; }
void orc_traction_kernels_nikkhoo15(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const int *i1, const int *i2,
                                    const int *i3, int nVe, const double *vert,
                                    double s, double d, double t, double G, double lambda,
                                    double *ts, double *td, double *tn)
{ for (int i = 0; i < ns; i++) ts[i] = td[i] = tn[i] = NAN; }
void orc_stress_kernels_nikkhoo15(const double x[3], const double sv[3], const double dv[3],
                                  const double nv[3], int ns, const int *i1, const int *i2,
                                  const int *i3, int nVe, const double *vert,
                                  double s, double d, double t, double G, double lambda,
                                  double *s11, double *s12, double *s13, double *s22,
                                  double *s23, double *s33)
{ for (int i = 0; i < ns; i++) s11[i] = s12[i] = s13[i] = s22[i] = s23[i] = s33[i] = NAN; }
