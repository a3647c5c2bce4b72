/* ORACLE — test infrastructure only (see orc_eos.c header). Internal decls. */
#ifndef ORC_INT_H
#define ORC_INT_H
#include "orc.h"

#define ORC_NPV 9

double orc_sndp2(const orc_cfg *c, double snd2, double uuu2);
void orc_tvd(const orc_cfg *c, const double (*x)[ORC_NPV], double *xl, double *xr);
void orc_mlp(const orc_cfg *c, const double (*x)[ORC_NPV], double *xl, double *xr, int *flags);
void orc_roe(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
             const double *dvl, const double *dvr, double *fx);
void orc_roem(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
              const double *dvl, const double *dvr, const double *sd, double *fx);
void orc_ausmpwp(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
                 const double *dvl, const double *dvr, const double *sd, double *fx);
void orc_ausmpup(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
                 const double *dvl, const double *dvr, double *fx, int *flag);
void orc_vsflux(const orc_cfg *c, const double *nx, const double *const e[4], const double *const t[4],
                const double *grdl, const double *grdr, const double (*x)[ORC_NPV],
                const double *dvl, const double *dvr, const double *tvl, const double *tvr, double *fx);
void orc_merkle(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                double *r_c, double *r_v, double *icav);
void orc_cavsource(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                double *r_c, double *r_v, double *icav);
double orc_h2o_sigma(double t);
void orc_kwsst_source(const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                      const double *dv, const double *tv, double *source, double *itt, double *omega_cut);
void orc_kwsst_source_des(int ides, const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                      const double *dv, const double *tv, double *source, double *itt, double *omega_cut);
double orc_kwsst_eddy(const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                      const double *dv, const double *tv);
double orc_eigenvis(const orc_cfg *c, const double *dv, const double *tv);
void orc_caljac(const orc_cfg *c, const double *n3, const double *grd, const double *pv,
                const double *dv, const double *tv, int sign, double *a);
void orc_getx(const orc_cfg *c, const double *const f[6], const double *grd, const double *pv,
              const double *dv, const double *tv, double dt, const double *cc, const double *tt4,
              const double *res, double *x);
double orc_localdt(const orc_cfg *c, const double *const f[6], const double *grd, const double *pv,
                   const double *dv, const double *tv);
#endif
