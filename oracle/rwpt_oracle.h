/*
 * CPU oracle for the MODPATH-RW random-walk hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (upc-ghs/modpath-rw, Fortran) used as the checker for the HIP
 * product path and as the CPU baseline timed by bench.py.  Nothing in the
 * product library links, calls or loads it.  Only tests/, the smoke check and
 * the cpu_baseline leg of bench.py may use it.
 *
 * Conventions follow the reference: 1-based MODFLOW cell numbers
 * (k-1)*nrow*ncol + (i-1)*ncol + j, MODFLOW-2005 face-flow signs, faces
 * 1..6 = left, right, front, back, bottom, top, corners 1..8 = 000..111.
 */
#ifndef RWPT_ORACLE_H
#define RWPT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* grid (ModflowRectangularGrid / RectangularGridDis) */
  int ncol, nrow, nlay;
  const double *delr;        /* ncol */
  const double *delc;        /* nrow */
  const double *top;         /* ncell: cell top elevation */
  const double *bot;         /* ncell: cell bottom elevation */
  /* flow model data (FlowModelData) */
  const int *ibound;         /* ncell, IBoundTS; NULL => all active */
  const double *frf, *fff, *flf; /* FlowsRightFace/FrontFace/LowerFace */
  const double *bflow;       /* 6*ncell boundary flows, NULL => 0 */
  const double *porosity;    /* ncell */
  const double *retardation; /* ncell, NULL => 1 */
  const double *head;        /* ncell, NULL => head = top */
  const int *layer_type;     /* ncell, NULL => 0 (confined) */
  const double *source_flow, *sink_flow; /* ncell or NULL */
  const int *zone;           /* ncell or NULL */
  /* transport model data */
  const int *icbound;        /* ncell, NULL => 0 */
  int default_icbound;
  int disp_model;            /* 0 linear isotropic, 1 axisymmetric */
  int disp_uniform;          /* 1 => parameter arrays have one entry */
  const double *aL, *aT, *aLV, *aTV, *dmeff;
  /* tracking options (ParticleTrackingOptionsType) */
  int advection_kind;        /* 1 exponential, 2 eulerian */
  int time_step_kind;        /* 1 ADV, 2 DISP, 3 MIN_ADV_DISP, 4 FIXED */
  double ts_params[2];
  int dim_mask[3];
  int uniform_porosity;      /* isUniformPorosity */
  int backward;
  int stop_weak_sinks, stop_weak_sources, stop_zone, steady_state;
} orc_problem;

/* xoroshiro128+ (rng_par_zig.f90) */
uint64_t orc_xoroshiro_next(uint64_t s[2]);
double orc_uniform(uint64_t s[2]);
double orc_normal12(uint64_t s[2]);
void orc_rng_jump(uint64_t s[2]);
/* stream p = jump^(first+p)(base) for p = 0..n-1, written as 2n words */
void orc_rng_streams(const uint64_t base[2], long first, long n, uint64_t *out);

/* Advance every particle with status 1 to t_stop, following
 * pr_RWPTrackPath + TrackCell + TrackSubCell.  Particles with status 0
 * are released following MPathRW.F90:1256-1302 with t as release time.
 * rng holds 2 words per particle (its stream state).  substeps[p] (optional)
 * is incremented by the number of RW time-loop iterations performed.
 * Returns 0 or a negative error code. */
int orc_advance(const orc_problem *P, long n, int *cell, double *lx,
                double *ly, double *lz, double *t, int *status,
                uint64_t *rng, double t_engine, double t_stop,
                long long *substeps, int nthreads);

/* Per-cell assembled quantities for checking the device cell tables.
 * out has ORC_CELL_FIELDS doubles per cell, see rwpt_oracle.c. */
#define ORC_CELL_FIELDS 104
int orc_cell_tables(const orc_problem *P, double *out);

/* Global coordinates (ConvertToModelXYZ) */
void orc_global_xyz(const orc_problem *P, long n, const int *cell,
                    const double *lx, const double *ly, const double *lz,
                    double *gx, double *gy, double *gz);

/* gpkde histogram (Histogram.F90:276-350): counts per bin, column-major
 * (x fastest), points outside dropped. */
void orc_histogram(long n, const double *x, const double *y, const double *z,
                   const double *w, const double origin[3],
                   const double binsize[3], const int nbins[3],
                   double *counts);

/* gpkde ComputeDensity with nOptimizationLoops = 0 and the brute-force
 * kernel: histogram (weights optional, effective-weight format 0), the
 * Silverman bandwidth over all points, one scatter of the reflected
 * multi-Gaussian of every active bin, divided by the bin volume.  Grids are
 * column-major (x fastest); dimensions with one bin are inactive.  hout
 * (optional) receives the bandwidth used. */
int orc_kde_density(long n, const double *x, const double *y,
                    const double *z, const double *w, const double origin[3],
                    const double binsize[3], const int nbins[3],
                    double *density, double *counts, double hout[3]);
/* as orc_kde_density with the initial smoothing selection (0 Silverman
 * times the factor when it is > 0 and != 1, 1 factor * bin distance,
 * GridProjectedKDE.F90:631-650, 4109-4117) */
int orc_kde_density_sel(long n, const double *x, const double *y,
                        const double *z, const double *w,
                        const double origin[3], const double binsize[3],
                        const int nbins[3], int selection, double factor,
                        double *density, double *counts, double hout[3]);

/* exit-time solvers for one axis (Newton: TrackSubCell.F90:2451-2554;
 * quadratic: one axis of :1576-1892, exit faces 1/2) */
double orc_newton_axis(double dt, double v, double v1, double v2, double dx,
                       double dI, double divD, double dB);
double orc_quadratic_axis(double dt, double n, double p, double d, double v,
                          double divD, double dB, int *exit_face);

/* Unstructured grids of rectangular cells (MODFLOW-USG, MODFLOW 6 DISV
 * smoothed quadtrees): the grid as CSR connectivity plus cell geometry,
 * FLOW-JA-FACE, and the options and per-cell arrays of `base` (top, bot,
 * ibound, bflow, porosity, retardation, head, layer_type, source_flow,
 * sink_flow, zone, icbound, dispersion and tracking options; its ncol,
 * nrow, nlay, delr, delc, frf, fff, flf are not used). */
typedef struct {
  orc_problem base;
  int ncell, nlay;
  long nja;
  const int *ia;         /* ncell+1, 1-based (MODFLOW IA)                   */
  const int *ja;         /* nja, 1-based, each row led by the cell itself    */
  const int *layer;      /* ncell, 1-based layer of each cell                */
  const double *cell_x, *cell_y, *del_x, *del_y; /* ncell                     */
  const double *flowja;  /* nja: flow into the row's cell from JA's cell     */
  const double *storage; /* ncell or NULL                                    */
} orc_usg_problem;

/* Sub-cell tables, ORC_USG_FIELDS doubles per sub-cell (rwpt_oracle_usg.c);
 * returns the record count (out NULL: count only) or a negative error. */
#define ORC_USG_FIELDS 126
long orc_usg_tables(const orc_usg_problem *U, double *out);
/* orc_advance on an unstructured grid; locations are those of the parent
 * cell (what pr_RWPTrackPath reports). */
int orc_usg_advance(const orc_usg_problem *U, long n, int *cell, double *lx,
                    double *ly, double *lz, double *t, int *status,
                    uint64_t *rng, double t_engine, double t_stop,
                    long long *substeps, int nthreads);
void orc_usg_global_xyz(const orc_usg_problem *U, long n, const int *cell,
                        const double *lx, const double *ly, const double *lz,
                        double *gx, double *gy, double *gz);

#ifdef __cplusplus
}
#endif
#endif
