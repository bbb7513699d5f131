/*
 * rwpt.h -- C ABI of the MI355X random-walk particle-tracking engine.
 *
 * Drop-in boundary for MODPATH-RW's hot path: the per-particle
 * TrackPath/TrackCell/TrackSubCell random walk and the gpkde histogram and
 * density reconstruction.  All arguments are plain pointers and sizes
 * (host memory unless a name says "device"); no C++ or torch types cross
 * this boundary.  Every entry point returns 0 on success or a negative
 * RWPT_E* code; rwpt_last_error() then holds a message.  The engine never
 * exits the process (the reference calls ustop, UtilMisc.f90:33-49).
 *
 * Conventions (the reference's):
 *  - cells are 1-based MODFLOW numbers (k-1)*nrow*ncol + (i-1)*ncol + j and
 *    per-cell arrays are ordered the same way (layer slowest, column fastest);
 *  - face flows use the MODFLOW-2005 signs ModpathCellData.f90:1315-1351
 *    consumes: frf = flow from column j to j+1, fff = row i to i+1,
 *    flf = layer k to k+1; bflow holds 6 boundary flows per cell;
 *  - local coordinates are in [0,1]; faces 1..6 = x-, x+, y-(front),
 *    y+(back), z-(bottom), z+(top);
 *  - particle status codes are those of MPathRW.F90:1256-1403:
 *    0 pending, 1 active, 2 boundary face, 3 weak sink, 4 weak source,
 *    5 no exit possible, 6 stop zone, 7 inactive cell, 8 unreleased,
 *    9 undefined.
 *
 * Threading: one host thread per engine; calls on one engine do not
 * overlap.  State stays resident in HBM between calls.
 */
#ifndef RWPT_H
#define RWPT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RWPT_OK 0
#define RWPT_E_ARG (-1)      /* invalid argument */
#define RWPT_E_STATE (-2)    /* call out of order (e.g. advance before flow) */
#define RWPT_E_DEVICE (-3)   /* HIP runtime error */
#define RWPT_E_NOMEM (-4)    /* device allocation failed */
#define RWPT_E_UNSUPPORTED (-5)
#define RWPT_E_IO (-6)       /* a file could not be opened or written */

typedef struct rwpt_engine rwpt_engine;

/* Structured grid (RectangularGridDis.f90 DIS arrays; the origin is the
 * left/front edge at x = 0, y = 0 as in RectangularGridDis.f90:295-361). */
typedef struct {
  int32_t ncol, nrow, nlay;
  const double *delr; /* ncol column widths (DX)                        */
  const double *delc; /* nrow row widths (DY)                           */
  const double *top;  /* ncell cell top elevations                      */
  const double *bot;  /* ncell cell bottom elevations                   */
} rwpt_grid;

/* Unstructured grid of rectangular cells: MODFLOW-USG or a MODFLOW 6 DISV
 * smoothed quadtree (RectangularGridDisvMf6.f90:47-574,
 * RectangularGridDisuMfusg.f90): MODFLOW's 1-based CSR connectivity, each
 * row led by the cell itself, and every cell's centre, sizes, layer and
 * elevations, as ModflowRectangularGridType holds them after ReadGRB /
 * CheckRectangular.  A horizontal face meets one cell of its size, one
 * twice its size or two of half its size; cells are then tracked in 2 x 2
 * sub-cells as the reference does. */
typedef struct {
  int32_t ncell, nlay;
  int64_t nja;
  const int32_t *ia;        /* ncell+1, 1-based (IA)                      */
  const int32_t *ja;        /* nja, 1-based (JA)                          */
  const int32_t *layer;     /* ncell, 1-based layer of each cell          */
  const double *cell_x, *cell_y;  /* ncell cell centres (CellX, CellY)    */
  const double *del_x, *del_y;    /* ncell cell sizes (DelX, DelY)        */
  const double *top, *bot;        /* ncell elevations                      */
  const int32_t *cell_type; /* ncell ICELLTYPE (GetDZRW's layer type) or NULL */
} rwpt_usg_grid;

/* Tracking options: the random-walk fields of ParticleTrackingOptionsType
 * (ParticleTrackingOptions.f90:20-31) read from RWOPTS
 * (ModpathSimulationData.f90:1770-2021). */
typedef struct {
  int32_t advection_kind;     /* 1 EXPONENTIAL, 2 EULERIAN                */
  int32_t time_step_kind;     /* 1 ADV, 2 DISP, 3 MIN_ADV_DISP, 4 FIXED   */
  double time_step_params[2]; /* timeStepParameters(1:2)                  */
  int32_t dimension_mask[3];  /* dimensionMask, 0 or 1 each              */
  int32_t backward_tracking;  /* BackwardTracking                        */
  int32_t stop_zone;          /* StopZone (0: none)                      */
  int32_t stop_at_weak_sinks;
  int32_t stop_at_weak_sources;
  int32_t steady_state;       /* FlowModelData%SteadyState               */
  int32_t precision;          /* 0: plain FP64 (the reference's numbers);
                                 1: FP32 cell-local positions and sub-step
                                    arithmetic, FP64 times and Newton exit
                                    solves, the FP64 run's random draws;
                                 2: as 1 with normals from 4 draws (16-bit
                                    Irwin-Hall) instead of 12            */
  uint64_t base_seed[2];      /* baseseed, TrackSubCell.F90:553-555      */
} rwpt_options;

/* Options of the gpkde reconstruction (ParticleTrackingOptions.f90:33-63,
 * ReadGPKDEData ModpathSimulationData.f90:816-1273). */
typedef struct {
  double origin[3];      /* gpkdeDomainOrigin                            */
  double bin_size[3];    /* gpkdeBinSize                                 */
  int32_t nbins[3];      /* domain size / bin size                       */
  int32_t mode;          /* 0 histogram only; 1 one Gaussian density pass
                            with the Silverman bandwidth (uniform h)     */
  int32_t weighted;      /* 1: particles carry mass                     */
  double scaling_factor; /* multiplies the density (gpkdeScalingFactor)  */
  int32_t unit_volume;   /* mode 1: 1 returns the density times the bin
                            volume before scaling_factor, as
                            ComputeDensity(unitVolume=.true.) does for
                            MPathRW.F90:1626-1642; 0 the density itself */
  int32_t n_opt_loops;   /* gpkdeNOptLoops: 0 one pass with the Silverman
                            bandwidth; > 0 the bandwidth optimisation of
                            ComputeDensityOptimization with brute-force
                            kernels (GridProjectedKDE.F90:4396-5160), 2-D
                            and 3-D reconstructions                      */
  int32_t isotropic;     /* gpkdeIsotropicKernels                        */
  int32_t skip_error_convergence; /* gpkdeSkipError: run all loops        */
  double rel_error_convergence;   /* gpkdeRelErrorConvergence (<= 0: 0.02) */
  int32_t initial_smoothing;      /* gpkdeInitialSmoothingFormat
                                     (GridProjectedKDE.F90:631-650, 4109-4117):
                                     0 Silverman, times the factor when it
                                     is > 0 and != 1; 1 factor * bin
                                     distance (bin volume ^ (1/dims with a
                                     bin size))                          */
  double initial_smoothing_factor; /* gpkdeBinSizeFactor (0 acts as 1)   */
  int32_t bound_kernel_size;      /* gpkdeBoundKernelSize (:788-858): 0 by
                                     the domain, 1 by min/max h over bin
                                     size, 2 unbounded                   */
  double min_h_over_delta;        /* gpkdeKDBParams(1): smallest h/delta
                                     of the kernel database and the lower
                                     bound of bound format 1              */
  double max_h_over_delta;        /* gpkdeKDBParams(3): largest h/delta   */
  int32_t effective_weight_format; /* gpkdeEffectiveWeightFormat
                                     (Histogram.F90:276-490): 0 Kish
                                     effective mass, 1 average mass, 2
                                     point counts drive the loop and the
                                     masses the last pass, 3 as 2 with the
                                     per-bin Kish counts                 */
  int32_t kernel_database;        /* gpkdeKernelDatabase: 1 kernels come
                                     from the log-spaced database of
                                     InitializeKernelDatabaseFlat
                                     (GridProjectedKDE.F90:3138-3526) with
                                     its lookups (:5727-6221); 0 brute force */
  double delta_h_over_delta;      /* gpkdeKDBParams(2): database spacing   */
  int32_t member;                 /* ensemble engines: the member whose
                                     particles are reconstructed (0-based),
                                     -1 all members together; ignored with
                                     one member                          */
  int32_t initial;                /* 1: the particles pending release
                                     (status 0) whose release time is 0, at
                                     their loaded locations: MPathRW's
                                     initial-condition reconstruction
                                     (MPathRW.F90:1043-1100); 0: the active
                                     particles (status 1)                */
} rwpt_kde_options;

/* Fills *opt with the reference defaults (EULERIAN, FIXED, 3-D, forward,
 * steady, FP64, baseseed). */
void rwpt_default_options(rwpt_options *opt);

/* ParticleTrackingEngine%Initialize, ParticleTrackingEngine.f90:393-492 */
int rwpt_create(const rwpt_grid *grid, const rwpt_options *opt,
                rwpt_engine **out);

/* An ensemble of `members` independent copies of the grid in one engine
 * (no counterpart in the reference, which runs one model per process): the
 * members are stacked in HBM as one grid of members * nlay layers whose
 * vertical connections are cut between members, so one rwpt_advance walks
 * every member's particles in the same launches.  Cells are numbered
 * member-major: cell c of member m (both 0-based) is m * ncell + c + 1 in
 * rwpt_load_particles / rwpt_get_particles, and every per-cell array of
 * rwpt_set_flow, rwpt_set_cell_extras and rwpt_set_dispersion (uniform 0)
 * holds members * ncell values.  A member's results are those of the same
 * member run alone with stream_offset = its first particle's index.
 * rwpt_create is rwpt_create_ensemble with one member. */
int rwpt_create_ensemble(const rwpt_grid *grid, const rwpt_options *opt,
                         int32_t members, rwpt_engine **out);
/* ParticleTrackingEngine%Initialize on an unstructured grid
 * (ParticleTrackingEngine.f90:393-492, grid types 2 and 4: the unstructured
 * cell fill pr_FillTransportCellBufferUnstructured :711-748 and neighbour
 * fill pr_FillNeighborCellDataUnstructured :1257-2005).  Needs precision 0
 * (FP64).  The flow comes from rwpt_set_flow_usg; rwpt_set_cell_extras,
 * rwpt_set_dispersion (uniform 0 or 1, per cell), the particle calls,
 * rwpt_advance, rwpt_global_xyz, rwpt_kde and rwpt_kde_write work as on a
 * structured grid (cells numbered as the grid's).  Particle locations are
 * local coordinates of the cell, as pr_RWPTrackPath reports them. */
int rwpt_create_usg(const rwpt_usg_grid *grid, const rwpt_options *opt,
                    rwpt_engine **out);
/* FlowModelData's arrays on an unstructured grid (FlowModelData.f90:217-622):
 * flowja (nja) is FLOW-JA-FACE, entry q of row n the flow into n from
 * ja[q]; bflow (6*ncell, IFACE boundary flows) and storage (ncell) may be
 * NULL; the others as in rwpt_set_flow. */
int rwpt_set_flow_usg(rwpt_engine *e, const double *flowja, const double *bflow,
                      const double *storage, const double *porosity,
                      const double *retardation, const int32_t *ibound,
                      const int32_t *icbound, int32_t default_icbound);
/* Tracking units of an unstructured engine (one per unrefined cell, four
 * per refined cell), -1 for a structured engine. */
int64_t rwpt_usg_unit_count(const rwpt_engine *e);
/* Checking copy of the unit records (no counterpart in the reference):
 * RWPT_USG_TABLE_FIELDS doubles per unit: cell, sub-row, sub-column,
 * sub-cells of the cell, connection and mass boundary of faces 1-6,
 * sub-cell has exit, cell has exit, dry, partially dry, VX1..VZ2, DX DY DZ,
 * offsets x0 x1 y0 y1, porosity, retardation, time step, source, sink,
 * corner discharge (8 corners x 4), corner porosity (8), Dxxx Dyyy Dzzz
 * Dxyx Dxzx Dyzy (8 each). */
#define RWPT_USG_TABLE_FIELDS 126
int rwpt_get_usg_tables(rwpt_engine *e, double *out);
int rwpt_destroy(rwpt_engine *e);
const char *rwpt_last_error(const rwpt_engine *e);

/* Flow and cell properties (FlowModelData, ModpathBasicData, IMP file):
 * frf/fff/flf ncell each; bflow 6*ncell or NULL; retardation, ibound,
 * icbound may be NULL (1.0, all active, 0).  default_icbound is the mass
 * boundary of faces without a neighbour (TransportModelData.f90:194-204). */
int rwpt_set_flow(rwpt_engine *e, const double *frf, const double *fff,
                  const double *flf, const double *bflow,
                  const double *porosity, const double *retardation,
                  const int32_t *ibound, const int32_t *icbound,
                  int32_t default_icbound);

/* Optional per-cell data for convertible layers, zones and weak sinks
 * (head/layer_type drive GetDZRW, ModpathCellData.f90:3208-3256). */
int rwpt_set_cell_extras(rwpt_engine *e, const double *head,
                         const int32_t *layer_type, const int32_t *zone,
                         const double *source_flow, const double *sink_flow);

/* Grid%SaturatedTop of the loaded time step (FlowModelData.f90:255-300:
 * the head in a wet convertible cell whose head lies between its bottom
 * and top, the bottom of a dry MODFLOW 6 cell whose head is below it, the
 * top otherwise), ncell values, or NULL for the cell tops.  It is the top
 * of ConvertToModelXYZ (ModflowRectangularGrid.f90:524-549): the global z
 * of rwpt_global_xyz, the KDE's coordinates and the written records.  The
 * tracking itself uses the heads of rwpt_set_cell_extras. */
int rwpt_set_saturated_top(rwpt_engine *e, const double *saturated_top);

/* Dispersion (DSP file, TransportModelData.f90:583-1040): model 0 linear
 * isotropic, 1 axisymmetric.  uniform=1 => each array holds one value,
 * uniform=2 => one value per ensemble member, uniform=0 => one value per
 * cell.  aLV/aTV may be NULL for model 0. */
int rwpt_set_dispersion(rwpt_engine *e, int32_t model, int32_t uniform,
                        const double *aL, const double *aT, const double *aLV,
                        const double *aTV, const double *dmeff);

/* Particles as structure of arrays.  rng holds 2 words per particle (the
 * xoroshiro128+ state); NULL derives stream p = jump^(stream_offset+p) of
 * base_seed, the stream rng_zigset gives OpenMP thread p
 * (rng_par_zig.f90:183-191). */
int rwpt_load_particles(rwpt_engine *e, int64_t n, const int32_t *cell,
                        const double *lx, const double *ly, const double *lz,
                        const double *t, const int32_t *status,
                        const uint64_t *rng, int64_t stream_offset);
int rwpt_get_particles(rwpt_engine *e, int32_t *cell, double *lx, double *ly,
                       double *lz, double *t, int32_t *status, uint64_t *rng);
int64_t rwpt_particle_count(const rwpt_engine *e);

/* Advances every active particle to t_stop (one TRACKING_INTERVAL_LOOP
 * pass, MPathRW.F90:1119-1725, with pr_RWPTrackPath per particle).
 * Pending particles are released first.  Outputs (optional): particles
 * active after the pass and RW sub-steps performed. */
int rwpt_advance(rwpt_engine *e, double t_stop, int64_t *n_active,
                 int64_t *n_substeps);

/* Sets the engine time (the start of the next tracking interval, the
 * `time` of MPathRW.F90:1256-1262 used for the release test). */
int rwpt_set_time(rwpt_engine *e, double t);

/* FlowModelData%SteadyState of the time step whose flow is loaded: the
 * reference copies it into TrackCell%SteadyState at every time step
 * (ParticleTrackingEngine.f90:893, 1091), after FlowModelData%LoadTimeStep
 * has set it from the step's storage flows.  It decides, with stop at weak
 * sinks (forward) or weak sources (backward), whether a cell without an
 * exit face stops its particles (TrackCell.f90:455-497).  rwpt_create
 * takes the first value from rwpt_options.steady_state. */
int rwpt_set_steady_state(rwpt_engine *e, int32_t steady_state);

/* Launch shape of the walk (no counterpart in the reference: the
 * reference walks each particle to t_stop in one call).  budget = RW
 * iterations per particle per launch (the loop state is saved between
 * launches, so results do not depend on it), block = workgroup size,
 * sort = re-sort unfinished particles by cell between launches.
 * block must be 64, 128 or 256; min_waves (1, 3 or 4) selects a build of
 * the walk register-limited for that many waves per SIMD.
 * Defaults 64, 256, 1, 1. */
int rwpt_set_tuning(rwpt_engine *e, int32_t budget, int32_t block,
                    int32_t sort, int32_t min_waves);
/* Graph mode of rwpt_advance (no counterpart in the reference): with
 * launches = K > 0 (even, at most 256) the walk of an interval continues,
 * once its list of unfinished particles holds at most max_list of them
 * (max_list <= 0: from the first launch), as replays of a HIP graph of K
 * walk launches captured on the engine stream, each launch reading its
 * work-list length from device memory, so that the host waits once per K
 * launches instead of once per launch; those lists are not sorted by cell.
 * The launches before stay eager and cell-sorted (large lists walk faster
 * sorted).  Results are bit-identical to the eager walk (they depend on no
 * launch shape).  launches 0 (the default) turns it off. */
int rwpt_set_graph(rwpt_engine *e, int32_t launches, int64_t max_list);
/* Launch-shape autotuner (no counterpart in the reference): times trial
 * advances from the current state to t_probe over workgroup size and
 * occupancy variant, per-launch iteration budget and work-list sort (a
 * staged search of rwpt_set_tuning's space), restores the particle state
 * and engine time, and keeps the fastest shape.  Results never depend on
 * the shape (the walk is bit-identical under every tuning), only the time.
 * Any output pointer may be NULL; ms_best is the wall time of the best
 * trial, trial_substeps the RW sub-steps of all trials (their walk
 * launches count in rwpt_kernel_time like those of rwpt_advance).  Needs
 * twice the particle state in device memory while it runs. */
int rwpt_autotune(rwpt_engine *e, double t_probe, int32_t *budget, int32_t *block,
                  int32_t *sort, int32_t *min_waves, double *ms_best,
                  int64_t *trial_substeps);
/* Walk launches the last rwpt_advance used. */
int rwpt_last_launches(const rwpt_engine *e);

/* Cumulative RW sub-steps per particle since rwpt_load_particles. */
int rwpt_get_substeps(rwpt_engine *e, int64_t *substeps);

/* Global coordinates of the particles (ConvertToModelXYZ,
 * ModflowRectangularGrid.f90:524-549). */
int rwpt_global_xyz(rwpt_engine *e, double *gx, double *gy, double *gz);

/* Concentration reconstruction over the active particles
 * (GridProjectedKDE%ComputeDensity, GridProjectedKDE.F90:3598-4392, with
 * brute-force kernels: one pass with the initial bandwidth when
 * n_opt_loops = 0, the bandwidth optimisation otherwise).  density and
 * histogram are nbins[0]*nbins[1]*nbins[2] doubles, x fastest (either may
 * be NULL); mass (one per particle) may be NULL for unit masses.  The
 * histogram holds the per-bin sums of the masses; mode 0 returns it
 * divided by the bin volume as the density.  A dimension of one bin may
 * have bin size 0 (compressed: no check of the coordinate); with a bin
 * size it drops the points outside that bin, as the reference does. */
int rwpt_kde(rwpt_engine *e, const rwpt_kde_options *opt, const double *mass,
             double *density, double *histogram);

/* Largest grid (bins) and 1-D kernel storage (weights over the active
 * dimensions, sum of ceiling(3h/bin) + 1) of rwpt_kde_device. */
#define RWPT_KDE_DEVICE_MAX_CELLS 4194304
#define RWPT_KDE_DEVICE_MAX_TAPS 8192

/* rwpt_kde without any device allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * one-pass reconstruction of rwpt_kde (n_opt_loops 0, mode 0 or 1) over the
 * engine's particles with unit masses, written into caller-owned DEVICE
 * buffers density_dev and histogram_dev of nbins[0]*nbins[1]*nbins[2]
 * doubles (x fastest).  Results equal rwpt_kde's bit for bit.  Its scratch
 * memory is a fixed device array of about 32 MiB per GPU, held until the
 * process ends and shared by every engine on that GPU: calls that overlap
 * on one GPU race.  Grids above RWPT_KDE_DEVICE_MAX_CELLS bins and kernels
 * above RWPT_KDE_DEVICE_MAX_TAPS weights return RWPT_E_UNSUPPORTED;
 * weighted particles (opt->weighted) and n_opt_loops > 0 return
 * RWPT_E_UNSUPPORTED (their sorts and per-bin arrays need allocations).
 * rwpt_kde_info reports the call like rwpt_kde. */
int rwpt_kde_device(rwpt_engine *e, const rwpt_kde_options *opt, double *density_dev,
                    double *histogram_dev);

/* Quantities of the last rwpt_kde call: the Silverman bandwidth used, the
 * number of active particles passed, the histogram's nPoints, the
 * effective mass (Histogram.F90:332-341) and the optimisation loops run. */
int rwpt_kde_info(const rwpt_engine *e, double *bandwidth, double *points,
                  double *npoints, double *eff_mass, int32_t *loops);

/* Density rows as MODPATH-RW's gpkde output file holds them
 * (GridProjectedKDE%ExportDensityUnit / %ExportDensityUnitBinary,
 * GridProjectedKDE.F90:6504-7021, in the branch MPathRW.F90:1626-1642
 * reaches: time index, time and particle group present): one row per bin
 * whose density is not <= 0, x fastest.  format 0 text (I8 and ES18.9E3
 * fields), 1 binary stream (int32 and float64 fields, no record markers);
 * column_format (gpkdeOutColFormat) 0 bin indices, 1 bin indices and bin
 * centre coordinates, 2 centre coordinates.  density and histogram are the
 * rwpt_kde outputs for the grid of opt; the histogram column is written
 * times histogram_scale (MPathRW passes gpkdeScalingFactor as the
 * histogramScalingFactor).  append 0 replaces the file, 1 appends (the
 * reference keeps one unit open for the run).  Needs no engine or device. */
int rwpt_kde_write(const char *path, int32_t append, int32_t format,
                   int32_t column_format, int32_t time_id, double time,
                   int32_t group_id, const rwpt_kde_options *opt,
                   const double *density, const double *histogram,
                   double histogram_scale);

/* MODPATH-RW's timeseries file (no counterpart needs an engine there: the
 * reference writes it from its ParticleType array): the header of
 * WriteTimeseriesHeader (ParticleManager.F90:208-226), which replaces the
 * file, and the records of WriteTimeseriesRecord (:229-250) for the
 * engine's particles at one time point, appended in particle order:
 * (2I8,ES18.9E3,I10,I5,2I10,6ES18.9E3,I10) = time point index, time step,
 * time, sequence number, group, particle ID, cell, local x y z, model x y,
 * global z, layer.  sequence, id and group hold one value per particle;
 * select (NULL: the particles with status 1) picks the records written.
 * Cell numbers and layers are those of the particle's ensemble member. */
int rwpt_write_timeseries_header(const char *path, int32_t tracking_direction,
                                 double reference_time, double origin_x,
                                 double origin_y, double rotation);
int rwpt_write_timeseries(rwpt_engine *e, const char *path, int32_t time_point,
                          int32_t time_step, double time, const int32_t *sequence,
                          const int32_t *id, const int32_t *group,
                          const uint8_t *select);

/* MODPATH-RW's endpoint file for the engine's particles: format 0
 * WriteEndpoints (MODPATH_ENDPOINT_FILE 7 2, ParticleManager.F90:51-125),
 * format 1 WriteEndpointsMassParticles (MODPATHRW_ENDPOINT_FILE 1 0,
 * :128-205).  Particles are listed group by group as the engine holds them;
 * statuses 0 and 8 are counted but not written.  group (1-based),
 * sequence, id and the initial cell, local coordinates, time and face
 * (format 0) hold one value per particle; group_names (written padded to
 * 20 characters) and group_solute (format 1) one per group; mass (format
 * 1) one per particle.  Zones are those of rwpt_set_cell_extras (1 when
 * none were set).  Replaces the file. */
int rwpt_write_endpoints(rwpt_engine *e, const char *path, int32_t format,
                         int32_t tracking_direction, double reference_time,
                         double origin_x, double origin_y, double rotation,
                         int32_t n_groups, const char *const *group_names,
                         const int32_t *group_solute, const int32_t *group,
                         const int32_t *sequence, const int32_t *id,
                         const int32_t *initial_cell, const double *initial_lx,
                         const double *initial_ly, const double *initial_lz,
                         const double *initial_time, const int32_t *initial_face,
                         const double *mass);

/* MODPATH-RW's observation records (WriteResidentObsRecord[Binary],
 * WriteSinkObsRecord[Binary], Observation.f90:114-221, as MPathRW.F90
 * :1344-1392 and :1449-1530 writes them) for the engine's particles in the
 * observation's cells (n_cells engine cell numbers): style 1 resident, the
 * particles with status 1, at a time point after rwpt_advance; style 2
 * sink, the particles stopped at a weak sink (status 3; weak source, 4, in
 * backward tracking) with flow_rate the observation's summed sink flow.
 * format 0 text, (2I8,ES18.9E3,I10,ES18.9E3,2I5,2I10,5ES18.9E3) resident
 * and (...,ES18.9E3) sink: time point, time step, time, ID, mass, group,
 * solute, cell, layer, then the cell's retardation, water volume (porosity
 * DX DY and the saturated thickness; 0 when dry) and the model x, y, global
 * z, or the flow rate; format 1 the same fields as a stream of int32 and
 * float64.  id, group (1-based) and mass (NULL: 1) per particle,
 * group_solute per group; select (may be NULL): style 1 the particles
 * written instead of those with status 1 (MPathRW's TimeseriesOutputOption
 * 1 writes every particle in the cells), style 2 a limit on the stopped
 * particles, e.g. to those stopped in the last advance.  append 0 replaces
 * the file. */
int rwpt_write_observation(rwpt_engine *e, const char *path, int32_t append,
                           int32_t style, int32_t format, int32_t n_cells,
                           const int32_t *cells, int32_t time_point, int32_t time_step,
                           const int32_t *id, const int32_t *group,
                           const int32_t *group_solute, const double *mass,
                           double flow_rate, const uint8_t *select);

/* Per-cell tables as built on the device (checking aid), ncell * 104
 * doubles per cell: face velocities (6), DX DY DZ, porosity, R, dt_cell,
 * corner discharge qx qy qz |q| (8 x 4), corner porosity (8), corner phi*D
 * Dxx Dyy Dzz Dxy Dxz Dyz (6 x 8), dry flag, sub-cell exit flag, and the
 * cell's source and sink flows (boundary flows on faces with an active
 * neighbour included). */
int rwpt_get_cell_tables(rwpt_engine *e, double *out);

/* Pathwise adjoint of an observation misfit (no counterpart in the
 * reference; the gradient of a calibration objective with respect to the
 * transport parameters).  Advances the engine through t_windows[0..n_windows-1]
 * exactly as n_windows rwpt_advance calls would (the windows are also the
 * checkpoints of the reverse sweep), then evaluates at the final time
 *   conc[o] = sum over active particles of m_p prod_d N((X_pd - x_od)/h_d)/h_d
 *   misfit  = sum_o (conc[o] - target[o])^2
 * with X_p the global coordinates (rwpt_global_xyz), N the standard normal
 * density, obs_xyz 3 * n_obs global coordinates, bandwidth h[3] (dimensions
 * with h_d = 0 are left out), mass per particle (NULL: 1).  With any
 * gradient pointer given, also the derivative of misfit with common random
 * numbers: grad_al, grad_at, grad_dm (ncell each, may be NULL) with respect
 * to each cell's aLH, aTH and Dm, and grad_porosity_scale with respect to a
 * factor s multiplying every cell's porosity (at s = 1).  The derivative is
 * the pathwise one: each particle's branches (cells crossed, jump-limiter
 * cuts, faces) held fixed.  Gradients need an FP64 engine, EULERIAN
 * advection, the linear dispersion model and a 2-D or 3-D walk.  ms (may be
 * NULL) receives the wall time of the call. */
int rwpt_adjoint(rwpt_engine *e, int32_t n_windows, const double *t_windows,
                 int32_t n_obs, const double *obs_xyz, const double *target,
                 const double *bandwidth, const double *mass, double *misfit,
                 double *conc, double *grad_al, double *grad_at, double *grad_dm,
                 double *grad_porosity_scale, double *ms);

/* Options of rwpt_adjoint's reverse sweep (no counterpart in the
 * reference).  The exact pathwise gradient of the walk grows along a path
 * by a random factor at every cell exit: a sub-step that ends at a face at
 * time tau is followed by one over the rest of the time step with fresh
 * draws, so the final position depends on tau through both random
 * displacements, B xi1 sqrt(tau) and B xi2 sqrt(dt - tau), and the
 * Jacobian of one crossing is (v + B xi2 / (2 sqrt(dt - tau))) /
 * (v + B xi1 / (2 sqrt(tau))) in one dimension (1 for the continuous
 * process).  These options trade that variance for a bias; all zero (the
 * default) is the exact pathwise gradient. */
typedef struct {
  /* exits with kappa = |2 A m + B| / (|2 A m| + |B|) < graze_cut lose the
   * derivative of their exit time (the exit root m = sqrt(dt) of
   * A m^2 + B m - dI = 0 on the exit axis; kappa near 0 is a grazing,
   * double root); 0 keeps every term */
  double graze_cut;
  /* K > 0 sweeps only the last K windows: the gradient holds the
   * sensitivities of those windows, the state at the start of window
   * n_windows - K counting as fixed */
  int32_t reverse_windows;
  /* L > 0 restarts the adjoint at the start of every L-th window counted
   * from the last (straight-through): each particle's adjoint becomes its
   * final-time adjoint of the global position, expressed in the cell it
   * starts that window in, with a zero time adjoint, as if the transport
   * from there to the final time moved every displacement unchanged */
  int32_t tangent_windows;
  /* 1 drops the derivative of every random displacement B xi sqrt(dt) with
   * respect to the sub-step duration dt (exit roots keep theirs), which
   * makes the crossing Jacobian above v / (v + B xi1 / (2 sqrt(tau))) */
  int32_t frozen_noise;
  /* 1 makes the next gradients tally the sweep's statistics for
   * rwpt_adjoint_stats (a slower sweep) */
  int32_t diagnostics;
} rwpt_adjoint_opts;

/* Sets the options of the engine's following rwpt_adjoint calls. */
int rwpt_adjoint_options(rwpt_engine *e, const rwpt_adjoint_opts *opt);

/* Statistics of the last rwpt_adjoint gradient taken with diagnostics on,
 * RWPT_ADJ_NSTATS doubles.  Every reverse iteration is one of five kinds
 * k: 0 a sub-step inside the cell, 1 a sub-step cut by the jump limiter,
 * 2 a cell exit with kappa >= 1e-3, 3 a cell exit with kappa < 1e-3,
 * 4 a rare iteration (restart, rebound; dual-number Jacobians).  Its
 * growth is |adjoint of its input| / |adjoint of its output| (L1 norm of
 * the x, y, z, t adjoints).
 *   out[k]       iterations of kind k
 *   out[5 + k]   sum of log10(growth) over those with growth > 1
 *   out[10 + k]  largest log10(growth) (0 when none grows)
 *   out[15 + k]  iterations with growth > 10
 *   out[20 .. 26] exits by kappa: < 1e-10, [1e-10, 1e-8), [1e-8, 1e-6),
 *                [1e-6, 1e-4), [1e-4, 1e-2), [1e-2, 1e-1), >= 1e-1
 *   out[27]      exits whose exit-time derivative graze_cut dropped
 *   out[28]      particles swept; out[29] the largest log10 of a
 *                particle's adjoint growth from the final time to the
 *                start of the first window swept; out[30 .. 32] particles
 *                whose growth exceeds 1e3, 1e6, 1e9 */
#define RWPT_ADJ_NSTATS 33
int rwpt_adjoint_stats(rwpt_engine *e, double *out);

/* Device timing of the last rwpt_advance/rwpt_kde calls, measured with HIP
 * events on the engine stream: which = 0 advance kernel, 1 histogram,
 * 2 density.  Returns the summed milliseconds and the launch count. */
int rwpt_kernel_time(rwpt_engine *e, int32_t which, double *ms,
                     int64_t *launches);
int rwpt_reset_kernel_time(rwpt_engine *e);

#ifdef __cplusplus
}
#endif
#endif
