// Internal declarations shared by the HIP translation units of librwpt.
//
// Device data layout (all resident in HBM for the engine's lifetime):
//  * particle state, structure of arrays: cell (i32), status (i32),
//    lx/ly/lz (f64), t (f64), rng (2 x u64), substeps (i64);
//  * one CellRec (768 B, 6 x 128-B lines) per grid cell holding everything a
//    random-walk sub-step reads: face pore velocities, sizes, porosity,
//    retardation, dispersivities, the cell time step, corner velocities and
//    the 9 dispersion-derivative terms in the form the trilinear derivative
//    consumes, connections and mass-boundary flags.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rwpt.h"

namespace rwpt {

// Field offsets (in doubles) inside a CellRec.
enum CellField : int {
  F_VX1 = 0, F_VX2, F_VY1, F_VY2, F_VZ1, F_VZ2,  // 0..5 face pore velocities
  F_DX = 6, F_DY, F_DZ,                          // 6..8
  F_PHI = 9, F_R = 10, F_DTCELL = 11,
  F_DVDX = 12, F_DVDY, F_DVDZ,                   // exponential gradients
  F_ALH = 15, F_ATH, F_ALV, F_ATV, F_DMEFF,      // 15..19
  F_VC = 20,        // 20..43: corner velocity [component][corner]
  F_DD = 44,        // 44..79: derivative terms [term][4]
  F_CONN = 80,      // 80..82: 6 x int32 connections
  F_MASSBND = 83,   // 83..85: 6 x int32 mass-boundary flags
  F_FLAGS = 86,     // int32 ibound, int32 flags
  F_ZONE = 87,      // int32 zone
  F_SOURCE = 88, F_SINK = 89,
  F_RDX = 90, F_RDY, F_RDZ, F_RR,  // 1/DX.. 1/R when a power of two, else 0
  CELL_DOUBLES = 96
};

// flag bits in the second int32 of F_FLAGS
enum : int32_t {
  CF_DRY = 1,             // GetDZRW dry (TrackSubCell returns InactiveCell)
  CF_SUBCELL_EXIT = 2,    // sub-cell HasExitFace
  CF_CELL_EXIT = 4,       // cell HasExitFace
  CF_VERIFY_PDRY = 8,     // VerifyDryCell partiallyDry (release of status 7)
  CF_SOURCE = 16,         // source flow != 0 (weak source)
  CF_SINK = 32,           // sink flow != 0 (weak sink)
  CF_DT_NONPOS = 64,      // cell time step <= 0
  CF_NB_INACTIVE = 256,   // << face-1 (6 bits): the neighbour across that
                          // face is a cell with IBOUND 0
};

// FP32-local cell record (precision 1 and 2), in the forms the FP32 walk
// evaluates with fused multiply-adds:
//   0..11   as in the FP64 record (face velocities, DX DY DZ, phi, R, dt_cell);
//   12..14  G_DV: VX2 - VX1, VY2 - VY1, VZ2 - VZ1 (the face velocity slope;
//           the exponential gradient is G_DV times 1/DX);
//   15..19  G_S2ALH..G_S2DM: 2 aLH / R, 2 aTH / R, 2 aLV / R, 2 aTV / R,
//           2 Dm / R (the dispersion factors under the square roots);
//   20..43  per velocity component, the 8 coefficients c0..c7 of the
//           trilinear corner interpolant c0 + c1 x + c2 y + c3 z + c4 xy +
//           c5 xz + c6 yz + c7 xyz;
//   44..79  per drift term, the 4 coefficients e0..e3 of the bilinear
//           interpolant e0 + e1 u + e2 w + e3 uw of its edge differences,
//           divided by phi R (so that the drift is their plain sum);
// then the integer fields one per slot and full reciprocals.
enum CellField32 : int {
  G_DV = 12,
  G_S2ALH = 15, G_S2ATH, G_S2ALV, G_S2ATV, G_S2DM,  // 15..19
  G_VC = 20,        // 20..43: trilinear coefficients [component][8]
  G_DD = 44,        // 44..79: bilinear coefficients [term][4]
  G_CONN = 80,      // 80..85: 6 x int32 connections
  G_MASSBND = 86,   // 86..91: 6 x int32 mass-boundary flags
  G_FLAGS = 92,     // int32 ibound, int32 flags
  G_ZONE = 94,
  G_SOURCE = 95, G_SINK = 96,
  G_RDX = 97, G_RDY, G_RDZ, G_RR, G_RPHI,  // 1/DX 1/DY 1/DZ 1/R 1/phi
  CELL_FLOATS = 104
};

// Record layout by arithmetic type: indices of the integer fields in int32
// units from the record start, of the other fields in R units.
template <typename R> struct Lay;
template <> struct Lay<double> {
  static constexpr int STRIDE = CELL_DOUBLES;
  static constexpr int CONN = 2 * F_CONN, MASSBND = 2 * F_MASSBND,
                       FLAGS = 2 * F_FLAGS, ZONE = 2 * F_ZONE;
  static constexpr int SOURCE = F_SOURCE, SINK = F_SINK, RDX = F_RDX,
                       RDY = F_RDY, RDZ = F_RDZ, RR = F_RR;
};
template <> struct Lay<float> {
  static constexpr int STRIDE = CELL_FLOATS;
  static constexpr int CONN = G_CONN, MASSBND = G_MASSBND, FLAGS = G_FLAGS,
                       ZONE = G_ZONE;
  static constexpr int SOURCE = G_SOURCE, SINK = G_SINK, RDX = G_RDX,
                       RDY = G_RDY, RDZ = G_RDZ, RR = G_RR;
};

// Layout of the checking copy of the cell tables (104 doubles per cell).
constexpr int DEBUG_FIELDS = 104;

struct KernelTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct DeviceProblem {
  // an ensemble of `members` copies of one grid is stored as one grid of
  // members * nlay_member layers whose vertical connections are cut at
  // every member boundary; cells are numbered member-major
  int ncol = 0, nrow = 0, nlay = 0, nlay_member = 0, members = 1;
  int64_t ncell = 0;           // all members
  int64_t ncell_member = 0;    // one member
  // device inputs
  double *delr = nullptr, *delc = nullptr, *top = nullptr, *bot = nullptr;
  double *frf = nullptr, *fff = nullptr, *flf = nullptr, *bflow = nullptr;
  double *porosity = nullptr, *retardation = nullptr;
  int32_t *ibound = nullptr, *icbound = nullptr;
  double *head = nullptr;
  int32_t *layer_type = nullptr, *zone = nullptr;
  double *source = nullptr, *sink = nullptr;
  double *disp = nullptr;  // 5 arrays of ndisp: aL aT aLV aTV dmeff
  double *sat_top = nullptr;  // SaturatedTop of the time step, null: top
  int64_t ndisp = 0;
  int32_t default_icbound = 0;
  int32_t disp_model = 0;
  int32_t uniform_porosity = 0;      // every member's porosity uniform
  int32_t *uniform_member = nullptr;  // per member: its porosity is uniform
  // assembled per-cell scratch of the table build
  double *q6 = nullptr;     // 6 * ncell assembled face flows
  double *dzrw = nullptr;   // ncell
  // tables
  double *cells = nullptr;  // ncell * CELL_DOUBLES
  float *cells32 = nullptr; // FP32-local tables (ncell * CELL_FLOATS)
  bool pow2_cells = false;  // every cell's DX, DY, DZ, R a power of two
};

// the top of ConvertToModelXYZ: SaturatedTop when it was set, else Top
inline const double *model_top(const DeviceProblem &P) {
  return P.sat_top ? P.sat_top : P.top;
}

struct Particles {
  int64_t n = 0, cap = 0;
  int32_t *cell = nullptr, *status = nullptr;
  double *lx = nullptr, *ly = nullptr, *lz = nullptr, *t = nullptr;
  uint64_t *rng = nullptr;
  int64_t *substeps = nullptr;
  // resumable walk state (see rwpt_walk.hip)
  int32_t *aux = nullptr;
  double *ix = nullptr, *iy = nullptr, *iz = nullptr, *t0 = nullptr;
  // FP32-local positions (precision 1, 2): the walk's copies of lx ly lz and
  // of the visit start; lx ly lz above are refreshed after every advance
  float *flx = nullptr, *fly = nullptr, *flz = nullptr;
  float *fix = nullptr, *fiy = nullptr, *fiz = nullptr;
  // work lists sorted by cell between launches
  int32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
};

struct Tuning {
  int32_t budget = 64;    // time-loop iterations per particle and launch
  int32_t block = 256;    // threads per workgroup of the walk
  int32_t sort = 1;       // sort the work list by cell between launches
  int32_t min_waves = 1;  // occupancy hint variant of the walk (1, 3, 4)
};

struct WalkParams {
  int32_t advection_kind, time_step_kind, ndim, dims[3];
  int32_t disp_model, backward, stop_zone, stop_weak_sinks, stop_weak_sources;
  int32_t steady_state;
  double t_engine, t_stop;
};

// Unstructured grids (rwpt_usg.hip).  A tracking unit is an unrefined
// cell or one of the 2 x 2 sub-cells of a refined one; units are numbered
// cell by cell (sub-cells row-major) and each has a CellRec.
//  per cell, USG_TOPO int32: potential connection count per face, sub-face
//    count per face (1 without connections), sub-face connections [6][4]
//    and their FLOW-JA-FACE positions [6][4] (-1 none);
//  per unit, USG_SIDE int32: cell (1-based), sub-row, sub-column, sub-cells
//    of the cell, then per face: the sibling unit across an internal face,
//    or for a face with a neighbour cell the face of that neighbour which
//    meets this cell plus 8 when ConvertFromNeighbor's full-face copy
//    applies (one connection on both sides), else 0;
//  per unit, 4 doubles: the sub-cell's offsets x0 x1 y0 y1 in its cell;
//  per unit, 18 x 2 int32: how each stencil entry is formed.
enum : int {
  UT_PCOUNT = 0, UT_SFC = 6, UT_CONN = 12, UT_QIDX = 36, USG_TOPO = 60,
  US_CELL = 0, US_ROW = 1, US_COL = 2, US_NSUB = 3, US_FACE = 4, USG_SIDE = 10,
  USG_DEBUG_FIELDS = 126,
};

struct UsgHost {
  int32_t nunit = 0;
  std::vector<int32_t> cell_unit, topo, side, slots;
  std::vector<double> off;
};

struct UsgData {
  int32_t ncell = 0, nunit = 0, nlay = 0;
  int64_t nja = 0;
  int32_t *topo = nullptr, *cell_unit = nullptr, *side = nullptr, *slots = nullptr;
  int32_t *layer = nullptr, *cell_type = nullptr;
  double *off = nullptr, *cx = nullptr, *cy = nullptr, *dx = nullptr, *dy = nullptr;
  double *flowja = nullptr, *storage = nullptr, *casm = nullptr;
  // particle state of the unstructured walk: the unit of a visit in
  // progress, the cell-entry point (parent-cell coordinates and time) and
  // the last pathline point before the entry (cell and coordinates)
  int32_t *unit = nullptr, *pcell = nullptr;
  double *ex = nullptr, *ey = nullptr, *ez = nullptr, *et = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;
  int64_t pcap = 0;
};

std::string usg_build_topology(const rwpt_usg_grid &g, UsgHost &H);

}  // namespace rwpt

struct rwpt_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  rwpt_options opt{};
  rwpt::DeviceProblem prob;
  rwpt::Particles part;
  rwpt::UsgData *usg = nullptr;  // unstructured grid (rwpt_create_usg), else null
  bool have_flow = false, have_disp = false, tables_dirty = true;
  bool tables32_dirty = true;  // FP32-local records behind the FP64 ones
  double t_engine = 0.0;
  uint64_t *jump_mats = nullptr;  // device copies of J^(2^k), k < 64
  unsigned long long *counters = nullptr;  // device counters
  double *kde_scratch = nullptr;
  int64_t kde_scratch_size = 0;
  double *kde_weights = nullptr;
  int64_t kde_weights_size = 0;
  double kde_last[4] = {0, 0, 0, 0};  // bandwidth, points, nPoints, effMass
  int32_t kde_loops = 0;              // optimisation loops of the last kde
  // options of rwpt_adjoint's reverse sweep (rwpt_adjoint_options) and the
  // statistics of its last gradient taken with diagnostics on
  rwpt_adjoint_opts adj_opt{0.0, 0, 0, 0, 0};
  double adj_stats[RWPT_ADJ_NSTATS] = {};
  double *xface = nullptr, *yface = nullptr;  // DIS face coordinates X, Y
  // grow-only device buffers of the gpkde optimisation loop (active bins,
  // per-bin arrays, curvature grids, kernel factors, tile lists), kept for
  // the engine's lifetime so that a reconstruction allocates nothing once
  // they are large enough
  static constexpr int kGkBuffers = 12;
  void *gk_mem[kGkBuffers] = {};
  size_t gk_bytes[kGkBuffers] = {};
  rwpt::KernelTimer timers[3];
  rwpt::Tuning tune;
  std::vector<hipEvent_t> walk_events;  // start/stop pairs per launch
  int64_t last_launches = 0;
  // timed event pairs of the last advance and the walk launches each spans
  // (one per eager launch, a graph's launches per replay)
  int64_t last_timed = 0;
  std::vector<int32_t> walk_event_launches;
  // graph mode (rwpt_set_graph): walk launches per captured graph, its
  // executable, and the device list length captured launches read
  int32_t graph_launches = 0;
  int64_t graph_max_list = 0;  // replays take over at this list length (<= 0: at once)
  hipGraphExec_t graph_exec = nullptr;
  const unsigned long long *nwork_dev = nullptr;
  std::string err;
};

namespace rwpt {

// launchers (return hipError_t as int)
hipError_t build_cell_tables(rwpt_engine *e, double *debug_out);
hipError_t launch_seed_streams(rwpt_engine *e, int64_t n, int64_t offset);
hipError_t launch_advance(rwpt_engine *e, double t_stop);
// the walk's parameters for an advance from the engine time to t_stop
WalkParams walk_params(const rwpt_engine *e, double t_stop);
// d(record)/d(parameter) of every cell, fields 0..F_CONN-1, for the
// adjoint: 0 aLH, 1 aTH, 2 Dm (each cell's own value), 3 a common scale s
// of all porosities (at s = 1); out[(cell * F_CONN + field) * 4 + param]
hipError_t build_record_tangents(rwpt_engine *e, double *out);
hipError_t launch_global_xyz(rwpt_engine *e, double *gx, double *gy,
                             double *gz);
// unstructured grids: unit records (debug: USG_DEBUG_FIELDS per unit,
// optional) and global coordinates
hipError_t build_usg_tables(rwpt_engine *e, double *debug_out);
hipError_t launch_usg_global_xyz(rwpt_engine *e, double *gx, double *gy, double *gz);
hipError_t launch_kde(rwpt_engine *e, const rwpt_kde_options *opt,
                      const double *mass_dev, double *density_dev,
                      double *hist_dev, double *stats_host);

// rwpt_kde_device: the device address of the file-scope scratch array on
// `device` (looked up once per device, with that device current), and the
// one-pass reconstruction that uses it instead of allocating
hipError_t kde_device_scratch_init(int device);
hipError_t launch_kde_device(rwpt_engine *e, const rwpt_kde_options *opt,
                             double *density_dev, double *hist_dev, double *stats_host);

hipError_t ensure_sort_storage(rwpt_engine *e);
// walk launches by arithmetic type: FP64 (rwpt_walk.hip) and FP32-local
// (rwpt_walk32.hip, a translation unit of its own built with contraction)
hipError_t walk_launch64(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                         const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                         int32_t *nv);
hipError_t walk_launch32(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                         const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                         int32_t *nv);
hipError_t walk_prepare32(rwpt_engine *e, const WalkParams &W);
hipError_t walk_terminate32(rwpt_engine *e, const int32_t *work, int64_t cnt);
hipError_t launch_gpkde_opt(rwpt_engine *e, const rwpt_kde_options *opt,
                            const double *counts_dev, double hinit,
                            double nPoints, const double *final_weights_dev,
                            double *density_dev, int *loops_run);
// FP32-local state: float record table from the FP64 one, positions to and
// from the double arrays
hipError_t build_cell_tables32(rwpt_engine *e);
hipError_t launch_positions_to32(rwpt_engine *e);
hipError_t launch_positions_from32(rwpt_engine *e);

// host helper: GF(2) matrices of the xoroshiro128+ jump, J^(2^k)
void jump_power_matrices(int kmax, std::vector<uint64_t> &out);

// gpkde kernel database (InitializeKernelDatabaseFlat,
// GridProjectedKDE.F90:3138-3526, log form): the h/delta grid and, for
// each lookup, the bandwidths of the database entry it returns.  Every
// entry is a brute-force kernel at grid bandwidths, so the device builds
// that kernel directly instead of storing the database.
struct KernelDb {
  int n = 0;
  double min = 0.0, delta = 0.0;
  std::vector<double> hod;  // h/delta of entries 1..n
  double bin[3] = {0, 0, 0};
  int active[3] = {0, 0, 0};
  int ndim = 0, dims[3] = {0, 0, 0};
  void build(double min_hod, double delta_hod, double max_hod, const double binsz[3],
             const int nbins[3]);
  int index(double h, int d) const;
  // density kernel (prSetKernelFromDatabase :5965-6003): bandwidths and
  // the dimension left one bin wide by the x-y transpose of an x-z or y-z
  // lookup (-1 none)
  int w(const double h[3], double out[3]) const;
  // sigma kernel (prSetKernelSigmaFromDatabase :6006-6043)
  void sigma(double s, double out[3]) const;
  // second-derivative kernel along der (:6081-6124, :6184-6221)
  void sd(double cb, int der, double out[3]) const;
};

}  // namespace rwpt
