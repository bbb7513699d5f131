// C ABI of librwpt (include/rwpt.h): engine lifetime, host<->device copies,
// launches and error reporting.  Inputs are validated on the host so that
// every kernel only sees indices inside its buffers.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "rwpt_internal.h"

using namespace rwpt;

namespace {

int fail(rwpt_engine *e, int code, const char *msg) {
  if (e) e->err = msg;
  return code;
}

int hip_fail(rwpt_engine *e, hipError_t err, const char *where) {
  if (e) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(err));
    e->err = buf;
  }
  return err == hipErrorOutOfMemory ? RWPT_E_NOMEM : RWPT_E_DEVICE;
}

template <class T>
hipError_t dev_alloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <class T>
void dev_free(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <class T>
hipError_t upload(rwpt_engine *e, T **dst, const T *src, size_t count) {
  dev_free(*dst);
  if (!src) return hipSuccess;
  hipError_t err = dev_alloc(dst, count);
  if (err != hipSuccess) return err;
  return hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice,
                        e->stream);
}

void free_particles(Particles &P) {
  dev_free(P.cell);
  dev_free(P.status);
  dev_free(P.lx);
  dev_free(P.ly);
  dev_free(P.lz);
  dev_free(P.t);
  dev_free(P.rng);
  dev_free(P.substeps);
  dev_free(P.aux);
  dev_free(P.ix);
  dev_free(P.iy);
  dev_free(P.iz);
  dev_free(P.t0);
  dev_free(P.flx);
  dev_free(P.fly);
  dev_free(P.flz);
  dev_free(P.fix);
  dev_free(P.fiy);
  dev_free(P.fiz);
  for (int k = 0; k < 2; ++k) {
    dev_free(P.keys[k]);
    dev_free(P.vals[k]);
  }
  if (P.sort_tmp) (void)hipFree(P.sort_tmp);
  P.sort_tmp = nullptr;
  P.sort_tmp_bytes = 0;
  P.n = P.cap = 0;
}

void free_usg_particles(UsgData &U) {
  dev_free(U.unit); dev_free(U.pcell);
  dev_free(U.ex); dev_free(U.ey); dev_free(U.ez); dev_free(U.et);
  dev_free(U.px); dev_free(U.py); dev_free(U.pz);
  U.pcap = 0;
}

void free_usg(UsgData *&U) {
  if (!U) return;
  free_usg_particles(*U);
  dev_free(U->topo); dev_free(U->cell_unit); dev_free(U->side); dev_free(U->slots);
  dev_free(U->layer); dev_free(U->cell_type); dev_free(U->off);
  dev_free(U->cx); dev_free(U->cy); dev_free(U->dx); dev_free(U->dy);
  dev_free(U->flowja); dev_free(U->storage); dev_free(U->casm);
  delete U;
  U = nullptr;
}

int ensure_tables(rwpt_engine *e) {
  if (!e->have_flow) return fail(e, RWPT_E_STATE, "flow not set");
  if (!e->have_disp) return fail(e, RWPT_E_STATE, "dispersion not set");
  hipError_t err = hipSuccess;
  if (e->tables_dirty) {
    err = e->usg ? build_usg_tables(e, nullptr) : build_cell_tables(e, nullptr);
    if (err != hipSuccess) return hip_fail(e, err, "cell tables");
    e->tables_dirty = false;
    e->tables32_dirty = true;
  }
  if (e->opt.precision != 0 && e->tables32_dirty) {
    if (!e->prob.cells32)
      err = dev_alloc(&e->prob.cells32, e->prob.ncell * CELL_FLOATS);
    if (err == hipSuccess) err = build_cell_tables32(e);
    if (err != hipSuccess) return hip_fail(e, err, "FP32 cell tables");
    e->tables32_dirty = false;
  }
  return RWPT_OK;
}

}  // namespace

extern "C" {

void rwpt_default_options(rwpt_options *opt) {
  if (!opt) return;
  memset(opt, 0, sizeof *opt);
  opt->advection_kind = 2;
  opt->time_step_kind = 4;
  opt->time_step_params[0] = 1.0;
  opt->time_step_params[1] = 0.0;
  opt->dimension_mask[0] = opt->dimension_mask[1] = opt->dimension_mask[2] = 1;
  opt->steady_state = 1;
  opt->precision = 0;
  opt->base_seed[0] = 0x1DADBEEFBAADD0D0ull;
  opt->base_seed[1] = 0x5BADD0D0DEADBEEFull;
}

int rwpt_create(const rwpt_grid *grid, const rwpt_options *opt,
                rwpt_engine **out) {
  return rwpt_create_ensemble(grid, opt, 1, out);
}

int rwpt_create_ensemble(const rwpt_grid *grid, const rwpt_options *opt,
                         int32_t members, rwpt_engine **out) {
  if (!out) return RWPT_E_ARG;
  *out = nullptr;
  if (!grid || !opt || members < 1) return RWPT_E_ARG;
  if (grid->ncol <= 0 || grid->nrow <= 0 || grid->nlay <= 0 || !grid->delr ||
      !grid->delc || !grid->top || !grid->bot)
    return RWPT_E_ARG;
  const int64_t ncell_member = (int64_t)grid->ncol * grid->nrow * grid->nlay;
  const int64_t ncell = ncell_member * members;
  if (ncell > 0x7fffffffLL || (int64_t)grid->nlay * members > 0x7fffffffLL)
    return RWPT_E_ARG;
  if (opt->advection_kind != 1 && opt->advection_kind != 2) return RWPT_E_ARG;
  if (opt->time_step_kind < 1 || opt->time_step_kind > 4) return RWPT_E_ARG;
  int nd = 0;
  for (int d = 0; d < 3; ++d) {
    if (opt->dimension_mask[d] != 0 && opt->dimension_mask[d] != 1)
      return RWPT_E_ARG;
    nd += opt->dimension_mask[d];
  }
  if (nd == 0) return RWPT_E_ARG;
  if (opt->precision < 0 || opt->precision > 2) return RWPT_E_ARG;
  for (int j = 0; j < grid->ncol; ++j)
    if (!(grid->delr[j] > 0.0)) return RWPT_E_ARG;
  for (int i = 0; i < grid->nrow; ++i)
    if (!(grid->delc[i] > 0.0)) return RWPT_E_ARG;

  rwpt_engine *e = new rwpt_engine();
  e->opt = *opt;
  hipError_t err = hipGetDevice(&e->device);
  // rwpt_kde_device's scratch array on this device
  if (err == hipSuccess) err = kde_device_scratch_init(e->device);
  if (err == hipSuccess) err = hipStreamCreate(&e->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(e, err, "create");
    fprintf(stderr, "rwpt_create: %s\n", e->err.c_str());
    delete e;
    return rc;
  }
  DeviceProblem &P = e->prob;
  P.ncol = grid->ncol;
  P.nrow = grid->nrow;
  P.nlay = grid->nlay * members;
  P.nlay_member = grid->nlay;
  P.members = members;
  P.ncell = ncell;
  P.ncell_member = ncell_member;
  err = upload(e, &P.delr, grid->delr, P.ncol);
  if (err == hipSuccess) err = upload(e, &P.delc, grid->delc, P.nrow);
  {
    // every member has the geometry of the grid given
    std::vector<double> tb(ncell);
    for (int64_t c = 0; c < ncell; ++c) tb[c] = grid->top[c % ncell_member];
    if (err == hipSuccess) err = upload(e, &P.top, tb.data(), ncell);
    for (int64_t c = 0; c < ncell; ++c) tb[c] = grid->bot[c % ncell_member];
    if (err == hipSuccess) err = upload(e, &P.bot, tb.data(), ncell);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  }
  if (err == hipSuccess) err = dev_alloc(&P.uniform_member, members);
  if (err == hipSuccess) err = dev_alloc(&P.q6, 6 * ncell);
  if (err == hipSuccess) err = dev_alloc(&P.dzrw, ncell);
  if (err == hipSuccess) err = dev_alloc(&P.cells, ncell * CELL_DOUBLES);
  if (err == hipSuccess) err = dev_alloc(&e->counters, 4);
  {
    // face coordinates X(1..ncol+1), Y(1..nrow+1) of RectangularGridDis.f90
    std::vector<double> X(P.ncol + 1), Y(P.nrow + 1);
    X[0] = 0.0;
    for (int j = 0; j < P.ncol; ++j) X[j + 1] = X[j] + grid->delr[j];
    Y[0] = 0.0;
    for (int i = 0; i < P.nrow; ++i) Y[0] = Y[0] + grid->delc[i];
    for (int i = 0; i < P.nrow; ++i) Y[i + 1] = Y[i] - grid->delc[i];
    if (err == hipSuccess) err = upload(e, &e->xface, X.data(), X.size());
    if (err == hipSuccess) err = upload(e, &e->yface, Y.data(), Y.size());
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  }
  std::vector<uint64_t> mats;
  jump_power_matrices(64, mats);
  if (err == hipSuccess) err = upload(e, &e->jump_mats, mats.data(), mats.size());
  for (int k = 0; k < 3 && err == hipSuccess; ++k) {
    err = hipEventCreate(&e->timers[k].start);
    if (err == hipSuccess) err = hipEventCreate(&e->timers[k].stop);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(e, err, "create");
    rwpt_destroy(e);
    return rc;
  }
  *out = e;
  return RWPT_OK;
}

int rwpt_destroy(rwpt_engine *e) {
  if (!e) return RWPT_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  DeviceProblem &P = e->prob;
  dev_free(P.delr); dev_free(P.delc); dev_free(P.top); dev_free(P.bot);
  dev_free(P.frf); dev_free(P.fff); dev_free(P.flf); dev_free(P.bflow);
  dev_free(P.porosity); dev_free(P.retardation); dev_free(P.ibound);
  dev_free(P.icbound); dev_free(P.head); dev_free(P.layer_type);
  dev_free(P.zone); dev_free(P.source); dev_free(P.sink); dev_free(P.disp);
  dev_free(P.sat_top);
  dev_free(P.q6); dev_free(P.dzrw); dev_free(P.cells); dev_free(P.cells32);
  dev_free(P.uniform_member);
  free_particles(e->part);
  free_usg(e->usg);
  dev_free(e->jump_mats);
  dev_free(e->counters);
  dev_free(e->kde_scratch);
  dev_free(e->kde_weights);
  dev_free(e->xface);
  dev_free(e->yface);
  for (void *m : e->gk_mem)
    if (m) (void)hipFree(m);
  for (auto &T : e->timers) {
    if (T.start) (void)hipEventDestroy(T.start);
    if (T.stop) (void)hipEventDestroy(T.stop);
  }
  for (auto ev : e->walk_events) (void)hipEventDestroy(ev);
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return RWPT_OK;
}

const char *rwpt_last_error(const rwpt_engine *e) {
  return e ? e->err.c_str() : "null engine";
}

int rwpt_set_flow(rwpt_engine *e, const double *frf, const double *fff,
                  const double *flf, const double *bflow,
                  const double *porosity, const double *retardation,
                  const int32_t *ibound, const int32_t *icbound,
                  int32_t default_icbound) {
  if (!e) return RWPT_E_ARG;
  if (e->usg) return fail(e, RWPT_E_STATE, "unstructured grid: use rwpt_set_flow_usg");
  if (!frf || !fff || !flf || !porosity)
    return fail(e, RWPT_E_ARG, "frf, fff, flf and porosity are required");
  DeviceProblem &P = e->prob;
  const int64_t n = P.ncell;
  for (int64_t c = 0; c < n; ++c) {
    if (!(porosity[c] > 0.0)) return fail(e, RWPT_E_ARG, "porosity must be > 0");
    if (retardation && !(retardation[c] > 0.0))
      return fail(e, RWPT_E_ARG, "retardation must be > 0");
  }
  // isUniformPorosity (ModpathBasicData.f90:208), per ensemble member: a
  // member whose cells share one porosity takes its corner porosities from
  // the cell (TrackSubCell.F90:4227-4283)
  std::vector<int32_t> uni(P.members, 1);
  for (int m = 0; m < P.members; ++m) {
    const double *pm = porosity + (int64_t)m * P.ncell_member;
    for (int64_t c = 1; c < P.ncell_member; ++c)
      if (pm[c] != pm[0]) {
        uni[m] = 0;
        break;
      }
  }
  hipError_t err = hipMemcpyAsync(P.uniform_member, uni.data(), P.members * 4,
                                  hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess) err = upload(e, &P.frf, frf, n);
  if (err == hipSuccess) err = upload(e, &P.fff, fff, n);
  if (err == hipSuccess) err = upload(e, &P.flf, flf, n);
  if (err == hipSuccess) err = upload(e, &P.bflow, bflow, 6 * n);
  if (err == hipSuccess) err = upload(e, &P.porosity, porosity, n);
  if (err == hipSuccess) err = upload(e, &P.retardation, retardation, n);
  if (err == hipSuccess) err = upload(e, &P.ibound, ibound, n);
  if (err == hipSuccess) err = upload(e, &P.icbound, icbound, n);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "set_flow");
  P.default_icbound = default_icbound;
  P.uniform_porosity = 1;
  for (int m = 0; m < P.members; ++m) P.uniform_porosity &= uni[m];
  e->have_flow = true;
  e->tables_dirty = true;
  return RWPT_OK;
}

int rwpt_set_cell_extras(rwpt_engine *e, const double *head,
                         const int32_t *layer_type, const int32_t *zone,
                         const double *source_flow, const double *sink_flow) {
  if (!e) return RWPT_E_ARG;
  DeviceProblem &P = e->prob;
  const int64_t n = P.ncell;
  hipError_t err = upload(e, &P.head, head, n);
  if (err == hipSuccess) err = upload(e, &P.layer_type, layer_type, n);
  if (err == hipSuccess) err = upload(e, &P.zone, zone, n);
  if (err == hipSuccess) err = upload(e, &P.source, source_flow, n);
  if (err == hipSuccess) err = upload(e, &P.sink, sink_flow, n);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "set_cell_extras");
  e->tables_dirty = true;
  return RWPT_OK;
}

int rwpt_set_saturated_top(rwpt_engine *e, const double *saturated_top) {
  if (!e) return RWPT_E_ARG;
  DeviceProblem &P = e->prob;
  hipError_t err = upload(e, &P.sat_top, saturated_top, P.ncell);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "set_saturated_top");
  return RWPT_OK;
}

int rwpt_set_dispersion(rwpt_engine *e, int32_t model, int32_t uniform,
                        const double *aL, const double *aT, const double *aLV,
                        const double *aTV, const double *dmeff) {
  if (!e) return RWPT_E_ARG;
  if (model != 0 && model != 1)
    return fail(e, RWPT_E_UNSUPPORTED, "dispersion model must be 0 or 1");
  if (!aL || !aT || !dmeff) return fail(e, RWPT_E_ARG, "aL, aT, dmeff required");
  if (model == 1 && (!aLV || !aTV))
    return fail(e, RWPT_E_ARG, "axisymmetric model needs aLV and aTV");
  if (uniform < 0 || uniform > 2)
    return fail(e, RWPT_E_ARG, "uniform must be 0 (per cell), 1 (one value) or 2 "
                               "(per ensemble member)");
  DeviceProblem &P = e->prob;
  // per member values become per cell values of that member's cells
  const int64_t nd = uniform == 1 ? 1 : P.ncell;
  std::vector<double> h(5 * nd, 0.0);
  for (int64_t i = 0; i < nd; ++i) {
    const int64_t s = uniform == 2 ? i / P.ncell_member : i;
    h[i] = aL[s];
    h[nd + i] = aT[s];
    h[2 * nd + i] = model == 1 ? aLV[s] : 0.0;
    h[3 * nd + i] = model == 1 ? aTV[s] : 0.0;
    h[4 * nd + i] = dmeff[s];
  }
  hipError_t err = upload(e, &P.disp, h.data(), h.size());
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "set_dispersion");
  P.ndisp = nd;
  P.disp_model = model;
  e->have_disp = true;
  e->tables_dirty = true;
  return RWPT_OK;
}

int rwpt_load_particles(rwpt_engine *e, int64_t n, const int32_t *cell,
                        const double *lx, const double *ly, const double *lz,
                        const double *t, const int32_t *status,
                        const uint64_t *rng, int64_t stream_offset) {
  if (!e) return RWPT_E_ARG;
  if (n < 0 || (n > 0 && (!cell || !lx || !ly || !lz || !t || !status)))
    return fail(e, RWPT_E_ARG, "particle arrays required");
  if (stream_offset < 0) return fail(e, RWPT_E_ARG, "stream_offset < 0");
  const int64_t ncell = e->prob.ncell;
  for (int64_t p = 0; p < n; ++p) {
    if (cell[p] < 1 || cell[p] > ncell)
      return fail(e, RWPT_E_ARG, "particle cell outside the grid");
    if (status[p] < 0 || status[p] > 9)
      return fail(e, RWPT_E_ARG, "particle status outside 0..9");
  }
  Particles &P = e->part;
  hipError_t err = hipSuccess;
  if (n > P.cap) {
    free_particles(P);
    err = dev_alloc(&P.cell, n);
    if (err == hipSuccess) err = dev_alloc(&P.status, n);
    if (err == hipSuccess) err = dev_alloc(&P.lx, n);
    if (err == hipSuccess) err = dev_alloc(&P.ly, n);
    if (err == hipSuccess) err = dev_alloc(&P.lz, n);
    if (err == hipSuccess) err = dev_alloc(&P.t, n);
    if (err == hipSuccess) err = dev_alloc(&P.rng, 2 * n);
    if (err == hipSuccess) err = dev_alloc(&P.substeps, n);
    if (err == hipSuccess) err = dev_alloc(&P.aux, n);
    if (err == hipSuccess) err = dev_alloc(&P.ix, n);
    if (err == hipSuccess) err = dev_alloc(&P.iy, n);
    if (err == hipSuccess) err = dev_alloc(&P.iz, n);
    if (err == hipSuccess) err = dev_alloc(&P.t0, n);
    if (e->opt.precision != 0) {
      if (err == hipSuccess) err = dev_alloc(&P.flx, n);
      if (err == hipSuccess) err = dev_alloc(&P.fly, n);
      if (err == hipSuccess) err = dev_alloc(&P.flz, n);
      if (err == hipSuccess) err = dev_alloc(&P.fix, n);
      if (err == hipSuccess) err = dev_alloc(&P.fiy, n);
      if (err == hipSuccess) err = dev_alloc(&P.fiz, n);
    }
    for (int k = 0; k < 2 && err == hipSuccess; ++k) {
      err = dev_alloc(&P.keys[k], n);
      if (err == hipSuccess) err = dev_alloc(&P.vals[k], n);
    }
    if (e->usg && err == hipSuccess) {
      UsgData &U = *e->usg;
      free_usg_particles(U);
      err = dev_alloc(&U.unit, n);
      if (err == hipSuccess) err = dev_alloc(&U.pcell, n);
      for (double **a : {&U.ex, &U.ey, &U.ez, &U.et, &U.px, &U.py, &U.pz})
        if (err == hipSuccess) err = dev_alloc(a, n);
      U.pcap = n;
    }
    P.cap = n;
    if (err == hipSuccess) err = ensure_sort_storage(e);
    if (err != hipSuccess) {
      free_particles(P);
      return hip_fail(e, err, "load_particles alloc");
    }
  }
  P.n = n;
  if (n == 0) return RWPT_OK;
  hipStream_t s = e->stream;
  err = hipMemcpyAsync(P.cell, cell, n * 4, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemcpyAsync(P.status, status, n * 4, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemcpyAsync(P.lx, lx, n * 8, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemcpyAsync(P.ly, ly, n * 8, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemcpyAsync(P.lz, lz, n * 8, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemcpyAsync(P.t, t, n * 8, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipMemsetAsync(P.substeps, 0, n * 8, s);
  if (err == hipSuccess) {
    if (rng) err = hipMemcpyAsync(P.rng, rng, n * 16, hipMemcpyHostToDevice, s);
    else err = launch_seed_streams(e, n, stream_offset);
  }
  if (err == hipSuccess && e->opt.precision != 0) err = launch_positions_to32(e);
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) return hip_fail(e, err, "load_particles");
  return RWPT_OK;
}

int rwpt_get_particles(rwpt_engine *e, int32_t *cell, double *lx, double *ly,
                       double *lz, double *t, int32_t *status, uint64_t *rng) {
  if (!e) return RWPT_E_ARG;
  Particles &P = e->part;
  const int64_t n = P.n;
  if (n == 0) return RWPT_OK;
  hipStream_t s = e->stream;
  hipError_t err = hipSuccess;
  if (cell) err = hipMemcpyAsync(cell, P.cell, n * 4, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && lx) err = hipMemcpyAsync(lx, P.lx, n * 8, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && ly) err = hipMemcpyAsync(ly, P.ly, n * 8, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && lz) err = hipMemcpyAsync(lz, P.lz, n * 8, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && t) err = hipMemcpyAsync(t, P.t, n * 8, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && status) err = hipMemcpyAsync(status, P.status, n * 4, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess && rng) err = hipMemcpyAsync(rng, P.rng, n * 16, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) return hip_fail(e, err, "get_particles");
  return RWPT_OK;
}

int rwpt_get_substeps(rwpt_engine *e, int64_t *substeps) {
  if (!e || !substeps) return RWPT_E_ARG;
  Particles &P = e->part;
  if (P.n == 0) return RWPT_OK;
  hipError_t err = hipMemcpy(substeps, P.substeps, P.n * 8, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return hip_fail(e, err, "get_substeps");
  return RWPT_OK;
}

int64_t rwpt_particle_count(const rwpt_engine *e) { return e ? e->part.n : -1; }

int rwpt_advance(rwpt_engine *e, double t_stop, int64_t *n_active,
                 int64_t *n_substeps) {
  if (!e) return RWPT_E_ARG;
  if (!std::isfinite(t_stop)) return fail(e, RWPT_E_ARG, "t_stop not finite");
  int rc = ensure_tables(e);
  if (rc != RWPT_OK) return rc;
  hipError_t err = launch_advance(e, t_stop);
  if (err != hipSuccess) return hip_fail(e, err, "advance launch");
  if (e->opt.precision != 0) {
    err = launch_positions_from32(e);
    if (err != hipSuccess) return hip_fail(e, err, "advance positions");
  }
  unsigned long long h[4] = {0, 0, 0, 0};
  err = hipMemcpyAsync(h, e->counters, sizeof h, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "advance");
  for (int64_t k = 0; k < e->last_timed; ++k) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e->walk_events[2 * k], e->walk_events[2 * k + 1]);
    e->timers[0].total_ms += ms;
    e->timers[0].launches += e->walk_event_launches[k];
  }
  e->t_engine = t_stop;
  if (n_active) *n_active = (int64_t)h[0];
  if (n_substeps) *n_substeps = (int64_t)h[1];
  return RWPT_OK;
}

int rwpt_set_tuning(rwpt_engine *e, int32_t budget, int32_t block,
                    int32_t sort, int32_t min_waves) {
  if (!e) return RWPT_E_ARG;
  // the walk kernel is compiled for at most 256 threads per workgroup
  if (budget < 1 || (block != 64 && block != 128 && block != 256))
    return fail(e, RWPT_E_ARG, "budget >= 1, block one of 64, 128, 256");
  if (min_waves != 1 && min_waves != 3 && min_waves != 4)
    return fail(e, RWPT_E_ARG, "min_waves one of 1, 3, 4");
  e->tune.min_waves = min_waves;
  e->tune.budget = budget;
  e->tune.block = block;
  e->tune.sort = sort ? 1 : 0;
  return RWPT_OK;
}

int rwpt_set_graph(rwpt_engine *e, int32_t launches, int64_t max_list) {
  if (!e) return RWPT_E_ARG;
  if (launches < 0 || launches > 256 || (launches & 1))
    return fail(e, RWPT_E_ARG, "graph launches must be 0 or an even number up to 256");
  if (launches != e->graph_launches && e->graph_exec) {
    (void)hipGraphExecDestroy(e->graph_exec);
    e->graph_exec = nullptr;
  }
  e->graph_launches = launches;
  e->graph_max_list = max_list;
  return RWPT_OK;
}

namespace {

// Device copy of every particle array an advance writes, so that trial
// advances can start from the same state.
struct StateCopy {
  std::vector<std::pair<void *, size_t>> arrays;  // (device array, bytes)
  char *buf = nullptr;
  size_t bytes = 0;
  hipError_t save(rwpt_engine *e) {
    Particles &S = e->part;
    const size_t n = (size_t)S.n;
    arrays = {{S.cell, 4 * n},  {S.status, 4 * n}, {S.lx, 8 * n},  {S.ly, 8 * n},
              {S.lz, 8 * n},    {S.t, 8 * n},      {S.rng, 16 * n}, {S.substeps, 8 * n},
              {S.aux, 4 * n},   {S.ix, 8 * n},     {S.iy, 8 * n},   {S.iz, 8 * n},
              {S.t0, 8 * n}};
    if (S.flx)
      for (float *f : {S.flx, S.fly, S.flz, S.fix, S.fiy, S.fiz}) arrays.push_back({f, 4 * n});
    if (e->usg) {
      // the unstructured walk's visit state
      UsgData &U = *e->usg;
      arrays.push_back({U.unit, 4 * n});
      arrays.push_back({U.pcell, 4 * n});
      for (double *d : {U.ex, U.ey, U.ez, U.et, U.px, U.py, U.pz}) arrays.push_back({d, 8 * n});
    }
    bytes = 0;
    for (auto &a : arrays)
      if (a.first) bytes += a.second;
    hipError_t err = hipMalloc(&buf, bytes);
    size_t off = 0;
    for (auto &a : arrays) {
      if (err != hipSuccess) break;
      if (!a.first) continue;
      err = hipMemcpyAsync(buf + off, a.first, a.second, hipMemcpyDeviceToDevice, e->stream);
      off += a.second;
    }
    return err;
  }
  hipError_t restore(rwpt_engine *e) {
    size_t off = 0;
    hipError_t err = hipSuccess;
    for (auto &a : arrays) {
      if (err != hipSuccess) break;
      if (!a.first) continue;
      err = hipMemcpyAsync(a.first, buf + off, a.second, hipMemcpyDeviceToDevice, e->stream);
      off += a.second;
    }
    return err;
  }
  ~StateCopy() {
    if (buf) (void)hipFree(buf);
  }
};

}  // namespace

int rwpt_autotune(rwpt_engine *e, double t_probe, int32_t *budget, int32_t *block,
                  int32_t *sort, int32_t *min_waves, double *ms_best,
                  int64_t *trial_substeps) {
  if (!e) return RWPT_E_ARG;
  if (!std::isfinite(t_probe) || t_probe == e->t_engine)
    return fail(e, RWPT_E_ARG, "t_probe must differ from the engine time");
  int rc = ensure_tables(e);
  if (rc != RWPT_OK) return rc;
  if (e->part.n == 0) return fail(e, RWPT_E_STATE, "no particles loaded");
  StateCopy snap;
  hipError_t err = snap.save(e);
  if (err != hipSuccess) return hip_fail(e, err, "autotune snapshot");
  const double t0 = e->t_engine;
  const Tuning keep = e->tune;
  // one trial advance from the snapshot, wall time of the whole launch loop;
  // its walk launches count in the advance kernel's timing like any others
  int64_t steps = 0;
  auto trial = [&](const Tuning &T, double *ms) -> hipError_t {
    hipError_t r = snap.restore(e);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    if (r != hipSuccess) return r;
    e->tune = T;
    e->t_engine = t0;
    const auto a = std::chrono::steady_clock::now();
    r = launch_advance(e, t_probe);
    if (r == hipSuccess && e->opt.precision != 0) r = launch_positions_from32(e);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    const auto b = std::chrono::steady_clock::now();
    *ms = std::chrono::duration<double, std::milli>(b - a).count();
    unsigned long long h[2] = {0, 0};
    if (r == hipSuccess)
      r = hipMemcpy(h, e->counters, sizeof h, hipMemcpyDeviceToHost);
    steps += (int64_t)h[1];
    for (int64_t k = 0; r == hipSuccess && k < e->last_timed; ++k) {
      float km = 0.f;
      (void)hipEventElapsedTime(&km, e->walk_events[2 * k], e->walk_events[2 * k + 1]);
      e->timers[0].total_ms += km;
      e->timers[0].launches += e->walk_event_launches[k];
    }
    return r;
  };
  // each shape is timed twice and keeps its faster run
  auto timed = [&](const Tuning &T, double *ms) -> hipError_t {
    double a = 0.0, b = 0.0;
    hipError_t r = trial(T, &a);
    if (r == hipSuccess) r = trial(T, &b);
    *ms = std::min(a, b);
    return r;
  };
  Tuning best = keep;
  double bestMs = 0.0;
  err = trial(best, &bestMs);  // warm-up (first launches compile and cache)
  if (err == hipSuccess) err = timed(best, &bestMs);
  auto consider = [&](Tuning T) -> hipError_t {
    double ms = 0.0;
    hipError_t r = timed(T, &ms);
    if (r == hipSuccess && ms < bestMs) {
      bestMs = ms;
      best = T;
    }
    return r;
  };
  // staged search: workgroup size and occupancy variant, then the
  // per-launch iteration budget, then the cell sort of the work list
  for (int bl : {64, 128, 256})
    for (int mw : {1, 3, 4}) {
      if (err != hipSuccess) break;
      Tuning T = best;
      T.block = bl;
      T.min_waves = mw;
      err = consider(T);
    }
  for (int bu : {16, 32, 64, 128, 256}) {
    if (err != hipSuccess) break;
    Tuning T = best;
    T.budget = bu;
    err = consider(T);
  }
  for (int so : {0, 1}) {
    if (err != hipSuccess) break;
    Tuning T = best;
    T.sort = so;
    err = consider(T);
  }
  // back to the state before the call, with the tuning found
  if (err == hipSuccess) err = snap.restore(e);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  e->t_engine = t0;
  if (err != hipSuccess) {
    e->tune = keep;
    return hip_fail(e, err, "autotune");
  }
  e->tune = best;
  if (budget) *budget = best.budget;
  if (block) *block = best.block;
  if (sort) *sort = best.sort;
  if (min_waves) *min_waves = best.min_waves;
  if (ms_best) *ms_best = bestMs;
  if (trial_substeps) *trial_substeps = steps;
  return RWPT_OK;
}

int rwpt_last_launches(const rwpt_engine *e) {
  return e ? (int)e->last_launches : -1;
}

int rwpt_set_time(rwpt_engine *e, double t) {
  if (!e) return RWPT_E_ARG;
  e->t_engine = t;
  return RWPT_OK;
}

int rwpt_set_steady_state(rwpt_engine *e, int32_t steady_state) {
  if (!e) return RWPT_E_ARG;
  if (steady_state != 0 && steady_state != 1)
    return fail(e, RWPT_E_ARG, "steady_state must be 0 or 1");
  // read by the walk's entry checks only: the cell tables do not change
  e->opt.steady_state = steady_state;
  return RWPT_OK;
}

int rwpt_global_xyz(rwpt_engine *e, double *gx, double *gy, double *gz) {
  if (!e || !gx || !gy || !gz) return RWPT_E_ARG;
  const int64_t n = e->part.n;
  if (n == 0) return RWPT_OK;
  double *d = nullptr;
  hipError_t err = dev_alloc(&d, 3 * n);
  if (err == hipSuccess) err = launch_global_xyz(e, d, d + n, d + 2 * n);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err == hipSuccess) err = hipMemcpy(gx, d, n * 8, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(gy, d + n, n * 8, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(gz, d + 2 * n, n * 8, hipMemcpyDeviceToHost);
  dev_free(d);
  if (err != hipSuccess) return hip_fail(e, err, "global_xyz");
  return RWPT_OK;
}

int rwpt_get_cell_tables(rwpt_engine *e, double *out) {
  if (!e || !out) return RWPT_E_ARG;
  if (e->usg) return fail(e, RWPT_E_STATE, "unstructured grid: use rwpt_get_usg_tables");
  if (!e->have_flow || !e->have_disp)
    return fail(e, RWPT_E_STATE, "flow and dispersion must be set");
  double *d = nullptr;
  const int64_t n = e->prob.ncell * DEBUG_FIELDS;
  hipError_t err = dev_alloc(&d, n);
  if (err == hipSuccess) err = build_cell_tables(e, d);
  if (err == hipSuccess) err = hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  dev_free(d);
  if (err != hipSuccess) return hip_fail(e, err, "cell tables");
  e->tables_dirty = false;
  e->tables32_dirty = true;
  return RWPT_OK;
}

int rwpt_create_usg(const rwpt_usg_grid *grid, const rwpt_options *opt,
                    rwpt_engine **out) {
  if (!out) return RWPT_E_ARG;
  *out = nullptr;
  if (!grid || !opt) return RWPT_E_ARG;
  const rwpt_usg_grid &g = *grid;
  if (g.ncell <= 0 || g.nlay <= 0 || g.nja < g.ncell || g.nja > 0x7fffffffLL || !g.ia ||
      !g.ja || !g.layer || !g.cell_x || !g.cell_y || !g.del_x || !g.del_y || !g.top ||
      !g.bot)
    return RWPT_E_ARG;
  if (opt->precision != 0) return RWPT_E_UNSUPPORTED;
  if (opt->advection_kind != 1 && opt->advection_kind != 2) return RWPT_E_ARG;
  if (opt->time_step_kind < 1 || opt->time_step_kind > 4) return RWPT_E_ARG;
  int nd = 0;
  for (int d = 0; d < 3; ++d) {
    if (opt->dimension_mask[d] != 0 && opt->dimension_mask[d] != 1) return RWPT_E_ARG;
    nd += opt->dimension_mask[d];
  }
  if (nd == 0) return RWPT_E_ARG;
  // the connectivity must be a CSR of the cells, each row led by its cell
  if (g.ia[0] != 1 || g.ia[g.ncell] != g.nja + 1) return RWPT_E_ARG;
  for (int c = 1; c <= g.ncell; ++c) {
    const int32_t a = g.ia[c - 1], b = g.ia[c];
    // a removed cell (IDOMAIN < 1) has an empty row
    if (b < a || (b > a && g.ja[a - 1] != c)) return RWPT_E_ARG;
    if (g.layer[c - 1] < 1 || g.layer[c - 1] > g.nlay) return RWPT_E_ARG;
    if (!(g.del_x[c - 1] > 0.0) || !(g.del_y[c - 1] > 0.0)) return RWPT_E_ARG;
    for (int32_t q = a; q < b - 1; ++q)
      if (g.ja[q] < 1 || g.ja[q] > g.ncell) return RWPT_E_ARG;
  }
  rwpt_engine *e = new rwpt_engine();
  e->opt = *opt;
  UsgHost H;
  const std::string terr = usg_build_topology(g, H);
  if (!terr.empty()) {
    fprintf(stderr, "rwpt_create_usg: %s\n", terr.c_str());
    delete e;
    return RWPT_E_UNSUPPORTED;
  }
  hipError_t err = hipGetDevice(&e->device);
  if (err == hipSuccess) err = kde_device_scratch_init(e->device);
  if (err == hipSuccess) err = hipStreamCreate(&e->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(e, err, "create_usg");
    fprintf(stderr, "rwpt_create_usg: %s\n", e->err.c_str());
    delete e;
    return rc;
  }
  e->usg = new UsgData();
  UsgData &U = *e->usg;
  U.ncell = g.ncell;
  U.nunit = H.nunit;
  U.nlay = g.nlay;
  U.nja = g.nja;
  DeviceProblem &P = e->prob;
  P.ncol = 0;
  P.nrow = 0;
  P.nlay = g.nlay;
  P.nlay_member = g.nlay;
  P.members = 1;
  P.ncell = g.ncell;
  P.ncell_member = g.ncell;
  const int64_t n = g.ncell;
  err = upload(e, &P.top, g.top, n);
  if (err == hipSuccess) err = upload(e, &P.bot, g.bot, n);
  if (err == hipSuccess) err = upload(e, &U.cx, g.cell_x, n);
  if (err == hipSuccess) err = upload(e, &U.cy, g.cell_y, n);
  if (err == hipSuccess) err = upload(e, &U.dx, g.del_x, n);
  if (err == hipSuccess) err = upload(e, &U.dy, g.del_y, n);
  if (err == hipSuccess) err = upload(e, &U.layer, g.layer, n);
  if (err == hipSuccess) err = upload(e, &U.cell_type, g.cell_type, n);
  if (err == hipSuccess) err = upload(e, &U.topo, H.topo.data(), H.topo.size());
  if (err == hipSuccess) err = upload(e, &U.cell_unit, H.cell_unit.data(), H.cell_unit.size());
  if (err == hipSuccess) err = upload(e, &U.side, H.side.data(), H.side.size());
  if (err == hipSuccess) err = upload(e, &U.slots, H.slots.data(), H.slots.size());
  if (err == hipSuccess) err = upload(e, &U.off, H.off.data(), H.off.size());
  if (err == hipSuccess) err = dev_alloc(&U.casm, 28 * (size_t)n);
  if (err == hipSuccess) err = dev_alloc(&P.uniform_member, 1);
  if (err == hipSuccess) err = dev_alloc(&P.cells, (int64_t)U.nunit * CELL_DOUBLES);
  if (err == hipSuccess) err = dev_alloc(&e->counters, 4);
  std::vector<uint64_t> mats;
  jump_power_matrices(64, mats);
  if (err == hipSuccess) err = upload(e, &e->jump_mats, mats.data(), mats.size());
  for (int k = 0; k < 3 && err == hipSuccess; ++k) {
    err = hipEventCreate(&e->timers[k].start);
    if (err == hipSuccess) err = hipEventCreate(&e->timers[k].stop);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    int rc = hip_fail(e, err, "create_usg");
    rwpt_destroy(e);
    return rc;
  }
  *out = e;
  return RWPT_OK;
}

int rwpt_set_flow_usg(rwpt_engine *e, const double *flowja, const double *bflow,
                      const double *storage, const double *porosity,
                      const double *retardation, const int32_t *ibound,
                      const int32_t *icbound, int32_t default_icbound) {
  if (!e) return RWPT_E_ARG;
  if (!e->usg) return fail(e, RWPT_E_STATE, "not an unstructured grid: use rwpt_set_flow");
  if (!flowja || !porosity) return fail(e, RWPT_E_ARG, "flowja and porosity are required");
  UsgData &U = *e->usg;
  DeviceProblem &P = e->prob;
  const int64_t n = U.ncell;
  int32_t uni = 1;
  for (int64_t c = 0; c < n; ++c) {
    if (!(porosity[c] > 0.0)) return fail(e, RWPT_E_ARG, "porosity must be > 0");
    if (retardation && !(retardation[c] > 0.0))
      return fail(e, RWPT_E_ARG, "retardation must be > 0");
    if (porosity[c] != porosity[0]) uni = 0;
  }
  hipError_t err = hipMemcpyAsync(P.uniform_member, &uni, 4, hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess) err = upload(e, &U.flowja, flowja, U.nja);
  if (err == hipSuccess) err = upload(e, &P.bflow, bflow, 6 * n);
  if (err == hipSuccess) err = upload(e, &U.storage, storage, n);
  if (err == hipSuccess) err = upload(e, &P.porosity, porosity, n);
  if (err == hipSuccess) err = upload(e, &P.retardation, retardation, n);
  if (err == hipSuccess) err = upload(e, &P.ibound, ibound, n);
  if (err == hipSuccess) err = upload(e, &P.icbound, icbound, n);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return hip_fail(e, err, "set_flow_usg");
  P.default_icbound = default_icbound;
  P.uniform_porosity = uni;
  e->have_flow = true;
  e->tables_dirty = true;
  return RWPT_OK;
}

int64_t rwpt_usg_unit_count(const rwpt_engine *e) {
  return (e && e->usg) ? e->usg->nunit : -1;
}

int rwpt_get_usg_tables(rwpt_engine *e, double *out) {
  if (!e || !out) return RWPT_E_ARG;
  if (!e->usg) return fail(e, RWPT_E_STATE, "not an unstructured grid");
  if (!e->have_flow || !e->have_disp)
    return fail(e, RWPT_E_STATE, "flow and dispersion must be set");
  double *d = nullptr;
  const int64_t n = (int64_t)e->usg->nunit * USG_DEBUG_FIELDS;
  hipError_t err = dev_alloc(&d, n);
  if (err == hipSuccess) err = build_usg_tables(e, d);
  if (err == hipSuccess) err = hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  dev_free(d);
  if (err != hipSuccess) return hip_fail(e, err, "usg tables");
  e->tables_dirty = false;
  return RWPT_OK;
}

int rwpt_kernel_time(rwpt_engine *e, int32_t which, double *ms,
                     int64_t *launches) {
  if (!e || which < 0 || which > 2) return RWPT_E_ARG;
  if (ms) *ms = e->timers[which].total_ms;
  if (launches) *launches = e->timers[which].launches;
  return RWPT_OK;
}

int rwpt_reset_kernel_time(rwpt_engine *e) {
  if (!e) return RWPT_E_ARG;
  for (auto &T : e->timers) {
    T.total_ms = 0.0;
    T.launches = 0;
  }
  return RWPT_OK;
}

namespace {
// rwpt_kde's option checks; the bin count goes to *ncells_out
int check_kde_options(rwpt_engine *e, const rwpt_kde_options *opt, int64_t *ncells_out) {
  if (opt->mode != 0 && opt->mode != 1)
    return fail(e, RWPT_E_ARG, "kde mode must be 0 or 1");
  if (opt->n_opt_loops < 0) return fail(e, RWPT_E_ARG, "n_opt_loops < 0");
  if (opt->initial_smoothing != 0 && opt->initial_smoothing != 1)
    return fail(e, RWPT_E_ARG, "initial_smoothing must be 0 or 1");
  if (opt->initial_smoothing == 1 && !(opt->initial_smoothing_factor > 0.0))
    return fail(e, RWPT_E_ARG, "initial_smoothing 1 needs a factor > 0");
  if (opt->bound_kernel_size < 0 || opt->bound_kernel_size > 2)
    return fail(e, RWPT_E_ARG, "bound_kernel_size must be 0, 1 or 2");
  // Initialize stops on these (GridProjectedKDE.F90:693-718)
  if (opt->bound_kernel_size == 1 &&
      !(opt->min_h_over_delta > 0.0 && opt->max_h_over_delta > opt->min_h_over_delta))
    return fail(e, RWPT_E_ARG, "bound_kernel_size 1 needs 0 < min_h_over_delta < "
                               "max_h_over_delta");
  if (opt->effective_weight_format < 0 || opt->effective_weight_format > 3)
    return fail(e, RWPT_E_ARG, "effective_weight_format must be 0..3");
  if (opt->initial != 0 && opt->initial != 1)
    return fail(e, RWPT_E_ARG, "initial must be 0 or 1");
  // InitializeKernelDatabaseFlat's parameter checks (GridProjectedKDE.F90:693-740)
  if (opt->kernel_database &&
      !(opt->min_h_over_delta > 0.0 && opt->max_h_over_delta > opt->min_h_over_delta &&
        opt->delta_h_over_delta > 0.0 && opt->delta_h_over_delta < opt->max_h_over_delta))
    return fail(e, RWPT_E_ARG, "kernel_database needs 0 < min_h_over_delta < "
                               "max_h_over_delta and 0 < delta_h_over_delta < "
                               "max_h_over_delta");
  int64_t ncells = 1;
  for (int d = 0; d < 3; ++d) {
    if (opt->nbins[d] < 1) return fail(e, RWPT_E_ARG, "nbins must be >= 1");
    // a zero bin size compresses a single-bin dimension (Histogram.F90:121)
    if (!(opt->bin_size[d] > 0.0) && !(opt->bin_size[d] == 0.0 && opt->nbins[d] == 1))
      return fail(e, RWPT_E_ARG, "bin_size must be > 0 (or 0 for a single bin)");
    ncells *= opt->nbins[d];
  }
  if (ncells >= 0x7fffffffLL) return fail(e, RWPT_E_ARG, "kde grid too large");
  if (!e->have_flow) return fail(e, RWPT_E_STATE, "flow not set");
  *ncells_out = ncells;
  return RWPT_OK;
}

void add_kde_times(rwpt_engine *e) {
  for (int k = 1; k <= 2; ++k) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e->timers[k].start, e->timers[k].stop);
    e->timers[k].total_ms += ms;
    e->timers[k].launches += 1;
  }
}
}  // namespace

int rwpt_kde(rwpt_engine *e, const rwpt_kde_options *opt, const double *mass,
             double *density, double *histogram) {
  if (!e || !opt) return RWPT_E_ARG;
  int64_t ncells = 0;
  int rc = check_kde_options(e, opt, &ncells);
  if (rc != RWPT_OK) return rc;
  const int64_t n = e->part.n;
  double *dens = nullptr, *hist = nullptr, *m = nullptr;
  hipError_t err = dev_alloc(&dens, ncells);
  if (err == hipSuccess) err = dev_alloc(&hist, ncells);
  if (err == hipSuccess && mass && n > 0) {
    err = dev_alloc(&m, n);
    if (err == hipSuccess)
      err = hipMemcpyAsync(m, mass, n * 8, hipMemcpyHostToDevice, e->stream);
  }
  if (err == hipSuccess) err = launch_kde(e, opt, m, dens, hist, e->kde_last);
  if (err == hipErrorNotSupported) {
    dev_free(dens);
    dev_free(hist);
    dev_free(m);
    return fail(e, RWPT_E_UNSUPPORTED,
                "bandwidth optimisation needs a 2-D or 3-D grid and kernels narrower "
                "than 2048 bins");
  }
  if (err == hipSuccess && density)
    err = hipMemcpyAsync(density, dens, ncells * 8, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess && histogram)
    err = hipMemcpyAsync(histogram, hist, ncells * 8, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err == hipSuccess) add_kde_times(e);
  dev_free(dens);
  dev_free(hist);
  dev_free(m);
  if (err != hipSuccess) return hip_fail(e, err, "kde");
  return RWPT_OK;
}

int rwpt_kde_device(rwpt_engine *e, const rwpt_kde_options *opt, double *density_dev,
                    double *histogram_dev) {
  if (!e || !opt) return RWPT_E_ARG;
  if (!density_dev || !histogram_dev)
    return fail(e, RWPT_E_ARG, "rwpt_kde_device needs both device output buffers");
  int64_t ncells = 0;
  int rc = check_kde_options(e, opt, &ncells);
  if (rc != RWPT_OK) return rc;
  if (opt->n_opt_loops != 0)
    return fail(e, RWPT_E_UNSUPPORTED, "rwpt_kde_device: one pass only (n_opt_loops 0)");
  if (opt->weighted)
    return fail(e, RWPT_E_UNSUPPORTED, "rwpt_kde_device: unit masses only");
  if (opt->initial)
    return fail(e, RWPT_E_UNSUPPORTED, "rwpt_kde_device: active particles only");
  if (ncells > RWPT_KDE_DEVICE_MAX_CELLS)
    return fail(e, RWPT_E_UNSUPPORTED,
                "rwpt_kde_device: grid above RWPT_KDE_DEVICE_MAX_CELLS bins");
  if (e->part.n > 0 && !e->part.keys[0])
    return fail(e, RWPT_E_STATE, "particles not loaded");
  hipError_t err = launch_kde_device(e, opt, density_dev, histogram_dev, e->kde_last);
  if (err == hipErrorInvalidConfiguration)
    return fail(e, RWPT_E_UNSUPPORTED,
                "rwpt_kde_device: kernel above RWPT_KDE_DEVICE_MAX_TAPS weights");
  if (err == hipErrorNotInitialized)
    return fail(e, RWPT_E_STATE, "rwpt_kde_device: scratch not set up on this device");
  if (err != hipSuccess) return hip_fail(e, err, "kde_device");
  add_kde_times(e);
  return RWPT_OK;
}

int rwpt_kde_info(const rwpt_engine *e, double *bandwidth, double *points,
                  double *npoints, double *eff_mass, int32_t *loops) {
  if (!e) return RWPT_E_ARG;
  if (loops) *loops = e->kde_loops;
  if (bandwidth) *bandwidth = e->kde_last[0];
  if (points) *points = e->kde_last[1];
  if (npoints) *npoints = e->kde_last[2];
  if (eff_mass) *eff_mass = e->kde_last[3];
  return RWPT_OK;
}

}  // extern "C"
