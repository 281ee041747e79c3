// C ABI of the MI355X RK4 time step (include/ajet_capi.h): context, device buffers in HBM,
// stream-ordered launches.  The step never needs the host: the time step, the stale
// temperature of the CFL rule and the accumulated time live on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ajet_internal.h"
#include "ajet_kernels.h"

using namespace ajet;

struct ajet_ctx {
  ajet_params p{};
  Setup s;
  Geometry g{};
  PhysConst pc{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;             // boundary tiles run here beside the interior tiles
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double *Q0 = nullptr, *Q1 = nullptr, *Q2 = nullptr, *A = nullptr, *S = nullptr, *tgt = nullptr;
  double *lad = nullptr, *ladA = nullptr, *ladB = nullptr;  // mu*, beta*, kappa* and two 3-field scratch sets
  double *dzl = nullptr, *drl = nullptr, *ramp = nullptr;   // LAD grid spacings and ramp (1-D)
  uint32_t* meta = nullptr;
  uint8_t* codes = nullptr;  // low byte of meta (the LAD filter's row codes)
  double* rinv = nullptr;  // 1/r(j), 0 on the axis (interior-tile stage kernels)
  double *zc = nullptr, *rc = nullptr, *dxidz = nullptr, *ddxidxidz = nullptr, *detadr = nullptr, *ddetadetadr = nullptr;
  PatchDev* patches = nullptr;
  uint8_t* seq = nullptr;
  int *zcuts = nullptr, *rcuts = nullptr;
  int *int_tiles = nullptr, *bnd_tiles = nullptr;  // interior and boundary tile lists
  int n_int = 0, n_bnd = 0;
  int* mix_tiles = nullptr;  // both in one list (mixed_tile_list), for the LAD-mode stage kernel
  int n_mix = 0;
  int4 *int_box = nullptr, *bnd_box = nullptr, *mix_box = nullptr;  // the three lists' tile bounds
  int* lad_tiles = nullptr;  // the source kernel's tiles: fused, or reading the strain kernel's values
  int n_lad_tiles = 0;
  int* lad_str_tiles = nullptr;  // the strain kernel's tiles: interior or boundary
  int n_lad_str = 0;
  int* lad_groups = nullptr;                       // line groups of the fused z/r filter
  int n_lad_groups = 0;
  int lad_seg_first = 0, lad_src_imin = -1, lad_str_imin = -1;  // zero-ramp work skipped
  int lad_seg = 64;                                               // filter segment length
  StepScalars* sc = nullptr;
  double* dtlog = nullptr;
  long long logcap = AJET_DTLOG_CAP;
  bool dt_valid = false;     // sc->dt holds a time step
  bool fused_valid = false;  // sc->dmax_bits holds the CFL reduction of the current state
  bool pending_finalize = false;  // the next stage 1 turns that reduction into dt
  bool have_state = false;
  bool forcing_set = false;  // ajet_set_forcing was called
  int parity = 0;            // step parity (which StepScalars.tpar slot holds the step start time)
  ForcingDev fdev{};
  // learned LAD coefficient (ajet_set_learned_lad)
  bool learned = false;
  double* lad_net = nullptr;               // the network's mu*, beta*, kappa* (3 fields), kept between refreshes
  bool lad_restore = false;                // lad holds the reference's LAD in the blocks flagged at gen_written
  bool lad_full_restore = false;           // lad was overwritten everywhere (test hooks): copy lad_net back
  uint32_t gen_written = 0;                // stamp of the blocks the reference's LAD last wrote
  __bf16* mlpW = nullptr;                  // packed BF16 weights
  float *mlpB = nullptr, *mlpShift = nullptr, *mlpScale = nullptr;
  float out_scale[2] = {1.0f, 1.0f};
  double sensor_thr = 0.0, ramp_min = 0.0;
  int fallback_parts = 0;                  // parts a flagged filter segment is walked in (0: eight)
  int refresh = 1;                         // network every refresh steps
  long long learned_steps = 0;             // steps since the network last ran (0: run it next)
  uint32_t gen = 0;                        // stamp of the sensor's flags (one per step)
  bool sensor_ready = false;               // the sensor of the current state ran after the last step
  bool skip_lad = false;                   // this step flagged no block: no LAD launches
  bool last_skip = false;                  // the last step flagged no block
  uint32_t* any_dev = nullptr;             // the sensor's stamp when it flagged a block
  int sensor_i0 = 0;                       // first column whose ramp exceeds ramp_min
  std::vector<double> ramp_h;              // the ramp, on the host
  uint32_t* sensor_host = nullptr;         // pinned, fine-grained: [0] stamp, [1] any block flagged
  uint32_t* sensor_host_dev = nullptr;     // ... its device address
  hipEvent_t ev_sensor = nullptr;          // the sensor's result is published
  std::vector<int> lad_groups_h;           // host copy of the line groups
  int nseg = 0;                            // z segments of the filter blocks
  uint32_t* zr_flags = nullptr;            // flagged filter blocks (segment + nseg * group)
  uint8_t *src_mask = nullptr, *str_mask = nullptr;
  int* line_group = nullptr;               // group whose output lines hold line j
  int infer_blocks = 0;
  float* feat = nullptr;                   // feature hook scratch
  int* feat_pts = nullptr;
  int feat_cap = 0;
  int* probe_pts = nullptr;                // ajet_probe: points and values on the device
  double* probe_out = nullptr;
  int probe_n = 0;
  std::string err;
};

static const char* kVersion = "axisymmetricjet_amd 0.1 (gfx950, FP64)";

static int fail(ajet_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

static int hip_fail(ajet_ctx* c, hipError_t e, const char* what) {
  if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
  return (int)e;
}

#define HIPCHK(c, expr)                                   \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) return hip_fail((c), e_, #expr); \
  } while (0)

extern "C" {

const char* ajet_version(void) { return kVersion; }

void ajet_params_default(ajet_params* p, int32_t nz, int32_t nr) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->nz = nz;
  p->nr = nr;
  p->viscous = 1;          // do_viscous_terms (ModuleVariables.f90:71)
  p->shock_capturing = 0;  // ShockCapturing (ModuleVariables.f90:152)
  p->forcing = 0;          // ControlForcing (ModuleVariables.f90:185)
  p->gam = 1.4;            // ModuleVariables.f90:66
  p->Re = 1000.0;          // ModuleVariables.f90:73
  p->Pr = 0.72;
  p->CFL = 0.5;            // ModuleVariables.f90:177
  p->C_mu = 0.0;           // ModuleVariables.f90:164
  p->C_beta = 0.0;
  p->C_kappa = 0.0;
  p->sponge_start_top = 12.5;  // ModuleVariables.f90:12-13
  p->sponge_A_top = 0.5;
  p->sponge_end_left = 6.0;
  p->sponge_A_left = 5.0;
}

int ajet_build_patches(int32_t nz, int32_t nr, const ajet_bc* lines, int32_t nlines, ajet_patch* out,
                       int32_t max_out) {
  if ((!lines && nlines > 0) || nlines < 0 || nz < 1 || nr < 1) return AJET_EINVAL;
  std::vector<ajet_patch> v;
  const int n = build_patches(nz, nr, lines, nlines, v);
  if (n < 0) return n;
  for (int k = 0; k < n && k < max_out; ++k) out[k] = v[k];
  return n;
}

const char* ajet_last_error(const ajet_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* ajet_stream(ajet_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int ajet_destroy(ajet_ctx* c) {
  if (!c) return AJET_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->Q0, c->Q1, c->Q2, c->A, c->S, c->tgt, c->lad, c->ladA, c->ladB, c->dzl, c->drl, c->ramp,
                  c->meta, c->zc, c->rc, c->dxidz,
                  c->ddxidxidz, c->detadr, c->ddetadetadr, c->patches, c->seq, c->zcuts, c->rcuts,
                  c->int_tiles, c->bnd_tiles, c->mix_tiles, c->lad_tiles, c->lad_str_tiles, c->lad_groups,
                  c->codes, c->sc, c->dtlog,
                  c->mlpW, c->mlpB, c->mlpShift, c->mlpScale, c->zr_flags, c->src_mask, c->str_mask,
                  c->line_group, c->feat, c->feat_pts, c->probe_pts, c->probe_out, c->any_dev, c->lad_net, c->rinv,
                  c->int_box, c->bnd_box, c->mix_box};
  if (c->aux) (void)hipStreamSynchronize(c->aux);
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->sensor_host) (void)hipHostFree(c->sensor_host);
  if (c->ev_sensor) (void)hipEventDestroy(c->ev_sensor);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AJET_OK;
}

int ajet_create(const ajet_params* p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int32_t nlines, ajet_ctx** out) {
  if (!out) return AJET_EINVAL;
  *out = nullptr;
  if (!p || !z || !r || (!lines && nlines > 0) || nlines < 0) return AJET_EINVAL;
  ajet_ctx* c = new ajet_ctx();
  // the caller gets the context on every failure path below (for ajet_last_error) and destroys it
  *out = c;
  c->p = *p;
  int rc = build_setup(*p, z, r, iblank, lines, nlines, c->s, c->err);
  if (rc == AJET_OK) rc = build_tiling(c->s, p->nz, p->nr, c->err);
  if (rc == AJET_OK && (p->forcing & ~(AJET_FORCING_FEEDBACK | AJET_FORCING_EXCITATION)) != 0)
    rc = fail(c, AJET_EINVAL, "unknown forcing bits");
  int ndev = 0;
  if (rc == AJET_OK && (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1))
    rc = fail(c, AJET_ENODEV, "no HIP device");
  if (rc == AJET_OK) {
    (void)hipGetDevice(&c->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      rc = fail(c, AJET_ENODEV, std::string("device is not gfx950: ") + prop.gcnArchName);
  }
  if (rc != AJET_OK) {
    std::fprintf(stderr, "ajet_create: %s\n", c->err.c_str());
    return rc;
  }
  const int nz = p->nz, nr = p->nr;
  Geometry& g = c->g;
  g.nz = nz;
  g.nr = nr;
  g.pitch = (nz + 15) / 16 * 16;
  g.fstride = (long long)g.pitch * nr;
  g.ntz = (int)c->s.zcuts.size() - 1;
  g.ntr = (int)c->s.rcuts.size() - 1;
  g.any_blank = c->s.any_blank;
  g.iT = c->s.iT;
  g.jT = c->s.jT;
  g.nseq = c->s.nseq;
  g.jtop = nr;
  for (int j = nr - 1; j >= 0; --j)
    if (c->s.rc[(size_t)j] >= p->sponge_start_top) g.jtop = j;
  c->pc.gam = p->gam;
  c->pc.Re = p->Re;
  c->pc.Pr = p->Pr;
  c->pc.sponge_start_top = p->sponge_start_top;
  c->pc.sponge_A_top = p->sponge_A_top;
  c->pc.sponge_end_left = p->sponge_end_left;
  c->pc.sponge_A_left = p->sponge_A_left;
  c->pc.r_top = c->s.rc[nr - 1];
  c->pc.inv_gam = 1.0 / p->gam;
  c->pc.gm1 = p->gam - 1.0;
  c->pc.inv_Re = 1.0 / p->Re;
  c->pc.g_RePr = p->gam / (p->Re * p->Pr);
  c->pc.g_gm1_Re = p->gam * (p->gam - 1.0) / p->Re;
  c->pc.cfl_visc = std::max(1.0 / p->Re, 1.0 / (p->Re * p->Pr));
  if (5ull * (unsigned long long)g.fstride * sizeof(double) >= (1ull << 31)) {
    std::fprintf(stderr, "ajet_create: grid too large\n");
    return fail(c, AJET_EGRID, "a state of the grid needs 2 GiB or more (the kernels address it with 31-bit offsets)");
  }

  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  const size_t sbytes = sizeof(double) * 5 * (size_t)g.fstride;
  for (double** b : {&c->Q0, &c->Q1, &c->Q2, &c->A, &c->S, &c->tgt}) {
    HIPCHK(c, hipMalloc((void**)b, sbytes));
    HIPCHK(c, hipMemsetAsync(*b, 0, sbytes, c->stream));
  }
  if (p->shock_capturing == 1) {
    for (double** b : {&c->lad, &c->ladA, &c->ladB}) {
      HIPCHK(c, hipMalloc((void**)b, sizeof(double) * 3 * (size_t)g.fstride));
      HIPCHK(c, hipMemsetAsync(*b, 0, sizeof(double) * 3 * (size_t)g.fstride, c->stream));
    }
  }
  std::vector<uint32_t> meta((size_t)g.fstride, ROW_BLANK | (ROW_BLANK << 4));
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < nz; ++i) meta[(size_t)i + (size_t)g.pitch * j] = c->s.meta[(size_t)i + (size_t)nz * j];
  HIPCHK(c, hipMalloc((void**)&c->meta, sizeof(uint32_t) * meta.size()));
  HIPCHK(c, hipMemcpyAsync(c->meta, meta.data(), sizeof(uint32_t) * meta.size(), hipMemcpyHostToDevice, c->stream));
  {
    std::vector<uint8_t> codes(meta.size());
    for (size_t k = 0; k < meta.size(); ++k) codes[k] = (uint8_t)(meta[k] & 255u);
    HIPCHK(c, hipMalloc((void**)&c->codes, codes.size()));
    HIPCHK(c, hipMemcpy(c->codes, codes.data(), codes.size(), hipMemcpyHostToDevice));
  }
  std::vector<double> dzl, drl, ramp;
  lad_tables(c->s, nz, nr, dzl, drl, ramp);
  c->ramp_h = ramp;
  std::vector<double> rinv((size_t)nr);
  for (int j = 0; j < nr; ++j) rinv[(size_t)j] = c->s.rc[(size_t)j] > 0.0 ? 1.0 / c->s.rc[(size_t)j] : 0.0;
  struct V { double** d; const std::vector<double>* h; } vs[] = {
      {&c->zc, &c->s.zc}, {&c->rc, &c->s.rc}, {&c->dxidz, &c->s.dxidz}, {&c->ddxidxidz, &c->s.ddxidxidz},
      {&c->detadr, &c->s.detadr}, {&c->ddetadetadr, &c->s.ddetadetadr}, {&c->dzl, &dzl}, {&c->drl, &drl},
      {&c->ramp, &ramp}, {&c->rinv, &rinv}};
  for (auto& v : vs) {
    HIPCHK(c, hipMalloc((void**)v.d, sizeof(double) * v.h->size()));
    HIPCHK(c, hipMemcpyAsync(*v.d, v.h->data(), sizeof(double) * v.h->size(), hipMemcpyHostToDevice, c->stream));
  }
  std::vector<PatchDev> pd(MAX_PATCH);
  for (size_t k = 0; k < c->s.patches.size(); ++k) {
    const ajet_patch& P = c->s.patches[k];
    pd[k] = PatchDev{(int32_t)P.bc, (int32_t)P.bc_dir, (int32_t)P.i_start, (int32_t)P.i_end, (int32_t)P.j_start,
                     (int32_t)P.j_end};
  }
  HIPCHK(c, hipMalloc((void**)&c->patches, sizeof(PatchDev) * MAX_PATCH));
  HIPCHK(c, hipMemcpyAsync(c->patches, pd.data(), sizeof(PatchDev) * MAX_PATCH, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMalloc((void**)&c->seq, c->s.seq.size()));
  HIPCHK(c, hipMemcpyAsync(c->seq, c->s.seq.data(), c->s.seq.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMalloc((void**)&c->zcuts, sizeof(int) * c->s.zcuts.size()));
  HIPCHK(c, hipMemcpyAsync(c->zcuts, c->s.zcuts.data(), sizeof(int) * c->s.zcuts.size(), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMalloc((void**)&c->rcuts, sizeof(int) * c->s.rcuts.size()));
  HIPCHK(c, hipMemcpyAsync(c->rcuts, c->s.rcuts.data(), sizeof(int) * c->s.rcuts.size(), hipMemcpyHostToDevice,
                           c->stream));
  {
    std::vector<int> it, bt;
    classify_tiles(c->s, nz, nr, it, bt);
    c->n_int = (int)it.size();
    c->n_bnd = (int)bt.size();
    HIPCHK(c, hipMalloc((void**)&c->int_tiles, sizeof(int) * (it.size() + 1)));
    HIPCHK(c, hipMalloc((void**)&c->bnd_tiles, sizeof(int) * (bt.size() + 1)));
    if (!it.empty())
      HIPCHK(c, hipMemcpy(c->int_tiles, it.data(), sizeof(int) * it.size(), hipMemcpyHostToDevice));
    if (!bt.empty())
      HIPCHK(c, hipMemcpy(c->bnd_tiles, bt.data(), sizeof(int) * bt.size(), hipMemcpyHostToDevice));
    // the step scalars are written by the first boundary tile of a launch
    if (bt.empty()) return fail(c, AJET_EINVAL, "the grid has no boundary tile");
    std::vector<int> mix;
    mixed_tile_list(it, bt, mix);
    c->n_mix = (int)mix.size();
    HIPCHK(c, hipMalloc((void**)&c->mix_tiles, sizeof(int) * mix.size()));
    HIPCHK(c, hipMemcpy(c->mix_tiles, mix.data(), sizeof(int) * mix.size(), hipMemcpyHostToDevice));
    // each list entry's owned columns and rows, so that a block needs one load to find its tile
    const int ntz = g.ntz;
    auto boxes = [&](const std::vector<int>& list, int4** dst) -> hipError_t {
      std::vector<int4> b(list.size() + 1, make_int4(0, 0, 0, 0));
      for (size_t k = 0; k < list.size(); ++k) {
        const int t = list[k] >= 0 ? list[k] : ~list[k];
        b[k] = make_int4(c->s.zcuts[(size_t)(t % ntz)], c->s.zcuts[(size_t)(t % ntz) + 1], c->s.rcuts[(size_t)(t / ntz)],
                         c->s.rcuts[(size_t)(t / ntz) + 1]);
      }
      hipError_t e = hipMalloc((void**)dst, sizeof(int4) * b.size());
      return e == hipSuccess ? hipMemcpy(*dst, b.data(), sizeof(int4) * b.size(), hipMemcpyHostToDevice) : e;
    };
    HIPCHK(c, boxes(it, &c->int_box));
    HIPCHK(c, boxes(bt, &c->bnd_box));
    HIPCHK(c, boxes(mix, &c->mix_box));
    std::vector<int> si, sb;
    classify_lad_source(c->s, nz, nr, it, bt, si, sb);
    auto upload = [&](const std::vector<int>& list, int** dst, int* n) {
      *n = (int)list.size();
      hipError_t e = hipMalloc((void**)dst, sizeof(int) * (list.size() + 1));
      if (e == hipSuccess && !list.empty())
        e = hipMemcpy(*dst, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice);
      return e;
    };
    mixed_tile_list(it, bt, mix);
    HIPCHK(c, upload(mix, &c->lad_tiles, &c->n_lad_tiles));
    mixed_tile_list(si, sb, mix);
    HIPCHK(c, upload(mix, &c->lad_str_tiles, &c->n_lad_str));
    std::vector<int> gr;
    build_lad_line_groups(c->s, nz, nr, LAD_LINES, gr);
    c->n_lad_groups = (int)gr.size() / 3;
    c->lad_groups_h = gr;
    HIPCHK(c, hipMalloc((void**)&c->lad_groups, sizeof(int) * gr.size()));
    HIPCHK(c, hipMemcpy(c->lad_groups, gr.data(), sizeof(int) * gr.size(), hipMemcpyHostToDevice));
  }
  c->lad_seg = lad_zfilter_segment(ramp.data(), nz, c->n_lad_groups, lad_zrfilter_slots(c->device));
  c->nseg = (nz + c->lad_seg - 1) / c->lad_seg;
  lad_skip_bounds(ramp.data(), nz, c->lad_seg, c->lad_seg_first, c->lad_src_imin, c->lad_str_imin);
  HIPCHK(c, hipMalloc((void**)&c->sc, sizeof(StepScalars)));
  HIPCHK(c, hipMemsetAsync(c->sc, 0, sizeof(StepScalars), c->stream));
  HIPCHK(c, hipMalloc((void**)&c->dtlog, sizeof(double) * (size_t)c->logcap));
  double A[9][6], B[9][6], F[9][11];
  sbp_tables(A, B, F);
  HIPCHK(c, upload_tables(A, B, F, c->stream));
  HIPCHK(c, upload_lad_tables(A, B, F, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

static void copy_state_in(ajet_ctx* c, double* dst, const double* src) {
  (void)hipMemcpy2DAsync(dst, sizeof(double) * c->g.pitch, src, sizeof(double) * c->g.nz, sizeof(double) * c->g.nz,
                         (size_t)c->g.nr * 5, hipMemcpyHostToDevice, c->stream);
}

int ajet_set_state(ajet_ctx* c, const double* pv, const double* pv_target) {
  if (!c || !c->stream || !pv) return AJET_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  copy_state_in(c, c->Q0, pv);
  if (pv_target) copy_state_in(c, c->tgt, pv_target);
  c->learned_steps = 0;
  c->sensor_ready = false;
  c->last_skip = false;
  StepScalars z{};
  HIPCHK(c, hipMemcpyAsync(c->sc, &z, sizeof z, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dt_valid = false;
  c->fused_valid = false;
  c->pending_finalize = false;
  c->have_state = true;
  c->parity = 0;
  return AJET_OK;
}

int ajet_set_forcing(ajet_ctx* c, const ajet_forcing* f) {
  if (!c || !f) return AJET_EINVAL;
  if (c->p.forcing == 0) return fail(c, AJET_EINVAL, "the context was created without forcing");
  if (!(f->lz > 0.0) || !(f->lr > 0.0)) return fail(c, AJET_EINVAL, "forcing widths lz, lr must be positive");
  ForcingDev d{};
  const double alpha = 6.0;  // ModuleVariables.f90:184
  for (int e = 0; e < 5; ++e)
    for (int k = 0; k < 5; ++k) d.aF[e][k] = alpha * f->F[e][k];
  d.z0 = f->z0;
  d.r0 = f->r0;
  d.lz2 = f->lz * f->lz;
  d.lr2 = f->lr * f->lr;
  d.au = f->amp * f->uj;
  d.w = 2.0 * 3.14159265358979323846 * f->strouhal * f->uj;
  d.t0 = f->t0;
  c->fdev = d;
  c->forcing_set = true;
  return AJET_OK;
}

int ajet_get_state(ajet_ctx* c, double* pv) {
  if (!c || !c->stream || !pv) return AJET_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpy2DAsync(pv, sizeof(double) * c->g.nz, c->Q0, sizeof(double) * c->g.pitch,
                             sizeof(double) * c->g.nz, (size_t)c->g.nr * 5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

static StageArgs stage_args(ajet_ctx* c) {
  StageArgs a{};
  a.tgt = c->tgt;
  a.meta = c->meta;
  a.zc = c->zc;
  a.dxidz = c->dxidz;
  a.ddxidxidz = c->ddxidxidz;
  a.rc = c->rc;
  a.rinv = c->rinv;
  a.detadr = c->detadr;
  a.ddetadetadr = c->ddetadetadr;
  a.lad = c->lad;
  a.patches = c->patches;
  a.seq = c->seq;
  a.zcuts = c->zcuts;
  a.rcuts = c->rcuts;
  a.sc = c->sc;
  a.g = c->g;
  a.pc = c->pc;
  a.dtlog = c->dtlog;
  a.logcap = c->logcap;
  a.CFL = c->p.CFL;
  a.dt_from_dmax = 0;
  a.forcing = c->p.forcing;
  a.parity = c->parity;
  a.fc = c->fdev;
  return a;
}

// One stage over the whole grid: the boundary tiles (general kernel; it also updates the
// step scalars) and then the interior tiles (specialised kernel).
static int physics_mode(const ajet_params& p) {
  if (p.viscous != 1) return MODE_INVISCID;
  if (p.shock_capturing != 1) return MODE_PLAIN;
  return (p.C_mu != 0.0 || p.C_beta != 0.0 || p.C_kappa != 0.0) ? MODE_LAD : MODE_SC;
}

static LadArgs lad_args(ajet_ctx* c, const double* q) {
  LadArgs l{};
  l.q = q;
  l.meta = c->meta;
  l.codes = c->codes;
  l.dxidz = c->dxidz;
  l.detadr = c->detadr;
  l.rc = c->rc;
  l.dzl = c->dzl;
  l.drl = c->drl;
  l.ramp = c->ramp;
  l.A = c->ladA;
  l.B = c->ladB;
  l.out = c->lad;
  l.zcuts = c->zcuts;
  l.rcuts = c->rcuts;
  l.tiles = c->lad_tiles;
  l.ntiles = c->n_lad_tiles;
  l.str_tiles = c->lad_str_tiles;
  l.nstr = c->n_lad_str;
  l.rgroups = c->lad_groups;
  l.nrgroups = c->n_lad_groups;
  l.seg_first = c->lad_seg_first;
  l.zseg = c->lad_seg;
  l.nseg = c->nseg;
  l.src_imin = c->lad_src_imin;
  l.str_imin = c->lad_str_imin;
  l.zsub = 1;
  l.g = c->g;
  l.gam = c->p.gam;
  // ModShockCapturing.f90:66, 72, 79: C_mu Re, C_beta Re, C_kappa Re Pr
  l.cmu = c->p.C_mu * c->p.Re;
  l.cbeta = c->p.C_beta * c->p.Re;
  l.ckappa = c->p.C_kappa * c->p.Re * c->p.Pr;
  return l;
}

// The artificial coefficients of the stage input come first (ComputeRHS calls
// ComputeArtTransCoeffs after the first derivatives, ModRHS.f90:46-47).  The two stage
// launches then write disjoint points from the same input, so the boundary tiles run on a
// second stream beside the interior tiles; the main stream waits for both before going on.
static MlpInfer mlp_args(ajet_ctx* c, const double* q) {
  MlpInfer m{};
  m.q = q;
  m.g = c->g;
  m.dzl = c->dzl;
  m.drl = c->drl;
  m.ramp = c->ramp;
  m.rc = c->rc;
  m.W = c->mlpW;
  m.bias = c->mlpB;
  m.shift = c->mlpShift;
  m.scale = c->mlpScale;
  m.out_scale[0] = c->out_scale[0];
  m.out_scale[1] = c->out_scale[1];
  m.ramp_min = c->ramp_min;
  // the network's coefficients are kept in lad_net; lad, which the stage kernels read, gets the
  // same values and the reference's LAD in the blocks the sensor flags
  m.out = c->lad_net;
  m.out2 = c->lad;
  return m;
}

// Learned mode: the network's coefficients at the first stage of every refresh-th step; the
// dilatation sensor at the first stage of every step; at every stage the reference's LAD on the
// filter blocks it flagged.
// The dilatation sensor for the state q under a new stamp, and the publication of whether it
// flagged a block to pinned host memory (ev_sensor behind it).
static hipError_t learned_sensor(ajet_ctx* c, const double* q, hipStream_t s) {
  hipError_t e;
  LadArgs l = lad_args(c, q);
  l.nseg = c->nseg;
  l.gen = ++c->gen;
  if ((e = launch_lad_sensor(l, c->zr_flags, c->any_dev, c->line_group, c->sensor_thr, c->ramp_min, c->sensor_i0,
                             c->sensor_host_dev, s)) != hipSuccess)
    return e;
  return hipEventRecord(c->ev_sensor, s);
}

// Waits for the sensor result of the current stamp.  The host polls the pinned word the GPU
// stores last: a blocking wait on the event took ~30 us to wake the host, GPU time left idle
// each step.  After 200 ms of polling (a long queue ahead) it waits on the event instead, which
// also reports a failed launch.
static hipError_t wait_sensor(ajet_ctx* c, bool* any) {
  const uint32_t gen = c->gen;
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1;; ++it) {
    if (__atomic_load_n(&c->sensor_host[0], __ATOMIC_ACQUIRE) == gen) {
      *any = __atomic_load_n(&c->sensor_host[1], __ATOMIC_RELAXED) != 0u;
      return hipSuccess;
    }
    if ((it & 1023u) == 0u && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
  const hipError_t e = hipEventSynchronize(c->ev_sensor);
  if (e != hipSuccess) return e;
  if (__atomic_load_n(&c->sensor_host[0], __ATOMIC_ACQUIRE) != gen) return hipErrorUnknown;
  *any = __atomic_load_n(&c->sensor_host[1], __ATOMIC_RELAXED) != 0u;
  return hipSuccess;
}

// After the last stage of a step: the sensor of the new state, so that the next step knows
// before its first launch whether any block falls back to the reference's LAD.
static hipError_t learned_after_step(ajet_ctx* c, hipStream_t s) {
  hipError_t e = learned_sensor(c, c->Q0, s);
  c->sensor_ready = e == hipSuccess;
  return e;
}

// Learned mode, stage 1, everything before the sensor's verdict: the network at a refresh, the
// sensor of this state if it did not run after the last step, and the network's values put back
// in the blocks the reference's LAD wrote last step.
static hipError_t learned_stage1_pre(ajet_ctx* c, const double* q, hipStream_t s) {
  hipError_t e;
  if (c->learned_steps % c->refresh == 0) {
    // the network: lad_net and lad
    if ((e = launch_mlp_infer(mlp_args(c, q), c->infer_blocks, s)) != hipSuccess) return e;
    c->lad_full_restore = c->lad_restore = false;
  } else if (c->lad_full_restore) {
    if ((e = hipMemcpyAsync(c->lad, c->lad_net, sizeof(double) * 3 * (size_t)c->g.fstride, hipMemcpyDeviceToDevice,
                            s)) != hipSuccess)
      return e;
    c->lad_full_restore = c->lad_restore = false;
  }
  // the sensor of this state: taken after the previous step (wait for its result), or now
  if (!c->sensor_ready && (e = learned_sensor(c, q, s)) != hipSuccess) return e;
  c->sensor_ready = false;
  if (c->lad_restore) {
    // blocks the reference's LAD wrote in the previous step and the sensor no longer flags
    // get the network's values back (blocks flagged again are rewritten below)
    LadArgs l = lad_args(c, q);
    if ((e = launch_lad_restore(l, c->zr_flags, c->gen_written, c->lad_net, c->lad, s)) != hipSuccess) return e;
    c->lad_restore = false;
  }
  return hipSuccess;
}

// The reference's LAD on the blocks this step's sensor flagged (new_step: derive the tile masks
// of the flagged blocks first, once per step).
static hipError_t learned_fallback(ajet_ctx* c, const double* q, bool new_step, hipStream_t s) {
  hipError_t e;
  LadArgs l = lad_args(c, q);
  l.nseg = c->nseg;
  l.gen = c->gen;  // the stamp of this step's sensor
  // the tile masks of the flagged blocks, for the four stages of this step
  if (new_step && (e = launch_lad_masks(l, c->zr_flags, c->src_mask, c->str_mask, s)) != hipSuccess) return e;
  l.zr_flags = c->zr_flags;
  l.src_mask = c->src_mask;
  l.str_mask = c->str_mask;
  // a few flagged blocks: each walks its segment in parts (whole segments when every block is
  // flagged on purpose, a negative threshold, so that the run is the reference's bit for bit)
  l.zsub = c->sensor_thr < 0.0 ? 1 : lad_fallback_zsub(c->lad_seg, c->fallback_parts);
  c->lad_restore = true;
  c->gen_written = c->gen;
  return launch_lad(l, s);
}

static hipError_t launch_learned(ajet_ctx* c, int stage, const double* q, hipStream_t s) {
  hipError_t e;
  if (stage == 4) ++c->learned_steps;
  bool any = false;
  if (stage == 1) {
    if ((e = learned_stage1_pre(c, q, s)) != hipSuccess) return e;
    if ((e = wait_sensor(c, &any)) != hipSuccess) return e;
    c->skip_lad = !any;
    c->last_skip = !any;
  }
  if (c->skip_lad) return hipSuccess;  // no block flagged: the LAD launches would all return at once
  return learned_fallback(c, q, any, s);
}

// The artificial coefficients of one stage of a run in LAD mode (the reference's or learned).
static hipError_t launch_coefficients(ajet_ctx* c, int stage, const double* q, hipStream_t s) {
  return c->learned ? launch_learned(c, stage, q, s) : launch_lad(lad_args(c, q), s);
}

static hipError_t launch_stages(ajet_ctx* c, int stage, StageArgs a) {
  const int mode = physics_mode(c->p);
  hipError_t e;
  if (mode == MODE_LAD && c->learned && stage == 1 && c->last_skip && c->sensor_ready && stage_mixed_supported(1, mode)) {
    // Learned mode, the previous step flagged no block: stage 1 is launched at once on the guess
    // that this step flags none either, and the host waits for the sensor's verdict while the GPU
    // runs it (waiting first left the GPU idle ~22 us per step).  A step that flags blocks runs
    // their reference LAD and stage 1 again; the first launch has logged dt and left it in the
    // step scalars, which the second reads.
    if ((e = learned_stage1_pre(c, a.qin, c->stream)) != hipSuccess) return e;
    a.tiles = c->mix_tiles;
    a.boxes = c->mix_box;
    a.ntiles = c->n_mix;
    if ((e = launch_stage_mixed(stage, mode, a, c->stream)) != hipSuccess) return e;
    bool any = false;
    if ((e = wait_sensor(c, &any)) != hipSuccess) return e;
    c->skip_lad = !any;
    c->last_skip = !any;
    if (!any) return hipSuccess;
    if ((e = learned_fallback(c, a.qin, true, c->stream)) != hipSuccess) return e;
    a.dt_from_dmax = 0;
    return launch_stage_mixed(stage, mode, a, c->stream);
  }
  if (mode == MODE_LAD && stage == 0) {
    // the test hook's RHS takes the reference's coefficients of its input; in learned mode that
    // overwrites lad, which the next step restores from the network's values in lad_net
    if ((e = launch_lad(lad_args(c, a.qin), c->stream)) != hipSuccess) return e;
    if (c->learned) c->lad_full_restore = true;
  } else if (mode == MODE_LAD && (e = launch_coefficients(c, stage, a.qin, c->stream)) != hipSuccess) {
    return e;
  }
  if (stage_mixed_supported(stage, mode)) {
    // one launch for both tile classes: the boundary tiles alone leave most of the GPU idle for
    // ~30 us, and an event between two streams costs ~6 us of idle GPU each way (plain mode at
    // 4400 x 640: stages 1-3 115.6 -> 105.7, 127 -> 119, 141 -> 133 us)
    a.tiles = c->mix_tiles;
    a.boxes = c->mix_box;
    a.ntiles = c->n_mix;
    return launch_stage_mixed(stage, mode, a, c->stream);
  }
  if (stage <= 2) {
    // the interior tiles, then the boundary tiles, on the one stream (the other physics modes'
    // stages 1-2): the boundary launch alone takes less than the two events between streams
    // leave the GPU idle; stage 4, whose boundary tiles take longer, keeps them beside the
    // interior tiles on the second stream
    a.tiles = c->int_tiles;
    a.boxes = c->int_box;
    a.ntiles = c->n_int;
    if ((e = launch_stage(stage, mode, true, a, c->stream)) != hipSuccess) return e;
    a.tiles = c->bnd_tiles;
    a.boxes = c->bnd_box;
    a.ntiles = c->n_bnd;
    return launch_stage(stage, mode, false, a, c->stream);
  }
  if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(c->aux, c->ev_fork, 0)) != hipSuccess) return e;
  a.tiles = c->bnd_tiles;
  a.boxes = c->bnd_box;
  a.ntiles = c->n_bnd;
  if ((e = launch_stage(stage, mode, false, a, c->aux)) != hipSuccess) return e;
  if ((e = hipEventRecord(c->ev_join, c->aux)) != hipSuccess) return e;
  a.tiles = c->int_tiles;
  a.boxes = c->int_box;
  a.ntiles = c->n_int;
  if ((e = launch_stage(stage, mode, true, a, c->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess) return e;
  if (mode == MODE_LAD && c->learned && stage == 4) return learned_after_step(c, c->stream);
  return hipSuccess;
}

static int read_dt(ajet_ctx* c, double* dt) {
  StepScalars h{};
  HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (dt) *dt = h.dt;
  if (h.nonfinite) return fail(c, AJET_ENONFINITE, "non-finite value in the time-step reduction");
  if (h.dt == 100.0) return fail(c, AJET_EBLOWUP, "Time step blow up");
  return AJET_OK;
}

int ajet_compute_dt(ajet_ctx* c, double* dt) {
  if (!c || !c->stream) return AJET_EINVAL;
  if (!c->have_state) return fail(c, AJET_EINVAL, "no state loaded");
  HIPCHK(c, hipSetDevice(c->device));
  hipError_t e = hipSuccess;
  if (c->fused_valid && !dt) {
    // the previous step's last stage holds the reduction; the next stage 1 finishes it
    c->fused_valid = false;
    c->pending_finalize = true;
    c->dt_valid = true;
    return AJET_OK;
  }
  c->pending_finalize = false;
  if (c->fused_valid) {
    e = launch_finalize(c->sc, c->p.CFL, c->g.any_blank, c->dtlog, c->logcap, c->stream);
  } else {
    CflArgs a{};
    a.q = c->Q0;
    a.meta = c->meta;
    a.dxidz = c->dxidz;
    a.detadr = c->detadr;
    a.sc = c->sc;
    a.g = c->g;
    a.Re = c->p.Re;
    a.Pr = c->p.Pr;
    e = launch_cfl(a, c->p.CFL, c->g.any_blank, c->dtlog, c->logcap, c->stream);
  }
  if (e != hipSuccess) return hip_fail(c, e, "time-step reduction launch");
  c->fused_valid = false;
  c->dt_valid = true;
  if (dt) return read_dt(c, dt);
  return AJET_OK;
}

// Buffers of the four stages: q^n in Q0, q(1) in Q1, q(2) in Q2, q(3) in Q1, q^{n+1} in Q0.
int ajet_rk4_stage(ajet_ctx* c, int32_t stage) {
  if (!c || !c->stream || stage < 1 || stage > 4) return AJET_EINVAL;
  if (!c->dt_valid) return fail(c, AJET_EINVAL, "no time step: call ajet_compute_dt first");
  if (c->p.forcing && !c->forcing_set) return fail(c, AJET_EINVAL, "forcing parameters not set: call ajet_set_forcing");
  StageArgs a = stage_args(c);
  a.qn = c->Q0;
  a.acc_a = c->A;
  a.acc_s = c->S;
  switch (stage) {
    case 1: a.qin = c->Q0; a.qout = c->Q1; break;
    case 2: a.qin = c->Q1; a.qout = c->Q2; break;
    case 3: a.qin = c->Q2; a.qout = c->Q1; break;
    default: a.qin = c->Q1; a.qout = c->Q0; break;
  }
  if (stage == 1 && c->pending_finalize) a.dt_from_dmax = 1;
  HIPCHK(c, launch_stages(c, stage, a));
  if (stage == 1) c->pending_finalize = false;
  c->fused_valid = stage == 4;
  if (stage == 4) c->parity ^= 1;
  return AJET_OK;
}

int ajet_rk4_step(ajet_ctx* c) {
  if (!c || !c->stream) return AJET_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  for (int s = 1; s <= 4; ++s) {
    const int rc = ajet_rk4_stage(c, s);
    if (rc) return rc;
  }
  return AJET_OK;
}

int ajet_run(ajet_ctx* c, int32_t nsteps) {
  if (!c || nsteps < 0) return AJET_EINVAL;
  for (int32_t k = 0; k < nsteps; ++k) {
    int rc = ajet_compute_dt(c, nullptr);
    if (rc) return rc;
    rc = ajet_rk4_step(c);
    if (rc) return rc;
  }
  return AJET_OK;
}

int ajet_get_dt_log(ajet_ctx* c, double* dts, int32_t n, double* time) {
  if (!c || !c->stream) return AJET_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  StepScalars h{};
  HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const long long have = h.nlog < c->logcap ? h.nlog : c->logcap;
  if (dts && n > 0) {
    const long long m = have < n ? have : n;
    if (m > 0) HIPCHK(c, hipMemcpy(dts, c->dtlog, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost));
  }
  if (time) *time = h.time;
  if (h.nonfinite) return fail(c, AJET_ENONFINITE, "non-finite value in the time-step reduction");
  return (int)(h.nlog > 0x7fffffff ? 0x7fffffff : h.nlog);
}

int ajet_rhs(ajet_ctx* c, const double* q, double* rhs) {
  if (!c || !c->stream || !q || !rhs) return AJET_EINVAL;
  if (c->p.forcing && !c->forcing_set) return fail(c, AJET_EINVAL, "forcing parameters not set: call ajet_set_forcing");
  HIPCHK(c, hipSetDevice(c->device));
  copy_state_in(c, c->Q2, q);
  StageArgs a = stage_args(c);
  a.qin = c->Q2;
  a.qn = c->Q2;
  a.rhs_out = c->S;
  HIPCHK(c, launch_stages(c, 0, a));
  HIPCHK(c, hipMemcpy2DAsync(rhs, sizeof(double) * c->g.nz, c->S, sizeof(double) * c->g.pitch,
                             sizeof(double) * c->g.nz, (size_t)c->g.nr * 5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

int ajet_lad_coeffs(ajet_ctx* c, const double* q, double* mustar, double* betastar, double* kappastar) {
  if (!c || !c->stream || !q || !mustar || !betastar || !kappastar) return AJET_EINVAL;
  if (c->p.shock_capturing != 1) return fail(c, AJET_EINVAL, "shock capturing is off in this context");
  HIPCHK(c, hipSetDevice(c->device));
  copy_state_in(c, c->Q2, q);
  HIPCHK(c, launch_lad(lad_args(c, c->Q2), c->stream));
  // in learned mode the next step restores lad from the network's values in lad_net
  if (c->learned) c->lad_full_restore = true;
  double* outs[3] = {mustar, betastar, kappastar};
  for (int f = 0; f < 3; ++f)
    HIPCHK(c, hipMemcpy2DAsync(outs[f], sizeof(double) * c->g.nz, c->lad + (size_t)f * c->g.fstride,
                               sizeof(double) * c->g.pitch, sizeof(double) * c->g.nz, (size_t)c->g.nr,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

int ajet_set_probes(ajet_ctx* c, const int32_t* points, int32_t n) {
  if (!c || !c->stream || n < 0 || (n > 0 && !points)) return AJET_EINVAL;
  for (int32_t k = 0; k < n; ++k)
    if (points[2 * k] < 0 || points[2 * k] >= c->g.nz || points[2 * k + 1] < 0 || points[2 * k + 1] >= c->g.nr)
      return fail(c, AJET_EINVAL, "probe outside the grid");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->probe_pts) (void)hipFree(c->probe_pts);
  if (c->probe_out) (void)hipFree(c->probe_out);
  c->probe_pts = nullptr;
  c->probe_out = nullptr;
  c->probe_n = n;
  if (n == 0) return AJET_OK;
  HIPCHK(c, hipMalloc((void**)&c->probe_pts, sizeof(int) * 2 * (size_t)n));
  HIPCHK(c, hipMalloc((void**)&c->probe_out, sizeof(double) * 5 * (size_t)n));
  HIPCHK(c, hipMemcpy(c->probe_pts, points, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice));
  return AJET_OK;
}

int ajet_probe(ajet_ctx* c, double* values) {
  if (!c || !c->stream || !values) return AJET_EINVAL;
  if (c->probe_n == 0) return AJET_OK;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_probe(c->Q0, c->g, c->probe_pts, c->probe_n, c->probe_out, c->stream));
  HIPCHK(c, hipMemcpyAsync(values, c->probe_out, sizeof(double) * 5 * (size_t)c->probe_n, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

}  // extern "C"

bool ajet::ctx_probe_view(ajet_ctx* c, ProbeView* v) {
  if (!c || !c->stream) return false;
  v->device = c->device;
  v->stream = c->stream;
  v->q = c->Q0;
  v->g = c->g;
  return true;
}

void ajet::ctx_set_error(ajet_ctx* c, const std::string& msg) {
  if (c) c->err = msg;
}

extern "C" {

// ---------------------------------------------------------------- learned LAD coefficient

int ajet_set_learned_lad(ajet_ctx* c, const ajet_mlp* model, double sensor_threshold, double ramp_min,
                         int32_t refresh) {
  if (!c || !c->stream) return AJET_EINVAL;
  c->learned_steps = 0;  // the next step runs the network, which writes lad_net and lad everywhere
  c->sensor_ready = false;
  c->lad_restore = c->lad_full_restore = false;
  if (!model) {
    c->learned = false;
    return AJET_OK;
  }
  if (physics_mode(c->p) != MODE_LAD)
    return fail(c, AJET_EINVAL, "the learned coefficient needs shock capturing with non-zero C_mu, C_beta, C_kappa");
  if (!(ramp_min >= 0.0) || sensor_threshold != sensor_threshold) return fail(c, AJET_EINVAL, "bad threshold");
  if (refresh < 1) return fail(c, AJET_EINVAL, "refresh must be at least 1");
  // AJET_FALLBACK_PARTS: the most parts a flagged filter segment is walked in (unset: eight;
  // 1: whole segments, the order of the reference-LAD launches)
  int parts = 0;
  if (const char* env = std::getenv("AJET_FALLBACK_PARTS")) {
    char* end = nullptr;
    const long v = std::strtol(env, &end, 10);
    if (end == env || *end != '\0' || (v != 1 && v != 2 && v != 4 && v != 8))
      return fail(c, AJET_EINVAL, "AJET_FALLBACK_PARTS must be 1, 2, 4 or 8");
    parts = (int)v;
  }
  HIPCHK(c, hipSetDevice(c->device));
  const Geometry& g = c->g;
  if (!c->mlpW) {
    HIPCHK(c, hipMalloc((void**)&c->mlpW, sizeof(uint16_t) * (3 * MLP_H + 16) * MLP_K));
    HIPCHK(c, hipMalloc((void**)&c->mlpB, sizeof(float) * (3 * MLP_H + MLP_NOUT)));
    HIPCHK(c, hipMalloc((void**)&c->mlpShift, sizeof(float) * MLP_NIN));
    HIPCHK(c, hipMalloc((void**)&c->mlpScale, sizeof(float) * MLP_NIN));
    HIPCHK(c, hipMalloc((void**)&c->zr_flags, sizeof(uint32_t) * (size_t)c->nseg * c->n_lad_groups));
    HIPCHK(c, hipMemset(c->zr_flags, 0, sizeof(uint32_t) * (size_t)c->nseg * c->n_lad_groups));
    HIPCHK(c, hipMalloc((void**)&c->any_dev, sizeof(uint32_t)));
    HIPCHK(c, hipMemset(c->any_dev, 0, sizeof(uint32_t)));
    HIPCHK(c, hipHostMalloc((void**)&c->sensor_host, 2 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
    c->sensor_host[0] = 0u;
    c->sensor_host[1] = 0u;
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->sensor_host_dev, c->sensor_host, 0));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_sensor, hipEventDisableTiming));
    HIPCHK(c, hipMalloc((void**)&c->src_mask, (size_t)g.ntz * g.ntr));
    HIPCHK(c, hipMalloc((void**)&c->str_mask, (size_t)g.ntz * g.ntr));
    HIPCHK(c, hipMemset(c->src_mask, 0, (size_t)g.ntz * g.ntr));
    HIPCHK(c, hipMemset(c->str_mask, 0, (size_t)g.ntz * g.ntr));
    HIPCHK(c, hipMalloc((void**)&c->lad_net, sizeof(double) * 3 * (size_t)g.fstride));
    HIPCHK(c, hipMemset(c->lad_net, 0, sizeof(double) * 3 * (size_t)g.fstride));
    std::vector<int> lg((size_t)g.nr, 0);
    for (int k = 0; k < c->n_lad_groups; ++k)
      for (int j = c->lad_groups_h[3 * k + 1]; j < c->lad_groups_h[3 * k + 2]; ++j) lg[(size_t)j] = k;
    HIPCHK(c, hipMalloc((void**)&c->line_group, sizeof(int) * lg.size()));
    HIPCHK(c, hipMemcpy(c->line_group, lg.data(), sizeof(int) * lg.size(), hipMemcpyHostToDevice));
    int cus = 0;
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    c->infer_blocks = cus > 0 ? cus : 256;
  }
  std::vector<uint16_t> W;
  std::vector<float> bias;
  mlp_pack(*model, W, bias);
  HIPCHK(c, hipMemcpy(c->mlpW, W.data(), sizeof(uint16_t) * W.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->mlpB, bias.data(), sizeof(float) * bias.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->mlpShift, model->shift, sizeof(float) * MLP_NIN, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->mlpScale, model->scale, sizeof(float) * MLP_NIN, hipMemcpyHostToDevice));
  c->out_scale[0] = model->out_scale[0];
  c->out_scale[1] = model->out_scale[1];
  c->sensor_thr = sensor_threshold;
  c->ramp_min = ramp_min;
  c->fallback_parts = parts;
  // the sensor starts at the first column whose ramp exceeds ramp_min (the ramp rises with z)
  c->sensor_i0 = g.nz;
  for (int i = 0; i < g.nz; ++i)
    if (c->ramp_h[(size_t)i] > ramp_min) {
      c->sensor_i0 = i;
      break;
    }
  c->refresh = refresh;
  c->learned = true;
  return AJET_OK;
}

int ajet_lad_features(ajet_ctx* c, const double* q, const int32_t* points, int32_t n, float* x) {
  if (!c || !c->stream || !q || (n > 0 && (!points || !x)) || n < 0) return AJET_EINVAL;
  for (int32_t k = 0; k < n; ++k)
    if (points[2 * k] < 0 || points[2 * k] >= c->g.nz || points[2 * k + 1] < 0 || points[2 * k + 1] >= c->g.nr)
      return fail(c, AJET_EINVAL, "point outside the grid");
  if (n == 0) return AJET_OK;
  HIPCHK(c, hipSetDevice(c->device));
  if (n > c->feat_cap) {
    if (c->feat) (void)hipFree(c->feat);
    if (c->feat_pts) (void)hipFree(c->feat_pts);
    c->feat = nullptr;
    c->feat_pts = nullptr;
    HIPCHK(c, hipMalloc((void**)&c->feat, sizeof(float) * (size_t)n * MLP_NIN));
    HIPCHK(c, hipMalloc((void**)&c->feat_pts, sizeof(int) * 2 * (size_t)n));
    c->feat_cap = n;
  }
  copy_state_in(c, c->Q2, q);
  HIPCHK(c, hipMemcpyAsync(c->feat_pts, points, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_mlp_features(mlp_args(c, c->Q2), c->feat_pts, n, c->feat, c->stream));
  HIPCHK(c, hipMemcpyAsync(x, c->feat, sizeof(float) * (size_t)n * MLP_NIN, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

int ajet_lad_learned_coeffs(ajet_ctx* c, const double* q, double* mustar, double* betastar, double* kappastar) {
  if (!c || !c->stream || !q || !mustar || !betastar || !kappastar) return AJET_EINVAL;
  if (!c->learned) return fail(c, AJET_EINVAL, "no learned model set");
  HIPCHK(c, hipSetDevice(c->device));
  copy_state_in(c, c->Q2, q);
  // into the LAD kernels' scratch set, so that the context's coefficients stay as they are
  MlpInfer m = mlp_args(c, c->Q2);
  m.out = c->ladA;
  m.out2 = nullptr;
  HIPCHK(c, launch_mlp_infer(m, c->infer_blocks, c->stream));
  double* outs[3] = {mustar, betastar, kappastar};
  for (int f = 0; f < 3; ++f)
    HIPCHK(c, hipMemcpy2DAsync(outs[f], sizeof(double) * c->g.nz, c->ladA + (size_t)f * c->g.fstride,
                               sizeof(double) * c->g.pitch, sizeof(double) * c->g.nz, (size_t)c->g.nr,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AJET_OK;
}

int ajet_lad_fallback_fraction(ajet_ctx* c, double* fraction) {
  if (!c || !fraction) return AJET_EINVAL;
  if (!c->learned) return fail(c, AJET_EINVAL, "no learned model set");
  std::vector<uint32_t> f((size_t)c->nseg * c->n_lad_groups);
  HIPCHK(c, hipMemcpyAsync(f.data(), c->zr_flags, sizeof(uint32_t) * f.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t k = 0;
  for (uint32_t v : f) k += v == c->gen && c->gen != 0u;
  *fraction = f.empty() ? 0.0 : (double)k / (double)f.size();
  return AJET_OK;
}

// ---------------------------------------------------------------- ensemble

struct ajet_batch {
  std::vector<ajet_ctx*> runs;
  int device = 0;
  int mode = MODE_PLAIN;
  bool force = false;
  hipStream_t stream = nullptr, aux = nullptr;  // those of the first run
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_done = nullptr;
  StageArgs* args[4] = {nullptr, nullptr, nullptr, nullptr};  // per stage, one entry per run
  uint32_t *d_int = nullptr, *d_bnd = nullptr;
  int n_int = 0, n_bnd = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // one step, by step parity
  std::string err;
};

static int bfail(ajet_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIPCHK(b, expr)                                                         \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      if (b) (b)->err = std::string(#expr) + ": " + hipGetErrorString(e_);       \
      return (int)e_;                                                            \
    }                                                                            \
  } while (0)

const char* ajet_batch_last_error(const ajet_batch* b) { return b ? b->err.c_str() : "null batch"; }

int ajet_batch_destroy(ajet_batch* b) {
  if (!b) return AJET_EINVAL;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->aux) (void)hipStreamSynchronize(b->aux);
  for (auto& g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto* a : b->args)
    if (a) (void)hipFree(a);
  if (b->d_int) (void)hipFree(b->d_int);
  if (b->d_bnd) (void)hipFree(b->d_bnd);
  for (hipEvent_t e : {b->ev_fork, b->ev_join, b->ev_done})
    if (e) (void)hipEventDestroy(e);
  delete b;
  return AJET_OK;
}

int ajet_batch_create(ajet_ctx* const* runs, int32_t nruns, ajet_batch** out) {
  if (!out) return AJET_EINVAL;
  *out = nullptr;
  if (!runs || nruns < 1 || nruns > 2047) return AJET_EINVAL;
  ajet_batch* b = new ajet_batch();
  *out = b;  // so ajet_batch_last_error can report; the caller destroys it
  for (int r = 0; r < nruns; ++r) {
    ajet_ctx* c = runs[r];
    if (!c || !c->stream) return bfail(b, AJET_EINVAL, "null context");
    if (c->device != runs[0]->device) return bfail(b, AJET_EINVAL, "runs on different devices");
    const int m = physics_mode(c->p);
    if (m == MODE_SC)
      return bfail(b, AJET_EUNSUPPORTED, "batched shock capturing needs non-zero C_mu, C_beta, C_kappa");
    if (m != physics_mode(runs[0]->p))
      return bfail(b, AJET_EINVAL, "runs differ in do_viscous_terms or shock capturing");
    if (c->p.forcing && !c->forcing_set) return bfail(b, AJET_EINVAL, "a run's forcing parameters are not set");
    if ((int)c->s.zcuts.size() * (int)c->s.rcuts.size() >= (1 << 20)) return bfail(b, AJET_EINVAL, "grid too large");
    b->runs.push_back(c);
    if (c->p.forcing) b->force = true;
  }
  ajet_ctx* c0 = runs[0];
  b->device = c0->device;
  b->mode = physics_mode(c0->p);
  b->stream = c0->stream;
  b->aux = c0->aux;
  BHIPCHK(b, hipSetDevice(b->device));
  BHIPCHK(b, hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  BHIPCHK(b, hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  BHIPCHK(b, hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming));
  // per-stage arguments of every run: the buffer rotation of ajet_rk4_stage
  for (int st = 1; st <= 4; ++st) {
    std::vector<StageArgs> v;
    for (ajet_ctx* c : b->runs) {
      StageArgs a = stage_args(c);
      a.qn = c->Q0;
      a.acc_a = c->A;
      a.acc_s = c->S;
      switch (st) {
        case 1: a.qin = c->Q0; a.qout = c->Q1; break;
        case 2: a.qin = c->Q1; a.qout = c->Q2; break;
        case 3: a.qin = c->Q2; a.qout = c->Q1; break;
        default: a.qin = c->Q1; a.qout = c->Q0; break;
      }
      a.tiles = nullptr;
      a.ntiles = 0;
      v.push_back(a);
    }
    BHIPCHK(b, hipMalloc((void**)&b->args[st - 1], sizeof(StageArgs) * v.size()));
    BHIPCHK(b, hipMemcpy(b->args[st - 1], v.data(), sizeof(StageArgs) * v.size(), hipMemcpyHostToDevice));
  }
  // work lists: interior and boundary tiles of all runs; the first boundary tile of a run writes
  // its step scalars
  std::vector<uint32_t> li, lb;
  for (int r = 0; r < nruns; ++r) {
    ajet_ctx* c = b->runs[r];
    std::vector<int> it, bt;
    classify_tiles(c->s, c->p.nz, c->p.nr, it, bt);
    if (bt.empty()) return bfail(b, AJET_EINVAL, "a run has no boundary tile");
    for (int t : it) li.push_back((uint32_t)t | ((uint32_t)r << 20));
    for (size_t k = 0; k < bt.size(); ++k)
      lb.push_back((uint32_t)bt[k] | ((uint32_t)r << 20) | (k == 0 ? 0x80000000u : 0u));
  }
  b->n_int = (int)li.size();
  b->n_bnd = (int)lb.size();
  BHIPCHK(b, hipMalloc((void**)&b->d_int, sizeof(uint32_t) * (li.size() + 1)));
  BHIPCHK(b, hipMalloc((void**)&b->d_bnd, sizeof(uint32_t) * (lb.size() + 1)));
  if (!li.empty()) BHIPCHK(b, hipMemcpy(b->d_int, li.data(), sizeof(uint32_t) * li.size(), hipMemcpyHostToDevice));
  BHIPCHK(b, hipMemcpy(b->d_bnd, lb.data(), sizeof(uint32_t) * lb.size(), hipMemcpyHostToDevice));
  return AJET_OK;
}

// One RK stage of every run: the artificial coefficients of each LAD run (its own launches, in
// run order), then the stage kernels of all runs, boundary tiles on the second stream beside
// the interior tiles.
static hipError_t batch_stage(ajet_batch* b, int stage, int dt_from_dmax, int parity) {
  hipError_t e;
  if (b->mode == MODE_LAD) {
    for (ajet_ctx* c : b->runs) {
      const double* q = stage == 1 ? c->Q0 : (stage == 3 ? c->Q2 : c->Q1);
      if ((e = launch_coefficients(c, stage, q, b->stream)) != hipSuccess) return e;
    }
  }
  if (stage <= 2) {
    // interior, then boundary tiles on the one stream (as a single run's stages 1 and 2)
    if ((e = launch_stage_batch(stage, b->mode, true, b->force, b->args[stage - 1], b->d_int, b->n_int, dt_from_dmax,
                                parity, b->stream)) != hipSuccess)
      return e;
    return launch_stage_batch(stage, b->mode, false, b->force, b->args[stage - 1], b->d_bnd, b->n_bnd, dt_from_dmax,
                              parity, b->stream);
  }
  if ((e = hipEventRecord(b->ev_fork, b->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(b->aux, b->ev_fork, 0)) != hipSuccess) return e;
  if ((e = launch_stage_batch(stage, b->mode, false, b->force, b->args[stage - 1], b->d_bnd, b->n_bnd, dt_from_dmax,
                              parity, b->aux)) != hipSuccess)
    return e;
  if ((e = hipEventRecord(b->ev_join, b->aux)) != hipSuccess) return e;
  if ((e = launch_stage_batch(stage, b->mode, true, b->force, b->args[stage - 1], b->d_int, b->n_int, dt_from_dmax,
                              parity, b->stream)) != hipSuccess)
    return e;
  if ((e = hipStreamWaitEvent(b->stream, b->ev_join, 0)) != hipSuccess) return e;
  if (b->mode == MODE_LAD && stage == 4)
    for (ajet_ctx* c : b->runs)
      if (c->learned && (e = learned_after_step(c, b->stream)) != hipSuccess) return e;
  return hipSuccess;
}

int ajet_batch_run(ajet_batch* b, int32_t nsteps) {
  if (!b || nsteps < 0) return AJET_EINVAL;
  if (nsteps == 0) return AJET_OK;
  BHIPCHK(b, hipSetDevice(b->device));
  const int parity0 = b->runs[0]->parity;
  for (ajet_ctx* c : b->runs) {
    if (!c->have_state) return bfail(b, AJET_EINVAL, "a run has no state loaded");
    if (c->parity != parity0) return bfail(b, AJET_EINVAL, "runs differ in step parity");
  }
  // work already queued on the runs' own streams comes first
  for (ajet_ctx* c : b->runs)
    if (c->stream != b->stream) BHIPCHK(b, hipStreamSynchronize(c->stream));
  // the first step's time steps, run by run: the reduction fused into a previous last stage
  // if there is one, else the stand-alone sweep (ajet_compute_dt)
  for (ajet_ctx* c : b->runs) {
    hipError_t e;
    if (c->fused_valid) {
      e = launch_finalize(c->sc, c->p.CFL, c->g.any_blank, c->dtlog, c->logcap, b->stream);
    } else {
      CflArgs a{};
      a.q = c->Q0;
      a.meta = c->meta;
      a.dxidz = c->dxidz;
      a.detadr = c->detadr;
      a.sc = c->sc;
      a.g = c->g;
      a.Re = c->p.Re;
      a.Pr = c->p.Pr;
      e = launch_cfl(a, c->p.CFL, c->g.any_blank, c->dtlog, c->logcap, b->stream);
    }
    if (e != hipSuccess) return bfail(b, (int)e, std::string("time-step reduction: ") + hipGetErrorString(e));
  }
  int parity = parity0;
  for (int st = 1; st <= 4; ++st) BHIPCHK(b, batch_stage(b, st, 0, parity));
  parity ^= 1;
  // later steps: stage 1 finishes the previous step's fused reduction; one graph per parity,
  // except with LAD runs (launched directly: the learned mode's network runs every refresh-th
  // step, and each stage's LAD work dwarfs its launch cost)
  for (int k = 1; k < nsteps; ++k) {
    if (b->mode == MODE_LAD) {
      for (int st = 1; st <= 4; ++st) BHIPCHK(b, batch_stage(b, st, 1, parity));
      parity ^= 1;
      continue;
    }
    if (!b->graph[parity]) {
      hipGraph_t g = nullptr;
      BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (int st = 1; st <= 4 && e == hipSuccess; ++st) e = batch_stage(b, st, 1, parity);
      hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      if (e != hipSuccess) return bfail(b, (int)e, std::string("capture: ") + hipGetErrorString(e));
      BHIPCHK(b, e2);
      e = hipGraphInstantiate(&b->graph[parity], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      BHIPCHK(b, e);
    }
    BHIPCHK(b, hipGraphLaunch(b->graph[parity], b->stream));
    parity ^= 1;
  }
  BHIPCHK(b, hipEventRecord(b->ev_done, b->stream));
  for (ajet_ctx* c : b->runs) {
    if (c->stream != b->stream) BHIPCHK(b, hipStreamWaitEvent(c->stream, b->ev_done, 0));
    c->parity = parity;
    c->fused_valid = true;
    c->pending_finalize = false;
    c->dt_valid = true;
  }
  return AJET_OK;
}

}  // extern "C"
