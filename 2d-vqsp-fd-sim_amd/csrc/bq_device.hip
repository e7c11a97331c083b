// bq_device.hip - CDNA4 (gfx950) kernels and C ABI of the Biot time loop.
//
// One time step of the reference (time_loop.F90:91-190, serial build) is two
// launches on the handle's stream:
//
//   k_step      the fused interior tiles (full, half and quarter width: both
//               half-steps of every interior-formula cell in one march), the
//               stress half-step of every other cell (HPLANE_*_STRESS: PML
//               strips, surface and bottom rows) and the receiver samples of
//               the previous step
//   k_velocity  the velocity half-step of every other cell (HPLANE_*_VELOCITY,
//               INH_*; the previous filtration velocity is the cell's own
//               value before the update, so the QXP/QZP copies of the
//               reference are not stored), with row 1 (HPLANE_1_VELOCITY) and
//               row PMLZH (HPLANE_PMLZH0_VELOCITY) and the bottom history in
//               its edge blocks
//
// k_edges takes rows 1 and PMLZH in a launch of their own where the velocity
// launch cannot (two steps per sweep; a PML too wide for the edge blocks);
// k_sample takes the samples of a run's last step; k_sweep2 and
// k_rest_stress are the two-step sweep.
//
// Each cell is evaluated with the reference's operation order and no
// floating-point contraction (the file is built with -ffp-contract=off), so
// the fields are bit-identical to the source-order binary32 evaluation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/biotfd.h"
#include "bq_handle.h"

namespace {

// The velocity half-step of the cells outside the fused blocks: the edge
// blocks (rows 1-3 and PMLZH-2..PMLZH, where the split folds the edge rows
// in), then nrest blocks of cells one per thread (per-cell path), then nruns
// blocks of runs.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_velocity(KParams k, Split sp, int nrest, int nruns) {
  __shared__ float lds[kEdgeLds];
  velocity_block<NR>(k, sp, blockIdx.x, nrest, nruns, lds);
}

// First launch of a step: fused interior blocks (long-running, first), then
// the stress half-step of every other cell (per-cell items and runs), then
// the receiver samples of the previous step, taken from the current set,
// which this launch only reads; the items after the fused blocks take
// kBlock threads each, two to a block. All parts are independent of each
// other.
struct Launch1 {
  int nfused, nrest, nruns, nsample, sample_step;
};

template <int NR, int FR>
__global__ __launch_bounds__(kStepBlock) void k_step(KParams k, Fused fu, Split sp, Launch1 ln,
                                                     int itile) {
  __shared__ FusedLds L;
  int b = blockIdx.x;
  if (b < ln.nfused) {
    const int item = fused_item(fu, b);
    if (item >= 0) fused_block<FR>(k, fu, item, itile, L);
    return;
  }
  // the other parts work in kBlock threads: two of their items per block
  const int tid = threadIdx.x & (kBlock - 1);
  b = 2 * (b - ln.nfused) + (threadIdx.x >= kBlock);
  if (b < ln.nrest) {
    int i, l;
    if (rest_cell(sp, b * kBlock + tid, &i, &l)) stress_cell(k, i, l, itile);
    return;
  }
  b -= ln.nrest;
  if (b < ln.nruns) {
    int tile;
    const int r = run_rect(sp, xcd_tile(b, ln.nruns), &tile);
    stress_run<NR>(k, sp.qx0[r], sp.qxl[r], sp.qntx[r], sp.qz0[r], sp.qz1[r], tile, tid >> 6,
                   tid & 63, itile);
    return;
  }
  b -= ln.nruns;
  if (b < ln.nsample) sample_receivers(k, b * kBlock + tid, ln.sample_step, true);
}

// Two-step sweep of steps itile, itile+1 over the region: set A -> set C,
// and the rim's level n+1 into the scratch set T.
__global__ __launch_bounds__(kF2T, 2) void k_sweep2(KParams kac, Fused2 f2, F8 rim, int nregion,
                                                    int itile) {
  __shared__ Fused2Lds L;
  fused2_block(kac, f2, rim, xcd_tile(blockIdx.x, nregion), itile, L);
}

// The stress half-step of the cells outside a two-step region (per-cell
// blocks, then runs), then the receiver samples of step sample_step from the
// current set; no LDS, so these blocks fit beside the two-step tiles.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_rest_stress(KParams k, Split sp, Launch1 ln, int itile) {
  int b = blockIdx.x;
  if (b < ln.nrest) {
    int i, l;
    if (rest_cell(sp, b * kBlock + threadIdx.x, &i, &l)) stress_cell(k, i, l, itile);
    return;
  }
  b -= ln.nrest;
  if (b < ln.nruns) {
    int tile;
    const int r = run_rect(sp, xcd_tile(b, ln.nruns), &tile);
    stress_run<NR>(k, sp.qx0[r], sp.qxl[r], sp.qntx[r], sp.qz0[r], sp.qz1[r], tile, threadIdx.x >> 6,
                   threadIdx.x & 63, itile);
    return;
  }
  b -= ln.nruns;
  sample_receivers(k, b * kBlock + threadIdx.x, ln.sample_step, true);
}

// ------------------------------------------------------------------- edges
// Row 1, stress imaging (HPLANE_1_VELOCITY, hplane_1_velocity.F90:61-76):
// INH_V_QX_INT_1_SI (inh_v_qx_int_1_si.F90:26-31) for all x first, then
// INH_W_QZ_INT_1_SI (inh_w_qz_int_1_si.F90:30-44), which reads the new row-1
// U and QX at i+1: each block recomputes the one halo column it needs.
// Row PMLZH (HPLANE_PMLZH0_VELOCITY, hplane_pmlzh0_velocity.F90:34-63) from
// the new rows PMLZH-1, PMLZH-2 and the two stored levels, then the history
// is advanced, as the store of time_loop.F90:133-155 sees it next step.
__global__ __launch_bounds__(kEdgeBlock) void k_edges(KParams k) {
  __shared__ float su[kEdgeBlock + 1], sq[kEdgeBlock + 1];
  edges_block(k, blockIdx.x, su, sq);
}

// Receiver samples of the last step of a run, from the set it wrote.
__global__ void k_sample(KParams k, int itile) {
  sample_receivers(k, blockIdx.x * blockDim.x + threadIdx.x, itile, false);
}

#undef FA

}  // namespace

// ===================================================================== host

namespace {


int fail(bq_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                             \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(h, -100, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

template <class Tp>
int dalloc(bq_handle *h, Tp **p, size_t n) {
  void *q = nullptr;
  size_t bytes = n ? n * sizeof(Tp) : 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(h, -101, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemsetAsync(q, 0, bytes, h->stream);
  if (e != hipSuccess) return fail(h, -101, std::string("hipMemset: ") + hipGetErrorString(e));
  h->allocs.push_back(q);
  *p = (Tp *)q;
  return 0;
}

// Frees one buffer of the handle (after the stream has drained).
void dfree(bq_handle *h, void *p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it == h->allocs.end()) return;
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(p);
  h->allocs.erase(it);
}

// The partition with fused tiles of fr rows per block.
bq::Plan tiles_plan(const bq_handle *h, int fr) {
  return bq::plan_cells(h->g, h->kp.OFF, kFW, fr, kRun, h->z0, h->z0);
}

// The partition bq_run uses: the one with fused tiles for BQ_TILES_FUSED, the
// one without for BQ_TILES_CELLS, and for BQ_TILES_AUTO the one with tiles
// where bq::fused_worth finds enough of them.
void select_tiles(bq_handle *h) {
  const bool fused = h->tile_mode == BQ_TILES_FUSED || (h->tile_mode == BQ_TILES_AUTO && h->tiles_auto);
  const bq::Plan &p = fused ? h->tiles : h->cells;
  h->fu = p.fu;
  h->ss = p.ss;
  h->sv = p.sv;
  h->fr = h->fr_tiles;
  h->fused_cells = p.fused_cells;
  h->stress_cells = p.stress_cells;
  h->velocity_cells = p.velocity_cells;
}

}  // namespace

// Message of the last failed bq_create on this thread (there is no handle to
// hold it); bq_last_error(NULL) returns it.
static thread_local std::string g_create_err;

static int create_fail(int rc, const std::string &msg) {
  g_create_err = "bq_create: " + msg;
  return rc;
}

extern "C" {

int bq_create(const bq_desc *d, bq_handle **out) {
  if (!d || !out) return create_fail(-1, "null descriptor or output pointer");
  *out = nullptr;
  g_create_err.clear();
  if (d->mx < 8 || d->mz < 6 || d->tpml < 1 || d->tpml > 256 || d->jmnum < 1 || d->mt2 < 1)
    return create_fail(-2, "need mx >= 8, mz >= 6, 1 <= tpml <= 256, jmnum >= 1, mt2 >= 1");
  if (d->lin_z < 2 || d->lin_z > d->mz - 2 || d->lin_x < 2 - d->tpml ||
      d->lin_x > d->mx - 1 + d->tpml)
    return create_fail(-4, "source cell outside lin_x in [2-tpml, mx-1+tpml], lin_z in [2, mz-2]");
  if (d->variant & ~(BQ_AFDA_ZZ | BQ_AFDA_XZ | BQ_SECOND))
    return create_fail(-1, "unknown variant bits");
  const long long nx = (long long)d->mx + 2LL * d->tpml;
  const long long nz = (long long)d->mz + d->tpml + 1;
  const long long pitch = (nx + 63) / 64 * 64;
  if (pitch * nz >= (1LL << 31)) return create_fail(-2, "grid of 2^31 cells or more per field");
  bq_handle *h = new bq_handle();
  GridDesc &g = h->g;
  g.MX = d->mx; g.MZ = d->mz; g.T = d->tpml; g.PMLL = 1 - g.T; g.PMLXH = g.MX + g.T;
  g.PMLZH = g.MZ + g.T; g.NX = (int)nx; g.NZ = (int)nz; g.LIN_X = d->lin_x; g.LIN_Z = d->lin_z;
  g.KTBO = d->ktbo; g.MT2 = d->mt2; g.JMNUM = d->jmnum;
  g.H = (float)d->h; g.DT = (float)d->dt;
  h->P = (int)pitch;
  (void)hipGetDevice(&h->device);
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return create_fail(-100, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
  }
  KParams &k = h->kp;
  k.MX = g.MX; k.MZ = g.MZ; k.T = g.T; k.PMLL = g.PMLL; k.PMLZH = g.PMLZH; k.NX = g.NX;
  k.P = h->P; k.OFF = -g.PMLL; k.LIN_X = g.LIN_X; k.LIN_Z = g.LIN_Z; k.MR = 0;
  k.SI = d->stress_imaging != 0;
  k.AFDA_ZZ = (d->variant & BQ_AFDA_ZZ) != 0;
  k.AFDA_XZ = (d->variant & BQ_AFDA_XZ) != 0;
  k.SECOND = (d->variant & BQ_SECOND) != 0;
  // MODULE AUXIL constants (mod_auxil.F90:19-20) and AUX1 (aux1.F90:44-46)
  k.DT = g.DT;
  k.A = k.SECOND ? 0.0f : -1.0f / 24.0f;  // mod_auxil.F90:15-21
  k.B = k.SECOND ? 1.0f : 9.0f / 8.0f;
  k.AH = k.A / g.H;
  k.BH = k.B / g.H;
  k.CH = 1.0f / g.H;
  k.BC = k.BH / k.CH;  // B/C of strain_uw_zz.F90:43
  const int T = g.T, NX = g.NX, NZ = g.NZ;
  int rc = 0;
  const size_t fsz = (size_t)h->P * NZ;
  rc |= dalloc(h, &h->fields[0], 8 * fsz);
  rc |= dalloc(h, &h->fields[1], 8 * fsz);
  if (rc) {
    create_fail(rc, h->err);
    bq_destroy(h);
    return rc;
  }
  const float **f0[8] = {&k.U0, &k.W0, &k.QX0, &k.QZ0, &k.TXX0, &k.TZZ0, &k.TXZ0, &k.PR0};
  float **f1[8] = {&k.U1, &k.W1, &k.QX1, &k.QZ1, &k.TXX1, &k.TZZ1, &k.TXZ1, &k.PR1};
  for (int f = 0; f < 8; ++f) {
    *f0[f] = h->fields[0] + f * fsz;
    *f1[f] = h->fields[1] + f * fsz;
  }
  // the material index in 16 bits where the table allows (2 B per cell
  // instead of 4 on every step's reads)
  int *jmd = nullptr;
  rc |= dalloc(h, &jmd, fsz);
  const bool jm16 = g.JMNUM <= 65535;
  k.JM = jm16 ? nullptr : jmd;
  k.JM16 = jm16 ? reinterpret_cast<const unsigned short *>(jmd) : nullptr;
  float4 *st = nullptr, *vt = nullptr;
  float2 *at = nullptr;
  rc |= dalloc(h, &st, 3 * (size_t)g.JMNUM);
  rc |= dalloc(h, &vt, 3 * (size_t)g.JMNUM);
  rc |= dalloc(h, &at, (size_t)g.JMNUM);
  k.stab = st; k.vtab = vt; k.atab = at;
  for (int t = 0; t < 8; ++t) {
    float *a = nullptr, *b = nullptr;
    rc |= dalloc(h, &a, (size_t)(T + 1) * NZ);
    rc |= dalloc(h, &b, (size_t)(T + 2) * NZ);
    k.x1[t] = a; k.x2[t] = b;
  }
  for (int t = 0; t < 4; ++t) {
    float *a = nullptr, *b = nullptr;
    rc |= dalloc(h, &a, (size_t)(T + 1) * NX);
    rc |= dalloc(h, &b, (size_t)(T + 2) * NX);
    k.z1[t] = a; k.z2[t] = b;
  }
  // PML memories, shapes of ppml.F90:88-112
  rc |= dalloc(h, &k.PZXL, (size_t)(T + 2) * NZ);
  rc |= dalloc(h, &k.PZXH, (size_t)(T + 1) * NZ);
  rc |= dalloc(h, &k.PXXL, (size_t)(T + 1) * NZ);
  rc |= dalloc(h, &k.PXXH, (size_t)(T + 1) * NZ);
  rc |= dalloc(h, &k.OXXL, (size_t)(T + 1) * NZ);
  rc |= dalloc(h, &k.OXXH, (size_t)(T + 1) * NZ);
  rc |= dalloc(h, &k.PXZH, (size_t)(T + 2) * NX);
  rc |= dalloc(h, &k.PZZH, (size_t)(T + 1) * NX);
  rc |= dalloc(h, &k.OZZH, (size_t)(T + 1) * NX);
  float **l3[4] = {&k.RXXL, &k.SXXL, &k.GXXL, &k.HXXL};
  for (auto p : l3) rc |= dalloc(h, p, (size_t)(T + 3) * NZ);
  float **r2[4] = {&k.RXXH, &k.SXXH, &k.GXXH, &k.HXXH};
  for (auto p : r2) rc |= dalloc(h, p, (size_t)(T + 2) * NZ);
  float **z2m[4] = {&k.RZXL, &k.GZXL, &k.RZXH, &k.GZXH};
  for (auto p : z2m) rc |= dalloc(h, p, (size_t)(T + 2) * NZ);
  float **zb[6] = {&k.RXZH, &k.GXZH, &k.RZZH, &k.SZZH, &k.GZZH, &k.HZZH};
  for (auto p : zb) rc |= dalloc(h, p, (size_t)(T + 1) * NX);
  rc |= dalloc(h, &k.SZc, (size_t)4 * NX * 3);
  rc |= dalloc(h, &k.SZp, (size_t)4 * NX * 3);
  float2 *ds = nullptr;
  rc |= dalloc(h, &ds, (size_t)g.MT2);
  k.dsrc = ds;
  if (rc) {
    create_fail(-101, h->err);
    bq_destroy(h);
    return -101;
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    create_fail(-100, "stream synchronize after clearing the grid failed");
    bq_destroy(h);
    return -100;
  }
  // with stress imaging the cells of both half-steps start at row 2: rows 0, 1
  // of the stresses are images written by rows 2-4, row 1 of the velocities is
  // k_edges', row 0 of the velocities is never updated; without it every row
  // from 0 is a per-cell row
  const int z0 = k.SI ? 2 : 0;
  // rows per fused block: the height whose launch bq::fused_estimate finds
  // shortest. A block marches its rows one after another, so on a grid with
  // few tiles shorter blocks finish sooner; on a large grid taller ones
  // recompute fewer halo rows.
  h->z0 = z0;
  bq::Plan plan = tiles_plan(h, bq::kTileRows[0]);
  const int fr = bq::fused_pick(plan.fu);
  bq::fused_rows(plan.fu, fr);
  h->tiles = plan;
  h->fr_tiles = fr;
  h->tiles_auto = bq::fused_worth(g, plan.fu);
  // runs of 2 rows per wavefront without fused tiles: these grids are small,
  // and twice the run blocks finish sooner (4 rows: 3-15 % slower there)
  h->cells = bq::plan_cells(g, k.OFF, kFW, fr, kRunCells, z0, z0, true, false);
  select_tiles(h);
  // two steps per sweep where the grid has a region for it: its partition and
  // the scratch set (cleared, like the others)
  const bq::Plan2 plan2 = bq::plan_cells2(g, k.OFF, kF2W, kF2R, kRun, z0, z0);
  if (plan2.f2.nfx > 0) {
    if (dalloc(h, &h->scratch_set, 8 * fsz) || hipStreamSynchronize(h->stream) != hipSuccess) {
      create_fail(-101, h->err);
      bq_destroy(h);
      return -101;
    }
    h->f2 = plan2.f2;
    h->ss2 = plan2.ss;
    h->sv2 = plan2.sv;
    h->region_cells = plan2.region_cells;
    h->stress_cells2 = plan2.stress_cells;
    h->velocity_cells2 = plan2.velocity_cells;
  }
  *out = h;
  return 0;
}

void bq_destroy(bq_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *p : h->allocs) (void)hipFree(p);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char *bq_last_error(bq_handle *h) {
  if (h) return h->err.c_str();
  return g_create_err.empty() ? "null handle" : g_create_err.c_str();
}

int bq_grid_shape(bq_handle *h, int32_t *nx, int32_t *nz) {
  if (!h) return -1;
  if (nx) *nx = h->g.NX;
  if (nz) *nz = h->g.NZ;
  return 0;
}

int bq_upload_model(bq_handle *h, const int32_t *jm, const float *tab) {
  if (!h || !jm || !tab) return fail(h, -1, "null argument");
  const GridDesc &g = h->g;
  const size_t ncell = (size_t)g.NX * g.NZ;
  for (size_t c = 0; c < ncell; ++c)
    if (jm[c] < 1 || jm[c] > g.JMNUM) return fail(h, -5, "material index out of range");
  h->jm_host.assign(jm, jm + ncell);
  h->tab_host.assign(tab, tab + (size_t)g.JMNUM * 24);
  // jm padded to the device pitch
  std::vector<int32_t> jp((size_t)h->P * g.NZ, 1);
  for (int l = 0; l < g.NZ; ++l)
    std::memcpy(&jp[(size_t)l * h->P], jm + (size_t)l * g.NX, sizeof(int32_t) * g.NX);
  std::vector<float> st(12 * (size_t)g.JMNUM, 0.0f), vt(12 * (size_t)g.JMNUM, 0.0f);
  std::vector<float> at(2 * (size_t)g.JMNUM, 0.0f);
  using namespace bq;
  for (int m = 0; m < g.JMNUM; ++m) {
    const float *t = tab + (size_t)m * 24;
    float *s = &st[12 * (size_t)m];
    s[0] = t[T_MU]; s[1] = t[T_SXX1]; s[2] = t[T_SXX2]; s[3] = t[T_SXX3];
    s[4] = t[T_SZZ1]; s[5] = t[T_SZZ2]; s[6] = t[T_SZZ3]; s[7] = t[T_P1];
    s[8] = t[T_P2]; s[9] = t[T_P3];
    // friction factors in the reference's binary32 order (inh_qx_int.F90:44,
    // inh_v_uv_int.F90:46, inh_qz_int.F90:45, inh_v_w_int.F90:48)
    const float eqx = expf(-t[T_REL3_QX]);
    const float eqz = expf(-t[T_REL3_QZ]);
    const float fu = (t[T_REL2_U] / t[T_REL3_QX]) * (1.0f - expf(-t[T_REL3_QX]));
    const float fw = (t[T_REL2_W] / t[T_REL3_QZ]) * (1.0f - expf(-t[T_REL3_QZ]));
    float *v = &vt[12 * (size_t)m];
    v[0] = eqx; v[1] = t[T_REL1_QX]; v[2] = t[T_REL2_QX]; v[3] = fu;
    v[4] = t[T_REL1_U]; v[5] = t[T_REL3_U]; v[6] = eqz; v[7] = t[T_REL1_QZ];
    v[8] = t[T_REL2_QZ]; v[9] = fw; v[10] = t[T_REL1_W]; v[11] = t[T_REL3_W];
    at[2 * m] = t[T_AUXW];
    at[2 * m + 1] = t[T_AUXQZ];
  }
  std::vector<uint16_t> jp16;
  if (h->kp.JM16) {
    jp16.assign(jp.begin(), jp.end());   // 1 <= jm <= JMNUM <= 65535, checked above
    HIPCHK(h, hipMemcpyAsync((void *)h->kp.JM16, jp16.data(), jp16.size() * 2, hipMemcpyHostToDevice,
                             h->stream));
  } else {
    HIPCHK(h, hipMemcpyAsync((void *)h->kp.JM, jp.data(), jp.size() * 4, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipMemcpyAsync((void *)h->kp.stab, st.data(), st.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((void *)h->kp.vtab, vt.data(), vt.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((void *)h->kp.atab, at.data(), at.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->model_ok = true;
  return 0;
}

static int upload_pml_host(bq_handle *h, const bq::PmlHost &p) {
  KParams &k = h->kp;
  for (int t = 0; t < 8; ++t) {
    HIPCHK(h, hipMemcpyAsync((void *)k.x1[t], p.x1[t].data(), p.x1[t].size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync((void *)k.x2[t], p.x2[t].data(), p.x2[t].size() * 4, hipMemcpyHostToDevice, h->stream));
  }
  // bottom coefficients: (0:T[+1], NXc) on the host, [j][ix] on the device
  const int NX = h->g.NX, T = h->g.T;
  std::vector<float> zt[8];
  for (int t = 0; t < 8; ++t) {
    const std::vector<float> &src = t < 4 ? p.z1[t] : p.z2[t - 4];
    const int nj = t < 4 ? T + 1 : T + 2;
    zt[t].resize(src.size());
    for (int ix = 0; ix < NX; ++ix)
      for (int j = 0; j < nj; ++j) zt[t][(size_t)j * NX + ix] = src[(size_t)ix * nj + j];
    const float *dst = t < 4 ? k.z1[t] : k.z2[t - 4];
    HIPCHK(h, hipMemcpyAsync((void *)dst, zt[t].data(), zt[t].size() * 4, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->pml_ok = true;
  return 0;
}

int bq_build_pml(bq_handle *h, float r, float tau, float pow_, float sf, float wc) {
  if (!h) return -1;
  if (!h->model_ok) return fail(h, -6, "bq_build_pml needs bq_upload_model first");
  bq::PmlHost p;
  bq::build_pml(h->g, h->jm_host.data(), h->tab_host.data(), r, tau, pow_, sf, wc, p);
  return upload_pml_host(h, p);
}

// MODULE PML order -> (group, index) of PmlHost: XL/XH 0..7 = a b c d of XL,
// then XH; ZH 0..3 = a b c d.
static void pml_slot(int id, int *grp, int *idx) {
  static const int map[24][2] = {
      {0, 0}, {0, 1}, {0, 2}, {0, 4}, {0, 5}, {0, 6}, {0, 3}, {0, 7},  // ODA..ODD _XL/_XH
      {2, 0}, {2, 1}, {2, 2}, {2, 3},                                  // ODA_ZH..ODD_ZH
      {1, 0}, {1, 1}, {1, 2}, {1, 4}, {1, 5}, {1, 6}, {1, 3}, {1, 7},  // ODA2..ODD2 _XL/_XH
      {3, 0}, {3, 1}, {3, 2}, {3, 3}};                                 // ODA2_ZH..ODD2_ZH
  *grp = map[id][0];
  *idx = map[id][1];
}

int bq_upload_pml(bq_handle *h, const float *const *coef) {
  if (!h || !coef) return fail(h, -1, "null argument");
  const GridDesc &g = h->g;
  bq::PmlHost p;
  p.resize(g.T, g.NX, g.NZ);
  for (int id = 0; id < 24; ++id) {
    int grp, idx;
    pml_slot(id, &grp, &idx);
    std::vector<float> &v = grp == 0 ? p.x1[idx] : grp == 1 ? p.x2[idx] : grp == 2 ? p.z1[idx] : p.z2[idx];
    if (!coef[id]) return fail(h, -1, "null PML array");
    std::memcpy(v.data(), coef[id], v.size() * 4);
  }
  return upload_pml_host(h, p);
}

int bq_get_pml(bq_handle *h, int id, float *out) {
  if (!h || id < 0 || id >= 24) return fail(h, -1, "bad PML id");
  int grp, idx;
  pml_slot(id, &grp, &idx);
  const KParams &k = h->kp;
  const float *src = grp == 0 ? k.x1[idx] : grp == 1 ? k.x2[idx] : grp == 2 ? k.z1[idx] : k.z2[idx];
  const GridDesc &g = h->g;
  const size_t n = grp == 0 ? (size_t)(g.T + 1) * g.NZ : grp == 1 ? (size_t)(g.T + 2) * g.NZ
                 : grp == 2 ? (size_t)(g.T + 1) * g.NX : (size_t)(g.T + 2) * g.NX;
  if (out) {
    if (grp < 2) {
      HIPCHK(h, hipMemcpyAsync(out, src, n * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    } else {  // [j][ix] on the device -> (0:T[+1], NXc)
      std::vector<float> t(n);
      HIPCHK(h, hipMemcpyAsync(t.data(), src, n * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      const int nj = grp == 2 ? g.T + 1 : g.T + 2;
      for (int ix = 0; ix < g.NX; ++ix)
        for (int j = 0; j < nj; ++j) out[(size_t)ix * nj + j] = t[(size_t)j * g.NX + ix];
    }
  }
  return (int)n;
}

int bq_upload_bottom_nr(bq_handle *h, const float abw[8], const float abh[8]) {
  if (!h || !abw || !abh) return fail(h, -1, "null argument");
  for (int c = 0; c < 8; ++c) { h->kp.abw[c] = abw[c]; h->kp.abh[c] = abh[c]; }
  return 0;
}

int bq_build_bottom_nr(bq_handle *h, const float prm[4], float abw[8], float abh[8]) {
  if (!h || !prm) return fail(h, -1, "null argument");
  if (!h->model_ok) return fail(h, -6, "bq_build_bottom_nr needs bq_upload_model first");
  const bq::NrParams p{prm[0], prm[1], prm[2], prm[3]};
  float w[8], hh[8];
  if (bq::bottom_nr(h->g, h->jm_host.data(), h->tab_host.data(), p, w, hh))
    return fail(h, -2, "unknown bottom boundary key KTBO");
  for (int c = 0; c < 8; ++c) {
    h->kp.abw[c] = w[c];
    h->kp.abh[c] = hh[c];
    if (abw) abw[c] = w[c];
    if (abh) abh[c] = hh[c];
  }
  return 0;
}

int bq_set_source(bq_handle *h, const float *sourtf) {
  if (!h || !sourtf) return fail(h, -1, "null argument");
  const int n = h->g.MT2;
  const float DT = h->g.DT;
  // STRESS_LIN_NN injects -(S-SOURSP)/DT into TXX, TZZ and +(F-SOURFP)/DT into
  // PRES, SOURSP/SOURFP being the previous step's values starting from 0
  // (stress_lin_nn.F90:58-65, alloc_t_seis.F90:69-70).
  std::vector<float> ds(2 * (size_t)n);
  float sp = 0.0f, fp = 0.0f;
  for (int s = 0; s < n; ++s) {
    const float S = sourtf[2 * (size_t)s], F = sourtf[2 * (size_t)s + 1];
    ds[2 * (size_t)s] = (S - sp) / DT;
    ds[2 * (size_t)s + 1] = (F - fp) / DT;
    sp = S;
    fp = F;
  }
  HIPCHK(h, hipMemcpyAsync((void *)h->kp.dsrc, ds.data(), ds.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->src_ok = true;
  return 0;
}

int bq_set_receivers(bq_handle *h, int mr, const int32_t *irecx, const int32_t *lrecx,
                     const int32_t *irecz, const int32_t *lrecz) {
  if (!h || mr < 0 || (mr > 0 && (!irecx || !lrecx || !irecz || !lrecz)))
    return fail(h, -1, "bad receiver arguments");
  // the trace buffer holds every step from the first: receivers are part of
  // the set-up, as in the reference (read_open.F90:234-236)
  if (h->next_step > 1) return fail(h, -6, "bq_set_receivers after the first step");
  std::vector<int4> r(mr > 0 ? mr : 1);
  for (int j = 0; j < mr; ++j) r[j] = make_int4(irecx[j], lrecx[j], irecz[j], lrecz[j]);
  // a second call before the first step replaces the previous buffers
  dfree(h, (void *)h->kp.rec);
  dfree(h, h->kp.traces);
  h->kp.rec = nullptr;
  h->kp.traces = nullptr;
  h->kp.MR = 0;
  h->mr = 0;
  int4 *rd = nullptr;
  float *tr = nullptr;
  if (dalloc(h, &rd, (size_t)(mr > 0 ? mr : 1))) return -101;
  if (dalloc(h, &tr, (size_t)h->g.MT2 * (mr > 0 ? mr : 1) * 6)) return -101;
  HIPCHK(h, hipMemcpyAsync(rd, r.data(), r.size() * sizeof(int4), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->kp.rec = rd;
  h->kp.traces = tr;
  h->kp.MR = mr;
  h->mr = mr;
  // two-step sweeps keep level n+1 only outside the region and in its rim:
  // every sampled node must lie there
  const bq::Fused2 &f = h->f2;
  auto ok = [&](int x, int z) {
    const bool in = f.nfx > 0 && x >= f.x0 && x < f.x1 && z >= f.z0 && z < f.z1;
    const bool rim = x < f.x0 + 2 || x >= f.x1 - 2 || z < f.z0 + 2 || z >= f.z1 - 2;
    return !in || rim;
  };
  h->rec_sweep_ok = true;
  for (int j = 0; j < mr; ++j)
    if (!ok(irecx[j], lrecx[j]) || !ok(irecz[j], lrecz[j]) || !ok(irecz[j], lrecx[j]))
      h->rec_sweep_ok = false;
  return 0;
}

int bq_set_tiles(bq_handle *h, int mode) {
  if (!h) return -1;
  if (mode != BQ_TILES_AUTO && mode != BQ_TILES_FUSED && mode != BQ_TILES_CELLS)
    return fail(h, -2, "bq_set_tiles: mode must be BQ_TILES_AUTO, BQ_TILES_FUSED or BQ_TILES_CELLS");
  h->tile_mode = mode;
  select_tiles(h);
  return 0;
}

int bq_get_tiles(bq_handle *h) {
  if (!h) return -1;
  return bq::fused_blocks(h->fu) ? BQ_TILES_FUSED : BQ_TILES_CELLS;
}

int bq_get_tile_rows(bq_handle *h) {
  if (!h) return -1;
  return h->fr_tiles;
}

int bq_autotune(bq_handle *h, int trial_steps, double *ms_per_step) {
  if (!h) return -1;
  if (trial_steps < 1) return fail(h, -2, "bq_autotune: trial_steps must be at least 1");
  if (!h->model_ok || !h->pml_ok || !h->src_ok)
    return fail(h, -6, "model, PML and source must be set before bq_autotune");
  if (h->next_step - 1 + (bq::kNumRows + 1LL) * (trial_steps + 1) > h->g.MT2)
    return fail(h, -7, "bq_autotune: the trial steps run beyond MT2");
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto &f : ev)
        if (f) (void)hipEventDestroy(f);
      return fail(h, -100, "bq_autotune: hipEventCreate failed");
    }
  // the tiles apply to single steps, and the trials keep out of bq_kernel_times;
  // a failed trial leaves the partition as it was
  const int sweep = h->sweep, timing = h->timing;
  const bq::Plan tiles0 = h->tiles;
  const int fr0 = h->fr_tiles, mode0 = h->tile_mode;
  h->sweep = 1;
  h->timing = 0;
  double ms[bq::kNumRows + 1] = {};
  int rc = 0;
  for (int c = 0; c <= bq::kNumRows && rc == 0; ++c) {
    if (c < bq::kNumRows) {
      h->tiles = tiles_plan(h, bq::kTileRows[c]);
      h->fr_tiles = bq::kTileRows[c];
    }
    h->tile_mode = c < bq::kNumRows ? BQ_TILES_FUSED : BQ_TILES_CELLS;
    select_tiles(h);
    // one step first, so each candidate's timed steps start warm
    rc = bq_run(h, h->next_step, h->next_step);
    if (rc == 0 && hipEventRecord(ev[0], h->stream) != hipSuccess) rc = fail(h, -100, "hipEventRecord");
    if (rc == 0) rc = bq_run(h, h->next_step, h->next_step + trial_steps - 1);
    float t = 0.f;
    if (rc == 0 && (hipEventRecord(ev[1], h->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                    hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess))
      rc = fail(h, -100, "bq_autotune: event timing failed");
    ms[c] = (double)t / trial_steps;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  h->sweep = sweep;
  h->timing = timing;
  if (rc) {
    h->tiles = tiles0;
    h->fr_tiles = fr0;
    h->tile_mode = mode0;
    select_tiles(h);
    return rc;
  }
  int best = 0;
  for (int c = 1; c < bq::kNumRows; ++c)
    if (ms[c] < ms[best]) best = c;
  h->tiles = tiles_plan(h, bq::kTileRows[best]);
  h->fr_tiles = bq::kTileRows[best];
  h->tile_mode = ms[bq::kNumRows] < ms[best] ? BQ_TILES_CELLS : BQ_TILES_FUSED;
  select_tiles(h);
  if (ms_per_step)
    for (int c = 0; c <= bq::kNumRows; ++c) ms_per_step[c] = ms[c];
  return rc;
}

int bq_set_sweep(bq_handle *h, int steps) {
  if (!h) return -1;
  if (steps != 1 && steps != 2) return fail(h, -2, "bq_set_sweep: steps per sweep must be 1 or 2");
  h->sweep = steps;
  return 0;
}

int bq_get_sweep(bq_handle *h) {
  if (!h) return -1;
  return (h->sweep == 2 && h->f2.nfx > 0 && h->scratch_set && h->rec_sweep_ok && !h->lb) ? 2 : 1;
}

int bq_enable_timing(bq_handle *h, int every) {
  if (!h) return -1;
  if (every < 0) return fail(h, -2, "bq_enable_timing: negative step stride");
  h->timing = every;
  return 0;
}

int bq_run(bq_handle *h, int itile_begin, int itile_end) {
  if (!h) return -1;
  if (!h->model_ok || !h->pml_ok || !h->src_ok)
    return fail(h, -6, "model, PML and source must be set before bq_run");
  if (itile_begin != h->next_step) return fail(h, -6, "steps must run in order");
  if (itile_end > h->g.MT2) return fail(h, -7, "step beyond MT2");
  if (itile_end < itile_begin) return 0;
  const GridDesc &g = h->g;
  const dim3 blk(kBlock);
  Launch1 ln{};
  ln.nfused = bq::fused_blocks(h->fu) ? bq::fused_launch_blocks(h->fu) : 0;
  ln.nrest = (h->ss.cum[kMaxCells] + kBlock - 1) / kBlock;
  ln.nruns = h->ss.qcum[kMaxRuns];
  const int vrest = (h->sv.cum[kMaxCells] + kBlock - 1) / kBlock;
  const int vruns = h->sv.qcum[kMaxRuns];
  const int vgrid = h->sv.etop + h->sv.ebot + vrest + vruns;
  const int nsb = (h->mr + kBlock - 1) / kBlock;
  const dim3 ge((g.NX + kEdgeBlock - 1) / kEdgeBlock);
  // two steps per sweep (bq_get_sweep says when): the partition of the other cells
  const bool two = bq_get_sweep(h) == 2;
  Launch1 l2{};
  l2.nrest = (h->ss2.cum[kMaxCells] + kBlock - 1) / kBlock;
  l2.nruns = h->ss2.qcum[kMaxRuns];
  const int vrest2 = (h->sv2.cum[kMaxCells] + kBlock - 1) / kBlock;
  const int vruns2 = h->sv2.qcum[kMaxRuns];
  const int vgrid2 = h->sv2.etop + h->sv2.ebot + vrest2 + vruns2;
  const int nregion = h->f2.nfx * h->f2.nfz;
  F8 rim{};
  for (int f = 0; f < 8; ++f)
    rim.f[f] = h->scratch_set ? h->scratch_set + (size_t)f * h->P * g.NZ : nullptr;
  const int nsteps = itile_end - itile_begin + 1;
  // events of a timed sweep: before the first launch and after each of up to 7
  constexpr int kEv = 8;
  const int every = h->timing;
  const int ntimed = every ? (nsteps + every - 1) / every : 0;
  if (every) {
    const size_t need = (size_t)ntimed * kEv;
    while (h->ev.size() < need) {
      hipEvent_t e;
      HIPCHK(h, hipEventCreate(&e));
      h->ev.push_back(e);
    }
  }
  std::vector<int> timed_kind;  // 1: single step, 2: two-step sweep, per timed slot
  KParams last{};
  int it = itile_begin;
  while (it <= itile_end) {
    const int rel = it - itile_begin;
    const bool sweep2 = two && it + 1 <= itile_end;
    // timed: the sweeps that hold a step rel = 0, every, 2*every, ...
    const bool timed = every && (rel % every == 0 || (sweep2 && (rel + 1) % every == 0)) &&
                       (int)timed_kind.size() < ntimed;
    hipEvent_t *e = timed ? &h->ev[timed_kind.size() * kEv] : nullptr;
    if (e) {
      timed_kind.push_back(sweep2 ? 2 : 1);
      (void)hipEventRecord(e[0], h->stream);
    }
    if (!sweep2) {
      const KParams k = step_params(h);
      h->parity ^= 1;
      // the samples of step it-1 ride along with the first launch of step it
      ln.nsample = (it > itile_begin && k.MR > 0) ? nsb : 0;
      ln.sample_step = it - 1;
      const int g1 = ln.nfused + (ln.nrest + ln.nruns + ln.nsample + 1) / 2;
      if (ln.nfused == 0) {
        // no fused tiles: one item per 256-thread block, twice the blocks
        const int gs = ln.nrest + ln.nruns + ln.nsample;
        if (gs) {
          if (h->ss.nr == kRunCells)
            hipLaunchKernelGGL(k_rest_stress<kRunCells>, dim3(gs), blk, 0, h->stream, k, h->ss, ln, it);
          else
            hipLaunchKernelGGL(k_rest_stress<kRun>, dim3(gs), blk, 0, h->stream, k, h->ss, ln, it);
        }
      } else if (g1) {
        auto launch = [&](auto kern) {
          hipLaunchKernelGGL(kern, dim3(g1), dim3(kStepBlock), 0, h->stream, k, h->fu, h->ss, ln, it);
        };
        switch (h->fr) {
          case 16: launch(k_step<kRun, 16>); break;
          case 24: launch(k_step<kRun, 24>); break;
          case 32: launch(k_step<kRun, 32>); break;
          case 48: launch(k_step<kRun, 48>); break;
          case 64: launch(k_step<kRun, 64>); break;
          default: launch(k_step<kRun, 128>); break;
        }
      }
      if (e) (void)hipEventRecord(e[1], h->stream);
      if (vgrid) {
        if (h->sv.nr == kRunCells)
          hipLaunchKernelGGL(k_velocity<kRunCells>, dim3(vgrid), blk, 0, h->stream, k, h->sv, vrest, vruns);
        else
          hipLaunchKernelGGL(k_velocity<kRun>, dim3(vgrid), blk, 0, h->stream, k, h->sv, vrest, vruns);
      }
      if (e) (void)hipEventRecord(e[2], h->stream);
      if (!h->sv.fold) {
        hipLaunchKernelGGL(k_edges, ge, dim3(kEdgeBlock), 0, h->stream, k);
        if (e) (void)hipEventRecord(e[3], h->stream);
      }
      if (h->lb && bq::lb_after_step(h, k, it)) return fail(h, -100, "learned boundary launch failed");
      last = k;
      it += 1;
      continue;
    }
    // steps it, it+1: A = current set, C = the other one, T = scratch set
    const KParams kac = step_params(h);
    float *A = (float *)kac.U0, *C = kac.U1, *T = h->scratch_set;
    const KParams kat = params_sets(h, A, T), ktc = params_sets(h, T, C);
    h->parity ^= 1;
    hipLaunchKernelGGL(k_sweep2, dim3(nregion), dim3(kF2T), 0, h->stream, kac, h->f2, rim, nregion,
                       it);
    if (e) (void)hipEventRecord(e[1], h->stream);
    // the other cells, step it (A -> T), with the samples of step it-1 from A
    l2.nsample = (it > itile_begin && kac.MR > 0) ? nsb : 0;
    l2.sample_step = it - 1;
    int gs = l2.nrest + l2.nruns + l2.nsample;
    if (gs) hipLaunchKernelGGL(k_rest_stress<kRun>, dim3(gs), blk, 0, h->stream, kat, h->ss2, l2, it);
    if (e) (void)hipEventRecord(e[2], h->stream);
    if (vgrid2)
      hipLaunchKernelGGL(k_velocity<kRun>, dim3(vgrid2), blk, 0, h->stream, kat, h->sv2, vrest2,
                         vruns2);
    if (e) (void)hipEventRecord(e[3], h->stream);
    hipLaunchKernelGGL(k_edges, ge, dim3(kEdgeBlock), 0, h->stream, kat);
    if (e) (void)hipEventRecord(e[4], h->stream);
    // step it+1 (T -> C), with the samples of step it from T
    l2.nsample = kac.MR > 0 ? nsb : 0;
    l2.sample_step = it;
    gs = l2.nrest + l2.nruns + l2.nsample;
    if (gs)
      hipLaunchKernelGGL(k_rest_stress<kRun>, dim3(gs), blk, 0, h->stream, ktc, h->ss2, l2, it + 1);
    if (e) (void)hipEventRecord(e[5], h->stream);
    if (vgrid2)
      hipLaunchKernelGGL(k_velocity<kRun>, dim3(vgrid2), blk, 0, h->stream, ktc, h->sv2, vrest2,
                         vruns2);
    if (e) (void)hipEventRecord(e[6], h->stream);
    hipLaunchKernelGGL(k_edges, ge, dim3(kEdgeBlock), 0, h->stream, ktc);
    if (e) (void)hipEventRecord(e[7], h->stream);
    last = ktc;
    it += 2;
  }
  if (last.MR > 0)
    hipLaunchKernelGGL(k_sample, dim3(nsb), dim3(kBlock), 0, h->stream, last, itile_end);
  HIPCHK(h, hipGetLastError());
  h->next_step = itile_end + 1;
  h->last_sweep = two ? 2 : 1;
  if (every && !timed_kind.empty()) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    auto dt = [&](size_t slot, int a, int b) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, h->ev[slot * kEv + a], h->ev[slot * kEv + b]);
      return (double)ms;
    };
    double s1[3] = {0, 0, 0}, s2[4] = {0, 0, 0, 0};
    int n1 = 0, n2 = 0;
    for (size_t q = 0; q < timed_kind.size(); ++q) {
      if (timed_kind[q] == 1) {
        // (with the edge rows folded into k_velocity there is no k_edges)
        for (int c = 0; c < 3; ++c) s1[c] += (c == 2 && h->sv.fold) ? 0.0 : dt(q, c, c + 1);
        ++n1;
      } else {
        s2[0] += dt(q, 0, 1);                          // the two-step sweep
        s2[1] += dt(q, 1, 2) + dt(q, 4, 5);            // other cells, stress (two launches)
        s2[2] += dt(q, 2, 3) + dt(q, 5, 6);            // velocity
        s2[3] += dt(q, 3, 4) + dt(q, 6, 7);            // edges
        ++n2;
      }
    }
    if (n1)
      for (int c = 0; c < 3; ++c) h->ktime[c] = s1[c] / n1;
    if (n2) {
      h->ktime2[0] = s2[0] / n2;
      for (int c = 1; c < 4; ++c) h->ktime2[c] = s2[c] / (2.0 * n2);
    }
    h->timed_steps = n1 + n2;
    h->last_sweep = n2 ? 2 : 1;
  }
  return 0;
}

// Durations (ms per launch) and cells of the last timed run. steps_per_sweep
// = 2: ms = {two-step sweep, stress of the other cells, velocity, edges},
// cells = {region, other stress cells, other velocity cells, 0};
// 1: ms = {k_step, k_velocity, k_edges (0 where k_velocity takes the edge
// rows), 0}, cells = {fused, stress, velocity, 0}.
int bq_kernel_stats(bq_handle *h, double ms[4], long long cells[4], int *steps_per_sweep) {
  if (!h || !ms || !cells || !steps_per_sweep) return -1;
  if (!h->timed_steps) return fail(h, -6, "no timed run");
  *steps_per_sweep = h->last_sweep;
  if (h->last_sweep == 2) {
    for (int c = 0; c < 4; ++c) ms[c] = h->ktime2[c];
    cells[0] = h->region_cells;
    cells[1] = h->stress_cells2;
    cells[2] = h->velocity_cells2;
  } else {
    for (int c = 0; c < 3; ++c) ms[c] = h->ktime[c];
    ms[3] = 0;
    cells[0] = h->fused_cells;
    cells[1] = h->stress_cells;
    cells[2] = h->velocity_cells;
  }
  cells[3] = 0;
  return 0;
}

int bq_kernel_cells(bq_handle *h, long long out[3]) {
  if (!h || !out) return -1;
  out[0] = h->fused_cells;
  out[1] = h->stress_cells;
  out[2] = h->velocity_cells;
  return 0;
}

int bq_kernel_times(bq_handle *h, double out[3]) {
  if (!h || !out) return -1;
  if (!h->timed_steps) return fail(h, -6, "no timed run");
  if (h->last_sweep == 2) {
    out[0] = h->ktime2[0];
    out[1] = h->ktime2[2];
    out[2] = h->ktime2[3];
  } else {
    for (int q = 0; q < 3; ++q) out[q] = h->ktime[q];
  }
  return 0;
}

int bq_sync(bq_handle *h) {
  if (!h) return -1;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int bq_get_traces(bq_handle *h, float *out, int first_step, int nsteps) {
  if (!h || !out) return fail(h, -1, "null argument");
  if (!h->kp.traces || first_step < 1 || nsteps < 0 || first_step + nsteps - 1 > h->g.MT2)
    return fail(h, -9, "trace range");
  HIPCHK(h, hipMemcpyAsync(out, h->kp.traces + (size_t)(first_step - 1) * h->mr * 6,
                           (size_t)nsteps * h->mr * 6 * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// The latest values of a field: the current set of the next step.
static float *field_ptr(bq_handle *h, int id, int which) {
  if (id < 0 || id >= 8) return nullptr;
  return h->fields[which] + (size_t)id * h->P * h->g.NZ;
}

int bq_get_field(bq_handle *h, int field_id, float *host) {
  if (!h || !host) return fail(h, -1, "null argument");
  float *d = field_ptr(h, field_id, h->parity);
  if (!d) return fail(h, -8, "bad field id");
  HIPCHK(h, hipMemcpy2DAsync(host, (size_t)h->g.NX * 4, d, (size_t)h->P * 4, (size_t)h->g.NX * 4,
                             h->g.NZ, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int bq_set_field(bq_handle *h, int field_id, const float *host) {
  if (!h || !host) return fail(h, -1, "null argument");
  // every set (the scratch set of two-step sweeps too): cells no step writes
  // must hold the same value in each
  for (int which = 0; which < 3; ++which) {
    if (which == 2 && !h->scratch_set) break;
    float *d = which < 2 ? field_ptr(h, field_id, which)
                         : (field_id >= 0 && field_id < 8
                                ? h->scratch_set + (size_t)field_id * h->P * h->g.NZ
                                : nullptr);
    if (!d) return fail(h, -8, "bad field id");
    HIPCHK(h, hipMemcpy2DAsync(d, (size_t)h->P * 4, host, (size_t)h->g.NX * 4, (size_t)h->g.NX * 4,
                               h->g.NZ, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"
